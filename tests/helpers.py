"""Shared helpers for the parity tests (CPU oracle vs HIP path)."""
import numpy as np

from axonexa_amd import abi

FLOAT_FIELDS = ("pos", "disp", "r0", "E", "lf", "mom")
INT_FIELDS = ("comp", "comp0", "li", "ncoll", "ntran", "ncap", "step")


def mask_moments(st, n_walkers, n_profiles, axes):
    """Copy of a state with the moments of axes outside the bit mask zeroed
    (a walk that keeps every axis vs one that keeps only those in use)."""
    st = np.array(st, copy=True)
    m = abi.state_view(st, n_walkers, n_profiles)["mom"]
    for a in range(3):
        if not (axes >> a) & 1:
            m[:, :, a] = 0.0
    return st


def compare_states(gst, ost, n_walkers, n_profiles, rtol=1e-5):
    """Integer fields must be identical; float fields within rtol relative
    (scale: max(|x|, 1)).  Returns a dict of diagnostics."""
    g = abi.state_view(gst, n_walkers, n_profiles)
    o = abi.state_view(ost, n_walkers, n_profiles)
    out = {}
    for k in INT_FIELDS:
        diff = np.nonzero(np.asarray(g[k]) != np.asarray(o[k]))[0]
        out[k] = diff.size
        assert diff.size == 0, f"{k}: {diff.size} walkers differ, first {diff[:5]}"
    worst = 0.0
    for k in FLOAT_FIELDS:
        a = np.asarray(g[k], np.float64)
        b = np.asarray(o[k], np.float64)
        err = np.abs(a - b) / np.maximum(np.abs(b), 1.0)
        m = float(err.max()) if err.size else 0.0
        worst = max(worst, m)
        assert m <= rtol, f"{k}: max relative error {m:.3g} > {rtol}"
    bit = np.all(gst.reshape(-1, n_walkers) == ost.reshape(-1, n_walkers), axis=0)
    out["max_rel"] = worst
    out["bitwise_fraction"] = float(bit.mean())
    return out


def fmt_case(name, **kw):
    return name + " " + " ".join(f"{k}={v}" for k, v in kw.items())
