"""Substrate builders: periodic cells with cylinders, spheres or planes.

Equivalents of how the reference drivers assemble a ``Lattice`` and its
``basis[]``:
  * one cylinder at the cell centre  — examples/Cos_OneCyl.cu:103-114
  * hexagonal cylinder packing       — examples/hexPhase.cu:108-125
  * gamma-distributed radii, random sequential packing with periodic edge
    copies                           — src/substrate/RPSinitializer.h:76-186,
                                       examples/gammadist.cu:118-149
  * an empty cell (basisSize = 0)    — examples/Cos_Empty_Mel.cu:139-141
  * two parallel planes              — examples/planePhase.cu:114-117
Periodic images are added by the engine itself, so builders return each
primitive once (centres may lie anywhere; they are wrapped into the cell).
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from .abi import PRIM_CYLINDER_XY, PRIM_PLANE, PRIM_SPHERE, SubstrateSpec


def empty_box(L: float, D: float, T2: float = float("inf")) -> SubstrateSpec:
    return SubstrateSpec(a=L, b=L, c=L, D_lat=D, T2_lat=T2)


def single_cylinder(cell: float, R: float, D_in: float, D_ex: float,
                    T2_in: float = float("inf"), T2_ex: float = float("inf"),
                    kappa: float = 0.0, center=None) -> SubstrateSpec:
    c = [cell / 2.0, cell / 2.0, 0.0] if center is None else list(center)
    return SubstrateSpec(a=cell, b=cell, c=cell, D_lat=D_ex, T2_lat=T2_ex,
                         types=[PRIM_CYLINDER_XY], centers=[c], radii=[R],
                         D=D_in, T2=T2_in, kappa=kappa)


def single_sphere(cell: float, R: float, D_in: float, D_ex: float,
                  T2_in: float = float("inf"), T2_ex: float = float("inf"),
                  kappa: float = 0.0) -> SubstrateSpec:
    c = [cell / 2.0] * 3
    return SubstrateSpec(a=cell, b=cell, c=cell, D_lat=D_ex, T2_lat=T2_ex,
                         types=[PRIM_SPHERE], centers=[c], radii=[R],
                         D=D_in, T2=T2_in, kappa=kappa)


def hex_lattice(R: float, spacing: float, nx: int, ny: int, D_in: float, D_ex: float,
                T2_in: float = float("inf"), T2_ex: float = float("inf"),
                kappa: float = 0.0, c: Optional[float] = None) -> SubstrateSpec:
    """Hexagonal packing in a rectangular periodic cell of nx x ny hex unit
    cells (width s, height s*sqrt(3)), two cylinders per unit cell."""
    s = float(spacing)
    h = s * math.sqrt(3.0)
    a, b = nx * s, ny * h
    cs = []
    for j in range(ny):
        for i in range(nx):
            x0, y0 = i * s + 0.25 * s, j * h + 0.25 * h
            cs.append([x0, y0, 0.0])
            cs.append([x0 + 0.5 * s, y0 + 0.5 * h, 0.0])
    n = len(cs)
    return SubstrateSpec(a=a, b=b, c=(c if c is not None else min(a, b)), D_lat=D_ex,
                         T2_lat=T2_ex, types=np.full(n, PRIM_CYLINDER_XY), centers=cs,
                         radii=np.full(n, R), D=D_in, T2=T2_in, kappa=kappa)


def parallel_planes(cell: float, lo: float, hi: float, D: float, T2: float = float("inf"),
                    axis: int = 0) -> SubstrateSpec:
    """Two reflecting planes normal to `axis` at lo and hi in a cubic periodic
    cell, normals pointing into the slab between them (planePhase.cu:114-117
    puts them at 0 and at the cell length)."""
    n = np.zeros((2, 3))
    n[0, axis], n[1, axis] = 1.0, -1.0
    cs = np.zeros((2, 3))
    cs[0, axis], cs[1, axis] = lo, hi
    return SubstrateSpec(a=cell, b=cell, c=cell, D_lat=D, T2_lat=T2, types=[PRIM_PLANE] * 2,
                         centers=cs, radii=[0.0, 0.0], normals=n)


_TYPE_NAMES = {PRIM_CYLINDER_XY: "cylinder_xy", PRIM_SPHERE: "sphere", PRIM_PLANE: "plane"}


def _num(v: float):
    v = float(v)
    return v if math.isfinite(v) else None


def save_json(sub: SubstrateSpec, path: str, provenance: Optional[dict] = None) -> None:
    """Write a substrate file (the format of include/axonexa_compat.hpp's
    write_substrate_json: cell, lattice D / T2, one record per primitive;
    an infinite T2 is null; floats round-trip exactly)."""
    import json

    prims = []
    for k in range(sub.n_prims):
        prims.append({
            "type": _TYPE_NAMES[int(sub.types[k])],
            "center": [float(x) for x in sub.centers[k]],
            "radius": float(sub.radii[k]),
            "D": float(sub.D[k]),
            "T2": _num(sub.T2[k]),
            "T1": 0.0,
            "kappa": float(sub.kappa[k]),
            "region": k,
            "normal": [float(x) for x in sub.normals[k]],
        })
    doc = {"format": "axonexa-substrate", "version": 1,
           "cell": [float(sub.a), float(sub.b), float(sub.c)],
           "lattice": {"D": float(sub.D_lat), "T2": _num(sub.T2_lat), "T1": 0.0},
           "primitives": prims, "provenance": provenance or {}}
    with open(path, "w") as f:
        json.dump(doc, f, indent=1, allow_nan=False)


def load_json(path: str) -> SubstrateSpec:
    """Read a substrate file written by save_json or by the C++ header."""
    import json

    with open(path) as f:
        doc = json.load(f)
    if doc.get("format") != "axonexa-substrate":
        raise ValueError(f"{path}: not an axonexa substrate file")
    names = {v: k for k, v in _TYPE_NAMES.items()}
    prims = doc.get("primitives", [])
    inf = float("inf")

    def real(v, dflt):
        return dflt if v is None else float(v)

    lat = doc.get("lattice", {})
    return SubstrateSpec(
        a=float(doc["cell"][0]), b=float(doc["cell"][1]), c=float(doc["cell"][2]),
        D_lat=real(lat.get("D"), 0.0), T2_lat=real(lat.get("T2"), inf),
        types=[names[p.get("type", "cylinder_xy")] for p in prims],
        centers=np.array([p["center"] for p in prims], np.float64).reshape(-1, 3),
        radii=[real(p.get("radius"), 0.0) for p in prims],
        D=np.array([real(p.get("D"), 0.0) for p in prims]),
        T2=np.array([real(p.get("T2"), inf) for p in prims]),
        kappa=np.array([real(p.get("kappa"), 0.0) for p in prims]),
        normals=np.array([p.get("normal", [0.0, 0.0, 0.0]) for p in prims], np.float64).reshape(-1, 3))


def hex_intra_fraction(R: float, spacing: float) -> float:
    return 2.0 * math.pi * R * R / (math.sqrt(3.0) * spacing * spacing)


def kappa_for_probability(p: float, D: float, dt: float) -> float:
    """Permeability giving transmission probability p per collision for a
    walker of diffusivity D: p = 4*kappa/v, v = sqrt(6 D / dt)
    (src/kernels/boundaryCheck.cuh:94)."""
    return p * math.sqrt(6.0 * D / dt) / 4.0


def sample_gamma(n: int, k: float, theta: float, rng: np.random.Generator,
                 rmin: float = 0.0, rmax: float = float("inf")) -> np.ndarray:
    """Truncated Gamma(k, theta) radii by rejection (RPSinitializer.h:76-101)."""
    out = []
    while len(out) < n:
        x = rng.gamma(k, theta, size=2 * n)
        x = x[(x > rmin) & (x < rmax)]
        out.extend(x.tolist())
    return np.sort(np.array(out[:n]))


def _min_image(d: np.ndarray, L: np.ndarray) -> np.ndarray:
    return d - L * np.round(d / L)


def pack_cylinders(radii: np.ndarray, box: float, rng: np.random.Generator,
                   tries: int = 2000, gap: float = 0.0, relax_iters: int = 4000):
    """Place discs of the given radii in a periodic square box without overlap.

    Random sequential addition from the largest radius down (as
    RPSinitializer.h:103-129), with the full minimum-image overlap test
    (fixes defect D8).  Discs that cannot be placed by RSA are inserted at the
    point of least overlap and the packing is relaxed by pairwise repulsion
    until no overlaps remain (needed above the RSA jamming limit).
    Returns centres [n, 2] or raises RuntimeError."""
    r = np.sort(np.asarray(radii, dtype=np.float64))[::-1]
    n = r.size
    L = np.array([box, box])
    pos = np.zeros((n, 2))
    placed = 0
    for i in range(n):
        ok = False
        for _ in range(tries):
            p = rng.random(2) * box
            if placed:
                d = _min_image(pos[:placed] - p, L)
                dist = np.hypot(d[:, 0], d[:, 1])
                if np.any(dist < r[:placed] + r[i] + gap):
                    continue
            ok = True
            break
        if not ok:
            # least-overlap candidate; fixed by relaxation below
            cand = rng.random((256, 2)) * box
            best, bo = None, None
            for p in cand:
                d = _min_image(pos[:placed] - p, L)
                ov = np.maximum(r[:placed] + r[i] + gap - np.hypot(d[:, 0], d[:, 1]), 0).sum()
                if bo is None or ov < bo:
                    best, bo = p, ov
            p = best
        pos[placed] = p
        placed += 1
    # repulsive relaxation (soft-disc gradient descent, deterministic given rng)
    for it in range(relax_iters):
        d = pos[:, None, :] - pos[None, :, :]
        d = _min_image(d, L)
        dist = np.hypot(d[..., 0], d[..., 1])
        np.fill_diagonal(dist, np.inf)
        ov = (r[:, None] + r[None, :] + gap) - dist
        mask = ov > 0
        if not mask.any():
            return pos % box, r
        with np.errstate(invalid="ignore"):
            f = np.where(mask, ov / np.maximum(dist, 1e-12), 0.0)
        pos += 0.5 * np.einsum("ij,ijk->ik", f, d) * 0.5
        pos %= box
    raise RuntimeError("packing did not converge to a non-overlapping state")


def gamma_packing(n: int, k: float, theta: float, fraction: float, D_in: float, D_ex: float,
                  seed: int = 1, T2_in: float = float("inf"), T2_ex: float = float("inf"),
                  kappa: float = 0.0, rmin: float = 0.0, gap: float = 0.0) -> SubstrateSpec:
    """Gamma-distributed cylinder radii packed to a target area fraction."""
    rng = np.random.default_rng(seed)
    radii = sample_gamma(n, k, theta, rng, rmin=rmin)
    box = math.sqrt(np.sum(math.pi * radii ** 2) / fraction)
    pos, r = pack_cylinders(radii, box, rng, gap=gap)
    cs = np.zeros((n, 3))
    cs[:, :2] = pos
    return SubstrateSpec(a=box, b=box, c=box, D_lat=D_ex, T2_lat=T2_ex,
                         types=np.full(n, PRIM_CYLINDER_XY), centers=cs, radii=r,
                         D=D_in, T2=T2_in, kappa=kappa)


def area_fraction(sub: SubstrateSpec) -> float:
    """Exact intra-cylinder area fraction (each cylinder counted once;
    replaces RPSinitializer::getVI, which de-duplicates by radius)."""
    cyl = sub.types == PRIM_CYLINDER_XY
    return float(np.sum(math.pi * sub.radii[cyl] ** 2) / (sub.a * sub.b))
