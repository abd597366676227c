"""The substrate file format (f1): Python save/load round trips exactly,
and the C++ header's reader/writer (include/axonexa_compat.hpp) reads what
Python writes and writes what Python reads back, bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from axonexa_amd import abi
from axonexa_amd import substrates as sb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CXX = r'''
#define SPECULAR_REFLECTION
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"
int main(int argc, char** argv) {
    Lattice l0(1, 1, 1, 0, 0, 0, 0);
    RPSLatticeInitializer<Cylinder_XY> init(l0, 0);
    init.readJSON(argv[1]);
    init.writeJSON(argv[2], "{\"tool\": \"roundtrip\"}");
    std::printf("%d\n", (int)init.basis.size());
    return 0;
}
'''


def _same(a: abi.SubstrateSpec, b: abi.SubstrateSpec):
    assert (a.a, a.b, a.c, a.D_lat, a.T2_lat) == (b.a, b.b, b.c, b.D_lat, b.T2_lat)
    for f in ("types", "centers", "radii", "D", "T2", "kappa", "normals"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def _subs():
    return [
        sb.gamma_packing(20, 5.0, 0.4, 0.6, D_in=1.7, D_ex=2.0, seed=3, T2_in=70.0, T2_ex=50.0,
                         kappa=0.01, rmin=0.3),
        sb.hex_lattice(2.5, 6.0, 2, 2, D_in=1.7, D_ex=2.0),
        sb.parallel_planes(10.0, 2.5, 7.5, 2.0),
        sb.single_sphere(16.0, 5.0, D_in=1.0, D_ex=2.5, T2_in=40.0),
    ]


def test_python_round_trip_is_exact(tmp_path):
    for k, sub in enumerate(_subs()):
        p = tmp_path / f"s{k}.json"
        sb.save_json(sub, str(p), provenance={"builder": "test", "k": k})
        _same(sub, sb.load_json(str(p)))


def test_cpp_header_reads_and_writes_the_same_file(tmp_path):
    src = tmp_path / "rt.cpp"
    src.write_text(CXX)
    exe = tmp_path / "rt"
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr[-2000:])
    for k, sub in enumerate(_subs()):
        a, b = tmp_path / f"a{k}.json", tmp_path / f"b{k}.json"
        sb.save_json(sub, str(a))
        out = subprocess.run([str(exe), str(a), str(b)], capture_output=True, text=True, check=True)
        assert int(out.stdout) == sub.n_prims
        _same(sub, sb.load_json(str(b)))


def test_rejects_other_files(tmp_path):
    p = tmp_path / "x.json"
    p.write_text('{"format": "something-else"}')
    with pytest.raises(ValueError):
        sb.load_json(str(p))
