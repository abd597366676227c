"""include/axonexa_compat.hpp on the CPU: periodic copies of primitives that
a reference driver lists explicitly are passed once (the engine builds the
images itself), with the basis-to-primitive map that region indices use."""
import os
import subprocess
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#define GAMMA 267500.0
typedef double real;
#include "axonexa_compat.hpp"
int main() {
    const real a = 2.0;
    // the 14 spheres of a face-centred cubic cell: corners and face centres
    const real pos[14][3] = {{0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, {0, 1, 0}, {1, 1, 0}, {1, 1, 1},
                             {0, 1, 1}, {.5, 0, .5}, {1, .5, .5}, {.5, 1, .5}, {0, .5, .5}, {.5, .5, 1},
                             {.5, .5, 0}};
    std::vector<Sphere> s(14);
    for (int k = 0; k < 14; ++k) s[k] = Sphere(pos[k][0] * a, pos[k][1] * a, pos[k][2] * a, 0.3, 0, 0, 1, k + 1, 0);
    Lattice lat(a, a, a, 0, 0, 1, 14);
    std::vector<int> map;
    auto u = axx_compat::unique_modulo_cell(&s[0], lat, map);
    std::printf("%zu", u.size());
    for (int m : map) std::printf(" %d", m);
    std::printf("\n");
    // cylinders: z is ignored; a different radius is a different primitive
    std::vector<Cylinder_XY> c(4);
    c[0] = Cylinder_XY(0, 0, 0.2, 0, 0, 1, 1, 0);
    c[1] = Cylinder_XY(a, a, 0.2, 0, 0, 1, 2, 0);
    c[2] = Cylinder_XY(a, 0, 0.25, 0, 0, 1, 3, 0);
    c[3] = Cylinder_XY(1.0, 1.0, 0.2, 0, 0, 1, 4, 0);
    Lattice l2(a, a, a, 0, 0, 1, 4);
    auto u2 = axx_compat::unique_modulo_cell(&c[0], l2, map);
    std::printf("%zu", u2.size());
    for (int m : map) std::printf(" %d", m);
    std::printf("\n");
    // planes: x = 0 and x = a (opposite normals) are one plane; x = a/2 is another
    std::vector<Plane> pl(3);
    pl[0] = Plane(Vector3(0, 0, 0), Vector3(1, 0, 0));
    pl[1] = Plane(Vector3(a, 0.3, 0), Vector3(-1, 0, 0));
    pl[2] = Plane(Vector3(a / 2, 0, 0), Vector3(1, 0, 0));
    Lattice l3(a, a, a, 0, 0, 1, 3);
    auto u3 = axx_compat::unique_modulo_cell(&pl[0], l3, map);
    std::printf("%zu", u3.size());
    for (int m : map) std::printf(" %d", m);
    std::printf("\n");
    return 0;
}
'''


def test_unique_modulo_cell():
    with tempfile.TemporaryDirectory() as d:
        cf = os.path.join(d, "t.cpp")
        open(cf, "w").write(SRC)
        exe = os.path.join(d, "t")
        # header-only part: no engine call, so no library to link
        subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), cf, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    fcc = [int(x) for x in out[0].split()]
    # four distinct spheres: the corner and three face centres
    assert fcc[0] == 4
    assert fcc[1:] == [0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 1, 2, 3, 3]
    cyl = [int(x) for x in out[1].split()]
    assert cyl == [3, 0, 0, 1, 2]
    pl = [int(x) for x in out[2].split()]
    assert pl == [2, 0, 0, 1]
