// Host-only checks of include/axonexa_compat.hpp that need no GPU and no
// library: the substrate-file JSON parser and reader (valid documents,
// every truncation of one, malformed and hostile text, which must throw
// std::runtime_error and never read out of bounds) and
// RPSLatticeInitializer<B> (src/substrate/RPSinitializer.h: radius
// distributions, overlap-free random sequential packing under the minimum
// image, volume fraction, JSON round trip).  tests/test_sanitizers.py builds
// it with AddressSanitizer and UndefinedBehaviorSanitizer and runs it; it
// prints one line per group and exits non-zero if any check failed.
//
//   compat_check <scratch directory>
#define SPECULAR_REFLECTION
typedef double real;
#include "axonexa_compat.hpp"

#include <fstream>

namespace {

int g_checks = 0, g_fails = 0;

#define CHECK(cond)                                                                    \
    do {                                                                               \
        ++g_checks;                                                                    \
        if (!(cond)) {                                                                 \
            ++g_fails;                                                                 \
            std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
        }                                                                              \
    } while (0)

bool parse_throws(const std::string& text) {
    try {
        axx_compat::JsonParser(text).parse();
    } catch (const std::runtime_error&) {
        return true;
    }
    return false;
}

void write_file(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    f << text;
}

bool read_throws(const std::string& path) {
    Lattice lat;
    std::vector<axx_compat::Prim> prims;
    try {
        axx_compat::read_substrate_json(path, lat, prims);
    } catch (const std::runtime_error&) {
        return true;
    }
    return false;
}

const char* kDoc =
    "{\"format\": \"axonexa-substrate\", \"version\": 1,\n"
    " \"cell\": [20, 17.320508075688775, 1],\n"
    " \"lattice\": {\"D\": 2.5, \"T2\": null, \"T1\": 0},\n"
    " \"primitives\": [\n"
    "  {\"type\": \"cylinder_xy\", \"center\": [5, 4.25, 0], \"radius\": 2.5, \"D\": 1.7, \"T2\": 70,"
    " \"T1\": 0, \"kappa\": 0.01, \"region\": 1, \"normal\": [0, 0, 0]},\n"
    "  {\"type\": \"sphere\", \"center\": [10, 10, 0.5], \"radius\": 0.25, \"D\": 1, \"T2\": null,"
    " \"T1\": 0, \"kappa\": 0, \"region\": 2},\n"
    "  {\"type\": \"plane\", \"center\": [0, 0, 0.5], \"radius\": 0, \"normal\": [0, 0, 1]}\n"
    " ],\n"
    " \"provenance\": {\"tool\": \"compat_check\", \"list\": [1, [2, [3]], {\"x\": \"a\\\"b\\\\c\"}]}}\n";

void parser_checks(const std::string& dir) {
    const std::string doc = kDoc;
    const axx_compat::Json v = axx_compat::JsonParser(doc).parse();
    CHECK(v.kind == axx_compat::Json::OBJ);
    CHECK(v.get("version") && v.get("version")->num == 1.0);
    CHECK(v.get("cell")->arr.size() == 3 && v.get("cell")->arr[1].num == 17.320508075688775);
    CHECK(v.get("lattice")->get("T2")->kind == axx_compat::Json::NUL);
    CHECK(v.get("primitives")->arr.size() == 3);
    const axx_compat::Json* prov = v.get("provenance");
    CHECK(prov && prov->get("list")->arr[2].get("x")->str == "a\"b\\c");
    CHECK(!v.get("missing"));

    // every proper prefix of the document is malformed
    int prefixes = 0;
    for (size_t n = 0; n < doc.size(); ++n) {
        // a prefix that ends inside trailing whitespace after the last brace
        // does not exist: the document ends with "}\n"
        if (n == doc.size() - 1) continue;   // "...}" without the newline is complete
        CHECK(parse_throws(doc.substr(0, n)));
        ++prefixes;
    }
    CHECK(!parse_throws(doc.substr(0, doc.size() - 1)));

    const char* bad[] = {
        "", " ", "{", "}", "[", "]", "{\"a\" 1}", "{\"a\": }", "{\"a\": 1,}", "{,}", "[1 2]", "[1,]", "[,]",
        "\"abc", "\"ab\\", "tru", "nul", "fals", "-", "--1", "1e", "{}x", "[]]", "{\"a\": [1, 2}", "{1: 2}",
        "{\"a\": 1 \"b\": 2}", "\"\\", "[\"a\", ]", "{\"a\"}", "01x",
    };
    for (const char* b : bad) CHECK(parse_throws(b));
    // nesting within the limit parses; beyond it throws instead of recursing on
    std::string deep_ok(axx_compat::JsonParser::kMaxDepth, '['), deep_bad(100000, '[');
    deep_ok += std::string(axx_compat::JsonParser::kMaxDepth, ']');
    CHECK(!parse_throws(deep_ok));
    CHECK(parse_throws(deep_bad));
    std::string deep_obj;
    for (int k = 0; k < 5000; ++k) deep_obj += "{\"a\": ";
    CHECK(parse_throws(deep_obj));

    // the reader: the document itself, then files it must refuse
    const std::string good = dir + "/good.json";
    write_file(good, doc);
    Lattice lat;
    std::vector<axx_compat::Prim> prims;
    axx_compat::read_substrate_json(good, lat, prims);
    CHECK(prims.size() == 3 && lat.getBasisSize() == 3);
    CHECK(lat.a == 20.0 && lat.b == 17.320508075688775 && std::isinf(lat.T2) && lat.D == 2.5);
    CHECK(prims[0].type == AXX_PRIM_CYLINDER_XY && prims[0].R == 2.5 && prims[0].kappa == 0.01 && prims[0].region == 1);
    CHECK(prims[1].type == AXX_PRIM_SPHERE && std::isinf(prims[1].T2) && prims[1].cz == 0.5);
    CHECK(prims[2].type == AXX_PRIM_PLANE && prims[2].nz == 1.0);

    const std::string cases[] = {
        "{\"format\": \"other\", \"cell\": [1, 1, 1]}",
        "{\"cell\": [1, 1, 1]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, \"x\", 1]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": 3}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": 7}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": [{\"radius\": 1}]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": [{\"center\": [0, 0]}]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": "
        "[{\"center\": [0, 0, 0], \"radius\": \"big\"}]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": "
        "[{\"center\": [0, 0, 0], \"region\": 1e300}]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": "
        "[{\"center\": [0, 0, 0], \"region\": -3}]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": "
        "[{\"center\": [0, 0, 0], \"region\": 1.5}]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1], \"primitives\": "
        "[{\"center\": [0, 0, 0], \"normal\": [1, 0]}]}",
        "{\"format\": \"axonexa-substrate\", \"cell\": [1, 1, 1]",
    };
    int k = 0;
    for (const std::string& c : cases) {
        const std::string path = dir + "/bad" + std::to_string(k++) + ".json";
        write_file(path, c);
        CHECK(read_throws(path));
    }
    CHECK(read_throws(dir + "/no_such_file.json"));
    std::printf("parser: %d prefixes, %zu malformed texts, %d refused files\n", prefixes,
                sizeof(bad) / sizeof(bad[0]), k + 1);
}

template <class B>
bool any_overlap(const RPSLatticeInitializer<B>& init, bool three_d) {
    const Lattice& L = init.lat;
    for (size_t i = 0; i < init.basis.size(); ++i)
        for (size_t j = i + 1; j < init.basis.size(); ++j) {
            const Vector3 ci = init.basis[i].getCenter(), cj = init.basis[j].getCenter();
            double d[3] = {ci.x - cj.x, ci.y - cj.y, three_d ? ci.z - cj.z : 0.0};
            const double len[3] = {L.getA(), L.getB(), L.getC()};
            double r2 = 0;
            for (int a = 0; a < 3; ++a) {
                d[a] -= len[a] * std::nearbyint(d[a] / len[a]);
                r2 += d[a] * d[a];
            }
            const double rr = init.basis[i].getRadius() + init.basis[j].getRadius();
            if (r2 < rr * rr) return true;
        }
    return false;
}

template <class B>
bool in_cell(const RPSLatticeInitializer<B>& init, bool three_d) {
    for (const B& b : init.basis) {
        const Vector3 c = b.getCenter();
        if (c.x < 0 || c.x > init.lat.getA() || c.y < 0 || c.y > init.lat.getB()) return false;
        if (three_d && (c.z < 0 || c.z > init.lat.getC())) return false;
    }
    return true;
}

void initializer_checks(const std::string& dir) {
    // gamma radii (RPSinitializer.h:76-101: alpha = shape, beta = rate), packed
    Lattice lat(40.0, 40.0, 1.0, 70.0, 0.0, 2.0, 40);
    RPSLatticeInitializer<Cylinder_XY> g(lat, 0);
    g.gammaRadialDist(7, 5.0, 1.0 / 0.4, 0.3, 4.0);
    bool sorted_in_range = true;
    for (size_t i = 0; i < g.basis.size(); ++i) {
        const double r = g.basis[i].getRadius();
        sorted_in_range &= r > 0.3 && r < 4.0 && (i == 0 || g.basis[i - 1].getRadius() <= r);
    }
    CHECK(g.basis.size() == 40 && sorted_in_range);
    CHECK(g.uniformCenterDist(11, 2000));
    CHECK(in_cell(g, false) && !any_overlap(g, false));
    double vi = 0;
    for (const auto& b : g.basis) vi += PI * b.getRadius() * b.getRadius();
    vi /= 40.0 * 40.0;
    CHECK(std::fabs(g.getVI() - vi) <= 1e-15 * vi && vi > 0.1 && vi < 0.9);
    g.setRegions();
    CHECK(g.basis[39].getRegion() == 39);

    // the packing through a file and back, bit for bit
    const std::string path = dir + "/packing.json";
    g.writeJSON(path, "{\"tool\": \"compat_check\"}");
    Lattice l0(1, 1, 1, 0, 0, 0, 0);
    RPSLatticeInitializer<Cylinder_XY> h(l0, 0);
    h.readJSON(path);
    bool same = h.basis.size() == g.basis.size() && h.lat.getA() == 40.0 && h.lat.T2 == 70.0;
    for (size_t i = 0; same && i < g.basis.size(); ++i) {
        const axx_compat::Prim &a = g.basis[i].p, &b = h.basis[i].p;
        same = a.type == b.type && a.cx == b.cx && a.cy == b.cy && a.cz == b.cz && a.R == b.R && a.D == b.D &&
               a.T2 == b.T2 && a.T1 == b.T1 && a.kappa == b.kappa && a.region == b.region && a.nx == b.nx &&
               a.ny == b.ny && a.nz == b.nz;
    }
    CHECK(same);

    // uniform and Gaussian radii; spheres packed in 3-d
    Lattice l3(30.0, 30.0, 30.0, 50.0, 0.0, 2.0, 40);
    RPSLatticeInitializer<Sphere> s(l3, 1);
    s.uniformRadialDist(3, 2.5, 1.0);
    bool uni = true;
    for (const auto& b : s.basis) uni &= b.getRadius() >= 1.0 && b.getRadius() <= 2.5;
    CHECK(uni);
    CHECK(s.uniformCenterDist(5, 2000));
    CHECK(in_cell(s, true) && !any_overlap(s, true));
    s.gaussianRadialDist(9, 1.5, 0.3);
    bool pos = true;
    for (const auto& b : s.basis) pos &= b.getRadius() > 0;
    CHECK(pos);

    // an impossible packing gives up (returns false) instead of looping on
    Lattice tight(4.0, 4.0, 1.0, 70.0, 0.0, 2.0, 20);
    RPSLatticeInitializer<Cylinder_XY> t(tight, 0);
    t.uniformRadialDist(1, 1.5, 1.4);
    CHECK(!t.uniformCenterDist(2, 50));
    std::printf("initializer: 40 gamma cylinders at %.3f, 40 spheres, file round trip, refused packing\n", vi);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    try {
        parser_checks(dir);
        initializer_checks(dir);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 2;
    }
    std::printf("compat_check: %d checks, %d failed\n", g_checks, g_fails);
    return g_fails ? 1 : 0;
}
