// axonexa_compat.hpp — the reference's C++ driver interface over the C ABI.
//
// Drivers written against the reference's headers
//   Vector3 (src/util/vector3.cuh), Cylinder_XY (src/primitives/cylinderXY.cuh),
//   Sphere (src/primitives/Sphere.cuh), Lattice (src/substrate/lattice.cuh),
//   RectGrad / CosGFunc / SinGFunc / ApodCosGFunc / SWOGSEFunc
//   (src/functors/gfunctors.cuh), phaseAcquisition<F>
//   (src/acquisition/phaseAcquisition.cuh), phaseAcquisitionStream<F>
//   (src/acquisition/phaseAcquisitionStream.cuh), RPSLatticeInitializer<B>
//   (src/substrate/RPSinitializer.h)
// compile against this header instead: same class names, constructors and
// methods.  run*() methods sample the functors on the host at t_i = i*dt,
// factor them into (unit profile, amplitude, direction), and call
// axx_run_phase (include/axonexa.h) on device 0; blocks/threads/numOfSM are
// accepted and ignored (the library picks its launch shape).
//
// As in the reference, the driver defines `real` (typedef) and may define
// GAMMA, PI, SPECULAR_REFLECTION, USE_PERMEABLE, USE_RELAXATION and
// USE_INITIALIZE_IN_REGION before including this header.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "axonexa.h"

#ifndef GAMMA
#define GAMMA 267500.0
#endif
#ifndef PI
#define PI 3.1415926535897932384626433832795
#endif

// ---------------------------------------------------------------- Vector3
class Vector3 {
  public:
    real x, y, z;
    Vector3() : x(0), y(0), z(0) {}
    Vector3(real x_, real y_, real z_) : x(x_), y(y_), z(z_) {}
    real magnitude() const { return std::sqrt(x * x + y * y + z * z); }
    real squareMagnitude() const { return x * x + y * y + z * z; }
    void normalize() {
        const real l = magnitude();
        x /= l; y /= l; z /= l;
    }
    void invert() { x = -x; y = -y; z = -z; }
    Vector3 operator+(const Vector3& o) const { return Vector3(x + o.x, y + o.y, z + o.z); }
    Vector3 operator-(const Vector3& o) const { return Vector3(x - o.x, y - o.y, z - o.z); }
    Vector3 operator*(real s) const { return Vector3(x * s, y * s, z * s); }
    Vector3 operator/(real s) const { return Vector3(x / s, y / s, z / s); }
    real operator*(const Vector3& o) const { return x * o.x + y * o.y + z * o.z; }      // dot
    Vector3 operator%(const Vector3& o) const {                                        // cross
        return Vector3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x);
    }
    void operator+=(const Vector3& o) { x += o.x; y += o.y; z += o.z; }
    void operator-=(const Vector3& o) { x -= o.x; y -= o.y; z -= o.z; }
    void operator*=(real s) { x *= s; y *= s; z *= s; }
};
inline std::ostream& operator<<(std::ostream& os, const Vector3& v) {
    return os << v.x << " " << v.y << " " << v.z;
}

// -------------------------------------------------------------- primitives
namespace axx_compat {
struct Prim {
    int type = AXX_PRIM_CYLINDER_XY;
    real cx = 0, cy = 0, cz = 0, R = 0, D = 0, T2 = 0, T1 = 0, kappa = 0, eps = 0;
    real nx = 0, ny = 0, nz = 0;   // planes: normal
    int region = 0;
};
}  // namespace axx_compat

class Cylinder_XY {
  public:
    axx_compat::Prim p;
    Cylinder_XY() {}
    Cylinder_XY(real cx, real cy, real R, real T2, real T1, real D, int region, real perm, real eps = 0) {
        p.type = AXX_PRIM_CYLINDER_XY;
        p.cx = cx; p.cy = cy; p.R = R; p.T2 = T2; p.T1 = T1; p.D = D;
        p.region = region; p.kappa = perm; p.eps = eps;
    }
    bool inside(const Vector3& r) const {
        return (r.x - p.cx) * (r.x - p.cx) + (r.y - p.cy) * (r.y - p.cy) < p.R * p.R;
    }
    Vector3 getCenter() const { return Vector3(p.cx, p.cy, 0.0); }
    void setCenter(Vector3 c) { p.cx = c.x; p.cy = c.y; }
    real getRadius() const { return p.R; }
    void setRadius(real r) { p.R = r; }
    void setEPS(real e) { p.eps = e; }
    void setRegion(int r) { p.region = r; }
    int getRegion() const { return p.region; }
    real getD() const { return p.D; }
    real getT2() const { return p.T2; }
    real getT1() const { return p.T1; }
    real getPermeability() const { return p.kappa; }
};

class Sphere {
  public:
    axx_compat::Prim p;
    Sphere() { p.type = AXX_PRIM_SPHERE; }
    Sphere(real cx, real cy, real cz, real R, real T2, real T1, real D, int region, real perm, real eps = 0) {
        p.type = AXX_PRIM_SPHERE;
        p.cx = cx; p.cy = cy; p.cz = cz; p.R = R; p.T2 = T2; p.T1 = T1; p.D = D;
        p.region = region; p.kappa = perm; p.eps = eps;
    }
    bool inside(const Vector3& r) const {
        return (r - getCenter()).squareMagnitude() < p.R * p.R;
    }
    Vector3 getCenter() const { return Vector3(p.cx, p.cy, p.cz); }
    void setCenter(Vector3 c) { p.cx = c.x; p.cy = c.y; p.cz = c.z; }
    real getRadius() const { return p.R; }
    void setRadius(real r) { p.R = r; }
    void setEPS(real e) { p.eps = e; }
    void setRegion(int r) { p.region = r; }
    int getRegion() const { return p.region; }
    real getD() const { return p.D; }
    real getT2() const { return p.T2; }
    real getPermeability() const { return p.kappa; }
};

// Infinite reflecting plane through `center` with normal `normal`
// (src/primitives/plane.cuh:1-40): bounds no region, never transmits.
class Plane {
  public:
    axx_compat::Prim p;
    Plane() { p.type = AXX_PRIM_PLANE; p.nx = 1; }
    Plane(Vector3 center, Vector3 normal) {
        p.type = AXX_PRIM_PLANE;
        p.cx = center.x; p.cy = center.y; p.cz = center.z;
        p.nx = normal.x; p.ny = normal.y; p.nz = normal.z;
    }
    Vector3 getNormal(const Vector3&) const { return Vector3(p.nx, p.ny, p.nz); }
    int getRegion(const Vector3&) const { return 0; }
    real getT2(const Vector3&) const { return -1.0; }
    real getD(const Vector3&) const { return -1.0; }
    real getPermeability() const { return 0.0; }
};

class Lattice {
  public:
    real a = 0, b = 0, c = 0, T2 = 0, T1 = 0, D = 0;
    int basisSize = 0;
    Lattice() {}
    Lattice(real a_, real b_, real c_, real T2_, real T1_, real D_, int n)
        : a(a_), b(b_), c(c_), T2(T2_), T1(T1_), D(D_), basisSize(n) {}
    real getA() const { return a; }
    real getB() const { return b; }
    real getC() const { return c; }
    int getBasisSize() const { return basisSize; }
    void setBasisSize(int n) { basisSize = n; }
    bool inLatticeCell(const Vector3& r) const {
        return !(r.x > a || r.x < 0 || r.y > b || r.y < 0 || r.z > c || r.z < 0);
    }
};

// ------------------------------------------------ substrate file (JSON)
// {"format": "axonexa-substrate", "version": 1, "cell": [a, b, c],
//  "lattice": {"D": D, "T2": T2, "T1": T1},
//  "primitives": [{"type": "cylinder_xy" | "sphere" | "plane",
//                  "center": [x, y, z], "radius": R, "D": D, "T2": T2, "T1": T1,
//                  "kappa": k, "region": r, "normal": [nx, ny, nz]}, ...],
//  "provenance": {...}}
// An infinite T2 is written as null.  Numbers carry 17 significant digits,
// so a file round-trips doubles exactly.  The same format is read and
// written by axonexa_amd/substrates.py (save_json / load_json).
namespace axx_compat {
struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;
    const Json* get(const std::string& k) const {
        for (auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

// Recursive descent over the whole text; malformed text throws
// std::runtime_error (never reads past the end).  Nesting deeper than
// kMaxDepth throws as well, so a hostile file cannot exhaust the stack.
class JsonParser {
  public:
    static constexpr int kMaxDepth = 64;
    explicit JsonParser(const std::string& t) : s(t) {}
    Json parse() {
        Json v = value();
        ws();
        if (i != s.size()) fail("trailing characters");
        return v;
    }

  private:
    const std::string& s;
    size_t i = 0;
    int depth = 0;
    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("substrate JSON: ") + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\t' || s[i] == '\r')) ++i;
    }
    bool lit(const char* w) {
        const size_t n = std::strlen(w);
        if (s.compare(i, n, w) == 0) { i += n; return true; }
        return false;
    }
    std::string string_() {
        if (peek() != '"') fail("expected a string");
        ++i;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\') {
                ++i;
                if (i >= s.size()) fail("bad escape");
                const char c = s[i];
                out += c == 'n' ? '\n' : c == 't' ? '\t' : c;
            } else {
                out += s[i];
            }
            ++i;
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    char peek() const {
        if (i >= s.size()) fail("unexpected end");
        return s[i];
    }
    struct Nest {
        JsonParser& p;
        explicit Nest(JsonParser& q) : p(q) {
            if (++p.depth > kMaxDepth) p.fail("nesting too deep");
        }
        ~Nest() { --p.depth; }
    };
    Json value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        Json v;
        const char c = s[i];
        if (c == '{' || c == '[') {
            Nest guard(*this);
            return c == '{' ? object() : array();
        }
        if (c == '"') { v.kind = Json::STR; v.str = string_(); return v; }
        if (lit("null")) return v;
        if (lit("true")) { v.kind = Json::BOOL; v.b = true; return v; }
        if (lit("false")) { v.kind = Json::BOOL; return v; }
        char* end = nullptr;
        v.num = std::strtod(s.c_str() + i, &end);
        if (end == s.c_str() + i) fail("bad value");
        i = (size_t)(end - s.c_str());
        v.kind = Json::NUM;
        return v;
    }
    Json object() {
        Json v;
        v.kind = Json::OBJ;
        ++i;
        ws();
        if (peek() == '}') { ++i; return v; }
        for (;;) {
            ws();
            std::string k = string_();
            ws();
            if (peek() != ':') fail("expected ':'");
            ++i;
            v.obj.emplace_back(k, value());
            ws();
            if (peek() == ',') { ++i; continue; }
            if (peek() == '}') { ++i; return v; }
            fail("expected ',' or '}'");
        }
    }
    Json array() {
        Json v;
        v.kind = Json::ARR;
        ++i;
        ws();
        if (peek() == ']') { ++i; return v; }
        for (;;) {
            v.arr.push_back(value());
            ws();
            if (peek() == ',') { ++i; continue; }
            if (peek() == ']') { ++i; return v; }
            fail("expected ',' or ']'");
        }
    }
};

inline std::string json_num(double v) {
    if (!std::isfinite(v)) return "null";
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

inline double json_real(const Json* v, double dflt) {
    if (!v || v->kind == Json::NUL) return dflt;
    if (v->kind != Json::NUM) throw std::runtime_error("substrate JSON: expected a number");
    return v->num;
}

// The three numbers of an array member ("cell", "center", "normal").
inline void json_vec3(const Json* v, double out[3], const std::string& what) {
    if (!v || v->kind != Json::ARR || v->arr.size() != 3) throw std::runtime_error(what + " must hold 3 numbers");
    for (int k = 0; k < 3; ++k) {
        if (v->arr[k].kind != Json::NUM) throw std::runtime_error(what + " must hold 3 numbers");
        out[k] = v->arr[k].num;
    }
}

inline void write_substrate_json(const std::string& path, const Lattice& lat, const std::vector<Prim>& prims,
                                 const std::string& provenance_json = "{}") {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fprintf(f, "{\"format\": \"axonexa-substrate\", \"version\": 1,\n \"cell\": [%s, %s, %s],\n",
                 json_num(lat.a).c_str(), json_num(lat.b).c_str(), json_num(lat.c).c_str());
    std::fprintf(f, " \"lattice\": {\"D\": %s, \"T2\": %s, \"T1\": %s},\n \"primitives\": [",
                 json_num(lat.D).c_str(), json_num(lat.T2).c_str(), json_num(lat.T1).c_str());
    for (size_t k = 0; k < prims.size(); ++k) {
        const Prim& p = prims[k];
        const char* type = p.type == AXX_PRIM_SPHERE ? "sphere" : p.type == AXX_PRIM_PLANE ? "plane" : "cylinder_xy";
        std::fprintf(f, "%s\n  {\"type\": \"%s\", \"center\": [%s, %s, %s], \"radius\": %s, \"D\": %s, "
                        "\"T2\": %s, \"T1\": %s, \"kappa\": %s, \"region\": %d, \"normal\": [%s, %s, %s]}",
                     k ? "," : "", type, json_num(p.cx).c_str(), json_num(p.cy).c_str(), json_num(p.cz).c_str(),
                     json_num(p.R).c_str(), json_num(p.D).c_str(), json_num(p.T2).c_str(), json_num(p.T1).c_str(),
                     json_num(p.kappa).c_str(), p.region, json_num(p.nx).c_str(), json_num(p.ny).c_str(),
                     json_num(p.nz).c_str());
    }
    std::fprintf(f, "\n ],\n \"provenance\": %s}\n", provenance_json.c_str());
    std::fclose(f);
}

inline void read_substrate_json(const std::string& path, Lattice& lat, std::vector<Prim>& prims) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw std::runtime_error("cannot read " + path);
    std::string text;
    char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, n);
    std::fclose(f);
    const Json root = JsonParser(text).parse();
    const Json* fmt = root.get("format");
    if (!fmt || fmt->str != "axonexa-substrate") throw std::runtime_error(path + ": not an axonexa substrate file");
    double cell[3];
    json_vec3(root.get("cell"), cell, path + ": cell");
    const Json* L = root.get("lattice");
    const double inf = std::numeric_limits<double>::infinity();
    lat = Lattice(cell[0], cell[1], cell[2], L ? json_real(L->get("T2"), inf) : inf,
                  L ? json_real(L->get("T1"), 0.0) : 0.0, L ? json_real(L->get("D"), 0.0) : 0.0, 0);
    prims.clear();
    const Json* P = root.get("primitives");
    if (P && P->kind != Json::ARR) throw std::runtime_error(path + ": primitives must be an array");
    if (P)
        for (const Json& e : P->arr) {
            Prim p;
            const Json* t = e.get("type");
            const std::string ty = t ? t->str : "cylinder_xy";
            p.type = ty == "sphere" ? AXX_PRIM_SPHERE : ty == "plane" ? AXX_PRIM_PLANE : AXX_PRIM_CYLINDER_XY;
            double c[3];
            json_vec3(e.get("center"), c, path + ": primitive center");
            p.cx = c[0]; p.cy = c[1]; p.cz = c[2];
            p.R = json_real(e.get("radius"), 0.0);
            p.D = json_real(e.get("D"), 0.0);
            p.T2 = json_real(e.get("T2"), inf);
            p.T1 = json_real(e.get("T1"), 0.0);
            p.kappa = json_real(e.get("kappa"), 0.0);
            const double region = json_real(e.get("region"), 0.0);
            if (!(region >= 0 && region <= 1e9) || region != std::floor(region))
                throw std::runtime_error(path + ": region must be a non-negative integer");
            p.region = (int)region;
            if (const Json* nrm = e.get("normal")) {
                double n3[3];
                json_vec3(nrm, n3, path + ": primitive normal");
                p.nx = n3[0]; p.ny = n3[1]; p.nz = n3[2];
            }
            prims.push_back(p);
        }
    lat.setBasisSize((int)prims.size());
}
}  // namespace axx_compat

// --------------------------------------------------------------- functors
// Two-lobe waveform G(t) = G m(t) ghat on [0, T), -G m(t - gap - T) ghat on
// [gap + T, 2T + gap) (gfunctors.cuh:36-54, 94-107).
namespace axx_compat {
class TwoLobe {
  public:
    real G = 0, T = 0, gap = 0, N = 1;
    Vector3 ghat = Vector3(1, 0, 0);
    int kind = 0;
    virtual ~TwoLobe() {}
    virtual double mod(double t) const = 0;
    double unit(double t) const {
        if (t < T) return mod(t);
        if (t >= gap + T && t < 2.0 * T + gap) return -mod(t - gap - T);
        return 0.0;
    }
    Vector3 operator()(real t) const {
        return ghat * (G * (real)unit(t));
    }
    std::tuple<int, double, double, double> key() const { return std::make_tuple(kind, (double)T, (double)gap, (double)N); }
    real getG() const { return G; }
};
}  // namespace axx_compat

class RectGrad : public axx_compat::TwoLobe {
  public:
    RectGrad() { kind = 0; }
    RectGrad(real g, real gd, real gs) { kind = 0; G = g; T = gd; gap = gs; }
    // b-value form (gfunctors.cuh:14-23)
    RectGrad(real bval, real gd, real gs, Vector3& gh) {
        kind = 0; T = gd; gap = gs; ghat = gh;
        G = std::sqrt(bval / (gd * gd * GAMMA * GAMMA * ((2.0 / 3.0) * gd + gs)));
    }
    double mod(double) const override { return 1.0; }
    real getB() const { return G * G * T * T * GAMMA * GAMMA * ((2.0 / 3.0) * T + gap); }
    real getDiffusionTime() const { return gap + T; }
    real getTau() const { return T + gap * 0.5; }
    real getFreq() const { return 0; }
};

class CosGFunc : public axx_compat::TwoLobe {
  public:
    CosGFunc() { kind = 1; }
    CosGFunc(real g, real gd, real gs, real nper, const Vector3& gh) {
        kind = 1; G = g; T = gd; gap = gs; N = nper; ghat = gh;
    }
    double mod(double t) const override { return std::cos(2.0 * PI * t * N / T); }
    real getFreq() const { return N / T; }
    real getB() const { return GAMMA * GAMMA * 0.25 * G * G * T * T * T / (PI * PI * N * N); }
};

class SinGFunc : public axx_compat::TwoLobe {
  public:
    SinGFunc() { kind = 2; }
    SinGFunc(real g, real gd, real gs, real nper, const Vector3& gh) {
        kind = 2; G = g; T = gd; gap = gs; N = nper; ghat = gh;
    }
    double mod(double t) const override { return std::sin(2.0 * PI * t * N / T); }
    real getFreq() const { return N / T; }
    real getB() const { return GAMMA * GAMMA * 0.75 * G * G * T * T * T / (PI * PI * N * N); }
};

class ApodCosGFunc : public axx_compat::TwoLobe {
  public:
    ApodCosGFunc() { kind = 3; }
    ApodCosGFunc(real g, real gd, real gs, real nper, const Vector3& gh) {
        kind = 3; G = g; T = gd; gap = gs; N = nper; ghat = gh;
    }
    double mod(double t) const override {
        const double period = T / N;
        if (t >= period / 4.0 && t < T - period / 4.0) return std::cos(2.0 * PI * t / period);
        return std::sin(4.0 * PI * t / period);
    }
    real getFreq() const { return N / T; }
    real getB() const {
        return GAMMA * GAMMA * 0.25 * G * G * T * T * T / (PI * PI * N * N) * (1.0 - 1.0 / (8.0 * N));
    }
};

class SWOGSEFunc : public axx_compat::TwoLobe {
  public:
    SWOGSEFunc() { kind = 4; }
    SWOGSEFunc(real g, real gd, real gs, real nper, const Vector3& gh) {
        kind = 4; G = g; T = gd; gap = gs; N = nper; ghat = gh;
    }
    // (-1)^floor(2 v t); the reference's pow(-1, 2 v t) is NaN off the grid
    double mod(double t) const override {
        const long long k = (long long)std::floor(2.0 * (N / T) * t);
        return (k % 2 == 0) ? 1.0 : -1.0;
    }
    real getFreq() const { return N / T; }
    real getB() const { const double v = N / T; return G * G * GAMMA * GAMMA * T / (6.0 * v * v); }
};

// ------------------------------------------------------------ acquisitions
inline double compat_linear_regression(const std::vector<real>& y, const std::vector<real>& x) {
    double n = (double)y.size(), sxy = 0, sx = 0, sy = 0, sxx = 0;
    for (size_t i = 0; i < y.size(); ++i) {
        sxy += x[i] * y[i]; sx += x[i]; sy += y[i]; sxx += x[i] * x[i];
    }
    return (n * sxy - sx * sy) / (n * sxx - sx * sx);
}

template <class F>
class phaseAcquisition {
  public:
    int numOfMeasurements = 0, numOfSteps = 0, seed = 0, last = 0, points = 0;
    std::vector<real> mx, my;
    std::vector<F> functors;
    real ADC = 0, ADCmeansquare = 0;

    phaseAcquisition() {}
    phaseAcquisition(int nMeas, int nSteps, int nParticles, int seed_)
        : numOfMeasurements(nMeas), numOfSteps(nSteps), seed(seed_), mx(nMeas, 0), my(nMeas, 0), functors(nMeas) {
        (void)nParticles;
    }
    void addMeasurement(F& f) { functors[last++] = f; }
    std::vector<F>& getGradientFunctors() { return functors; }
    std::vector<real>& getMx() { return mx; }
    std::vector<real>& getMy() { return my; }
    real getSignal(int i) { return std::sqrt(mx[i] * mx[i] + my[i] * my[i]); }
    int getNumOfMeas() { return numOfMeasurements; }
    int getNumOfSteps() { return numOfSteps; }
    int& getSeed() { return seed; }
    // ADC from a linear fit of ln|S| against b (phaseAcquisition.cuh:54-77)
    void calcADC() {
        points++;
        std::vector<real> ln, b;
        for (int i = 0; i < numOfMeasurements; ++i) {
            b.push_back(functors[i].getB());
            ln.push_back(std::log(getSignal(i)));
        }
        const real a = -compat_linear_regression(ln, b);
        ADC = (ADC * (points - 1) + a) / points;
        ADCmeansquare = (ADCmeansquare * (points - 1) + a * a) / points;
    }
    real& getADC() { return ADC; }
    real getADCError() { return std::sqrt(std::max<real>(ADCmeansquare - ADC * ADC, 0)); }
};

namespace axx_compat {
inline axx_ctx* context() {
    static axx_ctx* ctx = nullptr;
    if (!ctx) {
        const int rc = axx_create(&ctx, 0);
        if (rc) throw std::runtime_error(std::string("axx_create: ") + axx_strerror(rc));
    }
    return ctx;
}

template <class F>
struct Sampled {
    std::vector<float> prof;
    std::vector<int> wf_p;
    std::vector<double> amp, dir;
    int P = 0;
};

template <class F>
Sampled<F> sample(std::vector<F>& fs, int n, double dt) {
    Sampled<F> S;
    std::vector<std::tuple<int, double, double, double>> keys;
    std::vector<const F*> reps;
    for (auto& f : fs) {
        auto k = f.key();
        auto it = std::find(keys.begin(), keys.end(), k);
        int idx = (int)(it - keys.begin());
        if (it == keys.end()) { keys.push_back(k); reps.push_back(&f); }
        S.wf_p.push_back(idx);
        S.amp.push_back(f.G);
        const double nrm = std::sqrt((double)(f.ghat * f.ghat));
        S.dir.push_back(f.ghat.x / nrm); S.dir.push_back(f.ghat.y / nrm); S.dir.push_back(f.ghat.z / nrm);
    }
    S.P = (int)reps.size();
    S.prof.resize((size_t)n * S.P);
    for (int i = 0; i < n; ++i)
        for (int p = 0; p < S.P; ++p) S.prof[(size_t)i * S.P + p] = (float)reps[p]->unit(dt * i);
    return S;
}

// The reference's lattices list periodic copies of a primitive explicitly
// (all eight corner spheres of a cubic cell, say: its lattice test only sees
// the basis it is given); the engine builds periodic images itself, so
// copies that coincide modulo the cell (same type and radius, centres equal
// to 1e-9 of the cell under the minimum image; z ignored for cylinders)
// are passed once; so are planes that coincide up to a cell translation
// (a slab bounded by x = 0 and x = l in a cell of length l is one plane).
// map[i] = index of basis i's primitive among those kept.
template <class B>
std::vector<Prim> unique_modulo_cell(const B* basis, const Lattice& lat, std::vector<int>& map) {
    std::vector<Prim> out;
    const int n = lat.getBasisSize();
    map.assign(n, -1);
    const real L[3] = {lat.getA(), lat.getB(), lat.getC()};
    const real tol = 1e-9 * std::max(L[0], std::max(L[1], L[2]));
    for (int i = 0; i < n; ++i) {
        const Prim& p = basis[i].p;
        for (int k = 0; k < (int)out.size() && map[i] < 0; ++k) {
            const Prim& q = out[k];
            if (p.type != q.type) continue;
            if (p.type == AXX_PRIM_PLANE) {
                // parallel normals (either sense) and no offset along them
                // up to a cell translation
                const real lp = std::sqrt(p.nx * p.nx + p.ny * p.ny + p.nz * p.nz);
                const real lq = std::sqrt(q.nx * q.nx + q.ny * q.ny + q.nz * q.nz);
                if (!(lp > 0 && lq > 0) ||
                    std::fabs((p.nx * q.nx + p.ny * q.ny + p.nz * q.nz) / (lp * lq)) < 1.0 - 1e-12)
                    continue;
                const real d[3] = {p.cx - q.cx, p.cy - q.cy, p.cz - q.cz};
                for (int sx = -1; sx <= 1 && map[i] < 0; ++sx)
                    for (int sy = -1; sy <= 1 && map[i] < 0; ++sy)
                        for (int sz = -1; sz <= 1 && map[i] < 0; ++sz) {
                            const real off = ((d[0] - sx * L[0]) * q.nx + (d[1] - sy * L[1]) * q.ny +
                                              (d[2] - sz * L[2]) * q.nz) / lq;
                            if (std::fabs(off) <= tol) map[i] = k;
                        }
                continue;
            }
            if (std::fabs(p.R - q.R) > tol) continue;
            const real d[3] = {p.cx - q.cx, p.cy - q.cy, p.type == AXX_PRIM_CYLINDER_XY ? 0.0 : p.cz - q.cz};
            real r2 = 0;
            for (int a = 0; a < 3; ++a) {
                const real m = d[a] - L[a] * std::nearbyint(d[a] / L[a]);
                r2 += m * m;
            }
            if (r2 <= tol * tol) map[i] = k;
        }
        if (map[i] < 0) {
            map[i] = (int)out.size();
            out.push_back(p);
        }
    }
    return out;
}

inline void check(int rc, const char* what) {
    if (rc) {
        std::string msg = std::string(what) + ": " + axx_strerror(rc) + " — " + axx_last_error(context());
        throw std::runtime_error(msg);
    }
}
}  // namespace axx_compat

template <class F>
class phaseAcquisitionStream {
  public:
    int number_of_particles = 0;
    std::vector<phaseAcquisition<F>> acqSet;
    phaseAcquisitionStream() {}
    explicit phaseAcquisitionStream(int n) : number_of_particles(n) {}
    void addAcquisition(phaseAcquisition<F> a) { acqSet.push_back(a); }
    phaseAcquisition<F>& getAcquisition(int i) { return acqSet[i]; }

    // Periodic lattice of basis[0 .. lat.getBasisSize()) (phaseAcquisitionStream.cuh:126-190)
    template <class B>
    void runAcquisitionLattice(int num, B* basis, Lattice& lat, real dt, int blocks, int threads, int numOfSM) {
        (void)blocks; (void)threads; (void)numOfSM;
        std::vector<int> map;
        const std::vector<axx_compat::Prim> prims = axx_compat::unique_modulo_cell(basis, lat, map);
        run(num, prims, lat, dt, init_mode(), init_arg(map));
    }
    // Single pore: walkers start inside basis (phaseAcquisitionStream.cuh:37-88);
    // the pore sits in a cell three radii wider than it.
    template <class B>
    void runAcquisition(int num, B& basis, real dt, int blocks, int threads, int numOfSM) {
        (void)blocks; (void)threads; (void)numOfSM;
        Lattice lat;
        const axx_compat::Prim p = pore(basis, lat);
        run(num, std::vector<axx_compat::Prim>{p}, lat, dt, AXX_INIT_INSIDE_PRIM, 0);
    }
    // The reference's serial CPU path for a single pore
    // (phaseAcquisitionStream.cuh:90-124, CPUkernels.cuh:80-172): here the
    // same walk runs on the engine (its results equal the CPU oracle's bit
    // for bit).
    template <class B>
    void runAcquisitionCPU(int num, B& basis, real dt) { runAcquisition(num, basis, dt, 0, 0, 0); }
    // Gaussian steps with rejection at the pore wall (kernelWC.cuh:25-89,
    // phaseAcquisitionStream.cuh:192-252).
    template <class B>
    void runAcquisitionWC(int num, B& basis, real dt, int blocks, int threads) {
        (void)blocks; (void)threads;
        Lattice lat;
        const axx_compat::Prim p = pore(basis, lat);
        run(num, std::vector<axx_compat::Prim>{p}, lat, dt, AXX_INIT_INSIDE_PRIM, 0, AXX_STEP_GAUSSIAN,
            AXX_BOUNDARY_REJECTION);
    }
    // Every acquisition, single pore (phaseAcquisitionStream.cuh:254-423):
    // one batched persistent launch for all of them (each acquisition gets
    // exactly what runAcquisition(i, ...) gives it).
    template <class B>
    void runAcquisitionStream(B& basis, real dt, int blocks, int threads, int numOfDevices, std::vector<int>& plan,
                              std::vector<int>& numOfSMPerDevice) {
        (void)blocks; (void)threads; (void)numOfDevices; (void)plan; (void)numOfSMPerDevice;
        Lattice lat;
        const std::vector<axx_compat::Prim> prims{pore(basis, lat)};
        run_all(prims, lat, dt, AXX_INIT_INSIDE_PRIM, 0);
    }
    // All acquisitions, one device (phaseAcquisitionStream.cuh:425-613): one
    // batched persistent launch, results equal to runAcquisitionLattice(i, ...)
    template <class B>
    void runAcquisitionStreamLattice(B* basis, Lattice& lat, real dt, int blocks, int threads, int numOfDevices,
                                     std::vector<int>& plan, std::vector<int>& numOfSMPerDevice) {
        (void)blocks; (void)threads; (void)numOfDevices; (void)plan; (void)numOfSMPerDevice;
        std::vector<int> map;
        const std::vector<axx_compat::Prim> prims = axx_compat::unique_modulo_cell(basis, lat, map);
        run_all(prims, lat, dt, init_mode(), init_arg(map));
    }
    // acquisition i gets seed s (i + 1) (phaseAcquisitionStream.cuh:623-627)
    void changeSeeds(long s) {
        for (int i = 0; i < (int)acqSet.size(); ++i) acqSet[i].getSeed() = (int)(s * (i + 1));
    }
    void calcEveryADC() { for (auto& a : acqSet) a.calcADC(); }
    void flushADC() { for (auto& a : acqSet) { a.ADC = 0; a.ADCmeansquare = 0; a.points = 0; } }

  private:
    static int init_mode() {
#if defined USE_INITIALIZE_IN_REGION
#if USE_INITIALIZE_IN_REGION == 0
        return AXX_INIT_COMPARTMENT;
#else
        return AXX_INIT_INSIDE_PRIM;
#endif
#else
        return AXX_INIT_UNIFORM;
#endif
    }
    // basis index of the starting primitive, renumbered by map when periodic
    // copies were merged
    static int init_arg(const std::vector<int>& map = std::vector<int>()) {
#if defined USE_INITIALIZE_IN_REGION && USE_INITIALIZE_IN_REGION > 0
        const int k = USE_INITIALIZE_IN_REGION - 1;
        return k < (int)map.size() ? map[k] : k;
#else
        (void)map;
        return 0;
#endif
    }
    static int step_mode() {
#if defined LATTICE_PICKING
        return AXX_STEP_AXES;    // +-sqrt(2 D dt) per axis (the reference's +-L per axis is defect D4)
#else
        return AXX_STEP_ISOTROPIC;
#endif
    }
    // A single pore centred in a periodic cell six radii wide (so the pore's
    // images never matter); the cell takes the pore's D and T2.
    template <class B>
    static axx_compat::Prim pore(B& basis, Lattice& lat) {
        axx_compat::Prim p = basis.p;
        const real L = 6.0 * p.R;
        p.cx = L / 2; p.cy = L / 2; p.cz = L / 2;
        lat = Lattice(L, L, L, p.T2, p.T1, p.D, 1);
        return p;
    }
    static int boundary_mode() {
#if defined USE_PERMEABLE
        return AXX_BOUNDARY_PERMEABLE;
#elif defined SPECULAR_REFLECTION
        return AXX_BOUNDARY_SPECULAR;
#else
        return AXX_BOUNDARY_REJECTION;
#endif
    }
    // Acquisition num described for the C ABI; the descriptors point into
    // the vectors, so a Desc is neither copied nor moved once filled.
    struct Desc {
        axx_compat::Sampled<F> S;
        std::vector<int> type;
        std::vector<double> center, radius, D, T2, kappa, normal, mx, my;
        axx_substrate s{};
        axx_acq q{};
        axx_run r{};
        axx_result out{};
        Desc() = default;
        Desc(const Desc&) = delete;
        Desc& operator=(const Desc&) = delete;
    };
    void describe(Desc& d, int num, const std::vector<axx_compat::Prim>& prims, Lattice& lat, real dt, int imode,
                  int iarg, int smode, int boundary) {
        phaseAcquisition<F>& acq = acqSet[num];
        const int n = acq.getNumOfSteps();
        d.S = axx_compat::sample(acq.getGradientFunctors(), n, (double)dt);
        for (auto& p : prims) {
            d.type.push_back(p.type);
            d.center.push_back(p.cx); d.center.push_back(p.cy); d.center.push_back(p.cz);
            d.radius.push_back(p.R); d.D.push_back(p.D); d.T2.push_back(p.T2); d.kappa.push_back(p.kappa);
            d.normal.push_back(p.nx); d.normal.push_back(p.ny); d.normal.push_back(p.nz);
        }
        axx_substrate& s = d.s;
        s.a = lat.a; s.b = lat.b; s.c = lat.c; s.D_lat = lat.D; s.T2_lat = lat.T2;
        s.n_prims = (int)prims.size();
        s.type = d.type.data(); s.center = d.center.data(); s.radius = d.radius.data();
        s.D = d.D.data(); s.T2 = d.T2.data(); s.kappa = d.kappa.data(); s.normal = d.normal.data();
        axx_acq& q = d.q;
        q.n_steps = n; q.dt = dt; q.gamma = GAMMA;
        q.n_profiles = d.S.P; q.profiles = d.S.prof.data();
        q.n_waveforms = (int)d.S.wf_p.size(); q.wf_profile = d.S.wf_p.data();
        q.wf_amplitude = d.S.amp.data(); q.wf_dir = d.S.dir.data();
        axx_run& r = d.r;
        r.n_walkers = number_of_particles;
        r.seed = (uint64_t)(uint32_t)acq.getSeed();
        r.instance = (uint32_t)num;
        r.init_mode = imode; r.init_arg = iarg;
        r.boundary_mode = boundary >= 0 ? boundary : boundary_mode();
#if defined USE_RELAXATION
        r.relaxation = 1;
#endif
        r.step_mode = smode;
        d.mx.assign(q.n_waveforms, 0.0);
        d.my.assign(q.n_waveforms, 0.0);
        d.out.sum_cos = d.mx.data(); d.out.sum_sin = d.my.data();
    }
    void store(const Desc& d, int num) {
        phaseAcquisition<F>& acq = acqSet[num];
        for (int j = 0; j < d.q.n_waveforms; ++j) { acq.getMx()[j] = (real)d.mx[j]; acq.getMy()[j] = (real)d.my[j]; }
    }
    void run(int num, const std::vector<axx_compat::Prim>& prims, Lattice& lat, real dt, int imode, int iarg,
             int smode = step_mode(), int boundary = -1) {
        Desc d;
        describe(d, num, prims, lat, dt, imode, iarg, smode, boundary);
        axx_compat::check(axx_run_phase(axx_compat::context(), &d.s, &d.q, &d.r, &d.out), "axx_run_phase");
        store(d, num);
    }
    // Every acquisition in one batched launch (axx_run_batch) when the batch
    // takes them (at most AXX_BATCH_MAX_PROFILES distinct waveform shapes
    // each); otherwise one run per acquisition.
    void run_all(const std::vector<axx_compat::Prim>& prims, Lattice& lat, real dt, int imode, int iarg) {
        const int n = (int)acqSet.size();
        if (n == 0) return;
        std::vector<std::unique_ptr<Desc>> ds;
        bool batched = true;
        for (int i = 0; i < n; ++i) {
            ds.emplace_back(new Desc);
            describe(*ds.back(), i, prims, lat, dt, imode, iarg, step_mode(), -1);
            batched = batched && ds.back()->q.n_profiles <= AXX_BATCH_MAX_PROFILES;
        }
        if (batched) {
            std::vector<axx_instance> inst(n);
            std::vector<axx_result> out(n);
            for (int i = 0; i < n; ++i) {
                inst[i].substrate = &ds[i]->s; inst[i].acq = &ds[i]->q; inst[i].run = &ds[i]->r;
                out[i] = ds[i]->out;
            }
            axx_compat::check(axx_run_batch(axx_compat::context(), n, inst.data(), out.data()), "axx_run_batch");
        } else {
            for (int i = 0; i < n; ++i) {
                Desc& d = *ds[i];
                axx_compat::check(axx_run_phase(axx_compat::context(), &d.s, &d.q, &d.r, &d.out), "axx_run_phase");
            }
        }
        for (int i = 0; i < n; ++i) store(*ds[i], i);
    }
};

// ------------------------------------------------------------------ timer
// CPUtimer (src/util/timer.cuh): start / stop / getTime in seconds.  Wall
// clock here: the reference's clock() counts process CPU time, which is not
// the duration of a run the host thread spends waiting on the GPU.
class CPUtimer {
  public:
    void start() { t0 = std::chrono::steady_clock::now(); }
    void stop() { t1 = std::chrono::steady_clock::now(); }
    double getTime() const { return std::chrono::duration<double>(t1 - t0).count(); }
    void display() const { std::cout << " Elapsed time = " << getTime() << " s " << std::endl; }

  private:
    std::chrono::steady_clock::time_point t0{}, t1{};
};

// ------------------------------------------------- magnetic-field functors
// bfunctors.cuh:5-476.  Each field is affine in the position:
// B(r, t) = b0(t) + J(t) r; affine(t, b0, J) gives the decomposition (J row
// major) that the engine samples, operator() the field itself.
namespace axx_compat {
struct AffineField {
    virtual ~AffineField() {}
    virtual void affine(real t, double b0[3], double J[9]) const = 0;
    Vector3 operator()(Vector3 r, real t) const {
        double b0[3], J[9];
        affine(t, b0, J);
        return Vector3(b0[0] + J[0] * r.x + J[1] * r.y + J[2] * r.z, b0[1] + J[3] * r.x + J[4] * r.y + J[5] * r.z,
                       b0[2] + J[6] * r.x + J[7] * r.y + J[8] * r.z);
    }
    static void clear(double b0[3], double J[9]) {
        for (int k = 0; k < 3; ++k) b0[k] = 0;
        for (int k = 0; k < 9; ++k) J[k] = 0;
    }
};
}  // namespace axx_compat

// saddle coil, gradient echo: (0, G z, B0 + G y), negative gradient in the second lobe
class QSC_GRE : public axx_compat::AffineField {
  public:
    real B0 = 0, G = 0, gradientDuration = 0, gradientSpacing = 0;
    QSC_GRE() {}
    QSC_GRE(real b0, real g, real d, real s) : B0(b0), G(g), gradientDuration(d), gradientSpacing(s) {}
    void affine(real t, double b0[3], double J[9]) const override {
        clear(b0, J);
        b0[2] = B0;
        const real d = gradientDuration, s = gradientSpacing;
        const real g = t < d ? G : ((t >= s + d && t < 2.0 * d + s) ? -G : 0.0);
        J[5] = g;   // By = g z
        J[7] = g;   // Bz = B0 + g y
    }
    real getB() const {
        return G * G * gradientDuration * gradientDuration * GAMMA * GAMMA *
               ((2.0 / 3.0) * gradientDuration + gradientSpacing);
    }
    real getDiffusionTime() const { return gradientSpacing + gradientDuration; }
};

// saddle-coil PGSE with a 180-degree RF pulse of rfDuration at TE/2
class QSC_PGSE : public axx_compat::AffineField {
  public:
    real B0 = 0, G = 0, gradientDuration = 0, gradientSpacing = 0, rfDuration = 0, dtmin = 0, dtmax = 0, TE = 0,
         B1 = 0;
    QSC_PGSE() {}
    QSC_PGSE(real b0, real g, real d, real s, real rf, real dmin, real dmax, real te)
        : B0(b0), G(g), gradientDuration(d), gradientSpacing(s), rfDuration(rf), dtmin(dmin), dtmax(dmax), TE(te) {
        B1 = PI / std::fabs(GAMMA * rfDuration);
    }
    void affine(real t, double b0[3], double J[9]) const override {
        clear(b0, J);
        b0[2] = B0;
        const real d = gradientDuration, s = gradientSpacing;
        if (t < d) {
            J[5] = G; J[7] = G;
        } else if (t >= TE / 2.0 && t < TE / 2.0 + rfDuration) {
            const real w = GAMMA * B0;
            b0[0] = B1 * std::cos(w * t);
            b0[1] = -B1 * std::sin(w * t);
        } else if (t >= s + d && t < 2.0 * d + s) {
            J[5] = G; J[7] = G;
        }
    }
    real getStepTimeMin(real t) const { return (t >= TE / 2.0 && t < TE / 2.0 + rfDuration) ? dtmin : dtmax; }
    real getB() const {
        return G * G * gradientDuration * gradientDuration * GAMMA * GAMMA *
               ((2.0 / 3.0) * gradientDuration + gradientSpacing);
    }
    real getDiffusionTime() const { return gradientSpacing + gradientDuration; }
};

// saddle coil, cosine-modulated gradient echo
class QSC_COS_GRE : public axx_compat::AffineField {
  public:
    real B0 = 0, G = 0, gradientDuration = 0, gradientSpacing = 0, period = 1, b = 0;
    QSC_COS_GRE() {}
    QSC_COS_GRE(real b0, real g, real d, real s, real n) : B0(b0), G(g), gradientDuration(d), gradientSpacing(s) {
        b = GAMMA * GAMMA * 0.25 * G * G * d * d * d / (PI * PI * n * n);
        period = d / n;
    }
    void affine(real t, double b0[3], double J[9]) const override {
        clear(b0, J);
        b0[2] = B0;
        const real d = gradientDuration, s = gradientSpacing;
        real g = 0;
        if (t < d) g = G * std::cos(2.0 * PI * t / period);
        else if (t >= s + d && t < 2.0 * d + s) g = -G * std::cos(2.0 * PI * (t - s - d) / period);
        J[5] = g; J[7] = g;
    }
    real getB() const { return b; }
    real getFreq() const { return 1.0 / period; }
};

// saddle coil with trapezoidal lobes, as the reference defines them
// (ramps scale by (G/ramp) t, ramp-down over [ramp + d, 3 ramp + d), second
// lobe of the same sign)
class QSC_GRE_TRAP : public axx_compat::AffineField {
  public:
    real B0 = 0, G = 0, gradientDuration = 0, gradientSpacing = 0, gradientRampTime = 1;
    QSC_GRE_TRAP() {}
    QSC_GRE_TRAP(real b0, real g, real d, real s, real ramp)
        : B0(b0), G(g), gradientDuration(d), gradientSpacing(s), gradientRampTime(ramp) {}
    void affine(real t, double b0[3], double J[9]) const override {
        clear(b0, J);
        b0[2] = B0;
        const real r = gradientRampTime, t1 = r, t2 = r + gradientDuration, t3 = 2.0 * t1 + t2,
                   t5 = gradientSpacing + t3;
        auto up = [&](real u) { return (G / r) * u; };
        auto down = [&](real u) { return G - (G / r) * u; };
        real g = 0;
        if (t < t1) g = G * up(t);
        else if (t >= t1 && t < t2) g = G;
        else if (t >= t2 && t < t3) g = G * down(t - t2);
        else if (t >= t5 && t < t5 + t1) g = G * up(t - t5);
        else if (t >= t5 + t1 && t < t5 + t2) g = G;
        else if (t >= t5 + t2 && t < t5 + t3) g = G * down(t - t5 - t2);
        J[5] = g; J[7] = g;
    }
};

// B = (0, 0, B0 + m(t) G.r) in the lobes (SIN_GRE, COS_GRE, RECT_GRE)
namespace axx_compat {
class LongitudinalLobes : public AffineField {
  public:
    real B0 = 0, gradientDuration = 0, gradientSpacing = 0;
    Vector3 G;
    virtual real mod1(real) const { return 1.0; }
    virtual real mod2(real) const { return 1.0; }
    void affine(real t, double b0[3], double J[9]) const override {
        clear(b0, J);
        b0[2] = B0;
        const real d = gradientDuration, s = gradientSpacing;
        real m = 0;
        if (t < d) m = mod1(t);
        else if (t >= s + d && t < 2.0 * d + s) m = -mod2(t);
        J[6] = m * G.x; J[7] = m * G.y; J[8] = m * G.z;
    }
};
}  // namespace axx_compat

class SIN_GRE : public axx_compat::LongitudinalLobes {
  public:
    real period = 1;
    SIN_GRE() {}
    SIN_GRE(real b0, Vector3 g, real d, real s, real T) {
        B0 = b0; G = g; gradientDuration = d; gradientSpacing = s; period = T;
    }
    real mod1(real t) const override { return std::sin(2.0 * PI * t / period); }
    real mod2(real t) const override {
        return std::sin(2.0 * PI * (t - gradientSpacing - gradientDuration) / period);
    }
};

class COS_GRE : public axx_compat::LongitudinalLobes {
  public:
    real period = 1, b = 0;
    COS_GRE() {}
    COS_GRE(real b0, Vector3 g, real d, real s, int n) {
        B0 = b0; G = g; gradientDuration = d; gradientSpacing = s;
        period = d / (real)n;
        b = GAMMA * GAMMA * 0.25 * g.magnitude() * g.magnitude() * d * d * d / (PI * PI * n * n);
    }
    COS_GRE(real b0, real g, Vector3 ghat, real d, real s, int n) : COS_GRE(b0, ghat * g, d, s, n) {
        b = GAMMA * GAMMA * 0.25 * g * g * d * d * d / (PI * PI * n * n);
    }
    real mod1(real t) const override { return std::cos(2.0 * PI * t / period); }
    real mod2(real t) const override {
        return std::cos(2.0 * PI * (t - gradientSpacing - gradientDuration) / period);
    }
    real getB() const { return b; }
    real getFreq() const { return 1.0 / period; }
};

class RECT_GRE : public axx_compat::LongitudinalLobes {
  public:
    RECT_GRE() {}
    RECT_GRE(real b0, Vector3 g, real d, real s) {
        B0 = b0; G = g; gradientDuration = d; gradientSpacing = s;
    }
};

// low-field cell with a pi/2 RF pulse over [0, rfDuration)
class UCN_lowfield : public axx_compat::AffineField {
  public:
    real B0 = 0, G = 0, rfDuration = 0, dtmin = 0, dtmax = 0, w = 0, B1 = 0;
    UCN_lowfield() {}
    UCN_lowfield(real b0, real g, real rf, real dmin, real dmax)
        : B0(b0), G(g), rfDuration(rf), dtmin(dmin), dtmax(dmax) {
        B1 = PI / (2.0 * std::fabs(GAMMA * rfDuration));
        w = GAMMA * B0;
    }
    void affine(real t, double b0[3], double J[9]) const override {
        clear(b0, J);
        if (t < rfDuration) {
            b0[0] = B1 * std::cos(w * t);
            b0[1] = -B1 * std::sin(w * t);
        }
        b0[2] = B0;
        J[0] = -0.5 * G; J[4] = -0.5 * G; J[8] = G;
    }
    real getStepTimeMin(real t) const { return t <= rfDuration ? dtmin : dtmax; }
};

// --------------------------------------------------- magnetization acquisitions
// magAcquisition<F> (magAcquisition.cuh:1-79): the fields of one acquisition
// and the summed magnetization per field.
template <class F>
class magAcquisition {
  public:
    int numOfMeasurements = 0, numOfSteps = 0, lastAlloc = 0, seed = 0;
    std::vector<real> mx, my, mz;
    std::vector<F> fieldFunctors;
    real ADC = 0, ADCmeansquare = 0;
    int points = 0;
    magAcquisition(int nmeas, int nsteps, int s)
        : numOfMeasurements(nmeas), numOfSteps(nsteps), seed(s), mx(nmeas, 0.0), my(nmeas, 0.0), mz(nmeas, 0.0),
          fieldFunctors(nmeas) {}
    void addMeasurement(F& f) { fieldFunctors[lastAlloc++] = f; }
    std::vector<F>& getFieldFunctors() { return fieldFunctors; }
    std::vector<real>& getMx() { return mx; }
    std::vector<real>& getMy() { return my; }
    std::vector<real>& getMz() { return mz; }
    int getNumOfMeas() { return numOfMeasurements; }
    int getNumOfSteps() { return numOfSteps; }
    int& getSeed() { return seed; }
    void calcADC() {
        points++;
        std::vector<real> ln, b;
        for (int i = 0; i < numOfMeasurements; ++i) {
            b.push_back(fieldFunctors[i].getB());
            ln.push_back(std::log(std::sqrt(mx[i] * mx[i] + my[i] * my[i])));
        }
        const real a = -compat_linear_regression(ln, b);
        ADC = (ADC * (points - 1) + a) / points;
        ADCmeansquare = (ADCmeansquare * (points - 1) + a * a) / points;
    }
    real getADC() { return ADC; }
    real getADCError() { return std::sqrt(std::max<real>(ADCmeansquare - ADC * ADC, 0)); }
};

namespace axx_compat {
template <class F, class = void>
struct has_step_time_min : std::false_type {};
template <class F>
struct has_step_time_min<F, decltype((void)std::declval<const F&>().getStepTimeMin(real(0)))> : std::true_type {};

template <class F>
int substeps_at(const std::vector<F>& fs, real t, real dt, std::true_type) {
    int s = 1;
    for (auto& f : fs) s = std::max(s, (int)std::ceil(dt / f.getStepTimeMin(t) - 1e-9));
    return std::min(s, 4096);
}
template <class F>
int substeps_at(const std::vector<F>&, real, real, std::false_type) { return 1; }
}  // namespace axx_compat

// magAcquisitionStream<F> (magAcquisitionStream.cuh:1-342);
// runAcquisitionStream runs every acquisition in one batched launch
// (axx_run_mag_batch).  The reference walks inside one basis primitive; here the pore sits in a periodic cell
// six radii wide (as the phase path's single-pore runs) and the fields are
// evaluated at the walker's position relative to the pore's own centre.
// Integrator macros as in the reference (kernelMag.cuh:29-68, 93-136):
// RK4_RELAXATION / USE_RELAXATION (RK4 with T1, T2 of the pore),
// RK4_NORELAXATION, DORP_NORELAXATION (Dormand-Prince, substeps from
// getStepTimeMin), ROTATE_MULTISTEP (rotations, substeps from
// getStepTimeMin); none: one exact rotation per step.
template <class F>
class magAcquisitionStream {
  public:
    std::vector<magAcquisition<F>> acqSet;
    int number_of_particles = 0;
    explicit magAcquisitionStream(int num) : number_of_particles(num) {}
    void addAcquisition(magAcquisition<F> a) { acqSet.push_back(a); }
    magAcquisition<F>& getAcquisition(int num) { return acqSet[num]; }

    template <class B>
    void runAcquisition(int num, B& basis, Vector3& initialM, real timestep, int blocks, int threads, int devNum,
                        int numOfSM) {
        (void)blocks; (void)threads; (void)devNum; (void)numOfSM;
        run(num, basis, initialM, timestep);
    }
    // The reference's serial CPU loop (updateWalkersMagCPU); the same walk runs
    // on the engine here.
    template <class B>
    void runCPUAcquisition(int num, B& basis, Vector3& initialM, real timestep) {
        run(num, basis, initialM, timestep);
    }
    template <class B>
    void runAcquisitionStream(B& basis, Vector3& initialM, real timestep, int blocks, int threads, int numOfDevices,
                              std::vector<int>& plan, std::vector<int>& numOfSMPerDevice) {
        (void)blocks; (void)threads; (void)numOfDevices; (void)plan; (void)numOfSMPerDevice;
        run_all(basis, initialM, timestep);
    }
    void changeSeeds(long s) {
        for (int i = 0; i < (int)acqSet.size(); ++i) acqSet[i].getSeed() = (int)(s * (i + 1));
    }
    void calcEveryADC() { for (auto& a : acqSet) a.calcADC(); }

  private:
    static int integrator() {
#if defined RK4_RELAXATION || defined USE_RELAXATION || defined RK4_NORELAXATION
        return AXX_MAG_RK4;
#elif defined DORP_NORELAXATION
        return AXX_MAG_DOPRI;
#else
        return AXX_MAG_ROTATE;
#endif
    }
    static bool relaxing() {
#if defined RK4_RELAXATION || defined USE_RELAXATION
        return true;
#else
        return false;
#endif
    }
    static bool multistep() {
#if defined ROTATE_MULTISTEP || defined DORP_NORELAXATION
        return true;
#else
        return false;
#endif
    }
    // Acquisition num described for the C ABI (the descriptors point into
    // the vectors, so a MagDesc is neither copied nor moved once filled).
    struct MagDesc {
        std::vector<int> sub;
        std::vector<double> samples;
        int type = 0;
        double center[3] = {0, 0, 0}, radius = 0, D = 0, T2 = 0, kappa = 0, normal[3] = {0, 0, 0}, T1[2] = {0, 0};
        axx_substrate s{};
        axx_field f{};
        axx_run r{};
        std::vector<double> mx, my, mz;
        axx_mag_result out{};
        MagDesc() = default;
        MagDesc(const MagDesc&) = delete;
        MagDesc& operator=(const MagDesc&) = delete;
    };
    template <class B>
    void describe(MagDesc& d, int num, B& basis, Vector3& initialM, real dt) {
        magAcquisition<F>& acq = acqSet[num];
        const int n = acq.getNumOfSteps();
        const int nf = acq.getNumOfMeas();
        const int integ = integrator();
        axx_compat::Prim p = basis.p;
        const real L = 6.0 * p.R;
        const double origin[3] = {L / 2 - p.cx, L / 2 - p.cy, L / 2 - p.cz};
        p.cx = L / 2; p.cy = L / 2; p.cz = L / 2;
        // substeps and sample times per step
        d.sub.assign(n, 1);
        if (multistep())
            for (int i = 0; i < n; ++i)
                d.sub[i] = axx_compat::substeps_at(acq.getFieldFunctors(), dt * i, dt,
                                                   axx_compat::has_step_time_min<F>());
        std::vector<double> times;
        static const double nodes[6] = {0.0, 1.0 / 5.0, 3.0 / 10.0, 4.0 / 5.0, 8.0 / 9.0, 1.0};
        for (int i = 0; i < n; ++i) {
            const double ti = (double)i * dt, h = dt / d.sub[i];
            if (integ == AXX_MAG_ROTATE)
                for (int k = 0; k < d.sub[i]; ++k) times.push_back(ti + k * h);
            else if (integ == AXX_MAG_RK4)
                for (int k = 0; k <= 2 * d.sub[i]; ++k) times.push_back(ti + k * (h / 2.0));
            else
                for (int k = 0; k < d.sub[i]; ++k)
                    for (int c = 0; c < 6; ++c) times.push_back(ti + (k + nodes[c]) * h);
        }
        d.samples.assign(times.size() * nf * 12, 0.0);
        for (size_t k = 0; k < times.size(); ++k)
            for (int j = 0; j < nf; ++j) {
                double* rec = &d.samples[(k * nf + j) * 12];
                acq.getFieldFunctors()[j].affine(times[k], rec, rec + 3);
            }
        d.type = p.type;
        d.center[0] = p.cx; d.center[1] = p.cy; d.center[2] = p.cz;
        d.radius = p.R; d.D = p.D; d.T2 = p.T2; d.kappa = p.kappa;
        d.normal[0] = p.nx; d.normal[1] = p.ny; d.normal[2] = p.nz;
        axx_substrate& s = d.s;
        s.a = L; s.b = L; s.c = L; s.D_lat = p.D; s.T2_lat = p.T2;
        s.n_prims = 1;
        s.type = &d.type; s.center = d.center; s.radius = &d.radius; s.D = &d.D; s.T2 = &d.T2;
        s.kappa = &d.kappa; s.normal = d.normal;
        const double t1 = p.T1 > 0 ? (double)p.T1 : INFINITY;
        d.T1[0] = t1; d.T1[1] = t1;
        axx_field& f = d.f;
        f.n_steps = n; f.dt = dt; f.gamma = GAMMA; f.n_fields = nf; f.integrator = integ;
        f.substeps = d.sub.data(); f.samples = d.samples.data();
        for (int a = 0; a < 3; ++a) f.origin[a] = origin[a];
        f.m0[0] = initialM.x; f.m0[1] = initialM.y; f.m0[2] = initialM.z;
        f.T1 = d.T1;
        axx_run& r = d.r;
        r.n_walkers = number_of_particles;
        r.seed = (uint64_t)(uint32_t)acq.getSeed();
        r.instance = (uint32_t)num;
        r.init_mode = AXX_INIT_INSIDE_PRIM;
        r.init_arg = 0;
#if defined SPECULAR_REFLECTION
        r.boundary_mode = AXX_BOUNDARY_SPECULAR;
#else
        r.boundary_mode = AXX_BOUNDARY_REJECTION;
#endif
        r.relaxation = relaxing() ? 1 : 0;
        d.mx.assign(nf, 0.0); d.my.assign(nf, 0.0); d.mz.assign(nf, 0.0);
        d.out.mx = d.mx.data(); d.out.my = d.my.data(); d.out.mz = d.mz.data();
    }
    void store(const MagDesc& d, int num) {
        magAcquisition<F>& acq = acqSet[num];
        for (int j = 0; j < d.f.n_fields; ++j) {
            acq.getMx()[j] = (real)d.mx[j];
            acq.getMy()[j] = (real)d.my[j];
            acq.getMz()[j] = (real)d.mz[j];
        }
    }
    template <class B>
    void run(int num, B& basis, Vector3& initialM, real dt) {
        MagDesc d;
        describe(d, num, basis, initialM, dt);
        axx_compat::check(axx_run_mag(axx_compat::context(), &d.s, &d.f, &d.r, &d.out), "axx_run_mag");
        store(d, num);
    }
    // Every acquisition in one batched launch when they carry the same number
    // of fields (at most AXX_MAG_GROUP); otherwise one run each.
    template <class B>
    void run_all(B& basis, Vector3& initialM, real dt) {
        const int n = (int)acqSet.size();
        if (n == 0) return;
        std::vector<std::unique_ptr<MagDesc>> ds;
        bool batched = true;
        for (int i = 0; i < n; ++i) {
            ds.emplace_back(new MagDesc);
            describe(*ds.back(), i, basis, initialM, dt);
            batched = batched && ds.back()->f.n_fields <= AXX_MAG_GROUP && ds.back()->f.n_fields == ds[0]->f.n_fields;
        }
        if (batched) {
            std::vector<axx_mag_instance> inst(n);
            std::vector<axx_mag_result> out(n);
            for (int i = 0; i < n; ++i) {
                inst[i].substrate = &ds[i]->s; inst[i].field = &ds[i]->f; inst[i].run = &ds[i]->r;
                out[i] = ds[i]->out;
            }
            axx_compat::check(axx_run_mag_batch(axx_compat::context(), n, inst.data(), out.data()),
                              "axx_run_mag_batch");
        } else {
            for (int i = 0; i < n; ++i) {
                MagDesc& d = *ds[i];
                axx_compat::check(axx_run_mag(axx_compat::context(), &d.s, &d.f, &d.r, &d.out), "axx_run_mag");
            }
        }
        for (int i = 0; i < n; ++i) store(*ds[i], i);
    }
};

// ------------------------------------------------- random packing (host)
// RPSLatticeInitializer<B> (src/substrate/RPSinitializer.h): radii from a
// distribution, random sequential placement from the largest down with the
// full minimum-image overlap test.  The engine adds periodic images itself,
// so correctEdges() keeps one copy per primitive.
template <class B>
class RPSLatticeInitializer {
  public:
    std::vector<B> basis;
    Lattice lat;
    int basisType = 0;
    bool verbose = false;
    RPSLatticeInitializer(Lattice& l, int type, bool v = false) : basis(l.getBasisSize()), lat(l), basisType(type), verbose(v) {}
    void uniformRadialDist(int seed, real maxr, real minr) {
        std::mt19937_64 g((uint64_t)seed);
        std::uniform_real_distribution<double> u(minr, maxr);
        std::vector<real> r(basis.size());
        for (auto& x : r) x = u(g);
        assign(r);
    }
    void gaussianRadialDist(int seed, double mean, double sd) {
        std::mt19937_64 g((uint64_t)seed);
        std::normal_distribution<double> nd(mean, sd);
        std::vector<real> r;
        while (r.size() < basis.size()) { double x = nd(g); if (x > 0) r.push_back(x); }
        assign(r);
    }
    // alpha = shape k, beta = rate 1/theta (RPSinitializer.h:76-101)
    void gammaRadialDist(int seed, double alpha, double beta, real minimum, real maximum) {
        std::mt19937_64 g((uint64_t)seed);
        std::gamma_distribution<double> gd(alpha, 1.0 / beta);
        std::vector<real> r;
        while (r.size() < basis.size()) { double x = gd(g); if (x > minimum && x < maximum) r.push_back(x); }
        assign(r);
    }
    bool uniformCenterDist(int seed, int triesMax) {
        std::mt19937_64 g((uint64_t)seed);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        const int n = (int)basis.size();
        for (int i = n - 1; i >= 0; --i) {
            int tries = 0;
            do {
                if (tries++ > triesMax * (n - i)) return false;
                basis[i].setCenter(Vector3(u(g) * lat.getA(), u(g) * lat.getB(), basisType == 1 ? u(g) * lat.getC() : 0.0));
            } while (overlaps(i));
        }
        return true;
    }
    void setRegions() { for (size_t i = 0; i < basis.size(); ++i) basis[i].setRegion((int)i); }
    // The packing as a substrate file (format above) and back; reading
    // replaces the basis (its primitives must be of this class's type).
    void writeJSON(const std::string& path, const std::string& provenance_json = "{}") const {
        std::vector<axx_compat::Prim> prims;
        for (auto& b : basis) prims.push_back(b.p);
        axx_compat::write_substrate_json(path, lat, prims, provenance_json);
    }
    void readJSON(const std::string& path) {
        std::vector<axx_compat::Prim> prims;
        axx_compat::read_substrate_json(path, lat, prims);
        basis.assign(prims.size(), B());
        for (size_t i = 0; i < prims.size(); ++i) basis[i].p = prims[i];
    }
    Lattice& getLattice() { return lat; }
    void correctEdges() { lat.setBasisSize((int)basis.size()); }
    real getVI() {
        real v = 0;
        for (auto& b : basis) v += PI * b.getRadius() * b.getRadius();
        return v / (lat.getA() * lat.getB());
    }

  private:
    void assign(std::vector<real> r) {
        std::sort(r.begin(), r.end());
        for (size_t i = 0; i < basis.size(); ++i) basis[i].setRadius(r[i]);
    }
    bool overlaps(int i) const {
        const Vector3 ci = basis[i].getCenter();
        for (size_t j = i + 1; j < basis.size(); ++j) {
            Vector3 d = ci - basis[j].getCenter();
            d.x -= lat.getA() * std::nearbyint(d.x / lat.getA());
            d.y -= lat.getB() * std::nearbyint(d.y / lat.getB());
            if (basisType == 1) d.z -= lat.getC() * std::nearbyint(d.z / lat.getC());
            else d.z = 0;
            const real rr = basis[i].getRadius() + basis[j].getRadius();
            if (d.squareMagnitude() < rr * rr) return true;
        }
        return false;
    }
};
