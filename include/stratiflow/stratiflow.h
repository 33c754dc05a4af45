// stratiflow.h -- C++14 facade with the reference's class and function names
// (Field.h, Stratiflow.h, IMEXRK.h, StateVector.h, Integration.h, Parameters.h)
// over the C ABI of libsfhip.so (include/sfhip.h).
//
// Differences from the reference that a caller sees:
//  * gridParams is a runtime global (constexpr in Parameters.h:42-43); set it,
//    and flowParams, before the first field or solver is constructed.
//  * Fields live on the GPU.  Element access (operator()) goes through a host
//    mirror that is fetched on first access and uploaded before the next
//    device operation; modal fields store only the active (dealiased) modes,
//    so writing an inactive mode is dropped as ModalField::Filter would.
//  * Inner products, norms and energies are accumulated in FP64 on the GPU.
//  * Errors are reported by throwing std::runtime_error (the C ABI returns
//    status codes; the reference asserts or throws std::string).
#pragma once

// The standard headers the reference's StateVector.h makes visible to its
// callers through IMEXRK.h:12-22 and OSUtils.cpp (streams, directory listing,
// assert), so that drivers written against it compile unchanged.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sfhip.h"

#ifdef USE_DOUBLE
typedef double stratifloat;
#define SF_FACADE_PRECISION 8
#else
typedef float stratifloat;
#define SF_FACADE_PRECISION 4
#endif

using complex = std::complex<stratifloat>;
constexpr stratifloat pi = 3.141592653589793238462;  // Constants.h:63
constexpr std::complex<stratifloat> i(0, 1);          // Constants.h:64

enum class BoundaryCondition { Dirichlet, Neumann };                        // Constants.h:78-82
enum class Dimensionality { ThreeDimensional, TwoDimensional, TwoAndAHalf };  // Constants.h:84-89

struct GridParams {  // Parameters.h:6-18
    int N1;
    int N2;
    int N3;
    Dimensionality dimensionality;
    bool ThirdDimension() const {
        return dimensionality == Dimensionality::ThreeDimensional || dimensionality == Dimensionality::TwoAndAHalf;
    }
};

struct FlowParams {  // Parameters.h:21-30
    stratifloat L1, L2, L3, Re, Ri, Pr;
    bool EvolveBackground;
};

inline stratifloat InitialU(stratifloat z) { return std::tanh(z); }  // Parameters.h:33-36

// physicists' Hermite polynomial H_n(x) (Field.cpp:91-131)
inline stratifloat Hermite(unsigned int n, stratifloat x) {
    double hm = 1, hc = 2 * double(x);
    if (n == 0) return 1;
    for (unsigned int m = 1; m < n; ++m) {
        const double hn = 2 * double(x) * hc - 2 * double(m) * hm;
        hm = hc;
        hc = hn;
    }
    return stratifloat(hc);
}

// mkdir -p (OSUtils.cpp:14-70)
inline void MakeCleanDir(const std::string& dir) {
    std::string path;
    for (size_t i = 0; i <= dir.size(); ++i) {
        if (i == dir.size() || (dir[i] == '/' && i > 0)) {
            path = dir.substr(0, i);
            mkdir(path.c_str(), S_IRWXU);
        }
    }
}

namespace sfacade {

template <typename D = void>
struct Globals {
    static GridParams grid;
    static FlowParams flow;
    static sf_ctx* ctx;
    static bool ctx_three;
    static std::mt19937* rng;
};
// defaults of Parameters.h:42-43 and Parameters.cpp:6-7
template <typename D>
GridParams Globals<D>::grid = {48, 1, 512, Dimensionality::TwoDimensional};
template <typename D>
FlowParams Globals<D>::flow = {stratifloat(8.885765876 / 2), stratifloat(1.570795), 10, 1000, stratifloat(0.16), 7,
                               false};
template <typename D>
sf_ctx* Globals<D>::ctx = nullptr;
template <typename D>
bool Globals<D>::ctx_three = false;
template <typename D>
std::mt19937* Globals<D>::rng = nullptr;

// the generator behind ExciteLowWavenumbers; the reference draws a fresh
// std::random_device seed on every call (Field.h:1074-1076), the facade keeps
// one generator that SeedRandom makes reproducible
inline std::mt19937& rng() {
    if (!Globals<>::rng) Globals<>::rng = new std::mt19937(std::random_device{}());
    return *Globals<>::rng;
}

inline void check(int rc, const sf_ctx* c) {
    if (rc != SF_OK) throw std::runtime_error(std::string("sfhip: ") + sf_last_error(c));
}

inline int dim_code(Dimensionality d) {
    return d == Dimensionality::ThreeDimensional ? SF_THREE_D
                                                 : (d == Dimensionality::TwoDimensional ? SF_TWO_D : SF_TWO_AND_A_HALF);
}
inline int bc_code(BoundaryCondition b) { return b == BoundaryCondition::Neumann ? SF_NEUMANN : SF_DIRICHLET; }

inline sf_flow flow_c() {
    const FlowParams& f = Globals<>::flow;
    sf_flow c;
    c.L1 = f.L1;
    c.L2 = f.L2;
    c.L3 = f.L3;
    c.Re = f.Re;
    c.Ri = f.Ri;
    c.Pr = f.Pr;
    c.evolve_background = f.EvolveBackground ? 1 : 0;
    return c;
}

// the process-wide solver context (the reference's StateVector::solver
// singleton, StateVector.cpp:238)
inline sf_ctx* ctx() {
    if (!Globals<>::ctx) {
        const GridParams& g = Globals<>::grid;
        sf_grid gc = {g.N1, g.N2, g.N3, dim_code(g.dimensionality)};
        sf_flow fc = flow_c();
        sf_ctx* c = nullptr;
        check(sf_ctx_create(&gc, &fc, SF_FACADE_PRECISION, &c), nullptr);
        Globals<>::ctx = c;
        Globals<>::ctx_three = g.ThirdDimension();
    }
    return Globals<>::ctx;
}

}  // namespace sfacade

static GridParams& gridParams = sfacade::Globals<>::grid;
static FlowParams& flowParams = sfacade::Globals<>::flow;

// ---------------------------------------------------------------------------
// Scalar files and parameter files (OSUtils.cpp:72-95, Parameters.cpp:55-111)
// ---------------------------------------------------------------------------
inline bool FileExists(const std::string& filename) {  // OSUtils.cpp:78-82
    std::ifstream f(filename, std::ios::in | std::ios::binary);
    return f.good();
}
inline bool EndsWith(const std::string& value, const std::string& ending) {  // OSUtils.h:24-28
    return ending.size() <= value.size() && value.compare(value.size() - ending.size(), ending.size(), ending) == 0;
}
// a .params file: the value as text at 30 significant digits
inline void SaveValueToFile(stratifloat value, const std::string& filename) {
    std::ofstream out(filename);
    out << std::setprecision(30) << value;
}
inline void LoadValueFromFile(stratifloat& value, const std::string& filename) {
    std::ifstream in(filename);
    in >> value;
}

namespace sfacade {
// The lines of params.dat: each parameter's C++ expression as the name, then
// its value (booleans as 0/1).  Readers match a line by its leading name.
struct ParamLine {
    std::string name;
    double value;
    bool integral;
};
inline std::vector<ParamLine> param_lines(bool twoAndAHalfFlag) {
    const GridParams& g = Globals<>::grid;
    const FlowParams& f = Globals<>::flow;
    std::vector<ParamLine> v = {
        {"gridParams.N1", double(g.N1), true},
        {"gridParams.N2", double(g.N2), true},
        {"gridParams.N3", double(g.N3), true},
        {"(gridParams.dimensionality == Dimensionality::ThreeDimensional)",
         double(g.dimensionality == Dimensionality::ThreeDimensional), true}};
    if (twoAndAHalfFlag)
        v.push_back({"(gridParams.dimensionality == Dimensionality::TwoAndAHalf)",
                     double(g.dimensionality == Dimensionality::TwoAndAHalf), true});
    const std::vector<ParamLine> flow = {{"flowParams.L1", double(f.L1), false},
                                         {"flowParams.L2", double(f.L2), false},
                                         {"flowParams.L3", double(f.L3), false},
                                         {"flowParams.Re", double(f.Re), false},
                                         {"flowParams.Ri", double(f.Ri), false},
                                         {"flowParams.Pr", double(f.Pr), false},
                                         {"flowParams.EvolveBackground", double(f.EvolveBackground), true}};
    v.insert(v.end(), flow.begin(), flow.end());
    return v;
}
inline void print_params(std::ostream& out, bool twoAndAHalfFlag) {
    for (const ParamLine& p : param_lines(twoAndAHalfFlag)) {
        out << p.name << " ";
        if (p.integral) out << int(p.value);
        else out << stratifloat(p.value);
        out << std::endl;
    }
}
// the text after "<name> " on the first line that starts with name
inline bool find_param(std::istream& in, const std::string& name, std::string& text) {
    in.clear();
    in.seekg(0);
    for (std::string line; std::getline(in, line);) {
        if (line.compare(0, name.size(), name) == 0) {
            text = line.size() > name.size() ? line.substr(name.size() + 1) : std::string();
            return true;
        }
    }
    std::cerr << "Could not load parameter " << name << std::endl;
    return false;
}
inline void check_param(std::istream& in, const std::string& name, int expected) {
    std::string text;
    if (!find_param(in, name, text) || std::stoi(text) != expected)
        throw std::string("Error: parameter ") + name + " not correct";
}
}  // namespace sfacade

// params.dat in the working directory (Parameters.cpp:55-75)
inline void DumpParameters() {
    std::ofstream out("params.dat", std::fstream::out);
    out << std::setprecision(30);
    sfacade::print_params(out, true);
}
// Parameters.cpp:95-111
inline void PrintParameters() { sfacade::print_params(std::cout, false); }
// Parameters.cpp:77-93: the grid and EvolveBackground must equal this run's
// (a mismatch throws std::string), L3 is taken from the file; call
// StateVector::ResetForParams() afterwards
inline void LoadParameters(const std::string& file) {
    std::ifstream in(file, std::fstream::in);
    const GridParams& g = gridParams;
    sfacade::check_param(in, "gridParams.N1", g.N1);
    sfacade::check_param(in, "gridParams.N2", g.N2);
    sfacade::check_param(in, "gridParams.N3", g.N3);
    sfacade::check_param(in, "(gridParams.dimensionality == Dimensionality::ThreeDimensional)",
                         g.dimensionality == Dimensionality::ThreeDimensional);
    std::string text;
    if (sfacade::find_param(in, "flowParams.L3", text)) flowParams.L3 = std::stof(text);
    sfacade::check_param(in, "flowParams.EvolveBackground", flowParams.EvolveBackground);
}

// root of a x^2 + b x + c (Integration.cpp:3-21): 0 with a message when the
// roots are complex
inline stratifloat SolveQuadratic(stratifloat a, stratifloat b, stratifloat c, bool positiveSign = false) {
    const stratifloat disc = b * b - 4 * a * c;
    if (disc < 0) {
        std::cout << "Quadratic gives complex solutions!" << std::endl;
        std::cout << "a=" << a << ", b=" << b << ", c=" << c << std::endl;
        return 0;
    }
    const stratifloat root = std::sqrt(disc);
    return (positiveSign ? (-b + root) : (-b - root)) / (2 * a);
}

// ---------------------------------------------------------------------------
// Fields
// ---------------------------------------------------------------------------
class DeviceField {
public:
    DeviceField(int kind, BoundaryCondition bc) : kind_(kind), bc_(bc) {
        sfacade::check(sf_field_alloc(sfacade::ctx(), kind, sfacade::bc_code(bc), &f_), sfacade::ctx());
        owned_ = true;
    }
    // borrow a solver-owned field
    DeviceField(sf_field* f, int kind, BoundaryCondition bc) : f_(f), kind_(kind), bc_(bc), owned_(false) {}
    DeviceField(const DeviceField& o) : DeviceField(o.kind_, o.bc_) { copy_from(o); }
    DeviceField& operator=(const DeviceField& o) {
        copy_from(o);
        return *this;
    }
    virtual ~DeviceField() {
        if (owned_ && f_) sf_field_free(f_);
    }

    BoundaryCondition BC() const { return bc_; }

    // device handle after pending host writes are uploaded; 'write' drops the mirror
    sf_field* dev(bool write) const {
        if (dirty_) {
            upload();
            dirty_ = false;
        }
        if (write) valid_ = false;
        return f_;
    }

    void Zero() { sfacade::check(sf_field_zero(dev(true)), sfacade::ctx()); }

    DeviceField& operator*=(stratifloat a) {
        sfacade::check(sf_field_scale(dev(true), a), sfacade::ctx());
        return *this;
    }
    DeviceField& operator+=(const DeviceField& o) {
        axpby(1, o, 1);
        return *this;
    }
    DeviceField& operator-=(const DeviceField& o) {
        axpby(-1, o, 1);
        return *this;
    }
    // this = a*x + b*this
    void axpby(stratifloat a, const DeviceField& x, stratifloat b) {
        sfacade::check(sf_field_axpby(dev(true), a, x.dev(false), b), sfacade::ctx());
    }

protected:
    void copy_from(const DeviceField& o) {
        if (&o == this) return;
        if (o.bc_ != bc_) throw std::runtime_error("assignment between fields of different boundary conditions");
        sfacade::check(sf_field_copy(dev(true), o.dev(false)), sfacade::ctx());
    }
    virtual size_t host_doubles() const = 0;
    void fetch() const {
        if (valid_) return;
        hostd_.resize(host_doubles());
        sfacade::check(sf_field_download(f_, hostd_.data()), sfacade::ctx());
        valid_ = true;
    }
    void upload() const { sfacade::check(sf_field_upload(f_, hostd_.data()), sfacade::ctx()); }
    double* host_write() {
        if (dirty_ == false) fetch();
        dirty_ = true;
        return hostd_.data();
    }
    const double* host_read() const {
        fetch();
        return hostd_.data();
    }

    sf_field* f_ = nullptr;
    int kind_;
    BoundaryCondition bc_;
    bool owned_ = false;
    mutable std::vector<double> hostd_;
    mutable bool valid_ = false;
    mutable bool dirty_ = false;
};

// element proxies (the reference returns T&; the facade returns a proxy that
// writes through the host mirror)
class ModalRef {
public:
    ModalRef(double* p) : p_(p) {}
    operator complex() const { return complex(stratifloat(p_[0]), stratifloat(p_[1])); }
    ModalRef& operator=(complex v) {
        p_[0] = v.real();
        p_[1] = v.imag();
        return *this;
    }
    ModalRef& operator+=(complex v) { return *this = complex(*this) + v; }
    complex value() const { return *this; }

private:
    double* p_;
};

class NodalField;
class Nodal1D;

class ModalField : public DeviceField {  // Field.h:925-1207
public:
    ModalField(BoundaryCondition bc) : DeviceField(SF_MODAL, bc) {}
    ModalField(sf_field* f, BoundaryCondition bc) : DeviceField(f, SF_MODAL, bc) {}
    ModalField(const ModalField&) = default;
    ModalField& operator=(const ModalField&) = default;
    using DeviceField::operator=;

    ModalRef operator()(int j1, int j2, int j3) { return ModalRef(host_write() + 2 * index(j1, j2, j3)); }
    complex operator()(int j1, int j2, int j3) const {
        const double* p = host_read() + 2 * index(j1, j2, j3);
        return complex(stratifloat(p[0]), stratifloat(p[1]));
    }

    void ToNodal(NodalField& other) const;
    // low-wavenumber Hermite-Gaussian excitation (Field.h:1072-1116)
    void ExciteLowWavenumbers(stratifloat L);
    void ZeroEnds() { sfacade::check(sf_field_zero_ends(dev(true)), sfacade::ctx()); }
    void NeumannEnds() { sfacade::check(sf_field_neumann_ends(dev(true)), sfacade::ctx()); }
    void Filter() { sfacade::check(sf_field_filter(dev(true)), sfacade::ctx()); }

    // ModalField::PhaseShift (Field.h:1168-1174)
    void PhaseShift(stratifloat shift) { sfacade::check(sf_phase_shift(dev(true), shift), sfacade::ctx()); }
    // ModalField::MakeMode2 (Field.h:1118-1127): zero the odd j1 stacks
    void MakeMode2() {
        const GridParams& g = gridParams;
        double* p = host_write();
        for (int j2 = 0; j2 < g.N2; ++j2)
            for (int j1 = 1; j1 < g.N1 / 2 + 1; j1 += 2)
                for (int j3 = 0; j3 < g.N3; ++j3) {
                    double* q = p + 2 * index(j1, j2, j3);
                    q[0] = q[1] = 0;
                }
    }

protected:
    size_t host_doubles() const override {
        const GridParams& g = gridParams;
        return size_t(2) * (g.N1 / 2 + 1) * g.N2 * g.N3;
    }
    static size_t index(int j1, int j2, int j3) {
        const GridParams& g = gridParams;
        return (size_t(g.N1 / 2 + 1) * j2 + j1) * g.N3 + j3;  // Field.h:399-410
    }
};

class NodalField : public DeviceField {  // Field.h:699-923
public:
    NodalField(BoundaryCondition bc) : DeviceField(SF_NODAL, bc) {}
    NodalField(sf_field* f, BoundaryCondition bc) : DeviceField(f, SF_NODAL, bc) {}
    NodalField(const NodalField&) = default;
    NodalField& operator=(const NodalField&) = default;
    using DeviceField::operator=;

    double& operator()(int n1, int n2, int n3) { return host_write()[index(n1, n2, n3)]; }
    stratifloat operator()(int n1, int n2, int n3) const { return stratifloat(host_read()[index(n1, n2, n3)]); }

    void ToModal(ModalField& other) const {
        sfacade::check(sf_to_modal(dev(false), other.dev(true)), sfacade::ctx());
    }
    // NodalField +/- Nodal1D, broadcast over every stack (Field.h:1215-1253)
    NodalField& operator+=(const Nodal1D& profile);
    NodalField& operator-=(const Nodal1D& profile);
    using DeviceField::operator+=;
    using DeviceField::operator-=;
    stratifloat Max() const {  // Field.h:771-786
        const double* p = host_read();
        double m = 0;
        for (size_t k = 0; k < host_doubles(); ++k) m = std::max(m, std::fabs(p[k]));
        return stratifloat(m);
    }
    void Save(std::ofstream& s) const {  // Field.h:470-473: raw stratifloat values
        const double* p = host_read();
        std::vector<stratifloat> v(p, p + host_doubles());
        s.write(reinterpret_cast<const char*>(v.data()), sizeof(stratifloat) * v.size());
    }
    void Load(std::ifstream& s, bool twoDimensional = false) {  // Field.h:903-920
        const GridParams& g = gridParams;
        std::vector<stratifloat> v(host_doubles());
        if (twoDimensional) {
            const size_t plane = size_t(g.N1) * g.N3;
            s.read(reinterpret_cast<char*>(v.data()), sizeof(stratifloat) * plane);
            for (int n = 1; n < g.N2; ++n) std::memcpy(&v[n * plane], &v[(n - 1) * plane], sizeof(stratifloat) * plane);
        } else {
            s.read(reinterpret_cast<char*>(v.data()), sizeof(stratifloat) * v.size());
        }
        double* p = host_write();
        for (size_t k = 0; k < v.size(); ++k) p[k] = v[k];
    }

protected:
    size_t host_doubles() const override {
        const GridParams& g = gridParams;
        return size_t(g.N1) * g.N2 * g.N3;
    }
    static size_t index(int n1, int n2, int n3) {
        const GridParams& g = gridParams;
        return (size_t(g.N1) * n2 + n1) * g.N3 + n3;
    }
};

inline void ModalField::ToNodal(NodalField& other) const {
    sfacade::check(sf_to_nodal(dev(false), other.dev(true)), sfacade::ctx());
}

// Stratiflow.h:8-35
class NeumannModal : public ModalField {
public:
    NeumannModal() : ModalField(BoundaryCondition::Neumann) {}
    NeumannModal(sf_field* f) : ModalField(f, BoundaryCondition::Neumann) {}
    NeumannModal(const NeumannModal&) = default;
    NeumannModal& operator=(const NeumannModal&) = default;
    using ModalField::operator=;
};
class DirichletModal : public ModalField {
public:
    DirichletModal() : ModalField(BoundaryCondition::Dirichlet) {}
    DirichletModal(sf_field* f) : ModalField(f, BoundaryCondition::Dirichlet) {}
    DirichletModal(const DirichletModal&) = default;
    DirichletModal& operator=(const DirichletModal&) = default;
    using ModalField::operator=;
};
class NeumannNodal : public NodalField {
public:
    NeumannNodal() : NodalField(BoundaryCondition::Neumann) {}
    NeumannNodal(sf_field* f) : NodalField(f, BoundaryCondition::Neumann) {}
    NeumannNodal(const NeumannNodal&) = default;
    NeumannNodal& operator=(const NeumannNodal&) = default;
    using NodalField::operator=;
};
class DirichletNodal : public NodalField {
public:
    DirichletNodal() : NodalField(BoundaryCondition::Dirichlet) {}
    DirichletNodal(sf_field* f) : NodalField(f, BoundaryCondition::Dirichlet) {}
    DirichletNodal(const DirichletNodal&) = default;
    DirichletNodal& operator=(const DirichletNodal&) = default;
    using NodalField::operator=;
};

// Neumann1D / Dirichlet1D (Stratiflow.h:37-49) are host arrays here; the
// solver uploads U_ when it is set.
class Nodal1D {
public:
    Nodal1D(BoundaryCondition bc) : bc_(bc), v_(gridParams.N3, 0) {}
    BoundaryCondition BC() const { return bc_; }
    std::vector<stratifloat>& Get() { return v_; }
    const std::vector<stratifloat>& Get() const { return v_; }
    // Nodal1D::SetValue (Field.h:655-682)
    void SetValue(std::function<stratifloat(stratifloat)> f, stratifloat L3);
    stratifloat Max() const { return *std::max_element(v_.begin(), v_.end()); }
    stratifloat Min() const { return *std::min_element(v_.begin(), v_.end()); }

private:
    BoundaryCondition bc_;
    std::vector<stratifloat> v_;
};
class Neumann1D : public Nodal1D {
public:
    Neumann1D() : Nodal1D(BoundaryCondition::Neumann) {}
};
class Dirichlet1D : public Nodal1D {
public:
    Dirichlet1D() : Nodal1D(BoundaryCondition::Dirichlet) {}
};

// VerticalPoints / VerticalPointsFractional (Field.cpp:3-31) from the library
inline void VerticalPointsBoth(std::vector<double>& z, std::vector<double>& zf) {
    z.assign(gridParams.N3, 0.0);
    zf.assign(gridParams.N3, 0.0);
    sfacade::check(sf_vertical_points(sfacade::ctx(), z.data(), zf.data()), sfacade::ctx());
}

// HeatPlot, as the reference's IMEXRK.h includes its Graph.h
#include "Graph.h"

// Field.h:1072-1116: coefficients (j1^2+j2^2+h^2+1)^(-5/3) (U(-1,1) + i U(-1,1))
// on the profiles exp(-2 z^2) H_h(2 z), h < 4, for j1 < N1/6 and |j2| <= M2
// (M2 = N2/6 in 3D, N2/2 otherwise), drawn from the facade's generator.
// The context's L3 (flowParams.L3) fixes the vertical points.
inline void ModalField::ExciteLowWavenumbers(stratifloat L) {
    (void)L;
    std::vector<double> z, zf;
    VerticalPointsBoth(z, zf);
    const std::vector<double>& zz = (bc_ == BoundaryCondition::Dirichlet) ? z : zf;
    const GridParams& g = gridParams;
    const int M2 = g.ThirdDimension() ? g.N2 / 6 : g.N2 / 2;
    // the draws and the profile are evaluated in double whatever stratifloat
    // is, so a seeded state is the same (to rounding) in FP32 and FP64 builds
    std::uniform_real_distribution<double> dist(-1.0, 1.0);
    std::mt19937& gen = sfacade::rng();
    for (int j1 = 0; j1 < g.N1 / 6; j1++) {
        for (int j2 = -M2; j2 <= M2; j2++) {
            std::complex<double> coeff[4];
            for (int h = 0; h < 4; h++) {
                const double re = dist(gen);
                const double im = dist(gen);
                coeff[h] = std::pow(double(j1 * j1 + j2 * j2 + h * h + 1), -5.0 / 3.0) * std::complex<double>(re, im);
            }
            const int actualj2 = j2 < 0 ? j2 + g.N2 : j2;
            for (int j3 = 0; j3 < g.N3; j3++) {
                std::complex<double> v = 0;
                const double zj = zz[j3];
                for (int h = 0; h < 4; h++) v += coeff[h] * std::exp(-zj * zj * 2) * double(Hermite(h, stratifloat(zj * 2)));
                (*this)(j1, actualj2, j3) = complex(v);
            }
        }
    }
}

// reproducible ExciteLowWavenumbers draws
inline void SeedRandom(unsigned int seed) { sfacade::rng().seed(seed); }

inline void Nodal1D::SetValue(std::function<stratifloat(stratifloat)> f, stratifloat L3) {
    (void)L3;  // the context's L3 is flowParams.L3
    std::vector<double> z, zf;
    VerticalPointsBoth(z, zf);
    const int N = gridParams.N3;
    for (int j = 0; j < N; ++j) v_[j] = f(stratifloat(zf[j]));
    if (bc_ == BoundaryCondition::Neumann) {
        v_[1] = v_[2];
        v_[0] = v_[1];
        v_[N - 2] = v_[N - 3];
        v_[N - 1] = v_[N - 2];
    } else {
        v_[1] = -v_[2];
        v_[0] = v_[1];
        v_[N - 1] = -v_[N - 2];
    }
}

inline NodalField& NodalField::operator+=(const Nodal1D& profile) {
    std::vector<double> v(profile.Get().begin(), profile.Get().end());
    sfacade::check(sf_field_add_profile(dev(true), v.data()), sfacade::ctx());
    return *this;
}
inline NodalField& NodalField::operator-=(const Nodal1D& profile) {
    std::vector<double> v(profile.Get().size());
    for (size_t k = 0; k < v.size(); ++k) v[k] = -double(profile.Get()[k]);
    sfacade::check(sf_field_add_profile(dev(true), v.data()), sfacade::ctx());
    return *this;
}

// HorizontalAverage (Integration.h:37-42): the (0,0) stack, real part
inline void HorizontalAverage(const ModalField& integrand, Nodal1D& into) {
    std::vector<double> v(gridParams.N3);
    sfacade::check(sf_field_mean_profile(integrand.dev(false), v.data()), sfacade::ctx());
    for (int k = 0; k < gridParams.N3; ++k) into.Get()[k] = stratifloat(v[k]);
}
// RemoveHorizontalAverage (Integration.h:44-48)
inline void RemoveHorizontalAverage(ModalField& field) {
    for (int k = 0; k < gridParams.N3; ++k) field(0, 0, k) = complex(0);
}
// IntegrateVertically (Integration.h:5-35): Neumann profiles on rows 1..N3-2
// with weights z[k+1]-z[k], Dirichlet profiles on rows 2..N3-2 with zf[k]-zf[k-1]
inline stratifloat IntegrateVertically(const Nodal1D& U, stratifloat L3) {
    (void)L3;  // the context's L3 is flowParams.L3
    std::vector<double> z, zf;
    VerticalPointsBoth(z, zf);
    const int N = gridParams.N3;
    stratifloat result = 0;
    if (U.BC() == BoundaryCondition::Neumann) {
        for (int k = 1; k < N - 1; k++) result += stratifloat(z[k + 1] - z[k]) * U.Get()[k];
    } else {
        for (int k = 2; k < N - 1; k++) result += stratifloat(zf[k] - zf[k - 1]) * U.Get()[k];
    }
    return result;
}
// RemoveAverage (Integration.h:72-75)
inline void RemoveAverage(ModalField& field, stratifloat L3) {
    (void)L3;  // the context's L3 is flowParams.L3
    sfacade::check(sf_remove_average(field.dev(true)), sfacade::ctx());
}
// IntegrateAllSpace (Integration.h:50-57)
inline stratifloat IntegrateAllSpace(const ModalField& u, stratifloat L1, stratifloat L2, stratifloat L3) {
    Nodal1D horzAve(u.BC());
    HorizontalAverage(u, horzAve);
    return IntegrateVertically(horzAve, L3) * L1 * L2;
}

// ddx / ddy / ddz (Stratiflow.h:51-80) evaluated into a new field
inline NeumannModal ddx(const NeumannModal& f) {
    NeumannModal r;
    sfacade::check(sf_ddx(f.dev(false), r.dev(true)), sfacade::ctx());
    return r;
}
inline DirichletModal ddx(const DirichletModal& f) {
    DirichletModal r;
    sfacade::check(sf_ddx(f.dev(false), r.dev(true)), sfacade::ctx());
    return r;
}
inline NeumannModal ddy(const NeumannModal& f) {
    NeumannModal r;
    sfacade::check(sf_ddy(f.dev(false), r.dev(true)), sfacade::ctx());
    return r;
}
inline DirichletModal ddy(const DirichletModal& f) {
    DirichletModal r;
    sfacade::check(sf_ddy(f.dev(false), r.dev(true)), sfacade::ctx());
    return r;
}
inline DirichletModal ddz(const NeumannModal& f) {
    DirichletModal r;
    sfacade::check(sf_ddz(f.dev(false), r.dev(true)), sfacade::ctx());
    return r;
}
inline NeumannModal ddz(const DirichletModal& f) {
    NeumannModal r;
    sfacade::check(sf_ddz(f.dev(false), r.dev(true)), sfacade::ctx());
    return r;
}

// InnerProd (Integration.h:117-146), FP64 on the GPU
inline stratifloat InnerProd(const ModalField& a, const ModalField& b, stratifloat L3) {
    (void)L3;
    double out = 0;
    sfacade::check(sf_inner_prod(a.dev(false), b.dev(false), &out), sfacade::ctx());
    return stratifloat(out);
}
inline double InnerProdF64(const ModalField& a, const ModalField& b) {
    double out = 0;
    sfacade::check(sf_inner_prod(a.dev(false), b.dev(false), &out), sfacade::ctx());
    return out;
}

// ---------------------------------------------------------------------------
// IMEXRK (IMEXRK.h:31-670): a view on the context's solver state
// ---------------------------------------------------------------------------
class IMEXRK {
public:
    stratifloat deltaT = 0.01f;  // IMEXRK.h:34

    IMEXRK()
        : u1(state(SF_U1)), u2(state(SF_U2)), b(state(SF_B)), p(state(SF_P)), u3(state(SF_U3)),
          U1(state(SF_NODAL_U1)), U2(state(SF_NODAL_U2)), B(state(SF_NODAL_B)), U3(state(SF_NODAL_U3)) {
        UpdateForTimestep();
    }

    void UpdateForTimestep() { sfacade::check(sf_set_dt(c(), deltaT), c()); }  // IMEXRK.h:502-545
    void TimeStep() {                                                          // IMEXRK.cpp:8-24
        sync_dt();
        sfacade::check(sf_step(c(), 1), c());
        sync_background();
    }
    void TimeStepLinear() {  // IMEXRK.cpp:26-41
        sync_dt();
        sfacade::check(sf_step_linear(c(), 1), c());
        sync_background();
    }
    // U_ as the steps left it when flowParams.EvolveBackground is set
    // (IMEXRK.cpp:100-104)
    void sync_background() {
        if (!flowParams.EvolveBackground) return;
        std::vector<double> u(gridParams.N3);
        sfacade::check(sf_get_background(c(), u.data()), c());
        for (int k = 0; k < gridParams.N3; ++k) U_.Get()[k] = stratifloat(u[k]);
    }
    // IMEXRK.h:86-126: one backward step of the adjoint state in u1..b with the
    // direct state at the end (Below) and start (Above) of the step
    void TimeStepAdjoint(const NeumannModal& u1Below, const NeumannModal& u2Below, const DirichletModal& u3Below,
                         const NeumannModal& bBelow, const NeumannModal& u1Above, const NeumannModal& u2Above,
                         const DirichletModal& u3Above, const NeumannModal& bAbove) {
        sync_dt();
        const bool three = gridParams.ThirdDimension();
        const sf_field* below[4] = {u1Below.dev(false), three ? u2Below.dev(false) : nullptr, u3Below.dev(false),
                                    bBelow.dev(false)};
        const sf_field* above[4] = {u1Above.dev(false), three ? u2Above.dev(false) : nullptr, u3Above.dev(false),
                                    bAbove.dev(false)};
        sfacade::check(sf_step_adjoint(c(), below, above), c());
        sync_background();
        U1.dev(true);
        U2.dev(true);
        U3.dev(true);
        B.dev(true);
    }
    // IMEXRK.h:445-459: mixing diagnostic from the nodal B and modal u3
    stratifloat JoverK() {
        flush_state();
        double v = 0;
        sfacade::check(sf_jover_k(c(), &v), c());
        return stratifloat(v);
    }
    // IMEXRK.h:183-199 without the image folders
    void PrepareRunAdjoint(const std::string& imageDir = "") {
        (void)imageDir;
        p.Zero();
        PopulateNodalVariables();
    }
    void RemoveDivergence(stratifloat pressureMultiplier = 1.0) {  // IMEXRK.cpp:43-73
        flush_state();
        sfacade::check(sf_remove_divergence(c(), pressureMultiplier), c());
        invalidate_state();
    }
    void FilterAll() { sfacade::check(sf_filter_all(c()), c()); }  // IMEXRK.h:128-139
    void PopulateNodalVariables() {                                 // IMEXRK.h:141-162
        flush_state();
        sfacade::check(sf_populate_nodal(c()), c());
        U1.dev(true);
        U2.dev(true);
        U3.dev(true);
        B.dev(true);
    }
    void SetBackground(const Neumann1D& velocity) {  // IMEXRK.h:321-324
        U_ = velocity;
        std::vector<double> u(U_.Get().begin(), U_.Get().end());
        sfacade::check(sf_set_background(c(), u.data(), S_.empty() ? nullptr : S_.data()), c());
    }
    // Background stratification dB/dz at the Neumann points, which the
    // reference fixes at 1 (IMEXRK.cpp:145-147); kept across SetBackground
    // calls.  An empty function restores 1.
    void SetBackgroundStratification(std::function<stratifloat(stratifloat)> dBdz) {
        S_.clear();
        if (dBdz) {
            std::vector<double> z, zf;
            VerticalPointsBoth(z, zf);
            for (double v : zf) S_.push_back(double(dBdz(stratifloat(v))));
        }
        std::vector<double> u(U_.Get().begin(), U_.Get().end());
        sfacade::check(sf_set_background(c(), u.data(), S_.empty() ? nullptr : S_.data()), c());
    }
    void SetBackground(std::function<stratifloat(stratifloat)> velocity) {  // IMEXRK.h:326-332
        Neumann1D Ubar;
        Ubar.SetValue(velocity, flowParams.L3);
        SetBackground(Ubar);
    }
    // IMEXRK.h:334-352
    void SetBackground(const NeumannModal& velocity1, const NeumannModal& velocity2, const DirichletModal& velocity3,
                       const NeumannModal& buoyancy) {
        sfacade::check(sf_set_linear_background(c(), velocity1.dev(false),
                                                gridParams.ThirdDimension() ? velocity2.dev(false) : nullptr,
                                                velocity3.dev(false), buoyancy.dev(false)),
                       c());
    }
    stratifloat CFL() {  // IMEXRK.h:355-374
        sync_dt();
        double cfl = 0;
        sfacade::check(sf_cfl(c(), &cfl), c());
        constexpr stratifloat targetCFL = 0.8;
        deltaT *= targetCFL / stratifloat(cfl);
        UpdateForTimestep();
        return stratifloat(cfl);
    }
    stratifloat CFLlinear() {  // IMEXRK.h:376-393
        sync_dt();
        double cfl = 0;
        sfacade::check(sf_cfl_linear(c(), &cfl), c());
        constexpr stratifloat targetCFL = 0.4;
        deltaT *= targetCFL / stratifloat(cfl);
        UpdateForTimestep();
        return stratifloat(cfl);
    }
    // IMEXRK.h:164-218: PrepareRun, PrepareRunLinear and PrepareRunAdjoint
    // populate the nodal fields and, when makeDirs, make the image folders the
    // Plot* members write into
    void PrepareRun(const std::string& imageDir = "", bool makeDirs = true) {
        imageDirectory = imageDir;
        PopulateNodalVariables();
        if (makeDirs) {
            for (const char* d : {"/u1", "/u2", "/u3", "/buoyancy", "/pressure", "/vorticity", "/perturbvorticity",
                                  "/buoyancyBG"})
                MakeCleanDir(imageDirectory + d);
        }
    }
    void PrepareRunLinear(const std::string& imageDir = "", bool makeDirs = true) { PrepareRun(imageDir, makeDirs); }
    stratifloat KE() const {
        double ke = 0, pe = 0;
        sfacade::check(sf_energy(c(), &ke, &pe), c());
        return stratifloat(ke);
    }
    stratifloat PE() const {
        double ke = 0, pe = 0;
        sfacade::check(sf_energy(c(), &ke, &pe), c());
        return stratifloat(pe);
    }
    // .fields files: nodal U1, U2, U3, B (IMEXRK.h:415-443)
    void SaveFlow(const std::string& filename) {
        U1.dev(true);
        U2.dev(true);
        U3.dev(true);
        B.dev(true);
        flush_state();
        sfacade::check(sf_refresh_nodal(c()), c());
        std::ofstream s(filename, std::ios::out | std::ios::binary);
        U1.Save(s);
        U2.Save(s);
        U3.Save(s);
        B.Save(s);
    }
    void LoadFlow(const std::string& filename, bool twoDimensional) {
        std::ifstream s(filename, std::ios::in | std::ios::binary);
        U1.Load(s, twoDimensional);
        U2.Load(s, twoDimensional);
        U3.Load(s, twoDimensional);
        B.Load(s, twoDimensional);
        if (twoDimensional) U2.Zero();
        U1.ToModal(u1);
        if (gridParams.ThirdDimension()) U2.ToModal(u2);
        U3.ToModal(u3);
        B.ToModal(b);
    }

    // IMEXRK.h:221-303: heat maps of slice j2 (Graph.h); the nodal fields are
    // refreshed from the modal state first
    void PlotBuoyancy(std::string filename, int j2, bool includeBackground = true) {
        if (includeBackground) {
            Neumann1D B_;
            B_.SetValue([](stratifloat z) { return z; }, flowParams.L3);
            NeumannNodal nn;
            nodal(b, nn);
            nn += B_;
            NeumannModal nm;
            nn.ToModal(nm);
            HeatPlot(nm, flowParams.L1, flowParams.L3, j2, filename);
        } else {
            HeatPlot(b, flowParams.L1, flowParams.L3, j2, filename);
        }
    }
    void PlotPressure(std::string filename, int j2) { HeatPlot(p, flowParams.L1, flowParams.L3, j2, filename); }
    void PlotVerticalVelocity(std::string filename, int j2) { HeatPlot(u3, flowParams.L1, flowParams.L3, j2, filename); }
    void PlotSpanwiseVelocity(std::string filename, int j2) {
        if (gridParams.ThirdDimension()) HeatPlot(u2, flowParams.L1, flowParams.L3, j2, filename);
    }
    // -d/dz (u1 + U_) + d/dx u3
    void PlotSpanwiseVorticity(std::string filename, int j2) {
        NeumannNodal nn;
        nodal(u1, nn);
        nn += U_;
        NeumannModal nm;
        nn.ToModal(nm);
        DirichletModal w = ddz(nm);
        w *= -1;
        w += ddx(u3);
        HeatPlot(w, flowParams.L1, flowParams.L3, j2, filename);
    }
    // d/dz u1 - d/dx u3
    void PlotPerturbationVorticity(std::string filename, int j2) {
        DirichletModal w = ddz(u1);
        w -= ddx(u3);
        HeatPlot(w, flowParams.L1, flowParams.L3, j2, filename);
    }
    void PlotStreamwiseVelocity(std::string filename, int j2, bool includeBackground = true) {
        if (includeBackground) {
            NeumannNodal nn;
            nodal(u1, nn);
            nn += U_;
            NeumannModal nm;
            nn.ToModal(nm);
            HeatPlot(nm, flowParams.L1, flowParams.L3, j2, filename);
        } else {
            HeatPlot(u1, flowParams.L1, flowParams.L3, j2, filename);
        }
    }
    void PlotAll(std::string filename, bool includeBackground) {
        const int j2 = gridParams.N2 / 2;
        PlotPressure(imageDirectory + "/pressure/" + filename, j2);
        PlotBuoyancy(imageDirectory + "/buoyancy/" + filename, j2, includeBackground);
        PlotVerticalVelocity(imageDirectory + "/u3/" + filename, j2);
        PlotSpanwiseVelocity(imageDirectory + "/u2/" + filename, j2);
        PlotStreamwiseVelocity(imageDirectory + "/u1/" + filename, j2, includeBackground);
        PlotPerturbationVorticity(imageDirectory + "/perturbvorticity/" + filename, j2);
        if (includeBackground) PlotSpanwiseVorticity(imageDirectory + "/vorticity/" + filename, j2);
    }
    std::string imageDirectory;  // IMEXRK.h:612

    void flush_state() {
        u1.dev(false);
        u2.dev(false);
        u3.dev(false);
        b.dev(false);
        p.dev(false);
    }
    void invalidate_state() {
        u1.dev(true);
        u2.dev(true);
        u3.dev(true);
        b.dev(true);
        p.dev(true);
    }

    NeumannModal u1, u2, b, p;  // IMEXRK.h:605-606
    DirichletModal u3;
    Neumann1D U_;                       // IMEXRK.h:609
    std::vector<double> S_;             // dB/dz (empty: 1)
    NeumannNodal U1, U2, B;             // IMEXRK.h:618-619
    DirichletNodal U3;

private:
    static void nodal(const ModalField& m, NodalField& n) { m.ToNodal(n); }
    static sf_ctx* c() { return sfacade::ctx(); }
    static sf_field* state(int which) {
        sf_field* f = nullptr;
        sfacade::check(sf_state_field(sfacade::ctx(), which, &f), sfacade::ctx());
        return f;
    }
    void sync_dt() {
        double d = 0;
        sfacade::check(sf_get_dt(c(), &d), c());
        if (stratifloat(d) != deltaT) UpdateForTimestep();
        flush_state();
        invalidate_state();
    }
};

// ---------------------------------------------------------------------------
// StateVector (StateVector.h:8-552)
// ---------------------------------------------------------------------------
class StateVector {
public:
    StateVector() { Zero(); }
    StateVector(const StateVector& other) {  // StateVector.h:16-23 (p is not copied)
        u1 = other.u1;
        u2 = other.u2;
        u3 = other.u3;
        b = other.b;
        EnforceBCs();
    }
    const StateVector& operator=(const StateVector& other) {  // StateVector.h:93-101
        u1 = other.u1;
        u2 = other.u2;
        u3 = other.u3;
        b = other.b;
        EnforceBCs();
        return *this;
    }

    NeumannModal u1;
    NeumannModal u2;
    DirichletModal u3;
    NeumannModal b;
    NeumannModal p;

    const StateVector& operator+=(const StateVector& o) { return MulAdd(1, o); }
    const StateVector& operator-=(const StateVector& o) { return MulAdd(-1, o); }
    const StateVector& MulAdd(stratifloat a, const StateVector& B) {  // StateVector.h:67-78
        u1.axpby(a, B.u1, 1);
        if (gridParams.ThirdDimension()) u2.axpby(a, B.u2, 1);
        u3.axpby(a, B.u3, 1);
        b.axpby(a, B.b, 1);
        EnforceBCs();
        return *this;
    }
    const StateVector& operator*=(stratifloat a) {
        u1 *= a;
        if (gridParams.ThirdDimension()) u2 *= a;
        u3 *= a;
        b *= a;
        EnforceBCs();
        return *this;
    }
    // Dot (StateVector.h:122-134), accumulated in FP64
    double DotF64(const StateVector& o) const {
        double prod = InnerProdF64(u1, o.u1) + InnerProdF64(u3, o.u3) + flowParams.Ri * InnerProdF64(b, o.b);
        if (gridParams.ThirdDimension()) prod += InnerProdF64(u2, o.u2);
        return prod;
    }
    stratifloat Dot(const StateVector& o) const { return stratifloat(DotF64(o)); }
    stratifloat Norm2() const { return Dot(*this); }
    stratifloat Energy() const { return stratifloat(0.5 * DotF64(*this)); }
    stratifloat Norm() const { return stratifloat(std::sqrt(DotF64(*this))); }

    // StateVector.h:151-157: <w, w> of the vorticity w = ddz(u1) - ddx(u3)
    stratifloat Enstrophy() const {
        DirichletModal w = ddz(u1);
        w.axpby(-1, ddx(u3), 1);
        return InnerProd(w, w, flowParams.L3);
    }
    // StateVector.h:159-169: InnerProd(u3, s) with s(1,0,j) = i, under u3's
    // (Dirichlet) weights
    stratifloat OtherAxis() const {
        std::vector<double> z, zf;
        VerticalPointsBoth(z, zf);
        const DirichletModal& w = u3;
        stratifloat result = 0;
        for (int k = 2; k < gridParams.N3 - 1; k++) {
            const stratifloat sum = 2 * std::real(w(1, 0, k) * std::conj(complex(0, 1)));
            result += stratifloat(zf[k] - zf[k - 1]) * sum;
        }
        return result;
    }
    // StateVector.h:171-197: minimum over z of the gradient Richardson number
    // of the horizontally averaged flow plus the background (which it sets to
    // InitialU, as the reference does)
    stratifloat MinimumRi() const {
        solver().SetBackground(InitialU);
        Neumann1D u_ave, b_ave;
        HorizontalAverage(u1, u_ave);
        HorizontalAverage(b, b_ave);
        const int N = gridParams.N3;
        for (int k = 0; k < N; k++) u_ave.Get()[k] = u_ave.Get()[k] + solver().U_.Get()[k];
        // ddz of a Neumann profile (Differentiation.cpp:36-47): rows 0, 1 are zero
        std::vector<double> z, zf;
        VerticalPointsBoth(z, zf);
        auto ddz1 = [&](const Neumann1D& f, int k) -> stratifloat {
            if (k < 2) return 0;
            const stratifloat inv = stratifloat(1.0 / (zf[k] - zf[k - 1]));
            return inv * f.Get()[k] - inv * f.Get()[k - 1];
        };
        Neumann1D Ri_g;
        for (int k = 0; k < N; k++) {
            const stratifloat dbdz = ddz1(b_ave, k), dudz = ddz1(u_ave, k);
            Ri_g.Get()[k] = flowParams.Ri * (dbdz + 1) / (dudz * dudz);
        }
        return Ri_g.Min();
    }
    // StateVector.h:468-474
    void MakeMode2() {
        u1.MakeMode2();
        u2.MakeMode2();
        u3.MakeMode2();
        b.MakeMode2();
    }
    // StateVector.h:423-466: add or remove U_ = tanh(z) and B_ = z through a
    // nodal round trip
    void AddBackground() { Background(1); }
    void RemoveBackground() { Background(-1); }
    // StateVector.h:476-500: heat maps of u1 (u2), u3, b and the spanwise
    // vorticity -d/dz u1 + d/dx u3 at j2 = 0; TwoAndAHalf adds the vorticity
    // of every spanwise slice
    void PlotAll(std::string directory) const {
        MakeCleanDir(directory);
        const FlowParams& f = flowParams;
        HeatPlot(u1, f.L1, f.L3, 0, directory + "/u1.png");
        if (gridParams.ThirdDimension()) HeatPlot(u2, f.L1, f.L3, 0, directory + "/u2.png");
        HeatPlot(u3, f.L1, f.L3, 0, directory + "/u3.png");
        HeatPlot(b, f.L1, f.L3, 0, directory + "/b.png");
        DirichletModal w = ddz(u1);
        w *= -1;
        w += ddx(u3);
        HeatPlot(w, f.L1, f.L3, 0, directory + "/vorticity.png");
        if (gridParams.dimensionality == Dimensionality::TwoAndAHalf)
            for (int n = 0; n < gridParams.N2; n++)
                HeatPlot(w, f.L1, f.L3, n, directory + "/vorticity" + std::to_string(n) + ".png");
    }

    void Zero() {
        u1.Zero();
        if (gridParams.ThirdDimension()) u2.Zero();
        u3.Zero();
        b.Zero();
        p.Zero();
    }
    void EnforceBCs() {  // StateVector.h:415-421
        u3.ZeroEnds();
        u1.NeumannEnds();
        u2.NeumannEnds();
        b.NeumannEnds();
    }

    stratifloat RemovePhaseShift(stratifloat shift) {
        u1.PhaseShift(shift);
        if (gridParams.ThirdDimension()) u2.PhaseShift(shift);
        u3.PhaseShift(shift);
        b.PhaseShift(shift);
        return shift;
    }
    stratifloat RemovePhaseShift() {
        const stratifloat shift = -std::arg(complex(u1(1, 0, gridParams.N3 / 2))) + pi / 2;
        return RemovePhaseShift(shift);
    }

    // StateVector.cpp:85-105
    void FixedEvolve(stratifloat deltaT, int steps, std::vector<StateVector>& result) const {
        result.resize(steps);
        CopyToSolver();
        solver().FilterAll();
        solver().PopulateNodalVariables();
        solver().RemoveDivergence(0.0f);
        solver().deltaT = deltaT;
        solver().UpdateForTimestep();
        solver().PrepareRun("blah", false);
        for (int step = 0; step < steps; ++step) {
            CopyFromSolver(result[step]);
            solver().TimeStep();
        }
    }

    // StateVector.cpp:158-187: steps backward steps of the adjoint from this
    // state, with intermediate[0..steps] the direct states of the interval
    void AdjointEvolve(stratifloat deltaT, int steps, const std::vector<StateVector>& intermediate,
                       StateVector& result) const {
        CopyToSolver();
        solver().SetBackground(InitialU);
        solver().FilterAll();
        solver().PopulateNodalVariables();
        solver().RemoveDivergence(0.0f);
        solver().PrepareRunAdjoint();
        solver().deltaT = deltaT;
        solver().UpdateForTimestep();
        for (int step = 0; step < steps; ++step) {
            const StateVector& below = intermediate[steps - 1 - step];
            const StateVector& above = intermediate[steps - step];
            solver().TimeStepAdjoint(below.u1, below.u2, below.u3, below.b, above.u1, above.u2, above.u3, above.b);
        }
        CopyFromSolver(result);
    }

    // StateVector.cpp:3-83: CFL-adaptive steps to time T that end exactly at T,
    // a progress line every 100 steps, .fields snapshots in snapshots/ (every
    // 100 steps and at the end) when snapshot is set, and the time integral of
    // JoverK returned when calcmixing is set, and when screenshot is set the
    // solver's heat maps at the same steps in images-<run>/ (PlotAll).
    stratifloat FullEvolve(stratifloat T, StateVector& result, bool snapshot = false, bool screenshot = true,
                           bool calcmixing = false) const {
        CopyToSolver();
        solver().SetBackground(InitialU);
        solver().FilterAll();
        solver().PopulateNodalVariables();
        solver().RemoveDivergence(0.0f);
        stratifloat t = 0.0f;
        int step = 0;
        static int runnum = 0;
        runnum++;
        solver().PrepareRun(std::string("images-") + std::to_string(runnum) + "/", screenshot);
        MakeCleanDir("snapshots");
        const int stepinterval = 100;
        stratifloat mixing = 0;
        auto save = [&]() {
            solver().SaveFlow("snapshots/" + std::to_string(step) + "-" + std::to_string(t) + ".fields");
        };
        auto shot = [&]() { solver().PlotAll(std::to_string(step) + "-" + std::to_string(t) + ".png", true); };
        while (t + 0.0001 < T) {
            if (step % stepinterval == 0) {
                solver().CFL();
                std::cout << step << " " << t << " " << (solver().KE() + solver().PE()) << std::endl;
                // finish exactly for the last step
                const stratifloat remaining = T - t;
                const int remainingSteps = int(remaining / solver().deltaT) + 1;
                if (remainingSteps < stepinterval) {
                    solver().deltaT = remaining / remainingSteps;
                    solver().UpdateForTimestep();
                }
                if (screenshot) shot();
                if (snapshot) save();
            }
            solver().TimeStep();
            if (calcmixing) mixing += solver().JoverK() * solver().deltaT;
            t += solver().deltaT;
            step++;
            if (t + 0.0001 >= T) {
                if (screenshot) shot();
                if (snapshot) save();
            }
        }
        CopyFromSolver(result);
        return mixing;
    }

    // StateVector.cpp:107-156: linearised steps about `about` to time T, dt
    // from CFLlinear every 100 steps (starting from 0.01) and shortened so
    // that the last step ends at T
    void LinearEvolve(stratifloat T, const StateVector& about, StateVector& result) const {
        CopyToSolver();
        solver().SetBackground(InitialU);
        solver().SetBackground(about.u1, about.u2, about.u3, about.b);
        solver().FilterAll();
        solver().PopulateNodalVariables();
        solver().RemoveDivergence(0.0f);
        stratifloat t = 0.0f;
        int step = 0;
        solver().PrepareRunLinear("", false);
        solver().deltaT = 0.01f;
        solver().UpdateForTimestep();
        const int stepinterval = 100;
        while (t + 0.0001 < T) {
            if (step % stepinterval == 0) {
                solver().CFLlinear();
                std::cout << step << " " << t << std::endl;
                const stratifloat remaining = T - t;
                const int remainingSteps = int(remaining / solver().deltaT) + 1;
                if (remainingSteps < stepinterval) {
                    solver().deltaT = remaining / remainingSteps;
                    solver().UpdateForTimestep();
                }
            }
            solver().TimeStepLinear();
            t += solver().deltaT;
            step++;
        }
        CopyFromSolver(result);
    }

    // StateVector.cpp:189-226
    void Rescale(stratifloat energy) {
        CopyToSolver();
        solver().FilterAll();
        solver().PopulateNodalVariables();
        solver().RemoveDivergence(0.0f);
        CopyFromSolver();
        const stratifloat before = Energy();
        const stratifloat scale = before != 0.0f ? std::sqrt(energy / before) : 0.0f;
        u1 *= scale;
        if (gridParams.ThirdDimension()) u2 *= scale;
        u3 *= scale;
        b *= scale;
        CopyToSolver();
        solver().FilterAll();
        solver().PopulateNodalVariables();
        solver().RemoveDivergence(0.0f);
        CopyFromSolver();
    }

    // StateVector.cpp:227-235
    void ExciteLowWavenumbers(stratifloat energy) {
        u1.ExciteLowWavenumbers(flowParams.L3);
        u2.ExciteLowWavenumbers(flowParams.L3);
        u3.ExciteLowWavenumbers(flowParams.L3);
        b.ExciteLowWavenumbers(flowParams.L3);
        Rescale(energy);
    }

    void SaveToFile(const std::string& filename) const {  // StateVector.h:260-273
        CopyToSolver();
        solver().PopulateNodalVariables();
        const bool ext = filename.size() >= 7 && filename.compare(filename.size() - 7, 7, ".fields") == 0;
        solver().SaveFlow(ext ? filename : filename + ".fields");
    }
    void LoadFromFile(const std::string& filename, bool twoDimensional = false) {  // StateVector.h:245-258
        solver().LoadFlow(filename, twoDimensional);
        CopyFromSolver();
    }

    static IMEXRK& solver() {  // StateVector.h:546
        static IMEXRK s;
        return s;
    }
    static void ResetForParams() {  // StateVector.h:505-508
        sf_flow f = sfacade::flow_c();
        sfacade::check(sf_set_flow(sfacade::ctx(), &f), sfacade::ctx());
        solver().UpdateForTimestep();
    }

    // StateVector.h:275-413: load a 2D .fields file of grid K1 x K2 x K3 written
    // with L3 = 10 and interpolate it linearly in z onto this grid (modes
    // j1 < min(K1/2+1, N1/2+1); clamped beyond the old points; Neumann ends
    // copied).  The file's ToModal runs in a temporary context of its grid.
    template <int K1, int K2, int K3>
    void LoadAndInterpolate(const std::string& filename) {
        static_assert(K2 == 1, "LoadAndInterpolate is two-dimensional (StateVector.h:303-306)");
        if (gridParams.N2 != 1 || gridParams.ThirdDimension())
            throw std::runtime_error("LoadAndInterpolate needs a 2D grid (StateVector.h:303-306)");
        Zero();
        std::ifstream in(filename, std::ios::in | std::ios::binary);
        if (!in) throw std::runtime_error("cannot open " + filename);
        sf_grid g = {K1, 1, K3, SF_TWO_D};
        sf_flow f = sfacade::flow_c();
        f.L3 = 10;
        sf_ctx* t = nullptr;
        sfacade::check(sf_ctx_create(&g, &f, SF_FACADE_PRECISION, &t), nullptr);
        std::unique_ptr<sf_ctx, int (*)(sf_ctx*)> hold(t, sf_ctx_destroy);
        const int M1old = K1 / 2 + 1;
        std::vector<double> loaded[4];
        const int bcs[4] = {SF_NEUMANN, SF_NEUMANN, SF_DIRICHLET, SF_NEUMANN};
        for (int i = 0; i < 4; ++i) {
            const size_t count = size_t(K1) * size_t(K3);
            std::vector<stratifloat> raw(count);
            in.read(reinterpret_cast<char*>(raw.data()), sizeof(stratifloat) * raw.size());
            if (!in) throw std::runtime_error(filename + " holds fewer than four " + std::to_string(K1) + "x1x" +
                                              std::to_string(K3) + " fields");
            std::vector<double> nod(raw.begin(), raw.end());
            sf_field *fn = nullptr, *fm = nullptr;
            sfacade::check(sf_field_alloc(t, SF_NODAL, bcs[i], &fn), t);
            std::unique_ptr<sf_field, int (*)(sf_field*)> hn(fn, sf_field_free);
            sfacade::check(sf_field_alloc(t, SF_MODAL, bcs[i], &fm), t);
            std::unique_ptr<sf_field, int (*)(sf_field*)> hm(fm, sf_field_free);
            sfacade::check(sf_field_upload(fn, nod.data()), t);
            sfacade::check(sf_to_modal(fn, fm), t);
            loaded[i].assign(size_t(2) * M1old * K3, 0.0);
            sfacade::check(sf_field_download(fm, loaded[i].data()), t);
        }
        std::vector<double> zo, zfo, zn, zfn;
        zo.assign(K3, 0.0);
        zfo.assign(K3, 0.0);
        sfacade::check(sf_vertical_points(t, zo.data(), zfo.data()), t);
        VerticalPointsBoth(zn, zfn);
        auto at = [&](int i, int j1, int k3) {
            const double* q = &loaded[i][2 * (size_t(j1) * K3 + k3)];
            return complex(stratifloat(q[0]), stratifloat(q[1]));
        };
        // bracketing old points and weights for a new point z (StateVector.h:316-345)
        auto bracket = [&](const std::vector<double>& old, int last, stratifloat zz, int& kb, int& ka,
                           stratifloat& wb, stratifloat& wa) {
            ka = 0;
            stratifloat zb, za;
            do {
                ka++;
                kb = ka - 1;
                zb = stratifloat(old[kb]);
                za = stratifloat(old[ka]);
            } while (za < zz && ka < last);
            wa = (zz - zb) / (za - zb);
            wb = (za - zz) / (za - zb);
            if (wa > 1) {
                wa = 1;
                wb = 0;
            }
            if (wb > 1) {
                wb = 1;
                wa = 0;
            }
        };
        const int N3 = gridParams.N3;
        for (int j1 = 0; j1 < std::min(M1old, gridParams.N1 / 2 + 1); j1++) {
            for (int j3 = 0; j3 < N3; j3++) {  // Neumann fields u1, u2, b
                int kb, ka;
                stratifloat wb, wa;
                bracket(zfo, K3 - 2, stratifloat(zfn[j3]), kb, ka, wb, wa);
                u1(j1, 0, j3) = wb * at(0, j1, kb) + wa * at(0, j1, ka);
                u2(j1, 0, j3) = wb * at(1, j1, kb) + wa * at(1, j1, ka);
                b(j1, 0, j3) = wb * at(3, j1, kb) + wa * at(3, j1, ka);
            }
            u1(j1, 0, 0) = complex(u1(j1, 0, 1));
            u2(j1, 0, 0) = complex(u2(j1, 0, 1));
            b(j1, 0, 0) = complex(b(j1, 0, 1));
            u1(j1, 0, N3 - 1) = complex(u1(j1, 0, N3 - 2));
            u2(j1, 0, N3 - 1) = complex(u2(j1, 0, N3 - 2));
            b(j1, 0, N3 - 1) = complex(b(j1, 0, N3 - 2));
            for (int j3 = 1; j3 < N3; j3++) {  // Dirichlet field u3
                int kb, ka;
                stratifloat wb, wa;
                bracket(zo, K3 - 1, stratifloat(zn[j3]), kb, ka, wb, wa);
                u3(j1, 0, j3) = wb * at(2, j1, kb) + wa * at(2, j1, ka);
            }
        }
        EnforceBCs();
    }

private:
    void Background(int sign) {
        NeumannNodal U1, B;
        Neumann1D U_, B_;
        U_.SetValue(InitialU, flowParams.L3);
        B_.SetValue([](stratifloat z) { return z; }, flowParams.L3);
        u1.ToNodal(U1);
        b.ToNodal(B);
        if (sign > 0) {
            U1 += U_;
            B += B_;
        } else {
            U1 -= U_;
            B -= B_;
        }
        U1.ToModal(u1);
        B.ToModal(b);
    }
    void CopyToSolver() const {  // StateVector.h:511-525
        solver().u1 = u1;
        if (gridParams.ThirdDimension()) solver().u2 = u2;
        else solver().u2.Zero();
        solver().u3 = u3;
        solver().b = b;
        solver().p = p;
    }
    void CopyFromSolver() { CopyFromSolver(*this); }
    void CopyFromSolver(StateVector& into) const {  // StateVector.h:532-543
        into.u1 = solver().u1;
        if (gridParams.ThirdDimension()) into.u2 = solver().u2;
        into.u3 = solver().u3;
        into.b = solver().b;
        into.p = solver().p;
        into.EnforceBCs();
    }
};

inline StateVector operator+(const StateVector& l, const StateVector& r) {
    StateVector ret = l;
    ret += r;
    return ret;
}
inline StateVector operator-(const StateVector& l, const StateVector& r) {
    StateVector ret = l;
    ret -= r;
    return ret;
}
inline StateVector operator*(stratifloat s, const StateVector& v) {
    StateVector ret = v;
    ret *= s;
    return ret;
}
inline StateVector operator*(const StateVector& v, stratifloat s) { return s * v; }
