// sf_oracle.cpp -- CPU restatement of Stratiflow's IMEX Runge-Kutta timestep.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
// compared against; only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  Nothing in stratiflow_amd/ links or calls it.
//
// It restates, in C++14 templated on float/double (the reference's
// `stratifloat`, Constants.h:5-61), the arithmetic of the reference's hot path:
//   * stretched vertical grid             Field.cpp:3-31, 62-83
//   * banded vertical operators           Differentiation.cpp:7-111 (dense there, banded here)
//   * Fourier wavenumbers                 Differentiation.cpp:113-160
//   * Neumannify / Dirichlify             Differentiation.h:17-59
//   * Thomas algorithm                    Tridiagonal.h:14-53
//   * 2D r2c / c2r transforms per z-level Field.h:734-769, 965-994 (own mixed-radix FFT, not FFTW)
//   * 2/3 filter and active-stack loops   Field.h:996-1021, 1129-1166
//   * ZeroEnds / NeumannEnds              Field.h:475-500
//   * IMEXRK::TimeStep and its stages     IMEXRK.cpp:8-24, 43-193; IMEXRK.h:37-81, 128-162, 502-594
//   * InnerProd / KE / PE                 Integration.h:5-35, 117-146; IMEXRK.h:395-411
//   * TimeStepLinear / BuildRHSLinear     IMEXRK.cpp:26-41, 195-258; IMEXRK.h:334-352
//   * TimeStepAdjoint / BuildRHSAdjoint   IMEXRK.h:86-126, 183-199, 461-500; IMEXRK.cpp:260-379
//   * JoverK                              IMEXRK.h:445-459, 470-490
//
// Deliberate differences from the reference (all tolerance-level):
//   * grid and operator coefficients are computed in double and rounded to T;
//   * the FFT is this file's own Stockham mixed-radix transform;
//   * Thomas factors are rebuilt per solve instead of stored (same operations);
//   * optional background stratification profile S(z)=dB/dz (default 1, which
//     is exactly the reference's hard-wired term at IMEXRK.cpp:145-147).
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace sfo {

// ---------------------------------------------------------------------------
// Complex FFT: Stockham autosort, mixed radix (4, 2, 3, 5, generic odd prime).
// sign = -1 forward (FFTW_FORWARD), +1 backward; unnormalised like FFTW.
// ---------------------------------------------------------------------------
template <typename T>
class CFFT {
public:
    using C = std::complex<T>;
    explicit CFFT(int n) : n_(n) {
        if (n < 1) throw std::runtime_error("fft length must be positive");
        int m = n;
        while (m % 4 == 0) { rad_.push_back(4); m /= 4; }
        while (m % 2 == 0) { rad_.push_back(2); m /= 2; }
        while (m % 3 == 0) { rad_.push_back(3); m /= 3; }
        while (m % 5 == 0) { rad_.push_back(5); m /= 5; }
        for (int p = 7; m > 1; p += 2)
            while (m % p == 0) { rad_.push_back(p); m /= p; }
        w_.resize(n);
        for (int k = 0; k < n; ++k) {
            const double a = -2.0 * M_PI * double(k) / double(n);
            w_[k] = C(T(std::cos(a)), T(std::sin(a)));
        }
    }
    int size() const { return n_; }

    // data: n contiguous complex values; scratch: n values.
    void run(C* data, C* scratch, int sign) const {
        C* in = data;
        C* out = scratch;
        int Ns = 1;
        C v[64];
        for (int R : rad_) {
            const int m = n_ / R;
            const int tstep = n_ / (Ns * R);
            for (int j = 0; j < m; ++j) {
                const int k = j % Ns;
                for (int r = 0; r < R; ++r) v[r] = in[j + r * m];
                for (int r = 1; r < R; ++r) {
                    C w = w_[(r * k * tstep) % n_];
                    if (sign > 0) w = std::conj(w);
                    v[r] *= w;
                }
                dft(v, R, sign);
                const int base = (j / Ns) * Ns * R + k;
                for (int r = 0; r < R; ++r) out[base + r * Ns] = v[r];
            }
            std::swap(in, out);
            Ns *= R;
        }
        if (in != data) std::memcpy(data, in, sizeof(C) * n_);
    }

private:
    void dft(C* v, int R, int sign) const {
        if (R == 1) return;
        if (R == 2) {
            C a = v[0], b = v[1];
            v[0] = a + b; v[1] = a - b;
            return;
        }
        if (R == 4) {
            C a0 = v[0] + v[2], a1 = v[0] - v[2];
            C a2 = v[1] + v[3], a3 = v[1] - v[3];
            // multiply a3 by -i (forward) or +i (backward)
            C a3r = (sign < 0) ? C(a3.imag(), -a3.real()) : C(-a3.imag(), a3.real());
            v[0] = a0 + a2; v[2] = a0 - a2;
            v[1] = a1 + a3r; v[3] = a1 - a3r;
            return;
        }
        // generic O(R^2) DFT with twiddles from the length-n table
        C t[64];
        const int step = n_ / R;
        for (int s = 0; s < R; ++s) {
            C acc(0, 0);
            for (int r = 0; r < R; ++r) {
                C w = w_[((r * s) % R) * step];
                if (sign > 0) w = std::conj(w);
                acc += v[r] * w;
            }
            t[s] = acc;
        }
        for (int s = 0; s < R; ++s) v[s] = t[s];
    }

    int n_;
    std::vector<int> rad_;
    std::vector<C> w_;
};

enum Dim { THREE_D = 0, TWO_D = 1, TWO_AND_A_HALF = 2 };
enum BC { NEUMANN = 0, DIRICHLET = 1 };

// ---------------------------------------------------------------------------
// Vertical grid, Field.cpp:3-31 (VerticalPoints, VerticalPointsFractional) and
// Field.cpp:62-83 (dz, dzFractional).  Slots the reference leaves
// uninitialised (z(0), dz(0)) are given extrapolated values; they are never
// read on the hot path.
// ---------------------------------------------------------------------------
struct VGrid {
    int N;
    double L;
    std::vector<double> z, zf, DY, DYF;
};

inline VGrid make_vgrid(double L, int N) {
    if (N < 8 || N % 4) throw std::runtime_error("N3 must be a multiple of 4 and >= 8 (Field.cpp:7)");
    VGrid g;
    g.N = N;
    g.L = L;
    g.z.assign(N, 0.0);
    g.zf.assign(N, 0.0);
    g.DY.assign(N, 0.0);
    g.DYF.assign(N, 0.0);
    const double cs = 2.0;
    const int M = N - 1;  // LinSpaced(N-1, -1, 1)
    const double step = 2.0 / double(M - 1);
    for (int t = 0; t < M; ++t) {
        // symmetric evaluation keeps x exactly antisymmetric
        const double x = (t < M / 2) ? -1.0 + t * step : 1.0 - (M - 1 - t) * step;
        g.z[1 + t] = L * (cs * std::pow(x, 7) + x) / (cs + 1.0);
    }
    const double offset = 0.5 * (g.z[1] + g.z[2]);
    for (int t = 1; t < N; ++t) g.z[t] *= -L / offset;
    g.z[0] = 2.0 * g.z[1] - g.z[2];
    for (int j = 1; j <= N - 2; ++j) g.zf[j] = 0.5 * (g.z[j] + g.z[j + 1]);
    g.zf[0] = 2.0 * g.zf[1] - g.zf[2];
    g.zf[N - 1] = 2.0 * g.zf[N - 2] - g.zf[N - 3];
    for (int j = 1; j < N; ++j) g.DY[j] = g.zf[j] - g.zf[j - 1];
    g.DY[0] = g.DY[1];
    for (int j = 1; j <= N - 2; ++j) g.DYF[j] = g.z[j + 1] - g.z[j];
    g.DYF[N - 1] = g.DYF[N - 2];
    g.DYF[0] = g.DYF[1];
    return g;
}

// Fourier index -> integer wavenumber, Differentiation.cpp:113-132
inline int kindex(int n, int j) {
    if (n == 1) return 0;
    if (j <= n / 3) return j;
    if (j < n - n / 3) return 0;
    return j - n;
}

// ---------------------------------------------------------------------------
// The restated solver.
// ---------------------------------------------------------------------------
template <typename T>
class Solver {
public:
    using C = std::complex<T>;

    Solver(int N1, int N2, int N3, int dim, double L1, double L2, double L3, double Re, double Ri, double Pr)
        : N1(N1), N2(N2), N3(N3), M1(N1 / 2 + 1), dim(dim), L1(L1), L2(L2), L3(L3), Re(Re), Ri(Ri), Pr(Pr),
          g(make_vgrid(L3, N3)), fx(N1), fy(N2) {
        if (dim == TWO_D && N2 != 1) throw std::runtime_error("2D requires N2 == 1 (IMEXRK.h:43)");
        third = (dim == THREE_D || dim == TWO_AND_A_HALF);
        filterSpanwise = (dim == THREE_D);
        nm = size_t(M1) * N2 * N3;
        nn = size_t(N1) * N2 * N3;
        kx.resize(M1);
        ky.resize(N2);
        kx2.resize(M1);
        ky2.resize(N2);
        for (int j = 0; j < M1; ++j) {
            const double k = 2.0 * M_PI * kindex(N1, j) / L1;
            kx[j] = T(k);
            kx2[j] = T(-k * k);
        }
        for (int j = 0; j < N2; ++j) {
            const double k = 2.0 * M_PI * kindex(N2, j) / L2;
            ky[j] = T(k);
            ky2[j] = T(-k * k);
        }
        build_operators();
        for (auto* f : modal_fields()) f->assign(nm, C(0));
        for (auto* f : nodal_fields()) f->assign(nn, T(0));
        Ubar.assign(N3, T(0));
        S.assign(N3, T(1));
        set_dt(0.01);
    }

    // ---- public state (reference layout) ----
    int N1, N2, N3, M1, dim;
    double L1, L2, L3, Re, Ri, Pr;
    bool third, filterSpanwise;
    size_t nm, nn;
    VGrid g;
    std::vector<T> kx, ky, kx2, ky2;

    std::vector<C> u1, u2, u3, b, p;             // prognostic (IMEXRK.h:605-606)
    std::vector<C> R1, R2, R3, RB, r1, r2, r3, rB; // IMEXRK.h:635-641
    std::vector<C> u1t, u2t, u3t, bt;            // linearisation state (IMEXRK.h:614-615)
    std::vector<C> neumannTemp, dirichletTemp, divergence, q;
    std::vector<T> U1, U2, U3, B, U1_tot, U2_tot, U3_tot, B_tot, nnTemp, prod;
    std::vector<T> u1F, u2F, u3F, bF, nnTemp2, ndTemp, ndTemp2;  // IMEXRK.h:625-647
    std::vector<T> Ubar;  // Neumann1D U_ (IMEXRK.h:609)
    bool evolveBackground = false;  // FlowParams::EvolveBackground (Parameters.h:29)
    std::vector<T> S;     // dB/dz profile at Neumann points (1 in the reference)
    double deltaT = 0.01;
    T h[3];

    static constexpr double beta[3] = {1.0, 25.0 / 8.0, 9.0 / 4.0};   // IMEXRK.h:631
    static constexpr double zeta[3] = {0.0, -17.0 / 8.0, -5.0 / 4.0}; // IMEXRK.h:632

    std::vector<std::vector<C>*> modal_fields() {
        return {&u1, &u2, &u3, &b, &p, &R1, &R2, &R3, &RB, &r1, &r2, &r3, &rB,
                &u1t, &u2t, &u3t, &bt, &neumannTemp, &dirichletTemp, &divergence, &q};
    }
    std::vector<std::vector<T>*> nodal_fields() {
        return {&U1, &U2, &U3, &B, &U1_tot, &U2_tot, &U3_tot, &B_tot, &nnTemp, &prod,
                &u1F, &u2F, &u3F, &bF, &nnTemp2, &ndTemp, &ndTemp2};
    }

    size_t mi(int j1, int j2, int j3) const { return (size_t(M1) * j2 + j1) * N3 + j3; }
    size_t ni(int n1, int n2, int n3) const { return (size_t(N1) * n2 + n1) * N3 + n3; }

    // active stacks of ModalField::ParallelPerStack, Field.h:1129-1166
    bool active(int j1, int j2) const {
        if (j1 >= N1 / 3) return false;
        if (N2 > 1 && filterSpanwise) return (j2 < N2 / 3) || (j2 >= N2 - N2 / 3 + 1);
        return true;
    }
    template <typename F>
    void per_active(F f) const {
#pragma omp parallel for collapse(2) schedule(static)
        for (int j2 = 0; j2 < N2; ++j2)
            for (int j1 = 0; j1 < M1; ++j1)
                if (active(j1, j2)) f(j1, j2);
    }

    // UpdateForTimestep, IMEXRK.h:502-509 (factors are rebuilt per solve)
    void set_dt(double dt) {
        deltaT = dt;
        h[0] = T(dt * 8.0 / 15.0);
        h[1] = T(dt * 2.0 / 15.0);
        h[2] = T(dt * 5.0 / 15.0);
    }

    // ---- banded vertical operators (rows not listed are zero) ----
    // Dzz Neumann rows 1..N-2 and Dirichlet rows 2..N-2, Differentiation.cpp:7-34
    std::vector<T> dzzN_lo, dzzN_di, dzzN_up, dzzD_lo, dzzD_di, dzzD_up;
    // Dz N->D rows 2..N-1 ; Dz D->N rows 1..N-1, Differentiation.cpp:36-61
    std::vector<T> dzND, dzDN;
    // reinterpolations, Differentiation.cpp:63-111
    std::vector<T> ifull_lo, ifull_di, itil_lo, itil_di;

    void build_operators() {
        const int N = N3;
        dzzN_lo.assign(N, 0); dzzN_di.assign(N, 0); dzzN_up.assign(N, 0);
        dzzD_lo.assign(N, 0); dzzD_di.assign(N, 0); dzzD_up.assign(N, 0);
        dzND.assign(N, 0); dzDN.assign(N, 0);
        ifull_lo.assign(N, 0); ifull_di.assign(N, 0); itil_lo.assign(N, 0); itil_di.assign(N, 0);
        const auto& DY = g.DY;
        const auto& DYF = g.DYF;
        for (int j = 1; j < N - 1; ++j) {
            const double lo = 1.0 / DY[j] / DYF[j], up = 1.0 / DY[j + 1] / DYF[j];
            dzzN_lo[j] = T(lo); dzzN_up[j] = T(up); dzzN_di[j] = T(-lo - up);
        }
        for (int j = 2; j < N - 1; ++j) {
            const double lo = 1.0 / DYF[j - 1] / DY[j], up = 1.0 / DYF[j] / DY[j];
            dzzD_lo[j] = T(lo); dzzD_up[j] = T(up); dzzD_di[j] = T(-lo - up);
        }
        for (int j = 2; j <= N - 1; ++j) dzND[j] = T(1.0 / DY[j]);
        for (int j = 1; j <= N - 1; ++j) dzDN[j] = T(1.0 / DYF[j]);
        for (int j = 1; j < N; ++j) {
            ifull_lo[j] = T(DYF[j] / (2.0 * DY[j]));
            ifull_di[j] = T(DYF[j - 1] / (2.0 * DY[j]));
            itil_lo[j] = T(DYF[j - 1] / (2.0 * DY[j]));
            itil_di[j] = T(DYF[j] / (2.0 * DY[j]));
        }
    }

    // column operators; U may be T or C.  out and in must differ.
    template <typename U>
    void op_dz_ND(const U* f, U* o) const {  // Neumann in, Dirichlet out
        o[0] = U(0); o[1] = U(0);
        for (int j = 2; j < N3; ++j) o[j] = (f[j] - f[j - 1]) * dzND[j];
    }
    template <typename U>
    void op_dz_DN(const U* f, U* o) const {  // Dirichlet in, Neumann out
        o[0] = U(0);
        for (int j = 1; j < N3 - 1; ++j) o[j] = (f[j + 1] - f[j]) * dzDN[j];
        o[N3 - 1] = (-f[N3 - 1]) * dzDN[N3 - 1];
    }
    template <typename U>
    void op_ifull(const U* f, U* o) const {
        o[0] = U(0);
        for (int j = 1; j < N3; ++j) o[j] = ifull_lo[j] * f[j - 1] + ifull_di[j] * f[j];
    }
    template <typename U>
    void op_itilde(const U* f, U* o) const {
        o[0] = U(0);
        for (int j = 1; j < N3; ++j) o[j] = itil_lo[j] * f[j - 1] + itil_di[j] * f[j];
    }
    template <typename U>
    void op_ibar(const U* f, U* o) const {
        o[0] = U(0);
        for (int j = 1; j < N3; ++j) o[j] = T(0.5) * f[j - 1] + T(0.5) * f[j];
    }
    template <typename U>
    void op_idir(const U* f, U* o) const {  // Dirichlet -> Neumann
        for (int j = 0; j < N3 - 1; ++j) o[j] = T(0.5) * f[j] + T(0.5) * f[j + 1];
        o[N3 - 1] = T(0.5) * f[N3 - 1];
    }
    template <typename U>
    void op_dzz(int bc, const U* f, U* o) const {
        const auto& lo = bc == NEUMANN ? dzzN_lo : dzzD_lo;
        const auto& di = bc == NEUMANN ? dzzN_di : dzzD_di;
        const auto& up = bc == NEUMANN ? dzzN_up : dzzD_up;
        o[0] = U(0);
        o[N3 - 1] = U(0);
        for (int j = 1; j < N3 - 1; ++j) o[j] = lo[j] * f[j - 1] + di[j] * f[j] + up[j] * f[j + 1];
    }

    // ---- tridiagonal systems, Tridiagonal.h + Differentiation.h:17-59 ----
    struct Tri {
        std::vector<T> a, b, c;
    };
    // kind: 0 velocity Neumann, 1 velocity Dirichlet, 2 buoyancy Neumann, 3 Laplacian
    // raw diagonals of the dense matrix the reference builds (IMEXRK.h:55-78, 502-545)
    void build_raw(int kind, int stage, int j1, int j2, Tri& t) const {
        const int N = N3;
        t.a.assign(N, 0); t.b.assign(N, 0); t.c.assign(N, 0);
        const bool dir = (kind == 1);
        const auto& lo = dir ? dzzD_lo : dzzN_lo;
        const auto& di = dir ? dzzD_di : dzzN_di;
        const auto& up = dir ? dzzD_up : dzzN_up;
        const T kk = kx2[j1];
        const T kl = ky2[j2];
        // dense rows: laplacian = Dzz + kx2 I + ky2 I  (IMEXRK.h:521-523)
        for (int j = 0; j < N; ++j) {
            T l_lo = lo[j], l_up = up[j];
            T l_di = (di[j] + kk) + kl;
            if (kind == 3) {
                t.a[j] = l_lo; t.b[j] = l_di; t.c[j] = l_up;
            } else {
                // solve = I - 0.5*h*laplacian/Re (/Pr)   (IMEXRK.h:525, 529, 538)
                const T hh = T(0.5) * h[stage];
                auto coef = [&](T v) {
                    T r = (hh * v) / T(Re);
                    if (kind == 2) r = r / T(Pr);
                    return r;
                };
                t.a[j] = -coef(l_lo);
                t.b[j] = T(1) - coef(l_di);
                t.c[j] = -coef(l_up);
            }
        }
        t.a[0] = 0;
        t.c[N - 1] = 0;
        if (!dir) {
            // Neumannify, Differentiation.h:17-35
            t.b[0] = 1; t.c[0] = 0;
            t.a[1] = 0; t.b[1] = 1; t.c[1] = -1;
            t.a[N - 2] = 1; t.b[N - 2] = -1; t.c[N - 2] = 0;
            t.a[N - 1] = 0; t.b[N - 1] = 1;
        } else {
            // Dirichlify, Differentiation.h:37-59
            t.b[0] = 1; t.c[0] = 0;
            t.a[1] = 0; t.b[1] = 1; t.c[1] = 0;
            t.a[2] = 0; t.b[2] = 1; t.c[2] = 0;
            t.a[N - 1] = 0; t.b[N - 1] = 1;
            t.b[N - 2] = 1; t.c[N - 2] = 0; t.a[N - 2] = 0;
        }
        if (kind == 3 && j1 == 0 && j2 == 0) {
            // pressure pin, IMEXRK.h:68-74
            t.a[0] = 0; t.b[0] = 1; t.c[0] = 0;
            t.a[1] = 0; t.b[1] = 1; t.c[1] = 0;
        }
    }
    // Tridiagonal::compute, Tridiagonal.h:14-27 (c becomes C)
    void build_tri(int kind, int stage, int j1, int j2, Tri& t) const {
        const int N = N3;
        build_raw(kind, stage, j1, j2, t);
        t.c[0] = t.c[0] / t.b[0];
        for (int j = 1; j < N - 1; ++j) t.c[j] = t.c[j] / (t.b[j] - t.a[j] * t.c[j - 1]);
    }
    // Tridiagonal::solve, Tridiagonal.h:29-53, applied to real and imaginary
    // parts separately (Field.h:520-531)
    void tri_solve(const Tri& t, const C* d, C* x) const {
        const int N = N3;
        std::vector<T> D(N), dr(N), xr(N);
        for (int part = 0; part < 2; ++part) {
            for (int j = 0; j < N; ++j) dr[j] = part ? d[j].imag() : d[j].real();
            D[0] = dr[0] / t.b[0];
            for (int j = 1; j < N; ++j) D[j] = (dr[j] - t.a[j] * D[j - 1]) / (t.b[j] - t.a[j] * t.c[j - 1]);
            xr[N - 1] = D[N - 1];
            for (int j = N - 2; j >= 0; --j) xr[j] = D[j] - t.c[j] * xr[j + 1];
            for (int j = 0; j < N; ++j) {
                if (part) x[j].imag(xr[j]); else x[j].real(xr[j]);
            }
        }
    }

    // ---- field-level helpers ----
    void zero(std::vector<C>& f) { std::fill(f.begin(), f.end(), C(0)); }
    // ZeroEnds, Field.h:475-486 (whole planes)
    void zero_ends(std::vector<C>& f, int bc) {
        for (int j2 = 0; j2 < N2; ++j2)
            for (int j1 = 0; j1 < M1; ++j1) {
                C* s = &f[mi(j1, j2, 0)];
                s[0] = s[1] = s[N3 - 1] = s[N3 - 2] = C(0);
                if (bc == DIRICHLET) s[2] = C(0);
            }
    }
    // NeumannEnds, Field.h:488-500
    void neumann_ends(std::vector<C>& f) {
        for (int j2 = 0; j2 < N2; ++j2)
            for (int j1 = 0; j1 < M1; ++j1) {
                C* s = &f[mi(j1, j2, 0)];
                s[0] = s[1];
                s[N3 - 1] = s[N3 - 2];
            }
    }
    // ModalField::Filter, Field.h:996-1021
    void filter(std::vector<C>& f) {
        if (N1 > 2)
            for (int j2 = 0; j2 < N2; ++j2)
                for (int j1 = N1 / 3; j1 < M1; ++j1)
                    for (int j3 = 0; j3 < N3; ++j3) f[mi(j1, j2, j3)] = C(0);
        if (N2 > 2 && filterSpanwise)
            for (int j2 = N2 / 3; j2 <= N2 - N2 / 3; ++j2)
                for (int j1 = 0; j1 < M1; ++j1)
                    for (int j3 = 0; j3 < N3; ++j3) f[mi(j1, j2, j3)] = C(0);
    }

    // ---- transforms ----
    // NodalField::ToModal, Field.h:734-769: r2c over (x,y) per z-level,
    // then *1/(N1*N2) on active stacks and Filter.
    void to_modal(const std::vector<T>& in, std::vector<C>& out) {
        std::vector<C>& o = out;
        const int A1 = N1 / 3 > 0 ? N1 / 3 : M1;
#pragma omp parallel
        {
            std::vector<C> line(std::max(N1, N2)), scr(std::max(N1, N2));
            std::vector<C> plane(size_t(M1) * N2);
#pragma omp for schedule(static)
            for (int j3 = 0; j3 < N3; ++j3) {
                for (int n2 = 0; n2 < N2; ++n2) {
                    for (int n1 = 0; n1 < N1; ++n1) line[n1] = C(in[ni(n1, n2, j3)], 0);
                    fx.run(line.data(), scr.data(), -1);
                    for (int j1 = 0; j1 < M1; ++j1) plane[size_t(n2) * M1 + j1] = line[j1];
                }
                for (int j1 = 0; j1 < M1; ++j1) {
                    // columns that the filter removes are not transformed along y
                    if (N1 > 2 && j1 >= A1) {
                        for (int j2 = 0; j2 < N2; ++j2) o[mi(j1, j2, j3)] = C(0);
                        continue;
                    }
                    for (int n2 = 0; n2 < N2; ++n2) line[n2] = plane[size_t(n2) * M1 + j1];
                    if (N2 > 1) fy.run(line.data(), scr.data(), -1);
                    for (int j2 = 0; j2 < N2; ++j2) o[mi(j1, j2, j3)] = line[j2];
                }
            }
        }
        const T s = T(1) / T(N1 * N2);
        per_active([&](int j1, int j2) {
            for (int j3 = 0; j3 < N3; ++j3) o[mi(j1, j2, j3)] *= s;
        });
        filter(o);
    }

    // ModalField::ToNodal, Field.h:965-994: unnormalised c2r.  c2c along y,
    // then a Hermitian c2r along x that ignores the imaginary parts of the
    // DC (and Nyquist) entries, as FFTW's c2r does.
    void to_nodal(const std::vector<C>& in, std::vector<T>& out) {
#pragma omp parallel
        {
            std::vector<C> line(std::max(N1, N2)), scr(std::max(N1, N2));
            std::vector<C> plane(size_t(M1) * N2);
            std::vector<char> nz(M1);
#pragma omp for schedule(static)
            for (int j3 = 0; j3 < N3; ++j3) {
                for (int j1 = 0; j1 < M1; ++j1) {
                    bool any = false;
                    for (int j2 = 0; j2 < N2; ++j2) {
                        line[j2] = in[mi(j1, j2, j3)];
                        any = any || line[j2] != C(0);
                    }
                    nz[j1] = any;
                    if (any && N2 > 1) fy.run(line.data(), scr.data(), +1);
                    for (int n2 = 0; n2 < N2; ++n2) plane[size_t(n2) * M1 + j1] = any ? line[n2] : C(0);
                }
                for (int n2 = 0; n2 < N2; ++n2) {
                    for (int j = 0; j < N1; ++j) line[j] = C(0);
                    for (int j1 = 0; j1 < M1; ++j1) {
                        C v = plane[size_t(n2) * M1 + j1];
                        const bool selfconj = (j1 == 0) || (2 * j1 == N1);
                        if (selfconj) v = C(v.real(), 0);
                        line[j1] = v;
                        if (!selfconj) line[N1 - j1] = std::conj(v);
                    }
                    fx.run(line.data(), scr.data(), +1);
                    for (int n1 = 0; n1 < N1; ++n1) out[ni(n1, n2, j3)] = line[n1].real();
                }
            }
        }
    }

    // ---- stages, IMEXRK.cpp ----
    void filter_all() {  // IMEXRK.h:128-139
        filter(u1);
        if (third) filter(u2);
        filter(u3);
        filter(b);
        filter(p);
    }
    void populate_nodal() {  // IMEXRK.h:141-162
        to_nodal(u1, U1);
        if (third) to_nodal(u2, U2);
        to_nodal(u3, U3);
        to_nodal(b, B);
    }

    // RemoveDivergence, IMEXRK.cpp:43-73
    void remove_divergence(double pressureMultiplier) {
        const T mult = T(pressureMultiplier);
        per_active([&](int j1, int j2) {
            std::vector<C> t(N3);
            const C* a1 = &u1[mi(j1, j2, 0)];
            const C* a2 = &u2[mi(j1, j2, 0)];
            const C* a3 = &u3[mi(j1, j2, 0)];
            op_dz_DN(a3, t.data());
            const C ikx(0, kx[j1]), iky(0, ky[j2]);
            C* d = &divergence[mi(j1, j2, 0)];
            for (int j = 0; j < N3; ++j) {
                if (third) d[j] = (ikx * a1[j] + iky * a2[j]) + t[j];
                else d[j] = ikx * a1[j] + t[j];
            }
        });
        zero_ends(divergence, NEUMANN);
        per_active([&](int j1, int j2) {
            Tri t;
            build_tri(3, 0, j1, j2, t);
            tri_solve(t, &divergence[mi(j1, j2, 0)], &q[mi(j1, j2, 0)]);
        });
        per_active([&](int j1, int j2) {
            std::vector<C> t(N3);
            const C ikx(0, kx[j1]), iky(0, ky[j2]);
            C* qq = &q[mi(j1, j2, 0)];
            C* a1 = &u1[mi(j1, j2, 0)];
            C* a2 = &u2[mi(j1, j2, 0)];
            C* a3 = &u3[mi(j1, j2, 0)];
            C* pp = &p[mi(j1, j2, 0)];
            for (int j = 0; j < N3; ++j) a1[j] -= ikx * qq[j];
            if (third)
                for (int j = 0; j < N3; ++j) a2[j] -= iky * qq[j];
            op_dz_ND(qq, t.data());
            for (int j = 0; j < N3; ++j) a3[j] -= t[j];
            for (int j = 0; j < N3; ++j) pp[j] += mult * qq[j];
        });
    }

    // ExplicitRK, IMEXRK.cpp:119-134
    void explicit_rk(int k) {
        const T hz = h[k] * T(zeta[k]);
        const T mh = -h[k];
        per_active([&](int j1, int j2) {
            std::vector<C> t(N3);
            const C ikx(0, kx[j1]), iky(0, ky[j2]);
            const size_t o = mi(j1, j2, 0);
            for (int j = 0; j < N3; ++j) R1[o + j] = (u1[o + j] + hz * r1[o + j]) + mh * (ikx * p[o + j]);
            if (third)
                for (int j = 0; j < N3; ++j) R2[o + j] = (u2[o + j] + hz * r2[o + j]) + mh * (iky * p[o + j]);
            op_dz_ND(&p[o], t.data());
            for (int j = 0; j < N3; ++j) R3[o + j] = (u3[o + j] + hz * r3[o + j]) + mh * t[j];
            for (int j = 0; j < N3; ++j) RB[o + j] = b[o + j] + hz * rB[o + j];
        });
        zero(r1); zero(r2); zero(r3); zero(rB);
    }

    // elementwise nodal helper over all N1*N2 stacks (Field::ParallelPerStack)
    template <typename F>
    void per_nodal_stack(F f) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int n2 = 0; n2 < N2; ++n2)
            for (int n1 = 0; n1 < N1; ++n1) f(ni(n1, n2, 0));
    }

    // prod = A*B (InterpolateProduct, Stratiflow.h:112-117) -> modal
    void prod_nn(const std::vector<T>& A, const std::vector<T>& Bf, std::vector<C>& to) {
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) prod[o + j] = A[o + j] * Bf[o + j];
        });
        to_modal(prod, to);
    }
    // prod = A1*B1 + A2*B2 (Stratiflow.h:141-148)
    void prod_nn2(const std::vector<T>& A1, const std::vector<T>& A2, const std::vector<T>& B1,
                  const std::vector<T>& B2, std::vector<C>& to) {
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) prod[o + j] = A1[o + j] * B1[o + j] + A2[o + j] * B2[o + j];
        });
        to_modal(prod, to);
    }
    // prod = ddz(Ibar(A)*B) (DifferentiateProductBar, Stratiflow.h:119-124)
    void prod_dbar(const std::vector<T>& A, const std::vector<T>& Bf, std::vector<C>& to) {
        per_nodal_stack([&](size_t o) {
            std::vector<T> ia(N3), g(N3);
            op_ibar(&A[o], ia.data());
            for (int j = 0; j < N3; ++j) g[j] = ia[j] * Bf[o + j];
            op_dz_DN(g.data(), &prod[o]);
        });
        to_modal(prod, to);
    }
    // prod = ddz(Ibar(A1)*B1 + Ibar(A2)*B2) (Stratiflow.h:149-156)
    void prod_dbar2(const std::vector<T>& A1, const std::vector<T>& A2, const std::vector<T>& B1,
                    const std::vector<T>& B2, std::vector<C>& to) {
        per_nodal_stack([&](size_t o) {
            std::vector<T> i1(N3), i2(N3), gg(N3);
            op_ibar(&A1[o], i1.data());
            op_ibar(&A2[o], i2.data());
            for (int j = 0; j < N3; ++j) gg[j] = i1[j] * B1[o + j] + i2[j] * B2[o + j];
            op_dz_DN(gg.data(), &prod[o]);
        });
        to_modal(prod, to);
    }
    // prod = Itilde(A)*B (InterpolateProductTilde, Stratiflow.h:127-132)
    void prod_tilde(const std::vector<T>& A, const std::vector<T>& Bf, std::vector<C>& to) {
        per_nodal_stack([&](size_t o) {
            std::vector<T> ia(N3);
            op_itilde(&A[o], ia.data());
            for (int j = 0; j < N3; ++j) prod[o + j] = ia[j] * Bf[o + j];
        });
        to_modal(prod, to);
    }
    // prod = Itilde(A1)*B1 + Itilde(A2)*B2 (Stratiflow.h:157-164)
    void prod_tilde2(const std::vector<T>& A1, const std::vector<T>& A2, const std::vector<T>& B1,
                     const std::vector<T>& B2, std::vector<C>& to) {
        per_nodal_stack([&](size_t o) {
            std::vector<T> i1(N3), i2(N3);
            op_itilde(&A1[o], i1.data());
            op_itilde(&A2[o], i2.data());
            for (int j = 0; j < N3; ++j) prod[o + j] = i1[j] * B1[o + j] + i2[j] * B2[o + j];
        });
        to_modal(prod, to);
    }
    // prod = Idir(A)*Idir(B) (Stratiflow.h:134-139)
    void prod_dd(const std::vector<T>& A, const std::vector<T>& Bf, std::vector<C>& to) {
        per_nodal_stack([&](size_t o) {
            std::vector<T> ia(N3), ib(N3);
            op_idir(&A[o], ia.data());
            op_idir(&Bf[o], ib.data());
            for (int j = 0; j < N3; ++j) prod[o + j] = ia[j] * ib[j];
        });
        to_modal(prod, to);
    }

    // r op= f(stack) over active stacks
    template <typename F>
    void modal_update(std::vector<C>& r, F f) {
        per_active([&](int j1, int j2) {
            std::vector<C> t(N3);
            f(j1, j2, mi(j1, j2, 0), t.data());
            C* rr = &r[mi(j1, j2, 0)];
            for (int j = 0; j < N3; ++j) rr[j] -= t[j];
        });
    }

    // Linear buoyancy and stratification terms shared by BuildRHS and
    // BuildRHSLinear (IMEXRK.cpp:140-147 and 199-206)
    void linear_terms() {
        // neumannTemp = b; RemoveHorizontalAverage; r3 += Ri*ReinterpolateFull
        per_active([&](int j1, int j2) {
            const size_t o = mi(j1, j2, 0);
            for (int j = 0; j < N3; ++j) neumannTemp[o + j] = b[o + j];
        });
        for (int j = 0; j < N3; ++j) neumannTemp[mi(0, 0, j)] = C(0);
        const T ri = T(Ri);
        per_active([&](int j1, int j2) {
            std::vector<C> t(N3);
            const size_t o = mi(j1, j2, 0);
            op_ifull(&neumannTemp[o], t.data());
            for (int j = 0; j < N3; ++j) r3[o + j] += ri * t[j];
        });
        // dirichletTemp = u3; rB -= S*ReinterpolateDirichlet(dirichletTemp)
        per_active([&](int j1, int j2) {
            std::vector<C> t(N3);
            const size_t o = mi(j1, j2, 0);
            op_idir(&u3[o], t.data());
            for (int j = 0; j < N3; ++j) rB[o + j] -= S[j] * t[j];
        });
    }

    // BuildRHS, IMEXRK.cpp:136-193
    void build_rhs() {
        linear_terms();
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) U1_tot[o + j] = U1[o + j] + Ubar[j];
        });
        prod_nn(U1_tot, U1_tot, neumannTemp);
        modal_update(r1, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
        });
        prod_dbar(U1_tot, U3, neumannTemp);
        modal_update(r1, [&](int, int, size_t o, C* t) {
            for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
        });
        prod_tilde(U1_tot, U3, dirichletTemp);
        prod_dd(U3, U3, neumannTemp);
        modal_update(r3, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            std::vector<C> dz(N3);
            op_dz_ND(&neumannTemp[o], dz.data());
            for (int j = 0; j < N3; ++j) t[j] = ikx * dirichletTemp[o + j] + dz[j];
        });
        if (third) {
            prod_nn(U2, U2, neumannTemp);
            modal_update(r2, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
            });
            prod_dbar(U2, U3, neumannTemp);
            modal_update(r2, [&](int, int, size_t o, C* t) {
                for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
            });
            prod_tilde(U2, U3, dirichletTemp);
            modal_update(r3, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * dirichletTemp[o + j];
            });
            prod_nn(U1_tot, U2, neumannTemp);
            modal_update(r1, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
            });
            modal_update(r2, [&](int j1, int, size_t o, C* t) {
                const C ikx(0, kx[j1]);
                for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
            });
        }
        prod_dbar(B, U3, neumannTemp);
        modal_update(rB, [&](int, int, size_t o, C* t) {
            for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
        });
        if (third) {
            prod_nn(U2, B, neumannTemp);
            modal_update(rB, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
            });
        }
        prod_nn(U1_tot, B, neumannTemp);
        modal_update(rB, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
        });
    }

    // SetBackground(u1,u2,u3,b), IMEXRK.h:334-352
    void set_linear_background() {
        to_nodal(u1t, U1_tot);
        if (dim == THREE_D) to_nodal(u2t, U2_tot);
        to_nodal(u3t, U3_tot);
        to_nodal(bt, B_tot);
    }

    // BuildRHSLinear, IMEXRK.cpp:195-258
    void build_rhs_linear() {
        linear_terms();
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) nnTemp[o + j] = U1_tot[o + j] + Ubar[j];
        });
        const T two = T(2);
        prod_nn(U1, nnTemp, neumannTemp);
        modal_update(r1, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            for (int j = 0; j < N3; ++j) t[j] = two * (ikx * neumannTemp[o + j]);
        });
        prod_dbar2(U1, nnTemp, U3_tot, U3, neumannTemp);
        modal_update(r1, [&](int, int, size_t o, C* t) {
            for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
        });
        prod_tilde2(U1, nnTemp, U3_tot, U3, dirichletTemp);
        // InterpolateProduct(U3, U3_tot) takes two DirichletNodal arguments,
        // i.e. the Idir*Idir overload of Stratiflow.h:134-139.
        prod_dd(U3, U3_tot, neumannTemp);
        modal_update(r3, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            std::vector<C> dz(N3);
            op_dz_ND(&neumannTemp[o], dz.data());
            for (int j = 0; j < N3; ++j) t[j] = ikx * dirichletTemp[o + j] + two * dz[j];
        });
        if (third) {
            prod_dbar2(U2, U2_tot, U3_tot, U3, neumannTemp);
            modal_update(r2, [&](int, int, size_t o, C* t) {
                for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
            });
            if (dim == THREE_D) {
                prod_nn(U2, U2_tot, neumannTemp);
                modal_update(r2, [&](int, int j2, size_t o, C* t) {
                    const C iky(0, ky[j2]);
                    for (int j = 0; j < N3; ++j) t[j] = two * (iky * neumannTemp[o + j]);
                });
                prod_tilde2(U2, U2_tot, U3_tot, U3, dirichletTemp);
                modal_update(r3, [&](int, int j2, size_t o, C* t) {
                    const C iky(0, ky[j2]);
                    for (int j = 0; j < N3; ++j) t[j] = iky * dirichletTemp[o + j];
                });
            }
            prod_nn2(U1, nnTemp, U2_tot, U2, neumannTemp);
            if (dim == THREE_D)
                modal_update(r1, [&](int, int j2, size_t o, C* t) {
                    const C iky(0, ky[j2]);
                    for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
                });
            modal_update(r2, [&](int j1, int, size_t o, C* t) {
                const C ikx(0, kx[j1]);
                for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
            });
        }
        prod_dbar2(B, B_tot, U3_tot, U3, neumannTemp);
        modal_update(rB, [&](int, int, size_t o, C* t) {
            for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
        });
        if (third) {
            prod_nn2(B, B_tot, U2_tot, U2, neumannTemp);
            modal_update(rB, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
            });
        }
        prod_nn2(B, B_tot, nnTemp, U1, neumannTemp);
        modal_update(rB, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
        });
    }

    // FinishRHS, IMEXRK.cpp:107-117
    void finish_rhs(int k) {
        const T hb = h[k] * T(beta[k]);
        per_active([&](int j1, int j2) {
            const size_t o = mi(j1, j2, 0);
            for (int j = 0; j < N3; ++j) R1[o + j] += hb * r1[o + j];
            if (third)
                for (int j = 0; j < N3; ++j) R2[o + j] += hb * r2[o + j];
            for (int j = 0; j < N3; ++j) R3[o + j] += hb * r3[o + j];
            for (int j = 0; j < N3; ++j) RB[o + j] += hb * rB[o + j];
        });
    }

    // CrankNicolson, IMEXRK.cpp:75-105 and CNSolve*, IMEXRK.h:548-564
    void crank_nicolson(int k) {
        const T cv = (T(0.5) * h[k]) / T(Re);
        const T cb = ((T(0.5) * h[k]) / T(Re)) / T(Pr);
        auto explicit_half = [&](std::vector<C>& R, const std::vector<C>& u, int bc, T c) {
            per_active([&](int j1, int j2) {
                std::vector<C> t(N3);
                const size_t o = mi(j1, j2, 0);
                op_dzz(bc, &u[o], t.data());
                for (int j = 0; j < N3; ++j) R[o + j] += c * ((kx2[j1] * u[o + j] + ky2[j2] * u[o + j]) + t[j]);
            });
        };
        auto solve = [&](std::vector<C>& R, std::vector<C>& u, int kind, int bc) {
            zero_ends(R, bc);
            per_active([&](int j1, int j2) {
                Tri t;
                build_tri(kind, k, j1, j2, t);
                tri_solve(t, &R[mi(j1, j2, 0)], &u[mi(j1, j2, 0)]);
            });
        };
        explicit_half(R1, u1, NEUMANN, cv);
        solve(R1, u1, 0, NEUMANN);
        if (third) {
            explicit_half(R2, u2, NEUMANN, cv);
            solve(R2, u2, 0, NEUMANN);
        }
        explicit_half(R3, u3, DIRICHLET, cv);
        solve(R3, u3, 1, DIRICHLET);
        explicit_half(RB, b, NEUMANN, cb);
        solve(RB, b, 2, NEUMANN);
        // IMEXRK.cpp:100-104: RU_ = U_ + (h/2Re) Dzz U_, ZeroEnds of the 1D
        // field (slots 0, 1, N3-2, N3-1, Field.h:625-636), then the (0,0)
        // velocity Neumann matrix of this stage (CNSolve1D, IMEXRK.h:584-588)
        if (evolveBackground) {
            std::vector<C> Uc(N3), t(N3), RU(N3);
            for (int j = 0; j < N3; ++j) Uc[j] = C(Ubar[j]);
            op_dzz(NEUMANN, Uc.data(), t.data());
            for (int j = 0; j < N3; ++j) RU[j] = Uc[j] + cv * t[j];
            RU[0] = RU[1] = RU[N3 - 2] = RU[N3 - 1] = C(0);
            Tri tr;
            build_tri(0, k, 0, 0, tr);
            tri_solve(tr, RU.data(), Uc.data());
            for (int j = 0; j < N3; ++j) Ubar[j] = Uc[j].real();
        }
    }

    void stage(int k, bool linear) {
        explicit_rk(k);
        if (linear) build_rhs_linear(); else build_rhs();
        finish_rhs(k);
        crank_nicolson(k);
        remove_divergence(1.0 / double(h[k]));
        filter_all();
        populate_nodal();
    }
    void time_step() { for (int k = 0; k < 3; ++k) stage(k, false); }        // IMEXRK.cpp:8-24
    void time_step_linear() { for (int k = 0; k < 3; ++k) stage(k, true); }  // IMEXRK.cpp:26-41

    // ---- adjoint stepping ----
    // r += f(stack) over active stacks (BuildRHSAdjoint adds its terms)
    template <typename F>
    void modal_add(std::vector<C>& r, F f) {
        per_active([&](int j1, int j2) {
            std::vector<C> t(N3);
            f(j1, j2, mi(j1, j2, 0), t.data());
            C* rr = &r[mi(j1, j2, 0)];
            for (int j = 0; j < N3; ++j) rr[j] += t[j];
        });
    }
    // out = ddx(in), ddy(in) (Stratiflow.h:52-66) or ddz(in) (Stratiflow.h:68-81,
    // Neumann in -> Dirichlet out and the reverse) of a modal field
    void modal_deriv(int axis, int bc_in, const std::vector<C>& in, std::vector<C>& out) {
        per_active([&](int j1, int j2) {
            const size_t o = mi(j1, j2, 0);
            if (axis == 2) {
                if (bc_in == NEUMANN) op_dz_ND(&in[o], &out[o]);
                else op_dz_DN(&in[o], &out[o]);
                return;
            }
            const C ik(0, axis == 0 ? kx[j1] : ky[j2]);
            for (int j = 0; j < N3; ++j) out[o + j] = ik * in[o + j];
        });
    }

    // UpdateAdjointVariables, IMEXRK.h:461-500 (the J integrand is a
    // diagnostic and is not formed)
    void update_adjoint_variables() {
        to_nodal(u1t, U1_tot);
        to_nodal(u2t, U2_tot);
        to_nodal(u3t, U3_tot);
        to_nodal(bt, B_tot);
        std::vector<T> bAve(N3), wAve(N3);  // HorizontalAverage, Integration.h:37-42
        for (int j = 0; j < N3; ++j) {
            bAve[j] = bt[mi(0, 0, j)].real();
            wAve[j] = u3t[mi(0, 0, j)].real();
        }
        const T K = T(2);
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) {
                u3F[o + j] = (T(1) / K) * (B_tot[o + j] + T(-1) * bAve[j]);
                bF[o + j] = (T(1) / K) * (U3_tot[o + j] + T(-1) * wAve[j]);
                u1F[o + j] = T(0);
                u2F[o + j] = T(0);
            }
        });
    }

    // BuildRHSAdjoint, IMEXRK.cpp:260-379
    void build_rhs_adjoint() {
        const T ri = T(Ri);
        per_nodal_stack([&](size_t o) {
            std::vector<T> t(N3);
            op_idir(&U3[o], t.data());
            for (int j = 0; j < N3; ++j) bF[o + j] += ri * t[j];
        });
        // advection of the adjoint fields by the direct flow
        prod_nn(U1, U1_tot, neumannTemp);
        modal_add(r1, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
        });
        prod_dbar(U1, U3_tot, neumannTemp);
        modal_add(r1, [&](int, int, size_t o, C* t) {
            for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
        });
        prod_tilde(U1_tot, U3, dirichletTemp);
        prod_dd(U3, U3_tot, neumannTemp);
        modal_add(r3, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            std::vector<C> dz(N3);
            op_dz_ND(&neumannTemp[o], dz.data());
            for (int j = 0; j < N3; ++j) t[j] = ikx * dirichletTemp[o + j] + dz[j];
        });
        if (third) {
            prod_nn(U2, U2_tot, neumannTemp);
            modal_add(r2, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
            });
            prod_dbar(U2, U3_tot, neumannTemp);
            modal_add(r2, [&](int, int, size_t o, C* t) {
                for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
            });
            prod_tilde(U2_tot, U3, dirichletTemp);
            modal_add(r3, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * dirichletTemp[o + j];
            });
            prod_nn(U1, U2_tot, neumannTemp);
            modal_add(r1, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
            });
            prod_nn(U2, U1_tot, neumannTemp);
            modal_add(r2, [&](int j1, int, size_t o, C* t) {
                const C ikx(0, kx[j1]);
                for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
            });
        }
        prod_dbar(B, U3_tot, neumannTemp);
        modal_add(rB, [&](int, int, size_t o, C* t) {
            for (int j = 0; j < N3; ++j) t[j] = neumannTemp[o + j];
        });
        if (third) {
            prod_nn(B, U2_tot, neumannTemp);
            modal_add(rB, [&](int, int j2, size_t o, C* t) {
                const C iky(0, ky[j2]);
                for (int j = 0; j < N3; ++j) t[j] = iky * neumannTemp[o + j];
            });
        }
        prod_nn(B, U1_tot, neumannTemp);
        modal_add(rB, [&](int j1, int, size_t o, C* t) {
            const C ikx(0, kx[j1]);
            for (int j = 0; j < N3; ++j) t[j] = ikx * neumannTemp[o + j];
        });

        // extra adjoint terms: gradients of the direct flow acting on the adjoint
        // horizontal derivatives (u1Forcing, and u2Forcing with a third dimension)
        for (int axis = 0; axis < (third ? 2 : 1); ++axis) {
            std::vector<T>& F = axis == 0 ? u1F : u2F;
            modal_deriv(axis, NEUMANN, u1t, neumannTemp);
            to_nodal(neumannTemp, nnTemp);
            per_nodal_stack([&](size_t o) {
                for (int j = 0; j < N3; ++j) nnTemp2[o + j] = nnTemp[o + j] * U1[o + j];
            });
            if (third) {
                modal_deriv(axis, NEUMANN, u2t, neumannTemp);
                to_nodal(neumannTemp, nnTemp);
                per_nodal_stack([&](size_t o) {
                    for (int j = 0; j < N3; ++j) nnTemp2[o + j] += nnTemp[o + j] * U2[o + j];
                });
            }
            modal_deriv(axis, DIRICHLET, u3t, dirichletTemp);
            to_nodal(dirichletTemp, ndTemp);
            per_nodal_stack([&](size_t o) {
                std::vector<T> pr(N3), ip(N3);
                for (int j = 0; j < N3; ++j) pr[j] = ndTemp[o + j] * U3[o + j];
                op_idir(pr.data(), ip.data());
                for (int j = 0; j < N3; ++j) F[o + j] -= nnTemp2[o + j] + ip[j];
            });
        }
        // vertical derivatives (u3Forcing)
        modal_deriv(2, NEUMANN, u1t, dirichletTemp);
        to_nodal(dirichletTemp, ndTemp);
        per_nodal_stack([&](size_t o) {
            std::vector<T> iu(N3);
            op_ifull(&U1[o], iu.data());
            for (int j = 0; j < N3; ++j) ndTemp2[o + j] = ndTemp[o + j] * iu[j];
        });
        if (third) {
            modal_deriv(2, NEUMANN, u2t, dirichletTemp);
            to_nodal(dirichletTemp, ndTemp);
            per_nodal_stack([&](size_t o) {
                std::vector<T> iu(N3);
                op_ifull(&U2[o], iu.data());
                for (int j = 0; j < N3; ++j) ndTemp2[o + j] += ndTemp[o + j] * iu[j];
            });
        }
        modal_deriv(2, DIRICHLET, u3t, neumannTemp);
        to_nodal(neumannTemp, nnTemp);
        per_nodal_stack([&](size_t o) {
            std::vector<T> in(N3);
            op_ifull(&nnTemp[o], in.data());
            for (int j = 0; j < N3; ++j) u3F[o + j] -= ndTemp2[o + j] + in[j] * U3[o + j];
        });
        // buoyancy gradients
        modal_deriv(0, NEUMANN, bt, neumannTemp);
        to_nodal(neumannTemp, nnTemp);
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) u1F[o + j] -= nnTemp[o + j] * B[o + j];
        });
        if (third) {
            modal_deriv(1, NEUMANN, bt, neumannTemp);
            to_nodal(neumannTemp, nnTemp);
            per_nodal_stack([&](size_t o) {
                for (int j = 0; j < N3; ++j) u2F[o + j] -= nnTemp[o + j] * B[o + j];
            });
        }
        modal_deriv(2, NEUMANN, bt, dirichletTemp);
        to_nodal(dirichletTemp, ndTemp);
        per_nodal_stack([&](size_t o) {
            std::vector<T> ib(N3);
            op_ifull(&B[o], ib.data());
            for (int j = 0; j < N3; ++j) u3F[o + j] -= ndTemp[o + j] * ib[j];
        });
        // all forcing terms
        auto add_modal = [&](std::vector<C>& r, const std::vector<T>& F, std::vector<C>& tmp) {
            to_modal(F, tmp);
            modal_add(r, [&](int, int, size_t o, C* t) {
                for (int j = 0; j < N3; ++j) t[j] = tmp[o + j];
            });
        };
        add_modal(r1, u1F, neumannTemp);
        if (third) add_modal(r2, u2F, neumannTemp);
        add_modal(r3, u3F, dirichletTemp);
        add_modal(rB, bF, neumannTemp);
    }

    // the direct state at fraction f of the backward step, with the background
    // velocity added through a nodal round trip (IMEXRK.h:96-107)
    void adjoint_totals(const std::vector<C>* const below[4], const std::vector<C>* const above[4], T f) {
        std::vector<C>* tot[4] = {&u1t, &u2t, &u3t, &bt};
        for (int i = 0; i < 4; ++i)
            for (size_t e = 0; e < nm; ++e) (*tot[i])[e] = (T(1) - f) * (*above[i])[e] + f * (*below[i])[e];
        to_nodal(u1t, U1_tot);
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) U1_tot[o + j] += Ubar[j];
        });
        to_modal(U1_tot, u1t);
    }
    // explicit terms alone, for tests: r = BuildRHSAdjoint with the nodal
    // fields as they are; the state is left untouched
    void adjoint_rhs(const std::vector<C>* const below[4], const std::vector<C>* const above[4], T f) {
        adjoint_totals(below, above, f);
        update_adjoint_variables();
        zero(r1); zero(r2); zero(r3); zero(rB);
        build_rhs_adjoint();
    }
    // r = BuildRHSLinear about the SetBackground state, for tests
    void linear_rhs() {
        zero(r1); zero(r2); zero(r3); zero(rB);
        build_rhs_linear();
    }

    // JoverK, IMEXRK.h:445-459, with the J integrand of UpdateAdjointVariables
    // (IMEXRK.h:470-490): the nodal buoyancy B and the modal u3 of the state
    T jover_k() {
        std::vector<C> btot(nm);
        std::vector<T> nd(nn);
        to_modal(B, btot);      // nnTemp = B; ToModal(b_total)
        to_nodal(btot, B_tot);  // b_total.ToNodal(B_tot)
        to_nodal(u3, U3_tot);   // u3_total.ToNodal(U3_tot)
        std::vector<T> bAve(N3);
        for (int j = 0; j < N3; ++j) bAve[j] = btot[mi(0, 0, j)].real();
        per_nodal_stack([&](size_t o) {
            for (int j = 0; j < N3; ++j) nd[o + j] = (T(-1) * (B_tot[o + j] + T(-1) * bAve[j])) * U3_tot[o + j];
        });
        to_modal(nd, dirichletTemp);
        // IntegrateVertically of the Dirichlet profile, Integration.h:22-34
        T J = 0;
        for (int k = 2; k < N3 - 1; ++k) J += T(g.zf[k] - g.zf[k - 1]) * dirichletTemp[mi(0, 0, k)].real();
        const T K = 2;
        return J / K;
    }

    // NodalField::Max (Field.h:771-786): the largest |value| over every z level
    // (ghost levels included) and every horizontal point; `add` is a profile
    // broadcast over the stacks first (U1 + U_ of IMEXRK.h:359)
    T nodal_max(const std::vector<T>& f, const std::vector<T>* add) const {
        T m = 0;
        for (size_t o = 0; o < nn; o += size_t(N3))
            for (int j = 0; j < N3; ++j) {
                const T v = add ? T(f[o + j] + (*add)[j]) : f[o + j];
                m = std::max(m, std::abs(v));
            }
        return m;
    }
    // the three grid spacings of CFL/CFLlinear (IMEXRK.h:359-364, 380-384):
    // L1/N1, L2/N2 and the smallest vertical gap z(N3/2+1) - z(N3/2)
    void cfl_spacings(T& d1, T& d2, T& d3) const {
        d1 = T(L1) / T(N1);
        d2 = T(L2) / T(N2);
        d3 = T(g.z[N3 / 2 + 1]) - T(g.z[N3 / 2]);
    }
    // IMEXRK::CFL (IMEXRK.h:355-374) before its dt update:
    //   dt * (max|U1 + U_|/d1 + max|U2|/d2 + max|U3|/d3)
    // from the nodal fields as they are (U2 stays zero in 2D)
    T cfl() const {
        T d1, d2, d3;
        cfl_spacings(d1, d2, d3);
        T c = nodal_max(U1, &Ubar) / d1 + nodal_max(U2, nullptr) / d2 + nodal_max(U3, nullptr) / d3;
        return c * T(deltaT);
    }
    // IMEXRK::CFLlinear (IMEXRK.h:376-393) before its dt update: the nodal
    // linearisation state of SetBackground(u1..b) (IMEXRK.h:334-352), with 1 as
    // the bound on |U_|:  dt * ((1 + max|U1_tot|)/d1 + max|U2_tot|/d2 + max|U3_tot|/d3)
    T cfl_linear() const {
        T d1, d2, d3;
        cfl_spacings(d1, d2, d3);
        T c = (T(1) + nodal_max(U1_tot, nullptr)) / d1 + nodal_max(U2_tot, nullptr) / d2 +
              nodal_max(U3_tot, nullptr) / d3;
        return c * T(deltaT);
    }

    // PrepareRunAdjoint, IMEXRK.h:183-199 (without the image folders)
    void prepare_adjoint() {
        zero(p);
        populate_nodal();
    }

    // TimeStepAdjoint, IMEXRK.h:86-126.  below/above: the direct state at the
    // end and at the start of the (backward) step, u1, u2, u3, b.
    void time_step_adjoint(const std::vector<C>* const below[4], const std::vector<C>* const above[4]) {
        T interpFrac = 0;
        for (int k = 0; k < 3; ++k) {
            adjoint_totals(below, above, interpFrac);
            update_adjoint_variables();
            explicit_rk(k);
            build_rhs_adjoint();
            finish_rhs(k);
            crank_nicolson(k);
            remove_divergence(1.0 / double(h[k]));
            filter_all();
            populate_nodal();
            interpFrac += h[k] / T(deltaT);
        }
    }

    // InnerProd (modal), Integration.h:117-146 with IntegrateVertically,
    // Integration.h:5-35.  Accumulates in T like the reference.
    T inner_prod(const std::vector<C>& a, const std::vector<C>& bb, int bc) const {
        T result = 0;
        for (int j3 = 0; j3 < N3; ++j3) {
            T sum = 0;
            for (int j1 = 0; j1 < M1; ++j1)
                for (int j2 = 0; j2 < N2; ++j2) {
                    const T v = std::real(a[mi(j1, j2, j3)] * std::conj(bb[mi(j1, j2, j3)]));
                    sum += (j1 == 0 && j2 == 0) ? v : T(2) * v;
                }
            if (bc == NEUMANN) {
                if (j3 >= 1 && j3 < N3 - 1) result += T(g.z[j3 + 1] - g.z[j3]) * sum;
            } else {
                if (j3 >= 2 && j3 < N3 - 1) result += T(g.zf[j3] - g.zf[j3 - 1]) * sum;
            }
        }
        return result;
    }
    T ke() const {  // IMEXRK.h:395-406
        T e = T(0.5) * (inner_prod(u1, u1, NEUMANN) + inner_prod(u3, u3, DIRICHLET));
        if (third) e += T(0.5) * inner_prod(u2, u2, NEUMANN);
        return e;
    }
    T pe() const { return T(Ri) * T(0.5) * inner_prod(b, b, NEUMANN); }  // IMEXRK.h:408-411

private:
    CFFT<T> fx, fy;
};

template <typename T>
constexpr double Solver<T>::beta[3];
template <typename T>
constexpr double Solver<T>::zeta[3];

}  // namespace sfo

// ---------------------------------------------------------------------------
// C ABI used by the Python test harness (ctypes).  All arrays are double in
// the reference layout: modal fields (M1*N2*N3 complex, interleaved re/im),
// nodal fields (N1*N2*N3), 1D fields (N3).
// ---------------------------------------------------------------------------
namespace {
struct Handle {
    int precision;
    std::unique_ptr<sfo::Solver<float>> f;
    std::unique_ptr<sfo::Solver<double>> d;
};

template <typename Fn>
auto dispatch(Handle* h, Fn fn) -> decltype(fn(*h->d)) {
    if (h->precision == 4) return fn(*h->f);
    return fn(*h->d);
}

template <typename S>
std::vector<std::complex<typename S::C::value_type>>* modal_by_id(S& s, int id) {
    switch (id) {
        case 0: return &s.u1;
        case 1: return &s.u2;
        case 2: return &s.u3;
        case 3: return &s.b;
        case 4: return &s.p;
        case 5: return &s.r1;
        case 6: return &s.r2;
        case 7: return &s.r3;
        case 8: return &s.rB;
        case 9: return &s.u1t;
        case 10: return &s.u2t;
        case 11: return &s.u3t;
        case 12: return &s.bt;
        case 13: return &s.q;
        case 14: return &s.divergence;
        default: return nullptr;
    }
}
template <typename S>
std::vector<typename S::C::value_type>* nodal_by_id(S& s, int id) {
    switch (id) {
        case 0: return &s.U1;
        case 1: return &s.U2;
        case 2: return &s.U3;
        case 3: return &s.B;
        default: return nullptr;
    }
}
}  // namespace

extern "C" {

void* sfo_create(int N1, int N2, int N3, int dim, double L1, double L2, double L3, double Re, double Ri, double Pr,
                 int precision) {
    try {
        auto* h = new Handle();
        h->precision = precision == 4 ? 4 : 8;
        if (h->precision == 4)
            h->f.reset(new sfo::Solver<float>(N1, N2, N3, dim, L1, L2, L3, Re, Ri, Pr));
        else
            h->d.reset(new sfo::Solver<double>(N1, N2, N3, dim, L1, L2, L3, Re, Ri, Pr));
        return h;
    } catch (...) {
        return nullptr;
    }
}

void sfo_destroy(void* p) { delete static_cast<Handle*>(p); }

void sfo_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

void sfo_set_dt(void* p, double dt) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.set_dt(dt); return 0; });
}

void sfo_set_flow(void* p, double Re, double Ri, double Pr) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.Re = Re; s.Ri = Ri; s.Pr = Pr; return 0; });
}

// U: N3 values of the Neumann1D background velocity; S: N3 values of dB/dz
// at Neumann points (NULL keeps 1).
void sfo_set_background(void* p, const double* U, const double* S) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        for (int j = 0; j < s.N3; ++j) {
            s.Ubar[j] = T(U[j]);
            s.S[j] = S ? T(S[j]) : T(1);
        }
        return 0;
    });
}

int sfo_set_modal(void* p, int id, const double* data) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        auto* f = modal_by_id(s, id);
        if (!f) return -1;
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        for (size_t i = 0; i < s.nm; ++i) (*f)[i] = std::complex<T>(T(data[2 * i]), T(data[2 * i + 1]));
        return 0;
    });
}

int sfo_get_modal(void* p, int id, double* data) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        auto* f = modal_by_id(s, id);
        if (!f) return -1;
        for (size_t i = 0; i < s.nm; ++i) {
            data[2 * i] = (*f)[i].real();
            data[2 * i + 1] = (*f)[i].imag();
        }
        return 0;
    });
}

int sfo_set_nodal(void* p, int id, const double* data) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        auto* f = nodal_by_id(s, id);
        if (!f) return -1;
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        for (size_t i = 0; i < s.nn; ++i) (*f)[i] = T(data[i]);
        return 0;
    });
}

int sfo_get_nodal(void* p, int id, double* data) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        auto* f = nodal_by_id(s, id);
        if (!f) return -1;
        for (size_t i = 0; i < s.nn; ++i) data[i] = (*f)[i];
        return 0;
    });
}

void sfo_filter_all(void* p) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.filter_all(); return 0; });
}
void sfo_populate_nodal(void* p) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.populate_nodal(); return 0; });
}
void sfo_remove_divergence(void* p, double mult) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.remove_divergence(mult); return 0; });
}
void sfo_set_linear_background(void* p) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.set_linear_background(); return 0; });
}
void sfo_stage(void* p, int k, int linear) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.stage(k, linear != 0); return 0; });
}
void sfo_step(void* p, int n) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { for (int i = 0; i < n; ++i) s.time_step(); return 0; });
}
void sfo_step_linear(void* p, int n) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { for (int i = 0; i < n; ++i) s.time_step_linear(); return 0; });
}
void sfo_prepare_adjoint(void* p) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { s.prepare_adjoint(); return 0; });
}
// One TimeStepAdjoint (mode 0), or only the explicit terms of the adjoint at
// fraction frac (mode 1) or of the linearised equations (mode 2) into r1..rB.
// below/above: 4 modal fields each (u1, u2, u3, b) in the reference layout;
// a null field is zero.
void sfo_adjoint(void* p, int mode, const double* const* below, const double* const* above, double frac) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        if (mode == 2) {
            s.linear_rhs();
            return 0;
        }
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        using C = std::complex<T>;
        std::vector<C> bl[4], ab[4];
        const std::vector<C>* pb[4];
        const std::vector<C>* pa[4];
        for (int i = 0; i < 4; ++i) {
            bl[i].assign(s.nm, C(0));
            ab[i].assign(s.nm, C(0));
            for (size_t e = 0; e < s.nm; ++e) {
                if (below[i]) bl[i][e] = C(T(below[i][2 * e]), T(below[i][2 * e + 1]));
                if (above[i]) ab[i][e] = C(T(above[i][2 * e]), T(above[i][2 * e + 1]));
            }
            pb[i] = &bl[i];
            pa[i] = &ab[i];
        }
        using T2 = typename std::decay<decltype(s)>::type::C::value_type;
        if (mode == 1) s.adjoint_rhs(pb, pa, T2(frac));
        else s.time_step_adjoint(pb, pa);
        return 0;
    });
}
double sfo_jover_k(void* p) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) { return double(s.jover_k()); });
}
// FlowParams::EvolveBackground: U_ diffuses by one Crank-Nicolson solve per
// stage (IMEXRK.cpp:100-104)
void sfo_set_evolve_background(void* p, int on) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        s.evolveBackground = on != 0;
        return 0;
    });
}
// the current U_ (N3 values)
void sfo_get_background(void* p, double* U) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (int j = 0; j < s.N3; ++j) U[j] = double(s.Ubar[j]);
        return 0;
    });
}
// CFL (linear == 0, IMEXRK.h:355-374) or CFLlinear (IMEXRK.h:376-393) for the
// current dt, without its dt update
double sfo_cfl(void* p, int linear) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) { return double(linear ? s.cfl_linear() : s.cfl()); });
}
void sfo_energy(void* p, double* ke, double* pe) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) { *ke = s.ke(); *pe = s.pe(); return 0; });
}

// InnerProd of two modal arrays given in the reference layout
double sfo_inner_prod(void* p, int bc, const double* a, const double* b) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        std::vector<std::complex<T>> A(s.nm), B(s.nm);
        for (size_t i = 0; i < s.nm; ++i) {
            A[i] = std::complex<T>(T(a[2 * i]), T(a[2 * i + 1]));
            B[i] = std::complex<T>(T(b[2 * i]), T(b[2 * i + 1]));
        }
        return double(s.inner_prod(A, B, bc));
    });
}

void sfo_to_modal(void* p, const double* nodal, double* modal) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        std::vector<T> in(s.nn);
        std::vector<std::complex<T>> out(s.nm);
        for (size_t i = 0; i < s.nn; ++i) in[i] = T(nodal[i]);
        s.to_modal(in, out);
        for (size_t i = 0; i < s.nm; ++i) {
            modal[2 * i] = out[i].real();
            modal[2 * i + 1] = out[i].imag();
        }
        return 0;
    });
}

void sfo_to_nodal(void* p, const double* modal, double* nodal) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        std::vector<std::complex<T>> in(s.nm);
        std::vector<T> out(s.nn);
        for (size_t i = 0; i < s.nm; ++i) in[i] = std::complex<T>(T(modal[2 * i]), T(modal[2 * i + 1]));
        s.to_nodal(in, out);
        for (size_t i = 0; i < s.nn; ++i) nodal[i] = out[i];
        return 0;
    });
}

// Apply one banded vertical operator to a real column of N3 values:
// 0 Dz N->D, 1 Dz D->N, 2 ReinterpolateFull, 3 ReinterpolateTilde,
// 4 ReinterpolateBar, 5 ReinterpolateDirichlet, 6 Dzz Neumann, 7 Dzz Dirichlet
int sfo_apply_op(void* p, int op, const double* in, double* out) {
    return dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using T = typename std::decay<decltype(s)>::type::C::value_type;
        std::vector<T> f(s.N3), o(s.N3);
        for (int j = 0; j < s.N3; ++j) f[j] = T(in[j]);
        switch (op) {
            case 0: s.op_dz_ND(f.data(), o.data()); break;
            case 1: s.op_dz_DN(f.data(), o.data()); break;
            case 2: s.op_ifull(f.data(), o.data()); break;
            case 3: s.op_itilde(f.data(), o.data()); break;
            case 4: s.op_ibar(f.data(), o.data()); break;
            case 5: s.op_idir(f.data(), o.data()); break;
            case 6: s.op_dzz(sfo::NEUMANN, f.data(), o.data()); break;
            case 7: s.op_dzz(sfo::DIRICHLET, f.data(), o.data()); break;
            default: return -1;
        }
        for (int j = 0; j < s.N3; ++j) out[j] = o[j];
        return 0;
    });
}

// z, zf, DY, DYF (each N3 values, double)
void sfo_grid(void* p, double* z, double* zf, double* DY, double* DYF) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        for (int j = 0; j < s.N3; ++j) {
            z[j] = s.g.z[j]; zf[j] = s.g.zf[j]; DY[j] = s.g.DY[j]; DYF[j] = s.g.DYF[j];
        }
        return 0;
    });
}

// Solve one Thomas system of the given kind (0 vel-N, 1 vel-D, 2 buoy-N,
// 3 Laplacian) for mode (j1,j2) at RK stage k; d and x are N3 complex.
// Also returns the three diagonals (before Tridiagonal::compute) if non-NULL.
void sfo_tri_solve(void* p, int kind, int stage, int j1, int j2, const double* d, double* x, double* a, double* bdiag,
                   double* c) {
    dispatch(static_cast<Handle*>(p), [&](auto& s) {
        using S = typename std::decay<decltype(s)>::type;
        using T = typename S::C::value_type;
        typename S::Tri raw, t;
        s.build_raw(kind, stage, j1, j2, raw);
        s.build_tri(kind, stage, j1, j2, t);
        std::vector<std::complex<T>> D(s.N3), X(s.N3);
        for (int j = 0; j < s.N3; ++j) D[j] = std::complex<T>(T(d[2 * j]), T(d[2 * j + 1]));
        s.tri_solve(t, D.data(), X.data());
        for (int j = 0; j < s.N3; ++j) {
            x[2 * j] = X[j].real();
            x[2 * j + 1] = X[j].imag();
            if (a) a[j] = raw.a[j];
            if (bdiag) bdiag[j] = raw.b[j];
            if (c) c[j] = raw.c[j];
        }
        return 0;
    });
}

}  // extern "C"
