// sf_api.hip -- context, fields and the extern "C" boundary (include/sfhip.h).
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sfhip.h"
#include "sf_internal.h"
#include "sf_kernels.h"

namespace sf {

// ---------------------------------------------------------------------------
// Vertical grid (Field.cpp:3-31, 62-83) in double.  z[0] and DY[0], which
// the reference leaves uninitialised, get extrapolated values.
// ---------------------------------------------------------------------------
struct VGrid {
    std::vector<double> z, zf, DY, DYF;
};

static VGrid vertical_grid(double L, int N) {
    VGrid g;
    g.z.assign(N, 0.0);
    g.zf.assign(N, 0.0);
    g.DY.assign(N, 0.0);
    g.DYF.assign(N, 0.0);
    const int M = N - 1;  // points 1..N-1 from LinSpaced(N-1, -1, 1)
    const double step = 2.0 / double(M - 1);
    for (int t = 0; t < M; ++t) {
        const double x = (t < M / 2) ? -1.0 + t * step : 1.0 - (M - 1 - t) * step;
        g.z[1 + t] = L * (2.0 * std::pow(x, 7) + x) / 3.0;
    }
    const double scale = -L / (0.5 * (g.z[1] + g.z[2]));
    for (int t = 1; t < N; ++t) g.z[t] *= scale;
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

// integer wavenumber of Fourier index j (Differentiation.cpp:113-132)
static int kindex(int n, int j) {
    if (n == 1) return 0;
    if (j <= n / 3) return j;
    if (j < n - n / 3) return 0;
    return j - n;
}

struct FieldBase {
    virtual ~FieldBase() {}
};

}  // namespace sf

struct sf_batch {
    int precision;
    std::string err;
    std::shared_ptr<void> impl;  // sf::Batch<T>
    void* raw;
};

namespace sf {

static thread_local std::string g_create_error;

template <typename T>
struct Ctx {
    using V = cx<T>;
    sf_ctx* owner = nullptr;
    int N1 = 0, N2 = 0, N3 = 0, M1 = 0, A1 = 0, A2 = 0, dim = 0, N3p = 0;
    bool three = false;
    int zch = 4;   // z levels per x-pass block
    int device = 0;  // the device current at creation
    double L1 = 0, L2 = 0, L3 = 0, Re = 0, Ri = 0, Pr = 0;
    bool evolve_bg = false;  // FlowParams::EvolveBackground
    double cfl_d[3] = {1, 1, 1};  // L1/N1, L2/N2, z(N3/2+1) - z(N3/2) (IMEXRK.h:359-364)
    double dt = 0.01;
    T h[3] = {0, 0, 0};
    size_t nmodal = 0;  // complex values per modal field
    size_t nnodal = 0;  // real values per nodal field
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;

    T* zc = nullptr;
    T *kx = nullptr, *ky = nullptr, *kx2 = nullptr, *ky2 = nullptr;
    V *tw1 = nullptr, *tw2 = nullptr;
    V* tw1_r8 = nullptr;  // x twiddles of the plan with first radix 8 (made on first use)
    double* partial = nullptr;
    double* scal = nullptr;
    double* scal_host = nullptr;
    std::vector<double> zc_host;  // double copy of the coefficient table

    sf_field state[17];
    V* rbuf[2][4] = {{nullptr}};
    int rcur = 0;
    V* nl[4] = {nullptr};
    V* Y[4] = {nullptr};
    V* Yb[4] = {nullptr};
    V* X[8] = {nullptr};
    // X[] (the 3D x-pass outputs) in tiles of 16 modes where the paired
    // forward y transform reads them as contiguous runs (yfwd_tiles)
    bool ytiles = false;
    // Spectra the nodal fields were last populated from.  The reference's
    // products read the nodal fields U1..B, which PopulateNodalVariables
    // refreshes at the end of every stage but which RemoveDivergence does not
    // touch (StateVector.cpp:9-11 populates before projecting).  After an
    // explicit sf_populate_nodal the next stage reads these copies; every
    // stage ends with nodal fields consistent with the state again.
    V* snap[4] = {nullptr};
    bool use_snap = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // per-launch events for sf_time_steps_detailed
    bool prof = false;
    std::vector<hipEvent_t> pev;
    std::vector<int> pcls;
    size_t pn = 0;
    void mark(int cls) {
        if (!prof) return;
        if (pn >= pev.size()) {
            hipEvent_t e;
            SF_HIP(hipEventCreate(&e));
            pev.push_back(e);
            pcls.push_back(-1);
        }
        SF_HIP(hipEventRecord(pev[pn], stream));
        pcls[pn] = cls;
        ++pn;
    }

    template <typename P>
    P* dalloc(size_t count) {
        void* p = nullptr;
        const size_t bytes = count * sizeof(P);
        SF_HIP(hipMalloc(&p, bytes > 0 ? bytes : 16));
        allocs.push_back(p);
        SF_HIP(hipMemsetAsync(p, 0, bytes > 0 ? bytes : 16, stream));
        return static_cast<P*>(p);
    }

    ~Ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void* p : allocs) (void)hipFree(p);
        if (scal_host) (void)hipHostFree(scal_host);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (hipEvent_t e : pev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void init(const sf_grid& g, const sf_flow& f) {
        N1 = g.N1;
        N2 = g.N2;
        N3 = g.N3;
        dim = g.dimensionality;
        SF_REQUIRE(dim == SF_THREE_D || dim == SF_TWO_D || dim == SF_TWO_AND_A_HALF, SF_ERR_ARG,
                   "dimensionality must be SF_THREE_D, SF_TWO_D or SF_TWO_AND_A_HALF");
        // TwoAndAHalf resolves the spanwise direction like ThreeDimensional
        // (GridParams::ThirdDimension, Parameters.h:13-17)
        three = (dim != SF_TWO_D);
        SF_REQUIRE(N3 >= 8 && N3 % 4 == 0 && N3 <= 64 * max_rows_per_lane(), SF_ERR_ARG,
                   "N3 must be a multiple of 4 in [8, 512] (Field.cpp:7)");
        SF_REQUIRE(tz_supported(N1), SF_ERR_UNSUPPORTED, "unsupported N1 (FFT sizes: 6..512 of the form 2^a*3^b)");
        if (three) {
            SF_REQUIRE(N2 >= 6 && tz_supported(N2), SF_ERR_UNSUPPORTED, "unsupported N2 for 3D");
        } else {
            SF_REQUIRE(N2 == 1, SF_ERR_ARG, "2D requires N2 == 1 (IMEXRK.h:43)");
        }
        SF_REQUIRE(!(N1 > 384 && sizeof(T) == 8), SF_ERR_UNSUPPORTED, "FP64 contexts support N1 <= 384");
        SF_HIP(hipGetDevice(&device));
        M1 = N1 / 2 + 1;
        A1 = N1 / 3;
        A2 = three ? 2 * (N2 / 3) - 1 : 1;
        N3p = N3;
        zch = xpass_levels(N1, N2, N3, three, int(sizeof(T)));
        nmodal = size_t(A1) * A2 * N3;
        nnodal = size_t(N1) * N2 * N3;
        SF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        SF_HIP(hipEventCreate(&ev0));
        SF_HIP(hipEventCreate(&ev1));
        SF_HIP(hipHostMalloc((void**)&scal_host, 16 * sizeof(double), hipHostMallocDefault));

        zc = dalloc<T>(size_t(ZC_COUNT) * N3p);
        kx = dalloc<T>(A1);
        kx2 = dalloc<T>(A1);
        ky = dalloc<T>(A2);
        ky2 = dalloc<T>(A2);
        tw1 = dalloc<V>(2 * N1);
        tw2 = dalloc<V>(2 * N2);
        partial = dalloc<double>(2048);
        scal = dalloc<double>(16);

        // solver fields
        const int mk[17][2] = {{SF_MODAL, SF_NEUMANN},   {SF_MODAL, SF_NEUMANN},   {SF_MODAL, SF_DIRICHLET},
                               {SF_MODAL, SF_NEUMANN},   {SF_MODAL, SF_NEUMANN},   {SF_NODAL, SF_NEUMANN},
                               {SF_NODAL, SF_NEUMANN},   {SF_NODAL, SF_DIRICHLET}, {SF_NODAL, SF_NEUMANN},
                               {SF_MODAL, SF_NEUMANN},   {SF_MODAL, SF_NEUMANN},   {SF_MODAL, SF_DIRICHLET},
                               {SF_MODAL, SF_NEUMANN},   {SF_MODAL, SF_NEUMANN},   {SF_MODAL, SF_NEUMANN},
                               {SF_MODAL, SF_DIRICHLET}, {SF_MODAL, SF_NEUMANN}};
        for (int i = 0; i < 17; ++i) {
            sf_field& s = state[i];
            s.ctx = owner;
            s.kind = mk[i][0];
            s.bc = mk[i][1];
            s.owned = false;
            s.n = s.kind == SF_MODAL ? 2 * nmodal : nnodal;
            if (i >= SF_R1 && i <= SF_RB) {
                s.d = nullptr;
                continue;
            }
            s.d = (s.kind == SF_MODAL) ? (void*)dalloc<V>(nmodal) : (void*)dalloc<T>(nnodal);
        }
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < 4; ++i) rbuf[k][i] = dalloc<V>(nmodal);
        rcur = 0;
        bind_r();
        for (int i = 0; i < 4; ++i) nl[i] = dalloc<V>(nmodal);
        if (three) {
            const size_t ny = size_t(N2) * A1 * N3;
            for (int i = 0; i < 4; ++i) Y[i] = dalloc<V>(ny);
            for (int i = 0; i < 8; ++i) X[i] = dalloc<V>(ny);
            ytiles = yfwd_tiles(N2, A1, int(sizeof(T)));
        }
        set_flow(f);
        std::vector<double> U(N3, 0.0);
        set_background(U.data(), nullptr);
        set_dt(0.01);
        SF_HIP(hipStreamSynchronize(stream));
    }

    void bind_r() {
        for (int i = 0; i < 4; ++i) state[SF_R1 + i].d = rbuf[rcur][i];
    }

    void upload_coefficients() {
        std::vector<T> h(zc_host.size());
        for (size_t i = 0; i < h.size(); ++i) h[i] = T(zc_host[i]);
        SF_HIP(hipMemcpyAsync(zc, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream));
        SF_HIP(hipStreamSynchronize(stream));
    }

    void set_flow(const sf_flow& f) {
        SF_REQUIRE(f.L1 > 0 && f.L2 > 0 && f.L3 > 0 && f.Re > 0 && f.Pr > 0, SF_ERR_ARG, "bad flow parameters");
        // the device holds U_ as it evolves; keep it across the re-upload below
        if (zc && evolve_bg) fetch_background();
        evolve_bg = f.evolve_background != 0;
        L1 = f.L1;
        L2 = f.L2;
        L3 = f.L3;
        Re = f.Re;
        Ri = f.Ri;
        Pr = f.Pr;
        // vertical operators (Differentiation.cpp:7-111), banded
        const VGrid g = vertical_grid(L3, N3);
        cfl_d[0] = L1 / N1;
        cfl_d[1] = L2 / N2;
        cfl_d[2] = g.z[N3 / 2 + 1] - g.z[N3 / 2];
        std::vector<double> old = zc_host;
        zc_host.assign(size_t(ZC_COUNT) * N3p, 0.0);
        auto Z = [&](int c, int j) -> double& { return zc_host[size_t(c) * N3p + j]; };
        const int N = N3;
        for (int j = 1; j < N - 1; ++j) {
            const double lo = 1.0 / g.DY[j] / g.DYF[j], up = 1.0 / g.DY[j + 1] / g.DYF[j];
            Z(ZC_DZZN_LO, j) = lo;
            Z(ZC_DZZN_UP, j) = up;
            Z(ZC_DZZN_DI, j) = -lo - up;
        }
        for (int j = 2; j < N - 1; ++j) {
            const double lo = 1.0 / g.DYF[j - 1] / g.DY[j], up = 1.0 / g.DYF[j] / g.DY[j];
            Z(ZC_DZZD_LO, j) = lo;
            Z(ZC_DZZD_UP, j) = up;
            Z(ZC_DZZD_DI, j) = -lo - up;
        }
        for (int j = 2; j <= N - 1; ++j) Z(ZC_DZND, j) = 1.0 / g.DY[j];
        for (int j = 1; j <= N - 1; ++j) Z(ZC_DZDN, j) = 1.0 / g.DYF[j];
        for (int j = 1; j < N; ++j) {
            Z(ZC_IFULL_LO, j) = g.DYF[j] / (2.0 * g.DY[j]);
            Z(ZC_IFULL_DI, j) = g.DYF[j - 1] / (2.0 * g.DY[j]);
            Z(ZC_ITIL_LO, j) = g.DYF[j - 1] / (2.0 * g.DY[j]);
            Z(ZC_ITIL_DI, j) = g.DYF[j] / (2.0 * g.DY[j]);
        }
        // InnerProd weights, IntegrateVertically (Integration.h:5-35)
        for (int j = 1; j < N - 1; ++j) Z(ZC_WN, j) = g.z[j + 1] - g.z[j];
        for (int j = 2; j < N - 1; ++j) Z(ZC_WD, j) = g.zf[j] - g.zf[j - 1];
        // keep the background profiles
        for (int j = 0; j < N; ++j) {
            Z(ZC_UBAR, j) = old.empty() ? 0.0 : old[size_t(ZC_UBAR) * N3p + j];
            Z(ZC_S, j) = old.empty() ? 1.0 : old[size_t(ZC_S) * N3p + j];
        }
        upload_coefficients();
        // Fourier wavenumbers of the active modes (Differentiation.cpp:113-160)
        std::vector<T> hkx(A1), hkx2(A1), hky(A2), hky2(A2);
        for (int j = 0; j < A1; ++j) {
            const double k = 2.0 * M_PI * kindex(N1, j) / L1;
            hkx[j] = T(k);
            hkx2[j] = T(-k * k);
        }
        for (int ja = 0; ja < A2; ++ja) {
            const int j2 = three ? ((ja < N2 / 3) ? ja : ja + N2 - 2 * (N2 / 3) + 1) : 0;
            const double k = 2.0 * M_PI * kindex(N2, j2) / L2;
            hky[ja] = T(k);
            hky2[ja] = T(-k * k);
        }
        SF_HIP(hipMemcpyAsync(kx, hkx.data(), A1 * sizeof(T), hipMemcpyHostToDevice, stream));
        SF_HIP(hipMemcpyAsync(kx2, hkx2.data(), A1 * sizeof(T), hipMemcpyHostToDevice, stream));
        SF_HIP(hipMemcpyAsync(ky, hky.data(), A2 * sizeof(T), hipMemcpyHostToDevice, stream));
        SF_HIP(hipMemcpyAsync(ky2, hky2.data(), A2 * sizeof(T), hipMemcpyHostToDevice, stream));
        // FFT twiddles, computed in double
        auto tw = [&](int n, V* dst) {
            std::vector<double> re(2 * n), im(2 * n);
            fft_twiddles(n, re.data(), im.data());
            std::vector<V> w(2 * n);
            for (int k = 0; k < 2 * n; ++k) {
                w[k].x = T(re[k]);
                w[k].y = T(im[k]);
            }
            SF_HIP(hipMemcpyAsync(dst, w.data(), 2 * n * sizeof(V), hipMemcpyHostToDevice, stream));
            SF_HIP(hipStreamSynchronize(stream));
        };
        tw(N1, tw1);
        tw(N2, tw2);
        SF_HIP(hipStreamSynchronize(stream));
    }

    // x-direction twiddles for the FFT plan with first radix r1f (0: default)
    const V* x_twiddles(int r1f) {
        if (r1f == 0 || fft_first_radix(N1, r1f) == fft_first_radix(N1, 0)) return tw1;
        SF_REQUIRE(r1f == 8, SF_ERR_ARG, "x FFT plans: first radix 0 (default) or 8");
        if (!tw1_r8) {
            std::vector<double> re(2 * N1), im(2 * N1);
            fft_twiddles_plan(N1, 8, re.data(), im.data());
            std::vector<V> w(2 * N1);
            for (int k = 0; k < 2 * N1; ++k) {
                w[k].x = T(re[k]);
                w[k].y = T(im[k]);
            }
            V* d = dalloc<V>(2 * N1);
            SF_HIP(hipMemcpyAsync(d, w.data(), 2 * N1 * sizeof(V), hipMemcpyHostToDevice, stream));
            SF_HIP(hipStreamSynchronize(stream));
            tw1_r8 = d;
        }
        return tw1_r8;
    }

    // U_ from the device into the host copy of the coefficient table
    void fetch_background() {
        std::vector<T> u(N3);
        SF_HIP(hipMemcpyAsync(u.data(), zc + size_t(ZC_UBAR) * N3p, N3 * sizeof(T), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        for (int j = 0; j < N3; ++j) zc_host[size_t(ZC_UBAR) * N3p + j] = double(u[j]);
    }

    void set_background(const double* U, const double* S) {
        for (int j = 0; j < N3; ++j) {
            zc_host[size_t(ZC_UBAR) * N3p + j] = U ? U[j] : 0.0;
            zc_host[size_t(ZC_S) * N3p + j] = S ? S[j] : 1.0;
        }
        upload_coefficients();
    }

    // UpdateForTimestep (IMEXRK.h:506-508); the implicit matrices are rebuilt
    // inside k_modal, so nothing else is refactorised.
    void set_dt(double d) {
        SF_REQUIRE(d > 0 && std::isfinite(d), SF_ERR_ARG, "dt must be positive");
        dt = d;
        h[0] = T(d * 8.0 / 15.0);
        h[1] = T(d * 2.0 / 15.0);
        h[2] = T(d * 5.0 / 15.0);
    }

    V* mf(int i) { return static_cast<V*>(state[i].d); }

    void check(hipError_t e, const char* what) {
        if (e != hipSuccess) throw Error(SF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }

    const V* src(int i) { return use_snap ? snap[i] : mf(SF_U1 + i); }


    // nonlinear (or linearised) explicit terms of the current state into nl[]
    void transform_path(bool linear) {
        if (three) {
            YArgs<T> a{};
            a.N3 = N3;
            a.A1 = A1;
            a.A2 = A2;
            a.nf = 4;
            for (int i = 0; i < 4; ++i) a.in[i] = src(i);
            for (int i = 0; i < 4; ++i) a.out[i] = Y[i];
            a.tw = tw2;
            a.scale = T(1);
            check(launch_yinv<T>(a, N2, stream), "yinv");
            mark(SF_K_YINV);
            if (linear) {
                if (!Yb[0])
                    for (int i = 0; i < 4; ++i) Yb[i] = dalloc<V>(size_t(N2) * A1 * N3);
                a.in[0] = mf(SF_U1_TOT);
                a.in[1] = mf(SF_U2_TOT);
                a.in[2] = mf(SF_U3_TOT);
                a.in[3] = mf(SF_B_TOT);
                for (int i = 0; i < 4; ++i) a.out[i] = Yb[i];
                check(launch_yinv<T>(a, N2, stream), "yinv(base)");
                mark(SF_K_YINV);
            }
            XArgs<T> x{};
            x.N1 = N1;
            x.N3 = N3;
            x.N3p = N3p;
            x.A1 = A1;
            x.zch = zch;
            for (int i = 0; i < 4; ++i) x.in[i] = Y[i];
            for (int i = 0; i < 8; ++i) x.out[i] = X[i];
            x.zc = zc;
            x.kx = kx;
            x.tw = tw1;
            x.scale = T(1);
            x.linear = linear ? 1 : 0;
            x.N2 = N2;
            x.tiled = ytiles ? 1 : 0;
            for (int i = 0; i < 4; ++i) x.base[i] = linear ? Yb[i] : nullptr;
            check(launch_xpass<T>(x, N1, N2, true, stream), "xpass");
            mark(SF_K_XPASS);
            YArgs<T> f{};
            f.N3 = N3;
            f.A1 = A1;
            f.A2 = A2;
            f.nf = 4;
            f.paired = 1;
            f.tiled = ytiles ? 1 : 0;
            for (int r = 0; r < 4; ++r) {
                f.in[2 * r] = X[r];
                f.in[2 * r + 1] = X[4 + r];
                f.out[r] = nl[r];
            }
            f.ky = ky;
            f.tw = tw2;
            f.scale = T(1.0 / (double(N1) * double(N2)));
            check(launch_yfwd<T>(f, N2, stream), "yfwd");
            mark(SF_K_YFWD);
        } else {
            const XArgs<T> x = xargs_2d(linear);
            check(launch_xpass<T>(x, N1, 1, false, stream), "xpass");
            mark(SF_K_XPASS);
        }
    }

    // arguments of the 2D x-pass (IMEXRK.cpp:136-193, or :195-258 when linear);
    // from_state: read the prognostic fields even while the populated nodal
    // snapshot is pending (for arguments reused by every later step)
    XArgs<T> xargs_2d(bool linear, bool from_state = false) {
        auto src = [&](int i) -> const V* { return from_state ? mf(SF_U1 + i) : this->src(i); };
        XArgs<T> x{};
        x.N1 = N1;
        x.N3 = N3;
        x.N3p = N3p;
        x.A1 = A1;
        x.zch = zch;
        x.in[0] = src(0);
        x.in[1] = nullptr;
        x.in[2] = src(2);
        x.in[3] = src(3);
        x.out[0] = nl[0];
        x.out[1] = nl[1];
        x.out[2] = nl[2];
        x.out[3] = nl[3];
        for (int i = 4; i < 8; ++i) x.out[i] = nullptr;
        x.zc = zc;
        x.kx = kx;
        x.tw = tw1;
        x.scale = T(1.0 / double(N1));
        x.linear = linear ? 1 : 0;
        x.base[0] = mf(SF_U1_TOT);
        x.base[1] = nullptr;
        x.base[2] = mf(SF_U3_TOT);
        x.base[3] = mf(SF_B_TOT);
        return x;
    }

    // arguments of the modal kernel for stage k with the explicit terms of the
    // previous stage in rbuf[rc] (IMEXRK.cpp:8-24; ExplicitRK :119-134,
    // FinishRHS :107-117, CrankNicolson :75-105, RemoveDivergence :43-73)
    ModalArgs<T> stage_modal_args(int k, bool store_r, int rc) {
        static const double beta[3] = {1.0, 25.0 / 8.0, 9.0 / 4.0};    // IMEXRK.h:631
        static const double zeta[3] = {0.0, -17.0 / 8.0, -5.0 / 4.0};  // IMEXRK.h:632
        ModalArgs<T> m = modal_args();
        const int nxt = 1 - rc;
        const bool read_old = (k > 0);
        m.ro1 = read_old ? rbuf[rc][0] : nullptr;
        m.ro2 = (read_old && three) ? rbuf[rc][1] : nullptr;
        m.ro3 = read_old ? rbuf[rc][2] : nullptr;
        m.roB = read_old ? rbuf[rc][3] : nullptr;
        m.rn1 = store_r ? rbuf[nxt][0] : nullptr;
        m.rn2 = (store_r && three) ? rbuf[nxt][1] : nullptr;
        m.rn3 = store_r ? rbuf[nxt][2] : nullptr;
        m.rnB = store_r ? rbuf[nxt][3] : nullptr;
        m.n1 = nl[0];
        m.n2 = three ? nl[1] : nullptr;
        m.n3 = nl[2];
        m.nB = nl[3];
        m.hz = h[k] * T(zeta[k]);
        m.mh = -h[k];
        m.hb = h[k] * T(beta[k]);
        m.cv = (T(0.5) * h[k]) / T(Re);
        m.cb = ((T(0.5) * h[k]) / T(Re)) / T(Pr);
        m.Ri = T(Ri);
        m.pmult = T(1) / h[k];
        m.mode = 0;
        // CrankNicolson evolves U_ whenever the flag is set (IMEXRK.cpp:100-104),
        // in the linearised and adjoint stages as well
        m.ubar = evolve_bg ? zc + size_t(ZC_UBAR) * N3p : nullptr;
        return m;
    }

    ModalArgs<T> modal_args() {
        ModalArgs<T> m{};
        m.N3 = N3;
        m.N3p = N3p;
        m.A1 = A1;
        m.A2 = A2;
        m.nmodes = A1 * A2;
        m.threeD = three ? 1 : 0;
        m.zc = zc;
        m.kx = kx;
        m.ky = ky;
        m.kx2 = kx2;
        m.ky2 = ky2;
        m.u1 = mf(SF_U1);
        m.u2 = three ? mf(SF_U2) : nullptr;
        m.u3 = mf(SF_U3);
        m.b = mf(SF_B);
        m.p = mf(SF_P);
        m.ubar = nullptr;
        return m;
    }

    // one stage k; IMEXRK.cpp:8-24 (and :26-41 when linear)
    void stage(int k, bool linear, bool store_r) {
        transform_path(linear);
        ModalArgs<T> m = stage_modal_args(k, store_r, rcur);
        const int nxt = 1 - rcur;
        check(launch_modal<T>(m, stream), "modal");
        mark(SF_K_MODAL);
        use_snap = false;
        if (store_r) {
            rcur = nxt;
            bind_r();
        }
    }

    void step(int n, bool linear) {
        for (int s = 0; s < n; ++s)
            for (int k = 0; k < 3; ++k) stage(k, linear, k < 2);
    }

    void remove_divergence(double mult) {
        ModalArgs<T> m = modal_args();
        m.pmult = T(mult);
        m.mode = 1;
        check(launch_modal<T>(m, stream), "remove_divergence");
    }

    // ModalField::ToNodal (Field.h:965-994)
    void to_nodal(const V* modal, T* nodal) {
        XTArgs<T> x{};
        x.N3 = N3;
        x.A1 = A1;
        x.tw = tw1;
        x.nodal_out = nodal;
        x.scale = T(1);
        x.nzt = 16;
        if (three) {
            YArgs<T> a{};
            a.N3 = N3;
            a.A1 = A1;
            a.A2 = A2;
            a.nf = 1;
            a.in[0] = modal;
            a.out[0] = X[0];
            a.tw = tw2;
            a.scale = T(1);
            check(launch_yinv<T>(a, N2, stream), "to_nodal:y");
            x.spec_in = X[0];
        } else {
            x.spec_in = modal;
        }
        check(launch_x_c2r<T>(x, N1, N2, stream), "to_nodal:x");
    }

    // NodalField::ToModal (Field.h:734-769): scale 1/(N1*N2), keep active modes
    void to_modal(const T* nodal, V* modal) {
        XTArgs<T> x{};
        x.N3 = N3;
        x.A1 = A1;
        x.tw = tw1;
        x.nodal_in = nodal;
        x.nzt = 16;
        if (three) {
            x.spec_out = X[0];
            x.scale = T(1);
            check(launch_x_r2c<T>(x, N1, N2, stream), "to_modal:x");
            YArgs<T> a{};
            a.N3 = N3;
            a.A1 = A1;
            a.A2 = A2;
            a.nf = 1;
            a.paired = 0;
            a.in[0] = X[0];
            a.out[0] = modal;
            a.tw = tw2;
            a.ky = ky;
            a.scale = T(1.0 / (double(N1) * double(N2)));
            check(launch_yfwd<T>(a, N2, stream), "to_modal:y");
        } else {
            x.spec_out = modal;
            x.scale = T(1.0 / double(N1));
            check(launch_x_r2c<T>(x, N1, N2, stream), "to_modal:x");
        }
    }

    void populate_nodal() {
        for (int i = 0; i < 4; ++i) {
            if (i == 1 && !three) continue;
            if (!snap[i]) snap[i] = dalloc<V>(nmodal);
            SF_HIP(hipMemcpyAsync(snap[i], mf(SF_U1 + i), nmodal * sizeof(V), hipMemcpyDeviceToDevice, stream));
        }
        use_snap = true;
        to_nodal(mf(SF_U1), static_cast<T*>(state[SF_NODAL_U1].d));
        if (three) to_nodal(mf(SF_U2), static_cast<T*>(state[SF_NODAL_U2].d));
        to_nodal(mf(SF_U3), static_cast<T*>(state[SF_NODAL_U3].d));
        to_nodal(mf(SF_B), static_cast<T*>(state[SF_NODAL_B].d));
    }

    // the nodal fields U1..B as the reference holds them between calls: the
    // transforms of the last PopulateNodalVariables while its snapshot is
    // pending, else of the state (every stage ends with PopulateNodalVariables,
    // IMEXRK.cpp:8-24).  The step itself never writes them.
    void refresh_nodal() {
        for (int i = 0; i < 4; ++i) {
            if (i == 1 && !three) continue;
            to_nodal(src(i), static_cast<T*>(state[SF_NODAL_U1 + i].d));
        }
    }

    // ---- adjoint step (IMEXRK.h:86-126) ----
    // nodal: direct totals, their horizontal derivatives, explicit terms;
    // modal: transformed explicit terms and one derivative scratch field.
    // Allocated on the first adjoint step.
    T* adjT[4] = {nullptr};
    T* adjDX[4] = {nullptr};
    T* adjDY[4] = {nullptr};
    T* adjQ[12] = {nullptr};
    V* adjM[12] = {nullptr};
    V* adjD = nullptr;

    int adj_terms() const { return three ? 12 : 6; }
    bool has_field(int i) const { return three || i != 1; }

    void adjoint_alloc() {
        if (adjD) return;
        for (int i = 0; i < 4; ++i) {
            if (!has_field(i)) continue;
            adjT[i] = dalloc<T>(nnodal);
            adjDX[i] = dalloc<T>(nnodal);
            if (three) adjDY[i] = dalloc<T>(nnodal);
        }
        for (int q = 0; q < adj_terms(); ++q) {
            adjQ[q] = dalloc<T>(nnodal);
            adjM[q] = dalloc<V>(nmodal);
        }
        adjD = dalloc<V>(nmodal);
    }

    // below/above: the direct state (u1, u2, u3, b; device modal fields of this
    // context, u2 null in 2D) at the end and at the start of the backward step
    void step_adjoint(const V* const below[4], const V* const above[4]) {
        adjoint_alloc();
        V* tot[4] = {mf(SF_U1_TOT), mf(SF_U2_TOT), mf(SF_U3_TOT), mf(SF_B_TOT)};
        T* nod[4] = {static_cast<T*>(state[SF_NODAL_U1].d), static_cast<T*>(state[SF_NODAL_U2].d),
                     static_cast<T*>(state[SF_NODAL_U3].d), static_cast<T*>(state[SF_NODAL_B].d)};
        const T* ubar = zc + size_t(ZC_UBAR) * N3p;
        T frac = T(0);
        for (int k = 0; k < 3; ++k) {
            // the direct state at this stage, with the background velocity
            // added through a nodal round trip (IMEXRK.h:96-107)
            for (int i = 0; i < 4; ++i)
                if (has_field(i)) check(launch_lerp<T>(tot[i], above[i], below[i], frac, nmodal, stream), "lerp");
            to_nodal(tot[0], adjT[0]);
            check(launch_add_profile<T>(adjT[0], ubar, nnodal, N3, stream), "add_profile");
            to_modal(adjT[0], tot[0]);
            // UpdateAdjointVariables (IMEXRK.h:461-500) and the nodal
            // derivatives of BuildRHSAdjoint (IMEXRK.cpp:312-360)
            for (int i = 0; i < 4; ++i) {
                if (!has_field(i)) continue;
                to_nodal(tot[i], adjT[i]);
                check(launch_dmul<T>(adjD, tot[i], A1, A2, N3, kx, ky, 0, zc, N3p, stream), "ddx");
                to_nodal(adjD, adjDX[i]);
                if (three) {
                    check(launch_dmul<T>(adjD, tot[i], A1, A2, N3, kx, ky, 1, zc, N3p, stream), "ddy");
                    to_nodal(adjD, adjDY[i]);
                }
            }
            AdjArgs<T> a{};
            a.N3 = N3;
            a.N3p = N3p;
            a.A1 = A1;
            a.n = nnodal;
            for (int i = 0; i < 4; ++i) {
                a.U[i] = has_field(i) ? nod[i] : nullptr;
                a.Tt[i] = adjT[i];
                a.DX[i] = adjDX[i];
                a.DY[i] = adjDY[i];
            }
            a.u3t = tot[2];
            a.bt = tot[3];
            a.zc = zc;
            a.Ri = T(Ri);
            for (int q = 0; q < adj_terms(); ++q) a.Q[q] = adjQ[q];
            check(launch_adj_products<T>(a, three, stream), "adj_products");
            for (int q = 0; q < adj_terms(); ++q) to_modal(adjQ[q], adjM[q]);
            AdjCombine<T> c{};
            c.N3 = N3;
            c.A1 = A1;
            c.A2 = A2;
            c.three = three ? 1 : 0;
            c.kx = kx;
            c.ky = ky;
            for (int q = 0; q < adj_terms(); ++q) c.M[q] = adjM[q];
            for (int i = 0; i < 4; ++i) c.out[i] = has_field(i) ? nl[i] : nullptr;
            check(launch_adj_combine<T>(c, stream), "adj_combine");
            // ExplicitRK, FinishRHS, CrankNicolson, RemoveDivergence(1/h), FilterAll
            const bool store_r = k < 2;
            ModalArgs<T> m = stage_modal_args(k, store_r, rcur);
            m.skip_linear = 1;
            check(launch_modal<T>(m, stream), "modal");
            if (store_r) {
                rcur = 1 - rcur;
                bind_r();
            }
            // PopulateNodalVariables: the nodal fields follow the new state
            for (int i = 0; i < 4; ++i)
                if (has_field(i)) to_nodal(mf(SF_U1 + i), nod[i]);
            use_snap = false;
            frac += h[k] / T(dt);
        }
    }

    // ---- diagnostics ----
    T* dprof = nullptr;  // one N3 profile on the device
    T* jkN[3] = {nullptr};
    V* jkM[2] = {nullptr};

    // nodal += profile(z) (a NodalField plus a Nodal1D, Field.h:1215-1253)
    void add_profile(T* nodal, const double* host) {
        if (!dprof) dprof = dalloc<T>(N3p);
        std::vector<T> h(N3);
        for (int j = 0; j < N3; ++j) h[j] = T(host[j]);
        SF_HIP(hipMemcpyAsync(dprof, h.data(), N3 * sizeof(T), hipMemcpyHostToDevice, stream));
        check(launch_add_profile<T>(nodal, dprof, nnodal, N3, stream), "add_profile");
        SF_HIP(hipStreamSynchronize(stream));
    }

    // HorizontalAverage (Integration.h:37-42): real part of the (0,0) stack
    void mean_profile(const V* modal, double* out) {
        check(launch_mean_profile<T>(modal, partial, N3, A1, stream), "mean_profile");
        std::vector<double> h(N3);
        SF_HIP(hipMemcpyAsync(h.data(), partial, N3 * sizeof(double), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        for (int j = 0; j < N3; ++j) out[j] = h[j];
    }

    // JoverK (IMEXRK.h:445-459) with the J integrand of UpdateAdjointVariables
    // (IMEXRK.h:470-490): J = IntegrateVertically(<-(b - <b>) w>), K = 2.
    // The reference reads the nodal B, which every stage leaves equal to
    // ToNodal(b); the step here does not write nodal fields, so B is formed
    // from the modal b of the state.
    double jover_k() {
        if (!jkM[0]) {
            for (int i = 0; i < 3; ++i) jkN[i] = dalloc<T>(nnodal);
            for (int i = 0; i < 2; ++i) jkM[i] = dalloc<V>(nmodal);
        }
        to_nodal(mf(SF_B), jkN[0]);  // B
        to_modal(jkN[0], jkM[0]);    // b_total
        to_nodal(jkM[0], jkN[0]);                                        // B_tot
        to_nodal(mf(SF_U3), jkN[1]);                                     // U3_tot
        check(launch_mix_product<T>(jkN[2], jkN[0], jkN[1], jkM[0], nnodal, N3, A1, stream), "mix_product");
        to_modal(jkN[2], jkM[1]);
        std::vector<double> Jint(N3);
        mean_profile(jkM[1], Jint.data());
        // IntegrateVertically of a Dirichlet profile (Integration.h:22-34), in T
        const VGrid g = vertical_grid(L3, N3);
        T J = 0;
        for (int k = 2; k < N3 - 1; ++k) J += T(g.zf[k] - g.zf[k - 1]) * T(Jint[k]);
        const T K = 2;
        return double(J / K);
    }

    // dot products into scal[slot]
    void dot_async(const V* a, const V* b, int bc, int slot) {
        const T* w = zc + size_t(bc == SF_NEUMANN ? ZC_WN : ZC_WD) * N3p;
        check(launch_dot<T>(a, b, A1, A2, N3, w, partial, scal, slot, stream), "dot");
    }
    void read_scalars(int n, double* out) {
        SF_HIP(hipMemcpyAsync(scal_host, scal, n * sizeof(double), hipMemcpyDeviceToHost, stream));
        SF_HIP(hipStreamSynchronize(stream));
        for (int i = 0; i < n; ++i) out[i] = scal_host[i];
    }

    // IMEXRK::CFL (IMEXRK.h:355-374) for the current dt, without changing it
    double cfl() {
        // the reference reads whatever U1..U3 hold (IMEXRK.h:359-365): the
        // populated snapshot if one is pending, else the current state
        to_nodal(src(0), static_cast<T*>(state[SF_NODAL_U1].d));
        if (three) to_nodal(src(1), static_cast<T*>(state[SF_NODAL_U2].d));
        to_nodal(src(2), static_cast<T*>(state[SF_NODAL_U3].d));
        const double d1 = cfl_d[0], d2 = cfl_d[1], d3 = cfl_d[2];
        const T* ubar = zc + size_t(ZC_UBAR) * N3p;
        check(launch_maxabs<T>(static_cast<T*>(state[SF_NODAL_U1].d), nnodal, N3, ubar, partial, scal, 0, stream),
              "cfl");
        if (three)
            check(launch_maxabs<T>(static_cast<T*>(state[SF_NODAL_U2].d), nnodal, N3, nullptr, partial, scal, 1,
                                   stream),
                  "cfl");
        check(launch_maxabs<T>(static_cast<T*>(state[SF_NODAL_U3].d), nnodal, N3, nullptr, partial, scal, 2, stream),
              "cfl");
        double m[3];
        read_scalars(3, m);
        if (!three) m[1] = 0;
        return (m[0] / d1 + m[1] / d2 + m[2] / d3) * dt;
    }

    // IMEXRK::CFLlinear (IMEXRK.h:376-393) for the current dt, without changing
    // it: dt*((1 + max|U1_tot|)/d1 + max|U2_tot|/d2 + max|U3_tot|/d3) over the
    // nodal linearisation state (SetBackground(u1..b), IMEXRK.h:334-352); 1
    // bounds |U_|.  The nodal totals go through one nodal scratch field in
    // turn (allocated on the first call), the maxima through the context's
    // reduction partials.
    T* ntmp = nullptr;
    double cfl_linear() {
        if (!ntmp) ntmp = dalloc<T>(nnodal);
        return cfl_linear_with(ntmp, partial);
    }
    // the same with caller-provided scratch: a nodal field and the partial
    // sums of the max reductions (at most max_reduction_blocks() doubles)
    double cfl_linear_with(T* nodal, double* part) {
        double m[3] = {0, 0, 0};
        for (int i = 0; i < 3; ++i) {
            if (i == 1 && !three) continue;
            to_nodal(mf(SF_U1_TOT + i), nodal);
            check(launch_maxabs<T>(nodal, nnodal, N3, nullptr, part, scal, 0, stream), "cfl_linear");
            read_scalars(1, &m[i]);
        }
        return ((1.0 + m[0]) / cfl_d[0] + m[1] / cfl_d[1] + m[2] / cfl_d[2]) * dt;
    }

    void energy(double* ke, double* pe) {
        dot_async(mf(SF_U1), mf(SF_U1), SF_NEUMANN, 0);
        dot_async(mf(SF_U3), mf(SF_U3), SF_DIRICHLET, 1);
        if (three) dot_async(mf(SF_U2), mf(SF_U2), SF_NEUMANN, 2);
        dot_async(mf(SF_B), mf(SF_B), SF_NEUMANN, 3);
        double s[4];
        read_scalars(4, s);
        double e = 0.5 * (s[0] + s[1]);  // IMEXRK.h:395-406
        if (three) e += 0.5 * s[2];
        *ke = e;
        *pe = Ri * 0.5 * s[3];  // IMEXRK.h:408-411
    }

    // ---- host <-> device with layout conversion ----
    void upload(sf_field* f, const double* host) {
        if (f->kind == SF_MODAL) {
            std::vector<V> h(nmodal);
            for (int ja = 0; ja < A2; ++ja) {
                const int j2 = three ? ((ja < N2 / 3) ? ja : ja + N2 - 2 * (N2 / 3) + 1) : 0;
                for (int j1 = 0; j1 < A1; ++j1)
                    for (int j3 = 0; j3 < N3; ++j3) {
                        const size_t src = (size_t(M1) * j2 + j1) * N3 + j3;
                        const size_t dst = sidx(ja, j3, j1, N3, A1);
                        h[dst].x = T(host[2 * src]);
                        h[dst].y = T(host[2 * src + 1]);
                    }
            }
            SF_HIP(hipMemcpyAsync(f->d, h.data(), nmodal * sizeof(V), hipMemcpyHostToDevice, stream));
            SF_HIP(hipStreamSynchronize(stream));
        } else {
            std::vector<T> h(f->n);
            for (size_t i = 0; i < f->n; ++i) h[i] = T(host[i]);
            SF_HIP(hipMemcpyAsync(f->d, h.data(), f->n * sizeof(T), hipMemcpyHostToDevice, stream));
            SF_HIP(hipStreamSynchronize(stream));
        }
    }
    void download(const sf_field* f, double* host) {
        if (f->kind == SF_MODAL) {
            std::vector<V> h(nmodal);
            SF_HIP(hipMemcpyAsync(h.data(), f->d, nmodal * sizeof(V), hipMemcpyDeviceToHost, stream));
            SF_HIP(hipStreamSynchronize(stream));
            const size_t full = size_t(M1) * N2 * N3;
            std::memset(host, 0, 2 * full * sizeof(double));
            for (int ja = 0; ja < A2; ++ja) {
                const int j2 = three ? ((ja < N2 / 3) ? ja : ja + N2 - 2 * (N2 / 3) + 1) : 0;
                for (int j1 = 0; j1 < A1; ++j1)
                    for (int j3 = 0; j3 < N3; ++j3) {
                        const size_t dst = (size_t(M1) * j2 + j1) * N3 + j3;
                        const size_t src = sidx(ja, j3, j1, N3, A1);
                        host[2 * dst] = h[src].x;
                        host[2 * dst + 1] = h[src].y;
                    }
            }
        } else {
            std::vector<T> h(f->n);
            SF_HIP(hipMemcpyAsync(h.data(), f->d, f->n * sizeof(T), hipMemcpyDeviceToHost, stream));
            SF_HIP(hipStreamSynchronize(stream));
            for (size_t i = 0; i < f->n; ++i) host[i] = h[i];
        }
    }

    sf_field* new_field(int kind, int bc) {
        SF_REQUIRE(kind == SF_NODAL || kind == SF_MODAL || kind == SF_1D, SF_ERR_ARG, "bad field kind");
        SF_REQUIRE(bc == SF_NEUMANN || bc == SF_DIRICHLET, SF_ERR_ARG, "bad boundary condition");
        std::unique_ptr<sf_field> f(new sf_field());
        f->ctx = owner;
        f->kind = kind;
        f->bc = bc;
        f->owned = true;
        f->n = kind == SF_MODAL ? 2 * nmodal : (kind == SF_NODAL ? nnodal : size_t(N3));
        void* p = nullptr;
        SF_HIP(hipMalloc(&p, f->n * sizeof(T)));
        f->d = p;
        SF_HIP(hipMemsetAsync(p, 0, f->n * sizeof(T), stream));
        return f.release();
    }
};

// ---------------------------------------------------------------------------
// Batched ensemble: many independent 2D runs (each its own context, grid and
// parameters) advanced together.  Each phase of a stage has two plans:
//   x-pass: one launch per distinct N1 (k_xpass_batch), or one launch for all
//           runs (k_xpass_mixed, each work item dispatching on its run's N1)
//           with the default FFT plans or first radix 8 where N1 allows it;
//   modal:  one launch per distinct rows-per-lane M (k_modal_batch), or one
//           launch for all runs with 4 or 8 modes per block (k_modal_mixed).
// Both kernels loop over the work items of all their runs (grid-stride, so a
// fixed number of resident blocks drains every run).  The autotuner times
// every candidate when the batch is built -- the x-pass plans with their
// levels per block and resident-block counts on the runs themselves (the
// x-pass only rewrites the nonlinear terms), the modal plans on scratch
// copies of the state -- and keeps the fastest of each phase.  Every plan does
// the same arithmetic per mode and per level, so the choice changes the time
// only.  One step of all runs is captured in a HIP graph and replayed.
// ---------------------------------------------------------------------------
template <typename T>
struct Batch {
    using V = cx<T>;
    struct Group {
        int key = 0;  // N1 (x-pass), rows per lane M (grouped modal), modes per block (mixed modal)
        std::vector<int> members;
        int total = 0;
        int grid = 1;
        int zch = 2;
        int span = 1;  // mixed launches: the largest N1 (x-pass) or rows per lane (modal) of the members
        void* args = nullptr;
        int* start = nullptr;
    };
    std::vector<Ctx<T>*> runs;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<Group> xg;  // x-pass, one group per N1
    Group xmix;             // x-pass, all runs
    std::vector<Group> mg[3];
    Group mmix[3];
    int xplan = SF_BATCH_GROUPED;
    int mplan = SF_BATCH_GROUPED;
    // x grouped, x mixed, x mixed with first radix 8, modal grouped, modal
    // mixed with 4 and with 8 modes per block
    std::vector<double> tuned_ms;
    std::vector<double> dts;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int ncu = 256;
    int launches_per_step = 0;

    ~Batch() {
        release_graph();
        for (auto& g : xg) free_group(g);
        free_group(xmix);
        for (int k = 0; k < 3; ++k) {
            for (auto& g : mg[k]) free_group(g);
            free_group(mmix[k]);
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    static void free_group(Group& g) {
        if (g.args) (void)hipFree(g.args);
        if (g.start) (void)hipFree(g.start);
        g.args = nullptr;
        g.start = nullptr;
    }
    void release_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    void init(const std::vector<Ctx<T>*>& r) {
        runs = r;
        SF_REQUIRE(!runs.empty(), SF_ERR_ARG, "empty batch");
        for (auto* c : runs) SF_REQUIRE(!c->three, SF_ERR_UNSUPPORTED, "batched runs must be 2D (N2 == 1)");
        int dev = 0;
        SF_HIP(hipGetDevice(&dev));
        SF_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        SF_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        SF_HIP(hipEventCreate(&ev0));
        SF_HIP(hipEventCreate(&ev1));
        std::map<int, std::vector<int>> byN1, byM;
        std::vector<int> all;
        for (int i = 0; i < int(runs.size()); ++i) {
            byN1[runs[i]->N1].push_back(i);
            byM[(runs[i]->N3 + 63) / 64].push_back(i);
            all.push_back(i);
        }
        for (auto& kv : byN1) {
            Group g;
            g.key = kv.first;
            g.members = kv.second;
            xg.push_back(g);
        }
        xmix.members = all;
        int nmax = 0, mmax = 0;
        for (auto* c : runs) {
            nmax = std::max(nmax, c->N1);
            mmax = std::max(mmax, (c->N3 + 63) / 64);
        }
        xmix.span = nmax;
        for (int k = 0; k < 3; ++k) {
            for (auto& kv : byM) {
                Group g;
                g.key = kv.first;
                g.members = kv.second;
                mg[k].push_back(g);
            }
            mmix[k].members = all;
            mmix[k].key = 4;
            mmix[k].span = mmax;
        }
        sync_runs();
        tune();
        build();
    }

    void sync_runs() {
        for (auto* c : runs) SF_HIP(hipStreamSynchronize(c->stream));
    }

    template <typename A>
    void upload_group(Group& g, const std::vector<A>& a, const std::vector<int>& st) {
        free_group(g);
        SF_HIP(hipMalloc(&g.args, a.size() * sizeof(A)));
        SF_HIP(hipMalloc(reinterpret_cast<void**>(&g.start), st.size() * sizeof(int)));
        SF_HIP(hipMemcpy(g.args, a.data(), a.size() * sizeof(A), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(g.start, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice));
    }

    // r1f: first radix of the FFT plans (mixed launches; 0 = default)
    void build_xgroup(Group& g, int zch, int grid, int r1f = 0) {
        std::vector<XArgs<T>> a;
        std::vector<int> st;
        int total = 0;
        for (int r : g.members) {
            XArgs<T> x = runs[r]->xargs_2d(false, true);
            x.zch = zch;
            x.tw = runs[r]->x_twiddles(r1f);
            a.push_back(x);
            st.push_back(total);
            total += (runs[r]->N3 + zch - 1) / zch;
        }
        st.push_back(total);
        g.zch = zch;
        g.total = total;
        g.grid = std::max(1, std::min(total, grid));
        upload_group(g, a, st);
    }

    // modal group: arguments of run r from args(r), mb modes per block for
    // every member (mixed) or the kernel's own modes per block for M = key
    template <typename Fn>
    void build_mgroup(Group& g, bool mixed, Fn args) {
        std::vector<ModalArgs<T>> a;
        std::vector<int> st;
        int total = 0;
        for (int r : g.members) {
            Ctx<T>* c = runs[r];
            a.push_back(args(r));
            st.push_back(total);
            const int mb = mixed ? g.key : modal_modes_per_block(g.key, int(sizeof(T)));
            total += ((c->A1 + mb - 1) / mb) * c->A2;
        }
        st.push_back(total);
        g.total = total;
        g.grid = std::max(1, std::min(total, 8 * ncu));
        upload_group(g, a, st);
    }

    // the explicit terms ping-pong as in Ctx::stage: stage 0 writes
    // rbuf[1-rcur], stage 1 reads it and writes rbuf[rcur], stage 2 reads
    // rbuf[rcur]; after a step rcur is unchanged
    ModalArgs<T> stage_args(int r, int k) {
        Ctx<T>* c = runs[r];
        const int rc = (k == 1) ? 1 - c->rcur : c->rcur;
        return c->stage_modal_args(k, k < 2, rc);
    }

    void build() {
        release_graph();
        if (xplan != SF_BATCH_GROUPED) {
            build_xgroup(xmix, xmix.zch, xmix.grid, xr1f());
        } else {
            for (auto& g : xg) build_xgroup(g, g.zch, g.grid);
        }
        for (int k = 0; k < 3; ++k) {
            if (mplan == SF_BATCH_GROUPED) {
                for (auto& g : mg[k]) build_mgroup(g, false, [&](int r) { return stage_args(r, k); });
            } else {
                mmix[k].key = (mplan == SF_BATCH_MIXED8) ? 8 : 4;
                build_mgroup(mmix[k], true, [&](int r) { return stage_args(r, k); });
            }
        }
        dts.clear();
        for (auto* c : runs) dts.push_back(c->dt);
        launches_per_step = 3 * ((xplan != SF_BATCH_GROUPED ? 1 : int(xg.size())) +
                                 (mplan == SF_BATCH_GROUPED ? int(mg[0].size()) : 1));
    }

    int xr1f() const { return xplan == SF_BATCH_MIXED_R8 ? 8 : 0; }

    void launch_x(bool mixed) {
        if (mixed) {
            const hipError_t e = launch_xpass_mixed<T>(static_cast<const XArgs<T>*>(xmix.args), xmix.start,
                                                       int(xmix.members.size()), xmix.total, xmix.span, xr1f(),
                                                       xmix.grid, stream);
            if (e != hipSuccess) throw Error(SF_ERR_HIP, std::string("batched xpass: ") + hipGetErrorString(e));
            return;
        }
        for (auto& g : xg) {
            const hipError_t e = launch_xpass_batch<T>(static_cast<const XArgs<T>*>(g.args), g.start,
                                                       int(g.members.size()), g.total, g.key, g.grid, stream);
            if (e != hipSuccess) throw Error(SF_ERR_HIP, std::string("batched xpass: ") + hipGetErrorString(e));
        }
    }

    void launch_m(int plan, std::vector<Group>& groups, Group& mix) {
        if (plan != SF_BATCH_GROUPED) {
            const hipError_t e =
                launch_modal_mixed<T>(static_cast<const ModalArgs<T>*>(mix.args), mix.start, int(mix.members.size()),
                                      mix.total, mix.key, mix.span, mix.grid, stream);
            if (e != hipSuccess) throw Error(SF_ERR_HIP, std::string("batched modal: ") + hipGetErrorString(e));
            return;
        }
        for (auto& g : groups) {
            const hipError_t e = launch_modal_batch<T>(static_cast<const ModalArgs<T>*>(g.args), g.start,
                                                       int(g.members.size()), g.total, g.key, g.grid, stream);
            if (e != hipSuccess) throw Error(SF_ERR_HIP, std::string("batched modal: ") + hipGetErrorString(e));
        }
    }

    void launch_stage(int k) {
        launch_x(xplan != SF_BATCH_GROUPED);
        launch_m(mplan, mg[k], mmix[k]);
    }

    // GPU time of `reps` calls of fn after one untimed call
    template <typename Fn>
    double time_ms(Fn fn, int reps = 3) {
        fn();
        SF_HIP(hipEventRecord(ev0, stream));
        for (int i = 0; i < reps; ++i) fn();
        SF_HIP(hipEventRecord(ev1, stream));
        SF_HIP(hipEventSynchronize(ev1));
        float ms = 0;
        SF_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        return double(ms) / reps;
    }

    void tune() {
        tuned_ms.assign(6, 0.0);
        // x-pass: levels per block and resident blocks of each N1 group, and of
        // the single launch over all runs under each FFT plan
        auto best_x = [&](Group& g, bool mixed, int r1f) {
            double best = 1e30;
            int bz = 2, bgrid = 8 * ncu;
            for (int zch : {1, 2, 4, 8, 16}) {
                for (int mult : {4, 8, 16, 0}) {
                    const int grid = mult ? mult * ncu : (1 << 30);
                    build_xgroup(g, zch, grid, r1f);
                    const double ms = time_ms([&]() {
                        hipError_t e;
                        if (mixed)
                            e = launch_xpass_mixed<T>(static_cast<const XArgs<T>*>(g.args), g.start,
                                                      int(g.members.size()), g.total, g.span, r1f, g.grid, stream);
                        else
                            e = launch_xpass_batch<T>(static_cast<const XArgs<T>*>(g.args), g.start,
                                                      int(g.members.size()), g.total, g.key, g.grid, stream);
                        if (e != hipSuccess) throw Error(SF_ERR_HIP, "batched xpass (tuning)");
                    });
                    if (ms < best) {
                        best = ms;
                        bz = zch;
                        bgrid = grid;
                    }
                }
            }
            g.zch = bz;
            g.grid = bgrid;
            return best;
        };
        for (auto& g : xg) tuned_ms[0] += best_x(g, false, 0);
        Group x8 = xmix;
        tuned_ms[2] = best_x(x8, true, 8);
        free_group(x8);
        tuned_ms[1] = best_x(xmix, true, 0);
        xplan = SF_BATCH_GROUPED;
        if (tuned_ms[1] < tuned_ms[0]) xplan = SF_BATCH_MIXED;
        if (tuned_ms[2] < std::min(tuned_ms[0], tuned_ms[1])) {
            xplan = SF_BATCH_MIXED_R8;
            xmix.zch = x8.zch;
            xmix.grid = x8.grid;
        }

        // modal: stage-0 launches on scratch copies of each run's state (the
        // fields the stage rewrites and its explicit-term outputs)
        std::vector<V*> scratch(runs.size(), nullptr);
        auto release = [&]() {
            for (V*& p : scratch)
                if (p) (void)hipFree(p);
        };
        try {
            std::vector<ModalArgs<T>> sa(runs.size());
            for (size_t r = 0; r < runs.size(); ++r) {
                Ctx<T>* c = runs[r];
                const size_t n = c->nmodal;
                SF_HIP(hipMalloc(reinterpret_cast<void**>(&scratch[r]), 9 * n * sizeof(V)));
                V* w = scratch[r];
                ModalArgs<T> m = stage_args(int(r), 0);
                V** rw[5] = {&m.u1, &m.u2, &m.u3, &m.b, &m.p};
                for (int i = 0; i < 5; ++i)
                    if (*rw[i]) {
                        SF_HIP(hipMemcpyAsync(w + i * n, *rw[i], n * sizeof(V), hipMemcpyDeviceToDevice, stream));
                        *rw[i] = w + i * n;
                    }
                V** wo[4] = {&m.rn1, &m.rn2, &m.rn3, &m.rnB};
                for (int i = 0; i < 4; ++i)
                    if (*wo[i]) *wo[i] = w + (5 + i) * n;
                m.ubar = nullptr;
                sa[r] = m;
            }
            auto sargs = [&](int r) { return sa[r]; };
            std::vector<Group> tg = mg[0];
            for (auto& g : tg) {
                g.args = nullptr;
                g.start = nullptr;
                build_mgroup(g, false, sargs);
            }
            tuned_ms[3] = time_ms([&]() { launch_m(SF_BATCH_GROUPED, tg, mmix[0]); });
            for (auto& g : tg) free_group(g);
            Group tm = mmix[0];
            tm.args = nullptr;
            tm.start = nullptr;
            for (int q = 0; q < 2; ++q) {
                tm.key = q ? 8 : 4;
                build_mgroup(tm, true, sargs);
                tuned_ms[4 + q] = time_ms([&]() { launch_m(q ? SF_BATCH_MIXED8 : SF_BATCH_MIXED, tg, tm); });
            }
            free_group(tm);
            SF_HIP(hipStreamSynchronize(stream));
        } catch (...) {
            release();
            throw;
        }
        release();
        mplan = SF_BATCH_GROUPED;
        double best = tuned_ms[3];
        if (tuned_ms[4] < best) {
            best = tuned_ms[4];
            mplan = SF_BATCH_MIXED;
        }
        if (tuned_ms[5] < best) mplan = SF_BATCH_MIXED8;
    }

    // a plan chosen by the caller instead of the tuner (-1 keeps a phase's)
    void select(int x, int m) {
        SF_REQUIRE(x == -1 || x == SF_BATCH_GROUPED || x == SF_BATCH_MIXED || x == SF_BATCH_MIXED_R8, SF_ERR_ARG,
                   "bad x-pass plan");
        SF_REQUIRE(m == -1 || m == SF_BATCH_GROUPED || m == SF_BATCH_MIXED || m == SF_BATCH_MIXED8, SF_ERR_ARG,
                   "bad modal plan");
        if (x != -1) xplan = x;
        if (m != -1) mplan = m;
        build();
    }

    void capture() {
        release_graph();
        SF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int k = 0; k < 3; ++k) launch_stage(k);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SF_HIP(hipStreamEndCapture(stream, &graph));
        SF_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }

    // advance every run by n TimeSteps (IMEXRK.cpp:8-24)
    void step(int n, double* ms_out = nullptr) {
        SF_REQUIRE(n >= 0, SF_ERR_ARG, "negative step count");
        sync_runs();
        bool stale = false;
        for (size_t i = 0; i < runs.size(); ++i) stale = stale || (runs[i]->dt != dts[i]);
        if (stale) build();
        int first = 0;
        bool snap = false;
        for (auto* c : runs) snap = snap || c->use_snap;
        if (snap && n > 0) {
            // the first step after PopulateNodalVariables reads the populated
            // nodal snapshot (StateVector.cpp:9-11): take it run by run
            for (auto* c : runs) c->step(1, false);
            sync_runs();
            first = 1;
        }
        if (!exec) capture();
        SF_HIP(hipEventRecord(ev0, stream));
        for (int i = first; i < n; ++i) SF_HIP(hipGraphLaunch(exec, stream));
        SF_HIP(hipEventRecord(ev1, stream));
        SF_HIP(hipEventSynchronize(ev1));
        if (ms_out) {
            float ms = 0;
            SF_HIP(hipEventElapsedTime(&ms, ev0, ev1));
            *ms_out = ms;
        }
    }
};

template <typename F>
static int guarded(sf_ctx* c, F fn) {
    try {
        fn();
        if (c) c->err.clear();
        return SF_OK;
    } catch (const Error& e) {
        if (c) c->err = e.what();
        else g_create_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (c) c->err = "host allocation failed";
        return SF_ERR_ALLOC;
    } catch (const std::exception& e) {
        if (c) c->err = e.what();
        return SF_ERR_STATE;
    }
}

template <typename Fn>
static void with_ctx(sf_ctx* c, Fn fn) {
    if (c->precision == 4) fn(*static_cast<Ctx<float>*>(c->raw));
    else fn(*static_cast<Ctx<double>*>(c->raw));
}

}  // namespace sf

using namespace sf;

#define SF_NULLCHK(p)                                                  \
    do {                                                               \
        if (!(p)) return SF_ERR_ARG;                                   \
    } while (0)

template <typename Fn>
static int batch_call(sf_batch* b, Fn fn) {
    try {
        if (b->precision == 4) fn(*static_cast<Batch<float>*>(b->raw));
        else fn(*static_cast<Batch<double>*>(b->raw));
        b->err.clear();
        return SF_OK;
    } catch (const Error& e) {
        b->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        b->err = e.what();
        return SF_ERR_STATE;
    }
}

sf::CtxInfo sf_ctx_info(const sf_ctx* c) {
    sf::CtxInfo info{};
    info.precision = c->precision;
    with_ctx(const_cast<sf_ctx*>(c), [&](auto& C) {
        info.N1 = C.N1;
        info.N2 = C.N2;
        info.N3 = C.N3;
        info.A1 = C.A1;
        info.A2 = C.A2;
        info.three = C.three;
        info.stream = C.stream;
        info.kx = C.kx;
        info.ky = C.ky;
    });
    return info;
}

// ---------------------------------------------------------------------------
// Scratch of sf_cfl_linear_noalloc: one array per GPU, held until the process
// ends, shared by every call on that GPU (overlapping calls on one GPU race).
// Layout: the partial maxima of the reductions, then one nodal field.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
constexpr size_t NOALLOC_PARTIAL_BYTES = 8192;  // >= max_reduction_blocks() doubles
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, unsigned char*> g_scratch_ptr;  // device -> address of its scratch
static std::map<int, std::string> g_scratch_err;

// Resolves the device address of `scratch` on `dev`, which must be current;
// once per device.  Host code uses only the resolved address, never the
// array's name (a host-side stand-in).
static void resolve_scratch(int dev) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr.count(dev) || g_scratch_err.count(dev)) return;
    unsigned char* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        g_scratch_err[dev] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    else
        g_scratch_ptr[dev] = ptr;
}

static unsigned char* scratch_of(int dev) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    auto er = g_scratch_err.find(dev);
    throw Error(SF_ERR_HIP, er != g_scratch_err.end() ? er->second : "scratch not resolved for this device");
}

extern "C" {

int sf_ctx_create(const sf_grid* grid, const sf_flow* flow, int precision, sf_ctx** out) {
    if (!grid || !flow || !out) {
        g_create_error = "null argument";
        return SF_ERR_ARG;
    }
    *out = nullptr;
    std::unique_ptr<sf_ctx> c(new sf_ctx());
    c->precision = precision == 8 ? 8 : 4;
    if (precision != 4 && precision != 8) {
        g_create_error = "precision must be 4 or 8";
        return SF_ERR_ARG;
    }
    const int rc = guarded(nullptr, [&]() {
        if (c->precision == 4) {
            auto p = std::make_shared<Ctx<float>>();
            p->owner = c.get();
            p->init(*grid, *flow);
            c->raw = p.get();
            c->impl = p;
        } else {
            auto p = std::make_shared<Ctx<double>>();
            p->owner = c.get();
            p->init(*grid, *flow);
            c->raw = p.get();
            c->impl = p;
        }
    });
    if (rc != SF_OK) return rc;
    // the context was created on the current device
    with_ctx(c.get(), [&](auto& C) { resolve_scratch(C.device); });
    *out = c.release();
    return SF_OK;
}

int sf_ctx_destroy(sf_ctx* ctx) {
    delete ctx;
    return SF_OK;
}

const char* sf_last_error(const sf_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int sf_set_dt(sf_ctx* ctx, double dt) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.set_dt(dt); }); });
}

int sf_get_dt(const sf_ctx* ctx, double* dt) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(dt);
    with_ctx(const_cast<sf_ctx*>(ctx), [&](auto& C) { *dt = C.dt; });
    return SF_OK;
}

int sf_set_flow(sf_ctx* ctx, const sf_flow* flow) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(flow);
    return guarded(ctx, [&]() {
        with_ctx(ctx, [&](auto& C) {
            C.set_flow(*flow);
        });
    });
}

int sf_set_background(sf_ctx* ctx, const double* U, const double* dBdz) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(U);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.set_background(U, dBdz); }); });
}

int sf_vertical_points(const sf_ctx* ctx, double* z, double* zf) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(z);
    SF_NULLCHK(zf);
    with_ctx(const_cast<sf_ctx*>(ctx), [&](auto& C) {
        const VGrid g = vertical_grid(C.L3, C.N3);
        for (int j = 0; j < C.N3; ++j) {
            z[j] = g.z[j];
            zf[j] = g.zf[j];
        }
    });
    return SF_OK;
}

int sf_field_alloc(sf_ctx* ctx, int kind, int bc, sf_field** out) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(out);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { *out = C.new_field(kind, bc); }); });
}

int sf_field_free(sf_field* f) {
    if (!f) return SF_OK;
    if (!f->owned) return SF_ERR_ARG;
    sf_synchronize(f->ctx);
    (void)hipFree(f->d);
    delete f;
    return SF_OK;
}

int sf_field_kind(const sf_field* f, int* kind, int* bc) {
    SF_NULLCHK(f);
    if (kind) *kind = f->kind;
    if (bc) *bc = f->bc;
    return SF_OK;
}

int sf_field_upload(sf_field* f, const double* host) {
    SF_NULLCHK(f);
    SF_NULLCHK(host);
    return guarded(f->ctx, [&]() { with_ctx(f->ctx, [&](auto& C) { C.upload(f, host); }); });
}

int sf_field_download(const sf_field* f, double* host) {
    SF_NULLCHK(f);
    SF_NULLCHK(host);
    return guarded(f->ctx, [&]() { with_ctx(f->ctx, [&](auto& C) { C.download(f, host); }); });
}

static bool same_shape(const sf_field* a, const sf_field* b) {
    return a->ctx == b->ctx && a->kind == b->kind && a->n == b->n;
}

int sf_field_copy(sf_field* dst, const sf_field* src) {
    SF_NULLCHK(dst);
    SF_NULLCHK(src);
    return guarded(dst->ctx, [&]() {
        SF_REQUIRE(same_shape(dst, src) && dst->bc == src->bc, SF_ERR_ARG, "field shape/BC mismatch (Field.h:285)");
        with_ctx(dst->ctx, [&](auto& C) {
            using TT = typename std::remove_reference<decltype(C)>::type;
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            (void)sizeof(TT);
            SF_HIP(hipMemcpyAsync(dst->d, src->d, dst->n * sizeof(T), hipMemcpyDeviceToDevice, C.stream));
        });
    });
}

int sf_field_zero(sf_field* f) {
    SF_NULLCHK(f);
    return guarded(f->ctx, [&]() {
        with_ctx(f->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            SF_HIP(hipMemsetAsync(f->d, 0, f->n * sizeof(T), C.stream));
        });
    });
}

int sf_field_axpby(sf_field* y, double a, const sf_field* x, double b) {
    SF_NULLCHK(y);
    return guarded(y->ctx, [&]() {
        if (x) SF_REQUIRE(same_shape(x, y) && x->bc == y->bc, SF_ERR_ARG, "field shape/BC mismatch (Field.h:352)");
        with_ctx(y->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.check(launch_axpby<T>(static_cast<cx<T>*>(y->d), T(a), x ? static_cast<const cx<T>*>(x->d) : nullptr,
                                    T(b), y->n / 2, C.stream),
                    "axpby");
        });
    });
}

int sf_field_scale(sf_field* f, double a) { return sf_field_axpby(f, 0.0, nullptr, a); }

static int ends_op(sf_field* f, int op) {
    SF_NULLCHK(f);
    return guarded(f->ctx, [&]() {
        SF_REQUIRE(f->kind == SF_MODAL, SF_ERR_UNSUPPORTED, "end conditions are implemented for modal fields");
        with_ctx(f->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.check(launch_ends<T>(static_cast<cx<T>*>(f->d), C.A1, C.A2, C.N3, op, C.stream), "ends");
        });
    });
}

int sf_field_zero_ends(sf_field* f) { return ends_op(f, f && f->bc == SF_DIRICHLET ? 1 : 0); }

int sf_field_neumann_ends(sf_field* f) {
    if (f && f->bc != SF_NEUMANN) {
        f->ctx->err = "NeumannEnds requires a Neumann field (Field.h:493-499)";
        return SF_ERR_ARG;
    }
    return ends_op(f, 2);
}

int sf_field_filter(sf_field* f) {
    SF_NULLCHK(f);
    return SF_OK;
}

int sf_to_modal(const sf_field* nodal, sf_field* modal) {
    SF_NULLCHK(nodal);
    SF_NULLCHK(modal);
    return guarded(modal->ctx, [&]() {
        SF_REQUIRE(nodal->kind == SF_NODAL && modal->kind == SF_MODAL && nodal->bc == modal->bc, SF_ERR_ARG,
                   "ToModal needs a nodal input and a modal output with the same BC (Field.h:736)");
        with_ctx(modal->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.to_modal(static_cast<const T*>(nodal->d), static_cast<cx<T>*>(modal->d));
        });
    });
}

int sf_to_nodal(const sf_field* modal, sf_field* nodal) {
    SF_NULLCHK(nodal);
    SF_NULLCHK(modal);
    return guarded(modal->ctx, [&]() {
        SF_REQUIRE(nodal->kind == SF_NODAL && modal->kind == SF_MODAL && nodal->bc == modal->bc, SF_ERR_ARG,
                   "ToNodal needs a modal input and a nodal output with the same BC (Field.h:967)");
        with_ctx(modal->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.to_nodal(static_cast<const cx<T>*>(modal->d), static_cast<T*>(nodal->d));
        });
    });
}

static int deriv(const sf_field* in, sf_field* out, int axis) {
    SF_NULLCHK(in);
    SF_NULLCHK(out);
    return guarded(out->ctx, [&]() {
        SF_REQUIRE(in->kind == SF_MODAL && out->kind == SF_MODAL && in->d != out->d, SF_ERR_ARG,
                   "derivatives act on distinct modal fields");
        int ax = axis;
        if (axis == 2) {
            SF_REQUIRE(out->bc != in->bc, SF_ERR_ARG, "ddz changes the boundary condition (Stratiflow.h:67-80)");
            ax = in->bc == SF_NEUMANN ? 2 : 3;
        } else {
            SF_REQUIRE(out->bc == in->bc, SF_ERR_ARG, "ddx/ddy keep the boundary condition");
        }
        with_ctx(out->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.check(launch_dmul<T>(static_cast<cx<T>*>(out->d), static_cast<const cx<T>*>(in->d), C.A1, C.A2, C.N3,
                                   C.kx, C.ky, ax, C.zc, C.N3p, C.stream),
                    "deriv");
        });
    });
}

int sf_ddx(const sf_field* in, sf_field* out) { return deriv(in, out, 0); }
int sf_ddy(const sf_field* in, sf_field* out) { return deriv(in, out, 1); }
int sf_ddz(const sf_field* in, sf_field* out) { return deriv(in, out, 2); }

int sf_field_diffusion_precondition(const sf_field* in, sf_field* out, double tnu, double delta, double kmax,
                                    double tadv) {
    SF_NULLCHK(in);
    SF_NULLCHK(out);
    return guarded(out->ctx, [&]() {
        SF_REQUIRE(in->kind == SF_MODAL && same_shape(in, out) && in->bc == out->bc && in->d != out->d, SF_ERR_ARG,
                   "the diffusion preconditioner maps a modal field to a distinct one of the same BC");
        SF_REQUIRE(std::isfinite(tnu) && tnu >= 0 && std::isfinite(delta) && delta > 0 && std::isfinite(kmax) &&
                       std::isfinite(tadv),
                   SF_ERR_ARG, "needs tnu >= 0, delta > 0, a finite kmax and a finite tadv");
        with_ctx(out->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            DiffPrecArgs<T> p{};
            p.N3 = C.N3;
            p.N3p = C.N3p;
            p.A1 = C.A1;
            p.A2 = C.A2;
            p.dirichlet = in->bc == SF_DIRICHLET ? 1 : 0;
            p.zc = C.zc;
            p.kx = C.kx;
            p.kx2 = C.kx2;
            p.ky2 = C.ky2;
            p.in = static_cast<const cx<T>*>(in->d);
            p.out = static_cast<cx<T>*>(out->d);
            p.tnu = T(tnu);
            p.delta = T(delta);
            p.k2max = kmax < 0 ? T(-1) : T(kmax * kmax);
            p.tadv = T(tadv);
            SF_REQUIRE(tadv == 0 || size_t(C.N3) * ADVPREC_MODES * sizeof(cx<T>) <= ADVPREC_LDS, SF_ERR_UNSUPPORTED,
                       "the advected preconditioner supports N3 <= 512 (FP32) or 256 (FP64)");
            C.check(launch_diffprec<T>(p, C.stream), "diffusion preconditioner");
        });
    });
}

int sf_inner_prod(const sf_field* a, const sf_field* b, double* out) {
    SF_NULLCHK(a);
    SF_NULLCHK(b);
    SF_NULLCHK(out);
    return guarded(a->ctx, [&]() {
        SF_REQUIRE(a->kind == SF_MODAL && same_shape(a, b) && a->bc == b->bc, SF_ERR_ARG,
                   "InnerProd needs two modal fields with the same BC (Integration.h:285)");
        with_ctx(a->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.dot_async(static_cast<const cx<T>*>(a->d), static_cast<const cx<T>*>(b->d), a->bc, 0);
            C.read_scalars(1, out);
        });
    });
}

int sf_state_field(sf_ctx* ctx, int which, sf_field** out) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(out);
    return guarded(ctx, [&]() {
        SF_REQUIRE(which >= 0 && which <= SF_B_TOT, SF_ERR_ARG, "bad state field id");
        with_ctx(ctx, [&](auto& C) { *out = &C.state[which]; });
    });
}

int sf_step(sf_ctx* ctx, int nsteps) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.step(nsteps, false); }); });
}

int sf_step_linear(sf_ctx* ctx, int nsteps) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.step(nsteps, true); }); });
}

int sf_set_linear_background(sf_ctx* ctx, const sf_field* u1, const sf_field* u2, const sf_field* u3,
                             const sf_field* b) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(u1);
    SF_NULLCHK(u3);
    SF_NULLCHK(b);
    return guarded(ctx, [&]() {
        with_ctx(ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            const sf_field* src[4] = {u1, u2, u3, b};
            for (int i = 0; i < 4; ++i) {
                sf_field& d = C.state[SF_U1_TOT + i];
                if (!src[i]) {
                    SF_HIP(hipMemsetAsync(d.d, 0, d.n * sizeof(T), C.stream));
                    continue;
                }
                SF_REQUIRE(src[i]->kind == SF_MODAL && src[i]->n == d.n, SF_ERR_ARG, "background must be modal");
                SF_HIP(hipMemcpyAsync(d.d, src[i]->d, d.n * sizeof(T), hipMemcpyDeviceToDevice, C.stream));
            }
        });
    });
}

int sf_step_adjoint(sf_ctx* ctx, const sf_field* const below[4], const sf_field* const above[4]) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(below);
    SF_NULLCHK(above);
    return guarded(ctx, [&]() {
        with_ctx(ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            const cx<T>* b[4] = {nullptr};
            const cx<T>* a[4] = {nullptr};
            for (int i = 0; i < 4; ++i) {
                if (!C.has_field(i)) continue;
                for (int s = 0; s < 2; ++s) {
                    const sf_field* f = s ? above[i] : below[i];
                    SF_REQUIRE(f != nullptr, SF_ERR_ARG, "TimeStepAdjoint needs u1, u3, b (and u2 in 3D) below and above");
                    SF_REQUIRE(f->ctx == ctx && f->kind == SF_MODAL && f->n == 2 * C.nmodal, SF_ERR_ARG,
                               "adjoint direct states must be modal fields of this context");
                    SF_REQUIRE(f->owned, SF_ERR_ARG,
                               "adjoint direct states must be fields of their own, not the solver's state");
                    (s ? a : b)[i] = static_cast<const cx<T>*>(f->d);
                }
            }
            C.step_adjoint(b, a);
        });
    });
}

int sf_field_add_profile(sf_field* nodal, const double* profile) {
    SF_NULLCHK(nodal);
    SF_NULLCHK(profile);
    return guarded(nodal->ctx, [&]() {
        SF_REQUIRE(nodal->kind == SF_NODAL, SF_ERR_ARG, "a profile is added to a nodal field");
        with_ctx(nodal->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.add_profile(static_cast<T*>(nodal->d), profile);
        });
    });
}

int sf_field_mean_profile(const sf_field* modal, double* out) {
    SF_NULLCHK(modal);
    SF_NULLCHK(out);
    return guarded(modal->ctx, [&]() {
        SF_REQUIRE(modal->kind == SF_MODAL, SF_ERR_ARG, "HorizontalAverage reads a modal field (Integration.h:38)");
        with_ctx(modal->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.mean_profile(static_cast<const cx<T>*>(modal->d), out);
        });
    });
}

int sf_dot_f64(const sf_field* a, const sf_field* b, double* out) { return sf_inner_prod(a, b, out); }

int sf_field_enforce_bcs(sf_field* f) {
    SF_NULLCHK(f);
    return f->bc == SF_DIRICHLET ? sf_field_zero_ends(f) : sf_field_neumann_ends(f);
}

int sf_phase_shift(sf_field* f, double shift) {
    SF_NULLCHK(f);
    return guarded(f->ctx, [&]() {
        SF_REQUIRE(f->kind == SF_MODAL, SF_ERR_ARG, "PhaseShift acts on modal fields (Field.h:1168)");
        with_ctx(f->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            C.check(launch_phase_shift<T>(static_cast<cx<T>*>(f->d), C.nmodal, C.A1, shift, C.stream), "phase_shift");
        });
    });
}

int sf_remove_average(sf_field* f) {
    SF_NULLCHK(f);
    return guarded(f->ctx, [&]() {
        SF_REQUIRE(f->kind == SF_MODAL, SF_ERR_ARG, "RemoveAverage acts on modal fields (Integration.h:72)");
        with_ctx(f->ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            std::vector<double> m(C.N3);
            C.mean_profile(static_cast<const cx<T>*>(f->d), m.data());
            // IntegrateAllSpace(field, 1, 1, L3) (Integration.h:5-35, 50-57), in T
            const VGrid g = vertical_grid(C.L3, C.N3);
            T integral = 0;
            if (f->bc == SF_NEUMANN) {
                for (int k = 1; k < C.N3 - 1; ++k) integral += T(g.z[k + 1] - g.z[k]) * T(m[k]);
            } else {
                for (int k = 2; k < C.N3 - 1; ++k) integral += T(g.zf[k] - g.zf[k - 1]) * T(m[k]);
            }
            const T avg = integral / T(C.L3) / T(2);
            C.check(launch_add_mean<T>(static_cast<cx<T>*>(f->d), C.N3, C.A1, T(-avg), C.stream), "remove_average");
        });
    });
}

int sf_state_bind(sf_ctx* ctx, const sf_field* u1, const sf_field* u2, const sf_field* u3, const sf_field* b,
                  const sf_field* p) {
    SF_NULLCHK(ctx);
    const sf_field* src[5] = {u1, u2, u3, b, p};
    for (int i = 0; i < 5; ++i) {
        sf_field* st = nullptr;
        int rc = sf_state_field(ctx, SF_U1 + i, &st);
        if (rc == SF_OK) rc = src[i] ? sf_field_copy(st, src[i]) : sf_field_zero(st);
        if (rc != SF_OK) return rc;
    }
    return SF_OK;
}

int sf_state_fetch(sf_ctx* ctx, sf_field* u1, sf_field* u2, sf_field* u3, sf_field* b, sf_field* p) {
    SF_NULLCHK(ctx);
    sf_field* dst[5] = {u1, u2, u3, b, p};
    for (int i = 0; i < 5; ++i) {
        if (!dst[i]) continue;
        sf_field* st = nullptr;
        int rc = sf_state_field(ctx, SF_U1 + i, &st);
        if (rc == SF_OK) rc = sf_field_copy(dst[i], st);
        if (rc == SF_OK && i < 4) rc = sf_field_enforce_bcs(dst[i]);
        if (rc != SF_OK) return rc;
    }
    return SF_OK;
}

int sf_energy_f64(sf_ctx* ctx, double* ke, double* pe) { return sf_energy(ctx, ke, pe); }

int sf_jover_k(sf_ctx* ctx, double* out) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(out);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { *out = C.jover_k(); }); });
}

int sf_stage(sf_ctx* ctx, int k, int linear) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() {
        SF_REQUIRE(k >= 0 && k < 3, SF_ERR_ARG, "stage must be 0, 1 or 2");
        with_ctx(ctx, [&](auto& C) { C.stage(k, linear != 0, true); });
    });
}

int sf_remove_divergence(sf_ctx* ctx, double mult) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.remove_divergence(mult); }); });
}

int sf_filter_all(sf_ctx* ctx) {
    SF_NULLCHK(ctx);
    return SF_OK;  // compact storage holds active modes only
}

int sf_populate_nodal(sf_ctx* ctx) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.populate_nodal(); }); });
}

int sf_energy(sf_ctx* ctx, double* ke, double* pe) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(ke);
    SF_NULLCHK(pe);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.energy(ke, pe); }); });
}

int sf_cfl(sf_ctx* ctx, double* cfl) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(cfl);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { *cfl = C.cfl(); }); });
}

int sf_get_background(sf_ctx* ctx, double* U) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(U);
    return guarded(ctx, [&]() {
        with_ctx(ctx, [&](auto& C) {
            C.fetch_background();
            for (int j = 0; j < C.N3; ++j) U[j] = C.zc_host[size_t(ZC_UBAR) * C.N3p + j];
        });
    });
}

int sf_cfl_linear_noalloc(sf_ctx* ctx, double* cfl) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(cfl);
    return guarded(ctx, [&]() {
        with_ctx(ctx, [&](auto& C) {
            using T = typename std::remove_reference<decltype(C.h[0])>::type;
            const size_t need = C.nnodal * sizeof(T);
            SF_REQUIRE(need <= NOALLOC_SCRATCH_BYTES - NOALLOC_PARTIAL_BYTES, SF_ERR_ARG,
                       "sf_cfl_linear_noalloc: a nodal field of " + std::to_string(need) +
                           " bytes exceeds the scratch of " +
                           std::to_string(NOALLOC_SCRATCH_BYTES - NOALLOC_PARTIAL_BYTES) +
                           " bytes (use sf_cfl_linear)");
            SF_REQUIRE(size_t(max_reduction_blocks()) * sizeof(double) <= NOALLOC_PARTIAL_BYTES, SF_ERR_STATE,
                       "sf_cfl_linear_noalloc: reduction partials exceed their scratch");
            unsigned char* base = scratch_of(C.device);
            *cfl = C.cfl_linear_with(reinterpret_cast<T*>(base + NOALLOC_PARTIAL_BYTES),
                                     reinterpret_cast<double*>(base));
        });
    });
}

int sf_refresh_nodal(sf_ctx* ctx) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { C.refresh_nodal(); }); });
}

int sf_cfl_linear(sf_ctx* ctx, double* cfl) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(cfl);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { *cfl = C.cfl_linear(); }); });
}

int sf_synchronize(sf_ctx* ctx) {
    SF_NULLCHK(ctx);
    return guarded(ctx, [&]() { with_ctx(ctx, [&](auto& C) { SF_HIP(hipStreamSynchronize(C.stream)); }); });
}

int sf_time_steps(sf_ctx* ctx, int nsteps, double* ms) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(ms);
    return guarded(ctx, [&]() {
        with_ctx(ctx, [&](auto& C) {
            SF_HIP(hipEventRecord(C.ev0, C.stream));
            C.step(nsteps, false);
            SF_HIP(hipEventRecord(C.ev1, C.stream));
            SF_HIP(hipEventSynchronize(C.ev1));
            float t = 0;
            SF_HIP(hipEventElapsedTime(&t, C.ev0, C.ev1));
            *ms = t;
        });
    });
}

int sf_time_steps_detailed(sf_ctx* ctx, int nsteps, double* total_ms, double* kernel_ms, int* launches) {
    SF_NULLCHK(ctx);
    SF_NULLCHK(total_ms);
    SF_NULLCHK(kernel_ms);
    SF_NULLCHK(launches);
    return guarded(ctx, [&]() {
        with_ctx(ctx, [&](auto& C) {
            for (int c = 0; c < SF_K_COUNT; ++c) {
                kernel_ms[c] = 0;
                launches[c] = 0;
            }
            C.prof = true;
            C.pn = 0;
            try {
                C.mark(-1);
                C.step(nsteps, false);
            } catch (...) {
                C.prof = false;
                throw;
            }
            C.prof = false;
            SF_HIP(hipEventSynchronize(C.pev[C.pn - 1]));
            float t = 0;
            SF_HIP(hipEventElapsedTime(&t, C.pev[0], C.pev[C.pn - 1]));
            *total_ms = t;
            for (size_t i = 1; i < C.pn; ++i) {
                SF_HIP(hipEventElapsedTime(&t, C.pev[i - 1], C.pev[i]));
                const int c = C.pcls[i];
                if (c >= 0 && c < SF_K_COUNT) {
                    kernel_ms[c] += t;
                    launches[c] += 1;
                }
            }
        });
    });
}

// ---- batched ensemble ------------------------------------------------------
int sf_batch_create(sf_ctx* const* runs, int nruns, sf_batch** out) {
    SF_NULLCHK(runs);
    SF_NULLCHK(out);
    *out = nullptr;
    if (nruns <= 0) return SF_ERR_ARG;
    const int prec = runs[0] ? runs[0]->precision : 0;
    for (int i = 0; i < nruns; ++i)
        if (!runs[i] || runs[i]->precision != prec) return SF_ERR_ARG;
    std::unique_ptr<sf_batch> b(new sf_batch());
    b->precision = prec;
    const int rc = guarded(nullptr, [&]() {
        if (prec == 4) {
            auto impl = std::make_shared<Batch<float>>();
            std::vector<Ctx<float>*> r;
            for (int i = 0; i < nruns; ++i) r.push_back(static_cast<Ctx<float>*>(runs[i]->raw));
            impl->init(r);
            b->raw = impl.get();
            b->impl = impl;
        } else {
            auto impl = std::make_shared<Batch<double>>();
            std::vector<Ctx<double>*> r;
            for (int i = 0; i < nruns; ++i) r.push_back(static_cast<Ctx<double>*>(runs[i]->raw));
            impl->init(r);
            b->raw = impl.get();
            b->impl = impl;
        }
    });
    if (rc != SF_OK) return rc;
    *out = b.release();
    return SF_OK;
}

int sf_batch_destroy(sf_batch* b) {
    delete b;
    return SF_OK;
}

const char* sf_batch_last_error(const sf_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int sf_batch_step(sf_batch* b, int nsteps) {
    SF_NULLCHK(b);
    return batch_call(b, [&](auto& B) { B.step(nsteps); });
}

int sf_batch_time_steps(sf_batch* b, int nsteps, double* total_ms) {
    SF_NULLCHK(b);
    SF_NULLCHK(total_ms);
    return batch_call(b, [&](auto& B) { B.step(nsteps, total_ms); });
}

int sf_batch_info(const sf_batch* b, int* launches_per_step, int* xpass_groups, int* modal_groups) {
    SF_NULLCHK(b);
    return batch_call(const_cast<sf_batch*>(b), [&](auto& B) {
        if (launches_per_step) *launches_per_step = B.launches_per_step;
        if (xpass_groups) *xpass_groups = int(B.xg.size());
        if (modal_groups) *modal_groups = int(B.mg[0].size());
    });
}

int sf_batch_plan(const sf_batch* b, int* xpass_plan, int* modal_plan, double* tuned_ms) {
    SF_NULLCHK(b);
    return batch_call(const_cast<sf_batch*>(b), [&](auto& B) {
        if (xpass_plan) *xpass_plan = B.xplan;
        if (modal_plan) *modal_plan = B.mplan;
        if (tuned_ms)
            for (int i = 0; i < 6; ++i) tuned_ms[i] = B.tuned_ms[i];
    });
}

int sf_batch_select(sf_batch* b, int xpass_plan, int modal_plan) {
    SF_NULLCHK(b);
    return batch_call(b, [&](auto& B) { B.select(xpass_plan, modal_plan); });
}

}  // extern "C"
