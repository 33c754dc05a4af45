// sf_adjoint.hip -- kernels of the adjoint time step (IMEXRK::TimeStepAdjoint,
// IMEXRK.h:86-126; UpdateAdjointVariables, IMEXRK.h:461-500; BuildRHSAdjoint,
// IMEXRK.cpp:260-379).
//
// The context drives the step from the host (sf_api.hip): it interpolates the
// direct state (k_lerp), adds the background velocity through a nodal round
// trip (k_add_profile), transforms the direct state and its horizontal
// derivatives to nodal space, forms every explicit adjoint term as a nodal
// field in one pass (k_adj_products), transforms those to modal space and
// combines them with their horizontal derivatives into the explicit terms
// (k_adj_combine), which the stage kernel then consumes with its linear
// buoyancy terms switched off.  All of these kernels are HBM-bound streams.
// The file also holds the small kernels of the diagnostics (JoverK, horizontal
// averages) and of the field helpers (PhaseShift, RemoveAverage).
#include "sf_kernels.h"

namespace sf {

namespace {
constexpr int ABS = 256;

template <typename T>
__host__ __device__ inline unsigned grid_of(size_t n) {
    size_t b = (n + ABS - 1) / ABS;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return unsigned(b);
}
}  // namespace

// out = (1 - f) * above + f * below (IMEXRK.h:98-101); a null input is zero
template <typename T>
__global__ __launch_bounds__(ABS) void k_lerp(cx<T>* out, const cx<T>* above, const cx<T>* below, T f, size_t n) {
    const T g = T(1) - f;
    for (size_t i = size_t(blockIdx.x) * ABS + threadIdx.x; i < n; i += size_t(gridDim.x) * ABS) {
        const cx<T> a = above ? above[i] : czero<cx<T>>();
        const cx<T> b = below ? below[i] : czero<cx<T>>();
        cx<T> r;
        r.x = g * a.x + f * b.x;
        r.y = g * a.y + f * b.y;
        out[i] = r;
    }
}
template <typename T>
hipError_t launch_lerp(cx<T>* out, const cx<T>* above, const cx<T>* below, T f, size_t n, hipStream_t s) {
    hipLaunchKernelGGL((k_lerp<T>), dim3(grid_of<T>(n)), dim3(ABS), 0, s, out, above, below, f, n);
    return hipGetLastError();
}

// nodal += profile(z) (U1_tot += U_, IMEXRK.h:104)
template <typename T>
__global__ __launch_bounds__(ABS) void k_add_profile(T* f, const T* prof, size_t n, int N3) {
    for (size_t i = size_t(blockIdx.x) * ABS + threadIdx.x; i < n; i += size_t(gridDim.x) * ABS)
        f[i] += prof[i % size_t(N3)];
}
template <typename T>
hipError_t launch_add_profile(T* f, const T* prof, size_t n, int N3, hipStream_t s) {
    hipLaunchKernelGGL((k_add_profile<T>), dim3(grid_of<T>(n)), dim3(ABS), 0, s, f, prof, n, N3);
    return hipGetLastError();
}

// Every explicit adjoint term as a nodal field, one thread per nodal point
// (stack, z) with its vertical neighbours read from the same column (z is the
// fastest index of nodal fields).  With U = adjoint fields, T = direct totals,
// DX/DY = nodal horizontal derivatives of the totals:
//   Q0 = U1 T1                          (ddx -> r1)
//   Q1 = ddz(Ibar(U1) T3) + F1           (r1)
//   Q2 = Itilde(T1) U3                   (ddx -> r3)
//   Q3 = ddz(Idir(U3) Idir(T3)) + F3     (r3)
//   Q4 = B T1                            (ddx -> rB)
//   Q5 = ddz(Ibar(B) T3) + FB            (rB)
// and with a third dimension
//   Q6 = U1 T2 (ddy -> r1), Q7 = U2 T2 (ddy -> r2), Q8 = U2 T1 (ddx -> r2),
//   Q9 = ddz(Ibar(U2) T3) + F2 (r2), Q10 = Itilde(T2) U3 (ddy -> r3),
//   Q11 = B T2 (ddy -> rB).
// The forcing fields follow UpdateAdjointVariables and the extra terms of
// BuildRHSAdjoint (IMEXRK.cpp:312-367):
//   F1 = -(DX1 U1 + DX2 U2 + Idir(DX3 U3)) - DXb B
//   F2 = -(DY1 U1 + DY2 U2 + Idir(DY3 U3)) - DYb B
//   F3 = (TB - <b>)/2 - (ddz(T1) Ifull(U1) + ddz(T2) Ifull(U2) + Ifull(ddz T3) U3) - ddz(TB) Ifull(B)
//   FB = (T3 - <w>)/2 + Ri Idir(U3)
// The vertical derivatives of the totals act on the nodal totals (they commute
// with the horizontal transform); the ddz of the products is taken before the
// transform to modal space for the same reason.
template <typename T, bool THREE>
__global__ __launch_bounds__(ABS) void k_adj_products(const AdjArgs<T> A) {
    const int N = A.N3;
    const T* zc = A.zc;
    const int P = A.N3p;
    for (size_t e = size_t(blockIdx.x) * ABS + threadIdx.x; e < A.n; e += size_t(gridDim.x) * ABS) {
        const size_t o = e - e % size_t(N);
        const int j = int(e - o);
        auto at = [&](const T* f, int jj) -> T { return f[o + jj]; };
        auto ibar = [&](const T* f, int jj) -> T {
            return jj == 0 ? T(0) : T(0.5) * at(f, jj - 1) + T(0.5) * at(f, jj);
        };
        auto idir = [&](const T* f, int jj) -> T {
            return jj < N - 1 ? T(0.5) * at(f, jj) + T(0.5) * at(f, jj + 1) : T(0.5) * at(f, N - 1);
        };
        auto ifull = [&](const T* f, int jj) -> T {
            return jj == 0 ? T(0) : zc[ZC_IFULL_LO * P + jj] * at(f, jj - 1) + zc[ZC_IFULL_DI * P + jj] * at(f, jj);
        };
        auto itil = [&](const T* f, int jj) -> T {
            return jj == 0 ? T(0) : zc[ZC_ITIL_LO * P + jj] * at(f, jj - 1) + zc[ZC_ITIL_DI * P + jj] * at(f, jj);
        };
        auto dz_nd = [&](const T* f, int jj) -> T {  // Neumann -> Dirichlet points
            return jj < 2 ? T(0) : (at(f, jj) - at(f, jj - 1)) * zc[ZC_DZND * P + jj];
        };
        // Dirichlet -> Neumann points of a column given point-wise by g
        auto dz_dn = [&](auto g, int jj) -> T {
            if (jj == 0) return T(0);
            if (jj < N - 1) return (g(jj + 1) - g(jj)) * zc[ZC_DZDN * P + jj];
            return (-g(N - 1)) * zc[ZC_DZDN * P + jj];
        };
        const T* U1 = A.U[0];
        const T* U2 = A.U[1];
        const T* U3 = A.U[2];
        const T* Bf = A.U[3];
        const T* T1 = A.Tt[0];
        const T* T2 = A.Tt[1];
        const T* T3 = A.Tt[2];
        const T* TB = A.Tt[3];
        const T u1 = at(U1, j), u3 = at(U3, j), bb = at(Bf, j);
        const T t1 = at(T1, j), t3 = at(T3, j), tb = at(TB, j);
        const T u2 = THREE ? at(U2, j) : T(0);
        const T t2 = THREE ? at(T2, j) : T(0);

        // horizontal averages of the direct buoyancy and vertical velocity:
        // the real part of their (0,0) modes (Integration.h:37-42)
        const T bave = A.bt[size_t(j) * A.A1].x;
        const T wave = A.u3t[size_t(j) * A.A1].x;

        // F1 (and F2)
        auto fh = [&](const T* const* D) -> T {
            T nn = at(D[0], j) * u1;
            if (THREE) nn += at(D[1], j) * u2;
            const T* D3 = D[2];
            auto pr = [&](int jj) -> T { return at(D3, jj) * at(U3, jj); };
            const T id = j < N - 1 ? T(0.5) * pr(j) + T(0.5) * pr(j + 1) : T(0.5) * pr(N - 1);
            T F = T(0) - (nn + id);
            return F - at(D[3], j) * bb;
        };
        // F3
        T nd2 = dz_nd(T1, j) * ifull(U1, j);
        if (THREE) nd2 += dz_nd(T2, j) * ifull(U2, j);
        auto dzt3 = [&](int jj) -> T { return dz_dn([&](int q) { return at(T3, q); }, jj); };
        const T ifdz3 = j == 0 ? T(0) : zc[ZC_IFULL_LO * P + j] * dzt3(j - 1) + zc[ZC_IFULL_DI * P + j] * dzt3(j);
        T F3 = (T(1) / T(2)) * (tb + T(-1) * bave);
        F3 -= nd2 + ifdz3 * u3;
        F3 -= dz_nd(TB, j) * ifull(Bf, j);
        // FB
        T FB = (T(1) / T(2)) * (t3 + T(-1) * wave);
        FB += A.Ri * idir(U3, j);

        A.Q[0][e] = u1 * t1;
        A.Q[1][e] = dz_dn([&](int q) { return ibar(U1, q) * at(T3, q); }, j) + fh(A.DX);
        A.Q[2][e] = itil(T1, j) * u3;
        auto dd = [&](int q) -> T { return idir(U3, q) * idir(T3, q); };
        A.Q[3][e] = (j < 2 ? T(0) : (dd(j) - dd(j - 1)) * zc[ZC_DZND * P + j]) + F3;
        A.Q[4][e] = bb * t1;
        A.Q[5][e] = dz_dn([&](int q) { return ibar(Bf, q) * at(T3, q); }, j) + FB;
        if (THREE) {
            A.Q[6][e] = u1 * t2;
            A.Q[7][e] = u2 * t2;
            A.Q[8][e] = u2 * t1;
            A.Q[9][e] = dz_dn([&](int q) { return ibar(U2, q) * at(T3, q); }, j) + fh(A.DY);
            A.Q[10][e] = itil(T2, j) * u3;
            A.Q[11][e] = bb * t2;
        }
    }
}

template <typename T>
hipError_t launch_adj_products(const AdjArgs<T>& a, bool three, hipStream_t s) {
    if (three)
        hipLaunchKernelGGL((k_adj_products<T, true>), dim3(grid_of<T>(a.n)), dim3(ABS), 0, s, a);
    else
        hipLaunchKernelGGL((k_adj_products<T, false>), dim3(grid_of<T>(a.n)), dim3(ABS), 0, s, a);
    return hipGetLastError();
}

// Explicit adjoint terms from the transformed products (modal layout
// [ja][z][j1]):
//   r1 = ikx M0 + M1 + iky M6,   r2 = iky M7 + ikx M8 + M9,
//   r3 = ikx M2 + M3 + iky M10,  rB = ikx M4 + M5 + iky M11.
template <typename T>
__global__ __launch_bounds__(ABS) void k_adj_combine(AdjCombine<T> C) {
    const size_t n = size_t(C.A1) * C.A2 * C.N3;
    for (size_t e = size_t(blockIdx.x) * ABS + threadIdx.x; e < n; e += size_t(gridDim.x) * ABS) {
        const size_t rowz = e / size_t(C.A1);
        const int j1 = int(e - rowz * size_t(C.A1));
        const int ja = int(rowz / size_t(C.N3));
        const T kx = C.kx[j1], ky = C.ky[ja];
        auto term = [&](int i, int kind) -> cx<T> {  // kind 0: ikx, 1: iky, 2: plain
            const cx<T> v = C.M[i][e];
            return kind == 0 ? cik(kx, v) : (kind == 1 ? cik(ky, v) : v);
        };
        cx<T> r1 = cadd(term(0, 0), term(1, 2));
        cx<T> r3 = cadd(term(2, 0), term(3, 2));
        cx<T> rB = cadd(term(4, 0), term(5, 2));
        if (C.three) {
            r1 = cadd(r1, term(6, 1));
            r3 = cadd(r3, term(10, 1));
            rB = cadd(rB, term(11, 1));
            C.out[1][e] = cadd(cadd(term(7, 1), term(8, 0)), term(9, 2));
        }
        C.out[0][e] = r1;
        C.out[2][e] = r3;
        C.out[3][e] = rB;
    }
}
template <typename T>
hipError_t launch_adj_combine(const AdjCombine<T>& c, hipStream_t s) {
    const size_t n = size_t(c.A1) * c.A2 * c.N3;
    hipLaunchKernelGGL((k_adj_combine<T>), dim3(grid_of<T>(n)), dim3(ABS), 0, s, c);
    return hipGetLastError();
}

// J integrand of UpdateAdjointVariables (IMEXRK.h:478-483): -(B - <b>) W with
// <b> the real part of the (0,0) mode of the modal buoyancy
template <typename T>
__global__ __launch_bounds__(ABS) void k_mix_product(T* out, const T* Bt, const T* Wt, const cx<T>* bt, size_t n,
                                                     int N3, int A1) {
    for (size_t i = size_t(blockIdx.x) * ABS + threadIdx.x; i < n; i += size_t(gridDim.x) * ABS) {
        const int z = int(i % size_t(N3));
        const T bave = bt[size_t(z) * A1].x;
        out[i] = (T(-1) * (Bt[i] + T(-1) * bave)) * Wt[i];
    }
}
template <typename T>
hipError_t launch_mix_product(T* out, const T* Bt, const T* Wt, const cx<T>* bt, size_t n, int N3, int A1,
                              hipStream_t s) {
    hipLaunchKernelGGL((k_mix_product<T>), dim3(grid_of<T>(n)), dim3(ABS), 0, s, out, Bt, Wt, bt, n, N3, A1);
    return hipGetLastError();
}

// real part of the (0,0) stack of a modal field (HorizontalAverage,
// Integration.h:37-42) into N3 doubles
template <typename T>
__global__ __launch_bounds__(ABS) void k_mean_profile(const cx<T>* f, double* out, int N3, int A1) {
    const int z = blockIdx.x * ABS + threadIdx.x;
    if (z < N3) out[z] = double(f[size_t(z) * A1].x);
}
template <typename T>
hipError_t launch_mean_profile(const cx<T>* f, double* out, int N3, int A1, hipStream_t s) {
    hipLaunchKernelGGL((k_mean_profile<T>), dim3((N3 + ABS - 1) / ABS), dim3(ABS), 0, s, f, out, N3, A1);
    return hipGetLastError();
}

// ---- field helpers ----
// stack(j1, j2) *= exp(i j1 shift) over the active modes (ModalField::PhaseShift,
// Field.h:1168-1174); the active j1 equal their wavenumbers
template <typename T>
__global__ __launch_bounds__(ABS) void k_phase_shift(cx<T>* f, size_t n, int A1, double shift) {
    for (size_t e = size_t(blockIdx.x) * ABS + threadIdx.x; e < n; e += size_t(gridDim.x) * ABS) {
        const int j1 = int(e % size_t(A1));
        double sn, cs;
        sincos(double(j1) * shift, &sn, &cs);
        const T c = T(cs), s = T(sn);
        const cx<T> v = f[e];
        cx<T> r;
        r.x = v.x * c - v.y * s;
        r.y = v.x * s + v.y * c;
        f[e] = r;
    }
}
template <typename T>
hipError_t launch_phase_shift(cx<T>* f, size_t n, int A1, double shift, hipStream_t s) {
    hipLaunchKernelGGL((k_phase_shift<T>), dim3(grid_of<T>(n)), dim3(ABS), 0, s, f, n, A1, shift);
    return hipGetLastError();
}

// stack(0, 0) += value on every z (the real constant of RemoveAverage,
// Integration.h:72-75)
template <typename T>
__global__ __launch_bounds__(ABS) void k_add_mean(cx<T>* f, int N3, int A1, T value) {
    const int z = blockIdx.x * ABS + threadIdx.x;
    if (z < N3) f[size_t(z) * A1].x += value;
}
template <typename T>
hipError_t launch_add_mean(cx<T>* f, int N3, int A1, T value, hipStream_t s) {
    hipLaunchKernelGGL((k_add_mean<T>), dim3((N3 + ABS - 1) / ABS), dim3(ABS), 0, s, f, N3, A1, value);
    return hipGetLastError();
}

#define SF_ADJ_INST(T)                                                                                     \
    template hipError_t launch_lerp<T>(cx<T>*, const cx<T>*, const cx<T>*, T, size_t, hipStream_t);       \
    template hipError_t launch_add_profile<T>(T*, const T*, size_t, int, hipStream_t);                    \
    template hipError_t launch_adj_products<T>(const AdjArgs<T>&, bool, hipStream_t);                     \
    template hipError_t launch_adj_combine<T>(const AdjCombine<T>&, hipStream_t);                         \
    template hipError_t launch_mix_product<T>(T*, const T*, const T*, const cx<T>*, size_t, int, int, hipStream_t); \
    template hipError_t launch_mean_profile<T>(const cx<T>*, double*, int, int, hipStream_t);             \
    template hipError_t launch_phase_shift<T>(cx<T>*, size_t, int, double, hipStream_t);                    \
    template hipError_t launch_add_mean<T>(cx<T>*, int, int, T, hipStream_t);

SF_ADJ_INST(float)
SF_ADJ_INST(double)

}  // namespace sf
