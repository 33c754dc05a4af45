// sf_kernels.hip -- gfx950 kernels for the IMEX-RK timestep.
//
// Per RK stage the reference runs ExplicitRK, BuildRHS, FinishRHS,
// CrankNicolson, RemoveDivergence, FilterAll and PopulateNodalVariables
// (IMEXRK.cpp:8-24).  Here a stage is
//   3D: k_yinv -> k_xpass -> k_yfwd -> k_modal
//   2D:           k_xpass ->           k_modal
// * k_yinv   inverse y transform of the active spectra (y part of ToNodal)
// * k_xpass  inverse x transform, all physical-space products of BuildRHS
//            (Stratiflow.h:112-164) and forward x transform, with the
//            x-derivative factors and signs applied so that only 8 (3D) or
//            3 (2D) x-spectral fields leave the kernel instead of 11 or 6
// * k_yfwd   forward y transform, i*ky factors, 1/(N1*N2) scale (ToModal)
// * k_modal  everything per Fourier mode: explicit RK terms, linear
//            buoyancy and stratification terms, FinishRHS, Crank-Nicolson
//            matvec and implicit solves, divergence, Poisson solve,
//            projection and pressure update, all on one wavefront.
// Modal fields hold only the active (dealiased) modes, [j2a][j1][z].
#include "sf_fft.h"
#include "sf_kernels.h"
#include "sf_tri.h"

#include <type_traits>

namespace sf {

constexpr int BS = 256;

// ---------------------------------------------------------------------------
// k_modal
// ---------------------------------------------------------------------------
template <typename T, int M>
__device__ __forceinline__ void nb_prev(const cx<T> (&x)[M], cx<T> (&xm)[M], int lane) {
    using V = cx<T>;
    V s;
    s.x = __shfl(x[M - 1].x, (lane - 1) & 63, 64);
    s.y = __shfl(x[M - 1].y, (lane - 1) & 63, 64);
    if (lane == 0) s = czero<V>();
    xm[0] = s;
#pragma unroll
    for (int t = 1; t < M; ++t) xm[t] = x[t - 1];
}
template <typename T, int M>
__device__ __forceinline__ void nb_next(const cx<T> (&x)[M], cx<T> (&xp)[M], int lane) {
    using V = cx<T>;
    V s;
    s.x = __shfl(x[0].x, (lane + 1) & 63, 64);
    s.y = __shfl(x[0].y, (lane + 1) & 63, 64);
    if (lane == 63) s = czero<V>();
#pragma unroll
    for (int t = 0; t < M - 1; ++t) xp[t] = x[t + 1];
    xp[M - 1] = s;
}

// kind: 0 velocity Neumann, 1 velocity Dirichlet, 2 buoyancy Neumann,
// 3 pressure Laplacian.  Matrix rows as built in IMEXRK.h:55-78, 502-545
// with Neumannify/Dirichlify (Differentiation.h:17-59).
template <typename T, int M>
__device__ __forceinline__ void build_rows(int kind, const ModalArgs<T>& A, const int (&row)[M], T kk, T kl,
                                           bool zero_mode, T coef, T (&a)[M], T (&b)[M], T (&c)[M]) {
    const int N = A.N3;
    const bool dir = (kind == 1);
#pragma unroll
    for (int t = 0; t < M; ++t) {
        const int j = row[t];
        if (j >= N) {
            a[t] = 0;
            b[t] = 1;
            c[t] = 0;
            continue;
        }
        const T lo = A.zc[(dir ? ZC_DZZD_LO : ZC_DZZN_LO) * A.N3p + j];
        const T di = A.zc[(dir ? ZC_DZZD_DI : ZC_DZZN_DI) * A.N3p + j];
        const T up = A.zc[(dir ? ZC_DZZD_UP : ZC_DZZN_UP) * A.N3p + j];
        const T ldi = (di + kk) + kl;
        T aa, bb, cc;
        if (kind == 3) {
            aa = lo;
            bb = ldi;
            cc = up;
        } else {
            aa = -(coef * lo);
            bb = T(1) - coef * ldi;
            cc = -(coef * up);
        }
        if (j == 0) aa = 0;
        if (j == N - 1) cc = 0;
        if (!dir) {
            if (j == 0) {
                bb = 1;
                cc = 0;
            } else if (j == 1) {
                aa = 0;
                bb = 1;
                cc = -1;
            } else if (j == N - 2) {
                aa = 1;
                bb = -1;
                cc = 0;
            } else if (j == N - 1) {
                aa = 0;
                bb = 1;
            }
        } else if (j <= 2 || j >= N - 2) {
            aa = 0;
            bb = 1;
            cc = 0;
        }
        if (kind == 3 && zero_mode && j <= 1) {
            aa = 0;
            bb = 1;
            cc = 0;
        }
        a[t] = aa;
        b[t] = bb;
        c[t] = cc;
    }
}

template <typename T, int M>
__global__ __launch_bounds__(BS) void k_modal(ModalArgs<T> A) {
    using V = cx<T>;
    const int lane = threadIdx.x & 63;
    const int mode = blockIdx.x * (BS / 64) + (threadIdx.x >> 6);
    if (mode >= A.nmodes) return;  // wave-uniform
    const int j1 = mode % A.A1;
    const int j2 = mode / A.A1;
    const bool zero_mode = (mode == 0);
    const T kx = A.kx[j1], ky = A.ky[j2], kk = A.kx2[j1], kl = A.ky2[j2];
    const int N = A.N3;
    const size_t base = size_t(mode) * size_t(N);
    const bool three = A.threeD != 0;

    int row[M];
    bool ok[M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
        row[t] = lane * M + t;
        ok[t] = row[t] < N;
    }
    auto coefz = [&](int c, int t) -> T { return ok[t] ? A.zc[c * A.N3p + row[t]] : T(0); };
    auto load = [&](const V* f, V (&x)[M]) {
#pragma unroll
        for (int t = 0; t < M; ++t) x[t] = (f != nullptr && ok[t]) ? f[base + row[t]] : czero<V>();
    };
    auto store = [&](V* f, const V (&x)[M]) {
        if (f == nullptr) return;
#pragma unroll
        for (int t = 0; t < M; ++t)
            if (ok[t]) f[base + row[t]] = x[t];
    };
    // ZeroEnds, Field.h:475-486
    auto zero_ends = [&](V (&x)[M], bool dir) {
#pragma unroll
        for (int t = 0; t < M; ++t) {
            const int j = row[t];
            if (j <= 1 || j >= N - 2 || (dir && j == 2)) x[t] = czero<V>();
        }
    };

    V u1[M], u2[M], u3[M], bb[M], pp[M];
    load(A.u1, u1);
    if (three) load(A.u2, u2);
    else
#pragma unroll
        for (int t = 0; t < M; ++t) u2[t] = czero<V>();
    load(A.u3, u3);
    load(A.p, pp);

    if (A.mode == 0) {
        load(A.b, bb);
        V rn1[M], rn2[M], rn3[M], rnB[M];
        load(A.n1, rn1);
        if (three) load(A.n2, rn2);
        load(A.n3, rn3);
        load(A.nB, rnB);
        // BuildRHS linear terms, IMEXRK.cpp:140-147:
        //   r3 += Ri * ReinterpolateFull(b - <b>),  rB -= S * ReinterpolateDirichlet(u3)
        {
            V bm[M];
            nb_prev<T, M>(bb, bm, lane);
            if (!zero_mode) {
#pragma unroll
                for (int t = 0; t < M; ++t) {
                    const V ifb = cadd(cscale(coefz(ZC_IFULL_LO, t), bm[t]), cscale(coefz(ZC_IFULL_DI, t), bb[t]));
                    rn3[t] = cadd(cscale(A.Ri, ifb), rn3[t]);
                }
            }
            V u3p[M];
            nb_next<T, M>(u3, u3p, lane);
#pragma unroll
            for (int t = 0; t < M; ++t) {
                const V idr = cadd(cscale(T(0.5), u3[t]), cscale(T(0.5), u3p[t]));
                rnB[t] = csub(rnB[t], cscale(coefz(ZC_S, t), idr));
            }
        }
        store(A.rn1, rn1);
        if (three) store(A.rn2, rn2);
        store(A.rn3, rn3);
        store(A.rnB, rnB);

        // ExplicitRK + FinishRHS + CrankNicolson explicit half, IMEXRK.cpp:75-134
        auto assemble = [&](V (&R)[M], const V (&u)[M], const V* ro, const V (&rn)[M], const V (&gp)[M], bool dir,
                            T c) {
            V ro_[M];
            load(ro, ro_);
            V um[M], up[M];
            nb_prev<T, M>(u, um, lane);
            nb_next<T, M>(u, up, lane);
            const int zlo = dir ? ZC_DZZD_LO : ZC_DZZN_LO;
#pragma unroll
            for (int t = 0; t < M; ++t) {
                V r = cadd(cadd(u[t], cscale(A.hz, ro_[t])), cscale(A.mh, gp[t]));
                r = cadd(r, cscale(A.hb, rn[t]));
                const V dzz = cadd(cadd(cscale(coefz(zlo, t), um[t]), cscale(coefz(zlo + 1, t), u[t])),
                                   cscale(coefz(zlo + 2, t), up[t]));
                const V lap = cadd(cadd(cscale(kk, u[t]), cscale(kl, u[t])), dzz);
                R[t] = cadd(r, cscale(c, lap));
            }
            zero_ends(R, dir);
        };

        V R[M], gp[M];
        TriFactor<T, M> F;
        T ta[M], tb[M], tc[M];
        // u1 and u2 share the velocity Neumann matrix
        build_rows<T, M>(0, A, row, kk, kl, zero_mode, A.cv, ta, tb, tc);
        tri_factor<T, M>(F, ta, tb, tc, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) gp[t] = cik(kx, pp[t]);
        assemble(R, u1, A.ro1, rn1, gp, false, A.cv);
        tri_apply<T, M>(F, R, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) u1[t] = R[t];
        if (three) {
#pragma unroll
            for (int t = 0; t < M; ++t) gp[t] = cik(ky, pp[t]);
            assemble(R, u2, A.ro2, rn2, gp, false, A.cv);
            tri_apply<T, M>(F, R, lane);
#pragma unroll
            for (int t = 0; t < M; ++t) u2[t] = R[t];
        }
        // u3: velocity Dirichlet matrix; pressure gradient ddz(p) (Neumann -> Dirichlet)
        {
            V pm[M];
            nb_prev<T, M>(pp, pm, lane);
#pragma unroll
            for (int t = 0; t < M; ++t) gp[t] = cscale(coefz(ZC_DZND, t), csub(pp[t], pm[t]));
        }
        build_rows<T, M>(1, A, row, kk, kl, zero_mode, A.cv, ta, tb, tc);
        tri_factor<T, M>(F, ta, tb, tc, lane);
        assemble(R, u3, A.ro3, rn3, gp, true, A.cv);
        tri_apply<T, M>(F, R, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) u3[t] = R[t];
        // b: buoyancy Neumann matrix, no pressure term
#pragma unroll
        for (int t = 0; t < M; ++t) gp[t] = czero<V>();
        build_rows<T, M>(2, A, row, kk, kl, zero_mode, A.cb, ta, tb, tc);
        tri_factor<T, M>(F, ta, tb, tc, lane);
        assemble(R, bb, A.roB, rnB, gp, false, A.cb);
        tri_apply<T, M>(F, R, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) bb[t] = R[t];
        store(A.b, bb);
    }

    // RemoveDivergence, IMEXRK.cpp:43-73
    {
        V u3p[M], dv[M];
        nb_next<T, M>(u3, u3p, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) {
            const V dz = cscale(coefz(ZC_DZDN, t), csub(u3p[t], u3[t]));
            V h = cik(kx, u1[t]);
            if (three) h = cadd(h, cik(ky, u2[t]));
            dv[t] = cadd(h, dz);
        }
        zero_ends(dv, false);
        if (std::is_same<T, float>::value && zero_mode) {
            // The pinned Neumann Poisson problem of the (0,0) mode
            // (IMEXRK.h:68-74) has a condition number ~N3^2 * (L3/dz)^2;
            // it is solved in FP64 so that q(0,0) is not left with FP32
            // rounding amplified by that conditioning.
            TriFactor<double, M> F;
            double ta[M], tb[M], tc[M];
            T fa[M], fb[M], fc[M];
            build_rows<T, M>(3, A, row, kk, kl, zero_mode, T(0), fa, fb, fc);
#pragma unroll
            for (int t = 0; t < M; ++t) {
                ta[t] = fa[t];
                tb[t] = fb[t];
                tc[t] = fc[t];
            }
            tri_factor<double, M>(F, ta, tb, tc, lane);
            cx<double> dd[M];
#pragma unroll
            for (int t = 0; t < M; ++t) dd[t] = mkc<double>(double(dv[t].x), double(dv[t].y));
            tri_apply<double, M>(F, dd, lane);
            // finish the projection in FP64 for this mode as well
            cx<double> qm[M];
            nb_prev<double, M>(dd, qm, lane);
#pragma unroll
            for (int t = 0; t < M; ++t) {
                const double dzq = double(coefz(ZC_DZND, t)) * (dd[t].x - qm[t].x);
                const double dzqi = double(coefz(ZC_DZND, t)) * (dd[t].y - qm[t].y);
                u3[t].x = T(double(u3[t].x) - dzq);
                u3[t].y = T(double(u3[t].y) - dzqi);
                pp[t].x = T(double(pp[t].x) + double(A.pmult) * dd[t].x);
                pp[t].y = T(double(pp[t].y) + double(A.pmult) * dd[t].y);
            }
            // kx = ky = 0: u1 and u2 are unchanged
        } else {
            TriFactor<T, M> F;
            T ta[M], tb[M], tc[M];
            build_rows<T, M>(3, A, row, kk, kl, zero_mode, T(0), ta, tb, tc);
            tri_factor<T, M>(F, ta, tb, tc, lane);
            tri_apply<T, M>(F, dv, lane);  // dv now holds q
            V qm[M];
            nb_prev<T, M>(dv, qm, lane);
#pragma unroll
            for (int t = 0; t < M; ++t) {
                u1[t] = csub(u1[t], cik(kx, dv[t]));
                if (three) u2[t] = csub(u2[t], cik(ky, dv[t]));
                u3[t] = csub(u3[t], cscale(coefz(ZC_DZND, t), csub(dv[t], qm[t])));
                pp[t] = cadd(pp[t], cscale(A.pmult, dv[t]));
            }
        }
    }
    store(A.u1, u1);
    if (three) store(A.u2, u2);
    store(A.u3, u3);
    store(A.p, pp);
}

template <typename T>
hipError_t launch_modal(const ModalArgs<T>& a, hipStream_t s) {
    const int M = (a.N3 + 63) / 64;
    dim3 grid((a.nmodes + BS / 64 - 1) / (BS / 64));
    switch (M) {
#define SF_M(m)                                                     \
    case m:                                                         \
        hipLaunchKernelGGL((k_modal<T, m>), grid, dim3(BS), 0, s, a); \
        break;
        SF_M(1) SF_M(2) SF_M(3) SF_M(4) SF_M(5) SF_M(6) SF_M(7) SF_M(8)
#undef SF_M
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int max_rows_per_lane() { return 8; }

// ---------------------------------------------------------------------------
// k_xpass: inverse x FFT, products, forward x FFT
// ---------------------------------------------------------------------------
template <int N, typename T, int NI>
constexpr int xpass_tz() {
    // LDS: NI*(TZ+2) + 3*TZ lines of N complex values, kept under ~80 KB
    constexpr int bytes = 80 * 1024;
    constexpr int lines = bytes / (N * int(sizeof(cx<T>)));
    constexpr int tz = (lines - 2 * NI) / (NI + 3);
    return tz < 1 ? 1 : (tz > 8 ? 8 : tz);
}

template <typename T, int N1, int TZ, bool THREED, bool LIN>
__global__ __launch_bounds__(BS) void k_xpass(XArgs<T> A) {
    using V = cx<T>;
    constexpr int NLEV = TZ + 2;
    constexpr int NI = LIN ? 4 : 2;      // inverse pair lines per level
    constexpr int NLI = NI * NLEV;
    constexpr int NLP = 3 * TZ;          // product pair lines per half
    __shared__ V bufI[NLI * N1];
    __shared__ V bufP[NLP * N1];

    const int zt = blockIdx.x;
    const int n2 = blockIdx.y;
    const int z0 = zt * TZ;
    const int N3 = A.N3, A1 = A.A1;
    const int tid = threadIdx.x;

    // 1) Hermitian-extended spectra, two real fields per complex line
    for (int idx = tid; idx < NLI * N1; idx += BS) {
        const int lev = idx % NLEV;
        const int rest = idx / NLEV;
        const int j = rest % N1;
        const int pair = rest / N1;
        const int zz = z0 - 1 + lev;
        V val = czero<V>();
        if (zz >= 0 && zz < N3) {
            int jj = -1;
            bool cj = false;
            if (j < A1) jj = j;
            else if (j > N1 - A1) {
                jj = N1 - j;
                cj = true;
            }
            if (jj >= 0) {
                const int src = pair & 1;
                const bool bas = pair >= 2;
                const V* const* in = bas ? A.base : A.in;
                const V* fa;
                const V* fb;
                if (src == 0) {
                    fa = in[0];
                    fb = in[2];
                } else if (THREED) {
                    fa = in[1];
                    fb = in[3];
                } else {
                    fa = in[3];
                    fb = nullptr;
                }
                const size_t off = size_t(n2) * A.in_n2_stride + size_t(jj) * N3 + zz;
                V X = fa[off];
                V Y = fb ? fb[off] : czero<V>();
                if (jj == 0) {  // c2r ignores the imaginary part of the DC term
                    X.y = 0;
                    Y.y = 0;
                }
                if (cj) {
                    X.y = -X.y;
                    Y.y = -Y.y;
                }
                val.x = X.x - Y.y;
                val.y = X.y + Y.x;
            }
        }
        bufI[(pair * NLEV + lev) * N1 + j] = val;
    }
    __syncthreads();
    lds_fft<T, N1, NLI, N1, BS, +1>(bufI, A.tw);

    // nodal accessors (zero outside 0..N3-1 because those lines were zero)
    auto nod = [&](int pair, int lev, int n, bool im) -> T {
        const V v = bufI[(pair * NLEV + lev) * N1 + n];
        return im ? v.y : v.x;
    };
    auto ubar = [&](int j) -> T { return (j >= 0 && j < N3) ? A.zc[ZC_UBAR * A.N3p + j] : T(0); };

    const int nhalf = THREED ? 2 : 1;
    for (int half = 0; half < nhalf; ++half) {
        // 2) products for TZ levels
        for (int idx = tid; idx < TZ * N1; idx += BS) {
            const int n = idx % N1;
            const int t = idx / N1;
            const int j = z0 + t;
            const int lc = t + 1, lm = t, lp = t + 2;
            V o0 = czero<V>(), o1 = czero<V>(), o2 = czero<V>();
            if (j < N3) {
                const T itl = A.zc[ZC_ITIL_LO * A.N3p + j], itd = A.zc[ZC_ITIL_DI * A.N3p + j];
                const T dDN = A.zc[ZC_DZDN * A.N3p + j], dND = A.zc[ZC_DZND * A.N3p + j];
                const T hc = (j >= 1) ? T(0.5) : T(0);  // ReinterpolateBar row j
                const T u3m = nod(0, lm, n, true), u3c = nod(0, lc, n, true), u3p = nod(0, lp, n, true);
                const T B_m = nod(1, lm, n, THREED), B_c = nod(1, lc, n, THREED), B_p = nod(1, lp, n, THREED);
                // total streamwise velocity U1 + U_
                const T u1m = nod(0, lm, n, false) + ubar(j - 1);
                const T u1c = nod(0, lc, n, false) + ubar(j);
                const T u1p = nod(0, lp, n, false) + ubar(j + 1);
                T u2m = 0, u2c = 0, u2p = 0;
                if (THREED) {
                    u2m = nod(1, lm, n, false);
                    u2c = nod(1, lc, n, false);
                    u2p = nod(1, lp, n, false);
                }
                if (!LIN) {
                    if (half == 0) {
                        // P1 = U1t^2 ; P2 = ddz(Ibar(U1t)*U3)          (IMEXRK.cpp:155-159)
                        const T P1 = u1c * u1c;
                        const T gc = (hc * u1m + hc * u1c) * u3c, gp = (T(0.5) * u1c + T(0.5) * u1p) * u3p;
                        const T P2 = (gp - gc) * dDN;
                        // P3 = Itilde(U1t)*U3 ; Q4 = ddz(Idir(U3)^2)    (IMEXRK.cpp:161-163)
                        const T P3 = (itl * u1m + itd * u1c) * u3c;
                        const T ic = T(0.5) * u3c + T(0.5) * u3p, im = T(0.5) * u3m + T(0.5) * u3c;
                        const T Q4 = (ic * ic - im * im) * dND;
                        // P5 = ddz(Ibar(B)*U3) ; P6 = U1t*B            (IMEXRK.cpp:182-192)
                        const T gbc = (hc * B_m + hc * B_c) * u3c, gbp = (T(0.5) * B_c + T(0.5) * B_p) * u3p;
                        const T P5 = (gbp - gbc) * dDN;
                        const T P6 = u1c * B_c;
                        o0 = mkc<T>(P1, P2);
                        o1 = mkc<T>(P3, Q4);
                        o2 = mkc<T>(P5, P6);
                    } else {
                        // P7 = U2^2 ; P8 = ddz(Ibar(U2)*U3)             (IMEXRK.cpp:167-171)
                        const T P7 = u2c * u2c;
                        const T gc = (hc * u2m + hc * u2c) * u3c, gp = (T(0.5) * u2c + T(0.5) * u2p) * u3p;
                        const T P8 = (gp - gc) * dDN;
                        // P9 = Itilde(U2)*U3 ; P10 = U1t*U2 ; P11 = U2*B (IMEXRK.cpp:173-188)
                        const T P9 = (itl * u2m + itd * u2c) * u3c;
                        const T P10 = u1c * u2c;
                        const T P11 = u2c * B_c;
                        o0 = mkc<T>(P7, P8);
                        o1 = mkc<T>(P9, P10);
                        o2 = mkc<T>(P11, T(0));
                    }
                } else {
                    // BuildRHSLinear, IMEXRK.cpp:195-258; base state U*_tot, nnTemp = U1_tot + U_
                    const T v3m = nod(2, lm, n, true), v3c = nod(2, lc, n, true), v3p = nod(2, lp, n, true);
                    const T w1m = nod(2, lm, n, false) + ubar(j - 1);
                    const T w1c = nod(2, lc, n, false) + ubar(j);
                    const T w1p = nod(2, lp, n, false) + ubar(j + 1);
                    const T Bb_m = nod(3, lm, n, THREED), Bb_c = nod(3, lc, n, THREED), Bb_p = nod(3, lp, n, THREED);
                    T w2m = 0, w2c = 0, w2p = 0;
                    if (THREED) {
                        w2m = nod(3, lm, n, false);
                        w2c = nod(3, lc, n, false);
                        w2p = nod(3, lp, n, false);
                    }
                    // perturbation U1 without background
                    const T p1m = nod(0, lm, n, false), p1c = nod(0, lc, n, false), p1p = nod(0, lp, n, false);
                    (void)u1m;
                    (void)u1c;
                    (void)u1p;
                    if (half == 0) {
                        const T P1 = T(2) * (p1c * w1c);
                        const T gc = (hc * p1m + hc * p1c) * v3c + (hc * w1m + hc * w1c) * u3c;
                        const T gp = (T(0.5) * p1c + T(0.5) * p1p) * v3p + (T(0.5) * w1c + T(0.5) * w1p) * u3p;
                        const T P2 = (gp - gc) * dDN;
                        const T P3 = (itl * p1m + itd * p1c) * v3c + (itl * w1m + itd * w1c) * u3c;
                        const T ic = (T(0.5) * u3c + T(0.5) * u3p) * (T(0.5) * v3c + T(0.5) * v3p);
                        const T im = (T(0.5) * u3m + T(0.5) * u3c) * (T(0.5) * v3m + T(0.5) * v3c);
                        const T Q4 = T(2) * ((ic - im) * dND);
                        const T gbc = (hc * B_m + hc * B_c) * v3c + (hc * Bb_m + hc * Bb_c) * u3c;
                        const T gbp = (T(0.5) * B_c + T(0.5) * B_p) * v3p + (T(0.5) * Bb_c + T(0.5) * Bb_p) * u3p;
                        const T P5 = (gbp - gbc) * dDN;
                        const T P6 = B_c * w1c + Bb_c * p1c;
                        o0 = mkc<T>(P1, P2);
                        o1 = mkc<T>(P3, Q4);
                        o2 = mkc<T>(P5, P6);
                    } else {
                        const T P7 = T(2) * (u2c * w2c);
                        const T gc = (hc * u2m + hc * u2c) * v3c + (hc * w2m + hc * w2c) * u3c;
                        const T gp = (T(0.5) * u2c + T(0.5) * u2p) * v3p + (T(0.5) * w2c + T(0.5) * w2p) * u3p;
                        const T P8 = (gp - gc) * dDN;
                        const T P9 = (itl * u2m + itd * u2c) * v3c + (itl * w2m + itd * w2c) * u3c;
                        const T P10 = p1c * w2c + w1c * u2c;
                        const T P11 = B_c * w2c + Bb_c * u2c;
                        o0 = mkc<T>(P7, P8);
                        o1 = mkc<T>(P9, P10);
                        o2 = mkc<T>(P11, T(0));
                    }
                }
            }
            bufP[(0 * TZ + t) * N1 + n] = o0;
            bufP[(1 * TZ + t) * N1 + n] = o1;
            bufP[(2 * TZ + t) * N1 + n] = o2;
        }
        __syncthreads();
        lds_fft<T, N1, NLP, N1, BS, -1>(bufP, A.tw);

        // 3) unpack the two real spectra of each line, keep k < A1, combine
        for (int idx = tid; idx < TZ * A1; idx += BS) {
            const int t = idx % TZ;
            const int k = idx / TZ;
            const int j = z0 + t;
            if (j >= N3) continue;
            const T kxv = A.kx[k];
            V Fa[3], Fb[3];
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const V Z = bufP[(q * TZ + t) * N1 + k];
                const V W = bufP[(q * TZ + t) * N1 + ((N1 - k) % N1)];
                // Fa = (Z + conj W)/2 ; Fb = (Z - conj W)/(2i)
                Fa[q].x = T(0.5) * (Z.x + W.x);
                Fa[q].y = T(0.5) * (Z.y - W.y);
                Fb[q].x = T(0.5) * (Z.y + W.y);
                Fb[q].y = T(-0.5) * (Z.x - W.x);
            }
            const size_t off = size_t(n2) * A.out_n2_stride + size_t(k) * N3 + j;
            const T sc = A.scale;
            if (half == 0) {
                // r1 -= ikx P1 + P2 ; r3 -= ikx P3 + ddz P4 ; rB -= P5 + ikx P6
                const V r1 = cadd(cik(kxv, Fa[0]), Fb[0]);
                const V r3 = cadd(cik(kxv, Fa[1]), Fb[1]);
                const V rb = cadd(Fa[2], cik(kxv, Fb[2]));
                A.out[0][off] = cscale(-sc, r1);
                A.out[2][off] = cscale(-sc, r3);
                A.out[3][off] = cscale(-sc, rb);
            } else {
                // r2 -= iky P7 + P8 + ikx P10 ; r3 -= iky P9 ; r1 -= iky P10 ; rB -= iky P11
                const V a2 = cadd(Fb[0], cik(kxv, Fb[1]));
                A.out[1][off] = cscale(-sc, a2);
                A.out[4][off] = cscale(-sc, Fb[1]);
                A.out[5][off] = cscale(-sc, Fa[0]);
                A.out[6][off] = cscale(-sc, Fa[1]);
                A.out[7][off] = cscale(-sc, Fa[2]);
            }
        }
        __syncthreads();
    }
}

#define SF_SIZES(X) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

bool tz_supported(int n) {
    switch (n) {
#define SF_CASE(n) case n:
        SF_SIZES(SF_CASE)
#undef SF_CASE
        return true;
        default:
            return false;
    }
}

template <typename T, int N1, bool THREED, bool LIN>
static hipError_t launch_xpass_n(const XArgs<T>& a, int N2, hipStream_t s) {
    constexpr int TZ = xpass_tz<N1, T, LIN ? 4 : 2>();
    XArgs<T> b = a;
    b.nzt = (a.N3 + TZ - 1) / TZ;
    dim3 grid(b.nzt, THREED ? N2 : 1);
    hipLaunchKernelGGL((k_xpass<T, N1, TZ, THREED, LIN>), grid, dim3(BS), 0, s, b);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_xpass(const XArgs<T>& a, int N1, int N2, bool threeD, hipStream_t s) {
    switch (N1) {
#define SF_CASE(n)                                                                                     \
    case n:                                                                                            \
        if (a.linear)                                                                                  \
            return threeD ? launch_xpass_n<T, n, true, true>(a, N2, s)                                 \
                          : launch_xpass_n<T, n, false, true>(a, N2, s);                               \
        return threeD ? launch_xpass_n<T, n, true, false>(a, N2, s) : launch_xpass_n<T, n, false, false>(a, N2, s);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// y transforms (3D)
// ---------------------------------------------------------------------------
template <int N, typename T>
constexpr int ypass_tz() {
    constexpr int lines = (48 * 1024) / (N * int(sizeof(cx<T>)));
    constexpr int tz = lines / 2;
    return tz < 1 ? 1 : (tz > 16 ? 16 : tz);
}

// inverse: out_f[n2][j1][z] = sum_{j2 active} in_f[j2a][j1][z] e^{+2 pi i j2 n2/N2}
template <typename T, int N2, int TZ>
__global__ __launch_bounds__(BS) void k_yinv(YArgs<T> A) {
    using V = cx<T>;
    constexpr int NL = 2 * TZ;
    __shared__ V buf[NL * N2];
    const int z0 = blockIdx.x * TZ;
    const int j1 = blockIdx.y;
    const int N3 = A.N3, A1 = A.A1;
    const int h = N2 / 3;
    for (int f0 = 0; f0 < A.nf; f0 += 2) {
        for (int idx = threadIdx.x; idx < NL * N2; idx += BS) {
            const int t = idx % TZ;
            const int rest = idx / TZ;
            const int j2 = rest % N2;
            const int fl = rest / N2;
            const int f = f0 + fl;
            const int z = z0 + t;
            V v = czero<V>();
            int j2a = -1;
            if (j2 < h) j2a = j2;
            else if (j2 >= N2 - h + 1) j2a = j2 - N2 + 2 * h - 1;
            if (f < A.nf && z < N3 && j2a >= 0) v = A.in[f][(size_t(j2a) * A1 + j1) * N3 + z];
            buf[(fl * TZ + t) * N2 + j2] = v;
        }
        __syncthreads();
        lds_fft<T, N2, NL, N2, BS, +1>(buf, A.tw);
        for (int idx = threadIdx.x; idx < NL * N2; idx += BS) {
            const int t = idx % TZ;
            const int rest = idx / TZ;
            const int n2 = rest % N2;
            const int fl = rest / N2;
            const int f = f0 + fl;
            const int z = z0 + t;
            if (f < A.nf && z < N3) A.out[f][(size_t(n2) * A1 + j1) * N3 + z] = buf[(fl * TZ + t) * N2 + n2];
        }
        __syncthreads();
    }
}

// forward: X_f[n2][j1][z] -> active j2.  paired: out[r] = scale*(F(in[2r]) + i ky F(in[2r+1]))
template <typename T, int N2, int TZ>
__global__ __launch_bounds__(BS) void k_yfwd(YArgs<T> A) {
    using V = cx<T>;
    constexpr int NL = 2 * TZ;
    __shared__ V buf[NL * N2];
    const int z0 = blockIdx.x * TZ;
    const int j1 = blockIdx.y;
    const int N3 = A.N3, A1 = A.A1, A2 = A.A2;
    const int h = N2 / 3;
    const int nin = A.paired ? 2 * A.nf : A.nf;
    for (int f0 = 0; f0 < nin; f0 += 2) {
        for (int idx = threadIdx.x; idx < NL * N2; idx += BS) {
            const int t = idx % TZ;
            const int rest = idx / TZ;
            const int n2 = rest % N2;
            const int fl = rest / N2;
            const int f = f0 + fl;
            const int z = z0 + t;
            V v = czero<V>();
            if (f < nin && z < N3) v = A.in[f][(size_t(n2) * A1 + j1) * N3 + z];
            buf[(fl * TZ + t) * N2 + n2] = v;
        }
        __syncthreads();
        lds_fft<T, N2, NL, N2, BS, -1>(buf, A.tw);
        if (A.paired) {
            const int r = f0 / 2;
            for (int idx = threadIdx.x; idx < TZ * A2; idx += BS) {
                const int t = idx % TZ;
                const int j2a = idx / TZ;
                const int z = z0 + t;
                if (z >= N3) continue;
                const int j2 = (j2a < h) ? j2a : j2a + N2 - 2 * h + 1;
                const V Fa = buf[(0 * TZ + t) * N2 + j2];
                const V Fb = buf[(1 * TZ + t) * N2 + j2];
                const V v = cadd(Fa, cik(A.ky[j2a], Fb));
                A.out[r][(size_t(j2a) * A1 + j1) * N3 + z] = cscale(A.scale, v);
            }
        } else {
            for (int idx = threadIdx.x; idx < 2 * TZ * A2; idx += BS) {
                const int t = idx % TZ;
                const int rest = idx / TZ;
                const int j2a = rest % A2;
                const int fl = rest / A2;
                const int f = f0 + fl;
                const int z = z0 + t;
                if (f >= nin || z >= N3) continue;
                const int j2 = (j2a < h) ? j2a : j2a + N2 - 2 * h + 1;
                A.out[f][(size_t(j2a) * A1 + j1) * N3 + z] = cscale(A.scale, buf[(fl * TZ + t) * N2 + j2]);
            }
        }
        __syncthreads();
    }
}

template <typename T, int N2>
static hipError_t launch_y_n(const YArgs<T>& a, bool inv, hipStream_t s) {
    constexpr int TZ = ypass_tz<N2, T>();
    YArgs<T> b = a;
    b.nzt = (a.N3 + TZ - 1) / TZ;
    dim3 grid(b.nzt, a.A1);
    if (inv)
        hipLaunchKernelGGL((k_yinv<T, N2, TZ>), grid, dim3(BS), 0, s, b);
    else
        hipLaunchKernelGGL((k_yfwd<T, N2, TZ>), grid, dim3(BS), 0, s, b);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_yinv(const YArgs<T>& a, int N2, hipStream_t s) {
    switch (N2) {
#define SF_CASE(n) \
    case n:        \
        return launch_y_n<T, n>(a, true, s);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return hipErrorInvalidValue;
    }
}
template <typename T>
hipError_t launch_yfwd(const YArgs<T>& a, int N2, hipStream_t s) {
    switch (N2) {
#define SF_CASE(n) \
    case n:        \
        return launch_y_n<T, n>(a, false, s);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// single-field x transforms (ToNodal / ToModal for API fields)
// ---------------------------------------------------------------------------
template <typename T, int N1, int TZ>
__global__ __launch_bounds__(BS) void k_x_c2r(XTArgs<T> A) {
    using V = cx<T>;
    __shared__ V buf[TZ * N1];
    const int z0 = blockIdx.x * TZ;
    const int n2 = blockIdx.y;
    const int N3 = A.N3, A1 = A.A1;
    for (int idx = threadIdx.x; idx < TZ * N1; idx += BS) {
        const int t = idx % TZ;
        const int j = idx / TZ;
        const int z = z0 + t;
        V v = czero<V>();
        if (z < N3) {
            int jj = -1;
            bool cj = false;
            if (j < A1) jj = j;
            else if (j > N1 - A1) {
                jj = N1 - j;
                cj = true;
            }
            if (jj >= 0) {
                v = A.spec_in[size_t(n2) * A.spec_n2_stride + size_t(jj) * N3 + z];
                if (jj == 0) v.y = 0;
                if (cj) v.y = -v.y;
            }
        }
        buf[t * N1 + j] = v;
    }
    __syncthreads();
    lds_fft<T, N1, TZ, N1, BS, +1>(buf, A.tw);
    for (int idx = threadIdx.x; idx < TZ * N1; idx += BS) {
        const int t = idx % TZ;
        const int n = idx / TZ;
        const int z = z0 + t;
        if (z < N3) A.nodal_out[(size_t(n2) * N1 + n) * N3 + z] = buf[t * N1 + n].x;
    }
}

template <typename T, int N1, int TZ>
__global__ __launch_bounds__(BS) void k_x_r2c(XTArgs<T> A) {
    using V = cx<T>;
    __shared__ V buf[TZ * N1];
    const int z0 = blockIdx.x * TZ;
    const int n2 = blockIdx.y;
    const int N3 = A.N3, A1 = A.A1;
    for (int idx = threadIdx.x; idx < TZ * N1; idx += BS) {
        const int t = idx % TZ;
        const int n = idx / TZ;
        const int z = z0 + t;
        V v = czero<V>();
        if (z < N3) v.x = A.nodal_in[(size_t(n2) * N1 + n) * N3 + z];
        buf[t * N1 + n] = v;
    }
    __syncthreads();
    lds_fft<T, N1, TZ, N1, BS, -1>(buf, A.tw);
    for (int idx = threadIdx.x; idx < TZ * A1; idx += BS) {
        const int t = idx % TZ;
        const int k = idx / TZ;
        const int z = z0 + t;
        if (z < N3) A.spec_out[size_t(n2) * A.spec_n2_stride + size_t(k) * N3 + z] = cscale(A.scale, buf[t * N1 + k]);
    }
}

template <typename T, int N1>
static hipError_t launch_xt_n(const XTArgs<T>& a, int N2, bool c2r, hipStream_t s) {
    constexpr int TZ = ypass_tz<N1, T>();
    XTArgs<T> b = a;
    b.nzt = (a.N3 + TZ - 1) / TZ;
    dim3 grid(b.nzt, N2);
    if (c2r)
        hipLaunchKernelGGL((k_x_c2r<T, N1, TZ>), grid, dim3(BS), 0, s, b);
    else
        hipLaunchKernelGGL((k_x_r2c<T, N1, TZ>), grid, dim3(BS), 0, s, b);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_x_c2r(const XTArgs<T>& a, int N1, int N2, hipStream_t s) {
    switch (N1) {
#define SF_CASE(n) \
    case n:        \
        return launch_xt_n<T, n>(a, N2, true, s);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return hipErrorInvalidValue;
    }
}
template <typename T>
hipError_t launch_x_r2c(const XTArgs<T>& a, int N1, int N2, hipStream_t s) {
    switch (N1) {
#define SF_CASE(n) \
    case n:        \
        return launch_xt_n<T, n>(a, N2, false, s);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// FP64 inner product, InnerProd (Integration.h:117-146):
//   sum_z w_z sum_modes c_m Re(a conj b),  c = 1 for mode (0,0), 2 otherwise.
// Deterministic two-pass reduction.
// ---------------------------------------------------------------------------
constexpr int DOT_BLOCKS = 1024;

template <typename T>
__global__ __launch_bounds__(BS) void k_dot_partial(const cx<T>* __restrict__ a, const cx<T>* __restrict__ b,
                                                    size_t n, int N3, const T* __restrict__ w,
                                                    double* __restrict__ partial) {
    __shared__ double red[BS];
    double acc = 0;
    for (size_t e = size_t(blockIdx.x) * BS + threadIdx.x; e < n; e += size_t(gridDim.x) * BS) {
        const int z = int(e % size_t(N3));
        const size_t mode = e / size_t(N3);
        const cx<T> x = a[e], y = b[e];
        const double re = double(x.x) * double(y.x) + double(x.y) * double(y.y);
        acc += (mode == 0 ? 1.0 : 2.0) * double(w[z]) * re;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(BS) void k_dot_final(const double* __restrict__ partial, int n, double* out, int slot) {
    __shared__ double red[BS];
    double acc = 0;
    for (int i = threadIdx.x; i < n; i += BS) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = red[0];
}

template <typename T>
hipError_t launch_dot(const cx<T>* a, const cx<T>* b, int nmodes, int N3, const T* w, double* partial, double* out,
                      int slot, hipStream_t s) {
    const size_t n = size_t(nmodes) * size_t(N3);
    int blocks = int((n + BS - 1) / BS);
    if (blocks > DOT_BLOCKS) blocks = DOT_BLOCKS;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_dot_partial<T>), dim3(blocks), dim3(BS), 0, s, a, b, n, N3, w, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(BS), 0, s, partial, blocks, out, slot);
    return hipGetLastError();
}

// max |f + add[z]| over a nodal field (NodalField::Max, Field.h:771-786)
template <typename T>
__global__ __launch_bounds__(BS) void k_maxabs_partial(const T* __restrict__ f, size_t n, int N3,
                                                       const T* __restrict__ add, double* __restrict__ partial) {
    __shared__ double red[BS];
    double m = 0;
    for (size_t e = size_t(blockIdx.x) * BS + threadIdx.x; e < n; e += size_t(gridDim.x) * BS) {
        T v = f[e];
        if (add) v += add[e % size_t(N3)];
        const double a = fabs(double(v));
        m = a > m ? a : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(BS) void k_max_final(const double* __restrict__ partial, int n, double* out, int slot) {
    __shared__ double red[BS];
    double m = 0;
    for (int i = threadIdx.x; i < n; i += BS) m = fmax(m, partial[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = red[0];
}

template <typename T>
hipError_t launch_maxabs(const T* f, size_t n, int N3, const T* add, double* partial, double* out, int slot,
                         hipStream_t s) {
    int blocks = int((n + BS - 1) / BS);
    if (blocks > DOT_BLOCKS) blocks = DOT_BLOCKS;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_maxabs_partial<T>), dim3(blocks), dim3(BS), 0, s, f, n, N3, add, partial);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(BS), 0, s, partial, blocks, out, slot);
    return hipGetLastError();
}

// y = a*x + b*y over n complex values (x may equal y)
template <typename T>
__global__ __launch_bounds__(BS) void k_axpby(cx<T>* y, T a, const cx<T>* x, T b, size_t n) {
    for (size_t i = size_t(blockIdx.x) * BS + threadIdx.x; i < n; i += size_t(gridDim.x) * BS) {
        const cx<T> xv = x ? x[i] : czero<cx<T>>();
        const cx<T> yv = y[i];
        cx<T> r;
        r.x = a * xv.x + b * yv.x;
        r.y = a * xv.y + b * yv.y;
        y[i] = r;
    }
}
template <typename T>
hipError_t launch_axpby(cx<T>* y, T a, const cx<T>* x, T b, size_t n, hipStream_t s) {
    size_t blocks = (n + BS - 1) / BS;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_axpby<T>), dim3(unsigned(blocks)), dim3(BS), 0, s, y, a, x, b, n);
    return hipGetLastError();
}

// bc_op: 0 ZeroEnds Neumann, 1 ZeroEnds Dirichlet, 2 NeumannEnds (Field.h:475-500)
template <typename T>
__global__ __launch_bounds__(BS) void k_ends(cx<T>* f, int nstacks, int N3, int op) {
    const int s = blockIdx.x * BS + threadIdx.x;
    if (s >= nstacks) return;
    cx<T>* c = f + size_t(s) * N3;
    if (op == 2) {
        c[0] = c[1];
        c[N3 - 1] = c[N3 - 2];
    } else {
        c[0] = c[1] = c[N3 - 1] = c[N3 - 2] = czero<cx<T>>();
        if (op == 1) c[2] = czero<cx<T>>();
    }
}
template <typename T>
hipError_t launch_ends(cx<T>* f, int nstacks, int N3, int op, hipStream_t s) {
    hipLaunchKernelGGL((k_ends<T>), dim3((nstacks + BS - 1) / BS), dim3(BS), 0, s, f, nstacks, N3, op);
    return hipGetLastError();
}

// ddx (axis 0), ddy (axis 1), ddz (axis 2: Neumann->Dirichlet when zc row is
// ZC_DZND, Dirichlet->Neumann when ZC_DZDN; the caller passes the row).
template <typename T>
__global__ __launch_bounds__(BS) void k_dmul(cx<T>* out, const cx<T>* in, int A1, int A2, int N3, const T* kx,
                                             const T* ky, int axis, const T* zc, int N3p) {
    const size_t n = size_t(A1) * A2 * N3;
    for (size_t e = size_t(blockIdx.x) * BS + threadIdx.x; e < n; e += size_t(gridDim.x) * BS) {
        const int z = int(e % size_t(N3));
        const size_t mode = e / size_t(N3);
        const int j1 = int(mode % size_t(A1));
        const int j2 = int(mode / size_t(A1));
        cx<T> r;
        if (axis == 0) r = cik(kx[j1], in[e]);
        else if (axis == 1) r = cik(ky[j2], in[e]);
        else if (axis == 2) {  // Neumann in -> Dirichlet out
            const cx<T> fm = z > 0 ? in[e - 1] : czero<cx<T>>();
            r = cscale(zc[ZC_DZND * N3p + z], csub(in[e], fm));
        } else {  // Dirichlet in -> Neumann out
            const cx<T> fp = z + 1 < N3 ? in[e + 1] : czero<cx<T>>();
            r = cscale(zc[ZC_DZDN * N3p + z], csub(fp, in[e]));
        }
        out[e] = r;
    }
}
template <typename T>
hipError_t launch_dmul(cx<T>* out, const cx<T>* in, int A1, int A2, int N3, const T* kx, const T* ky, int axis,
                       const T* zc, int N3p, hipStream_t s) {
    const size_t n = size_t(A1) * A2 * N3;
    size_t blocks = (n + BS - 1) / BS;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_dmul<T>), dim3(unsigned(blocks)), dim3(BS), 0, s, out, in, A1, A2, N3, kx, ky, axis, zc,
                       N3p);
    return hipGetLastError();
}

#define SF_INST(T)                                                                                                  \
    template hipError_t launch_modal<T>(const ModalArgs<T>&, hipStream_t);                                           \
    template hipError_t launch_xpass<T>(const XArgs<T>&, int, int, bool, hipStream_t);                               \
    template hipError_t launch_yinv<T>(const YArgs<T>&, int, hipStream_t);                                           \
    template hipError_t launch_yfwd<T>(const YArgs<T>&, int, hipStream_t);                                           \
    template hipError_t launch_x_c2r<T>(const XTArgs<T>&, int, int, hipStream_t);                                   \
    template hipError_t launch_x_r2c<T>(const XTArgs<T>&, int, int, hipStream_t);                                   \
    template hipError_t launch_dot<T>(const cx<T>*, const cx<T>*, int, int, const T*, double*, double*, int,         \
                                      hipStream_t);                                                                  \
    template hipError_t launch_axpby<T>(cx<T>*, T, const cx<T>*, T, size_t, hipStream_t);                          \
    template hipError_t launch_maxabs<T>(const T*, size_t, int, const T*, double*, double*, int, hipStream_t);       \
    template hipError_t launch_ends<T>(cx<T>*, int, int, int, hipStream_t);                                          \
    template hipError_t launch_dmul<T>(cx<T>*, const cx<T>*, int, int, int, const T*, const T*, int, const T*, int, \
                                       hipStream_t);

SF_INST(float)
SF_INST(double)

}  // namespace sf
