// sf_kernels_impl.h -- gfx950 kernels for the IMEX-RK timestep.  The .hip
// files sf_kernels_{f32,f64}.hip and sf_batch_{f32,f64}.hip instantiate them
// (four translation units, compiled in parallel).
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
#pragma once
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
// zcoef(c, t): coefficient row c at the lane's row t (zero beyond N3).
template <typename T, int M, typename ZC>
__device__ __forceinline__ void build_rows(int kind, const ModalArgs<T>& A, const int (&row)[M], T kk, T kl,
                                           bool zero_mode, T coef, ZC zcoef, T (&a)[M], T (&b)[M], T (&c)[M]) {
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
        const T lo = zcoef(dir ? ZC_DZZD_LO : ZC_DZZN_LO, t);
        const T di = zcoef(dir ? ZC_DZZD_DI : ZC_DZZN_DI, t);
        const T up = zcoef(dir ? ZC_DZZD_UP : ZC_DZZN_UP, t);
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

// largest rows per lane of the modal kernel (N3 <= 64 * MODAL_MAX_M)
constexpr int MODAL_MAX_M = 8;

// modes per block of the modal kernel: 4 consecutive j1, one wavefront each
// (two such blocks share a CU at the register use of M = 6)
template <typename T, int M>
constexpr int modal_mb() {
    return (std::is_same<T, float>::value ? M <= 6 : M <= 3) ? 8 : 4;
}

// One block owns MB modes (ja, j10 .. j10+MB-1), one wavefront each.  Every
// field access goes through an LDS tile of N3 rows x MB modes so that global
// memory sees runs of MB consecutive j1 ([ja][z][j1] layout); wave w then
// holds the column of mode j10+w with lane l owning rows l*M .. l*M+M-1.
// Waves past A1 run on zeros and store nothing; they keep the block's barriers.
// FULL (N3 == 64*M): every tile row exists, so a thread keeps one base offset
// per table and reaches its row t through the scalar offset (fewer live
// registers than an offset per row).
template <typename T, int M, int MB, bool FULL = false>
__device__ __forceinline__ void modal_body(const ModalArgs<T>& A, const int ja, const int j10, cx<T>* const tile) {
    using V = cx<T>;
    constexpr int TP = MB + 1;  // tile row pitch (tile: 64*M*TP values of LDS)
    const int lane = threadIdx.x & 63;
    // wave-uniform: the mode's scalars (j1, wavenumbers, flags) live in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int A1 = A.A1;
    const bool active = j10 + wave < A1;
    const int j1 = active ? j10 + wave : A1 - 1;
    const int j2 = ja;
    const bool zero_mode = (ja == 0 && j10 + wave == 0);
    const T kx = A.kx[j1], ky = A.ky[j2], kk = A.kx2[j1], kl = A.ky2[j2];
    const int N = A.N3;
    const bool three = A.threeD != 0;
    (void)active;

    int row[M];
#pragma unroll
    for (int t = 0; t < M; ++t) row[t] = lane * M + t;
    // coefficient table: row c of the lane's row t through the scalar offset
    // c * N3p; rows at or beyond N3 read zero
    const rsrc_t zr = make_rsrc(A.zc, size_t(ZC_COUNT) * A.N3p * sizeof(T));
    int zoff[FULL ? 1 : M];
    if constexpr (FULL) {
        zoff[0] = lane * M * int(sizeof(T));
    } else {
#pragma unroll
        for (int t = 0; t < M; ++t) zoff[t] = row[t] < N ? row[t] * int(sizeof(T)) : BUF_OOB;
    }
    auto coefz = [&](int c, int t) -> T {
        if constexpr (FULL) return bld<T>(zr, zoff[0], (c * A.N3p + t) * int(sizeof(T)));
        else return bld<T>(zr, zoff[t], c * A.N3p * int(sizeof(T)));
    };
    // tile element owned by this thread in the cooperative (coalesced) phase:
    // rows z = tid/MB + t*64 for t < M (N3 <= 64*M), mode w = tid % MB.  Every
    // modal field has the same layout, so one table of byte offsets serves all
    // of them; elements outside the field (z >= N3, j1 >= A1) get an
    // out-of-range offset and read zero / are not written.
    const int cw = int(threadIdx.x) % MB;
    const int cz = int(threadIdx.x) / MB;
    const int cj1 = j10 + cw;
    const size_t fbytes = size_t(A1) * size_t(A.A2) * size_t(N) * sizeof(V);
    int foff[FULL ? 1 : M];
    if constexpr (FULL) {
        foff[0] = (cj1 < A1) ? int(sidx(ja, cz, cj1, N, A1) * sizeof(V)) : BUF_OOB;
    } else {
#pragma unroll
        for (int t = 0; t < M; ++t) {
            const int z = cz + t * 64;
            foff[t] = (z < N && cj1 < A1) ? int(sidx(ja, z, cj1, N, A1) * sizeof(V)) : BUF_OOB;
        }
    }
    // element t: vector offset fo(t) plus scalar offset fs(t) (64 rows of A1 modes per t)
    const int fstride = 64 * A1 * int(sizeof(V));
    auto fo = [&](int t) -> int {
        if constexpr (FULL) return foff[0];
        else return foff[t];
    };
    auto fs = [&](int t) -> int { return FULL ? t * fstride : 0; };
    // issue: loads of this thread's tile elements into g (left in flight);
    // finish: through the LDS tile into the wave's column x.  Several issues
    // before their finishes overlap the memory latency of those fields.
    auto issue = [&](const V* f, V (&g)[M]) {
        if (f == nullptr) return;  // block-uniform
        const rsrc_t r = make_rsrc(f, fbytes);
#pragma unroll
        for (int t = 0; t < M; ++t) g[t] = bld<V>(r, fo(t), fs(t));
    };
    auto finish = [&](const V* f, const V (&g)[M], V (&x)[M]) {
        if (f == nullptr) {  // block-uniform
#pragma unroll
            for (int t = 0; t < M; ++t) x[t] = czero<V>();
            return;
        }
        sync_lds();  // the previous tile has been read by every wave
        // rows z >= N3 receive the zeros their loads returned, so the padding
        // rows of every column are zero
#pragma unroll
        for (int t = 0; t < M; ++t) tile[(cz + t * 64) * TP + cw] = g[t];
        sync_lds();
#pragma unroll
        for (int t = 0; t < M; ++t) x[t] = tile[row[t] * TP + wave];
    };
    auto store = [&](V* f, const V (&x)[M]) {
        if (f == nullptr) return;
        const rsrc_t r = make_rsrc(f, fbytes);
        sync_lds();
#pragma unroll
        for (int t = 0; t < M; ++t) tile[row[t] * TP + wave] = x[t];
        sync_lds();
#pragma unroll
        for (int t = 0; t < M; ++t) bst<V>(tile[(cz + t * 64) * TP + cw], r, fo(t), fs(t));
    };
    // ZeroEnds, Field.h:475-486
    auto zero_ends = [&](V (&x)[M], bool dir) {
#pragma unroll
        for (int t = 0; t < M; ++t) {
            const int j = row[t];
            if (j <= 1 || j >= N - 2 || (dir && j == 2)) x[t] = czero<V>();
        }
    };

    V g0[M], g1[M], g2[M], g3[M];
    V u1[M], u2[M], u3[M], bb[M], pp[M];
    const V* pu2 = three ? A.u2 : nullptr;
    issue(A.u1, g0);
    issue(pu2, g1);
    issue(A.u3, g2);
    issue(A.p, g3);
    finish(A.u1, g0, u1);
    finish(pu2, g1, u2);
    finish(A.u3, g2, u3);
    finish(A.p, g3, pp);

    if (A.mode == 0) {
        V rn1[M], rn2[M], rn3[M], rnB[M];
        const V* pn2 = three ? A.n2 : nullptr;
        issue(A.b, g0);
        issue(A.n1, g1);
        issue(pn2, g2);
        issue(A.n3, g3);
        finish(A.b, g0, bb);
        finish(A.n1, g1, rn1);
        finish(pn2, g2, rn2);
        finish(A.n3, g3, rn3);
        issue(A.nB, g0);
        finish(A.nB, g0, rnB);
        issue(A.ro1, g0);  // first explicit term of the previous stage, in flight from here
        // BuildRHS linear terms, IMEXRK.cpp:140-147:
        //   r3 += Ri * ReinterpolateFull(b - <b>),  rB -= S * ReinterpolateDirichlet(u3)
        // (BuildRHSAdjoint has none: its explicit terms arrive complete)
        if (!A.skip_linear) {
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

        // ExplicitRK + FinishRHS + CrankNicolson explicit half, IMEXRK.cpp:75-134;
        // ro (the previous stage's explicit term) was issued into g earlier
        auto assemble = [&](V (&R)[M], const V (&u)[M], const V* ro, const V (&g)[M], const V (&rn)[M],
                            const V (&gp)[M], bool dir, T c) {
            V ro_[M];
            finish(ro, g, ro_);
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
        build_rows<T, M>(0, A, row, kk, kl, zero_mode, A.cv, coefz, ta, tb, tc);
        tri_factor<T, M>(F, ta, tb, tc, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) gp[t] = cik(kx, pp[t]);
        assemble(R, u1, A.ro1, g0, rn1, gp, false, A.cv);
        const V* pro2 = three ? A.ro2 : nullptr;
        issue(pro2, g1);
        issue(A.ro3, g2);
        tri_apply<T, M>(F, R, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) u1[t] = R[t];
        // EvolveBackground (IMEXRK.cpp:100-104): the (0,0) wave, which holds the
        // factor of that mode's velocity Neumann matrix (CNSolve1D, IMEXRK.h:584-588),
        // diffuses U_ in place: RU = U_ + cv Dzz U_, ZeroEnds of the 1D field
        // (Field.h:625-636), U_ = T^-1 RU.  No other wave of the launch reads U_.
        if (A.ubar != nullptr && zero_mode) {  // wave-uniform
            V ub[M], um[M], up[M];
#pragma unroll
            for (int t = 0; t < M; ++t) ub[t] = mkc<T>(coefz(ZC_UBAR, t), T(0));
            nb_prev<T, M>(ub, um, lane);
            nb_next<T, M>(ub, up, lane);
#pragma unroll
            for (int t = 0; t < M; ++t) {
                const int j = row[t];
                const V dzz = cadd(cadd(cscale(coefz(ZC_DZZN_LO, t), um[t]), cscale(coefz(ZC_DZZN_DI, t), ub[t])),
                                   cscale(coefz(ZC_DZZN_UP, t), up[t]));
                R[t] = (j <= 1 || j >= N - 2) ? czero<V>() : cadd(ub[t], cscale(A.cv, dzz));
            }
            tri_apply<T, M>(F, R, lane);
#pragma unroll
            for (int t = 0; t < M; ++t)
                if (row[t] < N) A.ubar[row[t]] = R[t].x;
        }
        if (three) {
#pragma unroll
            for (int t = 0; t < M; ++t) gp[t] = cik(ky, pp[t]);
            assemble(R, u2, pro2, g1, rn2, gp, false, A.cv);
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
        build_rows<T, M>(1, A, row, kk, kl, zero_mode, A.cv, coefz, ta, tb, tc);
        tri_factor<T, M>(F, ta, tb, tc, lane);
        assemble(R, u3, A.ro3, g2, rn3, gp, true, A.cv);
        issue(A.roB, g3);
        tri_apply<T, M>(F, R, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) u3[t] = R[t];
        // b: buoyancy Neumann matrix, no pressure term
#pragma unroll
        for (int t = 0; t < M; ++t) gp[t] = czero<V>();
        build_rows<T, M>(2, A, row, kk, kl, zero_mode, A.cb, coefz, ta, tb, tc);
        tri_factor<T, M>(F, ta, tb, tc, lane);
        assemble(R, bb, A.roB, g3, rnB, gp, false, A.cb);
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
            build_rows<T, M>(3, A, row, kk, kl, zero_mode, T(0), coefz, fa, fb, fc);
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
            build_rows<T, M>(3, A, row, kk, kl, zero_mode, T(0), coefz, ta, tb, tc);
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

// Workgroups go to the XCDs round-robin (block b to XCD b % 8), each XCD with
// its own L2.  Blocks whose tiles share cache lines (neighbouring j1 tiles)
// are renumbered so that they land on the same XCD at about the same time:
// XCD x takes the x-th contiguous eighth of the tile sequence.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    constexpr int NXCD = 8;
    if (nb % NXCD) return b;
    return (b % NXCD) * (nb / NXCD) + b / NXCD;
}

template <typename T, int M, bool FULL, int MB>
__global__ __launch_bounds__(MB * 64) void k_modal(ModalArgs<T> A) {
    __shared__ cx<T> tile[64 * M * (MB + 1)];
    const int nt = gridDim.x;  // j1 tiles per ja
    const int t = xcd_tile(blockIdx.x + nt * blockIdx.y, nt * gridDim.y);
    modal_body<T, M, MB, FULL>(A, t / nt, (t % nt) * MB, tile);
}

// index r of the run owning work item `item`: start[r] <= item < start[r+1]
__device__ __forceinline__ int find_run(const int* __restrict__ start, int nruns, int item) {
    int lo = 0, hi = nruns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Many independent runs in one launch: a grid-stride (persistent) loop over
// the work items of all runs, one item = one block's tile of modes.  Every
// run has the same rows-per-lane M; each reads its own arguments.
template <typename T, int M>
__global__ __launch_bounds__((modal_mb<T, M>() * 64)) void k_modal_batch(const ModalArgs<T>* __restrict__ args,
                                                                      const int* __restrict__ start, int nruns,
                                                                      int total) {
    constexpr int MB = modal_mb<T, M>();
    __shared__ cx<T> tile[64 * M * (MB + 1)];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int r = find_run(start, nruns, item);
        const ModalArgs<T>& a = args[r];
        const int nt = (a.A1 + MB - 1) / MB;
        const int local = item - start[r];
        modal_body<T, M, MB>(a, local / nt, (local % nt) * MB, tile);
    }
}

// All runs in one launch whatever their N3: each work item dispatches on its
// run's rows per lane M; the block has MB waves for every M and one LDS tile
// sized for the largest M.  Resources (registers, LDS) are those of the most
// demanding case compiled in.
template <typename T, int MB, int MMAX>
__global__ __launch_bounds__(MB * 64) void k_modal_mixed(const ModalArgs<T>* __restrict__ args,
                                                        const int* __restrict__ start, int nruns, int total) {
    __shared__ cx<T> tile[64 * MMAX * (MB + 1)];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int r = find_run(start, nruns, item);
        const ModalArgs<T>& a = args[r];
        const int nt = (a.A1 + MB - 1) / MB;
        const int local = item - start[r];
        const int ja = local / nt, j10 = (local % nt) * MB;
        switch ((a.N3 + 63) / 64) {
#define SF_M(m)                                                       \
    case m:                                                           \
        if constexpr (m <= MMAX) modal_body<T, m, MB>(a, ja, j10, tile); \
        break;
            SF_M(1) SF_M(2) SF_M(3) SF_M(4) SF_M(5) SF_M(6) SF_M(7) SF_M(8)
#undef SF_M
            default:
                break;
        }
    }
}

// the variant compiled for rows per lane up to 3, 6 or 8 that covers mmax
template <typename T, int MB>
static void launch_modal_mixed_mb(const ModalArgs<T>* args, const int* start, int nruns, int total, int mmax, int grid,
                                  hipStream_t s) {
    if (mmax <= 3)
        hipLaunchKernelGGL((k_modal_mixed<T, MB, 3>), dim3(grid), dim3(MB * 64), 0, s, args, start, nruns, total);
    else if (mmax <= 6)
        hipLaunchKernelGGL((k_modal_mixed<T, MB, 6>), dim3(grid), dim3(MB * 64), 0, s, args, start, nruns, total);
    else
        hipLaunchKernelGGL((k_modal_mixed<T, MB, 8>), dim3(grid), dim3(MB * 64), 0, s, args, start, nruns, total);
}

template <typename T>
hipError_t launch_modal_mixed(const ModalArgs<T>* args, const int* start, int nruns, int total, int mb, int mmax,
                              int grid, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    if (mmax < 1 || mmax > MODAL_MAX_M) return hipErrorInvalidValue;
    if (mb == 4) launch_modal_mixed_mb<T, 4>(args, start, nruns, total, mmax, grid, s);
    else if (mb == 8) launch_modal_mixed_mb<T, 8>(args, start, nruns, total, mmax, grid, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

#ifdef SF_KERNELS_HOST_DEFS
int modal_modes_per_block(int M, int precision) {
    switch (M) {
#define SF_M(m) \
    case m:     \
        return precision == 4 ? modal_mb<float, m>() : modal_mb<double, m>();
        SF_M(1) SF_M(2) SF_M(3) SF_M(4) SF_M(5) SF_M(6) SF_M(7) SF_M(8)
#undef SF_M
        default:
            return 4;
    }
}
#endif

template <typename T>
hipError_t launch_modal_batch(const ModalArgs<T>* args, const int* start, int nruns, int total, int M, int grid,
                              hipStream_t s) {
    if (total <= 0) return hipSuccess;
    switch (M) {
#define SF_M(m)                                                                                               \
    case m:                                                                                                   \
        hipLaunchKernelGGL((k_modal_batch<T, m>), dim3(grid), dim3(modal_mb<T, m>() * 64), 0, s, args, start, \
                           nruns, total);                                                                     \
        break;
        SF_M(1) SF_M(2) SF_M(3) SF_M(4) SF_M(5) SF_M(6) SF_M(7) SF_M(8)
#undef SF_M
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int M, int MB>
static void launch_modal_mb(const ModalArgs<T>& a, hipStream_t s) {
    const dim3 grid((a.A1 + MB - 1) / MB, a.A2);
    if (a.N3 == 64 * M) hipLaunchKernelGGL((k_modal<T, M, true, MB>), grid, dim3(MB * 64), 0, s, a);
    else hipLaunchKernelGGL((k_modal<T, M, false, MB>), grid, dim3(MB * 64), 0, s, a);
}

// Grids with at most this many modes (one per CU) take one mode per block:
// the GPU is far from full and latency sets the time, so each mode's wave
// runs alone on its CU (2D 256 x 1 x 256 FP32: 82 against 112 us per step
// with 8 modes per block); larger grids keep the wide tiles that make the
// loads coalesce.
inline int modal_small_grid() {
    static const int n = []() {
        int dev = 0, ncu = 256;
        if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        return ncu;
    }();
    return n;
}

template <typename T>
hipError_t launch_modal(const ModalArgs<T>& a, hipStream_t s) {
    const int M = (a.N3 + 63) / 64;
    switch (M) {
#define SF_M(m)                                                                                              \
    case m:                                                                                                  \
        if (a.A1 * a.A2 <= modal_small_grid())                                                               \
            launch_modal_mb<T, m, 1>(a, s);                                                                  \
        else                                                                                                 \
            launch_modal_mb<T, m, modal_mb<T, m>()>(a, s);                                                   \
        break;
        SF_M(1) SF_M(2) SF_M(3) SF_M(4) SF_M(5) SF_M(6) SF_M(7) SF_M(8)
#undef SF_M
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

#ifdef SF_KERNELS_HOST_DEFS
int max_rows_per_lane() { return MODAL_MAX_M; }
#endif

// ---------------------------------------------------------------------------
// k_xpass: inverse x FFT, products, forward x FFT
//
// A block owns one spanwise plane n2 and A.zch consecutive z levels, which it
// visits one level at a time.  A ring of three nodal levels (z-1, z, z+1, the
// +-1 halo of the vertical stencils) stays in LDS.  Each step computes all
// products of level z from the ring, then runs ONE batched FFT over the
// product lines of level z and the spectra of level z+2 (which replaces z-1 in
// the ring): every level is inverse-transformed once per block and no
// separate halo pass is needed.  The inverse transform uses
// ifft(x) = conj(fft(conj(x))), so both directions share the forward kernel.
// ---------------------------------------------------------------------------
#define SF_SIZES(X) X(6) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

#ifdef SF_KERNELS_HOST_DEFS
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
#endif

// line pitch of the x-pass: the smallest that holds the FFT's padded pass-1
// table, rounded to 8 mod 32 complex values (keeps four 3D blocks per CU)
template <int N, int R1F = 0>
constexpr int x_ldp() {
    int l = FftPlan<N, R1F>::two ? N + FftPlan<N, R1F>::R2 : N + 1;
    while (l % 32 != 8) ++l;
    return l;
}

// waves per SIMD the x-pass is compiled for: four in FP32 3D without the
// linearised products (register budget 128), else what the registers allow
template <typename T, bool THREED, bool LIN>
constexpr int xpass_waves() {
    return (std::is_same<T, float>::value && THREED && !LIN) ? 4 : 1;
}

// threads per x-pass block (a multiple of 64)
template <typename T, bool THREED, bool LIN>
constexpr int xpass_bs() {
    return BS;
}

// LDS values of one x-pass block: NP product lines and a ring of 3 x NI
// input lines at the line pitch x_ldp
template <int N1, bool THREED, bool LIN, int R1F = 0>
constexpr int xpass_lds() {
    return ((THREED ? 6 : 3) + 3 * (LIN ? 4 : 2)) * x_ldp<N1, R1F>();
}

// R1F: the FFT plan's first radix (0: the default; A.tw must be the table of
// the same plan)
template <typename T, int N1, bool THREED, bool LIN, int R1F = 0>
__device__ __forceinline__ void xpass_body(const XArgs<T>& A, const int bx, const int n2, cx<T>* const lds) {
    using V = cx<T>;
    constexpr int LDP = x_ldp<N1, R1F>();
    constexpr int NI = LIN ? 4 : 2;        // input pair lines per level
    constexpr int NP = THREED ? 6 : 3;     // product pair lines per level
    constexpr int NB = NP + NI;            // lines per batched FFT
    constexpr int AMAX = N1 / 3;           // A1 <= N1/3
    constexpr int XB = xpass_bs<T, THREED, LIN>();
    constexpr int FI = (NI * AMAX + XB - 1) / XB;
    // lds: [NP product lines][ring slot 0: NI lines][slot 1][slot 2]
    V* const prd = lds;
    V* const ring = lds + NP * LDP;

    const int N3 = A.N3, A1 = A.A1;
    const int zbeg = bx * A.zch;
    const int zend = min(zbeg + A.zch, N3);
    // Rounds that fill only part of the block (the products of 384 points,
    // the FFT passes, the unpacking of A1 modes) fall to the first waves of
    // the work-item order; rotating that order by whole waves from block to
    // block spreads the extra rounds over the SIMDs of a CU.
    const int tid = (int(threadIdx.x) + 64 * ((bx + n2) % (XB / 64))) % XB;
    auto slot = [&](int z) { return (z + 1 - zbeg) % 3; };

    // spectra of one level into registers: item = pair * A1 + j1
    V X[FI], Y[FI];
    auto load = [&](int z) {
#pragma unroll
        for (int i = 0; i < FI; ++i) {
            const int item = tid + i * XB;
            X[i] = czero<V>();
            Y[i] = czero<V>();
            if (item < NI * A1 && z >= 0 && z < N3) {
                const int pair = item / A1;
                const int j1 = item - pair * A1;
                const V* const* in = (pair >= 2) ? A.base : A.in;
                const size_t off = sidx(n2, z, j1, N3, A1);
                if ((pair & 1) == 0) {
                    X[i] = in[0][off];
                    Y[i] = in[2][off];
                } else if (THREED) {
                    X[i] = in[1][off];
                    Y[i] = in[3][off];
                } else {
                    X[i] = in[3][off];
                }
            }
        }
    };
    // conj of the Hermitian-extended pair line Z = F(a) + i F(b) into ring slot s
    auto store = [&](int s) {
        V* R = ring + s * NI * LDP;
        const int nband = N1 - 2 * A1 + 1;  // j = A1 .. N1-A1 is zero
        for (int idx = tid; idx < NI * nband; idx += XB) {
            const int pair = idx / nband;
            R[pair * LDP + A1 + (idx - pair * nband)] = czero<V>();
        }
#pragma unroll
        for (int i = 0; i < FI; ++i) {
            const int item = tid + i * XB;
            if (item < NI * A1) {
                const int pair = item / A1;
                const int j1 = item - pair * A1;
                V* L = R + pair * LDP;
                const V x = X[i], y = Y[i];
                if (j1 == 0) {  // c2r ignores the imaginary part of the DC term
                    L[0] = mkc<T>(x.x, -y.x);
                } else {
                    L[j1] = mkc<T>(x.x - y.y, -(x.y + y.x));
                    L[N1 - j1] = mkc<T>(x.x + y.y, x.y - y.x);
                }
            }
        }
    };

    // prologue: levels zbeg-1, zbeg, zbeg+1 into slots 0, 1, 2
    for (int l = 0; l < 3; ++l) {
        load(zbeg - 1 + l);
        store(l);
    }
    // the spectra of level z+2 are loaded one whole level ahead: issued after
    // the previous level's store, consumed by this level's store
    if (zbeg + 1 < zend) load(zbeg + 2);
    sync_lds();
    lds_fft_lines<T, N1, 3 * NI, XB, -1, 1, false, R1F, true>(3 * NI, [ring](int line) { return ring + line * LDP; }, A.tw,
                                                        tid);

    for (int z = zbeg; z < zend; ++z) {
        const bool more = z + 1 < zend;

        // ---- all products of level z --------------------------------------
        {
            const T itl = A.zc[ZC_ITIL_LO * A.N3p + z], itd = A.zc[ZC_ITIL_DI * A.N3p + z];
            const T dDN = A.zc[ZC_DZDN * A.N3p + z], dND = A.zc[ZC_DZND * A.N3p + z];
            const T hc = (z >= 1) ? T(0.5) : T(0);  // ReinterpolateBar row z (row 0 is empty)
            const T h = T(0.5);
            T ub[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int zz = z - 1 + d;
                ub[d] = (zz >= 0 && zz < N3) ? A.zc[ZC_UBAR * A.N3p + zz] : T(0);
            }
            const V* Ls[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) Ls[d] = ring + slot(z - 1 + d) * NI * LDP;
            for (int n = tid; n < N1; n += XB) {
                T u1[3], u2[3], u3[3], bb[3], w1[3], w2[3], w3[3], wb[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const V a = Ls[d][n];           // conj(U1 + i U3)
                    const V c = Ls[d][LDP + n];     // conj(U2 + i B) or conj(B)
                    u1[d] = a.x + (LIN ? T(0) : ub[d]);
                    u3[d] = -a.y;
                    u2[d] = THREED ? c.x : T(0);
                    bb[d] = THREED ? -c.y : c.x;
                    if (LIN) {
                        const V e = Ls[d][2 * LDP + n];
                        const V f = Ls[d][3 * LDP + n];
                        w1[d] = e.x + ub[d];
                        w3[d] = -e.y;
                        w2[d] = THREED ? f.x : T(0);
                        wb[d] = THREED ? -f.y : f.x;
                    } else {
                        w1[d] = w2[d] = w3[d] = wb[d] = T(0);
                    }
                }
                T P[12];
                if (!LIN) {
                    // BuildRHS products (IMEXRK.cpp:136-193), U1t = U1 + U_
                    P[0] = u1[1] * u1[1];                                                   // P1
                    P[1] = ((h * u1[1] + h * u1[2]) * u3[2] - (hc * u1[0] + hc * u1[1]) * u3[1]) * dDN;  // P2
                    P[2] = (itl * u1[0] + itd * u1[1]) * u3[1];                            // P3
                    {
                        const T ic = h * u3[1] + h * u3[2], im = h * u3[0] + h * u3[1];
                        P[3] = (ic * ic - im * im) * dND;                                   // Q4
                    }
                    P[4] = ((h * bb[1] + h * bb[2]) * u3[2] - (hc * bb[0] + hc * bb[1]) * u3[1]) * dDN;  // P5
                    P[5] = u1[1] * bb[1];                                                   // P6
                    if (THREED) {
                        P[6] = ((h * u2[1] + h * u2[2]) * u3[2] - (hc * u2[0] + hc * u2[1]) * u3[1]) * dDN;  // P8
                        P[7] = u1[1] * u2[1];                                               // P10
                        P[8] = u2[1] * u2[1];                                               // P7
                        P[9] = (itl * u2[0] + itd * u2[1]) * u3[1];                        // P9
                        P[10] = u2[1] * bb[1];                                              // P11
                        P[11] = T(0);
                    }
                } else {
                    // BuildRHSLinear products (IMEXRK.cpp:195-258) about the base w
                    // (w1 includes U_); u1 is the perturbation without U_
                    P[0] = T(2) * (u1[1] * w1[1]);
                    {
                        const T gc = (hc * u1[0] + hc * u1[1]) * w3[1] + (hc * w1[0] + hc * w1[1]) * u3[1];
                        const T gp = (h * u1[1] + h * u1[2]) * w3[2] + (h * w1[1] + h * w1[2]) * u3[2];
                        P[1] = (gp - gc) * dDN;
                    }
                    P[2] = (itl * u1[0] + itd * u1[1]) * w3[1] + (itl * w1[0] + itd * w1[1]) * u3[1];
                    {
                        const T ic = (h * u3[1] + h * u3[2]) * (h * w3[1] + h * w3[2]);
                        const T im = (h * u3[0] + h * u3[1]) * (h * w3[0] + h * w3[1]);
                        P[3] = T(2) * ((ic - im) * dND);
                    }
                    {
                        const T gc = (hc * bb[0] + hc * bb[1]) * w3[1] + (hc * wb[0] + hc * wb[1]) * u3[1];
                        const T gp = (h * bb[1] + h * bb[2]) * w3[2] + (h * wb[1] + h * wb[2]) * u3[2];
                        P[4] = (gp - gc) * dDN;
                    }
                    P[5] = bb[1] * w1[1] + wb[1] * u1[1];
                    if (THREED) {
                        const T gc = (hc * u2[0] + hc * u2[1]) * w3[1] + (hc * w2[0] + hc * w2[1]) * u3[1];
                        const T gp = (h * u2[1] + h * u2[2]) * w3[2] + (h * w2[1] + h * w2[2]) * u3[2];
                        P[6] = (gp - gc) * dDN;
                        P[7] = u1[1] * w2[1] + w1[1] * u2[1];
                        P[8] = T(2) * (u2[1] * w2[1]);
                        P[9] = (itl * u2[0] + itd * u2[1]) * w3[1] + (itl * w2[0] + itd * w2[1]) * u3[1];
                        P[10] = bb[1] * w2[1] + wb[1] * u2[1];
                        P[11] = T(0);
                    }
                }
#pragma unroll
                for (int q = 0; q < NP; ++q) prd[q * LDP + n] = mkc<T>(P[2 * q], P[2 * q + 1]);
            }
        }
        sync_lds();
        if (more) {
            store(slot(z + 2));  // replaces level z-1
            if (z + 2 < zend) load(z + 3);
        }
        sync_lds();
        {
            const int s2 = slot(z + 2) * NI;
            lds_fft_lines<T, N1, NB, XB, -1, 1, false, R1F, true>(more ? NB : NP,
                                             [prd, s2](int line) {
                                                 return prd + (line + (line >= NP ? s2 : 0)) * LDP;
                                             },
                                             A.tw, tid);
        }
        // ---- separate the two real spectra of each product line, combine ---
        for (int k = tid; k < A1; k += XB) {
            V Fa[NP], Fb[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const V Z = prd[q * LDP + k];
                const V W = prd[q * LDP + ((N1 - k) % N1)];
                // Fa = (Z + conj W)/2, Fb = (Z - conj W)/(2i)
                Fa[q].x = T(0.5) * (Z.x + W.x);
                Fa[q].y = T(0.5) * (Z.y - W.y);
                Fb[q].x = T(0.5) * (Z.y + W.y);
                Fb[q].y = T(-0.5) * (Z.x - W.x);
            }
            const T kxv = A.kx[k];
            const T ms = -A.scale;
            const size_t off = (THREED && A.tiled) ? ytile_idx(n2, z, k, A.N2, A1) : sidx(n2, z, k, N3, A1);
            // r1 -= ikx P1 + P2 ; r3 -= ikx P3 + ddz P4 ; rB -= P5 + ikx P6
            A.out[0][off] = cscale(ms, cadd(cik(kxv, Fa[0]), Fb[0]));
            A.out[2][off] = cscale(ms, cadd(cik(kxv, Fa[1]), Fb[1]));
            A.out[3][off] = cscale(ms, cadd(Fa[2], cik(kxv, Fb[2])));
            if constexpr (THREED) {
                // r2 -= P8 + ikx P10 ; r1 -= iky P10 ; r2 -= iky P7 ; r3 -= iky P9 ; rB -= iky P11
                A.out[1][off] = cscale(ms, cadd(Fa[3], cik(kxv, Fb[3])));
                A.out[4][off] = cscale(ms, Fb[3]);
                A.out[5][off] = cscale(ms, Fa[4]);
                A.out[6][off] = cscale(ms, Fb[4]);
                A.out[7][off] = cscale(ms, Fa[5]);
            }
        }
        sync_lds();
    }
}

template <typename T, int N1, bool THREED, bool LIN>
__global__ __launch_bounds__((xpass_bs<T, THREED, LIN>())) __attribute__((amdgpu_waves_per_eu(xpass_waves<T, THREED, LIN>()))) void k_xpass(
    XArgs<T> A) {
    __shared__ cx<T> lds[xpass_lds<N1, THREED, LIN>()];
    xpass_body<T, N1, THREED, LIN>(A, blockIdx.x, blockIdx.y, lds);
}

// 2D runs with the same N1 in one launch (grid-stride over the z-blocks of
// all runs; see k_modal_batch)
template <typename T, int N1>
__global__ __launch_bounds__(BS) void k_xpass_batch(const XArgs<T>* __restrict__ args, const int* __restrict__ start,
                                                    int nruns, int total) {
    static_assert(xpass_bs<T, false, false>() == BS, "the batched 2D x-pass runs blocks of BS threads");
    __shared__ cx<T> lds[xpass_lds<N1, false, false>()];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int r = find_run(start, nruns, item);
        xpass_body<T, N1, false, false>(args[r], item - start[r], 0, lds);
    }
}

// 2D runs of every N1 up to NMAX in one launch: each work item dispatches on
// its run's N1 and all cases share one LDS allocation sized for the largest.
// R1F picks the FFT plan of every N1 that has it (fft_plan_ok), the default
// plan of the others; each run's tw is the table of its plan.
template <int NMAX, int R1F>
constexpr int xpass_mixed_lds() {
    int m = 0;
#define SF_CASE(n) \
    if (n <= NMAX && xpass_lds<n, false, false, R1F>() > m) m = xpass_lds<n, false, false, R1F>();
    SF_SIZES(SF_CASE)
#undef SF_CASE
    return m;
}

template <typename T, int NMAX, int R1F>
__global__ __launch_bounds__(BS) void k_xpass_mixed(const XArgs<T>* __restrict__ args, const int* __restrict__ start,
                                                    int nruns, int total) {
    __shared__ cx<T> lds[xpass_mixed_lds<NMAX, R1F>()];
    for (int item = blockIdx.x; item < total; item += gridDim.x) {
        const int r = find_run(start, nruns, item);
        const XArgs<T>& a = args[r];
        switch (a.N1) {
#define SF_CASE(n)                                                                              \
    case n:                                                                                     \
        if constexpr (n <= NMAX) xpass_body<T, n, false, false, R1F>(a, item - start[r], 0, lds); \
        break;
            SF_SIZES(SF_CASE)
#undef SF_CASE
            default:
                break;
        }
    }
}

template <typename T, int R1F>
static void launch_xpass_mixed_r(const XArgs<T>* args, const int* start, int nruns, int total, int nmax, int grid,
                                 hipStream_t s) {
    if (nmax <= 128) hipLaunchKernelGGL((k_xpass_mixed<T, 128, R1F>), dim3(grid), dim3(BS), 0, s, args, start, nruns, total);
    else if (nmax <= 256) hipLaunchKernelGGL((k_xpass_mixed<T, 256, R1F>), dim3(grid), dim3(BS), 0, s, args, start, nruns, total);
    else hipLaunchKernelGGL((k_xpass_mixed<T, 512, R1F>), dim3(grid), dim3(BS), 0, s, args, start, nruns, total);
}

// the variant compiled for N1 up to 128, 256 or 512 that covers nmax, with
// the default FFT plans (r1f = 0) or first radix r1f = 8 where possible
template <typename T>
hipError_t launch_xpass_mixed(const XArgs<T>* args, const int* start, int nruns, int total, int nmax, int r1f,
                              int grid, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    if (nmax > 512) return hipErrorInvalidValue;
    if (r1f == 0) launch_xpass_mixed_r<T, 0>(args, start, nruns, total, nmax, grid, s);
    else if (r1f == 8) launch_xpass_mixed_r<T, 8>(args, start, nruns, total, nmax, grid, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

template <typename T>
hipError_t launch_xpass_batch(const XArgs<T>* args, const int* start, int nruns, int total, int N1, int grid,
                              hipStream_t s) {
    if (total <= 0) return hipSuccess;
    switch (N1) {
#define SF_CASE(n)                                                                                        \
    case n:                                                                                               \
        hipLaunchKernelGGL((k_xpass_batch<T, n>), dim3(grid), dim3(BS), 0, s, args, start, nruns, total); \
        return hipGetLastError();
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return hipErrorInvalidValue;
    }
}

template <typename T, int N1, bool THREED, bool LIN>
static hipError_t launch_xpass_v(const XArgs<T>& a, int N2, hipStream_t s) {
    constexpr int lds = xpass_lds<N1, THREED, LIN>() * int(sizeof(cx<T>));
    if constexpr (lds > 160 * 1024) {
        return hipErrorInvalidValue;  // not instantiated: does not fit the LDS of one CU
    } else {
        if (a.zch < 1 || a.A1 > N1 / 3) return hipErrorInvalidValue;
        const int nblk = (a.N3 + a.zch - 1) / a.zch;
        dim3 grid(nblk, THREED ? N2 : 1);
        hipLaunchKernelGGL((k_xpass<T, N1, THREED, LIN>), grid, dim3(xpass_bs<T, THREED, LIN>()), 0, s, a);
        return hipGetLastError();
    }
}

template <typename T>
hipError_t launch_xpass(const XArgs<T>& a, int N1, int N2, bool threeD, hipStream_t s) {
    switch (N1) {
#define SF_CASE(n)                                                                                   \
    case n:                                                                                          \
        if (a.linear)                                                                                \
            return threeD ? launch_xpass_v<T, n, true, true>(a, N2, s)                               \
                          : launch_xpass_v<T, n, false, true>(a, N2, s);                             \
        return threeD ? launch_xpass_v<T, n, true, false>(a, N2, s) : launch_xpass_v<T, n, false, false>(a, N2, s);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return hipErrorInvalidValue;
    }
}

// Levels per x-pass block.  Each block pays two extra ring levels, and the
// last round of blocks may leave CUs idle; pick the height in [1, 48] that
// minimises (rounds of resident blocks) x (levels per block + 2), given how
// many blocks of this kernel the GPU keeps resident (small grids take one
// level per block: more blocks in flight for the same single round).
template <typename T, int N1, bool THREED, bool LIN>
static int xpass_resident(int ncu) {
    constexpr int lds = xpass_lds<N1, THREED, LIN>() * int(sizeof(cx<T>));
    if constexpr (lds > 160 * 1024) {
        return 0;
    } else {
        int per = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_xpass<T, N1, THREED, LIN>, xpass_bs<T, THREED, LIN>(), 0) != hipSuccess)
            per = 1;
        return (per < 1 ? 1 : per) * ncu;
    }
}

// resident x-pass blocks on ncu CUs (each precision's translation unit
// instantiates its own)
template <typename T>
int xpass_slots(int N1, bool threeD, int ncu) {
    switch (N1) {
#define SF_CASE(n) \
    case n:        \
        return threeD ? xpass_resident<T, n, true, false>(ncu) : xpass_resident<T, n, false, false>(ncu);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return ncu;
    }
}

#ifdef SF_KERNELS_HOST_DEFS
extern template int xpass_slots<double>(int, bool, int);  // in the FP64 translation unit

int xpass_levels(int N1, int N2, int N3, bool threeD, int precision) {
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    int slots = precision == 4 ? xpass_slots<float>(N1, threeD, ncu) : xpass_slots<double>(N1, threeD, ncu);
    if (slots < 1) slots = ncu;
    const int rows = threeD ? N2 : 1;
    long best = -1;
    int pick = 2;
    for (int zc = 1; zc <= 48; ++zc) {
        const long nblk = long((N3 + zc - 1) / zc) * rows;
        const long rounds = (nblk + slots - 1) / slots;
        const long cost = rounds * (zc + 2);
        if (best < 0 || cost < best) {
            best = cost;
            pick = zc;
        }
    }
    return pick;
}
#endif

#ifdef SF_KERNELS_HOST_DEFS
void fft_twiddles_plan(int n, int r1f, double* re, double* im) {
    switch (n) {
#define SF_CASE(n)                              \
    case n:                                     \
        if (r1f == 8) fft_twiddle_table<n, 8>(re, im);   \
        else if (r1f == 16) fft_twiddle_table<n, 16>(re, im); \
        else fft_twiddle_table<n>(re, im);      \
        return;
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return;
    }
}

int fft_first_radix(int n, int r1f) {
    switch (n) {
#define SF_CASE(n)                                                                       \
    case n:                                                                              \
        return r1f == 8 ? FftPlan<n, 8>::R1 : (r1f == 16 ? FftPlan<n, 16>::R1 : FftPlan<n>::R1);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return 0;
    }
}

void fft_twiddles(int n, double* re, double* im) {
    switch (n) {
#define SF_CASE(n)                           \
    case n:                                  \
        fft_twiddle_table<n>(re, im);        \
        return;
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return;
    }
}
#endif

// ---------------------------------------------------------------------------
// y transforms (3D), one block per (z, tile of YJ consecutive j1)
//
// The j1 tile makes both sides of the transpose between the modal layout
// ([ja][z][j1]) and the x-spectral layout ([n2][z][j1]) read and write runs of
// YJ consecutive values.  Lines (one per j1 of the tile and field) sit in LDS
// at a pitch of 1 mod 32 complex values, so that the YJ lanes that fill or
// drain the same position of YJ lines hit distinct banks.
// ---------------------------------------------------------------------------
// line pitch: the residues mod 32 complex values (12 for tiles of 8 j1, 14
// for 16) gave the fewest LDS bank conflicts over the transpose fill, both FFT
// passes and the drain at N2 = 192 (3% faster forward pass than a pitch of
// 64/YJ mod 16)
template <int N, int YJ>
constexpr int y_ldp() {
    int l = FftPlan<N>::two ? N + FftPlan<N>::R2 : N;
    while (l % 32 != (YJ == 8 ? 12 : 14)) ++l;
    return l;
}

// inverse: out_f(n2, z, j1) = sum_{j2 active} in_f(ja, z, j1) e^{+2 pi i j2 n2/N2}
// The loads of the next field are issued before the FFT of the current one.
template <typename T, int N2, int YJ, int FR>
__global__ __launch_bounds__(BS) void k_yinv(YArgs<T> A) {
    using V = cx<T>;
    constexpr int LDP = y_ldp<N2, YJ>();
    constexpr int NL = YJ * FR;
    constexpr int AMAX = 2 * (N2 / 3);  // bound on A2
    constexpr int PG = (FR * AMAX * YJ + BS - 1) / BS;
    __shared__ V buf[NL * LDP];
    // dispatch index -> (z, j1 tile) with neighbouring j1 tiles on one XCD
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int z = tile / gridDim.y;
    const int j10 = (tile % gridDim.y) * YJ;
    const int N3 = A.N3, A1 = A.A1, A2 = A.A2;
    const int h = N2 / 3;
    auto lp = [&](int line) { return buf + line * LDP; };
    V g[PG];
    auto issue = [&](int f0) {
#pragma unroll
        for (int i = 0; i < PG; ++i) {
            const int idx = int(threadIdx.x) + i * BS;
            const int jj = idx % YJ;
            const int rest = idx / YJ;
            const int ja = rest % A2;
            const int fl = rest / A2;
            const int f = f0 + fl;
            const int j1 = j10 + jj;
            g[i] = (idx < FR * A2 * YJ && f < A.nf && j1 < A1) ? A.in[f][sidx(ja, z, j1, N3, A1)] : czero<V>();
        }
    };
    issue(0);
    for (int f0 = 0; f0 < A.nf; f0 += FR) {
        // zero band j2 = h .. N2-h of every line, then place the active modes
        const int nband = N2 - 2 * h + 1;
        for (int idx = threadIdx.x; idx < NL * nband; idx += BS) {
            const int line = idx / nband;
            buf[line * LDP + h + (idx - line * nband)] = czero<V>();
        }
#pragma unroll
        for (int i = 0; i < PG; ++i) {
            const int idx = int(threadIdx.x) + i * BS;
            if (idx < FR * A2 * YJ) {
                const int jj = idx % YJ;
                const int rest = idx / YJ;
                const int ja = rest % A2;
                const int fl = rest / A2;
                const int j2 = (ja < h) ? ja : ja + N2 - 2 * h + 1;
                buf[(fl * YJ + jj) * LDP + j2] = g[i];
            }
        }
        sync_lds();
        if (f0 + FR < A.nf) issue(f0 + FR);
        lds_fft_lines<T, N2, NL, BS, +1>(NL, lp, A.tw);
        for (int idx = threadIdx.x; idx < FR * N2 * YJ; idx += BS) {
            const int jj = idx % YJ;
            const int rest = idx / YJ;
            const int n2 = rest % N2;
            const int fl = rest / N2;
            const int f = f0 + fl;
            const int j1 = j10 + jj;
            if (f < A.nf && j1 < A1) A.out[f][sidx(n2, z, j1, N3, A1)] = buf[(fl * YJ + jj) * LDP + n2];
        }
        sync_lds();
    }
}

// forward: X_f(n2, z, j1) -> active j2.  paired: out[r] = scale*(F(in[2r]) + i ky F(in[2r+1]))
template <typename T, int N2, int YJ, int FR>
__global__ __launch_bounds__(BS) void k_yfwd(YArgs<T> A) {
    using V = cx<T>;
    constexpr int LDP = y_ldp<N2, YJ>();
    constexpr int NL = YJ * FR;
    constexpr int PG = (FR * N2 * YJ + BS - 1) / BS;
    __shared__ V buf[NL * LDP];
    // dispatch index -> (z, j1 tile) with neighbouring j1 tiles on one XCD
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int z = tile / gridDim.y;
    const int j10 = (tile % gridDim.y) * YJ;
    const int N3 = A.N3, A1 = A.A1, A2 = A.A2;
    const int h = N2 / 3;
    const int nin = A.paired ? 2 * A.nf : A.nf;
    auto lp = [&](int line) { return buf + line * LDP; };
    V g[PG];
    auto issue = [&](int f0) {
#pragma unroll
        for (int i = 0; i < PG; ++i) {
            const int idx = int(threadIdx.x) + i * BS;
            const int jj = idx % YJ;
            const int rest = idx / YJ;
            const int n2 = rest % N2;
            const int fl = rest / N2;
            const int f = f0 + fl;
            const int j1 = j10 + jj;
            g[i] = (idx < FR * N2 * YJ && f < nin && j1 < A1) ? A.in[f][sidx(n2, z, j1, N3, A1)] : czero<V>();
        }
    };
    issue(0);
    for (int f0 = 0; f0 < nin; f0 += FR) {
#pragma unroll
        for (int i = 0; i < PG; ++i) {
            const int idx = int(threadIdx.x) + i * BS;
            if (idx < FR * N2 * YJ) {
                const int jj = idx % YJ;
                const int rest = idx / YJ;
                const int n2 = rest % N2;
                const int fl = rest / N2;
                buf[(fl * YJ + jj) * LDP + n2] = g[i];
            }
        }
        sync_lds();
        if (f0 + FR < nin) issue(f0 + FR);
        lds_fft_lines<T, N2, NL, BS, -1>(NL, lp, A.tw);
        if (A.paired) {
            for (int idx = threadIdx.x; idx < (FR / 2) * A2 * YJ; idx += BS) {
                const int jj = idx % YJ;
                const int rest = idx / YJ;
                const int ja = rest % A2;
                const int pr = rest / A2;
                const int r = f0 / 2 + pr;
                const int j1 = j10 + jj;
                if (r >= A.nf || j1 >= A1) continue;
                const int j2 = (ja < h) ? ja : ja + N2 - 2 * h + 1;
                const V Fa = buf[((2 * pr) * YJ + jj) * LDP + j2];
                const V Fb = buf[((2 * pr + 1) * YJ + jj) * LDP + j2];
                A.out[r][sidx(ja, z, j1, N3, A1)] = cscale(A.scale, cadd(Fa, cik(A.ky[ja], Fb)));
            }
        } else {
            for (int idx = threadIdx.x; idx < FR * A2 * YJ; idx += BS) {
                const int jj = idx % YJ;
                const int rest = idx / YJ;
                const int ja = rest % A2;
                const int fl = rest / A2;
                const int f = f0 + fl;
                const int j1 = j10 + jj;
                if (f >= nin || j1 >= A1) continue;
                const int j2 = (ja < h) ? ja : ja + N2 - 2 * h + 1;
                A.out[f][sidx(ja, z, j1, N3, A1)] = cscale(A.scale, buf[(fl * YJ + jj) * LDP + j2]);
            }
        }
        sync_lds();
    }
}

// y transforms with the rows loaded straight into LDS (buffer_load ... lds,
// no VGPR staging), one block per (z, tile of YJ j1).  A field arrives as N2
// rows of YJ consecutive j1, row pitch YJ: the YJ lines are interleaved
// (element n2 of line jj at n2*YJ + jj) and the FFT runs over them
// line-fastest, so lanes touch consecutive LDS words.  Global reads and writes
// move runs of YJ values.  One field region of LDS per block.
template <typename T, int N2, int YJ>
struct YDma {
    using V = cx<T>;
    static constexpr int R2 = FftPlan<N2>::two ? FftPlan<N2>::R2 : 1;
    static constexpr int ROWS = N2 + R2;                   // positions per line, pass-1 padding included
    static constexpr int FIELD = ROWS * YJ;                // values per field region
    static constexpr int VPL = 16 / int(sizeof(V));        // values per lane per 16-byte load
    static constexpr int LPR = YJ / VPL;                   // lanes per row
    static constexpr int RPI = 64 / LPR;                   // rows per wave load
    static constexpr size_t LDS = size_t(FIELD) * sizeof(V);
    static bool fits(int A1, int nin) {
        return N2 % RPI == 0 && (A1 * int(sizeof(V))) % 16 == 0 && nin % 2 == 0 && LDS <= 160 * 1024;
    }
};

// Paired forward y transform: F(A) is kept in registers at the active rows
// while B is loaded into the same LDS.  The inputs are in the sidx layout, or
// in tiles of 16 modes (A.tiled, ytile_idx) where a block's tile of each field
// is one contiguous run.
template <typename T, int N2, int YJ>
__global__ __launch_bounds__(BS) void k_yfwd_seq(YArgs<T> A) {
    using V = cx<T>;
    using D = YDma<T, N2, YJ>;
    constexpr int PI = (2 * (N2 / 3) * YJ + BS - 1) / BS;  // active rows x YJ per thread
    __shared__ __align__(16) V buf[D::FIELD];
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int z = tile / gridDim.y;
    const int j10 = (tile % gridDim.y) * YJ;
    const int N3 = A.N3, A1 = A.A1, A2 = A.A2;
    const int h = N2 / 3;
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    const size_t fbytes = size_t(N2) * size_t(N3) * size_t(A1) * sizeof(V);
    const int col = (lane % D::LPR) * D::VPL;
    const int lrow = lane / D::LPR;
    const bool jok = j10 + col < A1;
    for (int f0 = 0; f0 < 2 * A.nf; f0 += 2) {
        V Fa[PI];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const rsrc_t r = make_rsrc(A.in[f0 + q], fbytes);
            for (int i = wave; i < N2 / D::RPI; i += BS / 64) {
                const int row = i * D::RPI + lrow;
                const size_t e = A.tiled ? ytile_idx(row, z, j10 + col, N2, A1) : sidx(row, z, j10 + col, N3, A1);
                const int off = jok ? int(e * sizeof(V)) : BUF_OOB;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    r, (__attribute__((address_space(3))) void*)(buf + i * D::RPI * YJ), 16, off, 0, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            lds_fft_lines<T, N2, YJ, BS, -1, YJ, true>(YJ, [&](int line) { return buf + line; }, A.tw);
            if (q == 0) {
#pragma unroll
                for (int i = 0; i < PI; ++i) {
                    const int idx = int(threadIdx.x) + i * BS;
                    const int ja = idx / YJ;
                    const int j2 = (ja < h) ? ja : ja + N2 - 2 * h + 1;
                    Fa[i] = (idx < A2 * YJ) ? buf[j2 * YJ + idx % YJ] : czero<V>();
                }
            } else {
                const rsrc_t o = make_rsrc(A.out[f0 / 2], size_t(A2) * size_t(N3) * size_t(A1) * sizeof(V));
#pragma unroll
                for (int i = 0; i < PI; ++i) {
                    const int idx = int(threadIdx.x) + i * BS;
                    if (idx < A2 * YJ) {
                        const int jj = idx % YJ;
                        const int ja = idx / YJ;
                        const int j2 = (ja < h) ? ja : ja + N2 - 2 * h + 1;
                        const V Fb = buf[j2 * YJ + jj];
                        const int off = (j10 + jj < A1) ? int(sidx(ja, z, j10 + jj, N3, A1) * sizeof(V)) : BUF_OOB;
                        bst<V>(cscale(A.scale, cadd(Fa[i], cik(A.ky[ja], Fb))), o, off);
                    }
                }
            }
            __syncthreads();  // every wave is done with buf before the next loads
        }
    }
}

// Inverse y transform with the modal rows loaded straight into LDS, one block
// per (z, tile of YJ j1), the nf fields one after the other.  Row ja of a
// field lands at its line position j2(ja) (interleaved lines as in
// k_yfwd_seq); the dealiased band j2 = h .. N2-h is cleared in LDS.  A wave
// load covers RPI consecutive rows, so h must be a multiple of RPI for no load
// to straddle the band.
template <typename T, int N2, int YJ>
__global__ __launch_bounds__(BS) void k_yinv_dma(YArgs<T> A) {
    using V = cx<T>;
    using D = YDma<T, N2, YJ>;
    __shared__ __align__(16) V buf[D::FIELD];
    const int tile = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int z = tile / gridDim.y;
    const int j10 = (tile % gridDim.y) * YJ;
    const int N3 = A.N3, A1 = A.A1, A2 = A.A2;
    const int h = N2 / 3;
    const int nband = N2 - 2 * h + 1;
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    const int col = (lane % D::LPR) * D::VPL;
    const int lrow = lane / D::LPR;
    const bool jok = j10 + col < A1;
    const size_t mbytes = size_t(A2) * size_t(N3) * size_t(A1) * sizeof(V);
    const size_t ybytes = size_t(N2) * size_t(N3) * size_t(A1) * sizeof(V);
    for (int f = 0; f < A.nf; ++f) {
        for (int idx = threadIdx.x; idx < nband * YJ; idx += BS) buf[(h + idx / YJ) * YJ + idx % YJ] = czero<V>();
        const rsrc_t r = make_rsrc(A.in[f], mbytes);
        for (int i = wave; i * D::RPI < A2; i += BS / 64) {
            const int ja0 = i * D::RPI;
            const int ja = ja0 + lrow;
            const int pos0 = (ja0 < h) ? ja0 : ja0 + nband;
            const int off = (jok && ja < A2) ? int(sidx(ja, z, j10 + col, N3, A1) * sizeof(V)) : BUF_OOB;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)(buf + pos0 * YJ), 16,
                                                     off, 0, 0, 0);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        lds_fft_lines<T, N2, YJ, BS, +1, YJ, true>(YJ, [&](int line) { return buf + line; }, A.tw);
        const rsrc_t o = make_rsrc(A.out[f], ybytes);
        for (int idx = threadIdx.x; idx < N2 * YJ; idx += BS) {
            const int jj = idx % YJ;
            const int n2 = idx / YJ;
            const int off = (j10 + jj < A1) ? int(sidx(n2, z, j10 + jj, N3, A1) * sizeof(V)) : BUF_OOB;
            bst<V>(buf[n2 * YJ + jj], o, off);
        }
        __syncthreads();  // every wave is done with buf before the next field's loads
    }
}

template <typename T, int N2, int YJ, int FR>
static hipError_t launch_y_v(const YArgs<T>& a, bool inv, hipStream_t s) {
    constexpr int lds = YJ * FR * y_ldp<N2, YJ>() * int(sizeof(cx<T>));
    if constexpr (lds > 160 * 1024) {
        return hipErrorInvalidValue;
    } else {
        dim3 grid(a.N3, (a.A1 + YJ - 1) / YJ);
        if (inv) hipLaunchKernelGGL((k_yinv<T, N2, YJ, FR>), grid, dim3(BS), 0, s, a);
        else hipLaunchKernelGGL((k_yfwd<T, N2, YJ, FR>), grid, dim3(BS), 0, s, a);
        return hipGetLastError();
    }
}

// the paired forward transform of nin fields runs k_yfwd_seq (16-mode tiles
// straight into LDS)
template <typename T, int N2>
static bool yfwd_seq_fits(int A1, int nin) {
    if constexpr (FftPlan<N2>::two && YDma<T, N2, 16>::LDS <= 160 * 1024) {
        return YDma<T, N2, 16>::fits(A1, nin);
    } else {
        (void)A1;
        (void)nin;
        return false;
    }
}

template <typename T, int N2>
static hipError_t launch_y_n(const YArgs<T>& a, bool inv, hipStream_t s) {
    // tiled inputs exist only for the paired forward transform on k_yfwd_seq
    if (a.tiled && (inv || !a.paired || !yfwd_seq_fits<T, N2>(a.A1, 2 * a.nf) || a.A1 % 16 != 0))
        return hipErrorInvalidValue;
    // the inverse loads 32-wide tiles (256-byte runs) straight into LDS, the
    // paired forward pass 16-wide tiles of one field at a time (LDS for six
    // blocks per CU in FP32; 32-wide tiles, or both fields of a pair in LDS,
    // measured no faster), where the grid allows it; else 16 lines per block
    // through registers: 16 j1 x 1 field (inverse and single fields) or 8 j1 x
    // 2 fields (paired forward)
    if constexpr (FftPlan<N2>::two && YDma<T, N2, 32>::LDS <= 160 * 1024) {
        using D = YDma<T, N2, 32>;
        // the last row load may run RPI-1 rows past A2: those land in the
        // pass-1 padding rows, which must hold them
        const int last = (a.A2 + D::RPI - 1) / D::RPI * D::RPI + (N2 - 2 * (N2 / 3) + 1);
        if (inv && N2 % D::RPI == 0 && (N2 / 3) % D::RPI == 0 && last <= D::ROWS &&
            (a.A1 * int(sizeof(cx<T>))) % 16 == 0) {
            dim3 grid(a.N3, (a.A1 + 31) / 32);
            hipLaunchKernelGGL((k_yinv_dma<T, N2, 32>), grid, dim3(BS), 0, s, a);
            return hipGetLastError();
        }
    }
    if (inv || !a.paired) return launch_y_v<T, N2, 16, 1>(a, inv, s);
    if (yfwd_seq_fits<T, N2>(a.A1, 2 * a.nf)) {
        if constexpr (FftPlan<N2>::two && YDma<T, N2, 16>::LDS <= 160 * 1024) {
            dim3 grid(a.N3, (a.A1 + 15) / 16);
            hipLaunchKernelGGL((k_yfwd_seq<T, N2, 16>), grid, dim3(BS), 0, s, a);
            return hipGetLastError();
        }
    }
    return launch_y_v<T, N2, 8, 2>(a, inv, s);
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

#ifdef SF_KERNELS_HOST_DEFS
bool yfwd_tiles(int N2, int A1, int precision) {
    if (A1 % 16 != 0) return false;
    switch (N2) {
#define SF_CASE(n) \
    case n:        \
        return precision == 8 ? yfwd_seq_fits<double, n>(A1, 8) : yfwd_seq_fits<float, n>(A1, 8);
        SF_SIZES(SF_CASE)
#undef SF_CASE
        default:
            return false;
    }
}
#endif

// ---------------------------------------------------------------------------
// single-field x transforms (ToNodal / ToModal for API fields), one block per
// (range of nzt z levels, n2)
// ---------------------------------------------------------------------------
template <typename T, int N1, int NL>
__global__ __launch_bounds__(BS) void k_x_c2r(XTArgs<T> A) {
    using V = cx<T>;
    constexpr int LDP = fft_ldp<N1>();
    __shared__ V buf[NL * LDP];
    const int n2 = blockIdx.y;
    const int N3 = A.N3, A1 = A.A1;
    for (int z0 = blockIdx.x * A.nzt; z0 < min(N3, (int(blockIdx.x) + 1) * A.nzt); z0 += NL) {
        const int nband = N1 - 2 * A1 + 1;
        for (int idx = threadIdx.x; idx < NL * nband; idx += BS)
            buf[(idx / nband) * LDP + (A1 + idx % nband)] = czero<V>();
        for (int idx = threadIdx.x; idx < NL * A1; idx += BS) {
            const int j1 = idx % A1;
            const int t = idx / A1;
            const int z = z0 + t;
            V X = czero<V>();
            if (z < N3) X = A.spec_in[sidx(n2, z, j1, N3, A1)];
            V* L = buf + t * LDP;
            if (j1 == 0) {
                L[(0)] = mkc<T>(X.x, T(0));
            } else {
                L[(j1)] = X;
                L[(N1 - j1)] = mkc<T>(X.x, -X.y);
            }
        }
        sync_lds();
        lds_fft<T, N1, NL, BS, +1>(buf, A.tw);
        for (int idx = threadIdx.x; idx < NL * N1; idx += BS) {
            const int t = idx % NL;
            const int n = idx / NL;
            const int z = z0 + t;
            if (z < N3) A.nodal_out[(size_t(n2) * N1 + n) * N3 + z] = buf[t * LDP + (n)].x;
        }
        sync_lds();
    }
}

template <typename T, int N1, int NL>
__global__ __launch_bounds__(BS) void k_x_r2c(XTArgs<T> A) {
    using V = cx<T>;
    constexpr int LDP = fft_ldp<N1>();
    __shared__ V buf[NL * LDP];
    const int n2 = blockIdx.y;
    const int N3 = A.N3, A1 = A.A1;
    for (int z0 = blockIdx.x * A.nzt; z0 < min(N3, (int(blockIdx.x) + 1) * A.nzt); z0 += NL) {
        for (int idx = threadIdx.x; idx < NL * N1; idx += BS) {
            const int t = idx % NL;
            const int n = idx / NL;
            const int z = z0 + t;
            V v = czero<V>();
            if (z < N3) v.x = A.nodal_in[(size_t(n2) * N1 + n) * N3 + z];
            buf[t * LDP + (n)] = v;
        }
        sync_lds();
        lds_fft<T, N1, NL, BS, -1>(buf, A.tw);
        for (int idx = threadIdx.x; idx < NL * A1; idx += BS) {
            const int k = idx % A1;
            const int t = idx / A1;
            const int z = z0 + t;
            if (z < N3) A.spec_out[sidx(n2, z, k, N3, A1)] = cscale(A.scale, buf[t * LDP + (k)]);
        }
        sync_lds();
    }
}

template <typename T, int N1>
static hipError_t launch_xt_n(const XTArgs<T>& a, int N2, bool c2r, hipStream_t s) {
    // z levels per LDS batch (a block sweeps a.nzt levels): 16 levels make the
    // nodal side's z runs 16 values long (4 levels: 115.5 -> 100.5 us for one
    // 384x192x384 FP32 field)
    constexpr int NL = 16;
    XTArgs<T> b = a;
    if (b.nzt % NL) b.nzt = ((b.nzt + NL - 1) / NL) * NL;
    dim3 grid((a.N3 + b.nzt - 1) / b.nzt, N2);
    if (c2r)
        hipLaunchKernelGGL((k_x_c2r<T, N1, NL>), grid, dim3(BS), 0, s, b);
    else
        hipLaunchKernelGGL((k_x_r2c<T, N1, NL>), grid, dim3(BS), 0, s, b);
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
                                                    size_t n, int N3, int A1, const T* __restrict__ w,
                                                    double* __restrict__ partial) {
    __shared__ double red[BS];
    double acc = 0;
    for (size_t e = size_t(blockIdx.x) * BS + threadIdx.x; e < n; e += size_t(gridDim.x) * BS) {
        const size_t rowz = e / size_t(A1);  // [ja][z][j1]
        const int j1 = int(e - rowz * size_t(A1));
        const int z = int(rowz % size_t(N3));
        const bool m00 = (j1 == 0) && (rowz < size_t(N3));  // ja == 0 and j1 == 0
        const cx<T> x = a[e], y = b[e];
        const double re = double(x.x) * double(y.x) + double(x.y) * double(y.y);
        acc += (m00 ? 1.0 : 2.0) * double(w[z]) * re;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = BS / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

static __global__ __launch_bounds__(BS) void k_dot_final(const double* __restrict__ partial, int n, double* out, int slot) {
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
hipError_t launch_dot(const cx<T>* a, const cx<T>* b, int A1, int A2, int N3, const T* w, double* partial,
                      double* out, int slot, hipStream_t s) {
    const size_t n = size_t(A1) * size_t(A2) * size_t(N3);
    int blocks = int((n + BS - 1) / BS);
    if (blocks > DOT_BLOCKS) blocks = DOT_BLOCKS;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL((k_dot_partial<T>), dim3(blocks), dim3(BS), 0, s, a, b, n, N3, A1, w, partial);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(BS), 0, s, partial, blocks, out, slot);
    return hipGetLastError();
}

#ifdef SF_KERNELS_HOST_DEFS
int max_reduction_blocks() { return DOT_BLOCKS; }
#endif

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

static __global__ __launch_bounds__(BS) void k_max_final(const double* __restrict__ partial, int n, double* out, int slot) {
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
__global__ __launch_bounds__(BS) void k_ends(cx<T>* f, int A1, int A2, int N3, int op) {
    const int s = blockIdx.x * BS + threadIdx.x;  // stack (ja, j1), j1 fastest
    if (s >= A1 * A2) return;
    const int ja = s / A1, j1 = s - (s / A1) * A1;
    auto at = [&](int z) -> cx<T>& { return f[sidx(ja, z, j1, N3, A1)]; };
    if (op == 2) {
        at(0) = at(1);
        at(N3 - 1) = at(N3 - 2);
    } else {
        at(0) = at(1) = at(N3 - 1) = at(N3 - 2) = czero<cx<T>>();
        if (op == 1) at(2) = czero<cx<T>>();
    }
}
template <typename T>
hipError_t launch_ends(cx<T>* f, int A1, int A2, int N3, int op, hipStream_t s) {
    hipLaunchKernelGGL((k_ends<T>), dim3((A1 * A2 + BS - 1) / BS), dim3(BS), 0, s, f, A1, A2, N3, op);
    return hipGetLastError();
}

// ddx (axis 0), ddy (axis 1), ddz (axis 2: Neumann->Dirichlet when zc row is
// ZC_DZND, Dirichlet->Neumann when ZC_DZDN; the caller passes the row).
template <typename T>
__global__ __launch_bounds__(BS) void k_dmul(cx<T>* out, const cx<T>* in, int A1, int A2, int N3, const T* kx,
                                             const T* ky, int axis, const T* zc, int N3p) {
    const size_t n = size_t(A1) * A2 * N3;
    for (size_t e = size_t(blockIdx.x) * BS + threadIdx.x; e < n; e += size_t(gridDim.x) * BS) {
        const size_t rowz = e / size_t(A1);  // [ja][z][j1]
        const int j1 = int(e - rowz * size_t(A1));
        const int z = int(rowz % size_t(N3));
        const int j2 = int(rowz / size_t(N3));
        cx<T> r;
        if (axis == 0) r = cik(kx[j1], in[e]);
        else if (axis == 1) r = cik(ky[j2], in[e]);
        else if (axis == 2) {  // Neumann in -> Dirichlet out
            const cx<T> fm = z > 0 ? in[e - size_t(A1)] : czero<cx<T>>();
            r = cscale(zc[ZC_DZND * N3p + z], csub(in[e], fm));
        } else {  // Dirichlet in -> Neumann out
            const cx<T> fp = z + 1 < N3 ? in[e + size_t(A1)] : czero<cx<T>>();
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

// Diffusion preconditioner, one wavefront per active mode (block = ja*A1 + j1).
// Modes with kx^2 + ky^2 <= k2max:
//   out = (tnu*(Dzz - k^2) - delta)^-1 ZeroEnds(in) - in/2
// with the Neumannified or Dirichletified rows of the implicit solves
// (IMEXRK.h:502-545), written as (I - (tnu/delta)*(Dzz - k^2)) x = -ZeroEnds(in)/delta
// so that the Crank-Nicolson row builder and the column solver serve it;
// other modes: out = -in.  (T L)^-1 - 1/2 is the Cayley approximation of
// (exp(T L) - I)^-1 for the diffusion operator L = nu (Dzz - k^2).
template <typename T, int M>
__global__ __launch_bounds__(64) void k_diffprec(DiffPrecArgs<T> P) {
    using V = cx<T>;
    const int lane = threadIdx.x;
    const int ja = blockIdx.x / P.A1;
    const int j1 = blockIdx.x - ja * P.A1;
    const int N = P.N3;
    int row[M];
    V x[M];
#pragma unroll
    for (int t = 0; t < M; ++t) {
        row[t] = lane * M + t;
        x[t] = row[t] < N ? P.in[sidx(ja, row[t], j1, N, P.A1)] : czero<V>();
    }
    const T kk = P.kx2[j1], kl = P.ky2[ja];
    V y[M];
    if (-(kk + kl) <= P.k2max) {  // wave-uniform
        const bool dir = P.dirichlet != 0;
        const T rs = T(-1) / P.delta;
#pragma unroll
        for (int t = 0; t < M; ++t) {
            const int j = row[t];
            const bool end = j >= N || j <= 1 || j >= N - 2 || (dir && j == 2);
            y[t] = end ? czero<V>() : cscale(rs, x[t]);
        }
        ModalArgs<T> A{};
        A.N3 = N;
        A.N3p = P.N3p;
        A.zc = P.zc;
        T a[M], b[M], c[M];
        auto zcoef = [&](int cr, int t) -> T { return row[t] < N ? P.zc[cr * P.N3p + row[t]] : T(0); };
        build_rows<T, M>(dir ? 1 : 0, A, row, kk, kl, false, P.tnu / P.delta, zcoef, a, b, c);
        TriFactor<T, M> F;
        tri_factor<T, M>(F, a, b, c, lane);
        tri_apply<T, M>(F, y, lane);
#pragma unroll
        for (int t = 0; t < M; ++t) y[t] = csub(y[t], cscale(T(0.5), x[t]));
    } else {
#pragma unroll
        for (int t = 0; t < M; ++t) y[t] = cscale(T(-1), x[t]);
    }
#pragma unroll
    for (int t = 0; t < M; ++t)
        if (row[t] < N) P.out[sidx(ja, row[t], j1, N, P.A1)] = y[t];
}

// The same preconditioner with the advection of each mode by the background,
// L = nu (Dzz - k^2) - i kx U(z): complex coefficients, so one thread per mode
// runs a Thomas sweep (the system is strictly diagonally dominant for
// delta > 0) with its modified super-diagonal in LDS and the modified
// right-hand side in out; ADV_T modes per block.
constexpr int ADV_T = ADVPREC_MODES;
template <typename T>
__global__ __launch_bounds__(ADV_T) void k_advprec(DiffPrecArgs<T> P) {
    using V = cx<T>;
    extern __shared__ unsigned char adv_lds[];
    V* cp = reinterpret_cast<V*>(adv_lds);  // [z][ADV_T]
    const int t = threadIdx.x;
    const int m = blockIdx.x * ADV_T + t;
    if (m >= P.A1 * P.A2) return;  // no barriers below
    const int ja = m / P.A1;
    const int j1 = m - ja * P.A1;
    const int N = P.N3;
    const T kk = P.kx2[j1], kl = P.ky2[ja];
    auto at = [&](int z) { return sidx(ja, z, j1, N, P.A1); };
    if (-(kk + kl) > P.k2max) {
        for (int z = 0; z < N; ++z) P.out[at(z)] = cscale(T(-1), P.in[at(z)]);
        return;
    }
    const bool dir = P.dirichlet != 0;
    const T coef = P.tnu / P.delta;
    const T rs = T(-1) / P.delta;
    const T adv = P.tadv * P.kx[j1] / P.delta;
    const int lo = dir ? ZC_DZZD_LO : ZC_DZZN_LO;
    // rows of (I - coef (Dzz - k^2) + i adv U) with the Neumannify/Dirichletify
    // replacements of the implicit solves
    auto rowc = [&](int j, T& a, V& b, T& c) {
        a = -(coef * P.zc[lo * P.N3p + j]);
        b.x = T(1) - coef * ((P.zc[(lo + 1) * P.N3p + j] + kk) + kl);
        b.y = adv * P.zc[ZC_UBAR * P.N3p + j];
        c = -(coef * P.zc[(lo + 2) * P.N3p + j]);
        if (j == 0) a = 0;
        if (j == N - 1) c = 0;
        const bool id = dir ? (j <= 2 || j >= N - 2) : (j == 0 || j == N - 1);
        if (id) {
            a = 0;
            b = mkc<T>(1, 0);
            c = 0;
        } else if (!dir && j == 1) {
            a = 0;
            b = mkc<T>(1, 0);
            c = -1;
        } else if (!dir && j == N - 2) {
            a = 1;
            b = mkc<T>(-1, 0);
            c = 0;
        }
    };
    auto rhs = [&](int j) {
        const bool end = j <= 1 || j >= N - 2 || (dir && j == 2);
        return end ? czero<V>() : cscale(rs, P.in[at(j)]);
    };
    auto cdiv = [](V a, V b) {
        const T r = T(1) / (b.x * b.x + b.y * b.y);
        return cscale(r, cmulc(a, b));
    };
    V cprev = czero<V>(), dprev = czero<V>();
    for (int j = 0; j < N; ++j) {
        T a, c;
        V b;
        rowc(j, a, b, c);
        const V den = csub(b, cscale(a, cprev));
        const V cj = cdiv(mkc<T>(c, 0), den);
        const V dj = cdiv(csub(rhs(j), cscale(a, dprev)), den);
        cp[j * ADV_T + t] = cj;
        P.out[at(j)] = dj;
        cprev = cj;
        dprev = dj;
    }
    V x = dprev;
    P.out[at(N - 1)] = csub(x, cscale(T(0.5), P.in[at(N - 1)]));
    for (int j = N - 2; j >= 0; --j) {
        x = csub(P.out[at(j)], cmul(cp[j * ADV_T + t], x));
        P.out[at(j)] = csub(x, cscale(T(0.5), P.in[at(j)]));
    }
}

template <typename T>
hipError_t launch_diffprec(const DiffPrecArgs<T>& p, hipStream_t s) {
    const int M = (p.N3 + 63) / 64;
    const unsigned nb = unsigned(p.A1) * unsigned(p.A2);
    if (nb == 0 || p.in == p.out || !(p.delta > T(0))) return hipErrorInvalidValue;
    if (p.tadv != T(0)) {
        const size_t lds = size_t(ADV_T) * size_t(p.N3) * sizeof(cx<T>);
        if (lds > ADVPREC_LDS) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_advprec<T>), dim3((nb + ADV_T - 1) / ADV_T), dim3(ADV_T), lds, s, p);
        return hipGetLastError();
    }
    switch (M) {
#define SF_M(m)                                                                   \
    case m:                                                                       \
        hipLaunchKernelGGL((k_diffprec<T, m>), dim3(nb), dim3(64), 0, s, p);      \
        break;
        SF_M(1) SF_M(2) SF_M(3) SF_M(4) SF_M(5) SF_M(6) SF_M(7) SF_M(8)
#undef SF_M
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

#define SF_INST_CORE(T) \
    template hipError_t launch_modal<T>(const ModalArgs<T>&, hipStream_t); \
    template hipError_t launch_xpass<T>(const XArgs<T>&, int, int, bool, hipStream_t); \
    template hipError_t launch_yinv<T>(const YArgs<T>&, int, hipStream_t); \
    template hipError_t launch_yfwd<T>(const YArgs<T>&, int, hipStream_t); \
    template hipError_t launch_x_c2r<T>(const XTArgs<T>&, int, int, hipStream_t); \
    template hipError_t launch_x_r2c<T>(const XTArgs<T>&, int, int, hipStream_t); \
    template hipError_t launch_dot<T>(const cx<T>*, const cx<T>*, int, int, int, const T*, double*, double*, int, hipStream_t); \
    template hipError_t launch_axpby<T>(cx<T>*, T, const cx<T>*, T, size_t, hipStream_t); \
    template hipError_t launch_maxabs<T>(const T*, size_t, int, const T*, double*, double*, int, hipStream_t); \
    template hipError_t launch_ends<T>(cx<T>*, int, int, int, int, hipStream_t); \
    template hipError_t launch_dmul<T>(cx<T>*, const cx<T>*, int, int, int, const T*, const T*, int, const T*, int, hipStream_t); \
    template hipError_t launch_diffprec<T>(const DiffPrecArgs<T>&, hipStream_t); \
    template int xpass_slots<T>(int, bool, int);

#define SF_INST_BATCH(T) \
    template hipError_t launch_modal_batch<T>(const ModalArgs<T>*, const int*, int, int, int, int, hipStream_t); \
    template hipError_t launch_xpass_batch<T>(const XArgs<T>*, const int*, int, int, int, int, hipStream_t); \
    template hipError_t launch_xpass_mixed<T>(const XArgs<T>*, const int*, int, int, int, int, int, hipStream_t); \
    template hipError_t launch_modal_mixed<T>(const ModalArgs<T>*, const int*, int, int, int, int, int, hipStream_t);

}  // namespace sf
