// sf_tri.h -- one-wavefront tridiagonal solver for a vertical column.
//
// The reference solves every vertical system with a sequential Thomas sweep
// (Tridiagonal.h:29-53).  On a 64-wide wavefront the column of Np = 64*M rows
// is split into 64 chunks of M consecutive rows, one chunk per lane:
//   * rows 0..M-2 of a chunk are "interior"; the last row is the chunk's
//     interface unknown s_l;
//   * each lane eliminates its interior block with a local Thomas sweep and
//     expresses its interior unknowns as  x = g - v*s_{l-1} - w*s_l;
//   * substituting into the interface rows gives a tridiagonal Schur system
//     in s_0..s_63, solved by parallel cyclic reduction across the lanes
//     (6 levels of cross-lane shuffles);
//   * each lane then back-substitutes its interior rows.
// Matrix-only quantities (local factors, spikes v and w, PCR multipliers) are
// computed once per matrix and reused for every right-hand side that shares
// it.  Coefficients are real; right-hand sides are complex (real and
// imaginary parts see identical arithmetic, as in Field.h:520-531).  In FP32
// the pivots are inverted with the hardware reciprocal (1 ulp).
// Padding rows beyond N3 must be identity rows with zero right-hand side.
#pragma once
#include "sf_common.h"

namespace sf {

template <typename T>
__device__ __forceinline__ T shfl_idx(T v, int src) {
    return __shfl(v, src, 64);
}

template <typename T, int M>
struct TriFactor {
    T a[M], c[M];      // original sub/super diagonals of this lane's rows
    T rden[M];         // interior Thomas: 1/den
    T cc[M];           // interior Thomas: modified super diagonal
    T v[M], w[M];      // spikes
    T alpha[6], gamma[6];
    T rB;              // 1 / final PCR diagonal
    T aI, cI;          // interface row's sub/super diagonal
    T vn0, gn_dummy;   // next lane's v[0]
};

// Build factors.  a,b,c: the lane's M rows (a[0] of lane 0 and c[M-1] of
// lane 63 must be zero).
template <typename T, int M>
__device__ __forceinline__ void tri_factor(TriFactor<T, M>& F, const T (&a)[M], const T (&b)[M], const T (&c)[M],
                                           int lane) {
#pragma unroll
    for (int t = 0; t < M; ++t) {
        F.a[t] = a[t];
        F.c[t] = c[t];
    }
    T As, Bs, Cs;
    if constexpr (M == 1) {
        As = a[0];
        Bs = b[0];
        Cs = c[0];
    } else {
        // interior rows 0..M-2
        T den = b[0];
        F.rden[0] = recip(den);
#pragma unroll
        for (int t = 0; t < M - 2; ++t) {
            F.cc[t] = c[t] * F.rden[t];
            den = b[t + 1] - a[t + 1] * F.cc[t];
            F.rden[t + 1] = recip(den);
        }
        // spike v: A_int^{-1} (a_0 e_0)
        {
            T D[M];
            D[0] = a[0] * F.rden[0];
#pragma unroll
            for (int t = 1; t <= M - 2; ++t) D[t] = (-a[t] * D[t - 1]) * F.rden[t];
            F.v[M - 2] = D[M - 2];
#pragma unroll
            for (int t = M - 3; t >= 0; --t) F.v[t] = D[t] - F.cc[t] * F.v[t + 1];
        }
        // spike w: A_int^{-1} (c_{M-2} e_{M-2})
        {
            F.w[M - 2] = c[M - 2] * F.rden[M - 2];
#pragma unroll
            for (int t = M - 3; t >= 0; --t) F.w[t] = -F.cc[t] * F.w[t + 1];
        }
        const T aI = a[M - 1], cI = c[M - 1];
        F.aI = aI;
        F.cI = cI;
        T vn0 = shfl_idx(F.v[0], (lane + 1) & 63);
        T wn0 = shfl_idx(F.w[0], (lane + 1) & 63);
        if (lane == 63) {
            vn0 = 0;
            wn0 = 0;
        }
        F.vn0 = vn0;
        As = -aI * F.v[M - 2];
        Bs = b[M - 1] - aI * F.w[M - 2] - cI * vn0;
        Cs = -cI * wn0;
    }
    // parallel cyclic reduction on the 64-unknown Schur system
#pragma unroll
    for (int lev = 0; lev < 6; ++lev) {
        const int s = 1 << lev;
        T Am = shfl_idx(As, (lane - s) & 63), Bm = shfl_idx(Bs, (lane - s) & 63), Cm = shfl_idx(Cs, (lane - s) & 63);
        T Ap = shfl_idx(As, (lane + s) & 63), Bp = shfl_idx(Bs, (lane + s) & 63), Cp = shfl_idx(Cs, (lane + s) & 63);
        if (lane - s < 0) {
            Am = 0;
            Bm = 1;
            Cm = 0;
        }
        if (lane + s > 63) {
            Ap = 0;
            Bp = 1;
            Cp = 0;
        }
        const T al = -As * recip(Bm);
        const T ga = -Cs * recip(Bp);
        F.alpha[lev] = al;
        F.gamma[lev] = ga;
        const T nA = al * Am;
        const T nC = ga * Cp;
        Bs = Bs + al * Cm + ga * Ap;
        As = nA;
        Cs = nC;
    }
    F.rB = recip(Bs);
}

// Solve in place: d holds the lane's M right-hand-side rows on entry and the
// solution on exit.
template <typename T, int M>
__device__ __forceinline__ void tri_apply(const TriFactor<T, M>& F, cx<T> (&d)[M], int lane) {
    using V = cx<T>;
    V Ds;
    V g[M];
    if constexpr (M == 1) {
        Ds = d[0];
    } else {
        // interior forward/back sweep for g
        g[0] = cscale(F.rden[0], d[0]);
#pragma unroll
        for (int t = 1; t <= M - 2; ++t) g[t] = cscale(F.rden[t], csub(d[t], cscale(F.a[t], g[t - 1])));
#pragma unroll
        for (int t = M - 3; t >= 0; --t) g[t] = csub(g[t], cscale(F.cc[t], g[t + 1]));
        V gn0;
        gn0.x = shfl_idx(g[0].x, (lane + 1) & 63);
        gn0.y = shfl_idx(g[0].y, (lane + 1) & 63);
        if (lane == 63) gn0 = czero<V>();
        Ds = csub(csub(d[M - 1], cscale(F.aI, g[M - 2])), cscale(F.cI, gn0));
    }
#pragma unroll
    for (int lev = 0; lev < 6; ++lev) {
        const int s = 1 << lev;
        V Dm, Dp;
        Dm.x = shfl_idx(Ds.x, (lane - s) & 63);
        Dm.y = shfl_idx(Ds.y, (lane - s) & 63);
        Dp.x = shfl_idx(Ds.x, (lane + s) & 63);
        Dp.y = shfl_idx(Ds.y, (lane + s) & 63);
        if (lane - s < 0) Dm = czero<V>();
        if (lane + s > 63) Dp = czero<V>();
        Ds = cadd(cadd(Ds, cscale(F.alpha[lev], Dm)), cscale(F.gamma[lev], Dp));
    }
    const V sl = cscale(F.rB, Ds);
    if constexpr (M == 1) {
        d[0] = sl;
    } else {
        V sm;
        sm.x = shfl_idx(sl.x, (lane - 1) & 63);
        sm.y = shfl_idx(sl.y, (lane - 1) & 63);
        if (lane == 0) sm = czero<V>();
#pragma unroll
        for (int t = 0; t <= M - 2; ++t) d[t] = csub(csub(g[t], cscale(F.v[t], sm)), cscale(F.w[t], sl));
        d[M - 1] = sl;
    }
}

}  // namespace sf
