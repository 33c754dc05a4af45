// sf_fft.h -- batched complex FFTs staged in LDS for gfx950.
//
// A batch of NL lines of length N lives in LDS (line stride LD).  Each radix-R
// Stockham pass loads its butterflies into registers, synchronises, and stores
// them back in autosorted order, so a single LDS buffer suffices.  The radix
// plan (8, then 4, 2, 3, 5) is fixed at compile time from N.  Twiddles come
// from a table w[k] = exp(-2*pi*i*k/N) that the host fills in double precision.
// SIGN = -1 is FFTW_FORWARD, +1 FFTW_BACKWARD; both are unnormalised.
#pragma once
#include "sf_common.h"

namespace sf {

constexpr int pick_radix(int rem) {
    return (rem % 8 == 0) ? 8 : (rem % 4 == 0) ? 4 : (rem % 2 == 0) ? 2 : (rem % 3 == 0) ? 3 : (rem % 5 == 0) ? 5 : rem;
}

template <typename V, typename T, int SIGN>
__device__ __forceinline__ V mul_si(V a) {  // (SIGN*i) * a
    V r;
    r.x = -T(SIGN) * a.y;
    r.y = T(SIGN) * a.x;
    return r;
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void small_dft(cx<T>* v, const cx<T>* __restrict__ tw, int N) {
    using V = cx<T>;
    if constexpr (R == 2) {
        V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T s3 = T(SIGN) * T(0.86602540378443864676);
        V t1 = cadd(v[1], v[2]);
        V t2 = csub(v[0], cscale(T(0.5), t1));
        V d = csub(v[1], v[2]);
        V t3 = cscale(s3, d);
        v[0] = cadd(v[0], t1);
        V it3;
        it3.x = -t3.y;
        it3.y = t3.x;
        v[1] = cadd(t2, it3);
        v[2] = csub(t2, it3);
    } else if constexpr (R == 4) {
        V a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        V a2 = cadd(v[1], v[3]), a3 = mul_si<V, T, SIGN>(csub(v[1], v[3]));
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    } else if constexpr (R == 8) {
        const T c = T(0.70710678118654752440);
        V e0 = cadd(v[0], v[4]), e1 = csub(v[0], v[4]);
        V e2 = cadd(v[2], v[6]), e3 = mul_si<V, T, SIGN>(csub(v[2], v[6]));
        V E0 = cadd(e0, e2), E2 = csub(e0, e2), E1 = cadd(e1, e3), E3 = csub(e1, e3);
        V o0 = cadd(v[1], v[5]), o1 = csub(v[1], v[5]);
        V o2 = cadd(v[3], v[7]), o3 = mul_si<V, T, SIGN>(csub(v[3], v[7]));
        V O0 = cadd(o0, o2), O2 = csub(o0, o2), O1 = cadd(o1, o3), O3 = csub(o1, o3);
        // W8^1 = c(1 + SIGN i), W8^2 = SIGN i, W8^3 = c(-1 + SIGN i)
        V w1O1, w2O2, w3O3;
        w1O1.x = c * (O1.x - T(SIGN) * O1.y);
        w1O1.y = c * (O1.y + T(SIGN) * O1.x);
        w2O2 = mul_si<V, T, SIGN>(O2);
        w3O3.x = c * (-O3.x - T(SIGN) * O3.y);
        w3O3.y = c * (-O3.y + T(SIGN) * O3.x);
        v[0] = cadd(E0, O0);
        v[4] = csub(E0, O0);
        v[1] = cadd(E1, w1O1);
        v[5] = csub(E1, w1O1);
        v[2] = cadd(E2, w2O2);
        v[6] = csub(E2, w2O2);
        v[3] = cadd(E3, w3O3);
        v[7] = csub(E3, w3O3);
    } else {
        // generic small prime: O(R^2) DFT with twiddles from the length-N table
        V t[R];
        const int step = N / R;
#pragma unroll
        for (int s = 0; s < R; ++s) {
            V acc = czero<V>();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                V w = tw[((r * s) % R) * step];
                if (SIGN > 0) w.y = -w.y;
                acc = cadd(acc, cmul(v[r], w));
            }
            t[s] = acc;
        }
#pragma unroll
        for (int s = 0; s < R; ++s) v[s] = t[s];
    }
}

// One Stockham pass over NL lines.  Called by all BS threads of the block.
template <typename T, int N, int R, int NS, int NL, int LD, int BS, int SIGN>
__device__ __forceinline__ void fft_pass(cx<T>* buf, const cx<T>* __restrict__ tw) {
    using V = cx<T>;
    constexpr int NB = N / R;
    constexpr int TOT = NL * NB;
    constexpr int PER = (TOT + BS - 1) / BS;
    constexpr int TSTEP = N / (NS * R);
    V v[PER][R];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = int(threadIdx.x) + i * BS;
        if ((TOT % BS == 0) || g < TOT) {
            const int line = g / NB;
            const int j = g - line * NB;
            const int k = j % NS;
            const V* L = buf + line * LD;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = L[j + r * NB];
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    V w = tw[r * k * TSTEP];
                    if (SIGN > 0) w.y = -w.y;
                    v[i][r] = cmul(v[i][r], w);
                }
            }
            small_dft<T, R, SIGN>(v[i], tw, N);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = int(threadIdx.x) + i * BS;
        if ((TOT % BS == 0) || g < TOT) {
            const int line = g / NB;
            const int j = g - line * NB;
            const int k = j % NS;
            V* L = buf + line * LD;
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) L[o + r * NS] = v[i][r];
        }
    }
    __syncthreads();
}

template <typename T, int N, int NS, int NL, int LD, int BS, int SIGN>
__device__ __forceinline__ void fft_run(cx<T>* buf, const cx<T>* __restrict__ tw) {
    if constexpr (NS < N) {
        constexpr int R = pick_radix(N / NS);
        fft_pass<T, N, R, NS, NL, LD, BS, SIGN>(buf, tw);
        fft_run<T, N, NS * R, NL, LD, BS, SIGN>(buf, tw);
    }
}

// In-place FFT of NL lines in LDS.  The caller must have synchronised after
// filling buf; on return the results are visible to all threads.
template <typename T, int N, int NL, int LD, int BS, int SIGN>
__device__ __forceinline__ void lds_fft(cx<T>* buf, const cx<T>* __restrict__ tw) {
    fft_run<T, N, 1, NL, LD, BS, SIGN>(buf, tw);
}

}  // namespace sf
