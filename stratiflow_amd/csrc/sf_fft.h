// sf_fft.h -- batched complex FFTs staged in LDS for gfx950.
//
// A batch of NL lines of length N lives in LDS at line stride fft_ldp<N>().
// Lines are in natural order (element i at offset i) before and after a
// transform.  N = R1 * R2 (R1 = 16 or 8, R2 <= 32) takes two Stockham passes:
//   pass 1: butterfly j < R2 reads x[j + R2*r], r < R1, computes a DFT-R1 in
//           registers and writes it to row j of a padded R2 x (R1+1) table, so
//           that the stride-R1 stores do not collide in the LDS banks;
//   pass 2: butterfly j < R1 reads the table entries (r, j), r < R2, applies
//           the twiddles w^(r*j), computes a DFT-R2 and writes x[j + R1*r].
// Every LDS access is a compile-time offset from one per-thread base.  N <= 32
// is a single DFT-N.  The register DFTs are Cooley-Tukey compositions of
// 2/3/4/6/8-point kernels with compile-time twiddles; the pass-2 twiddles come
// from a table tw[j*R2 + r] = exp(-2*pi*i*r*j/N) that the host fills in double
// precision (fft_twiddle_table), so that each butterfly reads its twiddles at
// compile-time offsets from one base; entries N..2N-1 hold the same values in
// the order of pass 2 on lane pairs, tw[N + j*R2 + (r%2)*R2/2 + r/2].  SIGN = -1 is FFTW_FORWARD, +1 FFTW_BACKWARD; both unnormalised.
#pragma once
#include "sf_common.h"

#include <cmath>

namespace sf {

// ---- compile-time cos(2*pi*m/R) -------------------------------------------
constexpr double cpi = 3.14159265358979323846264338327950288;
constexpr double csin_series(double x) {  // |x| <= pi/4
    double term = x, sum = x;
    for (int n = 1; n < 12; ++n) {
        term *= -x * x / double((2 * n) * (2 * n + 1));
        sum += term;
    }
    return sum;
}
constexpr double ccos_series(double x) {  // |x| <= pi/4
    double term = 1, sum = 1;
    for (int n = 1; n < 12; ++n) {
        term *= -x * x / double((2 * n - 1) * (2 * n));
        sum += term;
    }
    return sum;
}
constexpr double ccos2pi(int m, int R) {
    m = ((m % R) + R) % R;
    const double a = 2 * cpi * double(m) / double(R);
    if (8 * m <= R) return ccos_series(a);
    if (8 * m <= 3 * R) return csin_series(cpi / 2 - a);
    if (8 * m <= 5 * R) return -ccos_series(cpi - a);
    if (8 * m <= 7 * R) return csin_series(a - 3 * cpi / 2);
    return ccos_series(2 * cpi - a);
}
// sin(2 pi m / R) = cos(2 pi m / R - pi/2)
constexpr double csin2pi(int m, int R) { return ccos2pi(4 * m - R, 4 * R); }

template <int R>
struct TwTab {
    double c[R];
    double s[R];
    constexpr TwTab() : c(), s() {
        for (int m = 0; m < R; ++m) {
            c[m] = ccos2pi(m, R);
            s[m] = csin2pi(m, R);
        }
    }
};
template <int R>
inline constexpr TwTab<R> kTw{};

// ---- register DFTs ---------------------------------------------------------
template <typename V, typename T, int SIGN>
__device__ __forceinline__ V mul_si(V a) {  // (SIGN*i) * a
    V r;
    r.x = -T(SIGN) * a.y;
    r.y = T(SIGN) * a.x;
    return r;
}

// a * exp(SIGN * 2*pi*i*m/R); m is a constant once the caller is unrolled
template <typename T, int SIGN, int R>
__device__ __forceinline__ cx<T> twv(cx<T> a, int m) {
    m = m % R;
    if (m == 0) return a;
    if (4 * m == R) return mul_si<cx<T>, T, SIGN>(a);
    if (2 * m == R) return mkc<T>(-a.x, -a.y);
    if (4 * m == 3 * R) return mul_si<cx<T>, T, -SIGN>(a);
    const T c = T(kTw<R>.c[m]);
    const T s = T(SIGN) * T(kTw<R>.s[m]);
    return mkc<T>(c * a.x - s * a.y, c * a.y + s * a.x);
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft3(cx<T>& a0, cx<T>& a1, cx<T>& a2) {
    using V = cx<T>;
    const T s3 = T(SIGN) * T(0.86602540378443864676);
    const V t1 = cadd(a1, a2);
    const V t2 = csub(a0, cscale(T(0.5), t1));
    const V t3 = cscale(s3, csub(a1, a2));
    V it3;
    it3.x = -t3.y;
    it3.y = t3.x;
    a0 = cadd(a0, t1);
    a1 = cadd(t2, it3);
    a2 = csub(t2, it3);
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void reg_dft(cx<T>* v);

// R = P*Q in registers: n = Q*n1 + n2, k = k1 + P*k2
template <typename T, int P, int Q, int SIGN>
__device__ __forceinline__ void reg_dft_ct(cx<T>* v) {
    using V = cx<T>;
    constexpr int R = P * Q;
    V t[R];
#pragma unroll
    for (int n2 = 0; n2 < Q; ++n2) {
        V c[P];
#pragma unroll
        for (int n1 = 0; n1 < P; ++n1) c[n1] = v[Q * n1 + n2];
        reg_dft<T, P, SIGN>(c);
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) t[k1 * Q + n2] = twv<T, SIGN, R>(c[k1], n2 * k1);
    }
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
        V c[Q];
#pragma unroll
        for (int n2 = 0; n2 < Q; ++n2) c[n2] = t[k1 * Q + n2];
        reg_dft<T, Q, SIGN>(c);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) v[k1 + P * k2] = c[k2];
    }
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void reg_dft(cx<T>* v) {
    using V = cx<T>;
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        dft3<T, SIGN>(v[0], v[1], v[2]);
    } else if constexpr (R == 4) {
        V a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        V a2 = cadd(v[1], v[3]), a3 = mul_si<V, T, SIGN>(csub(v[1], v[3]));
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    } else if constexpr (R == 6) {
        // 6 = 2 x 3: E = DFT3(v0,v2,v4), O = DFT3(v1,v3,v5), y_k = E_k + W6^k O_k
        V e0 = v[0], e1 = v[2], e2 = v[4];
        V o0 = v[1], o1 = v[3], o2 = v[5];
        dft3<T, SIGN>(e0, e1, e2);
        dft3<T, SIGN>(o0, o1, o2);
        const V w1o1 = twv<T, SIGN, 6>(o1, 1);
        const V w2o2 = twv<T, SIGN, 6>(o2, 2);
        v[0] = cadd(e0, o0);
        v[3] = csub(e0, o0);
        v[1] = cadd(e1, w1o1);
        v[4] = csub(e1, w1o1);
        v[2] = cadd(e2, w2o2);
        v[5] = csub(e2, w2o2);
    } else if constexpr (R == 8) {
        V e0 = cadd(v[0], v[4]), e1 = csub(v[0], v[4]);
        V e2 = cadd(v[2], v[6]), e3 = mul_si<V, T, SIGN>(csub(v[2], v[6]));
        V E0 = cadd(e0, e2), E2 = csub(e0, e2), E1 = cadd(e1, e3), E3 = csub(e1, e3);
        V o0 = cadd(v[1], v[5]), o1 = csub(v[1], v[5]);
        V o2 = cadd(v[3], v[7]), o3 = mul_si<V, T, SIGN>(csub(v[3], v[7]));
        V O0 = cadd(o0, o2), O2 = csub(o0, o2), O1 = cadd(o1, o3), O3 = csub(o1, o3);
        const V w1O1 = twv<T, SIGN, 8>(O1, 1);
        const V w2O2 = mul_si<V, T, SIGN>(O2);
        const V w3O3 = twv<T, SIGN, 8>(O3, 3);
        v[0] = cadd(E0, O0);
        v[4] = csub(E0, O0);
        v[1] = cadd(E1, w1O1);
        v[5] = csub(E1, w1O1);
        v[2] = cadd(E2, w2O2);
        v[6] = csub(E2, w2O2);
        v[3] = cadd(E3, w3O3);
        v[7] = csub(E3, w3O3);
    } else if constexpr (R == 12) {
        reg_dft_ct<T, 4, 3, SIGN>(v);
    } else if constexpr (R == 16) {
        reg_dft_ct<T, 4, 4, SIGN>(v);
    } else if constexpr (R == 24) {
        reg_dft_ct<T, 8, 3, SIGN>(v);
    } else if constexpr (R == 32) {
        reg_dft_ct<T, 8, 4, SIGN>(v);
    } else {
        static_assert(R == 0, "unsupported register DFT size");
    }
}

// ---- plan ------------------------------------------------------------------
// R1F = 0: the default first radix (16 where N/16 <= 32, else 8); R1F = 8
// or 16 asks for that first radix where it gives R2 <= 32 (fft_plan_ok), the
// default otherwise.
template <int N, int R1F>
constexpr bool fft_plan_ok() {
    return N > 32 && (R1F == 8 || R1F == 16) && N % R1F == 0 && N / R1F <= 32;
}
template <int N, int R1F = 0>
struct FftPlan {
    static constexpr bool two = N > 32;
    static constexpr int R1 = !two ? N
                              : fft_plan_ok<N, R1F>() ? R1F
                                                      : ((N % 16 == 0 && N / 16 <= 32) ? 16 : 8);
    static constexpr int R2 = N / R1;
    static_assert(!two || (R1 * R2 == N && R2 <= 32), "unsupported FFT length");
};

// LDS line stride: room for the padded pass-1 table, staggered by half a bank
// cycle (16 complex values) between consecutive lines
template <int N, int R1F = 0>
constexpr int fft_ldp() {
    if constexpr (N <= 32) {
        return N + 1;
    } else {
        int l = N + FftPlan<N, R1F>::R2;
        while (l % 32 != 16) ++l;
        return l;
    }
}

// ES: element stride of a line (1: contiguous lines; lines interleaved with
// ES = line count when element i of line l sits at lp(l) + i*ES).  IL: work
// items run over the lines first (consecutive lanes take consecutive lines,
// which is what keeps interleaved lines free of LDS bank conflicts).
template <typename T, int N, int NLMAX, int BS, int SIGN, int ES = 1, bool IL = false, int R1F = 0, typename LP>
__device__ __forceinline__ void fft_pass1(int nl, LP lp, const int tid) {
    using V = cx<T>;
    constexpr int R1 = FftPlan<N, R1F>::R1, R2 = FftPlan<N, R1F>::R2;
    constexpr int PER = (NLMAX * R2 + BS - 1) / BS;
    const int tot = nl * R2;
    V v[PER][R1];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = tid + i * BS;
        if (g < tot) {
            const int line = IL ? g % nl : g / R2;
            const int j = IL ? g / nl : g - line * R2;
            const V* L = lp(line) + j * ES;
#pragma unroll
            for (int r = 0; r < R1; ++r) v[i][r] = L[R2 * r * ES];
            reg_dft<T, R1, SIGN>(v[i]);
        }
    }
    sync_lds();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = tid + i * BS;
        if (g < tot) {
            const int line = IL ? g % nl : g / R2;
            const int j = IL ? g / nl : g - line * R2;
            V* Lp = lp(line) + j * (R1 + 1) * ES;
#pragma unroll
            for (int r = 0; r < R1; ++r) Lp[r * ES] = v[i][r];
        }
    }
    sync_lds();
}

// pass 2 twiddles: tw[j*R2 + r] = exp(-2 pi i r j / N), j < R1, r < R2 (each
// lane reads its R2 - 1 twiddles from one contiguous run, in wide loads)
template <typename T, int N, int NLMAX, int BS, int SIGN, int ES = 1, bool IL = false, int R1F = 0, typename LP>
__device__ __forceinline__ void fft_pass2(int nl, LP lp, const cx<T>* __restrict__ tw, const int tid) {
    using V = cx<T>;
    constexpr int R1 = FftPlan<N, R1F>::R1, R2 = FftPlan<N, R1F>::R2;
    constexpr int PER = (NLMAX * R1 + BS - 1) / BS;
    const int tot = nl * R1;
    V v[PER][R2];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = tid + i * BS;
        if (g < tot) {
            const int line = IL ? g % nl : g / R1;
            const int j = IL ? g / nl : g - line * R1;
            const V* Lp = lp(line) + j * ES;
            const V* twj = tw + j * R2;
#pragma unroll
            for (int r = 0; r < R2; ++r) v[i][r] = Lp[r * (R1 + 1) * ES];
#pragma unroll
            for (int r = 1; r < R2; ++r) {
                V w = twj[r];
                if (SIGN > 0) w.y = -w.y;
                v[i][r] = cmul(v[i][r], w);
            }
            reg_dft<T, R2, SIGN>(v[i]);
        }
    }
    sync_lds();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = tid + i * BS;
        if (g < tot) {
            const int line = IL ? g % nl : g / R1;
            const int j = IL ? g / nl : g - line * R1;
            V* L = lp(line) + j * ES;
#pragma unroll
            for (int r = 0; r < R2; ++r) L[R1 * r * ES] = v[i][r];
        }
    }
    sync_lds();
}

// Pass 2 on lane pairs: butterfly (line, j) of the radix-R2 pass is split over
// two adjacent lanes (h = lane & 1).  Lane h takes the inputs r = 2m + h, applies
// their twiddles and a DFT of R2/2 points; lane 1 multiplies by W^k, the pair
// exchanges its halves through DPP (quad_perm [1,0,3,2]) and lane h writes the
// outputs k + h*R2/2: X_k = E_k + W^k O_k, X_{k+R2/2} = E_k - W^k O_k.  Every
// pair is whole: nl*R1*2 work items, even, lanes of a pair in one wave.  Each
// lane reads its R2/2 twiddles as one contiguous run of the pair-ordered half
// of the table.  The two lanes of a pair store R1*R2/2 values apart, in one
// bank when that is a multiple of 32; pairing lanes l and l+32 instead
// (v_permlane32_swap) avoids that conflict but costs a copy per exchanged value
// and measured no faster.
__device__ __forceinline__ float dpp_swap_pair(float x) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ double dpp_swap_pair(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_mov_dpp(int(b & 0xffffffff), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

template <typename T, int N, int NLMAX, int BS, int SIGN, int ES = 1, int R1F = 0, typename LP>
__device__ __forceinline__ void fft_pass2_pairs(int nl, LP lp, const cx<T>* __restrict__ tw, const int tid) {
    using V = cx<T>;
    constexpr int R1 = FftPlan<N, R1F>::R1, R2 = FftPlan<N, R1F>::R2, H = R2 / 2;
    static_assert(R2 % 2 == 0, "pass 2 on lane pairs needs an even radix");
    static_assert(BS % 64 == 0, "pass 2 on lane pairs needs whole waves");
    constexpr int PER = (NLMAX * R1 * 2 + BS - 1) / BS;
    const int tot = nl * R1 * 2;
    V v[PER][H];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = tid + i * BS;
        if (g < tot) {
            const int h = g & 1;
            const int line = (g >> 1) / R1;
            const int j = (g >> 1) - line * R1;
            const V* Lp = lp(line) + (j + h * (R1 + 1)) * ES;
            const V* twj = tw + N + j * R2 + h * H;
#pragma unroll
            for (int m = 0; m < H; ++m) {
                V w = twj[m];
                if (SIGN > 0) w.y = -w.y;
                v[i][m] = cmul(Lp[2 * m * (R1 + 1) * ES], w);
            }
            reg_dft<T, H, SIGN>(v[i]);
            // lane 1 of the pair: W^k O_k
#pragma unroll
            for (int k = 1; k < H; ++k) {
                const V t = twv<T, SIGN, R2>(v[i][k], k);
                v[i][k].x = h ? t.x : v[i][k].x;
                v[i][k].y = h ? t.y : v[i][k].y;
            }
            const T sg = h ? T(-1) : T(1);
#pragma unroll
            for (int k = 0; k < H; ++k) {
                const T px = dpp_swap_pair(v[i][k].x), py = dpp_swap_pair(v[i][k].y);
                v[i][k].x = px + sg * v[i][k].x;
                v[i][k].y = py + sg * v[i][k].y;
            }
        }
    }
    sync_lds();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int g = tid + i * BS;
        if (g < tot) {
            const int h = g & 1;
            const int line = (g >> 1) / R1;
            const int j = (g >> 1) - line * R1;
            V* L = lp(line) + (j + R1 * H * h) * ES;
#pragma unroll
            for (int k = 0; k < H; ++k) L[R1 * k * ES] = v[i][k];
        }
    }
    sync_lds();
}

template <typename T, int N, int NLMAX, int BS, int SIGN, int ES = 1, typename LP>
__device__ __forceinline__ void fft_single(int nl, LP lp, const int tid) {
    using V = cx<T>;
    constexpr int PER = (NLMAX + BS - 1) / BS;
    V v[PER][N];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int line = tid + i * BS;
        if (line < nl) {
            const V* L = lp(line);
#pragma unroll
            for (int r = 0; r < N; ++r) v[i][r] = L[r * ES];
            reg_dft<T, N, SIGN>(v[i]);
        }
    }
    sync_lds();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int line = tid + i * BS;
        if (line < nl) {
            V* L = lp(line);
#pragma unroll
            for (int r = 0; r < N; ++r) L[r * ES] = v[i][r];
        }
    }
    sync_lds();
}

// In-place FFT of nl <= NLMAX lines in LDS, each in natural order at lp(line)
// with room for fft_ldp<N>() values (element stride ES: room for that many
// positions, N plus the pass-1 padding of R2).  The caller must have synchronised after
// filling the lines; on return the results are visible to all threads.  tid
// is the thread's index in [0, BS) for the distribution of the work items (a
// block may permute it to spread partially filled rounds over its waves).
// PAIRS (contiguous lines, even R2) runs pass 2 on lane pairs; it needs tid
// and tid ^ 1 on adjacent lanes of one wave, which a rotation of tid by whole
// waves keeps, and all BS threads calling.
template <typename T, int N, int NLMAX, int BS, int SIGN, int ES = 1, bool IL = false, int R1F = 0,
          bool PAIRS = false, typename LP>
__device__ __forceinline__ void lds_fft_lines(int nl, LP lp, const cx<T>* __restrict__ tw,
                                              const int tid = int(threadIdx.x)) {
    if constexpr (FftPlan<N>::two) {
        fft_pass1<T, N, NLMAX, BS, SIGN, ES, IL, R1F>(nl, lp, tid);
        if constexpr (PAIRS && !IL && FftPlan<N, R1F>::R2 % 2 == 0)
            fft_pass2_pairs<T, N, NLMAX, BS, SIGN, ES, R1F>(nl, lp, tw, tid);
        else
            fft_pass2<T, N, NLMAX, BS, SIGN, ES, IL, R1F>(nl, lp, tw, tid);
    } else {
        fft_single<T, N, NLMAX, BS, SIGN, ES>(nl, lp, tid);
        (void)tw;
    }
}

// NL contiguous lines at stride fft_ldp<N>()
template <typename T, int N, int NL, int BS, int SIGN>
__device__ __forceinline__ void lds_fft(cx<T>* buf, const cx<T>* __restrict__ tw, const int tid = int(threadIdx.x)) {
    constexpr int LDP = fft_ldp<N>();
    lds_fft_lines<T, N, NL, BS, SIGN>(NL, [buf](int line) { return buf + line * LDP; }, tw, tid);
}

// host: twiddle table of 2N entries for the plan FftPlan<N, R1F>: the layout
// fft_pass2 reads, then the same values in the order fft_pass2_pairs reads
template <int N, int R1F = 0>
void fft_twiddle_table(double* re, double* im) {
    constexpr int R1 = FftPlan<N, R1F>::R1, R2 = FftPlan<N, R1F>::R2;
    for (int k = 0; k < N; ++k) {
        re[k] = 1.0;
        im[k] = 0.0;
    }
    if constexpr (FftPlan<N, R1F>::two) {
        for (int j = 0; j < R1; ++j)
            for (int r = 0; r < R2; ++r) {
                const double a = -2.0 * cpi * double((r * j) % N) / double(N);
                re[j * R2 + r] = std::cos(a);
                im[j * R2 + r] = std::sin(a);
            }
    }
    for (int k = 0; k < N; ++k) {
        re[N + k] = re[k];
        im[N + k] = im[k];
    }
    if constexpr (FftPlan<N, R1F>::two && R2 % 2 == 0) {
        for (int j = 0; j < R1; ++j)
            for (int r = 0; r < R2; ++r) {
                re[N + j * R2 + (r & 1) * (R2 / 2) + r / 2] = re[j * R2 + r];
                im[N + j * R2 + (r & 1) * (R2 / 2) + r / 2] = im[j * R2 + r];
            }
    }
}

}  // namespace sf
