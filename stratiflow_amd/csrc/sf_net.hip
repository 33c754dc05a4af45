// sf_net.hip -- learned preconditioner: a 1D convolutional network over the
// vertical profile of every horizontal Fourier mode, trained and applied on
// the GPU with BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate).
//
// Network: L layers of Conv1D(kernel 5, zero padding) with ReLU between them,
// H hidden channels.  Input channels per mode (ja, j1) and height z: Re and Im
// of every prognostic field (u1, u2 in 3D, u3, b) divided by the RMS of the
// input vector, kx(j1)/kx_max and ky(ja)/ky_max.  Output channels: the
// correction to the input, so that the preconditioner is
//     M^-1 v = v + s * net(v / s),   s = RMS(v),
// which is the identity while the last layer is zero (its initial value) and
// positively homogeneous in v.  Training pairs (w, q) with A q = w (w = A q is
// what every Krylov step of GMRES computes) teach net(w/s) ~ q/RMS(q) - w/s:
// M^-1 w then points along A^-1 w, which is all a (flexible) GMRES direction
// needs, and the targets stay O(1) however small the eigenvalues of A are.
//
// Tensors are [position][channel] with channels fastest, position p = profile
// * N3 + z; channel counts are padded to multiples of 32 (the MFMA K step).
// One wavefront computes 16 positions x all output channels of a layer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "sf_internal.h"
#include "sf_kernels.h"

namespace sf {
namespace net {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int KT = 5;    // taps
constexpr int NET_DMAX = 8;  // largest dilation of a layer
constexpr int NB = 256;  // threads per block

__device__ __forceinline__ bf16x8 ld8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = bf16(0.0f);
    return z;
}

// out[p][co] = bias[co] + sum_{t,ci} W[t][co][ci] * in[p+(t-2)*dil][ci], in
// taken as zero outside the profile of p.  Optional: mask (out = 0 where mask
// <= 0), FP32 copy, BF16 copy (with ReLU when relu != 0).  W is [KT][Cout][Cin].
template <int NT>
__global__ __launch_bounds__(NB) void k_conv(const bf16* __restrict__ in, int Cin, const bf16* __restrict__ W,
                                             const float* __restrict__ bias, int P, int Z, float* __restrict__ out_f,
                                             bf16* __restrict__ out_b, int relu, const float* __restrict__ mask,
                                             int dil) {
    constexpr int Cout = NT * 16;
    const int lane = threadIdx.x & 63;
    const int p0 = (blockIdx.x * (NB / 64) + (threadIdx.x >> 6)) * 16;
    if (p0 >= P) return;  // wave-uniform; this kernel has no block barriers
    const int m = lane & 15, kg = lane >> 4;
    const int p = p0 + m;
    const int b = p / Z, z = p - (p / Z) * Z;
    f32x4 acc[NT];
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < KT; ++t) {
        const int zz = z + (t - 2) * dil;
        const bool ok = (p < P) && zz >= 0 && zz < Z;
        const bf16* arow = in + (size_t(b) * Z + (ok ? zz : 0)) * Cin + kg * 8;
        const bf16* wrow = W + size_t(t) * Cout * Cin + size_t(m) * Cin + kg * 8;
        for (int c0 = 0; c0 < Cin; c0 += 32) {
            const bf16x8 a = ok ? ld8(arow + c0) : zero8();
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const bf16x8 w = ld8(wrow + size_t(j) * 16 * Cin + c0);
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, w, acc[j], 0, 0, 0);
            }
        }
    }
    // C layout: lane (n = lane & 15, g = lane >> 4) holds rows g*4 .. g*4+3 of column n
    const int n = lane & 15, g = lane >> 4;
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int pp = p0 + g * 4 + i;
            const int co = j * 16 + n;
            if (pp >= P) continue;
            float v = acc[j][i] + (bias ? bias[co] : 0.f);
            const size_t o = size_t(pp) * Cout + co;
            if (mask) v = mask[o] > 0.f ? v : 0.f;
            if (out_f) out_f[o] = v;
            if (out_b) out_b[o] = bf16(relu ? fmaxf(v, 0.f) : v);
        }
}

// Partial weight gradients dW[t][co][ci] = sum_p g[p][co] * a[p+(t-2)*dil][ci]
// (same profile) over a grid-stride range of 64-position chunks, one partial
// per block: part[block][KT][Cout][Cin].  Chunks go through LDS transposed, so
// that the MFMA fragments read 8 consecutive positions.  dil <= NET_DMAX.
template <int NTO, int NTI>
__global__ __launch_bounds__(NB) void k_conv_dw(const bf16* __restrict__ g, const bf16* __restrict__ a, int P, int Z,
                                                float* __restrict__ part, int dil) {
    constexpr int Cout = NTO * 16, Cin = NTI * 16;
    constexpr int PC = 64;                    // positions per chunk
    constexpr int AW = PC + 4 * NET_DMAX + 8;  // staged a positions p0-2*dil .. p0+PC-1+2*dil (+pad)
    constexpr int NTILE = KT * NTO * NTI;
    constexpr int TPW = (NTILE + 3) / 4;  // tiles per wave
    __shared__ bf16 gT[Cout * PC];
    __shared__ bf16 aT[Cin * AW];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kg = lane >> 4;
    f32x4 acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int p0 = blockIdx.x * PC; p0 < P; p0 += gridDim.x * PC) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < PC * Cout; idx += NB) {
            const int q = idx / Cout, co = idx - (idx / Cout) * Cout;
            const int p = p0 + q;
            gT[co * PC + q] = p < P ? g[size_t(p) * Cout + co] : bf16(0.f);
        }
        for (int idx = threadIdx.x; idx < (PC + 4 * dil) * Cin; idx += NB) {
            const int q = idx / Cin, ci = idx - (idx / Cin) * Cin;
            const int p = p0 - 2 * dil + q;
            aT[ci * AW + q] = (p >= 0 && p < P) ? a[size_t(p) * Cin + ci] : bf16(0.f);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < TPW; ++k) {
            const int tile = wave + 4 * k;
            if (tile >= NTILE) break;
            const int t = tile / (NTO * NTI);
            const int jo = (tile / NTI) % NTO;
            const int ji = tile % NTI;
#pragma unroll
            for (int kb = 0; kb < PC; kb += 32) {
                const bf16x8 av = ld8(&gT[(jo * 16 + m) * PC + kb + kg * 8]);
                bf16x8 bv;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int q = kb + kg * 8 + i;  // position p0+q
                    const int p = p0 + q;
                    const int z = p - (p / Z) * Z + (t - 2) * dil;
                    bv[i] = (z >= 0 && z < Z) ? aT[(ji * 16 + m) * AW + q + t * dil] : bf16(0.f);
                }
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[k], 0, 0, 0);
            }
        }
    }
    float* out = part + size_t(blockIdx.x) * KT * Cout * Cin;
    const int n = lane & 15, gg = lane >> 4;
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
        const int tile = wave + 4 * k;
        if (tile >= NTILE) break;
        const int t = tile / (NTO * NTI);
        const int jo = (tile / NTI) % NTO;
        const int ji = tile % NTI;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int co = jo * 16 + gg * 4 + i, ci = ji * 16 + n;
            out[(size_t(t) * Cout + co) * Cin + ci] = acc[k][i];
        }
    }
}

// sum over blocks of partial results: dst[e] = sum_b part[b][e]
__global__ __launch_bounds__(NB) void k_sum_parts(const float* __restrict__ part, int nblk, int n, float* __restrict__ dst) {
    const int e = blockIdx.x * NB + threadIdx.x;
    if (e >= n) return;
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[size_t(b) * n + e];
    dst[e] = s;
}

// partial bias gradients: part[block][co] = sum over the block's positions of g[p][co]
__global__ __launch_bounds__(NB) void k_bias_part(const bf16* __restrict__ g, int P, int Cout, float* __restrict__ part) {
    __shared__ float red[NB];
    const int co = threadIdx.x % Cout;
    const int lanes = NB / Cout;
    const int sub = threadIdx.x / Cout;
    float s = 0.f;
    if (sub < lanes)
        for (int p = blockIdx.x * lanes + sub; p < P; p += gridDim.x * lanes) s += float(g[size_t(p) * Cout + co]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < Cout) {
        float t = 0.f;
        for (int k = 0; k < lanes; ++k) t += red[k * Cout + threadIdx.x];
        part[size_t(blockIdx.x) * Cout + threadIdx.x] = t;
    }
}

// mean squared error over the real output channels and its gradient (BF16)
__global__ __launch_bounds__(NB) void k_mse(const float* __restrict__ out, const float* __restrict__ tgt, int P, int Cout,
                                            int Creal, bf16* __restrict__ g, double* __restrict__ part) {
    __shared__ double red[NB];
    const size_t n = size_t(P) * Cout;
    const double inv = 1.0 / (double(P) * double(Creal));
    double s = 0;
    for (size_t e = size_t(blockIdx.x) * NB + threadIdx.x; e < n; e += size_t(gridDim.x) * NB) {
        const int c = int(e % size_t(Cout));
        float d = 0.f;
        if (c < Creal) d = out[e] - tgt[e];
        s += double(d) * double(d);
        g[e] = bf16(float(2.0 * double(d) * inv));
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = NB / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv;
}

// Adam on the FP32 master parameters; padded entries (mask 0) stay zero
__global__ __launch_bounds__(NB) void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m,
                                             float* __restrict__ v, const unsigned char* __restrict__ mask, size_t n,
                                             float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t e = size_t(blockIdx.x) * NB + threadIdx.x;
    if (e >= n || !mask[e]) return;
    const float g = gr[e];
    const float mm = b1 * m[e] + (1.f - b1) * g;
    const float vv = b2 * v[e] + (1.f - b2) * g * g;
    m[e] = mm;
    v[e] = vv;
    w[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 copies of one layer's weights: W [KT][Cout][Cin] and the flipped
// transpose Wt[KT-1-t][ci][co] = W[t][co][ci] used by the input gradient
__global__ __launch_bounds__(NB) void k_wcast(const float* __restrict__ w, int Cout, int Cin, bf16* __restrict__ Wb,
                                              bf16* __restrict__ Wt) {
    const int e = blockIdx.x * NB + threadIdx.x;
    if (e >= KT * Cout * Cin) return;
    const int t = e / (Cout * Cin), r = e - t * Cout * Cin;
    const int co = r / Cin, ci = r - (r / Cin) * Cin;
    const bf16 x = bf16(w[e]);
    Wb[e] = x;
    Wt[(size_t(KT - 1 - t) * Cin + ci) * Cout + co] = x;
}

// sum of |f|^2 over the active modes of up to 4 complex fields (FP64 partials)
template <typename T>
__global__ __launch_bounds__(NB) void k_sumsq(const cx<T>* f0, const cx<T>* f1, const cx<T>* f2, const cx<T>* f3, size_t n,
                                              double* __restrict__ part) {
    __shared__ double red[NB];
    double s = 0;
    const cx<T>* fs[4] = {f0, f1, f2, f3};
    for (int k = 0; k < 4; ++k) {
        if (!fs[k]) continue;
        for (size_t e = size_t(blockIdx.x) * NB + threadIdx.x; e < n; e += size_t(gridDim.x) * NB) {
            const cx<T> v = fs[k][e];
            s += double(v.x) * double(v.x) + double(v.y) * double(v.y);
        }
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = NB / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

struct PackArgs {
    const void* f[4];  // present fields (null: absent), modal layout [ja][z][j1]
    int nf;            // number of present fields
    int A1, A2, N3, Cin;
    const void* kx;
    const void* ky;
    double kxmax, kymax, sinv;
};

// network input: in[p = (ja*A1 + j1)*N3 + z][c]
template <typename T>
__global__ __launch_bounds__(NB) void k_pack(PackArgs A, bf16* __restrict__ in) {
    const size_t P = size_t(A.A1) * A.A2 * A.N3;
    const size_t n = P * A.Cin;
    for (size_t e = size_t(blockIdx.x) * NB + threadIdx.x; e < n; e += size_t(gridDim.x) * NB) {
        const int c = int(e % size_t(A.Cin));
        const size_t p = e / size_t(A.Cin);
        const int z = int(p % size_t(A.N3));
        const size_t mode = p / size_t(A.N3);
        const int j1 = int(mode % size_t(A.A1)), ja = int(mode / size_t(A.A1));
        float v = 0.f;
        if (c < 2 * A.nf) {
            const cx<T> x = static_cast<const cx<T>*>(A.f[c >> 1])[sidx(ja, z, j1, A.N3, A.A1)];
            v = float(double((c & 1) ? x.y : x.x) * A.sinv);
        } else if (c == 2 * A.nf) {
            v = float(double(static_cast<const T*>(A.kx)[j1]) / A.kxmax);
        } else if (c == 2 * A.nf + 1 && A.kymax > 0) {
            v = float(double(static_cast<const T*>(A.ky)[ja]) / A.kymax);
        }
        in[e] = bf16(v);
    }
}

// training target: q / RMS(q) - w / RMS(w) on the real output channels (the
// direction of A^-1 w at the scale of w, minus w; W.sinv and Q.sinv hold the
// two inverse RMS values)
template <typename T>
__global__ __launch_bounds__(NB) void k_pack_target(PackArgs W, PackArgs Q, int Cout, float* __restrict__ tgt) {
    const size_t P = size_t(W.A1) * W.A2 * W.N3;
    const size_t n = P * Cout;
    for (size_t e = size_t(blockIdx.x) * NB + threadIdx.x; e < n; e += size_t(gridDim.x) * NB) {
        const int c = int(e % size_t(Cout));
        const size_t p = e / size_t(Cout);
        const int z = int(p % size_t(W.N3));
        const size_t mode = p / size_t(W.N3);
        const int j1 = int(mode % size_t(W.A1)), ja = int(mode / size_t(W.A1));
        float v = 0.f;
        if (c < 2 * W.nf) {
            const size_t o = sidx(ja, z, j1, W.N3, W.A1);
            const cx<T> w = static_cast<const cx<T>*>(W.f[c >> 1])[o];
            const cx<T> q = static_cast<const cx<T>*>(Q.f[c >> 1])[o];
            v = float((c & 1) ? double(q.y) * Q.sinv - double(w.y) * W.sinv : double(q.x) * Q.sinv - double(w.x) * W.sinv);
        }
        tgt[e] = v;
    }
}

// out_f = in_f + s * (y[p][2f] + i y[p][2f+1])
template <typename T>
__global__ __launch_bounds__(NB) void k_unpack(PackArgs A, const float* __restrict__ y, int Cout, double s, cx<T>* o0,
                                               cx<T>* o1, cx<T>* o2, cx<T>* o3) {
    const size_t P = size_t(A.A1) * A.A2 * A.N3;
    cx<T>* outs[4] = {o0, o1, o2, o3};
    for (size_t p = size_t(blockIdx.x) * NB + threadIdx.x; p < P; p += size_t(gridDim.x) * NB) {
        const int z = int(p % size_t(A.N3));
        const size_t mode = p / size_t(A.N3);
        const int j1 = int(mode % size_t(A.A1)), ja = int(mode / size_t(A.A1));
        const size_t o = sidx(ja, z, j1, A.N3, A.A1);
        for (int k = 0; k < A.nf; ++k) {
            const cx<T> x = static_cast<const cx<T>*>(A.f[k])[o];
            cx<T> r;
            r.x = T(double(x.x) + s * double(y[p * Cout + 2 * k]));
            r.y = T(double(x.y) + s * double(y[p * Cout + 2 * k + 1]));
            outs[k][o] = r;
        }
    }
}

// gather training profiles idx[] into a contiguous batch
__global__ __launch_bounds__(NB) void k_gather(const bf16* __restrict__ src_in, const float* __restrict__ src_t,
                                               const int* __restrict__ idx, int nb, int Z, int Cin, int Cout,
                                               bf16* __restrict__ in, float* __restrict__ tgt) {
    const size_t ni = size_t(nb) * Z * Cin, nt = size_t(nb) * Z * Cout;
    for (size_t e = size_t(blockIdx.x) * NB + threadIdx.x; e < ni + nt; e += size_t(gridDim.x) * NB) {
        if (e < ni) {
            const size_t per = size_t(Z) * Cin;
            in[e] = src_in[size_t(idx[e / per]) * per + e % per];
        } else {
            const size_t f = e - ni, per = size_t(Z) * Cout;
            tgt[f] = src_t[size_t(idx[f / per]) * per + f % per];
        }
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static unsigned grid_for(size_t n, unsigned cap = 4096) {
    size_t b = (n + NB - 1) / NB;
    if (b > cap) b = cap;
    return unsigned(b < 1 ? 1 : b);
}

template <typename P>
static P* dalloc(size_t n) {
    void* p = nullptr;
    SF_HIP(hipMalloc(&p, (n ? n : 1) * sizeof(P)));
    SF_HIP(hipMemset(p, 0, (n ? n : 1) * sizeof(P)));
    return static_cast<P*>(p);
}

struct Net {
    sf_ctx* ctx = nullptr;
    CtxInfo info{};
    int nf = 0, L = 4, H = 64, Z = 0;
    int Creal_in = 0, Creal_out = 0;
    std::vector<int> C;  // padded channels: C[0] input .. C[L] output
    std::vector<int> dil;  // dilation of each layer's taps (1: neighbours)
    std::vector<size_t> woff, boff;
    size_t nparam = 0;
    float *w = nullptr, *gr = nullptr, *am = nullptr, *av = nullptr;
    unsigned char* pmask = nullptr;
    std::vector<bf16*> Wb, Wt;
    int adam_t = 0;
    std::mt19937_64 rng;
    // training set (profiles)
    bf16* set_in = nullptr;
    float* set_t = nullptr;
    int set_n = 0, set_cap = 0;
    // work buffers for up to cap_pos positions
    size_t cap_pos = 0;
    std::vector<bf16*> act;     // act[l]: input of layer l, [P][C[l]]
    std::vector<float*> pre;    // pre[l]: output of layer l (FP32), [P][C[l+1]]
    std::vector<bf16*> gbuf;    // gbuf[l]: gradient w.r.t. the output of layer l (BF16)
    float* tgt = nullptr;
    float* part = nullptr;
    size_t part_n = 0;
    double* dpart = nullptr;
    int* didx = nullptr;
    int didx_n = 0;
    static constexpr int DW_BLOCKS = 240;

    ~Net() {
        auto F = [](void* p) {
            if (p) (void)hipFree(p);
        };
        F(w);
        F(gr);
        F(am);
        F(av);
        F(pmask);
        for (auto* p : Wb) F(p);
        for (auto* p : Wt) F(p);
        F(set_in);
        F(set_t);
        free_work();
        F(dpart);
        F(didx);
    }
    void free_work() {
        for (auto* p : act) (void)hipFree(p);
        for (auto* p : pre) (void)hipFree(p);
        for (auto* p : gbuf) (void)hipFree(p);
        if (tgt) (void)hipFree(tgt);
        if (part) (void)hipFree(part);
        act.clear();
        pre.clear();
        gbuf.clear();
        tgt = nullptr;
        part = nullptr;
        cap_pos = 0;
    }

    hipStream_t st() const { return info.stream; }
    size_t profiles() const { return size_t(info.A1) * info.A2; }

    void init(sf_ctx* c, int hidden, int layers, unsigned long long seed) {
        ctx = c;
        info = sf_ctx_info(c);
        SF_REQUIRE(layers >= 2 && layers <= 8, SF_ERR_ARG, "layers must be in [2, 8]");
        SF_REQUIRE(hidden == 32 || hidden == 64, SF_ERR_ARG, "hidden channels must be 32 or 64");
        L = layers;
        H = hidden;
        Z = info.N3;
        nf = info.three ? 4 : 3;
        Creal_in = 2 * nf + 2;
        Creal_out = 2 * nf;
        C.assign(L + 1, H);
        C[0] = 32;
        C[L] = 32;
        dil.assign(L, 1);
        size_t off = 0;
        woff.resize(L);
        boff.resize(L);
        for (int l = 0; l < L; ++l) {
            woff[l] = off;
            off += size_t(KT) * C[l + 1] * C[l];
            boff[l] = off;
            off += size_t(C[l + 1]);
        }
        nparam = off;
        std::vector<float> hw(nparam, 0.f);
        std::vector<unsigned char> hm(nparam, 0);
        rng.seed(seed);
        std::normal_distribution<double> nd(0.0, 1.0);
        for (int l = 0; l < L; ++l) {
            const int cin = (l == 0) ? Creal_in : C[l];
            const int cout = (l == L - 1) ? Creal_out : C[l + 1];
            const double sd = std::sqrt(2.0 / (double(KT) * cin));  // He initialisation
            for (int t = 0; t < KT; ++t)
                for (int co = 0; co < cout; ++co)
                    for (int ci = 0; ci < cin; ++ci) {
                        const size_t e = woff[l] + (size_t(t) * C[l + 1] + co) * C[l] + ci;
                        hm[e] = 1;
                        // the last layer starts at zero: M^-1 = I until trained
                        hw[e] = (l == L - 1) ? 0.f : float(sd * nd(rng));
                    }
            for (int co = 0; co < cout; ++co) hm[boff[l] + co] = 1;
        }
        w = dalloc<float>(nparam);
        gr = dalloc<float>(nparam);
        am = dalloc<float>(nparam);
        av = dalloc<float>(nparam);
        pmask = dalloc<unsigned char>(nparam);
        SF_HIP(hipMemcpy(w, hw.data(), nparam * sizeof(float), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(pmask, hm.data(), nparam, hipMemcpyHostToDevice));
        for (int l = 0; l < L; ++l) {
            Wb.push_back(dalloc<bf16>(size_t(KT) * C[l + 1] * C[l]));
            Wt.push_back(dalloc<bf16>(size_t(KT) * C[l + 1] * C[l]));
        }
        dpart = dalloc<double>(4096);
        cast_weights();
        SF_HIP(hipStreamSynchronize(st()));
    }

    void cast_weights() {
        for (int l = 0; l < L; ++l) {
            const size_t n = size_t(KT) * C[l + 1] * C[l];
            hipLaunchKernelGGL(k_wcast, dim3(grid_for(n, 1u << 20)), dim3(NB), 0, st(), w + woff[l], C[l + 1], C[l],
                               Wb[l], Wt[l]);
        }
        SF_HIP(hipGetLastError());
    }

    void ensure_work(size_t P) {
        if (P <= cap_pos) return;
        free_work();
        cap_pos = P;
        for (int l = 0; l <= L; ++l) act.push_back(dalloc<bf16>(P * C[l]));
        for (int l = 0; l < L; ++l) pre.push_back(dalloc<float>(P * C[l + 1]));
        for (int l = 0; l < L; ++l) gbuf.push_back(dalloc<bf16>(P * C[l + 1]));
        tgt = dalloc<float>(P * C[L]);
        part_n = size_t(DW_BLOCKS) * KT * H * H;
        part = dalloc<float>(part_n);
    }

    void conv(int NT, const bf16* in, int Cin, const bf16* W, const float* bias, int P, float* out_f, bf16* out_b,
              int relu, const float* mask, int d) {
        const dim3 grid((P + 63) / 64);
        switch (NT) {
            case 2:
                hipLaunchKernelGGL(k_conv<2>, grid, dim3(NB), 0, st(), in, Cin, W, bias, P, Z, out_f, out_b, relu, mask,
                                   d);
                break;
            case 4:
                hipLaunchKernelGGL(k_conv<4>, grid, dim3(NB), 0, st(), in, Cin, W, bias, P, Z, out_f, out_b, relu, mask,
                                   d);
                break;
            default:
                throw Error(SF_ERR_ARG, "unsupported channel count");
        }
        SF_HIP(hipGetLastError());
    }

    // act[0] holds the input of P positions; the output is pre[L-1]
    void forward(int P) {
        for (int l = 0; l < L; ++l) {
            const bool last = (l == L - 1);
            conv(C[l + 1] / 16, act[l], C[l], Wb[l], w + boff[l], P, pre[l], last ? nullptr : act[l + 1], last ? 0 : 1,
                 nullptr, dil[l]);
        }
    }

    void dw(int l, int P) {
        const int NTO = C[l + 1] / 16, NTI = C[l] / 16;
        const dim3 grid(DW_BLOCKS);
        const bf16* g = gbuf[l];
        const bf16* a = act[l];
#define SF_DW(o, i)                                                                               \
    if (NTO == o && NTI == i) {                                                                   \
        hipLaunchKernelGGL((k_conv_dw<o, i>), grid, dim3(NB), 0, st(), g, a, P, Z, part, dil[l]); \
    } else
        SF_DW(4, 2) SF_DW(4, 4) SF_DW(2, 4) SF_DW(2, 2) { throw Error(SF_ERR_ARG, "unsupported layer shape"); }
#undef SF_DW
        SF_HIP(hipGetLastError());
        const int n = KT * C[l + 1] * C[l];
        hipLaunchKernelGGL(k_sum_parts, dim3((n + NB - 1) / NB), dim3(NB), 0, st(), part, DW_BLOCKS, n, gr + woff[l]);
        // bias
        hipLaunchKernelGGL(k_bias_part, dim3(DW_BLOCKS), dim3(NB), 0, st(), g, P, C[l + 1], part);
        hipLaunchKernelGGL(k_sum_parts, dim3((C[l + 1] + NB - 1) / NB), dim3(NB), 0, st(), part, DW_BLOCKS, C[l + 1],
                           gr + boff[l]);
        SF_HIP(hipGetLastError());
    }

    // gradient of the MSE (pre[L-1] against tgt) with respect to every
    // parameter, into gr; returns the loss
    double backward(int P) {
        const unsigned gb = grid_for(size_t(P) * C[L], 1024);
        hipLaunchKernelGGL(k_mse, dim3(gb), dim3(NB), 0, st(), pre[L - 1], tgt, P, C[L], Creal_out, gbuf[L - 1], dpart);
        SF_HIP(hipGetLastError());
        for (int l = L - 1; l >= 0; --l) {
            dw(l, P);
            if (l > 0)  // gradient w.r.t. the ReLU input of layer l-1
                conv(C[l] / 16, gbuf[l], C[l + 1], Wt[l], nullptr, P, nullptr, gbuf[l - 1], 0, pre[l - 1], dil[l]);
        }
        std::vector<double> h(gb);
        SF_HIP(hipMemcpyAsync(h.data(), dpart, gb * sizeof(double), hipMemcpyDeviceToHost, st()));
        SF_HIP(hipStreamSynchronize(st()));
        double loss = 0;
        for (double x : h) loss += x;
        return loss;
    }

    void adam(double lr) {
        ++adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = float(1.0 - std::pow(0.9, adam_t)), c2 = float(1.0 - std::pow(0.999, adam_t));
        hipLaunchKernelGGL(k_adam, dim3(grid_for(nparam, 1u << 20)), dim3(NB), 0, st(), w, gr, am, av, pmask, nparam,
                           float(lr), b1, b2, 1e-8f, c1, c2);
        SF_HIP(hipGetLastError());
        cast_weights();
    }

    template <typename T>
    PackArgs pack_args(sf_field* const* f, double sinv) const {
        PackArgs A{};
        int k = 0;
        for (int i = 0; i < 4; ++i) {
            if (!info.three && i == 1) continue;
            SF_REQUIRE(f[i] != nullptr, SF_ERR_ARG, "missing field");
            A.f[k++] = f[i]->d;
        }
        A.nf = k;
        A.A1 = info.A1;
        A.A2 = info.A2;
        A.N3 = info.N3;
        A.Cin = C[0];
        A.kx = info.kx;
        A.ky = info.ky;
        A.kxmax = kmax<T>(info.kx, info.A1);
        A.kymax = info.three ? kmax<T>(info.ky, info.A2) : 0.0;
        A.sinv = sinv;
        return A;
    }

    template <typename T>
    double kmax(const void* k, int n) const {
        std::vector<T> h(n);
        SF_HIP(hipMemcpy(h.data(), k, n * sizeof(T), hipMemcpyDeviceToHost));
        double m = 0;
        for (T x : h) m = std::max(m, std::fabs(double(x)));
        return m > 0 ? m : 1.0;
    }

    template <typename T>
    double rms(sf_field* const* f) {
        const size_t n = size_t(info.A1) * info.A2 * info.N3;
        const cx<T>* p[4] = {nullptr, nullptr, nullptr, nullptr};
        int nf_ = 0;
        for (int i = 0; i < 4; ++i)
            if (f[i] && (info.three || i != 1)) {
                p[i] = static_cast<const cx<T>*>(f[i]->d);
                ++nf_;
            }
        const unsigned g = grid_for(n, 1024);
        hipLaunchKernelGGL((k_sumsq<T>), dim3(g), dim3(NB), 0, st(), p[0], p[1], p[2], p[3], n, dpart);
        SF_HIP(hipGetLastError());
        std::vector<double> h(g);
        SF_HIP(hipMemcpyAsync(h.data(), dpart, g * sizeof(double), hipMemcpyDeviceToHost, st()));
        SF_HIP(hipStreamSynchronize(st()));
        double s = 0;
        for (double x : h) s += x;
        return std::sqrt(s / double(n * size_t(nf_)));
    }

    template <typename T>
    void apply(sf_field* const* in, sf_field* const* out) {
        const int P = int(profiles()) * Z;
        ensure_work(size_t(P));
        const double s = rms<T>(in);
        PackArgs A = pack_args<T>(in, s > 0 ? 1.0 / s : 0.0);
        hipLaunchKernelGGL((k_pack<T>), dim3(grid_for(size_t(P) * C[0])), dim3(NB), 0, st(), A, act[0]);
        SF_HIP(hipGetLastError());
        forward(P);
        cx<T>* o[4] = {nullptr, nullptr, nullptr, nullptr};
        int k = 0;
        for (int i = 0; i < 4; ++i) {
            if (!info.three && i == 1) continue;
            SF_REQUIRE(out[i] != nullptr, SF_ERR_ARG, "missing output field");
            o[k++] = static_cast<cx<T>*>(out[i]->d);
        }
        hipLaunchKernelGGL((k_unpack<T>), dim3(grid_for(size_t(P))), dim3(NB), 0, st(), A, pre[L - 1], C[L], s, o[0],
                           o[1], o[2], o[3]);
        SF_HIP(hipGetLastError());
    }

    void grow_set(int add) {
        if (set_n + add <= set_cap) return;
        int cap = std::max(set_cap * 2, set_n + add);
        bf16* ni = dalloc<bf16>(size_t(cap) * Z * C[0]);
        float* nt = dalloc<float>(size_t(cap) * Z * C[L]);
        if (set_n) {
            SF_HIP(hipMemcpyAsync(ni, set_in, size_t(set_n) * Z * C[0] * sizeof(bf16), hipMemcpyDeviceToDevice, st()));
            SF_HIP(hipMemcpyAsync(nt, set_t, size_t(set_n) * Z * C[L] * sizeof(float), hipMemcpyDeviceToDevice, st()));
            SF_HIP(hipStreamSynchronize(st()));
        }
        if (set_in) (void)hipFree(set_in);
        if (set_t) (void)hipFree(set_t);
        set_in = ni;
        set_t = nt;
        set_cap = cap;
    }

    // a pair (w, q) with A q = w; every mode's profile becomes one sample
    template <typename T>
    void add_pair(sf_field* const* wf, sf_field* const* qf) {
        const int np = int(profiles());
        grow_set(np);
        const double s = rms<T>(wf), sq = rms<T>(qf);
        if (!(s > 0) || !(sq > 0)) return;
        PackArgs W = pack_args<T>(wf, 1.0 / s);
        PackArgs Q = pack_args<T>(qf, 1.0 / sq);
        bf16* di = set_in + size_t(set_n) * Z * C[0];
        float* dt = set_t + size_t(set_n) * Z * C[L];
        const size_t P = size_t(np) * Z;
        hipLaunchKernelGGL((k_pack<T>), dim3(grid_for(P * C[0])), dim3(NB), 0, st(), W, di);
        hipLaunchKernelGGL((k_pack_target<T>), dim3(grid_for(P * C[L])), dim3(NB), 0, st(), W, Q, C[L], dt);
        SF_HIP(hipGetLastError());
        set_n += np;
    }

    // mini-batch Adam over the stored profiles
    void train(int epochs, int batch, double lr, double* first, double* last) {
        SF_REQUIRE(set_n > 0, SF_ERR_STATE, "no training pairs");
        batch = std::max(1, std::min(batch, set_n));
        ensure_work(size_t(batch) * Z);
        if (didx_n < batch) {
            if (didx) (void)hipFree(didx);
            didx = dalloc<int>(batch);
            didx_n = batch;
        }
        std::vector<int> order(set_n);
        for (int i = 0; i < set_n; ++i) order[i] = i;
        double l0 = -1, l1 = 0;
        for (int ep = 0; ep < epochs; ++ep) {
            std::shuffle(order.begin(), order.end(), rng);
            double sum = 0;
            int nbt = 0;
            for (int b0 = 0; b0 + batch <= set_n; b0 += batch) {
                SF_HIP(hipMemcpyAsync(didx, order.data() + b0, batch * sizeof(int), hipMemcpyHostToDevice, st()));
                const size_t P = size_t(batch) * Z;
                hipLaunchKernelGGL(k_gather, dim3(grid_for(P * (C[0] + C[L]))), dim3(NB), 0, st(), set_in, set_t, didx,
                                   batch, Z, C[0], C[L], act[0], tgt);
                SF_HIP(hipGetLastError());
                forward(int(P));
                const double loss = backward(int(P));
                adam(lr);
                sum += loss;
                ++nbt;
            }
            const double mean = nbt ? sum / nbt : 0.0;
            if (l0 < 0) l0 = mean;
            l1 = mean;
        }
        if (first) *first = l0;
        if (last) *last = l1;
    }
};

}  // namespace net
}  // namespace sf

struct sf_net {
    int precision;
    std::string err;
    sf::net::Net impl;
};

using sf::Error;

template <typename Fn>
static int net_call(sf_net* n, Fn fn) {
    try {
        fn(n->impl);
        n->err.clear();
        return SF_OK;
    } catch (const Error& e) {
        n->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        n->err = e.what();
        return SF_ERR_STATE;
    }
}

static thread_local std::string g_net_error;

extern "C" {

int sf_net_create(sf_ctx* ctx, int hidden, int layers, unsigned long long seed, sf_net** out) {
    if (!ctx || !out) return SF_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<sf_net> n(new sf_net());
    n->precision = ctx->precision;
    try {
        n->impl.init(ctx, hidden, layers, seed);
    } catch (const Error& e) {
        g_net_error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_net_error = e.what();
        return SF_ERR_STATE;
    }
    *out = n.release();
    return SF_OK;
}

int sf_net_destroy(sf_net* n) {
    delete n;
    return SF_OK;
}

const char* sf_net_last_error(const sf_net* n) { return n ? n->err.c_str() : g_net_error.c_str(); }

int sf_net_apply(sf_net* n, sf_field* const* in, sf_field* const* out) {
    if (!n || !in || !out) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) {
        if (n->precision == 4) N.apply<float>(in, out);
        else N.apply<double>(in, out);
    });
}

int sf_net_add_pair(sf_net* n, sf_field* const* w, sf_field* const* q) {
    if (!n || !w || !q) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) {
        if (n->precision == 4) N.add_pair<float>(w, q);
        else N.add_pair<double>(w, q);
    });
}

int sf_net_clear_pairs(sf_net* n) {
    if (!n) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) { N.set_n = 0; });
}

int sf_net_samples(const sf_net* n, int* count) {
    if (!n || !count) return SF_ERR_ARG;
    *count = n->impl.set_n;
    return SF_OK;
}

int sf_net_train(sf_net* n, int epochs, int batch, double lr, double* loss_first, double* loss_last) {
    if (!n || epochs < 0 || batch <= 0 || !(lr > 0)) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) { N.train(epochs, batch, lr, loss_first, loss_last); });
}

int sf_net_param_count(const sf_net* n, size_t* count) {
    if (!n || !count) return SF_ERR_ARG;
    *count = n->impl.nparam;
    return SF_OK;
}

int sf_net_get_params(sf_net* n, float* host) {
    if (!n || !host) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) {
        SF_HIP(hipMemcpy(host, N.w, N.nparam * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int sf_net_set_params(sf_net* n, const float* host) {
    if (!n || !host) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) {
        SF_HIP(hipMemcpy(N.w, host, N.nparam * sizeof(float), hipMemcpyHostToDevice));
        N.cast_weights();
        SF_HIP(hipStreamSynchronize(N.st()));
    });
}

int sf_net_set_dilations(sf_net* n, const int* dilations) {
    if (!n || !dilations) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) {
        for (int l = 0; l < N.L; ++l)
            SF_REQUIRE(dilations[l] >= 1 && dilations[l] <= sf::net::NET_DMAX, SF_ERR_ARG,
                       "dilations must be in [1, 8]");
        for (int l = 0; l < N.L; ++l) N.dil[l] = dilations[l];
    });
}

int sf_net_geometry(const sf_net* n, int* layers, int* channels, int* z) {
    if (!n || !layers || !channels || !z) return SF_ERR_ARG;
    *layers = n->impl.L;
    for (int l = 0; l <= n->impl.L; ++l) channels[l] = n->impl.C[l];
    *z = n->impl.Z;
    return SF_OK;
}

int sf_net_forward_raw(sf_net* n, const float* in, int nprof, float* out) {
    if (!n || !in || !out || nprof <= 0) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) {
        const size_t P = size_t(nprof) * N.Z;
        N.ensure_work(P);
        std::vector<sf::net::bf16> h(P * N.C[0]);
        for (size_t i = 0; i < h.size(); ++i) h[i] = sf::net::bf16(in[i]);
        SF_HIP(hipMemcpy(N.act[0], h.data(), h.size() * sizeof(sf::net::bf16), hipMemcpyHostToDevice));
        N.forward(int(P));
        SF_HIP(hipMemcpyAsync(out, N.pre[N.L - 1], P * N.C[N.L] * sizeof(float), hipMemcpyDeviceToHost, N.st()));
        SF_HIP(hipStreamSynchronize(N.st()));
    });
}

int sf_net_grad_raw(sf_net* n, const float* in, const float* target, int nprof, float* grad, double* loss) {
    if (!n || !in || !target || !grad || !loss || nprof <= 0) return SF_ERR_ARG;
    return net_call(n, [&](sf::net::Net& N) {
        const size_t P = size_t(nprof) * N.Z;
        N.ensure_work(P);
        std::vector<sf::net::bf16> h(P * N.C[0]);
        for (size_t i = 0; i < h.size(); ++i) h[i] = sf::net::bf16(in[i]);
        SF_HIP(hipMemcpy(N.act[0], h.data(), h.size() * sizeof(sf::net::bf16), hipMemcpyHostToDevice));
        SF_HIP(hipMemcpy(N.tgt, target, P * N.C[N.L] * sizeof(float), hipMemcpyHostToDevice));
        N.forward(int(P));
        *loss = N.backward(int(P));
        SF_HIP(hipMemcpy(grad, N.gr, N.nparam * sizeof(float), hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
