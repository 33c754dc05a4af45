// sf_common.h -- shared types and small device helpers for the gfx950 path.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sf {

template <typename T>
struct CxT;
template <>
struct CxT<float> {
    using t = float2;
};
template <>
struct CxT<double> {
    using t = double2;
};
template <typename T>
using cx = typename CxT<T>::t;

template <typename T>
__host__ __device__ __forceinline__ cx<T> mkc(T r, T i) {
    cx<T> c;
    c.x = r;
    c.y = i;
    return c;
}
template <typename V>
__host__ __device__ __forceinline__ V cadd(V a, V b) {
    V r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <typename V>
__host__ __device__ __forceinline__ V csub(V a, V b) {
    V r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}
template <typename V>
__host__ __device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
// a * conj(b)
template <typename V>
__host__ __device__ __forceinline__ V cmulc(V a, V b) {
    V r;
    r.x = a.x * b.x + a.y * b.y;
    r.y = a.y * b.x - a.x * b.y;
    return r;
}
template <typename V>
__host__ __device__ __forceinline__ V cconj(V a) {
    V r;
    r.x = a.x;
    r.y = -a.y;
    return r;
}
template <typename V, typename T>
__host__ __device__ __forceinline__ V cscale(T s, V a) {
    V r;
    r.x = s * a.x;
    r.y = s * a.y;
    return r;
}
// (i*k) * a  for real k
template <typename V, typename T>
__host__ __device__ __forceinline__ V cik(T k, V a) {
    V r;
    r.x = -k * a.y;
    r.y = k * a.x;
    return r;
}
// a + s*b
template <typename V, typename T>
__host__ __device__ __forceinline__ V caxpy(T s, V b, V a) {
    V r;
    r.x = a.x + s * b.x;
    r.y = a.y + s * b.y;
    return r;
}
// Workgroup barrier for LDS traffic only.  __syncthreads() also waits for
// every outstanding global load and store of the wave (vmcnt(0)), which would
// expose the full HBM latency at each barrier of a level-marching loop; here
// global loads and stores stay in flight and the compiler's own waits before
// the first use of a loaded register remain in place.  The "memory" clobber
// keeps the compiler from moving LDS accesses across the barrier.
__device__ __forceinline__ void sync_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <typename V>
__host__ __device__ __forceinline__ V czero() {
    V r;
    r.x = 0;
    r.y = 0;
    return r;
}

// ---- raw buffer access ------------------------------------------------------
// Loads and stores through a buffer resource (base and size in SGPRs) with a
// 32-bit byte offset per lane: no 64-bit address arithmetic per access, and an
// offset at or beyond the size reads zero and drops a store, so masked
// accesses need no branches.  Every array addressed this way is below 2 GiB.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int BUF_OOB = int(0x80000000u);  // an offset outside every buffer

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, int(unsigned(bytes)), 0x00020000);
}
template <typename V>
__device__ __forceinline__ V bld(rsrc_t r, int voff, int soff = 0) {
    if constexpr (sizeof(V) == 4) {
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
    } else if constexpr (sizeof(V) == 8) {
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    } else {
        static_assert(sizeof(V) == 16, "buffer access of 4, 8 or 16 bytes");
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
    }
}
template <typename V>
__device__ __forceinline__ void bst(V v, rsrc_t r, int voff, int soff = 0) {
    if constexpr (sizeof(V) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
    } else if constexpr (sizeof(V) == 8) {
        using U2 = unsigned __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, voff, soff, 0);
    } else {
        static_assert(sizeof(V) == 16, "buffer access of 4, 8 or 16 bytes");
        using U4 = unsigned __attribute__((ext_vector_type(4)));
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U4, v), r, voff, soff, 0);
    }
}

// 1/x: the hardware reciprocal (1 ulp) in FP32, a correctly rounded division
// in FP64
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

// Per-z coefficient table rows (filled by the context set-up in sf_api.hip).
enum ZCoef {
    ZC_DZZN_LO = 0,
    ZC_DZZN_DI,
    ZC_DZZN_UP,
    ZC_DZZD_LO,
    ZC_DZZD_DI,
    ZC_DZZD_UP,
    ZC_DZND,   // 1/DY_j   rows 2..N-1  (Neumann -> Dirichlet derivative)
    ZC_DZDN,   // 1/DYF_j  rows 1..N-1  (Dirichlet -> Neumann derivative)
    ZC_IFULL_LO,
    ZC_IFULL_DI,
    ZC_ITIL_LO,
    ZC_ITIL_DI,
    ZC_UBAR,   // background velocity U_(z)
    ZC_S,      // background stratification dB/dz at Neumann points
    ZC_WN,     // InnerProd weight, Neumann
    ZC_WD,     // InnerProd weight, Dirichlet
    ZC_COUNT
};

}  // namespace sf
