// sf_kernels.h -- kernel argument blocks and host launchers.
#pragma once
#include "sf_common.h"

namespace sf {


// Fused per-mode stage kernel (one wavefront per Fourier mode).
template <typename T>
struct ModalArgs {
    int N3, N3p, A1, A2, nmodes, threeD;
    const T* zc;  // [ZC_COUNT][N3p]
    const T *kx, *ky, *kx2, *ky2;
    cx<T> *u1, *u2, *u3, *b, *p;
    const cx<T> *ro1, *ro2, *ro3, *roB;  // explicit terms of the previous stage (null: not read)
    cx<T> *rn1, *rn2, *rn3, *rnB;        // explicit terms of this stage (null: not stored)
    const cx<T> *n1, *n2, *n3, *nB;      // nonlinear terms from the transform path
    T hz, mh, hb, cv, cb, Ri, pmult;
    int mode;         // 0: full stage, 1: projection only (RemoveDivergence)
    int skip_linear;  // 1: n1..nB are the complete explicit terms (BuildRHSAdjoint)
    T* ubar;          // non-null: EvolveBackground, U_ (the ZC_UBAR row of zc) diffuses in place
};

// Position of element (row, z, j1) of every spectral array: row is the active
// y-mode ja of a modal field ([ja][z][j1]) or the spanwise point n2 of an
// x-spectral intermediate ([n2][z][j1]).  j1 is fastest, so the x transforms
// read and write whole contiguous rows of A1 values, the y transforms read and
// write runs of a j1 tile, and the modal kernels stage a tile of consecutive
// j1 modes through LDS.
__host__ __device__ inline size_t sidx(int row, int z, int j1, int N3, int A1) {
    return (size_t(row) * size_t(N3) + size_t(z)) * size_t(A1) + size_t(j1);
}

// Fused inverse-x / products / forward-x kernel.
template <typename T>
struct XArgs {
    int N1, N3, N3p, A1, zch;            // zch: z levels per block
    int N2;                              // 3D: y points (rows of the outputs)
    int tiled;                           // 3D: outputs in tiles of 16 modes (ytile_idx)
    const cx<T>* in[4];                  // u1, u2, u3, b spectra
    cx<T>* out[8];                       // 3D: A1,A2,A3,AB,B1,B2,B3,BB ; 2D: Nl1,-,Nl3,NlB
    const T* zc;
    const T* kx;
    const cx<T>* tw;
    T scale;
    int linear;                          // 1: linearised products (BuildRHSLinear)
    const cx<T>* base[4];                // linear: u1t, u2t, u3t, bt spectra (same layout as in)
};

// y-direction transforms (3D): per (z, tile of consecutive j1).
template <typename T>
struct YArgs {
    int N3, A1, A2, nf;
    const cx<T>* in[8];
    cx<T>* out[8];
    const T* ky;      // A2 active wavenumbers (forward: i*ky*B term)
    const cx<T>* tw;
    T scale;
    int paired;       // forward: out[r] = scale*(F(in[2r]) + i ky F(in[2r+1]))
    int tiled;        // paired forward: inputs in tiles of 16 modes (ytile_idx)
};

// Offset of (n2, z, j1) in an x-spectral field stored in tiles of 16 modes,
// [z][j1/16][n2][j1%16] (A1 a multiple of 16): what the 3D x-pass writes and
// the paired forward y transform reads when the grid allows it
// (yfwd_tiles), so that a y-transform block's 16-mode tile is one
// contiguous run of N2*16 values.
__host__ __device__ inline size_t ytile_idx(int n2, int z, int j1, int N2, int A1) {
    return ((size_t(z) * size_t(A1 >> 4) + size_t(j1 >> 4)) * size_t(N2) + size_t(n2)) * 16 + size_t(j1 & 15);
}

// Single-field x transforms between nodal and x-spectral layouts.
template <typename T>
struct XTArgs {
    int N3, A1, nzt;
    const cx<T>* spec_in;
    cx<T>* spec_out;
    const T* nodal_in;
    T* nodal_out;
    const cx<T>* tw;
    T scale;
};

template <typename T>
hipError_t launch_modal(const ModalArgs<T>& a, hipStream_t s);
template <typename T>
hipError_t launch_xpass(const XArgs<T>& a, int N1, int N2, bool threeD, hipStream_t s);
// many independent 2D runs per launch: args/start live in device memory,
// start[r] = first work item of run r (x-pass: z-blocks; modal: tiles of
// modes), total = start[nruns]; grid-stride loop over `grid` blocks.  The
// _batch launches take runs of one N1 (x-pass) or one rows-per-lane M
// (modal); the _mixed launches take runs of any N1 / N3, the modal one with
// mb (4 or 8) modes per block for every run.
template <typename T>
hipError_t launch_modal_batch(const ModalArgs<T>* args, const int* start, int nruns, int total, int M, int grid,
                              hipStream_t s);
template <typename T>
hipError_t launch_xpass_batch(const XArgs<T>* args, const int* start, int nruns, int total, int N1, int grid,
                              hipStream_t s);
// (mmax: the largest rows per lane of the runs; nmax: the largest N1)
template <typename T>
hipError_t launch_modal_mixed(const ModalArgs<T>* args, const int* start, int nruns, int total, int mb, int mmax,
                              int grid, hipStream_t s);
// (r1f: first radix of the FFT plans, 0 for the defaults or 8; each run's tw
// must be the table of its N1 under that plan, fft_twiddles_plan)
template <typename T>
hipError_t launch_xpass_mixed(const XArgs<T>* args, const int* start, int nruns, int total, int nmax, int r1f,
                              int grid, hipStream_t s);
template <typename T>
hipError_t launch_yinv(const YArgs<T>& a, int N2, hipStream_t s);
template <typename T>
hipError_t launch_yfwd(const YArgs<T>& a, int N2, hipStream_t s);
template <typename T>
hipError_t launch_x_c2r(const XTArgs<T>& a, int N1, int N2, hipStream_t s);
template <typename T>
hipError_t launch_x_r2c(const XTArgs<T>& a, int N1, int N2, hipStream_t s);
template <typename T>
hipError_t launch_dot(const cx<T>* a, const cx<T>* b, int A1, int A2, int N3, const T* w, double* partial,
                      double* out, int slot, hipStream_t s);
template <typename T>
hipError_t launch_maxabs(const T* f, size_t n, int N3, const T* add, double* partial, double* out, int slot,
                         hipStream_t s);
template <typename T>
hipError_t launch_axpby(cx<T>* y, T a, const cx<T>* x, T b, size_t n, hipStream_t s);
template <typename T>
hipError_t launch_ends(cx<T>* f, int A1, int A2, int N3, int bc_op, hipStream_t s);
template <typename T>
hipError_t launch_dmul(cx<T>* out, const cx<T>* in, int A1, int A2, int N3, const T* kx, const T* ky, int axis,
                       const T* zc, int N3p, hipStream_t s);

// diffusion preconditioner of the Newton-Krylov solves (one wavefront per mode)
template <typename T>
struct DiffPrecArgs {
    int N3, N3p, A1, A2, dirichlet;
    const T* zc;
    const T *kx, *kx2, *ky2;  // kx, -kx^2, -ky^2 of the active modes
    const cx<T>* in;
    cx<T>* out;
    T tnu, delta, k2max;
    T tadv;  // T of the advection term -i kx U(z) (0: diffusion only)
};
template <typename T>
hipError_t launch_diffprec(const DiffPrecArgs<T>& p, hipStream_t s);
// the advected variant runs ADVPREC_MODES modes per block, each with N3 complex
// values of LDS, within ADVPREC_LDS bytes
constexpr int ADVPREC_MODES = 16;
constexpr size_t ADVPREC_LDS = 64 * 1024;

// adjoint step (sf_adjoint.hip)
template <typename T>
struct AdjArgs {
    int N3, N3p, A1;
    size_t n;            // nodal points N1*N2*N3
    const T* U[4];       // adjoint nodal u1, u2, u3, b
    const T* Tt[4];      // direct totals, nodal
    const T* DX[4];      // nodal ddx of the totals
    const T* DY[4];      // nodal ddy of the totals (3D)
    const cx<T>* u3t;    // modal totals, for the horizontal averages
    const cx<T>* bt;
    const T* zc;
    T Ri;
    T* Q[12];            // nodal explicit terms
};
template <typename T>
struct AdjCombine {
    int N3, A1, A2, three;
    const T *kx, *ky;
    const cx<T>* M[12];  // modal transforms of the nodal explicit terms
    cx<T>* out[4];       // explicit terms r1, r2, r3, rB
};
template <typename T>
hipError_t launch_lerp(cx<T>* out, const cx<T>* above, const cx<T>* below, T f, size_t n, hipStream_t s);
template <typename T>
hipError_t launch_add_profile(T* f, const T* prof, size_t n, int N3, hipStream_t s);
template <typename T>
hipError_t launch_adj_products(const AdjArgs<T>& a, bool three, hipStream_t s);
template <typename T>
hipError_t launch_adj_combine(const AdjCombine<T>& c, hipStream_t s);
template <typename T>
hipError_t launch_mix_product(T* out, const T* Bt, const T* Wt, const cx<T>* bt, size_t n, int N3, int A1,
                              hipStream_t s);
template <typename T>
hipError_t launch_mean_profile(const cx<T>* f, double* out, int N3, int A1, hipStream_t s);
template <typename T>
hipError_t launch_phase_shift(cx<T>* f, size_t n, int A1, double shift, hipStream_t s);
template <typename T>
hipError_t launch_add_mean(cx<T>* f, int N3, int A1, T value, hipStream_t s);

bool tz_supported(int n);
// z levels per x-pass block for this grid (the block height that minimises
// rounds of resident blocks x (levels + 2))
int xpass_levels(int N1, int N2, int N3, bool threeD, int precision);
// twiddle table of an n-point LDS FFT: 2n entries (the layout its second pass
// reads, then the same values in the order of the second pass on lane pairs)
void fft_twiddles(int n, double* re, double* im);
// the same for the plan with first radix r1f (8 or 16 where n allows it,
// else the default), and that plan's first radix
void fft_twiddles_plan(int n, int r1f, double* re, double* im);
// the paired forward y transform of this grid runs on tiles of 16 modes read
// as contiguous runs (A1 a multiple of 16 and k_yfwd_seq's LDS tile fits)
bool yfwd_tiles(int N2, int A1, int precision);
int fft_first_radix(int n, int r1f);
int max_rows_per_lane();
// blocks (and so partial sums) of the two-pass FP64 reductions (dot, max-abs)
int max_reduction_blocks();
// modes per block of the modal kernels for M rows per lane
int modal_modes_per_block(int M, int precision);

}  // namespace sf
