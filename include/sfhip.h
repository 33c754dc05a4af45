/* sfhip.h -- C ABI of the MI355X (gfx950) IMEX Runge-Kutta timestep.
 *
 * The reference has no plugin or FFI layer: its boundary is the C++ class API
 * of IMEXRK (IMEXRK.h:31-670), StateVector (StateVector.h:8-552) and the Field
 * family (Field.h, Stratiflow.h:8-49).  Each entry point below replaces the
 * reference member function cited next to it; the C++14 facade in
 * include/stratiflow/ maps those class names and signatures onto these calls.
 *
 * Conventions
 *  - Every function returns an int status: SF_OK (0) or an SF_ERR_* code; the
 *    message is available from sf_last_error(ctx).  No C++ exception crosses
 *    this boundary.
 *  - Host arrays use the reference's layouts (Field.h:399-410) in double:
 *      modal : complex, interleaved re/im, index ((N1/2+1)*j2 + j1)*N3 + j3
 *      nodal : real,    index (N1*n2 + n1)*N3 + n3
 *      1D    : real,    index j3
 *    Device storage of modal fields keeps only the active (2/3-dealiased)
 *    modes (ModalField::ParallelPerStack, Field.h:1129-1166); uploading a
 *    modal array therefore applies ModalField::Filter (Field.h:996-1021).
 *  - One context owns one HIP stream; calls are stream-ordered and
 *    asynchronous except downloads, reductions and sf_synchronize.
 */
#ifndef SFHIP_H
#define SFHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SF_OK 0
#define SF_ERR_ARG 1
#define SF_ERR_HIP 2
#define SF_ERR_UNSUPPORTED 3
#define SF_ERR_ALLOC 4
#define SF_ERR_STATE 5

/* Dimensionality, Constants.h:84-89 */
#define SF_THREE_D 0
#define SF_TWO_D 1
#define SF_TWO_AND_A_HALF 2

/* BoundaryCondition, Constants.h:78-82 */
#define SF_NEUMANN 0
#define SF_DIRICHLET 1

/* field kinds: NodalField, ModalField, Nodal1D (Field.h:541-1207) */
#define SF_NODAL 0
#define SF_MODAL 1
#define SF_1D 2

/* solver-owned fields for sf_state_field (IMEXRK.h:605-619, 640-641) */
#define SF_U1 0
#define SF_U2 1
#define SF_U3 2
#define SF_B 3
#define SF_P 4
#define SF_NODAL_U1 5
#define SF_NODAL_U2 6
#define SF_NODAL_U3 7
#define SF_NODAL_B 8
#define SF_R1 9
#define SF_R2 10
#define SF_R3 11
#define SF_RB 12
#define SF_U1_TOT 13 /* linearisation state u1_tot .. b_tot (IMEXRK.h:614-615) */
#define SF_U2_TOT 14
#define SF_U3_TOT 15
#define SF_B_TOT 16

/* GridParams, Parameters.h:6-18 (runtime here, constexpr in the reference) */
typedef struct sf_grid {
    int N1, N2, N3;
    int dimensionality;
} sf_grid;

/* FlowParams, Parameters.h:21-30.  evolve_background != 0: every stage's
 * Crank-Nicolson step also diffuses the background U_ (IMEXRK.cpp:100-104),
 * in the linearised and adjoint steps as well; sf_get_background reads it. */
typedef struct sf_flow {
    double L1, L2, L3, Re, Ri, Pr;
    int evolve_background;
} sf_flow;

typedef struct sf_ctx sf_ctx;
typedef struct sf_field sf_field;

/* ---- context --------------------------------------------------------- */
/* IMEXRK::IMEXRK (IMEXRK.h:37-81).  precision: 4 (float, the reference's
 * default stratifloat) or 8 (double, -DDOUBLE=On). */
int sf_ctx_create(const sf_grid* grid, const sf_flow* flow, int precision, sf_ctx** out);
/* Fields, networks and batches made from ctx must be freed first: they use
 * its stream, and sf_field_free on a field of a destroyed context is a
 * use after free. */
int sf_ctx_destroy(sf_ctx* ctx);
/* message of the last failed call on ctx (or of the last failed
 * sf_ctx_create when ctx is NULL) */
const char* sf_last_error(const sf_ctx* ctx);
/* IMEXRK::deltaT + UpdateForTimestep (IMEXRK.h:34, 502-545) */
int sf_set_dt(sf_ctx* ctx, double dt);
int sf_get_dt(const sf_ctx* ctx, double* dt);
/* StateVector::ResetForParams (StateVector.h:505-508) after flowParams changes */
int sf_set_flow(sf_ctx* ctx, const sf_flow* flow);
/* IMEXRK::SetBackground(const Neumann1D&) (IMEXRK.h:321-324): U is the N3
 * values of U_.  dBdz (N3 values at Neumann points, may be NULL = 1) is the
 * background stratification gradient that IMEXRK.cpp:145-147 fixes at 1. */
int sf_set_background(sf_ctx* ctx, const double* U, const double* dBdz);
/* the current U_ (IMEXRK.h:609), N3 values: what sf_set_background set, as
 * the steps have evolved it when evolve_background is set */
int sf_get_background(sf_ctx* ctx, double* U);
/* VerticalPoints and VerticalPointsFractional (Field.cpp:3-31): N3 values
 * each (z[0] is extrapolated; the reference leaves it uninitialised) */
int sf_vertical_points(const sf_ctx* ctx, double* z, double* zf);

/* ---- fields ---------------------------------------------------------- */
/* NeumannNodal/DirichletNodal/NeumannModal/DirichletModal/Neumann1D
 * constructors (Stratiflow.h:8-49) */
int sf_field_alloc(sf_ctx* ctx, int kind, int bc, sf_field** out);
int sf_field_free(sf_field* f);
int sf_field_kind(const sf_field* f, int* kind, int* bc);
/* host <-> device in the reference layout (Field::Raw, Field.h:418-431) */
int sf_field_upload(sf_field* f, const double* host);
int sf_field_download(const sf_field* f, double* host);
/* Field::operator= (Field.h:285-305) */
int sf_field_copy(sf_field* dst, const sf_field* src);
/* Field::Zero (Field.h:433-440) */
int sf_field_zero(sf_field* f);
/* Field::operator*= (Field.h:342-350) */
int sf_field_scale(sf_field* f, double a);
/* y = a*x + b*y; covers Field::operator+=, -= and StateVector::MulAdd
 * (Field.h:352-377, StateVector.h:67-78) */
int sf_field_axpby(sf_field* y, double a, const sf_field* x, double b);
/* ZeroEnds (Field.h:475-486), NeumannEnds (Field.h:488-500) */
int sf_field_zero_ends(sf_field* f);
int sf_field_neumann_ends(sf_field* f);
/* ModalField::Filter (Field.h:996-1021): a no-op on compact storage */
int sf_field_filter(sf_field* f);
/* NodalField::ToModal (Field.h:734-769), ModalField::ToNodal (Field.h:965-994) */
int sf_to_modal(const sf_field* nodal, sf_field* modal);
int sf_to_nodal(const sf_field* modal, sf_field* nodal);
/* ddx, ddy, ddz (Stratiflow.h:51-80); ddz maps Neumann->Dirichlet and
 * Dirichlet->Neumann; out must have the resulting boundary condition */
int sf_ddx(const sf_field* in, sf_field* out);
int sf_ddy(const sf_field* in, sf_field* out);
int sf_ddz(const sf_field* in, sf_field* out);
/* Diffusion preconditioner for Newton-Krylov solves of F(x) = Phi_T(x) - x
 * (no counterpart in the reference, whose GMRES is unpreconditioned:
 * NewtonKrylov.h:138-235).  For each active mode with kx^2 + ky^2 <= kmax^2:
 *   out = (tnu*(Dzz - k^2) - i*tadv*kx*U(z) - delta)^-1 ZeroEnds(in) - in/2,
 * Dzz Neumannified or Dirichletified by the field's BC as in the implicit
 * solves (IMEXRK.h:502-545); tnu = T*nu (nu = 1/Re, or 1/(Re*Pr) for b),
 * tadv = T to include the advection by the background U_(z) (0: diffusion
 * only), delta > 0 regularises the conserved Neumann mean.  Other modes (all
 * of them when kmax < 0): out = -in.  (T L)^-1 - 1/2 approximates
 * (exp(T L) - I)^-1 for L = nu*(Dzz - k^2) - i*kx*U.  in and out: distinct
 * modal fields of one context and BC.  With tadv != 0: N3 <= 512 (FP32) or
 * 256 (FP64). */
int sf_field_diffusion_precondition(const sf_field* in, sf_field* out, double tnu, double delta, double kmax,
                                    double tadv);
/* InnerProd (Integration.h:117-146) accumulated in FP64 on the GPU */
int sf_inner_prod(const sf_field* a, const sf_field* b, double* out);
/* NodalField += Nodal1D (the per-stack broadcast of the expression templates,
 * Field.h:1215-1253, as in U1_tot = U1 + U_, IMEXRK.cpp:150): profile holds
 * N3 values */
int sf_field_add_profile(sf_field* nodal, const double* profile);
/* HorizontalAverage (Integration.h:37-42): the real part of the (0,0) stack
 * of a modal field into N3 doubles */
int sf_field_mean_profile(const sf_field* modal, double* out);
/* InnerProd under the name of the FP64 dot product (same as sf_inner_prod) */
int sf_dot_f64(const sf_field* a, const sf_field* b, double* out);
/* the per-field part of StateVector::EnforceBCs (StateVector.h:415-421):
 * ZeroEnds for a Dirichlet field, NeumannEnds for a Neumann field */
int sf_field_enforce_bcs(sf_field* f);
/* ModalField::PhaseShift (Field.h:1168-1174): stack(j1,j2) *= exp(i j1 shift) */
int sf_phase_shift(sf_field* modal, double shift);
/* RemoveAverage (Integration.h:72-75): stack(0,0) -= IntegrateAllSpace(f,1,1,L3)/L3/2 */
int sf_remove_average(sf_field* modal);

/* ---- stepper (the IMEXRK solver state) ------------------------------- */
/* borrow one of the solver's own fields (SF_U1 .. SF_B_TOT); the handle is
 * owned by the context and must not be freed */
int sf_state_field(sf_ctx* ctx, int which, sf_field** out);
/* StateVector::CopyToSolver (StateVector.h:511-525): copy u1..p into the
 * solver state (a NULL field zeroes its state field, as 2D zeroes u2) */
int sf_state_bind(sf_ctx* ctx, const sf_field* u1, const sf_field* u2, const sf_field* u3, const sf_field* b,
                  const sf_field* p);
/* StateVector::CopyFromSolver (StateVector.h:532-543): copy the solver state
 * into the non-NULL fields and enforce their boundary conditions */
int sf_state_fetch(sf_ctx* ctx, sf_field* u1, sf_field* u2, sf_field* u3, sf_field* b, sf_field* p);
/* IMEXRK::TimeStep (IMEXRK.cpp:8-24), nsteps times */
int sf_step(sf_ctx* ctx, int nsteps);
/* IMEXRK::TimeStepLinear (IMEXRK.cpp:26-41) about the state set with
 * sf_set_linear_background */
int sf_step_linear(sf_ctx* ctx, int nsteps);
/* IMEXRK::SetBackground(u1,u2,u3,b) (IMEXRK.h:334-352): copies SF_U1_TOT..
 * SF_B_TOT from the given modal fields (u2 may be NULL in 2D) */
int sf_set_linear_background(sf_ctx* ctx, const sf_field* u1, const sf_field* u2, const sf_field* u3,
                             const sf_field* b);
/* IMEXRK::TimeStepAdjoint (IMEXRK.h:86-126, with UpdateAdjointVariables
 * IMEXRK.h:461-500 and BuildRHSAdjoint IMEXRK.cpp:260-379): one backward
 * step of the adjoint state held in SF_U1..SF_B, with the direct state
 * below[] = {u1, u2, u3, b} at the end and above[] at the start of the step
 * (reference argument order u1Below..bBelow, u1Above..bAbove).  The direct
 * states are modal fields allocated with sf_field_alloc (u2 may be NULL in
 * 2D).  Overwrites SF_U1_TOT..SF_B_TOT with the interpolated direct state, as
 * the reference overwrites u1_tot..b_tot.  Call after the set-up of
 * StateVector::AdjointEvolve (StateVector.cpp:158-172): FilterAll,
 * PopulateNodalVariables, RemoveDivergence(0), zero pressure,
 * PopulateNodalVariables (PrepareRunAdjoint, IMEXRK.h:183-199). */
int sf_step_adjoint(sf_ctx* ctx, const sf_field* const below[4], const sf_field* const above[4]);
/* IMEXRK::JoverK (IMEXRK.h:445-459): the mixing diagnostic J/K of
 * UpdateAdjointVariables (IMEXRK.h:470-495), J = IntegrateVertically of the
 * horizontal mean of -(B - <b>) W, K = 2, with B = ToNodal(b) (what every
 * stage leaves in the reference's nodal B) and W = ToNodal(u3) */
int sf_jover_k(sf_ctx* ctx, double* jover_k);
/* one RK stage k of TimeStep (0..2) or TimeStepLinear (linear != 0) */
int sf_stage(sf_ctx* ctx, int k, int linear);
/* IMEXRK::RemoveDivergence (IMEXRK.cpp:43-73) */
int sf_remove_divergence(sf_ctx* ctx, double pressure_multiplier);
/* IMEXRK::FilterAll (IMEXRK.h:128-139) */
int sf_filter_all(sf_ctx* ctx);
/* IMEXRK::PopulateNodalVariables (IMEXRK.h:141-162) */
int sf_populate_nodal(sf_ctx* ctx);
/* The nodal fields SF_NODAL_U1..SF_NODAL_B as the reference holds them at
 * this point (what IMEXRK::SaveFlow writes, IMEXRK.h:415-427): the step never
 * writes physical-space fields, so they are transformed on request from the
 * spectra of the last sf_populate_nodal while it is pending (before the next
 * stage), else from the state that the last stage left. */
int sf_refresh_nodal(sf_ctx* ctx);
/* IMEXRK::KE and PE (IMEXRK.h:395-411), FP64 reductions */
int sf_energy(sf_ctx* ctx, double* ke, double* pe);
/* KE and PE under the name of their FP64 reductions (same as sf_energy) */
int sf_energy_f64(sf_ctx* ctx, double* ke, double* pe);
/* IMEXRK::CFL (IMEXRK.h:355-374) without the dt update: returns the CFL
 * number for the current dt from the nodal fields (populates them) */
int sf_cfl(sf_ctx* ctx, double* cfl);
/* IMEXRK::CFLlinear (IMEXRK.h:376-393) without the dt update: the CFL number
 * for the current dt of the linearised step, from the state set with
 * sf_set_linear_background and 1 as the bound on |U_| */
int sf_cfl_linear(sf_ctx* ctx, double* cfl);
/* sf_cfl_linear without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc), since
 * sf_cfl_linear allocates its nodal scratch field on its first call.  The
 * scratch (one nodal field and the reduction partials) is a fixed 32 MiB
 * device array per GPU, held until the process ends: grids whose nodal field
 * exceeds 32 MiB - 8 KiB (N1*N2*N3*precision bytes) are rejected with
 * SF_ERR_ARG.  Every call on one GPU shares that array, so calls that overlap
 * on one GPU (contexts on different streams or threads) race. */
int sf_cfl_linear_noalloc(sf_ctx* ctx, double* cfl);
int sf_synchronize(sf_ctx* ctx);
/* nsteps of sf_step bracketed by HIP events on the context's stream;
 * *ms = elapsed milliseconds */
int sf_time_steps(sf_ctx* ctx, int nsteps, double* ms);
/* as sf_time_steps, with an event after every kernel launch: kernel_ms[c]
 * and launches[c] accumulate per kernel class c (SF_K_* below) */
#define SF_K_YINV 0
#define SF_K_XPASS 1
#define SF_K_YFWD 2
#define SF_K_MODAL 3
#define SF_K_COUNT 4
int sf_time_steps_detailed(sf_ctx* ctx, int nsteps, double* total_ms, double* kernel_ms, int* launches);

/* ---- batched ensemble ----------------------------------------------------
 * The reference advances one solver per process (the StateVector::solver
 * singleton, StateVector.h:546); an ensemble of independent runs is a loop of
 * FullEvolve/FixedEvolve calls, each TimeStep (IMEXRK.cpp:8-24) on its own.
 * A batch advances many independent 2D contexts (any N1, N3, Re, Ri, Pr, dt)
 * together, and one step of all runs is a HIP graph replay.  Each phase of a
 * stage runs under one of two plans: one launch per distinct N1 (x-pass) or
 * per vertical tile (modal), or ONE launch for all runs (SF_BATCH_MIXED; the
 * x-pass with either FFT plan, the modal kernel with 4 or 8 modes per block).
 * sf_batch_create times every
 * plan, with its levels per block and resident blocks, and keeps the fastest
 * (sf_batch_plan reports the choice and the timings); every plan gives the
 * same results bit for bit.  Each context keeps its own state, set up through the
 * usual calls (sf_set_background, uploads, sf_filter_all, sf_populate_nodal,
 * sf_remove_divergence, sf_set_dt); sf_batch_step(b, n) then has the effect of
 * sf_step(ctx, n) on every member.  The member contexts must outlive the
 * batch, and all members must have the same precision and be 2D. */
typedef struct sf_batch sf_batch;
int sf_batch_create(sf_ctx* const* runs, int nruns, sf_batch** out);
int sf_batch_destroy(sf_batch* b);
const char* sf_batch_last_error(const sf_batch* b);
int sf_batch_step(sf_batch* b, int nsteps);
/* as sf_batch_step, returning the GPU time of the graph replays */
int sf_batch_time_steps(sf_batch* b, int nsteps, double* total_ms);
int sf_batch_info(const sf_batch* b, int* launches_per_step, int* xpass_groups, int* modal_groups);
#define SF_BATCH_GROUPED 0   /* one launch per N1 (x-pass) or per vertical tile (modal) */
#define SF_BATCH_MIXED 1     /* one launch for all runs (x-pass: default FFT plans; modal: 4 modes per block) */
#define SF_BATCH_MIXED_R8 2  /* x-pass: one launch, FFT plans with first radix 8 where N1 allows (N1/8 <= 32) */
#define SF_BATCH_MIXED8 2    /* modal: one launch, 8 modes per block */
/* the plans in use and the tuner's GPU times per stage in ms (six values):
 * x-pass grouped, mixed, mixed radix 8; modal grouped, mixed (4), mixed (8) */
int sf_batch_plan(const sf_batch* b, int* xpass_plan, int* modal_plan, double* tuned_ms);
/* replace the tuner's choice (-1 keeps a phase's plan) */
int sf_batch_select(sf_batch* b, int xpass_plan, int modal_plan);

/* ---- learned preconditioner ----------------------------------------------
 * No counterpart in the reference, whose Newton-GMRES is unpreconditioned
 * (NewtonKrylov.h:138-235).  A small 1D convolutional network over the
 * vertical profile of every horizontal mode (layers x Conv1D(kernel 5),
 * `hidden` channels, ReLU), BF16 weights and activations with FP32
 * accumulation on the matrix cores, trained on the GPU with Adam.  Applied as
 * a right preconditioner:  out = in + s * net(in / s),  s = RMS(in), the
 * identity until trained (the last layer starts at zero).  Field arrays hold
 * u1, u2, u3, b (u2 ignored in 2D).  Training pairs (w, q) satisfy A q = w
 * for the operator being preconditioned; each pair stores one sample per
 * horizontal mode. */
typedef struct sf_net sf_net;
int sf_net_create(sf_ctx* ctx, int hidden, int layers, unsigned long long seed, sf_net** out);
int sf_net_destroy(sf_net* n);
const char* sf_net_last_error(const sf_net* n);
int sf_net_apply(sf_net* n, sf_field* const* in, sf_field* const* out);
int sf_net_add_pair(sf_net* n, sf_field* const* w, sf_field* const* q);
int sf_net_clear_pairs(sf_net* n);
int sf_net_samples(const sf_net* n, int* count);
int sf_net_train(sf_net* n, int epochs, int batch, double lr, double* loss_first, double* loss_last);
/* parameters: FP32 master copy, per layer l: W[5][C[l+1]][C[l]] then bias[C[l+1]] */
int sf_net_param_count(const sf_net* n, size_t* count);
int sf_net_get_params(sf_net* n, float* host);
int sf_net_set_params(sf_net* n, const float* host);
/* layers, padded channel counts C[0..layers] and profile length (N3) */
int sf_net_geometry(const sf_net* n, int* layers, int* channels, int* z);
/* dilation of each layer's five taps (layers values in [1, 8]; default 1): tap t
 * of layer l reads position p + (t-2)*dilations[l], so dilations 1, 2, 4, 8
 * widen the receptive field from 17 to 61 vertical points */
int sf_net_set_dilations(sf_net* n, const int* dilations);
/* raw tensors [profile][z][channel] (checks): forward of nprof profiles, and
 * the mean-squared-error loss against target with its parameter gradient */
int sf_net_forward_raw(sf_net* n, const float* in, int nprof, float* out);
int sf_net_grad_raw(sf_net* n, const float* in, const float* target, int nprof, float* grad, double* loss);

#ifdef __cplusplus
}
#endif

#endif /* SFHIP_H */
