/*
 * nsgpu.h -- C ABI of the MI355X-native right-hand-side + SSP-RK3 path of the compressible
 * Navier-Stokes block solver.
 *
 * Plain C: pointers, sizes and status codes; no C++ types or exceptions cross it.  One context
 * binds one GPU and one HIP stream and owns all device memory of one block.  Host arrays are
 * caller-owned and use the reference Array4d layout: row-major [NX][NY][NZ][5], q fastest,
 * NX = nx + 2*3 (containers/array_4d.hpp:116-124, array_4d_dimensions.hpp:40-49); they are read
 * or written synchronously before a call returns.  Variables: conservative U =
 * [rho, rhoE, rho*u, rho*v, rho*w], primitive W = [rho, p, u, v, w] (eq_of_state.cpp:32-87).
 *
 * Reference interfaces replaced (paths relative to the reference's src/):
 *   nsgpu_apply_boundary    Domain::apply_boundary_conditions + Communicator::communicate
 *                           (domain.hpp:34, domain.cpp:18-28; communicator.hpp:21,
 *                           communicator.cpp:6-144) on one block: Neumann faces, periodic wrap
 *   nsgpu_convection_solve  ConvectionScheme::solve(grid, W, U, R) (convection_scheme.hpp:26;
 *                           Weno5 weno-5.hpp:19-22, Weno5cw weno-5cw.hpp): adds into R
 *   nsgpu_diffusion_solve   DiffusionScheme::solve(grid, comm, W, R) (diffusion_scheme.hpp:14-17;
 *                           CenteredDiffusion centered_diffusion.hpp): adds into R
 *   nsgpu_rhs               one stage body of TimeScheme::step (time_scheme.cpp:33-41)
 *   nsgpu_step              TimeScheme::step (time_scheme.hpp:25-29, time_scheme.cpp:3-70),
 *                           state resident on the GPU between calls
 *   nsgpu_get_dt            TimeScheme::get_dt (time_scheme.hpp:31)
 *   nsgpu_set_state /       Domain::get_NS_primitive_array / get_NS_conservative_array
 *   nsgpu_get_state         (domain.hpp:36-40): upload / download of the block state
 *   nsgpu_reduce            no reference counterpart: conservation sums, CFL minimum and health
 *                           flags for callers that monitor a run
 */
#ifndef NSGPU_H
#define NSGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nsgpu_ctx nsgpu_ctx;

/* status codes: 0 = OK, < 0 = error */
enum {
    NSGPU_OK              = 0,
    NSGPU_ERR_INVALID     = -1, /* bad configuration or argument */
    NSGPU_ERR_HIP         = -2, /* a HIP runtime call failed */
    NSGPU_ERR_NODEVICE    = -3, /* no GPU / device index out of range */
    NSGPU_ERR_NONFINITE   = -4, /* state holds NaN / Inf (nsgpu_step, nsgpu_reduce) */
    NSGPU_ERR_POSITIVITY  = -5, /* rho <= 0 or p <= 0 (nsgpu_step, nsgpu_reduce) */
    NSGPU_ERR_ALPHA       = -6, /* max |u|+c exceeds the Lax-Friedrichs constant 3.0 */
    NSGPU_ERR_UNSUPPORTED = -7  /* valid request the library does not implement */
};

/* boundary kind per face; faces in the reference order W,E,N,S,F,B (domain.cpp:18-28):
 * West = i-low, East = i-high, North = j-low, South = j-high, Front = k-high, Back = k-low.
 * An axis is periodic when its West / South / Front entry says so (decomposition.cpp:145-161);
 * each face of a non-periodic axis must then be NEUMANN (a reference face with another kind has
 * no boundary object, domain.cpp:87-137). */
enum { NSGPU_BC_PERIODIC = 0, NSGPU_BC_NEUMANN = 1 };
enum { NSGPU_CONV_NONE = 0, NSGPU_CONV_WENO5 = 1, NSGPU_CONV_WENO5CW = 2 };
/* LEARNED (extension): the WENO-5 nonlinear weights come from a small network, see
 * nsgpu_set_mlp below; the classic Shu weights are kept on smooth and flagged stencils */
enum { NSGPU_WEIGHTS_SHU = 0, NSGPU_WEIGHTS_BORGES = 1, NSGPU_WEIGHTS_LEARNED = 2 };
/* arithmetic of the kernels:
 *   EXACT  the reference's FP64 operation order, no contraction, correctly rounded division and
 *          sqrt: results are bitwise those of a strict (-ffp-contract=off) build of the reference
 *   FAST   FP64 with fused multiply-add, hoisted reciprocals and one division per WENO
 *          reconstruction; agrees with EXACT to round-off (tolerances in the tests)
 *   FP32   FAST with the reconstruction and the viscous fluxes in single precision; state, stage
 *          accumulation, equation of state and reductions in FP64 */
enum { NSGPU_MATH_EXACT = 0, NSGPU_MATH_FAST = 1, NSGPU_MATH_FP32 = 2 };
enum { NSGPU_STATE_CONSERVATIVE = 0, NSGPU_STATE_PRIMITIVE = 1 };
/* viscosity law: CONSTANT mu (CenteredDiffusion as the reference runs it, dmu = 0,
 * centered_diffusion.cpp:96-103), or SUTHERLAND mu(T) = mu (1 + C)/(T + C) T^1.5 with
 * T = gamma p / rho (Viscosity::mu_sutherland, viscosity.cpp:5-11, which the reference defines
 * but never calls) feeding the dmu slots of the same residual */
enum { NSGPU_VISC_CONSTANT = 0, NSGPU_VISC_SUTHERLAND = 1 };
/* Lax-Friedrichs constant of the flux split (extension): CONST3 is the reference's alpha = 3.0
 * (ConvectionFlux::max_eigenvalue, convection_flux.cpp:241); LOCAL_MAX takes, per face i+1/2, the
 * largest max(|q_n - c|, |q_n + c|) over the cells i-2 .. i+2 (the eigenvalue that line leaves
 * commented out, maximised by Weno5::get_max_alpha, weno-5.cpp:324-330).  With LOCAL_MAX the
 * split stays upwind for any sound speed, and nsgpu_reduce raises no alpha flag. */
enum { NSGPU_ALPHA_CONST3 = 0, NSGPU_ALPHA_LOCAL_MAX = 1 };
/* corrections of reference behaviour (extension, bit mask; 0 = the reference):
 *   CONVECTION  the characteristic projection of Weno5cw with the eigenvectors of the true flux
 *               Jacobian: c^2 = gamma p / rho (convection_flux.cpp:139 has gamma / (p rho)),
 *               L and R in the order of the stored conserved vector (the reference builds them
 *               for [rho, rho u, rho v, rho w, rho E] and applies them to [rho, rho E, rho u,
 *               rho v, rho w], weno-5cw.cpp:92-93), and the Roe-averaged face state instead of
 *               the mean of the primitives (weno-5cw.cpp:67-70)
 *   VISCOUS     the full 3-D viscous and heat-flux terms (3-D blocks only): what
 *               CenteredDiffusion::compute_residual (centered_diffusion.cpp:70-182) computes for x
 *               and y, for all three axes, with each second derivative on its own axis
 *               (cd-n.cpp:74-86 runs the x stencil), Tauyy from dv/dy (:143), the viscous work
 *               tau_ij du_i/dx_j (:147-148 multiply Tauxy by dv/dy), and the heat flux from the
 *               derivatives of T = p / rho (:161-166 use the value 1 for p in the y term and
 *               +2 / 1 where the expansion of d2(p/rho) has -2 / 2).  Not in ensembles
 *               (NSGPU_ERR_UNSUPPORTED) */
enum { NSGPU_PHYS_REFERENCE = 0, NSGPU_PHYS_CONVECTION = 1, NSGPU_PHYS_VISCOUS = 2,
       NSGPU_PHYS_CORRECTED = 3 };
enum { NSGPU_RHS_CONVECTION = 1, NSGPU_RHS_DIFFUSION = 2, NSGPU_RHS_BOTH = 3 };

typedef struct {
    int    nx, ny, nz;  /* interior cells of the block (each >= 3) */
    int    ngc;         /* halo layers: must be 3 (params.hpp:16-18) */
    int    dim;         /* 3, or 2: the z sweep is skipped (weno-5.cpp:28) */
    double Lx, Ly, Lz;  /* block lengths: dx = Lx / nx (grid.cpp:14-16) */
    double gamma, mu, Pr, cfl;
    int    conv;        /* NSGPU_CONV_* */
    int    weights;     /* NSGPU_WEIGHTS_* */
    int    diff_order;  /* 0: no diffusion; 2, 4, 6: CenteredDiffusion order */
    int    bc[6];       /* NSGPU_BC_*, faces W,E,N,S,F,B */
    int    math;        /* NSGPU_MATH_* */
    int    viscosity;   /* NSGPU_VISC_*: CONSTANT (the reference path) or SUTHERLAND */
    double suth_C;      /* Sutherland constant over the reference temperature (T0 = 1) */
    int    alpha_mode;  /* NSGPU_ALPHA_*; 0 = the reference */
    int    physics;     /* NSGPU_PHYS_* bits; 0 = the reference.  The learned weights run with
                           alpha_mode 0 and physics 0 only (NSGPU_ERR_UNSUPPORTED otherwise) */
} nsgpu_config;

/* Create a context on GPU `device` (HIP ordinal).  Allocates the block's device arrays. */
int nsgpu_create(const nsgpu_config* cfg, int device, nsgpu_ctx** out);
/* Frees the context.  A context that is a member of a live ensemble is not freed (the ensemble
 * holds its device buffers): the call prints an error and returns; destroy the ensemble first.
 * Members must outlive their ensembles. */
void nsgpu_destroy(nsgpu_ctx* ctx);
const char* nsgpu_strerror(int status);

/* number of doubles of one padded host array: (nx+6)*(ny+6)*(nz+6)*5 */
size_t nsgpu_array_len(const nsgpu_ctx* ctx);

/* Upload / download the block state (whole padded array).  PRIMITIVE is the state TimeScheme::step
 * advances (the reference keeps W between steps and re-derives U from it, time_scheme.cpp:25);
 * CONSERVATIVE reads the U of the last step (interior).  Setting either form derives the other
 * on the device (p2c / c2p on the interior and face halos, eq_of_state.cpp:32-87), so
 * nsgpu_get_state, nsgpu_reduce and the next step all see the state just set. */
int nsgpu_set_state(nsgpu_ctx* ctx, const double* aos, int kind);
int nsgpu_get_state(nsgpu_ctx* ctx, double* aos, int kind);

/* Boundary conditions + single-block halo exchange on a 5-variable array, in place.
 * This call and the three operator-level calls below (convection_solve, diffusion_solve, rhs)
 * work in scratch arrays of the context: the resident state that nsgpu_step advances,
 * nsgpu_get_state returns and nsgpu_reduce reads is left as it was. */
int nsgpu_apply_boundary(nsgpu_ctx* ctx, double* aos);

/* ConvectionScheme::solve: R += convective residual of (W, U), halos of W and U as given. */
int nsgpu_convection_solve(nsgpu_ctx* ctx, const double* W_aos, const double* U_aos, double* R_aos);
/* DiffusionScheme::solve: R += viscous residual of W (halos as given); the first-derivative
 * arrays get the periodic exchange, or zero halos at physical walls. */
int nsgpu_diffusion_solve(nsgpu_ctx* ctx, const double* W_aos, double* R_aos);

/* Stage body: U (interior) -> boundary + exchange in place -> W = c2p(U) -> R = 0 + terms of
 * `which`.  W_out may be NULL. */
int nsgpu_rhs(nsgpu_ctx* ctx, double* U_aos, double* W_out, double* R_out, int which);

/* nsteps x TimeScheme::step on the resident primitive state.  Adds each step's dt to *time_io
 * (may be NULL) and stores the last dt in *dt_last (may be NULL).  Synchronises once at the end.
 * Health: every stage update checks the state it writes; the call returns NSGPU_ERR_NONFINITE
 * if any stage of any of its steps produced a NaN / Inf, else NSGPU_ERR_POSITIVITY if one
 * produced rho <= 0 or p <= 0 (the steps still all run; the reference has no such check and
 * would carry on).  The alpha check stays with nsgpu_reduce. */
int nsgpu_step(nsgpu_ctx* ctx, int nsteps, double* time_io, double* dt_last);
/* dt of the last step taken (the prescribed one after nsgpu_step_fixed). */
int nsgpu_get_dt(nsgpu_ctx* ctx, double* dt);
/* No reference counterpart (the reference always takes the CFL step, time_scheme.cpp:16):
 * nsteps steps with a prescribed dt, for convergence and resolution studies that compare runs
 * at equal times.  Every stage and the time use dt itself (the CFL does not enter); health as
 * nsgpu_step. */
int nsgpu_step_fixed(nsgpu_ctx* ctx, int nsteps, double dt, double* time_io);

/* Conservation sums over the interior of the resident conservative state (sum of U_q * dx*dy*dz),
 * the CFL time-step minimum of the resident primitive state, and health flags
 * (bit 0 non-finite, bit 1 rho<=0 or p<=0, bit 2 max|u|+c > 3). */
int nsgpu_reduce(nsgpu_ctx* ctx, double sums[5], double* dt_min, unsigned* flags);

/* No reference counterpart: checks on GPU `device` the device arithmetic the EXACT path relies on.
 * Divides n pairs of random operands (exponents in [-300, 300]) with the kernels' division
 * routine and with the IEEE operator and stores the number of differing quotients.  For callers
 * that validate a new GPU, driver or compiler before trusting bitwise results. */
int nsgpu_selftest(int device, long n, unsigned long long seed, long* mismatches);

/* The context's HIP stream (a hipStream_t), for callers that time or order work around it.  The
 * step runs some launches on a second internal stream that joins back into this one within each
 * stage, so work ordered on this stream after a call sees all of its results. */
void* nsgpu_stream(nsgpu_ctx* ctx);

/* Per-kernel timing with HIP events on the context's stream: enable, then read the accumulated
 * milliseconds and launch counts of the kernel families (names via nsgpu_timer_name). */
int nsgpu_timing(nsgpu_ctx* ctx, int enable);
int nsgpu_timer_count(void);
const char* nsgpu_timer_name(int i);
int nsgpu_timer_read(nsgpu_ctx* ctx, int i, double* ms, long* launches);

/* ---- Learned WENO-5 weights -------------------------------------------------------------------
 * No reference counterpart (the reference's WENO5_weights_shu / _borges, weno-5.cpp:222-283, are
 * what it replaces).  An 8-32-32-3 MLP (GELU, softmax) maps the five stencil values normalised
 * by their range, (v - min) / (max - min), and their three smoothness indicators to the three
 * weights.  Stencils with (max beta + 1e-6) / (min beta + 1e-6) below 3 (smooth) or above 1e3
 * (flagged discontinuity) keep the classic Shu weights, bitwise as the SHU path computes them.
 * Parameters: NSGPU_MLP_NPARAM floats, row-major W1[32][8], b1[32], W2[32][32], b2[32],
 * W3[3][32], b3[3].  The kernels use them as bfloat16 with FP32 accumulation (matrix cores). */
#define NSGPU_MLP_NPARAM 1443

/* Network of a context created with weights = NSGPU_WEIGHTS_LEARNED (required before stepping). */
int nsgpu_set_mlp(nsgpu_ctx* ctx, const float* params);

/* Batch inference on GPU `device`: softmax weights W[n][3] of the features X[n][8]. */
int nsgpu_mlp_forward(int device, const float* params, long n, const float* X, float* W);
/* Mean over n samples of (sum_k w_k Q[s][k] - T[s])^2 and its parameter gradient. */
int nsgpu_mlp_loss_grad(int device, const float* params, long n, const float* X, const float* Q,
                        const float* T, double* loss, float* grad);
/* Adam training on the GPU (lr, beta 0.9 / 0.999, eps 1e-8): `steps` steps on consecutive
 * mini-batches of `batch` samples of (X, Q, T), wrapping around; params updated in place,
 * loss_hist[step] = mini-batch loss before the step (may be NULL). */
int nsgpu_mlp_train(int device, float* params, long n, const float* X, const float* Q,
                    const float* T, int steps, long batch, float lr, float* loss_hist);
/* The learned reconstruction of the face-flux kernels of math mode `math` on n stencils
 * S[n][5]: side 0 = WENO5_left, 1 = WENO5_right; out[n], regime[n] (0 smooth, 1 network,
 * 2 flagged). */
int nsgpu_mlp_recon(int device, int math, const float* params, long n, const double* S, int side,
                    double* out, int* regime);

/* No counterpart in the reference.  For callers that cannot allocate device memory: code under
 * HIP graph capture, training loops that forbid allocation between steps.  The mean loss and
 * its parameter gradient as nsgpu_mlp_loss_grad computes them (same kernels, same summation
 * order, bitwise the same floats), but every pointer is DEVICE memory of the caller: params
 * [NSGPU_MLP_NPARAM], X[n][8], Q[n][3], T[n], out[NSGPU_MLP_NPARAM + 1] (gradient, then the loss).
 * The launches are enqueued on `stream` (a hipStream_t; NULL = the null stream) and the call
 * returns without waiting: read `out` after synchronising that stream.  It calls no hipMalloc,
 * hipFree or allocating library routine: the packed network and the per-workgroup partial
 * gradients live in a fixed device array of the library (6 MB per GPU, held until the process
 * ends), so n is limited to NSGPU_MLP_NOALLOC_MAX_N samples (NSGPU_ERR_INVALID beyond).  All
 * calls on one GPU share that array: calls that overlap on one GPU (different streams, or
 * threads) race and must be serialised by the caller. */
#define NSGPU_MLP_NOALLOC_MAX_N 262144L
int nsgpu_mlp_loss_grad_noalloc(int device, const float* params, long n, const float* X,
                                const float* Q, const float* T, float* out, void* stream);

/* ---- Ensembles: many independent blocks of different sizes on one GPU ----------------------
 * No reference counterpart: the reference advances one block per MPI rank, each with its own
 * driver loop (driver.cpp:45-64 calling TimeScheme::step, time_scheme.cpp:3-70).  An ensemble
 * advances n member contexts together: each kernel family of an SSP-RK3 step runs as ONE launch
 * over every block, and the whole step (33 launches) is captured once into a hipGraph and
 * replayed.  Each member keeps its own dt (its own CFL minimum), time and boundary conditions,
 * and its results are those of nsgpu_step on that member alone.  Members must share the device,
 * the math mode, the convection scheme and weights, the diffusion order, the viscosity law and
 * dim; sizes, lengths, gamma, mu, Pr, CFL and boundary kinds may differ.  The member contexts
 * stay usable (set_state, get_state, reduce, step) between ensemble calls. */
typedef struct nsgpu_ensemble nsgpu_ensemble;

/* The members stay owned by the caller and must outlive the ensemble: nsgpu_destroy refuses a
 * context while an ensemble holds it. */
int  nsgpu_ensemble_create(nsgpu_ctx* const* members, int n, nsgpu_ensemble** out);
void nsgpu_ensemble_destroy(nsgpu_ensemble* ens);

/* Tile shape per member, 3 ints each: kw_flux (lanes along k of a face-flux wave: 16, 32 or 64),
 * kw_cell (lanes along k of the 256-thread cell kernels: 16, 32 or 64), march (planes per
 * marching chunk of the x-stencil kernels, 1..1024).  Defaults: 64, 64, 16. */
int nsgpu_ensemble_set_tiles(nsgpu_ensemble* ens, const int* tiles);
int nsgpu_ensemble_get_tiles(const nsgpu_ensemble* ens, int* tiles);

/* nsteps SSP-RK3 steps of every member; use_graph != 0 replays the captured step graph,
 * otherwise the launches are enqueued directly.  Synchronises once at the end.  Health as
 * nsgpu_step, over all members (nsgpu_reduce on each member tells which one failed). */
int nsgpu_ensemble_step(nsgpu_ensemble* ens, int nsteps, int use_graph);

/* Autotuner: coordinate descent over the tile shapes, kw_flux and kw_cell per class of members
 * with equal nz and march per class of equal nx, timing `steps` graph-replayed steps per trial
 * (best of 2).  The member states (primitive and conservative state, time and last dt) are
 * restored after every trial, also after a failed one, so tuning leaves them as they were.
 * Stores the step time of the default tiles and of the chosen tiles (ms, may be NULL) and
 * leaves the chosen tiles set. */
int nsgpu_ensemble_tune(nsgpu_ensemble* ens, int steps, double* ms_default, double* ms_tuned);

/* ---- Decomposed block: several linked blocks on one GPU --------------------------------------
 * Replaces the reference's multi-rank run of one block per MPI rank: Decomposition
 * (decomposition.cpp:26-57: n / p cells per block, the remainder to block 0; neighbours from the
 * Cartesian topology, periodic where West / South / Front say so, :116-161), Grid's global spacing
 * (grid.cpp:14-16), Communicator::communicate (communicator.cpp:6-144: the 3-layer face halos of
 * each block from its neighbours, faces only) and the dt minimum over ranks (time_scheme.cpp:131-135,
 * MPI_Allreduce MIN).  The blocks live on one GPU: halo kernels copy the neighbouring block's
 * interior planes, the derivative halos of the viscous terms are exchanged the same way
 * (centered_diffusion.cpp:22-24), and one SSP-RK3 step of all blocks is one launch per kernel
 * family, replayed as one hipGraph.  The result of every block is bitwise that of the same cells
 * advanced as one block.  nproc[3] = blocks per axis (x, y, z); dim = 2 needs nproc[2] = 1;
 * every block at least 3 cells thick along a split axis; the full 3-D viscous option is not
 * supported (NSGPU_ERR_UNSUPPORTED). */
typedef struct nsgpu_domain nsgpu_domain;

int  nsgpu_domain_create(const nsgpu_config* global, const int nproc[3], int device,
                         nsgpu_domain** out);
void nsgpu_domain_destroy(nsgpu_domain* dom);
/* number of blocks (nproc[0] * nproc[1] * nproc[2]) */
int  nsgpu_domain_size(const nsgpu_domain* dom);
/* Block `rank` (rank = (k * nproc[1] + j) * nproc[0] + i for block coordinates i, j, k, the
 * order of the reference's Cartesian communicator): its first global cell index and its cell
 * counts per axis, and its context, owned by the domain.  Use nsgpu_set_state / nsgpu_get_state /
 * nsgpu_reduce on it with the block's own padded arrays (Domain::get_NS_primitive_array of that
 * rank); the operator-level calls and nsgpu_step refuse it (NSGPU_ERR_INVALID). */
nsgpu_ctx* nsgpu_domain_block(nsgpu_domain* dom, int rank, int start[3], int count[3]);
/* nsteps SSP-RK3 steps of all blocks, as nsgpu_step (time, last dt, health over all blocks). */
int nsgpu_domain_step(nsgpu_domain* dom, int nsteps, double* time_io, double* dt_last);

#ifdef __cplusplus
}
#endif
#endif
