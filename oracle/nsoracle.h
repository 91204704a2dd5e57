/*
 * nsoracle.h -- CPU restatement (oracle) of the reference solver's right-hand-side and SSP-RK3
 * path, in plain C with FP64 arithmetic in the reference's own operation order.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker.  The product (lib/libnsgpu.so) never links or
 * calls it.
 *
 * Arrays use the reference Array4d layout (containers/array_4d.hpp:64-71, :116-124):
 * row-major [NX][NY][NZ][5] with q fastest, NX = nx + 2*3 (array_4d_dimensions.hpp:40-49),
 * 3 halo layers (params.hpp:16-18).  Variables: U = [rho, rhoE, rho*u, rho*v, rho*w],
 * W = [rho, p, u, v, w] (eq_of_state.cpp:32-87).
 */
#ifndef NSORACLE_H
#define NSORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary kinds per face, faces in the reference order W,E,N,S,F,B (domain.cpp:18-28).
 * x: West = i-low, East = i-high; y: North = j-low, South = j-high; z: Back = k-low,
 * Front = k-high (boundary.cpp:11-16).  Periodicity of an axis is decided by its West / South /
 * Front entry (decomposition.cpp:145-161). */
enum { NSO_BC_PERIODIC = 0, NSO_BC_NEUMANN = 1 };
enum { NSO_CONV_NONE = 0, NSO_CONV_WENO5 = 1, NSO_CONV_WENO5CW = 2 };
enum { NSO_W_SHU = 0, NSO_W_BORGES = 1, NSO_W_LEARNED = 2 };

typedef struct {
    int    nx, ny, nz; /* interior cells of the block */
    int    dim;        /* 3, or 2: z sweep skipped (weno-5.cpp:28) */
    double Lx, Ly, Lz; /* dx = Lx/nx (grid.cpp:14-16) */
    double gamma, mu, Pr, cfl;
    int    conv;       /* NSO_CONV_* */
    int    weights;    /* NSO_W_* */
    int    diff_order; /* 0 = no diffusion, else 2, 4 or 6 (cd_constants.hpp) */
    int    bc[6];      /* NSO_BC_*, W,E,N,S,F,B */
    int    viscosity;  /* 0: constant mu (the reference path); 1: Sutherland law (extension) */
    double suth_C;     /* Sutherland constant over the reference temperature (T0 = 1) */
    int    alpha_mode; /* NSO_ALPHA_*: the Lax-Friedrichs constant of the flux split */
    int    physics;    /* NSO_PHYS_* bits: corrections of reference behaviour (extension) */
} nso_cfg;

/* CONST3: alpha = 3.0 everywhere (convection_flux.cpp:241, the reference).  LOCAL_MAX: the
 * eigenvalue that line has commented out, max(|q_n - c|, |q_n + c|) with c = sqrt(gamma p / rho),
 * taken per face as the maximum over the cells i-2 .. i+2 (Weno5::get_max_alpha,
 * weno-5.cpp:324-330, weno-5cw.cpp:60-64). */
enum { NSO_ALPHA_CONST3 = 0, NSO_ALPHA_LOCAL_MAX = 1 };
/* CONVECTION: the characteristic projection of Weno5cw uses the eigenvectors of the true flux
 * Jacobian: c^2 = gamma p / rho (convection_flux.cpp:139 writes gamma / (p rho)), L and R in the
 * order of the conserved vector they act on (the reference builds them for
 * [rho, rho u, rho v, rho w, rho E] and applies them to [rho, rho E, rho u, rho v, rho w]), and the
 * Roe-averaged face state instead of the mean of the primitives (weno-5cw.cpp:67-70). */
/* VISCOUS: the full 3-D viscous and heat-flux terms instead of the reference's 2-D subset with
 * its quirks (centered_diffusion.cpp:70-182); 3-D blocks only. */
enum { NSO_PHYS_CONVECTION = 1, NSO_PHYS_VISCOUS = 2 };

/* the face state, L and R that the characteristic sweep applies at the face between the cells
 * with primitives Wl and Wr, L and R acting on vectors in the stored order [rho, rho E, rho u,
 * rho v, rho w] (row-major 5x5) */
void nso_face_eigensystem(const nso_cfg* c, const double Wl[5], const double Wr[5], int dir,
                          double Wface[5], double L[25], double R[25]);
/* the Lax-Friedrichs constant of the face i+1/2 along dir (padded indices of cell i) */
double nso_face_alpha(const nso_cfg* c, const double* W, int dir, long i, long j, long k);

/* number of doubles of one padded 5-variable array */
size_t nso_array_len(const nso_cfg* c);

/* eq_of_state.cpp:32-59 / :61-87, over all cells including halos */
void nso_p2c(const nso_cfg* c, const double* W, double* U);
void nso_c2p(const nso_cfg* c, const double* U, double* W);

/* Domain::apply_boundary_conditions (domain.cpp:18-28) followed by Communicator::communicate
 * (communicator.cpp:6-14) on a single block: Neumann faces, then periodic face wrap x, y, z. */
void nso_fill_halos(const nso_cfg* c, double* A);

/* One RHS evaluation as the stage body of TimeScheme::step (time_scheme.cpp:33-41): U interior in,
 * halos filled in place, W = c2p(U) over all cells, then R = 0 + convection + diffusion.
 * which: bit 0 = convection, bit 1 = diffusion.  Returns 0, or -1 on a bad config. */
int nso_rhs(const nso_cfg* c, double* U, double* W, double* R, int which);

/* TimeScheme::set_dt (time_scheme.cpp:98-136): CFL * min over interior of local_dt */
double nso_dt(const nso_cfg* c, const double* W);

/* nsteps x TimeScheme::step (time_scheme.cpp:3-70) from primitive W (interior used).
 * U receives the last conservative state, dts[s] the dt of step s (may be NULL). */
int nso_step(const nso_cfg* c, double* W, double* U, int nsteps, double* dts);

/* Building blocks, exported for the unit tests. */
void   nso_weno5_left(const double f[5], int weights, double out[1]);
void   nso_weno5_right(const double f[5], int weights, double out[1]);
void   nso_eigenvectors(const double Wavg[5], double gamma, int dir, double L[25], double R[25]);
/* CDN::generic_fd_{x,y,z} (cd-n.cpp:91-165): interior df = sum_k f(.+s+k)*c_k/d along axis dir;
 * second selects the second-derivative coefficients; order is c->diff_order. */
void   nso_fd(const nso_cfg* c, const double* f, double* df, int dir, int second, double d);
int    nso_num_threads(void);
double nso_mu_sutherland(double mu0, double C, double gamma, double rho, double p);

/* ---- learned WENO-5 weights (extension, nsmlp.c) -------------------------------------------
 * An 8-32-32-3 MLP (GELU, softmax) maps the range-normalised stencil and its three smoothness
 * indicators to the three WENO weights.  Faces whose indicator ratio
 * (max beta + EPS) / (min beta + EPS) is below SMOOTH (smooth) or above FLAG (flagged
 * discontinuity) keep the classic Shu weights; in between the network's reconstruction enters
 * with the share nso_mlp_blend(ratio).  Parameter layout (row-major):
 * W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32], b3[3]. */
#define NSO_MLP_HIDDEN 32
#define NSO_MLP_NPARAM (32 * 8 + 32 + 32 * 32 + 32 + 3 * 32 + 3)
#define NSO_MLP_EPS 1e-6
#define NSO_MLP_SMOOTH 3.0
#define NSO_MLP_FLAG 1e3
/* a stencil whose range is at most this share of its largest magnitude counts as constant; a
 * momentum stencil of a component-wise face also when its range is at most this share of the
 * face's momentum scale (nso_mlp_face_scale) */
#define NSO_MLP_RANGE (1.0 / 8192.0)
enum { NSO_REGIME_SMOOTH = 0, NSO_REGIME_MLP = 1, NSO_REGIME_FLAGGED = 2 };

double nso_bf16(double v);
/* x[0..4] = (v - min) / (max - min) (0 when max == min), x[5..7] = beta_0..2 of x */
void   nso_mlp_features(const double v[5], double x[8]);
/* the same with a scale (>= 0) the range is also measured against */
void   nso_mlp_features_s(const double v[5], double scale, double x[8]);
/* momentum scale of a component-wise face: the largest |split value| of the momentum
 * components (stored positions 2..4) at the face's two cells, fl[p][2], fl[p][3] (fL of cells i,
 * i+1) and fr[p][1], fr[p][2] (fR of cells i, i+1) */
double nso_mlp_face_scale(const double fl[5][5], const double fr[5][5]);
int    nso_mlp_regime(const double x[8]);
/* (max beta + EPS) / (min beta + EPS) and the network's share of the reconstruction at it:
 * min(1, (r - SMOOTH) / SMOOTH) * min(1, (FLAG - r) / (0.75 FLAG)), clipped at 0 */
double nso_mlp_ratio(const double x[8]);
double nso_mlp_blend(double ratio);
/* the three left-biased ENO candidates of the normalised stencil x[0..4] */
void   nso_mlp_candidates(const double x[8], double q[3]);
/* softmax weights; bf16 != 0 rounds the operands of the matrix products like the GPU kernels */
void   nso_mlp_forward(const double* params, const double x[8], int bf16, double w[3]);
/* mean squared error of sum_k w_k Q[s][k] against T[s] over n samples, and its gradient */
double nso_mlp_loss_grad(const double* params, long n, const double* X, const double* Q,
                         const double* T, int bf16, double* grad);
/* learned reconstruction of one stencil, side 0 = WENO5_left, 1 = WENO5_right */
void   nso_weno5_learned(const double* params, int bf16, const double f[5], int side,
                         double out[1], int* regime);
/* network used by nso_rhs / nso_step when cfg.weights == NSO_W_LEARNED */
void   nso_set_mlp(const double* params, int bf16);
double nso_weno5_learned_g(const double f[5], int side);
/* the same with the constant-stencil scale of nso_mlp_features_s */
double nso_weno5_learned_gs(const double f[5], int side, double scale);

#ifdef __cplusplus
}
#endif
#endif
