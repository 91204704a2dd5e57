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
enum { NSO_W_SHU = 0, NSO_W_BORGES = 1 };

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
} nso_cfg;

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

#ifdef __cplusplus
}
#endif
#endif
