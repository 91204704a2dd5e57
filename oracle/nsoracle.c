/*
 * nsoracle.c -- CPU restatement (oracle) of the reference RHS + SSP-RK3 path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg as the checker.  Never linked into or called by the product library.
 *
 * Every function restates the reference algorithm in the reference's own floating-point operation
 * order (left-to-right evaluation, one rounding per operation, no contraction: the recipe compiles
 * this file with -O2 -ffp-contract=off and without -ffast-math), so that it agrees bitwise with a
 * strict build of the reference.  Citations are paths relative to the reference's src/ directory.
 *
 * Layout: reference Array4d, [NX][NY][NZ][5], q fastest, 3 halo layers (array_4d.hpp:116-124).
 */
#include "nsoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NGC 3
#define NQ 5

typedef struct {
    long NX, NY, NZ;
} pad_t;

static pad_t pad_of(const nso_cfg* c) {
    pad_t p = {c->nx + 2 * NGC, c->ny + 2 * NGC, c->nz + 2 * NGC};
    return p;
}

#define AT(A, P, i, j, k, q)                                                                       \
    (A)[((((size_t)(i)) * (size_t)(P).NY + (size_t)(j)) * (size_t)(P).NZ + (size_t)(k)) * NQ +     \
        (size_t)(q)]

size_t nso_array_len(const nso_cfg* c) {
    pad_t p = pad_of(c);
    return (size_t)p.NX * p.NY * p.NZ * NQ;
}

int nso_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double* alloc_zero(const nso_cfg* c) { return (double*)calloc(nso_array_len(c), sizeof(double)); }

/* ------------------------------------------------------------------------------------------ */
/* Equation of state                                                                          */
/* ------------------------------------------------------------------------------------------ */

/* EqOfState::primitive_to_conservative, eq_of_state.cpp:32-59 (all cells, incl. halos) */
void nso_p2c(const nso_cfg* c, const double* W, double* U) {
    pad_t  P = pad_of(c);
    double g = c->gamma;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < P.NX; i++)
        for (long j = 0; j < P.NY; j++)
            for (long k = 0; k < P.NZ; k++) {
                const double* w = &AT(W, P, i, j, k, 0);
                double*       u = &AT(U, P, i, j, k, 0);
                u[0]            = w[0];
                u[2]            = w[0] * w[2];
                u[3]            = w[0] * w[3];
                u[4]            = w[0] * w[4];
                double kin      = 0.5 * w[0] * (w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
                u[1]            = w[1] / (g - 1.0) + kin;
            }
}

/* EqOfState::conservative_to_primitive, eq_of_state.cpp:61-87 (all cells, incl. halos) */
void nso_c2p(const nso_cfg* c, const double* U, double* W) {
    pad_t  P = pad_of(c);
    double g = c->gamma;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < P.NX; i++)
        for (long j = 0; j < P.NY; j++)
            for (long k = 0; k < P.NZ; k++) {
                const double* u = &AT(U, P, i, j, k, 0);
                double*       w = &AT(W, P, i, j, k, 0);
                w[0]            = u[0];
                w[2]            = u[2] / u[0];
                w[3]            = u[3] / u[0];
                w[4]            = u[4] / u[0];
                double kin      = 0.5 * (w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
                w[1]            = (g - 1.0) * (u[1] - u[0] * kin);
            }
}

/* EqOfState::local_dt, eq_of_state.cpp:89-101 */
static double local_dt(double g, double rho, double p, double u, double v, double w, double dx,
                       double dy, double dz) {
    double c  = sqrt(g * p / rho);
    double ax = fabs(u + c), bx = fabs(u - c);
    double ay = fabs(v + c), by = fabs(v - c);
    double az = fabs(w + c), bz = fabs(w - c);
    /* std::max(a,b) = (a<b)?b:a, std::min(a,b) = (b<a)?b:a */
    double ex  = dx / ((ax < bx) ? bx : ax);
    double ey  = dy / ((ay < by) ? by : ay);
    double ez  = dz / ((az < bz) ? bz : az);
    double tmp = (ey < ex) ? ey : ex;
    return (ez < tmp) ? ez : tmp;
}

/* TimeScheme::get_local_max_dt + set_dt, time_scheme.cpp:98-136 (single block: the MIN
 * all-reduce is the identity) */
double nso_dt(const nso_cfg* c, const double* W) {
    pad_t  P  = pad_of(c);
    double dx = c->Lx / c->nx, dy = c->Ly / c->ny, dz = c->Lz / c->nz;
    double m  = 1E10;
#pragma omp parallel for schedule(static) reduction(min : m)
    for (long i = NGC; i < NGC + c->nx; i++)
        for (long j = NGC; j < NGC + c->ny; j++)
            for (long k = NGC; k < NGC + c->nz; k++) {
                const double* w = &AT(W, P, i, j, k, 0);
                double        d = local_dt(c->gamma, w[0], w[1], w[2], w[3], w[4], dx, dy, dz);
                if (d < m) m = d;
            }
    return m * c->cfl;
}

/* ------------------------------------------------------------------------------------------ */
/* Boundary conditions and the single-block halo exchange                                     */
/* ------------------------------------------------------------------------------------------ */

static int axis_periodic(const nso_cfg* c, int axis) {
    /* decomposition.cpp:145-161: x from West, y from South, z from Front */
    static const int face_of_axis[3] = {0, 3, 4};
    return c->bc[face_of_axis[axis]] == NSO_BC_PERIODIC;
}

/* Neumann::apply_* (boundary_neumann.cpp:22-132): the 3 halo layers of a physical face copy the
 * first interior cell; tangential loops over the interior only (faces, no edges). */
static void neumann_face(const nso_cfg* c, double* A, int face) {
    pad_t P = pad_of(c);
    int   nx = c->nx, ny = c->ny, nz = c->nz;
    switch (face) {
    case 0: /* west */
        for (int h = 0; h < NGC; h++)
            for (long j = NGC; j < NGC + ny; j++)
                for (long k = NGC; k < NGC + nz; k++)
                    for (int q = 0; q < NQ; q++) AT(A, P, h, j, k, q) = AT(A, P, NGC, j, k, q);
        break;
    case 1: /* east */
        for (int h = 0; h < NGC; h++)
            for (long j = NGC; j < NGC + ny; j++)
                for (long k = NGC; k < NGC + nz; k++)
                    for (int q = 0; q < NQ; q++)
                        AT(A, P, nx + NGC + h, j, k, q) = AT(A, P, nx + NGC - 1, j, k, q);
        break;
    case 2: /* north = j low */
        for (long i = NGC; i < NGC + nx; i++)
            for (int h = 0; h < NGC; h++)
                for (long k = NGC; k < NGC + nz; k++)
                    for (int q = 0; q < NQ; q++) AT(A, P, i, h, k, q) = AT(A, P, i, NGC, k, q);
        break;
    case 3: /* south = j high */
        for (long i = NGC; i < NGC + nx; i++)
            for (int h = 0; h < NGC; h++)
                for (long k = NGC; k < NGC + nz; k++)
                    for (int q = 0; q < NQ; q++)
                        AT(A, P, i, ny + NGC + h, k, q) = AT(A, P, i, ny + NGC - 1, k, q);
        break;
    case 4: /* front = k high */
        for (long i = NGC; i < NGC + nx; i++)
            for (long j = NGC; j < NGC + ny; j++)
                for (int h = 0; h < NGC; h++)
                    for (int q = 0; q < NQ; q++)
                        AT(A, P, i, j, nz + NGC + h, q) = AT(A, P, i, j, nz + NGC - 1, q);
        break;
    case 5: /* back = k low */
        for (long i = NGC; i < NGC + nx; i++)
            for (long j = NGC; j < NGC + ny; j++)
                for (int h = 0; h < NGC; h++)
                    for (int q = 0; q < NQ; q++) AT(A, P, i, j, h, q) = AT(A, P, i, j, NGC, q);
        break;
    }
}

/* Communicator::communicate_{x,y,z}_direction (communicator.cpp:16-144) with the face slicers of
 * array_4d.cpp:39-164 on one periodic rank (MPI_Cart_shift returns self): halo_lo[h] =
 * interior[n + h], halo_hi[h] = interior[NGC + h]; faces only.  A non-periodic axis exchanges
 * with MPI_PROC_NULL, i.e. does nothing. */
static void communicate(const nso_cfg* c, double* A) {
    pad_t P = pad_of(c);
    int   nx = c->nx, ny = c->ny, nz = c->nz;
    if (axis_periodic(c, 0))
        for (int h = 0; h < NGC; h++)
            for (long j = NGC; j < NGC + ny; j++)
                for (long k = NGC; k < NGC + nz; k++)
                    for (int q = 0; q < NQ; q++) {
                        AT(A, P, h, j, k, q)            = AT(A, P, nx + h, j, k, q);
                        AT(A, P, NGC + nx + h, j, k, q) = AT(A, P, NGC + h, j, k, q);
                    }
    if (axis_periodic(c, 1))
        for (long i = NGC; i < NGC + nx; i++)
            for (int h = 0; h < NGC; h++)
                for (long k = NGC; k < NGC + nz; k++)
                    for (int q = 0; q < NQ; q++) {
                        AT(A, P, i, h, k, q)            = AT(A, P, i, ny + h, k, q);
                        AT(A, P, i, NGC + ny + h, k, q) = AT(A, P, i, NGC + h, k, q);
                    }
    if (axis_periodic(c, 2))
        for (long i = NGC; i < NGC + nx; i++)
            for (long j = NGC; j < NGC + ny; j++)
                for (int h = 0; h < NGC; h++)
                    for (int q = 0; q < NQ; q++) {
                        AT(A, P, i, j, h, q)            = AT(A, P, i, j, nz + h, q);
                        AT(A, P, i, j, NGC + nz + h, q) = AT(A, P, i, j, NGC + h, q);
                    }
}

void nso_fill_halos(const nso_cfg* c, double* A) {
    /* Domain::apply_boundary_conditions, domain.cpp:18-28: W, E, N, S, F, B; a face gets a
     * Neumann object only when its axis is not periodic (domain.cpp:30-67) */
    static const int order[6] = {0, 1, 2, 3, 4, 5};
    static const int axis[6]  = {0, 0, 1, 1, 2, 2};
    for (int f = 0; f < 6; f++) {
        int face = order[f];
        if (!axis_periodic(c, axis[face]) && c->bc[face] == NSO_BC_NEUMANN) neumann_face(c, A, face);
    }
    communicate(c, A);
}

/* ------------------------------------------------------------------------------------------ */
/* Convection                                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* convection_constants.hpp:4-12 */
static const double epsilon_shu       = 1E-6;
static const double epsilon_borges    = 1E-14;
static const double onepersix         = 1. / 6.;
static const double thirteenpertwelve = 13. / 12.;
static const double oneperfour        = 1. / 4.;
static const double d_left[3]         = {0.1, 0.6, 0.3};
static const double d_right[3]        = {0.3, 0.6, 0.1};

/* Weno5::WENO5_weights_shu / _borges, weno-5.cpp:222-283 */
static void weno_weights(double fm2, double fm1, double f0, double fp1, double fp2,
                         const double d[3], int type, double w[3]) {
    double t1, t2, b0, b1, b2;
    t1 = fm2 - 2. * fm1 + f0;
    t2 = fm2 - 4. * fm1 + 3. * f0;
    b0 = thirteenpertwelve * t1 * t1 + oneperfour * t2 * t2;
    t1 = fm1 - 2. * f0 + fp1;
    t2 = fm1 - fp1;
    b1 = thirteenpertwelve * t1 * t1 + oneperfour * t2 * t2;
    t1 = f0 - 2. * fp1 + fp2;
    t2 = 3. * f0 - 4. * fp1 + fp2;
    b2 = thirteenpertwelve * t1 * t1 + oneperfour * t2 * t2;
    double a0, a1, a2;
    if (type == NSO_W_BORGES) {
        double tau5 = fabs(b0 - b2);
        a0          = d[0] * (1.0 + (tau5 / (b0 + epsilon_borges)));
        a1          = d[1] * (1.0 + (tau5 / (b1 + epsilon_borges)));
        a2          = d[2] * (1.0 + (tau5 / (b2 + epsilon_borges)));
    } else {
        a0 = d[0] / ((epsilon_shu + b0) * (epsilon_shu + b0));
        a1 = d[1] / ((epsilon_shu + b1) * (epsilon_shu + b1));
        a2 = d[2] / ((epsilon_shu + b2) * (epsilon_shu + b2));
    }
    double a = a0 + a1 + a2;
    w[0]     = a0 / a;
    w[1]     = a1 / a;
    w[2]     = a2 / a;
}

/* Weno5::WENO5_left, weno-5.cpp:305-322; scale: the learned weights' constant-stencil scale
 * (nso_mlp_features_s) */
static double weno_left_s(double a, double b, double c, double d, double e, int type,
                          double scale) {
    if (type == NSO_W_LEARNED) {
        const double f[5] = {a, b, c, d, e};
        return nso_weno5_learned_gs(f, 0, scale);
    }
    double f0 = (2. * a - 7. * b + 11. * c) * onepersix;
    double f1 = (-b + 5. * c + 2. * d) * onepersix;
    double f2 = (2. * c + 5. * d - e) * onepersix;
    double w[3];
    weno_weights(a, b, c, d, e, d_left, type, w);
    return w[0] * f0 + w[1] * f1 + w[2] * f2;
}

/* Weno5::WENO5_right, weno-5.cpp:285-303 */
static double weno_right_s(double a, double b, double c, double d, double e, int type,
                           double scale) {
    if (type == NSO_W_LEARNED) {
        const double f[5] = {a, b, c, d, e};
        return nso_weno5_learned_gs(f, 1, scale);
    }
    double f0 = (-a + 5. * b + 2. * c) * onepersix;
    double f1 = (2. * b + 5. * c - d) * onepersix;
    double f2 = (11. * c - 7. * d + 2. * e) * onepersix;
    double w[3];
    weno_weights(a, b, c, d, e, d_right, type, w);
    return w[0] * f0 + w[1] * f1 + w[2] * f2;
}

static double weno_left(double a, double b, double c, double d, double e, int type) {
    return weno_left_s(a, b, c, d, e, type, 0.0);
}
static double weno_right(double a, double b, double c, double d, double e, int type) {
    return weno_right_s(a, b, c, d, e, type, 0.0);
}

void nso_weno5_left(const double f[5], int weights, double out[1]) {
    out[0] = weno_left(f[0], f[1], f[2], f[3], f[4], weights);
}
void nso_weno5_right(const double f[5], int weights, double out[1]) {
    out[0] = weno_right(f[0], f[1], f[2], f[3], f[4], weights);
}

static const double unit_normal[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};

/* ConvectionFlux::form_flux, convection_flux.cpp:23-61 (all cells) */
static void form_flux(const nso_cfg* c, const double* W, int dir, double* F) {
    pad_t        P  = pad_of(c);
    double       g  = c->gamma;
    const double nx = unit_normal[dir][0], ny = unit_normal[dir][1], nz = unit_normal[dir][2];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < P.NX; i++)
        for (long j = 0; j < P.NY; j++)
            for (long k = 0; k < P.NZ; k++) {
                const double* w   = &AT(W, P, i, j, k, 0);
                double*       f   = &AT(F, P, i, j, k, 0);
                double        rho = w[0], p = w[1], u = w[2], v = w[3], ww = w[4];
                double        q_n = u * nx + v * ny + ww * nz;
                double        cs  = sqrt(g * p / rho);
                double        H   = cs * cs / (g - 1.0) + 0.5 * (u * u + v * v + ww * ww);
                f[0]              = rho * q_n;
                f[1]              = rho * q_n * H;
                f[2]              = rho * q_n * u + p * nx;
                f[3]              = rho * q_n * v + p * ny;
                f[4]              = rho * q_n * ww + p * nz;
            }
}

/* ConvectionFlux::max_eigenvalue(Array4d...), convection_flux.cpp:213-249: the stored value is
 * the constant 3.0, so Weno5::get_max_alpha (weno-5.cpp:324-330) of five of them is 3.0. */
static const double lf_alpha = 3.0;

/* the eigenvalue convection_flux.cpp:235-238 computes and leaves unused (alpha_mode LOCAL_MAX):
 * max(|q_n - c|, |q_n + c|), std::max(a, b) = (a < b) ? b : a */
static double cell_eig(double g, const double* w, int dir) {
    double q_n  = w[2] * unit_normal[dir][0] + w[3] * unit_normal[dir][1] + w[4] * unit_normal[dir][2];
    double c    = sqrt(g * w[1] / w[0]);
    double eig1 = fabs(q_n - c);
    double eig2 = fabs(q_n + c);
    return (eig1 < eig2) ? eig2 : eig1;
}

static double max_d(double a, double b) { return (a < b) ? b : a; }

/* Weno5::get_max_alpha (weno-5.cpp:324-330) over the cells i-2 .. i+2 of face i+1/2 */
static double face_alpha(const nso_cfg* c, const double* W, size_t base, size_t s, int dir) {
    if (c->alpha_mode != NSO_ALPHA_LOCAL_MAX) return lf_alpha;
    double d[5];
    for (int ii = 0; ii < 5; ii++) d[ii] = cell_eig(c->gamma, &W[base + (ii - 2) * s], dir);
    return max_d(max_d(max_d(max_d(d[0], d[1]), d[2]), d[3]), d[4]);
}

/* ConvectionFlux::orthogonal_unit_vectors, convection_flux.cpp:63-115 */
static void orthogonal_unit_vectors(const double n[3], double out[6]) {
    double nx = n[0], ny = n[1], nz = n[2];
    double tempx = ny * ny + nz * nz;
    double tempy = nz * nz + nx * nx;
    double tempz = nx * nx + ny * ny;
    double lx = 0, ly = 0, lz = 0, mx, my, mz;
    if (tempx >= tempy && tempx >= tempz) {
        lx = 0.0;
        ly = -nz;
        lz = ny;
    } else if (tempy >= tempx && tempy >= tempz) {
        lx = -nz;
        ly = 0.0;
        lz = nx;
    } else if (tempz >= tempx && tempz >= tempy) {
        lx = -ny;
        ly = nx;
        lz = 0.0;
    }
    double temp = sqrt(lx * lx + ly * ly + lz * lz);
    lx          = lx / temp;
    ly          = ly / temp;
    lz          = lz / temp;
    mx          = ny * lz - nz * ly;
    my          = nz * lx - nx * lz;
    mz          = nx * ly - ny * lx;
    double am   = sqrt(mx * mx + my * my + mz * mz);
    mx          = mx / am;
    my          = my / am;
    mz          = mz / am;
    out[0] = lx, out[1] = ly, out[2] = lz, out[3] = mx, out[4] = my, out[5] = mz;
}

/* ConvectionFlux::eigenvectors, convection_flux.cpp:117-175 (c2 = gamma/(p*rho) as written;
 * true_c: gamma*p/rho) */
static void eigenvectors(double rho, double p, double u, double v, double w, double gamma,
                         const double n[3], const double l[3], const double m[3], double* L,
                         double* R, int true_c) {
    double nx = n[0], ny = n[1], nz = n[2];
    double lx = l[0], ly = l[1], lz = l[2];
    double mx = m[0], my = m[1], mz = m[2];
    double c2   = true_c ? gamma * p / rho : gamma / (p * rho);
    double c    = sqrt(c2);
    double qn   = u * nx + v * ny + w * nz;
    double ql   = u * lx + v * ly + w * lz;
    double qm   = u * mx + v * my + w * mz;
    double q2   = u * u + v * v + w * w;
    double H    = c2 / (gamma - 1.0) + 0.5 * q2;
    double K    = gamma - 1.0;
    double K_pc = K / (2.0 * c2);
    double RR[25] = {1.0,        0.0, 1.0,      1.0,        0.0,
                     u - c * nx, mx,  u,        u + c * nx, lx,
                     v - c * ny, my,  v,        v + c * ny, ly,
                     w - c * nz, mz,  w,        w + c * nz, lz,
                     H - qn * c, qm,  0.5 * q2, H + qn * c, ql};
    double LL[25] = {0.5 * K_pc * q2 + qn / (2.0 * c),
                     -(K_pc * u + nx / (2.0 * c)),
                     -(K_pc * v + ny / (2.0 * c)),
                     -(K_pc * w + nz / (2.0 * c)),
                     K_pc,
                     -qm,
                     mx,
                     my,
                     mz,
                     0.0,
                     1.0 - K_pc * q2,
                     K * u / c2,
                     K * v / c2,
                     K * w / c2,
                     -K / c2,
                     0.5 * K_pc * q2 - qn / (2.0 * c),
                     -(K_pc * u - nx / (2.0 * c)),
                     -(K_pc * v - ny / (2.0 * c)),
                     -(K_pc * w - nz / (2.0 * c)),
                     K_pc,
                     -ql,
                     lx,
                     ly,
                     lz,
                     0.0};
    memcpy(R, RR, sizeof RR);
    memcpy(L, LL, sizeof LL);
}

void nso_eigenvectors(const double Wavg[5], double gamma, int dir, double L[25], double R[25]) {
    double o[6];
    orthogonal_unit_vectors(unit_normal[dir], o);
    eigenvectors(Wavg[0], Wavg[1], Wavg[2], Wavg[3], Wavg[4], gamma, unit_normal[dir], o, o + 3, L,
                 R, 0);
}

/* Roe average of two primitive states (physics CONVECTION): sqrt(rho)-weighted velocity and
 * total enthalpy, rho = sqrt(rho_l rho_r), c^2 = (gamma - 1)(H - |u|^2 / 2), returned as the
 * primitive state [rho, p, u, v, w] with p = rho c^2 / gamma */
static void roe_state(double g, const double* l, const double* r, double* out) {
    double sl  = sqrt(l[0]), sr = sqrt(r[0]);
    double den = sl + sr;
    double hl  = g / (g - 1.0) * (l[1] / l[0]) + 0.5 * (l[2] * l[2] + l[3] * l[3] + l[4] * l[4]);
    double hr  = g / (g - 1.0) * (r[1] / r[0]) + 0.5 * (r[2] * r[2] + r[3] * r[3] + r[4] * r[4]);
    double u   = (sl * l[2] + sr * r[2]) / den;
    double v   = (sl * l[3] + sr * r[3]) / den;
    double w   = (sl * l[4] + sr * r[4]) / den;
    double H   = (sl * hl + sr * hr) / den;
    double c2  = (g - 1.0) * (H - 0.5 * (u * u + v * v + w * w));
    double rho = sl * sr;
    out[0] = rho, out[1] = rho * c2 / g, out[2] = u, out[3] = v, out[4] = w;
}

/* conserved index p of [rho, rho u, rho v, rho w, rho E] -> stored index [rho, rho E, ...] */
static const int cons_to_stored[5] = {0, 2, 3, 4, 1};

/* the face state and eigenvectors of the characteristic sweep, with L and R permuted to act on
 * stored-order vectors when the convection is corrected */
static void face_eigensystem(const nso_cfg* c, const double* wl, const double* wr, int dir,
                             const double o[6], double* Wf, double* L, double* R) {
    int corr = (c->physics & NSO_PHYS_CONVECTION) != 0;
    if (corr) roe_state(c->gamma, wl, wr, Wf);
    else
        for (int q = 0; q < NQ; q++) Wf[q] = 0.5 * (wl[q] + wr[q]);
    double Lc[25], Rc[25];
    eigenvectors(Wf[0], Wf[1], Wf[2], Wf[3], Wf[4], c->gamma, unit_normal[dir], o, o + 3, Lc, Rc,
                 corr);
    if (!corr) {
        memcpy(L, Lc, sizeof Lc), memcpy(R, Rc, sizeof Rc);
        return;
    }
    /* L_s[q][cons_to_stored[p]] = L[q][p]; R_s[cons_to_stored[p]][q] = R[p][q] */
    for (int q = 0; q < 5; q++)
        for (int p = 0; p < 5; p++) {
            L[q * 5 + cons_to_stored[p]] = Lc[q * 5 + p];
            R[cons_to_stored[p] * 5 + q] = Rc[p * 5 + q];
        }
}

void nso_face_eigensystem(const nso_cfg* c, const double Wl[5], const double Wr[5], int dir,
                          double Wface[5], double L[25], double R[25]) {
    double o[6];
    orthogonal_unit_vectors(unit_normal[dir], o);
    face_eigensystem(c, Wl, Wr, dir, o, Wface, L, R);
}

/* Weno5cw::compute_num_flux, weno-5cw.cpp:4-28, with matrix::square_matrix_matrix_product and
 * square_matrix_vector_product (matrix.cpp:4-20) */
static void compute_num_flux(const double fL[25], const double fR[25], const double L[25],
                             const double R[25], int type, double* flux) {
    double fL_[25], fR_[25];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            fL_[i * 5 + j] = 0.0;
            for (int k = 0; k < 5; k++) fL_[i * 5 + j] += L[i * 5 + k] * fL[k * 5 + j];
        }
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            fR_[i * 5 + j] = 0.0;
            for (int k = 0; k < 5; k++) fR_[i * 5 + j] += L[i * 5 + k] * fR[k * 5 + j];
        }
    double ft[5];
    for (int q = 0; q < 5; q++) {
        const double* a  = &fL_[q * 5];
        const double* b  = &fR_[q * 5];
        double        fl = weno_left(a[0], a[1], a[2], a[3], a[4], type);
        double        fr = weno_right(b[0], b[1], b[2], b[3], b[4], type);
        ft[q]            = fl + fr;
    }
    for (int i = 0; i < 5; i++) {
        flux[i] = 0.0;
        for (int j = 0; j < 5; j++) flux[i] += R[i * 5 + j] * ft[j];
    }
}

/* stride (in doubles) of one step along axis dir */
static size_t axis_stride(pad_t P, int dir) {
    if (dir == 0) return (size_t)P.NY * P.NZ * NQ;
    if (dir == 1) return (size_t)P.NZ * NQ;
    return NQ;
}

/* Weno5::{x,y,z}_sweep (weno-5.cpp:75-220) and Weno5cw::{x,y,z}_sweep (weno-5cw.cpp:30-262):
 * face value f(i) = numerical flux at i+1/2, for the sweep index from begin-1 to end-1 and the
 * other two indices over the interior. */
static void sweep(const nso_cfg* c, int dir, int charwise, const double* F, const double* U,
                  const double* W, double* f) {
    pad_t  P  = pad_of(c);
    size_t s  = axis_stride(P, dir);
    long   lo[3] = {NGC, NGC, NGC};
    long   hi[3] = {NGC + c->nx, NGC + c->ny, NGC + c->nz};
    lo[dir] -= 1;
    double o[6];
    orthogonal_unit_vectors(unit_normal[dir], o);
#pragma omp parallel for schedule(static)
    for (long i = lo[0]; i < hi[0]; i++)
        for (long j = lo[1]; j < hi[1]; j++)
            for (long k = lo[2]; k < hi[2]; k++) {
                size_t base = ((((size_t)i) * P.NY + j) * P.NZ + k) * NQ;
                const double alpha = face_alpha(c, W, base, s, dir);
                if (!charwise) {
                    double fl[NQ][5], fr[NQ][5];
                    for (int q = 0; q < NQ; q++)
                        for (int ii = 0; ii < 5; ii++) {
                            size_t a  = base + (size_t)(ii - 2) * s + q; /* idx = i-2+ii */
                            size_t b  = a + s;                           /* idx + 1 */
                            fl[q][ii] = (F[a] + alpha * U[a]) * 0.5;
                            fr[q][ii] = (F[b] - alpha * U[b]) * 0.5;
                        }
                    /* learned weights: momentum stencils are also measured against the face's
                     * momentum scale (no-op for the classic weights) */
                    const double fs = nso_mlp_face_scale(fl, fr);
                    for (int q = 0; q < NQ; q++) {
                        const double* l  = fl[q];
                        const double* r  = fr[q];
                        const double  sc = q >= 2 ? fs : 0.0;
                        double left  = weno_left_s(l[0], l[1], l[2], l[3], l[4], c->weights, sc);
                        double right = weno_right_s(r[0], r[1], r[2], r[3], r[4], c->weights, sc);
                        f[base + q]  = left + right;
                    }
                } else if (c->physics & NSO_PHYS_CONVECTION) {
                    /* the same product in the conserved order: fL[p][ii] holds the stored
                     * component cons_to_stored[p], the result lands back in stored order */
                    double Wf[5], L[25], R[25], fL[25], fR[25], out[5];
                    double wl[5], wr[5];
                    for (int q = 0; q < NQ; q++) wl[q] = W[base + q], wr[q] = W[base + s + q];
                    roe_state(c->gamma, wl, wr, Wf);
                    eigenvectors(Wf[0], Wf[1], Wf[2], Wf[3], Wf[4], c->gamma, unit_normal[dir], o,
                                 o + 3, L, R, 1);
                    for (int ii = 0; ii < 5; ii++)
                        for (int p = 0; p < NQ; p++) {
                            size_t a       = base + (size_t)(ii - 2) * s + cons_to_stored[p];
                            size_t b       = a + s;
                            fL[p * 5 + ii] = (F[a] + alpha * U[a]) * 0.5;
                            fR[p * 5 + ii] = (F[b] - alpha * U[b]) * 0.5;
                        }
                    compute_num_flux(fL, fR, L, R, c->weights, out);
                    for (int p = 0; p < NQ; p++) f[base + cons_to_stored[p]] = out[p];
                } else {
                    double Wavg[5], L[25], R[25], fL[25], fR[25], out[5];
                    for (int q = 0; q < NQ; q++) Wavg[q] = 0.5 * (W[base + q] + W[base + s + q]);
                    eigenvectors(Wavg[0], Wavg[1], Wavg[2], Wavg[3], Wavg[4], c->gamma,
                                 unit_normal[dir], o, o + 3, L, R, 0);
                    for (int ii = 0; ii < 5; ii++)
                        for (int q = 0; q < NQ; q++) {
                            size_t a       = base + (size_t)(ii - 2) * s + q;
                            size_t b       = a + s;
                            fL[q * 5 + ii] = (F[a] + alpha * U[a]) * 0.5;
                            fR[q * 5 + ii] = (F[b] - alpha * U[b]) * 0.5;
                        }
                    compute_num_flux(fL, fR, L, R, c->weights, out);
                    for (int q = 0; q < NQ; q++) f[base + q] = out[q];
                }
            }
}

/* Weno5::compute_residual, weno-5.cpp:39-73 */
static void conv_residual(const nso_cfg* c, const double* fx, const double* fy, const double* fz,
                          double* R) {
    pad_t  P  = pad_of(c);
    double dx = c->Lx / c->nx, dy = c->Ly / c->ny, dz = c->Lz / c->nz;
    size_t sx = axis_stride(P, 0), sy = axis_stride(P, 1), sz = axis_stride(P, 2);
#pragma omp parallel for schedule(static)
    for (long i = NGC; i < NGC + c->nx; i++)
        for (long j = NGC; j < NGC + c->ny; j++)
            for (long k = NGC; k < NGC + c->nz; k++)
                for (int q = 0; q < NQ; q++) {
                    size_t a = ((((size_t)i) * P.NY + j) * P.NZ + k) * NQ + q;
                    R[a] += (fx[a] - fx[a - sx]) / dx;
                    R[a] += (fy[a] - fy[a - sy]) / dy;
                    R[a] += (fz[a] - fz[a - sz]) / dz;
                }
}

/* Weno5::solve (weno-5.cpp:3-37) / Weno5cw: x, y, then z only when the grid is 3-D */
static void convection(const nso_cfg* c, const double* W, const double* U, double* R) {
    double* F  = alloc_zero(c);
    double* fx = alloc_zero(c);
    double* fy = alloc_zero(c);
    double* fz = alloc_zero(c);
    int     cw = (c->conv == NSO_CONV_WENO5CW);
    form_flux(c, W, 0, F);
    sweep(c, 0, cw, F, U, W, fx);
    form_flux(c, W, 1, F);
    sweep(c, 1, cw, F, U, W, fy);
    if (c->dim == 3) {
        form_flux(c, W, 2, F);
        sweep(c, 2, cw, F, U, W, fz);
    }
    conv_residual(c, fx, fy, fz, R);
    free(F), free(fx), free(fy), free(fz);
}

/* ------------------------------------------------------------------------------------------ */
/* Diffusion                                                                                  */
/* ------------------------------------------------------------------------------------------ */

/* cd_constants.hpp:9-27 */
static const double cd2_df[3]  = {-1. / 2, 0., 1. / 2.};
static const double cd2_ddf[3] = {1., -2., 1.};
static const double cd4_df[5]  = {1. / 12., -2. / 3., 0., 2. / 3., -1. / 12.};
static const double cd4_ddf[5] = {-1. / 12., 4. / 3., -5. / 2., 4. / 3., -1. / 12.};
static const double cd6_df[7]  = {-1. / 60., 3. / 20., -3. / 4., 0., 3. / 4., -3. / 20., 1. / 60.};
static const double cd6_ddf[7] = {1. / 90., -3. / 20., 3. / 2., -49. / 18., 3. / 2., -3. / 20., 1. / 90.};

static const double* fd_coeffs(int order, int second, int* width) {
    *width = order + 1;
    if (order == 2) return second ? cd2_ddf : cd2_df;
    if (order == 4) return second ? cd4_ddf : cd4_df;
    return second ? cd6_ddf : cd6_df;
}

void nso_fd(const nso_cfg* c, const double* f, double* df, int dir, int second, double d) {
    pad_t         P = pad_of(c);
    int           width;
    const double* co    = fd_coeffs(c->diff_order, second, &width);
    long          start = -(long)(width / 2);
    size_t        s     = axis_stride(P, dir);
#pragma omp parallel for schedule(static)
    for (long i = NGC; i < NGC + c->nx; i++)
        for (long j = NGC; j < NGC + c->ny; j++)
            for (long k = NGC; k < NGC + c->nz; k++)
                for (int q = 0; q < NQ; q++) {
                    size_t a   = ((((size_t)i) * P.NY + j) * P.NZ + k) * NQ + q;
                    double sum = 0.;
                    for (int idx = 0; idx < width; idx++)
                        sum += f[a + (size_t)(start + idx) * s] * co[idx] / d;
                    df[a] = sum;
                }
}

/* DiffusionScheme helpers, diffusion_scheme.cpp:3-58 */
static double tauxx(double mu, double du_dx, double dv_dy, double lamb) {
    return 2.0 * mu * du_dx + lamb * (du_dx + dv_dy);
}
static double tauxy(double mu, double du_dy, double dv_dx) { return mu * (du_dy + dv_dx); }
static double dtauxx_dx(double mu, double dmu_dx, double du_dx, double dv_dy, double du_dxdx,
                        double dv_dydx, double lamb, double dlamb_dx) {
    double ret = 2.0 * dmu_dx * du_dx + 2.0 * mu * du_dxdx + dlamb_dx * (du_dx + dv_dy);
    ret += lamb * (du_dxdx + dv_dydx);
    return ret;
}
static double dtauyy_dy(double mu, double dmu_dy, double dv_dy, double dv_dydy, double du_dx,
                        double du_dxdy, double lamb, double dlamb_dy) {
    double ret = 2.0 * dmu_dy * dv_dy + 2.0 * mu * dv_dydy + dlamb_dy * (du_dx + dv_dy);
    ret += lamb * (du_dxdy + dv_dydy);
    return ret;
}
static double dtauxy_dx(double mu, double dmu_dx, double du_dy, double dv_dx, double du_dydx,
                        double dv_dxdx) {
    return dmu_dx * (du_dy + dv_dx) + mu * (du_dydx + dv_dxdx);
}
static double dtauyx_dy(double mu, double dmu_dy, double du_dy, double dv_dx, double du_dydy,
                        double dv_dxdy) {
    return dmu_dy * (du_dy + dv_dx) + mu * (du_dydy + dv_dxdy);
}

/* CenteredDiffusion::compute_residual, centered_diffusion.cpp:70-182 (2-D terms, quirks kept:
 * Tauyy = tauxx(...), Tauxy*dv_dy, the heat-flux y term uses the enum value p = 1) */
static void diff_residual(const nso_cfg* c, const double* Wp, const double* Dx, const double* Dy,
                          const double* Dxx, const double* Dyy, const double* Dxy,
                          const double* Dyx, const double* MU, const double* DMUX,
                          const double* DMUY, double* Rs) {
    pad_t        P = pad_of(c);
    const int    rho = 0, p = 1, u = 2, v = 3, w = 4;
    const double m_mu = c->mu, m_gamma = c->gamma, m_Pr = c->Pr;
#pragma omp parallel for schedule(static)
    for (long i = NGC; i < NGC + c->nx; i++)
        for (long j = NGC; j < NGC + c->ny; j++)
            for (long k = NGC; k < NGC + c->nz; k++) {
                size_t b        = ((((size_t)i) * P.NY + j) * P.NZ + k) * NQ;
                double mu       = MU ? MU[b] : m_mu;
                double dmu_dx   = MU ? DMUX[b] : 0.0;
                double dmu_dy   = MU ? DMUY[b] : 0.0;
                double lamb     = (-2. / 3.) * mu;
                double dlamb_dx = -(2. / 3.) * dmu_dx;
                double dlamb_dy = -(2. / 3.) * dmu_dy;
                double U = Wp[b + u], V = Wp[b + v], Pp = Wp[b + p], RHO = Wp[b + rho];
                double du_dx = Dx[b + u], du_dy = Dy[b + u], du_dxdx = Dxx[b + u];
                double du_dydy = Dyy[b + u], du_dxdy = Dxy[b + u], du_dydx = Dyx[b + u];
                double dv_dy = Dy[b + v], dv_dx = Dx[b + v], dv_dxdx = Dxx[b + v];
                double dv_dydy = Dyy[b + v], dv_dydx = Dyx[b + v], dv_dxdy = Dxy[b + v];
                double drho_dx = Dx[b + rho], dp_dx = Dx[b + p], drho_dxdx = Dxx[b + rho];
                double dp_dxdx = Dxx[b + p];
                double drho_dy = Dy[b + rho], dp_dy = Dy[b + p], drho_dydy = Dyy[b + rho];
                double dp_dydy = Dyy[b + p];

                double Dtauxx_dx =
                    dtauxx_dx(mu, dmu_dx, du_dx, dv_dy, du_dxdx, dv_dydx, lamb, dlamb_dx);
                double Dtauyy_dy =
                    dtauyy_dy(mu, dmu_dy, dv_dy, dv_dydy, du_dx, du_dxdy, lamb, dlamb_dy);
                double Dtauxy_dx = dtauxy_dx(mu, dmu_dx, du_dy, dv_dx, du_dydx, dv_dxdx);
                double Dtauyx_dy = dtauyx_dy(mu, dmu_dy, du_dy, dv_dx, du_dydy, dv_dxdy);
                double Tauxx     = tauxx(mu, du_dx, dv_dy, lamb);
                double Tauyy     = tauxx(mu, du_dx, dv_dy, lamb);
                double Tauxy     = tauxy(mu, du_dy, dv_dx);
                double Tauyx     = Tauxy;

                double Dtauvx_dx = U * Dtauxx_dx + V * Dtauxy_dx + Tauxx * du_dx + Tauxy * dv_dy;
                double Dtauvy_dy = U * Dtauyx_dy + V * Dtauyy_dy + Tauyy * dv_dy + Tauyx * du_dy;

                double one_rho  = 1. / RHO;
                double one_rho2 = 1. / (RHO * RHO);
                double one_rho3 = 1. / (RHO * RHO * RHO);

                double temp1 = dmu_dx * (one_rho * dp_dx - Pp * one_rho2 * drho_dx);
                double temp2 = mu * (2.0 * one_rho2 * drho_dx * dp_dx + one_rho * dp_dxdx +
                                     one_rho3 * Pp * drho_dx * drho_dx - one_rho2 * Pp * drho_dxdx);
                double temp3  = m_gamma / (m_Pr * (m_gamma - 1.0));
                double Dqx_dx = -temp3 * (temp1 + temp2);

                temp1 = dmu_dy * (one_rho * dp_dy - Pp * one_rho2 * drho_dy);
                temp2 = mu * (2.0 * one_rho2 * drho_dy * dp_dy + one_rho * dp_dydy +
                              one_rho3 * p * drho_dy * drho_dy - one_rho2 * Pp * drho_dydy);
                double Dqy_dy = -temp3 * (temp1 + temp2);

                double x_mom = -Dtauxx_dx - Dtauyx_dy;
                double y_mom = -Dtauyy_dy - Dtauxy_dx;
                double ene   = -Dtauvx_dx - Dtauvy_dy + Dqx_dx + Dqy_dy;

                Rs[b + p] += ene;
                Rs[b + u] += x_mom;
                Rs[b + v] += y_mom;
                Rs[b + w] -= 0.0;
            }
}

/* CenteredDiffusion::solve, centered_diffusion.cpp:3-68.  Only the derivative arrays that
 * compute_residual reads are formed; the z-derivatives are computed there but never used.
 * ddf_dydy runs the x stencil scaled by dy*dy (cd-n.cpp:74-79).  The first-derivative arrays are
 * fresh (zero halos) and get the periodic face exchange (centered_diffusion.cpp:22-24), so their
 * halos are the periodic wrap, or zero beyond a physical wall. */
/* Sutherland law, Viscosity::mu_sutherland (viscosity.cpp:5-11): mu0 (T0 + C)/(T + C) (T/T0)^1.5
 * with T = gamma p / rho and T0 = 1.  The reference defines the law but never calls it
 * (viscosity.cpp:3 has an empty body); here T^1.5 is written T*sqrt(T) (two IEEE operations, so
 * the GPU reproduces it bit for bit) and the viscosity derivatives feed the dmu slots that
 * compute_residual already has (centered_diffusion.cpp:96-103). */
double nso_mu_sutherland(double mu0, double C, double gamma, double rho, double p) {
    double T = gamma * p / rho;
    return mu0 * (1.0 + C) / (T + C) * (T * sqrt(T));
}

static void diffusion(const nso_cfg* c, const double* W, double* R) {
    double  dx = c->Lx / c->nx, dy = c->Ly / c->ny;
    double* Dxx = alloc_zero(c);
    double* Dyy = alloc_zero(c);
    double* Dx  = alloc_zero(c);
    double* Dy  = alloc_zero(c);
    double* Dxy = alloc_zero(c);
    double* Dyx = alloc_zero(c);
    nso_fd(c, W, Dxx, 0, 1, dx * dx);
    nso_fd(c, W, Dyy, 0, 1, dy * dy);
    nso_fd(c, W, Dx, 0, 0, dx);
    nso_fd(c, W, Dy, 1, 0, dy);
    communicate(c, Dx);
    communicate(c, Dy);
    nso_fd(c, Dx, Dxy, 1, 0, dy); /* dW_dxdy = df_dy(dW_dx) */
    nso_fd(c, Dy, Dyx, 0, 0, dx); /* dW_dydx = df_dx(dW_dy) */
    double *MU = NULL, *DMUX = NULL, *DMUY = NULL;
    if (c->viscosity == 1) {
        /* mu on the interior and the face halos (from W there), its derivatives on the interior
         * with the first-derivative stencil; stored in slot 0 of 5-variable arrays */
        pad_t P = pad_of(c);
        MU      = alloc_zero(c);
        DMUX    = alloc_zero(c);
        DMUY    = alloc_zero(c);
        for (long i = 0; i < P.NX; i++)
            for (long j = 0; j < P.NY; j++)
                for (long k = 0; k < P.NZ; k++) {
                    int out = (i < NGC || i >= NGC + c->nx) + (j < NGC || j >= NGC + c->ny) +
                              (k < NGC || k >= NGC + c->nz);
                    if (out > 1) continue;
                    const double* w  = &AT(W, P, i, j, k, 0);
                    AT(MU, P, i, j, k, 0) = nso_mu_sutherland(c->mu, c->suth_C, c->gamma, w[0], w[1]);
                }
        nso_fd(c, MU, DMUX, 0, 0, dx);
        nso_fd(c, MU, DMUY, 1, 0, dy);
    }
    diff_residual(c, W, Dx, Dy, Dxx, Dyy, Dxy, Dyx, MU, DMUX, DMUY, R);
    free(Dxx), free(Dyy), free(Dx), free(Dy), free(Dxy), free(Dyx);
    free(MU), free(DMUX), free(DMUY);
}

/* Full 3-D viscous terms (physics VISCOUS): what CenteredDiffusion::compute_residual
 * (centered_diffusion.cpp:70-182) computes in 2-D, for all three axes and without the quirks of
 * the reference's expressions (each second derivative on its own axis, Tauyy from dv/dy, the
 * viscous work from tau_ij du_i/dx_j, the heat flux from the temperature p/rho with its true
 * derivatives).  Expanded like the reference's: with g[a][d] = du_a/dx_d, div = tr g,
 * lamb = -2/3 mu,
 *   tau[a][d]  = mu (g[a][d] + g[d][a]) + [a == d] lamb div
 *   divtau[a]  = sum_d dmu_d (g[a][d] + g[d][a]) + mu (lap u_a + d_a div)
 *                + dlamb_a div + lamb d_a div
 *   energy     = -(sum_a u_a divtau[a]) - sum_{a,d} tau[a][d] g[a][d] + sum_d dq_d/dx_d,
 *                dq_d/dx_d = -heat (dmu_d dT/dx_d + mu d2T/dx_d2),  T = p / rho
 * Derivative arrays: first derivatives of (u, v, w, T) on the interior, exchanged like
 * centered_diffusion.cpp:22-24 (periodic wrap, zero beyond a wall), pure second derivatives on
 * their own axis, mixed second derivatives d_a d_d u_d = df_da(du_d/dx_d). */
static void visc3d_residual(const nso_cfg* c, const double* W, const double* G, const double* D[3],
                            const double* S[3], const double* X[3][3], const double* MU,
                            const double* DMU[3], double* R) {
    pad_t        P    = pad_of(c);
    const double heat = c->gamma / (c->Pr * (c->gamma - 1.0));
#pragma omp parallel for schedule(static)
    for (long i = NGC; i < NGC + c->nx; i++)
        for (long j = NGC; j < NGC + c->ny; j++)
            for (long k = NGC; k < NGC + c->nz; k++) {
                size_t b  = ((((size_t)i) * P.NY + j) * P.NZ + k) * NQ;
                double mu = MU ? MU[b] : c->mu;
                double dmu[3] = {0.0, 0.0, 0.0};
                if (MU)
                    for (int d = 0; d < 3; d++) dmu[d] = DMU[d][b];
                double lamb = (-2. / 3.) * mu;
                double g[3][3], lap[3], ddiv[3];
                for (int a = 0; a < 3; a++)
                    for (int d = 0; d < 3; d++) g[a][d] = D[d][b + a];
                double div = (g[0][0] + g[1][1]) + g[2][2];
                for (int a = 0; a < 3; a++) {
                    lap[a]  = (S[0][b + a] + S[1][b + a]) + S[2][b + a];
                    double x[3];
                    for (int d = 0; d < 3; d++) x[d] = (d == a) ? S[a][b + a] : X[a][d][b + d];
                    ddiv[a] = (x[0] + x[1]) + x[2];
                }
                double divtau[3];
                for (int a = 0; a < 3; a++) {
                    double t = mu * (lap[a] + ddiv[a]);
                    if (MU) {
                        double m = dmu[0] * (g[a][0] + g[0][a]);
                        m += dmu[1] * (g[a][1] + g[1][a]);
                        m += dmu[2] * (g[a][2] + g[2][a]);
                        t = (m + t) + (-(2. / 3.) * dmu[a]) * div;
                    }
                    divtau[a] = t + lamb * ddiv[a];
                }
                const double* w    = &W[b];
                double        work = (w[2] * divtau[0] + w[3] * divtau[1]) + w[4] * divtau[2];
                double        phi  = 0.0;
                for (int a = 0; a < 3; a++)
                    for (int d = 0; d < 3; d++) {
                        double tau = mu * (g[a][d] + g[d][a]);
                        if (a == d) tau = tau + lamb * div;
                        phi = (a == 0 && d == 0) ? tau * g[a][d] : phi + tau * g[a][d];
                    }
                double q[3];
                for (int d = 0; d < 3; d++) {
                    double Tg = D[d][b + 3], Tdd = S[d][b + 3];
                    double t  = mu * Tdd;
                    if (MU) t = dmu[d] * Tg + t;
                    q[d] = -heat * t;
                }
                double ene = ((-work) - phi) + ((q[0] + q[1]) + q[2]);
                R[b + 1] += ene;
                R[b + 2] += -divtau[0];
                R[b + 3] += -divtau[1];
                R[b + 4] += -divtau[2];
            }
    (void)G;
}

static void diffusion_3d(const nso_cfg* c, const double* W, double* R) {
    pad_t  P     = pad_of(c);
    double dd[3] = {c->Lx / c->nx, c->Ly / c->ny, c->Lz / c->nz};
    /* G = (u, v, w, p / rho) wherever W is defined (interior and face halos) */
    double* G = alloc_zero(c);
    for (size_t a = 0; a < nso_array_len(c); a += NQ) {
        G[a + 0] = W[a + 2], G[a + 1] = W[a + 3], G[a + 2] = W[a + 4];
        G[a + 3] = (W[a] != 0.0) ? W[a + 1] / W[a] : 0.0;
    }
    double *D[3], *S[3], *X[3][3] = {{0}}, *MU = NULL, *DMU[3] = {NULL, NULL, NULL};
    for (int d = 0; d < 3; d++) {
        D[d] = alloc_zero(c);
        S[d] = alloc_zero(c);
        nso_fd(c, G, D[d], d, 0, dd[d]);
        nso_fd(c, G, S[d], d, 1, dd[d] * dd[d]);
        communicate(c, D[d]);
    }
    /* X[a][d] = df_da(dG/dx_d), a != d */
    for (int a = 0; a < 3; a++)
        for (int d = 0; d < 3; d++)
            if (a != d) {
                X[a][d] = alloc_zero(c);
                nso_fd(c, D[d], X[a][d], a, 0, dd[a]);
            }
    if (c->viscosity == 1) {
        MU = alloc_zero(c);
        for (long i = 0; i < P.NX; i++)
            for (long j = 0; j < P.NY; j++)
                for (long k = 0; k < P.NZ; k++) {
                    int out = (i < NGC || i >= NGC + c->nx) + (j < NGC || j >= NGC + c->ny) +
                              (k < NGC || k >= NGC + c->nz);
                    if (out > 1) continue;
                    const double* w       = &AT(W, P, i, j, k, 0);
                    AT(MU, P, i, j, k, 0) = nso_mu_sutherland(c->mu, c->suth_C, c->gamma, w[0], w[1]);
                }
        for (int d = 0; d < 3; d++) {
            DMU[d] = alloc_zero(c);
            nso_fd(c, MU, DMU[d], d, 0, dd[d]);
        }
    }
    visc3d_residual(c, W, G, (const double**)D, (const double**)S,
                    (const double* (*)[3])X, MU, (const double**)DMU, R);
    free(G), free(MU);
    for (int d = 0; d < 3; d++) {
        free(D[d]), free(S[d]), free(DMU[d]);
        for (int a = 0; a < 3; a++) free(X[a][d]);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* RHS and time stepping                                                                      */
/* ------------------------------------------------------------------------------------------ */

double nso_face_alpha(const nso_cfg* c, const double* W, int dir, long i, long j, long k) {
    pad_t P = pad_of(c);
    return face_alpha(c, W, ((((size_t)i) * P.NY + j) * P.NZ + k) * NQ, axis_stride(P, dir), dir);
}

static int cfg_ok(const nso_cfg* c) {
    if (c->nx < 1 || c->ny < 1 || c->nz < 1) return 0;
    if (c->conv < 0 || c->conv > 2) return 0;
    if (c->alpha_mode < 0 || c->alpha_mode > 1) return 0;
    if (c->physics & ~(NSO_PHYS_CONVECTION | NSO_PHYS_VISCOUS)) return 0;
    if ((c->physics & NSO_PHYS_VISCOUS) && c->dim != 3) return 0;
    if (c->diff_order != 0 && c->diff_order != 2 && c->diff_order != 4 && c->diff_order != 6)
        return 0;
    return 1;
}

int nso_rhs(const nso_cfg* c, double* U, double* W, double* R, int which) {
    if (!cfg_ok(c)) return -1;
    nso_fill_halos(c, U);
    nso_c2p(c, U, W);
    memset(R, 0, nso_array_len(c) * sizeof(double));
    if ((which & 1) && c->conv != NSO_CONV_NONE) convection(c, W, U, R);
    if ((which & 2) && c->diff_order != 0) {
        if (c->physics & NSO_PHYS_VISCOUS) diffusion_3d(c, W, R);
        else diffusion(c, W, R);
    }
    return 0;
}

/* TimeScheme::substep, time_scheme.cpp:72-96 */
static void substep(const nso_cfg* c, double* now, const double* prev, const double* inc,
                    double m1, double m2, double m3) {
    pad_t P = pad_of(c);
#pragma omp parallel for schedule(static)
    for (long i = NGC; i < NGC + c->nx; i++)
        for (long j = NGC; j < NGC + c->ny; j++)
            for (long k = NGC; k < NGC + c->nz; k++)
                for (int q = 0; q < NQ; q++) {
                    size_t a = ((((size_t)i) * P.NY + j) * P.NZ + k) * NQ + q;
                    now[a]   = (prev[a] * m1) + (now[a] * m2) + (inc[a] * m3);
                }
}

/* TimeScheme::step, time_scheme.cpp:3-70 */
int nso_step(const nso_cfg* c, double* W, double* U, int nsteps, double* dts) {
    if (!cfg_ok(c)) return -1;
    pad_t   P     = pad_of(c);
    double* R     = alloc_zero(c);
    double* Uprev = alloc_zero(c);
    for (int s = 0; s < nsteps; s++) {
        memset(Uprev, 0, nso_array_len(c) * sizeof(double));
        double dt = nso_dt(c, W);
        if (dts) dts[s] = dt;
        nso_p2c(c, W, U);
        /* assign_array (time_scheme.cpp:138-156), interior only */
        for (long i = NGC; i < NGC + c->nx; i++)
            for (long j = NGC; j < NGC + c->ny; j++)
                memcpy(&AT(Uprev, P, i, j, NGC, 0), &AT(U, P, i, j, NGC, 0),
                       sizeof(double) * NQ * c->nz);
        const double m[3][3] = {{1.0, 0.0, -dt},
                                {3.0 / 4.0, 1.0 / 4.0, -(1. / 4.) * dt},
                                {1. / 3., 2. / 3., -(2. / 3.) * dt}};
        for (int st = 0; st < 3; st++) {
            nso_rhs(c, U, W, R, 3);
            substep(c, U, Uprev, R, m[st][0], m[st][1], m[st][2]);
        }
        nso_c2p(c, U, W);
    }
    free(R), free(Uprev);
    return 0;
}
