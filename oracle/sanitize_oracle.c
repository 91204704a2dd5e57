/* sanitize_oracle.c -- drives the oracle (nsoracle.c, nsmlp.c) through every scheme, boundary and
 * option combination on small blocks, for the AddressSanitizer / UndefinedBehaviorSanitizer build
 * of oracle/Makefile (target `sanitize`).  TEST INFRASTRUCTURE ONLY.  Exits 0 when every call
 * returns normally and the results are finite; the sanitizers abort on the first finding. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "nsoracle.h"

static double frand(unsigned* s) {
    *s = *s * 1664525u + 1013904223u;
    return (double)(*s >> 8) / 16777216.0;
}

/* smooth state with noise: rho ~ 1, p ~ c0^2 / gamma, |u| < 0.5 */
static void fill(const nso_cfg* c, double* W, unsigned seed) {
    const int NX = c->nx + 6, NY = c->ny + 6, NZ = c->nz + 6;
    for (int i = 3; i < NX - 3; i++)
        for (int j = 3; j < NY - 3; j++)
            for (int k = 3; k < NZ - 3; k++) {
                double* w = W + (((size_t)i * NY + j) * NZ + k) * 5;
                w[0]      = 1.0 + 0.1 * sin(0.7 * i + 0.3 * j) + 0.01 * frand(&seed);
                w[1]      = 2.2 * 2.2 / 1.4 * (1.0 + 0.05 * cos(0.5 * j + 0.2 * k));
                w[2]      = 0.3 * sin(0.4 * k) + 0.02 * frand(&seed);
                w[3]      = 0.2 * cos(0.3 * i) + 0.02 * frand(&seed);
                w[4]      = 0.1 * sin(0.6 * j) + 0.02 * frand(&seed);
            }
}

static int finite_all(const double* a, size_t n) {
    for (size_t t = 0; t < n; t++)
        if (!isfinite(a[t])) return 0;
    return 1;
}

int main(void) {
    int fails = 0, runs = 0;
    const int convs[3] = {NSO_CONV_NONE, NSO_CONV_WENO5, NSO_CONV_WENO5CW};
    for (int conv = 0; conv < 3; conv++)
        for (int wt = 0; wt < 2; wt++)
            for (int ord = 0; ord <= 6; ord += 2)
                for (int bcs = 0; bcs < 3; bcs++)
                    for (int dim = 2; dim <= 3; dim++) {
                        nso_cfg c = {0};
                        c.nx = 7, c.ny = 6, c.nz = dim == 3 ? 5 : 1, c.dim = dim;
                        c.Lx = 1.0, c.Ly = 1.2, c.Lz = 0.8;
                        c.gamma = 1.4, c.mu = 2e-3, c.Pr = 0.72, c.cfl = 0.4;
                        c.conv = convs[conv], c.weights = wt, c.diff_order = ord;
                        for (int f = 0; f < 6; f++)
                            c.bc[f] = (bcs == 0) ? NSO_BC_PERIODIC
                                      : (bcs == 1) ? NSO_BC_NEUMANN
                                                   : ((f == 2 || f == 3) ? NSO_BC_NEUMANN
                                                                         : NSO_BC_PERIODIC);
                        c.viscosity = (ord == 4 && conv == 2) ? 1 : 0;
                        c.suth_C    = 0.368;
                        c.alpha_mode = (wt == 1 && conv > 0) ? NSO_ALPHA_LOCAL_MAX : 0;
                        c.physics    = (conv == 2 && ord == 2 && dim == 3) ? 3 : 0;
                        size_t  n    = nso_array_len(&c);
                        double* W    = calloc(n, sizeof(double));
                        double* U    = calloc(n, sizeof(double));
                        double* R    = calloc(n, sizeof(double));
                        double  dts[3];
                        fill(&c, W, 7u + (unsigned)runs);
                        nso_p2c(&c, W, U);
                        if (nso_rhs(&c, U, W, R, 3) != 0) fails++;
                        if (nso_step(&c, W, U, 3, dts) != 0) fails++;
                        if (!finite_all(dts, 3) || !(dts[2] > 0)) fails++;
                        runs++;
                        free(W), free(U), free(R);
                    }
    /* WENO kernels, eigenvectors and the learned weights' oracle */
    const double f[5] = {1.0, 1.3, 0.7, 2.0, 1.1};
    double       o[1], L[25], Rm[25], x[8], q[3], w[3];
    for (int wt = 0; wt < 2; wt++) nso_weno5_left(f, wt, o), nso_weno5_right(f, wt, o);
    const double Wavg[5] = {1.0, 0.7, 0.2, -0.1, 0.05};
    for (int d = 0; d < 3; d++) nso_eigenvectors(Wavg, 1.4, d, L, Rm);
    double* params = calloc(NSO_MLP_NPARAM, sizeof(double));
    unsigned s     = 3u;
    for (int t = 0; t < NSO_MLP_NPARAM; t++) params[t] = 0.2 * (frand(&s) - 0.5);
    nso_mlp_features(f, x);
    (void)nso_mlp_regime(x), (void)nso_mlp_blend(nso_mlp_ratio(x));
    nso_mlp_candidates(x, q);
    nso_mlp_forward(params, x, 1, w);
    const long n = 100;
    double *X = malloc(sizeof(double) * 8 * n), *Q = malloc(sizeof(double) * 3 * n),
           *T = malloc(sizeof(double) * n), *G = malloc(sizeof(double) * NSO_MLP_NPARAM);
    for (long i = 0; i < 8 * n; i++) X[i] = frand(&s);
    for (long i = 0; i < 3 * n; i++) Q[i] = frand(&s);
    for (long i = 0; i < n; i++) T[i] = frand(&s);
    double loss = nso_mlp_loss_grad(params, n, X, Q, T, 1, G);
    if (!isfinite(loss) || !finite_all(G, NSO_MLP_NPARAM)) fails++;
    for (int side = 0; side < 2; side++) nso_weno5_learned(params, 1, f, side, o, NULL);
    free(X), free(Q), free(T), free(G), free(params);
    printf("oracle sanitizer run: %d configurations, %d failures\n", runs, fails);
    return fails ? 1 : 0;
}
