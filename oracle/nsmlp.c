/*
 * nsmlp.c -- CPU oracle of the learned WENO-5 weights (an extension with no reference
 * counterpart): features, regime classification, the 8-32-32-3 GELU/softmax network in FP64
 * (optionally rounding its operands to bfloat16 exactly where the GPU kernels feed their BF16
 * matrix instructions), its loss and parameter gradient, and the learned reconstruction with its
 * classic-WENO fallback.
 *
 * TEST INFRASTRUCTURE ONLY (see nsoracle.h).
 *
 * Method.  The network replaces the nonlinear weights of WENO-5 (Jiang & Shu 1996) by the output
 * of a small MLP, as in WENO-NN (Stevens & Colonius, "Enhancement of shock-capturing methods via
 * machine learning", Theor. Comput. Fluid Dyn. 34, 2020): inputs are the five stencil values
 * normalised by their local range plus the three smoothness indicators of the normalised
 * values; the softmax output keeps the weights positive and summing to one (consistency).
 * Classic WENO-5 is kept on faces classified smooth (all indicators alike: the classic weights
 * are already the ideal ones) and on faces flagged as discontinuous (indicators far apart); in
 * between, the reconstruction is classic + blend(ratio) (network - classic), with a blend that
 * rises and falls continuously so the weights do not jump where a face changes regime.
 */
#include <math.h>
#include <string.h>

#include "nsoracle.h"

/* parameter layout (row-major): W1[32][8], b1[32], W2[32][32], b2[32], W3[3][32], b3[3] */
#define H NSO_MLP_HIDDEN
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + H * 8)
#define OFF_W2 (OFF_B1 + H)
#define OFF_B2 (OFF_W2 + H * H)
#define OFF_W3 (OFF_B2 + H)
#define OFF_B3 (OFF_W3 + 3 * H)

/* round to the nearest bfloat16, ties to even (the conversion of the GPU's v_cvt_pk_bf16_f32);
 * values are first rounded to float, as the GPU holds them in FP32 */
double nso_bf16(double v) {
    float    f = (float)v;
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return v; /* inf / nan unchanged */
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double rnd(double v, int bf16) { return bf16 ? nso_bf16(v) : v; }

/* GELU, tanh form (Hendrycks & Gimpel 2016): 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3))) */
static const double GELU_K = 0.7978845608028654; /* sqrt(2/pi) */
static const double GELU_A = 0.044715;
static double gelu(double x) { return 0.5 * x * (1.0 + tanh(GELU_K * (x + GELU_A * x * x * x))); }
static double gelu_d(double x) {
    double t = tanh(GELU_K * (x + GELU_A * x * x * x));
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_K * (1.0 + 3.0 * GELU_A * x * x);
}

/* smoothness indicators of a left-biased stencil (weno-5.cpp:231-244 formulas) */
static void betas(const double v[5], double b[3]) {
    double t1, t2;
    t1   = v[0] - 2. * v[1] + v[2];
    t2   = v[0] - 4. * v[1] + 3. * v[2];
    b[0] = (13. / 12.) * t1 * t1 + (1. / 4.) * t2 * t2;
    t1   = v[1] - 2. * v[2] + v[3];
    t2   = v[1] - v[3];
    b[1] = (13. / 12.) * t1 * t1 + (1. / 4.) * t2 * t2;
    t1   = v[2] - 2. * v[3] + v[4];
    t2   = 3. * v[2] - 4. * v[3] + v[4];
    b[2] = (13. / 12.) * t1 * t1 + (1. / 4.) * t2 * t2;
}

void nso_mlp_features_s(const double v[5], double scale, double x[8]) {
    double m = v[0], M = v[0];
    for (int k = 1; k < 5; k++) {
        m = v[k] < m ? v[k] : m;
        M = v[k] > M ? v[k] : M;
    }
    double r    = M - m;
    double am   = fabs(m), aM = fabs(M);
    double mg   = am > aM ? am : aM;
    mg          = scale > mg ? scale : mg;
    int    vary = r > NSO_MLP_RANGE * mg;
    for (int k = 0; k < 5; k++) x[k] = vary ? (v[k] - m) / r : 0.0;
    betas(x, x + 5);
}

void nso_mlp_features(const double v[5], double x[8]) { nso_mlp_features_s(v, 0.0, x); }

double nso_mlp_face_scale(const double fl[5][5], const double fr[5][5]) {
    double s = 0.0;
    for (int p = 2; p < 5; p++) {
        const double a[4] = {fl[p][2], fl[p][3], fr[p][1], fr[p][2]};
        for (int t = 0; t < 4; t++) s = fabs(a[t]) > s ? fabs(a[t]) : s;
    }
    return s;
}

double nso_mlp_ratio(const double x[8]) {
    double bmin = x[5], bmax = x[5];
    for (int k = 6; k < 8; k++) {
        bmin = x[k] < bmin ? x[k] : bmin;
        bmax = x[k] > bmax ? x[k] : bmax;
    }
    return (bmax + NSO_MLP_EPS) / (bmin + NSO_MLP_EPS);
}

/* share of the network: 0 at SMOOTH rising to 1 at 2 SMOOTH, 1 at FLAG / 4 falling to 0 at FLAG */
double nso_mlp_blend(double ratio) {
    double up   = (ratio - NSO_MLP_SMOOTH) / NSO_MLP_SMOOTH;
    double down = (NSO_MLP_FLAG - ratio) / (0.75 * NSO_MLP_FLAG);
    up          = up < 1.0 ? up : 1.0;
    down        = down < 1.0 ? down : 1.0;
    return (up > 0.0 ? up : 0.0) * (down > 0.0 ? down : 0.0);
}

int nso_mlp_regime(const double x[8]) {
    double bmin = x[5], bmax = x[5];
    for (int k = 6; k < 8; k++) {
        bmin = x[k] < bmin ? x[k] : bmin;
        bmax = x[k] > bmax ? x[k] : bmax;
    }
    double ratio = (bmax + NSO_MLP_EPS) / (bmin + NSO_MLP_EPS);
    if (ratio < NSO_MLP_SMOOTH) return NSO_REGIME_SMOOTH;
    if (ratio > NSO_MLP_FLAG) return NSO_REGIME_FLAGGED;
    return NSO_REGIME_MLP;
}

void nso_mlp_candidates(const double x[8], double q[3]) {
    q[0] = (2. * x[0] - 7. * x[1] + 11. * x[2]) / 6.;
    q[1] = (-x[1] + 5. * x[2] + 2. * x[3]) / 6.;
    q[2] = (2. * x[2] + 5. * x[3] - x[4]) / 6.;
}

typedef struct {
    double x[8], z1[H], h1[H], z2[H], h2[H], z3[3], w[3];
} mlp_cache;

/* forward pass; with bf16 the inputs of every matrix product (parameters, features, hidden
 * activations) are rounded to bfloat16 and the products accumulate exactly (the GPU accumulates
 * in FP32; biases enter the accumulator in FP32) */
static void forward(const double* p, const double xin[8], int bf16, mlp_cache* c) {
    for (int k = 0; k < 8; k++) c->x[k] = rnd(xin[k], bf16);
    for (int u = 0; u < H; u++) {
        double s = bf16 ? (double)(float)p[OFF_B1 + u] : p[OFF_B1 + u];
        for (int k = 0; k < 8; k++) s += rnd(p[OFF_W1 + u * 8 + k], bf16) * c->x[k];
        c->z1[u] = s;
        c->h1[u] = rnd(gelu(s), bf16);
    }
    for (int u = 0; u < H; u++) {
        double s = bf16 ? (double)(float)p[OFF_B2 + u] : p[OFF_B2 + u];
        for (int k = 0; k < H; k++) s += rnd(p[OFF_W2 + u * H + k], bf16) * c->h1[k];
        c->z2[u] = s;
        c->h2[u] = rnd(gelu(s), bf16);
    }
    double zmax = -1e300;
    for (int o = 0; o < 3; o++) {
        double s = bf16 ? (double)(float)p[OFF_B3 + o] : p[OFF_B3 + o];
        for (int k = 0; k < H; k++) s += rnd(p[OFF_W3 + o * H + k], bf16) * c->h2[k];
        c->z3[o] = s;
        zmax     = s > zmax ? s : zmax;
    }
    double e[3], se = 0;
    for (int o = 0; o < 3; o++) se += (e[o] = exp(c->z3[o] - zmax));
    for (int o = 0; o < 3; o++) c->w[o] = e[o] / se;
}

void nso_mlp_forward(const double* params, const double x[8], int bf16, double w[3]) {
    mlp_cache c;
    forward(params, x, bf16, &c);
    for (int o = 0; o < 3; o++) w[o] = c.w[o];
}

/* mean over n samples of (sum_k w_k Q_k - T)^2 and its gradient (grad may be NULL).  With bf16,
 * the backward pass rounds the operands of its matrix products like the GPU: dz3, dz2, dz1 and
 * the transposed parameters to bfloat16. */
double nso_mlp_loss_grad(const double* p, long n, const double* X, const double* Q,
                         const double* T, int bf16, double* grad) {
    double loss = 0;
    if (grad) memset(grad, 0, sizeof(double) * NSO_MLP_NPARAM);
    for (long s = 0; s < n; s++) {
        mlp_cache c;
        forward(p, X + 8 * s, bf16, &c);
        const double* q = Q + 3 * s;
        double        e = c.w[0] * q[0] + c.w[1] * q[1] + c.w[2] * q[2] - T[s];
        loss += e * e;
        if (!grad) continue;
        /* gradient of the sum (scaled by 1/n at the end, as the GPU does): dL/dw_k = 2 e q_k;
         * softmax: dz_o = w_o (dw_o - sum_k w_k dw_k) */
        double dw[3], dot = 0, dz3[3], dz2[H], dz1[H];
        for (int o = 0; o < 3; o++) dot += c.w[o] * (dw[o] = 2.0 * e * q[o]);
        for (int o = 0; o < 3; o++) dz3[o] = rnd(c.w[o] * (dw[o] - dot), bf16);
        for (int o = 0; o < 3; o++) {
            grad[OFF_B3 + o] += dz3[o];
            for (int k = 0; k < H; k++) grad[OFF_W3 + o * H + k] += dz3[o] * c.h2[k];
        }
        for (int k = 0; k < H; k++) {
            double s = 0;
            for (int o = 0; o < 3; o++) s += rnd(p[OFF_W3 + o * H + k], bf16) * dz3[o];
            dz2[k] = rnd(s * gelu_d(c.z2[k]), bf16);
        }
        for (int u = 0; u < H; u++) {
            grad[OFF_B2 + u] += dz2[u];
            for (int k = 0; k < H; k++) grad[OFF_W2 + u * H + k] += dz2[u] * c.h1[k];
        }
        for (int k = 0; k < H; k++) {
            double s = 0;
            for (int u = 0; u < H; u++) s += rnd(p[OFF_W2 + u * H + k], bf16) * dz2[u];
            dz1[k] = rnd(s * gelu_d(c.z1[k]), bf16);
        }
        for (int u = 0; u < H; u++) {
            grad[OFF_B1 + u] += dz1[u];
            for (int k = 0; k < 8; k++) grad[OFF_W1 + u * 8 + k] += dz1[u] * c.x[k];
        }
    }
    if (grad)
        for (int i = 0; i < NSO_MLP_NPARAM; i++) grad[i] /= (double)n;
    return loss / (double)n;
}

/* parameters used by the solver oracle (nso_rhs / nso_step) when weights == NSO_W_LEARNED */
static double g_params[NSO_MLP_NPARAM];
static int    g_bf16 = 1;

void nso_set_mlp(const double* params, int bf16) {
    memcpy(g_params, params, sizeof g_params);
    g_bf16 = bf16;
}

/* learned reconstruction of a left- (side 0) or right-biased (side 1) stencil; the right-biased
 * stencil is fed to the network mirrored (WENO5_right(a..e) is WENO5_left(e..a) with the weight
 * order reversed) and the classic Shu reconstruction is used on smooth and flagged faces */
static double learned_ps(const double* params, int bf16, const double f[5], int side,
                         double scale, int* regime) {
    double v[5], x[8];
    for (int k = 0; k < 5; k++) v[k] = side ? f[4 - k] : f[k];
    nso_mlp_features_s(v, scale, x);
    int rg = nso_mlp_regime(x);
    if (regime) *regime = rg;
    double out;
    if (rg != NSO_REGIME_MLP) {
        if (side) nso_weno5_right(f, NSO_W_SHU, &out);
        else nso_weno5_left(f, NSO_W_SHU, &out);
        return out;
    }
    double w[3], classic;
    nso_mlp_forward(params, x, bf16, w);
    if (side) nso_weno5_right(f, NSO_W_SHU, &classic);
    else nso_weno5_left(f, NSO_W_SHU, &classic);
    /* candidate polynomials of the (mirrored) left-biased stencil in the caller's units */
    double c0  = (2. * v[0] - 7. * v[1] + 11. * v[2]) / 6.;
    double c1  = (-v[1] + 5. * v[2] + 2. * v[3]) / 6.;
    double c2  = (2. * v[2] + 5. * v[3] - v[4]) / 6.;
    double net = w[0] * c0 + w[1] * c1 + w[2] * c2;
    return classic + nso_mlp_blend(nso_mlp_ratio(x)) * (net - classic);
}

double nso_weno5_learned_p(const double* params, int bf16, const double f[5], int side,
                           int* regime) {
    return learned_ps(params, bf16, f, side, 0.0, regime);
}

void nso_weno5_learned(const double* params, int bf16, const double f[5], int side, double out[1],
                       int* regime) {
    out[0] = nso_weno5_learned_p(params, bf16, f, side, regime);
}

double nso_weno5_learned_g(const double f[5], int side) {
    return learned_ps(g_params, g_bf16, f, side, 0.0, 0);
}

double nso_weno5_learned_gs(const double f[5], int side, double scale) {
    return learned_ps(g_params, g_bf16, f, side, scale, 0);
}
