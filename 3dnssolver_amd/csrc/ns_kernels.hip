// ns_kernels.hip -- HIP kernels (gfx950) of the RHS + SSP-RK3 path.
//
// This translation unit is compiled twice:
//   NS_FAST=0 with -ffp-contract=off   -> namespace nsgpu::exact
//       every expression is evaluated in the reference's order with one rounding per operation
//       (IEEE division and square root), so results are bitwise those of a strict build of the
//       reference.  Multiplications by the structural zeros / ones of the axis-aligned
//       eigenvector bases are skipped: x + (+-0) == x and 1*x == x exactly, so only the sign of
//       an exact zero can differ.
//   NS_FAST=1 with -ffp-contract=on    -> namespace nsgpu::fast
//       fused multiply-add (within each source expression, so every instantiation of a body
//       fuses alike), reciprocals hoisted out of the stencils, no square root in the flux split,
//       one division per WENO reconstruction.
//   NS_FAST=1 NS_FP32=1                 -> namespace nsgpu::fp32
//       the FAST arithmetic with the reconstruction (flux split, eigen system, projections,
//       WENO) and the viscous fluxes in single precision; the state, the stage accumulation,
//       the equation of state and the reductions stay FP64.
//
// Citations are paths relative to the reference's src/ directory.
#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "ns_common.h"
#include "ns_mlp.h"

#ifndef NS_FAST
#error "compile with -DNS_FAST=0 or -DNS_FAST=1"
#endif
#ifndef NS_ROW_FENCE
#define NS_ROW_FENCE 1
#endif
// arithmetic type of the reconstruction and the viscous fluxes: double for the EXACT and FAST
// builds; float for the FP32 build (state, stage accumulation and reductions stay double)
#ifndef NS_FP32
#define NS_FP32 0
#endif
#ifndef NS_STATE_COPY
#define NS_STATE_COPY 1  // the single-precision stage kernel writes the float copies of its output
#endif
#if NS_FP32
#define NS_MODE fp32
#define NS_REAL float
#elif NS_FAST
#define NS_MODE fast
#define NS_REAL double
#else
#define NS_MODE exact
#define NS_REAL double
#endif

namespace nsgpu {
namespace NS_MODE {

typedef NS_REAL real;
#define RL(x) ((real)(x))
// face fluxes are stored in the precision they are formed in: FP64 planes in the EXACT and FAST
// builds, FP32 planes in the FP32 build (half the bytes the sweeps write and the residual reads;
// the values are the ones the FP64 planes held, float to double is exact)
typedef real flux_t;

// ------------------------------------------------------------------------------------------
// constants (convection_constants.hpp:4-12, cd_constants.hpp:9-27)
// ------------------------------------------------------------------------------------------
#define EPS_SHU RL(1E-6)
#define EPS_BORGES RL(1E-14)
#define ONEPERSIX RL(1. / 6.)
#define THIRTEENPERTWELVE RL(13. / 12.)
#define ONEPERFOUR RL(1. / 4.)
#define LF_ALPHA RL(3.0)  // ConvectionFlux::max_eigenvalue stores 3.0 (convection_flux.cpp:241)

__constant__ double c_fd1[3][7] = {
    {-1. / 2, 0., 1. / 2.},
    {1. / 12., -2. / 3., 0., 2. / 3., -1. / 12.},
    {-1. / 60., 3. / 20., -3. / 4., 0., 3. / 4., -3. / 20., 1. / 60.}};
__constant__ double c_fd2[3][7] = {
    {1., -2., 1.},
    {-1. / 12., 4. / 3., -5. / 2., 4. / 3., -1. / 12.},
    {1. / 90., -3. / 20., 3. / 2., -49. / 18., 3. / 2., -3. / 20., 1. / 90.}};

// ------------------------------------------------------------------------------------------
// FP64 division.  IEEE division on gfx950 is div_scale x2, rcp, 2 Newton steps, q = n r,
// remainder, div_fmas, div_fixup.  div_scale leaves its operands unchanged and div_fmas is a plain
// fma unless an exponent is near the ends of the range (|exponents| < ~900 and an exponent
// difference < ~700 never scale), and div_fixup only rewrites NaN / Inf / zero / overflowing
// quotients.  On the operands of this path (physical states, grid spacings, smoothness indicators)
// the sequence below is therefore the same arithmetic without the 3 range-fixing instructions:
// the correctly rounded quotient.  nsgpu_selftest checks it against the IEEE operator on random
// operands over exponents [-300, 300]; only the sign of an exact-zero quotient can differ.
// ------------------------------------------------------------------------------------------
#ifndef NS_DIV_IEEE
#define NS_DIV_IEEE 0
#endif
// refined reciprocal of a divisor: depends on d only, so quotients by one divisor share it
struct Rcp {
    double d, r;
};
__device__ __forceinline__ Rcp rcp(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, r, 1.0);
    r        = __builtin_fma(r, e, r);
    e        = __builtin_fma(-d, r, 1.0);
    r        = __builtin_fma(r, e, r);
    return {d, r};
}
// n / d from the refined reciprocal: q = n r, remainder n - d q (exact), corrected quotient
__device__ __forceinline__ double dv(double n, const Rcp& R) {
#if NS_DIV_IEEE
    return n / R.d;
#endif
    double q   = n * R.r;
    double rem = __builtin_fma(-R.d, q, n);
    return __builtin_fma(rem, R.r, q);
}
__device__ __forceinline__ double dv(double n, double d) { return dv(n, rcp(d)); }
// single precision (FP32 build): hardware reciprocal times the numerator, ~1 ulp
struct RcpF {
    float d, r;
};
__device__ __forceinline__ RcpF  rcp(float d) { return {d, __builtin_amdgcn_rcpf(d)}; }
__device__ __forceinline__ float dv(float n, const RcpF& R) { return n * R.r; }
__device__ __forceinline__ float dv(float n, float d) { return n * __builtin_amdgcn_rcpf(d); }

// an empty instruction that reads and may rewrite x: x (a load or a computed value) is complete
// here and cannot be moved below this point (the compiler otherwise sinks loads and arithmetic
// towards their first use, which can keep many intermediate values live)
template <class T> __device__ __forceinline__ void pin(T& x) { asm volatile("" : "+v"(x)); }

// ------------------------------------------------------------------------------------------
// WENO-5 (weno-5.cpp:222-322)
// ------------------------------------------------------------------------------------------
// bo (optional): the smoothness indicators b0, b1, b2 of the stencil
template <bool BORGES>
__device__ __forceinline__ real weno_combine(real a, real b, real c, real d, real e,
                                               real d0, real d1, real d2, real f0,
                                               real f1, real f2, real* bo = nullptr) {
    real t1, t2;
    t1        = a - RL(2.) * b + c;
    t2        = a - RL(4.) * b + RL(3.) * c;
    real b0 = THIRTEENPERTWELVE * t1 * t1 + ONEPERFOUR * t2 * t2;
    t1        = b - RL(2.) * c + d;
    t2        = b - d;
    real b1 = THIRTEENPERTWELVE * t1 * t1 + ONEPERFOUR * t2 * t2;
    t1        = c - RL(2.) * d + e;
    t2        = RL(3.) * c - RL(4.) * d + e;
    real b2 = THIRTEENPERTWELVE * t1 * t1 + ONEPERFOUR * t2 * t2;
    if (bo) bo[0] = b0, bo[1] = b1, bo[2] = b2;
#if !NS_FAST
    real a0, a1, a2;
    if constexpr (BORGES) {
        real tau5 = fabs(b0 - b2);
        a0          = d0 * (RL(1.0) + dv(tau5, b0 + EPS_BORGES));
        a1          = d1 * (RL(1.0) + dv(tau5, b1 + EPS_BORGES));
        a2          = d2 * (RL(1.0) + dv(tau5, b2 + EPS_BORGES));
    } else {
        a0 = dv(d0, (EPS_SHU + b0) * (EPS_SHU + b0));
        a1 = dv(d1, (EPS_SHU + b1) * (EPS_SHU + b1));
        a2 = dv(d2, (EPS_SHU + b2) * (EPS_SHU + b2));
    }
    real al = a0 + a1 + a2;
    auto   ra = rcp(al);
    return dv(a0, ra) * f0 + dv(a1, ra) * f1 + dv(a2, ra) * f2;
#else
    // one division: with s_k = (eps + b_k)^2 the Shu weights are
    // w_k = d_k prod_{j!=k} s_j / sum_m d_m prod_{j!=m} s_j; Borges a_k = d_k (1 + tau/e_k)
    // = d_k (e_k + tau) / e_k with e_k = b_k + eps, same clearing of denominators.
    real p0, p1, p2;
    if constexpr (BORGES) {
        real tau5 = fabs(b0 - b2);
        real e0 = b0 + EPS_BORGES, e1 = b1 + EPS_BORGES, e2 = b2 + EPS_BORGES;
        p0 = d0 * (e0 + tau5) * (e1 * e2);
        p1 = d1 * (e1 + tau5) * (e0 * e2);
        p2 = d2 * (e2 + tau5) * (e0 * e1);
    } else {
        real s0 = (EPS_SHU + b0) * (EPS_SHU + b0);
        real s1 = (EPS_SHU + b1) * (EPS_SHU + b1);
        real s2 = (EPS_SHU + b2) * (EPS_SHU + b2);
        p0        = d0 * (s1 * s2);
        p1        = d1 * (s0 * s2);
        p2        = d2 * (s0 * s1);
    }
    return dv(p0 * f0 + p1 * f1 + p2 * f2, p0 + p1 + p2);
#endif
}

// Weno5::WENO5_left (weno-5.cpp:305-322), ideal weights d_left = {0.1, 0.6, 0.3}
template <bool BORGES>
__device__ __forceinline__ real weno_left_c(real a, real b, real c, real d, real e,
                                            real* bo = nullptr) {
    real f0 = (RL(2.) * a - RL(7.) * b + RL(11.) * c) * ONEPERSIX;
    real f1 = (-b + RL(5.) * c + RL(2.) * d) * ONEPERSIX;
    real f2 = (RL(2.) * c + RL(5.) * d - e) * ONEPERSIX;
    return weno_combine<BORGES>(a, b, c, d, e, RL(0.1), RL(0.6), RL(0.3), f0, f1, f2, bo);
}

// Weno5::WENO5_right (weno-5.cpp:285-303), ideal weights d_right = {0.3, 0.6, 0.1}
template <bool BORGES>
__device__ __forceinline__ real weno_right_c(real a, real b, real c, real d, real e,
                                             real* bo = nullptr) {
    real f0 = (-a + RL(5.) * b + RL(2.) * c) * ONEPERSIX;
    real f1 = (RL(2.) * b + RL(5.) * c - d) * ONEPERSIX;
    real f2 = (RL(11.) * c - RL(7.) * d + RL(2.) * e) * ONEPERSIX;
    return weno_combine<BORGES>(a, b, c, d, e, RL(0.3), RL(0.6), RL(0.1), f0, f1, f2, bo);
}


// ------------------------------------------------------------------------------------------
// learned weights (ns_mlp.h; no reference counterpart).  The network sees the stencil of the
// left-biased reconstruction; WENO5_right(a..e) is fed mirrored, (e..a).  Smooth and flagged
// stencils keep the classic Shu reconstruction, computed by the same code as the classic path;
// in between, the network's reconstruction enters with the share mlp_blend(indicator ratio).
// Every lane of the wavefront must call this (the network uses matrix instructions); `active`
// marks the lanes whose face is real.
// ------------------------------------------------------------------------------------------
// q: this wavefront's queue of network samples in LDS (LearnQueue), or null: the face kernels
// batch the network samples of all reconstructions of a face (face_flux_learned_cw)
// e: emit mode (single-block component-wise tiles): the network stencils go to the sample
// buffers of the launch (LearnSink) and a separate kernel evaluates them; fc is this thread's
// face (flux index of variable 0), fP the stride between variables, region its workgroup's
// region of the buffers
struct LearnEmit {
    real*     v[2];     // per side: planes [5][cap] (stencil in left-biased order)
    unsigned* dest[2];  // per side: flux index of each sample
    unsigned* cnt;      // counters [side][LS_REGIONS] at stride LS_STRIDE
    unsigned* zero;     // the set's other counters, zeroed by the launch's first workgroup
    long      cap;      // samples per side
    long      capr;     // samples per region of this launch
};

// the first workgroup of an emit-mode launch zeroes the counters of the set's next launch
__device__ __forceinline__ void learn_zero_next(const LearnEmit& e) {
    if (blockIdx.x != 0 || blockIdx.y != 0 || blockIdx.z != 0) return;
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    if (t < 2 * LS_REGIONS) e.zero[(long)t * LS_STRIDE] = 0u;
}
struct Learn {
    const MlpPacked* P;
    bool             active;
    real*            q = nullptr;
    LearnEmit        e = {};
    unsigned         fc = 0, fP = 0;
    int              region = 0;
};

// the classic Shu reconstruction of the stencil v in left-biased order (side 1: mirrored);
// bo (optional): its smoothness indicators
__device__ __forceinline__ real learned_classic(const real v[5], int side, real* bo = nullptr) {
    return side ? weno_right_c<false>(v[4], v[3], v[2], v[1], v[0], bo)
                : weno_left_c<false>(v[0], v[1], v[2], v[3], v[4], bo);
}

// min, max and largest magnitude of three values.  Single precision: one v_min3_f32 /
// v_max3_f32 each (the magnitudes through the operands' abs modifiers).  For operands that are
// not NaN they return the values the comparison chains of double precision return, up to the
// sign of a zero result, which no regime decision below depends on (a zero range, magnitude or
// indicator compares the same with either sign).
__device__ __forceinline__ float min3r(float a, float b, float c) {
    float r;
    __asm__("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float max3r(float a, float b, float c) {
    float r;
    __asm__("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ float maxabs3r(float a, float b, float c) {  // max(|a|, |b|, c)
    float r;
    __asm__("v_max3_f32 %0, |%1|, |%2|, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ __forceinline__ double min3r(double a, double b, double c) {
    const double t = b < a ? b : a;
    return c < t ? c : t;
}
__device__ __forceinline__ double max3r(double a, double b, double c) {
    const double t = b > a ? b : a;
    return c > t ? c : t;
}
__device__ __forceinline__ double maxabs3r(double a, double b, double c) {
    const double x = a < 0.0 ? -a : a, y = b < 0.0 ? -b : b;
    const double t = x > y ? x : y;
    return c > t ? c : t;
}

// regime of the stencil v (0 smooth, 1 network, 2 flagged discontinuity) from the smoothness
// indicators b of its classic reconstruction: the indicators of the range-normalised stencil
// (v - min) / r are b / r^2, so their ratio (max + EPS) / (min + EPS) is (max b + EPS r^2) /
// (min b + EPS r^2), compared here without a division.  A stencil whose range r is at most
// MLP_RANGE of its largest magnitude counts as constant (smooth).  The same classification as
// mlp_features' up to rounding, for the price of a few operations on values the classic
// reconstruction has formed anyway.  scale (>= 0): a magnitude the range is also measured
// against (the face's momentum scale for a momentum component, learned_face_scale)
__device__ __forceinline__ int learned_regime(const real v[5], const real b[3],
                                              real scale = RL(0.0)) {
    const real m  = min3r(min3r(v[0], v[1], v[2]), v[3], v[4]);
    const real M  = max3r(max3r(v[0], v[1], v[2]), v[3], v[4]);
    const real r  = M - m;
    const real mg = maxabs3r(m, M, scale);
    if (!(r > (real)MLP_RANGE * mg)) return 0;
    const real e   = (real)MLP_EPS * (r * r);
    const real lo  = min3r(b[0], b[1], b[2]), hi = max3r(b[0], b[1], b[2]);
    const real num = hi + e, den = lo + e;
    return num < (real)MLP_SMOOTH * den ? 0 : (num > (real)MLP_FLAG * den ? 2 : 1);
}

// the blend of the network's reconstruction (weights w of the three ENO candidates of v) with
// the classic one
__device__ __forceinline__ real learned_finish(const real v[5], const float w[3], real classic,
                                               float ratio) {
    const real f0 = (RL(2.) * v[0] - RL(7.) * v[1] + RL(11.) * v[2]) * ONEPERSIX;
    const real f1 = (-v[1] + RL(5.) * v[2] + RL(2.) * v[3]) * ONEPERSIX;
    const real f2 = (RL(2.) * v[2] + RL(5.) * v[3] - v[4]) * ONEPERSIX;
    const real net = (real)w[0] * f0 + (real)w[1] * f1 + (real)w[2] * f2;
    return classic + (real)mlp_blend(ratio) * (net - classic);
}

// scale: as in learned_regime
template <int SIDE>
__device__ __forceinline__ real weno_learned(real a, real b, real c, real d, real e,
                                             const Learn& L, int* regime = nullptr,
                                             real scale = RL(0.0)) {
    const real v[5]    = {SIDE ? e : a, SIDE ? d : b, c, SIDE ? b : d, SIDE ? a : e};
    real       bs[3];
    const real classic = learned_classic(v, SIDE, bs);
    const int  rg      = learned_regime(v, bs, scale);
    const bool need    = L.active && rg == 1;
    if (regime) *regime = rg;
    if (__ballot(need) == 0) return classic;  // wavefront-uniform
    float x[8], ratio;
    mlp_features(v, x, &ratio);
    float xs[8], w[3];
#pragma unroll
    for (int k = 0; k < 8; k++) xs[k] = need ? x[k] : 0.0f;
    mlp_weights_lane(L.P, xs, w);
    if (!need) return classic;
    return learned_finish(v, w, classic, ratio);
}

// ---- batched network samples of one face (component-wise sweeps) ---------------------------
// A wavefront's queue holds up to 64 stencils that need the network: their five values
// (left-biased order) and their classic reconstruction, as planes [6][64] of `real` in LDS.  The reconstructions of a
// face enqueue their network stencils; when the next one would overflow the queue, and after the
// last, the wavefront evaluates the queued samples in one 64-sample network call (lane l takes
// sample l), writes each result over the sample's first value, and every lane takes back the
// results of its own samples.  A sample's arithmetic (features, classic reconstruction, matrix
// columns, blend) is that of weno_learned, so the results are bitwise the same; only the number
// of network calls changes (one per 64 samples instead of one per reconstruction site).
constexpr int LQ_REALS = 6 * 64;  // per wavefront

// LDS writes of one lane become visible to the other lanes of its wavefront
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// evaluates the n <= 64 queued samples; sample s's result replaces its value 0 (q[s])
__device__ __forceinline__ void learn_flush(const MlpPacked* P, real* q, int n) {
    const int  l   = threadIdx.x & 63;
    const bool has = l < n;
    real       v[5];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 5; k++) v[k] = has ? q[64 * k + l] : RL(0.0);
    const real classic = has ? q[64 * 5 + l] : RL(0.0);
    float      x[8], xs[8], ratio, w[3];
    mlp_features(v, x, &ratio);
#pragma unroll
    for (int k = 0; k < 8; k++) xs[k] = has ? x[k] : 0.0f;
    mlp_weights_lane(P, xs, w);
    if (has) q[l] = learned_finish(v, w, classic, ratio);
    wave_sync();
}

// momentum scale of a component-wise face: the largest split value (fL and fR) of the three
// momentum components (stored positions 2..4) at the face's two cells.  A momentum stencil also
// counts as constant when its range is at most MLP_RANGE of this scale: a momentum component
// that vanishes in exact arithmetic (the spanwise velocity of a plane shear flow) holds only
// round-off, whose stencils are not structure for the network, however they look normalised.
template <class S>
__device__ __forceinline__ real learned_face_scale(const S& st) {
    real s = RL(0.0);
#pragma unroll
    for (int q = 2; q < 5; q++) {
#pragma unroll
        for (int ii = 2; ii < 4; ii++) s = maxabs3r(st(ii, q), st(ii, 5 + q), s);
    }
    return s;
}

// stencil of reconstruction r = 2 q + side of a component-wise face: WENO5_left of fL of cells
// i-2 .. i+2, or WENO5_right of fR of cells i-1 .. i+3 mirrored (left-biased order)
template <class S>
__device__ __forceinline__ void learned_stencil(const S& st, int r, real v[5]) {
    const int qv = r >> 1;
    if ((r & 1) == 0) {
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = st(k, qv);
    } else {
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = st(5 - k, 5 + qv);
    }
}

// component-wise face flux with the learned weights, emit mode.  Every reconstruction is first
// formed with the classic weights.  The stencils in the network regime are appended to the
// launch's sample buffers (one counter reservation per side and wavefront), and the face flux is
// stored with -0 in place of each such reconstruction: -0 + x == x for every x, so once the
// network kernel has added the learned left reconstructions and then the learned right ones
// (flux = flux + value, in that order), every flux holds exactly left_q + right_q as
// face_flux_learned_cw forms it.  A sample's network arithmetic does not depend on the other
// samples of its matrix call, so the results are bitwise those of the in-kernel queue.
template <class S>
__device__ __forceinline__ void face_flux_learned_emit(const S& st, real out[5], const Learn& L) {
    const LearnEmit&         E     = L.e;
    const int                lane  = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long       m[10];  // wavefront masks of the network samples per reconstruction
    int                      n0 = 0, n1 = 0;
    const real               fs = learned_face_scale(st);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        // a compiler fence per variable: its stencils are read from LDS here and not kept in
        // registers from an earlier read or for a later one (few live registers)
        __asm__ __volatile__("" ::: "memory");
        real half[2];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const int r = 2 * q + side;
            real      v[5], bs[3];
            learned_stencil(st, r, v);
            const real classic = learned_classic(v, side, bs);
            const bool need = L.active && learned_regime(v, bs, q >= 2 ? fs : RL(0.0)) == 1;
            m[r]            = __ballot(need);
            half[side]      = need ? RL(-0.0) : classic;
            if (side) n1 += __popcll(m[r]);
            else n0 += __popcll(m[r]);
        }
        out[q] = half[0] + half[1];
        pin(out[q]);  // formed here, not after the sample writes below
    }
    if (n0 + n1 == 0) return;  // wavefront-uniform
    int t0 = 0, t1 = 0;
    if (lane == 0) {
        if (n0) t0 = (int)atomicAdd(E.cnt + (long)L.region * LS_STRIDE, (unsigned)n0);
        if (n1)
            t1 = (int)atomicAdd(E.cnt + (long)(LS_REGIONS + L.region) * LS_STRIDE, (unsigned)n1);
    }
    const unsigned s0 = (unsigned)__shfl(t0, 0), s1 = (unsigned)__shfl(t1, 0);
    // the buffers hold every reconstruction of the region's faces, so a reservation never ends
    // past its region while the counters start each launch at zero; a wavefront whose
    // reservation would end past it writes none of that side's samples
    const bool ok0 = (long)s0 + n0 <= E.capr, ok1 = (long)s1 + n1 <= E.capr;
    long b0 = (long)L.region * E.capr + s0;
    long b1 = (long)L.region * E.capr + s1;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        __asm__ __volatile__("" ::: "memory");  // (as above: the stencil is read again)
        const int side = r & 1;
        if (m[r] == 0 || !(side ? ok1 : ok0)) continue;  // wavefront-uniform
        if ((m[r] >> lane) & 1ull) {
            real v[5];
            learned_stencil(st, r, v);
            const long s = (side ? b1 : b0) + __popcll(m[r] & below);
            real*      V = E.v[side];
#pragma unroll
            for (int k = 0; k < 5; k++) V[k * E.cap + s] = v[k];
            E.dest[side][s] = L.fc + (unsigned)(r >> 1) * L.fP;
        }
        if (side) b1 += __popcll(m[r]);
        else b0 += __popcll(m[r]);
    }
}

// the network pass of emit mode: every wavefront evaluates 64 samples of one side per matrix
// call and adds each result to its flux.  A launch handles one side; each flux receives at most
// one sample per side, so the launch writes disjoint fluxes.  Chunks of 64 samples are numbered
// region-fastest (chunk c: region c % LS_REGIONS, local chunk c / LS_REGIONS), so the occupied
// chunks come first; the wavefronts stride over them.
__device__ __forceinline__ void learn_net_body(const MlpPacked* __restrict__ P,
                                               const LearnEmit& E, int side,
                                               flux_t* __restrict__ f) {
    const int      l    = threadIdx.x & 63;
    const long     gw   = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long     nw   = ((long)gridDim.x * blockDim.x) >> 6;
    const unsigned* cnt = E.cnt + (long)side * LS_REGIONS * LS_STRIDE;
    const unsigned cr   = (unsigned)E.capr;  // (a region's count never exceeds it)
    unsigned       mx   = (l < LS_REGIONS) ? min(cnt[(long)l * LS_STRIDE], cr) : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned t = (unsigned)__shfl_xor((int)mx, o, 64);
        mx               = t > mx ? t : mx;
    }
    const long      limit = (long)((mx + 63u) / 64u) * LS_REGIONS;
    const real*     V     = E.v[side];
    const unsigned* D     = E.dest[side];
    for (long ch = gw; ch < limit; ch += nw) {
        const int      r  = (int)(ch % LS_REGIONS);
        const long     lc = ch / LS_REGIONS;
        const unsigned n  = min(cnt[(long)r * LS_STRIDE], cr);
        if (lc * 64 >= (long)n) continue;  // wavefront-uniform
        const bool has = lc * 64 + l < (long)n;
        const long s   = (long)r * E.capr + lc * 64 + l;
        real       v[5];
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = has ? V[k * E.cap + s] : RL(0.0);
        const real classic = learned_classic(v, side);
        float      x[8], xs[8], ratio, w[3];
        mlp_features(v, x, &ratio);
#pragma unroll
        for (int k = 0; k < 8; k++) xs[k] = has ? x[k] : 0.0f;
        mlp_weights_lane(P, xs, w);
        if (has) {
            const unsigned d = D[s];
            f[d]             = f[d] + (flux_t)learned_finish(v, w, classic, ratio);
        }
    }
}

// component-wise face flux with the learned weights: out[q] = left_q + right_q as in face_flux,
// each reconstruction as weno_learned computes it, the network samples batched per face
template <class S>
__device__ __forceinline__ void face_flux_learned_cw(const S& st, real out[5], const Learn& L) {
    real val[10];
    int  slot[10];
    int  base = 0;  // queued samples (wavefront-uniform)
    const unsigned long long below = (1ull << (threadIdx.x & 63)) - 1ull;
    const real               fs    = learned_face_scale(st);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const int side = r & 1;
        real      v[5];
        learned_stencil(st, r, v);
        real       bs[3];
        const real classic = learned_classic(v, side, bs);
        const bool need    = L.active && learned_regime(v, bs, r >= 4 ? fs : RL(0.0)) == 1;
        const unsigned long long m = __ballot(need);
        const int cnt = __popcll(m);
        if (base + cnt > 64) {  // wavefront-uniform
            learn_flush(L.P, L.q, base);
#pragma unroll
            for (int t = 0; t < r; t++)
                if (slot[t] >= 0) val[t] = L.q[slot[t]], slot[t] = -1;
            wave_sync();  // results read before the queue is refilled
            base = 0;
        }
        val[r]  = classic;
        slot[r] = -1;
        if (need) {
            const int s = base + __popcll(m & below);
#pragma unroll
            for (int k = 0; k < 5; k++) L.q[64 * k + s] = v[k];
            L.q[64 * 5 + s] = classic;
            slot[r]         = s;
        }
        base += cnt;
    }
    if (base > 0) {
        learn_flush(L.P, L.q, base);
#pragma unroll
        for (int t = 0; t < 10; t++)
            if (slot[t] >= 0) val[t] = L.q[slot[t]];
        wave_sync();
    }
#pragma unroll
    for (int q = 0; q < 5; q++) out[q] = val[2 * q] + val[2 * q + 1];
}

// weights WT: 0 Shu, 1 Borges, 2 learned
// scale: the learned weights' constant-stencil scale (learned_regime)
template <int WT>
__device__ __forceinline__ real weno_left(real a, real b, real c, real d, real e, const Learn& L,
                                          real scale = RL(0.0)) {
    if constexpr (WT == 2) return weno_learned<0>(a, b, c, d, e, L, nullptr, scale);
    else return weno_left_c<WT == 1>(a, b, c, d, e);
}
template <int WT>
__device__ __forceinline__ real weno_right(real a, real b, real c, real d, real e, const Learn& L,
                                           real scale = RL(0.0)) {
    if constexpr (WT == 2) return weno_learned<1>(a, b, c, d, e, L, nullptr, scale);
    else return weno_right_c<WT == 1>(a, b, c, d, e);
}

// ------------------------------------------------------------------------------------------
// Euler flux and Lax-Friedrichs split (convection_flux.cpp:23-61, weno-5.cpp:94-110):
// s[0..4] = fL = (F + alpha U)/2, s[5..9] = fR = (F - alpha U)/2 of one cell, direction DIR
// ------------------------------------------------------------------------------------------
template <int DIR, class TS>
__device__ __forceinline__ void split_cell(const TS* __restrict__ U, const TS* __restrict__ W,
                                           long c, long P, real gamma, real s[10]) {
    real rho = W[c], p = W[c + P], u = W[c + 2 * P], v = W[c + 3 * P], w = W[c + 4 * P];
    real qn  = (DIR == 0) ? u : (DIR == 1) ? v : w;  // u*nx + v*ny + w*nz
#if !NS_FAST
    real cs = sqrt(dv(gamma * p, rho));
    real H  = dv(cs * cs, gamma - RL(1.0)) + RL(0.5) * (u * u + v * v + w * w);
#else
    real H = dv(gamma, gamma - RL(1.0)) * dv(p, rho) + RL(0.5) * (u * u + v * v + w * w);
#endif
    real F[5];
    real m = rho * qn;
    F[0]     = m;
    F[1]     = m * H;
    F[2]     = (DIR == 0) ? m * u + p : m * u;
    F[3]     = (DIR == 1) ? m * v + p : m * v;
    F[4]     = (DIR == 2) ? m * w + p : m * w;
#pragma unroll
    for (int q = 0; q < 5; q++) {
        real uq = U[c + q * P];
        s[q]      = (F[q] + LF_ALPHA * uq) * RL(0.5);
        s[5 + q]  = (F[q] - LF_ALPHA * uq) * RL(0.5);
    }
}

// the per-face Lax-Friedrichs constant (alpha_mode LOCAL_MAX): the split depends on the face, so
// a cell stores s[0..4] = F, s[5..9] = U and s[10] = its eigenvalue max(|q_n - c|, |q_n + c|)
// (the value convection_flux.cpp:235-238 computes and leaves unused)
template <int DIR, class TS>
__device__ __forceinline__ void split_cell_la(const TS* __restrict__ U,
                                              const TS* __restrict__ W, long c, long P,
                                              real gamma, real s[11]) {
    real rho = W[c], p = W[c + P], u = W[c + 2 * P], v = W[c + 3 * P], w = W[c + 4 * P];
    real qn  = (DIR == 0) ? u : (DIR == 1) ? v : w;
    real cs  = sqrt(dv(gamma * p, rho));
#if !NS_FAST
    real H = dv(cs * cs, gamma - RL(1.0)) + RL(0.5) * (u * u + v * v + w * w);
#else
    real H = dv(gamma, gamma - RL(1.0)) * dv(p, rho) + RL(0.5) * (u * u + v * v + w * w);
#endif
    real m = rho * qn;
    s[0]   = m;
    s[1]   = m * H;
    s[2]   = (DIR == 0) ? m * u + p : m * u;
    s[3]   = (DIR == 1) ? m * v + p : m * v;
    s[4]   = (DIR == 2) ? m * w + p : m * w;
#pragma unroll
    for (int q = 0; q < 5; q++) s[5 + q] = U[c + q * P];
    real e1 = fabs(qn - cs), e2 = fabs(qn + cs);
    s[10]   = (e1 < e2) ? e2 : e1;  // std::max
}

// split values of one face from the stored F, U of its cells and the face's constant a:
// (F + a U) / 2 and (F - a U) / 2 as in weno-5.cpp:94-110
template <class S>
struct AlphaSplit {
    const S& raw;
    real     a;
    __device__ real operator()(int ii, int v) const {
        return (v < 5) ? (raw(ii, v) + a * raw(ii, v + 5)) * RL(0.5)
                       : (raw(ii, v - 5) - a * raw(ii, v)) * RL(0.5);
    }
};

// Weno5::get_max_alpha (weno-5.cpp:324-330) over the cells i-2 .. i+2 of the face
template <class S>
__device__ __forceinline__ real face_alpha(const S& raw) {
    real m = raw(0, 10);
#pragma unroll
    for (int ii = 1; ii < 5; ii++) {
        real d = raw(ii, 10);
        m      = (m < d) ? d : m;
    }
    return m;
}

// ------------------------------------------------------------------------------------------
// Characteristic projection (convection_flux.cpp:63-175, weno-5cw.cpp:4-28)
// ------------------------------------------------------------------------------------------

// axis bases of ConvectionFlux::orthogonal_unit_vectors: n = e_DIR,
// x: l = (0,0,1), m = (0,-1,0); y: l = (0,0,1), m = (1,0,0); z: l = (0,-1,0), m = (1,0,0)
template <int DIR, int A> struct Basis;
template <int A> struct Basis<0, A> {
    static constexpr int n = (A == 0), l = (A == 2), m = (A == 1) ? -1 : 0;
};
template <int A> struct Basis<1, A> {
    static constexpr int n = (A == 1), l = (A == 2), m = (A == 0);
};
template <int A> struct Basis<2, A> {
    static constexpr int n = (A == 2), l = (A == 1) ? -1 : 0, m = (A == 0);
};

// acc + K*x for a structural constant K in {-1, 0, 1}
template <int K> __device__ __forceinline__ real cadd(real acc, real x) {
    if constexpr (K == 1) return acc + x;
    else if constexpr (K == -1) return acc - x;
    else return acc;
}

// the eigenvector entries that depend on the face state (ConvectionFlux::eigenvectors,
// convection_flux.cpp:117-175, c^2 = gamma / (p rho) as written there)
struct Eig {
    real vel[3];
    real c, H, qn, ql, qm, q2h;
    real L0[5], L2[5], L3[5];  // rows 0, 2, 3 of L (all entries depend on the state)
};

template <int DIR, bool TRUEC>
__device__ __forceinline__ void eigen(real rho, real p, real u, real v, real w,
                                      real gamma, Eig& E) {
    E.vel[0]  = u;
    E.vel[1]  = v;
    E.vel[2]  = w;
    real qn = E.vel[DIR];
    // ql = u*lx + v*ly + w*lz, qm = u*mx + v*my + w*mz with one nonzero entry each
    real ql = (DIR == 2) ? -v : w;
    real qm = (DIR == 0) ? -v : u;
    real q2 = u * u + v * v + w * w;
    real K  = gamma - RL(1.0);
#if !NS_FAST
    real c2   = TRUEC ? dv(gamma * p, rho) : dv(gamma, p * rho);
    real c    = sqrt(c2);
    real H    = dv(c2, gamma - RL(1.0)) + RL(0.5) * q2;
    real K_pc = dv(K, RL(2.0) * c2);
    auto   r2c  = rcp(RL(2.0) * c);
    real qn2c = dv(qn, r2c);
    real i2c  = dv(RL(1.0), r2c);  // n_a / (2c) for the normal component
    auto   rc2  = rcp(c2);
#pragma unroll
    for (int a = 0; a < 3; a++) E.L2[1 + a] = dv(K * E.vel[a], rc2);
    E.L2[4] = dv(-K, rc2);
#else
    real c2   = TRUEC ? dv(gamma * p, rho) : dv(gamma, p * rho);
    real ic2  = TRUEC ? dv(rho, gamma * p) : dv(p * rho, gamma);  // 1 / c^2
    real c    = sqrt(c2);
    real H    = dv(c2, K) + RL(0.5) * q2;
    real K_pc = RL(0.5) * K * ic2;
    real i2c  = RL(0.5) * c * ic2;    // 1 / (2c)
    real qn2c = qn * i2c;
#pragma unroll
    for (int a = 0; a < 3; a++) E.L2[1 + a] = K * E.vel[a] * ic2;
    E.L2[4] = -K * ic2;
#endif
    real t = RL(0.5) * K_pc * q2;
    E.c = c, E.H = H, E.qn = qn, E.ql = ql, E.qm = qm, E.q2h = RL(0.5) * q2;
    E.L0[0] = t + qn2c;
    E.L3[0] = t - qn2c;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        real kv = K_pc * E.vel[a];
        if (a == DIR) {
            E.L0[1 + a] = -(kv + i2c);
            E.L3[1 + a] = -(kv - i2c);
        } else {
            E.L0[1 + a] = -(kv);
            E.L3[1 + a] = -(kv);
        }
    }
    E.L0[4] = K_pc;
    E.L3[4] = K_pc;
    E.L2[0] = RL(1.0) - K_pc * q2;
}

// row Q of L times the split stencil values s[p] (p in conservative order), summed in p order
// from 0.0 (matrix::square_matrix_matrix_product, matrix.cpp:12-20)
template <int DIR, int Q>
__device__ __forceinline__ real Lrow(const Eig& E, const real s[5]) {
    if constexpr (Q == 0) {
        return E.L0[0] * s[0] + E.L0[1] * s[1] + E.L0[2] * s[2] + E.L0[3] * s[3] + E.L0[4] * s[4];
    } else if constexpr (Q == 2) {
        return E.L2[0] * s[0] + E.L2[1] * s[1] + E.L2[2] * s[2] + E.L2[3] * s[3] + E.L2[4] * s[4];
    } else if constexpr (Q == 3) {
        return E.L3[0] * s[0] + E.L3[1] * s[1] + E.L3[2] * s[2] + E.L3[3] * s[3] + E.L3[4] * s[4];
    } else if constexpr (Q == 1) {  // {-qm, mx, my, mz, 0}
        real acc = (-E.qm) * s[0];
        acc        = cadd<Basis<DIR, 0>::m>(acc, s[1]);
        acc        = cadd<Basis<DIR, 1>::m>(acc, s[2]);
        acc        = cadd<Basis<DIR, 2>::m>(acc, s[3]);
        return acc;
    } else {  // Q == 4: {-ql, lx, ly, lz, 0}
        real acc = (-E.ql) * s[0];
        acc        = cadd<Basis<DIR, 0>::l>(acc, s[1]);
        acc        = cadd<Basis<DIR, 1>::l>(acc, s[2]);
        acc        = cadd<Basis<DIR, 2>::l>(acc, s[3]);
        return acc;
    }
}

// R times the characteristic fluxes ft (matrix::square_matrix_vector_product, matrix.cpp:4-10)
template <int DIR, int A>
__device__ __forceinline__ real Rvel(const Eig& E, const real ft[5]) {
    // row 1+A: {vel_A - c n_A, m_A, vel_A, vel_A + c n_A, l_A}
    constexpr int n   = Basis<DIR, A>::n;
    real        va  = E.vel[A];
    real        e0  = n ? va - E.c : va;
    real        e3  = n ? va + E.c : va;
    real        acc = e0 * ft[0];
    acc               = cadd<Basis<DIR, A>::m>(acc, ft[1]);
    acc               = acc + va * ft[2];
    acc               = acc + e3 * ft[3];
    acc               = cadd<Basis<DIR, A>::l>(acc, ft[4]);
    return acc;
}

template <int DIR>
__device__ __forceinline__ void Rmul(const Eig& E, const real ft[5], real out[5]) {
    out[0]    = ft[0] + ft[2] + ft[3];  // {1, 0, 1, 1, 0}
    out[1]    = Rvel<DIR, 0>(E, ft);
    out[2]    = Rvel<DIR, 1>(E, ft);
    out[3]    = Rvel<DIR, 2>(E, ft);
    real qc = E.qn * E.c;             // {H - qn c, qm, q2/2, H + qn c, ql}
    out[4] = (E.H - qc) * ft[0] + E.qm * ft[1] + E.q2h * ft[2] + (E.H + qc) * ft[3] + E.ql * ft[4];
}

// ------------------------------------------------------------------------------------------
// one face flux from the split values of its six stencil cells.  S(ii, v) returns split value v
// (0..4 fL, 5..9 fR) of cell i-2+ii, ii = 0..5; Wl, Wr are the primitives of cells i and i+1
// (Weno5::{x,y,z}_sweep weno-5.cpp:75-220, Weno5cw::{x,y,z}_sweep weno-5cw.cpp:30-262)
// ------------------------------------------------------------------------------------------
// stored position [rho, rho E, rho u, rho v, rho w] of conserved component p of
// [rho, rho u, rho v, rho w, rho E]
__device__ constexpr int stored_of(int p) { return p == 0 ? 0 : (p == 4 ? 1 : p + 1); }

// CORR: L acts on the conserved order (physics CONVECTION); otherwise on the stored order as the
// reference applies it
template <int DIR, int WT, int Q, bool CORR, class S>
__device__ __forceinline__ real char_component(const Eig& E, const S& st, const Learn& L) {
    // stencil ii uses fL at cell i-2+ii and fR at cell i-1+ii (weno-5cw.cpp:85-97)
#if NS_ROW_FENCE
    // re-read the stencil per row (per pair of rows with NS_ROW_FENCE 2): fewer live registers
    if constexpr (NS_ROW_FENCE == 1 || Q % 2 == 0) __asm__ __volatile__("" ::: "memory");
#endif
    real a[5], b[5];
#pragma unroll
    for (int ii = 0; ii < 5; ii++) {
        real sl[5], sr[5];
#pragma unroll
        for (int p = 0; p < 5; p++) {
            const int sp = CORR ? stored_of(p) : p;
            sl[p]        = st(ii, sp);
            sr[p]        = st(ii + 1, 5 + sp);
        }
        a[ii] = Lrow<DIR, Q>(E, sl);
        b[ii] = Lrow<DIR, Q>(E, sr);
    }
    return weno_left<WT>(a[0], a[1], a[2], a[3], a[4], L) +
           weno_right<WT>(b[0], b[1], b[2], b[3], b[4], L);
}

// Roe average of the primitives of the two cells of a face (physics CONVECTION), returned as
// primitives [rho, p, u, v, w] with p = rho c^2 / gamma
__device__ __forceinline__ void roe_state(const real l[5], const real r[5], real gamma,
                                          real out[5]) {
    real sl = sqrt(l[0]), sr = sqrt(r[0]);
    real den = sl + sr;
    real gg  = dv(gamma, gamma - RL(1.0));
    real hl  = gg * dv(l[1], l[0]) + RL(0.5) * (l[2] * l[2] + l[3] * l[3] + l[4] * l[4]);
    real hr  = gg * dv(r[1], r[0]) + RL(0.5) * (r[2] * r[2] + r[3] * r[3] + r[4] * r[4]);
    real u   = dv(sl * l[2] + sr * r[2], den);
    real v   = dv(sl * l[3] + sr * r[3], den);
    real w   = dv(sl * l[4] + sr * r[4], den);
    real H   = dv(sl * hl + sr * hr, den);
    real c2  = (gamma - RL(1.0)) * (H - RL(0.5) * (u * u + v * v + w * w));
    real rho = sl * sr;
    out[0] = rho, out[1] = dv(rho * c2, gamma), out[2] = u, out[3] = v, out[4] = w;
}

// OPT bit 0: per-face alpha (st holds F, U and the eigenvalue per cell); bit 1: corrected
// characteristic projection
// EMIT: the learned component-wise flux in emit mode (face_flux_learned_emit)
template <int DIR, int CONV, int WT, int OPT, bool EMIT = false, class S>
__device__ __forceinline__ void face_flux(const S& raw, const real Wl[5], const real Wr[5],
                                          real gamma, real out[5], const Learn& L) {
    if constexpr (OPT & 1) {
        AlphaSplit<S> st{raw, face_alpha(raw)};
        face_flux<DIR, CONV, WT, OPT & ~1>(st, Wl, Wr, gamma, out, L);
        return;
    } else {
    const S& st = raw;
    constexpr bool CORR = (OPT & 2) != 0;
    if constexpr (CONV == 1 && WT == 2 && EMIT) {
        face_flux_learned_emit(st, out, L);
    } else if constexpr (CONV == 1) {
        if constexpr (WT == 2) {
            if (L.q) {
                face_flux_learned_cw(st, out, L);
                return;
            }
        }
        real fs = RL(0.0);
        if constexpr (WT == 2) fs = learned_face_scale(st);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const real sc = q >= 2 ? fs : RL(0.0);
            out[q] = weno_left<WT>(st(0, q), st(1, q), st(2, q), st(3, q), st(4, q), L, sc) +
                     weno_right<WT>(st(1, 5 + q), st(2, 5 + q), st(3, 5 + q), st(4, 5 + q),
                                    st(5, 5 + q), L, sc);
        }
    } else {
        real Wa[5];
        if constexpr (CORR) {
            roe_state(Wl, Wr, gamma, Wa);
        } else {
#pragma unroll
            for (int q = 0; q < 5; q++) Wa[q] = RL(0.5) * (Wl[q] + Wr[q]);
        }
        Eig E;
        eigen<DIR, CORR>(Wa[0], Wa[1], Wa[2], Wa[3], Wa[4], gamma, E);
        real ft[5];
        ft[0] = char_component<DIR, WT, 0, CORR>(E, st, L);
        ft[1] = char_component<DIR, WT, 1, CORR>(E, st, L);
        ft[2] = char_component<DIR, WT, 2, CORR>(E, st, L);
        ft[3] = char_component<DIR, WT, 3, CORR>(E, st, L);
        ft[4] = char_component<DIR, WT, 4, CORR>(E, st, L);
        Rmul<DIR>(E, ft, out);
        if constexpr (CORR) {  // conserved order back to the stored order
            real m0 = out[1], m1 = out[2], m2 = out[3], e = out[4];
            out[1] = e, out[2] = m0, out[3] = m1, out[4] = m2;
        }
    }
    }
}

template <int NV>
struct RegSplit {  // split values held in registers: s[ii][v]
    const real (*s)[NV];
    __device__ real operator()(int ii, int v) const { return s[ii][v]; }
};

template <int LANES, int NV>
struct LdsSplit {  // split values in LDS, layout [cell][v][lane]
    const real* base;  // &tile[first cell][0][lane]
    __device__ real operator()(int ii, int v) const { return base[(ii * NV + v) * LANES]; }
};

// values stored per cell: the split (10), or F, U and the eigenvalue (11) for a per-face alpha
template <int OPT> constexpr int split_nv() { return (OPT & 1) ? 11 : 10; }

// TS: the type of the stored state the cell is read from (double, or the float copies of the
// single-precision build: every value is converted to real on its load either way)
template <int DIR, int OPT, class TS>
__device__ __forceinline__ void stage_cell(const TS* __restrict__ U,
                                           const TS* __restrict__ W, long c, long P,
                                           real gamma, real* sv) {
    if constexpr (OPT & 1) split_cell_la<DIR>(U, W, c, P, gamma, sv);
    else split_cell<DIR>(U, W, c, P, gamma, sv);
}

// ------------------------------------------------------------------------------------------
// face-flux kernels.  Face i+1/2 of the sweep direction is stored at cell i.  The sweep index
// of the faces runs from NGC-1 to NGC+n-1; the tile kernels cover the first n faces, the plane
// kernel the last one (or, for operator-level use, any plane range).
// ------------------------------------------------------------------------------------------
// region of the sample buffers (LearnSink) of this workgroup: its linear id modulo LS_REGIONS
__device__ __forceinline__ int learn_region() {
    const unsigned id = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    return (int)(id % LS_REGIONS);
}

constexpr int TILE_S = 8;    // x/y tiles: faces along the sweep per block (one wave each)
constexpr int TILE_L = 64;   // lanes along k
// sweep tile of the single-block x/y kernels of convection type CONV: 8 faces per block (one wave
// each); the single-precision characteristic sweeps take 16 (1024-thread blocks), where the 32 KB
// tile of 8 faces leaves room and the 5 extra cells per tile weigh more (the component-wise
// sweeps, with less arithmetic per face, run faster with 8)
template <int CONV> constexpr int tile_s1() {
#ifdef NS_TILE_S1
    return NS_TILE_S1;
#else
    return (NS_FP32 && CONV == 2) ? 16 : 8;
#endif
}
#ifndef NS_TILE_WAVES
#define NS_TILE_WAVES 1      // minimum waves per SIMD requested for the tile kernels
#endif
#ifndef NS_TILE1_WAVES
#define NS_TILE1_WAVES NS_TILE_WAVES  // the same for the single-block x/y tile kernel
#endif
#ifndef NS_LEARN_WAVES
#define NS_LEARN_WAVES 0  // the same for the single-block tiles with learned weights (0: as above)
#endif
// waves per SIMD requested for a single-block tile kernel with weights WT (base: the classic value)
template <int WT> constexpr int tile_waves(int base) {
    return (WT == 2 && NS_LEARN_WAVES > 0) ? NS_LEARN_WAVES : base;
}

// x or y sweep: 8 waves at 8 consecutive sweep positions; a wave covers kw consecutive k (unit
// stride) of 64/kw consecutive rows of the other tangential index (kw = 64: one row).  The split
// values of the 13 cells the tile needs are computed once into LDS.  (bx, by, bz) = (k tile,
// row tile, sweep tile).  [lo, hi): for the x sweep the faces of the launch (face f between cells
// f - 1 and f; hi <= nx), for the y sweep its planes i (a slab of the block); the full range is
// [0, n) of the sweep resp. of the other index.
// EMIT (learned weights, component-wise): the network samples go to *le (emit mode)
template <int DIR, int CONV, int WT, int OPT, int TS = TILE_S, bool EMIT = false,
          class ST = double>
__device__ __forceinline__ void flux_tile_xy_body(const ST* __restrict__ U,
                                                  const ST* __restrict__ W,
                                                  flux_t* __restrict__ f, const Geo& g,
                                                  real gamma, int bx, int by, int bz, int kw,
                                                  real* __restrict__ tile,
                                                  const MlpPacked* __restrict__ mlp, int lo,
                                                  int hi, real* lq = nullptr,
                                                  const LearnEmit* le = nullptr) {
    constexpr int NC = TS + 5;
    constexpr int NV = split_nv<OPT>();
    const int lane = threadIdx.x, w = threadIdx.y;
    const int rows = TILE_L / kw;
    const int k    = NGC + bx * kw + (lane & (kw - 1));
    // j for the x sweep, i for the y sweep
    const int o    = NGC + (DIR == 1 ? lo : 0) + by * rows + lane / kw;
    const int s0   = (DIR == 0 ? lo : 0) + bz * TS;  // first face of the tile (0-based)
    const int n    = (DIR == 0) ? hi : g.ny;          // end of the faces of this launch
    const int no   = (DIR == 0) ? g.ny : hi;          // end of the other index
    const int  nsw = (DIR == 0) ? g.nx : g.ny;      // cells along the sweep
    const long ss  = (DIR == 0) ? g.sx : g.sy;      // stride along the sweep
    const long so  = (DIR == 0) ? g.sy : g.sx;      // stride of the other index
    const bool kin = (k < NGC + g.nz) && (o < NGC + no);
    // cells s0 .. s0+12 (padded sweep index) feed faces s0+2 .. s0+9
    for (int t = w; t < NC; t += TS) {
        int cs = s0 + t;
        if (kin && cs <= nsw + 2 * NGC - 1) {
            real sv[NV];
            stage_cell<DIR, OPT>(U, W, (long)cs * ss + (long)o * so + k, g.P, gamma, sv);
#pragma unroll
            for (int v = 0; v < NV; v++) tile[(t * NV + v) * TILE_L + lane] = sv[v];
        }
    }
    __syncthreads();
    const int  fidx  = s0 + w;  // face index 0..n-1
    const bool valid = kin && fidx < n;
    // the learned weights need every lane of the wavefront: lanes without a face compute the
    // first face of the block from in-bounds data and store nothing
    if constexpr (WT != 2) {
        if (!valid) return;
    }
    const int  i = (valid ? fidx : (DIR == 0 ? lo : 0)) + NGC - 1;  // face i+1/2
    const long c = (long)i * ss + (long)(valid ? o : NGC + (DIR == 1 ? lo : 0)) * so +
                   (valid ? k : NGC);
    real     Wl[5], Wr[5];
    if constexpr (CONV == 2) {
#pragma unroll
        for (int q = 0; q < 5; q++) Wl[q] = W[c + q * g.P], Wr[q] = W[c + ss + q * g.P];
    }
    LdsSplit<TILE_L, NV> st{&tile[(w * NV) * TILE_L + lane]};
    real               out[5];
    if constexpr (EMIT) {
        face_flux<DIR, CONV, WT, OPT, true>(st, Wl, Wr, gamma, out,
                                            Learn{mlp, valid, nullptr, *le, (unsigned)c,
                                                  (unsigned)g.P, learn_region()});
    } else {
        face_flux<DIR, CONV, WT, OPT>(st, Wl, Wr, gamma, out,
                                      Learn{mlp, valid, lq ? lq + w * LQ_REALS : nullptr});
    }
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < 5; q++) f[c + q * g.P] = out[q];
}

template <int DIR, int CONV, int WT, int OPT, class ST>
__global__ __launch_bounds__(64 * tile_s1<CONV>(), tile_waves<WT>(NS_TILE1_WAVES)) void k_flux_tile_xy(
    const ST* __restrict__ U, const ST* __restrict__ W, flux_t* __restrict__ f, Geo g,
    real gamma, const MlpPacked* mlp, int lo, int hi) {
    constexpr int TS = tile_s1<CONV>();
    __shared__ real tile[(TS + 5) * split_nv<OPT>() * TILE_L];
    __shared__ real lq[WT == 2 ? TS * LQ_REALS : 1];  // network queues (learned weights)
    flux_tile_xy_body<DIR, CONV, WT, OPT, TS>(U, W, f, g, gamma, blockIdx.x, blockIdx.y,
                                                   blockIdx.z, TILE_L, tile, mlp, lo, hi,
                                                   WT == 2 ? lq : nullptr);
}

// z sweep: lanes along k = the sweep.  A wave covers 64/kw rows (j) with a segment of kw faces
// each, a block 8 waves; each row stages the split values of its kw + 5 cells in LDS.
// (bx, by, bz) = (segment, row tile, i).
// the largest per-wave need (kw = 16: 4 rows of 21 cells)
template <int OPT> constexpr int zlds_wave_max() { return 4 * (16 + 5) * split_nv<OPT>(); }
template <int CONV, int WT, int OPT, bool EMIT = false, class ST = double>
__device__ __forceinline__ void flux_tile_z_body(const ST* __restrict__ U,
                                                 const ST* __restrict__ W,
                                                 flux_t* __restrict__ f, const Geo& g,
                                                 real gamma, int bx, int by, int bz, int kw,
                                                 real* __restrict__ lds,
                                                 const MlpPacked* __restrict__ mlp,
                                                 real* lq = nullptr,
                                                 const LearnEmit* le = nullptr) {
    constexpr int NV = split_nv<OPT>();
    const int NC   = kw + 5;
    const int sub  = TILE_L / kw;                   // rows per wave
    const int lane = threadIdx.x, r = threadIdx.y;
    const int kk   = lane & (kw - 1);
    const int sr   = lane / kw;
    const int j    = NGC + (by * 8 + r) * sub + sr;
    const int i    = NGC + bz;
    const int f0   = bx * kw;                       // first face of the segment (0-based)
    const bool rowok = (j < NGC + g.ny);            // every wave reaches the barrier below
    const long row = (long)i * g.sx + (long)j * g.sy;
    real*      T   = lds + (r * sub + sr) * (NC * NV);
    for (int t = kk; t < NC; t += kw) {
        int cs = f0 + t;  // padded k of the cell
        if (rowok && cs <= g.nz + 2 * NGC - 1) {
            real sv[NV];
            stage_cell<2, OPT>(U, W, row + cs, g.P, gamma, sv);
#pragma unroll
            for (int v = 0; v < NV; v++) T[v * NC + t] = sv[v];
        }
    }
    __syncthreads();  // rows are independent; each wave only reads its own rows
    const int  fidx  = f0 + kk;
    const bool valid = rowok && fidx < g.nz;
    if constexpr (WT != 2) {  // (see flux_tile_xy_body)
        if (!valid) return;
    }
    const int  kf = (valid ? fidx : 0) + NGC - 1;
    const long c  = (valid ? row : (long)i * g.sx + (long)NGC * g.sy) + kf;
    real     Wl[5], Wr[5];
    if constexpr (CONV == 2) {
#pragma unroll
        for (int q = 0; q < 5; q++) Wl[q] = W[c + q * g.P], Wr[q] = W[c + 1 + q * g.P];
    }
    struct {
        const real* b;
        int         nc;
        __device__ real operator()(int ii, int v) const { return b[v * nc + ii]; }
    } st{&T[kk], NC};
    real out[5];
    if constexpr (EMIT) {
        face_flux<2, CONV, WT, OPT, true>(st, Wl, Wr, gamma, out,
                                          Learn{mlp, valid, nullptr, *le, (unsigned)c,
                                                (unsigned)g.P, learn_region()});
    } else {
        face_flux<2, CONV, WT, OPT>(st, Wl, Wr, gamma, out,
                                    Learn{mlp, valid, lq ? lq + r * LQ_REALS : nullptr});
    }
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < 5; q++) f[c + q * g.P] = out[q];
}

// planes i = lo + blockIdx.z
// waves per SIMD requested for the single-block z tile: the single-precision tiles without options
// and without the in-kernel network fit 8 (the characteristic one in 64 VGPRs instead of 66, no
// spills: its sweep 3 % faster)
template <int WT, int OPT> constexpr int ztile_waves() {
    return (NS_FP32 && WT != 2 && OPT == 0) ? 8 : tile_waves<WT>(NS_TILE_WAVES);
}
template <int CONV, int WT, int OPT, class ST>
__global__ __launch_bounds__(512, (ztile_waves<WT, OPT>())) void k_flux_tile_z(const ST* __restrict__ U,
                                                     const ST* __restrict__ W,
                                                     flux_t* __restrict__ f, Geo g,
                                                     real gamma, const MlpPacked* mlp, int lo) {
    __shared__ real tile[8 * (TILE_L + 5) * split_nv<OPT>()];
    __shared__ real lq[WT == 2 ? 8 * LQ_REALS : 1];
    flux_tile_z_body<CONV, WT, OPT>(U, W, f, g, gamma, blockIdx.x, blockIdx.y, lo + blockIdx.z,
                                    TILE_L, tile, mlp, WT == 2 ? lq : nullptr);
}

// emit-mode tiles of the learned component-wise sweeps (no network queue in LDS) and the
// network pass over one side of their samples
template <int DIR, class ST>
__global__ __launch_bounds__(64 * tile_s1<1>(), NS_TILE1_WAVES) void k_flux_tile_xy_emit(
    const ST* __restrict__ U, const ST* __restrict__ W, flux_t* __restrict__ f, Geo g,
    real gamma, int lo, int hi, LearnEmit e) {
    __shared__ real tile[(tile_s1<1>() + 5) * split_nv<0>() * TILE_L];
    learn_zero_next(e);
    flux_tile_xy_body<DIR, 1, 2, 0, tile_s1<1>(), true>(U, W, f, g, gamma, blockIdx.x, blockIdx.y,
                                                   blockIdx.z, TILE_L, tile, nullptr, lo, hi,
                                                   nullptr, &e);
}
template <class ST>
__global__ __launch_bounds__(512, NS_TILE_WAVES) void k_flux_tile_z_emit(
    const ST* __restrict__ U, const ST* __restrict__ W, flux_t* __restrict__ f, Geo g,
    real gamma, int lo, LearnEmit e) {
    __shared__ real tile[8 * (TILE_L + 5) * split_nv<0>()];
    learn_zero_next(e);
    flux_tile_z_body<1, 2, 0, true>(U, W, f, g, gamma, blockIdx.x, blockIdx.y, lo + blockIdx.z,
                                    TILE_L, tile, nullptr, nullptr, &e);
}
#ifndef NS_NET_WAVES
#define NS_NET_WAVES 5  // minimum waves per SIMD requested for the network pass (95 VGPRs)
#endif
__global__ __launch_bounds__(256, NS_NET_WAVES) void k_learn_net(const MlpPacked* __restrict__ P, LearnEmit e,
                                                   int side, flux_t* __restrict__ f) {
    learn_net_body(P, e, side, f);
}

// one plane of faces (sweep index `i`), one thread per face, split values in registers;
// block (64, 4) over the tangential indices, (bx, by) = (fast, slow) tile
template <int DIR, int CONV, int WT, int OPT, class ST = double>
__device__ __forceinline__ void flux_plane_body(const ST* __restrict__ U,
                                                const ST* __restrict__ W,
                                                flux_t* __restrict__ f, const Geo& g, real gamma,
                                                int i, int bx, int by,
                                                const MlpPacked* __restrict__ mlp,
                                                real* lq = nullptr) {
    // tangential indices (t1 slow, t2 fast) over the interior
    int        t2    = threadIdx.x + bx * 64;
    int        t1    = threadIdx.y + by * 4;
    int        n1    = (DIR == 0) ? g.ny : g.nx;
    int        n2    = (DIR == 2) ? g.ny : g.nz;
    const bool valid = t1 < n1 && t2 < n2;
    if constexpr (WT != 2) {  // (see flux_tile_xy_body)
        if (!valid) return;
    }
    if (!valid) t1 = 0, t2 = 0;
    long c, s;
    if (DIR == 0) c = (long)i * g.sx + (long)(t1 + NGC) * g.sy + (t2 + NGC), s = g.sx;
    else if (DIR == 1) c = (long)(t1 + NGC) * g.sx + (long)i * g.sy + (t2 + NGC), s = g.sy;
    else c = (long)(t1 + NGC) * g.sx + (long)(t2 + NGC) * g.sy + i, s = 1;
    constexpr int NV = split_nv<OPT>();
    real          sv[6][NV];
#pragma unroll
    for (int ii = 0; ii < 6; ii++) stage_cell<DIR, OPT>(U, W, c + (ii - 2) * s, g.P, gamma, sv[ii]);
    real Wl[5], Wr[5];
    if constexpr (CONV == 2) {
#pragma unroll
        for (int q = 0; q < 5; q++) Wl[q] = W[c + q * g.P], Wr[q] = W[c + s + q * g.P];
    }
    RegSplit<NV> st{sv};
    real         out[5];
    face_flux<DIR, CONV, WT, OPT>(st, Wl, Wr, gamma, out,
                                  Learn{mlp, valid, lq ? lq + threadIdx.y * LQ_REALS : nullptr});
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < 5; q++) f[c + q * g.P] = out[q];
}

// by0: first row tile (of 4 rows of the slow tangential index) of the launch
template <int DIR, int CONV, int WT, int OPT, class ST>
__global__ __launch_bounds__(256) void k_flux_plane(const ST* __restrict__ U,
                                                    const ST* __restrict__ W,
                                                    flux_t* __restrict__ f, Geo g, real gamma,
                                                    int i, const MlpPacked* mlp, int by0) {
    __shared__ real lq[WT == 2 ? 4 * LQ_REALS : 1];
    flux_plane_body<DIR, CONV, WT, OPT>(U, W, f, g, gamma, i, blockIdx.x, by0 + blockIdx.y, mlp,
                                        WT == 2 ? lq : nullptr);
}

// ------------------------------------------------------------------------------------------
// equation of state (eq_of_state.cpp:32-101)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ void c2p(const double u[5], double gamma, double w[5]) {
    w[0]       = u[0];
    Rcp r0     = rcp(u[0]);
    w[2]       = dv(u[2], r0);
    w[3]       = dv(u[3], r0);
    w[4]       = dv(u[4], r0);
    double kin = 0.5 * (w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
    w[1]       = (gamma - 1.0) * (u[1] - u[0] * kin);
}

__device__ __forceinline__ void p2c(const double w[5], double gamma, double u[5]) {
    u[0]       = w[0];
    u[2]       = w[0] * w[2];
    u[3]       = w[0] * w[3];
    u[4]       = w[0] * w[4];
    double kin = 0.5 * w[0] * (w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
    u[1]       = dv(w[1], gamma - 1.0) + kin;
}

// health of a new stage state (the step's failure detection): bit 0 a non-finite conserved or
// primitive value, bit 1 rho <= 0 or p <= 0 (finite).  The wavefront's verdicts are combined with
// ballots and its first active lane sets them with one atomic OR, only in a wavefront that holds
// such a cell, so a healthy stage adds compares and two ballots per cell and no memory traffic.
__device__ __forceinline__ void stage_health(unsigned* flags, const double u[5], const double w[5]) {
    bool nf = false;
#pragma unroll
    for (int q = 0; q < 5; q++)
        nf = nf || !(__builtin_fabs(u[q]) <= 1.7976931348623157e308) ||
             !(__builtin_fabs(w[q]) <= 1.7976931348623157e308);
    const bool np = !nf && !(w[0] > 0.0 && w[1] > 0.0);
    const unsigned long long a = __ballot(nf), b = __ballot(np);
    if ((a | b) == 0ull) return;
    const unsigned long long act = __ballot(1);
    if (__lane_id() == (unsigned)(__ffsll((unsigned long long)act) - 1))
        atomicOr(flags, (a ? 1u : 0u) | (b ? 2u : 0u));
}

__device__ __forceinline__ double local_dt(const double w[5], double gamma, double dx, double dy,
                                           double dz) {
    double c  = sqrt(dv(gamma * w[1], w[0]));
    double ax = fabs(w[2] + c), bx = fabs(w[2] - c);
    double ay = fabs(w[3] + c), by = fabs(w[3] - c);
    double az = fabs(w[4] + c), bz = fabs(w[4] - c);
    double ex  = dv(dx, (ax < bx) ? bx : ax);  // std::max
    double ey  = dv(dy, (ay < by) ? by : ay);
    double ez  = dv(dz, (az < bz) ? bz : az);
    double tmp = (ey < ex) ? ey : ex;  // std::min
    return (ez < tmp) ? ez : tmp;
}

// min over a workgroup of positive doubles, then one atomicMin on the bit pattern (positive
// IEEE doubles order like their unsigned 64-bit patterns; NaN sorts above every finite value,
// matching `if (dt < max_dt)` in time_scheme.cpp:119 which never picks a NaN)
__device__ __forceinline__ void block_min_atomic(double v, unsigned long long* out) {
    __shared__ double red[16];
    for (int o = 32; o > 0; o >>= 1) {
        double t = __shfl_xor(v, o, 64);
        v        = (t < v) ? t : v;
    }
    int tid  = threadIdx.x + threadIdx.y * blockDim.x;
    int lane = tid & 63;
    int wid  = tid >> 6;
    int nw   = (blockDim.x * blockDim.y + 63) >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (wid == 0) {
        double x = (lane < nw) ? red[lane] : 1E300;
        for (int o = 8; o > 0; o >>= 1) {
            double t = __shfl_xor(x, o, 64);
            x        = (t < x) ? t : x;
        }
        if (lane == 0 && x < 1E300) atomicMin(out, (unsigned long long)__double_as_longlong(x));
    }
}

// the same from the calling lanes of one wavefront, without a barrier (the lanes of a workgroup
// need not all be present): a butterfly when every lane takes part, else one atomic per lane
__device__ __forceinline__ void wave_min_atomic(double v, unsigned long long* out) {
    if (__ballot(1) == ~0ull) {
        for (int o = 32; o > 0; o >>= 1) {
            double t = __shfl_xor(v, o, 64);
            v        = (t < v) ? t : v;
        }
        if ((threadIdx.x & 63) == 0 && v < 1E300)
            atomicMin(out, (unsigned long long)__double_as_longlong(v));
    } else if (v < 1E300) {
        atomicMin(out, (unsigned long long)__double_as_longlong(v));
    }
}

// the next step's prologue of one cell from this step's final primitives w (step_begin_body's
// arithmetic): U and c2p(U) stored at cell cc, the cell's CFL limit folded into dtl
// CP: also their float copies into st.Uf, st.Wf
template <bool CP>
__device__ __forceinline__ void next_prologue(const StageArgs& st, const double w[5], long cc,
                                              long P, const Geo& g, double gamma, double& dtl) {
    const double d = local_dt(w, gamma, g.dx, g.dy, g.dz);
    dtl            = (d < dtl) ? d : dtl;
    double u[5], w2[5];
    p2c(w, gamma, u);
    c2p(u, gamma, w2);
#pragma unroll
    for (int q = 0; q < 5; q++) {
        st.Unext[q * P + cc] = u[q], st.Wnext[q * P + cc] = w2[q];
        if constexpr (CP) st.Uf[q * P + cc] = (float)u[q], st.Wf[q * P + cc] = (float)w2[q];
    }
}

__global__ void k_reset_min(unsigned long long* bits) {
    *bits = (unsigned long long)__double_as_longlong(1E10);  // max_dt = 1E10 (time_scheme.cpp:104)
}

// a prescribed time step: the stages read dt = bits * CFL
__global__ void k_set_dt(unsigned long long* bits, double dt_over_cfl) {
    *bits = (unsigned long long)__double_as_longlong(dt_over_cfl);
}

__global__ void k_zero_flags(unsigned* flags) { *flags = 0u; }

// step prologue: dt from W (time_scheme.cpp:16, :98-136), U = p2c(W) (:25), Uprev = U (:27),
// and the stage-1 primitives c2p(U) (:36).  Workgroup `wg` of `nwg` strides over the interior.
// cp (may be null): float copies of Un and Wst (single-precision build)
__device__ __forceinline__ void step_begin_body(const double* __restrict__ W,
                                                double* __restrict__ Ua, double* __restrict__ Un,
                                                double* __restrict__ Wst,
                                                unsigned long long* dtmin, const Geo& g,
                                                double gamma, int wg, int nwg,
                                                float* __restrict__ Uf = nullptr,
                                                float* __restrict__ Wf = nullptr) {
    long   n  = (long)g.nx * g.ny * g.nz;
    double dt = 1E300;
    for (long t = threadIdx.x + (long)wg * blockDim.x; t < n; t += (long)nwg * blockDim.x) {
        int    k = (int)(t % g.nz);
        long   r = t / g.nz;
        int    j = (int)(r % g.ny);
        int    i = (int)(r / g.ny);
        long   c = (long)(i + NGC) * g.sx + (long)(j + NGC) * g.sy + (k + NGC);
        double w[5], u[5], w2[5];
#pragma unroll
        for (int q = 0; q < 5; q++) w[q] = W[c + q * g.P];
        double d = local_dt(w, gamma, g.dx, g.dy, g.dz);
        dt       = (d < dt) ? d : dt;
        p2c(w, gamma, u);
        c2p(u, gamma, w2);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            if (Ua) Ua[c + q * g.P] = u[q];
            Un[c + q * g.P] = u[q];
            Wst[c + q * g.P] = w2[q];
            if (Uf) Uf[c + q * g.P] = (float)u[q], Wf[c + q * g.P] = (float)w2[q];
        }
    }
    block_min_atomic(dt, dtmin);
}

__global__ __launch_bounds__(256) void k_step_begin(const double* __restrict__ W,
                                                    double* __restrict__ Ua,
                                                    double* __restrict__ Un,
                                                    double* __restrict__ Wst,
                                                    unsigned long long* dtmin, Geo g, double gamma,
                                                    float* Uf, float* Wf) {
    step_begin_body(W, Ua, Un, Wst, dtmin, g, gamma, blockIdx.x, gridDim.x, Uf, Wf);
}

// ------------------------------------------------------------------------------------------
// halo fill: Neumann faces (boundary_neumann.cpp:22-132) and the periodic face exchange of one
// rank (communicator.cpp:16-144 with the slicers of array_4d.cpp:39-164).  The six face regions
// are disjoint and read interior cells only, so one launch covers them in any order.  The same
// copy applied to the primitive planes gives c2p(U) on the halo bit for bit (c2p is per cell).
// ------------------------------------------------------------------------------------------
// block (64, 4); (bx, by, bz): bz = face, by = (row tile, layer), bx = lane tile
// A linked face (FACE_LINK) copies the neighbour's interior planes next to it (`lk`: its arrays and
// extent along the normal; same tangential layout), as Communicator::communicate does between
// ranks (communicator.cpp:16-144): faces only, like every other face.
// T: double, or float for the float copies of the single-precision build (no linked faces)
template <class T>
__device__ __forceinline__ void halo_body(T* A, int nA, T* B, int nB, const Geo& g,
                                          const Bnd& b, int bx, int by, int bz,
                                          const HaloLink* lk = nullptr) {
    int face = bz;
    int axis = face >> 1;
    const bool link = b.face_kind[face] == FACE_LINK;
    if (link && !(lk && lk->A)) return;  // a linked face is filled by its decomposition only
    // a 2-D block thinner than the halo (nz < 3): its z halos are never read (no z sweep, 2-D
    // viscous terms) and their periodic copy would read halo cells being written: left as is
    if (axis == 2 && g.nz < NGC) return;
    // face order W,E,N,S,F,B: low side for W (0), N (2), B (5); high side for E, S, F
    bool low = (face == 0 || face == 2 || face == 5);
    int  n   = (axis == 0) ? g.nx : (axis == 1) ? g.ny : g.nz;
    int  t1, t2, h;
    if (axis < 2) {
        // x / y faces: lanes along k (unit stride), one layer per `by` slice
        t2 = threadIdx.x + bx * 64;
        t1 = threadIdx.y + (by / NGC) * 4;
        h  = by % NGC;
        int n1 = (axis == 0) ? g.ny : g.nx;
        if (t1 >= n1 || t2 >= g.nz) return;
    } else {
        // z faces: the 3 halo cells of a row are contiguous; lanes over (row j, layer h)
        int lin = threadIdx.x + bx * 64;  // j * NGC + h
        t2      = lin / NGC;
        h       = lin % NGC;
        t1      = threadIdx.y + (by / NGC) * 4;
        if ((by % NGC) != 0 || t1 >= g.nx || t2 >= g.ny) return;
    }
    int dst = low ? h : NGC + n + h;
    int src;
    const T* sA = A;
    const T* sB = B;
    int      sx = g.sx, sy = g.sy;  // strides of the source
    long     sP = g.P;
    if (link) {
        if constexpr (std::is_same<T, double>::value) {
            src = low ? lk->n + h : NGC + h, sA = lk->A, sB = lk->B;
            sx = lk->sx, sy = lk->sy, sP = lk->P;
        } else {
            return;
        }
    } else if (b.periodic[axis]) {
        src = low ? n + h : NGC + h;
    } else {
        src = low ? NGC : NGC + n - 1;
    }
    long cd, cs;
    if (axis == 0) {
        cd = (long)dst * g.sx + (long)(t1 + NGC) * g.sy + (t2 + NGC);
        cs = (long)src * sx + (long)(t1 + NGC) * sy + (t2 + NGC);
    } else if (axis == 1) {
        cd = (long)(t1 + NGC) * g.sx + (long)dst * g.sy + (t2 + NGC);
        cs = (long)(t1 + NGC) * sx + (long)src * sy + (t2 + NGC);
    } else {
        cd = (long)(t1 + NGC) * g.sx + (long)(t2 + NGC) * g.sy + dst;
        cs = (long)(t1 + NGC) * sx + (long)(t2 + NGC) * sy + src;
    }
    for (int q = 0; q < nA; q++) A[cd + q * g.P] = sA[cs + q * sP];
    for (int q = 0; q < nB; q++) B[cd + q * g.P] = sB[cs + q * sP];
}

// faces f0 .. f0 + gridDim.z - 1
__global__ void k_halo(double* A, int nA, double* B, int nB, Geo g, Bnd b, int f0) {
    halo_body(A, nA, B, nB, g, b, blockIdx.x, blockIdx.y, f0 + (int)blockIdx.z);
}
// the same, and the same copies on the float copies Af, Bf (5 planes each) of A and B
__global__ void k_halo_cp(double* A, int nA, double* B, int nB, float* Af, float* Bf, Geo g,
                          Bnd b, int f0) {
    halo_body(A, nA, B, nB, g, b, blockIdx.x, blockIdx.y, f0 + (int)blockIdx.z);
    halo_body(Af, 5, Bf, 5, g, b, blockIdx.x, blockIdx.y, f0 + (int)blockIdx.z);
}

// W = c2p(U) on the interior and the face halos (edges and corners are never read)
__global__ __launch_bounds__(256) void k_c2p_faces(const double* __restrict__ U,
                                                   double* __restrict__ W, Geo g, double gamma) {
    int k = threadIdx.x + blockIdx.x * blockDim.x;
    int j = threadIdx.y + blockIdx.y * blockDim.y;
    int i = blockIdx.z;
    if (k >= g.NZ || j >= g.NY) return;
    int outside = (i < NGC || i >= NGC + g.nx) + (j < NGC || j >= NGC + g.ny) +
                  (k < NGC || k >= NGC + g.nz);
    if (outside > 1) return;
    long   c = (long)i * g.sx + (long)j * g.sy + k;
    double u[5], w[5];
#pragma unroll
    for (int q = 0; q < 5; q++) u[q] = U[c + q * g.P];
    c2p(u, gamma, w);
#pragma unroll
    for (int q = 0; q < 5; q++) W[c + q * g.P] = w[q];
}

// U = p2c(W) on the interior and the face halos (EqOfState::primitive_to_conservative,
// eq_of_state.cpp:32-59): the conservative state that goes with a primitive state set from the
// host
__global__ __launch_bounds__(256) void k_p2c_faces(const double* __restrict__ W,
                                                   double* __restrict__ U, Geo g, double gamma) {
    int k = threadIdx.x + blockIdx.x * blockDim.x;
    int j = threadIdx.y + blockIdx.y * blockDim.y;
    int i = blockIdx.z;
    if (k >= g.NZ || j >= g.NY) return;
    int outside = (i < NGC || i >= NGC + g.nx) + (j < NGC || j >= NGC + g.ny) +
                  (k < NGC || k >= NGC + g.nz);
    if (outside > 1) return;
    long   c = (long)i * g.sx + (long)j * g.sy + k;
    double w[5], u[5];
#pragma unroll
    for (int q = 0; q < 5; q++) w[q] = W[c + q * g.P];
    p2c(w, gamma, u);
#pragma unroll
    for (int q = 0; q < 5; q++) U[c + q * g.P] = u[q];
}

// ------------------------------------------------------------------------------------------
// viscous terms (centered_diffusion.cpp:3-182, cd-n.cpp:91-165, diffusion_scheme.cpp:3-58)
// ------------------------------------------------------------------------------------------

// one stencil sum of the first- (SECOND=0) or second-derivative coefficients; f(t) returns the
// value at offset t - H, t = 0 .. 2H.  EXACT: sum of f * c_k / d in k order, the zero centre term
// of the first derivative skipped (an exact +-0 addend); FAST: sum of f * (c_k / d) with the
// reciprocal of d hoisted.
template <int ORDER, bool SECOND, class T, class F>
__device__ __forceinline__ T fd_sum_t(const F& f, T d, T id) {
    constexpr int H     = ORDER / 2;
    constexpr int OI    = ORDER / 2 - 1;
    T             s     = T(0.);
    bool          first = true;
#if !NS_FAST
    // id is the correctly rounded 1 / d (host): the corrected quotient below is then the
    // correctly rounded one (the same value as the device's refined reciprocal gives), and a
    // loop-invariant divisor costs no reciprocal in the kernel
    const Rcp rd{d, id};
#endif
#pragma unroll
    for (int t = 0; t <= 2 * H; t++) {
        if (!SECOND && t == H) continue;
        T co = SECOND ? c_fd2[OI][t] : c_fd1[OI][t];
#if !NS_FAST
        T v = dv(f(t) * co, rd);
        (void)id;
#else
        T v = f(t) * (co * id);
        (void)d;
#endif
        if (first) {
            s = v, first = false;
        } else {
            s += v;
        }
    }
    return s;
}
template <int ORDER, bool SECOND, class F>
__device__ __forceinline__ real fd_sum(const F& f, real d, real id) {
    return fd_sum_t<ORDER, SECOND, real>(f, d, id);
}

// Kernels with x-direction stencils march along i: each thread owns one (j, k) column of a
// MARCH-plane chunk and keeps the x-window of the values it needs in registers, so every plane is
// fetched from HBM once instead of once per stencil point.
#ifndef NS_MARCH
#define NS_MARCH 16
#endif
#ifndef NS_KW_CELL
#define NS_KW_CELL 64
#endif
#ifndef NS_CELL_WAVES
#define NS_CELL_WAVES 1  // minimum waves per SIMD requested for the marching cell kernels
#endif
#ifndef NS_XCD_CELL
#define NS_XCD_CELL 1
#endif
constexpr int MARCH   = NS_MARCH;
constexpr int KW_CELL = NS_KW_CELL;  // lanes along k of the single-block cell kernels

// first derivatives of rho, p, u, v along x and y on the interior.  256 threads: kw lanes along
// k times 256/kw rows j; (bx, by, bz) = (k tile, j tile, chunk of `march` planes i).  FP64 in
// every build: the kernel is bound by its HBM traffic, and single-precision arithmetic on FP64
// planes only adds conversions (it ran 1.5x slower)
// DY = false: dW/dx of u and v only (planes 2, 3 of Dx; the residual kernel forms the rest)
template <int ORDER, bool DY = true>
__device__ __forceinline__ void deriv1_body(const double* __restrict__ W, double* __restrict__ Dx,
                                            double* __restrict__ Dy, const Geo& g, int bx, int by,
                                            int bz, int kw, int march, int nthreads = 256) {
    constexpr int H   = ORDER / 2;
    constexpr int Q0  = DY ? 0 : 2;  // first variable formed
    const int     tid = threadIdx.x + threadIdx.y * 64;
    const int     k   = (tid & (kw - 1)) + bx * kw;
    const int     j   = tid / kw + by * (nthreads / kw);
    const int     i0  = bz * march;
    if (k >= g.nz || j >= g.ny) return;
    const int  i1 = (i0 + march < g.nx) ? i0 + march : g.nx;
    const long c0 = (long)(j + NGC) * g.sy + (k + NGC);
    double   win[4][2 * H + 1];
#pragma unroll
    for (int t = 0; t < 2 * H; t++)
#pragma unroll
        for (int q = Q0; q < 4; q++) win[q][t] = W[q * g.P + c0 + (long)(i0 + NGC - H + t) * g.sx];
    for (int i = i0; i < i1; i++) {
        const long c = c0 + (long)(i + NGC) * g.sx;
#pragma unroll
        for (int q = Q0; q < 4; q++) win[q][2 * H] = W[q * g.P + c + (long)H * g.sx];
#pragma unroll
        for (int q = Q0; q < 4; q++) {
            const double* fy = W + q * g.P + c;
            Dx[c + q * g.P] =
                fd_sum_t<ORDER, false, double>([&](int t) { return win[q][t]; }, g.dx, g.idx);
            if constexpr (DY)
                Dy[c + q * g.P] = fd_sum_t<ORDER, false, double>(
                    [&](int t) { return fy[(long)(t - H) * g.sy]; }, g.dy, g.idy);
        }
#pragma unroll
        for (int q = Q0; q < 4; q++)
#pragma unroll
            for (int t = 0; t < 2 * H; t++) win[q][t] = win[q][t + 1];
    }
}

// tile of a single-block cell kernel.  With NS_XCD_CELL the workgroups, which the hardware
// hands round-robin to the 8 XCDs, are renumbered so that XCD x takes the contiguous tile range
// [x N/8, (x+1) N/8): neighbouring j tiles, which read each other's halo rows, then run on one
// XCD and share its L2
__device__ __forceinline__ void cell_tile(int& bx, int& by, int& bz) {
#if NS_XCD_CELL
    const unsigned gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
    unsigned       t  = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (n % 8 == 0) t = (t % 8) * (n / 8) + t / 8;
    bx = (int)(t % gx);
    t /= gx;
    by = (int)(t % gy);
    bz = (int)(t / gy);
#else
    bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
#endif
}

#ifndef NS_DERIV_ROWS
#define NS_DERIV_ROWS 4  // rows j per block of the single-block derivative kernel (64 k each)
#endif
constexpr int DERIV_ROWS = NS_DERIV_ROWS;
// DY = false: dW/dx of u and v only (the residual kernel forms the rest)
template <int ORDER, bool DY>
__global__ __launch_bounds__(64 * DERIV_ROWS, NS_CELL_WAVES) void k_deriv1(const double* __restrict__ W,
                                                double* __restrict__ Dx, double* __restrict__ Dy,
                                                Geo g, int z0) {
    int bx, by, bz;
    cell_tile(bx, by, bz);
    deriv1_body<ORDER, DY>(W, Dx, Dy, g, bx, by, z0 + bz, 64, MARCH, 64 * DERIV_ROWS);
}

// periodic exchange of the derivative halos that the cross derivatives read: the j halos of
// dW/dx and the i halos of dW/dy (centered_diffusion.cpp:22-24).  At a physical wall the halos
// stay zero: the arrays are zeroed once at allocation and never written there.
// block (64, 4); (bx, by, bz) = (k tile, row tile, (side, low/high, layer))
// A linked face copies the neighbour's derivative planes (`lk` of that face: its Dx in A, its Dy
// in B, its extent along the normal), as the derivative exchange between ranks does.
__device__ __forceinline__ void deriv_halo_body(double* Dx, double* Dy, const Geo& g,
                                                const Bnd& b, int bx, int by, int bz, int q0 = 0,
                                                const HaloLink* lk = nullptr) {
    int  k    = threadIdx.x + bx * 64;
    int  t1   = threadIdx.y + by * 4;
    int  side = bz / (2 * NGC);  // 0: Dx j-halo, 1: Dy i-halo
    int  h    = bz % NGC;
    bool low  = ((bz / NGC) & 1) == 0;
    if (k >= g.nz) return;
    // faces N (2), S (3) for the j halos, W (0), E (1) for the i halos
    const int  face = side == 0 ? (low ? 2 : 3) : (low ? 0 : 1);
    const bool link = b.face_kind[face] == FACE_LINK && lk && lk->A;
    // strides of the source (the neighbour's own, see HaloLink)
    const int  sx = link ? lk->sx : g.sx, sy = link ? lk->sy : g.sy;
    const long sP = link ? lk->P : g.P;
    if (side == 0) {
        if (!(b.periodic[1] || link) || t1 >= g.nx) return;
        int  dst = low ? h : NGC + g.ny + h;
        int  src = low ? (link ? lk->n : g.ny) + h : NGC + h;
        long cd  = (long)(t1 + NGC) * g.sx + (long)dst * g.sy + (k + NGC);
        long cs  = (long)(t1 + NGC) * sx + (long)src * sy + (k + NGC);
        const double* S = link ? lk->A : Dx;
        for (int q = q0; q < 4; q++) Dx[cd + q * g.P] = S[cs + q * sP];
    } else {
        if (!(b.periodic[0] || link) || t1 >= g.ny) return;
        int  dst = low ? h : NGC + g.nx + h;
        int  src = low ? (link ? lk->n : g.nx) + h : NGC + h;
        long cd  = (long)dst * g.sx + (long)(t1 + NGC) * g.sy + (k + NGC);
        long cs  = (long)src * sx + (long)(t1 + NGC) * sy + (k + NGC);
        const double* S = link ? lk->B : Dy;
        for (int q = 0; q < 4; q++) Dy[cd + q * g.P] = S[cs + q * sP];
    }
}

// Dy null: the j halos of dW/dx of u and v only; by0: the first row tile (4 planes i) of the launch
__global__ void k_deriv_halo(double* Dx, double* Dy, Geo g, Bnd b, int by0) {
    deriv_halo_body(Dx, Dy, g, b, blockIdx.x, by0 + blockIdx.y, blockIdx.z, Dy ? 0 : 2);
}

// CenteredDiffusion::compute_residual (centered_diffusion.cpp:70-182) at one cell, with the
// constant-viscosity zero terms (dmu = dlamb = 0) dropped: each is an exact +-0 addend.
// wx[q][t]: primitives rho, p, u, v at x offsets t - H; dyx[a][t]: dW/dy of u (a = 0), v (a = 1)
// at x offsets t - H; Dx, Dy: derivative planes at the cell.
// Sutherland law (extension, see nsgpu.h): ((mu (1 + C)) / (T + C)) (T sqrt(T)), T = gamma p / rho
__device__ __forceinline__ real mu_suth(real rho, real p, const Phys& ph) {
    real T = dv((real)ph.gamma * p, rho);
    return dv((real)ph.suth_k, T + (real)ph.suth_C) * (T * sqrt(T));
}

// dxy[a][t]: dW/dx of u (a = 0), v (a = 1) at y offsets t - H (FP64, as the derivative planes
// hold them); dxr, dxp, dyr, dyp: dW/dx and dW/dy of rho and p at the cell; dmu_dy (Sutherland):
// the y derivative of the viscosity at the cell
// muw (Sutherland, optional): the viscosity mu_suth(wx[0][t], wx[1][t]) at the x offsets,
// formed by a caller that carries it along its march
template <int ORDER, bool SQ, bool SUTH>
__device__ __forceinline__ void viscous_v(const real (&wx)[4][ORDER + 1],
                                          const real (&dyx)[2][ORDER + 1],
                                          const double (&dxy)[2][ORDER + 1], double dxr,
                                          double dxp, double dyr, double dyp, real dmu_dy_c,
                                          const Geo& g, const Phys& ph, real& ene, real& x_mom,
                                          real& y_mom, const real* muw = nullptr) {
    constexpr int H   = ORDER / 2;
    const int     rho = 0, p = 1, u = 2, v = 3;
    // second derivatives: ddf_dxdx and ddf_dydy both run the x stencil (cd-n.cpp:67-79); when
    // dx*dx == dy*dy the two sums are the same operations on the same operands
    real du_dxdx   = fd_sum<ORDER, true>([&](int t) { return wx[u][t]; }, g.dxx, g.idxx);
    real dv_dxdx   = fd_sum<ORDER, true>([&](int t) { return wx[v][t]; }, g.dxx, g.idxx);
    real drho_dxdx = fd_sum<ORDER, true>([&](int t) { return wx[rho][t]; }, g.dxx, g.idxx);
    real dp_dxdx   = fd_sum<ORDER, true>([&](int t) { return wx[p][t]; }, g.dxx, g.idxx);
    real du_dydy, dv_dydy, drho_dydy, dp_dydy;
    if constexpr (SQ) {
        du_dydy = du_dxdx, dv_dydy = dv_dxdx, drho_dydy = drho_dxdx, dp_dydy = dp_dxdx;
    } else {
        du_dydy   = fd_sum<ORDER, true>([&](int t) { return wx[u][t]; }, g.dyy, g.idyy);
        dv_dydy   = fd_sum<ORDER, true>([&](int t) { return wx[v][t]; }, g.dyy, g.idyy);
        drho_dydy = fd_sum<ORDER, true>([&](int t) { return wx[rho][t]; }, g.dyy, g.idyy);
        dp_dydy   = fd_sum<ORDER, true>([&](int t) { return wx[p][t]; }, g.dyy, g.idyy);
    }
    // cross derivatives: dW_dxdy = df_dy(dW_dx), dW_dydx = df_dx(dW_dy)
    real du_dxdy = fd_sum<ORDER, false>([&](int t) { return dxy[0][t]; }, g.dy, g.idy);
    real dv_dxdy = fd_sum<ORDER, false>([&](int t) { return dxy[1][t]; }, g.dy, g.idy);
    real du_dydx = fd_sum<ORDER, false>([&](int t) { return dyx[0][t]; }, g.dx, g.idx);
    real dv_dydx = fd_sum<ORDER, false>([&](int t) { return dyx[1][t]; }, g.dx, g.idx);

    real du_dx = dxy[0][H], dv_dx = dxy[1][H], drho_dx = dxr, dp_dx = dxp;
    real du_dy = dyx[0][H], dv_dy = dyx[1][H], drho_dy = dyr, dp_dy = dyp;
    real U = wx[u][H], V = wx[v][H], Pp = wx[p][H], RHO = wx[rho][H];

    real Dtauxx_dx, Dtauyy_dy, Dtauxy_dx, Dtauyx_dy, Tauxx, Tauxy, mu, dmu_dx = 0, dmu_dy = 0;
    if constexpr (!SUTH) {
        mu          = (real)ph.mu;
        real lamb   = (real)ph.lamb;
        real two_mu = (real)ph.two_mu;
        Dtauxx_dx   = two_mu * du_dxdx + lamb * (du_dxdx + dv_dydx);
        Dtauyy_dy   = two_mu * dv_dydy + lamb * (du_dxdy + dv_dydy);
        Dtauxy_dx = mu * (du_dydx + dv_dxdx);
        Dtauyx_dy = mu * (du_dydy + dv_dxdy);
        Tauxx     = two_mu * du_dx + lamb * (du_dx + dv_dy);
        Tauxy     = mu * (du_dy + dv_dx);
    } else {
        // viscosity and its first derivatives from the primitive stencils; the full expressions
        // of diffusion_scheme.cpp:3-58 with the dmu / dlamb terms (centered_diffusion.cpp:96-145)
        if (muw) {
            mu     = muw[H];
            dmu_dx = fd_sum<ORDER, false>([&](int t) { return muw[t]; }, g.dx, g.idx);
        } else {
            mu     = mu_suth(RHO, Pp, ph);
            dmu_dx = fd_sum<ORDER, false>(
                [&](int t) { return mu_suth(wx[rho][t], wx[p][t], ph); }, g.dx, g.idx);
        }
        dmu_dy = dmu_dy_c;
        real lamb     = (-RL(2.) / RL(3.)) * mu;
        real dlamb_dx = -(RL(2.) / RL(3.)) * dmu_dx;
        real dlamb_dy = -(RL(2.) / RL(3.)) * dmu_dy;
        Dtauxx_dx = RL(2.0) * dmu_dx * du_dx + RL(2.0) * mu * du_dxdx + dlamb_dx * (du_dx + dv_dy);
        Dtauxx_dx += lamb * (du_dxdx + dv_dydx);
        Dtauyy_dy = RL(2.0) * dmu_dy * dv_dy + RL(2.0) * mu * dv_dydy + dlamb_dy * (du_dx + dv_dy);
        Dtauyy_dy += lamb * (du_dxdy + dv_dydy);
        Dtauxy_dx = dmu_dx * (du_dy + dv_dx) + mu * (du_dydx + dv_dxdx);
        Dtauyx_dy = dmu_dy * (du_dy + dv_dx) + mu * (du_dydy + dv_dxdy);
        Tauxx     = RL(2.0) * mu * du_dx + lamb * (du_dx + dv_dy);
        Tauxy     = mu * (du_dy + dv_dx);
    }
    real Tauyy     = Tauxx;  // tauxx(...) as written at centered_diffusion.cpp:143
    real Tauyx     = Tauxy;
    real Dtauvx_dx = U * Dtauxx_dx + V * Dtauxy_dx + Tauxx * du_dx + Tauxy * dv_dy;
    real Dtauvy_dy = U * Dtauyx_dy + V * Dtauyy_dy + Tauyy * dv_dy + Tauyx * du_dy;
#if !NS_FAST
    real one_rho  = dv(RL(1.), RHO);
    real one_rho2 = dv(RL(1.), RHO * RHO);
    real one_rho3 = dv(RL(1.), RHO * RHO * RHO);
#else
    real one_rho  = dv(RL(1.), RHO);
    real one_rho2 = one_rho * one_rho;
    real one_rho3 = one_rho2 * one_rho;
#endif
    real temp2x = mu * (RL(2.0) * one_rho2 * drho_dx * dp_dx + one_rho * dp_dxdx +
                          one_rho3 * Pp * drho_dx * drho_dx - one_rho2 * Pp * drho_dxdx);
    // the y heat-flux term multiplies by the enum value p = 1 (centered_diffusion.cpp:166)
    real temp2y = mu * (RL(2.0) * one_rho2 * drho_dy * dp_dy + one_rho * dp_dydy +
                          one_rho3 * drho_dy * drho_dy - one_rho2 * Pp * drho_dydy);
    real Dqx_dx, Dqy_dy;
    if constexpr (!SUTH) {
        Dqx_dx = -(real)ph.heat * temp2x;
        Dqy_dy = -(real)ph.heat * temp2y;
    } else {
        real temp1x = dmu_dx * (one_rho * dp_dx - Pp * one_rho2 * drho_dx);
        real temp1y = dmu_dy * (one_rho * dp_dy - Pp * one_rho2 * drho_dy);
        Dqx_dx        = -(real)ph.heat * (temp1x + temp2x);
        Dqy_dy        = -(real)ph.heat * (temp1y + temp2y);
    }
    x_mom = -Dtauxx_dx - Dtauyx_dy;
    y_mom = -Dtauyy_dy - Dtauxy_dx;
    ene   = -Dtauvx_dx - Dtauvy_dy + Dqx_dx + Dqy_dy;
}

// the y derivative of the Sutherland viscosity at a cell; f(t) returns (rho, p) at y offset t - H
template <int ORDER, class F>
__device__ __forceinline__ real dmu_dy_of(const F& f, const Geo& g, const Phys& ph) {
    return fd_sum<ORDER, false>(
        [&](int t) {
            double r, q;
            f(t, r, q);
            return mu_suth(r, q, ph);
        },
        g.dy, g.idy);
}

// viscous_v with its inputs read from the derivative planes Dx, Dy (4 planes each: rho, p, u, v)
// and, for the Sutherland law, the y stencil of rho and p from W
template <int ORDER, bool SQ, bool SUTH>
__device__ __forceinline__ void viscous(const real (&wx)[4][ORDER + 1],
                                        const real (&dyx)[2][ORDER + 1],
                                        const double* __restrict__ W,
                                        const double* __restrict__ Dx,
                                        const double* __restrict__ Dy, long c, const Geo& g,
                                        const Phys& ph, real& ene, real& x_mom,
                                        real& y_mom) {
    constexpr int H = ORDER / 2;
    const long    P = g.P;
    double        dxy[2][ORDER + 1];
#pragma unroll
    for (int t = 0; t <= ORDER; t++) {
        dxy[0][t] = Dx[2 * P + c + (long)(t - H) * g.sy];
        dxy[1][t] = Dx[3 * P + c + (long)(t - H) * g.sy];
    }
    real dmu_dy = 0;
    if constexpr (SUTH)
        dmu_dy = dmu_dy_of<ORDER>(
            [&](int t, double& r, double& q) {
                r = W[c + (long)(t - H) * g.sy];
                q = W[P + c + (long)(t - H) * g.sy];
            },
            g, ph);
    viscous_v<ORDER, SQ, SUTH>(wx, dyx, dxy, Dx[c], Dx[P + c], Dy[c], Dy[P + c], dmu_dy, g, ph,
                               ene, x_mom, y_mom);
}

// residual + stage update, marching along i:
// R = R_in + (fx_i - fx_{i-1})/dx + (fy_j - fy_{j-1})/dy + (fz_k - fz_{k-1})/dz
// (Weno5::compute_residual, weno-5.cpp:39-73), then the viscous terms (centered_diffusion.cpp:
// 173-176), then now = prev*m1 + now*m2 + R*m3 (TimeScheme::substep, time_scheme.cpp:72-96)
// and the primitives of the new state (eq_of_state.cpp:61-87).  Without a stage (st.Uout null)
// R is stored instead (operator-level calls).
// 256 threads as in deriv1_body: (bx, by, bz) = (k tile, j tile, chunk of `march` planes i)
// VC (full 3-D viscous terms): the viscous terms were formed by k_visc3d and Dx holds them
// (4 planes: energy, x, y, z momentum); Dy is unused
template <int ORDER, bool SQ, bool SUTH, bool VC = false>
__device__ __forceinline__ void residual_body(
    const flux_t* __restrict__ fx, const flux_t* __restrict__ fy, const flux_t* __restrict__ fz,
    const double* __restrict__ W, const double* __restrict__ Dx, const double* __restrict__ Dy,
    const double* __restrict__ Rin, double* __restrict__ Rout, const StageArgs& st,
    const unsigned long long* __restrict__ dtmin, int which, const Geo& g, const Phys& ph, int bx,
    int by, int bz, int kw, int march) {
    constexpr int O   = (ORDER > 0) ? ORDER : 2;  // window size - 1 (unused without diffusion)
    constexpr int H   = O / 2;
    const int     tid = threadIdx.x + threadIdx.y * 64;
    const int     k   = (tid & (kw - 1)) + bx * kw;
    const int     j   = tid / kw + by * (256 / kw);
    const int     i0  = bz * march;
    if (k >= g.nz || j >= g.ny) return;
    const int  i1   = (i0 + march < g.nx) ? i0 + march : g.nx;
    const long P    = g.P;
    const long c0   = (long)(j + NGC) * g.sy + (k + NGC);
    const bool conv = (which & 1) && ph.conv;
    const bool visc = (ORDER > 0) && (which & 2);
    real       wx[4][O + 1], dyx[2][O + 1];
    double     fxm[5];
    if (visc && !VC) {
#pragma unroll
        for (int t = 0; t < 2 * H; t++) {
            const long ct = c0 + (long)(i0 + NGC - H + t) * g.sx;
#pragma unroll
            for (int q = 0; q < 4; q++) wx[q][t] = W[q * P + ct];
            dyx[0][t] = Dy[2 * P + ct];
            dyx[1][t] = Dy[3 * P + ct];
        }
    }
    if (conv) {
#pragma unroll
        for (int q = 0; q < 5; q++) fxm[q] = fx[q * P + c0 + (long)(i0 + NGC - 1) * g.sx];
    }
    double dt = 0, m1 = 0, m2 = 0, m3 = 0;
    if (st.Uout) {
        // SSP-RK3 stage (time_scheme.cpp:37-63): dt = CFL * min local dt (time_scheme.cpp:129)
        dt = __longlong_as_double((long long)*dtmin) * ph.cfl;
        if (st.stage == 0) {
            m1 = 1.0, m2 = 0.0, m3 = -dt;
        } else if (st.stage == 1) {
            m1 = 3.0 / 4.0, m2 = 1.0 / 4.0, m3 = -(1. / 4.) * dt;
        } else {
            m1 = 1. / 3., m2 = 2. / 3., m3 = -(2. / 3.) * dt;
        }
    }
    for (int i = i0; i < i1; i++) {
        const long c = c0 + (long)(i + NGC) * g.sx;
        double     R[5];
        bool       have[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            have[q] = (Rin != nullptr);
            R[q]    = Rin ? Rin[c + q * P] : 0.0;
        }
        if (conv) {
#pragma unroll
            for (int q = 0; q < 5; q++) {
                double fc = fx[c + q * P];
#if !NS_FAST
                double a = dv(fc - fxm[q], Rcp{g.dx, g.idx});
                double b = dv((double)fy[c + q * P] - (double)fy[c - g.sy + q * P], Rcp{g.dy, g.idy});
#else
                double a = (fc - fxm[q]) * g.idx;
                double b = ((double)fy[c + q * P] - (double)fy[c - g.sy + q * P]) * g.idy;
#endif
                fxm[q] = fc;
                R[q]   = have[q] ? R[q] + a : a;
                R[q] += b;
                if (g.dim == 3) {
#if !NS_FAST
                    R[q] += dv((double)fz[c + q * P] - (double)fz[c - 1 + q * P], Rcp{g.dz, g.idz});
#else
                    R[q] += ((double)fz[c + q * P] - (double)fz[c - 1 + q * P]) * g.idz;
#endif
                }
                have[q] = true;
            }
        }
        if constexpr (ORDER > 0 && VC) {
            if (visc) {
                const double ene = Dx[c], xm = Dx[P + c], ym = Dx[2 * P + c], zm = Dx[3 * P + c];
                R[1]    = have[1] ? R[1] + ene : ene;
                R[2]    = have[2] ? R[2] + xm : xm;
                R[3]    = have[3] ? R[3] + ym : ym;
                R[4]    = have[4] ? R[4] + zm : zm;
                have[1] = have[2] = have[3] = have[4] = true;
            }
        } else if constexpr (ORDER > 0) {
            if (visc) {
                const long cn = c + (long)H * g.sx;  // plane entering the window
#pragma unroll
                for (int q = 0; q < 4; q++) wx[q][2 * H] = W[q * P + cn];
                dyx[0][2 * H] = Dy[2 * P + cn];
                dyx[1][2 * H] = Dy[3 * P + cn];
                real ene, xm, ym;
                viscous<ORDER, SQ, SUTH>(wx, dyx, W, Dx, Dy, c, g, ph, ene, xm, ym);
                R[1]    = have[1] ? R[1] + ene : ene;
                R[2]    = have[2] ? R[2] + xm : xm;
                R[3]    = have[3] ? R[3] + ym : ym;
                have[1] = have[2] = have[3] = true;
#pragma unroll
                for (int t = 0; t < 2 * H; t++) {
#pragma unroll
                    for (int q = 0; q < 4; q++) wx[q][t] = wx[q][t + 1];
                    dyx[0][t] = dyx[0][t + 1];
                    dyx[1][t] = dyx[1][t + 1];
                }
            }
        }
        if (st.Uout == nullptr) {
#pragma unroll
            for (int q = 0; q < 5; q++) Rout[c + q * P] = have[q] ? R[q] : 0.0;
            continue;
        }
        double un[5], w[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            double r = have[q] ? R[q] : 0.0;
#if !NS_FAST
            un[q] = (st.Un[c + q * P] * m1) + (st.Ucur[c + q * P] * m2) + (r * m3);
#else
            un[q] = st.Un[c + q * P] * m1 + st.Ucur[c + q * P] * m2 + r * m3;
#endif
            st.Uout[c + q * P] = un[q];
        }
        c2p(un, ph.gamma, w);
#pragma unroll
        for (int q = 0; q < 5; q++) st.Wout[c + q * P] = w[q];
        if (st.flags) stage_health(st.flags, un, w);
    }
}

template <int ORDER, bool SQ, bool SUTH, bool VC = false>
__global__ __launch_bounds__(256, NS_CELL_WAVES) void k_residual(
    const flux_t* __restrict__ fx, const flux_t* __restrict__ fy, const flux_t* __restrict__ fz,
    const double* __restrict__ W, const double* __restrict__ Dx, const double* __restrict__ Dy,
    const double* __restrict__ Rin, double* __restrict__ Rout, StageArgs st,
    const unsigned long long* __restrict__ dtmin, int which, Geo g, Phys ph) {
    int bx, by, bz;
    cell_tile(bx, by, bz);
    residual_body<ORDER, SQ, SUTH, VC>(fx, fy, fz, W, Dx, Dy, Rin, Rout, st, dtmin, which, g, ph,
                                       bx, by, bz, KW_CELL, MARCH);
}

// ------------------------------------------------------------------------------------------
// residual + stage update of a single block with the y derivatives of the 2-D viscous terms
// formed in the same kernel.  The separate path writes dW/dx and dW/dy of (rho, p, u, v) to 8
// derivative planes, fills their halos and reads them back in the residual; with this kernel
// k_deriv1 writes only dW/dx of u and v (an x march, no y stencil), which the cross derivatives
// read on the rows j +- 1 .. H, and the residual forms the rest itself: dW/dx at the cell from its
// x window, and dW/dy from the primitives of the neighbouring rows (L1 / L2 hits: the rows belong
// to the other waves of the block and to the neighbouring tiles of the same XCD), of u and v on
// the plane entering the x window, of rho and p at the cell.  Each is the sum of deriv1_body over
// the same operands, and on the x-halo planes the window reads, the value the halo exchange of
// the dW/dy planes puts there (centered_diffusion.cpp:22-24): across a periodic x the
// derivative at the wrapped interior plane, beyond a wall zero.  The results are bitwise those
// of k_deriv1 + k_deriv_halo + k_residual.  Offsets within a plane are 32-bit (the host runs
// this kernel only for planes below 2^31 doubles): uniform plane bases plus one offset register
// per access.
// ------------------------------------------------------------------------------------------
#ifndef NS_FUSED_WAVES
#define NS_FUSED_WAVES 2  // waves per SIMD the fused kernel is compiled for (CD2 / CD4)
#endif
// the sixth-order stencils and the second derivatives on unequal dx, dy need more registers than
// three waves per SIMD leave: two there
template <int ORDER, bool SQ, bool SUTH>
constexpr int fused_waves() { return (ORDER >= 6 || !SQ) ? 2 : NS_FUSED_WAVES; }
#ifndef NS_FUSED_MARCH
#define NS_FUSED_MARCH NS_MARCH  // planes i per block of the fused kernel
#endif
constexpr int FUSED_MARCH = NS_FUSED_MARCH;
#ifndef NS_FUSE_DIFF
#define NS_FUSE_DIFF 1  // single blocks and ensembles take the derivative split (2-D viscous terms)
#endif

// z faces formed in the residual kernel (ZC = convection type of the sweep, Shu weights, no
// kernel options): the LDS of a block, sized for every row segment kw in {64, 128, 256}
// (256 / kw rows): the split values of the kw + 5 cells of each row, and the kw + 1 face fluxes
template <int ZC> constexpr int zfd_split_reals() { return ZC ? 4 * (64 + 5) * split_nv<0>() : 1; }
template <int ZC> constexpr int zfd_face_reals() { return ZC ? 4 * 5 * (64 + 1) : 1; }

// The march of a stage (convection and diffusion of a 3-D block, no residual input or output):
// every value a plane needs from memory (face fluxes, the plane entering the x window, the rows
// j +- 1 .. H of the primitives and of dW/dx, the stage inputs) is loaded at the top of the
// plane and held (pin) before any arithmetic, and nothing between the loads and the stores
// branches: a wavefront waits for memory once per plane.  The x window (rho, p, u, v, dW/dy of
// u and v and, for the Sutherland law, mu at the 2H + 1 planes) lives in LDS when `ring` is
// given ([variable][slot][thread], a ring over the planes, each thread its own column: no
// barrier), else in registers; the registers then go to the loads in flight.  The arithmetic
// is that of the general march below, operation for operation.
#ifndef NS_RING
#define NS_RING 1  // the single-block stage march keeps its x window in LDS
#endif
template <int ORDER, bool SUTH> constexpr int ring_vars() { return SUTH ? 7 : 6; }
template <int ORDER, bool SUTH> constexpr int ring_doubles() {
    return ring_vars<ORDER, SUTH>() * (ORDER + 1) * 256;
}
// NX: this stage also forms the next step's prologue (st.Unext, next_prologue).  CP: float
// copies of what the next stage's face sweeps read (the stage output U and c2p(U), or with NX
// the next step's U and stage-1 primitives) go to st.Uf and st.Wf
template <int ORDER, bool SQ, bool SUTH, bool LW, bool NX = false, bool CP = false>
__device__ __forceinline__ void residual_fd_stage_march(
    const flux_t* __restrict__ fx, const flux_t* __restrict__ fy, const flux_t* __restrict__ fz,
    const double* __restrict__ W, const double* __restrict__ Dx, const StageArgs& st,
    const Geo& g, const Phys& ph, unsigned row, int i0, int i1,
    double (&wd)[4][ORDER + 1], real (&dyx)[2][ORDER + 1], double (&fxm)[5], double m1,
    double m2, double m3, double* __restrict__ ring) {
    constexpr int H  = ORDER / 2;
    constexpr int NS = 2 * H + 1;  // planes of the window
    const long    P  = g.P;
    const unsigned sx = (unsigned)g.sx;
    const int      sy = g.sy;
    const int      tid = threadIdx.x + threadIdx.y * 64;
    double dtl = 1E300;  // the next step's CFL minimum over this thread's cells (st.Unext)
    const double* __restrict__ Un = st.Un;
    const double* __restrict__ Uc = st.Ucur;
    double* __restrict__       Uo = st.Uout;
    double* __restrict__       Wo = st.Wout;
    // window variable v (0..3 rho, p, u, v; 4, 5 dW/dy of u, v; 6 mu) at ring slot sl
    auto rs = [&](int v, int sl) -> double& { return ring[(v * NS + sl) * 256 + tid]; };
    // Sutherland: the viscosity along the x window, carried like the primitives
    real muw[NS];
    if constexpr (SUTH) {
#pragma unroll
        for (int t = 0; t < 2 * H; t++) muw[t] = mu_suth((real)wd[0][t], (real)wd[1][t], ph);
    }
    if constexpr (LW) {  // the first 2H planes of the window into slots 0 .. 2H-1
#pragma unroll
        for (int t = 0; t < 2 * H; t++) {
#pragma unroll
            for (int q = 0; q < 4; q++) rs(q, t) = wd[q][t];
            rs(4, t) = dyx[0][t], rs(5, t) = dyx[1][t];
            if constexpr (SUTH) rs(6, t) = muw[t];
        }
    }
    int sl0 = 0;  // ring slot of window offset 0
    for (int i = i0; i < i1; i++) {
        const unsigned cc = row + (unsigned)(i + NGC) * sx;
        const unsigned cn = cc + (unsigned)H * sx;
        // plane e = i + H enters the window; dW/dy of u, v there (across a periodic x the
        // wrapped plane's; the caller keeps this march off planes beyond an x wall)
        const int      e  = i + H;
        const unsigned eb = row + (unsigned)((e >= g.nx ? e - g.nx : e) + NGC) * sx;
        // ---- loads ----
        double fxc[5], fyc[5], fym[5], fzc[5], fzm[5], unv[5], ucv[5], wn[4];
        double uy[2][NS], dxn[2][NS], ry[NS], py[NS];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            fxc[q] = fx[q * P + cc];
            fyc[q] = fy[q * P + cc];
            fym[q] = fy[q * P + cc - sy];
            fzc[q] = fz[q * P + cc];
            fzm[q] = fz[q * P + cc - 1];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) wn[q] = W[q * P + cn];
#pragma unroll
        for (int t = 0; t < NS; t++) {
            if (t == H) continue;
            const int o = (t - H) * sy;
            uy[0][t]    = W[2 * P + eb + o];
            uy[1][t]    = W[3 * P + eb + o];
            dxn[0][t]   = Dx[2 * P + cc + o];
            dxn[1][t]   = Dx[3 * P + cc + o];
            ry[t]       = W[cc + o];
            py[t]       = W[P + cc + o];
        }
#pragma unroll
        for (int q = 0; q < 5; q++) {
            ucv[q] = Uc[q * P + cc];
            unv[q] = Un[q * P + cc];
        }
        if constexpr (LW) {
#pragma unroll
            for (int q = 0; q < 5; q++) {
                pin(fxc[q]), pin(fyc[q]), pin(fym[q]), pin(fzc[q]), pin(fzm[q]);
                pin(ucv[q]), pin(unv[q]);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) pin(wn[q]);
#pragma unroll
            for (int t = 0; t < NS; t++) {
                if (t == H) continue;
                pin(uy[0][t]), pin(uy[1][t]), pin(dxn[0][t]), pin(dxn[1][t]), pin(ry[t]),
                    pin(py[t]);
            }
        }
        // ---- convection: x, then y, then z flux differences (weno-5.cpp:60-66) ----
        double R[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const double fc = fxc[q];
#if !NS_FAST
            double a = dv(fc - fxm[q], Rcp{g.dx, g.idx});
            double b = dv(fyc[q] - fym[q], Rcp{g.dy, g.idy});
#else
            double a = (fc - fxm[q]) * g.idx;
            double b = (fyc[q] - fym[q]) * g.idy;
#endif
            fxm[q] = fc;
            R[q]   = a;
            R[q] += b;
            double zr = fzc[q], zl = fzm[q];
#if !NS_FAST
            R[q] += dv(zr - zl, Rcp{g.dz, g.idz});
#else
            R[q] += (zr - zl) * g.idz;
#endif
        }
        // ---- diffusion: the window at this plane ----
        const real dyu = (real)fd_sum_t<ORDER, false, double>([&](int t) { return uy[0][t]; },
                                                              g.dy, g.idy);
        const real dyv = (real)fd_sum_t<ORDER, false, double>([&](int t) { return uy[1][t]; },
                                                              g.dy, g.idy);
        double     wl[4][NS];
        real       wx[4][NS], dy2[2][NS], mw[NS];
        if constexpr (LW) {
            const int sn = (sl0 + 2 * H < NS) ? sl0 + 2 * H : sl0 + 2 * H - NS;
#pragma unroll
            for (int q = 0; q < 4; q++) rs(q, sn) = wn[q];
            rs(4, sn) = dyu, rs(5, sn) = dyv;
            if constexpr (SUTH) rs(6, sn) = mu_suth((real)wn[0], (real)wn[1], ph);
#pragma unroll
            for (int t = 0; t < NS; t++) {
                const int sl = (sl0 + t < NS) ? sl0 + t : sl0 + t - NS;
#pragma unroll
                for (int q = 0; q < 4; q++) wl[q][t] = rs(q, sl), wx[q][t] = (real)wl[q][t];
                dy2[0][t] = (real)rs(4, sl), dy2[1][t] = (real)rs(5, sl);
                if constexpr (SUTH) mw[t] = (real)rs(6, sl);
            }
            sl0 = (sl0 + 1 < NS) ? sl0 + 1 : 0;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) wd[q][2 * H] = wn[q];
            dyx[0][2 * H] = dyu, dyx[1][2 * H] = dyv;
#pragma unroll
            for (int t = 0; t < NS; t++) {
#pragma unroll
                for (int q = 0; q < 4; q++) wl[q][t] = wd[q][t], wx[q][t] = (real)wd[q][t];
                dy2[0][t] = dyx[0][t], dy2[1][t] = dyx[1][t];
            }
            if constexpr (SUTH) {
                muw[2 * H] = mu_suth(wx[0][2 * H], wx[1][2 * H], ph);
#pragma unroll
                for (int t = 0; t < NS; t++) mw[t] = muw[t];
            }
        }
        double dxo[4];
#pragma unroll
        for (int q = 0; q < 4; q++)
            dxo[q] = fd_sum_t<ORDER, false, double>([&](int t) { return wl[q][t]; }, g.dx, g.idx);
        double dxy[2][NS];
#pragma unroll
        for (int t = 0; t < NS; t++) {
            if (t == H) continue;
            dxy[0][t] = dxn[0][t];
            dxy[1][t] = dxn[1][t];
        }
        dxy[0][H]        = dxo[2];
        dxy[1][H]        = dxo[3];
        const double dxr = dxo[0], dxp = dxo[1];
        double dyr = fd_sum_t<ORDER, false, double>([&](int t) { return ry[t]; }, g.dy, g.idy);
        double dyp = fd_sum_t<ORDER, false, double>([&](int t) { return py[t]; }, g.dy, g.idy);
        real   dmu_dy = 0;
        if constexpr (SUTH)
            dmu_dy = dmu_dy_of<ORDER>(
                [&](int t, double& r, double& q) {
                    r = ry[t];
                    q = py[t];
                },
                g, ph);
        real ene, xm, ym;
        viscous_v<ORDER, SQ, SUTH>(wx, dy2, dxy, dxr, dxp, dyr, dyp, dmu_dy, g, ph, ene, xm, ym,
                                   SUTH ? mw : nullptr);
        R[1] = R[1] + ene;
        R[2] = R[2] + xm;
        R[3] = R[3] + ym;
        if constexpr (!LW) {
#pragma unroll
            for (int t = 0; t < 2 * H; t++) {
#pragma unroll
                for (int q = 0; q < 4; q++) wd[q][t] = wd[q][t + 1];
                dyx[0][t] = dyx[0][t + 1];
                dyx[1][t] = dyx[1][t + 1];
                if constexpr (SUTH) muw[t] = muw[t + 1];
            }
        }
        // ---- SSP-RK3 stage (time_scheme.cpp:72-96), primitives of the result ----
        double un[5], w[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            double r = R[q];
#if !NS_FAST
            un[q] = (unv[q] * m1) + (ucv[q] * m2) + (r * m3);
#else
            un[q] = unv[q] * m1 + ucv[q] * m2 + r * m3;
#endif
            if constexpr (!NX) Uo[q * P + cc] = un[q];
            if constexpr (!NX && CP) st.Uf[q * P + cc] = (float)un[q];
        }
        c2p(un, ph.gamma, w);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            Wo[q * P + cc] = w[q];
            if constexpr (!NX && CP) st.Wf[q * P + cc] = (float)w[q];
        }
        if (st.flags) stage_health(st.flags, un, w);
        if constexpr (NX) next_prologue<CP>(st, w, cc, P, g, ph.gamma, dtl);
    }
    if constexpr (NX) wave_min_atomic(dtl, st.dtnext);
}

// 256 threads: kw lanes along k times 256/kw rows j; (bx, by, bz) = (k tile, j tile, chunk of
// `march` planes i).  ZC > 0: the z face fluxes of the plane are formed here (convection type
// ZC, Shu weights) instead of read from fz: the lanes of a row stage the split values of their
// cells, and of the two cells before and three after the row segment, in LDS (zt); lane k forms
// face k + 1/2 into zf, and takes face k - 1/2 from lane k - 1, the first lane of the segment
// from fz, where the plane launch of that face stored it.  The same face_flux on the same split
// values as the z sweep: the residual is bitwise the one the sweep's faces give.  Every thread
// of the block reaches the two barriers of each plane; threads without a cell skip the work.
template <int ORDER, bool SQ, bool SUTH, int ZC = 0, bool LW = false, bool NX = false,
          bool CP = false>
__device__ __forceinline__ void residual_fd_body(
    const flux_t* __restrict__ fx, const flux_t* __restrict__ fy, const flux_t* __restrict__ fz,
    const double* __restrict__ W, const double* __restrict__ Dx, const double* __restrict__ Rin,
    double* __restrict__ Rout, const StageArgs& st, const unsigned long long* __restrict__ dtmin,
    int which, const Geo& g, const Phys& ph, const Bnd& bnd, int bx, int by, int bz, int kw,
    int march, real* __restrict__ zt = nullptr, real* __restrict__ zf = nullptr,
    double* __restrict__ ring = nullptr) {
    constexpr int H   = ORDER / 2;
    const int     tid = threadIdx.x + threadIdx.y * 64;
    const int     kr  = tid & (kw - 1);  // lane within the row segment
    const int     k   = kr + bx * kw;
    const int     j   = tid / kw + by * (256 / kw);
    const int     i0  = bz * march;
    const bool    live = k < g.nz && j < g.ny;
    if constexpr (ZC == 0) {
        if (!live) return;
    }
    const int       i1   = (i0 + march < g.nx) ? i0 + march : g.nx;
    const long      P    = g.P;
    const unsigned  sx   = (unsigned)g.sx, sy = (unsigned)g.sy;
    const unsigned  row  = (unsigned)(j + NGC) * sy + (unsigned)(k + NGC);
    const bool      xper = bnd.periodic[0] != 0;
    const bool      conv = (which & 1) && ph.conv;
    const double*   Wr = W;
    const double*   Wp = W + P;
    const double*   Wu = W + 2 * P;
    const double*   Wv = W + 3 * P;
    double wd[4][2 * H + 1];   // primitives of row j at x offsets -H .. H
    real   dyx[2][2 * H + 1];  // dW/dy of u, v at x offsets -H .. H
    double fxm[5];
    // dW/dy of u (a = 0), v (a = 1) on plane e: zero beyond an x wall, the wrapped plane's
    // across a periodic x
    auto dy_uv = [&](int e, int a) -> real {
        if (!xper && (e < 0 || e >= g.nx)) return real(0.0);
        const int      ep = (e < 0) ? e + g.nx : (e >= g.nx ? e - g.nx : e);
        const unsigned b  = row + (unsigned)(ep + NGC) * sx;
        const double*  f  = a ? Wv : Wu;
        return (real)fd_sum_t<ORDER, false, double>(
            [&](int t) { return f[b + (t - H) * (int)sy]; }, g.dy, g.idy);
    };
    if (live) {
#pragma unroll
    for (int t = 0; t < 2 * H; t++) {
        const int      e = i0 - H + t;
        const unsigned a = row + (unsigned)(e + NGC) * sx;
#pragma unroll
        for (int q = 0; q < 4; q++) wd[q][t] = (W + q * P)[a];
    }
#pragma unroll
    for (int t = 0; t < 2 * H; t++) {
        dyx[0][t] = dy_uv(i0 - H + t, 0);
        dyx[1][t] = dy_uv(i0 - H + t, 1);
    }
    if (conv) {
#pragma unroll
        for (int q = 0; q < 5; q++) fxm[q] = (fx + q * P)[row + (unsigned)(i0 + NGC - 1) * sx];
    }
    }  // live
    // z faces (ZC): the row's split tile and face row in LDS, the cells valid in the segment
    constexpr int NVZ  = split_nv<0>();
    const int     ncz  = kw + 5;
    real*         T    = zt + (tid / kw) * (NVZ * ncz);
    real*         Fz   = zf + (tid / kw) * (5 * (kw + 1));
    const bool    rowz = j < g.ny;
    const int     mz   = (g.nz - bx * kw < kw) ? g.nz - bx * kw : kw;
    double dt = 0, m1 = 0, m2 = 0, m3 = 0;
    double dtl = 1E300;  // the next step's CFL minimum over this thread's cells (st.Unext)
    if (st.Uout) {
        dt = __longlong_as_double((long long)*dtmin) * ph.cfl;
        if (st.stage == 0) {
            m1 = 1.0, m2 = 0.0, m3 = -dt;
        } else if (st.stage == 1) {
            m1 = 3.0 / 4.0, m2 = 1.0 / 4.0, m3 = -(1. / 4.) * dt;
        } else {
            m1 = 1. / 3., m2 = 2. / 3., m3 = -(2. / 3.) * dt;
        }
    }
    if (ZC == 0 && st.Uout && !Rin && !Rout && conv && g.dim == 3 && (xper || i1 + H <= g.nx)) {
        residual_fd_stage_march<ORDER, SQ, SUTH, LW, NX, CP>(fx, fy, fz, W, Dx, st, g, ph, row, i0, i1,
                                                     wd, dyx, fxm, m1, m2, m3, ring);
        return;
    }
    for (int i = i0; i < i1; i++) {
        const unsigned cc = row + (unsigned)(i + NGC) * sx;
        real           zp[5], zm[5];  // z faces k + 1/2 and k - 1/2 (ZC)
        if constexpr (ZC > 0) {
            const unsigned rb  = (unsigned)(j + NGC) * sy + (unsigned)(i + NGC) * sx;
            const real     gam = (real)ph.gamma;
            if (live) {
                real sv[NVZ];
                stage_cell<2, 0>(st.Ucur, W, (long)cc, P, gam, sv);
#pragma unroll
                for (int v = 0; v < NVZ; v++) T[v * ncz + kr + 2] = sv[v];
            }
            if (rowz && kr < 5) {  // cells 2 and 1 before the segment, 1 .. 3 after its end
                const int t = kr < 2 ? kr : mz + kr;
                real      sv[NVZ];
                stage_cell<2, 0>(st.Ucur, W, (long)(rb + (unsigned)(bx * kw + NGC - 2 + t)), P,
                                 gam, sv);
#pragma unroll
                for (int v = 0; v < NVZ; v++) T[v * ncz + t] = sv[v];
            }
            __syncthreads();
            if (live) {
                real Wl[5], Wr[5];
                if constexpr (ZC == 2) {
#pragma unroll
                    for (int q = 0; q < 5; q++) Wl[q] = (W + q * P)[cc], Wr[q] = (W + q * P)[cc + 1];
                }
                struct {
                    const real* b;
                    int         nc;
                    __device__ real operator()(int ii, int v) const { return b[v * nc + ii]; }
                } sz{&T[kr], ncz};
                face_flux<2, ZC, 0, 0>(sz, Wl, Wr, gam, zp, Learn{nullptr, false});
#pragma unroll
                for (int q = 0; q < 5; q++) Fz[q * (kw + 1) + kr + 1] = zp[q];
                if (kr == 0) {
#pragma unroll
                    for (int q = 0; q < 5; q++) Fz[q * (kw + 1)] = (real)(fz + q * P)[cc - 1];
                }
            }
            __syncthreads();
            if (!live) continue;
#pragma unroll
            for (int q = 0; q < 5; q++) zm[q] = Fz[q * (kw + 1) + kr];
        }
        double         R[5];
        bool           have[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            have[q] = (Rin != nullptr);
            R[q]    = Rin ? (Rin + q * P)[cc] : 0.0;
        }
        if (conv) {
#pragma unroll
            for (int q = 0; q < 5; q++) {
                double fc = (fx + q * P)[cc];
#if !NS_FAST
                double a = dv(fc - fxm[q], Rcp{g.dx, g.idx});
                double b = dv((double)(fy + q * P)[cc] - (double)(fy + q * P)[cc - sy], Rcp{g.dy, g.idy});
#else
                double a = (fc - fxm[q]) * g.idx;
                double b = ((double)(fy + q * P)[cc] - (double)(fy + q * P)[cc - sy]) * g.idy;
#endif
                fxm[q] = fc;
                R[q]   = have[q] ? R[q] + a : a;
                R[q] += b;
                if (g.dim == 3) {
                    double zr, zl;
                    if constexpr (ZC > 0) {
                        zr = (double)zp[q], zl = (double)zm[q];
                    } else {
                        zr = (fz + q * P)[cc], zl = (fz + q * P)[cc - 1];
                    }
#if !NS_FAST
                    R[q] += dv(zr - zl, Rcp{g.dz, g.idz});
#else
                    R[q] += (zr - zl) * g.idz;
#endif
                }
                have[q] = true;
            }
        }
        {
            // the plane entering the windows
            const unsigned cn = cc + (unsigned)H * sx;
#pragma unroll
            for (int q = 0; q < 4; q++) wd[q][2 * H] = (W + q * P)[cn];
            dyx[0][2 * H] = dy_uv(i + H, 0);
            dyx[1][2 * H] = dy_uv(i + H, 1);
            // dW/dx: at the cell from the window, of u and v on the rows j +- 1 .. H from the
            // derivative planes (their halos filled like the separate path's)
            double dxo[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                dxo[q] = fd_sum_t<ORDER, false, double>([&](int t) { return wd[q][t]; }, g.dx,
                                                        g.idx);
            double dxy[2][2 * H + 1];
#pragma unroll
            for (int t = 0; t <= 2 * H; t++) {
                if (t == H) continue;
                dxy[0][t] = (Dx + 2 * P)[cc + (t - H) * (int)sy];
                dxy[1][t] = (Dx + 3 * P)[cc + (t - H) * (int)sy];
            }
            dxy[0][H]        = dxo[2];
            dxy[1][H]        = dxo[3];
            const double dxr = dxo[0], dxp = dxo[1];
            // dW/dy of rho and p at the cell, and of the viscosity
            double dyr = fd_sum_t<ORDER, false, double>(
                [&](int t) { return Wr[cc + (t - H) * (int)sy]; }, g.dy, g.idy);
            double dyp = fd_sum_t<ORDER, false, double>(
                [&](int t) { return Wp[cc + (t - H) * (int)sy]; }, g.dy, g.idy);
            real dmu_dy = 0;
            if constexpr (SUTH)
                dmu_dy = dmu_dy_of<ORDER>(
                    [&](int t, double& r, double& q) {
                        r = Wr[cc + (t - H) * (int)sy];
                        q = Wp[cc + (t - H) * (int)sy];
                    },
                    g, ph);
            real wx[4][2 * H + 1];
#pragma unroll
            for (int t = 0; t <= 2 * H; t++)
#pragma unroll
                for (int q = 0; q < 4; q++) wx[q][t] = (real)wd[q][t];
            real ene, xm, ym;
            viscous_v<ORDER, SQ, SUTH>(wx, dyx, dxy, dxr, dxp, dyr, dyp, dmu_dy, g, ph, ene, xm,
                                       ym);
            R[1]    = have[1] ? R[1] + ene : ene;
            R[2]    = have[2] ? R[2] + xm : xm;
            R[3]    = have[3] ? R[3] + ym : ym;
            have[1] = have[2] = have[3] = true;
#pragma unroll
            for (int t = 0; t < 2 * H; t++) {
#pragma unroll
                for (int q = 0; q < 4; q++) wd[q][t] = wd[q][t + 1];
                dyx[0][t] = dyx[0][t + 1];
                dyx[1][t] = dyx[1][t + 1];
            }
        }
        if (st.Uout == nullptr) {
#pragma unroll
            for (int q = 0; q < 5; q++) (Rout + q * P)[cc] = have[q] ? R[q] : 0.0;
            continue;
        }
        double un[5], w[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            double r = have[q] ? R[q] : 0.0;
#if !NS_FAST
            un[q] = ((st.Un + q * P)[cc] * m1) + ((st.Ucur + q * P)[cc] * m2) + (r * m3);
#else
            un[q] = (st.Un + q * P)[cc] * m1 + (st.Ucur + q * P)[cc] * m2 + r * m3;
#endif
            if constexpr (!NX) (st.Uout + q * P)[cc] = un[q];
            if constexpr (!NX && CP) (st.Uf + q * P)[cc] = (float)un[q];
        }
        c2p(un, ph.gamma, w);
#pragma unroll
        for (int q = 0; q < 5; q++) {
            (st.Wout + q * P)[cc] = w[q];
            if constexpr (!NX && CP) (st.Wf + q * P)[cc] = (float)w[q];
        }
        if (st.flags) stage_health(st.flags, un, w);
        if constexpr (NX) next_prologue<CP>(st, w, cc, P, g, ph.gamma, dtl);
    }
    if constexpr (NX) wave_min_atomic(dtl, st.dtnext);
}

template <int ORDER, bool SQ, bool SUTH, bool NX = false>
__global__ __launch_bounds__(256, (fused_waves<ORDER, SQ, SUTH>())) void k_residual_fd(
    const flux_t* __restrict__ fx, const flux_t* __restrict__ fy, const flux_t* __restrict__ fz,
    const double* __restrict__ W, const double* __restrict__ Dx, const double* __restrict__ Rin,
    double* __restrict__ Rout, StageArgs st, const unsigned long long* __restrict__ dtmin,
    int which, Geo g, Phys ph, Bnd bnd, int z0) {
    // the x window of the stage march (one thread's column per slot: no barrier)
    __shared__ double ring[NS_RING ? ring_doubles<ORDER, SUTH>() : 1];
    int bx, by, bz;
    cell_tile(bx, by, bz);
    constexpr bool CP = NS_FP32 && NS_STATE_COPY;
    residual_fd_body<ORDER, SQ, SUTH, 0, NS_RING != 0, NX, CP>(fx, fy, fz, W, Dx, Rin, Rout, st,
                                                             dtmin, which, g, ph, bnd, bx, by,
                                                             z0 + bz, KW_CELL, FUSED_MARCH,
                                                             nullptr, nullptr, ring);
}

// the residual + stage update with the z faces formed in the kernel (convection type ZC, Shu
// weights, no kernel options; the host launches it for CD4 only).  kw in {64, 128, 256}: the
// row segment a block's 256 threads cover (256 / kw rows); fz holds the first face of every
// segment (face plane k = bx kw - 1/2, written by the plane launch before this kernel)
#ifndef NS_ZFD_WAVES
#define NS_ZFD_WAVES 2  // waves per SIMD the kernel is compiled for
#endif
template <int ORDER, bool SQ, bool SUTH, int ZC>
__global__ __launch_bounds__(256, NS_ZFD_WAVES) void k_residual_zfd(
    const flux_t* __restrict__ fx, const flux_t* __restrict__ fy, const flux_t* __restrict__ fz,
    const double* __restrict__ W, const double* __restrict__ Dx, StageArgs st,
    const unsigned long long* __restrict__ dtmin, Geo g, Phys ph, Bnd bnd, int z0, int kw) {
    __shared__ real zt[zfd_split_reals<ZC>()];
    __shared__ real zf[zfd_face_reals<ZC>()];
    int bx, by, bz;
    cell_tile(bx, by, bz);
    residual_fd_body<ORDER, SQ, SUTH, ZC>(fx, fy, fz, W, Dx, nullptr, nullptr, st, dtmin,
                                          3 /* convection and diffusion */, g, ph, bnd, bx,
                                          by, z0 + bz, kw,
                                          FUSED_MARCH, zt, zf);
}

// ------------------------------------------------------------------------------------------
// full 3-D viscous terms (physics VISCOUS, extension): CenteredDiffusion::compute_residual
// (centered_diffusion.cpp:70-182) for all three axes, each second derivative on its own axis,
// the viscous work tau_ij du_i/dx_j and the heat flux from T = p / rho.  With g[a][d] =
// du_a/dx_d, div = tr g, lamb = -2/3 mu:
//   divtau[a] = sum_d dmu_d (g[a][d] + g[d][a]) + mu (lap u_a + d_a div) + dlamb_a div
//               + lamb d_a div
//   energy    = -(sum_a u_a divtau[a]) - sum_{a,d} tau[a][d] g[a][d]
//               - heat sum_d (dmu_d dT/dx_d + mu d2T/dx_d2)
// Three passes: first derivatives of (u, v, w, T) along x, y, z; their periodic halos across
// the other two axes (zero beyond a wall, like centered_diffusion.cpp:22-24); the terms per cell
// (mixed derivatives d_a d_d u_d = df_da(du_d/dx_d) from the derivative planes).
// ------------------------------------------------------------------------------------------

// u, v, w or T = p / rho (f = 0..3) of the cell at offset o
__device__ __forceinline__ real visc_field(const double* __restrict__ W, long P, long o, int f) {
    return (f < 3) ? (real)W[(2 + f) * P + o] : dv((real)W[P + o], (real)W[o]);
}

// one thread per interior cell: block (64, 4) over (k, j), blockIdx.z = i
template <int ORDER>
__global__ __launch_bounds__(256) void k_deriv3d(const double* __restrict__ W,
                                                 double* __restrict__ Dx, double* __restrict__ Dy,
                                                 double* __restrict__ Dz, Geo g) {
    constexpr int H = ORDER / 2;
    const int     k = threadIdx.x + blockIdx.y * 64, j = threadIdx.y + blockIdx.z * 4;
    const int     i = blockIdx.x;  // consecutive workgroups step along i: x stencils meet in L2
    if (k >= g.nz || j >= g.ny) return;
    const long P = g.P;
    const long c = (long)(i + NGC) * g.sx + (long)(j + NGC) * g.sy + (k + NGC);
#pragma unroll
    for (int f = 0; f < 4; f++) {
        Dx[c + f * P] = fd_sum<ORDER, false>(
            [&](int t) { return visc_field(W, P, c + (long)(t - H) * g.sx, f); }, g.dx, g.idx);
        Dy[c + f * P] = fd_sum<ORDER, false>(
            [&](int t) { return visc_field(W, P, c + (long)(t - H) * g.sy, f); }, g.dy, g.idy);
        Dz[c + f * P] = fd_sum<ORDER, false>(
            [&](int t) { return visc_field(W, P, c + (t - H), f); }, g.dz, g.idz);
    }
}

// periodic halos of the derivative planes the mixed derivatives read: du/dx across y and z,
// dv/dy across x and z, dw/dz across x and y.  blockIdx.z = (combination, low/high, layer);
// block (64, 4) over the face (fast, slow tangential index)
__global__ void k_deriv3d_halo(double* Dx, double* Dy, double* Dz, Geo g, Bnd b) {
    const int combo = blockIdx.z / (2 * NGC);
    const int h     = blockIdx.z % NGC;
    const bool low  = ((blockIdx.z / NGC) & 1) == 0;
    int d, a;  // derivative plane set (and its field), exchange axis
    switch (combo) {
    case 0: d = 0, a = 1; break;
    case 1: d = 0, a = 2; break;
    case 2: d = 1, a = 0; break;
    case 3: d = 1, a = 2; break;
    case 4: d = 2, a = 0; break;
    default: d = 2, a = 1; break;
    }
    if (!b.periodic[a]) return;
    const int  n[3]  = {g.nx, g.ny, g.nz};
    const long st[3] = {g.sx, g.sy, 1};
    const int  ta = (a == 0) ? 1 : 0, tb = (a == 2) ? 1 : 2;  // slow, fast tangential axes
    const int  f = threadIdx.x + blockIdx.x * 64, sl = threadIdx.y + blockIdx.y * 4;
    if (sl >= n[ta] || f >= n[tb]) return;
    double*    D    = (d == 0) ? Dx : (d == 1) ? Dy : Dz;
    const long base = (long)d * g.P + (long)(sl + NGC) * st[ta] + (long)(f + NGC) * st[tb];
    const int  dst  = low ? h : NGC + n[a] + h;
    const int  src  = low ? n[a] + h : NGC + h;
    D[base + (long)dst * st[a]] = D[base + (long)src * st[a]];
}

template <int ORDER, bool SUTH>
__global__ __launch_bounds__(256) void k_visc3d(const double* __restrict__ W,
                                                const double* __restrict__ Dx,
                                                const double* __restrict__ Dy,
                                                const double* __restrict__ Dz,
                                                double* __restrict__ Rv, Geo g, Phys ph) {
    constexpr int H = ORDER / 2;
    const int     k = threadIdx.x + blockIdx.y * 64, j = threadIdx.y + blockIdx.z * 4;
    const int     i = blockIdx.x;  // consecutive workgroups step along i: x stencils meet in L2
    if (k >= g.nz || j >= g.ny) return;
    const long    P     = g.P;
    const long    c     = (long)(i + NGC) * g.sx + (long)(j + NGC) * g.sy + (k + NGC);
    const long    st[3] = {g.sx, g.sy, 1};
    const double* D[3]  = {Dx, Dy, Dz};
    const double  dd[3] = {g.dx, g.dy, g.dz}, idd[3] = {g.idx, g.idy, g.idz};
    const double  d2[3] = {g.dxx, g.dyy, g.dz * g.dz};
    const double  id2[3] = {g.idxx, g.idyy, g.idzz};
    real gr[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) gr[a][d] = (real)D[d][c + a * P];
    const real div = (gr[0][0] + gr[1][1]) + gr[2][2];
    real S[3][4];  // d2 f / dx_d2, f = u, v, w, T
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int f = 0; f < 4; f++)
            S[d][f] = fd_sum<ORDER, true>(
                [&](int t) { return visc_field(W, P, c + (long)(t - H) * st[d], f); }, d2[d],
                id2[d]);
    real lap[3], ddiv[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        lap[a] = (S[0][a] + S[1][a]) + S[2][a];
        real x[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (d == a) {
                x[d] = S[a][a];
            } else {
                const double* Dd = D[d] + (long)d * P + c;
                x[d] = fd_sum<ORDER, false>([&](int t) { return (real)Dd[(long)(t - H) * st[a]]; },
                                            dd[a], idd[a]);
            }
        }
        ddiv[a] = (x[0] + x[1]) + x[2];
    }
    real mu, dmu[3] = {RL(0.), RL(0.), RL(0.)};
    if constexpr (SUTH) {
        mu = mu_suth(W[c], W[P + c], ph);
#pragma unroll
        for (int d = 0; d < 3; d++)
            dmu[d] = fd_sum<ORDER, false>(
                [&](int t) {
                    const long o = c + (long)(t - H) * st[d];
                    return mu_suth(W[o], W[P + o], ph);
                },
                dd[d], idd[d]);
    } else {
        mu = (real)ph.mu;
    }
    const real lamb = (-RL(2.) / RL(3.)) * mu;
    real       divtau[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        real t = mu * (lap[a] + ddiv[a]);
        if constexpr (SUTH) {
            real m = dmu[0] * (gr[a][0] + gr[0][a]);
            m += dmu[1] * (gr[a][1] + gr[1][a]);
            m += dmu[2] * (gr[a][2] + gr[2][a]);
            t = (m + t) + (-(RL(2.) / RL(3.)) * dmu[a]) * div;
        }
        divtau[a] = t + lamb * ddiv[a];
    }
    const real u = (real)W[2 * P + c], v = (real)W[3 * P + c], w = (real)W[4 * P + c];
    const real work = (u * divtau[0] + v * divtau[1]) + w * divtau[2];
    real       phi  = RL(0.);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            real tau = mu * (gr[a][d] + gr[d][a]);
            if (a == d) tau = tau + lamb * div;
            phi = (a == 0 && d == 0) ? tau * gr[a][d] : phi + tau * gr[a][d];
        }
    real q[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        real t = mu * S[d][3];
        if constexpr (SUTH) t = dmu[d] * (real)D[d][c + 3 * P] + t;
        q[d] = -(real)ph.heat * t;
    }
    const real ene = ((-work) - phi) + ((q[0] + q[1]) + q[2]);
    Rv[c]         = ene;
    Rv[P + c]     = -divtau[0];
    Rv[2 * P + c] = -divtau[1];
    Rv[3 * P + c] = -divtau[2];
}

__global__ void k_time_accum(double* t, const unsigned long long* dtmin, double cfl) {
    *t = *t + __longlong_as_double((long long)*dtmin) * cfl;  // time = time + dt (driver.cpp:57)
}

// ------------------------------------------------------------------------------------------
// host layout conversion: reference AoS [NX][NY][NZ][5] <-> device SoA planes
// ------------------------------------------------------------------------------------------
// cell c of the dense host order (i, j, k) -> its offset in the padded device planes
__device__ __forceinline__ long dev_offset(long c, const Geo& g) {
    const int  k = (int)(c % g.NZ);
    const long t = c / g.NZ;
    const int  j = (int)(t % g.NY);
    const int  i = (int)(t / g.NY);
    return (long)i * g.sx + (long)j * g.sy + k;
}
__global__ void k_aos_to_soa(const double* __restrict__ aos, double* __restrict__ soa, Geo g) {
    long c = threadIdx.x + (long)blockIdx.x * blockDim.x;
    if (c >= (long)g.NX * g.NY * g.NZ) return;
    const long d = dev_offset(c, g);
#pragma unroll
    for (int q = 0; q < 5; q++) soa[d + q * g.P] = aos[c * 5 + q];
}
__global__ void k_soa_to_aos(const double* __restrict__ soa, double* __restrict__ aos, Geo g) {
    long c = threadIdx.x + (long)blockIdx.x * blockDim.x;
    if (c >= (long)g.NX * g.NY * g.NZ) return;
    const long d = dev_offset(c, g);
#pragma unroll
    for (int q = 0; q < 5; q++) aos[c * 5 + q] = soa[d + q * g.P];
}

// ------------------------------------------------------------------------------------------
// monitoring reductions: conservation sums (fixed-order two-pass tree, deterministic), CFL
// minimum and health flags
// ------------------------------------------------------------------------------------------
constexpr int RED_THREADS = 256;

// sums of five doubles over a 256-thread workgroup: butterfly over the 64 lanes of each
// wavefront (__shfl_xor, DPP / ds_swizzle under the hood), then the 4 wave sums in a fixed order;
// the result is in every lane of wave 0
__device__ __forceinline__ void block_sum5(double acc[5]) {
    __shared__ double ws[5][RED_THREADS / 64];
#pragma unroll
    for (int q = 0; q < 5; q++)
        for (int o = 32; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) ws[q][wid] = acc[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 5; q++) acc[q] = (ws[q][0] + ws[q][1]) + (ws[q][2] + ws[q][3]);
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce(const double* __restrict__ U,
                                                        const double* __restrict__ W,
                                                        double* __restrict__ partial,
                                                        unsigned long long* dtmin,
                                                        unsigned* flags, Geo g, Phys ph) {
    long     n      = (long)g.nx * g.ny * g.nz;
    double   acc[5] = {0, 0, 0, 0, 0};
    double   dtl    = 1E300;
    unsigned fl     = 0;
    for (long t = threadIdx.x + (long)blockIdx.x * RED_THREADS; t < n;
         t += (long)gridDim.x * RED_THREADS) {
        int    k = (int)(t % g.nz);
        long   r = t / g.nz;
        int    j = (int)(r % g.ny);
        int    i = (int)(r / g.ny);
        long   c = (long)(i + NGC) * g.sx + (long)(j + NGC) * g.sy + (k + NGC);
        double w[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            double u = U[c + q * g.P];
            acc[q] += u;
            w[q] = W[c + q * g.P];
            if (!isfinite(u) || !isfinite(w[q])) fl |= 1u;
        }
        if (!(w[0] > 0.0) || !(w[1] > 0.0)) fl |= 2u;
        else {
            double cs = sqrt(ph.gamma * w[1] / w[0]);
            double m  = fmax(fabs(w[2]), fmax(fabs(w[3]), fabs(w[4]))) + cs;
            if (ph.alpha_mode == 0 && m > LF_ALPHA) fl |= 4u;
            double d = local_dt(w, ph.gamma, g.dx, g.dy, g.dz);
            dtl      = (d < dtl) ? d : dtl;
        }
    }
    block_sum5(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) partial[blockIdx.x * 5 + q] = acc[q];
    if (fl) atomicOr(flags, fl);
    block_min_atomic(dtl, dtmin);
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double* __restrict__ partial,
                                                              int nb, double* __restrict__ sums) {
    double acc[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += RED_THREADS)
        for (int q = 0; q < 5; q++) acc[q] += partial[b * 5 + q];
    block_sum5(acc);
    if (threadIdx.x == 0)
#pragma unroll
        for (int q = 0; q < 5; q++) sums[q] = acc[q];
}

// ------------------------------------------------------------------------------------------
// self-test of dv() against the IEEE division operator on random operands: sign, exponent in
// [-300, 300] and mantissa drawn from a counter-based generator (splitmix64 of seed + index)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double random_double(unsigned long long r) {
    unsigned long long mant = r & ((1ull << 52) - 1);
    unsigned long long e    = 1023ull - 300ull + ((r >> 52) % 601ull);
    unsigned long long sign = (r >> 63) & 1ull;
    return __longlong_as_double((long long)((sign << 63) | (e << 52) | mant));
}

__global__ void k_selftest_div(long n, unsigned long long seed, unsigned long long* count) {
    unsigned long long bad = 0;
    for (long t = threadIdx.x + (long)blockIdx.x * blockDim.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        double a = random_double(splitmix64(seed + 2 * (unsigned long long)t));
        double b = random_double(splitmix64(seed + 2 * (unsigned long long)t + 1));
        if (t % 97 == 0) a = 0.0;                     // exact-zero numerators
        if (t % 89 == 0) b = (double)((t % 13) + 1);  // small integers (grid-like divisors)
        double x = dv(a, b), y = a / b;
        if (!(x == y)) bad++;
    }
    if (bad) atomicAdd(count, bad);
}

// ------------------------------------------------------------------------------------------
// ensemble (batched) kernels: one launch per family covers every block of an ensemble; each
// workgroup finds its block by a binary search over the family's workgroup prefix sums and runs
// the same body as the single-block kernel on its tile.  Stage buffers follow the single-block
// step: stage s reads U_s = (Un, Ua, Ub)[s] with primitives (Ws0, Ws1, Ws0)[s] and writes
// (Ua, Ub, Ua)[s] with primitives (Ws1, Ws0, W)[s].
// ------------------------------------------------------------------------------------------
// returns the block of this workgroup.  The kernels copy the descriptor fields they use into
// locals before any store, so the compiler keeps them in scalar registers.
__device__ __forceinline__ int ens_locate(const EnsTable& T, int fam, int& bx, int& by, int& bz) {
    const int* st = T.start + fam * (T.n + 1);
    const int  w  = blockIdx.x;
    int        lo = 0, hi = T.n;  // st[0] = 0 <= w < st[n]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (st[mid] <= w) lo = mid;
        else hi = mid;
    }
    const int l  = w - st[lo];
    const int gx = T.blk[lo].grid[fam][0], gy = T.blk[lo].grid[fam][1];
    bx = l % gx;
    by = (l / gx) % gy;
    bz = l / (gx * gy);
    return lo;
}

__device__ __forceinline__ double* ens_u(const EnsBlock* E, int s) {
    return s == 0 ? E->Un : (s == 1 ? E->Ua : E->Ub);
}
__device__ __forceinline__ double* ens_w(const EnsBlock* E, int s) {
    return (s & 1) ? E->Ws[1] : E->Ws[0];
}

__global__ void kb_reset_min(EnsTable T) {
    int b = threadIdx.x + blockIdx.x * blockDim.x;
    if (b < T.n) *T.blk[b].dtmin = (unsigned long long)__double_as_longlong(1E10);
}

__global__ void kb_time_accum(EnsTable T) {
    int b = threadIdx.x + blockIdx.x * blockDim.x;
    if (b >= T.n) return;
    const EnsBlock* E = T.blk + b;
    *E->dtime = *E->dtime + __longlong_as_double((long long)*E->dtmin) * E->ph.cfl;
}

// CPY: the blocks' float copies of the stage state are written (step prologue, residual),
// filled (halo) and read by the sweeps
template <bool CPY>
__global__ __launch_bounds__(256) void kb_step_begin(EnsTable T, double* Ua) {
    int             bx, by, bz;
    const EnsBlock* E   = T.blk + ens_locate(T, EF_STEP_BEGIN, bx, by, bz);
    const Geo       g   = E->g;
    const double    gam = E->ph.gamma;
    const int       nwg = E->grid[EF_STEP_BEGIN][0];
    double *W = E->W, *Un = E->Un, *Ws = E->Ws[0];
    unsigned long long* dtmin = E->dtmin;
    step_begin_body(W, Ua, Un, Ws, dtmin, g, gam, bx, nwg, CPY ? E->Uc : nullptr,
                    CPY ? E->Wc : nullptr);
}

// the link of face `f` of block E: the neighbour's stage arrays and its extent along the normal
__device__ __forceinline__ HaloLink ens_link(const EnsTable& T, const EnsBlock* E, int f, int s) {
    const int nb = E->nbr[f];
    if (nb < 0) return HaloLink{nullptr, nullptr, 0, 0, 0, 0};
    const EnsBlock* N = T.blk + nb;
    const int       a = f >> 1;
    return HaloLink{ens_u(N, s), ens_w(N, s), a == 0 ? N->g.nx : (a == 1 ? N->g.ny : N->g.nz),
                    N->g.sx, N->g.sy, N->g.P};
}

// CPY: the copies' halos, and of the double arrays only those of the primitives across x and y
// (the residual and the derivative pass read no others; the sweeps read the copies)
template <bool CPY>
__global__ void kb_halo(EnsTable T, int s) {
    int             bx, by, bz;
    const EnsBlock* E  = T.blk + ens_locate(T, EF_HALO, bx, by, bz);
    const Geo       g  = E->g;
    const Bnd       bd = E->b;
    double *U = ens_u(E, s), *Wp = ens_w(E, s);
    if constexpr (CPY) {  // (no linked blocks)
        halo_body(U, 0, Wp, bz < 4 ? 5 : 0, g, bd, bx, by, bz);
        halo_body(E->Uc, 5, E->Wc, 5, g, bd, bx, by, bz);
        return;
    }
    if (E->linked) {
        const HaloLink lk = ens_link(T, E, bz, s);  // bz is the face
        halo_body(U, 5, Wp, 5, g, bd, bx, by, bz, &lk);
    } else {
        halo_body(U, 5, Wp, 5, g, bd, bx, by, bz);
    }
}

template <int DIR, int CONV, int WT, int OPT, bool CPY>
__global__ __launch_bounds__(512, NS_TILE_WAVES) void kb_flux_tile_xy(EnsTable T, int s) {
    __shared__ real tile[(TILE_S + 5) * split_nv<OPT>() * TILE_L];
    __shared__ real lq[WT == 2 ? TILE_S * LQ_REALS : 1];
    int             bx, by, bz;
    const EnsBlock* E   = T.blk + ens_locate(T, DIR == 0 ? EF_FLUX_X : EF_FLUX_Y, bx, by, bz);
    const Geo       g   = E->g;
    const real      gam = (real)E->ph.gamma;
    const int       kw  = E->t.kw_flux;
    const double *U = ens_u(E, s), *Wp = ens_w(E, s);
    flux_t*       f = (flux_t*)E->F[DIR];
    const MlpPacked* mlp = E->ph.mlp;
    // the whole block: faces 0 .. nx-1 of the x sweep, planes 0 .. nx-1 of the y sweep
    if constexpr (CPY)
        flux_tile_xy_body<DIR, CONV, WT, OPT>(E->Uc, E->Wc, f, g, gam, bx, by, bz, kw, tile, mlp,
                                              0, g.nx, WT == 2 ? lq : nullptr);
    else
        flux_tile_xy_body<DIR, CONV, WT, OPT>(U, Wp, f, g, gam, bx, by, bz, kw, tile, mlp, 0, g.nx,
                                              WT == 2 ? lq : nullptr);
}

template <int CONV, int WT, int OPT, bool CPY>
__global__ __launch_bounds__(512, NS_TILE_WAVES) void kb_flux_tile_z(EnsTable T, int s) {
    __shared__ real tile[8 * zlds_wave_max<OPT>()];
    __shared__ real lq[WT == 2 ? 8 * LQ_REALS : 1];
    int             bx, by, bz;
    const EnsBlock* E   = T.blk + ens_locate(T, EF_FLUX_Z, bx, by, bz);
    const Geo       g   = E->g;
    const real      gam = (real)E->ph.gamma;
    const int       kw  = E->t.kw_flux;
    const double *U = ens_u(E, s), *Wp = ens_w(E, s);
    flux_t*       f = (flux_t*)E->F[2];
    const MlpPacked* mlp = E->ph.mlp;
    if constexpr (CPY)
        flux_tile_z_body<CONV, WT, OPT>(E->Uc, E->Wc, f, g, gam, bx, by, bz, kw, tile, mlp,
                                        WT == 2 ? lq : nullptr);
    else
        flux_tile_z_body<CONV, WT, OPT>(U, Wp, f, g, gam, bx, by, bz, kw, tile, mlp,
                                        WT == 2 ? lq : nullptr);
}

template <int DIR, int CONV, int WT, int OPT, bool CPY>
__global__ __launch_bounds__(256) void kb_flux_plane(EnsTable T, int s) {
    int             bx, by, bz;
    const EnsBlock* E   = T.blk + ens_locate(T, EF_PLANE_X + DIR, bx, by, bz);
    const Geo       g   = E->g;
    const real      gam = (real)E->ph.gamma;
    const int       n   = (DIR == 0) ? g.nx : (DIR == 1) ? g.ny : g.nz;
    const double *U = ens_u(E, s), *Wp = ens_w(E, s);
    flux_t*       f = (flux_t*)E->F[DIR];
    const MlpPacked* mlp = E->ph.mlp;
    __shared__ real lq[WT == 2 ? 4 * LQ_REALS : 1];
    if constexpr (CPY)
        flux_plane_body<DIR, CONV, WT, OPT>(E->Uc, E->Wc, f, g, gam, NGC + n - 1, bx, by, mlp,
                                            WT == 2 ? lq : nullptr);
    else
        flux_plane_body<DIR, CONV, WT, OPT>(U, Wp, f, g, gam, NGC + n - 1, bx, by, mlp,
                                            WT == 2 ? lq : nullptr);
}

// ensemble members take the same derivative split as single blocks (k_residual_fd): dW/dx of
// u and v from the derivative pass, the rest formed in the residual
// (blocks of a decomposition take the separate path: their derivative halos are exchanged)
__device__ __forceinline__ bool ens_fused(const Geo& g, int linked = 0) {
    return NS_FUSE_DIFF && (long)g.NX * g.sx < (1L << 31) && !linked;
}

template <int ORDER>
__global__ __launch_bounds__(256) void kb_deriv1(EnsTable T, int s) {
    int             bx, by, bz;
    const EnsBlock* E  = T.blk + ens_locate(T, EF_DERIV1, bx, by, bz);
    const Geo       g  = E->g;
    const int       kw = E->t.kw_cell, march = E->t.march;
    const double*   Wp = ens_w(E, s);
    double *Dx = E->Dx, *Dy = E->Dy;
    if (ens_fused(g, E->linked)) deriv1_body<ORDER, false>(Wp, Dx, Dy, g, bx, by, bz, kw, march);
    else deriv1_body<ORDER>(Wp, Dx, Dy, g, bx, by, bz, kw, march);
}

__global__ void kb_deriv_halo(EnsTable T) {
    int             bx, by, bz;
    const EnsBlock* E  = T.blk + ens_locate(T, EF_DERIV_HALO, bx, by, bz);
    const Geo       g  = E->g;
    const Bnd       bd = E->b;
    double *Dx = E->Dx, *Dy = E->Dy;
    if (E->linked) {
        // the face of this (side, low/high) slice, and its neighbour's derivative planes
        const int side = bz / (2 * NGC), low = ((bz / NGC) & 1) == 0;
        const int f    = side == 0 ? (low ? 2 : 3) : (low ? 0 : 1);
        const int nb   = E->nbr[f];
        HaloLink  lk{nullptr, nullptr, 0, 0, 0, 0};
        if (nb >= 0) {
            const EnsBlock* N = T.blk + nb;
            lk = HaloLink{N->Dx, N->Dy, side == 0 ? N->g.ny : N->g.nx, N->g.sx, N->g.sy, N->g.P};
        }
        deriv_halo_body(Dx, Dy, g, bd, bx, by, bz, 0, &lk);
    } else if (!ens_fused(g)) {
        deriv_halo_body(Dx, Dy, g, bd, bx, by, bz);
    } else if (bz < 2 * NGC) {
        deriv_halo_body(Dx, Dy, g, bd, bx, by, bz, 2);  // j halos of du/dx, dv/dx
    }
}

// which / Rin / Rout are launch arguments (3, null, null) rather than constants, so that the body
// keeps the control flow of the single-block kernel
template <int ORDER, bool SUTH, bool CPY>
__global__ __launch_bounds__(256) void kb_residual(EnsTable T, int s, int which,
                                                   const double* Rin, double* Rout) {
    int             bx, by, bz;
    const EnsBlock* E  = T.blk + ens_locate(T, EF_RESIDUAL, bx, by, bz);
    const Geo       g  = E->g;
    const Phys      ph = E->ph;
    const int       kw = E->t.kw_cell, march = E->t.march;
    const flux_t *fx = (const flux_t*)E->F[0], *fy = (const flux_t*)E->F[1],
                 *fz = (const flux_t*)E->F[2];
    const double *Dx = E->Dx, *Dy = E->Dy;
    const double* Wp = ens_w(E, s);
    StageArgs     a{E->Un, ens_u(E, s), s == 1 ? E->Ub : E->Ua,
                    s == 0 ? E->Ws[1] : (s == 1 ? E->Ws[0] : E->W), s, E->flags};
    a.Uf = E->Uc, a.Wf = E->Wc;
    const unsigned long long* dtmin = E->dtmin;
    if constexpr (ORDER > 0) {
        if (ens_fused(g, E->linked)) {
            const Bnd bd = E->b;
            if (g.square_xy)
                residual_fd_body<ORDER, true, SUTH, 0, false, false, CPY>(
                    fx, fy, fz, Wp, Dx, Rin, Rout, a, dtmin, which, g, ph, bd, bx, by, bz, kw,
                    march);
            else
                residual_fd_body<ORDER, false, SUTH, 0, false, false, CPY>(
                    fx, fy, fz, Wp, Dx, Rin, Rout, a, dtmin, which, g, ph, bd, bx, by, bz, kw,
                    march);
            return;
        }
    }
    if (g.square_xy)
        residual_body<ORDER, true, SUTH>(fx, fy, fz, Wp, Dx, Dy, Rin, Rout, a, dtmin, which, g,
                                         ph, bx, by, bz, kw, march);
    else
        residual_body<ORDER, false, SUTH>(fx, fy, fz, Wp, Dx, Dy, Rin, Rout, a, dtmin, which, g,
                                          ph, bx, by, bz, kw, march);
}

// ------------------------------------------------------------------------------------------
// the learned reconstruction on a batch of stencils (checks of the device function the face
// kernels use); one stencil per lane, 64 lanes per block
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_mlp_recon(const MlpPacked* __restrict__ P, long n,
                                                  const double* __restrict__ S, int side,
                                                  double* __restrict__ out, int* __restrict__ regime) {
    const long s     = threadIdx.x + (long)blockIdx.x * 64;
    const bool valid = s < n;
    const long sc    = valid ? s : 0;
    real       v[5];
#pragma unroll
    for (int k = 0; k < 5; k++) v[k] = (real)S[sc * 5 + k];
    int  rg = 0;
    real r  = side ? weno_learned<1>(v[0], v[1], v[2], v[3], v[4], Learn{P, valid}, &rg)
                   : weno_learned<0>(v[0], v[1], v[2], v[3], v[4], Learn{P, valid}, &rg);
    if (valid) out[s] = (double)r, regime[s] = rg;
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------
static dim3 cell_grid(int n2, int n1, int n0, dim3 blk) {
    return dim3((n2 + blk.x - 1) / blk.x, (n1 + blk.y - 1) / blk.y, n0);
}
static const dim3 BLK(64, 4, 1);

static void l_reset_min(unsigned long long* bits, hipStream_t s) {
    k_reset_min<<<1, 1, 0, s>>>(bits);
}
static void l_zero_flags(unsigned* flags, hipStream_t s) { k_zero_flags<<<1, 1, 0, s>>>(flags); }
static void l_set_dt(unsigned long long* bits, double dt_over_cfl, hipStream_t s) {
    k_set_dt<<<1, 1, 0, s>>>(bits, dt_over_cfl);
}

static void l_step_begin(const double* W, double* Ua, double* Un, double* Wst,
                         unsigned long long* dtmin, Geo g, Phys ph, hipStream_t s,
                         const StateCopy* cp) {
    long n  = (long)g.nx * g.ny * g.nz;
    long nb = (n + 255) / 256;
    if (nb > 2048) nb = 2048;
    float* Uf = (cp && cp->U) ? cp->U : nullptr;
    k_step_begin<<<(unsigned)nb, 256, 0, s>>>(W, Ua, Un, Wst, dtmin, g, ph.gamma, Uf,
                                              Uf ? cp->W : nullptr);
}

static void l_halo_faces(double* A, int nA, double* B, int nB, Geo g, Bnd b, int f0, int nf,
                         hipStream_t s, const StateCopy* cp) {
    // x, y faces: (k, tangential row, layer); z faces: ((j, layer), i)
    int  m1 = g.ny > g.nx ? g.ny : g.nx;
    int  mx = g.nz > NGC * g.ny ? g.nz : NGC * g.ny;
    dim3 blk(64, 4, 1);
    dim3 grd((mx + 63) / 64, NGC * ((m1 + 3) / 4), nf);
    if (nf <= 0) return;
    if (cp && cp->U)
        k_halo_cp<<<grd, blk, 0, s>>>(A, nA, B, B ? nB : 0, (float*)cp->U, (float*)cp->W, g, b,
                                      f0);
    else k_halo<<<grd, blk, 0, s>>>(A, nA, B, B ? nB : 0, g, b, f0);
}

static void l_halo(double* A, int nA, double* B, int nB, Geo g, Bnd b, hipStream_t s) {
    l_halo_faces(A, nA, B, nB, g, b, 0, 6, s, nullptr);
}

static void l_p2c_faces(const double* W, double* U, Geo g, Phys ph, hipStream_t s) {
    k_p2c_faces<<<cell_grid(g.NZ, g.NY, g.NX, BLK), BLK, 0, s>>>(W, U, g, ph.gamma);
}

static void l_c2p_faces(const double* U, double* W, Geo g, Phys ph, hipStream_t s) {
    k_c2p_faces<<<cell_grid(g.NZ, g.NY, g.NX, BLK), BLK, 0, s>>>(U, W, g, ph.gamma);
}

// parts: 1 the tile kernel (faces 0 .. n-1 of the sweep), 2 the last face plane, 3 both.
// [lo, hi): for the x sweep the faces f of the launch (f between cells f-1 and f, 0 <= f <= nx:
// the tile kernel takes f < nx, the last plane f = nx), for the y and z sweeps the planes i (both
// kernels); the whole block is [0, nx + 1) resp. [0, nx).  lo is a multiple of 4 for y and z.
// samples per side of the emit-mode buffers a launch of nb workgroups of fb faces needs: every
// region holds up to 5 samples (one per variable) of each face of its workgroups
static long learn_capr(long nb, long fb) { return (nb + LS_REGIONS - 1) / LS_REGIONS * fb * 5; }

// the tile launch of a learned component-wise sweep in emit mode: reset the counters of the
// sweep's sample set, the tiles, then the network pass over the left and then the right samples
template <int DIR, class ST>
static bool flux_emit(const ST* U, const ST* W, flux_t* f, Geo g, double gamma,
                      const MlpPacked* mlp, const LearnSink* sink, dim3 grd, dim3 blk, int lo,
                      int hi, hipStream_t s) {
    const long nb   = (long)grd.x * grd.y * grd.z;
    const long capr = learn_capr(nb, (long)blk.x * blk.y);
    if (!sink || capr * LS_REGIONS > sink->cap || 4 * g.P + (long)g.NX * g.sx > 0xffffffffL)
        return false;
    const int set = (DIR == 2) ? 1 : 0;
    const int par = sink->parity[set];
    sink->parity[set] ^= 1;
    LearnEmit e{{(real*)sink->v[set][0], (real*)sink->v[set][1]},
                {sink->dest[set][0], sink->dest[set][1]}, sink->cnt[set][par],
                sink->cnt[set][par ^ 1], sink->cap, capr};
    if constexpr (DIR == 2) k_flux_tile_z_emit<<<grd, blk, 0, s>>>(U, W, f, g, gamma, lo, e);
    else k_flux_tile_xy_emit<DIR><<<grd, blk, 0, s>>>(U, W, f, g, gamma, lo, hi, e);
    // enough wavefronts to cover the occupied chunks of a typical launch several times over;
    // each strides over the rest
    static const long max_blk = [] {
        const char* e = std::getenv("NSGPU_NET_BLOCKS");
        const long  v = e ? std::atol(e) : 0;
        return v > 0 ? v : 2048L;
    }();
    const long chunks = (capr * LS_REGIONS + 63) / 64;
    const long nblk   = std::min<long>(max_blk, std::max<long>(1, (chunks + 3) / 4));
    k_learn_net<<<(unsigned)nblk, 256, 0, s>>>(mlp, e, 0, f);
    k_learn_net<<<(unsigned)nblk, 256, 0, s>>>(mlp, e, 1, f);
    return true;
}

template <int DIR, int CONV, int B, int OPT, class ST>
static void flux_dir_st(const ST* U, const ST* W, flux_t* f, Geo g, double gamma,
                        const MlpPacked* mlp, int parts, int lo, int hi, hipStream_t s,
                        const LearnSink* sink) {
    constexpr bool EMIT = (CONV == 1 && B == 2 && OPT == 0);
    if (!(parts & 1)) {
    } else if constexpr (DIR == 2) {
        dim3 blk(TILE_L, 8, 1);
        dim3 grd((g.nz + TILE_L - 1) / TILE_L, (g.ny + 7) / 8, hi - lo);
        if (hi > lo && !(EMIT && flux_emit<DIR>(U, W, f, g, gamma, mlp, sink, grd, blk, lo, hi, s)))
            k_flux_tile_z<CONV, B, OPT><<<grd, blk, 0, s>>>(U, W, f, g, gamma, mlp, lo);
    } else if constexpr (DIR == 0) {
        const int fe = hi < g.nx ? hi : g.nx;  // faces of the tile kernel: [lo, fe)
        constexpr int TS = tile_s1<CONV>();
        dim3      blk(TILE_L, TS, 1);
        dim3      grd((g.nz + TILE_L - 1) / TILE_L, g.ny, (fe - lo + TS - 1) / TS);
        if (fe > lo && !(EMIT && flux_emit<DIR>(U, W, f, g, gamma, mlp, sink, grd, blk, lo, fe, s)))
            k_flux_tile_xy<DIR, CONV, B, OPT><<<grd, blk, 0, s>>>(U, W, f, g, gamma, mlp, lo, fe);
    } else {
        constexpr int TS = tile_s1<CONV>();
        dim3 blk(TILE_L, TS, 1);
        dim3 grd((g.nz + TILE_L - 1) / TILE_L, hi - lo, (g.ny + TS - 1) / TS);
        if (hi > lo && !(EMIT && flux_emit<DIR>(U, W, f, g, gamma, mlp, sink, grd, blk, lo, hi, s)))
            k_flux_tile_xy<DIR, CONV, B, OPT><<<grd, blk, 0, s>>>(U, W, f, g, gamma, mlp, lo, hi);
    }
    if (!(parts & 2)) return;
    // the last face plane (sweep index NGC + n - 1)
    int  n  = (DIR == 0) ? g.nx : (DIR == 1) ? g.ny : g.nz;
    int  n2 = (DIR == 2) ? g.ny : g.nz;
    dim3 pb(64, 4, 1);
    if constexpr (DIR == 0) {
        if (hi <= g.nx) return;  // face nx is not in this launch
        dim3 pg((n2 + 63) / 64, (g.ny + 3) / 4, 1);
        k_flux_plane<DIR, CONV, B, OPT><<<pg, pb, 0, s>>>(U, W, f, g, gamma, NGC + n - 1, mlp, 0);
    } else {
        if (hi <= lo) return;
        dim3 pg((n2 + 63) / 64, (hi - lo + 3) / 4, 1);
        k_flux_plane<DIR, CONV, B, OPT><<<pg, pb, 0, s>>>(U, W, f, g, gamma, NGC + n - 1, mlp,
                                                          lo / 4);
    }
}

// cp: the float copies of (U, W) with their halos (single-precision build, null: none).  The
// face kernels convert every value they load to real, so reading the copies gives the same
// fluxes from half the bytes.  The per-face alpha and corrected projection kernels read the
// double arrays.
template <int DIR, int CONV, int B, int OPT>
static void flux_dir(const double* U, const double* W, flux_t* f, Geo g, double gamma,
                     const MlpPacked* mlp, int parts, int lo, int hi, hipStream_t s,
                     const LearnSink* sink = nullptr, const StateCopy* cp = nullptr) {
#if NS_FP32
    if constexpr (OPT == 0) {
        if (cp && cp->U) {
            flux_dir_st<DIR, CONV, B, OPT>(cp->U, cp->W, f, g, gamma, mlp, parts, lo, hi, s, sink);
            return;
        }
    }
#endif
    flux_dir_st<DIR, CONV, B, OPT>(U, W, f, g, gamma, mlp, parts, lo, hi, s, sink);
}

// kernel options of a configuration: bit 0 per-face alpha, bit 1 corrected characteristic
// projection (characteristic sweeps only).  The learned weights run with neither (the host
// refuses the combination).
static int flux_opt(const Phys& ph) {
    return (ph.alpha_mode == 1 ? 1 : 0) | ((ph.conv == 2 && (ph.physics & 1)) ? 2 : 0);
}

template <int DIR, int CONV, int B>
static void flux_dir_opt(const double* U, const double* W, flux_t* f, Geo g, double gamma,
                         const MlpPacked* m, int opt, int parts, int lo, int hi, hipStream_t s,
                         const LearnSink* sink, const StateCopy* cp) {
    if constexpr (B == 2) {
        flux_dir<DIR, CONV, 2, 0>(U, W, f, g, gamma, m, parts, lo, hi, s, sink, cp);
    } else if constexpr (CONV == 1) {
        if (opt & 1) flux_dir<DIR, 1, B, 1>(U, W, f, g, gamma, m, parts, lo, hi, s);
        else flux_dir<DIR, 1, B, 0>(U, W, f, g, gamma, m, parts, lo, hi, s, nullptr, cp);
    } else {
        if (opt == 1) flux_dir<DIR, 2, B, 1>(U, W, f, g, gamma, m, parts, lo, hi, s);
        else if (opt == 2) flux_dir<DIR, 2, B, 2>(U, W, f, g, gamma, m, parts, lo, hi, s);
        else if (opt == 3) flux_dir<DIR, 2, B, 3>(U, W, f, g, gamma, m, parts, lo, hi, s);
        else flux_dir<DIR, 2, B, 0>(U, W, f, g, gamma, m, parts, lo, hi, s, nullptr, cp);
    }
}

template <int DIR>
static void face_flux_dir(const double* U, const double* W, flux_t* f, Geo g, Phys ph,
                          int parts, int lo, int hi, hipStream_t s, const StateCopy* cp) {
    const MlpPacked* m   = ph.mlp;
    const int        opt = flux_opt(ph);
    const LearnSink* sk  = ph.sink;
    if (ph.conv == 1) {
        if (ph.weights == 1)
            flux_dir_opt<DIR, 1, 1>(U, W, f, g, ph.gamma, m, opt, parts, lo, hi, s, nullptr, cp);
        else if (ph.weights == 2)
            flux_dir_opt<DIR, 1, 2>(U, W, f, g, ph.gamma, m, opt, parts, lo, hi, s, sk, cp);
        else flux_dir_opt<DIR, 1, 0>(U, W, f, g, ph.gamma, m, opt, parts, lo, hi, s, nullptr, cp);
    } else {
        if (ph.weights == 1)
            flux_dir_opt<DIR, 2, 1>(U, W, f, g, ph.gamma, m, opt, parts, lo, hi, s, nullptr, cp);
        else if (ph.weights == 2)
            flux_dir_opt<DIR, 2, 2>(U, W, f, g, ph.gamma, m, opt, parts, lo, hi, s, nullptr, cp);
        else flux_dir_opt<DIR, 2, 0>(U, W, f, g, ph.gamma, m, opt, parts, lo, hi, s, nullptr, cp);
    }
}

static void l_face_flux(int dir, const double* U, const double* W, double* f, Geo g, Phys ph,
                        int parts, int lo, int hi, hipStream_t s, const StateCopy* cp) {
    flux_t* ff = (flux_t*)f;  // the build's flux precision (see flux_t)
    if (dir == 0) face_flux_dir<0>(U, W, ff, g, ph, parts, lo, hi, s, cp);
    else if (dir == 1) face_flux_dir<1>(U, W, ff, g, ph, parts, lo, hi, s, cp);
    else face_flux_dir<2>(U, W, ff, g, ph, parts, lo, hi, s, cp);
}

// planes [lo, hi) (lo a multiple of MARCH)
static void l_deriv1(const double* W, double* Dx, double* Dy, Geo g, Phys ph, int lo, int hi,
                     hipStream_t s) {
    if (hi <= lo) return;
    const dim3 blk(64, DERIV_ROWS, 1);
    const dim3 grd((g.nz + 63) / 64, (g.ny + DERIV_ROWS - 1) / DERIV_ROWS,
                   (hi - lo + MARCH - 1) / MARCH);
    const int  z0 = lo / MARCH;
    if (Dy) {
        if (ph.diff_order == 2) k_deriv1<2, true><<<grd, blk, 0, s>>>(W, Dx, Dy, g, z0);
        else if (ph.diff_order == 4) k_deriv1<4, true><<<grd, blk, 0, s>>>(W, Dx, Dy, g, z0);
        else k_deriv1<6, true><<<grd, blk, 0, s>>>(W, Dx, Dy, g, z0);
    } else {
        if (ph.diff_order == 2) k_deriv1<2, false><<<grd, blk, 0, s>>>(W, Dx, Dy, g, z0);
        else if (ph.diff_order == 4) k_deriv1<4, false><<<grd, blk, 0, s>>>(W, Dx, Dy, g, z0);
        else k_deriv1<6, false><<<grd, blk, 0, s>>>(W, Dx, Dy, g, z0);
    }
}

// without Dy: the j halos of dW/dx of u and v on planes [lo, hi) (lo a multiple of 4); with Dy
// both halo families of the whole block (lo, hi ignored)
static void l_deriv_halo(double* Dx, double* Dy, Geo g, Bnd b, int lo, int hi, hipStream_t s) {
    int  m1 = g.nx > g.ny ? g.nx : g.ny;
    dim3 blk(64, 4, 1);
    if (Dy) {
        dim3 grd((g.nz + 63) / 64, (m1 + 3) / 4, 4 * NGC);
        k_deriv_halo<<<grd, blk, 0, s>>>(Dx, Dy, g, b, 0);
    } else if (hi > lo) {
        dim3 grd((g.nz + 63) / 64, (hi - lo + 3) / 4, 2 * NGC);
        k_deriv_halo<<<grd, blk, 0, s>>>(Dx, Dy, g, b, lo / 4);
    }
}

template <int ORDER>
static void residual_order(const flux_t* fx, const flux_t* fy, const flux_t* fz, const double* W,
                           const double* Dx, const double* Dy, const double* Rin, double* Rout,
                           StageArgs a, const unsigned long long* dtmin, int which, Geo g,
                           Phys ph, hipStream_t s) {
    dim3 grd((g.nz + KW_CELL - 1) / KW_CELL, (g.ny + 256 / KW_CELL - 1) / (256 / KW_CELL),
             (g.nx + MARCH - 1) / MARCH);
    if (ORDER > 0 && (ph.physics & 2)) {
        k_residual<ORDER, false, false, true><<<grd, BLK, 0, s>>>(fx, fy, fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph);
    } else if (ph.suth) {
        if (g.square_xy)
            k_residual<ORDER, true, true><<<grd, BLK, 0, s>>>(fx, fy, fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph);
        else
            k_residual<ORDER, false, true><<<grd, BLK, 0, s>>>(fx, fy, fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph);
    } else {
        if (g.square_xy)
            k_residual<ORDER, true, false><<<grd, BLK, 0, s>>>(fx, fy, fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph);
        else
            k_residual<ORDER, false, false><<<grd, BLK, 0, s>>>(fx, fy, fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph);
    }
}

// planes [lo, hi) (lo a multiple of FUSED_MARCH)
template <int ORDER>
static void residual_fd_order(const flux_t* fx, const flux_t* fy, const flux_t* fz,
                              const double* W, const double* Dx, const double* Rin, double* Rout,
                              StageArgs a,
                              const unsigned long long* dtmin, int which, Geo g, Phys ph, Bnd b,
                              int lo, int hi, hipStream_t s) {
    if (hi <= lo) return;
    const dim3 blk(64, 4, 1);
    const dim3 grd((g.nz + KW_CELL - 1) / KW_CELL, (g.ny + 256 / KW_CELL - 1) / (256 / KW_CELL),
                   (hi - lo + FUSED_MARCH - 1) / FUSED_MARCH);
    const int  z0 = lo / FUSED_MARCH;
    auto go = [&](auto nx) {
        constexpr bool NX = decltype(nx)::value;
        if (ph.suth) {
            if (g.square_xy)
                k_residual_fd<ORDER, true, true, NX><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, Rin, Rout, a, dtmin, which, g, ph, b, z0);
            else
                k_residual_fd<ORDER, false, true, NX><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, Rin, Rout, a, dtmin, which, g, ph, b, z0);
        } else {
            if (g.square_xy)
                k_residual_fd<ORDER, true, false, NX><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, Rin, Rout, a, dtmin, which, g, ph, b, z0);
            else
                k_residual_fd<ORDER, false, false, NX><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, Rin, Rout, a, dtmin, which, g, ph, b, z0);
        }
    };
    // the last stage of a step that another follows forms the next step's prologue
    if (a.Unext && a.Uout) go(std::true_type{});
    else go(std::false_type{});
}

static void l_residual_fd(const double* fx, const double* fy, const double* fz, const double* W,
                          const double* Dx, const double* Rin, double* Rout, const StageArgs* st,
                          const unsigned long long* dtmin, int which, Geo g, Phys ph, Bnd b,
                          int lo, int hi, hipStream_t s) {
    StageArgs a{};
    if (st) a = *st;
    switch (ph.diff_order) {
    case 2: residual_fd_order<2>((const flux_t*)fx, (const flux_t*)fy, (const flux_t*)fz, W, Dx, Rin, Rout, a, dtmin, which, g, ph, b, lo, hi, s); break;
    case 4: residual_fd_order<4>((const flux_t*)fx, (const flux_t*)fy, (const flux_t*)fz, W, Dx, Rin, Rout, a, dtmin, which, g, ph, b, lo, hi, s); break;
    default: residual_fd_order<6>((const flux_t*)fx, (const flux_t*)fy, (const flux_t*)fz, W, Dx, Rin, Rout, a, dtmin, which, g, ph, b, lo, hi, s); break;
    }
}

#ifndef NS_ZFD
#define NS_ZFD 1  // build k_residual_zfd (the host runs it when asked, $NSGPU_ZFD=1)
#endif
// segment width of k_residual_zfd for the configuration, 0 where it does not cover it (no z
// convection, learned or Borges weights, kernel options, other orders than CD4, 3-D viscous
// terms)
static int l_zfd_kw(Geo g, Phys ph) {
    if (!NS_ZFD || g.dim != 3 || !ph.conv || ph.weights != 0 || flux_opt(ph) != 0 ||
        ph.diff_order != 4 || (ph.physics & 2))
        return 0;
    return g.nz <= 64 ? 64 : (g.nz <= 128 ? 128 : 256);
}

template <int ZC>
static void residual_zfd_conv(const double* U, const flux_t* fx, const flux_t* fy, flux_t* fz,
                              const double* W, const double* Dx, const StageArgs& a,
                              const unsigned long long* dtmin, Geo g, Phys ph, Bnd b, int kw,
                              int lo, int hi, hipStream_t s) {
    // the first face of every segment, planes i in [lo, hi)
    const int  nseg = (g.nz + kw - 1) / kw;
    const dim3 pb(64, 4, 1), pg((g.ny + 63) / 64, (hi - lo + 3) / 4, 1);
    for (int m = 0; m < nseg; m++)
        k_flux_plane<2, ZC, 0, 0><<<pg, pb, 0, s>>>(U, W, fz, g, ph.gamma, NGC - 1 + m * kw,
                                                    nullptr, lo / 4);
    const dim3 blk(64, 4, 1);
    const dim3 grd(nseg, (g.ny + 256 / kw - 1) / (256 / kw),
                   (hi - lo + FUSED_MARCH - 1) / FUSED_MARCH);
    const int  z0 = lo / FUSED_MARCH;
    if (ph.suth) {
        if (g.square_xy)
            k_residual_zfd<4, true, true, ZC><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, a, dtmin, g, ph, b, z0, kw);
        else
            k_residual_zfd<4, false, true, ZC><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, a, dtmin, g, ph, b, z0, kw);
    } else {
        if (g.square_xy)
            k_residual_zfd<4, true, false, ZC><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, a, dtmin, g, ph, b, z0, kw);
        else
            k_residual_zfd<4, false, false, ZC><<<grd, blk, 0, s>>>(fx, fy, fz, W, Dx, a, dtmin, g, ph, b, z0, kw);
    }
}

// residual + stage update of planes [lo, hi) (lo a multiple of FUSED_MARCH) with the z faces
// formed in the kernel; U, W: the stage's input state (st->Ucur and the primitives)
static void l_residual_zfd(const double* fx, const double* fy, double* fz, const double* W,
                           const double* Dx, const StageArgs* st,
                           const unsigned long long* dtmin, Geo g, Phys ph, Bnd b, int lo,
                           int hi, hipStream_t s) {
    const int kw = l_zfd_kw(g, ph);
    if (hi <= lo || !kw || !st) return;
    if (ph.conv == 1)
        residual_zfd_conv<1>(st->Ucur, (const flux_t*)fx, (const flux_t*)fy, (flux_t*)fz, W, Dx, *st, dtmin, g, ph, b, kw, lo, hi, s);
    else
        residual_zfd_conv<2>(st->Ucur, (const flux_t*)fx, (const flux_t*)fy, (flux_t*)fz, W, Dx, *st, dtmin, g, ph, b, kw, lo, hi, s);
}

static void l_residual(const double* fx, const double* fy, const double* fz, const double* W,
                       const double* Dx, const double* Dy, const double* Rin, double* Rout,
                       const StageArgs* st, const unsigned long long* dtmin, int which, Geo g,
                       Phys ph, hipStream_t s) {
    StageArgs a{};
    if (st) a = *st;
    switch (ph.diff_order) {
    case 2: residual_order<2>((const flux_t*)fx, (const flux_t*)fy, (const flux_t*)fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph, s); break;
    case 4: residual_order<4>((const flux_t*)fx, (const flux_t*)fy, (const flux_t*)fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph, s); break;
    case 6: residual_order<6>((const flux_t*)fx, (const flux_t*)fy, (const flux_t*)fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph, s); break;
    default: residual_order<0>((const flux_t*)fx, (const flux_t*)fy, (const flux_t*)fz, W, Dx, Dy, Rin, Rout, a, dtmin, which, g, ph, s); break;
    }
}

template <int ORDER>
static void visc3d_order(const double* W, double* Dx, double* Dy, double* Dz, double* Rv, Geo g,
                         Phys ph, Bnd b, hipStream_t s) {
    dim3 grd(g.nx, (g.nz + 63) / 64, (g.ny + 3) / 4);
    k_deriv3d<ORDER><<<grd, BLK, 0, s>>>(W, Dx, Dy, Dz, g);
    int  m2 = g.nz > g.ny ? g.nz : g.ny, m1 = g.ny > g.nx ? g.ny : g.nx;
    dim3 hg((m2 + 63) / 64, (m1 + 3) / 4, 6 * 2 * NGC);
    k_deriv3d_halo<<<hg, BLK, 0, s>>>(Dx, Dy, Dz, g, b);
    if (ph.suth) k_visc3d<ORDER, true><<<grd, BLK, 0, s>>>(W, Dx, Dy, Dz, Rv, g, ph);
    else k_visc3d<ORDER, false><<<grd, BLK, 0, s>>>(W, Dx, Dy, Dz, Rv, g, ph);
}

static void l_visc3d(const double* W, double* Dx, double* Dy, double* Dz, double* Rv, Geo g,
                     Phys ph, Bnd b, hipStream_t s) {
    if (ph.diff_order == 2) visc3d_order<2>(W, Dx, Dy, Dz, Rv, g, ph, b, s);
    else if (ph.diff_order == 4) visc3d_order<4>(W, Dx, Dy, Dz, Rv, g, ph, b, s);
    else visc3d_order<6>(W, Dx, Dy, Dz, Rv, g, ph, b, s);
}

static void l_aos_to_soa(const double* aos, double* soa, Geo g, hipStream_t s) {
    const long n = (long)g.NX * g.NY * g.NZ;
    k_aos_to_soa<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(aos, soa, g);
}
static void l_soa_to_aos(const double* soa, double* aos, Geo g, hipStream_t s) {
    const long n = (long)g.NX * g.NY * g.NZ;
    k_soa_to_aos<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(soa, aos, g);
}

static void l_reduce(const double* U, const double* W, double* partial, int nblocks_max,
                     unsigned long long* dtmin, unsigned* flags, Geo g, Phys ph, hipStream_t s,
                     int* nblocks_out) {
    long n  = (long)g.nx * g.ny * g.nz;
    long nb = (n + RED_THREADS - 1) / RED_THREADS;
    if (nb > nblocks_max) nb = nblocks_max;
    k_reduce<<<(unsigned)nb, RED_THREADS, 0, s>>>(U, W, partial, dtmin, flags, g, ph);
    *nblocks_out = (int)nb;
}

static void l_reduce_final(const double* partial, int nb, double* sums, hipStream_t s) {
    k_reduce_final<<<1, RED_THREADS, 0, s>>>(partial, nb, sums);
}

static void l_time_accum(double* t, const unsigned long long* dtmin, double cfl, hipStream_t s) {
    k_time_accum<<<1, 1, 0, s>>>(t, dtmin, cfl);
}

static void l_selftest_div(long n, unsigned long long seed, unsigned long long* count,
                           hipStream_t s) {
    k_selftest_div<<<1024, 256, 0, s>>>(n, seed, count);
}

template <int CONV, int B, int OPT, bool CPY>
static void ens_fluxes_t(const EnsTable& T, const int* total, int s, hipStream_t st) {
    dim3 tb(TILE_L, 8, 1), pb(64, 4, 1);
    if (total[EF_FLUX_X])
        kb_flux_tile_xy<0, CONV, B, OPT, CPY><<<total[EF_FLUX_X], tb, 0, st>>>(T, s);
    if (total[EF_PLANE_X])
        kb_flux_plane<0, CONV, B, OPT, CPY><<<total[EF_PLANE_X], pb, 0, st>>>(T, s);
    if (total[EF_FLUX_Y])
        kb_flux_tile_xy<1, CONV, B, OPT, CPY><<<total[EF_FLUX_Y], tb, 0, st>>>(T, s);
    if (total[EF_PLANE_Y])
        kb_flux_plane<1, CONV, B, OPT, CPY><<<total[EF_PLANE_Y], pb, 0, st>>>(T, s);
    if (total[EF_FLUX_Z]) kb_flux_tile_z<CONV, B, OPT, CPY><<<total[EF_FLUX_Z], tb, 0, st>>>(T, s);
    if (total[EF_PLANE_Z])
        kb_flux_plane<2, CONV, B, OPT, CPY><<<total[EF_PLANE_Z], pb, 0, st>>>(T, s);
}

// the copies serve the kernels without options (as for single blocks)
template <int CONV, int B, bool CPY>
static void ens_fluxes(const EnsTable& T, const int* total, int s, int opt, hipStream_t st) {
    if constexpr (B == 2) {
        ens_fluxes_t<CONV, 2, 0, CPY>(T, total, s, st);
    } else if constexpr (CONV == 1) {
        if (opt & 1) ens_fluxes_t<1, B, 1, false>(T, total, s, st);
        else ens_fluxes_t<1, B, 0, CPY>(T, total, s, st);
    } else {
        if (opt == 1) ens_fluxes_t<2, B, 1, false>(T, total, s, st);
        else if (opt == 2) ens_fluxes_t<2, B, 2, false>(T, total, s, st);
        else if (opt == 3) ens_fluxes_t<2, B, 3, false>(T, total, s, st);
        else ens_fluxes_t<2, B, 0, CPY>(T, total, s, st);
    }
}

template <int ORDER, bool SUTH, bool CPY>
static void ens_stage(const EnsTable& T, const int* total, const Phys& ph, int s, hipStream_t st) {
    kb_halo<CPY><<<total[EF_HALO], dim3(64, 4, 1), 0, st>>>(T, s);
    const int opt = flux_opt(ph);
    if (ph.conv == 1) {
        if (ph.weights == 1) ens_fluxes<1, 1, CPY>(T, total, s, opt, st);
        else if (ph.weights == 2) ens_fluxes<1, 2, CPY>(T, total, s, opt, st);
        else ens_fluxes<1, 0, CPY>(T, total, s, opt, st);
    } else if (ph.conv == 2) {
        if (ph.weights == 1) ens_fluxes<2, 1, CPY>(T, total, s, opt, st);
        else if (ph.weights == 2) ens_fluxes<2, 2, CPY>(T, total, s, opt, st);
        else ens_fluxes<2, 0, CPY>(T, total, s, opt, st);
    }
    if constexpr (ORDER > 0) {
        kb_deriv1<ORDER><<<total[EF_DERIV1], 256, 0, st>>>(T, s);
        kb_deriv_halo<<<total[EF_DERIV_HALO], dim3(64, 4, 1), 0, st>>>(T);
    }
    kb_residual<ORDER, SUTH, CPY><<<total[EF_RESIDUAL], 256, 0, st>>>(T, s, 3, nullptr, nullptr);
}

template <int ORDER, bool SUTH, bool CPY>
static void ens_step_c(const EnsTable& T, const int* total, const Phys& ph, hipStream_t st) {
    const unsigned nb = (unsigned)((T.n + 63) / 64);
    kb_reset_min<<<nb, 64, 0, st>>>(T);
    kb_step_begin<CPY><<<total[EF_STEP_BEGIN], 256, 0, st>>>(T, nullptr);
    for (int s = 0; s < 3; s++) ens_stage<ORDER, SUTH, CPY>(T, total, ph, s, st);
    kb_time_accum<<<nb, 64, 0, st>>>(T);
}

// the float copies: single-precision build, diffusion on (the fused residual writes them)
template <int ORDER, bool SUTH>
static void ens_step_t(const EnsTable& T, const int* total, const Phys& ph, hipStream_t st) {
#if NS_FP32
    if constexpr (ORDER > 0) {
        if (T.copies) {
            ens_step_c<ORDER, SUTH, true>(T, total, ph, st);
            return;
        }
    }
#endif
    ens_step_c<ORDER, SUTH, false>(T, total, ph, st);
}

static void l_ens_step(const EnsTable& T, const int* total, const Phys& ph, hipStream_t st) {
    const bool su = ph.suth != 0;
    switch (ph.diff_order) {
    case 2: su ? ens_step_t<2, true>(T, total, ph, st) : ens_step_t<2, false>(T, total, ph, st); break;
    case 4: su ? ens_step_t<4, true>(T, total, ph, st) : ens_step_t<4, false>(T, total, ph, st); break;
    case 6: su ? ens_step_t<6, true>(T, total, ph, st) : ens_step_t<6, false>(T, total, ph, st); break;
    default: su ? ens_step_t<0, true>(T, total, ph, st) : ens_step_t<0, false>(T, total, ph, st); break;
    }
}

static void l_mlp_recon(const MlpPacked* P, long n, const double* S, int side, double* out,
                        int* regime, hipStream_t s) {
    k_mlp_recon<<<(unsigned)((n + 63) / 64), 64, 0, s>>>(P, n, S, side, out, regime);
}

// samples per side of the emit-mode buffers that the full-block sweeps of g need, and the bytes
// of one sample value (the build's real type)
static long l_learn_cap(Geo g, int* value_bytes) {
    if (value_bytes) *value_bytes = (int)sizeof(real);
    const long kt = (g.nz + TILE_L - 1) / TILE_L;
    constexpr long TS = tile_s1<1>();  // (the emit-mode sweeps are component-wise)
    const long fx = learn_capr(kt * g.ny * ((g.nx + TS - 1) / TS), 64L * TS);
    const long fy = learn_capr(kt * g.nx * ((g.ny + TS - 1) / TS), 64L * TS);
    const long fz = learn_capr(kt * ((g.ny + 7) / 8) * g.nx, 512L);
    return std::max(fx, std::max(fy, fz)) * LS_REGIONS;
}

const Launchers& launchers() {
    static const Launchers L = {l_reset_min,    l_step_begin, l_halo,       l_c2p_faces,
                                l_face_flux,    l_deriv1,     l_deriv_halo, l_residual,
                                l_aos_to_soa,   l_soa_to_aos, l_reduce,     l_reduce_final,
                                l_time_accum,   l_zero_flags, l_selftest_div, l_ens_step,
                                l_mlp_recon,    l_set_dt,     l_visc3d,     l_p2c_faces,
                                l_residual_fd,  NS_FUSE_DIFF,
                                // plane ranges of the cell kernels start on a march boundary
                                (MARCH > FUSED_MARCH ? MARCH : FUSED_MARCH),
                                l_zfd_kw,       l_residual_zfd, l_learn_cap,
                                l_halo_faces};
    return L;
}

}  // namespace NS_MODE
}  // namespace nsgpu
