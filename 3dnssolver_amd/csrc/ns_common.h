// ns_common.h -- block geometry, device layout and the kernel launch table shared by the C-ABI
// (nsgpu_capi.hip) and the kernel translation units (ns_kernels.hip, compiled once per math mode).
//
// Device layout: rows of one variable.  A 5-variable field of the padded block holds, for every
// (i, j), the five rows of k (k fastest; the reference Array4d order, containers/array_4d.hpp:
// 116-124, with q moved out of the innermost index) side by side: [NX][NY][5][row], NX = nx +
// 2*NGC, so P (the stride between variables) is one row, sy = 5 rows and sx = NY*sy.  (The
// planar layout [5][NX][NY][row], P = one plane, is selected at run time; the kernels only use
// P, sx and sy.)  Rows are padded to NZ rounded up to KALIGN doubles and every array starts KOFF
// doubles into its allocation, so the first interior cell of every row (k = NGC) sits on a
// 128-byte boundary: a wavefront reading 64 consecutive interior k of one variable reads exactly
// four 128-byte lines.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsgpu {

constexpr int NGC = 3;
constexpr int NQ  = 5;
#ifndef NS_KALIGN
#define NS_KALIGN 16
#endif
constexpr int KALIGN = NS_KALIGN;                             // row alignment in doubles
constexpr int KOFF   = (KALIGN - NGC % KALIGN) % KALIGN;      // array start offset in doubles

struct Geo {
    int    nx, ny, nz;  // interior cells
    int    NX, NY, NZ;  // padded extents
    int    sx, sy;      // strides of i and j (k stride is 1)
    long   P;           // stride between the variables of an array (one row or one plane)
    double dx, dy, dz;  // Grid::dx() etc. (grid.cpp:14-16)
    double dxx, dyy;    // dx*dx, dy*dy (CDN second-derivative divisors, cd-n.cpp:67-79)
    double idx, idy, idz, idxx, idyy, idzz;  // correctly rounded reciprocals 1.0 / dx etc.
                                             // (FAST: factors; EXACT: the divisions' reciprocals)
    int    dim;
    int    square_xy;   // dx*dx == dy*dy bitwise: the two x-stencil second derivatives coincide
};

// per-face boundary action (resolved on the host from nsgpu_config::bc)
constexpr int FACE_PERIODIC = 0;  // the block's own opposite interior planes (single-rank wrap)
constexpr int FACE_NEUMANN  = 1;  // the first interior plane (a physical wall)
constexpr int FACE_LINK     = 2;  // the neighbouring block of a decomposition (its interior planes)
struct Bnd {
    int periodic[3];   // the axis wraps onto the block itself
    int face_kind[6];  // W,E,N,S,F,B: FACE_*
};

// source of a linked face: the neighbour's arrays, its extent along the face normal and its
// strides (a neighbour has the block's tangential extents, but its rows and planes are as long as
// its own extents make them); unused for the other kinds
struct HaloLink {
    const double* A;
    const double* B;
    int           n;
    int           sx, sy;
    long          P;
};

struct MlpPacked;  // ns_mlp.h

// Network samples of the learned weights, emitted by the single-block component-wise face tiles
// and evaluated by a separate network kernel.  Two sets (0: the x and y sweeps, which run one
// after the other on one stream; 1: the z sweep, which may run beside them), each with two sides
// (0: left-biased reconstructions, 1: right-biased).  A side holds `cap` samples: planes
// [5][cap] of the build's real type (the stencil in left-biased order) and the flux index of
// each sample.  A launch splits its workgroups over
// LS_REGIONS regions (workgroup id modulo LS_REGIONS), each with its own counter and a fixed
// share of `cap`, so no counter serialises the whole launch.
constexpr int LS_REGIONS = 64;
constexpr int LS_STRIDE  = 32;  // counters one 128-byte line apart
struct LearnSink {
    void*     v[2][2];     // [set][side]: samples, planes [5][cap] of real
    unsigned* dest[2][2];  // [set][side]: flux index of each sample
    unsigned* cnt[2][2];   // [set][parity]: counters [side][LS_REGIONS] at stride LS_STRIDE
    long      cap;         // samples per side and set
    // the counters of a set alternate between two arrays from one launch to the next: a launch
    // counts in cnt[set][parity[set]] and its first workgroup zeroes the other array, which the
    // previous launch's network pass has finished reading (one stream per set)
    mutable int parity[2];
};

struct Phys {
    double gamma, mu, Pr, cfl;
    double lamb;        // (-2./3.) * mu            (centered_diffusion.cpp:101)
    double two_mu;      // 2.0 * mu                 (diffusion_scheme.cpp:5, :27)
    double heat;        // gamma / (Pr * (gamma - 1.0))  (centered_diffusion.cpp:161)
    int    conv;        // 0 none, 1 component-wise, 2 characteristic
    int    weights;     // 0 Shu, 1 Borges, 2 learned (network `mlp`)
    int    alpha_mode;  // 0 alpha = 3 (the reference), 1 per-face maximum eigenvalue
    int    physics;     // bit 0: corrected characteristic projection (true c, order, Roe state);
                        // bit 1: full 3-D viscous terms
    int    diff_order;  // 0, 2, 4, 6
    int    suth;        // Sutherland viscosity
    double suth_C;      // Sutherland constant (T0 = 1)
    double suth_k;      // mu * (1.0 + C)
    const MlpPacked* mlp;  // packed network of the learned weights (device), or null
    const LearnSink* sink; // host-side descriptor of the sample buffers of the single-block
                           // learned face tiles, or null (the tiles then evaluate the network
                           // themselves)
};

// SSP-RK3 stage (time_scheme.cpp:37-63); the coefficients are formed on the device from dt
struct StageArgs {
    const double* Un;     // step-start state (interior)
    const double* Ucur;   // stage input state (halo-filled)
    double*       Uout;   // stage output (interior)
    double*       Wout;   // c2p(Uout) (interior)
    int           stage;  // 0, 1, 2
    unsigned*     flags;  // health of the stage output (may be null): bit 0 a non-finite value,
                          // bit 1 rho <= 0 or p <= 0; bits are only ever set (atomic OR)
    // the next step's prologue, formed by the last stage of a step that another follows (null:
    // none): U = p2c(Wout) into Unext, c2p(U) into Wnext, the CFL minimum of Wout into dtnext
    // (atomic); Uout is then not written (the next step starts from Wout, as the reference's
    // does, time_scheme.cpp:15-27)
    double*             Unext = nullptr;
    double*             Wnext = nullptr;
    unsigned long long* dtnext = nullptr;
    // single-precision build: float copies of what the next stage's face sweeps read (Uout and
    // Wout, or Unext and Wnext), interior cells; required for every stage of that build's
    // single-block stage kernel
    float* Uf = nullptr;
    float* Wf = nullptr;
};

// float copies of a stage's U and primitives W (single-precision build), halos included, in the
// element layout of the double arrays (P, sy, sx); the face kernels of that build convert every
// loaded value to float, so they read these in place of the double arrays (same fluxes, half the
// bytes)
struct StateCopy {
    float* U;
    float* W;
};

// ------------------------------------------------------------------------------------------
// ensembles: independent blocks of different sizes advanced by batched launches.  Each kernel
// family runs once for all blocks; workgroup w of a launch belongs to the block b with
// start[fam][b] <= w < start[fam][b + 1] and is tile (w - start) of that block's local grid.
// ------------------------------------------------------------------------------------------
enum EnsFam {
    EF_STEP_BEGIN, EF_HALO, EF_FLUX_X, EF_FLUX_Y, EF_FLUX_Z, EF_PLANE_X, EF_PLANE_Y, EF_PLANE_Z,
    EF_DERIV1, EF_DERIV_HALO, EF_RESIDUAL, EF_COUNT
};

// tile shape of one block: lanes along k of the face-flux waves (64 / kw_flux rows per wave),
// lanes along k of the 256-thread cell kernels (256 / kw_cell rows), planes per marching chunk
struct EnsTiles {
    int kw_flux, kw_cell, march;
};

struct EnsBlock {
    Geo   g;
    Bnd   b;
    Phys  ph;
    double* W;       // primitive state: read by the step prologue, written by the last stage
    double* Un;      // step-start conservative state
    double* Ua;      // stage states
    double* Ub;
    double* Ws[2];   // stage primitives
    double* F[3];    // face fluxes
    double* Dx;      // first derivatives
    double* Dy;
    unsigned long long* dtmin;
    double*  dtime;
    unsigned* flags;  // health word of the ensemble step (shared by all blocks)
    float*   Uc;      // float copies of the stage U and W (single-precision build, StateCopy)
    float*   Wc;
    int      nbr[6];  // member index of the neighbour behind each linked face, -1 elsewhere
    int      linked;  // the block has linked faces: the separate derivative path, exchanges
    EnsTiles t;
    int      grid[EF_COUNT][3];  // local grid of each family (x fastest); 0 = no work
};

struct EnsTable {
    const EnsBlock* blk;    // device array [n]
    const int*      start;  // device array [EF_COUNT][n + 1] of workgroup prefix sums
    int             n;
    int             copies;  // the sweeps read the blocks' float copies (every block has them
                             // and takes the fused residual, which writes them)
};

inline int ens_cdiv(int a, int b) { return (a + b - 1) / b; }

// local grid of every family for one block (must match the index decode of the kernel bodies)
inline void ens_grids(EnsBlock& e) {
    const Geo& g  = e.g;
    const int  kf = e.t.kw_flux, kc = e.t.kw_cell;
    auto set = [&](int f, int x, int y, int z) {
        e.grid[f][0] = x, e.grid[f][1] = y, e.grid[f][2] = z;
    };
    long cells = (long)g.nx * g.ny * g.nz;
    set(EF_STEP_BEGIN, (int)((cells + 2047) / 2048), 1, 1);
    int m1 = g.ny > g.nx ? g.ny : g.nx;
    int mx = g.nz > 3 * g.ny ? g.nz : 3 * g.ny;
    set(EF_HALO, ens_cdiv(mx, 64), 3 * ens_cdiv(m1, 4), 6);
    const bool conv = e.ph.conv != 0;
    set(EF_FLUX_X, conv ? ens_cdiv(g.nz, kf) : 0, ens_cdiv(g.ny, 64 / kf), ens_cdiv(g.nx, 8));
    set(EF_FLUX_Y, conv ? ens_cdiv(g.nz, kf) : 0, ens_cdiv(g.nx, 64 / kf), ens_cdiv(g.ny, 8));
    set(EF_FLUX_Z, (conv && g.dim == 3) ? ens_cdiv(g.nz, kf) : 0, ens_cdiv(g.ny, 8 * (64 / kf)),
        g.nx);
    set(EF_PLANE_X, conv ? ens_cdiv(g.nz, 64) : 0, ens_cdiv(g.ny, 4), 1);
    set(EF_PLANE_Y, conv ? ens_cdiv(g.nz, 64) : 0, ens_cdiv(g.nx, 4), 1);
    set(EF_PLANE_Z, (conv && g.dim == 3) ? ens_cdiv(g.ny, 64) : 0, ens_cdiv(g.nx, 4), 1);
    const bool diff = e.ph.diff_order != 0;
    set(EF_DERIV1, diff ? ens_cdiv(g.nz, kc) : 0, ens_cdiv(g.ny, 256 / kc),
        ens_cdiv(g.nx, e.t.march));
    set(EF_DERIV_HALO, diff ? ens_cdiv(g.nz, 64) : 0, ens_cdiv(m1, 4), 4 * NGC);
    set(EF_RESIDUAL, ens_cdiv(g.nz, kc), ens_cdiv(g.ny, 256 / kc), ens_cdiv(g.nx, e.t.march));
}

// kernel launchers of one math mode (EXACT or FAST); all enqueue on `s`
struct Launchers {
    void (*reset_min)(unsigned long long* bits, hipStream_t s);
    // dt from W, U = p2c(W) -> Ua and Un, Wst = c2p(U) (interior)
    void (*step_begin)(const double* W, double* Ua, double* Un, double* Wst,
                       unsigned long long* dtmin, Geo g, Phys ph, hipStream_t s,
                       const StateCopy* cp);  // cp: float copies of Un and Wst (or null)
    // boundary conditions + periodic exchange on nA planes of A and nB planes of B (B may be null)
    void (*halo)(double* A, int nA, double* B, int nB, Geo g, Bnd b, hipStream_t s);
    void (*c2p_faces)(const double* U, double* W, Geo g, Phys ph, hipStream_t s);
    // parts: 1 the tile kernel (all faces but the last plane), 2 the last face plane, 3 both;
    // [lo, hi): the faces of the x sweep (the whole block: [0, nx + 1)), the planes i of the y
    // and z sweeps (the whole block: [0, nx); lo a multiple of 4)
    // cp (single-precision build, or null): float copies of U and W that the kernels read instead
    void (*face_flux)(int dir, const double* U, const double* W, double* f, Geo g, Phys ph,
                      int parts, int lo, int hi, hipStream_t s, const StateCopy* cp);
    // planes i in [lo, hi) (lo a multiple of the march length, 16).  Dy null: dW/dx of u and v
    // (Dx planes 2, 3) only (for residual_fd)
    void (*deriv1)(const double* W, double* Dx, double* Dy, Geo g, Phys ph, int lo, int hi,
                   hipStream_t s);
    // Dy null: the j halos of dW/dx of u and v on planes [lo, hi); else all halos (lo, hi unused)
    void (*deriv_halo)(double* Dx, double* Dy, Geo g, Bnd b, int lo, int hi, hipStream_t s);
    // R (interior) = R_in (or 0) + convection(fx,fy,fz) + diffusion(W, Dx, Dy); then either
    // stores R (st == nullptr) or applies the SSP-RK3 stage update.  With the full 3-D viscous
    // terms (physics bit 1) Dx is the 4-plane output of visc3d and Dy is unused.
    void (*residual)(const double* fx, const double* fy, const double* fz, const double* W,
                     const double* Dx, const double* Dy, const double* Rin, double* Rout,
                     const StageArgs* st, const unsigned long long* dtmin, int which, Geo g,
                     Phys ph, hipStream_t s);
    void (*aos_to_soa)(const double* aos, double* soa, Geo g, hipStream_t s);
    void (*soa_to_aos)(const double* soa, double* aos, Geo g, hipStream_t s);
    void (*reduce)(const double* U, const double* W, double* partial, int nblocks_max,
                   unsigned long long* dtmin, unsigned* flags, Geo g, Phys ph, hipStream_t s,
                   int* nblocks_out);
    void (*reduce_final)(const double* partial, int nblocks, double* sums, hipStream_t s);
    void (*time_accum)(double* t, const unsigned long long* dtmin, double cfl, hipStream_t s);
    void (*zero_flags)(unsigned* flags, hipStream_t s);
    // counts the operand pairs where the division routine of the kernels differs from `/`
    void (*selftest_div)(long n, unsigned long long seed, unsigned long long* count, hipStream_t s);
    // one SSP-RK3 step of every block of an ensemble; total[f] = workgroups of family f; the
    // template choices (convection, weights, diffusion order, viscosity law) come from `ph`
    void (*ens_step)(const EnsTable& t, const int* total, const Phys& ph, hipStream_t s);
    // the learned reconstruction of the face-flux kernels on n stencils S[n][5] (side 0 left-,
    // 1 right-biased): out[n] and the regime of each stencil (0 smooth, 1 network, 2 flagged)
    void (*mlp_recon)(const MlpPacked* P, long n, const double* S, int side, double* out,
                      int* regime, hipStream_t s);
    // overwrite the step's CFL minimum with a prescribed dt / CFL
    void (*set_dt)(unsigned long long* bits, double dt_over_cfl, hipStream_t s);
    // full 3-D viscous terms (physics bit 1): first derivatives of (u, v, w, p/rho) into Dx, Dy,
    // Dz (4 planes each) with their periodic halos, then the terms into Rv (energy, x, y, z
    // momentum), which the residual reads in place of Dx
    void (*visc3d)(const double* W, double* Dx, double* Dy, double* Dz, double* Rv, Geo g,
                   Phys ph, Bnd b, hipStream_t s);
    // U = p2c(W) on the interior and the face halos
    void (*p2c_faces)(const double* W, double* U, Geo g, Phys ph, hipStream_t s);
    // residual as above for the 2-D viscous terms (diff_order > 0, physics bit 1 clear) with
    // their y derivatives formed in the same kernel: Dx from deriv1 / deriv_halo with Dy null
    // on planes [lo, hi) (lo a multiple of 16)
    void (*residual_fd)(const double* fx, const double* fy, const double* fz, const double* W,
                        const double* Dx, const double* Rin, double* Rout, const StageArgs* st,
                        const unsigned long long* dtmin, int which, Geo g, Phys ph, Bnd b,
                        int lo, int hi, hipStream_t s);
    int fuse_diff;  // 1: single blocks use residual_fd (ensembles take the same split)
    int slab_align; // slabs of a pipelined stage start on multiples of this many planes i
    // the stage's residual with the z face fluxes formed in the kernel (fz never materialised
    // but for the first face of each row segment): the segment width it runs with, 0 where it
    // does not cover the configuration (then the z sweep + residual_fd run instead)
    int (*zfd_kw)(Geo g, Phys ph);
    // residual_fd + stage update of planes [lo, hi) for the stage st (its Ucur and the
    // primitives W are the input state), writing only the first face of each segment into fz
    void (*residual_zfd)(const double* fx, const double* fy, double* fz, const double* W,
                         const double* Dx, const StageArgs* st, const unsigned long long* dtmin,
                         Geo g, Phys ph, Bnd b, int lo, int hi, hipStream_t s);
    // samples per side and set of the LearnSink the full-block sweeps of g need, and the bytes of
    // one sample value in this build
    long (*learn_cap)(Geo g, int* value_bytes);
    // halo as above on the faces f0 .. f0 + nf - 1 only (W, E, N, S, F, B = 0 .. 5): the faces
    // are disjoint and read interior cells only, so they may be filled in separate launches.
    // cp (or null): the same faces of the float copies of A and B (5 planes each)
    void (*halo_faces)(double* A, int nA, double* B, int nB, Geo g, Bnd b, int f0, int nf,
                       hipStream_t s, const StateCopy* cp);
};
// packs FP32 parameters (device) into the lane fragments of the network (ns_mlp.hip)
void mlp_pack(const float* params, MlpPacked* P, hipStream_t s);

namespace exact {
const Launchers& launchers();
}
namespace fast {
const Launchers& launchers();
}
namespace fp32 {
const Launchers& launchers();
}

}  // namespace nsgpu
