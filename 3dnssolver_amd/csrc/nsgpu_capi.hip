// nsgpu_capi.hip -- C ABI (include/nsgpu.h): context, device buffers, host<->device layout
// conversion and the orchestration of TimeScheme::step on the GPU.
//
// Step sequence (time_scheme.cpp:3-70), all enqueued on the context's stream, no host sync
// inside a call to nsgpu_step:
//   reset dt accumulator; step_begin: dt = CFL * min local dt(W), U = p2c(W), Uprev = U
//   3 x stage: halo(U) -> W = c2p(U) on interior + faces -> face fluxes x, y, (z)
//              -> first derivatives + derivative halos -> residual + stage update
//   (the last stage also writes W = c2p(U); time += dt on the device)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/nsgpu.h"
#include "ns_common.h"
#include "ns_mlp.h"

using namespace nsgpu;

namespace {

enum Timer { T_STEP_BEGIN, T_HALO, T_C2P, T_FLUX_X, T_FLUX_Y, T_FLUX_Z, T_DERIV, T_RESIDUAL,
             T_COUNT };
const char* kTimerNames[T_COUNT] = {"step_begin", "halo",   "c2p_faces", "face_flux_x",
                                    "face_flux_y", "face_flux_z", "deriv1", "residual_update"};

}  // namespace

struct nsgpu_ctx {
    nsgpu_config       cfg;
    int                device = 0;
    hipStream_t        stream = nullptr;
    // second stream of the step: the small launches (last face planes, derivative pass) run on it
    // beside the face tiles; fork / join events order it against `stream`
    hipStream_t        aux      = nullptr;
    hipEvent_t         ev_fork  = nullptr;
    hipEvent_t         ev_join  = nullptr;
    // pipelined stages: slab s covers planes [slab[s], slab[s + 1]); one event per slab
    static constexpr int MAX_SLABS = 8;
    int                nslab = 1;
    int                slab[MAX_SLABS + 1] = {};
    hipEvent_t         ev_slab[MAX_SLABS] = {};
    // the stage forms its z faces in the residual kernel where the launchers cover the case:
    // an option ($NSGPU_ZFD=1 at creation), bitwise the same as z sweep + residual and slower at
    // 256^3 (the fused kernel runs 2 waves per SIMD: 3.18 ms per stage against 1.27 + 1.68 ms)
    bool               zfd_on = false;
    Geo                g{};
    Bnd                b{};
    Phys               ph{};
    const Launchers*   L = nullptr;
    size_t             plane_bytes = 0;  // one padded plane
    size_t             arr_bytes   = 0;  // one 5-variable array (4-variable arrays take as much)
    // device buffers (SoA)
    double*            W  = nullptr;  // primitive state (5 planes)
    double*            W2 = nullptr;  // next primitive state (written by the last stage)
    double*            Ws[2] = {nullptr, nullptr};  // stage primitives (ping-pong)
    double*            Un = nullptr;  // step-start conservative state
    double*            Ua = nullptr;  // stage states (ping-pong)
    double*            Ub = nullptr;
    double*            F[3] = {nullptr, nullptr, nullptr};  // face fluxes per direction
    double*            Dx = nullptr;  // first derivatives of rho, p, u, v (4 planes each)
    double*            Dy = nullptr;
    double*            Dz = nullptr;  // full 3-D viscous terms: d/dz of u, v, w, p/rho
    double*            Rv = nullptr;  // full 3-D viscous terms per cell (4 planes)
    double*            R  = nullptr;  // residual (operator-level calls)
    double*            stage_aos = nullptr;  // host-layout staging (5 planes)
    double*            partial = nullptr;    // reduction partials
    int                partial_max = 0;
    unsigned long long* dtmin = nullptr;  // [0] step dt accumulator, [1] reduce accumulator
    unsigned*          flags = nullptr;
    unsigned*          sflags = nullptr;  // health of the stage outputs of the current step call
    double*            dtime = nullptr;   // device time accumulator
    double*            sums  = nullptr;   // [5]
    double*            Ulast = nullptr;   // buffer holding the last step's U
    double             dt_scale = 0.0;    // dt = dtmin[0] bits * dt_scale: the CFL, or 1 after a
                                          // prescribed-dt step (dtmin[0] then holds dt itself)
    int                enrolled = 0;      // ensembles this context is a member of
    int                linked   = 0;      // a block of a decomposition (nsgpu_domain): faces
                                          // linked to its neighbours; stepped by its domain only
    float*             mlp_params = nullptr;  // learned weights: FP32 parameters
    MlpPacked*         mlp_packed = nullptr;  // and their packed BF16 fragments
    LearnSink          sink{};                // sample buffers of the learned component-wise
    void*              sink_mem   = nullptr;  // sweeps (emit mode), allocated at the first
    bool               sink_tried = false;    // single-block learned sweep (ensure_sink)
    // single-precision build: float copies of the stage state U and its primitives W, halos
    // included, written by the stage kernel (and the step prologue) beside the double arrays
    // and read by the next stage's face sweeps (use_copies)
    StateCopy          cp{};
    void*              cp_mem  = nullptr;
    bool               cp_read = true;  // $NSGPU_STATE_COPY=0: the sweeps read the double arrays
    bool               timing = false;
    hipEvent_t         ev[2 * T_COUNT] = {};
    double             ms[T_COUNT]     = {};
    long               launches[T_COUNT] = {};
};

namespace {

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            fprintf(stderr, "nsgpu: %s failed: %s\n", #x, hipGetErrorString(e_));             \
            return NSGPU_ERR_HIP;                                                              \
        }                                                                                      \
    } while (0)

struct Scope {  // times one kernel family when timing is enabled
    nsgpu_ctx* c;
    int        t;
    Scope(nsgpu_ctx* c_, int t_) : c(c_), t(t_) {
        if (c->timing) hipEventRecord(c->ev[2 * t], c->stream);
    }
    ~Scope() {
        if (!c->timing) return;
        hipEventRecord(c->ev[2 * t + 1], c->stream);
        hipEventSynchronize(c->ev[2 * t + 1]);
        float m = 0.f;
        hipEventElapsedTime(&m, c->ev[2 * t], c->ev[2 * t + 1]);
        c->ms[t] += m;
        c->launches[t] += 1;
    }
};

int validate(const nsgpu_config* c) {
    if (!c) return NSGPU_ERR_INVALID;
    if (c->ngc != NGC) return NSGPU_ERR_INVALID;
    // a 2-D block (the z sweep skipped, weno-5.cpp:28) may be 1 or 2 cells thick in z, as the
    // reference's 2-D inputs are (Nz = 1)
    if (c->nx < 3 || c->ny < 3 || c->nz < (c->dim == 2 ? 1 : 3)) return NSGPU_ERR_INVALID;
    if (c->dim != 2 && c->dim != 3) return NSGPU_ERR_INVALID;
    if (!(c->Lx > 0) || !(c->Ly > 0) || !(c->Lz > 0)) return NSGPU_ERR_INVALID;
    if (!(c->gamma > 1.0) || !(c->cfl > 0)) return NSGPU_ERR_INVALID;
    if (c->conv < NSGPU_CONV_NONE || c->conv > NSGPU_CONV_WENO5CW) return NSGPU_ERR_INVALID;
    if (c->weights != NSGPU_WEIGHTS_SHU && c->weights != NSGPU_WEIGHTS_BORGES &&
        c->weights != NSGPU_WEIGHTS_LEARNED)
        return NSGPU_ERR_INVALID;
    if (c->diff_order != 0 && c->diff_order != 2 && c->diff_order != 4 && c->diff_order != 6)
        return NSGPU_ERR_INVALID;
    if (c->math != NSGPU_MATH_EXACT && c->math != NSGPU_MATH_FAST && c->math != NSGPU_MATH_FP32)
        return NSGPU_ERR_INVALID;
    if (c->viscosity != NSGPU_VISC_CONSTANT && c->viscosity != NSGPU_VISC_SUTHERLAND)
        return NSGPU_ERR_INVALID;
    if (c->viscosity == NSGPU_VISC_SUTHERLAND && !(c->suth_C > 0)) return NSGPU_ERR_INVALID;
    if (c->alpha_mode != NSGPU_ALPHA_CONST3 && c->alpha_mode != NSGPU_ALPHA_LOCAL_MAX)
        return NSGPU_ERR_INVALID;
    if (c->physics & ~(NSGPU_PHYS_CONVECTION | NSGPU_PHYS_VISCOUS)) return NSGPU_ERR_INVALID;
    if ((c->physics & NSGPU_PHYS_VISCOUS) && c->dim != 3) return NSGPU_ERR_INVALID;
    if (c->weights == NSGPU_WEIGHTS_LEARNED && (c->alpha_mode || c->physics))
        return NSGPU_ERR_UNSUPPORTED;
    // periodicity per axis from West / South / Front (decomposition.cpp:145-161); every face of a
    // non-periodic axis must be Neumann (domain.cpp:87-137 leaves other kinds without a boundary)
    static const int lead[3] = {0, 3, 4};
    static const int pair[3][2] = {{0, 1}, {2, 3}, {4, 5}};
    for (int a = 0; a < 3; a++) {
        if (c->bc[lead[a]] == NSGPU_BC_PERIODIC) continue;
        for (int f : pair[a])
            if (c->bc[f] != NSGPU_BC_NEUMANN) return NSGPU_ERR_INVALID;
    }
    long P = (long)(c->nx + 6) * (c->ny + 6) * (c->nz + 6);
    if (P * 5 > (1L << 40)) return NSGPU_ERR_INVALID;
    return NSGPU_OK;
}

// zero-filled device buffer.  The fill is ordered on the context's stream: that stream is
// non-blocking, so a fill on the null stream could still be running when the first upload on
// the context's stream writes the buffer (nsgpu_create synchronises the stream before returning).
int alloc(double** p, size_t bytes, hipStream_t s) {
    HIPCHK(hipMalloc((void**)p, bytes));
    HIPCHK(hipMemsetAsync(*p, 0, bytes, s));
    return NSGPU_OK;
}

// a padded-layout array of `bytes`: KOFF doubles into a zero-filled allocation, so the first
// interior cell of each row is 128-byte aligned (hipMalloc returns at least 256-byte alignment);
// free with free_array
int alloc_array(double** p, size_t bytes, hipStream_t s) {
    double* raw = nullptr;
    int     rc  = alloc(&raw, bytes + KOFF * sizeof(double), s);
    if (rc) return rc;
    *p = raw + KOFF;
    return NSGPU_OK;
}
void free_array(double* p) {
    if (p) hipFree(p - KOFF);
}

// doubles of one 5-variable array in the reference's dense host layout
size_t host_len(const Geo& g) { return (size_t)g.NX * g.NY * g.NZ * 5; }

int upload_soa(nsgpu_ctx* c, const double* aos, double* dst) {
    HIPCHK(hipMemcpyAsync(c->stage_aos, aos, host_len(c->g) * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
    c->L->aos_to_soa(c->stage_aos, dst, c->g, c->stream);
    return NSGPU_OK;
}

int download_soa(nsgpu_ctx* c, const double* src, double* aos) {
    c->L->soa_to_aos(src, c->stage_aos, c->g, c->stream);
    HIPCHK(hipMemcpyAsync(aos, c->stage_aos, host_len(c->g) * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

// grid spacings and the quantities derived from them
void set_spacing(Geo& g, double dx, double dy, double dz) {
    g.dx = dx, g.dy = dy, g.dz = dz;
    g.dxx = g.dx * g.dx;  // CDN::ddf_dxdx / ddf_dydy divisors (cd-n.cpp:67-79)
    g.dyy = g.dy * g.dy;
    g.idx = 1.0 / g.dx, g.idy = 1.0 / g.dy, g.idz = 1.0 / g.dz;
    g.idxx = 1.0 / g.dxx, g.idyy = 1.0 / g.dyy, g.idzz = 1.0 / (g.dz * g.dz);
    g.square_xy = (g.dxx == g.dyy);
}

// status of a health word (bit 0 non-finite, bit 1 positivity, bit 2 alpha)
int health_status(unsigned fl) {
    if (fl & 1u) return NSGPU_ERR_NONFINITE;
    if (fl & 2u) return NSGPU_ERR_POSITIVITY;
    if (fl & 4u) return NSGPU_ERR_ALPHA;
    return NSGPU_OK;
}

// the learned weights need their network
bool net_ready(const nsgpu_ctx* c) { return c->ph.weights != NSGPU_WEIGHTS_LEARNED || c->ph.mlp; }

// Sample buffers of the learned component-wise sweeps (LearnSink): the face tiles then leave the
// network to a separate dense kernel instead of evaluating it themselves.  Sized for the worst
// case (every reconstruction of every face in the network regime), allocated once at the first
// single-block learned sweep of the context, and only up to SINK_MAX_BYTES; blocks of a
// decomposition (stepped inside captured graphs) and contexts without them keep the in-kernel
// network.  Costs at 256^3: 8.1 GB with the FP32 build, 14.8 GB in FP64.
constexpr size_t SINK_MAX_BYTES = 32ull << 30;
void ensure_sink(nsgpu_ctx* c) {
    if (c->sink_tried || c->ph.weights != NSGPU_WEIGHTS_LEARNED || c->ph.conv != 1 || c->linked)
        return;
    c->sink_tried = true;
    if (const char* e = std::getenv("NSGPU_LEARN_SINK"))  // 0: no buffers (in-kernel network)
        if (std::atoi(e) == 0) return;
    int        vb  = 8;
    const long cap = c->L->learn_cap(c->g, &vb);
    const size_t vbytes = ((size_t)cap * 5 * vb + 255) / 256 * 256;
    const size_t dbytes = ((size_t)cap * 4 + 255) / 256 * 256;
    const size_t cbytes = (size_t)2 * LS_REGIONS * LS_STRIDE * sizeof(unsigned);
    const size_t total  = 4 * (vbytes + dbytes) + 4 * cbytes;
    if (total > SINK_MAX_BYTES) return;
    char* m = nullptr;
    if (hipMalloc((void**)&m, total) != hipSuccess) {
        (void)hipGetLastError();  // the in-kernel network stays in use
        return;
    }
    c->sink_mem = m;
    for (int set = 0; set < 2; set++) {
        for (int side = 0; side < 2; side++) {
            c->sink.v[set][side] = m, m += vbytes;
            c->sink.dest[set][side] = (unsigned*)m, m += dbytes;
        }
        c->sink.cnt[set][0] = (unsigned*)m, m += cbytes;
        c->sink.cnt[set][1] = (unsigned*)m, m += cbytes;
    }
    // every counter starts at zero; from then on each emit launch zeroes the next launch's
    for (int set = 0; set < 2; set++) {  // (the two arrays of a set are adjacent)
        if (hipMemsetAsync(c->sink.cnt[set][0], 0, 2 * cbytes, c->stream) != hipSuccess) {
            (void)hipGetLastError();
            hipFree(c->sink_mem);
            c->sink_mem = nullptr;
            return;
        }
    }
    c->sink.parity[0] = c->sink.parity[1] = 0;
    c->sink.cap = cap;
    c->ph.sink  = &c->sink;
}

// the whole block as a plane range of the launchers: faces [0, nx + 1) of the x sweep, planes
// [0, nx) of the others
inline int full_hi(const nsgpu_ctx* c, int dir) { return dir == 0 ? c->g.nx + 1 : c->g.nx; }

// convection face fluxes of the current (U, W); fz stays zero for dim == 2
// cp: float copies of (U, W) the sweeps read instead (or null)
void convection_fluxes(nsgpu_ctx* c, const double* U, const double* W, int parts = 3,
                       hipStream_t st = nullptr, bool skip_z = false,
                       const StateCopy* cp = nullptr) {
    if (!c->ph.conv) return;
    ensure_sink(c);
    if (!st) st = c->stream;
    {
        Scope s(c, T_FLUX_X);
        c->L->face_flux(0, U, W, c->F[0], c->g, c->ph, parts, 0, full_hi(c, 0), st, cp);
    }
    {
        Scope s(c, T_FLUX_Y);
        c->L->face_flux(1, U, W, c->F[1], c->g, c->ph, parts, 0, full_hi(c, 1), st, cp);
    }
    if (c->g.dim == 3 && !skip_z) {
        Scope s(c, T_FLUX_Z);
        c->L->face_flux(2, U, W, c->F[2], c->g, c->ph, parts, 0, full_hi(c, 2), st, cp);
    }
}

// the residual kernel forms the y derivatives of the 2-D viscous terms itself
bool fused_diffusion(const nsgpu_ctx* c) {
    return c->L->fuse_diff && c->ph.diff_order && !(c->ph.physics & NSGPU_PHYS_VISCOUS) &&
           (long)c->g.NX * c->g.sx < (1L << 31) && !c->linked;
}

// the stage's residual forms the z faces (k_residual_zfd): the z sweep is left out
bool zfd(const nsgpu_ctx* c) {
    return c->zfd_on && c->ph.conv && fused_diffusion(c) && c->L->zfd_kw(c->g, c->ph) > 0;
}

void diffusion_derivs(nsgpu_ctx* c, const double* W, hipStream_t st = nullptr) {
    if (!c->ph.diff_order) return;
    if (!st) st = c->stream;
    Scope s(c, T_DERIV);
    if (c->ph.physics & NSGPU_PHYS_VISCOUS) {
        c->L->visc3d(W, c->Dx, c->Dy, c->Dz, c->Rv, c->g, c->ph, c->b, st);
        return;
    }
    double* Dy = fused_diffusion(c) ? nullptr : c->Dy;
    c->L->deriv1(W, c->Dx, Dy, c->g, c->ph, 0, c->g.nx, st);
    c->L->deriv_halo(c->Dx, Dy, c->g, c->b, 0, c->g.nx, st);
}

// the inputs of a stage's residual: face fluxes and derivatives of (U, W), which all read U and W
// and write disjoint memory.  The x and y face tiles run on the context's stream; the z sweep,
// the last face planes of x and y and the derivative pass run on the auxiliary stream beside
// them (a z tile block fits beside an x/y tile block on one CU: 44 + 66.5 KB of LDS); the
// context's stream waits for both before the residual.  Timed runs keep everything on one
// stream so that each kernel family is timed alone.
// zsplit: the z-face halos of U and W are still to be filled; they go on the auxiliary stream,
// ahead of the z sweep (only it and the work after the join read them)
// cp: the float copies of (U, W) (use_copies), or null.  With copies only the sweeps read the z
// halos, and only from the copies: the z-face fill then covers the copies alone.
void stage_inputs(nsgpu_ctx* c, const double* U, const double* W, bool zsplit = false,
                  const StateCopy* cp = nullptr) {
    ensure_sink(c);
    const int nz5 = cp ? 0 : 5;  // planes of U and W whose z halos are filled
    if (c->timing || !c->aux) {
        if (zsplit)
            c->L->halo_faces((double*)U, nz5, (double*)W, nz5, c->g, c->b, 4, 2, c->stream, cp);
        convection_fluxes(c, U, W, 3, nullptr, zfd(c), cp);
        diffusion_derivs(c, W);
        return;
    }
    hipEventRecord(c->ev_fork, c->stream);
    hipStreamWaitEvent(c->aux, c->ev_fork, 0);
    const Geo& g = c->g;
    if (zsplit) c->L->halo_faces((double*)U, nz5, (double*)W, nz5, g, c->b, 4, 2, c->aux, cp);
    if (c->ph.conv && g.dim == 3 && !zfd(c))
        c->L->face_flux(2, U, W, c->F[2], g, c->ph, 3, 0, full_hi(c, 2), c->aux, cp);
    if (c->ph.conv) {
        c->L->face_flux(0, U, W, c->F[0], g, c->ph, 2, 0, full_hi(c, 0), c->aux, cp);
        c->L->face_flux(1, U, W, c->F[1], g, c->ph, 2, 0, full_hi(c, 1), c->aux, cp);
    }
    diffusion_derivs(c, W, c->aux);
    hipEventRecord(c->ev_join, c->aux);
    if (c->ph.conv) {
        c->L->face_flux(0, U, W, c->F[0], g, c->ph, 1, 0, full_hi(c, 0), c->stream, cp);
        c->L->face_flux(1, U, W, c->F[1], g, c->ph, 1, 0, full_hi(c, 1), c->stream, cp);
    }
    hipStreamWaitEvent(c->stream, c->ev_join, 0);
}

// Slabs of the pipelined stage: $NSGPU_STAGE_SLABS of them (at most MAX_SLABS), every slab
// starting on a multiple of the cell kernels' march length.  Off (one slab) by default: at 256^3
// the overlap gains nothing measurable (15.92 ms per step with 4 slabs, 16.38 with 8, 15.84
// without), the sweeps and the residual do not run faster side by side than one after another.
void plan_slabs(nsgpu_ctx* c) {
    const int nx = c->g.nx, al = c->L->slab_align;
    int       n  = 1;
    if (const char* e = std::getenv("NSGPU_STAGE_SLABS")) n = std::atoi(e);
    n = std::max(1, std::min(n, nsgpu_ctx::MAX_SLABS));
    int k = 0;
    c->slab[0] = 0;
    for (int s = 1; s < n; s++) {
        const int b = (int)((long)nx * s / n) / al * al;
        if (b > c->slab[k] && b < nx) c->slab[++k] = b;
    }
    c->slab[++k] = nx;
    c->nslab     = k;
}

bool pipelined(const nsgpu_ctx* c) {
    return c->nslab > 1 && c->aux && !c->timing && c->ph.conv && fused_diffusion(c);
}

// a context whose sweeps may read its float copies: the single-precision build's fused stage
// kernel writes them for every stage, and the sweep kernels without options read them (the
// per-face alpha and the corrected projection read the double arrays, whose halos then have to
// be filled in full)
bool copies_ok(const nsgpu_ctx* c) {
    const bool opt = c->ph.alpha_mode == NSGPU_ALPHA_LOCAL_MAX ||
                     (c->ph.conv == 2 && (c->ph.physics & NSGPU_PHYS_CONVECTION));
    return c->cp.U && c->cp_read && fused_diffusion(c) && !opt;
}

// the face sweeps of a step read the float copies (not with the z faces formed in the residual,
// which reads the double arrays; a pipelined stage's residual writes the copies while the next
// slab's sweeps still read, so those read the double arrays)
bool use_copies(const nsgpu_ctx* c) { return copies_ok(c) && !zfd(c) && !pipelined(c); }

// One stage in slabs of planes i.  The face sweeps and the derivative pass of slab s run on the
// context's stream; the residual + stage update of slab s runs on the auxiliary stream as soon as
// they are done, beside the sweeps of slab s + 1: the HBM-bound residual overlaps the FP64-bound
// sweeps.  The residual of cell i reads the x faces i and i + 1 (face f lies between cells f - 1
// and f), so slab [lo, hi) computes the x faces [lo + 1, hi + 1) (the first slab also face 0),
// and every face a slab's residual reads is computed by that slab or an earlier one; y and z
// faces, derivatives and their j halos are plane-local.  The streams write disjoint memory: the
// sweeps write faces and derivative planes of later slabs, the residual writes the stage output.
// The context's stream waits for the last residual before the stage ends.
void stage_pipelined(nsgpu_ctx* c, const double* U, const double* W, const StageArgs& a) {
    ensure_sink(c);
    const Launchers& L = *c->L;
    const Geo&       g = c->g;
    for (int sl = 0; sl < c->nslab; sl++) {
        const int lo = c->slab[sl], hi = c->slab[sl + 1];
        L.face_flux(0, U, W, c->F[0], g, c->ph, 3, sl == 0 ? 0 : lo + 1, hi + 1, c->stream,
                    nullptr);
        L.face_flux(1, U, W, c->F[1], g, c->ph, 3, lo, hi, c->stream, nullptr);
        if (g.dim == 3 && !zfd(c))
            L.face_flux(2, U, W, c->F[2], g, c->ph, 3, lo, hi, c->stream, nullptr);
        L.deriv1(W, c->Dx, nullptr, g, c->ph, lo, hi, c->stream);
        L.deriv_halo(c->Dx, nullptr, g, c->b, lo, hi, c->stream);
        hipEventRecord(c->ev_slab[sl], c->stream);
        hipStreamWaitEvent(c->aux, c->ev_slab[sl], 0);
        if (zfd(c))
            L.residual_zfd(c->F[0], c->F[1], c->F[2], W, c->Dx, &a, &c->dtmin[0], g, c->ph, c->b,
                           lo, hi, c->aux);
        else
            L.residual_fd(c->F[0], c->F[1], c->F[2], W, c->Dx, nullptr, nullptr, &a,
                          &c->dtmin[0], NSGPU_RHS_BOTH, g, c->ph, c->b, lo, hi, c->aux);
    }
    hipEventRecord(c->ev_join, c->aux);
    hipStreamWaitEvent(c->stream, c->ev_join, 0);
}

// what the residual reads as its viscous input: the derivative planes, or the precomputed full
// 3-D viscous terms
const double* visc_in(const nsgpu_ctx* c) {
    return (c->ph.physics & NSGPU_PHYS_VISCOUS) ? c->Rv : c->Dx;
}

// residual (+ stage update) of the single block on the face fluxes F and the primitives W
void residual(nsgpu_ctx* c, const double* W, const double* Rin, double* Rout, const StageArgs* a,
              const unsigned long long* dtmin, int which) {
    if (a && !Rin && !Rout && which == NSGPU_RHS_BOTH && zfd(c))  // a stage (stage_inputs)
        c->L->residual_zfd(c->F[0], c->F[1], c->F[2], W, c->Dx, a, dtmin, c->g, c->ph, c->b, 0,
                           c->g.nx, c->stream);
    else if ((which & NSGPU_RHS_DIFFUSION) && fused_diffusion(c))
        c->L->residual_fd(c->F[0], c->F[1], c->F[2], W, c->Dx, Rin, Rout, a, dtmin, which, c->g,
                          c->ph, c->b, 0, c->g.nx, c->stream);
    else
        c->L->residual(c->F[0], c->F[1], c->F[2], W, visc_in(c), c->Dy, Rin, Rout, a, dtmin,
                       which, c->g, c->ph, c->stream);
}

}  // namespace

extern "C" {

const char* nsgpu_strerror(int s) {
    switch (s) {
    case NSGPU_OK: return "ok";
    case NSGPU_ERR_INVALID: return "invalid configuration or argument";
    case NSGPU_ERR_HIP: return "HIP runtime error";
    case NSGPU_ERR_NODEVICE: return "no GPU device";
    case NSGPU_ERR_NONFINITE: return "non-finite state";
    case NSGPU_ERR_POSITIVITY: return "non-positive density or pressure";
    case NSGPU_ERR_ALPHA: return "max |u|+c exceeds the Lax-Friedrichs constant 3";
    case NSGPU_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}

int nsgpu_create(const nsgpu_config* cfg, int device, nsgpu_ctx** out) {
    if (!out) return NSGPU_ERR_INVALID;
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NSGPU_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return NSGPU_ERR_NODEVICE;
    HIPCHK(hipSetDevice(device));
    nsgpu_ctx* c = new (std::nothrow) nsgpu_ctx();
    if (!c) return NSGPU_ERR_INVALID;
    c->cfg    = *cfg;
    c->device = device;
    Geo& g    = c->g;
    g.nx = cfg->nx, g.ny = cfg->ny, g.nz = cfg->nz;
    g.NX = g.nx + 2 * NGC, g.NY = g.ny + 2 * NGC, g.NZ = g.nz + 2 * NGC;
    const int row = (g.NZ + KALIGN - 1) / KALIGN * KALIGN;  // one padded row of one variable
    // Device layout of a 5-variable array: the five rows of one (i, j) side by side,
    // [i][j][q][k] (variable stride P = one row), so the values a cell needs from one array lie
    // in one 5-row stretch of memory; $NSGPU_LAYOUT=0 selects whole planes per variable,
    // [q][i][j][k] (P = one plane).  The kernels only see P, sx and sy.  At 256^3 the interleaved
    // rows take the residual kernel from 1.30 to 1.22 ms and the step from 14.7 to 14.5 ms.
    if (const char* e = std::getenv("NSGPU_LAYOUT"); e && std::atoi(e) == 0) {
        g.sy = row;
        g.sx = g.NY * g.sy;
        g.P  = (long)g.NX * g.sx;
    } else {
        g.P  = row;
        g.sy = 5 * row;
        g.sx = g.NY * g.sy;
    }
    // Grid::compute_step_sizes (grid.cpp:14-16)
    set_spacing(g, cfg->Lx / cfg->nx, cfg->Ly / cfg->ny, cfg->Lz / cfg->nz);
    g.dim = cfg->dim;
    static const int lead[3] = {0, 3, 4};
    for (int a = 0; a < 3; a++) c->b.periodic[a] = (cfg->bc[lead[a]] == NSGPU_BC_PERIODIC);
    for (int f = 0; f < 6; f++) c->b.face_kind[f] = cfg->bc[f];
    c->ph.gamma = cfg->gamma, c->ph.mu = cfg->mu, c->ph.Pr = cfg->Pr, c->ph.cfl = cfg->cfl;
    c->ph.conv = cfg->conv, c->ph.weights = cfg->weights;
    c->ph.alpha_mode = cfg->alpha_mode, c->ph.physics = cfg->physics;
    c->ph.mlp  = nullptr;  // set by nsgpu_set_mlp
    c->ph.diff_order = cfg->diff_order;
    c->ph.lamb       = (-2. / 3.) * cfg->mu;                           // centered_diffusion.cpp:101
    c->ph.two_mu     = 2.0 * cfg->mu;
    c->ph.heat       = cfg->gamma / (cfg->Pr * (cfg->gamma - 1.0));    // centered_diffusion.cpp:161
    c->ph.suth       = (cfg->viscosity == NSGPU_VISC_SUTHERLAND);
    c->ph.suth_C     = cfg->suth_C;
    c->ph.suth_k     = cfg->mu * (1.0 + cfg->suth_C);
    c->L = (cfg->math == NSGPU_MATH_FAST)   ? &fast::launchers()
           : (cfg->math == NSGPU_MATH_FP32) ? &fp32::launchers()
                                            : &exact::launchers();
    c->plane_bytes = (size_t)g.NX * g.NY * row * sizeof(double);
    c->arr_bytes   = 5 * c->plane_bytes;
    size_t five = c->arr_bytes, four = (g.P == row) ? five : 4 * c->plane_bytes;
    rc = NSGPU_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc) rc = alloc_array(&c->W, five, c->stream);
    if (!rc) rc = alloc_array(&c->W2, five, c->stream);
    for (int t = 0; t < 2 && !rc; t++) rc = alloc_array(&c->Ws[t], five, c->stream);
    if (!rc) rc = alloc_array(&c->Un, five, c->stream);
    if (!rc) rc = alloc_array(&c->Ua, five, c->stream);
    if (!rc) rc = alloc_array(&c->Ub, five, c->stream);
    for (int d = 0; d < 3 && !rc; d++) rc = alloc_array(&c->F[d], five, c->stream);
    if (!rc) rc = alloc_array(&c->Dx, four, c->stream);
    if (!rc) rc = alloc_array(&c->Dy, four, c->stream);
    if (!rc && (cfg->physics & NSGPU_PHYS_VISCOUS)) {
        rc = alloc_array(&c->Dz, four, c->stream);
        if (!rc) rc = alloc_array(&c->Rv, four, c->stream);
    }
    if (!rc) rc = alloc_array(&c->R, five, c->stream);
    if (!rc) rc = alloc_array(&c->stage_aos, five, c->stream);
    if (!rc && cfg->math == NSGPU_MATH_FP32) {
        // two arrays of `five / 8` floats, each starting so that its first interior cell of a row
        // (k = NGC) lies on a 128-byte line like the double arrays'
        const size_t n = (five / sizeof(double) + 31) / 32 * 32;
        rc = alloc((double**)&c->cp_mem, (2 * n + 64) * sizeof(float), c->stream);
        if (!rc) {
            c->cp.U = (float*)c->cp_mem + (32 - NGC);
            c->cp.W = c->cp.U + n;
        }
    }
    c->partial_max = 4096;
    if (!rc) rc = alloc(&c->partial, sizeof(double) * 5 * c->partial_max, c->stream);
    if (!rc) rc = alloc((double**)&c->dtmin, 3 * sizeof(unsigned long long), c->stream);
    if (!rc) rc = alloc((double**)&c->flags, sizeof(double), c->stream);
    if (!rc) rc = alloc((double**)&c->sflags, sizeof(double), c->stream);
    if (!rc) rc = alloc(&c->dtime, sizeof(double), c->stream);
    if (!rc) rc = alloc(&c->sums, 5 * sizeof(double), c->stream);
    if (!rc && cfg->weights == NSGPU_WEIGHTS_LEARNED) {
        rc = alloc((double**)&c->mlp_params, sizeof(float) * MLP_NPARAM, c->stream);
        if (!rc) rc = alloc((double**)&c->mlp_packed, sizeof(MlpPacked), c->stream);
    }
    for (int t = 0; t < 2 * T_COUNT && !rc; t++)
        if (hipEventCreate(&c->ev[t]) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc && hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    if (!rc && hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    for (int t = 0; t < nsgpu_ctx::MAX_SLABS && !rc; t++)
        if (hipEventCreateWithFlags(&c->ev_slab[t], hipEventDisableTiming) != hipSuccess)
            rc = NSGPU_ERR_HIP;
    if (!rc) plan_slabs(c);
    if (const char* e = std::getenv("NSGPU_ZFD")) c->zfd_on = std::atoi(e) != 0;
    if (const char* e = std::getenv("NSGPU_STATE_COPY")) c->cp_read = std::atoi(e) != 0;
    c->Ulast    = c->Ua;
    c->dt_scale = cfg->cfl;
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (rc) {
        nsgpu_destroy(c);
        return rc;
    }
    *out = c;
    return NSGPU_OK;
}

void nsgpu_destroy(nsgpu_ctx* c) {
    if (!c) return;
    if (c->enrolled > 0) {
        // an ensemble holds this context's device buffers in its block table and its graph
        fprintf(stderr, "nsgpu: nsgpu_destroy refused: the context is a member of %d live "
                        "ensemble(s); destroy them first\n", c->enrolled);
        return;
    }
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->aux) hipStreamSynchronize(c->aux);
    double* arrays[] = {c->W,     c->W2,   c->Ws[0], c->Ws[1], c->Un, c->Ua, c->Ub,
                        c->F[0],  c->F[1], c->F[2],  c->Dx,    c->Dy, c->Dz, c->Rv,
                        c->R,     c->stage_aos};
    for (double* p : arrays) free_array(p);
    double* bufs[] = {c->partial, (double*)c->dtmin, (double*)c->flags, (double*)c->sflags,
                      c->dtime, c->sums,
                      (double*)c->mlp_params, (double*)c->mlp_packed, (double*)c->sink_mem,
                      (double*)c->cp_mem};
    for (double* p : bufs)
        if (p) hipFree(p);
    for (int t = 0; t < 2 * T_COUNT; t++)
        if (c->ev[t]) hipEventDestroy(c->ev[t]);
    if (c->ev_fork) hipEventDestroy(c->ev_fork);
    if (c->ev_join) hipEventDestroy(c->ev_join);
    for (hipEvent_t e : c->ev_slab)
        if (e) hipEventDestroy(e);
    if (c->aux) hipStreamDestroy(c->aux);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

size_t nsgpu_array_len(const nsgpu_ctx* c) { return c ? host_len(c->g) : 0; }

void* nsgpu_stream(nsgpu_ctx* c) { return c ? (void*)c->stream : nullptr; }

int nsgpu_set_state(nsgpu_ctx* c, const double* aos, int kind) {
    if (!c || !aos) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    // the other form follows on the device, so get_state, nsgpu_reduce and the next step see
    // one consistent state whichever form was set (the step itself starts from W,
    // time_scheme.cpp:25)
    if (kind == NSGPU_STATE_PRIMITIVE) {
        int rc = upload_soa(c, aos, c->W);
        if (rc) return rc;
        c->L->p2c_faces(c->W, c->Ua, c->g, c->ph, c->stream);
        c->Ulast = c->Ua;
    } else if (kind == NSGPU_STATE_CONSERVATIVE) {
        int rc = upload_soa(c, aos, c->Ua);
        if (rc) return rc;
        c->L->c2p_faces(c->Ua, c->W, c->g, c->ph, c->stream);
        c->Ulast = c->Ua;
    } else {
        return NSGPU_ERR_INVALID;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

int nsgpu_get_state(nsgpu_ctx* c, double* aos, int kind) {
    if (!c || !aos) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    if (kind == NSGPU_STATE_PRIMITIVE) return download_soa(c, c->W, aos);
    if (kind == NSGPU_STATE_CONSERVATIVE) return download_soa(c, c->Ulast, aos);
    return NSGPU_ERR_INVALID;
}

int nsgpu_apply_boundary(nsgpu_ctx* c, double* aos) {
    if (!c || !aos) return NSGPU_ERR_INVALID;
    if (c->linked) return NSGPU_ERR_INVALID;  // a decomposition's block: through its domain
    HIPCHK(hipSetDevice(c->device));
    int rc = upload_soa(c, aos, c->R);
    if (rc) return rc;
    c->L->halo(c->R, 5, nullptr, 0, c->g, c->b, c->stream);
    return download_soa(c, c->R, aos);
}

// The operator-level calls below work in the stage scratch arrays (Ws[0] for W, Ub for U): a
// step rewrites both before reading them, and the resident state (W, Ulast = Ua between steps)
// stays as it was, so stepping, get_state and reduce are unaffected by them.
int nsgpu_convection_solve(nsgpu_ctx* c, const double* W_aos, const double* U_aos, double* R_aos) {
    if (!c || !W_aos || !U_aos || !R_aos || !net_ready(c)) return NSGPU_ERR_INVALID;
    if (c->linked) return NSGPU_ERR_INVALID;  // a decomposition's block: through its domain
    HIPCHK(hipSetDevice(c->device));
    double* Wt = c->Ws[0];
    int     rc = upload_soa(c, W_aos, Wt);
    if (!rc) rc = upload_soa(c, U_aos, c->Ub);
    if (!rc) rc = upload_soa(c, R_aos, c->R);
    if (rc) return rc;
    convection_fluxes(c, c->Ub, Wt);
    {
        Scope s(c, T_RESIDUAL);
        residual(c, Wt, c->R, c->R, nullptr, nullptr, NSGPU_RHS_CONVECTION);
    }
    return download_soa(c, c->R, R_aos);
}

int nsgpu_diffusion_solve(nsgpu_ctx* c, const double* W_aos, double* R_aos) {
    if (!c || !W_aos || !R_aos) return NSGPU_ERR_INVALID;
    if (c->linked) return NSGPU_ERR_INVALID;  // a decomposition's block: through its domain
    HIPCHK(hipSetDevice(c->device));
    double* Wt = c->Ws[0];
    int     rc = upload_soa(c, W_aos, Wt);
    if (!rc) rc = upload_soa(c, R_aos, c->R);
    if (rc) return rc;
    diffusion_derivs(c, Wt);
    {
        Scope s(c, T_RESIDUAL);
        residual(c, Wt, c->R, c->R, nullptr, nullptr, NSGPU_RHS_DIFFUSION);
    }
    return download_soa(c, c->R, R_aos);
}

int nsgpu_rhs(nsgpu_ctx* c, double* U_aos, double* W_out, double* R_out, int which) {
    if (!c || !U_aos || !R_out || which < 1 || which > 3 || !net_ready(c)) return NSGPU_ERR_INVALID;
    if (c->linked) return NSGPU_ERR_INVALID;  // a decomposition's block: through its domain
    HIPCHK(hipSetDevice(c->device));
    double* Ut = c->Ub;
    double* Wt = c->Ws[0];
    int     rc = upload_soa(c, U_aos, Ut);
    if (rc) return rc;
    {
        Scope s(c, T_HALO);
        c->L->halo(Ut, 5, nullptr, 0, c->g, c->b, c->stream);
    }
    {
        Scope s(c, T_C2P);
        c->L->c2p_faces(Ut, Wt, c->g, c->ph, c->stream);
    }
    if (which & 1) convection_fluxes(c, Ut, Wt);
    if (which & 2) diffusion_derivs(c, Wt);
    {
        Scope s(c, T_RESIDUAL);
        residual(c, Wt, nullptr, c->R, nullptr, nullptr, which);
    }
    rc = download_soa(c, Ut, U_aos);
    if (!rc && W_out) rc = download_soa(c, Wt, W_out);
    if (!rc) rc = download_soa(c, c->R, R_out);
    return rc;
}

namespace {
int step_impl(nsgpu_ctx* c, int nsteps, double fixed_dt, double* time_io, double* dt_last);
}

int nsgpu_step(nsgpu_ctx* c, int nsteps, double* time_io, double* dt_last) {
    return step_impl(c, nsteps, 0.0, time_io, dt_last);
}

int nsgpu_step_fixed(nsgpu_ctx* c, int nsteps, double dt, double* time_io) {
    if (!(dt > 0.0)) return NSGPU_ERR_INVALID;
    return step_impl(c, nsteps, dt, time_io, nullptr);
}

namespace {
// restores the context's CFL after a prescribed-dt call, on every exit path
struct CflGuard {
    nsgpu_ctx* c;
    double     cfl;
    ~CflGuard() { c->ph.cfl = cfl; }
};

// nsteps SSP-RK3 steps; fixed_dt > 0 replaces the CFL time step.  The stages and the time
// accumulation form dt = dtmin[0] * ph.cfl; a prescribed dt is stored as dtmin[0] = fixed_dt
// with ph.cfl = 1 for the call, so every stage and the time use fixed_dt itself (bitwise).
int step_impl(nsgpu_ctx* c, int nsteps, double fixed_dt, double* time_io, double* dt_last) {
    if (!c || nsteps < 0 || !net_ready(c) || c->linked) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    CflGuard guard{c, c->ph.cfl};
    if (fixed_dt > 0.0) c->ph.cfl = 1.0;
    if (time_io) HIPCHK(hipMemcpyAsync(c->dtime, time_io, sizeof(double), hipMemcpyHostToDevice, c->stream));
    const Launchers& L = *c->L;
    L.zero_flags(c->sflags, c->stream);
    // the last stage of a step that another follows forms the next step's prologue (dt and U
    // from its primitives, the stage-1 primitives c2p(U)) in place of the prologue kernel: the
    // same arithmetic per cell, on values the stage holds already
    const bool fuse = fused_diffusion(c) && !zfd(c) && !pipelined(c) && !c->timing;
    const StateCopy* cpp = use_copies(c) ? &c->cp : nullptr;
    bool       pro  = false;     // this step's prologue was formed by the previous step
    double*    wfirst = c->Ws[0];  // the stage-1 primitives of this step
    for (int s = 0; s < nsteps; s++) {
        const bool nx = fuse && s + 1 < nsteps;
        {
            Scope sc(c, T_STEP_BEGIN);
            if (pro) {
                HIPCHK(hipMemcpyAsync(&c->dtmin[0], &c->dtmin[2], sizeof(unsigned long long),
                                      hipMemcpyDeviceToDevice, c->stream));
            } else {
                wfirst = c->Ws[0];
                L.reset_min(&c->dtmin[0], c->stream);
                L.step_begin(c->W, nullptr, c->Un, wfirst, &c->dtmin[0], c->g, c->ph, c->stream,
                             cpp);
            }
            if (fixed_dt > 0.0) L.set_dt(&c->dtmin[0], fixed_dt, c->stream);
        }
        // stage inputs Un -> Ua -> Ub: Un (Uprev of time_scheme.cpp:27) is read by every stage
        // and overwritten (cell by cell, after its read) only by a last stage that forms the
        // next step's prologue
        double* cur  = c->Un;
        double* nxt  = c->Ua;
        double* wcur = wfirst;
        double* wnxt = (wfirst == c->Ws[0]) ? c->Ws[1] : c->Ws[0];
        for (int st = 0; st < 3; st++) {
            // the z faces of the halos go beside the x/y tiles (stage_inputs), which read only
            // x and y halos
            const bool zsplit = c->aux && !c->timing && !pipelined(c) && c->ph.conv &&
                                c->g.dim == 3 && !zfd(c);
            {
                // halos of U and of its primitives: the same copies (c2p acts per cell)
                Scope sc(c, T_HALO);
                if (cpp) {
                    // the copies' halos for the sweeps; of the double arrays only those of the
                    // primitives across x and y, which the residual and the derivative pass read
                    L.halo_faces(cur, 0, wcur, 5, c->g, c->b, 0, 4, c->stream, cpp);
                    if (!zsplit) L.halo_faces(cur, 0, wcur, 0, c->g, c->b, 4, 2, c->stream, cpp);
                } else if (zsplit) {
                    L.halo_faces(cur, 5, wcur, 5, c->g, c->b, 0, 4, c->stream, nullptr);
                } else {
                    L.halo(cur, 5, wcur, 5, c->g, c->b, c->stream);
                }
            }
            // the stage writes U and c2p(U) of the next stage; the last stage writes the
            // primitive state into the other state buffer (the residual reads wcur meanwhile)
            StageArgs a{c->Un, cur, nxt, (st == 2) ? c->W2 : wnxt, st, c->sflags};
            a.Uf = c->cp.U, a.Wf = c->cp.W;  // (the single-precision stage kernel needs them)
            if (st == 2 && nx) {  // wnxt is free in the last stage
                L.reset_min(&c->dtmin[2], c->stream);
                a.Unext  = c->Un;
                a.Wnext  = wnxt;
                a.dtnext = &c->dtmin[2];
                wfirst   = wnxt;
            }
            if (pipelined(c)) {
                hipEventRecord(c->ev_fork, c->stream);  // orders aux after the halo fill
                hipStreamWaitEvent(c->aux, c->ev_fork, 0);
                stage_pipelined(c, cur, wcur, a);
            } else {
                stage_inputs(c, cur, wcur, zsplit, cpp);
                Scope sc(c, T_RESIDUAL);
                residual(c, wcur, nullptr, nullptr, &a, &c->dtmin[0], NSGPU_RHS_BOTH);
            }
            double* t = cur;
            cur       = nxt;
            nxt       = (st == 0) ? c->Ub : t;
            t         = wcur;
            wcur      = wnxt;
            wnxt      = t;
        }
        c->Ulast = cur;  // (the last step of a call forms no prologue: Ulast is its U)
        double* t = c->W;
        c->W      = c->W2;
        c->W2     = t;
        L.time_accum(c->dtime, &c->dtmin[0], c->ph.cfl, c->stream);
        pro = nx;
    }
    HIPCHK(hipGetLastError());
    if (nsteps > 0) c->dt_scale = c->ph.cfl;
    unsigned fl = 0;
    if (time_io) HIPCHK(hipMemcpyAsync(time_io, c->dtime, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&fl, c->sflags, sizeof fl, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (dt_last) {
        int rc = nsgpu_get_dt(c, dt_last);
        if (rc) return rc;
    }
    return health_status(fl);
}
}  // namespace

int nsgpu_get_dt(nsgpu_ctx* c, double* dt) {
    if (!c || !dt) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, &c->dtmin[0], sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double m;
    std::memcpy(&m, &bits, sizeof m);
    *dt = m * c->dt_scale;
    return NSGPU_OK;
}

int nsgpu_reduce(nsgpu_ctx* c, double sums[5], double* dt_min, unsigned* flags) {
    if (!c) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    c->L->reset_min(&c->dtmin[1], c->stream);
    c->L->zero_flags(c->flags, c->stream);
    int nb = 0;
    c->L->reduce(c->Ulast, c->W, c->partial, c->partial_max, &c->dtmin[1], c->flags, c->g, c->ph,
                 c->stream, &nb);
    c->L->reduce_final(c->partial, nb, c->sums, c->stream);
    double             h[5];
    unsigned long long bits;
    unsigned           fl = 0;
    HIPCHK(hipMemcpyAsync(h, c->sums, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&bits, &c->dtmin[1], sizeof bits, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(&fl, c->flags, sizeof fl, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double vol = c->g.dx * c->g.dy * c->g.dz;
    if (sums)
        for (int q = 0; q < 5; q++) sums[q] = h[q] * vol;
    if (dt_min) {
        double m;
        std::memcpy(&m, &bits, sizeof m);
        *dt_min = m * c->ph.cfl;
    }
    if (flags) *flags = fl;
    return health_status(fl);
}

int nsgpu_selftest(int device, long n, unsigned long long seed, long* mismatches) {
    if (n < 0 || !mismatches) return NSGPU_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NSGPU_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return NSGPU_ERR_NODEVICE;
    HIPCHK(hipSetDevice(device));
    unsigned long long* d = nullptr;
    HIPCHK(hipMalloc((void**)&d, sizeof(*d)));
    unsigned long long h = 0;
    int                rc = NSGPU_OK;
    if (hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc) exact::launchers().selftest_div(n, seed, d, nullptr);
    if (!rc && hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost) != hipSuccess) rc = NSGPU_ERR_HIP;
    hipFree(d);
    *mismatches = (long)h;
    return rc;
}

int nsgpu_set_mlp(nsgpu_ctx* c, const float* params) {
    if (!c || !params || c->ph.weights != NSGPU_WEIGHTS_LEARNED) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->mlp_params, params, sizeof(float) * MLP_NPARAM,
                          hipMemcpyHostToDevice, c->stream));
    mlp_pack(c->mlp_params, c->mlp_packed, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ph.mlp = c->mlp_packed;
    return NSGPU_OK;
}

int nsgpu_mlp_recon(int device, int math, const float* params, long n, const double* S, int side,
                    double* out, int* regime) {
    if (!params || !S || !out || !regime || n < 1 || n > (1L << 30) || side < 0 || side > 1 ||
        math < NSGPU_MATH_EXACT || math > NSGPU_MATH_FP32)
        return NSGPU_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return NSGPU_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return NSGPU_ERR_NODEVICE;
    HIPCHK(hipSetDevice(device));
    const Launchers& L = (math == NSGPU_MATH_FAST)   ? fast::launchers()
                         : (math == NSGPU_MATH_FP32) ? fp32::launchers()
                                                     : exact::launchers();
    float*     dp = nullptr;
    MlpPacked* P  = nullptr;
    double *   dS = nullptr, *dout = nullptr;
    int*       drg = nullptr;
    int        rc  = NSGPU_OK;
    if (hipMalloc((void**)&dp, sizeof(float) * MLP_NPARAM) != hipSuccess ||
        hipMalloc((void**)&P, sizeof(MlpPacked)) != hipSuccess ||
        hipMalloc((void**)&dS, sizeof(double) * 5 * n) != hipSuccess ||
        hipMalloc((void**)&dout, sizeof(double) * n) != hipSuccess ||
        hipMalloc((void**)&drg, sizeof(int) * n) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    if (!rc && hipMemcpy(dp, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    if (!rc && hipMemcpy(dS, S, sizeof(double) * 5 * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    if (!rc) {
        mlp_pack(dp, P, nullptr);
        L.mlp_recon(P, n, dS, side, dout, drg, nullptr);
        if (hipGetLastError() != hipSuccess) rc = NSGPU_ERR_HIP;
    }
    if (!rc && hipMemcpy(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    if (!rc && hipMemcpy(regime, drg, sizeof(int) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    for (void* q : {(void*)dp, (void*)P, (void*)dS, (void*)dout, (void*)drg})
        if (q) hipFree(q);
    return rc;
}

int nsgpu_timing(nsgpu_ctx* c, int enable) {
    if (!c) return NSGPU_ERR_INVALID;
    c->timing = enable != 0;
    for (int t = 0; t < T_COUNT; t++) c->ms[t] = 0.0, c->launches[t] = 0;
    return NSGPU_OK;
}

int nsgpu_timer_count(void) { return T_COUNT; }

const char* nsgpu_timer_name(int i) { return (i >= 0 && i < T_COUNT) ? kTimerNames[i] : ""; }

int nsgpu_timer_read(nsgpu_ctx* c, int i, double* ms, long* launches) {
    if (!c || i < 0 || i >= T_COUNT) return NSGPU_ERR_INVALID;
    if (ms) *ms = c->ms[i];
    if (launches) *launches = c->launches[i];
    return NSGPU_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// ensembles (include/nsgpu.h): batched launches over the member blocks, one step per graph
// ------------------------------------------------------------------------------------------
struct nsgpu_ensemble {
    std::vector<nsgpu_ctx*> m;
    std::vector<EnsTiles>   tiles;
    std::vector<EnsBlock>   hblk;
    std::vector<int>        hstart;
    int                     total[EF_COUNT] = {};
    int                     device = 0;
    hipStream_t             stream = nullptr;
    EnsBlock*               dblk   = nullptr;
    int*                    dstart = nullptr;
    hipGraph_t              graph  = nullptr;
    hipGraphExec_t          exec   = nullptr;
    hipEvent_t              ev[2]  = {nullptr, nullptr};
    unsigned*               flags  = nullptr;  // health word of the stage outputs (all blocks)
    // decompositions (nsgpu_domain): neighbour member per face and the shared CFL minimum
    std::vector<std::array<int, 6>> nbr;
    unsigned long long*             shared_dt = nullptr;
    int                             copies    = 0;  // EnsTable::copies
};

namespace {

void ens_drop_graph(nsgpu_ensemble* e) {
    if (e->exec) hipGraphExecDestroy(e->exec);
    if (e->graph) hipGraphDestroy(e->graph);
    e->exec  = nullptr;
    e->graph = nullptr;
}

// descriptors and workgroup prefix sums from the members' current buffers and the tiles
int ens_build(nsgpu_ensemble* e) {
    const int n = (int)e->m.size();
    e->hblk.assign(n, EnsBlock{});
    e->hstart.assign((size_t)EF_COUNT * (n + 1), 0);
    for (int b = 0; b < n; b++) {
        nsgpu_ctx* c = e->m[b];
        EnsBlock&  k = e->hblk[b];
        k.g = c->g, k.b = c->b, k.ph = c->ph;
        k.W = c->W, k.Un = c->Un, k.Ua = c->Ua, k.Ub = c->Ub;
        k.Ws[0] = c->Ws[0], k.Ws[1] = c->Ws[1];
        for (int d = 0; d < 3; d++) k.F[d] = c->F[d];
        k.Dx = c->Dx, k.Dy = c->Dy;
        k.dtmin = e->shared_dt ? e->shared_dt : &c->dtmin[0];
        k.dtime = c->dtime;
        k.flags = e->flags;
        k.linked = c->linked;
        for (int f = 0; f < 6; f++) k.nbr[f] = e->nbr.empty() ? -1 : e->nbr[b][f];
        k.Uc = c->cp.U, k.Wc = c->cp.W;
        k.t     = e->tiles[b];
        ens_grids(k);
    }
    // the sweeps read the float copies when every block has them and takes the fused residual
    // (which writes them), as for a single block (use_copies)
    e->copies = 1;
    for (nsgpu_ctx* c : e->m)
        if (!copies_ok(c)) e->copies = 0;
    for (int f = 0; f < EF_COUNT; f++) {
        long acc = 0;
        for (int b = 0; b < n; b++) {
            e->hstart[(size_t)f * (n + 1) + b] = (int)acc;
            const int* gr = e->hblk[b].grid[f];
            acc += (long)gr[0] * gr[1] * gr[2];
            if (acc > 0x7fffffffL) return NSGPU_ERR_INVALID;
        }
        e->hstart[(size_t)f * (n + 1) + n] = (int)acc;
        e->total[f]                        = (int)acc;
    }
    HIPCHK(hipMemcpyAsync(e->dblk, e->hblk.data(), sizeof(EnsBlock) * n, hipMemcpyHostToDevice,
                          e->stream));
    HIPCHK(hipMemcpyAsync(e->dstart, e->hstart.data(), sizeof(int) * e->hstart.size(),
                          hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return NSGPU_OK;
}

EnsTable ens_table(const nsgpu_ensemble* e) {
    return EnsTable{e->dblk, e->dstart, (int)e->m.size(), e->copies};
}

void ens_enqueue(nsgpu_ensemble* e) {
    const nsgpu_ctx* c0 = e->m[0];
    c0->L->ens_step(ens_table(e), e->total, c0->ph, e->stream);
}

int ens_capture(nsgpu_ensemble* e) {
    ens_drop_graph(e);
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    ens_enqueue(e);
    hipGraph_t g  = nullptr;
    hipError_t rc = hipStreamEndCapture(e->stream, &g);
    if (rc != hipSuccess) {
        fprintf(stderr, "nsgpu: ensemble capture failed: %s\n", hipGetErrorString(rc));
        return NSGPU_ERR_HIP;
    }
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return NSGPU_OK;
}

// members may have swapped their state buffers (nsgpu_step alternates W and W2): re-read them
int ens_refresh(nsgpu_ensemble* e) {
    for (nsgpu_ctx* c : e->m) HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t b = 0; b < e->m.size(); b++)
        if (e->hblk[b].W != e->m[b]->W) return ens_build(e);
    return NSGPU_OK;
}

int ens_run(nsgpu_ensemble* e, int nsteps, bool graph) {
    if (graph && !e->exec) {
        int rc = ens_capture(e);
        if (rc) return rc;
    }
    for (int s = 0; s < nsteps; s++) {
        if (graph) HIPCHK(hipGraphLaunch(e->exec, e->stream));
        else ens_enqueue(e);
    }
    HIPCHK(hipGetLastError());
    return NSGPU_OK;
}

int ens_set_tiles(nsgpu_ensemble* e, const std::vector<EnsTiles>& t) {
    e->tiles = t;
    ens_drop_graph(e);
    return ens_build(e);
}

// state kept across autotuner trials: W in the member's operator-level array R, the stage state
// Ua (the conservative state the member reports) in its staging array, and on the host the time,
// the step's dt word (what nsgpu_get_dt reports) and its scale
struct EnsSaved {
    std::vector<double>             time, dt_scale;
    std::vector<unsigned long long> dtbits;
};

int ens_save(nsgpu_ensemble* e, EnsSaved& sv) {
    const size_t n = e->m.size();
    sv.time.resize(n), sv.dt_scale.resize(n), sv.dtbits.resize(n);
    for (size_t b = 0; b < n; b++) {
        nsgpu_ctx* c = e->m[b];
        HIPCHK(hipMemcpyAsync(c->R, c->W, c->arr_bytes, hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(c->stage_aos, c->Ua, c->arr_bytes, hipMemcpyDeviceToDevice,
                              e->stream));
        HIPCHK(hipMemcpyAsync(&sv.time[b], c->dtime, sizeof(double), hipMemcpyDeviceToHost,
                              e->stream));
        HIPCHK(hipMemcpyAsync(&sv.dtbits[b], &c->dtmin[0], sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, e->stream));
        sv.dt_scale[b] = c->dt_scale;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return NSGPU_OK;
}

int ens_restore(nsgpu_ensemble* e, const EnsSaved& sv) {
    for (size_t b = 0; b < e->m.size(); b++) {
        nsgpu_ctx* c = e->m[b];
        HIPCHK(hipMemcpyAsync(c->W, c->R, c->arr_bytes, hipMemcpyDeviceToDevice, e->stream));
        HIPCHK(hipMemcpyAsync(c->Ua, c->stage_aos, c->arr_bytes, hipMemcpyDeviceToDevice,
                              e->stream));
        HIPCHK(hipMemcpyAsync(c->dtime, &sv.time[b], sizeof(double), hipMemcpyHostToDevice,
                              e->stream));
        HIPCHK(hipMemcpyAsync(&c->dtmin[0], &sv.dtbits[b], sizeof(unsigned long long),
                              hipMemcpyHostToDevice, e->stream));
        c->dt_scale = sv.dt_scale[b];
        c->Ulast    = c->Ua;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return NSGPU_OK;
}

// ms per step of `steps` graph replays with tiles t (best of 2); the state is restored on every
// exit path (a failed restore is reported before a failed trial)
int ens_trial(nsgpu_ensemble* e, const std::vector<EnsTiles>& t, int steps, const EnsSaved& sv,
              double* ms) {
    int    rc   = ens_set_tiles(e, t);
    double best = 1e300;
    if (!rc) rc = ens_run(e, 1, true);
    for (int r = 0; r < 2 && !rc; r++) {
        float m = 0.f;
        if (hipEventRecord(e->ev[0], e->stream) != hipSuccess) rc = NSGPU_ERR_HIP;
        if (!rc) rc = ens_run(e, steps, true);
        if (!rc && hipEventRecord(e->ev[1], e->stream) != hipSuccess) rc = NSGPU_ERR_HIP;
        if (!rc && hipEventSynchronize(e->ev[1]) != hipSuccess) rc = NSGPU_ERR_HIP;
        if (!rc && hipEventElapsedTime(&m, e->ev[0], e->ev[1]) != hipSuccess) rc = NSGPU_ERR_HIP;
        if (!rc) best = std::min(best, (double)m / steps);
    }
    const int rr = ens_restore(e, sv);
    *ms = best;
    return rr ? rr : rc;
}

// ensemble creation; `nbr` (per member, per face: the neighbour member or -1) and `shared_dt`
// link the members into one decomposed block (nsgpu_domain), else both are empty / null
int ens_create(nsgpu_ctx* const* members, int n, const std::vector<std::array<int, 6>>& nbr,
               unsigned long long* shared_dt, nsgpu_ensemble** out) {
    if (!out) return NSGPU_ERR_INVALID;
    *out = nullptr;
    if (!members || n < 1 || n > 65536) return NSGPU_ERR_INVALID;
    const nsgpu_ctx* c0 = members[0];
    if (!c0) return NSGPU_ERR_INVALID;
    for (int b = 0; b < n; b++)
        if (!members[b] || (members[b]->linked != 0) != !nbr.empty()) return NSGPU_ERR_INVALID;
    for (int b = 0; b < n; b++) {
        const nsgpu_ctx* c = members[b];
        if (!c || c->device != c0->device || c->L != c0->L || c->ph.conv != c0->ph.conv ||
            c->ph.weights != c0->ph.weights || c->ph.diff_order != c0->ph.diff_order ||
            (c->ph.weights == NSGPU_WEIGHTS_LEARNED && !c->ph.mlp) ||
            c->ph.suth != c0->ph.suth || c->g.dim != c0->g.dim ||
            c->ph.alpha_mode != c0->ph.alpha_mode || c->ph.physics != c0->ph.physics)
            return NSGPU_ERR_INVALID;
        for (int a = 0; a < b; a++)
            if (members[a] == c) return NSGPU_ERR_INVALID;
    }
    // the batched kernels carry the reference's viscous terms only
    if (c0->ph.physics & NSGPU_PHYS_VISCOUS) return NSGPU_ERR_UNSUPPORTED;
    HIPCHK(hipSetDevice(c0->device));
    nsgpu_ensemble* e = new (std::nothrow) nsgpu_ensemble();
    if (!e) return NSGPU_ERR_INVALID;
    e->m.assign(members, members + n);
    e->device    = c0->device;
    e->nbr       = nbr;
    e->shared_dt = shared_dt;
    e->tiles.assign(n, EnsTiles{64, 64, 16});
    int rc = NSGPU_OK;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc && hipMalloc((void**)&e->dblk, sizeof(EnsBlock) * n) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc && hipMalloc((void**)&e->dstart, sizeof(int) * EF_COUNT * (n + 1)) != hipSuccess)
        rc = NSGPU_ERR_HIP;
    for (int t = 0; t < 2 && !rc; t++)
        if (hipEventCreate(&e->ev[t]) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc && hipMalloc((void**)&e->flags, sizeof(unsigned)) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc) rc = ens_build(e);
    // members are enrolled from here on (nsgpu_ensemble_destroy releases them)
    for (nsgpu_ctx* c : e->m) c->enrolled++;
    if (rc) {
        nsgpu_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return NSGPU_OK;
}
}  // namespace

extern "C" {

int nsgpu_ensemble_create(nsgpu_ctx* const* members, int n, nsgpu_ensemble** out) {
    // blocks of a decomposition belong to their domain's own ensemble
    return ens_create(members, n, {}, nullptr, out);
}

void nsgpu_ensemble_destroy(nsgpu_ensemble* e) {
    if (!e) return;
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    ens_drop_graph(e);
    if (e->dblk) hipFree(e->dblk);
    if (e->dstart) hipFree(e->dstart);
    if (e->flags) hipFree(e->flags);
    for (nsgpu_ctx* c : e->m) c->enrolled--;
    for (hipEvent_t v : e->ev)
        if (v) hipEventDestroy(v);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

int nsgpu_ensemble_set_tiles(nsgpu_ensemble* e, const int* tiles) {
    if (!e || !tiles) return NSGPU_ERR_INVALID;
    std::vector<EnsTiles> t(e->m.size());
    for (size_t b = 0; b < t.size(); b++) {
        t[b] = EnsTiles{tiles[3 * b], tiles[3 * b + 1], tiles[3 * b + 2]};
        for (int kw : {t[b].kw_flux, t[b].kw_cell})
            if (kw != 16 && kw != 32 && kw != 64) return NSGPU_ERR_INVALID;
        if (t[b].march < 1 || t[b].march > 1024) return NSGPU_ERR_INVALID;
    }
    HIPCHK(hipSetDevice(e->device));
    int rc = ens_refresh(e);
    return rc ? rc : ens_set_tiles(e, t);
}

int nsgpu_ensemble_get_tiles(const nsgpu_ensemble* e, int* tiles) {
    if (!e || !tiles) return NSGPU_ERR_INVALID;
    for (size_t b = 0; b < e->m.size(); b++) {
        tiles[3 * b]     = e->tiles[b].kw_flux;
        tiles[3 * b + 1] = e->tiles[b].kw_cell;
        tiles[3 * b + 2] = e->tiles[b].march;
    }
    return NSGPU_OK;
}

int nsgpu_ensemble_step(nsgpu_ensemble* e, int nsteps, int use_graph) {
    if (!e || nsteps < 0) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(e->device));
    int rc = ens_refresh(e);
    if (!rc) e->m[0]->L->zero_flags(e->flags, e->stream);
    if (!rc) rc = ens_run(e, nsteps, use_graph != 0);
    unsigned fl = 0;
    HIPCHK(hipMemcpyAsync(&fl, e->flags, sizeof fl, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    // the last stage leaves the conservative state in Ua and the primitive state in W
    for (nsgpu_ctx* c : e->m) {
        c->Ulast = c->Ua;
        if (nsteps > 0) c->dt_scale = c->ph.cfl;
    }
    return rc ? rc : health_status(fl);
}

int nsgpu_ensemble_tune(nsgpu_ensemble* e, int steps, double* ms_default, double* ms_tuned) {
    if (!e || steps < 1) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(e->device));
    int rc = ens_refresh(e);
    if (rc) return rc;
    EnsSaved saved;
    rc = ens_save(e, saved);
    if (rc) return rc;
    const size_t n = e->m.size();
    std::vector<EnsTiles> best(n, EnsTiles{64, 64, 16});
    double                tbest = 0;
    rc = ens_trial(e, best, steps, saved, &tbest);
    if (rc) return rc;
    if (ms_default) *ms_default = tbest;
    // classes: members with equal nz share kw_flux / kw_cell, members with equal nx share march
    std::map<int, std::vector<size_t>> by_nz, by_nx;
    for (size_t b = 0; b < n; b++) {
        by_nz[e->m[b]->g.nz].push_back(b);
        by_nx[e->m[b]->g.nx].push_back(b);
    }
    const int kws[3] = {64, 32, 16}, marches[4] = {16, 8, 32, 4};
    for (int pass = 0; pass < 2; pass++) {
        bool improved = false;
        auto try_field = [&](const std::vector<size_t>& cls, int EnsTiles::*field, const int* vals,
                             int nv) -> int {
            for (int v = 0; v < nv; v++) {
                if (best[cls[0]].*field == vals[v]) continue;
                std::vector<EnsTiles> cand = best;
                for (size_t b : cls) cand[b].*field = vals[v];
                double t  = 0;
                int    r2 = ens_trial(e, cand, steps, saved, &t);
                if (r2) return r2;
                if (t < tbest * 0.99) best = cand, tbest = t, improved = true;
            }
            return NSGPU_OK;
        };
        for (auto& kv : by_nz) {
            if ((rc = try_field(kv.second, &EnsTiles::kw_flux, kws, 3))) return rc;
            if ((rc = try_field(kv.second, &EnsTiles::kw_cell, kws, 3))) return rc;
        }
        for (auto& kv : by_nx)
            if ((rc = try_field(kv.second, &EnsTiles::march, marches, 4))) return rc;
        if (!improved) break;
    }
    rc = ens_set_tiles(e, best);
    if (ms_tuned) *ms_tuned = tbest;
    return rc;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// decomposed block on one GPU (include/nsgpu.h): the reference's Cartesian decomposition
// (decomposition.cpp:26-57) into member contexts whose faces are linked to their neighbours, and
// the exchange of Communicator::communicate (communicator.cpp:16-144) as halo kernels that read
// the neighbouring member's interior planes; dt is one minimum over all blocks
// (time_scheme.cpp:131-135, MPI_Allreduce MIN).  The members are advanced by a linked ensemble:
// one launch per kernel family over every block, one hipGraph per step.
// ------------------------------------------------------------------------------------------
struct nsgpu_domain {
    nsgpu_config                    global{};
    int                             p[3]   = {1, 1, 1};
    int                             device = 0;
    std::vector<nsgpu_ctx*>         blocks;  // rank = (bk * py + bj) * px + bi
    std::vector<std::array<int, 3>> start, count;
    nsgpu_ensemble*                 ens = nullptr;
    unsigned long long*             dt  = nullptr;  // shared CFL minimum (device word)
};

namespace {

// block sizes of one axis: n / p each, the remainder to block 0 (decomposition.cpp:41-52), and
// the start indices (decomposition.cpp:84-90)
void split_axis(int n, int p, int b, int* start, int* count) {
    const int local = n / p, rem = n % p;
    *count = local + (b == 0 ? rem : 0);
    *start = b * local + (b > 0 ? rem : 0);
}

}  // namespace

extern "C" {

int nsgpu_domain_create(const nsgpu_config* cfg, const int nproc[3], int device,
                        nsgpu_domain** out) {
    if (!out) return NSGPU_ERR_INVALID;
    *out = nullptr;
    if (!cfg || !nproc) return NSGPU_ERR_INVALID;
    int rc = validate(cfg);
    if (rc) return rc;
    const int n[3] = {cfg->nx, cfg->ny, cfg->nz};
    long      nb   = 1;
    for (int a = 0; a < 3; a++) {
        if (nproc[a] < 1 || nproc[a] > 1024) return NSGPU_ERR_INVALID;
        nb *= nproc[a];
        // every block at least as thick as the halo it hands to its neighbour
        if (nproc[a] > 1 && n[a] / nproc[a] < NGC) return NSGPU_ERR_INVALID;
    }
    if (nb > 65536 || (cfg->dim == 2 && nproc[2] > 1)) return NSGPU_ERR_INVALID;
    if (cfg->physics & NSGPU_PHYS_VISCOUS) return NSGPU_ERR_UNSUPPORTED;  // ensembles carry the
                                                                        // reference's terms only
    nsgpu_domain* d = new (std::nothrow) nsgpu_domain();
    if (!d) return NSGPU_ERR_INVALID;
    d->global = *cfg;
    d->device = device;
    for (int a = 0; a < 3; a++) d->p[a] = nproc[a];
    static const int lead[3] = {0, 3, 4};
    bool periodic[3];
    for (int a = 0; a < 3; a++) periodic[a] = cfg->bc[lead[a]] == NSGPU_BC_PERIODIC;
    // grid spacing of every block: the global one (grid.cpp:14-16)
    const double dx = cfg->Lx / cfg->nx, dy = cfg->Ly / cfg->ny, dz = cfg->Lz / cfg->nz;
    std::vector<std::array<int, 6>> nbr;
    auto rank_of = [&](int i, int j, int k) { return (k * nproc[1] + j) * nproc[0] + i; };
    for (int k = 0; k < nproc[2] && !rc; k++)
        for (int j = 0; j < nproc[1] && !rc; j++)
            for (int i = 0; i < nproc[0] && !rc; i++) {
                const int           bc[3] = {i, j, k};
                std::array<int, 3> st, ct;
                for (int a = 0; a < 3; a++) split_axis(n[a], nproc[a], bc[a], &st[a], &ct[a]);
                nsgpu_config m = *cfg;
                m.nx = ct[0], m.ny = ct[1], m.nz = ct[2];
                // the block's length at the global spacing; the spacing is set exactly below
                m.Lx = dx * ct[0], m.Ly = dy * ct[1], m.Lz = dz * ct[2];
                nsgpu_ctx* c = nullptr;
                rc = nsgpu_create(&m, device, &c);
                if (rc) break;
                d->blocks.push_back(c);
                d->start.push_back(st);
                d->count.push_back(ct);
                set_spacing(c->g, dx, dy, dz);
                // faces W, E (axis 0), N, S (1), F (high), B (low) (2): linked where a neighbouring
                // block lies behind (including the wrap of a periodic axis over several
                // blocks), the block's own wrap for a periodic axis of one block, walls otherwise
                std::array<int, 6> nb6;
                nb6.fill(-1);
                bool linked = false;
                for (int a = 0; a < 3; a++) {
                    const int faces[2] = {a == 2 ? 5 : 2 * a, a == 2 ? 4 : 2 * a + 1};  // low, high
                    c->b.periodic[a]   = periodic[a] && nproc[a] == 1;
                    for (int side = 0; side < 2; side++) {
                        int       o    = bc[a] + (side ? 1 : -1);
                        const int f    = faces[side];
                        bool      edge = (o < 0 || o >= nproc[a]);
                        if (edge && periodic[a] && nproc[a] > 1) o = (o + nproc[a]) % nproc[a], edge = false;
                        if (nproc[a] == 1 || edge) {
                            c->b.face_kind[f] = periodic[a] ? FACE_PERIODIC : FACE_NEUMANN;
                        } else {
                            int nbc[3] = {i, j, k};
                            nbc[a]     = o;
                            nb6[f]     = rank_of(nbc[0], nbc[1], nbc[2]);
                            c->b.face_kind[f] = FACE_LINK;
                            linked            = true;
                        }
                    }
                }
                c->linked = linked ? 1 : 0;
                nbr.push_back(nb6);
            }
    if (!rc && hipSetDevice(device) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc && hipMalloc((void**)&d->dt, sizeof(unsigned long long)) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc) {
        // a decomposition of one block has no links: an ordinary one-member ensemble
        const bool any = std::any_of(d->blocks.begin(), d->blocks.end(),
                                     [](const nsgpu_ctx* c) { return c->linked != 0; });
        rc = ens_create(d->blocks.data(), (int)d->blocks.size(),
                        any ? nbr : std::vector<std::array<int, 6>>{}, d->dt, &d->ens);
    }
    if (rc) {
        nsgpu_domain_destroy(d);
        return rc;
    }
    *out = d;
    return NSGPU_OK;
}

void nsgpu_domain_destroy(nsgpu_domain* d) {
    if (!d) return;
    if (d->ens) nsgpu_ensemble_destroy(d->ens);
    for (nsgpu_ctx* c : d->blocks) nsgpu_destroy(c);
    if (d->dt) {
        hipSetDevice(d->device);
        hipFree(d->dt);
    }
    delete d;
}

int nsgpu_domain_size(const nsgpu_domain* d) { return d ? (int)d->blocks.size() : 0; }

nsgpu_ctx* nsgpu_domain_block(nsgpu_domain* d, int rank, int start[3], int count[3]) {
    if (!d || rank < 0 || rank >= (int)d->blocks.size()) return nullptr;
    for (int a = 0; a < 3; a++) {
        if (start) start[a] = d->start[rank][a];
        if (count) count[a] = d->count[rank][a];
    }
    return d->blocks[rank];
}

int nsgpu_domain_step(nsgpu_domain* d, int nsteps, double* time_io, double* dt_last) {
    if (!d || nsteps < 0) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(d->device));
    if (time_io)
        for (nsgpu_ctx* c : d->blocks) {
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipMemcpy(c->dtime, time_io, sizeof(double), hipMemcpyHostToDevice));
        }
    const int rc = nsgpu_ensemble_step(d->ens, nsteps, 1);
    if (rc != NSGPU_OK && rc != NSGPU_ERR_NONFINITE && rc != NSGPU_ERR_POSITIVITY) return rc;
    if (time_io) HIPCHK(hipMemcpy(time_io, d->blocks[0]->dtime, sizeof(double), hipMemcpyDeviceToHost));
    if (dt_last) {
        unsigned long long bits = 0;
        HIPCHK(hipMemcpy(&bits, d->dt, sizeof bits, hipMemcpyDeviceToHost));
        double m;
        std::memcpy(&m, &bits, sizeof m);
        *dt_last = m * d->global.cfl;
    }
    return rc;
}

}  // extern "C"
