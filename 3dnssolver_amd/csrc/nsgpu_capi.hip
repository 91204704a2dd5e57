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

#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/nsgpu.h"
#include "ns_common.h"

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
    Geo                g{};
    Bnd                b{};
    Phys               ph{};
    const Launchers*   L = nullptr;
    size_t             plane_bytes = 0;  // one padded plane
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
    double*            R  = nullptr;  // residual (operator-level calls)
    double*            stage_aos = nullptr;  // host-layout staging (5 planes)
    double*            partial = nullptr;    // reduction partials
    int                partial_max = 0;
    unsigned long long* dtmin = nullptr;  // [0] step dt accumulator, [1] reduce accumulator
    unsigned*          flags = nullptr;
    double*            dtime = nullptr;   // device time accumulator
    double*            sums  = nullptr;   // [5]
    double*            Ulast = nullptr;   // buffer holding the last step's U
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
    if (c->nx < 3 || c->ny < 3 || c->nz < 3) return NSGPU_ERR_INVALID;
    if (c->dim != 2 && c->dim != 3) return NSGPU_ERR_INVALID;
    if (!(c->Lx > 0) || !(c->Ly > 0) || !(c->Lz > 0)) return NSGPU_ERR_INVALID;
    if (!(c->gamma > 1.0) || !(c->cfl > 0)) return NSGPU_ERR_INVALID;
    if (c->conv < NSGPU_CONV_NONE || c->conv > NSGPU_CONV_WENO5CW) return NSGPU_ERR_INVALID;
    if (c->weights != NSGPU_WEIGHTS_SHU && c->weights != NSGPU_WEIGHTS_BORGES) return NSGPU_ERR_INVALID;
    if (c->diff_order != 0 && c->diff_order != 2 && c->diff_order != 4 && c->diff_order != 6)
        return NSGPU_ERR_INVALID;
    if (c->math != NSGPU_MATH_EXACT && c->math != NSGPU_MATH_FAST && c->math != NSGPU_MATH_FP32)
        return NSGPU_ERR_INVALID;
    if (c->viscosity != NSGPU_VISC_CONSTANT && c->viscosity != NSGPU_VISC_SUTHERLAND)
        return NSGPU_ERR_INVALID;
    if (c->viscosity == NSGPU_VISC_SUTHERLAND && !(c->suth_C > 0)) return NSGPU_ERR_INVALID;
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

int alloc(double** p, size_t bytes) {
    HIPCHK(hipMalloc((void**)p, bytes));
    HIPCHK(hipMemset(*p, 0, bytes));
    return NSGPU_OK;
}

int upload_soa(nsgpu_ctx* c, const double* aos, double* dst) {
    HIPCHK(hipMemcpyAsync(c->stage_aos, aos, 5 * c->plane_bytes, hipMemcpyHostToDevice, c->stream));
    c->L->aos_to_soa(c->stage_aos, dst, c->g, c->stream);
    return NSGPU_OK;
}

int download_soa(nsgpu_ctx* c, const double* src, double* aos) {
    c->L->soa_to_aos(src, c->stage_aos, c->g, c->stream);
    HIPCHK(hipMemcpyAsync(aos, c->stage_aos, 5 * c->plane_bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return NSGPU_OK;
}

// convection face fluxes of the current (U, W); fz stays zero for dim == 2
void convection_fluxes(nsgpu_ctx* c, const double* U, const double* W) {
    if (!c->ph.conv) return;
    {
        Scope s(c, T_FLUX_X);
        c->L->face_flux(0, U, W, c->F[0], c->g, c->ph, c->stream);
    }
    {
        Scope s(c, T_FLUX_Y);
        c->L->face_flux(1, U, W, c->F[1], c->g, c->ph, c->stream);
    }
    if (c->g.dim == 3) {
        Scope s(c, T_FLUX_Z);
        c->L->face_flux(2, U, W, c->F[2], c->g, c->ph, c->stream);
    }
}

void diffusion_derivs(nsgpu_ctx* c, const double* W) {
    if (!c->ph.diff_order) return;
    Scope s(c, T_DERIV);
    c->L->deriv1(W, c->Dx, c->Dy, c->g, c->ph, c->stream);
    c->L->deriv_halo(c->Dx, c->Dy, c->g, c->b, c->stream);
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
    g.sy = g.NZ;
    g.sx = g.NY * g.NZ;
    g.P  = (long)g.NX * g.NY * g.NZ;
    g.dx = cfg->Lx / cfg->nx;  // Grid::compute_step_sizes (grid.cpp:14-16)
    g.dy = cfg->Ly / cfg->ny;
    g.dz = cfg->Lz / cfg->nz;
    g.dxx = g.dx * g.dx;  // CDN::ddf_dxdx / ddf_dydy divisors (cd-n.cpp:67-79)
    g.dyy = g.dy * g.dy;
    g.idx = 1.0 / g.dx, g.idy = 1.0 / g.dy, g.idz = 1.0 / g.dz;
    g.idxx = 1.0 / g.dxx, g.idyy = 1.0 / g.dyy;
    g.dim       = cfg->dim;
    g.square_xy = (g.dxx == g.dyy);
    static const int lead[3] = {0, 3, 4};
    for (int a = 0; a < 3; a++) c->b.periodic[a] = (cfg->bc[lead[a]] == NSGPU_BC_PERIODIC);
    for (int f = 0; f < 6; f++) c->b.face_kind[f] = cfg->bc[f];
    c->ph.gamma = cfg->gamma, c->ph.mu = cfg->mu, c->ph.Pr = cfg->Pr, c->ph.cfl = cfg->cfl;
    c->ph.conv = cfg->conv, c->ph.borges = (cfg->weights == NSGPU_WEIGHTS_BORGES);
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
    c->plane_bytes = (size_t)g.P * sizeof(double);
    size_t five = 5 * c->plane_bytes, four = 4 * c->plane_bytes;
    rc = NSGPU_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = NSGPU_ERR_HIP;
    if (!rc) rc = alloc(&c->W, five);
    if (!rc) rc = alloc(&c->W2, five);
    for (int t = 0; t < 2 && !rc; t++) rc = alloc(&c->Ws[t], five);
    if (!rc) rc = alloc(&c->Un, five);
    if (!rc) rc = alloc(&c->Ua, five);
    if (!rc) rc = alloc(&c->Ub, five);
    for (int d = 0; d < 3 && !rc; d++) rc = alloc(&c->F[d], five);
    if (!rc) rc = alloc(&c->Dx, four);
    if (!rc) rc = alloc(&c->Dy, four);
    if (!rc) rc = alloc(&c->R, five);
    if (!rc) rc = alloc(&c->stage_aos, five);
    c->partial_max = 4096;
    if (!rc) rc = alloc(&c->partial, sizeof(double) * 5 * c->partial_max);
    if (!rc) rc = alloc((double**)&c->dtmin, 2 * sizeof(unsigned long long));
    if (!rc) rc = alloc((double**)&c->flags, sizeof(double));
    if (!rc) rc = alloc(&c->dtime, sizeof(double));
    if (!rc) rc = alloc(&c->sums, 5 * sizeof(double));
    for (int t = 0; t < 2 * T_COUNT && !rc; t++)
        if (hipEventCreate(&c->ev[t]) != hipSuccess) rc = NSGPU_ERR_HIP;
    c->Ulast = c->Ua;
    if (rc) {
        nsgpu_destroy(c);
        return rc;
    }
    *out = c;
    return NSGPU_OK;
}

void nsgpu_destroy(nsgpu_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    double* bufs[] = {c->W,  c->W2, c->Ws[0], c->Ws[1], c->Un, c->Ua, c->Ub, c->F[0], c->F[1], c->F[2], c->Dx, c->Dy,
                      c->R,  c->stage_aos, c->partial, (double*)c->dtmin, (double*)c->flags,
                      c->dtime, c->sums};
    for (double* p : bufs)
        if (p) hipFree(p);
    for (int t = 0; t < 2 * T_COUNT; t++)
        if (c->ev[t]) hipEventDestroy(c->ev[t]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

size_t nsgpu_array_len(const nsgpu_ctx* c) { return c ? (size_t)c->g.P * 5 : 0; }

void* nsgpu_stream(nsgpu_ctx* c) { return c ? (void*)c->stream : nullptr; }

int nsgpu_set_state(nsgpu_ctx* c, const double* aos, int kind) {
    if (!c || !aos) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    if (kind == NSGPU_STATE_PRIMITIVE) {
        int rc = upload_soa(c, aos, c->W);
        if (rc) return rc;
    } else if (kind == NSGPU_STATE_CONSERVATIVE) {
        int rc = upload_soa(c, aos, c->Ua);
        if (rc) return rc;
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
    HIPCHK(hipSetDevice(c->device));
    int rc = upload_soa(c, aos, c->R);
    if (rc) return rc;
    c->L->halo(c->R, 5, nullptr, 0, c->g, c->b, c->stream);
    return download_soa(c, c->R, aos);
}

int nsgpu_convection_solve(nsgpu_ctx* c, const double* W_aos, const double* U_aos, double* R_aos) {
    if (!c || !W_aos || !U_aos || !R_aos) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    int rc = upload_soa(c, W_aos, c->W);
    if (!rc) rc = upload_soa(c, U_aos, c->Ub);
    if (!rc) rc = upload_soa(c, R_aos, c->R);
    if (rc) return rc;
    convection_fluxes(c, c->Ub, c->W);
    {
        Scope s(c, T_RESIDUAL);
        c->L->residual(c->F[0], c->F[1], c->F[2], c->W, c->Dx, c->Dy, c->R, c->R, nullptr,
                       nullptr, NSGPU_RHS_CONVECTION, c->g, c->ph, c->stream);
    }
    return download_soa(c, c->R, R_aos);
}

int nsgpu_diffusion_solve(nsgpu_ctx* c, const double* W_aos, double* R_aos) {
    if (!c || !W_aos || !R_aos) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    int rc = upload_soa(c, W_aos, c->W);
    if (!rc) rc = upload_soa(c, R_aos, c->R);
    if (rc) return rc;
    diffusion_derivs(c, c->W);
    {
        Scope s(c, T_RESIDUAL);
        c->L->residual(c->F[0], c->F[1], c->F[2], c->W, c->Dx, c->Dy, c->R, c->R, nullptr,
                       nullptr, NSGPU_RHS_DIFFUSION, c->g, c->ph, c->stream);
    }
    return download_soa(c, c->R, R_aos);
}

int nsgpu_rhs(nsgpu_ctx* c, double* U_aos, double* W_out, double* R_out, int which) {
    if (!c || !U_aos || !R_out || which < 1 || which > 3) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    int rc = upload_soa(c, U_aos, c->Ua);
    if (rc) return rc;
    {
        Scope s(c, T_HALO);
        c->L->halo(c->Ua, 5, nullptr, 0, c->g, c->b, c->stream);
    }
    {
        Scope s(c, T_C2P);
        c->L->c2p_faces(c->Ua, c->W, c->g, c->ph, c->stream);
    }
    if (which & 1) convection_fluxes(c, c->Ua, c->W);
    if (which & 2) diffusion_derivs(c, c->W);
    {
        Scope s(c, T_RESIDUAL);
        c->L->residual(c->F[0], c->F[1], c->F[2], c->W, c->Dx, c->Dy, nullptr, c->R, nullptr,
                       nullptr, which, c->g, c->ph, c->stream);
    }
    rc = download_soa(c, c->Ua, U_aos);
    if (!rc && W_out) rc = download_soa(c, c->W, W_out);
    if (!rc) rc = download_soa(c, c->R, R_out);
    c->Ulast = c->Ua;
    return rc;
}

int nsgpu_step(nsgpu_ctx* c, int nsteps, double* time_io, double* dt_last) {
    if (!c || nsteps < 0) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    if (time_io) HIPCHK(hipMemcpyAsync(c->dtime, time_io, sizeof(double), hipMemcpyHostToDevice, c->stream));
    const Launchers& L = *c->L;
    for (int s = 0; s < nsteps; s++) {
        {
            Scope sc(c, T_STEP_BEGIN);
            L.reset_min(&c->dtmin[0], c->stream);
            L.step_begin(c->W, nullptr, c->Un, c->Ws[0], &c->dtmin[0], c->g, c->ph, c->stream);
        }
        // stage inputs Un -> Ua -> Ub: Un (Uprev of time_scheme.cpp:27) is never overwritten
        double* cur  = c->Un;
        double* nxt  = c->Ua;
        double* wcur = c->Ws[0];
        double* wnxt = c->Ws[1];
        for (int st = 0; st < 3; st++) {
            {
                // halos of U and of its primitives: the same copies (c2p acts per cell)
                Scope sc(c, T_HALO);
                L.halo(cur, 5, wcur, 5, c->g, c->b, c->stream);
            }
            convection_fluxes(c, cur, wcur);
            diffusion_derivs(c, wcur);
            // the stage writes U and c2p(U) of the next stage; the last stage writes the
            // primitive state into the other state buffer (the residual reads wcur meanwhile)
            StageArgs a{c->Un, cur, nxt, (st == 2) ? c->W2 : wnxt, st};
            {
                Scope sc(c, T_RESIDUAL);
                L.residual(c->F[0], c->F[1], c->F[2], wcur, c->Dx, c->Dy, nullptr, nullptr, &a,
                           &c->dtmin[0], NSGPU_RHS_BOTH, c->g, c->ph, c->stream);
            }
            double* t = cur;
            cur       = nxt;
            nxt       = (st == 0) ? c->Ub : t;
            t         = wcur;
            wcur      = wnxt;
            wnxt      = t;
        }
        c->Ulast = cur;
        double* t = c->W;
        c->W      = c->W2;
        c->W2     = t;
        L.time_accum(c->dtime, &c->dtmin[0], c->ph.cfl, c->stream);
    }
    HIPCHK(hipGetLastError());
    if (time_io) HIPCHK(hipMemcpyAsync(time_io, c->dtime, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (dt_last) return nsgpu_get_dt(c, dt_last);
    return NSGPU_OK;
}

int nsgpu_get_dt(nsgpu_ctx* c, double* dt) {
    if (!c || !dt) return NSGPU_ERR_INVALID;
    HIPCHK(hipSetDevice(c->device));
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, &c->dtmin[0], sizeof(bits), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double m;
    std::memcpy(&m, &bits, sizeof m);
    *dt = m * c->ph.cfl;
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
    if (fl & 1u) return NSGPU_ERR_NONFINITE;
    if (fl & 2u) return NSGPU_ERR_POSITIVITY;
    if (fl & 4u) return NSGPU_ERR_ALPHA;
    return NSGPU_OK;
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
