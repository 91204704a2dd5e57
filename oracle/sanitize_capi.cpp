// sanitize_capi.cpp -- drives the host side of the C ABI (include/nsgpu.h: argument and
// configuration validation, status strings, refusals before any device work) for the
// AddressSanitizer / UndefinedBehaviorSanitizer build of oracle/Makefile (target `sanitize`),
// which compiles nsgpu_capi.hip with the sanitizers on its host code only.  TEST INFRASTRUCTURE
// ONLY.  Without a GPU every valid request ends in NSGPU_ERR_NODEVICE; with one, the contexts it
// creates are stepped once and destroyed.  Exits 0 when every call returns the expected status.
#include <cstdio>
#include <cstring>
#include <vector>

#include "nsgpu.h"

static int fails = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what);
        fails++;
    }
}

static nsgpu_config base() {
    nsgpu_config c;
    std::memset(&c, 0, sizeof c);
    c.nx = 8, c.ny = 7, c.nz = 6, c.ngc = 3, c.dim = 3;
    c.Lx = c.Ly = c.Lz = 1.0;
    c.gamma = 1.4, c.mu = 1e-3, c.Pr = 0.72, c.cfl = 0.5;
    c.conv = NSGPU_CONV_WENO5CW, c.weights = NSGPU_WEIGHTS_SHU, c.diff_order = 4;
    c.math = NSGPU_MATH_EXACT, c.viscosity = NSGPU_VISC_CONSTANT, c.suth_C = 0.368;
    return c;
}

int main() {
    for (int s = 1; s >= -8; s--) expect(nsgpu_strerror(s) != nullptr, "strerror");
    nsgpu_ctx* ctx = nullptr;
    expect(nsgpu_create(nullptr, 0, &ctx) == NSGPU_ERR_INVALID && !ctx, "null config");
    expect(nsgpu_create(nullptr, 0, nullptr) == NSGPU_ERR_INVALID, "null out");
    // every configuration check of the validation, one field wrong at a time
    std::vector<void (*)(nsgpu_config&)> bad = {
        [](nsgpu_config& c) { c.ngc = 2; },       [](nsgpu_config& c) { c.nx = 2; },
        [](nsgpu_config& c) { c.nz = 0; },        [](nsgpu_config& c) { c.dim = 4; },
        [](nsgpu_config& c) { c.Lx = 0.0; },      [](nsgpu_config& c) { c.gamma = 1.0; },
        [](nsgpu_config& c) { c.cfl = -1.0; },    [](nsgpu_config& c) { c.conv = 7; },
        [](nsgpu_config& c) { c.weights = 9; },   [](nsgpu_config& c) { c.diff_order = 3; },
        [](nsgpu_config& c) { c.math = 5; },      [](nsgpu_config& c) { c.viscosity = 2; },
        [](nsgpu_config& c) { c.viscosity = 1, c.suth_C = 0.0; },
        [](nsgpu_config& c) { c.alpha_mode = 2; }, [](nsgpu_config& c) { c.physics = 8; },
        [](nsgpu_config& c) { c.physics = NSGPU_PHYS_VISCOUS, c.dim = 2, c.nz = 1; },
        [](nsgpu_config& c) { c.bc[0] = NSGPU_BC_NEUMANN; },  // x not periodic, East periodic
        [](nsgpu_config& c) { c.nx = 1 << 20, c.ny = 1 << 20; },
    };
    for (auto f : bad) {
        nsgpu_config c = base();
        f(c);
        expect(nsgpu_create(&c, 0, &ctx) == NSGPU_ERR_INVALID && !ctx, "invalid config");
    }
    {
        nsgpu_config c = base();
        c.weights      = NSGPU_WEIGHTS_LEARNED, c.alpha_mode = NSGPU_ALPHA_LOCAL_MAX;
        expect(nsgpu_create(&c, 0, &ctx) == NSGPU_ERR_UNSUPPORTED, "learned + options");
    }
    // the valid configuration: no device here, or a context to step on a GPU
    nsgpu_config c  = base();
    int          rc = nsgpu_create(&c, 0, &ctx);
    expect(rc == NSGPU_OK || rc == NSGPU_ERR_NODEVICE, "valid config");
    expect(nsgpu_create(&c, -1, &ctx) == NSGPU_ERR_NODEVICE || rc == NSGPU_ERR_NODEVICE,
           "bad device");
    if (rc == NSGPU_OK && ctx) {
        std::vector<double> W(nsgpu_array_len(ctx), 0.0);
        for (size_t t = 0; t < W.size(); t += 5) W[t] = 1.0, W[t + 1] = 1.0;
        expect(nsgpu_set_state(ctx, W.data(), NSGPU_STATE_PRIMITIVE) == NSGPU_OK, "set_state");
        expect(nsgpu_set_state(ctx, W.data(), 7) == NSGPU_ERR_INVALID, "bad kind");
        expect(nsgpu_step(ctx, -1, nullptr, nullptr) == NSGPU_ERR_INVALID, "negative steps");
        nsgpu_destroy(ctx);
    }
    nsgpu_destroy(nullptr);
    // the null-argument refusals of the other entry points
    expect(nsgpu_set_state(nullptr, nullptr, 0) == NSGPU_ERR_INVALID, "set_state null");
    expect(nsgpu_get_state(nullptr, nullptr, 0) == NSGPU_ERR_INVALID, "get_state null");
    expect(nsgpu_step(nullptr, 1, nullptr, nullptr) == NSGPU_ERR_INVALID, "step null");
    expect(nsgpu_step_fixed(nullptr, 1, -1.0, nullptr) == NSGPU_ERR_INVALID, "step_fixed dt");
    expect(nsgpu_reduce(nullptr, nullptr, nullptr, nullptr) == NSGPU_ERR_INVALID, "reduce null");
    expect(nsgpu_rhs(nullptr, nullptr, nullptr, nullptr, 3) == NSGPU_ERR_INVALID, "rhs null");
    expect(nsgpu_timing(nullptr, 1) == NSGPU_ERR_INVALID, "timing null");
    expect(nsgpu_timer_read(nullptr, 0, nullptr, nullptr) == NSGPU_ERR_INVALID, "timer null");
    expect(std::strlen(nsgpu_timer_name(-1)) == 0, "timer name range");
    nsgpu_ensemble* e = nullptr;
    expect(nsgpu_ensemble_create(nullptr, 1, &e) == NSGPU_ERR_INVALID && !e, "ensemble null");
    nsgpu_ensemble_destroy(nullptr);
    nsgpu_domain* d     = nullptr;
    const int     p3[3] = {2, 1, 1}, pz[3] = {1, 1, 0};
    expect(nsgpu_domain_create(nullptr, p3, 0, &d) == NSGPU_ERR_INVALID && !d, "domain null");
    expect(nsgpu_domain_create(&c, pz, 0, &d) == NSGPU_ERR_INVALID && !d, "domain nproc 0");
    nsgpu_domain_destroy(nullptr);
    expect(nsgpu_mlp_loss_grad_noalloc(0, nullptr, 1, nullptr, nullptr, nullptr, nullptr,
                                       nullptr) == NSGPU_ERR_INVALID, "noalloc null");
    expect(nsgpu_mlp_recon(0, 0, nullptr, 1, nullptr, 0, nullptr, nullptr) == NSGPU_ERR_INVALID,
           "recon null");
    long mism = -1;
    rc        = nsgpu_selftest(0, -5, 1, &mism);
    expect(rc == NSGPU_ERR_INVALID, "selftest n");
    std::printf("C-ABI sanitizer run: %d failures\n", fails);
    return fails ? 1 : 0;
}
