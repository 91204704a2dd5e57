// adapter_harness.cpp -- runs the reference solver's own time loop with the reference-side GPU
// adapters of 3dnssolver_amd/adapters/gpu_schemes.hpp, compiled together with the reference's own
// translation units (oracle/Makefile, target `adapters`), so that the drop-in is exercised as a
// maintainer would wire it (driver.cpp:74-105 returning the adapters; the loop of
// driver.cpp:43-68 holding GpuTimeScheme).
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/, loaded by tests/test_gpu_adapters.py.  It
// links the product library (libnsgpu.so) the way the reference's build would; the product never
// links it.
#include <cstring>
#include <exception>
#include <mpi.h>

#include "communicator.hpp"
#include "decomposition.hpp"
#include "domain.hpp"
#include "grid.hpp"
#include "time_scheme.hpp"

#include "gpu_schemes.hpp"
#include "ref_harness.hpp"

using namespace ref_harness;

namespace {

// the adapters for the schemes Params names; the reference's null schemes otherwise
// (Driver::create_convection_operator / create_diffusion_operator, driver.cpp:74-105)
ConvectionScheme* gpu_convection(const Params& p, const Array4d& like) {
    if (p.convection_scheme == "Weno") return new nsgpu_adapters::GpuConvection(p, like);
    return new ConvectionScheme("null", p.gamma);
}

DiffusionScheme* gpu_diffusion(const Params& p, const Array4d& like) {
    if (p.diffusion_scheme == "Centered") return new nsgpu_adapters::GpuDiffusion(p, like);
    return new DiffusionScheme();
}

}  // namespace

extern "C" {

// nsteps x the reference's TimeScheme::step (time_scheme.cpp:3-70) from primitive W, with
// GpuConvection / GpuDiffusion as its scheme objects.  dts[s] = get_dt() after step s.
// Returns 0, -1 for a configuration the reference path does not have, -2 if an adapter threw.
int adapt_step_schemes(const nso_cfg* c, double* W_io, int nsteps, double* dts) {
    if (!reference_semantics(c)) return -1;
    ensure_mpi();
    try {
        Params        p = make_params(c);
        Decomposition dec(p);
        Grid          grid(p, dec);
        Communicator  comm(dec);
        Domain        domain(p, dec);
        Array4d&      W = domain.get_NS_primitive_array();
        std::memcpy(&W(0, 0, 0, 0), W_io, len(W) * sizeof(double));
        ConvectionScheme* conv = gpu_convection(p, W);
        DiffusionScheme*  diff = gpu_diffusion(p, W);
        TimeScheme        ts(p);
        for (int s = 0; s < nsteps; s++) {
            ts.step(domain, grid, comm, conv, diff);
            if (dts) dts[s] = ts.get_dt();
        }
        std::memcpy(W_io, &W(0, 0, 0, 0), len(W) * sizeof(double));
        delete conv;
        delete diff;
    } catch (const std::exception&) {
        return -2;
    }
    return 0;
}

// nsteps with GpuTimeScheme in place of TimeScheme (the state stays on the GPU; the Domain's
// arrays are brought up to date by sync_to_host at the end, before they are read).
int adapt_step_timescheme(const nso_cfg* c, double* W_io, int nsteps, double* dts) {
    if (!reference_semantics(c)) return -1;
    ensure_mpi();
    try {
        Params        p = make_params(c);
        Decomposition dec(p);
        Grid          grid(p, dec);
        Communicator  comm(dec);
        Domain        domain(p, dec);
        Array4d&      W = domain.get_NS_primitive_array();
        std::memcpy(&W(0, 0, 0, 0), W_io, len(W) * sizeof(double));
        ConvectionScheme*             conv = make_convection(p);  // passed, not called
        DiffusionScheme*              diff = make_diffusion(p);
        nsgpu_adapters::GpuTimeScheme ts(p, domain);
        for (int s = 0; s < nsteps; s++) {
            ts.step(domain, grid, comm, conv, diff);
            if (dts) dts[s] = ts.get_dt();
        }
        ts.sync_to_host(domain);
        std::memcpy(W_io, &W(0, 0, 0, 0), len(W) * sizeof(double));
        delete conv;
        delete diff;
    } catch (const std::exception&) {
        return -2;
    }
    return 0;
}

// make_config on a Params with convection_scheme `scheme`: 0 when it returns, 1 when it throws
// std::invalid_argument (CenteredConvection has no GPU counterpart), -1 otherwise
int adapt_make_config(const nso_cfg* c, const char* scheme) {
    Params p             = make_params(c);
    p.convection_scheme  = scheme;
    Array4d like(c->nx, c->ny, c->nz, 5, 3);
    try {
        nsgpu_adapters::make_config(p, like, NSGPU_MATH_EXACT);
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (...) {
        return -1;
    }
    return 0;
}

}  // extern "C"
