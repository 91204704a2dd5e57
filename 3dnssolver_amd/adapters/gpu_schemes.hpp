// gpu_schemes.hpp -- reference-side adapters: scheme classes with the reference's own interfaces
// whose work runs through the C ABI (include/nsgpu.h) on an MI355X.
//
// Drop-in for the reference solver: compile this header inside the reference tree (it includes
// the reference's own headers) and let Driver::create_convection_operator /
// create_diffusion_operator (driver.cpp:74-105) return these classes; the time loop
// (driver.cpp:43-68) can hold a GpuTimeScheme in place of TimeScheme (time_scheme.hpp:18-31).
//
//   GpuConvection : ConvectionScheme   solve(grid, W, U, R) adds the convective residual
//                                      (convection_scheme.hpp:26), Weno5 / Weno5cw semantics
//   GpuDiffusion  : DiffusionScheme    solve(grid, comm, W, R) adds the viscous residual
//                                      (diffusion_scheme.hpp:14-17)
//   GpuTimeScheme                      step(domain, grid, comm, conv, diff) / get_dt()
//                                      (time_scheme.hpp:25-31), state resident on the GPU
//
// The operator-level adapters copy W, U, R over PCIe on every call (exact but slow); the time
// scheme keeps the block on the GPU and only copies the primitive state back when the caller
// asks (sync_to_host), as the reference's writers would.
#pragma once

#include <stdexcept>
#include <string>

#include "convection_schemes/convection_scheme.hpp"
#include "diffusion_schemes/diffusion_scheme.hpp"
#include "domain.hpp"
#include "grid.hpp"
#include "params.hpp"
#include "time_scheme.hpp"

#include "nsgpu.h"

namespace nsgpu_adapters {

// Driver::create_convection_operator (driver.cpp:74-94) builds Weno5 / Weno5cw for "Weno",
// CenteredConvection for "Centered" and the null scheme for any other name.  The GPU path has no
// centred convection: such a configuration is refused rather than run without convection.
inline nsgpu_config make_config(const Params& p, const Array4d& like, int math) {
    if (p.convection_scheme == "Centered")
        throw std::invalid_argument("nsgpu: convection_scheme \"Centered\" (CenteredConvection) "
                                    "has no GPU implementation; use the reference's scheme");
    nsgpu_config c{};
    c.nx = (int)like.nx(), c.ny = (int)like.ny(), c.nz = (int)like.nz(), c.ngc = (int)like.ngc();
    c.dim = p.dim;
    // one block per rank: the block length is the domain length, so L / n is the reference's
    // dx = p.Lx / p.nx (grid.cpp:14-16) bit for bit; a sub-block uses its share of the length
    c.Lx = (c.nx == p.nx) ? p.Lx : p.Lx / p.nx * c.nx;
    c.Ly = (c.ny == p.ny) ? p.Ly : p.Ly / p.ny * c.ny;
    c.Lz = (c.nz == p.nz) ? p.Lz : p.Lz / p.nz * c.nz;
    c.gamma = p.gamma, c.mu = p.mu, c.Pr = p.Pr, c.cfl = p.CFL_max;
    if (p.convection_scheme == "Weno")
        c.conv = p.convection_in_characteristic_space ? NSGPU_CONV_WENO5CW : NSGPU_CONV_WENO5;
    else
        c.conv = NSGPU_CONV_NONE;  // the reference's null ConvectionScheme (driver.cpp:91-93)
    c.weights    = (p.weight_type == "Borges") ? NSGPU_WEIGHTS_BORGES : NSGPU_WEIGHTS_SHU;
    c.diff_order = (p.diffusion_scheme == "Centered") ? p.diffusion_scheme_order : 0;
    static const char* names[6] = {"West", "East", "North", "South", "Front", "Back"};
    for (int f = 0; f < 6; f++)
        c.bc[f] = (p.boundary_conditions.at(names[f]) == "Periodic") ? NSGPU_BC_PERIODIC
                                                                      : NSGPU_BC_NEUMANN;
    c.math      = math;
    c.viscosity = NSGPU_VISC_CONSTANT;  // the reference's CenteredDiffusion uses a constant mu
    return c;
}

inline void check(int rc) {
    if (rc != NSGPU_OK) throw std::runtime_error(std::string("nsgpu: ") + nsgpu_strerror(rc));
}

inline double* data(Array4d& a) { return &a(0, 0, 0, 0); }
// the const accessor of Array4d returns by value (array_4d.hpp:117-120): take the address through
// the non-const one; the C ABI only reads through const pointers
inline const double* data(const Array4d& a) { return &const_cast<Array4d&>(a)(0, 0, 0, 0); }

class GpuConvection : public ConvectionScheme {
public:
    GpuConvection(const Params& p, const Array4d& like, int device = 0,
                  int math = NSGPU_MATH_EXACT)
        : ConvectionScheme("Weno-5 (MI355X)", p.gamma) {
        nsgpu_config c = make_config(p, like, math);
        c.diff_order   = 0;
        check(nsgpu_create(&c, device, &m_ctx));
    }
    ~GpuConvection() override { nsgpu_destroy(m_ctx); }
    void solve(const Grid&, const Array4d& W, const Array4d& U, Array4d& R) override {
        check(nsgpu_convection_solve(m_ctx, data(W), data(U), data(R)));
    }

private:
    nsgpu_ctx* m_ctx = nullptr;
};

class GpuDiffusion : public DiffusionScheme {
public:
    GpuDiffusion(const Params& p, const Array4d& like, int device = 0,
                 int math = NSGPU_MATH_EXACT)
        : DiffusionScheme("Centered (MI355X)", p.gamma, p.mu, p.Pr) {
        nsgpu_config c = make_config(p, like, math);
        c.conv         = NSGPU_CONV_NONE;
        check(nsgpu_create(&c, device, &m_ctx));
    }
    ~GpuDiffusion() override { nsgpu_destroy(m_ctx); }
    void solve(const Grid&, const Communicator&, const Array4d& W, Array4d& R) override {
        check(nsgpu_diffusion_solve(m_ctx, data(W), data(R)));
    }

private:
    nsgpu_ctx* m_ctx = nullptr;
};

// Same public surface as TimeScheme (time_scheme.hpp:25-31).  The scheme objects passed to
// step() are not called: the whole step runs on the GPU with the schemes named in Params.
class GpuTimeScheme {
public:
    GpuTimeScheme(const Params& p, Domain& domain, int device = 0, int math = NSGPU_MATH_EXACT) {
        nsgpu_config c = make_config(p, domain.get_NS_primitive_array(), math);
        check(nsgpu_create(&c, device, &m_ctx));
        check(nsgpu_set_state(m_ctx, data(domain.get_NS_primitive_array()),
                              NSGPU_STATE_PRIMITIVE));
    }
    ~GpuTimeScheme() { nsgpu_destroy(m_ctx); }
    void step(Domain&, const Grid&, const Communicator&, ConvectionScheme*, DiffusionScheme*) {
        check(nsgpu_step(m_ctx, 1, nullptr, &m_dt));
    }
    double get_dt() const { return m_dt; }
    // copy the resident state back into the Domain (before output)
    void sync_to_host(Domain& domain) {
        check(nsgpu_get_state(m_ctx, data(domain.get_NS_primitive_array()), NSGPU_STATE_PRIMITIVE));
        check(nsgpu_get_state(m_ctx, data(domain.get_NS_conservative_array()),
                              NSGPU_STATE_CONSERVATIVE));
    }

private:
    nsgpu_ctx* m_ctx = nullptr;
    double     m_dt  = 0.0;
};

}  // namespace nsgpu_adapters
