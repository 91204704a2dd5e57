// ref_harness.cpp -- drives the reference solver's own classes (compiled from the reference's
// src/ tree by oracle/Makefile, never copied) on caller-provided arrays, so that the C oracle
// (nsoracle.c) and the golden fixtures can be checked against the real reference.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/_ref/, loaded by tests/ and bench.py's
// cpu_baseline leg.  The product never links it.
//
// The harness replaces Driver::run (driver.cpp:3-72) and the XML reader: it fills Params
// directly (params.hpp:20-43), builds Decomposition/Grid/Communicator/Domain like the driver does
// and calls the scheme objects that Driver::create_convection_operator /
// create_diffusion_operator (driver.cpp:74-105) would create.  MPI is the image's MPICH, run as a
// single self-initialised rank.
#include <cstdlib>
#include <cstring>
#include <sys/stat.h>
#include <unistd.h>
#include <mpi.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "communicator.hpp"
#include "convection_schemes/weno-5.hpp"
#include "convection_schemes/weno-5cw.hpp"
#include "decomposition.hpp"
#include "diffusion_schemes/centered_diffusion.hpp"
#include "domain.hpp"
#include "grid.hpp"
#include "params.hpp"
#include "io/io_binary.hpp"
#include "testcases.hpp"
// eq_of_state.hpp comes in through time_scheme.hpp only: its include guard is mismatched
// (eq_of_state.hpp:1-2), so a second inclusion would redefine EqOfState.
#include "time_scheme.hpp"

#include "nsoracle.h"
#include "ref_harness.hpp"

namespace ref_harness {

void ensure_mpi() {
    int flag = 0;
    MPI_Initialized(&flag);
    if (!flag) MPI_Init(nullptr, nullptr);
}

Params make_params(const nso_cfg* c) {
    Params p;
    p.dim                                = c->dim;
    p.io_interval                        = 0;
    p.nx                                 = c->nx;
    p.ny                                 = c->ny;
    p.nz                                 = c->nz;
    p.nproc_x                            = 1;
    p.nproc_y                            = 1;
    p.nproc_z                            = 1;
    p.Lx                                 = c->Lx;
    p.Ly                                 = c->Ly;
    p.Lz                                 = c->Lz;
    p.mu                                 = c->mu;
    p.Pr                                 = c->Pr;
    p.gamma                              = c->gamma;
    p.CFL_max                            = c->cfl;
    p.end_t                              = 1.0;
    p.testcase                           = "none";
    p.riemann_solver                     = "none";
    p.convection_scheme                  = (c->conv == NSO_CONV_NONE) ? "none" : "Weno";
    p.convection_scheme_order            = 5;
    p.convection_in_characteristic_space = (c->conv == NSO_CONV_WENO5CW) ? 1 : 0;
    p.weight_type                        = (c->weights == NSO_W_BORGES) ? "Borges" : "Shu";
    p.diffusion_scheme                   = (c->diff_order == 0) ? "none" : "Centered";
    p.diffusion_scheme_order             = c->diff_order ? c->diff_order : 2;
    static const char* names[6]          = {"West", "East", "North", "South", "Front", "Back"};
    for (int f = 0; f < 6; f++)
        p.boundary_conditions[names[f]] = (c->bc[f] == NSO_BC_PERIODIC) ? "Periodic" : "Neumann";
    return p;
}

ConvectionScheme* make_convection(const Params& p) {
    if (p.convection_scheme == "Weno") {
        if (p.convection_in_characteristic_space)
            return new Weno5cw("Weno-5cw", p.gamma, p.weight_type);
        return new Weno5("Weno-5cw", p.gamma, p.weight_type);
    }
    return new ConvectionScheme("null", p.gamma);
}

DiffusionScheme* make_diffusion(const Params& p) {
    if (p.diffusion_scheme == "Centered")
        return new CenteredDiffusion("Centered", p.gamma, p.mu, p.Pr, p.diffusion_scheme_order);
    return new DiffusionScheme();
}

size_t len(const Array4d& a) { return a.NX() * a.NY() * a.NZ() * a.NQ(); }

// the options the reference's own code path has: constant mu, alpha = 3, its eigen system
bool reference_semantics(const nso_cfg* c) {
    return c->viscosity == 0 && c->alpha_mode == NSO_ALPHA_CONST3 && c->physics == 0 &&
           c->weights != NSO_W_LEARNED;
}

} // namespace ref_harness

using namespace ref_harness;

extern "C" {

// One stage body of TimeScheme::step (time_scheme.cpp:33-41): BC, exchange, c2p, R = conv+diff.
int ref_rhs(const nso_cfg* c, double* U_io, double* W_out, double* R_out, int which) {
    if (!reference_semantics(c)) return -1;
    ensure_mpi();
    Params        p = make_params(c);
    Decomposition dec(p);
    Grid          grid(p, dec);
    Communicator  comm(dec);
    Domain        domain(p, dec);
    Array4d&      U = domain.get_NS_conservative_array();
    Array4d&      W = domain.get_NS_primitive_array();
    std::memcpy(&U(0, 0, 0, 0), U_io, len(U) * sizeof(double));
    domain.apply_boundary_conditions(U);
    comm.communicate(U);
    EqOfState eos(p.gamma);
    eos.conservative_to_primitive(U, W);
    Array4d           R(W.get_dimensions());
    ConvectionScheme* conv = make_convection(p);
    DiffusionScheme*  diff = make_diffusion(p);
    if (which & 1) conv->solve(grid, W, U, R);
    if (which & 2) diff->solve(grid, comm, W, R);
    std::memcpy(U_io, &U(0, 0, 0, 0), len(U) * sizeof(double));
    std::memcpy(W_out, &W(0, 0, 0, 0), len(W) * sizeof(double));
    std::memcpy(R_out, &R(0, 0, 0, 0), len(R) * sizeof(double));
    delete conv;
    delete diff;
    return 0;
}

// nsteps x TimeScheme::step (time_scheme.cpp:3-70) from primitive W.
int ref_step(const nso_cfg* c, double* W_io, double* U_out, int nsteps, double* dts) {
    if (!reference_semantics(c)) return -1;
    ensure_mpi();
    Params        p = make_params(c);
    Decomposition dec(p);
    Grid          grid(p, dec);
    Communicator  comm(dec);
    Domain        domain(p, dec);
    Array4d&      U = domain.get_NS_conservative_array();
    Array4d&      W = domain.get_NS_primitive_array();
    std::memcpy(&W(0, 0, 0, 0), W_io, len(W) * sizeof(double));
    ConvectionScheme* conv = make_convection(p);
    DiffusionScheme*  diff = make_diffusion(p);
    TimeScheme        ts(p);
    for (int s = 0; s < nsteps; s++) {
        ts.step(domain, grid, comm, conv, diff);
        if (dts) dts[s] = ts.get_dt();
    }
    std::memcpy(W_io, &W(0, 0, 0, 0), len(W) * sizeof(double));
    std::memcpy(U_out, &U(0, 0, 0, 0), len(U) * sizeof(double));
    delete conv;
    delete diff;
    return 0;
}

// The reference's initial conditions (testcases.cpp:5-344) on the block of `c`: id 0..2 Sod along
// x, y, z; 3..8 the 2-D Riemann configurations 3, 4, 6, 12, 15, 17; 9 shear_layer_xy, 10
// shear_layer_3d (their rand() draws from the sequence of a fresh process: srand(1) first).
// W_out gets the whole padded primitive array (halos as a fresh Domain leaves them: zero).
int ref_testcase(const nso_cfg* c, int id, double* W_out) {
    ensure_mpi();
    Params        p = make_params(c);
    Decomposition dec(p);
    Grid          grid(p, dec);
    Domain        domain(p, dec);
    typedef void (*ic_fn)(Domain&, const Grid&, const Params&);
    static const ic_fn fns[] = {testcases::sod_xdir, testcases::sod_ydir, testcases::sod_zdir,
                                testcases::case3,    testcases::case4,    testcases::case6,
                                testcases::case12,   testcases::case15,   testcases::case17,
                                testcases::shear_layer_xy, testcases::shear_layer_3d};
    if (id < 0 || id >= (int)(sizeof(fns) / sizeof(fns[0]))) return -1;
    srand(1);
    fns[id](domain, grid, p);
    Array4d& W = domain.get_NS_primitive_array();
    std::memcpy(W_out, &W(0, 0, 0, 0), len(W) * sizeof(double));
    return 0;
}

// BinaryWriter::write_mesh and write_fields (io/io_binary.cpp:3-75) of the primitive array W_in
// for iteration `iter`, into <dir>/results/ (the writer's relative path); returns 0 on success.
int ref_write_binary(const nso_cfg* c, const double* W_in, int iter, const char* dir) {
    ensure_mpi();
    Params        p = make_params(c);
    Decomposition dec(p);
    Grid          grid(p, dec);
    Domain        domain(p, dec);
    Array4d&      W = domain.get_NS_primitive_array();
    std::memcpy(&W(0, 0, 0, 0), W_in, len(W) * sizeof(double));
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd) || chdir(dir) != 0) return -1;
    mkdir("results", 0755);
    BinaryWriter writer;
    writer.write_mesh(grid);
    writer.write_fields(domain, (size_t)iter);
    return chdir(cwd) == 0 ? 0 : -1;
}

int ref_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
}
