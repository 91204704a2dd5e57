// ref_harness.hpp -- helpers of the reference harness (ref_harness.cpp) shared with the adapter
// harness (adapter_harness.cpp).  TEST INFRASTRUCTURE ONLY, like both harnesses.
#pragma once
#include <cstddef>

#include "convection_schemes/convection_scheme.hpp"
#include "diffusion_schemes/diffusion_scheme.hpp"
#include "params.hpp"

#include "nsoracle.h"

namespace ref_harness {
void              ensure_mpi();
Params            make_params(const nso_cfg* c);
ConvectionScheme* make_convection(const Params& p);
DiffusionScheme*  make_diffusion(const Params& p);
size_t            len(const Array4d& a);
bool              reference_semantics(const nso_cfg* c);
}  // namespace ref_harness
