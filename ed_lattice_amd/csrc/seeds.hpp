#pragma once
#include "edl_internal.hpp"

namespace edl {

// One GF/chi seed: op 1 = c, 2 = c^dg, 3 = S^z; io, jo 1-based sites
// (jo = 0: single operator, jo > 0: the sum op_io + op_jo).
struct SeedSpec {
  int op, ispin, io, jo;
};

void launch_seed(const SeedSpec& s, const double* gs, int gs_dimup, const Sector& target, double* out, hipStream_t st);
void device_observables(Ctx& c, const Sector& s, const double* gs, double* dens_up, double* dens_dw, double* docc);

}  // namespace edl
