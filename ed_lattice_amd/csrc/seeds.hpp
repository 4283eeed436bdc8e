#pragma once
#include "edl_internal.hpp"

namespace edl {

// One GF/chi seed: op 1 = c, 2 = c^dg, 3 = S^z; io, jo 1-based sites
// (jo = 0: single operator, jo > 0: the sum op_io + op_jo).
struct SeedSpec {
  int op, ispin, io, jo;
};

void launch_seed(const SeedSpec& s, const double* gs, int gs_dimup, const Sector& target, double* out, hipStream_t st);

// Current-operator seed of the optical conductivity of one orbital
// (lanc_ed_build_oc, OC_ELECTRONS.f90:103-213): terms coef * c+_a c_b of one
// spin along the ring order of the orbital's sites, in the ground-state
// sector.  The reference multiplies the seed by i (OC_ELECTRONS.f90:205);
// with a real H that factor leaves norm2, alpha and beta unchanged, so the
// seed is built real.
struct CurrentTerm {
  uint8_t a, b, spin;  // 0-based sites, spin 1 up / 2 down
  double coef;
};
constexpr int kMaxCurrentTerms = 128;
struct CurrentSeed {
  int n;
  CurrentTerm t[kMaxCurrentTerms];
};
void launch_current_seed(const CurrentSeed& J, const double* gs, const Sector& s, double* out, hipStream_t st);
void device_observables(Ctx& c, const Sector& s, const double* gs, double* dens_up, double* dens_dw, double* docc);
// <Sz_i Sz_j> and <n_i n_j> (Ns x Ns, column-major) of one state of weight 1.
void device_correlations(Ctx& c, const Sector& s, const double* gs, double* sz2, double* n2);
// lanc_energy (ENERGY.f90:70-353) of one state of weight 1, in the column
// order of energy_last.ed (ENERGY.f90:654-657): <K>, <Hi>, <V> = <Hi> - Ehf,
// <Eloc>, <Ehf>, Dust, Dund, Dse, Dph, Dkxy, Dkz (the last two are never
// accumulated by the reference's Lanczos path and stay 0).
void device_energy(Ctx& c, const Sector& s, const double* gs, double* e11);

}  // namespace edl
