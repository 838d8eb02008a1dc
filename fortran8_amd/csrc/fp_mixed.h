// Mixed-precision path: FP32 storage and stencils, FP64 where accuracy
// demands it (Peterlin factor near the pole, Poisson defect correction,
// reductions).  Entry points used by the C-ABI host code.
#pragma once
#include "fp_internal.h"

namespace fp {
void mixed_create(Ctx& c);
void mixed_destroy(Ctx& c);
void mixed_load_from_f64(Ctx& c);   // FP64 state buffers -> mixed state
void mixed_store_to_f64(Ctx& c);    // mixed state -> FP64 state buffers
void mixed_phi_to_f64(Ctx& c);      // the FP64 potential only
void mixed_set_force(Ctx& c);
void mixed_set_markers(Ctx& c);
const double* mixed_marker_forces(Ctx& c);
double mixed_last_ndc(Ctx& c);
int mixed_stats(Ctx& c, double* out3);
int mixed_check_selection(Ctx& c, double* out3);
int mixed_get_selection(Ctx& c, uint8_t* out);
// tape entries of the adjoint's reverse sweep from the mixed step
void mixed_tape_in(Ctx& c, double* E, uint8_t* sel_out);
void mixed_tape_out(Ctx& c, double* E);
void mixed_widen_c(Ctx& c, double* dst);
void mixed_step(Ctx& c);
void mixed_join(Ctx& c);
void mixed_physics(Ctx& c);   // after Wi, beta or Lp changed
void mixed_velocity(Ctx& c, const float** u, const float** v, const float** w);  // FP32 state
}  // namespace fp
