// Free-vibration analysis of the context's model (modal.cpp).
#pragma once
#include "slx_internal.hpp"

namespace slx {
// The lowest eigenpairs of (K_ff, M_ff) with the context's assembled K_ff and
// AMG hierarchy; d_mscale: per-element mass factors on the device (or null).
void modal_solve(Context& c, const slx_modal_options& o, const double* d_mscale, double total_mass, double* eig_out,
                 double* res_out, double* shapes, slx_modal_result& R);
}  // namespace slx
