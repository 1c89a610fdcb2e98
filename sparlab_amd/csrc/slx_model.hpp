// Host-side model construction (model.cpp).
#pragma once

#include "slx_internal.hpp"

namespace slx {

std::vector<double> hex8_ke0(const double* xe, double E, double nu);
std::vector<double> quad4_ke0(const double* xe, double E, double nu, double t);
void build_structured_grid(Context& c);
std::vector<double> element0_coords(const Context& c);
double element_volume(const Context& c);
HCsr structured_kff_pattern(const Context& c);
void build_filter(Context& c, double radius);
void identity_filter(Context& c);

}  // namespace slx
