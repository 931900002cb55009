// drift_internal.hpp -- declarations shared by the host-side translation units
// of libdrift_hip.so.
#pragma once

#include <string>

#include "../../include/drift_hip.h"
#include "drift_device.hpp"

namespace drift {

// Milky-Way spherical table (spherical_table.cpp).
void build_mw_params(const drift_potential& p, PotParams& out);
void build_mw_table(const drift_potential& p, double* coeffs);
double mw_table_max_rel_error(const drift_potential& p, const double* coeffs, int samples_per_seg);
double mw_spherical_exact(const drift_potential& p, double r2);

// Thread-local error message (drift_last_error).
void set_last_error(const std::string& msg);

}  // namespace drift
