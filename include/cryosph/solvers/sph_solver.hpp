// Solver, SolverConfig, ManeuverProfile, Diagnostics: the reference's sph_solver.hpp:20-232 at the same include path, provided by
// cryosph/api.hpp over the C ABI of libcryosph.so.
#pragma once

#include "cryosph/api.hpp"
