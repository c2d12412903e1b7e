// AABB, Geometry, the tank shapes, FillSpec, build_tank_particles, make_block: the reference's geometry.hpp:17-182 at the same include path, provided by
// cryosph/api.hpp over the C ABI of libcryosph.so.
#pragma once

#include "cryosph/api.hpp"
