// SpatialHash: the reference's spatial_hash.hpp:58-92 at the same include path, provided by
// cryosph/api.hpp over the C ABI of libcryosph.so.
#pragma once

#include "cryosph/api.hpp"
