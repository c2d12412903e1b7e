// FluidProperties and its presets: the reference's fluid_properties.hpp:11-38 at the same include path, provided by
// cryosph/api.hpp over the C ABI of libcryosph.so.
#pragma once

#include "cryosph/api.hpp"
