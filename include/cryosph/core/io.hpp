// write_vtk, write_json_frame, make_frame: the reference's io.hpp:12-22 at the same include path, provided by
// cryosph/api.hpp over the C ABI of libcryosph.so.
#pragma once

#include "cryosph/api.hpp"
