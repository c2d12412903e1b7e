// TelemetryMode, TelemetryFrame, TelemetryRecord, TelemetryStream: the reference's telemetry.hpp:12-70 at the same include path, provided by
// cryosph/api.hpp over the C ABI of libcryosph.so.
#pragma once

#include "cryosph/api.hpp"
