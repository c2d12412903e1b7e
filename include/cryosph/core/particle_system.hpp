// Phase, pflag, HostParticles, ParticleSystem: the reference's particle_system.hpp:13-142 at the same include path, provided by
// cryosph/api.hpp over the C ABI of libcryosph.so.
#pragma once

#include "cryosph/api.hpp"
