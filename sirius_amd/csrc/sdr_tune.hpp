// sdr_tune.hpp — current launch-shape knobs for the launchers.
#pragma once

#include "../../include/sdr/decode.h"
#include "sdr_kernels.hpp"

#include <stdint.h>

namespace sdr {

// Current value of an sdr_tune_key (the key must be valid).
uint32_t tune(uint32_t key);

}  // namespace sdr
