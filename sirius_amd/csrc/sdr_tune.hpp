// sdr_tune.hpp — current launch-shape knobs for the launchers.
#pragma once

#include "../../include/sdr/decode.h"
#include "sdr_kernels.hpp"

#include <stdint.h>

namespace sdr {

// Current value of an sdr_tune_key (the key must be valid) for values of
// `ts` bytes (1, 2, 4 or 8; the 4-byte value for any other).
uint32_t tune(uint32_t key, uint32_t ts = 4);

}  // namespace sdr
