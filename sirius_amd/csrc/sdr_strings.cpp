// sdr_strings.cpp — C ABI of the VARCHAR decode path (FSST).
#include "sdr_internal.hpp"

#include <cstdint>

using namespace sdr;

extern "C" {

uint64_t sdr_strings_upper_bound(const sdr_run* runs, uint32_t n_runs)
{
  uint64_t total = 0;
  for (uint32_t r = 0; runs && r < n_runs; ++r)
    for (uint32_t k = 0; runs[r].segments && k < runs[r].n_segments; ++k) {
      const sdr_segment& g = runs[r].segments[k];
      if (!g.row_count) continue;
      if (g.max_string_length == 0) return UINT64_MAX;
      total += uint64_t{g.row_count} * g.max_string_length;
    }
  return total;
}

int sdr_decode_strings_column_mode(sdr_string_column* col, sdr_stream_t stream, uint32_t launch_mode)
{
  (void)col;
  (void)stream;
  (void)launch_mode;
  clear_error();
  return fail(SDR_E_UNSUPPORTED_CODEC, "string decode is not built in this library yet");
}

int sdr_decode_strings_column(sdr_string_column* col, sdr_stream_t stream)
{
  return sdr_decode_strings_column_mode(col, stream, SDR_LAUNCH_BATCHED);
}

}  // extern "C"
