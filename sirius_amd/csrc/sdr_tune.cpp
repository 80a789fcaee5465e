// sdr_tune.cpp — process-wide launch-shape knobs (sdr_tune_*).
#include "sdr_fsst.hpp"
#include "sdr_internal.hpp"
#include "sdr_tune.hpp"

#include <atomic>
#include <cstdio>
#include <cstring>

namespace sdr {
namespace {

struct Knob {
  const char* name;
  uint64_t lo, hi, def;
  std::atomic<uint64_t> value;
};

Knob g_knobs[] = {
  {"rle_slice_rows", 1024, 1u << 20, RLE_SLICE_ROWS, {RLE_SLICE_ROWS}},
  {"bp_groups_per_wave", 1, 16, 1, {1}},
  {"bp_chunk_rows", 512, 2048, 2048, {2048}},
  {"alp_vectors_per_wave", 1, 16, 1, {1}},
  {"fsst_spin_limit", 0, 1u << 31, FSST_SPIN_LIMIT, {FSST_SPIN_LIMIT}},
};
constexpr uint32_t N_KNOBS = sizeof(g_knobs) / sizeof(g_knobs[0]);

bool valid(uint32_t key, uint64_t v)
{
  const Knob& k = g_knobs[key - 1];
  if (v < k.lo || v > k.hi) return false;
  if (key == SDR_TUNE_RLE_SLICE_ROWS && v % 64u) return false;
  if (key == SDR_TUNE_BP_CHUNK_ROWS && v != 512 && v != 1024 && v != 2048) return false;
  return true;
}

}  // namespace

uint32_t tune(uint32_t key)
{
  return static_cast<uint32_t>(g_knobs[key - 1].value.load(std::memory_order_relaxed));
}

}  // namespace sdr

using namespace sdr;

extern "C" {

int sdr_tune_set(uint32_t key, uint64_t value)
{
  clear_error();
  if (key < 1 || key > N_KNOBS) return fail(SDR_E_INVALID_ARG, "unknown tuning key %u", key);
  if (!valid(key, value))
    return fail(SDR_E_INVALID_ARG, "%s = %llu out of range", g_knobs[key - 1].name,
                static_cast<unsigned long long>(value));
  g_knobs[key - 1].value.store(value, std::memory_order_relaxed);
  return SDR_OK;
}

uint64_t sdr_tune_get(uint32_t key)
{
  if (key < 1 || key > N_KNOBS) return 0;
  return g_knobs[key - 1].value.load(std::memory_order_relaxed);
}

int sdr_tune_load(const char* path)
{
  clear_error();
  if (!path) return fail(SDR_E_INVALID_ARG, "path is NULL");
  FILE* f = std::fopen(path, "r");
  if (!f) return fail(SDR_E_NOT_FOUND, "cannot open tuning table %s", path);
  char line[256];
  int rc     = SDR_OK;
  int lineno = 0;
  while (rc == SDR_OK && std::fgets(line, sizeof line, f)) {
    ++lineno;
    if (char* h = std::strchr(line, '#')) *h = '\0';
    char name[64];
    unsigned long long v = 0;
    const int got = std::sscanf(line, "%63s %llu", name, &v);
    if (got <= 0) continue;  // blank or comment
    uint32_t key = 0;
    for (uint32_t k = 0; k < N_KNOBS; ++k)
      if (got == 2 && std::strcmp(name, g_knobs[k].name) == 0) key = k + 1;
    if (!key) {
      rc = fail(SDR_E_INVALID_ARG, "%s:%d: bad tuning line", path, lineno);
      break;
    }
    rc = sdr_tune_set(key, v);
  }
  std::fclose(f);
  return rc;
}

}  // extern "C"
