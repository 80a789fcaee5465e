// sdr_tune.cpp — process-wide launch-shape knobs (sdr_tune_*).
#include "sdr_fsst.hpp"
#include "sdr_internal.hpp"
#include "sdr_tune.hpp"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace sdr {
namespace {

// A knob holds one value per value width (1, 2, 4 and 8 bytes); setting it
// without a width sets all four.
struct Knob {
  const char* name;
  uint64_t lo, hi, def;
  std::atomic<uint64_t> value[4];
};

#define SDR_KNOB(name, lo, hi, def) {name, lo, hi, def, {{def}, {def}, {def}, {def}}}
Knob g_knobs[] = {
  SDR_KNOB("rle_slice_rows", 1024, 1u << 20, RLE_SLICE_ROWS),
  SDR_KNOB("bp_groups_per_wave", 1, 16, 1),
  SDR_KNOB("bp_chunk_rows", 512, 2048, 2048),
  SDR_KNOB("alp_vectors_per_wave", 1, 16, 1),
  SDR_KNOB("fsst_spin_limit", 0, 1u << 31, FSST_SPIN_LIMIT),
  SDR_KNOB("batch_units_bitpacking", 1, 1u << 20, 16),
  SDR_KNOB("batch_units_rle", 1, 1u << 20, 2),
  SDR_KNOB("batch_units_alp", 1, 1u << 20, 16),
  SDR_KNOB("batch_units_alprd", 1, 1u << 20, 16),
  SDR_KNOB("batch_units_copy", 1, 1u << 20, 4),
};
#undef SDR_KNOB
constexpr uint32_t N_KNOBS = sizeof(g_knobs) / sizeof(g_knobs[0]);

bool valid(uint32_t key, uint64_t v)
{
  const Knob& k = g_knobs[key - 1];
  if (v < k.lo || v > k.hi) return false;
  if (key == SDR_TUNE_RLE_SLICE_ROWS && v % 64u) return false;
  if (key == SDR_TUNE_BP_CHUNK_ROWS && v != 512 && v != 1024 && v != 2048) return false;
  return true;
}

// Index of a value width; -1 if it is none of 1, 2, 4, 8.
int width_index(uint32_t ts)
{
  switch (ts) {
    case 1: return 0;
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    default: return -1;
  }
}

}  // namespace

uint32_t tune(uint32_t key, uint32_t ts)
{
  const int w = width_index(ts);
  return static_cast<uint32_t>(g_knobs[key - 1].value[w < 0 ? 2 : w].load(std::memory_order_relaxed));
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
  for (auto& v : g_knobs[key - 1].value) v.store(value, std::memory_order_relaxed);
  return SDR_OK;
}

int sdr_tune_set_width(uint32_t key, uint32_t type_size, uint64_t value)
{
  clear_error();
  if (key < 1 || key > N_KNOBS) return fail(SDR_E_INVALID_ARG, "unknown tuning key %u", key);
  const int w = width_index(type_size);
  if (w < 0) return fail(SDR_E_BAD_WIDTH, "tuning width %u is not 1, 2, 4 or 8", type_size);
  if (!valid(key, value))
    return fail(SDR_E_INVALID_ARG, "%s = %llu out of range", g_knobs[key - 1].name,
                static_cast<unsigned long long>(value));
  g_knobs[key - 1].value[w].store(value, std::memory_order_relaxed);
  return SDR_OK;
}

uint64_t sdr_tune_get(uint32_t key) { return sdr_tune_get_width(key, 4); }

uint64_t sdr_tune_get_width(uint32_t key, uint32_t type_size)
{
  const int w = width_index(type_size);
  if (key < 1 || key > N_KNOBS || w < 0) return 0;
  return g_knobs[key - 1].value[w].load(std::memory_order_relaxed);
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
    uint32_t width = 0;      // "name@ts value": one value width
    if (char* at = std::strchr(name, '@')) {
      *at = '\0';
      char* endp = nullptr;
      width      = static_cast<uint32_t>(std::strtoul(at + 1, &endp, 10));
      if (!endp || *endp || width_index(width) < 0) width = ~0u;
    }
    uint32_t key = 0;
    for (uint32_t k = 0; k < N_KNOBS; ++k)
      if (got == 2 && std::strcmp(name, g_knobs[k].name) == 0) key = k + 1;
    if (!key || width == ~0u) {
      rc = fail(SDR_E_INVALID_ARG, "%s:%d: bad tuning line", path, lineno);
      break;
    }
    rc = width ? sdr_tune_set_width(key, width, v) : sdr_tune_set(key, v);
  }
  std::fclose(f);
  return rc;
}

}  // extern "C"
