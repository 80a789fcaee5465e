// sdr_arrow.cpp — decoded columns as Arrow C device arrays (host only).
//
// The reference wraps its decode outputs in cuDF columns without a copy
// (gpu_native_decode.cu:546-608, gpu_decode_strings.cu:100-402).  Here the
// same buffers are described by an ArrowSchema and an ArrowDeviceArray; the
// structs own only their small host-side state (format and name strings,
// the buffer pointer list), and releasing the array calls the buffers'
// owner back.
#include "../../include/sdr/arrow.h"
#include "sdr_internal.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>

namespace sdr {
namespace {

struct SchemaState {
  char format[32];
  char* name;
};

struct ArrayState {
  const void* buffers[3];
  void (*release)(void*);
  void* owner;
};

void release_schema(ArrowSchema* s)
{
  if (!s || !s->release) return;
  auto* st = static_cast<SchemaState*>(s->private_data);
  if (st) {
    std::free(st->name);
    delete st;
  }
  s->release = nullptr;
}

void release_array(ArrowArray* a)
{
  if (!a || !a->release) return;
  auto* st = static_cast<ArrayState*>(a->private_data);
  if (st) {
    if (st->release) st->release(st->owner);
    delete st;
  }
  a->release = nullptr;
}

// Arrow format string of a column; false for an unknown type.
bool format_of(const sdr_arrow_column& c, char (&f)[32])
{
  const char* fixed = nullptr;
  switch (c.type) {
    case SDR_ARROW_INT8: fixed = "c"; break;
    case SDR_ARROW_INT16: fixed = "s"; break;
    case SDR_ARROW_INT32: fixed = "i"; break;
    case SDR_ARROW_INT64: fixed = "l"; break;
    case SDR_ARROW_UINT8: fixed = "C"; break;
    case SDR_ARROW_UINT16: fixed = "S"; break;
    case SDR_ARROW_UINT32: fixed = "I"; break;
    case SDR_ARROW_UINT64: fixed = "L"; break;
    case SDR_ARROW_FLOAT32: fixed = "f"; break;
    case SDR_ARROW_FLOAT64: fixed = "g"; break;
    case SDR_ARROW_DATE32: fixed = "tdD"; break;
    case SDR_ARROW_TIMESTAMP_US: fixed = "tsu:"; break;
    case SDR_ARROW_TIME_US: fixed = "ttu"; break;
    case SDR_ARROW_UTF8: fixed = "u"; break;
    case SDR_ARROW_TIMESTAMP_S: fixed = "tss:"; break;
    case SDR_ARROW_TIMESTAMP_MS: fixed = "tsm:"; break;
    case SDR_ARROW_TIMESTAMP_NS: fixed = "tsn:"; break;
    case SDR_ARROW_TIMESTAMP_US_UTC: fixed = "tsu:UTC"; break;
    case SDR_ARROW_DECIMAL:
      // decimal128 takes no bit width; the 32- and 64-bit layouts name theirs
      if (c.value_bytes == 16) std::snprintf(f, sizeof f, "d:%u,%u", c.precision, c.scale);
      else std::snprintf(f, sizeof f, "d:%u,%u,%u", c.precision, c.scale, c.value_bytes * 8u);
      return true;
    default: return false;
  }
  std::snprintf(f, sizeof f, "%s", fixed);
  return true;
}

}  // namespace
}  // namespace sdr

using namespace sdr;

extern "C" int sdr_arrow_export(const sdr_arrow_column* col, ArrowSchema* schema, ArrowDeviceArray* array)
{
  clear_error();
  if (!col || !schema || !array) return fail(SDR_E_INVALID_ARG, "column, schema or array is NULL");
  const sdr_arrow_column& c = *col;
  if (c.type < SDR_ARROW_INT8 || c.type > SDR_ARROW_TIMESTAMP_US_UTC) return fail(SDR_E_INVALID_ARG, "unknown Arrow type %u", c.type);
  if (c.type == SDR_ARROW_DECIMAL && c.value_bytes != 4 && c.value_bytes != 8 && c.value_bytes != 16)
    return fail(SDR_E_BAD_WIDTH, "DECIMAL of %u bytes has no Arrow layout (4, 8 or 16)", c.value_bytes);
  if (c.type == SDR_ARROW_DECIMAL && (c.precision == 0 || c.scale > c.precision))
    return fail(SDR_E_INVALID_ARG, "DECIMAL(%u,%u)", c.precision, c.scale);
  if (c.length < 0) return fail(SDR_E_INVALID_ARG, "negative length");
  if (c.length > INT32_MAX && c.type == SDR_ARROW_UTF8)
    return fail(SDR_E_INVALID_ARG, "UTF8 of %lld rows passes int32 offsets", static_cast<long long>(c.length));
  if (c.null_count < -1 || c.null_count > c.length) return fail(SDR_E_INVALID_ARG, "null_count out of range");
  if (!c.validity && c.null_count > 0) return fail(SDR_E_INVALID_ARG, "nulls without a validity bitmap");
  // UTF8 always has offsets[0]; fixed width needs values once there are rows
  if (!c.values && (c.length > 0 || c.type == SDR_ARROW_UTF8)) return fail(SDR_E_INVALID_ARG, "values is NULL");
  if (c.type == SDR_ARROW_UTF8 && !c.chars) return fail(SDR_E_INVALID_ARG, "UTF8 chars is NULL");
  if (c.device_type != ARROW_DEVICE_ROCM && c.device_type != ARROW_DEVICE_CPU)
    return fail(SDR_E_INVALID_ARG, "device type %d (ROCm or CPU)", c.device_type);

  auto* ss = new (std::nothrow) SchemaState();
  if (!ss) return fail(SDR_E_CAPACITY, "out of host memory");
  if (!format_of(c, ss->format)) {  // unreachable after the range check
    delete ss;
    return fail(SDR_E_INVALID_ARG, "unknown Arrow type %u", c.type);
  }
  ss->name = c.name ? strdup(c.name) : nullptr;
  if (c.name && !ss->name) {
    delete ss;
    return fail(SDR_E_CAPACITY, "out of host memory");
  }
  auto* as = new (std::nothrow) ArrayState();
  if (!as) {
    std::free(ss->name);
    delete ss;
    return fail(SDR_E_CAPACITY, "out of host memory");
  }
  as->buffers[0] = c.validity;
  as->buffers[1] = c.values;
  as->buffers[2] = c.chars;
  as->release    = c.release;
  as->owner      = c.owner;

  *schema              = ArrowSchema{};
  schema->format       = ss->format;
  schema->name         = ss->name;
  schema->metadata     = nullptr;
  schema->flags        = ARROW_FLAG_NULLABLE;
  schema->n_children   = 0;
  schema->children     = nullptr;
  schema->dictionary   = nullptr;
  schema->release      = release_schema;
  schema->private_data = ss;

  *array                    = ArrowDeviceArray{};
  ArrowArray& a             = array->array;
  a.length                  = c.length;
  a.null_count              = c.validity ? c.null_count : 0;
  a.offset                  = 0;
  a.n_buffers               = c.type == SDR_ARROW_UTF8 ? 3 : 2;
  a.n_children              = 0;
  a.buffers                 = as->buffers;
  a.children                = nullptr;
  a.dictionary              = nullptr;
  a.release                 = release_array;
  a.private_data            = as;
  array->device_id          = c.device_id;
  array->device_type        = c.device_type;
  array->sync_event         = c.sync_event;
  return SDR_OK;
}
