/* sdr/arrow.h — decoded columns handed to consumers as Arrow C device arrays.
 *
 * The reference hands its decoded buffers downstream as cuDF columns: the
 * fixed-width decode returns one cudf::column per input column
 * (gpu_native_decode.cu:546-608), the strings decode a strings column built
 * on its offsets and chars (gpu_decode_strings.cu:100-402), and the scan
 * hands the table to filter and projection
 * (duckdb_native_gpu_ingestible.cpp:345-388).  Without cuDF, the zero-copy
 * hand-off here is the Arrow C Device Data Interface: an ArrowSchema plus an
 * ArrowDeviceArray whose buffers are the decode's own device buffers
 * (validity bitmap, values or int32 offsets, chars), on ROCm device memory.
 * Nothing is copied; the consumer's release of the array calls back the
 * buffers' owner.
 *
 * The struct layouts below are the ones the Arrow specification fixes
 * (C Data Interface and C Device Data Interface, ABI-stable), under the
 * include guards the specification asks for, so they coexist with Arrow's
 * own headers.
 */
#ifndef SDR_ARROW_H
#define SDR_ARROW_H

#include <stdint.h>

#include "decode.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef ARROW_C_DATA_INTERFACE
#define ARROW_C_DATA_INTERFACE

#define ARROW_FLAG_DICTIONARY_ORDERED 1
#define ARROW_FLAG_NULLABLE 2
#define ARROW_FLAG_MAP_KEYS_SORTED 4

struct ArrowSchema {
  const char* format;
  const char* name;
  const char* metadata;
  int64_t flags;
  int64_t n_children;
  struct ArrowSchema** children;
  struct ArrowSchema* dictionary;
  void (*release)(struct ArrowSchema*);
  void* private_data;
};

struct ArrowArray {
  int64_t length;
  int64_t null_count;
  int64_t offset;
  int64_t n_buffers;
  int64_t n_children;
  const void** buffers;
  struct ArrowArray** children;
  struct ArrowArray* dictionary;
  void (*release)(struct ArrowArray*);
  void* private_data;
};

#endif /* ARROW_C_DATA_INTERFACE */

#ifndef ARROW_C_DEVICE_DATA_INTERFACE
#define ARROW_C_DEVICE_DATA_INTERFACE

typedef int32_t ArrowDeviceType;
#define ARROW_DEVICE_CPU 1
#define ARROW_DEVICE_ROCM 10

struct ArrowDeviceArray {
  struct ArrowArray array;
  int64_t device_id;
  ArrowDeviceType device_type;
  void* sync_event;
  int64_t reserved[3];
};

#endif /* ARROW_C_DEVICE_DATA_INTERFACE */

/* Logical types of decoded DuckDB columns and their Arrow formats. */
typedef enum sdr_arrow_type {
  SDR_ARROW_INT8 = 1,      /* "c"   TINYINT                         */
  SDR_ARROW_INT16 = 2,     /* "s"   SMALLINT                        */
  SDR_ARROW_INT32 = 3,     /* "i"   INTEGER                         */
  SDR_ARROW_INT64 = 4,     /* "l"   BIGINT                          */
  SDR_ARROW_UINT8 = 5,     /* "C"   UTINYINT (and BOOLEAN's bytes)  */
  SDR_ARROW_UINT16 = 6,    /* "S"   USMALLINT                       */
  SDR_ARROW_UINT32 = 7,    /* "I"   UINTEGER                        */
  SDR_ARROW_UINT64 = 8,    /* "L"   UBIGINT                         */
  SDR_ARROW_FLOAT32 = 9,   /* "f"   FLOAT                           */
  SDR_ARROW_FLOAT64 = 10,  /* "g"   DOUBLE                          */
  SDR_ARROW_DATE32 = 11,   /* "tdD" DATE (int32 days)               */
  SDR_ARROW_TIMESTAMP_US = 12, /* "tsu:" TIMESTAMP (int64 us)       */
  SDR_ARROW_TIME_US = 13,  /* "ttu" TIME (int64 us)                 */
  SDR_ARROW_DECIMAL = 14,  /* "d:P,S[,bits]" DECIMAL, 4/8/16 bytes  */
  SDR_ARROW_UTF8 = 15,     /* "u"   VARCHAR: int32 offsets + chars  */
  SDR_ARROW_TIMESTAMP_S = 16,  /* "tss:" TIMESTAMP_S                */
  SDR_ARROW_TIMESTAMP_MS = 17, /* "tsm:" TIMESTAMP_MS               */
  SDR_ARROW_TIMESTAMP_NS = 18, /* "tsn:" TIMESTAMP_NS               */
  SDR_ARROW_TIMESTAMP_US_UTC = 19 /* "tsu:UTC" TIMESTAMP WITH TIME ZONE */
} sdr_arrow_type;

/* One decoded column, described for export.  Buffers are the decode's
 * outputs as they are: `validity` a bitmap with 1 = valid, least significant
 * bit first (the null mask sdr_decode_table and sdr_decode_strings_column
 * write), `values` the fixed-width values or, for UTF8, the int32
 * offsets[length + 1], `chars` the UTF8 bytes. */
typedef struct sdr_arrow_column {
  const char* name;             /* copied; may be NULL                        */
  uint32_t type;                /* sdr_arrow_type                             */
  uint32_t value_bytes;         /* DECIMAL: 4, 8 or 16; otherwise ignored     */
  uint32_t precision, scale;    /* DECIMAL                                    */
  int64_t length;               /* rows                                       */
  int64_t null_count;           /* -1: unknown                                */
  const void* validity;         /* NULL: no nulls (null_count must be 0 or -1) */
  const void* values;
  const void* chars;            /* UTF8 only                                  */
  int32_t device_type;          /* ARROW_DEVICE_ROCM, or ARROW_DEVICE_CPU     */
  int32_t reserved;
  int64_t device_id;
  void* sync_event;             /* a hipEvent_t* the consumer waits on, or NULL */
  void (*release)(void* owner); /* called once, when the consumer releases    */
  void* owner;
} sdr_arrow_column;

/* Fill *schema and *array for one column (host only: no device call).  On
 * success the consumer owns both and releases each through its release
 * callback; the array's release calls col->release(col->owner).  On error
 * neither is touched (SDR_E_INVALID_ARG for a NULL pointer, an unknown type,
 * a negative length, missing buffers or nulls without a bitmap;
 * SDR_E_BAD_WIDTH for a DECIMAL width other than 4, 8 or 16 bytes). */
int sdr_arrow_export(const sdr_arrow_column* col, struct ArrowSchema* schema, struct ArrowDeviceArray* array);

#ifdef __cplusplus
}
#endif

#endif /* SDR_ARROW_H */
