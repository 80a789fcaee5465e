/*
 * sdr/decode.h — C ABI of the MI355X DuckDB segment decoder.
 *
 * Plain C, no STL, no torch or HIP types: a stream is passed as an opaque
 * `sdr_stream_t` (a hipStream_t).  Every pointer named `d_*` is device memory
 * on the current HIP device; the caller owns every output buffer.
 *
 * Each entry point replaces one function of the reference's decode-operator
 * interface (ducndh/sirius); the file:line it replaces is given beside it.
 * Conventions shared by every entry:
 *   - the return value is an sdr_status; SDR_OK is 0, errors are negative and
 *     `sdr_last_error()` returns a thread-local message for the last failure;
 *   - viability violations (unsupported codec or width, rows out of bounds,
 *     misaligned validity, string offsets past INT32_MAX) are returned as
 *     errors where the reference throws std::runtime_error;
 *   - malformed segment bytes never fault: the affected rows decode to zero
 *     (strings: to empty), as in the reference kernels;
 *   - all device work is enqueued on `stream`; the entries may be called
 *     from several host threads at once, on the same or different streams.
 *     The pinned staging and descriptor slots they share are chosen under
 *     short internal locks that make no blocking HIP call (waits, growth and
 *     copies run on a slot the caller holds), so one caller's waits never
 *     stall another's.  A batch, plan or server object is used from one
 *     thread at a time.
 */
#ifndef SDR_DECODE_H
#define SDR_DECODE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDR_ABI_VERSION 1

typedef void* sdr_stream_t; /* hipStream_t */

typedef enum sdr_status {
  SDR_OK                  = 0,
  SDR_E_INVALID_ARG       = -1, /* null pointer, zero type size, bad enum      */
  SDR_E_UNSUPPORTED_CODEC = -2, /* codec not decodable on this path            */
  SDR_E_BAD_WIDTH         = -3, /* type_size not supported by the codec        */
  SDR_E_BOUNDS            = -4, /* row range outside [0,total_rows), short buf */
  SDR_E_OVERFLOW          = -5, /* chars past capacity or INT32_MAX            */
  SDR_E_ALIGNMENT         = -6, /* validity row_offset not a multiple of 8     */
  SDR_E_HIP               = -7, /* HIP runtime error (message has the name)    */
  SDR_E_CAPACITY          = -8, /* queue / pool / scratch capacity exceeded    */
  SDR_E_NOT_FOUND         = -9, /* unknown handle                              */
  SDR_E_STATE             = -10 /* call not valid in the object's state        */
} sdr_status;

/* Codec ids.  The numbers follow DuckDB's CompressionType order; callers that
 * read DuckDB's pragma_storage_info get the id from the codec's name through
 * sdr_codec_from_name ("Uncompressed", "Constant", "RLE", "Dictionary",
 * "BitPacking", "FSST", "ALP", "ALPRD", "DICT_FSST"). */
typedef enum sdr_codec {
  SDR_CODEC_UNCOMPRESSED = 1,
  SDR_CODEC_CONSTANT     = 2,
  SDR_CODEC_RLE          = 3,
  SDR_CODEC_DICTIONARY   = 4,
  SDR_CODEC_BITPACKING   = 6,
  SDR_CODEC_FSST         = 7,
  SDR_CODEC_ALP          = 10,
  SDR_CODEC_ALPRD        = 11,
  SDR_CODEC_DICT_FSST    = 15
} sdr_codec;

/* One DuckDB column segment already staged on the device.
 * Replaces gpu_segment_desc (src/include/cuda/scan/gpu_native_decode.cuh:50-55)
 * and gpu_string_segment_desc (src/include/cuda/scan/gpu_decode_strings.cuh:41-49).
 * `bytes_size` is an upper bound on the segment's bytes (it may include block
 * slack); decoders bound every read by it.  `max_string_length` is the
 * segment-statistics bound on decoded chars per row (strings only; 0 means
 * unknown).
 *
 * The reference's string descriptor also carries `seg_row_start`, rows
 * skipped at the head of a segment that a caller slices into chunks.  Its
 * one production caller sets it to 0 (duckdb_native_decoder.cpp:914), and
 * only the DICT_FSST and UNCOMPRESSED kernels read it.  This ABI has no such
 * field: the decoders split segments into chunks themselves, and a caller
 * that wants rows [k, n) of a segment decodes the whole segment and takes
 * the rows from the output (offsets rebased by the chars before row k). */
typedef struct sdr_segment {
  const uint8_t* d_bytes;
  uint32_t bytes_size;
  uint32_t row_offset;
  uint32_t row_count;
  uint32_t max_string_length;
} sdr_segment;

/* A run of segments sharing one codec.
 * Replaces gpu_codec_run (gpu_native_decode.cuh:60-63) and
 * gpu_string_codec_run (gpu_decode_strings.cuh:52-55). */
typedef struct sdr_run {
  uint32_t codec; /* sdr_codec */
  uint32_t n_segments;
  const sdr_segment* segments;
} sdr_run;

/* One fixed-width column: inputs, caller-owned outputs, and the null count.
 * Replaces gpu_column_decode_input (gpu_native_decode.cuh:67-73) plus the
 * cudf::column the reference allocates for it.
 *   d_out        total_rows * type_size bytes
 *   d_null_mask  ceil(total_rows/8) bytes (bit set = row valid), or NULL;
 *                written only when has_nulls and validity runs are given
 *   null_count   out: number of null rows (0 when no mask is built) */
typedef struct sdr_column {
  uint32_t type_size; /* bytes per value: 1, 2, 4, 8 (16 for UNCOMPRESSED/CONSTANT) */
  uint32_t total_rows;
  const sdr_run* data;
  uint32_t n_data;
  uint32_t has_nulls;
  const sdr_run* validity;
  uint32_t n_validity;
  uint32_t reserved;
  void* d_out;
  uint8_t* d_null_mask;
  int64_t null_count;
} sdr_column;

/* One VARCHAR column.  Replaces gpu_string_column_decode_input
 * (gpu_decode_strings.cuh:60-65) plus the strings column the reference builds.
 *   d_offsets    int32[total_rows + 1]
 *   d_chars      chars_capacity bytes
 *   chars_used   out: total decoded chars (offsets[total_rows]); when it
 *                exceeds chars_capacity the call returns SDR_E_OVERFLOW
 *                after writing the offsets, and d_chars is left untouched */
typedef struct sdr_string_column {
  uint32_t total_rows;
  uint32_t has_nulls;
  const sdr_run* data;
  uint32_t n_data;
  uint32_t n_validity;
  const sdr_run* validity;
  int32_t* d_offsets;
  uint8_t* d_chars;
  uint64_t chars_capacity;
  uint8_t* d_null_mask;
  uint64_t chars_used;
  int64_t null_count;
} sdr_string_column;

/* Launch strategies for the fixed-width and string paths. */
typedef enum sdr_launch_mode {
  SDR_LAUNCH_BATCHED     = 0, /* one launch per codec and type size (default)  */
  SDR_LAUNCH_PER_SEGMENT = 1, /* one launch per segment (the plain baseline)   */
  SDR_LAUNCH_PERSISTENT  = 2  /* one persistent mixed-codec launch (sdr_batch) */
} sdr_launch_mode;

/* ---- library ----------------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none). */
const char* sdr_last_error(void);
/* Codec id for a DuckDB storage-info name, or SDR_E_UNSUPPORTED_CODEC. */
int sdr_codec_from_name(const char* name);
/* DuckDB storage-info name of a codec id, or "" */
const char* sdr_codec_name(uint32_t codec);
/* ABI version and gfx target the library was built for. */
int sdr_abi_version(void);
const char* sdr_build_target(void);

/* ---- fixed-width decode ------------------------------------------------ */

/* Decode every column (all must share total_rows).  Replaces
 * gpu_decode_table (src/include/cuda/scan/gpu_native_decode.cuh:84-86,
 * src/cuda/scan/gpu_native_decode.cu:546-608).  Synchronises `stream` once
 * when a null mask is built (to return null_count), as the reference does in
 * its batched null count (gpu_native_decode.cu:590-591); otherwise the call
 * is asynchronous. */
int sdr_decode_table(sdr_column* cols, uint32_t n_cols, sdr_stream_t stream);
/* Same, with an explicit launch strategy. */
int sdr_decode_table_mode(sdr_column* cols, uint32_t n_cols, sdr_stream_t stream,
                          uint32_t launch_mode);

/* Per-codec entries: decode one run into d_out (sized for the column; each
 * segment writes row_count values at row_offset).  Replace
 * decode_bitpacking_data (src/include/cuda/scan/gpu_decode_bitpacking.cuh:74-79),
 * decode_rle_data (gpu_decode_rle.cuh:35-40), decode_alp_data and
 * decode_alprd_data (gpu_decode_alp.cuh:50-62). */
int sdr_decode_bitpacking_data(const sdr_run* run, void* d_out, uint32_t type_size,
                               sdr_stream_t stream);
int sdr_decode_rle_data(const sdr_run* run, void* d_out, uint32_t type_size, sdr_stream_t stream);
int sdr_decode_alp_data(const sdr_run* run, void* d_out, uint32_t type_size, sdr_stream_t stream);
int sdr_decode_alprd_data(const sdr_run* run, void* d_out, uint32_t type_size,
                          sdr_stream_t stream);

/* ---- prepared plans ---------------------------------------------------- */

/* A plan uploads the segment descriptors of a set of fixed-width columns
 * once; sdr_plan_launch then only enqueues kernels (no allocation, no host
 * copy, no synchronisation), so it can be repeated and captured in a
 * hipGraph.  It has no counterpart in the reference, whose gpu_decode_table
 * rebuilds and uploads descriptors on every call
 * (src/cuda/scan/gpu_native_decode.cu:546-608); a plan's launches decode
 * exactly what sdr_decode_table_mode would.  The column structs, segment
 * bytes and outputs must stay valid while the plan is used.  Null masks
 * are rebuilt by every launch; null counts stay on the device until
 * sdr_plan_null_counts (which synchronises `stream`). */
typedef struct sdr_plan sdr_plan;

int sdr_plan_create(const sdr_column* cols, uint32_t n_cols, uint32_t launch_mode, sdr_stream_t stream,
                    sdr_plan** out);
int sdr_plan_launch(sdr_plan* plan, sdr_stream_t stream);
/* Launch only the kernels of one codec (codec 0 = all). */
int sdr_plan_launch_codec(sdr_plan* plan, uint32_t codec, sdr_stream_t stream);
/* Number of kernel launches one sdr_plan_launch enqueues. */
int sdr_plan_num_launches(const sdr_plan* plan);
int sdr_plan_null_counts(sdr_plan* plan, int64_t* counts, uint32_t n_cols, sdr_stream_t stream);
int sdr_plan_destroy(sdr_plan* plan);

/* ---- persistent batches ------------------------------------------------- */

/* A batch collects the work of many columns (fixed-width and VARCHAR, any
 * codec mix) into one launch: the fixed-width segments of each codec and
 * value width decode in one per-codec grid over descriptors in the batch's
 * arena (or, with SDR_BATCH_FIXED_KERNELS = 1, in persistent group kernels
 * that drain a work queue), and the VARCHAR chunks in a kernel of one
 * block per chunk that claims them from the queue in order.  The arena is
 * uploaded from pinned memory with one copy ahead of the kernels.  A batch of fixed-width segments alone
 * (no string column, validity mask or queue entry) passes its descriptors as
 * sdr_decode_table does instead: up to 256 per per-codec grid in the kernel
 * arguments, more through a reused descriptor buffer; it is never replayed
 * from a graph.  It has no
 * counterpart in the reference, which launches one kernel per codec run
 * (src/cuda/scan/gpu_native_decode.cu:356-386, gpu_decode_strings.cu:257-366);
 * every output is identical to sdr_decode_table / sdr_decode_strings_column.
 * Columns, segment bytes and outputs must stay valid until sdr_batch_finish.
 * A batch is used from one thread; its launches are ordered on the stream
 * of the previous launch. */
typedef struct sdr_batch sdr_batch;

typedef enum sdr_batch_option {
  SDR_BATCH_USE_GRAPH          = 1,  /* 1: replay a recorded hipGraph; 0 (default) */
                                     /*   enqueue the kernels directly (faster on  */
                                     /*   gfx950: no graph-launch overhead)        */
  SDR_BATCH_GRID               = 2,  /* blocks of each fixed-width group kernel;   */
                                     /*   0 (default): resident capacity (the FSST */
                                     /*   kernel runs one block per chunk)         */
  SDR_BATCH_FIXED_KERNELS      = 3,  /* fixed-width segments: 0 (default) the      */
                                     /*   per-codec kernels, one grid per codec    */
                                     /*   and value width over descriptors in the  */
                                     /*   batch; 1 persistent group kernels that   */
                                     /*   drain queue entries (SDR_BATCH_UNITS_*,  */
                                     /*   SDR_BATCH_GRID apply)                    */
  SDR_BATCH_UNITS_BITPACKING   = 16, /* work units per queue entry, per codec:     */
  SDR_BATCH_UNITS_RLE          = 17, /*   2048-row groups (default 16), RLE slices */
  SDR_BATCH_UNITS_ALP          = 18, /*   of rle_slice_rows (2), 1024-row vectors  */
  SDR_BATCH_UNITS_ALPRD        = 19, /*   (16 each), 16 KiB byte ranges (4 each)   */
  SDR_BATCH_UNITS_UNCOMPRESSED = 20,
  SDR_BATCH_UNITS_CONSTANT     = 21
} sdr_batch_option;

int sdr_batch_create(sdr_batch** out);
int sdr_batch_set_option(sdr_batch* batch, uint32_t key, int64_t value);
int sdr_batch_add_table(sdr_batch* batch, sdr_column* cols, uint32_t n_cols);
/* A column with DICT_FSST segments is refused (SDR_E_UNSUPPORTED_CODEC):
 * its dictionaries are sized by host round trips a queued launch cannot
 * make; sdr_decode_strings_column decodes it. */
int sdr_batch_add_strings(sdr_batch* batch, sdr_string_column* col);
int sdr_batch_num_entries(const sdr_batch* batch);
/* Upload the queue and enqueue the persistent decode on `stream`. */
int sdr_batch_launch(sdr_batch* batch, sdr_stream_t stream);
/* Wait for the last launch, then write chars_used and null counts back into
 * the added columns; reports string overflow like sdr_decode_strings_column.
 * (`stream` is unused: the wait is on the launch itself.) */
int sdr_batch_finish(sdr_batch* batch, sdr_stream_t stream);
/* Forget the added columns (the device arena and graph are kept); the
 * write-back of a launch that was not finished is dropped. */
int sdr_batch_clear(sdr_batch* batch);
int sdr_batch_destroy(sdr_batch* batch);

/* ---- serving ------------------------------------------------------------ */

/* A decode server: requests from concurrent queries arrive over time and
 * decode segments that stay resident on the device.  The server
 *   - keeps one device memory pool for resident segments and request
 *     outputs: slabs reserved when the pool grows, carved on the host, and
 *     ranges released behind an event that rejoin the pool once it has
 *     passed, so a warm server makes no allocation call,
 *   - stages new segment bytes from host memory through a ring of pinned
 *     buffers with asynchronous copies on its own stream
 *     (sdr_server_stage); its launches wait for those copies on the device,
 *   - admits queued requests in arrival order while resident bytes plus the
 *     outputs of admitted, unreleased requests fit the HBM budget (a request
 *     larger than the budget is admitted alone, when nothing else is held),
 *   - and, in SDR_LAUNCH_PERSISTENT mode, coalesces every admitted request
 *     into the next persistent launch (the batch's group kernels, enqueued
 *     directly on the slot's stream) while up to `max_inflight`
 *     launches run on their own streams: requests that arrive while others
 *     decode join the next launch as soon as a launch slot and budget are
 *     free (continuous batching).  SDR_LAUNCH_BATCHED and
 *     SDR_LAUNCH_PER_SEGMENT decode one request per launch slot with the
 *     drop-in entries (the baselines).
 * No counterpart in the reference (it decodes one split per call:
 * src/op/scan/duckdb_native_decoder.cpp:972-1217, staging each split's
 * segments with its own copies at :712-855); each request's output is
 * what sdr_decode_table / sdr_decode_strings_column return for it.  A server
 * is used from one thread. */
typedef struct sdr_server sdr_server;

typedef struct sdr_server_config {
  uint64_t budget_bytes;       /* HBM admission budget; 0 = 200 GB            */
  uint32_t max_inflight;       /* launches in flight at once; 0 = 2           */
  uint32_t mode;               /* sdr_launch_mode                             */
  uint32_t max_batch_requests; /* requests per persistent launch; 0 = no cap  */
  uint32_t reserved;
} sdr_server_config;

/* One decode request: a fixed-width or VARCHAR column over resident
 * segments.  The descriptors are copied at submit. */
typedef struct sdr_request {
  uint64_t tag;             /* caller's id (unique among pending requests)       */
  uint32_t is_string;       /* 0 fixed-width, 1 VARCHAR                          */
  uint32_t type_size;       /* fixed-width bytes per value                       */
  uint32_t total_rows;
  uint32_t n_data;
  const sdr_run* data;
  uint64_t chars_capacity;  /* VARCHAR chars to reserve; 0 = sdr_strings_upper_bound */
} sdr_request;

/* A finished request.  Times are milliseconds on the server clock (0 at
 * creation): submitted, launch start and launch end on the GPU. */
typedef struct sdr_completion {
  uint64_t tag;
  int32_t status;           /* sdr_status of the decode                          */
  uint32_t reserved;
  void* d_out;              /* values, or int32 offsets for VARCHAR              */
  uint8_t* d_chars;         /* VARCHAR chars                                     */
  uint64_t chars_used;
  uint64_t out_bytes;       /* output bytes held until sdr_server_release        */
  double submit_ms, start_ms, done_ms;
} sdr_completion;

/* Where a server's time went since it was created: host milliseconds spent
 * in each phase of its calls on the calling thread, and the GPU time of its
 * launches (start to end event; `gpu_union_ms` counts time during which at
 * least one launch ran once).  Decoded bytes over gpu_union_ms is the rate
 * the device sustained while it had work. */
typedef struct sdr_server_stats {
  uint64_t launches;
  uint64_t requests;          /* completed                                       */
  uint64_t queue_entries;     /* persistent-queue entries launched               */
  double host_submit_ms;      /* sdr_server_submit                               */
  double host_collect_ms;     /* polling and completing launches                 */
  double host_alloc_ms;       /* output allocation                               */
  double host_launch_ms;      /* queue build, upload and kernel enqueue          */
  double host_release_ms;     /* sdr_server_release                              */
  double host_stage_ms;       /* sdr_server_stage / sdr_server_unstage           */
  double gpu_busy_ms;         /* sum of launch durations                         */
  double gpu_union_ms;        /* union of launch intervals                       */
} sdr_server_stats;

int sdr_server_create(const sdr_server_config* config, sdr_server** out);
/* Phase times and launch counts since creation. */
int sdr_server_get_stats(const sdr_server* server, sdr_server_stats* out);
/* Copy n host bytes into the server's pool, where they stay until
 * sdr_server_unstage (counted against the budget); *d_out receives the
 * device address.  The host buffer may be reused when the call returns; the
 * device copy is asynchronous: the server's own launches are ordered after
 * it, any other user of *d_out calls sdr_server_stage_wait first. */
int sdr_server_stage(sdr_server* server, const void* host_bytes, uint64_t n, const uint8_t** d_out);
/* Make `stream` wait (on the device) for every staging copy issued so far. */
int sdr_server_stage_wait(sdr_server* server, sdr_stream_t stream);
/* Give a staged range back: it rejoins the pool once the launches in flight
 * have finished.  Requests still queued must not reference it. */
int sdr_server_unstage(sdr_server* server, const uint8_t* d_ptr);
/* Queue a request (host only); SDR_E_CAPACITY if its outputs exceed the
 * budget left after the resident segments. */
int sdr_server_submit(sdr_server* server, const sdr_request* request);
/* Collect finished launches, admit, and launch what slots allow.
 * *n_launched (optional) receives the number of launches started. */
int sdr_server_progress(sdr_server* server, uint32_t* n_launched);
/* Completions since the last poll, at most `cap`; *n receives the count. */
int sdr_server_poll(sdr_server* server, sdr_completion* out, uint32_t cap, uint32_t* n);
/* Return a completed request's outputs to the pool, ordered after the work
 * already enqueued on `stream` (the consumer's stream): the range is reused
 * only once that work has finished.  No HIP call is made here. */
int sdr_server_release(sdr_server* server, uint64_t tag, sdr_stream_t stream);
/* Requests queued, admitted-not-launched and in flight. */
int sdr_server_pending(const sdr_server* server);
/* Bytes currently held against the budget (resident + outputs). */
uint64_t sdr_server_held_bytes(const sdr_server* server);
/* Milliseconds on the server clock. */
double sdr_server_now_ms(const sdr_server* server);
int sdr_server_destroy(sdr_server* server);

/* ---- launch-shape tuning ------------------------------------------------ */

/* Process-wide launch-shape knobs, read whenever kernels are launched, a
 * plan is created or work is added to a batch.  They change speed only,
 * never results.  No counterpart in the reference, whose launch shapes are
 * compile-time constants (src/cuda/scan/gpu_decode_bitpacking.cu:394-427,
 * gpu_decode_rle.cu:518-560, gpu_decode_alp.cu:680-743).  The autotuner
 * (sirius_amd/autotune.py) measures candidates on the GPU and writes a
 * table that sdr_tune_load applies. */
typedef enum sdr_tune_key {
  SDR_TUNE_RLE_SLICE_ROWS      = 1, /* rows one 256-thread block expands:
                                       1024..1048576, multiple of 64 (16384) */
  SDR_TUNE_BP_GROUPS_PER_WAVE  = 2, /* 2048-row groups one wavefront decodes
                                       in turn, 1..16 (1)                    */
  SDR_TUNE_BP_CHUNK_ROWS       = 3, /* most rows a wavefront stages per LDS
                                       round: 512, 1024 or 2048 (2048)       */
  SDR_TUNE_ALP_VECTORS_PER_WAVE = 4, /* 1024-row ALP/ALPRD vectors one
                                       wavefront decodes in turn, 1..16 (1)  */
  SDR_TUNE_FSST_SPIN_LIMIT     = 5, /* sleeps a VARCHAR chunk waits for its
                                       predecessors' chars offsets before its
                                       decode reports SDR_E_HIP, 0..2^31
                                       (2^24, never reached); 0 reports at
                                       once without waiting (for tests)      */
  /* work units per queue entry of a persistent batch, for every batch that
   * does not set SDR_BATCH_UNITS_* itself: 1..2^20 */
  SDR_TUNE_BATCH_UNITS_BITPACKING = 6, /* 2048-row groups (16)               */
  SDR_TUNE_BATCH_UNITS_RLE      = 7,   /* RLE slices (2)                     */
  SDR_TUNE_BATCH_UNITS_ALP      = 8,   /* 1024-row vectors (16)              */
  SDR_TUNE_BATCH_UNITS_ALPRD    = 9,   /* 1024-row vectors (16)              */
  SDR_TUNE_BATCH_UNITS_COPY     = 10   /* 16 KiB UNCOMPRESSED / CONSTANT
                                          byte ranges (4)                    */
} sdr_tune_key;

/* Every knob holds one value per value width (1, 2, 4 and 8 bytes): the
 * value used for a column of that type size.  sdr_tune_set sets all four;
 * SDR_E_INVALID_ARG for an unknown key or a value out of range. */
int sdr_tune_set(uint32_t key, uint64_t value);
/* Set one width's value (SDR_E_BAD_WIDTH for a width other than 1, 2, 4, 8). */
int sdr_tune_set_width(uint32_t key, uint32_t type_size, uint64_t value);
/* Current value of a knob for 4-byte values, 0 for an unknown key. */
uint64_t sdr_tune_get(uint32_t key);
/* Current value for `type_size`-byte values, 0 for an unknown key or width. */
uint64_t sdr_tune_get_width(uint32_t key, uint32_t type_size);
/* Apply a table of "name value" lines (every width) and "name@ts value"
 * lines (one width; names: rle_slice_rows, bp_groups_per_wave,
 * bp_chunk_rows, alp_vectors_per_wave, fsst_spin_limit,
 * batch_units_{bitpacking,rle,alp,alprd,copy}; '#' starts a comment).
 * Stops at the first bad line with SDR_E_INVALID_ARG. */
int sdr_tune_load(const char* path);

/* ---- string decode ----------------------------------------------------- */

/* Upper bound on decoded chars: sum(row_count * max_string_length), or
 * UINT64_MAX if any segment reports max_string_length == 0.  Mirrors the
 * reference's host sizing (src/cuda/scan/gpu_decode_strings.cu:183-191). */
uint64_t sdr_strings_upper_bound(const sdr_run* runs, uint32_t n_runs);

/* Decode one VARCHAR column.  Replaces gpu_decode_strings_column
 * (src/include/cuda/scan/gpu_decode_strings.cuh:71-73,
 * src/cuda/scan/gpu_decode_strings.cu:100-402).  Runs may mix FSST,
 * DICTIONARY, UNCOMPRESSED and DICT_FSST segments (src/cuda/scan/strings/
 * fsst.cu, dictionary.cu, uncompressed.cu, dict_fsst.cu).  Synchronises
 * `stream` once to read back chars_used; a column with DICT_FSST segments
 * synchronises twice more, for their headers and then their dictionaries'
 * decoded sizes (the reference's prepare_dict_fsst makes the same two
 * round trips, dict_fsst.cu:471-690), and decodes through the per-column
 * passes in every launch mode.  DICT_FSST rows that select entry 0 are
 * NULL when any of the column's DICT_FSST segments has inline NULLs
 * (dict_fsst.cu:238-243, :426-446): the mask is then built (validity runs,
 * if any, and those rows) and d_null_mask is required, else the call
 * returns SDR_E_INVALID_ARG before writing anything. */
int sdr_decode_strings_column(sdr_string_column* col, sdr_stream_t stream);
int sdr_decode_strings_column_mode(sdr_string_column* col, sdr_stream_t stream,
                                   uint32_t launch_mode);

/* Bytes of the fixed per-device scratch array of
 * sdr_decode_strings_column_noalloc. */
#define SDR_NOALLOC_SCRATCH_BYTES (16u << 20)

/* sdr_decode_strings_column that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Outputs are the caller's, as in sdr_decode_strings_column;
 * the decode's own device memory (range descriptors, chunk descriptors,
 * symbol tables, look-back state, null counter) lives in one array of
 * SDR_NOALLOC_SCRATCH_BYTES in the library's code object, which each GPU the
 * library runs on holds until the process ends.  A column needing more
 * (about 2.3 KiB per segment plus 56 bytes per 1024 rows) returns
 * SDR_E_CAPACITY before any device work.  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU, on any streams or threads,
 * race and must be serialised by the caller.  Synchronises `stream` once,
 * as sdr_decode_strings_column does.  DICT_FSST segments, whose
 * dictionaries are sized at run time, are refused (SDR_E_UNSUPPORTED_CODEC). */
int sdr_decode_strings_column_noalloc(sdr_string_column* col, sdr_stream_t stream);

/* ---- row ids --------------------------------------------------------------- */

/* Rows [first_row, first_row + count) of a table, written as int64 to
 * d_out[out_offset, out_offset + count). */
typedef struct sdr_row_range {
  int64_t first_row;
  uint64_t out_offset;
  uint32_t count;
  uint32_t reserved;
} sdr_row_range;

/* Fill the row-id column of a scan (one range per row group).  Replaces
 * build_rowid_column (src/op/scan/duckdb_native_decoder.cpp:924-949), a
 * sequence per row group then a concatenation; here one launch writes every
 * range in place. */
int sdr_rowids(const sdr_row_range* ranges, uint32_t n, int64_t* d_out, sdr_stream_t stream);

/* ---- staging segment bytes from a file ----------------------------------- */

/* One byte range of a file (a segment).  device_offset is an output: the
 * range's position in the staged device image. */
typedef struct sdr_extent {
  uint64_t file_offset;
  uint64_t bytes;
  uint64_t device_offset;
} sdr_extent;

/* Lay out `n` extents: sorted by file offset, ranges whose gap is at most
 * `max_gap` bytes share a piece, pieces are packed 16-byte aligned.  Fills
 * every device_offset and returns the image size in bytes.  Host only. */
uint64_t sdr_stage_plan(sdr_extent* extents, uint32_t n, uint64_t max_gap);

/* Read the extents of the file at `path` into pinned host memory (one read
 * per piece) and copy the image to `d_dst` (dst_capacity bytes) with one
 * host-to-device copy on `stream`; device_offset as sdr_stage_plan.  The
 * reads complete before return; the copy is asynchronous.  Replaces the
 * reference's submit_and_await (src/op/scan/duckdb_native_decoder.cpp:
 * 675-855), which copies each segment separately. */
int sdr_stage_file(const char* path, sdr_extent* extents, uint32_t n, uint64_t max_gap, void* d_dst,
                   uint64_t dst_capacity, sdr_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* SDR_DECODE_H */
