// sdr_batch.cpp — host side of the persistent mixed-codec decode: builds the
// device work queue (one arena: header, entries, FSST columns, chunks,
// segments, look-back state), uploads it from pinned memory with one copy,
// and launches the persistent kernel, optionally as a replayed hipGraph.
#include "sdr_batch.hpp"
#include "sdr_internal.hpp"
#include "sdr_tune.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

struct sdr_batch {
  // host-side description
  std::vector<sdr::BatchItem> items;
  struct Col {
    uint32_t n_chunks = 0;
    std::vector<sdr::FsstSeg> segs;  // row ranges: segments and the gaps between them
    sdr_string_column* user;  // chars_used / null_count written back at finish
    int32_t* offsets;
    uint8_t* chars;
    uint64_t cap;
    uint32_t total_rows;
    uint32_t item_first;  // first queue entry of this column
    int status = SDR_OK;  // this column's outcome, set by batch_collect
  };
  std::vector<Col> cols;
  struct MaskJob {
    sdr_column* fixed;        // one of the two is set
    sdr_string_column* str;
  };
  std::vector<MaskJob> masks;
  // device arena and pinned staging
  uint8_t* d_arena     = nullptr;
  size_t arena_bytes   = 0;
  uint8_t* h_stage     = nullptr;
  size_t stage_bytes   = 0;
  uint8_t* h_readback  = nullptr;  // pinned copy of the look-back state and null counters
  size_t readback_cap  = 0;
  size_t readback_bytes = 0;
  size_t used_bytes    = 0;
  size_t off_items = 0, off_cols = 0, off_chunks = 0, off_segs = 0, off_state = 0, off_nulls = 0;
  uint32_t n_status    = 0;
  hipEvent_t upload_done = nullptr;
  hipEvent_t launch_done = nullptr;
  bool upload_pending  = false;
  // launch configuration
  // work units per queue entry, per BatchKind.  Each entry costs one claim
  // (an atomic on the queue's one counter and two barriers); on the
  // integer/date mix 16 groups and 2 RLE slices per entry took a step from
  // 392 to 277 us against 4 and 1, and 16 ALP vectors 198 -> 152 us against 4.
  uint32_t units[8]    = {0, 16, 2, 16, 16, 4, 4, 1};
  uint32_t grid        = 0;                         // 0 = resident capacity
  bool use_graph       = true;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph     = nullptr;
  uint8_t* graph_arena = nullptr;
  uint32_t graph_grid  = 0;
  bool graph_strings   = false;
  hipStream_t last_stream = nullptr;
  bool launched           = false;  // a launch whose results are not collected yet
  ~sdr_batch();
};

sdr_batch::~sdr_batch()
{
  if (upload_done) {
    (void)hipEventSynchronize(upload_done);
    (void)hipEventDestroy(upload_done);
  }
  if (launch_done) {
    (void)hipEventSynchronize(launch_done);
    (void)hipEventDestroy(launch_done);
  }
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  if (d_arena) (void)hipFree(d_arena);
  if (h_stage) (void)hipHostFree(h_stage);
  if (h_readback) (void)hipHostFree(h_readback);
}

namespace sdr {

namespace {

BatchKind kind_of(uint32_t codec)
{
  switch (codec) {
    case SDR_CODEC_BITPACKING: return BK_BITPACKING;
    case SDR_CODEC_RLE: return BK_RLE;
    case SDR_CODEC_ALP: return BK_ALP;
    case SDR_CODEC_ALPRD: return BK_ALPRD;
    case SDR_CODEC_UNCOMPRESSED: return BK_COPY;
    case SDR_CODEC_CONSTANT: return BK_CONSTANT;
    default: return BatchKind(0);
  }
}

uint32_t units_in_segment(BatchKind k, uint32_t rows, uint32_t ts, uint32_t rle_slice)
{
  switch (k) {
    case BK_BITPACKING: return (rows + 2047u) / 2048u;
    case BK_RLE: return (rows + rle_slice - 1u) / rle_slice;
    case BK_ALP:
    case BK_ALPRD: return (rows + 1023u) / 1024u;
    default: return static_cast<uint32_t>((uint64_t{rows} * ts + COPY_UNIT_BYTES - 1u) / COPY_UNIT_BYTES);
  }
}

int wait_upload(sdr_batch* b)
{
  if (b->upload_pending) {
    HIP_TRY(hipEventSynchronize(b->upload_done));
    b->upload_pending = false;
  }
  return SDR_OK;
}

}  // namespace

int batch_add_fixed(sdr_batch* b, sdr_column* cols, uint32_t n_cols)
{
  for (uint32_t c = 0; c < n_cols; ++c) {
    const sdr_column& col = cols[c];
    for (uint32_t r = 0; r < col.n_data; ++r) {
      const sdr_run& run = col.data[r];
      const BatchKind k  = kind_of(run.codec);
      const uint32_t per = std::max<uint32_t>(1, b->units[k]);
      const uint32_t rle_slice = tune(SDR_TUNE_RLE_SLICE_ROWS);
      for (uint32_t s = 0; s < run.n_segments; ++s) {
        const sdr_segment& g = run.segments[s];
        if (!g.row_count) continue;
        const uint32_t nu = units_in_segment(k, g.row_count, col.type_size, rle_slice);
        for (uint32_t u = 0; u < nu; u += per) {
          BatchItem it{};
          it.seg        = g.d_bytes;
          it.out        = static_cast<uint8_t*>(col.d_out) + uint64_t{g.row_offset} * col.type_size;
          it.bytes_size = g.bytes_size;
          it.row_count  = g.row_count;
          it.unit0      = u;
          it.n_units    = std::min(per, nu - u);
          it.kind       = k;
          it.ts         = static_cast<uint8_t>(col.type_size);
          if (k == BK_RLE) it.reserved = static_cast<uint16_t>(rle_slice / 64u);
          b->items.push_back(it);
        }
      }
    }
    if (col.has_nulls && col.n_validity && col.total_rows) b->masks.push_back({&cols[c], nullptr});
  }
  return SDR_OK;
}

int batch_add_strings(sdr_batch* b, sdr_string_column* col)
{
  const uint32_t n = col->total_rows;
  sdr_batch::Col bc;
  bc.user       = col;
  bc.offsets    = col->d_offsets;
  bc.chars      = col->d_chars;
  bc.cap        = col->d_chars ? col->chars_capacity : 0;
  bc.total_rows = n;
  std::vector<std::pair<const sdr_segment*, uint32_t>> segs;  // segment, codec
  for (uint32_t r = 0; r < col->n_data; ++r)
    for (uint32_t k = 0; k < col->data[r].n_segments; ++k)
      if (col->data[r].segments[k].row_count) segs.push_back({&col->data[r].segments[k], col->data[r].codec});
  std::stable_sort(segs.begin(), segs.end(),
                   [](const auto& a, const auto& c) { return a.first->row_offset < c.first->row_offset; });
  auto add_range = [&](const uint8_t* bytes, uint32_t size, uint32_t col_row, uint32_t rows, uint32_t codec) {
    bc.segs.push_back({bytes, size, rows, bc.n_chunks, col_row, codec, 0});
    bc.n_chunks += fsst_chunk_count(rows);
  };
  uint32_t cursor = 0;
  for (const auto& sc : segs) {
    const sdr_segment* g = sc.first;
    if (g->row_offset > cursor) add_range(nullptr, 0, cursor, g->row_offset - cursor, 0);
    add_range(g->d_bytes, g->bytes_size, g->row_offset, g->row_count, sc.second);
    cursor = g->row_offset + g->row_count;
  }
  if (cursor < n) add_range(nullptr, 0, cursor, n - cursor, 0);
  bc.item_first = static_cast<uint32_t>(b->items.size());
  const uint32_t ci = static_cast<uint32_t>(b->cols.size());
  for (uint32_t k = 0; k < bc.n_chunks; ++k) {
    BatchItem it{};
    it.kind  = BK_FSST;
    it.chunk = k;
    it.unit0 = ci;  // resolved to a device column pointer at upload
    b->items.push_back(it);
  }
  if (col->has_nulls && col->d_null_mask) b->masks.push_back({nullptr, col});
  b->cols.push_back(std::move(bc));
  return SDR_OK;
}

// Lay out the arena, fill the pinned staging copy, upload it.
int batch_upload(sdr_batch* b, hipStream_t s)
{
  auto al = [](size_t v) { return (v + 255u) & ~size_t{255}; };
  size_t n_chunks = 0, n_segs = 0;
  for (auto& c : b->cols) {
    n_chunks += c.n_chunks;
    n_segs += c.segs.size();
  }
  // FsstState = {ticket, error, total, status[n]}: 16 + 8 n bytes per column
  size_t state_words = 0;
  for (auto& c : b->cols) state_words += 2 + c.n_chunks;
  size_t off                  = al(sizeof(BatchQueue));
  b->off_items                = off;
  off                         = al(off + b->items.size() * sizeof(BatchItem));
  b->off_cols                 = off;
  off                         = al(off + b->cols.size() * sizeof(FsstColumn));
  b->off_segs                 = off;
  off                         = al(off + n_segs * sizeof(FsstSeg));
  b->off_state                = off;
  off                         = al(off + state_words * sizeof(uint64_t));
  b->off_nulls                = off;
  off                         = al(off + std::max<size_t>(1, b->masks.size()) * sizeof(unsigned long long));
  const size_t readback_end   = off;
  // written on the device by the FSST segment pre-pass: neither uploaded nor read back
  b->off_chunks               = off;
  off                         = al(off + n_chunks * sizeof(FsstChunk));
  const size_t off_info       = off;
  off                         = al(off + n_chunks * sizeof(FsstChunkInfo));
  const size_t off_tabs       = off;
  off                         = al(off + n_segs * sizeof(FsstSymTab));
  const size_t upload_bytes   = b->off_state;  // the state and null counters are cleared on the device
  b->used_bytes               = off;
  if (off > b->arena_bytes) {
    if (b->d_arena) HIP_TRY(hipFree(b->d_arena));
    b->d_arena     = nullptr;
    b->arena_bytes = 0;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&b->d_arena), off + off / 4));
    b->arena_bytes = off + off / 4;
  }
  if (upload_bytes > b->stage_bytes) {
    if (b->h_stage) HIP_TRY(hipHostFree(b->h_stage));
    b->h_stage     = nullptr;
    b->stage_bytes = 0;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&b->h_stage), upload_bytes + upload_bytes / 4, hipHostMallocDefault));
    b->stage_bytes = upload_bytes + upload_bytes / 4;
  }
  b->readback_bytes = readback_end - b->off_state;  // state words and null counters
  if (b->readback_bytes > b->readback_cap) {
    if (b->h_readback) HIP_TRY(hipHostFree(b->h_readback));
    b->h_readback   = nullptr;
    b->readback_cap = 0;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&b->h_readback), b->readback_bytes, hipHostMallocDefault));
    b->readback_cap = b->readback_bytes;
  }
  uint8_t* h = b->h_stage;
  uint8_t* d = b->d_arena;
  BatchQueue q{};
  q.n_items       = static_cast<uint32_t>(b->items.size());
  q.n_fsst_status = static_cast<uint32_t>(state_words);
  q.fsst_status   = reinterpret_cast<uint64_t*>(d + b->off_state);
  q.n_fsst_segs   = static_cast<uint32_t>(n_segs);
  q.spin_limit    = tune(SDR_TUNE_FSST_SPIN_LIMIT);
  q.fsst_segs     = reinterpret_cast<const FsstSeg*>(d + b->off_segs);
  q.fsst_chunks   = reinterpret_cast<FsstChunk*>(d + b->off_chunks);
  q.fsst_info     = reinterpret_cast<FsstChunkInfo*>(d + off_info);
  q.fsst_tabs     = reinterpret_cast<FsstSymTab*>(d + off_tabs);
  std::memcpy(h, &q, sizeof(q));
  // FSST columns, chunks, segments, state
  size_t ck = 0, sg = 0, st = 0;
  std::vector<FsstColumn> dcols(b->cols.size());
  for (size_t i = 0; i < b->cols.size(); ++i) {
    auto& c = b->cols[i];
    FsstColumn fc{};
    fc.chunks     = reinterpret_cast<const FsstChunk*>(d + b->off_chunks) + ck;
    fc.segs       = reinterpret_cast<const FsstSeg*>(d + b->off_segs) + sg;
    fc.info       = reinterpret_cast<const FsstChunkInfo*>(d + off_info) + ck;
    fc.tabs       = reinterpret_cast<const FsstSymTab*>(d + off_tabs);  // chunk seg indices are batch-wide
    fc.state      = reinterpret_cast<FsstState*>(reinterpret_cast<uint64_t*>(d + b->off_state) + st);
    fc.offsets    = c.offsets;
    fc.chars      = c.chars;
    fc.chars_cap  = c.cap;
    fc.total_rows = c.total_rows;
    fc.n_chunks   = c.n_chunks;
    dcols[i]      = fc;
    auto* hs = reinterpret_cast<FsstSeg*>(h + b->off_segs) + sg;
    for (size_t k = 0; k < c.segs.size(); ++k) {  // first_chunk: an index into the batch's chunk array
      hs[k] = c.segs[k];
      hs[k].first_chunk += static_cast<uint32_t>(ck);
    }
    ck += c.n_chunks;
    sg += c.segs.size();
    st += 2 + c.n_chunks;
  }
  if (!dcols.empty()) std::memcpy(h + b->off_cols, dcols.data(), dcols.size() * sizeof(FsstColumn));
  auto* hi = reinterpret_cast<BatchItem*>(h + b->off_items);
  std::memcpy(hi, b->items.data(), b->items.size() * sizeof(BatchItem));
  for (size_t i = 0; i < b->items.size(); ++i)
    if (hi[i].kind == BK_FSST) hi[i].column = reinterpret_cast<const FsstColumn*>(d + b->off_cols) + hi[i].unit0;
  HIP_TRY(hipMemcpyAsync(d, h, upload_bytes, hipMemcpyHostToDevice, s));
  if (!b->upload_done) HIP_TRY(hipEventCreateWithFlags(&b->upload_done, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(b->upload_done, s));
  b->upload_pending = true;
  return SDR_OK;
}

int batch_launch(sdr_batch* b, hipStream_t s)
{
  clear_error();
  int rc = wait_upload(b);
  if (rc) return rc;
  if (b->launch_done && b->last_stream && b->last_stream != s) HIP_TRY(hipStreamWaitEvent(s, b->launch_done, 0));
  rc = batch_upload(b, s);
  if (rc) return rc;
  auto* q     = reinterpret_cast<BatchQueue*>(b->d_arena);
  auto* items = reinterpret_cast<const BatchItem*>(b->d_arena + b->off_items);
  const bool strings = std::any_of(b->items.begin(), b->items.end(),
                                   [](const BatchItem& it) { return it.kind == BK_FSST; });
  uint32_t grid = b->grid;
  if (grid == 0) {
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    grid = static_cast<uint32_t>(cus * batch_occupancy_blocks_per_cu(strings));
  }
  grid = std::max<uint32_t>(1, std::min<uint32_t>(grid, std::max<uint32_t>(1, static_cast<uint32_t>(b->items.size()))));
  if (b->use_graph) {
    // the graph's arguments are the arena addresses; entries change through
    // the upload above, so one instantiation replays every batch
    if (!b->gexec || b->graph_arena != b->d_arena || b->graph_grid != grid || b->graph_strings != strings) {
      if (b->gexec) HIP_TRY(hipGraphExecDestroy(b->gexec));
      if (b->graph) HIP_TRY(hipGraphDestroy(b->graph));
      b->gexec = nullptr;
      b->graph = nullptr;
      hipStream_t cs;
      HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) e = launch_batch_reset(q, cs);
      if (e == hipSuccess && strings) e = launch_batch_fsst_prep(q, cs);
      if (e == hipSuccess) e = launch_batch(items, q, grid, strings, cs);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(cs, &g);
      (void)hipStreamDestroy(cs);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(SDR_E_HIP, "batch graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
      }
      b->graph = g;
      HIP_TRY(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
      b->graph_arena = b->d_arena;
      b->graph_grid    = grid;
      b->graph_strings = strings;
    }
    HIP_TRY(hipGraphLaunch(b->gexec, s));
  } else {
    HIP_TRY(launch_batch_reset(q, s));
    if (strings) HIP_TRY(launch_batch_fsst_prep(q, s));
    HIP_TRY(launch_batch(items, q, grid, strings, s));
  }
  // a column of no rows has no chunk to write its closing offset
  for (const auto& c : b->cols)
    if (c.total_rows == 0 && c.offsets) HIP_TRY(launch_fill_u8(reinterpret_cast<uint8_t*>(c.offsets), 0, sizeof(int32_t), s));
  // validity masks (UNCOMPRESSED overlays) after the data decode
  auto* d_nulls = reinterpret_cast<unsigned long long*>(b->d_arena + b->off_nulls);
  if (!b->masks.empty()) {
    HIP_TRY(launch_fill_u8(reinterpret_cast<uint8_t*>(d_nulls), 0, b->masks.size() * sizeof(unsigned long long), s));
    for (size_t i = 0; i < b->masks.size(); ++i) {
      const auto& m = b->masks[i];
      if (m.fixed)
        rc = build_null_mask(m.fixed->validity, m.fixed->n_validity, m.fixed->total_rows, m.fixed->d_null_mask, d_nulls + i, s);
      else
        rc = build_null_mask(m.str->validity, m.str->n_validity, m.str->total_rows, m.str->d_null_mask, d_nulls + i, s);
      if (rc) return rc;
    }
  }
  // the host reads string totals and null counts from a pinned copy once
  // the launch has finished (batch_collect)
  HIP_TRY(hipMemcpyAsync(b->h_readback, b->d_arena + b->off_state, b->readback_bytes, hipMemcpyDeviceToHost, s));
  if (!b->launch_done) HIP_TRY(hipEventCreateWithFlags(&b->launch_done, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(b->launch_done, s));
  b->last_stream = s;
  b->launched    = true;
  return SDR_OK;
}

// 1 when the last launch has finished, 0 while it runs.
int batch_ready(sdr_batch* b)
{
  if (!b->launched || !b->launch_done) return 1;
  const hipError_t e = hipEventQuery(b->launch_done);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(SDR_E_HIP, "batch launch: %s", hipGetErrorString(e));
}

// After the last launch finished: write string totals and null counts back
// into the added columns and report look-back failures and overflow.
int batch_collect(sdr_batch* b)
{
  if (!b->launched) return SDR_OK;
  b->launched = false;
  const auto* st = reinterpret_cast<const uint64_t*>(b->h_readback);
  const auto* nulls =
    reinterpret_cast<const unsigned long long*>(b->h_readback + (b->off_nulls - b->off_state));
  int rc = SDR_OK;
  size_t w = 0;
  for (auto& c : b->cols) {
    const uint32_t err   = static_cast<uint32_t>(st[w] >> 32);
    const uint64_t total = st[w + 1];
    w += 2 + c.n_chunks;
    c.user->chars_used = total;
    c.status           = err ? SDR_E_HIP : (total > 0x7FFFFFFFull || total > c.cap) ? SDR_E_OVERFLOW : SDR_OK;
    if (c.status && !rc)
      rc = c.status == SDR_E_HIP
             ? fail(SDR_E_HIP, "string decode: offset look-back did not complete")
             : fail(SDR_E_OVERFLOW, "total_chars (%llu) exceeds capacity (%llu) or int32 max",
                    static_cast<unsigned long long>(total), static_cast<unsigned long long>(c.cap));
  }
  for (size_t i = 0; i < b->masks.size(); ++i) {
    if (b->masks[i].fixed) b->masks[i].fixed->null_count = static_cast<int64_t>(nulls[i]);
    else b->masks[i].str->null_count = static_cast<int64_t>(nulls[i]);
  }
  return rc;
}

// Outcome of one string column of the last collected launch (SDR_OK for a
// column the batch does not hold).
int batch_column_status(const sdr_batch* b, const sdr_string_column* col)
{
  for (const auto& c : b->cols)
    if (c.user == col) return c.status;
  return SDR_OK;
}

// Wait for the last launch and write back string totals and null counts.
int batch_finish(sdr_batch* b, hipStream_t s)
{
  clear_error();
  (void)s;
  if (!b->launched) return SDR_OK;
  HIP_TRY(hipEventSynchronize(b->launch_done));
  return batch_collect(b);
}

namespace {
sdr_batch* thread_batch()
{
  thread_local std::unique_ptr<sdr_batch> b(new sdr_batch());
  return b.get();
}
}  // namespace

int decode_table_persistent(sdr_column* cols, uint32_t n_cols, hipStream_t s)
{
  sdr_batch* b = thread_batch();
  int rc = wait_upload(b);
  if (rc) return rc;
  b->items.clear();
  b->cols.clear();
  b->masks.clear();
  b->launched = false;  // results of a launch nobody collected are dropped
  for (uint32_t c = 0; c < n_cols; ++c) cols[c].null_count = 0;
  rc = batch_add_fixed(b, cols, n_cols);
  if (!rc) rc = batch_launch(b, s);
  if (!rc && !b->masks.empty()) rc = batch_finish(b, s);  // null counts need the one sync
  return rc;
}

int decode_strings_persistent(sdr_string_column* col, hipStream_t s)
{
  sdr_batch* b = thread_batch();
  int rc = wait_upload(b);
  if (rc) return rc;
  b->items.clear();
  b->cols.clear();
  b->masks.clear();
  b->launched = false;  // results of a launch nobody collected are dropped
  col->chars_used = 0;
  col->null_count = 0;
  rc = batch_add_strings(b, col);
  if (!rc) rc = batch_launch(b, s);
  if (!rc) rc = batch_finish(b, s);
  return rc;
}

}  // namespace sdr

using namespace sdr;

extern "C" {

int sdr_batch_create(sdr_batch** out)
{
  clear_error();
  if (!out) return fail(SDR_E_INVALID_ARG, "out is NULL");
  *out = new sdr_batch();
  return SDR_OK;
}

int sdr_batch_set_option(sdr_batch* b, uint32_t key, int64_t value)
{
  clear_error();
  if (!b) return fail(SDR_E_INVALID_ARG, "batch is NULL");
  if (key >= SDR_BATCH_UNITS_BITPACKING && key <= SDR_BATCH_UNITS_CONSTANT) {
    if (value < 1 || value > 1 << 20) return fail(SDR_E_INVALID_ARG, "units per entry out of range");
    b->units[key - SDR_BATCH_UNITS_BITPACKING + 1] = static_cast<uint32_t>(value);
  } else if (key == SDR_BATCH_GRID) {
    if (value < 0 || value > 1 << 24) return fail(SDR_E_INVALID_ARG, "grid out of range");
    b->grid = static_cast<uint32_t>(value);
  } else if (key == SDR_BATCH_USE_GRAPH) {
    b->use_graph = value != 0;
  } else {
    return fail(SDR_E_INVALID_ARG, "unknown batch option %u", key);
  }
  return SDR_OK;
}

int sdr_batch_add_table(sdr_batch* b, sdr_column* cols, uint32_t n_cols)
{
  clear_error();
  if (!b) return fail(SDR_E_INVALID_ARG, "batch is NULL");
  int rc = validate_fixed_columns(cols, n_cols);
  if (rc) return rc;
  rc = wait_upload(b);
  if (rc) return rc;
  for (uint32_t c = 0; c < n_cols; ++c) cols[c].null_count = 0;
  return batch_add_fixed(b, cols, n_cols);
}

int sdr_batch_add_strings(sdr_batch* b, sdr_string_column* col)
{
  clear_error();
  if (!b) return fail(SDR_E_INVALID_ARG, "batch is NULL");
  int rc = validate_string_column(col);
  if (rc) return rc;
  rc = wait_upload(b);
  if (rc) return rc;
  col->chars_used = 0;
  col->null_count = 0;
  return batch_add_strings(b, col);
}

int sdr_batch_num_entries(const sdr_batch* b) { return b ? static_cast<int>(b->items.size()) : SDR_E_INVALID_ARG; }

int sdr_batch_launch(sdr_batch* b, sdr_stream_t stream)
{
  if (!b) return SDR_E_INVALID_ARG;
  return batch_launch(b, static_cast<hipStream_t>(stream));
}

int sdr_batch_finish(sdr_batch* b, sdr_stream_t stream)
{
  if (!b) return SDR_E_INVALID_ARG;
  return batch_finish(b, static_cast<hipStream_t>(stream));
}

int sdr_batch_clear(sdr_batch* b)
{
  clear_error();
  if (!b) return fail(SDR_E_INVALID_ARG, "batch is NULL");
  int rc = wait_upload(b);
  if (rc) return rc;
  b->items.clear();
  b->cols.clear();
  b->masks.clear();
  b->launched = false;  // results of a launch nobody collected are dropped
  return SDR_OK;
}

int sdr_batch_destroy(sdr_batch* b)
{
  delete b;
  return SDR_OK;
}

}  // extern "C"
