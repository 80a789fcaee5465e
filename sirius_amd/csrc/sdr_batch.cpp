// sdr_batch.cpp — host side of the persistent mixed-codec decode: builds the
// device work queue (one arena: header, entries grouped by decoder and
// width, FSST columns, chunks, segments, look-back state), uploads it from
// pinned memory with one copy, and launches the groups' persistent kernels,
// optionally as a replayed hipGraph whose group kernels are parallel
// branches.
#include "sdr_batch.hpp"
#include "sdr_internal.hpp"
#include "sdr_kernels.hpp"
#include "sdr_tune.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

struct sdr_batch {
  // host-side description
  std::vector<sdr::BatchItem> items;
  std::vector<uint8_t> item_group;  // BatchGroup of each entry
  struct Col {
    uint32_t n_chunks = 0;
    std::vector<sdr::FsstSeg> segs;  // row ranges: segments and the gaps between them
    sdr_string_column* user;  // chars_used / null_count written back at finish
    int32_t* offsets;
    uint8_t* chars;
    uint64_t cap;
    uint32_t total_rows;
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
  hipEvent_t upload_done = nullptr;
  bool upload_pending  = false;
  uint8_t* h_readback  = nullptr;  // pinned copy of the look-back state and null counters
  size_t readback_cap  = 0;
  size_t readback_bytes = 0;
  size_t used_bytes    = 0;
  size_t off_items = 0, off_cols = 0, off_chunks = 0, off_segs = 0, off_state = 0, off_nulls = 0;
  uint32_t n_status    = 0;
  hipEvent_t launch_done = nullptr;
  size_t n_entries     = 0;                         // queue entries of the last upload
  uint32_t entry_count[sdr::BG_COUNT] = {};         // ... per group
  // launch configuration
  // work units per queue entry, per BatchKind, when the batch sets them
  // (SDR_BATCH_UNITS_*); otherwise the SDR_TUNE_BATCH_UNITS_* knob of the
  // kind and value width
  uint32_t units[8]    = {0, 16, 2, 16, 16, 4, 4, 1};
  bool units_set[8]    = {};
  uint32_t grid        = 0;                         // 0 = resident capacity
  bool use_graph       = false;
  // fixed-width segments: per-codec grids over descriptors in the arena
  // (default), or persistent group kernels draining queue entries
  bool fixed_grids     = true;
  size_t off_fsegs     = 0;
  std::vector<sdr::FixedLaunch> fixed;              // the per-codec grids of the last upload
  std::vector<sdr::FixedLaunch> graph_fixed;        // ... that the recorded graph launches
  size_t graph_fsegs   = 0;                         // ... and where their descriptors were
  uint32_t graph_fsst  = 0;                         // ... and its FSST grid (one block per chunk)
  // where the per-codec grids of the last upload read their descriptors: the
  // arena, or (a batch of fixed-width segments alone, no graph) the kernel
  // arguments or a descriptor slot, as the batched table call passes them
  enum FixedSrc { FS_ARENA, FS_INLINE, FS_SLOT } fixed_src = FS_ARENA;
  std::vector<sdr::FixedSeg> h_fsegs;               // ... the descriptors of the last upload, by grid
  hipStream_t capture  = nullptr;                   // records the graph of a launch with grids
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph     = nullptr;
  uint8_t* graph_arena = nullptr;
  uint32_t graph_mask  = 0;  // groups the recorded graph launches
  uint32_t graph_gridopt = 0;
  uint32_t group_cap[sdr::BG_COUNT] = {};  // resident blocks of each group's kernel (0 = not queried)
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
  if (capture) (void)hipStreamDestroy(capture);
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

// Queue group of an entry of kind k over values of ts bytes.
BatchGroup group_of(BatchKind k, uint32_t ts)
{
  const uint32_t w = ts == 1 ? 0u : ts == 2 ? 1u : ts == 4 ? 2u : 3u;
  switch (k) {
    case BK_BITPACKING: return BatchGroup(BG_BP1 + w);
    case BK_RLE: return BatchGroup(BG_RLE1 + w);
    case BK_ALP: return ts == 4 ? BG_ALP4 : BG_ALP8;
    case BK_ALPRD: return ts == 4 ? BG_ALPRD4 : BG_ALPRD8;
    case BK_FSST: return BG_FSST;
    default: return BG_COPY;
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

// Wait until the staging copy is no longer read by a queued upload.
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
      const uint32_t rle_slice = tune(SDR_TUNE_RLE_SLICE_ROWS, col.type_size);
      for (uint32_t s = 0; s < run.n_segments; ++s) {
        const sdr_segment& g = run.segments[s];
        if (!g.row_count) continue;
        // one work item per segment; upload cuts it into queue entries
        BatchItem it{};
        it.seg        = g.d_bytes;
        it.out        = static_cast<uint8_t*>(col.d_out) + uint64_t{g.row_offset} * col.type_size;
        it.bytes_size = g.bytes_size;
        it.row_count  = g.row_count;
        it.unit0      = 0;
        it.n_units    = units_in_segment(k, g.row_count, col.type_size, rle_slice);
        it.kind       = k;
        it.ts         = static_cast<uint8_t>(col.type_size);
        if (k == BK_RLE) it.reserved = static_cast<uint16_t>(rle_slice / 64u);
        b->items.push_back(it);
        b->item_group.push_back(static_cast<uint8_t>(group_of(k, col.type_size)));
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
  // the column's chunks become FSST queue entries at upload
  if (col->has_nulls && col->d_null_mask) b->masks.push_back({nullptr, col});
  b->cols.push_back(std::move(bc));
  return SDR_OK;
}

// Resident blocks of the kernels of the groups in `mask` (queried once).
int batch_group_caps(sdr_batch* b, uint32_t mask)
{
  int dev = 0, cus = 0;
  for (uint32_t g = 0; g < BG_COUNT; ++g) {
    if (!(mask >> g & 1u) || b->group_cap[g]) continue;
    if (!cus) {
      HIP_TRY(hipGetDevice(&dev));
      HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    const void* f = nullptr;
    uint32_t block = 0, per = 0;
    HIP_TRY(batch_group_launch_params(g, &f, &block, &per));
    b->group_cap[g] = static_cast<uint32_t>(cus) * per;
  }
  return SDR_OK;
}

// Units per queue entry of kind k over values of ts bytes.
uint32_t batch_units(const sdr_batch* b, uint32_t k, uint32_t ts)
{
  if (b->units_set[k]) return b->units[k];
  switch (k) {
    case BK_BITPACKING: return tune(SDR_TUNE_BATCH_UNITS_BITPACKING, ts);
    case BK_RLE: return tune(SDR_TUNE_BATCH_UNITS_RLE, ts);
    case BK_ALP: return tune(SDR_TUNE_BATCH_UNITS_ALP, ts);
    case BK_ALPRD: return tune(SDR_TUNE_BATCH_UNITS_ALPRD, ts);
    case BK_COPY:
    case BK_CONSTANT: return tune(SDR_TUNE_BATCH_UNITS_COPY, ts);
    default: return b->units[k];
  }
}

// How the work items become queue entries.  A fixed-width group's items
// are cut into entries of per[g] units: enough entries to give every
// resident block of the group's kernel work (a multiple of the four
// wavefronts for the wave-parallel codecs), at most the batch's units per
// entry for the codec, whose claims then amortise.  An FSST item is one
// entry.  Returns the entries per group and the groups present.
int batch_plan(sdr_batch* b, uint32_t per[BG_COUNT], uint32_t entries[BG_COUNT], uint32_t* mask)
{
  uint64_t tot[BG_COUNT]  = {};
  uint32_t maxu[BG_COUNT] = {};
  *mask = 0;
  for (size_t i = 0; i < b->items.size(); ++i) {
    const uint32_t g = b->item_group[i];
    *mask |= 1u << g;
    tot[g] += b->items[i].n_units;
    maxu[g] = std::max<uint32_t>(1, batch_units(b, b->items[i].kind, b->items[i].ts));
  }
  uint64_t fsst_chunks = 0;  // one FSST entry per chunk of every string column
  for (const auto& c : b->cols) fsst_chunks += c.n_chunks;
  if (fsst_chunks) *mask |= 1u << BG_FSST;
  if (b->fixed_grids) {  // fixed-width groups become per-codec grids: one work entry per segment
    for (uint32_t g = 0; g < BG_COUNT; ++g) {
      entries[g] = 0;
      per[g]     = 1;
    }
    for (size_t i = 0; i < b->items.size(); ++i) ++entries[b->item_group[i]];
    entries[BG_FSST] = static_cast<uint32_t>(fsst_chunks);
    *mask &= 1u << BG_FSST;
    return batch_group_caps(b, *mask);
  }
  int rc = batch_group_caps(b, *mask);
  if (rc) return rc;
  for (uint32_t g = 0; g < BG_COUNT; ++g) {
    entries[g] = 0;
    per[g]     = 1;
    if (!tot[g] || g == BG_FSST) continue;
    const uint64_t cap  = std::max<uint32_t>(1, b->group_cap[g]);
    uint64_t want       = (tot[g] + cap - 1) / cap;
    const bool wave_par = !(g >= BG_RLE1 && g <= BG_RLE8);
    if (wave_par) want = (want + BATCH_WAVES - 1) / BATCH_WAVES * BATCH_WAVES;
    per[g] = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(want, 1), maxu[g]));
  }
  for (size_t i = 0; i < b->items.size(); ++i) {
    const uint32_t g = b->item_group[i];
    entries[g] += (b->items[i].n_units + per[g] - 1) / per[g];
  }
  entries[BG_FSST] = static_cast<uint32_t>(fsst_chunks);
  return SDR_OK;
}

// Lay out the arena, fill the pinned staging copy, upload it.
int batch_upload(sdr_batch* b, hipStream_t s)
{
  uint32_t per[BG_COUNT], entries[BG_COUNT], mask = 0;
  int prc = batch_plan(b, per, entries, &mask);
  if (prc) return prc;
  size_t n_entries = 0;  // fixed-width queue entries (BatchItem); the FSST ones are FsstEntry
  for (uint32_t g = 0; g < BG_COUNT; ++g) {
    if (!(mask >> g & 1u)) entries[g] = 0;  // no queue entries: per-codec grids (or nothing)
    if (g != BG_FSST) n_entries += entries[g];
    b->entry_count[g] = entries[g];
  }
  b->n_entries = n_entries;
  // the per-codec grids' descriptors, by codec kind and value width, in the
  // order the items were added
  std::vector<FixedSeg>& fsegs = b->h_fsegs;  // kept between launches: no allocation once warm
  fsegs.clear();
  b->fixed.clear();
  if (b->fixed_grids) {
    // value widths 1, 2, 4, 8 and 16 (UNCOMPRESSED and CONSTANT)
    uint32_t count[8][5] = {};
    auto wi = [](uint32_t ts) { return ts == 1 ? 0u : ts == 2 ? 1u : ts == 4 ? 2u : ts == 8 ? 3u : 4u; };
    for (size_t i = 0; i < b->items.size(); ++i)
      if (b->item_group[i] != BG_FSST) ++count[b->items[i].kind][wi(b->items[i].ts)];
    uint32_t at_k[8][5] = {}, base = 0;
    for (uint32_t k = 0; k < 8; ++k)
      for (uint32_t w = 0; w < 5; ++w) {
        at_k[k][w] = base;
        if (count[k][w]) b->fixed.push_back({k, 1u << w, base, count[k][w], 0u});
        base += count[k][w];
      }
    fsegs.resize(base);
    for (size_t i = 0; i < b->items.size(); ++i) {
      if (b->item_group[i] == BG_FSST) continue;
      const BatchItem& it = b->items[i];
      fsegs[at_k[it.kind][wi(it.ts)]++] = FixedSeg{it.seg, it.out, it.bytes_size, it.row_count};
    }
    for (auto& fl : b->fixed)
      for (uint32_t j = 0; j < fl.n; ++j) fl.max_rows = std::max(fl.max_rows, fsegs[fl.base + j].row_count);
  }
  b->fixed_src = sdr_batch::FS_ARENA;
  if (b->fixed_grids && mask == 0 && b->cols.empty() && b->masks.empty() && !b->fixed.empty()) {
    // nothing on the device reads the arena: up to two launches of
    // SEGS_INLINE descriptors per grid travel in the kernel arguments,
    // larger grids read a descriptor slot uploaded at launch
    const bool fits = std::all_of(b->fixed.begin(), b->fixed.end(),
                                  [](const FixedLaunch& fl) { return fl.n <= 2u * SEGS_INLINE; });
    b->fixed_src = fits ? sdr_batch::FS_INLINE : sdr_batch::FS_SLOT;
    return SDR_OK;
  }
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
  off                         = al(off + n_entries * sizeof(BatchItem));
  const size_t off_fentries   = off;
  off                         = al(off + size_t{entries[BG_FSST]} * sizeof(FsstEntry));
  b->off_fsegs                = off;
  off                         = al(off + fsegs.size() * sizeof(FixedSeg));
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
  int wrc = wait_upload(b);
  if (wrc) return wrc;
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
  // fixed-width entries laid out group by group, in the order added within a
  // group; the FSST group's entries are an array of their own
  uint32_t at[BG_COUNT] = {};
  for (uint32_t g = 0, o = 0; g < BG_COUNT; ++g) {
    q.count[g] = entries[g];
    q.first[g] = g == BG_FSST ? 0u : o;
    at[g]      = q.first[g];
    if (g != BG_FSST) o += entries[g];
  }
  q.fsst_entries = reinterpret_cast<const FsstEntry*>(d + off_fentries);
  q.fsst_cols    = reinterpret_cast<const FsstColumn*>(d + b->off_cols);
  q.n_items       = static_cast<uint32_t>(n_entries);
  for (uint32_t w = 0; w < 4; ++w) q.bp_chunk_rows[w] = tune(SDR_TUNE_BP_CHUNK_ROWS, 1u << w);
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
  if (!fsegs.empty()) std::memcpy(h + b->off_fsegs, fsegs.data(), fsegs.size() * sizeof(FixedSeg));
  if (mask & (1u << BG_FSST)) {  // every chunk of every string column, column by column
    auto* fe = reinterpret_cast<FsstEntry*>(h + off_fentries);
    size_t j = 0;
    for (uint32_t ci = 0; ci < b->cols.size(); ++ci)
      for (uint32_t k = 0; k < b->cols[ci].n_chunks; ++k) fe[j++] = FsstEntry{ci, k};
  }
  auto* hi = reinterpret_cast<BatchItem*>(h + b->off_items);
  for (size_t i = 0; i < b->items.size(); ++i) {
    const uint32_t g  = b->item_group[i];
    if (!(mask >> g & 1u)) continue;  // a per-codec grid's segment
    const BatchItem& it = b->items[i];
    for (uint32_t u = 0; u < it.n_units; u += per[g]) {
      BatchItem& dst = hi[at[g]++];
      dst            = it;
      dst.unit0      = it.unit0 + u;
      dst.n_units    = std::min(per[g], it.n_units - u);
    }
  }
  HIP_TRY(hipMemcpyAsync(d, h, upload_bytes, hipMemcpyHostToDevice, s));
  if (!b->upload_done) HIP_TRY(hipEventCreateWithFlags(&b->upload_done, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(b->upload_done, s));
  b->upload_pending = true;
  return SDR_OK;
}

// The per-codec grids of the batch's fixed-width segments, one launch per
// codec kind and value width: descriptors in the arena, in the kernel
// arguments (FS_INLINE) or at `slot` (FS_SLOT).
hipError_t launch_fixed_grids(const sdr_batch* b, hipStream_t s, const FixedSeg* slot = nullptr)
{
  const bool host  = b->fixed_src == sdr_batch::FS_INLINE;
  const auto* segs = host                                 ? b->h_fsegs.data()
                     : b->fixed_src == sdr_batch::FS_SLOT ? slot
                                                          : reinterpret_cast<const FixedSeg*>(b->d_arena + b->off_fsegs);
  for (const auto& fl : b->fixed) {
    hipError_t e = hipSuccess;
    switch (fl.kind) {
      case BK_BITPACKING: e = launch_bitpacking(segs, fl.base, fl.n, fl.max_rows, fl.ts, s, host); break;
      case BK_RLE: e = launch_rle(segs, fl.base, fl.n, fl.max_rows, fl.ts, s, host); break;
      case BK_ALP: e = launch_alp(segs, fl.base, fl.n, fl.max_rows, fl.ts, false, s, host); break;
      case BK_ALPRD: e = launch_alp(segs, fl.base, fl.n, fl.max_rows, fl.ts, true, s, host); break;
      case BK_COPY: e = launch_uncompressed(segs, fl.base, fl.n, fl.max_rows, fl.ts, s, host); break;
      default: e = launch_constant(segs, fl.base, fl.n, fl.max_rows, fl.ts, s, host); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Every kernel of a launch in stream order: the reset (when the batch holds
// string columns or queue entries), the FSST pre-pass, the group kernels,
// then the per-codec grids.
hipError_t enqueue_launch(const sdr_batch* b, BatchQueue* q, const BatchItem* items, uint32_t mask,
                          const uint32_t* grid, const uint32_t* cnt, hipStream_t s)
{
  hipError_t e = hipSuccess;
  if (mask || !b->cols.empty()) e = launch_batch_reset(q, s);
  if (e == hipSuccess && (mask & (1u << BG_FSST))) e = launch_batch_fsst_prep(q, s);
  for (uint32_t g = 0; e == hipSuccess && g < BG_COUNT; ++g)
    if (mask >> g & 1u) e = launch_batch_group(items, q, g, std::min(grid[g], cnt[g]), s);
  if (e == hipSuccess) e = launch_fixed_grids(b, s);
  return e;
}

int batch_launch(sdr_batch* b, hipStream_t s)
{
  clear_error();
  int rc = SDR_OK;
  if (b->launch_done && b->last_stream && b->last_stream != s) HIP_TRY(hipStreamWaitEvent(s, b->launch_done, 0));
  rc = batch_upload(b, s);
  if (rc) return rc;
  if (b->fixed_src != sdr_batch::FS_ARENA) {  // the launches the batched table call makes
    void* d     = nullptr;
    void* token = nullptr;
    if (b->fixed_src == sdr_batch::FS_SLOT) {
      rc = desc_acquire(b->h_fsegs.data(), b->h_fsegs.size() * sizeof(FixedSeg), s, &d, &token);
      if (rc) return rc;
    }
    const hipError_t e = launch_fixed_grids(b, s, static_cast<const FixedSeg*>(d));
    rc                 = desc_release(token, s);
    HIP_TRY(e);
    if (rc) return rc;
    if (!b->launch_done) HIP_TRY(hipEventCreateWithFlags(&b->launch_done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(b->launch_done, s));
    b->last_stream = s;
    b->launched    = true;
    return SDR_OK;
  }
  auto* q     = reinterpret_cast<BatchQueue*>(b->d_arena);
  auto* items = reinterpret_cast<const BatchItem*>(b->d_arena + b->off_items);
  // each group kernel's grid: its resident capacity (or the SDR_BATCH_GRID
  // cap), so a block loops over many entries
  uint32_t cnt[BG_COUNT] = {}, mask = 0;
  for (uint32_t g = 0; g < BG_COUNT; ++g) {
    cnt[g] = b->entry_count[g];
    if (cnt[g]) mask |= 1u << g;
  }
  uint32_t grid[BG_COUNT] = {};
  for (uint32_t g = 0; g < BG_COUNT; ++g) {
    if (!(mask >> g & 1u)) continue;
    grid[g] = std::max<uint32_t>(1, b->grid ? std::min(b->grid, b->group_cap[g]) : b->group_cap[g]);
  }
  grid[BG_FSST] = cnt[BG_FSST];  // the FSST kernel decodes one chunk per block
  const bool strings = mask & (1u << BG_FSST);
  if (b->use_graph && !b->fixed.empty()) {
    // with per-codec grids the graph is recorded from the stream order (on
    // the batch's own stream), and recorded again when the grids change
    if (!b->gexec || b->graph_arena != b->d_arena || b->graph_mask != mask || b->graph_gridopt != b->grid ||
        !(b->graph_fixed == b->fixed) || b->graph_fsegs != b->off_fsegs || b->graph_fsst != cnt[BG_FSST]) {
      if (b->gexec) HIP_TRY(hipGraphExecDestroy(b->gexec));
      if (b->graph) HIP_TRY(hipGraphDestroy(b->graph));
      b->gexec = nullptr;
      b->graph = nullptr;
      if (!b->capture) HIP_TRY(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
      HIP_TRY(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
      const hipError_t e = enqueue_launch(b, q, items, mask, grid, cnt, b->capture);
      hipGraph_t gr      = nullptr;
      const hipError_t ee = hipStreamEndCapture(b->capture, &gr);
      if ((e != hipSuccess || ee != hipSuccess) && gr) (void)hipGraphDestroy(gr);
      HIP_TRY(e);
      HIP_TRY(ee);
      b->graph = gr;
      HIP_TRY(hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0));
      b->graph_arena   = b->d_arena;
      b->graph_mask    = mask;
      b->graph_gridopt = b->grid;
      b->graph_fixed   = b->fixed;
      b->graph_fsegs   = b->off_fsegs;
      b->graph_fsst    = cnt[BG_FSST];
    }
    HIP_TRY(hipGraphLaunch(b->gexec, s));
  } else if (b->use_graph && mask) {
    // nodes: reset -> [FSST pre-pass] -> one persistent kernel per group, the
    // group kernels parallel branches; node arguments are the arena
    // addresses, entries change through the upload above, so one
    // instantiation replays every batch with the same groups
    if (!b->gexec || b->graph_arena != b->d_arena || b->graph_mask != mask || b->graph_gridopt != b->grid ||
        !b->graph_fixed.empty() || b->graph_fsst != cnt[BG_FSST]) {
      if (b->gexec) HIP_TRY(hipGraphExecDestroy(b->gexec));
      if (b->graph) HIP_TRY(hipGraphDestroy(b->graph));
      b->gexec = nullptr;
      b->graph = nullptr;
      hipGraph_t gr = nullptr;
      HIP_TRY(hipGraphCreate(&gr, 0));
      b->graph = gr;
      BatchQueue* qa = q;
      const BatchItem* ia = items;
      void* reset_args[] = {&qa};
      hipKernelNodeParams kp{};
      kp.func           = const_cast<void*>(batch_reset_kernel());
      kp.gridDim        = dim3(BATCH_RESET_GRID);
      kp.blockDim       = dim3(256);
      kp.kernelParams   = reset_args;
      hipGraphNode_t n_reset = nullptr, n_prep = nullptr;
      HIP_TRY(hipGraphAddKernelNode(&n_reset, gr, nullptr, 0, &kp));
      if (strings) {
        kp.func     = const_cast<void*>(batch_fsst_prep_kernel());
        kp.gridDim  = dim3(BATCH_PREP_GRID);
        kp.blockDim = dim3(BATCH_PREP_THREADS);
        HIP_TRY(hipGraphAddKernelNode(&n_prep, gr, &n_reset, 1, &kp));
      }
      uint32_t gids[BG_COUNT];
      for (uint32_t g = 0; g < BG_COUNT; ++g) {
        if (!(mask >> g & 1u)) continue;
        gids[g] = g;
        const void* f = nullptr;
        uint32_t block = 0, per = 0;
        HIP_TRY(batch_group_launch_params(g, &f, &block, &per));
        void* args[] = {&ia, &qa, &gids[g]};
        kp.func         = const_cast<void*>(f);
        kp.gridDim      = dim3(grid[g]);
        kp.blockDim     = dim3(block);
        kp.kernelParams = args;
        hipGraphNode_t n = nullptr;
        hipGraphNode_t dep = g == BG_FSST ? n_prep : n_reset;
        HIP_TRY(hipGraphAddKernelNode(&n, gr, &dep, 1, &kp));
      }
      HIP_TRY(hipGraphInstantiate(&b->gexec, gr, nullptr, nullptr, 0));
      b->graph_arena   = b->d_arena;
      b->graph_mask    = mask;
      b->graph_gridopt = b->grid;
      b->graph_fixed.clear();
      b->graph_fsst    = cnt[BG_FSST];
    }
    HIP_TRY(hipGraphLaunch(b->gexec, s));
  } else {  // also a batch with no entries: its columns' look-back state is cleared
    HIP_TRY(enqueue_launch(b, q, items, mask, grid, cnt, s));
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
  if (!b->cols.empty() || !b->masks.empty())  // otherwise there is nothing to write back
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
  int rc = SDR_OK;
  b->items.clear();
  b->item_group.clear();
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
  int rc = SDR_OK;
  b->items.clear();
  b->item_group.clear();
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
    b->units[key - SDR_BATCH_UNITS_BITPACKING + 1]     = static_cast<uint32_t>(value);
    b->units_set[key - SDR_BATCH_UNITS_BITPACKING + 1] = true;
  } else if (key == SDR_BATCH_GRID) {
    if (value < 0 || value > 1 << 24) return fail(SDR_E_INVALID_ARG, "grid out of range");
    b->grid = static_cast<uint32_t>(value);
  } else if (key == SDR_BATCH_USE_GRAPH) {
    b->use_graph = value != 0;
  } else if (key == SDR_BATCH_FIXED_KERNELS) {
    if (value < 0 || value > 1) return fail(SDR_E_INVALID_ARG, "fixed-width kernel choice out of range");
    b->fixed_grids = value == 0;
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
  for (uint32_t c = 0; c < n_cols; ++c) cols[c].null_count = 0;
  return batch_add_fixed(b, cols, n_cols);
}

int sdr_batch_add_strings(sdr_batch* b, sdr_string_column* col)
{
  clear_error();
  if (!b) return fail(SDR_E_INVALID_ARG, "batch is NULL");
  int rc = validate_string_column(col);
  if (rc) return rc;
  if (has_dict_fsst(col))
    return fail(SDR_E_UNSUPPORTED_CODEC,
                "DICT_FSST strings need host round trips for their dictionaries; decode them with "
                "sdr_decode_strings_column");
  col->chars_used = 0;
  col->null_count = 0;
  return batch_add_strings(b, col);
}

int sdr_batch_num_entries(const sdr_batch* b)
{
  if (!b) return SDR_E_INVALID_ARG;
  // entries the added work makes with the current options
  uint32_t per[BG_COUNT], entries[BG_COUNT], mask = 0;
  if (batch_plan(const_cast<sdr_batch*>(b), per, entries, &mask)) return SDR_E_HIP;
  size_t n = 0;
  for (uint32_t g = 0; g < BG_COUNT; ++g) n += entries[g];
  return static_cast<int>(n);
}

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
  b->items.clear();
  b->item_group.clear();
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
