// sdr_serve.cpp — the decode server (sdr_server_*): resident segments,
// HBM-budget admission, pool-allocated outputs and continuous batching over
// persistent launches (or one drop-in launch per request, the baselines).  Runs on the calling thread; progress() does all device work.
//
// Device memory (resident segments and request outputs) comes from one
// server-owned pool (DeviceArena): slabs reserved with hipMalloc when it
// grows, carved on the host.  A range that device work may still use is
// freed behind an event and rejoins the pool once the event has passed, so
// steady-state serving makes no allocation call.  New segments are staged
// through a ring of pinned buffers with hipMemcpyAsync on a staging stream;
// launches wait for the staging copies on the device, never on the host.
#include "sdr_arena.hpp"
#include "sdr_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <unordered_map>
#include <vector>

namespace {

constexpr uint64_t DEFAULT_BUDGET = 200ull * 1000 * 1000 * 1000;
constexpr size_t STAGE_CHUNK      = 16u << 20;  // one pinned staging buffer
constexpr uint32_t STAGE_RING     = 4;          // staging buffers in flight
constexpr uint64_t SLAB_MIN       = 64ull << 20;
constexpr uint64_t SLAB_MAX       = 4ull << 30;

struct Req {
  sdr_request r{};
  std::vector<sdr_segment> segs;
  std::vector<sdr_run> runs;
  sdr_column col{};
  sdr_string_column scol{};
  uint64_t out_bytes  = 0;
  uint64_t chars_cap  = 0;
  uint8_t* d_alloc    = nullptr;  // outputs, from the server's pool
  double t_submit     = 0;
  int status          = SDR_OK;
  bool done           = false;
};

struct Slot {
  hipStream_t s        = nullptr;
  hipEvent_t ev_start  = nullptr;
  hipEvent_t ev_end    = nullptr;
  sdr_batch* batch     = nullptr;
  bool busy            = false;
  int launch_status    = SDR_OK;  // synchronous status of a drop-in launch
  uint64_t stage_seen  = 0;       // staging copies this slot's stream already waits for
  std::vector<Req*> reqs;
};

// Pool ranges waiting for an event before they can be reused.
struct Fenced {
  std::vector<uint8_t*> ptrs;
  hipEvent_t ev;
};

using Clock = std::chrono::steady_clock;

// Adds the host time of a scope to one of the server's phase counters.
struct PhaseTimer {
  double& acc;
  Clock::time_point t = Clock::now();
  explicit PhaseTimer(double& a) : acc(a) {}
  ~PhaseTimer() { acc += std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }
};

}  // namespace

struct sdr_server {
  sdr_server_config cfg{};
  int device            = 0;
  sdr::DeviceArena arena;
  uint64_t slab_bytes   = SLAB_MIN;
  std::vector<hipEvent_t> spare_events;
  std::deque<Fenced> fenced;  // released ranges behind an event
  std::unordered_map<hipStream_t, std::vector<uint8_t*>> open_release;  // released, event not recorded yet
  hipStream_t aux       = nullptr;  // staging copies
  hipStream_t fence_s   = nullptr;  // fences that wait for in-flight launches
  uint8_t* ring[STAGE_RING] = {};
  hipEvent_t ring_ev[STAGE_RING] = {};
  bool ring_used[STAGE_RING] = {};
  uint32_t ring_next    = 0;
  hipEvent_t stage_done = nullptr;  // recorded on aux after the last staging copy
  uint64_t stage_seq    = 0;
  hipEvent_t t0_event   = nullptr;
  double t0_gpu_ms      = 0;  // server clock at t0_event
  Clock::time_point t0;
  std::vector<Slot> slots;
  std::deque<Req*> queue;     // submitted, not admitted
  std::deque<Req*> admitted;  // admitted, not launched
  std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs;
  std::vector<sdr_completion> done;
  std::unordered_map<const uint8_t*, uint64_t> staged;
  uint64_t resident     = 0;
  uint64_t outputs_held = 0;
  uint32_t running      = 0;
  sdr_server_stats stats{};
  double union_end = 0;  // end of the launch intervals counted in gpu_union_ms

  double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
  double gpu_ms(hipEvent_t ev) const
  {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t0_event, ev) != hipSuccess) return 0;
    return t0_gpu_ms + ms;
  }
  ~sdr_server();
};

sdr_server::~sdr_server()
{
  // everything that may still touch the pool's memory finishes first; the
  // arena's destructor then returns the slabs
  for (auto& sl : slots) {
    if (sl.s) (void)hipStreamSynchronize(sl.s);
    if (sl.batch) (void)sdr_batch_destroy(sl.batch);
    if (sl.ev_start) (void)hipEventDestroy(sl.ev_start);
    if (sl.ev_end) (void)hipEventDestroy(sl.ev_end);
  }
  for (auto& f : fenced) {
    (void)hipEventSynchronize(f.ev);
    (void)hipEventDestroy(f.ev);
  }
  if (aux) (void)hipStreamSynchronize(aux);
  if (fence_s) (void)hipStreamSynchronize(fence_s);
  for (hipEvent_t e : spare_events) (void)hipEventDestroy(e);
  for (uint32_t i = 0; i < STAGE_RING; ++i) {
    if (ring_ev[i]) (void)hipEventDestroy(ring_ev[i]);
    if (ring[i]) (void)hipHostFree(ring[i]);
  }
  if (stage_done) (void)hipEventDestroy(stage_done);
  for (auto& sl : slots)
    if (sl.s) (void)hipStreamDestroy(sl.s);
  if (aux) (void)hipStreamDestroy(aux);
  if (fence_s) (void)hipStreamDestroy(fence_s);
  if (t0_event) (void)hipEventDestroy(t0_event);
}

namespace sdr {
namespace {

uint64_t held(const sdr_server* sv) { return sv->resident + sv->outputs_held; }

int get_event(sdr_server* sv, hipEvent_t* ev)
{
  if (!sv->spare_events.empty()) {
    *ev = sv->spare_events.back();
    sv->spare_events.pop_back();
    return SDR_OK;
  }
  HIP_TRY(hipEventCreateWithFlags(ev, hipEventDisableTiming));
  return SDR_OK;
}

// Record one event per consumer stream behind the ranges released on it
// since the last call (releases are grouped, so a release costs no HIP call).
int close_releases(sdr_server* sv)
{
  for (auto& kv : sv->open_release) {
    if (kv.second.empty()) continue;
    hipEvent_t ev = nullptr;
    int rc        = get_event(sv, &ev);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ev, kv.first));
    sv->fenced.push_back({std::move(kv.second), ev});
    kv.second.clear();
  }
  return SDR_OK;
}

// Return every fenced range whose event has passed to the pool; with
// `wait_one`, first wait for the oldest fence.
int reclaim(sdr_server* sv, bool wait_one)
{
  if (wait_one && !sv->fenced.empty()) HIP_TRY(hipEventSynchronize(sv->fenced.front().ev));
  for (auto it = sv->fenced.begin(); it != sv->fenced.end();) {
    const hipError_t e = hipEventQuery(it->ev);
    if (e == hipErrorNotReady) {
      ++it;
      continue;
    }
    if (e != hipSuccess) return fail(SDR_E_HIP, "pool fence: %s", hipGetErrorString(e));
    for (uint8_t* p : it->ptrs) (void)sv->arena.put(p);
    sv->spare_events.push_back(it->ev);
    it = sv->fenced.erase(it);
  }
  return SDR_OK;
}

// `bytes` of device memory from the pool.  Reuses released ranges first;
// while the pool's reservation would pass the budget, waits for released
// ranges still behind a fence rather than reserve more; otherwise reserves a
// new slab.
int pool_alloc(sdr_server* sv, uint64_t bytes, uint8_t** out)
{
  *out = sv->arena.take(bytes);
  if (*out) return SDR_OK;
  int rc = close_releases(sv);
  if (!rc) rc = reclaim(sv, false);
  if (rc) return rc;
  while (!(*out = sv->arena.take(bytes)) && !sv->fenced.empty() &&
         sv->arena.reserved() + std::max(bytes, sv->slab_bytes) > sv->cfg.budget_bytes) {
    rc = reclaim(sv, true);
    if (rc) return rc;
  }
  if (*out) return SDR_OK;
  const hipError_t e = sv->arena.grow(bytes, sv->slab_bytes);
  if (e != hipSuccess)
    return fail(SDR_E_CAPACITY, "device pool: reserving %llu bytes: %s",
                static_cast<unsigned long long>(std::max(bytes, sv->slab_bytes)), hipGetErrorString(e));
  *out = sv->arena.take(bytes);
  return *out ? SDR_OK : fail(SDR_E_CAPACITY, "device pool: no range after growth");
}

// Allocate the request's outputs from the pool and point its column struct
// at them.
int allocate_outputs(sdr_server* sv, Req* q)
{
  int rc = pool_alloc(sv, std::max<uint64_t>(q->out_bytes, 256), &q->d_alloc);
  if (rc) return rc;
  if (q->r.is_string) {
    const uint64_t off_bytes = (uint64_t{q->r.total_rows} + 1u) * 4u;
    q->scol.d_offsets        = reinterpret_cast<int32_t*>(q->d_alloc);
    q->scol.d_chars          = q->d_alloc + ((off_bytes + 255u) & ~uint64_t{255});
    q->scol.chars_capacity   = q->chars_cap;
  } else {
    q->col.d_out = q->d_alloc;
  }
  return SDR_OK;
}

// Finish every request of a slot.  `rc` is the launch's own outcome (a HIP
// failure applies to all of its requests); a persistent launch's string
// requests then take their own column's outcome, so one request's overflow
// never marks the others of its batch.
void complete(sdr_server* sv, Slot& sl, int rc, bool timed = true)
{
  const double t_start = timed ? sv->gpu_ms(sl.ev_start) : sv->now_ms();
  const double t_end   = timed ? sv->gpu_ms(sl.ev_end) : t_start;
  if (timed) {
    sv->stats.gpu_busy_ms += t_end - t_start;
    const double from = std::max(t_start, sv->union_end);
    if (t_end > from) sv->stats.gpu_union_ms += t_end - from;
    sv->union_end = std::max(sv->union_end, t_end);
  }
  sv->stats.requests += sl.reqs.size();
  for (Req* q : sl.reqs) {
    int st = rc;
    if (st == SDR_OK && sl.batch && q->r.is_string) st = batch_column_status(sl.batch, &q->scol);
    if (q->r.is_string && st == SDR_OK && q->scol.chars_used > q->chars_cap) st = SDR_E_OVERFLOW;
    q->status = st;
    q->done   = true;
    sdr_completion c{};
    c.tag        = q->r.tag;
    c.status     = st;
    c.d_out      = q->r.is_string ? static_cast<void*>(q->scol.d_offsets) : q->col.d_out;
    c.d_chars    = q->r.is_string ? q->scol.d_chars : nullptr;
    c.chars_used = q->r.is_string ? q->scol.chars_used : 0;
    c.out_bytes  = q->out_bytes;
    c.submit_ms  = q->t_submit;
    c.start_ms   = t_start;
    c.done_ms    = t_end;
    sv->done.push_back(c);
  }
  sl.reqs.clear();
  sl.busy = false;
  --sv->running;
}

int collect(sdr_server* sv)
{
  for (auto& sl : sv->slots) {
    if (!sl.busy) continue;
    int rc = SDR_OK;
    // the end event is the slot's last work (after a batch's read-back), so
    // once it has passed the launch is done and its times are readable
    const hipError_t e = hipEventQuery(sl.ev_end);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) return fail(SDR_E_HIP, "server launch: %s", hipGetErrorString(e));
    if (sl.batch) {
      const int r = batch_ready(sl.batch);
      if (r < 0) return r;
      (void)batch_collect(sl.batch);  // per-column outcomes: batch_column_status
    } else {
      rc = sl.launch_status;
    }
    complete(sv, sl, rc);
  }
  return SDR_OK;
}

void admit(sdr_server* sv)
{
  while (!sv->queue.empty()) {
    Req* q = sv->queue.front();
    // a request larger than the budget runs alone
    if (held(sv) + q->out_bytes > sv->cfg.budget_bytes && sv->outputs_held != 0) break;
    sv->outputs_held += q->out_bytes;
    sv->admitted.push_back(q);
    sv->queue.pop_front();
  }
}

int launch_slot(sdr_server* sv, Slot& sl);

// Move admitted requests into a free slot and launch them.  A failure before
// the launch is enqueued completes every request of the slot with that
// error (their outputs go back to the pool), so the server keeps serving.
int launch(sdr_server* sv, Slot& sl, uint32_t* n_launched)
{
  const uint32_t cap = sv->cfg.mode == SDR_LAUNCH_PERSISTENT
                         ? (sv->cfg.max_batch_requests ? sv->cfg.max_batch_requests : UINT32_MAX)
                         : 1u;
  while (!sv->admitted.empty() && sl.reqs.size() < cap) {
    sl.reqs.push_back(sv->admitted.front());
    sv->admitted.pop_front();
  }
  const int rc = launch_slot(sv, sl);
  if (rc == SDR_OK) {
    sl.busy = true;
    ++sv->running;
    ++sv->stats.launches;
    if (n_launched) ++*n_launched;
    return SDR_OK;
  }
  for (Req* q : sl.reqs) {  // nothing was enqueued on them
    if (q->d_alloc) (void)sv->arena.put(q->d_alloc);
    q->d_alloc        = nullptr;
    q->col.d_out      = nullptr;
    q->scol.d_offsets = nullptr;
    q->scol.d_chars   = nullptr;
  }
  if (sl.batch) (void)sdr_batch_clear(sl.batch);
  ++sv->running;  // complete() counts the slot down
  complete(sv, sl, rc, false);
  return SDR_OK;
}

int launch_slot(sdr_server* sv, Slot& sl)
{
  {
    PhaseTimer pt(sv->stats.host_alloc_ms);
    for (Req* q : sl.reqs) {
      int rc = allocate_outputs(sv, q);
      if (rc) return rc;
    }
  }
  PhaseTimer pt(sv->stats.host_launch_ms);
  // outputs may be ranges other streams released: their fences have passed
  // on the host (reclaim), so no device wait is needed for them; segments
  // staged since this slot's last launch are waited for on the device
  if (sv->stage_done && sl.stage_seen != sv->stage_seq) {
    HIP_TRY(hipStreamWaitEvent(sl.s, sv->stage_done, 0));
    sl.stage_seen = sv->stage_seq;
  }
  HIP_TRY(hipEventRecord(sl.ev_start, sl.s));
  int rc = SDR_OK;
  if (sl.batch) {
    rc = sdr_batch_clear(sl.batch);
    for (Req* q : sl.reqs) {
      if (rc) break;
      rc = q->r.is_string ? batch_add_strings(sl.batch, &q->scol) : batch_add_fixed(sl.batch, &q->col, 1);
    }
    if (!rc) rc = batch_launch(sl.batch, sl.s);
    if (!rc) sv->stats.queue_entries += static_cast<uint64_t>(sdr_batch_num_entries(sl.batch));
  } else {
    Req* q = sl.reqs.front();
    if (q->r.is_string) rc = sdr_decode_strings_column_mode(&q->scol, sl.s, sv->cfg.mode);
    else rc = sdr_decode_table_mode(&q->col, 1, sl.s, sv->cfg.mode);
    sl.launch_status = rc;
    if (rc == SDR_E_OVERFLOW) rc = SDR_OK;  // reported on the request, not the server
  }
  if (rc) return rc;
  HIP_TRY(hipEventRecord(sl.ev_end, sl.s));
  return SDR_OK;
}

}  // namespace
}  // namespace sdr

using namespace sdr;

extern "C" {

int sdr_server_create(const sdr_server_config* config, sdr_server** out)
{
  clear_error();
  if (!out) return fail(SDR_E_INVALID_ARG, "out is NULL");
  *out = nullptr;
  auto sv = std::make_unique<sdr_server>();
  if (config) sv->cfg = *config;
  if (!sv->cfg.budget_bytes) sv->cfg.budget_bytes = DEFAULT_BUDGET;
  if (!sv->cfg.max_inflight) sv->cfg.max_inflight = 2;
  if (sv->cfg.max_inflight > 64) return fail(SDR_E_INVALID_ARG, "max_inflight %u > 64", sv->cfg.max_inflight);
  if (sv->cfg.mode > SDR_LAUNCH_PERSISTENT) return fail(SDR_E_INVALID_ARG, "unknown launch mode %u", sv->cfg.mode);
  sv->t0 = Clock::now();
  HIP_TRY(hipGetDevice(&sv->device));
  // slabs of 1/32 of the budget, between 64 MiB (or the whole budget, if
  // smaller) and 4 GiB
  sv->slab_bytes = std::min(SLAB_MAX, std::max(std::min(SLAB_MIN, sv->cfg.budget_bytes), sv->cfg.budget_bytes / 32));
  HIP_TRY(hipStreamCreateWithFlags(&sv->aux, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&sv->fence_s, hipStreamNonBlocking));
  sv->slots.resize(sv->cfg.max_inflight);
  for (auto& sl : sv->slots) {
    HIP_TRY(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&sl.ev_start));
    HIP_TRY(hipEventCreate(&sl.ev_end));
    if (sv->cfg.mode == SDR_LAUNCH_PERSISTENT) {
      int rc = sdr_batch_create(&sl.batch);
      // a slot's launches enqueue the group kernels directly: consecutive
      // launches differ in the groups they hold, and a graph would be
      // destroyed and instantiated again right after the last launch of
      // the previous one completed
      if (!rc) rc = sdr_batch_set_option(sl.batch, SDR_BATCH_USE_GRAPH, 0);
      if (rc) return rc;
    }
  }
  HIP_TRY(hipEventCreate(&sv->t0_event));
  HIP_TRY(hipEventRecord(sv->t0_event, sv->aux));
  HIP_TRY(hipEventSynchronize(sv->t0_event));
  sv->t0_gpu_ms = sv->now_ms();
  *out = sv.release();
  return SDR_OK;
}

int sdr_server_stage(sdr_server* sv, const void* host_bytes, uint64_t n, const uint8_t** d_out)
{
  clear_error();
  if (!sv || !d_out || (n && !host_bytes)) return fail(SDR_E_INVALID_ARG, "NULL argument");
  *d_out = nullptr;
  PhaseTimer pt(sv->stats.host_stage_ms);
  if (held(sv) + n > sv->cfg.budget_bytes)
    return fail(SDR_E_CAPACITY, "staging %llu bytes exceeds the budget", static_cast<unsigned long long>(n));
  uint8_t* d = nullptr;
  int rc     = pool_alloc(sv, std::max<uint64_t>(n, 16), &d);
  if (rc) return rc;
  // chunks go through the ring of pinned buffers: the host copies into a
  // buffer once the copy that last used it has finished (a wait only when
  // the ring is full), the device copies on the staging stream
  const auto* src = static_cast<const uint8_t*>(host_bytes);
  for (uint64_t o = 0; o < n; o += STAGE_CHUNK) {
    const size_t m   = static_cast<size_t>(std::min<uint64_t>(STAGE_CHUNK, n - o));
    const uint32_t k = sv->ring_next;
    sv->ring_next    = (k + 1) % STAGE_RING;
    hipError_t e     = hipSuccess;
    if (!sv->ring[k]) {
      e = hipHostMalloc(reinterpret_cast<void**>(&sv->ring[k]), STAGE_CHUNK, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->ring_ev[k], hipEventDisableTiming);
    } else if (sv->ring_used[k]) {
      e = hipEventSynchronize(sv->ring_ev[k]);
    }
    if (e == hipSuccess) {
      std::memcpy(sv->ring[k], src + o, m);
      e = hipMemcpyAsync(d + o, sv->ring[k], m, hipMemcpyHostToDevice, sv->aux);
    }
    if (e == hipSuccess) e = hipEventRecord(sv->ring_ev[k], sv->aux);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(sv->aux);  // nothing may still write the range
      (void)sv->arena.put(d);
      return fail(SDR_E_HIP, "stage: %s", hipGetErrorString(e));
    }
    sv->ring_used[k] = true;
  }
  if (!sv->stage_done) HIP_TRY(hipEventCreateWithFlags(&sv->stage_done, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(sv->stage_done, sv->aux));
  ++sv->stage_seq;
  sv->staged[d] = n;
  sv->resident += n;
  *d_out = d;
  return SDR_OK;
}

int sdr_server_stage_wait(sdr_server* sv, sdr_stream_t stream)
{
  clear_error();
  if (!sv) return fail(SDR_E_INVALID_ARG, "server is NULL");
  if (sv->stage_done) HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(stream), sv->stage_done, 0));
  return SDR_OK;
}

int sdr_server_unstage(sdr_server* sv, const uint8_t* d_ptr)
{
  clear_error();
  if (!sv) return fail(SDR_E_INVALID_ARG, "server is NULL");
  auto it = sv->staged.find(d_ptr);
  if (it == sv->staged.end()) return fail(SDR_E_NOT_FOUND, "pointer was not staged by this server");
  PhaseTimer pt(sv->stats.host_stage_ms);
  // the range rejoins the pool once every launch in flight (which may read
  // it) and its own staging copy have finished: a fence on a stream that
  // waits for both, so the host never blocks here
  for (auto& sl : sv->slots)
    if (sl.busy) HIP_TRY(hipStreamWaitEvent(sv->fence_s, sl.ev_end, 0));
  if (sv->stage_done) HIP_TRY(hipStreamWaitEvent(sv->fence_s, sv->stage_done, 0));
  hipEvent_t ev = nullptr;
  int rc        = get_event(sv, &ev);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(ev, sv->fence_s));
  sv->fenced.push_back({{const_cast<uint8_t*>(d_ptr)}, ev});
  sv->resident -= it->second;
  sv->staged.erase(it);
  return SDR_OK;
}

int sdr_server_submit(sdr_server* sv, const sdr_request* r)
{
  clear_error();
  if (!sv || !r) return fail(SDR_E_INVALID_ARG, "NULL argument");
  PhaseTimer pt(sv->stats.host_submit_ms);
  if (sv->reqs.count(r->tag)) return fail(SDR_E_INVALID_ARG, "tag %llu is pending", static_cast<unsigned long long>(r->tag));
  if (r->n_data && !r->data) return fail(SDR_E_INVALID_ARG, "data is NULL");
  auto q = std::make_unique<Req>();
  q->r = *r;
  size_t nseg = 0;
  for (uint32_t i = 0; i < r->n_data; ++i) nseg += r->data[i].segments ? r->data[i].n_segments : 0;
  q->segs.reserve(nseg);
  for (uint32_t i = 0; i < r->n_data; ++i) {
    const sdr_run& in = r->data[i];
    const size_t first = q->segs.size();
    for (uint32_t k = 0; in.segments && k < in.n_segments; ++k) q->segs.push_back(in.segments[k]);
    q->runs.push_back({in.codec, static_cast<uint32_t>(q->segs.size() - first), nullptr});
  }
  for (size_t i = 0, o = 0; i < q->runs.size(); ++i) {
    q->runs[i].segments = q->segs.data() + o;
    o += q->runs[i].n_segments;
  }
  q->r.data = q->runs.data();
  // validate with placeholder outputs (validation never dereferences them)
  uint8_t* const placeholder = reinterpret_cast<uint8_t*>(uintptr_t{256});
  int rc;
  if (r->is_string) {
    q->chars_cap = r->chars_capacity ? r->chars_capacity : sdr_strings_upper_bound(q->runs.data(), r->n_data);
    if (q->chars_cap == UINT64_MAX)
      return fail(SDR_E_INVALID_ARG, "VARCHAR request without chars_capacity and with an unknown max_string_length");
    auto& s = q->scol;
    s.total_rows = r->total_rows;
    s.data       = q->runs.data();
    s.n_data     = r->n_data;
    s.d_offsets  = reinterpret_cast<int32_t*>(placeholder);
    rc           = validate_string_column(&s);
    if (!rc && sv->cfg.mode == SDR_LAUNCH_PERSISTENT && has_dict_fsst(&s))
      rc = fail(SDR_E_UNSUPPORTED_CODEC,
                "DICT_FSST requests need host round trips for their dictionaries: serve them in a drop-in mode");
    q->out_bytes = (((uint64_t{r->total_rows} + 1u) * 4u + 255u) & ~uint64_t{255}) + q->chars_cap;
  } else {
    auto& c = q->col;
    c.type_size  = r->type_size;
    c.total_rows = r->total_rows;
    c.data       = q->runs.data();
    c.n_data     = r->n_data;
    c.d_out      = placeholder;
    rc           = validate_fixed_columns(&c, 1);
    q->out_bytes = uint64_t{r->total_rows} * r->type_size;
  }
  if (rc) return rc;
  q->t_submit = sv->now_ms();
  sv->queue.push_back(q.get());
  sv->reqs[r->tag] = std::move(q);
  return SDR_OK;
}

int sdr_server_progress(sdr_server* sv, uint32_t* n_launched)
{
  clear_error();
  if (!sv) return fail(SDR_E_INVALID_ARG, "server is NULL");
  if (n_launched) *n_launched = 0;
  int rc;
  {
    PhaseTimer pt(sv->stats.host_collect_ms);
    rc = collect(sv);
    if (!rc) rc = close_releases(sv);
    if (!rc) rc = reclaim(sv, false);
  }
  if (rc) return rc;
  admit(sv);
  for (auto& sl : sv->slots) {
    if (sv->admitted.empty()) break;
    if (sl.busy) continue;
    rc = launch(sv, sl, n_launched);
    if (rc) return rc;
  }
  return SDR_OK;
}

int sdr_server_poll(sdr_server* sv, sdr_completion* out, uint32_t cap, uint32_t* n)
{
  clear_error();
  if (!sv || !n || (cap && !out)) return fail(SDR_E_INVALID_ARG, "NULL argument");
  const uint32_t m = static_cast<uint32_t>(std::min<size_t>(cap, sv->done.size()));
  std::copy(sv->done.begin(), sv->done.begin() + m, out);
  sv->done.erase(sv->done.begin(), sv->done.begin() + m);
  *n = m;
  return SDR_OK;
}

int sdr_server_release(sdr_server* sv, uint64_t tag, sdr_stream_t stream)
{
  clear_error();
  if (!sv) return fail(SDR_E_INVALID_ARG, "server is NULL");
  auto it = sv->reqs.find(tag);
  if (it == sv->reqs.end()) return fail(SDR_E_NOT_FOUND, "unknown tag %llu", static_cast<unsigned long long>(tag));
  Req* q = it->second.get();
  PhaseTimer pt(sv->stats.host_release_ms);
  if (!q->done) return fail(SDR_E_STATE, "request %llu has not completed", static_cast<unsigned long long>(tag));
  // the range rejoins the pool behind an event on the consumer's stream,
  // recorded once for all the releases on that stream before the next
  // progress or allocation
  if (q->d_alloc) sv->open_release[static_cast<hipStream_t>(stream)].push_back(q->d_alloc);
  sv->outputs_held -= q->out_bytes;
  sv->reqs.erase(it);
  return SDR_OK;
}

int sdr_server_pending(const sdr_server* sv)
{
  if (!sv) return SDR_E_INVALID_ARG;
  size_t inflight = 0;
  for (auto& sl : sv->slots) inflight += sl.reqs.size();
  return static_cast<int>(sv->queue.size() + sv->admitted.size() + inflight);
}

uint64_t sdr_server_held_bytes(const sdr_server* sv) { return sv ? held(sv) : 0; }

double sdr_server_now_ms(const sdr_server* sv) { return sv ? sv->now_ms() : 0.0; }

int sdr_server_get_stats(const sdr_server* sv, sdr_server_stats* out)
{
  clear_error();
  if (!sv || !out) return fail(SDR_E_INVALID_ARG, "NULL argument");
  *out = sv->stats;
  return SDR_OK;
}

int sdr_server_destroy(sdr_server* sv)
{
  delete sv;
  return SDR_OK;
}

}  // extern "C"
