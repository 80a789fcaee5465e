"""Served mixed-codec decode stream: the workload and driver of the serving
loop (sdr_server_*).

Resident pool: for each codec of the mix, one source column of whole
122880-row row groups (enough for the largest request) and a ladder of tail
columns of log-spaced row counts below one row group, DuckDB-format
segments (self-encoded), staged into the server's device memory through its
pinned staging ring.  Each staged column is decoded once into a reference
copy and checked bit-exact against the generated values.  The pool can hold
`replicas` copies of every column at distinct device addresses: each
request reads one copy, drawn at random, so a large pool makes requests read
their segments from HBM rather than from a cache that earlier requests
warmed.

Requests: sizes drawn log-uniform from [min_rows, max_rows]; a request of R
rows decodes consecutive source segments from a random start while they
fit in R, then every segment of the largest tail column that fits the rest,
so every request is a real multi-segment column.  Codec mix 40% BITPACKING INT32 / 20% RLE INT64
/ 25% ALP DOUBLE / 15% FSST VARCHAR by default.  Arrivals are Poisson at a
fixed rate (or all at time 0 to measure capacity).

The driver submits each request at its arrival time, drives
sdr_server_progress, collects completions, checks every `verify_every`-th
output against the reference copy on the device, and releases outputs.
Latency is completion (GPU end-of-launch event) minus scheduled arrival.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from . import encode as E
from . import lib as L

SEG_ROWS = E.ROW_GROUP_ROWS
MIX = {"bp_int32": 0.40, "rle_int64": 0.20, "alp_double": 0.25, "fsst_varchar": 0.15}
_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("libamdhip64.so")
        _HIP.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _HIP.hipMemcpyAsync.restype = C.c_int
    return _HIP


def _gen(kind: str, seed: int, n: int):
    if kind == "bp_int32":
        return E.gen_bp_int32(seed, n)
    if kind == "rle_int64":
        return E.gen_runs_int64(seed, n, 16.0, 1 << 40)
    if kind == "alp_double":
        return E.gen_alp_doubles(seed, n)
    return E.gen_zipf_strings(seed, n)


def _encode(kind: str, values):
    if kind == "bp_int32":
        return E.encode_bitpacking(values)
    if kind == "rle_int64":
        return E.encode_rle(values)
    if kind == "alp_double":
        return E.encode_alp(values)
    return E.encode_fsst(*values)


@dataclass
class Column:
    """One staged column and its device reference decode."""
    kind: str
    enc: E.Encoded
    d_base: int  # device address of the staged blob
    ref: object = None  # torch uint8 (fixed) or (offsets int32, chars uint8)
    ref_offs: object = None  # VARCHAR: host int64 copy of the reference offsets

    @property
    def is_string(self) -> bool:
        return self.enc.codec == A.FSST

    @property
    def cum(self) -> np.ndarray:
        """First row of each segment, then the row count (n_segments + 1)."""
        return np.append(self.enc.segs["row_offset"].astype(np.int64), self.enc.n_rows)

    def seg(self, i: int, row_offset: int, base: int = 0) -> A.Segment:
        s = self.enc.segs[i]
        return A.Segment((base or self.d_base) + int(s["offset"]), int(s["bytes_size"]), row_offset,
                         int(s["row_count"]), int(s["max_string_length"]))


class Pool:
    """Resident source and tail columns per codec, staged into `server`."""

    def __init__(self, lib, server, kinds=tuple(MIX), max_rows=16 << 20, n_tails=48, seed=0xC0DE0005, log=None,
                 replicas=1, resident_bytes=0):
        """`replicas` copies of every column, or, when `resident_bytes` is
        given, as many as it takes to hold about that many bytes."""
        import torch

        self.lib, self.server = lib, server
        self.src, self.tails = {}, {}
        self.copies = {}  # column id -> device bases of copies 1..replicas-1
        self.n_src_segs = -(-max_rows // SEG_ROWS)  # source rows: whole row groups covering max_rows
        self.staged_bytes = 0
        tail_sizes = np.unique(np.round(np.geomspace(2048, SEG_ROWS - 1, n_tails)).astype(np.int64))
        for ki, kind in enumerate(kinds):
            t = time.time()
            vals = _gen(kind, seed + 16 * ki, self.n_src_segs * SEG_ROWS)
            self.src[kind] = self._stage(kind, vals)
            self.tails[kind] = []
            for ti, n in enumerate(tail_sizes):
                tv = _gen(kind, seed + 16 * ki + 1 + 7919 * ti, int(n))
                self.tails[kind].append(self._stage(kind, tv))
            torch.cuda.synchronize()
            if log:
                log(f"[serve] pool {kind}: source {self.n_src_segs} segments + {len(tail_sizes)} tails, "
                    f"{time.time() - t:.1f}s")
        self.copy_bytes = self.staged_bytes
        self.replicas = max(1, int(replicas))
        if resident_bytes:
            self.replicas = max(1, -(-int(resident_bytes) // self.copy_bytes))
        t = time.time()
        cols = [c for k in self.src for c in [self.src[k], *self.tails[k]]]
        for col in cols:
            self.copies[id(col)] = []
        for _ in range(self.replicas - 1):
            for col in cols:
                self.copies[id(col)].append(self._stage_bytes(col.enc))
        torch.cuda.synchronize()
        if log and self.replicas > 1:
            log(f"[serve] pool: {self.replicas} copies at distinct addresses, {self.staged_bytes / 1e9:.1f} GB "
                f"resident, copies staged in {time.time() - t:.1f}s")
        self.tail_rows = [int(x) for x in tail_sizes]

    def _stage_bytes(self, enc) -> int:
        d = C.c_void_p()
        L.check(self.lib.sdr_server_stage(self.server, enc.blob.ctypes.data, enc.blob.size, C.byref(d)))
        self.staged_bytes += enc.blob.size
        return d.value

    def base(self, col: Column, replica: int) -> int:
        """Device address of copy `replica` of a staged column."""
        return col.d_base if replica == 0 else self.copies[id(col)][replica - 1]

    def _stage(self, kind, values) -> Column:
        import torch

        enc = _encode(kind, values)
        col = Column(kind, enc, self._stage_bytes(enc))
        # reference decode of the whole column, checked against the values
        # (on torch's stream, after the staging copy)
        dev = torch.device("cuda:0")
        sh = torch.cuda.current_stream().cuda_stream
        L.check(self.lib.sdr_server_stage_wait(self.server, sh))
        runs, keep = A.make_runs([(enc.codec, [col.seg(i, int(s["row_offset"])) for i, s in enumerate(enc.segs)])])
        if col.is_string:
            chars, offs = values
            d_offs = torch.empty(enc.n_rows + 1, dtype=torch.int32, device=dev)
            d_chars = torch.empty(max(1, len(chars)), dtype=torch.uint8, device=dev)
            sc = A.StringColumn()
            sc.total_rows, sc.data, sc.n_data = enc.n_rows, runs, 1
            sc.d_offsets, sc.d_chars, sc.chars_capacity = d_offs.data_ptr(), d_chars.data_ptr(), len(chars)
            L.check(self.lib.sdr_decode_strings_column(C.byref(sc), sh))
            ok = (sc.chars_used == len(chars)
                  and torch.equal(d_offs.cpu(), torch.from_numpy(offs.astype(np.int32)))
                  and torch.equal(d_chars[: len(chars)].cpu(), torch.from_numpy(chars)))
            col.ref = (d_offs, d_chars)
            col.ref_offs = offs.astype(np.int64)  # host copy: slicing bounds without a device sync
        else:
            out = torch.empty(enc.n_rows * enc.type_size, dtype=torch.uint8, device=dev)
            c = A.Column()
            c.type_size, c.total_rows, c.data, c.n_data, c.d_out = enc.type_size, enc.n_rows, runs, 1, out.data_ptr()
            arr = (A.Column * 1)(c)
            L.check(self.lib.sdr_decode_table(arr, 1, sh))
            ok = torch.equal(out.cpu(), torch.from_numpy(values.view(np.uint8)))
            col.ref = out
        if not ok:
            raise RuntimeError(f"reference decode of a {kind} pool column is not bit-exact")
        return col


@dataclass
class Spec:
    """One request: which pool segments, in order, and its expected output."""
    tag: int
    kind: str
    rows: int
    src_first: int  # first source segment
    n_src: int      # source segments
    tail: int       # tail column index or -1
    replica: int = 0  # which copy of the pool's columns it reads
    arrival_ms: float = 0.0
    comp_bytes: int = 0
    keep: list = field(default_factory=list)

    def request(self, pool: Pool) -> A.Request:
        src = pool.src[self.kind]
        segs, row = [], 0
        sb = pool.base(src, self.replica)
        for i in range(self.src_first, self.src_first + self.n_src):
            seg = src.seg(i, row, sb)
            segs.append(seg)
            row += seg.row_count
        if self.tail >= 0:
            tcol = pool.tails[self.kind][self.tail]
            tb = pool.base(tcol, self.replica)
            for j in range(len(tcol.enc.segs)):
                seg = tcol.seg(j, row, tb)
                segs.append(seg)
                row += seg.row_count
        runs, keep = A.make_runs([(src.enc.codec, segs)])
        self.keep = [runs, keep]
        self.comp_bytes = sum(int(g.bytes_size) for g in segs)
        r = A.Request()
        r.tag, r.is_string, r.type_size, r.total_rows = self.tag, int(src.is_string), src.enc.type_size, row
        r.n_data, r.data = 1, runs
        return r


def make_specs(pool: Pool, n: int, seed: int = 0xC0DE0005, min_rows: int = 2048, max_rows: int = 16 << 20,
               mix=None) -> list:
    """Requests of log-uniform size: whole source segments from a random
    start while they fit, then the largest tail column that still fits."""
    mix = mix or MIX
    rng = np.random.default_rng(seed)
    kinds = [k for k in mix if k in pool.src]
    p = np.array([mix[k] for k in kinds], float)
    p /= p.sum()
    sizes = np.exp(rng.uniform(np.log(min_rows), np.log(max_rows), n)).astype(np.int64)
    which = rng.choice(len(kinds), size=n, p=p)
    cums = {k: pool.src[k].cum for k in kinds}
    specs = []
    for i in range(n):
        k, r = kinds[which[i]], int(sizes[i])
        cum = cums[k]
        nseg = len(cum) - 1
        per = cum[-1] / nseg
        want = min(int(r // per), nseg)
        first = int(rng.integers(0, nseg - want + 1))
        last = int(np.searchsorted(cum, cum[first] + r, side="right")) - 1  # segments [first, last) fit
        last = max(first, min(last, nseg))
        rem = r - int(cum[last] - cum[first])
        tail = int(np.searchsorted(pool.tail_rows, rem, side="right")) - 1 if rem >= pool.tail_rows[0] else -1
        if last == first and tail < 0:
            tail = 0
        rows = int(cum[last] - cum[first]) + (pool.tail_rows[tail] if tail >= 0 else 0)
        replica = int(rng.integers(0, pool.replicas)) if pool.replicas > 1 else 0
        specs.append(Spec(i, k, rows, first, last - first, tail, replica))
    return specs


def poisson_arrivals(specs: list, rate_per_s: float, seed: int = 0xC0DE0005, t0_ms: float = 0.0):
    """Arrival times (ms) at `rate_per_s` requests/s; rate <= 0: all at t0."""
    if rate_per_s <= 0:
        for s in specs:
            s.arrival_ms = t0_ms
        return
    rng = np.random.default_rng(seed ^ 0x5A5A)
    t = t0_ms + np.cumsum(rng.exponential(1000.0 / rate_per_s, len(specs)))
    for s, a in zip(specs, t):
        s.arrival_ms = float(a)


class _DeviceView:
    """Zero-copy torch view of device bytes through __cuda_array_interface__."""

    def __init__(self, ptr: int, nbytes: int):
        self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False),
                                         "version": 3, "strides": None}


_ZERO_COPY = True


def _view(ptr: int, nbytes: int):
    """uint8 tensor over device bytes: a zero-copy view when torch accepts
    the array interface, else a stream-ordered copy."""
    global _ZERO_COPY
    import torch

    if nbytes == 0:
        return torch.empty(0, dtype=torch.uint8, device="cuda:0")
    if _ZERO_COPY:
        try:
            return torch.as_tensor(_DeviceView(ptr, nbytes), device="cuda:0")
        except (TypeError, RuntimeError, ValueError):
            _ZERO_COPY = False
    t = torch.empty(nbytes, dtype=torch.uint8, device="cuda:0")
    if _hip().hipMemcpyAsync(t.data_ptr(), ptr, nbytes, 3, torch.cuda.current_stream().cuda_stream) != 0:
        raise RuntimeError("hipMemcpyAsync failed")
    return t


def _differs(out, ref):
    """0-dim bool device tensor: some element of `out` differs from `ref`
    (equal shapes; no host synchronisation)."""
    return (out != ref).any()


def _word_view(t, ts: int):
    """Byte tensor as ts-byte words where ts is 4 or 8 (fewer elements to
    compare); other widths stay bytes."""
    import torch

    return t.view({4: torch.int32, 8: torch.int64}[ts]) if ts in (4, 8) else t


def mismatches(pool: Pool, spec: Spec, comp: A.Completion):
    """0-dim bool device tensor, True when the request's output differs
    from the pool's reference decode anywhere (enqueued on torch's current
    stream, no host synchronisation).  The source rows and the tail rows
    are compared as two slices of the output, in words of the value width."""
    import torch

    src = pool.src[spec.kind]
    cum = src.cum
    r0 = int(cum[spec.src_first])
    ns = int(cum[spec.src_first + spec.n_src]) - r0
    tcol = pool.tails[spec.kind][spec.tail] if spec.tail >= 0 else None
    if src.is_string:
        ro_h = src.ref_offs
        ro, rc = src.ref
        c0, c1 = int(ro_h[r0]), int(ro_h[r0 + ns])
        n_tail_chars = int(tcol.ref_offs[-1]) if tcol is not None else 0
        if comp.chars_used != (c1 - c0) + n_tail_chars:
            return torch.ones((), dtype=torch.bool, device="cuda:0")
        offs = _view(comp.d_out, (spec.rows + 1) * 4).view(torch.int32)
        chars = _view(comp.d_chars, comp.chars_used)
        d = _differs(offs[:ns + 1] + c0, ro[r0:r0 + ns + 1])
        d = d | _differs(chars[:c1 - c0], rc[c0:c1])
        if tcol is not None:
            to, tc = tcol.ref
            d = d | _differs(offs[ns:] - (c1 - c0), to)
            d = d | _differs(chars[c1 - c0:], tc[:n_tail_chars])
        return d
    ts = src.enc.type_size
    out = _view(comp.d_out, spec.rows * ts)
    d = _differs(_word_view(out[:ns * ts], ts), _word_view(src.ref[r0 * ts:(r0 + ns) * ts], ts))
    if tcol is not None:
        d = d | _differs(_word_view(out[ns * ts:], ts), _word_view(tcol.ref, ts))
    return d


def verify(pool: Pool, spec: Spec, comp: A.Completion, stream=None) -> bool:
    """True when the request's output equals the pool's reference decode."""
    return not bool(mismatches(pool, spec, comp).item())


def run(lib, pool: Pool, specs: list, mode: int, max_inflight: int = 2, budget: int = 0,
        max_batch_requests: int = 0, verify_every: int = 8, server=None, log=None) -> dict:
    """Serve `specs` (arrival_ms relative to the start of the run) and return
    throughput, latency percentiles and verification counts.

    Request descriptors are built before the clock starts; each request is
    submitted when its arrival time has passed (all at once when every
    arrival is 0, which measures capacity).  The outputs of every
    `verify_every`-th request are held until the last request has
    completed, then compared with the reference on torch's stream (so the
    comparison does not share the GPU with the timed stream) and released
    after the comparison in stream order; `verified_ok` counts the sampled
    requests whose output equals the reference."""
    import torch

    own = server is None
    if own:
        cfg = A.ServerConfig(budget, max_inflight, mode, max_batch_requests, 0)
        server = C.c_void_p()
        L.check(lib.sdr_server_create(C.byref(cfg), C.byref(server)))
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    comps = (A.Completion * 256)()
    n_out = C.c_uint32()
    launched = C.c_uint32()
    by_tag = {s.tag: s for s in specs}
    order = sorted(specs, key=lambda s: s.arrival_ms)
    reqs = [s.request(pool) for s in order]
    is_str = {k: pool.src[k].is_string for k in pool.src}
    lat, done_ms, statuses = [], [], []
    bytes_out = bytes_in = 0
    n_verified = 0
    checks = []  # (tag, 0-dim bool device tensor: output differs) per verified request
    sampled = []  # (spec, completion) of the requests to verify, outputs held until then
    n_launches = 0
    i = 0
    st0, st1 = A.ServerStats(), A.ServerStats()
    t_verify = 0.0
    L.check(lib.sdr_server_get_stats(server, C.byref(st0)))
    try:
        t_base = lib.sdr_server_now_ms(server)
        if all(s.arrival_ms <= 0 for s in order):  # capacity: queue everything first
            for r in reqs:
                L.check(lib.sdr_server_submit(server, C.byref(r)))
            i = len(order)
            t_base = lib.sdr_server_now_ms(server)
        while len(lat) < len(specs):
            now = lib.sdr_server_now_ms(server) - t_base
            while i < len(order) and order[i].arrival_ms <= now:
                L.check(lib.sdr_server_submit(server, C.byref(reqs[i])))
                i += 1
            L.check(lib.sdr_server_progress(server, C.byref(launched)))
            n_launches += launched.value
            L.check(lib.sdr_server_poll(server, comps, 256, C.byref(n_out)))
            for k in range(n_out.value):
                c = comps[k]
                s = by_tag[c.tag]
                statuses.append(c.status)
                lat.append(c.done_ms - t_base - max(s.arrival_ms, 0.0))
                done_ms.append(c.done_ms - t_base)
                bytes_out += (s.rows + 1) * 4 + c.chars_used if is_str[s.kind] else c.out_bytes
                bytes_in += s.comp_bytes
                if verify_every and s.tag % verify_every == 0 and c.status == 0:
                    kept = A.Completion()
                    C.pointer(kept)[0] = c
                    sampled.append((s, kept))  # checked and released once the stream is served
                    continue
                L.check(lib.sdr_server_release(server, c.tag, sh))
            if n_out.value == 0 and launched.value == 0:
                nxt = order[i].arrival_ms if i < len(order) else None
                if nxt is not None and nxt - now > 0.2:
                    time.sleep(min(0.001, (nxt - now) / 2000.0))
        L.check(lib.sdr_server_get_stats(server, C.byref(st1)))
        tv = time.perf_counter()
        for s, c in sampled:
            checks.append((s.tag, mismatches(pool, s, c)))
            L.check(lib.sdr_server_release(server, c.tag, sh))
        n_verified = len(sampled)
        t_verify = time.perf_counter() - tv
    finally:
        torch.cuda.synchronize()
        if own:
            lib.sdr_server_destroy(server)
    d0, d1 = st0.as_dict(), st1.as_dict()
    server_stats = {k: d1[k] - d0[k] for k in d1}
    server_stats["driver_verify_ms"] = t_verify * 1e3
    for s in specs:
        s.keep = []
    first = max(min(s.arrival_ms for s in specs), 0.0)
    span_ms = max(done_ms) - first
    lat = np.asarray(lat)
    flags = torch.stack([m for _, m in checks]).cpu().numpy() if checks else np.zeros(0, bool)
    n_bad = int(flags.sum())
    failed = []
    for (tag, _), f in zip(checks, flags):
        if f:
            sp = by_tag[tag]
            failed.append({"tag": tag, "kind": sp.kind, "rows": sp.rows, "src_first": sp.src_first,
                           "n_src": sp.n_src, "tail": sp.tail})
    return {"failed": failed[:20], "requests": len(specs), "launches": n_launches, "decoded_bytes": int(bytes_out),
            "compressed_bytes": int(bytes_in), "span_ms": span_ms, "decoded_gbs": bytes_out / (span_ms * 1e-3) / 1e9,
            "achieved_gbs": (bytes_out + bytes_in) / (span_ms * 1e-3) / 1e9,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "errors": int(sum(1 for x in statuses if x != 0)),
            "verified": n_verified, "verified_ok": n_verified - n_bad, "mismatched_requests": n_bad,
            "server_stats": server_stats, "latencies_ms": lat}


def merge_runs(runs: list) -> dict:
    """One result over several passes of run(): bytes, spans and server
    phase times add up, percentiles are over every request of every pass."""
    lat = np.concatenate([r["latencies_ms"] for r in runs])
    span = sum(r["span_ms"] for r in runs)
    dec = sum(r["decoded_bytes"] for r in runs)
    comp = sum(r["compressed_bytes"] for r in runs)
    stats = {k: sum(r["server_stats"][k] for r in runs) for k in runs[0]["server_stats"]}
    verified = sum(r["verified"] for r in runs)
    return {"failed": [f for r in runs for f in r["failed"]][:20], "requests": sum(r["requests"] for r in runs),
            "launches": sum(r["launches"] for r in runs), "decoded_bytes": dec, "compressed_bytes": comp,
            "span_ms": span, "decoded_gbs": dec / (span * 1e-3) / 1e9,
            "achieved_gbs": (dec + comp) / (span * 1e-3) / 1e9,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "errors": sum(r["errors"] for r in runs), "verified": verified,
            "verified_ok": sum(r["verified_ok"] for r in runs),
            "mismatched_requests": sum(r["mismatched_requests"] for r in runs),
            "server_stats": stats}
