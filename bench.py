#!/usr/bin/env python3
"""Decode throughput of DuckDB storage segments on one MI355X.

Headline workload ("served_mixed_10k", the served mixed-codec stream the
target is quoted on): 10,000 decode requests of log-uniform size (2 Ki to 16
Mi rows; 40% BITPACKING INT32, 20% RLE INT64, 25% ALP DOUBLE, 15% FSST
VARCHAR) over DuckDB-format segments (self-encoded; DuckDB is not installed)
resident in the decode server's device pool, Poisson arrivals at 0.8x the
measured capacity, a 200 GB admission budget.  A step is one pass of the
whole stream through the server (persistent group-kernel launches, continuous
batching); `value` is decoded bytes per second over the timed passes.

Beside it the run reports the rate sweep (0.5x, 0.95x), the baselines (one
drop-in launch per request, one launch per segment, each at 0.8x its own
capacity), configs 1-4 (8M-row bitpacking; the 64M-row integer/date mix;
ALP/ALP-RD doubles; FSST strings) through the batched, persistent and
one-launch-per-segment paths with per-kernel durations, the PCIe-inclusive
staging leg, and the CPU baseline per config.  The default run measures the
served stream and the other legs in two child processes, each with a
runtime and device memory of its own; this process only merges their
results.  `--workload NAME` makes one of configs 1-4 the headline instead.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "decoded GB/s and % HBM roofline vs DuckDB CPU; p50/p99 latency batched vs plain"
HBM_PEAK_GBS = 8000.0
SERVED = "served_mixed_10k"
WORKLOADS = (SERVED, "int_date_mix_64M", "copy_256M", "bp_int32_8M", "alp_double_32M", "fsst_varchar_16M")
CONFIGS = ("bp_int32_8M", "int_date_mix_64M", "alp_double_32M", "fsst_varchar_16M")  # configs 1-4
# Requests per persistent launch of the served stream.  Two launches of at
# most this many run side by side on the server's two slots, so one's FSST
# chunks overlap the other's fixed-width groups; on the 10,000-request
# stream, 128 gave 1200 GB/s capacity against 1040 for one launch of every
# queued request (64: 1120).
SERVED_MAX_BATCH = 128


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads: (name, values, Encoded) per column; strings carry (chars, offsets)
# ---------------------------------------------------------------------------
def build_workload(name: str, rows: int | None = None):
    import sirius_amd.encode as E

    t = time.time()
    if name == "copy_256M":  # calibration: UNCOMPRESSED INT64, a plain copy of known size
        rows = rows or 32 << 20
        cols = [("uncompressed_int64", E.gen_u64(0xC0DE00FF, rows).view(np.int64), E.encode_uncompressed)]
    elif name == "int_date_mix_64M":
        rows = rows or 64 << 20
        cols = [
            ("rle_int64_runs", E.gen_runs_int64(0xC0DE0002, rows, 16.0, 1 << 40), E.encode_rle),
            ("bp_int32_quantity", E.gen_uniform_int32(0xC0DE0012, rows, 1, 50), E.encode_bitpacking),
            ("bp_date_sorted", E.gen_dates_sorted(0xC0DE0022, rows), E.encode_bitpacking),
        ]
    elif name == "bp_int32_8M":
        rows = rows or 8 << 20
        cols = [("bp_int32_w17_w5", E.gen_bp_int32(0xC0DE0001, rows), E.encode_bitpacking)]
    elif name == "alp_double_32M":
        rows = rows or 32 << 20
        cols = [("alp_double", E.gen_alp_doubles(0xC0DE0003, rows), E.encode_alp),
                ("alprd_double", E.gen_unit_doubles(0xC0DE0013, rows // 4), E.encode_alprd)]
    elif name == "fsst_varchar_16M":
        rows = rows or 16 << 20
        chars, offs = E.gen_zipf_strings(0xC0DE0004, rows)
        enc = E.encode_fsst(chars, offs)
        log(f"[bench] encoded fsst_varchar: {len(enc.segs)} segments, {enc.compressed_bytes / 1e6:.1f} MB "
            f"compressed, {len(chars) / 1e6:.1f} MB of chars")
        return [("fsst_varchar", (chars, offs), enc)]
    else:
        raise ValueError(name)
    out = []
    for cname, values, enc_f in cols:
        enc = enc_f(values)
        out.append((cname, values, enc))
        log(f"[bench] encoded {cname}: {len(enc.segs)} segments, {enc.compressed_bytes / 1e6:.1f} MB "
            f"compressed, {values.nbytes / 1e6:.1f} MB decoded")
    log(f"[bench] workload {name} built in {time.time() - t:.1f}s")
    return out


class Resident:
    """Encoded columns uploaded to HBM with their output buffers."""

    def __init__(self, work, dev):
        import torch

        import sirius_amd._abi as A

        self.keep, self.fixed, self.outs, self.values = [], [], [], []
        self.strings = []  # (StringColumn, d_offs, d_chars, (chars, offs))
        self.comp_bytes = self.dec_bytes = 0
        self.per_codec = {}  # codec -> [compressed, decoded] bytes
        for _, values, enc in work:
            blob = torch.from_numpy(enc.blob).to(dev)
            self.keep.append(blob)
            runs, k = A.make_runs([(enc.codec, enc.segments(blob.data_ptr()))])
            self.keep.append(k)
            self.comp_bytes += enc.compressed_bytes
            if enc.codec == A.FSST:
                chars, offs = values
                n = enc.n_rows
                d_offs = torch.empty(n + 1, dtype=torch.int32, device=dev)
                d_chars = torch.empty(len(chars) + 64, dtype=torch.uint8, device=dev)
                col = A.StringColumn()
                col.total_rows, col.data, col.n_data = n, runs, 1
                col.d_offsets, col.d_chars, col.chars_capacity = d_offs.data_ptr(), d_chars.data_ptr(), len(chars) + 64
                self.strings.append((col, d_offs, d_chars, values))
                dec = len(chars) + 4 * (n + 1)
            else:
                out = torch.empty(enc.n_rows * enc.type_size, dtype=torch.uint8, device=dev)
                col = A.Column()
                col.type_size, col.total_rows, col.data, col.n_data, col.d_out = (
                    enc.type_size, enc.n_rows, runs, 1, out.data_ptr())
                self.fixed.append(col)
                self.outs.append(out)
                self.values.append(values)
                dec = enc.n_rows * enc.type_size
            self.dec_bytes += dec
            cb = self.per_codec.setdefault(enc.codec, [0, 0])
            cb[0] += enc.compressed_bytes
            cb[1] += dec
        self.n_fixed = len(self.fixed)
        # a table call decodes columns of one row count: one table per count
        self.tables = []
        for n in sorted({c.total_rows for c in self.fixed}):
            cs = [c for c in self.fixed if c.total_rows == n]
            self.tables.append(((A.Column * len(cs))(*cs), len(cs)))

    def decode(self, lib, sh, mode):
        from sirius_amd import lib as L

        for arr, n in self.tables:
            L.check(lib.sdr_decode_table_mode(arr, n, sh, mode))
        for col, *_ in self.strings:
            L.check(lib.sdr_decode_strings_column_mode(C.byref(col), sh, mode))

    def bit_exact(self) -> bool:
        ok = all(np.array_equal(o.cpu().numpy(), v.view(np.uint8)) for o, v in zip(self.outs, self.values))
        for col, d_offs, d_chars, (chars, offs) in self.strings:
            ok &= col.chars_used == len(chars)
            ok &= np.array_equal(d_offs.cpu().numpy().astype(np.uint64), offs)
            ok &= np.array_equal(d_chars[: len(chars)].cpu().numpy(), chars)
        return bool(ok)


def time_steps(fn, steps, warmup, stream):
    """Per-step durations (ms) from HIP events on `stream`, plus the total."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    evs[0].record(stream)
    for i in range(steps):
        fn()
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(steps)]
    return per, evs[0].elapsed_time(evs[-1])


def pct(xs, q):
    return float(np.percentile(np.asarray(xs), q))


def kernel_times(lib, res, sh, stream, steps, warmup):
    """Average duration (ms) of each codec's kernels inside a timed loop of
    whole decode steps (a prepared plan, one codec's launches bracketed by
    HIP events on the launch stream)."""
    import torch

    import sirius_amd._abi as A
    from sirius_amd import lib as L

    plans = []
    try:
        for arr, n in res.tables:
            plan = C.c_void_p()
            L.check(lib.sdr_plan_create(arr, n, A.LAUNCH_BATCHED, sh, C.byref(plan)))
            plans.append(plan)
        codecs = sorted(c for c in res.per_codec if c != A.FSST)
        for _ in range(warmup):
            for plan in plans:
                L.check(lib.sdr_plan_launch(plan, sh))
        torch.cuda.synchronize()
        evs = {c: [] for c in codecs}
        for _ in range(steps):
            for c in codecs:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for plan in plans:
                    L.check(lib.sdr_plan_launch_codec(plan, c, sh))
                e1.record(stream)
                evs[c].append((e0, e1))
        torch.cuda.synchronize()
        return {c: float(np.mean([a.elapsed_time(b) for a, b in evs[c]])) for c in codecs}
    finally:
        for plan in plans:
            lib.sdr_plan_destroy(plan)


def persistent_times(lib, res, sh, stream, steps, warmup):
    """The same columns through one persistent batch launch (sdr_batch, default options)."""
    from sirius_amd import lib as L

    b = C.c_void_p()
    L.check(lib.sdr_batch_create(C.byref(b)))
    try:
        def step():
            L.check(lib.sdr_batch_clear(b))
            for arr, n in res.tables:
                L.check(lib.sdr_batch_add_table(b, arr, n))
            for col, *_ in res.strings:
                L.check(lib.sdr_batch_add_strings(b, C.byref(col)))
            L.check(lib.sdr_batch_launch(b, sh))

        per, _ = time_steps(step, steps, warmup, stream)
        L.check(lib.sdr_batch_finish(b, sh))
        return per
    finally:
        lib.sdr_batch_destroy(b)


def measure(lib, work, dev, steps, warmup, plain_steps, full=True):
    import torch

    import sirius_amd._abi as A

    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    res = Resident(work, dev)
    r = {"rows": int(sum(e.n_rows for _, _, e in work)), "segments": int(sum(len(e.segs) for _, _, e in work)),
         "compressed_bytes": res.comp_bytes, "decoded_bytes": res.dec_bytes}
    per, tot = time_steps(lambda: res.decode(lib, sh, A.LAUNCH_BATCHED), steps, warmup, stream)
    r["bit_exact"] = res.bit_exact()
    ms = tot / steps
    r.update(ms_per_step=ms, decoded_gbs=res.dec_bytes / (ms * 1e-3) / 1e9,
             achieved_gbs=(res.comp_bytes + res.dec_bytes) / (ms * 1e-3) / 1e9,
             p50_ms=pct(per, 50), p99_ms=pct(per, 99))
    if not full:
        return r, res
    pp = persistent_times(lib, res, sh, stream, steps, warmup)
    r["persistent"] = dict(ms_per_step=float(np.mean(pp)), p50_ms=pct(pp, 50), p99_ms=pct(pp, 99),
                           decoded_gbs=res.dec_bytes / (np.mean(pp) * 1e-3) / 1e9, bit_exact=res.bit_exact())
    pl, tot = time_steps(lambda: res.decode(lib, sh, A.LAUNCH_PER_SEGMENT), plain_steps, 1, stream)
    r["plain"] = dict(ms_per_step=tot / plain_steps, p50_ms=pct(pl, 50), p99_ms=pct(pl, 99),
                      decoded_gbs=res.dec_bytes / (tot / plain_steps * 1e-3) / 1e9, bit_exact=res.bit_exact())
    r["batched_vs_plain"] = r["plain"]["ms_per_step"] / ms
    if res.n_fixed:
        kt = kernel_times(lib, res, sh, stream, steps, warmup)
        r["kernels"] = {A.CODEC_NAMES[c]: dict(us=kt[c] * 1e3, alg_bytes=sum(res.per_codec[c]),
                                               gbs=sum(res.per_codec[c]) / (kt[c] * 1e-3) / 1e9)
                        for c in kt}
    return r, res


def apply_tuning(lib):
    """Apply the autotuner's table (sirius_amd/tuning/gfx950.txt) if present;
    returns the knob values in effect."""
    import sirius_amd._abi as A
    from sirius_amd import lib as L
    from sirius_amd.autotune import DEFAULT_TABLE

    if DEFAULT_TABLE.exists():
        L.check(lib.sdr_tune_load(str(DEFAULT_TABLE).encode()))
        log(f"[bench] tuning table {DEFAULT_TABLE.name} applied")
    out = {}
    for k, name in A.TUNE_NAMES.items():
        vals = {w: int(lib.sdr_tune_get_width(k, w)) for w in A.TUNE_WIDTHS}
        out[name] = vals[4] if len(set(vals.values())) == 1 else {f"@{w}": v for w, v in vals.items()}
    return out


def rounded(x, nd=4):
    if isinstance(x, float):
        return round(x, nd)
    if isinstance(x, dict):
        return {k: rounded(v, nd) for k, v in x.items() if k != "latencies_ms"}
    return x


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=None)
    ap.add_argument("--workload", choices=WORKLOADS, default=SERVED)
    ap.add_argument("--only", choices=("all", "batched", "served"), default="all",
                    help="batched: the drop-in steps and the per-kernel loop only; served: the served "
                         "mixed-codec stream only (both for profiler runs)")
    ap.add_argument("--no-served-baselines", action="store_true",
                    help="serve the stream through the persistent path only")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other workloads")
    ap.add_argument("--served-resident-gb", type=float, default=32.0,
                    help="resident pool size of the served stream (copies of its columns at distinct addresses)")
    ap.add_argument("--served-max-batch", type=int, default=SERVED_MAX_BATCH,
                    help="requests per persistent launch of the served stream (0: every admitted request)")
    ap.add_argument("--served-requests", type=int, default=10000,
                    help="requests of the served mixed-codec stream (0: skip it)")
    ap.add_argument("--leg", choices=("served", "extra"), default=None,
                    help="run one leg of the default run and write its JSON to --leg-out")
    ap.add_argument("--leg-out", default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the decode path runs on a "
              f"single MI355X. Across GPUs it shards by placement (one process per GPU decoding its own "
              f"segments; decode never communicates)", file=sys.stderr)
        return 2

    if args.leg is None and args.workload == SERVED and args.only == "all":
        return run_legs(args)

    import torch

    import sirius_amd._abi as A
    from sirius_amd import lib as L

    lib = L.load()
    tuning = apply_tuning(lib)
    dev = torch.device("cuda:0")
    if args.leg is not None:
        res = (served_leg if args.leg == "served" else extra_leg)(args, lib, tuning, dev)
        Path(args.leg_out).write_text(json.dumps(res))
        return 0
    if args.workload == SERVED or args.only == "served":
        s = served_leg(args, lib, tuning, dev)
        print(json.dumps({"workload": SERVED, **s["served"], "cpu_baseline": s["cpu"]}), flush=True)
        return 0
    work = build_workload(args.workload, args.rows)

    if args.only == "batched":
        r, res = measure(lib, work, dev, args.steps, args.warmup, 0, full=False)
        stream = torch.cuda.current_stream()
        kt = kernel_times(lib, res, stream.cuda_stream, stream, args.steps, args.warmup) if res.n_fixed else {}
        r["kernels_us"] = {A.CODEC_NAMES[c]: v * 1e3 for c, v in kt.items()}
        print(json.dumps(rounded({"workload": args.workload, **r})), flush=True)
        return 0

    log(f"[bench] measuring {args.workload}")
    r, res = measure(lib, work, dev, args.steps, args.warmup, max(3, args.steps // 10))
    del res
    log(f"[bench] {args.workload}: batched {r['ms_per_step']:.3f} ms/step, plain {r['plain']['ms_per_step']:.3f}, "
        f"persistent {r['persistent']['ms_per_step']:.3f}, bit-exact={r['bit_exact']}")

    dom_name, dom = max(r["kernels"].items(), key=lambda kv: kv[1]["us"]) if "kernels" in r else (None, None)
    traffic = None
    tf = ROOT / "profiles" / f"traffic_{args.workload}.json"
    if dom_name and tf.exists():
        traffic = json.loads(tf.read_text()).get(dom_name)

    extra = {}
    if not args.no_extra and args.workload == "int_date_mix_64M":
        for wl in ("bp_int32_8M", "alp_double_32M", "fsst_varchar_16M"):
            log(f"[bench] measuring {wl}")
            w = build_workload(wl)
            er, eres = measure(lib, w, dev, max(5, args.steps // 2), 2, 3)
            del eres
            extra[wl] = er
            log(f"[bench] {wl}: batched {er['ms_per_step']:.3f} ms/step, {er['decoded_gbs']:.0f} GB/s, "
                f"bit-exact={er['bit_exact']}")
            del w

    staged_r = None
    if not args.no_extra:
        staged_r = staged(lib, work, dev, max(5, args.steps // 5))

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_configs({args.workload: work})

    cols = {"int_date_mix_64M": ["RLE INT64 runs(mean 16)", "BITPACKING INT32 1..50", "BITPACKING DATE sorted/2048"],
            "copy_256M": ["UNCOMPRESSED INT64"]}.get(args.workload, [n for n, _, _ in work])
    line = {
        "metric": METRIC,
        "value": round(r["decoded_gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"copy_256M": "int64"}.get(args.workload, "int64+int32"),
        "data": "synthetic (DuckDB-format segments, self-encoded; DuckDB absent)",
        "config": {"workload": args.workload, "rows": r["rows"] // len(work), "columns": cols,
                   "segments": r["segments"],
                   "compressed_bytes": r["compressed_bytes"], "decoded_bytes": r["decoded_bytes"],
                   "path": "sdr_decode_table (batched launches)"},
        "rows_per_s": round(r["rows"] / len(work) / (r["ms_per_step"] * 1e-3), 1),
        "job_achieved_gbs": round(r["achieved_gbs"], 2),
        "job_frac_hbm": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
        "latency_ms": {"batched": {"p50": round(r["p50_ms"], 4), "p99": round(r["p99_ms"], 4)},
                       "persistent": {"p50": round(r["persistent"]["p50_ms"], 4),
                                      "p99": round(r["persistent"]["p99_ms"], 4)},
                       "plain": {"p50": round(r["plain"]["p50_ms"], 4), "p99": round(r["plain"]["p99_ms"], 4)}},
        "persistent": rounded(r["persistent"]),
        "plain": rounded(r["plain"]),
        "batched_vs_plain": round(r["batched_vs_plain"], 3),
        "kernels": rounded(r.get("kernels", {}), 2),
        "bit_exact": r["bit_exact"] and r["plain"]["bit_exact"] and r["persistent"]["bit_exact"],
        "tuning": tuning,
        "workloads": {k: rounded({kk: vv for kk, vv in v.items() if kk != "kernels"}) for k, v in extra.items()},
        "staged_from_file": staged_r,
        "roofline": None if dom is None else {
            "bound": "hbm", "kernel": dom_name, "achieved": round(dom["gbs"], 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
            "alg_bytes": dom["alg_bytes"], "avg_us": round(dom["us"], 2)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def served_leg(args, lib, tuning, dev) -> dict:
    """The served mixed-codec stream (config 5) and its CPU baseline."""
    sr = served(lib, args.served_requests, steps=args.steps, warmup=args.warmup,
                baselines=not args.no_served_baselines, resident_gb=args.served_resident_gb,
                max_batch=args.served_max_batch)
    pool, specs = sr.pop("_pool"), sr.pop("_specs")
    cpu = None if args.no_cpu_baseline else cpu_served(pool, specs)
    return {"served": sr, "cpu": cpu, "tuning": tuning}


def extra_leg(args, lib, tuning, dev) -> dict:
    """Configs 1-4 (batched, persistent and one-launch-per-segment paths, per-kernel durations), the
    PCIe-inclusive staging leg and the configs' CPU baselines."""
    extra, works, staged_r = {}, {}, None
    if not args.no_extra:
        for wl in CONFIGS:
            log(f"[bench] measuring {wl}")
            w = build_workload(wl)
            er, eres = measure(lib, w, dev, max(5, args.steps // 2), max(5, args.warmup), 3)
            del eres
            extra[wl] = er
            works[wl] = w
            log(f"[bench] {wl}: batched {er['ms_per_step']:.3f} ms/step ({er['decoded_gbs']:.0f} GB/s), persistent "
                f"{er['persistent']['ms_per_step']:.3f} (p50 {er['persistent']['p50_ms']:.3f}), plain "
                f"{er['plain']['ms_per_step']:.3f}, bit-exact={er['bit_exact']}")
        staged_r = staged(lib, works["int_date_mix_64M"], dev, max(5, args.steps // 5))
    elif not args.no_cpu_baseline:
        works = {wl: build_workload(wl) for wl in CONFIGS}
    cpu = None if args.no_cpu_baseline else cpu_configs(works)
    return {"workloads": {k: rounded(v) for k, v in extra.items()}, "staged_from_file": staged_r, "cpu": cpu,
            "tuning": tuning}


def run_legs(args) -> int:
    """The default run: the served stream, then configs 1-4 with the staging leg, each in a child
    process of its own (this process never touches the GPU), merged into one JSON line."""
    import subprocess
    import tempfile

    common = ["--steps", str(args.steps), "--warmup", str(args.warmup),
              "--served-requests", str(args.served_requests), "--served-resident-gb", str(args.served_resident_gb),
              "--served-max-batch", str(args.served_max_batch)]
    for flag in ("no_served_baselines", "no_cpu_baseline", "no_extra"):
        if getattr(args, flag):
            common.append("--" + flag.replace("_", "-"))
    legs = {}
    for leg in ("served", "extra"):
        if leg == "extra" and args.no_extra and args.no_cpu_baseline:
            continue
        fd, path = tempfile.mkstemp(prefix=f"sdr_bench_{leg}_", suffix=".json")
        os.close(fd)
        try:
            log(f"[bench] leg {leg}: child process")
            rc = subprocess.run([sys.executable, "-u", str(Path(__file__).resolve()), "--leg", leg,
                                 "--leg-out", path, *common], stdout=sys.stderr).returncode
            legs[leg] = json.loads(Path(path).read_text()) if rc == 0 else {"rc": rc}
        finally:
            os.unlink(path)
        if rc != 0:
            log(f"[bench] leg {leg} exited with status {rc}")
            if leg == "served":
                return 1
    print(json.dumps(headline_line(args, legs["served"], legs.get("extra"))), flush=True)
    return 0


def headline_line(args, s: dict, x: dict | None) -> dict:
    """The JSON line: the served stream is the value, the other legs ride along."""
    sr = s["served"]
    main = sr["persistent"]
    st = main["server_stats"]
    alg = main["decoded_bytes"] + main["compressed_bytes"]
    launch_ms = st["gpu_union_ms"] / max(1, st["launches"])
    x = x or {}
    ik = (x.get("workloads") or {}).get("int_date_mix_64M", {}).get("kernels", {})
    dom = max(ik, key=lambda k: ik[k]["us"]) if ik else None
    per = {**((x.get("cpu") or {}).get("per_config") or {}), **((s.get("cpu") or {}).get("per_config") or {})}
    cpu = None
    if s.get("cpu"):
        cpu = {**s["cpu"], "per_config": per,
               "sample": s["cpu"]["sample"] + ("; configs 1-4: " + x["cpu"].get("sample", "") if x.get("cpu") else "")}
    return {
        "metric": METRIC,
        "value": round(main["decoded_gbs"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": sr["passes"],
        "warmup": args.warmup,
        "ms_per_step": round(main["span_ms"] / sr["passes"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+int64+f64+varchar",
        "data": "synthetic (DuckDB-format segments, self-encoded; DuckDB absent)",
        "config": {"workload": SERVED, "requests": main["requests"] // sr["passes"],
                   "rows": "log-uniform 2Ki..16Mi per request",
                   "mix": "40% BITPACKING INT32, 20% RLE INT64, 25% ALP DOUBLE, 15% FSST VARCHAR",
                   "arrivals": f"Poisson at {sr['rate_req_per_s']} req/s (0.8x measured capacity)",
                   "budget_bytes": 200 * 10 ** 9, "resident_gb": sr["resident_gb"],
                   "path": "sdr_server, persistent batch launches on 2 slot streams (continuous batching)",
                   "max_batch_requests": sr.get("max_batch_requests")},
        "job_achieved_gbs": round(main["achieved_gbs"], 2),
        "job_frac_hbm": sr["frac_hbm"],
        "latency_ms": {"served": {"p50": round(main["p50_ms"], 3), "p99": round(main["p99_ms"], 3)},
                       **{k: {"p50": v["p50_ms"], "p99": v["p99_ms"]} for k, v in sr["baselines"].items()}},
        "bit_exact": main["verified_ok"] == main["verified"] and main["mismatched_requests"] == 0,
        "tuning": s["tuning"],
        "served": sr,
        "workloads": x.get("workloads"),
        "staged_from_file": x.get("staged_from_file"),
        "legs": {"extra": "ok" if x and "rc" not in x else (f"exit status {x['rc']}" if x else "skipped")},
        "roofline": {
            "bound": "hbm", "kernel": "persistent batch launches (per-codec grids and the FSST group kernel), served stream",
            "achieved": round(alg / (st["gpu_union_ms"] * 1e-3) / 1e9, 1) if st["gpu_union_ms"] > 0 else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(alg / (st["gpu_union_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if st["gpu_union_ms"] > 0 else None,
            "traffic": None,
            "alg_bytes_per_launch": round(alg / max(1, st["launches"])), "avg_ms_per_launch": round(launch_ms, 4),
            "note": "algorithmic bytes (compressed read + decoded written) over the time at least one launch ran, "
                    "from each launch's start/end HIP events on its slot stream"},
        "roofline_int_date_mix_64M": None if dom is None else {
            "kernel": dom, **{k: round(v, 2) if isinstance(v, float) else v for k, v in ik[dom].items()}},
        "cpu_baseline": cpu,
    }


def staged(lib, work, dev, steps):
    """PCIe-inclusive rate (reported beside the headline, never as its value): the workload's
    segments written to one file, then per step sdr_stage_file (reads into pinned memory,
    one host-to-device copy) and sdr_decode_table on the staged image."""
    import os
    import tempfile

    import torch

    import sirius_amd._abi as A
    from sirius_amd import lib as L
    from sirius_amd import stage as S

    fixed = [(v, e) for _, v, e in work if e.codec != A.FSST]
    parts, ranges, pos = [], [], 0
    for _, e in fixed:
        for o, bs, ro, rc, msl in e.segs:
            parts.append(e.blob[int(o):int(o) + int(bs)].tobytes())
            ranges.append((pos, int(bs)))
            pos += int(bs)
            parts.append(bytes((-pos) % 8))
            pos += (-pos) % 8
    fd, path = tempfile.mkstemp(prefix="sdr_stage_", dir="/tmp")
    try:
        with os.fdopen(fd, "wb") as fh:
            fh.write(b"".join(parts))
        del parts
        outs = [torch.empty(e.n_rows * e.type_size, dtype=torch.uint8, device=dev) for _, e in fixed]
        stream = torch.cuda.current_stream()
        sh = stream.cuda_stream

        def step(decode=True):
            buf, offs = S.stage_file(path, ranges, dev)
            if not decode:
                return
            cols, keep, k = [], [], 0
            for (_, e), out in zip(fixed, outs):
                segs = []
                for o, bs, ro, rc, msl in e.segs:
                    segs.append(A.Segment(buf.data_ptr() + offs[k], int(bs), int(ro), int(rc), 0))
                    k += 1
                runs, kp = A.make_runs([(e.codec, segs)])
                keep.append(kp)
                c = A.Column()
                c.type_size, c.total_rows, c.data, c.n_data, c.d_out = e.type_size, e.n_rows, runs, 1, out.data_ptr()
                cols.append(c)
            arr = (A.Column * len(cols))(*cols)
            L.check(lib.sdr_decode_table(arr, len(cols), sh))

        def timed(fn):
            for _ in range(2):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / steps

        t_all = timed(step)
        ok = all(np.array_equal(o.cpu().numpy(), v.view(np.uint8)) for o, (v, _) in zip(outs, fixed))
        t_stage = timed(lambda: step(decode=False))
    finally:
        os.unlink(path)
    dec = sum(e.n_rows * e.type_size for _, e in fixed)
    log(f"[bench] staged from a file: {t_all * 1e3:.2f} ms/step ({dec / t_all / 1e9:.0f} GB/s decoded), "
        f"staging alone {t_stage * 1e3:.2f} ms ({pos / t_stage / 1e9:.1f} GB/s of file bytes)")
    return {"file_bytes": pos, "segments": len(ranges), "ms_per_step": round(t_all * 1e3, 3),
            "decoded_gbs": round(dec / t_all / 1e9, 1), "stage_ms": round(t_stage * 1e3, 3),
            "stage_gbs": round(pos / t_stage / 1e9, 1), "bit_exact": bool(ok),
            "note": "file in the page cache; host-clock timing around sdr_stage_file + sdr_decode_table"}


def served(lib, n_requests, steps=0, warmup=0, baselines=True, resident_gb=32.0, max_batch=0):
    """The served mixed-codec stream: requests of log-uniform size
    (2 Ki..16 Mi rows), 40% BITPACKING INT32 / 20% RLE INT64 / 25% ALP DOUBLE
    / 15% FSST VARCHAR over segments resident in the server's pool (copies at
    distinct addresses, about `resident_gb` GB), 200 GB admission budget.

    Capacity is the throughput with a prefix of the stream queued at once.
    The headline serves the whole stream with Poisson arrivals at 0.8x that
    capacity: `warmup` untimed passes, then `steps` timed passes (a step is
    one pass of the stream), every 16th request checked bit-exact on the
    device.  A sweep serves a prefix at 0.5x and 0.95x.  Each baseline (one
    drop-in launch per request, one launch per segment) is served at 0.8x
    of its own capacity, so its p50/p99 are those of a load it sustains."""
    import copy

    import torch

    import sirius_amd._abi as A
    from sirius_amd import lib as L
    from sirius_amd import serve as S

    t = time.time()
    cfg = A.ServerConfig(200 * 10 ** 9, 2, A.LAUNCH_PERSISTENT, max_batch, 0)
    server = C.c_void_p()
    L.check(lib.sdr_server_create(C.byref(cfg), C.byref(server)))

    def capacity(mode, m, srv=None, verify_every=16):
        sp = copy.deepcopy(specs[:m])
        S.poisson_arrivals(sp, 0)
        r = S.run(lib, pool, sp, mode, server=srv, verify_every=verify_every)
        return r, m / (r["span_ms"] * 1e-3)

    try:
        pool = S.Pool(lib, server, max_rows=16 << 20, n_tails=48, log=log, resident_bytes=int(resident_gb * 1e9))
        specs = S.make_specs(pool, n_requests)
        log(f"[bench] served: pool staged ({pool.staged_bytes / 1e9:.2f} GB resident, {pool.replicas} copies) "
            f"in {time.time() - t:.1f}s")
        m = min(len(specs), 1000)
        capacity(A.LAUNCH_PERSISTENT, m, server)  # reserves the pool's output slabs
        cap, cap_rps = capacity(A.LAUNCH_PERSISTENT, m, server)
        rate = 0.8 * cap_rps
        log(f"[bench] served: capacity {cap_rps:.1f} req/s ({cap['decoded_gbs']:.0f} GB/s decoded, "
            f"{cap['achieved_gbs'] / HBM_PEAK_GBS:.3f} of HBM); serving {len(specs)} requests at {rate:.1f} req/s, "
            f"{warmup} warm-up + {steps} timed passes")
        run_specs = copy.deepcopy(specs)
        S.poisson_arrivals(run_specs, rate)
        for _ in range(warmup):
            S.run(lib, pool, run_specs, A.LAUNCH_PERSISTENT, server=server, verify_every=16)
        passes = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(1, steps)):
            passes.append(S.run(lib, pool, run_specs, A.LAUNCH_PERSISTENT, server=server, verify_every=16))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        main = S.merge_runs(passes)
        log(f"[bench] served: {main['decoded_gbs']:.0f} GB/s decoded, {main['achieved_gbs'] / HBM_PEAK_GBS:.3f} "
            f"of HBM, p50 {main['p50_ms']:.2f} ms, p99 {main['p99_ms']:.2f} ms, verified "
            f"{main['verified_ok']}/{main['verified']}; GPU busy {main['server_stats']['gpu_union_ms']:.1f} of "
            f"{main['span_ms']:.1f} ms")
        for f in main["failed"]:
            log(f"[bench] served MISMATCH {f}")
        sweep = {}
        n_sweep = min(len(specs), 3000)
        for fr in (0.5, 0.95):
            sp = copy.deepcopy(specs[:n_sweep])
            S.poisson_arrivals(sp, fr * cap_rps)
            sr = S.run(lib, pool, sp, A.LAUNCH_PERSISTENT, server=server, verify_every=16)
            sweep[f"{fr:.2f}"] = {k: sr[k] for k in ("decoded_gbs", "achieved_gbs", "p50_ms", "p99_ms", "requests")}
            log(f"[bench] served at {fr:.2f}x capacity: {sr['decoded_gbs']:.0f} GB/s, p50 {sr['p50_ms']:.2f} ms, "
                f"p99 {sr['p99_ms']:.2f} ms")
        sweep["0.80"] = {k: main[k] for k in ("decoded_gbs", "achieved_gbs", "p50_ms", "p99_ms", "requests")}
        base = {}
        for name, mode in (("one_launch_per_request", A.LAUNCH_BATCHED),
                           ("one_launch_per_segment", A.LAUNCH_PER_SEGMENT)) if baselines else ():
            mb = min(len(specs), 200)
            bcap, bcap_rps = capacity(mode, mb, None)
            nb = min(len(specs), 1000)
            bs = copy.deepcopy(specs[:nb])
            S.poisson_arrivals(bs, 0.8 * bcap_rps)
            br = S.run(lib, pool, bs, mode, verify_every=16)
            base[name] = {**rounded(br), "capacity_req_per_s": round(bcap_rps, 2),
                          "capacity_gbs": round(bcap["decoded_gbs"], 2), "rate_req_per_s": round(0.8 * bcap_rps, 2)}
            log(f"[bench] served baseline {name}: capacity {bcap['decoded_gbs']:.0f} GB/s; at 0.8x it "
                f"{br['decoded_gbs']:.0f} GB/s, p50 {br['p50_ms']:.1f} ms, p99 {br['p99_ms']:.1f} ms")
        torch.cuda.synchronize()
    finally:
        lib.sdr_server_destroy(server)
    st = main["server_stats"]
    alg = main["decoded_bytes"] + main["compressed_bytes"]
    union_gbs = alg / (st["gpu_union_ms"] * 1e-3) / 1e9 if st["gpu_union_ms"] > 0 else 0.0
    out = {
        "workload": "served mixed-codec stream (40% BP INT32, 20% RLE INT64, 25% ALP DOUBLE, 15% FSST VARCHAR; "
                    "rows log-uniform 2Ki..16Mi; Poisson arrivals at 0.8x capacity; 200 GB budget)",
        "resident_gb": round(pool.staged_bytes / 1e9, 2), "pool_copies": pool.replicas,
        "max_batch_requests": max_batch,
        "capacity_req_per_s": round(cap_rps, 2), "capacity_gbs": round(cap["decoded_gbs"], 2),
        "capacity_frac_hbm": round(cap["achieved_gbs"] / HBM_PEAK_GBS, 4),
        "mean_request_bytes": round(cap["decoded_bytes"] / m), "rate_req_per_s": round(rate, 2),
        "passes": len(passes), "wall_s": round(wall, 3),
        "persistent": rounded(main),
        "frac_hbm": round(main["achieved_gbs"] / HBM_PEAK_GBS, 4),
        "gpu_busy_gbs": round(union_gbs, 1),
        "sweep": rounded(sweep),
        "baselines": base,
    }
    for name, b in base.items():
        out[f"vs_{name}"] = {"capacity_ratio": round(cap["decoded_gbs"] / b["capacity_gbs"], 3),
                             "p99_ms_at_0.8": [round(main["p99_ms"], 3), round(b["p99_ms"], 3)]}
    out["_pool"], out["_specs"] = pool, specs  # host data for the CPU baseline (not printed)
    return out


def host_cores():
    """Logical CPUs, physical cores (from /proc/cpuinfo) and the CPUs this
    process may run on."""
    logical = os.cpu_count() or 1
    phys = set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
            elif cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
        if cur:
            phys.add((cur.get("physical id"), cur.get("core id")))
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = logical
    return logical, len(phys) if phys and None not in next(iter(phys)) else None, avail


CPU_BUDGET_S = 2.0  # seconds of CPU decoding per config and thread count


class _CpuDecoder:
    """The oracle's multithreaded C decoders (a restatement of the
    reference's decoders, labelled 'port': DuckDB is not installed) over
    host copies of the segments, with all the CPUs this process may use (at
    most 16, the host share of one GPU here) or with one thread."""

    def __init__(self):
        from oracle import oracle as O

        self.lib = O.load()
        self.logical, self.physical, self.avail = host_cores()
        self.n_all = max(1, min(self.avail, 16))

    def fixed_col(self, enc, segs, n):
        import sirius_amd._abi as A

        out = np.empty(n * enc.type_size, np.uint8)
        runs, k = A.make_runs([(enc.codec, segs)])
        col = A.Column()
        col.type_size, col.total_rows, col.data, col.n_data, col.d_out = enc.type_size, n, runs, 1, out.ctypes.data
        return col, [k, out], n * enc.type_size

    def str_col(self, enc, segs, n, n_chars):
        import sirius_amd._abi as A

        offs = np.empty(n + 1, np.int32)
        chars = np.empty(max(1, n_chars) + 64, np.uint8)
        runs, k = A.make_runs([(enc.codec, segs)])
        col = A.StringColumn()
        col.total_rows, col.data, col.n_data = n, runs, 1
        col.d_offsets, col.d_chars, col.chars_capacity = offs.ctypes.data, chars.ctypes.data, len(chars)
        return col, [k, offs, chars], 4 * (n + 1) + n_chars

    @staticmethod
    def host_segs(enc, row0=0):
        import sirius_amd._abi as A

        b = enc.blob.ctypes.data
        return [A.Segment(b + int(s["offset"]), int(s["bytes_size"]), row0 + int(s["row_offset"]),
                          int(s["row_count"]), int(s["max_string_length"])) for s in enc.segs]

    def decode(self, col, is_str, threads):
        if is_str:
            rc = self.lib.oracle_decode_strings_column_mt(C.byref(col), threads)
        else:
            rc = self.lib.oracle_decode_table_mt(C.byref(col), 1, threads)
        assert rc == 0, rc

    def summary(self, per: dict, head: str, sample: str) -> dict:
        return {"value": per[head]["gbs_all"], "unit": "GB/s", "cores": self.n_all, "kind": "port",
                "sample": sample, "threads_all": self.n_all, "logical_cpus": self.logical,
                "physical_cores": self.physical, "usable_cpus": self.avail, "l3_bytes_per_core": l3_bytes(),
                "per_config": per}


def l3_bytes() -> int | None:
    """Bytes of the last-level cache one core sees, from sysfs."""
    try:
        txt = Path("/sys/devices/system/cpu/cpu0/cache/index3/size").read_text().strip()
        return int(txt[:-1]) * {"K": 1 << 10, "M": 1 << 20}[txt[-1]] if txt[-1] in "KM" else int(txt)
    except (OSError, ValueError, KeyError):
        return None


CPU_MIN_SET = 1 << 30  # working set of one timed CPU run: well past any host's last-level cache
CPU_RUNS = 5


def cpu_configs(works, runs=CPU_RUNS):
    """CPU baseline per config: copies of the whole config (encoded bytes and
    outputs at distinct addresses, at least CPU_MIN_SET bytes in all, so the
    working set is far larger than the last-level cache) decoded from host
    memory, the median of `runs` runs, on all usable CPUs and on one."""
    import sirius_amd._abi as A

    cd = _CpuDecoder()

    def median_gbs(jobs, threads):
        rates = []
        for _ in range(runs):
            t0 = time.perf_counter()
            nbytes = 0
            for col, is_str, nb in jobs:
                cd.decode(col, is_str, threads)
                nbytes += nb
            rates.append(nbytes / (time.perf_counter() - t0) / 1e9)
        return float(np.median(rates))

    per = {}
    for name, work in works.items():
        one = sum(e.compressed_bytes + (len(v[0]) + 4 * (e.n_rows + 1) if e.codec == A.FSST else e.n_rows * e.type_size)
                  for _, v, e in work)
        copies = max(1, -(-CPU_MIN_SET // one))
        jobs, keep = [], []
        for _ in range(copies):
            for _, values, enc in work:
                blob = enc.blob.copy()  # a copy at its own address
                segs = [A.Segment(blob.ctypes.data + int(s["offset"]), int(s["bytes_size"]), int(s["row_offset"]),
                                  int(s["row_count"]), int(s["max_string_length"])) for s in enc.segs]
                if enc.codec == A.FSST:
                    col, k, nb = cd.str_col(enc, segs, enc.n_rows, len(values[0]))
                    jobs.append((col, True, nb))
                else:
                    col, k, nb = cd.fixed_col(enc, segs, enc.n_rows)
                    jobs.append((col, False, nb))
                keep += [k, blob]
        g_all, g_one = median_gbs(jobs, cd.n_all), median_gbs(jobs, 1)
        per[name] = {"gbs_all": round(g_all, 3), "gbs_1": round(g_one, 3),
                     "decoded_bytes_per_run": int(sum(nb for _, _, nb in jobs)), "copies": copies,
                     "working_set_bytes": int(one * copies)}
        log(f"[bench] cpu {name}: {g_all:.2f} GB/s on {cd.n_all} threads, {g_one:.2f} GB/s on 1 "
            f"({copies} copies, {one * copies / 1e9:.2f} GB working set)")
    return cd.summary(per, next(iter(per)), f"each config decoded whole, as copies at distinct addresses "
                      f"totalling >= {CPU_MIN_SET / (1 << 30):g} GiB (past the last-level cache), median of {runs} runs; "
                      "duckdb not installed, so the repo's multithreaded C restatement of the decoders")


def cpu_served(pool, specs, budget_s=CPU_BUDGET_S):
    """CPU baseline of the served stream: its requests decoded in order
    from host copies of the pool's segments for about `budget_s`, on all
    usable CPUs and on one."""
    import sirius_amd._abi as A

    cd = _CpuDecoder()
    jobs, keep = [], []
    for sp in specs[:400]:
        src = pool.src[sp.kind]
        segs, row = [], 0
        for i in range(sp.src_first, sp.src_first + sp.n_src):
            s_ = src.enc.segs[i]
            segs.append(A.Segment(src.enc.blob.ctypes.data + int(s_["offset"]), int(s_["bytes_size"]), row,
                                  int(s_["row_count"]), int(s_["max_string_length"])))
            row += int(s_["row_count"])
        if sp.tail >= 0:
            tenc = pool.tails[sp.kind][sp.tail].enc
            segs += cd.host_segs(tenc, row0=row)
            row += tenc.n_rows
        if src.is_string:
            cum = src.cum
            r0, r1 = int(cum[sp.src_first]), int(cum[sp.src_first + sp.n_src])
            n_chars = int(src.ref_offs[r1] - src.ref_offs[r0])
            if sp.tail >= 0:
                n_chars += int(pool.tails[sp.kind][sp.tail].ref_offs[-1])
            col, k, _ = cd.str_col(src.enc, segs, row, n_chars)
            jobs.append((col, True, None))
        else:
            col, k, nb = cd.fixed_col(src.enc, segs, row)
            jobs.append((col, False, nb))
        keep.append(k)

    def timed(threads):
        nbytes, t0, i = 0, time.perf_counter(), 0
        while time.perf_counter() - t0 < budget_s and i < len(jobs):
            col, is_str, nb = jobs[i]
            cd.decode(col, is_str, threads)
            nbytes += (4 * (col.total_rows + 1) + col.chars_used) if is_str else nb
            i += 1
        return nbytes / (time.perf_counter() - t0) / 1e9, i

    g_all, n_req = timed(cd.n_all)
    g_one, n_req1 = timed(1)
    log(f"[bench] cpu served stream: {g_all:.2f} GB/s on {cd.n_all} threads ({n_req} requests), "
        f"{g_one:.2f} GB/s on 1 ({n_req1})")
    per = {SERVED: {"gbs_all": round(g_all, 3), "gbs_1": round(g_one, 3), "requests": n_req, "requests_1": n_req1}}
    return cd.summary(per, SERVED, f"served stream: its requests decoded in order for about {budget_s:.0f} s; "
                      "duckdb not installed, so the repo's multithreaded C restatement of the decoders")


if __name__ == "__main__":
    sys.exit(main())
