#!/usr/bin/env python3
"""Decode throughput of DuckDB storage segments on one MI355X.

Default workload (the configuration the headline metric is quoted on):
"int_date_mix_64M" — 64 Mi rows of three columns, DuckDB-format segments
(self-encoded; DuckDB is not installed): an INT64 column of runs (geometric
length, mean 16, values in [0, 2^40)) encoded RLE, an INT32 quantity column
(1..50) and a DATE column (7 years, sorted in 2048-row groups), both
BITPACKING.  Inputs are resident in HBM before timing starts.

A step is one decode of the three columns through sdr_decode_table (the
drop-in entry, batched launches).  Prints one JSON line.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


WORKLOADS = ("int_date_mix_64M", "copy_256M")


def build_workload(name: str, rows: int):
    import sirius_amd.encode as E

    t = time.time()
    if name == "copy_256M":
        # calibration: UNCOMPRESSED INT64 (a plain copy of known size)
        cols = [("uncompressed_int64", E.gen_u64(0xC0DE00FF, rows).view(np.int64), E.encode_uncompressed)]
    else:
        cols = [
            ("rle_int64_runs", E.gen_runs_int64(0xC0DE0002, rows, 16.0, 1 << 40), E.encode_rle),
            ("bp_int32_quantity", E.gen_uniform_int32(0xC0DE0012, rows, 1, 50), E.encode_bitpacking),
            ("bp_date_sorted", E.gen_dates_sorted(0xC0DE0022, rows), E.encode_bitpacking),
        ]
    out = []
    for name, values, enc_f in cols:
        enc = enc_f(values)
        out.append((name, values, enc))
        log(f"[bench] encoded {name}: {len(enc.segs)} segments, {enc.compressed_bytes / 1e6:.1f} MB "
            f"compressed, {values.nbytes / 1e6:.1f} MB decoded")
    log(f"[bench] workload built in {time.time() - t:.1f}s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=None)
    ap.add_argument("--workload", choices=WORKLOADS, default="int_date_mix_64M")
    ap.add_argument("--only", choices=("all", "batched"), default="all",
                    help="batched: time only the drop-in path (for profiler runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.rows is None:
        args.rows = (32 << 20) if args.workload == "copy_256M" else (64 << 20)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this build decodes on a "
              f"single MI355X; multi-GPU sharding is described in DESIGN.md)", file=sys.stderr)
        return 2

    import torch

    import sirius_amd._abi as A
    from sirius_amd import lib as L

    lib = L.load()
    tuning = apply_tuning(lib)
    dev = torch.device("cuda:0")
    work = build_workload(args.workload, args.rows)
    rows = args.rows

    keep, cols, outs = [], [], []
    comp_bytes = dec_bytes = 0
    per_codec_bytes = {}
    for name, values, enc in work:
        blob = torch.from_numpy(enc.blob).to(dev)
        out = torch.empty(rows * enc.type_size, dtype=torch.uint8, device=dev)
        runs, k = A.make_runs([(enc.codec, enc.segments(blob.data_ptr()))])
        col = A.Column()
        col.type_size, col.total_rows, col.data, col.n_data, col.d_out = enc.type_size, rows, runs, 1, out.data_ptr()
        keep += [blob, k]
        cols.append(col)
        outs.append(out)
        comp_bytes += enc.compressed_bytes
        dec_bytes += rows * enc.type_size
        key = (enc.codec, enc.type_size)
        cb, db = per_codec_bytes.get(key, (0, 0))
        per_codec_bytes[key] = (cb + enc.compressed_bytes, db + rows * enc.type_size)
    arr = (A.Column * len(cols))(*cols)
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(steps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
        return ev0.elapsed_time(ev1) / steps, wall

    def step_batched():
        L.check(lib.sdr_decode_table_mode(arr, len(cols), sh, A.LAUNCH_BATCHED))

    def step_plain():
        L.check(lib.sdr_decode_table_mode(arr, len(cols), sh, A.LAUNCH_PER_SEGMENT))

    log("[bench] timing batched path")
    ms_batched, wall_batched = timed(step_batched, args.steps, args.warmup)
    # correctness of what was timed: bit-exact against the generated values
    torch.cuda.synchronize()
    exact = all(np.array_equal(out.cpu().numpy(), values.view(np.uint8)) for (_, values, _), out in zip(work, outs))
    log(f"[bench] batched {ms_batched:.3f} ms/step, bit-exact={exact}")
    if args.only == "batched":
        print(json.dumps({"workload": args.workload, "ms_per_step": ms_batched, "bit_exact": bool(exact),
                          "decoded_bytes": dec_bytes, "compressed_bytes": comp_bytes}), flush=True)
        return 0
    log("[bench] timing plain (one launch per segment) path")
    ms_plain, wall_plain = timed(step_plain, max(5, args.steps // 5), 2)
    log(f"[bench] plain {ms_plain:.3f} ms/step")

    # prepared plan: descriptors resident, launch-only steps; per-kernel timing
    plan = C.c_void_p()
    L.check(lib.sdr_plan_create(arr, len(cols), A.LAUNCH_BATCHED, sh, C.byref(plan)))
    ms_plan, _ = timed(lambda: L.check(lib.sdr_plan_launch(plan, sh)), args.steps, args.warmup)
    kern = {}
    for (codec, ts), (cb, db) in per_codec_bytes.items():
        ms_k, _ = timed(lambda c=codec: L.check(lib.sdr_plan_launch_codec(plan, c, sh)), args.steps, args.warmup)
        kern[A.CODEC_NAMES[codec]] = dict(ms=ms_k, bytes=cb + db, gbs=(cb + db) / (ms_k * 1e-3) / 1e9)
    lib.sdr_plan_destroy(plan)
    dom = max(kern.items(), key=lambda kv: kv[1]["ms"])
    log(f"[bench] plan {ms_plan:.3f} ms/step; kernels: " +
        ", ".join(f"{k} {v['ms']*1e3:.1f}us {v['gbs']:.0f}GB/s" for k, v in kern.items()))

    value = dec_bytes / (ms_batched * 1e-3) / 1e9
    achieved_job = (comp_bytes + dec_bytes) / (ms_batched * 1e-3) / 1e9

    traffic = None
    tf = ROOT / "profiles" / "traffic_int_date_mix_64M.json"
    if tf.exists():
        try:
            traffic = json.loads(tf.read_text()).get(dom[0])
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(work, rows)

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batched, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+int32",
        "data": "synthetic (DuckDB-format segments, self-encoded; DuckDB absent)",
        "config": {"workload": args.workload, "rows": rows,
                   "columns": ["RLE INT64 runs(mean 16)", "BITPACKING INT32 1..50", "BITPACKING DATE sorted/2048"],
                   "segments": int(sum(len(e.segs) for _, _, e in work)),
                   "compressed_bytes": comp_bytes, "decoded_bytes": dec_bytes},
        "rows_per_s": round(rows / (ms_batched * 1e-3), 1),
        "job_achieved_gbs": round(achieved_job, 2),
        "job_frac_hbm": round(achieved_job / HBM_PEAK_GBS, 4),
        "plain_ms_per_step": round(ms_plain, 4),
        "plain_gbs": round(dec_bytes / (ms_plain * 1e-3) / 1e9, 2),
        "batched_vs_plain": round(ms_plain / ms_batched, 3),
        "plan_ms_per_step": round(ms_plan, 4),
        "plan_gbs": round(dec_bytes / (ms_plan * 1e-3) / 1e9, 2),
        "wall_ms_per_step": round(wall_batched, 4),
        "kernels": {k: {"us": round(v["ms"] * 1e3, 2), "alg_bytes": v["bytes"], "gbs": round(v["gbs"], 1)}
                    for k, v in kern.items()},
        "bit_exact": bool(exact),
        "tuning": tuning,
        "roofline": {"bound": "hbm", "kernel": dom[0], "achieved": round(dom[1]["gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(dom[1]["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def apply_tuning(lib):
    """Apply the autotuner's table (sirius_amd/tuning/gfx950.txt) if present;
    returns the knob values in effect."""
    import sirius_amd._abi as A
    from sirius_amd import lib as L
    from sirius_amd.autotune import DEFAULT_TABLE

    if DEFAULT_TABLE.exists():
        L.check(lib.sdr_tune_load(str(DEFAULT_TABLE).encode()))
        log(f"[bench] tuning table {DEFAULT_TABLE.name} applied")
    return {name: int(lib.sdr_tune_get(k)) for k, name in A.TUNE_NAMES.items()}


def cpu_baseline(work, rows):
    """The oracle (C restatement of the decoders, 'port') on the host cores,
    on a bounded sample of the same workload: the first segments of each
    column up to ~8M rows, repeated until ~10 s of CPU time."""
    import sirius_amd._abi as A
    from oracle import oracle as O

    lib = O.load()
    threads = os.cpu_count() or 1
    sample_rows = min(rows, 8 << 20)
    cols, keep, nbytes = [], [], 0
    for name, values, enc in work:
        sel = enc.segs[enc.segs["row_offset"] < sample_rows]
        segs = []
        for s in sel:
            segs.append(A.Segment(enc.blob.ctypes.data + int(s["offset"]), int(s["bytes_size"]),
                                  int(s["row_offset"]), int(s["row_count"]), 0))
        n = int(max(int(s["row_offset"]) + int(s["row_count"]) for s in sel))
        out = np.empty(n * enc.type_size, np.uint8)
        runs, k = A.make_runs([(enc.codec, segs)])
        col = A.Column()
        col.type_size, col.total_rows, col.data, col.n_data, col.d_out = enc.type_size, n, runs, 1, out.ctypes.data
        keep += [k, out]
        cols.append(col)
        nbytes += n * enc.type_size
    # all columns must share total_rows for the table call: decode per column
    reps, elapsed = 0, 0.0
    t0 = time.perf_counter()
    while elapsed < 10.0 and reps < 2000:
        for col in cols:
            rc = lib.oracle_decode_table_mt(C.byref(col), 1, threads)
            assert rc == 0
        reps += 1
        elapsed = time.perf_counter() - t0
    gbs = nbytes * reps / elapsed / 1e9
    duck = None
    try:
        import duckdb  # noqa: F401

        duck = "duckdb importable on this host (not timed in this build)"
    except Exception:
        duck = "duckdb not installed: baseline is the repo's C restatement (not DuckDB)"
    log(f"[bench] cpu baseline {gbs:.2f} GB/s on {threads} threads ({reps} reps, {elapsed:.1f}s)")
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": threads, "kind": "port",
            "sample": f"first {sample_rows} rows of each of the 3 columns, decoded {reps}x; {duck}"}


if __name__ == "__main__":
    sys.exit(main())
