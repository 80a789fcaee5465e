"""Launch-shape autotuner for the decode kernels.

Encodes synthetic columns of the benchmark shapes at every value width the
decoders specialise on (DuckDB-format segments, self-encoded), uploads them
once, and for every tuning knob and width times the kernels the knob shapes
under each candidate value: the per-codec kernels (HIP events around
repeated launches of a prepared plan) for the launch-shape knobs, and the
persistent batch with its fixed-width segments in the group kernels
(replayed launches of one batch, SDR_BATCH_FIXED_KERNELS = 1) for the
units-per-entry knobs, which shape only those kernels.  The fastest value of each (knob, width) is kept before the next one
is tried, and the result is written as a table that ``sdr_tune_load``
applies: ``name@width value`` lines, timings as comments.  Knobs change
launch shapes only, never results.

    python -m sirius_amd.autotune [--out PATH] [--rows N] [--reps R]
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time
from pathlib import Path

import numpy as np

from . import _abi as A
from . import encode as E
from . import lib as L

DEFAULT_TABLE = Path(__file__).resolve().parent / "tuning" / "gfx950.txt"

# knob -> (candidate values, the codecs whose kernels it shapes, timed path)
CANDIDATES = {
    A.TUNE_RLE_SLICE_ROWS: ([4096, 8192, 12288, 16384, 24576, 32768, 61440], [A.RLE], "plan"),
    A.TUNE_BP_CHUNK_ROWS: ([1024, 2048], [A.BITPACKING], "plan"),
    A.TUNE_BP_GROUPS_PER_WAVE: ([1, 2, 4], [A.BITPACKING], "plan"),
    A.TUNE_ALP_VECTORS_PER_WAVE: ([1, 2, 4], [A.ALP, A.ALPRD], "plan"),
    A.TUNE_BATCH_UNITS_BITPACKING: ([4, 8, 16, 32], [A.BITPACKING], "persistent"),
    A.TUNE_BATCH_UNITS_RLE: ([1, 2, 4], [A.RLE], "persistent"),
    A.TUNE_BATCH_UNITS_ALP: ([4, 8, 16, 32], [A.ALP], "persistent"),
    A.TUNE_BATCH_UNITS_ALPRD: ([4, 8, 16, 32], [A.ALPRD], "persistent"),
    A.TUNE_BATCH_UNITS_COPY: ([1, 2, 4, 8], [A.UNCOMPRESSED], "persistent"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def columns(rows: int):
    """(name, encoded) pairs: the benchmark column shapes (configs 1-3) and
    the same distributions at the other widths each decoder specialises on."""
    q = max(1, rows // 4)
    return [
        ("rle_int64_runs", E.encode_rle(E.gen_runs_int64(0xC0DE0002, rows, 16.0, 1 << 40))),
        ("rle_int32_runs", E.encode_rle(E.gen_runs_int64(0xC0DE0032, q, 16.0, 1 << 30).astype(np.int32))),
        ("bp_int32_quantity", E.encode_bitpacking(E.gen_uniform_int32(0xC0DE0012, rows, 1, 50))),
        ("bp_date_sorted", E.encode_bitpacking(E.gen_dates_sorted(0xC0DE0022, rows))),
        ("bp_int32_w17", E.encode_bitpacking(E.gen_bp_int32(0xC0DE0001, rows // 8))),
        ("bp_int16", E.encode_bitpacking(E.gen_uniform_int32(0xC0DE0042, q, 0, 1000).astype(np.int16))),
        ("bp_int64", E.encode_bitpacking((E.gen_u64(0xC0DE0052, q) % (1 << 24)).astype(np.int64))),
        ("alp_double", E.encode_alp(E.gen_alp_doubles(0xC0DE0003, rows // 2))),
        ("alp_float", E.encode_alp(E.gen_alp_doubles(0xC0DE0063, q).astype(np.float32))),
        ("alprd_double", E.encode_alprd(E.gen_unit_doubles(0xC0DE0013, rows // 8))),
        ("alprd_float", E.encode_alprd(E.gen_unit_doubles(0xC0DE0073, q).astype(np.float32))),
        ("uncompressed_int64", E.encode_uncompressed(E.gen_u64(0xC0DE00FF, q).view(np.int64))),
        ("uncompressed_int32", E.encode_uncompressed(E.gen_u64(0xC0DE00FE, q).astype(np.uint32).view(np.int32))),
    ]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", type=Path, default=DEFAULT_TABLE)
    ap.add_argument("--rows", type=int, default=64 << 20)
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)

    import torch

    lib = L.load()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    keep, cols, groups = [], [], {}  # groups: (codec, width) -> column indices
    names = []
    for name, enc in columns(args.rows):
        blob = torch.from_numpy(enc.blob).to(dev)
        out = torch.empty(enc.n_rows * enc.type_size, dtype=torch.uint8, device=dev)
        runs, k = A.make_runs([(enc.codec, enc.segments(blob.data_ptr()))])
        col = A.Column()
        col.type_size, col.total_rows, col.data, col.n_data, col.d_out = (
            enc.type_size, enc.n_rows, runs, 1, out.data_ptr())
        keep += [blob, out, k]
        cols.append(col)
        names.append(name)
        groups.setdefault((enc.codec, enc.type_size), []).append(len(cols) - 1)
        log(f"[autotune] {name}: {enc.n_rows} rows, {len(enc.segs)} segments, {enc.type_size}-byte values")

    def timed(launch) -> float:
        for _ in range(3):
            launch()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.reps):
            launch()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / args.reps

    def time_plan(idx) -> float:
        """ms per launch of the per-codec kernels of the columns `idx`."""
        total = 0.0
        for ci in idx:
            arr = (A.Column * 1)(cols[ci])
            plan = C.c_void_p()
            L.check(lib.sdr_plan_create(arr, 1, A.LAUNCH_BATCHED, sh, C.byref(plan)))
            try:
                total += timed(lambda: L.check(lib.sdr_plan_launch(plan, sh)))
            finally:
                lib.sdr_plan_destroy(plan)
        return total

    def time_persistent(idx) -> float:
        """ms per persistent batch launch over the columns `idx`."""
        b = C.c_void_p()
        L.check(lib.sdr_batch_create(C.byref(b)))
        L.check(lib.sdr_batch_set_option(b, A.BATCH_FIXED_KERNELS, 1))
        arrs = [(A.Column * 1)(cols[ci]) for ci in idx]

        def launch():
            L.check(lib.sdr_batch_clear(b))
            for arr in arrs:
                L.check(lib.sdr_batch_add_table(b, arr, 1))
            L.check(lib.sdr_batch_launch(b, sh))
        try:
            ms = timed(launch)
            L.check(lib.sdr_batch_finish(b, sh))
            return ms
        finally:
            lib.sdr_batch_destroy(b)

    lines = [f"# decode launch shapes for gfx950, measured {time.strftime('%Y-%m-%d')} by sirius_amd.autotune",
             f"# columns: {', '.join(names)}; rows {args.rows} (other widths {max(1, args.rows // 4)})",
             "# name@width value: the value used for columns of that many bytes per value"]
    for knob, (values, codecs, path) in CANDIDATES.items():
        name = A.TUNE_NAMES[knob]
        for w in A.TUNE_WIDTHS:
            idx = [ci for c in codecs for ci in groups.get((c, w), [])]
            if not idx:
                continue
            results = []
            for v in values:
                L.check(lib.sdr_tune_set_width(knob, w, v))
                ms = time_plan(idx) if path == "plan" else time_persistent(idx)
                results.append((ms, v))
                log(f"[autotune] {name}@{w}={v}: {ms * 1e3:.1f} us")
            best_ms, best = min(results)
            L.check(lib.sdr_tune_set_width(knob, w, best))
            lines.append(f"# {name}@{w} ({path}): " + ", ".join(f"{v}={ms * 1e3:.1f}us" for ms, v in results))
            lines.append(f"{name}@{w} {best}")
    args.out.parent.mkdir(parents=True, exist_ok=True)
    args.out.write_text("\n".join(lines) + "\n")
    log(f"[autotune] wrote {args.out}")
    print(args.out.read_text(), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
