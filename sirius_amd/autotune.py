"""Launch-shape autotuner for the decode kernels.

Encodes synthetic columns of the benchmark shapes (DuckDB-format segments,
self-encoded), uploads them once, and for every tuning knob times the
kernels the knob affects under each candidate value (HIP events around
repeated launches of a prepared plan, one codec at a time).  The fastest
value of each knob is kept before the next knob is tried, and the result is
written as a table that ``sdr_tune_load`` applies (``name value`` lines,
timings as comments).  Knobs change launch shapes only, never results.

    python -m sirius_amd.autotune [--out PATH] [--rows N] [--reps R]
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time
from pathlib import Path

from . import _abi as A
from . import encode as E
from . import lib as L

DEFAULT_TABLE = Path(__file__).resolve().parent / "tuning" / "gfx950.txt"

# knob -> candidate values; the codecs whose kernels it shapes
CANDIDATES = {
    A.TUNE_RLE_SLICE_ROWS: ([4096, 8192, 12288, 16384, 24576, 32768, 61440], [A.RLE]),
    A.TUNE_BP_CHUNK_ROWS: ([1024, 2048], [A.BITPACKING]),
    A.TUNE_BP_GROUPS_PER_WAVE: ([1, 2, 4], [A.BITPACKING]),
    A.TUNE_ALP_VECTORS_PER_WAVE: ([1, 2, 4], [A.ALP, A.ALPRD]),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


COLUMN_NAMES = ["rle_int64_runs", "bp_int32_quantity", "bp_date_sorted", "bp_int32_w17", "alp_double",
                "alprd_double"]


def columns(rows: int):
    """(name, encoded) pairs of the benchmark column shapes."""
    return [
        ("rle_int64_runs", E.encode_rle(E.gen_runs_int64(0xC0DE0002, rows, 16.0, 1 << 40))),
        ("bp_int32_quantity", E.encode_bitpacking(E.gen_uniform_int32(0xC0DE0012, rows, 1, 50))),
        ("bp_date_sorted", E.encode_bitpacking(E.gen_dates_sorted(0xC0DE0022, rows))),
        ("bp_int32_w17", E.encode_bitpacking(E.gen_bp_int32(0xC0DE0001, rows // 8))),
        ("alp_double", E.encode_alp(E.gen_alp_doubles(0xC0DE0003, rows // 2))),
        ("alprd_double", E.encode_alprd(E.gen_unit_doubles(0xC0DE0013, rows // 8))),
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
    keep, cols, per_codec = [], [], {}
    for name, enc in columns(args.rows):
        blob = torch.from_numpy(enc.blob).to(dev)
        out = torch.empty(enc.n_rows * enc.type_size, dtype=torch.uint8, device=dev)
        runs, k = A.make_runs([(enc.codec, enc.segments(blob.data_ptr()))])
        col = A.Column()
        col.type_size, col.total_rows, col.data, col.n_data, col.d_out = (
            enc.type_size, enc.n_rows, runs, 1, out.data_ptr())
        keep += [blob, out, k]
        cols.append(col)
        per_codec.setdefault(enc.codec, []).append(len(cols) - 1)
        log(f"[autotune] {name}: {enc.n_rows} rows, {len(enc.segs)} segments")

    def time_codec(codec: int) -> float:
        """ms per launch of `codec`'s kernels over every column using it."""
        total = 0.0
        for ci in per_codec.get(codec, []):
            arr = (A.Column * 1)(cols[ci])
            plan = C.c_void_p()
            L.check(lib.sdr_plan_create(arr, 1, A.LAUNCH_BATCHED, sh, C.byref(plan)))
            try:
                for _ in range(3):
                    L.check(lib.sdr_plan_launch(plan, sh))
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(args.reps):
                    L.check(lib.sdr_plan_launch(plan, sh))
                e1.record(stream)
                torch.cuda.synchronize()
                total += e0.elapsed_time(e1) / args.reps
            finally:
                lib.sdr_plan_destroy(plan)
        return total

    lines = [f"# decode launch shapes for gfx950, measured {time.strftime('%Y-%m-%d')} by sirius_amd.autotune",
             f"# columns: {', '.join(COLUMN_NAMES)}; rows {args.rows}"]
    for knob, (values, codecs) in CANDIDATES.items():
        name = A.TUNE_NAMES[knob]
        results = []
        for v in values:
            L.check(lib.sdr_tune_set(knob, v))
            ms = sum(time_codec(c) for c in codecs)
            results.append((ms, v))
            log(f"[autotune] {name}={v}: {ms * 1e3:.1f} us")
        best_ms, best = min(results)
        L.check(lib.sdr_tune_set(knob, best))
        lines.append(f"# {name}: " + ", ".join(f"{v}={ms * 1e3:.1f}us" for ms, v in results))
        lines.append(f"{name} {best}")
    args.out.parent.mkdir(parents=True, exist_ok=True)
    args.out.write_text("\n".join(lines) + "\n")
    log(f"[autotune] wrote {args.out}")
    print(args.out.read_text(), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
