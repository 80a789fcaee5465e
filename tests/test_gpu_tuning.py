"""What ships is what is tested: every candidate value the autotuner may
choose for every knob (set at every value width), and the shipped table
sirius_amd/tuning/gfx950.txt, decode bit-exact.  Under each, columns of
every codec and value width the decoders specialise on go through the
per-codec kernels (sdr_decode_table) and through one persistent batch, with
its fixed-width segments in per-codec grids and in the group kernels that
the units-per-entry knobs shape."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
from oracle import oracle as O
from sirius_amd import autotune as T
from sirius_amd import lib as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = [(knob, v) for knob, (values, _, _) in T.CANDIDATES.items() for v in values] + [("table", 0)]
IDS = [f"{A.TUNE_NAMES[k]}={v}" if k != "table" else "gfx950.txt" for k, v in CASES]


@pytest.fixture(scope="module")
def cols():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    dev = torch.device("cuda:0")
    out = []
    for name, enc in T.columns(300_000):
        blob = torch.from_numpy(enc.blob).to(dev)
        out.append((name, enc, blob, O.decode_encoded(enc, threads=8)))
    return out


def _column(enc, blob, dev):
    d_out = torch.full((enc.n_rows * enc.type_size,), 0xAB, dtype=torch.uint8, device=dev)
    runs, keep = A.make_runs([(enc.codec, enc.segments(blob.data_ptr()))])
    col = A.Column()
    col.type_size, col.total_rows, col.data, col.n_data, col.d_out = enc.type_size, enc.n_rows, runs, 1, d_out.data_ptr()
    return col, d_out, keep


@pytest.mark.parametrize("knob,value", CASES, ids=IDS)
def test_decode_under_tuned_shape(cols, knob, value):
    lib = L.load()
    dev = torch.device("cuda:0")
    sh = torch.cuda.current_stream().cuda_stream
    saved = {(k, w): lib.sdr_tune_get_width(k, w) for k in A.TUNE_NAMES for w in A.TUNE_WIDTHS}
    try:
        if knob == "table":
            assert T.DEFAULT_TABLE.exists()
            L.check(lib.sdr_tune_load(str(T.DEFAULT_TABLE).encode()))
        else:
            L.check(lib.sdr_tune_set(knob, value))
        # per-codec kernels, one call per column
        for name, enc, blob, want in cols:
            col, d_out, keep = _column(enc, blob, dev)
            arr = (A.Column * 1)(col)
            L.check(lib.sdr_decode_table(arr, 1, sh))
            got = d_out.cpu().numpy()
            assert np.array_equal(got, want), f"{name}: batched path differs"
        # one persistent batch over every column, in both fixed-width forms
        for fixed in (0, 1):
            made = [_column(enc, blob, dev) for _, enc, blob, _ in cols]
            b = C.c_void_p()
            L.check(lib.sdr_batch_create(C.byref(b)))
            try:
                L.check(lib.sdr_batch_set_option(b, A.BATCH_FIXED_KERNELS, fixed))
                arrs = [(A.Column * 1)(col) for col, _, _ in made]
                for arr in arrs:
                    L.check(lib.sdr_batch_add_table(b, arr, 1))
                L.check(lib.sdr_batch_launch(b, sh))
                L.check(lib.sdr_batch_finish(b, sh))
            finally:
                lib.sdr_batch_destroy(b)
            for (name, _, _, want), (_, d_out, _) in zip(cols, made):
                assert np.array_equal(d_out.cpu().numpy(), want), f"{name}: persistent path (fixed={fixed}) differs"
    finally:
        for (k, w), v in saved.items():
            L.check(lib.sdr_tune_set_width(k, w, v))
