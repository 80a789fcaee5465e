"""The persistent mixed-codec path (sdr_batch, SDR_LAUNCH_PERSISTENT):
identical outputs to the oracle for mixed fixed-width and FSST columns,
with and without hipGraph replay, with the fixed-width segments in per-codec
grids (the default) and in persistent group kernels at several entry
granularities, and on repeated replays of one instantiated graph with new
queues."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from oracle import oracle as O
from sirius_amd import lib as L
from tests.helpers import fixed_column, string_column

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def sh():
    return torch.cuda.current_stream().cuda_stream


class Workload:
    """A mixed set of encoded columns resident on the device."""

    def __init__(self, n, seed):
        rng = np.random.default_rng(seed)
        self.n = n
        self.fixed = [
            E.encode_bitpacking(E.gen_bp_int32(seed + 1, n)),
            E.encode_bitpacking(E.gen_dates_sorted(seed + 2, n)),
            E.encode_rle(E.gen_runs_int64(seed + 3, n)),
            E.encode_alp(E.gen_alp_doubles(seed + 4, n)),
            E.encode_alprd(E.gen_unit_doubles(seed + 5, n)),
            E.encode_bitpacking(rng.integers(0, 1 << 12, n).astype(np.int16)),
            E.encode_rle(E.gen_runs_int64(seed + 6, n, 3.0, 5).astype(np.uint8)),
            E.encode_uncompressed(rng.integers(0, 1 << 60, n, dtype=np.int64)),
        ]
        self.str_chars, self.str_offs = E.gen_zipf_strings(seed + 7, n)
        self.fsst = E.encode_fsst(self.str_chars, self.str_offs)
        self.keep = []
        self.cols, self.outs = [], []
        for enc in self.fixed:
            blob = torch.from_numpy(enc.blob).to(_dev())
            out = torch.full((n * enc.type_size,), 0xCC, dtype=torch.uint8, device=_dev())
            col, k = fixed_column([(enc.codec, enc.segments(blob.data_ptr()))], enc.type_size, n, out.data_ptr())
            self.keep += [blob, k]
            self.cols.append(col)
            self.outs.append(out)
        self.arr = (A.Column * len(self.cols))(*self.cols)
        blob = torch.from_numpy(self.fsst.blob).to(_dev())
        self.d_offs = torch.full((n + 1,), -1, dtype=torch.int32, device=_dev())
        self.d_chars = torch.full((len(self.str_chars) + 64,), 0xCC, dtype=torch.uint8, device=_dev())
        self.scol, k = string_column([(A.FSST, self.fsst.segments(blob.data_ptr()))], n, self.d_offs.data_ptr(),
                                     self.d_chars.data_ptr(), len(self.str_chars) + 64)
        self.keep += [blob, k]

    def wipe(self):
        for o in self.outs:
            o.fill_(0xCC)
        self.d_offs.fill_(-1)
        self.d_chars.fill_(0xCC)

    def check(self):
        for enc, out in zip(self.fixed, self.outs):
            np.testing.assert_array_equal(out.cpu().numpy(), O.decode_encoded(enc, threads=8),
                                          err_msg=A.CODEC_NAMES[enc.codec])
        assert self.scol.chars_used == len(self.str_chars)
        np.testing.assert_array_equal(self.d_offs.cpu().numpy().astype(np.uint64), self.str_offs)
        np.testing.assert_array_equal(self.d_chars.cpu().numpy()[: len(self.str_chars)], self.str_chars)


@pytest.fixture(scope="module")
def work():
    return Workload(300_000, 100)


@pytest.mark.parametrize("graph", [1, 0], ids=["graph", "nograph"])
@pytest.mark.parametrize("fixed,units", [(0, None), (1, 1), (1, 4), (1, 16)],
                         ids=["grids", "groups-units1", "groups-units4", "groups-units16"])
def test_batch_mixed_codecs(work, graph, fixed, units):
    lib = L.load()
    b = C.c_void_p()
    L.check(lib.sdr_batch_create(C.byref(b)))
    try:
        L.check(lib.sdr_batch_set_option(b, A.BATCH_USE_GRAPH, graph))
        L.check(lib.sdr_batch_set_option(b, A.BATCH_FIXED_KERNELS, fixed))
        for codec, key in A.BATCH_UNITS.items() if units else ():
            L.check(lib.sdr_batch_set_option(b, key, units))
        work.wipe()
        L.check(lib.sdr_batch_add_table(b, work.arr, len(work.cols)))
        L.check(lib.sdr_batch_add_strings(b, C.byref(work.scol)))
        assert lib.sdr_batch_num_entries(b) > 0
        L.check(lib.sdr_batch_launch(b, sh()))
        L.check(lib.sdr_batch_finish(b, sh()))
        work.check()
    finally:
        lib.sdr_batch_destroy(b)


@pytest.mark.parametrize("graph", [1, 0], ids=["graph", "nograph"])
@pytest.mark.parametrize("fixed", [0, 1], ids=["grids", "groups"])
def test_graph_replays_with_new_queues(work, fixed, graph):
    """Graph replays with different queue contents (a new recording where
    the launch shapes change), eager work and synchronisation between
    replays; every replay checked.  Without the graph the same sequence of
    direct launches."""
    lib = L.load()
    other = Workload(100_000, 200)
    b = C.c_void_p()
    L.check(lib.sdr_batch_create(C.byref(b)))
    try:
        L.check(lib.sdr_batch_set_option(b, A.BATCH_USE_GRAPH, graph))
        L.check(lib.sdr_batch_set_option(b, A.BATCH_FIXED_KERNELS, fixed))
        for w in (work, other, work, work):
            w.wipe()
            torch.cuda.synchronize()
            L.check(lib.sdr_batch_clear(b))
            L.check(lib.sdr_batch_add_table(b, w.arr, len(w.cols)))
            L.check(lib.sdr_batch_add_strings(b, C.byref(w.scol)))
            L.check(lib.sdr_batch_launch(b, sh()))
            L.check(lib.sdr_batch_finish(b, sh()))
            _ = torch.ones(1024, device=_dev()).sum().item()  # eager work between replays
            w.check()
    finally:
        lib.sdr_batch_destroy(b)


def test_persistent_launch_mode_entries(work):
    lib = L.load()
    work.wipe()
    L.check(lib.sdr_decode_table_mode(work.arr, len(work.cols), sh(), A.LAUNCH_PERSISTENT))
    L.check(lib.sdr_decode_strings_column_mode(C.byref(work.scol), sh(), A.LAUNCH_PERSISTENT))
    torch.cuda.synchronize()
    work.check()


def test_fixed_kernel_option_is_checked():
    lib = L.load()
    b = C.c_void_p()
    L.check(lib.sdr_batch_create(C.byref(b)))
    try:
        assert lib.sdr_batch_set_option(b, A.BATCH_FIXED_KERNELS, 2) == A.SDR_E_INVALID_ARG
    finally:
        lib.sdr_batch_destroy(b)


def test_batch_validity_null_counts():
    n = 50_000
    enc = E.encode_bitpacking(E.gen_bp_int32(9, n))
    valid = np.random.default_rng(1).random(n) > 0.3
    mb = np.packbits(valid, bitorder="little")
    blob = torch.from_numpy(enc.blob).to(_dev())
    mblob = torch.from_numpy(mb).to(_dev())
    out = torch.zeros(n * 4, dtype=torch.uint8, device=_dev())
    mask = torch.zeros(len(mb), dtype=torch.uint8, device=_dev())
    col, keep = fixed_column([(enc.codec, enc.segments(blob.data_ptr()))], 4, n, out.data_ptr(),
                             validity=[(A.UNCOMPRESSED, [A.Segment(mblob.data_ptr(), len(mb), 0, n, 0)])],
                             d_mask=mask.data_ptr())
    arr = (A.Column * 1)(col)
    L.check(L.load().sdr_decode_table_mode(arr, 1, sh(), A.LAUNCH_PERSISTENT))
    assert arr[0].null_count == int((~valid).sum())
    np.testing.assert_array_equal(out.cpu().numpy(), O.decode_encoded(enc))
    np.testing.assert_array_equal(mask.cpu().numpy(), mb)


@pytest.mark.parametrize("graph", [1, 0], ids=["graph", "nograph"])
def test_fixed_only_descriptor_forms_alternate(graph):
    """A batch of fixed-width segments alone passes its descriptors like the
    batched table call: at most 256 per per-codec grid in the kernel
    arguments, more through a descriptor slot (neither uploads the arena or
    replays a graph); a batch that adds a string column goes through the
    arena.  Launches of one batch object alternate between the three forms,
    each checked."""
    lib = L.load()
    rows = 2048

    def table(n_segs, seed):
        n = n_segs * rows
        encs = [E.encode_bitpacking(E.gen_bp_int32(seed, n), rg_rows=rows),
                E.encode_rle(E.gen_runs_int64(seed + 1, n), rg_rows=rows)]
        cols, outs, keep = [], [], []
        for enc in encs:
            blob = torch.from_numpy(enc.blob).to(_dev())
            out = torch.full((enc.n_rows * enc.type_size,), 0xCC, dtype=torch.uint8, device=_dev())
            segs = enc.segments(blob.data_ptr())
            col, k = fixed_column([(enc.codec, segs)], enc.type_size, enc.n_rows, out.data_ptr())
            keep += [blob, k]
            cols.append(col)
            outs.append(out)
        return encs, outs, (A.Column * len(cols))(*cols), keep, len(encs[0].segments(0))

    small = table(256, 31)
    big = table(257, 41)
    assert small[4] <= 256 < big[4]
    strings = Workload(20_000, 51)
    b = C.c_void_p()
    L.check(lib.sdr_batch_create(C.byref(b)))
    try:
        L.check(lib.sdr_batch_set_option(b, A.BATCH_USE_GRAPH, graph))
        for form in ("inline", "slot", "arena", "inline", "slot", "slot", "arena", "inline"):
            encs, outs, arr, _, _ = big if form == "slot" else small
            for o in outs:
                o.fill_(0xCC)
            strings.wipe()
            torch.cuda.synchronize()
            L.check(lib.sdr_batch_clear(b))
            L.check(lib.sdr_batch_add_table(b, arr, len(encs)))
            if form == "arena":
                L.check(lib.sdr_batch_add_strings(b, C.byref(strings.scol)))
            L.check(lib.sdr_batch_launch(b, sh()))
            L.check(lib.sdr_batch_finish(b, sh()))
            for enc, out in zip(encs, outs):
                np.testing.assert_array_equal(out.cpu().numpy(), O.decode_encoded(enc),
                                              err_msg=f"{form} {A.CODEC_NAMES[enc.codec]}")
            if form == "arena":
                assert strings.scol.chars_used == len(strings.str_chars)
                np.testing.assert_array_equal(strings.d_chars.cpu().numpy()[: len(strings.str_chars)],
                                              strings.str_chars)
    finally:
        lib.sdr_batch_destroy(b)
