"""Concurrent callers: eight host threads, each on its own stream, decode
fixed-width tables, FSST and DICT_FSST string columns and file-staged
segments at the same time through the shared descriptor, staging and
pinned-buffer slots; every result is bit-exact to the single-threaded
expectation."""
import ctypes as C
import os
import tempfile
import threading

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from oracle import oracle as O
from sirius_amd import lib as L
from sirius_amd import stage as S
from tests.helpers import string_column

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

THREADS = 8
ROUNDS = 6


@pytest.fixture(scope="module")
def work():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    dev = torch.device("cuda:0")
    fixed = []
    for values, enc_f in ((E.gen_bp_int32(0xC0C0, 200_000), E.encode_bitpacking),
                          (E.gen_runs_int64(0xC0C1, 200_000, 16.0, 1 << 40), E.encode_rle),
                          (E.gen_alp_doubles(0xC0C2, 100_000), E.encode_alp)):
        enc = enc_f(values)
        d = torch.from_numpy(enc.blob).to(dev)
        fixed.append((enc, d, np.ascontiguousarray(values).view(np.uint8)))
    strs = []
    chars, offs = E.gen_zipf_strings(0xC0C3, 50_000)
    enc = E.encode_fsst(chars, offs)
    strs.append((enc, torch.from_numpy(enc.blob).to(dev), O.decode_strings_masked(enc)))
    chars, offs = E.gen_zipf_strings(0xC0C4, 20_000)
    valid = (np.arange(20_000) % 13 != 1).astype(np.uint8)
    enc = E.encode_dict_fsst(chars, offs, 1, valid=valid, seg_rows=4096)
    strs.append((enc, torch.from_numpy(enc.blob).to(dev), O.decode_strings_masked(enc)))
    fd, path = tempfile.mkstemp(prefix="sdr_conc_", dir="/tmp")
    enc0 = fixed[0][0]
    ranges = []
    with os.fdopen(fd, "wb") as fh:
        pos = 0
        for s in enc0.segs:
            b = enc0.blob[int(s["offset"]):int(s["offset"]) + int(s["bytes_size"])].tobytes()
            fh.write(b + bytes((-len(b)) % 8))
            ranges.append((pos, len(b)))
            pos += len(b) + (-len(b)) % 8
    yield dev, fixed, strs, path, ranges
    os.unlink(path)


def _fixed(lib, enc, base, sh, dev):
    out = torch.empty(enc.n_rows * enc.type_size, dtype=torch.uint8, device=dev)
    runs, keep = A.make_runs([(enc.codec, enc.segments(base))])
    col = A.Column()
    col.type_size, col.total_rows, col.data, col.n_data, col.d_out = enc.type_size, enc.n_rows, runs, 1, out.data_ptr()
    arr = (A.Column * 1)(col)
    L.check(lib.sdr_decode_table(arr, 1, sh))
    return out


def _strings(lib, enc, base, sh, dev, cap):
    n = enc.n_rows
    offs = torch.empty(n + 1, dtype=torch.int32, device=dev)
    chars = torch.empty(cap, dtype=torch.uint8, device=dev)
    mask = torch.full(((n + 7) // 8,), 0xFF, dtype=torch.uint8, device=dev)
    col, keep = string_column([(enc.codec, enc.segments(base))], n, offs.data_ptr(), chars.data_ptr(), cap)
    col.d_null_mask = mask.data_ptr()
    L.check(lib.sdr_decode_strings_column(C.byref(col), sh))
    return offs, chars, mask, col


def test_threads_on_their_own_streams(work):
    dev, fixed, strs, path, ranges = work
    lib = L.load()
    errors = []
    barrier = threading.Barrier(THREADS)

    def worker(t):
        try:
            st = torch.cuda.Stream(device=dev)
            sh = st.cuda_stream
            with torch.cuda.stream(st):
                barrier.wait()
                for k in range(ROUNDS):
                    which = (t + k) % 4
                    if which == 0:
                        res = [(_fixed(lib, enc, d.data_ptr(), sh, dev), want) for enc, d, want in fixed]
                        st.synchronize()
                        for out, want in res:
                            if not np.array_equal(out.cpu().numpy(), want):
                                errors.append(f"thread {t} round {k}: fixed mismatch")
                    elif which in (1, 2):
                        enc, d, (o_offs, o_chars, o_valid, o_nulls) = strs[which - 1]
                        offs, chars, mask, col = _strings(lib, enc, d.data_ptr(), sh, dev, len(o_chars) + 64)
                        st.synchronize()
                        valid = np.unpackbits(mask.cpu().numpy(), bitorder="little")[:enc.n_rows].astype(bool)
                        if not (np.array_equal(offs.cpu().numpy(), o_offs) and col.chars_used == len(o_chars)
                                and np.array_equal(chars.cpu().numpy()[:col.chars_used], o_chars)
                                and np.array_equal(valid, o_valid) and col.null_count == o_nulls):
                            errors.append(f"thread {t} round {k}: strings {which} mismatch")
                    else:
                        enc, _, want = fixed[0]
                        buf, offs = S.stage_file(path, ranges, dev, stream=st)
                        segs = [A.Segment(buf.data_ptr() + o, int(s["bytes_size"]), int(s["row_offset"]),
                                          int(s["row_count"]), 0) for o, s in zip(offs, enc.segs)]
                        out = torch.empty(enc.n_rows * 4, dtype=torch.uint8, device=dev)
                        runs, keep = A.make_runs([(enc.codec, segs)])
                        col = A.Column()
                        col.type_size, col.total_rows, col.data, col.n_data = 4, enc.n_rows, runs, 1
                        col.d_out = out.data_ptr()
                        arr = (A.Column * 1)(col)
                        L.check(lib.sdr_decode_table(arr, 1, sh))
                        st.synchronize()
                        if not np.array_equal(out.cpu().numpy(), want):
                            errors.append(f"thread {t} round {k}: staged mismatch")
        except Exception as ex:  # reported below with the thread that raised it
            errors.append(f"thread {t}: {type(ex).__name__}: {ex}")

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(THREADS)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "a decode thread did not finish"
    torch.cuda.synchronize()
    assert not errors, errors[:10]


def test_batches_in_threads_share_descriptor_slots(work):
    """Eight threads, each with its own sdr_batch on its own stream, launch in
    turn a table of one 257-segment grid (descriptors through the shared
    descriptor slots), a two-segment table (descriptors in the kernel
    arguments) and a table with an FSST column (the batch's arena); every
    output is checked."""
    dev, fixed, strs, path, ranges = work
    lib = L.load()
    rows = 2048
    big = E.encode_bitpacking(E.gen_bp_int32(0xC0C5, 257 * rows), rg_rows=rows)
    assert len(big.segs) == 257
    d_big = torch.from_numpy(big.blob).to(dev)
    want_big = O.decode_encoded(big)
    small, d_small, want_small = fixed[0]
    senc, d_str, (o_offs, o_chars, _, _) = strs[0]
    errors = []
    barrier = threading.Barrier(THREADS)

    def table(enc, d, out):
        runs, keep = A.make_runs([(enc.codec, enc.segments(d.data_ptr()))])
        col = A.Column()
        col.type_size, col.total_rows, col.data, col.n_data = enc.type_size, enc.n_rows, runs, 1
        col.d_out = out.data_ptr()
        return (A.Column * 1)(col), keep

    def worker(t):
        b = C.c_void_p()
        try:
            L.check(lib.sdr_batch_create(C.byref(b)))
            st = torch.cuda.Stream(device=dev)
            sh = st.cuda_stream
            with torch.cuda.stream(st):
                barrier.wait()
                for k in range(ROUNDS):
                    form = (t + k) % 3  # 0: slot, 1: kernel arguments, 2: with strings (arena)
                    enc, d, want = (small, d_small, want_small) if form == 1 else (big, d_big, want_big)
                    out = torch.full((enc.n_rows * enc.type_size,), 0xCC, dtype=torch.uint8, device=dev)
                    arr, keep = table(enc, d, out)
                    L.check(lib.sdr_batch_clear(b))
                    L.check(lib.sdr_batch_add_table(b, arr, 1))
                    if form == 2:
                        offs = torch.full((senc.n_rows + 1,), -1, dtype=torch.int32, device=dev)
                        chars = torch.full((len(o_chars) + 64,), 0xCC, dtype=torch.uint8, device=dev)
                        scol, skeep = string_column([(senc.codec, senc.segments(d_str.data_ptr()))], senc.n_rows,
                                                    offs.data_ptr(), chars.data_ptr(), len(o_chars) + 64)
                        L.check(lib.sdr_batch_add_strings(b, C.byref(scol)))
                    L.check(lib.sdr_batch_launch(b, sh))
                    L.check(lib.sdr_batch_finish(b, sh))
                    st.synchronize()
                    if not np.array_equal(out.cpu().numpy(), want):
                        errors.append(f"thread {t} round {k}: form {form} fixed mismatch")
                    if form == 2 and not (scol.chars_used == len(o_chars)
                                          and np.array_equal(offs.cpu().numpy(), o_offs)
                                          and np.array_equal(chars.cpu().numpy()[:len(o_chars)], o_chars)):
                        errors.append(f"thread {t} round {k}: strings mismatch")
        except Exception as ex:  # reported below with the thread that raised it
            errors.append(f"thread {t}: {type(ex).__name__}: {ex}")
        finally:
            if b.value:
                lib.sdr_batch_destroy(b)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(THREADS)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "a batch thread did not finish"
    torch.cuda.synchronize()
    assert not errors, errors[:10]
