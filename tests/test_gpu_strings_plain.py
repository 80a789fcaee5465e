"""GPU decode of DICTIONARY and UNCOMPRESSED string segments, alone and mixed
with FSST segments and uncovered rows in one column, through the C ABI,
bit-exact against the oracle (offsets, chars, null mask)."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from sirius_amd import lib as L
from tests.helpers import (StringBlob, dictionary_string_segment, duckdb_segments, oracle_string_runs,
                           string_column, uncompressed_string_segment)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
MODES = [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT, A.LAUNCH_PERSISTENT]
MODE_IDS = ["batched", "plain", "persistent"]


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _words(n, seed, vocab=300, max_len=40):
    rng = np.random.default_rng(seed)
    words = [bytes(rng.integers(32, 127, rng.integers(0, max_len)).astype(np.uint8)) for _ in range(vocab)]
    return [words[i] for i in rng.integers(0, vocab, n)]


def gpu_decode(blob, table, total, mode, chars_skew=0, valid=None):
    d = torch.zeros(len(blob) + 256, dtype=torch.uint8, device=_dev())
    d[:len(blob)] = torch.from_numpy(blob).to(_dev())
    cap = int(sum(int(t[1]) for t in table)) * 4 + total * 64 + 1024
    offs = torch.full((total + 1,), -7, dtype=torch.int32, device=_dev())
    chars = torch.full((cap + chars_skew + 32,), 0xCC, dtype=torch.uint8, device=_dev())
    validity, mask, keep_v = None, None, None
    if valid is not None:
        vb = np.packbits(valid, bitorder="little")
        keep_v = torch.from_numpy(np.concatenate([vb, np.zeros(16, np.uint8)])).to(_dev())
        validity = [(A.UNCOMPRESSED, [A.Segment(keep_v.data_ptr(), len(vb), 0, total, 0)])]
        mask = torch.zeros((total + 7) // 8, dtype=torch.uint8, device=_dev())
    col, keep = string_column(duckdb_segments(table, d.data_ptr()), total, offs.data_ptr(),
                              chars.data_ptr() + chars_skew, cap, validity,
                              mask.data_ptr() if mask is not None else None)
    print(f"[gpu] strings total={total} segs={len(table)} mode={mode}", flush=True)
    rc = L.load().sdr_decode_strings_column_mode(C.byref(col), torch.cuda.current_stream().cuda_stream, mode)
    torch.cuda.synchronize()
    L.check(rc)
    out_mask = mask.cpu().numpy() if mask is not None else None
    return offs.cpu().numpy(), chars.cpu().numpy()[chars_skew:chars_skew + col.chars_used], col, out_mask


def check(blob, table, total, mode, chars_skew=0, valid=None):
    table = np.array(table, np.int64)
    o_offs, o_chars = oracle_string_runs(blob, table, total)
    g_offs, g_chars, col, mask = gpu_decode(blob, table, total, mode, chars_skew, valid)
    np.testing.assert_array_equal(g_offs, o_offs)
    np.testing.assert_array_equal(g_chars, o_chars)
    if valid is not None:
        assert col.null_count == int((~valid).sum())
        np.testing.assert_array_equal(np.unpackbits(mask, bitorder="little")[:total], valid.astype(np.uint8))
    return o_offs, o_chars


@pytest.mark.parametrize("codec", [A.DICTIONARY, A.UNCOMPRESSED], ids=["dictionary", "uncompressed"])
@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_plain_codec_column(codec, mode):
    sizes = [1, 1023, 1024, 1025, 4096, 20000, 777, 122880, 3]
    strings = _words(sum(sizes), 11)
    if codec == A.DICTIONARY:
        strings = [None if i % 13 == 5 else s for i, s in enumerate(strings)]
    enc = dictionary_string_segment if codec == A.DICTIONARY else uncompressed_string_segment
    b, row, gap = StringBlob(), 0, 0
    for k, n in enumerate(sizes):
        b.add(codec, enc(strings[row - gap:row - gap + n]), row, n)
        row += n
        if k == 3:  # rows no segment covers
            row += 5000
            gap += 5000
    offs, chars = check(b.blob, b.table, row + 17, mode, chars_skew=3)
    assert len(chars) > 0


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_mixed_codecs_with_nulls(mode):
    b, row = StringBlob(), 0
    fs_chars, fs_offs = E.gen_zipf_strings(77, 30000)
    fs = E.encode_fsst(fs_chars, fs_offs)
    for o, bs, ro, rc, msl in fs.segs:
        b.add(A.FSST, fs.blob[int(o):int(o) + int(bs)].tobytes(), row + int(ro), int(rc), max_len=int(msl))
    row += fs.n_rows
    words = _words(50000, 12)
    b.add(A.DICTIONARY, dictionary_string_segment(words[:20000]), row, 20000)
    row += 20000 + 1000
    b.add(A.UNCOMPRESSED, uncompressed_string_segment(words[20000:], negate={5, 6}), row, 30000)
    row += 30000
    valid = np.random.default_rng(3).random(row) > 0.1
    check(b.blob, b.table, row, mode, valid=valid)


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_long_rows_cross_output_pieces(mode):
    rng = np.random.default_rng(5)
    strings = [bytes(rng.integers(0, 256, int(n)).astype(np.uint8)) for n in rng.integers(0, 30000, 300)]
    b = StringBlob()
    b.add(A.UNCOMPRESSED, uncompressed_string_segment(strings), 0, 300)
    b.add(A.DICTIONARY, dictionary_string_segment(strings[:40] * 10), 300, 400)
    check(b.blob, b.table, 700, mode, chars_skew=7)


def test_malformed_plain_segments_match_oracle():
    words = _words(3000, 13)
    b = StringBlob()
    seg = uncompressed_string_segment(words[:1000])
    b.add(A.UNCOMPRESSED, seg, 0, 1000, bytes_size=8 + 4 * 999)  # offsets do not fit: all empty
    d = bytearray(dictionary_string_segment(words[1000:2000]))
    d[12:16] = (5).to_bytes(4, "little")  # selections >= 5 are out of range: empty
    b.add(A.DICTIONARY, bytes(d), 1000, 1000)
    u = bytearray(uncompressed_string_segment(words[2000:]))
    u[8 + 4 * 10:12 + 4 * 10] = (1).to_bytes(4, "little", signed=True)  # row 10 ends below its start
    b.add(A.UNCOMPRESSED, bytes(u), 2000, 1000)
    d2 = bytearray(dictionary_string_segment(words[:50]))
    d2[4:8] = (len(d2) + 100).to_bytes(4, "little")  # dict_end past the segment: header invalid
    b.add(A.DICTIONARY, bytes(d2), 3000, 50)
    for mode in MODES:
        check(b.blob, b.table, 3050, mode)
