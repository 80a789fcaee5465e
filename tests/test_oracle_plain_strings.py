"""The oracle's DICTIONARY and UNCOMPRESSED string decoders on segments built
in DuckDB's layout (tests/helpers.py), against the strings they were built
from, plus the malformed-input rules.  The layout itself is pinned by the
DuckDB-written fixtures (tests/test_duckdb_segments.py)."""
import numpy as np
import pytest

import sirius_amd._abi as A
from tests.helpers import StringBlob, dictionary_string_segment, oracle_string_runs, uncompressed_string_segment


def _words(n, seed, vocab=50):
    rng = np.random.default_rng(seed)
    words = [bytes(rng.integers(97, 123, rng.integers(0, 20)).astype(np.uint8)) for _ in range(vocab)]
    return [words[i] for i in rng.integers(0, vocab, n)]


def _expect(strings):
    offs = np.concatenate([[0], np.cumsum([len(s or b"") for s in strings])]).astype(np.int32)
    return offs, np.frombuffer(b"".join(s or b"" for s in strings), np.uint8)


@pytest.mark.parametrize("codec", [A.UNCOMPRESSED, A.DICTIONARY])
def test_round_trip(codec):
    strings = _words(5000, 1)
    if codec == A.DICTIONARY:
        strings = [None if i % 7 == 3 else s for i, s in enumerate(strings)]
    b = StringBlob()
    enc = uncompressed_string_segment if codec == A.UNCOMPRESSED else dictionary_string_segment
    b.add(codec, enc(strings), 0, len(strings))
    offs, chars = oracle_string_runs(b.blob, np.array(b.table), len(strings))
    eo, ec = _expect(strings)
    np.testing.assert_array_equal(offs, eo)
    np.testing.assert_array_equal(chars, ec)


def test_negated_offsets_read_as_magnitudes():
    strings = _words(100, 2)
    b = StringBlob()
    b.add(A.UNCOMPRESSED, uncompressed_string_segment(strings, negate={3, 50, 99}), 0, 100)
    offs, chars = oracle_string_runs(b.blob, np.array(b.table), 100)
    np.testing.assert_array_equal(chars, _expect(strings)[1])


def test_malformed_segments_decode_empty():
    strings = _words(64, 3)
    seg = uncompressed_string_segment(strings)
    b = StringBlob()
    b.add(A.UNCOMPRESSED, seg, 0, 64, bytes_size=8 + 4 * 63)  # offsets do not fit
    d = bytearray(dictionary_string_segment(strings))
    d[16:20] = (33).to_bytes(4, "little")  # selection width > 32
    b.add(A.DICTIONARY, bytes(d), 64, 64)
    offs, chars = oracle_string_runs(b.blob, np.array(b.table), 128)
    assert (offs == 0).all() and len(chars) == 0


def test_out_of_range_rows_decode_empty():
    strings = [b"abc", b"defg", b"hi"]
    d = bytearray(dictionary_string_segment(strings))
    ibc_pos = 12
    d[ibc_pos:ibc_pos + 4] = (2).to_bytes(4, "little")  # only entry 1 is in range
    u = bytearray(uncompressed_string_segment(strings))
    u[12:16] = (1).to_bytes(4, "little", signed=True)  # row 1 ends before it starts
    b = StringBlob()
    b.add(A.DICTIONARY, bytes(d), 0, 3)
    b.add(A.UNCOMPRESSED, bytes(u), 3, 3)
    offs, chars = oracle_string_runs(b.blob, np.array(b.table), 6)
    # row 1 of the UNCOMPRESSED segment is empty; row 2 then spans from row 1's stored end (1)
    # to its own (9), as |off[i]| - |off[i-1]| gives in the reference
    assert list(np.diff(offs)) == [3, 0, 0, 3, 0, 8]
    assert chars.tobytes() == b"abc" + b"abc" + b"hidefgab"
