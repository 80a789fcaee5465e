"""The decode server (sdr_server_*): every served output is bit-exact to the
pool's reference decode, whether a request is coalesced with others into a
persistent launch or decoded alone; admission keeps held bytes within the
HBM budget; the baseline modes serve the same outputs."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
from sirius_amd import lib as L
from sirius_amd import serve as S

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def setup():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = L.load()
    cfg = A.ServerConfig(0, 2, A.LAUNCH_PERSISTENT, 0, 0)
    server = C.c_void_p()
    L.check(lib.sdr_server_create(C.byref(cfg), C.byref(server)))
    pool = S.Pool(lib, server, max_rows=1 << 21, n_tails=8, seed=77)
    yield lib, server, pool
    torch.cuda.synchronize()
    lib.sdr_server_destroy(server)


def test_mixed_and_alone_are_identical(setup):
    lib, server, pool = setup
    specs = S.make_specs(pool, 60, seed=5, max_rows=1 << 21)
    assert {s.kind for s in specs} == set(S.MIX)
    S.poisson_arrivals(specs, 0)  # all at once: coalesced into few launches
    mixed = S.run(lib, pool, specs, A.LAUNCH_PERSISTENT, verify_every=1, server=server)
    assert mixed["errors"] == 0 and mixed["verified"] == 60 and mixed["verified_ok"] == 60
    assert mixed["launches"] < 60
    # one request per launch, one at a time
    alone = S.run(lib, pool, specs[:24], A.LAUNCH_PERSISTENT, max_inflight=1, max_batch_requests=1,
                  verify_every=1)
    assert alone["errors"] == 0 and alone["verified_ok"] == 24 and alone["launches"] == 24


def test_poisson_stream_overlaps_launches(setup):
    lib, server, pool = setup
    specs = S.make_specs(pool, 80, seed=9, max_rows=1 << 21)
    S.poisson_arrivals(specs, 2000.0)
    r = S.run(lib, pool, specs, A.LAUNCH_PERSISTENT, verify_every=1, server=server)
    assert r["errors"] == 0 and r["verified_ok"] == 80
    assert r["p99_ms"] >= r["p50_ms"] > 0


@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT], ids=["per_request", "per_segment"])
def test_baseline_modes(setup, mode):
    lib, server, pool = setup
    specs = S.make_specs(pool, 20, seed=13, max_rows=1 << 20)
    S.poisson_arrivals(specs, 0)
    r = S.run(lib, pool, specs, mode, verify_every=1)
    assert r["errors"] == 0 and r["verified_ok"] == 20 and r["launches"] == 20


def test_admission_holds_budget(setup):
    lib, server, pool = setup
    specs = [s for s in S.make_specs(pool, 200, seed=21, max_rows=1 << 21) if not pool.src[s.kind].is_string][:30]
    need = max(s.rows * pool.src[s.kind].enc.type_size for s in specs)
    budget = 2 * need + 1
    cfg = A.ServerConfig(budget, 4, A.LAUNCH_PERSISTENT, 0, 0)
    sv = C.c_void_p()
    L.check(lib.sdr_server_create(C.byref(cfg), C.byref(sv)))
    try:
        reqs = [s.request(pool) for s in specs]
        for r in reqs:
            L.check(lib.sdr_server_submit(sv, C.byref(r)))
        comps = (A.Completion * 64)()
        n = C.c_uint32()
        got, peak = 0, 0
        stream = torch.cuda.current_stream()
        while got < len(specs):
            L.check(lib.sdr_server_progress(sv, None))
            peak = max(peak, lib.sdr_server_held_bytes(sv))
            L.check(lib.sdr_server_poll(sv, comps, 64, C.byref(n)))
            for k in range(n.value):
                c = comps[k]
                spec = next(s for s in specs if s.tag == c.tag)
                assert c.status == 0 and S.verify(pool, spec, c, stream)
                L.check(lib.sdr_server_release(sv, c.tag, stream.cuda_stream))
                got += 1
        assert peak <= budget
        assert lib.sdr_server_pending(sv) == 0 and lib.sdr_server_held_bytes(sv) == 0
    finally:
        torch.cuda.synchronize()
        lib.sdr_server_destroy(sv)


def _drain(lib, sv, n_expect, stream):
    """Progress and poll until n_expect completions: {tag: Completion copy}."""
    comps = (A.Completion * 64)()
    n = C.c_uint32()
    out = {}
    while len(out) < n_expect:
        L.check(lib.sdr_server_progress(sv, None))
        L.check(lib.sdr_server_poll(sv, comps, 64, C.byref(n)))
        for k in range(n.value):
            c = A.Completion()
            C.pointer(c)[0] = comps[k]
            out[c.tag] = c
    return out


def test_overflow_is_reported_on_its_own_request(setup):
    """One VARCHAR request with too small a chars_capacity, coalesced into one
    persistent launch with good requests: only it reports SDR_E_OVERFLOW and
    the others' outputs are bit-exact."""
    lib, _, pool = setup
    specs = S.make_specs(pool, 40, seed=31, max_rows=1 << 20)
    strs = [s for s in specs if pool.src[s.kind].is_string]
    assert strs
    bad = strs[0]
    cfg = A.ServerConfig(0, 1, A.LAUNCH_PERSISTENT, 0, 0)
    sv = C.c_void_p()
    L.check(lib.sdr_server_create(C.byref(cfg), C.byref(sv)))
    stream = torch.cuda.current_stream()
    try:
        reqs = {s.tag: s.request(pool) for s in specs}
        reqs[bad.tag].chars_capacity = 7
        for s in specs:
            L.check(lib.sdr_server_submit(sv, C.byref(reqs[s.tag])))
        got = _drain(lib, sv, len(specs), stream)
        for s in specs:
            c = got[s.tag]
            if s.tag == bad.tag:
                assert c.status == A.SDR_E_OVERFLOW and c.chars_used > 7
            else:
                assert c.status == 0 and S.verify(pool, s, c), s.tag
            L.check(lib.sdr_server_release(sv, s.tag, stream.cuda_stream))
    finally:
        torch.cuda.synchronize()
        lib.sdr_server_destroy(sv)


def test_overlapping_string_segments_refused_at_submit(setup):
    lib, server, pool = setup
    spec = next(s for s in S.make_specs(pool, 200, seed=41, max_rows=1 << 20)
                if pool.src[s.kind].is_string and s.n_src >= 2)
    r = spec.request(pool)
    s0, s1 = r.data[0].segments[0], r.data[0].segments[1]
    s1.row_offset = s0.row_offset + s0.row_count - 1  # its first row is segment 0's last
    assert lib.sdr_server_submit(server, C.byref(r)) == A.SDR_E_BOUNDS
    assert lib.sdr_server_pending(server) == 0


@pytest.mark.parametrize("mode", [A.LAUNCH_PERSISTENT, A.LAUNCH_BATCHED], ids=["persistent", "per_request"])
def test_zero_row_string_request(setup, mode):
    """A VARCHAR request of no rows completes with offsets[0] == 0."""
    lib, _, pool = setup
    cfg = A.ServerConfig(0, 1, mode, 0, 0)
    sv = C.c_void_p()
    L.check(lib.sdr_server_create(C.byref(cfg), C.byref(sv)))
    stream = torch.cuda.current_stream()
    try:
        # the pool recycles memory: fill a block first so stale bytes would show
        junk = S.make_specs(pool, 50, seed=3, max_rows=1 << 16)
        first = next(s for s in junk if pool.src[s.kind].is_string)
        L.check(lib.sdr_server_submit(sv, C.byref(first.request(pool))))
        c0 = _drain(lib, sv, 1, stream)[first.tag]
        assert c0.status == 0
        L.check(lib.sdr_server_release(sv, first.tag, stream.cuda_stream))
        runs, keep = A.make_runs([(A.FSST, [])])
        r = A.Request()
        r.tag, r.is_string, r.total_rows, r.n_data, r.data, r.chars_capacity = 999_001, 1, 0, 1, runs, 16
        L.check(lib.sdr_server_submit(sv, C.byref(r)))
        c = _drain(lib, sv, 1, stream)[999_001]
        assert c.status == 0 and c.chars_used == 0
        off0 = S._view(c.d_out, 4).view(torch.int32)
        torch.cuda.synchronize()
        assert int(off0.item()) == 0
        L.check(lib.sdr_server_release(sv, 999_001, stream.cuda_stream))
    finally:
        torch.cuda.synchronize()
        lib.sdr_server_destroy(sv)


def test_verify_detects_a_changed_byte(setup):
    """The served-output check itself: after one byte of a completed
    request's output changes, in its source rows or in its tail rows
    (values, or chars and offsets of a VARCHAR request), verify() fails;
    restored, it passes again."""
    lib, _, pool = setup
    pick = {}
    for s in S.make_specs(pool, 400, seed=41, max_rows=1 << 21):
        if s.tail >= 0 and s.n_src > 0:
            pick.setdefault(pool.src[s.kind].is_string, s)
    assert set(pick) == {False, True}
    cfg = A.ServerConfig(0, 1, A.LAUNCH_PERSISTENT, 0, 0)
    sv = C.c_void_p()
    L.check(lib.sdr_server_create(C.byref(cfg), C.byref(sv)))
    stream = torch.cuda.current_stream()
    try:
        reqs = {s.tag: s.request(pool) for s in pick.values()}
        for r in reqs.values():
            L.check(lib.sdr_server_submit(sv, C.byref(r)))
        got = _drain(lib, sv, len(reqs), stream)
        for is_str, s in pick.items():
            c = got[s.tag]
            assert c.status == 0 and S.verify(pool, s, c)
            src = pool.src[s.kind]
            ns = int(src.cum[s.src_first + s.n_src] - src.cum[s.src_first])
            if is_str:
                c_src = int(src.ref_offs[int(src.cum[s.src_first]) + ns] - src.ref_offs[int(src.cum[s.src_first])])
                chars = S._view(c.d_chars, c.chars_used)
                offs = S._view(c.d_out, (s.rows + 1) * 4)
                targets = [(chars, 0), (chars, c_src), (offs, 4 * (ns + 1))]
            else:
                ts = src.enc.type_size
                out = S._view(c.d_out, s.rows * ts)
                targets = [(out, 0), (out, ns * ts + ts - 1)]
            assert S._ZERO_COPY
            for t, i in targets:
                t[i] ^= 1
                assert not S.verify(pool, s, c), (s.kind, i)
                t[i] ^= 1
                assert S.verify(pool, s, c), (s.kind, i)
            L.check(lib.sdr_server_release(sv, s.tag, stream.cuda_stream))
    finally:
        torch.cuda.synchronize()
        lib.sdr_server_destroy(sv)
