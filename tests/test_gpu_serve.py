"""GPU: the serving loop (resident boards, HBM-budget admission, continuous
batching) through the C ABI.

Every request's paths must equal the CPU oracle's alone search for the same
nets (window pass, then full-board retry), whatever else was on the device at
the same time: each field is the unique fixpoint of its own relaxation, so
solving a request alone and in a mix of other requests on other boards gives
the same bytes.
"""
from __future__ import annotations

import numpy as np
import pytest

import mrcpu
import mrfield

pytestmark = pytest.mark.gpu

# (side, layers, nets, seed): boards of the config-2 generator at small sizes
BOARDS = [(64, 2, 24, 11), (96, 4, 16, 12), (128, 3, 24, 13)]
PER_REQ = 8


def oracle_spec(s):
    g = s.grid
    og = mrcpu.Grid(g.w, g.h, g.layers, g.cost, board=g.board)
    oe = mrcpu.Edges(s.edges.x, s.edges.y, s.edges.via)
    return mrcpu.Spec(og, oe, additive=s.additive)


def oracle_paths(s, nets, windows=None, retry=True):
    g = s.grid
    wins = windows if windows is not None else [(0, 0, g.w - 1, g.h - 1)] * len(nets)
    on = [mrcpu.Net(n.src, n.dst) for n in nets]
    paths, costs = mrcpu.alone_batch(oracle_spec(s), on, wins, retry_full=retry, threads=8)
    return [None if p is None else (p, c) for p, c in zip(paths, costs)]


@pytest.fixture(scope="module")
def boards():
    return [mrfield.synth_board(side, L, n, seed) for side, L, n, seed in BOARDS]


@pytest.fixture(scope="module")
def requests(boards):
    """(board index, nets) in a fixed interleaved order across boards."""
    per_board = []
    for bi, s in enumerate(boards):
        nets = s.nets()
        per_board.append([(bi, nets[k:k + PER_REQ]) for k in range(0, len(nets), PER_REQ)])
    out = []
    for k in range(max(len(p) for p in per_board)):
        for p in per_board:
            if k < len(p):
                out.append(p[k])
    return out


@pytest.fixture(scope="module")
def expected(boards, requests):
    return [oracle_paths(boards[bi], nets) for bi, nets in requests]


def make_server(boards, **kw):
    srv = mrfield.Server(0, **kw)
    ids = [srv.add_board(s.grid, s.edges, s.pricing()) for s in boards]
    return srv, ids


def submit_all(srv, ids, requests, **kw):
    rids = [srv.submit(ids[bi], nets, **kw) for bi, nets in requests]
    return [srv.wait(r) for r in rids]


def test_requests_match_oracle(boards, requests, expected):
    srv, ids = make_server(boards, label_bytes=1 << 30)
    try:
        got = submit_all(srv, ids, requests)
        for k, ((paths, st), exp) in enumerate(zip(got, expected)):
            assert paths == exp, k
            assert st["total_ms"] >= st["solve_ms"] >= 0.0
        s = srv.stats()
        assert s["completed"] == len(requests) and s["failed"] == 0
        assert s["inflight_peak"] >= 2  # requests from several boards relaxed together
    finally:
        srv.close()


def test_alone_equals_mixed(boards, requests, expected):
    srv, ids = make_server(boards, label_bytes=1 << 30)
    try:
        alone = []
        for bi, nets in requests:
            alone.append(srv.wait(srv.submit(ids[bi], nets))[0])
        mixed = [p for p, _ in submit_all(srv, ids, requests)]
        assert alone == mixed
        assert alone == expected
    finally:
        srv.close()


def test_admission_by_label_budget(boards, requests, expected):
    # room for about one request's labels at a time
    budget = 4 << 20
    srv, ids = make_server(boards, label_bytes=budget)
    try:
        got = submit_all(srv, ids, requests)
        assert [p for p, _ in got] == expected
        s = srv.stats()
        assert 0 < s["label_bytes_peak"] <= budget
        assert s["label_bytes"] == budget
    finally:
        srv.close()


def test_windows_then_full_retry(boards):
    s = boards[2]
    nets = s.nets()[:16]
    g = s.grid
    # even nets: the endpoints' bounding box; odd nets: a window holding the
    # source column only, so their target is reachable on the full board only
    wins = []
    for k, n in enumerate(nets):
        ax, ay = n.src % g.w, (n.src % (g.w * g.h)) // g.w
        bx, by = n.dst % g.w, (n.dst % (g.w * g.h)) // g.w
        if k % 2 == 0:
            wins.append((min(ax, bx), min(ay, by), max(ax, bx), max(ay, by)))
        else:
            wins.append((ax, ay, ax, ay))
    srv, ids = make_server([s], label_bytes=1 << 28)
    try:
        with_retry = srv.wait(srv.submit(ids[0], nets, windows=wins, retry_full=True))[0]
        one_pass = srv.wait(srv.submit(ids[0], nets, windows=wins, retry_full=False))[0]
    finally:
        srv.close()
    assert with_retry == oracle_paths(s, nets, wins, retry=True)
    assert one_pass == oracle_paths(s, nets, wins, retry=False)
    assert sum(p is None for p in one_pass) > sum(p is None for p in with_retry)


def test_errors_fail_closed(boards):
    s = boards[0]
    # room for one 64x64x2 field (74 KB of labels) but not for eight
    srv, ids = make_server([s], label_bytes=256 << 10)
    try:
        with pytest.raises(mrfield.MrfError):
            srv.submit(ids[0] + 100, s.nets()[:1])
        bad = mrfield.Net(int(np.flatnonzero(s.grid.cost == mrfield.MAX)[0]), s.nets()[0].dst)
        with pytest.raises(mrfield.MrfError) as e:
            srv.submit(ids[0], [bad])
        assert e.value.code == mrfield.INVALID_ENDPOINT
        # a request whose labels exceed the whole budget fails on its own ...
        big = srv.submit(ids[0], s.nets()[:8])
        with pytest.raises(mrfield.MrfError) as e:
            srv.wait(big)
        assert e.value.code == mrfield.BACKEND_UNAVAILABLE
        # ... and the server keeps serving requests that fit
        small = srv.wait(srv.submit(ids[0], s.nets()[:1]))[0]
        assert small == oracle_paths(s, s.nets()[:1])
        srv.remove_board(ids[0])
        with pytest.raises(mrfield.MrfError):
            srv.submit(ids[0], s.nets()[:1])
    finally:
        srv.close()


def test_small_requests_go_ahead_of_a_large_one():
    """Size-aware admission: with the device paced by a large request (pacing
    threshold 1 tile), small requests that arrive after a large queued one
    are admitted ahead of it and finish first; every result still equals the
    oracle's.  With admit_small_tiles = 1 arrival order holds."""
    big = mrfield.synth_board(384, 4, 48, 21)
    small = mrfield.synth_board(64, 2, 32, 22)
    big_nets = big.nets()
    small_reqs = [small.nets()[k:k + 8] for k in range(0, 32, 8)]
    exp_small = [oracle_paths(small, nets) for nets in small_reqs]
    # small requests: 8 full 64x64 fields (at most 128 tiles); the large ones:
    # 24 full 384x384 fields (thousands of tiles)
    for admit_small, expect_ahead in ((1024, True), (1, False)):
        # pacing threshold 1: while anything runs, only small requests enter
        srv = mrfield.Server(0, label_bytes=8 << 30, admit_below_tiles=1, admit_small_tiles=admit_small)
        try:
            ib = srv.add_board(big.grid, big.edges, big.pricing())
            isml = srv.add_board(small.grid, small.edges, small.pricing())
            r0 = srv.submit(ib, big_nets[:24])
            r1 = srv.submit(ib, big_nets[24:])  # queued behind the running request
            rs = [srv.submit(isml, nets) for nets in small_reqs]
            got_small = [srv.wait(r) for r in rs]
            got_big = [srv.wait(r0), srv.wait(r1)]
            st = srv.stats()
        finally:
            srv.close()
        for (paths, _), exp in zip(got_small, exp_small):
            assert paths == exp
        if expect_ahead:
            assert st["admitted_ahead"] >= 1
            # the small requests did not wait for the second large one
            assert max(s["total_ms"] for _, s in got_small) < got_big[1][1]["total_ms"]
        else:
            assert st["admitted_ahead"] == 0
    ref_big = oracle_paths(big, big_nets[:24:4])
    assert [got_big[0][0][i] for i in range(0, 24, 4)] == ref_big


def test_packed_certificate_falls_back_to_wide_labels(boards):
    """Requests run on packed u32 labels where their boards allow.  A board
    whose distances pass the 16-bit packed limit (gap prices of 700) fails
    the certificate: its request runs again on u64 labels, after the packed
    requests beside it drain.  Every result equals the oracle's."""
    w, h, L = 96, 80, 2
    rng = np.random.default_rng(31)
    cost = np.ones(w * h * L, dtype=np.uint32)
    cost[rng.random(w * h * L) < 0.15] = 0xFFFFFFFF
    free = np.flatnonzero(cost != 0xFFFFFFFF)
    pairs = rng.choice(free, size=(8, 2), replace=False)
    nets = [mrfield.Net(int(a), int(b)) for a, b in pairs]
    full = [(0, 0, w - 1, h - 1)] * len(nets)
    og = mrcpu.Grid(w, h, L, cost)
    oe = mrcpu.Edges([700] * (w - 1), [700] * (h - 1), [2100])
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), [mrcpu.Net(n.src, n.dst) for n in nets], full,
                                     retry_full=True)
    assert max(c for c in ref_c if c != 0xFFFFFFFF) > 0xFFFF  # the case is past the packed limit
    high = [None if p is None else (p, c) for p, c in zip(ref_p, ref_c)]
    small = boards[0]
    srv = mrfield.Server(0, label_bytes=1 << 28)
    try:
        hid = srv.add_board(mrfield.Grid(w, h, L, cost), mrfield.Edges([700] * (w - 1), [700] * (h - 1), [2100]))
        sid = srv.add_board(small.grid, small.edges, small.pricing())
        a = small.nets()[:8]
        b = small.nets()[8:16]
        r1 = srv.submit(sid, a)
        r2 = srv.submit(hid, nets)
        r3 = srv.submit(sid, b)
        got1, got2, got3 = srv.wait(r1)[0], srv.wait(r2)[0], srv.wait(r3)[0]
        assert got2 == high
        assert got1 == oracle_paths(small, a)
        assert got3 == oracle_paths(small, b)
        assert srv.stats()["failed"] == 0
    finally:
        srv.close()
