"""GPU parity at the benchmark configurations' own sizes.

The small-board cases of test_gpu_parity.py exercise every feature; these
run the headline and the other configurations at full scale and compare with
the CPU oracle, the way the reference's GPU == CPU stress tests run at their
own sizes (crates/mr-metal/src/lib.rs:1652-1701, :1882-1933):

* config 2 (512x512x4, history plane, 12-gon outline masks, 256 nets): the
  whole batch solved on the device; the paths and search costs of every net
  against the oracle's alone search, and 4 complete dist / hops fields
  against the oracle's complete lexicographic fields;
* config 4 (1024x1024x6, 512 multi-pin nets, history plus the +40 halo on
  cells next to another net's pin): 4 complete multi-source fields against
  the oracle, every cell;
* config 5 (serving): requests of 16 to 64 nets on 512- and 1024-side
  8-layer boards of the config-2 generator, submitted together so they relax
  in the same rounds, sampled nets against the oracle;
* config 3 (2048x2048x2, 1024 windowed nets): every net against the
  oracle's window pass with full-board retry.

Integer work: the bar is bit-exact.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import mrcpu
import mrfield
from cases import MAX, oracle_parts

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


# The workgroup form (four wavefronts per tile visit) and the wave form (one
# wavefront per tile visit up to four layers, two up to eight: unit-weight
# narrow batches without pads, overrides or legal planes).
FORMS = {"workgroup": (16, 16, 256, 32), "wave": (16, 16, 64, 32)}


@pytest.fixture(scope="module", params=list(FORMS))
def eng(request):
    e = mrfield.Engine(0)
    e.set_tiles(*FORMS[request.param])
    e.set_mode(mrfield.MODE_PLAIN)
    e.form = request.param
    yield e
    e.close()


def assert_form_ran(eng, wave_label_bytes=None):
    st = eng.stats()
    if eng.form == "wave":
        assert st["wave_launches"] == st["relax_launches"] > 0, st
        if wave_label_bytes is not None:
            assert st["label_bytes"] == wave_label_bytes, st
    else:
        assert st["wave_launches"] == 0 and st["label_bytes"] == 8, st


def spec_of(s, **kw):
    og, oe = oracle_parts(s)
    via = [v for v in s.edges.via if v is not None]
    return mrcpu.Spec(og, oe, additive=s.additive, halo_owner=s.halo_owner, halo_penalty=s.halo_penalty,
                      use_heuristic=True, min_via_cost=min(via) if via else 0, **kw)


def assert_paths(got, ref_p, ref_c, idx):
    for k, i in enumerate(idx):
        if ref_p[k] is None:
            assert got.path(i) is None, i
        else:
            gp = got.path(i)
            assert gp is not None, i
            assert gp[0] == ref_p[k], i
            assert gp[1] == ref_c[k], i


def test_config2_full_board(eng):
    s = mrfield.synth_config(2, seed=1)
    g = s.grid
    assert (g.w, g.h, g.layers, len(s.src)) == (512, 512, 4, 256)
    board = mrfield.Board(eng, g, s.edges, s.pricing())
    try:
        got = board.solve_arrays(s.nets())
        assert_form_ran(eng, 4)  # packed labels: every C2 distance is far below the 16-bit limit
        assert int(np.count_nonzero(got.routed)) > 200
        idx = list(range(256))
        full = [(0, 0, g.w - 1, g.h - 1)] * len(idx)
        nets = [mrcpu.Net(int(s.src[i]), int(s.dst[i])) for i in idx]
        ref_p, ref_c = mrcpu.alone_batch(spec_of(s), nets, full, retry_full=True, threads=THREADS)
        assert_paths(got, ref_p, ref_c, idx)
        # complete fields of four nets, every cell
        fi = [3, 77, 141, 250]
        board.solve_fields([[int(s.src[i])] for i in fi], dsts=[int(s.dst[i]) for i in fi])
        fnets = [mrcpu.Net(int(s.src[i]), int(s.dst[i])) for i in fi]
        og, oe = oracle_parts(s)
        ref_d, ref_h = mrcpu.fields_batch(mrcpu.Spec(og, oe, additive=s.additive), fnets,
                                          [[n.src] for n in fnets], threads=THREADS)
        for k in range(len(fi)):
            d, h = board.read_field(k)
            assert np.array_equal(d, ref_d[k]), (fi[k], np.flatnonzero(d != ref_d[k])[:8])
            assert np.array_equal(h, ref_h[k]), fi[k]
            assert (d != MAX).sum() > g.n // 4  # the field covers the board inside its outline
    finally:
        board.close()


def test_config4_multisource_fields_full_board(eng):
    """Six layers with a halo-owner plane: the wave form runs two wavefronts
    per tile on u64 labels."""
    s = mrfield.synth_config(4, seed=1)
    g = s.grid
    assert (g.w, g.h, g.layers, len(s.pins)) == (1024, 1024, 6, 512)
    ids = list(range(len(s.pins)))
    board = mrfield.Board(eng, g, s.edges, s.pricing(ids))
    try:
        fi = [0, 101, 333, 511]
        board.solve_fields([s.pins[i] for i in fi], net_ids=fi)
        assert_form_ran(eng, 8)
        og, oe = oracle_parts(s)
        spec = mrcpu.Spec(og, oe, additive=s.additive, halo_owner=s.halo_owner, halo_penalty=s.halo_penalty)
        nets = [mrcpu.Net(0, MAX, own_id=i) for i in fi]
        ref_d, ref_h = mrcpu.fields_batch(spec, nets, [s.pins[i] for i in fi], threads=THREADS)
        for k, i in enumerate(fi):
            d, h = board.read_field(k)
            assert np.array_equal(d, ref_d[k]), (i, np.flatnonzero(d != ref_d[k])[:8])
            assert np.array_equal(h, ref_h[k]), i
            assert (d == 0).sum() == len(set(s.pins[i]))  # every pin is a source
        # the halo penalty is live: a foreign pin's neighbour costs more to enter
        assert np.count_nonzero((s.halo_owner >= 0) & (s.halo_owner != fi[0])) > 1000
    finally:
        board.close()


def test_config3_windowed_batch(eng):
    s = mrfield.synth_config(3, seed=1)
    g = s.grid
    assert (g.w, g.h, g.layers, len(s.src)) == (2048, 2048, 2, 1024)
    board = mrfield.Board(eng, g, s.edges, s.pricing())
    try:
        got = board.solve_arrays(s.nets(), windows=s.windows, retry_full=True)
        assert_form_ran(eng)
    finally:
        board.close()
    spec = spec_of(s)
    for lo in range(0, 1024, 256):  # bounded path capacity per oracle batch
        idx = list(range(lo, lo + 256))
        nets = [mrcpu.Net(int(s.src[i]), int(s.dst[i])) for i in idx]
        ref_p, ref_c = mrcpu.alone_batch(spec, nets, [s.windows[i] for i in idx], retry_full=True,
                                         threads=THREADS)
        assert_paths(got, ref_p, ref_c, idx)


# (side, layers, nets, seed): 8-layer boards of the config-2 generator
SERVE_BOARDS = [(512, 8, 64, 31), (1024, 8, 48, 32)]
SAMPLE = 4


def test_config5_serving_large_boards():
    boards = [mrfield.synth_board(side, L, n, seed) for side, L, n, seed in SERVE_BOARDS]
    srv = mrfield.Server(0, label_bytes=24 << 30, max_layers=8)
    try:
        ids = [srv.add_board(b.grid, b.edges, b.pricing()) for b in boards]
        # requests of 16, 48 (512 side) and 16, 32 (1024 side) nets, interleaved
        reqs = [(0, 0, 16), (1, 0, 16), (0, 16, 64), (1, 16, 48)]
        rids = [srv.submit(ids[bi], boards[bi].nets()[a:b]) for bi, a, b in reqs]
        got = [srv.wait(r, arrays=True)[0] for r in rids]
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["inflight_peak"] >= 2
    finally:
        srv.close()
    for (bi, a, b), res in zip(reqs, got):
        s = boards[bi]
        g = s.grid
        idx = list(np.linspace(0, b - a - 1, SAMPLE).astype(int))
        nets = [mrcpu.Net(int(s.src[a + k]), int(s.dst[a + k])) for k in idx]
        full = [(0, 0, g.w - 1, g.h - 1)] * len(idx)
        ref_p, ref_c = mrcpu.alone_batch(spec_of(s), nets, full, retry_full=True, threads=THREADS)
        assert_paths(res, ref_p, ref_c, idx)
