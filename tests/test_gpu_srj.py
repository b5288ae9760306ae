"""GPU: real boards end to end.  Each committed SimpleRouteJson fixture
(tests/golden/srj/: the reference's CLI payloads and corpus boards) goes
through the library's producer (mrf_srj_rasterize, the CLI's rule selection,
crates/mr-cli/src/lib.rs:751-834), then every two-point net through the
provider seam (mrf_route_isolated_batch: window pass, full-board retry) with
planar prices from the line geometry (edge_cost over line gaps,
negotiated.rs:1197-1201, :1268-1272) and a via price of 160 per layer step.
Paths and search costs must equal the oracle's alone search on the same
grid, bit for bit, in the per-round and the work-queue forms.
"""
from __future__ import annotations

import glob
import os

import pytest

import mrcpu
import mrfield
import srjref

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "srj", "*.json")))
VIA = 160


# the workgroup form, and the wave form (own pads and via-exempt pads applied
# per tile)
@pytest.fixture(scope="module", params=[256, 64], ids=["workgroup", "wave"])
def eng(request):
    e = mrfield.Engine(0)
    e.set_tiles(16, 16, request.param, 32)
    e.form = "wave" if request.param == 64 else "workgroup"
    yield e
    e.close()


def board(path):
    p = mrfield.srj_rasterize(srjref.load(path), mode=mrfield.SRJ_ROUTE)
    og = mrcpu.Grid(p.w, p.h, p.layers, p.cost, p.via_forbidden, p.board)
    ex = mrcpu.axis_edge_costs(list(p.x_lines), p.w)
    ey = mrcpu.axis_edge_costs(list(p.y_lines), p.h)
    return p, og, mrcpu.Edges(ex, ey, [VIA] * (p.layers - 1))


@pytest.mark.parametrize("mode", [mrfield.MODE_PLAIN, mrfield.MODE_QUEUE], ids=["plain", "queue"])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def test_srj_board_routes_match_oracle(eng, path, mode):
    p, og, oe = board(path)
    assert p.nets, path
    nets = [mrcpu.Net(n.src, n.dst, list(n.pads), list(n.via_pads)) for n in p.nets]
    wins = [mrcpu.net_window(og, n.src, n.dst, n.pads) for n in nets]
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), nets, wins, retry_full=True, threads=8)
    eng.set_mode(mode)
    try:
        got = eng.route_isolated_batch(p.grid(), [n.net() for n in p.nets], wins,
                                       mrfield.Edges(list(oe.x), list(oe.y), list(oe.via)))
        st = eng.stats()
    finally:
        eng.set_mode(mrfield.MODE_PLAIN)
    if eng.form == "wave" and mode == mrfield.MODE_PLAIN:
        assert st["wave_launches"] == st["relax_launches"] > 0, st  # unit weights: the wave form runs
    routed = 0
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), (os.path.basename(path), p.nets[i].name)
        if r is not None:
            assert r == (ref_p[i], ref_c[i]), (os.path.basename(path), p.nets[i].name)
            routed += 1
    assert routed > 0, path


MULTI = [f for f in FIXTURES if srjref.load(f)["layerCount"] > 1]


@pytest.mark.parametrize("path", MULTI, ids=[os.path.basename(f) for f in MULTI])
def test_srj_exact_model_via_masks_match_oracle(eng, path):
    """The exact clearance model on the multi-layer boards: via keepouts
    (Grid::via_forbidden) and own-pad via exemptions reach the device."""
    doc = srjref.load(path)
    b = doc["bounds"]
    res = max(b["maxX"] - b["minX"], b["maxY"] - b["minY"]) / 64.0
    p = mrfield.srj_rasterize(doc, mode=mrfield.SRJ_LAYERS, resolution=res, clearance_cells=1,
                              min_clearance_mm=0.15, via_pad_mm=0.45, model=2)
    assert p.via_forbidden is not None
    og = mrcpu.Grid(p.w, p.h, p.layers, p.cost, p.via_forbidden, p.board)
    oe = mrcpu.Edges(mrcpu.axis_edge_costs(list(p.x_lines), p.w), mrcpu.axis_edge_costs(list(p.y_lines), p.h),
                     [VIA] * (p.layers - 1))
    nets = [mrcpu.Net(n.src, n.dst, list(n.pads), list(n.via_pads)) for n in p.nets]
    wins = [mrcpu.net_window(og, n.src, n.dst, n.pads) for n in nets]
    ref_p, ref_c = mrcpu.alone_batch(mrcpu.Spec(og, oe), nets, wins, retry_full=True, threads=8)
    got = eng.route_isolated_batch(p.grid(), [n.net() for n in p.nets], wins,
                                   mrfield.Edges(list(oe.x), list(oe.y), list(oe.via)))
    for i, r in enumerate(got):
        assert (r is None) == (ref_p[i] is None), (os.path.basename(path), i)
        if r is not None:
            assert r == (ref_p[i], ref_c[i]), (os.path.basename(path), i)
