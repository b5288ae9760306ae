"""SimpleRouteJson boards: the library's host producer (mrf_srj_rasterize,
metal-route_amd/csrc/srj.cpp) against the asserted values of the reference's
own unit tests (crates/mr-srj/src/lib.rs, crates/mr-grid/src/lib.rs; each test
names the one it restates) and against the brute-force restatement in
oracle/srjref.py on every committed SRJ fixture and on random boards.

Fixtures (tests/golden/srj/): the reference's CLI test payloads
(crates/mr-cli/tests/fixtures/srj/*.json) and the corpus boards its unit tests
read (benchmarks/corpus/bug-reports/bugreport20-*, bugreport21-*,
benchmarks/corpus/srj15/sample21-*, sample22-*), unchanged.

NaN cannot be written in JSON; where a reference test sets a field to NaN,
the restatement writes 1e999, which the producer reads as infinity: the same
non-finite rejection path.
"""
from __future__ import annotations

import copy
import glob
import json
import math
import os

import numpy as np
import pytest

import mrfield
import srjref
from mrfield import MAX, SRJ_LAYERS, SRJ_PHYSICAL, SRJ_ROUTE

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "srj", "*.json")))

SAMPLE = {
    "layerCount": 2,
    "bounds": {"minX": 0, "maxX": 10, "minY": 0, "maxY": 10},
    "obstacles": [{"type": "rect", "center": {"x": 5, "y": 5}, "width": 2, "height": 2}],
    "connections": [
        {"name": "VCC", "pointsToConnect": [{"x": 1, "y": 1}, {"x": 9, "y": 1}]},
        {"name": "GND", "pointsToConnect": [{"x": 1, "y": 9}, {"x": 5, "y": 9}, {"x": 9, "y": 9}]},
    ],
}


def rasterize(doc, res):
    """rasterize() (lib.rs:1323-1334): no clearance, the problem's layers."""
    return mrfield.srj_rasterize(doc, mode=SRJ_LAYERS, resolution=res)


def with_layers(doc, res, layers, cells, min_clear, via_pad, model=0):
    """rasterize_with_layers (lib.rs:1364-1384)."""
    return mrfield.srj_rasterize(doc, mode=SRJ_LAYERS, resolution=res, layers=layers, clearance_cells=cells,
                                 min_clearance_mm=min_clear, via_pad_mm=via_pad, model=model)


def xy(p, point):
    return srjref.nearest_line(list(p.x_lines), point[0]), srjref.nearest_line(list(p.y_lines), point[1])


def cell(p, point, layer=0):
    x, y = xy(p, point)
    return p.idx(x, y, min(layer, p.layers - 1))


def n_obstacles(p):
    return int(np.count_nonzero(p.cost == MAX))


def via_forbidden(p, c):
    return p.via_forbidden is not None and bool(p.via_forbidden[c])


# ---- lib.rs:2498-2641 ----------------------------------------------------------
def test_sample_dims_lines_and_nets():
    """rasterize_dims_and_net_count, two_point_connection_keeps_bare_name,
    k3_connection_decomposes_into_two_chained_nets (lib.rs:2509-2614)."""
    p = rasterize(SAMPLE, 1.0)
    assert (p.w, p.h, p.layers) == (15, 13, 2)
    assert list(p.x_lines) == [0.0, 0.5, 1.0, 2.0, 3.0, 4.0, 4.5, 5.0, 5.5, 6.0, 7.0, 8.0, 9.0, 9.5, 10.0]
    assert list(p.y_lines) == [0.0, 0.5, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 9.5, 10.0]
    assert len(p.nets) == 3
    vcc = [n for n in p.nets if n.name.startswith("VCC")]
    assert len(vcc) == 1 and vcc[0].name == "VCC"
    assert (vcc[0].src, vcc[0].dst) == (p.idx(2, 2), p.idx(12, 2))
    gnd = [n for n in p.nets if n.name.startswith("GND")]
    assert [n.name for n in gnd] == ["GND#0", "GND#1"]
    assert (gnd[0].src, gnd[0].dst, gnd[1].src, gnd[1].dst) == (p.idx(2, 10), p.idx(7, 10), p.idx(7, 10),
                                                                 p.idx(12, 10))


def test_sample_obstacle_cells():
    """rasterize_marks_expected_obstacle_cells (lib.rs:2541-2573)."""
    p = rasterize(SAMPLE, 1.0)
    for y in range(5, 8):
        for x in range(5, 10):
            assert p.is_obstacle(p.idx(x, y))
    for c in (p.idx(10, 6), p.idx(4, 6), p.idx(7, 8), p.idx(0, 0)):
        assert not p.is_obstacle(c)
    assert p.is_obstacle(p.idx(5, 5, 1)) and p.is_obstacle(p.idx(9, 7, 1))
    assert n_obstacles(p) == 30


def test_root_connection_name_groups():
    """root_connection_name_collapses_into_one_group (lib.rs:2615-2641)."""
    doc = {"layerCount": 1, "bounds": {"minX": 0, "maxX": 10, "minY": 0, "maxY": 10}, "obstacles": [],
           "connections": [
               {"name": "n_mst0", "rootConnectionName": "NET", "pointsToConnect": [{"x": 1, "y": 1}, {"x": 4, "y": 1}]},
               {"name": "n_mst1", "rootConnectionName": "NET", "pointsToConnect": [{"x": 6, "y": 8}, {"x": 9, "y": 8}]},
           ]}
    p = rasterize(doc, 1.0)
    assert len(p.nets) == 2
    assert [n.name.split("#")[0] for n in p.nets] == ["NET", "NET"]
    assert p.nets[0].name != p.nets[1].name


# ---- lib.rs:2729-2815 ----------------------------------------------------------
def test_obstacle_layers():
    """rasterize_obstacle_layer_isolated, rasterize_src_resolves_endpoint_layer,
    rasterize_obstacle_no_layers_blocks_all (lib.rs:2729-2815)."""
    b = {"minX": 0, "maxX": 10, "minY": 0, "maxY": 10}
    p = rasterize({"layerCount": 2, "bounds": b, "connections": [], "obstacles": [
        {"type": "rect", "layers": ["bottom"], "center": {"x": 5, "y": 5}, "width": 2, "height": 2}]}, 1.0)
    assert p.layers == 2
    for y in range(4, 7):
        for x in range(4, 7):
            assert p.is_obstacle(p.idx(x, y, 1)) and not p.is_obstacle(p.idx(x, y, 0))
    assert n_obstacles(p) == 9
    p = rasterize({"layerCount": 2, "bounds": b, "obstacles": [], "connections": [
        {"name": "a", "pointsToConnect": [{"x": 1, "y": 1, "layer": "top"}, {"x": 9, "y": 9, "layer": "bottom"}]}]},
        1.0)
    assert (p.nets[0].src, p.nets[0].dst) == (p.idx(2, 2, 0), p.idx(10, 10, 1))
    p = rasterize({"layerCount": 3, "bounds": b, "connections": [], "obstacles": [
        {"type": "rect", "center": {"x": 5, "y": 5}, "width": 2, "height": 2}]}, 1.0)
    assert all(p.is_obstacle(p.idx(4, 4, l)) for l in range(3))
    assert n_obstacles(p) == 27


def test_real_harness_fields_parse():
    """parses_real_harness_fields (lib.rs:3088-3117): unknown point fields are
    ignored."""
    doc = {"minTraceWidth": 0.1, "layerCount": 1, "bounds": {"minX": 0, "maxX": 5, "minY": 0, "maxY": 5},
           "obstacles": [{"type": "rect", "layers": ["top"], "center": {"x": 2, "y": 2}, "width": 0.38,
                          "height": 0.38, "connectedTo": ["pcb_smtpad_3"]}],
           "connections": [{"name": "source_trace_0", "pointsToConnect": [
               {"x": 1, "y": 1, "layer": "top", "pcb_port_id": "pcb_port_15"},
               {"x": 4, "y": 4, "layer": "top", "pcb_port_id": "pcb_port_16"}]}]}
    p = rasterize(doc, 0.5)
    assert len(p.nets) == 1 and p.nets[0].name == "source_trace_0"
    assert p.pins[p.nets[0].src] == (1.0, 1.0)


# ---- lib.rs:3120-3208 ----------------------------------------------------------
def test_own_pad_is_obstacle_but_passable_decoy_is_not():
    """rasterize_pad_is_obstacle_but_in_own_net_passable_pads,
    rasterize_keeps_decoy_pad_obstacle (lib.rs:3120-3208)."""
    doc = {"minTraceWidth": 0.1, "layerCount": 1, "bounds": {"minX": 0, "maxX": 10, "minY": 0, "maxY": 10},
           "obstacles": [
               {"type": "rect", "center": {"x": 5, "y": 5}, "width": 0.4, "height": 0.4, "connectedTo": ["pcb_smtpad_0"]},
               {"type": "rect", "center": {"x": 2, "y": 2}, "width": 0.4, "height": 0.4,
                "connectedTo": ["pcb_smtpad_decoy"]}],
           "connections": [{"name": "n", "pointsToConnect": [{"x": 5, "y": 5}, {"x": 9, "y": 9}]}]}
    p = rasterize(doc, 0.1)
    src = p.nets[0].src
    assert p.is_obstacle(src) and src in p.nets[0].pads
    cx, cy = xy(p, (5.0, 5.0))
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            c = p.idx(cx + dx, cy + dy)
            assert p.is_obstacle(c) and c in p.nets[0].pads
    decoy = cell(p, (2.0, 2.0))
    assert p.is_obstacle(decoy) and all(decoy not in n.pads for n in p.nets)
    doc2 = {"layerCount": 1, "bounds": {"minX": 0, "maxX": 10, "minY": 0, "maxY": 10},
            "obstacles": [{"type": "rect", "center": {"x": 5, "y": 5}, "width": 1, "height": 1}],
            "connections": [{"name": "n", "pointsToConnect": [{"x": 1, "y": 1}, {"x": 9, "y": 9}]}]}
    p = rasterize(doc2, 1.0)
    decoy = cell(p, (5.0, 5.0))
    assert p.is_obstacle(decoy) and all(decoy not in n.pads for n in p.nets)


# ---- typed physical rules (lib.rs:3277-3743) -------------------------------------
TYPED = {"layerCount": 2, "minTraceWidth": 0.1, "nominalTraceWidth": 0.1, "defaultObstacleMargin": 0.04,
         "minTraceToPadEdgeClearance": 0.07, "minViaEdgeToPadEdgeClearance": 0.09, "minViaHoleDiameter": 0.2,
         "minViaPadDiameter": 0.4, "bounds": {"minX": 0.0, "maxX": 4.0, "minY": 0.0, "maxY": 4.0}}


def typed(**kw):
    d = copy.deepcopy(TYPED)
    d.update(kw)
    return d


def pad(x, y, conn, shape="rect", layers=("top",), w=0.2, h=0.2, **kw):
    o = {"type": "rect", "shape": shape, "center": {"x": x, "y": y}, "width": w, "height": h,
         "layers": list(layers), "connectedTo": list(conn)}
    o.update(kw)
    return o


def test_typed_rules_accept_both_via_spellings():
    """typed_physical_rules_accept_modern_and_legacy_via_spellings (lib.rs:3277-3343)."""
    doc = {"layerCount": 2, "minTraceWidth": 0.1, "nominalTraceWidth": 0.12, "defaultObstacleMargin": 0.05,
           "minTraceToPadEdgeClearance": 0.06, "minViaEdgeToPadEdgeClearance": 0.08,
           "minViaHoleEdgeToViaHoleEdgeClearance": 0.09, "minBoardEdgeClearance": 0.2, "minViaHoleDiameter": 0.2,
           "min_via_hole_diameter": 0.2, "minViaPadDiameter": 0.45, "min_via_pad_diameter": 0.45,
           "allowViaInPad": False,
           "outline": [{"x": 0.0, "y": 0.0}, {"x": 4.0, "y": 0.0}, {"x": 4.0, "y": 4.0}, {"x": 0.0, "y": 4.0}],
           "bounds": {"minX": 0.0, "maxX": 4.0, "minY": 0.0, "maxY": 4.0},
           "obstacles": [pad(0.5, 0.5, ["n"]), pad(3.5, 3.5, ["n"])],
           "connections": [{"name": "n", "nominalTraceWidth": 0.12, "width": 0.12,
                            "pointsToConnect": [{"x": 0.5, "y": 0.5}, {"x": 3.5, "y": 3.5}]}]}
    r = mrfield.srj_uniform_rules(doc)
    assert r["trace_width_mm"] == 0.12 and r["via_hole_diameter_mm"] == 0.2
    assert r["via_pad_diameter_mm"] == 0.45 and r["via_hole_to_hole_clearance_mm"] == 0.09


def test_typed_rules_reject_mixed_widths_and_invalid_optionals():
    """typed_physical_rules_reject_mixed_connection_widths,
    _reject_invalid_optional_supported_rules (lib.rs:3392-3453)."""
    doc = typed(minTraceToPadEdgeClearance=0.05, minViaEdgeToPadEdgeClearance=0.08, minViaPadDiameter=0.45,
                defaultObstacleMargin=0.05, connections=[
                    {"name": "a", "nominalTraceWidth": 0.1, "pointsToConnect": [{"x": 0.5, "y": 0.5}, {"x": 3.5, "y": 0.5}]},
                    {"name": "b", "nominalTraceWidth": 0.2, "pointsToConnect": [{"x": 0.5, "y": 3.5}, {"x": 3.5, "y": 3.5}]}])
    assert mrfield.srj_uniform_rules(doc) is None
    valid = typed(minPadEdgeToPadEdgeClearance=0.11, minViaHoleEdgeToViaHoleEdgeClearance=0.1)
    assert mrfield.srj_uniform_rules(valid) is not None
    assert mrfield.srj_uniform_rules(dict(valid, minPadEdgeToPadEdgeClearance=-0.01)) is None
    bad = json.dumps(valid).replace('"minViaHoleEdgeToViaHoleEdgeClearance": 0.1', '"minViaHoleEdgeToViaHoleEdgeClearance": 1e999')
    assert mrfield.srj_uniform_rules(bad) is None


def test_typed_rules_validate_every_alias():
    """typed_physical_rules_validate_every_supplied_alias (lib.rs:3455-3502)."""
    valid = typed(min_via_hole_diameter=0.2, min_via_pad_diameter=0.4, minViaDiameter=0.4,
                  obstacles=[pad(0.5, 0.5, ["n"]), pad(3.5, 3.5, ["n"])],
                  connections=[{"name": "n", "nominalTraceWidth": 0.1, "width": 0.1,
                                "pointsToConnect": [{"x": 0.5, "y": 0.5}, {"x": 3.5, "y": 3.5}]}])
    assert mrfield.srj_uniform_rules(valid) is not None
    assert mrfield.srj_uniform_rules(dict(valid, min_via_pad_diameter=-0.4)) is None
    assert mrfield.srj_uniform_rules(dict(valid, min_via_hole_diameter=0.21)) is None
    for w in ("1e999", "0.11"):
        bad = json.dumps(valid).replace('"width": 0.1,', f'"width": {w},')
        assert mrfield.srj_uniform_rules(bad) is None, w


def test_typed_rules_annular_ring_and_fixed_holes():
    """typed_physical_rules_resolve_annular_minimum_and_conservatively_cover_fixed_holes (lib.rs:3503-3533)."""
    doc = typed(minViaHoleEdgeToViaHoleEdgeClearance=0.1, minViaPadDiameter=0.2)
    assert mrfield.srj_uniform_rules(doc) is None
    r = mrfield.srj_uniform_rules(dict(doc, minViaPadDiameter=0.3))
    assert r["via_hole_diameter_mm"] == 0.2 and r["via_pad_diameter_mm"] == 0.3
    assert mrfield.srj_uniform_rules(dict(doc, minViaPadDiameter=0.3, minViaHoleEdgeToViaHoleEdgeClearance=0.2)) is None


def test_typed_rules_unclassified_obstacles():
    """typed_physical_rules_reject_unclassified_obstacles (lib.rs:3535-3581)."""
    base = typed(defaultObstacleMargin=0.08, minTraceToPadEdgeClearance=0.08, minViaEdgeToPadEdgeClearance=0.08,
                 obstacles=[{"type": "rect", "shape": "rect", "center": {"x": 2.0, "y": 2.0}, "width": 0.2,
                             "height": 0.2}])
    assert mrfield.srj_uniform_rules(base) is None
    ok = copy.deepcopy(base)
    ok["obstacles"][0]["connectedTo"] = ["pad_1"]
    assert mrfield.srj_uniform_rules(ok) is not None
    assert mrfield.srj_uniform_rules(dict(ok, minTraceToPadEdgeClearance=0.07)) is None
    assert mrfield.srj_uniform_rules(dict(ok, minViaEdgeToPadEdgeClearance=0.07)) is None
    assert mrfield.srj_uniform_rules(dict(ok, minTraceToPadEdgeClearance=0.09)) is None
    assert mrfield.srj_uniform_rules(dict(ok, minPadEdgeToPadEdgeClearance=0.07)) is None


def test_typed_rules_terminal_pads():
    """typed_physical_rules_require_unambiguous_same_net_terminal_pads (lib.rs:3583-3683)."""
    conn = {"name": "edge", "rootConnectionName": "root", "pointsToConnect": [
        {"x": 0.5, "y": 0.5, "layer": "top"}, {"x": 3.5, "y": 3.5, "layer": "bottom"}]}
    doc = typed(connections=[conn])
    assert mrfield.srj_uniform_rules(doc) is None
    doc["obstacles"] = [pad(0.5, 0.5, ["root", "pcb_port_7"]), pad(3.5, 3.5, ["edge"], shape="circle", layers=["bottom"])]
    assert mrfield.srj_uniform_rules(doc) is not None
    unknown = copy.deepcopy(doc)
    unknown["obstacles"][0]["connectedTo"].append("mystery")
    assert mrfield.srj_uniform_rules(unknown) is None
    same_root = copy.deepcopy(doc)
    same_root["connections"].append(dict(copy.deepcopy(conn), name="sibling"))
    assert mrfield.srj_uniform_rules(same_root) is not None
    amb = copy.deepcopy(same_root)
    amb["connections"][1]["rootConnectionName"] = "other"
    for o in amb["obstacles"]:
        o["connectedTo"].append("other")
    assert mrfield.srj_uniform_rules(amb) is None
    corner = copy.deepcopy(doc)
    corner["connections"][0]["pointsToConnect"][1].update(x=3.6, y=3.6)
    assert mrfield.srj_uniform_rules(corner) is None
    foreign = copy.deepcopy(doc)
    foreign["obstacles"].append(pad(0.5, 0.5, ["foreign"], w=0.1, h=0.1))
    assert mrfield.srj_uniform_rules(foreign) is None
    unk_layer = copy.deepcopy(doc)
    unk_layer["connections"][0]["pointsToConnect"][0]["layer"] = "unknown"
    assert mrfield.srj_uniform_rules(unk_layer) is None


def test_typed_rules_geometry_gate_and_legacy_precedence():
    """typed_physical_rules_reject_geometry_the_axis_aligned_raster_cannot_model,
    legacy_min_clearance_precedes_typed_default_obstacle_margin (lib.rs:3685-3743)."""
    doc = typed(obstacles=[pad(2.0, 2.0, ["pad_1"], layers=(), h=0.4, ccwRotationDegrees=0.0)])
    del doc["obstacles"][0]["layers"]
    assert mrfield.srj_uniform_rules(doc) is not None
    for key, val, ok in (("ccwRotationDegrees", 45.0, False), ("shape", "oval", False), ("shape", "circle", True)):
        d = copy.deepcopy(doc)
        d["obstacles"][0][key] = val
        if key == "shape":
            d["obstacles"][0].pop("ccwRotationDegrees")
        assert (mrfield.srj_uniform_rules(d) is not None) == ok, (key, val)
    nan_x = json.dumps(doc).replace('"x": 2.0', '"x": 1e999')
    assert mrfield.srj_uniform_rules(nan_x) is None
    prec = typed(minClearance=0.04, defaultObstacleMargin=0.05, minTraceToPadEdgeClearance=0.06,
                 minViaEdgeToPadEdgeClearance=0.08, minViaPadDiameter=0.45)
    assert mrfield.srj_uniform_rules(prec)["obstacle_margin_mm"] == 0.04


def test_typed_raster_keeps_clearances_distinct():
    """typed_raster_keeps_generic_trace_and_pair_pad_clearances_distinct,
    typed_via_pad_clearance_uses_an_exact_edge_boundary (lib.rs:3744-3874)."""
    doc = typed(obstacles=[pad(2.0, 2.0, ["pad_probe"])], connections=[
        {"name": "probe", "pointsToConnect": [{"x": 2.24, "y": 2.0}, {"x": 2.38, "y": 2.0}, {"x": 3.5, "y": 3.5}]}])
    rules = mrfield.srj_uniform_rules(dict(doc, connections=[]))
    assert (rules["obstacle_margin_mm"], rules["trace_to_pad_clearance_mm"], rules["via_to_pad_clearance_mm"]) == \
        (0.04, 0.07, 0.09)
    p = mrfield.srj_rasterize(doc, mode=SRJ_PHYSICAL, resolution=0.5, layers=2, rules=rules)
    assert p.is_obstacle(cell(p, (2.24, 2.0)))
    c = cell(p, (2.38, 2.0))
    assert not p.is_obstacle(c) and via_forbidden(p, c)
    fill = typed(bounds={"minX": 0.0, "maxX": 0.2, "minY": 0.0, "maxY": 1.0}, connections=[
        {"name": "fill", "pointsToConnect": [{"x": 0.0, "y": 0.5}, {"x": 0.2, "y": 0.5}]}])
    p = mrfield.srj_rasterize(fill, mode=SRJ_PHYSICAL, resolution=0.5, layers=2,
                              rules=mrfield.srj_uniform_rules(dict(fill, connections=[])))
    assert any(abs(x - 0.1) <= 1e-6 for x in p.x_lines)
    doc = typed(defaultObstacleMargin=0.05, minTraceToPadEdgeClearance=0.05, minViaEdgeToPadEdgeClearance=0.2,
                obstacles=[pad(2.0, 2.0, ["pad_boundary"])], connections=[
                    {"name": "probes", "pointsToConnect": [{"x": 2.499999, "y": 2.0}, {"x": 2.5, "y": 2.0}]}])
    p = mrfield.srj_rasterize(doc, mode=SRJ_PHYSICAL, resolution=0.5, layers=2,
                              rules=mrfield.srj_uniform_rules(dict(doc, connections=[])))
    assert via_forbidden(p, cell(p, (2.499999, 2.0)))
    assert not via_forbidden(p, cell(p, (2.5, 2.0)))


def test_physical_mode_fails_closed_without_coherent_rules():
    with pytest.raises(mrfield.MrfError):
        mrfield.srj_rasterize(SAMPLE, mode=SRJ_PHYSICAL, resolution=1.0)


# ---- clearance rollout (lib.rs:3875-4065) -----------------------------------------
def test_clearance_model_boundary():
    """clearance_rollout_trigger_has_physical_boundary_and_safe_zero_rule (lib.rs:3875-3925)."""
    t = 0.15 + 0.45
    cases = [(t - 2.0e-9, 1, 0.15, 0.45, False), (t - 0.5e-9, 1, 0.15, 0.45, True), (t, 1, 0.15, 0.45, True),
             (0.15, 1, 0.15, 0.45, False), (1.0, 0, 0.0, 0.45, True), (1.0, 0, 0.0, 0.0, False),
             (0.15, 1, 0.15, 0.0, True), (1.0, 1, float("nan"), 0.45, False)]
    for res, cells, mc, vp, exact in cases:
        p = with_layers(SAMPLE, res, 2, cells, mc, vp)
        assert p.exact_model == exact, (res, cells, mc, vp)


PROBE = {"layerCount": 2, "minClearance": 0.15, "minTraceWidth": 0.15,
         "bounds": {"minX": -2.0, "maxX": 2.0, "minY": -2.0, "maxY": 2.0},
         "obstacles": [{"type": "rect", "center": {"x": 0.0, "y": 0.0}, "width": 0.2, "height": 0.2,
                        "layers": ["top"], "connectedTo": ["connectivity_net_foreign"]}],
         "connections": [{"name": "probe", "pointsToConnect": [{"x": 0.4, "y": 0.0, "layer": "top"},
                                                               {"x": 1.5, "y": 0.0, "layer": "top"}]}]}


def same_problem(a, b):
    assert (a.w, a.h, a.layers) == (b.w, b.h, b.layers)
    assert np.array_equal(a.x_lines, b.x_lines) and np.array_equal(a.y_lines, b.y_lines)
    assert np.array_equal(a.cost, b.cost)
    for k in ("via_forbidden", "board"):
        u, v = getattr(a, k), getattr(b, k)
        assert (u is None) == (v is None) and (u is None or np.array_equal(u, v)), k
    assert a.nets == b.nets and a.pins == b.pins


def test_low_quantization_is_identical_to_forced_legacy():
    """low_quantization_raster_is_identical_to_forced_legacy_model,
    low_quantization_real_boards_are_identical_to_forced_legacy_model (lib.rs:3936-4065)."""
    auto = with_layers(PROBE, 0.2, 2, 1, 0.15, 0.45)
    same_problem(auto, with_layers(PROBE, 0.2, 2, 1, 0.15, 0.45, model=1))
    assert auto.via_forbidden is None and all(not n.via_pads for n in auto.nets)
    assert with_layers(PROBE, 0.2, 2, 1, 0.15, 0.45, model=2).via_forbidden is not None
    for name in ("sample21-region-reroute.srj.json", "sample22-region-reroute.srj.json"):
        doc = srjref.load(os.path.join(HERE, "golden", "srj", name))
        b = doc["bounds"]
        res = max(b["maxX"] - b["minX"], b["maxY"] - b["minY"]) / 64.0
        cells = math.ceil(0.15 / res)
        a = with_layers(doc, res, doc["layerCount"], cells, 0.15, 0.45)
        assert not a.exact_model, name
        same_problem(a, with_layers(doc, res, doc["layerCount"], cells, 0.15, 0.45, model=1))


# ---- pad clearance (lib.rs:4067-4497) -------------------------------------------------
CLEARANCE_SRJ = {"layerCount": 1, "bounds": {"minX": 0, "maxX": 7, "minY": 0, "maxY": 7},
                 "obstacles": [{"type": "rect", "center": {"x": 1.5, "y": 3.5}, "width": 0.5, "height": 0.5,
                                "connectedTo": ["pad_a"]},
                               {"type": "rect", "center": {"x": 4.5, "y": 3.5}, "width": 0.5, "height": 0.5,
                                "connectedTo": ["pad_b"]}],
                 "connections": [{"name": "a", "pointsToConnect": [{"x": 1.5, "y": 3.5}, {"x": 1.5, "y": 0.5}]},
                                 {"name": "b", "pointsToConnect": [{"x": 4.5, "y": 3.5}, {"x": 4.5, "y": 0.5}]}]}


def test_pad_clearance_halo_and_own_escape():
    """rasterize_pad_clearance_reserves_halo_and_unmasks_own (lib.rs:4067-4124)."""
    p = with_layers(CLEARANCE_SRJ, 1.0, 1, 1, 0.0, 0.0)
    ax, ay = xy(p, (1.5, 3.5))
    pad_a, pad_b = p.idx(ax, ay), cell(p, (4.5, 3.5))
    assert p.is_obstacle(pad_a) and p.is_obstacle(pad_b)
    for ny in range(p.h):
        for nx in range(p.w):
            if abs(p.x_lines[nx] - p.x_lines[ax]) <= 1.0 and abs(p.y_lines[ny] - p.y_lines[ay]) <= 1.0:
                assert p.is_obstacle(p.idx(nx, ny))
    net_a = next(n for n in p.nets if n.name == "a")
    for x, y in ((max(ax - 1, 0), ay), (ax + 1, ay), (ax, max(ay - 1, 0)), (ax, ay + 1)):
        if x < p.w and y < p.h:
            assert p.idx(x, y) in net_a.pads
    assert pad_b not in net_a.pads


def test_coarse_exact_clearance_probe():
    """true_clearance_does_not_round_up_to_coarse_fill_pitch,
    via_pad_reservation_is_distinct_from_planar_pad_reservation (lib.rs:4131-4180)."""
    p = with_layers(PROBE, 1.0, 2, 1, 0.15, 0.45)
    c = cell(p, (0.4, 0.0))
    assert not p.is_obstacle(c) and via_forbidden(p, c)


def via_probe(point, owned_pad):
    obs = [{"type": "rect", "center": {"x": 0.0, "y": 0.0}, "width": 0.2, "height": 0.2, "layers": ["top"],
            "connectedTo": ["foreign"]}]
    if owned_pad:
        obs.append({"type": "rect", "center": {"x": point[0], "y": point[1]}, "width": 0.01, "height": 0.01,
                    "layers": ["top"], "connectedTo": ["probe"]})
    doc = {"layerCount": 2, "minTraceWidth": 0.15, "minClearance": 0.15,
           "bounds": {"minX": -1.0, "maxX": 2.0, "minY": -1.0, "maxY": 2.0}, "obstacles": obs,
           "connections": [{"name": "probe", "pointsToConnect": [{"x": point[0], "y": point[1], "layer": "top"},
                                                                 {"x": 1.5, "y": 1.5, "layer": "top"}]}]}
    p = with_layers(doc, 1.0, 2, 1, 0.15, 0.45)
    return p, cell(p, point)


def test_euclidean_via_halo():
    """via_halo_diagonal_just_outside_is_legal_and_preserves_own_exemption,
    via_halo_diagonal_just_inside_is_blocked_and_clips_own_exemption,
    via_halo_axis_boundary_is_legal_but_two_eps_inside_is_blocked (lib.rs:4248-4318)."""
    margin = 0.15 + 0.45 / 2.0
    out = 0.1 + margin / math.sqrt(2.0) + 1e-6
    p, c = via_probe((out, out), False)
    assert not via_forbidden(p, c)
    p, c = via_probe((out, out), True)
    assert via_forbidden(p, c) and c in p.nets[0].pads and c in p.nets[0].via_pads
    inn = 0.1 + margin / math.sqrt(2.0) - 1e-6
    p, c = via_probe((inn, inn), False)
    assert via_forbidden(p, c)
    p, c = via_probe((inn, inn), True)
    assert via_forbidden(p, c) and c in p.nets[0].pads and c not in p.nets[0].via_pads
    edge = 0.1 + margin
    assert not via_forbidden(*via_probe((edge, 0.0), False))
    assert via_forbidden(*via_probe((edge - 2.0e-9, 0.0), False))
    assert not via_forbidden(*via_probe((edge + 1e-6, 0.0), False))


def test_via_pad_exemptions():
    """foreign_via_halo_wins_when_it_overlaps_another_pad_core,
    via_pad_permissions_are_layer_local_for_real_rasterized_smd_pads,
    matching_layer_endpoints_authorize_a_real_through_pad_via,
    zero_clearance_still_reserves_via_radius_and_exempts_only_own_core (lib.rs:4319-4496)."""
    base = {"layerCount": 2, "minClearance": 0.15, "minTraceWidth": 0.15}
    doc = dict(base, bounds={"minX": -2.0, "maxX": 2.0, "minY": -2.0, "maxY": 2.0}, obstacles=[
        {"type": "rect", "center": {"x": 0.0, "y": 0.0}, "width": 0.2, "height": 0.2, "layers": ["top"],
         "connectedTo": ["connectivity_net_foreign"]},
        {"type": "rect", "center": {"x": 0.4, "y": 0.0}, "width": 0.1, "height": 0.1, "layers": ["top"],
         "connectedTo": ["connectivity_net_own"]}],
        connections=[{"name": "own", "pointsToConnect": [{"x": 0.4, "y": 0.0, "layer": "top"},
                                                         {"x": 1.5, "y": 0.0, "layer": "top"}]}])
    p = with_layers(doc, 1.0, 2, 1, 0.15, 0.45)
    core = cell(p, (0.4, 0.0))
    own = next(n for n in p.nets if n.name == "own")
    assert via_forbidden(p, core) and core in own.pads and core not in own.via_pads
    assert set(own.via_pads) <= set(own.pads)
    doc = dict(base, bounds={"minX": -1.0, "maxX": 2.0, "minY": -1.0, "maxY": 1.0}, obstacles=[
        {"type": "rect", "center": {"x": 0.0, "y": 0.0}, "width": 0.3, "height": 0.3, "layers": ["top"]},
        {"type": "rect", "center": {"x": 0.0, "y": 0.0}, "width": 0.3, "height": 0.3, "layers": ["bottom"]},
        {"type": "rect", "center": {"x": 1.0, "y": 0.0}, "width": 0.3, "height": 0.3, "layers": ["bottom"]}],
        connections=[{"name": "smd", "pointsToConnect": [{"x": 0.0, "y": 0.0, "layer": "top"},
                                                         {"x": 1.0, "y": 0.0, "layer": "bottom"}]}])
    p = with_layers(doc, 1.0, 2, 1, 0.15, 0.45)
    top, bottom = cell(p, (0.0, 0.0), 0), cell(p, (0.0, 0.0), 1)
    n = p.nets[0]
    assert top in n.via_pads and bottom not in n.pads and bottom not in n.via_pads
    assert via_forbidden(p, top) and via_forbidden(p, bottom)
    doc = dict(base, bounds={"minX": -1.0, "maxX": 1.0, "minY": -1.0, "maxY": 1.0}, obstacles=[
        {"type": "rect", "center": {"x": 0.0, "y": 0.0}, "width": 0.4, "height": 0.4, "layers": ["top", "bottom"]}],
        connections=[{"name": "through", "pointsToConnect": [{"x": 0.0, "y": 0.0, "layer": "top"},
                                                             {"x": 0.0, "y": 0.0, "layer": "bottom"}]}])
    p = with_layers(doc, 1.0, 2, 1, 0.15, 0.45)
    n = p.nets[0]
    assert n.src != n.dst
    for c in (n.src, n.dst):
        assert via_forbidden(p, c) and c in n.pads and c in n.via_pads
    zero = {"layerCount": 2, "minClearance": 0.0, "minTraceWidth": 0.15,
            "bounds": {"minX": -1.0, "maxX": 2.0, "minY": -1.0, "maxY": 1.0},
            "obstacles": [{"type": "rect", "center": {"x": 0.0, "y": 0.0}, "width": 0.2, "height": 0.2,
                           "layers": ["top"]}]}
    p = with_layers(dict(zero, connections=[{"name": "probe", "pointsToConnect": [
        {"x": 0.3, "y": 0.0, "layer": "top"}, {"x": 1.0, "y": 0.0, "layer": "top"}]}]), 1.0, 2, 0, 0.0, 0.45)
    c = cell(p, (0.3, 0.0))
    assert via_forbidden(p, c) and c not in p.nets[0].via_pads
    p = with_layers(dict(zero, connections=[{"name": "own", "pointsToConnect": [
        {"x": 0.0, "y": 0.0, "layer": "top"}, {"x": 1.0, "y": 0.0, "layer": "top"}]}]), 1.0, 2, 0, 0.0, 0.45)
    c = cell(p, (0.0, 0.0))
    assert via_forbidden(p, c) and c in p.nets[0].via_pads


def test_clearance_zero_is_rasterize():
    """rasterize_clearance_zero_is_byte_identical (lib.rs:4497-4531)."""
    base = rasterize(CLEARANCE_SRJ, 1.0)
    zero = with_layers(CLEARANCE_SRJ, 1.0, 1, 0, 0.0, 0.0)
    assert np.array_equal(base.cost, zero.cost) and base.nets == zero.nets
    assert n_obstacles(zero) == 18
    ax, _ = xy(zero, (1.5, 3.5))
    top = xy(zero, (1.5, 3.25))[1]
    if top >= 1:
        assert not zero.is_obstacle(zero.idx(ax, top - 1))


TRACK_GT = {"minTraceWidth": 0.3, "minClearance": 0.1, "layerCount": 1,
            "bounds": {"minX": 0, "maxX": 6, "minY": 0, "maxY": 6},
            "obstacles": [{"type": "rect", "center": {"x": 2.0, "y": 2.0}, "width": 0.6, "height": 0.6,
                           "connectedTo": ["pad_a"]},
                          {"type": "rect", "center": {"x": 3.3, "y": 2.0}, "width": 0.6, "height": 0.6,
                           "connectedTo": ["pad_b"]}],
            "connections": [{"name": "a", "pointsToConnect": [{"x": 2.0, "y": 2.0}, {"x": 2.0, "y": 5.0}]},
                            {"name": "b", "pointsToConnect": [{"x": 3.3, "y": 2.0}, {"x": 3.3, "y": 5.0}]}]}


def test_track_wider_than_clearance():
    """track_gt_clearance_reserves_centreline_margin_zero_free_nodes,
    track_gt_clearance_keeps_own_pad_escapable (lib.rs:4577-4723)."""
    cells = math.ceil(0.1 / 0.05)
    p = with_layers(TRACK_GT, 0.05, 1, cells, 0.0, 0.0)
    margin = 0.1 + 0.3 / 2.0
    pads = [srjref.rect(o) for o in TRACK_GT["obstacles"]]
    offending = 0
    for ny in range(p.h):
        for nx in range(p.w):
            if p.is_obstacle(p.idx(nx, ny)):
                continue
            x, y = p.x_lines[nx], p.y_lines[ny]
            for r in pads:
                if r[0] <= x <= r[1] and r[2] <= y <= r[3]:
                    continue
                d = max(max(r[0] - x, x - r[1], 0.0), max(r[2] - y, y - r[3], 0.0))
                offending += d < margin - 1e-9
    assert offending == 0 and int(np.count_nonzero(p.cost != MAX)) > 0
    a = next(n for n in p.nets if n.name == "a")
    ax, ay = xy(p, (2.0, 2.0))
    for x, y in ((max(ax - 1, 0), ay), (ax + 1, ay), (ax, max(ay - 1, 0)), (ax, ay + 1)):
        if x < p.w and y < p.h:
            assert p.idx(x, y) in a.pads
    assert cell(p, (3.3, 2.0)) not in a.pads


HANAN = {"minTraceWidth": 0.15, "minClearance": 0.2, "layerCount": 1,
         "bounds": {"minX": 0, "maxX": 20, "minY": 0, "maxY": 20},
         "obstacles": [{"type": "rect", "center": {"x": 7.3, "y": 4.1}, "width": 1.2, "height": 0.8},
                       {"type": "rect", "center": {"x": 12.6, "y": 13.9}, "width": 2.0, "height": 1.0}],
         "connections": [{"name": "net1", "pointsToConnect": [{"x": 3.14, "y": 2.72}, {"x": 16.5, "y": 17.25}]}]}


def test_hanan_lines_through_features():
    """pad_center_maps_to_exact_node, lines_pass_through_endpoints_and_obstacle_edges (lib.rs:4725-4798)."""
    p = rasterize(HANAN, 1.0)

    def has(lines, v):
        return any(abs(l - v) <= 1e-6 for l in lines)

    for c in HANAN["connections"]:
        for q in c["pointsToConnect"]:
            x, y = xy(p, (q["x"], q["y"]))
            assert abs(p.x_lines[x] - q["x"]) <= 1e-6 and abs(p.y_lines[y] - q["y"]) <= 1e-6
            assert has(p.x_lines, q["x"]) and has(p.y_lines, q["y"])
    assert has(p.x_lines, 0.0) and has(p.x_lines, 20.0) and has(p.y_lines, 0.0) and has(p.y_lines, 20.0)
    for o in HANAN["obstacles"]:
        x0, x1, y0, y1 = srjref.rect(o)
        assert has(p.x_lines, x0) and has(p.x_lines, x1) and has(p.y_lines, y0) and has(p.y_lines, y1)
    x0 = xy(p, (3.14, 2.72))[0]
    assert x0 + 2 < p.w


# ---- fill lanes and the cell budget (lib.rs:4855-5352, :5836-5855) --------------------------
def pads_board(centres, w, bounds, track=0.1, clear=0.1):
    return {"layerCount": 1, "minTraceWidth": track, "minClearance": clear,
            "bounds": dict(zip(("minX", "maxX", "minY", "maxY"), bounds)),
            "obstacles": [{"type": "rect", "center": {"x": x, "y": y}, "width": w, "height": w} for x, y in centres],
            "connections": []}


def grid_lines(doc, track, clearance, escape, res=None):
    """build_grid_lines(srj, 1, track, clearance, escape) through the legacy
    projection: clearance = cells x resolution, escape = the true rule."""
    p = with_layers(dict(doc, minTraceWidth=track), clearance if res is None else res, 1, 1, escape, 0.0, model=1)
    return list(p.x_lines), list(p.y_lines)


def test_fill_channel_between_pads():
    """fill_inserts_routable_channel_between_adjacent_pads (lib.rs:4855-4968)."""
    xs, _ = grid_lines(pads_board([(0.0, 0.0), (1.0, 0.0)], 0.4, (-1.0, 2.0, -1.0, 1.0)), 0.1, 0.1, 0.1)
    assert any(0.3 + 1e-6 < x < 0.7 - 1e-6 for x in xs)
    xs, _ = grid_lines(pads_board([(-0.36, 0.0), (0.36, 0.0)], 0.4, (-1.0, 2.0, -1.0, 1.0)), 0.1, 0.1, 0.1)
    assert any(abs(x) <= 0.06 + 1e-6 for x in xs)


def test_escape_lanes():
    """escape_lanes_added_for_sub_pitch_pad_gaps, escape_lanes_inert_when_true_clearance_not_finer
    (lib.rs:4970-5101)."""
    board = pads_board([(0.0, 0.0), (2.54, 0.0)], 1.6, (-2.0, 4.54, -2.0, 2.0), track=0.15, clear=0.15)
    xs, _ = grid_lines(board, 0.15, 0.618, 0.618)
    assert not any(0.8 + 1e-6 < x < 1.74 - 1e-6 for x in xs)
    xs, _ = grid_lines(board, 0.15, 0.618, 0.15)
    assert any(abs(x - 1.27) <= 1e-6 for x in xs)
    xs, _ = grid_lines(dict(board, minClearance=0.618), 0.15, 0.618, 0.618)
    assert not any(0.8 + 1e-6 < x < 1.74 - 1e-6 for x in xs)


def test_no_lane_in_sub_track_gap():
    """fill_no_lane_when_pads_too_tight_but_go_around_exists (lib.rs:5102-5174)."""
    xs, _ = grid_lines(pads_board([(0.0, 0.0), (0.42, 0.0)], 0.4, (-2.0, 2.0, -1.0, 1.0)), 0.1, 0.1, 0.1)
    assert not any(0.2 + 1e-6 < x < 0.22 - 1e-6 for x in xs)
    assert any(0.62 + 1e-6 < x < 2.0 - 1e-6 for x in xs)


def test_cell_budget_enforced(monkeypatch):
    """cell_budget_is_enforced_by_thinning_fill, cell_budget_default_is_floor (lib.rs:5176-5352)."""
    monkeypatch.delenv("MR_CELL_BUDGET", raising=False)
    n, pitch = 60, 0.5
    span = (n - 1) * pitch
    board = pads_board([(i * pitch, j * pitch) for i in range(n) for j in range(n)], 0.4,
                       (-1.0, span + 1.0, -1.0, span + 1.0))
    xs, ys = grid_lines(board, 0.1, 0.1, 0.1)
    assert len(xs) * len(ys) <= 160_000
    for i in range(n):
        for e in (i * pitch - 0.2, i * pitch + 0.2):
            assert any(abs(x - e) <= 1e-6 for x in xs)
    # a ceiling the full grid exceeds: fill lanes go, the most crowded first,
    # in the oracle's drop order; every feature line stays
    p = mrfield.srj_rasterize(dict(board, minTraceWidth=0.1), mode=SRJ_LAYERS, resolution=0.1, clearance_cells=1,
                              min_clearance_mm=0.1, model=1, cell_budget=50_000)
    assert p.fill_dropped > 0 and p.w * p.h <= 50_000
    for i in range(n):
        for e in (i * pitch - 0.2, i * pitch + 0.2):
            assert any(abs(x - e) <= 1e-6 for x in p.x_lines)
    ox, oy = srjref.grid_lines(dict(board, minTraceWidth=0.1), 0.1, 0.1, 0.1, 50_000)
    assert list(p.x_lines) == ox and list(p.y_lines) == oy


def test_enforce_budget_direct():
    """adaptive_budget_retains_fill_when_over_floor_but_under_headroom,
    enforce_budget_drops_all_fill_when_features_alone_over_budget,
    enforce_budget_keeps_features_and_drops_redundant_fill (lib.rs:5266-5339, :5836-5855)."""
    f = [float(k) for k in range(200)]
    fill = [k + 0.5 for k in range(199)]
    xs, ys = mrfield.srj_enforce_budget(f, f, fill, fill, 200 * 200 * 4)
    assert len(xs) == 199 and len(ys) == 199
    f = [float(k) for k in range(50)]
    fill = [k + 0.5 for k in range(49)]
    assert mrfield.srj_enforce_budget(f, f, fill, fill, 100) == ([], [])
    xf, yf = [0.0, 10.0], [0.0, 1.0, 2.0]
    xs, ys = mrfield.srj_enforce_budget(xf, yf, [k * 0.001 for k in range(1, 5001)], [], 160_000)
    assert (len(xf) + len(xs)) * (len(yf) + len(ys)) <= 160_000
    # the oracle's drop order gives the same survivors
    ox, oy = srjref.enforce_budget(xf, yf, [k * 0.001 for k in range(1, 5001)], [], 160_000)
    assert xs == ox and ys == oy


def test_cell_budget_env(monkeypatch):
    """MR_CELL_BUDGET (lib.rs:741-792): a pinned budget thins the fill, 0 and an
    unparseable value leave the default (no thinning below the floor here)."""
    board = pads_board([(0.0, 0.0), (5.0, 5.0)], 0.4, (-5.0, 10.0, -5.0, 10.0))
    monkeypatch.delenv("MR_CELL_BUDGET", raising=False)
    base = grid_lines(board, 0.1, 0.1, 0.1)
    assert len(base[0]) * len(base[1]) > 400
    monkeypatch.setenv("MR_CELL_BUDGET", "400")
    xs, ys = grid_lines(board, 0.1, 0.1, 0.1)
    assert len(xs) * len(ys) <= 400
    assert (xs, ys) == tuple(srjref.grid_lines(dict(board, minTraceWidth=0.1), 0.1, 0.1, 0.1))
    for v in ("0", "not-a-number", " adaptive "):
        monkeypatch.setenv("MR_CELL_BUDGET", v)
        assert grid_lines(board, 0.1, 0.1, 0.1) == base, v


# ---- corpus fixtures and the board outline (lib.rs:5441-5834) --------------------------
def fixture(name):
    return srjref.load(os.path.join(HERE, "golden", "srj", name))


def test_corpus_fixture_deterministic():
    """checked_in_corpus_fixture_parses_and_rasterizes_deterministically (lib.rs:5441-5461)."""
    doc = fixture("bugreport20-obstacle-clipping.srj.json")
    assert doc["layerCount"] == 2 and len(doc["obstacles"]) == 4 and len(doc["connections"]) == 1
    a, b = rasterize(doc, 0.15), rasterize(doc, 0.15)
    same_problem(a, b)
    assert len(a.nets) == 1 and a.layers == 2 and a.nets[0].src != a.nets[0].dst


def test_bugreport21_outline_mask_equals_continuous_predicates():
    """bugreport21_raster_mask_matches_exact_concave_outline_geometry (lib.rs:5462-5517)."""
    doc = fixture("bugreport21-board-outline.srj.json")
    p = with_layers(doc, 0.15, 2, 1, 0.15, 0.45)
    assert p.board is not None
    assert -2.0 not in list(p.x_lines) and 2.0 not in list(p.x_lines) and 2.0 not in list(p.y_lines)
    o = srjref.outline_from(doc, 0.15, 0.45)
    plane = o.mask_plane(list(p.x_lines), list(p.y_lines))
    for l in range(p.layers):
        assert np.array_equal(p.board[l * p.w * p.h:(l + 1) * p.w * p.h], plane)


def lcg(seed):
    s = [seed]

    def nxt(lo, hi):
        s[0] = (s[0] * 6364136223846793005 + 1442695040888963407) % 2 ** 64
        return lo + (hi - lo) * ((s[0] >> 11) * (1.0 / 2 ** 53))

    return nxt


def test_outline_mask_random_concave_polygons():
    """indexed_board_outline_mask_matches_naive_random_concave_polygons (lib.rs:5597-5707):
    random concave outlines with probes at exact, epsilon- and keepout-adjacent
    distances from a sloped edge; the producer's indexed mask equals the
    continuous predicates over every edge."""
    rng = lcg(0x4D45_5441_4C52_4F55)
    for case in range(32):
        nv = 6 + case % 13
        rot = rng(-0.4, 0.4)
        verts = []
        for k in range(nv):
            ang = rot + 2 * math.pi * k / nv
            r = (8.0 if k % 2 == 0 else 3.5) + rng(-0.45, 0.45)
            verts.append((r * math.cos(ang), r * math.sin(ang)))
        o = srjref.Outline(verts, 0.21, 0.18 / 2.0, 0.46 / 2.0)
        xl, yl = [-10.0, 0.0, 10.0], [-10.0, 0.0, 10.0]
        for _ in range(18):
            xl.append(rng(-9.8, 9.8))
            yl.append(rng(-9.8, 9.8))
        xl += [v[0] for v in verts]
        yl += [v[1] for v in verts]
        a, b = verts[case % nv], verts[(case + 1) % nv]
        mid = ((a[0] + b[0]) / 2.0, (a[1] + b[1]) / 2.0)
        ln = math.hypot(b[0] - a[0], b[1] - a[1])
        inward = (-(b[1] - a[1]) / ln, (b[0] - a[0]) / ln)
        tk, vk = 0.21 + 0.09, 0.21 + 0.23
        for d in (0.0, 0.5e-9, 1e-9, 2e-9, tk - 2e-9, tk - 1e-9, tk, tk + 1e-9, vk - 1e-9, vk, vk + 1e-9):
            for sg in (-1.0, 1.0):
                xl.append(mid[0] + inward[0] * d * sg)
                yl.append(mid[1] + inward[1] * d * sg)
        xl, yl = sorted(set(xl)), sorted(set(yl))
        doc = {"layerCount": 1 + case % 4, "minTraceWidth": 0.18, "minBoardEdgeClearance": 0.21,
               "bounds": {"minX": -10.0, "maxX": 10.0, "minY": -10.0, "maxY": 10.0},
               "outline": [{"x": x, "y": y} for x, y in verts]}
        got = mrfield.srj_outline_plane(doc, 0.18, 0.46, xl, yl)
        assert np.array_equal(got, o.mask_plane(xl, yl)), case


def test_malformed_outline_fails_closed():
    """malformed_active_outline_raster_fails_closed (lib.rs:5809-5834)."""
    doc = {"layerCount": 1, "minTraceWidth": 0.2, "bounds": {"minX": 0.0, "maxX": 10.0, "minY": 0.0, "maxY": 10.0},
           "outline": [{"x": 0.0, "y": 0.0}, {"x": 10.0, "y": 10.0}, {"x": 0.0, "y": 10.0}, {"x": 10.0, "y": 0.0}],
           "connections": [{"name": "n", "pointsToConnect": [{"x": 1.0, "y": 1.0}, {"x": 9.0, "y": 9.0}]}]}
    p = rasterize(doc, 0.2)
    assert p.board is not None and bool(np.all(p.board & mrfield.BOARD_TRACE_NODE))
    assert bool(np.all(p.board == 63))


# ---- the producer against the brute-force oracle ------------------------------------
def oracle_raster(doc, res, layers, cells, mc, vp, exact):
    return srjref.rasterize(doc, res, layers, cells, mc, mc, mc, vp, None, exact)


def same_as_oracle(p, r, tag):
    assert (p.w, p.h, p.layers) == (r.w, r.h, r.layers), tag
    assert list(p.x_lines) == r.x and list(p.y_lines) == r.y, tag
    assert np.array_equal(p.cost, r.cost), tag
    for k in ("via_forbidden", "board"):
        u, v = getattr(p, k), getattr(r, k)
        assert (u is None) == (v is None), (tag, k)
        if u is not None:
            assert np.array_equal(u, v), (tag, k)
    got = [{"name": n.name, "src": n.src, "dst": n.dst, "pads": n.pads, "via_pads": n.via_pads} for n in p.nets]
    assert got == r.nets, tag
    assert p.pins == r.pins, tag


SMALL = [f for f in FIXTURES if os.path.getsize(f) < 20000]


@pytest.mark.parametrize("path", SMALL, ids=[os.path.basename(f) for f in SMALL])
def test_fixtures_match_oracle(path):
    """Every small committed fixture at its CLI resolution under the legacy and
    the exact model, with and without a via pad."""
    doc = srjref.load(path)
    b = doc["bounds"]
    res = max(b["maxX"] - b["minX"], b["maxY"] - b["minY"]) / 64.0
    L = max(doc["layerCount"], 1)
    for cells, mc, vp, exact in ((math.ceil(0.15 / res), 0.15, 0.45, False), (1, 0.15, 0.45, True),
                                 (0, 0.0, 0.0, False), (2, 0.1, 0.0, True)):
        p = with_layers(doc, res, L, cells, mc, vp, model=2 if exact else 1)
        same_as_oracle(p, oracle_raster(doc, res, L, cells, mc, vp, exact), (os.path.basename(path), cells, exact))


def random_doc(rng, case):
    L = 1 + case % 3
    names = srjref.layer_names(L)
    side = rng.uniform(4.0, 12.0)
    obs = []
    for _ in range(rng.integers(0, 12)):
        o = {"type": "rect", "center": {"x": float(rng.uniform(0, side)), "y": float(rng.uniform(0, side))},
             "width": float(rng.uniform(0.05, 1.5)), "height": float(rng.uniform(0.05, 1.5))}
        if rng.random() < 0.6:
            o["layers"] = [str(rng.choice(names))] if rng.random() < 0.7 else ["nowhere"]
        obs.append(o)
    conns = []
    for k in range(rng.integers(1, 5)):
        pts = []
        for _ in range(rng.integers(1, 4)):
            if obs and rng.random() < 0.6:
                o = obs[rng.integers(0, len(obs))]
                q = {"x": o["center"]["x"], "y": o["center"]["y"]}
            else:
                q = {"x": float(rng.uniform(0, side)), "y": float(rng.uniform(0, side))}
            if rng.random() < 0.5:
                q["layer"] = str(rng.choice(names))
            pts.append(q)
        c = {"name": f"c{k}", "pointsToConnect": pts}
        if rng.random() < 0.3:
            c["rootConnectionName"] = "root"
        conns.append(c)
    doc = {"layerCount": L, "bounds": {"minX": 0.0, "maxX": side, "minY": 0.0, "maxY": side},
           "obstacles": obs, "connections": conns}
    if rng.random() < 0.7:
        doc["minTraceWidth"] = float(rng.uniform(0.05, 0.4))
    if rng.random() < 0.5:
        doc["minClearance"] = float(rng.uniform(0.0, 0.3))
    if rng.random() < 0.3:
        n = int(rng.integers(3, 8))
        doc["outline"] = [{"x": side / 2 + side * 0.45 * math.cos(2 * math.pi * k / n),
                           "y": side / 2 + side * 0.45 * math.sin(2 * math.pi * k / n)} for k in range(n)]
    elif rng.random() < 0.2:
        doc["minBoardEdgeClearance"] = float(rng.uniform(0.0, 0.5))
    return doc


def test_random_boards_match_oracle():
    rng = np.random.default_rng(11)
    for case in range(40):
        doc = random_doc(rng, case)
        side = doc["bounds"]["maxX"]
        res = float(rng.choice([side / 64.0, side / 16.0, 0.5]))
        cells = int(rng.integers(0, 3))
        mc = float(rng.choice([0.0, 0.1, 0.15]))
        vp = float(rng.choice([0.0, 0.45]))
        exact = bool(rng.random() < 0.5)
        p = with_layers(doc, res, doc["layerCount"], cells, mc, vp, model=2 if exact else 1)
        same_as_oracle(p, oracle_raster(doc, res, doc["layerCount"], cells, mc, vp, exact), case)


def test_route_mode_is_the_cli_rule_selection():
    """route_problem_impl (crates/mr-cli/src/lib.rs:751-834): the legacy
    branch is rasterize_with_layers at ceil(minClearance / resolution) cells
    (default 0.15 mm) and a 0.45 mm via pad, with the CLI's 0.15 mm trace
    width when a board-edge contract is active (:688-704); a coherent typed
    profile goes through rasterize_with_uniform_physical_rules."""
    for path in SMALL + [os.path.join(HERE, "golden", "srj", "bugreport21-board-outline.srj.json")]:
        doc = srjref.load(path)
        got = mrfield.srj_rasterize(doc, mode=SRJ_ROUTE)
        if got.physical:
            same_problem(got, mrfield.srj_rasterize(doc, mode=SRJ_PHYSICAL))
            continue
        b = doc["bounds"]
        res = max(b["maxX"] - b["minX"], b["maxY"] - b["minY"]) / 64.0
        mc = doc.get("minClearance")
        mc = 0.15 if mc is None else max(mc, 0.0)
        raster_doc = doc
        if (doc.get("outline") or doc.get("minBoardEdgeClearance") is not None) and doc.get("minTraceWidth") != 0.15:
            raster_doc = dict(doc, minTraceWidth=0.15)
        want = with_layers(raster_doc, res, doc["layerCount"], math.ceil(mc / res) if mc > 0 else 0, mc, 0.45)
        same_problem(got, want)
        assert got.clearance_cells == want.clearance_cells
    doc = typed(obstacles=[pad(0.5, 0.5, ["n"]), pad(3.5, 3.5, ["n"])],
                connections=[{"name": "n", "pointsToConnect": [{"x": 0.5, "y": 0.5}, {"x": 3.5, "y": 3.5}]}])
    got = mrfield.srj_rasterize(doc, mode=SRJ_ROUTE)
    assert got.physical and got.exact_model and got.rules["trace_to_pad_clearance_mm"] == 0.07
    same_problem(got, mrfield.srj_rasterize(doc, mode=SRJ_PHYSICAL))


def test_malformed_json_fails_closed():
    for bad in ('{"layerCount": 1}', '{"layerCount": 1, "bounds": {"minX": 0}}', "[1, 2", "",
                '{"layerCount": 1, "bounds": {"minX": 0, "maxX": 1, "minY": 0, "maxY": 1}, "obstacles": 3}'):
        with pytest.raises(mrfield.MrfError):
            mrfield.srj_rasterize(bad, mode=SRJ_LAYERS, resolution=1.0)
