"""CPU-only checks of the product library: it loads, exports every symbol the
C header declares, fails cleanly without a device, and its synthetic
workload generators produce what their documentation states."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import mrfield

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mrfield.h")).read()
    return sorted(set(re.findall(r"\b(mrf_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = mrfield.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert L.mrf_abi_version() == 3
    assert len(declared_symbols()) >= 20


def test_open_without_device_fails_closed():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(mrfield.MrfError) as e:
        mrfield.Engine(0)
    assert e.value.code == mrfield.BACKEND_UNAVAILABLE


def test_server_without_device_fails_closed():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    with pytest.raises(mrfield.MrfError) as e:
        mrfield.Server(0)
    assert e.value.code == mrfield.BACKEND_UNAVAILABLE


def _blocked_fraction(s):
    return float((s.grid.cost == mrfield.MAX).mean())


def test_synth_config1_shape_and_density():
    s = mrfield.synth_config(1, seed=1)
    g = s.grid
    assert (g.w, g.h, g.layers) == (128, 128, 2)
    assert len(s.src) == 64
    assert 0.15 <= _blocked_fraction(s) < 0.17
    assert s.edges.x == [10] * 127 and s.edges.via == [50]
    assert all(g.cost[c] == 1 for c in list(s.src) + list(s.dst))
    assert s.windows is None and s.additive is None


def test_synth_config2_outline_and_history():
    s = mrfield.synth_config(2, seed=1, scale_nets=32)
    g = s.grid
    assert (g.w, g.h, g.layers) == (512, 512, 4)
    assert 0.20 <= _blocked_fraction(s) < 0.23
    add = s.additive
    assert add.min() >= 1 and add.max() <= 64
    # log-uniform: median near sqrt(64) = 8
    assert 6 <= np.median(add) <= 10
    b = g.board
    assert b is not None
    trace = (b & 1) != 0
    assert 0.1 < trace.mean() < 0.6          # outside of the 12-gon
    assert ((b & 0x3C) != 0).any()            # directed edge bits present
    assert ((b & 2) != 0).mean() > trace.mean()  # via keep-out is wider
    for c in list(s.src) + list(s.dst):
        assert g.cost[c] != mrfield.MAX and not (b[c] & 1)
    assert s.edges.via == [100, 100, 100]


def test_synth_config3_windows_contain_endpoints():
    s = mrfield.synth_config(3, seed=1, scale_nets=200)
    g = s.grid
    assert (g.w, g.h, g.layers) == (2048, 2048, 2)
    plane = g.w * g.h
    sides = []
    for (x0, y0, x1, y1), a, b in zip(s.windows, s.src, s.dst):
        assert x1 - x0 == y1 - y0
        sides.append(x1 - x0 + 1)
        for c in (a, b):
            r = int(c) % plane
            assert x0 <= r % g.w <= x1 and y0 <= r // g.w <= y1
    assert 16 <= min(sides) and max(sides) <= 512
    assert 50 < np.median(sides) < 140       # log-uniform median ~ 90.5
    assert s.additive.min() >= 1 and s.additive.max() <= 32


def test_synth_config4_pins_and_halo():
    s = mrfield.synth_config(4, seed=1, scale_nets=64)
    g = s.grid
    assert (g.w, g.h, g.layers) == (1024, 1024, 6)
    assert all(2 <= len(p) <= 12 for p in s.pins)
    assert s.halo_penalty == 40
    h = s.halo_owner
    plane = g.w * g.h
    # every pin's 4-neighbours carry its owner id, or MIXED
    for i, pins in enumerate(s.pins[:16]):
        for p in pins:
            x, y = p % plane % g.w, p % plane // g.w
            if x + 1 < g.w:
                assert h[p + 1] in (i, -2)
    assert np.all(s.dst == mrfield.MAX)


def test_synth_reference_bench_board():
    """mr-metal/src/lib.rs:2761-2800 bench_board: walls leave row 0 open."""
    s = mrfield.synth_config(0)
    g = s.grid
    assert (g.w, g.h, g.layers) == (128, 128, 1)
    walls = [min(2 + k * 5, 126) for k in range(16)]
    for x in walls:
        assert g.cost[x] == 1                     # row 0 open
        assert g.cost[g.w * 63 + x] == mrfield.MAX
        assert g.cost[g.w * 64 + x] == 1
    assert len(s.src) == 64


# mr-core/src/lib.rs:856-891: congestion counts occupying routes, once per
# route and cell, whatever the path's shape.
CONGESTION_VECTORS = [
    ((3, 1, 1), [[0, 1], [1, 2]], [1, 2, 1]),
    ((4, 1, 1), [[0, 1, 2, 1, 0, 1], [1, 2, 3, 2]], [1, 2, 2, 1]),
]


@pytest.mark.parametrize("dims,paths,expect", CONGESTION_VECTORS)
def test_congestion_from_reference_vectors(dims, paths, expect):
    assert mrfield.congestion_from(dims, paths).tolist() == expect


def test_congestion_from_skips_unrouted_and_matches_the_oracle():
    import mrcpu

    rng = np.random.default_rng(3)
    dims = (7, 5, 2)
    n = 7 * 5 * 2
    for _ in range(50):
        paths = [None if rng.random() < 0.2 else [int(c) for c in rng.integers(0, n, size=int(rng.integers(1, 30)))]
                 for _ in range(int(rng.integers(0, 12)))]
        want = np.zeros(n, dtype=np.uint32)
        for p in paths:
            for c in set(p or []):
                want[c] += 1
        assert np.array_equal(mrfield.congestion_from(dims, paths), want)
        assert np.array_equal(mrcpu.congestion_from(dims, paths), want)
    with pytest.raises(mrfield.MrfError):
        mrfield.congestion_from((2, 1, 1), [[0, 5]])
