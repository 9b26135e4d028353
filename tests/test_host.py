"""CPU tests of the product's host side: the C-ABI library, the synthetic
workload generator and the batcher.  No device work."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(gr):
    hdr = open(os.path.join(ROOT, "include", "iitmgr.h")).read()
    names = sorted(set(re.findall(r"\b(gr_[a-z0-9_]+)\s*\(", hdr)))
    assert len(names) >= 20
    lib = ctypes.CDLL(gr.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_no_device_fails_loudly(gr):
    if gr.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(gr.RouterError):
        gr.Router(8, 8, 4, np.ones(8 * 8 * 4, np.int32))


def _dec(X, Y, v):
    z = v // (X * Y)
    r = v - z * X * Y
    odd = (z & 1) == 1
    x = np.where(odd, r // Y, r % X)
    y = np.where(odd, r % Y, r // X)
    return z, x, y


@pytest.mark.parametrize("kind", [1, 2, 3, 5])
def test_generator_shape_and_determinism(gr, kind):
    X, Y, Z = (128, 128, 6) if kind == 1 else (300, 260, 10)
    n = 4000
    pfx, pins = gr.gen_nets(kind, 11, X, Y, Z, n)
    pfx2, pins2 = gr.gen_nets(kind, 11, X, Y, Z, n)
    assert np.array_equal(pfx, pfx2) and np.array_equal(pins, pins2)
    pfx3, _ = gr.gen_nets(kind, 12, X, Y, Z, n)
    assert not np.array_equal(pfx, pfx3)
    deg = np.diff(pfx)
    assert deg.min() >= 2 and len(deg) <= n
    assert pins.min() >= 0 and pins.max() < X * Y * Z
    for i in range(0, len(deg), 97):
        p = pins[pfx[i]:pfx[i + 1]]
        assert np.all(np.diff(p) > 0)  # ascending, deduplicated
    z, _, _ = _dec(X, Y, pins.astype(np.int64))
    assert set(np.unique(z)) <= {0, 1}
    if kind == 1:
        # pin mix before dedupe: 60% 2, 30% 3-5, 10% 6-16 -> dedupe only lowers counts
        assert deg.max() <= 16
        assert 0.55 < np.mean(deg == 2) < 0.75


@pytest.mark.parametrize("kind", [1, 2, 5])
def test_grid_capacities(gr, kind):
    X, Y, Z = (128, 128, 6) if kind == 1 else (400, 300, 10)
    cap = gr.gen_grid(kind, 0, X, Y, Z)
    assert np.array_equal(cap, gr.gen_grid(kind, 0, X, Y, Z))
    if kind == 1:
        assert cap.min() == 2 and cap.max() == 8
    elif kind == 2:
        assert cap.max() == 16
        assert 0.045 < np.mean(cap == 0) < 0.055
        assert cap[cap > 0].min() == 4
    else:
        assert cap.max() <= 9  # floor(16 * 0.6)


def test_batches_are_disjoint_partitions(gr):
    X, Y, Z = 128, 128, 6
    pfx, pins = gr.gen_nets(1, 0, X, Y, Z, 5000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    n = len(pfx) - 1
    assert ofst[0] == 0 and ofst[-1] == n
    assert np.array_equal(np.sort(ids), np.arange(n))
    deg = np.diff(pfx)
    z, x, y = _dec(X, Y, pins.astype(np.int64))
    net = np.repeat(np.arange(n), deg)
    minx = np.full(n, 1 << 30); maxx = np.full(n, -1); miny = np.full(n, 1 << 30); maxy = np.full(n, -1)
    np.minimum.at(minx, net, x); np.maximum.at(maxx, net, x)
    np.minimum.at(miny, net, y); np.maximum.at(maxy, net, y)
    for b in range(len(ofst) - 1):
        bt = ids[ofst[b]:ofst[b + 1]]
        assert len(bt) > 0
        # in-batch order: (pin count, net id)
        key = list(zip(deg[bt], bt))
        assert key == sorted(key)
        # pairwise disjoint closed boxes
        ax0, ax1, ay0, ay1 = minx[bt], maxx[bt], miny[bt], maxy[bt]
        ov = (ax0[:, None] <= ax1[None, :]) & (ax0[None, :] <= ax1[:, None]) & \
             (ay0[:, None] <= ay1[None, :]) & (ay0[None, :] <= ay1[:, None])
        np.fill_diagonal(ov, False)
        assert not ov.any(), b


def test_batch_count_workload2(gr):
    X, Y, Z = 800, 800, 10
    pfx, pins = gr.gen_nets(2, 0, X, Y, Z, 150000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    assert len(pfx) - 1 == 146516
    assert len(ofst) - 1 == 63
