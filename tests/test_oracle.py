"""Pinning of the CPU oracle (test infrastructure) against known answers,
closed forms, brute force and invariants.

Pinning map (function -> checks):
  cost LUT (bit_util.hpp:13-18)            glibc expf within 1 ulp; closed forms of the 3 models
  Fenwick prefix Q (bit_util.hpp:19-38)    exact on unit cost; naive double prefix within fp32 bound
  pattern cost (bit_util.hpp:131-179)      hand KAT; brute-force cell walk in double (every pattern,
                                           random pairs, odd and even Z)
  pattern selection                        exhaustive min over p with lowest-p ties
  root-pair Boruvka (gmst_util.hpp)        hand KAT; spanning tree; tree cost = true MST weight;
                                           independent Python restatement on random small nets
  commit cells (mst_bit_commit.hpp)        hand KAT (with and without the quirk); route/commit consistency
"""
import json
import math
import os
import random

import numpy as np
import pytest

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_hot_path.json")))


def _ulp(x):
    return np.spacing(np.float32(abs(x)))


def test_lut_matches_expf(orc):
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.expf.restype = ctypes.c_float
    libm.expf.argtypes = [ctypes.c_float]
    lut = orc.lut(0)
    rmin = orc.lib().orc_lut_rmin()
    for i, v in enumerate(lut):
        r = rmin + i
        if r < -138:  # saturated region: exp(-r/2) > 1e30
            assert v == np.float32(1e30)
            continue
        ref = np.float32(libm.expf(ctypes.c_float(-0.5 * r)))
        assert abs(float(v) - float(ref)) <= float(_ulp(ref)), (r, v, ref)
    assert np.all(np.diff(lut.astype(np.float64)) <= 0)  # cost falls as remaining capacity grows


def test_lut_models_closed_form(orc):
    rmin = orc.lib().orc_lut_rmin()
    inc, lit = orc.lut(1), orc.lut(2)
    for i in range(len(inc)):
        r = rmin + i
        e0 = np.float32(math.exp(-0.5 * (r - 1))) if -0.5 * (r - 1) < 88 else np.float32(np.inf)
        e1 = np.float32(math.exp(-0.5 * r)) if -0.5 * r < 88 else np.float32(np.inf)
        want = np.float32(e0 - e1)
        if not np.isfinite(want) or want >= 1e30:
            assert inc[i] == np.float32(1e30)
        else:
            assert inc[i] == want
        w2 = math.exp(1.0 - r) if 1.0 - r < 700 else math.inf
        assert lit[i] == (np.float32(1e30) if w2 >= 1e30 else np.float32(w2))


def test_prefix_unit_cost_exact(orc):
    X, Y, Z = 37, 29, 5
    cap = np.zeros(X * Y * Z, np.int32)  # r = 0 -> cost 1.0 exactly
    o = orc.Oracle(X, Y, Z, cap)
    q = o.prefix()
    for z in range(Z):
        L, lines = (Y, X) if z & 1 else (X, Y)
        blk = q[z * X * Y:(z + 1) * X * Y].reshape(lines, L)
        assert np.array_equal(blk, np.tile(np.arange(1, L + 1, dtype=np.float32), (lines, 1)))


def test_prefix_vs_naive(orc):
    rng = np.random.default_rng(1)
    X, Y, Z = 200, 123, 4
    cap = rng.integers(-20, 30, X * Y * Z).astype(np.int32)
    o = orc.Oracle(X, Y, Z, cap)
    lut = orc.lut(0).astype(np.float64)
    c = lut[np.clip(cap, -256, 255) + 256]
    q = o.prefix().astype(np.float64)
    for z in range(Z):
        L, lines = (Y, X) if z & 1 else (X, Y)
        cc = c[z * X * Y:(z + 1) * X * Y].reshape(lines, L)
        naive = np.cumsum(cc, axis=1)
        got = q[z * X * Y:(z + 1) * X * Y].reshape(lines, L)
        bound = naive * (np.log2(L) + 2) * 2.0 ** -23
        assert np.all(np.abs(got - naive) <= bound)


def _enc(X, Y, z, x, y):
    return z * X * Y + (x * Y + y if z & 1 else y * X + x)


def test_pattern_cost_kat(orc):
    kat = GOLD["pattern_cost"]
    X, Y, Z = kat["grid"]
    o = orc.Oracle(X, Y, Z, np.zeros(X * Y * Z, np.int32))
    src, dst = _enc(X, Y, *kat["src"]), _enc(X, Y, *kat["dst"])
    Z2 = (Z + 1) // 2
    for p in range(Z * Z2):
        c = o.pattern_cost(src, dst, p)
        w = o.pattern_cost_walk(src, dst, p)
        if p in kat["invalid"]:
            assert c is None and w is None
        else:
            assert c == kat["costs"][str(p)] and w == kat["costs"][str(p)]
    assert o.best_pattern(src, dst) == (kat["argmin"], 9.0)


@pytest.mark.parametrize("Z", [4, 6, 7, 10])
def test_pattern_cost_vs_cell_walk(orc, Z):
    """Fenwick evaluation vs an explicit double-precision walk of every cell."""
    rng = np.random.default_rng(Z)
    X, Y = 40, 33
    cap = rng.integers(2, 12, X * Y * Z).astype(np.int32)
    o = orc.Oracle(X, Y, Z, cap)
    q = np.abs(o.prefix()).max()
    Z2 = (Z + 1) // 2
    for _ in range(150):
        z1, z2 = rng.integers(0, Z, 2)
        x1, x2 = rng.integers(0, X, 2)
        y1, y2 = rng.integers(0, Y, 2)
        if (z1, x1, y1) == (z2, x2, y2):
            continue
        s, d = _enc(X, Y, z1, x1, y1), _enc(X, Y, z2, x2, y2)
        best = None
        for p in range(Z * Z2):
            c = o.pattern_cost(s, d, p)
            w = o.pattern_cost_walk(s, d, p)
            assert (c is None) == (w is None), p
            if c is None:
                continue
            # each term is a difference of two prefix values: <= 2 ulp(max Q) per term, ~3Z+2 terms
            tol = (3 * Z + 2) * 3 * float(_ulp(q)) + 1e-6 * w
            assert abs(c - w) <= tol, (p, c, w)
            if best is None or c < best[1]:
                best = (p, c)
        assert best is not None
        assert o.best_pattern(s, d) == (best[0], np.float32(best[1]))


def test_no_pair_without_valid_pattern(orc):
    for Z in (3, 4, 5, 10):
        X = Y = 5
        o = orc.Oracle(X, Y, Z, np.full(X * Y * Z, 3, np.int32))
        for s in range(X * Y * Z):
            for d in range(0, X * Y * Z, 7):
                if s != d:
                    assert o.best_pattern(s, d)[0] >= 0


def test_boruvka_kat(orc):
    kat = GOLD["boruvka"]
    edges, cost, rounds = orc.boruvka(kat["n"], kat["pair_costs_by_o"])
    assert [list(e) for e in edges] == kat["edges_on1_on2"]
    assert sorted([max(e), min(e)] for e in edges) == sorted(kat["edges_hi_lo"])
    assert cost == kat["net_cost"]
    assert rounds == 2


def _py_root_pair_boruvka(n, costs):
    """Independent restatement of the root-pair Boruvka contract (lowest (cost, o) per
    component, candidates in ascending root order, smaller root hooked under larger)."""
    parent = list(range(n))

    def find(v):
        while parent[v] != v:
            v = parent[v]
        return v
    pairs = [(j, k) for j in range(1, n) for k in range(j)]
    out, total = [], np.float32(0)
    while len(out) < n - 1:
        best = {}
        for o, (j, k) in enumerate(pairs):
            rj, rk = find(j), find(k)
            if rj == rk:
                continue
            key = (float(costs[o]), o)
            for r in (rj, rk):
                if r not in best or key < best[r]:
                    best[r] = key
        partner = {}
        for r, (_, o) in best.items():
            j, k = pairs[o]
            partner[r] = find(k) if find(j) == r else find(j)
        for r in sorted(partner):
            a, b = find(r), find(partner[r])
            if a == b:
                continue
            parent[min(a, b)] = max(a, b)
            total = np.float32(total + np.float32(costs[best[r][1]]))
            out.append((r, partner[r]))
    return out, total


def test_boruvka_random_vs_restatement_and_mst(orc):
    from scipy.sparse.csgraph import minimum_spanning_tree
    rng = np.random.default_rng(5)
    for trial in range(300):
        n = int(rng.integers(2, 24))
        P = n * (n - 1) // 2
        if trial % 3 == 0:
            costs = rng.integers(0, 4, P).astype(np.float32)  # many ties
        else:
            costs = rng.random(P).astype(np.float32) + 0.01
        edges, cost, _ = orc.boruvka(n, costs)
        want, wtot = _py_root_pair_boruvka(n, costs)
        assert edges == want
        assert np.float32(cost) == wtot
        # spanning tree over the recorded root pairs
        par = list(range(n))

        def f(v):
            while par[v] != v:
                v = par[v]
            return v
        for a, b in edges:
            ra, rb = f(a), f(b)
            assert ra != rb
            par[ra] = rb
        if trial % 3 != 0:
            m = np.zeros((n, n))
            o = 0
            for j in range(1, n):
                for k in range(j):
                    m[j, k] = costs[o]
                    o += 1
            assert abs(minimum_spanning_tree(m).sum() - cost) <= 1e-4 * max(1.0, cost)


@pytest.mark.parametrize("quirk", [True, False])
def test_commit_cells_kat(orc, quirk):
    kat = GOLD["commit_cells"]
    X, Y, Z = kat["grid"]
    o = orc.Oracle(X, Y, Z, np.full(X * Y * Z, 5, np.int32), 0, 0 if quirk else 1)
    cells = o.edge_cells(_enc(X, Y, *kat["src"]), _enc(X, Y, *kat["dst"]), kat["pattern"])
    want = [_enc(X, Y, *c) for c in kat["cells"]]
    if quirk:
        want += [_enc(X, Y, *c) for c in kat["quirk_extra"]]
    assert sorted(cells.tolist()) == sorted(want)


def test_route_batch_commit_consistency(orc, gr):
    """Demand after a batch = union of the A.7 cells of the emitted edges; each cell +1 at most."""
    X, Y, Z = 64, 48, 6
    cap = gr.gen_grid(1, 3, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 3, X, Y, Z, 600)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    prev = o.demand()
    for b in range(3):
        batch = ids[ofst[b]:ofst[b + 1]]
        e, l, c = o.route_batch(batch)
        cells = set()
        pos = 0
        for i, nid in enumerate(batch):
            d = pfx[nid + 1] - pfx[nid]
            slot = e[pos:pos + d]
            assert slot[-1] == -1 and l[pos + d - 1] == -1
            for t in range(d - 1):
                v = int(slot[t])
                assert v >> 30 == i
                on1, on2 = (v >> 15) & 0x7FFF, v & 0x7FFF
                hi, lo = max(on1, on2), min(on1, on2)
                cells.update(o.edge_cells(pins[pfx[nid] + hi], pins[pfx[nid] + lo], l[pos + t]).tolist())
            pos += d
        cur = o.demand()
        delta = cur - prev
        assert set(np.nonzero(delta)[0].tolist()) == cells
        assert delta.max() <= 1 and delta.min() >= 0
        prev = cur


def test_quirk_cells_below_the_grid_are_skipped(orc):
    """With X*Y < Z the quirk's cells before the first wire can fall below
    enc 0 (the reference writes before its flag array there); they are
    skipped.  1x1x6 grid, src layer 5 -> dst layer 0, pattern a = 1: interior
    vias 2..4, the wire cell 1, then quirk cells 1-1 = 0 (kept), -1, -2
    (skipped)."""
    X, Y, Z = 1, 1, 6
    o = orc.Oracle(X, Y, Z, np.full(X * Y * Z, 4, np.int32), 0, 0)
    cells = o.edge_cells(5, 0, 1 * 3 + 0)
    assert sorted(cells.tolist()) == [0, 1, 2, 3, 4]
    o_fixed = orc.Oracle(X, Y, Z, np.full(X * Y * Z, 4, np.int32), 0, 1)
    assert sorted(o_fixed.edge_cells(5, 0, 3).tolist()) == [1, 2, 3, 4]
    # routed: every cell of the stack gets demand once, nothing outside the grid
    o.set_nets(np.array([0, 2], np.int64), np.array([0, 5], np.int32))
    o.route_batch(np.array([0], np.int32))
    assert o.demand().tolist() == [1, 1, 1, 1, 1, 0]
