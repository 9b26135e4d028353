"""The non-BIT ("prefix") mode of the oracle: prefix-sum L-pattern costing
(l_pattern_files/lpattern.hpp:13-160, util.hpp:133-353) and the difference-
array commit (mst_files/mst_commit.hpp:17-364), pinned by

  wire prefix        a step-by-step Python transliteration of the
                     reference's shared-memory scan kernel
                     (d_get_epfx_cost_cube_inplace, util.hpp:277-353), bit
                     for bit; unit costs exact
  via cost           the contract: the increase of e^(-r/2) over exactly the
                     capacity the commit's via charge takes, in double
  pattern cost       an explicit double-precision cell walk of every pattern
                     of random pairs (odd and even Z, straight, stacked and
                     border pairs), validity equal
  commit             hand-worked known answers (interior, first and last run
                     positions, a straight pair routed through b'), and the
                     remaining capacity after routed batches equal to the
                     emitted edges' wires and via charges; the per-run
                     refresh equal to a full rebuild
"""
import math

import numpy as np
import pytest

PREFIX = 2  # commit mode of the non-BIT path


def _enc(X, Y, z, x, y):
    return z * X * Y + (x * Y + y if z & 1 else y * X + x)


def _dec(X, Y, v):
    z = v // (X * Y)
    r = v - z * X * Y
    return (z, r // Y, r % Y) if z & 1 else (z, r % X, r // X)


def _f(r):
    return math.exp(-0.5 * r)


def _scan_kernel(w, M):
    """d_get_epfx_cost_cube_inplace for one run of len(w) <= M cells, nth = 128
    scanning threads, shared array of M floats, executed step by step (one
    pass: M >= run length).  Cells past the run hold arbitrary values."""
    f32 = np.float32
    nth = 128
    T = -(-M // nth)
    L = len(w)
    sm = [f32(v) for v in np.random.default_rng(L).normal(0, 1e3, M)]
    for i in range(L):
        sm[i] = f32(w[i])
    start = [f32(0) if t == 0 else sm[max(0, min(t * T - 1, M - 1))] for t in range(nth)]
    for t in range(nth):
        pfx = start[t]
        for i in range(t * T, min(t * T + T, M)):
            prev = sm[i]
            sm[i] = pfx
            pfx = f32(pfx + prev)
    pfx = sm[T - 1]
    i = 2 * T - 1
    while i < M - 1:
        pfx = f32(pfx + sm[i])
        sm[i] = pfx
        i += T
    if T - 1 != M - 1:
        pfx = f32(pfx + sm[M - 1])
        sm[M - 1] = pfx
    carry = [sm[min(t * T - 1, M - 1)] if t else None for t in range(nth)]
    for t in range(1, nth):
        for i in range(t * T, min(t * T + T - 1, M - 1)):
            sm[i] = f32(sm[i] + carry[t])
    return np.array(sm[:L], np.float32)


def _run(X, Y, z, line):
    L = Y if z & 1 else X
    base = z * X * Y + line * L
    return base, L


@pytest.mark.parametrize("X,Y,Z", [(300, 40, 3), (1000, 17, 4), (129, 130, 3)])
def test_wire_prefix_equals_reference_scan_steps(orc, X, Y, Z):
    rng = np.random.default_rng(X)
    cap = rng.integers(-3, 12, X * Y * Z).astype(np.int32)
    o = orc.Oracle(X, Y, Z, cap, 1, PREFIX)
    P = o.prefix()
    M = max(X, Y)
    for z in range(Z):
        for line in rng.choice((X if z & 1 else Y), size=4, replace=False):
            base, L = _run(X, Y, z, int(line))
            w = [orc.inc_cost(cap[base + i], 1.0) for i in range(L)]
            want = _scan_kernel(w, M)
            assert np.array_equal(P[base:base + L].view(np.int32), want.view(np.int32)), (z, line)


def test_unit_wire_cost_prefix_is_exact(orc):
    """Every cell with r = 1 costs f(0) - f(1) alike; the chunked fp32 sums
    of equal values are then exact multiples up to 2^24 / value."""
    X, Y, Z = 700, 300, 3
    o = orc.Oracle(X, Y, Z, np.full(X * Y * Z, 1, np.int32), 1, PREFIX)
    w = orc.inc_cost(1.0, 1.0)
    P = o.prefix()
    for z in range(Z):
        L = Y if z & 1 else X
        base = z * X * Y
        got = P[base:base + L].astype(np.float64)
        want = np.arange(L) * float(w)
        assert np.allclose(got, want, rtol=2e-6, atol=0), z


def test_via_cost_is_the_increase_the_commit_charges(orc):
    """V at (i, x, y) equals, in double, the growth of e^(-r/2) over the
    capacity a one-layer via stack there takes in the commit (the halves the
    commit charges): the contract that replaces the racing via cube."""
    X, Y, Z = 9, 7, 4
    rng = np.random.default_rng(4)
    cap = rng.integers(-2, 9, X * Y * Z).astype(np.int32)
    o = orc.Oracle(X, Y, Z, cap, 1, PREFIX)
    V = o.via_terms()
    for i in range(Z - 1):
        for x in range(X):
            for y in range(Y):
                # a pair stacked at (x, y) from layer i to layer i + 1 commits
                # one via stack over layer i there (pattern a = i + 1 needs a
                # >= 1: use src layer i, dst layer i + 1 and read the halves of
                # the first stack only)
                c = _enc(X, Y, i, x, y)
                _, halves = o.edge_commit_prefix(_enc(X, Y, i + 1, x, y), _enc(X, Y, i, x, y), (i + 1) * ((Z + 1) // 2))
                first = [(cell, h) for cell, h in halves if _dec(X, Y, cell)[0] == i]
                want = sum(_f(cap[cell] - 0.5 * h) - _f(cap[cell]) for cell, h in first)
                assert abs(float(V[c]) - want) <= 4e-7 * abs(want) + 1e-30, (i, x, y, V[c], want)


def _walk(X, Y, Z, cap, src, dst, p):
    """Cost of pattern p in double by walking its cells per the non-BIT
    costing's route: first via stack over [min, max) layers, corner and final
    stacks over the open interval, wires over [min, max) run cells."""
    Z2 = (Z + 1) // 2
    z1, x1, y1 = _dec(X, Y, src)
    z2, x2, y2 = _dec(X, Y, dst)
    a, b = divmod(p, Z2)
    if a == 0:
        return None

    def w(z, x, y):
        r = float(cap[_enc(X, Y, z, x, y)])
        return _f(r - 1) - _f(r)

    def v(z, x, y):
        L = Y if z & 1 else X
        pos = y if z & 1 else x
        r = float(cap[_enc(X, Y, z, x, y)])
        if pos == 0:
            return _f(r - 1) - _f(r)
        prev = (x, y - 1) if z & 1 else (x - 1, y)
        rp = float(cap[_enc(X, Y, z, *prev)])
        if pos == L - 1:
            return _f(rp - 1) - _f(rp)
        return (_f(rp - 0.5) - _f(rp)) + (_f(r - 0.5) - _f(r))

    s = 0.0
    zs, xs, ys = z1, x1, y1

    def wire(z):
        nonlocal xs, ys, s
        if z & 1:
            for t in range(min(ys, y2), max(ys, y2)):
                s += w(z, xs, t)
            ys = y2
        else:
            for t in range(min(xs, x2), max(xs, x2)):
                s += w(z, t, ys)
            xs = x2

    for i in range(min(zs, a), max(zs, a)):
        s += v(i, xs, ys)
    zs = a
    wire(zs)
    bp = 2 * b if zs & 1 else 2 * b + 1
    if bp >= Z:
        return None
    if (zs, xs, ys) != (z2, x2, y2):
        if bp > 0:
            for i in range(min(zs, bp) + 1, max(zs, bp)):
                s += v(i, xs, ys)
            zs = bp
            wire(zs)
        if zs != z2:
            for i in range(min(zs, z2) + 1, max(zs, z2)):
                s += v(i, xs, ys)
            zs = z2
    if (zs, xs, ys) != (z2, x2, y2):
        return None
    return s


@pytest.mark.parametrize("X,Y,Z", [(23, 19, 4), (31, 12, 7), (16, 29, 10), (600, 9, 5)])
def test_pattern_cost_equals_cell_walk(orc, X, Y, Z):
    rng = np.random.default_rng(Z)
    cap = rng.integers(0, 9, X * Y * Z).astype(np.int32)
    o = orc.Oracle(X, Y, Z, cap, 1, PREFIX)
    P = np.abs(o.prefix().astype(np.float64))
    Z2 = (Z + 1) // 2
    pairs = []
    for _ in range(60):
        pairs.append((int(rng.integers(0, X * Y * Z)), int(rng.integers(0, X * Y * Z))))
    # straight along x, along y, stacked, at the borders
    pairs += [(_enc(X, Y, 0, 0, 3), _enc(X, Y, 1, X - 1, 3)), (_enc(X, Y, 1, 2, 0), _enc(X, Y, 0, 2, Y - 1)),
              (_enc(X, Y, Z - 1, 5, 5), _enc(X, Y, 0, 5, 5)), (_enc(X, Y, 0, X - 1, Y - 1), _enc(X, Y, 2, 0, 0))]
    n_valid = 0
    for src, dst in pairs:
        if src == dst:
            continue
        for p in range(Z * Z2):
            got = o.pattern_cost(src, dst, p)
            want = _walk(X, Y, Z, cap, src, dst, p)
            assert (got is None) == (want is None), (src, dst, p, got, want)
            if want is None:
                continue
            n_valid += 1
            tol = 3e-5 * (P.max() + 1) * 1e-2 + 2e-6 * want
            assert abs(got - want) <= tol, (src, dst, p, got, want)
    assert n_valid > 100


def test_selection_is_lowest_p_minimum(orc):
    """Uniform capacities: many patterns tie; the choice is the minimum cost
    with the lowest p (lpattern.hpp:155-157, strict <)."""
    X, Y, Z = 14, 11, 6
    o = orc.Oracle(X, Y, Z, np.full(X * Y * Z, 3, np.int32), 1, PREFIX)
    rng = np.random.default_rng(2)
    Z2 = (Z + 1) // 2
    for _ in range(200):
        s, d = (int(v) for v in rng.integers(0, X * Y * Z, 2))
        if s == d:
            continue
        p, c = o.best_pattern(s, d)
        costs = [(o.pattern_cost(s, d, q), q) for q in range(Z * Z2)]
        costs = [(np.float32(cc), q) for cc, q in costs if cc is not None]
        assert (np.float32(c), p) == min(costs)


def test_commit_known_answers(orc):
    """Worked by hand from mst_commit.hpp:17-251 on a 6x5x4 grid.
    Edge 1: src (0,4,3), dst (0,1,1), a = 1, b = 1 (b' = 2).  Wires: layer 1
    at x = 4, y in [1, 3); layer 2 at y = 1, x in [1, 4).  Vias: 0->1 at
    (4,3) (layer 0, x = 4 interior: halves at x = 3 and 4), 1->2 at (4,1)
    (layer 1, y = 1 interior: y = 0 and 1), 2->0 at (1,1) (layers 0 and 1,
    both interior).
    Edge 2: src (0,0,2), dst (0,5,2), straight along x, a = 2, b = 0 (b' = 1,
    taken although the wire on 2 reaches x = 5).  Vias: 0->2 at (0,2): layer
    0 at x = 0 (first position: 2 halves to the cell), layer 1 at y = 2
    (interior); wire on 2, x in [0, 5); zero-length wire on 1; via 2->1 at
    (5,2) over layer 1 (interior); via 1->0 at (5,2) over layer 0 (x = 5 =
    last position: 2 halves to x = 4)."""
    X, Y, Z = 6, 5, 4
    o = orc.Oracle(X, Y, Z, np.full(X * Y * Z, 4, np.int32), 1, PREFIX)
    e = lambda z, x, y: _enc(X, Y, z, x, y)  # noqa: E731
    w, h = o.edge_commit_prefix(e(0, 4, 3), e(0, 1, 1), 1 * 2 + 1)
    assert sorted(w.tolist()) == sorted([e(1, 4, 1), e(1, 4, 2), e(2, 1, 1), e(2, 2, 1), e(2, 3, 1)])
    assert sorted(h) == sorted([(e(0, 3, 3), 1), (e(0, 4, 3), 1), (e(1, 4, 0), 1), (e(1, 4, 1), 1),
                                (e(0, 0, 1), 1), (e(0, 1, 1), 1), (e(1, 1, 0), 1), (e(1, 1, 1), 1)])
    w, h = o.edge_commit_prefix(e(0, 0, 2), e(0, 5, 2), 2 * 2 + 0)
    assert sorted(w.tolist()) == sorted([e(2, x, 2) for x in range(5)])
    assert sorted(h) == sorted([(e(0, 0, 2), 2), (e(1, 0, 1), 1), (e(1, 0, 2), 1), (e(1, 5, 1), 1), (e(1, 5, 2), 1),
                                (e(0, 4, 2), 2)])


def test_batches_commit_their_edges_and_refresh_equals_rebuild(orc, gr):
    """After each batch the remaining capacity fell by 1.0 on every cell some
    emitted edge's wire covers and by 0.5 per via half charged; the cost state
    refreshed run by run equals a rebuild from the remaining capacities."""
    X, Y, Z = 64, 48, 6
    cap = gr.gen_grid(1, 3, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 3, X, Y, Z, 700)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    o = orc.Oracle(X, Y, Z, cap, 1, PREFIX)
    o.set_nets(pfx, pins)
    prev = o.remaining_f32().astype(np.float64)
    for b in range(4):
        batch = ids[ofst[b]:ofst[b + 1]]
        e_, l_, c_ = o.route_batch(batch)
        wires, halves = set(), {}
        pos = 0
        for i, nid in enumerate(batch):
            d = pfx[nid + 1] - pfx[nid]
            for t in range(d - 1):
                v = int(e_[pos + t])
                hi, lo = max((v >> 15) & 0x7FFF, v & 0x7FFF), min((v >> 15) & 0x7FFF, v & 0x7FFF)
                w, hv = o.edge_commit_prefix(pins[pfx[nid] + hi], pins[pfx[nid] + lo], l_[pos + t])
                wires.update(w.tolist())
                for cell, n in hv:
                    halves[cell] = halves.get(cell, 0) + n
            pos += d
        want = prev.copy()
        for cell in wires:
            want[cell] -= 1.0
        for cell, n in halves.items():
            want[cell] -= 0.5 * n
        cur = o.remaining_f32().astype(np.float64)
        assert np.array_equal(cur, want), b
        prev = cur
    fresh = orc.Oracle(X, Y, Z, o.remaining_f32(), 1, PREFIX)
    assert np.array_equal(fresh.prefix().view(np.int32), o.prefix().view(np.int32))
    assert np.array_equal(fresh.via_terms().view(np.int32), o.via_terms().view(np.int32))
