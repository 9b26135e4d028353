"""GPU parity on the cases random workloads reach rarely.

* A uniform-capacity grid: every cell costs the same, so pattern costs tie
  across patterns and pair costs tie across pairs (squares, lines), and the
  lowest-index tie rules of the L-pattern minimum and of Boruvka decide.
* Degenerate pairs: pure via stacks (all pins at one GCELL), straight lines
  along x or y, adjacent cells, pins on the grid border and on the bottom and
  top layers; in quirk mode the first wire's extra cells then lie next to the
  run ends.
* Empty batches, and nets of every tree class (2 pins, groups, workgroups).

Every output is compared bit for bit with the oracle after every batch.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _encode(X, Y, Z, nets):
    pfx, pins = [0], []
    for net in nets:
        e = sorted({z * X * Y + (x * Y + y if z & 1 else y * X + x) for z, x, y in net})
        if len(e) < 2:
            continue
        pins.extend(e)
        pfx.append(len(pins))
    return np.array(pfx, np.int64), np.array(pins, np.int32)


def _edge_case_nets(X, Y, Z, seed, n_rand):
    rng = np.random.default_rng(seed)
    ri = lambda lo, hi: int(rng.integers(lo, hi))  # noqa: E731
    nets = []
    # pure via stacks, at corners, border midpoints and inside
    for x, y in [(0, 0), (X - 1, Y - 1), (X - 1, 0), (0, Y - 1), (X // 2, 0), (0, Y // 2), (3, 5)]:
        zs = rng.choice(Z, size=ri(2, Z + 1), replace=False)
        nets.append([(int(z), x, y) for z in zs])
    nets.append([(0, 9, 9), (Z - 1, 9, 9)])  # bottom to top
    # straight lines along y (same x) and along x (same y), on the border too
    for x in [0, X - 1, X // 3]:
        nets.append([(ri(0, Z), x, int(y)) for y in rng.choice(Y, size=5, replace=False)])
    for y in [0, Y - 1, Y // 3]:
        nets.append([(ri(0, Z), int(x), y) for x in rng.choice(X, size=5, replace=False)])
    # adjacent cells, same layer and across layers
    nets += [[(0, 1, 1), (0, 2, 1)], [(1, 4, 4), (1, 4, 5)], [(Z - 1, 6, 7), (0, 7, 7)], [(2, 0, 0), (3, 1, 1)]]
    # squares and crosses on one layer: pair costs tie exactly on a uniform grid
    for s in (2, 3, 5):
        x0, y0, z = ri(0, X - s), ri(0, Y - s), ri(0, Z)
        nets.append([(z, x0, y0), (z, x0 + s, y0), (z, x0, y0 + s), (z, x0 + s, y0 + s)])
        nets.append([(z, x0 + s // 2, y0), (z, x0 + s // 2, y0 + s), (z, x0, y0 + s // 2), (z, x0 + s, y0 + s // 2)])
    # random nets in small boxes: 2..40 pins (2-pin lanes and wavefront groups)
    for _ in range(n_rand):
        w, h = ri(1, 12), ri(1, 12)
        x0, y0 = ri(0, X - w + 1), ri(0, Y - h + 1)
        nets.append([(ri(0, Z), x0 + ri(0, w), y0 + ri(0, h)) for _ in range(ri(2, 41))])
    # a few nets above the wavefront-group sizes (workgroup trees)
    for d in (70, 130):
        nets.append([(ri(0, Z), ri(0, 20), ri(0, 20)) for _ in range(d)])
    return _encode(X, Y, Z, nets)


def _run(gr, orc, X, Y, Z, cap, pfx, pins, cost_model, commit_mode, tag):
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap, cost_model, commit_mode)
    r.set_nets(pfx, pins)
    o = orc.Oracle(X, Y, Z, cap, cost_model, commit_mode)
    o.set_nets(pfx, pins)
    for b in range(len(ofst) - 1):
        batch = ids[ofst[b]:ofst[b + 1]]
        print(f"{tag} batch {b} nets {len(batch)}", flush=True)
        e1, l1, c1 = r.route_batch(batch)
        e2, l2, c2 = o.route_batch(batch)
        pc1, pp1 = r.pair_costs()
        pc2, pp2 = o.pair_costs()
        assert _same_f32(pc1, pc2), f"{tag} batch {b}: pair costs differ"
        assert np.array_equal(pp1, pp2), f"{tag} batch {b}: pair patterns differ"
        assert np.array_equal(e1, e2), f"{tag} batch {b}: tree edges differ"
        assert np.array_equal(l1, l2), f"{tag} batch {b}: edge patterns differ"
        assert _same_f32(c1, c2), f"{tag} batch {b}: net costs differ"
        assert np.array_equal(r.demand(), o.demand()), f"{tag} batch {b}: demand differs"
    assert _same_f32(r.prefix(), o.prefix()), f"{tag}: final prefix state differs"
    return r, o


@pytest.mark.parametrize("Z,cost_model,commit_mode", [(6, 0, 0), (5, 0, 0), (6, 2, 1), (7, 1, 0)])
def test_uniform_grid_ties_and_degenerate_pairs(gr, orc, gpu, Z, cost_model, commit_mode):
    X, Y = 40, 32
    cap = np.full(X * Y * Z, 3, np.int32)
    pfx, pins = _edge_case_nets(X, Y, Z, seed=Z, n_rand=300)
    deg = np.diff(pfx)
    assert (deg == 2).any() and ((deg > 2) & (deg <= 32)).any() and (deg > 64).any()
    r, o = _run(gr, orc, X, Y, Z, cap, pfx, pins, cost_model, commit_mode, f"uniform Z={Z}")
    # the grid runs out of capacity on the busiest cells: negative remainders too
    assert (r.demand() > 3).any()


def test_degenerate_pairs_on_random_grid(gr, orc, gpu):
    X, Y, Z = 33, 47, 8
    cap = gr.gen_grid(1, 5, X, Y, Z)
    pfx, pins = _edge_case_nets(X, Y, Z, seed=11, n_rand=200)
    _run(gr, orc, X, Y, Z, cap, pfx, pins, 0, 0, "random grid")


def test_empty_batches(gr, orc, gpu):
    X, Y, Z = 24, 24, 4
    cap = np.full(X * Y * Z, 2, np.int32)
    pfx, pins = _edge_case_nets(X, Y, Z, seed=3, n_rand=20)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    d0 = r.demand().copy()
    for h in (r, o):
        e, lp, c = h.route_batch(np.zeros(0, np.int32))
        assert e.size == 0 and lp.size == 0 and c.size == 0
    assert np.array_equal(r.demand(), d0) and np.array_equal(o.demand(), d0)
    # a schedule with empty batches between full ones equals batch-by-batch routing
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 2)
    ofst2 = np.concatenate([[0], np.repeat(ofst[1:], 2)]).astype(np.int32)  # every batch followed by an empty one
    r.set_schedule(ofst2, ids)
    r.route_schedule()
    r.sync()
    for b in range(len(ofst) - 1):
        o.route_batch(ids[ofst[b]:ofst[b + 1]])
    assert np.array_equal(r.demand(), o.demand())
    edges, lpat = r.schedule_outputs()[:2]
    r2 = gr.Router(X, Y, Z, cap)
    r2.set_nets(pfx, pins)
    e_ref = np.concatenate([r2.route_batch(ids[ofst[b]:ofst[b + 1]])[0] for b in range(len(ofst) - 1)])
    assert np.array_equal(edges, e_ref)


@pytest.mark.parametrize("X,Y,Z", [(1, 1, 6), (1, 2, 12), (2, 1, 9)])
def test_quirk_cells_below_the_grid(gr, orc, gpu, X, Y, Z):
    """Grids with X*Y < Z: the quirk's extra cells before the first wire can
    fall below enc 0; both sides skip them (no write outside the grid)."""
    nets = []
    for x in range(X):
        for y in range(Y):
            nets.append([(Z - 1, x, y), (0, x, y)])
            nets.append([(Z - 1, x, y), (Z // 2, x, y), (1, x, y)])
    if X * Y > 1:
        nets.append([(Z - 1, 0, 0), (0, X - 1, Y - 1)])
    pfx, pins = _encode(X, Y, Z, nets)
    cap = np.full(X * Y * Z, 2, np.int32)
    _run(gr, orc, X, Y, Z, cap, pfx, pins, 0, 0, f"tiny {X}x{Y}x{Z}")
