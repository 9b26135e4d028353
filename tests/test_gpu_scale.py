"""GPU parity at scale: nets of hundreds of pins (pair costs streamed from
global memory in the tree stage) and the full 2000x2000x10 / 1.5M-net
workloads (heavy-tailed pin counts; and the same nets on capacities cut by 40%
with 30 hotspots of capacity 1-2, where overflow is heavy), where sampled
batches are re-routed by the oracle from the GPU's own state and must match
bit-exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _enc(X, Y, z, x, y):
    return z * X * Y + (x * Y + y if z & 1 else y * X + x)


def test_huge_nets(gr, orc, gpu):
    X, Y, Z = 256, 256, 10
    cap = gr.gen_grid(2, 9, X, Y, Z)
    rng = np.random.default_rng(9)
    nets = []
    for npins in (2500, 900, 620, 333, 170, 90, 40, 7, 2):
        s = set()
        while len(s) < npins:
            s.add(_enc(X, Y, int(rng.integers(0, 2)), int(rng.integers(0, X)), int(rng.integers(0, Y))))
        nets.append(sorted(s))
    pfx = np.cumsum([0] + [len(n) for n in nets]).astype(np.int64)
    pins = np.concatenate([np.array(n, np.int32) for n in nets])
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    # every box spans most of the grid: one net per batch, largest first
    for nid in range(len(nets)):
        print(f"net {nid} pins {len(nets[nid])}", flush=True)
        e1, l1, c1 = r.route_batch([nid])
        e2, l2, c2 = o.route_batch([nid])
        pc1, pp1 = r.pair_costs()
        pc2, pp2 = o.pair_costs()
        assert _same_f32(pc1, pc2) and np.array_equal(pp1, pp2)
        assert np.array_equal(e1, e2) and np.array_equal(l1, l2) and _same_f32(c1, c2)
    assert np.array_equal(r.demand(), o.demand())


def test_mixed_batches_with_huge_nets(gr, orc, gpu):
    """Batches holding large and small nets together run as one mixed tree
    launch: with a net whose tree needs most of the LDS (one workgroup per CU,
    up to 15 small-net wavefronts per workgroup) and with moderate nets (two
    workgroups per CU).  Bit-exact with the oracle, and identical to the
    second-stream form (fork = 1)."""
    X, Y, Z = 256, 256, 10
    cap = gr.gen_grid(2, 11, X, Y, Z)
    rng = np.random.default_rng(11)
    nets = []
    for npins in (2500, 900, 90, 40, 7, 2, 333, 170, 60, 12, 3, 170, 90, 33, 9, 2):
        s = set()
        while len(s) < npins:
            s.add(_enc(X, Y, int(rng.integers(0, 2)), int(rng.integers(0, X)), int(rng.integers(0, Y))))
        nets.append(sorted(s))
    pfx = np.cumsum([0] + [len(n) for n in nets]).astype(np.int64)
    pins = np.concatenate([np.array(n, np.int32) for n in nets])
    batches = [list(range(0, 6)), list(range(6, 11)), list(range(11, 16))]
    outs = {}
    for fork in (3, 1):
        r = gr.Router(X, Y, Z, cap)
        r.set_tuning(fork=fork)
        r.set_nets(pfx, pins)
        o = orc.Oracle(X, Y, Z, cap) if fork == 3 else None
        if o is not None:
            o.set_nets(pfx, pins)
        res = []
        for b, batch in enumerate(batches):
            print(f"fork {fork} batch {b} pins {[len(nets[i]) for i in batch]}", flush=True)
            e1, l1, c1 = r.route_batch(batch)
            res.append((e1, l1, c1.view(np.int32), r.demand().copy()))
            if o is not None:
                e2, l2, c2 = o.route_batch(batch)
                assert np.array_equal(e1, e2) and np.array_equal(l1, l2) and _same_f32(c1, c2), b
                assert np.array_equal(r.demand(), o.demand()), b
        outs[fork] = res
    for a, b in zip(outs[3], outs[1]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("kind,X,nets", [(2, 800, 150000), (3, 2000, 1500000), (5, 2000, 1500000)])
def test_full_workload_sampled_batches(gr, orc, gpu, kind, X, nets):
    Y, Z = X, 10
    cap = gr.gen_grid(kind, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(kind, 0, X, Y, Z, nets)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    nb = len(ofst) - 1
    deg = np.diff(pfx)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    done = 0
    slot0 = np.concatenate([[0], np.cumsum(deg[ids])])
    for b in (1, nb // 2, nb - 1):
        print(f"workload{kind} batch {b} of {nb}: {ofst[b + 1] - ofst[b]} nets", flush=True)
        r.route_schedule(done, b - done)
        r.sync()
        o.set_rem(cap - r.demand())
        r.route_schedule(b, 1)
        done = b + 1
        e1, l1, c1 = r.schedule_outputs()
        batch = ids[ofst[b]:ofst[b + 1]]
        e2, l2, c2 = o.route_batch(batch)
        s0, s1 = slot0[ofst[b]], slot0[ofst[b + 1]]
        pc1, pp1 = r.pair_costs()
        pc2, pp2 = o.pair_costs()
        assert _same_f32(pc1, pc2) and np.array_equal(pp1, pp2), b
        assert np.array_equal(e1[s0:s1], e2) and np.array_equal(l1[s0:s1], l2), b
        assert _same_f32(c1[ofst[b]:ofst[b + 1]], c2), b
        assert np.array_equal(r.demand(), o.demand()), b
    # whole pass: every net has deg-1 edges and one sentinel in its slots
    r.reset_demand()
    r.route_schedule()
    e, l, c = r.schedule_outputs()
    last = slot0[1:] - 1
    assert np.all(e[last] == -1) and np.all(l[last] == -1)
    assert int((e >= 0).sum()) == int((deg - 1).sum())
    # every net's edges connect all of its pins: over the whole pass the edge
    # graph (node = schedule slot) has exactly one component per net
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    pos = np.repeat(np.arange(len(ids)), deg[ids])  # schedule position of every slot
    ok = e >= 0
    src = slot0[pos[ok]] + ((e[ok] >> 15) & 0x7FFF)
    dst = slot0[pos[ok]] + (e[ok] & 0x7FFF)
    assert np.all((e[ok] >> 30) == np.arange(len(ids))[pos[ok]] - ofst[np.searchsorted(ofst, pos[ok], "right") - 1])
    n = int(slot0[-1])
    g = coo_matrix((np.ones(len(src), np.int8), (src, dst)), shape=(n, n))
    assert connected_components(g, directed=False)[0] == len(ids)
    if kind == 5:  # the stress grid really overflows
        assert int(np.maximum(r.demand() - cap, 0).sum()) > 0


def test_giant_nets(gr, orc, gpu):
    """Nets of more than 4096 pins (up to the 32767 of the 15-bit pin index):
    their trees keep the node arrays in device memory.  Two such nets in
    disjoint boxes route in one batch, then one alone, each bit-exact with the
    oracle; a smaller net follows to check the state."""
    X, Y, Z = 300, 160, 4
    cap = gr.gen_grid(2, 21, X, Y, Z)
    rng = np.random.default_rng(21)
    nets = []
    for npins, x0 in ((4300, 0), (5200, 150), (4097, 40), (300, 0)):
        s = set()
        while len(s) < npins:
            s.add(_enc(X, Y, int(rng.integers(0, 2)), int(rng.integers(x0, x0 + 140)), int(rng.integers(0, Y))))
        nets.append(sorted(s))
    pfx = np.cumsum([0] + [len(n) for n in nets]).astype(np.int64)
    pins = np.concatenate([np.array(n, np.int32) for n in nets])
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    for batch in ([0, 1], [2], [3]):
        print(f"giant batch {batch} pins {[len(nets[i]) for i in batch]}", flush=True)
        e1, l1, c1 = r.route_batch(batch)
        e2, l2, c2 = o.route_batch(batch)
        pc1, pp1 = r.pair_costs()
        pc2, pp2 = o.pair_costs()
        assert _same_f32(pc1, pc2) and np.array_equal(pp1, pp2), batch
        assert np.array_equal(e1, e2) and np.array_equal(l1, l2) and _same_f32(c1, c2), batch
        assert np.array_equal(r.demand(), o.demand()), batch
    with pytest.raises(gr.RouterError, match="4096"):
        r.route_batch_static([2])


def test_streamed_key_layouts_agree(gr, gpu):
    """A batch whose large nets stream their pair keys: the trees read the
    square key matrices the lane costing kernel writes; the wavefront costing
    kernel writes them too; the allocation-free entry point has none and
    streams the triangular pair costs.  All three route identically."""
    X, Y, Z = 256, 256, 10
    cap = gr.gen_grid(2, 13, X, Y, Z)
    rng = np.random.default_rng(13)
    nets = []
    for npins in (1700, 640, 95, 40, 7, 2):
        s = set()
        while len(s) < npins:
            s.add(_enc(X, Y, int(rng.integers(0, 2)), int(rng.integers(0, X)), int(rng.integers(0, Y))))
        nets.append(sorted(s))
    pfx = np.cumsum([0] + [len(n) for n in nets]).astype(np.int64)
    pins = np.concatenate([np.array(n, np.int32) for n in nets])
    outs = []
    for form in ("lane", "wave", "static"):
        print(f"form {form}", flush=True)
        r = gr.Router(X, Y, Z, cap)
        if form == "wave":
            r.set_tuning(lpat_wave=1)
        r.set_nets(pfx, pins)
        res = []
        for batch in ([0, 1, 2, 3, 4, 5], [1, 2], [0]):
            e, l, c = r.route_batch_static(batch) if form == "static" else r.route_batch(batch)
            res += [e, l, np.asarray(c, np.float32).view(np.int32), r.demand().copy()]
        outs.append(res)
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert np.array_equal(a, b)
