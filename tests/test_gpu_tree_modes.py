"""True-MST tree mode on the GPU against the oracle in the same mode,
bit-exact (pair costs, edges, patterns, net costs, demand): the group,
workgroup-LDS and streamed tree paths, and graph replays across mode changes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _pair(gr, orc, X, Y, Z, cap, pfx, pins):
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_tree_mode(True)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    o.set_tree_mode(True)
    return r, o


def _compare(r, o, batches, label):
    for b, batch in enumerate(batches):
        print(f"{label} batch {b} nets {len(batch)}", flush=True)
        e1, l1, c1 = r.route_batch(batch)
        e2, l2, c2 = o.route_batch(batch)
        pc1, pp1 = r.pair_costs()
        pc2, pp2 = o.pair_costs()
        assert _same_f32(pc1, pc2) and np.array_equal(pp1, pp2), (label, b)
        assert np.array_equal(e1, e2) and np.array_equal(l1, l2) and _same_f32(c1, c2), (label, b)
    assert np.array_equal(r.demand(), o.demand()), label


@pytest.mark.parametrize("kind,X,Y,Z,nets", [(1, 128, 128, 6, 5000), (2, 800, 800, 10, 6000)])
def test_true_mst_parity(gr, orc, gpu, kind, X, Y, Z, nets):
    cap = gr.gen_grid(kind, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(kind, 0, X, Y, Z, nets)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r, o = _pair(gr, orc, X, Y, Z, cap, pfx, pins)
    _compare(r, o, [ids[ofst[b]:ofst[b + 1]] for b in range(len(ofst) - 1)], f"mst w{kind}")


def test_true_mst_huge_nets(gr, orc, gpu):
    X, Y, Z = 256, 256, 10
    cap = gr.gen_grid(2, 9, X, Y, Z)
    rng = np.random.default_rng(4)
    nets = []
    for npins in (700, 300, 120, 40, 9):
        s = set()
        while len(s) < npins:
            z = int(rng.integers(0, 2))
            x, y = int(rng.integers(0, X)), int(rng.integers(0, Y))
            s.add(z * X * Y + (x * Y + y if z & 1 else y * X + x))
        nets.append(sorted(s))
    pfx = np.cumsum([0] + [len(n) for n in nets]).astype(np.int64)
    pins = np.concatenate([np.array(n, np.int32) for n in nets])
    r, o = _pair(gr, orc, X, Y, Z, cap, pfx, pins)
    _compare(r, o, [[i] for i in range(len(nets))], "mst huge")


def test_mode_switch_with_graph_replay(gr, orc, gpu):
    X, Y, Z = 128, 128, 6
    cap = gr.gen_grid(1, 2, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 2, X, Y, Z, 4000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    outs = {}
    for mode in (False, True, False, True):
        r.set_tree_mode(mode)
        r.reset_demand()
        r.route_schedule()  # captured once per mode change, then replayed
        e, l, c = r.schedule_outputs()
        d = r.demand()
        if mode in outs:
            assert all(np.array_equal(u, v) for u, v in zip(outs[mode], (e, l, d)))
        outs[mode] = (e, l, d)
    for mode in (False, True):
        o = orc.Oracle(X, Y, Z, cap)
        o.set_nets(pfx, pins)
        o.set_tree_mode(mode)
        E, L = [], []
        for b in range(len(ofst) - 1):
            e, l, _ = o.route_batch(ids[ofst[b]:ofst[b + 1]])
            E.append(e)
            L.append(l)
        assert np.array_equal(outs[mode][0], np.concatenate(E)) and np.array_equal(outs[mode][1], np.concatenate(L))
        assert np.array_equal(outs[mode][2], o.demand())
    assert not np.array_equal(outs[False][0], outs[True][0])
