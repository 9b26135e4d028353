"""True-MST tree mode (oracle, CPU): the recorded pairs are the minimum
spanning tree of the pair costs — checked against the hand KAT and against a
restatement of the reference's CPU Kruskal twin (MST_UTIL::t_get_mst,
mst_util.hpp:34-102, ties by pair index) — while the rounds, unions and net
cost are those of the root-pair mode."""
import json
import os

import numpy as np

from oracle import ispd_ref

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_hot_path.json")))


def _kruskal(n, costs):
    """mst_util.hpp:34-102: pairs sorted by cost (ties by pair index), DSU, first n-1 joins."""
    pairs = [(j, k) for j in range(1, n) for k in range(j)]
    order = sorted(range(len(pairs)), key=lambda o: (float(costs[o]), o))
    par = list(range(n))

    def find(v):
        while par[v] != v:
            par[v] = par[par[v]]
            v = par[v]
        return v
    out = []
    for o in order:
        j, k = pairs[o]
        a, b = find(j), find(k)
        if a != b:
            par[a] = b
            out.append((j, k))
        if len(out) == n - 1:
            break
    return out


def test_true_mst_kat(orc):
    kat = GOLD["boruvka"]
    edges, cost, rounds = orc.boruvka(kat["n"], kat["pair_costs_by_o"], true_mst=True)
    assert sorted([max(e), min(e)] for e in edges) == sorted(kat["true_mst_hi_lo"])
    assert cost == kat["net_cost"] and rounds == 2


def test_true_mst_equals_kruskal(orc):
    rng = np.random.default_rng(11)
    for trial in range(300):
        n = int(rng.integers(2, 26))
        P = n * (n - 1) // 2
        costs = (rng.integers(0, 3, P) if trial % 2 else rng.random(P) + 0.01).astype(np.float32)
        e_mst, c_mst, r_mst = orc.boruvka(n, costs, true_mst=True)
        e_rp, c_rp, r_rp = orc.boruvka(n, costs)
        assert sorted((max(e), min(e)) for e in e_mst) == sorted(_kruskal(n, costs)), trial
        # same rounds and net cost; recorded endpoints lie in the two components the root pair joins
        assert (c_mst, r_mst) == (c_rp, r_rp)
        w_mst = sum(float(costs[max(e) * (max(e) - 1) // 2 + min(e)]) for e in e_mst)
        w_rp = sum(float(costs[max(e) * (max(e) - 1) // 2 + min(e)]) for e in e_rp)
        assert w_mst <= w_rp + 1e-6 * max(1.0, w_rp)


def test_true_mst_routes(gr, orc, tmp_path):
    """Routed in true-MST mode: first-batch trees weigh no more than the
    root-pair trees at the same state, demand rises by 0 or 1 per batch, and
    the written routes connect every net."""
    X, Y, Z = 80, 70, 6
    cap = gr.gen_grid(1, 6, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 6, X, Y, Z, 1200)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    a = orc.Oracle(X, Y, Z, cap)
    b = orc.Oracle(X, Y, Z, cap)
    b.set_tree_mode(True)
    for o in (a, b):
        o.set_nets(pfx, pins)
    batch = ids[ofst[0]:ofst[1]]
    ea, la, ca = a.route_batch(batch)
    pc, _ = a.pair_costs()
    eb, lb, cb = b.route_batch(batch)
    assert np.array_equal(ca, cb)
    deg = np.diff(pfx)[batch]
    s0 = np.concatenate([[0], np.cumsum(deg)])
    p0 = np.concatenate([[0], np.cumsum(deg * (deg - 1) // 2)])
    for i in range(len(batch)):
        def weight(E):
            w = 0.0
            for e in E[s0[i]:s0[i + 1] - 1]:
                j, k = (int(e) >> 15) & 0x7FFF, int(e) & 0x7FFF
                hi, lo = max(j, k), min(j, k)
                w += float(pc[p0[i] + hi * (hi - 1) // 2 + lo])
            return w
        assert weight(eb) <= weight(ea) + 1e-5
    E, L = [eb], [lb]
    before = b.demand()
    for k in range(1, len(ofst) - 1):
        e, l, _ = b.route_batch(ids[ofst[k]:ofst[k + 1]])
        d = b.demand() - before
        before = b.demand()
        assert d.min() >= 0 and d.max() <= 1
        E.append(e)
        L.append(l)
    d = gr.Design(X, Y, Z, cap.astype(np.float32), pfx, pins)
    d.write_routes(str(tmp_path / "m.txt"), ofst, ids, np.concatenate(E), np.concatenate(L))
    text = (tmp_path / "m.txt").read_text()
    for block in text.split(")\n"):
        if not block.strip():
            continue
        name, body = block.split("\n(\n")
        k = int(name[3:])
        par = {}

        def find(v):
            while par.setdefault(v, v) != v:
                v = par[v]
            return v
        for ln in body.split("\n"):
            if not ln:
                continue
            z1, x1, y1, z2, x2, y2 = map(int, ln.split())
            if (x1, y1, z1) == (x2, y2, z2):
                par[find((z1, x1, y1))] = find((z1 + 1, x1, y1))
            elif x1 == x2:
                for t in range(min(y1, y2), max(y1, y2)):
                    par[find((z1, x1, t))] = find((z1, x1, t + 1))
            else:
                for t in range(min(x1, x2), max(x1, x2)):
                    par[find((z1, t, y1))] = find((z1, t + 1, y1))
        assert len({find(ispd_ref.dec(X, Y, int(c))) for c in pins[pfx[k]:pfx[k + 1]]}) == 1, k
