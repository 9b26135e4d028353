"""Route evaluator restatement (CPU): a hand-computed route, and the counts
from the oracle's difference-array/run-scan formulation against counts taken
from the writer's printed segments."""
import math

import numpy as np

from oracle import ispd_ref


def test_eval_known_answer(orc):
    """X=5, Y=4, Z=4: pins (0,1,0) and (0,3,2), pattern a=1, b=1 (b'=2)."""
    X, Y, Z = 5, 4, 4
    A, B = ispd_ref.enc(X, Y, 0, 1, 0), ispd_ref.enc(X, Y, 0, 3, 2)
    pfx = np.array([0, 2], np.int64)
    pins = np.array([A, B], np.int32)
    edges = np.array([(1 << 15) | 0, -1], np.int64)
    lpat = np.array([1 * 2 + 1, -1], np.int32)
    cap = np.ones(X * Y * Z, np.float32)
    dem = np.zeros(X * Y * Z, np.int32)
    w, v, layers = orc.evaluate(X, Y, Z, cap, dem, [0, 1], [0], pfx, pins, edges, lpat)
    e = lambda z, x, y: ispd_ref.enc(X, Y, z, x, y)
    want_w = {e(1, 3, 0): 1, e(1, 3, 1): 1, e(2, 1, 0): 1, e(2, 2, 0): 1}
    want_v = {e(0, 2, 2): 1, e(0, 3, 2): 1, e(1, 3, 0): 2, e(0, 0, 0): 1, e(0, 1, 0): 1, e(1, 1, 0): 2}
    assert {int(i): int(w[i]) for i in np.nonzero(w)[0]} == want_w
    assert {int(i): int(v[i]) for i in np.nonzero(v)[0]} == want_v
    L1 = layers[1]
    assert (L1["wire"], L1["via"], L1["max_wire"], L1["max_via"], L1["max_overflow"]) == (2, 4, 1, 2, 1.0)
    # of per cell on layer 1: (1,3,0): 2w+v-2cap = 2; (1,3,1): 0; (1,1,0): 0; the other 17 cells: -2
    want = np.exp(0.5) + 2 * 1.0 + 17 * np.exp(-0.5)
    assert abs(L1["of_cost"] - want) < 1e-6 * want
    L3 = layers[3]
    assert L3["wire"] == 0 and L3["via"] == 0 and abs(L3["of_cost"] - 20 * math.exp(-0.5)) < 1e-5


def _counts_from_text(X, Y, Z, text, batch_of_net):
    """Counts from the writer's grid-form lines: wire [min, max) once per batch, via charges."""
    w = np.zeros(X * Y * Z, np.int32)
    v = np.zeros(X * Y * Z, np.int32)
    seen = {}
    for block in text.split(")\n"):
        if not block.strip():
            continue
        name, body = block.split("\n(\n")
        b = batch_of_net[name]
        for ln in body.split("\n"):
            if not ln:
                continue
            z1, x1, y1, z2, x2, y2 = map(int, ln.split())
            if (x1, y1, z1) == (x2, y2, z2):
                pos, L = (y1, Y) if z1 & 1 else (x1, X)
                c = ispd_ref.enc(X, Y, z1, x1, y1)
                if pos == 0:
                    v[c] += 2
                elif pos == L - 1:
                    v[c - 1] += 2
                else:
                    v[c - 1] += 1
                    v[c] += 1
                continue
            cells = ([ispd_ref.enc(X, Y, z1, x1, t) for t in range(min(y1, y2), max(y1, y2))] if x1 == x2 else
                     [ispd_ref.enc(X, Y, z1, t, y1) for t in range(min(x1, x2), max(x1, x2))])
            for c in cells:
                if seen.get(c) != b:
                    seen[c] = b
                    w[c] += 1
    return w, v


def test_eval_counts_equal_written_routes(gr, orc, tmp_path):
    X, Y, Z = 60, 50, 6
    cap = gr.gen_grid(1, 9, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 9, X, Y, Z, 800)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    E, L = [], []
    for b in range(len(ofst) - 1):
        e, l, _ = o.route_batch(ids[ofst[b]:ofst[b + 1]])
        E.append(e)
        L.append(l)
    E, L = np.concatenate(E), np.concatenate(L)
    w, v, layers = orc.evaluate(X, Y, Z, cap.astype(np.float32), o.demand(), ofst, ids, pfx, pins, E, L)
    d = gr.Design(X, Y, Z, cap.astype(np.float32), pfx, pins)
    d.write_routes(str(tmp_path / "r.txt"), ofst, ids, E, L)
    batch_of = {f"net{ids[q]}": b for b in range(len(ofst) - 1) for q in range(ofst[b], ofst[b + 1])}
    w2, v2 = _counts_from_text(X, Y, Z, (tmp_path / "r.txt").read_text(), batch_of)
    assert np.array_equal(w, w2) and np.array_equal(v, v2)
    assert w.sum() > 0 and v.sum() > 0
    dem = o.demand()
    for z in range(Z):
        sl = slice(z * X * Y, (z + 1) * X * Y)
        assert layers[z]["wire"] == w[sl].sum() and layers[z]["via"] == v[sl].sum()
        assert layers[z]["demand"] == dem[sl].sum()
        assert layers[z]["overflow"] == np.maximum(dem[sl] - cap[sl], 0).sum()
