"""GPU: the allocation-free entry point (gr_route_batch_static) against
gr_route_batch.

gr_route_batch_static keeps every device buffer of a batch in the library's
static per-GPU array instead of allocating; its results must equal
gr_route_batch's bit for bit.  Two contexts on the same grid and nets route
the same batches, one through each entry point: first a large batch, then a
smaller one (the array keeps the first call's contents), then a mid-size one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _pairs(pfx, batch):
    d = np.diff(pfx)[batch]
    return int((d * (d - 1) // 2).sum())


def test_static_equals_route_batch(gr, gpu):
    X, Y, Z = 128, 128, 6
    cap = gr.gen_grid(1, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 0, X, Y, Z, 5000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    batches = [ids[ofst[b]:ofst[b + 1]] for b in range(len(ofst) - 1)]
    order = sorted(range(len(batches)), key=lambda b: -_pairs(pfx, batches[b]))
    big, small, mid = order[0], order[-3], order[len(order) // 2]
    assert _pairs(pfx, batches[small]) < _pairs(pfx, batches[big])
    r1 = gr.Router(X, Y, Z, cap)
    r2 = gr.Router(X, Y, Z, cap)
    for r in (r1, r2):
        r.set_nets(pfx, pins)
    assert gr.load_library().gr_static_scratch_bytes() == 64 << 20
    for tag, b in (("large", big), ("smaller", small), ("mid", mid)):
        print(f"static vs route_batch: {tag} batch {b} nets {len(batches[b])} pairs {_pairs(pfx, batches[b])}",
              flush=True)
        e1, l1, c1 = r1.route_batch(batches[b])
        e2, l2, c2 = r2.route_batch_static(batches[b])
        assert np.array_equal(e1, e2), f"{tag}: edges differ"
        assert np.array_equal(l1, l2), f"{tag}: edge patterns differ"
        assert _same_f32(c1, c2), f"{tag}: net costs differ"
        pc1, pp1 = r1.pair_costs()
        pc2, pp2 = r2.pair_costs()
        assert _same_f32(pc1, pc2) and np.array_equal(pp1, pp2), f"{tag}: pair results differ"
        assert np.array_equal(r1.demand(), r2.demand()), f"{tag}: demand differs"
    assert _same_f32(r1.prefix(), r2.prefix())


def test_static_rejects_oversized_batch(gr, gpu):
    """Four 2500-pin nets hold 12.5M pin pairs, 75 MB of pair buffers: more
    than the 64 MiB array, so GR_ERANGE before any device work and the state
    is untouched; one such net alone (3.1M pairs) fits and routes as
    gr_route_batch routes it."""
    X, Y, Z = 512, 128, 6
    rng = np.random.default_rng(5)
    nets = []
    for x0 in (0, 128, 256, 384):
        cells = set()
        while len(cells) < 2500:
            x, y, z = int(rng.integers(x0, x0 + 100)), int(rng.integers(0, 128)), int(rng.integers(0, 2))
            cells.add(z * X * Y + (x * Y + y if z & 1 else y * X + x))
        nets.append(sorted(cells))
    pfx = np.arange(0, 4 * 2500 + 1, 2500, dtype=np.int64)
    pins = np.array(sum(nets, []), np.int32)
    cap = np.full(X * Y * Z, 4, np.int32)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    print("static: oversized batch of four 2500-pin nets", flush=True)
    with pytest.raises(gr.RouterError, match="scratch"):
        r.route_batch_static(np.arange(4, dtype=np.int32))
    assert not r.demand().any()
    print("static: one 2500-pin net", flush=True)
    e, l, c = r.route_batch_static(np.array([0], np.int32))
    r2 = gr.Router(X, Y, Z, cap)
    r2.set_nets(pfx, pins)
    e2, l2, c2 = r2.route_batch(np.array([0], np.int32))
    assert np.array_equal(e, e2) and np.array_equal(l, l2) and _same_f32(c, c2)
    assert np.array_equal(r.demand(), r2.demand())


def test_static_refuses_proposal(gr, gpu):
    X, Y, Z = 48, 40, 6
    cap = gr.gen_grid(1, 7, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 7, X, Y, Z, 200)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    m = gr.Model(r, hidden=256, seed=0, max_samples=1024)
    r.set_proposal(m, k=4, threshold=0.6)
    with pytest.raises(gr.RouterError, match="exhaustive"):
        r.route_batch_static(np.array([0], np.int32))
    r.set_proposal(None)
    e, l, c = r.route_batch_static(np.array([0], np.int32))
    assert e[-1] == -1
    m.close()
