"""GPU: one launch per batch for costing, trees and commit (tuning fork = 4).

The fused launch splits each batch into one queue per XCD (costing chunks,
large-net trees that wait for their own chunks, small-net tree blocks that
wait for their XCD's small-net chunks).  Its routes must equal the oracle's
bit for bit, on every replay of the captured pass and launched batch by
batch, and no bounded wait may time out (gr_sync reports one).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _oracle_pass(orc, X, Y, Z, cap, pfx, pins, ofst, ids):
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    E, L, C = [], [], []
    for b in range(len(ofst) - 1):
        e, l, c = o.route_batch(ids[ofst[b]:ofst[b + 1]])
        E.append(e); L.append(l); C.append(c)
    return np.concatenate(E), np.concatenate(L), np.concatenate(C), o.demand()


def _check(r, want, label):
    print(label, flush=True)
    r.reset_demand()
    r.route_schedule()
    r.sync()  # raises if a fused launch's wait timed out
    e, l, c = r.schedule_outputs()
    E, L, C, D = want
    assert np.array_equal(e, E), f"{label}: edges differ"
    assert np.array_equal(l, L), f"{label}: edge patterns differ"
    assert _same_f32(c, C), f"{label}: net costs differ"
    assert np.array_equal(r.demand(), D), f"{label}: demand differs"


@pytest.mark.parametrize("group_max,threads", [(32, 1024), (64, 512)])
def test_fused_workload1(gr, orc, gpu, group_max, threads):
    """128x128x6, 5,000 nets: every batch through the fused launch."""
    W = gr.WORKLOADS[1]
    X, Y, Z = W["X"], W["Y"], W["Z"]
    cap = gr.gen_grid(1, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 0, X, Y, Z, W["nets"])
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.set_tuning(large_threads=threads, group_max=group_max, fork=4, refresh_blocks=4096, lpat_wave=0, lpat_block=0)
    want = _oracle_pass(orc, X, Y, Z, cap, pfx, pins, ofst, ids)
    for rep in range(3):
        _check(r, want, f"workload 1 fused, replay {rep}")


def test_fused_800_slice(gr, orc, gpu):
    """800x800x10: the first 4 and the last 12 batches of the 150k-net
    workload as one schedule (bulk and tail batches, nets up to ~200 pins)."""
    W = gr.WORKLOADS[2]
    X, Y, Z = W["X"], W["Y"], W["Z"]
    cap = gr.gen_grid(2, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(2, 0, X, Y, Z, W["nets"])
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    nb = len(ofst) - 1
    keep = list(range(4)) + list(range(nb - 12, nb))
    parts = [ids[ofst[b]:ofst[b + 1]] for b in keep]
    o1 = np.concatenate([[0], np.cumsum([len(p) for p in parts])]).astype(np.int32)
    i1 = np.concatenate(parts).astype(np.int32)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(o1, i1)
    r.set_tuning(large_threads=1024, group_max=32, fork=4, refresh_blocks=8192, lpat_wave=0, lpat_block=0)
    want = _oracle_pass(orc, X, Y, Z, cap, pfx, pins, o1, i1)
    _check(r, want, "800x800x10 slice fused")
    _check(r, want, "800x800x10 slice fused, replay")
    # batch by batch (no captured graph): the same routes
    r.reset_demand()
    for b in range(len(o1) - 1):
        r.route_schedule(b, 1)
    r.sync()
    e, l, c = r.schedule_outputs()
    E, L, C, D = want
    assert np.array_equal(e, E) and np.array_equal(l, L) and _same_f32(c, C)
    assert np.array_equal(r.demand(), D)
