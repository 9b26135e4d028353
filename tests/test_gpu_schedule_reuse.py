"""GPU: the captured whole-pass graph stays consistent when the schedule or
the context's pair buffers change under it.

A resident schedule's pass is captured once as a HIP graph.  Rebuilding the
schedule (a second gr_set_schedule on the same nets) or growing the pair
buffers that every schedule shares (a gr_route_batch with more pin pairs than
any scheduled batch) must drop that capture; the next full pass then equals
the oracle's routing of the new schedule, bit for bit.
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


def _check_pass(r, want, label):
    print(label, flush=True)
    r.reset_demand()
    r.route_schedule()
    e, l, c = r.schedule_outputs()
    E, L, C, D = want
    assert np.array_equal(e, E), f"{label}: edges differ"
    assert np.array_equal(l, L), f"{label}: edge patterns differ"
    assert _same_f32(c, C), f"{label}: net costs differ"
    assert np.array_equal(r.demand(), D), f"{label}: demand differs"


def test_set_schedule_twice(gr, orc, gpu):
    """A second, larger schedule on the same nets replaces the captured pass."""
    X, Y, Z = 96, 80, 6
    cap = gr.gen_grid(1, 41, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 41, X, Y, Z, 1200)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    # first schedule: the first half of the batches
    h = (len(ofst) - 1) // 2
    o1, i1 = ofst[:h + 1], ids[:ofst[h]]
    r.set_schedule(o1, i1)
    _check_pass(r, _oracle_pass(orc, X, Y, Z, cap, pfx, pins, o1, i1), "half schedule")
    _check_pass(r, _oracle_pass(orc, X, Y, Z, cap, pfx, pins, o1, i1), "half schedule replay")
    # second schedule: every batch (more positions, slots, chunks and pairs)
    r.set_schedule(ofst, ids)
    _check_pass(r, _oracle_pass(orc, X, Y, Z, cap, pfx, pins, ofst, ids), "full schedule")
    # and back to a smaller one (buffers keep their size, contents change)
    r.set_schedule(o1, i1)
    _check_pass(r, _oracle_pass(orc, X, Y, Z, cap, pfx, pins, o1, i1), "half schedule again")


def test_route_batch_between_passes(gr, orc, gpu):
    """A single batch larger than any scheduled one grows the shared pair
    buffers; the next scheduled pass must not use the freed ones."""
    X, Y, Z = 96, 80, 6
    cap = gr.gen_grid(1, 43, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 43, X, Y, Z, 1200)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    # schedule of the small batches only (the last ones, few nets each)
    nb = len(ofst) - 1
    first = nb - 6
    o1 = ofst[first:] - ofst[first]
    i1 = ids[ofst[first]:]
    r.set_schedule(o1, i1)
    want = _oracle_pass(orc, X, Y, Z, cap, pfx, pins, o1, i1)
    _check_pass(r, want, "tail schedule")
    # one batch with many more pairs than any scheduled batch
    big = ids[ofst[0]:ofst[1]]
    deg = np.diff(pfx)
    assert (deg[big] * (deg[big] - 1) // 2).sum() > max(
        (deg[i1[o1[b]:o1[b + 1]]] * (deg[i1[o1[b]:o1[b + 1]]] - 1) // 2).sum() for b in range(len(o1) - 1))
    r.reset_demand()
    e, l, c = r.route_batch(big)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    e2, l2, c2 = o.route_batch(big)
    assert np.array_equal(e, e2) and np.array_equal(l, l2) and _same_f32(c, c2)
    _check_pass(r, want, "tail schedule after a large single batch")
    _check_pass(r, want, "tail schedule replay")
