"""GPU parity of the non-BIT ("prefix") mode: prefix-sum L-pattern costing
(lpattern.hpp:13-160) and the difference-array commit (mst_commit.hpp:17-364)
against the oracle, bit for bit on every batch: per-pair minimum cost and
pattern, tree edges and edge patterns, net costs, the fp32 remaining
capacity, the wire-cost prefix P and the per-cell via cost V.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PREFIX = 2
INC = 1


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _compare(gr, orc, X, Y, Z, cap, pfx, pins, label, every_state=True, parts=4):
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, parts)
    r = gr.Router(X, Y, Z, cap, INC, PREFIX)
    r.set_nets(pfx, pins)
    o = orc.Oracle(X, Y, Z, cap, INC, PREFIX)
    o.set_nets(pfx, pins)
    assert _same_f32(r.prefix(), o.prefix()) and _same_f32(r.via_terms(), o.via_terms()), f"{label}: initial state"
    nb = len(ofst) - 1
    for b in range(nb):
        batch = ids[ofst[b]:ofst[b + 1]]
        print(f"{label} batch {b} nets {len(batch)}", flush=True)
        e1, l1, c1 = r.route_batch(batch)
        e2, l2, c2 = o.route_batch(batch)
        pc1, pp1 = r.pair_costs()
        pc2, pp2 = o.pair_costs()
        assert _same_f32(pc1, pc2), f"{label} batch {b}: pair costs differ"
        assert np.array_equal(pp1, pp2), f"{label} batch {b}: pair patterns differ"
        assert np.array_equal(e1, e2), f"{label} batch {b}: tree edges differ"
        assert np.array_equal(l1, l2), f"{label} batch {b}: edge patterns differ"
        assert _same_f32(c1, c2), f"{label} batch {b}: net costs differ"
        if every_state or b == nb - 1:
            assert _same_f32(r.remaining_f32(), o.remaining_f32()), f"{label} batch {b}: remaining capacity differs"
            assert _same_f32(r.prefix(), o.prefix()), f"{label} batch {b}: wire prefix differs"
            assert _same_f32(r.via_terms(), o.via_terms()), f"{label} batch {b}: via costs differ"
    return r, o, ofst, ids


def test_workload1_full(gr, orc, gpu):
    """128x128x6, 5,000 nets, seed 0, every batch."""
    X, Y, Z = 128, 128, 6
    cap = gr.gen_grid(1, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 0, X, Y, Z, 5000)
    r, o, ofst, ids = _compare(gr, orc, X, Y, Z, cap, pfx, pins, "prefix w1")
    with pytest.raises(gr.RouterError, match="remaining"):
        r.demand()


def test_workload2_slice(gr, orc, gpu):
    """800x800x10 with 5% blocked cells and the 2-200-pin mix (workgroup
    trees), 6,000 nets: every batch's outputs, state at the end."""
    X, Y, Z = 800, 800, 10
    cap = gr.gen_grid(2, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(2, 0, X, Y, Z, 6000)
    _compare(gr, orc, X, Y, Z, cap, pfx, pins, "prefix w2", every_state=False)


@pytest.mark.parametrize("Z", [3, 4, 5, 7, 12, 13])
def test_layer_counts(gr, orc, gpu, Z):
    """Every lane-kernel instantiation family (odd and even Z) and the
    wavefront-per-pair kernel (Z = 13), on a non-square grid."""
    X, Y = 72, 56
    cap = gr.gen_grid(2, 40 + Z, X, Y, Z)
    pfx, pins = gr.gen_nets(2, 40 + Z, X, Y, Z, 900)
    _compare(gr, orc, X, Y, Z, cap, pfx, pins, f"prefix z{Z}")


def test_fractional_capacities(gr, orc, gpu):
    """fp32 capacities with fractions (the .cap file's values): via halves on
    non-half-integral remainders round in fp32 exactly as the oracle's."""
    X, Y, Z = 96, 80, 6
    rng = np.random.default_rng(8)
    cap = (rng.integers(0, 40, X * Y * Z) * 0.3 + 0.1).astype(np.float32)
    pfx, pins = gr.gen_nets(1, 9, X, Y, Z, 1500)
    _compare(gr, orc, X, Y, Z, cap, pfx, pins, "prefix f32")


def test_uniform_grid_ties_and_long_runs(gr, orc, gpu):
    """Uniform capacities (pattern and pair costs tie everywhere, lowest index
    wins) on a grid longer than 128 cells per run, so the wire prefix has
    several cells per scan chunk, driven past its capacity."""
    X, Y, Z = 300, 140, 5
    cap = np.full(X * Y * Z, 2, np.int32)
    pfx, pins = gr.gen_nets(2, 12, X, Y, Z, 2500)
    r, o, _, _ = _compare(gr, orc, X, Y, Z, cap, pfx, pins, "prefix uniform", every_state=False)
    assert (r.remaining_f32() < 0).any()


def test_schedule_replay_and_static_entry(gr, orc, gpu):
    """Whole-schedule graph replays after resets, and the allocation-free
    entry point, route the non-BIT mode exactly as batch-by-batch calls."""
    X, Y, Z = 128, 128, 6
    cap = gr.gen_grid(1, 2, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 2, X, Y, Z, 3000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    a = gr.Router(X, Y, Z, cap, INC, PREFIX)
    a.set_nets(pfx, pins)
    outs = [a.route_batch(ids[ofst[b]:ofst[b + 1]]) for b in range(len(ofst) - 1)]
    b_ = gr.Router(X, Y, Z, cap, INC, PREFIX)
    b_.set_nets(pfx, pins)
    b_.set_schedule(ofst, ids)
    for rep in range(3):
        print(f"prefix replay {rep}", flush=True)
        b_.route_schedule()
        e, l, c = b_.schedule_outputs()
        assert np.array_equal(e, np.concatenate([x[0] for x in outs]))
        assert np.array_equal(l, np.concatenate([x[1] for x in outs]))
        assert _same_f32(c, np.concatenate([x[2] for x in outs]))
        assert _same_f32(b_.remaining_f32(), a.remaining_f32())
        b_.reset_demand()
    s = gr.Router(X, Y, Z, cap, INC, PREFIX)
    s.set_nets(pfx, pins)
    for b in range(len(ofst) - 1):
        e, l, c = s.route_batch_static(ids[ofst[b]:ofst[b + 1]])
        assert np.array_equal(e, outs[b][0]) and np.array_equal(l, outs[b][1]) and _same_f32(c, outs[b][2])
    assert _same_f32(s.remaining_f32(), a.remaining_f32())


def test_mode_needs_incremental_costs(gr, gpu):
    with pytest.raises(gr.RouterError):
        gr.Router(16, 16, 4, np.full(16 * 16 * 4, 3, np.int32), 0, PREFIX)
