"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact on every output: per-pair minimum cost and pattern, tree edges,
per-edge patterns, per-net tree cost, integer demand and the materialised
Fenwick prefix state.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _setup(gr, orc, kind, seed, X, Y, Z, nets, cost_model=0, commit_mode=0, parts=4):
    cap = gr.gen_grid(kind, seed, X, Y, Z)
    pfx, pins = gr.gen_nets(kind, seed, X, Y, Z, nets)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, parts)
    r = gr.Router(X, Y, Z, cap, cost_model, commit_mode)
    r.set_nets(pfx, pins)
    o = orc.Oracle(X, Y, Z, cap, cost_model, commit_mode)
    o.set_nets(pfx, pins)
    return r, o, ofst, ids


def _compare_batches(r, o, ofst, ids, every_demand=True, label=""):
    for b in range(len(ofst) - 1):
        batch = ids[ofst[b]:ofst[b + 1]]
        print(f"{label} batch {b} nets {len(batch)}", flush=True)
        e1, l1, c1 = r.route_batch(batch)
        e2, l2, c2 = o.route_batch(batch)
        pc1, pp1 = r.pair_costs()
        pc2, pp2 = o.pair_costs()
        assert _same_f32(pc1, pc2), f"batch {b}: pair costs differ"
        assert np.array_equal(pp1, pp2), f"batch {b}: pair patterns differ"
        assert np.array_equal(e1, e2), f"batch {b}: tree edges differ"
        assert np.array_equal(l1, l2), f"batch {b}: edge patterns differ"
        assert _same_f32(c1, c2), f"batch {b}: net costs differ"
        if every_demand or b == len(ofst) - 2:
            assert np.array_equal(r.demand(), o.demand()), f"batch {b}: demand differs"
    assert _same_f32(r.prefix(), o.prefix()), "final prefix state differs"


def test_initial_state(gr, orc, gpu):
    for (X, Y, Z) in [(128, 128, 6), (37, 29, 7), (64, 200, 10)]:
        cap = gr.gen_grid(2, 3, X, Y, Z)
        r = gr.Router(X, Y, Z, cap)
        o = orc.Oracle(X, Y, Z, cap)
        assert np.array_equal(r.lut(), orc.lut(0))
        assert _same_f32(r.prefix(), o.prefix()), (X, Y, Z)
        assert not r.demand().any()


@pytest.mark.parametrize("cost_model", [0, 2])
def test_workload1_full(gr, orc, gpu, cost_model):
    """128x128x6, 5,000 nets, seed 0: every batch bit-exact, with the BIT cost
    exp(-r/2) and with the literal exp(d + 1 - c) = exp(1 - r) the workload
    names."""
    r, o, ofst, ids = _setup(gr, orc, 1, 0, 128, 128, 6, 5000, cost_model)
    _compare_batches(r, o, ofst, ids, label=f"w1m{cost_model}")


@pytest.mark.parametrize("cost_model,commit_mode", [(1, 0), (2, 0), (0, 1)])
def test_models_and_commit_modes(gr, orc, gpu, cost_model, commit_mode):
    r, o, ofst, ids = _setup(gr, orc, 1, 5, 96, 80, 6, 1500, cost_model, commit_mode)
    _compare_batches(r, o, ofst, ids, every_demand=False, label=f"m{cost_model}c{commit_mode}")


def test_odd_layers_and_rectangular(gr, orc, gpu):
    """Odd Z (b' >= Z patterns are invalid) on a non-square grid."""
    r, o, ofst, ids = _setup(gr, orc, 2, 11, 90, 140, 7, 2500)
    _compare_batches(r, o, ofst, ids, every_demand=False, label="z7")


@pytest.mark.parametrize("Z", [3, 4, 5, 8, 9, 11, 12, 13, 16])
def test_layer_counts(gr, orc, gpu, Z):
    """Every lane-kernel instantiation (Z = 3..12) and the wavefront-per-pair
    kernel (Z > 12), every batch bit-exact on a small non-square grid."""
    r, o, ofst, ids = _setup(gr, orc, 2, 20 + Z, 72, 56, Z, 900)
    _compare_batches(r, o, ofst, ids, every_demand=False, label=f"z{Z}")


def test_workload2_slice(gr, orc, gpu):
    """800x800x10 grid with 5% blocked cells and the workload-2 pin mix (incl. 21-200 pins)."""
    r, o, ofst, ids = _setup(gr, orc, 2, 0, 800, 800, 10, 6000)
    _compare_batches(r, o, ofst, ids, every_demand=False, label="w2")


def test_schedule_equals_route_batch(gr, orc, gpu):
    X, Y, Z = 128, 128, 6
    cap = gr.gen_grid(1, 2, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 2, X, Y, Z, 3000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    a = gr.Router(X, Y, Z, cap)
    a.set_nets(pfx, pins)
    outs = [a.route_batch(ids[ofst[b]:ofst[b + 1]]) for b in range(len(ofst) - 1)]
    b_ = gr.Router(X, Y, Z, cap)
    b_.set_nets(pfx, pins)
    b_.set_schedule(ofst, ids)
    for rep in range(3):  # graph replays after eager resets reproduce the eager result
        b_.route_schedule()
        e, l, c = b_.schedule_outputs()
        assert np.array_equal(e, np.concatenate([x[0] for x in outs]))
        assert np.array_equal(l, np.concatenate([x[1] for x in outs]))
        assert _same_f32(c, np.concatenate([x[2] for x in outs]))
        assert np.array_equal(b_.demand(), a.demand())
        b_.reset_demand()


def test_tunings_are_bit_identical(gr, gpu):
    """Every schedule parameter choice routes identically; the autotuner keeps one of them."""
    X, Y, Z = 256, 256, 10
    cap = gr.gen_grid(2, 4, X, Y, Z)
    pfx, pins = gr.gen_nets(2, 4, X, Y, Z, 5000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    ref = None
    for t in ({}, dict(large_threads=256, group_max=32), dict(large_threads=512, fork=0, refresh_blocks=333),
              dict(lpat_wave=1, group_max=64, large_threads=1024, fork=1), dict(fork=2),
              dict(fork=2, large_threads=512, group_max=32), dict(fork=2, large_threads=256, group_max=64),
              dict(fork=3, large_threads=1024, group_max=32), dict(fork=3, large_threads=256, group_max=32),
              dict(fork=1, large_threads=1024, group_max=64), dict(fork=3, lpat_block=64),
              dict(fork=3, lpat_block=128, group_max=32), dict(fork=1, lpat_block=256)):
        r.set_tuning(**t)
        r.reset_demand()
        r.route_schedule()
        out = [np.asarray(a).view(np.int32) for a in r.schedule_outputs()] + [r.demand()]
        if ref is None:
            ref = out
        for a, b in zip(out, ref):
            assert np.array_equal(a, b), t
    best, ms = r.autotune(reps=1)
    assert ms > 0 and best.large_threads in (256, 512, 1024)
    r.route_schedule()
    out = [np.asarray(a).view(np.int32) for a in r.schedule_outputs()] + [r.demand()]
    for a, b in zip(out, ref):
        assert np.array_equal(a, b)
