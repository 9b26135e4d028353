"""fp32-capacity grids (capacities as read from .cap files).

CPU: the oracle's e^x against the math library, its cost function against the
integer-capacity table, the fp32 grid against the integer grid on integral
capacities, and the remaining capacity against sequential fp32 decrements.
GPU: the HIP path on fractional capacities against the oracle, bit-exact.
"""
import math

import numpy as np
import pytest


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _frac_grid(gr, kind, seed, X, Y, Z):
    """Integer workload capacities plus fractional parts that are not exact in fp32."""
    cap = gr.gen_grid(kind, seed, X, Y, Z).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    frac = rng.choice([0.0, 0.1, 0.25, 0.3, 0.5, 0.7, 0.9, 0.05], size=cap.size)
    return (cap + frac).astype(np.float32)


def test_cw_exp_accuracy(orc):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-105, 88, 20000), np.arange(-100, 88, 0.5), [0.0, -0.0, 1e-300]])
    worst = 0.0
    for x in xs:
        a, b = orc.cw_exp(x), math.exp(x)
        worst = max(worst, abs(a - b) / b)
    assert worst <= 4 * 2.0 ** -53, worst
    assert orc.cw_exp(0.0) == 1.0
    assert orc.cw_exp(200.0) == math.inf and orc.cw_exp(-200.0) == 0.0


def test_cost_f32_equals_table_on_integers(orc):
    for model in (0, 1, 2):
        table = orc.lut(model)
        got = np.array([orc.cost_f32(model, r) for r in range(-256, 256)], np.float32)
        assert _same_f32(got, table), model
    # saturation and the incremental model's fp32 r - 1
    assert orc.cost_f32(0, -400.0) == np.float32(1e30)
    r = np.float32(0.3)
    want = np.float32(math.exp(-0.5 * float(r - np.float32(1.0)))) - np.float32(math.exp(-0.5 * float(r)))
    assert abs(float(orc.cost_f32(1, r)) - float(want)) <= 2 * float(np.spacing(np.float32(want)))


def test_float_grid_equals_int_grid_on_integral_caps(gr, orc):
    X, Y, Z = 64, 56, 6
    cap = gr.gen_grid(1, 4, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 4, X, Y, Z, 900)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    for model in (0, 1):
        a = orc.Oracle(X, Y, Z, cap, model, 0)
        b = orc.Oracle(X, Y, Z, cap.astype(np.float32), model, 0)
        assert _same_f32(a.prefix(), b.prefix())
        for o in (a, b):
            o.set_nets(pfx, pins)
        for k in range(len(ofst) - 1):
            batch = ids[ofst[k]:ofst[k + 1]]
            e1, l1, c1 = a.route_batch(batch)
            e2, l2, c2 = b.route_batch(batch)
            assert np.array_equal(e1, e2) and np.array_equal(l1, l2) and _same_f32(c1, c2), (model, k)
        assert np.array_equal(a.demand(), b.demand())
        assert _same_f32(a.prefix(), b.prefix())
        assert _same_f32(b.remaining_f32(), (cap - a.demand()).astype(np.float32))


def test_float_remaining_is_sequential_fp32_decrement(gr, orc):
    X, Y, Z = 48, 40, 6
    cap = _frac_grid(gr, 1, 2, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 2, X, Y, Z, 600)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    for k in range(len(ofst) - 1):
        o.route_batch(ids[ofst[k]:ofst[k + 1]])
    d = o.demand()
    assert d.max() >= 2
    want = cap.copy()
    for _ in range(int(d.max())):
        m = d > 0
        want[m] = want[m] - np.float32(1.0)
        d = d - m
    assert _same_f32(o.remaining_f32(), want)


@pytest.mark.gpu
def test_float_grid_parity(gr, orc, gpu):
    """Fractional capacities: initial prefix, every batch's pair costs, trees,
    patterns, net costs, demand and fp32 remaining capacity bit-exact."""
    for model, (X, Y, Z, n) in ((0, (96, 80, 6, 1500)), (1, (70, 90, 7, 1200)), (2, (64, 64, 10, 800))):
        cap = _frac_grid(gr, 2, model, X, Y, Z)
        pfx, pins = gr.gen_nets(2, model, X, Y, Z, n)
        ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
        r = gr.Router(X, Y, Z, cap, model, 0)
        o = orc.Oracle(X, Y, Z, cap, model, 0)
        assert _same_f32(r.prefix(), o.prefix()), model
        r.set_nets(pfx, pins)
        o.set_nets(pfx, pins)
        for k in range(len(ofst) - 1):
            print(f"fp32 grid model {model} batch {k}", flush=True)
            batch = ids[ofst[k]:ofst[k + 1]]
            e1, l1, c1 = r.route_batch(batch)
            e2, l2, c2 = o.route_batch(batch)
            pc1, pp1 = r.pair_costs()
            pc2, pp2 = o.pair_costs()
            assert _same_f32(pc1, pc2) and np.array_equal(pp1, pp2), (model, k)
            assert np.array_equal(e1, e2) and np.array_equal(l1, l2) and _same_f32(c1, c2), (model, k)
        assert np.array_equal(r.demand(), o.demand())
        assert _same_f32(r.remaining_f32(), o.remaining_f32())
        assert _same_f32(r.prefix(), o.prefix())
        r.reset_demand()
        assert not r.demand().any() and _same_f32(r.remaining_f32(), cap)


@pytest.mark.gpu
def test_float_grid_wide_range_equals_int_grid(gr, orc, gpu):
    """Integral capacities up to 230 (costs down to fp32 subnormals) and
    negative ones: the fp32 grid and the integer grid give the same bits, and
    both equal the oracle."""
    X, Y, Z = 80, 72, 6
    rng = np.random.default_rng(5)
    cap = rng.integers(-3, 231, X * Y * Z).astype(np.int32)
    pfx, pins = gr.gen_nets(1, 6, X, Y, Z, 1200)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    a = gr.Router(X, Y, Z, cap)
    b = gr.Router(X, Y, Z, cap.astype(np.float32))
    o = orc.Oracle(X, Y, Z, cap)
    assert _same_f32(a.prefix(), o.prefix()) and _same_f32(b.prefix(), o.prefix())
    for x in (a, b, o):
        x.set_nets(pfx, pins)
    for k in range(len(ofst) - 1):
        batch = ids[ofst[k]:ofst[k + 1]]
        ra, rb, ro = a.route_batch(batch), b.route_batch(batch), o.route_batch(batch)
        for u, v, w in zip(ra, rb, ro):
            assert _same_f32(u, w) if u.dtype == np.float32 else np.array_equal(u, w), k
            assert _same_f32(v, w) if v.dtype == np.float32 else np.array_equal(v, w), k
    assert np.array_equal(a.demand(), o.demand()) and np.array_equal(b.demand(), o.demand())
    assert _same_f32(a.prefix(), o.prefix()) and _same_f32(b.prefix(), o.prefix())
