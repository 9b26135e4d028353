"""Route evaluator on the GPU against the oracle's restatement: per-cell wire
and via counts bit-exact, integer layer figures exact, maxima exact, sums of
fp64 terms within 1e-12 relative (the device adds them in another order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXACT = ("wire", "via", "max_wire", "max_via", "max_cap", "max_overflow", "demand")
REL = ("of_cost", "overflow")


def _check(gr, orc, r, X, Y, Z, cap, pfx, pins, ofst, ids):
    layers, tot = r.evaluate()
    w, v = r.eval_counts()
    E, L, _ = r.schedule_outputs()
    capf = np.asarray(cap, np.float32)
    w2, v2, ol = orc.evaluate(X, Y, Z, capf, r.demand(), ofst, ids, pfx, pins, E, L)
    assert np.array_equal(w, w2) and np.array_equal(v, v2)
    for z in range(Z):
        for k in EXACT:
            assert layers[z][k] == ol[z][k], (z, k, layers[z][k], ol[z][k])
        for k in REL:
            assert abs(layers[z][k] - ol[z][k]) <= 1e-12 * max(1.0, abs(ol[z][k])), (z, k)
    assert tot["wire"] == sum(ol[z]["wire"] for z in range(1, Z))
    assert tot["via"] == sum(ol[z]["via"] for z in range(1, Z))
    assert tot["demand"] == int(r.demand().sum())
    return layers, tot


def test_eval_workload1(gr, orc, gpu):
    X, Y, Z = 128, 128, 6
    cap = gr.gen_grid(1, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 0, X, Y, Z, 5000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.route_schedule()
    _check(gr, orc, r, X, Y, Z, cap, pfx, pins, ofst, ids)
    # a second pass on the same context evaluates the same routes to the same figures
    a = r.evaluate()
    r.reset_demand()
    r.route_schedule()
    assert r.evaluate() == a


def test_eval_fp32_grid_and_overflow(gr, orc, gpu):
    """Fractional capacities, many cells over capacity."""
    X, Y, Z = 90, 70, 8
    rng = np.random.default_rng(2)
    cap = np.round(rng.uniform(0, 3, X * Y * Z), 1).astype(np.float32)
    pfx, pins = gr.gen_nets(2, 1, X, Y, Z, 3000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.route_schedule()
    layers, tot = _check(gr, orc, r, X, Y, Z, cap, pfx, pins, ofst, ids)
    assert tot["overflow"] > 0 and max(l["max_overflow"] for l in layers) > 0


def test_eval_workload2(gr, orc, gpu):
    """800x800x10, 150,000 nets: the full pass evaluated on the device."""
    X, Y, Z = 800, 800, 10
    cap = gr.gen_grid(2, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(2, 0, X, Y, Z, 150000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.route_schedule()
    print("workload2 routed, evaluating", flush=True)
    _check(gr, orc, r, X, Y, Z, cap, pfx, pins, ofst, ids)
