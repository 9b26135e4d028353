"""GPU tests of the learned proposal: MFMA forward and gradients against the
numpy CPU reference, features, top-k costing against the oracle, and what the
trained network gains over the untrained one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _router(gr, X=64, Y=64, Z=10, kind=2, seed=3, nets=1500):
    cap = gr.gen_grid(kind, seed, X, Y, Z)
    pfx, pins = gr.gen_nets(kind, seed, X, Y, Z, nets)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    return r, cap, pfx, pins, ofst, ids


def _rand_params(m, rng, scale=1.0):
    p = rng.standard_normal(m.n_params).astype(np.float32) * 0.05 * scale
    return p


def test_forward_matches_cpu(gr, gpu):
    from oracle import mlp_ref
    r, *_ = _router(gr)
    m = gr.Model(r, hidden=256, seed=1)
    rng = np.random.default_rng(0)
    p = _rand_params(m, rng)
    m.set_params(p)
    x = rng.standard_normal((300, m.k_in)).astype(np.float32)
    got = m.forward(x)
    want, _ = mlp_ref.forward(mlp_ref.unpack(p, r.Z), x)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 2e-2 * scale, np.abs(got - want).max()
    assert np.abs(got - want).mean() <= 2e-3 * scale


def test_gradient_matches_cpu(gr, gpu):
    from oracle import mlp_ref
    r, *_ = _router(gr)
    m = gr.Model(r, hidden=256, seed=2)
    rng = np.random.default_rng(1)
    p = _rand_params(m, rng, 2.0)
    m.set_params(p)
    x = rng.standard_normal((100, m.k_in)).astype(np.float32)
    y = rng.integers(0, m.n_out, 100).astype(np.int32)
    g, loss = m.gradient(x, y)
    gw, lw = mlp_ref.gradient(mlp_ref.unpack(p, r.Z), x, y, m.n_out)
    assert abs(loss - lw) <= 1e-3 * max(1.0, abs(lw))
    G, W = mlp_ref.unpack(g, r.Z), mlp_ref.unpack(gw, r.Z)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        err = np.linalg.norm(G[k] - W[k]) / max(np.linalg.norm(W[k]), 1e-12)
        assert err < 3e-2, (k, err)


def test_features_match_cpu(gr, orc, gpu):
    from oracle import mlp_ref
    r, cap, pfx, pins, ofst, ids = _router(gr, X=96, Y=80, nets=2500)
    m = gr.Model(r, seed=0)
    for b in range(3):  # some congestion first
        r.route_batch(ids[ofst[b]:ofst[b + 1]])
    rem = cap - r.demand()
    rng = np.random.default_rng(4)
    src = rng.integers(0, len(pins), 200)
    dst = rng.integers(0, len(pins), 200)
    s, d = pins[src], pins[dst]
    got = m.features(s, d)
    want = mlp_ref.features(r.X, r.Y, r.Z, rem, r.lut(), s, d)
    # equal up to one bf16 ulp (host logf vs numpy log of the same LUT entry)
    assert np.all(np.abs(got - want) <= np.abs(want) * 2.0 ** -7 + 1e-6)


# Z = 3..11 route through the fused kernel (one instantiation per Z); Z = 12
# scores 66 patterns and takes the feature / GEMM / top-k kernel sequence
@pytest.mark.parametrize("Z", [4, 10, 11, 12])
def test_always_fallback_is_exhaustive(gr, gpu, Z):
    a, cap, pfx, pins, ofst, ids = _router(gr, X=96, Y=96, Z=Z, nets=2500)
    b = gr.Router(96, 96, Z, cap)
    b.set_nets(pfx, pins)
    m = gr.Model(b, seed=5)
    b.set_proposal(m, k=4, threshold=2.0)  # top-1 probability never reaches 2
    for i in range(len(ofst) - 1):
        e1 = a.route_batch(ids[ofst[i]:ofst[i + 1]])
        e2 = b.route_batch(ids[ofst[i]:ofst[i + 1]])
        for u, v in zip(e1, e2):
            assert np.array_equal(np.asarray(u).view(np.int32), np.asarray(v).view(np.int32))
    assert np.array_equal(a.demand(), b.demand())
    pairs, fb = b.proposal_stats()
    assert pairs > 0 and fb == pairs


@pytest.mark.parametrize("Z", [4, 10, 11, 12])
def test_topk_choice_matches_oracle(gr, orc, gpu, Z):
    """With no fallback, each pair's pattern is the (cost, p) minimum over the
    model's top-k candidates, costed by the oracle at the same state."""
    X, Y = 96, 96
    r, cap, pfx, pins, ofst, ids = _router(gr, X=X, Y=Y, Z=Z, nets=2500)
    m = gr.Model(r, seed=7)
    rng = np.random.default_rng(3)
    m.set_params(_rand_params(m, rng, 4.0))
    o = orc.Oracle(X, Y, Z, cap)
    k = {4: 2, 10: 3, 11: 8, 12: 5}[Z]
    r.set_proposal(m, k=k, threshold=0.0)
    Z2 = (Z + 1) // 2
    # several batches: the later ones gather their features from the per-cell
    # copy the refresh keeps, which must equal the state the oracle sees
    for b in range(4):
        batch = ids[ofst[b]:ofst[b + 1]]
        src, dst = [], []
        for nid in batch:
            p = pins[pfx[nid]:pfx[nid + 1]]
            for j in range(1, len(p)):
                for kk in range(j):
                    src.append(p[j]); dst.append(p[kk])
        logits = m.forward(m.features(np.array(src), np.array(dst)))[:, :m.n_out]
        o.set_rem(cap - r.demand())
        r.route_batch(batch)
        cost, pat = r.pair_costs()
        for i in range(len(src)):
            order = sorted(range(m.n_out), key=lambda q: (-logits[i, q], q))[:k]
            best = None
            for q in order:
                c = o.pattern_cost(src[i], dst[i], q + Z2)
                if c is not None and (best is None or (np.float32(c), q + Z2) < best):
                    best = (np.float32(c), q + Z2)
            if best is None:
                pp, cc = o.best_pattern(src[i], dst[i])
                best = (np.float32(cc), pp)
            assert (pat[i], cost[i]) == (best[1], best[0]), i
    r.set_proposal(None)


def test_training_gains(gr, gpu):
    """Samples from exhaustive routing of 4 seeds; the trained model routes a
    held-out seed with top-4 candidates closer to exhaustive than the
    untrained model, and its loss falls."""
    X, Y, Z, nets = 200, 200, 10, 6000
    cap = gr.gen_grid(2, 0, X, Y, Z)
    r = gr.Router(X, Y, Z, cap)
    m = gr.Model(r, seed=11, max_samples=200000)

    def load(seed):
        pfx, pins = gr.gen_nets(2, seed, X, Y, Z, nets)
        ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
        r.set_nets(pfx, pins)
        r.set_schedule(ofst, ids)

    def route_pass():
        r.reset_demand()
        r.route_schedule()
        r.sync()
        return gr.route_cost(cap, r.demand())["total"]

    m.collect(4000)
    for seed in (1, 2, 3, 4):
        load(seed)
        route_pass()
    m.collect(0)
    assert m.samples() > 20000
    load(0)
    exhaustive = route_pass()
    r.set_proposal(m, k=4, threshold=0.0)
    untrained = route_pass()
    r.set_proposal(None)
    l0 = m.train(epochs=1, batch=512, lr=0.05, seed=1)
    l1 = m.train(epochs=6, batch=512, lr=0.05, seed=2)
    print(f"loss {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < l0
    r.set_proposal(m, k=4, threshold=0.0)
    trained = route_pass()
    r.set_proposal(None)
    gap_u = (untrained - exhaustive) / exhaustive
    gap_t = (trained - exhaustive) / exhaustive
    print(f"route cost exhaustive {exhaustive} untrained {untrained} ({gap_u:.4%}) trained {trained} ({gap_t:.4%})",
          flush=True)
    assert gap_t < gap_u


def test_model_outlives_router(gr, gpu):
    """A router destroyed first detaches its models: their calls then fail
    cleanly and closing them frees only their own buffers."""
    r, *_ = _router(gr)
    m = gr.Model(r, seed=0)
    r.close()
    with pytest.raises(gr.RouterError):
        m.params()
    with pytest.raises(gr.RouterError):
        m.forward(np.zeros((4, m.k_in), np.float32))
    m.close()


def test_proposal_graph_replay_matches_route_batch(gr, gpu):
    """Whole-schedule passes with the proposal (captured once, then replayed)
    route exactly as batch-by-batch calls; changing k re-captures."""
    X, Y, Z = 96, 96, 10
    r, cap, pfx, pins, ofst, ids = _router(gr, X=X, Y=Y, nets=2500)
    m = gr.Model(r, seed=7)
    m.set_params(_rand_params(m, np.random.default_rng(5), 4.0))
    b_ = gr.Router(X, Y, Z, cap)
    b_.set_nets(pfx, pins)
    b_.set_schedule(ofst, ids)
    mb = gr.Model(b_, seed=7)
    mb.set_params(m.params())
    for k in (4, 2):
        r.set_proposal(m, k=k, threshold=0.3)
        r.reset_demand()
        outs = [r.route_batch(ids[ofst[b]:ofst[b + 1]]) for b in range(len(ofst) - 1)]
        b_.set_proposal(mb, k=k, threshold=0.3)
        for rep in range(3):
            b_.reset_demand()
            b_.route_schedule()
            e, l, c = b_.schedule_outputs()
            assert np.array_equal(e, np.concatenate([x[0] for x in outs])), (k, rep)
            assert np.array_equal(l, np.concatenate([x[1] for x in outs])), (k, rep)
            assert np.array_equal(b_.demand(), r.demand()), (k, rep)
    r.set_proposal(None)
    b_.set_proposal(None)


def test_capture_under_proposal_keeps_fallback_labels(gr, gpu):
    """Capture while routing with the proposal keeps labels only where the pair
    fell back to every pattern: with every pair falling back (threshold 2) it
    equals capture from exhaustive routing, so training on either gives the
    same parameters bit for bit; with no low-confidence fallback almost no
    sample is labelled."""
    X, Y, Z = 96, 96, 10
    a, cap, pfx, pins, ofst, ids = _router(gr, X=X, Y=Y, nets=2500)
    b = gr.Router(X, Y, Z, cap)
    b.set_nets(pfx, pins)
    ma = gr.Model(a, seed=3, max_samples=1 << 15)
    mb = gr.Model(b, seed=3, max_samples=1 << 15)
    ma.collect(200)
    mb.collect(200)
    b.set_proposal(mb, k=4, threshold=2.0)
    for i in range(len(ofst) - 1):
        a.route_batch(ids[ofst[i]:ofst[i + 1]])
        b.route_batch(ids[ofst[i]:ofst[i + 1]])
    ma.collect(0)
    mb.collect(0)
    assert ma.samples() == mb.samples() > 256
    la = ma.train(epochs=1, batch=64, lr=0.05, seed=1)
    lb = mb.train(epochs=1, batch=64, lr=0.05, seed=1)
    assert abs(la - lb) <= 1e-6 * abs(la)  # the reported loss is summed with atomics
    assert np.array_equal(ma.params(), mb.params())
    # confident proposal: labels only on the rare pairs without a valid candidate
    mb.clear_samples()
    mb.collect(200)
    b.set_proposal(mb, k=4, threshold=0.0)
    b.proposal_stats()  # counters from here on
    b.reset_demand()
    for i in range(len(ofst) - 1):
        b.route_batch(ids[ofst[i]:ofst[i + 1]])
    mb.collect(0)
    pairs, fb = b.proposal_stats()
    assert mb.samples() > 256
    if fb < 32:
        with pytest.raises(gr.RouterError):
            mb.train(epochs=1, batch=32, lr=0.05, seed=1)
    b.set_proposal(None)


def test_always_fallback_full_workload(gr, gpu):
    """800x800x10, 150,000 nets requested, whole passes as graph replays:
    through the fused kernel with every pair falling back (threshold 2) the
    routes and demand equal plain routing bit for bit, on every replay."""
    X, Y, Z = 800, 800, 10
    cap = gr.gen_grid(2, 0, X, Y, Z)
    pfx, pins = gr.gen_nets(2, 0, X, Y, Z, 150000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.route_schedule()
    r.sync()
    e0, l0, c0 = r.schedule_outputs()
    d0 = r.demand()
    m = gr.Model(r, seed=1)
    r.set_proposal(m, k=4, threshold=2.0)
    r.proposal_stats()
    for rep in range(2):
        r.reset_demand()
        r.route_schedule()
        r.sync()
        e, l, c = r.schedule_outputs()
        assert np.array_equal(e, e0) and np.array_equal(l, l0), rep
        assert np.array_equal(np.asarray(c).view(np.int32), np.asarray(c0).view(np.int32)), rep
        assert np.array_equal(r.demand(), d0), rep
    pairs, fb = r.proposal_stats()
    assert pairs == 2 * int((np.diff(pfx) * (np.diff(pfx) - 1) // 2).sum()) and fb == pairs
    r.set_proposal(None)


def _sampled_topk_check(r, m, o, cap, pfx, pins, ofst, ids, b, k, rng, n_sample=3000, thr=0.6):
    """Route batch b (the router holds the state before it, synced) and check
    sampled pairs against the oracle's costing at that state: the (cost, p)
    minimum over the model's top-k candidates when the top-1 softmax
    probability is at least thr, else the exhaustive minimum.  Returns the
    number of sampled pairs of each kind (top-k, fell back)."""
    Z2 = (r.Z + 1) // 2
    o.set_rem(cap - r.demand())
    batch = ids[ofst[b]:ofst[b + 1]]
    deg = np.diff(pfx)[batch]
    P = deg * (deg - 1) // 2
    first = np.concatenate([[0], np.cumsum(P)])
    tot = int(first[-1])
    sample = np.sort(rng.choice(tot, size=min(tot, n_sample), replace=False))
    owner = np.searchsorted(first, sample, side="right") - 1
    src, dst = [], []
    for w, q in zip(sample, owner):
        o_ = int(w - first[q])
        j = int((1 + np.sqrt(1 + 8 * o_)) // 2)
        while j * (j - 1) // 2 > o_:
            j -= 1
        while (j + 1) * j // 2 <= o_:
            j += 1
        kk = o_ - j * (j - 1) // 2
        p = pins[pfx[batch[q]]:pfx[batch[q] + 1]]
        src.append(int(p[j])); dst.append(int(p[kk]))
    logits = m.forward(m.features(np.array(src, np.int32), np.array(dst, np.int32)))[:, :m.n_out]
    r.route_schedule(b, 1)
    cost, pat = r.pair_costs()
    seen_topk = seen_fb = 0
    for i, w in enumerate(sample):
        lg = logits[i].astype(np.float64)
        p1 = 1.0 / np.exp(lg - lg.max()).sum()
        order = sorted(range(m.n_out), key=lambda q: (-logits[i, q], q))[:k]
        top = None
        for q in order:
            c = o.pattern_cost(src[i], dst[i], q + Z2)
            if c is not None and (top is None or (np.float32(c), q + Z2) < top):
                top = (np.float32(c), q + Z2)
        pp, cc = o.best_pattern(src[i], dst[i])
        full = (np.float32(cc), pp)
        got = (np.float32(cost[w]), int(pat[w]))
        if abs(p1 - thr) < 1e-4:
            assert got in (top, full), (k, b, i)
        elif p1 < thr or top is None:
            assert got == full, (k, b, i, got, full)
            seen_fb += 1
        else:
            assert got == top, (k, b, i, got, top)
            seen_topk += 1
    return seen_topk, seen_fb


def test_trained_topk_full_workload_sampled(gr, orc, gpu):
    """The learned workload at its full size: 800x800x10, 150,000 nets
    requested, a model trained on exhaustive routing of two other seeds,
    top-k in {4, 8} with the 0.6 confidence threshold.  On sampled pairs of
    a bulk, a middle and a tail batch, each pair's (cost, pattern) equals the
    oracle's costing at the GPU's own state before the batch: the (cost, p)
    minimum over the model's top-k candidates when the top-1 softmax
    probability is at least 0.6, else the exhaustive minimum.  Both kinds of
    pair must occur."""
    X, Y, Z = 800, 800, 10
    cap = gr.gen_grid(2, 0, X, Y, Z)
    r = gr.Router(X, Y, Z, cap)
    m = gr.Model(r, hidden=256, seed=0, max_samples=1 << 18)
    m.collect(1000)
    for seed in (1, 2):
        pf, pn = gr.gen_nets(2, seed, X, Y, Z, 150000)
        of, ii = gr.partition_nets(X, Y, Z, pf, pn, 4)
        r.set_nets(pf, pn)
        r.set_schedule(of, ii)
        r.reset_demand()
        r.route_schedule()
        r.sync()
    m.collect(0)
    for e in range(3):
        m.train(epochs=1, batch=1024, lr=0.05, seed=e)
    pfx, pins = gr.gen_nets(2, 0, X, Y, Z, 150000)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    o = orc.Oracle(X, Y, Z, cap)
    rng = np.random.default_rng(9)
    nb = len(ofst) - 1
    seen_fb = seen_topk = 0
    for k in (4, 8):
        r.set_proposal(m, k=k, threshold=0.6)
        for b in (0, nb // 3, nb - 3):
            print(f"trained top-{k} batch {b}", flush=True)
            r.reset_demand()
            if b:
                r.route_schedule(0, b)
            r.sync()
            t, f = _sampled_topk_check(r, m, o, cap, pfx, pins, ofst, ids, b, k, rng)
            seen_topk += t
            seen_fb += f
    r.set_proposal(None)
    print(f"sampled pairs: {seen_topk} from top-k, {seen_fb} fell back", flush=True)
    assert seen_fb > 0 and seen_topk > 0
    m.close()


def test_stress_online_finetune_sampled(gr, orc, gpu):
    """The combined-stress workload at its full size (2000x2000x10,
    capacities cut 40% plus 30 hotspots, 1.5M nets requested) with the
    online fine-tuning of the learned pass: a model trained on exhaustive
    routing of another seed routes batches 0..9 with k = 4 and threshold
    0.6 while it captures fallback-judged samples, is fine-tuned on them
    for one epoch (as the learned pass does every 50 batches), and then
    routes batches 10 and 11.  The fine-tuning must change the model, and
    on sampled pairs of batches 10 and 11 (bulk batches, where both kinds
    of pair occur; later batches fall back almost entirely) each
    pair's (cost, pattern) equals the oracle's costing with the fine-tuned
    model at the GPU's own state before the batch."""
    W = gr.WORKLOADS[5]
    X, Y, Z = W["X"], W["Y"], W["Z"]
    cap = gr.gen_grid(5, 0, X, Y, Z)
    r = gr.Router(X, Y, Z, cap)
    m = gr.Model(r, hidden=256, seed=0, max_samples=1 << 20)
    m.collect(500)
    pf, pn = gr.gen_nets(5, 1, X, Y, Z, W["nets"])
    of, ii = gr.partition_nets(X, Y, Z, pf, pn, 4)
    r.set_nets(pf, pn)
    r.set_schedule(of, ii)
    r.reset_demand()
    r.route_schedule()
    r.sync()
    m.collect(0)
    for e in range(2):
        m.train(epochs=1, batch=1024, lr=0.05, seed=e)
    m.clear_samples()
    pfx, pins = gr.gen_nets(5, 0, X, Y, Z, W["nets"])
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    assert len(ofst) - 1 > 12
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.set_proposal(m, k=4, threshold=0.6)
    m.collect(500)
    r.reset_demand()
    print("stress: batches 0..9 with sample capture", flush=True)
    r.route_schedule(0, 10)
    r.sync()
    before = m.params().copy()
    m.train(epochs=1, batch=1024, lr=0.02, seed=1000)
    m.clear_samples()
    m.collect(0)
    assert not np.array_equal(m.params(), before), "fine-tuning left the model unchanged"
    o = orc.Oracle(X, Y, Z, cap)
    rng = np.random.default_rng(11)
    seen_topk = seen_fb = 0
    for b in (10, 11):
        print(f"stress fine-tuned, batch {b}", flush=True)
        r.sync()
        t, f = _sampled_topk_check(r, m, o, cap, pfx, pins, ofst, ids, b, 4, rng, n_sample=6000)
        seen_topk += t
        seen_fb += f
    r.set_proposal(None)
    print(f"sampled pairs: {seen_topk} from top-k, {seen_fb} fell back", flush=True)
    assert seen_fb > 0 and seen_topk > 0
    m.close()
    r.close()
