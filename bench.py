"""Benchmark: nets routed per second on one MI355X.

A step is one full routing pass over the workload: reset demand, then route
every batch of the schedule in order (L-pattern costing -> root-pair Boruvka
-> Fenwick commit -> cost refresh per batch).  The default workload is the
800x800x10 GCELL grid with 150,000 nets.  Inputs are resident in HBM before
the timed region, which replays each pass as one captured HIP graph.
Per-kernel durations come from HIP events recorded on the library's launch
stream around every launch, in an instrumented repeat of the same passes.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "nets routed/sec on 1 MI355X, learned+batched vs plain, at ≤0.5% route-cost gap"
HBM_PEAK_GBS = 8000.0
KNAMES = ["k_lpat", "k_tree", "k_refresh"]


def footprint(X, Y, Z, pfx, pins):
    """Per-net pin count n, pair count P and unique-cell footprint F of the L-pattern stage."""
    n_nets = len(pfx) - 1
    deg = np.diff(pfx)
    net = np.repeat(np.arange(n_nets, dtype=np.int64), deg)
    p = pins.astype(np.int64)
    z = p // (X * Y)
    r = p - z * X * Y
    odd = (z & 1) == 1
    x = np.where(odd, r // Y, r % X)
    y = np.where(odd, r % Y, r // X)
    xs = np.bincount(net[np.unique(net * X + x, return_index=True)[1]], minlength=n_nets)
    ys = np.bincount(net[np.unique(net * Y + y, return_index=True)[1]], minlength=n_nets)
    w = np.zeros(n_nets, np.int64)
    h = np.zeros(n_nets, np.int64)
    np.maximum.at(w, net, x); np.maximum.at(h, net, y)
    wmin = np.full(n_nets, 1 << 40, np.int64); hmin = np.full(n_nets, 1 << 40, np.int64)
    np.minimum.at(wmin, net, x); np.minimum.at(hmin, net, y)
    W = w - wmin + 1
    H = h - hmin + 1
    E = (Z - 1) // 2
    O = Z // 2
    F = Z * xs * ys + E * ys * (W - xs) + O * xs * (H - ys)
    P = deg * (deg - 1) // 2
    return deg, P, F


def cpu_baseline(X, Y, Z, cap, pfx, pins, ofst, ids, budget_s):
    """The oracle (C++, OpenMP over the nets of a batch, batches in order) on
    the host cores: whole passes on all threads, the median of up to 5 that
    fit in 80% of the budget; then one thread over the batches that fit in a
    fifth of the budget, with the all-thread rate over the same batches
    beside it."""
    import oracle
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count(), os.cpu_count())
    nb = len(ofst) - 1
    n_nets = len(pfx) - 1

    def one_pass(nthreads, stop_s=None):
        o = oracle.Oracle(X, Y, Z, cap)
        o.set_nets(pfx, pins)
        per = []
        t0 = time.perf_counter()
        for b in range(nb):
            tb = time.perf_counter()
            o.route_batch(ids[ofst[b]:ofst[b + 1]], nthreads=nthreads)
            per.append(time.perf_counter() - tb)
            if stop_s is not None and time.perf_counter() - t0 > stop_s:
                break
        return per

    # all threads: whole passes (the first one cut at half the budget: then
    # the batches it reached are the sample), up to 5 within 80% of the budget
    runs = [one_pass(threads, stop_s=budget_s / 2)]
    nb_s = len(runs[0])
    pass_s = sum(runs[0])
    while nb_s == nb and len(runs) < 5 and sum(sum(r) for r in runs) + pass_s < budget_s * 0.8:
        runs.append(one_pass(threads))
    times = sorted(sum(r) for r in runs)
    med = times[len(times) // 2]
    nets_s = int(ofst[nb_s])
    single = one_pass(1, stop_s=budget_s / 5)[:nb_s]
    k = len(single)
    nets_k = int(ofst[k])
    multi_k = sorted(sum(r[:k]) for r in runs)[len(runs) // 2]
    what = (f"whole pass, all {nb} batches ({n_nets} nets)" if nb_s == nb else
            f"batches 0..{nb_s - 1} ({nets_s} nets) of {nb}")
    return {"value": round(nets_s / med, 1), "unit": "nets/s", "cores": threads, "kind": "port",
            "sample": f"{what} of the same workload, C++ oracle with OpenMP over the nets of each batch; median of "
                      f"{len(runs)} run(s) ({', '.join(f'{t:.2f}' for t in times)} s)",
            "single_thread": {"value": round(nets_k / sum(single), 1), "unit": "nets/s", "cores": 1,
                              "sample": f"batches 0..{k - 1} ({nets_k} nets), one pass, {sum(single):.1f} s",
                              "all_threads_same_sample": round(nets_k / multi_k, 1)}}


def io_eval(gr, r, X, Y, Z, cap, pfx, pins, ofst, ids, cpu=True):
    """The path's callers either side, on the routes of the last timed pass:
    route evaluator (device) with the oracle's CPU evaluation beside it,
    route writer and design reader/saver (host, all threads and one thread).
    The reader and writer are host code in the reference too (one reader
    thread, read_data.hpp; four writer threads, writer.hpp)."""
    import tempfile
    out = {}
    # evaluator: per-layer report of main.cpp:103-135 on the device
    r.sync()
    r.evaluate()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        layers, tot = r.evaluate()
    ev_ms = (time.perf_counter() - t0) * 1e3 / reps
    cells = X * Y * Z
    E, L, _ = r.schedule_outputs()
    ev = {"ms": round(ev_ms, 3), "cells": cells, "slots": int(len(E)), "total": tot,
          "note": "gr_evaluate end to end (clears, one path launch per batch, layer reduction, D2H)"}
    if cpu:
        import oracle
        t0 = time.perf_counter()
        oracle.evaluate(X, Y, Z, np.asarray(cap, np.float32), r.demand(), ofst, ids, pfx, pins, E, L)
        ev["cpu_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        ev["cpu_kind"] = "port: oracle difference-array evaluator, 1 thread"
    out["evaluator"] = ev
    d = gr.Design(X, Y, Z, np.asarray(cap, np.float32), pfx, pins)
    with tempfile.TemporaryDirectory() as tmp:
        wr = {}
        for flags, name in ((0, "grid"), (gr.WRITE_MERGE | gr.WRITE_DBU, "merged_dbu")):
            path = os.path.join(tmp, f"routes_{name}.txt")
            for threads in (0, 1):
                t0 = time.perf_counter()
                nbytes = d.write_routes(path, ofst, ids, E, L, flags=flags, threads=threads)
                dt = time.perf_counter() - t0
                key = f"{name}_{'all' if threads == 0 else '1'}_threads"
                wr[key] = {"s": round(dt, 3), "MB": round(nbytes / 1e6, 1), "MBps": round(nbytes / 1e6 / dt, 1),
                           "nets_per_s": round(len(ids) / dt, 1)}
        out["writer"] = wr
        capf, netf = os.path.join(tmp, "d.cap"), os.path.join(tmp, "d.net")
        t0 = time.perf_counter()
        d.save(capf, netf)
        save_s = time.perf_counter() - t0
        size = os.path.getsize(capf) + os.path.getsize(netf)
        rd = {"MB": round(size / 1e6, 1), "save_s": round(save_s, 3)}
        for threads in (0, 1):
            t0 = time.perf_counter()
            e = gr.Design.read(capf, netf, 0, threads)
            dt = time.perf_counter() - t0
            rd[f"{'all' if threads == 0 else '1'}_threads"] = {"s": round(dt, 3), "MBps": round(size / 1e6 / dt, 1)}
            e.close()
        out["reader"] = rd
    out["threads_all"] = min(os.cpu_count() or 1, 16)
    return out


def mst_side(gr, r, cap, n_nets, args):
    """The pass with minimum spanning trees (GR_TREE_TRUE_MST): time and route
    cost beside the root-pair pass's; the context is left in root-pair mode
    holding a root-pair pass."""
    rp_cost = gr.route_cost(cap, r.demand())
    r.set_tree_mode(True)
    r.reset_demand()
    r.route_schedule()
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.reset_demand()
        r.route_schedule()
    r.sync()
    mst_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    mst_cost = gr.route_cost(cap, r.demand())
    out = {"ms_per_step": round(mst_ms, 4), "nets_per_s": round(n_nets / (mst_ms / 1e3), 1),
           "route_cost": mst_cost, "root_pair_route_cost": rp_cost,
           "route_cost_change": round((mst_cost["total"] - rp_cost["total"]) / rp_cost["total"], 5)}
    r.set_tree_mode(False)
    r.reset_demand()
    r.route_schedule()
    r.sync()
    print(f"# true-MST trees {out}", file=sys.stderr, flush=True)
    return out


def learned_eval(gr, r, X, Y, Z, cap, n_req, args, n_nets, plain_ms, seed0):
    """Learned proposal on the same generator: samples from exhaustive routing of
    seeds 1..train_seeds train the MLP on the GPU; seed 0 is then routed with the
    model's top-k candidates (fallback to full enumeration below 0.6 top-1
    probability) and judged against exhaustive costing of the same seed."""
    t0 = time.perf_counter()
    m = gr.Model(r, hidden=256, seed=0, max_samples=args.train_seeds * 70 * args.train_per_batch)
    m.collect(args.train_per_batch)
    for s in range(1, args.train_seeds + 1):
        pf, pn = gr.gen_nets(args.workload, s, X, Y, Z, n_req)
        of, ids = gr.partition_nets(X, Y, Z, pf, pn, 4)
        r.set_nets(pf, pn)
        r.set_schedule(of, ids)
        r.reset_demand()
        r.route_schedule()
        r.sync()
    m.collect(0)
    t1 = time.perf_counter()
    losses = [m.train(epochs=1, batch=1024, lr=0.05, seed=e) for e in range(args.train_epochs)]
    t2 = time.perf_counter()
    pfx, pins, ofst, ids = seed0
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.reset_demand()
    r.route_schedule()
    r.sync()
    exhaustive = gr.route_cost(cap, r.demand())
    P = float((np.diff(pfx) * (np.diff(pfx) - 1) // 2).sum())
    # algorithmic MLP FLOPs per scored pair: 2 (254*256 + 256*256 + 256*45) = 284,160 at Z = 10
    flop_pair = 2.0 * (m.n_in * m.hidden + m.hidden * m.hidden + m.hidden * m.n_out)
    out = {"train": {"seeds": args.train_seeds, "samples": int(m.samples()), "epochs": args.train_epochs,
                     "loss_per_epoch": [round(v, 4) for v in losses], "collect_s": round(t1 - t0, 2),
                     "train_s": round(t2 - t1, 2)},
           "exhaustive_route_cost": exhaustive}
    for k in (4, 8):
        r.set_proposal(m, k=k, threshold=0.6)
        r.reset_demand()
        r.route_schedule()  # warm-up; captures the pass as one graph
        r.sync()
        r.proposal_stats()
        steps = 3
        r.sync()
        ta = time.perf_counter()
        for _ in range(steps):
            r.reset_demand()
            r.route_schedule()
        r.sync()
        tb = time.perf_counter()
        pairs, fb = r.proposal_stats()
        cost = gr.route_cost(cap, r.demand())
        # costing kernel time: HIP events around each launch, in separate passes
        r.enable_timing(True)
        for _ in range(2):
            r.reset_demand()
            r.route_schedule()
        r.sync()
        n, ms = r.kernel_stats()
        r.enable_timing(False)
        r.set_proposal(None)
        ms_l = (tb - ta) * 1e3 / steps
        cost_ms = float(ms[0]) / 2  # features + MLP + top-k costing
        out[f"k{k}"] = {
            "nets_per_s": round(n_nets / (ms_l / 1e3), 1), "ms_per_step": round(ms_l, 3),
            "vs_plain": round(plain_ms / ms_l, 4), "route_cost": cost,
            "route_cost_gap": round((cost["total"] - exhaustive["total"]) / exhaustive["total"], 6),
            "fallback_rate": round(fb / max(pairs, 1), 4), "costing_ms_per_step": round(cost_ms, 3),
            "mlp_TFLOPs_lower_bound": round(P * flop_pair / (cost_ms / 1e3) / 1e12, 2),
            # the fused costing kernel against the dense bf16 MFMA peak: its MLP FLOPs over its whole time
            # (gather, top-k and candidate costing included), HIP events on the launch stream
            "roofline": {"kernel": "k_prop", "bound": "mfma",
                         "achieved": round(P * flop_pair / (cost_ms / 1e3) / 1e12, 2), "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": round(P * flop_pair / (cost_ms / 1e3) / 1e12 / 2500.0, 4)}}
    m.close()
    return out


def stress_eval(gr, args, tuning):
    """Combined stress (2000x2000x10, capacities cut 40% plus 30 Gaussian
    hotspots, 1.5M nets requested): the plain pass against the learned
    proposal (k=4, 0.6 confidence threshold), first trained on exhaustive
    routing of other seeds, then fine-tuned online every 50 batches on
    fallback-judged decisions (pairs whose choice came from every pattern).
    The learned pass is timed whole, fine-tuning included."""
    W = gr.WORKLOADS[5]
    X, Y, Z = W["X"], W["Y"], W["Z"]
    t = time.perf_counter()
    cap = gr.gen_grid(5, args.seed, X, Y, Z)
    pfx, pins = gr.gen_nets(5, args.seed, X, Y, Z, W["nets"])
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    n_nets, nb = len(pfx) - 1, len(ofst) - 1
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    if tuning:
        keys = ["large_threads", "group_max", "fork", "refresh_blocks", "lpat_wave", "lpat_block"]
        r.set_tuning(**{k: tuning[k] for k in keys})
    print(f"# stress setup {time.perf_counter() - t:.2f}s nets {n_nets} batches {nb}", file=sys.stderr, flush=True)
    r.reset_demand()
    r.route_schedule()  # warm-up; captures the pass
    r.sync()
    steps = 2
    ta = time.perf_counter()
    for _ in range(steps):
        r.reset_demand()
        r.route_schedule()
    r.sync()
    plain_ms = (time.perf_counter() - ta) * 1e3 / steps
    plain_cost = gr.route_cost(cap, r.demand())
    # initial model from exhaustive routing of seeds 1..stress_seeds
    t = time.perf_counter()
    m = gr.Model(r, hidden=256, seed=0, max_samples=1 << 20)
    m.collect(args.train_per_batch)
    for sd in range(1, args.stress_seeds + 1):
        pf, pn = gr.gen_nets(5, sd, X, Y, Z, W["nets"])
        of, ii = gr.partition_nets(X, Y, Z, pf, pn, 4)
        r.set_nets(pf, pn)
        r.set_schedule(of, ii)
        r.reset_demand()
        r.route_schedule()
        r.sync()
    m.collect(0)
    for e in range(args.train_epochs):
        m.train(epochs=1, batch=1024, lr=0.05, seed=e)
    m.clear_samples()
    pretrain_s = time.perf_counter() - t
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.set_proposal(m, k=4, threshold=0.6)
    m.collect(500)
    r.reset_demand()
    r.proposal_stats()
    r.sync()
    ta = time.perf_counter()
    rounds = skipped = 0
    for b0 in range(0, nb, 50):
        r.route_schedule(b0, min(50, nb - b0))
        if b0 + 50 < nb:  # fine-tune on this chunk's fallback-judged samples
            try:
                m.train(epochs=1, batch=1024, lr=0.02, seed=1000 + b0)
                rounds += 1
            except gr.RouterError as e:  # fewer labelled samples than one minibatch
                skipped += 1
                print(f"# stress fine-tune after batch {b0 + 50} skipped: {e}", file=sys.stderr, flush=True)
            m.clear_samples()
    r.sync()
    learned_ms = (time.perf_counter() - ta) * 1e3
    m.collect(0)
    pairs, fb = r.proposal_stats()
    learned_cost = gr.route_cost(cap, r.demand())
    r.set_proposal(None)
    m.close()
    r.close()
    return {
        "workload": f"{X}x{Y}x{Z} GCELLs, capacities cut 40% plus 30 hotspots, {W['nets']} nets requested "
                    f"(seed {args.seed}), {n_nets} routed, {nb} batches",
        "plain": {"nets_per_s": round(n_nets / (plain_ms / 1e3), 1), "ms_per_pass": round(plain_ms, 3),
                  "route_cost": plain_cost},
        "learned_online": {
            "nets_per_s": round(n_nets / (learned_ms / 1e3), 1), "ms_per_pass": round(learned_ms, 3),
            "vs_plain": round(plain_ms / learned_ms, 4), "route_cost": learned_cost,
            "route_cost_gap": round((learned_cost["total"] - plain_cost["total"]) / plain_cost["total"], 6),
            "fallback_rate": round(fb / max(pairs, 1), 4), "finetune_rounds": rounds,
            "finetune_skipped": skipped, "finetune_points": len(range(50, nb, 50)),
            "pretrain": {"seeds": args.stress_seeds, "s": round(pretrain_s, 2)}},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=50.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no per-kernel HIP events")
    ap.add_argument("--learned", type=int, default=1, help="also evaluate the learned proposal (0: skip)")
    ap.add_argument("--fused", type=int, default=1, help="also time the one-launch-per-batch form (0: skip)")
    ap.add_argument("--autotune", type=int, default=1, help="pick schedule parameters by timing (0: defaults)")
    ap.add_argument("--tuning", default="",
                    help="explicit large_threads,group_max,fork,refresh_blocks,lpat_wave[,lpat_block]")
    ap.add_argument("--io", type=int, default=1, help="also measure evaluator, writer and reader (0: skip)")
    ap.add_argument("--mst", type=int, default=1, help="also time true-MST trees for route quality (0: skip)")
    ap.add_argument("--stress", type=int, default=1,
                    help="also run the combined-stress workload (learned proposal with online fine-tuning; 0: skip)")
    ap.add_argument("--stress-seeds", type=int, default=2, help="seeds of exhaustive routing for the stress model")
    ap.add_argument("--train-seeds", type=int, default=20)
    ap.add_argument("--train-per-batch", type=int, default=1500)
    ap.add_argument("--train-epochs", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the routing path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import iitmglobalrouter_amd as gr
    W = gr.WORKLOADS[args.workload]
    X, Y, Z = W["X"], W["Y"], W["Z"]
    t = time.perf_counter()
    cap = gr.gen_grid(args.workload, args.seed, X, Y, Z)
    pfx, pins = gr.gen_nets(args.workload, args.seed, X, Y, Z, W["nets"])
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    print(f"# setup {time.perf_counter() - t:.2f}s nets {len(pfx) - 1} batches {len(ofst) - 1}",
          file=sys.stderr, flush=True)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    n_nets = len(pfx) - 1
    tuning = None
    if args.tuning:
        keys = ["large_threads", "group_max", "fork", "refresh_blocks", "lpat_wave", "lpat_block"]
        r.set_tuning(**dict(zip(keys, (int(v) for v in args.tuning.split(",")))))
        tuning = r.tuning().as_dict()
    elif args.autotune:
        t, tms = r.autotune(reps=2)
        tuning = dict(t.as_dict(), autotune_ms_per_pass=round(tms, 3))
        print(f"# autotuned {tuning}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        r.reset_demand()
        r.route_schedule()
        r.sync()
    print("# warmup done", file=sys.stderr, flush=True)
    # timed region: the whole pass replays as one captured HIP graph per step
    r.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.reset_demand()
        r.route_schedule()
    r.sync()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = n_nets / ((t1 - t0) / args.steps)
    print(f"# timed {args.steps} steps: {ms_step:.3f} ms/step", file=sys.stderr, flush=True)
    # the same passes launched kernel by kernel (no graph): the per-kernel-launch
    # form the captured graph replaces (a pass split in two ranges is not replayed)
    nb_all = len(ofst) - 1
    r.sync()
    te = time.perf_counter()
    for _ in range(args.steps):
        r.reset_demand()
        r.route_schedule(0, nb_all - 1)
        r.route_schedule(nb_all - 1, 1)
    r.sync()
    eager_ms = (time.perf_counter() - te) * 1e3 / args.steps
    print(f"# per-kernel launches: {eager_ms:.3f} ms/step", file=sys.stderr, flush=True)
    # the one-launch-per-batch form (tuning fork 4: costing, trees and commit
    # in one launch per batch, the refresh a second) over the same passes, as
    # graph replays and kernel by kernel; its demand must equal the pass above
    fused = None
    if args.fused:
        tkeys = ["large_threads", "group_max", "fork", "refresh_blocks", "lpat_wave", "lpat_block"]
        base = {k: r.tuning().as_dict()[k] for k in tkeys}
        d_ref = r.demand()
        form = "one launch per batch for costing, trees and commit (tuning fork 4), plus the refresh"
        try:
            r.set_tuning(**dict(base, fork=4))
            r.reset_demand()
            r.route_schedule()
            r.sync()  # raises if a fused launch's bounded wait timed out
            same = bool(np.array_equal(r.demand(), d_ref))
            tf = time.perf_counter()
            for _ in range(args.steps):
                r.reset_demand()
                r.route_schedule()
            r.sync()
            f_ms = (time.perf_counter() - tf) * 1e3 / args.steps
            tf = time.perf_counter()
            for _ in range(args.steps):
                r.reset_demand()
                r.route_schedule(0, nb_all - 1)
                r.route_schedule(nb_all - 1, 1)
            r.sync()
            fe_ms = (time.perf_counter() - tf) * 1e3 / args.steps
            fused = {"form": form, "ms_per_step": round(f_ms, 4), "nets_per_s": round(n_nets / (f_ms / 1e3), 1),
                     "per_kernel_launches_ms_per_step": round(fe_ms, 4), "graph_speedup": round(fe_ms / f_ms, 3),
                     "same_demand_as_default": same}
            print(f"# fused form: {f_ms:.3f} ms/step ({fe_ms:.3f} kernel by kernel), same demand: {same}",
                  file=sys.stderr, flush=True)
        except gr.RouterError as e:  # reported in the line; the default form's figures stand
            fused = {"form": form, "error": str(e)}
            print(f"# fused form failed: {e}", file=sys.stderr, flush=True)
        r.set_tuning(**base)
        r.reset_demand()
        r.route_schedule()
        r.sync()

    deg, P, F = footprint(X, Y, Z, pfx, pins)
    G = int(r.demand().sum())  # committed cells over the pass = total demand bumps
    n_b = len(ofst) - 1
    # algorithmic bytes per pass of each kernel, summed per net: costing reads
    # the pins (4 B each) and the distinct cost cells its pairs depend on
    # (4 B each) and writes a cost and a pattern per pair (6 B); the trees read
    # every pair cost, the pins, the edge patterns and slots, and stamp 4 B per
    # committed cell
    kbytes = {
        0: float((4 * deg + 4 * F + 6 * P).sum()),                        # pins + cost footprint + (cost, pattern) out
        1: float((4 * P + 4 * deg + 2 * (deg - 1) + 12 * deg + 4).sum() + 4 * G),  # costs, pins, patterns, slots, stamps
        2: None,
    }
    bytes_path = float((4 * deg + 4 * F + 10 * (deg - 1)).sum() + 8 * G)
    roof = None
    kstats = {}
    if not args.no_timing:
        # per-kernel durations: HIP events on the launch stream around every
        # launch, over the same number of passes (events cannot sit inside the
        # captured graph, so these passes launch kernel by kernel)
        r.enable_timing(True)
        t2 = time.perf_counter()
        for _ in range(args.steps):
            r.reset_demand()
            r.route_schedule()
        r.sync()
        t3 = time.perf_counter()
        n, ms = r.kernel_stats()
        r.enable_timing(False)
        # k_refresh also counts the reset's one full rebuild per step
        for k in range(3):
            kstats[KNAMES[k]] = {"launches": int(n[k]), "ms_per_step": float(ms[k]) / args.steps}
            if kbytes.get(k):
                kstats[KNAMES[k]]["GBps"] = round(kbytes[k] * args.steps / (ms[k] / 1e3) / 1e9, 2)
        kstats["instrumented_ms_per_step"] = (t3 - t2) * 1e3 / args.steps
        # dominant kernel: the largest share of the pass (the refresh count
        # includes the reset's full rebuild, one launch per step)
        nl = [int(n[0]), int(n[1]), int(n[2]) - args.steps]
        mk = [float(ms[0]), float(ms[1]), float(ms[2])]
        dom = int(np.argmax(mk))
        avg_s = mk[dom] / 1e3 / nl[dom]
        # algorithmic bytes of the whole path per batch (per routed net
        # 4n + 4F + 10(n-1) + 8G, summed over the pass, over the batches),
        # against the dominant kernel's average launch: one launch per batch
        ach = bytes_path / n_b / avg_s / 1e9
        roof = {"kernel": KNAMES[dom], "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None,
                "avg_launch_us": round(avg_s * 1e6, 2),
                "algorithmic_bytes_per_launch": round(bytes_path / n_b),
                "kernel_split": {KNAMES[k]: {"bytes_per_launch": round(kbytes[k] / n_b) if kbytes.get(k) else None,
                                             "avg_launch_us": round(mk[k] / nl[k] * 1e3, 2)} for k in range(3)}}
        # HBM bytes per launch measured by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
        # (calibrated on a copy kernel) for this workload; see profiles/
        prof = os.path.join(ROOT, "profiles", f"pmc_traffic_workload{args.workload}.json")
        if os.path.exists(prof):
            with open(prof) as f:
                roof["traffic"] = json.load(f)["per_launch"][KNAMES[dom]]["traffic_bytes"]
    # side figure: the same pass with minimum spanning trees (GR_TREE_TRUE_MST)
    # instead of the reference's root-pair trees, for route quality
    true_mst = mst_side(gr, r, cap, n_nets, args) if args.mst else None
    io = None
    if args.io:

        io = io_eval(gr, r, X, Y, Z, cap, pfx, pins, ofst, ids, cpu=not args.no_cpu)
        print(f"# io {io}", file=sys.stderr, flush=True)
    learned = None
    if args.learned:
        learned = learned_eval(gr, r, X, Y, Z, cap, W["nets"], args, n_nets, ms_step, (pfx, pins, ofst, ids))
    stress = None
    if args.stress:
        stress = stress_eval(gr, args, tuning)
        print(f"# stress {stress}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(X, Y, Z, cap, pfx, pins, ofst, ids, args.cpu_seconds)
    head = {}
    if learned and "k4" in learned:
        head = {"learned_nets_per_s": learned["k4"]["nets_per_s"], "learned_vs_plain": learned["k4"]["vs_plain"],
                "learned_route_cost_gap": learned["k4"]["route_cost_gap"],
                "learned_fallback_rate": learned["k4"]["fallback_rate"]}
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "nets/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"plain GPU path; {X}x{Y}x{Z} GCELLs, {W['nets']} nets requested "
                               f"(seed {args.seed}), {n_nets} routed after pin dedupe, "
                               f"{len(ofst) - 1} batches from 4 x-buckets",
                   "grid": [X, Y, Z], "nets": n_nets, "pairs": int(P.sum()), "batches": len(ofst) - 1,
                   "mode": "plain", "tuning": tuning},
        "roofline": roof,
        "cpu_baseline": cpu,
        **head,
        "kernels": kstats,
        "learned": learned,
        "stress": stress,
        "true_mst": true_mst,
        "callers": io,
        "per_kernel_launches": {"ms_per_step": round(eager_ms, 4), "nets_per_s": round(n_nets / (eager_ms / 1e3), 1),
                                "graph_speedup": round(eager_ms / ms_step, 3)},
        "fused_form": fused,
        "path_bytes_per_step": bytes_path,
        "path_GBps": round(bytes_path / (ms_step / 1e3) / 1e9, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
