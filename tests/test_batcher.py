"""Batch generation (part_net/net_part.hpp:167-339): the product's
gr_partition_nets against the oracle's restatement, and the restatement
pinned by a hand-worked example, brute force and the invariants of its steps.

Batch membership is part of the routing result (it fixes which demand each
net sees), so the product must produce exactly the oracle's batches.
"""
import numpy as np
import pytest


def _boxes(X, Y, pfx, pins):
    n = len(pfx) - 1
    p = pins.astype(np.int64)
    z = p // (X * Y)
    r = p - z * X * Y
    odd = (z & 1) == 1
    x = np.where(odd, r // Y, r % X)
    y = np.where(odd, r % Y, r // X)
    net = np.repeat(np.arange(n), np.diff(pfx))
    b = np.empty((n, 4), np.int64)
    b[:, 0] = 1 << 40; b[:, 1] = -1; b[:, 2] = 1 << 40; b[:, 3] = -1
    np.minimum.at(b[:, 0], net, x); np.maximum.at(b[:, 1], net, x)
    np.minimum.at(b[:, 2], net, y); np.maximum.at(b[:, 3], net, y)
    return b


def _meet(a, b):
    return not (a[1] < b[0] or b[1] < a[0] or a[3] < b[2] or b[3] < a[2])


def test_hand_worked_example(orc, gr):
    """Six 2-pin nets, two x-buckets, worked by hand from net_part.hpp:167-235.
    Boxes (x range, y range): 0 [0,3]x[0,0], 1 [1,1]x[0,3], 2 [2,2]x[1,1],
    3 [5,6]x[0,0], 4 [4,7]x[2,2], 5 [3,5]x[1,1].
    Batch 1 (targ 3) walks nets by maxx: 1, 2, 0 fill bucket 0 and set the cut
    to 3; net 5 (minx 3) is left of the cut; 3 and 4 go to bucket 1.  Bucket 0
    is listed in reverse walking order, 0, 2, 1, and sorted by half perimeter
    (2: 2, 0: 5, 1: 5, a tie kept in that order), so 2 and 0 are taken and 1,
    which meets 0, is not (in walking order 1 would win the tie).  Bucket 1
    takes 3 and 4.  Batch 2 (targ 1): 1 fills bucket 0 (cut 1), 5 bucket 1."""
    X, Y, Z = 8, 4, 3

    def e(z, x, y):
        return z * X * Y + (x * Y + y if z & 1 else y * X + x)

    nets = [[e(0, 0, 0), e(0, 3, 0)], [e(0, 1, 0), e(0, 1, 3)], [e(0, 2, 1), e(1, 2, 1)],
            [e(0, 5, 0), e(0, 6, 0)], [e(0, 4, 2), e(0, 7, 2)], [e(0, 3, 1), e(0, 5, 1)]]
    pfx = np.array([0, 2, 4, 6, 8, 10, 12], np.int64)
    pins = np.array([v for n in nets for v in sorted(n)], np.int32)
    want_ofst, want_ids = [0, 4, 6], [0, 2, 3, 4, 1, 5]
    o, i = orc.partition(X, Y, pfx, pins, 2)
    assert o.tolist() == want_ofst and i.tolist() == want_ids
    o2, i2 = gr.partition_nets(X, Y, Z, pfx, pins, 2)
    assert o2.tolist() == want_ofst and i2.tolist() == want_ids


@pytest.mark.parametrize("seed,n,side", [(1, 400, 6), (2, 900, 20), (3, 2000, 3)])
def test_tiles_equal_brute_force(orc, seed, n, side):
    """The oracle's tiled box test against every-box brute force, on dense
    random boxes (many touching and nested ones)."""
    X, Y = 97, 61
    rng = np.random.default_rng(seed)
    pfx, pins = [0], []
    for _ in range(n):
        x0, y0 = int(rng.integers(0, X - side)), int(rng.integers(0, Y - side))
        cells = {(int(x0 + rng.integers(0, side + 1)), int(y0 + rng.integers(0, side + 1))) for _ in range(3)}
        enc = sorted({y * X + x for x, y in cells})
        if len(enc) < 2:
            continue
        pins += enc
        pfx.append(len(pins))
    pfx, pins = np.array(pfx, np.int64), np.array(pins, np.int32)
    ref = orc.partition(X, Y, pfx, pins, 4, tile=0)
    for tile in (1, 7, 32):
        got = orc.partition(X, Y, pfx, pins, 4, tile=tile)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), tile


def test_oracle_batch_invariants(orc, gr):
    """Properties of every step, on the 128x128 workload's first batches:
    bucket sizes (targ each but the last), x-disjoint buckets, nets left of a
    cut in no bucket, taken boxes pairwise disjoint, and the greedy is
    maximal: every bucket net not taken meets a box taken in its bucket."""
    X, Y, Z = 128, 128, 6
    pfx, pins = gr.gen_nets(1, 0, X, Y, Z, 5000)
    bx = _boxes(X, Y, pfx, pins)
    ofst, ids = orc.partition(X, Y, pfx, pins, 4)
    for b in range(6):
        _, _, bucket = orc.partition(X, Y, pfx, pins, 4, trace_batch=b)
        taken = set(ids[ofst[b]:ofst[b + 1]].tolist())
        remaining = int((bucket != -2).sum())
        assert remaining == len(pfx) - 1 - ofst[b]
        targ = -(-remaining // 4)
        nk = bucket.max() + 1
        sizes = [int((bucket == k).sum()) for k in range(nk)]
        assert all(s == targ for s in sizes[:-1]) and 0 < sizes[-1] <= targ, sizes
        cuts = [int(bx[bucket == k, 1].max()) for k in range(nk)]
        for k in range(1, nk):
            assert bx[bucket == k, 0].min() > cuts[k - 1]
        for u in np.nonzero(bucket == -1)[0]:
            assert any(bx[u, 0] <= c <= bx[u, 1] for c in cuts[:-1]), u
        assert all(bucket[t] >= 0 for t in taken)
        for k in range(nk):
            members = np.nonzero(bucket == k)[0]
            tk = [m for m in members if m in taken]
            for a in range(len(tk)):
                for c in range(a + 1, len(tk)):
                    assert not _meet(bx[tk[a]], bx[tk[c]])
            for m in members:
                if m not in taken:
                    assert any(_meet(bx[m], bx[t]) for t in tk), (b, k, m)


@pytest.mark.parametrize("kind,X,Y,Z,nets", [(1, 128, 128, 6, 5000), (2, 800, 800, 10, 150000),
                                             (3, 2000, 2000, 10, 1500000)])
def test_product_batches_equal_oracle(gr, orc, kind, X, Y, Z, nets):
    """gr_partition_nets gives exactly the oracle's batches (membership and
    order) on the three workloads' nets."""
    pfx, pins = gr.gen_nets(kind, 0, X, Y, Z, nets)
    o1, i1 = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    o2, i2 = orc.partition(X, Y, pfx, pins, 4)
    assert len(o1) == len(o2)
    assert np.array_equal(o1, o2)
    assert np.array_equal(i1, i2)
