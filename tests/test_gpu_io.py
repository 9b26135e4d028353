"""End to end on the GPU: ISPD files -> reader -> HIP router on the fp32
capacities -> route writer, against the oracle (fp32 grid) routing the same
design and the restated writer; and the GPU's commit against the written
route geometry, batch by batch."""
import numpy as np
import pytest

from oracle import ispd_ref

pytestmark = pytest.mark.gpu


def _design_files(tmp_path, rng, X, Y, Z, n):
    cap = np.round(rng.uniform(0, 14, X * Y * Z), 1).astype(np.float32)
    lines = [f"{Z} {X} {Y}", "0.5 4", " ".join(["1"] * Z), " ".join(["2000"] * (X - 1)), " ".join(["2000"] * (Y - 1))]
    for z in range(Z):
        lines.append(f"Metal{z + 1} {z & 1} 1")
        for y in range(Y):
            lines.append(" ".join(f"{cap[ispd_ref.enc(X, Y, z, x, y)]:.1f}" for x in range(X)))
    (tmp_path / "g.cap").write_text("\n".join(lines) + "\n")
    out = []
    for i in range(n):
        k = int(rng.choice([1, 2, 2, 3, 4, 6, 12]))
        cx, cy = int(rng.integers(0, X)), int(rng.integers(0, Y))
        out.append(f"n{i}\n(\n")
        for _ in range(k):
            x = int(np.clip(cx + rng.integers(-6, 7), 0, X - 1))
            y = int(np.clip(cy + rng.integers(-6, 7), 0, Y - 1))
            z = int(rng.integers(0, 2))
            out.append(f"[({z}, {x}, {y}), ({z}, {min(x + 1, X - 1)}, {y})]\n")
        out.append(")\n")
    (tmp_path / "g.net").write_text("".join(out))


def test_files_route_write_against_oracle(gr, orc, gpu, tmp_path):
    X, Y, Z = 72, 64, 8
    _design_files(tmp_path, np.random.default_rng(3), X, Y, Z, 900)
    d = gr.Design.read(str(tmp_path / "g.cap"), str(tmp_path / "g.net"), gr.FMT_FIRST_ACCESS_POINT)
    cap = d.capacity()
    pfx, pins = d.nets()
    assert d.n_single > 0 and cap.dtype == np.float32
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap)
    r.set_nets(pfx, pins)
    r.set_schedule(ofst, ids)
    r.route_schedule()
    E, L, C = r.schedule_outputs()
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    oe, ol = [], []
    for b in range(len(ofst) - 1):
        e, l, _ = o.route_batch(ids[ofst[b]:ofst[b + 1]])
        oe.append(e)
        ol.append(l)
    assert np.array_equal(E, np.concatenate(oe)) and np.array_equal(L, np.concatenate(ol))
    assert np.array_equal(r.demand(), o.demand())
    routed, single = ispd_ref.read_net((tmp_path / "g.net").read_bytes(), X, Y, Z, "first")
    for flags in (0, 1, 2, 7):
        path = tmp_path / f"out{flags}.txt"
        d.write_routes(str(path), ofst, ids, E, L, flags=flags)
        want = ispd_ref.routes_text(X, Y, Z, routed, single, ofst, ids, E, L,
                                    merge=bool(flags & 2), patch=bool(flags & 4), dbu=bool(flags & 1))
        assert path.read_text() == want, flags


def test_gpu_commit_equals_written_route(gr, gpu, tmp_path):
    X, Y, Z = 96, 80, 8
    cap = gr.gen_grid(2, 4, X, Y, Z)
    pfx, pins = gr.gen_nets(2, 4, X, Y, Z, 2500)
    d = gr.Design(X, Y, Z, cap.astype(np.float32), pfx, pins)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    r = gr.Router(X, Y, Z, cap, 0, gr.COMMIT_FIXED)
    r.set_nets(pfx, pins)
    before = r.demand()
    for b in range(len(ofst) - 1):
        batch = ids[ofst[b]:ofst[b + 1]]
        print(f"commit vs route batch {b}", flush=True)
        e, l, _ = r.route_batch(batch)
        after = r.demand()
        delta = after - before
        assert delta.min() >= 0 and delta.max() <= 1
        marked = set(np.nonzero(delta)[0].tolist())
        before = after
        path = tmp_path / "b.txt"
        d.write_routes(str(path), np.array([0, len(batch)]), batch, e, l)
        covered = set()
        for line in path.read_text().split("\n"):
            f = line.split()
            if len(f) != 6:
                continue
            z1, x1, y1, z2, x2, y2 = map(int, f)
            if (x1, y1, z1) == (x2, y2, z2):
                covered |= {ispd_ref.enc(X, Y, z1, x1, y1), ispd_ref.enc(X, Y, z1 + 1, x1, y1)}
            elif x1 == x2:
                covered |= {ispd_ref.enc(X, Y, z1, x1, t) for t in range(min(y1, y2), max(y1, y2) + 1)}
            else:
                covered |= {ispd_ref.enc(X, Y, z1, t, y1) for t in range(min(x1, x2), max(x1, x2) + 1)}
        pin_cells = set(np.concatenate([pins[pfx[n]:pfx[n + 1]] for n in batch]).tolist())
        assert marked | pin_cells == covered | pin_cells, b
