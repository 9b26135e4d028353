"""Design files and route writer (host code of the library, no GPU).

The reader is checked against the restatement of READ_DATA's tokenizer and
net dialects (oracle/ispd_ref.py) on files with every number form the
tokenizer accepts, plus known answers for the fp32 digit accumulation; the
writer against the restatement of ROUTE_WRITER_UTIL on routes from the
oracle, in all four text forms, plus connectivity of every written net.
"""
import numpy as np
import pytest

from oracle import ispd_ref


def _same_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.int32), np.asarray(b, np.float32).view(np.int32))


def _num(rng, v):
    """A text form of v that the tokenizer reads: integer, decimal, exponent."""
    k = rng.integers(0, 5)
    if k == 0:
        return str(int(v))
    if k == 1:
        return f"{v:.1f}"
    if k == 2:
        return f"{v:.3f}"
    if k == 3:
        return f"{v * 10:.0f}e-1"
    return f"{v / 10:g}e+1" if v else "0"


def _write_cap(path, rng, X, Y, Z, new_fmt=False):
    vals = np.round(rng.uniform(0, 20, X * Y * Z), 2)
    vals[rng.random(vals.size) < 0.1] = 0
    lines = [f"{Z} {X} {Y}", "0.5 4", " ".join(_num(rng, v) for v in rng.uniform(0, 3, Z))]
    if new_fmt:
        lines.append("1 2.5 3e1 0.125")
    lines.append(" ".join(str(int(v)) for v in rng.integers(1000, 5000, X - 1)))
    lines.append(" ".join(str(int(v)) for v in rng.integers(1000, 5000, Y - 1)))
    k = 0
    for z in range(Z):
        lines.append(f"Metal{z + 1} {z & 1} {int(rng.integers(0, 9))}")
        for y in range(Y):
            row = []
            for x in range(X):
                row.append(_num(rng, vals[k]))
                k += 1
            lines.append(" ".join(row))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _write_net(path, rng, X, Y, Z, n, dialect):
    out = []
    for i in range(n):
        npins = int(rng.choice([1, 2, 3, 5, 9]))
        out.append(f"net_{i} x{i}\n(\n" if i % 3 == 0 else f"n{i}\n(\n")
        for q in range(npins):
            z, x, y = int(rng.integers(0, min(Z, 3))), int(rng.integers(0, X)), int(rng.integers(0, Y))
            if q and rng.random() < 0.2:  # duplicate pin
                z, x, y = last
            last = (z, x, y)
            if dialect == "new":
                out.append(f"[p{q}, c{i}, ({z}, {x}, {y})]\n")
            elif dialect == "first":
                extra = ", ".join(f"({z}, {min(x + t, X - 1)}, {y})" for t in range(1, int(rng.integers(1, 4))))
                out.append(f"[({z}, {x}, {y})" + (", " + extra if extra else "") + "]\n")
            else:
                out.append(f"[({z}, {x}, {y})]\n")
        out.append(")\n")
    with open(path, "w") as f:
        f.write("".join(out))


def test_tokenizer_known_answers(gr, tmp_path):
    """fp32 digit accumulation: integers and one-digit fractions exact where
    fp32 holds them; longer fractions within 2 ulp of the decimal value."""
    X, Y, Z = 4, 3, 2
    text = ["0.1", "2.5", "-1.5", "1e3", "25e-1", "7", "0", "3.7", "0.001", "123.456", "1e+2", "-0.25"] * 2
    lines = [f"{Z} {X} {Y}", "1 1", "0 0", "1 1 1", "1 1", "M1 0 0", " ".join(text[:12]), "M2 1 0", " ".join(text[12:])]
    p = tmp_path / "k.cap"
    p.write_text("\n".join(lines) + "\n")
    (tmp_path / "k.net").write_text("a\n(\n[(0, 0, 0)]\n[(1, 1, 1)]\n)\n")
    d = gr.Design.read(str(p), str(tmp_path / "k.net"))
    cap = d.capacity()
    got = [cap[ispd_ref.enc(X, Y, 0, k % X, k // X)] for k in range(12)]
    exact = {0: np.float32(0.1), 1: 2.5, 2: -1.5, 3: 1000.0, 4: 2.5, 5: 7.0, 6: 0.0, 10: 100.0, 11: -0.25}
    for k, v in exact.items():
        assert got[k] == np.float32(v), (text[k], got[k])
    for k in (7, 8, 9):
        want = np.float32(float(text[k]))
        assert abs(int(got[k].view(np.int32)) - int(want.view(np.int32))) <= 2, (text[k], got[k])
    ref = ispd_ref.read_cap(p.read_bytes())
    assert _same_f32(cap, ref["cap"])


@pytest.mark.parametrize("dialect,flags,new_fmt", [("single", 0, False), ("first", 1, False), ("new", 2, True)])
def test_reader_matches_restatement(gr, tmp_path, dialect, flags, new_fmt):
    rng = np.random.default_rng(7)
    X, Y, Z = 23, 17, 5
    _write_cap(tmp_path / "d.cap", rng, X, Y, Z, new_fmt)
    _write_net(tmp_path / "d.net", rng, X, Y, Z, 60, dialect)
    for threads in (1, 3, 0):
        d = gr.Design.read(str(tmp_path / "d.cap"), str(tmp_path / "d.net"), flags, threads)
        ref = ispd_ref.read_cap((tmp_path / "d.cap").read_bytes(), new_fmt)
        assert (d.X, d.Y, d.Z) == (X, Y, Z)
        assert _same_f32(d.capacity(), ref["cap"])
        lay = d.layers()
        assert lay["name"] == ref["layer_name"] and np.array_equal(lay["dir"], ref["layer_dir"])
        assert np.array_equal(lay["min_len"], ref["layer_min_len"]) and _same_f32(lay["unit_short_cost"], ref["short_cost"])
        h, v = d.gcell_lengths()
        assert np.array_equal(h, ref["hlen"]) and np.array_equal(v, ref["vlen"])
        assert (d.unit_length_cost, d.unit_via_cost) == (0.5, 4.0)
        routed, single = ispd_ref.read_net((tmp_path / "d.net").read_bytes(), X, Y, Z, dialect)
        pfx, pins = d.nets()
        assert d.n_nets == len(routed) and d.n_single == len(single)
        assert np.array_equal(pfx, np.cumsum([0] + [len(c) for _, c in routed]))
        assert np.array_equal(pins, np.concatenate([c for _, c in routed]))
        assert np.array_equal(d.single_nets(), [c for _, c in single])
        for i, (nm, _) in enumerate(routed):
            assert d.net_name(i) == nm.replace("\n", "").replace(" ", "")
        for j, (nm, _) in enumerate(single):
            assert d.net_name(d.n_nets + j) == nm.replace("\n", "").replace(" ", "")


def test_reader_errors(gr, tmp_path):
    (tmp_path / "a.cap").write_text("2 3 3\n1 1\n1 1\n1 1\n1 1\nM1 0 0\n1 2 3\n")  # too few capacities
    (tmp_path / "a.net").write_text("n\n(\n[(0, 0, 0)]\n)\n")
    with pytest.raises(gr.RouterError):
        gr.Design.read(str(tmp_path / "a.cap"), str(tmp_path / "a.net"))
    rng = np.random.default_rng(1)
    _write_cap(tmp_path / "b.cap", rng, 5, 4, 3)
    for bad in ("n\n(\n[(0, 9, 0)]\n)\n",          # outside the grid
                "n\n(\n[(0, 1, 1), (0, 2, 1)]\n)\n",  # two access points, single dialect
                "n\n(\n[(0, 1, 1)]\n",                # unbalanced
                "n\n(\n)\n"):                         # no pins
        (tmp_path / "b.net").write_text(bad)
        with pytest.raises(gr.RouterError):
            gr.Design.read(str(tmp_path / "b.cap"), str(tmp_path / "b.net"))
    with pytest.raises(gr.RouterError):
        gr.Design.read(str(tmp_path / "missing.cap"), str(tmp_path / "b.net"))


def test_reader_survives_mangled_files(gr, tmp_path):
    """Truncated, byte-flipped, token-mangled and span-cut copies of a saved
    design either read (possibly as a different design) or fail with an error
    code; none may crash the process.  The same mangles ran without reports
    under AddressSanitizer and UBSan on a host-only build."""
    import random
    X, Y, Z = 12, 9, 4
    cap = gr.gen_grid(1, 2, X, Y, Z).astype(np.float32)
    pfx, pins = gr.gen_nets(1, 2, X, Y, Z, 40)
    gr.Design(X, Y, Z, cap, pfx, pins).save(str(tmp_path / "d.cap"), str(tmp_path / "d.net"))
    texts = [(tmp_path / "d.cap").read_bytes(), (tmp_path / "d.net").read_bytes()]
    rng = random.Random(11)
    outcomes = {"read": 0, "error": 0}
    for trial in range(150):
        which = trial % 2
        t = bytearray(texts[which])
        mode = rng.randrange(4)
        if mode == 0:
            t = t[:rng.randrange(len(t))]
        elif mode == 1:
            for _ in range(rng.randrange(1, 8)):
                t[rng.randrange(len(t))] = rng.randrange(256)
        elif mode == 2:
            toks = t.split(b" ")
            toks[rng.randrange(len(toks))] = rng.choice([b"-1", b"99999999999", b"1e40", b"nan", b"", b"(", b")", b"x"])
            t = bytearray(b" ".join(toks))
        else:
            i = rng.randrange(len(t))
            t = t[:i] + t[i + rng.randrange(1, 40):]
        files = [bytes(t) if k == which else texts[k] for k in range(2)]
        (tmp_path / "f.cap").write_bytes(files[0])
        (tmp_path / "f.net").write_bytes(files[1])
        try:
            d = gr.Design.read(str(tmp_path / "f.cap"), str(tmp_path / "f.net"), threads=rng.choice([0, 1, 3]))
            assert d.X > 0 and d.Y > 0 and d.Z > 0
            outcomes["read"] += 1
        except gr.RouterError:
            outcomes["error"] += 1
    assert outcomes["error"] > 0


def test_save_round_trip(gr, tmp_path):
    X, Y, Z = 40, 30, 6
    cap = gr.gen_grid(1, 3, X, Y, Z).astype(np.float32)
    cap[::7] += np.float32(0.5)
    cap[::11] = np.float32(0.25)
    pfx, pins = gr.gen_nets(1, 3, X, Y, Z, 400)
    d = gr.Design(X, Y, Z, cap, pfx, pins)
    d.save(str(tmp_path / "s.cap"), str(tmp_path / "s.net"))
    e = gr.Design.read(str(tmp_path / "s.cap"), str(tmp_path / "s.net"))
    assert _same_f32(e.capacity(), cap)
    for a, b in zip(e.nets(), (pfx, pins)):
        assert np.array_equal(a, b)
    assert [e.net_name(i) for i in range(5)] == [f"net{i}" for i in range(5)]


def _oracle_routes(gr, orc, X, Y, Z, kind, seed, nets, singles=()):
    cap = gr.gen_grid(kind, seed, X, Y, Z)
    pfx, pins = gr.gen_nets(kind, seed, X, Y, Z, nets)
    # interleave one-pin nets into the design's net list
    alln = []
    for i in range(len(pfx) - 1):
        alln.append(list(pins[pfx[i]:pfx[i + 1]]))
        if i in singles:
            alln.append([int(pins[pfx[i]])])
    apfx = np.cumsum([0] + [len(n) for n in alln]).astype(np.int64)
    d = gr.Design(X, Y, Z, cap.astype(np.float32), apfx, np.concatenate([np.array(n, np.int32) for n in alln]))
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    o = orc.Oracle(X, Y, Z, cap)
    o.set_nets(pfx, pins)
    E, L = [], []
    for b in range(len(ofst) - 1):
        e, l, _ = o.route_batch(ids[ofst[b]:ofst[b + 1]])
        E.append(e)
        L.append(l)
    return d, alln, pfx, pins, ofst, ids, np.concatenate(E), np.concatenate(L), o


@pytest.mark.parametrize("flags", [0, 1, 2, 3, 6, 7])
def test_writer_matches_restatement(gr, orc, tmp_path, flags):
    X, Y, Z = 60, 50, 6
    d, alln, pfx, pins, ofst, ids, E, L, _ = _oracle_routes(gr, orc, X, Y, Z, 1, 2, 700, singles={3, 50, 400})
    assert d.n_single == 3
    path = tmp_path / "r.txt"
    n = d.write_routes(str(path), ofst, ids, E, L, flags=flags, threads=3)
    text = path.read_text()
    assert n == len(text.encode())
    routed = [(f"net{k}", c) for k, c in enumerate(alln) if len(c) > 1]
    single = [(f"net{k}", c[0]) for k, c in enumerate(alln) if len(c) == 1]
    want = ispd_ref.routes_text(X, Y, Z, routed, single, ofst, ids, E, L,
                                merge=bool(flags & 2), patch=bool(flags & 4), dbu=bool(flags & 1))
    assert text == want


def _cells_of_route_text(X, Y, text):
    """Per net name: set of grid cells and union-find connectivity of the grid-form lines."""
    nets = {}
    for block in text.split(")\n"):
        if not block.strip():
            continue
        name, body = block.split("\n(\n")
        segs = [tuple(map(int, ln.split())) for ln in body.strip("\n").split("\n") if ln]
        nets[name] = segs
    return nets


def test_written_routes_connect_all_pins(gr, orc, tmp_path):
    X, Y, Z = 60, 50, 6
    d, alln, pfx, pins, ofst, ids, E, L, _ = _oracle_routes(gr, orc, X, Y, Z, 1, 5, 500)
    for flags in (0, 2, 6):
        path = tmp_path / f"r{flags}.txt"
        d.write_routes(str(path), ofst, ids, E, L, flags=flags)
        nets = _cells_of_route_text(X, Y, path.read_text())
        for k in range(len(pfx) - 1):
            par = {}

            def find(a):
                while par.setdefault(a, a) != a:
                    par[a] = par[par[a]]
                    a = par[a]
                return a

            def union(a, b):
                par[find(a)] = find(b)

            for z1, x1, y1, z2, x2, y2 in nets[f"net{k}"]:
                if (x1, y1) == (x2, y2) and z1 == z2:  # via line: layer z1 to z1+1
                    union((z1, x1, y1), (z1 + 1, x1, y1))
                    continue
                assert z1 == z2 and (x1 == x2 or y1 == y2)
                if x1 == x2:
                    for t in range(min(y1, y2), max(y1, y2)):
                        union((z1, x1, t), (z1, x1, t + 1))
                else:
                    for t in range(min(x1, x2), max(x1, x2)):
                        union((z1, t, y1), (z1, t + 1, y1))
            roots = {find(ispd_ref.dec(X, Y, int(c))) for c in pins[pfx[k]:pfx[k + 1]]}
            assert len(roots) == 1, (flags, k)


def test_commit_equals_written_route(gr, orc, tmp_path):
    """Route geometry written by the writer against the commit (fixed mode,
    oracle): per batch, the cells whose demand rose, plus the batch's pins,
    equal the cells the written segments cover, plus the batch's pins."""
    X, Y, Z = 60, 50, 6
    cap = gr.gen_grid(1, 8, X, Y, Z)
    pfx, pins = gr.gen_nets(1, 8, X, Y, Z, 600)
    d = gr.Design(X, Y, Z, cap.astype(np.float32), pfx, pins)
    ofst, ids = gr.partition_nets(X, Y, Z, pfx, pins, 4)
    o = orc.Oracle(X, Y, Z, cap, 0, 1)
    o.set_nets(pfx, pins)
    before = o.demand()
    for b in range(len(ofst) - 1):
        batch = ids[ofst[b]:ofst[b + 1]]
        e, l, _ = o.route_batch(batch)
        after = o.demand()
        marked = set(np.nonzero(after - before)[0].tolist())
        before = after
        path = tmp_path / "b.txt"
        d.write_routes(str(path), np.array([0, len(batch)]), batch, e, l)
        covered = set()
        for segs in _cells_of_route_text(X, Y, path.read_text()).values():
            for z1, x1, y1, z2, x2, y2 in segs:
                if (x1, y1, z1) == (x2, y2, z2):
                    covered |= {ispd_ref.enc(X, Y, z1, x1, y1), ispd_ref.enc(X, Y, z1 + 1, x1, y1)}
                elif x1 == x2:
                    covered |= {ispd_ref.enc(X, Y, z1, x1, t) for t in range(min(y1, y2), max(y1, y2) + 1)}
                else:
                    covered |= {ispd_ref.enc(X, Y, z1, t, y1) for t in range(min(x1, x2), max(x1, x2) + 1)}
        pin_cells = set(np.concatenate([pins[pfx[n]:pfx[n + 1]] for n in batch]).tolist())
        assert marked | pin_cells == covered | pin_cells, b


def test_writer_rejects_bad_edges(gr, orc, tmp_path):
    X, Y, Z = 60, 50, 6
    d, alln, pfx, pins, ofst, ids, E, L, _ = _oracle_routes(gr, orc, X, Y, Z, 1, 2, 200)
    E2 = E.copy()
    E2[0] = (5 << 30) | (E2[0] & ((1 << 30) - 1))  # wrong net-in-batch field
    with pytest.raises(gr.RouterError):
        d.write_routes(str(tmp_path / "x.txt"), ofst, ids, E2, L)
    L2 = L.copy()
    L2[0] = 0  # a = 0 never reaches the other pin for a pair in different columns/rows
    first = ids[0]
    p = pins[pfx[first]:pfx[first + 1]]
    n1, n2 = int(E[0]) & 0x7FFF, (int(E[0]) >> 15) & 0x7FFF
    (za, xa, ya), (zb, xb, yb) = ispd_ref.dec(X, Y, int(p[max(n1, n2)])), ispd_ref.dec(X, Y, int(p[min(n1, n2)]))
    if (xa, ya) != (xb, yb):
        with pytest.raises(gr.RouterError):
            d.write_routes(str(tmp_path / "x.txt"), ofst, ids, E, L2)
