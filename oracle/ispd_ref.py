"""CPU restatement of the router's file formats (TEST INFRASTRUCTURE).

Pure-Python restatement, for small files, of
  * the ISPD .cap / .net reader READ_DATA (read_data.hpp): tokenizer
    get_long / get_string / get_float (:15-93, fp32 arithmetic emulated with
    numpy float32), read_net_metadata / read_hlen_data / read_vlen_data /
    read_cube_data (:94-183), get_num_nets_temp (:215-236), the three net
    dialects read_nt / read_ispd24_net / read_ispd24_net_single (:262-333),
    remove_redundant_pins (:413-432) and remove_single_deg_nets (:366-412);
  * the route writer ROUTE_WRITER_UTIL (writer_files/writer_util.hpp):
    write_route (:115-165), mrg_route_write (:290-364) with
    merge_st_edges_nnew (:166-205) and oprd_ppatch (:234-289),
    write_odeg_route (:101-114), line formats (:17-53, :206-220), and the
    batch order of ROUTE_WRITER::write_route (writer.hpp:12-51).
Used by tests/ to check the library's reader and writer.
"""
import numpy as np

F32 = np.float32


def enc(X, Y, z, x, y):
    return z * X * Y + (x * Y + y if z & 1 else y * X + x)


def dec(X, Y, v):
    z = v // (X * Y)
    r = v - z * X * Y
    return (z, r // Y, r % Y) if z & 1 else (z, r % X, r // X)


class _Tok:
    def __init__(self, data):
        self.b = data + b"\0"
        self.i = 0
        self.n = len(data)

    def _dig(self, k):
        c = self.b[k]
        return 48 <= c <= 57

    def get_long(self):  # read_data.hpp:15-22
        while self.i < self.n and not self._dig(self.i):
            self.i += 1
        v = 0
        while self._dig(self.i):
            v = v * 10 + self.b[self.i] - 48
            self.i += 1
        return v

    def get_string(self):  # read_data.hpp:31-40
        sep = (32, 10, 44)
        while self.i < self.n and self.b[self.i] in sep:
            self.i += 1
        s = self.i
        while self.i < self.n and self.b[self.i] not in sep:
            self.i += 1
        return self.b[s:self.i].decode()

    def get_float(self):  # read_data.hpp:41-66 (get_double :67-93 is the same for gcb_dt = float)
        val = F32(0.0)
        while self.b[self.i] != 45 and not self._dig(self.i):
            if self.i >= self.n:
                raise ValueError("file ends early")
            self.i += 1
        neg = False
        if self.b[self.i] == 45:
            neg = True
            self.i += 1
        while self._dig(self.i):
            ch = self.b[self.i]
            self.i += 1
            val = F32(float(val) * 10.0 + float(ch) - 48.0)  # double arithmetic, stored to float
        if self.b[self.i] == 46:
            mul = F32(0.1)
            self.i += 1
            while self._dig(self.i):
                d = F32(self.b[self.i] - 48)
                self.i += 1
                val = F32(val + F32(d * mul))
                mul = F32(float(mul) * 0.1)
        if self.b[self.i] == 101:
            self.i += 1
            mneg = False
            pw = 0
            if self.b[self.i] == 45:
                mneg = True
                self.i += 1
            if self.b[self.i] == 43:
                self.i += 1
            while self._dig(self.i):
                pw = pw * 10 + self.b[self.i] - 48
                self.i += 1
            for _ in range(pw):
                val = F32(float(val) * 1.0 / 10.0) if mneg else F32(float(val) * 10.0)
        return F32(-val) if neg else val


def read_cap(data, new_fmt=False):
    t = _Tok(data)
    Z, X, Y = t.get_long(), t.get_long(), t.get_long()
    out = dict(X=X, Y=Y, Z=Z)
    out["unit_length_cost"] = t.get_float()
    out["unit_via_cost"] = t.get_float()
    out["short_cost"] = np.array([t.get_float() for _ in range(Z)], F32)
    if new_fmt:
        for _ in range(4):
            t.get_float()
    out["hlen"] = np.array([t.get_long() for _ in range(X - 1)], np.int64)
    out["vlen"] = np.array([t.get_long() for _ in range(Y - 1)], np.int64)
    cap = np.zeros(X * Y * Z, F32)
    names, dirs, mins = [], [], []
    for z in range(Z):
        names.append(t.get_string())
        dirs.append(t.get_long())
        mins.append(t.get_long())
        for y in range(Y):
            for x in range(X):
                cap[enc(X, Y, z, x, y)] = t.get_float()
    out.update(cap=cap, layer_name=names, layer_dir=np.array(dirs), layer_min_len=np.array(mins))
    return out


def _cd(b, i, stop):  # read_cd_till_comma / read_cd_till_cbrck (:250-261)
    s = ""
    while i < len(b) and b[i] != stop:
        if 48 <= b[i] <= 57:
            s += chr(b[i])
        i += 1
    return int(s), i + 1


def read_net(data, X, Y, Z, dialect="single"):
    """Returns (routed nets [(raw name, [cells])], one-pin nets [(raw name, cell)])."""
    b = data
    spans, depth = [], 0
    for i, c in enumerate(b):  # get_num_nets_temp
        if c == 40:
            if depth == 0:
                spans.append([i, -1])
            depth += 1
        elif c == 41:
            depth -= 1
            if depth == 0:
                spans[-1][1] = i
    nets = []
    for k, (op, cl) in enumerate(spans):
        st = 0 if k == 0 else spans[k - 1][1] + 1
        name = b[st:op].decode()
        ed = cl
        pins = []
        while True:
            if dialect == "new":  # read_nt
                for _ in range(2):
                    while st < ed and b[st] != 44:
                        st += 1
                    st += 1
                if st >= ed:
                    break
                z, st = _cd(b, st, 44)
                x, st = _cd(b, st, 44)
                y, st = _cd(b, st, 41)
                pins.append((z, x, y))
                continue
            while st < ed and b[st] != 40:  # read_obrckt
                st += 1
            st += 1
            if st + 5 >= ed:
                break
            aps = []
            while True:
                z, st = _cd(b, st, 44)
                x, st = _cd(b, st, 44)
                y, st = _cd(b, st, 41)
                aps.append((z, x, y))
                if b[st] == 93:
                    break
                if dialect == "single":
                    raise ValueError("second access point")
                st += 1
            pins.append(aps[0])
        cells = sorted({enc(X, Y, *p) for p in pins})  # remove_redundant_pins
        nets.append((name, cells))
    routed = [(n, c) for n, c in nets if len(c) != 1]
    single = [(n, c[0]) for n, c in nets if len(c) == 1]
    return routed, single


def _name(raw):
    return raw.replace("\n", "").replace(" ", "")


def _gran(c):
    return 4200 * c + 2100


def _via_lines(x, y, z1, z2, dbu):  # add_vias_cpu
    z1, z2 = min(z1, z2), max(z1, z2)
    out = []
    for i in range(z1, z2):
        if dbu:
            out.append(f"{_gran(x)} {_gran(y)} metal{i + 1} {_gran(x)} {_gran(y)} metal{i + 2}")
        else:
            out.append(f"{i} {x} {y} {i} {x} {y}")
    return out


def _edge_path(X, Y, Z, src, dst, p):
    """The segments of write_route's walk: ('v', x, y, za, zb) and ('w', z, x1, y1, x2, y2)."""
    Z2 = (Z + 1) // 2
    z1, x1, y1 = dec(X, Y, src)
    z2, x2, y2 = dec(X, Y, dst)
    via1, via2 = p // Z2, p % Z2
    bth = x1 != x2 and y1 != y2
    segs = []
    if z1 != via1:
        segs.append(("v", x1, y1, z1, via1))
        z1 = via1
    if z1 > 0:
        if z1 & 1:
            segs.append(("w", z1, x1, y1, x1, y2))
            y1 = y2
        else:
            segs.append(("w", z1, x1, y1, x2, y1))
            x1 = x2
    via2 = 2 * via2 if z1 & 1 else 2 * via2 + 1
    if bth and via2 != 0 and z1 != via2:
        segs.append(("v", x1, y1, z1, via2))
        z1 = via2
        if z1 & 1:
            segs.append(("w", z1, x1, y1, x1, y2))
            y1 = y2
        else:
            segs.append(("w", z1, x1, y1, x2, y1))
            x1 = x2
    if z1 != z2:
        segs.append(("v", x1, y1, z1, z2))
    return segs


def net_text(X, Y, Z, raw_name, pins, edges, lpat, merge=False, patch=False, dbu=False):
    """Text of one routed net (write_route or mrg_route_write)."""
    lines = []
    ivs = []
    for e, p in zip(edges, lpat):
        if e == -1:
            continue
        n1, n2 = int(e) & 0x7FFF, (int(e) >> 15) & 0x7FFF
        if n1 < n2:
            n1, n2 = n2, n1
        for s in _edge_path(X, Y, Z, pins[n1], pins[n2], int(p)):
            if s[0] == "v":
                _, x, y, za, zb = s
                if merge:
                    for i in range(min(za, zb), max(za, zb)):
                        c = enc(X, Y, i, x, y)
                        ivs.append((c, c))
                else:
                    lines += _via_lines(x, y, za, zb, dbu)
            else:
                _, z, xa, ya, xb, yb = s
                if (xa, ya) == (xb, yb):
                    continue
                if merge:
                    c1, c2 = enc(X, Y, z, xa, ya), enc(X, Y, z, xb, yb)
                    ivs.append((min(c1, c2), max(c1, c2)))
                else:
                    xa, xb = min(xa, xb), max(xa, xb)
                    ya, yb = min(ya, yb), max(ya, yb)
                    if dbu:
                        lines.append(f"{_gran(xa)} {_gran(ya)} metal{z + 1} {_gran(xb)} {_gran(yb)} metal{z + 2}")
                    else:
                        lines.append(f"{z} {xa} {ya} {z} {xb} {yb}")
    if merge:
        pts = sorted({a for a, b in ivs if a == b})
        wires = sorted((a, b) for a, b in ivs if a != b)
        merged = []
        for a, b in wires:
            if merged and merged[-1][1] >= a:
                merged[-1][1] = max(merged[-1][1], b)
            else:
                merged.append([a, b])
        merged = [tuple(m) for m in merged]
        if patch:
            ends = set()
            for c in pts:
                z, x, y = dec(X, Y, c)
                ends.add(c)
                ends.add(enc(X, Y, z + 1, x, y))
            cst = set(merged)
            for a, b in merged:
                ends.update((a, b))
            for c in pins:
                if c < X * Y or c in ends:
                    continue
                below = [iv for iv in sorted(cst) if iv <= (c, c)]
                a, b = below[-1]
                assert a < c < b
                cst.discard((a, b))
                cst.update({(a, c), (c, b)})
            order = [(c, c) for c in pts] + sorted(cst)
        else:
            order = merged + [(c, c) for c in pts]
        for a, b in order:
            z1, x1, y1 = dec(X, Y, a)
            z2, x2, y2 = dec(X, Y, b)
            if dbu:
                zz = z1 + 2 if a == b else z1 + 1
                lines.append(f"{_gran(x1)} {_gran(y1)} metal{z1 + 1} {_gran(x2)} {_gran(y2)} metal{zz}")
            else:
                lines.append(f"{z1} {x1} {y1} {z2} {x2} {y2}")
    return _name(raw_name) + "\n(\n" + "".join(l + "\n" for l in lines) + ")\n"


def routes_text(X, Y, Z, routed, single, batch_ofst, net_ids, edges, lpat, merge=False, patch=False, dbu=False):
    """Whole route file: batches in order, one-pin nets after the first batch."""
    out = []
    slot = 0
    for b in range(len(batch_ofst) - 1):
        for q in range(batch_ofst[b], batch_ofst[b + 1]):
            name, pins = routed[net_ids[q]]
            d = len(pins)
            out.append(net_text(X, Y, Z, name, pins, edges[slot:slot + d], lpat[slot:slot + d], merge, patch, dbu))
            slot += d
        if b == 0:
            for name, c in single:
                z, x, y = dec(X, Y, c)
                out.append(_name(name) + "\n(\n" + "".join(l + "\n" for l in _via_lines(x, y, z, z + 1, dbu)) + ")\n")
    if len(batch_ofst) == 1:
        for name, c in single:
            z, x, y = dec(X, Y, c)
            out.append(_name(name) + "\n(\n" + "".join(l + "\n" for l in _via_lines(x, y, z, z + 1, dbu)) + ")\n")
    return "".join(out)
