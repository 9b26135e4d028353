"""CPU oracle of the per-batch routing hot path (TEST INFRASTRUCTURE).

A ctypes wrapper over liboracle.so (oracle.cpp, a C++ restatement of the
reference's algorithm).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64

_SIGS = {
    "orc_lut": (None, [ctypes.c_int, _P]),
    "orc_lut_rmin": (ctypes.c_int, []),
    "orc_lut_size": (ctypes.c_int, []),
    "orc_create": (_P, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]),
    "orc_create_f32": (_P, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]),
    "orc_get_remf": (ctypes.c_int, [_P, _P]),
    "orc_set_remf": (ctypes.c_int, [_P, _P, _P]),
    "orc_cw_exp": (ctypes.c_double, [ctypes.c_double]),
    "orc_cost_f32": (ctypes.c_float, [ctypes.c_int, ctypes.c_float]),
    "orc_evaluate": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_int, _P, _P, _P, _P,
                                    _P, _P, _P, _P, _P]),
    "orc_destroy": (None, [_P]),
    "orc_set_nets": (ctypes.c_int, [_P, ctypes.c_int, _P, _P]),
    "orc_get_demand": (ctypes.c_int, [_P, _P]),
    "orc_get_pfx": (ctypes.c_int, [_P, _P]),
    "orc_set_rem": (ctypes.c_int, [_P, _P]),
    "orc_pattern_cost": (ctypes.c_float, [_P, _I32, _I32, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "orc_pattern_cost_walk": (ctypes.c_double, [_P, _I32, _I32, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "orc_best_pattern": (ctypes.c_int, [_P, _I32, _I32, ctypes.POINTER(ctypes.c_float)]),
    "orc_edge_cells": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_int, _P, ctypes.c_int]),
    "orc_boruvka": (ctypes.c_int, [ctypes.c_int, _P, _P, _P, ctypes.c_int]),
    "orc_set_tree_mode": (None, [_P, ctypes.c_int]),
    "orc_route_batch": (_I64, [_P, _P, ctypes.c_int, _P, _P, _P, ctypes.c_int]),
    "orc_last_pairs": (_I64, [_P]),
    "orc_get_vterm": (ctypes.c_int, [_P, _P]),
    "orc_inc_cost": (ctypes.c_float, [ctypes.c_float, ctypes.c_float]),
    "orc_edge_commit_prefix": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_int, _P, ctypes.c_int, _P, _P, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]),
    "orc_partition": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_int, ctypes.c_int,
                                     _P, _P, ctypes.c_int, _P]),
    "orc_get_pair_costs": (ctypes.c_int, [_P, _P, _P]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def lut(model):
    out = np.empty(lib().orc_lut_size(), dtype=np.float32)
    lib().orc_lut(model, _ptr(out))
    return out


def boruvka(n, costs, true_mst=False):
    """Root-pair (or true-MST) Boruvka on explicit pair costs: (edges[(on1, on2)], net_cost, rounds)."""
    c = np.ascontiguousarray(costs, dtype=np.float32)
    out = np.empty(n, dtype=np.int64)
    nc = ctypes.c_float()
    rounds = lib().orc_boruvka(n, _ptr(c), _ptr(out), ctypes.byref(nc), int(true_mst))
    edges = [(int(v >> 15), int(v & 0x7FFF)) for v in out[: n - 1]]
    assert out[n - 1] == -1
    return edges, nc.value, rounds


def partition(X, Y, pin_pfx, pin_enc, parts=4, tile=32, trace_batch=-1):
    """Batches of the reference's batch mode (net_part.hpp:167-339):
    (batch_ofst, net_ids[, per-net bucket in trace_batch]).  tile = 0 tests
    boxes by brute force."""
    pfx = np.ascontiguousarray(pin_pfx, np.int64)
    pins = np.ascontiguousarray(pin_enc, np.int32)
    n = len(pfx) - 1
    ofst = np.empty(n + 1, np.int32)
    ids = np.empty(max(n, 1), np.int32)
    trace = np.empty(max(n, 1), np.int32)
    nb = lib().orc_partition(X, Y, n, _ptr(pfx), _ptr(pins), parts, tile, _ptr(ofst), _ptr(ids), trace_batch,
                             _ptr(trace))
    assert nb >= 0
    res = (ofst[:nb + 1].copy(), ids[:n].copy())
    return res + (trace[:n].copy(),) if trace_batch >= 0 else res


def inc_cost(r, d):
    """f(r - d) - f(r) in fp32 with f(r) = e^(-r/2) (the non-BIT mode's cell costs)."""
    return np.float32(lib().orc_inc_cost(float(np.float32(r)), float(np.float32(d))))


def cw_exp(x):
    """The oracle's e^x in double (Cody-Waite reduction + Taylor polynomial)."""
    return lib().orc_cw_exp(float(x))


def cost_f32(model, r):
    """Cell cost of fp32 remaining capacity r under a cost model."""
    return np.float32(lib().orc_cost_f32(int(model), float(np.float32(r))))


EVAL_FIELDS = ("wire", "via", "max_wire", "max_via", "max_cap", "max_overflow", "of_cost", "demand", "overflow")


def evaluate(X, Y, Z, cap, demand, batch_ofst, net_ids, pin_pfx, pins, edges, lpat):
    """Route evaluation: (per-cell wire counts, via counts, per-layer dict list)."""
    cs = X * Y * Z
    w = np.empty(cs, np.int32)
    v = np.empty(cs, np.int32)
    out = np.empty(9 * Z, np.float64)
    args = [np.ascontiguousarray(cap, np.float32), np.ascontiguousarray(demand, np.int32)]
    ofst = np.ascontiguousarray(batch_ofst, np.int32)
    rest = [ofst, np.ascontiguousarray(net_ids, np.int32), np.ascontiguousarray(pin_pfx, np.int64),
            np.ascontiguousarray(pins, np.int32), np.ascontiguousarray(edges, np.int64),
            np.ascontiguousarray(lpat, np.int32)]
    lib().orc_evaluate(X, Y, Z, _ptr(args[0]), _ptr(args[1]), len(ofst) - 1, *[_ptr(a) for a in rest],
                       _ptr(w), _ptr(v), _ptr(out))
    layers = [dict(zip(EVAL_FIELDS, out[9 * z:9 * z + 9])) for z in range(Z)]
    return w, v, layers


class Oracle:
    def __init__(self, X, Y, Z, cap, cost_model=0, commit_mode=0):
        """An integer capacity array gives the integer-capacity grid, a float
        array the fp32 one."""
        self.X, self.Y, self.Z = X, Y, Z
        cap = np.asarray(cap)
        self.f32 = cap.dtype.kind == "f"
        if self.f32:
            cap = np.ascontiguousarray(cap, dtype=np.float32)
            self.h = lib().orc_create_f32(X, Y, Z, cost_model, commit_mode, _ptr(cap))
        else:
            cap = np.ascontiguousarray(cap, dtype=np.int32)
            self.h = lib().orc_create(X, Y, Z, cost_model, commit_mode, _ptr(cap))
        if not self.h:
            raise ValueError("bad oracle grid")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def enc(self, z, x, y):
        X, Y = self.X, self.Y
        return z * X * Y + (x * Y + y if z & 1 else y * X + x)

    def set_nets(self, pin_pfx, pin_enc):
        self.pin_pfx = np.ascontiguousarray(pin_pfx, dtype=np.int64)
        self.pin_enc = np.ascontiguousarray(pin_enc, dtype=np.int32)
        self.deg = np.diff(self.pin_pfx)
        lib().orc_set_nets(self.h, len(self.pin_pfx) - 1, _ptr(self.pin_pfx), _ptr(self.pin_enc))

    def set_tree_mode(self, true_mst):
        lib().orc_set_tree_mode(self.h, int(bool(true_mst)))

    def set_rem(self, rem):
        r = np.ascontiguousarray(rem, dtype=np.int32)
        lib().orc_set_rem(self.h, _ptr(r))

    def remaining_f32(self):
        out = np.empty(self.X * self.Y * self.Z, dtype=np.float32)
        assert lib().orc_get_remf(self.h, _ptr(out)) == 0
        return out

    def set_remaining_f32(self, remf, demand):
        r = np.ascontiguousarray(remf, dtype=np.float32)
        d = np.ascontiguousarray(demand, dtype=np.int32)
        assert lib().orc_set_remf(self.h, _ptr(r), _ptr(d)) == 0

    def route_batch(self, net_ids, nthreads=0):
        ids = np.ascontiguousarray(net_ids, dtype=np.int32)
        slots = int(self.deg[ids].sum()) if len(ids) else 0
        edges = np.empty(slots, dtype=np.int64)
        lpat = np.empty(slots, dtype=np.int32)
        cost = np.empty(len(ids), dtype=np.float32)
        lib().orc_route_batch(self.h, _ptr(ids), len(ids), _ptr(edges), _ptr(lpat), _ptr(cost), nthreads)
        return edges, lpat, cost

    def demand(self):
        out = np.empty(self.X * self.Y * self.Z, dtype=np.int32)
        lib().orc_get_demand(self.h, _ptr(out))
        return out

    def prefix(self):
        out = np.empty(self.X * self.Y * self.Z, dtype=np.float32)
        lib().orc_get_pfx(self.h, _ptr(out))
        return out

    def pair_costs(self):
        n = lib().orc_last_pairs(self.h)
        cost = np.empty(n, dtype=np.float32)
        pat = np.empty(n, dtype=np.int16)
        lib().orc_get_pair_costs(self.h, _ptr(cost), _ptr(pat))
        return cost, pat

    def pattern_cost(self, src, dst, p):
        v = ctypes.c_int()
        c = lib().orc_pattern_cost(self.h, int(src), int(dst), int(p), ctypes.byref(v))
        return (c if v.value else None)

    def pattern_cost_walk(self, src, dst, p):
        v = ctypes.c_int()
        c = lib().orc_pattern_cost_walk(self.h, int(src), int(dst), int(p), ctypes.byref(v))
        return (c if v.value else None)

    def best_pattern(self, src, dst):
        c = ctypes.c_float()
        p = lib().orc_best_pattern(self.h, int(src), int(dst), ctypes.byref(c))
        return p, c.value

    def via_terms(self):
        """Via cost per cell (non-BIT mode)."""
        out = np.empty(self.X * self.Y * self.Z, dtype=np.float32)
        assert lib().orc_get_vterm(self.h, _ptr(out)) == 0
        return out

    def edge_commit_prefix(self, src, dst, p):
        """Non-BIT commit of one edge: (wire cells with repeats, [(cell, halves)])."""
        w = np.empty(1 << 16, np.int64)
        hc = np.empty(1 << 12, np.int64)
        hn = np.empty(1 << 12, np.int32)
        nh = ctypes.c_int()
        nw = lib().orc_edge_commit_prefix(self.h, int(src), int(dst), int(p), _ptr(w), len(w), _ptr(hc), _ptr(hn),
                                          len(hc), ctypes.byref(nh))
        assert nw <= len(w) and nh.value <= len(hc)
        return w[:nw].copy(), list(zip(hc[:nh.value].tolist(), hn[:nh.value].tolist()))

    def edge_cells(self, src, dst, p):
        buf = np.empty(4096, dtype=np.int64)
        n = lib().orc_edge_cells(self.h, int(src), int(dst), int(p), _ptr(buf), len(buf))
        assert n <= len(buf)
        return buf[:n].copy()
