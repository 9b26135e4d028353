"""MI355X (gfx950) routing hot path of the IITM 3-D global router.

The per-batch path (L-pattern costing -> root-pair Boruvka tree -> Fenwick
demand commit) runs as hand-written HIP kernels inside ``libiitmgr.so``; this
module is the host-side mirror of the reference's router loop
(``TEST_LPAT::test_gpu_lpat``, l_pattern_files/test_lpat.hpp:231-476) over the
C ABI in ``include/iitmgr.h``.  There is no CPU fallback: without the built
library or a gfx950 device every routing call raises.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "load_library", "RouterError", "Router", "Model", "Tuning", "Design", "route_cost", "gen_grid", "gen_nets",
    "partition_nets", "FMT_FIRST_ACCESS_POINT", "FMT_NEW", "WRITE_DBU", "WRITE_MERGE", "WRITE_PATCH",
    "COST_REF_BIT", "COST_REF_INC", "COST_LIT_EXP", "COMMIT_REF_QUIRK", "COMMIT_FIXED", "COMMIT_PREFIX",
    "WORKLOADS", "device_count", "TREE_ROOT_PAIR", "TREE_TRUE_MST",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libiitmgr.so")

COST_REF_BIT, COST_REF_INC, COST_LIT_EXP = 0, 1, 2
COMMIT_REF_QUIRK, COMMIT_FIXED, COMMIT_PREFIX = 0, 1, 2
TREE_ROOT_PAIR, TREE_TRUE_MST = 0, 1
K_LPAT, K_TREE, K_REFRESH = 0, 1, 2

# Workload families of the synthetic generator: grid shape and net count.
WORKLOADS = {
    1: dict(X=128, Y=128, Z=6, nets=5000),        # plain method, CPU-runnable
    2: dict(X=800, Y=800, Z=10, nets=150000),     # mid-chip, 5% blocked cells
    3: dict(X=2000, Y=2000, Z=10, nets=1500000),  # heavy-tailed pin counts
    4: dict(X=800, Y=800, Z=10, nets=150000),     # same generator as 2
    5: dict(X=2000, Y=2000, Z=10, nets=1500000),  # 2000^2 with cut capacity + hotspots
}


class RouterError(RuntimeError):
    """A C-ABI call returned an error code."""


class Tuning(ctypes.Structure):
    """Schedule parameters (speed only, never results)."""
    _fields_ = [("large_threads", ctypes.c_int32), ("group_max", ctypes.c_int32), ("fork", ctypes.c_int32),
                ("refresh_blocks", ctypes.c_int32), ("lpat_wave", ctypes.c_int32), ("lpat_block", ctypes.c_int32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class _Desc(ctypes.Structure):
    _fields_ = [("X", ctypes.c_int32), ("Y", ctypes.c_int32), ("Z", ctypes.c_int32),
                ("cost_model", ctypes.c_int32), ("commit_mode", ctypes.c_int32),
                ("device", ctypes.c_int32)]


_lib = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64

_SIGS = {
    "gr_version": (ctypes.c_char_p, []),
    "gr_device_count": (ctypes.c_int, []),
    "gr_lut_rmin": (ctypes.c_int, []),
    "gr_lut_size": (ctypes.c_int, []),
    "gr_last_error": (ctypes.c_char_p, [_P]),
    "gr_create": (ctypes.c_int, [ctypes.POINTER(_Desc), _P, ctypes.POINTER(_P)]),
    "gr_create_f32": (ctypes.c_int, [ctypes.POINTER(_Desc), _P, ctypes.POINTER(_P)]),
    "gr_get_remaining_f32": (ctypes.c_int, [_P, _P]),
    "gr_destroy": (None, [_P]),
    "gr_set_nets": (ctypes.c_int, [_P, _I32, _P, _P]),
    "gr_route_batch": (ctypes.c_int, [_P, _P, _I32, _P, _P, _P]),
    "gr_route_batch_static": (ctypes.c_int, [_P, _P, _I32, _P, _P, _P]),
    "gr_static_scratch_bytes": (_I64, []),
    "gr_set_schedule": (ctypes.c_int, [_P, _I32, _P, _P]),
    "gr_route_schedule": (ctypes.c_int, [_P, _I32, _I32]),
    "gr_sync": (ctypes.c_int, [_P]),
    "gr_get_schedule_outputs": (ctypes.c_int, [_P, _P, _P, _P]),
    "gr_schedule_slots": (_I64, [_P]),
    "gr_reset_demand": (ctypes.c_int, [_P]),
    "gr_get_demand": (ctypes.c_int, [_P, _P]),
    "gr_get_prefix": (ctypes.c_int, [_P, _P]),
    "gr_last_pairs": (_I64, [_P]),
    "gr_get_via_terms": (ctypes.c_int, [_P, _P]),
    "gr_get_pair_costs": (ctypes.c_int, [_P, _P, _P]),
    "gr_get_lut": (ctypes.c_int, [_P, _P]),
    "gr_enable_timing": (ctypes.c_int, [_P, ctypes.c_int]),
    "gr_kernel_stats": (ctypes.c_int, [_P, _P, _P]),
    "gr_gen_grid": (ctypes.c_int, [_I32, _U64, _I32, _I32, _I32, _P]),
    "gr_gen_nets": (ctypes.c_int, [_I32, _U64, _I32, _I32, _I32, _I32, _P, _P, _P, _P]),
    "gr_partition_nets": (ctypes.c_int, [_I32, _I32, _I32, _I32, _P, _P, _I32, _P, _P]),
    "gr_mlp_create": (ctypes.c_int, [_P, _I32, _U64, _I32, ctypes.POINTER(_P)]),
    "gr_mlp_destroy": (None, [_P]),
    "gr_mlp_shape": (ctypes.c_int, [_P, _P]),
    "gr_mlp_get_params": (ctypes.c_int, [_P, _P]),
    "gr_mlp_set_params": (ctypes.c_int, [_P, _P]),
    "gr_mlp_collect": (ctypes.c_int, [_P, _P, _I32]),
    "gr_mlp_samples": (_I64, [_P]),
    "gr_mlp_clear_samples": (ctypes.c_int, [_P]),
    "gr_mlp_train": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_float, _U64, ctypes.POINTER(ctypes.c_float)]),
    "gr_mlp_features": (ctypes.c_int, [_P, _P, _P, _P, _I32, _P]),
    "gr_mlp_forward": (ctypes.c_int, [_P, _P, _I32, _P]),
    "gr_mlp_gradient": (ctypes.c_int, [_P, _P, _P, _I32, _P, ctypes.POINTER(ctypes.c_float)]),
    "gr_set_proposal": (ctypes.c_int, [_P, _P, _I32, ctypes.c_float]),
    "gr_set_tuning": (ctypes.c_int, [_P, _P]),
    "gr_get_tuning": (ctypes.c_int, [_P, _P]),
    "gr_autotune": (ctypes.c_int, [_P, _I32, _P, ctypes.POINTER(ctypes.c_double)]),
    "gr_proposal_stats": (ctypes.c_int, [_P, _P, _P]),
    "gr_evaluate": (ctypes.c_int, [_P, _P, _P]),
    "gr_set_tree_mode": (ctypes.c_int, [_P, _I32]),
    "gr_eval_counts": (ctypes.c_int, [_P, _P, _P]),
    "gr_design_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _I32, _I32, ctypes.POINTER(_P)]),
    "gr_design_create": (ctypes.c_int, [_I32, _I32, _I32, _P, _I32, _P, _P, ctypes.POINTER(_P)]),
    "gr_design_free": (None, [_P]),
    "gr_design_info": (ctypes.c_int, [_P, _P]),
    "gr_design_capacity": (ctypes.c_int, [_P, _P]),
    "gr_design_layers": (ctypes.c_int, [_P, _P, _P, _P]),
    "gr_design_layer_name": (ctypes.c_int, [_P, _I32, ctypes.c_char_p, _I32]),
    "gr_design_gcell_lengths": (ctypes.c_int, [_P, _P, _P]),
    "gr_design_nets": (ctypes.c_int, [_P, _P, _P]),
    "gr_design_single_nets": (ctypes.c_int, [_P, _P]),
    "gr_design_net_name": (ctypes.c_int, [_P, _I32, ctypes.c_char_p, _I32]),
    "gr_design_save": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_char_p]),
    "gr_write_routes": (ctypes.c_int, [_P, ctypes.c_char_p, _I32, _I32, _P, _P, _P, _P, _I32,
                                       ctypes.POINTER(_I64)]),
}


def load_library(path=LIB_PATH):
    """Load libiitmgr.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _check(rc, ctx=None):
    if rc < 0:
        lib = load_library()
        msg = lib.gr_last_error(ctx).decode()
        raise RouterError(f"error {rc}: {msg}")
    return rc


def device_count():
    return load_library().gr_device_count()


def gen_grid(kind, seed, X, Y, Z):
    """Capacity cube (int32, WIRE_ENC order) of a synthetic workload family."""
    lib = load_library()
    cap = np.empty(X * Y * Z, dtype=np.int32)
    _check(lib.gr_gen_grid(kind, seed, X, Y, Z, _ptr(cap)))
    return cap


def gen_nets(kind, seed, X, Y, Z, n_nets):
    """Nets of a synthetic workload: (pin_pfx int64[n+1], pin_enc int32[...])."""
    lib = load_library()
    nn = ctypes.c_int32()
    npins = ctypes.c_int64()
    _check(lib.gr_gen_nets(kind, seed, X, Y, Z, n_nets, ctypes.byref(nn), ctypes.byref(npins), None, None))
    pfx = np.empty(nn.value + 1, dtype=np.int64)
    pins = np.empty(npins.value, dtype=np.int32)
    _check(lib.gr_gen_nets(kind, seed, X, Y, Z, n_nets, ctypes.byref(nn), ctypes.byref(npins),
                           _ptr(pfx), _ptr(pins)))
    return pfx, pins


def partition_nets(X, Y, Z, pin_pfx, pin_enc, num_xdim_parts=4):
    """Batches of nets with pairwise disjoint bounding boxes (net_part.hpp:167-379)."""
    lib = load_library()
    n = len(pin_pfx) - 1
    ofst = np.empty(n + 1, dtype=np.int32)
    ids = np.empty(max(n, 1), dtype=np.int32)
    nb = _check(lib.gr_partition_nets(X, Y, Z, n, _ptr(np.ascontiguousarray(pin_pfx, dtype=np.int64)),
                                      _ptr(np.ascontiguousarray(pin_enc, dtype=np.int32)), num_xdim_parts,
                                      _ptr(ofst), _ptr(ids)))
    return ofst[: nb + 1].copy(), ids[:n].copy()


def route_cost(cap, demand):
    """Route quality of a finished pass: wirelength = committed cell-uses (sum
    of demand), overflow = sum over cells of max(0, demand - capacity), and the
    total route cost wirelength + overflow."""
    cap = np.asarray(cap, np.int64)
    d = np.asarray(demand, np.int64)
    wl = int(d.sum())
    of = int(np.maximum(d - cap, 0).sum())
    return {"wirelength": wl, "overflow": of, "total": wl + of}


class _EvalLayer(ctypes.Structure):
    _fields_ = [("wire", ctypes.c_int64), ("via", ctypes.c_int64), ("max_wire", ctypes.c_int32),
                ("max_via", ctypes.c_int32), ("max_cap", ctypes.c_float), ("max_overflow", ctypes.c_double),
                ("of_cost", ctypes.c_double), ("demand", ctypes.c_int64), ("overflow", ctypes.c_double)]


class _EvalTotal(ctypes.Structure):
    _fields_ = [("wire", ctypes.c_int64), ("via", ctypes.c_int64), ("of_cost", ctypes.c_double),
                ("demand", ctypes.c_int64), ("overflow", ctypes.c_double)]


def _as_dict(st):
    return {k: getattr(st, k) for k, _ in st._fields_}


FMT_FIRST_ACCESS_POINT, FMT_NEW = 1, 2
WRITE_DBU, WRITE_MERGE, WRITE_PATCH = 1, 2, 4


class _DesignDesc(ctypes.Structure):
    _fields_ = [("X", ctypes.c_int32), ("Y", ctypes.c_int32), ("Z", ctypes.c_int32), ("n_nets", ctypes.c_int32),
                ("n_pins", ctypes.c_int64), ("n_single", ctypes.c_int32), ("unit_length_cost", ctypes.c_float),
                ("unit_via_cost", ctypes.c_float)]


class Design:
    """Grid capacities and named nets of a design (host only).

    Mirrors READ_DATA (read_data.hpp): ``Design.read(cap, net)`` parses an
    ISPD .cap/.net pair; ``Design(X, Y, Z, cap, pin_pfx, pin_enc)`` builds one
    from arrays.  Routed nets (>= 2 distinct pins) keep file order and are the
    router's net ids; one-pin nets are kept aside for the writer.
    """

    def __init__(self, X=None, Y=None, Z=None, cap=None, pin_pfx=None, pin_enc=None, _handle=None):
        self.lib = load_library()
        if _handle is not None:
            self.h = _handle
        else:
            h = ctypes.c_void_p()
            pfx = np.ascontiguousarray(pin_pfx, dtype=np.int64)
            pins = np.ascontiguousarray(pin_enc, dtype=np.int32)
            c = None if cap is None else np.ascontiguousarray(cap, dtype=np.float32)
            _check(self.lib.gr_design_create(X, Y, Z, _ptr(c), len(pfx) - 1, _ptr(pfx), _ptr(pins), ctypes.byref(h)))
            self.h = h
        info = _DesignDesc()
        _check(self.lib.gr_design_info(self.h, ctypes.byref(info)))
        self.X, self.Y, self.Z = info.X, info.Y, info.Z
        self.n_nets, self.n_pins, self.n_single = info.n_nets, info.n_pins, info.n_single
        self.unit_length_cost, self.unit_via_cost = info.unit_length_cost, info.unit_via_cost

    @classmethod
    def read(cls, cap_path, net_path, flags=0, threads=0):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.gr_design_read(os.fsencode(cap_path), os.fsencode(net_path), flags, threads, ctypes.byref(h)))
        return cls(_handle=h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.gr_design_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def capacity(self):
        out = np.empty(self.X * self.Y * self.Z, dtype=np.float32)
        _check(self.lib.gr_design_capacity(self.h, _ptr(out)))
        return out

    def layers(self):
        d = np.empty(self.Z, np.int32)
        m = np.empty(self.Z, np.int32)
        c = np.empty(self.Z, np.float32)
        _check(self.lib.gr_design_layers(self.h, _ptr(d), _ptr(m), _ptr(c)))
        names = []
        buf = ctypes.create_string_buffer(256)
        for z in range(self.Z):
            _check(self.lib.gr_design_layer_name(self.h, z, buf, 256))
            names.append(buf.value.decode())
        return {"name": names, "dir": d, "min_len": m, "unit_short_cost": c}

    def gcell_lengths(self):
        h = np.empty(max(self.X - 1, 1), np.int64)
        v = np.empty(max(self.Y - 1, 1), np.int64)
        _check(self.lib.gr_design_gcell_lengths(self.h, _ptr(h), _ptr(v)))
        return h[: self.X - 1], v[: self.Y - 1]

    def nets(self):
        pfx = np.empty(self.n_nets + 1, np.int64)
        pins = np.empty(max(self.n_pins, 1), np.int32)
        _check(self.lib.gr_design_nets(self.h, _ptr(pfx), _ptr(pins)))
        return pfx, pins[: self.n_pins]

    def single_nets(self):
        out = np.empty(max(self.n_single, 1), np.int32)
        _check(self.lib.gr_design_single_nets(self.h, _ptr(out)))
        return out[: self.n_single]

    def net_name(self, i):
        buf = ctypes.create_string_buffer(1024)
        n = _check(self.lib.gr_design_net_name(self.h, i, buf, 1024))
        if n >= 1024:
            buf = ctypes.create_string_buffer(n + 1)
            _check(self.lib.gr_design_net_name(self.h, i, buf, n + 1))
        return buf.value.decode()

    def save(self, cap_path=None, net_path=None):
        _check(self.lib.gr_design_save(self.h, None if cap_path is None else os.fsencode(cap_path),
                                       None if net_path is None else os.fsencode(net_path)))

    def write_routes(self, path, batch_ofst, net_ids, edges, lpat, flags=0, threads=0):
        """Write the routes of a schedule (gr_write_routes); returns bytes written."""
        ofst = np.ascontiguousarray(batch_ofst, dtype=np.int32)
        ids = np.ascontiguousarray(net_ids, dtype=np.int32)
        e = np.ascontiguousarray(edges, dtype=np.int64)
        l = np.ascontiguousarray(lpat, dtype=np.int32)
        nbytes = ctypes.c_int64()
        _check(self.lib.gr_write_routes(self.h, os.fsencode(path), flags, len(ofst) - 1, _ptr(ofst), _ptr(ids),
                                        _ptr(e), _ptr(l), threads, ctypes.byref(nbytes)))
        return nbytes.value


class Router:
    """One routing context on one gfx950 device.

    Mirrors the router thread of the reference (test_lpat.hpp:231-476): the
    grid and nets are loaded once, then batches are routed in order.
    """

    def __init__(self, X, Y, Z, cap, cost_model=COST_REF_BIT, commit_mode=COMMIT_REF_QUIRK, device=0):
        """cap: capacity cube in WIRE_ENC order; an integer array gives the
        integer-capacity grid (gr_create), a float array the fp32 one
        (gr_create_f32, capacities as read from a .cap file)."""
        self.lib = load_library()
        self.X, self.Y, self.Z = X, Y, Z
        cap = np.asarray(cap)
        self.f32 = cap.dtype.kind == "f"
        cap = np.ascontiguousarray(cap, dtype=np.float32 if self.f32 else np.int32)
        if cap.size != X * Y * Z:
            raise ValueError("capacity cube has the wrong size")
        d = _Desc(X, Y, Z, cost_model, commit_mode, device)
        h = ctypes.c_void_p()
        create = self.lib.gr_create_f32 if self.f32 else self.lib.gr_create
        rc = create(ctypes.byref(d), _ptr(cap), ctypes.byref(h))
        if rc < 0:
            raise RouterError(f"gr_create error {rc}: {self.lib.gr_last_error(None).decode()}")
        self.h = h
        self.deg = None

    def close(self):
        if getattr(self, "h", None):
            self.lib.gr_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _c(self, rc):
        return _check(rc, self.h)

    def set_nets(self, pin_pfx, pin_enc):
        self.pin_pfx = np.ascontiguousarray(pin_pfx, dtype=np.int64)
        self.pin_enc = np.ascontiguousarray(pin_enc, dtype=np.int32)
        self.deg = np.diff(self.pin_pfx)
        self._c(self.lib.gr_set_nets(self.h, len(self.pin_pfx) - 1, _ptr(self.pin_pfx), _ptr(self.pin_enc)))

    def route_batch(self, net_ids):
        """Route one batch; returns (edges int64[sum deg], lpat int32[sum deg], net_cost f32[n])."""
        ids = np.ascontiguousarray(net_ids, dtype=np.int32)
        slots = int(self.deg[ids].sum()) if len(ids) else 0
        edges = np.empty(slots, dtype=np.int64)
        lpat = np.empty(slots, dtype=np.int32)
        cost = np.empty(len(ids), dtype=np.float32)
        self._c(self.lib.gr_route_batch(self.h, _ptr(ids), len(ids), _ptr(edges), _ptr(lpat), _ptr(cost)))
        return edges, lpat, cost

    def route_batch_static(self, net_ids):
        """route_batch without any device allocation (gr_route_batch_static):
        the batch's device buffers live in the library's static per-GPU array."""
        ids = np.ascontiguousarray(net_ids, dtype=np.int32)
        slots = int(self.deg[ids].sum()) if len(ids) else 0
        edges = np.empty(slots, dtype=np.int64)
        lpat = np.empty(slots, dtype=np.int32)
        cost = np.empty(len(ids), dtype=np.float32)
        self._c(self.lib.gr_route_batch_static(self.h, _ptr(ids), len(ids), _ptr(edges), _ptr(lpat), _ptr(cost)))
        return edges, lpat, cost

    def set_schedule(self, batch_ofst, net_ids):
        self.sched_ofst = np.ascontiguousarray(batch_ofst, dtype=np.int32)
        self.sched_ids = np.ascontiguousarray(net_ids, dtype=np.int32)
        self._c(self.lib.gr_set_schedule(self.h, len(self.sched_ofst) - 1, _ptr(self.sched_ofst),
                                         _ptr(self.sched_ids)))

    def route_schedule(self, first=0, count=None):
        if count is None:
            count = len(self.sched_ofst) - 1 - first
        self._c(self.lib.gr_route_schedule(self.h, first, count))

    def sync(self):
        self._c(self.lib.gr_sync(self.h))

    def schedule_outputs(self):
        slots = self.lib.gr_schedule_slots(self.h)
        edges = np.empty(slots, dtype=np.int64)
        lpat = np.empty(slots, dtype=np.int32)
        cost = np.empty(len(self.sched_ids), dtype=np.float32)
        self._c(self.lib.gr_get_schedule_outputs(self.h, _ptr(edges), _ptr(lpat), _ptr(cost)))
        return edges, lpat, cost

    def reset_demand(self):
        self._c(self.lib.gr_reset_demand(self.h))

    def demand(self):
        out = np.empty(self.X * self.Y * self.Z, dtype=np.int32)
        self._c(self.lib.gr_get_demand(self.h, _ptr(out)))
        return out

    def set_tree_mode(self, true_mst):
        """False: root-pair trees (the reference); True: minimum spanning trees."""
        self._c(self.lib.gr_set_tree_mode(self.h, TREE_TRUE_MST if true_mst else TREE_ROOT_PAIR))

    def evaluate(self):
        """Route report of the resident schedule's routes (gr_evaluate):
        (per-layer dicts, totals dict)."""
        layers = (_EvalLayer * self.Z)()
        tot = _EvalTotal()
        self._c(self.lib.gr_evaluate(self.h, layers, ctypes.byref(tot)))
        return [_as_dict(l) for l in layers], _as_dict(tot)

    def eval_counts(self):
        """Per-cell (wire, via) counts of the last evaluate()."""
        w = np.empty(self.X * self.Y * self.Z, dtype=np.int32)
        v = np.empty(self.X * self.Y * self.Z, dtype=np.int32)
        self._c(self.lib.gr_eval_counts(self.h, _ptr(w), _ptr(v)))
        return w, v

    def remaining_f32(self):
        """Remaining fp32 capacity per cell (fp32 grids only)."""
        out = np.empty(self.X * self.Y * self.Z, dtype=np.float32)
        self._c(self.lib.gr_get_remaining_f32(self.h, _ptr(out)))
        return out

    def prefix(self):
        out = np.empty(self.X * self.Y * self.Z, dtype=np.float32)
        self._c(self.lib.gr_get_prefix(self.h, _ptr(out)))
        return out

    def via_terms(self):
        """Per-cell via cost term (BIT: Q(pos) - Q(pos-1); non-BIT: the layer's via cost)."""
        out = np.empty(self.X * self.Y * self.Z, dtype=np.float32)
        self._c(self.lib.gr_get_via_terms(self.h, _ptr(out)))
        return out

    def pair_costs(self):
        n = self.lib.gr_last_pairs(self.h)
        cost = np.empty(n, dtype=np.float32)
        pat = np.empty(n, dtype=np.int16)
        self._c(self.lib.gr_get_pair_costs(self.h, _ptr(cost), _ptr(pat)))
        return cost, pat

    def lut(self):
        out = np.empty(self.lib.gr_lut_size(), dtype=np.float32)
        self._c(self.lib.gr_get_lut(self.h, _ptr(out)))
        return out

    def tuning(self):
        t = Tuning()
        self._c(self.lib.gr_get_tuning(self.h, ctypes.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._c(self.lib.gr_set_tuning(self.h, ctypes.byref(t)))

    def autotune(self, reps=2):
        """Time every candidate tuning on the resident schedule; keep the fastest."""
        t = Tuning()
        ms = ctypes.c_double()
        self._c(self.lib.gr_autotune(self.h, reps, ctypes.byref(t), ctypes.byref(ms)))
        return t, ms.value

    def set_proposal(self, model, k=8, threshold=0.6):
        """Cost only the model's top-k patterns (None: exhaustive costing)."""
        self._c(self.lib.gr_set_proposal(self.h, model.h if model is not None else None, k, threshold))

    def proposal_stats(self):
        p = ctypes.c_int64()
        f = ctypes.c_int64()
        self._c(self.lib.gr_proposal_stats(self.h, ctypes.byref(p), ctypes.byref(f)))
        return p.value, f.value

    def enable_timing(self, on=True):
        self._c(self.lib.gr_enable_timing(self.h, 1 if on else 0))

    def kernel_stats(self):
        n = np.zeros(3, dtype=np.int64)
        ms = np.zeros(3, dtype=np.float64)
        self._c(self.lib.gr_kernel_stats(self.h, _ptr(n), _ptr(ms)))
        return n, ms


class Model:
    """Learned proposal: MLP (4 + 25 Z) -> hidden -> hidden -> (Z-1) ceil(Z/2)
    scoring L-pattern candidates; bf16 MFMA kernels, trained on the GPU."""

    def __init__(self, router, hidden=256, seed=0, max_samples=1 << 20):
        self.router = router
        self.lib = router.lib
        h = ctypes.c_void_p()
        _check(self.lib.gr_mlp_create(router.h, hidden, seed, max_samples, ctypes.byref(h)), router.h)
        self.h = h
        s = np.zeros(6, np.int32)
        self.lib.gr_mlp_shape(self.h, _ptr(s))
        self.n_in, self.k_in, self.hidden, self.n_out, self.n_out_pad, self.n_params = (int(v) for v in s)

    def close(self):
        if getattr(self, "h", None):
            self.lib.gr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _c(self, rc):
        return _check(rc, self.router.h)

    def params(self):
        out = np.empty(self.n_params, np.float32)
        self._c(self.lib.gr_mlp_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self._c(self.lib.gr_mlp_set_params(self.h, _ptr(p)))

    def collect(self, per_batch):
        self._c(self.lib.gr_mlp_collect(self.router.h, self.h if per_batch else None, per_batch))

    def samples(self):
        return self.lib.gr_mlp_samples(self.h)

    def clear_samples(self):
        self._c(self.lib.gr_mlp_clear_samples(self.h))

    def train(self, epochs=1, batch=1024, lr=0.05, seed=0):
        loss = ctypes.c_float()
        self._c(self.lib.gr_mlp_train(self.h, epochs, batch, lr, seed, ctypes.byref(loss)))
        return loss.value

    def features(self, src, dst):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        out = np.empty((len(src), self.k_in), np.float32)
        self._c(self.lib.gr_mlp_features(self.router.h, self.h, _ptr(src), _ptr(dst), len(src), _ptr(out)))
        return out

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((x.shape[0], self.n_out_pad), np.float32)
        self._c(self.lib.gr_mlp_forward(self.h, _ptr(x), x.shape[0], _ptr(out)))
        return out

    def gradient(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.int32)
        g = np.empty(self.n_params, np.float32)
        loss = ctypes.c_float()
        self._c(self.lib.gr_mlp_gradient(self.h, _ptr(x), _ptr(y), x.shape[0], _ptr(g), ctypes.byref(loss)))
        return g, loss.value
