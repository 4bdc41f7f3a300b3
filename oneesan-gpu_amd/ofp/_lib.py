"""ctypes binding of libofp.so (include/ofp.h).

The product path is the HIP library; there is no CPU fallback.  Loading fails
loudly when the in-tree library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

PKG = Path(__file__).resolve().parents[1]
LIB_PATH = Path(os.environ.get("OFP_LIB", PKG / "lib" / "libofp.so"))

OFP_OK = 0
OFP_E_ARG = -1
OFP_E_DEVICE = -2
OFP_E_MEMORY = -5
OFP_E_RUNTIME = -6
OFP_E_RANGE = -7
OFP_MAX_N = 27

# Exported names, in the order include/ofp.h declares them.
EXPORTS = (
    "ofp_create", "ofp_plan", "ofp_solve", "ofp_apply_cells", "ofp_prepare",
    "ofp_row_sweep", "ofp_read_main", "ofp_set_launch", "ofp_set_variant", "ofp_set_low_cells", "ofp_set_layout", "ofp_state_counts",
    "ofp_last_error", "ofp_version", "ofp_destroy",
    "ofp_job_create", "ofp_job_enqueue", "ofp_job_info_get", "ofp_job_result", "ofp_job_destroy",
    "ofp_shard_plan", "ofp_shard_sweep", "ofp_group_plan", "ofp_group_sweep", "ofp_set_height_caps", "ofp_height_cap", "ofp_set_xcd_pieces",
    "ofp_set_sorted_cells", "ofp_time_cells", "ofp_set_index_width", "ofp_set_max_batch",
    "ofp_moduli_groups", "ofp_solve_static", "ofp_set_pair_passes", "ofp_group_exchange",
    "ofp_set_group_launch",
)


class OfpError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"ofp status {status}: {message}")
        self.status = status


class Config(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("threads", ctypes.c_int),
        ("blocks_per_cu", ctypes.c_int),
        ("reserve_bytes", ctypes.c_size_t),
        ("use_graphs", ctypes.c_int),
        ("layout", ctypes.c_int),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("main_states", ctypes.c_uint64),
        ("blocked_states", ctypes.c_uint64),
        ("state_bytes", ctypes.c_uint64),
        ("transitions", ctypes.c_uint64),
        ("free_bytes", ctypes.c_uint64),
        ("fits", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("wall_s", ctypes.c_double),
        ("kernel_s", ctypes.c_double),
        ("transitions", ctypes.c_uint64),
        ("main_states", ctypes.c_uint64),
        ("blocked_states", ctypes.c_uint64),
        ("rows", ctypes.c_int),
        ("cells_per_row", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("full_kernel_s", ctypes.c_double),
        ("capped_kernel_s", ctypes.c_double),
        ("full_launches", ctypes.c_int),
        ("capped_launches", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class JobInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("k", ctypes.c_int),
        ("rows_total", ctypes.c_int),
        ("rows_enqueued", ctypes.c_int),
        ("idle", ctypes.c_int),
        ("state_bytes", ctypes.c_uint64),
        ("transitions_per_row", ctypes.c_uint64),
    ]


class ShardSpec(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("k", ctypes.c_int),
        ("moduli", ctypes.POINTER(ctypes.c_uint32)),
        ("slices", ctypes.c_int),
        ("slice", ctypes.c_int),
        ("low_cells", ctypes.c_int),
        ("seed", ctypes.c_uint64),
    ]


class ShardInfo(ctypes.Structure):
    _fields_ = [
        ("first_blocked", ctypes.c_uint64),
        ("blocked_count", ctypes.c_uint64),
        ("first_main", ctypes.c_uint64),
        ("main_count", ctypes.c_uint64),
        ("state_bytes", ctypes.c_uint64),
        ("low_cells", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("kernel_s", ctypes.c_double),
        ("stray_stores", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class GroupSpec(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("k", ctypes.c_int),
        ("moduli", ctypes.POINTER(ctypes.c_uint32)),
        ("low", ctypes.c_int),
        ("high", ctypes.c_int),
        ("first_group", ctypes.c_uint32),
        ("groups", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
    ]


class GroupInfo(ctypes.Structure):
    _fields_ = [
        ("main_count", ctypes.c_uint64),
        ("blocked_count", ctypes.c_uint64),
        ("largest_group", ctypes.c_uint64),
        ("state_bytes", ctypes.c_uint64),
        ("low", ctypes.c_int),
        ("cells", ctypes.c_int),
        ("launches", ctypes.c_int),
        ("kernel_s", ctypes.c_double),
        ("cell_ms", ctypes.c_float * OFP_MAX_N),
        ("steps", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        d = {name: getattr(self, name) for name, _ in self._fields_ if name != "cell_ms"}
        d["cell_ms"] = [float(self.cell_ms[i]) for i in range(min(self.cells, OFP_MAX_N))]
        return d


OFP_MAX_SHARES = 16


class ExchangeInfo(ctypes.Structure):
    _fields_ = [
        ("main_count", ctypes.c_uint64),
        ("blocked_count", ctypes.c_uint64),
        ("shares", ctypes.c_int),
        ("share_first", ctypes.c_uint32 * (OFP_MAX_SHARES + 1)),
        ("main_records", ctypes.c_uint64 * OFP_MAX_SHARES),
        ("blocked_records", ctypes.c_uint64 * OFP_MAX_SHARES),
        ("staging_bytes", ctypes.c_uint64),
        ("moved_bytes", ctypes.c_uint64),
        ("kernel_s", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        s = self.shares
        return {"main_count": self.main_count, "blocked_count": self.blocked_count, "shares": s,
                "share_first": [int(self.share_first[i]) for i in range(s + 1)],
                "main_records": [int(self.main_records[i]) for i in range(s)],
                "blocked_records": [int(self.blocked_records[i]) for i in range(s)],
                "staging_bytes": self.staging_bytes, "moved_bytes": self.moved_bytes,
                "kernel_s": self.kernel_s}


_lib = None


def load() -> ctypes.CDLL:
    """Load libofp.so (once).  Raises if the library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python __graft_entry__.py or "
            "python oneesan-gpu_amd/build.py); there is no CPU fallback")
    lib = ctypes.CDLL(str(LIB_PATH))
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    vp = ctypes.c_void_p
    sig = {
        "ofp_create": (ctypes.c_int32, [ctypes.POINTER(Config), ctypes.POINTER(vp)]),
        "ofp_plan": (ctypes.c_int32, [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(PlanInfo)]),
        "ofp_solve": (ctypes.c_int32, [vp, ctypes.c_int, u32p, ctypes.c_int, u32p, ctypes.POINTER(Stats)]),
        "ofp_apply_cells": (ctypes.c_int32, [vp, ctypes.c_int, u32p, ctypes.c_int, u32p, u32p,
                                             ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "ofp_prepare": (ctypes.c_int32, [vp, ctypes.c_int, u32p, ctypes.c_int]),
        "ofp_row_sweep": (ctypes.c_int32, [vp, ctypes.c_int, ctypes.POINTER(Stats)]),
        "ofp_read_main": (ctypes.c_int32, [vp, ctypes.c_uint64, u32p, ctypes.c_int]),
        "ofp_time_cells": (ctypes.c_int32, [vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
                                            ctypes.c_int]),
        "ofp_set_index_width": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_set_max_batch": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_set_pair_passes": (ctypes.c_int32, [vp, ctypes.c_int, ctypes.c_uint32]),
        "ofp_set_group_launch": (ctypes.c_int32, [vp, ctypes.c_int, ctypes.c_longlong]),
        "ofp_solve_static": (ctypes.c_int32, [vp, ctypes.c_int, u32p, ctypes.c_int, u32p]),
        "ofp_moduli_groups": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                             ctypes.c_int]),
        "ofp_set_launch": (ctypes.c_int32, [vp, ctypes.c_int, ctypes.c_int]),
        "ofp_set_layout": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_set_low_cells": (ctypes.c_int32, [vp, ctypes.c_int, ctypes.c_int]),
        "ofp_set_variant": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_set_height_caps": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_set_xcd_pieces": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_set_sorted_cells": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_height_cap": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, u64p]),
        "ofp_state_counts": (ctypes.c_int32, [ctypes.c_int, u64p, u64p]),
        "ofp_last_error": (ctypes.c_char_p, [vp]),
        "ofp_version": (ctypes.c_char_p, []),
        "ofp_destroy": (None, [vp]),
        "ofp_job_create": (ctypes.c_int32, [vp, ctypes.c_int, u32p, ctypes.c_int, ctypes.POINTER(vp)]),
        "ofp_job_enqueue": (ctypes.c_int32, [vp, ctypes.c_int]),
        "ofp_job_info_get": (ctypes.c_int32, [vp, ctypes.POINTER(JobInfo)]),
        "ofp_job_result": (ctypes.c_int32, [vp, u32p]),
        "ofp_job_destroy": (None, [vp]),
        "ofp_shard_plan": (ctypes.c_int32, [vp, ctypes.POINTER(ShardSpec), ctypes.POINTER(ShardInfo)]),
        "ofp_shard_sweep": (ctypes.c_int32, [vp, ctypes.POINTER(ShardSpec), ctypes.POINTER(ShardInfo), u32p, u32p]),
        "ofp_group_plan": (ctypes.c_int32, [vp, ctypes.POINTER(GroupSpec), ctypes.POINTER(GroupInfo), u64p, u64p]),
        "ofp_group_sweep": (ctypes.c_int32, [vp, ctypes.POINTER(GroupSpec), ctypes.POINTER(GroupInfo), u32p, u32p]),
        "ofp_group_exchange": (ctypes.c_int32, [vp, ctypes.POINTER(GroupSpec), ctypes.c_int,
                                                ctypes.POINTER(ExchangeInfo), u64p, u32p, u64p, u32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _u32_array(values: Sequence[int]):
    arr = (ctypes.c_uint32 * len(values))()
    for i, v in enumerate(values):
        arr[i] = int(v)
    return arr


def state_counts(n: int) -> tuple[int, int]:
    lib = load()
    m, b = ctypes.c_uint64(), ctypes.c_uint64()
    st = lib.ofp_state_counts(n, ctypes.byref(m), ctypes.byref(b))
    if st != OFP_OK:
        raise OfpError(st, f"state_counts(n={n})")
    return m.value, b.value


DEFAULT_CAP_PERCENT = 85


def moduli_groups(k: int, max_per_sweep: int) -> list[int]:
    """Sizes of the sweeps a solve of k moduli makes with at most max_per_sweep each
    (host computation, no GPU needed)."""
    lib = load()
    buf = (ctypes.c_int * max(k, 1))()
    cnt = lib.ofp_moduli_groups(k, max_per_sweep, buf, max(k, 1))
    if cnt < 0:
        raise OfpError(-1, f"moduli_groups(k={k}, max_per_sweep={max_per_sweep})")
    return list(buf[:cnt])


def height_cap(n: int, row: int, p: int, max_percent: int = DEFAULT_CAP_PERCENT) -> tuple[int, int]:
    """(cap, blocked words swept) of cell p in row `row` of an n-solve; cap 0 =
    the whole state.  Host computation, no GPU needed."""
    lib = load()
    cnt = ctypes.c_uint64()
    cap = lib.ofp_height_cap(n, row, p, max_percent, ctypes.byref(cnt))
    if cap < 0:
        raise OfpError(-1, f"height_cap(n={n})")
    return cap, cnt.value


@dataclass
class Plan:
    main_states: int
    blocked_states: int
    state_bytes: int
    transitions: int
    free_bytes: int
    fits: bool


class Context:
    """One GPU context (a device, a stream, one device-resident state)."""

    def __init__(self, device: int = 0, threads: int = 0, blocks_per_cu: int = 0,
                 reserve_bytes: int = 0, use_graphs: bool = True, layout: int | None = None):
        self._lib = load()
        cfg = Config(device, threads, blocks_per_cu, reserve_bytes, 1 if use_graphs else 0,
                     0 if layout is None else layout + 1)
        h = ctypes.c_void_p()
        st = self._lib.ofp_create(ctypes.byref(cfg), ctypes.byref(h))
        if st != OFP_OK:
            raise OfpError(st, "ofp_create failed (no usable GPU?)")
        self._h = h

    def _check(self, st: int, what: str) -> None:
        if st != OFP_OK:
            msg = self._lib.ofp_last_error(self._h)
            raise OfpError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ofp_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, n: int, k: int) -> Plan:
        info = PlanInfo()
        self._check(self._lib.ofp_plan(self._h, n, k, ctypes.byref(info)), "ofp_plan")
        return Plan(info.main_states, info.blocked_states, info.state_bytes, info.transitions,
                    info.free_bytes, bool(info.fits))

    def solve(self, n: int, moduli: Sequence[int]) -> tuple[list[int], dict]:
        k = len(moduli)
        mods = _u32_array(moduli)
        out = (ctypes.c_uint32 * k)()
        stats = Stats()
        self._check(self._lib.ofp_solve(self._h, n, mods, k, out, ctypes.byref(stats)), "ofp_solve")
        return list(out), stats.as_dict()

    def solve_static(self, n: int, moduli: Sequence[int]) -> list[int]:
        """solve() with the state in the library's static device array (no device
        allocation); for small states only (see include/ofp.h)."""
        k = len(moduli)
        out = (ctypes.c_uint32 * k)()
        self._check(self._lib.ofp_solve_static(self._h, n, _u32_array(moduli), k, out), "ofp_solve_static")
        return list(out)

    def apply_cells(self, n: int, moduli: Sequence[int], main, blocked, p_hi: int, p_lo: int,
                    blocked_valid: bool = True) -> None:
        """main / blocked: contiguous numpy uint32 arrays of shape (states, k), updated in place."""
        import numpy as np
        k = len(moduli)
        if main.dtype != np.uint32 or blocked.dtype != np.uint32:
            raise TypeError("state arrays must be uint32")
        if not (main.flags.c_contiguous and blocked.flags.c_contiguous):
            raise ValueError("state arrays must be C-contiguous")
        nm, nb = state_counts(n)
        if main.shape != (nm, k) or blocked.shape != (nb, k):
            raise ValueError(f"expected shapes {(nm, k)} and {(nb, k)}")
        u32p = ctypes.POINTER(ctypes.c_uint32)
        st = self._lib.ofp_apply_cells(self._h, n, _u32_array(moduli), k,
                                       main.ctypes.data_as(u32p), blocked.ctypes.data_as(u32p),
                                       p_hi, p_lo, 1 if blocked_valid else 0)
        self._check(st, "ofp_apply_cells")

    def prepare(self, n: int, moduli: Sequence[int]) -> None:
        self._check(self._lib.ofp_prepare(self._h, n, _u32_array(moduli), len(moduli)), "ofp_prepare")

    def row_sweep(self, rows: int) -> dict:
        stats = Stats()
        self._check(self._lib.ofp_row_sweep(self._h, rows, ctypes.byref(stats)), "ofp_row_sweep")
        return stats.as_dict()

    def read_main(self, index: int, k: int) -> list[int]:
        """Residues of main word `index` of the prepared state; k must be its moduli
        count (the library refuses a smaller buffer rather than overrun it)."""
        out = (ctypes.c_uint32 * k)()
        self._check(self._lib.ofp_read_main(self._h, index, out, k), "ofp_read_main")
        return list(out)

    def time_cells(self, n: int) -> list[tuple[int, float, str]]:
        """One whole-state row of the prepared n-state, each cell or pair pass launched
        and timed on its own: [(p, ms, kernel)] for p = n..1, kernel 'pair_kernel' for
        the first cell of a pair pass (the pass's time) and 'pair_second' (0 ms) for
        its second (advances the state by one row)."""
        cells = n
        ms = (ctypes.c_float * cells)()
        kind = (ctypes.c_int * cells)()
        self._check(self._lib.ofp_time_cells(self._h, ms, kind, cells), "ofp_time_cells")
        names = ("cell_kernel", "cell_sorted_kernel", "pair_kernel", "pair_second")
        return [(cells - i, float(ms[i]), names[kind[i]]) for i in range(cells)]

    def set_index_width(self, bits: int) -> None:
        """0 = by n (32-bit ranks up to n=22), 64 = 64-bit ranks at every n."""
        self._check(self._lib.ofp_set_index_width(self._h, bits), "ofp_set_index_width")

    def set_pair_passes(self, mode: bool | int, cell_mask: int = 0) -> None:
        """Whole-state rows as pair passes (two cells per in-place pass) instead of
        one launch per cell: False/0 off, True/1 on, an even 2..16 = on with at most
        that many moduli per launch; cells p and p-1 pair up where bit p of
        cell_mask is set (0: the library's default).  Bits 8..9 of an int mode pick
        a kernel variant."""
        m = int(mode) if not isinstance(mode, bool) else (1 if mode else 0)
        self._check(self._lib.ofp_set_pair_passes(self._h, m, cell_mask & 0xFFFFFFFF), "ofp_set_pair_passes")

    def set_group_launch(self, streams: int = 0, pair_mask: int = -1) -> None:
        """Group sweeps: streams the group launches of a step are dealt to (1..4,
        0 = 3); cells p and p-1 as one pass per group where bit p of pair_mask is
        set (0 = never, -1 = the library's default: from cell 7 up)."""
        self._check(self._lib.ofp_set_group_launch(self._h, streams, int(pair_mask)), "ofp_set_group_launch")

    def set_max_batch(self, max_moduli: int) -> None:
        """At most max_moduli moduli per sweep of solve() (0 = as many as fit HBM)."""
        self._check(self._lib.ofp_set_max_batch(self._h, max_moduli), "ofp_set_max_batch")

    def set_launch(self, threads: int, blocks_per_cu: int) -> None:
        self._check(self._lib.ofp_set_launch(self._h, threads, blocks_per_cu), "ofp_set_launch")

    def _shard_spec(self, n, moduli, slices, slice_, low_cells, seed):
        mods = _u32_array(moduli)
        spec = ShardSpec(n, len(moduli), ctypes.cast(mods, ctypes.POINTER(ctypes.c_uint32)), slices, slice_,
                         low_cells, seed)
        return spec, mods

    def shard_plan(self, n: int, moduli: Sequence[int], slices: int, slice_: int, low_cells: int = 0,
                   seed: int = 0) -> dict:
        spec, _keep = self._shard_spec(n, moduli, slices, slice_, low_cells, seed)
        info = ShardInfo()
        self._check(self._lib.ofp_shard_plan(self._h, ctypes.byref(spec), ctypes.byref(info)), "ofp_shard_plan")
        return info.as_dict()

    def shard_sweep(self, n: int, moduli: Sequence[int], slices: int, slice_: int, low_cells: int = 0,
                    seed: int = 0, copy_out: bool = False):
        """Sweep cells low_cells..1 over one shard; returns (info, main, blocked) (arrays if copy_out)."""
        import numpy as np
        spec, _keep = self._shard_spec(n, moduli, slices, slice_, low_cells, seed)
        info = ShardInfo()
        self._check(self._lib.ofp_shard_plan(self._h, ctypes.byref(spec), ctypes.byref(info)), "ofp_shard_plan")
        k = len(moduli)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        mo = bo = None
        if copy_out:
            mo = np.zeros((info.main_count, k), np.uint32)
            bo = np.zeros((info.blocked_count, k), np.uint32)
        st = self._lib.ofp_shard_sweep(self._h, ctypes.byref(spec), ctypes.byref(info),
                                       mo.ctypes.data_as(u32p) if copy_out else None,
                                       bo.ctypes.data_as(u32p) if copy_out else None)
        self._check(st, "ofp_shard_sweep")
        return info.as_dict(), mo, bo

    def group_sweep(self, n: int, moduli: Sequence[int], first_group: int, groups: int, low: int = 0,
                    seed: int = 0, main_in=None, blocked_out: bool = False, high: bool = False, blocked_in=None):
        """One half of a row over the occupancy groups first_group..first_group+
        groups-1: the high cells W-1..low+1 over LOW groups (high=False) or the
        low cells low..1 over HIGH groups (high=True).  main_in: None for seeded
        inputs, else a (main_count, k) uint32 array in window order; blocked_in
        (high only): the blocked words before cell low, None for seeded ones.
        Returns (info, main, blocked): main is the result when main_in was
        given, blocked the window's blocked words after cell low+1 when
        blocked_out is set (high=False only; else None)."""
        import numpy as np
        k = len(moduli)
        mods = _u32_array(moduli)
        spec = GroupSpec(n, k, ctypes.cast(mods, ctypes.POINTER(ctypes.c_uint32)), low, 1 if high else 0,
                         first_group, groups, seed)
        info = GroupInfo()
        self._check(self._lib.ofp_group_plan(self._h, ctypes.byref(spec), ctypes.byref(info), None, None),
                    "ofp_group_plan")
        u32p = ctypes.POINTER(ctypes.c_uint32)
        mo = bo = None
        if main_in is not None:
            mo = np.ascontiguousarray(np.asarray(main_in, dtype=np.uint32).reshape(info.main_count, k)).copy()
        if high and blocked_out:
            raise ValueError("the low cells leave no blocked words")
        if blocked_out:
            bo = np.zeros((info.blocked_count, k), np.uint32)
        elif high and blocked_in is not None:
            bo = np.ascontiguousarray(np.asarray(blocked_in, dtype=np.uint32).reshape(info.blocked_count, k)).copy()
        st = self._lib.ofp_group_sweep(self._h, ctypes.byref(spec), ctypes.byref(info),
                                       mo.ctypes.data_as(u32p) if mo is not None else None,
                                       bo.ctypes.data_as(u32p) if bo is not None else None)
        self._check(st, "ofp_group_sweep")
        return info.as_dict(), mo, (bo if blocked_out else None)

    def group_exchange(self, n: int, moduli: Sequence[int], first_group: int, groups: int, low: int = 0,
                       shares: int = 8, download: bool = False):
        """Write the words of the LOW window this context swept last (groups
        first_group..first_group+groups-1, the high half of a row) into staging
        buckets, one per share of the HIGH groups (ofp_group_exchange).  Returns
        (info, buckets): buckets is None unless download is set, else a dict with
        'main_index', 'main_res', 'blocked_index', 'blocked_res' (all buckets in
        order; bucket s of kind x holds info[x + '_records'][s] records)."""
        import numpy as np
        k = len(moduli)
        mods = _u32_array(moduli)
        spec = GroupSpec(n, k, ctypes.cast(mods, ctypes.POINTER(ctypes.c_uint32)), low, 0, first_group, groups, 0)
        ginfo = GroupInfo()
        self._check(self._lib.ofp_group_plan(self._h, ctypes.byref(spec), ctypes.byref(ginfo), None, None),
                    "ofp_group_plan")
        u32p = ctypes.POINTER(ctypes.c_uint32)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        out = None
        args = [None, None, None, None]
        if download:
            out = {"main_index": np.zeros(ginfo.main_count, np.uint64),
                   "main_res": np.zeros((ginfo.main_count, k), np.uint32),
                   "blocked_index": np.zeros(ginfo.blocked_count, np.uint64),
                   "blocked_res": np.zeros((ginfo.blocked_count, k), np.uint32)}
            args = [out["main_index"].ctypes.data_as(u64p), out["main_res"].ctypes.data_as(u32p),
                    out["blocked_index"].ctypes.data_as(u64p), out["blocked_res"].ctypes.data_as(u32p)]
        info = ExchangeInfo()
        self._check(self._lib.ofp_group_exchange(self._h, ctypes.byref(spec), shares, ctypes.byref(info), *args),
                    "ofp_group_exchange")
        return info.as_dict(), out

    def set_variant(self, variant: int) -> None:
        """Cell-kernel variant (state-major layout): 0 = early orbit loads, 1 = late,
        2 = one branch-free orbit path, 3 = as 2 with early stores, 4 = as 2 with the
        sorted assignment on every cell."""
        self._check(self._lib.ofp_set_variant(self._h, variant), "ofp_set_variant")

    def set_xcd_pieces(self, log2_chunks: int) -> None:
        """Deal the blocked-word chunks to the 8 XCDs in pieces of 2**log2_chunks
        chunks, round-robin (-1: one contiguous eighth per XCD)."""
        self._check(self._lib.ofp_set_xcd_pieces(self._h, log2_chunks), "ofp_set_xcd_pieces")

    def set_sorted_cells(self, from_cell: int) -> None:
        """Full-state cells p >= from_cell run the sorted kernel (x=N words dealt to
        their own waves); 0 = never, -1 = the library's choice by moduli per launch."""
        self._check(self._lib.ofp_set_sorted_cells(self._h, from_cell), "ofp_set_sorted_cells")

    def set_height_caps(self, max_percent: int | bool) -> None:
        """Skip, in the first and last rows of a solve, the blocked words whose
        height exceeds what the row can hold or still use, in cells where that
        leaves at most max_percent % of them (True or -1 = the library's default: 75 at
        one modulus, 85 with more; False/0 = off)."""
        if max_percent is True:
            max_percent = -1
        self._check(self._lib.ofp_set_height_caps(self._h, int(max_percent)), "ofp_set_height_caps")

    def set_low_cells(self, low_cells: int, low_kc: int) -> None:
        """Cells p <= low_cells advance at most low_kc moduli per launch."""
        self._check(self._lib.ofp_set_low_cells(self._h, low_cells, low_kc), "ofp_set_low_cells")

    def set_layout(self, layout: int) -> None:
        """0 = 64-state tiles, 1 = state-major; invalidates the prepared state."""
        self._check(self._lib.ofp_set_layout(self._h, layout), "ofp_set_layout")


class Job:
    """A device-resident solve running asynchronously on its own stream."""

    def __init__(self, ctx: Context, n: int, moduli: Sequence[int]):
        self.ctx, self.n, self.moduli = ctx, n, list(moduli)
        h = ctypes.c_void_p()
        ctx._check(ctx._lib.ofp_job_create(ctx._h, n, _u32_array(self.moduli), len(self.moduli),
                                           ctypes.byref(h)), "ofp_job_create")
        self._h = h

    def enqueue(self, rows: int) -> None:
        self.ctx._check(self.ctx._lib.ofp_job_enqueue(self._h, rows), "ofp_job_enqueue")

    def info(self) -> JobInfo:
        inf = JobInfo()
        self.ctx._check(self.ctx._lib.ofp_job_info_get(self._h, ctypes.byref(inf)), "ofp_job_info_get")
        return inf

    def result(self) -> list[int]:
        out = (ctypes.c_uint32 * len(self.moduli))()
        self.ctx._check(self.ctx._lib.ofp_job_result(self._h, out), "ofp_job_result")
        return list(out)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.ctx._lib.ofp_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def shard_plan(n: int, k: int, slices: int, slice_: int, low_cells: int = 0) -> dict:
    """Window of one shard (host computation, no GPU needed)."""
    lib = load()
    mods = _u32_array([4294967291] * k)
    spec = ShardSpec(n, k, ctypes.cast(mods, ctypes.POINTER(ctypes.c_uint32)), slices, slice_, low_cells, 0)
    info = ShardInfo()
    st = lib.ofp_shard_plan(None, ctypes.byref(spec), ctypes.byref(info))
    if st != OFP_OK:
        raise OfpError(st, f"shard_plan(n={n}, k={k}, slices={slices}, slice={slice_})")
    return info.as_dict()


def group_plan(n: int, k: int, first_group: int, groups: int, low: int = 0, high: bool = False):
    """Window totals and per-group (main, blocked) counts of occupancy groups,
    LOW (high=False) or HIGH (host computation, no GPU needed).  Returns
    (info, main_counts, blocked_counts)."""
    import numpy as np
    lib = load()
    mods = _u32_array([4294967291] * k)
    spec = GroupSpec(n, k, ctypes.cast(mods, ctypes.POINTER(ctypes.c_uint32)), low, 1 if high else 0,
                     first_group, groups, 0)
    info = GroupInfo()
    mc = np.zeros(max(groups, 1), np.uint64)
    bc = np.zeros(max(groups, 1), np.uint64)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    st = lib.ofp_group_plan(None, ctypes.byref(spec), ctypes.byref(info), mc.ctypes.data_as(u64p),
                            bc.ctypes.data_as(u64p))
    if st != OFP_OK:
        raise OfpError(st, f"group_plan(n={n}, k={k}, first_group={first_group}, groups={groups}, low={low}, "
                           f"high={high})")
    return info.as_dict(), mc[:groups], bc[:groups]


def version() -> str:
    return load().ofp_version().decode()
