"""ctypes view of liboracle.so, the CPU restatement of the contact-tracing path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

# tree maintenance between ticks (oracle.h ORACLE_SCHEMA_*)
SCHEMA_STATIC, SCHEMA_DYNAMIC, SCHEMA_GRID_REBUILD = 0, 1, 2


class OracleConfig(ctypes.Structure):
    _fields_ = [
        ("num_objects", ctypes.c_uint32),
        ("grid_capacity", ctypes.c_uint32),
        ("min_meet_time", ctypes.c_uint32),
        ("schema_update_delay", ctypes.c_uint32),
        ("dynamic_schema", ctypes.c_uint32),
        ("complete_mode", ctypes.c_uint32),
        ("num_threads", ctypes.c_uint32),
        ("start_time", ctypes.c_uint32),
        ("low_x", ctypes.c_double),
        ("low_y", ctypes.c_double),
        ("high_x", ctypes.c_double),
        ("high_y", ctypes.c_double),
        ("reach_distance", ctypes.c_double),
        ("geography", ctypes.c_uint32),
        ("pad_", ctypes.c_uint32),
    ]


class OracleStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "num_leaves", "num_nodes", "num_border", "num_units", "num_tests", "num_contacts",
        "num_active", "num_meetings", "max_leaf_count", "ref_drop_leaves", "splits", "merges",
        "ref_drop_objects")] + [
        (n, ctypes.c_double) for n in ("t_filter", "t_refine", "t_evaluate", "t_schema")]


class OracleKvConfig(ctypes.Structure):
    _fields_ = [("split_num", ctypes.c_uint32), ("max_meet_time", ctypes.c_uint32), ("area_limit", ctypes.c_double)]


class OracleKvSummary(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("kept_keys", "oversize_keys", "packed_bytes", "logged_meetings")] + [
        (n, ctypes.c_uint32) for n in ("ctf_count", "sid_count", "oversize_oid_count", "zero_one_count",
                                       "end_time_min", "end_time_max", "o_start_time_min", "o_start_time_max")]


# one file (CTF) record, oracle_ctf in oracle.h
CTF_DTYPE = np.dtype([("mbr", "<f4", (4,)), ("key_offset", "<u8"), ("packed_offset", "<u8")] +
                     [(n, "<u4") for n in ("kv_count", "start_time_min", "start_time_max", "end_time_min",
                                           "end_time_max", "bitmap_bytes")] +
                     [(n, "<u2") for n in ("key_bit", "id_bit", "duration_bit", "end_bit", "low_x_bit",
                                           "low_y_bit", "edge_bit", "mbr_bit", "x_grid", "y_grid")] +
                     [("reserved", "<u4")])
assert CTF_DTYPE.itemsize == 80
BITMAP_BYTES = 8192

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return HERE / "liboracle.so"


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = HERE / "liboracle.so"
        if not path.exists():
            build()
        L = ctypes.CDLL(str(path))
        vp = ctypes.c_void_p
        L.oracle_create.argtypes = [ctypes.POINTER(OracleConfig)]
        L.oracle_create.restype = vp
        L.oracle_destroy.argtypes = [vp]
        L.oracle_tick.argtypes = [vp, ctypes.c_void_p, ctypes.c_uint32]
        L.oracle_build.argtypes = [vp, ctypes.c_void_p]
        L.oracle_grid_cells.argtypes = [ctypes.POINTER(OracleConfig), ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_grid_cells.restype = ctypes.c_int
        for f in ("oracle_num_contacts", "oracle_num_meetings", "oracle_num_active", "oracle_num_leaves"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = ctypes.c_uint64
        L.oracle_contacts.argtypes = [vp]
        L.oracle_contacts.restype = ctypes.POINTER(ctypes.c_uint64)
        L.oracle_meetings.argtypes = [vp]
        L.oracle_meetings.restype = ctypes.POINTER(ctypes.c_uint32)
        L.oracle_get_stats.argtypes = [vp, ctypes.POINTER(OracleStats)]
        L.oracle_get_active.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_flush.argtypes = [vp, ctypes.c_void_p, ctypes.c_uint64]
        L.oracle_flush.restype = ctypes.c_uint64
        L.oracle_get_leaves.argtypes = [vp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        L.oracle_get_leaves.restype = ctypes.c_uint64
        L.oracle_locate.argtypes = [vp, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                    ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]
        L.oracle_locate.restype = ctypes.c_int64
        L.oracle_evaluate_contacts.argtypes = [vp, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32]
        L.oracle_update_schema_counts.argtypes = [vp, ctypes.c_void_p, ctypes.c_uint64]
        L.oracle_point_within.argtypes = [ctypes.c_double] * 5
        L.oracle_point_within.restype = ctypes.c_int
        L.oracle_box_distance.argtypes = [ctypes.c_double] * 6
        L.oracle_box_distance.restype = ctypes.c_double
        L.oracle_degree_per_meter_longitude.argtypes = [ctypes.c_double]
        L.oracle_degree_per_meter_longitude.restype = ctypes.c_double
        L.oracle_point_within_geo.argtypes = [ctypes.c_double] * 5
        L.oracle_point_within_geo.restype = ctypes.c_int
        L.oracle_box_distance_geo.argtypes = [ctypes.c_double] * 6
        L.oracle_box_distance_geo.restype = ctypes.c_double
        L.oracle_kv_enable.argtypes = [vp, ctypes.POINTER(OracleKvConfig)]
        L.oracle_kv_pending.argtypes = [vp]
        L.oracle_kv_pending.restype = ctypes.c_uint64
        L.oracle_kv_flush.argtypes = [vp, ctypes.POINTER(OracleKvSummary)]
        for f in ("oracle_kv_keys", "oracle_kv_boxes"):
            getattr(L, f).argtypes = [vp, ctypes.c_int]
            getattr(L, f).restype = ctypes.c_void_p
        for f in ("oracle_kv_sids", "oracle_kv_ctfs", "oracle_kv_bitmaps", "oracle_kv_obitmap", "oracle_kv_packed"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = ctypes.c_void_p
        _lib = L
    return _lib


class Oracle:
    """One deterministic reference-semantics run (partition -> refine -> evaluate)."""

    def __init__(self, num_objects: int, box, reach_distance: float = 2.0, grid_capacity: int = 200,
                 min_meet_time: int = 5, schema_update_delay: int = 1, dynamic_schema: int = 1,
                 complete: bool = False, num_threads: int = 0, start_time: int = 0, geography: bool = False):
        lx, ly, hx, hy = box
        self.cfg = OracleConfig(num_objects, grid_capacity, min_meet_time, schema_update_delay,
                                int(dynamic_schema), int(complete), num_threads, start_time,
                                lx, ly, hx, hy, reach_distance, int(geography), 0)
        self.h = lib().oracle_create(ctypes.byref(self.cfg))
        if not self.h:
            raise ValueError("invalid oracle configuration")
        self.n = num_objects

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid_cells(self, xy: np.ndarray) -> np.ndarray:
        """Cell x + 16384*y of every point, as the grid rebuild takes them."""
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        cell = np.zeros(self.n, dtype=np.uint32)
        if lib().oracle_grid_cells(ctypes.byref(self.cfg), xy.ctypes.data, cell.ctypes.data):
            raise ValueError("a point has no cell")
        return cell

    def tick(self, xy: np.ndarray, t: int) -> None:
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        assert xy.size == 2 * self.n
        rc = lib().oracle_tick(self.h, xy.ctypes.data, t)
        if rc:
            raise ValueError(f"oracle_tick failed ({rc})")

    def build(self, xy: np.ndarray) -> None:
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        lib().oracle_build(self.h, xy.ctypes.data)

    def contacts(self) -> np.ndarray:
        n = lib().oracle_num_contacts(self.h)
        if n == 0:
            return np.zeros(0, dtype=np.uint64)
        return np.ctypeslib.as_array(lib().oracle_contacts(self.h), shape=(n,)).copy()

    def meetings(self) -> np.ndarray:
        n = lib().oracle_num_meetings(self.h)
        if n == 0:
            return np.zeros((0, 4), dtype=np.uint32)
        return np.ctypeslib.as_array(lib().oracle_meetings(self.h), shape=(n * 4,)).reshape(n, 4).copy()

    def active(self) -> tuple[np.ndarray, np.ndarray]:
        n = lib().oracle_num_active(self.h)
        k = np.zeros(n, dtype=np.uint64)
        s = np.zeros(n, dtype=np.uint32)
        if n:
            lib().oracle_get_active(self.h, k.ctypes.data, s.ctypes.data)
        return k, s

    def flush(self) -> np.ndarray:
        n = lib().oracle_num_active(self.h)
        out = np.zeros((max(n, 1), 4), dtype=np.uint32)
        k = lib().oracle_flush(self.h, out.ctypes.data, n)
        return out[:k]

    def stats(self) -> dict:
        s = OracleStats()
        lib().oracle_get_stats(self.h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in OracleStats._fields_}

    def leaves(self):
        n = lib().oracle_num_leaves(self.h)
        boxes = np.zeros((n, 4), dtype=np.float64)
        levels = np.zeros(n, dtype=np.uint32)
        counts = np.zeros(n, dtype=np.uint32)
        k = lib().oracle_get_leaves(self.h, boxes.ctypes.data, levels.ctypes.data, counts.ctypes.data, n)
        assert k == n
        return boxes, levels, counts

    def locate(self, x: float, y: float):
        lo = ctypes.c_int64()
        hi = ctypes.c_int64()
        nc = ctypes.c_uint32()
        cands = np.zeros(4096, dtype=np.uint32)
        leaf = lib().oracle_locate(self.h, x, y, ctypes.byref(lo), ctypes.byref(hi), cands.ctypes.data,
                                   4096, ctypes.byref(nc))
        return int(leaf), (lo.value, hi.value), sorted(int(v) for v in cands[:nc.value])

    def evaluate_contacts(self, keys, t: int) -> None:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        lib().oracle_evaluate_contacts(self.h, keys.ctypes.data, keys.size, t)

    def kv_enable(self, split_num: int = 10, max_meet_time: int = 4096, area_limit: float = 100.0) -> None:
        kc = OracleKvConfig(split_num, max_meet_time, area_limit)
        if lib().oracle_kv_enable(self.h, ctypes.byref(kc)):
            raise ValueError("invalid key-value configuration (or enabled after the first tick)")

    def kv_pending(self) -> int:
        return int(lib().oracle_kv_pending(self.h))

    def kv_flush(self) -> dict:
        """Organise the keys collected since the last flush; returns every
        output of the flush as numpy arrays."""
        sm = OracleKvSummary()
        rc = lib().oracle_kv_flush(self.h, ctypes.byref(sm))
        summary = {f: int(getattr(sm, f)) for f, _ in OracleKvSummary._fields_}
        if rc:
            raise ValueError(f"key-value flush failed: {summary}")
        L = lib()

        def arr(ptr, dtype, count):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            buf = (ctypes.c_uint8 * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dtype).copy()

        nk, no, C = summary["kept_keys"], summary["oversize_keys"], summary["ctf_count"]
        return {
            "summary": summary,
            "keys": arr(L.oracle_kv_keys(self.h, 0), np.uint64, 2 * nk).reshape(-1, 2),
            "boxes": arr(L.oracle_kv_boxes(self.h, 0), np.float32, 4 * nk).reshape(-1, 4),
            "okeys": arr(L.oracle_kv_keys(self.h, 1), np.uint64, 2 * no).reshape(-1, 2),
            "oboxes": arr(L.oracle_kv_boxes(self.h, 1), np.float32, 4 * no).reshape(-1, 4),
            "sids": arr(L.oracle_kv_sids(self.h), np.uint16, self.n),
            "ctfs": arr(L.oracle_kv_ctfs(self.h), CTF_DTYPE, C),
            "bitmaps": arr(L.oracle_kv_bitmaps(self.h), np.uint8, C * BITMAP_BYTES),
            "obitmap": arr(L.oracle_kv_obitmap(self.h), np.uint8, BITMAP_BYTES),
            "packed": arr(L.oracle_kv_packed(self.h), np.uint8, summary["packed_bytes"]),
        }

    def update_schema_counts(self, counts) -> None:
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        rc = lib().oracle_update_schema_counts(self.h, counts.ctypes.data, counts.size)
        if rc:
            raise ValueError("leaf count mismatch")


def point_within(x1, y1, x2, y2, r) -> bool:
    return bool(lib().oracle_point_within(x1, y1, x2, y2, r))


def degree_per_meter_longitude(latitude: float) -> float:
    return lib().oracle_degree_per_meter_longitude(latitude)


def point_within_geo(x1, y1, x2, y2, r) -> bool:
    return bool(lib().oracle_point_within_geo(x1, y1, x2, y2, r))


def box_distance_geo(lx, ly, hx, hy, px, py) -> float:
    return lib().oracle_box_distance_geo(lx, ly, hx, hy, px, py)


def box_distance(lx, ly, hx, hy, px, py) -> float:
    return lib().oracle_box_distance(lx, ly, hx, hy, px, py)


def pair_key(a, b) -> int:
    a, b = (a, b) if a < b else (b, a)
    return (int(a) << 32) | int(b)
