"""ctypes bindings of libglint.so (include/glint.h) and a Python mirror of the
reference's stage interface.

The reference drives one tick as (src/tracing/process.cpp:452-460, 608-612):

    bench->filter();          // partition + border lookup
    bench->reachability();    // distance join + meeting-table upsert
    bench->update_meetings(); // emit ended meetings
    bench->update_schema();   // one level of split/merge

``Workbench`` exposes the same four methods over the HIP implementation, and
``Tracer.process`` runs the tick loop of tracer::process (process.cpp:378-634).
Nothing here falls back to a CPU path: if libglint.so is missing or no GPU is
visible the calls raise.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import gen as _gen

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libglint.so"

GLINT_OK = 0
STATUS_NAMES = {
    0: "GLINT_OK", 1: "GLINT_ERR_INVALID_ARGUMENT", 2: "GLINT_ERR_OUT_OF_MEMORY", 3: "GLINT_ERR_HIP",
    4: "GLINT_ERR_BAD_INPUT", 5: "GLINT_ERR_TABLE_FULL", 6: "GLINT_ERR_NODE_CAPACITY",
    7: "GLINT_ERR_UNIT_CAPACITY", 8: "GLINT_ERR_OUTPUT_CAPACITY", 9: "GLINT_ERR_STATE",
    10: "GLINT_ERR_INJECTED",
}
globals().update({name: code for code, name in STATUS_NAMES.items()})
CANDIDATES_REFERENCE, CANDIDATES_COMPLETE = 0, 1
PRECISION_FP64, PRECISION_FP32_GUARD = 0, 1
SCHEMA_STATIC, SCHEMA_DYNAMIC, SCHEMA_GRID_REBUILD = 0, 1, 2


class GlintError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}" + (f" ({detail})" if detail else ""))


class Config(ctypes.Structure):
    _fields_ = [
        ("num_objects", ctypes.c_uint32),
        ("grid_capacity", ctypes.c_uint32),
        ("min_meet_time", ctypes.c_uint32),
        ("max_meet_time", ctypes.c_uint32),
        ("schema_update_delay", ctypes.c_uint32),
        ("dynamic_schema", ctypes.c_uint32),
        ("start_time", ctypes.c_uint32),
        ("candidate_mode", ctypes.c_uint32),
        ("precision", ctypes.c_uint32),
        ("device", ctypes.c_uint32),
        ("low_x", ctypes.c_double),
        ("low_y", ctypes.c_double),
        ("high_x", ctypes.c_double),
        ("high_y", ctypes.c_double),
        ("reach_distance", ctypes.c_double),
        ("guard_epsilon", ctypes.c_double),
        ("num_meeting_buckets", ctypes.c_uint64),
        ("unit_capacity", ctypes.c_uint64),
        ("meeting_capacity", ctypes.c_uint64),
        ("contact_capacity", ctypes.c_uint64),
        ("node_capacity", ctypes.c_uint32),
        ("record_contacts", ctypes.c_uint32),
        ("graph_ticks", ctypes.c_uint32),
        ("kv_split_num", ctypes.c_uint32),
        ("kv_capacity", ctypes.c_uint64),
        ("kv_area_limit", ctypes.c_double),
        ("meeting_log_capacity", ctypes.c_uint64),
        ("join_tile", ctypes.c_uint32),
        ("table_threads", ctypes.c_uint32),
        ("geography", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class Meeting(ctypes.Structure):
    _fields_ = [("pid_a", ctypes.c_uint32), ("pid_b", ctypes.c_uint32), ("start", ctypes.c_uint32),
                ("end", ctypes.c_uint32)]


class TickStats(ctypes.Structure):
    _fields_ = [
        ("tick", ctypes.c_uint32), ("num_leaves", ctypes.c_uint32), ("num_nodes", ctypes.c_uint32),
        ("num_border", ctypes.c_uint32), ("max_leaf_count", ctypes.c_uint32), ("splits", ctypes.c_uint32),
        ("merges", ctypes.c_uint32), ("error_flags", ctypes.c_uint32),
        ("num_units", ctypes.c_uint64), ("num_contacts", ctypes.c_uint64), ("num_active", ctypes.c_uint64),
        ("num_meetings", ctypes.c_uint64), ("num_rechecks", ctypes.c_uint64), ("table_slots", ctypes.c_uint64),
        ("ms_partition", ctypes.c_float), ("ms_lookup", ctypes.c_float), ("ms_refine", ctypes.c_float),
        ("ms_evaluate", ctypes.c_float), ("ms_schema", ctypes.c_float), ("num_stored", ctypes.c_uint64),
        ("ref_drop_objects", ctypes.c_uint64), ("ref_drop_leaves", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
        ("num_tests", ctypes.c_uint64), ("num_evaluated", ctypes.c_uint64),
    ]


class Totals(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "ticks", "contacts", "contacts_border", "distinct_pairs", "meetings", "units", "border_objects",
        "rechecks", "max_contacts_tick", "max_distinct_tick", "max_meetings_tick", "stored", "max_stored_tick",
        "ref_drop_leaves", "ref_drop_objects", "max_ref_drop_objects_tick", "max_leaf_count", "tests", "evaluated")]


class ServerConfig(ctypes.Structure):
    _fields_ = [("arena_bytes", ctypes.c_uint64), ("max_running", ctypes.c_uint32),
                ("estimate_percent", ctypes.c_uint32), ("min_hits", ctypes.c_uint64),
                ("index_block_ticks", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class QueryInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("state", ctypes.c_uint32), ("hops", ctypes.c_uint32),
                ("hits1", ctypes.c_uint64), ("hits2", ctypes.c_uint64), ("footprint_bytes", ctypes.c_uint64),
                ("submit_tick", ctypes.c_uint32), ("admitted_tick", ctypes.c_uint32), ("retries", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double)]


QUERY_PENDING, QUERY_RUNNING, QUERY_DONE, QUERY_FAILED = range(4)


class KvSummary(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("kept_keys", "oversize_keys", "packed_bytes", "logged_meetings")] + [
        (n, ctypes.c_uint32) for n in ("ctf_count", "sid_count", "oversize_oid_count", "zero_one_count",
                                       "end_time_min", "end_time_max", "o_start_time_min", "o_start_time_max")]


# glint_ctf, one file of the sorted key-value output
CTF_DTYPE = np.dtype([("mbr", "<f4", (4,)), ("key_offset", "<u8"), ("packed_offset", "<u8")] +
                     [(n, "<u4") for n in ("kv_count", "start_time_min", "start_time_max", "end_time_min",
                                           "end_time_max", "bitmap_bytes")] +
                     [(n, "<u2") for n in ("key_bit", "id_bit", "duration_bit", "end_bit", "low_x_bit",
                                           "low_y_bit", "edge_bit", "mbr_bit", "x_grid", "y_grid")] +
                     [("reserved", "<u4")])
KV_KEYS, KV_BOXES, KV_OVERSIZE_KEYS, KV_OVERSIZE_BOXES, KV_SIDS, KV_CTFS, KV_BITMAPS, KV_OVERSIZE_BITMAP, \
    KV_PACKED = range(9)

# (name, restype, argtypes) of every entry point of include/glint.h
class TrHeader(ctypes.Structure):
    _fields_ = [("num_objects", ctypes.c_uint32), ("duration", ctypes.c_uint32), ("low_x", ctypes.c_double),
                ("low_y", ctypes.c_double), ("high_x", ctypes.c_double), ("high_y", ctypes.c_double)]


class Projection(ctypes.Structure):
    _fields_ = [("lon0", ctypes.c_double), ("lat0", ctypes.c_double), ("deg_per_m_lon", ctypes.c_double),
                ("deg_per_m_lat", ctypes.c_double)]


_VP = ctypes.c_void_p
_U32, _U64, _I = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
SIGNATURES = {
    "glint_status_string": (ctypes.c_char_p, [_I]),
    "glint_reach_threshold": (ctypes.c_double, [ctypes.c_double]),
    "glint_last_error_message": (ctypes.c_char_p, [_VP]),
    "glint_config_default": (None, [ctypes.POINTER(Config)]),
    "glint_create": (_I, [ctypes.POINTER(Config), ctypes.POINTER(_VP)]),
    "glint_destroy": (_I, [_VP]),
    "glint_set_positions": (_I, [_VP, _VP, _U32, _U32]),
    "glint_set_positions_device": (_I, [_VP, _VP, _U32, _U32]),
    "glint_partition": (_I, [_VP]),
    "glint_refine": (_I, [_VP]),
    "glint_evaluate": (_I, [_VP, _VP, _U64, ctypes.POINTER(_U64)]),
    "glint_update_schema": (_I, [_VP]),
    "glint_process_tick": (_I, [_VP, _VP, _U32]),
    "glint_run_ticks": (_I, [_VP, _VP, _U32, _U32, _VP]),
    "glint_run_ticks_noalloc": (_I, [_VP, _VP, _U32, _U32, _VP]),
    "glint_synchronize": (_I, [_VP]),
    "glint_stats": (_I, [_VP, ctypes.POINTER(TickStats)]),
    "glint_contacts": (_I, [_VP, _VP, _U64, ctypes.POINTER(_U64)]),
    "glint_leaves": (_I, [_VP, _VP, _VP, _VP, _U64, ctypes.POINTER(_U64)]),
    "glint_active_pairs": (_I, [_VP, _VP, _VP, _U64, ctypes.POINTER(_U64)]),
    "glint_device_meetings": (_I, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_U64)]),
    "glint_debug_inject_failure": (_I, [_VP, ctypes.c_char_p]),
    "glint_ingest_updates": (_I, [_VP, _VP, _VP, _U32, _U32]),
    "glint_server_create": (_I, [_VP, ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)]),
    "glint_server_destroy": (_I, [_VP]),
    "glint_server_last_error": (ctypes.c_char_p, [_VP]),
    "glint_server_tick": (_I, [_VP, _VP, _VP, _VP, _U32, _U32]),
    "glint_server_submit": (_I, [_VP, _VP, _U32, _U32, _U32, _U32, ctypes.POINTER(_U64)]),
    "glint_server_poll": (_I, [_VP, ctypes.POINTER(_U32), ctypes.POINTER(_U32)]),
    "glint_server_index_info": (_I, [_VP, ctypes.POINTER(_U64), ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    "glint_server_drain": (_I, [_VP]),
    "glint_server_query_info": (_I, [_VP, _U64, ctypes.POINTER(QueryInfo)]),
    "glint_server_result": (_I, [_VP, _U64, _VP, _U64, ctypes.POINTER(_U64)]),
    "glint_gen_walk_updates_device": (_I, [ctypes.POINTER(_gen.GenConfig), ctypes.POINTER(_gen.Walkers), _U32, _VP,
                                           _VP, _VP, _U32]),
    "glint_batch_create": (_I, [ctypes.POINTER(Config), _U32, _U32, ctypes.POINTER(_VP)]),
    "glint_batch_destroy": (_I, [_VP]),
    "glint_batch_context": (_I, [_VP, _U32, ctypes.POINTER(_VP)]),
    "glint_batch_last_error": (ctypes.c_char_p, [_VP]),
    "glint_batch_groups": (_I, [_VP, ctypes.POINTER(_U32)]),
    "glint_tr_header_read": (_I, [ctypes.c_char_p, ctypes.POINTER(TrHeader)]),
    "glint_tr_read": (_I, [ctypes.c_char_p, _U32, _U32, _U32, _VP]),
    "glint_tr_read_chunk": (_I, [ctypes.c_char_p, _U32, _U32, _VP]),
    "glint_projection_for_box": (_I, [ctypes.c_double, ctypes.c_double, ctypes.POINTER(Projection)]),
    "glint_project_device": (_I, [ctypes.POINTER(Projection), _VP, _VP, _U64, _U32]),
    "glint_longitude_scale_device": (_I, [_VP, _VP, _U64, _U32]),
    "glint_batch_run_ticks": (_I, [_VP, _VP, _U32, _U32, _VP]),
    "glint_meeting_log_info": (_I, [_VP, ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    "glint_meeting_log_read": (_I, [_VP, _U64, _VP, _U64, ctypes.POINTER(_U64)]),
    "glint_snapshot_size": (_I, [_VP, ctypes.POINTER(_U64)]),
    "glint_snapshot": (_I, [_VP, _VP, _U64]),
    "glint_restore": (_I, [_VP, _VP, _U64]),
    "glint_kv_pending": (_I, [_VP, ctypes.POINTER(_U64)]),
    "glint_kv_flush": (_I, [_VP, ctypes.POINTER(KvSummary)]),
    "glint_kv_part_size": (_I, [_VP, _U32, ctypes.POINTER(_U64)]),
    "glint_kv_copy": (_I, [_VP, _U32, _VP, _U64]),
    "glint_totals_get": (_I, [_VP, ctypes.POINTER(Totals), _I]),
    "glint_set_kernel_profiling": (_I, [_VP, _I]),
    "glint_kernel_times": (_I, [_VP, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(_U64), _U32, ctypes.POINTER(_U32), _I]),
    "glint_device_malloc": (_I, [ctypes.POINTER(_VP), _U64, _U32]),
    "glint_device_free": (_I, [_VP]),
    "glint_device_copy": (_I, [_VP, _VP, _U64]),
    "glint_memcpy_to_device": (_I, [_VP, _VP, _U64]),
    "glint_memcpy_to_host": (_I, [_VP, _VP, _U64]),
    "glint_device_count": (_I, [ctypes.POINTER(_I)]),
    "glint_gen_walk_device": (_I, [ctypes.POINTER(_gen.GenConfig), ctypes.POINTER(_gen.Walkers), _U32, _U32,
                                   _VP, _U32]),
    "glint_gen_validate": (_I, [ctypes.POINTER(_gen.GenConfig)]),
    "glint_gen_init": (_I, [ctypes.POINTER(_gen.GenConfig), ctypes.POINTER(_gen.Walkers), _U32, _U32]),
    "glint_gen_step": (_I, [ctypes.POINTER(_gen.GenConfig), ctypes.POINTER(_gen.Walkers), _U32, _U32, _U32]),
    "glint_gen_hash": (_U64, [_U64, _U64, _U64, _U64]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libglint.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def device_count() -> int:
    n = ctypes.c_int(0)
    lib().glint_device_count(ctypes.byref(n))
    return n.value


def _check(st: int, where: str, ctx=None) -> None:
    if st != GLINT_OK:
        detail = lib().glint_last_error_message(ctx)
        raise GlintError(st, where, detail.decode() if detail else "")


def tr_header(path) -> dict:
    """Header of a reference trace file (objects, ticks, box)."""
    h = TrHeader()
    _check(lib().glint_tr_header_read(str(path).encode(), ctypes.byref(h)), "glint_tr_header_read")
    return {f: getattr(h, f) for f, _ in TrHeader._fields_}


def tr_read(path, num_objects: int, st: int, duration: int) -> np.ndarray:
    """Ticks st..st+duration-1 of a reference trace file as (duration, N, 2)
    FP64 points, loaded as the reference loads them."""
    out = np.zeros((duration, num_objects, 2), dtype=np.float64)
    _check(lib().glint_tr_read(str(path).encode(), num_objects, st, duration, out.ctypes.data), "glint_tr_read")
    return out


def tr_read_chunk(prefix, st: int, num_objects: int) -> np.ndarray:
    """The 100 ticks of a reference chunk file prefix + st + "_" + N + ".tr"."""
    out = np.zeros((100, num_objects, 2), dtype=np.float64)
    _check(lib().glint_tr_read_chunk(str(prefix).encode(), st, num_objects, out.ctypes.data), "glint_tr_read_chunk")
    return out


def projection_for_box(lon_lo: float, lat_lo: float) -> Projection:
    p = Projection()
    _check(lib().glint_projection_for_box(lon_lo, lat_lo, ctypes.byref(p)), "glint_projection_for_box")
    return p


def project_device(proj: Projection, d_lonlat: int, d_xy: int, n: int, device: int = 0) -> None:
    """lon/lat degrees (FP64 AoS, device memory) to planar metres on the GPU."""
    _check(lib().glint_project_device(ctypes.byref(proj), d_lonlat, d_xy, n, device), "glint_project_device")


def longitude_scale_device(d_latitudes: int, d_scale: int, n: int, device: int = 0) -> None:
    """Degrees per metre of longitude at n latitudes (FP64, device memory),
    computed by the device function the geography predicates use."""
    _check(lib().glint_longitude_scale_device(d_latitudes, d_scale, n, device), "glint_longitude_scale_device")


class DeviceBuffer:
    """Raw device allocation owned by Python (no framework needed)."""

    def __init__(self, nbytes: int, device: int = 0):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        _check(lib().glint_device_malloc(ctypes.byref(self.ptr), self.nbytes, device), "glint_device_malloc")

    def upload(self, arr: np.ndarray, offset: int = 0) -> None:
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        _check(lib().glint_memcpy_to_device(ctypes.c_void_p(self.ptr.value + offset), arr.ctypes.data, arr.nbytes),
               "glint_memcpy_to_device")

    def download(self, arr: np.ndarray, offset: int = 0) -> np.ndarray:
        assert arr.flags.c_contiguous and offset + arr.nbytes <= self.nbytes
        _check(lib().glint_memcpy_to_host(arr.ctypes.data, ctypes.c_void_p(self.ptr.value + offset), arr.nbytes),
               "glint_memcpy_to_host")
        return arr

    def address(self, offset: int = 0) -> int:
        return self.ptr.value + offset

    def free(self) -> None:
        if self.ptr.value:
            lib().glint_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def make_config(num_objects: int, box, *, reach_distance: float = 2.0, grid_capacity: int = 50,
                min_meet_time: int = 5, schema_update_delay: int = 1, dynamic_schema: int = SCHEMA_DYNAMIC,
                start_time: int = 0, complete: bool = False, precision: int = PRECISION_FP64,
                guard_epsilon: float = 1e-3, num_meeting_buckets: int = 0, unit_capacity: int = 0,
                meeting_capacity: int = 0, node_capacity: int = 0, record_contacts: bool = False,
                contact_capacity: int = 0, device: int = 0, max_meet_time: int = 4096,
                graph_ticks: int = 0, kv_capacity: int = 0, kv_split_num: int = 10,
                kv_area_limit: float = 100.0, meeting_log_capacity: int = 0, join_tile: int = 0,
                table_threads: int = 0, geography: bool = False) -> Config:
    c = Config()
    lib().glint_config_default(ctypes.byref(c))
    c.num_objects = num_objects
    c.low_x, c.low_y, c.high_x, c.high_y = (float(v) for v in box)
    c.reach_distance = reach_distance
    c.grid_capacity = grid_capacity
    c.min_meet_time = min_meet_time
    c.max_meet_time = max_meet_time
    c.schema_update_delay = schema_update_delay
    c.dynamic_schema = int(dynamic_schema)
    c.start_time = start_time
    c.candidate_mode = CANDIDATES_COMPLETE if complete else CANDIDATES_REFERENCE
    c.precision = precision
    c.guard_epsilon = guard_epsilon
    c.num_meeting_buckets = num_meeting_buckets
    c.unit_capacity = unit_capacity
    c.meeting_capacity = meeting_capacity
    c.node_capacity = node_capacity
    c.record_contacts = int(record_contacts)
    c.contact_capacity = contact_capacity
    c.device = device
    c.graph_ticks = graph_ticks
    c.kv_capacity = kv_capacity
    c.kv_split_num = kv_split_num
    c.kv_area_limit = kv_area_limit
    c.meeting_log_capacity = meeting_log_capacity
    c.join_tile = join_tile
    c.table_threads = table_threads
    c.geography = int(geography)
    return c


class Workbench:
    """One device context; the four stage methods of the reference's workbench."""

    def __init__(self, cfg: Config, handle: ctypes.c_void_p | None = None):
        self.cfg = cfg
        self.n = cfg.num_objects
        self._owned = handle is None
        if handle is None:
            handle = ctypes.c_void_p()
            _check(lib().glint_create(ctypes.byref(cfg), ctypes.byref(handle)), "glint_create")
        self.h = handle
        self._meet_cap = max(1024, 2 * self.n)
        self._meet_buf = (Meeting * self._meet_cap)()

    def close(self) -> None:
        if self._owned and self.h and self.h.value:
            lib().glint_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, st, where):
        _check(st, where, self.h)

    # -- the reference's stage methods ------------------------------------
    def set_points(self, xy: np.ndarray, tick: int) -> None:
        """bench->points = trace + t*N; bench->cur_time = st + t (process.cpp:434-435)."""
        xy = np.ascontiguousarray(xy, dtype=np.float64)
        assert xy.size == 2 * self.n
        self._xy_keep = xy
        self._chk(lib().glint_set_positions(self.h, xy.ctypes.data, self.n, tick), "glint_set_positions")

    def set_points_device(self, d_xy: int, tick: int) -> None:
        self._chk(lib().glint_set_positions_device(self.h, ctypes.c_void_p(d_xy), self.n, tick),
                  "glint_set_positions_device")

    def filter(self) -> None:
        """workbench::filter (filter.cpp:290-313)."""
        self._chk(lib().glint_partition(self.h), "glint_partition")

    def reachability(self) -> None:
        """workbench::reachability (refinement.cpp:66-85)."""
        self._chk(lib().glint_refine(self.h), "glint_refine")

    def update_meetings(self, copy_out: bool = True) -> np.ndarray | None:
        """workbench::update_meetings (meeting.cpp:166-189); returns the tick's
        meetings as an (M, 4) uint32 array (pid_a, pid_b, start, end), sorted."""
        if not copy_out:
            self._chk(lib().glint_evaluate(self.h, None, 0, None), "glint_evaluate")
            return None
        n = ctypes.c_uint64(0)
        st = lib().glint_evaluate(self.h, self._meet_buf, self._meet_cap, ctypes.byref(n))
        if st == 8 and n.value > self._meet_cap:   # output capacity: grow and fetch
            self._meet_cap = int(n.value) * 2
            self._meet_buf = (Meeting * self._meet_cap)()
            ptr = ctypes.c_void_p()
            cnt = ctypes.c_uint64()
            self._chk(lib().glint_device_meetings(self.h, ctypes.byref(ptr), ctypes.byref(cnt)),
                      "glint_device_meetings")
            arr = np.zeros((cnt.value, 4), dtype=np.uint32)
            _check(lib().glint_memcpy_to_host(arr.ctypes.data, ptr, arr.nbytes), "glint_memcpy_to_host")
            order = np.lexsort((arr[:, 2], arr[:, 1], arr[:, 0]))
            return arr[order]
        self._chk(st, "glint_evaluate")
        arr = np.ctypeslib.as_array(self._meet_buf).view(np.uint32).reshape(-1, 4)[: n.value]
        return arr.copy()

    def update_schema(self) -> None:
        """workbench::update_schema (filter.cpp:428-449)."""
        self._chk(lib().glint_update_schema(self.h), "glint_update_schema")

    # -- convenience --------------------------------------------------------
    def tick(self, xy: np.ndarray, t: int) -> np.ndarray:
        self.set_points(xy, t)
        self.filter()
        self.reachability()
        m = self.update_meetings()
        self.update_schema()
        return m

    def run_ticks(self, d_trace: int, first_tick: int, nticks: int, latencies: bool = True) -> np.ndarray | None:
        ms = np.zeros(nticks, dtype=np.float32) if latencies else None
        self._chk(lib().glint_run_ticks(self.h, ctypes.c_void_p(d_trace), first_tick, nticks,
                                        ms.ctypes.data if latencies else None), "glint_run_ticks")
        return ms

    def synchronize(self) -> None:
        self._chk(lib().glint_synchronize(self.h), "glint_synchronize")

    def stats(self) -> dict:
        s = TickStats()
        self._chk(lib().glint_stats(self.h, ctypes.byref(s)), "glint_stats")
        return {f: getattr(s, f) for f, _ in TickStats._fields_}

    def contacts(self) -> np.ndarray:
        n = ctypes.c_uint64()
        self._chk(lib().glint_contacts(self.h, None, 0, ctypes.byref(n)), "glint_contacts")
        out = np.zeros(max(1, n.value), dtype=np.uint64)
        self._chk(lib().glint_contacts(self.h, out.ctypes.data, out.size, ctypes.byref(n)), "glint_contacts")
        return out[: n.value]

    def leaves(self):
        n = ctypes.c_uint64()
        self._chk(lib().glint_leaves(self.h, None, None, None, 0, ctypes.byref(n)), "glint_leaves")
        boxes = np.zeros((max(1, n.value), 4), dtype=np.float64)
        levels = np.zeros(max(1, n.value), dtype=np.uint32)
        counts = np.zeros(max(1, n.value), dtype=np.uint32)
        self._chk(lib().glint_leaves(self.h, boxes.ctypes.data, levels.ctypes.data, counts.ctypes.data,
                                     n.value, ctypes.byref(n)), "glint_leaves")
        k = n.value
        return boxes[:k], levels[:k], counts[:k]

    def active_pairs(self):
        n = ctypes.c_uint64()
        self._chk(lib().glint_active_pairs(self.h, None, None, 0, ctypes.byref(n)), "glint_active_pairs")
        keys = np.zeros(max(1, n.value), dtype=np.uint64)
        starts = np.zeros(max(1, n.value), dtype=np.uint32)
        self._chk(lib().glint_active_pairs(self.h, keys.ctypes.data, starts.ctypes.data, keys.size,
                                           ctypes.byref(n)), "glint_active_pairs")
        return keys[: n.value], starts[: n.value]

    def totals(self, reset: bool = False) -> dict:
        t = Totals()
        self._chk(lib().glint_totals_get(self.h, ctypes.byref(t), int(reset)), "glint_totals_get")
        return {f: getattr(t, f) for f, _ in Totals._fields_}

    def set_kernel_profiling(self, on: bool) -> None:
        self._chk(lib().glint_set_kernel_profiling(self.h, int(on)), "glint_set_kernel_profiling")

    def kernel_times(self, reset: bool = False) -> dict:
        """{kernel name: (total ms, launches)} since the last reset."""
        cap = 64
        names = (ctypes.c_char_p * cap)()
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_uint64 * cap)()
        n = ctypes.c_uint32()
        self._chk(lib().glint_kernel_times(self.h, names, ms, cnt, cap, ctypes.byref(n), int(reset)),
                  "glint_kernel_times")
        return {names[k].decode(): (ms[k], cnt[k]) for k in range(min(n.value, cap))}

    def meeting_log_info(self) -> tuple[int, int]:
        """(meetings ever logged, index of the oldest one retained)."""
        tot, first = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().glint_meeting_log_info(self.h, ctypes.byref(tot), ctypes.byref(first)),
               "glint_meeting_log_info", self.h)
        return int(tot.value), int(first.value)

    def meeting_log(self, start: int | None = None, sort: bool = True) -> np.ndarray:
        """Retained log entries from `start` (default: the oldest retained) as
        an (m, 4) array of (pid_a, pid_b, start, end), sorted if asked."""
        tot, first = self.meeting_log_info()
        start = first if start is None else start
        out = np.zeros((max(tot - start, 1), 4), dtype=np.uint32)
        n = ctypes.c_uint64()
        _check(lib().glint_meeting_log_read(self.h, start, out.ctypes.data, tot - start, ctypes.byref(n)),
               "glint_meeting_log_read", self.h)
        out = out[:n.value]
        if sort and len(out):
            out = out[np.lexsort((out[:, 3], out[:, 2], out[:, 1], out[:, 0]))]
        return out

    def ingest_updates(self, d_pids: int, d_xy: int, k: int, tick: int) -> None:
        """Location updates (device arrays) into the resident positions; starts tick."""
        self._chk(lib().glint_ingest_updates(self.h, ctypes.c_void_p(d_pids), ctypes.c_void_p(d_xy), k, tick),
                  "glint_ingest_updates")

    def kv_pending(self) -> int:
        """Directed keys logged since the last key-value flush."""
        n = ctypes.c_uint64()
        _check(lib().glint_kv_pending(self.h, ctypes.byref(n)), "glint_kv_pending", self.h)
        return int(n.value)

    def kv_flush(self, copy: bool = True) -> dict:
        """Flush the sorted key-value output after the last evaluated tick;
        returns every part as numpy arrays (keys as (lo, hi) u64 pairs), or
        only the summary when copy is False (the parts stay on the device)."""
        sm = KvSummary()
        _check(lib().glint_kv_flush(self.h, ctypes.byref(sm)), "glint_kv_flush", self.h)
        summary = {f: int(getattr(sm, f)) for f, _ in KvSummary._fields_}
        if not copy:
            return {"summary": summary}

        def part(p, dtype, shape=None):
            nb = ctypes.c_uint64()
            _check(lib().glint_kv_part_size(self.h, p, ctypes.byref(nb)), "glint_kv_part_size", self.h)
            out = np.zeros(nb.value // np.dtype(dtype).itemsize, dtype=dtype)
            _check(lib().glint_kv_copy(self.h, p, out.ctypes.data if out.size else None, nb.value),
                   "glint_kv_copy", self.h)
            return out.reshape(shape) if shape else out

        return {
            "summary": summary,
            "keys": part(KV_KEYS, np.uint64, (-1, 2)),
            "boxes": part(KV_BOXES, np.float32, (-1, 4)),
            "okeys": part(KV_OVERSIZE_KEYS, np.uint64, (-1, 2)),
            "oboxes": part(KV_OVERSIZE_BOXES, np.float32, (-1, 4)),
            "sids": part(KV_SIDS, np.uint16),
            "ctfs": part(KV_CTFS, CTF_DTYPE),
            "bitmaps": part(KV_BITMAPS, np.uint8),
            "obitmap": part(KV_OVERSIZE_BITMAP, np.uint8),
            "packed": part(KV_PACKED, np.uint8),
        }

    def snapshot(self) -> np.ndarray:
        """The state the next tick reads, as bytes in host memory (at a tick boundary)."""
        n = ctypes.c_uint64()
        self._chk(lib().glint_snapshot_size(self.h, ctypes.byref(n)), "glint_snapshot_size")
        buf = np.zeros(n.value, dtype=np.uint8)
        self._chk(lib().glint_snapshot(self.h, buf.ctypes.data, buf.size), "glint_snapshot")
        return buf

    def restore(self, buf: np.ndarray) -> None:
        """Continue from a snapshot of a context with the same configuration."""
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        self._chk(lib().glint_restore(self.h, buf.ctypes.data, buf.size), "glint_restore")

    def inject_failure(self, stage: str) -> None:
        self._chk(lib().glint_debug_inject_failure(self.h, stage.encode()), "glint_debug_inject_failure")


class DeviceWalkers:
    """Walker state of the seeded generator held in device memory."""

    def __init__(self, g: _gen.Generator, device: int = 0):
        n = g.w.num_objects
        self.cfg = g.cfg
        self.n = n
        self.device = device
        self.bufs = {
            "xy": DeviceBuffer(16 * n, device), "speed": DeviceBuffer(8 * n, device),
            "tx": DeviceBuffer(4 * n, device), "ty": DeviceBuffer(4 * n, device),
            "dir": DeviceBuffer(max(1, n), device),
        }
        self.bufs["xy"].upload(g.xy)
        self.bufs["speed"].upload(g.speed)
        self.bufs["tx"].upload(g.tx)
        self.bufs["ty"].upload(g.ty)
        self.bufs["dir"].upload(g.dir)
        b = self.bufs
        self.w = _gen.Walkers(ctypes.cast(ctypes.c_void_p(b["xy"].address()), ctypes.POINTER(ctypes.c_double)),
                              ctypes.cast(ctypes.c_void_p(b["speed"].address()), ctypes.POINTER(ctypes.c_double)),
                              ctypes.cast(ctypes.c_void_p(b["tx"].address()), ctypes.POINTER(ctypes.c_int32)),
                              ctypes.cast(ctypes.c_void_p(b["ty"].address()), ctypes.POINTER(ctypes.c_int32)),
                              ctypes.cast(ctypes.c_void_p(b["dir"].address()), ctypes.POINTER(ctypes.c_uint8)))
        self.tick = g.tick

    def walk(self, nticks: int, trace: DeviceBuffer | None = None, trace_offset: int = 0) -> None:
        """Advance nticks ticks; tick k's positions go to trace at trace_offset + k*16N."""
        ptr = ctypes.c_void_p(trace.address(trace_offset)) if trace is not None else None
        _check(lib().glint_gen_walk_device(ctypes.byref(self.cfg), ctypes.byref(self.w), self.tick + 1, nticks,
                                           ptr, self.device), "glint_gen_walk_device")
        self.tick += nticks

    def positions(self) -> np.ndarray:
        out = np.zeros(2 * self.n, dtype=np.float64)
        return self.bufs["xy"].download(out)

    def step_updates(self, pids: DeviceBuffer, xy: DeviceBuffer, count: DeviceBuffer) -> int:
        """One tick; the objects that moved go to (pids, xy) as a location-update
        list; returns its length."""
        _check(lib().glint_gen_walk_updates_device(
            ctypes.byref(self.cfg), ctypes.byref(self.w), self.tick + 1, ctypes.c_void_p(pids.address()),
            ctypes.c_void_p(xy.address()), ctypes.c_void_p(count.address()), self.device),
            "glint_gen_walk_updates_device")
        self.tick += 1
        k = np.zeros(1, dtype=np.uint32)
        count.download(k)
        return int(k[0])


class Tracer:
    """tracer::process (process.cpp:378-634) over a host generator: per tick
    set the positions, then filter, reachability, update_meetings and
    update_schema; collects every emitted meeting."""

    def __init__(self, workload: _gen.Workload, **cfg_kw):
        self.workload = workload
        kw = dict(reach_distance=workload.reach_distance, grid_capacity=workload.grid_capacity,
                  min_meet_time=workload.min_meet_time)
        kw.update(cfg_kw)
        self.bench = Workbench(make_config(workload.num_objects, workload.box(), **kw))
        self.generator = _gen.Generator(workload)

    def process(self, ticks: int | None = None) -> np.ndarray:
        T = self.workload.ticks if ticks is None else ticks
        out = []
        for t in range(T):
            xy = self.generator.positions() if t == 0 else self.generator.step()
            out.append(self.bench.tick(xy, t))
        return np.concatenate(out) if out else np.zeros((0, 4), dtype=np.uint32)

    def close(self):
        self.bench.close()


class Batch:
    """Independent scenarios ticked together on one device (glint_batch_*):
    one context per scenario; with graph_ticks, that many ticks of every
    scenario are replayed as one captured graph."""

    def __init__(self, cfgs: list, graph_ticks: int = 0):
        arr = (Config * len(cfgs))(*cfgs)
        h = ctypes.c_void_p()
        st = lib().glint_batch_create(arr, len(cfgs), graph_ticks, ctypes.byref(h))
        if st != GLINT_OK:
            msg = lib().glint_batch_last_error(None)
            raise GlintError(st, "glint_batch_create", msg.decode() if msg else "")
        self.h = h
        self.cfgs = list(cfgs)
        self.contexts = []
        for i, c in enumerate(cfgs):
            ch = ctypes.c_void_p()
            _check(lib().glint_batch_context(self.h, i, ctypes.byref(ch)), "glint_batch_context")
            self.contexts.append(Workbench(c, handle=ch))

    def groups(self) -> int:
        """Groups of scenarios that share fused launches."""
        g = ctypes.c_uint32()
        _check(lib().glint_batch_groups(self.h, ctypes.byref(g)), "glint_batch_groups")
        return g.value

    def run_ticks(self, d_traces: list, first_tick: int, nticks: int, latencies: bool = True):
        ptrs = (ctypes.c_void_p * len(d_traces))(*[ctypes.c_void_p(p) for p in d_traces])
        ms = np.zeros((len(d_traces), nticks), dtype=np.float32) if latencies else None
        st = lib().glint_batch_run_ticks(self.h, ptrs, first_tick, nticks, ms.ctypes.data if latencies else None)
        if st != GLINT_OK:
            msg = lib().glint_batch_last_error(self.h)
            raise GlintError(st, "glint_batch_run_ticks", msg.decode() if msg else "")
        return ms

    def close(self) -> None:
        for c in self.contexts:
            c.close()
        self.contexts = []
        if self.h and self.h.value:
            lib().glint_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving loop over one context (glint_server_*): ticks of location
    updates, and tracing queries admitted by HBM footprint."""

    def __init__(self, wb: Workbench, arena_bytes: int = 1 << 30, max_running: int = 8, min_hits: int = 4096,
                 estimate_percent: int = 100, index_block_ticks: int = 0):
        self.wb = wb
        sc = ServerConfig(arena_bytes, max_running, estimate_percent, min_hits, index_block_ticks, 0)
        h = ctypes.c_void_p()
        st = lib().glint_server_create(wb.h, ctypes.byref(sc), ctypes.byref(h))
        if st != GLINT_OK:
            raise GlintError(st, "glint_server_create", "the context needs meeting_log_capacity > 0")
        self.h = h

    def _chk(self, st, where):
        if st != GLINT_OK:
            msg = lib().glint_server_last_error(self.h)
            ctx = lib().glint_last_error_message(self.wb.h)
            raise GlintError(st, where, " / ".join(m.decode() for m in (msg, ctx) if m))

    def tick(self, tick: int, *, d_full_xy: int | None = None, d_pids: int = 0, d_xy: int = 0, k: int = 0) -> None:
        self._chk(lib().glint_server_tick(self.h, ctypes.c_void_p(d_full_xy) if d_full_xy else None,
                                          ctypes.c_void_p(d_pids) if d_pids else None,
                                          ctypes.c_void_p(d_xy) if d_xy else None, k, tick), "glint_server_tick")

    def submit(self, seeds, t_start: int, t_end: int, hops: int = 1) -> int:
        seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
        qid = ctypes.c_uint64()
        self._chk(lib().glint_server_submit(self.h, seeds.ctypes.data, seeds.size, t_start, t_end, hops,
                                            ctypes.byref(qid)), "glint_server_submit")
        return int(qid.value)

    def poll(self) -> tuple[int, int]:
        r, p = ctypes.c_uint32(), ctypes.c_uint32()
        self._chk(lib().glint_server_poll(self.h, ctypes.byref(r), ctypes.byref(p)), "glint_server_poll")
        return int(r.value), int(p.value)

    def drain(self) -> None:
        self._chk(lib().glint_server_drain(self.h), "glint_server_drain")

    def index_info(self) -> dict:
        """Blocks of the per-object meeting index, their entries, and the log
        entries below which every meeting is in a block."""
        b, e, sl = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._chk(lib().glint_server_index_info(self.h, ctypes.byref(b), ctypes.byref(e), ctypes.byref(sl)),
                  "glint_server_index_info")
        return {"blocks": b.value, "entries": e.value, "sealed": sl.value}

    def info(self, qid: int) -> dict:
        qi = QueryInfo()
        self._chk(lib().glint_server_query_info(self.h, qid, ctypes.byref(qi)), "glint_server_query_info")
        return {f: getattr(qi, f) for f, _ in QueryInfo._fields_}

    def result(self, qid: int) -> tuple[np.ndarray, np.ndarray]:
        """(hop-1 hits, hop-2 hits) as (m, 4) arrays of (source, target, start, end); releases the query."""
        info = self.info(qid)
        n = ctypes.c_uint64()
        self._chk(lib().glint_server_result(self.h, qid, None, 0, ctypes.byref(n)), "glint_server_result")
        out = np.zeros((max(1, n.value), 4), dtype=np.uint32)
        self._chk(lib().glint_server_result(self.h, qid, out.ctypes.data, out.shape[0], ctypes.byref(n)),
                  "glint_server_result")
        out = out[:n.value]
        return out[:info["hits1"]], out[info["hits1"]:]

    def close(self) -> None:
        if self.h and self.h.value:
            lib().glint_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
