"""Python view of the seeded trace generator (libglint_gen.so, include/glint_gen.h).

The workloads below are the synthetic inputs of the contact-tracing benchmark:
objects walking a 50 m Manhattan street grid at U[1, 15) m/s.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent


class GenConfig(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("num_objects", ctypes.c_uint32),
        ("num_hotspots", ctypes.c_uint32),
        ("side", ctypes.c_double),
        ("street_gap", ctypes.c_double),
        ("speed_min", ctypes.c_double),
        ("speed_max", ctypes.c_double),
        ("hotspot_fraction", ctypes.c_double),
        ("hotspot_sigma", ctypes.c_double),
        ("zipf_s", ctypes.c_double),
        ("move_fraction", ctypes.c_double),
    ]


class Walkers(ctypes.Structure):
    _fields_ = [
        ("xy", ctypes.POINTER(ctypes.c_double)),
        ("speed", ctypes.POINTER(ctypes.c_double)),
        ("tx", ctypes.POINTER(ctypes.c_int32)),
        ("ty", ctypes.POINTER(ctypes.c_int32)),
        ("dir", ctypes.POINTER(ctypes.c_uint8)),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = _PKG / "libglint_gen.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(str(path))
        L.glint_gen_init.argtypes = [ctypes.POINTER(GenConfig), ctypes.POINTER(Walkers),
                                     ctypes.c_uint32, ctypes.c_uint32]
        L.glint_gen_step.argtypes = [ctypes.POINTER(GenConfig), ctypes.POINTER(Walkers),
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.glint_gen_validate.argtypes = [ctypes.POINTER(GenConfig)]
        L.glint_gen_hash.argtypes = [ctypes.c_uint64] * 4
        L.glint_gen_hash.restype = ctypes.c_uint64
        _lib = L
    return _lib


@dataclass
class Workload:
    """One synthetic scenario: generator parameters plus pipeline parameters."""

    name: str
    num_objects: int
    ticks: int
    side: float
    seed: int = 1
    num_hotspots: int = 0
    hotspot_fraction: float = 0.0
    hotspot_sigma: float = 200.0
    zipf_s: float = 1.1
    move_fraction: float = 1.0
    street_gap: float = 50.0
    speed_min: float = 1.0
    speed_max: float = 15.0
    reach_distance: float = 2.0
    min_meet_time: int = 5
    grid_capacity: int = 200

    def gen_config(self) -> GenConfig:
        return GenConfig(self.seed, self.num_objects, self.num_hotspots, self.side, self.street_gap,
                         self.speed_min, self.speed_max, self.hotspot_fraction, self.hotspot_sigma,
                         self.zipf_s, self.move_fraction)

    def box(self) -> tuple[float, float, float, float]:
        return (0.0, 0.0, float(self.side), float(self.side))


def uniform_city(n: int, ticks: int, side: float, seed: int = 1, **kw) -> Workload:
    return Workload(name=f"uniform_{n}", num_objects=n, ticks=ticks, side=side, seed=seed, **kw)


def plain_cpu_reference() -> Workload:
    """100k objects, 100 ticks, 10 km box, r=2, min=5, leaf capacity 200."""
    return Workload("plain_cpu_100k", 100_000, 100, 10_000.0, seed=1, grid_capacity=200)


def plain_gpu_form() -> Workload:
    """1M objects, 1000 ticks, 30 km box, FP64 throughout."""
    return Workload("plain_gpu_1m", 1_000_000, 1000, 30_000.0, seed=2, grid_capacity=200)


def skewed_density(n: int = 10_000_000, ticks: int = 300) -> Workload:
    """10M objects, 300 ticks, 30 km box; 80% in 64 Gaussian hotspots (sigma 200 m,
    Zipf s=1.1), 20% uniform."""
    return Workload("skewed_10m", n, ticks, 30_000.0, seed=3, num_hotspots=64, hotspot_fraction=0.8,
                    hotspot_sigma=200.0, zipf_s=1.1, grid_capacity=200)


def city_side(n: int, gap: float = 50.0) -> float:
    """Box side sqrt(N)*30 m rounded up to a whole number of street blocks."""
    return float(math.ceil(math.sqrt(n) * 30.0 / gap) * gap)


class Generator:
    """Host-side walker state for one workload."""

    def __init__(self, w: Workload):
        self.w = w
        self.cfg = w.gen_config()
        n = w.num_objects
        rc = lib().glint_gen_validate(ctypes.byref(self.cfg))
        if rc:
            raise ValueError(f"invalid generator configuration ({rc})")
        self.xy = np.zeros(2 * n, dtype=np.float64)
        self.speed = np.zeros(n, dtype=np.float64)
        self.tx = np.zeros(n, dtype=np.int32)
        self.ty = np.zeros(n, dtype=np.int32)
        self.dir = np.zeros(n, dtype=np.uint8)
        self._w = Walkers(self.xy.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                          self.speed.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                          self.tx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          self.ty.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          self.dir.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        rc = lib().glint_gen_init(ctypes.byref(self.cfg), ctypes.byref(self._w), 0, n)
        if rc:
            raise ValueError(f"generator init failed ({rc})")
        self.tick = 0

    def positions(self) -> np.ndarray:
        """Current positions, AoS float64 [x0, y0, x1, y1, ...] (a view)."""
        return self.xy

    def step(self) -> np.ndarray:
        self.tick += 1
        rc = lib().glint_gen_step(ctypes.byref(self.cfg), ctypes.byref(self._w), self.tick, 0,
                                  self.w.num_objects)
        if rc:
            raise RuntimeError(f"generator step failed ({rc})")
        return self.xy

    def state(self) -> dict:
        return {"xy": self.xy.copy(), "speed": self.speed.copy(), "tx": self.tx.copy(),
                "ty": self.ty.copy(), "dir": self.dir.copy()}


def batched_cities(count: int = 32, ticks: int = 200, seed: int = 4, n_min: int = 50_000,
                   n_max: int = 2_000_000, grid_capacity: int = 200) -> list[Workload]:
    """Independent cities of uneven size: N log-uniform in [n_min, n_max],
    side sqrt(N)*30 m (whole blocks), reach in {1.5, 2, 3} m, minimum
    meeting in {3, 5, 10} ticks, one seed per city."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(math.floor(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        r = float(rng.choice([1.5, 2.0, 3.0]))
        m = int(rng.choice([3, 5, 10]))
        out.append(Workload(f"city_{i}", n, ticks, city_side(n), seed=1000 * seed + i, reach_distance=r,
                            min_meet_time=m, grid_capacity=grid_capacity))
    return out
