"""Ensembles of independent structured grids of uneven sizes on one GPU.

``Ensemble`` wraps ``rkg_ensemble_*`` (include/rakali_amd.h): all members
advance together, one kernel launch per RK stage for the whole ensemble,
with a flat or persistent (atomic tile queue) schedule, optional hipGraph
replay of the steps, and an autotuner over schedule and tile height.  Each
member keeps its own clock and physics; its results are bit-identical to a
``Grid`` with the same configuration.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import FIELDS, Config, check
from .grid import make_config

SCHEDULES = {"flat": 0, "persistent": 1}


class Ensemble:
    """``members`` is a sequence of ``Config`` or of ``make_config`` keyword
    dicts (which must include nx, ny, dx, dy)."""

    def __init__(self, members: Sequence, stream: int | None = None):
        self._L = _lib.load()
        cfgs = [m if isinstance(m, Config) else make_config(**m) for m in members]
        if not cfgs:
            raise ValueError("an ensemble needs at least one member")
        arr = (Config * len(cfgs))(*cfgs)
        self._h = ctypes.c_void_p()
        check("rkg_ensemble_create",
              self._L.rkg_ensemble_create(arr, len(cfgs), ctypes.c_void_p(stream), ctypes.byref(self._h)))
        self.shapes = [(c.nx, c.ny) for c in cfgs]

    def __len__(self):
        return len(self.shapes)

    def _ok(self):
        if not self._h:
            raise RuntimeError("ensemble has been closed")

    @property
    def cells(self) -> int:
        self._ok()
        c = ctypes.c_longlong()
        check("rkg_ensemble_size", self._L.rkg_ensemble_size(self._h, None, ctypes.byref(c), None))
        return c.value

    @property
    def tiles(self) -> int:
        self._ok()
        t = ctypes.c_int()
        check("rkg_ensemble_size", self._L.rkg_ensemble_size(self._h, None, None, ctypes.byref(t)))
        return t.value

    def _shape(self, member: int):
        self._ok()
        if not 0 <= member < len(self.shapes):
            raise IndexError(f"member {member} out of range [0, {len(self.shapes)})")
        return self.shapes[member]

    def set_field(self, member: int, name: str, arr) -> None:
        nx, ny = self._shape(member)
        a = np.asfortranarray(arr, dtype=np.float64)
        if a.shape != (nx, ny):
            raise ValueError(f"member {member}: expected shape {(nx, ny)}, got {a.shape}")
        check("rkg_ensemble_upload", self._L.rkg_ensemble_upload(
            self._h, member, FIELDS[name], a.ctypes.data_as(ctypes.c_void_p), nx))

    def get_field(self, member: int, name: str) -> np.ndarray:
        nx, ny = self._shape(member)
        out = np.empty((nx, ny), dtype=np.float64, order="F")
        check("rkg_ensemble_download", self._L.rkg_ensemble_download(
            self._h, member, FIELDS[name], out.ctypes.data_as(ctypes.c_void_p), nx))
        return out

    def step(self, n: int = 1) -> None:
        self._ok()
        check("rkg_ensemble_step", self._L.rkg_ensemble_step(self._h, int(n)))

    def sync(self) -> None:
        self._ok()
        check("rkg_ensemble_sync", self._L.rkg_ensemble_sync(self._h))

    def clock(self, member: int):
        """(t, dt, step_count) of one member."""
        self._ok()
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_longlong()
        check("rkg_ensemble_get_clock", self._L.rkg_ensemble_get_clock(
            self._h, member, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)))
        return t.value, dt.value, n.value

    def total_mass(self, member: int) -> float:
        self._ok()
        m = ctypes.c_double()
        check("rkg_ensemble_total_mass", self._L.rkg_ensemble_total_mass(self._h, member, ctypes.byref(m)))
        return m.value

    def set_schedule(self, schedule: str = "flat", tile_rows: int = 0, persistent_blocks: int = 0,
                     use_graph: bool = False) -> None:
        self._ok()
        check("rkg_ensemble_set_schedule", self._L.rkg_ensemble_set_schedule(
            self._h, SCHEDULES[schedule], int(tile_rows), int(persistent_blocks), int(bool(use_graph))))

    @property
    def schedule(self) -> dict:
        self._ok()
        v = [ctypes.c_int() for _ in range(4)]
        check("rkg_ensemble_get_schedule", self._L.rkg_ensemble_get_schedule(self._h, *[ctypes.byref(x) for x in v]))
        names = {b: a for a, b in SCHEDULES.items()}
        return dict(schedule=names[v[0].value], tile_rows=v[1].value, persistent_blocks=v[2].value,
                    use_graph=bool(v[3].value))

    def autotune(self, steps: int = 5) -> float:
        """Times the candidate schedules, keeps the fastest, restores the
        state; returns the best ms per step."""
        self._ok()
        ms = ctypes.c_double()
        check("rkg_ensemble_autotune", self._L.rkg_ensemble_autotune(self._h, int(steps), ctypes.byref(ms)))
        return ms.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.rkg_ensemble_destroy(ctypes.byref(self._h))
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def from_scenarios(scns, **override) -> Ensemble:
    """Ensemble of scenario members with their fields uploaded."""
    cfgs = []
    for s in scns:
        kw = dict(s.grid_kw)
        kw.update(override)
        cfgs.append(make_config(s.nx, s.ny, s.dx, s.dy, **kw))
    e = Ensemble(cfgs)
    for m, s in enumerate(scns):
        for k in ("b", "h", "hu", "hv"):
            e.set_field(m, k, getattr(s, k))
    return e
