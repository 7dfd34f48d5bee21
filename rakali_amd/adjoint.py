"""Gradients of a gauge misfit through the GPU step, and their torch binding.

``GaugeMisfit`` wraps ``rkg_adjoint_*`` (include/rakali_amd.h) on a ``Grid``:

    L = sum_{steps s} sum_{gauges k} (h + b - target_{s,k})^2,

with eta = h + b read at the gauge cells after every step, or, with
``stations=(x, y)`` in metres, sampled bilinearly at those points as the
gauge recorder does (rki_gauges.F90).  ``forward``
runs the steps from the grid's current state, keeping a checkpoint every
``checkpoint_interval`` steps; ``backward`` replays each segment and
reverses it, giving dL/db, dL/dn (per-cell Manning n) and dL/d(h, hu, hv)
of the initial state.  The gradient is that of the discrete scheme with its
dt rule: the CFL dt of each step (through the cell of maximum speed), the
t_end bound, or a fixed dt, and the boundary times that follow from them.
``forward_many`` / ``backward_many`` run several sessions (one grid and
stream each) with their launches interleaved step by step.

``BathymetryRoughnessLoss`` exposes the same computation as a
``torch.autograd.Function`` of the bed and Manning fields, starting from
still water.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import FIELDS, check


class GaugeMisfit:
    def __init__(self, grid, n_steps: int, checkpoint_interval: int, gauges=None, targets=None, stations=None):
        self._L = _lib.load()
        self.grid = grid
        self.n_steps = int(n_steps)
        self._h = ctypes.c_void_p()
        if stations is not None:
            if gauges is not None:
                raise ValueError("give gauge cells or stations, not both")
            x = np.ascontiguousarray(stations[0], dtype=np.float64).ravel()
            y = np.ascontiguousarray(stations[1], dtype=np.float64).ravel()
            if x.shape != y.shape:
                raise ValueError("station x and y must have the same length")
            self.n_gauges = len(x)
            t = None if targets is None else self._targets(targets)
            check("rkg_adjoint_create_stations", self._L.rkg_adjoint_create_stations(
                grid.handle, self.n_steps, int(checkpoint_interval), self.n_gauges, x.ctypes.data, y.ctypes.data,
                None if t is None else t.ctypes.data_as(ctypes.c_void_p), ctypes.byref(self._h)))
            return
        gi = np.ascontiguousarray([int(i) for i, _ in gauges], dtype=np.int32)
        gj = np.ascontiguousarray([int(j) for _, j in gauges], dtype=np.int32)
        self.n_gauges = len(gi)
        t = None if targets is None else self._targets(targets)
        check("rkg_adjoint_create", self._L.rkg_adjoint_create(
            grid.handle, self.n_steps, int(checkpoint_interval), self.n_gauges,
            gi.ctypes.data_as(ctypes.c_void_p), gj.ctypes.data_as(ctypes.c_void_p),
            None if t is None else t.ctypes.data_as(ctypes.c_void_p), ctypes.byref(self._h)))

    def _targets(self, targets) -> np.ndarray:
        t = np.ascontiguousarray(targets, dtype=np.float64)
        if t.shape != (self.n_steps, self.n_gauges):
            raise ValueError(f"targets: expected shape {(self.n_steps, self.n_gauges)}, got {t.shape}")
        return t

    def set_targets(self, targets) -> None:
        t = self._targets(targets)
        check("rkg_adjoint_set_targets", self._L.rkg_adjoint_set_targets(self._h, t.ctypes.data_as(ctypes.c_void_p)))

    def forward(self) -> float:
        loss = ctypes.c_double()
        check("rkg_adjoint_forward", self._L.rkg_adjoint_forward(self._h, ctypes.byref(loss)))
        return loss.value

    def gauge_eta(self) -> np.ndarray:
        out = np.empty((self.n_steps, self.n_gauges), dtype=np.float64)
        check("rkg_adjoint_gauge_eta", self._L.rkg_adjoint_gauge_eta(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def backward(self, scale: float = 1.0, tile_rows: int = 0) -> None:
        check("rkg_adjoint_backward", self._L.rkg_adjoint_backward(self._h, float(scale), int(tile_rows)))

    def fold_still_water(self, eta0: float = 0.0) -> None:
        check("rkg_adjoint_fold_still_water", self._L.rkg_adjoint_fold_still_water(self._h, float(eta0)))

    def gradient(self, name: str) -> np.ndarray:
        nx, ny = self.grid.nx, self.grid.ny
        out = np.empty((nx, ny), dtype=np.float64, order="F")
        check("rkg_adjoint_gradient", self._L.rkg_adjoint_gradient(
            self._h, FIELDS[name], out.ctypes.data_as(ctypes.c_void_p), nx, 0))
        return out

    def gradient_device(self, name: str, ptr: int, ld: int) -> None:
        """The gradient copied device to device into `ptr` ((nx, ny) interior,
        x fastest, `ld` >= nx elements per row); returns once it is there."""
        check("rkg_adjoint_gradient", self._L.rkg_adjoint_gradient(
            self._h, FIELDS[name], ctypes.c_void_p(ptr), int(ld), 1))

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        check("rkg_adjoint_timing", self._L.rkg_adjoint_timing(self._h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.rkg_adjoint_destroy(ctypes.byref(self._h))
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


def _handles(sessions):
    arr = (ctypes.c_void_p * len(sessions))(*[s._h.value for s in sessions])
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), arr


def forward_many(sessions) -> np.ndarray:
    """Forward runs of several sessions (each on its own grid and stream)
    queued step by step across them; returns their losses."""
    L = _lib.load()
    ptr, keep = _handles(sessions)
    out = np.zeros(len(sessions))
    check("rkg_adjoint_forward_many", L.rkg_adjoint_forward_many(
        len(sessions), ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return out


def backward_many(sessions, scale: float = 1.0) -> None:
    """Reverse sweeps of several sessions, queued step by step across them;
    each session's gradients equal its own backward()."""
    L = _lib.load()
    ptr, keep = _handles(sessions)
    check("rkg_adjoint_backward_many", L.rkg_adjoint_backward_many(len(sessions), ptr, float(scale)))


def start_still_water(grid, b, n, eta0: float = 0.0) -> None:
    """Bed `b`, Manning field `n`, still water h = max(eta0 - b, 0), at rest,
    and the grid's clock back to t = 0."""
    b = np.asfortranarray(b, dtype=np.float64)
    grid.set_field("b", b)
    grid.set_field("h", np.maximum(eta0 - b, 0.0))
    z = np.zeros_like(b)
    grid.set_field("hu", z)
    grid.set_field("hv", z)
    grid.use_manning_field(True)
    grid.set_field("n", np.asfortranarray(n, dtype=np.float64))
    grid.set_clock(0.0, 0)


def _start_still_water_device(grid, b, n, eta0: float) -> None:
    """start_still_water from float64 device tensors of shape (nx, ny):
    transposed copies (x fastest) go to the grid device to device, and
    h = max(eta0 - b, 0) is formed on the device with the same IEEE
    subtraction and the select of fmax0 (eta0 - b > 0 ? eta0 - b : 0)."""
    import torch
    nx = grid.nx
    bt = b.detach().t().contiguous()                       # (ny, nx), x fastest
    d = eta0 - bt
    ht = torch.where(d > 0.0, d, torch.zeros_like(d))
    zt = torch.zeros_like(bt)
    nt = n.detach().to(torch.float64).t().contiguous()
    torch.cuda.current_stream(b.device).synchronize()      # the grid copies on its own stream
    grid.set_field_device("b", bt.data_ptr(), nx)
    grid.set_field_device("h", ht.data_ptr(), nx)
    grid.set_field_device("hu", zt.data_ptr(), nx)
    grid.set_field_device("hv", zt.data_ptr(), nx)
    grid.use_manning_field(True)
    grid.set_field_device("n", nt.data_ptr(), nx)
    grid.set_clock(0.0, 0)


def _torch_function():
    import torch

    class BathymetryRoughnessLoss(torch.autograd.Function):
        """loss = L(b, n): still water over bed b, per-cell Manning n,
        ``session.n_steps`` steps, gauge misfit against the session's
        targets.  Gradients come from the discrete adjoint.  Tensors on the
        GPU move device to device (no host copy); CPU tensors go through
        host memory."""

        @staticmethod
        def forward(ctx, b, n, session: GaugeMisfit, eta0: float = 0.0):
            on_gpu = b.is_cuda and b.dtype == torch.float64
            if on_gpu:
                _start_still_water_device(session.grid, b, n, eta0)
            else:
                start_still_water(session.grid, b.detach().cpu().numpy(), n.detach().cpu().numpy(), eta0)
            loss = session.forward()
            ctx.session, ctx.eta0, ctx.on_gpu = session, eta0, on_gpu
            ctx.dev, ctx.dtype = b.device, b.dtype
            return torch.tensor(loss, dtype=b.dtype, device=b.device)

        @staticmethod
        def backward(ctx, grad):
            s = ctx.session
            s.backward(scale=float(grad))
            s.fold_still_water(ctx.eta0)
            if ctx.on_gpu:
                nx, ny = s.grid.nx, s.grid.ny
                out = []
                for name in ("b", "n"):
                    t = torch.empty((ny, nx), dtype=torch.float64, device=ctx.dev)
                    torch.cuda.current_stream(ctx.dev).synchronize()   # t is allocated before the copy
                    s.gradient_device(name, t.data_ptr(), nx)
                    out.append(t.t())
                return out[0], out[1], None, None
            gb = torch.from_numpy(np.ascontiguousarray(s.gradient("b"))).to(device=ctx.dev, dtype=ctx.dtype)
            gn = torch.from_numpy(np.ascontiguousarray(s.gradient("n"))).to(device=ctx.dev, dtype=ctx.dtype)
            return gb, gn, None, None

    return BathymetryRoughnessLoss


def bathymetry_roughness_loss(b, n, session: GaugeMisfit, eta0: float = 0.0):
    """torch entry point: ``loss = bathymetry_roughness_loss(b, n, session)``
    with b, n float64 tensors of shape (nx, ny); ``loss.backward()`` fills
    b.grad and n.grad from the adjoint."""
    return _torch_function().apply(b, n, session, eta0)
