"""FP32 field storage with FP64 residuals and errors.

The stencil path is bandwidth-bound, so storing fields in FP32 halves the
bytes moved per cell update.  What FP32 storage costs in accuracy is measured,
not assumed: a *twin* runs the same kernel on the same inputs twice, once with
``float`` meaning FP32 and once FP64, and reports for each field the L2 and
Linf norms of (FP32 - FP64) over its interior, computed on the device in FP64
(``norms.diff_norm``), together with the FP64 residual of the pressure
Poisson equation of both runs (``norms.poisson_residual``).

The FP64 run is the reference: it is bit-exact with the oracle (the restated
reference CPU path) wherever the test suite compares them.

No counterpart in the reference, whose ``precision`` option is process-wide
(``xgrid/util/init.py:35-36``); the twin switches it around each build.
"""
from __future__ import annotations

import numpy as np

from . import config as _config_mod
from . import norms
from .backend import build_kernel
from .grid import Grid


class precision_scope:
    """Temporarily select what ``float`` means (kernel parse/build and Grid dtype)."""

    def __init__(self, precision: str):
        self.precision = precision

    def __enter__(self):
        self.saved = _config_mod._config
        base = self.saved
        kw = {} if base is None else dict(cacheroot=base.cacheroot, device=base.device, graphs=base.graphs,
                                          fuse=base.fuse, blocked=base.blocked)
        _config_mod.init(precision=self.precision, **kw)
        return self

    def __exit__(self, *exc):
        _config_mod._config = self.saved
        return False


def _built(op):
    op.native, op.depth = build_kernel(op)
    return op


def _rel(e: dict, ref: dict) -> dict:
    out = dict(e)
    out["rel_l2"] = e["l2"] / ref["l2"] if ref["l2"] else (0.0 if e["l2"] == 0 else float("inf"))
    out["rel_linf"] = e["linf"] / ref["linf"] if ref["linf"] else (0.0 if e["linf"] == 0 else float("inf"))
    return out


class CavityTwin:
    """The lid-driven cavity in FP32 and FP64 on the same inputs."""

    FIELDS = "bpuv"

    def __init__(self, n: int, dt: float = 2e-6, nu: float = 0.1, rho: float = 1.0):
        from . import workloads as W
        self.n = n
        self.cfg = W.cavity_config(n, dt=dt, nu=nu, rho=rho)
        masks = W.cavity_masks(n)
        self.kernels, self.fields = {}, {}
        for prec in ("double", "float"):
            with precision_scope(prec):
                self.kernels[prec] = _built(W.make_cavity())
                grids = {}
                for c in self.FIELDS:
                    g = Grid((n, n), float)
                    g.boundary[...] = masks[c]
                    grids[c] = g
                self.fields[prec] = grids
        self.f64, self.f32 = self.fields["double"], self.fields["float"]
        self.steps = 0

    def run(self, steps: int, which=("double", "float")) -> None:
        for prec in which:
            k, g = self.kernels[prec], self.fields[prec]
            for _ in range(steps):
                k(*(g[c] for c in self.FIELDS), self.cfg)
        self.steps += steps

    def residual(self, prec: str) -> dict:
        g = self.fields[prec]
        cy, cx = 1.0 / self.cfg.dy ** 2, 1.0 / self.cfg.dx ** 2
        return norms.poisson_residual(g["p"], g["b"], [cy, cx])

    def report(self) -> dict:
        err = {}
        for c in "puv":
            e = norms.diff_norm(self.f32[c], self.f64[c])
            err[c] = _rel(e, norms.diff_norm(self.f64[c]))
        return {"n": self.n, "steps": self.steps, "error": err,
                "residual": {"f64": self.residual("double"), "f32": self.residual("float")}}


class Heat3DTwin:
    """The 3-D heat equation (7-point FTCS) in FP32 and FP64 on the same seeded input."""

    def __init__(self, n: int, nu: float = 0.1, seed: int = 0):
        from . import workloads as W
        self.n = n
        self.nu = nu
        self.kernels, self.fields = {}, {}
        for prec, dt_np in (("double", np.float64), ("float", np.float32)):
            with precision_scope(prec):
                u0, mask, dx = W.heat_3d_setup(n, dt_np, seed=seed)
                self.dx = dx
                self.dt = 0.1 * dx * dx / nu
                self.kernels[prec] = _built(W.make_heat_3d())
                g = Grid((n, n, n), float)
                g.now[...] = u0
                g.boundary[...] = mask
                self.fields[prec] = g
        self.steps = 0

    def run(self, steps: int, which=("double", "float")) -> None:
        for prec in which:
            k, g = self.kernels[prec], self.fields[prec]
            for _ in range(steps):
                k(g, self.nu, self.dt, self.dx)
        self.steps += steps

    def report(self) -> dict:
        e = norms.diff_norm(self.fields["float"], self.fields["double"])
        return {"n": self.n, "steps": self.steps,
                "error": _rel(e, norms.diff_norm(self.fields["double"]))}
