"""The stencil workloads of the reference, written as xgrid kernels.

* ``diffusion_1d`` - FTCS heat equation with fixed ends (``test.py:254-280``);
* ``convection_1d`` / ``convection_1d_nonlinear`` / ``convection_2d``
  (``test.py:195-251,283-314``, with the single-argument ``boundary(k)`` the
  reference parser accepts, ``xgrid/lang/parser.py:225-227``);
* ``cavity`` - 2-D lid-driven cavity: pressure source, 1 + 50 Jacobi sweeps with
  four boundary classes, momentum, walls and lid (``examples/cavity.py:74-142``);
* ``heat_3d`` - 7-point FTCS heat equation with zero walls (configured 3-D case);
* ``diffusion_2d`` - 5-point FTCS with fixed walls (batched case).

A kernel's text fixes its arithmetic (association and constants), so the
workloads the reference defines keep the reference's expressions.  Each maker
returns a fresh Operator: the meaning of ``float`` is fixed when a kernel is
first parsed, under the precision ``xgrid.init`` set at that time.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .lang import boundary
from .operator import kernel
from .types import grid

float1d = grid[float, 1]
float2d = grid[float, 2]
float3d = grid[float, 3]


@dataclass
class CavityConfig:
    rho: float
    nu: float
    dt: float
    dx: float
    dy: float


def make_diffusion_1d():
    @kernel()
    def diffusion_1d(u: float1d, nu: float, dt: float, dx: float) -> None:
        u[0] = u[0] + nu * dt / dx ** 2.0 * (u[1] - 2.0 * u[0] + u[-1])
        with boundary(1):
            u[0] = 1.0
    return diffusion_1d


def make_convection_1d():
    @kernel()
    def convection_1d(u: float1d, c: float, dt: float, dx: float) -> None:
        u[0] = u[0] - c * dt / dx * (u[0] - u[-1])
        with boundary(1):
            u[0] = 1.0
    return convection_1d


def make_convection_1d_nonlinear():
    @kernel()
    def convection_1d_nonlinear(u: float1d, dt: float, dx: float) -> None:
        u[0] = u[0] - u[0] * dt / dx * (u[0] - u[-1])
        with boundary(1):
            u[0] = 1.0
    return convection_1d_nonlinear


def make_convection_2d():
    @kernel()
    def convection_2d(u: float2d, c: float, dt: float, dx: float, dy: float) -> None:
        cdx = c * dt / dx
        cdy = c * dt / dy
        u[0, 0] = u[0, 0] + cdx * (u[0, 0] - u[-1, 0]) - cdy * (u[0, 0] - u[0, -1])
        with boundary(1):
            u[0, 0] = 1.0
    return convection_2d


def make_cavity(sweeps: int = 50):
    """Lid-driven cavity step; ``sweeps`` Jacobi iterations after the first one."""
    if sweeps != 50:
        return _make_cavity_n(sweeps)

    @kernel()
    def cavity_kernel(b: float2d, p: float2d, u: float2d, v: float2d, cfg: CavityConfig) -> None:
        b[0, 0] = (cfg.rho * (1.0 / cfg.dt *
                              ((u[0, 1] - u[0, -1]) /
                               (2.0 * cfg.dx) + (v[1, 0] - v[-1, 0]) / (2.0 * cfg.dy)) -
                              ((u[0, 1] - u[0, -1]) / (2.0 * cfg.dx))**2.0 -
                              2.0 * ((u[1, 0] - u[-1, 0]) / (2.0 * cfg.dy) *
                                     (v[0, 1] - v[0, -1]) / (2.0 * cfg.dx)) -
                              ((v[1, 0] - v[-1, 0]) / (2.0 * cfg.dy))**2.0))

        p[0, 0] = (((p[0, 1] + p[0, -1]) * cfg.dy**2.0 +
                    (p[1, 0] + p[-1, 0]) * cfg.dx**2.0) /
                   (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) -
                   cfg.dx**2.0 * cfg.dy**2.0 / (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) *
                   b[0, 0][0])

        with boundary(1):
            p[0, 0] = p[0, -1][0]
        with boundary(2):
            p[0, 0] = p[1, 0][0]
        with boundary(3):
            p[0, 0] = p[0, 1][0]
        with boundary(4):
            p[0, 0] = 0.0

        for _ in range(0, 50):
            p[0, 0] = (((p[0, 1][0] + p[0, -1][0]) * cfg.dy**2.0 +
                        (p[1, 0][0] + p[-1, 0][0]) * cfg.dx**2.0) /
                       (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) -
                       cfg.dx**2.0 * cfg.dy**2.0 / (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) *
                       b[0, 0][0])

            with boundary(1):
                p[0, 0] = p[0, -1][0]
            with boundary(2):
                p[0, 0] = p[1, 0][0]
            with boundary(3):
                p[0, 0] = p[0, 1][0]
            with boundary(4):
                p[0, 0] = 0.0

        u[0, 0] = (u[0, 0] -
                   u[0, 0] * cfg.dt / cfg.dx *
                   (u[0, 0] - u[0, -1]) -
                   v[0, 0] * cfg.dt / cfg.dy *
                   (u[0, 0] - u[-1, 0]) -
                   cfg.dt / (2.0 * cfg.rho * cfg.dx) * (p[0, 1][0] - p[0, -1][0]) +
                   cfg.nu * (cfg.dt / cfg.dx**2.0 *
                             (u[0, 1] - 2.0 * u[0, 0] + u[0, -1]) +
                             cfg.dt / cfg.dy**2.0 *
                             (u[1, 0] - 2.0 * u[0, 0] + u[-1, 0])))

        v[0, 0] = (v[0, 0] -
                   u[0, 0] * cfg.dt / cfg.dx *
                   (v[0, 0] - v[0, -1]) -
                   v[0, 0] * cfg.dt / cfg.dy *
                   (v[0, 0] - v[-1, 0]) -
                   cfg.dt / (2.0 * cfg.rho * cfg.dy) * (p[1, 0][0] - p[-1, 0][0]) +
                   cfg.nu * (cfg.dt / cfg.dx**2.0 *
                             (v[0, 1] - 2.0 * v[0, 0] + v[0, -1]) +
                             cfg.dt / cfg.dy**2.0 *
                             (v[1, 0] - 2.0 * v[0, 0] + v[-1, 0])))

        with boundary(1):
            u[0, 0] = 0.0
            v[0, 0] = 0.0

        with boundary(2):
            u[0, 0] = 1.0

    return cavity_kernel


def _make_cavity_n(sweeps: int):
    """Cavity step with a configurable sweep count (the count is a kernel argument)."""
    @kernel()
    def cavity_kernel_n(b: float2d, p: float2d, u: float2d, v: float2d, cfg: CavityConfig, nit: int) -> None:
        b[0, 0] = (cfg.rho * (1.0 / cfg.dt *
                              ((u[0, 1] - u[0, -1]) /
                               (2.0 * cfg.dx) + (v[1, 0] - v[-1, 0]) / (2.0 * cfg.dy)) -
                              ((u[0, 1] - u[0, -1]) / (2.0 * cfg.dx))**2.0 -
                              2.0 * ((u[1, 0] - u[-1, 0]) / (2.0 * cfg.dy) *
                                     (v[0, 1] - v[0, -1]) / (2.0 * cfg.dx)) -
                              ((v[1, 0] - v[-1, 0]) / (2.0 * cfg.dy))**2.0))

        p[0, 0] = (((p[0, 1] + p[0, -1]) * cfg.dy**2.0 +
                    (p[1, 0] + p[-1, 0]) * cfg.dx**2.0) /
                   (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) -
                   cfg.dx**2.0 * cfg.dy**2.0 / (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) *
                   b[0, 0][0])

        with boundary(1):
            p[0, 0] = p[0, -1][0]
        with boundary(2):
            p[0, 0] = p[1, 0][0]
        with boundary(3):
            p[0, 0] = p[0, 1][0]
        with boundary(4):
            p[0, 0] = 0.0

        for _ in range(0, nit):
            p[0, 0] = (((p[0, 1][0] + p[0, -1][0]) * cfg.dy**2.0 +
                        (p[1, 0][0] + p[-1, 0][0]) * cfg.dx**2.0) /
                       (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) -
                       cfg.dx**2.0 * cfg.dy**2.0 / (2.0 * (cfg.dx**2.0 + cfg.dy**2.0)) *
                       b[0, 0][0])

            with boundary(1):
                p[0, 0] = p[0, -1][0]
            with boundary(2):
                p[0, 0] = p[1, 0][0]
            with boundary(3):
                p[0, 0] = p[0, 1][0]
            with boundary(4):
                p[0, 0] = 0.0

        u[0, 0] = (u[0, 0] -
                   u[0, 0] * cfg.dt / cfg.dx *
                   (u[0, 0] - u[0, -1]) -
                   v[0, 0] * cfg.dt / cfg.dy *
                   (u[0, 0] - u[-1, 0]) -
                   cfg.dt / (2.0 * cfg.rho * cfg.dx) * (p[0, 1][0] - p[0, -1][0]) +
                   cfg.nu * (cfg.dt / cfg.dx**2.0 *
                             (u[0, 1] - 2.0 * u[0, 0] + u[0, -1]) +
                             cfg.dt / cfg.dy**2.0 *
                             (u[1, 0] - 2.0 * u[0, 0] + u[-1, 0])))

        v[0, 0] = (v[0, 0] -
                   u[0, 0] * cfg.dt / cfg.dx *
                   (v[0, 0] - v[0, -1]) -
                   v[0, 0] * cfg.dt / cfg.dy *
                   (v[0, 0] - v[-1, 0]) -
                   cfg.dt / (2.0 * cfg.rho * cfg.dy) * (p[1, 0][0] - p[-1, 0][0]) +
                   cfg.nu * (cfg.dt / cfg.dx**2.0 *
                             (v[0, 1] - 2.0 * v[0, 0] + v[0, -1]) +
                             cfg.dt / cfg.dy**2.0 *
                             (v[1, 0] - 2.0 * v[0, 0] + v[-1, 0])))

        with boundary(1):
            u[0, 0] = 0.0
            v[0, 0] = 0.0

        with boundary(2):
            u[0, 0] = 1.0

    return cavity_kernel_n


def make_heat_3d():
    @kernel()
    def heat_3d(u: float3d, nu: float, dt: float, dx: float) -> None:
        u[0, 0, 0] = u[0, 0, 0] + nu * dt / dx ** 2.0 * (u[1, 0, 0] + u[-1, 0, 0] + u[0, 1, 0] + u[0, -1, 0] +
                                                        u[0, 0, 1] + u[0, 0, -1] - 6.0 * u[0, 0, 0])
        with boundary(1):
            u[0, 0, 0] = 0.0
    return heat_3d


def make_diffusion_2d():
    @kernel()
    def diffusion_2d(u: float2d, nu: float, dt: float, dx: float) -> None:
        u[0, 0] = u[0, 0] + nu * dt / dx ** 2.0 * (u[1, 0] + u[-1, 0] + u[0, 1] + u[0, -1] - 4.0 * u[0, 0])
        with boundary(1):
            u[0, 0] = 1.0
    return diffusion_2d


# --------------------------------------------------------------------------- setups
def cavity_masks(n: int) -> dict:
    """Boundary classes of the cavity grids (``examples/cavity.py:56-68``)."""
    u = np.zeros((n, n), np.int32)
    u[0, :] = 1
    u[:, 0] = 1
    u[:, -1] = 1
    u[-1, :] = 2
    v = np.ones((n, n), np.int32)
    v[1:-1, 1:-1] = 0
    p = np.zeros((n, n), np.int32)
    p[:, -1] = 1
    p[0, :] = 2
    p[:, 0] = 3
    p[-1, :] = 4
    b = np.ones((n, n), np.int32)
    b[1:-1, 1:-1] = 0
    return {"b": b, "p": p, "u": u, "v": v}


def cavity_config(n: int, dt: float = 1e-5, nu: float = 0.1, rho: float = 1.0) -> CavityConfig:
    return CavityConfig(rho, nu, dt, 2 / (n - 1), 2 / (n - 1))


def diffusion_1d_setup(nx: int):
    """Hat profile and fixed ends (``test.py:264-267``): u=2 on [0.5, 1], else 1."""
    dx = 2 / (nx - 1)
    u = np.ones(nx, np.float64)
    u[int(.5 / dx):int(1 / dx + 1)] = 2
    mask = np.zeros(nx, np.int32)
    mask[0] = mask[-1] = 1
    return u, mask, dx


def heat_3d_setup(n: int, dtype=np.float32, seed: int = 0):
    """Seeded uniform [0, 1) interior, zero walls with class 1 on the six faces."""
    rng = np.random.default_rng(seed)
    u = np.zeros((n, n, n), dtype)
    u[1:-1, 1:-1, 1:-1] = rng.random((n - 2,) * 3, dtype=np.float32).astype(dtype)
    mask = np.ones((n, n, n), np.int32)
    mask[1:-1, 1:-1, 1:-1] = 0
    dx = 1 / (n - 1)
    return u, mask, dx


def box_walls(shape) -> np.ndarray:
    m = np.ones(shape, np.int32)
    m[tuple(slice(1, -1) for _ in shape)] = 0
    return m


# --------------------------------------------------------------------------- many instances
@dataclass
class InstanceSpec:
    """One independent solve of a batched or served workload."""
    kind: str          # "cavity" | "diffusion"
    n: int             # grid side
    nu: float
    dt: float
    steps: int = 0     # kernel calls of the solve (serving)
    arrival: float = 0.0   # seconds after the start of the trace (serving)
    rid: int = 0


def stable_dt(kind: str, n: int, nu: float) -> float:
    """0.2 dx^2 / nu (explicit diffusion limit with margin); the cavity also keeps
    the lid's CFL number u dt / dx at most 0.2."""
    dx = 2.0 / (n - 1)
    dt = 0.2 * dx * dx / nu
    return min(dt, 0.2 * dx) if kind == "cavity" else dt


def batch_specs(count: int = 256, seed: int = 1, lo: int = 64, hi: int = 1024) -> list:
    """Independent cavity / 2-D diffusion instances: side log-uniform in [lo, hi],
    nu uniform in [0.01, 0.1], kind a fair coin, stable dt."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        nu = float(rng.uniform(0.01, 0.1))
        kind = "cavity" if rng.random() < 0.5 else "diffusion"
        out.append(InstanceSpec(kind, n, nu, stable_dt(kind, n, nu), rid=i))
    return out


def serving_specs(count: int = 400, seed: int = 2, rate: float = 200.0, lo: int = 128, hi: int = 2048,
                  steps_lo: int = 100, steps_hi: int = 2000) -> list:
    """A finite Poisson-arrival trace of cavity solves: exponential gaps at `rate`
    per second, side log-uniform in [lo, hi], nu uniform in [0.001, 0.1], steps
    uniform in [steps_lo, steps_hi]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(round(float(np.exp(rng.uniform(np.log(lo), np.log(hi))))))
        nu = float(rng.uniform(0.001, 0.1))
        steps = int(rng.integers(steps_lo, steps_hi + 1))
        out.append(InstanceSpec("cavity", n, nu, stable_dt("cavity", n, nu), steps, t, i))
    return out


def instance_grids(spec: InstanceSpec, Grid):
    """Fresh grids and the scalar arguments of one instance (zero initial fields;
    diffusion starts from a hat: 2 inside the middle half, 1 elsewhere)."""
    n = spec.n
    if spec.kind == "cavity":
        masks = cavity_masks(n)
        grids = []
        for c in "bpuv":
            g = Grid((n, n), float)
            g.boundary[...] = masks[c]
            grids.append(g)
        return grids, (CavityConfig(1.0, spec.nu, spec.dt, 2 / (n - 1), 2 / (n - 1)),)
    g = Grid((n, n), float)
    u = g.now
    u[...] = 1.0
    u[n // 4:3 * n // 4, n // 4:3 * n // 4] = 2.0
    g.boundary[...] = box_walls((n, n))
    return [g], (spec.nu, spec.dt, 2 / (n - 1))
