"""numpy restatements of the workloads (ORACLE - test infrastructure only).

Written from the workload definitions, independently of the xgrid frontend,
so they pin ``cpu_path`` (and, through it, the device backend).  Each step
uses the kernel's own association of every operation and IEEE double (or
float) arithmetic without FMA, which is what gcc emits for the reference's C
on x86-64 without ``-march``; ``x ** 2.0`` is ``x * x`` (gcc folds
``pow(x, 2.0)``).  Time levels follow the reference: loads default to the
previous level, ``[0]`` reads the level being written
(``xgrid/lang/parser.py:512-525``).
"""
from __future__ import annotations

import numpy as np


def _s(a, o0, o1=None, o2=None):
    """Interior view of `a` shifted by the stencil offset (interior = 1..n-2)."""
    if o1 is None:
        return a[1 + o0: a.shape[0] - 1 + o0]
    if o2 is None:
        return a[1 + o0: a.shape[0] - 1 + o0, 1 + o1: a.shape[1] - 1 + o1]
    return a[1 + o0: a.shape[0] - 1 + o0, 1 + o1: a.shape[1] - 1 + o1, 1 + o2: a.shape[2] - 1 + o2]


# --------------------------------------------------------------- 1-D (test.py:254-280)
def diffusion_1d(u0: np.ndarray, nu: float, dt: float, dx: float, steps: int) -> np.ndarray:
    """Ends are class 1 (fixed at 1.0), the rest class 0."""
    u = u0.astype(np.float64).copy()
    coef = (nu * dt) / (dx * dx)
    for _ in range(steps):
        un = u
        u = np.empty_like(un)
        u[1:-1] = un[1:-1] + coef * ((un[2:] - 2.0 * un[1:-1]) + un[:-2])
        u[0] = 1.0
        u[-1] = 1.0
    return u


def convection_1d(u0, c, dt, dx, steps, nonlinear=False):
    """u[0] = u[0] - c*dt/dx*(u[0] - u[-1]); class 1 (index 0) fixed at 1.0."""
    u = u0.astype(np.float64).copy()
    for _ in range(steps):
        un = u
        u = np.empty_like(un)
        if nonlinear:
            u[1:] = un[1:] - ((un[1:] * dt) / dx) * (un[1:] - un[:-1])
        else:
            u[1:] = un[1:] - ((c * dt) / dx) * (un[1:] - un[:-1])
        u[0] = 1.0
    return u


# --------------------------------------------------------------- 2-D
def convection_2d(u0, c, dt, dx, dy, steps):
    """cdx, cdy on the host; class 1 = row 0 and column 0 (fixed at 1.0)."""
    u = u0.astype(np.float64).copy()
    cdx = (c * dt) / dx
    cdy = (c * dt) / dy
    for _ in range(steps):
        un = u
        u = np.empty_like(un)
        u[1:, 1:] = (un[1:, 1:] + cdx * (un[1:, 1:] - un[:-1, 1:])) - cdy * (un[1:, 1:] - un[1:, :-1])
        u[0, :] = 1.0
        u[:, 0] = 1.0
    return u


def diffusion_2d(u0, nu, dt, dx, steps, dtype=np.float64):
    """Class 1 = the frame, fixed at 1.0."""
    u = u0.astype(dtype).copy()
    coef = (nu * dt) / (dx * dx)
    for _ in range(steps):
        un = u
        u = np.empty_like(un)
        c = _s(un, 0, 0)
        u[1:-1, 1:-1] = c + coef * ((((_s(un, 1, 0) + _s(un, -1, 0)) + _s(un, 0, 1)) + _s(un, 0, -1)) - 4.0 * c)
        u[0, :] = u[-1, :] = 1.0
        u[:, 0] = u[:, -1] = 1.0
    return u


def cavity(n: int, steps: int, rho: float, nu: float, dt: float, dx: float, dy: float, sweeps: int = 50):
    """Lid-driven cavity as the xgrid kernel runs it (examples/cavity.py:74-142).

    One step: b from (u, v) of the previous step; p from the previous p and
    the new b, then ``sweeps`` Jacobi sweeps, each followed by the four ordered
    boundary statements; u, v from the previous u, v and the new p; walls;
    lid.  b's frame is never written (stays 0).
    """
    b = np.zeros((n, n))
    p = np.zeros((n, n))
    u = np.zeros((n, n))
    v = np.zeros((n, n))
    dx2, dy2 = dx * dx, dy * dy
    den = 2.0 * (dx2 + dy2)
    kb = (dx2 * dy2) / den

    def p_bc(q):
        q[1:-1, -1] = q[1:-1, -2]      # class 1: column n-1, rows 1..n-2
        q[0, 1:] = q[1, 1:]            # class 2: row 0, columns 1..n-1
        q[:-1, 0] = q[:-1, 1]          # class 3: column 0, rows 0..n-2
        q[-1, :] = 0.0                 # class 4: row n-1

    for _ in range(steps):
        un, vn, pn = u, v, p
        # b
        X = (_s(un, 0, 1) - _s(un, 0, -1)) / (2.0 * dx)
        Y = (_s(vn, 1, 0) - _s(vn, -1, 0)) / (2.0 * dy)
        Z = (((_s(un, 1, 0) - _s(un, -1, 0)) / (2.0 * dy)) * (_s(vn, 0, 1) - _s(vn, 0, -1))) / (2.0 * dx)
        b = np.zeros((n, n))
        b[1:-1, 1:-1] = rho * (((((1.0 / dt) * (X + Y)) - X * X) - 2.0 * Z) - Y * Y)
        # p: first statement reads the previous p
        p = np.empty((n, n))
        p[1:-1, 1:-1] = (((_s(pn, 0, 1) + _s(pn, 0, -1)) * dy2 + (_s(pn, 1, 0) + _s(pn, -1, 0)) * dx2) / den
                         - kb * b[1:-1, 1:-1])
        p_bc(p)
        for _ in range(sweeps):
            q = p.copy()
            p[1:-1, 1:-1] = (((_s(q, 0, 1) + _s(q, 0, -1)) * dy2 + (_s(q, 1, 0) + _s(q, -1, 0)) * dx2) / den
                             - kb * b[1:-1, 1:-1])
            p_bc(p)
        # u, v from the previous step's u, v and this step's p
        U, V = _s(un, 0, 0), _s(vn, 0, 0)
        u = np.empty((n, n))
        v = np.empty((n, n))
        u[1:-1, 1:-1] = ((((U - ((U * dt) / dx) * (U - _s(un, 0, -1)))
                           - ((V * dt) / dy) * (U - _s(un, -1, 0)))
                          - (dt / ((2.0 * rho) * dx)) * (_s(p, 0, 1) - _s(p, 0, -1)))
                         + nu * ((dt / dx2) * ((_s(un, 0, 1) - 2.0 * U) + _s(un, 0, -1))
                                 + (dt / dy2) * ((_s(un, 1, 0) - 2.0 * U) + _s(un, -1, 0))))
        v[1:-1, 1:-1] = ((((V - ((U * dt) / dx) * (V - _s(vn, 0, -1)))
                           - ((V * dt) / dy) * (V - _s(vn, -1, 0)))
                          - (dt / ((2.0 * rho) * dy)) * (_s(p, 1, 0) - _s(p, -1, 0)))
                         + nu * ((dt / dx2) * ((_s(vn, 0, 1) - 2.0 * V) + _s(vn, 0, -1))
                                 + (dt / dy2) * ((_s(vn, 1, 0) - 2.0 * V) + _s(vn, -1, 0))))
        # walls (u class 1, v class 1) then lid (u class 2 = last row)
        u[0, :] = 0.0
        u[:, 0] = 0.0
        u[:-1, -1] = 0.0
        v[0, :] = v[-1, :] = 0.0
        v[:, 0] = v[:, -1] = 0.0
        u[-1, :] = 1.0
    return b, p, u, v


# --------------------------------------------------------------- 3-D heat
def heat_3d(u0, nu, dt, dx, steps, dtype=np.float64):
    """Zero walls (class 1).  With float storage the C arithmetic is mixed:
    ``nu*dt/dx**2`` is float, ``6.0*u`` and everything after it double, and the
    result is rounded to float when stored."""
    u = u0.astype(dtype).copy()
    if dtype == np.float64:
        coef = (nu * dt) / (dx * dx)
        for _ in range(steps):
            un = u
            u = np.zeros_like(un)
            c = _s(un, 0, 0, 0)
            s = ((((_s(un, 1, 0, 0) + _s(un, -1, 0, 0)) + _s(un, 0, 1, 0)) + _s(un, 0, -1, 0))
                 + _s(un, 0, 0, 1)) + _s(un, 0, 0, -1)
            u[1:-1, 1:-1, 1:-1] = c + coef * (s - 6.0 * c)
        return u
    f = np.float32
    coef = f(f(f(nu) * f(dt)) / f(f(dx) * f(dx)))
    for _ in range(steps):
        un = u
        u = np.zeros_like(un)
        c = _s(un, 0, 0, 0)
        s = ((((_s(un, 1, 0, 0) + _s(un, -1, 0, 0)) + _s(un, 0, 1, 0)) + _s(un, 0, -1, 0))
             + _s(un, 0, 0, 1)) + _s(un, 0, 0, -1)                                # float32 sums
        d = s.astype(np.float64) - 6.0 * c.astype(np.float64)                    # double
        u[1:-1, 1:-1, 1:-1] = (c.astype(np.float64) + np.float64(coef) * d).astype(np.float32)
    return u
