"""ORACLE -- test infrastructure only.

The Stokes dual-splitting step of examples/stokes_2d/stokes_2d.h:358-386
(advance_timestep) restated on the CPU, for the example's analytic problem.
It runs on the oracle's operators (oracle/dg_oracle.cpp through a Problem
object of tests/oracle_lib.py: grad, div, laplace_bc, laplace_inhomogeneous)
and assembles each matrix by unit vectors as the reference does
(stokes_2d.h:388-422), then solves densely. Only tests/ may use it, as the
checker of the HIP path.

Two departures, shared with the HIP path:
* the boundary data enter through the whole affine part of the BR1 operator
  (oracle_laplace_apply_data on a zero field). The reference's
  laplace_2d::apply_inhomogeneous_as_rhs lifts the data on the boundary faces
  only and leaves out the interior faces of boundary cells, which makes the
  Laplacian with inhomogeneous Dirichlet data inconsistent (its error grows as
  the mesh is refined); for Neumann data the two agree;
* the reference factorises the singular pure-Neumann pressure matrix directly
  (cusolverSpDcsrlsvluHost, stokes_2d.h:452-526; not available here); this
  restatement solves -M Lap_h p = -M w with the right-hand side projected onto
  the complement of the constants, taking the solution orthogonal to them.
Layout: cell-major u[c * Np + j].
"""
from __future__ import annotations

import numpy as np

A = 2.883356
LAMBDA = 1 + A * A


def exact(x, y, t):
    # stokes_2d.h:297-298
    ux = np.sin(x) * (A * np.sin(A * y) - np.cos(A) * np.sinh(y)) * np.exp(-LAMBDA * t)
    uy = np.cos(x) * (np.cos(A * y) + np.cos(A) * np.cosh(y)) * np.exp(-LAMBDA * t)
    return ux, uy


class Geometry:
    """Per-cell geometry from the cell corners (src/mapping_2d.h:101-158,
    src/basic_geom_2d.h:110-153) and the boundary nodes in the order the
    reference calls its BC objects (cell by cell, face by face)."""

    def __init__(self, cx, cy, nb_cell, r, s, fmask):
        K = len(cx) // 3
        x = np.asarray(cx).reshape(K, 3)
        y = np.asarray(cy).reshape(K, 3)
        xr, xs = 0.5 * (x[:, 1] - x[:, 0]), 0.5 * (x[:, 2] - x[:, 0])
        yr, ys = 0.5 * (y[:, 1] - y[:, 0]), 0.5 * (y[:, 2] - y[:, 0])
        det = xr * ys - xs * yr
        self.rx, self.ry, self.sx, self.sy = ys / det, -xs / det, -yr / det, xr / det
        self.J = 0.25 * ((x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0]))
        self.K, self.x, self.y = K, x, y
        self.fmask = np.asarray(fmask)
        nb = np.asarray(nb_cell).reshape(K, 3)
        self.bnd = nb == np.arange(K)[:, None]
        bx, by, bnx, bny, bcell = [], [], [], [], []
        self.nx = np.zeros((K, 3))
        self.ny = np.zeros((K, 3))
        for c in range(K):
            for e in range(3):
                a, b, o = e, (e + 1) % 3, (e + 2) % 3
                ex, ey = x[c, b] - x[c, a], y[c, b] - y[c, a]
                if ey * (x[c, o] - x[c, a]) - ex * (y[c, o] - y[c, a]) > 0.0:
                    ex, ey = -ex, -ey
                l = np.hypot(ex, ey)
                self.nx[c, e], self.ny[c, e] = ey / l, -ex / l
                if self.bnd[c, e]:
                    for d in range(len(self.fmask[e])):
                        n = self.fmask[e][d]
                        l1, l2, l3 = 0.5 * (s[n] + 1.0), -0.5 * (r[n] + s[n]), 0.5 * (r[n] + 1.0)
                        bx.append(l2 * x[c, 0] + l3 * x[c, 1] + l1 * x[c, 2])
                        by.append(l2 * y[c, 0] + l3 * y[c, 1] + l1 * y[c, 2])
        self.bx, self.by = np.array(bx), np.array(by)


def _assemble(apply, n):
    """Columns A e_i by unit vectors (stokes_2d.h:388-422)."""
    A = np.zeros((n, n))
    e = np.zeros(n)
    for i in range(n):
        e[i] = 1.0
        A[:, i] = apply(e)
        e[i] = 0.0
    return A


def _mass(geo, Mref):
    K, Np = geo.K, Mref.shape[0]
    M = np.zeros((K * Np, K * Np))
    for c in range(K):
        M[c * Np:(c + 1) * Np, c * Np:(c + 1) * Np] = geo.J[c] * Mref
    return M


def _pressure_data(geo, Dr, Ds, ux, uy, uxp, uyp, t_next):
    """2 h_p per boundary node (stokes_2d.h:186-260)."""
    K, Np = geo.K, Dr.shape[0]
    out = []
    for c in range(K):
        if not geo.bnd[c].any():
            continue
        Dx = Dr * geo.rx[c] + Ds * geo.sx[c]
        Dy = Dr * geo.ry[c] + Ds * geo.sy[c]
        sl = slice(c * Np, (c + 1) * Np)
        curl_n = Dx @ uy[sl] - Dy @ ux[sl]
        ccx = Dy @ curl_n
        ccy = -(Dx @ curl_n)
        curl_p = Dx @ uyp[sl] - Dy @ uxp[sl]
        ccx = Dy @ curl_p - 2 * ccx
        ccy = -(Dx @ curl_p) - 2 * ccy
        for e in range(3):
            if not geo.bnd[c, e]:
                continue
            nx, ny = geo.nx[c, e], geo.ny[c, e]
            for node in geo.fmask[e]:
                out.append((ccx[node] * nx + ccy[node] * ny, nx, ny))
    # the boundary node coordinates are in the same (cell, face, node) order
    hq = np.zeros(len(out))
    for b, (curl_term, nx, ny) in enumerate(out):
        x, y = geo.bx[b], geo.by[b]
        gx, gy = exact(x, y, t_next)
        hq[b] = 2 * (LAMBDA * gx * nx + LAMBDA * gy * ny + curl_term)
    return hq


def advance_timestep(Pw, Ps, geo, Mref, u_prev, t_prev, u, t, dt):
    """u_prev, u: (ux, uy) cell-major at t_prev, t; returns u at t + dt.
    Pw / Ps: oracle Problems with the weak / strong operators."""
    n = geo.K * Mref.shape[0]
    ut = [(u[i] * 4.0 - u_prev[i]) / 3.0 for i in range(2)]
    gtx, gty = exact(geo.bx, geo.by, t)
    gpx, gpy = exact(geo.bx, geo.by, t_prev)
    gt = (2.0 * (4.0 * gtx - gpx) / 3.0, 2.0 * (4.0 * gty - gpy) / 3.0)
    gnx, gny = exact(geo.bx, geo.by, t + dt)
    gu = (2.0 * gnx, 2.0 * gny)
    hq = _pressure_data(geo, Ps.Dr, Ps.Ds, u[0], u[1], u_prev[0], u_prev[1], t + dt)
    M = _mass(geo, Mref)
    # pressure (stokes_2d.h:322-338)
    zero = np.zeros(n)
    w = Pw.div(ut[0], ut[1], gt[0], gt[1], -1.0) * (1.5 / dt)
    w = w - Ps.laplace_data(zero, 1, None, hq)  # Lap_h p + affine part = (3 / 2dt) div u~
    S = -M @ _assemble(lambda v: Ps.laplace_bc(v, 1), n)
    b = -M @ w
    b -= b.mean()
    p = np.linalg.solve(S + np.ones((n, n)) / n, b)
    # projection (stokes_2d.h:340-349)
    gx, gy = Pw.grad(p, None, 1.0)
    ut = [ut[0] - 2.0 * dt * gx / 3.0, ut[1] - 2.0 * dt * gy / 3.0]
    # viscous (stokes_2d.h:351-368)
    sigma = 1.5 / dt
    H = -M @ _assemble(lambda v: Ps.laplace_bc(v, 0), n) + sigma * M
    out = []
    for i in range(2):
        w = -sigma * ut[i] - Ps.laplace_data(zero, 0, gu[i], None)
        out.append(np.linalg.solve(H, -M @ w))
    return out[0], out[1]
