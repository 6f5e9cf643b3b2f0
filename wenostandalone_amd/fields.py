"""Host-side input set-up: the harness's synthetic field and coefficients, and
the seeded field recipes used for parity tests and benchmarks.

All arrays are numpy, Fortran order, with the reference's shapes:
  v_vf (nx+6, ny+6, nz+6, v)   weno_standalone.f90:163-165
"""
from __future__ import annotations

import numpy as np

DIR_OFFSETS = (0.03, 0.05, 0.07)   # weno_standalone.f90:190-192


def extents(m: int, n: int, p: int, buff: int):
    return m + 1 + 2 * buff, n + 1 + 2 * buff, p + 1 + 2 * buff


def dir_extents(m, n, p, buff, d):
    """(ns, nk, nl) of direction d: x:(nx,ny,nz) y:(ny,nx,nz) z:(nz,ny,nx)."""
    nx, ny, nz = extents(m, n, p, buff)
    return [(nx, ny, nz), (ny, nx, nz), (nz, ny, nx)][d]


def harness_field(m: int, n: int, p: int, v: int, buff: int) -> np.ndarray:
    """s_fill_input_field (weno_standalone.f90:221-244)."""
    nx, ny, nz = extents(m, n, p, buff)
    ib = -buff
    i = np.arange(ib - 3, ib - 3 + nx + 6)
    j = np.arange(ib - 3, ib - 3 + ny + 6)
    k = np.arange(ib - 3, ib - 3 + nz + 6)
    x = (i - ib + 1).astype(np.float64) / float(max(1, (m + buff) - ib + 3))
    y = (j - ib + 1).astype(np.float64) / float(max(1, (n + buff) - ib + 3))
    z = (k - ib + 1).astype(np.float64) / float(max(1, (p + buff) - ib + 3))
    out = np.empty((nx + 6, ny + 6, nz + 6, v), dtype=np.float64, order="F")
    cy = np.cos(3.7 * y)[None, :, None]
    sz = 0.25 * np.sin(4.3 * z)[None, None, :]
    ii = i[:, None, None]
    jj = j[None, :, None]
    kk = k[None, None, :]
    for l in range(1, v + 1):
        sx = np.sin(2.1 * x * float(l))[:, None, None]
        out[..., l - 1] = sx + cy + sz + 0.02 * (ii + 2 * jj - kk + l).astype(np.float64)
    return out


def harness_coefficients(nc: int, offset: float):
    """s_fill_coefficients (weno_standalone.f90:246-279): polyL, polyR (nc,5,3),
    dL, dR (5,nc), beta (nc,5,6); constant in j."""
    baseL = np.array([0.30, 0.27, 0.20, 0.14, 0.09])
    baseR = np.array([0.09, 0.14, 0.20, 0.27, 0.30])
    r = np.arange(5)[:, None]
    s = np.arange(3)[None, :]
    pl = offset + 0.015 * ((r + 1) * (s + 1)).astype(np.float64)
    pr = -offset + 0.012 * (2 * r + s + 1).astype(np.float64)
    polyL = np.asfortranarray(np.broadcast_to(pl[None], (nc, 5, 3)))
    polyR = np.asfortranarray(np.broadcast_to(pr[None], (nc, 5, 3)))
    dL = np.asfortranarray(np.broadcast_to((baseL + offset * 0.10)[:, None], (5, nc)))
    dR = np.asfortranarray(np.broadcast_to((baseR + offset * 0.10)[:, None], (5, nc)))
    b = np.zeros((5, 6))
    rr = np.arange(5).astype(np.float64)
    b[:, 0] = 0.8 + offset + 0.05 * (rr + 1)
    b[:, 2] = 0.9 + offset + 0.04 * (rr + 1)
    b[:, 3] = 1.0 + offset + 0.03 * (rr + 1)
    b[:, 5] = 1.1 + offset + 0.02 * (rr + 1)
    beta = np.asfortranarray(np.broadcast_to(b[None], (nc, 5, 6)))
    return polyL, polyR, dL, dR, beta


def random_coefficients(nc: int, seed: int):
    """j-varying coefficients (what a stretched grid produces) for tests that
    must not depend on the harness's constant-in-j values."""
    g = np.random.default_rng([seed, 77])
    polyL = np.asfortranarray(g.uniform(-0.6, 0.6, (nc, 5, 3)))
    polyR = np.asfortranarray(g.uniform(-0.6, 0.6, (nc, 5, 3)))
    dL = np.asfortranarray(g.uniform(0.05, 0.6, (5, nc)))
    dR = np.asfortranarray(g.uniform(0.05, 0.6, (5, nc)))
    beta = np.asfortranarray(g.uniform(0.2, 2.0, (nc, 5, 6)))
    return polyL, polyR, dL, dR, beta


def psd_coefficients(nc: int, order: int, seed: int):
    """j-varying coefficients whose smoothness indicators are positive
    semi-definite quadratic forms in the stencil's differences, as every real
    WENO scheme's are: beta_coef(j, r, :) holds the upper triangle of
    M = A A^T (A random) in the reference's term order -- WENO3 (m00),
    WENO5 (m00, 2 m01, m11), WENO7 (m00, 2 m01, 2 m02, m11, 2 m12, m22)
    (weno_standalone.f90:427-428, :545-550, :1462-1473).  Branches whose
    weights are linear in 1/beta (WENO-Z) need beta >= eps to be well
    conditioned; random_coefficients' independent cross terms do not
    guarantee it."""
    polyL, polyR, dL, dR, _ = random_coefficients(nc, seed)
    g = np.random.default_rng([seed, 79, order])
    n = {3: 1, 5: 2, 7: 3}[order]
    beta = np.zeros((nc, 5, 6), order="F")
    for r in range(5):
        A = g.uniform(-1.0, 1.0, (nc, n, n))
        M = np.einsum("jab,jcb->jac", A, A)
        col = 0
        for a in range(n):
            for c in range(a, n):
                beta[:, r, col] = M[:, a, c] * (1.0 if a == c else 2.0)
                col += 1
    return polyL, polyR, dL, dR, np.asfortranarray(beta)


def seeded_field(m: int, n: int, p: int, v: int, buff: int, seed: int, recipe: int = 0) -> np.ndarray:
    """Seeded synthetic field.

    recipe 0: per variable 1 + 0.3*sum_{3 modes} sin(2*pi*(k.x) + phi), k in
      U{1..4}^3, phi in U[0,2pi); plus a planar step along x of height
      U(0.5,2) at x_s in U(0.2,0.8); plus U(-1e-3,1e-3) noise per cell.
      Coordinates span the full v_vf extent, [0,1] per axis.
    recipe 2: recipe 0, then values at y >= y_s (y_s in U(0.3,0.7)) scaled by
      1e3 (jump ratio 1e3), then two boxes per variable set exactly constant
      (weights dominated by weno_eps there).
    Deterministic in (seed, variable)."""
    nx, ny, nz = extents(m, n, p, buff)
    NX, NY, NZ = nx + 6, ny + 6, nz + 6
    x = (np.arange(NX) / max(1, NX - 1))[:, None, None]
    y = (np.arange(NY) / max(1, NY - 1))[None, :, None]
    z = (np.arange(NZ) / max(1, NZ - 1))[None, None, :]
    out = np.empty((NX, NY, NZ, v), dtype=np.float64, order="F")
    for i in range(v):
        g = np.random.default_rng([seed, i, recipe])
        f = np.ones((NX, NY, NZ))
        for _ in range(3):
            kx, ky, kz = g.integers(1, 5, size=3)
            phi = g.uniform(0.0, 2.0 * np.pi)
            f += 0.3 * np.sin(2.0 * np.pi * (kx * x + ky * y + kz * z) + phi)
        h = g.uniform(0.5, 2.0)
        xs = g.uniform(0.2, 0.8)
        f += h * (x >= xs)
        f += g.uniform(-1e-3, 1e-3, size=(NX, NY, NZ))
        if recipe == 2:
            ys = g.uniform(0.3, 0.7)
            f = np.where(y >= ys, f * 1e3, f)
            for _ in range(2):
                lo = [int(g.uniform(0.0, 0.7) * N) for N in (NX, NY, NZ)]
                sz = [max(4, int(g.uniform(0.15, 0.3) * N)) for N in (NX, NY, NZ)]
                c = g.uniform(0.5, 2.0)
                f[lo[0]:lo[0] + sz[0], lo[1]:lo[1] + sz[1], lo[2]:lo[2] + sz[2]] = c
        out[..., i] = f
    return out
