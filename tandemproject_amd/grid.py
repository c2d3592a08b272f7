"""Algebraic H-grids for the reference's 3x2 block layout, and the raw grid format.

The reference generates its grid with Hermite-patch blending around a
wavy-leading-edge aerofoil (src/Grid.f90, outside the hot path).  This module
writes grids of the same topology with a simpler algebraic construction:

* columns in x: upstream [-xu, 0], chord [0, 1], wake [1, xd];
* rows in y: bottom [-yb, surface], top [surface, yt]; the surface is the
  aerofoil (y = -t(x) / +t(x)) over the chord and the line y = 0 elsewhere;
* z uniform over [-span/2, span/2] (both ends are periodic images).

The thickness t(x) is the NACA 4-digit distribution blended to zero slope at
the leading and trailing edges (``blend``): an H-grid cannot wrap a blunt
nose, so the profile is smoothed where it meets the stagnation and wake
lines.  Grids are returned per sub-block as ss(0:lmx, 1:3) in Fortran order,
the layout of the reference's raw grid file (src/Grid.f90:452-459,
src/rpt.f90:673-696).
"""
import numpy as np


def naca_thickness(x, t=0.12, blend=0.04):
    x = np.clip(x, 0.0, 1.0)
    y = 5.0 * t * (0.2969 * np.sqrt(x) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3 - 0.1036 * x ** 4)
    if blend > 0:
        y = y * (1.0 - np.exp(-(x / blend) ** 2)) * (1.0 - np.exp(-((1.0 - x) / blend) ** 2))
    return y


def _stretch(n, first):
    """n+1 points on [0, 1] with first spacing ~first (clustered at 0)."""
    s = np.linspace(0.0, 1.0, n + 1)
    uni = 1.0 / n
    if first >= uni:
        return s
    lo, hi = 1e-6, 50.0
    for _ in range(200):
        b = 0.5 * (lo + hi)
        d = np.sinh(b / n) / np.sinh(b)
        if d > first:
            lo = b
        else:
            hi = b
    b = 0.5 * (lo + hi)
    return np.sinh(b * s) / np.sinh(b)


def _chord(n, c=0.6):
    s = np.linspace(0.0, 1.0, n + 1)
    return s - c * np.sin(2 * np.pi * s) / (2 * np.pi)


def h_grid_blocks(lxibk, letbk, lzebk, xu=7.0, xd=10.0, yb=7.0, yt=7.0, span=0.2,
                  thick=0.12, blend=0.04, wall_spacing=2e-3):
    """Grid of each parent block (bkx=3, bky=2, bkz=1): list of (x, y, z) arrays
    shaped (nz, ny, nx)."""
    if len(lxibk) != 3 or len(letbk) != 2 or len(lzebk) != 1:
        raise ValueError("the H-grid generator builds the reference's 3x2x1 block layout")
    nc = lxibk[1]
    xc = _chord(nc)
    dx_le = xc[1] - xc[0]
    xup = -xu * _stretch(lxibk[0], dx_le / xu)[::-1]
    xup[-1] = 0.0
    xwk = 1.0 + (xd - 1.0) * _stretch(lxibk[2], dx_le / (xd - 1.0))
    xcols = [xup, xc, xwk]
    etab = _stretch(letbk[0], wall_spacing / yb)   # from surface outwards
    etat = _stretch(letbk[1], wall_spacing / yt)
    nz = lzebk[0] + 1
    zs = -0.5 * span + span * np.arange(nz) / lzebk[0]
    blocks = []
    for j in range(2):
        for i in range(3):
            x = xcols[i]
            ts = naca_thickness(x, thick, blend) if i == 1 else np.zeros_like(x)
            if j == 0:
                ysurf = -ts
                e = etab[::-1]  # eta runs from far field (e=1) to surface (e=0)
                y2 = ysurf[None, :] + (-yb - ysurf[None, :]) * e[:, None]
            else:
                ysurf = ts
                y2 = ysurf[None, :] + (yt - ysurf[None, :]) * etat[:, None]
            ny, nx = y2.shape
            X = np.broadcast_to(x[None, None, :], (nz, ny, nx)).copy()
            Y = np.broadcast_to(y2[None, :, :], (nz, ny, nx)).copy()
            Z = np.broadcast_to(zs[:, None, None], (nz, ny, nx)).copy()
            blocks.append((X, Y, Z))
    return blocks


def subblock_grids(layout, blocks):
    """Slice parent-block grids into the layout's sub-blocks; returns a list of
    flat ss arrays (x, y, z concatenated, Fortran order within each)."""
    out = []
    for r in range(layout.nb):
        X, Y, Z = blocks[layout.parent[r]]
        i0, j0, k0 = layout.start[r]
        nx, ny, nz = layout.dims[r] + 1
        sl = (slice(k0, k0 + nz), slice(j0, j0 + ny), slice(i0, i0 + nx))
        out.append(np.concatenate([X[sl].ravel(), Y[sl].ravel(), Z[sl].ravel()]))
    return out


def perturbed_box(lxi, let, lze, amp=0.05, seed=0, lengths=(1.0, 1.0, 1.0), periodic_z=True):
    """Smoothly perturbed Cartesian grid of one block, flat ss(0:lmx, 1:3).

    With periodic_z the z=const planes at k=0 and k=lze are periodic images
    (z differs by lengths[2], x and y identical)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = lxi + 1, let + 1, lze + 1
    s = np.linspace(0, 1, nx)[None, None, :]
    t = np.linspace(0, 1, ny)[None, :, None]
    u = np.linspace(0, 1, nz)[:, None, None]
    ph = rng.uniform(0, 2 * np.pi, 6)
    uz = 2 * np.pi * u
    x = lengths[0] * (s + amp * np.sin(np.pi * s) * np.sin(2 * np.pi * t + ph[0]) * (1 + 0.3 * np.sin(uz + ph[1])))
    y = lengths[1] * (t + amp * np.sin(np.pi * t) * np.sin(2 * np.pi * s + ph[2]) * (1 + 0.3 * np.cos(uz + ph[3])))
    z = lengths[2] * (u + 0.5 * amp * np.sin(2 * np.pi * s + ph[4]) * np.sin(2 * np.pi * t + ph[5]) *
                      (np.sin(np.pi * u) if not periodic_z else 1.0))
    if periodic_z:
        z = lengths[2] * u + 0.5 * amp * lengths[2] * np.sin(2 * np.pi * s + ph[4]) * np.sin(2 * np.pi * t + ph[5])
    X = np.broadcast_to(x, (nz, ny, nx)); Y = np.broadcast_to(y, (nz, ny, nx)); Z = np.broadcast_to(z, (nz, ny, nx))
    return np.concatenate([X.ravel(), Y.ravel(), Z.ravel()])


def write_raw_grid(path, ss):
    """Raw FP64 block grid as written by src/Grid.f90:452-459 (x, y, z planes)."""
    np.asarray(ss, "<f8").tofile(path)


def read_raw_grid(path, npoints):
    a = np.fromfile(path, "<f8")
    if a.size != 3 * npoints:
        raise ValueError(f"{path}: expected {3 * npoints} values, found {a.size}")
    return a
