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


def tandem_grid_blocks(lxibk, letbk, lzebk, spacing=3.0, xu=7.0, xd=10.0, ym=1.0, yb=7.0, yt=7.0, span=0.2,
                       thick=0.12, blend=0.04, wall_spacing=2e-3):
    """Grid of each parent block of two aerofoils in tandem (bkx = 6, bky = 4,
    bkz = 1): columns upstream [-xu, 0], chord 1 [0, 1], the gap in two
    columns [1, spacing] (clustered at the trailing and the leading edge),
    chord 2 [spacing, spacing + 1], wake [spacing + 1, spacing + 1 + xd];
    rows [-yb, -ym], [-ym, surface], [surface, ym], [ym, yt], the surface the
    aerofoil over the chord columns and y = 0 elsewhere, clustered at the
    surface (first spacing wall_spacing) and at y = +-ym from the far rows.
    Returns (x, y, z) arrays shaped (nz, ny, nx) per parent block, row-major
    over the columns as the layout numbers them."""
    if len(lxibk) != 6 or len(letbk) != 4 or len(lzebk) != 1:
        raise ValueError("the tandem grid has 6 x 4 x 1 blocks")
    x1, x2 = 0.0, spacing
    c1 = x1 + _chord(lxibk[1])
    c2 = x2 + _chord(lxibk[4])
    d1 = c1[1] - c1[0]
    d2 = c2[1] - c2[0]
    mid = 0.5 * (1.0 + x2)
    xup = -xu * _stretch(lxibk[0], d1 / xu)[::-1]
    xup[-1] = 0.0
    ga = 1.0 + (mid - 1.0) * _stretch(lxibk[2], d1 / (mid - 1.0))
    gb = x2 - (x2 - mid) * _stretch(lxibk[3], d2 / (x2 - mid))[::-1]
    ga[-1] = gb[0] = mid
    wk = x2 + 1.0 + xd * _stretch(lxibk[5], d2 / xd)
    xcols = [xup, c1, ga, gb, c2, wk]
    chord = [None, x1, None, None, x2, None]
    en = [_stretch(letbk[1], wall_spacing / ym), _stretch(letbk[2], wall_spacing / ym)]   # surface -> ym
    last = [en[0][-1] - en[0][-2], en[1][-1] - en[1][-2]]
    ef = [_stretch(letbk[0], ym * last[0] / (yb - ym)), _stretch(letbk[3], ym * last[1] / (yt - ym))]   # ym -> far
    nz = lzebk[0] + 1
    zs = -0.5 * span + span * np.arange(nz) / lzebk[0]
    blocks = []
    for j in range(4):
        for i in range(6):
            x = xcols[i]
            ts = naca_thickness(x - chord[i], thick, blend) if chord[i] is not None else np.zeros_like(x)
            if j == 0:
                y1 = -ym - (yb - ym) * ef[0][::-1]
                y2 = np.broadcast_to(y1[:, None], (y1.size, x.size))
            elif j == 1:
                e = en[0][::-1]   # far side (y = -ym, e = 1) to the surface (e = 0)
                y2 = -ts[None, :] + (-ym + ts[None, :]) * e[:, None]
            elif j == 2:
                y2 = ts[None, :] + (ym - ts[None, :]) * en[1][:, None]
            else:
                y1 = ym + (yt - ym) * ef[1]
                y2 = np.broadcast_to(y1[:, None], (y1.size, x.size))
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
