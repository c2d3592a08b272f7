"""Small cases shared by the CPU and GPU tests."""
import numpy as np

from tandemproject_amd import grid as G
from tandemproject_amd import layout as L

BASE_PARAMS = dict(reoo=2.5e4, tempoo=273.0, amach1=0.4, amach2=0.0, amach3=0.0, wtemp=1.0,
                   cfl=1.0, tmax=1e9, timf=2.0, fltk=0.85, fltkbc=0.85, dto=-1.0, domlen=4.0,
                   szth1=1.0, szth2=1.0, szxt=1.0, szco=12.0, span=0.2, wlew=0.11, wlea=0.015,
                   nkrk=4, nviscous=1, nsmf=0, les=0, nextrabc=0, nextgcic=0, aoa0=0.0, aoa1=0.0,
                   tps=0.0, tp=0.0, smago1=0.16, smago2=0.0066)


def aerofoil_case(lxibk=(24, 32, 24), letbk=(20, 20), lzebk=(16,), npcz=None, npcx=None,
                  nviscous=1, span=0.2, **over):
    lay = L.h_layout(list(lxibk), list(letbk), list(lzebk), npcx=npcx, npcz=npcz, nviscous=nviscous)
    blocks = G.h_grid_blocks(list(lxibk), list(letbk), list(lzebk), xu=4, xd=5, yb=4, yt=4, span=span)
    ss = G.subblock_grids(lay, blocks)
    p = dict(BASE_PARAMS); p["nviscous"] = nviscous; p["span"] = span; p.update(over)
    return lay, ss, p


def c1_case():
    """The plain CPU-runnable configuration: NACA0012 at AoA 0, M = 0.4,
    Re_c = 1e4 (reoo = Re_c / M = 2.5e4: the reference scales by the speed
    of sound), laminar, uniform-flow start, spanwise periodic over 0.1c.
    The 4 blocks of 96 x 64 x 8 become the reference's 3x2 H layout of the
    same in-plane size with 13 spanwise points, since the periodic closure
    reads 12 points beside the face (LMD + 1, src/Subroutines3D.f90:39): 6
    blocks, 207,662 points."""
    return aerofoil_case(lxibk=(40, 80, 40), letbk=(48, 48), lzebk=(12,), span=0.1,
                         amach1=0.4, reoo=2.5e4, nviscous=1)


def seeded_field(n_points_shape, seed, nmodes=8, noise=1e-3):
    """Sum of sinusoids in index space plus white noise (shape (nz, ny, nx))."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = n_points_shape
    k = np.arange(nz)[:, None, None]; j = np.arange(ny)[None, :, None]; i = np.arange(nx)[None, None, :]
    f = np.zeros((nz, ny, nx))
    for _ in range(nmodes):
        a, b, c = rng.uniform(0, 0.5, 3)
        ph = rng.uniform(0, 2 * np.pi)
        f += rng.uniform(0.2, 1.0) * np.sin(a * i + b * j + c * k + ph)
    f += noise * rng.standard_normal(f.shape)
    return f.ravel()
