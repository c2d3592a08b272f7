"""Block topology of the reference solver, built once on the host.

Restates the reference's domain decomposition so that one process can own
every sub-block:

* block boundary codes and neighbours: src/SingleTurbulence3D.f90:121-164
  (``domdcomp``; 10 characteristic far field, 20/25 inviscid/viscous wall,
  30 characteristic interface, 45 periodic);
* splitting of a block into ``npc`` sub-blocks with closure interfaces
  (code 40) and the neighbour sub-block of every face:
  src/Main3D.f90:92-98, 132-178;
* the 4-block corner lines averaged by ``junction``:
  src/SingleTurbulence3D.f90:378-443.

A sub-block here is what the reference runs as one MPI rank.  The arrays
produced are the plain int32 tables the C ABI takes (see include/tandem_hip.h).
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Layout:
    nb: int
    dims: np.ndarray            # (nb, 3) lxi, let, lze (points - 1)
    nbc: np.ndarray             # (nb, 3, 2) boundary code per face
    ncd: np.ndarray             # (nb, 3, 2) neighbour sub-block per face (-1: none)
    nsz: np.ndarray             # (nb, 3, 2) sponge flag per face of the parent block
    jct: np.ndarray             # (njct, 4, 3) (sub-block, ip, jp) of each corner member
    jvel: np.ndarray            # (njct, 2) members whose velocities are averaged
    parent: np.ndarray = None   # (nb,) parent block index
    pos: np.ndarray = None      # (nb, 3) sub-block coordinates (ip, jp, kp)
    start: np.ndarray = None    # (nb, 3) first point of the sub-block in its parent
    bk: tuple = (1, 1, 1)       # (bkx, bky, bkz)
    blockdims: np.ndarray = None  # (nblocks, 3) parent block dims
    extra: dict = field(default_factory=dict)

    @property
    def njct(self):
        return int(self.jct.shape[0])

    def npoints(self, b=None):
        d = self.dims + 1
        if b is None:
            return int(np.prod(d, axis=1).sum())
        return int(np.prod(d[b]))


def _domdcomp(mb, bkx, bky, mbk, nviscous, walls):
    """Boundary codes and neighbour blocks of block mb (SingleTurbulence3D.f90:121-164).

    The reference's 3x2 layout has its one aerofoil between rows 0 and 1 of
    column 1 (walls = [(0, 1)]); a wall (r, c) here puts an aerofoil surface
    between rows r and r + 1 of column c: the eta end of block (r, c) and the
    eta start of block (r + 1, c) are the wall (20 + 5 nviscous), each with
    the block across the aerofoil as neighbour."""
    bcw = 20 + 5 * nviscous
    ms = [0, 0, 0]; me = [0, 0, 0]; nbcs = [0, 0, 0]; nbce = [0, 0, 0]
    ll = mb % bkx
    if ll == 0:
        ms[0], me[0], nbcs[0], nbce[0] = mb + (bkx - 1), mb + 1, 10, 30
    elif ll == bkx - 1:
        ms[0], me[0], nbcs[0], nbce[0] = mb - 1, mb - (bkx - 1), 30, 10
    else:
        ms[0], me[0], nbcs[0], nbce[0] = mb - 1, mb + 1, 30, 30
    if mb < bkx:
        ms[1], me[1], nbcs[1], nbce[1] = mb + bkx * (bky - 1), mb + bkx, 10, 30
    elif mb > mbk - bkx:
        ms[1], me[1], nbcs[1], nbce[1] = mb - bkx, mb - bkx * (bky - 1), 30, 10
    else:
        ms[1], me[1], nbcs[1], nbce[1] = mb - bkx, mb + bkx, 30, 30
    nbcs[2] = nbce[2] = 45
    ms[2] = me[2] = mb
    row = mb // bkx
    for (r, c) in walls:
        if ll == c and row == r:
            nbce[1] = bcw
        if ll == c and row == r + 1:
            nbcs[1] = bcw
    return ms, me, nbcs, nbce


def h_layout(lxibk, letbk, lzebk, npcx=None, npcy=None, npcz=None, nviscous=1, walls=None):
    """The reference's bkx x bky x bkz H-type block layout split into sub-blocks.

    lxibk/letbk/lzebk are the per-column/row/plane last indices as in
    inputp.dat lines 1-3; npcx/npcy/npcz the sub-block counts (lines 6-8).
    walls: aerofoil surfaces as (row, column) pairs, each between rows row
    and row + 1 of its column (default: the reference's 3x2 layout, one
    aerofoil between the two rows of the middle column).
    """
    bkx, bky, bkz = len(lxibk), len(letbk), len(lzebk)
    if walls is None:
        walls = [(0, 1)]
    if bkz != 1:
        raise ValueError("one block plane in zeta (split it with npcz)")
    for (r, c) in walls:
        if not (0 <= r < bky - 1 and 0 < c < bkx - 1):
            raise ValueError(f"wall {(r, c)} must lie between two rows of an inner column")
    npcx = list(npcx or [1] * bkx)
    npcy = list(npcy or [1] * bky)
    npcz = list(npcz or [1] * bkz)
    nblk = bkx * bky * bkz
    mbk = nblk - 1
    bdims = np.zeros((nblk, 3), np.int64)
    npc = np.zeros((nblk, 3), np.int64)
    for k in range(bkz):
        for j in range(bky):
            for i in range(bkx):
                l = k * (bkx * bky) + j * bkx + i
                bdims[l] = (lxibk[i], letbk[j], lzebk[k])
                npc[l] = (npcx[i], npcy[j], npcz[k])
    mo = np.zeros(nblk + 1, np.int64)
    for mm in range(1, nblk + 1):
        mo[mm] = mo[mm - 1] + npc[mm - 1].prod()
    nranks = int(mo[nblk])
    dims = np.zeros((nranks, 3), np.int32)
    nbc = np.zeros((nranks, 3, 2), np.int32)
    ncd = np.full((nranks, 3, 2), -1, np.int32)
    nsz = np.zeros((nranks, 3, 2), np.int32)
    parent = np.zeros(nranks, np.int32)
    pos = np.zeros((nranks, 3), np.int32)
    start = np.zeros((nranks, 3), np.int32)
    for myid in range(nranks):
        mb = int(np.searchsorted(mo, myid, side="right") - 1)
        ms, me, nbcs, nbce = _domdcomp(mb, bkx, bky, mbk, nviscous, walls)
        rel = myid - mo[mb]
        ip = rel % npc[mb, 0]
        jp = (rel // npc[mb, 0]) % npc[mb, 1]
        kp = (rel // (npc[mb, 0] * npc[mb, 1])) % npc[mb, 2]
        ncds = [0, 0, 0]; ncde = [0, 0, 0]
        ncds[0] = mo[ms[0]] + kp * npc[ms[0], 1] * npc[ms[0], 0] + jp * npc[ms[0], 0] + npc[ms[0], 0] - 1
        ncde[0] = mo[me[0]] + kp * npc[me[0], 1] * npc[me[0], 0] + jp * npc[me[0], 0]
        ncds[1] = mo[ms[1]] + kp * npc[ms[1], 1] * npc[ms[1], 0] + (npc[ms[1], 1] - 1) * npc[ms[1], 0] + ip
        ncde[1] = mo[me[1]] + kp * npc[me[1], 1] * npc[me[1], 0] + ip
        ncds[2] = mo[ms[2]] + (npc[ms[2], 2] - 1) * npc[ms[2], 1] * npc[ms[2], 0] + jp * npc[ms[2], 0] + ip
        ncde[2] = mo[me[2]] + jp * npc[me[2], 0] + ip
        lpv = (ip, jp, kp)
        mpv = (1, npc[mb, 0], npc[mb, 0] * npc[mb, 1])
        for nn in range(3):
            ll = int(bdims[mb, nn]); lp = int(lpv[nn]); mp = int(mpv[nn]); ma = int(npc[mb, nn])
            if ma == 1:
                l = ll
                nbc[myid, nn] = (nbcs[nn], nbce[nn]); ncd[myid, nn] = (ncds[nn], ncde[nn])
                st = 0
            else:
                first = ll - ((ll + 1) // ma) * (ma - 1)
                if lp == 0:
                    l = first
                    nbc[myid, nn] = (nbcs[nn], 40); ncd[myid, nn] = (ncds[nn], myid + mp)
                    st = 0
                else:
                    l = (ll + 1) // ma - 1
                    st = first + 1 + (lp - 1) * ((ll + 1) // ma)
                    if lp < ma - 1:
                        nbc[myid, nn] = (40, 40); ncd[myid, nn] = (myid - mp, myid + mp)
                    else:
                        nbc[myid, nn] = (40, nbce[nn]); ncd[myid, nn] = (myid - mp, ncde[nn])
            dims[myid, nn] = l
            start[myid, nn] = st
            nsz[myid, nn] = (1 if nbcs[nn] == 10 else 0, 1 if nbce[nn] == 10 else 0)
        parent[myid] = mb
        pos[myid] = lpv
    # characteristic faces never exchange closure sums; keep their neighbour only
    # where the stage exchanges (30 interface, 45 periodic, 40 split)
    for r in range(nranks):
        for nn in range(3):
            for s in range(2):
                if nbc[r, nn, s] in (10, 20, 25):
                    ncd[r, nn, s] = -1 if nbc[r, nn, s] == 10 else ncd[r, nn, s]
    # junctions (SingleTurbulence3D.f90:378-443): the corner lines where an
    # aerofoil's leading or trailing edge meets the blocks beside it, per
    # spanwise sub-block: members (ip, jp, block) bottom-left, bottom-right,
    # top-left, top-right; the velocity is averaged over the two members
    # whose faces are the wall (right ones at a leading edge, left ones at a
    # trailing edge)
    wallset = set(walls)
    corners = []
    for (r, c) in sorted(walls, key=lambda w: (w[0], w[1])):
        if (r, c - 1) not in wallset:
            corners.append((r, c - 1, 1))   # leading edge: walls on the right
        if (r, c + 1) not in wallset:
            corners.append((r, c, 0))       # trailing edge: walls on the left
    jct = []; jvel = []
    for kpl in range(int(npc[0, 2])):
        for (r, cl, le) in corners:
            mem = []
            base = r * bkx + cl
            for (ip, jp, mm) in ((1, 1, base), (0, 1, base + 1), (1, 0, base + bkx), (0, 0, base + bkx + 1)):
                rr = mo[mm] + kpl * npc[mm, 1] * npc[mm, 0] + jp * (npc[mm, 1] - 1) * npc[mm, 0] + ip * (npc[mm, 0] - 1)
                mem.append((int(rr), ip, jp))
            jct.append(mem)
            jvel.append((1, 3) if le else (0, 2))
    jct = np.array(jct, np.int32).reshape(-1, 4, 3)
    jvel = np.array(jvel, np.int32).reshape(-1, 2)
    return Layout(nb=nranks, dims=dims, nbc=nbc, ncd=ncd, nsz=nsz, jct=jct, jvel=jvel,
                  parent=parent, pos=pos, start=start, bk=(bkx, bky, bkz),
                  blockdims=bdims.astype(np.int32), extra={"walls": list(walls)})


def tandem_layout(lxibk, letbk, lzebk, chord_cols=(1, 4), npcz=None, nviscous=1):
    """Two aerofoils in tandem on a bkx x bky H layout: each chord column
    carries an aerofoil between the two middle rows (no counterpart in the
    reference, whose decks hold one aerofoil; the block codes, neighbours,
    sub-blocks and junctions follow the same rules as h_layout)."""
    bky = len(letbk)
    r = bky // 2 - 1
    return h_layout(lxibk, letbk, lzebk, npcz=npcz, nviscous=nviscous, walls=[(r, c) for c in chord_cols])


def single_block(lxi, let, lze, nbc, ncd=None):
    """One block with explicit boundary codes (``nbc[nn][ip]``) and neighbours.

    Periodic (45) and closure (35/40) faces default to the block itself as
    neighbour, which is the reference's self-exchange (Main3D.f90:145-146).
    """
    nbc = np.array(nbc, np.int32).reshape(1, 3, 2)
    if ncd is None:
        ncd = np.where(np.isin(nbc, (30, 35, 40, 45)), 0, -1).astype(np.int32)
    ncd = np.array(ncd, np.int32).reshape(1, 3, 2)
    nsz = (nbc == 10).astype(np.int32)
    return Layout(nb=1, dims=np.array([[lxi, let, lze]], np.int32), nbc=nbc, ncd=ncd, nsz=nsz,
                  jct=np.zeros((0, 4, 3), np.int32), jvel=np.zeros((0, 2), np.int32),
                  parent=np.zeros(1, np.int32), pos=np.zeros((1, 3), np.int32),
                  start=np.zeros((1, 3), np.int32), blockdims=np.array([[lxi, let, lze]], np.int32))


def concat(layouts):
    """Several independent cases side by side (block indices renumbered)."""
    off = 0
    parts = {k: [] for k in ("dims", "nbc", "ncd", "nsz", "jct", "jvel")}
    case_of = []
    for ci, L in enumerate(layouts):
        parts["dims"].append(L.dims); parts["nbc"].append(L.nbc); parts["nsz"].append(L.nsz)
        parts["ncd"].append(np.where(L.ncd >= 0, L.ncd + off, -1).astype(np.int32))
        j = L.jct.copy(); j[:, :, 0] += off
        parts["jct"].append(j); parts["jvel"].append(L.jvel)
        case_of += [ci] * L.nb
        off += L.nb
    return Layout(nb=off, dims=np.concatenate(parts["dims"]), nbc=np.concatenate(parts["nbc"]),
                  ncd=np.concatenate(parts["ncd"]), nsz=np.concatenate(parts["nsz"]),
                  jct=np.concatenate(parts["jct"]).reshape(-1, 4, 3),
                  jvel=np.concatenate(parts["jvel"]).reshape(-1, 2),
                  extra={"case_of": np.array(case_of, np.int32)})
