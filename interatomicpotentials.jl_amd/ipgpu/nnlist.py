"""neighborlist(s, rcutoff) on the GPU (replaces src/nnlist.jl:5-46)."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import IPG_LIST_FULL, IPG_LIST_HALF, check


class NeighborList:
    """src/nnlist.jl:5-15: per-atom neighbor indices j (0-based here, the
    Julia wrapper adds 1) and displacements r = get_displacement(x_i, x_j)."""

    def __init__(self, offsets, nbr, disp):
        self.offsets = offsets
        self.nbr = nbr
        self.disp = disp
        n = len(offsets) - 1
        self.j = [nbr[offsets[i]:offsets[i + 1]] for i in range(n)]
        self.r = [disp[offsets[i]:offsets[i + 1]] for i in range(n)]

    def __len__(self):
        return len(self.j)


def make_csystem(pos, box, periodic, species=None):
    """ctypes ipg_system over host arrays (kept alive by the returned tuple)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    s = _lib.System()
    s.natoms = pos.shape[0]
    s.positions = pos.ctypes.data
    sp = None
    if species is not None:
        sp = np.ascontiguousarray(species, dtype=np.int32)
        s.species = sp.ctypes.data
    for d in range(3):
        s.box[d] = float(box[d]) if periodic[d] else 0.0
        s.periodic[d] = 1 if periodic[d] else 0
    return s, (pos, sp)


def neighborlist_arrays(pos, box, periodic, rcut, kind=IPG_LIST_HALF, want_disp=True):
    """Raw CSR list (offsets, j, disp) from the GPU, positions in any unit."""
    L = _lib.load()
    ctx = _lib.default_context()
    s, keep = make_csystem(pos, box, periodic)
    n = s.natoms
    off = np.zeros(n + 1, dtype=np.int64)
    cap = max(64, int(n) * 64)
    for _ in range(2):
        nbr = np.zeros(max(cap, 1), dtype=np.int32)
        disp = np.zeros((max(cap, 1), 3)) if want_disp else None
        npairs = C.c_int64()
        rc = L.ipg_neighborlist(ctx, C.byref(s), float(rcut), kind, off.ctypes.data, nbr.ctypes.data,
                                disp.ctypes.data if want_disp else None, cap, C.byref(npairs))
        if rc == _lib.IPG_ERR_CAPACITY:
            cap = int(npairs.value)
            continue
        check(rc)
        m = int(npairs.value)
        return off, nbr[:m].copy(), (disp[:m].copy() if want_disp else None)
    raise _lib.IPGError(_lib.IPG_ERR_CAPACITY, "neighbor list capacity retry failed")


def neighborlist(s, rcutoff, full=False):
    """src/nnlist.jl:23-46: half list (j > i) on austrip'ed coordinates with
    the PeriodicEuclidean membership; rcutoff in bohr (a plain number, as the
    reference's Float64 argument)."""
    pos = s.austrip_positions()
    box = s.austrip_box()
    off, nbr, disp = neighborlist_arrays(pos, box, s.periodic, rcutoff, IPG_LIST_FULL if full else IPG_LIST_HALF)
    return NeighborList(off, nbr, disp)
