"""SNAP descriptors and linear SNAP potentials on the GPU.

Mirrors src/BasisSystems/SNAP/snap.jl (constructor flags, compute_local /
force / all descriptors), utilities/utl.jl (the SNAP neighbor list) and
src/BasisPotentials (LBasisPotential energy/force).  SNAP is unit-naive as in
the reference: coordinates are the system's raw numbers (ustrip).
"""
import collections
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (IPG_FP32, IPG_FP64, IPG_SNAP_ADJOINT_FP32, IPG_SNAP_ADJOINT_FP64, IPG_SNAP_LIST_CUTOFF,
                   IPG_SNAP_LIST_REFERENCE, IPG_SNAP_PLAIN_FP64, check)
from .nnlist import make_csystem
from .system import species_index, symbol_codes

EnergyForce = collections.namedtuple("EnergyForce", ["e", "f"])
SnapNeighborList = collections.namedtuple("SnapNeighborList", ["offsets", "j", "R", "r"])

_MODES = {"reference": IPG_SNAP_LIST_REFERENCE, "cutoff": IPG_SNAP_LIST_CUTOFF}
_PATHS = {"plain": IPG_SNAP_PLAIN_FP64, "adjoint": IPG_SNAP_ADJOINT_FP64, "adjoint_fp64": IPG_SNAP_ADJOINT_FP64,
          "adjoint_fp32": IPG_SNAP_ADJOINT_FP32}


def get_num_snap_coeffs(twojmax, n_elems, chem_flag):
    """utl.jl:19-37."""
    J = twojmax
    if J % 2 == 0:
        m = J // 2 + 1
        nc = m * (m + 1) * (2 * m + 1) // 6
    else:
        m = (J + 1) // 2
        nc = m * (m + 1) * (m + 2) // 3
    return nc * n_elems ** 3 if chem_flag else nc


class SNAP:
    """SNAP basis (snap.jl:20-36 flag semantics).  list_mode 'reference'
    reproduces utl.jl's neighbor list as-is (inrange radius sqrt(cutmax));
    'cutoff' uses the full cutoff."""

    def __init__(self, n_atoms, twojmax, species, rcutfac, rmin0, rfac0, radii, weight, chem_flag=False,
                 bzero_flag=False, bnorm_flag=False, switch_flag=False, wselfall_flag=False,
                 list_mode="reference"):
        if bool(chem_flag) != bool(bnorm_flag):
            raise ValueError("chem_flag and bnorm_flag should be the same, setting bnorm_flag == chem_flag")
        self.n_atoms = int(n_atoms)
        self.twojmax = int(twojmax)
        self.species = [str(s) for s in species]
        self.rcutfac, self.rmin0, self.rfac0 = float(rcutfac), float(rmin0), float(rfac0)
        self.radii = np.ascontiguousarray(radii, dtype=np.float64)
        self.weight = np.ascontiguousarray(weight, dtype=np.float64)
        self.chem_flag, self.bzero_flag, self.bnorm_flag = bool(chem_flag), bool(bzero_flag), bool(bnorm_flag)
        self.switch_flag, self.wselfall_flag = bool(switch_flag), bool(wselfall_flag)
        # runtime-table kernels (FP64, beta . dB route only)
        self.generic = self.chem_flag or self.twojmax > 8
        if list_mode not in _MODES:
            raise ValueError("list_mode must be 'reference' or 'cutoff'")
        self.list_mode = list_mode
        p = _lib.SnapParams()
        p.twojmax, p.nspecies = self.twojmax, len(self.species)
        p.rcutfac, p.rmin0, p.rfac0 = self.rcutfac, self.rmin0, self.rfac0
        p.radii = self.radii.ctypes.data_as(C.POINTER(C.c_double))
        p.weight = self.weight.ctypes.data_as(C.POINTER(C.c_double))
        p.chem_flag, p.bzero_flag, p.bnorm_flag = int(chem_flag), int(bzero_flag), int(bnorm_flag)
        p.switch_flag, p.wselfall_flag = int(switch_flag), int(wselfall_flag)
        self._params = p
        L = _lib.load()
        h = C.c_void_p()
        check(L.ipg_snap_create(_lib.default_context(), C.byref(p), C.byref(h)))
        self._h = h
        nc = C.c_int32()
        check(L.ipg_snap_info(h, C.byref(nc), None, None, None))
        self.ncoeff = nc.value

    def __len__(self):
        return get_num_snap_coeffs(self.twojmax, len(self.species), self.chem_flag)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ipg_snap_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def get_species(self):
        return self.species

    def graph_info(self):
        """Captured-step state of the energy/force calls: dict with state
        ("none", "ready", "disabled"), captures, replays, overflows, reason."""
        st = C.c_int32()
        cap, rep, ovf = C.c_int64(), C.c_int64(), C.c_int64()
        buf = C.create_string_buffer(512)
        check(_lib.load().ipg_snap_graph_info(self._h, C.byref(st), C.byref(cap), C.byref(rep), C.byref(ovf),
                                              buf, 512))
        return {"state": ("none", "ready", "disabled")[st.value], "captures": cap.value, "replays": rep.value,
                "overflows": ovf.value, "reason": buf.value.decode()}

    def get_rcutoff(self):
        return float(np.max(2.0 * self.radii * self.rcutfac))

    # helpers ------------------------------------------------------------
    def _sys(self, s):
        uniq, appear = symbol_codes(s)
        sp = species_index(uniq, self.species)[appear]
        cs, keep = make_csystem(s.positions, s.box, s.periodic, sp)
        return cs, keep, appear

    def _mode(self, list_mode):
        return _MODES[list_mode or self.list_mode]


def snap_neighborlist(s, snap, list_mode=None):
    """utl.jl:38-101 neighborlist(A, snap): full list, sorted j, R, r = x_j - x_i."""
    L = _lib.load()
    cs, keep, appear = snap._sys(s)
    n = cs.natoms
    off = np.zeros(n + 1, dtype=np.int64)
    cap = max(64, n * 96)
    for _ in range(2):
        nbr = np.zeros(max(cap, 1), dtype=np.int32)
        R = np.zeros(max(cap, 1))
        r = np.zeros((max(cap, 1), 3))
        np_ = C.c_int64()
        rc = L.ipg_snap_neighborlist(snap._h, C.byref(cs), snap._mode(list_mode), appear.ctypes.data,
                                     off.ctypes.data, nbr.ctypes.data, R.ctypes.data, r.ctypes.data, cap,
                                     C.byref(np_))
        if rc == _lib.IPG_ERR_CAPACITY:
            cap = int(np_.value)
            continue
        check(rc)
        m = int(np_.value)
        return SnapNeighborList(off, nbr[:m].copy(), R[:m].copy(), r[:m].copy())
    raise _lib.IPGError(_lib.IPG_ERR_CAPACITY, "capacity retry failed")


def compute_local_descriptors(s, snap, list_mode=None):
    """snap.jl:103-124: B, shape (N, ncoeff)."""
    L = _lib.load()
    cs, keep, appear = snap._sys(s)
    B = np.zeros((cs.natoms, snap.ncoeff))
    check(L.ipg_snap_local_descriptors(snap._h, C.byref(cs), snap._mode(list_mode), appear.ctypes.data,
                                       B.ctypes.data))
    return B


def compute_force_descriptors(s, snap, list_mode=None):
    """snap.jl:126-167 (row fix of :261): dB, shape (N, 3, ncoeff*nspecies)."""
    return compute_all_descriptors(s, snap, list_mode, want=("dB",))[1]


def compute_all_descriptors(s, snap, list_mode=None, want=("B", "dB", "W")):
    """snap.jl:226-285: (B (N, ncoeff), dB (N, 3, M), W (6, M)), M = ncoeff*nspecies."""
    L = _lib.load()
    cs, keep, appear = snap._sys(s)
    n, M = cs.natoms, snap.ncoeff * len(snap.species)
    B = np.zeros((n, snap.ncoeff)) if "B" in want else None
    dB = np.zeros((n, M, 3)) if "dB" in want else None
    W = np.zeros((M, 6)) if "W" in want else None
    check(L.ipg_snap_all_descriptors(snap._h, C.byref(cs), snap._mode(list_mode), appear.ctypes.data,
                                     None if B is None else B.ctypes.data, None if dB is None else dB.ctypes.data,
                                     None if W is None else W.ctypes.data))
    return (B, None if dB is None else np.ascontiguousarray(dB.transpose(0, 2, 1)),
            None if W is None else np.ascontiguousarray(W.T))


BatchDescriptors = collections.namedtuple("BatchDescriptors", ["B", "dB", "erow", "W", "start"])


def compute_batch_descriptors(systems, snap, precision="fp64", want=("B", "dB", "erow", "W")):
    """Descriptors of a data set in one GPU call (a fitting loop over
    snap.jl:103-124 and :226-285, CUTOFF list mode).  Returns packed arrays:
    B (Ntot, ncoeff), dB (Ntot, 3, M), erow (nconf, M) = per-configuration
    sums of B in species blocks (M = ncoeff*nspecies), W (nconf, 6, M) =
    each configuration's virial descriptors as compute_all_descriptors
    returns them, and start (nconf+1), the first packed atom of each
    configuration.  precision 'fp32' runs U, Z
    and the per-pair dB in FP32 with FP64 sums."""
    L = _lib.load()
    systems = list(systems)
    if not systems:
        raise ValueError("no configurations")
    arr = (_lib.System * len(systems))()
    keep = []
    for c, s in enumerate(systems):
        uniq, codes = symbol_codes(s)
        cs, k = make_csystem(s.positions, s.box, s.periodic, species_index(uniq, snap.species)[codes])
        arr[c] = cs
        keep.append(k)
    natoms = np.array([a.natoms for a in arr], dtype=np.int64)
    start = np.concatenate([[0], np.cumsum(natoms)]).astype(np.int64)
    n, M = int(start[-1]), snap.ncoeff * len(snap.species)
    B = np.zeros((n, snap.ncoeff)) if "B" in want else None
    dB = np.zeros((n, M, 3)) if "dB" in want else None
    erow = np.zeros((len(systems), M)) if "erow" in want else None
    W = np.zeros((len(systems), M, 6)) if "W" in want else None
    prec = IPG_FP32 if precision == "fp32" else IPG_FP64
    check(L.ipg_snap_batch_descriptors(snap._h, len(systems), arr, prec, None if B is None else B.ctypes.data,
                                       None if dB is None else dB.ctypes.data,
                                       None if erow is None else erow.ctypes.data,
                                       None if W is None else W.ctypes.data))
    return BatchDescriptors(B, None if dB is None else np.ascontiguousarray(dB.transpose(0, 2, 1)), erow,
                            None if W is None else np.ascontiguousarray(W.transpose(0, 2, 1)), start)


def batch_info(snap, precision="fp64"):
    """Launch shapes the batch autotuner chose and the captured-graph state."""
    v = [C.c_int32() for _ in range(4)]
    rep, npairs = C.c_int64(), C.c_int64()
    check(_lib.load().ipg_snap_batch_info(snap._h, IPG_FP32 if precision == "fp32" else IPG_FP64,
                                          *[C.byref(x) for x in v], C.byref(rep), C.byref(npairs)))
    return {"scan_lanes": v[0].value, "ui_atoms_per_wave": v[1].value, "dbdr_threads": v[2].value,
            "graph": ("none", "ready", "disabled")[v[3].value], "replays": rep.value, "npairs": npairs.value}


class LBasisPotential:
    """linear_potential.jl:6-49: E = beta0 + sum_i beta_s(i) . B_i, F = dB . beta.
    beta has ncoeff entries (shared by every species, as the reference) or
    ncoeff*nspecies (one block per species)."""

    def __init__(self, beta, beta0, basis):
        self.beta = np.ascontiguousarray(beta, dtype=np.float64).reshape(-1)
        b0 = np.asarray(beta0, dtype=np.float64).reshape(-1)
        self.beta0 = float(b0[0]) if b0.size else 0.0
        self.basis = basis

    def get_rcutoff(self):
        return self.basis.get_rcutoff()

    def get_species(self):
        return self.basis.get_species()

    def expanded_beta(self):
        nc, ns = self.basis.ncoeff, len(self.basis.species)
        return np.tile(self.beta, ns) if self.beta.size == nc else self.beta.copy()


def potential_energy_from_descriptors(B, lbp, species=None):
    """linear_potential.jl:20-26 (per-species blocks when species is given)."""
    B = np.asarray(B)
    if species is None or lbp.beta.size == lbp.basis.ncoeff:
        return lbp.beta0 + float(B.sum(0) @ lbp.beta[:B.shape[1]])
    nc = lbp.basis.ncoeff
    be = lbp.expanded_beta().reshape(-1, nc)
    return lbp.beta0 + float(np.einsum("nc,nc->", B, be[np.asarray(species)]))


def force_from_descriptors(dB, lbp):
    """linear_potential.jl:28-35 with the intended contraction over every block."""
    return np.einsum("nkm,m->nk", np.asarray(dB), lbp.expanded_beta())


def _path(snap, path):
    """Default path: 'adjoint', or 'plain' for the configurations that only
    have the reference's beta . dB route (chem_flag, twojmax > 8)."""
    if path is None:
        return "plain" if snap.generic else "adjoint"
    return path


def energy_and_force(s, lbp, path=None, list_mode=None):
    """basis_potentials.jl:32-41.  path: 'plain' (FP64 dB then beta, as the
    reference), 'adjoint' (FP64), 'adjoint_fp32' (FP32, FP64 reductions);
    None: 'adjoint', or 'plain' where that is the only route (chem_flag,
    twojmax > 8)."""
    L = _lib.load()
    snap = lbp.basis
    cs, keep, appear = snap._sys(s)
    e = C.c_double()
    f = np.zeros((cs.natoms, 3))
    check(L.ipg_snap_energy_forces(snap._h, C.byref(cs), snap._mode(list_mode), appear.ctypes.data,
                                   lbp.beta.ctypes.data, lbp.beta.size, lbp.beta0, _PATHS[_path(snap, path)], C.byref(e),
                                   f.ctypes.data))
    return EnergyForce(e.value, f)


def virial_stress(s, lbp, path=None, list_mode=None):
    """basis_potentials.jl:43-50 with linear_potential.jl:37-42: W * beta,
    W the virial descriptors of compute_all_descriptors (snap.jl:267-281,
    raw positions), rows [xx, yy, zz, zy, zx, yx].  path 'plain' forms W
    from the per-pair dB as the reference does; the adjoint paths contract
    the per-pair dE/dr with x_i - x_j on the device."""
    L = _lib.load()
    snap = lbp.basis
    cs, keep, appear = snap._sys(s)
    v = np.zeros(6)
    check(L.ipg_snap_virial_stress(snap._h, C.byref(cs), snap._mode(list_mode), appear.ctypes.data,
                                   lbp.beta.ctypes.data, lbp.beta.size, _PATHS[_path(snap, path)], v.ctypes.data))
    return v


def virial(s, lbp, **kw):
    """basis_potentials.jl:52-57: the sum of the six virial_stress entries."""
    return float(np.sum(virial_stress(s, lbp, **kw)))


EnergyForceVirial = collections.namedtuple("EnergyForceVirial", ["e", "f", "v"])


def energy_force_virial(s, lbp, path=None, list_mode=None):
    """Energy, forces and virial stress from one evaluation."""
    L = _lib.load()
    snap = lbp.basis
    cs, keep, appear = snap._sys(s)
    e = C.c_double()
    f = np.zeros((cs.natoms, 3))
    v = np.zeros(6)
    check(L.ipg_snap_energy_forces_virial(snap._h, C.byref(cs), snap._mode(list_mode), appear.ctypes.data,
                                          lbp.beta.ctypes.data, lbp.beta.size, lbp.beta0, _PATHS[_path(snap, path)],
                                          C.byref(e), f.ctypes.data, v.ctypes.data))
    return EnergyForceVirial(e.value, f, v)


def potential_energy(s, lbp, **kw):
    return energy_and_force(s, lbp, **kw).e


def force(s, lbp, **kw):
    return energy_and_force(s, lbp, **kw).f


def force_vjp(s, lbp, gF, precision="fp64", list_mode=None):
    """Reverse mode of F(x, beta): returns (gF . dF/dbeta, gF . dF/dx)."""
    L = _lib.load()
    snap = lbp.basis
    cs, keep, appear = snap._sys(s)
    g = np.ascontiguousarray(gF, dtype=np.float64).reshape(-1, 3)
    gb = np.zeros(lbp.beta.size)
    gp = np.zeros((cs.natoms, 3))
    prec = IPG_FP32 if precision == "fp32" else IPG_FP64
    check(L.ipg_snap_force_vjp(snap._h, C.byref(cs), snap._mode(list_mode), appear.ctypes.data,
                               lbp.beta.ctypes.data, lbp.beta.size, g.ctypes.data, prec, gb.ctypes.data,
                               gp.ctypes.data))
    return gb, gp


def force_rrule(s, lbp, path=None, precision="fp64", list_mode=None):
    """ChainRulesCore-style rrule of force(s, lbp): returns (F, pullback),
    pullback(gF) -> (gbeta, gpositions)."""
    F = energy_and_force(s, lbp, path=path, list_mode=list_mode).f

    def pullback(gF):
        return force_vjp(s, lbp, gF, precision=precision, list_mode=list_mode)

    return F, pullback
