"""ctypes wrapper of the CPU oracle (liboracle_ipg.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  Every function
forwards to the C restatement in ipg_oracle.c, which cites the reference
file:line it follows.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle_ipg.so")

PAIR_LJ, PAIR_MORSE, PAIR_ZBL, PAIR_BM, PAIR_COULOMB, PAIR_MOCK = range(6)
LIST_HALF, LIST_FULL = 0, 1
SNAP_LIST_REFERENCE, SNAP_LIST_CUTOFF = 0, 1
ALL_SPECIES = (1 << 64) - 1


class PairTerm(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("pad_", C.c_int32),
        ("coeff", C.c_double),
        ("rcut", C.c_double),
        ("p", C.c_double * 6),
        ("species_mask", C.c_uint64),
    ]


class SnapParams(C.Structure):
    _fields_ = [
        ("twojmax", C.c_int32),
        ("nspecies", C.c_int32),
        ("rcutfac", C.c_double),
        ("rmin0", C.c_double),
        ("rfac0", C.c_double),
        ("radii", C.POINTER(C.c_double)),
        ("weight", C.POINTER(C.c_double)),
        ("chem_flag", C.c_int32),
        ("bzero_flag", C.c_int32),
        ("bnorm_flag", C.c_int32),
        ("switch_flag", C.c_int32),
        ("wselfall_flag", C.c_int32),
    ]


class ZIdx(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("j1", "j2", "j", "ma1min", "ma2max", "mb1min", "mb2max", "na", "nb", "jju")]


class SnapTables(C.Structure):
    _fields_ = [
        ("twojmax", C.c_int),
        ("num_elements", C.c_int),
        ("idxcg_max", C.c_int),
        ("idxu_max", C.c_int),
        ("idxb_max", C.c_int),
        ("idxz_max", C.c_int),
        ("idxcg_block", C.POINTER(C.c_int)),
        ("idxu_block", C.POINTER(C.c_int)),
        ("idxb_block", C.POINTER(C.c_int)),
        ("idxz_block", C.POINTER(C.c_int)),
        ("idxb", C.POINTER(C.c_int)),
        ("idxz", C.POINTER(ZIdx)),
        ("cglist", C.POINTER(C.c_double)),
        ("rootpq", C.POINTER(C.c_double)),
        ("wself", C.c_double),
        ("bzero", C.POINTER(C.c_double)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(_LIBPATH)
        d, i64, i32 = C.c_double, C.c_int64, C.c_int32
        P = C.c_void_p
        L.orc_jl_mod.restype = d
        L.orc_jl_mod.argtypes = [d, d]
        L.orc_pe_sumsq.restype = d
        L.orc_pe_sumsq.argtypes = [P, P, P]
        L.orc_get_displacement.argtypes = [P, P, P, P]
        for f in (L.orc_nlist_brute, L.orc_nlist_cells):
            f.restype = i64
            f.argtypes = [P, i64, P, d, C.c_int, P, P, P, i64]
        L.orc_pair_V.restype = d
        L.orc_pair_V.argtypes = [C.POINTER(PairTerm), d]
        L.orc_pair_F.argtypes = [C.POINTER(PairTerm), d, P, P]
        L.orc_pair_energy_force_virial.restype = C.c_int
        L.orc_pair_energy_force_virial.argtypes = [P, i64, P, P, P, C.c_int, C.c_int, P, P, P]
        L.orc_pair_energy_force_virial_mt.restype = C.c_int
        L.orc_pair_energy_force_virial_mt.argtypes = [P, i64, P, P, P, C.c_int, C.c_int, C.c_int, P, P, P]
        L.orc_snap_tables_build.restype = C.POINTER(SnapTables)
        L.orc_snap_tables_build.argtypes = [C.c_int] * 5
        L.orc_snap_tables_free.argtypes = [C.POINTER(SnapTables)]
        L.orc_snap_num_coeffs.restype = C.c_int
        L.orc_snap_num_coeffs.argtypes = [C.c_int] * 3
        L.orc_snap_neighborlist.restype = i64
        L.orc_snap_neighborlist.argtypes = [P, i64, P, P, P, C.POINTER(SnapParams), C.c_int, P, P, P, P, i64]
        L.orc_snap_local_descriptors.restype = C.c_int
        L.orc_snap_local_descriptors.argtypes = [C.POINTER(SnapParams), C.POINTER(SnapTables), i64, P, P, P, P, P]
        L.orc_snap_all_descriptors.restype = C.c_int
        L.orc_snap_all_descriptors.argtypes = [C.POINTER(SnapParams), C.POINTER(SnapTables), i64,
                                               P, P, P, P, P, P, P, P]
        L.orc_snap_adjoint_energy_forces.restype = C.c_int
        L.orc_snap_adjoint_energy_forces.argtypes = [C.POINTER(SnapParams), C.POINTER(SnapTables), i64,
                                                     P, P, P, P, P, d, P, P, P, P]
        L.orc_snap_plain_energy_forces.restype = C.c_int
        L.orc_snap_plain_energy_forces.argtypes = [C.POINTER(SnapParams), C.POINTER(SnapTables), i64, i64,
                                                   P, P, P, P, P, C.c_int, P, P]
        L.orc_snap_atom_uzb.restype = C.c_int
        L.orc_snap_atom_uzb.argtypes = [C.POINTER(SnapParams), C.POINTER(SnapTables), C.c_int, C.c_int,
                                        P, P, P, P, P, P]
        L.orc_snap_uarray.restype = C.c_int
        L.orc_snap_uarray.argtypes = [C.POINTER(SnapTables), d, d, d, P, P]
        L.orc_snap_duarray.restype = C.c_int
        L.orc_snap_duarray.argtypes = [C.POINTER(SnapParams), C.POINTER(SnapTables), d, d, P, P]
        L.orc_wtime.restype = d
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _box(L):
    return np.ascontiguousarray(np.asarray(L, dtype=np.float64).reshape(3))


def jl_mod(x, y):
    return lib().orc_jl_mod(float(x), float(y))


def pe_dist(a, b, L):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return float(np.sqrt(lib().orc_pe_sumsq(_p(a), _p(b), _p(_box(L)))))


def get_displacement(L, x, y):
    out = np.zeros(3)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    lib().orc_get_displacement(_p(_box(L)), _p(x), _p(y), _p(out))
    return out


def neighborlist(pos, L, rcut, kind=LIST_HALF, cells=False):
    """Half (or full) list as (offsets, j, disp); j is 0-based, sorted per atom."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    f = lib().orc_nlist_cells if cells else lib().orc_nlist_brute
    off = np.zeros(n + 1, dtype=np.int64)
    Lb = _box(L)
    npairs = f(_p(pos), n, _p(Lb), float(rcut), kind, _p(off), None, None, 0)
    if npairs < 0:
        raise ValueError("invalid neighborlist input")
    nbr = np.zeros(max(npairs, 1), dtype=np.int32)
    disp = np.zeros((max(npairs, 1), 3))
    f(_p(pos), n, _p(Lb), float(rcut), kind, _p(off), _p(nbr), _p(disp), npairs)
    return off, nbr[:npairs], disp[:npairs]


def make_terms(terms):
    """terms: list of dicts {kind, coeff, rcut, p, species_mask}."""
    arr = (PairTerm * max(len(terms), 1))()
    for k, t in enumerate(terms):
        arr[k].kind = t["kind"]
        arr[k].coeff = t.get("coeff", 1.0)
        arr[k].rcut = t["rcut"]
        for q, v in enumerate(t["p"]):
            arr[k].p[q] = v
        arr[k].species_mask = t.get("species_mask", ALL_SPECIES)
    return arr


def pair_energy_force_virial(pos, L, species, terms, cells=False, nthreads=1):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    sp = np.ascontiguousarray(species, dtype=np.int32)
    arr = make_terms(terms)
    e = np.zeros(1)
    f = np.zeros((n, 3))
    v = np.zeros(6)
    rc = lib().orc_pair_energy_force_virial_mt(_p(pos), n, _p(_box(L)), _p(sp), C.cast(arr, C.c_void_p),
                                               len(terms), int(cells), int(nthreads), _p(e), _p(f), _p(v))
    if rc != 0:
        raise ValueError("invalid pair input")
    return e[0], f, v


def pair_V(term, R):
    arr = make_terms([term])
    return lib().orc_pair_V(arr, float(R))


def pair_F(term, R, r):
    arr = make_terms([term])
    out = np.zeros(3)
    r = np.ascontiguousarray(r, dtype=np.float64)
    lib().orc_pair_F(arr, float(R), _p(r), _p(out))
    return out


class Snap:
    """Oracle SNAP basis (snap.jl:20-36 14-argument constructor semantics)."""

    def __init__(self, twojmax, nspecies, rcutfac, rmin0, rfac0, radii, weight, chem_flag=False,
                 bzero_flag=False, bnorm_flag=False, switch_flag=False, wselfall_flag=False):
        self.radii = np.ascontiguousarray(radii, dtype=np.float64)
        self.weight = np.ascontiguousarray(weight, dtype=np.float64)
        p = SnapParams()
        p.twojmax, p.nspecies = twojmax, nspecies
        p.rcutfac, p.rmin0, p.rfac0 = rcutfac, rmin0, rfac0
        p.radii = self.radii.ctypes.data_as(C.POINTER(C.c_double))
        p.weight = self.weight.ctypes.data_as(C.POINTER(C.c_double))
        p.chem_flag, p.bzero_flag, p.bnorm_flag = int(chem_flag), int(bzero_flag), int(bnorm_flag)
        p.switch_flag, p.wselfall_flag = int(switch_flag), int(wselfall_flag)
        self.p = p
        self.t = lib().orc_snap_tables_build(twojmax, nspecies, int(chem_flag), int(bzero_flag), int(bnorm_flag))
        self.ncoeff = lib().orc_snap_num_coeffs(twojmax, nspecies, int(chem_flag))
        self.nspecies = nspecies
        T = self.t.contents
        self.idxu_max, self.idxz_max, self.idxb_max, self.idxcg_max = T.idxu_max, T.idxz_max, T.idxb_max, T.idxcg_max

    def __del__(self):
        try:
            lib().orc_snap_tables_free(self.t)
        except Exception:
            pass

    # table views -----------------------------------------------------
    def cglist(self):
        return np.ctypeslib.as_array(self.t.contents.cglist, shape=(self.idxcg_max,)).copy()

    def idxb(self):
        return np.ctypeslib.as_array(self.t.contents.idxb, shape=(self.idxb_max, 3)).copy()

    def idxu_block(self):
        J = self.p.twojmax
        return np.ctypeslib.as_array(self.t.contents.idxu_block, shape=(J + 1,)).copy()

    def idxcg_block(self):
        J = self.p.twojmax
        return np.ctypeslib.as_array(self.t.contents.idxcg_block, shape=((J + 1) ** 3,)).copy().reshape(J + 1, J + 1, J + 1)

    def idxz(self):
        z = self.t.contents.idxz
        return np.array([[getattr(z[k], f) for f, _ in ZIdx._fields_] for k in range(self.idxz_max)], dtype=np.int64)

    def rootpq(self):
        J = self.p.twojmax
        return np.ctypeslib.as_array(self.t.contents.rootpq, shape=((J + 2) * (J + 2),)).copy().reshape(J + 2, J + 2)

    # computations ----------------------------------------------------
    def neighborlist(self, pos, L, species, appear=None, mode=SNAP_LIST_CUTOFF):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        sp = np.ascontiguousarray(species, dtype=np.int32)
        ap = sp if appear is None else np.ascontiguousarray(appear, dtype=np.int32)
        off = np.zeros(n + 1, dtype=np.int64)
        Lb = _box(L)
        m = lib().orc_snap_neighborlist(_p(pos), n, _p(Lb), _p(sp), _p(ap), C.byref(self.p), mode, _p(off),
                                        None, None, None, 0)
        if m < 0:
            raise ValueError("invalid SNAP neighborlist input (%d)" % m)
        nbr = np.zeros(max(m, 1), dtype=np.int32)
        R = np.zeros(max(m, 1))
        r = np.zeros((max(m, 1), 3))
        lib().orc_snap_neighborlist(_p(pos), n, _p(Lb), _p(sp), _p(ap), C.byref(self.p), mode, _p(off),
                                    _p(nbr), _p(R), _p(r), m)
        return off, nbr[:m], R[:m], r[:m]

    def local_descriptors(self, species, off, nbr, rvec):
        n = len(species)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        B = np.zeros((n, self.ncoeff))
        rv = np.ascontiguousarray(rvec, dtype=np.float64)
        lib().orc_snap_local_descriptors(C.byref(self.p), self.t, n, _p(sp), _p(off), _p(nbr), _p(rv), _p(B))
        return B

    def all_descriptors(self, pos, species, off, nbr, rvec):
        n = len(species)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        M = self.ncoeff * self.nspecies
        B = np.zeros((n, self.ncoeff))
        dB = np.zeros((n, M, 3))
        W = np.zeros((M, 6))
        rv = np.ascontiguousarray(rvec, dtype=np.float64)
        lib().orc_snap_all_descriptors(C.byref(self.p), self.t, n, _p(pos), _p(sp), _p(off), _p(nbr), _p(rv),
                                       _p(B), _p(dB), _p(W))
        return B, dB, W

    def plain_energy_forces(self, species, off, nbr, rvec, beta, nthreads=1, count=None):
        """Reference-route E, F (B and per-pair dB contracted with beta) over
        the first `count` atoms (all by default); nthreads > 1 uses OpenMP.
        Returns (energy, forces, seconds)."""
        n = len(species)
        count = n if count is None else int(count)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        beta = np.ascontiguousarray(beta, dtype=np.float64).reshape(-1)
        if beta.size == self.ncoeff and self.nspecies > 1:
            beta = np.tile(beta, self.nspecies)
        e = np.zeros(1)
        F = np.zeros((n, 3))
        rv = np.ascontiguousarray(rvec, dtype=np.float64)
        t0 = lib().orc_wtime()
        lib().orc_snap_plain_energy_forces(C.byref(self.p), self.t, n, count, _p(sp), _p(off), _p(nbr), _p(rv),
                                           _p(beta), int(nthreads), _p(e), _p(F))
        return e[0], F, lib().orc_wtime() - t0

    def adjoint_energy_forces(self, species, off, nbr, rvec, beta, beta0=0.0, want_y=False, want_dedr=False):
        n = len(species)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        beta = np.ascontiguousarray(beta, dtype=np.float64).reshape(-1)
        if beta.size == self.ncoeff and self.nspecies > 1:
            beta = np.tile(beta, self.nspecies)
        e = np.zeros(1)
        F = np.zeros((n, 3))
        Y = np.zeros((n, self.idxu_max, 2)) if want_y else None
        D = np.zeros((len(nbr), 3)) if want_dedr else None
        rv = np.ascontiguousarray(rvec, dtype=np.float64)
        rc = lib().orc_snap_adjoint_energy_forces(C.byref(self.p), self.t, n, _p(sp), _p(off), _p(nbr), _p(rv),
                                                  _p(beta), float(beta0), _p(e), _p(F), _p(Y), _p(D))
        if rc != 0:
            raise ValueError("adjoint unsupported for these flags")
        out = {"energy": e[0], "forces": F}
        if want_y:
            out["Y"] = Y
        if want_dedr:
            out["dEdr"] = D
        return out

    def atom_uzb(self, rij, wj, rcutij, ielem=1):
        rij = np.ascontiguousarray(rij, dtype=np.float64).reshape(-1, 3)
        nnb = rij.shape[0]
        wj = np.ascontiguousarray(wj, dtype=np.float64)
        rc = np.ascontiguousarray(rcutij, dtype=np.float64)
        U = np.zeros((self.idxu_max, 2))
        Z = np.zeros((self.idxz_max, 2))
        B = np.zeros(self.idxb_max)
        lib().orc_snap_atom_uzb(C.byref(self.p), self.t, ielem, nnb, _p(rij), _p(wj), _p(rc), _p(U), _p(Z), _p(B))
        return U[:, 0] + 1j * U[:, 1], Z[:, 0] + 1j * Z[:, 1], B

    def uarray(self, r3, rcut):
        r3 = np.ascontiguousarray(r3, dtype=np.float64)
        out = np.zeros((self.idxu_max, 2))
        lib().orc_snap_uarray(self.t, self.p.rfac0, self.p.rmin0, float(rcut), _p(r3), _p(out))
        return out[:, 0] + 1j * out[:, 1]

    def duarray(self, r3, rcut, wj=1.0):
        r3 = np.ascontiguousarray(r3, dtype=np.float64)
        out = np.zeros((self.idxu_max, 3, 2))
        lib().orc_snap_duarray(C.byref(self.p), self.t, float(rcut), float(wj), _p(r3), _p(out))
        return out[..., 0] + 1j * out[..., 1]


def wtime():
    return lib().orc_wtime()
