"""Empirical pair potentials on the GPU.

Mirrors src/empirical/{lj,morse,zbl,bm,coulomb}.jl (constructors with the
reference's austrip semantics), src/types/empirical_potential.jl:49-77
(energy_and_force, virial_stress, force, potential_energy) and
src/mixed/linear_combination_potential.jl (algebra and sums).  Numbers are
in atomic units (hartree, bohr); multiply by ipgpu.eV / ipgpu.angstrom to
pass eV / angstrom, as austrip would convert them.
"""
import collections
import ctypes as C

import numpy as np

from . import _lib
from ._lib import IPG_FP32, IPG_FP64, check
from .nnlist import make_csystem
from .system import symbol_codes

EnergyForce = collections.namedtuple("EnergyForce", ["e", "f"])

ALL_SPECIES = (1 << 64) - 1


class AbstractPotential:
    def __add__(self, other):
        return LinearCombinationPotential._of(self) + LinearCombinationPotential._of(other)

    def __sub__(self, other):
        return self + (-1.0) * other

    def __neg__(self):
        return (-1.0) * self

    def __pos__(self):
        return 1.0 * self

    def __rmul__(self, a):
        lc = LinearCombinationPotential._of(self)
        return LinearCombinationPotential(lc.potentials, [a * c for c in lc.coefficients])

    def __truediv__(self, a):
        return (1.0 / a) * self


class EmpiricalPotential(AbstractPotential):
    kind = None

    def __init__(self, rcutoff, species):
        self.rcutoff = float(rcutoff)
        self.species = None if species is None else tuple(str(s) for s in species)

    def get_rcutoff(self):
        return self.rcutoff

    def get_species(self):
        return self.species

    def params(self):
        raise NotImplementedError

    def _term(self, coeff, table):
        t = _lib.PairTerm()
        t.kind = self.kind
        t.coeff = float(coeff)
        t.rcut = self.rcutoff
        for k, v in enumerate(self.params()):
            t.p[k] = float(v)
        if self.species is None:
            t.species_mask = ALL_SPECIES
        else:
            m = 0
            for k, sym in enumerate(table):
                if sym in self.species:
                    m |= 1 << k
            t.species_mask = m
        return t


class LennardJones(EmpiricalPotential):
    """lj.jl:18-36; epsilon (hartree), sigma, rcutoff (bohr)."""

    kind = _lib.IPG_PAIR_LJ

    def __init__(self, eps, sigma, rcutoff, species):
        super().__init__(rcutoff, species)
        self.eps, self.sigma = float(eps), float(sigma)

    def params(self):
        return (self.eps, self.sigma)


class Morse(EmpiricalPotential):
    """morse.jl:18-39; alpha is not austripped by the reference (pass bohr^-1)."""

    kind = _lib.IPG_PAIR_MORSE

    def __init__(self, D, alpha, sigma, rcutoff, species):
        super().__init__(rcutoff, species)
        self.D, self.alpha, self.sigma = float(D), float(alpha), float(sigma)

    def params(self):
        return (self.D, self.alpha, self.sigma)


class ZBL(EmpiricalPotential):
    """zbl.jl:9-36; e is the charge parameter used as the exponential base."""

    kind = _lib.IPG_PAIR_ZBL

    def __init__(self, Z1, Z2, e, rcutoff, species):
        super().__init__(rcutoff, species)
        self.Z1, self.Z2, self.e = int(Z1), int(Z2), float(e)

    def params(self):
        return (float(self.Z1), float(self.Z2), self.e)


class BornMayer(EmpiricalPotential):
    """bm.jl:21-39."""

    kind = _lib.IPG_PAIR_BORN_MAYER

    def __init__(self, A, rho, sigma, Cc, D, rcutoff, species):
        super().__init__(rcutoff, species)
        self.A, self.rho, self.sigma, self.C, self.D = map(float, (A, rho, sigma, Cc, D))

    def params(self):
        return (self.A, self.rho, self.sigma, self.C, self.D)


class Coulomb(EmpiricalPotential):
    """coulomb.jl:18-37."""

    kind = _lib.IPG_PAIR_COULOMB

    def __init__(self, q1, q2, rcutoff, species):
        super().__init__(rcutoff, species)
        self.q1, self.q2 = float(q1), float(q2)

    def params(self):
        return (self.q1, self.q2)


class MockEmpiricalPotential(EmpiricalPotential):
    """test/mocks.jl:22-34: V = R^2 seed, F = R^2 seed r (for the driver tests)."""

    kind = _lib.IPG_PAIR_MOCK

    def __init__(self, seed, rcutoff, species):
        super().__init__(rcutoff, species)
        self.seed = float(seed)

    def params(self):
        return (self.seed,)


class LinearCombinationPotential(AbstractPotential):
    """linear_combination_potential.jl:7-39."""

    def __init__(self, potentials, coefficients):
        self.potentials = list(potentials)
        self.coefficients = [float(c) for c in coefficients]

    @staticmethod
    def _of(p):
        if isinstance(p, LinearCombinationPotential):
            return p
        return LinearCombinationPotential([p], [1.0])

    def __add__(self, other):
        o = LinearCombinationPotential._of(other)
        return LinearCombinationPotential(self.potentials + o.potentials, self.coefficients + o.coefficients)

    def __rmul__(self, a):
        return LinearCombinationPotential(self.potentials, [a * c for c in self.coefficients])

    def get_rcutoff(self):
        return max(p.get_rcutoff() for p in self.potentials)

    def get_species(self):
        return None

    def flat_terms(self, scale=1.0):
        out = []
        for p, c in zip(self.potentials, self.coefficients):
            if isinstance(p, LinearCombinationPotential):
                out.extend(p.flat_terms(scale * c))
            else:
                out.append((p, scale * c))
        return out


def _terms_of(p):
    if isinstance(p, LinearCombinationPotential):
        return p.flat_terms()
    return [(p, 1.0)]


def _eval(s, terms, precision):
    """One fused GPU evaluation of (potential, coefficient) terms."""
    L = _lib.load()
    ctx = _lib.default_context()
    table, sp = symbol_codes(s)
    if len(table) > 64:
        raise ValueError("more than 64 species")
    cs, keep = make_csystem(s.austrip_positions(), s.austrip_box(), s.periodic, sp)
    arr = (_lib.PairTerm * len(terms))()
    for k, (p, c) in enumerate(terms):
        arr[k] = p._term(c, table)
    n = len(s)
    e = C.c_double()
    v = np.zeros(6)
    f = np.zeros((n, 3))
    prec = IPG_FP32 if precision in ("fp32", IPG_FP32) else IPG_FP64
    check(L.ipg_pair_energy_forces(ctx, C.byref(cs), C.cast(arr, C.c_void_p), len(terms), prec, C.byref(e),
                                   f.ctypes.data, v.ctypes.data))
    return e.value, f, v


def _evaluate(s, p, precision="fp64", fused=True):
    terms = _terms_of(p)
    if fused or len(terms) == 1:
        e, f, v = _eval(s, terms, precision)
        return e, f, v
    # one evaluation per sub-potential, summed with the coefficients, as the
    # reference's LinearCombinationPotential does
    e = 0.0
    f = np.zeros((len(s), 3))
    v = np.zeros(6)
    for q, c in terms:
        ei, fi, vi = _eval(s, [(q, 1.0)], precision)
        e += c * ei
        f += c * fi
        v += c * vi
    return e, f, v


def energy_and_force(s, p, precision="fp64", fused=True):
    """(e [hartree], f [N x 3, hartree/bohr]) -- empirical_potential.jl:49-58."""
    e, f, _ = _evaluate(s, p, precision, fused)
    return EnergyForce(e, f)


def potential_energy(s, p, **kw):
    return energy_and_force(s, p, **kw).e


def force(s, p, **kw):
    return energy_and_force(s, p, **kw).f


def virial_stress(s, p, precision="fp64", fused=True):
    """[xx, yy, zz, zy, zx, yx] in hartree -- empirical_potential.jl:60-67."""
    return _evaluate(s, p, precision, fused)[2]


def virial(s, p, **kw):
    """sum(virial_stress[1:3]) -- abstract_potential.jl:10."""
    return float(np.sum(virial_stress(s, p, **kw)[:3]))


class VerletEvaluator:
    """Energy, forces and virial of one potential over a device-resident
    Verlet-skin list, for MD loops that move the atoms of one system: the
    list holds every pair within rcut + skin of the positions at its last
    build and is rebuilt on the GPU when an atom has moved more than skin/2
    (the reference instead rebuilds the list on every call,
    empirical_potential.jl:21).  skin in bohr."""

    def __init__(self, s, p, skin, precision="fp64"):
        from .device import DeviceArray, dev_system

        L = _lib.load()
        self._ctx = _lib.default_context()
        self.table, sp = symbol_codes(s)
        terms = _terms_of(p)
        self._arr = (_lib.PairTerm * len(terms))()
        for k, (q, c) in enumerate(terms):
            self._arr[k] = q._term(c, self.table)
        self._nterms = len(terms)
        self._prec = IPG_FP32 if precision in ("fp32", IPG_FP32) else IPG_FP64
        self.n = len(s)
        self.pos = DeviceArray.from_host(s.austrip_positions())
        self._sp = DeviceArray.from_host(sp)
        self._dsys = dev_system(self.pos, self._sp, self.n, s.austrip_box(), s.periodic)
        self._e = DeviceArray((1,), np.float64)
        self._f = DeviceArray((self.n, 3), np.float64)
        self._v = DeviceArray((6,), np.float64)
        h = C.c_void_p()
        check(L.ipg_nlist_build_skin_dev(self._ctx, C.byref(self._dsys), p.get_rcutoff(), float(skin), 1,
                                         C.byref(h)))
        self._h = h
        self.builds = 1

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _lib.load().ipg_nlist_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def npairs(self):
        n = C.c_int64()
        check(_lib.load().ipg_nlist_info(self._h, None, C.byref(n), None, None))
        return n.value

    def step(self, positions):
        """New positions (N x 3, bohr) -> (e, f, virial, rebuilt)."""
        L = _lib.load()
        self.pos.upload(np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3))
        r = C.c_int32()
        check(L.ipg_nlist_update_dev(self._h, C.byref(self._dsys), C.byref(r)))
        self.builds += r.value
        check(L.ipg_pair_eval_dev(self._ctx, self._h, C.byref(self._dsys), C.cast(self._arr, C.c_void_p),
                                  self._nterms, self._prec, C.c_void_p(self._e.ptr), C.c_void_p(self._f.ptr),
                                  C.c_void_p(self._v.ptr)))
        return float(self._e.download()[0]), self._f.download(), self._v.download(), bool(r.value)
