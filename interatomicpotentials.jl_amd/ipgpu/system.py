"""Atomic systems and units (the parts of AtomsBase/Unitful the hot path reads).

The reference takes an AtomsBase AbstractSystem and strips units in two ways
(src/nnlist.jl:25 `austrip` for empirical potentials, utl.jl:60 `ustrip`
for SNAP).  An AtomsSystem stores raw coordinates plus their length unit so
both conversions are exact restatements:
  - austrip: coordinates in bohr (src/unit_convention.jl:6-10),
  - ustrip: the raw numbers in the system's own unit.
"""
import numpy as np

# CODATA 2018 values used by UnitfulAtomic.
BOHR_IN_ANGSTROM = 0.529177210903
HARTREE_IN_EV = 27.211386245988

# Multiply a quantity in the named unit by these to get atomic units.
angstrom = 1.0 / BOHR_IN_ANGSTROM
bohr = 1.0
eV = 1.0 / HARTREE_IN_EV
hartree = 1.0

_LENGTH = {"bohr": 1.0, "angstrom": angstrom, "Å": angstrom, "A": angstrom}


class AtomsSystem:
    """Orthorhombic system: positions (N, 3), box edges, periodicity, symbols.

    Mirrors the AtomsBase accessors the reference uses: position(s),
    bounding_box(s)[i][i], periodicity(s), atomic_symbol(s).
    """

    def __init__(self, positions, box, periodic, symbols, unit="bohr"):
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
        if unit not in _LENGTH:
            raise ValueError("unknown length unit %r" % unit)
        b = np.asarray(box, dtype=np.float64)
        if b.shape == (3, 3):
            off = b - np.diag(np.diag(b))
            if np.any(off != 0):
                raise ValueError("only orthorhombic boxes are supported (box must be diagonal)")
            b = np.diag(b).copy()
        if b.shape != (3,):
            raise ValueError("box must be 3 edge lengths or a 3x3 diagonal matrix")
        per = tuple(bool(x) for x in periodic)
        if len(per) != 3:
            raise ValueError("periodic must have 3 entries")
        syms = [str(s) for s in symbols]
        if len(syms) != pos.shape[0]:
            raise ValueError("one symbol per atom required")
        self.positions = pos
        self.box = b
        self.periodic = per
        self.symbols = syms
        self.unit = unit

    # symbols are kept as a tuple, so the per-atom species codes derived
    # from them can be formed once per system
    @property
    def symbols(self):
        return self._symbols

    @symbols.setter
    def symbols(self, v):
        self._symbols = tuple(str(x) for x in v)
        self._codes = None

    def symbol_codes(self):
        """(unique_symbols(), the index of every atom's symbol in that list
        as a read-only int32 array), formed on first use."""
        if self._codes is None:
            first = {}
            codes = np.fromiter((first.setdefault(x, len(first)) for x in self._symbols), dtype=np.int32,
                                count=len(self._symbols))
            codes.flags.writeable = False
            self._codes = (tuple(first), codes)
        return list(self._codes[0]), self._codes[1]

    def __len__(self):
        return self.positions.shape[0]

    # AtomsBase-like accessors
    def atomic_symbol(self):
        return list(self.symbols)

    def periodicity(self):
        return self.periodic

    def austrip_positions(self):
        return self.positions * _LENGTH[self.unit]

    def austrip_box(self):
        return self.box * _LENGTH[self.unit]

    def unique_symbols(self):
        """unique(atomic_symbol(s)) in first-appearance order."""
        seen = []
        for s in self.symbols:
            if s not in seen:
                seen.append(s)
        return seen


def periodic_system(atoms, box, unit="bohr"):
    """AtomsBase periodic_system([sym => pos, ...], box)."""
    syms = [a[0] for a in atoms]
    pos = [a[1] for a in atoms]
    return AtomsSystem(pos, box, (True, True, True), syms, unit)


def symbol_codes(s):
    """(unique symbols in first-appearance order, int32 index of every
    atom's symbol among them) of a system; cached by AtomsSystem."""
    f = getattr(s, "symbol_codes", None)
    if f is not None:
        return f()
    uniq = s.unique_symbols()
    lut = {x: i for i, x in enumerate(uniq)}
    return uniq, np.array([lut[x] for x in s.symbols], dtype=np.int32)


def species_index(symbols, table):
    """Index of every atom's symbol in `table` (raises like findall(...)[1])."""
    lut = {s: i for i, s in enumerate(table)}
    out = np.empty(len(symbols), dtype=np.int32)
    for k, s in enumerate(symbols):
        if s not in lut:
            raise ValueError("species %r not in %r" % (s, list(table)))
        out[k] = lut[s]
    return out
