"""The C-ABI library loads and exports every entry point include/ipgpu.h
declares (no GPU calls), and the host-side logic of the Python mirror."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import REPO


def header_symbols():
    txt = open(os.path.join(REPO, "include", "ipgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ipg_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    import ipgpu

    L = ipgpu.load()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert sorted(ipgpu._lib.EXPORTED) == syms
    assert L.ipg_get_version() == 100


def test_status_codes_without_gpu_work():
    import ipgpu

    L = ipgpu.load()
    # NULL arguments are rejected before any device call
    assert L.ipg_ctx_create(0, None) == ipgpu._lib.IPG_ERR_INVALID
    assert b"NULL" in L.ipg_last_error()
    assert L.ipg_nlist_info(None, None, None, None, None) == ipgpu._lib.IPG_ERR_INVALID


def test_units_and_system():
    import ipgpu

    assert ipgpu.angstrom == pytest.approx(1.8897261246257702)
    assert ipgpu.eV == pytest.approx(1 / 27.211386245988)
    s = ipgpu.AtomsSystem([[1, 2, 3]], np.diag([4.0, 5, 6]), (1, 0, 1), ["X"], unit="angstrom")
    np.testing.assert_allclose(s.austrip_positions(), np.array([[1, 2, 3]]) * ipgpu.angstrom)
    np.testing.assert_allclose(s.box, [4, 5, 6])
    with pytest.raises(ValueError):
        ipgpu.AtomsSystem([[0, 0, 0]], [[1, 1, 0], [0, 1, 0], [0, 0, 1]], (1, 1, 1), ["X"])
    s2 = ipgpu.AtomsSystem(np.zeros((4, 3)), [1, 1, 1], (1, 1, 1), ["B", "A", "B", "C"])
    assert s2.unique_symbols() == ["B", "A", "C"]


def test_symbol_codes():
    """The per-atom species codes the wrappers pass to the library: each
    atom's index among the unique symbols in first-appearance order, formed
    once per system and redone when the symbols are replaced."""
    import ipgpu
    from ipgpu.system import species_index, symbol_codes

    s = ipgpu.AtomsSystem(np.zeros((6, 3)), [1, 1, 1], (1, 1, 1), ["W", "Mo", "Mo", "W", "Ta", "Mo"])
    uniq, codes = symbol_codes(s)
    assert uniq == s.unique_symbols() == ["W", "Mo", "Ta"]
    assert codes.dtype == np.int32 and codes.tolist() == [0, 1, 1, 0, 2, 1]
    assert symbol_codes(s)[1] is codes  # cached
    with pytest.raises(ValueError):
        codes[0] = 1  # read-only
    # species table order differs from first appearance: the table index per atom
    assert species_index(uniq, ["Mo", "Ta", "W"])[codes].tolist() == species_index(s.symbols, ["Mo", "Ta", "W"]).tolist()
    with pytest.raises(ValueError):
        species_index(uniq, ["W", "Mo"])  # Ta is not in the table
    s.symbols = ["Ta"] * 6
    assert symbol_codes(s)[0] == ["Ta"] and symbol_codes(s)[1].tolist() == [0] * 6

    class Duck:  # any system with symbols and unique_symbols()
        symbols = ["B", "A", "B"]

        def unique_symbols(self):
            return ["B", "A"]

    u2, c2 = symbol_codes(Duck())
    assert u2 == ["B", "A"] and c2.tolist() == [0, 1, 0]


def test_linear_combination_algebra_and_terms():
    import ipgpu

    lj = ipgpu.LennardJones(1.0, 2.0, 5.0, ["A"])
    mo = ipgpu.Morse(0.1, 0.5, 1.0, 4.0, ["A", "B"])
    p = 2.0 * lj - mo / 4.0 + lj
    terms = p.flat_terms()
    assert [c for _, c in terms] == [2.0, -0.25, 1.0]
    assert p.get_rcutoff() == 5.0 and p.get_species() is None
    t = lj._term(3.0, ["B", "A"])
    assert t.species_mask == 0b10 and t.coeff == 3.0 and t.kind == ipgpu._lib.IPG_PAIR_LJ
    assert ipgpu.ZBL(1, 2, 1.0, 3.0, None)._term(1.0, ["A"]).species_mask == (1 << 64) - 1


def test_snap_num_coeffs():
    from ipgpu.snap import get_num_snap_coeffs

    assert [get_num_snap_coeffs(J, 1, False) for J in (4, 6, 8)] == [14, 30, 55]
    assert get_num_snap_coeffs(3, 1, False) == 8
    assert get_num_snap_coeffs(4, 2, True) == 14 * 8


def test_default_context_returns_promptly():
    """Regression: creating the default context must not deadlock (it either
    returns a context or raises IPGError when no GPU is present)."""
    import subprocess
    import sys

    code = ("import sys; sys.path.insert(0, %r)\n"
            "import ipgpu\n"
            "try:\n    ipgpu._lib.default_context(); print('ctx')\n"
            "except ipgpu.IPGError as e:\n    print('err', e.code)\n") % os.path.join(REPO, "interatomicpotentials.jl_amd")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("ctx") or r.stdout.startswith("err -3")


def test_struct_layouts_match_the_header():
    """The ctypes mirrors (and the Julia structs, which use the same natural
    alignment) match the sizes and offsets the library asserts at compile
    time (static_assert in csrc/ipg_core.hip)."""
    from ipgpu import _lib

    want = {
        _lib.System: (64, {"natoms": 0, "positions": 8, "species": 16, "box": 24, "periodic": 48}),
        _lib.PairTerm: (80, {"kind": 0, "reserved": 4, "coeff": 8, "rcut": 16, "p": 24, "species_mask": 72}),
        _lib.SnapParams: (72, {"twojmax": 0, "nspecies": 4, "rcutfac": 8, "rmin0": 16, "rfac0": 24, "radii": 32,
                               "weight": 40, "chem_flag": 48, "bzero_flag": 52, "bnorm_flag": 56,
                               "switch_flag": 60, "wselfall_flag": 64}),
    }
    for cls, (size, offs) in want.items():
        assert C.sizeof(cls) == size, cls
        for name, off in offs.items():
            assert getattr(cls, name).offset == off, (cls, name)
    src = open(os.path.join(REPO, "interatomicpotentials.jl_amd", "csrc", "ipg_core.hip")).read()
    for cls, (size, offs) in want.items():
        cname = {"System": "ipg_system", "PairTerm": "ipg_pair_term", "SnapParams": "ipg_snap_params"}[cls.__name__]
        assert "sizeof(%s) == %d" % (cname, size) in src
        for name, off in offs.items():
            assert "offsetof(%s, %s) == %d" % (cname, name, off) in src
