"""GPU pair potentials vs the reference's known answers and the oracle.

Tolerance (FP64): 1e-12 relative on energy, forces and virial (forces and
virial relative to their largest component).  FP32 path: errors reported,
bounded loosely.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_xyz

pytestmark = pytest.mark.gpu

INF = np.inf
RTOL = 1e-12


def vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


def oterm(orc, p, table, coeff=1.0):
    mask = orc.ALL_SPECIES if p.species is None else sum(1 << k for k, s in enumerate(table) if s in p.species)
    return dict(kind=p.kind, coeff=coeff, rcut=p.rcutoff, p=list(p.params()), species_mask=mask)


def oracle_eval(orc, s, pot):
    from ipgpu.empirical import _terms_of

    table = s.unique_symbols()
    sp = np.array([table.index(x) for x in s.symbols], dtype=np.int32)
    L = [s.austrip_box()[d] if s.periodic[d] else INF for d in range(3)]
    terms = [oterm(orc, p, table, c) for p, c in _terms_of(pot)]
    return orc.pair_energy_force_virial(s.austrip_positions(), L, sp, terms, cells=True)


def close(a, b, rtol=RTOL):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rtol * scale, (np.abs(a - b).max(), scale)


def test_reference_driver_unit(ipg):
    v = vectors()["empirical_unit"]
    s = ipg.periodic_system(list(zip(v["symbols"], v["positions_bohr"])), v["box_bohr"])
    m = v["mock"]
    p = ipg.MockEmpiricalPotential(m["seed"], m["rcutoff"], m["species"])
    ef = ipg.energy_and_force(s, p)
    assert ef.e == v["e_hartree"]
    np.testing.assert_array_equal(ef.f, np.array(v["f_hartree_per_bohr"]))
    np.testing.assert_array_equal(ipg.virial_stress(s, p), np.array(v["virial_stress_hartree"]))


@pytest.mark.parametrize("name", ["lj_150", "lj_1000"])
def test_lj_cluster_known_answers(ipg, name):
    v = vectors()[name]
    X = golden_xyz(v["xyz"])
    s = ipg.AtomsSystem(X, [8.0, 8.0, 8.0], (False,) * 3, ["Ar"] * len(X), unit="angstrom")
    p = ipg.LennardJones(v["eps_eV"] * ipg.eV, v["sigma_A"] * ipg.angstrom, v["rcutoff_A"] * ipg.angstrom, ["Ar"])
    e, f = ipg.energy_and_force(s, p)
    assert abs(e / ipg.eV - v["e_eV"]) <= v["e_atol_eV"]
    fe = f / (ipg.eV / ipg.angstrom)
    if "f_atol_eV_per_A" in v:
        assert np.abs(fe).max() <= v["f_atol_eV_per_A"]
        assert abs(ipg.virial(s, p) / ipg.eV) <= v["virial_atol_eV"]
    assert np.abs(fe.sum(0)).max() <= v.get("sumf_atol_eV_per_A", 1e-10)


def test_lj_argon_fcc_vs_oracle(ipg, orc):
    from ipgpu import workloads

    s, p = workloads.lj_argon_fcc()
    e, f = ipg.energy_and_force(s, p)
    vir = ipg.virial_stress(s, p)
    eo, fo, vo = oracle_eval(orc, s, p)
    close(e, eo)
    close(f, fo)
    close(vir, vo)


def _mixed_small(seed=5):
    from ipgpu import workloads

    return workloads.mixed_pairs(n_side=(12, 12, 10), seed=seed)


def test_mixed_fused_fp64_vs_oracle(ipg, orc):
    s, pot = _mixed_small()
    e, f = ipg.energy_and_force(s, pot)
    vir = ipg.virial_stress(s, pot)
    eo, fo, vo = oracle_eval(orc, s, pot)
    close(e, eo)
    close(f, fo)
    close(vir, vo)
    # unfused (one evaluation per sub-potential, as the reference) agrees too
    e2, f2 = ipg.energy_and_force(s, pot, fused=False)
    close(e2, eo)
    close(f2, fo)


def test_mixed_fp32_error(ipg):
    s, pot = _mixed_small()
    e64, f64 = ipg.energy_and_force(s, pot)
    e32, f32 = ipg.energy_and_force(s, pot, precision="fp32")
    rel_e = abs(e32 - e64) / abs(e64)
    rel_f = np.abs(f32 - f64).max() / np.abs(f64).max()
    assert rel_e < 1e-5 and rel_f < 1e-4, (rel_e, rel_f)


def test_per_term_cutoffs_and_other_terms(ipg, orc):
    rng = np.random.default_rng(9)
    L = np.array([14.0, 13.0, 12.0])
    X = rng.random((300, 3)) * L
    syms = [["A", "B", "C"][k] for k in rng.integers(0, 3, 300)]
    s = ipg.AtomsSystem(X, L, (True, True, False), syms, unit="bohr")
    pot = (ipg.LennardJones(0.01, 1.2, 4.0, ["A", "B"]) + 0.5 * ipg.BornMayer(0.3, 0.4, 1.0, 0.8, 0.6, 3.0, ["C"])
           - ipg.Coulomb(0.3, -0.2, 5.0, ["A", "C"]) + 2.0 * ipg.Morse(0.02, 1.1, 1.5, 2.5, ["B"]))
    e, f = ipg.energy_and_force(s, pot)
    vir = ipg.virial_stress(s, pot)
    eo, fo, vo = oracle_eval(orc, s, pot)
    close(e, eo)
    close(f, fo)
    close(vir, vo)


def test_zbl_and_morse_as_is(ipg, orc):
    rng = np.random.default_rng(4)
    L = np.array([10.0, 10.0, 10.0])
    X = rng.random((150, 3)) * L
    s = ipg.AtomsSystem(X, L, (True,) * 3, ["B"] * 150, unit="bohr")
    for p in (ipg.ZBL(18, 54, 1.0, 3.5, ["B"]), ipg.ZBL(6, 8, 1.3, 3.5, ["B"]), ipg.Morse(0.1, 0.8, 2.0, 4.0, ["B"])):
        e, f = ipg.energy_and_force(s, p)
        eo, fo, _ = oracle_eval(orc, s, p)
        close(e, eo)
        close(f, fo)


def test_deterministic_repeat(ipg):
    s, pot = _mixed_small(7)
    a = ipg.energy_and_force(s, pot)
    b = ipg.energy_and_force(s, pot)
    assert a.e == b.e
    np.testing.assert_array_equal(a.f, b.f)


def test_full_size_mixed_properties(ipg):
    """256,000-atom mixed system: pairwise antisymmetry (sum F = 0),
    virial trace consistency and FP32 vs FP64 agreement."""
    from ipgpu import workloads

    s, pot = workloads.mixed_pairs()
    e, f = ipg.energy_and_force(s, pot)
    assert np.isfinite(e) and np.all(np.isfinite(f))
    assert np.abs(f.sum(0)).max() <= 1e-9 * np.abs(f).sum()
    e32, f32 = ipg.energy_and_force(s, pot, precision="fp32")
    assert abs(e32 - e) / abs(e) < 1e-5
    assert np.abs(f32 - f).max() / np.abs(f).max() < 1e-4


def test_dense_clump_overflows_neighbor_rows(ipg, orc):
    """A dense clump in a dilute periodic box: the clump's rows outgrow the
    row width estimated from the mean density, so the call falls back to
    the counted CSR list; energy, forces and virial still match the
    oracle."""
    rng = np.random.default_rng(21)
    L = 60.0
    dilute = rng.uniform(0, L, size=(600, 3))
    clump = rng.uniform(25.0, 35.0, size=(300, 3))  # ~270 neighbors each, mean density gives ~4
    X = np.concatenate([dilute, clump])
    s = ipg.AtomsSystem(X, [L, L, L], (True, True, True), ["Ar"] * len(X), unit="angstrom")
    p = ipg.LennardJones(0.0104 * ipg.eV, 2.0 * ipg.angstrom, 6.0 * ipg.angstrom, ["Ar"])
    e, f = ipg.energy_and_force(s, p)
    vir = ipg.virial_stress(s, p)
    eo, fo, vo = oracle_eval(orc, s, p)
    close(e, eo)
    close(f, fo)
    close(vir, vo)
