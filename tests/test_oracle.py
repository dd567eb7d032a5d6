"""Pins the CPU oracle (no GPU): the reference's golden vectors and known
answers, plus closed forms, invariants, brute force and finite differences
chosen so that a dropped term, wrong sign or transposed index fails."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_xyz

INF = np.inf
A = 1.0 / 0.529177210903
EV = 1.0 / 27.211386245988


def vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


# ---------------------------------------------------------------- lists ----
def test_nnlist_golden(orc):
    v = vectors()["nnlist_unit"]
    L = [1.0, INF, 1.0]
    for cells in (False, True):
        off, j, d = orc.neighborlist(np.array(v["positions_bohr"]), L, v["rcutoff"], cells=cells)
        assert [list(j[off[i]:off[i + 1]] + 1) for i in range(3)] == v["j_1based"]
        np.testing.assert_allclose(d, np.array(v["r"][0]), atol=1e-14)


def test_jl_mod_and_min_image(orc):
    # Julia mod: result has the sign of the divisor
    assert orc.jl_mod(-0.2, 1.0) == pytest.approx(0.8)
    assert orc.jl_mod(7.0, 1.0) == 0.0 and np.copysign(1, orc.jl_mod(-7.0, 1.0)) == 1
    assert orc.jl_mod(2.5, 1.0) == 0.5
    # displacement at the exact half box picks d - L
    np.testing.assert_array_equal(orc.get_displacement([1, 1, 1], [0, 0, 0], [0.5, 0, 0]), [-0.5, 0, 0])
    np.testing.assert_array_equal(orc.get_displacement([1, INF, 1], [0, 0, 0], [0.25, 3, 0]), [-0.25, -3, 0])
    # PeriodicEuclidean vs brute force over images
    rng = np.random.default_rng(0)
    for _ in range(200):
        a, b = rng.normal(size=3) * 5, rng.normal(size=3) * 5
        L = np.array([2.0, 3.0, 1.5])
        ks = np.arange(-60, 61)
        best = np.sqrt(sum(np.min((a[d] - b[d] + ks * L[d]) ** 2) for d in range(3)))
        assert orc.pe_dist(a, b, L) == pytest.approx(best, rel=1e-12)


@pytest.mark.parametrize("per", [(True, True, True), (True, False, True), (False, False, False)])
def test_cells_equal_brute(orc, per):
    rng = np.random.default_rng(3)
    box = np.array([9.0, 11.0, 10.0])
    X = rng.random((400, 3)) * box
    X[::5] += rng.integers(-2, 3, size=X[::5].shape) * box
    L = [box[d] if per[d] else INF for d in range(3)]
    for kind in (orc.LIST_HALF, orc.LIST_FULL):
        a = orc.neighborlist(X, L, 2.7, kind=kind, cells=False)
        b = orc.neighborlist(X, L, 2.7, kind=kind, cells=True)
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y)


# ---------------------------------------------------------------- pairs ----
def test_empirical_driver_golden(orc):
    v = vectors()["empirical_unit"]
    X = np.array(v["positions_bohr"])
    sp = np.array([0, 0, 0, 1])
    t = dict(kind=orc.PAIR_MOCK, rcut=v["mock"]["rcutoff"], p=[v["mock"]["seed"]], species_mask=1)
    assert orc.pair_V(t, np.sqrt(3.0)) == pytest.approx(v["potential_energy_r111"])
    np.testing.assert_allclose(orc.pair_F(t, np.sqrt(3.0), [1, 1, 1]), v["force_r111"])
    e, f, vir = orc.pair_energy_force_virial(X, [1, 1, 1], sp, [t])
    assert e == v["e_hartree"]
    np.testing.assert_array_equal(f, v["f_hartree_per_bohr"])
    np.testing.assert_array_equal(vir, v["virial_stress_hartree"])


@pytest.mark.parametrize("name", ["lj_150", "lj_1000"])
def test_lj_clusters_known_answers(orc, name):
    v = vectors()[name]
    X = golden_xyz(v["xyz"]) * A
    t = dict(kind=orc.PAIR_LJ, rcut=v["rcutoff_A"] * A, p=[v["eps_eV"] * EV, v["sigma_A"] * A])
    e, f, vir = orc.pair_energy_force_virial(X, [INF] * 3, np.zeros(len(X), np.int32), [t], cells=True)
    assert abs(e / EV - v["e_eV"]) <= v["e_atol_eV"]
    fe = f / (EV / A)
    assert np.abs(fe.sum(0)).max() <= v.get("sumf_atol_eV_per_A", v.get("f_atol_eV_per_A"))
    if "f_atol_eV_per_A" in v:
        assert np.abs(fe).max() <= v["f_atol_eV_per_A"]
        assert abs(vir[:3].sum() / EV) <= v["virial_atol_eV"]


def _fd_force(orc, t, R):
    h = 1e-6 * R
    dV = (orc.pair_V(t, R + h) - orc.pair_V(t, R - h)) / (2 * h)
    F = orc.pair_F(t, R, [R, 0, 0])[0]
    return dV, F


def test_pair_closed_forms(orc):
    # LJ: V(sigma) = 0, minimum at 2^(1/6) sigma with V = -eps, F = -dV/dR (r = R e_x)
    t = dict(kind=orc.PAIR_LJ, rcut=10, p=[0.3, 1.7])
    assert abs(orc.pair_V(t, 1.7)) < 1e-15
    assert orc.pair_V(t, 2 ** (1 / 6) * 1.7) == pytest.approx(-0.3, rel=1e-14)
    for R in (1.5, 2.0, 3.3):
        dV, F = _fd_force(orc, t, R)
        assert F == pytest.approx(-dV, rel=1e-7)
    # Morse: the reference's force is +dV/dR (morse.jl:39, sign as-is)
    t = dict(kind=orc.PAIR_MORSE, rcut=10, p=[0.2, 1.3, 2.1])
    assert orc.pair_V(t, 2.1) == 0.0
    for R in (1.6, 2.5, 3.7):
        dV, F = _fd_force(orc, t, R)
        assert F == pytest.approx(+dV, rel=1e-7)
    # ZBL with e = 1: phi = sum(c) = 1, dphi = -sum(x c) (zbl.jl:27-36)
    t = dict(kind=orc.PAIR_ZBL, rcut=10, p=[18, 54, 1.0])
    assert orc.pair_V(t, 2.0) == pytest.approx(18 * 54 / 2.0, rel=1e-14)
    a = 0.8854 * 0.529 / (18 ** 0.23 + 54 ** 0.23)
    d = 2.0 / a
    dphi = -sum(x * c for x, c in zip([3.19980, 0.94229, 0.40290, 0.20162], [0.18175, 0.50986, 0.28022, 0.02817]))
    assert orc.pair_F(t, 2.0, [2.0, 0, 0])[0] == pytest.approx(18 * 54 * (1 + d * dphi) / 8.0 * 2.0, rel=1e-13)
    # Coulomb: F = -dV/dR
    t = dict(kind=orc.PAIR_COULOMB, rcut=10, p=[0.5, -0.7])
    dV, F = _fd_force(orc, t, 1.9)
    assert F == pytest.approx(-dV, rel=1e-7)


def test_species_filter_and_linear_combination(orc):
    rng = np.random.default_rng(1)
    X = rng.random((60, 3)) * 8
    sp = rng.integers(0, 2, 60).astype(np.int32)
    t_all = dict(kind=orc.PAIR_LJ, rcut=3.0, p=[0.1, 1.0])
    t_a = dict(kind=orc.PAIR_LJ, rcut=3.0, p=[0.1, 1.0], species_mask=1)
    t_b = dict(kind=orc.PAIR_LJ, rcut=3.0, p=[0.1, 1.0], species_mask=2)
    t_ab = dict(kind=orc.PAIR_LJ, rcut=3.0, p=[0.1, 1.0], species_mask=3)
    L = [8, 8, 8]
    e_all, f_all, _ = orc.pair_energy_force_virial(X, L, sp, [t_all])
    e_ab, f_ab, _ = orc.pair_energy_force_virial(X, L, sp, [t_ab])
    assert e_all == e_ab
    e_a, _, _ = orc.pair_energy_force_virial(X, L, sp, [t_a])
    e_b, _, _ = orc.pair_energy_force_virial(X, L, sp, [t_b])
    # brute force: a pair counts iff both atoms are in the species set
    off, j, d = orc.neighborlist(X, L, 3.0)
    ea = sum(orc.pair_V(t_a, np.linalg.norm(d[q])) for i in range(60) for q in range(off[i], off[i + 1])
             if sp[i] == 0 and sp[j[q]] == 0)
    assert e_a == pytest.approx(ea, rel=1e-13)
    e2, f2, _ = orc.pair_energy_force_virial(X, L, sp, [dict(t_a, coeff=2.0), dict(t_b, coeff=-0.5)])
    assert e2 == pytest.approx(2 * e_a - 0.5 * e_b, rel=1e-13)


# ----------------------------------------------------------------- SNAP ----
def test_snap_table_sizes(orc):
    # (idxcg_max, idxu_max, idxz_max, idxb_max, sum na*nb)
    want = {4: (305, 55, 185, 14, 1206), 6: (1321, 140, 779, 30, 8836), 8: (4098, 285, 2386, 55, 41889)}
    for J, w in want.items():
        s = orc.Snap(J, 1, 1.0, 0.0, 0.99, [1.0], [1.0])
        z = s.idxz()
        assert (s.idxcg_max, s.idxu_max, s.idxz_max, s.idxb_max, int((z[:, 7] * z[:, 8]).sum())) == w
        # ncoeff = get_num_snap_coeffs (utl.jl:19-37)
        assert s.ncoeff == s.idxb_max


def test_clebsch_gordan_exact(orc):
    from sympy import S
    from sympy.physics.wigner import clebsch_gordan

    J = 6
    s = orc.Snap(J, 1, 1.0, 0.0, 0.99, [1.0], [1.0])
    cg, blk = s.cglist(), s.idxcg_block()
    err = 0.0
    for j1 in range(J + 1):
        for j2 in range(j1 + 1):
            for j in range(j1 - j2, min(J, j1 + j2) + 1, 2):
                for m1 in range(j1 + 1):
                    for m2 in range(j2 + 1):
                        a, b = 2 * m1 - j1, 2 * m2 - j2
                        ex = 0.0 if abs(a + b) > j else float(
                            clebsch_gordan(S(j1) / 2, S(j2) / 2, S(j) / 2, S(a) / 2, S(b) / 2, S(a + b) / 2))
                        err = max(err, abs(cg[blk[j1, j2, j] + m1 * (j2 + 1) + m2] - ex))
    assert err < 1e-14


def test_rootpq_closed_form(orc):
    s = orc.Snap(5, 1, 1.0, 0.0, 0.99, [1.0], [1.0])
    R = s.rootpq()
    for p in range(1, 6):
        for q in range(1, 6):
            assert R[p, q] == np.sqrt(p / q)


def wigner_D_closed(a, b, n):
    """Spin-n/2 Wigner matrix as the normalized symmetric power of the
    SU(2) matrix Q = [[conj a, -conj b], [b, a]] built from the Cayley-Klein
    pair: D[k, l] = [x^(n-k) y^k] (Q00 x + Q10 y)^(n-l) (Q01 x + Q11 y)^l
    * sqrt(C(n,l) / C(n,k)), rows = mb, columns = ma."""
    from math import comb

    Q = np.array([[np.conj(a), -np.conj(b)], [b, a]])
    D = np.zeros((n + 1, n + 1), complex)
    for l in range(n + 1):
        p = np.array([1.0 + 0j])
        for _ in range(n - l):
            p = np.convolve(p, [Q[0, 0], Q[1, 0]])
        for _ in range(l):
            p = np.convolve(p, [Q[0, 1], Q[1, 1]])
        for k in range(n + 1):
            D[k, l] = p[k] * np.sqrt(comb(n, l) / comb(n, k))
    return D


def test_wigner_u_unitary_and_closed_form(orc):
    J = 6
    s = orc.Snap(J, 1, 1.0, 0.0, 0.99363, [1.0], [1.0])
    ub = s.idxu_block()
    rng = np.random.default_rng(5)
    for _ in range(5):
        r3 = rng.normal(size=3)
        r3 *= rng.uniform(0.3, 1.9) / np.linalg.norm(r3)
        U = s.uarray(r3, 2.0)
        r = np.linalg.norm(r3)
        theta0 = r * 0.99363 * np.pi / 2.0
        z0 = r / np.tan(theta0)
        r0 = np.sqrt(r * r + z0 * z0)
        a = (z0 - 1j * r3[2]) / r0
        b = (r3[1] - 1j * r3[0]) / r0
        for j in range(J + 1):
            M = U[ub[j]:ub[j] + (j + 1) ** 2].reshape(j + 1, j + 1)
            assert np.abs(M @ M.conj().T - np.eye(j + 1)).max() < 1e-13
            Dc = wigner_D_closed(a, b, j)
            assert np.abs(M - Dc).max() < 1e-12, j


def brute_bispectrum(s_or, rij, wj, rcut, J):
    """B from the definition with full (j+1)^2 matrices and dense CG sums."""
    ub = s_or.idxu_block()
    Utot = {j: np.eye(j + 1, dtype=complex) for j in range(J + 1)}  # w_self = 1
    for r3, w in zip(rij, wj):
        U = s_or.uarray(r3, rcut)
        for j in range(J + 1):
            Utot[j] = Utot[j] + w * U[ub[j]:ub[j] + (j + 1) ** 2].reshape(j + 1, j + 1)
    from sympy import S
    from sympy.physics.wigner import clebsch_gordan

    def C(j1, m1, j2, m2, j, m):
        if abs(m1 + m2) > j or m1 + m2 != m:
            return 0.0
        return float(clebsch_gordan(S(j1) / 2, S(j2) / 2, S(j) / 2, S(m1) / 2, S(m2) / 2, S(m) / 2))

    out = []
    for (j1, j2, j) in s_or.idxb():
        tot = 0.0
        for mb in range(j + 1):
            for ma in range(j + 1):
                z = 0.0
                for mb1 in range(j1 + 1):
                    mb2 = mb - mb1 + (j1 + j2 - j) // 2
                    if not 0 <= mb2 <= j2:
                        continue
                    cb = C(j1, 2 * mb1 - j1, j2, 2 * mb2 - j2, j, 2 * mb - j)
                    if cb == 0:
                        continue
                    for ma1 in range(j1 + 1):
                        ma2 = ma - ma1 + (j1 + j2 - j) // 2
                        if not 0 <= ma2 <= j2:
                            continue
                        ca = C(j1, 2 * ma1 - j1, j2, 2 * ma2 - j2, j, 2 * ma - j)
                        z += cb * ca * Utot[j1][mb1, ma1] * Utot[j2][mb2, ma2]
                tot += (np.conj(Utot[j][mb, ma]) * z).real
        out.append(tot)
    return np.array(out)


def test_bispectrum_brute_force(orc):
    J = 4
    s = orc.Snap(J, 1, 1.2, 0.0, 0.989, [1.5], [1.0])
    rng = np.random.default_rng(2)
    rij = rng.normal(size=(5, 3))
    rij *= (rng.uniform(0.4, 3.0, size=(5, 1)) / np.linalg.norm(rij, axis=1, keepdims=True))
    wj = rng.uniform(0.5, 1.5, 5)
    _, _, B = s.atom_uzb(rij, wj, np.full(5, 3.6))
    Bb = brute_bispectrum(s, rij, wj, 3.6, J)
    assert np.abs(B - Bb).max() < 1e-10 * np.abs(Bb).max()


def test_bispectrum_invariances(orc):
    from scipy.spatial.transform import Rotation

    s = orc.Snap(6, 1, 2.0, 0.0, 0.99363, [1.0], [1.0])
    rng = np.random.default_rng(7)
    rij = rng.normal(size=(9, 3))
    rij *= (rng.uniform(0.5, 3.8, size=(9, 1)) / np.linalg.norm(rij, axis=1, keepdims=True))
    wj, rc = np.ones(9), np.full(9, 4.0)
    _, _, B0 = s.atom_uzb(rij, wj, rc)
    _, _, B1 = s.atom_uzb(rij @ Rotation.random(random_state=3).as_matrix().T, wj, rc)
    _, _, B2 = s.atom_uzb(rij[::-1].copy(), wj, rc)
    assert np.abs(B1 - B0).max() < 1e-12 * np.abs(B0).max()
    assert np.abs(B2 - B0).max() < 1e-12 * np.abs(B0).max()


def _bcc(orc, seed, switch, J=6):
    rng = np.random.default_rng(seed)
    b = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.array(list(np.ndindex(3, 3, 3)))
    X = ((g[:, None] + b[None]) * 3.16).reshape(-1, 3) + rng.normal(scale=0.05, size=(54, 3))
    L = np.full(3, 9.48)
    sp = rng.integers(0, 2, 54).astype(np.int32)
    s = orc.Snap(J, 2, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], switch_flag=switch)
    return s, X, L, sp


@pytest.mark.parametrize("switch", [False, True])
def test_dB_finite_differences(orc, switch):
    s, X, L, sp = _bcc(orc, 3, switch)

    def Btot(Xc):
        off, nb, R, r = s.neighborlist(Xc, L, sp, mode=orc.SNAP_LIST_CUTOFF)
        B = s.local_descriptors(sp, off, nb, r)
        return np.concatenate([B[sp == t].sum(0) for t in (0, 1)])

    off, nb, R, r = s.neighborlist(X, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    B, dB, W = s.all_descriptors(X, sp, off, nb, r)
    h = 1e-5
    for (a, k) in [(0, 0), (5, 1), (17, 2)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[a, k] += h
        Xm[a, k] -= h
        fd = -(Btot(Xp) - Btot(Xm)) / (2 * h)
        assert np.abs(fd - dB[a, :, k]).max() < 1e-7 * np.abs(dB[a, :, k]).max()


def test_virial_descriptors_identity(orc):
    """W rows are sum over pairs of dB (x) (x_i - x_j): with raw positions
    equal to the minimum-image ones (no wrap) W.beta = sum_i r_i (x) F_i."""
    s, X, L, sp = _bcc(orc, 4, True, J=4)
    Lbig = np.full(3, np.inf)  # non-periodic: raw and min-image positions agree
    off, nb, R, r = s.neighborlist(X, Lbig, sp, mode=orc.SNAP_LIST_CUTOFF)
    B, dB, W = s.all_descriptors(X, sp, off, nb, r)
    beta = np.random.default_rng(0).normal(size=2 * s.ncoeff)
    F = np.einsum("nmk,m->nk", dB, beta)
    Wb = W.T @ beta  # rows [xx, yy, zz, zy, zx, yx]
    want = [np.sum(X[:, 0] * F[:, 0]), np.sum(X[:, 1] * F[:, 1]), np.sum(X[:, 2] * F[:, 2]),
            np.sum(X[:, 2] * F[:, 1]), np.sum(X[:, 2] * F[:, 0]), np.sum(X[:, 1] * F[:, 0])]
    np.testing.assert_allclose(Wb, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())


@pytest.mark.parametrize("switch", [False, True])
def test_adjoint_equals_plain_and_fd_energy(orc, switch):
    s, X, L, sp = _bcc(orc, 5, switch)
    off, nb, R, r = s.neighborlist(X, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    beta = np.random.default_rng(1).normal(scale=0.1, size=2 * s.ncoeff)
    B, dB, W = s.all_descriptors(X, sp, off, nb, r)
    Fp = np.einsum("nmk,m->nk", dB, beta)
    ad = s.adjoint_energy_forces(sp, off, nb, r, beta, 0.25)
    assert np.abs(ad["forces"] - Fp).max() < 1e-12 * np.abs(Fp).max()
    Ep = 0.25 + sum(beta[sp[i] * s.ncoeff:(sp[i] + 1) * s.ncoeff] @ B[i] for i in range(len(X)))
    assert ad["energy"] == pytest.approx(Ep, rel=1e-13)
    h = 1e-5
    for (a, k) in [(1, 0), (30, 2)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[a, k] += h
        Xm[a, k] -= h
        Ef = []
        for Xc in (Xp, Xm):
            o2, n2, _, r2 = s.neighborlist(Xc, L, sp, mode=orc.SNAP_LIST_CUTOFF)
            Ef.append(s.adjoint_energy_forces(sp, o2, n2, r2, beta)["energy"])
        assert -(Ef[0] - Ef[1]) / (2 * h) == pytest.approx(ad["forces"][a, k], rel=1e-6, abs=1e-8)


@pytest.mark.parametrize("nthreads", [1, 3])
def test_plain_energy_forces_equals_descriptors(orc, nthreads):
    """The CPU-baseline routine (plain route, optionally OpenMP) against the
    descriptor route beta . dB and beta . B, over all atoms and over a
    leading subset of atoms."""
    s, X, L, sp = _bcc(orc, 5, True)
    off, nb, R, r = s.neighborlist(X, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    beta = np.random.default_rng(2).normal(scale=0.1, size=2 * s.ncoeff)
    B, dB, _ = s.all_descriptors(X, sp, off, nb, r)
    Fp = np.einsum("nmk,m->nk", dB, beta)
    Ep = sum(beta[sp[i] * s.ncoeff:(sp[i] + 1) * s.ncoeff] @ B[i] for i in range(len(X)))
    e, F, _ = s.plain_energy_forces(sp, off, nb, r, beta, nthreads=nthreads)
    assert e == pytest.approx(Ep, rel=1e-12)
    assert np.abs(F - Fp).max() < 1e-12 * np.abs(Fp).max()
    # leading subset: the pairs centred on the first k atoms, i.e. the
    # adjoint oracle on the same list with the rows of atoms k.. emptied
    k = 40
    e_k, F_k, _ = s.plain_energy_forces(sp, off, nb, r, beta, nthreads=nthreads, count=k)
    assert e_k == pytest.approx(sum(beta[sp[i] * s.ncoeff:(sp[i] + 1) * s.ncoeff] @ B[i] for i in range(k)),
                                rel=1e-12)
    off_k = np.minimum(off, off[k])
    ad = s.adjoint_energy_forces(sp, off_k, nb, r, beta)
    assert np.abs(F_k - ad["forces"]).max() < 1e-12 * np.abs(ad["forces"]).max()


# ------------------------------------------------- SNAP list (as-is) ----
def snap_list_fixture():
    with open(os.path.join(GOLDEN, "snap_list_hand.json")) as f:
        return json.load(f)


def first_appearance(symbols):
    seen = []
    for s in symbols:
        if s not in seen:
            seen.append(s)
    return np.array([seen.index(s) for s in symbols], dtype=np.int32)


@pytest.mark.parametrize("mode", ["reference", "cutoff"])
def test_snap_list_hand_derived(orc, mode):
    """utl.jl:38-101 on a 6-atom, two-species, mixed-boundary system whose
    lists were worked out by hand (tests/golden/snap_list_hand.json): the
    sqrt(cutmax) radius, the first-appearance element order and the
    enumerate-index element lookup each change at least one row."""
    v = snap_list_fixture()
    X = np.array(v["positions"])
    sp = np.array([v["snap_species"].index(s) for s in v["symbols"]], dtype=np.int32)
    ap = first_appearance(v["symbols"])
    L = [v["box"][d] if v["periodic"][d] else INF for d in range(3)]
    s = orc.Snap(2, 2, v["rcutfac"], 0.0, 0.99363, v["radii"], [1.0, 1.0])
    m = orc.SNAP_LIST_REFERENCE if mode == "reference" else orc.SNAP_LIST_CUTOFF
    off, j, R, r = s.neighborlist(X, L, sp, appear=ap, mode=m)
    want = v[mode]
    assert [list(j[off[i]:off[i + 1]] + 1) for i in range(len(X))] == want["j_1based"]
    rw = np.array([x for row in want["r"] for x in row]).reshape(-1, 3)
    np.testing.assert_array_equal(r, rw)
    np.testing.assert_array_equal(R, np.sqrt([x for row in want["rsq"] for x in row]))


# ----------------------------------------------------------- bzero ----
@pytest.mark.parametrize("J", [2, 5, 8])
def test_bzero_isolated_atom_is_zero(orc, J):
    """bzero (prebuilt_arrays.jl:52-63, compute_b.jl:70-93) is the bispectrum
    of an isolated atom: its U_j are identity matrices (w_self = 1 on the
    diagonal, compute_ui.jl:131-152), so Z_j = I by CG orthogonality and
    B(j1, j2, j) = tr I = j + 1.  With bzero_flag every entry of an isolated
    atom's B is therefore 0; a shift read at j1 or j2 instead of j, or a
    dropped (j+1), leaves entries nonzero.  Without the flag B = j + 1."""
    s0 = orc.Snap(J, 1, 2.0, 0.0, 0.99363, [1.0], [1.0])
    s1 = orc.Snap(J, 1, 2.0, 0.0, 0.99363, [1.0], [1.0], bzero_flag=True)
    n = 3
    off = np.zeros(n + 1, dtype=np.int64)
    sp = np.zeros(n, dtype=np.int32)
    nb, rv = np.zeros(0, np.int32), np.zeros((0, 3))
    assert np.abs(s1.local_descriptors(sp, off, nb, rv)).max() < 1e-13
    B0 = s0.local_descriptors(sp, off, nb, rv)
    np.testing.assert_allclose(B0, np.broadcast_to(s0.idxb()[:, 2] + 1.0, B0.shape), rtol=1e-14)


@pytest.mark.parametrize("J", [2, 5, 8])
@pytest.mark.parametrize("bzero", [False, True])
def test_chem_isolated_atoms_as_is(orc, J, bzero):
    """chem_flag = true (bnorm forced on) for isolated atoms of a two-element
    basis, worked out by hand from the as-is loops.  Only the own element's
    U is the identity (zero_uarraytot, compute_ui.jl:131-152).
    compute_zi (compute_z.jl:5-61) clears slot idouble*idxz_max + jjz but
    accumulates into slot jjz and divides it by j+1 after every element
    pair (:46-56), so only Z block 0 is ever nonzero: for element 0 it ends
    at I/(j+1)^4 (pairs (0,0), (0,1), (1,0), (1,1) each divide once), for
    element 1 at I/(j+1) (only pair (1,1) adds).  compute_bi (compute_b.jl:
    17-93) pairs Z block (e1, e2) with U of element e3, so the nonzero
    triple is (0, 0, e) with B = tr(Z) = 1/(j+1)^3 (element 0) or 1
    (element 1).  Its outer ielem loop recomputes every triple and then
    shifts triple (ielem-1)(nel+1)nel + (ielem-1), so only the last shift,
    triple (1, 1, 1) = 7, survives: B = -bzero[j] = -1 there (bnorm)."""
    s = orc.Snap(J, 2, 2.0, 0.0, 0.99363, [1.0, 1.0], [1.0, 1.0], chem_flag=True, bnorm_flag=True,
                 bzero_flag=bzero)
    nb_ = s.idxb_max
    assert s.ncoeff == 8 * nb_
    sp = np.array([0, 1], dtype=np.int32)
    B = s.local_descriptors(sp, np.zeros(3, dtype=np.int64), np.zeros(0, np.int32), np.zeros((0, 3)))
    j = s.idxb()[:, 2].astype(float)
    want = np.zeros((2, 8, nb_))
    want[0, 0] = 1.0 / (j + 1) ** 3
    want[1, 1] = 1.0
    if bzero:
        want[:, 7] -= 1.0
    np.testing.assert_allclose(B, want.reshape(2, -1), rtol=1e-13, atol=1e-15)


def test_bzero_shift_with_neighbors(orc):
    """With neighbors, B(bzero) + bzero[j] equals B without the flag
    entry by entry (the shift is a subtraction at the triple's j only)."""
    s0, X, L, sp = _bcc(orc, 6, True, J=8)
    s1 = orc.Snap(8, 2, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], switch_flag=True, bzero_flag=True)
    off, nb, R, r = s0.neighborlist(X, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    B0 = s0.local_descriptors(sp, off, nb, r)
    B1 = s1.local_descriptors(sp, off, nb, r)
    jb = s0.idxb()[:, 2]
    np.testing.assert_allclose(B1 + (jb + 1.0), B0, rtol=1e-13, atol=1e-13 * np.abs(B0).max())


# ------------------------------------------- Born-Mayer and ZBL e != 1 ----
def test_born_mayer_pins(orc):
    """bm.jl:38-39.  The exponential term's force is the gradient force
    (F.R/R = -dV/dR, checked by central differences with C = D = 0); the
    power terms keep the reference's non-gradient form F = V_pow / R^2 r,
    checked at a point where every value is exact: C = 2, D = 1, R = 2
    gives (C/R)^6 = 1, (D/R)^8 = 1/256."""
    t = dict(kind=orc.PAIR_BM, rcut=10, p=[0.7, 0.3, 1.1, 0.0, 0.0])
    for R in (0.9, 1.4, 2.6):
        dV, F = _fd_force(orc, t, R)
        assert F == pytest.approx(-dV, rel=1e-7)
    t = dict(kind=orc.PAIR_BM, rcut=10, p=[0.0, 0.3, 1.1, 2.0, 1.0])
    assert orc.pair_V(t, 2.0) == 1.0 - 1.0 / 256.0
    np.testing.assert_array_equal(orc.pair_F(t, 2.0, [2.0, 0.0, 0.0]), [(1.0 - 1.0 / 256.0) / 4.0 * 2.0, 0, 0])
    # both parts together: V and F add (a dropped term fails this)
    tA = dict(kind=orc.PAIR_BM, rcut=10, p=[0.7, 0.3, 1.1, 0.0, 0.0])
    tB = dict(kind=orc.PAIR_BM, rcut=10, p=[0.7, 0.3, 1.1, 2.0, 1.0])
    assert orc.pair_V(tB, 2.0) == pytest.approx(orc.pair_V(tA, 2.0) + 1.0 - 1.0 / 256.0, rel=1e-15)


ZBL_TEXTBOOK_C = (0.18175, 0.50986, 0.28022, 0.02817)
ZBL_TEXTBOOK_X = (3.19980, 0.94229, 0.40290, 0.20162)


def zbl_universal_screening(x):
    """The ZBL universal screening function (Ziegler, Biersack, Littmark,
    The Stopping and Range of Ions in Solids, 1985)."""
    return sum(c * np.exp(-k * x) for c, k in zip(ZBL_TEXTBOOK_C, ZBL_TEXTBOOK_X))


@pytest.mark.parametrize("e", [np.e, 2.0, 1.7])
def test_zbl_charge_base(orc, e):
    """zbl.jl:27-36 with e != 1: e^(-x d) = exp(-x d ln e), so
    phi(d, e) = Phi(d ln e) with Phi the textbook screening function, and
    dphi(d, e) = Phi'(d ln e); at e = Euler's number these are the
    textbook sums.  V = k Z1 Z2 e^2 Phi(d ln e) / R and the as-is force
    F.R/R = k Z1 Z2 e^2 (Phi + d Phi') / R^2, with Phi' from central
    differences of Phi (not from the oracle's dphi)."""
    Z1, Z2 = 18, 54
    t = dict(kind=orc.PAIR_ZBL, rcut=10, p=[Z1, Z2, e])
    a = 0.8854 * 0.529 / (Z1 ** 0.23 + Z2 ** 0.23)
    cf = Z1 * Z2 * e * e
    for R in (0.05, 0.3, 1.2):
        d = R / a
        x = d * np.log(e)
        assert orc.pair_V(t, R) == pytest.approx(cf * zbl_universal_screening(x) / R, rel=1e-12)
        h = 1e-6 * max(x, 1e-3)
        dPhi = (zbl_universal_screening(x + h) - zbl_universal_screening(x - h)) / (2 * h)
        F = orc.pair_F(t, R, [R, 0.0, 0.0])[0]  # r = R e_x, so F_x = F.r/R
        assert F == pytest.approx(cf * (zbl_universal_screening(x) + d * dPhi) / R ** 2, rel=1e-7)


def test_threaded_pair_path_equals_serial(orc):
    """The OpenMP pair path (the CPU baseline's) gives the serial path's
    list and, up to the per-thread summation order, its results."""
    rng = np.random.default_rng(8)
    X = rng.random((900, 3)) * 14.0
    sp = rng.integers(0, 2, 900).astype(np.int32)
    L = [14.0, 14.0, INF]
    terms = [dict(kind=orc.PAIR_LJ, rcut=3.0, p=[0.1, 1.0], species_mask=1),
             dict(kind=orc.PAIR_MORSE, rcut=2.5, p=[0.2, 1.3, 1.1], coeff=0.5)]
    e1, f1, v1 = orc.pair_energy_force_virial(X, L, sp, terms, cells=True)
    e0, f0, v0 = orc.pair_energy_force_virial(X, L, sp, terms, cells=False)
    for nt in (1, 3, 8):
        e, f, v = orc.pair_energy_force_virial(X, L, sp, terms, cells=True, nthreads=nt)
        if nt == 1:
            assert e == e1 and np.array_equal(f, f1) and np.array_equal(v, v1)
        assert e == pytest.approx(e0, rel=1e-12)
        np.testing.assert_allclose(f, f0, rtol=0, atol=1e-12 * np.abs(f0).max())
        np.testing.assert_allclose(v, v0, rtol=0, atol=1e-12 * np.abs(v0).max())


# ------------------------------------- chem_flag and 2J = 9, 10 pins ----
def _one_element(orc, seed, J, chem):
    rng = np.random.default_rng(seed)
    b = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.array(list(np.ndindex(3, 3, 3)))
    X = ((g[:, None] + b[None]) * 3.16).reshape(-1, 3) + rng.normal(scale=0.05, size=(54, 3))
    L = np.full(3, 9.48)
    s = orc.Snap(J, 1, 2.5, 0.0, 0.99363, [1.0], [1.0], chem_flag=chem, bnorm_flag=chem, switch_flag=True)
    return s, X, L, np.zeros(54, dtype=np.int32)


@pytest.mark.parametrize("J", [3, 6])
def test_chem_single_element_is_plain_over_jplus1(orc, J):
    """With one element, chem_flag = true (bnorm forced on) divides each Z
    block by its own j+1 once (compute_z.jl:53-56): B(j1, j2, j) is the plain
    B / (j+1).  dB/dr too: its three half-sums use the blocks (j1, j2, j),
    (j, j2, j1), (j, j1, j2), now divided by j+1, j1+1, j2+1, where the plain
    path multiplies the last two by (j+1)/(j1+1) and (j+1)/(j2+1)
    (compute_b.jl:268-270, 333-335).  W is linear in dB.  A j1 or j2 read
    for j in either factor fails this."""
    s0, X, L, sp = _one_element(orc, 9, J, False)
    s1, _, _, _ = _one_element(orc, 9, J, True)
    off, nb, R, r = s0.neighborlist(X, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    B0, dB0, W0 = s0.all_descriptors(X, sp, off, nb, r)
    B1, dB1, W1 = s1.all_descriptors(X, sp, off, nb, r)
    jf = s0.idxb()[:, 2] + 1.0
    np.testing.assert_allclose(B1, B0 / jf, rtol=1e-13, atol=0)
    np.testing.assert_allclose(dB1, dB0 / jf[None, :, None], rtol=0, atol=1e-13 * np.abs(dB0).max())
    np.testing.assert_allclose(W1, W0 / jf[:, None], rtol=0, atol=1e-13 * np.abs(W0).max())


class _CG:
    """Dense Clebsch-Gordan coefficients from sympy, cached (arguments doubled)."""

    def __init__(self):
        self.c = {}

    def __call__(self, j1, a, j2, b, j, m):
        k = (j1, a, j2, b, j, m)
        if k not in self.c:
            from sympy import S
            from sympy.physics.wigner import clebsch_gordan

            self.c[k] = 0.0 if (abs(a + b) > j or a + b != m) else float(
                clebsch_gordan(S(j1) / 2, S(j2) / 2, S(j) / 2, S(a) / 2, S(b) / 2, S(m) / 2))
        return self.c[k]


def _brute_Z(U1, U2, j1, j2, j, C):
    """Z(j1, j2, j)[mb, ma] from the definition, full (j+1) x (j+1)."""
    Z = np.zeros((j + 1, j + 1), complex)
    for mb in range(j + 1):
        for ma in range(j + 1):
            z = 0.0
            for mb1 in range(j1 + 1):
                mb2 = mb - mb1 + (j1 + j2 - j) // 2
                if not 0 <= mb2 <= j2:
                    continue
                cb = C(j1, 2 * mb1 - j1, j2, 2 * mb2 - j2, j, 2 * mb - j)
                if cb == 0:
                    continue
                for ma1 in range(j1 + 1):
                    ma2 = ma - ma1 + (j1 + j2 - j) // 2
                    if not 0 <= ma2 <= j2:
                        continue
                    z += cb * C(j1, 2 * ma1 - j1, j2, 2 * ma2 - j2, j, 2 * ma - j) * U1[j1][mb1, ma1] * U2[j2][mb2, ma2]
            Z[mb, ma] = z
    return Z


def test_chem_mixed_elements_brute_force(orc):
    """chem_flag = true with two elements on a 5-atom open cluster, from
    the definitions: per-element U sums (U from the closed-form-pinned
    uarray, wself on the center's element, compute_ui.jl:131-175), Z of
    every element pair by dense sympy CG sums, the as-is Z0 chain
    (compute_z.jl:46-56: every pair added to block 0, divided by j+1 after
    each), B of the triples (0, 0, e3) as full-matrix sums (twice the
    half-sum), all other triples zero; dB/dr from central differences of
    sfac w U of the pair contracted with the frozen Z0 at the three levels
    (compute_b.jl:145-337) into the triples (0, 0, e3), (e3, 0, 0),
    (0, e3, 0) of the neighbor's element e3, gathered with +/- as
    snap.jl:264-265.  Swapping two placements or dropping the chain's
    division fails it."""
    J, nel = 3, 2
    C = _CG()
    rng = np.random.default_rng(21)
    X = rng.uniform(-1.2, 1.2, size=(5, 3))
    L = np.array([INF, INF, INF])
    sp = np.array([0, 1, 1, 0, 1], dtype=np.int32)
    radii, weight, rfac, rcutfac = [1.1, 1.3], [1.0, 0.7], 0.99363, 1.4
    s = orc.Snap(J, nel, rcutfac, 0.0, rfac, radii, weight, chem_flag=True, bnorm_flag=True, switch_flag=True)
    off, nb, R, r = s.neighborlist(X, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    B, dB, W = s.all_descriptors(X, sp, off, nb, r)
    assert off[-1] > 10
    ub, idxb, nb_ = s.idxu_block(), s.idxb(), s.idxb_max
    ntr = nel ** 3

    def sfac(rr, rc):
        return 0.5 * (np.cos(rr * np.pi / rc) + 1.0)

    def U_of(r3, rc, w):
        u = s.uarray(r3, rc)
        rr = np.linalg.norm(r3)
        return {j: sfac(rr, rc) * w * u[ub[j]:ub[j] + (j + 1) ** 2].reshape(j + 1, j + 1) for j in range(J + 1)}

    n = len(sp)
    Bb = np.zeros((n, ntr, nb_))
    dBb = np.zeros((n, nel, ntr, nb_, 3))
    for i in range(n):
        si = sp[i]
        Ue = [{j: (np.eye(j + 1, dtype=complex) if e == si else np.zeros((j + 1, j + 1), complex))
               for j in range(J + 1)} for e in range(nel)]
        pairs = []
        for q in range(off[i], off[i + 1]):
            jn = nb[q]
            rc = (radii[si] + radii[sp[jn]]) * rcutfac
            u = U_of(r[q], rc, weight[sp[jn]])
            for j in range(J + 1):
                Ue[sp[jn]][j] = Ue[sp[jn]][j] + u[j]
            pairs.append((jn, r[q], rc))
        Z0 = {}

        def z0(j1, j2, j):
            if (j1, j2, j) not in Z0:
                z = np.zeros((j + 1, j + 1), complex)
                for e1 in range(nel):
                    for e2 in range(nel):
                        z = (z + _brute_Z(Ue[e1], Ue[e2], j1, j2, j, C)) / (j + 1)
                Z0[(j1, j2, j)] = z
            return Z0[(j1, j2, j)]

        for k, (j1, j2, j) in enumerate(idxb):
            for e3 in range(nel):
                Bb[i, e3, k] = (np.conj(Ue[e3][j]) * z0(j1, j2, j)).real.sum()
        h = 1e-6
        for (jn, r3, rc) in pairs:
            e3 = sp[jn]
            for kk in range(3):
                rp, rm = r3.copy(), r3.copy()
                rp[kk] += h
                rm[kk] -= h
                up, um = U_of(rp, rc, weight[e3]), U_of(rm, rc, weight[e3])
                du = {j: (up[j] - um[j]) / (2 * h) for j in range(J + 1)}
                for k, (j1, j2, j) in enumerate(idxb):
                    t1 = (np.conj(du[j]) * z0(j1, j2, j)).real.sum()
                    t2 = (np.conj(du[j1]) * z0(j, j2, j1)).real.sum()
                    t3 = (np.conj(du[j2]) * z0(j, j1, j2)).real.sum()
                    row = np.zeros(ntr)
                    row[e3] += t1
                    row[e3 * nel * nel] += t2
                    row[e3 * nel] += t3
                    dBb[i, si, :, k, kk] += row
                    dBb[jn, si, :, k, kk] -= row
    np.testing.assert_allclose(B, Bb.reshape(n, -1), rtol=0, atol=1e-12 * np.abs(Bb).max())
    dBr = dBb.reshape(n, nel * ntr * nb_, 3)
    assert np.abs(dBr).max() > 0
    np.testing.assert_allclose(dB, dBr, rtol=0, atol=1e-7 * np.abs(dBr).max())


def test_high_levels_tables_and_wigner(orc):
    """2J = 9, 10 (the largest twojmax the reference's Int64 Clebsch-Gordan
    factorials support): closed-form table sizes, exact CG on the top
    levels against sympy, and U against the closed-form Wigner matrices."""
    from sympy import S
    from sympy.physics.wigner import clebsch_gordan

    for J in (9, 10):
        s = orc.Snap(J, 1, 1.0, 0.0, 0.99363, [1.0], [1.0])
        assert s.idxu_max == sum((j + 1) ** 2 for j in range(J + 1))
        m = J // 2 + 1 if J % 2 == 0 else (J + 1) // 2
        assert s.ncoeff == s.idxb_max == (m * (m + 1) * (2 * m + 1) // 6 if J % 2 == 0 else m * (m + 1) * (m + 2) // 3)
    J = 10
    s = orc.Snap(J, 1, 1.0, 0.0, 0.99363, [1.0], [1.0])
    cg, blk = s.cglist(), s.idxcg_block()
    err = 0.0
    for (j1, j2, j) in [(10, 10, 10), (10, 9, 9), (10, 6, 6), (9, 8, 3), (10, 10, 0)]:
        for m1 in range(j1 + 1):
            for m2 in range(j2 + 1):
                a, b = 2 * m1 - j1, 2 * m2 - j2
                ex = 0.0 if abs(a + b) > j else float(
                    clebsch_gordan(S(j1) / 2, S(j2) / 2, S(j) / 2, S(a) / 2, S(b) / 2, S(a + b) / 2))
                err = max(err, abs(cg[blk[j1, j2, j] + m1 * (j2 + 1) + m2] - ex))
    assert err < 1e-13
    ub = s.idxu_block()
    rng = np.random.default_rng(15)
    for _ in range(3):
        r3 = rng.normal(size=3)
        r3 *= rng.uniform(0.3, 1.9) / np.linalg.norm(r3)
        U = s.uarray(r3, 2.0)
        r = np.linalg.norm(r3)
        z0 = r / np.tan(r * 0.99363 * np.pi / 2.0)
        r0 = np.sqrt(r * r + z0 * z0)
        a, b = (z0 - 1j * r3[2]) / r0, (r3[1] - 1j * r3[0]) / r0
        for j in (9, 10):
            M = U[ub[j]:ub[j] + (j + 1) ** 2].reshape(j + 1, j + 1)
            assert np.abs(M - wigner_D_closed(a, b, j)).max() < 1e-11, j


def test_high_levels_dB_finite_differences(orc):
    """dB/dr at 2J = 10 against central differences of sum B (as
    test_dB_finite_differences, at the top level the reference supports)."""
    s, X, L, sp = _bcc(orc, 5, True, J=10)

    def Btot(Xc):
        off, nb, R, r = s.neighborlist(Xc, L, sp, mode=orc.SNAP_LIST_CUTOFF)
        B = s.local_descriptors(sp, off, nb, r)
        return np.concatenate([B[sp == t].sum(0) for t in (0, 1)])

    off, nb, R, r = s.neighborlist(X, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    B, dB, W = s.all_descriptors(X, sp, off, nb, r)
    h = 1e-5
    for (a, k) in [(3, 0), (40, 2)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[a, k] += h
        Xm[a, k] -= h
        fd = -(Btot(Xp) - Btot(Xm)) / (2 * h)
        assert np.abs(fd - dB[a, :, k]).max() < 1e-6 * np.abs(dB[a, :, k]).max()
