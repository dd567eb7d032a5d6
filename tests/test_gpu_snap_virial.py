"""virial_stress / virial of the linear SNAP potential (basis_potentials.jl
:43-57, linear_potential.jl:37-49: W * beta with W of
compute_all_descriptors, snap.jl:267-281) through the C ABI.

Pins: the oracle's W (itself pinned by W.beta = sum x (x) F) contracted with
beta, within 1e-10 relative for the FP64 plain and adjoint routes and 1e-4
for the FP32 adjoint; and, for a cluster (no periodic images, so the raw
positions W uses are the pair vectors), central differences of the energy
under a homogeneous strain: row (a, b) of W * beta equals -dE/d eps_ab with
x -> (I + eps e_a e_b^T) x (h = 1e-6, tolerance 1e-6 relative)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ["W", "Mo"]
# Voigt rows [xx, yy, zz, zy, zx, yx] as (derivative component a, position component b)
ROWS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def setup(ipg, orc, periodic=True, J=6, seed=3):
    from ipgpu import workloads

    s = workloads.snap_bcc(ncell=3, a=3.16, jitter=0.05, nspecies=2, seed=seed)
    if not periodic:
        s = ipg.AtomsSystem(s.positions, s.box, (False,) * 3, s.symbols, unit="angstrom")
    g = ipg.snap.SNAP(len(s), J, NAMES, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], switch_flag=True,
                      list_mode="cutoff")
    o = orc.Snap(J, 2, 2.5, 0.0, 0.99363, [1.0, 1.0], [1.0, 0.8], switch_flag=True)
    sp = np.array([NAMES.index(x) for x in s.symbols], np.int32)
    L = [s.box[d] if s.periodic[d] else np.inf for d in range(3)]
    off, nbr, R, r = o.neighborlist(s.positions, L, sp, mode=orc.SNAP_LIST_CUTOFF)
    beta = np.random.default_rng(seed).normal(scale=0.1, size=2 * g.ncoeff)
    return s, g, o, sp, (off, nbr, r), beta


@pytest.mark.parametrize("periodic", [True, False])
def test_virial_stress_vs_oracle(ipg, orc, periodic):
    s, g, o, sp, (off, nbr, r), beta = setup(ipg, orc, periodic)
    _, _, Wo = o.all_descriptors(s.positions, sp, off, nbr, r)
    want = Wo.T @ beta
    lbp = ipg.snap.LBasisPotential(beta, [0.5], g)
    for path, tol in (("plain", 1e-10), ("adjoint", 1e-10), ("adjoint_fp32", 1e-4)):
        v = ipg.snap.virial_stress(s, lbp, path=path)
        assert rel(v, want) < tol, path
        assert ipg.snap.virial(s, lbp, path=path) == pytest.approx(want.sum(), rel=tol, abs=tol * np.abs(want).max())
    # one evaluation for energy, forces and virial equals the separate calls
    # (the forces bit for bit; the energy up to its summation order, which
    # the energy/force call takes in blocks of atoms)
    e, f = ipg.snap.energy_and_force(s, lbp, path="adjoint")
    efv = ipg.snap.energy_force_virial(s, lbp, path="adjoint")
    assert abs(efv.e - e) <= 1e-12 * abs(e) and np.array_equal(efv.f, f)
    assert rel(efv.v, want) < 1e-10


def test_virial_stress_strain_derivative(ipg, orc):
    s, g, o, sp, _, beta = setup(ipg, orc, periodic=False, seed=4)
    lbp = ipg.snap.LBasisPotential(beta, [0.0], g)
    v = ipg.snap.virial_stress(s, lbp, path="adjoint")
    h = 1e-6
    X0 = s.positions
    for row, (a, b) in enumerate(ROWS):
        E = []
        for sgn in (1.0, -1.0):
            D = np.eye(3)
            D[a, b] += sgn * h
            sx = ipg.AtomsSystem(X0 @ D.T, s.box, s.periodic, s.symbols, unit="angstrom")
            E.append(ipg.snap.energy_and_force(sx, lbp, path="adjoint").e)
        dE = (E[0] - E[1]) / (2 * h)
        assert v[row] == pytest.approx(-dE, rel=1e-6, abs=1e-6 * np.abs(v).max()), (row, v[row], -dE)
