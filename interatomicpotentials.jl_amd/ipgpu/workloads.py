"""Seeded synthetic systems for the benchmark configurations.

  lj_argon_fcc        FCC 5x5x10 cells, a = 5.26 A, 1000 Ar atoms, periodic,
                      jitter N(0, 0.05 A); LJ eps = 0.0104 eV, sigma = 3.4 A,
                      rcut = 8.5 A.
  mixed_pairs         simple cubic lattice at 0.084 atoms/A^3, jitter
                      N(0, 0.1 A), species A/B 50/50, rcut = 6.0 A;
                      LJ(A) + Morse(A,B) + ZBL(B) (256,000 atoms by default).
  snap_bcc            bcc supercell with jitter, one or two species.
  snap_batch          many periodic configurations of uneven atom counts.
"""
import numpy as np

from .system import AtomsSystem, angstrom, eV
from . import empirical as E


def fcc_positions(nx, ny, nz, a):
    basis = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return ((g[:, None, :] + basis[None]) * a).reshape(-1, 3), np.array([nx, ny, nz], float) * a


def bcc_positions(nx, ny, nz, a):
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    return ((g[:, None, :] + basis[None]) * a).reshape(-1, 3), np.array([nx, ny, nz], float) * a


def lj_argon_fcc(seed=1):
    rng = np.random.default_rng(seed)
    X, L = fcc_positions(5, 5, 10, 5.26)
    X = X + rng.normal(scale=0.05, size=X.shape)
    s = AtomsSystem(X, L, (True, True, True), ["Ar"] * len(X), unit="angstrom")
    p = E.LennardJones(0.0104 * eV, 3.4 * angstrom, 8.5 * angstrom, ("Ar",))
    return s, p


def mixed_pairs(n_side=(80, 80, 40), seed=2, density=0.084):
    rng = np.random.default_rng(seed)
    sp = (1.0 / density) ** (1.0 / 3.0)
    nx, ny, nz = n_side
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    X = g * sp + rng.normal(scale=0.1, size=g.shape)
    L = np.array([nx, ny, nz], float) * sp
    syms = np.where(rng.random(len(X)) < 0.5, "A", "B").tolist()
    s = AtomsSystem(X, L, (True, True, True), syms, unit="angstrom")
    rc = 6.0 * angstrom
    lj = E.LennardJones(0.0104 * eV, 2.0 * angstrom, rc, ("A",))
    mo = E.Morse(0.1 * eV, 1.5 / angstrom, 2.3 * angstrom, rc, ("A", "B"))
    zb = E.ZBL(18, 54, 1.0, rc, ("B",))
    return s, lj + mo + zb


def snap_bcc(ncell=10, a=3.32, jitter=0.03, nspecies=1, seed=3):
    rng = np.random.default_rng(seed)
    X, L = bcc_positions(ncell, ncell, ncell, a)
    X = X + rng.normal(scale=jitter, size=X.shape)
    if nspecies == 1:
        syms = ["W"] * len(X)
    else:
        syms = np.where(rng.random(len(X)) < 0.5, "W", "Mo").tolist()
    return AtomsSystem(X, L, (True, True, True), syms, unit="angstrom")


def snap_batch(nconfig=512, nmin=50, nmax=2000, rcut=5.0, seed=4):
    """Configurations with N_i ~ U{nmin..nmax}: random bcc/fcc lattices with
    a ~ U[3.0, 4.2] A and jitter ~ U[0.01, 0.1] A; cells are chosen so every
    box edge is >= 2 rcut, then N_i random sites are kept."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nconfig):
        n = int(rng.integers(nmin, nmax + 1))
        a = rng.uniform(3.0, 4.2)
        jit = rng.uniform(0.01, 0.1)
        fcc = rng.random() < 0.5
        per = 4 if fcc else 2
        m = max(int(np.ceil(2 * rcut / a)), int(np.ceil((n / per) ** (1 / 3))))
        X, L = (fcc_positions if fcc else bcc_positions)(m, m, m, a)
        keep = np.sort(rng.choice(len(X), size=min(n, len(X)), replace=False))
        X = X[keep] + rng.normal(scale=jit, size=(len(keep), 3))
        syms = np.where(rng.random(len(X)) < 0.5, "W", "Mo").tolist()
        out.append(AtomsSystem(X, L, (True, True, True), syms, unit="angstrom"))
    return out
