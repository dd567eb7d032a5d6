"""The reverse-mode Wigner sweep of the FP32 adjoint force kernel
(csrc/ipg_wigner_adj.h), checked on the host against the oracle.

tests/host/wigner_adj_check.cpp runs the sweep (in FP64 and FP32) for one
neighbor and random per-entry seeds s_e and returns
    dE/dr = dsfac r_hat Phi + sfac grad_r Phi,  Phi = 2 sum_half w(e) Re(conj(u_e) s_e),
plus the forward recursion's entries.  The oracle's compute_uarray and
compute_duarray restatements (compute_ui.jl:50-129, :186-347) give the same
quantities independently: u_e directly, and dE/dr as the half-sum of
Re(conj(du_e/dr) s_e).  A wrong sign, a dropped mirror adjoint or a wrong
rescaling factor changes dE/dr at O(1).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(REPO, "tests", "host", "wigner_adj_check.cpp")
CSRC = os.path.join(REPO, "interatomicpotentials.jl_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("wadj") / "wigner_adj_check")
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17", "-I" + CSRC, SRC, "-o", exe],
                   check=True, capture_output=True)
    return exe


def _ub(j):
    return j * (j + 1) * (2 * j + 1) // 6


def _left_half(J):
    for j in range(J + 1):
        for mb in range(j + 1):
            for ma in range(j + 1):
                if 2 * mb < j or (2 * mb == j and ma <= mb):
                    w = 1.0 if (2 * mb == j and ma == mb) else 2.0
                    yield _ub(j) + (j + 1) * mb + ma, w


@pytest.mark.parametrize("switch", [True, False])
def test_sweep_equals_oracle_du_contraction(checker, orc, switch):
    rng = np.random.default_rng(11 if switch else 12)
    rfac0, rmin0, rcut = 0.99363, 0.3, 4.7
    lines, cases = [], []
    for J in range(0, 9):
        o = orc.Snap(J, 1, 1.0, rmin0, rfac0, [1.0], [1.0], switch_flag=switch)
        for t in range(5):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            r3 = d * rng.uniform(rmin0 + 0.05, rcut - 0.05)
            if t == 0:
                r3 = np.array([0.0, 0.0, 1.7])  # on the z axis: b = 0
            seed = rng.normal(size=(o.idxu_max, 2))
            lines.append("%d %.17g %.17g %.17g %d %.17g %.17g %.17g " % (J, rfac0, rmin0, rcut, int(switch), *r3) +
                         " ".join("%.17g" % v for v in seed.ravel()))
            cases.append((J, o, r3, seed))
    out = subprocess.run([checker], input="\n".join(lines) + "\n", capture_output=True, text=True,
                         check=True).stdout.splitlines()
    assert len(out) == len(cases)
    for (J, o, r3, seed), ln in zip(cases, out):
        v = np.array(ln.split(), float)
        d64, d32, u = v[:3], v[3:6], v[6:].reshape(-1, 2)
        U = o.uarray(r3, rcut)
        dU = o.duarray(r3, rcut, 1.0)  # [jju][k], switching function included
        S = seed[:, 0] + 1j * seed[:, 1]
        exp = np.zeros(3)
        idx = []
        for e, w in _left_half(J):
            idx.append(e)
            exp += w * (dU[e].real * S[e].real + dU[e].imag * S[e].imag)
        idx = np.array(idx)
        scale = np.abs(exp).max()
        np.testing.assert_allclose(d64, exp, rtol=0, atol=1e-12 * scale, err_msg="J=%d r=%s" % (J, r3))
        np.testing.assert_allclose(d32, exp, rtol=0, atol=2e-5 * scale, err_msg="J=%d r=%s" % (J, r3))
        np.testing.assert_allclose(u[idx, 0] + 1j * u[idx, 1], U[idx], rtol=0, atol=1e-13)
