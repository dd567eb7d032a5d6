"""Writes the golden fixtures of tests/golden from the reference's own test data.

Run from the repository root with the reference mounted at /root/reference:
    python tests/golden/make_golden.py

It copies the LJ cluster coordinate files the reference's integration tests
read (test/integration/lj_clusters/{150,1000}.xyz) and writes
reference_vectors.json with the known answers those tests and the unit tests
assert (values transcribed with their file:line).  Nothing of the reference
is executed.
"""
import json
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

VECTORS = {
    "nnlist_unit": {
        "source": "test/unit/nnlist.jl:2-28",
        "positions_bohr": [[0.5, 0.1, 0.3], [0.7, 2.5, -2.1], [1.1, -3.1, 7.3]],
        "box_bohr": [1.0, 1.0, 1.0],
        "periodic": [True, False, True],
        "rcutoff": 5.0,
        "j_1based": [[2, 3], [], []],
        "r": [[[-0.2, -2.4, 0.4], [0.4, 3.2, 0.0]], [], []],
    },
    "empirical_unit": {
        "source": "test/unit/types/empirical_potential.jl:2-25 with test/mocks.jl:22-34",
        "symbols": ["Ar", "Ar", "Ar", "H"],
        "positions_bohr": [[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.5, 0.5, 0.0]],
        "box_bohr": [1.0, 1.0, 1.0],
        "mock": {"seed": 1.0, "rcutoff": 0.5, "species": ["Ar"]},
        "potential_energy_r111": 3.0,
        "force_r111": [3.0, 3.0, 3.0],
        "e_hartree": 0.5,
        "f_hartree_per_bohr": [[-0.125, -0.125, 0.0], [0.125, 0.0, 0.0], [0.0, 0.125, 0.0], [0.0, 0.0, 0.0]],
        "virial_stress_hartree": [0.0625, 0.0625, 0.0, 0.0, 0.0, 0.0],
    },
    "lj_150": {
        "source": "test/integration/lj_clusters/lj_150.jl:1-30",
        "xyz": "lj_150.xyz",
        "eps_eV": 1.0, "sigma_A": 1.0, "rcutoff_A": 100.0,
        "e_eV": -893.310258, "e_atol_eV": 1e-5,
        "f_atol_eV_per_A": 1e-10, "virial_atol_eV": 1e-5,
    },
    "lj_1000": {
        "source": "test/integration/lj_clusters/lj_1000.jl:1-23",
        "xyz": "lj_1000.xyz",
        "eps_eV": 1.0, "sigma_A": 1.0, "rcutoff_A": 100.0,
        "e_eV": -7128.821829, "e_atol_eV": 1e-5,
        "sumf_atol_eV_per_A": 1e-3,
    },
}


def main():
    shutil.copyfile(os.path.join(REF, "test/integration/lj_clusters/150.xyz"), os.path.join(HERE, "lj_150.xyz"))
    shutil.copyfile(os.path.join(REF, "test/integration/lj_clusters/1000.xyz"), os.path.join(HERE, "lj_1000.xyz"))
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(VECTORS, f, indent=1)


if __name__ == "__main__":
    main()
