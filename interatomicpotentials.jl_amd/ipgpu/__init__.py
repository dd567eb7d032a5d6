"""ipgpu -- host mirror of InteratomicPotentials.jl's hot-path API over
libipgpu.so (HIP kernels for MI355X / gfx950).

Neighbor lists (src/nnlist.jl), empirical pair potentials (src/empirical,
src/types/empirical_potential.jl, src/mixed) and SNAP descriptors/potentials
(src/BasisSystems/SNAP, src/BasisPotentials).  Every compute call runs on the
GPU; there is no CPU fallback.
"""
from ._lib import IPGError, LIB_PATH, load
from .system import AtomsSystem, periodic_system, angstrom, bohr, eV, hartree, BOHR_IN_ANGSTROM, HARTREE_IN_EV
from .nnlist import NeighborList, neighborlist
from .empirical import (LennardJones, Morse, ZBL, BornMayer, Coulomb, MockEmpiricalPotential,
                        LinearCombinationPotential, energy_and_force, potential_energy, force, virial_stress,
                        virial)
from . import snap

__all__ = [
    "IPGError", "LIB_PATH", "load", "AtomsSystem", "periodic_system", "angstrom", "bohr", "eV", "hartree",
    "BOHR_IN_ANGSTROM", "HARTREE_IN_EV", "NeighborList", "neighborlist", "LennardJones", "Morse", "ZBL",
    "BornMayer", "Coulomb", "MockEmpiricalPotential", "LinearCombinationPotential", "energy_and_force",
    "potential_energy", "force", "virial_stress", "virial", "snap",
]
