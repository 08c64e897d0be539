"""Molecular dynamics on the stealthy potential: NVE velocity Verlet on the GPU.

``integrators.verlet_step`` (one host-level step with any force function),
``run.make_force_fn`` (GPU forces at s = r inv(B)), ``run.init_state`` /
``run.simulate`` (batched lanes stepped on the device with the drift guard),
``thermal.maxwell_boltzmann`` / ``thermal.kinetic_energy`` / ``thermal.harmonic_sample``
(a canonical draw in the harmonic basin of a ground state, from the GPU Hessian).
"""

from xla_cco.dynamics.integrators import verlet_step
from xla_cco.dynamics.run import MDResult, MDSettings, MDState, init_state, make_force_fn, simulate
from xla_cco.dynamics.thermal import cartesian_hessian, harmonic_sample, kinetic_energy, maxwell_boltzmann

__all__ = [
    "verlet_step",
    "MDSettings",
    "MDState",
    "MDResult",
    "init_state",
    "make_force_fn",
    "simulate",
    "kinetic_energy",
    "maxwell_boltzmann",
    "cartesian_hessian",
    "harmonic_sample",
]
