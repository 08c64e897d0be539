"""Molecular dynamics on the stealthy potential, stepped on the GPU.

Velocity Verlet (NVE), Andersen resets (count or Bernoulli) and Langevin
BAOAB in Cartesian coordinates (mass 1, k_B = 1), with the tracked density
waves, winding, snapshots and the rho trace updated on the device each step
(src/xla_cco/dynamics/__init__.py).  ``state`` holds the containers and
conventions, ``run`` the driver, ``integrators`` / ``observables`` /
``thermal`` the single-lane host forms, ``checkpoint`` atomic state files
and ``rng`` the counter-based draws shared by the host and the kernels.
"""

from xla_cco.dynamics.checkpoint import CheckpointError, load_state, read_meta, save_state
from xla_cco.dynamics.integrators import andersen_reset, baoab_step, verlet_step
from xla_cco.dynamics.observables import center_of_mass, com_correct, wave_densities, winding_update
from xla_cco.dynamics.rng import prng_key
from xla_cco.dynamics.run import (
    MDSession,
    begin_segment,
    init_state,
    make_force_fn,
    make_waves,
    md_chunk,
    select_dt,
    simulate,
)
from xla_cco.dynamics.state import MDResult, MDSettings, MDState, RhoTrace, Snapshots, Waves, lane, stack
from xla_cco.dynamics.thermal import (
    cartesian_hessian,
    harmonic_sample,
    kinetic_energy,
    kinetic_temperature,
    maxwell_boltzmann,
)

__all__ = [
    "MDResult", "MDSettings", "MDState", "RhoTrace", "Snapshots", "Waves",
    "andersen_reset", "baoab_step", "begin_segment", "center_of_mass",
    "com_correct", "harmonic_sample", "init_state", "kinetic_energy",
    "kinetic_temperature", "load_state", "make_force_fn", "make_waves",
    "maxwell_boltzmann", "md_chunk", "save_state", "select_dt", "simulate",
    "verlet_step", "wave_densities", "winding_update", "CheckpointError",
    "read_meta", "prng_key", "cartesian_hessian", "lane", "stack", "MDSession",
]
