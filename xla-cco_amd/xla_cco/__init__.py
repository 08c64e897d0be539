"""xla_cco (MI355X): stealthy collective-coordinate energy, batched relax and
MD steps on gfx950 HIP kernels, with the host interface of sdawley1/xla-cco.

All arrays are float64.  The compute path is the HIP library
``_native/libcco_hip.so`` (C ABI in include/cco.h); there is no CPU fallback.
"""

from xla_cco._types import require_float64
from xla_cco.analysis.pairs import g2_from_histogram, nearest_neighbor_distances, pair_distance_histogram
from xla_cco.analysis.structure import isotropic_structure_factor, structure_factor
from xla_cco.box import Box, cubic_box, wrap_fractional
from xla_cco.energy import (
    AbstractEnergy,
    StealthyEnergy,
    collective_density,
    make_stealthy_energy,
)
from xla_cco.io.configpack import PackedConfig, read_configpack, stack_fractional, write_configpack
from xla_cco.io.tables import write_function_table
from xla_cco.kvectors import (
    KSet,
    chi_from_count,
    enumerate_kvectors,
    k_outer_from_ka,
    ka_from_chi,
    unit_ball_volume,
)
from xla_cco.minimize import relax, relax_uneven
from xla_cco.minimize.descent import DescentSettings, adaptive_descent
from xla_cco.minimize.lbfgs import LBFGSSettings, lbfgs
from xla_cco.minimize.result import RelaxResult, Trajectory

__all__ = [
    "require_float64",
    "Box",
    "cubic_box",
    "wrap_fractional",
    "AbstractEnergy",
    "StealthyEnergy",
    "collective_density",
    "make_stealthy_energy",
    "KSet",
    "chi_from_count",
    "enumerate_kvectors",
    "k_outer_from_ka",
    "ka_from_chi",
    "unit_ball_volume",
    "relax",
    "relax_uneven",
    "DescentSettings",
    "adaptive_descent",
    "LBFGSSettings",
    "lbfgs",
    "RelaxResult",
    "Trajectory",
    "PackedConfig",
    "read_configpack",
    "stack_fractional",
    "write_configpack",
    "write_function_table",
    "g2_from_histogram",
    "nearest_neighbor_distances",
    "pair_distance_histogram",
    "isotropic_structure_factor",
    "structure_factor",
]
