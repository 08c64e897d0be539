"""The package's top-level names and the MD API are the reference's (CPU only).

src/xla_cco/__init__.py and src/xla_cco/dynamics/__init__.py list the public
names a user of the reference imports; each must resolve here (extras allowed).
"""

from __future__ import annotations

import pytest

import xla_cco
import xla_cco.dynamics as dyn

TOP = [
    "AbstractEnergy", "Box", "DescentSettings", "KSet", "LBFGSSettings", "PackedConfig",
    "RelaxResult", "StealthyEnergy", "Trajectory", "adaptive_descent", "chi_from_count",
    "collective_density", "cubic_box", "enumerate_kvectors", "g2_from_histogram",
    "isotropic_structure_factor", "k_outer_from_ka", "ka_from_chi", "lbfgs",
    "make_stealthy_energy", "nearest_neighbor_distances", "pair_distance_histogram",
    "read_configpack", "relax", "require_float64", "stack_fractional", "structure_factor",
    "unit_ball_volume", "wrap_fractional", "write_configpack", "write_function_table",
]
DYNAMICS = [
    "MDResult", "MDSettings", "MDState", "RhoTrace", "Snapshots", "Waves",
    "andersen_reset", "baoab_step", "begin_segment", "center_of_mass",
    "com_correct", "harmonic_sample", "init_state", "kinetic_energy",
    "kinetic_temperature", "load_state", "make_force_fn", "make_waves",
    "maxwell_boltzmann", "md_chunk", "save_state", "select_dt", "simulate",
    "verlet_step", "wave_densities", "winding_update",
]


def test_top_level_names() -> None:
    missing = [n for n in TOP if not hasattr(xla_cco, n)]
    assert not missing, missing
    assert set(TOP) <= set(xla_cco.__all__)


def test_dynamics_names() -> None:
    missing = [n for n in DYNAMICS if not hasattr(dyn, n)]
    assert not missing, missing
    assert set(DYNAMICS) <= set(dyn.__all__)


def test_auto_chunk_size_and_cli_helpers() -> None:
    import numpy as np
    import pytest

    from xla_cco.cli.relax import select_device, wrap_fractional_batch
    from xla_cco.energy import auto_chunk_size

    assert auto_chunk_size(1000, 894) is None  # the (M, N) matrix fits 1 GiB
    assert auto_chunk_size(100_000, 20_000) == (1 << 30) // (100_000 * 96)
    assert auto_chunk_size(10, 10, budget_bytes=96 * 10 * 3) == 3
    w = wrap_fractional_batch(np.array([[[1.5, -0.25], [2.0, 0.999]]]))
    assert np.array_equal(w, [[[0.5, 0.75], [0.0, 0.999]]])
    with pytest.raises(ValueError):
        select_device("cpu")


def test_flow_precision_settings_are_validated_before_device_work() -> None:
    """precision='mixed' needs DOPRI5; bad names and a negative switch raise ValueError."""
    import numpy as np

    from xla_cco.minimize.flow import FlowSettings, _validate

    s0 = np.zeros((1, 4, 2))
    basis = np.eye(2)
    _validate(s0, basis, FlowSettings(precision="mixed"))
    for bad in (FlowSettings(precision="half"), FlowSettings(precision="mixed", method="rosenbrock"),
                FlowSettings(precision="mixed", precision_switch=-1.0)):
        with pytest.raises(ValueError):
            _validate(s0, basis, bad)
