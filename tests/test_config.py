"""Host-side problem description: units and workload presets."""
import ctypes
import math

from conftest import gcs_config


def test_unit_conversions_match_reference_normalisation():
    # 4.5846 nT per code unit (cubedsphere_mi_dipole.c:70-76)
    assert abs(gcs_config.B_NORM_T - 4.5846e-9) < 1e-12
    assert abs(gcs_config.nt_to_code(-5.0) + 1.0906) < 1e-4
    assert abs(gcs_config.TIME_NORM_S - 63.71) < 1e-12
    # n = 5 /cc, T = 1e5 K proton pressure in code units
    assert abs(gcs_config.pressure_from_temperature(5.0, 1e5) - 0.41272) < 1e-4


def test_presets_have_the_configured_sizes():
    c0 = gcs_config.workload("config0")
    assert c0.active == (48, 24, 24) and c0.active_cells == 165888
    c1 = gcs_config.workload("config1")
    assert c1.active == (192, 96, 96) and c1.active_cells == 10616832
    assert c1.perturbation == 0.01 and c1.mi_enabled == 1
    c4 = gcs_config.workload("config4")
    assert c4.active_cells == 84934656


def test_log_radial_map_and_uniform_wind():
    c = gcs_config.workload("config0")
    assert math.isclose(c.radial_stretch, math.log(12.0))
    s = c.startup
    # wind pressure rho u^2 / (gamma M^2) reproduces n k T
    p = s.density * s.speed ** 2 / (s.gamma * s.mach ** 2)
    assert math.isclose(p, gcs_config.pressure_from_temperature(5.0, 1e5))
    assert s.front_x > c.outer_radius * 1.5  # front beyond the ghost shells


def test_problem_config_layout_is_stable():
    # 3 size_t + int + 3 double + startup + ... (include/gamera_cs_problem.h)
    assert ctypes.sizeof(gcs_config.ProblemConfig) == 488
