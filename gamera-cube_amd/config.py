"""Problem configuration for the six-face Earth step.

`ProblemConfig` is the ctypes mirror of `gcs_problem_config`
(include/gamera_cs_problem.h); the same struct is accepted by the HIP
library, by the CPU oracle and by the harness around the compiled reference,
so every implementation builds its problem from identical inputs.

Units are the reference's Earth code units
(cubedsphere/tests/test_cubedsphere_dipole_background.c:467-470,
cubedsphere/cubedsphere_mi_dipole.c:70-76): length 1 R_E = 6.371e6 m,
speed 100 km/s, mass density 1 proton/cm^3, so one code time is 63.71 s and
one code B is 4.5846 nT.

Frame: the reference's driver blows the wind along +X with its upstream
front at negative x (test_cubedsphere_dipole_background.c:285-291). A GSM
wind of Vx = -400 km/s is the same problem rotated by 180 degrees about Z
(x -> -x, y -> -y): the aligned dipole and an IMF Bz are unchanged by that
rotation, so the presets below use +4 code speed along +X.
"""
import ctypes
import math

MAX_WIND = 8
FACE_COUNT = 6

# Physical normalisation (reference values quoted above).
LENGTH_NORM_M = 6371000.0
SPEED_NORM_M_S = 100000.0
RHO_NORM_KG_M3 = 1.67262192369e-21
MU0 = 4.0e-7 * math.pi
B_NORM_T = SPEED_NORM_M_S * math.sqrt(MU0 * RHO_NORM_KG_M3)
TIME_NORM_S = LENGTH_NORM_M / SPEED_NORM_M_S
K_BOLTZMANN = 1.380649e-23
PROTON_MASS_KG = 1.67262192369e-27
# 0.31 G equatorial surface dipole in code units (driver :293-295).
EARTH_DIPOLE_MOMENT = -6761.730688424097


class Vec3(ctypes.Structure):
    _fields_ = [("value", ctypes.c_double * 3)]


class EarthStartup(ctypes.Structure):
    """gamera_cs_earth_startup (cubedsphere_earth_startup.h:10-24)."""
    _fields_ = [
        ("inner_radius", ctypes.c_double),
        ("density", ctypes.c_double),
        ("speed", ctypes.c_double),
        ("mach", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("ambient_density", ctypes.c_double),
        ("ambient_pressure", ctypes.c_double),
        ("front_x", ctypes.c_double),
        ("front_width", ctypes.c_double),
        ("cut_radius", ctypes.c_double),
        ("cut_length", ctypes.c_double),
        ("imf", Vec3),
        ("dipole", Vec3),
    ]


class MiConfig(ctypes.Structure):
    """gamera_cs_mi_dipole_config (cubedsphere_mi_dipole.h:16-28)."""
    _fields_ = [
        ("azimuth_points", ctypes.c_size_t),
        ("colatitude_points", ctypes.c_size_t),
        ("current_shell", ctypes.c_size_t),
        ("ionosphere_radius_re", ctypes.c_double),
        ("pedersen_siemens", ctypes.c_double),
        ("coupling_interval_s", ctypes.c_double),
        ("length_norm_m", ctypes.c_double),
        ("speed_norm_m_s", ctypes.c_double),
        ("mass_density_norm_kg_m3", ctypes.c_double),
        ("maximum_iterations", ctypes.c_int),
        ("relative_tolerance", ctypes.c_double),
    ]


class ProblemConfig(ctypes.Structure):
    """gcs_problem_config (include/gamera_cs_problem.h)."""
    _fields_ = [
        ("radial_cells", ctypes.c_size_t),
        ("angular_cells", ctypes.c_size_t),
        ("ghost", ctypes.c_size_t),
        ("radial_map_version", ctypes.c_int),
        ("inner_radius", ctypes.c_double),
        ("outer_radius", ctypes.c_double),
        ("radial_stretch", ctypes.c_double),
        ("startup", EarthStartup),
        ("perturbation", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("use_boris", ctypes.c_int),
        ("strict_inner_gas_wall", ctypes.c_int),
        ("dt_fraction", ctypes.c_double),
        ("cfl", ctypes.c_double),
        ("wind_count", ctypes.c_size_t),
        ("wind_linear", ctypes.c_int),
        ("wind_time", ctypes.c_double * MAX_WIND),
        ("wind_bz", ctypes.c_double * MAX_WIND),
        ("wind_reference_x", ctypes.c_double),
        ("mi_enabled", ctypes.c_int),
        ("hall_siemens", ctypes.c_double),
        ("mi", MiConfig),
    ]

    @property
    def active(self):
        return (self.radial_cells, self.angular_cells, self.angular_cells)

    @property
    def extent(self):
        g = 2 * self.ghost
        return (self.radial_cells + g, self.angular_cells + g,
                self.angular_cells + g)

    @property
    def active_cells(self):
        return FACE_COUNT * self.radial_cells * self.angular_cells ** 2


def nt_to_code(b_nt):
    return b_nt * 1e-9 / B_NORM_T


def pressure_from_temperature(n_cm3, t_kelvin):
    """Proton thermal pressure n k_B T in code units."""
    pressure_pa = n_cm3 * 1e6 * K_BOLTZMANN * t_kelvin
    return pressure_pa / (RHO_NORM_KG_M3 * SPEED_NORM_M_S ** 2)


def earth_problem(radial_cells=48, angular_cells=24, *, n_cm3=5.0,
                  speed_km_s=400.0, t_kelvin=1e5, bz_nt=-5.0,
                  outer_radius=30.0, inner_radius=2.5, perturbation=0.0,
                  seed=0, mi=False, hall_siemens=0.0, pedersen_siemens=10.0,
                  mi_azimuth=256, mi_colatitude=128):
    """Six-face Earth problem on a log-spaced radial grid.

    The radial map is the reference's version-1 exponential with stretch
    ln(r_out / r_in), i.e. r = r_in (r_out / r_in)^f exactly log-spaced
    (nonorthogonal_radial_map.c:152-156). The initial state is a uniform
    wind of density n, speed v and proton temperature T: the startup front
    is placed beyond the outer boundary, so its tanh fraction is 1 on every
    cell (cubedsphere_earth_startup.c:21-23), plus the cut dipole of the
    driver (:284-295) and the IMF from the startup vector potential.
    """
    c = ProblemConfig()
    c.radial_cells = radial_cells
    c.angular_cells = angular_cells
    c.ghost = 4
    c.radial_map_version = 1
    c.inner_radius = inner_radius
    c.outer_radius = outer_radius
    c.radial_stretch = math.log(outer_radius / inner_radius)
    speed = speed_km_s * 1e3 / SPEED_NORM_M_S
    pressure = pressure_from_temperature(n_cm3, t_kelvin)
    gamma = 5.0 / 3.0
    s = c.startup
    s.inner_radius = inner_radius
    s.density = n_cm3
    s.speed = speed
    # wind pressure = rho u^2 / (gamma M^2)  (ref_harness / driver :459-465)
    s.mach = math.sqrt(n_cm3 * speed * speed / (gamma * pressure))
    s.gamma = gamma
    s.ambient_density = 1.0
    s.ambient_pressure = 0.001
    s.front_x = 100.0
    s.front_width = 1.5
    s.cut_radius = 16.0
    s.cut_length = 8.0
    s.imf.value[:] = (0.0, 0.0, nt_to_code(bz_nt))
    s.dipole.value[:] = (0.0, 0.0, EARTH_DIPOLE_MOMENT)
    c.perturbation = perturbation
    c.seed = seed
    c.use_boris = 1
    c.strict_inner_gas_wall = 1
    c.dt_fraction = 0.5
    c.cfl = 0.3
    c.wind_count = 1
    c.wind_linear = 1
    c.wind_reference_x = -outer_radius
    c.mi_enabled = 1 if mi else 0
    c.hall_siemens = hall_siemens
    m = c.mi
    m.azimuth_points = mi_azimuth
    m.colatitude_points = mi_colatitude
    m.current_shell = 1
    m.ionosphere_radius_re = 1.0
    m.pedersen_siemens = pedersen_siemens
    m.coupling_interval_s = 10.0
    m.length_norm_m = LENGTH_NORM_M
    m.speed_norm_m_s = SPEED_NORM_M_S
    m.mass_density_norm_kg_m3 = RHO_NORM_KG_M3
    m.maximum_iterations = 20000
    m.relative_tolerance = 1e-10
    return c


def bz_switch(c, t_switch, bz_after_nt):
    """A piecewise-constant wind Bz series: the startup IMF until code time
    t_switch, then bz_after_nt."""
    c.wind_count = 2
    c.wind_linear = 0
    c.wind_time[0] = 0.0
    c.wind_bz[0] = c.startup.imf.value[2]
    c.wind_time[1] = t_switch
    c.wind_bz[1] = nt_to_code(bz_after_nt)
    return c


def workload(name):
    """The named workloads: config0 .. config4 and small parity cases."""
    if name == "config0":      # 6 x 24^2 x 48, 200 steps, potential 0
        return earth_problem(48, 24)
    if name == "config1":      # 6 x 96^2 x 192, 1% perturbation, MI on
        return earth_problem(192, 96, perturbation=0.01, seed=1, mi=True,
                             hall_siemens=5.0)
    if name == "config1_nomi":
        return earth_problem(192, 96, perturbation=0.01, seed=1)
    if name == "config4":      # 6 x 192^2 x 384, Bz +5 nT (bz_switch -> -5)
        return earth_problem(384, 192, bz_nt=5.0, perturbation=0.01, seed=1, mi=True,
                             hall_siemens=5.0)
    if name == "tiny":         # 4 ghosts need n > 8 (cubedsphere_mesh.c:191)
        return earth_problem(8, 12)
    if name == "small":
        return earth_problem(16, 12, perturbation=0.01, seed=7)
    if name == "ragged":       # odd extents: every tile and block range partial
        return earth_problem(11, 13, perturbation=0.01, seed=3)
    if name == "ghost5":       # five ghost layers (the reference allows ghost >= 4)
        c = earth_problem(16, 12, perturbation=0.01, seed=7)
        c.ghost = 5
        return c
    if name == "fast_wind":    # 1600 km/s wind = 16 code speed >= 1.5 x the Boris
        # light speed 10: emergency interface diffusion fires on most faces
        return earth_problem(16, 12, perturbation=0.01, seed=7, speed_km_s=1600.0)
    if name == "small_mi":     # MI on a 32 x 12 polar grid (Pedersen only)
        return earth_problem(16, 12, perturbation=0.01, seed=7, mi=True,
                             mi_azimuth=32, mi_colatitude=12)
    if name == "small_mi_hall":
        return earth_problem(16, 12, perturbation=0.01, seed=7, mi=True,
                             hall_siemens=5.0, mi_azimuth=32, mi_colatitude=12)
    if name == "small_mi_net":  # a 64 x 32 polar grid the warm-start U-Net fits
        return earth_problem(16, 12, perturbation=0.01, seed=7, mi=True,
                             hall_siemens=5.0, mi_azimuth=64, mi_colatitude=32)
    if name == "small_mi_polar1":  # config 1's 2 x 128 x 256 polar grids, small MHD
        return earth_problem(16, 12, perturbation=0.01, seed=7, mi=True,
                             hall_siemens=5.0, mi_azimuth=256, mi_colatitude=128)
    raise KeyError(name)
