/*
 * Data types of the GAMERA-Cube six-face step boundary.
 *
 * These structs are an ABI: a caller written against the reference library
 * builds them and hands them to gamera_cs_advance_six_with_background_planned
 * (and friends, see gamera_cs_step.h). Field order, types and meaning follow
 * the reference headers cited beside each type, so that the same caller
 * source compiles and links against this library unchanged.
 *
 * Index convention (nonorthogonal_grid.c:7-10): element (i, j, k) of an
 * extent {n0, n1, n2} lives at (i * n1 + j) * n2 + k, k contiguous.
 * i is radial, j and k are the two face angles (cubedsphere_mesh.c:198-221).
 */
#ifndef GAMERA_CS_TYPES_H
#define GAMERA_CS_TYPES_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GAMERA_NO_I = 0,
  GAMERA_NO_J = 1,
  GAMERA_NO_K = 2,
  GAMERA_NO_DIM = 3,
  GAMERA_NO_LOWER = 0,
  GAMERA_NO_UPPER = 1
};

/* Conserved-variable order (nonorthogonal_flux.h:12-19). */
enum {
  GAMERA_NO_FLUX_DENSITY = 0,
  GAMERA_NO_FLUX_MOMENTUM_X,
  GAMERA_NO_FLUX_MOMENTUM_Y,
  GAMERA_NO_FLUX_MOMENTUM_Z,
  GAMERA_NO_FLUX_ENERGY,
  GAMERA_NO_FLUX_COUNT
};

/* Cube faces and face edges (cubedsphere_mesh.h:13-29). */
enum {
  GAMERA_CS_POS_X = 0,
  GAMERA_CS_NEG_X,
  GAMERA_CS_POS_Y,
  GAMERA_CS_NEG_Y,
  GAMERA_CS_POS_Z,
  GAMERA_CS_NEG_Z,
  GAMERA_CS_FACE_COUNT
};
enum {
  GAMERA_CS_A_LOWER = 0,
  GAMERA_CS_A_UPPER,
  GAMERA_CS_B_LOWER,
  GAMERA_CS_B_UPPER,
  GAMERA_CS_EDGE_COUNT
};

/* nonorthogonal_geometry.h:17-42 */
typedef struct {
  double value[GAMERA_NO_DIM];
} gamera_no_vec3;

typedef struct {
  double area;
  gamera_no_vec3 area_vector;
  gamera_no_vec3 centroid;
  gamera_no_vec3 normal;
  gamera_no_vec3 tangent1;
  gamera_no_vec3 tangent2;
} gamera_no_face_geometry;

typedef struct {
  double volume;
  gamera_no_vec3 centroid;
  gamera_no_face_geometry face[GAMERA_NO_DIM][2];
  double cfl_length[GAMERA_NO_DIM];
} gamera_no_cell_geometry;

typedef struct {
  double length;
  gamera_no_vec3 normal;
  gamera_no_vec3 tangent1;
  gamera_no_vec3 tangent2;
} gamera_no_edge_geometry;

/* nonorthogonal_grid.h:12-30 */
typedef struct {
  size_t extent[3];
  gamera_no_face_geometry *value;
} gamera_no_face_grid;

typedef struct {
  size_t extent[3];
  gamera_no_edge_geometry *value;
  unsigned char *valid;
} gamera_no_edge_grid;

typedef struct {
  size_t cell_extent[3];
  size_t vertex_extent[3];
  gamera_no_vec3 *vertex;
  gamera_no_cell_geometry *cell;
  gamera_no_face_grid face[3];
  gamera_no_edge_grid edge[3];
} gamera_no_grid;

/* nonorthogonal_flux.h:21-25 */
typedef struct {
  double density;
  gamera_no_vec3 velocity;
  double pressure;
} gamera_no_primitive;

/* nonorthogonal_storage.h:12-47: AoS state of one face block. */
typedef struct {
  size_t cell_extent[3];
  size_t face_extent[3][3];
  size_t edge_extent[3][3];
  double *conserved;               /* [cell][5] */
  double *old_conserved;           /* [cell][5] */
  double *predicted_conserved;     /* [cell][5] */
  gamera_no_vec3 *cell_magnetic;   /* residual B1 per cell */
  gamera_no_vec3 *old_cell_magnetic;
  gamera_no_vec3 *predicted_cell_magnetic;
  double *face_flux[3];            /* residual magnetic face flux */
  double *old_face_flux[3];
  double *predicted_face_flux[3];
  double *edge_emf[3];             /* line-integrated edge EMF */
  double *fluid_face_flux[3];      /* [face][5], area scaled */
  gamera_no_vec3 *maxwell_face_flux[3];
  double *hydro_rate;              /* [cell][5] */
  gamera_no_vec3 *maxwell_rate;
  size_t emergency_diffusion_face_count;
  double emergency_diffusion_max_interface_speed;
  size_t inner_wall_clamped_face_count[2];
  double inner_wall_positive_mass_max[2];
  double inner_wall_positive_energy_max[2];
  double inner_wall_fluid_momentum_max[2];
  double inner_wall_maxwell_momentum_max[2];
  double inner_wall_density_max[2];
  double inner_wall_pressure_max[2];
  double inner_wall_pressure_gradient_max[2];
  double inner_wall_speed_max[2];
  double inner_wall_residual_magnetic_max[2];
  gamera_no_vec3 inner_wall_residual_magnetic_location[2];
  double inner_wall_lowlat_pressure_max[2];
  double inner_wall_lowlat_pressure_gradient_max[2];
  double inner_wall_lowlat_speed_max[2];
  double inner_wall_lowlat_residual_magnetic_max[2];
} gamera_no_storage;

/* nonorthogonal_sweep.h:13-56 */
typedef struct {
  double gamma;
  double density_floor;
  double pressure_floor;
  double pdm_coefficient;
  double hydro_hogs_coefficient;
  double magnetic_hogs_coefficient;
  double light_speed;
  bool use_mhd;
  bool use_hogs;
  bool use_boris;
  bool use_background;
} gamera_no_sweep_options;

typedef struct {
  const gamera_no_vec3 *cell_magnetic;
  const gamera_no_vec3 *face_magnetic[3];
  const double *face_flux[3];
  const gamera_no_vec3 *edge_magnetic[3];
  const gamera_no_vec3 *cell_force;
} gamera_no_background_field;

typedef struct {
  double density_floor;
  double pdm_coefficient;
  double diffusion_coefficient;
  double light_speed;
  double cfl;
  double dt;
  bool use_boris;
  bool use_background;
} gamera_no_emf_options;

/* nonorthogonal_step.h:59-78 */
typedef struct {
  double gamma;
  double density_floor;
  double pressure_floor;
  double light_speed;
  bool use_mhd;
  bool use_boris;
  bool use_background;
} gamera_no_update_options;

typedef struct {
  double gamma;
  double density_floor;
  double pressure_floor;
  double cfl;
  double light_speed;
  bool use_mhd;
  bool use_boris;
  bool use_background;
} gamera_no_timestep_options;

/* nonorthogonal_advance.h:10-51 */
typedef struct {
  gamera_no_sweep_options stress;
  gamera_no_emf_options emf;
  gamera_no_update_options update;
  int (*cell_source)(gamera_no_vec3 point,
                     const double predicted[GAMERA_NO_FLUX_COUNT],
                     double time, void *context,
                     double rate[GAMERA_NO_FLUX_COUNT]);
  void *source_context;
  int (*indexed_cell_source)(size_t cell, gamera_no_vec3 point,
                             const double predicted[GAMERA_NO_FLUX_COUNT],
                             double time, void *context,
                             double rate[GAMERA_NO_FLUX_COUNT]);
  void *indexed_source_context;
  double source_time;
  int (*fluid_flux_sync)(gamera_no_storage *storage,
                         const gamera_no_grid *grid,
                         const size_t active_lower[3],
                         const size_t active_upper[3], void *context);
  void *fluid_flux_context;
  int (*edge_emf_sync)(gamera_no_storage *storage, const gamera_no_grid *grid,
                       const size_t active_lower[3],
                       const size_t active_upper[3], void *context);
  void *edge_emf_context;
} gamera_no_advance_options;

/* cubedsphere_step.h:23-31 */
typedef int (*gamera_cs_radial_boundary_function)(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double current_time, double previous_time, void *context);

/* cubedsphere_halo.h:12: immutable donor plan, opaque to callers. */
typedef struct gamera_cs_halo_plan gamera_cs_halo_plan;

/* nonorthogonal_background.h:21-45 */
typedef struct {
  gamera_no_background_field field;
  gamera_no_vec3 *cell_magnetic;
  gamera_no_vec3 *face_magnetic[GAMERA_NO_DIM];
  double *face_flux[GAMERA_NO_DIM];
  gamera_no_vec3 *edge_magnetic[GAMERA_NO_DIM];
  gamera_no_vec3 *cell_force;
} gamera_no_background_data;

typedef struct {
  gamera_no_vec3 moment;
} gamera_no_dipole;

/* solar_wind.h:10-35 */
enum {
  GAMERA_WIND_X = 0,
  GAMERA_WIND_Y = 1,
  GAMERA_WIND_Z = 2,
  GAMERA_WIND_DIM = 3
};

typedef struct {
  double density;
  double pressure;
  double velocity[GAMERA_WIND_DIM];
  double magnetic[GAMERA_WIND_DIM];
} gamera_solar_wind_state;

typedef struct {
  size_t count;
  double *time;
  gamera_solar_wind_state *state;
  double reference[GAMERA_WIND_DIM];
  double time_offset;
  double by_coefficient;
  double bz_coefficient;
  double bx_offset;
  int enforce_bx_relation;
  int linear_interpolation;
} gamera_solar_wind_series;

/* nonorthogonal_radial_map.h:6-26 */
#define GAMERA_NO_RADIAL_MAP_LEGACY_EXPONENTIAL 1
#define GAMERA_NO_RADIAL_MAP_EARTH_PRODUCTION_V4 4
typedef struct {
  int version;
  double inner_radius;
  double outer_radius;
  double legacy_stretch;
  double inner_logical_knot;
  double outer_logical_knot;
  double inner_physical_knot;
  double outer_physical_knot;
  double middle_stretch;
  double outer_stretch;
} gamera_no_radial_map;

/* cubedsphere_earth_startup.h:10-24 */
typedef struct {
  double inner_radius;
  double density;
  double speed;
  double mach;
  double gamma;
  double ambient_density;
  double ambient_pressure;
  double front_x;
  double front_width;
  double cut_radius;
  double cut_length;
  gamera_no_vec3 imf;
  gamera_no_dipole dipole;
} gamera_cs_earth_startup;

/* NO/nonorthogonal_mi.h:11-14, :54-72: polar ionosphere solver. Arrays are
 * colatitude-major with longitude contiguous; row 0 is the pole. */
enum { GAMERA_MI_NORTH = -1, GAMERA_MI_SOUTH = 1 };
typedef struct {
  size_t longitude_count;
  size_t colatitude_count;
  double maximum_colatitude;
  double ionosphere_radius_m;
  double pedersen_siemens;
  double hall_siemens;
  double low_latitude_potential_v;
  int hemisphere;
  int maximum_iterations;
  double relative_tolerance;
  double absolute_tolerance;
} gamera_mi_solver_config;

typedef struct {
  int iterations;
  double initial_residual;
  double final_residual;
  int converged;
} gamera_mi_solver_stats;

/* cubedsphere_mi_dipole.h:14-28; the solver state is opaque. */
typedef struct gamera_cs_mi_dipole gamera_cs_mi_dipole;
typedef struct gamera_cs_mi_dipole_config {
  size_t azimuth_points;
  size_t colatitude_points;
  size_t current_shell;
  double ionosphere_radius_re;
  double pedersen_siemens;
  double coupling_interval_s;
  double length_norm_m;
  double speed_norm_m_s;
  double mass_density_norm_kg_m3;
  int maximum_iterations;
  double relative_tolerance;
} gamera_cs_mi_dipole_config;

/* cubedsphere_earth_boundary.h:13-24: context of the Earth callbacks. */
typedef struct {
  const gamera_cs_earth_startup *startup;
  const gamera_solar_wind_series *wind;
  const gamera_no_background_field *background; /* six faces */
  const struct gamera_cs_mi_dipole *mi;         /* NULL: E_t = 0 wall */
  double density_floor;
  double pressure_floor;
  double step_time;
  double step_dt;
  double cfl;
  int strict_inner_gas_wall;
} gamera_cs_earth_boundary;

/* cubedsphere_decomposition.h:11-18: one rank's tile of a cube face. */
typedef struct {
  int face;
  int coord[3];
  int proc_dims[3];
  size_t global_active[3];
  size_t local_active[3];
  size_t global_start[3];
} gamera_cs_decomposition;

/* cubedsphere_earth_boundary.h:53-56: context of gamera_cs_earth_tile_emf. */
typedef struct {
  const gamera_cs_decomposition *tile;
  gamera_cs_earth_boundary *earth;
} gamera_cs_earth_tile_emf_context;

#ifdef __cplusplus
}
#endif

#endif
