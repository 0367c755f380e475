/*
 * Problem description shared by every implementation of the six-face Earth
 * step in this repository (the HIP library, the CPU oracle and the harness
 * around the compiled reference), so that each builds the same inputs.
 *
 * A problem is the reference's serial Earth driver setup
 * (cubedsphere/tests/test_cubedsphere_dipole_background.c:260-509): six
 * cubed-sphere faces of nr x n x n active cells with `ghost` layers, the
 * Problem-21 startup state, a split dipole background, a solar-wind series
 * feeding the Earth radial boundary, and optionally the aligned-dipole
 * ionosphere (MI) coupling. Its step loop is the driver's dt rule (:589-699):
 *   dt_0 = dt_fraction * min_faces local CFL dt,
 *   dt_n = min(dt_fraction * min_faces local CFL dt, 1.25 * dt_{n-1}),
 *   predictor_ratio = 0 at step 0, else 0.5 * dt / dt_prev.
 *
 * Additions that the reference driver does not have, each off by default:
 *  - perturbation: density and pressure of every cell are multiplied by
 *    (1 + perturbation * u) with u uniform in [-1, 1) from a counter-based
 *    splitmix64 hash of (seed, face, i, j, k, variable), applied to the
 *    initial primitive state before the old time level is saved;
 *  - a piecewise-constant or linear solar-wind Bz series (wind_count > 1);
 *  - wind_reference_x: the driver ties the ballistic reference point to
 *    startup.front_x (:491); here it is a separate field.
 */
#ifndef GAMERA_CS_PROBLEM_H
#define GAMERA_CS_PROBLEM_H

#include "gamera_cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { GCS_MAX_WIND_SAMPLES = 8 };

typedef struct {
  size_t radial_cells;        /* nr */
  size_t angular_cells;       /* n (both face angles) */
  size_t ghost;               /* >= 4 */
  int radial_map_version;     /* 1: r_in + (r_out-r_in) expm1(s f)/expm1(s); 4: Earth v4 */
  double inner_radius;
  double outer_radius;
  double radial_stretch;      /* s for version 1 */
  gamera_cs_earth_startup startup;
  double perturbation;        /* 0 disables */
  uint64_t seed;
  int use_boris;
  int strict_inner_gas_wall;
  double dt_fraction;         /* driver default 0.5 */
  double cfl;                 /* 0.3 */
  size_t wind_count;          /* 1: steady wind from startup */
  int wind_linear;            /* series interpolation */
  double wind_time[GCS_MAX_WIND_SAMPLES];   /* code time */
  double wind_bz[GCS_MAX_WIND_SAMPLES];     /* code B */
  double wind_reference_x;    /* ballistic-delay reference point (x, 0, 0) */
  int mi_enabled;
  double hall_siemens;        /* 0: constant-Pedersen PCG (reference adapter) */
  gamera_cs_mi_dipole_config mi;
} gcs_problem_config;

/* Identifiers of the arrays a problem exposes (count in doubles). */
enum {
  GCS_FIELD_CONSERVED = 0,
  GCS_FIELD_OLD_CONSERVED,
  GCS_FIELD_PREDICTED_CONSERVED,
  GCS_FIELD_CELL_MAGNETIC,
  GCS_FIELD_OLD_CELL_MAGNETIC,
  GCS_FIELD_PREDICTED_CELL_MAGNETIC,
  GCS_FIELD_FACE_FLUX,            /* + direction */
  GCS_FIELD_OLD_FACE_FLUX = GCS_FIELD_FACE_FLUX + 3,
  GCS_FIELD_PREDICTED_FACE_FLUX = GCS_FIELD_OLD_FACE_FLUX + 3,
  GCS_FIELD_EDGE_EMF = GCS_FIELD_PREDICTED_FACE_FLUX + 3,
  GCS_FIELD_FLUID_FACE_FLUX = GCS_FIELD_EDGE_EMF + 3,
  GCS_FIELD_MAXWELL_FACE_FLUX = GCS_FIELD_FLUID_FACE_FLUX + 3,
  GCS_FIELD_HYDRO_RATE = GCS_FIELD_MAXWELL_FACE_FLUX + 3,
  GCS_FIELD_MAXWELL_RATE,
  GCS_FIELD_VERTEX,
  GCS_FIELD_CELL_GEOMETRY,        /* gamera_no_cell_geometry as doubles */
  GCS_FIELD_FACE_GEOMETRY,        /* + direction */
  GCS_FIELD_EDGE_GEOMETRY = GCS_FIELD_FACE_GEOMETRY + 3,
  GCS_FIELD_BG_CELL_MAGNETIC = GCS_FIELD_EDGE_GEOMETRY + 3,
  GCS_FIELD_BG_CELL_FORCE,
  GCS_FIELD_BG_FACE_MAGNETIC,     /* + direction */
  GCS_FIELD_BG_FACE_FLUX = GCS_FIELD_BG_FACE_MAGNETIC + 3,
  GCS_FIELD_BG_EDGE_MAGNETIC = GCS_FIELD_BG_FACE_FLUX + 3,
  GCS_FIELD_COUNT = GCS_FIELD_BG_EDGE_MAGNETIC + 3
};

#ifdef __cplusplus
}
#endif

#endif
