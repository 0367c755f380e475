/*
 * Test-infrastructure driver around the compiled reference library
 * (oracle/_ref/libgamera_ref.so). It builds a gcs_problem_config problem
 * with the reference's own setup functions and advances it with the
 * reference's own step, following the serial Earth driver
 * (cubedsphere/tests/test_cubedsphere_dipole_background.c:330-699).
 *
 * It is compiled against the reference headers only; the problem config is
 * restated here field by field (same layout as include/gamera_cs_problem.h,
 * checked by a size probe from Python) so that this file needs no header
 * from the rest of the repository.
 *
 * Only tests and the CPU-baseline leg of bench.py load this library.
 */
#include "cubedsphere_earth_boundary.h"
#include "cubedsphere_earth_startup.h"
#include "cubedsphere_halo.h"
#include "cubedsphere_mi_dipole.h"
#include "cubedsphere_step.h"
#include "nonorthogonal_background.h"
#include "nonorthogonal_initialization.h"
#include "nonorthogonal_radial_map.h"
#include "nonorthogonal_state.h"
#include "nonorthogonal_step.h"
#include "solar_wind.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { REFH_MAX_WIND = 8 };

typedef struct {
  size_t radial_cells, angular_cells, ghost;
  int radial_map_version;
  double inner_radius, outer_radius, radial_stretch;
  gamera_cs_earth_startup startup;
  double perturbation;
  uint64_t seed;
  int use_boris, strict_inner_gas_wall;
  double dt_fraction, cfl;
  size_t wind_count;
  int wind_linear;
  double wind_time[REFH_MAX_WIND];
  double wind_bz[REFH_MAX_WIND];
  double wind_reference_x;
  int mi_enabled;
  double hall_siemens;
  gamera_cs_mi_dipole_config mi;
} refh_config;

/* The prototype of gamera_cs_advance_six_with_background_planned
 * (cubedsphere_step.h:58-67). */
typedef int (*refh_step_fn)(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context);

/* The prototype of gamera_no_local_timestep (nonorthogonal_step.h:79-84). */
typedef int (*refh_timestep_fn)(
    const gamera_no_grid *grid, const double *conserved,
    const gamera_no_vec3 *cell_magnetic, const size_t active_lower[3],
    const size_t active_upper[3], const gamera_no_timestep_options *options,
    const gamera_no_vec3 *background_cell_magnetic, double *local_dt);

typedef struct {
  refh_config cfg;
  size_t active[3], ghost, extent[3];
  gamera_no_grid grid[GAMERA_CS_FACE_COUNT];
  gamera_no_storage storage[GAMERA_CS_FACE_COUNT];
  gamera_no_background_data owned[GAMERA_CS_FACE_COUNT];
  gamera_no_background_field background[GAMERA_CS_FACE_COUNT];
  gamera_solar_wind_series wind;
  gamera_cs_earth_boundary boundary;
  gamera_no_advance_options options;
  gamera_no_timestep_options timestep_options;
  gamera_cs_halo_plan *halo_plan;
  gamera_cs_mi_dipole *mi;
  double time, previous_dt, dt;
  long steps;
  int mi_updates;
  refh_step_fn step;                          /* the six-face step called */
  refh_timestep_fn timestep;                  /* the local CFL dt called */
  gamera_cs_radial_boundary_function radial;  /* its radial boundary */
  /* an ionosphere of another implementation (refh_use_mi): its state and
   * its gamera_cs_mi_dipole_maybe_update */
  void *ext_mi;
  int (*ext_update)(void *mi, double time_code);
  /* the MPI Earth driver's EMF callback on one radial tile (refh_use_tile_emf) */
  gamera_cs_decomposition tile;
  gamera_cs_earth_tile_emf_context tile_emf;
} refh_problem;

/* Wrappers of the reference's Earth callbacks at addresses of their own:
 * same results, but a step implementation cannot recognise them. */
static int wrapped_radial(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                          gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                          const size_t active[3], size_t ghost,
                          double current_time, double previous_time,
                          void *context) {
  return gamera_cs_earth_fill_radial_ghosts(grid, storage, active, ghost,
                                            current_time, previous_time,
                                            context);
}
static int wrapped_flux(gamera_no_storage *storage, const gamera_no_grid *grid,
                        const size_t lower[3], const size_t upper[3],
                        void *context) {
  return gamera_cs_earth_inner_fluid_flux(storage, grid, lower, upper, context);
}
static int wrapped_emf(gamera_no_storage *storage, const gamera_no_grid *grid,
                       const size_t lower[3], const size_t upper[3],
                       void *context) {
  return gamera_cs_earth_no_mi_emf(storage, grid, lower, upper, context);
}

size_t refh_config_size(void) { return sizeof(refh_config); }

/* splitmix64 finaliser; the perturbation hash of gcs_problem_config. */
static uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static double hashed_unit(uint64_t seed, int face, size_t i, size_t j,
                          size_t k, int variable) {
  uint64_t x = mix64(seed + (uint64_t)face);
  x = mix64(x + (uint64_t)i);
  x = mix64(x + (uint64_t)j);
  x = mix64(x + (uint64_t)k);
  x = mix64(x + (uint64_t)variable);
  return 2.0 * ((double)(x >> 11) * 0x1.0p-53) - 1.0;
}

static int perturb(refh_problem *p, int face) {
  const double amp = p->cfg.perturbation;
  if (amp == 0.0) return 0;
  const double gamma = p->cfg.startup.gamma;
  for (size_t i = 0; i < p->extent[0]; ++i)
    for (size_t j = 0; j < p->extent[1]; ++j)
      for (size_t k = 0; k < p->extent[2]; ++k) {
        const size_t cell = gamera_no_index3(p->extent, i, j, k);
        double *u = &p->storage[face].conserved[cell * GAMERA_NO_FLUX_COUNT];
        gamera_no_primitive w;
        if (gamera_no_conserved_to_primitive(u, gamma, 1e-6, 1e-8, &w) != 0)
          return -1;
        w.density *= 1.0 + amp * hashed_unit(p->cfg.seed, face, i, j, k, 0);
        w.pressure *= 1.0 + amp * hashed_unit(p->cfg.seed, face, i, j, k, 1);
        if (gamera_no_primitive_to_conserved(&w, gamma, 1e-6, 1e-8, u) != 0)
          return -1;
      }
  return 0;
}

static double cfl_limit(refh_problem *p) {
  const size_t lower[3] = {p->ghost, p->ghost, p->ghost};
  const size_t upper[3] = {p->ghost + p->active[0], p->ghost + p->active[1],
                           p->ghost + p->active[2]};
  double limit = INFINITY;
  for (int face = 0; face < GAMERA_CS_FACE_COUNT; ++face) {
    double local;
    const refh_timestep_fn fn =
        p->timestep != NULL ? p->timestep : gamera_no_local_timestep;
    if (fn(&p->grid[face], p->storage[face].conserved,
           p->storage[face].cell_magnetic, lower, upper, &p->timestep_options,
           p->owned[face].cell_magnetic, &local) != 0)
      return NAN;
    limit = fmin(limit, local);
  }
  return limit;
}

void refh_destroy(refh_problem *p) {
  if (p == NULL) return;
  gamera_cs_halo_plan_destroy(p->halo_plan);
  gamera_cs_mi_dipole_destroy(p->mi);
  for (int face = 0; face < GAMERA_CS_FACE_COUNT; ++face) {
    gamera_no_background_destroy(&p->owned[face]);
    gamera_no_storage_destroy(&p->storage[face]);
    gamera_no_grid_destroy(&p->grid[face]);
  }
  gamera_solar_wind_destroy(&p->wind);
  free(p);
}

refh_problem *refh_create(const refh_config *cfg) {
  if (cfg == NULL || cfg->ghost < 4 || cfg->wind_count < 1 ||
      cfg->wind_count > REFH_MAX_WIND)
    return NULL;
  refh_problem *p = calloc(1, sizeof(*p));
  if (p == NULL) return NULL;
  p->cfg = *cfg;
  p->step = gamera_cs_advance_six_with_background_planned;
  p->radial = gamera_cs_earth_fill_radial_ghosts;
  p->ghost = cfg->ghost;
  p->active[0] = cfg->radial_cells;
  p->active[1] = p->active[2] = cfg->angular_cells;
  for (int d = 0; d < 3; ++d) p->extent[d] = p->active[d] + 2 * p->ghost;
  gamera_solar_wind_init(&p->wind);
  const gamera_cs_earth_startup *startup = &p->cfg.startup;
  const double gamma = startup->gamma;

  gamera_no_radial_map map;
  if (gamera_no_radial_map_init(&map, cfg->radial_map_version,
                                cfg->inner_radius, cfg->outer_radius,
                                cfg->radial_stretch) != 0)
    goto fail;
  const size_t vertex_count = (p->extent[0] + 1) * (p->extent[1] + 1) *
                              (p->extent[2] + 1);
  for (int face = 0; face < GAMERA_CS_FACE_COUNT; ++face) {
    gamera_no_vec3 *vertices = calloc(vertex_count, sizeof(*vertices));
    const int ok = vertices != NULL &&
        gamera_cs_generate_vertices_radial_map(face, p->active, p->ghost,
                                               &map, vertices) == 0 &&
        gamera_no_grid_create(p->extent, vertices, &p->grid[face]) == 0 &&
        gamera_cs_calibrate_face_frames(&p->grid[face]) == 0 &&
        gamera_no_storage_create(p->extent, &p->storage[face]) == 0 &&
        gamera_no_background_create(&p->grid[face],
                                    gamera_cs_earth_startup_background,
                                    (void *)startup, &p->owned[face]) == 0;
    free(vertices);
    if (!ok) goto fail;
    p->background[face] = p->owned[face].field;
    if (gamera_no_initialize_primitives(&p->grid[face], &p->storage[face],
                                        gamera_cs_earth_startup_primitive,
                                        (void *)startup, false, gamma, 1e-6,
                                        1e-8) != 0 ||
        perturb(p, face) != 0 ||
        gamera_no_initialize_flux_from_vector_potential(
            &p->grid[face], &p->storage[face],
            gamera_cs_earth_startup_potential, (void *)startup) != 0 ||
        gamera_no_save_current_as_old(&p->storage[face]) != 0 ||
        gamera_cs_earth_background_adjust(&p->grid[face], &p->owned[face],
                                          startup) != 0 ||
        gamera_cs_earth_total_to_residual(&p->grid[face], &p->storage[face],
                                          &p->owned[face]) != 0)
      goto fail;
  }

  p->timestep_options = (gamera_no_timestep_options){
      gamma, 1e-6, 1e-8, cfg->cfl, 10.0, true, cfg->use_boris != 0, true};
  p->options = (gamera_no_advance_options){
      .stress = {gamma, 1e-6, 1e-8, 1.0, 0.25, 0.25, 10.0, true, true,
                 cfg->use_boris != 0, true},
      .emf = {1e-6, 1.0, 0.5, 10.0, cfg->cfl, 0.0, cfg->use_boris != 0, true},
      .update = {gamma, 1e-6, 1e-8, 10.0, true, cfg->use_boris != 0, true}};

  gamera_solar_wind_state states[REFH_MAX_WIND];
  for (size_t n = 0; n < cfg->wind_count; ++n) {
    states[n] = (gamera_solar_wind_state){
        .density = startup->density,
        .pressure = startup->density * startup->speed * startup->speed /
                    (startup->gamma * startup->mach * startup->mach),
        .velocity = {startup->speed, 0.0, 0.0},
        .magnetic = {startup->imf.value[0], startup->imf.value[1],
                     startup->imf.value[2]}};
    if (cfg->wind_count > 1) states[n].magnetic[GAMERA_WIND_Z] =
        cfg->wind_bz[n];
  }
  p->wind.linear_interpolation = cfg->wind_count > 1 ? cfg->wind_linear : 1;
  if (gamera_solar_wind_set(&p->wind, cfg->wind_count, cfg->wind_time,
                            states) != 0)
    goto fail;
  p->wind.reference[0] = cfg->wind_reference_x;
  p->boundary = (gamera_cs_earth_boundary){
      .startup = startup, .wind = &p->wind, .background = p->background,
      .density_floor = 1e-6, .pressure_floor = 1e-8, .cfl = cfg->cfl,
      .strict_inner_gas_wall = cfg->strict_inner_gas_wall};
  p->options.fluid_flux_sync = gamera_cs_earth_inner_fluid_flux;
  p->options.edge_emf_sync = gamera_cs_earth_no_mi_emf;
  p->options.edge_emf_context = &p->boundary;
  if (cfg->mi_enabled) {
    if (cfg->hall_siemens != 0.0 ||
        gamera_cs_mi_dipole_create(p->grid, p->storage, p->background,
                                   startup, p->active, p->ghost, &cfg->mi,
                                   &p->mi) != 0)
      goto fail;
    p->boundary.mi = p->mi;
  }
  if (gamera_cs_halo_plan_create(p->grid, p->active, p->ghost,
                                 &p->halo_plan) != 0)
    goto fail;
  const double limit = cfl_limit(p);
  if (!(limit > 0.0)) goto fail;
  p->dt = cfg->dt_fraction * limit;
  return p;
fail:
  refh_destroy(p);
  return NULL;
}

/* The problem's arrays in the reference's layout, allocated as
 * gamera_no_grid_create, gamera_no_storage_create and
 * gamera_no_background_create allocate them (so the reference's destroy
 * functions release them), but left for the caller to fill: geometry,
 * background and state are copied in from another problem of the same
 * config, which skips the reference's setup (minutes of CPU at 6 x 96^2 x
 * 192) when only its step is to be timed. The driver options, wind and
 * Earth boundary are those of refh_create; the halo plan and the first dt
 * are made on the first step, from the copied geometry and state. No MI. */
refh_problem *refh_create_shell(const refh_config *cfg) {
  if (cfg == NULL || cfg->ghost < 4 || cfg->wind_count < 1 ||
      cfg->wind_count > REFH_MAX_WIND || cfg->mi_enabled)
    return NULL;
  refh_problem *p = calloc(1, sizeof(*p));
  if (p == NULL) return NULL;
  p->cfg = *cfg;
  p->step = gamera_cs_advance_six_with_background_planned;
  p->radial = gamera_cs_earth_fill_radial_ghosts;
  p->ghost = cfg->ghost;
  p->active[0] = cfg->radial_cells;
  p->active[1] = p->active[2] = cfg->angular_cells;
  for (int d = 0; d < 3; ++d) p->extent[d] = p->active[d] + 2 * p->ghost;
  gamera_solar_wind_init(&p->wind);
  const gamera_cs_earth_startup *startup = &p->cfg.startup;
  const double gamma = startup->gamma;
  for (int face = 0; face < GAMERA_CS_FACE_COUNT; ++face) {
    gamera_no_grid *g = &p->grid[face];
    gamera_no_background_data *b = &p->owned[face];
    for (int a = 0; a < 3; ++a) {
      g->cell_extent[a] = p->extent[a];
      g->vertex_extent[a] = p->extent[a] + 1;
    }
    for (int d = 0; d < 3; ++d)
      for (int a = 0; a < 3; ++a) {
        g->face[d].extent[a] = p->extent[a] + (a == d ? 1U : 0U);
        g->edge[d].extent[a] = p->extent[a] + (a == d ? 0U : 1U);
      }
    const size_t cells = gamera_no_element_count3(g->cell_extent);
    g->vertex = calloc(gamera_no_element_count3(g->vertex_extent), sizeof(*g->vertex));
    g->cell = calloc(cells, sizeof(*g->cell));
    b->cell_magnetic = calloc(cells, sizeof(*b->cell_magnetic));
    b->cell_force = calloc(cells, sizeof(*b->cell_force));
    int ok = g->vertex != NULL && g->cell != NULL && b->cell_magnetic != NULL &&
             b->cell_force != NULL;
    for (int d = 0; d < 3; ++d) {
      const size_t faces = gamera_no_element_count3(g->face[d].extent);
      const size_t edges = gamera_no_element_count3(g->edge[d].extent);
      g->face[d].value = calloc(faces, sizeof(*g->face[d].value));
      g->edge[d].value = calloc(edges, sizeof(*g->edge[d].value));
      g->edge[d].valid = calloc(edges, 1);
      b->face_magnetic[d] = calloc(faces, sizeof(*b->face_magnetic[d]));
      b->face_flux[d] = calloc(faces, sizeof(*b->face_flux[d]));
      b->edge_magnetic[d] = calloc(edges, sizeof(*b->edge_magnetic[d]));
      ok = ok && g->face[d].value != NULL && g->edge[d].value != NULL &&
           g->edge[d].valid != NULL && b->face_magnetic[d] != NULL &&
           b->face_flux[d] != NULL && b->edge_magnetic[d] != NULL;
      b->field.face_magnetic[d] = b->face_magnetic[d];
      b->field.face_flux[d] = b->face_flux[d];
      b->field.edge_magnetic[d] = b->edge_magnetic[d];
    }
    b->field.cell_magnetic = b->cell_magnetic;
    b->field.cell_force = b->cell_force;
    if (!ok || gamera_no_storage_create(p->extent, &p->storage[face]) != 0) goto fail;
    p->background[face] = b->field;
  }
  p->timestep_options = (gamera_no_timestep_options){
      gamma, 1e-6, 1e-8, cfg->cfl, 10.0, true, cfg->use_boris != 0, true};
  p->options = (gamera_no_advance_options){
      .stress = {gamma, 1e-6, 1e-8, 1.0, 0.25, 0.25, 10.0, true, true,
                 cfg->use_boris != 0, true},
      .emf = {1e-6, 1.0, 0.5, 10.0, cfg->cfl, 0.0, cfg->use_boris != 0, true},
      .update = {gamma, 1e-6, 1e-8, 10.0, true, cfg->use_boris != 0, true}};
  gamera_solar_wind_state states[REFH_MAX_WIND];
  for (size_t n = 0; n < cfg->wind_count; ++n) {
    states[n] = (gamera_solar_wind_state){
        .density = startup->density,
        .pressure = startup->density * startup->speed * startup->speed /
                    (startup->gamma * startup->mach * startup->mach),
        .velocity = {startup->speed, 0.0, 0.0},
        .magnetic = {startup->imf.value[0], startup->imf.value[1],
                     startup->imf.value[2]}};
    if (cfg->wind_count > 1) states[n].magnetic[GAMERA_WIND_Z] = cfg->wind_bz[n];
  }
  p->wind.linear_interpolation = cfg->wind_count > 1 ? cfg->wind_linear : 1;
  if (gamera_solar_wind_set(&p->wind, cfg->wind_count, cfg->wind_time, states) != 0)
    goto fail;
  p->wind.reference[0] = cfg->wind_reference_x;
  p->boundary = (gamera_cs_earth_boundary){
      .startup = startup, .wind = &p->wind, .background = p->background,
      .density_floor = 1e-6, .pressure_floor = 1e-8, .cfl = cfg->cfl,
      .strict_inner_gas_wall = cfg->strict_inner_gas_wall};
  p->options.fluid_flux_sync = gamera_cs_earth_inner_fluid_flux;
  p->options.edge_emf_sync = gamera_cs_earth_no_mi_emf;
  p->options.edge_emf_context = &p->boundary;
  return p;
fail:
  refh_destroy(p);
  return NULL;
}

/* Couple the driver loop to another implementation's ionosphere state `mi`
 * (with the gamera_cs_mi_dipole_* prototypes): the Earth context's mi
 * becomes it and each step calls `update` in place of the reference's
 * gamera_cs_mi_dipole_maybe_update. NULL restores the reference's. */
void refh_use_mi(refh_problem *p, void *mi, int (*update)(void *, double)) {
  if (p == NULL) return;
  p->ext_mi = mi;
  p->ext_update = update;
  p->boundary.mi = mi != NULL ? (const struct gamera_cs_mi_dipole *)mi : p->mi;
}

/* The grid's background, the startup and the active box of the problem,
 * for creating an ionosphere state on it. */
const gamera_no_background_field *refh_background(const refh_problem *p) {
  return p->background;
}
const gamera_cs_earth_startup *refh_startup(const refh_problem *p) {
  return &p->cfg.startup;
}

/* Replace the wind series with the reference's own setter
 * (gamera_solar_wind_set, solar_wind.c:55-80): count samples,
 * state[8 n ..] = density, pressure, velocity, magnetic. */
int refh_set_wind(refh_problem *p, size_t count, const double *time,
                  const double *state, int linear) {
  if (p == NULL || count == 0 || time == NULL || state == NULL) return -1;
  gamera_solar_wind_state *w = calloc(count, sizeof(*w));
  if (w == NULL) return -1;
  for (size_t n = 0; n < count; ++n) {
    const double *x = &state[8 * n];
    w[n].density = x[0];
    w[n].pressure = x[1];
    for (int c = 0; c < 3; ++c) {
      w[n].velocity[c] = x[2 + c];
      w[n].magnetic[c] = x[5 + c];
    }
  }
  const int rc = gamera_solar_wind_set(&p->wind, count, time, w);
  free(w);
  if (rc != 0) return -1;
  p->wind.linear_interpolation = count > 1 ? (linear != 0) : 1;
  return 0;
}

/* 0: hydrodynamic steps (the step options' use_mhd = false in all three
 * option groups, and the dt rule without the Alfven speed); 1: MHD. */
void refh_set_mhd(refh_problem *p, int on) {
  if (p == NULL) return;
  p->options.stress.use_mhd = on != 0;
  p->options.update.use_mhd = on != 0;
  p->timestep_options.use_mhd = on != 0;
}

/* Clock of the driver loop (time, dt of the last step, previous dt, steps),
 * e.g. to continue a run another implementation started. */
void refh_set_clock(refh_problem *p, double time, double dt, double previous_dt,
                    long steps) {
  if (p == NULL) return;
  p->time = time;
  p->dt = dt;
  p->previous_dt = previous_dt;
  p->steps = steps;
}

/* One driver iteration: dt rule, MI cadence, one six-face step. */
static int advance_one(refh_problem *p) {
  if (p->halo_plan == NULL &&
      gamera_cs_halo_plan_create(p->grid, p->active, p->ghost, &p->halo_plan) != 0)
    return -1;
  if (p->steps == 0 && !(p->dt > 0.0)) {
    const double limit = cfl_limit(p);
    if (!(limit > 0.0)) return -1;
    p->dt = p->cfg.dt_fraction * limit;
  }
  if (p->steps > 0) {
    const double limit = cfl_limit(p);
    if (!(limit > 0.0)) return -1;
    p->dt = fmin(p->cfg.dt_fraction * limit, 1.25 * p->previous_dt);
  }
  if (p->ext_mi != NULL) {
    const int updated = p->ext_update(p->ext_mi, p->time);
    if (updated < 0) return -1;
    p->mi_updates += updated;
  } else if (p->mi != NULL) {
    const int updated = gamera_cs_mi_dipole_maybe_update(p->mi, p->time);
    if (updated < 0) return -1;
    p->mi_updates += updated;
  }
  p->boundary.step_dt = p->dt;
  p->boundary.step_time = p->time;
  const double ratio = p->steps == 0 ? 0.0 : 0.5 * p->dt / p->previous_dt;
  if (p->step(p->grid, p->storage, p->active, p->ghost, ratio, p->dt,
              &p->options, p->background, p->halo_plan, p->radial, p->time,
              p->steps == 0 ? p->time : p->time - p->previous_dt,
              &p->boundary) != 0)
    return -1;
  p->time += p->dt;
  p->previous_dt = p->dt;
  ++p->steps;
  return 0;
}

int refh_advance(refh_problem *p, long count) {
  if (p == NULL || count < 0) return -1;
  for (long n = 0; n < count; ++n)
    if (advance_one(p) != 0) return -1;
  return 0;
}

/* One step with caller-chosen dt and predictor ratio (no dt rule, no MI). */
int refh_step_explicit(refh_problem *p, double ratio, double dt,
                       double time, double previous_time) {
  if (p == NULL) return -1;
  p->boundary.step_dt = dt;
  p->boundary.step_time = time;
  return p->step(p->grid, p->storage, p->active, p->ghost, ratio, dt,
                 &p->options, p->background, p->halo_plan, p->radial, time,
                 previous_time, &p->boundary);
}

/* Run later steps through `step` (NULL: the reference's own step), e.g. a
 * drop-in implementation with the same prototype. */
void refh_set_step(refh_problem *p, void *step) {
  if (p == NULL) return;
  p->step = step != NULL ? (refh_step_fn)step
                         : gamera_cs_advance_six_with_background_planned;
}

/* One call of a local CFL dt function with gamera_no_local_timestep's
 * prototype (NULL: the reference's) on face `face`'s arrays. */
int refh_local_timestep(refh_problem *p, void *fn, int face, const size_t lower[3],
                        const size_t upper[3], const gamera_no_timestep_options *options,
                        double *dt) {
  if (p == NULL || face < 0 || face >= GAMERA_CS_FACE_COUNT) return -2;
  const refh_timestep_fn call = fn != NULL ? (refh_timestep_fn)fn : gamera_no_local_timestep;
  return call(&p->grid[face], p->storage[face].conserved, p->storage[face].cell_magnetic,
              lower, upper, options, p->owned[face].cell_magnetic, dt);
}

/* 1: the edge EMF callback is gamera_cs_earth_tile_emf
 * (cubedsphere_earth_boundary.c:453-464) with a decomposition of one tile
 * per cube face, as the MPI Earth driver sets it
 * (test_cubedsphere_mpi_earth_tile.c:369-379); 0: gamera_cs_earth_no_mi_emf. */
void refh_use_tile_emf(refh_problem *p, int on) {
  if (p == NULL) return;
  if (on) {
    memset(&p->tile, 0, sizeof(p->tile));
    for (int a = 0; a < 3; ++a) {
      p->tile.proc_dims[a] = 1;
      p->tile.global_active[a] = p->tile.local_active[a] = p->active[a];
    }
    p->tile_emf.tile = &p->tile;
    p->tile_emf.earth = &p->boundary;
    p->options.edge_emf_sync = gamera_cs_earth_tile_emf;
    p->options.edge_emf_context = &p->tile_emf;
  } else {
    p->options.edge_emf_sync = gamera_cs_earth_no_mi_emf;
    p->options.edge_emf_context = &p->boundary;
  }
}

/* The local CFL dt function of the driver's dt rule (the prototype of
 * gamera_no_local_timestep); NULL restores the reference's. */
void refh_set_timestep(refh_problem *p, void *timestep) {
  if (p == NULL) return;
  p->timestep = (refh_timestep_fn)timestep;
}

/* 1: pass the Earth callbacks through wrappers (see wrapped_radial). */
void refh_use_wrapped_callbacks(refh_problem *p, int on) {
  if (p == NULL) return;
  p->radial = on ? wrapped_radial : gamera_cs_earth_fill_radial_ghosts;
  p->options.fluid_flux_sync = on ? wrapped_flux : gamera_cs_earth_inner_fluid_flux;
  p->options.edge_emf_sync = on ? wrapped_emf : gamera_cs_earth_no_mi_emf;
}

/* One step of the plain entry point gamera_cs_advance_six
 * (cubedsphere_step.h:16-21) on this problem's arrays: no background, no
 * radial boundary (radial outflow ghosts), no callbacks. `step` has that
 * prototype; NULL calls the reference's own. The clock is left alone. */
typedef int (*refh_plain_fn)(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                             gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                             const size_t active[3], size_t ghost,
                             double predictor_ratio, double dt,
                             const gamera_no_advance_options *options);
int refh_step_plain(refh_problem *p, void *step, double ratio, double dt) {
  if (p == NULL) return -1;
  gamera_no_advance_options o = p->options;
  o.stress.use_background = false;
  o.emf.use_background = false;
  o.update.use_background = false;
  o.fluid_flux_sync = NULL;
  o.edge_emf_sync = NULL;
  o.fluid_flux_context = NULL;
  o.edge_emf_context = NULL;
  refh_plain_fn fn = step != NULL ? (refh_plain_fn)step : gamera_cs_advance_six;
  return fn(p->grid, p->storage, p->active, p->ghost, ratio, dt, &o);
}

/* The caller-owned arrays, for a step implementation's bookkeeping. */
const gamera_no_grid *refh_grid(const refh_problem *p) {
  return p ? p->grid : NULL;
}
gamera_no_storage *refh_storage(refh_problem *p) {
  return p ? p->storage : NULL;
}

void refh_clock(const refh_problem *p, double out[4]) {
  out[0] = p->time;
  out[1] = p->dt;
  out[2] = p->previous_dt;
  out[3] = (double)p->steps;
}

double refh_cfl_limit(refh_problem *p) { return cfl_limit(p); }

/* Pointer to one array of one face and its length in doubles. */
int refh_field(refh_problem *p, int face, int field, double **data,
               size_t *count) {
  if (p == NULL || face < 0 || face >= GAMERA_CS_FACE_COUNT || data == NULL ||
      count == NULL)
    return -1;
  gamera_no_grid *g = &p->grid[face];
  gamera_no_storage *s = &p->storage[face];
  gamera_no_background_data *b = &p->owned[face];
  const size_t cells = gamera_no_element_count3(g->cell_extent);
  size_t faces[3], edges[3];
  for (int d = 0; d < 3; ++d) {
    faces[d] = gamera_no_element_count3(g->face[d].extent);
    edges[d] = gamera_no_element_count3(g->edge[d].extent);
  }
  /* Field ids follow include/gamera_cs_problem.h. */
  if (field >= 0 && field <= 5) {
    double *arrays[6] = {s->conserved, s->old_conserved,
                         s->predicted_conserved,
                         s->cell_magnetic->value, s->old_cell_magnetic->value,
                         s->predicted_cell_magnetic->value};
    *data = arrays[field];
    *count = cells * (field < 3 ? 5 : 3);
    return 0;
  }
  if (field >= 6 && field < 24) {
    const int group = (field - 6) / 3, d = (field - 6) % 3;
    switch (group) {
      case 0: *data = s->face_flux[d]; *count = faces[d]; return 0;
      case 1: *data = s->old_face_flux[d]; *count = faces[d]; return 0;
      case 2: *data = s->predicted_face_flux[d]; *count = faces[d]; return 0;
      case 3: *data = s->edge_emf[d]; *count = edges[d]; return 0;
      case 4: *data = s->fluid_face_flux[d]; *count = 5 * faces[d]; return 0;
      default: *data = s->maxwell_face_flux[d]->value; *count = 3 * faces[d];
               return 0;
    }
  }
  switch (field) {
    case 24: *data = s->hydro_rate; *count = 5 * cells; return 0;
    case 25: *data = s->maxwell_rate->value; *count = 3 * cells; return 0;
    case 26: *data = g->vertex->value;
             *count = 3 * gamera_no_element_count3(g->vertex_extent); return 0;
    case 27: *data = &g->cell->volume;
             *count = cells * (sizeof(gamera_no_cell_geometry) / 8); return 0;
    default: break;
  }
  if (field >= 28 && field < 31) {
    const int d = field - 28;
    *data = &g->face[d].value->area;
    *count = faces[d] * (sizeof(gamera_no_face_geometry) / 8);
    return 0;
  }
  if (field >= 31 && field < 34) {
    const int d = field - 31;
    *data = &g->edge[d].value->length;
    *count = edges[d] * (sizeof(gamera_no_edge_geometry) / 8);
    return 0;
  }
  if (field == 34) { *data = b->cell_magnetic->value; *count = 3 * cells; return 0; }
  if (field == 35) { *data = b->cell_force->value; *count = 3 * cells; return 0; }
  if (field >= 36 && field < 39) {
    *data = b->face_magnetic[field - 36]->value;
    *count = 3 * faces[field - 36];
    return 0;
  }
  if (field >= 39 && field < 42) {
    *data = b->face_flux[field - 39];
    *count = faces[field - 39];
    return 0;
  }
  if (field >= 42 && field < 45) {
    *data = b->edge_magnetic[field - 42]->value;
    *count = 3 * edges[field - 42];
    return 0;
  }
  return -1;
}

int refh_edge_valid(refh_problem *p, int face, int direction,
                    unsigned char **data, size_t *count) {
  if (p == NULL || face < 0 || face >= GAMERA_CS_FACE_COUNT || direction < 0 ||
      direction > 2)
    return -1;
  *data = p->grid[face].edge[direction].valid;
  *count = gamera_no_element_count3(p->grid[face].edge[direction].extent);
  return 0;
}

/* Diagnostic counters written by the step into storage. */
void refh_counters(const refh_problem *p, int face, double out[2]) {
  out[0] = (double)p->storage[face].emergency_diffusion_face_count;
  out[1] = p->storage[face].emergency_diffusion_max_interface_speed;
}

int refh_mi_updates(const refh_problem *p) { return p->mi_updates; }

/* The driver's snapshot writers (ref_snapshot.c) on this problem's arrays,
 * with its own ionosphere state (if any) for the MI files. */
int refh_driver_write_snapshot(const char *path,
                               const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                               const gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                               const gamera_no_background_data owned[GAMERA_CS_FACE_COUNT],
                               const size_t active[3], size_t ghost, double gamma,
                               const gamera_cs_mi_dipole *mi);
int refh_driver_write_drive_frame(const char *prefix, long frame, long step, double seconds,
                                  const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                                  const gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                                  const gamera_no_background_data owned[GAMERA_CS_FACE_COUNT],
                                  const size_t active[3], size_t ghost, double gamma,
                                  const gamera_cs_mi_dipole *mi, const char *index_path);

int refh_driver_load_snapshot(const char *path, const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                              gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                              const gamera_no_background_data owned[GAMERA_CS_FACE_COUNT],
                              const size_t active[3], size_t ghost, double gamma);

/* The driver's earth-mi-replay input (:526-528): conserved state and
 * residual cell B of both levels from a snapshot. */
int refh_load_snapshot(refh_problem *p, const char *path) {
  if (p == NULL || path == NULL) return -1;
  return refh_driver_load_snapshot(path, p->grid, p->storage, p->owned, p->active, p->ghost,
                                   p->cfg.startup.gamma);
}

int refh_write_snapshot(refh_problem *p, const char *path) {
  if (p == NULL || path == NULL) return -1;
  return refh_driver_write_snapshot(path, p->grid, p->storage, p->owned, p->active, p->ghost,
                                    p->cfg.startup.gamma, p->mi);
}

int refh_write_drive_frame(refh_problem *p, const char *prefix, long frame, long step,
                           double seconds, const char *index_path) {
  if (p == NULL || prefix == NULL || index_path == NULL) return -1;
  return refh_driver_write_drive_frame(prefix, frame, step, seconds, p->grid, p->storage,
                                       p->owned, p->active, p->ghost, p->cfg.startup.gamma,
                                       p->mi, index_path);
}

/* The reference's own ionosphere state of the problem (NULL without MI). */
gamera_cs_mi_dipole *refh_mi(const refh_problem *p) {
  return p != NULL ? p->mi : NULL;
}

/* Both hemispheres' sampled field-aligned current [hemi][colatitude]
 * [azimuth] in A/m^2 (cubedsphere_mi_dipole.h:30-41). */
int refh_mi_fac(const refh_problem *p, double *out, size_t count) {
  gamera_cs_mi_analysis_view v;
  if (p == NULL || p->mi == NULL ||
      gamera_cs_mi_dipole_analysis_view(p->mi, &v) != 0)
    return -1;
  const size_t n = v.azimuth_points * v.colatitude_points;
  if (count != 2 * n) return -1;
  for (int h = 0; h < 2; ++h)
    memcpy(out + (size_t)h * n, v.fac_parallel_a_m2[h], n * sizeof(double));
  return 0;
}

/* Both hemispheres' potential [hemi][colatitude][azimuth] in volts. */
int refh_mi_potential(const refh_problem *p, double *out, size_t count) {
  if (p == NULL || p->mi == NULL) return -1;
  double epoch, next;
  return gamera_cs_mi_dipole_export_potential(p->mi, out, count, &epoch,
                                              &next);
}

/* The reference's gamera_no_apply_pressure_control
 * (NO/nonorthogonal_step.c:74-171) on every face's active cells, as the MPI
 * Earth driver applies it after a step
 * (cubedsphere/tests/test_cubedsphere_mpi_earth_tile.c:459-479), with the
 * counts and maxima accumulated over the faces. */
int refh_pressure_control(refh_problem *p, double dt, double gamma, double density_floor,
                          double pressure_floor, double light_speed,
                          double low_density_threshold, double low_density_sound_speed,
                          unsigned long long counts[2], double maxima[2]) {
  if (p == NULL || counts == NULL || maxima == NULL) return -1;
  counts[0] = counts[1] = 0;
  maxima[0] = maxima[1] = 0.0;
  const size_t lower[3] = {p->ghost, p->ghost, p->ghost};
  const size_t upper[3] = {p->ghost + p->active[0], p->ghost + p->active[1],
                           p->ghost + p->active[2]};
  int failed = 0;
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    size_t low = 0, high = 0;
    double sound = 0.0, pressure = 0.0;
    failed |= gamera_no_apply_pressure_control(
                  &p->storage[f], &p->grid[f], lower, upper, dt, gamma, density_floor,
                  pressure_floor, light_speed, low_density_threshold,
                  low_density_sound_speed, &low, &high, &sound, &pressure) != 0;
    counts[0] += low;
    counts[1] += high;
    maxima[0] = fmax(maxima[0], sound);
    maxima[1] = fmax(maxima[1], pressure);
  }
  return failed ? -1 : 0;
}
