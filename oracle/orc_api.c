/* Problem API of the oracle (test infrastructure; see orc.h).
 *
 * orc_create allocates a problem's reference-layout arrays from its
 * gcs_problem_config; geometry, background and state are then copied in by
 * the caller (tests copy them from the problem under test), so the oracle
 * steps exactly the inputs it is asked to check. The step loop is the
 * reference driver's (test_cubedsphere_dipole_background.c:589-699). */
#include "orc.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  gcs_problem_config cfg;
  size_t active[3], ghost, extent[3];
  gamera_no_grid grid[ORC_FACES];
  gamera_no_storage storage[ORC_FACES];
  gamera_no_background_data bg[ORC_FACES];
  gamera_no_background_field field[ORC_FACES];
  gamera_solar_wind_series wind;
  double wind_time[GCS_MAX_WIND_SAMPLES];
  gamera_solar_wind_state wind_state[GCS_MAX_WIND_SAMPLES];
  gamera_cs_earth_boundary boundary;
  gamera_no_advance_options options;
  gamera_no_timestep_options timestep;
  orc_halo_plan *halo;
  orc_edge_plan *edges;
  orc_mi *mi;           /* created on the first step when mi_enabled */
  double *series_time;  /* a series set by orc_set_wind (else NULL) */
  gamera_solar_wind_state *series_state;
  int mi_updates;
  double time, dt, previous_dt;
  long steps;
} orc_problem;

static void *zalloc(size_t n, size_t size) { return calloc(n ? n : 1, size); }

void orc_destroy(orc_problem *p) {
  if (p == NULL) return;
  for (int f = 0; f < ORC_FACES; ++f) {
    gamera_no_grid *g = &p->grid[f];
    gamera_no_storage *s = &p->storage[f];
    gamera_no_background_data *b = &p->bg[f];
    free(g->vertex);
    free(g->cell);
    free(s->conserved);
    free(s->old_conserved);
    free(s->predicted_conserved);
    free(s->cell_magnetic);
    free(s->old_cell_magnetic);
    free(s->predicted_cell_magnetic);
    free(s->hydro_rate);
    free(s->maxwell_rate);
    free(b->cell_magnetic);
    free(b->cell_force);
    for (int d = 0; d < 3; ++d) {
      free(g->face[d].value);
      free(g->edge[d].value);
      free(g->edge[d].valid);
      free(s->face_flux[d]);
      free(s->old_face_flux[d]);
      free(s->predicted_face_flux[d]);
      free(s->edge_emf[d]);
      free(s->fluid_face_flux[d]);
      free(s->maxwell_face_flux[d]);
      free(b->face_magnetic[d]);
      free(b->face_flux[d]);
      free(b->edge_magnetic[d]);
    }
  }
  orc_halo_plan_free(p->halo);
  orc_edge_plan_free(p->edges);
  orc_mi_destroy(p->mi);
  free(p->series_time);
  free(p->series_state);
  free(p);
}

orc_problem *orc_create(const gcs_problem_config *cfg) {
  if (cfg == NULL || cfg->ghost < 4 || cfg->wind_count < 1 ||
      cfg->wind_count > GCS_MAX_WIND_SAMPLES)
    return NULL;
  orc_problem *p = calloc(1, sizeof(*p));
  if (p == NULL) return NULL;
  p->cfg = *cfg;
  p->ghost = cfg->ghost;
  p->active[0] = cfg->radial_cells;
  p->active[1] = p->active[2] = cfg->angular_cells;
  for (int d = 0; d < 3; ++d) p->extent[d] = p->active[d] + 2 * p->ghost;
  int ok = 1;
  for (int f = 0; f < ORC_FACES; ++f) {
    gamera_no_grid *g = &p->grid[f];
    gamera_no_storage *s = &p->storage[f];
    gamera_no_background_data *b = &p->bg[f];
    for (int a = 0; a < 3; ++a) {
      g->cell_extent[a] = s->cell_extent[a] = p->extent[a];
      g->vertex_extent[a] = p->extent[a] + 1;
      for (int d = 0; d < 3; ++d) {
        g->face[d].extent[a] = s->face_extent[d][a] = p->extent[a] + (a == d);
        g->edge[d].extent[a] = s->edge_extent[d][a] = p->extent[a] + (a != d);
      }
    }
    const size_t nc = orc_count(p->extent);
    g->vertex = zalloc(orc_count(g->vertex_extent), sizeof(vec3));
    g->cell = zalloc(nc, sizeof(gamera_no_cell_geometry));
    s->conserved = zalloc(nc * ORC_NQ, sizeof(double));
    s->old_conserved = zalloc(nc * ORC_NQ, sizeof(double));
    s->predicted_conserved = zalloc(nc * ORC_NQ, sizeof(double));
    s->cell_magnetic = zalloc(nc, sizeof(vec3));
    s->old_cell_magnetic = zalloc(nc, sizeof(vec3));
    s->predicted_cell_magnetic = zalloc(nc, sizeof(vec3));
    s->hydro_rate = zalloc(nc * ORC_NQ, sizeof(double));
    s->maxwell_rate = zalloc(nc, sizeof(vec3));
    b->cell_magnetic = zalloc(nc, sizeof(vec3));
    b->cell_force = zalloc(nc, sizeof(vec3));
    ok = ok && g->vertex && g->cell && s->conserved && s->old_conserved &&
         s->predicted_conserved && s->cell_magnetic && s->old_cell_magnetic &&
         s->predicted_cell_magnetic && s->hydro_rate && s->maxwell_rate &&
         b->cell_magnetic && b->cell_force;
    for (int d = 0; d < 3; ++d) {
      const size_t nf = orc_count(s->face_extent[d]);
      const size_t ne = orc_count(s->edge_extent[d]);
      g->face[d].value = zalloc(nf, sizeof(gamera_no_face_geometry));
      g->edge[d].value = zalloc(ne, sizeof(gamera_no_edge_geometry));
      g->edge[d].valid = zalloc(ne, 1);
      s->face_flux[d] = zalloc(nf, sizeof(double));
      s->old_face_flux[d] = zalloc(nf, sizeof(double));
      s->predicted_face_flux[d] = zalloc(nf, sizeof(double));
      s->edge_emf[d] = zalloc(ne, sizeof(double));
      s->fluid_face_flux[d] = zalloc(nf * ORC_NQ, sizeof(double));
      s->maxwell_face_flux[d] = zalloc(nf, sizeof(vec3));
      b->face_magnetic[d] = zalloc(nf, sizeof(vec3));
      b->face_flux[d] = zalloc(nf, sizeof(double));
      b->edge_magnetic[d] = zalloc(ne, sizeof(vec3));
      ok = ok && g->face[d].value && g->edge[d].value && g->edge[d].valid &&
           s->face_flux[d] && s->old_face_flux[d] && s->predicted_face_flux[d] &&
           s->edge_emf[d] && s->fluid_face_flux[d] && s->maxwell_face_flux[d] &&
           b->face_magnetic[d] && b->face_flux[d] && b->edge_magnetic[d];
      b->field.face_magnetic[d] = b->face_magnetic[d];
      b->field.face_flux[d] = b->face_flux[d];
      b->field.edge_magnetic[d] = b->edge_magnetic[d];
    }
    b->field.cell_magnetic = b->cell_magnetic;
    b->field.cell_force = b->cell_force;
    p->field[f] = b->field;
  }
  if (!ok) {
    orc_destroy(p);
    return NULL;
  }
  const gamera_cs_earth_startup *su = &p->cfg.startup;
  const double gamma = su->gamma;
  const bool boris = cfg->use_boris != 0;
  p->timestep = (gamera_no_timestep_options){gamma, 1e-6, 1e-8, cfg->cfl,
                                             10.0,  true, boris, true};
  p->options = (gamera_no_advance_options){
      .stress = {gamma, 1e-6, 1e-8, 1.0, 0.25, 0.25, 10.0, true, true, boris,
                 true},
      .emf = {1e-6, 1.0, 0.5, 10.0, cfg->cfl, 0.0, boris, true},
      .update = {gamma, 1e-6, 1e-8, 10.0, true, boris, true}};
  /* The wind series of the config (solar_wind.c:142-171, no Bx relation). */
  for (size_t n = 0; n < cfg->wind_count; ++n) {
    gamera_solar_wind_state *w = &p->wind_state[n];
    w->density = su->density;
    w->pressure = su->density * su->speed * su->speed /
                  (su->gamma * su->mach * su->mach);
    w->velocity[0] = su->speed;
    w->velocity[1] = w->velocity[2] = 0.0;
    for (int c = 0; c < 3; ++c) w->magnetic[c] = su->imf.value[c];
    if (cfg->wind_count > 1) w->magnetic[2] = cfg->wind_bz[n];
    p->wind_time[n] = cfg->wind_time[n];
  }
  p->wind.count = cfg->wind_count;
  p->wind.time = p->wind_time;
  p->wind.state = p->wind_state;
  p->wind.linear_interpolation = cfg->wind_count > 1 ? cfg->wind_linear : 1;
  p->wind.reference[0] = cfg->wind_reference_x;
  p->boundary = (gamera_cs_earth_boundary){
      .startup = &p->cfg.startup, .wind = &p->wind, .background = p->field,
      .density_floor = 1e-6, .pressure_floor = 1e-8, .cfl = cfg->cfl,
      .strict_inner_gas_wall = cfg->strict_inner_gas_wall};
  return p;
}

static int prepare(orc_problem *p) {
  if (p->halo == NULL)
    p->halo = orc_halo_plan_build(p->grid, p->active, p->ghost);
  if (p->edges == NULL)
    p->edges = orc_edge_plan_build(p->grid, p->active, p->ghost);
  /* the MI sample shell is read from the copied-in geometry
   * (cubedsphere_mi_dipole.c:157-221) */
  if (p->cfg.mi_enabled && p->mi == NULL &&
      orc_mi_create(p->grid, p->storage, p->field, &p->cfg.startup, p->active,
                    p->ghost, &p->cfg.mi, p->cfg.hall_siemens, &p->mi) != 0)
    return -1;
  p->boundary.mi = (const struct gamera_cs_mi_dipole *)p->mi;
  return p->halo != NULL && p->edges != NULL ? 0 : -1;
}

double orc_cfl_limit(orc_problem *p) {
  const size_t lo[3] = {p->ghost, p->ghost, p->ghost};
  const size_t hi[3] = {p->ghost + p->active[0], p->ghost + p->active[1],
                        p->ghost + p->active[2]};
  double limit = INFINITY;
  for (int f = 0; f < ORC_FACES; ++f) {
    double local;
    if (orc_local_timestep(&p->grid[f], p->storage[f].conserved,
                           p->storage[f].cell_magnetic, lo, hi, &p->timestep,
                           p->bg[f].cell_magnetic, &local) != 0)
      return NAN;
    limit = fmin(limit, local);
  }
  return limit;
}

int orc_step_explicit(orc_problem *p, double ratio, double dt, double time,
                      double previous_time) {
  if (p == NULL || prepare(p) != 0) return -1;
  p->boundary.step_dt = dt;
  p->boundary.step_time = time;
  const orc_step_context ctx = {p->halo, p->edges, &p->boundary, p->mi};
  return orc_advance_six(p->grid, p->storage, p->active, p->ghost, ratio, dt,
                         &p->options, p->field, &ctx, time, previous_time);
}

/* One iteration of the driver loop (:593-699): dt rule, MI cadence, step. */
int orc_advance(orc_problem *p, long count) {
  if (p == NULL || count < 0 || prepare(p) != 0) return -1;
  for (long n = 0; n < count; ++n) {
    const double limit = orc_cfl_limit(p);
    if (!(limit > 0.0)) return -1;
    p->dt = p->steps == 0 ? p->cfg.dt_fraction * limit
                          : fmin(p->cfg.dt_fraction * limit,
                                 1.25 * p->previous_dt);
    if (p->mi != NULL) {
      const int updated = orc_mi_maybe_update(p->mi, p->time);
      if (updated < 0) return -1;
      p->mi_updates += updated;
    }
    const double ratio = p->steps == 0 ? 0.0 : 0.5 * p->dt / p->previous_dt;
    if (orc_step_explicit(p, ratio, p->dt, p->time,
                          p->steps == 0 ? p->time
                                        : p->time - p->previous_dt) != 0)
      return -1;
    p->time += p->dt;
    p->previous_dt = p->dt;
    ++p->steps;
  }
  return 0;
}

void orc_clock(const orc_problem *p, double out[4]) {
  out[0] = p->time;
  out[1] = p->dt;
  out[2] = p->previous_dt;
  out[3] = (double)p->steps;
}

/* Replace the problem's wind series (solar_wind.c:55-80, gamera_solar_wind_set
 * without its Bx relation, as the series of the config): count samples,
 * state[8 n ..] = density, pressure, velocity, magnetic. */
int orc_set_wind(orc_problem *p, size_t count, const double *time,
                 const double *state, int linear) {
  if (p == NULL || count == 0 || time == NULL || state == NULL) return -1;
  for (size_t n = 0; n < count; ++n) {
    if (!isfinite(time[n]) || (n > 0 && !(time[n] > time[n - 1]))) return -1;
    for (int q = 0; q < 8; ++q)
      if (!isfinite(state[8 * n + q])) return -1;
  }
  double *t = malloc(count * sizeof(*t));
  gamera_solar_wind_state *w = malloc(count * sizeof(*w));
  if (t == NULL || w == NULL) {
    free(t);
    free(w);
    return -1;
  }
  for (size_t n = 0; n < count; ++n) {
    const double *x = &state[8 * n];
    t[n] = time[n];
    w[n].density = x[0];
    w[n].pressure = x[1];
    for (int c = 0; c < 3; ++c) {
      w[n].velocity[c] = x[2 + c];
      w[n].magnetic[c] = x[5 + c];
    }
  }
  free(p->series_time);
  free(p->series_state);
  p->series_time = t;
  p->series_state = w;
  p->wind.count = count;
  p->wind.time = t;
  p->wind.state = w;
  p->wind.linear_interpolation = count > 1 ? (linear != 0) : 1;
  return 0;
}

void orc_set_clock(orc_problem *p, double time, double dt, double previous_dt,
                   long steps) {
  p->time = time;
  p->dt = dt;
  p->previous_dt = previous_dt;
  p->steps = steps;
}

/* Field ids of include/gamera_cs_problem.h. */
int orc_field(orc_problem *p, int face, int field, double **data,
              size_t *count) {
  if (p == NULL || face < 0 || face >= ORC_FACES || data == NULL ||
      count == NULL)
    return -1;
  gamera_no_grid *g = &p->grid[face];
  gamera_no_storage *s = &p->storage[face];
  gamera_no_background_data *b = &p->bg[face];
  const size_t nc = orc_count(g->cell_extent);
  size_t nf[3], ne[3];
  for (int d = 0; d < 3; ++d) {
    nf[d] = orc_count(g->face[d].extent);
    ne[d] = orc_count(g->edge[d].extent);
  }
  if (field >= 0 && field <= 5) {
    double *a[6] = {s->conserved, s->old_conserved, s->predicted_conserved,
                    s->cell_magnetic->value, s->old_cell_magnetic->value,
                    s->predicted_cell_magnetic->value};
    *data = a[field];
    *count = nc * (field < 3 ? ORC_NQ : 3);
    return 0;
  }
  if (field >= GCS_FIELD_FACE_FLUX && field < GCS_FIELD_HYDRO_RATE) {
    const int group = (field - GCS_FIELD_FACE_FLUX) / 3;
    const int d = (field - GCS_FIELD_FACE_FLUX) % 3;
    switch (group) {
      case 0: *data = s->face_flux[d]; *count = nf[d]; return 0;
      case 1: *data = s->old_face_flux[d]; *count = nf[d]; return 0;
      case 2: *data = s->predicted_face_flux[d]; *count = nf[d]; return 0;
      case 3: *data = s->edge_emf[d]; *count = ne[d]; return 0;
      case 4: *data = s->fluid_face_flux[d]; *count = ORC_NQ * nf[d]; return 0;
      default: *data = s->maxwell_face_flux[d]->value; *count = 3 * nf[d];
               return 0;
    }
  }
  switch (field) {
    case GCS_FIELD_HYDRO_RATE: *data = s->hydro_rate; *count = ORC_NQ * nc; return 0;
    case GCS_FIELD_MAXWELL_RATE: *data = s->maxwell_rate->value; *count = 3 * nc; return 0;
    case GCS_FIELD_VERTEX: *data = g->vertex->value;
      *count = 3 * orc_count(g->vertex_extent); return 0;
    case GCS_FIELD_CELL_GEOMETRY: *data = &g->cell->volume;
      *count = nc * (sizeof(gamera_no_cell_geometry) / sizeof(double)); return 0;
    case GCS_FIELD_BG_CELL_MAGNETIC: *data = b->cell_magnetic->value; *count = 3 * nc; return 0;
    case GCS_FIELD_BG_CELL_FORCE: *data = b->cell_force->value; *count = 3 * nc; return 0;
    default: break;
  }
  if (field >= GCS_FIELD_FACE_GEOMETRY && field < GCS_FIELD_FACE_GEOMETRY + 3) {
    const int d = field - GCS_FIELD_FACE_GEOMETRY;
    *data = &g->face[d].value->area;
    *count = nf[d] * (sizeof(gamera_no_face_geometry) / sizeof(double));
    return 0;
  }
  if (field >= GCS_FIELD_EDGE_GEOMETRY && field < GCS_FIELD_EDGE_GEOMETRY + 3) {
    const int d = field - GCS_FIELD_EDGE_GEOMETRY;
    *data = &g->edge[d].value->length;
    *count = ne[d] * (sizeof(gamera_no_edge_geometry) / sizeof(double));
    return 0;
  }
  if (field >= GCS_FIELD_BG_FACE_MAGNETIC && field < GCS_FIELD_BG_FACE_FLUX) {
    const int d = field - GCS_FIELD_BG_FACE_MAGNETIC;
    *data = b->face_magnetic[d]->value;
    *count = 3 * nf[d];
    return 0;
  }
  if (field >= GCS_FIELD_BG_FACE_FLUX && field < GCS_FIELD_BG_EDGE_MAGNETIC) {
    const int d = field - GCS_FIELD_BG_FACE_FLUX;
    *data = b->face_flux[d];
    *count = nf[d];
    return 0;
  }
  if (field >= GCS_FIELD_BG_EDGE_MAGNETIC && field < GCS_FIELD_COUNT) {
    const int d = field - GCS_FIELD_BG_EDGE_MAGNETIC;
    *data = b->edge_magnetic[d]->value;
    *count = 3 * ne[d];
    return 0;
  }
  return -1;
}

int orc_edge_valid(orc_problem *p, int face, int direction,
                   unsigned char **data, size_t *count) {
  if (p == NULL || face < 0 || face >= ORC_FACES || direction < 0 ||
      direction > 2)
    return -1;
  *data = p->grid[face].edge[direction].valid;
  *count = orc_count(p->grid[face].edge[direction].extent);
  return 0;
}

void orc_counters(const orc_problem *p, int face, double out[2]) {
  out[0] = (double)p->storage[face].emergency_diffusion_face_count;
  out[1] = p->storage[face].emergency_diffusion_max_interface_speed;
}

/* Pointwise entry points for unit tests. */
int orc_unit_up7_pdm(const double vol[8], const double q[8], double pdm,
                     double out[2]) {
  return orc_up7_pdm(vol, q, pdm, &out[0], &out[1]);
}

int orc_mi_updates(const orc_problem *p) { return p ? p->mi_updates : 0; }

/* Both hemispheres' potential [hemi][colatitude][azimuth] in volts, like
 * gamera_cs_mi_dipole_export_potential. */
int orc_mi_potential_export(const orc_problem *p, double *out, size_t count) {
  if (p == NULL || p->mi == NULL || out == NULL || !orc_mi_ready(p->mi)) return -1;
  for (int h = 0; h < 2; ++h) {
    const double *v;
    size_t n;
    if (orc_mi_potential_values(p->mi, h, &v, &n) != 0 || count != 2 * n) return -1;
    memcpy(out + (size_t)h * n, v, n * sizeof(double));
  }
  return 0;
}

int orc_mi_fac_export(const orc_problem *p, double *out, size_t count) {
  if (p == NULL || p->mi == NULL || out == NULL || !orc_mi_ready(p->mi)) return -1;
  for (int h = 0; h < 2; ++h) {
    const double *v;
    size_t n;
    if (orc_mi_fac_values(p->mi, h, &v, &n) != 0 || count != 2 * n) return -1;
    memcpy(out + (size_t)h * n, v, n * sizeof(double));
  }
  return 0;
}

/* NO/nonorthogonal_step.c:74-171 (gamera_no_apply_pressure_control) on one
 * face's active cells, the counters and maxima accumulated into the
 * caller's. */
static int pressure_control_face(gamera_no_storage *storage,
                                 const gamera_no_grid *grid,
                                 const size_t lower[3], const size_t upper[3],
                                 double dt, double gamma, double density_floor,
                                 double pressure_floor, double light_speed,
                                 double low_density_threshold,
                                 double low_density_sound_speed,
                                 unsigned long long counts[2], double maxima[2]) {
  int failed = 0;
  for (size_t i = lower[0]; i < upper[0]; ++i)
    for (size_t j = lower[1]; j < upper[1]; ++j)
      for (size_t k = lower[2]; k < upper[2]; ++k) {
        const size_t cell = orc_idx(storage->cell_extent, i, j, k);
        double *conserved = &storage->conserved[cell * ORC_NQ];
        gamera_no_primitive primitive;
        if (orc_cons_to_prim(conserved, gamma, density_floor, pressure_floor,
                             &primitive) != 0) {
          failed = 1;
          continue;
        }
        int changed = 0;
        if (primitive.density <= low_density_threshold) {
          primitive.pressure =
              fmax(pressure_floor, primitive.density * low_density_sound_speed *
                                       low_density_sound_speed / gamma);
          ++counts[0];
          changed = 1;
        }
        const double sound_speed = sqrt(gamma * primitive.pressure / primitive.density);
        maxima[0] = fmax(maxima[0], sound_speed);
        maxima[1] = fmax(maxima[1], primitive.pressure);
        if (sound_speed > 1.5 * light_speed) {
          const gamera_no_vec3 point = grid->cell[cell].centroid;
          const double radius_square = point.value[0] * point.value[0] +
                                       point.value[1] * point.value[1] +
                                       point.value[2] * point.value[2];
          const double cylindrical_square =
              point.value[0] * point.value[0] + point.value[1] * point.value[1];
          if (!(radius_square > 0.0) || !(cylindrical_square > DBL_MIN)) {
            failed = 1;
            continue;
          }
          const double radius = sqrt(radius_square);
          const double dipole_l = radius * radius_square / cylindrical_square;
          const double target_pressure =
              primitive.density * (1.5 * light_speed) * (1.5 * light_speed) / gamma;
          const double bounce_time = dipole_l / sound_speed;
          primitive.pressure -= (dt / bounce_time) * (primitive.pressure - target_pressure);
          ++counts[1];
          changed = 1;
        }
        if (changed && orc_prim_to_cons(&primitive, gamma, density_floor, pressure_floor,
                                        conserved) != 0)
          failed = 1;
      }
  return failed ? -1 : 0;
}

/* The post-step pressure control of the MPI Earth driver
 * (cubedsphere/tests/test_cubedsphere_mpi_earth_tile.c:459-479) on every
 * face's active cells; counts = (low-density cells, high-sound-speed cells),
 * maxima = (sound speed, pressure) as the reference reports them. */
int orc_pressure_control(orc_problem *p, double dt, double gamma, double density_floor,
                         double pressure_floor, double light_speed,
                         double low_density_threshold, double low_density_sound_speed,
                         unsigned long long counts[2], double maxima[2]) {
  if (p == NULL || counts == NULL || maxima == NULL || !(dt > 0.0) || !(gamma > 1.0) ||
      !(density_floor > 0.0) || !(pressure_floor > 0.0) || !(light_speed > 0.0) ||
      !(low_density_threshold > 0.0) || !(low_density_sound_speed >= 0.0))
    return -1;
  counts[0] = counts[1] = 0;
  maxima[0] = maxima[1] = 0.0;
  const size_t lower[3] = {p->ghost, p->ghost, p->ghost};
  const size_t upper[3] = {p->ghost + p->active[0], p->ghost + p->active[1],
                           p->ghost + p->active[2]};
  int failed = 0;
  for (int f = 0; f < ORC_FACES; ++f)
    failed |= pressure_control_face(&p->storage[f], &p->grid[f], lower, upper, dt, gamma,
                                    density_floor, pressure_floor, light_speed,
                                    low_density_threshold, low_density_sound_speed, counts,
                                    maxima) != 0;
  return failed ? -1 : 0;
}

/* The serial Earth driver's six-face CSV snapshot
 * (cubedsphere/tests/test_cubedsphere_dipole_background.c:120-156,
 * write_earth_snapshot, without the MI files it appends): a header, then
 * per active cell in face, i, j, k order (k fastest) the active-relative
 * indices, the centroid, the primitive state with floors 1e-12 and the
 * total cell B (background + residual), %.17g. Returns -1 at the first cell
 * whose state does not convert, leaving the rows before it, as the
 * reference does. */
int orc_write_snapshot(orc_problem *p, const char *path) {
  if (p == NULL || path == NULL) return -1;
  FILE *out = fopen(path, "w");
  if (out == NULL) return -1;
  fputs("face,i,j,k,x,y,z,rho,p,vx,vy,vz,bx,by,bz\n", out);
  const double gamma = p->cfg.startup.gamma;
  const size_t g = p->ghost;
  for (int face = 0; face < ORC_FACES; ++face)
    for (size_t i = g; i < g + p->active[0]; ++i)
      for (size_t j = g; j < g + p->active[1]; ++j)
        for (size_t k = g; k < g + p->active[2]; ++k) {
          const size_t cell = orc_idx(p->grid[face].cell_extent, i, j, k);
          gamera_no_primitive w;
          if (orc_cons_to_prim(&p->storage[face].conserved[cell * ORC_NQ], gamma, 1e-12,
                               1e-12, &w) != 0) {
            fclose(out);
            return -1;
          }
          const gamera_no_vec3 x = p->grid[face].cell[cell].centroid;
          const gamera_no_vec3 b1 = p->storage[face].cell_magnetic[cell];
          const gamera_no_vec3 b0 = p->bg[face].cell_magnetic[cell];
          fprintf(out,
                  "%d,%zu,%zu,%zu,%.17g,%.17g,%.17g,%.17g,%.17g,"
                  "%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n",
                  face, i - g, j - g, k - g, x.value[0], x.value[1], x.value[2], w.density,
                  w.pressure, w.velocity.value[0], w.velocity.value[1], w.velocity.value[2],
                  b0.value[0] + b1.value[0], b0.value[1] + b1.value[1],
                  b0.value[2] + b1.value[2]);
        }
  return fclose(out) == 0 ? 0 : -1;
}
