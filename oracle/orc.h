/*
 * CPU oracle: a plain-C99 restatement of the reference six-face step.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls
 * this code; tests/, __graft_entry__.smoke() and the CPU-baseline leg of
 * bench.py use it as the checker of the HIP step.
 *
 * Every function cites the reference function it restates (file:line, paths
 * relative to the reference root; NO/ = nonorthogonal/src/nonorthogonal/).
 * Expressions keep the reference's evaluation order and grouping so that,
 * compiled like the reference (gcc -O3, no FMA contraction), the oracle is
 * bitwise the reference; tests/test_oracle_vs_reference.py checks that.
 */
#ifndef ORC_H
#define ORC_H

#include "gamera_cs_problem.h"

#include <stdbool.h>
#include <stddef.h>

#define ORC_NQ GAMERA_NO_FLUX_COUNT
#define ORC_FACES GAMERA_CS_FACE_COUNT

typedef gamera_no_vec3 vec3;

static inline size_t orc_idx(const size_t e[3], size_t i, size_t j, size_t k) {
  return (i * e[1] + j) * e[2] + k;
}
static inline size_t orc_count(const size_t e[3]) { return e[0] * e[1] * e[2]; }

/* ---- orc_numerics.c: pointwise kernels ---------------------------------- */
int orc_cons_to_prim(const double u[ORC_NQ], double gamma, double rho_floor,
                     double p_floor, gamera_no_primitive *w);
int orc_prim_to_cons(const gamera_no_primitive *w, double gamma,
                     double rho_floor, double p_floor, double u[ORC_NQ]);
int orc_predict_cell(const double old_u[ORC_NQ], const double u[ORC_NQ],
                     double ratio, double gamma, double rho_floor,
                     double p_floor, double out[ORC_NQ]);
int orc_apply_boris(double u[ORC_NQ], const double old_u[ORC_NQ],
                    vec3 b_new, vec3 b_old, const double rate[ORC_NQ],
                    vec3 magnetic_rate, double c, double dt, double gamma,
                    double rho_floor, double p_floor);
int orc_apply_maxwell(double u[ORC_NQ], vec3 magnetic_rate, double dt,
                      double gamma, double rho_floor, double p_floor);
double orc_central8(const double s[8]);
double orc_central6(const double s[8]);
int orc_up7_pdm(const double volume[8], const double q[8], double pdm,
                double *left, double *right);

typedef struct {
  double flux[ORC_NQ];
  double jump[ORC_NQ];
  double velocity_jump[3];
  double normal_velocity[2];
} orc_fluid_flux;
typedef struct {
  double momentum[3];
  double alfven_speed;
  double pressure_sum;
} orc_maxwell_flux;

int orc_fluid_flux_kinetic(const gamera_no_primitive w[2], double gamma,
                           const gamera_no_face_geometry *face,
                           orc_fluid_flux *out);
int orc_maxwell_flux_kinetic(const gamera_no_primitive w[2], const vec3 b[2],
                             double bn, const gamera_no_face_geometry *face,
                             bool use_bg, vec3 b0, double b0n, bool boris,
                             double c, const double un[2],
                             orc_maxwell_flux *out);
int orc_hogs(orc_fluid_flux *f, orc_maxwell_flux *m, double ch, double cm,
             bool boris, double c);
int orc_emergency_diffusion(orc_fluid_flux *f, const gamera_no_primitive w[2],
                            const double lower[ORC_NQ],
                            const double upper[ORC_NQ], bool boris, double c,
                            bool *applied);
int orc_edge_field_solve(const double n1[2], const double n2[2], double f1,
                         double f2, double out[2]);
int orc_face_normal_at_edge(const double area[8], const vec3 normal[8],
                            const gamera_no_edge_geometry *edge,
                            double out[2]);
int orc_edge_emf_formula(double v1, double v2, double b1, double b2,
                         double jump, double rho,
                         const gamera_no_edge_geometry *edge, double diff,
                         bool boris, double c, double cfl, double dt,
                         double *emf, double *speed);
int orc_recover_cell(const gamera_no_cell_geometry *g,
                     const double flux[3][2], vec3 *b);

/* ---- orc_step.c: per-face sweeps and the six-face orchestration ---------- */
int orc_recover_field(const gamera_no_grid *g, const double *const flux[3],
                      vec3 *cell_b);
int orc_save_old(gamera_no_storage *s);
int orc_predict_storage(gamera_no_storage *s, const gamera_no_grid *g,
                        double ratio, double gamma, double rho_floor,
                        double p_floor);
int orc_face_sweep(const gamera_no_grid *g, gamera_no_storage *s,
                   const double *u, const vec3 *b, const double *const phi[3],
                   int dir, const size_t lo[3], const size_t hi[3],
                   const gamera_no_sweep_options *o,
                   const gamera_no_background_field *bg);
int orc_edge_sweep(const gamera_no_grid *g, gamera_no_storage *s,
                   const double *u, const double *const phi[3], int dir,
                   const size_t lo[3], const size_t hi[3],
                   const gamera_no_emf_options *o,
                   const gamera_no_background_field *bg);
int orc_stress_rates(const gamera_no_grid *g, gamera_no_storage *s,
                     const size_t lo[3], const size_t hi[3], bool mhd,
                     const vec3 *force);
int orc_active_update(gamera_no_storage *s, const gamera_no_grid *g,
                      const size_t lo[3], const size_t hi[3], double dt,
                      const gamera_no_update_options *o, const vec3 *b0);
int orc_local_timestep(const gamera_no_grid *g, const double *u,
                       const vec3 *b, const size_t lo[3], const size_t hi[3],
                       const gamera_no_timestep_options *o, const vec3 *b0,
                       double *dt);

/* ---- orc_halo.c: cube topology, ghost fill, seam reconciliation --------- */
int orc_cube_inverse(vec3 direction, int *face, double *a, double *b);
int orc_cube_neighbor(int face, int edge, int *nface, int *nedge,
                      int *reversed);
typedef struct orc_halo_plan orc_halo_plan;
orc_halo_plan *orc_halo_plan_build(const gamera_no_grid grid[ORC_FACES],
                                   const size_t active[3], size_t ghost);
void orc_halo_plan_free(orc_halo_plan *plan);
int orc_fill_cell_halos(const orc_halo_plan *plan, gamera_no_storage *st);
int orc_fill_face_halos(const orc_halo_plan *plan,
                        const gamera_no_grid grid[ORC_FACES],
                        gamera_no_storage *st);
int orc_radial_outflow(gamera_no_storage *st, const size_t active[3],
                       size_t ghost);
int orc_sync_faces(gamera_no_storage *st, const size_t active[3], size_t ghost,
                   int magnetic);
typedef struct orc_edge_plan orc_edge_plan;
orc_edge_plan *orc_edge_plan_build(const gamera_no_grid grid[ORC_FACES],
                                   const size_t active[3], size_t ghost);
void orc_edge_plan_free(orc_edge_plan *plan);
int orc_sync_edges(const orc_edge_plan *plan, gamera_no_storage *st);

/* ---- orc_earth.c: solar wind and the Earth radial boundary -------------- */
int orc_wind_sample(const gamera_solar_wind_series *w, const double x[3],
                    double t, gamera_solar_wind_state *out);
int orc_wind_weight(const gamera_solar_wind_series *w, const double x[3],
                    double t, double *weight);
int orc_earth_ghosts(const gamera_no_grid grid[ORC_FACES],
                     gamera_no_storage st[ORC_FACES], const size_t active[3],
                     size_t ghost, double t_now, double t_prev,
                     const gamera_cs_earth_boundary *b, const void *mi);
int orc_earth_inner_flux(gamera_no_storage *s, const size_t lo[3],
                         const size_t hi[3]);
int orc_earth_emf(gamera_no_storage *s, const gamera_no_grid *g,
                  const size_t lo[3], const size_t hi[3],
                  const gamera_cs_earth_boundary *b, const void *mi);

/* ---- orc_mi.c: ionosphere coupling --------------------------------------- */
int orc_mi_compose_ghost_velocity(vec3 source, vec3 b0, vec3 drift, vec3 *ghost);
int orc_mi_cell_current(const gamera_no_grid *g, const vec3 *b, size_t i,
                        size_t j, size_t k, vec3 *current);
double orc_mi_mapped_colatitude(double theta, double r, double ri);
double orc_mi_field_ratio(double theta, double r, double ri);
double orc_mi_cos_inclination(double theta, int hemisphere);
int orc_mi_apply_pedersen(const gamera_mi_solver_config *c, const double *x,
                          double *y);
int orc_mi_solve_pedersen(const gamera_mi_solver_config *c, const double *fac,
                          double *potential, gamera_mi_solver_stats *stats);
int orc_mi_apply_tensor(const gamera_mi_solver_config *c, const double *sp,
                        const double *sh, const double *x, double *y);
int orc_mi_solve_tensor(const gamera_mi_solver_config *c, const double *fac,
                        const double *sp, const double *sh, double *potential,
                        gamera_mi_solver_stats *stats);

typedef struct orc_mi orc_mi;
int orc_mi_create(const gamera_no_grid grid[ORC_FACES],
                  gamera_no_storage st[ORC_FACES],
                  const gamera_no_background_field bg[ORC_FACES],
                  const gamera_cs_earth_startup *startup, const size_t active[3],
                  size_t ghost, const gamera_cs_mi_dipole_config *cfg,
                  double hall_siemens, orc_mi **out);
void orc_mi_destroy(orc_mi *mi);
int orc_mi_update_due(const orc_mi *mi, double time_code);
int orc_mi_maybe_update(orc_mi *mi, double time_code);
int orc_mi_ready(const orc_mi *mi);
int orc_mi_potential(const orc_mi *mi, vec3 point, double *potential_v);
int orc_mi_ghost_velocity(const orc_mi *mi, vec3 point, vec3 source,
                          vec3 *ghost_velocity);
int orc_mi_inner_emf(const orc_mi *mi, gamera_no_storage *s,
                     const gamera_no_grid *g, size_t wall_i);
int orc_mi_potential_values(const orc_mi *mi, int hemi, const double **values,
                            size_t *count);
int orc_mi_fac_values(const orc_mi *mi, int hemi, const double **values,
                      size_t *count);
int orc_mi_solve_stats(const orc_mi *mi, int hemi, gamera_mi_solver_stats *out);

/* ---- the six-face step (cubedsphere_step.c:13-154, Earth callbacks) ------ */
typedef struct {
  const orc_halo_plan *halo;
  const orc_edge_plan *edges;
  const gamera_cs_earth_boundary *earth;  /* NULL: plain outflow problem */
  const orc_mi *mi;
} orc_step_context;

int orc_advance_six(const gamera_no_grid grid[ORC_FACES],
                    gamera_no_storage st[ORC_FACES], const size_t active[3],
                    size_t ghost, double ratio, double dt,
                    const gamera_no_advance_options *o,
                    const gamera_no_background_field bg[ORC_FACES],
                    const orc_step_context *ctx, double t_now, double t_prev);

#endif
