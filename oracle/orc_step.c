/* Per-face sweeps, update and six-face orchestration (test infrastructure;
 * see orc.h). Every parallel loop writes disjoint outputs and reduces only
 * with |, +, min or max, so the thread schedule cannot change a result. */
#include "orc.h"

#include <float.h>
#include <math.h>
#include <string.h>

static double dot3(const double a[3], const double b[3]) {
  double r = 0.0;
  for (int d = 0; d < 3; ++d) r += a[d] * b[d];
  return r;
}

static const size_t *cell_ext(const gamera_no_grid *g) { return g->cell_extent; }

/* NO/nonorthogonal_step.c:310-359 over every cell of the block. */
int orc_recover_field(const gamera_no_grid *g, const double *const phi[3],
                      vec3 *cell_b) {
  const size_t *e = cell_ext(g);
  int failed = 0;
#pragma omp parallel for collapse(3) reduction(| : failed) schedule(static)
  for (size_t i = 0; i < e[0]; ++i)
    for (size_t j = 0; j < e[1]; ++j)
      for (size_t k = 0; k < e[2]; ++k) {
        double f[3][2];
        f[0][0] = phi[0][orc_idx(g->face[0].extent, i, j, k)];
        f[0][1] = phi[0][orc_idx(g->face[0].extent, i + 1, j, k)];
        f[1][0] = phi[1][orc_idx(g->face[1].extent, i, j, k)];
        f[1][1] = phi[1][orc_idx(g->face[1].extent, i, j + 1, k)];
        f[2][0] = phi[2][orc_idx(g->face[2].extent, i, j, k)];
        f[2][1] = phi[2][orc_idx(g->face[2].extent, i, j, k + 1)];
        const size_t c = orc_idx(e, i, j, k);
        if (orc_recover_cell(&g->cell[c], f, &cell_b[c]) != 0) failed = 1;
      }
  return failed ? -1 : 0;
}

/* NO/nonorthogonal_step.c:361-377 */
int orc_save_old(gamera_no_storage *s) {
  const size_t n = orc_count(s->cell_extent);
  memcpy(s->old_conserved, s->conserved, n * ORC_NQ * sizeof(double));
  memcpy(s->old_cell_magnetic, s->cell_magnetic, n * sizeof(vec3));
  for (int d = 0; d < 3; ++d)
    memcpy(s->old_face_flux[d], s->face_flux[d],
           orc_count(s->face_extent[d]) * sizeof(double));
  return 0;
}

/* NO/nonorthogonal_step.c:379-419 */
int orc_predict_storage(gamera_no_storage *s, const gamera_no_grid *g,
                        double r, double gamma, double rho_floor,
                        double p_floor) {
  if (!isfinite(r)) return -1;
  const size_t n = orc_count(s->cell_extent);
  int failed = 0;
#pragma omp parallel for reduction(| : failed) schedule(static)
  for (size_t c = 0; c < n; ++c)
    if (orc_predict_cell(&s->old_conserved[c * ORC_NQ],
                         &s->conserved[c * ORC_NQ], r, gamma, rho_floor,
                         p_floor, &s->predicted_conserved[c * ORC_NQ]) != 0)
      failed = 1;
  if (failed) return -1;
  for (int d = 0; d < 3; ++d) {
    const size_t m = orc_count(s->face_extent[d]);
#pragma omp parallel for schedule(static)
    for (size_t f = 0; f < m; ++f)
      s->predicted_face_flux[d][f] =
          s->face_flux[d][f] + r * (s->face_flux[d][f] - s->old_face_flux[d][f]);
  }
  const double *phi[3] = {s->predicted_face_flux[0], s->predicted_face_flux[1],
                          s->predicted_face_flux[2]};
  return orc_recover_field(g, phi, s->predicted_cell_magnetic);
}

/* NO/nonorthogonal_sweep.c:123-189: 8-cell stencil -> interface states. */
static int face_states(const gamera_no_grid *g, const double *u,
                       const vec3 *b, int dir, const size_t at[3],
                       const gamera_no_sweep_options *o,
                       gamera_no_primitive w[2], vec3 bb[2]) {
  double vol[8];
  gamera_no_primitive ws[8];
  vec3 bs[8];
  for (int n = 0; n < 8; ++n) {
    size_t c3[3] = {at[0], at[1], at[2]};
    c3[dir] = at[dir] + (size_t)n - 4U;
    const size_t c = orc_idx(g->cell_extent, c3[0], c3[1], c3[2]);
    vol[n] = g->cell[c].volume;
    if (orc_cons_to_prim(&u[c * ORC_NQ], o->gamma, o->density_floor,
                         o->pressure_floor, &ws[n]) != 0)
      return -1;
    if (o->use_mhd) bs[n] = b[c];
  }
  for (int q = 0; q < ORC_NQ; ++q) {
    double s[8], l, r;
    for (int n = 0; n < 8; ++n)
      s[n] = q == 0 ? ws[n].density
           : q == 4 ? ws[n].pressure : ws[n].velocity.value[q - 1];
    if (orc_up7_pdm(vol, s, o->pdm_coefficient, &l, &r) != 0) return -1;
    if (q == 0) { w[0].density = l; w[1].density = r; }
    else if (q == 4) { w[0].pressure = l; w[1].pressure = r; }
    else { w[0].velocity.value[q - 1] = l; w[1].velocity.value[q - 1] = r; }
  }
  if (o->use_mhd) {
    for (int c = 0; c < 3; ++c) {
      double s[8];
      for (int n = 0; n < 8; ++n) s[n] = bs[n].value[c];
      if (orc_up7_pdm(vol, s, o->pdm_coefficient, &bb[0].value[c],
                      &bb[1].value[c]) != 0)
        return -1;
    }
  } else {
    memset(bb, 0, 2 * sizeof(vec3));
  }
  return 0;
}

/* NO/nonorthogonal_sweep.c:408-548: fluid and Maxwell face fluxes. */
int orc_face_sweep(const gamera_no_grid *g, gamera_no_storage *s,
                   const double *u, const vec3 *b, const double *const phi[3],
                   int dir, const size_t lo[3], const size_t hi[3],
                   const gamera_no_sweep_options *o,
                   const gamera_no_background_field *bg) {
  if (lo[dir] < 4 || hi[dir] > g->cell_extent[dir] - 4) return -1;
  size_t top[3] = {hi[0], hi[1], hi[2]};
  ++top[dir];
  size_t emergency = 0;
  double emergency_speed = 0.0;
  int failed = 0;
  const gamera_no_face_grid *fg = &g->face[dir];
#pragma omp parallel for collapse(3) reduction(| : failed) \
    reduction(+ : emergency) reduction(max : emergency_speed) schedule(static)
  for (size_t i = lo[0]; i < top[0]; ++i)
    for (size_t j = lo[1]; j < top[1]; ++j)
      for (size_t k = lo[2]; k < top[2]; ++k) {
        const size_t at[3] = {i, j, k};
        const size_t f = orc_idx(fg->extent, i, j, k);
        const gamera_no_face_geometry *geo = &fg->value[f];
        gamera_no_primitive w[2];
        vec3 bb[2];
        if (face_states(g, u, b, dir, at, o, w, bb) != 0) { failed = 1; continue; }
        orc_fluid_flux fl;
        if (orc_fluid_flux_kinetic(w, o->gamma, geo, &fl) != 0) { failed = 1; continue; }
        orc_maxwell_flux mx = {{0.0, 0.0, 0.0}, 0.0, 0.0};
        if (o->use_mhd) {
          const double bn = phi[dir][f] / geo->area;
          vec3 b0 = {{0.0, 0.0, 0.0}};
          double b0n = 0.0;
          if (o->use_background) {
            b0 = bg->face_magnetic[dir][f];
            b0n = bg->face_flux[dir][f] / geo->area;
          }
          if (orc_maxwell_flux_kinetic(w, bb, bn, geo, o->use_background, b0,
                                       b0n, o->use_boris, o->light_speed,
                                       fl.normal_velocity, &mx) != 0)
            { failed = 1; continue; }
          if (o->use_hogs &&
              orc_hogs(&fl, &mx, o->hydro_hogs_coefficient,
                       o->magnetic_hogs_coefficient, o->use_boris,
                       o->light_speed) != 0)
            { failed = 1; continue; }
          if (o->use_hogs && o->use_boris) {
            size_t l3[3] = {i, j, k};
            --l3[dir];
            const size_t cl = orc_idx(g->cell_extent, l3[0], l3[1], l3[2]);
            const size_t cu = orc_idx(g->cell_extent, i, j, k);
            bool applied = false;
            if (orc_emergency_diffusion(&fl, w, &u[cl * ORC_NQ],
                                        &u[cu * ORC_NQ], true, o->light_speed,
                                        &applied) != 0)
              { failed = 1; continue; }
            if (applied) {
              double s2 = 0.0;
              for (int side = 0; side < 2; ++side)
                for (int c = 0; c < 3; ++c)
                  s2 += w[side].velocity.value[c] * w[side].velocity.value[c];
              emergency_speed = fmax(emergency_speed, sqrt(0.5 * s2));
              ++emergency;
            }
          }
        }
        for (int q = 0; q < ORC_NQ; ++q)
          s->fluid_face_flux[dir][f * ORC_NQ + q] = geo->area * fl.flux[q];
        for (int c = 0; c < 3; ++c)
          s->maxwell_face_flux[dir][f].value[c] = geo->area * mx.momentum[c];
      }
  s->emergency_diffusion_face_count += emergency;
  s->emergency_diffusion_max_interface_speed =
      fmax(s->emergency_diffusion_max_interface_speed, emergency_speed);
  return failed ? -1 : 0;
}

/* NO/nonorthogonal_sweep.c:223-259: central8 cell velocity at a face. */
static int face_velocity(const gamera_no_grid *g, const double *u, int dir,
                         const size_t at[3], double rho_floor, double v[3]) {
  double vol[8], wv[3][8];
  for (int n = 0; n < 8; ++n) {
    size_t c3[3] = {at[0], at[1], at[2]};
    c3[dir] = at[dir] + (size_t)n - 4U;
    const size_t c = orc_idx(g->cell_extent, c3[0], c3[1], c3[2]);
    vol[n] = g->cell[c].volume;
    const double rho = fmax(u[c * ORC_NQ], rho_floor);
    for (int d = 0; d < 3; ++d) wv[d][n] = vol[n] * u[c * ORC_NQ + 1 + d] / rho;
  }
  const double iv = orc_central8(vol);
  if (!isfinite(iv) || fabs(iv) <= DBL_MIN) return -1;
  for (int d = 0; d < 3; ++d) v[d] = orc_central8(wv[d]) / iv;
  return 0;
}

/* NO/nonorthogonal_sweep.c:309-349: area-weighted central8 of face
 * velocities along the second transverse direction. */
static int edge_velocity(const gamera_no_grid *g, const double *u, int t1,
                         int t2, const size_t at[3], double rho_floor,
                         double v[3]) {
  double area[8], fv[8][3];
  for (int n = 0; n < 8; ++n) {
    size_t f3[3] = {at[0], at[1], at[2]};
    f3[t2] = at[t2] + (size_t)n - 4U;
    const size_t f = orc_idx(g->face[t1].extent, f3[0], f3[1], f3[2]);
    area[n] = g->face[t1].value[f].area;
    if (face_velocity(g, u, t1, f3, rho_floor, fv[n]) != 0) return -1;
  }
  const double ea = orc_central8(area);
  if (!isfinite(ea) || fabs(ea) <= DBL_MIN) return -1;
  for (int d = 0; d < 3; ++d) {
    double w[8];
    for (int n = 0; n < 8; ++n) w[n] = area[n] * fv[n][d];
    v[d] = orc_central8(w) / ea;
  }
  return 0;
}

/* NO/nonorthogonal_sweep.c:351-370 */
static int edge_density(const gamera_no_grid *g, const double *u, int t1,
                        int t2, const size_t at[3], double *rho) {
  *rho = 0.0;
  for (size_t a = 0; a < 2; ++a)
    for (size_t b = 0; b < 2; ++b) {
      size_t c3[3] = {at[0], at[1], at[2]};
      c3[t1] -= a;
      c3[t2] -= b;
      *rho += 0.25 * u[orc_idx(g->cell_extent, c3[0], c3[1], c3[2]) * ORC_NQ];
    }
  return isfinite(*rho) && *rho > DBL_MIN ? 0 : -1;
}

/* NO/nonorthogonal_sweep.c:372-406 */
static int transverse_field(const gamera_no_grid *g, const double *phi,
                            int fdir, int sdir, const size_t at[3],
                            double pdm, double *l, double *r, double n2[2]) {
  double area[8], field[8];
  vec3 normal[8];
  for (int n = 0; n < 8; ++n) {
    size_t f3[3] = {at[0], at[1], at[2]};
    f3[sdir] = at[sdir] + (size_t)n - 4U;
    const size_t f = orc_idx(g->face[fdir].extent, f3[0], f3[1], f3[2]);
    const gamera_no_face_geometry *geo = &g->face[fdir].value[f];
    area[n] = geo->area;
    field[n] = phi[f] / geo->area;
    normal[n] = geo->normal;
  }
  const int edir = 3 - fdir - sdir;
  const size_t e = orc_idx(g->edge[edir].extent, at[0], at[1], at[2]);
  if (orc_up7_pdm(area, field, pdm, l, r) != 0 ||
      orc_face_normal_at_edge(area, normal, &g->edge[edir].value[e], n2) != 0)
    return -1;
  return 0;
}

/* NO/nonorthogonal_sweep.c:624-752: constrained-transport edge EMF. */
int orc_edge_sweep(const gamera_no_grid *g, gamera_no_storage *s,
                   const double *u, const double *const phi[3], int dir,
                   const size_t lo[3], const size_t hi[3],
                   const gamera_no_emf_options *o,
                   const gamera_no_background_field *bg) {
  const int t1 = (dir + 1) % 3, t2 = (dir + 2) % 3;
  if (lo[t1] < 4 || lo[t2] < 4 || hi[t1] > g->cell_extent[t1] - 4 ||
      hi[t2] > g->cell_extent[t2] - 4)
    return -1;
  size_t top[3] = {hi[0], hi[1], hi[2]};
  ++top[t1];
  ++top[t2];
  const gamera_no_edge_grid *eg = &g->edge[dir];
  int failed = 0;
#pragma omp parallel for collapse(3) reduction(| : failed) schedule(static)
  for (size_t i = lo[0]; i < top[0]; ++i)
    for (size_t j = lo[1]; j < top[1]; ++j)
      for (size_t k = lo[2]; k < top[2]; ++k) {
        const size_t at[3] = {i, j, k};
        const size_t e = orc_idx(eg->extent, i, j, k);
        if (eg->valid[e] == 0) { failed = 1; continue; }
        const gamera_no_edge_geometry *geo = &eg->value[e];
        double v[3];
        if (edge_velocity(g, u, t1, t2, at, o->density_floor, v) != 0)
          { failed = 1; continue; }
        const double v1 = dot3(v, geo->tangent1.value);
        const double v2 = dot3(v, geo->tangent2.value);
        double fl, fr, sl, sr, n1[2], n2[2];
        if (transverse_field(g, phi[t1], t1, t2, at, o->pdm_coefficient, &fl,
                             &fr, n1) != 0 ||
            transverse_field(g, phi[t2], t2, t1, at, o->pdm_coefficient, &sl,
                             &sr, n2) != 0)
          { failed = 1; continue; }
        double bt[2];
        if (orc_edge_field_solve(n1, n2, 0.5 * (fl + fr), 0.5 * (sl + sr),
                                 bt) != 0)
          { failed = 1; continue; }
        if (o->use_background) {
          const double *b0 = bg->edge_magnetic[dir][e].value;
          bt[0] += dot3(b0, geo->tangent1.value);
          bt[1] += dot3(b0, geo->tangent2.value);
        }
        const double jump = (sr - sl) - (fr - fl);
        double rho, speed;
        if (edge_density(g, u, t1, t2, at, &rho) != 0) { failed = 1; continue; }
        if (orc_edge_emf_formula(v1, v2, bt[0], bt[1], jump, rho, geo,
                                 o->diffusion_coefficient, o->use_boris,
                                 o->light_speed, o->cfl, o->dt,
                                 &s->edge_emf[dir][e], &speed) != 0)
          failed = 1;
      }
  return failed ? -1 : 0;
}

/* NO/nonorthogonal_sweep.c:550-622: flux divergence into cell rates. */
int orc_stress_rates(const gamera_no_grid *g, gamera_no_storage *s,
                     const size_t lo[3], const size_t hi[3], bool mhd,
                     const vec3 *force) {
#pragma omp parallel for collapse(3) schedule(static)
  for (size_t i = lo[0]; i < hi[0]; ++i)
    for (size_t j = lo[1]; j < hi[1]; ++j)
      for (size_t k = lo[2]; k < hi[2]; ++k) {
        const size_t c = orc_idx(g->cell_extent, i, j, k);
        const double iv = 1.0 / g->cell[c].volume;
        size_t flo[3], fup[3];
        for (int d = 0; d < 3; ++d) {
          size_t u3[3] = {i, j, k};
          ++u3[d];
          flo[d] = orc_idx(g->face[d].extent, i, j, k);
          fup[d] = orc_idx(g->face[d].extent, u3[0], u3[1], u3[2]);
        }
        for (int q = 0; q < ORC_NQ; ++q) {
          double diff = 0.0;
          for (int d = 0; d < 3; ++d)
            diff += s->fluid_face_flux[d][flo[d] * ORC_NQ + q] -
                    s->fluid_face_flux[d][fup[d] * ORC_NQ + q];
          s->hydro_rate[c * ORC_NQ + q] = diff * iv;
        }
        vec3 mr = {{0.0, 0.0, 0.0}};
        if (mhd)
          for (int comp = 0; comp < 3; ++comp) {
            double diff = 0.0;
            for (int d = 0; d < 3; ++d)
              diff += s->maxwell_face_flux[d][flo[d]].value[comp] -
                      s->maxwell_face_flux[d][fup[d]].value[comp];
            mr.value[comp] = diff * iv;
            if (force != NULL) mr.value[comp] += force[c].value[comp];
          }
        s->maxwell_rate[c] = mr;
      }
  return 0;
}

/* NO/nonorthogonal_step.c:233-298: CT on the faces bounding the box. */
static void advance_ct(gamera_no_storage *s, const size_t lo[3],
                       const size_t hi[3], double dt) {
  double *const *E = s->edge_emf;
  const size_t *e0 = s->edge_extent[0], *e1 = s->edge_extent[1],
               *e2 = s->edge_extent[2];
#pragma omp parallel for collapse(3) schedule(static)
  for (size_t i = lo[0]; i <= hi[0]; ++i)
    for (size_t j = lo[1]; j < hi[1]; ++j)
      for (size_t k = lo[2]; k < hi[2]; ++k) {
        const double curl = E[2][orc_idx(e2, i, j + 1, k)] -
                            E[2][orc_idx(e2, i, j, k)] -
                            E[1][orc_idx(e1, i, j, k + 1)] +
                            E[1][orc_idx(e1, i, j, k)];
        s->face_flux[0][orc_idx(s->face_extent[0], i, j, k)] -= dt * curl;
      }
#pragma omp parallel for collapse(3) schedule(static)
  for (size_t i = lo[0]; i < hi[0]; ++i)
    for (size_t j = lo[1]; j <= hi[1]; ++j)
      for (size_t k = lo[2]; k < hi[2]; ++k) {
        const double curl = E[0][orc_idx(e0, i, j, k + 1)] -
                            E[0][orc_idx(e0, i, j, k)] -
                            E[2][orc_idx(e2, i + 1, j, k)] +
                            E[2][orc_idx(e2, i, j, k)];
        s->face_flux[1][orc_idx(s->face_extent[1], i, j, k)] -= dt * curl;
      }
#pragma omp parallel for collapse(3) schedule(static)
  for (size_t i = lo[0]; i < hi[0]; ++i)
    for (size_t j = lo[1]; j < hi[1]; ++j)
      for (size_t k = lo[2]; k <= hi[2]; ++k) {
        const double curl = E[1][orc_idx(e1, i + 1, j, k)] -
                            E[1][orc_idx(e1, i, j, k)] -
                            E[0][orc_idx(e0, i, j + 1, k)] +
                            E[0][orc_idx(e0, i, j, k)];
        s->face_flux[2][orc_idx(s->face_extent[2], i, j, k)] -= dt * curl;
      }
}

/* NO/nonorthogonal_step.c:527-610 */
int orc_active_update(gamera_no_storage *s, const gamera_no_grid *g,
                      const size_t lo[3], const size_t hi[3], double dt,
                      const gamera_no_update_options *o, const vec3 *b0) {
  if (!isfinite(dt)) return -1;
  orc_save_old(s);
#pragma omp parallel for collapse(3) schedule(static)
  for (size_t i = lo[0]; i < hi[0]; ++i)
    for (size_t j = lo[1]; j < hi[1]; ++j)
      for (size_t k = lo[2]; k < hi[2]; ++k) {
        const size_t c = orc_idx(s->cell_extent, i, j, k);
        for (int q = 0; q < ORC_NQ; ++q)
          s->conserved[c * ORC_NQ + q] += dt * s->hydro_rate[c * ORC_NQ + q];
      }
  if (!o->use_mhd) return 0;
  advance_ct(s, lo, hi, dt);
  const double *phi[3] = {s->face_flux[0], s->face_flux[1], s->face_flux[2]};
  if (orc_recover_field(g, phi, s->cell_magnetic) != 0) return -1;
  int failed = 0;
#pragma omp parallel for collapse(3) reduction(| : failed) schedule(static)
  for (size_t i = lo[0]; i < hi[0]; ++i)
    for (size_t j = lo[1]; j < hi[1]; ++j)
      for (size_t k = lo[2]; k < hi[2]; ++k) {
        const size_t c = orc_idx(s->cell_extent, i, j, k);
        double *uc = &s->conserved[c * ORC_NQ];
        if (o->use_boris) {
          vec3 bn = s->cell_magnetic[c], bo = s->old_cell_magnetic[c];
          if (o->use_background)
            for (int d = 0; d < 3; ++d) {
              bn.value[d] = bn.value[d] + b0[c].value[d];
              bo.value[d] = bo.value[d] + b0[c].value[d];
            }
          if (orc_apply_boris(uc, &s->old_conserved[c * ORC_NQ], bn, bo,
                              &s->hydro_rate[c * ORC_NQ], s->maxwell_rate[c],
                              o->light_speed, dt, o->gamma, o->density_floor,
                              o->pressure_floor) != 0)
            failed = 1;
        } else if (orc_apply_maxwell(uc, s->maxwell_rate[c], dt, o->gamma,
                                     o->density_floor,
                                     o->pressure_floor) != 0) {
          failed = 1;
        }
      }
  return failed ? -1 : 0;
}

/* NO/nonorthogonal_step.c:449-525 */
int orc_local_timestep(const gamera_no_grid *g, const double *u,
                       const vec3 *b, const size_t lo[3], const size_t hi[3],
                       const gamera_no_timestep_options *o, const vec3 *b0,
                       double *dt) {
  double best = DBL_MAX;
  int failed = 0;
#pragma omp parallel for collapse(3) reduction(min : best) \
    reduction(| : failed) schedule(static)
  for (size_t i = lo[0]; i < hi[0]; ++i)
    for (size_t j = lo[1]; j < hi[1]; ++j)
      for (size_t k = lo[2]; k < hi[2]; ++k) {
        const size_t c = orc_idx(g->cell_extent, i, j, k);
        gamera_no_primitive w;
        if (orc_cons_to_prim(&u[c * ORC_NQ], o->gamma, o->density_floor,
                             o->pressure_floor, &w) != 0) {
          failed = 1;
          continue;
        }
        const double cs = sqrt(o->gamma * w.pressure / w.density);
        double va = 0.0;
        if (o->use_mhd) {
          double m[3];
          for (int d = 0; d < 3; ++d)
            m[d] = o->use_background ? b[c].value[d] + b0[c].value[d]
                                     : b[c].value[d];
          va = sqrt(dot3(m, m)) / sqrt(w.density);
          if (o->use_boris) va = o->light_speed * va / hypot(o->light_speed, va);
        }
        const double speed =
            sqrt(dot3(w.velocity.value, w.velocity.value)) + hypot(cs, va);
        double len = g->cell[c].cfl_length[0];
        for (int d = 1; d < 3; ++d) len = fmin(len, g->cell[c].cfl_length[d]);
        const double cand = o->cfl * len / speed;
        if (!isfinite(cand) || cand <= 0.0) {
          failed = 1;
          continue;
        }
        best = fmin(best, cand);
      }
  if (failed) return -1;
  *dt = best;
  return 0;
}

static int same_value(double a, double b) {
  const double scale = fmax(1.0, fmax(fabs(a), fabs(b)));
  return isfinite(a) && isfinite(b) && fabs(a - b) <= 8e-15 * scale;
}

/* cubedsphere/cubedsphere_step.c:13-154 with the Earth callbacks
 * (cubedsphere_earth_boundary.c:212-220, :259-266, :446-451) in place of the
 * caller's function pointers when ctx->earth is set. */
int orc_advance_six(const gamera_no_grid grid[ORC_FACES],
                    gamera_no_storage st[ORC_FACES], const size_t active[3],
                    size_t ghost, double ratio, double dt,
                    const gamera_no_advance_options *o,
                    const gamera_no_background_field bg[ORC_FACES],
                    const orc_step_context *ctx, double t_now, double t_prev) {
  if (ghost < 4 || !isfinite(dt) || dt <= 0.0 || !isfinite(ratio) ||
      active[1] != active[2] || o->cell_source != NULL ||
      o->indexed_cell_source != NULL ||
      o->update.use_background != o->stress.use_background ||
      o->update.use_background != o->emf.use_background ||
      (o->update.use_background && bg == NULL) || !isfinite(o->source_time) ||
      !same_value(o->stress.gamma, o->update.gamma) ||
      !same_value(o->stress.density_floor, o->update.density_floor) ||
      !same_value(o->stress.pressure_floor, o->update.pressure_floor) ||
      !same_value(o->emf.density_floor, o->update.density_floor) ||
      o->stress.use_mhd != o->update.use_mhd ||
      o->stress.use_boris != o->update.use_boris ||
      o->emf.use_boris != o->update.use_boris ||
      (o->update.use_boris &&
       (!same_value(o->stress.light_speed, o->update.light_speed) ||
        !same_value(o->emf.light_speed, o->update.light_speed))))
    return -1;
  const size_t lo[3] = {ghost, ghost, ghost};
  const size_t hi[3] = {ghost + active[0], ghost + active[1],
                        ghost + active[2]};
  if (orc_sync_faces(st, active, ghost, 1) != 0 ||
      orc_fill_cell_halos(ctx->halo, st) != 0 ||
      orc_radial_outflow(st, active, ghost) != 0 ||
      orc_fill_face_halos(ctx->halo, grid, st) != 0)
    return -1;
  if (ctx->earth != NULL &&
      orc_earth_ghosts(grid, st, active, ghost, t_now, t_prev, ctx->earth,
                       ctx->mi) != 0)
    return -1;
  for (int f = 0; f < ORC_FACES; ++f)
    if (orc_predict_storage(&st[f], &grid[f], ratio, o->update.gamma,
                            o->update.density_floor,
                            o->update.pressure_floor) != 0)
      return -1;
  for (int f = 0; f < ORC_FACES; ++f) {
    const gamera_no_background_field *b =
        o->update.use_background ? &bg[f] : NULL;
    const double *phi[3] = {st[f].predicted_face_flux[0],
                            st[f].predicted_face_flux[1],
                            st[f].predicted_face_flux[2]};
    for (int d = 0; d < 3; ++d)
      if (orc_face_sweep(&grid[f], &st[f], st[f].predicted_conserved,
                         st[f].predicted_cell_magnetic, phi, d, lo, hi,
                         &o->stress, b) != 0)
        return -1;
  }
  if (orc_sync_faces(st, active, ghost, 0) != 0) return -1;
  if (ctx->earth != NULL)
    for (int f = 0; f < ORC_FACES; ++f)
      if (orc_earth_inner_flux(&st[f], lo, hi) != 0) return -1;
  if (o->update.use_mhd) {
    gamera_no_emf_options emf = o->emf;
    emf.dt = dt;
    for (int f = 0; f < ORC_FACES; ++f) {
      const gamera_no_background_field *b =
          o->update.use_background ? &bg[f] : NULL;
      const double *phi[3] = {st[f].predicted_face_flux[0],
                              st[f].predicted_face_flux[1],
                              st[f].predicted_face_flux[2]};
      for (int d = 0; d < 3; ++d)
        if (orc_edge_sweep(&grid[f], &st[f], st[f].predicted_conserved, phi,
                           d, lo, hi, &emf, b) != 0)
          return -1;
    }
    if (ctx->earth != NULL)
      for (int f = 0; f < ORC_FACES; ++f)
        if (orc_earth_emf(&st[f], &grid[f], lo, hi, ctx->earth, ctx->mi) != 0)
          return -1;
    if (orc_sync_edges(ctx->edges, st) != 0) return -1;
  }
  for (int f = 0; f < ORC_FACES; ++f) {
    const gamera_no_background_field *b =
        o->update.use_background ? &bg[f] : NULL;
    if (orc_stress_rates(&grid[f], &st[f], lo, hi, o->update.use_mhd,
                         b ? b->cell_force : NULL) != 0 ||
        orc_active_update(&st[f], &grid[f], lo, hi, dt, &o->update,
                          b ? b->cell_magnetic : NULL) != 0)
      return -1;
  }
  if (o->update.use_mhd) {
    if (orc_sync_faces(st, active, ghost, 1) != 0) return -1;
    for (int f = 0; f < ORC_FACES; ++f) {
      const double *phi[3] = {st[f].face_flux[0], st[f].face_flux[1],
                              st[f].face_flux[2]};
      if (orc_recover_field(&grid[f], phi, st[f].cell_magnetic) != 0)
        return -1;
    }
  }
  return 0;
}
