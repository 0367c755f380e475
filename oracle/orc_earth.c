/* Solar-wind sampling and the Earth radial boundary (test infrastructure;
 * see orc.h). */
#include "orc.h"

#include <float.h>
#include <math.h>

static int wind_finite(const gamera_solar_wind_state *s) {
  if (!isfinite(s->density) || !isfinite(s->pressure) || s->density <= 0.0 ||
      s->pressure <= 0.0)
    return 0;
  for (int c = 0; c < 3; ++c)
    if (!isfinite(s->velocity[c]) || !isfinite(s->magnetic[c])) return 0;
  return fabs(s->velocity[GAMERA_WIND_X]) > DBL_MIN;
}

/* nonorthogonal/src/solar_wind.c:147-188 (sample_time, with its binary
 * search :173-186 and linear interpolation :188-204). */
static int wind_at_time(const gamera_solar_wind_series *w, double t,
                        gamera_solar_wind_state *out) {
  if (w->count == 0 || !isfinite(t)) return -1;
  if (w->count == 1 || t <= w->time[0]) {
    *out = w->state[0];
    return 0;
  }
  if (t >= w->time[w->count - 1]) {
    *out = w->state[w->count - 1];
    return 0;
  }
  size_t lo = 0, hi = w->count - 1;
  while (hi - lo > 1) {
    const size_t mid = lo + (hi - lo) / 2;
    if (w->time[mid] <= t) lo = mid;
    else hi = mid;
  }
  if (!w->linear_interpolation) {
    *out = w->state[lo];
    return 0;
  }
  const double f = (t - w->time[lo]) / (w->time[lo + 1] - w->time[lo]);
  const gamera_solar_wind_state *a = &w->state[lo], *b = &w->state[lo + 1];
  out->density = a->density + f * (b->density - a->density);
  out->pressure = a->pressure + f * (b->pressure - a->pressure);
  for (int c = 0; c < 3; ++c) {
    out->velocity[c] = a->velocity[c] + f * (b->velocity[c] - a->velocity[c]);
    out->magnetic[c] = a->magnetic[c] + f * (b->magnetic[c] - a->magnetic[c]);
  }
  return wind_finite(out) ? 0 : -1;
}

/* nonorthogonal/src/solar_wind.c:234-275: planar-front ballistic delay. */
int orc_wind_sample(const gamera_solar_wind_series *w, const double x[3],
                    double t, gamera_solar_wind_state *out) {
  if (!isfinite(t)) return -1;
  const double mt = t + w->time_offset;
  gamera_solar_wind_state cur;
  if (wind_at_time(w, mt, &cur) != 0) return -1;
  const double c2 = w->by_coefficient * w->by_coefficient +
                    w->bz_coefficient * w->bz_coefficient;
  const double factor = cur.velocity[GAMERA_WIND_X] / (1.0 + c2);
  const double fv[3] = {factor, -factor * w->by_coefficient,
                        -factor * w->bz_coefficient};
  double s2 = 0.0, proj = 0.0;
  for (int c = 0; c < 3; ++c) {
    if (!isfinite(x[c])) return -1;
    s2 += fv[c] * fv[c];
    proj += (x[c] - w->reference[c]) * fv[c];
  }
  if (!(s2 > DBL_MIN) || !isfinite(s2)) return -1;
  return wind_at_time(w, mt - proj / s2, out);
}

/* nonorthogonal/src/solar_wind.c:277-311: upstream-to-tail weight. */
int orc_wind_weight(const gamera_solar_wind_series *w, const double x[3],
                    double t, double *weight) {
  gamera_solar_wind_state s;
  if (orc_wind_sample(w, x, t, &s) != 0) return -1;
  double r2 = 0.0, v2 = 0.0, cosine = 0.0;
  for (int c = 0; c < 3; ++c) {
    r2 += x[c] * x[c];
    v2 += s.velocity[c] * s.velocity[c];
    cosine -= x[c] * s.velocity[c];
  }
  if (!(r2 > DBL_MIN) || !(v2 > DBL_MIN)) return -1;
  cosine /= sqrt(r2 * v2);
  cosine = fmax(-1.0, fmin(1.0, cosine));
  const double angle = acos(cosine);
  const double half_pi = 0.5 * acos(-1.0);
  if (angle <= half_pi) {
    *weight = 1.0;
  } else if (angle >= 2.0 * half_pi) {
    *weight = 0.0;
  } else {
    const double s1 = (angle - half_pi) / half_pi;
    *weight = exp(1.0 - 1.0 / (1.0 - s1 * s1));
  }
  return 0;
}

/* cubedsphere_earth_boundary.c:80-89 */
static int sample_wind(const gamera_cs_earth_boundary *b, const double x[3],
                       double t, gamera_solar_wind_state *s, double *w) {
  return orc_wind_sample(b->wind, x, t, s) == 0 &&
                 orc_wind_weight(b->wind, x, t, w) == 0
             ? 0 : -1;
}

/* cubedsphere_earth_boundary.c:27-112: radial ghost cells of one level. */
static int fill_cells(const gamera_no_grid *g, gamera_no_storage *s,
                      const size_t active[3], size_t ghost,
                      const gamera_cs_earth_boundary *b, double t, int old,
                      const orc_mi *mi) {
  const size_t outer = ghost + active[0];
  const size_t *e = g->cell_extent;
  double *u = old ? s->old_conserved : s->conserved;
  const double gamma = b->startup->gamma;
  int failed = 0;
  for (size_t i = 0; i < e[0]; ++i) {
    if (i >= ghost && i < outer) continue;
    const int inner = i < ghost;
    const size_t si = inner ? 2U * ghost - 1U - i : outer - 1U;
#pragma omp parallel for collapse(2) reduction(| : failed) schedule(static)
    for (size_t j = 0; j < e[1]; ++j)
      for (size_t k = 0; k < e[2]; ++k) {
        const size_t tc = orc_idx(e, i, j, k), sc = orc_idx(e, si, j, k);
        gamera_no_primitive w;
        if (orc_cons_to_prim(&u[sc * ORC_NQ], gamma, b->density_floor,
                             b->pressure_floor, &w) != 0) {
          failed = 1;
          continue;
        }
        const double *x = g->cell[tc].centroid.value;
        const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        if (!(r > 0.0)) {
          failed = 1;
          continue;
        }
        if (inner) {
          if (orc_mi_ready(mi)) {
            vec3 gv;
            if (orc_mi_ghost_velocity(mi, g->cell[tc].centroid, w.velocity,
                                      &gv) != 0) {
              failed = 1;
              continue;
            }
            if (b->strict_inner_gas_wall)
              for (int c = 0; c < 3; ++c)
                gv.value[c] = 0.5 * (gv.value[c] + w.velocity.value[c]);
            w.velocity = gv;
          } else if (b->strict_inner_gas_wall) {
            w.velocity = (vec3){{0.0, 0.0, 0.0}};
          } else {
            double vn = 0.0;
            for (int c = 0; c < 3; ++c) vn += w.velocity.value[c] * x[c] / r;
            for (int c = 0; c < 3; ++c) w.velocity.value[c] -= 2.0 * vn * x[c] / r;
          }
        } else {
          gamera_solar_wind_state sw;
          double wt;
          if (sample_wind(b, x, t, &sw, &wt) != 0) {
            failed = 1;
            continue;
          }
          double vn = 0.0;
          for (int c = 0; c < 3; ++c) vn += w.velocity.value[c] * x[c] / r;
          const double cs = sqrt(gamma * w.pressure / w.density);
          if (vn < cs)
            for (int c = 0; c < 3; ++c) w.velocity.value[c] += (cs - vn) * x[c] / r;
          w.density = wt * sw.density + (1.0 - wt) * w.density;
          w.pressure = wt * sw.pressure + (1.0 - wt) * w.pressure;
          for (int c = 0; c < 3; ++c)
            w.velocity.value[c] =
                wt * sw.velocity[c] + (1.0 - wt) * w.velocity.value[c];
        }
        if (orc_prim_to_cons(&w, gamma, b->density_floor, b->pressure_floor,
                             &u[tc * ORC_NQ]) != 0)
          failed = 1;
      }
  }
  return failed ? -1 : 0;
}

/* cubedsphere_earth_boundary.c:114-183: radial ghost face fluxes of one
 * level, then cell B of the whole block, then dB1/dr = 0 inside the wall. */
static int fill_faces(const gamera_no_grid *g, gamera_no_storage *s,
                      const size_t active[3], size_t ghost,
                      const gamera_cs_earth_boundary *b,
                      const gamera_no_background_field *bg, double t,
                      int old) {
  const size_t outer = ghost + active[0];
  int failed = 0;
  for (int d = 0; d < 3; ++d) {
    double *flux = old ? s->old_face_flux[d] : s->face_flux[d];
    const size_t *fe = g->face[d].extent;
    for (size_t i = 0; i < fe[0]; ++i) {
      const int inner = i < ghost;
      const int beyond = i >= outer + (d == 0);
      if (!inner && !beyond) continue;
#pragma omp parallel for collapse(2) reduction(| : failed) schedule(static)
      for (size_t j = 0; j < fe[1]; ++j)
        for (size_t k = 0; k < fe[2]; ++k) {
          const size_t f = orc_idx(fe, i, j, k);
          const gamera_no_face_geometry *geo = &g->face[d].value[f];
          if (inner) {
            const size_t src = orc_idx(fe, ghost, j, k);
            const double sa = g->face[d].value[src].area;
            if (!(sa > 0.0)) {
              failed = 1;
              continue;
            }
            flux[f] = flux[src] * geo->area / sa;
          } else {
            gamera_solar_wind_state sw;
            double wt;
            if (sample_wind(b, geo->centroid.value, t, &sw, &wt) != 0) {
              failed = 1;
              continue;
            }
            double wf = -bg->face_flux[d][f];
            for (int c = 0; c < 3; ++c) wf += sw.magnetic[c] * geo->area_vector.value[c];
            const size_t src = orc_idx(fe, d == 0 ? outer : outer - 1U, j, k);
            flux[f] = wt * wf + (1.0 - wt) * flux[src];
          }
        }
    }
  }
  if (failed) return -1;
  const double *phi[3] = {old ? s->old_face_flux[0] : s->face_flux[0],
                          old ? s->old_face_flux[1] : s->face_flux[1],
                          old ? s->old_face_flux[2] : s->face_flux[2]};
  vec3 *cb = old ? s->old_cell_magnetic : s->cell_magnetic;
  if (orc_recover_field(g, phi, cb) != 0) return -1;
  for (size_t i = 0; i < ghost; ++i)
    for (size_t j = 0; j < g->cell_extent[1]; ++j)
      for (size_t k = 0; k < g->cell_extent[2]; ++k)
        cb[orc_idx(g->cell_extent, i, j, k)] =
            cb[orc_idx(g->cell_extent, ghost, j, k)];
  return 0;
}

/* cubedsphere_earth_boundary.c:212-231: both time levels of every face. */
int orc_earth_ghosts(const gamera_no_grid grid[ORC_FACES],
                     gamera_no_storage st[ORC_FACES], const size_t active[3],
                     size_t ghost, double t_now, double t_prev,
                     const gamera_cs_earth_boundary *b, const void *mi) {
  if (!isfinite(t_now) || !isfinite(t_prev)) return -1;
  for (int f = 0; f < ORC_FACES; ++f)
    if (fill_cells(&grid[f], &st[f], active, ghost, b, t_now, 0, mi) != 0 ||
        fill_cells(&grid[f], &st[f], active, ghost, b, t_prev, 1, mi) != 0 ||
        fill_faces(&grid[f], &st[f], active, ghost, b, &b->background[f],
                   t_now, 0) != 0 ||
        fill_faces(&grid[f], &st[f], active, ghost, b, &b->background[f],
                   t_prev, 1) != 0)
      return -1;
  return 0;
}

/* NO/nonorthogonal_step.c:35-71 via cubedsphere_earth_boundary.c:259-266:
 * no outward mass (then energy) flux through the inner wall face. */
int orc_earth_inner_flux(gamera_no_storage *s, const size_t lo[3],
                         const size_t hi[3]) {
  const size_t i = lo[0];
  for (size_t j = lo[1]; j < hi[1]; ++j)
    for (size_t k = lo[2]; k < hi[2]; ++k) {
      double *f =
          &s->fluid_face_flux[0][orc_idx(s->face_extent[0], i, j, k) * ORC_NQ];
      if (f[0] > 0.0) {
        f[0] = 0.0;
        if (f[4] > 0.0) f[4] = 0.0;
      }
    }
  return 0;
}

/* cubedsphere_earth_boundary.c:332-345 */
static void edge_midpoint(const gamera_no_grid *g, int d, size_t i, size_t j,
                          size_t k, double out[3]) {
  size_t u3[3] = {i, j, k};
  ++u3[d];
  const double *a = g->vertex[orc_idx(g->vertex_extent, i, j, k)].value;
  const double *z = g->vertex[orc_idx(g->vertex_extent, u3[0], u3[1], u3[2])].value;
  for (int c = 0; c < 3; ++c) out[c] = 0.5 * (a[c] + z[c]);
}

/* cubedsphere_earth_boundary.c:347-360 */
static int flux_density(const gamera_no_grid *g, const double *flux, int d,
                        size_t i, size_t j, size_t k, double *out) {
  const size_t *fe = g->face[d].extent;
  if (i >= fe[0] || j >= fe[1] || k >= fe[2]) return -1;
  const size_t f = orc_idx(fe, i, j, k);
  const double area = g->face[d].value[f].area;
  if (!(area > 0.0)) return -1;
  *out = flux[f] / area;
  return isfinite(*out) ? 0 : -1;
}

/* cubedsphere_earth_boundary.c:362-443: diffusive wind EMF on the outer
 * wall edges, from the predicted face fluxes. */
static int wind_diffusive_emf(gamera_no_storage *s, const gamera_no_grid *g,
                              const size_t lo[3], const size_t hi[3],
                              const gamera_cs_earth_boundary *b) {
  const size_t i = hi[0];
  if (i == 0 || !(b->step_dt > 0.0) || !(b->cfl > 0.0)) return -1;
  const size_t *ej = g->edge[1].extent, *ek = g->edge[2].extent;
  double *const *pf = s->predicted_face_flux;
  for (size_t j = lo[1]; j <= hi[1]; ++j)
    for (size_t k = lo[2]; k <= hi[2]; ++k) {
      if (j == 0 || k == 0 || i >= ej[0] || j >= ej[1] || k >= ej[2] ||
          i >= ek[0] || j >= ek[1] || k >= ek[2])
        continue;
      const size_t e_j = orc_idx(ej, i, j, k), e_k = orc_idx(ek, i, j, k);
      if (g->edge[1].valid[e_j] == 0 || g->edge[2].valid[e_k] == 0) continue;
      double bikp, bikm, bkip, bkim, bjip, bjim, bijp, bijm;
      if (flux_density(g, pf[0], 0, i, j, k, &bikp) != 0 ||
          flux_density(g, pf[0], 0, i, j, k - 1, &bikm) != 0 ||
          flux_density(g, pf[2], 2, i, j, k, &bkip) != 0 ||
          flux_density(g, pf[2], 2, i - 1, j, k, &bkim) != 0 ||
          flux_density(g, pf[1], 1, i, j, k, &bjip) != 0 ||
          flux_density(g, pf[1], 1, i - 1, j, k, &bjim) != 0 ||
          flux_density(g, pf[0], 0, i, j, k, &bijp) != 0 ||
          flux_density(g, pf[0], 0, i, j - 1, k, &bijm) != 0)
        return -1;
      const double cj = (bikp - bikm) - (bkip - bkim);
      const double ck = (bjip - bjim) - (bijp - bijm);
      const double dl =
          sqrt(g->edge[1].value[e_j].length * g->edge[2].value[e_k].length);
      double xj[3], xk[3];
      edge_midpoint(g, 1, i, j, k, xj);
      edge_midpoint(g, 2, i, j, k, xk);
      gamera_solar_wind_state wj, wk;
      double gj, gk;
      const double t = b->step_time + 0.5 * b->step_dt;
      if (!(dl > 0.0) || sample_wind(b, xj, t, &wj, &gj) != 0 ||
          sample_wind(b, xk, t, &wk, &gk) != 0)
        return -1;
      const double c2 = b->wind->by_coefficient * b->wind->by_coefficient +
                        b->wind->bz_coefficient * b->wind->bz_coefficient;
      const double front =
          fmax(fabs(wj.velocity[0]), fabs(wk.velocity[0])) / sqrt(1.0 + c2);
      const double speed = fmin(front, b->cfl * dl / b->step_dt);
      s->edge_emf[1][e_j] += gj * speed * cj * dl;
      s->edge_emf[2][e_k] += gk * speed * ck * dl;
    }
  return 0;
}

/* cubedsphere_earth_boundary.c:381-451: outer wind EMF blend, diffusive
 * wind EMF, inner E_t = 0 (or the MI potential EMF). */
int orc_earth_emf(gamera_no_storage *s, const gamera_no_grid *g,
                  const size_t lo[3], const size_t hi[3],
                  const gamera_cs_earth_boundary *b, const void *mi) {
  if (!isfinite(b->step_time) || !(b->step_dt > 0.0)) return -1;
  const size_t outer = hi[0];
  const size_t first = outer > 3U ? outer - 3U : 0U;
  const double t = b->step_time + 0.5 * b->step_dt;
  int failed = 0;
  for (int d = 0; d < 3; ++d) {
    const size_t *ee = g->edge[d].extent;
    for (size_t i = first; i <= outer && i < ee[0]; ++i)
#pragma omp parallel for collapse(2) reduction(| : failed) schedule(static)
      for (size_t j = 0; j < ee[1]; ++j)
        for (size_t k = 0; k < ee[2]; ++k) {
          const size_t e = orc_idx(ee, i, j, k);
          if (g->edge[d].valid[e] == 0) continue;
          double x[3];
          edge_midpoint(g, d, i, j, k, x);
          gamera_solar_wind_state sw;
          double wt;
          if (sample_wind(b, x, t, &sw, &wt) != 0) {
            failed = 1;
            continue;
          }
          if (i < outer) wt /= 1.0 + (double)(outer - i);
          const double *v = sw.velocity, *m = sw.magnetic;
          const double el[3] = {-(v[1] * m[2] - v[2] * m[1]),
                                -(v[2] * m[0] - v[0] * m[2]),
                                -(v[0] * m[1] - v[1] * m[0])};
          const gamera_no_edge_geometry *geo = &g->edge[d].value[e];
          double line = 0.0;
          for (int c = 0; c < 3; ++c) line += el[c] * geo->normal.value[c] * geo->length;
          s->edge_emf[d][e] = (1.0 - wt) * s->edge_emf[d][e] + wt * line;
        }
  }
  if (failed || wind_diffusive_emf(s, g, lo, hi, b) != 0) return -1;
  for (int d = 1; d < 3; ++d)
    for (size_t j = 0; j < g->edge[d].extent[1]; ++j)
      for (size_t k = 0; k < g->edge[d].extent[2]; ++k) {
        const size_t e = orc_idx(g->edge[d].extent, lo[0], j, k);
        if (g->edge[d].valid[e] != 0) s->edge_emf[d][e] = 0.0;
      }
  if (orc_mi_ready(mi) && orc_mi_inner_emf(mi, s, g, lo[0]) != 0) return -1;
  return 0;
}
