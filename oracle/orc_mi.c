/* Ionosphere (MI) coupling of the Earth boundary: field-aligned current from
 * the residual field on a shell, the polar potential solve, and the potential
 * at the inner wall (test infrastructure; see orc.h).
 *
 * Restates NO/nonorthogonal_mi.c and cubedsphere/cubedsphere_mi_dipole.c.
 * Sums keep the reference's sequential order, so the solvers' iterates, and
 * therefore their iteration counts, are the reference's bit for bit. */
#include "orc.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { MI_NORTH = 0, MI_SOUTH = 1, MI_HEMIS = 2 };

static double dot3(vec3 a, vec3 b) {
  double r = 0.0;
  for (int c = 0; c < 3; ++c) r += a.value[c] * b.value[c];
  return r;
}
static double norm3(vec3 a) { return sqrt(dot3(a, a)); }

/* ---- pointwise MI numerics (NO/nonorthogonal_mi.c) ----------------------- */

/* :20-41 */
int orc_mi_compose_ghost_velocity(vec3 source, vec3 b0, vec3 drift,
                                  vec3 *ghost) {
  if (ghost == NULL) return -1;
  const double b2 = dot3(b0, b0);
  if (!(b2 > 0.0) || !isfinite(b2)) return -1;
  for (int c = 0; c < 3; ++c) {
    ghost->value[c] = 2.0 * drift.value[c] - source.value[c];
    if (!isfinite(ghost->value[c])) return -1;
  }
  return 0;
}

/* :43-95: line integral of the four-cell average of B along one edge. */
static int edge_integral(const gamera_no_grid *g, const vec3 *b, int dir,
                         size_t i, size_t j, size_t k, double *out) {
  const size_t *ee = g->edge[dir].extent;
  if (i >= ee[0] || j >= ee[1] || k >= ee[2]) return -1;
  const size_t e = orc_idx(ee, i, j, k);
  if (!g->edge[dir].valid[e]) return -1;
  /* the four cells around the edge: (0,0), (-1,0), (0,-1), (-1,-1) over the
   * two transverse axes, in the reference's order */
  const int t1 = dir == 0 ? 1 : 0, t2 = dir == 2 ? 1 : 2;
  size_t c[4][3];
  for (int n = 0; n < 4; ++n) {
    c[n][0] = i;
    c[n][1] = j;
    c[n][2] = k;
  }
  --c[1][t1];
  --c[2][t2];
  --c[3][t1];
  --c[3][t2];
  vec3 avg = {{0.0, 0.0, 0.0}};
  for (int n = 0; n < 4; ++n) {
    for (int a = 0; a < 3; ++a)
      if (c[n][a] >= g->cell_extent[a]) return -1;
    const vec3 v = b[orc_idx(g->cell_extent, c[n][0], c[n][1], c[n][2])];
    for (int m = 0; m < 3; ++m) avg.value[m] += 0.25 * v.value[m];
  }
  const gamera_no_edge_geometry *eg = &g->edge[dir].value[e];
  *out = dot3(avg, eg->normal) * eg->length;
  return isfinite(*out) ? 0 : -1;
}

/* :97-136: circulation of B around one face. */
static int face_circulation(const gamera_no_grid *g, const vec3 *b, int dir,
                            size_t i, size_t j, size_t k, double *out) {
  double e1, e2, e3, e4;
  int bad;
  if (dir == 0)
    bad = edge_integral(g, b, 1, i, j, k, &e1) ||
          edge_integral(g, b, 2, i, j + 1, k, &e2) ||
          edge_integral(g, b, 1, i, j, k + 1, &e3) ||
          edge_integral(g, b, 2, i, j, k, &e4);
  else if (dir == 1)
    bad = edge_integral(g, b, 2, i, j, k, &e1) ||
          edge_integral(g, b, 0, i, j, k + 1, &e2) ||
          edge_integral(g, b, 2, i + 1, j, k, &e3) ||
          edge_integral(g, b, 0, i, j, k, &e4);
  else
    bad = edge_integral(g, b, 0, i, j, k, &e1) ||
          edge_integral(g, b, 1, i + 1, j, k, &e2) ||
          edge_integral(g, b, 0, i, j + 1, k, &e3) ||
          edge_integral(g, b, 1, i, j, k, &e4);
  if (bad) return -1;
  *out = e1 + e2 - e3 - e4;
  return isfinite(*out) ? 0 : -1;
}

/* :138-167: cell current J = curl B from the six face circulations, through
 * the calibrated flux-to-cell recovery (cubedsphere_recovery.c). */
int orc_mi_cell_current(const gamera_no_grid *g, const vec3 *b, size_t i,
                        size_t j, size_t k, vec3 *current) {
  if (g == NULL || b == NULL || current == NULL || i == 0 || j == 0 ||
      k == 0 || i + 1 >= g->cell_extent[0] || j + 1 >= g->cell_extent[1] ||
      k + 1 >= g->cell_extent[2])
    return -1;
  double flux[3][2];
  if (face_circulation(g, b, 0, i, j, k, &flux[0][0]) ||
      face_circulation(g, b, 0, i + 1, j, k, &flux[0][1]) ||
      face_circulation(g, b, 1, i, j, k, &flux[1][0]) ||
      face_circulation(g, b, 1, i, j + 1, k, &flux[1][1]) ||
      face_circulation(g, b, 2, i, j, k, &flux[2][0]) ||
      face_circulation(g, b, 2, i, j, k + 1, &flux[2][1]))
    return -1;
  return orc_recover_cell(&g->cell[orc_idx(g->cell_extent, i, j, k)],
                          (const double(*)[2])flux, current);
}

/* :169-181 */
double orc_mi_mapped_colatitude(double theta, double r, double ri) {
  if (!isfinite(theta) || !isfinite(r) || !isfinite(ri) || r <= 0.0 ||
      ri <= 0.0 || ri > r)
    return NAN;
  double s = sin(theta) * sqrt(ri / r);
  s = fmax(-1.0, fmin(1.0, s));
  return asin(s);
}

/* :183-196 */
double orc_mi_field_ratio(double theta, double r, double ri) {
  const double ti = orc_mi_mapped_colatitude(theta, r, ri);
  if (!isfinite(ti)) return NAN;
  const double q = r / ri;
  const double top = sqrt(1.0 + 3.0 * pow(cos(ti), 2.0));
  const double bottom = sqrt(1.0 + 3.0 * pow(cos(theta), 2.0));
  return q * q * q * top / bottom;
}

/* :198-207 */
double orc_mi_cos_inclination(double theta, int hemisphere) {
  if (!isfinite(theta) || theta < 0.0 || theta > 0.5 * acos(-1.0) ||
      (hemisphere != GAMERA_MI_NORTH && hemisphere != GAMERA_MI_SOUTH))
    return NAN;
  const double c = cos(theta);
  return (double)hemisphere * 2.0 * c / sqrt(1.0 + 3.0 * c * c);
}

/* ---- polar solvers ------------------------------------------------------- */

typedef gamera_mi_solver_config solver_cfg;

/* :209-228 */
static int valid_solver(const solver_cfg *c) {
  return c != NULL && c->longitude_count >= 4 && c->colatitude_count >= 3 &&
         isfinite(c->maximum_colatitude) && c->maximum_colatitude > 0.0 &&
         c->maximum_colatitude <= 0.5 * acos(-1.0) &&
         isfinite(c->ionosphere_radius_m) && c->ionosphere_radius_m > 0.0 &&
         isfinite(c->pedersen_siemens) && c->pedersen_siemens > 0.0 &&
         c->hall_siemens == 0.0 && isfinite(c->low_latitude_potential_v) &&
         (c->hemisphere == GAMERA_MI_NORTH || c->hemisphere == GAMERA_MI_SOUTH) &&
         c->maximum_iterations > 0 && isfinite(c->relative_tolerance) &&
         c->relative_tolerance > 0.0 && isfinite(c->absolute_tolerance) &&
         c->absolute_tolerance >= 0.0;
}

static size_t full_at(const solver_cfg *c, size_t i, size_t j) {
  return i * c->longitude_count + j;
}
/* reduced system: one pole unknown, then rings 1..nt-2 (:305-308) */
static size_t red_at(const solver_cfg *c, size_t i, size_t j) {
  return 1 + (i - 1) * c->longitude_count + j;
}

static double dtheta_of(const solver_cfg *c) {
  return c->maximum_colatitude / (double)(c->colatitude_count - 1);
}
static double dphi_of(const solver_cfg *c) {
  return 2.0 * acos(-1.0) / (double)c->longitude_count;
}

/* :234-241: Pedersen coefficient of a theta half-face, over sin(theta). */
static double theta_coeff(const solver_cfg *c, double half, double center,
                          double dt) {
  const double cs = cos(half);
  const double ci2 = 4.0 * cs * cs / (1.0 + 3.0 * cs * cs);
  return sin(half) * c->pedersen_siemens / (ci2 * sin(center) * dt * dt);
}

/* :243-282 */
int orc_mi_apply_pedersen(const solver_cfg *c, const double *x, double *y) {
  if (!valid_solver(c) || x == NULL || y == NULL) return -1;
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  const double dt = dtheta_of(c), dp = dphi_of(c);
  double ring = 0.0;
  for (size_t j = 0; j < np; ++j) ring += x[full_at(c, 1, j)] / (double)np;
  for (size_t j = 0; j < np; ++j) y[full_at(c, 0, j)] = x[full_at(c, 0, j)] - ring;
  for (size_t i = 1; i + 1 < nt; ++i) {
    const double th = (double)i * dt;
    const double lo = theta_coeff(c, th - 0.5 * dt, th, dt);
    const double hi = theta_coeff(c, th + 0.5 * dt, th, dt);
    const double lon = c->pedersen_siemens / (sin(th) * sin(th) * dp * dp);
    for (size_t j = 0; j < np; ++j) {
      const size_t jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
      const size_t at = full_at(c, i, j);
      y[at] = lo * (x[at] - x[full_at(c, i - 1, j)]) +
              hi * (x[at] - x[full_at(c, i + 1, j)]) +
              lon * (2.0 * x[at] - x[full_at(c, i, jm)] - x[full_at(c, i, jp)]);
    }
  }
  for (size_t j = 0; j < np; ++j)
    y[full_at(c, nt - 1, j)] = x[full_at(c, nt - 1, j)];
  return 0;
}

static double vdot(const double *a, const double *b, size_t n) {
  double r = 0.0;
  for (size_t i = 0; i < n; ++i) r += a[i] * b[i];
  return r;
}

/* :310-320 */
static double pole_coeff(const solver_cfg *c) {
  const double dt = dtheta_of(c);
  const double half = 0.5 * dt;
  const double cs = cos(half);
  const double ci2 = 4.0 * cs * cs / (1.0 + 3.0 * cs * cs);
  return sin(half) * c->pedersen_siemens / (ci2 * dt * dt);
}

/* :322-344 */
static void reduced_diagonal(const solver_cfg *c, double *d) {
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  const double dt = dtheta_of(c), dp = dphi_of(c);
  d[0] = (double)np * pole_coeff(c);
  for (size_t i = 1; i + 1 < nt; ++i) {
    const double th = (double)i * dt;
    const double w = sin(th);
    const double lo = theta_coeff(c, th - 0.5 * dt, th, dt);
    const double hi = theta_coeff(c, th + 0.5 * dt, th, dt);
    const double lon = c->pedersen_siemens / (sin(th) * sin(th) * dp * dp);
    for (size_t j = 0; j < np; ++j) d[red_at(c, i, j)] = w * (lo + hi + 2.0 * lon);
  }
}

/* :355-392: the sin(theta)-weighted SPD operator on the reduced unknowns. */
static void reduced_apply(const solver_cfg *c, const double *x, double *y) {
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  const double dt = dtheta_of(c), dp = dphi_of(c);
  const double pe = pole_coeff(c);
  double ring = 0.0;
  for (size_t j = 0; j < np; ++j) ring += x[red_at(c, 1, j)];
  y[0] = pe * ((double)np * x[0] - ring);
  for (size_t i = 1; i + 1 < nt; ++i) {
    const double th = (double)i * dt;
    const double w = sin(th);
    const double lo = theta_coeff(c, th - 0.5 * dt, th, dt);
    const double hi = theta_coeff(c, th + 0.5 * dt, th, dt);
    const double lon = c->pedersen_siemens / (sin(th) * sin(th) * dp * dp);
    for (size_t j = 0; j < np; ++j) {
      const size_t jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
      const size_t at = red_at(c, i, j);
      const double below = i == 1 ? x[0] : x[red_at(c, i - 1, j)];
      const double above = i + 2 < nt ? x[red_at(c, i + 1, j)] : 0.0;
      y[at] = w * (lo * (x[at] - below) + hi * (x[at] - above) +
                   lon * (2.0 * x[at] - x[red_at(c, i, jm)] - x[red_at(c, i, jp)]));
    }
  }
}

/* :394-410: residual norm of the unweighted (original) equations. */
static double original_norm(const solver_cfg *c, const double *r) {
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  const double dt = dtheta_of(c);
  const double ps = (double)np * pole_coeff(c);
  const double pr = r[0] / ps;
  double s = (double)np * pr * pr;
  for (size_t i = 1; i + 1 < nt; ++i) {
    const double w = sin((double)i * dt);
    for (size_t j = 0; j < np; ++j) {
      const double v = r[red_at(c, i, j)] / w;
      s += v * v;
    }
  }
  return sqrt(fmax(0.0, s));
}

static void expand(const solver_cfg *c, const double *x, double boundary,
                   double *potential) {
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  for (size_t j = 0; j < np; ++j) {
    potential[full_at(c, 0, j)] = x[0];
    potential[full_at(c, nt - 1, j)] = boundary;
  }
  for (size_t i = 1; i + 1 < nt; ++i)
    for (size_t j = 0; j < np; ++j)
      potential[full_at(c, i, j)] = x[red_at(c, i, j)];
}

/* Right-hand side and initial guess shared by both solvers (:428-470). */
static double setup_rhs(const solver_cfg *c, const double *fac, double *rhs,
                        int with_boundary, int *bad) {
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  const double dt = dtheta_of(c);
  const double r2 = c->ionosphere_radius_m * c->ionosphere_radius_m;
  double norm2 = with_boundary ? (double)np * c->low_latitude_potential_v *
                                     c->low_latitude_potential_v
                               : 0.0;
  *bad = 0;
  for (size_t i = 1; i + 1 < nt; ++i) {
    const double th = (double)i * dt;
    const double w = sin(th);
    const double inc = orc_mi_cos_inclination(th, c->hemisphere);
    const double hi = with_boundary ? theta_coeff(c, th + 0.5 * dt, th, dt) : 0.0;
    for (size_t j = 0; j < np; ++j) {
      const size_t fa = full_at(c, i, j), ra = red_at(c, i, j);
      if (!isfinite(fac[fa])) {
        *bad = 1;
        return 0.0;
      }
      const double u = -r2 * fac[fa] * inc;
      rhs[ra] = w * u;
      norm2 += u * u;
      if (with_boundary && i + 2 == nt) rhs[ra] += w * hi * c->low_latitude_potential_v;
    }
  }
  return norm2;
}

static void initial_guess(const solver_cfg *c, const double *potential,
                          double *x) {
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  double pole = 0.0;
  for (size_t j = 0; j < np; ++j) {
    const double v = potential[full_at(c, 0, j)];
    if (isfinite(v)) pole += v / (double)np;
  }
  x[0] = pole;
  for (size_t i = 1; i + 1 < nt; ++i)
    for (size_t j = 0; j < np; ++j) {
      const double v = potential[full_at(c, i, j)];
      x[red_at(c, i, j)] = isfinite(v) ? v : 0.0;
    }
}

/* :412-609: Jacobi-preconditioned CG on the reduced SPD system, with
 * true-residual checks (and restarts every 1024 iterations). */
int orc_mi_solve_pedersen(const solver_cfg *c, const double *fac,
                          double *potential, gamera_mi_solver_stats *stats) {
  if (!valid_solver(c) || fac == NULL || potential == NULL) return -1;
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  if (np > (SIZE_MAX - 1) / (nt - 2)) return -1;
  const size_t n = 1 + np * (nt - 2);
  double *work = calloc(7 * n, sizeof(double));
  if (work == NULL) return -1;
  double *rhs = work, *diag = rhs + n, *x = diag + n, *r = x + n, *z = r + n,
         *p = z + n, *ap = p + n;
  int bad;
  const double rhs2 = setup_rhs(c, fac, rhs, 1, &bad);
  if (bad) {
    free(work);
    return -1;
  }
  initial_guess(c, potential, x);
  reduced_diagonal(c, diag);
  reduced_apply(c, x, ap);
  for (size_t q = 0; q < n; ++q) r[q] = rhs[q] - ap[q];
  const double first = original_norm(c, r);
  const double target =
      fmax(c->absolute_tolerance,
           c->relative_tolerance * fmax(sqrt(fmax(0.0, rhs2)), 1.0));
  if (stats) *stats = (gamera_mi_solver_stats){0, first, first, first <= target};
  if (first <= target) {
    expand(c, x, c->low_latitude_potential_v, potential);
    free(work);
    return 0;
  }
  for (size_t q = 0; q < n; ++q) z[q] = r[q] / diag[q];
  memcpy(p, z, n * sizeof(double));
  double rz = vdot(r, z, n);
  double last = first;
  int iters = 0, ok = 0;
  for (int it = 1; it <= c->maximum_iterations; ++it) {
    if (!isfinite(rz) || rz <= DBL_MIN) break;
    reduced_apply(c, p, ap);
    const double pap = vdot(p, ap, n);
    if (!isfinite(pap) || pap <= DBL_MIN) break;
    const double alpha = rz / pap;
    for (size_t q = 0; q < n; ++q) {
      x[q] += alpha * p[q];
      r[q] -= alpha * ap[q];
    }
    last = original_norm(c, r);
    iters = it;
    if (last <= target || it % 1024 == 0) {
      reduced_apply(c, x, ap);
      for (size_t q = 0; q < n; ++q) r[q] = rhs[q] - ap[q];
      last = original_norm(c, r);
      if (last <= target) {
        ok = 1;
        break;
      }
      for (size_t q = 0; q < n; ++q) z[q] = r[q] / diag[q];
      memcpy(p, z, n * sizeof(double));
      rz = vdot(r, z, n);
      continue;
    }
    for (size_t q = 0; q < n; ++q) z[q] = r[q] / diag[q];
    const double rz_new = vdot(r, z, n);
    if (!isfinite(rz_new) || rz_new <= DBL_MIN) break;
    const double beta = rz_new / rz;
    for (size_t q = 0; q < n; ++q) p[q] = z[q] + beta * p[q];
    rz = rz_new;
  }
  expand(c, x, c->low_latitude_potential_v, potential);
  if (stats) *stats = (gamera_mi_solver_stats){iters, first, last, ok};
  free(work);
  return ok ? 0 : -1;
}

/* ---- Pedersen + Hall tensor (:611-1151) ---------------------------------- */

/* :611-639 */
static int valid_tensor(const solver_cfg *c, const double *sp, const double *sh) {
  if (c == NULL || sp == NULL || sh == NULL || c->longitude_count < 4 ||
      c->colatitude_count < 3 || !isfinite(c->maximum_colatitude) ||
      !(c->maximum_colatitude > 0.0) || c->maximum_colatitude > 0.5 * acos(-1.0) ||
      !isfinite(c->ionosphere_radius_m) || !(c->ionosphere_radius_m > 0.0) ||
      !isfinite(c->low_latitude_potential_v) ||
      fabs(c->low_latitude_potential_v) > DBL_MIN ||
      (c->hemisphere != GAMERA_MI_NORTH && c->hemisphere != GAMERA_MI_SOUTH) ||
      c->maximum_iterations <= 0 || !isfinite(c->relative_tolerance) ||
      !(c->relative_tolerance > 0.0) || !isfinite(c->absolute_tolerance) ||
      c->absolute_tolerance < 0.0 ||
      c->longitude_count > SIZE_MAX / c->colatitude_count)
    return 0;
  const size_t n = c->longitude_count * c->colatitude_count;
  for (size_t q = 0; q < n; ++q)
    if (!isfinite(sp[q]) || !(sp[q] > 0.0) || !isfinite(sh[q]) || sh[q] < 0.0)
      return 0;
  return 1;
}

static double incl2(double th) {
  const double cs = cos(th);
  return 4.0 * cs * cs / (1.0 + 3.0 * cs * cs);
}

/* :646-653: the Hall off-diagonal entry -SigmaH / cos(I). */
static double f12(const solver_cfg *c, const double *sh, size_t i, size_t j) {
  const double inc = orc_mi_cos_inclination((double)i * dtheta_of(c), c->hemisphere);
  return -sh[full_at(c, i, j)] / inc;
}

/* :655-733 */
int orc_mi_apply_tensor(const solver_cfg *c, const double *sp, const double *sh,
                        const double *x, double *y) {
  if (!valid_tensor(c, sp, sh) || x == NULL || y == NULL) return -1;
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  const double dt = dtheta_of(c), dp = dphi_of(c);
  double ring = 0.0;
  for (size_t j = 0; j < np; ++j) ring += x[full_at(c, 1, j)] / (double)np;
  for (size_t j = 0; j < np; ++j) y[full_at(c, 0, j)] = x[full_at(c, 0, j)] - ring;
  for (size_t i = 1; i + 1 < nt; ++i) {
    const double th = (double)i * dt;
    const double s = sin(th);
    const double lt = th - 0.5 * dt, ut = th + 0.5 * dt;
    for (size_t j = 0; j < np; ++j) {
      const size_t jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
      const size_t at = full_at(c, i, j);
      const double lo = sin(lt) * 0.5 * (sp[at] + sp[full_at(c, i - 1, j)]) /
                        (incl2(lt) * s * dt * dt);
      const double hi = sin(ut) * 0.5 * (sp[at] + sp[full_at(c, i + 1, j)]) /
                        (incl2(ut) * s * dt * dt);
      const double llo = 0.5 * (sp[at] + sp[full_at(c, i, jm)]) / (s * s * dp * dp);
      const double lhi = 0.5 * (sp[at] + sp[full_at(c, i, jp)]) / (s * s * dp * dp);
      const double ped = lo * (x[at] - x[full_at(c, i - 1, j)]) +
                         hi * (x[at] - x[full_at(c, i + 1, j)]) +
                         llo * (x[at] - x[full_at(c, i, jm)]) +
                         lhi * (x[at] - x[full_at(c, i, jp)]);
      const double dft = (f12(c, sh, i + 1, j) - f12(c, sh, i - 1, j)) / (2.0 * dt);
      const double dfp = (f12(c, sh, i, jp) - f12(c, sh, i, jm)) / (2.0 * dp);
      const double dxt = (x[full_at(c, i + 1, j)] - x[full_at(c, i - 1, j)]) / (2.0 * dt);
      const double dxp = (x[full_at(c, i, jp)] - x[full_at(c, i, jm)]) / (2.0 * dp);
      const double hall = (-dft * dxp + dfp * dxt) / s;
      y[at] = ped + hall;
    }
  }
  for (size_t j = 0; j < np; ++j)
    y[full_at(c, nt - 1, j)] = x[full_at(c, nt - 1, j)];
  return 0;
}

typedef struct {
  const solver_cfg *c;
  double *diag, *lo, *hi, *llo, *lhi, *pole, *fwd;
  double *mem;
} tensor_stencil;

/* :735-833: the reduced five-point stencil (sin(theta)-weighted rows). */
static int tensor_build(const solver_cfg *c, const double *sp, const double *sh,
                        double *diag, tensor_stencil *t) {
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  const size_t n = 1 + np * (nt - 2);
  if (n > (SIZE_MAX - np) / 5) return -1;
  double *mem = calloc(5 * n + np, sizeof(double));
  if (mem == NULL) return -1;
  *t = (tensor_stencil){c, diag, mem, mem + n, mem + 2 * n, mem + 3 * n,
                        mem + 4 * n, mem + 4 * n + np, mem};
  const double dt = dtheta_of(c), dp = dphi_of(c);
  double pole_mean = 0.0;
  for (size_t j = 0; j < np; ++j) pole_mean += sp[full_at(c, 0, j)] / (double)np;
  diag[0] = 0.0;
  for (size_t j = 0; j < np; ++j) {
    const double half = 0.5 * dt;
    const double face = 0.5 * (pole_mean + sp[full_at(c, 1, j)]);
    t->pole[j] = sin(half) * face / (incl2(half) * dt * dt);
    diag[0] += t->pole[j];
  }
  for (size_t i = 1; i + 1 < nt; ++i) {
    const double th = (double)i * dt;
    const double s = sin(th);
    const double lt = th - 0.5 * dt, ut = th + 0.5 * dt;
    for (size_t j = 0; j < np; ++j) {
      const size_t jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
      const size_t fa = full_at(c, i, j);
      const double lo = sin(lt) * 0.5 * (sp[fa] + sp[full_at(c, i - 1, j)]) /
                        (incl2(lt) * dt * dt);
      const double hi = sin(ut) * 0.5 * (sp[fa] + sp[full_at(c, i + 1, j)]) /
                        (incl2(ut) * dt * dt);
      const double llo = 0.5 * (sp[fa] + sp[full_at(c, i, jm)]) / (s * dp * dp);
      const double lhi = 0.5 * (sp[fa] + sp[full_at(c, i, jp)]) / (s * dp * dp);
      const size_t at = red_at(c, i, j);
      diag[at] = lo + hi + llo + lhi;
      const double dft = (f12(c, sh, i + 1, j) - f12(c, sh, i - 1, j)) / (2.0 * dt);
      const double dfp = (f12(c, sh, i, jp) - f12(c, sh, i, jm)) / (2.0 * dp);
      t->lo[at] = -lo - dfp / (2.0 * dt);
      t->hi[at] = -hi + dfp / (2.0 * dt);
      t->llo[at] = -llo + dft / (2.0 * dp);
      t->lhi[at] = -lhi - dft / (2.0 * dp);
    }
  }
  return 0;
}

/* :871-900 */
static void tensor_apply(const tensor_stencil *t, const double *x, double *y) {
  const solver_cfg *c = t->c;
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  y[0] = t->diag[0] * x[0];
  for (size_t j = 0; j < np; ++j) y[0] -= t->pole[j] * x[red_at(c, 1, j)];
  for (size_t i = 1; i + 1 < nt; ++i)
    for (size_t j = 0; j < np; ++j) {
      const size_t jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
      const size_t at = red_at(c, i, j);
      const double below = i == 1 ? x[0] : x[red_at(c, i - 1, j)];
      const double above = i + 2 < nt ? x[red_at(c, i + 1, j)] : 0.0;
      y[at] = t->diag[at] * x[at] + t->lo[at] * below + t->hi[at] * above +
              t->llo[at] * x[red_at(c, i, jm)] + t->lhi[at] * x[red_at(c, i, jp)];
    }
}

/* :902-957: one symmetric Gauss-Seidel (SSOR, omega = 1) sweep pair in the
 * reference's ordering: rows outward, longitudes ascending; then back. */
static void tensor_precondition(tensor_stencil *t, const double *in, double *out) {
  const solver_cfg *c = t->c;
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  double *f = t->fwd;
  f[0] = in[0] / t->diag[0];
  for (size_t i = 1; i + 1 < nt; ++i)
    for (size_t j = 0; j < np; ++j) {
      const size_t at = red_at(c, i, j);
      double v = in[at];
      const size_t below = i == 1 ? 0 : red_at(c, i - 1, j);
      v -= t->lo[at] * f[below];
      if (j > 0) v -= t->llo[at] * f[red_at(c, i, j - 1)];
      if (j + 1 == np) v -= t->lhi[at] * f[red_at(c, i, 0)];
      f[at] = v / t->diag[at];
    }
  for (size_t i = nt - 2; i > 0; --i)
    for (size_t jr = np; jr > 0; --jr) {
      const size_t j = jr - 1;
      const size_t at = red_at(c, i, j);
      double v = t->diag[at] * f[at];
      if (i + 2 < nt) v -= t->hi[at] * out[red_at(c, i + 1, j)];
      if (j + 1 < np) v -= t->lhi[at] * out[red_at(c, i, j + 1)];
      if (j == 0) v -= t->llo[at] * out[red_at(c, i, np - 1)];
      out[at] = v / t->diag[at];
    }
  double pole = t->diag[0] * f[0];
  for (size_t j = 0; j < np; ++j) pole += t->pole[j] * out[red_at(c, 1, j)];
  out[0] = pole / t->diag[0];
}

/* :959-1151: right-preconditioned BiCGStab with true-residual restarts
 * every 256 iterations. */
int orc_mi_solve_tensor(const solver_cfg *c, const double *fac, const double *sp,
                        const double *sh, double *potential,
                        gamera_mi_solver_stats *stats) {
  if (!valid_tensor(c, sp, sh) || fac == NULL || potential == NULL) return -1;
  const size_t np = c->longitude_count, nt = c->colatitude_count;
  if (np > (SIZE_MAX - 1) / (nt - 2)) return -1;
  const size_t n = 1 + np * (nt - 2);
  double *work = calloc(11 * n, sizeof(double));
  if (work == NULL) return -1;
  double *rhs = work, *diag = rhs + n, *x = diag + n, *r = x + n, *sh0 = r + n,
         *p = sh0 + n, *ap = p + n, *pp = ap + n, *s = pp + n, *ps = s + n,
         *as = ps + n;
  int bad;
  const double rhs2 = setup_rhs(c, fac, rhs, 0, &bad);
  if (bad) {
    free(work);
    return -1;
  }
  initial_guess(c, potential, x);
  tensor_stencil t = {0};
  if (tensor_build(c, sp, sh, diag, &t) != 0) {
    free(work);
    return -1;
  }
  tensor_apply(&t, x, ap);
  for (size_t q = 0; q < n; ++q) {
    r[q] = rhs[q] - ap[q];
    sh0[q] = r[q];
  }
  const double first = original_norm(c, r);
  const double target =
      fmax(c->absolute_tolerance,
           c->relative_tolerance * fmax(sqrt(fmax(0.0, rhs2)), 1.0));
  if (stats) *stats = (gamera_mi_solver_stats){0, first, first, first <= target};
  double last = first;
  int iters = 0, ok = first <= target;
  double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  memset(p, 0, n * sizeof(double));
  memset(ap, 0, n * sizeof(double));
  for (int it = 1; !ok && it <= c->maximum_iterations; ++it) {
    const double rho = vdot(sh0, r, n);
    if (!isfinite(rho) || fabs(rho) <= DBL_MIN || !isfinite(omega) ||
        fabs(omega) <= DBL_MIN)
      break;
    const double beta = (rho / rho_prev) * (alpha / omega);
    for (size_t q = 0; q < n; ++q) p[q] = r[q] + beta * (p[q] - omega * ap[q]);
    tensor_precondition(&t, p, pp);
    tensor_apply(&t, pp, ap);
    const double den = vdot(sh0, ap, n);
    if (!isfinite(den) || fabs(den) <= DBL_MIN) break;
    alpha = rho / den;
    for (size_t q = 0; q < n; ++q) s[q] = r[q] - alpha * ap[q];
    last = original_norm(c, s);
    if (last <= target) {
      for (size_t q = 0; q < n; ++q) x[q] += alpha * pp[q];
      iters = it;
      ok = 1;
      break;
    }
    tensor_precondition(&t, s, ps);
    tensor_apply(&t, ps, as);
    const double wden = vdot(as, as, n);
    if (!isfinite(wden) || wden <= DBL_MIN) break;
    omega = vdot(as, s, n) / wden;
    if (!isfinite(omega) || fabs(omega) <= DBL_MIN) break;
    for (size_t q = 0; q < n; ++q) {
      x[q] += alpha * pp[q] + omega * ps[q];
      r[q] = s[q] - omega * as[q];
    }
    last = original_norm(c, r);
    iters = it;
    if (last <= target || it % 256 == 0) {
      tensor_apply(&t, x, as);
      for (size_t q = 0; q < n; ++q) r[q] = rhs[q] - as[q];
      last = original_norm(c, r);
      if (last <= target) {
        ok = 1;
        break;
      }
      memcpy(sh0, r, n * sizeof(double));
      memset(p, 0, n * sizeof(double));
      memset(ap, 0, n * sizeof(double));
      rho_prev = 1.0;
      alpha = 1.0;
      omega = 1.0;
      continue;
    }
    rho_prev = rho;
  }
  expand(c, x, 0.0, potential);
  if (stats) *stats = (gamera_mi_solver_stats){iters, first, last, ok};
  free(t.mem);
  free(work);
  return ok ? 0 : -1;
}

/* ---- the cube adapter (cubedsphere/cubedsphere_mi_dipole.c) -------------- */

struct orc_mi {
  const gamera_no_grid *grid;
  gamera_no_storage *st;
  const gamera_no_background_field *bg;
  gamera_cs_earth_startup startup;
  gamera_cs_mi_dipole_config cfg;
  double hall_siemens;   /* > 0: Pedersen + Hall tensor solve */
  size_t active[3], ghost, stride;
  vec3 *shell_b[ORC_FACES];
  double *shell_fac[ORC_FACES];
  double *fac[MI_HEMIS], *potential[MI_HEMIS];
  double *sigma_p, *sigma_h;
  gamera_mi_solver_stats solve[MI_HEMIS];
  double sample_radius, max_colat, sample_max_colat;
  double b_norm, potential_norm, current_norm;
  double next_update, epoch;
  int ready;
};

/* :45-57 */
static int valid_mi_config(const gamera_cs_mi_dipole_config *c) {
  return c != NULL && c->azimuth_points >= 8 && c->colatitude_points >= 3 &&
         c->current_shell < 64 && isfinite(c->ionosphere_radius_re) &&
         c->ionosphere_radius_re > 0.0 && isfinite(c->pedersen_siemens) &&
         c->pedersen_siemens > 0.0 && isfinite(c->coupling_interval_s) &&
         c->coupling_interval_s > 0.0 && isfinite(c->length_norm_m) &&
         c->length_norm_m > 0.0 && isfinite(c->speed_norm_m_s) &&
         c->speed_norm_m_s > 0.0 && isfinite(c->mass_density_norm_kg_m3) &&
         c->mass_density_norm_kg_m3 > 0.0 && c->maximum_iterations > 0 &&
         isfinite(c->relative_tolerance) && c->relative_tolerance > 0.0;
}

void orc_mi_destroy(orc_mi *m) {
  if (m == NULL) return;
  for (int f = 0; f < ORC_FACES; ++f) {
    free(m->shell_b[f]);
    free(m->shell_fac[f]);
  }
  for (int h = 0; h < MI_HEMIS; ++h) {
    free(m->fac[h]);
    free(m->potential[h]);
  }
  free(m->sigma_p);
  free(m->sigma_h);
  free(m);
}

/* :59-85 and :157-221 */
int orc_mi_create(const gamera_no_grid grid[ORC_FACES],
                  gamera_no_storage st[ORC_FACES],
                  const gamera_no_background_field bg[ORC_FACES],
                  const gamera_cs_earth_startup *startup, const size_t active[3],
                  size_t ghost, const gamera_cs_mi_dipole_config *cfg,
                  double hall_siemens, orc_mi **out) {
  if (out == NULL) return -1;
  *out = NULL;
  if (grid == NULL || st == NULL || bg == NULL || startup == NULL ||
      active == NULL || ghost < 4 || active[0] < 3 || active[1] < 8 ||
      active[1] != active[2] || !valid_mi_config(cfg) ||
      cfg->current_shell + 1 >= active[0] ||
      !(cfg->ionosphere_radius_re < startup->inner_radius) ||
      cfg->azimuth_points > SIZE_MAX / cfg->colatitude_points ||
      !isfinite(hall_siemens) || hall_siemens < 0.0)
    return -1;
  orc_mi *m = calloc(1, sizeof(*m));
  if (m == NULL) return -1;
  m->startup = *startup;
  m->cfg = *cfg;
  m->hall_siemens = hall_siemens;
  m->max_colat = asin(sqrt(cfg->ionosphere_radius_re / startup->inner_radius));
  const double mu0 = 4.0e-7 * acos(-1.0);
  m->b_norm = cfg->speed_norm_m_s * sqrt(mu0 * cfg->mass_density_norm_kg_m3);
  m->current_norm = m->b_norm / (mu0 * cfg->length_norm_m);
  m->potential_norm = cfg->speed_norm_m_s * m->b_norm * cfg->length_norm_m;
  const size_t polar = cfg->azimuth_points * cfg->colatitude_points;
  int ok = m->potential_norm > 0.0 && isfinite(m->max_colat);
  for (int h = 0; ok && h < MI_HEMIS; ++h) {
    m->potential[h] = calloc(polar, sizeof(double));
    m->fac[h] = calloc(polar, sizeof(double));
    ok = m->potential[h] != NULL && m->fac[h] != NULL;
  }
  if (ok && hall_siemens > 0.0) {
    m->sigma_p = malloc(polar * sizeof(double));
    m->sigma_h = malloc(polar * sizeof(double));
    ok = m->sigma_p != NULL && m->sigma_h != NULL;
    for (size_t q = 0; ok && q < polar; ++q) {
      m->sigma_p[q] = cfg->pedersen_siemens;
      m->sigma_h[q] = hall_siemens;
    }
  }
  m->grid = grid;
  m->st = st;
  m->bg = bg;
  memcpy(m->active, active, sizeof(m->active));
  m->ghost = ghost;
  m->stride = active[1] + 2;
  for (int f = 0; ok && f < ORC_FACES; ++f) {
    m->shell_b[f] = calloc(orc_count(grid[f].cell_extent), sizeof(vec3));
    m->shell_fac[f] = calloc(m->stride * m->stride, sizeof(double));
    ok = m->shell_b[f] != NULL && m->shell_fac[f] != NULL;
  }
  if (ok) {
    /* mean radius of the sample shell's active cells */
    double sum = 0.0;
    size_t count = 0;
    const size_t si = ghost + cfg->current_shell;
    for (int f = 0; f < ORC_FACES; ++f)
      for (size_t j = ghost; j < ghost + active[1]; ++j)
        for (size_t k = ghost; k < ghost + active[2]; ++k) {
          sum += norm3(grid[f].cell[orc_idx(grid[f].cell_extent, si, j, k)].centroid);
          ++count;
        }
    ok = count > 0;
    if (ok) {
      m->sample_radius = sum / (double)count;
      m->sample_max_colat =
          asin(sqrt(cfg->ionosphere_radius_re / m->sample_radius));
      ok = isfinite(m->sample_radius) && m->sample_radius > startup->inner_radius;
    }
  }
  if (!ok) {
    orc_mi_destroy(m);
    return -1;
  }
  *out = m;
  return 0;
}

/* Donor position of a point on the active block of its cube face, in cell
 * units (cubedsphere_mi_dipole.c:250-260). */
static int donor_of(const orc_mi *m, vec3 point, int *face, double *x,
                    double *y) {
  double a, b;
  if (orc_cube_inverse(point, face, &a, &b) != 0) return -1;
  const double qp = 0.25 * acos(-1.0);
  *x = (a + qp) * (double)m->active[1] / (2.0 * qp) - 0.5;
  *y = (b + qp) * (double)m->active[2] / (2.0 * qp) - 0.5;
  return 0;
}

/* :223-342 */
static int fill_shell_fac(orc_mi *m) {
  const size_t si = m->ghost + m->cfg.current_shell;
  const size_t stride = m->stride;
  for (int f = 0; f < ORC_FACES; ++f)
    memcpy(m->shell_b[f], m->st[f].cell_magnetic,
           orc_count(m->grid[f].cell_extent) * sizeof(vec3));
  /* one-cell fringe of B on shells si-1..si+1: bilinear from the owner face */
  for (int f = 0; f < ORC_FACES; ++f)
    for (size_t i = si - 1; i <= si + 1; ++i)
      for (size_t lj = 0; lj < stride; ++lj)
        for (size_t lk = 0; lk < stride; ++lk) {
          if (lj > 0 && lj <= m->active[1] && lk > 0 && lk <= m->active[2]) continue;
          const size_t j = m->ghost - 1 + lj, k = m->ghost - 1 + lk;
          const size_t t = orc_idx(m->grid[f].cell_extent, i, j, k);
          int df;
          double x, y;
          if (donor_of(m, m->grid[f].cell[t].centroid, &df, &x, &y) != 0 || df == f)
            return -1;
          const size_t j0 = (size_t)fmax(0.0, fmin((double)m->active[1] - 2.0, floor(x)));
          const size_t k0 = (size_t)fmax(0.0, fmin((double)m->active[2] - 2.0, floor(y)));
          const double wx = x - (double)j0, wy = y - (double)k0;
          vec3 v = {{0.0, 0.0, 0.0}};
          for (size_t dj = 0; dj < 2; ++dj)
            for (size_t dk = 0; dk < 2; ++dk) {
              const size_t s = orc_idx(m->grid[df].cell_extent, i, m->ghost + j0 + dj,
                                       m->ghost + k0 + dk);
              const double w = (dj == 0 ? 1.0 - wx : wx) * (dk == 0 ? 1.0 - wy : wy);
              for (int a = 0; a < 3; ++a)
                v.value[a] += w * m->st[df].cell_magnetic[s].value[a];
            }
          m->shell_b[f][t] = v;
        }
  /* FAC of the active shell cells */
  for (int f = 0; f < ORC_FACES; ++f)
    for (size_t lj = 1; lj <= m->active[1]; ++lj)
      for (size_t lk = 1; lk <= m->active[2]; ++lk) {
        const size_t j = m->ghost - 1 + lj, k = m->ghost - 1 + lk;
        const size_t cell = orc_idx(m->grid[f].cell_extent, si, j, k);
        vec3 cur;
        if (orc_mi_cell_current(&m->grid[f], m->shell_b[f], si, j, k, &cur) != 0)
          return -1;
        const vec3 b0 = m->bg[f].cell_magnetic[cell];
        const vec3 x = m->grid[f].cell[cell].centroid;
        const double r = norm3(x), bm = norm3(b0);
        if (!(r > 0.0) || !(bm > 0.0)) return -1;
        const double colat = acos(fmin(1.0, fabs(x.value[2]) / r));
        const double ratio =
            orc_mi_field_ratio(colat, r, m->cfg.ionosphere_radius_re);
        const double fac = dot3(cur, b0) / bm * ratio * m->current_norm;
        if (!isfinite(fac)) return -1;
        m->shell_fac[f][lj * stride + lk] = fac;
      }
  /* scalar FAC fringe from the physical owner face */
  for (int f = 0; f < ORC_FACES; ++f)
    for (size_t lj = 0; lj < stride; ++lj)
      for (size_t lk = 0; lk < stride; ++lk) {
        if (lj > 0 && lj <= m->active[1] && lk > 0 && lk <= m->active[2]) continue;
        const size_t j = m->ghost - 1 + lj, k = m->ghost - 1 + lk;
        const size_t cell = orc_idx(m->grid[f].cell_extent, si, j, k);
        int df;
        double x, y;
        if (donor_of(m, m->grid[f].cell[cell].centroid, &df, &x, &y) != 0 || df == f)
          return -1;
        const size_t j0 = (size_t)fmax(0.0, fmin((double)m->active[1] - 2.0, floor(x)));
        const size_t k0 = (size_t)fmax(0.0, fmin((double)m->active[2] - 2.0, floor(y)));
        const double wx = x - (double)j0, wy = y - (double)k0;
        const double *src = m->shell_fac[df];
        const size_t lo = (j0 + 1) * stride + k0 + 1;
        const double low = (1.0 - wy) * src[lo] + wy * src[lo + 1];
        const double high = (1.0 - wy) * src[lo + stride] + wy * src[lo + stride + 1];
        const double fac = (1.0 - wx) * low + wx * high;
        if (!isfinite(fac)) return -1;
        m->shell_fac[f][lj * stride + lk] = fac;
      }
  return 0;
}

/* :344-370 */
static int sample_fac(const orc_mi *m, vec3 point, double *out) {
  int f;
  double a, b;
  if (out == NULL || orc_cube_inverse(point, &f, &a, &b) != 0) return -1;
  /* the scale is formed first here, unlike the fringe fill (:352-354) */
  const double qp = 0.25 * acos(-1.0);
  const double scale = (double)m->active[1] / (2.0 * qp);
  const double x = (a + qp) * scale - 0.5, y = (b + qp) * scale - 0.5;
  const long j = (long)floor(x), k = (long)floor(y);
  if (j < -1 || k < -1 || j >= (long)m->active[1] || k >= (long)m->active[2])
    return -1;
  const double wx = x - (double)j, wy = y - (double)k;
  const size_t row = (size_t)(j + 1), col = (size_t)(k + 1), s = m->stride;
  const double *src = m->shell_fac[f];
  const double lower = (1.0 - wy) * src[row * s + col] + wy * src[row * s + col + 1];
  const double upper =
      (1.0 - wy) * src[(row + 1) * s + col] + wy * src[(row + 1) * s + col + 1];
  *out = (1.0 - wx) * lower + wx * upper;
  return isfinite(*out) ? 0 : -1;
}

/* :372-378 */
int orc_mi_update_due(const orc_mi *m, double time_code) {
  if (m == NULL || !isfinite(time_code) || time_code < 0.0) return -1;
  const double ts = time_code * m->cfg.length_norm_m / m->cfg.speed_norm_m_s;
  return !m->ready || ts + 1e-10 >= m->next_update;
}

/* :380-436 (with the tensor solve when a Hall conductance is set) */
int orc_mi_maybe_update(orc_mi *m, double time_code) {
  const int due = orc_mi_update_due(m, time_code);
  if (due <= 0) return due;
  const double ts = time_code * m->cfg.length_norm_m / m->cfg.speed_norm_m_s;
  if (fill_shell_fac(m) != 0) return -1;
  const size_t np = m->cfg.azimuth_points, nt = m->cfg.colatitude_points;
  const double dt = m->max_colat / (double)(nt - 1);
  const double dp = 2.0 * acos(-1.0) / (double)np;
  for (int h = 0; h < MI_HEMIS; ++h) {
    const double sign = h == MI_NORTH ? 1.0 : -1.0;
    for (size_t t = 0; t < nt; ++t) {
      const double ti = (double)t * dt;
      if (ti > m->sample_max_colat + 4e-14) {
        memset(m->fac[h] + t * np, 0, np * sizeof(double));
        continue;
      }
      const double arg =
          fmin(1.0, sin(ti) * sqrt(m->sample_radius / m->cfg.ionosphere_radius_re));
      const double tm = asin(arg);
      for (size_t p = 0; p < np; ++p) {
        const double phi = (double)p * dp;
        const vec3 x = {{m->sample_radius * sin(tm) * cos(phi),
                         m->sample_radius * sin(tm) * sin(phi),
                         sign * m->sample_radius * cos(tm)}};
        if (sample_fac(m, x, &m->fac[h][t * np + p]) != 0) return -1;
      }
    }
    const solver_cfg sc = {
        .longitude_count = np, .colatitude_count = nt,
        .maximum_colatitude = m->max_colat,
        .ionosphere_radius_m = m->cfg.ionosphere_radius_re * m->cfg.length_norm_m,
        .pedersen_siemens = m->cfg.pedersen_siemens, .hall_siemens = 0.0,
        .low_latitude_potential_v = 0.0,
        .hemisphere = h == MI_NORTH ? GAMERA_MI_NORTH : GAMERA_MI_SOUTH,
        .maximum_iterations = m->cfg.maximum_iterations,
        .relative_tolerance = m->cfg.relative_tolerance,
        .absolute_tolerance = 0.0};
    const int status =
        m->hall_siemens > 0.0
            ? orc_mi_solve_tensor(&sc, m->fac[h], m->sigma_p, m->sigma_h,
                                  m->potential[h], &m->solve[h])
            : orc_mi_solve_pedersen(&sc, m->fac[h], m->potential[h], &m->solve[h]);
    if (status != 0 || !m->solve[h].converged) return -1;
  }
  m->ready = 1;
  m->epoch = ts;
  do {
    m->next_update += m->cfg.coupling_interval_s;
  } while (m->next_update <= ts + 1e-10);
  return 1;
}

/* :438-440 */
int orc_mi_ready(const orc_mi *m) { return m != NULL && m->ready; }

/* :442-476: bilinear in (mapped colatitude, longitude), zero beyond the
 * polar cap. */
int orc_mi_potential(const orc_mi *m, vec3 x, double *out) {
  if (!orc_mi_ready(m) || out == NULL) return -1;
  const double r = norm3(x);
  if (!(r >= m->cfg.ionosphere_radius_re)) return -1;
  const int h = x.value[2] >= 0.0 ? MI_NORTH : MI_SOUTH;
  const double tm = acos(fmin(1.0, fabs(x.value[2]) / r));
  const double ti = orc_mi_mapped_colatitude(tm, r, m->cfg.ionosphere_radius_re);
  if (!isfinite(ti)) return -1;
  if (ti >= m->max_colat) {
    *out = 0.0;
    return 0;
  }
  const size_t np = m->cfg.azimuth_points, nt = m->cfg.colatitude_points;
  const double tq = ti / m->max_colat * (double)(nt - 1);
  const size_t t = (size_t)floor(tq);
  const double wt = tq - (double)t;
  double phi = atan2(x.value[1], x.value[0]);
  if (phi < 0.0) phi += 2.0 * acos(-1.0);
  const double pq = phi * (double)np / (2.0 * acos(-1.0));
  const size_t p = (size_t)floor(pq) % np, q = (p + 1) % np;
  const double wp = pq - floor(pq);
  const double *v = m->potential[h];
  const double lower = (1.0 - wp) * v[t * np + p] + wp * v[t * np + q];
  const double upper = (1.0 - wp) * v[(t + 1) * np + p] + wp * v[(t + 1) * np + q];
  *out = (1.0 - wt) * lower + wt * upper;
  return isfinite(*out) ? 0 : -1;
}

/* :478-503: inner-wall tangential EMF = -(potential difference) along each
 * valid j and k edge of the wall shell. */
int orc_mi_inner_emf(const orc_mi *m, gamera_no_storage *s,
                     const gamera_no_grid *g, size_t wall_i) {
  if (!orc_mi_ready(m) || s == NULL || g == NULL) return -1;
  for (int d = 1; d <= 2; ++d)
    for (size_t j = 0; j < g->edge[d].extent[1]; ++j)
      for (size_t k = 0; k < g->edge[d].extent[2]; ++k) {
        const size_t e = orc_idx(g->edge[d].extent, wall_i, j, k);
        if (!g->edge[d].valid[e]) continue;
        size_t end[3] = {wall_i, j, k};
        ++end[d];
        const vec3 a = g->vertex[orc_idx(g->vertex_extent, wall_i, j, k)];
        const vec3 b = g->vertex[orc_idx(g->vertex_extent, end[0], end[1], end[2])];
        double pa, pb;
        if (orc_mi_potential(m, a, &pa) != 0 || orc_mi_potential(m, b, &pb) != 0)
          return -1;
        s->edge_emf[d][e] = -(pb - pa) / m->potential_norm;
      }
  return 0;
}

/* cubedsphere_earth_startup.c:7-19, :92-102 with nonorthogonal_background.c
 * :422-442: dipole field with the radial cut. */
static int startup_background(const gamera_cs_earth_startup *p, vec3 x, vec3 *out) {
  const vec3 mo = p->dipole.moment;
  const double r2 = dot3(x, x);
  if (!(r2 > DBL_MIN) || !isfinite(r2)) return -1;
  const double r = sqrt(r2);
  const double irc = 1.0 / (r2 * r);
  const double irf = irc / r2;
  const double pr = dot3(mo, x);
  for (int c = 0; c < 3; ++c)
    out->value[c] = (3.0 * pr * irf) * x.value[c] - irc * mo.value[c];
  for (int c = 0; c < 3; ++c)
    if (!isfinite(out->value[c])) return -1;
  const double rad = sqrt(x.value[0] * x.value[0] + x.value[1] * x.value[1] +
                          x.value[2] * x.value[2]);
  const double sc = (rad - p->cut_radius) / p->cut_length;
  const double w = sc <= 0.0 ? 1.0 : sc >= 1.0 ? 0.0 : 1.0 - 3.0 * sc * sc + 2.0 * sc * sc * sc;
  for (int c = 0; c < 3; ++c) out->value[c] *= w;
  return 0;
}

/* :505-543: E from the potential by central differences, E x B / B^2 drift
 * without its radial part, ghost velocity 2 v_EB - v. */
int orc_mi_ghost_velocity(const orc_mi *m, vec3 x, vec3 source, vec3 *ghost) {
  if (!orc_mi_ready(m) || ghost == NULL) return -1;
  const double r = norm3(x);
  if (!(r > 0.0)) return -1;
  const double eps = 1e-4 * fmax(1.0, r);
  vec3 e = {{0.0, 0.0, 0.0}};
  for (int c = 0; c < 3; ++c) {
    vec3 plus = x, minus = x;
    plus.value[c] += eps;
    minus.value[c] -= eps;
    double pp, pm;
    if (orc_mi_potential(m, plus, &pp) != 0 || orc_mi_potential(m, minus, &pm) != 0)
      return -1;
    e.value[c] = -(pp - pm) / (2.0 * eps * m->cfg.length_norm_m *
                               m->cfg.speed_norm_m_s * m->b_norm);
  }
  vec3 b;
  if (startup_background(&m->startup, x, &b) != 0) return -1;
  const double b2 = dot3(b, b);
  if (!(b2 > 0.0)) return -1;
  vec3 drift = {{(e.value[1] * b.value[2] - e.value[2] * b.value[1]) / b2,
                 (e.value[2] * b.value[0] - e.value[0] * b.value[2]) / b2,
                 (e.value[0] * b.value[1] - e.value[1] * b.value[0]) / b2}};
  const double radial = dot3(drift, x) / (r * r);
  for (int c = 0; c < 3; ++c) drift.value[c] -= radial * x.value[c];
  return orc_mi_compose_ghost_velocity(source, b, drift, ghost);
}

/* Inspection for tests: potential of one hemisphere and its last solve. */
int orc_mi_potential_values(const orc_mi *m, int hemi, const double **values,
                            size_t *count) {
  if (m == NULL || hemi < 0 || hemi >= MI_HEMIS || values == NULL || count == NULL)
    return -1;
  *values = m->potential[hemi];
  *count = m->cfg.azimuth_points * m->cfg.colatitude_points;
  return 0;
}

int orc_mi_fac_values(const orc_mi *m, int hemi, const double **values,
                      size_t *count) {
  if (m == NULL || hemi < 0 || hemi >= MI_HEMIS || values == NULL || count == NULL)
    return -1;
  *values = m->fac[hemi];
  *count = m->cfg.azimuth_points * m->cfg.colatitude_points;
  return 0;
}

int orc_mi_solve_stats(const orc_mi *m, int hemi, gamera_mi_solver_stats *out) {
  if (m == NULL || hemi < 0 || hemi >= MI_HEMIS || out == NULL) return -1;
  *out = m->solve[hemi];
  return 0;
}
