/* Pointwise numerics of the step (test infrastructure; see orc.h). */
#include "orc.h"

#include <float.h>
#include <math.h>

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* ((0 + a0 b0) + a1 b1) + a2 b2, the reference's vector_dot order. */
static double dot3(const double a[3], const double b[3]) {
  double r = 0.0;
  for (int d = 0; d < 3; ++d) r += a[d] * b[d];
  return r;
}

static int finite3(const double v[3]) {
  return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

static int valid_floors(double gamma, double rho_floor, double p_floor) {
  return isfinite(gamma) && gamma > 1.0 && isfinite(rho_floor) &&
         rho_floor > 0.0 && isfinite(p_floor) && p_floor > 0.0;
}

/* NO/nonorthogonal_state.c:60-93 */
int orc_cons_to_prim(const double u[ORC_NQ], double gamma, double rho_floor,
                     double p_floor, gamera_no_primitive *w) {
  if (!valid_floors(gamma, rho_floor, p_floor) || !isfinite(u[0]) ||
      u[0] <= DBL_MIN)
    return -1;
  const double rho = u[0];
  double v2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(u[1 + d])) return -1;
    w->velocity.value[d] = u[1 + d] / rho;
    v2 += w->velocity.value[d] * w->velocity.value[d];
  }
  if (!isfinite(u[4])) return -1;
  const double kinetic = 0.5 * rho * v2;
  w->density = fmax(rho, rho_floor);
  w->pressure = fmax((gamma - 1.0) * (u[4] - kinetic), p_floor);
  return 0;
}

/* NO/nonorthogonal_state.c:95-120 */
int orc_prim_to_cons(const gamera_no_primitive *w, double gamma,
                     double rho_floor, double p_floor, double u[ORC_NQ]) {
  if (!valid_floors(gamma, rho_floor, p_floor) || !isfinite(w->density) ||
      !isfinite(w->pressure))
    return -1;
  const double rho = fmax(w->density, rho_floor);
  const double p = fmax(w->pressure, p_floor);
  double v2 = 0.0;
  u[0] = rho;
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(w->velocity.value[d])) return -1;
    u[1 + d] = rho * w->velocity.value[d];
    v2 += w->velocity.value[d] * w->velocity.value[d];
  }
  u[4] = p / (gamma - 1.0) + 0.5 * rho * v2;
  return 0;
}

/* NO/nonorthogonal_state.c:122-159: primitive Adams-Bashforth predictor. */
int orc_predict_cell(const double old_u[ORC_NQ], const double u[ORC_NQ],
                     double r, double gamma, double rho_floor, double p_floor,
                     double out[ORC_NQ]) {
  if (!isfinite(r)) return -1;
  gamera_no_primitive a, b;
  if (orc_cons_to_prim(old_u, gamma, rho_floor, p_floor, &a) != 0 ||
      orc_cons_to_prim(u, gamma, rho_floor, p_floor, &b) != 0)
    return -1;
  gamera_no_primitive w = b;
  w.density = fmax(b.density + r * (b.density - a.density), rho_floor);
  w.pressure = fmax(b.pressure + r * (b.pressure - a.pressure), p_floor);
  for (int d = 0; d < 3; ++d)
    w.velocity.value[d] =
        b.velocity.value[d] + r * (b.velocity.value[d] - a.velocity.value[d]);
  return orc_prim_to_cons(&w, gamma, rho_floor, p_floor, out);
}

/* NO/nonorthogonal_state.c:169-203 (non-Boris magnetic momentum update). */
int orc_apply_maxwell(double u[ORC_NQ], vec3 mr, double dt, double gamma,
                      double rho_floor, double p_floor) {
  if (!isfinite(dt) || !valid_floors(gamma, rho_floor, p_floor)) return -1;
  const double rho = fmax(u[0], rho_floor);
  double m2_old = 0.0;
  for (int d = 0; d < 3; ++d) m2_old += u[1 + d] * u[1 + d];
  const double p = fmax((gamma - 1.0) * (u[4] - 0.5 * m2_old / rho), p_floor);
  double m2_new = 0.0;
  for (int d = 0; d < 3; ++d) {
    u[1 + d] += dt * mr.value[d];
    m2_new += u[1 + d] * u[1 + d];
  }
  u[0] = rho;
  u[4] = 0.5 * m2_new / rho + p / (gamma - 1.0);
  return 0;
}

static void scale3(double s, const double v[3], double out[3]) {
  for (int d = 0; d < 3; ++d) out[d] = s * v[d];
}

/* v - (v.e) e, evaluated as the reference's vector_perpendicular. */
static void perp3(const double v[3], const double e[3], double out[3]) {
  double par[3];
  scale3(dot3(v, e), e, par);
  for (int d = 0; d < 3; ++d) out[d] = v[d] - par[d];
}

/* NO/nonorthogonal_state.c:205-280: semi-relativistic Boris correction. */
int orc_apply_boris(double u[ORC_NQ], const double old_u[ORC_NQ], vec3 bn,
                    vec3 bo, const double rate[ORC_NQ], vec3 mr, double c,
                    double dt, double gamma, double rho_floor, double p_floor) {
  if (!isfinite(c) || c <= DBL_MIN || !isfinite(dt)) return -1;
  gamera_no_primitive w;
  if (orc_cons_to_prim(u, gamma, rho_floor, p_floor, &w) != 0 ||
      old_u[0] <= DBL_MIN)
    return -1;
  const double rho = fmax(w.density, rho_floor);
  const double p = fmax(w.pressure, p_floor);
  const double rho_old = old_u[0];
  double v_old[3];
  for (int d = 0; d < 3; ++d) v_old[d] = old_u[1 + d] / fmax(rho_old, rho_floor);
  double half_b[3];
  for (int d = 0; d < 3; ++d) half_b[d] = 0.5 * (bn.value[d] + bo.value[d]);
  const double alpha = dot3(half_b, half_b) / (rho * c * c);
  const double drho = dt * rate[0];
  double dm_gas[3], dm_mag[3], dm[3];
  for (int d = 0; d < 3; ++d) {
    dm_gas[d] = dt * rate[1 + d];
    dm_mag[d] = dt * mr.value[d];
  }
  if (alpha > 1.0e-12) {
    const double plasma = 1.0 / (1.0 + alpha);
    double unit[3] = {0.0, 0.0, 0.0};
    const double norm = sqrt(dot3(half_b, half_b));
    if (!(norm <= 1.0e-12)) scale3(1.0 / norm, half_b, unit);
    double total[3], par[3], perp_total[3], perp_v[3], t1[3], t2[3], sum[3];
    for (int d = 0; d < 3; ++d) total[d] = dm_gas[d] + dm_mag[d];
    scale3(dot3(dm_gas, unit), unit, par);
    perp3(total, unit, perp_total);
    perp3(v_old, unit, perp_v);
    scale3(plasma, perp_total, t1);
    scale3(alpha * plasma * drho, perp_v, t2);
    for (int d = 0; d < 3; ++d) sum[d] = t1[d] + t2[d];
    for (int d = 0; d < 3; ++d) dm[d] = par[d] + sum[d];
  } else {
    for (int d = 0; d < 3; ++d) dm[d] = dm_gas[d] + dm_mag[d];
  }
  gamera_no_primitive r = w;
  r.density = rho;
  r.pressure = p;
  for (int d = 0; d < 3; ++d)
    r.velocity.value[d] = (rho_old * v_old[d] + dm[d]) / rho;
  return orc_prim_to_cons(&r, gamma, rho_floor, p_floor, u);
}

/* NO/nonorthogonal_reconstruction.c:7-17: stencil weights. */
static const double C8[8] = {-3.0 / 840.0,  29.0 / 840.0, -139.0 / 840.0,
                             533.0 / 840.0, 533.0 / 840.0, -139.0 / 840.0,
                             29.0 / 840.0,  -3.0 / 840.0};
static const double C6[8] = {0.0,          1.0 / 60.0, -8.0 / 60.0,
                             37.0 / 60.0,  37.0 / 60.0, -8.0 / 60.0,
                             1.0 / 60.0,   0.0};
static const double U7[7] = {-3.0 / 420.0,  25.0 / 420.0, -101.0 / 420.0,
                             319.0 / 420.0, 214.0 / 420.0, -38.0 / 420.0,
                             4.0 / 420.0};

/* NO/nonorthogonal_reconstruction.c:19-26, 53-59: left-to-right sums. */
double orc_central8(const double s[8]) {
  double r = 0.0;
  for (int n = 0; n < 8; ++n) r += C8[n] * s[n];
  return r;
}
double orc_central6(const double s[8]) {
  double r = 0.0;
  for (int n = 0; n < 8; ++n) r += C6[n] * s[n];
  return r;
}

/* NO/nonorthogonal_reconstruction.c:39-51: PDM limiter, pdmb code A. */
static double pdm(double q0, double q1, double q2, double hi, double a) {
  const double bounded = fmax(fmin(q1, q2), fmin(hi, fmax(q1, q2)));
  const double d0 = a * (q1 - q0);
  const double d1 = a * (q2 - q1);
  const double s0 = copysign(1.0, d0);
  const double s1 = copysign(1.0, d1);
  const double same = fabs(s0 + s1);
  const double slope = bounded - q1;
  return bounded - s1 * fmax(0.0, fabs(slope) - same * fabs(d0));
}

/* NO/nonorthogonal_reconstruction.c:61-93: volume-weighted Up7 + PDM. */
int orc_up7_pdm(const double vol[8], const double q[8], double a,
                double *left, double *right) {
  if (!isfinite(a) || a < 0.0) return -1;
  double wq[8];
  for (int n = 0; n < 8; ++n) wq[n] = vol[n] * q[n];
  const double iv = orc_central8(vol);
  if (!isfinite(iv) || fabs(iv) <= DBL_MIN) return -1;
  double l = 0.0, r = 0.0;
  for (int n = 0; n < 7; ++n) l += U7[n] * wq[n];
  for (int n = 0; n < 7; ++n) r += U7[n] * wq[7 - n];
  *left = pdm(q[2], q[3], q[4], l / iv, a);
  *right = pdm(q[5], q[4], q[3], r / iv, a);
  return 0;
}

static int valid_prim(const gamera_no_primitive *w) {
  return isfinite(w->density) && isfinite(w->pressure) &&
         w->density > DBL_MIN && w->pressure > DBL_MIN &&
         finite3(w->velocity.value);
}

/* NO/nonorthogonal_flux.c:38-105: gas-kinetic half-space fluid flux. */
int orc_fluid_flux_kinetic(const gamera_no_primitive w[2], double gamma,
                           const gamera_no_face_geometry *f,
                           orc_fluid_flux *out) {
  if (!isfinite(gamma) || gamma <= 1.0 || !valid_prim(&w[0]) ||
      !valid_prim(&w[1]))
    return -1;
  for (int q = 0; q < ORC_NQ; ++q) out->flux[q] = out->jump[q] = 0.0;
  for (int d = 0; d < 3; ++d) out->velocity_jump[d] = 0.0;
  for (int side = 0; side < 2; ++side) {
    const double sign = side == 0 ? -1.0 : 1.0;
    const double rho = w[side].density, p = w[side].pressure;
    const double *v = w[side].velocity.value;
    const double v2 = dot3(v, v);
    const double e = 0.5 * rho * v2 + p / (gamma - 1.0);
    const double lambda = rho / (2.0 * p);
    const double un = dot3(f->normal.value, v);
    const double ut1 = dot3(f->tangent1.value, v);
    const double ut2 = dot3(f->tangent2.value, v);
    const double m0 = 0.5 * erfc(sign * sqrt(lambda) * un);
    const double m1 = un * m0 - 0.5 * sign * exp(-lambda * un * un) /
                                    sqrt(ORC_PI * lambda);
    out->normal_velocity[side] = un;
    out->flux[0] += rho * m1;
    out->flux[4] += (e + 0.5 * p) * m1 + 0.5 * p * m0 * un;
    const double mn = rho * un * m1 + p * m0;
    const double mt1 = rho * ut1 * m1;
    const double mt2 = rho * ut2 * m1;
    for (int d = 0; d < 3; ++d) {
      out->flux[1 + d] += f->normal.value[d] * mn +
                          f->tangent1.value[d] * mt1 +
                          f->tangent2.value[d] * mt2;
      out->velocity_jump[d] += sign * v[d];
    }
    out->jump[0] += sign * rho;
    for (int d = 0; d < 3; ++d) out->jump[1 + d] += sign * rho * v[d];
    out->jump[4] += sign * e;
  }
  return 0;
}

/* NO/nonorthogonal_flux.c:107-173: kinetic Maxwell-stress flux. */
int orc_maxwell_flux_kinetic(const gamera_no_primitive w[2], const vec3 b[2],
                             double bn, const gamera_no_face_geometry *f,
                             bool use_bg, vec3 b0, double b0n, bool boris,
                             double c, const double un[2],
                             orc_maxwell_flux *out) {
  if (!isfinite(bn) || (boris && (!isfinite(c) || c <= DBL_MIN))) return -1;
  for (int side = 0; side < 2; ++side) {
    if (!valid_prim(&w[side])) return -1;
    for (int d = 0; d < 3; ++d)
      if (!isfinite(b[side].value[d]) || (use_bg && !isfinite(b0.value[d])))
        return -1;
  }
  for (int d = 0; d < 3; ++d) out->momentum[d] = 0.0;
  out->alfven_speed = 0.0;
  out->pressure_sum = 0.0;
  for (int side = 0; side < 2; ++side) {
    const double *bs = b[side].value;
    const double r2 = dot3(bs, bs);
    const double pr = 0.5 * r2;
    double total[3];
    for (int d = 0; d < 3; ++d) total[d] = use_bg ? bs[d] + b0.value[d] : bs[d];
    const double t2 = dot3(total, total);
    const double va2 = t2 / w[side].density;
    const double va2c = boris ? va2 / (1.0 + va2 / (c * c)) : va2;
    const double lambda =
        1.0 / (2.0 * w[side].pressure / w[side].density + va2c);
    const double m0 =
        0.5 * erfc((side == 0 ? -1.0 : 1.0) * sqrt(lambda) * un[side]);
    out->alfven_speed += 0.5 * sqrt(va2c);
    out->pressure_sum += 0.5 * t2;
    const double b0_dot_b = use_bg ? dot3(b0.value, bs) : 0.0;
    for (int d = 0; d < 3; ++d) {
      double stress = pr * f->normal.value[d] - bs[d] * bn;
      if (use_bg)
        stress += -b0.value[d] * bn - bs[d] * b0n + b0_dot_b * f->normal.value[d];
      out->momentum[d] += m0 * stress;
    }
  }
  return 0;
}

/* NO/nonorthogonal_flux.c:175-204: high-order gradient stabilisation. */
int orc_hogs(orc_fluid_flux *f, orc_maxwell_flux *m, double ch, double cm,
             bool boris, double c) {
  if (!isfinite(ch) || !isfinite(cm) || ch < 0.0 || cm < 0.0 ||
      (boris && (!isfinite(c) || c <= DBL_MIN)))
    return -1;
  for (int q = 0; q < ORC_NQ; ++q) f->flux[q] -= ch * m->alfven_speed * f->jump[q];
  if (boris) {
    const double ic2 = 1.0 / (c * c);
    for (int d = 0; d < 3; ++d) {
      const double jump = m->pressure_sum * f->velocity_jump[d] * ic2;
      m->momentum[d] -= cm * m->alfven_speed * jump;
    }
  }
  return 0;
}

/* NO/nonorthogonal_flux.c:206-241: emergency interface diffusion. */
int orc_emergency_diffusion(orc_fluid_flux *f, const gamera_no_primitive w[2],
                            const double lo[ORC_NQ], const double up[ORC_NQ],
                            bool boris, double c, bool *applied) {
  if (boris && (!isfinite(c) || c <= DBL_MIN)) return -1;
  *applied = false;
  if (!boris) return 0;
  double s2 = 0.0;
  for (int side = 0; side < 2; ++side) {
    if (!valid_prim(&w[side])) return -1;
    s2 += dot3(w[side].velocity.value, w[side].velocity.value);
  }
  if (sqrt(0.5 * s2) < 1.5 * c) return 0;
  for (int q = 0; q < ORC_NQ; ++q) {
    if (!isfinite(lo[q]) || !isfinite(up[q])) return -1;
    f->flux[q] -= c * (up[q] - lo[q]);
  }
  *applied = true;
  return 0;
}

/* NO/nonorthogonal_operators.c:56-79: 2x2 edge-normal solve. */
int orc_edge_field_solve(const double n1[2], const double n2[2], double f1,
                         double f2, double out[2]) {
  const double det = n1[0] * n2[1] - n2[0] * n1[1];
  const double scale = fmax(fmax(fabs(n1[0]), fabs(n1[1])),
                            fmax(fabs(n2[0]), fabs(n2[1])));
  if (!isfinite(det) || scale <= DBL_MIN ||
      fabs(det) <= 64.0 * DBL_EPSILON * scale * scale)
    return -1;
  const double inv = 1.0 / det;
  out[0] = inv * (n2[1] * f1 - n1[1] * f2);
  out[1] = inv * (-n2[0] * f1 + n1[0] * f2);
  return 0;
}

/* NO/nonorthogonal_operators.c:81-113: ebGeom face normal at the edge. */
int orc_face_normal_at_edge(const double area[8], const vec3 normal[8],
                            const gamera_no_edge_geometry *e, double out[2]) {
  const double ia = orc_central6(area);
  if (!isfinite(ia) || fabs(ia) <= DBL_MIN) return -1;
  double n[3];
  for (int d = 0; d < 3; ++d) {
    double wn[8];
    for (int s = 0; s < 8; ++s) wn[s] = area[s] * normal[s].value[d];
    n[d] = orc_central6(wn) / ia;
  }
  const double p1 = dot3(n, e->tangent1.value);
  const double p2 = dot3(n, e->tangent2.value);
  const double norm = hypot(p1, p2);
  if (!isfinite(norm) || norm <= DBL_MIN) return -1;
  out[0] = p1 / norm;
  out[1] = p2 / norm;
  return 0;
}

/* NO/nonorthogonal_operators.c:144-182: local edge EMF, line integrated. */
int orc_edge_emf_formula(double v1, double v2, double b1, double b2,
                         double jump, double rho,
                         const gamera_no_edge_geometry *e, double diff,
                         bool boris, double c, double cfl, double dt,
                         double *emf, double *speed) {
  if (!isfinite(v1) || !isfinite(v2) || !isfinite(b1) || !isfinite(b2) ||
      !isfinite(jump) || !isfinite(rho) || rho <= DBL_MIN ||
      !isfinite(e->length) || e->length <= DBL_MIN || !isfinite(diff) ||
      !isfinite(cfl) || cfl < 0.0 || !isfinite(dt) || dt <= DBL_MIN ||
      (boris && (!isfinite(c) || c <= DBL_MIN)))
    return -1;
  const double flow = hypot(v1, v2);
  double va = hypot(b1, b2) / sqrt(rho);
  if (boris) va = c * va / hypot(c, va);
  *speed = flow + va;
  *speed = fmin(*speed, cfl * e->length / dt);
  const double ideal = -(v1 * b2 - v2 * b1);
  *emf = (ideal + diff * (*speed) * jump) * e->length;
  return 0;
}

/* NO/nonorthogonal_operators.c:17-54 (renamed uncalibrated recovery,
 * cubedsphere/CMakeLists.txt:47-49): face-moment cell field. */
static int moment_field(const gamera_no_cell_geometry *g,
                        const double flux[3][2], double out[3]) {
  if (!isfinite(g->volume) || g->volume <= DBL_MIN) return -1;
  double div = 0.0;
  for (int c = 0; c < 3; ++c) out[c] = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double lo = flux[d][0], up = flux[d][1];
    div += up - lo;
    for (int c = 0; c < 3; ++c)
      out[c] += up * g->face[d][1].centroid.value[c] -
                lo * g->face[d][0].centroid.value[c];
  }
  for (int c = 0; c < 3; ++c)
    out[c] = (out[c] - div * g->centroid.value[c]) / g->volume;
  return 0;
}

/* cubedsphere/cubedsphere_recovery.c:17-67: Gauss-Jordan with partial
 * pivoting on the augmented 3x4 system. */
static int gauss_jordan3(const double m[3][3], const double rhs[3],
                         double x[3]) {
  double a[3][4];
  double scale = 0.0;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      a[i][j] = m[i][j];
      scale = fmax(scale, fabs(m[i][j]));
    }
    a[i][3] = rhs[i];
  }
  if (!isfinite(scale) || scale <= DBL_MIN) return -1;
  for (int col = 0; col < 3; ++col) {
    int piv = col;
    for (int row = col + 1; row < 3; ++row)
      if (fabs(a[row][col]) > fabs(a[piv][col])) piv = row;
    if (!isfinite(a[piv][col]) ||
        fabs(a[piv][col]) <= 128.0 * DBL_EPSILON * scale)
      return -1;
    if (piv != col)
      for (int j = col; j < 4; ++j) {
        const double t = a[col][j];
        a[col][j] = a[piv][j];
        a[piv][j] = t;
      }
    const double inv = 1.0 / a[col][col];
    for (int j = col; j < 4; ++j) a[col][j] *= inv;
    for (int row = 0; row < 3; ++row) {
      if (row == col) continue;
      const double f = a[row][col];
      for (int j = col; j < 4; ++j) a[row][j] -= f * a[col][j];
    }
  }
  for (int i = 0; i < 3; ++i) {
    if (!isfinite(a[i][3])) return -1;
    x[i] = a[i][3];
  }
  return 0;
}

/* cubedsphere/cubedsphere_recovery.c:69-98: moment recovery calibrated by
 * its response to the three unit uniform fields. */
int orc_recover_cell(const gamera_no_cell_geometry *g,
                     const double flux[3][2], vec3 *b) {
  double raw[3];
  if (moment_field(g, flux, raw) != 0) return -1;
  double response[3][3];
  for (int basis = 0; basis < 3; ++basis) {
    double bf[3][2], col[3];
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s) bf[d][s] = g->face[d][s].area_vector.value[basis];
    if (moment_field(g, bf, col) != 0) return -1;
    for (int d = 0; d < 3; ++d) response[d][basis] = col[d];
  }
  return gauss_jordan3(response, raw, b->value);
}
