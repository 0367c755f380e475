// Pointwise FP64 numerics of the six-face step, as device functions.
//
// Each function states the reference routine whose result it reproduces.
// The translation units that include this header are compiled with
// -ffp-contract=off: every product and sum below rounds exactly where the
// reference's (FMA-free) C does, so the only departures from the reference
// are the last-ulp differences of ROCm's erfc/exp/hypot/tan/atan2/acos from
// glibc's.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace gcs {

#define GCS_DEV __device__ __forceinline__

constexpr double kPi = 3.14159265358979323846264338327950288;

struct Prim {
  double rho, v[3], p;
};

GCS_DEV double dot3(const double a[3], const double b[3]) {
  double r = 0.0;
  r += a[0] * b[0];
  r += a[1] * b[1];
  r += a[2] * b[2];
  return r;
}

// nonorthogonal_state.c:60-93 (validity of gamma and floors is checked once
// on the host when the step options are accepted).
GCS_DEV bool cons_to_prim(const double u[5], double gamma, double rf,
                          double pf, Prim &w) {
  const double rho = u[0];
  if (!isfinite(rho) || rho <= DBL_MIN) return false;
  double v2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(u[1 + d])) return false;
    w.v[d] = u[1 + d] / rho;
    v2 += w.v[d] * w.v[d];
  }
  if (!isfinite(u[4])) return false;
  const double kinetic = 0.5 * rho * v2;
  w.rho = fmax(rho, rf);
  w.p = fmax((gamma - 1.0) * (u[4] - kinetic), pf);
  return true;
}

// nonorthogonal_state.c:95-120
GCS_DEV bool prim_to_cons(const Prim &w, double gamma, double rf, double pf,
                          double u[5]) {
  if (!isfinite(w.rho) || !isfinite(w.p)) return false;
  const double rho = fmax(w.rho, rf);
  const double p = fmax(w.p, pf);
  double v2 = 0.0;
  u[0] = rho;
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(w.v[d])) return false;
    u[1 + d] = rho * w.v[d];
    v2 += w.v[d] * w.v[d];
  }
  u[4] = p / (gamma - 1.0) + 0.5 * rho * v2;
  return true;
}

// nonorthogonal_state.c:122-159
GCS_DEV bool predict_cell(const double uo[5], const double uc[5], double r,
                          double gamma, double rf, double pf, double out[5]) {
  Prim a, b;
  if (!cons_to_prim(uo, gamma, rf, pf, a) || !cons_to_prim(uc, gamma, rf, pf, b))
    return false;
  Prim w;
  w.rho = fmax(b.rho + r * (b.rho - a.rho), rf);
  w.p = fmax(b.p + r * (b.p - a.p), pf);
  for (int d = 0; d < 3; ++d) w.v[d] = b.v[d] + r * (b.v[d] - a.v[d]);
  return prim_to_cons(w, gamma, rf, pf, out);
}

// nonorthogonal_reconstruction.c:7-26: left-to-right stencil sums.
GCS_DEV double central8(const double s[8]) {
  double r = 0.0;
  r += (-3.0 / 840.0) * s[0];
  r += (29.0 / 840.0) * s[1];
  r += (-139.0 / 840.0) * s[2];
  r += (533.0 / 840.0) * s[3];
  r += (533.0 / 840.0) * s[4];
  r += (-139.0 / 840.0) * s[5];
  r += (29.0 / 840.0) * s[6];
  r += (-3.0 / 840.0) * s[7];
  return r;
}

GCS_DEV double central6(const double s[8]) {
  double r = 0.0;
  r += 0.0 * s[0];
  r += (1.0 / 60.0) * s[1];
  r += (-8.0 / 60.0) * s[2];
  r += (37.0 / 60.0) * s[3];
  r += (37.0 / 60.0) * s[4];
  r += (-8.0 / 60.0) * s[5];
  r += (1.0 / 60.0) * s[6];
  r += 0.0 * s[7];
  return r;
}

GCS_DEV double upwind7(double a0, double a1, double a2, double a3, double a4,
                       double a5, double a6) {
  double r = 0.0;
  r += (-3.0 / 420.0) * a0;
  r += (25.0 / 420.0) * a1;
  r += (-101.0 / 420.0) * a2;
  r += (319.0 / 420.0) * a3;
  r += (214.0 / 420.0) * a4;
  r += (-38.0 / 420.0) * a5;
  r += (4.0 / 420.0) * a6;
  return r;
}

// nonorthogonal_reconstruction.c:36-51: PDM limiter (Fortran SIGN(1, x)).
GCS_DEV double pdm(double q0, double q1, double q2, double hi, double a) {
  const double bounded = fmax(fmin(q1, q2), fmin(hi, fmax(q1, q2)));
  const double d0 = a * (q1 - q0);
  const double d1 = a * (q2 - q1);
  const double s0 = copysign(1.0, d0);
  const double s1 = copysign(1.0, d1);
  const double same = fabs(s0 + s1);
  const double slope = bounded - q1;
  return bounded - s1 * fmax(0.0, fabs(slope) - same * fabs(d0));
}

// nonorthogonal_reconstruction.c:61-93 with the interface volume (central8 of
// the stencil volumes) passed in, since every variable of a face shares it.
GCS_DEV void up7_pdm(const double vol[8], double iv, const double q[8],
                     double a, double &left, double &right) {
  double w[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) w[n] = vol[n] * q[n];
  const double hl = upwind7(w[0], w[1], w[2], w[3], w[4], w[5], w[6]) / iv;
  const double hr = upwind7(w[7], w[6], w[5], w[4], w[3], w[2], w[1]) / iv;
  left = pdm(q[2], q[3], q[4], hl, a);
  right = pdm(q[5], q[4], q[3], hr, a);
}

// a / b given y = 1 / b (IEEE-rounded): q = a y, then one correction with
// the exact remainder a - b q. With y the correctly rounded reciprocal this
// is the correctly rounded quotient (Markstein's theorem, P. Markstein, IBM
// J. Res. Dev. 34(1), 1990), bit for bit a / b, in three operations instead
// of a full division sequence. The theorem needs the remainder to be exact,
// i.e. a and the quotient well inside the normal range: zeros (whose sign
// the sequence does not keep), infinities, NaNs, subnormal dividends and
// quotients near underflow or overflow take the plain division instead.
GCS_DEV double div_by_reciprocal(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-q, b, a);
  const double x = fma(r, y, q);
  const double ax = fabs(a), xx = fabs(x);
  if (__builtin_expect(ax >= 0x1p-960 && xx >= 0x1p-960 && xx <= 0x1p1000, 1)) return x;
  return a / b;
}

// up7_pdm for the nine variables of one face, which share the interface
// volume: its reciprocal riv = 1 / iv is computed once by the caller.
GCS_DEV void up7_pdm_shared(const double vol[8], double iv, double riv, const double q[8],
                            double a, double &left, double &right) {
  double w[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) w[n] = vol[n] * q[n];
  const double hl = div_by_reciprocal(upwind7(w[0], w[1], w[2], w[3], w[4], w[5], w[6]), iv, riv);
  const double hr = div_by_reciprocal(upwind7(w[7], w[6], w[5], w[4], w[3], w[2], w[1]), iv, riv);
  left = pdm(q[2], q[3], q[4], hl, a);
  right = pdm(q[5], q[4], q[3], hr, a);
}

GCS_DEV bool valid_interface(double rho, const double v[3], double p) {
  return isfinite(rho) && isfinite(p) && rho > DBL_MIN && p > DBL_MIN &&
         isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

struct FaceFrame {
  double n[3], t1[3], t2[3];
};

struct FluidFlux {
  double f[5], jump[5], vjump[3], un[2];
};

// nonorthogonal_flux.c:38-105: gas-kinetic half-space fluid flux summed
// over the left (sign -1) and right (sign +1) states.
GCS_DEV void fluid_flux(const Prim w[2], double gamma, const FaceFrame &fr,
                        FluidFlux &out) {
  for (int q = 0; q < 5; ++q) out.f[q] = out.jump[q] = 0.0;
  for (int d = 0; d < 3; ++d) out.vjump[d] = 0.0;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const double sign = side == 0 ? -1.0 : 1.0;
    const double rho = w[side].rho, p = w[side].p;
    const double *v = w[side].v;
    const double v2 = dot3(v, v);
    const double e = 0.5 * rho * v2 + p / (gamma - 1.0);
    const double lambda = rho / (2.0 * p);
    const double un = dot3(fr.n, v);
    const double ut1 = dot3(fr.t1, v);
    const double ut2 = dot3(fr.t2, v);
    const double m0 = 0.5 * erfc(sign * sqrt(lambda) * un);
    const double m1 =
        un * m0 - 0.5 * sign * exp(-lambda * un * un) / sqrt(kPi * lambda);
    out.un[side] = un;
    out.f[0] += rho * m1;
    out.f[4] += (e + 0.5 * p) * m1 + 0.5 * p * m0 * un;
    const double mn = rho * un * m1 + p * m0;
    const double mt1 = rho * ut1 * m1;
    const double mt2 = rho * ut2 * m1;
    for (int d = 0; d < 3; ++d) {
      out.f[1 + d] += fr.n[d] * mn + fr.t1[d] * mt1 + fr.t2[d] * mt2;
      out.vjump[d] += sign * v[d];
    }
    out.jump[0] += sign * rho;
    for (int d = 0; d < 3; ++d) out.jump[1 + d] += sign * rho * v[d];
    out.jump[4] += sign * e;
  }
}

struct MaxwellFlux {
  double m[3], va, psum;
};

// nonorthogonal_flux.c:107-173 (background on, Boris optional).
GCS_DEV bool maxwell_flux(const Prim w[2], const double b[2][3], double bn,
                          const double n[3], const double b0[3], double b0n,
                          bool boris, double c, const double un[2],
                          MaxwellFlux &out) {
  if (!isfinite(bn)) return false;
  for (int side = 0; side < 2; ++side)
    for (int d = 0; d < 3; ++d)
      if (!isfinite(b[side][d]) || !isfinite(b0[d])) return false;
  out.m[0] = out.m[1] = out.m[2] = 0.0;
  out.va = 0.0;
  out.psum = 0.0;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const double *bs = b[side];
    const double pr = 0.5 * dot3(bs, bs);
    double tot[3];
    for (int d = 0; d < 3; ++d) tot[d] = bs[d] + b0[d];
    const double t2 = dot3(tot, tot);
    const double va2 = t2 / w[side].rho;
    const double va2c = boris ? va2 / (1.0 + va2 / (c * c)) : va2;
    const double lambda = 1.0 / (2.0 * w[side].p / w[side].rho + va2c);
    const double m0 =
        0.5 * erfc((side == 0 ? -1.0 : 1.0) * sqrt(lambda) * un[side]);
    out.va += 0.5 * sqrt(va2c);
    out.psum += 0.5 * t2;
    const double bdb = dot3(b0, bs);
    for (int d = 0; d < 3; ++d) {
      double stress = pr * n[d] - bs[d] * bn;
      stress += -b0[d] * bn - bs[d] * b0n + bdb * n[d];
      out.m[d] += m0 * stress;
    }
  }
  return true;
}

// nonorthogonal_flux.c:175-204
GCS_DEV void hogs(FluidFlux &f, MaxwellFlux &m, double ch, double cm,
                  bool boris, double c) {
  for (int q = 0; q < 5; ++q) f.f[q] -= ch * m.va * f.jump[q];
  if (boris) {
    const double ic2 = 1.0 / (c * c);
    for (int d = 0; d < 3; ++d) {
      const double jump = m.psum * f.vjump[d] * ic2;
      m.m[d] -= cm * m.va * jump;
    }
  }
}

// nonorthogonal_operators.c:56-79 with the geometry-only part (det, its
// well-posedness test, 1/det) done at setup: s = {n1x, n1y, n2x, n2y, 1/det}.
GCS_DEV void edge_field(const double s[5], double f1, double f2,
                        double out[2]) {
  out[0] = s[4] * (s[3] * f1 - s[1] * f2);
  out[1] = s[4] * (-s[2] * f1 + s[0] * f2);
}

// nonorthogonal_operators.c:144-182
GCS_DEV bool edge_emf(double v1, double v2, double b1, double b2,
                      double jump, double rho, double len, double diff,
                      bool boris, double c, double cfl, double dt,
                      double &emf) {
  if (!isfinite(v1) || !isfinite(v2) || !isfinite(b1) || !isfinite(b2) ||
      !isfinite(jump) || !isfinite(rho) || rho <= DBL_MIN ||
      !isfinite(len) || len <= DBL_MIN)
    return false;
  const double flow = hypot(v1, v2);
  double va = hypot(b1, b2) / sqrt(rho);
  if (boris) va = c * va / hypot(c, va);
  double speed = flow + va;
  speed = fmin(speed, cfl * len / dt);
  const double ideal = -(v1 * b2 - v2 * b1);
  emf = (ideal + diff * speed * jump) * len;
  return true;
}

// Calibrated face-flux -> cell-B recovery: the moment formula of
// nonorthogonal_operators.c:17-54 followed by the 3x3 Gauss-Jordan solve of
// cubedsphere/cubedsphere_recovery.c:17-98. The solve's matrix depends only
// on geometry, so setup records its pivots and multipliers (rec[] =
// inv0 f01 f02 inv1 f10 f12 inv2 f20 f21) and this applies the same row
// operations, in the same order, to the right-hand side only.
// flux = {lo_i, up_i, lo_j, up_j, lo_k, up_k}; fc = matching face centroids.
GCS_DEV bool recover_cell(const double flux[6], const double fc[6][3],
                          const double cen[3], double vol, const double rec[9],
                          int piv, double out[3]) {
  double div = 0.0;
  double r[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double lo = flux[2 * d], up = flux[2 * d + 1];
    div += up - lo;
    for (int c = 0; c < 3; ++c) r[c] += up * fc[2 * d + 1][c] - lo * fc[2 * d][c];
  }
  for (int c = 0; c < 3; ++c) r[c] = (r[c] - div * cen[c]) / vol;
  const int p0 = piv & 3, p1 = (piv >> 2) & 3;
  double t;
  if (p0 != 0) { t = r[0]; r[0] = r[p0]; r[p0] = t; }
  r[0] *= rec[0];
  r[1] -= rec[1] * r[0];
  r[2] -= rec[2] * r[0];
  if (p1 != 1) { t = r[1]; r[1] = r[p1]; r[p1] = t; }
  r[1] *= rec[3];
  r[0] -= rec[4] * r[1];
  r[2] -= rec[5] * r[1];
  r[2] *= rec[6];
  r[0] -= rec[7] * r[2];
  r[1] -= rec[8] * r[2];
  if (!isfinite(r[0]) || !isfinite(r[1]) || !isfinite(r[2])) return false;
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
  return true;
}

GCS_DEV void scale3(double s, const double v[3], double o[3]) {
  o[0] = s * v[0];
  o[1] = s * v[1];
  o[2] = s * v[2];
}

GCS_DEV void perp3(const double v[3], const double e[3], double o[3]) {
  double par[3];
  scale3(dot3(v, e), e, par);
  for (int d = 0; d < 3; ++d) o[d] = v[d] - par[d];
}

// nonorthogonal_state.c:205-280: semi-relativistic Boris split of the
// momentum increment; bn, bo are total (residual + B0) cell fields.
GCS_DEV bool boris_update(double u[5], const double uo[5], const double bn[3],
                          const double bo[3], const double rate[5],
                          const double mr[3], double c, double dt,
                          double gamma, double rf, double pf) {
  Prim w;
  if (!cons_to_prim(u, gamma, rf, pf, w) || uo[0] <= DBL_MIN) return false;
  const double rho = fmax(w.rho, rf);
  const double p = fmax(w.p, pf);
  const double rho_old = uo[0];
  double v_old[3];
  for (int d = 0; d < 3; ++d) v_old[d] = uo[1 + d] / fmax(rho_old, rf);
  double hb[3];
  for (int d = 0; d < 3; ++d) hb[d] = 0.5 * (bn[d] + bo[d]);
  const double alpha = dot3(hb, hb) / (rho * c * c);
  const double drho = dt * rate[0];
  double dg[3], dmag[3], dm[3];
  for (int d = 0; d < 3; ++d) {
    dg[d] = dt * rate[1 + d];
    dmag[d] = dt * mr[d];
  }
  if (alpha > 1.0e-12) {
    const double plasma = 1.0 / (1.0 + alpha);
    double e[3] = {0.0, 0.0, 0.0};
    const double norm = sqrt(dot3(hb, hb));
    if (!(norm <= 1.0e-12)) scale3(1.0 / norm, hb, e);
    double tot[3], par[3], pt[3], pv[3], a1[3], a2[3];
    for (int d = 0; d < 3; ++d) tot[d] = dg[d] + dmag[d];
    scale3(dot3(dg, e), e, par);
    perp3(tot, e, pt);
    perp3(v_old, e, pv);
    scale3(plasma, pt, a1);
    scale3(alpha * plasma * drho, pv, a2);
    for (int d = 0; d < 3; ++d) dm[d] = par[d] + (a1[d] + a2[d]);
  } else {
    for (int d = 0; d < 3; ++d) dm[d] = dg[d] + dmag[d];
  }
  Prim r;
  r.rho = rho;
  r.p = p;
  for (int d = 0; d < 3; ++d) r.v[d] = (rho_old * v_old[d] + dm[d]) / rho;
  return prim_to_cons(r, gamma, rf, pf, u);
}

// nonorthogonal_state.c:169-203 (update without the Boris correction).
GCS_DEV void maxwell_update(double u[5], const double mr[3], double dt,
                            double gamma, double rf, double pf) {
  const double rho = fmax(u[0], rf);
  double m2 = 0.0;
  for (int d = 0; d < 3; ++d) m2 += u[1 + d] * u[1 + d];
  const double p = fmax((gamma - 1.0) * (u[4] - 0.5 * m2 / rho), pf);
  double m2n = 0.0;
  for (int d = 0; d < 3; ++d) {
    u[1 + d] += dt * mr[d];
    m2n += u[1 + d] * u[1 + d];
  }
  u[0] = rho;
  u[4] = 0.5 * m2n / rho + p / (gamma - 1.0);
}

}  // namespace gcs
