// Reconstruction, face-flux and edge-EMF numerics templated on the compute
// type R, for the mixed-precision step: R = float computes Up7/PDM, the
// kinetic fluid and Maxwell fluxes, HOGS and the edge EMF in FP32 while the
// caller keeps the conserved-variable accumulation, the CT face-flux update,
// the CFL reduction and the ionosphere solve in FP64.
//
// The formulas are those of gcs_math.hpp (same reference routines, same
// order of operations); constants are converted to R so that no expression
// is silently promoted to FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>

#include "gcs_math.hpp"

namespace gcs {
namespace mixed {

template <typename R>
struct PrimR {
  R rho, v[3], p;
};

template <typename R>
GCS_DEV R tiny() { return std::numeric_limits<R>::min(); }

template <typename R>
GCS_DEV R dot3r(const R a[3], const R b[3]) {
  R r = R(0);
  r += a[0] * b[0];
  r += a[1] * b[1];
  r += a[2] * b[2];
  return r;
}

// nonorthogonal_state.c:60-93
template <typename R>
GCS_DEV bool cons_to_prim_r(const R u[5], R gamma, R rf, R pf, PrimR<R> &w) {
  const R rho = u[0];
  if (!isfinite(rho) || rho <= tiny<R>()) return false;
  R v2 = R(0);
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(u[1 + d])) return false;
    w.v[d] = u[1 + d] / rho;
    v2 += w.v[d] * w.v[d];
  }
  if (!isfinite(u[4])) return false;
  const R kinetic = R(0.5) * rho * v2;
  w.rho = fmax(rho, rf);
  w.p = fmax((gamma - R(1)) * (u[4] - kinetic), pf);
  return true;
}

// nonorthogonal_reconstruction.c:7-26
template <typename R>
GCS_DEV R central8_r(const R s[8]) {
  R r = R(0);
  r += R(-3.0 / 840.0) * s[0];
  r += R(29.0 / 840.0) * s[1];
  r += R(-139.0 / 840.0) * s[2];
  r += R(533.0 / 840.0) * s[3];
  r += R(533.0 / 840.0) * s[4];
  r += R(-139.0 / 840.0) * s[5];
  r += R(29.0 / 840.0) * s[6];
  r += R(-3.0 / 840.0) * s[7];
  return r;
}

template <typename R>
GCS_DEV R upwind7_r(R a0, R a1, R a2, R a3, R a4, R a5, R a6) {
  R r = R(0);
  r += R(-3.0 / 420.0) * a0;
  r += R(25.0 / 420.0) * a1;
  r += R(-101.0 / 420.0) * a2;
  r += R(319.0 / 420.0) * a3;
  r += R(214.0 / 420.0) * a4;
  r += R(-38.0 / 420.0) * a5;
  r += R(4.0 / 420.0) * a6;
  return r;
}

// nonorthogonal_reconstruction.c:36-51
template <typename R>
GCS_DEV R pdm_r(R q0, R q1, R q2, R hi, R a) {
  const R bounded = fmax(fmin(q1, q2), fmin(hi, fmax(q1, q2)));
  const R d0 = a * (q1 - q0);
  const R d1 = a * (q2 - q1);
  const R s0 = copysign(R(1), d0);
  const R s1 = copysign(R(1), d1);
  const R same = fabs(s0 + s1);
  const R slope = bounded - q1;
  return bounded - s1 * fmax(R(0), fabs(slope) - same * fabs(d0));
}

// nonorthogonal_reconstruction.c:61-93
template <typename R>
GCS_DEV void up7_pdm_r(const R vol[8], R iv, const R q[8], R a, R &left, R &right) {
  R w[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) w[n] = vol[n] * q[n];
  const R hl = upwind7_r(w[0], w[1], w[2], w[3], w[4], w[5], w[6]) / iv;
  const R hr = upwind7_r(w[7], w[6], w[5], w[4], w[3], w[2], w[1]) / iv;
  left = pdm_r(q[2], q[3], q[4], hl, a);
  right = pdm_r(q[5], q[4], q[3], hr, a);
}

// up7_pdm_r for the nine variables of one face, which share the interface
// volume: a / iv as a * riv (riv = 1 / iv, once per face) plus one fused
// correction, the correctly rounded quotient (gcs::div_by_reciprocal), with
// the plain division outside the range where the correction is exact.
template <typename R>
GCS_DEV void up7_pdm_shared_r(const R vol[8], R iv, R riv, const R q[8], R a, R &left,
                              R &right) {
  R w[8];
#pragma unroll
  for (int n = 0; n < 8; ++n) w[n] = vol[n] * q[n];
  auto div = [&](R x) {
    const R y = x * riv;
    const R z = fma(fma(-y, iv, x), riv, y);
    const R lo = (R)(sizeof(R) == 4 ? 0x1p-100 : 0x1p-960);
    const R hi = (R)(sizeof(R) == 4 ? 0x1p120 : 0x1p1000);
    if (__builtin_expect(fabs(x) >= lo && fabs(z) >= lo && fabs(z) <= hi, 1)) return z;
    return x / iv;
  };
  const R hl = div(upwind7_r(w[0], w[1], w[2], w[3], w[4], w[5], w[6]));
  const R hr = div(upwind7_r(w[7], w[6], w[5], w[4], w[3], w[2], w[1]));
  left = pdm_r(q[2], q[3], q[4], hl, a);
  right = pdm_r(q[5], q[4], q[3], hr, a);
}

template <typename R>
GCS_DEV bool valid_interface_r(R rho, const R v[3], R p) {
  return isfinite(rho) && isfinite(p) && rho > tiny<R>() && p > tiny<R>() &&
         isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]);
}

template <typename R>
struct FaceFrameR {
  R n[3], t1[3], t2[3];
};

template <typename R>
struct FluidFluxR {
  R f[5], jump[5], vjump[3], un[2];
};

// nonorthogonal_flux.c:38-105
template <typename R>
GCS_DEV void fluid_flux_r(const PrimR<R> w[2], R gamma, const FaceFrameR<R> &fr,
                          FluidFluxR<R> &out) {
  for (int q = 0; q < 5; ++q) out.f[q] = out.jump[q] = R(0);
  for (int d = 0; d < 3; ++d) out.vjump[d] = R(0);
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const R sign = side == 0 ? R(-1) : R(1);
    const R rho = w[side].rho, p = w[side].p;
    const R *v = w[side].v;
    const R v2 = dot3r(v, v);
    const R e = R(0.5) * rho * v2 + p / (gamma - R(1));
    const R lambda = rho / (R(2) * p);
    const R un = dot3r(fr.n, v);
    const R ut1 = dot3r(fr.t1, v);
    const R ut2 = dot3r(fr.t2, v);
    const R m0 = R(0.5) * erfc(sign * sqrt(lambda) * un);
    const R m1 = un * m0 - R(0.5) * sign * exp(-lambda * un * un) / sqrt(R(kPi) * lambda);
    out.un[side] = un;
    out.f[0] += rho * m1;
    out.f[4] += (e + R(0.5) * p) * m1 + R(0.5) * p * m0 * un;
    const R mn = rho * un * m1 + p * m0;
    const R mt1 = rho * ut1 * m1;
    const R mt2 = rho * ut2 * m1;
    for (int d = 0; d < 3; ++d) {
      out.f[1 + d] += fr.n[d] * mn + fr.t1[d] * mt1 + fr.t2[d] * mt2;
      out.vjump[d] += sign * v[d];
    }
    out.jump[0] += sign * rho;
    for (int d = 0; d < 3; ++d) out.jump[1 + d] += sign * rho * v[d];
    out.jump[4] += sign * e;
  }
}

template <typename R>
struct MaxwellFluxR {
  R m[3], va, psum;
};

// nonorthogonal_flux.c:107-173
template <typename R>
GCS_DEV bool maxwell_flux_r(const PrimR<R> w[2], const R b[2][3], R bn, const R n[3],
                            const R b0[3], R b0n, bool boris, R c, const R un[2],
                            MaxwellFluxR<R> &out) {
  if (!isfinite(bn)) return false;
  for (int side = 0; side < 2; ++side)
    for (int d = 0; d < 3; ++d)
      if (!isfinite(b[side][d]) || !isfinite(b0[d])) return false;
  out.m[0] = out.m[1] = out.m[2] = R(0);
  out.va = R(0);
  out.psum = R(0);
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const R *bs = b[side];
    const R pr = R(0.5) * dot3r(bs, bs);
    R tot[3];
    for (int d = 0; d < 3; ++d) tot[d] = bs[d] + b0[d];
    const R t2 = dot3r(tot, tot);
    const R va2 = t2 / w[side].rho;
    const R va2c = boris ? va2 / (R(1) + va2 / (c * c)) : va2;
    const R lambda = R(1) / (R(2) * w[side].p / w[side].rho + va2c);
    const R m0 = R(0.5) * erfc((side == 0 ? R(-1) : R(1)) * sqrt(lambda) * un[side]);
    out.va += R(0.5) * sqrt(va2c);
    out.psum += R(0.5) * t2;
    const R bdb = dot3r(b0, bs);
    for (int d = 0; d < 3; ++d) {
      R stress = pr * n[d] - bs[d] * bn;
      stress += -b0[d] * bn - bs[d] * b0n + bdb * n[d];
      out.m[d] += m0 * stress;
    }
  }
  return true;
}

// nonorthogonal_flux.c:175-204
template <typename R>
GCS_DEV void hogs_r(FluidFluxR<R> &f, MaxwellFluxR<R> &m, R ch, R cm, bool boris, R c) {
  for (int q = 0; q < 5; ++q) f.f[q] -= ch * m.va * f.jump[q];
  if (boris) {
    const R ic2 = R(1) / (c * c);
    for (int d = 0; d < 3; ++d) {
      const R jump = m.psum * f.vjump[d] * ic2;
      m.m[d] -= cm * m.va * jump;
    }
  }
}

// nonorthogonal_operators.c:56-79 (geometry-only part precomputed)
template <typename R>
GCS_DEV void edge_field_r(const R s[5], R f1, R f2, R out[2]) {
  out[0] = s[4] * (s[3] * f1 - s[1] * f2);
  out[1] = s[4] * (-s[2] * f1 + s[0] * f2);
}

// nonorthogonal_operators.c:144-182
template <typename R>
GCS_DEV bool edge_emf_r(R v1, R v2, R b1, R b2, R jump, R rho, R len, R diff, bool boris,
                        R c, R cfl, R dt, R &emf) {
  if (!isfinite(v1) || !isfinite(v2) || !isfinite(b1) || !isfinite(b2) ||
      !isfinite(jump) || !isfinite(rho) || rho <= tiny<R>() || !isfinite(len) ||
      len <= tiny<R>())
    return false;
  const R flow = hypot(v1, v2);
  R va = hypot(b1, b2) / sqrt(rho);
  if (boris) va = c * va / hypot(c, va);
  R speed = flow + va;
  speed = fmin(speed, cfl * len / dt);
  const R ideal = -(v1 * b2 - v2 * b1);
  emf = (ideal + diff * speed * jump) * len;
  return true;
}

}  // namespace mixed
}  // namespace gcs
