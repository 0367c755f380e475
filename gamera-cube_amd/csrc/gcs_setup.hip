// Device-side problem setup: cube vertices, hexahedral geometry with the
// 12-point Gauss-Legendre rule, face-frame calibration, split-field dipole
// background, Problem-21 initial state and the residual face fluxes.
//
// Sequence and arithmetic follow the reference setup of the serial Earth
// driver (cubedsphere/tests/test_cubedsphere_dipole_background.c:330-378):
// every quadrature sums in the reference's order and this file is compiled
// without FMA contraction, so the geometry and background are bitwise the
// reference's. Transcendentals (tan, expm1 of the radial map) are evaluated
// on the host with the C library, like the reference.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "gcs_device.hpp"
#include "gcs_math.hpp"

namespace gcs {

namespace {

// 12-point Gauss-Legendre nodes and weights on [-1, 1] (Abramowitz &
// Stegun table 25.4), positive half first: the order in which every
// reference quadrature sums (nonorthogonal_geometry.c:9-23).
__constant__ double kNode[12] = {
    0.1252334085114689154724,  0.3678314989981801937527,
    0.5873179542866174472967,  0.7699026741943046870369,
    0.9041172563704748566785,  0.9815606342467192506906,
    -0.1252334085114689154724, -0.3678314989981801937527,
    -0.5873179542866174472967, -0.7699026741943046870369,
    -0.9041172563704748566785, -0.9815606342467192506906};
__constant__ double kWeight[12] = {
    0.2491470458134027850006, 0.2334925365383548087610,
    0.2031674267230659217490, 0.1600783285433462263350,
    0.1069393259953184309603, 0.0471753363865118271946,
    0.2491470458134027850006, 0.2334925365383548087610,
    0.2031674267230659217490, 0.1600783285433462263350,
    0.1069393259953184309603, 0.0471753363865118271946};

// Cube-face frames {normal, tangent_a, tangent_b} (cubedsphere_mesh.c:8-14).
__constant__ int kFrame[6][3][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},   {{-1, 0, 0}, {0, -1, 0}, {0, 0, 1}},
    {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}},  {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}},   {{0, 0, -1}, {1, 0, 0}, {0, -1, 0}}};

// Hexahedron corner order and trilinear signs (nonorthogonal_geometry.h:
// 33-36, nonorthogonal_geometry.c:25-27).
__constant__ int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

struct V3 {
  double x[3];
};

GCS_DEV V3 vsub(const V3 &a, const V3 &b) {
  return {{a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]}};
}
GCS_DEV V3 vadd(const V3 &a, const V3 &b) {
  return {{a.x[0] + b.x[0], a.x[1] + b.x[1], a.x[2] + b.x[2]}};
}
GCS_DEV V3 vmul(const V3 &a, double s) {
  return {{a.x[0] * s, a.x[1] * s, a.x[2] * s}};
}
GCS_DEV double vdot(const V3 &a, const V3 &b) {
  double r = 0.0;
  r += a.x[0] * b.x[0];
  r += a.x[1] * b.x[1];
  r += a.x[2] * b.x[2];
  return r;
}
GCS_DEV V3 vcross(const V3 &a, const V3 &b) {
  return {{a.x[1] * b.x[2] - a.x[2] * b.x[1], a.x[2] * b.x[0] - a.x[0] * b.x[2],
           a.x[0] * b.x[1] - a.x[1] * b.x[0]}};
}
// nonorthogonal_geometry.c:83-90 (unit = v * (1 / |v|)).
GCS_DEV bool vunit(const V3 &v, V3 &u) {
  const double norm = sqrt(vdot(v, v));
  if (!isfinite(norm) || norm <= DBL_MIN) return false;
  u = vmul(v, 1.0 / norm);
  return true;
}

struct Ctx {
  Layout L;
  double *pl;
  uint8_t *by;
  StepScalars *sc;
  __device__ double *p(int q) const { return pl + (long)q * L.total; }
  __device__ uint8_t *b(int q) const { return by + (long)q * L.total; }
  __device__ V3 vertex(long at) const {
    return {{p(P_VERT)[at], p(P_VERT + 1)[at], p(P_VERT + 2)[at]}};
  }
  __device__ void fail(unsigned stage) const {
    atomicOr(&sc->failed, 1u);
    atomicCAS(&sc->fail_stage, 0u, stage);
  }
};

// Decode a 1-D thread index over an (f, i, j, k) box of extents e.
GCS_DEV bool box_index(long t, int e0, int e1, int e2, int &f, int &i, int &j,
                       int &k) {
  const long per = (long)e0 * e1 * e2;
  if (t >= 6 * per) return false;
  f = (int)(t / per);
  long r = t - f * per;
  i = (int)(r / ((long)e1 * e2));
  r -= (long)i * e1 * e2;
  j = (int)(r / e2);
  k = (int)(r - (long)j * e2);
  return true;
}

// cubedsphere_mesh.c:58-75, :198-221 with tan() of both angles precomputed.
__global__ void k_vertices(Ctx c, const double *radius, const double *tan_a,
                           const double *tan_b) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + 1,
                 L.NJ + 1, L.NK + 1, f, i, j, k))
    return;
  const double u = tan_a[j], v = tan_b[k];
  const double il = 1.0 / sqrt(1.0 + u * u + v * v);
  const long at = L.at(f, i, j, k);
  for (int d = 0; d < 3; ++d) {
    const double dir = il * ((double)kFrame[f][0][d] + u * (double)kFrame[f][1][d] +
                             v * (double)kFrame[f][2][d]);
    c.p(P_VERT + d)[at] = radius[i] * dir;
  }
}

// Corners of face (i, j, k) of direction d, in the order the reference's
// cell geometry passes them (nonorthogonal_geometry.c:29-33, :226-234).
GCS_DEV void face_corners(const Ctx &c, int d, int f, int i, int j, int k,
                          V3 q[4]) {
  const Layout &L = c.L;
  const int o[3][4][3] = {{{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}},
                          {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 1}},
                          {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}}};
  for (int n = 0; n < 4; ++n)
    q[n] = c.vertex(L.at(f, i + o[d][n][0], j + o[d][n][1], k + o[d][n][2]));
}

// nonorthogonal_geometry.c:92-103
GCS_DEV V3 bilinear(const V3 q[4], double eta, double psi) {
  const V3 de = vsub(q[1], q[0]);
  const V3 dp = vsub(q[2], q[0]);
  const V3 dc = vadd(vsub(q[3], q[2]), vsub(q[0], q[1]));
  return vadd(q[0], vadd(vmul(de, eta), vadd(vmul(dp, psi), vmul(dc, eta * psi))));
}

// nonorthogonal_geometry.c:105-160 then cubedsphere_mesh.c:256-284.
__global__ void k_face_geometry(Ctx c, int d) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (d == 0),
                 L.NJ + (d == 1), L.NK + (d == 2), f, i, j, k))
    return;
  V3 q[4];
  face_corners(c, d, f, i, j, k, q);
  const V3 de = vsub(q[1], q[0]);
  const V3 dp = vsub(q[2], q[0]);
  const V3 dc = vadd(vsub(q[3], q[2]), vsub(q[0], q[1]));
  double area = 0.0;
  V3 av = {{0.0, 0.0, 0.0}}, cen = {{0.0, 0.0, 0.0}};
  for (int a = 0; a < 12; ++a) {
    const double eta = 0.5 * (1.0 + kNode[a]);
    for (int b = 0; b < 12; ++b) {
      const double psi = 0.5 * (1.0 + kNode[b]);
      const double w = 0.25 * kWeight[a] * kWeight[b];
      const V3 te = vadd(de, vmul(dc, psi));
      const V3 tp = vadd(dp, vmul(dc, eta));
      const V3 oa = vcross(te, tp);
      const double aw = w * sqrt(vdot(oa, oa));
      const V3 pt = bilinear(q, eta, psi);
      area += aw;
      av = vadd(av, vmul(oa, w));
      cen = vadd(cen, vmul(pt, aw));
    }
  }
  const long at = L.at(f, i, j, k);
  const int base = P_FACE + FACE_PLANES * d;
  if (!isfinite(area) || area <= DBL_MIN) {
    c.fail(101);
    return;
  }
  cen = vmul(cen, 1.0 / area);
  const V3 ce = vmul(vadd(vsub(q[3], q[2]), vsub(q[1], q[0])), 0.5);
  const V3 cp = vmul(vadd(vsub(q[3], q[1]), vsub(q[2], q[0])), 0.5);
  V3 n0, t2, t1;
  if (!vunit(vcross(ce, cp), n0) || !vunit(cp, t2) || !vunit(vcross(t2, n0), t1)) {
    c.fail(102);
    return;
  }
  (void)t1;
  c.p(P_SCRATCH_AREA + d)[at] = area;
  // Calibration: the integrated area vector defines area and normal.
  const double mag = sqrt(av.x[0] * av.x[0] + av.x[1] * av.x[1] + av.x[2] * av.x[2]);
  if (!isfinite(mag) || mag <= 0.0) {
    c.fail(103);
    return;
  }
  V3 nn;
  for (int e = 0; e < 3; ++e) nn.x[e] = av.x[e] / mag;
  double pr = 0.0;
  for (int e = 0; e < 3; ++e) pr += t2.x[e] * nn.x[e];
  V3 ct2;
  double t2sq = 0.0;
  for (int e = 0; e < 3; ++e) {
    ct2.x[e] = t2.x[e] - pr * nn.x[e];
    t2sq += ct2.x[e] * ct2.x[e];
  }
  if (!isfinite(t2sq) || t2sq <= 0.0) {
    c.fail(104);
    return;
  }
  for (int e = 0; e < 3; ++e) ct2.x[e] /= sqrt(t2sq);
  const V3 ct1 = {{ct2.x[1] * nn.x[2] - ct2.x[2] * nn.x[1],
                   ct2.x[2] * nn.x[0] - ct2.x[0] * nn.x[2],
                   ct2.x[0] * nn.x[1] - ct2.x[1] * nn.x[0]}};
  c.p(base + FO_AREA)[at] = mag;
  for (int e = 0; e < 3; ++e) {
    c.p(base + FO_AV + e)[at] = av.x[e];
    c.p(base + FO_CEN + e)[at] = cen.x[e];
    c.p(base + FO_N + e)[at] = nn.x[e];
    c.p(base + FO_T1 + e)[at] = ct1.x[e];
    c.p(base + FO_T2 + e)[at] = ct2.x[e];
  }
}

// Moment part of the recovery (nonorthogonal_operators.c:17-54).
GCS_DEV void moment(const double flux[6], const double fc[6][3],
                    const double cen[3], double vol, double out[3]) {
  double div = 0.0;
  out[0] = out[1] = out[2] = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double lo = flux[2 * d], up = flux[2 * d + 1];
    div += up - lo;
    for (int e = 0; e < 3; ++e) out[e] += up * fc[2 * d + 1][e] - lo * fc[2 * d][e];
  }
  for (int e = 0; e < 3; ++e) out[e] = (out[e] - div * cen[e]) / vol;
}

// Cell volume and centroid (nonorthogonal_geometry.c:162-213), CFL length
// (:236-240, uncalibrated face areas) and the pivots and multipliers of the
// calibrated recovery's Gauss-Jordan solve (cubedsphere_recovery.c:17-98).
__global__ void k_cell_geometry(Ctx c) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI, L.NJ, L.NK,
                 f, i, j, k))
    return;
  V3 q[8];
  for (int n = 0; n < 8; ++n)
    q[n] = c.vertex(L.at(f, i + kCorner[n][0], j + kCorner[n][1], k + kCorner[n][2]));
  double vol = 0.0;
  V3 cen = {{0.0, 0.0, 0.0}};
  for (int a = 0; a < 12; ++a) {
    const double xi = kNode[a];
    for (int b = 0; b < 12; ++b) {
      const double eta = kNode[b];
      for (int e = 0; e < 12; ++e) {
        const double zeta = kNode[e];
        V3 pt = {{0.0, 0.0, 0.0}};
        V3 der[3] = {{{0.0, 0.0, 0.0}}, {{0.0, 0.0, 0.0}}, {{0.0, 0.0, 0.0}}};
        for (int n = 0; n < 8; ++n) {
          const double sx = kCorner[n][0] ? 1.0 : -1.0;
          const double sy = kCorner[n][1] ? 1.0 : -1.0;
          const double sz = kCorner[n][2] ? 1.0 : -1.0;
          const double fx = 1.0 + sx * xi;
          const double fy = 1.0 + sy * eta;
          const double fz = 1.0 + sz * zeta;
          const double shape = 0.125 * fx * fy * fz;
          const double ds[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz,
                                0.125 * fx * fy * sz};
          pt = vadd(pt, vmul(q[n], shape));
          for (int m = 0; m < 3; ++m) der[m] = vadd(der[m], vmul(q[n], ds[m]));
        }
        const double det = vdot(der[0], vcross(der[1], der[2]));
        const double wv = fabs(det) * kWeight[a] * kWeight[b] * kWeight[e];
        vol += wv;
        cen = vadd(cen, vmul(pt, wv));
      }
    }
  }
  const long at = L.at(f, i, j, k);
  if (!isfinite(vol) || vol <= DBL_MIN) {
    c.fail(111);
    return;
  }
  cen = vmul(cen, 1.0 / vol);
  c.p(P_VOL)[at] = vol;
  for (int e = 0; e < 3; ++e) c.p(P_CEN + e)[at] = cen.x[e];
  // CFL length: volume / larger uncalibrated area of each direction.
  const long step[3] = {L.di(), (long)L.SK, 1};
  for (int d = 0; d < 3; ++d) {
    const double *ua = c.p(P_SCRATCH_AREA + d);
    c.p(P_CFL + d)[at] = vol / fmax(ua[at], ua[at + step[d]]);
  }
}

// Calibration of the cell-field recovery (cubedsphere_recovery.c:12-98): the
// response of the moment formula to the three unit uniform fields, inverted
// by Gauss-Jordan elimination with partial pivoting. The pivots and
// multipliers depend on geometry only and are kept, so the per-step
// recovery replays the elimination on the moments alone.
__global__ void k_recovery_constants(Ctx c) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI, L.NJ, L.NK,
                 f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const long step[3] = {L.di(), (long)L.SK, 1};
  const double vol = c.p(P_VOL)[at];
  const V3 cen = {{c.p(P_CEN)[at], c.p(P_CEN + 1)[at], c.p(P_CEN + 2)[at]}};
  double fc[6][3], resp[3][3];
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int e = 0; e < 3; ++e)
        fc[2 * d + s][e] = c.p(P_FACE + FACE_PLANES * d + FO_CEN + e)[at + s * step[d]];
  for (int basis = 0; basis < 3; ++basis) {
    double bf[6], col[3];
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s)
        bf[2 * d + s] = c.p(P_FACE + FACE_PLANES * d + FO_AV + basis)[at + s * step[d]];
    moment(bf, fc, cen.x, vol, col);
    for (int r = 0; r < 3; ++r) resp[r][basis] = col[r];
  }
  double a[3][3], scale = 0.0;
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) {
      a[r][s] = resp[r][s];
      scale = fmax(scale, fabs(resp[r][s]));
    }
  double rec[9];
  int piv[2] = {0, 1};
  bool ok = isfinite(scale) && scale > DBL_MIN;
  for (int col = 0; col < 3 && ok; ++col) {
    int pv = col;
    for (int r = col + 1; r < 3; ++r)
      if (fabs(a[r][col]) > fabs(a[pv][col])) pv = r;
    if (!isfinite(a[pv][col]) || fabs(a[pv][col]) <= 128.0 * DBL_EPSILON * scale) {
      ok = false;
      break;
    }
    if (col < 2) piv[col] = pv;
    if (pv != col)
      for (int s = col; s < 3; ++s) {
        const double t = a[col][s];
        a[col][s] = a[pv][s];
        a[pv][s] = t;
      }
    const double inv = 1.0 / a[col][col];
    for (int s = col; s < 3; ++s) a[col][s] *= inv;
    rec[3 * col] = inv;
    int slot = 1;
    for (int r = 0; r < 3; ++r) {
      if (r == col) continue;
      const double fct = a[r][col];
      for (int s = col; s < 3; ++s) a[r][s] -= fct * a[col][s];
      rec[3 * col + slot++] = fct;
    }
  }
  if (!ok) {
    c.b(Q_PIV)[at] = 0xff;
    return;
  }
  for (int r = 0; r < 9; ++r) c.p(P_REC + r)[at] = rec[r];
  c.b(Q_PIV)[at] = (uint8_t)(piv[0] | (piv[1] << 2));
}

// nonorthogonal_grid.c:391-472 and nonorthogonal_geometry.c:245-262.
__global__ void k_edge_geometry(Ctx c, int d) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (d != 0),
                 L.NJ + (d != 1), L.NK + (d != 2), f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const bool interior =
      d == 0 ? (j > 0 && j < L.NJ && k > 0 && k < L.NK)
      : d == 1 ? (i > 0 && i < L.NI && k > 0 && k < L.NK)
               : (i > 0 && i < L.NI && j > 0 && j < L.NJ);
  c.b(Q_VALID + d)[at] = 0;
  if (!interior) return;
  const V3 s = c.vertex(at);
  const V3 e = c.vertex(L.at(f, i + (d == 0), j + (d == 1), k + (d == 2)));
  V3 first, second;
  if (d == 0) {
    first = vsub(c.vertex(L.at(f, i, j, k + 1)), c.vertex(L.at(f, i, j, k - 1)));
    second = vsub(c.vertex(L.at(f, i + 1, j, k + 1)), c.vertex(L.at(f, i + 1, j, k - 1)));
  } else if (d == 1) {
    first = vsub(c.vertex(L.at(f, i + 1, j, k)), c.vertex(L.at(f, i - 1, j, k)));
    second = vsub(c.vertex(L.at(f, i + 1, j + 1, k)), c.vertex(L.at(f, i - 1, j + 1, k)));
  } else {
    first = vsub(c.vertex(L.at(f, i, j + 1, k)), c.vertex(L.at(f, i, j - 1, k)));
    second = vsub(c.vertex(L.at(f, i, j + 1, k + 1)), c.vertex(L.at(f, i, j - 1, k + 1)));
  }
  const V3 avg = {{0.5 * (first.x[0] + second.x[0]), 0.5 * (first.x[1] + second.x[1]),
                   0.5 * (first.x[2] + second.x[2])}};
  const V3 ed = vsub(e, s);
  const double len = sqrt(vdot(ed, ed));
  V3 n0, t1, t2;
  if (!isfinite(len) || len <= DBL_MIN || !vunit(ed, n0) ||
      !vunit(vcross(avg, n0), t1) || !vunit(vcross(n0, t1), t2)) {
    c.fail(121);
    return;
  }
  const int base = P_EDGE + EDGE_PLANES * d;
  c.p(base + EO_LEN)[at] = len;
  for (int m = 0; m < 3; ++m) {
    c.p(base + EO_N + m)[at] = n0.x[m];
    c.p(base + EO_T1 + m)[at] = t1.x[m];
    c.p(base + EO_T2 + m)[at] = t2.x[m];
  }
  c.b(Q_VALID + d)[at] = 1;
}

// Geometry-only part of the edge-B solve: face normals of the two
// transverse directions interpolated to the edge (nonorthogonal_operators.c:
// 81-113) and the 2x2 determinant with its test (:56-79). Defined where the
// 8-face stencils exist; the EMF sweep only reads edges inside that range.
__global__ void k_edge_solve(Ctx c, int e) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (e != 0),
                 L.NJ + (e != 1), L.NK + (e != 2), f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  c.b(Q_EBOK + e)[at] = 0;
  if (!c.b(Q_VALID + e)[at]) return;
  const int t1 = (e + 1) % 3, t2 = (e + 2) % 3;
  const int pos[3] = {i, j, k};
  const int ext[3] = {L.NI, L.NJ, L.NK};
  if (pos[t1] < 4 || pos[t2] < 4 || pos[t1] > ext[t1] - 4 || pos[t2] > ext[t2] - 4)
    return;
  const long step[3] = {L.di(), (long)L.SK, 1};
  const int eb = P_EDGE + EDGE_PLANES * e;
  const double et1[3] = {c.p(eb + EO_T1)[at], c.p(eb + EO_T1 + 1)[at], c.p(eb + EO_T1 + 2)[at]};
  const double et2[3] = {c.p(eb + EO_T2)[at], c.p(eb + EO_T2 + 1)[at], c.p(eb + EO_T2 + 2)[at]};
  double nn[2][2];
  for (int side = 0; side < 2; ++side) {
    const int fd = side == 0 ? t1 : t2;     // face direction
    const int sd = side == 0 ? t2 : t1;     // stencil direction
    const int fb = P_FACE + FACE_PLANES * fd;
    double area[8], wn[3][8];
    for (int n = 0; n < 8; ++n) {
      const long fa = at + (long)(n - 4) * step[sd];
      area[n] = c.p(fb + FO_AREA)[fa];
      for (int m = 0; m < 3; ++m) wn[m][n] = area[n] * c.p(fb + FO_N + m)[fa];
    }
    const double ia = central6(area);
    if (!isfinite(ia) || fabs(ia) <= DBL_MIN) return;
    double ne[3];
    for (int m = 0; m < 3; ++m) ne[m] = central6(wn[m]) / ia;
    const double p1 = dot3(ne, et1), p2 = dot3(ne, et2);
    const double norm = hypot(p1, p2);
    if (!isfinite(norm) || norm <= DBL_MIN) return;
    nn[side][0] = p1 / norm;
    nn[side][1] = p2 / norm;
  }
  const double det = nn[0][0] * nn[1][1] - nn[1][0] * nn[0][1];
  const double scale = fmax(fmax(fabs(nn[0][0]), fabs(nn[0][1])),
                            fmax(fabs(nn[1][0]), fabs(nn[1][1])));
  if (!isfinite(det) || scale <= DBL_MIN ||
      fabs(det) <= 64.0 * DBL_EPSILON * scale * scale)
    return;
  c.p(eb + EO_SOLVE + 0)[at] = nn[0][0];
  c.p(eb + EO_SOLVE + 1)[at] = nn[0][1];
  c.p(eb + EO_SOLVE + 2)[at] = nn[1][0];
  c.p(eb + EO_SOLVE + 3)[at] = nn[1][1];
  c.p(eb + EO_SOLVE + 4)[at] = 1.0 / det;
  c.b(Q_EBOK + e)[at] = 1;
}

struct Startup {
  double inner_radius, density, speed, mach, gamma, ambient_density,
      ambient_pressure, front_x, front_width, cut_radius, cut_length;
  double imf[3], moment[3];
};

GCS_DEV double radius3(const V3 &x) {
  return sqrt(x.x[0] * x.x[0] + x.x[1] * x.x[1] + x.x[2] * x.x[2]);
}

// cubedsphere_earth_startup.c:52-58
GCS_DEV double ramp_down(double x, double start, double length) {
  const double s = (x - start) / length;
  if (s <= 0.0) return 1.0;
  if (s >= 1.0) return 0.0;
  return 1.0 - 3.0 * s * s + 2.0 * s * s * s;
}

// nonorthogonal_background.c:477-497 scaled by the radial cut
// (cubedsphere_earth_startup.c:131-141).
GCS_DEV bool background_field(const Startup &s, const V3 &x, V3 &out) {
  const V3 m = {{s.moment[0], s.moment[1], s.moment[2]}};
  const double r2 = vdot(x, x);
  if (!(r2 > DBL_MIN) || !isfinite(r2)) return false;
  const double r = sqrt(r2);
  const double irc = 1.0 / (r2 * r);
  const double irf = irc / r2;
  const double pr = vdot(m, x);
  for (int e = 0; e < 3; ++e) out.x[e] = 3.0 * pr * irf * x.x[e] - irc * m.x[e];
  const double w = ramp_down(radius3(x), s.cut_radius, s.cut_length);
  for (int e = 0; e < 3; ++e) out.x[e] *= w;
  return isfinite(out.x[0]) && isfinite(out.x[1]) && isfinite(out.x[2]);
}

// Background face average, flux and pure-B0 stress (nonorthogonal_background.c
// :174-224), on the calibrated scalar area.
__global__ void k_background_face(Ctx c, int d, Startup s) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (d == 0),
                 L.NJ + (d == 1), L.NK + (d == 2), f, i, j, k))
    return;
  V3 q[4];
  face_corners(c, d, f, i, j, k, q);
  const long at = L.at(f, i, j, k);
  const double sa = c.p(P_FACE + FACE_PLANES * d + FO_AREA)[at];
  if (!(sa > DBL_MIN) || !isfinite(sa)) {
    c.fail(131);
    return;
  }
  const V3 de = vsub(q[1], q[0]);
  const V3 dp = vsub(q[2], q[0]);
  const V3 dc = vadd(vsub(q[3], q[2]), vsub(q[0], q[1]));
  V3 integral = {{0.0, 0.0, 0.0}}, stress = {{0.0, 0.0, 0.0}};
  double flux = 0.0;
  for (int a = 0; a < 12; ++a) {
    const double eta = 0.5 * (1.0 + kNode[a]);
    for (int b = 0; b < 12; ++b) {
      const double psi = 0.5 * (1.0 + kNode[b]);
      const double w = 0.25 * kWeight[a] * kWeight[b];
      const V3 te = vadd(de, vmul(dc, psi));
      const V3 tp = vadd(dp, vmul(dc, eta));
      const V3 oa = vcross(te, tp);
      const double area = sqrt(vdot(oa, oa));
      const V3 pt = bilinear(q, eta, psi);
      V3 fld;
      if (!(area > DBL_MIN) || !isfinite(area) || !background_field(s, pt, fld)) {
        c.fail(132);
        return;
      }
      integral = vadd(integral, vmul(fld, w * area));
      const double nf = vdot(fld, oa);
      flux += w * nf;
      const double mp = 0.5 * vdot(fld, fld);
      stress = vadd(stress, vmul(vsub(vmul(oa, mp), vmul(fld, nf)), w));
    }
  }
  const V3 avg = vmul(integral, 1.0 / sa);
  for (int e = 0; e < 3; ++e) {
    c.p(P_B0F + 3 * d + e)[at] = avg.x[e];
    c.p(P_SCRATCH_STRESS + 3 * d + e)[at] = stress.x[e];
  }
  c.p(P_B0FLUX + d)[at] = flux;
  c.p(P_B0N + d)[at] = flux / sa;
}

// Background cell average and dpB0 force (nonorthogonal_background.c:226-281,
// :400-436).
__global__ void k_background_cell(Ctx c, Startup s) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI, L.NJ, L.NK,
                 f, i, j, k))
    return;
  V3 q[8];
  for (int n = 0; n < 8; ++n)
    q[n] = c.vertex(L.at(f, i + kCorner[n][0], j + kCorner[n][1], k + kCorner[n][2]));
  const long at = L.at(f, i, j, k);
  const double vol = c.p(P_VOL)[at];
  V3 integral = {{0.0, 0.0, 0.0}};
  double iv = 0.0;
  for (int a = 0; a < 12; ++a) {
    const double xi = kNode[a];
    for (int b = 0; b < 12; ++b) {
      const double eta = kNode[b];
      for (int e = 0; e < 12; ++e) {
        const double zeta = kNode[e];
        V3 pt = {{0.0, 0.0, 0.0}};
        V3 der[3] = {{{0.0, 0.0, 0.0}}, {{0.0, 0.0, 0.0}}, {{0.0, 0.0, 0.0}}};
        for (int n = 0; n < 8; ++n) {
          const double sx = kCorner[n][0] ? 1.0 : -1.0;
          const double sy = kCorner[n][1] ? 1.0 : -1.0;
          const double sz = kCorner[n][2] ? 1.0 : -1.0;
          const double fx = 1.0 + sx * xi;
          const double fy = 1.0 + sy * eta;
          const double fz = 1.0 + sz * zeta;
          const double shape = 0.125 * fx * fy * fz;
          const double ds[3] = {0.125 * sx * fy * fz, 0.125 * fx * sy * fz,
                                0.125 * fx * fy * sz};
          pt = vadd(pt, vmul(q[n], shape));
          for (int m = 0; m < 3; ++m) der[m] = vadd(der[m], vmul(q[n], ds[m]));
        }
        const double jac = fabs(vdot(der[0], vcross(der[1], der[2])));
        const double w = jac * kWeight[a] * kWeight[b] * kWeight[e];
        V3 fld;
        if (!isfinite(w) || !background_field(s, pt, fld)) {
          c.fail(141);
          return;
        }
        iv += w;
        integral = vadd(integral, vmul(fld, w));
      }
    }
  }
  if (!(iv > DBL_MIN) || fabs(iv - vol) > 2.0e-11 * vol) {
    c.fail(142);
    return;
  }
  const V3 avg = vmul(integral, 1.0 / vol);
  const long step[3] = {L.di(), (long)L.SK, 1};
  V3 uml = {{0.0, 0.0, 0.0}};
  for (int d = 0; d < 3; ++d) {
    V3 lo, up;
    for (int e = 0; e < 3; ++e) {
      lo.x[e] = c.p(P_SCRATCH_STRESS + 3 * d + e)[at];
      up.x[e] = c.p(P_SCRATCH_STRESS + 3 * d + e)[at + step[d]];
    }
    uml = vadd(uml, vsub(up, lo));
  }
  const V3 force = vmul(uml, -1.0 / vol);
  // Earth startup zeroes the force inside 0.75 cut radius
  // (cubedsphere_earth_startup.c:143-154).
  const V3 cen = {{c.p(P_CEN)[at], c.p(P_CEN + 1)[at], c.p(P_CEN + 2)[at]}};
  const bool inner = radius3(cen) <= 0.75 * s.cut_radius;
  for (int e = 0; e < 3; ++e) {
    c.p(P_B0C + e)[at] = avg.x[e];
    c.p(P_FORCE + e)[at] = inner ? 0.0 : force.x[e];
  }
  if (!isfinite(force.x[0]) || !isfinite(force.x[1]) || !isfinite(force.x[2]))
    c.fail(143);
}

// Edge-averaged B0 (nonorthogonal_background.c:283-303, :438-458) and its
// projections on the edge frame used by the EMF sweep.
__global__ void k_background_edge(Ctx c, int d, Startup s) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (d != 0),
                 L.NJ + (d != 1), L.NK + (d != 2), f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const V3 a = c.vertex(at);
  const V3 z = c.vertex(L.at(f, i + (d == 0), j + (d == 1), k + (d == 2)));
  const V3 disp = vsub(z, a);
  if (!(sqrt(vdot(disp, disp)) > DBL_MIN)) {
    c.fail(151);
    return;
  }
  V3 r = {{0.0, 0.0, 0.0}};
  for (int n = 0; n < 12; ++n) {
    const double t = 0.5 * (1.0 + kNode[n]);
    V3 fld;
    if (!background_field(s, vadd(a, vmul(disp, t)), fld)) {
      c.fail(152);
      return;
    }
    r = vadd(r, vmul(fld, 0.5 * kWeight[n]));
  }
  const int eb = P_EDGE + EDGE_PLANES * d;
  double t1[3], t2[3];
  for (int e = 0; e < 3; ++e) {
    c.p(P_B0E + 3 * d + e)[at] = r.x[e];
    t1[e] = c.p(eb + EO_T1 + e)[at];
    t2[e] = c.p(eb + EO_T2 + e)[at];
  }
  c.p(P_B0ET + 2 * d)[at] = dot3(r.x, t1);
  c.p(P_B0ET + 2 * d + 1)[at] = dot3(r.x, t2);
}

// splitmix64 finaliser (the perturbation hash of gcs_problem_config).
GCS_DEV uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
GCS_DEV double hashed_unit(uint64_t seed, int f, int i, int j, int k, int var) {
  uint64_t x = mix64(seed + (uint64_t)f);
  x = mix64(x + (uint64_t)i);
  x = mix64(x + (uint64_t)j);
  x = mix64(x + (uint64_t)k);
  x = mix64(x + (uint64_t)var);
  return 2.0 * ((double)(x >> 11) * 0x1.0p-53) - 1.0;
}

// Problem-21 startup primitive at the cell centroid
// (cubedsphere_earth_startup.c:89-107), its conserved form with floors
// 1e-6 / 1e-8 (nonorthogonal_initialization.c:149-184), then the optional
// density/pressure perturbation.
__global__ void k_initial_cells(Ctx c, Startup s, double perturbation,
                                uint64_t seed) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI, L.NJ, L.NK,
                 f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const V3 x = {{c.p(P_CEN)[at], c.p(P_CEN + 1)[at], c.p(P_CEN + 2)[at]}};
  const double frac = 0.5 * (1.0 - tanh((x.x[0] - s.front_x) / s.front_width));
  const double pin = s.density * s.speed * s.speed / (s.gamma * s.mach * s.mach);
  Prim w;
  w.rho = s.ambient_density + frac * (s.density - s.ambient_density);
  w.p = s.ambient_pressure + frac * (pin - s.ambient_pressure);
  double wf = (radius3(x) - s.inner_radius) / 2.5;
  wf = fmax(0.0, fmin(1.0, wf));
  const double ramp = wf * wf * (3.0 - 2.0 * wf);
  w.v[0] = frac * s.speed * ramp;
  w.v[1] = w.v[2] = 0.0;
  double u[5];
  if (!prim_to_cons(w, s.gamma, 1e-6, 1e-8, u)) {
    c.fail(161);
    return;
  }
  if (perturbation != 0.0) {
    Prim q;
    if (!cons_to_prim(u, s.gamma, 1e-6, 1e-8, q)) {
      c.fail(162);
      return;
    }
    q.rho *= 1.0 + perturbation * hashed_unit(seed, f, i, j, k, 0);
    q.p *= 1.0 + perturbation * hashed_unit(seed, f, i, j, k, 1);
    if (!prim_to_cons(q, s.gamma, 1e-6, 1e-8, u)) {
      c.fail(163);
      return;
    }
  }
  for (int m = 0; m < 5; ++m) c.p(plane_u(0, m))[at] = u[m];
}

// Startup vector potential (cubedsphere_earth_startup.c:60-73, :109-129).
GCS_DEV bool potential(const Startup &s, const V3 &x, V3 &a) {
  const double r = radius3(x);
  if (!(r > 0.0)) return false;
  const double w = ramp_down(r, s.cut_radius, s.cut_length);
  const double factor = w / (r * r * r);
  const double *m = s.moment;
  a.x[0] = factor * (m[1] * x.x[2] - m[2] * x.x[1]);
  a.x[1] = factor * (m[2] * x.x[0] - m[0] * x.x[2]);
  a.x[2] = factor * (m[0] * x.x[1] - m[1] * x.x[0]);
  a.x[1] -= 0.5 * s.imf[0] * x.x[2];
  a.x[2] += 0.5 * s.imf[0] * x.x[1];
  const double q = (x.x[0] - s.front_x) / s.front_width;
  const double aq = fabs(q);
  const double log_cosh = aq + log1p(exp(-2.0 * aq)) - log(2.0);
  const double front = 0.5 * (x.x[0] - s.front_width * log_cosh - s.front_x -
                              s.front_width * log(2.0));
  a.x[1] += s.imf[2] * front;
  a.x[2] -= s.imf[1] * front;
  return true;
}

// Edge line integrals of A (nonorthogonal_initialization.c:194-210).
__global__ void k_line_integrals(Ctx c, int d, Startup s) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (d != 0),
                 L.NJ + (d != 1), L.NK + (d != 2), f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const V3 a = c.vertex(at);
  const V3 z = c.vertex(L.at(f, i + (d == 0), j + (d == 1), k + (d == 2)));
  const V3 disp = vsub(z, a);
  const V3 mid = {{0.5 * (a.x[0] + z.x[0]), 0.5 * (a.x[1] + z.x[1]),
                   0.5 * (a.x[2] + z.x[2])}};
  double line = 0.0;
  for (int n = 0; n < 12; ++n) {
    const V3 pt = vadd(mid, vmul(disp, 0.5 * kNode[n]));
    V3 pot;
    if (!potential(s, pt, pot)) {
      c.fail(171);
      return;
    }
    line += 0.5 * kWeight[n] * vdot(pot, disp);
  }
  if (!isfinite(line)) c.fail(172);
  c.p(P_SCRATCH_LINE + d)[at] = line;
}

// Face flux as the circulation of A (nonorthogonal_initialization.c:254-313),
// minus the background flux (cubedsphere_earth_startup.c:156-172), both
// time levels.
__global__ void k_initial_faces(Ctx c, int d) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (d == 0),
                 L.NJ + (d == 1), L.NK + (d == 2), f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const long si = L.di(), sj = L.SK;
  const double *I = c.p(P_SCRATCH_LINE), *J = c.p(P_SCRATCH_LINE + 1),
               *K = c.p(P_SCRATCH_LINE + 2);
  double phi;
  if (d == 0) phi = J[at] - J[at + 1] + K[at + sj] - K[at];
  else if (d == 1) phi = -I[at] + I[at + 1] - K[at + si] + K[at];
  else phi = I[at] - I[at + sj] + J[at + si] - J[at];
  c.p(plane_phi(0, d))[at] = phi;
}

__global__ void k_subtract_background(Ctx c, int d) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + (d == 0),
                 L.NJ + (d == 1), L.NK + (d == 2), f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const double b = c.p(P_B0FLUX + d)[at];
  c.p(plane_phi(0, d))[at] -= b;
  c.p(plane_phi(1, d))[at] -= b;
}

}  // namespace

// Recover cell B of every cell of level `level` from its face fluxes
// (defined with the step kernels).
// B0 normal component on each face (flux over area, nonorthogonal_sweep.c:
// 470-474) and B0 edge components along the edge tangents (:719-725), from
// background planes that were uploaded rather than computed here.
__global__ void k_background_derived(Ctx c) {
  int f, i, j, k;
  const Layout &L = c.L;
  if (!box_index(blockIdx.x * (long)blockDim.x + threadIdx.x, L.NI + 1, L.NJ + 1,
                 L.NK + 1, f, i, j, k))
    return;
  const long at = L.at(f, i, j, k);
  const int ext[3] = {L.NI, L.NJ, L.NK};
  const int pos[3] = {i, j, k};
  for (int d = 0; d < 3; ++d) {
    bool face = true, edge = true;
    for (int a = 0; a < 3; ++a) {
      if (pos[a] > ext[a] - (a == d ? 0 : 1)) face = false;
      if (pos[a] > ext[a] - (a == d ? 1 : 0)) edge = false;
    }
    if (face)
      c.p(P_B0N + d)[at] = c.p(P_B0FLUX + d)[at] / c.p(P_FACE + FACE_PLANES * d + FO_AREA)[at];
    if (edge) {
      const int eb = P_EDGE + EDGE_PLANES * d;
      double r[3], t1[3], t2[3];
      for (int e = 0; e < 3; ++e) {
        r[e] = c.p(P_B0E + 3 * d + e)[at];
        t1[e] = c.p(eb + EO_T1 + e)[at];
        t2[e] = c.p(eb + EO_T2 + e)[at];
      }
      c.p(P_B0ET + 2 * d)[at] = dot3(r, t1);
      c.p(P_B0ET + 2 * d + 1)[at] = dot3(r, t2);
    }
  }
}

int launch_recover_all(Problem &P, int phi_first, int b_first);

static unsigned blocks_for(long n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

// gamera_no_radial_map_forward (nonorthogonal/src/nonorthogonal_radial_map.c
// :147-211) for the two map versions a config can name, with the v4
// stretch solved by the same bisection (:5-31, :58-81).
static int solve_stretch(double slope, double *stretch) {
  if (!std::isfinite(slope) || !(slope > 0.0) || !(slope < 1.0)) return -1;
  double lo = 0.0, hi = 1.0;
  while (hi / expm1(hi) > slope) {
    hi *= 2.0;
    if (!(hi < 128.0)) return -1;
  }
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid / expm1(mid) > slope) lo = mid;
    else hi = mid;
  }
  *stretch = 0.5 * (lo + hi);
  return std::isfinite(*stretch) && *stretch > 0.0 ? 0 : -1;
}

static int radial_radius(const gcs_problem_config &cfg, double fraction,
                         double *r) {
  const double rin = cfg.inner_radius, rout = cfg.outer_radius;
  if (cfg.radial_map_version == GAMERA_NO_RADIAL_MAP_LEGACY_EXPONENTIAL) {
    *r = rin + (rout - rin) * expm1(cfg.radial_stretch * fraction) /
                   expm1(cfg.radial_stretch);
  } else if (cfg.radial_map_version == GAMERA_NO_RADIAL_MAP_EARTH_PRODUCTION_V4) {
    const double knot = 23.0 / 72.0, pknot = 14.0;
    const double inner_derivative = (pknot - rin) / knot;
    double stretch;
    if (solve_stretch(inner_derivative * (1.0 - knot) / (rout - pknot),
                      &stretch) != 0)
      return -1;
    if (fraction < 0.0) {
      *r = rin * exp(inner_derivative / rin * fraction);
    } else if (fraction <= knot) {
      *r = rin + (pknot - rin) * fraction / knot;
    } else {
      const double x = (fraction - knot) / (1.0 - knot);
      *r = pknot + (rout - pknot) * expm1(stretch * x) / expm1(stretch);
    }
  } else {
    return -1;
  }
  return std::isfinite(*r) && *r > 0.0 ? 0 : -1;
}

int setup_problem(Problem &P) {
  const gcs_problem_config &cfg = P.cfg;
  const Layout &L = P.L;
  const double qp = 0.25 * acos(-1.0);
  const double da = 2.0 * qp / (double)L.n;
  if (qp + (double)L.ghost * da >= 2.0 * qp) {
    set_error("angular grid too coarse for the ghost layers");
    return -1;
  }
  std::vector<double> radius(L.NI + 1), ta(L.NJ + 1), tb(L.NK + 1);
  for (int i = 0; i <= L.NI; ++i) {
    const double logical = (double)i - (double)L.ghost;
    if (radial_radius(cfg, logical / (double)L.nr, &radius[i]) != 0) {
      set_error("radial map failed");
      return -1;
    }
  }
  for (int j = 0; j <= L.NJ; ++j) ta[j] = tan(-qp + ((double)j - (double)L.ghost) * da);
  for (int k = 0; k <= L.NK; ++k) tb[k] = tan(-qp + ((double)k - (double)L.ghost) * da);
  double *tables = nullptr;
  const size_t nt = radius.size() + ta.size() + tb.size();
  GCS_HIP(hipMalloc(&tables, nt * sizeof(double)));
  GCS_HIP(hipMemcpy(tables, radius.data(), radius.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  GCS_HIP(hipMemcpy(tables + radius.size(), ta.data(), ta.size() * sizeof(double),
                    hipMemcpyHostToDevice));
  GCS_HIP(hipMemcpy(tables + radius.size() + ta.size(), tb.data(),
                    tb.size() * sizeof(double), hipMemcpyHostToDevice));

  Ctx c{L, P.planes, P.bytes, P.scalars};
  const int T = 256;
  const long nv = 6L * (L.NI + 1) * (L.NJ + 1) * (L.NK + 1);
  const long nc = 6L * L.NI * L.NJ * L.NK;
  auto nface = [&](int d) {
    return 6L * (L.NI + (d == 0)) * (L.NJ + (d == 1)) * (L.NK + (d == 2));
  };
  auto nedge = [&](int d) {
    return 6L * (L.NI + (d != 0)) * (L.NJ + (d != 1)) * (L.NK + (d != 2));
  };
  hipStream_t s = P.stream;
  k_vertices<<<blocks_for(nv, T), T, 0, s>>>(c, tables, tables + radius.size(),
                                             tables + radius.size() + ta.size());
  for (int d = 0; d < 3; ++d)
    k_face_geometry<<<blocks_for(nface(d), 64), 64, 0, s>>>(c, d);
  k_cell_geometry<<<blocks_for(nc, 64), 64, 0, s>>>(c);
  k_recovery_constants<<<blocks_for(nc, 64), 64, 0, s>>>(c);
  for (int d = 0; d < 3; ++d)
    k_edge_geometry<<<blocks_for(nedge(d), T), T, 0, s>>>(c, d);
  for (int d = 0; d < 3; ++d)
    k_edge_solve<<<blocks_for(nedge(d), T), T, 0, s>>>(c, d);
  const gamera_cs_earth_startup &su = cfg.startup;
  Startup st{su.inner_radius, su.density,          su.speed,
             su.mach,         su.gamma,            su.ambient_density,
             su.ambient_pressure, su.front_x,      su.front_width,
             su.cut_radius,   su.cut_length,
             {su.imf.value[0], su.imf.value[1], su.imf.value[2]},
             {su.dipole.moment.value[0], su.dipole.moment.value[1],
              su.dipole.moment.value[2]}};
  for (int d = 0; d < 3; ++d)
    k_background_face<<<blocks_for(nface(d), 64), 64, 0, s>>>(c, d, st);
  k_background_cell<<<blocks_for(nc, 64), 64, 0, s>>>(c, st);
  for (int d = 0; d < 3; ++d)
    k_background_edge<<<blocks_for(nedge(d), T), T, 0, s>>>(c, d, st);
  k_initial_cells<<<blocks_for(nc, T), T, 0, s>>>(c, st, cfg.perturbation, cfg.seed);
  for (int d = 0; d < 3; ++d)
    k_line_integrals<<<blocks_for(nedge(d), T), T, 0, s>>>(c, d, st);
  for (int d = 0; d < 3; ++d)
    k_initial_faces<<<blocks_for(nface(d), T), T, 0, s>>>(c, d);
  if (launch_recover_all(P, plane_phi(0, 0), plane_b(0, 0)) != 0) return -1;
  // Old level := current level (nonorthogonal_step.c:361-377).
  for (int m = 0; m < 5; ++m)
    GCS_HIP(hipMemcpyAsync(P.plane(plane_u(1, m)), P.plane(plane_u(0, m)),
                           L.total * sizeof(double), hipMemcpyDeviceToDevice, s));
  for (int m = 0; m < 3; ++m) {
    GCS_HIP(hipMemcpyAsync(P.plane(plane_b(1, m)), P.plane(plane_b(0, m)),
                           L.total * sizeof(double), hipMemcpyDeviceToDevice, s));
    GCS_HIP(hipMemcpyAsync(P.plane(plane_phi(1, m)), P.plane(plane_phi(0, m)),
                           L.total * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  for (int d = 0; d < 3; ++d)
    k_subtract_background<<<blocks_for(nface(d), T), T, 0, s>>>(c, d);
  if (launch_recover_all(P, plane_phi(0, 0), plane_b(0, 0)) != 0 ||
      launch_recover_all(P, plane_phi(1, 0), plane_b(1, 0)) != 0)
    return -1;
  // The flux planes served as scratch; leave them zero like fresh storage.
  GCS_HIP(hipMemsetAsync(P.plane(P_FLUID), 0, 24 * L.total * sizeof(double), s));
  GCS_HIP(hipStreamSynchronize(s));
  GCS_HIP(hipFree(tables));
  StepScalars sc;
  GCS_HIP(hipMemcpy(&sc, P.scalars, sizeof(sc), hipMemcpyDeviceToHost));
  if (sc.failed) {
    set_error("setup kernel failed at stage " + std::to_string(sc.fail_stage));
    return -1;
  }
  P.level = 0;
  P.host_dirty = true;
  return 0;
}

int derive_constants(Problem &P) {
  const Layout &L = P.L;
  Ctx c{L, P.planes, P.bytes, P.scalars};
  hipStream_t s = P.stream;
  const long nc = 6L * L.NI * L.NJ * L.NK;
  const long nv = 6L * (L.NI + 1) * (L.NJ + 1) * (L.NK + 1);
  k_recovery_constants<<<blocks_for(nc, 64), 64, 0, s>>>(c);
  for (int d = 0; d < 3; ++d) {
    const long ne = 6L * (L.NI + (d != 0)) * (L.NJ + (d != 1)) * (L.NK + (d != 2));
    k_edge_solve<<<blocks_for(ne, 256), 256, 0, s>>>(c, d);
  }
  k_background_derived<<<blocks_for(nv, 256), 256, 0, s>>>(c);
  GCS_HIP(hipGetLastError());
  GCS_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace gcs
