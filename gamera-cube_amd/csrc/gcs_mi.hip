// Ionosphere (MI) coupling on the device: field-aligned current on the sample
// shell, its samples on the two polar grids, the potential solve, and the
// view of the potential that the step kernels read
// (cubedsphere/cubedsphere_mi_dipole.c:157-436, NO/nonorthogonal_mi.c).
//
// Geometry-only parts (donor faces and bilinear weights of the shell fringe
// and of every polar sample point, the sample-shell radius) are computed once
// on the host with the C library, exactly as the reference computes them on
// every update. The per-update work is five kernels plus the solve.
//
// Solvers. The Pedersen-only potential uses the reference's algorithm:
// Jacobi-preconditioned CG on the reduced sin(theta)-weighted system with
// true-residual checks (NO/nonorthogonal_mi.c:412-609). With a Hall
// conductance the operator is the reference's five-point tensor stencil
// solved by right-preconditioned BiCGStab (:959-1151); the reference's
// preconditioner is a symmetric Gauss-Seidel sweep, which is sequential over
// all unknowns, so here each colatitude ring is solved exactly instead (a
// periodic tridiagonal solve of the ring's longitude coupling, one thread per
// ring). Dot products are two-pass block reductions in a fixed order; the
// iterates therefore differ from the reference's by rounding, and the
// converged potential agrees with it to the solve's tolerance. Both
// hemispheres run in the same launches (blockIdx.y).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "gcs_device.hpp"
#include "gcs_math.hpp"
#include "gcs_unet.hpp"

namespace gcs {

namespace {

constexpr int kT = 256;

// value = (1-wx) [(1-wy) v[b] + wy v[b+1]] + wx [(1-wy) v[b+s] + wy v[b+s+1]]
// (cubedsphere_mi_dipole.c:326-331, :361-366)
struct Bilinear {
  int target, base, step, zero;
  double wx, wy;
};

enum Slot {  // solver scalars per hemisphere
  SL_RHS2, SL_FIRST, SL_TARGET, SL_RZ, SL_PAP, SL_NORM, SL_RHO, SL_DEN, SL_WNUM,
  SL_WDEN, SL_ALPHA, SL_OMEGA, SL_RHOPREV, SL_POLE, SL_COUNT = 16
};

}  // namespace

struct MiState {
  gamera_cs_mi_dipole_config cfg;
  double hall = 0.0;
  int np = 0, nt = 0, n = 0, S = 0, si = 0;
  double sample_radius = 0.0, sample_max_colat = 0.0, max_colat = 0.0;
  double b_norm = 0.0, current_norm = 0.0, potential_norm = 0.0;
  double next_update = 0.0, epoch = 0.0;
  bool ready = false;
  int updates = 0;
  int iterations[2] = {0, 0};
  double final_residual[2] = {0.0, 0.0};
  double solve_ms = 0.0;
  // device arrays
  double *shell_fac = nullptr;  // [6][S][S]
  double *fac = nullptr;        // [2][nt][np]
  double *potential = nullptr;  // [2][nt][np]
  double *work = nullptr;       // [2][kVectors][n] + ring scratch
  double *scal = nullptr;       // [2][SL_COUNT]
  double *partial = nullptr;    // [2][kMaxBlocks]
  double *host_scal = nullptr;  // pinned mirror of scal
  Bilinear *bfringe = nullptr, *ffringe = nullptr, *samples = nullptr;
  int n_bfringe = 0, n_ffringe = 0, n_samples = 0;
  int *active = nullptr;  // per-hemisphere "still iterating" flags
  // learned warm start (not owned): the initial guess becomes the previous
  // potential plus the network's answer for the current residual
  unet::Net *net = nullptr;
  double *warm_scale = nullptr;  // [2] RMS of the residual's current equivalent
};

namespace {

constexpr int kVectors = 18;   // solver vectors per hemisphere
constexpr int kMaxBlocks = 512;

struct MiK {
  Layout L;
  double *pl;
  uint8_t *by;
  StepScalars *sc;
  int cur;
  int np, nt, n, S, si;
  double ri, max_colat, current_norm, pedersen, r2;
  double *shell_fac, *fac, *potential, *work, *scal, *partial;
  __device__ double *p(int q) const { return pl + (long)q * L.total; }
  __device__ const uint8_t *b(int q) const { return by + (long)q * L.total; }
  __device__ double *vec(int h, int v) const { return work + ((long)h * kVectors + v) * n; }
  __device__ double *sl(int h) const { return scal + h * SL_COUNT; }
  __device__ void fail(unsigned stage, long index) const {
    if (atomicOr(&sc->failed, 1u) == 0u) {
      sc->fail_stage = stage;
      sc->fail_index = (int)index;
    }
  }
};

// vectors of one hemisphere's solve
enum Vec { V_RHS, V_DIAG, V_X, V_R, V_Z, V_P, V_AP, V_SH, V_PP, V_S, V_PS, V_AS,
           V_LO, V_HI, V_LLO, V_LHI, V_SCR, V_SCR2 };

GCS_DEV long gtid() { return blockIdx.x * (long)blockDim.x + threadIdx.x; }

// ---- shell FAC (cubedsphere_mi_dipole.c:223-342) ---------------------------------

// Current-level B of shells si-1..si+1 into the scratch planes.
__global__ void k_mi_copy_shell(MiK c) {
  const Layout &L = c.L;
  const long per = 3L * L.NJ * L.NK;
  const long t = gtid();
  if (t >= 6 * per) return;
  const int f = (int)(t / per);
  long r = t - f * per;
  const int di = (int)(r / ((long)L.NJ * L.NK));
  r -= (long)di * L.NJ * L.NK;
  const int j = (int)(r / L.NK), k = (int)(r - (long)j * L.NK);
  const long at = L.at(f, c.si - 1 + di, j, k);
  for (int d = 0; d < 3; ++d) c.p(P_FVEL + d)[at] = c.p(plane_b(c.cur, d))[at];
}

// One-cell B fringe from the owning face (:236-276).
__global__ void k_mi_fringe_b(MiK c, const Bilinear *e, int n) {
  const long t = gtid();
  if (t >= n) return;
  const Bilinear h = e[t];
  const double w[4] = {(1.0 - h.wx) * (1.0 - h.wy), (1.0 - h.wx) * h.wy,
                       h.wx * (1.0 - h.wy), h.wx * h.wy};
  const long src[4] = {h.base, h.base + 1, h.base + h.step, h.base + h.step + 1};
  for (int d = 0; d < 3; ++d) {
    double v = 0.0;
    for (int q = 0; q < 4; ++q) v += w[q] * c.p(plane_b(c.cur, d))[src[q]];
    c.p(P_FVEL + d)[h.target] = v;
  }
}

// NO/nonorthogonal_mi.c:43-95: 4-cell average of B along one edge.
GCS_DEV bool edge_line(const MiK &c, int dir, long at, double &out) {
  const Layout &L = c.L;
  if (!c.b(Q_VALID + dir)[at]) return false;
  const long step[3] = {L.di(), (long)L.SK, 1};
  const int t1 = dir == 0 ? 1 : 0, t2 = dir == 2 ? 1 : 2;
  const long cell[4] = {at, at - step[t1], at - step[t2], at - step[t1] - step[t2]};
  double avg[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < 4; ++q)
    for (int m = 0; m < 3; ++m) avg[m] += 0.25 * c.p(P_FVEL + m)[cell[q]];
  const int eb = P_EDGE + EDGE_PLANES * dir;
  double dot = 0.0;
  for (int m = 0; m < 3; ++m) dot += avg[m] * c.p(eb + EO_N + m)[at];
  out = dot * c.p(eb + EO_LEN)[at];
  return isfinite(out);
}

// :97-136: circulation of B around the face of direction dir at `at`.
GCS_DEV bool face_circulation(const MiK &c, int dir, long at, double &out) {
  const Layout &L = c.L;
  const long si = L.di(), sj = L.SK;
  double e1, e2, e3, e4;
  bool ok;
  if (dir == 0)
    ok = edge_line(c, 1, at, e1) && edge_line(c, 2, at + sj, e2) &&
         edge_line(c, 1, at + 1, e3) && edge_line(c, 2, at, e4);
  else if (dir == 1)
    ok = edge_line(c, 2, at, e1) && edge_line(c, 0, at + 1, e2) &&
         edge_line(c, 2, at + si, e3) && edge_line(c, 0, at, e4);
  else
    ok = edge_line(c, 0, at, e1) && edge_line(c, 1, at + si, e2) &&
         edge_line(c, 0, at + sj, e3) && edge_line(c, 1, at, e4);
  if (!ok) return false;
  out = e1 + e2 - e3 - e4;
  return isfinite(out);
}

// FAC of the active shell cells (:277-307).
__global__ void k_mi_shell_fac(MiK c) {
  const Layout &L = c.L;
  const int n = L.n;
  const long t = gtid();
  if (t >= 6L * n * n) return;
  const int f = (int)(t / ((long)n * n));
  const int r = (int)(t - (long)f * n * n);
  const int lj = 1 + r / n, lk = 1 + r % n;
  const int j = L.ghost - 1 + lj, k = L.ghost - 1 + lk;
  const long at = L.at(f, c.si, j, k);
  const long step[3] = {L.di(), (long)L.SK, 1};
  double flux[6];
  for (int d = 0; d < 3; ++d)
    if (!face_circulation(c, d, at, flux[2 * d]) ||
        !face_circulation(c, d, at + step[d], flux[2 * d + 1])) {
      c.fail(91, at);
      return;
    }
  double fc[6][3], cen[3], rec[9], cur[3];
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int m = 0; m < 3; ++m)
        fc[2 * d + s][m] = c.p(P_FACE + FACE_PLANES * d + FO_CEN + m)[at + s * step[d]];
  for (int m = 0; m < 3; ++m) cen[m] = c.p(P_CEN + m)[at];
  for (int q = 0; q < 9; ++q) rec[q] = c.p(P_REC + q)[at];
  const int piv = c.b(Q_PIV)[at];
  if (piv == 0xff || !recover_cell(flux, fc, cen, c.p(P_VOL)[at], rec, piv, cur)) {
    c.fail(92, at);
    return;
  }
  double b0[3], x[3];
  for (int m = 0; m < 3; ++m) {
    b0[m] = c.p(P_B0C + m)[at];
    x[m] = cen[m];
  }
  const double rad = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double bm = sqrt(b0[0] * b0[0] + b0[1] * b0[1] + b0[2] * b0[2]);
  if (!(rad > 0.0) || !(bm > 0.0)) {
    c.fail(93, at);
    return;
  }
  const double colat = acos(fmin(1.0, fabs(x[2]) / rad));
  // NO/nonorthogonal_mi.c:183-196
  double s = sin(colat) * sqrt(c.ri / rad);
  s = fmax(-1.0, fmin(1.0, s));
  const double ti = asin(s);
  const double q = rad / c.ri;
  const double ratio = q * q * q * sqrt(1.0 + 3.0 * pow(cos(ti), 2.0)) /
                       sqrt(1.0 + 3.0 * pow(cos(colat), 2.0));
  const double jb = cur[0] * b0[0] + cur[1] * b0[1] + cur[2] * b0[2];
  const double fac = jb / bm * ratio * c.current_norm;
  if (!isfinite(fac)) {
    c.fail(94, at);
    return;
  }
  c.shell_fac[(long)f * c.S * c.S + lj * c.S + lk] = fac;
}

// FAC fringe (:308-342) and polar samples (:344-370, :391-406).
__global__ void k_mi_bilinear(const Bilinear *e, int n, const double *src, double *dst) {
  const long t = gtid();
  if (t >= n) return;
  const Bilinear h = e[t];
  if (h.zero) {
    dst[h.target] = 0.0;
    return;
  }
  const double *v = src + h.base;
  const double lo = (1.0 - h.wy) * v[0] + h.wy * v[1];
  const double hi = (1.0 - h.wy) * v[h.step] + h.wy * v[h.step + 1];
  dst[h.target] = (1.0 - h.wx) * lo + h.wx * hi;
}

// ---- polar solver (NO/nonorthogonal_mi.c) -----------------------------------

GCS_DEV int red_at(int np, int i, int j) { return 1 + (i - 1) * np + j; }

GCS_DEV double incl2(double th) {
  const double cs = cos(th);
  return 4.0 * cs * cs / (1.0 + 3.0 * cs * cs);
}
GCS_DEV double cos_incl(double th, int hemisphere) {
  const double cs = cos(th);
  return (double)hemisphere * 2.0 * cs / sqrt(1.0 + 3.0 * cs * cs);
}
GCS_DEV int hemi_sign(int h) { return h == 0 ? GAMERA_MI_NORTH : GAMERA_MI_SOUTH; }

// :234-241
GCS_DEV double theta_coeff(double sigma, double half, double center, double dt) {
  return sin(half) * sigma / (incl2(half) * sin(center) * dt * dt);
}

// One interior row (colatitude index i) of the five-point tensor stencil
// (:777-833) for uniform conductances: the averages of neighbouring Pedersen
// values are the value itself; the Hall term enters through
// d(-SigmaH / cos I)/dtheta. y = diag x + lo x(i-1) + hi x(i+1)
// + llo x(j-1) + lhi x(j+1).
struct TensorRow {
  double diag, lo, hi, llo, lhi;
};
GCS_DEV TensorRow tensor_row(double sp, double hall, int i, double dt, double dp, int hemi) {
  const double th = (double)i * dt;
  const double w = sin(th);
  const double inc = cos_incl(th, hemi);
  const double lt = th - 0.5 * dt, ut = th + 0.5 * dt;
  const double lo = sin(lt) * 0.5 * (sp + sp) / (incl2(lt) * dt * dt);
  const double hi = sin(ut) * 0.5 * (sp + sp) / (incl2(ut) * dt * dt);
  const double llo = 0.5 * (sp + sp) / (w * dp * dp);
  const double lhi = 0.5 * (sp + sp) / (w * dp * dp);
  const double f_up = -hall / cos_incl((double)(i + 1) * dt, hemi);
  const double f_dn = -hall / cos_incl((double)(i - 1) * dt, hemi);
  const double dft = (f_up - f_dn) / (2.0 * dt);
  const double dfp = (-hall / inc - -hall / inc) / (2.0 * dp);
  TensorRow t;
  t.diag = lo + hi + llo + lhi;
  t.lo = -lo - dfp / (2.0 * dt);
  t.hi = -hi + dfp / (2.0 * dt);
  t.llo = -llo + dft / (2.0 * dp);
  t.lhi = -lhi - dft / (2.0 * dp);
  return t;
}

// Right-hand side (:428-448, weighted by sin theta), initial guess from the
// previous potential (:449-464), and the diagonal (Pedersen, :322-344) or the
// five-point tensor stencil (:777-833).
__global__ void k_mi_setup(MiK c, int tensor, double hall) {
  const int h = blockIdx.y;
  const long q = gtid();
  if (q >= c.n) return;
  const int np = c.np;
  const double dt = c.max_colat / (double)(c.nt - 1);
  const double dp = 2.0 * kPi / (double)np;
  double *rhs = c.vec(h, V_RHS), *x = c.vec(h, V_X), *diag = c.vec(h, V_DIAG);
  const double *pot = c.potential + (long)h * c.nt * np;
  const double *fac = c.fac + (long)h * c.nt * np;
  if (q == 0) {
    rhs[0] = 0.0;
    double pole = 0.0;
    for (int j = 0; j < np; ++j) {
      const double v = pot[j];
      if (isfinite(v)) pole += v / (double)np;
    }
    x[0] = pole;
    const double half = 0.5 * dt;
    if (!tensor) {
      diag[0] = (double)np * (sin(half) * c.pedersen / (incl2(half) * dt * dt));
    } else {
      // uniform Pedersen: every pole edge has the same coefficient
      double d0 = 0.0;
      for (int j = 0; j < np; ++j) d0 += sin(half) * c.pedersen / (incl2(half) * dt * dt);
      diag[0] = d0;
    }
    return;
  }
  const int i = 1 + (int)((q - 1) / np), j = (int)((q - 1) % np);
  const double th = (double)i * dt;
  const double w = sin(th);
  const double inc = cos_incl(th, hemi_sign(h));
  const double u = -c.r2 * fac[i * np + j] * inc;
  rhs[q] = w * u;
  const double v = pot[i * np + j];
  x[q] = isfinite(v) ? v : 0.0;
  if (!tensor) {
    const double lo = theta_coeff(c.pedersen, th - 0.5 * dt, th, dt);
    const double hi = theta_coeff(c.pedersen, th + 0.5 * dt, th, dt);
    const double lon = c.pedersen / (sin(th) * sin(th) * dp * dp);
    diag[q] = w * (lo + hi + 2.0 * lon);
    return;
  }
  const TensorRow t = tensor_row(c.pedersen, hall, i, dt, dp, hemi_sign(h));
  diag[q] = t.diag;
  c.vec(h, V_LO)[q] = t.lo;
  c.vec(h, V_HI)[q] = t.hi;
  c.vec(h, V_LLO)[q] = t.llo;
  c.vec(h, V_LHI)[q] = t.lhi;
}

// y = A x on the reduced unknowns: the sin(theta)-weighted Pedersen operator
// (:355-392) or the tensor stencil (:871-900).
__global__ void k_mi_apply(MiK c, int tensor, int xin, int yout, const int *active) {
  const int h = blockIdx.y;
  if (!active[h]) return;
  const long q = gtid();
  if (q >= c.n) return;
  const int np = c.np, nt = c.nt;
  const double *x = c.vec(h, xin);
  double *y = c.vec(h, yout);
  const double *diag = c.vec(h, V_DIAG);
  const double dt = c.max_colat / (double)(nt - 1);
  if (q == 0) {
    if (!tensor) {
      const double half = 0.5 * dt;
      const double pe = sin(half) * c.pedersen / (incl2(half) * dt * dt);
      double ring = 0.0;
      for (int j = 0; j < np; ++j) ring += x[red_at(np, 1, j)];
      y[0] = pe * ((double)np * x[0] - ring);
    } else {
      const double half = 0.5 * dt;
      const double pe = sin(half) * c.pedersen / (incl2(half) * dt * dt);
      double v = diag[0] * x[0];
      for (int j = 0; j < np; ++j) v -= pe * x[red_at(np, 1, j)];
      y[0] = v;
    }
    return;
  }
  const int i = 1 + (int)((q - 1) / np), j = (int)((q - 1) % np);
  const int jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
  const double below = i == 1 ? x[0] : x[red_at(np, i - 1, j)];
  const double above = i + 2 < nt ? x[red_at(np, i + 1, j)] : 0.0;
  if (!tensor) {
    const double dp = 2.0 * kPi / (double)np;
    const double th = (double)i * dt;
    const double w = sin(th);
    const double lo = theta_coeff(c.pedersen, th - 0.5 * dt, th, dt);
    const double hi = theta_coeff(c.pedersen, th + 0.5 * dt, th, dt);
    const double lon = c.pedersen / (sin(th) * sin(th) * dp * dp);
    y[q] = w * (lo * (x[q] - below) + hi * (x[q] - above) +
                lon * (2.0 * x[q] - x[red_at(np, i, jm)] - x[red_at(np, i, jp)]));
  } else {
    y[q] = diag[q] * x[q] + c.vec(h, V_LO)[q] * below + c.vec(h, V_HI)[q] * above +
           c.vec(h, V_LLO)[q] * x[red_at(np, i, jm)] +
           c.vec(h, V_LHI)[q] * x[red_at(np, i, jp)];
  }
}

// Two-pass dot products: per-block partial sums in a fixed order, then one
// block sums the partials in order. mode 0: a.b; mode 1: the residual norm of
// the unweighted equations (:394-410), sqrt taken in the finish.
__global__ void k_mi_dot(MiK c, int a, int b, int mode, const int *active) {
  const int h = blockIdx.y;
  if (!active[h]) return;
  __shared__ double red[kT];
  const int np = c.np;
  const double dt = c.max_colat / (double)(c.nt - 1);
  double s = 0.0;
  for (long q = blockIdx.x * (long)kT + threadIdx.x; q < c.n; q += (long)gridDim.x * kT) {
    if (mode == 0) {
      s += c.vec(h, a)[q] * c.vec(h, b)[q];
    } else if (q == 0) {
      const double half = 0.5 * dt;
      const double ps = (double)np * (sin(half) * c.pedersen / (incl2(half) * dt * dt));
      const double pr = c.vec(h, a)[0] / ps;
      s += (double)np * pr * pr;
    } else {
      const int i = 1 + (int)((q - 1) / np);
      const double v = c.vec(h, a)[q] / sin((double)i * dt);
      s += v * v;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) c.partial[h * kMaxBlocks + blockIdx.x] = red[0];
}

__global__ void k_mi_finish(MiK c, int nblocks, int slot, int mode, const int *active) {
  const int h = blockIdx.x;
  if (!active[h] || threadIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += c.partial[h * kMaxBlocks + b];
  c.sl(h)[slot] = mode == 1 ? sqrt(fmax(0.0, s)) : s;
}

// Elementwise steps of the Krylov loops; scalars come from device slots.
enum Op {
  OP_RESIDUAL,      // r = rhs - ap; shadow = r
  OP_JACOBI,        // z = r / diag
  OP_COPY_ZP,       // p = z
  OP_CG_XR,         // alpha = rz / pap; x += alpha p; r -= alpha ap
  OP_CG_P,          // p = z + (rz_new / rz) p   (rz_new in SL_NORM slot holder)
  OP_BI_P,          // p = r + beta (p - omega ap)
  OP_BI_S,          // alpha = rho / den; s = r - alpha ap
  OP_BI_XR,         // omega = wnum / wden; x += alpha pp + omega ps; r = s - omega as
  OP_BI_X_ONLY,     // x += alpha pp
  OP_ZERO_PAP,      // p = 0, ap = 0
  OP_EXPAND         // potential from x
};

__global__ void k_mi_vec(MiK c, int op, double extra, const int *active) {
  const int h = blockIdx.y;
  if (!active[h]) return;
  const long q = gtid();
  const double *sl = c.sl(h);
  if (op == OP_EXPAND) {
    const long np = c.np, nt = c.nt;
    if (q >= np * nt) return;
    const int i = (int)(q / np), j = (int)(q % np);
    const double *x = c.vec(h, V_X);
    double *pot = c.potential + (long)h * nt * np;
    pot[q] = i == 0 ? x[0] : i == nt - 1 ? extra : x[red_at((int)np, i, j)];
    return;
  }
  if (q >= c.n) return;
  double *x = c.vec(h, V_X), *r = c.vec(h, V_R), *z = c.vec(h, V_Z), *p = c.vec(h, V_P),
         *ap = c.vec(h, V_AP);
  switch (op) {
    case OP_RESIDUAL:
      r[q] = c.vec(h, V_RHS)[q] - ap[q];
      c.vec(h, V_SH)[q] = r[q];
      break;
    case OP_JACOBI: z[q] = r[q] / c.vec(h, V_DIAG)[q]; break;
    case OP_COPY_ZP: p[q] = z[q]; break;
    case OP_CG_XR: {
      const double alpha = sl[SL_RZ] / sl[SL_PAP];
      x[q] += alpha * p[q];
      r[q] -= alpha * ap[q];
      break;
    }
    case OP_CG_P: p[q] = z[q] + (sl[SL_RHO] / sl[SL_RZ]) * p[q]; break;
    case OP_BI_P: {
      const double beta = (sl[SL_RHO] / sl[SL_RHOPREV]) * (sl[SL_ALPHA] / sl[SL_OMEGA]);
      p[q] = r[q] + beta * (p[q] - sl[SL_OMEGA] * ap[q]);
      break;
    }
    case OP_BI_S: c.vec(h, V_S)[q] = r[q] - sl[SL_ALPHA] * ap[q]; break;
    case OP_BI_XR: {
      const double omega = sl[SL_OMEGA];
      x[q] += sl[SL_ALPHA] * c.vec(h, V_PP)[q] + omega * c.vec(h, V_PS)[q];
      r[q] = c.vec(h, V_S)[q] - omega * c.vec(h, V_AS)[q];
      break;
    }
    case OP_BI_X_ONLY: x[q] += sl[SL_ALPHA] * c.vec(h, V_PP)[q]; break;
    case OP_ZERO_PAP:
      p[q] = 0.0;
      ap[q] = 0.0;
      break;
    default: break;
  }
}

// Scalar updates of BiCGStab that must follow a reduction.
__global__ void k_mi_scalar(MiK c, int what, const int *active) {
  const int h = threadIdx.x;
  if (h >= 2 || !active[h]) return;
  double *sl = c.sl(h);
  if (what == 0) sl[SL_ALPHA] = sl[SL_RHO] / sl[SL_DEN];
  else if (what == 1) sl[SL_OMEGA] = sl[SL_WNUM] / sl[SL_WDEN];
  else if (what == 2) sl[SL_RHOPREV] = sl[SL_RHO];
  else if (what == 3) {  // restart
    sl[SL_RHOPREV] = 1.0;
    sl[SL_ALPHA] = 1.0;
    sl[SL_OMEGA] = 1.0;
  } else if (what == 4) {
    sl[SL_RZ] = sl[SL_RHO];  // CG: rz <- rz_new
  }
}

// Longitude-line preconditioner of the tensor system: out = M^-1 in with M
// the pole diagonal plus, on each colatitude ring, the ring's own periodic
// tridiagonal coupling (cyclic Thomas solve with a Sherman-Morrison
// correction). One thread per ring.
__global__ void k_mi_ring_solve(MiK c, int vin, int vout, const int *active) {
  const int h = blockIdx.y;
  if (!active[h]) return;
  const int ring = (int)gtid();
  const int np = c.np, rings = c.nt - 2;
  const double *in = c.vec(h, vin);
  double *out = c.vec(h, vout);
  const double *diag = c.vec(h, V_DIAG);
  if (ring == 0) out[0] = in[0] / diag[0];
  if (ring >= rings) return;
  const int base = 1 + ring * np;
  const double *a = c.vec(h, V_LLO) + base;  // coefficient of x[j-1]
  const double *b = diag + base;
  const double *cc = c.vec(h, V_LHI) + base; // coefficient of x[j+1]
  double *g = c.vec(h, V_SCR) + base;   // forward-sweep factors
  double *z = c.vec(h, V_SCR2) + base;  // correction vector
  double *x = out + base;
  const double *rhs = in + base;
  // Sherman-Morrison split of the periodic corners: gamma = -b[0]
  const double gam = -b[0];
  const double alpha = cc[np - 1];  // row np-1's coupling to x[0]
  const double beta = a[0];         // row 0's coupling to x[np-1]
  // T y = rhs into x[], T z = (gamma, 0, ..., alpha) into z[]; T has the
  // corner-corrected first and last diagonal entries
  double bet = b[0] - gam;
  x[0] = rhs[0] / bet;
  z[0] = gam / bet;
  g[0] = 0.0;
  for (int j = 1; j < np; ++j) {
    g[j] = cc[j - 1] / bet;
    const double bj = j == np - 1 ? b[j] - alpha * beta / gam : b[j];
    bet = bj - a[j] * g[j];
    x[j] = (rhs[j] - a[j] * x[j - 1]) / bet;
    z[j] = ((j == np - 1 ? alpha : 0.0) - a[j] * z[j - 1]) / bet;
  }
  for (int j = np - 2; j >= 0; --j) {
    x[j] -= g[j + 1] * x[j + 1];
    z[j] -= g[j + 1] * z[j + 1];
  }
  const double fact = (x[0] + beta * x[np - 1] / gam) / (1.0 + z[0] + beta * z[np - 1] / gam);
  for (int j = 0; j < np; ++j) x[j] -= fact * z[j];
}

template <typename T>
int upload_vec(const std::vector<T> &v, T **out) {
  *out = nullptr;
  if (v.empty()) return 0;
  GCS_HIP(hipMalloc(out, v.size() * sizeof(T)));
  GCS_HIP(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

unsigned blocks_of(long n, int t = kT) { return (unsigned)((n + t - 1) / t); }

MiK make_mik(Problem &P) {
  MiState &m = *P.mi;
  MiK c;
  c.L = P.L;
  c.pl = P.planes;
  c.by = P.bytes;
  c.sc = P.scalars;
  c.cur = P.level;
  c.np = m.np;
  c.nt = m.nt;
  c.n = m.n;
  c.S = m.S;
  c.si = m.si;
  c.ri = m.cfg.ionosphere_radius_re;
  c.max_colat = m.max_colat;
  c.current_norm = m.current_norm;
  c.pedersen = m.cfg.pedersen_siemens;
  const double rm = m.cfg.ionosphere_radius_re * m.cfg.length_norm_m;
  c.r2 = rm * rm;
  c.shell_fac = m.shell_fac;
  c.fac = m.fac;
  c.potential = m.potential;
  c.work = m.work;
  c.scal = m.scal;
  c.partial = m.partial;
  return c;
}

// ---- learned warm start ----------------------------------------------------------

// Current density equivalent to a residual row: the right-hand side of
// interior row (i, j) is -R^2 FAC cos(I) sin(theta) (:428-448), so a residual
// r there is what a current -r / (sin(theta) R^2 cos I) would add. The pole
// and outer rows carry none.
GCS_DEV double equivalent_current(double r, int i, int nt, double dt, double r2, int hemi) {
  if (i <= 0 || i >= nt - 1) return 0.0;
  const double th = (double)i * dt;
  return -(r / sin(th)) / (r2 * cos_incl(th, hemi));
}

// Network features of one point (16 floats): current / scale, SigmaP / 10,
// SigmaH / 10, colatitude index / (nt - 1), hemisphere sign, zeros.
GCS_DEV void write_features(float *f, double current, double scale, double sp, double sh,
                            int i, int nt, int hemi) {
  f[0] = scale > 0.0 ? (float)(current / scale) : 0.f;
  f[1] = (float)(sp / 10.0);
  f[2] = (float)(sh / 10.0);
  f[3] = (float)((double)i / (double)(nt - 1));
  f[4] = (float)hemi;
  for (int k = 5; k < unet::kInChannels; ++k) f[k] = 0.f;
}

// RMS of the residual's equivalent current over the interior rows, per
// hemisphere (one block each).
__global__ void k_warm_scale(MiK c, double *scale) {
  __shared__ double red[kT];
  const int h = blockIdx.y;
  const double dt = c.max_colat / (double)(c.nt - 1);
  const long count = (long)(c.nt - 2) * c.np;
  double sum = 0.0;
  for (long q = threadIdx.x; q < count; q += kT) {
    const int i = 1 + (int)(q / c.np), j = (int)(q % c.np);
    const double v = equivalent_current(c.vec(h, V_R)[red_at(c.np, i, j)], i, c.nt, dt, c.r2,
                                        hemi_sign(h));
    sum += v * v;
  }
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) scale[h] = sqrt(red[0] / (double)count);
}

__global__ void k_warm_features(MiK c, const double *scale, double hall, float *feat) {
  const int h = blockIdx.y;
  const long q = gtid();
  if (q >= (long)c.nt * c.np) return;
  const int i = (int)(q / c.np), j = (int)(q % c.np);
  const double dt = c.max_colat / (double)(c.nt - 1);
  const double r = (i > 0 && i < c.nt - 1) ? c.vec(h, V_R)[red_at(c.np, i, j)] : 0.0;
  write_features(feat + ((long)h * c.nt * c.np + q) * unet::kInChannels,
                 equivalent_current(r, i, c.nt, dt, c.r2, hemi_sign(h)), scale[h], c.pedersen,
                 hall, i, c.nt, hemi_sign(h));
}

// x += the network's potential correction (out * out_scale * R^2 * scale /
// SigmaP on the interior rows; the pole unknown takes the mean of row 0).
__global__ void k_warm_correct(MiK c, const double *scale, const float *out, double out_scale) {
  const int h = blockIdx.y;
  const long q = gtid();
  if (q >= c.n) return;
  const double s = scale[h];
  if (!(s > 0.0) || !isfinite(s)) return;
  const double k = out_scale * c.r2 * s / c.pedersen;
  const float *o = out + (long)h * c.nt * c.np;
  double *x = c.vec(h, V_X);
  if (q == 0) {
    double mean = 0.0;
    for (int j = 0; j < c.np; ++j) mean += (double)o[j];
    x[0] += k * (mean / (double)c.np);
    return;
  }
  const int i = 1 + (int)((q - 1) / c.np), j = (int)((q - 1) % c.np);
  x[q] += k * (double)o[(long)i * c.np + j];
}

// The guess in V_X (previous potential) becomes guess + N(residual): one
// operator product, the network's forward pass on both hemispheres, one
// update. The solve that follows recomputes the residual of the new guess.
int warm_start(Problem &P, const MiK &c, bool tensor) {
  MiState &m = *P.mi;
  unet::Net &n = *m.net;
  hipStream_t s = P.stream;
  const dim3 g2(blocks_of(m.n), 2);
  k_mi_apply<<<g2, kT, 0, s>>>(c, tensor ? 1 : 0, V_X, V_AP, m.active);
  k_mi_vec<<<g2, kT, 0, s>>>(c, OP_RESIDUAL, 0.0, m.active);
  k_warm_scale<<<dim3(1, 2), kT, 0, s>>>(c, m.warm_scale);
  k_warm_features<<<dim3(blocks_of((long)m.nt * m.np), 2), kT, 0, s>>>(c, m.warm_scale, m.hall,
                                                                       n.input);
  hipStream_t keep = n.stream;
  n.stream = s;
  const int rc = unet::forward_device(n, 2);
  n.stream = keep;
  if (rc != 0) return -1;
  k_warm_correct<<<g2, kT, 0, s>>>(c, m.warm_scale, n.out, n.out_scale);
  GCS_HIP(hipGetLastError());
  return 0;
}

// The polar solves of both hemispheres; 0 when both converge.
int solve(Problem &P, bool tensor) {
  MiState &m = *P.mi;
  MiK c = make_mik(P);
  hipStream_t s = P.stream;
  int *active = m.active;
  int act[2] = {1, 1};
  auto set_active = [&]() -> int {
    GCS_HIP(hipMemcpyAsync(active, act, sizeof(act), hipMemcpyHostToDevice, s));
    return 0;
  };
  const int nb = (int)std::min<long>(kMaxBlocks, blocks_of(m.n));
  const dim3 g2(blocks_of(m.n), 2);
  auto dot = [&](int a, int b, int mode, int slot) {
    k_mi_dot<<<dim3(nb, 2), kT, 0, s>>>(c, a, b, mode, active);
    k_mi_finish<<<2, 64, 0, s>>>(c, nb, slot, mode, active);
  };
  auto vec = [&](int op, double extra = 0.0) {
    const long count = op == OP_EXPAND ? (long)m.np * m.nt : m.n;
    k_mi_vec<<<dim3(blocks_of(count), 2), kT, 0, s>>>(c, op, extra, active);
  };
  auto apply = [&](int xin, int yout) {
    k_mi_apply<<<g2, kT, 0, s>>>(c, tensor ? 1 : 0, xin, yout, active);
  };
  auto precondition = [&](int vin, int vout) {
    if (tensor)
      k_mi_ring_solve<<<dim3(blocks_of(m.nt - 2, 64), 2), 64, 0, s>>>(c, vin, vout, active);
    else
      vec(OP_JACOBI);  // z = r / diag (the CG's only use)
  };
  auto read = [&]() -> int {
    GCS_HIP(hipMemcpyAsync(m.host_scal, m.scal, 2 * SL_COUNT * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    GCS_HIP(hipStreamSynchronize(s));
    return 0;
  };
  const auto t0 = std::chrono::steady_clock::now();
  if (set_active() != 0) return -1;
  k_mi_setup<<<g2, kT, 0, s>>>(c, tensor ? 1 : 0, m.hall);
  if (m.net != nullptr && warm_start(P, c, tensor) != 0) return -1;
  apply(V_X, V_AP);
  vec(OP_RESIDUAL);
  dot(V_R, V_R, 1, SL_FIRST);
  // right-hand side norm: sum of squares of the unweighted rows
  dot(V_RHS, V_RHS, 1, SL_RHS2);
  if (read() != 0) return -1;
  double target[2], last[2];
  int iters[2] = {0, 0};
  bool ok[2] = {false, false};
  for (int h = 0; h < 2; ++h) {
    const double *sl = m.host_scal + h * SL_COUNT;
    // SL_RHS2 holds sqrt(sum (rhs/w)^2) except for the pole row (zero rhs)
    const double rhs_norm = sl[SL_RHS2];
    target[h] = m.cfg.relative_tolerance * std::fmax(rhs_norm, 1.0);
    last[h] = sl[SL_FIRST];
    if (last[h] <= target[h]) {
      ok[h] = true;
      act[h] = 0;
    }
  }
  const int maxit = m.cfg.maximum_iterations;
  if (!tensor) {
    // :516-600
    if (set_active() != 0) return -1;
    precondition(V_R, V_Z);
    vec(OP_COPY_ZP);
    dot(V_R, V_Z, 0, SL_RZ);
    for (int it = 1; it <= maxit && (act[0] || act[1]); ++it) {
      apply(V_P, V_AP);
      dot(V_P, V_AP, 0, SL_PAP);
      if (read() != 0) return -1;
      for (int h = 0; h < 2; ++h) {
        if (!act[h]) continue;
        const double *sl = m.host_scal + h * SL_COUNT;
        if (!std::isfinite(sl[SL_RZ]) || sl[SL_RZ] <= DBL_MIN ||
            !std::isfinite(sl[SL_PAP]) || sl[SL_PAP] <= DBL_MIN)
          act[h] = 0;
      }
      if (set_active() != 0) return -1;
      vec(OP_CG_XR);
      dot(V_R, V_R, 1, SL_NORM);
      if (read() != 0) return -1;
      bool verify = false;
      for (int h = 0; h < 2; ++h)
        if (act[h]) {
          iters[h] = it;
          last[h] = m.host_scal[h * SL_COUNT + SL_NORM];
          if (last[h] <= target[h] || it % 1024 == 0) verify = true;
        }
      if (verify) {
        // true residual for every active hemisphere (restart when drifted)
        apply(V_X, V_AP);
        vec(OP_RESIDUAL);
        dot(V_R, V_R, 1, SL_NORM);
        if (read() != 0) return -1;
        for (int h = 0; h < 2; ++h)
          if (act[h]) {
            last[h] = m.host_scal[h * SL_COUNT + SL_NORM];
            if (last[h] <= target[h]) {
              ok[h] = true;
              act[h] = 0;
            }
          }
        if (set_active() != 0) return -1;
        precondition(V_R, V_Z);
        vec(OP_COPY_ZP);
        dot(V_R, V_Z, 0, SL_RZ);
        continue;
      }
      precondition(V_R, V_Z);
      dot(V_R, V_Z, 0, SL_RHO);  // rz_new
      vec(OP_CG_P);
      k_mi_scalar<<<1, 64, 0, s>>>(c, 4, active);
    }
  } else {
    // :1029-1131 with the ring preconditioner
    for (int h = 0; h < 2; ++h) {
      double *sl = m.host_scal + h * SL_COUNT;
      sl[SL_RHOPREV] = 1.0;
      sl[SL_ALPHA] = 1.0;
      sl[SL_OMEGA] = 1.0;
    }
    k_mi_scalar<<<1, 64, 0, s>>>(c, 3, active);
    vec(OP_ZERO_PAP);
    for (int it = 1; it <= maxit && (act[0] || act[1]); ++it) {
      dot(V_SH, V_R, 0, SL_RHO);
      if (read() != 0) return -1;
      for (int h = 0; h < 2; ++h) {
        if (!act[h]) continue;
        const double *sl = m.host_scal + h * SL_COUNT;
        if (!std::isfinite(sl[SL_RHO]) || std::fabs(sl[SL_RHO]) <= DBL_MIN ||
            !std::isfinite(sl[SL_OMEGA]) || std::fabs(sl[SL_OMEGA]) <= DBL_MIN)
          act[h] = 0;
      }
      if (set_active() != 0) return -1;
      vec(OP_BI_P);
      precondition(V_P, V_PP);
      apply(V_PP, V_AP);
      dot(V_SH, V_AP, 0, SL_DEN);
      k_mi_scalar<<<1, 64, 0, s>>>(c, 0, active);
      vec(OP_BI_S);
      dot(V_S, V_S, 1, SL_NORM);
      if (read() != 0) return -1;
      for (int h = 0; h < 2; ++h) {
        if (!act[h]) continue;
        const double *sl = m.host_scal + h * SL_COUNT;
        if (!std::isfinite(sl[SL_DEN]) || std::fabs(sl[SL_DEN]) <= DBL_MIN) {
          act[h] = 0;
          continue;
        }
        if (sl[SL_NORM] <= target[h]) {
          last[h] = sl[SL_NORM];
          iters[h] = it;
          ok[h] = true;
        }
      }
      // converged hemispheres take x += alpha pp and stop
      {
        int conv[2] = {act[0] && ok[0], act[1] && ok[1]};
        if (conv[0] || conv[1]) {
          int keep[2] = {act[0], act[1]};
          act[0] = conv[0];
          act[1] = conv[1];
          if (set_active() != 0) return -1;
          vec(OP_BI_X_ONLY);
          act[0] = keep[0] && !conv[0];
          act[1] = keep[1] && !conv[1];
        }
        if (set_active() != 0) return -1;
        if (!act[0] && !act[1]) break;
      }
      precondition(V_S, V_PS);
      apply(V_PS, V_AS);
      dot(V_AS, V_AS, 0, SL_WDEN);
      dot(V_AS, V_S, 0, SL_WNUM);
      k_mi_scalar<<<1, 64, 0, s>>>(c, 1, active);
      if (read() != 0) return -1;
      for (int h = 0; h < 2; ++h) {
        if (!act[h]) continue;
        const double *sl = m.host_scal + h * SL_COUNT;
        if (!std::isfinite(sl[SL_WDEN]) || sl[SL_WDEN] <= DBL_MIN ||
            !std::isfinite(sl[SL_OMEGA]) || std::fabs(sl[SL_OMEGA]) <= DBL_MIN)
          act[h] = 0;
      }
      if (set_active() != 0) return -1;
      vec(OP_BI_XR);
      dot(V_R, V_R, 1, SL_NORM);
      if (read() != 0) return -1;
      bool verify = false;
      for (int h = 0; h < 2; ++h)
        if (act[h]) {
          iters[h] = it;
          last[h] = m.host_scal[h * SL_COUNT + SL_NORM];
          if (last[h] <= target[h] || it % 256 == 0) verify = true;
        }
      if (verify) {
        // true residual r = rhs - A x (the shadow residual restarts from it)
        apply(V_X, V_AP);
        vec(OP_RESIDUAL);  // also resets the shadow residual to r
        dot(V_R, V_R, 1, SL_NORM);
        if (read() != 0) return -1;
        for (int h = 0; h < 2; ++h)
          if (act[h]) {
            last[h] = m.host_scal[h * SL_COUNT + SL_NORM];
            if (last[h] <= target[h]) {
              ok[h] = true;
              act[h] = 0;
            }
          }
        if (set_active() != 0) return -1;
        vec(OP_ZERO_PAP);
        k_mi_scalar<<<1, 64, 0, s>>>(c, 3, active);
        continue;
      }
      k_mi_scalar<<<1, 64, 0, s>>>(c, 2, active);
    }
  }
  act[0] = act[1] = 1;
  if (set_active() != 0) return -1;
  vec(OP_EXPAND, 0.0);
  GCS_HIP(hipStreamSynchronize(s));
  m.solve_ms = std::chrono::duration<double, std::milli>(
                   std::chrono::steady_clock::now() - t0).count();
  for (int h = 0; h < 2; ++h) {
    m.iterations[h] = iters[h];
    m.final_residual[h] = last[h];
  }
  if (!ok[0] || !ok[1]) {
    set_error("ionosphere potential solve did not converge");
    return -1;
  }
  return 0;
}

}  // namespace

int mi_create(Problem &P) {
  const gcs_problem_config &cfg = P.cfg;
  const gamera_cs_mi_dipole_config &mc = cfg.mi;
  const Layout &L = P.L;
  // cubedsphere_mi_dipole.c:45-57, :157-177
  if (mc.azimuth_points < 8 || mc.colatitude_points < 3 || mc.current_shell >= 64 ||
      !(mc.ionosphere_radius_re > 0.0) || !(mc.pedersen_siemens > 0.0) ||
      !(mc.coupling_interval_s > 0.0) || !(mc.length_norm_m > 0.0) ||
      !(mc.speed_norm_m_s > 0.0) || !(mc.mass_density_norm_kg_m3 > 0.0) ||
      mc.maximum_iterations <= 0 || !(mc.relative_tolerance > 0.0) ||
      L.nr < 3 || (int)mc.current_shell + 1 >= L.nr ||
      !(mc.ionosphere_radius_re < cfg.startup.inner_radius) ||
      !(cfg.hall_siemens >= 0.0)) {
    set_error("invalid MI configuration");
    return -1;
  }
  MiState *mp = new MiState();
  P.mi = mp;
  MiState &m = *mp;
  m.cfg = mc;
  m.hall = cfg.hall_siemens;
  m.np = (int)mc.azimuth_points;
  m.nt = (int)mc.colatitude_points;
  m.n = 1 + m.np * (m.nt - 2);
  m.S = L.n + 2;
  m.si = L.ghost + (int)mc.current_shell;
  m.max_colat = asin(sqrt(mc.ionosphere_radius_re / cfg.startup.inner_radius));
  const double mu0 = 4.0e-7 * acos(-1.0);
  m.b_norm = mc.speed_norm_m_s * sqrt(mu0 * mc.mass_density_norm_kg_m3);
  m.current_norm = m.b_norm / (mu0 * mc.length_norm_m);
  m.potential_norm = mc.speed_norm_m_s * m.b_norm * mc.length_norm_m;
  // centroids for the sample radius and the plans
  std::vector<double> cen(3 * L.total);
  for (int e = 0; e < 3; ++e)
    GCS_HIP(hipMemcpy(cen.data() + e * L.total, P.plane(P_CEN + e),
                      L.total * sizeof(double), hipMemcpyDeviceToHost));
  auto centroid = [&](long at, double x[3]) {
    for (int e = 0; e < 3; ++e) x[e] = cen[e * L.total + at];
  };
  double sum = 0.0;
  long count = 0;
  for (int f = 0; f < 6; ++f)
    for (int j = L.ghost; j < L.ghost + L.n; ++j)
      for (int k = L.ghost; k < L.ghost + L.n; ++k) {
        double x[3];
        centroid(L.at(f, m.si, j, k), x);
        sum += sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        ++count;
      }
  m.sample_radius = sum / (double)count;
  m.sample_max_colat = asin(sqrt(mc.ionosphere_radius_re / m.sample_radius));
  if (!std::isfinite(m.sample_radius) || !(m.sample_radius > cfg.startup.inner_radius)) {
    set_error("MI sample shell inside the inner radius");
    return -1;
  }
  const double qp = 0.25 * acos(-1.0);
  const int n = L.n, S = m.S, g = L.ghost;
  // B fringe on shells si-1..si+1 and the FAC fringe on si (:236-276, :308-342)
  std::vector<Bilinear> bf, ff;
  for (int f = 0; f < 6; ++f)
    for (int i = m.si - 1; i <= m.si + 1; ++i)
      for (int lj = 0; lj < S; ++lj)
        for (int lk = 0; lk < S; ++lk) {
          if (lj > 0 && lj <= n && lk > 0 && lk <= n) continue;
          const long at = L.at(f, i, g - 1 + lj, g - 1 + lk);
          double x[3], a, b;
          int df;
          centroid(at, x);
          if (!cube_inverse(x, df, a, b) || df == f) {
            set_error("MI fringe donor");
            return -1;
          }
          const double xx = (a + qp) * (double)n / (2.0 * qp) - 0.5;
          const double yy = (b + qp) * (double)n / (2.0 * qp) - 0.5;
          const int j0 = (int)fmax(0.0, fmin((double)n - 2.0, floor(xx)));
          const int k0 = (int)fmax(0.0, fmin((double)n - 2.0, floor(yy)));
          Bilinear e{};
          e.wx = xx - (double)j0;
          e.wy = yy - (double)k0;
          e.target = (int)at;
          e.base = (int)L.at(df, i, g + j0, g + k0);
          e.step = L.SK;
          bf.push_back(e);
          if (i == m.si) {
            Bilinear q = e;
            q.target = f * S * S + lj * S + lk;
            q.base = df * S * S + (j0 + 1) * S + k0 + 1;
            q.step = S;
            ff.push_back(q);
          }
        }
  // polar samples (:391-406, :344-370)
  std::vector<Bilinear> sm;
  const int np = m.np, nt = m.nt;
  const double dt = m.max_colat / (double)(nt - 1);
  const double dp = 2.0 * acos(-1.0) / (double)np;
  for (int h = 0; h < 2; ++h) {
    const double sign = h == 0 ? 1.0 : -1.0;
    for (int t = 0; t < nt; ++t) {
      const double ti = (double)t * dt;
      const bool zero = ti > m.sample_max_colat + 4e-14;
      const double arg = fmin(1.0, sin(ti) * sqrt(m.sample_radius / mc.ionosphere_radius_re));
      const double tm = asin(arg);
      for (int p = 0; p < np; ++p) {
        Bilinear e{};
        e.target = h * nt * np + t * np + p;
        if (zero) {
          e.zero = 1;
          sm.push_back(e);
          continue;
        }
        const double phi = (double)p * dp;
        const double x[3] = {m.sample_radius * sin(tm) * cos(phi),
                             m.sample_radius * sin(tm) * sin(phi),
                             sign * m.sample_radius * cos(tm)};
        int f;
        double a, b;
        if (!cube_inverse(x, f, a, b)) {
          set_error("MI sample point");
          return -1;
        }
        const double scale = (double)n / (2.0 * qp);
        const double xx = (a + qp) * scale - 0.5, yy = (b + qp) * scale - 0.5;
        const long jj = (long)floor(xx), kk = (long)floor(yy);
        if (jj < -1 || kk < -1 || jj >= n || kk >= n) {
          set_error("MI sample outside the shell");
          return -1;
        }
        e.wx = xx - (double)jj;
        e.wy = yy - (double)kk;
        e.base = f * S * S + (int)(jj + 1) * S + (int)(kk + 1);
        e.step = S;
        sm.push_back(e);
      }
    }
  }
  m.n_bfringe = (int)bf.size();
  m.n_ffringe = (int)ff.size();
  m.n_samples = (int)sm.size();
  if (upload_vec(bf, &m.bfringe) != 0 || upload_vec(ff, &m.ffringe) != 0 ||
      upload_vec(sm, &m.samples) != 0)
    return -1;
  const long polar = 2L * nt * np;
  GCS_HIP(hipMalloc(&m.shell_fac, 6L * S * S * sizeof(double)));
  GCS_HIP(hipMalloc(&m.fac, polar * sizeof(double)));
  GCS_HIP(hipMalloc(&m.potential, polar * sizeof(double)));
  GCS_HIP(hipMalloc(&m.work, 2L * kVectors * m.n * sizeof(double)));
  GCS_HIP(hipMalloc(&m.scal, 2 * SL_COUNT * sizeof(double)));
  GCS_HIP(hipMalloc(&m.partial, 2 * kMaxBlocks * sizeof(double)));
  GCS_HIP(hipHostMalloc(&m.host_scal, 2 * SL_COUNT * sizeof(double)));
  GCS_HIP(hipMalloc(&m.active, 2 * sizeof(int)));
  GCS_HIP(hipMemset(m.shell_fac, 0, 6L * S * S * sizeof(double)));
  GCS_HIP(hipMemset(m.fac, 0, polar * sizeof(double)));
  GCS_HIP(hipMemset(m.potential, 0, polar * sizeof(double)));
  GCS_HIP(hipMemset(m.work, 0, 2L * kVectors * m.n * sizeof(double)));
  GCS_HIP(hipMemset(m.scal, 0, 2 * SL_COUNT * sizeof(double)));
  // the view the step kernels read (potential set once ready)
  MiView &v = P.mi_view;
  v = MiView();
  v.np = np;
  v.nt = nt;
  v.ri = mc.ionosphere_radius_re;
  v.max_colat = m.max_colat;
  v.potential_norm = m.potential_norm;
  v.length_norm = mc.length_norm_m;
  v.speed_norm = mc.speed_norm_m_s;
  v.b_norm = m.b_norm;
  for (int e = 0; e < 3; ++e) v.moment[e] = cfg.startup.dipole.moment.value[e];
  v.cut_radius = cfg.startup.cut_radius;
  v.cut_length = cfg.startup.cut_length;
  return 0;
}

void mi_destroy(Problem &P) {
  MiState *m = P.mi;
  if (m == nullptr) return;
  (void)hipFree(m->shell_fac);
  (void)hipFree(m->fac);
  (void)hipFree(m->potential);
  (void)hipFree(m->work);
  (void)hipFree(m->scal);
  (void)hipFree(m->partial);
  (void)hipFree(m->bfringe);
  (void)hipFree(m->ffringe);
  (void)hipFree(m->samples);
  (void)hipFree(m->active);
  if (m->host_scal) (void)hipHostFree(m->host_scal);
  (void)hipFree(m->warm_scale);
  delete m;
  P.mi = nullptr;
}

// cubedsphere_mi_dipole.c:372-436
int mi_maybe_update(Problem &P, double time_code) {
  MiState *mp = P.mi;
  if (mp == nullptr) return 0;
  MiState &m = *mp;
  const double ts = time_code * m.cfg.length_norm_m / m.cfg.speed_norm_m_s;
  if (m.ready && !(ts + 1e-10 >= m.next_update)) return 0;
  MiK c = make_mik(P);
  hipStream_t s = P.stream;
  const Layout &L = P.L;
  mark_stage(P, S_MI);
  k_mi_copy_shell<<<blocks_of(6L * 3 * L.NJ * L.NK), kT, 0, s>>>(c);
  k_mi_fringe_b<<<blocks_of(m.n_bfringe), kT, 0, s>>>(c, m.bfringe, m.n_bfringe);
  k_mi_shell_fac<<<blocks_of(6L * L.n * L.n), kT, 0, s>>>(c);
  k_mi_bilinear<<<blocks_of(m.n_ffringe), kT, 0, s>>>(m.ffringe, m.n_ffringe, m.shell_fac,
                                                     m.shell_fac);
  k_mi_bilinear<<<blocks_of(m.n_samples), kT, 0, s>>>(m.samples, m.n_samples, m.shell_fac,
                                                      m.fac);
  GCS_HIP(hipGetLastError());
  StepScalars sc;
  if (read_scalars(P, sc) != 0) return -1;
  if (sc.failed) {
    set_error("MI field-aligned current failed at stage " + std::to_string(sc.fail_stage));
    return -1;
  }
  if (solve(P, m.hall > 0.0) != 0) return -1;
  m.ready = true;
  m.epoch = ts;
  ++m.updates;
  do {
    m.next_update += m.cfg.coupling_interval_s;
  } while (m.next_update <= ts + 1e-10);
  P.mi_view.potential = m.potential;
  return 1;
}

void mi_report(const Problem &P, MiReport &out) {
  std::memset(&out, 0, sizeof(out));
  if (P.mi == nullptr) return;
  const MiState &m = *P.mi;
  out.updates = m.updates;
  out.ready = m.ready;
  for (int h = 0; h < 2; ++h) {
    out.iterations[h] = m.iterations[h];
    out.final_residual[h] = m.final_residual[h];
  }
  out.epoch_s = m.epoch;
  out.solve_ms = m.solve_ms;
}

int mi_download(Problem &P, int what, std::vector<double> &out) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  const MiState &m = *P.mi;
  out.resize(2L * m.nt * m.np);
  GCS_HIP(hipStreamSynchronize(P.stream));
  GCS_HIP(hipMemcpy(out.data(), what == 0 ? m.potential : m.fac,
                    out.size() * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}


// ---- training pairs for the warm-start network ------------------------------------
//
// A pair is a potential and the field-aligned current for which it solves the
// polar system exactly: a random smooth potential (Gaussian bumps in
// colatitude times von Mises bumps in longitude, tapered to zero at the
// low-latitude edge, the pole row set to the mean of the first ring as the
// reduced system's pole equation requires) is multiplied by the tensor
// operator of a uniform (SigmaP, SigmaH), and the interior rows of the product
// are turned back into the current that produces them.

namespace {

constexpr int kBumps = 6;

struct GenSample {
  int hemi;
  double sp, sh;
  double amp[kBumps], th[kBumps], sth[kBumps], ph[kBumps], kap[kBumps];
};

__global__ void k_gen_potential(const GenSample *smp, int nt, int np, double max_colat,
                                double *phi) {
  const int b = blockIdx.y;
  const long q = gtid();
  if (q >= (long)nt * np) return;
  const int i = (int)(q / np), j = (int)(q % np);
  const double dt = max_colat / (double)(nt - 1), dp = 2.0 * kPi / (double)np;
  const double th = (double)i * dt, ph = (double)j * dp;
  const GenSample &g = smp[b];
  double v = 0.0;
  for (int k = 0; k < kBumps; ++k) {
    const double d = (th - g.th[k]) / g.sth[k];
    v += g.amp[k] * exp(-0.5 * d * d) * exp(g.kap[k] * (cos(ph - g.ph[k]) - 1.0));
  }
  const double taper = 1.0 - (th / max_colat) * (th / max_colat);
  phi[(long)b * nt * np + q] = i == nt - 1 ? 0.0 : taper * v;
}

// Pole row = mean of ring 1 (one block per sample).
__global__ void k_gen_pole(int nt, int np, double *phi) {
  __shared__ double red[kT];
  double *p = phi + (long)blockIdx.x * nt * np;
  double sum = 0.0;
  for (int j = threadIdx.x; j < np; j += kT) sum += p[np + j];
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double mean = red[0] / (double)np;
  for (int j = threadIdx.x; j < np; j += kT) p[j] = mean;
}

__global__ void k_gen_current(const GenSample *smp, int nt, int np, double max_colat,
                              double r2, const double *phi, double *fac) {
  const int b = blockIdx.y;
  const long q = gtid();
  if (q >= (long)nt * np) return;
  const int i = (int)(q / np), j = (int)(q % np);
  const long base = (long)b * nt * np;
  if (i == 0 || i == nt - 1) {
    fac[base + q] = 0.0;
    return;
  }
  const GenSample &g = smp[b];
  const double dt = max_colat / (double)(nt - 1), dp = 2.0 * kPi / (double)np;
  const TensorRow t = tensor_row(g.sp, g.sh, i, dt, dp, g.hemi);
  const double *p = phi + base;
  const int jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
  const double below = p[(long)(i - 1) * np + (i == 1 ? 0 : j)];
  const double above = i + 2 < nt ? p[(long)(i + 1) * np + j] : 0.0;
  const double y = t.diag * p[q] + t.lo * below + t.hi * above + t.llo * p[(long)i * np + jm] +
                   t.lhi * p[(long)i * np + jp];
  fac[base + q] = equivalent_current(y, i, nt, dt, r2, g.hemi);
}

// RMS of the current over the interior rows (one block per sample).
__global__ void k_gen_scale(int nt, int np, const double *fac, double *scale) {
  __shared__ double red[kT];
  const double *f = fac + (long)blockIdx.x * nt * np;
  const long count = (long)(nt - 2) * np;
  double sum = 0.0;
  for (long q = threadIdx.x; q < count; q += kT) sum += f[np + q] * f[np + q];
  red[threadIdx.x] = sum;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) scale[blockIdx.x] = sqrt(red[0] / (double)count);
}

__global__ void k_gen_features(const GenSample *smp, int nt, int np, double r2,
                               const double *fac, const double *phi, const double *scale,
                               double out_scale, float *feat, float *target) {
  const int b = blockIdx.y;
  const long q = gtid();
  if (q >= (long)nt * np) return;
  const int i = (int)(q / np);
  const GenSample &g = smp[b];
  const long at = (long)b * nt * np + q;
  write_features(feat + at * unet::kInChannels, fac[at], scale[b], g.sp, g.sh, i, nt, g.hemi);
  target[at] = (float)(phi[at] * g.sp / (r2 * scale[b] * out_scale));
}

// A batch of random pair parameters (host side, seeded).
void draw_samples(std::mt19937_64 &rng, int count, double max_colat, double sp_lo,
                  double sp_hi, double sh_lo, double sh_hi, GenSample *out) {
  std::uniform_real_distribution<double> u(0.0, 1.0);
  std::normal_distribution<double> z(0.0, 1.0);
  for (int b = 0; b < count; ++b) {
    GenSample &g = out[b];
    g.hemi = u(rng) < 0.5 ? GAMERA_MI_NORTH : GAMERA_MI_SOUTH;
    g.sp = sp_lo + (sp_hi - sp_lo) * u(rng);
    g.sh = sh_lo + (sh_hi - sh_lo) * u(rng);
    for (int k = 0; k < kBumps; ++k) {
      g.amp[k] = z(rng);
      g.th[k] = max_colat * (0.05 + 0.85 * u(rng));
      g.sth[k] = max_colat * (0.03 + 0.22 * u(rng));
      g.ph[k] = 2.0 * kPi * u(rng);
      g.kap[k] = 0.3 + 8.0 * u(rng);
    }
  }
}

}  // namespace

// Device buffers of a pair generator for `batch` samples on an nt x np grid.
struct PairBatch {
  GenSample *host = nullptr, *dev = nullptr;
  double *phi = nullptr, *fac = nullptr, *scale = nullptr;
  int batch = 0;
  ~PairBatch() {
    if (host) (void)hipHostFree(host);
    (void)hipFree(dev);
    (void)hipFree(phi);
    (void)hipFree(fac);
    (void)hipFree(scale);
  }
  int allocate(int b, int nt, int np) {
    batch = b;
    GCS_HIP(hipHostMalloc(&host, b * sizeof(GenSample)));
    GCS_HIP(hipMalloc(&dev, b * sizeof(GenSample)));
    GCS_HIP(hipMalloc(&phi, (size_t)b * nt * np * sizeof(double)));
    GCS_HIP(hipMalloc(&fac, (size_t)b * nt * np * sizeof(double)));
    GCS_HIP(hipMalloc(&scale, (size_t)b * sizeof(double)));
    return 0;
  }
};

// Potentials and currents of host->dev samples on stream s.
static int make_pairs(PairBatch &pb, int nt, int np, double max_colat, double r2,
                      hipStream_t s) {
  GCS_HIP(hipMemcpyAsync(pb.dev, pb.host, pb.batch * sizeof(GenSample), hipMemcpyHostToDevice,
                         s));
  const dim3 g(blocks_of((long)nt * np), pb.batch);
  k_gen_potential<<<g, kT, 0, s>>>(pb.dev, nt, np, max_colat, pb.phi);
  k_gen_pole<<<pb.batch, kT, 0, s>>>(nt, np, pb.phi);
  k_gen_current<<<g, kT, 0, s>>>(pb.dev, nt, np, max_colat, r2, pb.phi, pb.fac);
  k_gen_scale<<<pb.batch, kT, 0, s>>>(nt, np, pb.fac, pb.scale);
  GCS_HIP(hipGetLastError());
  return 0;
}

int mi_attach_warm_start(Problem &P, unet::Net *net) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  MiState &m = *P.mi;
  if (net != nullptr && (net->H != m.nt || net->W != m.np || net->max_batch < 2)) {
    set_error("warm-start network grid does not match the MI grid (colatitude x azimuth), "
              "or its batch is below 2");
    return -1;
  }
  if (m.warm_scale == nullptr) GCS_HIP(hipMalloc(&m.warm_scale, 2 * sizeof(double)));
  m.net = net;
  return 0;
}

int mi_make_pairs(Problem &P, uint64_t seed, int count, double sp, double sh, double *fac_out,
                  double *phi_out) {
  if (P.mi == nullptr || count <= 0) {
    set_error(P.mi == nullptr ? "MI is not enabled" : "count must be positive");
    return -1;
  }
  const MiState &m = *P.mi;
  const double rm = m.cfg.ionosphere_radius_re * m.cfg.length_norm_m;
  PairBatch pb;
  if (pb.allocate(count, m.nt, m.np) != 0) return -1;
  std::mt19937_64 rng(seed);
  draw_samples(rng, count, m.max_colat, sp, sp, sh, sh, pb.host);
  for (int b = 0; b < count; ++b) pb.host[b].hemi = b % 2 == 0 ? GAMERA_MI_NORTH : GAMERA_MI_SOUTH;
  if (make_pairs(pb, m.nt, m.np, m.max_colat, rm * rm, P.stream) != 0) return -1;
  GCS_HIP(hipStreamSynchronize(P.stream));
  const size_t bytes = (size_t)count * m.nt * m.np * sizeof(double);
  GCS_HIP(hipMemcpy(fac_out, pb.fac, bytes, hipMemcpyDeviceToHost));
  GCS_HIP(hipMemcpy(phi_out, pb.phi, bytes, hipMemcpyDeviceToHost));
  return 0;
}

int mi_train_warm_start(Problem &P, unet::Net &n, const gcs_warm_train_config &t,
                        float *loss_history) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  const MiState &m = *P.mi;
  if (n.H != m.nt || n.W != m.np || unet::check_batch(n, t.batch) != 0 || t.steps <= 0 ||
      !(t.learning_rate > 0.0) || !(t.pedersen_min > 0.0) ||
      !(t.pedersen_max >= t.pedersen_min) || !(t.hall_min >= 0.0) ||
      !(t.hall_max >= t.hall_min)) {
    set_error("invalid warm-start training configuration");
    return -1;
  }
  const double rm = m.cfg.ionosphere_radius_re * m.cfg.length_norm_m;
  const double r2 = rm * rm;
  hipStream_t s = n.stream;
  PairBatch pb;
  if (pb.allocate(t.batch, m.nt, m.np) != 0) return -1;
  float *host_loss = nullptr;
  GCS_HIP(hipHostMalloc(&host_loss, sizeof(float)));
  std::mt19937_64 rng(t.seed);
  const dim3 g(blocks_of((long)m.nt * m.np), t.batch);
  int rc = 0;
  for (long step = 0; step < t.steps && rc == 0; ++step) {
    draw_samples(rng, t.batch, m.max_colat, t.pedersen_min, t.pedersen_max, t.hall_min,
                 t.hall_max, pb.host);
    if (make_pairs(pb, m.nt, m.np, m.max_colat, r2, s) != 0) {
      rc = -1;
      break;
    }
    if (n.adam_t == 0 && step == 0) {
      // output scale: unit RMS of the first batch's targets
      k_gen_features<<<g, kT, 0, s>>>(pb.dev, m.nt, m.np, r2, pb.fac, pb.phi, pb.scale, 1.0,
                                      n.input, n.target);
      std::vector<float> tg((size_t)t.batch * m.nt * m.np);
      if (hipMemcpyAsync(tg.data(), n.target, tg.size() * sizeof(float),
                         hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("warm-start training: target readback failed");
        rc = -1;
        break;
      }
      double ss = 0.0;
      for (float v : tg) ss += (double)v * v;
      n.out_scale = std::sqrt(ss / (double)tg.size());
      if (!(n.out_scale > 0.0) || !std::isfinite(n.out_scale)) n.out_scale = 1.0;
    }
    k_gen_features<<<g, kT, 0, s>>>(pb.dev, m.nt, m.np, r2, pb.fac, pb.phi, pb.scale,
                                    n.out_scale, n.input, n.target);
    // cosine decay to 5% of the rate after a short linear warmup
    const double warm = std::min(1.0, (double)(step + 1) / 50.0);
    const double frac = (double)step / (double)std::max(1L, t.steps - 1);
    const double lr = t.learning_rate * warm * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(kPi * frac)));
    if (unet::forward_device(n, t.batch) != 0 || unet::loss_device(n, t.batch) != 0 ||
        unet::backward_device(n, t.batch) != 0 || unet::adam_device(n, (float)lr) != 0) {
      rc = -1;
      break;
    }
    if (loss_history != nullptr) {
      if (hipMemcpyAsync(host_loss, n.loss, sizeof(float), hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("warm-start training: loss readback failed");
        rc = -1;
        break;
      }
      loss_history[step] = *host_loss;
    }
  }
  (void)hipHostFree(host_loss);
  if (rc == 0 && hipStreamSynchronize(s) != hipSuccess) {
    set_error("warm-start training: stream error");
    rc = -1;
  }
  return rc;
}

// One polar solve of both hemispheres from host arrays ([2][nt][np] current
// and guess; the solution replaces the guess), with or without the attached
// network. Leaves the problem's MI state at the solved potential.
int mi_solve_current(Problem &P, const double *fac, double *potential, int use_net,
                     int iterations[2], double *ms) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  MiState &m = *P.mi;
  if (use_net && m.net == nullptr) {
    set_error("no warm-start network attached");
    return -1;
  }
  const size_t bytes = 2UL * m.nt * m.np * sizeof(double);
  GCS_HIP(hipStreamSynchronize(P.stream));
  GCS_HIP(hipMemcpy(m.fac, fac, bytes, hipMemcpyHostToDevice));
  GCS_HIP(hipMemcpy(m.potential, potential, bytes, hipMemcpyHostToDevice));
  unet::Net *keep = m.net;
  if (!use_net) m.net = nullptr;
  const int rc = solve(P, m.hall > 0.0);
  m.net = keep;
  iterations[0] = m.iterations[0];
  iterations[1] = m.iterations[1];
  *ms = m.solve_ms;
  if (rc != 0) return -1;
  GCS_HIP(hipMemcpy(potential, m.potential, bytes, hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace gcs
