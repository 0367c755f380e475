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
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
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
  SL_WDEN, SL_ALPHA, SL_OMEGA, SL_RHOPREV, SL_POLE, SL_ITER, SL_OK, SL_COUNT = 16
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
  double initial_residual[2] = {0.0, 0.0};
  double final_residual[2] = {0.0, 0.0};
  int converged[2] = {0, 0};
  double solve_ms = 0.0;
  // device arrays
  double *shell_fac = nullptr;  // [6][S][S]
  // J.b0/|b0| (A/m^2, before the dipole mapping), magnetic latitude and
  // longitude (degrees) of the active sample-shell cells at the last update,
  // [3][6][n][n]
  double *shell_current = nullptr;
  double *fac = nullptr;        // [2][nt][np]
  double *potential = nullptr;  // [2][nt][np]
  // E x B drift of the inner ghost cells for the current potential
  // (MiView::drift), refreshed by every writer of `potential`
  double *drift = nullptr;
  double *work = nullptr;       // [2][kVectors][n] + ring scratch
  double *scal = nullptr;       // [2][SL_COUNT]
  double *host_scal = nullptr;  // pinned mirror of scal
  Bilinear *bfringe = nullptr, *ffringe = nullptr, *samples = nullptr;
  int n_bfringe = 0, n_ffringe = 0, n_samples = 0;
  int *active = nullptr;  // per-hemisphere "still iterating" flags
  // learned warm start (not owned): the initial guess becomes the previous
  // potential plus the network's answer for the current residual
  unet::Net *net = nullptr;
  double *warm_scale = nullptr;  // [2] residual-current RMS, [2] step lengths
  double *coop_slots = nullptr;  // partial sums of the cooperative solve
  int *coop_flags = nullptr;     // its per-hemisphere liveness exchange
};

namespace {

constexpr int kVectors = 20;   // solver vectors per hemisphere

struct MiK {
  Layout L;
  double *pl;
  uint8_t *by;
  StepScalars *sc;
  int cur;
  int np, nt, n, S, si;
  double ri, max_colat, current_norm, pedersen, r2;
  double *shell_fac, *fac, *potential, *work, *scal;
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
// V_SCR, V_SCR2, V_IB, V_RDEN: the ring preconditioner's factors (forward
// factors g, Sherman-Morrison vector z, 1 / pivots, per-ring denominator).
enum Vec { V_RHS, V_DIAG, V_X, V_R, V_Z, V_P, V_AP, V_SH, V_PP, V_S, V_PS, V_AS,
           V_LO, V_HI, V_LLO, V_LHI, V_SCR, V_SCR2, V_IB, V_RDEN };

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
__global__ void k_mi_shell_fac(MiK c, double *shell_current) {
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
  // the shell's own J-parallel, magnetic latitude and longitude
  // (cubedsphere_mi_dipole.c:590-618)
  if (shell_current != nullptr) {
    const long q = ((long)f * n + (lj - 1)) * n + (lk - 1), per = 6L * n * n;
    double phi = atan2(x[1], x[0]) * 180.0 / kPi;
    if (phi < 0.0) phi += 360.0;
    shell_current[q] = jb / bm * c.current_norm;
    shell_current[per + q] = asin(x[2] / rad) * 180.0 / kPi;
    shell_current[2 * per + q] = phi;
  }
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

// Elementwise steps outside the fused solves.
enum Op {
  OP_RESIDUAL,  // r = rhs - ap; shadow = r
  OP_EXPAND     // potential from x (outer ring = extra)
};

__global__ void k_mi_vec(MiK c, int op, double extra, const int *active) {
  const int h = blockIdx.y;
  if (!active[h]) return;
  const long q = gtid();
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
  double *r = c.vec(h, V_R);
  r[q] = c.vec(h, V_RHS)[q] - c.vec(h, V_AP)[q];
  c.vec(h, V_SH)[q] = r[q];
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
  return c;
}

// ---- fused solves: one workgroup per hemisphere runs the whole Krylov loop -------
//
// The same iterations as the reference's solvers (CG: :516-600, BiCGStab:
// :1029-1131, with the ring preconditioner above), but every scalar stays on
// the device: the workgroup's threads stride over the unknowns, dot products
// are workgroup reductions in a fixed order, and the loop, its breakdown and
// convergence tests and the periodic true-residual checks run in the kernel.
// The host launches once and reads the iteration count and final residual.
// Every branch depends only on values all threads share, so the whole
// workgroup leaves the loop together, at the latest after maximum_iterations.

constexpr int kSolveThreads = 1024;

// Sum over the workgroup, returned to every thread (fixed order).
GCS_DEV double block_sum(double v, double *red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();  // red[] may still be read from the previous call
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < kSolveThreads / 64; ++w) s += red[w];
  return s;
}

// One term of the residual norm of the unweighted equations (:394-410).
GCS_DEV double orig_term(const MiK &c, double a, long q, double ps, double dt) {
  if (q == 0) {
    const double pr = a / ps;
    return (double)c.np * pr * pr;
  }
  const int i = 1 + (int)((q - 1) / c.np);
  const double v = a / sin((double)i * dt);
  return v * v;
}

// Rows of the Pedersen operator (w, lo, hi, lon per colatitude index),
// computed with k_mi_apply's expressions.
GCS_DEV void pedersen_rows(const MiK &c, double *rows) {
  const double dt = c.max_colat / (double)(c.nt - 1);
  const double dp = 2.0 * kPi / (double)c.np;
  for (int i = 1 + threadIdx.x; i + 1 < c.nt; i += blockDim.x) {
    const double th = (double)i * dt;
    rows[4 * i + 0] = sin(th);
    rows[4 * i + 1] = theta_coeff(c.pedersen, th - 0.5 * dt, th, dt);
    rows[4 * i + 2] = theta_coeff(c.pedersen, th + 0.5 * dt, th, dt);
    rows[4 * i + 3] = c.pedersen / (sin(th) * sin(th) * dp * dp);
  }
}

// (A x)[q], q >= 1, with k_mi_apply's arithmetic (Pedersen rows from `rows`).
GCS_DEV double apply_at(const MiK &c, int h, bool tensor, const double *x, long q,
                        const double *rows, double pe) {
  const int np = c.np, nt = c.nt;
  const int i = 1 + (int)((q - 1) / np), j = (int)((q - 1) % np);
  const int jm = j == 0 ? np - 1 : j - 1, jp = j + 1 == np ? 0 : j + 1;
  const double below = i == 1 ? x[0] : x[red_at(np, i - 1, j)];
  const double above = i + 2 < nt ? x[red_at(np, i + 1, j)] : 0.0;
  if (!tensor) {
    const double w = rows[4 * i], lo = rows[4 * i + 1], hi = rows[4 * i + 2],
                 lon = rows[4 * i + 3];
    return w * (lo * (x[q] - below) + hi * (x[q] - above) +
                lon * (2.0 * x[q] - x[red_at(np, i, jm)] - x[red_at(np, i, jp)]));
  }
  return c.vec(h, V_DIAG)[q] * x[q] + c.vec(h, V_LO)[q] * below +
         c.vec(h, V_HI)[q] * above + c.vec(h, V_LLO)[q] * x[red_at(np, i, jm)] +
         c.vec(h, V_LHI)[q] * x[red_at(np, i, jp)];
}

// Longitude-line preconditioner of the tensor system: out = M^-1 in with M
// the pole diagonal plus, on each colatitude ring, the ring's own periodic
// tridiagonal coupling, solved by the Thomas algorithm with a
// Sherman-Morrison correction for the periodic corners.
//
// The factors depend on the matrix only, so each solve computes them once
// (one thread per ring): g (forward factors), 1 / pivots, the correction
// vector z and the correction's denominator. Each application is then two
// first-order linear recurrences per ring, which one wavefront evaluates as
// a parallel prefix of affine maps (x -> A + B x) across its 64 lanes.
GCS_DEV void ring_factor(const MiK &c, int h, int ring) {
  const int np = c.np;
  const int base = 1 + ring * np;
  const double *__restrict__ a = c.vec(h, V_LLO) + base;  // coefficient of x[j-1]
  const double *__restrict__ b = c.vec(h, V_DIAG) + base;
  const double *__restrict__ cc = c.vec(h, V_LHI) + base; // coefficient of x[j+1]
  double *__restrict__ g = c.vec(h, V_SCR) + base;
  double *__restrict__ z = c.vec(h, V_SCR2) + base;
  double *__restrict__ ib = c.vec(h, V_IB) + base;
  const double gam = -b[0];
  const double alpha = cc[np - 1];  // row np-1's coupling to x[0]
  const double beta = a[0];         // row 0's coupling to x[np-1]
  double inv = 1.0 / (b[0] - gam);
  double zp = gam * inv;
  ib[0] = inv;
  g[0] = 0.0;
  z[0] = zp;
  for (int j = 1; j < np; ++j) {
    const double gj = cc[j - 1] * inv;
    g[j] = gj;
    const double bj = j == np - 1 ? b[j] - alpha * beta / gam : b[j];
    inv = 1.0 / (bj - a[j] * gj);
    ib[j] = inv;
    zp = ((j == np - 1 ? alpha : 0.0) - a[j] * zp) * inv;
    z[j] = zp;
  }
  const double zl = zp;
  for (int j = np - 2; j >= 0; --j) {
    zp = z[j] - g[j + 1] * zp;
    z[j] = zp;
  }
  c.vec(h, V_RDEN)[base] = 1.0 + zp + beta * zl / gam;
}

GCS_DEV void factor_rings(const MiK &c, int h) {
  for (int ring = threadIdx.x; ring < c.nt - 2; ring += blockDim.x) ring_factor(c, h, ring);
}

// One ring's solve by one wavefront (lane = threadIdx.x & 63): each lane
// holds kRingChunk consecutive points of the ring in registers, so a ring
// costs one load phase and one store phase.
constexpr int kRingChunk = 8;  // rings up to 64 * 8 = 512 points

GCS_DEV void ring_apply(const MiK &c, int h, const double *in, double *out, int ring,
                        int lane) {
  const int np = c.np;
  const int base = 1 + ring * np;
  const double *a = c.vec(h, V_LLO) + base;
  const double *g = c.vec(h, V_SCR) + base;
  const double *z = c.vec(h, V_SCR2) + base;
  const double *ib = c.vec(h, V_IB) + base;
  const double *rhs = in + base;
  double *x = out + base;
  const int k = (np + 63) / 64;
  const int j0 = min(np, lane * k);
  const int cnt = min(np, j0 + k) - j0;
  double fa[kRingChunk], fb[kRingChunk], gd[kRingChunk], zz[kRingChunk], v[kRingChunk];
#pragma unroll
  for (int t = 0; t < kRingChunk; ++t) {
    const int j = j0 + t;
    const bool in_range = t < cnt;
    const double ibj = in_range ? ib[j] : 0.0;
    fa[t] = in_range ? ibj * rhs[j] : 0.0;
    fb[t] = in_range && j > 0 ? -a[j] * ibj : 0.0;
    gd[t] = in_range && j < np - 1 ? -g[j + 1] : 0.0;
    zz[t] = in_range ? z[j] : 0.0;
  }
  // forward: x_j = fa_j + fb_j x_{j-1}, as a prefix of affine maps
  double A = 0.0, B = 1.0;
#pragma unroll
  for (int t = 0; t < kRingChunk; ++t)
    if (t < cnt) {
      A = fa[t] + fb[t] * A;
      B = fb[t] * B;
    }
  for (int off = 1; off < 64; off <<= 1) {
    const double ao = __shfl_up(A, off, 64), bo = __shfl_up(B, off, 64);
    if (lane >= off) {
      A = A + B * ao;
      B = B * bo;
    }
  }
  double w = __shfl_up(A, 1, 64);
#pragma unroll
  for (int t = 0; t < kRingChunk; ++t)
    if (t < cnt) {
      w = fa[t] + fb[t] * w;
      v[t] = w;
    }
  // backward: y_j = x_j + gd_j y_{j+1} (gd_j = -g_{j+1}; 0 for the last point)
  double C = 0.0, D = 1.0;
#pragma unroll
  for (int t = kRingChunk - 1; t >= 0; --t)
    if (t < cnt) {
      C = v[t] + gd[t] * C;
      D = gd[t] * D;
    }
  for (int off = 1; off < 64; off <<= 1) {
    const double co = __shfl_down(C, off, 64), dd = __shfl_down(D, off, 64);
    if (lane + off < 64) {
      C = C + D * co;
      D = D * dd;
    }
  }
  w = __shfl_down(C, 1, 64);
#pragma unroll
  for (int t = kRingChunk - 1; t >= 0; --t)
    if (t < cnt) {
      w = v[t] + gd[t] * w;
      v[t] = w;
    }
  // Sherman-Morrison correction with y_0 and y_{np-1}
  const double y0 = __shfl(v[0], 0, 64);
  const int owner = (np - 1) / k;
  double mine_last = v[0];
#pragma unroll
  for (int t = 0; t < kRingChunk; ++t)
    if (t == (np - 1) - owner * k) mine_last = v[t];
  const double yl = __shfl(mine_last, owner, 64);
  const double gam = -c.vec(h, V_DIAG)[base];
  const double fact = (y0 + a[0] * yl / gam) / c.vec(h, V_RDEN)[base];
#pragma unroll
  for (int t = 0; t < kRingChunk; ++t)
    if (t < cnt) x[j0 + t] = v[t] - fact * zz[t];
}

GCS_DEV void precondition_all(const MiK &c, int h, const double *in, double *out) {
  __syncthreads();  // `in` complete
  if (threadIdx.x == 0) out[0] = in[0] / c.vec(h, V_DIAG)[0];
  const int lane = threadIdx.x & 63, waves = blockDim.x >> 6;
  for (int ring = threadIdx.x >> 6; ring < c.nt - 2; ring += waves)
    ring_apply(c, h, in, out, ring, lane);
  __syncthreads();  // `out` complete
}

// (A x)[0], the pole row, by wavefront 0 (every lane returns it).
GCS_DEV double pole_row(const MiK &c, int h, bool tensor, const double *x, double pe) {
  const int np = c.np;
  double ring = 0.0;
  for (int j = threadIdx.x & 63; j < np; j += 64) ring += x[red_at(np, 1, j)];
  for (int off = 32; off > 0; off >>= 1) ring += __shfl_xor(ring, off, 64);
  if (!tensor) return pe * ((double)np * x[0] - ring);
  return c.vec(h, V_DIAG)[0] * x[0] - pe * ring;
}

struct SolveResult {
  int iters;
  bool ok;
  double last;
};

GCS_DEV void store_result(const MiK &c, int h, const SolveResult &r, double first) {
  if (threadIdx.x != 0) return;
  double *sl = c.sl(h);
  sl[SL_ITER] = (double)r.iters;
  sl[SL_OK] = r.ok ? 1.0 : 0.0;
  sl[SL_NORM] = r.last;
  sl[SL_FIRST] = first;
}

// Jacobi-preconditioned CG on the reduced Pedersen system (:516-600).
__global__ void __launch_bounds__(kSolveThreads) k_mi_cg(MiK c, double rel_tol, int maxit) {
  __shared__ double red[kSolveThreads / 64];
  extern __shared__ double rows[];
  const int h = blockIdx.x;
  const long n = c.n;
  const long t0 = threadIdx.x, st = blockDim.x;
  const double dt = c.max_colat / (double)(c.nt - 1);
  const double half = 0.5 * dt;
  const double pe = sin(half) * c.pedersen / (incl2(half) * dt * dt);
  const double ps = (double)c.np * pe;
  pedersen_rows(c, rows);
  __syncthreads();
  double *x = c.vec(h, V_X), *r = c.vec(h, V_R), *z = c.vec(h, V_Z), *p = c.vec(h, V_P),
         *ap = c.vec(h, V_AP);
  const double *rhs = c.vec(h, V_RHS), *diag = c.vec(h, V_DIAG);
  double a = 0.0, b = 0.0;
  const double pole = threadIdx.x < 64 ? pole_row(c, h, false, x, pe) : 0.0;
  for (long q = t0; q < n; q += st) {
    r[q] = rhs[q] - (q == 0 ? pole : apply_at(c, h, false, x, q, rows, pe));
    a += orig_term(c, r[q], q, ps, dt);
    b += q == 0 ? 0.0 : orig_term(c, rhs[q], q, ps, dt);
  }
  double norm = sqrt(fmax(0.0, block_sum(a, red)));
  const double first = norm;
  const double target = rel_tol * fmax(sqrt(fmax(0.0, block_sum(b, red))), 1.0);
  SolveResult res{0, norm <= target, norm};
  if (!res.ok) {
    a = 0.0;
    for (long q = t0; q < n; q += st) {
      z[q] = r[q] / diag[q];
      p[q] = z[q];
      a += r[q] * z[q];
    }
    double rz = block_sum(a, red);
    for (int it = 1; it <= maxit; ++it) {
      a = 0.0;
      const double pole = threadIdx.x < 64 ? pole_row(c, h, false, p, pe) : 0.0;
      for (long q = t0; q < n; q += st) {
        ap[q] = (q == 0 ? pole : apply_at(c, h, false, p, q, rows, pe));
        a += p[q] * ap[q];
      }
      const double pap = block_sum(a, red);
      if (!isfinite(rz) || rz <= DBL_MIN || !isfinite(pap) || pap <= DBL_MIN) break;
      const double alpha = rz / pap;
      a = 0.0;
      for (long q = t0; q < n; q += st) {
        x[q] += alpha * p[q];
        r[q] -= alpha * ap[q];
        a += orig_term(c, r[q], q, ps, dt);
      }
      norm = sqrt(fmax(0.0, block_sum(a, red)));
      res.iters = it;
      res.last = norm;
      if (norm <= target || it % 1024 == 0) {
        __syncthreads();  // x complete
        a = 0.0;
        const double pole = threadIdx.x < 64 ? pole_row(c, h, false, x, pe) : 0.0;
        for (long q = t0; q < n; q += st) {
          r[q] = rhs[q] - (q == 0 ? pole : apply_at(c, h, false, x, q, rows, pe));
          a += orig_term(c, r[q], q, ps, dt);
        }
        norm = sqrt(fmax(0.0, block_sum(a, red)));
        res.last = norm;
        if (norm <= target) {
          res.ok = true;
          break;
        }
        a = 0.0;
        for (long q = t0; q < n; q += st) {
          z[q] = r[q] / diag[q];
          p[q] = z[q];
          a += r[q] * z[q];
        }
        rz = block_sum(a, red);
        continue;
      }
      a = 0.0;
      for (long q = t0; q < n; q += st) {
        z[q] = r[q] / diag[q];
        a += r[q] * z[q];
      }
      const double rzn = block_sum(a, red);
      for (long q = t0; q < n; q += st) p[q] = z[q] + (rzn / rz) * p[q];
      rz = rzn;
      __syncthreads();  // p complete before the next product
    }
  }
  store_result(c, h, res, first);
}

// Right-preconditioned BiCGStab on the tensor system (:1029-1131).
__global__ void __launch_bounds__(kSolveThreads) k_mi_bicgstab(MiK c, double rel_tol,
                                                               int maxit) {
  __shared__ double red[kSolveThreads / 64];
  const int h = blockIdx.x;
  const long n = c.n;
  const long t0 = threadIdx.x, st = blockDim.x;
  const double dt = c.max_colat / (double)(c.nt - 1);
  const double half = 0.5 * dt;
  const double pe = sin(half) * c.pedersen / (incl2(half) * dt * dt);
  const double ps = (double)c.np * pe;
  double *x = c.vec(h, V_X), *r = c.vec(h, V_R), *p = c.vec(h, V_P), *ap = c.vec(h, V_AP),
         *sh = c.vec(h, V_SH), *pp = c.vec(h, V_PP), *s = c.vec(h, V_S), *psv = c.vec(h, V_PS),
         *as = c.vec(h, V_AS);
  const double *rhs = c.vec(h, V_RHS);
  factor_rings(c, h);
  __syncthreads();  // ring factors complete
  double a = 0.0, b = 0.0;
  const double pole = threadIdx.x < 64 ? pole_row(c, h, true, x, pe) : 0.0;
  for (long q = t0; q < n; q += st) {
    r[q] = rhs[q] - (q == 0 ? pole : apply_at(c, h, true, x, q, nullptr, pe));
    sh[q] = r[q];
    p[q] = 0.0;
    ap[q] = 0.0;
    a += orig_term(c, r[q], q, ps, dt);
    b += q == 0 ? 0.0 : orig_term(c, rhs[q], q, ps, dt);
  }
  double norm = sqrt(fmax(0.0, block_sum(a, red)));
  const double first = norm;
  const double target = rel_tol * fmax(sqrt(fmax(0.0, block_sum(b, red))), 1.0);
  SolveResult res{0, norm <= target, norm};
  double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  for (int it = 1; it <= maxit && !res.ok; ++it) {
    a = 0.0;
    for (long q = t0; q < n; q += st) a += sh[q] * r[q];
    const double rho = block_sum(a, red);
    if (!isfinite(rho) || fabs(rho) <= DBL_MIN || !isfinite(omega) || fabs(omega) <= DBL_MIN)
      break;
    const double beta = (rho / rho_prev) * (alpha / omega);
    for (long q = t0; q < n; q += st) p[q] = r[q] + beta * (p[q] - omega * ap[q]);
    precondition_all(c, h, p, pp);
    a = 0.0;
    const double pole = threadIdx.x < 64 ? pole_row(c, h, true, pp, pe) : 0.0;
    for (long q = t0; q < n; q += st) {
      ap[q] = (q == 0 ? pole : apply_at(c, h, true, pp, q, nullptr, pe));
      a += sh[q] * ap[q];
    }
    const double den = block_sum(a, red);
    alpha = rho / den;
    a = 0.0;
    for (long q = t0; q < n; q += st) {
      s[q] = r[q] - alpha * ap[q];
      a += orig_term(c, s[q], q, ps, dt);
    }
    const double snorm = sqrt(fmax(0.0, block_sum(a, red)));
    if (!isfinite(den) || fabs(den) <= DBL_MIN) break;
    if (snorm <= target) {
      res.last = snorm;
      res.iters = it;
      res.ok = true;
      for (long q = t0; q < n; q += st) x[q] += alpha * pp[q];
      break;
    }
    precondition_all(c, h, s, psv);
    a = 0.0;
    b = 0.0;
    const double pole_s = threadIdx.x < 64 ? pole_row(c, h, true, psv, pe) : 0.0;
    for (long q = t0; q < n; q += st) {
      as[q] = (q == 0 ? pole_s : apply_at(c, h, true, psv, q, nullptr, pe));
      a += as[q] * as[q];
      b += as[q] * s[q];
    }
    const double wden = block_sum(a, red);
    const double wnum = block_sum(b, red);
    omega = wnum / wden;
    if (!isfinite(wden) || wden <= DBL_MIN || !isfinite(omega) || fabs(omega) <= DBL_MIN)
      break;
    a = 0.0;
    for (long q = t0; q < n; q += st) {
      x[q] += alpha * pp[q] + omega * psv[q];
      r[q] = s[q] - omega * as[q];
      a += orig_term(c, r[q], q, ps, dt);
    }
    norm = sqrt(fmax(0.0, block_sum(a, red)));
    res.iters = it;
    res.last = norm;
    if (norm <= target || it % 256 == 0) {
      __syncthreads();  // x complete
      a = 0.0;
      const double pole = threadIdx.x < 64 ? pole_row(c, h, true, x, pe) : 0.0;
      for (long q = t0; q < n; q += st) {
        r[q] = rhs[q] - (q == 0 ? pole : apply_at(c, h, true, x, q, nullptr, pe));
        sh[q] = r[q];
        a += orig_term(c, r[q], q, ps, dt);
      }
      norm = sqrt(fmax(0.0, block_sum(a, red)));
      res.last = norm;
      if (norm <= target) {
        res.ok = true;
        break;
      }
      for (long q = t0; q < n; q += st) {
        p[q] = 0.0;
        ap[q] = 0.0;
      }
      rho_prev = alpha = omega = 1.0;
      continue;
    }
    rho_prev = rho;
  }
  store_result(c, h, res, first);
}

// Cooperative BiCGStab: each hemisphere's solve spread over kCoopBlocks
// workgroups (the same iterations as k_mi_bicgstab, with grid-wide barriers
// where one workgroup's barrier sufficed there). Dot products sum the
// workgroups' partial sums in workgroup order from two alternating slot
// banks (a bank is rewritten only after every workgroup passed the barrier
// that follows its last read), so every workgroup of a hemisphere computes
// the same sums and takes the same branches.
constexpr int kCoopBlocks = 16;

// Both hemispheres in one cooperative launch: workgroups [0, kCoopBlocks)
// solve the north, [kCoopBlocks, 2 kCoopBlocks) the south, concurrently. The
// grid barriers are shared, so every workgroup must pass the same sequence
// of barriers whatever its hemisphere does: each iteration runs a fixed
// schedule of ten barrier-carrying steps, and a hemisphere that has
// converged, broken down or needs no true-residual check this iteration
// skips the work of a step but not its barrier. No branch contains a
// barrier, the loop bound and its exit test are grid-uniform (the exit reads
// both hemispheres' flags after a barrier), so every workgroup leaves after
// the same number of barriers, at the latest after maximum_iterations.
// Per hemisphere the arithmetic is k_mi_bicgstab's.
__global__ void __launch_bounds__(kSolveThreads) k_mi_bicgstab_pair(MiK c, double rel_tol,
                                                                    int maxit, double *slots,
                                                                    int *flags) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  __shared__ double red[kSolveThreads / 64];
  const int h = blockIdx.x / kCoopBlocks, lb = blockIdx.x % kCoopBlocks;
  const long n = c.n;
  const long t0 = lb * (long)blockDim.x + threadIdx.x;
  const long st = (long)kCoopBlocks * blockDim.x;
  const bool pole_wave = lb == 0 && threadIdx.x < 64;
  int bank = 0;
  // sums of up to two values over this hemisphere's workgroups
  auto gsum2 = [&](double v, double w, double &sv, double &sw) {
    const double bv = block_sum(v, red);
    const double bw = block_sum(w, red);
    double *sl = slots + bank * (4 * kCoopBlocks) + h * (2 * kCoopBlocks);
    if (threadIdx.x == 0) {
      sl[lb] = bv;
      sl[kCoopBlocks + lb] = bw;
    }
    bank ^= 1;
    grid.sync();
    sv = 0.0;
    sw = 0.0;
    for (int i = 0; i < kCoopBlocks; ++i) {
      sv += sl[i];
      sw += sl[kCoopBlocks + i];
    }
  };
  auto gsum = [&](double v) {
    double sv, sw;
    gsum2(v, 0.0, sv, sw);
    return sv;
  };
  const double dt = c.max_colat / (double)(c.nt - 1);
  const double half = 0.5 * dt;
  const double pe = sin(half) * c.pedersen / (incl2(half) * dt * dt);
  const double ps = (double)c.np * pe;
  double *x = c.vec(h, V_X), *r = c.vec(h, V_R), *p = c.vec(h, V_P), *ap = c.vec(h, V_AP),
         *sh = c.vec(h, V_SH), *pp = c.vec(h, V_PP), *s = c.vec(h, V_S), *psv = c.vec(h, V_PS),
         *as = c.vec(h, V_AS);
  const double *rhs = c.vec(h, V_RHS);
  const int lane = threadIdx.x & 63;
  auto precondition = [&](bool work, const double *in, double *out) {
    grid.sync();  // `in` complete
    if (work) {
      if (t0 == 0) out[0] = in[0] / c.vec(h, V_DIAG)[0];
      for (long ring = t0 >> 6; ring < c.nt - 2; ring += st >> 6)
        ring_apply(c, h, in, out, (int)ring, lane);
    }
    grid.sync();  // `out` complete
  };
  for (long ring = t0; ring < c.nt - 2; ring += st) ring_factor(c, h, (int)ring);
  double a = 0.0, b = 0.0;
  {
    const double pole = pole_wave ? pole_row(c, h, true, x, pe) : 0.0;
    for (long q = t0; q < n; q += st) {
      r[q] = rhs[q] - (q == 0 ? pole : apply_at(c, h, true, x, q, nullptr, pe));
      sh[q] = r[q];
      p[q] = 0.0;
      ap[q] = 0.0;
      a += orig_term(c, r[q], q, ps, dt);
      b += q == 0 ? 0.0 : orig_term(c, rhs[q], q, ps, dt);
    }
  }
  double n0, b0;
  gsum2(a, b, n0, b0);  // also orders the ring factors before their first use
  double norm = sqrt(fmax(0.0, n0));
  const double first = norm;
  const double target = rel_tol * fmax(sqrt(fmax(0.0, b0)), 1.0);
  SolveResult res{0, norm <= target, norm};
  bool live = !res.ok;  // this hemisphere still iterates
  double rho_prev = 1.0, alpha = 1.0, omega = 1.0;
  bool other_live = true;
  for (int it = 1; it <= maxit && (live || other_live); ++it) {
    // 1: rho
    a = 0.0;
    if (live)
      for (long q = t0; q < n; q += st) a += sh[q] * r[q];
    const double rho = gsum(a);
    if (live && (!isfinite(rho) || fabs(rho) <= DBL_MIN || !isfinite(omega) ||
                 fabs(omega) <= DBL_MIN))
      live = false;
    // 2, 3: p and its preconditioned image
    if (live) {
      const double beta = (rho / rho_prev) * (alpha / omega);
      for (long q = t0; q < n; q += st) p[q] = r[q] + beta * (p[q] - omega * ap[q]);
    }
    precondition(live, p, pp);
    // 4: A pp and den
    a = 0.0;
    {
      const double pole = (live && pole_wave) ? pole_row(c, h, true, pp, pe) : 0.0;
      if (live)
        for (long q = t0; q < n; q += st) {
          ap[q] = q == 0 ? pole : apply_at(c, h, true, pp, q, nullptr, pe);
          a += sh[q] * ap[q];
        }
    }
    const double den = gsum(a);
    // 5: s and its norm
    a = 0.0;
    if (live) {
      alpha = rho / den;
      for (long q = t0; q < n; q += st) {
        s[q] = r[q] - alpha * ap[q];
        a += orig_term(c, s[q], q, ps, dt);
      }
    }
    const double snorm = sqrt(fmax(0.0, gsum(a)));
    if (live && (!isfinite(den) || fabs(den) <= DBL_MIN)) live = false;
    if (live && snorm <= target) {
      res.last = snorm;
      res.iters = it;
      res.ok = true;
      live = false;
      for (long q = t0; q < n; q += st) x[q] += alpha * pp[q];
    }
    // 6, 7: preconditioned s
    precondition(live, s, psv);
    // 8: A ps and omega
    a = 0.0;
    b = 0.0;
    {
      const double pole = (live && pole_wave) ? pole_row(c, h, true, psv, pe) : 0.0;
      if (live)
        for (long q = t0; q < n; q += st) {
          as[q] = q == 0 ? pole : apply_at(c, h, true, psv, q, nullptr, pe);
          a += as[q] * as[q];
          b += as[q] * s[q];
        }
    }
    double wden, wnum;
    gsum2(a, b, wden, wnum);
    if (live) {
      omega = wnum / wden;
      if (!isfinite(wden) || wden <= DBL_MIN || !isfinite(omega) || fabs(omega) <= DBL_MIN)
        live = false;
    }
    // 9: x, r and the recursive residual norm
    a = 0.0;
    if (live)
      for (long q = t0; q < n; q += st) {
        x[q] += alpha * pp[q] + omega * psv[q];
        r[q] = s[q] - omega * as[q];
        a += orig_term(c, r[q], q, ps, dt);
      }
    const double rn = gsum(a);
    bool verify = false;
    if (live) {
      norm = sqrt(fmax(0.0, rn));
      res.iters = it;
      res.last = norm;
      verify = norm <= target || it % 256 == 0;
    }
    // 10: true residual when checked; the exit flags of both hemispheres
    a = 0.0;
    if (verify) {
      const double pole = pole_wave ? pole_row(c, h, true, x, pe) : 0.0;
      for (long q = t0; q < n; q += st) {
        r[q] = rhs[q] - (q == 0 ? pole : apply_at(c, h, true, x, q, nullptr, pe));
        sh[q] = r[q];
        a += orig_term(c, r[q], q, ps, dt);
      }
    }
    const double tn = gsum(a);
    if (verify) {
      norm = sqrt(fmax(0.0, tn));
      res.last = norm;
      if (norm <= target) {
        res.ok = true;
        live = false;
      } else {
        for (long q = t0; q < n; q += st) {
          p[q] = 0.0;
          ap[q] = 0.0;
        }
        rho_prev = alpha = omega = 1.0;
      }
    } else if (live) {
      rho_prev = rho;
    }
    // exchange liveness: written now, read after the next iteration's first
    // barrier; one more round trip than needed only delays the exit
    if (lb == 0 && threadIdx.x == 0) flags[(it & 1) * 2 + h] = live ? 1 : 0;
    grid.sync();
    other_live = flags[(it & 1) * 2 + (1 - h)] != 0;
  }
  if (lb == 0) store_result(c, h, res, first);
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

// Network features of one point of its grid (kInChannels floats): the f x f
// fine currents of the point's block / scale in channels 0 .. f*f - 1 (row
// major; f <= 4), SigmaP / 10, SigmaH / 10, colatitude / maximum
// colatitude and hemisphere sign in channels 16 .. 19, zeros elsewhere.
constexpr int kFeatCurrent = 16, kFeatSp = 16, kFeatSh = 17, kFeatColat = 18, kFeatHemi = 19;

GCS_DEV void write_features(float *f, const double *current, int f2, double scale, double sp,
                            double sh, double colat_fraction, int hemi) {
  for (int k = 0; k < unet::kInChannels; ++k) f[k] = 0.f;
  for (int k = 0; k < f2; ++k) f[k] = scale > 0.0 ? (float)(current[k] / scale) : 0.f;
  f[kFeatSp] = (float)(sp / 10.0);
  f[kFeatSh] = (float)(sh / 10.0);
  f[kFeatColat] = (float)colat_fraction;
  f[kFeatHemi] = (float)hemi;
}

// The network runs on the polar grid pooled by f x f (f = nt / H): the
// Krylov iterations remove short-wavelength error quickly, the network
// supplies the long-wavelength part, and a coarse grid lets a 4-level
// U-Net's receptive field span the polar cap. Row of a coarse point, as a
// fraction of the fine grid's colatitude range:
GCS_DEV double coarse_row_fraction(int rc, int f, int nt) {
  return ((double)rc * f + 0.5 * (double)(f - 1)) / (double)(nt - 1);
}

// Bilinear sample of a coarse [Hc][Wc] field at fine point (i, j):
// rows clamped, columns periodic.
GCS_DEV double coarse_sample(const float *o, int Hc, int Wc, int f, int i, int j) {
  double y = ((double)i + 0.5) / f - 0.5, x = ((double)j + 0.5) / f - 0.5;
  y = fmin(fmax(y, 0.0), (double)(Hc - 1));
  const int y0 = min((int)floor(y), Hc - 2 < 0 ? 0 : Hc - 2);
  const double wy = Hc > 1 ? y - y0 : 0.0;
  const int y1 = min(y0 + 1, Hc - 1);
  const double xf = floor(x);
  const double wx = x - xf;
  int x0 = (int)xf % Wc;
  if (x0 < 0) x0 += Wc;
  const int x1 = (x0 + 1) % Wc;
  const double a = (1.0 - wx) * o[y0 * Wc + x0] + wx * o[y0 * Wc + x1];
  const double b = (1.0 - wx) * o[y1 * Wc + x0] + wx * o[y1 * Wc + x1];
  return (1.0 - wy) * a + wy * b;
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

// Features on the coarse grid: the residual's equivalent current of each
// point of the f x f block, one channel each (space to depth).
__global__ void k_warm_features(MiK c, const double *scale, double hall, int f, float *feat) {
  const int h = blockIdx.y;
  const int Hc = c.nt / f, Wc = c.np / f;
  const long q = gtid();
  if (q >= (long)Hc * Wc) return;
  const int rc = (int)(q / Wc), cc = (int)(q % Wc);
  const double dt = c.max_colat / (double)(c.nt - 1);
  double cur[kFeatCurrent];
  for (int a = 0; a < f; ++a)
    for (int b = 0; b < f; ++b) {
      const int i = rc * f + a, j = cc * f + b;
      const double r = (i > 0 && i < c.nt - 1) ? c.vec(h, V_R)[red_at(c.np, i, j)] : 0.0;
      cur[a * f + b] = equivalent_current(r, i, c.nt, dt, c.r2, hemi_sign(h));
    }
  write_features(feat + ((long)h * Hc * Wc + q) * unet::kInChannels, cur, f * f, scale[h],
                 c.pedersen, hall, coarse_row_fraction(rc, f, c.nt), hemi_sign(h));
}

// The network's potential correction d in V_Z, interpolated from the coarse
// grid (out * out_scale * R^2 * scale / SigmaP; the pole unknown takes the
// mean over row 0); zero where the residual was zero.
__global__ void k_warm_correction(MiK c, const double *scale, const float *out,
                                  double out_scale, int f) {
  const int h = blockIdx.y;
  const long q = gtid();
  if (q >= c.n) return;
  double *d = c.vec(h, V_Z);
  const double s = scale[h];
  if (!(s > 0.0) || !isfinite(s)) {
    d[q] = 0.0;
    return;
  }
  const double k = out_scale * c.r2 * s / c.pedersen;
  const int Hc = c.nt / f, Wc = c.np / f;
  const float *o = out + (long)h * Hc * Wc;
  if (q == 0) {
    double mean = 0.0;
    for (int j = 0; j < c.np; ++j) mean += coarse_sample(o, Hc, Wc, f, 0, j);
    d[0] = k * (mean / (double)c.np);
    return;
  }
  const int i = 1 + (int)((q - 1) / c.np), j = (int)((q - 1) % c.np);
  d[q] = k * coarse_sample(o, Hc, Wc, f, i, j);
}

// Step length along the correction in the operator's energy norm:
// alpha = (d . r) / (d . A d) (r the residual of the previous potential, A d
// in V_AS) minimises the A-norm of the guess's error for the symmetric
// Pedersen operator, and weighs the long-wavelength error the Krylov
// iterations reduce slowest more than the residual norm does. One block per
// hemisphere; alpha = 0 when d . A d is not positive or not finite.
__global__ void k_warm_alpha(MiK c, double *alpha) {
  __shared__ double rn[kT], rd[kT];
  const int h = blockIdx.y;
  const double *r = c.vec(h, V_R), *d = c.vec(h, V_Z), *ad = c.vec(h, V_AS);
  double num = 0.0, den = 0.0;
  for (long q = threadIdx.x; q < c.n; q += kT) {
    num += d[q] * r[q];
    den += d[q] * ad[q];
  }
  rn[threadIdx.x] = num;
  rd[threadIdx.x] = den;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      rn[threadIdx.x] += rn[threadIdx.x + w];
      rd[threadIdx.x] += rd[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double a = rn[0] / rd[0];
    alpha[h] = (rd[0] > 0.0 && isfinite(a)) ? a : 0.0;
  }
}

__global__ void k_warm_update(MiK c, const double *alpha) {
  const int h = blockIdx.y;
  const long q = gtid();
  if (q >= c.n) return;
  c.vec(h, V_X)[q] += alpha[h] * c.vec(h, V_Z)[q];
}

// The guess in V_X (previous potential) becomes guess + alpha N(residual):
// the residual, the network's forward pass on both hemispheres, the
// correction d, and the energy-norm step alpha along it (which shrinks a
// correction that does not point along the error). The solve that follows
// recomputes the residual of the new guess.
int warm_start(Problem &P, const MiK &c, bool tensor) {
  MiState &m = *P.mi;
  unet::Net &n = *m.net;
  hipStream_t s = P.stream;
  const dim3 g2(blocks_of(m.n), 2);
  k_mi_apply<<<g2, kT, 0, s>>>(c, tensor ? 1 : 0, V_X, V_AP, m.active);
  k_mi_vec<<<g2, kT, 0, s>>>(c, OP_RESIDUAL, 0.0, m.active);
  k_warm_scale<<<dim3(1, 2), kT, 0, s>>>(c, m.warm_scale);
  const int f = m.nt / n.H;
  k_warm_features<<<dim3(blocks_of((long)n.H * n.W), 2), kT, 0, s>>>(c, m.warm_scale, m.hall, f,
                                                                    n.input);
  hipStream_t keep = n.stream;
  n.stream = s;
  const int rc = unet::forward_device(n, 2);
  n.stream = keep;
  if (rc != 0) return -1;
  k_warm_correction<<<g2, kT, 0, s>>>(c, m.warm_scale, n.out, n.out_scale, f);
  k_mi_apply<<<g2, kT, 0, s>>>(c, tensor ? 1 : 0, V_Z, V_AS, m.active);
  k_warm_alpha<<<dim3(1, 2), kT, 0, s>>>(c, m.warm_scale + 2);
  k_warm_update<<<g2, kT, 0, s>>>(c, m.warm_scale + 2);
  GCS_HIP(hipGetLastError());
  return 0;
}

// The polar solves of both hemispheres; 0 when both converge. Setup (right-
// hand side, guess, stencil), the optional learned correction of the guess,
// then one fused Krylov kernel per solve (both hemispheres in one launch).
int solve(Problem &P, bool tensor) {
  MiState &m = *P.mi;
  MiK c = make_mik(P);
  hipStream_t s = P.stream;
  const auto t0 = std::chrono::steady_clock::now();
  const int act[2] = {1, 1};
  GCS_HIP(hipMemcpyAsync(m.active, act, sizeof(act), hipMemcpyHostToDevice, s));
  const dim3 g2(blocks_of(m.n), 2);
  k_mi_setup<<<g2, kT, 0, s>>>(c, tensor ? 1 : 0, m.hall);
  if (m.net != nullptr && warm_start(P, c, tensor) != 0) return -1;
  const double rel = m.cfg.relative_tolerance;
  const int maxit = m.cfg.maximum_iterations;
  if (tensor) {
    // cooperative form (both hemispheres, 2 kCoopBlocks workgroups) when the
    // device can hold its workgroups at once; one workgroup per hemisphere
    // otherwise
    bool coop = m.coop_slots != nullptr;
    if (coop) {
      double relv = rel;
      int maxv = maxit;
      double *slots = m.coop_slots;
      int *flags = m.coop_flags;
      void *args[] = {&c, &relv, &maxv, &slots, &flags};
      if (hipLaunchCooperativeKernel((const void *)k_mi_bicgstab_pair, dim3(2 * kCoopBlocks),
                                     dim3(kSolveThreads), args, 0, s) != hipSuccess) {
        (void)hipGetLastError();
        coop = false;
      }
    }
    if (!coop) k_mi_bicgstab<<<2, kSolveThreads, 0, s>>>(c, rel, maxit);
  } else
    k_mi_cg<<<2, kSolveThreads, 4 * (size_t)m.nt * sizeof(double), s>>>(c, rel, maxit);
  k_mi_vec<<<dim3(blocks_of((long)m.np * m.nt), 2), kT, 0, s>>>(c, OP_EXPAND, 0.0, m.active);
  GCS_HIP(hipGetLastError());
  GCS_HIP(hipMemcpyAsync(m.host_scal, m.scal, 2 * SL_COUNT * sizeof(double),
                         hipMemcpyDeviceToHost, s));
  GCS_HIP(hipStreamSynchronize(s));
  m.solve_ms = std::chrono::duration<double, std::milli>(
                   std::chrono::steady_clock::now() - t0).count();
  bool ok = true;
  for (int h = 0; h < 2; ++h) {
    const double *sl = m.host_scal + h * SL_COUNT;
    m.iterations[h] = (int)sl[SL_ITER];
    m.final_residual[h] = sl[SL_NORM];
    m.initial_residual[h] = sl[SL_FIRST];
    m.converged[h] = sl[SL_OK] != 0.0 ? 1 : 0;
    ok = ok && sl[SL_OK] != 0.0;
  }
  if (!ok) {
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
      !(cfg.hall_siemens >= 0.0) || mc.azimuth_points > 64 * kRingChunk) {
    set_error("invalid MI configuration (azimuth points must be 8..512)");
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
  GCS_HIP(hipMalloc(&m.shell_current, 18L * L.n * L.n * sizeof(double)));
  GCS_HIP(hipMemset(m.shell_current, 0, 18L * L.n * L.n * sizeof(double)));
  GCS_HIP(hipMalloc(&m.fac, polar * sizeof(double)));
  GCS_HIP(hipMalloc(&m.potential, polar * sizeof(double)));
  GCS_HIP(hipMalloc(&m.drift, 18L * L.ghost * L.NJ * L.NK * sizeof(double)));
  GCS_HIP(hipMalloc(&m.work, 2L * kVectors * m.n * sizeof(double)));
  {
    // the cooperative BiCGStab needs all 2 kCoopBlocks workgroups resident;
    // GCS_MI_COOPERATIVE=0 selects the single-workgroup solve instead (for
    // tools that cannot follow a cooperative launch, such as the kernel
    // tracer of rocprofv3 7.2, which faults at exit after one)
    const char *env = std::getenv("GCS_MI_COOPERATIVE");
    const bool allowed = env == nullptr || std::strcmp(env, "0") != 0;
    int dev = 0, coop = 0, per_cu = 0, cus = 0;
    if (allowed && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mi_bicgstab_pair, kSolveThreads,
                                                     0) == hipSuccess &&
        coop != 0 && per_cu * cus >= 2 * kCoopBlocks) {
      GCS_HIP(hipMalloc(&m.coop_slots, 8 * kCoopBlocks * sizeof(double)));
      GCS_HIP(hipMalloc(&m.coop_flags, 4 * sizeof(int)));
    }
    (void)hipGetLastError();
  }
  GCS_HIP(hipMalloc(&m.scal, 2 * SL_COUNT * sizeof(double)));
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
  (void)hipFree(m->shell_current);
  (void)hipFree(m->fac);
  (void)hipFree(m->potential);
  (void)hipFree(m->drift);
  (void)hipFree(m->work);
  (void)hipFree(m->scal);
  (void)hipFree(m->bfringe);
  (void)hipFree(m->ffringe);
  (void)hipFree(m->samples);
  (void)hipFree(m->active);
  if (m->host_scal) (void)hipHostFree(m->host_scal);
  (void)hipFree(m->warm_scale);
  (void)hipFree(m->coop_slots);
  (void)hipFree(m->coop_flags);
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
  mark_stage(P, S_MI);
  if (mi_sample_current(P) != 0) return -1;
  if (solve(P, m.hall > 0.0) != 0) return -1;
  m.ready = true;
  m.epoch = ts;
  ++m.updates;
  do {
    m.next_update += m.cfg.coupling_interval_s;
  } while (m.next_update <= ts + 1e-10);
  if (refresh_mi_drift(P, m.potential, m.drift) != 0) return -1;
  P.mi_view.potential = m.potential;
  return 1;
}

const double *mi_drift_cache(const Problem &P) {
  return P.mi != nullptr && P.mi->drift != nullptr && P.mi_view.potential != nullptr &&
                 P.mi_view.potential == P.mi->potential
             ? P.mi->drift
             : nullptr;
}

// The field-aligned current of the current state on the polar grids
// (:223-370): shell FAC, its fringe, bilinear samples.
int mi_sample_current(Problem &P) {
  MiState &m = *P.mi;
  MiK c = make_mik(P);
  hipStream_t s = P.stream;
  const Layout &L = P.L;
  k_mi_copy_shell<<<blocks_of(6L * 3 * L.NJ * L.NK), kT, 0, s>>>(c);
  k_mi_fringe_b<<<blocks_of(m.n_bfringe), kT, 0, s>>>(c, m.bfringe, m.n_bfringe);
  k_mi_shell_fac<<<blocks_of(6L * L.n * L.n), kT, 0, s>>>(c, m.shell_current);
  // the copy above used the face-velocity planes, where a fidelity step's
  // update leaves the stress rates
  P.rates_stored = false;
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
  return 0;
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

void mi_info(const Problem &P, MiInfo &out) {
  out = MiInfo();
  if (P.mi == nullptr) return;
  const MiState &m = *P.mi;
  out.ready = m.ready;
  out.updates = m.updates;
  out.epoch_s = m.epoch;
  out.next_update_s = m.next_update;
  out.sample_radius = m.sample_radius;
  out.max_colat = m.max_colat;
  out.potential_norm = m.potential_norm;
  out.potential = m.potential;
  for (int h = 0; h < 2; ++h) {
    out.iterations[h] = m.iterations[h];
    out.initial_residual[h] = m.initial_residual[h];
    out.final_residual[h] = m.final_residual[h];
    out.converged[h] = m.converged[h];
  }
}

void mi_view_enable(Problem &P, bool on) {
  P.mi_view.potential = on && P.mi != nullptr && P.mi->ready ? P.mi->potential : nullptr;
}

int mi_import_potential(Problem &P, const double *values, double epoch_s,
                        double next_update_s) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  MiState &m = *P.mi;
  GCS_HIP(hipStreamSynchronize(P.stream));
  GCS_HIP(hipMemcpy(m.potential, values, 2L * m.nt * m.np * sizeof(double),
                    hipMemcpyHostToDevice));
  m.epoch = epoch_s;
  m.next_update = next_update_s;
  m.ready = true;
  if (refresh_mi_drift(P, m.potential, m.drift) != 0) return -1;
  P.mi_view.potential = m.potential;
  return 0;
}

int mi_shell_current(Problem &P, std::vector<double> &out) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  out.resize(18L * P.L.n * P.L.n);
  GCS_HIP(hipStreamSynchronize(P.stream));
  GCS_HIP(hipMemcpy(out.data(), P.mi->shell_current, out.size() * sizeof(double),
                    hipMemcpyDeviceToHost));
  return 0;
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

// Features and targets on the coarse grid: the fine currents of each f x f
// block as channels, the block mean of the potential as the target.
__global__ void k_gen_features(const GenSample *smp, int nt, int np, int f, double r2,
                               const double *fac, const double *phi, const double *scale,
                               double out_scale, float *feat, float *target) {
  const int b = blockIdx.y;
  const int Hc = nt / f, Wc = np / f;
  const long q = gtid();
  if (q >= (long)Hc * Wc) return;
  const int rc = (int)(q / Wc), cc = (int)(q % Wc);
  const GenSample &g = smp[b];
  const long base = (long)b * nt * np;
  double cur[kFeatCurrent], pot = 0.0;
  for (int a = 0; a < f; ++a)
    for (int e = 0; e < f; ++e) {
      const long at = base + (long)(rc * f + a) * np + cc * f + e;
      cur[a * f + e] = fac[at];
      pot += phi[at];
    }
  pot /= (double)(f * f);
  const long out = (long)b * Hc * Wc + q;
  write_features(feat + out * unet::kInChannels, cur, f * f, scale[b], g.sp, g.sh,
                 coarse_row_fraction(rc, f, nt), g.hemi);
  target[out] = (float)(pot * g.sp / (r2 * scale[b] * out_scale));
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

// The pooling factor f of a network for this MI grid (nt = f H, np = f W),
// or 0 when the grids do not fit.
static int coarsening(const MiState &m, const unet::Net &n) {
  if (n.H <= 0 || n.W <= 0 || m.nt % n.H != 0 || m.np % n.W != 0) return 0;
  const int f = m.nt / n.H;
  return m.np / n.W == f && f * f <= kFeatCurrent ? f : 0;
}

int mi_attach_warm_start(Problem &P, unet::Net *net) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  MiState &m = *P.mi;
  if (net != nullptr && (coarsening(m, *net) <= 0 || net->max_batch < 2)) {
    set_error("warm-start network grid must be the MI grid (colatitude x azimuth) pooled "
              "by one factor in both directions, and its batch at least 2");
    return -1;
  }
  if (m.warm_scale == nullptr) GCS_HIP(hipMalloc(&m.warm_scale, 4 * sizeof(double)));
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
  if (coarsening(m, n) <= 0 || unet::check_batch(n, t.batch) != 0 || t.steps <= 0 ||
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
  const int f = coarsening(m, n);
  const dim3 g(blocks_of((long)n.H * n.W), t.batch);
  int rc = 0;
  for (long step = 0; step < t.steps && rc == 0; ++step) {
    // the previous step's copy of the pinned samples must finish before
    // they are redrawn
    if (step > 0 && loss_history == nullptr && hipStreamSynchronize(s) != hipSuccess) {
      set_error("warm-start training: stream error");
      rc = -1;
      break;
    }
    draw_samples(rng, t.batch, m.max_colat, t.pedersen_min, t.pedersen_max, t.hall_min,
                 t.hall_max, pb.host);
    if (make_pairs(pb, m.nt, m.np, m.max_colat, r2, s) != 0) {
      rc = -1;
      break;
    }
    if (n.adam_t == 0 && step == 0) {
      // output scale: unit RMS of the first batch's targets
      k_gen_features<<<g, kT, 0, s>>>(pb.dev, m.nt, m.np, f, r2, pb.fac, pb.phi, pb.scale,
                                      1.0, n.input, n.target);
      std::vector<float> tg((size_t)t.batch * n.H * n.W);
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
    k_gen_features<<<g, kT, 0, s>>>(pb.dev, m.nt, m.np, f, r2, pb.fac, pb.phi, pb.scale,
                                    n.out_scale, n.input, n.target);
    // cosine decay to 5% of the rate after a short linear warmup
    const double warm = std::min(1.0, (double)(step + 1) / 50.0);
    const double frac = (double)step / (double)std::max(1L, t.steps - 1);
    const double lr = t.learning_rate * warm * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(kPi * frac)));
    if (unet::forward_device(n, t.batch) != 0 || unet::loss_relative_device(n, t.batch) != 0 ||
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

// Training on given pairs (for instance currents logged from MHD runs and
// the potentials the plain solve gives for them): the same loss, Adam
// steps and learning-rate schedule as mi_train_warm_start, each batch drawn
// uniformly (seeded) from the `count` pairs, which stay on the device. The
// conductances are the problem's (Pedersen of its MI configuration, its
// Hall conductance).
int mi_train_warm_start_on(Problem &P, unet::Net &n, const gcs_warm_train_config &t,
                           const double *fac, const double *phi, const int *hemi, int count,
                           float *loss_history) {
  if (P.mi == nullptr) {
    set_error("MI is not enabled");
    return -1;
  }
  const MiState &m = *P.mi;
  if (coarsening(m, n) <= 0 || unet::check_batch(n, t.batch) != 0 || t.steps <= 0 ||
      !(t.learning_rate > 0.0) || count <= 0 || fac == nullptr || phi == nullptr ||
      hemi == nullptr) {
    set_error("invalid warm-start training configuration or data");
    return -1;
  }
  for (int b = 0; b < count; ++b)
    if (hemi[b] != GAMERA_MI_NORTH && hemi[b] != GAMERA_MI_SOUTH) {
      set_error("hemisphere must be GAMERA_MI_NORTH or GAMERA_MI_SOUTH");
      return -1;
    }
  const double rm = m.cfg.ionosphere_radius_re * m.cfg.length_norm_m;
  const double r2 = rm * rm;
  const size_t per = (size_t)m.nt * m.np;
  hipStream_t s = n.stream;
  PairBatch pb;
  if (pb.allocate(t.batch, m.nt, m.np) != 0) return -1;
  double *data_fac = nullptr, *data_phi = nullptr;
  float *host_loss = nullptr;
  int rc = 0;
  if (hipMalloc(&data_fac, count * per * sizeof(double)) != hipSuccess ||
      hipMalloc(&data_phi, count * per * sizeof(double)) != hipSuccess ||
      hipHostMalloc(&host_loss, sizeof(float)) != hipSuccess ||
      hipMemcpy(data_fac, fac, count * per * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(data_phi, phi, count * per * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess) {
    set_error("warm-start training: device buffers for the pairs");
    rc = -1;
  }
  std::mt19937_64 rng(t.seed);
  std::uniform_int_distribution<int> pick(0, count - 1);
  const int f = coarsening(m, n);
  const dim3 g(blocks_of((long)n.H * n.W), t.batch);
  for (long step = 0; step < t.steps && rc == 0; ++step) {
    // the pinned samples of the previous step must be copied before reuse
    if (hipStreamSynchronize(s) != hipSuccess) {
      set_error("warm-start training: stream error");
      rc = -1;
      break;
    }
    for (int b = 0; b < t.batch && rc == 0; ++b) {
      const int idx = pick(rng);
      GenSample &gs = pb.host[b];
      gs.hemi = hemi[idx];
      gs.sp = m.cfg.pedersen_siemens;
      gs.sh = m.hall;
      if (hipMemcpyAsync(pb.fac + b * per, data_fac + idx * per, per * sizeof(double),
                         hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(pb.phi + b * per, data_phi + idx * per, per * sizeof(double),
                         hipMemcpyDeviceToDevice, s) != hipSuccess) {
        set_error("warm-start training: batch gather");
        rc = -1;
      }
    }
    if (rc != 0) break;
    if (hipMemcpyAsync(pb.dev, pb.host, t.batch * sizeof(GenSample), hipMemcpyHostToDevice,
                       s) != hipSuccess) {
      set_error("warm-start training: sample upload");
      rc = -1;
      break;
    }
    k_gen_scale<<<t.batch, kT, 0, s>>>(m.nt, m.np, pb.fac, pb.scale);
    if (n.adam_t == 0 && step == 0) {
      // output scale: unit RMS of the first batch's targets
      k_gen_features<<<g, kT, 0, s>>>(pb.dev, m.nt, m.np, f, r2, pb.fac, pb.phi, pb.scale,
                                      1.0, n.input, n.target);
      std::vector<float> tg((size_t)t.batch * n.H * n.W);
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
    k_gen_features<<<g, kT, 0, s>>>(pb.dev, m.nt, m.np, f, r2, pb.fac, pb.phi, pb.scale,
                                    n.out_scale, n.input, n.target);
    const double warm = std::min(1.0, (double)(step + 1) / 50.0);
    const double frac = (double)step / (double)std::max(1L, t.steps - 1);
    const double lr = t.learning_rate * warm * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(kPi * frac)));
    if (unet::forward_device(n, t.batch) != 0 || unet::loss_relative_device(n, t.batch) != 0 ||
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
  if (hipStreamSynchronize(s) != hipSuccess && rc == 0) {
    set_error("warm-start training: stream error");
    rc = -1;
  }
  if (host_loss) (void)hipHostFree(host_loss);
  (void)hipFree(data_fac);
  (void)hipFree(data_phi);
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
  // a later step reads this potential
  if (refresh_mi_drift(P, m.potential, m.drift) != 0) return -1;
  iterations[0] = m.iterations[0];
  iterations[1] = m.iterations[1];
  *ms = m.solve_ms;
  if (rc != 0) return -1;
  GCS_HIP(hipMemcpy(potential, m.potential, bytes, hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace gcs
