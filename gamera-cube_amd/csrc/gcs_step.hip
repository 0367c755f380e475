// The six-face time step on the device.
//
// One step is the reference's gamera_cs_advance_six_with_background_planned
// (cubedsphere/cubedsphere_step.c:13-154) with the Earth callbacks
// (cubedsphere_earth_boundary.c:212-231, :259-266, :446-451), driven by the
// serial driver's dt rule (test_cubedsphere_dipole_background.c:593-699).
// Its stages run as a fixed sequence of kernels on one stream; dt, the
// predictor ratio and the clock live in device memory, so the sequence can be
// captured once as a graph and replayed.
//
// Differences in structure from the reference, none of which changes a
// computed value:
//  * the two time levels of U, cell B and face flux swap roles every step
//    instead of copying current into old (nonorthogonal_step.c:361-377);
//  * the update fuses stress divergence, Reynolds update, CT, B recovery and
//    the Boris correction per active cell; a face shared by two cells gets
//    the same CT value from both, and only one writes it;
//  * after the final seam-face average only cells touching a seam face are
//    recovered again: every other cell's six fluxes are unchanged;
//  * halo stencils carry only their 2 x 2 nonzero-weight donors, summed in
//    the reference's p-major / q-minor order (the other 60 terms are exact
//    zeros for finite donors, cubedsphere_halo_fastpath.h:1-24);
//  * the i- and j-face fluxes and the j- and i-edge EMFs stage their
//    stencils in LDS tiles (same arithmetic per face or edge);
//  * the next step's CFL dt is computed in the update and the seam
//    recovery from the values they store, and min-folded (exact);
//  * with fidelity 0 the work that only reproduces ghost values the step
//    never reads (ghost-cell B, the old level's unchanged ghost face
//    fluxes, post-update ghost copies) is skipped.
#include <hip/hip_runtime.h>

#include <cstring>

#include <algorithm>
#include <mutex>
#include <string>
#include <cfloat>
#include <cmath>
#include <type_traits>

#include "gcs_device.hpp"
#include "gcs_math.hpp"
#include "gcs_math_mixed.hpp"

namespace gcs {

namespace {



struct K {
  Layout L;
  double *pl;
  float *fl;  // FP32 planes of the mixed form (null until first used)
  uint8_t *by;
  StepScalars *sc;
  int cur, old;  // time-level slots of this step
  StepOptions o;
  MiView mi;
  __device__ double *p(int q) const { return pl + (long)q * L.total; }
  __device__ float *sp(int q) const { return fl + (long)q * L.total; }
  __device__ const uint8_t *b(int q) const { return by + (long)q * L.total; }
  __device__ void fail(unsigned stage, long index) const {
    if (atomicOr(&sc->failed, 1u) == 0u) {
      sc->fail_stage = stage;
      sc->fail_index = (int)index;
    }
  }
};

// (f, i, j, k) of thread t in the box lo <= (i, j, k) < hi on every face.
// Launch index -> (face, i, j, k) of a box. Every box here has fewer than
// 2^31 points over the six faces, so the decomposition runs in 32-bit
// unsigned arithmetic (64-bit division is a long software sequence).
GCS_DEV bool box(long t, const int lo[3], const int hi[3], int &f, int &i,
                 int &j, int &k) {
  const unsigned e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
  const unsigned e12 = e1 * e2;
  const unsigned per = (unsigned)(hi[0] - lo[0]) * e12;
  if (t >= 6L * per) return false;
  const unsigned u = (unsigned)t;
  const unsigned fu = u / per;
  unsigned r = u - fu * per;
  const unsigned iu = r / e12;
  r -= iu * e12;
  const unsigned ju = r / e2;
  f = (int)fu;
  i = (int)iu + lo[0];
  j = (int)ju + lo[1];
  k = (int)(r - ju * e2) + lo[2];
  return true;
}

// Launch index -> a cell outside the active box (radial shells of `ghost`
// cells at both ends, then the tangential frame of every active shell).
GCS_DEV bool ghost_cell(long t, const Layout &L, int &f, int &i, int &j, int &k) {
  const unsigned g = L.ghost, NJ = L.NJ, NK = L.NK, n = L.n;
  const unsigned plane = NJ * NK;
  const unsigned shell = 2 * g * plane;
  const unsigned frame = plane - n * n;
  const unsigned per = shell + (unsigned)L.nr * frame;
  if (t >= 6L * per) return false;
  const unsigned u = (unsigned)t;
  const unsigned fu = u / per;
  unsigned r = u - fu * per;
  f = (int)fu;
  if (r < shell) {
    const unsigned s = r / plane;
    r -= s * plane;
    i = (int)(s < g ? s : L.nr + s);
    j = (int)(r / NK);
    k = (int)(r - (unsigned)j * NK);
    return true;
  }
  r -= shell;
  const unsigned si = r / frame;
  r -= si * frame;
  i = (int)(g + si);
  const unsigned rows = 2 * g * NK;  // full rows j < g and j >= g + n
  if (r < rows) {
    const unsigned row = r / NK;
    j = (int)(row < g ? row : n + row);
    k = (int)(r - row * NK);
  } else {
    r -= rows;
    const unsigned mj = r / (2 * g), col = r - mj * (2 * g);
    j = (int)(g + mj);
    k = (int)(col < g ? col : n + col);
  }
  return true;
}

GCS_DEV long tid() { return blockIdx.x * (long)blockDim.x + threadIdx.x; }

// Launch index with an XCD-aware block order: workgroups are dispatched
// round-robin over the 8 XCDs, each with its own L2, so block b runs on XCD
// b % 8; giving each XCD one contiguous range of logical blocks keeps the
// stencil neighbours of a block in the same L2. (A bijection of [0, nb).)
// Used by the face-velocity and edge-EMF sweeps, whose 8-point stencils
// reach across k-rows; plan-driven kernels keep the plain order.
GCS_DEV unsigned block_xcd() {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned q = nb >> 3, r = nb & 7, x = b & 7, slot = b >> 3;
  return x * q + min(x, r) + slot;
}
GCS_DEV long tid_xcd() { return block_xcd() * (long)blockDim.x + threadIdx.x; }

constexpr int kCflThreads = 256;

// ---- dt rule ---------------------------------------------------------------

// Local CFL dt of one active cell from its conserved state u and residual
// cell B b (nonorthogonal_step.c:449-525); DBL_MAX, with the failure
// recorded, when the state is invalid.
// b0: the cell's background B; len: the minimum of its three CFL lengths.
GCS_DEV double cfl_dt_with(const K &c, long at, const double u[5], const double b[3],
                           const double b0[3], double len) {
  Prim w;
  if (!cons_to_prim(u, c.o.ts_gamma, c.o.ts_rho_floor, c.o.ts_p_floor, w)) {
    c.fail(1, at);
    return DBL_MAX;
  }
  const double cs = sqrt(c.o.ts_gamma * w.p / w.rho);
  double m[3];
  for (int d = 0; d < 3; ++d) m[d] = b[d] + b0[d];
  double va = sqrt(dot3(m, m)) / sqrt(w.rho);
  if (c.o.ts_boris) va = c.o.ts_c * va / hypot(c.o.ts_c, va);
  const double speed = sqrt(dot3(w.v, w.v)) + hypot(cs, va);
  const double cand = c.o.ts_cfl * len / speed;
  if (!isfinite(cand) || cand <= 0.0) {
    c.fail(2, at);
    return DBL_MAX;
  }
  return cand;
}

GCS_DEV double cfl_min_length(const K &c, long at) {
  double len = c.p(P_CFL)[at];
  len = fmin(len, c.p(P_CFL + 1)[at]);
  len = fmin(len, c.p(P_CFL + 2)[at]);
  return len;
}

GCS_DEV double cfl_dt(const K &c, long at, const double u[5], const double b[3]) {
  const double b0[3] = {c.p(P_B0C)[at], c.p(P_B0C + 1)[at], c.p(P_B0C + 2)[at]};
  return cfl_dt_with(c, at, u, b, b0, cfl_min_length(c, at));
}

// Minimum of v over the wavefront, folded by lane 0 into one of the slots.
// Every lane of the wave must call it.
GCS_DEV void cfl_fold(double v, unsigned long long *slots) {
  unsigned long long bits = __double_as_longlong(v);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long other = __shfl_down(bits, off, 64);
    bits = other < bits ? other : bits;
  }
  if ((threadIdx.x & 63) == 0) atomicMin(&slots[(tid() >> 6) % kCflSlots], bits);
}

__global__ void k_cfl_reset(unsigned long long *slots) {
  slots[threadIdx.x] = __double_as_longlong(DBL_MAX);
}

// The CFL dt of every active cell from the current level, min-reduced to
// one partial per block: the first step after creation or an upload (later
// steps take the minimum the previous step's update folded into the slots).
__global__ void __launch_bounds__(kCflThreads) k_cfl(K c, unsigned long long *partial) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr, L.ghost + L.n, L.ghost + L.n};
  int f, i, j, k;
  double best = DBL_MAX;
  if (box(tid(), lo, hi, f, i, j, k)) {
    const long at = L.at(f, i, j, k);
    double u[5], b[3];
    for (int q = 0; q < 5; ++q) u[q] = c.p(plane_u(c.cur, q))[at];
    for (int d = 0; d < 3; ++d) b[d] = c.p(plane_b(c.cur, d))[at];
    best = cfl_dt(c, at, u, b);
  }
  // wavefront min, block min through LDS, one partial per block
  unsigned long long bits = __double_as_longlong(best);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long other = __shfl_down(bits, off, 64);
    bits = other < bits ? other : bits;
  }
  __shared__ unsigned long long wave_min[kCflThreads / 64];
  if ((threadIdx.x & 63) == 0) wave_min[threadIdx.x >> 6] = bits;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = wave_min[0];
    for (int w = 1; w < kCflThreads / 64; ++w) m = wave_min[w] < m ? wave_min[w] : m;
    partial[blockIdx.x] = m;
  }
}

// Driver dt rule (test_cubedsphere_dipole_background.c:593-616, :669-670),
// after the min over the CFL kernel's per-block partials (one block).
__global__ void __launch_bounds__(kCflThreads) k_dt(K c, double fraction,
                                                    const unsigned long long *partial,
                                                    int nblock) {
  unsigned long long m = __double_as_longlong(DBL_MAX);
  for (int b = threadIdx.x; b < nblock; b += kCflThreads) m = partial[b] < m ? partial[b] : m;
  __shared__ unsigned long long red[kCflThreads];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kCflThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w && red[threadIdx.x + w] < red[threadIdx.x])
      red[threadIdx.x] = red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  StepScalars *s = c.sc;
  const double limit = __longlong_as_double(red[0]);
  if (s->steps == 0) {
    s->dt = fraction * limit;
    s->ratio = 0.0;
    s->previous_time = s->time;
  } else {
    s->dt = fmin(fraction * limit, 1.25 * s->previous_dt);
    s->ratio = 0.5 * s->dt / s->previous_dt;
    s->previous_time = s->time - s->previous_dt;
  }
  // The driver sets the Earth context's step clock to the step's own
  // (test_cubedsphere_dipole_background.c:677-678).
  s->wall_time = s->time;
  s->wall_dt = s->dt;
}

__global__ void k_clock(K c) {
  StepScalars *s = c.sc;
  s->time += s->dt;
  s->previous_dt = s->dt;
  s->steps += 1;
}

// ---- seams -----------------------------------------------------------------

// cubedsphere_interface.c:7-81 for the face flux of the given levels.
__global__ void k_sync_phi(K c, const SeamPair *pairs, int n, int both) {
  const long t = tid();
  if (t >= n) return;
  const SeamPair sp = pairs[t];
  for (int lv = 0; lv < (both ? 2 : 1); ++lv) {
    const int level = lv == 0 ? c.cur : c.old;
    double *a = c.p(plane_phi(level, sp.dir_a)) + sp.a;
    double *b = c.p(plane_phi(level, sp.dir_b)) + sp.b;
    const double common = 0.5 * (*a + sp.sign * *b);
    *a = common;
    *b = sp.sign * common;
  }
}

// Face-flux planes: FP64 (T = double) or the mixed form's FP32 planes.
template <typename T>
GCS_DEV T *fluid_plane(const K &c, int d, int q);
template <>
GCS_DEV double *fluid_plane<double>(const K &c, int d, int q) {
  return c.p(P_FLUID + 5 * d + q);
}
template <>
GCS_DEV float *fluid_plane<float>(const K &c, int d, int q) {
  return c.sp(M_FLUID + 5 * d + q);
}
template <typename T>
GCS_DEV T *maxwell_plane(const K &c, int d, int m);
template <>
GCS_DEV double *maxwell_plane<double>(const K &c, int d, int m) {
  return c.p(P_MAXWELL + 3 * d + m);
}
template <>
GCS_DEV float *maxwell_plane<float>(const K &c, int d, int m) {
  return c.sp(M_MAXWELL + 3 * d + m);
}

// Same reconciliation of the fluid (5) and Maxwell (3) face fluxes.
template <typename T>
__global__ void k_sync_flux(K c, const SeamPair *pairs, int n) {
  const long t = tid();
  if (t >= n) return;
  const SeamPair sp = pairs[t];
  for (int q = 0; q < 8; ++q) {
    T *a = q < 5 ? fluid_plane<T>(c, sp.dir_a, q) + sp.a
                 : maxwell_plane<T>(c, sp.dir_a, q - 5) + sp.a;
    T *b = q < 5 ? fluid_plane<T>(c, sp.dir_b, q) + sp.b
                 : maxwell_plane<T>(c, sp.dir_b, q - 5) + sp.b;
    const T sign = (T)sp.sign;
    const T common = T(0.5) * (*a + sign * *b);
    *a = common;
    *b = sign * common;
  }
}

// cubedsphere_interface.c:217-242: signed mean over each coincident group.
__global__ void k_sync_edges(K c, const int *groups, const EdgeMember *m,
                             int n) {
  const long g = tid();
  if (g >= n) return;
  const int a = groups[g], b = groups[g + 1];
  double common = 0.0;
  for (int t = a; t < b; ++t) common += m[t].sign * c.p(P_EMF + m[t].dir)[m[t].index];
  common /= (double)(b - a);
  for (int t = a; t < b; ++t) c.p(P_EMF + m[t].dir)[m[t].index] = m[t].sign * common;
}

// ---- ghost fill --------------------------------------------------------------

// cubedsphere_halo.c:103-115
GCS_DEV bool halo_valid(const double u[5]) {
  // Evaluated without early exits: the flags are combined with '&' so that
  // the caller's selection compiles to straight-line code (a short-circuit
  // form here was structurized incorrectly by the ROCm 7.2 compiler, which
  // then ignored the old-level test).
  const double rho = u[0];
  const double m2 = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
  const bool finite = isfinite(rho) & isfinite(u[1]) & isfinite(u[2]) &
                      isfinite(u[3]) & isfinite(u[4]);
  return finite & (rho > 0.0) & (u[4] > 0.5 * m2 / rho);
}

// 2 x 2 donor block sum in the reference's p-major / q-minor order.
GCS_DEV double block_sum(const double *plane, int donor, int sk,
                         const double w[4]) {
  double s = 0.0;
  s += w[0] * plane[donor];
  s += w[1] * plane[donor + 1];
  s += w[2] * plane[donor + sk];
  s += w[3] * plane[donor + sk + 1];
  return s;
}

// cubedsphere_halo.c:582-632: tangential ghost cells, both levels.
// Ghost cell B is written for the storage only: no step kernel reads the
// cell B of a ghost cell (the predictor recovers B* from face fluxes, halo
// donors and the MI shell are active cells), so without `with_b` (the
// problem API's fidelity 0) only U is filled.
__global__ void k_halo_cells(K c, const HaloCell *e, int n, int with_b) {
  const long t = tid();
  if (t >= n) return;
  const HaloCell h = e[t];
  const int sk = c.L.SK;
  double cu[5], ou[5];
  for (int q = 0; q < 5; ++q) {
    cu[q] = block_sum(c.p(plane_u(c.cur, q)), h.donor, sk, h.w);
    ou[q] = block_sum(c.p(plane_u(c.old, q)), h.donor, sk, h.w);
  }
  // Nearest-donor fallback when either interpolated level is non-physical.
  const bool ok = halo_valid(cu) & halo_valid(ou);
  const int src = h.nearest;
  for (int q = 0; q < 5; ++q) {
    const double nc = c.p(plane_u(c.cur, q))[src], no = c.p(plane_u(c.old, q))[src];
    c.p(plane_u(c.cur, q))[h.target] = ok ? cu[q] : nc;
    c.p(plane_u(c.old, q))[h.target] = ok ? ou[q] : no;
  }
  if (!with_b) return;
  for (int d = 0; d < 3; ++d) {
    const double cb = block_sum(c.p(plane_b(c.cur, d)), h.donor, sk, h.w);
    const double ob = block_sum(c.p(plane_b(c.old, d)), h.donor, sk, h.w);
    const double nc = c.p(plane_b(c.cur, d))[src], no = c.p(plane_b(c.old, d))[src];
    c.p(plane_b(c.cur, d))[h.target] = ok ? cb : nc;
    c.p(plane_b(c.old, d))[h.target] = ok ? ob : no;
  }
}

// cubedsphere_halo.c:651-686: ghost face flux = interpolated donor B . area.
// The donors are active cells. A level's active B does not change while it
// is the old level, so its ghost fluxes computed a step earlier, when it was
// the current level, are still exact once that step's ghost phase left the
// active B alone: with `both` = 0 (fidelity 0, Problem::old_halo_current)
// only the current level is filled.
__global__ void k_halo_faces(K c, const HaloFace *e, int n, int both) {
  const long t = tid();
  if (t >= n) return;
  const HaloFace h = e[t];
  const int sk = c.L.SK;
  const int base = P_FACE + FACE_PLANES * h.dir + FO_AV;
  double cf = 0.0, of = 0.0;
  for (int d = 0; d < 3; ++d) {
    const double av = c.p(base + d)[h.target];
    cf += block_sum(c.p(plane_b(c.cur, d)), h.donor, sk, h.w) * av;
    if (both) of += block_sum(c.p(plane_b(c.old, d)), h.donor, sk, h.w) * av;
  }
  c.p(plane_phi(c.cur, h.dir))[h.target] = cf;
  if (both) c.p(plane_phi(c.old, h.dir))[h.target] = of;
}

// cubedsphere_halo.c:217-245 (plain problems without the Earth boundary).
__global__ void k_radial_outflow(K c) {
  const Layout &L = c.L;
  const int lo[3] = {0, 0, 0};
  const int hi[3] = {2 * L.ghost, L.NJ, L.NK};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  const int ti = i < L.ghost ? i : L.ghost + L.nr + (i - L.ghost);
  const int si = i < L.ghost ? L.ghost : L.ghost + L.nr - 1;
  const long t = L.at(f, ti, j, k), s = L.at(f, si, j, k);
  for (int lv = 0; lv < 2; ++lv) {
    const int level = lv == 0 ? c.cur : c.old;
    for (int q = 0; q < 5; ++q) c.p(plane_u(level, q))[t] = c.p(plane_u(level, q))[s];
    for (int d = 0; d < 3; ++d) c.p(plane_b(level, d))[t] = c.p(plane_b(level, d))[s];
  }
}

// ---- solar wind (nonorthogonal/src/solar_wind.c) -------------------------------

GCS_DEV bool wind_ok(const double s[8]) {
  if (!isfinite(s[0]) || !isfinite(s[1]) || s[0] <= 0.0 || s[1] <= 0.0) return false;
  for (int q = 2; q < 8; ++q)
    if (!isfinite(s[q])) return false;
  return fabs(s[2]) > DBL_MIN;
}

// :206-232 (with :173-204)
GCS_DEV bool wind_at(const WindSeries &w, double t, double out[8]) {
  if (!isfinite(t)) return false;
  if (w.count == 1 || t <= w.time[0]) {
    for (int q = 0; q < 8; ++q) out[q] = w.state[q];
    return true;
  }
  if (t >= w.time[w.count - 1]) {
    for (int q = 0; q < 8; ++q) out[q] = w.state[8L * (w.count - 1) + q];
    return true;
  }
  int lo = 0, hi = w.count - 1;
  while (hi - lo > 1) {
    const int mid = lo + (hi - lo) / 2;
    if (w.time[mid] <= t) lo = mid;
    else hi = mid;
  }
  const double *a = w.state + 8L * lo, *b = a + 8;
  if (!w.linear) {
    for (int q = 0; q < 8; ++q) out[q] = a[q];
    return true;
  }
  const double fr = (t - w.time[lo]) / (w.time[lo + 1] - w.time[lo]);
  for (int q = 0; q < 8; ++q) out[q] = a[q] + fr * (b[q] - a[q]);
  return wind_ok(out);
}

// :234-275: planar-front ballistic delay.
GCS_DEV bool wind_sample(const WindSeries &w, const double x[3], double t,
                         double out[8]) {
  const double mt = t + w.time_offset;
  double cur[8];
  if (!wind_at(w, mt, cur)) return false;
  const double c2 = w.by * w.by + w.bz * w.bz;
  const double factor = cur[2] / (1.0 + c2);
  const double fv[3] = {factor, -factor * w.by, -factor * w.bz};
  double s2 = 0.0, pr = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(x[d])) return false;
    s2 += fv[d] * fv[d];
    pr += (x[d] - w.reference[d]) * fv[d];
  }
  if (!(s2 > DBL_MIN) || !isfinite(s2)) return false;
  return wind_at(w, mt - pr / s2, out);
}

// :277-311: upstream-to-tail weight.
GCS_DEV bool wind_weight(const WindSeries &w, const double x[3], double t,
                         double &weight) {
  double s[8];
  if (!wind_sample(w, x, t, s)) return false;
  double r2 = 0.0, v2 = 0.0, cosine = 0.0;
  for (int d = 0; d < 3; ++d) {
    r2 += x[d] * x[d];
    v2 += s[2 + d] * s[2 + d];
    cosine -= x[d] * s[2 + d];
  }
  if (!(r2 > DBL_MIN) || !(v2 > DBL_MIN)) return false;
  cosine /= sqrt(r2 * v2);
  cosine = fmax(-1.0, fmin(1.0, cosine));
  const double angle = acos(cosine);
  const double half_pi = 0.5 * acos(-1.0);
  if (angle <= half_pi) weight = 1.0;
  else if (angle >= 2.0 * half_pi) weight = 0.0;
  else {
    const double s1 = (angle - half_pi) / half_pi;
    weight = exp(1.0 - 1.0 / (1.0 - s1 * s1));
  }
  return true;
}

GCS_DEV bool sample_wind(const WindSeries &w, const double x[3], double t,
                         double s[8], double &weight) {
  return wind_sample(w, x, t, s) && wind_weight(w, x, t, weight);
}

// ---- ionosphere potential (cubedsphere_mi_dipole.c) ------------------------------

// NO/nonorthogonal_mi.c:169-181
GCS_DEV double mi_mapped_colatitude(double theta, double r, double ri) {
  double s = sin(theta) * sqrt(ri / r);
  s = fmax(-1.0, fmin(1.0, s));
  return asin(s);
}

// :442-476: bilinear in (mapped colatitude, longitude); zero beyond the cap.
GCS_DEV bool mi_potential(const MiView &m, const double x[3], double &out) {
  const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (!(r >= m.ri)) return false;
  const int h = x[2] >= 0.0 ? 0 : 1;
  const double tm = acos(fmin(1.0, fabs(x[2]) / r));
  const double ti = mi_mapped_colatitude(tm, r, m.ri);
  if (!isfinite(ti)) return false;
  if (ti >= m.max_colat) {
    out = 0.0;
    return true;
  }
  const double tq = ti / m.max_colat * (double)(m.nt - 1);
  const int t = (int)floor(tq);
  const double wt = tq - (double)t;
  double phi = atan2(x[1], x[0]);
  if (phi < 0.0) phi += 2.0 * kPi;
  const double pq = phi * (double)m.np / (2.0 * kPi);
  const int p = (int)((long)floor(pq) % m.np), q = (p + 1) % m.np;
  const double wp = pq - floor(pq);
  const double *v = m.potential + (long)h * m.nt * m.np;
  const double lower = (1.0 - wp) * v[t * m.np + p] + wp * v[t * m.np + q];
  const double upper = (1.0 - wp) * v[(t + 1) * m.np + p] + wp * v[(t + 1) * m.np + q];
  out = (1.0 - wt) * lower + wt * upper;
  return isfinite(out);
}

// :505-543 with NO/nonorthogonal_mi.c:20-41: E from the potential by
// central differences and the E x B drift of the startup dipole without its
// radial part (v_EB, the same for both time levels of a ghost cell).
GCS_DEV bool mi_drift(const MiView &m, const double x[3], double d[3]) {
  const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (!(r > 0.0)) return false;
  const double eps = 1e-4 * fmax(1.0, r);
  double e[3];
  for (int c = 0; c < 3; ++c) {
    double xp[3] = {x[0], x[1], x[2]}, xm[3] = {x[0], x[1], x[2]};
    xp[c] += eps;
    xm[c] -= eps;
    double pp, pm;
    if (!mi_potential(m, xp, pp) || !mi_potential(m, xm, pm)) return false;
    e[c] = -(pp - pm) / (2.0 * eps * m.length_norm * m.speed_norm * m.b_norm);
  }
  // cubedsphere_earth_startup.c:92-102, nonorthogonal_background.c:422-442
  double b[3];
  {
    double r2 = 0.0, pr = 0.0;
    for (int c = 0; c < 3; ++c) r2 += x[c] * x[c];
    if (!(r2 > DBL_MIN) || !isfinite(r2)) return false;
    const double rr = sqrt(r2);
    const double irc = 1.0 / (r2 * rr);
    const double irf = irc / r2;
    for (int c = 0; c < 3; ++c) pr += m.moment[c] * x[c];
    for (int c = 0; c < 3; ++c) b[c] = (3.0 * pr * irf) * x[c] - irc * m.moment[c];
    const double sc = (r - m.cut_radius) / m.cut_length;
    const double w = sc <= 0.0 ? 1.0 : sc >= 1.0 ? 0.0 : 1.0 - 3.0 * sc * sc + 2.0 * sc * sc * sc;
    for (int c = 0; c < 3; ++c) b[c] *= w;
  }
  double b2 = 0.0;
  for (int c = 0; c < 3; ++c) b2 += b[c] * b[c];
  if (!(b2 > 0.0) || !isfinite(b2)) return false;
  d[0] = (e[1] * b[2] - e[2] * b[1]) / b2;
  d[1] = (e[2] * b[0] - e[0] * b[2]) / b2;
  d[2] = (e[0] * b[1] - e[1] * b[0]) / b2;
  double rad = 0.0;
  for (int c = 0; c < 3; ++c) rad += d[c] * x[c];
  rad /= r * r;
  for (int c = 0; c < 3; ++c) d[c] -= rad * x[c];
  return true;
}

// The ghost velocity 2 v_EB - v from the drift.
GCS_DEV bool mi_ghost_from_drift(const double d[3], const double v[3], double g[3]) {
  for (int c = 0; c < 3; ++c) {
    g[c] = 2.0 * d[c] - v[c];
    if (!isfinite(g[c])) return false;
  }
  return true;
}

GCS_DEV bool mi_ghost_velocity(const MiView &m, const double x[3], const double v[3],
                               double g[3]) {
  double d[3];
  return mi_drift(m, x, d) && mi_ghost_from_drift(d, v, g);
}

// One point query of the potential (what 0) or the MI ghost velocity (what 1)
// for the drop-in's gamera_cs_mi_dipole_potential / _ghost_velocity;
// out = {value or velocity, ok}.
__global__ void k_mi_point(MiView m, int what, double x0, double x1, double x2, double v0,
                           double v1, double v2, double *out) {
  const double x[3] = {x0, x1, x2}, v[3] = {v0, v1, v2};
  double r[3] = {0.0, 0.0, 0.0};
  const bool ok = what == 0 ? mi_potential(m, x, r[0]) : mi_ghost_velocity(m, x, v, r);
  for (int c = 0; c < 3; ++c) out[c] = r[c];
  out[3] = ok ? 1.0 : 0.0;
}

// ---- Earth radial boundary -------------------------------------------------

// cubedsphere_earth_boundary.c:27-112: radial ghost cells of both levels.
__global__ void k_earth_cells(K c, WindSeries w) {
  const Layout &L = c.L;
  const int lo[3] = {0, 0, 0};
  const int hi[3] = {2 * L.ghost, L.NJ, L.NK};
  int f, r, j, k;
  if (!box(tid(), lo, hi, f, r, j, k)) return;
  const bool inner = r < L.ghost;
  const int outer = L.ghost + L.nr;
  const int i = inner ? r : outer + (r - L.ghost);
  const int si = inner ? 2 * L.ghost - 1 - i : outer - 1;
  const long t = L.at(f, i, j, k), s = L.at(f, si, j, k);
  const double x[3] = {c.p(P_CEN)[t], c.p(P_CEN + 1)[t], c.p(P_CEN + 2)[t]};
  const double rad = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double drift[3];
  const bool coupled = inner && c.mi.potential != nullptr;
  bool drift_ok = false;
  if (coupled && c.mi.drift != nullptr) {
    const long n = 6L * L.ghost * L.NJ * L.NK, q = (((long)f * L.ghost + i) * L.NJ + j) * L.NK + k;
    for (int e = 0; e < 3; ++e) drift[e] = c.mi.drift[e * n + q];
    drift_ok = true;  // a failed drift is NaN, which the ghost velocity rejects
  } else if (coupled) {
    drift_ok = mi_drift(c.mi, x, drift);
  }
  for (int lv = 0; lv < 2; ++lv) {
    const int level = lv == 0 ? c.cur : c.old;
    const double time = lv == 0 ? c.sc->time : c.sc->previous_time;
    double u[5];
    for (int q = 0; q < 5; ++q) u[q] = c.p(plane_u(level, q))[s];
    Prim pw;
    if (!cons_to_prim(u, c.o.wall_gamma, c.o.wall_rho_floor, c.o.wall_p_floor, pw) ||
        !(rad > 0.0)) {
      c.fail(11, t);
      return;
    }
    if (inner) {
      if (c.mi.potential != nullptr) {
        // cubedsphere_earth_boundary.c:58-68: the ionospheric drift
        double g[3];
        if (!drift_ok || !mi_ghost_from_drift(drift, pw.v, g)) {
          c.fail(14, t);
          return;
        }
        if (c.o.strict_wall)
          for (int d = 0; d < 3; ++d) g[d] = 0.5 * (g[d] + pw.v[d]);
        for (int d = 0; d < 3; ++d) pw.v[d] = g[d];
      } else if (c.o.strict_wall) {
        pw.v[0] = pw.v[1] = pw.v[2] = 0.0;
      } else {
        double vn = 0.0;
        for (int d = 0; d < 3; ++d) vn += pw.v[d] * x[d] / rad;
        for (int d = 0; d < 3; ++d) pw.v[d] -= 2.0 * vn * x[d] / rad;
      }
    } else {
      double sw[8], wt;
      if (!sample_wind(w, x, time, sw, wt)) {
        c.fail(12, t);
        return;
      }
      double vn = 0.0;
      for (int d = 0; d < 3; ++d) vn += pw.v[d] * x[d] / rad;
      const double cs = sqrt(c.o.wall_gamma * pw.p / pw.rho);
      if (vn < cs)
        for (int d = 0; d < 3; ++d) pw.v[d] += (cs - vn) * x[d] / rad;
      pw.rho = wt * sw[0] + (1.0 - wt) * pw.rho;
      pw.p = wt * sw[1] + (1.0 - wt) * pw.p;
      for (int d = 0; d < 3; ++d) pw.v[d] = wt * sw[2 + d] + (1.0 - wt) * pw.v[d];
    }
    if (!prim_to_cons(pw, c.o.wall_gamma, c.o.wall_rho_floor, c.o.wall_p_floor, u)) {
      c.fail(13, t);
      return;
    }
    for (int q = 0; q < 5; ++q) c.p(plane_u(level, q))[t] = u[q];
  }
}

// cubedsphere_earth_boundary.c:114-176: radial ghost faces, both levels.
// Thread box: radial slot r < ghost is inner face r; otherwise faces beyond
// the outer wall (r - ghost of them).
__global__ void k_earth_faces(K c, int d, WindSeries w) {
  const Layout &L = c.L;
  const int outer = L.ghost + L.nr;
  const int n_out = (L.NI + (d == 0)) - (outer + (d == 0));
  const int lo[3] = {0, 0, 0};
  const int hi[3] = {L.ghost + n_out, L.NJ + (d == 1), L.NK + (d == 2)};
  int f, r, j, k;
  if (!box(tid(), lo, hi, f, r, j, k)) return;
  const bool inner = r < L.ghost;
  const int i = inner ? r : outer + (d == 0) + (r - L.ghost);
  const long t = L.at(f, i, j, k);
  const int fb = P_FACE + FACE_PLANES * d;
  for (int lv = 0; lv < 2; ++lv) {
    const int level = lv == 0 ? c.cur : c.old;
    const double time = lv == 0 ? c.sc->time : c.sc->previous_time;
    double *flux = c.p(plane_phi(level, d));
    if (inner) {
      const long s = L.at(f, L.ghost, j, k);
      const double sa = c.p(fb + FO_AREA)[s];
      if (!(sa > 0.0)) {
        c.fail(14, t);
        return;
      }
      flux[t] = flux[s] * c.p(fb + FO_AREA)[t] / sa;
    } else {
      const double x[3] = {c.p(fb + FO_CEN)[t], c.p(fb + FO_CEN + 1)[t],
                           c.p(fb + FO_CEN + 2)[t]};
      double sw[8], wt;
      if (!sample_wind(w, x, time, sw, wt)) {
        c.fail(15, t);
        return;
      }
      double wf = -c.p(P_B0FLUX + d)[t];
      for (int e = 0; e < 3; ++e) wf += sw[5 + e] * c.p(fb + FO_AV + e)[t];
      const long s = L.at(f, d == 0 ? outer : outer - 1, j, k);
      flux[t] = wt * wf + (1.0 - wt) * flux[s];
    }
  }
}

// ---- recovery ----------------------------------------------------------------

GCS_DEV bool recover_at(const K &c, long at, int phi_level_plane,
                        double out[3]) {
  const Layout &L = c.L;
  const int piv = c.b(Q_PIV)[at];
  if (piv == 0xff) return false;
  const long step[3] = {L.di(), (long)L.SK, 1};
  double flux[6], fc[6][3], cen[3], rec[9];
  for (int d = 0; d < 3; ++d) {
    const double *ph = c.p(phi_level_plane + d);
    flux[2 * d] = ph[at];
    flux[2 * d + 1] = ph[at + step[d]];
    for (int e = 0; e < 3; ++e) {
      const double *fcp = c.p(P_FACE + FACE_PLANES * d + FO_CEN + e);
      fc[2 * d][e] = fcp[at];
      fc[2 * d + 1][e] = fcp[at + step[d]];
    }
  }
  for (int e = 0; e < 3; ++e) cen[e] = c.p(P_CEN + e)[at];
  for (int r = 0; r < 9; ++r) rec[r] = c.p(P_REC + r)[at];
  return recover_cell(flux, fc, cen, c.p(P_VOL)[at], rec, piv, out);
}

// Recover every cell of a level (nonorthogonal_step.c:310-359); with
// inner_copy the cells inside the wall take the first shell's value
// (cubedsphere_earth_boundary.c:235-245).
__global__ void k_recover(K c, int phi_plane, int b_plane, int inner_copy,
                          int ghost_only) {
  const Layout &L = c.L;
  const int lo[3] = {0, 0, 0};
  const int hi[3] = {L.NI, L.NJ, L.NK};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  if (ghost_only && i >= L.ghost && i < L.ghost + L.nr && j >= L.ghost &&
      j < L.ghost + L.n && k >= L.ghost && k < L.ghost + L.n)
    return;
  const long at = L.at(f, i, j, k);
  double b[3];
  if (!recover_at(c, at, phi_plane, b)) {
    c.fail(21, at);
    return;
  }
  if (inner_copy && i < L.ghost && !recover_at(c, L.at(f, L.ghost, j, k), phi_plane, b)) {
    c.fail(22, at);
    return;
  }
  for (int d = 0; d < 3; ++d) c.p(b_plane + d)[at] = b[d];
}

// k_recover over the cells outside the active box only.
__global__ void k_recover_ghosts(K c, int phi_plane, int b_plane, int inner_copy) {
  const Layout &L = c.L;
  int f, i, j, k;
  if (!ghost_cell(tid(), L, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  double b[3];
  if (!recover_at(c, at, phi_plane, b)) {
    c.fail(21, at);
    return;
  }
  if (inner_copy && i < L.ghost && !recover_at(c, L.at(f, L.ghost, j, k), phi_plane, b)) {
    c.fail(22, at);
    return;
  }
  for (int d = 0; d < 3; ++d) c.p(b_plane + d)[at] = b[d];
}

// Static data of the seam cells' recovery and CFL dt, gathered once per
// problem in the order of the seam cell list, field-major (kSeamFields
// arrays of n): the six face centroids, the cell centroid, the recovery
// factors, the volume, the pivot code, the background B and the minimum CFL
// length. Half of the seam cells lie alone on their row (the k seams), so
// reading these from the planes would cost a cache line per value.
constexpr int kSeamFields = 36;
enum { SG_FC = 0, SG_CEN = 18, SG_REC = 21, SG_VOL = 30, SG_PIV = 31, SG_B0 = 32, SG_LEN = 35 };

__global__ void k_seam_geometry(K c, const int *cells, int n, double *geom) {
  const long t = tid();
  if (t >= n) return;
  const Layout &L = c.L;
  const long at = cells[t];
  const long step[3] = {L.di(), (long)L.SK, 1};
  auto put = [&](int field, double v) { geom[(long)field * n + t] = v; };
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int e = 0; e < 3; ++e)
        put(SG_FC + 3 * (2 * d + s) + e, c.p(P_FACE + FACE_PLANES * d + FO_CEN + e)[at + s * step[d]]);
  for (int e = 0; e < 3; ++e) put(SG_CEN + e, c.p(P_CEN + e)[at]);
  for (int r = 0; r < 9; ++r) put(SG_REC + r, c.p(P_REC + r)[at]);
  put(SG_VOL, c.p(P_VOL)[at]);
  put(SG_PIV, (double)c.b(Q_PIV)[at]);
  for (int e = 0; e < 3; ++e) put(SG_B0 + e, c.p(P_B0C + e)[at]);
  put(SG_LEN, cfl_min_length(c, at));
}

// Recover the seam cells (active cells with a seam face) from the face
// fluxes of phi_plane, as recover_at does, with their static data from
// the gathered arrays (from the planes when geom is null: the plain-kernel
// switch, which tests hold the two to). With `cfl`, the cells are the
// current level's seam cells after the final seam sync, and their CFL dt is
// folded into the CFL slots.
__global__ void k_recover_list(K c, const int *cells, const double *geom, int n, int phi_plane,
                               int b_plane, unsigned long long *cfl) {
  const long t = tid();
  double dt_cell = DBL_MAX;
  if (t < n) {
    const long at = cells[t];
    auto get = [&](int field) { return geom[(long)field * n + t]; };
    double b[3];
    bool ok;
    if (geom != nullptr) {
      const Layout &L = c.L;
      const long step[3] = {L.di(), (long)L.SK, 1};
      const int piv = (int)get(SG_PIV);
      double flux[6], fc[6][3], cen[3], rec[9];
      for (int d = 0; d < 3; ++d) {
        const double *ph = c.p(phi_plane + d);
        flux[2 * d] = ph[at];
        flux[2 * d + 1] = ph[at + step[d]];
      }
      for (int q = 0; q < 6; ++q)
        for (int e = 0; e < 3; ++e) fc[q][e] = get(SG_FC + 3 * q + e);
      for (int e = 0; e < 3; ++e) cen[e] = get(SG_CEN + e);
      for (int r = 0; r < 9; ++r) rec[r] = get(SG_REC + r);
      ok = piv != 0xff && recover_cell(flux, fc, cen, get(SG_VOL), rec, piv, b);
    } else {
      ok = recover_at(c, at, phi_plane, b);
    }
    if (!ok) {
      c.fail(23, at);
    } else {
      for (int d = 0; d < 3; ++d) c.p(b_plane + d)[at] = b[d];
      if (cfl != nullptr) {
        double u[5];
        for (int q = 0; q < 5; ++q) u[q] = c.p(plane_u(c.cur, q))[at];
        if (geom != nullptr) {
          const double b0[3] = {get(SG_B0), get(SG_B0 + 1), get(SG_B0 + 2)};
          dt_cell = cfl_dt_with(c, at, u, b, b0, get(SG_LEN));
        } else {
          dt_cell = cfl_dt(c, at, u, b);
        }
      }
    }
  }
  if (cfl != nullptr) cfl_fold(dt_cell, cfl);
}


// ---- predictor -----------------------------------------------------------------

// nonorthogonal_step.c:379-419: primitive AB predictor of every cell,
// face-flux predictor of every face and B of the predicted fluxes. With
// Mixed the caches that only the FP32 kernels read (primitives, vol * v,
// B.n, predicted U and B) go to the FP32 planes instead of the FP64 ones.
// Consecutive blocks run on one XCD, so the neighbour faces a block reads
// (the next row's and plane's) are more often in that XCD's L2 (1.34 ->
// 1.32 ms at config 1; the update is slower that way, 1.43 -> 1.49).
template <bool Mixed>
__global__ void k_predict(K c) {
  const Layout &L = c.L;
  const int lo[3] = {0, 0, 0};
  const int hi[3] = {L.NI + 1, L.NJ + 1, L.NK + 1};
  int f, i, j, k;
  if (!box(tid_xcd(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  const double r = c.sc->ratio;
  const long step[3] = {L.di(), (long)L.SK, 1};
  const bool in[3] = {i < L.NI, j < L.NJ, k < L.NK};
  const bool cell = in[0] && in[1] && in[2];
  // Every load is issued before the first store: the planes share one
  // allocation, so a load after a store could not be moved above it.
  bool face[3];
  double cf[3], of[3], area[3];
  for (int d = 0; d < 3; ++d) {
    face[d] = (d == 0 || in[0]) && (d == 1 || in[1]) && (d == 2 || in[2]);
    if (!face[d]) continue;
    cf[d] = c.p(plane_phi(c.cur, d))[at];
    of[d] = c.p(plane_phi(c.old, d))[at];
    area[d] = c.p(P_FACE + FACE_PLANES * d + FO_AREA)[at];
  }
  double uo[5], uc[5], vol = 0.0, cu[3], ou[3], fc[6][3], cen[3], rec[9];
  int piv = 0xff;
  if (cell) {
    for (int q = 0; q < 5; ++q) {
      uo[q] = c.p(plane_u(c.old, q))[at];
      uc[q] = c.p(plane_u(c.cur, q))[at];
    }
    vol = c.p(P_VOL)[at];
    piv = c.b(Q_PIV)[at];
    for (int d = 0; d < 3; ++d) {
      cu[d] = c.p(plane_phi(c.cur, d))[at + step[d]];
      ou[d] = c.p(plane_phi(c.old, d))[at + step[d]];
      for (int s = 0; s < 2; ++s)
        for (int e = 0; e < 3; ++e)
          fc[2 * d + s][e] = c.p(P_FACE + FACE_PLANES * d + FO_CEN + e)[at + s * step[d]];
    }
    for (int e = 0; e < 3; ++e) cen[e] = c.p(P_CEN + e)[at];
    for (int q = 0; q < 9; ++q) rec[q] = c.p(P_REC + q)[at];
  }
  double pred[3];
  for (int d = 0; d < 3; ++d) {
    if (!face[d]) continue;
    pred[d] = cf[d] + r * (cf[d] - of[d]);
    // The predicted flux itself is storage output only (the step reads its
    // B.n below, and the outer wall recomputes it): written with fidelity.
    if (c.o.rates) c.p(P_PHIPRED + d)[at] = pred[d];
    // Face-normal B of the predicted flux, read by the face fluxes and the
    // edge EMF (both divide this flux by this area).
    const double bn = pred[d] / area[d];
    if (Mixed)
      c.sp(M_BN + d)[at] = (float)bn;
    else
      c.p(P_BN + d)[at] = bn;
  }
  if (!cell) return;
  double up[5];
  const double nan = __longlong_as_double(0x7ff8000000000000ll);
  auto wpred = [&](int q, double v) {
    if (Mixed)
      c.sp(M_WPRED + q)[at] = (float)v;
    else
      c.p(P_WPRED + q)[at] = v;
  };
  if (!predict_cell(uo, uc, r, c.o.upd_gamma, c.o.upd_rho_floor, c.o.upd_p_floor, up)) {
    for (int q = 0; q < 5; ++q) wpred(q, nan);
    c.fail(31, at);
    return;
  }
  // U*: the FP64 step reads only its density (the EMF's edge density); the
  // rest is storage output, written with fidelity.
  for (int q = 0; q < 5; ++q) {
    if (Mixed)
      c.sp(M_UPRED + q)[at] = (float)up[q];
    else if (q == 0 || c.o.rates)
      c.p(P_UPRED + q)[at] = up[q];
  }
  // The sweep's primitive conversion of this cell (nonorthogonal_state.c:60-93
  // with the sweep's floors), once per cell instead of once per stencil use;
  // the face fluxes read it back. NaN marks a failed conversion.
  {
    Prim w;
    const bool ok = cons_to_prim(up, c.o.gamma, c.o.rho_floor, c.o.p_floor, w);
    wpred(0, ok ? w.rho : nan);
    for (int e = 0; e < 3; ++e) wpred(1 + e, ok ? w.v[e] : nan);
    wpred(4, ok ? w.p : nan);
  }
  // Volume-weighted velocity with the EMF's density floor, the summand of the
  // face velocity's central8 (nonorthogonal_sweep.c:223-259).
  {
    const double rho = fmax(up[0], c.o.emf_rho_floor);
    for (int e = 0; e < 3; ++e) {
      const double wv = vol * up[1 + e] / rho;
      if (Mixed)
        c.sp(M_WV + e)[at] = (float)wv;
      else
        c.p(P_WV + e)[at] = wv;
    }
  }
  // B of the predicted fluxes (the lower faces' predictions above, the
  // upper faces' by the same expression).
  double flux[6], b[3];
  for (int d = 0; d < 3; ++d) {
    flux[2 * d] = pred[d];
    flux[2 * d + 1] = cu[d] + r * (cu[d] - ou[d]);
  }
  if (piv == 0xff || !recover_cell(flux, fc, cen, vol, rec, piv, b)) {
    c.fail(32, at);
    return;
  }
  for (int d = 0; d < 3; ++d) {
    if (Mixed)
      c.sp(M_BPRED + d)[at] = (float)b[d];
    else
      c.p(P_BPRED + d)[at] = b[d];
  }
}

// ---- face fluxes -----------------------------------------------------------------

// nonorthogonal_sweep.c:408-548 for one face of direction d bounding the
// active box: 8-cell Up7/PDM of the predicted state, kinetic fluid and
// Maxwell fluxes, HOGS, emergency diffusion, area scaling. cell(v, n) is
// stencil cell n (n - 4 cells along d from the face) of variable v: 0-4 the
// predicted primitives (P_WPRED), 5 the volume, 6-8 the predicted cell B.
template <class Cell>
GCS_DEV void face_flux_face(const K &c, int d, int f, long at, long st, const Cell &cell) {
  const StepOptions &o = c.o;
  // Primitives of the eight stencil cells come from the predictor's cache;
  // one variable at a time keeps a single 8-point column live.
  double q8[8];
  bool converted = true;
  for (int n = 0; n < 8; ++n) {
    q8[n] = cell(0, n);
    converted &= isfinite(q8[n]);
  }
  if (!converted) {
    c.fail(41, at);
    return;
  }
  double vol[8];
  for (int n = 0; n < 8; ++n) vol[n] = cell(5, n);
  const double iv = central8(vol);
  if (!isfinite(iv) || fabs(iv) <= DBL_MIN) {
    c.fail(42, at);
    return;
  }
  const double riv = 1.0 / iv;
  Prim w[2];
  up7_pdm_shared(vol, iv, riv, q8, o.pdm, w[0].rho, w[1].rho);
  for (int e = 0; e < 3; ++e) {
    for (int n = 0; n < 8; ++n) q8[n] = cell(1 + e, n);
    up7_pdm_shared(vol, iv, riv, q8, o.pdm, w[0].v[e], w[1].v[e]);
  }
  for (int n = 0; n < 8; ++n) q8[n] = cell(4, n);
  up7_pdm_shared(vol, iv, riv, q8, o.pdm, w[0].p, w[1].p);
  double bb[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  if (o.mhd) {
    for (int e = 0; e < 3; ++e) {
      double q8[8];
      for (int n = 0; n < 8; ++n) q8[n] = cell(6 + e, n);
      up7_pdm_shared(vol, iv, riv, q8, o.pdm, bb[0][e], bb[1][e]);
    }
  }
  if (!valid_interface(w[0].rho, w[0].v, w[0].p) ||
      !valid_interface(w[1].rho, w[1].v, w[1].p)) {
    c.fail(43, at);
    return;
  }
  const int fb = P_FACE + FACE_PLANES * d;
  const double area = c.p(fb + FO_AREA)[at];
  FaceFrame fr;
  for (int e = 0; e < 3; ++e) {
    fr.n[e] = c.p(fb + FO_N + e)[at];
    fr.t1[e] = c.p(fb + FO_T1 + e)[at];
    fr.t2[e] = c.p(fb + FO_T2 + e)[at];
  }
  FluidFlux fl;
  fluid_flux(w, o.gamma, fr, fl);
  // hydrodynamic steps: a zero Maxwell flux and neither HOGS nor emergency
  // diffusion, which the reference applies inside its MHD branch
  // (nonorthogonal_sweep.c:462-527)
  MaxwellFlux mx = {{0.0, 0.0, 0.0}, 0.0, 0.0};
  if (o.mhd) {
    const double bn = c.p(P_BN + d)[at];
    const double b0[3] = {c.p(P_B0F + 3 * d)[at], c.p(P_B0F + 3 * d + 1)[at],
                          c.p(P_B0F + 3 * d + 2)[at]};
    if (!maxwell_flux(w, bb, bn, fr.n, b0, c.p(P_B0N + d)[at], o.boris, o.c, fl.un, mx)) {
      c.fail(44, at);
      return;
    }
  }
  if (o.mhd && o.hogs) hogs(fl, mx, o.hogs_h, o.hogs_m, o.boris, o.c);
  if (o.mhd && o.hogs && o.boris) {
    // nonorthogonal_flux.c:206-241
    double s2 = 0.0;
    s2 += dot3(w[0].v, w[0].v);
    s2 += dot3(w[1].v, w[1].v);
    if (sqrt(0.5 * s2) >= 1.5 * o.c) {
      // U* of the two cells, by the predictor's own expression (the
      // predicted state is not kept in memory for this rare branch)
      double us[2][5];
      for (int side = 0; side < 2; ++side) {
        const long cell = at - (1 - side) * st;
        double uo[5], uc[5];
        for (int q = 0; q < 5; ++q) {
          uo[q] = c.p(plane_u(c.old, q))[cell];
          uc[q] = c.p(plane_u(c.cur, q))[cell];
        }
        if (!predict_cell(uo, uc, c.sc->ratio, o.upd_gamma, o.upd_rho_floor, o.upd_p_floor,
                          us[side])) {
          c.fail(45, at);
          return;
        }
      }
      for (int q = 0; q < 5; ++q) {
        const double lo_u = us[0][q], up_u = us[1][q];
        if (!isfinite(lo_u) || !isfinite(up_u)) {
          c.fail(45, at);
          return;
        }
        fl.f[q] -= o.c * (up_u - lo_u);
      }
      double t2 = 0.0;
      for (int side = 0; side < 2; ++side)
        for (int e = 0; e < 3; ++e) t2 += w[side].v[e] * w[side].v[e];
      atomicAdd(&c.sc->emergency_count[f], 1ull);
      atomicMax(&c.sc->emergency_speed_bits[f],
                (unsigned long long)__double_as_longlong(sqrt(0.5 * t2)));
    }
  }
  for (int q = 0; q < 5; ++q) c.p(P_FLUID + 5 * d + q)[at] = area * fl.f[q];
  for (int e = 0; e < 3; ++e) c.p(P_MAXWELL + 3 * d + e)[at] = area * mx.m[e];
}

// The face flux with the stencil read from global memory (one face per
// thread, 72 stencil loads per face).
__global__ void __launch_bounds__(128) k_face_flux(K c, int d) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr + (d == 0), L.ghost + L.n + (d == 1),
                     L.ghost + L.n + (d == 2)};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  const long st = d == 0 ? L.di() : d == 1 ? (long)L.SK : 1;
  face_flux_face(c, d, f, at, st, [&](int v, int n) {
    const int plane = v < 5 ? P_WPRED + v : v == 5 ? P_VOL : P_BPRED + v - 6;
    return c.p(plane)[at + (n - 4) * st];
  });
}

// LDS tiles of the face flux of direction D: TD faces along D by TT along a
// transverse axis (k, or j for D = 2) at one position of the third axis.
// The block stages the nine stencil variables of its cells (TD + 7 along D)
// once, so each face reads its 8 x 9 stencil values from LDS instead of
// issuing 72 global loads; faces are ordered so that a wave's faces are
// consecutive in k (coalesced geometry reads and flux writes). The
// arithmetic is face_flux_face's, bit for bit.
template <int D, int TD_, int TT_>
struct FluxTile {
  static constexpr int TD = TD_, TT = TT_, kThreads = TD * TT;
  static constexpr int CD = TD + 7;               // stencil cells along D
  static constexpr int TA = D == 2 ? 1 : 2;       // transverse axis
  static constexpr int XA = D == 0 ? 1 : 0;       // fixed axis
  static constexpr int kVars = 9;
  static constexpr int kSlots = CD * TT;
  // LDS slot of cell cd (along D) at transverse position t: consecutive
  // lanes of a wave read consecutive slots.
  GCS_DEV static int slot(int cd, int t) { return D == 2 ? t * CD + cd : cd * TT + t; }
  static constexpr int kStep = D == 2 ? 1 : TT;   // slot step along D
  // blocks for a layout with nr radial and n angular active cells
  static unsigned blocks(int nr, int n) {
    const int act[3] = {nr, n, n};
    return (unsigned)(6L * act[XA] * ((act[D] + 1 + TD - 1) / TD) * ((act[TA] + TT - 1) / TT));
  }
};

// Four waves per SIMD: unbounded, the kernel takes 129 VGPRs and runs three;
// bounded, it spills 8 registers a lane (0.89 / 0.91 -> 0.81 / 0.83 ms at
// config 1; k_face_flux_rows likewise, 0.97 -> 0.90 ms).
template <int D, int TD_, int TT_>
__global__ void __launch_bounds__(TD_ * TT_, 4) k_face_flux_tiled(K c) {
  using T = FluxTile<D, TD_, TT_>;
  constexpr int TD = T::TD, TT = T::TT, CD = T::CD, TA = T::TA, XA = T::XA;
  constexpr int kSlots = T::kSlots;
  __shared__ double s_cell[T::kVars * kSlots];
  const Layout &L = c.L;
  const int act[3] = {L.nr, L.n, L.n};
  const int nd = act[D] + 1, nt = act[TA], nx = act[XA];
  const int ntd = (nd + TD - 1) / TD, ntt = (nt + TT - 1) / TT;
  // Tiles that share stencil cells (neighbours along D) are consecutive
  // logical blocks, and consecutive logical blocks run on one XCD (one L2).
  unsigned b = block_xcd();
  const int td = (int)(b % (unsigned)ntd);
  b /= (unsigned)ntd;
  const int tt = (int)(b % (unsigned)ntt);
  b /= (unsigned)ntt;
  const int x = (int)(b % (unsigned)nx);
  const int f = (int)(b / (unsigned)nx);
  if (f >= 6) return;
  const int d0 = td * TD, t0 = tt * TT;             // tile origin (active-relative)
  const int nvd = min(TD, nd - d0), nvt = min(TT, nt - t0);
  auto at_of = [&](int ad, int at2) {
    int p[3];
    p[D] = L.ghost + ad;
    p[TA] = L.ghost + at2;
    p[XA] = L.ghost + x;
    return L.at(f, p[0], p[1], p[2]);
  };
  // Stencil cells of the tile's valid faces: along D from d0 - 4 to
  // d0 + nvd + 2 (inside the ghost layers), memory-contiguous index fastest.
  // Every load of the fill is issued before the first LDS write.
  constexpr int kFill = (kSlots + T::kThreads - 1) / T::kThreads;
  double fill[kFill][T::kVars];
  int slot[kFill];
#pragma unroll
  for (int r = 0; r < kFill; ++r) {
    const int q = (int)threadIdx.x + r * T::kThreads;
    const int cd = D == 2 ? q % CD : q / TT;
    const int t = D == 2 ? q / CD : q % TT;
    slot[r] = -1;
    if (q >= kSlots || cd >= nvd + 7 || t >= nvt) continue;
    const long cell = at_of(d0 - 4 + cd, t0 + t);
    slot[r] = T::slot(cd, t);
    for (int v = 0; v < 5; ++v) fill[r][v] = c.p(P_WPRED + v)[cell];
    fill[r][5] = c.p(P_VOL)[cell];
    for (int v = 0; v < 3; ++v) fill[r][6 + v] = c.p(P_BPRED + v)[cell];
  }
#pragma unroll
  for (int r = 0; r < kFill; ++r)
    if (slot[r] >= 0)
      for (int v = 0; v < T::kVars; ++v) s_cell[v * kSlots + slot[r]] = fill[r][v];
  __syncthreads();
  const int a = D == 2 ? (int)threadIdx.x % TD : (int)threadIdx.x / TT;
  const int t = D == 2 ? (int)threadIdx.x / TD : (int)threadIdx.x % TT;
  if (a >= nvd || t >= nvt) return;
  const long at = at_of(d0 + a, t0 + t);
  const long st = D == 0 ? L.di() : D == 1 ? (long)L.SK : 1;
  const double *base = s_cell + T::slot(a, t);
  face_flux_face(c, D, f, at, st,
                 [&](int v, int n) { return base[v * kSlots + n * T::kStep]; });
}

// Tile shapes of the FP64 face flux per direction.
using FluxTileI = FluxTile<0, 8, 16>;
using FluxTileJ = FluxTile<1, 8, 16>;

// The k-faces of the FP64 face flux in LDS rows. A block takes kRowFaces
// consecutive faces of the active box's k-faces in (cube face, i, j, k)
// order: the tail of one (i, j) row, whole rows and the head of another.
// It stages the nine stencil variables of those rows' cells (each row's
// faces plus 7, contiguous in k) in LDS, so each face reads its 8 x 9
// stencil values from LDS instead of issuing 72 global loads, and no lane
// idles on a row's end. A block spans at most kRowMax rows when a row has
// at least 24 faces (n >= 23); smaller problems use k_face_flux. The
// arithmetic is face_flux_face's, bit for bit.
constexpr int kRowFaces = 256, kRowMax = 12;
constexpr int kRowSlots = kRowFaces + 7 * kRowMax;
constexpr int kRowMinFaces = (kRowFaces - 1 + kRowMax - 2) / (kRowMax - 1);

__global__ void __launch_bounds__(kRowFaces, 4) k_face_flux_rows(K c) {
  __shared__ double s_cell[9 * kRowSlots];
  const Layout &L = c.L;
  const int nf = L.n + 1;  // k-faces per row
  const long rows = 6L * L.nr * L.n, total = rows * nf;
  const long q0 = (long)block_xcd() * kRowFaces;
  if (q0 >= total) return;
  const int nq = (int)min((long)kRowFaces, total - q0);  // faces of this block
  const long row0 = q0 / nf;
  const int kf0 = (int)(q0 - row0 * nf);
  const int head = min(nf - kf0, nq);  // faces in the first row
  const int nrows = 1 + (nq - head + nf - 1) / nf;
  // Slots: the first row's head + 7 cells from k-face kf0 - 4 on, then each
  // later row's faces + 7 cells from k-face -4 on; face ql of local row r
  // finds its stencil at slot ql + 7 r.
  const int nslots = nq + 7 * nrows, s1 = head + 7;
  auto row_cell = [&](int r, int ck) {
    const long row = row0 + r;
    const long fi = row / L.n;
    const int jj = (int)(row - fi * L.n);
    const int f = (int)(fi / L.nr), ii = (int)(fi - (long)f * L.nr);
    return L.at(f, L.ghost + ii, L.ghost + jj, L.ghost + ck);
  };
  constexpr int kFill = (kRowSlots + kRowFaces - 1) / kRowFaces;
  double fill[kFill][9];
  int slot[kFill];
#pragma unroll
  for (int m = 0; m < kFill; ++m) {
    const int sl = (int)threadIdx.x + m * kRowFaces;
    slot[m] = -1;
    if (sl >= nslots) continue;
    int r, ck;
    if (sl < s1) {
      r = 0;
      ck = kf0 - 4 + sl;
    } else {
      r = 1 + (sl - s1) / (nf + 7);
      ck = (sl - s1) - (r - 1) * (nf + 7) - 4;
    }
    const long cell = row_cell(r, ck);
    slot[m] = sl;
    for (int v = 0; v < 5; ++v) fill[m][v] = c.p(P_WPRED + v)[cell];
    fill[m][5] = c.p(P_VOL)[cell];
    for (int v = 0; v < 3; ++v) fill[m][6 + v] = c.p(P_BPRED + v)[cell];
  }
#pragma unroll
  for (int m = 0; m < kFill; ++m)
    if (slot[m] >= 0)
      for (int v = 0; v < 9; ++v) s_cell[v * kRowSlots + slot[m]] = fill[m][v];
  __syncthreads();
  const int ql = (int)threadIdx.x;
  if (ql >= nq) return;
  const int r = ql < head ? 0 : 1 + (ql - head) / nf;
  const int kf = r == 0 ? kf0 + ql : (ql - head) % nf;
  const long row = row0 + r;
  const int f = (int)(row / ((long)L.nr * L.n));
  const long at = row_cell(r, kf);
  const double *base = s_cell + ql + 7 * r;
  face_flux_face(c, 2, f, at, 1, [&](int v, int n) { return base[v * kRowSlots + n]; });
}

// ---- fused face velocity + edge EMF, LDS tiles ------------------------------------
//
// For edge directions e = 0 (t1 = j, t2 = k) and e = 1 (t1 = k, t2 = i) a
// block owns a tile of E1 x E2 edges in the (t1, t2) plane at one position
// along e, with the k direction the long one so that every LDS fill row is a
// contiguous run. It stages in LDS the cells the tile's t1-face velocities
// need (vol * m / rho summands and volumes, 4 cells beyond the tile on each
// side), the t1-faces' areas and B.n along t2 and the t2-faces' along t1,
// computes the t1-face velocities once per face (k_face_velocity_i's
// arithmetic), and then runs k_edge_emf's arithmetic per edge from LDS. Every
// value is computed by the same expressions in the same order as the two
// unfused kernels, so the EMF is unchanged bit for bit; what changes is that
// each cell and face value is read from HBM once per tile instead of once
// per stencil use, and the face velocities of these two directions are
// never written to HBM.
template <int E>
struct EmfTile {
  static constexpr int T1 = (E + 1) % 3, T2 = (E + 2) % 3;
  // Edges along t1 and t2: 33 along k by 7 along the other axis (231 of the
  // 256 threads). An axis of n cells has n + 1 edges, so tiles of 32 x 8
  // left a one-edge remainder tile on each axis at the drivers' sizes, which
  // cost nearly as much as a full tile; with 33 x 7, 96 + 1 = 33 + 33 + 31
  // and 192 + 1 = 27 x 7 + 4 leave no nearly empty tile.
  static constexpr int E1 = E == 1 ? 33 : 7;
  static constexpr int E2 = E == 1 ? 7 : 33;
  static constexpr int C1 = E1 + 7, C2 = E2 + 7;
  // Row stride of the t2-face arrays: when the edge loop walks t1 (= k)
  // fastest, an odd stride keeps a wave's reads on distinct LDS banks.
  static constexpr int S2 = T1 == 2 ? (E2 | 1) : E2;
  // (position along e, along t1, along t2) -> (i, j, k)
  GCS_DEV static void ijk(int ce, int a1, int a2, int &i, int &j, int &k) {
    int x[3];
    x[E] = ce;
    x[T1] = a1;
    x[T2] = a2;
    i = x[0];
    j = x[1];
    k = x[2];
  }
};

template <int E>
__global__ void __launch_bounds__(256) k_edge_emf_tiled(K c) {
  using T = EmfTile<E>;
  constexpr int E1 = T::E1, E2 = T::E2, C1 = T::C1, C2 = T::C2, S2 = T::S2;
  constexpr int t1 = T::T1, t2 = T::T2;
  __shared__ double s_vol[C1 * C2], s_wv[3][C1 * C2];
  __shared__ double s_fv[3][E1 * C2], s_a1[E1 * C2], s_b1[E1 * C2];
  __shared__ double s_a2[C1 * S2], s_b2[C1 * S2];
  const Layout &L = c.L;
  const int act[3] = {L.nr, L.n, L.n};
  // edge box: [g, g + act) along e, [g, g + act + 1) along t1 and t2
  const int lo1 = L.ghost, hi1 = L.ghost + act[t1] + 1;
  const int lo2 = L.ghost, hi2 = L.ghost + act[t2] + 1;
  const int nt1 = (hi1 - lo1 + E1 - 1) / E1, nt2 = (hi2 - lo2 + E2 - 1) / E2;
  // consecutive logical blocks, neighbours along t2 that share fill cells,
  // run on one XCD (one L2)
  long b = block_xcd();
  const int tile2 = (int)(b % nt2);
  b /= nt2;
  const int tile1 = (int)(b % nt1);
  b /= nt1;
  const int ce = L.ghost + (int)(b % act[E]);
  const int f = (int)(b / act[E]);
  if (f >= 6) return;
  const int o1 = lo1 + tile1 * E1, o2 = lo2 + tile2 * E2;  // tile origin
  const int n1 = min(E1, hi1 - o1), n2 = min(E2, hi2 - o2);  // valid edges
  const int tid = threadIdx.x;
  auto at_of = [&](int a1, int a2) {
    int i, j, k;
    T::ijk(ce, a1, a2, i, j, k);
    return L.at(f, i, j, k);
  };
  // This thread's edge (k fastest across the wave). Its own global data is
  // read first so that the loads overlap the tile fill; threads beyond a
  // partial tile read a valid edge's and leave after the tile phases.
  int e1, e2;
  if (t2 == 2) {
    e2 = tid % E2;
    e1 = tid / E2;
  } else {
    e1 = tid % E1;
    e2 = tid / E1;
  }
  const long at = at_of(o1 + min(e1, n1 - 1), o2 + min(e2, n2 - 1));
  const int ebp = P_EDGE + EDGE_PLANES * E;
  const bool edge_ok = c.b(Q_VALID + E)[at] && c.b(Q_EBOK + E)[at];
  double et1[3], et2[3], solve[5];
  for (int m = 0; m < 3; ++m) {
    et1[m] = c.p(ebp + EO_T1 + m)[at];
    et2[m] = c.p(ebp + EO_T2 + m)[at];
  }
  for (int m = 0; m < 5; ++m) solve[m] = c.p(ebp + EO_SOLVE + m)[at];
  const double b0t0 = c.p(P_B0ET + 2 * E)[at], b0t1 = c.p(P_B0ET + 2 * E + 1)[at];
  const double elen = c.p(ebp + EO_LEN)[at];
  // The fills walk k fastest across the threads (along t2 for e = 0, along
  // t1 for e = 1) so that each wave's loads are contiguous runs; the LDS
  // layouts do not depend on the walk.
  // cells: t1 in [o1 - 4, o1 + n1 + 3), t2 in [o2 - 4, o2 + n2 + 3)
  {
    const int w1 = n1 + 7, w2 = n2 + 7;
    for (int q = tid; q < w1 * w2; q += 256) {
      const int u1 = t1 == 2 ? q % w1 : q / w2, u2 = t1 == 2 ? q / w1 : q % w2;
      const long at = at_of(o1 - 4 + u1, o2 - 4 + u2);
      const int l = u1 * C2 + u2;
      s_vol[l] = c.p(P_VOL)[at];
      for (int m = 0; m < 3; ++m) s_wv[m][l] = c.p(P_WV + m)[at];
    }
  }
  // t1-faces: t1 in [o1, o1 + n1), t2 in [o2 - 4, o2 + n2 + 3)
  {
    const int w2 = n2 + 7;
    const double *area = c.p(P_FACE + FACE_PLANES * t1 + FO_AREA), *bn = c.p(P_BN + t1);
    for (int q = tid; q < n1 * w2; q += 256) {
      const int u1 = t1 == 2 ? q % n1 : q / w2, u2 = t1 == 2 ? q / n1 : q % w2;
      const long at = at_of(o1 + u1, o2 - 4 + u2);
      s_a1[u1 * C2 + u2] = area[at];
      s_b1[u1 * C2 + u2] = bn[at];
    }
  }
  // t2-faces: t1 in [o1 - 4, o1 + n1 + 3), t2 in [o2, o2 + n2)
  {
    const int w1 = n1 + 7;
    const double *area = c.p(P_FACE + FACE_PLANES * t2 + FO_AREA), *bn = c.p(P_BN + t2);
    for (int q = tid; q < w1 * n2; q += 256) {
      const int u1 = t1 == 2 ? q % w1 : q / n2, u2 = t1 == 2 ? q / w1 : q % n2;
      const long at = at_of(o1 - 4 + u1, o2 + u2);
      s_a2[u1 * S2 + u2] = area[at];
      s_b2[u1 * S2 + u2] = bn[at];
    }
  }
  __syncthreads();
  // t1-face velocities (k_face_velocity_i's central8, direction t1)
  {
    const int w2 = n2 + 7;
    for (int q = tid; q < n1 * w2; q += 256) {
      const int u1 = q / w2, u2 = q % w2;
      double vol[8], wv[3][8];
      for (int n = 0; n < 8; ++n) {
        const int l = (u1 + n) * C2 + u2;  // cell t1 = o1 + u1 + n - 4
        vol[n] = s_vol[l];
        for (int m = 0; m < 3; ++m) wv[m][n] = s_wv[m][l];
      }
      const double iv = central8(vol);
      const bool ok = isfinite(iv) && fabs(iv) > DBL_MIN;
      for (int m = 0; m < 3; ++m) s_fv[m][u1 * C2 + u2] = ok ? central8(wv[m]) / iv : NAN;
    }
  }
  __syncthreads();
  const StepOptions &o = c.o;
  const bool mine = e1 < n1 && e2 < n2;
  if (!mine) return;
  if (!edge_ok) {
    c.fail(51, at);
    return;
  }
  double a1[8], v[3];
  {
    double wv[3][8];
    for (int n = 0; n < 8; ++n) {
      const int l = e1 * C2 + e2 + n;  // t1-face at t2 = o2 + e2 + n - 4
      a1[n] = s_a1[l];
      for (int m = 0; m < 3; ++m) wv[m][n] = a1[n] * s_fv[m][l];
    }
    const double ea = central8(a1);
    if (!isfinite(ea) || fabs(ea) <= DBL_MIN) {
      c.fail(52, at);
      return;
    }
    for (int m = 0; m < 3; ++m) v[m] = central8(wv[m]) / ea;
  }
  const double v1 = dot3(v, et1), v2 = dot3(v, et2);
  double fl, fr, sl, sr;
  {
    double fld[8];
    for (int n = 0; n < 8; ++n) fld[n] = s_b1[e1 * C2 + e2 + n];
    up7_pdm(a1, central8(a1), fld, o.emf_pdm, fl, fr);
  }
  {
    double a2[8], fld[8];
    for (int n = 0; n < 8; ++n) {
      const int l = (e1 + n) * S2 + e2;  // t2-face at t1 = o1 + e1 + n - 4
      a2[n] = s_a2[l];
      fld[n] = s_b2[l];
    }
    const double ia = central8(a2);
    if (!isfinite(ia) || fabs(ia) <= DBL_MIN) {
      c.fail(53, at);
      return;
    }
    up7_pdm(a2, ia, fld, o.emf_pdm, sl, sr);
  }
  double bt[2];
  edge_field(solve, 0.5 * (fl + fr), 0.5 * (sl + sr), bt);
  bt[0] += b0t0;
  bt[1] += b0t1;
  const double jump = (sr - sl) - (fr - fl);
  const long step[3] = {L.di(), (long)L.SK, 1};
  double rho = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int bb = 0; bb < 2; ++bb) rho += 0.25 * c.p(P_UPRED)[at - a * step[t1] - bb * step[t2]];
  if (!isfinite(rho) || rho <= DBL_MIN) {
    c.fail(54, at);
    return;
  }
  double emf;
  if (!edge_emf(v1, v2, bt[0], bt[1], jump, rho, elen, o.emf_diff,
                o.emf_boris, o.emf_c, o.emf_cfl, c.sc->dt, emf)) {
    c.fail(55, at);
    return;
  }
  c.p(P_EMF + E)[at] = emf;
}

// nonorthogonal_step.c:35-71 on the inner wall face.
template <typename T>
__global__ void k_inner_trap(K c) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + 1, L.ghost + L.n, L.ghost + L.n};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  T *m = fluid_plane<T>(c, 0, 0) + at;
  T *e = fluid_plane<T>(c, 0, 4) + at;
  if (*m > T(0)) {
    *m = T(0);
    if (*e > T(0)) *e = T(0);
  }
}

// ---- edge EMF --------------------------------------------------------------------

// nonorthogonal_sweep.c:223-259: central8 cell velocity at the i-faces
// (direction 0) the k-edge EMF sweep reads (j and k from g - 4 to g + n + 4).
// A failure leaves NaN, which the EMF's finiteness test turns into the
// reference's error. Each thread walks a run of RM consecutive i-faces of
// one (j, k) column and keeps the RM + 7 cells their stencils share in
// registers, so a cell is loaded once per run instead of once per stencil
// use; each face is the same central8 expression.
constexpr int kFaceVelocityRun = 8;

template <int RM>
__global__ void __launch_bounds__(128) k_face_velocity_i(K c) {
  const Layout &L = c.L;
  const int nf = L.nr + 1;  // i-faces of the active box along a column
  const int lo[3] = {0, L.ghost - 4, L.ghost - 4};
  const int hi[3] = {(nf + RM - 1) / RM, min(L.ghost + L.n + 4, L.NJ),
                     min(L.ghost + L.n + 4, L.NK)};
  int f, r, j, k;
  if (!box(tid_xcd(), lo, hi, f, r, j, k)) return;
  const int i0 = L.ghost + r * RM;
  const int nv = min(RM, L.ghost + nf - i0);  // faces of this run
  const long si = L.di(), at0 = L.at(f, i0, j, k);
  // cells i0 - 4 .. i0 + nv + 2
  double vol[RM + 7], wv[3][RM + 7];
#pragma unroll
  for (int n = 0; n < RM + 7; ++n) {
    const bool in = n < nv + 7;
    const long cell = at0 + (n - 4) * si;
    vol[n] = in ? c.p(P_VOL)[cell] : 0.0;
    for (int e = 0; e < 3; ++e) wv[e][n] = in ? c.p(P_WV + e)[cell] : 0.0;
  }
#pragma unroll
  for (int m = 0; m < RM; ++m) {
    if (m < nv) {
      const double iv = central8(vol + m);
      const bool ok = isfinite(iv) && fabs(iv) > DBL_MIN;
      for (int e = 0; e < 3; ++e)
        c.p(P_FVEL + e)[at0 + m * si] = ok ? central8(wv[e] + m) / iv : NAN;
    }
  }
}

// nonorthogonal_sweep.c:223-259 one face per thread for the faces of
// direction d that the EMF sweeps read (d from g to g + n, the other axes
// from g - 4 to g + n + 4): the plain counterpart of k_face_velocity_i and
// of the face velocities inside k_edge_emf_tiled, same expression.
__global__ void __launch_bounds__(128) k_face_velocity(K c, int d) {
  const Layout &L = c.L;
  int lo[3], hi[3];
  const int ext[3] = {L.NI, L.NJ, L.NK};
  const int act[3] = {L.nr, L.n, L.n};
  for (int a = 0; a < 3; ++a) {
    if (a == d) {
      lo[a] = L.ghost;
      hi[a] = L.ghost + act[a] + 1;
    } else {
      lo[a] = L.ghost - 4;
      hi[a] = min(L.ghost + act[a] + 4, ext[a]);
    }
  }
  int f, i, j, k;
  if (!box(tid_xcd(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  const long st = d == 0 ? L.di() : d == 1 ? (long)L.SK : 1;
  double vol[8], wv[3][8];
  for (int n = 0; n < 8; ++n) {
    const long cell = at + (n - 4) * st;
    vol[n] = c.p(P_VOL)[cell];
    for (int e = 0; e < 3; ++e) wv[e][n] = c.p(P_WV + e)[cell];
  }
  const double iv = central8(vol);
  const bool ok = isfinite(iv) && fabs(iv) > DBL_MIN;
  for (int e = 0; e < 3; ++e) c.p(P_FVEL + 3 * d + e)[at] = ok ? central8(wv[e]) / iv : NAN;
}

// nonorthogonal_sweep.c:624-752 for edges of direction e bounding the
// active box.
__global__ void __launch_bounds__(128) k_edge_emf(K c, int e) {
  const Layout &L = c.L;
  const int t1 = (e + 1) % 3, t2 = (e + 2) % 3;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr + (e != 0), L.ghost + L.n + (e != 1),
                     L.ghost + L.n + (e != 2)};
  int f, i, j, k;
  if (!box(tid_xcd(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  if (!c.b(Q_VALID + e)[at] || !c.b(Q_EBOK + e)[at]) {
    c.fail(51, at);
    return;
  }
  const long step[3] = {L.di(), (long)L.SK, 1};
  const StepOptions &o = c.o;
  const int eb = P_EDGE + EDGE_PLANES * e;
  // Edge velocity: area-weighted central8 of t1-face velocities along t2.
  double a1[8], v[3];
  {
    double wv[3][8];
    for (int n = 0; n < 8; ++n) {
      const long fa = at + (n - 4) * step[t2];
      a1[n] = c.p(P_FACE + FACE_PLANES * t1 + FO_AREA)[fa];
      for (int m = 0; m < 3; ++m) wv[m][n] = a1[n] * c.p(P_FVEL + 3 * t1 + m)[fa];
    }
    const double ea = central8(a1);
    if (!isfinite(ea) || fabs(ea) <= DBL_MIN) {
      c.fail(52, at);
      return;
    }
    for (int m = 0; m < 3; ++m) v[m] = central8(wv[m]) / ea;
  }
  double et1[3], et2[3];
  for (int m = 0; m < 3; ++m) {
    et1[m] = c.p(eb + EO_T1 + m)[at];
    et2[m] = c.p(eb + EO_T2 + m)[at];
  }
  const double v1 = dot3(v, et1), v2 = dot3(v, et2);
  // Transverse face-normal B: Up7/PDM of flux / area along the other axis.
  double fl, fr, sl, sr;
  {
    double fld[8];
    const double *bn = c.p(P_BN + t1);
    for (int n = 0; n < 8; ++n) fld[n] = bn[at + (n - 4) * step[t2]];
    up7_pdm(a1, central8(a1), fld, o.emf_pdm, fl, fr);
  }
  {
    double a2[8], fld[8];
    const double *bn = c.p(P_BN + t2);
    for (int n = 0; n < 8; ++n) {
      const long fa = at + (n - 4) * step[t1];
      a2[n] = c.p(P_FACE + FACE_PLANES * t2 + FO_AREA)[fa];
      fld[n] = bn[fa];
    }
    const double ia = central8(a2);
    if (!isfinite(ia) || fabs(ia) <= DBL_MIN) {
      c.fail(53, at);
      return;
    }
    up7_pdm(a2, ia, fld, o.emf_pdm, sl, sr);
  }
  double solve[5], bt[2];
  for (int m = 0; m < 5; ++m) solve[m] = c.p(eb + EO_SOLVE + m)[at];
  edge_field(solve, 0.5 * (fl + fr), 0.5 * (sl + sr), bt);
  bt[0] += c.p(P_B0ET + 2 * e)[at];
  bt[1] += c.p(P_B0ET + 2 * e + 1)[at];
  const double jump = (sr - sl) - (fr - fl);
  double rho = 0.0;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) rho += 0.25 * c.p(P_UPRED)[at - a * step[t1] - b * step[t2]];
  if (!isfinite(rho) || rho <= DBL_MIN) {
    c.fail(54, at);
    return;
  }
  double emf;
  if (!edge_emf(v1, v2, bt[0], bt[1], jump, rho, c.p(eb + EO_LEN)[at], o.emf_diff,
                o.emf_boris, o.emf_c, o.emf_cfl, c.sc->dt, emf)) {
    c.fail(55, at);
    return;
  }
  c.p(P_EMF + e)[at] = emf;
}


// ---- mixed precision: FP32 reconstruction, fluxes and EMF ---------------------
//
// The same sweeps as k_face_flux, k_face_velocity_i and k_edge_emf with the
// arithmetic in float. Their per-step inputs (the predictor's caches, face
// velocities) and the face geometry they use are read from FP32 planes,
// already rounded; the face fluxes are stored to FP32 planes and the edge
// EMF to its FP64 plane, so the update, CT, CFL and ionosphere stay FP64.

// k_face_flux's face in FP32 (same stencil variables as face_flux_face,
// read from the FP32 planes).
template <class Cell>
GCS_DEV void face_flux_face_fp32(const K &c, int d, int f, long at, long st, const Cell &cell) {
  using R = float;
  using namespace mixed;
  const StepOptions &o = c.o;
  const R gamma = (R)o.gamma, pdm = (R)o.pdm;
  // Primitives of the stencil from the predictor's FP32 cache.
  R q8[8];
  bool converted = true;
  for (int n = 0; n < 8; ++n) {
    q8[n] = cell(0, n);
    converted &= isfinite(q8[n]);
  }
  if (!converted) {
    c.fail(41, at);
    return;
  }
  R vol[8];
  for (int n = 0; n < 8; ++n) vol[n] = cell(5, n);
  const R iv = central8_r(vol);
  if (!isfinite(iv) || fabs(iv) <= tiny<R>()) {
    c.fail(42, at);
    return;
  }
  const R riv = R(1) / iv;
  PrimR<R> w[2];
  up7_pdm_shared_r(vol, iv, riv, q8, pdm, w[0].rho, w[1].rho);
  for (int e = 0; e < 3; ++e) {
    for (int n = 0; n < 8; ++n) q8[n] = cell(1 + e, n);
    up7_pdm_shared_r(vol, iv, riv, q8, pdm, w[0].v[e], w[1].v[e]);
  }
  for (int n = 0; n < 8; ++n) q8[n] = cell(4, n);
  up7_pdm_shared_r(vol, iv, riv, q8, pdm, w[0].p, w[1].p);
  R bb[2][3] = {{R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
  if (o.mhd) {
    for (int e = 0; e < 3; ++e) {
      R q8[8];
      for (int n = 0; n < 8; ++n) q8[n] = cell(6 + e, n);
      up7_pdm_shared_r(vol, iv, riv, q8, pdm, bb[0][e], bb[1][e]);
    }
  }
  if (!valid_interface_r(w[0].rho, w[0].v, w[0].p) ||
      !valid_interface_r(w[1].rho, w[1].v, w[1].p)) {
    // The limiter's last subtraction cancels in FP32 when an interface
    // density or pressure is orders of magnitude below a neighbour's (gas
    // near its floors beside the wind): it can land on zero or below. Such
    // a face is reconstructed again in FP64 from the same stencil values
    // (widened exactly) and rounded once, which keeps it positive.
    double vd[8], qd[8];
    for (int n = 0; n < 8; ++n) vd[n] = (double)cell(5, n);
    const double ivd = central8(vd);
    double side[2][5];
    for (int v = 0; v < 5; ++v) {
      for (int n = 0; n < 8; ++n) qd[n] = (double)cell(v, n);
      up7_pdm(vd, ivd, qd, o.pdm, side[0][v], side[1][v]);
    }
    for (int s = 0; s < 2; ++s) {
      w[s].rho = (R)side[s][0];
      for (int e = 0; e < 3; ++e) w[s].v[e] = (R)side[s][1 + e];
      w[s].p = (R)side[s][4];
    }
    if (!valid_interface_r(w[0].rho, w[0].v, w[0].p) ||
        !valid_interface_r(w[1].rho, w[1].v, w[1].p)) {
      c.fail(43, at);
      return;
    }
  }
  const int fb = M_FACE + M_FACE_PLANES * d;
  const R area = c.sp(fb + MF_AREA)[at];
  FaceFrameR<R> fr;
  for (int e = 0; e < 3; ++e) {
    fr.n[e] = c.sp(fb + MF_N + e)[at];
    fr.t1[e] = c.sp(fb + MF_T1 + e)[at];
    fr.t2[e] = c.sp(fb + MF_T2 + e)[at];
  }
  FluidFluxR<R> fl;
  fluid_flux_r(w, gamma, fr, fl);
  const R cl = (R)o.c;
  MaxwellFluxR<R> mx;
  for (int e = 0; e < 3; ++e) mx.m[e] = R(0);
  mx.va = mx.psum = R(0);
  if (o.mhd) {
    const R bn = c.sp(M_BN + d)[at];
    const R b0[3] = {c.sp(fb + MF_B0)[at], c.sp(fb + MF_B0 + 1)[at], c.sp(fb + MF_B0 + 2)[at]};
    if (!maxwell_flux_r(w, bb, bn, fr.n, b0, c.sp(fb + MF_B0N)[at], o.boris != 0, cl, fl.un,
                        mx)) {
      c.fail(44, at);
      return;
    }
  }
  if (o.mhd && o.hogs) hogs_r(fl, mx, (R)o.hogs_h, (R)o.hogs_m, o.boris != 0, cl);
  if (o.mhd && o.hogs && o.boris) {
    R s2 = R(0);
    s2 += dot3r(w[0].v, w[0].v);
    s2 += dot3r(w[1].v, w[1].v);
    if (sqrt(R(0.5) * s2) >= R(1.5) * cl) {
      const long lc = at - st;
      for (int q = 0; q < 5; ++q) {
        const R lo_u = c.sp(M_UPRED + q)[lc], up_u = c.sp(M_UPRED + q)[at];
        if (!isfinite(lo_u) || !isfinite(up_u)) {
          c.fail(45, at);
          return;
        }
        fl.f[q] -= cl * (up_u - lo_u);
      }
      R t2 = R(0);
      for (int side = 0; side < 2; ++side)
        for (int e = 0; e < 3; ++e) t2 += w[side].v[e] * w[side].v[e];
      atomicAdd(&c.sc->emergency_count[f], 1ull);
      atomicMax(&c.sc->emergency_speed_bits[f],
                (unsigned long long)__double_as_longlong((double)sqrt(R(0.5) * t2)));
    }
  }
  for (int q = 0; q < 5; ++q) c.sp(M_FLUID + 5 * d + q)[at] = area * fl.f[q];
  for (int e = 0; e < 3; ++e) c.sp(M_MAXWELL + 3 * d + e)[at] = area * mx.m[e];
}

// Four waves per SIMD: the rare FP64 fallback reconstruction otherwise
// lifts the kernel to 154 VGPRs and three waves; bounded, it spills 84
// bytes a lane on that path only (0.75 -> 0.63 ms per direction at config 1).
__global__ void __launch_bounds__(128, 4) k_face_flux_fp32(K c, int d) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr + (d == 0), L.ghost + L.n + (d == 1),
                     L.ghost + L.n + (d == 2)};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  const long st = d == 0 ? L.di() : d == 1 ? (long)L.SK : 1;
  face_flux_face_fp32(c, d, f, at, st, [&](int v, int n) {
    const int plane = v < 5 ? M_WPRED + v : v == 5 ? M_VOL : M_BPRED + v - 6;
    return c.sp(plane)[at + (n - 4) * st];
  });
}

__global__ void k_face_velocity_fp32(K c, int d) {
  using R = float;
  using namespace mixed;
  const Layout &L = c.L;
  int lo[3], hi[3];
  const int ext[3] = {L.NI, L.NJ, L.NK};
  const int act[3] = {L.nr, L.n, L.n};
  for (int a = 0; a < 3; ++a) {
    if (a == d) {
      lo[a] = L.ghost;
      hi[a] = L.ghost + act[a] + 1;
    } else {
      lo[a] = L.ghost - 4;
      hi[a] = min(L.ghost + act[a] + 4, ext[a]);
    }
  }
  int f, i, j, k;
  if (!box(tid_xcd(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  const long st = d == 0 ? L.di() : d == 1 ? (long)L.SK : 1;
  R vol[8], wv[3][8];
  for (int n = 0; n < 8; ++n) {
    const long cell = at + (n - 4) * st;
    vol[n] = c.sp(M_VOL)[cell];
    for (int e = 0; e < 3; ++e) wv[e][n] = c.sp(M_WV + e)[cell];
  }
  const R iv = central8_r(vol);
  const bool ok = isfinite(iv) && fabs(iv) > tiny<R>();
  for (int e = 0; e < 3; ++e)
    c.sp(M_FVEL + 3 * d + e)[at] = ok ? central8_r(wv[e]) / iv : (R)NAN;
}

// The FP32 i-face velocities as the same register march as
// k_face_velocity_i (runs of RM faces of one (j, k) column), with this
// kernel's FP32 expressions.
template <int RM>
__global__ void __launch_bounds__(128) k_face_velocity_i_fp32(K c) {
  using R = float;
  using namespace mixed;
  const Layout &L = c.L;
  const int nf = L.nr + 1;
  const int lo[3] = {0, L.ghost - 4, L.ghost - 4};
  const int hi[3] = {(nf + RM - 1) / RM, min(L.ghost + L.n + 4, L.NJ),
                     min(L.ghost + L.n + 4, L.NK)};
  int f, r, j, k;
  if (!box(tid_xcd(), lo, hi, f, r, j, k)) return;
  const int i0 = L.ghost + r * RM;
  const int nv = min(RM, L.ghost + nf - i0);
  const long si = L.di(), at0 = L.at(f, i0, j, k);
  R vol[RM + 7], wv[3][RM + 7];
#pragma unroll
  for (int n = 0; n < RM + 7; ++n) {
    const bool in = n < nv + 7;
    const long cell = at0 + (n - 4) * si;
    vol[n] = in ? c.sp(M_VOL)[cell] : R(0);
    for (int e = 0; e < 3; ++e) wv[e][n] = in ? c.sp(M_WV + e)[cell] : R(0);
  }
#pragma unroll
  for (int m = 0; m < RM; ++m) {
    if (m < nv) {
      const R iv = central8_r(vol + m);
      const bool ok = isfinite(iv) && fabs(iv) > tiny<R>();
      for (int e = 0; e < 3; ++e)
        c.sp(M_FVEL + e)[at0 + m * si] = ok ? central8_r(wv[e] + m) / iv : (R)NAN;
    }
  }
}

__global__ void __launch_bounds__(128) k_edge_emf_fp32(K c, int e) {
  using R = float;
  using namespace mixed;
  const Layout &L = c.L;
  const int t1 = (e + 1) % 3, t2 = (e + 2) % 3;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr + (e != 0), L.ghost + L.n + (e != 1),
                     L.ghost + L.n + (e != 2)};
  int f, i, j, k;
  if (!box(tid_xcd(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  if (!c.b(Q_VALID + e)[at] || !c.b(Q_EBOK + e)[at]) {
    c.fail(51, at);
    return;
  }
  const long step[3] = {L.di(), (long)L.SK, 1};
  const StepOptions &o = c.o;
  const int eb = M_EDGE + M_EDGE_PLANES * e;
  R a1[8], v[3];
  {
    R wv[3][8];
    for (int n = 0; n < 8; ++n) {
      const long fa = at + (n - 4) * step[t2];
      a1[n] = c.sp(M_FACE + M_FACE_PLANES * t1 + MF_AREA)[fa];
      for (int m = 0; m < 3; ++m) wv[m][n] = a1[n] * c.sp(M_FVEL + 3 * t1 + m)[fa];
    }
    const R ea = central8_r(a1);
    if (!isfinite(ea) || fabs(ea) <= tiny<R>()) {
      c.fail(52, at);
      return;
    }
    for (int m = 0; m < 3; ++m) v[m] = central8_r(wv[m]) / ea;
  }
  R et1[3], et2[3];
  for (int m = 0; m < 3; ++m) {
    et1[m] = c.sp(eb + ME_T1 + m)[at];
    et2[m] = c.sp(eb + ME_T2 + m)[at];
  }
  const R v1 = dot3r(v, et1), v2 = dot3r(v, et2);
  const R pdm = (R)o.emf_pdm;
  R fl, fr, sl, sr;
  {
    R fld[8];
    const float *bn = c.sp(M_BN + t1);
    for (int n = 0; n < 8; ++n) fld[n] = bn[at + (n - 4) * step[t2]];
    up7_pdm_r(a1, central8_r(a1), fld, pdm, fl, fr);
  }
  {
    R a2[8], fld[8];
    const float *bn = c.sp(M_BN + t2);
    for (int n = 0; n < 8; ++n) {
      const long fa = at + (n - 4) * step[t1];
      a2[n] = c.sp(M_FACE + M_FACE_PLANES * t2 + MF_AREA)[fa];
      fld[n] = bn[fa];
    }
    const R ia = central8_r(a2);
    if (!isfinite(ia) || fabs(ia) <= tiny<R>()) {
      c.fail(53, at);
      return;
    }
    up7_pdm_r(a2, ia, fld, pdm, sl, sr);
  }
  R solve[5], bt[2];
  for (int m = 0; m < 5; ++m) solve[m] = c.sp(eb + ME_SOLVE + m)[at];
  edge_field_r(solve, R(0.5) * (fl + fr), R(0.5) * (sl + sr), bt);
  bt[0] += c.sp(eb + ME_B0T)[at];
  bt[1] += c.sp(eb + ME_B0T + 1)[at];
  const R jump = (sr - sl) - (fr - fl);
  R rho = R(0);
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      rho += R(0.25) * c.sp(M_UPRED)[at - a * step[t1] - b * step[t2]];
  if (!isfinite(rho) || rho <= tiny<R>()) {
    c.fail(54, at);
    return;
  }
  R emf;
  if (!edge_emf_r(v1, v2, bt[0], bt[1], jump, rho, c.sp(eb + ME_LEN)[at], (R)o.emf_diff,
                  o.emf_boris != 0, (R)o.emf_c, (R)o.emf_cfl, (R)c.sc->dt, emf)) {
    c.fail(55, at);
    return;
  }
  c.p(P_EMF + e)[at] = (double)emf;
}

// cubedsphere_earth_boundary.c:445-489: outer shells blend toward the wind
// line integral -(V x B) . l.
__global__ void k_earth_emf_outer(K c, int d, WindSeries w) {
  const Layout &L = c.L;
  const int outer = L.ghost + L.nr;
  const int lo[3] = {outer - 3, 0, 0};
  const int hi[3] = {min(outer + 1, L.NI + (d != 0)), L.NJ + (d != 1), L.NK + (d != 2)};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  if (!c.b(Q_VALID + d)[at]) return;
  const long up = at + (d == 0 ? L.di() : d == 1 ? (long)L.SK : 1);
  double x[3];
  for (int m = 0; m < 3; ++m) x[m] = 0.5 * (c.p(P_VERT + m)[at] + c.p(P_VERT + m)[up]);
  double sw[8], wt;
  if (!sample_wind(w, x, c.sc->wall_time + 0.5 * c.sc->wall_dt, sw, wt)) {
    c.fail(61, at);
    return;
  }
  if (i < outer) wt /= 1.0 + (double)(outer - i);
  const double *vv = sw + 2, *bb = sw + 5;
  const double el[3] = {-(vv[1] * bb[2] - vv[2] * bb[1]), -(vv[2] * bb[0] - vv[0] * bb[2]),
                        -(vv[0] * bb[1] - vv[1] * bb[0])};
  const int eb = P_EDGE + EDGE_PLANES * d;
  const double len = c.p(eb + EO_LEN)[at];
  double line = 0.0;
  for (int m = 0; m < 3; ++m) line += el[m] * c.p(eb + EO_N + m)[at] * len;
  double *emf = c.p(P_EMF + d) + at;
  *emf = (1.0 - wt) * *emf + wt * line;
}

// cubedsphere_earth_boundary.c:362-443: diffusive wind EMF on the outer wall.
__global__ void k_earth_emf_diffusive(K c, WindSeries w) {
  const Layout &L = c.L;
  const int i = L.ghost + L.nr;
  const int lo[3] = {i, L.ghost, L.ghost};
  const int hi[3] = {i + 1, L.ghost + L.n + 1, L.ghost + L.n + 1};
  int f, ii, j, k;
  if (!box(tid(), lo, hi, f, ii, j, k)) return;
  // Edge extents: J (NI+1, NJ, NK+1), K (NI+1, NJ+1, NK).
  if (j == 0 || k == 0 || j >= L.NJ || k >= L.NK) return;
  const long at = L.at(f, i, j, k);
  if (!c.b(Q_VALID + 1)[at] || !c.b(Q_VALID + 2)[at]) return;
  const long si = L.di(), sj = L.SK;
  auto dens = [&](int dir, long fa, double &out) {
    const double area = c.p(P_FACE + FACE_PLANES * dir + FO_AREA)[fa];
    if (!(area > 0.0)) return false;
    const double cf = c.p(plane_phi(c.cur, dir))[fa], of = c.p(plane_phi(c.old, dir))[fa];
    out = (cf + c.sc->ratio * (cf - of)) / area;  // the predictor's Phi*
    return (bool)isfinite(out);
  };
  double bikp, bikm, bkip, bkim, bjip, bjim, bijp, bijm;
  if (!dens(0, at, bikp) || !dens(0, at - 1, bikm) || !dens(2, at, bkip) ||
      !dens(2, at - si, bkim) || !dens(1, at, bjip) || !dens(1, at - si, bjim) ||
      !dens(0, at, bijp) || !dens(0, at - sj, bijm)) {
    c.fail(62, at);
    return;
  }
  const double cj = (bikp - bikm) - (bkip - bkim);
  const double ck = (bjip - bjim) - (bijp - bijm);
  const double dl = sqrt(c.p(P_EDGE + EDGE_PLANES * 1 + EO_LEN)[at] *
                         c.p(P_EDGE + EDGE_PLANES * 2 + EO_LEN)[at]);
  double xj[3], xk[3];
  for (int m = 0; m < 3; ++m) {
    xj[m] = 0.5 * (c.p(P_VERT + m)[at] + c.p(P_VERT + m)[at + sj]);
    xk[m] = 0.5 * (c.p(P_VERT + m)[at] + c.p(P_VERT + m)[at + 1]);
  }
  const double t = c.sc->wall_time + 0.5 * c.sc->wall_dt;
  double wj[8], wk[8], gj, gk;
  if (!(dl > 0.0) || !sample_wind(w, xj, t, wj, gj) || !sample_wind(w, xk, t, wk, gk)) {
    c.fail(63, at);
    return;
  }
  const double c2 = w.by * w.by + w.bz * w.bz;
  const double front = fmax(fabs(wj[2]), fabs(wk[2])) / sqrt(1.0 + c2);
  const double speed = fmin(front, c.o.wall_cfl * dl / c.sc->wall_dt);
  c.p(P_EMF + 1)[at] += gj * speed * cj * dl;
  c.p(P_EMF + 2)[at] += gk * speed * ck * dl;
}

// cubedsphere_earth_boundary.c:429-440: E_t = 0 on the inner wall shell, or
// with the ionosphere ready, -(potential difference) / potential norm along
// each wall edge (cubedsphere_mi_dipole.c:478-503).
__global__ void k_earth_emf_inner(K c) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, 0, 0};
  const int hi[3] = {L.ghost + 1, L.NJ + 1, L.NK + 1};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  // J edges span (NI+1, NJ, NK+1), K edges (NI+1, NJ+1, NK).
  for (int d = 1; d <= 2; ++d) {
    if ((d == 1 && j >= L.NJ) || (d == 2 && k >= L.NK) || !c.b(Q_VALID + d)[at])
      continue;
    double emf = 0.0;
    if (c.mi.potential != nullptr) {
      const long end = at + (d == 1 ? (long)L.SK : 1L);
      const double a[3] = {c.p(P_VERT)[at], c.p(P_VERT + 1)[at], c.p(P_VERT + 2)[at]};
      const double b[3] = {c.p(P_VERT)[end], c.p(P_VERT + 1)[end], c.p(P_VERT + 2)[end]};
      double pa, pb;
      if (!mi_potential(c.mi, a, pa) || !mi_potential(c.mi, b, pb)) {
        c.fail(64, at);
        return;
      }
      emf = -(pb - pa) / c.mi.potential_norm;
    }
    c.p(P_EMF + d)[at] = emf;
  }
}

// ---- update ------------------------------------------------------------------------

// Fused stress divergence (nonorthogonal_sweep.c:550-622), Reynolds update
// (nonorthogonal_step.c:542-553), CT on the cell's faces (:233-298), B
// recovery (:565) and Boris correction (:571-606) of one active cell,
// writing the new level. With Mixed the face fluxes come from the FP32
// planes (widened exactly); the arithmetic is FP64 either way.
// With want_cfl, returns the new state's CFL dt of the cell (cfl_dt) for the
// next step's dt rule; DBL_MAX otherwise, for a seam cell (its B changes
// again in the final seam sync; k_recover_list folds its dt) and for a
// failed cell.
template <bool Mixed>
GCS_DEV double update_cell(const K &c, int i, int j, int k, long at, const int hi[3],
                           bool want_cfl) {
  using Flux = typename std::conditional<Mixed, float, double>::type;
  const Layout &L = c.L;
  const long si = L.di(), sj = L.SK;
  const long step[3] = {si, sj, 1};
  const StepOptions &o = c.o;
  const double dt = c.sc->dt;
  const double vol = c.p(P_VOL)[at];
  const double ivol = 1.0 / vol;
  double rate[5], mr[3] = {0.0, 0.0, 0.0};
  for (int q = 0; q < 5; ++q) {
    double diff = 0.0;
    for (int d = 0; d < 3; ++d) {
      const Flux *fp = fluid_plane<Flux>(c, d, q);
      diff += (double)fp[at] - (double)fp[at + step[d]];
    }
    rate[q] = diff * ivol;
  }
  if (o.mhd) {
    for (int m = 0; m < 3; ++m) {
      double diff = 0.0;
      for (int d = 0; d < 3; ++d) {
        const Flux *mp = maxwell_plane<Flux>(c, d, m);
        diff += (double)mp[at] - (double)mp[at + step[d]];
      }
      mr[m] = diff * ivol;
      mr[m] += c.p(P_FORCE + m)[at];
    }
  }
  if (o.rates) {  // storage hydro_rate / maxwell_rate (nonorthogonal_sweep.c:585-619)
    for (int q = 0; q < 5; ++q) c.p(P_FVEL + q)[at] = rate[q];
    for (int m = 0; m < 3; ++m) c.p(P_FVEL + 5 + m)[at] = mr[m];
  }
  double u0[5], u[5];
  for (int q = 0; q < 5; ++q) {
    u0[q] = c.p(plane_u(c.cur, q))[at];
    u[q] = u0[q] + dt * rate[q];
  }
  if (!o.mhd) {
    // nonorthogonal_step.c:554-556: a hydrodynamic step ends with the
    // Reynolds update; cell B and the face fluxes carry over unchanged.
    for (int q = 0; q < 5; ++q) c.p(plane_u(c.old, q))[at] = u[q];
    for (int m = 0; m < 3; ++m) c.p(plane_b(c.old, m))[at] = c.p(plane_b(c.cur, m))[at];
    for (int d = 0; d < 3; ++d) {
      c.p(plane_phi(c.old, d))[at] = c.p(plane_phi(c.cur, d))[at];
      const int pos = d == 0 ? i : d == 1 ? j : k;
      if (pos == hi[d] - 1)
        c.p(plane_phi(c.old, d))[at + step[d]] = c.p(plane_phi(c.cur, d))[at + step[d]];
    }
    return DBL_MAX;
  }
  // CT: new flux of the six faces (identical arithmetic in both cells of a
  // shared face).
  const double *E0 = c.p(P_EMF), *E1 = c.p(P_EMF + 1), *E2 = c.p(P_EMF + 2);
  double flux[6];
  for (int s = 0; s < 2; ++s) {
    const long a0 = at + s * si;
    const double curl0 = E2[a0 + sj] - E2[a0] - E1[a0 + 1] + E1[a0];
    flux[0 + s] = c.p(plane_phi(c.cur, 0))[a0] - dt * curl0;
    const long a1 = at + s * sj;
    const double curl1 = E0[a1 + 1] - E0[a1] - E2[a1 + si] + E2[a1];
    flux[2 + s] = c.p(plane_phi(c.cur, 1))[a1] - dt * curl1;
    const long a2 = at + s;
    const double curl2 = E1[a2 + si] - E1[a2] - E0[a2 + sj] + E0[a2];
    flux[4 + s] = c.p(plane_phi(c.cur, 2))[a2] - dt * curl2;
  }
  double fc[6][3], cen[3], rec[9], bn[3];
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int m = 0; m < 3; ++m)
        fc[2 * d + s][m] = c.p(P_FACE + FACE_PLANES * d + FO_CEN + m)[at + s * step[d]];
  for (int m = 0; m < 3; ++m) cen[m] = c.p(P_CEN + m)[at];
  for (int q = 0; q < 9; ++q) rec[q] = c.p(P_REC + q)[at];
  const int piv = c.b(Q_PIV)[at];
  if (piv == 0xff || !recover_cell(flux, fc, cen, vol, rec, piv, bn)) {
    c.fail(71, at);
    return DBL_MAX;
  }
  if (o.upd_boris) {
    double btn[3], bto[3];
    for (int m = 0; m < 3; ++m) {
      const double b0 = c.p(P_B0C + m)[at];
      btn[m] = bn[m] + b0;
      bto[m] = c.p(plane_b(c.cur, m))[at] + b0;
    }
    if (!boris_update(u, u0, btn, bto, rate, mr, o.upd_c, dt, o.upd_gamma,
                      o.upd_rho_floor, o.upd_p_floor)) {
      c.fail(72, at);
      return DBL_MAX;
    }
  } else {
    maxwell_update(u, mr, dt, o.upd_gamma, o.upd_rho_floor, o.upd_p_floor);
  }
  for (int q = 0; q < 5; ++q) c.p(plane_u(c.old, q))[at] = u[q];
  for (int m = 0; m < 3; ++m) c.p(plane_b(c.old, m))[at] = bn[m];
  c.p(plane_phi(c.old, 0))[at] = flux[0];
  c.p(plane_phi(c.old, 1))[at] = flux[2];
  c.p(plane_phi(c.old, 2))[at] = flux[4];
  if (i == hi[0] - 1) c.p(plane_phi(c.old, 0))[at + si] = flux[1];
  if (j == hi[1] - 1) c.p(plane_phi(c.old, 1))[at + sj] = flux[3];
  if (k == hi[2] - 1) c.p(plane_phi(c.old, 2))[at + 1] = flux[5];
  const int g = L.ghost;
  if (!want_cfl || j == g || j == hi[1] - 1 || k == g || k == hi[2] - 1) return DBL_MAX;
  return cfl_dt(c, at, u, bn);
}

// One active cell per thread; with `cfl` (a driver step), every cell's dt
// of the new state is folded into the CFL slots.
template <bool Mixed>
__global__ void __launch_bounds__(128) k_update(K c, unsigned long long *cfl) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr, L.ghost + L.n, L.ghost + L.n};
  int f, i, j, k;
  double dt_cell = DBL_MAX;
  if (box(tid(), lo, hi, f, i, j, k))
    dt_cell = update_cell<Mixed>(c, i, j, k, L.at(f, i, j, k), hi, cfl != nullptr);
  if (cfl != nullptr) cfl_fold(dt_cell, cfl);
}

// Ghost values of the new level as the reference leaves them in storage:
// ghost U and ghost face fluxes keep the step's start values (the reference
// copies current into old before updating only active elements).
__global__ void k_keep_ghosts(K c) {
  const Layout &L = c.L;
  const int lo[3] = {0, 0, 0};
  const int hi[3] = {L.NI + 1, L.NJ + 1, L.NK + 1};
  int f, i, j, k;
  if (!box(tid(), lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  const int g = L.ghost;
  const bool ai = i >= g && i < g + L.nr, aj = j >= g && j < g + L.n,
             ak = k >= g && k < g + L.n;
  const bool cell = i < L.NI && j < L.NJ && k < L.NK;
  if (cell && !(ai && aj && ak)) {
    for (int q = 0; q < 5; ++q) c.p(plane_u(c.old, q))[at] = c.p(plane_u(c.cur, q))[at];
    // a hydrodynamic step recovers no B: ghost B carries over as well
    if (!c.o.mhd)
      for (int m = 0; m < 3; ++m) c.p(plane_b(c.old, m))[at] = c.p(plane_b(c.cur, m))[at];
  }
  const bool fi = i >= g && i <= g + L.nr, fj = j >= g && j <= g + L.n,
             fk = k >= g && k <= g + L.n;
  const bool face[3] = {i <= L.NI && j < L.NJ && k < L.NK && !(fi && aj && ak),
                        i < L.NI && j <= L.NJ && k < L.NK && !(ai && fj && ak),
                        i < L.NI && j < L.NJ && k <= L.NK && !(ai && aj && fk)};
  for (int d = 0; d < 3; ++d)
    if (face[d]) c.p(plane_phi(c.old, d))[at] = c.p(plane_phi(c.cur, d))[at];
}

}  // namespace

static unsigned nblocks(long n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

static K make_k(Problem &P) {
  K c;
  c.L = P.L;
  c.pl = P.planes;
  c.fl = P.planes32;
  c.by = P.bytes;
  c.sc = P.scalars;
  c.cur = P.level;
  c.old = 1 - P.level;
  c.o = P.opt;
  c.o.rates = P.fidelity ? 1 : 0;
  c.mi = P.mi_view;
  // the plain-kernel switch computes the drift in the step instead
  c.mi.drift = P.plain ? nullptr : mi_drift_cache(P);
  return c;
}

void driver_options(Problem &P) {
  // test_cubedsphere_dipole_background.c:405-449 (floors 1e-6 / 1e-8, PDM 1,
  // HOGS 0.25 / 0.25, Boris c = 10, EMF diffusion 0.5).
  StepOptions &o = P.opt;
  const gcs_problem_config &cfg = P.cfg;
  const double gamma = cfg.startup.gamma;
  const int boris = cfg.use_boris != 0;
  o.gamma = gamma;
  o.rho_floor = 1e-6;
  o.p_floor = 1e-8;
  o.pdm = 1.0;
  o.hogs_h = 0.25;
  o.hogs_m = 0.25;
  o.c = 10.0;
  o.boris = boris;
  o.hogs = 1;
  o.emf_rho_floor = 1e-6;
  o.emf_pdm = 1.0;
  o.emf_diff = 0.5;
  o.emf_c = 10.0;
  o.emf_cfl = cfg.cfl;
  o.emf_boris = boris;
  o.upd_gamma = gamma;
  o.upd_rho_floor = 1e-6;
  o.upd_p_floor = 1e-8;
  o.upd_c = 10.0;
  o.upd_boris = boris;
  o.ts_gamma = gamma;
  o.ts_rho_floor = 1e-6;
  o.ts_p_floor = 1e-8;
  o.ts_cfl = cfg.cfl;
  o.ts_c = 10.0;
  o.ts_boris = boris;
  o.wall_gamma = gamma;
  o.wall_rho_floor = 1e-6;
  o.wall_p_floor = 1e-8;
  o.wall_cfl = cfg.cfl;
  o.strict_wall = cfg.strict_inner_gas_wall != 0;
  P.bnd.outflow = false;
  P.bnd.earth_cells = P.bnd.inner_trap = P.bnd.earth_emf = true;
}

// The drift of every inner ghost cell for `potential` (MiView::drift).
__global__ void k_mi_drift_cache(K c, double *drift) {
  const Layout &L = c.L;
  const long n = 6L * L.ghost * L.NJ * L.NK, q = tid();
  if (q >= n) return;
  const int k = (int)(q % L.NK);
  const long r = q / L.NK;
  const int j = (int)(r % L.NJ);
  const int i = (int)((r / L.NJ) % L.ghost), f = (int)(r / ((long)L.NJ * L.ghost));
  const long t = L.at(f, i, j, k);
  const double x[3] = {c.p(P_CEN)[t], c.p(P_CEN + 1)[t], c.p(P_CEN + 2)[t]};
  double d[3];
  if (!mi_drift(c.mi, x, d)) d[0] = d[1] = d[2] = NAN;
  for (int e = 0; e < 3; ++e) drift[e * n + q] = d[e];
}

int refresh_mi_drift(Problem &P, const double *potential, double *drift) {
  if (drift == nullptr) return 0;
  K c = make_k(P);
  c.mi.potential = potential;
  c.mi.drift = nullptr;
  const long n = 6L * P.L.ghost * P.L.NJ * P.L.NK;
  k_mi_drift_cache<<<nblocks(n, 256), 256, 0, P.stream>>>(c, drift);
  GCS_HIP(hipGetLastError());
  return 0;
}

int gather_seam_geometry(Problem &P) {
  (void)hipFree(P.seam_geom);
  P.seam_geom = nullptr;
  if (P.n_seam_cells == 0) return 0;
  if (hipMalloc(&P.seam_geom, (size_t)kSeamFields * P.n_seam_cells * sizeof(double)) !=
      hipSuccess) {
    set_error("device allocation failed");
    return -1;
  }
  k_seam_geometry<<<nblocks(P.n_seam_cells, 256), 256, 0, P.stream>>>(
      make_k(P), P.seam_cells, P.n_seam_cells, P.seam_geom);
  GCS_HIP(hipGetLastError());
  return 0;
}

int launch_recover_all(Problem &P, int phi_first, int b_first) {
  K c = make_k(P);
  const Layout &L = P.L;
  const long n = 6L * L.NI * L.NJ * L.NK;
  k_recover<<<nblocks(n, 256), 256, 0, P.stream>>>(c, phi_first, b_first, 0, 0);
  GCS_HIP(hipGetLastError());
  return 0;
}

const char *stage_name(int stage) {
  static const char *names[kStages] = {
      "cfl", "dt", "sync_phi", "halo_cells", "radial_outflow", "halo_faces",
      "earth_cells", "earth_faces", "earth_recover", "predict", "face_flux_i",
      "face_flux_j", "face_flux_k", "sync_flux", "inner_trap", "face_velocity",
      "edge_emf_i", "edge_emf_j", "edge_emf_k", "earth_emf", "sync_edges",
      "update", "final_sync", "seam_recover", "fidelity", "clock", "end", "mi_update"};
  return stage >= 0 && stage < kStages ? names[stage] : "?";
}

// Record a stage boundary when profiling (events are preallocated).
void mark_stage(Problem &P, int stage) {
  if (!P.profiling || P.stage_cursor >= P.stage_events.size()) return;
  (void)hipEventRecord(P.stage_events[P.stage_cursor], P.stream);
  P.stage_ids[P.stage_cursor] = stage;
  ++P.stage_cursor;
}

int enqueue_phase(Problem &P, Phase phase) {
  K c = make_k(P);
  const Layout &L = P.L;
  hipStream_t s = P.stream;
  const int T = 256;
  const long n_active = 6L * L.nr * L.n * L.n;
  const long n_cells = 6L * L.NI * L.NJ * L.NK;
  const long n_box = 6L * (L.NI + 1) * (L.NJ + 1) * (L.NK + 1);
  const WindSeries &w = P.wind_series;
  const StepBoundary &bd = P.bnd;

  if (phase == PHASE_GHOSTS) {
    // cubedsphere_step.c:51-65
    mark_stage(P, S_SYNC_PHI);
    k_sync_phi<<<nblocks(P.n_seams, T), T, 0, s>>>(c, P.seams, P.n_seams, 1);
    mark_stage(P, S_HALO_CELLS);
    k_halo_cells<<<nblocks(P.n_halo_cells, T), T, 0, s>>>(c, P.halo_cells, P.n_halo_cells,
                                                         P.fidelity ? 1 : 0);
    if (bd.outflow) {
      mark_stage(P, S_OUTFLOW);
      k_radial_outflow<<<nblocks(6L * 2 * L.ghost * L.NJ * L.NK, T), T, 0, s>>>(c);
    }
    mark_stage(P, S_HALO_FACES);
    k_halo_faces<<<nblocks(P.n_halo_faces, T), T, 0, s>>>(
        c, P.halo_faces, P.n_halo_faces, (P.fidelity || !P.old_halo_current) ? 1 : 0);
    if (bd.earth_cells) {
      mark_stage(P, S_EARTH_CELLS);
      k_earth_cells<<<nblocks(6L * 2 * L.ghost * L.NJ * L.NK, T), T, 0, s>>>(c, w);
      mark_stage(P, S_EARTH_FACES);
      for (int d = 0; d < 3; ++d) {
        const long n_out = (L.NI + (d == 0)) - (L.ghost + L.nr + (d == 0));
        const long n = 6L * (L.ghost + n_out) * (L.NJ + (d == 1)) * (L.NK + (d == 2));
        k_earth_faces<<<nblocks(n, T), T, 0, s>>>(c, d, w);
      }
      mark_stage(P, S_EARTH_RECOVER);
      // cubedsphere_earth_boundary.c:169 recovers every cell of both levels.
      // After a completed step the six face fluxes of an active cell are
      // those its B was last recovered from (the step ends with the seam
      // sync and a recovery), so only the non-active cells change; and
      // ghost-cell B is never read by the step (see k_halo_cells), so with
      // fidelity 0 only the first step's recovery of every cell runs.
      for (int lev : {c.cur, c.old}) {
        if (!P.b_recovered)
          k_recover<<<nblocks(n_cells, T), T, 0, s>>>(c, plane_phi(lev, 0), plane_b(lev, 0), 1,
                                                      0);
        else if (P.fidelity)
          k_recover_ghosts<<<nblocks(n_cells - n_active, T), T, 0, s>>>(c, plane_phi(lev, 0),
                                                                        plane_b(lev, 0), 1);
      }
    }
  } else if (phase == PHASE_FLUX) {
    // cubedsphere_step.c:67-98
    mark_stage(P, S_PREDICT);
    if (P.mixed)
      k_predict<true><<<nblocks(n_box, T), T, 0, s>>>(c);
    else
      k_predict<false><<<nblocks(n_box, T), T, 0, s>>>(c);
    for (int d = 0; d < 3; ++d) {
      const long n = 6L * (L.nr + (d == 0)) * (L.n + (d == 1)) * (L.n + (d == 2));
      mark_stage(P, S_FACE_FLUX0 + d);
      // FP64: LDS tiles along i and j, LDS rows along k. The FP32 kernels
      // read half the bytes and are faster without tiles.
      if (P.mixed)
        k_face_flux_fp32<<<nblocks(n, 128), 128, 0, s>>>(c, d);
      else if (P.plain)
        k_face_flux<<<nblocks(n, 128), 128, 0, s>>>(c, d);
      else if (d == 0)
        k_face_flux_tiled<0, FluxTileI::TD, FluxTileI::TT>
            <<<FluxTileI::blocks(L.nr, L.n), FluxTileI::kThreads, 0, s>>>(c);
      else if (d == 1)
        k_face_flux_tiled<1, FluxTileJ::TD, FluxTileJ::TT>
            <<<FluxTileJ::blocks(L.nr, L.n), FluxTileJ::kThreads, 0, s>>>(c);
      else if (L.n + 1 >= kRowMinFaces)
        k_face_flux_rows<<<nblocks(n, kRowFaces), kRowFaces, 0, s>>>(c);
      else
        k_face_flux<<<nblocks(n, 128), 128, 0, s>>>(c, d);
    }
    mark_stage(P, S_SYNC_FLUX);
    if (P.mixed)
      k_sync_flux<float><<<nblocks(P.n_seams, T), T, 0, s>>>(c, P.seams, P.n_seams);
    else
      k_sync_flux<double><<<nblocks(P.n_seams, T), T, 0, s>>>(c, P.seams, P.n_seams);
    if (bd.inner_trap) {
      mark_stage(P, S_INNER_TRAP);
      if (P.mixed)
        k_inner_trap<float><<<nblocks(6L * L.n * L.n, T), T, 0, s>>>(c);
      else
        k_inner_trap<double><<<nblocks(6L * L.n * L.n, T), T, 0, s>>>(c);
    }
  } else if (phase == PHASE_EMF) {
    // cubedsphere_step.c:99-126 up to the seam EMF average; a hydrodynamic
    // step has no EMF
    if (!P.opt.mhd) return 0;
    mark_stage(P, S_FACE_VELOCITY);
    const int act[3] = {L.nr, L.n, L.n};
    // FP64: the i-face velocities for the k-edges only; the j- and i-edges
    // compute their t1-face velocities inside their LDS tiles.
    for (int d = 0; d < (P.mixed || P.plain ? 3 : 1); ++d) {
      long n = 6;
      const int ext[3] = {L.NI, L.NJ, L.NK};
      for (int a = 0; a < 3; ++a)
        n *= a == d ? act[a] + 1 : std::min(L.ghost + act[a] + 4, ext[a]) - (L.ghost - 4);
      if (P.plain && !P.mixed)
        k_face_velocity<<<nblocks(n, 128), 128, 0, s>>>(c, d);
      else if (P.mixed && d == 0 && !P.plain)  // 128 threads: the kernel's launch bound
        k_face_velocity_i_fp32<kFaceVelocityRun><<<nblocks(6L * ((L.nr + kFaceVelocityRun) / kFaceVelocityRun) *
                                                                (n / (6L * (L.nr + 1))), 128), 128, 0, s>>>(c);
      else if (P.mixed)
        k_face_velocity_fp32<<<nblocks(n, T), T, 0, s>>>(c, d);
      else
        // 128 threads: the kernel's launch bound
        k_face_velocity_i<kFaceVelocityRun><<<nblocks(6L * ((L.nr + kFaceVelocityRun) / kFaceVelocityRun) *
                                                           (n / (6L * (L.nr + 1))), 128), 128, 0, s>>>(c);
    }
    for (int e = 0; e < 3; ++e) {
      const long n = 6L * (L.nr + (e != 0)) * (L.n + (e != 1)) * (L.n + (e != 2));
      mark_stage(P, S_EDGE_EMF0 + e);
      if (P.mixed) {
        k_edge_emf_fp32<<<nblocks(n, 128), 128, 0, s>>>(c, e);
      } else if (e == 2 || P.plain) {
        k_edge_emf<<<nblocks(n, 128), 128, 0, s>>>(c, e);
      } else {
        const int t1 = (e + 1) % 3, t2 = (e + 2) % 3;
        const int e1 = e == 0 ? EmfTile<0>::E1 : EmfTile<1>::E1;
        const int e2 = e == 0 ? EmfTile<0>::E2 : EmfTile<1>::E2;
        const long tiles = 6L * act[e] * ((act[t1] + 1 + e1 - 1) / e1) *
                           ((act[t2] + 1 + e2 - 1) / e2);
        if (e == 0)
          k_edge_emf_tiled<0><<<(unsigned)tiles, 256, 0, s>>>(c);
        else
          k_edge_emf_tiled<1><<<(unsigned)tiles, 256, 0, s>>>(c);
      }
    }
    if (bd.earth_emf) {
      mark_stage(P, S_EARTH_EMF);
      for (int d = 0; d < 3; ++d) {
        const int outer = L.ghost + L.nr;
        const long n = 6L * (std::min(outer + 1, L.NI + (d != 0)) - (outer - 3)) *
                       (L.NJ + (d != 1)) * (L.NK + (d != 2));
        k_earth_emf_outer<<<nblocks(n, T), T, 0, s>>>(c, d, w);
      }
      k_earth_emf_diffusive<<<nblocks(6L * (L.n + 1) * (L.n + 1), T), T, 0, s>>>(c, w);
      k_earth_emf_inner<<<nblocks(6L * (L.NJ + 1) * (L.NK + 1), T), T, 0, s>>>(c);
    }
  } else {
    // cubedsphere_step.c:125-152
    if (P.opt.mhd) {
      mark_stage(P, S_SYNC_EDGES);
      k_sync_edges<<<nblocks(P.n_edge_groups, T), T, 0, s>>>(c, P.edge_groups, P.edge_members,
                                                             P.n_edge_groups);
    }
    mark_stage(P, S_UPDATE);
    // A driver step folds the new state's CFL dt into the slots that the
    // next step's dt rule reads (its seam cells after the final sync).
    unsigned long long *cfl = P.driver_cfl ? P.cfl_slots : nullptr;
    if (cfl != nullptr) k_cfl_reset<<<1, kCflSlots, 0, s>>>(cfl);
    if (P.mixed)
      k_update<true><<<nblocks(n_active, 128), 128, 0, s>>>(c, cfl);
    else
      k_update<false><<<nblocks(n_active, 128), 128, 0, s>>>(c, cfl);
    // Final seam average of the new face flux, then B of the touched cells.
    // A hydrodynamic step changes no face flux (cubedsphere_step.c:142).
    K n = c;
    n.cur = c.old;
    n.old = c.cur;
    if (P.opt.mhd) {
      mark_stage(P, S_FINAL_SYNC);
      k_sync_phi<<<nblocks(P.n_seams, T), T, 0, s>>>(n, P.seams, P.n_seams, 0);
      mark_stage(P, S_SEAM_RECOVER);
      k_recover_list<<<nblocks(P.n_seam_cells, T), T, 0, s>>>(
          n, P.seam_cells, P.plain ? nullptr : P.seam_geom, P.n_seam_cells, plane_phi(n.cur, 0),
          plane_b(n.cur, 0), cfl);
    }
    if (P.fidelity) {
      mark_stage(P, S_FIDELITY);
      k_keep_ghosts<<<nblocks(n_box, T), T, 0, s>>>(c);
      if (P.opt.mhd)
        k_recover_ghosts<<<nblocks(n_cells - n_active, T), T, 0, s>>>(n, plane_phi(n.cur, 0),
                                                                      plane_b(n.cur, 0), 0);
    }
    P.level = 1 - P.level;
    P.host_dirty = true;
    P.rates_stored = c.o.rates != 0;
    P.old_halo_current = P.b_recovered;
    P.b_recovered = true;
    P.cfl_folded = cfl != nullptr;
  }
  GCS_HIP(hipGetLastError());
  return 0;
}

// FP32 copies of the face and edge geometry the mixed form's kernels read
// (geometry and background are fixed after setup, so this runs once per
// problem).
__global__ void k_mixed_geometry(K c) {
  const long t = tid();
  if (t >= c.L.total) return;
  c.sp(M_VOL)[t] = (float)c.p(P_VOL)[t];
  for (int d = 0; d < 3; ++d) {
    const int fb = P_FACE + FACE_PLANES * d, mb = M_FACE + M_FACE_PLANES * d;
    c.sp(mb + MF_AREA)[t] = (float)c.p(fb + FO_AREA)[t];
    for (int e = 0; e < 3; ++e) {
      c.sp(mb + MF_N + e)[t] = (float)c.p(fb + FO_N + e)[t];
      c.sp(mb + MF_T1 + e)[t] = (float)c.p(fb + FO_T1 + e)[t];
      c.sp(mb + MF_T2 + e)[t] = (float)c.p(fb + FO_T2 + e)[t];
      c.sp(mb + MF_B0 + e)[t] = (float)c.p(P_B0F + 3 * d + e)[t];
    }
    c.sp(mb + MF_B0N)[t] = (float)c.p(P_B0N + d)[t];
    const int eb = P_EDGE + EDGE_PLANES * d, me = M_EDGE + M_EDGE_PLANES * d;
    c.sp(me + ME_LEN)[t] = (float)c.p(eb + EO_LEN)[t];
    for (int e = 0; e < 3; ++e) {
      c.sp(me + ME_T1 + e)[t] = (float)c.p(eb + EO_T1 + e)[t];
      c.sp(me + ME_T2 + e)[t] = (float)c.p(eb + EO_T2 + e)[t];
    }
    for (int m = 0; m < 5; ++m) c.sp(me + ME_SOLVE + m)[t] = (float)c.p(eb + EO_SOLVE + m)[t];
    for (int m = 0; m < 2; ++m) c.sp(me + ME_B0T + m)[t] = (float)c.p(P_B0ET + 2 * d + m)[t];
  }
}

__global__ void k_widen(const float *src, double *dst, long n) {
  const long t = tid();
  if (t < n) dst[t] = (double)src[t];
}

// max over active cells of |net total (B1 + B0) face flux| / volume, the
// reference driver's divergence gate (test_cubedsphere_dipole_background.c:
// 808-845); non-negative doubles order like their bit patterns.
// Reductions over the active cells run grid-stride with at most
// kReduceBlocks blocks of 256 threads and fold each block in LDS before one
// atomic per block: atomics of every wave (or thread) on one address
// serialise at the L2.
constexpr unsigned kReduceBlocks = 2048;

// Block extremum of u64 values (IEEE bits of non-negative doubles order
// like the doubles) into *out: want_max selects max, else min. Every thread
// of the 256-thread block calls it.
GCS_DEV void block_extremum_atomic(unsigned long long v, bool want_max,
                                   unsigned long long *out) {
  __shared__ unsigned long long s[4];
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_down(v, off, 64);
    v = want_max ? (o > v ? o : v) : (o < v ? o : v);
  }
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) v = want_max ? (s[w] > v ? s[w] : v) : (s[w] < v ? s[w] : v);
    if (want_max)
      atomicMax(out, v);
    else
      atomicMin(out, v);
  }
  __syncthreads();  // s is reused by a second call in the same kernel
}

__global__ void __launch_bounds__(256) k_div_b(K c, unsigned long long *out) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr, L.ghost + L.n, L.ghost + L.n};
  const long step[3] = {L.di(), (long)L.SK, 1};
  const long n = 6L * L.nr * L.n * L.n;
  unsigned long long best = 0;
  for (long t = tid(); t < n; t += (long)gridDim.x * blockDim.x) {
    int f, i, j, k;
    box(t, lo, hi, f, i, j, k);
    const long at = L.at(f, i, j, k);
    double net = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double *phi = c.p(plane_phi(c.cur, d)), *b0 = c.p(P_B0FLUX + d);
      net += (phi[at + step[d]] + b0[at + step[d]]) - (phi[at] + b0[at]);
    }
    const double v = fabs(net) / c.p(P_VOL)[at];
    const unsigned long long bits = __double_as_longlong(isfinite(v) ? v : DBL_MAX);
    best = bits > best ? bits : best;
  }
  block_extremum_atomic(best, true, out);
}

int prepare_mixed(Problem &P) {
  if (P.planes32 != nullptr) return 0;
  const size_t bytes = (size_t)PLANES32 * (size_t)P.L.total * sizeof(float);
  if (hipMalloc((void **)&P.planes32_alloc, bytes + plane_shift32(P.L) * sizeof(float)) !=
      hipSuccess) {
    P.planes32_alloc = P.planes32 = nullptr;
    set_error("mixed form: cannot allocate " + std::to_string(bytes >> 20) +
              " MiB of FP32 planes");
    return -1;
  }
  P.planes32 = P.planes32_alloc + plane_shift32(P.L);
  GCS_HIP(hipMemsetAsync(P.planes32, 0, bytes, P.stream));
  k_mixed_geometry<<<nblocks(P.L.total, 256), 256, 0, P.stream>>>(make_k(P));
  GCS_HIP(hipGetLastError());
  GCS_HIP(hipStreamSynchronize(P.stream));
  return 0;
}

int widen_mixed(Problem &P) {
  if (P.planes32 == nullptr) return 0;
  const long n = P.L.total;
  // Contiguous runs with the same order in both plane tables.
  const int runs[4][3] = {{M_UPRED, P_UPRED, 5}, {M_BPRED, P_BPRED, 3},
                          {M_FLUID, P_FLUID, 15}, {M_MAXWELL, P_MAXWELL, 9}};
  for (const auto &r : runs)
    k_widen<<<nblocks(n * r[2], 256), 256, 0, P.stream>>>(P.planes32 + (long)r[0] * n,
                                                         P.plane(r[1]), n * r[2]);
  GCS_HIP(hipGetLastError());
  GCS_HIP(hipStreamSynchronize(P.stream));
  return 0;
}

// gamera_no_apply_pressure_control (NO/nonorthogonal_step.c:74-171) on the
// active cells of the current level: a cell at or below the low-density
// threshold takes the pressure of the given sound speed; a cell whose sound
// speed exceeds 1.5 c relaxes toward that sound speed's pressure over a
// dipole bounce time. acc = {low-density cells, high-sound-speed cells,
// maximum sound speed bits, maximum pressure bits}, reduced per wave (the
// counts are sums and the maxima exact, so the order does not matter).
struct PressureControlArgs {
  double dt, gamma, density_floor, pressure_floor, light_speed, low_density_threshold,
      low_density_sound_speed;
};

// One active cell (thread t of the active box) of the pressure control:
// updates its U in place and reports whether each branch ran, its sound
// speed and its pressure (zeros for a thread outside the box).
GCS_DEV void pressure_control_cell(const K &c, const PressureControlArgs &a, long t,
                                   bool &low, bool &hot, double &sound, double &pressure) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr, L.ghost + L.n, L.ghost + L.n};
  int f, i, j, k;
  low = hot = false;
  sound = pressure = 0.0;
  if (!box(t, lo, hi, f, i, j, k)) return;
  const long at = L.at(f, i, j, k);
  double u[5];
  for (int q = 0; q < 5; ++q) u[q] = c.p(plane_u(c.cur, q))[at];
  Prim w;
  if (!cons_to_prim(u, a.gamma, a.density_floor, a.pressure_floor, w)) {
    c.fail(81, at);
    return;
  }
  bool changed = false, ok = true;
  if (w.rho <= a.low_density_threshold) {
    w.p = fmax(a.pressure_floor,
               w.rho * a.low_density_sound_speed * a.low_density_sound_speed / a.gamma);
    low = changed = true;
  }
  const double cs = sqrt(a.gamma * w.p / w.rho);
  sound = cs;
  pressure = w.p;
  if (cs > 1.5 * a.light_speed) {
    const double x = c.p(P_CEN)[at], y = c.p(P_CEN + 1)[at], z = c.p(P_CEN + 2)[at];
    const double radius_square = x * x + y * y + z * z;
    const double cylindrical_square = x * x + y * y;
    if (!(radius_square > 0.0) || !(cylindrical_square > DBL_MIN)) {
      c.fail(82, at);
      ok = false;
    } else {
      const double radius = sqrt(radius_square);
      const double dipole_l = radius * radius_square / cylindrical_square;
      const double target = w.rho * (1.5 * a.light_speed) * (1.5 * a.light_speed) / a.gamma;
      const double bounce_time = dipole_l / cs;
      w.p -= (a.dt / bounce_time) * (w.p - target);
      hot = changed = true;
    }
  }
  if (ok && changed) {
    if (!prim_to_cons(w, a.gamma, a.density_floor, a.pressure_floor, u))
      c.fail(83, at);
    else
      for (int q = 0; q < 5; ++q) c.p(plane_u(c.cur, q))[at] = u[q];
  }
}

// A wave's (low-density count, high-sound-speed count, maximum sound speed
// bits, maximum pressure bits), complete in lane 0.
GCS_DEV void pressure_control_wave(bool low, bool hot, double sound, double pressure,
                                   unsigned long long out[4]) {
  out[0] = __popcll(__ballot(low));
  out[1] = __popcll(__ballot(hot));
  unsigned long long s_bits = __double_as_longlong(sound), p_bits = __double_as_longlong(pressure);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long so = __shfl_down(s_bits, off, 64), po = __shfl_down(p_bits, off, 64);
    s_bits = so > s_bits ? so : s_bits;
    p_bits = po > p_bits ? po : p_bits;
  }
  out[2] = s_bits;
  out[3] = p_bits;
}


__global__ void k_zero4(unsigned long long *acc) { acc[threadIdx.x] = 0ull; }


// ---- pressure control without device allocation ---------------------------------
//
// gcs_problem_pressure_control_noalloc: the same cells and results as
// pressure_control, for callers that cannot allocate device memory. Its
// scratch is this file-scope array, one copy per GPU for the life of the
// process: 4 words of partial results per block of 256 active cells, then
// the 4 words of the total. Every block writes its own partials before the
// total is formed, so nothing needs clearing between calls; calls that
// overlap on one GPU share the array and race.
constexpr int kScratchBlocks = 65536;  // at most 16,777,216 active cells
static __device__ __align__(256) unsigned long long scratch[4 * (kScratchBlocks + 1)];

__global__ void __launch_bounds__(256) k_pressure_control_blocks(K c, PressureControlArgs a,
                                                                 unsigned long long *partial) {
  bool low, hot;
  double sound, pressure;
  pressure_control_cell(c, a, tid(), low, hot, sound, pressure);
  unsigned long long r[4];
  pressure_control_wave(low, hot, sound, pressure, r);
  __shared__ unsigned long long waves[4][4];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 4; ++q) waves[wave][q] = r[q];
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    unsigned long long v = waves[0][q];
    for (int w = 1; w < 4; ++w) v = q < 2 ? v + waves[w][q] : (waves[w][q] > v ? waves[w][q] : v);
    partial[4L * blockIdx.x + q] = v;
  }
}

// One block of 256 threads folds n blocks' partials into total[0..3].
__global__ void __launch_bounds__(256) k_pressure_control_total(
    const unsigned long long *partial, long n, unsigned long long *total) {
  unsigned long long r[4] = {0, 0, 0, 0};
  for (long b = threadIdx.x; b < n; b += blockDim.x) {
    r[0] += partial[4 * b];
    r[1] += partial[4 * b + 1];
    r[2] = partial[4 * b + 2] > r[2] ? partial[4 * b + 2] : r[2];
    r[3] = partial[4 * b + 3] > r[3] ? partial[4 * b + 3] : r[3];
  }
  for (int off = 32; off > 0; off >>= 1) {
    r[0] += __shfl_down(r[0], off, 64);
    r[1] += __shfl_down(r[1], off, 64);
    const unsigned long long so = __shfl_down(r[2], off, 64), po = __shfl_down(r[3], off, 64);
    r[2] = so > r[2] ? so : r[2];
    r[3] = po > r[3] ? po : r[3];
  }
  __shared__ unsigned long long waves[4][4];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 4; ++q) waves[wave][q] = r[q];
  __syncthreads();
  if (threadIdx.x < 4) {
    const int q = threadIdx.x;
    unsigned long long v = waves[0][q];
    for (int w = 1; w < 4; ++w) v = q < 2 ? v + waves[w][q] : (waves[w][q] > v ? waves[w][q] : v);
    total[q] = v;
  }
}

namespace {
// Device address of `scratch` on each GPU, looked up once per device with
// that device current (host code never passes the array's name itself).
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_lock;
unsigned long long *g_scratch[kMaxScratchDevices] = {};

int device_scratch(unsigned long long **out) {
  int dev = -1;
  GCS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) {
    set_error("pressure control (no allocation): device index out of range");
    return -1;
  }
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if (g_scratch[dev] == nullptr) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
      set_error(std::string("pressure control (no allocation): scratch address: ") +
                hipGetErrorString(e));
      return -1;
    }
    g_scratch[dev] = static_cast<unsigned long long *>(ptr);
  }
  *out = g_scratch[dev];
  return 0;
}
}  // namespace

int pressure_control_noalloc(Problem &P, const double args[7], unsigned long long counts[2],
                             double maxima[2]) {
  const PressureControlArgs a{args[0], args[1], args[2], args[3], args[4], args[5], args[6]};
  if (!(a.dt > 0.0) || !(a.gamma > 1.0) || !(a.density_floor > 0.0) ||
      !(a.pressure_floor > 0.0) || !(a.light_speed > 0.0) ||
      !(a.low_density_threshold > 0.0) || !(a.low_density_sound_speed >= 0.0)) {
    set_error("pressure control: invalid arguments");
    return -1;
  }
  const long cells = 6L * P.L.nr * P.L.n * P.L.n;
  const long blocks = (cells + 255) / 256;
  if (blocks > kScratchBlocks) {
    set_error("pressure control (no allocation): " + std::to_string(cells) +
              " active cells exceed the scratch limit of " +
              std::to_string(256L * kScratchBlocks) + " (use gcs_problem_pressure_control)");
    return -1;
  }
  unsigned long long *ptr = nullptr, host[4] = {0, 0, 0, 0};
  if (device_scratch(&ptr) != 0) return -1;
  unsigned long long *total = ptr + 4L * kScratchBlocks;
  k_pressure_control_blocks<<<(unsigned)blocks, 256, 0, P.stream>>>(make_k(P), a, ptr);
  k_pressure_control_total<<<1, 256, 0, P.stream>>>(ptr, blocks, total);
  const hipError_t e = hipGetLastError();
  const hipError_t m = hipMemcpyAsync(host, total, sizeof(host), hipMemcpyDeviceToHost, P.stream);
  const hipError_t w = hipStreamSynchronize(P.stream);
  if (e != hipSuccess || m != hipSuccess || w != hipSuccess) {
    set_error("pressure control (no allocation): device error");
    return -1;
  }
  P.cfl_folded = false;
  P.host_dirty = true;
  counts[0] = host[0];
  counts[1] = host[1];
  std::memcpy(&maxima[0], &host[2], sizeof(double));
  std::memcpy(&maxima[1], &host[3], sizeof(double));
  return 0;
}

// The allocating entry point: the same two kernels as the no-allocation one,
// with the block partials in a buffer of the problem (allocated on first use).
int pressure_control(Problem &P, const double args[7], unsigned long long counts[2],
                     double maxima[2]) {
  const PressureControlArgs a{args[0], args[1], args[2], args[3], args[4], args[5], args[6]};
  if (!(a.dt > 0.0) || !(a.gamma > 1.0) || !(a.density_floor > 0.0) ||
      !(a.pressure_floor > 0.0) || !(a.light_speed > 0.0) ||
      !(a.low_density_threshold > 0.0) || !(a.low_density_sound_speed >= 0.0)) {
    set_error("pressure control: invalid arguments");
    return -1;
  }
  const long cells = 6L * P.L.nr * P.L.n * P.L.n;
  const long blocks = (cells + 255) / 256;
  if (P.pc_partial == nullptr) {
    if (hipMalloc(&P.pc_partial, 4 * (blocks + 1) * sizeof(unsigned long long)) != hipSuccess) {
      P.pc_partial = nullptr;
      set_error("pressure control: device allocation failed");
      return -1;
    }
  }
  unsigned long long host[4] = {0, 0, 0, 0}, *total = P.pc_partial + 4 * blocks;
  k_pressure_control_blocks<<<(unsigned)blocks, 256, 0, P.stream>>>(make_k(P), a, P.pc_partial);
  k_pressure_control_total<<<1, 256, 0, P.stream>>>(P.pc_partial, blocks, total);
  const hipError_t e = hipGetLastError();
  const hipError_t m = hipMemcpyAsync(host, total, sizeof(host), hipMemcpyDeviceToHost, P.stream);
  const hipError_t w = hipStreamSynchronize(P.stream);
  if (e != hipSuccess || m != hipSuccess || w != hipSuccess) {
    set_error("pressure control: device error");
    return -1;
  }
  // The state changed: the next step's dt is no longer the folded one.
  P.cfl_folded = false;
  P.host_dirty = true;
  counts[0] = host[0];
  counts[1] = host[1];
  std::memcpy(&maxima[0], &host[2], sizeof(double));
  std::memcpy(&maxima[1], &host[3], sizeof(double));
  return 0;
}

// The Earth driver's end-of-run gate (test_cubedsphere_dipole_background.c:
// 807-893): minimum density and pressure of the active cells' primitive
// state (conversion floors 1e-12) and the maximum total div B. acc = {min
// density bits, min pressure bits, -, -}; the div B maximum goes through
// k_div_b. Positive doubles order like their bit patterns, so min and max
// are exact in any order.
__global__ void k_gate_init(unsigned long long *acc) {
  acc[threadIdx.x] = threadIdx.x < 2 ? __double_as_longlong(DBL_MAX) : 0ull;
}

__global__ void __launch_bounds__(256) k_gates(K c, double gamma, unsigned long long *acc) {
  const Layout &L = c.L;
  const int lo[3] = {L.ghost, L.ghost, L.ghost};
  const int hi[3] = {L.ghost + L.nr, L.ghost + L.n, L.ghost + L.n};
  const long n = 6L * L.nr * L.n * L.n;
  double rho = DBL_MAX, p = DBL_MAX;
  for (long t = tid(); t < n; t += (long)gridDim.x * blockDim.x) {
    int f, i, j, k;
    box(t, lo, hi, f, i, j, k);
    const long at = L.at(f, i, j, k);
    double u[5];
    for (int q = 0; q < 5; ++q) u[q] = c.p(plane_u(c.cur, q))[at];
    Prim w;
    if (!cons_to_prim(u, gamma, 1e-12, 1e-12, w)) {
      c.fail(91, at);
    } else {
      rho = fmin(rho, w.rho);
      p = fmin(p, w.p);
    }
  }
  block_extremum_atomic(__double_as_longlong(rho), false, &acc[0]);
  block_extremum_atomic(__double_as_longlong(p), false, &acc[1]);
}

int mi_point(const MiView &view, int what, const double x[3], const double v[3],
             double *result, double out[3]) {
  if (view.potential == nullptr) return -1;
  GCS_HIP(hipDeviceSynchronize());  // the potential is written on a problem's stream
  k_mi_point<<<1, 1>>>(view, what, x[0], x[1], x[2], v[0], v[1], v[2], result);
  GCS_HIP(hipGetLastError());
  double host[4];
  GCS_HIP(hipMemcpy(host, result, sizeof(host), hipMemcpyDeviceToHost));
  if (host[3] == 0.0) return -1;
  for (int c = 0; c < 3; ++c) out[c] = host[c];
  return 0;
}

int final_gates(Problem &P, double gamma, double out[3]) {
  if (!(gamma > 1.0)) {
    set_error("final gates: gamma must exceed 1");
    return -1;
  }
  unsigned long long *acc = P.acc4, host[4] = {0, 0, 0, 0};
  K c = make_k(P);
  const long n = 6L * P.L.nr * P.L.n * P.L.n;
  k_gate_init<<<1, 4, 0, P.stream>>>(acc);
  const unsigned blocks = std::min(kReduceBlocks, nblocks(n, 256));
  k_gates<<<blocks, 256, 0, P.stream>>>(c, gamma, acc);
  k_div_b<<<blocks, 256, 0, P.stream>>>(c, acc + 2);
  const hipError_t e = hipGetLastError();
  const hipError_t m = hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, P.stream);
  const hipError_t w = hipStreamSynchronize(P.stream);
  if (e != hipSuccess || m != hipSuccess || w != hipSuccess) {
    set_error("final gates: device error");
    return -1;
  }
  std::memcpy(&out[0], &host[0], sizeof(double));
  std::memcpy(&out[1], &host[1], sizeof(double));
  std::memcpy(&out[2], &host[2], sizeof(double));
  return 0;
}

// ---- gamera_no_local_timestep -----------------------------------------------------

// nonorthogonal_step.c:480-515 for one cell: the CFL dt from the conserved
// state u, the residual cell B b and background B b0 (read only as the
// options say) and the three CFL lengths; false where the reference fails.
GCS_DEV bool local_dt_cell(const double u[5], const double b[3], const double b0[3],
                           const double len[3], const LocalDtOptions &o, double &dt) {
  Prim w;
  if (!cons_to_prim(u, o.gamma, o.density_floor, o.pressure_floor, w)) return false;
  const double cs = sqrt(o.gamma * w.p / w.rho);
  double va = 0.0;
  if (o.mhd) {
    double m[3];
    for (int d = 0; d < 3; ++d) m[d] = o.background ? b[d] + b0[d] : b[d];
    va = sqrt(dot3(m, m)) / sqrt(w.rho);
    if (o.boris) va = o.light_speed * va / hypot(o.light_speed, va);
  }
  const double speed = sqrt(dot3(w.v, w.v)) + hypot(cs, va);
  double l = len[0];
  l = fmin(l, len[1]);
  l = fmin(l, len[2]);
  dt = o.cfl * l / speed;
  return isfinite(dt) && dt > 0.0;
}

// Block minimum of dt (as IEEE bits: every candidate is positive) into
// acc[0] and any failure into acc[1], one atomic per block: atomics of every
// wave on one address serialise at the L2. Every thread of the 256-thread
// block calls it.
constexpr int kLocalDtThreads = 256;
constexpr unsigned kLocalDtBlocks = 2048;  // grid-stride: 8 blocks per CU

GCS_DEV void local_dt_fold(double dt, bool bad, unsigned long long *acc) {
  __shared__ unsigned long long s_min[kLocalDtThreads / 64];
  __shared__ int s_bad[kLocalDtThreads / 64];
  unsigned long long bits = __double_as_longlong(dt);
  int any = bad ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long other = __shfl_down(bits, off, 64);
    bits = other < bits ? other : bits;
    any |= __shfl_down(any, off, 64);
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    s_min[wave] = bits;
    s_bad[wave] = any;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kLocalDtThreads / 64; ++w) {
      bits = s_min[w] < bits ? s_min[w] : bits;
      any |= s_bad[w];
    }
    atomicMin(&acc[0], bits);
    if (any) atomicOr(&acc[1], 1ull);
  }
}

__global__ void k_local_dt_init(unsigned long long *acc) {
  acc[0] = __double_as_longlong(DBL_MAX);
  acc[1] = 0ull;
}

// Box cell t -> (i, j, k), k fastest.
struct DtBox {
  int lo[3], n[3];
  GCS_DEV bool cell(long t, int &i, int &j, int &k) const {
    if (t >= (long)n[0] * n[1] * n[2]) return false;
    k = lo[2] + (int)(t % n[2]);
    const long r = t / n[2];
    j = lo[1] + (int)(r % n[1]);
    i = lo[0] + (int)(r / n[1]);
    return true;
  }
};

// Grid-stride over the box's cells; every thread reaches the block fold.
__global__ void __launch_bounds__(kLocalDtThreads) k_local_dt_planes(K c, int f, DtBox bx,
                                                                     LocalDtOptions o,
                                                                     unsigned long long *acc) {
  double best = DBL_MAX;
  bool bad = false;
  const long n = (long)bx.n[0] * bx.n[1] * bx.n[2];
  for (long t = tid(); t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    bx.cell(t, i, j, k);
    const long at = c.L.at(f, i, j, k);
    double u[5], b[3] = {0.0, 0.0, 0.0}, b0[3] = {0.0, 0.0, 0.0}, len[3];
    for (int q = 0; q < 5; ++q) u[q] = c.p(plane_u(c.cur, q))[at];
    for (int d = 0; d < 3; ++d) {
      if (o.mhd) b[d] = c.p(plane_b(c.cur, d))[at];
      if (o.mhd && o.background) b0[d] = c.p(P_B0C + d)[at];
      len[d] = c.p(P_CFL + d)[at];
    }
    double dt;
    if (local_dt_cell(u, b, b0, len, o, dt))
      best = fmin(best, dt);
    else
      bad = true;
  }
  local_dt_fold(best, bad, acc);
}

__global__ void __launch_bounds__(kLocalDtThreads) k_local_dt_arrays(
    const double *u, const double *b, const double *b0, const double *len, long ej, long ek,
    DtBox bx, LocalDtOptions o, unsigned long long *acc) {
  double best = DBL_MAX;
  bool bad = false;
  const long n = (long)bx.n[0] * bx.n[1] * bx.n[2];
  for (long t = tid(); t < n; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    bx.cell(t, i, j, k);
    const long h = ((long)i * ej + j) * ek + k;  // gamera_no_index3
    double uc[5], bc[3] = {0.0, 0.0, 0.0}, b0c[3] = {0.0, 0.0, 0.0}, lc[3];
    for (int q = 0; q < 5; ++q) uc[q] = u[5 * h + q];
    for (int d = 0; d < 3; ++d) {
      if (o.mhd) bc[d] = b[3 * h + d];
      if (o.mhd && o.background) b0c[d] = b0[3 * h + d];
      lc[d] = len[3 * t + d];
    }
    double dt;
    if (local_dt_cell(uc, bc, b0c, lc, o, dt))
      best = fmin(best, dt);
    else
      bad = true;
  }
  local_dt_fold(best, bad, acc);
}

namespace {
int finish_local_dt(unsigned long long *acc, hipStream_t s, double *dt, bool *invalid) {
  unsigned long long host[2] = {0, 0};
  const hipError_t e = hipGetLastError();
  const hipError_t m = hipMemcpyAsync(host, acc, sizeof(host), hipMemcpyDeviceToHost, s);
  const hipError_t w = hipStreamSynchronize(s);
  if (e != hipSuccess || m != hipSuccess || w != hipSuccess) {
    set_error("local timestep: device error");
    return -1;
  }
  std::memcpy(dt, &host[0], sizeof(double));
  *invalid = host[1] != 0;
  return 0;
}

DtBox dt_box(const int lo[3], const int hi[3]) {
  DtBox bx;
  for (int a = 0; a < 3; ++a) {
    bx.lo[a] = lo[a];
    bx.n[a] = hi[a] - lo[a];
  }
  return bx;
}
}  // namespace

int local_dt_planes(Problem &P, int face, const int lo[3], const int hi[3],
                    const LocalDtOptions &o, double *dt, bool *invalid) {
  const DtBox bx = dt_box(lo, hi);
  const long n = (long)bx.n[0] * bx.n[1] * bx.n[2];
  unsigned long long *acc = P.acc4;
  const unsigned blocks = (unsigned)std::min<long>(kLocalDtBlocks, nblocks(n, kLocalDtThreads));
  k_local_dt_init<<<1, 1, 0, P.stream>>>(acc);
  k_local_dt_planes<<<blocks, kLocalDtThreads, 0, P.stream>>>(make_k(P), face, bx, o, acc);
  return finish_local_dt(acc, P.stream, dt, invalid);
}

int local_dt_arrays(const double *u, const double *b, const double *b0, const double *len,
                    const long ext[3], const int lo[3], const int hi[3],
                    const LocalDtOptions &o, unsigned long long *acc, double *dt,
                    bool *invalid) {
  const DtBox bx = dt_box(lo, hi);
  const long n = (long)bx.n[0] * bx.n[1] * bx.n[2];
  const unsigned blocks = (unsigned)std::min<long>(kLocalDtBlocks, nblocks(n, kLocalDtThreads));
  k_local_dt_init<<<1, 1, 0, nullptr>>>(acc);
  k_local_dt_arrays<<<blocks, kLocalDtThreads, 0, nullptr>>>(u, b, b0, len, ext[1], ext[2], bx,
                                                             o, acc);
  return finish_local_dt(acc, nullptr, dt, invalid);
}

int max_div_b(Problem &P, double *out) {
  unsigned long long *dev = P.acc4, host = 0;
  k_zero4<<<1, 4, 0, P.stream>>>(dev);
  K c = make_k(P);
  k_div_b<<<std::min(kReduceBlocks, nblocks(6L * P.L.nr * P.L.n * P.L.n, 256)), 256, 0,
            P.stream>>>(c, dev);
  const hipError_t e = hipGetLastError();
  const hipError_t m = hipMemcpyAsync(&host, dev, sizeof(host), hipMemcpyDeviceToHost, P.stream);
  const hipError_t w = hipStreamSynchronize(P.stream);
  if (e != hipSuccess || m != hipSuccess || w != hipSuccess) {
    set_error("divergence check failed");
    return -1;
  }
  std::memcpy(out, &host, sizeof(double));
  return 0;
}

int enqueue_step(Problem &P, bool explicit_dt) {
  if (P.mixed && P.planes32 == nullptr) {
    set_error("mixed form selected without its FP32 planes");
    return -1;
  }
  P.last_step_mixed = P.mixed;
  K c = make_k(P);
  const Layout &L = P.L;
  hipStream_t s = P.stream;
  const long n_active = 6L * L.nr * L.n * L.n;
  if (!explicit_dt) {
    if (P.cfl_folded) {
      // the previous step's update folded this state's CFL dt
      mark_stage(P, S_DT);
      k_dt<<<1, kCflThreads, 0, s>>>(c, P.cfg.dt_fraction, P.cfl_slots, kCflSlots);
    } else {
      mark_stage(P, S_CFL);
      const int ncfl = (int)nblocks(n_active, kCflThreads);
      k_cfl<<<ncfl, kCflThreads, 0, s>>>(c, P.cfl_partial);
      mark_stage(P, S_DT);
      k_dt<<<1, kCflThreads, 0, s>>>(c, P.cfg.dt_fraction, P.cfl_partial, ncfl);
    }
  }
  for (Phase ph : {PHASE_GHOSTS, PHASE_FLUX, PHASE_EMF, PHASE_UPDATE})
    if (enqueue_phase(P, ph) != 0) return -1;
  if (!explicit_dt) {
    mark_stage(P, S_CLOCK);
    k_clock<<<1, 1, 0, s>>>(c);
  }
  mark_stage(P, S_END);
  GCS_HIP(hipGetLastError());
  return 0;
}

}  // namespace gcs

// ---- measurement: a plain copy kernel ----------------------------------------------

namespace gcs {

// Each thread copies four 16-byte elements 256 apart (every wave's accesses
// contiguous), all four loads issued before the first store.
constexpr int kCopyUnroll = 4;

__global__ void __launch_bounds__(256) k_copy(const double2 *__restrict__ a,
                                              double2 *__restrict__ b, long n) {
  const long base = (long)blockIdx.x * (256 * kCopyUnroll) + threadIdx.x;
  double2 v[kCopyUnroll];
#pragma unroll
  for (int r = 0; r < kCopyUnroll; ++r) {
    const long t = base + 256L * r;
    if (t < n) v[r] = a[t];
  }
#pragma unroll
  for (int r = 0; r < kCopyUnroll; ++r) {
    const long t = base + 256L * r;
    if (t < n) b[t] = v[r];
  }
}

int copy_bandwidth(size_t doubles, int iterations, double *gb_per_s) {
  if (doubles < 2 || iterations < 1 || gb_per_s == nullptr) {
    set_error("copy bandwidth: need at least 2 doubles and 1 iteration");
    return -1;
  }
  const long n2 = (long)(doubles / 2);
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t start = nullptr, stop = nullptr;
  int status = -1;
  float ms = 0.0f;
  if (hipMalloc(&a, n2 * sizeof(double2)) != hipSuccess ||
      hipMalloc(&b, n2 * sizeof(double2)) != hipSuccess ||
      hipEventCreate(&start) != hipSuccess || hipEventCreate(&stop) != hipSuccess) {
    set_error("copy bandwidth: allocation failed");
  } else {
    const unsigned blocks = (unsigned)((n2 + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll));
    k_copy<<<blocks, 256>>>(a, b, n2);  // first touch, not timed
    (void)hipEventRecord(start, nullptr);
    for (int r = 0; r < iterations; ++r) k_copy<<<blocks, 256>>>(a, b, n2);
    (void)hipEventRecord(stop, nullptr);
    if (hipGetLastError() == hipSuccess && hipEventSynchronize(stop) == hipSuccess &&
        hipEventElapsedTime(&ms, start, stop) == hipSuccess && ms > 0.0f) {
      *gb_per_s = 2.0 * sizeof(double2) * (double)n2 * iterations / (ms * 1e-3) / 1e9;
      status = 0;
    } else {
      set_error("copy bandwidth: device error");
    }
  }
  if (start) (void)hipEventDestroy(start);
  if (stop) (void)hipEventDestroy(stop);
  (void)hipFree(a);
  (void)hipFree(b);
  return status;
}

}  // namespace gcs

// ---- the drivers' CSV snapshot (the rows' values) ------------------------------------

namespace gcs {

// One cube face's snapshot rows (test_cubedsphere_dipole_background.c:128-150):
// per active cell in i, j, k order, 11 doubles: centroid, density,
// pressure and velocity with floors 1e-12, total cell B (background +
// residual). The index of the first cell whose state does not convert is
// folded into *first_bad.
__global__ void __launch_bounds__(256) k_snapshot_face(K c, int f, double gamma, double *rows,
                                                      unsigned long long *first_bad) {
  const Layout &L = c.L;
  const long n = (long)L.nr * L.n * L.n;
  const long t = tid();
  if (t >= n) return;
  const int k = L.ghost + (int)(t % L.n);
  const int j = L.ghost + (int)((t / L.n) % L.n);
  const int i = L.ghost + (int)(t / ((long)L.n * L.n));
  const long at = L.at(f, i, j, k);
  double u[5];
  for (int q = 0; q < 5; ++q) u[q] = c.p(plane_u(c.cur, q))[at];
  Prim w;
  if (!cons_to_prim(u, gamma, 1e-12, 1e-12, w)) {
    atomicMin(first_bad, (unsigned long long)t);
    return;
  }
  double *r = rows + 11 * t;
  for (int m = 0; m < 3; ++m) r[m] = c.p(P_CEN + m)[at];
  r[3] = w.rho;
  r[4] = w.p;
  for (int m = 0; m < 3; ++m) r[5 + m] = w.v[m];
  for (int m = 0; m < 3; ++m) r[8 + m] = c.p(P_B0C + m)[at] + c.p(plane_b(c.cur, m))[at];
}

int snapshot_face(Problem &P, int face, double gamma, std::vector<double> &rows,
                  long *first_bad) {
  const long n = (long)P.L.nr * P.L.n * P.L.n;
  rows.resize((size_t)n * 11);
  double *dev = nullptr;
  GCS_HIP(hipMalloc(&dev, rows.size() * sizeof(double)));
  unsigned long long *bad = P.acc4, host_bad = 0;
  k_gate_init<<<1, 4, 0, P.stream>>>(bad);  // bad[0] = max, as "none"
  k_snapshot_face<<<nblocks(n, 256), 256, 0, P.stream>>>(make_k(P), face, gamma, dev, bad);
  const hipError_t e = hipGetLastError();
  const hipError_t m1 = hipMemcpyAsync(rows.data(), dev, rows.size() * sizeof(double),
                                       hipMemcpyDeviceToHost, P.stream);
  const hipError_t m2 = hipMemcpyAsync(&host_bad, bad, sizeof(host_bad), hipMemcpyDeviceToHost,
                                       P.stream);
  const hipError_t w = hipStreamSynchronize(P.stream);
  (void)hipFree(dev);
  if (e != hipSuccess || m1 != hipSuccess || m2 != hipSuccess || w != hipSuccess) {
    set_error("snapshot: device error");
    return -1;
  }
  *first_bad = host_bad < (unsigned long long)n ? (long)host_bad : -1;
  return 0;
}

}  // namespace gcs

// ---- loading the drivers' CSV snapshot ------------------------------------------------

namespace gcs {

// Row r of a parsed snapshot (test_cubedsphere_dipole_background.c:61-118):
// idx[4r..] = face and active-relative i, j, k; rows[11r..] = centroid,
// density, pressure, velocity, total cell B. With apply 0 the rows are only
// checked (centroid within 1e-8 of the cell's, primitives convertible with
// floors 1e-12), the first bad row folded into *bad; with apply 1 the
// conserved state and residual cell B of both levels are written.
__global__ void __launch_bounds__(256) k_load_snapshot(K c, const double *rows, const int *idx,
                                                      long n, double gamma, int apply,
                                                      unsigned long long *bad) {
  const long r = tid();
  if (r >= n) return;
  const Layout &L = c.L;
  const int f = idx[4 * r], i = idx[4 * r + 1], j = idx[4 * r + 2], k = idx[4 * r + 3];
  const long at = L.at(f, L.ghost + i, L.ghost + j, L.ghost + k);
  const double *x = rows + 11 * r;
  Prim w;
  w.rho = x[3];
  w.p = x[4];
  for (int m = 0; m < 3; ++m) w.v[m] = x[5 + m];
  double u[5];
  bool ok = prim_to_cons(w, gamma, 1e-12, 1e-12, u);
  for (int m = 0; m < 3; ++m) ok = ok && !(fabs(c.p(P_CEN + m)[at] - x[m]) > 1e-8);
  if (!apply) {
    if (!ok) atomicMin(bad, (unsigned long long)r);
    return;
  }
  for (int q = 0; q < 5; ++q) {
    c.p(plane_u(c.cur, q))[at] = u[q];
    c.p(plane_u(c.old, q))[at] = u[q];
  }
  for (int m = 0; m < 3; ++m) {
    const double residual = x[8 + m] - c.p(P_B0C + m)[at];
    c.p(plane_b(c.cur, m))[at] = residual;
    c.p(plane_b(c.old, m))[at] = residual;
  }
}

int load_snapshot_rows(Problem &P, const std::vector<double> &rows, const std::vector<int> &idx,
                       double gamma, long *first_bad) {
  const long n = (long)idx.size() / 4;
  double *drows = nullptr;
  int *didx = nullptr;
  unsigned long long *bad = P.acc4, host_bad = 0;
  int status = -1;
  if (hipMalloc(&drows, rows.size() * sizeof(double)) == hipSuccess &&
      hipMalloc(&didx, idx.size() * sizeof(int)) == hipSuccess &&
      hipMemcpy(drows, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice) ==
          hipSuccess &&
      hipMemcpy(didx, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice) ==
          hipSuccess) {
    K c = make_k(P);
    k_gate_init<<<1, 4, 0, P.stream>>>(bad);  // bad[0] = "none"
    k_load_snapshot<<<nblocks(n, 256), 256, 0, P.stream>>>(c, drows, didx, n, gamma, 0, bad);
    if (hipMemcpyAsync(&host_bad, bad, sizeof(host_bad), hipMemcpyDeviceToHost, P.stream) ==
            hipSuccess &&
        hipStreamSynchronize(P.stream) == hipSuccess) {
      *first_bad = host_bad < (unsigned long long)n ? (long)host_bad : -1;
      status = 0;
      if (*first_bad < 0) {
        k_load_snapshot<<<nblocks(n, 256), 256, 0, P.stream>>>(c, drows, didx, n, gamma, 1, bad);
        status = hipGetLastError() == hipSuccess && hipStreamSynchronize(P.stream) == hipSuccess
                     ? 0
                     : -1;
      }
    }
  }
  (void)hipFree(drows);
  (void)hipFree(didx);
  if (status != 0) set_error("snapshot load: device error");
  return status;
}

}  // namespace gcs
