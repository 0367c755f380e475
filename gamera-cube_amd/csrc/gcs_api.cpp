// C ABI of the HIP library: the problem API (create / advance / inspect a
// device-resident six-face Earth problem). Declared in
// include/gamera_cs_step.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gcs_device.hpp"
#include "gcs_transfer.hpp"

namespace gcs {

static thread_local std::string g_error;

void set_error(const std::string &message) { g_error = message; }
const char *last_error() { return g_error.c_str(); }

namespace {

int check_config(const gcs_problem_config &c) {
  if (c.ghost < 4 || c.ghost > 16) return set_error("ghost must be 4 to 16"), -1;
  if (c.radial_cells < 2 || c.angular_cells <= 2 * c.ghost)
    return set_error("grid needs nr >= 2 and n > 2 ghost (cubedsphere_mesh.c:188-193)"), -1;
  if (c.wind_count < 1 || c.wind_count > GCS_MAX_WIND_SAMPLES)
    return set_error("wind_count out of range"), -1;
  if (!(c.startup.gamma > 1.0) || !(c.dt_fraction > 0.0) || !(c.cfl > 0.0))
    return set_error("invalid gamma, dt fraction or CFL"), -1;
  const long g2 = 2 * (long)c.ghost + 1;
  const long cells = ((long)c.radial_cells + g2) * ((long)c.angular_cells + g2) *
                     (((long)c.angular_cells + g2 + 7) / 8 * 8) * 6;
  if (cells >= (1L << 31)) return set_error("grid too large for 32-bit indices"), -1;
  return 0;
}

}  // namespace

int upload_wind(Problem &P, int count, const double *time, const double *state) {
  if (count < 1) return set_error("empty wind series"), -1;
  if ((size_t)count > P.wind_capacity) {
    (void)hipFree(P.wind_buffer);
    P.wind_buffer = nullptr;
    P.wind_capacity = 0;
    const size_t cap = std::max<size_t>((size_t)count, 16);
    if (hipMalloc(&P.wind_buffer, cap * 9 * sizeof(double)) != hipSuccess) {
      P.wind_buffer = nullptr;
      return set_error("wind series: device allocation failed"), -1;
    }
    P.wind_capacity = cap;
  }
  // the step works on its own stream; finish what may still read the buffer
  GCS_HIP(hipStreamSynchronize(P.stream));
  GCS_HIP(hipMemcpy(P.wind_buffer, time, (size_t)count * sizeof(double),
                    hipMemcpyHostToDevice));
  GCS_HIP(hipMemcpy(P.wind_buffer + P.wind_capacity, state,
                    (size_t)count * 8 * sizeof(double), hipMemcpyHostToDevice));
  P.wind_series.count = count;
  P.wind_series.time = P.wind_buffer;
  P.wind_series.state = P.wind_buffer + P.wind_capacity;
  return 0;
}

namespace {

// Wind series of the config (solar_wind.c:142-171 with the startup state).
int make_wind(Problem &P) {
  const gcs_problem_config &c = P.cfg;
  const gamera_cs_earth_startup &s = c.startup;
  WindSeries &w = P.wind_series;
  std::memset(&w, 0, sizeof(w));
  w.linear = c.wind_count > 1 ? c.wind_linear : 1;
  std::vector<double> time(c.wind_count), state(8 * c.wind_count);
  for (size_t n = 0; n < c.wind_count; ++n) {
    time[n] = c.wind_time[n];
    double *st = &state[8 * n];
    st[0] = s.density;
    st[1] = s.density * s.speed * s.speed / (s.gamma * s.mach * s.mach);
    st[2] = s.speed;
    st[3] = st[4] = 0.0;
    for (int d = 0; d < 3; ++d) st[5 + d] = s.imf.value[d];
    if (c.wind_count > 1) st[7] = c.wind_bz[n];
  }
  w.reference[0] = c.wind_reference_x;
  return upload_wind(P, (int)c.wind_count, time.data(), state.data());
}

}  // namespace

void free_problem(Problem *P) {
  if (P == nullptr) return;
  mi_destroy(*P);
  for (auto &g : P->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  (void)hipFree(P->planes_alloc);
  if (P->pinned != nullptr) (void)hipHostFree(P->pinned);
  (void)hipFree(P->planes32_alloc);
  (void)hipFree(P->bytes);
  (void)hipFree(P->scalars);
  (void)hipFree(P->cfl_partial);
  (void)hipFree(P->cfl_slots);
  (void)hipFree(P->acc4);
  (void)hipFree(P->pc_partial);
  (void)hipFree(P->halo_cells);
  (void)hipFree(P->halo_faces);
  (void)hipFree(P->seams);
  (void)hipFree(P->edge_members);
  (void)hipFree(P->edge_groups);
  (void)hipFree(P->seam_cells);
  (void)hipFree(P->seam_geom);
  (void)hipFree(P->wind_buffer);
  if (P->ev_start) (void)hipEventDestroy(P->ev_start);
  if (P->ev_stop) (void)hipEventDestroy(P->ev_stop);
  if (P->stream) (void)hipStreamDestroy(P->stream);
  delete P;
}

int allocate_problem(Problem &P) {
  const size_t nbytes = (size_t)PLANES * P.L.total * sizeof(double);
  const size_t nflags = (size_t)BYTE_PLANES * P.L.total;
  if (hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&P.planes_alloc, nbytes + plane_shift(P.L) * sizeof(double)) != hipSuccess) {
    set_error("device allocation failed");
    return -1;
  }
  P.planes = P.planes_alloc + plane_shift(P.L);
  if (hipMalloc(&P.bytes, nflags) != hipSuccess ||
      hipMalloc(&P.scalars, sizeof(StepScalars)) != hipSuccess ||
      hipMalloc(&P.cfl_partial,
                ((6L * P.L.nr * P.L.n * P.L.n + 255) / 256 + 1) * sizeof(unsigned long long)) !=
          hipSuccess ||
      hipMalloc(&P.cfl_slots, kCflSlots * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&P.acc4, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(P.planes, 0, nbytes) != hipSuccess ||
      hipMemset(P.bytes, 0, nflags) != hipSuccess ||
      hipEventCreate(&P.ev_start) != hipSuccess ||
      hipEventCreate(&P.ev_stop) != hipSuccess) {
    set_error("device allocation failed");
    return -1;
  }
  return init_scalars(P);
}

int init_scalars(Problem &P) {
  StepScalars s;
  std::memset(&s, 0, sizeof(s));
  const double big = DBL_MAX;
  std::memcpy(&s.cfl_min_bits, &big, sizeof(big));
  GCS_HIP(hipMemcpy(P.scalars, &s, sizeof(s), hipMemcpyHostToDevice));
  return 0;
}

int read_scalars(Problem &P, StepScalars &s) {
  GCS_HIP(hipStreamSynchronize(P.stream));
  GCS_HIP(hipMemcpy(&s, P.scalars, sizeof(s), hipMemcpyDeviceToHost));
  return 0;
}

int check_failure(Problem &P) {
  StepScalars s;
  if (read_scalars(P, s) != 0) return -1;
  if (s.failed) {
    char msg[160];
    const long per = P.L.face_stride;
    const long at = s.fail_index;
    const int f = (int)(at / per);
    const long r = at - f * per;
    std::snprintf(msg, sizeof(msg),
                  "MHD_FAILURE stage=%u face=%d local=%ld,%ld,%ld", s.fail_stage, f,
                  r / P.L.di(), (r % P.L.di()) / P.L.SK, r % P.L.SK);
    set_error(msg);
    return -1;
  }
  P.time = s.time;
  P.dt = s.dt;
  P.previous_dt = s.previous_dt;
  P.steps = (long)s.steps;
  return 0;
}

namespace {

// Copy one component plane of one cube face to the host in the
// reference's k-fastest order for extent e.
int fetch(Problem &P, int plane, int face, const size_t e[3], int ncomp,
          int comp, std::vector<double> &out) {
  const Layout &L = P.L;
  std::vector<double> buf(L.face_stride);
  GCS_HIP(hipMemcpy(buf.data(), P.plane(plane) + face * L.face_stride,
                    L.face_stride * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < e[0]; ++i)
    for (size_t j = 0; j < e[1]; ++j)
      for (size_t k = 0; k < e[2]; ++k)
        out[((i * e[1] + j) * e[2] + k) * ncomp + comp] =
            buf[(i * L.SJ + j) * L.SK + k];
  return 0;
}

}  // namespace

int download_field(Problem &P, int face, int field, std::vector<double> &out) {
  const Layout &L = P.L;
  const int cur = P.level, old = 1 - P.level;
  size_t ce[3] = {(size_t)L.NI, (size_t)L.NJ, (size_t)L.NK};
  size_t ve[3] = {(size_t)L.NI + 1, (size_t)L.NJ + 1, (size_t)L.NK + 1};
  const size_t ncell = ce[0] * ce[1] * ce[2];
  // The mixed form keeps the predicted state and the face fluxes in FP32.
  const bool widened = field == GCS_FIELD_PREDICTED_CONSERVED ||
                       field == GCS_FIELD_PREDICTED_CELL_MAGNETIC ||
                       (field >= GCS_FIELD_FACE_FLUX + 12 && field < GCS_FIELD_HYDRO_RATE);
  if (widened && P.last_step_mixed && widen_mixed(P) != 0) return -1;
  auto planes = [&](int first, int ncomp, const size_t e[3]) {
    out.assign(e[0] * e[1] * e[2] * ncomp, 0.0);
    for (int q = 0; q < ncomp; ++q)
      if (fetch(P, first + q, face, e, ncomp, q, out) != 0) return -1;
    return 0;
  };
  if (field == GCS_FIELD_CONSERVED) return planes(plane_u(cur, 0), 5, ce);
  if (field == GCS_FIELD_OLD_CONSERVED) return planes(plane_u(old, 0), 5, ce);
  if (field == GCS_FIELD_PREDICTED_CONSERVED) return planes(P_UPRED, 5, ce);
  if (field == GCS_FIELD_CELL_MAGNETIC) return planes(plane_b(cur, 0), 3, ce);
  if (field == GCS_FIELD_OLD_CELL_MAGNETIC) return planes(plane_b(old, 0), 3, ce);
  if (field == GCS_FIELD_PREDICTED_CELL_MAGNETIC) return planes(P_BPRED, 3, ce);
  if (field >= GCS_FIELD_FACE_FLUX && field < GCS_FIELD_HYDRO_RATE) {
    const int group = (field - GCS_FIELD_FACE_FLUX) / 3;
    const int d = (field - GCS_FIELD_FACE_FLUX) % 3;
    size_t fe[3], ee[3];
    face_extent(L, d, fe);
    edge_extent(L, d, ee);
    switch (group) {
      case 0: return planes(plane_phi(cur, d), 1, fe);
      case 1: return planes(plane_phi(old, d), 1, fe);
      case 2: return planes(P_PHIPRED + d, 1, fe);
      case 3: return planes(P_EMF + d, 1, ee);
      case 4: return planes(P_FLUID + 5 * d, 5, fe);
      default: return planes(P_MAXWELL + 3 * d, 3, fe);
    }
  }
  if (field == GCS_FIELD_HYDRO_RATE || field == GCS_FIELD_MAXWELL_RATE) {
    // active cells: what the last fidelity step's update stored in the
    // face-velocity planes; ghost cells (and steps without fidelity): zero
    const int nc = field == GCS_FIELD_HYDRO_RATE ? 5 : 3;
    out.assign(ncell * nc, 0.0);
    if (!P.rates_stored) return 0;
    const size_t lo[3] = {(size_t)L.ghost, (size_t)L.ghost, (size_t)L.ghost};
    const size_t hi[3] = {(size_t)(L.ghost + L.nr), (size_t)(L.ghost + L.n),
                          (size_t)(L.ghost + L.n)};
    return box_planes_to_host(P, nc == 5 ? P_FVEL : P_FVEL + 5, nc, face, ce, lo, hi,
                              out.data());
  }
  if (field == GCS_FIELD_VERTEX) return planes(P_VERT, 3, ve);
  if (field == GCS_FIELD_CELL_GEOMETRY) {
    // gamera_no_cell_geometry: volume, centroid, face[3][2], cfl_length[3].
    const int per = sizeof(gamera_no_cell_geometry) / sizeof(double);
    out.assign(ncell * per, 0.0);
    std::vector<double> tmp(ncell);
    auto put = [&](int plane, int slot, int di, int dj, int dk) {
      const Layout &LL = P.L;
      std::vector<double> buf(LL.face_stride);
      if (hipMemcpy(buf.data(), P.plane(plane) + face * LL.face_stride,
                    LL.face_stride * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
      for (size_t i = 0; i < ce[0]; ++i)
        for (size_t j = 0; j < ce[1]; ++j)
          for (size_t k = 0; k < ce[2]; ++k)
            out[((i * ce[1] + j) * ce[2] + k) * per + slot] =
                buf[((i + di) * LL.SJ + j + dj) * LL.SK + k + dk];
      return 0;
    };
    if (put(P_VOL, 0, 0, 0, 0) != 0) return -1;
    for (int e = 0; e < 3; ++e)
      if (put(P_CEN + e, 1 + e, 0, 0, 0) != 0) return -1;
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s)
        for (int q = 0; q < FACE_PLANES; ++q)
          if (put(P_FACE + FACE_PLANES * d + q, 4 + (2 * d + s) * FACE_PLANES + q,
                  s * (d == 0), s * (d == 1), s * (d == 2)) != 0)
            return -1;
    for (int d = 0; d < 3; ++d)
      if (put(P_CFL + d, 4 + 6 * FACE_PLANES + d, 0, 0, 0) != 0) return -1;
    return 0;
  }
  if (field >= GCS_FIELD_FACE_GEOMETRY && field < GCS_FIELD_FACE_GEOMETRY + 3) {
    const int d = field - GCS_FIELD_FACE_GEOMETRY;
    size_t fe[3];
    face_extent(L, d, fe);
    return planes(P_FACE + FACE_PLANES * d, FACE_PLANES, fe);
  }
  if (field >= GCS_FIELD_EDGE_GEOMETRY && field < GCS_FIELD_EDGE_GEOMETRY + 3) {
    const int d = field - GCS_FIELD_EDGE_GEOMETRY;
    size_t ee[3];
    edge_extent(L, d, ee);
    return planes(P_EDGE + EDGE_PLANES * d, 10, ee);
  }
  if (field == GCS_FIELD_BG_CELL_MAGNETIC) return planes(P_B0C, 3, ce);
  if (field == GCS_FIELD_BG_CELL_FORCE) return planes(P_FORCE, 3, ce);
  if (field >= GCS_FIELD_BG_FACE_MAGNETIC && field < GCS_FIELD_COUNT) {
    const int group = (field - GCS_FIELD_BG_FACE_MAGNETIC) / 3;
    const int d = (field - GCS_FIELD_BG_FACE_MAGNETIC) % 3;
    size_t fe[3], ee[3];
    face_extent(L, d, fe);
    edge_extent(L, d, ee);
    if (group == 0) return planes(P_B0F + 3 * d, 3, fe);
    if (group == 1) return planes(P_B0FLUX + d, 1, fe);
    return planes(P_B0E + 3 * d, 3, ee);
  }
  set_error("unknown field id");
  return -1;
}

int download_valid(Problem &P, int face, int dir, std::vector<uint8_t> &out) {
  const Layout &L = P.L;
  size_t ee[3];
  edge_extent(L, dir, ee);
  std::vector<uint8_t> buf(L.face_stride);
  GCS_HIP(hipMemcpy(buf.data(), P.byte_plane(Q_VALID + dir) + face * L.face_stride,
                    L.face_stride, hipMemcpyDeviceToHost));
  out.assign(ee[0] * ee[1] * ee[2], 0);
  for (size_t i = 0; i < ee[0]; ++i)
    for (size_t j = 0; j < ee[1]; ++j)
      for (size_t k = 0; k < ee[2]; ++k)
        out[(i * ee[1] + j) * ee[2] + k] = buf[(i * L.SJ + j) * L.SK + k];
  return 0;
}

}  // namespace gcs

namespace gcs {
// The driver step as one hipGraph launch. The graph of the current level
// parity is captured from enqueue_step on first use and again whenever a
// value that the capture fixed in the kernel arguments has changed.
static int launch_step_graph(Problem &P) {
  Problem::StepGraph &g = P.graphs[P.level];
  const bool match = g.exec != nullptr && g.mixed == P.mixed && g.fidelity == P.fidelity &&
                     g.plain == P.plain &&
                     g.old_halo_current == P.old_halo_current &&
                     g.cfl_folded == P.cfl_folded &&
                     g.b_recovered == P.b_recovered &&
                     g.mi_potential == P.mi_view.potential;
  if (!match) {
    if (g.exec != nullptr) {
      GCS_HIP(hipGraphExecDestroy(g.exec));
      g.exec = nullptr;
    }
    const int level = P.level;
    const bool recovered = P.b_recovered, halo_current = P.old_halo_current,
               folded = P.cfl_folded;
    hipGraph_t graph = nullptr;
    GCS_HIP(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step(P, false);
    const hipError_t end = hipStreamEndCapture(P.stream, &graph);
    // Capturing ran no kernel: the state is still that of the step's start.
    P.level = level;
    P.b_recovered = recovered;
    P.old_halo_current = halo_current;
    P.cfl_folded = folded;
    if (rc != 0 || end != hipSuccess) {
      if (graph != nullptr) (void)hipGraphDestroy(graph);
      if (rc == 0) set_error(std::string("step capture: ") + hipGetErrorString(end));
      return -1;
    }
    const hipError_t inst = hipGraphInstantiateWithFlags(&g.exec, graph, 0);
    (void)hipGraphDestroy(graph);
    if (inst != hipSuccess) {
      g.exec = nullptr;
      set_error(std::string("step graph instantiation: ") + hipGetErrorString(inst));
      return -1;
    }
    g.mixed = P.mixed;
    g.plain = P.plain;
    g.fidelity = P.fidelity;
    g.b_recovered = P.b_recovered;
    g.old_halo_current = P.old_halo_current;
    g.cfl_folded = P.cfl_folded;
    g.mi_potential = P.mi_view.potential;
  }
  GCS_HIP(hipGraphLaunch(g.exec, P.stream));
  P.level = 1 - P.level;
  P.host_dirty = true;
  P.old_halo_current = P.b_recovered;
  P.b_recovered = true;
  P.cfl_folded = P.driver_cfl;
  return 0;
}

// One driver iteration: the MI cadence check on the clock of the previous
// step (test_cubedsphere_dipole_background.c:617-625), then the step.
int driver_step(Problem &P) {
  if (P.mi != nullptr) {
    StepScalars s;
    if (read_scalars(P, s) != 0) return -1;
    if (s.failed) return check_failure(P);
    if (mi_maybe_update(P, s.time) < 0) return -1;
  }
  if (P.use_graph && !P.profiling) return launch_step_graph(P);
  return enqueue_step(P, false);
}
}  // namespace gcs

using gcs::Problem;

extern "C" {

const char *gcs_last_error(void) { return gcs::last_error(); }

size_t gcs_config_size(void) { return sizeof(gcs_problem_config); }

Problem *gcs_problem_create(const gcs_problem_config *cfg) {
  if (cfg == nullptr || gcs::check_config(*cfg) != 0) return nullptr;
  Problem *P = new Problem();
  P->cfg = *cfg;
  P->L = gcs::make_layout((int)cfg->radial_cells, (int)cfg->angular_cells,
                          (int)cfg->ghost);
  if (gcs::allocate_problem(*P) != 0) {
    gcs::free_problem(P);
    return nullptr;
  }
  if (gcs::make_wind(*P) != 0) {
    gcs::free_problem(P);
    return nullptr;
  }
  gcs::driver_options(*P);
  P->driver_cfl = true;
  if (gcs::setup_problem(*P) != 0 || gcs::build_plans(*P) != 0 ||
      (cfg->mi_enabled && gcs::mi_create(*P) != 0)) {
    gcs::free_problem(P);
    return nullptr;
  }
  return P;
}

void gcs_problem_destroy(Problem *P) { gcs::free_problem(P); }

int gcs_problem_advance(Problem *P, long steps) {
  if (P == nullptr || steps < 0) return -1;
  for (long n = 0; n < steps; ++n)
    if (gcs::driver_step(*P) != 0) return -1;
  return gcs::check_failure(*P);
}

// One step with caller-chosen dt, predictor ratio and boundary times; the
// clock is left unchanged (like refh_step_explicit / orc_step_explicit).
int gcs_problem_step_explicit(Problem *P, double ratio, double dt, double time,
                              double previous_time) {
  if (P == nullptr || !(dt > 0.0) || !std::isfinite(ratio)) return -1;
  gcs::StepScalars s;
  if (gcs::read_scalars(*P, s) != 0) return -1;
  s.dt = dt;
  s.ratio = ratio;
  s.time = time;
  s.previous_time = previous_time;
  s.wall_time = time;
  s.wall_dt = dt;
  if (hipMemcpy(P->scalars, &s, sizeof(s), hipMemcpyHostToDevice) != hipSuccess) return -1;
  if (gcs::enqueue_step(*P, true) != 0) return -1;
  return gcs::check_failure(*P);
}

void gcs_problem_clock(Problem *P, double out[4]) {
  gcs::StepScalars s;
  if (P == nullptr || gcs::read_scalars(*P, s) != 0) return;
  out[0] = s.time;
  out[1] = s.dt;
  out[2] = s.previous_dt;
  out[3] = (double)s.steps;
}

int gcs_problem_field(Problem *P, int face, int field, double **data,
                      size_t *count) {
  if (P == nullptr || face < 0 || face >= 6 || field < 0 ||
      field >= GCS_FIELD_COUNT || data == nullptr || count == nullptr)
    return -1;
  if (hipStreamSynchronize(P->stream) != hipSuccess) return -1;
  if (P->host_dirty) {
    P->host_fields.assign(6 * GCS_FIELD_COUNT, {});
    P->host_valid.assign(18, {});
    P->host_dirty = false;
  }
  std::vector<double> &v = P->host_fields[face * GCS_FIELD_COUNT + field];
  if (v.empty() && gcs::download_field(*P, face, field, v) != 0) return -1;
  *data = v.data();
  *count = v.size();
  return 0;
}

int gcs_problem_edge_valid(Problem *P, int face, int dir, unsigned char **data,
                           size_t *count) {
  if (P == nullptr || face < 0 || face >= 6 || dir < 0 || dir > 2) return -1;
  if (P->host_dirty) {
    P->host_fields.assign(6 * GCS_FIELD_COUNT, {});
    P->host_valid.assign(18, {});
    P->host_dirty = false;
  }
  std::vector<uint8_t> &v = P->host_valid[face * 3 + dir];
  if (v.empty() && gcs::download_valid(*P, face, dir, v) != 0) return -1;
  *data = v.data();
  *count = v.size();
  return 0;
}

// FP64 (0) or the mixed form (1); the first switch to the mixed form
// allocates its FP32 planes.
int gcs_problem_set_precision(Problem *P, int mixed) {
  if (P == nullptr || (mixed != 0 && mixed != 1)) return -1;
  if (mixed == 1 && gcs::prepare_mixed(*P) != 0) return -1;
  P->mixed = mixed == 1;
  return 0;
}

int gcs_problem_mi_sample_current(Problem *P) {
  if (P == nullptr || P->mi == nullptr) {
    gcs::set_error("MI is not enabled");
    return -1;
  }
  return gcs::mi_sample_current(*P);
}

int gcs_problem_max_div_b(Problem *P, double *out) {
  if (P == nullptr || out == nullptr) return -1;
  return gcs::max_div_b(*P, out);
}

int gcs_problem_pressure_control(Problem *P, double dt, double gamma, double density_floor,
                                 double pressure_floor, double light_speed,
                                 double low_density_threshold,
                                 double low_density_sound_speed, unsigned long long counts[2],
                                 double maxima[2]) {
  if (P == nullptr || counts == nullptr || maxima == nullptr) return -1;
  const double args[7] = {dt, gamma, density_floor, pressure_floor, light_speed,
                          low_density_threshold, low_density_sound_speed};
  if (gcs::pressure_control(*P, args, counts, maxima) != 0) return -1;
  gcs::StepScalars s;
  if (gcs::read_scalars(*P, s) != 0) return -1;
  return s.failed ? gcs::check_failure(*P) : 0;
}

int gcs_problem_pressure_control_noalloc(Problem *P, double dt, double gamma,
                                         double density_floor, double pressure_floor,
                                         double light_speed, double low_density_threshold,
                                         double low_density_sound_speed,
                                         unsigned long long counts[2], double maxima[2]) {
  if (P == nullptr || counts == nullptr || maxima == nullptr) return -1;
  const double args[7] = {dt, gamma, density_floor, pressure_floor, light_speed,
                          low_density_threshold, low_density_sound_speed};
  if (gcs::pressure_control_noalloc(*P, args, counts, maxima) != 0) return -1;
  gcs::StepScalars s;
  if (gcs::read_scalars(*P, s) != 0) return -1;
  return s.failed ? gcs::check_failure(*P) : 0;
}

int gcs_problem_final_gates(Problem *P, double out[3]) {
  if (P == nullptr || out == nullptr) return -1;
  if (gcs::final_gates(*P, P->cfg.startup.gamma, out) != 0) return -1;
  gcs::StepScalars s;
  if (gcs::read_scalars(*P, s) != 0) return -1;
  return s.failed ? gcs::check_failure(*P) : 0;
}

int gcs_problem_set_wind(Problem *P, size_t count, const double *time, const double *state,
                         int linear) {
  if (P == nullptr || count < 1 || count > (size_t)INT32_MAX || time == nullptr ||
      state == nullptr) {
    gcs::set_error("invalid wind series arguments");
    return -1;
  }
  // solar_wind.c:55-68: finite samples, strictly increasing times
  for (size_t n = 0; n < count; ++n) {
    if (!std::isfinite(time[n]) || (n > 0 && !(time[n] > time[n - 1]))) {
      gcs::set_error("wind series times must be finite and increasing");
      return -1;
    }
    for (int q = 0; q < 8; ++q)
      if (!std::isfinite(state[8 * n + q])) {
        gcs::set_error("wind series states must be finite");
        return -1;
      }
  }
  if (gcs::upload_wind(*P, (int)count, time, state) != 0) return -1;
  P->wind_series.linear = count > 1 ? (linear != 0) : 1;
  // captured steps hold the series' scalars: capture again
  for (auto &g : P->graphs)
    if (g.exec != nullptr) {
      (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
    }
  return 0;
}

int gcs_problem_write_snapshot(Problem *P, const char *path) {
  if (P == nullptr || path == nullptr) return -1;
  return gcs::write_snapshot(*P, path);
}

int gcs_problem_write_drive_frame(Problem *P, const char *prefix, long frame, long step,
                                  double seconds, const char *index_path) {
  if (P == nullptr) return -1;
  return gcs::write_drive_frame(*P, prefix, frame, step, seconds, index_path);
}

int gcs_problem_load_snapshot(Problem *P, const char *path) {
  if (P == nullptr || path == nullptr) return -1;
  return gcs::load_snapshot(*P, path);
}

int gcs_problem_mi_maybe_update(Problem *P, double time_code, int *updated) {
  if (P == nullptr || updated == nullptr || P->mi == nullptr) return -1;
  const int rc = gcs::mi_maybe_update(*P, time_code);
  if (rc < 0) return -1;
  *updated = rc;
  return 0;
}

int gcs_problem_write_mi_csv(Problem *P, const char *prefix) {
  if (P == nullptr || prefix == nullptr || P->mi == nullptr) return -1;
  return gcs::write_mi_csv(*P, prefix) == 0 && gcs::write_shell_current_csv(*P, prefix) == 0
             ? 0
             : -1;
}

int gcs_copy_bandwidth(size_t doubles, int iterations, double *gb_per_s) {
  return gcs::copy_bandwidth(doubles, iterations, gb_per_s);
}

int gcs_device_synchronize(void) {
  if (hipDeviceSynchronize() != hipSuccess) {
    gcs::set_error("device synchronize failed");
    return -1;
  }
  return 0;
}

int gcs_problem_set_plain_kernels(Problem *P, int on) {
  if (P == nullptr) return -1;
  P->plain = on != 0;
  return 0;
}

int gcs_problem_set_graph(Problem *P, int on) {
  if (P == nullptr) return -1;
  P->use_graph = on != 0;
  return 0;
}

// Keep (1) or skip (0) the work that only reproduces the reference's ghost
// values in storage; the step's active results are the same either way.
int gcs_problem_set_fidelity(Problem *P, int on) {
  if (P == nullptr) return -1;
  P->fidelity = on != 0;
  return 0;
}

int gcs_stage_count(void) { return gcs::kStages; }
const char *gcs_stage_name(int stage) { return gcs::stage_name(stage); }

// Run `steps` driver steps timed by HIP events on the problem's stream.
// elapsed_ms: start to end of all steps. With profile != 0, stage_ms[s]
// receives the summed time from the start of stage s to the start of the
// next recorded stage, over all steps (kStages entries).
int gcs_problem_time_steps(Problem *P, long steps, int profile,
                           double *elapsed_ms, double *stage_ms) {
  if (P == nullptr || steps <= 0 || elapsed_ms == nullptr) return -1;
  P->profiling = profile != 0;
  if (P->profiling) {
    const size_t need = (size_t)steps * gcs::kStages + 1;
    while (P->stage_events.size() < need) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      P->stage_events.push_back(e);
    }
    P->stage_ids.assign(P->stage_events.size(), -1);
    P->stage_cursor = 0;
  }
  if (hipEventRecord(P->ev_start, P->stream) != hipSuccess) return -1;
  for (long n = 0; n < steps; ++n)
    if (gcs::driver_step(*P) != 0) {
      P->profiling = false;
      return -1;
    }
  if (hipEventRecord(P->ev_stop, P->stream) != hipSuccess) return -1;
  if (gcs::check_failure(*P) != 0) {
    P->profiling = false;
    return -1;
  }
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, P->ev_start, P->ev_stop) != hipSuccess) return -1;
  *elapsed_ms = ms;
  if (P->profiling && stage_ms != nullptr) {
    for (int s = 0; s < gcs::kStages; ++s) stage_ms[s] = 0.0;
    for (size_t e = 0; e + 1 < P->stage_cursor; ++e) {
      if (P->stage_ids[e] == gcs::S_END) continue;
      float t = 0.0f;
      if (hipEventElapsedTime(&t, P->stage_events[e], P->stage_events[e + 1]) !=
          hipSuccess)
        return -1;
      stage_ms[P->stage_ids[e]] += t;
    }
  }
  P->profiling = false;
  return 0;
}

// Emergency-diffusion counters of one cube face (count, max speed).
void gcs_problem_counters(Problem *P, int face, double out[2]) {
  gcs::StepScalars s;
  if (P == nullptr || gcs::read_scalars(*P, s) != 0) return;
  out[0] = (double)s.emergency_count[face];
  double speed;
  std::memcpy(&speed, &s.emergency_speed_bits[face], sizeof(speed));
  out[1] = speed;
}

// Ionosphere coupling state: ints = {updates, ready, iterations north,
// iterations south}, doubles = {final residual north, south, epoch (s),
// last solve (ms)}.
int gcs_problem_mi_report(Problem *P, int ints[4], double doubles[4]) {
  if (P == nullptr || P->mi == nullptr || ints == nullptr || doubles == nullptr) {
    gcs::set_error("no ionosphere coupling on this problem");
    return -1;
  }
  gcs::MiReport r;
  gcs::mi_report(*P, r);
  ints[0] = r.updates;
  ints[1] = r.ready;
  ints[2] = r.iterations[0];
  ints[3] = r.iterations[1];
  doubles[0] = r.final_residual[0];
  doubles[1] = r.final_residual[1];
  doubles[2] = r.epoch_s;
  doubles[3] = r.solve_ms;
  return 0;
}

// what = 0: potential (V), 1: sampled field-aligned current (A/m^2); both
// [hemisphere][colatitude][azimuth]. Valid until the next call.
int gcs_problem_mi_field(Problem *P, int what, double **data, size_t *count) {
  if (P == nullptr || data == nullptr || count == nullptr || (what != 0 && what != 1))
    return -1;
  static thread_local std::vector<double> buf;
  if (gcs::mi_download(*P, what, buf) != 0) return -1;
  *data = buf.data();
  *count = buf.size();
  return 0;
}

}  // extern "C"
