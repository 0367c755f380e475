// Drop-in C ABI of the reference's serial six-face step
// (cubedsphere/cubedsphere_step.h:16-67): the four gamera_cs_advance_six*
// entry points with the reference's prototypes and error behaviour, run on
// the GPU. Declared in include/gamera_cs_dropin.h.
//
// The caller keeps owning grid, storage, background and plans in the
// reference's host layout. Each (grid, storage) pair gets a device mirror on
// its first call: the geometry and background are uploaded once (both must
// stay unchanged while the mirror lives, as the reference requires of its
// fixed grid and halo plan, cubedsphere_halo.c:561-570), the constants
// derived from them are computed on the device, and the halo and seam plans
// are built from the uploaded geometry. The caller's own halo plan is
// accepted and not read: both plans are functions of the same grid.
//
// State coherence. By default every call uploads the state the step reads
// and downloads everything the reference's step writes (all time levels,
// predicted arrays, face fluxes, EMFs, counters), so host code between
// calls (dt rule, output, pressure control) sees what it would see with the
// reference. In resident mode (gcs_dropin_set_resident) the state stays on
// the device between calls; gcs_dropin_sync_to_host and
// gcs_dropin_mark_host_modified move it explicitly.
//
// Callbacks. The reference's own Earth callbacks
// (cubedsphere_earth_boundary.h:26-58: gamera_cs_earth_fill_radial_ghosts,
// gamera_cs_earth_inner_fluid_flux, gamera_cs_earth_no_mi_emf) are
// recognised by address and run as device kernels, with an ionosphere
// state of this library's gamera_cs_mi_dipole_create as well. Any other
// callback, and the Earth ones when their context needs something the
// device path does not hold (another implementation's MI state, a different
// background), runs on the host at its place in the step: the state is
// downloaded, the callback called, and the state uploaded again.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gamera_cs_dropin.h"
#include "gamera_cs_mi_dropin.h"
#include "gamera_cs_mpi_dropin.h"
#include "gcs_device.hpp"
#include "gcs_transfer.hpp"

namespace gcs {
namespace {

using RadialFn = gamera_cs_radial_boundary_function;
using SyncFn = int (*)(gamera_no_storage *, const gamera_no_grid *,
                       const size_t *, const size_t *, void *);

struct Callbacks {
  RadialFn radial = nullptr;
  SyncFn flux = nullptr;
  SyncFn emf = nullptr;
  SyncFn tile_emf = nullptr;  // gamera_cs_earth_tile_emf (cubedsphere_earth_boundary.c:453-464)
  bool registered = false;
};

struct Mirror {
  const gamera_no_grid *grid = nullptr;
  gamera_no_storage *storage = nullptr;
  // Arrays the mirror was built from: a new allocation behind the same
  // struct addresses makes a new mirror.
  const void *arrays[GAMERA_CS_FACE_COUNT][3] = {};
  const gamera_no_background_field *background = nullptr;
  const void *background_arrays[GAMERA_CS_FACE_COUNT] = {};
  bool has_background = false;
  Problem *P = nullptr;
  bool device_current = false;  // resident mode: device state is newest
  std::vector<double> wind;     // the wind samples last uploaded
  struct MiHandle *mi = nullptr;  // its ionosphere state, if any
};

// A gamera_cs_mi_dipole of this library: the MI state of a mirror's problem
// (gamera_cs_mi_dipole_create), or a potential alone on the device
// (gamera_cs_mi_dipole_create_potential_view).
struct MiHandle {
  Mirror *m = nullptr;  // null for a potential view, or once the mirror is gone
  gamera_cs_mi_dipole_config cfg{};
  gamera_cs_earth_startup startup{};
  size_t active[3] = {0, 0, 0}, ghost = 0;
  // potential view: its potential on the device, clock and readiness
  double *view_potential = nullptr;
  MiView view;
  bool view_ready = false;
  double view_epoch = 0.0, view_next = 0.0;
  double *query = nullptr;  // 4 device doubles for point queries
  std::vector<double> host_fac, host_potential;  // analysis-view copies
};

std::mutex g_lock;
std::vector<MiHandle *> g_mi;

MiHandle *find_mi(const void *state) {
  for (MiHandle *h : g_mi)
    if (h == state) return h;
  return nullptr;
}
std::vector<Mirror *> g_mirrors;
Callbacks g_cb;
bool g_resident = false;
long g_host_rounds = 0;  // callbacks run on the host with the state moved there and back

// cubedsphere_step.c:7-11
bool same_value(double a, double b) {
  const double scale = std::fmax(1.0, std::fmax(std::fabs(a), std::fabs(b)));
  return std::isfinite(a) && std::isfinite(b) && std::fabs(a - b) <= 8e-15 * scale;
}

// cubedsphere_step.c:23-47: rejected before anything is written.
bool valid_call(const gamera_no_grid *grid, gamera_no_storage *storage,
                const size_t *active, size_t ghost, double ratio, double dt,
                const gamera_no_advance_options *o,
                const gamera_no_background_field *background,
                RadialFn radial, double current_time, double previous_time) {
  if (grid == nullptr || storage == nullptr || active == nullptr || o == nullptr ||
      ghost < 4 || !std::isfinite(dt) || dt <= 0.0 ||
      (radial != nullptr &&
       (!std::isfinite(current_time) || !std::isfinite(previous_time))) ||
      !std::isfinite(ratio) || active[1] != active[2] ||
      o->cell_source != nullptr || o->indexed_cell_source != nullptr ||
      o->update.use_background != o->stress.use_background ||
      o->update.use_background != o->emf.use_background ||
      (o->update.use_background && background == nullptr) ||
      !std::isfinite(o->source_time) ||
      !same_value(o->stress.gamma, o->update.gamma) ||
      !same_value(o->stress.density_floor, o->update.density_floor) ||
      !same_value(o->stress.pressure_floor, o->update.pressure_floor) ||
      !same_value(o->emf.density_floor, o->update.density_floor) ||
      o->stress.use_mhd != o->update.use_mhd ||
      o->stress.use_boris != o->update.use_boris ||
      o->emf.use_boris != o->update.use_boris ||
      (o->update.use_boris &&
       (!same_value(o->stress.light_speed, o->update.light_speed) ||
        !same_value(o->emf.light_speed, o->update.light_speed))))
    return false;
  return true;
}

void resolve_callbacks() {
  if (g_cb.registered) return;
  g_cb.radial = reinterpret_cast<RadialFn>(
      dlsym(RTLD_DEFAULT, "gamera_cs_earth_fill_radial_ghosts"));
  g_cb.flux = reinterpret_cast<SyncFn>(
      dlsym(RTLD_DEFAULT, "gamera_cs_earth_inner_fluid_flux"));
  g_cb.emf = reinterpret_cast<SyncFn>(
      dlsym(RTLD_DEFAULT, "gamera_cs_earth_no_mi_emf"));
  g_cb.tile_emf = reinterpret_cast<SyncFn>(
      dlsym(RTLD_DEFAULT, "gamera_cs_earth_tile_emf"));
}

// The Earth context of an EMF callback the device runs: the reference's
// gamera_cs_earth_no_mi_emf's own context, or, for gamera_cs_earth_tile_emf
// on the only radial tile (which applies both the outer and the inner wall
// EMF, cubedsphere_earth_boundary.c:453-464), its tile context's Earth
// context. Null otherwise (the callback then runs on the host).
const gamera_cs_earth_boundary *device_emf_context(SyncFn fn, void *context) {
  if (fn == nullptr) return nullptr;
  if (fn == g_cb.emf) return static_cast<const gamera_cs_earth_boundary *>(context);
  if (fn == g_cb.tile_emf && context != nullptr) {
    const auto *t = static_cast<const gamera_cs_earth_tile_emf_context *>(context);
    if (t->tile == nullptr || t->earth == nullptr) return nullptr;
    if (t->tile->coord[0] == 0 && t->tile->proc_dims[0] == 1) return t->earth;
  }
  return nullptr;
}

// ---- mirrors --------------------------------------------------------------------

bool extents_match(const gamera_no_grid *grid, const gamera_no_storage *storage,
                   const Layout &L) {
  const size_t ce[3] = {(size_t)L.NI, (size_t)L.NJ, (size_t)L.NK};
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    for (int a = 0; a < 3; ++a)
      if (grid[f].cell_extent[a] != ce[a] || storage[f].cell_extent[a] != ce[a] ||
          grid[f].vertex_extent[a] != ce[a] + 1)
        return false;
    for (int d = 0; d < 3; ++d) {
      size_t fe[3], ee[3];
      face_extent(L, d, fe);
      edge_extent(L, d, ee);
      for (int a = 0; a < 3; ++a)
        if (grid[f].face[d].extent[a] != fe[a] || storage[f].face_extent[d][a] != fe[a] ||
            grid[f].edge[d].extent[a] != ee[a] || storage[f].edge_extent[d][a] != ee[a])
          return false;
    }
    if (grid[f].vertex == nullptr || grid[f].cell == nullptr ||
        storage[f].conserved == nullptr || storage[f].old_conserved == nullptr ||
        storage[f].predicted_conserved == nullptr || storage[f].cell_magnetic == nullptr ||
        storage[f].old_cell_magnetic == nullptr ||
        storage[f].predicted_cell_magnetic == nullptr)
      return false;
    for (int d = 0; d < 3; ++d)
      if (grid[f].face[d].value == nullptr || grid[f].edge[d].value == nullptr ||
          grid[f].edge[d].valid == nullptr || storage[f].face_flux[d] == nullptr ||
          storage[f].old_face_flux[d] == nullptr ||
          storage[f].predicted_face_flux[d] == nullptr ||
          storage[f].edge_emf[d] == nullptr || storage[f].fluid_face_flux[d] == nullptr ||
          storage[f].maxwell_face_flux[d] == nullptr)
        return false;
  }
  return true;
}

void record_arrays(Mirror &m) {
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    m.arrays[f][0] = m.grid[f].vertex;
    m.arrays[f][1] = m.grid[f].cell;
    m.arrays[f][2] = m.storage[f].conserved;
  }
}

bool arrays_unchanged(const Mirror &m) {
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f)
    if (m.arrays[f][0] != m.grid[f].vertex || m.arrays[f][1] != m.grid[f].cell ||
        m.arrays[f][2] != m.storage[f].conserved)
      return false;
  return true;
}

int upload_geometry(Problem &P, const gamera_no_grid *grid) {
  const Layout &L = P.L;
  const size_t ce[3] = {(size_t)L.NI, (size_t)L.NJ, (size_t)L.NK};
  const size_t ve[3] = {(size_t)L.NI + 1, (size_t)L.NJ + 1, (size_t)L.NK + 1};
  const size_t per_cell = sizeof(gamera_no_cell_geometry) / sizeof(double);
  const size_t per_face = sizeof(gamera_no_face_geometry) / sizeof(double);
  const size_t per_edge = sizeof(gamera_no_edge_geometry) / sizeof(double);
  static_assert(sizeof(gamera_no_face_geometry) == FACE_PLANES * sizeof(double),
                "face geometry planes");
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    const gamera_no_grid &g = grid[f];
    if (host_to_planes(P, P_VERT, 3, f, ve, g.vertex[0].value) != 0) return -1;
    const double *cell = reinterpret_cast<const double *>(g.cell);
    if (host_to_plane(P, P_VOL, f, ce, cell, per_cell) != 0) return -1;
    for (int e = 0; e < 3; ++e) {
      if (host_to_plane(P, P_CEN + e, f, ce, cell + 1 + e, per_cell) != 0) return -1;
      const size_t cfl = offsetof(gamera_no_cell_geometry, cfl_length) / sizeof(double);
      if (host_to_plane(P, P_CFL + e, f, ce, cell + cfl + e, per_cell) != 0) return -1;
    }
    for (int d = 0; d < 3; ++d) {
      size_t fe[3], ee[3];
      face_extent(L, d, fe);
      edge_extent(L, d, ee);
      const double *face = reinterpret_cast<const double *>(g.face[d].value);
      for (int q = 0; q < (int)per_face; ++q)
        if (host_to_plane(P, P_FACE + FACE_PLANES * d + q, f, fe, face + q, per_face) != 0)
          return -1;
      const double *edge = reinterpret_cast<const double *>(g.edge[d].value);
      for (int q = 0; q < (int)per_edge; ++q)
        if (host_to_plane(P, P_EDGE + EDGE_PLANES * d + q, f, ee, edge + q, per_edge) != 0)
          return -1;
      if (host_to_byte_plane(P, Q_VALID + d, f, ee, g.edge[d].valid) != 0) return -1;
    }
  }
  return 0;
}

// Background planes; with no background they are zero, which gives the
// reference's use_background = false arithmetic (it adds nothing).
int upload_background(Problem &P, const gamera_no_background_field *bg) {
  const Layout &L = P.L;
  const size_t ce[3] = {(size_t)L.NI, (size_t)L.NJ, (size_t)L.NK};
  std::vector<double> zero;
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    const gamera_no_background_field *b = bg ? &bg[f] : nullptr;
    auto put = [&](int plane, int ncomp, const size_t e[3], const double *src) {
      if (src != nullptr) return host_to_planes(P, plane, ncomp, f, e, src);
      zero.assign(e[0] * e[1] * e[2] * ncomp, 0.0);
      return host_to_planes(P, plane, ncomp, f, e, zero.data());
    };
    if (b && (b->cell_magnetic == nullptr || b->cell_force == nullptr)) {
      set_error("background field arrays missing");
      return -1;
    }
    if (put(P_B0C, 3, ce, b ? b->cell_magnetic[0].value : nullptr) != 0 ||
        put(P_FORCE, 3, ce, b ? b->cell_force[0].value : nullptr) != 0)
      return -1;
    for (int d = 0; d < 3; ++d) {
      size_t fe[3], ee[3];
      face_extent(L, d, fe);
      edge_extent(L, d, ee);
      if (b && (b->face_magnetic[d] == nullptr || b->face_flux[d] == nullptr ||
                b->edge_magnetic[d] == nullptr)) {
        set_error("background field arrays missing");
        return -1;
      }
      if (put(P_B0F + 3 * d, 3, fe, b ? b->face_magnetic[d][0].value : nullptr) != 0 ||
          put(P_B0FLUX + d, 1, fe, b ? b->face_flux[d] : nullptr) != 0 ||
          put(P_B0E + 3 * d, 3, ee, b ? b->edge_magnetic[d][0].value : nullptr) != 0)
        return -1;
    }
  }
  return 0;
}

// The arrays one six-face step reads or writes, in both directions.
int state_to_device(Problem &P, const gamera_no_storage *st) {
  const Layout &L = P.L;
  P.b_recovered = false;
  P.old_halo_current = false;
  P.cfl_folded = false;
  const size_t ce[3] = {(size_t)L.NI, (size_t)L.NJ, (size_t)L.NK};
  const int cur = P.level, old = 1 - P.level;
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    const gamera_no_storage &s = st[f];
    if (host_to_planes(P, plane_u(cur, 0), 5, f, ce, s.conserved) != 0 ||
        host_to_planes(P, plane_u(old, 0), 5, f, ce, s.old_conserved) != 0 ||
        host_to_planes(P, P_UPRED, 5, f, ce, s.predicted_conserved) != 0 ||
        host_to_planes(P, plane_b(cur, 0), 3, f, ce, s.cell_magnetic[0].value) != 0 ||
        host_to_planes(P, plane_b(old, 0), 3, f, ce, s.old_cell_magnetic[0].value) != 0 ||
        host_to_planes(P, P_BPRED, 3, f, ce, s.predicted_cell_magnetic[0].value) != 0)
      return -1;
    for (int d = 0; d < 3; ++d) {
      size_t fe[3], ee[3];
      face_extent(L, d, fe);
      edge_extent(L, d, ee);
      if (host_to_plane(P, plane_phi(cur, d), f, fe, s.face_flux[d], 1) != 0 ||
          host_to_plane(P, plane_phi(old, d), f, fe, s.old_face_flux[d], 1) != 0 ||
          host_to_plane(P, P_PHIPRED + d, f, fe, s.predicted_face_flux[d], 1) != 0 ||
          host_to_plane(P, P_EMF + d, f, ee, s.edge_emf[d], 1) != 0 ||
          host_to_planes(P, P_FLUID + 5 * d, 5, f, fe, s.fluid_face_flux[d]) != 0 ||
          host_to_planes(P, P_MAXWELL + 3 * d, 3, f, fe,
                         s.maxwell_face_flux[d][0].value) != 0)
        return -1;
    }
  }
  return 0;
}

int state_to_host(Problem &P, gamera_no_storage *st) {
  const Layout &L = P.L;
  const size_t ce[3] = {(size_t)L.NI, (size_t)L.NJ, (size_t)L.NK};
  const int cur = P.level, old = 1 - P.level;
  GCS_HIP(hipStreamSynchronize(P.stream));
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    gamera_no_storage &s = st[f];
    if (planes_to_host(P, plane_u(cur, 0), 5, f, ce, s.conserved) != 0 ||
        planes_to_host(P, plane_u(old, 0), 5, f, ce, s.old_conserved) != 0 ||
        planes_to_host(P, P_UPRED, 5, f, ce, s.predicted_conserved) != 0 ||
        planes_to_host(P, plane_b(cur, 0), 3, f, ce, s.cell_magnetic[0].value) != 0 ||
        planes_to_host(P, plane_b(old, 0), 3, f, ce, s.old_cell_magnetic[0].value) != 0 ||
        planes_to_host(P, P_BPRED, 3, f, ce, s.predicted_cell_magnetic[0].value) != 0)
      return -1;
    for (int d = 0; d < 3; ++d) {
      size_t fe[3], ee[3];
      face_extent(L, d, fe);
      edge_extent(L, d, ee);
      if (plane_to_host(P, plane_phi(cur, d), f, fe, s.face_flux[d], 1) != 0 ||
          plane_to_host(P, plane_phi(old, d), f, fe, s.old_face_flux[d], 1) != 0 ||
          plane_to_host(P, P_PHIPRED + d, f, fe, s.predicted_face_flux[d], 1) != 0 ||
          plane_to_host(P, P_EMF + d, f, ee, s.edge_emf[d], 1) != 0 ||
          planes_to_host(P, P_FLUID + 5 * d, 5, f, fe, s.fluid_face_flux[d]) != 0 ||
          planes_to_host(P, P_MAXWELL + 3 * d, 3, f, fe,
                         s.maxwell_face_flux[d][0].value) != 0)
        return -1;
    }
    // stress rates of the active cells (nonorthogonal_sweep.c:550-622), which
    // the update stores in the face-velocity planes
    const size_t lo[3] = {(size_t)L.ghost, (size_t)L.ghost, (size_t)L.ghost};
    const size_t hi[3] = {(size_t)(L.ghost + L.nr), (size_t)(L.ghost + L.n),
                          (size_t)(L.ghost + L.n)};
    if (P.rates_stored &&
        (box_planes_to_host(P, P_FVEL, 5, f, ce, lo, hi, s.hydro_rate) != 0 ||
         box_planes_to_host(P, P_FVEL + 5, 3, f, ce, lo, hi, s.maxwell_rate[0].value) != 0))
      return -1;
  }
  return 0;
}

void release(Mirror *m) {
  if (m == nullptr) return;
  if (m->mi != nullptr) m->mi->m = nullptr;  // its MI state goes with the problem
  free_problem(m->P);
  delete m;
}

Mirror *build_mirror(const gamera_no_grid *grid, gamera_no_storage *storage,
                     const size_t active[3], size_t ghost,
                     const gamera_no_background_field *background) {
  if (ghost < 4 || active[0] < 2 || active[1] <= 2 * ghost) {
    set_error("grid needs ghost >= 4, at least 2 radial and more than 2 ghost angular "
              "active cells (cubedsphere_mesh.c:188-193)");
    return nullptr;
  }
  const long g2 = 2 * (long)ghost + 1;
  const long cells = ((long)active[0] + g2) * ((long)active[1] + g2) *
                     (((long)active[1] + g2 + 7) / 8 * 8) * 6;
  if (cells >= (1L << 31)) {
    set_error("grid too large for 32-bit indices");
    return nullptr;
  }
  Mirror *m = new Mirror();
  m->grid = grid;
  m->storage = storage;
  m->P = new Problem();
  Problem &P = *m->P;
  std::memset(&P.cfg, 0, sizeof(P.cfg));
  P.L = make_layout((int)active[0], (int)active[1], (int)ghost);
  if (!extents_match(grid, storage, P.L)) {
    set_error("grid/storage extents do not match active + 2 ghost");
    release(m);
    return nullptr;
  }
  P.bnd = StepBoundary();  // plans cover every non-physical face
  P.fidelity = true;
  if (allocate_problem(P) != 0 || upload_geometry(P, grid) != 0 ||
      upload_background(P, background) != 0 || derive_constants(P) != 0 ||
      build_plans(P) != 0) {
    release(m);
    return nullptr;
  }
  m->background = background;
  m->has_background = background != nullptr;
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f)
    m->background_arrays[f] = background ? background[f].cell_magnetic : nullptr;
  record_arrays(*m);
  return m;
}

Mirror *find(const gamera_no_grid *grid, const gamera_no_storage *storage) {
  for (Mirror *m : g_mirrors)
    if (m->grid == grid && m->storage == storage) return m;
  return nullptr;
}

Mirror *mirror_for(const gamera_no_grid *grid, gamera_no_storage *storage,
                   const size_t active[3], size_t ghost,
                   const gamera_no_background_field *background) {
  Mirror *m = find(grid, storage);
  if (m != nullptr) {
    const Layout &L = m->P->L;
    const bool same_shape = (size_t)L.nr == active[0] && (size_t)L.n == active[1] &&
                            (size_t)L.ghost == ghost && arrays_unchanged(*m) &&
                            extents_match(grid, storage, L);
    if (!same_shape) {
      for (size_t n = 0; n < g_mirrors.size(); ++n)
        if (g_mirrors[n] == m) g_mirrors.erase(g_mirrors.begin() + (long)n);
      release(m);
      m = nullptr;
    }
  }
  if (m == nullptr) {
    m = build_mirror(grid, storage, active, ghost, background);
    if (m == nullptr) return nullptr;
    g_mirrors.push_back(m);
    return m;
  }
  bool bg_same = m->background == background && m->has_background == (background != nullptr);
  for (int f = 0; bg_same && background && f < GAMERA_CS_FACE_COUNT; ++f)
    bg_same = m->background_arrays[f] == background[f].cell_magnetic;
  if (!bg_same) {
    if (upload_background(*m->P, background) != 0 || derive_constants(*m->P) != 0 ||
        gather_seam_geometry(*m->P) != 0)
      return nullptr;
    m->background = background;
    m->has_background = background != nullptr;
    for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f)
      m->background_arrays[f] = background ? background[f].cell_magnetic : nullptr;
  }
  return m;
}

// ---- one call -------------------------------------------------------------------

void set_options(Problem &P, const gamera_no_advance_options &o) {
  StepOptions &s = P.opt;
  s.gamma = o.stress.gamma;
  s.rho_floor = o.stress.density_floor;
  s.p_floor = o.stress.pressure_floor;
  s.pdm = o.stress.pdm_coefficient;
  s.hogs_h = o.stress.hydro_hogs_coefficient;
  s.hogs_m = o.stress.magnetic_hogs_coefficient;
  s.c = o.stress.light_speed;
  s.boris = o.stress.use_boris;
  s.hogs = o.stress.use_hogs;
  s.emf_rho_floor = o.emf.density_floor;
  s.emf_pdm = o.emf.pdm_coefficient;
  s.emf_diff = o.emf.diffusion_coefficient;
  s.emf_c = o.emf.light_speed;
  s.emf_cfl = o.emf.cfl;
  s.emf_boris = o.emf.use_boris;
  s.upd_gamma = o.update.gamma;
  s.upd_rho_floor = o.update.density_floor;
  s.upd_p_floor = o.update.pressure_floor;
  s.upd_c = o.update.light_speed;
  s.upd_boris = o.update.use_boris;
  s.mhd = o.update.use_mhd ? 1 : 0;
}

// The Earth context can run on the device when it matches what the mirror
// holds (cubedsphere_earth_boundary.c:9-14 for the rest of the checks).
bool earth_on_device(const Mirror &m, const gamera_cs_earth_boundary *b,
                     const gamera_no_background_field *background) {
  if (b == nullptr || b->startup == nullptr || b->wind == nullptr ||
      b->background == nullptr || b->background != background ||
      !m.has_background || !std::isfinite(b->density_floor) ||
      !(b->density_floor > 0.0) || !std::isfinite(b->pressure_floor) ||
      !(b->pressure_floor > 0.0))
    return false;
  // an ionosphere on the device only when it is this mirror's own
  if (b->mi != nullptr) {
    const MiHandle *h = find_mi(b->mi);
    if (h == nullptr || h->m != &m) return false;
  }
  const gamera_solar_wind_series &w = *b->wind;
  return w.count >= 1 && w.count <= (size_t)INT32_MAX && w.time != nullptr &&
         w.state != nullptr;
}

// The Earth context's wind series (any length) on the device; its samples
// are uploaded again only when they differ from the last upload.
int set_wind(Problem &P, std::vector<double> &last, const gamera_cs_earth_boundary &b) {
  const gamera_solar_wind_series &s = *b.wind;
  std::vector<double> samples(9 * s.count);
  for (size_t n = 0; n < s.count; ++n) {
    samples[n] = s.time[n];
    const gamera_solar_wind_state &x = s.state[n];
    double *st = &samples[s.count + 8 * n];
    st[0] = x.density;
    st[1] = x.pressure;
    for (int d = 0; d < 3; ++d) {
      st[2 + d] = x.velocity[d];
      st[5 + d] = x.magnetic[d];
    }
  }
  if (samples != last) {
    if (upload_wind(P, (int)s.count, samples.data(), samples.data() + s.count) != 0)
      return -1;
    last.swap(samples);
  }
  WindSeries &w = P.wind_series;
  w.linear = s.linear_interpolation;
  for (int d = 0; d < 3; ++d) w.reference[d] = s.reference[d];
  w.time_offset = s.time_offset;
  w.by = s.by_coefficient;
  w.bz = s.bz_coefficient;
  return 0;
}

int report_failure(Problem &P, const StepScalars &s) {
  const long per = P.L.face_stride;
  const long at = s.fail_index;
  const int f = (int)(at / per);
  const long r = at - f * per;
  const char *rank = std::getenv("PMI_RANK");
  if (rank == nullptr) rank = std::getenv("SLURM_PROCID");
  char msg[200];
  std::snprintf(msg, sizeof(msg),
                "MHD_FAILURE rank=%s stage=%u face=%d local=%ld,%ld,%ld",
                rank ? rank : "unknown", s.fail_stage, f, r / P.L.di(),
                (r % P.L.di()) / P.L.SK, r % P.L.SK);
  std::fprintf(stderr, "%s\n", msg);
  set_error(msg);
  return -1;
}

int advance(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
            gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
            size_t ghost, double ratio, double dt, const gamera_no_advance_options *o,
            const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
            RadialFn radial, double current_time, double previous_time,
            void *boundary_context) {
  if (!valid_call(grid, storage, active, ghost, ratio, dt, o, background, radial,
                  current_time, previous_time)) {
    set_error("invalid arguments (cubedsphere_step.c:23-47)");
    return -1;
  }
  std::lock_guard<std::mutex> guard(g_lock);
  resolve_callbacks();
  const gamera_no_background_field *bg = o->update.use_background ? background : nullptr;
  Mirror *m = mirror_for(grid, storage, active, ghost, bg);
  if (m == nullptr) return -1;
  Problem &P = *m->P;
  set_options(P, *o);

  // Boundary work: device kernels for the reference's Earth callbacks.
  const auto *rb = static_cast<const gamera_cs_earth_boundary *>(boundary_context);
  const gamera_cs_earth_boundary *eb = device_emf_context(o->edge_emf_sync, o->edge_emf_context);
  StepBoundary &bd = P.bnd;
  bd = StepBoundary();
  bool host_radial = false, host_flux = false, host_emf = false;
  const gamera_cs_earth_boundary *wall = nullptr;
  if (radial != nullptr) {
    if (radial == g_cb.radial && earth_on_device(*m, rb, bg)) {
      bd.earth_cells = true;  // rewrites every radial ghost the outflow fill writes
      wall = rb;
    } else {
      bd.outflow = true;
      host_radial = true;
    }
  } else {
    bd.outflow = true;
  }
  if (o->fluid_flux_sync != nullptr) {
    if (o->fluid_flux_sync == g_cb.flux) bd.inner_trap = true;
    else host_flux = true;
  }
  if (o->edge_emf_sync != nullptr) {
    if (eb != nullptr && earth_on_device(*m, eb, bg)) {
      bd.earth_emf = true;
      if (wall == nullptr) wall = eb;
    } else {
      host_emf = true;
    }
  }
  if (bd.earth_cells && bd.earth_emf && rb != eb) {
    // Two different Earth contexts: keep the EMF one on the host.
    bd.earth_emf = false;
    host_emf = true;
  }
  if (wall != nullptr) {
    StepOptions &s = P.opt;
    s.wall_gamma = wall->startup->gamma;
    s.wall_rho_floor = wall->density_floor;
    s.wall_p_floor = wall->pressure_floor;
    s.wall_cfl = wall->cfl;
    s.strict_wall = wall->strict_inner_gas_wall != 0;
    if (set_wind(P, m->wind, *wall) != 0) return -1;
  }
  // the wall reads the potential only when its context has this mirror's
  // ionosphere (cubedsphere_earth_boundary.c:58, :440)
  mi_view_enable(P, wall != nullptr && wall->mi != nullptr);

  // Scalars of this call.
  StepScalars sc;
  if (read_scalars(P, sc) != 0) return -1;
  sc.dt = dt;
  sc.ratio = ratio;
  sc.time = current_time;
  sc.previous_time = previous_time;
  sc.wall_time = wall ? wall->step_time : current_time;
  sc.wall_dt = wall ? wall->step_dt : dt;
  sc.failed = 0;
  sc.fail_stage = 0;
  sc.fail_index = 0;
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    sc.emergency_count[f] = 0;
    sc.emergency_speed_bits[f] = 0;
  }
  GCS_HIP(hipMemcpy(P.scalars, &sc, sizeof(sc), hipMemcpyHostToDevice));
  if (bd.earth_emf) {
    // cubedsphere_earth_boundary.c:304, :387
    if (!std::isfinite(sc.wall_time) || !(sc.wall_dt > 0.0) || !(P.opt.wall_cfl > 0.0)) {
      set_error("Earth EMF context needs finite step_time, step_dt > 0 and cfl > 0");
      return -1;
    }
  }

  if (!(g_resident && m->device_current) && state_to_device(P, storage) != 0)
    return -1;
  m->device_current = true;
  const size_t lower[3] = {ghost, ghost, ghost};
  const size_t upper[3] = {ghost + active[0], ghost + active[1], ghost + active[2]};
  auto host_round = [&](auto &&call) -> int {
    ++g_host_rounds;
    if (state_to_host(P, storage) != 0) return -1;
    const int status = call();
    if (state_to_device(P, storage) != 0) return -1;
    return status;
  };
  auto check = [&]() -> int {
    StepScalars now;
    if (read_scalars(P, now) != 0) return -1;
    return now.failed ? report_failure(P, now) : 0;
  };

  if (enqueue_phase(P, PHASE_GHOSTS) != 0) return -1;
  if (host_radial) {
    if (check() != 0) return -1;
    if (host_round([&] {
          return radial(grid, storage, active, ghost, current_time, previous_time,
                        boundary_context);
        }) != 0)
      return -1;
  }
  if (enqueue_phase(P, PHASE_FLUX) != 0) return -1;
  if (host_flux) {
    if (check() != 0) return -1;
    if (host_round([&] {
          for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f)
            if (o->fluid_flux_sync(&storage[f], &grid[f], lower, upper,
                                   o->fluid_flux_context) != 0)
              return -1;
          return 0;
        }) != 0)
      return -1;
  }
  if (enqueue_phase(P, PHASE_EMF) != 0) return -1;
  if (host_emf) {
    if (check() != 0) return -1;
    if (host_round([&] {
          for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f)
            if (o->edge_emf_sync(&storage[f], &grid[f], lower, upper,
                                 o->edge_emf_context) != 0)
              return -1;
          return 0;
        }) != 0)
      return -1;
  }
  if (enqueue_phase(P, PHASE_UPDATE) != 0) return -1;

  StepScalars after;
  if (read_scalars(P, after) != 0) return -1;
  const int status = after.failed ? report_failure(P, after) : 0;
  // nonorthogonal_sweep.c:541-544: counters accumulate in the storage.
  for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
    double speed;
    std::memcpy(&speed, &after.emergency_speed_bits[f], sizeof(speed));
    storage[f].emergency_diffusion_face_count += (size_t)after.emergency_count[f];
    storage[f].emergency_diffusion_max_interface_speed =
        std::fmax(storage[f].emergency_diffusion_max_interface_speed, speed);
  }
  if (!g_resident || status != 0) {
    if (state_to_host(P, storage) != 0) return -1;
  }
  return status;
}

}  // namespace
}  // namespace gcs

extern "C" {

int gamera_cs_advance_six_with_background_planned(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context) {
  (void)halo_plan;
  return gcs::advance(grid, storage, active, ghost, predictor_ratio, dt, options,
                      background, radial_boundary, current_time, previous_time,
                      boundary_context);
}

// cubedsphere/cubedsphere_mpi_step.h:36-91: the six-face MPI entry points
// step the whole cube on this process's GPU; the communicator and the plans
// are not read (include/gamera_cs_mpi_dropin.h).
int gamera_cs_mpi_advance_six_with_background_planned(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context, gcs_mpi_comm communicator) {
  (void)halo_plan;
  (void)communicator;
  return gcs::advance(grid, storage, active, ghost, predictor_ratio, dt, options,
                      background, radial_boundary, current_time, previous_time,
                      boundary_context);
}

int gamera_cs_mpi_advance_six_with_background_planned_cached_edges(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan, const void *edge_plan,
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context, gcs_mpi_comm communicator) {
  (void)edge_plan;
  return gamera_cs_mpi_advance_six_with_background_planned(
      grid, storage, active, ghost, predictor_ratio, dt, options, background, halo_plan,
      radial_boundary, current_time, previous_time, boundary_context, communicator);
}

int gamera_cs_mpi_advance_six_with_background_planned_cached_edges_pre_synced(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan, const void *edge_plan,
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context, gcs_mpi_comm communicator) {
  (void)edge_plan;
  return gamera_cs_mpi_advance_six_with_background_planned(
      grid, storage, active, ghost, predictor_ratio, dt, options, background, halo_plan,
      radial_boundary, current_time, previous_time, boundary_context, communicator);
}

int gamera_cs_mpi_advance_six_with_background_sparse_pre_synced(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan, const void *edge_plan,
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context, gcs_mpi_comm communicator) {
  (void)edge_plan;
  return gamera_cs_mpi_advance_six_with_background_planned(
      grid, storage, active, ghost, predictor_ratio, dt, options, background, halo_plan,
      radial_boundary, current_time, previous_time, boundary_context, communicator);
}

int gamera_cs_advance_six_with_background(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context) {
  return gcs::advance(grid, storage, active, ghost, predictor_ratio, dt, options,
                      background, radial_boundary, current_time, previous_time,
                      boundary_context);
}

int gamera_cs_advance_six_with_boundary(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT], const size_t active[3],
    size_t ghost, double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    gamera_cs_radial_boundary_function radial_boundary, double current_time,
    double previous_time, void *boundary_context) {
  return gcs::advance(grid, storage, active, ghost, predictor_ratio, dt, options,
                      nullptr, radial_boundary, current_time, previous_time,
                      boundary_context);
}

int gamera_cs_advance_six(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                          gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                          const size_t active[3], size_t ghost,
                          double predictor_ratio, double dt,
                          const gamera_no_advance_options *options) {
  return gcs::advance(grid, storage, active, ghost, predictor_ratio, dt, options,
                      nullptr, nullptr, 0.0, 0.0, nullptr);
}

void gcs_dropin_register_earth_callbacks(
    gamera_cs_radial_boundary_function radial,
    int (*inner_fluid_flux)(gamera_no_storage *, const gamera_no_grid *,
                            const size_t *, const size_t *, void *),
    int (*no_mi_emf)(gamera_no_storage *, const gamera_no_grid *, const size_t *,
                     const size_t *, void *)) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  gcs::g_cb.radial = radial;
  gcs::g_cb.flux = inner_fluid_flux;
  gcs::g_cb.emf = no_mi_emf;
  gcs::g_cb.registered = radial != nullptr || inner_fluid_flux != nullptr ||
                         no_mi_emf != nullptr;
}

int gcs_dropin_set_resident(int on) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  gcs::g_resident = on != 0;
  return 0;
}

int gcs_dropin_sync_to_host(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                            gamera_no_storage storage[GAMERA_CS_FACE_COUNT]) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  gcs::Mirror *m = gcs::find(grid, storage);
  if (m == nullptr) {
    gcs::set_error("no device mirror for this grid/storage");
    return -1;
  }
  return m->device_current ? gcs::state_to_host(*m->P, storage) : 0;
}

int gcs_dropin_mark_host_modified(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                                  gamera_no_storage storage[GAMERA_CS_FACE_COUNT]) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  gcs::Mirror *m = gcs::find(grid, storage);
  if (m != nullptr) m->device_current = false;
  return 0;
}

int gcs_dropin_release(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                       gamera_no_storage storage[GAMERA_CS_FACE_COUNT]) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  for (size_t n = 0; n < gcs::g_mirrors.size(); ++n)
    if (gcs::g_mirrors[n]->grid == grid && gcs::g_mirrors[n]->storage == storage) {
      gcs::release(gcs::g_mirrors[n]);
      gcs::g_mirrors.erase(gcs::g_mirrors.begin() + (long)n);
      return 0;
    }
  return 0;
}

int gcs_dropin_mirror_count(void) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  return (int)gcs::g_mirrors.size();
}

long gcs_dropin_host_round_trips(void) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  return gcs::g_host_rounds;
}

}  // extern "C"

// ---- the dipole ionosphere coupling (cubedsphere/cubedsphere_mi_dipole.h) -------

namespace gcs {
namespace {

// cubedsphere_mi_dipole.c:45-57
bool valid_mi_config(const gamera_cs_mi_dipole_config *c) {
  return c != nullptr && c->azimuth_points >= 8U && c->colatitude_points >= 3U &&
         c->current_shell < 64U && std::isfinite(c->ionosphere_radius_re) &&
         c->ionosphere_radius_re > 0.0 && std::isfinite(c->pedersen_siemens) &&
         c->pedersen_siemens > 0.0 && std::isfinite(c->coupling_interval_s) &&
         c->coupling_interval_s > 0.0 && std::isfinite(c->length_norm_m) &&
         c->length_norm_m > 0.0 && std::isfinite(c->speed_norm_m_s) &&
         c->speed_norm_m_s > 0.0 && std::isfinite(c->mass_density_norm_kg_m3) &&
         c->mass_density_norm_kg_m3 > 0.0 && c->maximum_iterations > 0 &&
         std::isfinite(c->relative_tolerance) && c->relative_tolerance > 0.0;
}

// The constants of a potential view (cubedsphere_mi_dipole.c:59-85) in the
// form the device point queries read.
bool make_view(const gamera_cs_earth_startup &st, const gamera_cs_mi_dipole_config &c,
               MiView &v) {
  if (!valid_mi_config(&c) || !(c.ionosphere_radius_re < st.inner_radius) ||
      c.azimuth_points > SIZE_MAX / c.colatitude_points)
    return false;
  const double mu0 = 4.0e-7 * std::acos(-1.0);
  const double b_norm = c.speed_norm_m_s * std::sqrt(mu0 * c.mass_density_norm_kg_m3);
  v = MiView();
  v.np = (int)c.azimuth_points;
  v.nt = (int)c.colatitude_points;
  v.ri = c.ionosphere_radius_re;
  v.max_colat = std::asin(std::sqrt(c.ionosphere_radius_re / st.inner_radius));
  v.potential_norm = c.speed_norm_m_s * b_norm * c.length_norm_m;
  v.length_norm = c.length_norm_m;
  v.speed_norm = c.speed_norm_m_s;
  v.b_norm = b_norm;
  for (int e = 0; e < 3; ++e) v.moment[e] = st.dipole.moment.value[e];
  v.cut_radius = st.cut_radius;
  v.cut_length = st.cut_length;
  return v.potential_norm > 0.0 && std::isfinite(v.max_colat);
}

void destroy_handle(MiHandle *h) {
  if (h == nullptr) return;
  if (h->m != nullptr) {
    mi_destroy(*h->m->P);
    h->m->P->mi_view = MiView();
    h->m->mi = nullptr;
  }
  (void)hipFree(h->view_potential);
  (void)hipFree(h->query);
  for (size_t n = 0; n < g_mi.size(); ++n)
    if (g_mi[n] == h) g_mi.erase(g_mi.begin() + (long)n);
  delete h;
}

MiHandle *new_handle() {
  MiHandle *h = new MiHandle();
  if (hipMalloc(&h->query, 4 * sizeof(double)) != hipSuccess) {
    h->query = nullptr;
    delete h;
    set_error("MI state: device allocation failed");
    return nullptr;
  }
  g_mi.push_back(h);
  return h;
}

bool handle_ready(const MiHandle *h) {
  if (h == nullptr) return false;
  if (h->m == nullptr) return h->view_potential != nullptr && h->view_ready;
  MiInfo info;
  mi_info(*h->m->P, info);
  return info.ready;
}

// The view point queries read: the mirror problem's or the potential view's.
bool query_view(const MiHandle *h, MiView &v) {
  if (!handle_ready(h)) return false;
  if (h->m != nullptr) {
    // the problem's view with its potential, whether or not a step reads it
    MiInfo info;
    mi_info(*h->m->P, info);
    v = h->m->P->mi_view;
    v.potential = info.potential;
    return v.potential != nullptr;
  }
  v = h->view;
  v.potential = h->view_potential;
  return true;
}

int point_query(const MiHandle *h, int what, const double x[3], const double vel[3],
                double out[3]) {
  MiView v;
  if (!query_view(h, v)) return -1;
  return mi_point(v, what, x, vel, h->query, out);
}

int polar_arrays(MiHandle *h, std::vector<double> &fac, std::vector<double> &pot) {
  if (h->m == nullptr) return -1;
  return mi_download(*h->m->P, 1, fac) != 0 || mi_download(*h->m->P, 0, pot) != 0 ? -1 : 0;
}

}  // namespace
}  // namespace gcs

extern "C" {

int gamera_cs_mi_dipole_create(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                               gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                               const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
                               const gamera_cs_earth_startup *startup, const size_t active[3],
                               size_t ghost, const gamera_cs_mi_dipole_config *config,
                               gamera_cs_mi_dipole **result) {
  using namespace gcs;
  if (result == nullptr) return -1;
  *result = nullptr;
  // cubedsphere_mi_dipole.c:165-170
  if (grid == nullptr || storage == nullptr || background == nullptr || startup == nullptr ||
      active == nullptr || ghost < 4U || active[0] < 3U || active[1] < 8U ||
      active[1] != active[2] || !valid_mi_config(config) ||
      config->current_shell + 1U >= active[0] ||
      !(config->ionosphere_radius_re < startup->inner_radius) ||
      config->azimuth_points > SIZE_MAX / config->colatitude_points)
    return -1;
  std::lock_guard<std::mutex> guard(g_lock);
  Mirror *m = mirror_for(grid, storage, active, ghost, background);
  if (m == nullptr) return -1;
  if (m->mi != nullptr) destroy_handle(m->mi);  // one ionosphere per mirror
  MiHandle *h = new_handle();
  if (h == nullptr) return -1;
  Problem &P = *m->P;
  P.cfg.mi = *config;
  P.cfg.hall_siemens = 0.0;  // the reference's adapter solves Pedersen only (:407-418)
  P.cfg.startup = *startup;
  P.cfg.radial_cells = active[0];
  P.cfg.angular_cells = active[1];
  P.cfg.ghost = ghost;
  P.cfg.mi_enabled = 1;
  if (mi_create(P) != 0) {
    destroy_handle(h);
    return -1;
  }
  P.mi_view.potential = nullptr;  // read by steps only once ready
  h->m = m;
  h->cfg = *config;
  h->startup = *startup;
  std::memcpy(h->active, active, sizeof(h->active));
  h->ghost = ghost;
  m->mi = h;
  *result = reinterpret_cast<gamera_cs_mi_dipole *>(h);
  return 0;
}

void gamera_cs_mi_dipole_destroy(gamera_cs_mi_dipole *state) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  gcs::destroy_handle(gcs::find_mi(state));
}

int gamera_cs_mi_dipole_create_potential_view(const gamera_cs_earth_startup *startup,
                                              const gamera_cs_mi_dipole_config *config,
                                              gamera_cs_mi_dipole **result) {
  using namespace gcs;
  if (result == nullptr) return -1;
  *result = nullptr;
  MiView v;
  if (startup == nullptr || !make_view(*startup, *config, v)) return -1;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = new_handle();
  if (h == nullptr) return -1;
  const size_t count = 2 * config->azimuth_points * config->colatitude_points;
  if (hipMalloc(&h->view_potential, count * sizeof(double)) != hipSuccess ||
      hipMemset(h->view_potential, 0, count * sizeof(double)) != hipSuccess) {
    destroy_handle(h);
    return -1;
  }
  h->view = v;
  h->cfg = *config;
  h->startup = *startup;
  *result = reinterpret_cast<gamera_cs_mi_dipole *>(h);
  return 0;
}

int gamera_cs_mi_dipole_export_potential(const gamera_cs_mi_dipole *state, double *values,
                                         size_t count, double *epoch_s,
                                         double *next_update_s) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  if (!handle_ready(h) || values == nullptr || epoch_s == nullptr || next_update_s == nullptr)
    return -1;
  const size_t per = h->cfg.azimuth_points * h->cfg.colatitude_points;
  if (count != 2 * per) return -1;
  if (h->m != nullptr) {
    std::vector<double> pot;
    if (mi_download(*h->m->P, 0, pot) != 0) return -1;
    std::memcpy(values, pot.data(), count * sizeof(double));
    MiInfo info;
    mi_info(*h->m->P, info);
    *epoch_s = info.epoch_s;
    *next_update_s = info.next_update_s;
    return 0;
  }
  if (hipMemcpy(values, h->view_potential, count * sizeof(double), hipMemcpyDeviceToHost) !=
      hipSuccess)
    return -1;
  *epoch_s = h->view_epoch;
  *next_update_s = h->view_next;
  return 0;
}

int gamera_cs_mi_dipole_import_potential(gamera_cs_mi_dipole *state, const double *values,
                                         size_t count, double epoch_s, double next_update_s) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  // cubedsphere_mi_dipole.c:138-147
  if (h == nullptr || values == nullptr || !std::isfinite(epoch_s) ||
      !std::isfinite(next_update_s) || epoch_s < 0.0 || next_update_s <= epoch_s)
    return -1;
  const size_t per = h->cfg.azimuth_points * h->cfg.colatitude_points;
  if (count != 2 * per) return -1;
  for (size_t q = 0; q < count; ++q)
    if (!std::isfinite(values[q])) return -1;
  if (h->m != nullptr) return mi_import_potential(*h->m->P, values, epoch_s, next_update_s);
  if (h->view_potential == nullptr ||
      hipMemcpy(h->view_potential, values, count * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess)
    return -1;
  h->view_epoch = epoch_s;
  h->view_next = next_update_s;
  h->view_ready = true;
  return 0;
}

// cubedsphere_mi_dipole.c:372-378
int gamera_cs_mi_dipole_update_due(const gamera_cs_mi_dipole *state, double time_code) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  const MiHandle *h = find_mi(state);
  if (h == nullptr || !std::isfinite(time_code) || time_code < 0.0) return -1;
  const double time_s = time_code * h->cfg.length_norm_m / h->cfg.speed_norm_m_s;
  if (h->m == nullptr) return !h->view_ready || time_s + 1e-10 >= h->view_next;
  MiInfo info;
  mi_info(*h->m->P, info);
  return !info.ready || time_s + 1e-10 >= info.next_update_s;
}

// cubedsphere_mi_dipole.c:380-436: the shell current, polar samples and
// Pedersen solves on the device, from the state the storage holds now.
int gamera_cs_mi_dipole_maybe_update(gamera_cs_mi_dipole *state, double time_code) {
  using namespace gcs;
  const int due = gamera_cs_mi_dipole_update_due(state, time_code);
  if (due <= 0) return due;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  if (h == nullptr || h->m == nullptr) return -1;
  Mirror *m = h->m;
  if (!(g_resident && m->device_current)) {
    if (state_to_device(*m->P, m->storage) != 0) return -1;
    m->device_current = true;
  }
  const int rc = mi_maybe_update(*m->P, time_code);
  m->P->mi_view.potential = nullptr;  // steps enable it per call
  return rc;
}

int gamera_cs_mi_dipole_ready(const gamera_cs_mi_dipole *state) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  return gcs::handle_ready(gcs::find_mi(state)) ? 1 : 0;
}

int gamera_cs_mi_dipole_potential(const gamera_cs_mi_dipole *state, gamera_no_vec3 point,
                                  double *potential_v) {
  using namespace gcs;
  if (potential_v == nullptr) return -1;
  std::lock_guard<std::mutex> guard(g_lock);
  const double zero[3] = {0.0, 0.0, 0.0};
  double out[3];
  if (point_query(find_mi(state), 0, point.value, zero, out) != 0) return -1;
  *potential_v = out[0];
  return 0;
}

int gamera_cs_mi_dipole_ghost_velocity(const gamera_cs_mi_dipole *state, gamera_no_vec3 point,
                                       gamera_no_vec3 source_velocity,
                                       gamera_no_vec3 *ghost_velocity) {
  using namespace gcs;
  if (ghost_velocity == nullptr) return -1;
  std::lock_guard<std::mutex> guard(g_lock);
  double out[3];
  if (point_query(find_mi(state), 1, point.value, source_velocity.value, out) != 0) return -1;
  for (int c = 0; c < 3; ++c) ghost_velocity->value[c] = out[c];
  return 0;
}

// cubedsphere_mi_dipole.c:478-503 on a face's host arrays: each valid wall
// edge's EMF from the potential at its two ends, evaluated on the device.
int gamera_cs_mi_dipole_apply_inner_emf(const gamera_cs_mi_dipole *state,
                                        gamera_no_storage *storage,
                                        const gamera_no_grid *grid, size_t wall_i) {
  using namespace gcs;
  if (storage == nullptr || grid == nullptr) return -1;
  std::lock_guard<std::mutex> guard(g_lock);
  const MiHandle *h = find_mi(state);
  MiView v;
  if (!query_view(h, v)) return -1;
  const double zero[3] = {0.0, 0.0, 0.0};
  for (int d = 1; d <= 2; ++d) {
    const gamera_no_edge_grid &eg = grid->edge[d];
    if (wall_i >= eg.extent[0]) return -1;
    for (size_t j = 0; j < eg.extent[1]; ++j)
      for (size_t k = 0; k < eg.extent[2]; ++k) {
        const size_t edge = (wall_i * eg.extent[1] + j) * eg.extent[2] + k;
        if (!eg.valid[edge]) continue;
        size_t end[3] = {wall_i, j, k};
        ++end[d];
        const size_t *ve = grid->vertex_extent;
        const gamera_no_vec3 &a = grid->vertex[(wall_i * ve[1] + j) * ve[2] + k];
        const gamera_no_vec3 &b = grid->vertex[(end[0] * ve[1] + end[1]) * ve[2] + end[2]];
        double pa[3], pb[3];
        if (mi_point(v, 0, a.value, zero, h->query, pa) != 0 ||
            mi_point(v, 0, b.value, zero, h->query, pb) != 0)
          return -1;
        storage->edge_emf[d][edge] = -(pb[0] - pa[0]) / v.potential_norm;
      }
  }
  return 0;
}

int gamera_cs_mi_dipole_stats(const gamera_cs_mi_dipole *state, double *fac_min,
                              double *fac_max, double *potential_min_v,
                              double *potential_max_v, gamera_mi_solver_stats solve[2]) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  // cubedsphere_mi_dipole.c:545-569
  if (!handle_ready(h) || h->m == nullptr || fac_min == nullptr || fac_max == nullptr ||
      potential_min_v == nullptr || potential_max_v == nullptr || solve == nullptr)
    return -1;
  std::vector<double> fac, pot;
  if (polar_arrays(h, fac, pot) != 0) return -1;
  *fac_min = INFINITY;
  *fac_max = -INFINITY;
  *potential_min_v = INFINITY;
  *potential_max_v = -INFINITY;
  for (size_t q = 0; q < fac.size(); ++q) {
    *fac_min = std::fmin(*fac_min, fac[q]);
    *fac_max = std::fmax(*fac_max, fac[q]);
    *potential_min_v = std::fmin(*potential_min_v, pot[q]);
    *potential_max_v = std::fmax(*potential_max_v, pot[q]);
  }
  MiInfo info;
  mi_info(*h->m->P, info);
  for (int hemi = 0; hemi < 2; ++hemi)
    solve[hemi] = gamera_mi_solver_stats{info.iterations[hemi], info.initial_residual[hemi],
                                         info.final_residual[hemi], info.converged[hemi]};
  return 0;
}

int gamera_cs_mi_dipole_analysis_view(const gamera_cs_mi_dipole *state,
                                      gamera_cs_mi_analysis_view *view) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  if (!handle_ready(h) || h->m == nullptr || view == nullptr) return -1;
  if (polar_arrays(h, h->host_fac, h->host_potential) != 0) return -1;
  MiInfo info;
  mi_info(*h->m->P, info);
  const size_t per = h->cfg.azimuth_points * h->cfg.colatitude_points;
  view->azimuth_points = h->cfg.azimuth_points;
  view->colatitude_points = h->cfg.colatitude_points;
  view->shell_j_cells = h->active[1];
  view->shell_k_cells = h->active[2];
  view->maximum_colatitude_rad = info.max_colat;
  view->epoch_s = info.epoch_s;
  view->sample_radius_re = info.sample_radius;
  view->pedersen_siemens = h->cfg.pedersen_siemens;
  for (int hemi = 0; hemi < 2; ++hemi) {
    view->fac_parallel_a_m2[hemi] = h->host_fac.data() + hemi * per;
    view->potential_v[hemi] = h->host_potential.data() + hemi * per;
  }
  return 0;
}

int gamera_cs_mi_dipole_export_shell_current(const gamera_cs_mi_dipole *state,
                                             double *jparallel_a_m2, double *mlat_deg,
                                             double *phi_deg, size_t count) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  if (!handle_ready(h) || h->m == nullptr || jparallel_a_m2 == nullptr ||
      mlat_deg == nullptr || phi_deg == nullptr ||
      count != (size_t)GAMERA_CS_FACE_COUNT * h->active[1] * h->active[2])
    return -1;
  std::vector<double> shell;
  if (mi_shell_current(*h->m->P, shell) != 0) return -1;
  std::memcpy(jparallel_a_m2, shell.data(), count * sizeof(double));
  std::memcpy(mlat_deg, shell.data() + count, count * sizeof(double));
  std::memcpy(phi_deg, shell.data() + 2 * count, count * sizeof(double));
  return 0;
}

// cubedsphere_mi_dipole.c:620-648: the reference's CSV format.
int gamera_cs_mi_dipole_write_csv(const gamera_cs_mi_dipole *state, const char *prefix) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  if (!handle_ready(h) || h->m == nullptr) return -1;
  return write_mi_csv(*h->m->P, prefix);
}

// cubedsphere_mi_dipole.c:650-689
int gamera_cs_mi_dipole_write_shell_current_csv(const gamera_cs_mi_dipole *state,
                                                const char *prefix) {
  using namespace gcs;
  std::lock_guard<std::mutex> guard(g_lock);
  MiHandle *h = find_mi(state);
  if (!handle_ready(h) || h->m == nullptr) return -1;
  return write_shell_current_csv(*h->m->P, prefix);
}

}  // extern "C"

// ---- the driver's local CFL dt (nonorthogonal_step.h:79-84) -------------------------

namespace gcs {
namespace {

// Device copies of the caller's arrays for gamera_no_local_timestep calls
// the mirror planes cannot serve, grown as needed.
struct DtStaging {
  double *u = nullptr, *b = nullptr, *b0 = nullptr, *len = nullptr;
  size_t cap[4] = {0, 0, 0, 0};  // doubles
  unsigned long long *acc = nullptr;
};
DtStaging g_dt;
long g_dt_device_calls = 0;  // served from a resident mirror's planes

int stage(double *&dev, size_t &cap, const void *host, size_t count) {
  if (count > cap) {
    (void)hipFree(dev);
    dev = nullptr;
    cap = 0;
    if (hipMalloc(&dev, count * sizeof(double)) != hipSuccess) {
      dev = nullptr;
      set_error("local timestep: device allocation failed");
      return -1;
    }
    cap = count;
  }
  if (host != nullptr &&
      hipMemcpy(dev, host, count * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("local timestep: upload failed");
    return -1;
  }
  return 0;
}

// A resident mirror whose device state is newest and whose face `f` is this
// grid with these state arrays: its planes hold what the host arrays would.
Mirror *resident_face(const gamera_no_grid *grid, const double *conserved,
                      const gamera_no_vec3 *cell_magnetic, const gamera_no_vec3 *background,
                      const LocalDtOptions &o, int &face) {
  if (!g_resident) return nullptr;
  for (Mirror *m : g_mirrors) {
    if (!m->device_current || m->P == nullptr) continue;
    for (int f = 0; f < GAMERA_CS_FACE_COUNT; ++f) {
      if (m->grid + f != grid) continue;
      const gamera_no_storage &st = m->storage[f];
      if (st.conserved != conserved) continue;
      if (o.mhd && st.cell_magnetic != cell_magnetic) continue;
      if (o.mhd && o.background &&
          !(m->has_background && m->background[f].cell_magnetic == background))
        continue;
      face = f;
      return m;
    }
  }
  return nullptr;
}

}  // namespace
}  // namespace gcs

extern "C" {

// nonorthogonal_step.c:449-525 on the GPU. A face of a resident mirror whose
// device state is newest is read from the device planes (no transfer);
// otherwise the caller's arrays are copied to the device and reduced there.
int gamera_no_local_timestep(const gamera_no_grid *grid, const double *conserved,
                             const gamera_no_vec3 *cell_magnetic, const size_t active_lower[3],
                             const size_t active_upper[3],
                             const gamera_no_timestep_options *options,
                             const gamera_no_vec3 *background_cell_magnetic,
                             double *local_dt) {
  // nonorthogonal_step.c:454-471
  if (grid == nullptr || conserved == nullptr || options == nullptr || local_dt == nullptr ||
      !std::isfinite(options->gamma) || options->gamma <= 1.0 ||
      !std::isfinite(options->density_floor) || options->density_floor <= 0.0 ||
      !std::isfinite(options->pressure_floor) || options->pressure_floor <= 0.0 ||
      !std::isfinite(options->cfl) || options->cfl <= 0.0 ||
      (options->use_mhd && cell_magnetic == nullptr) ||
      (options->use_background && background_cell_magnetic == nullptr) ||
      (options->use_boris &&
       (!std::isfinite(options->light_speed) || options->light_speed <= DBL_MIN)))
    return -1;
  if (active_lower == nullptr || active_upper == nullptr) return -1;
  for (int a = 0; a < 3; ++a)
    if (active_lower[a] >= active_upper[a] || active_upper[a] > grid->cell_extent[a]) return -1;
  gcs::LocalDtOptions o{options->gamma,       options->density_floor, options->pressure_floor,
                        options->cfl,         options->light_speed,   options->use_mhd ? 1 : 0,
                        options->use_boris ? 1 : 0, options->use_background ? 1 : 0};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = (int)active_lower[a];
    hi[a] = (int)active_upper[a];
  }
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  double dt = 0.0;
  bool invalid = false;
  int face = 0;
  if (gcs::Mirror *m =
          gcs::resident_face(grid, conserved, cell_magnetic, background_cell_magnetic, o, face)) {
    if (gcs::local_dt_planes(*m->P, face, lo, hi, o, &dt, &invalid) != 0 || invalid) return -1;
    ++gcs::g_dt_device_calls;
    *local_dt = dt;
    return 0;
  }
  if (grid->cell == nullptr) return -1;
  const size_t cells = grid->cell_extent[0] * grid->cell_extent[1] * grid->cell_extent[2];
  gcs::DtStaging &s = gcs::g_dt;
  const bool need_b = o.mhd != 0, need_b0 = o.mhd && o.background;
  std::vector<double> len;
  len.reserve(3 * (size_t)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]));
  for (int i = lo[0]; i < hi[0]; ++i)
    for (int j = lo[1]; j < hi[1]; ++j)
      for (int k = lo[2]; k < hi[2]; ++k) {
        const size_t h = ((size_t)i * grid->cell_extent[1] + j) * grid->cell_extent[2] + k;
        for (int d = 0; d < 3; ++d) len.push_back(grid->cell[h].cfl_length[d]);
      }
  if (s.acc == nullptr && hipMalloc(&s.acc, 2 * sizeof(unsigned long long)) != hipSuccess) {
    s.acc = nullptr;
    gcs::set_error("local timestep: device allocation failed");
    return -1;
  }
  if (gcs::stage(s.u, s.cap[0], conserved, 5 * cells) != 0 ||
      (need_b && gcs::stage(s.b, s.cap[1], cell_magnetic, 3 * cells) != 0) ||
      (need_b0 && gcs::stage(s.b0, s.cap[2], background_cell_magnetic, 3 * cells) != 0) ||
      gcs::stage(s.len, s.cap[3], len.data(), len.size()) != 0)
    return -1;
  const long ext[3] = {(long)grid->cell_extent[0], (long)grid->cell_extent[1],
                       (long)grid->cell_extent[2]};
  if (gcs::local_dt_arrays(s.u, need_b ? s.b : nullptr, need_b0 ? s.b0 : nullptr, s.len, ext,
                           lo, hi, o, s.acc, &dt, &invalid) != 0 ||
      invalid)
    return -1;
  *local_dt = dt;
  return 0;
}

long gcs_dropin_local_timestep_device_calls(void) {
  std::lock_guard<std::mutex> guard(gcs::g_lock);
  return gcs::g_dt_device_calls;
}

}  // extern "C"
