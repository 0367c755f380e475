// Host-side owner of one device-resident six-face problem.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "gamera_cs_problem.h"
#include "gcs_layout.hpp"
#include "gcs_planes.hpp"

namespace gcs {

#define GCS_HIP(call)                                                      \
  do {                                                                     \
    hipError_t err_ = (call);                                              \
    if (err_ != hipSuccess) {                                              \
      gcs::set_error(std::string(#call) + ": " + hipGetErrorString(err_)); \
      return -1;                                                           \
    }                                                                      \
  } while (0)

void set_error(const std::string &message);
const char *last_error();

// Scalars the step kernels read from device memory, so that a captured step
// graph replays with the dt the previous step's CFL reduction chose.
struct StepScalars {
  double dt, ratio, time, previous_time, previous_dt;
  double wall_time, wall_dt;  // Earth wall EMF clock (gamera_cs_earth_boundary)
  unsigned long long cfl_min_bits; // min local CFL dt (IEEE bits, > 0)
  long long steps;
  unsigned int failed;     // nonzero: a kernel hit a reference failure path
  unsigned int fail_stage; // first failing stage (kernel id)
  int fail_index;          // flat index of the first failing element
  int pad;
  unsigned long long emergency_count[6];
  unsigned long long emergency_speed_bits[6];  // max, as IEEE bits (>= 0)
};

// A solar-wind time series of any length for device sampling
// (gamera_solar_wind_series, solar_wind.h:24-35): the samples live in a
// device buffer of the problem.
struct WindSeries {
  int count, linear;
  const double *time;   // device, count sample times
  const double *state;  // device, count x 8: rho, p, v[3], b[3]
  double reference[3];
  double time_offset, by, bz;
};

// Halo plans on the device (built on the host from the cube geometry).
struct HaloCell {        // tangential ghost cell: 2 x 2 donor block
  int target;            // flat index (layout.at) of the receiving cell
  int donor;             // flat index of donor (p, q) = (0, 0) of the block
  int nearest;           // flat index of the nearest-donor fallback
  int pad;
  double w[4];           // wx[p] * wy[q] for (0,0), (0,1), (1,0), (1,1)
};
struct HaloFace {        // ghost magnetic face
  int target;            // flat index of the face
  int donor;             // flat index of the donor block corner cell
  int dir;
  int pad;
  double w[4];
};
struct SeamPair {        // one shared physical face seen from two cube faces
  int a, b;              // flat face indices
  int dir_a, dir_b;
  double sign;           // orientation of b relative to a
};
struct EdgeMember {
  int index;             // flat edge index
  int dir;
  double sign;
};

// Physics options of one step (gamera_no_advance_options after the
// consistency checks of cubedsphere_step.c:23-47) and of the Earth boundary
// context (cubedsphere_earth_boundary.h:13-24).
struct StepOptions {
  // sweep (gamera_no_sweep_options): reconstruction, fluxes, HOGS
  double gamma, rho_floor, p_floor, pdm, hogs_h, hogs_m, c;
  int boris, hogs;
  // edge EMF (gamera_no_emf_options)
  double emf_rho_floor, emf_pdm, emf_diff, emf_c, emf_cfl;
  int emf_boris;
  // predictor and update (gamera_no_update_options)
  double upd_gamma, upd_rho_floor, upd_p_floor, upd_c;
  int upd_boris;
  // dt rule of the problem driver (gamera_no_timestep_options)
  double ts_gamma, ts_rho_floor, ts_p_floor, ts_cfl, ts_c;
  int ts_boris;
  // Earth wall (startup gamma, context floors and cfl)
  double wall_gamma, wall_rho_floor, wall_p_floor, wall_cfl;
  int strict_wall;
  // 1: MHD step; 0: hydrodynamic only (use_mhd = false: no Maxwell flux,
  // EMF, CT, B recovery or Boris; B and face flux carried unchanged)
  int mhd = 1;
  // 1: the update also stores the stress rates (hydro 5, Maxwell 3) in the
  // face-velocity planes, for downloads of hydro_rate / maxwell_rate
  int rates = 0;
};


// The ionosphere potential as the step kernels read it
// (cubedsphere_mi_dipole.c:442-543); potential == nullptr: not ready.
struct MiView {
  const double *potential = nullptr;  // [2][colatitude][azimuth], volts
  int np = 0, nt = 0;
  double ri = 0.0, max_colat = 0.0;   // ionosphere radius (R_E), cap
  double potential_norm = 0.0, length_norm = 0.0, speed_norm = 0.0,
         b_norm = 0.0;
  double moment[3] = {0.0, 0.0, 0.0}; // startup dipole and radial cut
  double cut_radius = 0.0, cut_length = 0.0;
  // E x B drift of the inner radial ghost cells for `potential`, field-major
  // [3][6][ghost][NJ][NK] (NaN where it fails), or null: then the step
  // kernels compute it. Set by make_k from the problem's MI state.
  const double *drift = nullptr;
};

struct MiState;  // owner of the MI arrays and solver (gcs_mi.hip)

// Which boundary work runs on the device inside a step.
struct StepBoundary {
  bool outflow = false;     // gamera_cs_fill_radial_outflow_cells
  bool earth_cells = false; // gamera_cs_earth_fill_radial_ghosts
  bool inner_trap = false;  // gamera_cs_earth_inner_fluid_flux
  bool earth_emf = false;   // gamera_cs_earth_no_mi_emf
};

// Slots of the CFL minimum the update and the seam recovery fold their
// cells' dt into (positive doubles order like their bit patterns, and min
// is exact, so neither the order nor the slot of a contribution matters).
constexpr int kCflSlots = 1024;

struct Problem {
  gcs_problem_config cfg;
  StepOptions opt;
  StepBoundary bnd;
  Layout L;
  double *planes = nullptr;        // PLANES * L.total doubles (planes_alloc + plane_shift)
  double *planes_alloc = nullptr;
  float *planes32 = nullptr;       // PLANES32 * L.total floats (mixed form)
  float *planes32_alloc = nullptr; // its allocation (planes32 - plane_shift32)
  uint8_t *bytes = nullptr;        // BYTE_PLANES * L.total
  StepScalars *scalars = nullptr;  // device
  unsigned long long *cfl_partial = nullptr;  // per-block CFL minima
  unsigned long long *cfl_slots = nullptr;    // kCflSlots CFL minima folded by the update
  unsigned long long *acc4 = nullptr;         // 4 accumulator words (div B, gates, local dt)
  unsigned long long *pc_partial = nullptr;   // pressure-control block partials (on first use)
  // Driver problems fold the new state's CFL dt into cfl_slots during the
  // update (drop-in mirrors take dt from the caller); cfl_folded: the slots
  // hold the current state's minimum.
  bool driver_cfl = false, cfl_folded = false;
  int level = 0;                   // which U/B/PHI level is current
  long steps = 0;
  double time = 0.0, dt = 0.0, previous_dt = 0.0;
  // plans
  HaloCell *halo_cells = nullptr;
  int n_halo_cells = 0;
  HaloFace *halo_faces = nullptr;
  int n_halo_faces = 0;
  SeamPair *seams = nullptr;
  int n_seams = 0;
  EdgeMember *edge_members = nullptr;
  int *edge_groups = nullptr;      // CSR offsets, n_edge_groups + 1
  int n_edge_groups = 0;
  int n_edge_members = 0;
  // host mirrors of reference-layout arrays (allocated on first request)
  std::vector<std::vector<double>> host_fields;
  std::vector<std::vector<uint8_t>> host_valid;
  bool host_dirty = true;
  hipStream_t stream = nullptr;
  // Captured driver steps, one per time-level parity, with what the capture
  // baked into its kernel arguments (recaptured when any of it changes).
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    bool mixed = false, fidelity = false, b_recovered = false, old_halo_current = false,
         cfl_folded = false, plain = false;
    const double *mi_potential = nullptr;
  };
  bool use_graph = false;
  StepGraph graphs[2];
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  std::vector<double> kernel_ms;   // per-stage timing (profiling mode)
  int *seam_cells = nullptr;       // active cells with a seam face
  double *seam_geom = nullptr;     // their static recovery and CFL data
  int n_seam_cells = 0;
  WindSeries wind_series;
  double *wind_buffer = nullptr;   // device: times, then 8 values per sample
  size_t wind_capacity = 0;        // samples the buffer holds
  bool fidelity = true;            // reproduce ghost values of the storage
  bool mixed = false;              // FP32 reconstruction, fluxes and EMF
  // One face or edge per thread in every face-flux, face-velocity and EMF
  // sweep instead of the LDS tiles and register marches (same results).
  bool plain = false;
  // The last step stored the stress rates of the active cells in the
  // face-velocity planes (fidelity steps).
  bool rates_stored = false;
  // The predictor caches and face fluxes of the last step are in planes32.
  bool last_step_mixed = false;
  // Cell B of both levels equals the recovery from their face fluxes (true
  // after a completed step; false at creation and after a host upload).
  bool b_recovered = false;
  // The old level's ghost face fluxes are those of its active B (true after
  // a step whose ghost phase did not recover every cell): a level's active B
  // does not change while it is the old level, so fidelity 0 fills the
  // ghost fluxes of the current level only.
  bool old_halo_current = false;
  // Stage timing: when `profiling`, enqueue_step records an event before
  // every stage into `stage_events` (one row of kMaxStages per step).
  bool profiling = false;
  std::vector<hipEvent_t> stage_events;
  std::vector<int> stage_ids;      // stage id of each recorded event
  size_t stage_cursor = 0;
  double *pinned = nullptr;        // pinned host buffer for plane transfers
  size_t pinned_count = 0;         // its size in doubles
  MiState *mi = nullptr;           // ionosphere coupling (mi_enabled)
  MiView mi_view;                  // what the step kernels read

  double *plane(int p) const { return planes + (long)p * L.total; }
  uint8_t *byte_plane(int q) const { return bytes + (long)q * L.total; }
};

// Stages of one step, in launch order (names for timing reports).
enum Stage {
  S_CFL, S_DT, S_SYNC_PHI, S_HALO_CELLS, S_OUTFLOW, S_HALO_FACES,
  S_EARTH_CELLS, S_EARTH_FACES, S_EARTH_RECOVER, S_PREDICT, S_FACE_FLUX0,
  S_FACE_FLUX1, S_FACE_FLUX2, S_SYNC_FLUX, S_INNER_TRAP, S_FACE_VELOCITY,
  S_EDGE_EMF0, S_EDGE_EMF1, S_EDGE_EMF2, S_EARTH_EMF, S_SYNC_EDGES, S_UPDATE,
  S_FINAL_SYNC, S_SEAM_RECOVER, S_FIDELITY, S_CLOCK, S_END, S_MI, kStages
};
const char *stage_name(int stage);
// Record a stage boundary on the problem's stream when profiling.
void mark_stage(Problem &P, int stage);
int driver_step(Problem &P);

// Device buffers, stream and events of a problem whose layout L is set;
// scalars are initialised. free_problem releases everything.
int allocate_problem(Problem &P);
// Allocate planes32 and fill its geometry planes (first use of the mixed
// form); widen_mixed copies the mixed form's caches and face fluxes into
// their FP64 planes for download.
int prepare_mixed(Problem &P);
int widen_mixed(Problem &P);
// gamera_no_apply_pressure_control on the current state; args = dt, gamma,
// density floor, pressure floor, light speed, low-density threshold,
// low-density sound speed.
int pressure_control(Problem &P, const double args[7], unsigned long long counts[2],
                     double maxima[2]);
void free_problem(Problem *P);
int init_scalars(Problem &P);
int read_scalars(Problem &P, StepScalars &s);
// 0, or -1 with the first failing kernel stage and element in the reference's
// MHD_FAILURE wording; updates the host clock mirror.
int check_failure(Problem &P);

// Ionosphere coupling (gcs_mi.hip): create from P.cfg after setup; before
// a step at code time t, update the potential when the cadence is due
// (cubedsphere_mi_dipole.c:372-436). Returns 1 updated, 0 not due, -1.
int mi_create(Problem &P);
int mi_maybe_update(Problem &P, double time_code);
int mi_sample_current(Problem &P);
void mi_destroy(Problem &P);
struct MiReport {
  int updates, ready;
  int iterations[2];
  double final_residual[2], epoch_s, solve_ms;
};
void mi_report(const Problem &P, MiReport &out);
int mi_download(Problem &P, int what, std::vector<double> &out);  // 0 potential, 1 fac
// The MI state as the reference's gamera_cs_mi_dipole_* accessors report it.
struct MiInfo {
  bool ready = false;
  int updates = 0;
  double epoch_s = 0.0, next_update_s = 0.0, sample_radius = 0.0, max_colat = 0.0,
         potential_norm = 0.0;
  int iterations[2] = {0, 0};
  double initial_residual[2] = {0.0, 0.0}, final_residual[2] = {0.0, 0.0};
  int converged[2] = {0, 0};
  const double *potential = nullptr;  // device [2][nt][np]
};
void mi_info(const Problem &P, MiInfo &out);
// Let the step kernels read the potential (on: when the state is ready) or
// not (the Earth context of a call has no ionosphere).
void mi_view_enable(Problem &P, bool on);
// Replace the potential ([2][nt][np], V) and the update clock; ready after.
int mi_import_potential(Problem &P, const double *values, double epoch_s,
                        double next_update_s);
// J-parallel, magnetic latitude, longitude of the active sample-shell cells
// at the last update, [3][6][n][n].
int mi_shell_current(Problem &P, std::vector<double> &out);
// Evaluate gamera_cs_mi_dipole_potential (what 0) or _ghost_velocity (what
// 1) at one point on the device from a potential view; out = {value or
// velocity[3]}; returns 0, or -1 where the reference's function would.
int mi_point(const MiView &view, int what, const double x[3], const double v[3],
             double *result, double out[3]);

// Setup: geometry, background, initial state and plans.
int setup_problem(Problem &P);
// Cube face and angular coordinates of a direction, with the C library's
// atan2 like the reference's plans (cubedsphere_mesh.c:77-103).
bool cube_inverse(const double x[3], int &face, double &a, double &b);
int build_plans(Problem &P);
// Recompute the inner ghost cells' drift for `potential` into `drift`
// (the MiView::drift layout), on the problem's stream.
int refresh_mi_drift(Problem &P, const double *potential, double *drift);
// The problem's cached drift when the step view reads its own potential.
const double *mi_drift_cache(const Problem &P);
// Gather the seam cells' static recovery and CFL data (after build_plans'
// seam list and the geometry).
int gather_seam_geometry(Problem &P);
// Options and boundary of the serial Earth driver
// (test_cubedsphere_dipole_background.c:405-449).
void driver_options(Problem &P);
// One step of the driver loop (dt rule + six-face step + clock) on
// P.stream; with explicit_dt the step uses the dt, ratio and times already
// in the device scalars and leaves the clock alone.
int enqueue_step(Problem &P, bool explicit_dt);
// The same six-face step in four phases, so that host callbacks can run
// between them (cubedsphere_step.c:49-152): ghost fill (through the radial
// boundary), face fluxes (through fluid_flux_sync), edge EMFs (through
// edge_emf_sync), then seam EMF average, update and final seam sync. The
// last phase swaps the time levels.
enum Phase { PHASE_GHOSTS, PHASE_FLUX, PHASE_EMF, PHASE_UPDATE };
int max_div_b(Problem &P, double *out);
// {min density, min pressure, max total div B} of the current state's
// active cells (the Earth driver's end-of-run gate).
int final_gates(Problem &P, double gamma, double out[3]);
// The drivers' CSV output (gcs_output.cpp): the six-face snapshot
// (test_cubedsphere_dipole_background.c:120-161), an output frame with its
// index line (:177-193), and the ionosphere's CSV files
// (cubedsphere_mi_dipole.c:620-690).
int write_snapshot(Problem &P, const char *path);
int write_drive_frame(Problem &P, const char *prefix, long frame, long step, double seconds,
                      const char *index_path);
int write_mi_csv(Problem &P, const char *prefix);
int write_shell_current_csv(Problem &P, const char *prefix);
int load_snapshot(Problem &P, const char *path);
// Read + write bytes per second of a plain device copy of `doubles` doubles
// on the current device, timed with HIP events over `iterations` launches.
int copy_bandwidth(size_t doubles, int iterations, double *gb_per_s);
// The options of gamera_no_local_timestep (gamera_no_timestep_options,
// nonorthogonal_step.h:66-76).
struct LocalDtOptions {
  double gamma, density_floor, pressure_floor, cfl, light_speed;
  int mhd, boris, background;
};
// Minimum local CFL dt (nonorthogonal_step.c:449-525) over the cells
// lo <= (i, j, k) < hi of one cube face, from the current level's device
// planes; *invalid when a cell's state or dt is invalid (where the
// reference returns -1). Returns -1 on a device error.
int local_dt_planes(Problem &P, int face, const int lo[3], const int hi[3],
                    const LocalDtOptions &o, double *dt, bool *invalid);
// The same over device copies of arrays in the reference's host layout:
// u 5 and b, b0 3 values per cell of the extent ext (b, b0 may be null
// where the options do not read them), len the box cells' three CFL
// lengths (k fastest); acc is 2 device words of scratch.
int local_dt_arrays(const double *u, const double *b, const double *b0, const double *len,
                    const long ext[3], const int lo[3], const int hi[3],
                    const LocalDtOptions &o, unsigned long long *acc, double *dt,
                    bool *invalid);
// Upload a wind series (count samples: time[n], state[8 n .. 8 n + 7] =
// rho, p, v, B) into the problem's device buffer, growing it when needed,
// and point P.wind_series at it; the other fields are set by the caller.
int upload_wind(Problem &P, int count, const double *time, const double *state);
// pressure_control with no device allocation: scratch in a static device
// array (at most 16,777,216 active cells; calls overlapping on one GPU race).
int pressure_control_noalloc(Problem &P, const double args[7], unsigned long long counts[2],
                             double maxima[2]);
int enqueue_phase(Problem &P, Phase phase);
// Constants derived from uploaded geometry and background: recovery
// pivots, edge-B solve, B0 normal and edge-tangential components.
int derive_constants(Problem &P);
// Copy one reference-layout field of one cube face to the host.
int download_field(Problem &P, int face, int field, std::vector<double> &out);
int download_valid(Problem &P, int face, int dir, std::vector<uint8_t> &out);
int upload_field(Problem &P, int face, int field, const double *data,
                 size_t count);

}  // namespace gcs
