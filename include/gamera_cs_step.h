/*
 * C ABI of libgcs.so, the MI355X implementation of the GAMERA-Cube
 * six-face time step.
 *
 * Problem API: a device-resident Earth problem described by a
 * gcs_problem_config (gamera_cs_problem.h). It owns everything the
 * reference's serial Earth driver owns (test_cubedsphere_dipole_background.c
 * :260-699: grid, storage, background, wind, boundary and dt rule) with the
 * arrays in device memory; gcs_problem_field copies any of them back in the
 * reference's layout.
 *
 * Every function returns 0 on success and -1 on failure, like the
 * reference; gcs_last_error() describes the last failure of the calling
 * thread (numerical failures in the reference's MHD_FAILURE form).
 */
#ifndef GAMERA_CS_STEP_H
#define GAMERA_CS_STEP_H

#include "gamera_cs_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gcs_problem gcs_problem;

/* Description of the last error on this thread. */
const char *gcs_last_error(void);

/* sizeof(gcs_problem_config) as compiled into the library. */
size_t gcs_config_size(void);

/* Setup of the serial Earth driver (:330-546) on the device: cube
 * vertices, 12-point Gauss geometry, face-frame calibration, dipole
 * background, startup state, residual fluxes and the halo/seam plans. */
gcs_problem *gcs_problem_create(const gcs_problem_config *config);
void gcs_problem_destroy(gcs_problem *problem);

/* `steps` iterations of the driver loop (:589-699): CFL dt rule, then
 * gamera_cs_advance_six_with_background_planned (cubedsphere_step.c:13)
 * with the Earth radial-ghost, inner-flux and EMF callbacks. */
int gcs_problem_advance(gcs_problem *problem, long steps);

/* One six-face step with caller-chosen predictor ratio, dt and boundary
 * times (the arguments of cubedsphere_step.h:58-67); the clock is kept. */
int gcs_problem_step_explicit(gcs_problem *problem, double predictor_ratio,
                              double dt, double current_time,
                              double previous_time);

/* time, dt of the last step, previous dt, completed steps. */
void gcs_problem_clock(gcs_problem *problem, double out[4]);

/* Host copy of field `field` (GCS_FIELD_*) of cube face `face`, in the
 * reference's array-of-structs layout; valid until the next step. */
int gcs_problem_field(gcs_problem *problem, int face, int field,
                      double **data, size_t *count);
int gcs_problem_edge_valid(gcs_problem *problem, int face, int direction,
                           unsigned char **data, size_t *count);

/* Emergency-diffusion counters of one cube face: applied-face count and
 * maximum interface speed (nonorthogonal_storage.h:29-30). */
void gcs_problem_counters(gcs_problem *problem, int face, double out[2]);

/* 1 (default): also reproduce the ghost values the reference leaves in
 * storage after a step (ghost U and face flux keep their start-of-step
 * values, ghost B is recovered from them). 0: skip that work and the
 * ghost-cell B of the halo fill and of the Earth boundary's recovery (no
 * step computation reads ghost-cell B); ghost U and face fluxes are
 * rewritten at the start of each step, ghost B is not maintained, nor are
 * the predicted face flux, the predicted momentum and energy (the step
 * reads only the predicted density) and the stress rates; the active state,
 * clock and ionosphere potential are bit-identical to fidelity 1's. No
 * reference counterpart. */
int gcs_problem_set_fidelity(gcs_problem *problem, int on);

/* 0 (default): FP64 throughout, the drop-in's arithmetic. 1: the mixed
 * form: Up7/PDM reconstruction, face fluxes, face velocities and edge EMFs
 * computed in FP32, with the conserved-variable update, CT, CFL reduction
 * and ionosphere solve in FP64. The predictor's primitive caches, the face
 * and edge geometry the FP32 kernels read, the face velocities and the face
 * fluxes are held in FP32 arrays (allocated on the first switch, 137 floats
 * per grid point); edge EMFs stay FP64. Downloads of the predicted state and
 * the face fluxes widen the FP32 values. No reference counterpart. */
int gcs_problem_set_precision(gcs_problem *problem, int mixed);

/* gamera_no_apply_pressure_control (nonorthogonal_step.h:26-33,
 * NO/nonorthogonal_step.c:74-171) on the current state's active cells, as
 * the MPI Earth driver applies it after a step
 * (cubedsphere/tests/test_cubedsphere_mpi_earth_tile.c:459-479): cells at
 * or below low_density_threshold take the pressure of
 * low_density_sound_speed; cells whose sound speed exceeds 1.5 light_speed
 * relax toward it over a dipole bounce time. counts = (low-density cells,
 * high-sound-speed cells), maxima = (maximum sound speed, maximum pressure)
 * as the reference reports them (over all faces). Returns -1 on invalid
 * arguments (nothing changed) or on a cell the reference would reject
 * (MHD_FAILURE message). */
int gcs_problem_pressure_control(gcs_problem *problem, double dt, double gamma,
                                 double density_floor, double pressure_floor,
                                 double light_speed, double low_density_threshold,
                                 double low_density_sound_speed,
                                 unsigned long long counts[2], double maxima[2]);

/* gcs_problem_pressure_control without any device allocation, for callers
 * that cannot allocate (code under stream capture, tight driver loops that
 * apply the control after every step). No counterpart in the reference.
 * Same arguments, cells and results. Its scratch is a static device array
 * of 2 MiB per GPU, held until the process ends: at most 16,777,216 active
 * cells (6 x 96^2 x 192 fits; larger problems are rejected with an error and
 * use gcs_problem_pressure_control). Calls that overlap on one GPU share the
 * array and race. */
int gcs_problem_pressure_control_noalloc(
    gcs_problem *problem, double dt, double gamma, double density_floor,
    double pressure_floor, double light_speed, double low_density_threshold,
    double low_density_sound_speed, unsigned long long counts[2],
    double maxima[2]);

/* 1: run each driver step (outside profiled timing) as one hipGraph launch,
 * captured per time-level parity on first use and recaptured when the
 * precision, fidelity, ghost-recovery state or ionosphere view changes.
 * 0 (default): launch the step's kernels one by one. Results are identical
 * either way. No reference counterpart. */
int gcs_problem_set_graph(gcs_problem *problem, int on);

/* The serial Earth driver's end-of-run gate
 * (test_cubedsphere_dipole_background.c:807-893) on the current state:
 * out = {minimum density, minimum pressure (primitive conversion with
 * floors 1e-12, 1e-12), maximum total (B1 + B0) div B per volume} over the
 * active cells. The driver passes when out[0] > 0, out[1] > 0 and
 * out[2] < 1e-9. Returns -1 on a cell the conversion rejects. */
int gcs_problem_final_gates(gcs_problem *problem, double out[3]);

/* Replace the problem's solar-wind series by `count` samples (any number;
 * the config holds at most GCS_MAX_WIND_SAMPLES): time[n] in code time,
 * state[8 n .. 8 n + 7] = density, pressure, velocity (3), magnetic field
 * (3); linear != 0 interpolates between samples, else holds each sample
 * (gamera_solar_wind_set, nonorthogonal/src/solar_wind.c:55-80; sampled
 * with the ballistic delay of :234-275 at the series' reference point).
 * Times must be finite and increasing, states finite. */
int gcs_problem_set_wind(gcs_problem *problem, size_t count, const double *time,
                         const double *state, int linear);

/* 1: run every face-flux, face-velocity and edge-EMF sweep with one face or
 * edge per thread (the untiled kernels), in FP64 and in the mixed form; the
 * seam recovery reads its cells' static data from the planes, and the Earth
 * boundary computes the ionosphere drift of its ghost cells in every step.
 * 0 (default): the LDS-tiled and register-march kernels, the seam data
 * gathered once per problem and the drift cached per potential. Both compute
 * each value with the same expressions, so results are bit-identical; the
 * switch lets tests hold the fast kernels to the plain ones at any size. No
 * reference counterpart. */
int gcs_problem_set_plain_kernels(gcs_problem *problem, int on);

/* Sample the field-aligned current of the current state onto the MI polar
 * grids (cubedsphere_mi_dipole.c:223-370) without solving; the potential
 * and the update clock are unchanged. gcs_problem_mi_field(what = 1) then
 * returns it. No reference counterpart as a separate call. */
int gcs_problem_mi_sample_current(gcs_problem *problem);

/* max over active cells of |net total (B1 + B0) face flux| / volume of the
 * current level: the divergence gate of the reference's Earth driver
 * (test_cubedsphere_dipole_background.c:808-845), computed on the device. */
int gcs_problem_max_div_b(gcs_problem *problem, double *out);

/* Timing of `steps` driver steps with HIP events on the problem's stream:
 * total milliseconds, and with profile != 0 the milliseconds of each step
 * stage summed over the steps (gcs_stage_count() entries named by
 * gcs_stage_name). No reference counterpart (it times with MPI_Wtime,
 * cubedsphere_profile.h:7-38). */
int gcs_problem_time_steps(gcs_problem *problem, long steps, int profile,
                           double *elapsed_ms, double *stage_ms);
int gcs_stage_count(void);
const char *gcs_stage_name(int stage);

/* The serial Earth driver's six-face CSV snapshot of the current state
 * (cubedsphere/tests/test_cubedsphere_dipole_background.c:120-161,
 * write_earth_snapshot): header "face,i,j,k,x,y,z,rho,p,vx,vy,vz,bx,by,bz",
 * then per active cell in face, i, j, k order the active-relative indices,
 * centroid, primitive state with floors 1e-12 and total cell B, %.17g; with
 * the MI coupling on, the ionosphere's files path.mi_north.csv,
 * path.mi_south.csv and path.shell_j.csv follow (cubedsphere_mi_dipole.c:
 * 620-690). Values are computed on the device, formatted on the host. -1
 * when a cell's state does not convert (the rows before it are written, as
 * the reference leaves them) or a file cannot be written. */
int gcs_problem_write_snapshot(gcs_problem *problem, const char *path);

/* test_cubedsphere_dipole_background.c:177-193 (write_drive_frame): the
 * snapshot "<prefix>_<frame, 6 digits>.csv" and the line
 * "frame,step,simulation_seconds,path" appended to index_path, whose header
 * (:561) is written first when the file is new or empty. */
int gcs_problem_write_drive_frame(gcs_problem *problem, const char *prefix, long frame,
                                  long step, double seconds, const char *index_path);

/* test_cubedsphere_dipole_background.c:61-118 (load_mhd_snapshot), the
 * driver's earth-mi-replay input: every active cell's conserved state
 * (primitive conversion with floors 1e-12) and residual cell B (total minus
 * the dipole background) of both time levels from a snapshot written as
 * above. Same checks as the reference: exact header, every line parsed by
 * the same format naming an active cell whose centroid is within 1e-8, one
 * line per active cell. -1 when any check fails; unlike the reference,
 * which writes the lines before a bad one, the problem is then unchanged.
 * Face fluxes, the clock and the ionosphere are not touched (a later step
 * recovers cell B from the face fluxes, as the reference's boundary does). */
int gcs_problem_load_snapshot(gcs_problem *problem, const char *path);

/* gamera_cs_mi_dipole_maybe_update (cubedsphere_mi_dipole.h:59-60) of a
 * problem created with mi_enabled, at code time time_code: *updated = 1 when
 * the cadence called for a solve (which ran), 0 otherwise. With
 * gcs_problem_load_snapshot it replays the driver's earth-mi-replay
 * (:526-533). */
int gcs_problem_mi_maybe_update(gcs_problem *problem, double time_code, int *updated);

/* The ionosphere's CSV files of a problem created with mi_enabled
 * (gamera_cs_mi_dipole_write_csv and _write_shell_current_csv,
 * cubedsphere_mi_dipole.c:620-690): prefix.mi_north.csv,
 * prefix.mi_south.csv and prefix.shell_j.csv. */
int gcs_problem_write_mi_csv(gcs_problem *problem, const char *prefix);

/* Measurement helpers, no reference counterpart: the HBM bandwidth of a
 * plain device-to-device copy of `doubles` doubles on the current device
 * (read + write bytes per second over `iterations` launches, HIP events),
 * and a wait for all work queued on the current device. */
int gcs_copy_bandwidth(size_t doubles, int iterations, double *gb_per_s);
int gcs_device_synchronize(void);

/* Ionosphere coupling of a problem created with mi_enabled (the driver's
 * gamera_cs_mi_dipole_*, cubedsphere_mi_dipole.h:51-98): ints = {updates,
 * ready, solver iterations north, south}; doubles = {final residual north,
 * south, epoch in s, last solve in ms}. */
int gcs_problem_mi_report(gcs_problem *problem, int ints[4], double doubles[4]);
/* what = 0: potential in V, 1: sampled field-aligned current in A/m^2,
 * [hemisphere][colatitude][azimuth] like
 * gamera_cs_mi_dipole_export_potential; valid until the next call. */
int gcs_problem_mi_field(gcs_problem *problem, int what, double **data,
                         size_t *count);

#ifdef __cplusplus
}
#endif

#endif
