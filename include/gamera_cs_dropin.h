/*
 * Drop-in replacement of the reference's serial six-face step on the GPU.
 *
 * The four gamera_cs_advance_six* functions have the reference's prototypes,
 * argument meaning and return convention (0 success, -1 failure, with
 * invalid arguments rejected before anything is written), so a program
 * written against cubedsphere/cubedsphere_step.h links against libgcs.so in
 * place of cubedsphere_step.c. The other reference modules the caller uses
 * (grid and storage allocation, setup, dt rule, Earth boundary, output)
 * stay as they are; only the step moves to the GPU.
 *
 * libgcs.so keeps a device mirror of every (grid, storage) pair it is
 * called with. Geometry and background are uploaded on the first call and
 * must not change afterwards (gcs_dropin_release drops the mirror). By
 * default each call uploads the state and downloads every array the
 * reference's step writes; see gcs_dropin_set_resident for the faster mode.
 */
#ifndef GAMERA_CS_DROPIN_H
#define GAMERA_CS_DROPIN_H

#include "gamera_cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cubedsphere/cubedsphere_step.h:16-21 (cubedsphere_step.c:180-193): no
 * background, no radial boundary (radial outflow ghosts only). */
int gamera_cs_advance_six(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options);

/* cubedsphere/cubedsphere_step.h:33-40 (cubedsphere_step.c:167-178). */
int gamera_cs_advance_six_with_boundary(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    gamera_cs_radial_boundary_function radial_boundary,
    double current_time, double previous_time, void *boundary_context);

/* cubedsphere/cubedsphere_step.h:46-54 (cubedsphere_step.c:156-165). */
int gamera_cs_advance_six_with_background(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    gamera_cs_radial_boundary_function radial_boundary,
    double current_time, double previous_time, void *boundary_context);

/* cubedsphere/cubedsphere_step.h:58-67 (cubedsphere_step.c:13-154). The
 * halo plan is accepted and not read: the device builds its own plan from
 * the same grid. Any ghost >= 4; MHD and hydrodynamic
 * (options->update.use_mhd = false) steps. */
int gamera_cs_advance_six_with_background_planned(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    gamera_cs_radial_boundary_function radial_boundary,
    double current_time, double previous_time, void *boundary_context);

/* NO/nonorthogonal_step.h:79-84 (nonorthogonal_step.c:449-525): the local
 * CFL dt of one face's box, computed on the GPU. A face of a resident
 * mirror whose device state is newest (gcs_dropin_set_resident) is read
 * from the device, with no transfer; otherwise the caller's arrays are
 * copied to the device first. Same validation and return values. A program
 * that also links the reference's nonorthogonal_step.c keeps the
 * reference's definition; see the integration notes for taking this one. */
int gamera_no_local_timestep(const gamera_no_grid *grid, const double *conserved,
                             const gamera_no_vec3 *cell_magnetic,
                             const size_t active_lower[3], const size_t active_upper[3],
                             const gamera_no_timestep_options *options,
                             const gamera_no_vec3 *background_cell_magnetic,
                             double *local_dt);

/* ---- mirror control (no reference counterpart) ------------------------- */

/* Callbacks that run as device kernels. By default libgcs.so looks up the
 * reference's gamera_cs_earth_fill_radial_ghosts,
 * gamera_cs_earth_inner_fluid_flux, gamera_cs_earth_no_mi_emf and
 * gamera_cs_earth_tile_emf (the last on a single radial tile only;
 * cubedsphere_earth_boundary.h:26-63) in the global symbol table; a caller
 * that loads the first three privately names them here. Other callbacks
 * run on the host between device phases. */
void gcs_dropin_register_earth_callbacks(
    gamera_cs_radial_boundary_function radial,
    int (*inner_fluid_flux)(gamera_no_storage *, const gamera_no_grid *,
                            const size_t *, const size_t *, void *),
    int (*no_mi_emf)(gamera_no_storage *, const gamera_no_grid *,
                     const size_t *, const size_t *, void *));

/* 1: keep the state on the device between calls; the host storage is
 * current only after gcs_dropin_sync_to_host. Host changes to the state
 * must be announced with gcs_dropin_mark_host_modified. 0 (default):
 * upload and download on every call. */
int gcs_dropin_set_resident(int on);
int gcs_dropin_sync_to_host(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                            gamera_no_storage storage[GAMERA_CS_FACE_COUNT]);
int gcs_dropin_mark_host_modified(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT]);

/* Free the device mirror of a grid/storage pair (before freeing them). */
int gcs_dropin_release(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                       gamera_no_storage storage[GAMERA_CS_FACE_COUNT]);
int gcs_dropin_mirror_count(void);

/* Callbacks that ran on the host (the state downloaded before and uploaded
 * after) since the library was loaded; zero for a step whose boundary work
 * all runs on the device. No reference counterpart. */
long gcs_dropin_host_round_trips(void);

/* gamera_no_local_timestep calls served from a resident mirror's device
 * state (no transfer) since the library was loaded. No reference
 * counterpart. */
long gcs_dropin_local_timestep_device_calls(void);

#ifdef __cplusplus
}
#endif

#endif
