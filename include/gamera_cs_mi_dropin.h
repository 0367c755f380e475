/*
 * The reference's dipole ionosphere (MI) coupling, exported by libgcs.so
 * with the prototypes of cubedsphere/cubedsphere_mi_dipole.h:30-98, so that
 * a program links libgcs.so in place of cubedsphere/cubedsphere_mi_dipole.c
 * as well as cubedsphere/cubedsphere_step.c.
 *
 * A state made by gamera_cs_mi_dipole_create belongs to the device mirror of
 * its (grid, storage) pair (the same mirror the drop-in step uses): the shell
 * current, the polar samples and the Pedersen solve run on the GPU, and an
 * Earth context whose `mi` is such a state keeps the whole step on the
 * device (radial ghosts from the potential's E x B drift, wall EMF from its
 * differences). Return values and the failure cases are the reference's.
 * Differences:
 *  - the solve is the device's Jacobi-PCG: its dot products reduce in a
 *    different order, so potentials agree with the reference's to the
 *    solver tolerance and iteration counts can differ by a few;
 *  - maybe_update uploads the host state first unless the drop-in is in
 *    resident mode with the device state newest;
 *  - a potential view (create_potential_view) holds its potential on the
 *    device and answers potential / ghost_velocity from it.
 */
#ifndef GAMERA_CS_MI_DROPIN_H
#define GAMERA_CS_MI_DROPIN_H

#include "gamera_cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* cubedsphere_mi_dipole.h:30-39 */
typedef struct {
  size_t azimuth_points, colatitude_points;
  size_t shell_j_cells, shell_k_cells;
  double maximum_colatitude_rad, epoch_s, sample_radius_re;
  double pedersen_siemens;
  const double *fac_parallel_a_m2[2];
  const double *potential_v[2];
} gamera_cs_mi_analysis_view;

/* cubedsphere_mi_dipole.h:40-41: host copies of the last update's polar
 * FAC and potential, valid until the next update or destroy. */
int gamera_cs_mi_dipole_analysis_view(const gamera_cs_mi_dipole *state,
                                      gamera_cs_mi_analysis_view *view);
/* :44-46: J-parallel, magnetic latitude and longitude of the sample-shell
 * cells at the last update, [face][j][k]. */
int gamera_cs_mi_dipole_export_shell_current(const gamera_cs_mi_dipole *state,
                                             double *jparallel_a_m2, double *mlat_deg,
                                             double *phi_deg, size_t count);
/* :51-57 */
int gamera_cs_mi_dipole_create(const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                               gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                               const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
                               const gamera_cs_earth_startup *startup, const size_t active[3],
                               size_t ghost, const gamera_cs_mi_dipole_config *config,
                               gamera_cs_mi_dipole **result);
/* :58 */
void gamera_cs_mi_dipole_destroy(gamera_cs_mi_dipole *state);
/* :62-65 */
int gamera_cs_mi_dipole_create_potential_view(const gamera_cs_earth_startup *startup,
                                              const gamera_cs_mi_dipole_config *config,
                                              gamera_cs_mi_dipole **result);
/* :66-71 */
int gamera_cs_mi_dipole_export_potential(const gamera_cs_mi_dipole *state, double *values,
                                         size_t count, double *epoch_s,
                                         double *next_update_s);
int gamera_cs_mi_dipole_import_potential(gamera_cs_mi_dipole *state, const double *values,
                                         size_t count, double epoch_s,
                                         double next_update_s);
/* :75-83 */
int gamera_cs_mi_dipole_maybe_update(gamera_cs_mi_dipole *state, double time_code);
int gamera_cs_mi_dipole_update_due(const gamera_cs_mi_dipole *state, double time_code);
int gamera_cs_mi_dipole_ready(const gamera_cs_mi_dipole *state);
int gamera_cs_mi_dipole_potential(const gamera_cs_mi_dipole *state, gamera_no_vec3 point,
                                  double *potential_v);
/* :84-86: the wall EMF of one face's storage (host arrays) at shell wall_i */
int gamera_cs_mi_dipole_apply_inner_emf(const gamera_cs_mi_dipole *state,
                                        gamera_no_storage *storage,
                                        const gamera_no_grid *grid, size_t wall_i);
/* :87-89 */
int gamera_cs_mi_dipole_ghost_velocity(const gamera_cs_mi_dipole *state, gamera_no_vec3 point,
                                       gamera_no_vec3 source_velocity,
                                       gamera_no_vec3 *ghost_velocity);
/* :90-93: CSV files of the reference's format */
int gamera_cs_mi_dipole_write_csv(const gamera_cs_mi_dipole *state, const char *prefix);
int gamera_cs_mi_dipole_write_shell_current_csv(const gamera_cs_mi_dipole *state,
                                                const char *prefix);
/* :94-98 */
int gamera_cs_mi_dipole_stats(const gamera_cs_mi_dipole *state, double *fac_min,
                              double *fac_max, double *potential_min_v,
                              double *potential_max_v, gamera_mi_solver_stats solve[2]);

#ifdef __cplusplus
}
#endif

#endif
