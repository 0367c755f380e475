/*
 * The six-face entry points of the reference's MPI step
 * (cubedsphere/cubedsphere_mpi_step.h:31-91) for a cube that fits one GPU.
 *
 * The reference distributes the six faces over the ranks of a
 * communicator, one face per rank, and exchanges donor and interface data
 * between them. On an MI355X the whole cube of the Earth workloads (6 x 192^2
 * x 384 cells, 180 GB) fits one GPU's 288 GB, so these functions step all six
 * faces on the calling process's GPU with the serial drop-in
 * (gamera_cs_dropin.h) and never call MPI: the communicator is accepted and
 * not used. With a one-process communicator (the shim this header provides
 * when mpi.h is not included first, or MPI_COMM_SELF) the results are the
 * serial step's, which the reference's MPI step reproduces exactly (its
 * six-rank against serial gate, cubedsphere/tests/check_mpi_earth_parity.py:2).
 * A caller that does pass a multi-rank communicator gets every face stepped
 * on every rank, a superset of what each rank's storage holds after the
 * reference's distributed step.
 *
 * The tile-decomposed gamera_cs_mpi_advance_tiled
 * (cubedsphere_mpi_step.h:17-29) is not provided.
 *
 * Return convention, argument checks and the device mirror are those of
 * gamera_cs_advance_six_with_background_planned. The halo and edge-sync
 * plans are accepted and not read: the device builds its own from the grid.
 */
#ifndef GAMERA_CS_MPI_DROPIN_H
#define GAMERA_CS_MPI_DROPIN_H

#include "gamera_cs_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#if defined(MPI_VERSION)
typedef MPI_Comm gcs_mpi_comm;
#else
/* One-process shim: any value (NULL) stands for a one-rank communicator. */
typedef void *gcs_mpi_comm;
#endif

/* cubedsphere_mpi_step.h:36-46. */
int gamera_cs_mpi_advance_six_with_background_planned(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    gamera_cs_radial_boundary_function radial_boundary,
    double current_time, double previous_time, void *boundary_context,
    gcs_mpi_comm communicator);

/* cubedsphere_mpi_step.h:48-61 (edge_plan is a gamera_cs_edge_sync_plan). */
int gamera_cs_mpi_advance_six_with_background_planned_cached_edges(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    const void *edge_plan,
    gamera_cs_radial_boundary_function radial_boundary,
    double current_time, double previous_time, void *boundary_context,
    gcs_mpi_comm communicator);

/* cubedsphere_mpi_step.h:63-77. The device mirror keeps every face and
 * seam synchronized after each step, so no precondition applies. */
int gamera_cs_mpi_advance_six_with_background_planned_cached_edges_pre_synced(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    const void *edge_plan,
    gamera_cs_radial_boundary_function radial_boundary,
    double current_time, double previous_time, void *boundary_context,
    gcs_mpi_comm communicator);

/* cubedsphere_mpi_step.h:79-91. */
int gamera_cs_mpi_advance_six_with_background_sparse_pre_synced(
    const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
    gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
    const size_t active[3], size_t ghost,
    double predictor_ratio, double dt,
    const gamera_no_advance_options *options,
    const gamera_no_background_field background[GAMERA_CS_FACE_COUNT],
    const gamera_cs_halo_plan *halo_plan,
    const void *edge_plan,
    gamera_cs_radial_boundary_function radial_boundary,
    double current_time, double previous_time, void *boundary_context,
    gcs_mpi_comm communicator);

#ifdef __cplusplus
}
#endif

#endif
