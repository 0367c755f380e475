/*
 * The serial Earth driver's CSV snapshot reader and writers
 * (cubedsphere/tests/test_cubedsphere_dipole_background.c:61-193,
 * load_mhd_snapshot, write_earth_snapshot and write_drive_frame), compiled
 * from the driver's
 * source where it lies: the file is included here with its main renamed,
 * so that tests can call these static functions on a harness problem's
 * arrays. Test infrastructure, like ref_harness.c; built into
 * oracle/_ref/libgamera_ref.so only.
 */
#define main refh_dipole_background_main
#include "test_cubedsphere_dipole_background.c"
#undef main

int refh_driver_write_snapshot(const char *path,
                               const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                               const gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                               const gamera_no_background_data owned[GAMERA_CS_FACE_COUNT],
                               const size_t active[3], size_t ghost, double gamma,
                               const gamera_cs_mi_dipole *mi) {
  return write_earth_snapshot(path, grid, storage, owned, active, ghost, gamma, mi);
}

/* write_drive_frame with its index file opened for appending (the driver
 * writes the header when it creates the file, :557-562). */
int refh_driver_write_drive_frame(const char *prefix, long frame, long step, double seconds,
                                  const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                                  const gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                                  const gamera_no_background_data owned[GAMERA_CS_FACE_COUNT],
                                  const size_t active[3], size_t ghost, double gamma,
                                  const gamera_cs_mi_dipole *mi, const char *index_path) {
  FILE *index = fopen(index_path, "a");
  if (index == NULL) return -1;
  int ok = fseek(index, 0, SEEK_END) == 0;
  if (ok && ftell(index) == 0)
    ok = fputs("frame,step,simulation_seconds,mhd_csv\n", index) >= 0;
  ok = ok && write_drive_frame(prefix, frame, step, seconds, grid, storage, owned, active,
                               ghost, gamma, mi, index) == 0;
  return fclose(index) == 0 && ok ? 0 : -1;
}

int refh_driver_load_snapshot(const char *path, const gamera_no_grid grid[GAMERA_CS_FACE_COUNT],
                              gamera_no_storage storage[GAMERA_CS_FACE_COUNT],
                              const gamera_no_background_data owned[GAMERA_CS_FACE_COUNT],
                              const size_t active[3], size_t ghost, double gamma) {
  return load_mhd_snapshot(path, grid, storage, owned, active, ghost, gamma);
}
