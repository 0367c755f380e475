/* Ionosphere (MI) coupling of the Earth boundary (test infrastructure;
 * see orc.h). */
#include "orc.h"

struct orc_mi {
  int ready;
};

/* cubedsphere_mi_dipole.c:438-440 */
int orc_mi_ready(const orc_mi *mi) { return mi != NULL && mi->ready; }

int orc_mi_ghost_velocity(const orc_mi *mi, vec3 point, vec3 source,
                          vec3 *ghost_velocity) {
  (void)point;
  (void)source;
  (void)ghost_velocity;
  return orc_mi_ready(mi) ? -1 : -1;
}

int orc_mi_inner_emf(const orc_mi *mi, gamera_no_storage *s,
                     const gamera_no_grid *g, size_t wall_i) {
  (void)s;
  (void)g;
  (void)wall_i;
  return orc_mi_ready(mi) ? -1 : -1;
}
