/* Cube topology, six-face ghost fill and seam reconciliation (test
 * infrastructure; see orc.h). */
#include "orc.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Right-handed face frames {normal, tangent_a, tangent_b} of the six cube
 * faces; the table is the mesh definition (cubedsphere_mesh.c:8-14). */
static const int FRAME[ORC_FACES][3][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
    {{-1, 0, 0}, {0, -1, 0}, {0, 0, 1}},
    {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}},
    {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}},
    {{0, 0, -1}, {1, 0, 0}, {0, -1, 0}}};

static double proj(const double v[3], const int f[3]) {
  return v[0] * (double)f[0] + v[1] * (double)f[1] + v[2] * (double)f[2];
}

/* cubedsphere_mesh.c:58-75: equiangular gnomonic direction. */
static int cube_direction(int face, double a, double b, double out[3]) {
  const double qp = 0.25 * acos(-1.0);
  if (!isfinite(a) || !isfinite(b) || fabs(a) >= 2.0 * qp ||
      fabs(b) >= 2.0 * qp)
    return -1;
  const double u = tan(a), v = tan(b);
  const double il = 1.0 / sqrt(1.0 + u * u + v * v);
  for (int d = 0; d < 3; ++d)
    out[d] = il * ((double)FRAME[face][0][d] + u * (double)FRAME[face][1][d] +
                   v * (double)FRAME[face][2][d]);
  return 0;
}

/* cubedsphere_mesh.c:77-103: owning face and its two angles. */
int orc_cube_inverse(vec3 dir, int *face, double *a, double *b) {
  for (int d = 0; d < 3; ++d)
    if (!isfinite(dir.value[d])) return -1;
  int sel = -1;
  double dom = 0.0;
  for (int f = 0; f < ORC_FACES; ++f) {
    const double p = proj(dir.value, FRAME[f][0]);
    if (p > dom) {
      dom = p;
      sel = f;
    }
  }
  if (sel < 0 || !isfinite(dom)) return -1;
  *face = sel;
  *a = atan2(proj(dir.value, FRAME[sel][1]), dom);
  *b = atan2(proj(dir.value, FRAME[sel][2]), dom);
  return 0;
}

/* cubedsphere_mesh.c:105-162: neighbour across a face edge, found by
 * stepping just outside the edge and mapping back. */
int orc_cube_neighbor(int face, int edge, int *nface, int *nedge,
                      int *reversed) {
  const double qp = 0.25 * acos(-1.0);
  const double eps = 1.0e-7;
  double a = 0.0, b = 0.0;
  const int along_b = edge == GAMERA_CS_A_LOWER || edge == GAMERA_CS_A_UPPER;
  if (along_b)
    a = (edge == GAMERA_CS_A_LOWER ? -1.0 : 1.0) * (qp + eps);
  else
    b = (edge == GAMERA_CS_B_LOWER ? -1.0 : 1.0) * (qp + eps);
  double out[3];
  int target;
  double ta, tb;
  if (cube_direction(face, a, b, out) != 0) return -1;
  vec3 o3 = {{out[0], out[1], out[2]}};
  if (orc_cube_inverse(o3, &target, &ta, &tb) != 0 || target == face)
    return -1;
  const int dedge = fabs(ta) > fabs(tb)
                        ? (ta < 0.0 ? GAMERA_CS_A_LOWER : GAMERA_CS_A_UPPER)
                        : (tb < 0.0 ? GAMERA_CS_B_LOWER : GAMERA_CS_B_UPPER);
  double along[2];
  for (int n = 0; n < 2; ++n) {
    const double pos = n == 0 ? -0.2 : 0.2;
    double sa, sb;
    if (along_b) {
      sa = edge == GAMERA_CS_A_LOWER ? -qp : qp;
      sb = pos;
    } else {
      sa = pos;
      sb = edge == GAMERA_CS_B_LOWER ? -qp : qp;
    }
    if (cube_direction(face, sa, sb, out) != 0) return -1;
    const double p = proj(out, FRAME[target][0]);
    const double da = atan2(proj(out, FRAME[target][1]), p);
    const double db = atan2(proj(out, FRAME[target][2]), p);
    along[n] = dedge < GAMERA_CS_B_LOWER ? db : da;
  }
  *nface = target;
  *nedge = dedge;
  *reversed = along[1] < along[0];
  return 0;
}

/* cubedsphere_halo.c:25-30 */
static size_t stencil_start(double x, size_t n) {
  long s = (long)floor(x) - 3L;
  if (s < 0) s = 0;
  if (s > (long)n - 8L) s = (long)n - 8L;
  return (size_t)s;
}

/* cubedsphere_halo.c:32-57: two-point linear weights inside an 8-wide
 * stencil, continuing the end slope beyond the donor centres. */
static void linear_weights(double x, size_t start, double w[8]) {
  const double l = x - (double)start;
  int base;
  double frac;
  if (l < 0.0) {
    base = 0;
    frac = l;
  } else if (l >= 7.0) {
    base = 6;
    frac = l - 6.0;
  } else {
    base = (int)floor(l);
    frac = l - (double)base;
  }
  for (int p = 0; p < 8; ++p) w[p] = 0.0;
  w[base] = 1.0 - frac;
  w[base + 1] = frac;
}

typedef struct {
  int receiver, donor, direction;
  size_t target, nearest;
  size_t row[8];
  double wx[8], wy[8];
} halo_entry;

struct orc_halo_plan {
  size_t ncell, nface;
  halo_entry *cell, *face;
};

void orc_halo_plan_free(orc_halo_plan *p) {
  if (p == NULL) return;
  free(p->cell);
  free(p->face);
  free(p);
}

/* cubedsphere_halo.c:401-451 */
static int make_entry(halo_entry *e, const gamera_no_grid grid[ORC_FACES],
                      const size_t active[3], size_t ghost, int receiver,
                      int dir, size_t i, size_t j, size_t k, int magnetic) {
  const size_t *ext =
      magnetic ? grid[receiver].face[dir].extent : grid[receiver].cell_extent;
  const size_t t = orc_idx(ext, i, j, k);
  const vec3 point = magnetic ? grid[receiver].face[dir].value[t].centroid
                              : grid[receiver].cell[t].centroid;
  int donor;
  double a, b;
  if (orc_cube_inverse(point, &donor, &a, &b) != 0 ||
      (!magnetic && donor == receiver))
    return -1;
  const double qp = 0.25 * acos(-1.0);
  const double da = 2.0 * qp / (double)active[1];
  const double db = 2.0 * qp / (double)active[2];
  const double x = magnetic ? (a + qp) * (double)active[1] / (2.0 * qp) - 0.5
                            : (a + qp) / da - 0.5;
  const double y = magnetic ? (b + qp) * (double)active[2] / (2.0 * qp) - 0.5
                            : (b + qp) / db - 0.5;
  e->receiver = receiver;
  e->donor = donor;
  e->direction = dir;
  e->target = t;
  const size_t si =
      magnetic ? (i < ghost ? ghost
                  : i >= ghost + active[0] ? ghost + active[0] - 1U : i)
               : i;
  const size_t sx = stencil_start(x, active[1]);
  const size_t sy = stencil_start(y, active[2]);
  for (size_t p = 0; p < 8; ++p)
    e->row[p] = orc_idx(grid[donor].cell_extent, si, ghost + sx + p, ghost + sy);
  linear_weights(x, sx, e->wx);
  linear_weights(y, sy, e->wy);
  e->nearest = 0;
  if (!magnetic) {
    const size_t nx =
        (size_t)fmax(0.0, fmin((double)(active[1] - 1U), nearbyint(x)));
    const size_t ny =
        (size_t)fmax(0.0, fmin((double)(active[2] - 1U), nearbyint(y)));
    e->nearest = orc_idx(grid[donor].cell_extent, i, ghost + nx, ghost + ny);
  }
  return 0;
}

/* cubedsphere_halo.c:453-559 (entry order: receiver, then i, j, k). */
orc_halo_plan *orc_halo_plan_build(const gamera_no_grid grid[ORC_FACES],
                                   const size_t active[3], size_t ghost) {
  orc_halo_plan *p = calloc(1, sizeof(*p));
  if (p == NULL) return NULL;
  const size_t *ce = grid[0].cell_extent;
  const size_t per_cell = active[0] * (ce[1] * ce[2] - active[1] * active[2]);
  size_t per_face = 0;
  for (int d = 0; d < 3; ++d) {
    const size_t *fe = grid[0].face[d].extent;
    per_face += orc_count(fe) - (active[0] + (d == 0)) *
                                    (active[1] + (d == 1)) *
                                    (active[2] + (d == 2));
  }
  p->ncell = ORC_FACES * per_cell;
  p->nface = ORC_FACES * per_face;
  p->cell = calloc(p->ncell, sizeof(halo_entry));
  p->face = calloc(p->nface, sizeof(halo_entry));
  if (p->cell == NULL || p->face == NULL) goto fail;
  size_t nc = 0, nf = 0;
  for (int r = 0; r < ORC_FACES; ++r)
    for (size_t i = ghost; i < ghost + active[0]; ++i)
      for (size_t j = 0; j < ce[1]; ++j)
        for (size_t k = 0; k < ce[2]; ++k) {
          if (j >= ghost && j < ghost + active[1] && k >= ghost &&
              k < ghost + active[2])
            continue;
          if (make_entry(&p->cell[nc++], grid, active, ghost, r, 0, i, j, k,
                         0) != 0)
            goto fail;
        }
  for (int r = 0; r < ORC_FACES; ++r)
    for (int d = 0; d < 3; ++d) {
      const size_t *fe = grid[r].face[d].extent;
      for (size_t i = 0; i < fe[0]; ++i)
        for (size_t j = 0; j < fe[1]; ++j)
          for (size_t k = 0; k < fe[2]; ++k) {
            const size_t c3[3] = {i, j, k};
            int physical = 1;
            for (int a = 0; a < 3; ++a) {
              const size_t up = ghost + active[a] + (d == a ? 1U : 0U);
              if (c3[a] < ghost || c3[a] >= up) physical = 0;
            }
            if (physical) continue;
            if (make_entry(&p->face[nf++], grid, active, ghost, r, d, i, j, k,
                           1) != 0)
              goto fail;
          }
    }
  if (nc != p->ncell || nf != p->nface) goto fail;
  return p;
fail:
  orc_halo_plan_free(p);
  return NULL;
}

/* cubedsphere_halo.c:103-115 */
static int halo_valid(const double u[ORC_NQ]) {
  const double rho = u[0];
  if (!isfinite(rho) || rho <= 0.0) return 0;
  double m2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!isfinite(u[1 + d])) return 0;
    m2 += u[1 + d] * u[1 + d];
  }
  return isfinite(u[4]) && u[4] > 0.5 * m2 / rho;
}

/* cubedsphere_halo.c:582-632: all 64 stencil terms, p-major q-minor. */
int orc_fill_cell_halos(const orc_halo_plan *plan, gamera_no_storage *st) {
  if (plan == NULL) return -1;
#pragma omp parallel for schedule(static)
  for (size_t n = 0; n < plan->ncell; ++n) {
    const halo_entry *e = &plan->cell[n];
    const gamera_no_storage *src = &st[e->donor];
    double cur[ORC_NQ] = {0.0}, old[ORC_NQ] = {0.0};
    double cb[3] = {0.0, 0.0, 0.0}, ob[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < 8; ++p)
      for (int q = 0; q < 8; ++q) {
        const size_t s = e->row[p] + (size_t)q;
        const double w = e->wx[p] * e->wy[q];
        for (int m = 0; m < ORC_NQ; ++m) {
          cur[m] += w * src->conserved[s * ORC_NQ + m];
          old[m] += w * src->old_conserved[s * ORC_NQ + m];
        }
        for (int d = 0; d < 3; ++d) {
          cb[d] += w * src->cell_magnetic[s].value[d];
          ob[d] += w * src->old_cell_magnetic[s].value[d];
        }
      }
    if (!halo_valid(cur) || !halo_valid(old)) {
      memcpy(cur, &src->conserved[e->nearest * ORC_NQ], sizeof(cur));
      memcpy(old, &src->old_conserved[e->nearest * ORC_NQ], sizeof(old));
      memcpy(cb, src->cell_magnetic[e->nearest].value, sizeof(cb));
      memcpy(ob, src->old_cell_magnetic[e->nearest].value, sizeof(ob));
    }
    gamera_no_storage *dst = &st[e->receiver];
    memcpy(&dst->conserved[e->target * ORC_NQ], cur, sizeof(cur));
    memcpy(&dst->old_conserved[e->target * ORC_NQ], old, sizeof(old));
    memcpy(dst->cell_magnetic[e->target].value, cb, sizeof(cb));
    memcpy(dst->old_cell_magnetic[e->target].value, ob, sizeof(ob));
  }
  return 0;
}

/* cubedsphere_halo.c:651-686: ghost face flux = interpolated B . area. */
int orc_fill_face_halos(const orc_halo_plan *plan,
                        const gamera_no_grid grid[ORC_FACES],
                        gamera_no_storage *st) {
  if (plan == NULL) return -1;
#pragma omp parallel for schedule(static)
  for (size_t n = 0; n < plan->nface; ++n) {
    const halo_entry *e = &plan->face[n];
    const gamera_no_storage *src = &st[e->donor];
    double cb[3] = {0.0, 0.0, 0.0}, ob[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < 8; ++p)
      for (int q = 0; q < 8; ++q) {
        const size_t s = e->row[p] + (size_t)q;
        const double w = e->wx[p] * e->wy[q];
        for (int d = 0; d < 3; ++d) {
          cb[d] += w * src->cell_magnetic[s].value[d];
          ob[d] += w * src->old_cell_magnetic[s].value[d];
        }
      }
    const double *av =
        grid[e->receiver].face[e->direction].value[e->target].area_vector.value;
    double cf = 0.0, of = 0.0;
    for (int d = 0; d < 3; ++d) {
      cf += cb[d] * av[d];
      of += ob[d] * av[d];
    }
    st[e->receiver].face_flux[e->direction][e->target] = cf;
    st[e->receiver].old_face_flux[e->direction][e->target] = of;
  }
  return 0;
}

/* cubedsphere_halo.c:217-245: copy the boundary shell into radial ghosts. */
int orc_radial_outflow(gamera_no_storage *st, const size_t active[3],
                       size_t ghost) {
  for (int f = 0; f < ORC_FACES; ++f) {
    gamera_no_storage *s = &st[f];
    const size_t *e = s->cell_extent;
    for (size_t i = 0; i < e[0]; ++i) {
      if (i >= ghost && i < ghost + active[0]) continue;
      const size_t di = i < ghost ? ghost : ghost + active[0] - 1U;
      for (size_t j = 0; j < e[1]; ++j)
        for (size_t k = 0; k < e[2]; ++k) {
          const size_t t = orc_idx(e, i, j, k), d = orc_idx(e, di, j, k);
          memcpy(&s->conserved[t * ORC_NQ], &s->conserved[d * ORC_NQ],
                 ORC_NQ * sizeof(double));
          memcpy(&s->old_conserved[t * ORC_NQ], &s->old_conserved[d * ORC_NQ],
                 ORC_NQ * sizeof(double));
          s->cell_magnetic[t] = s->cell_magnetic[d];
          s->old_cell_magnetic[t] = s->old_cell_magnetic[d];
        }
    }
  }
  return 0;
}

/* cubedsphere_interface.c:7-81: average each shared seam face once. */
int orc_sync_faces(gamera_no_storage *st, const size_t active[3], size_t ghost,
                   int magnetic) {
  for (int f = 0; f < ORC_FACES; ++f)
    for (int edge = 0; edge < GAMERA_CS_EDGE_COUNT; ++edge) {
      int nf, ne, rev;
      if (orc_cube_neighbor(f, edge, &nf, &ne, &rev) != 0) return -1;
      if (f >= nf) continue;
      const int sa = edge < GAMERA_CS_B_LOWER ? 1 : 2;
      const int da = ne < GAMERA_CS_B_LOWER ? 1 : 2;
      const size_t sb = edge % 2 == 0 ? ghost : ghost + active[sa];
      const size_t db = ne % 2 == 0 ? ghost : ghost + active[da];
      const double sign = edge % 2 == ne % 2 ? -1.0 : 1.0;
      for (size_t i = ghost; i < ghost + active[0]; ++i)
        for (size_t along = 0; along < active[1]; ++along) {
          const size_t dal = rev ? active[1] - 1U - along : along;
          size_t s3[3] = {i, ghost + along, ghost + along};
          size_t t3[3] = {i, ghost + dal, ghost + dal};
          s3[sa] = sb;
          t3[da] = db;
          const size_t s = orc_idx(st[f].face_extent[sa], s3[0], s3[1], s3[2]);
          const size_t t = orc_idx(st[nf].face_extent[da], t3[0], t3[1], t3[2]);
          if (magnetic) {
            double *a0 = &st[f].face_flux[sa][s], *b0 = &st[nf].face_flux[da][t];
            const double c = 0.5 * (*a0 + sign * *b0);
            *a0 = c;
            *b0 = sign * c;
            double *a1 = &st[f].old_face_flux[sa][s],
                   *b1 = &st[nf].old_face_flux[da][t];
            const double oc = 0.5 * (*a1 + sign * *b1);
            *a1 = oc;
            *b1 = sign * oc;
          } else {
            for (int m = 0; m < ORC_NQ; ++m) {
              double *a = &st[f].fluid_face_flux[sa][s * ORC_NQ + m];
              double *b = &st[nf].fluid_face_flux[da][t * ORC_NQ + m];
              const double c = 0.5 * (*a + sign * *b);
              *a = c;
              *b = sign * c;
            }
            for (int d = 0; d < 3; ++d) {
              double *a = &st[f].maxwell_face_flux[sa][s].value[d];
              double *b = &st[nf].maxwell_face_flux[da][t].value[d];
              const double c = 0.5 * (*a + sign * *b);
              *a = c;
              *b = sign * c;
            }
          }
        }
    }
  return 0;
}

/* cubedsphere_mesh.c:36-56: integer identity of a seam vertex. */
static void vertex_key(int face, size_t n, const size_t v[3], int64_t key[4]) {
  const int64_t nn = (int64_t)n;
  const int64_t a = 2 * (int64_t)v[1] - nn;
  const int64_t b = 2 * (int64_t)v[2] - nn;
  key[0] = (int64_t)v[0];
  for (int d = 0; d < 3; ++d)
    key[d + 1] = (int64_t)FRAME[face][0][d] * nn +
                 (int64_t)FRAME[face][1][d] * a +
                 (int64_t)FRAME[face][2][d] * b;
}

typedef struct {
  int64_t key[8];
  size_t order;   /* insertion position: makes the sort a stable sort */
  int face, direction, sign;
  size_t index;
} seam_edge;

struct orc_edge_plan {
  size_t count;
  seam_edge *e;
};

static int key_cmp(const int64_t *a, const int64_t *b, int n) {
  for (int d = 0; d < n; ++d) {
    if (a[d] < b[d]) return -1;
    if (a[d] > b[d]) return 1;
  }
  return 0;
}

/* Ties keep insertion order, as the reference's glibc qsort (a merge sort
 * for this size) does; cubedsphere_interface.c:195 sorts by the 8 keys. */
static int edge_cmp(const void *l, const void *r) {
  const seam_edge *a = l, *b = r;
  const int c = key_cmp(a->key, b->key, 8);
  if (c != 0) return c;
  return a->order < b->order ? -1 : a->order > b->order ? 1 : 0;
}

void orc_edge_plan_free(orc_edge_plan *p) {
  if (p == NULL) return;
  free(p->e);
  free(p);
}

/* cubedsphere_interface.c:126-209 */
orc_edge_plan *orc_edge_plan_build(const gamera_no_grid grid[ORC_FACES],
                                   const size_t active[3], size_t ghost) {
  const size_t n = active[1];
  const size_t cap = 6U * (active[0] * (4U * n + 4U) + (active[0] + 1U) * 4U * n);
  orc_edge_plan *p = calloc(1, sizeof(*p));
  if (p == NULL) return NULL;
  p->e = malloc(cap * sizeof(seam_edge));
  if (p->e == NULL) goto fail;
  size_t count = 0;
  for (int f = 0; f < ORC_FACES; ++f)
    for (int d = 0; d < 3; ++d) {
      const size_t iu = ghost + active[0] + (d != 0);
      const size_t ju = ghost + active[1] + (d != 1);
      const size_t ku = ghost + active[2] + (d != 2);
      for (size_t i = ghost; i < iu; ++i)
        for (size_t j = ghost; j < ju; ++j)
          for (size_t k = ghost; k < ku; ++k) {
            const int on_a = j == ghost || j == ghost + n;
            const int on_b = k == ghost || k == ghost + n;
            const int shared = d == 0 ? (on_a || on_b) : d == 1 ? on_b : on_a;
            if (!shared) continue;
            if (count >= cap) goto fail;
            seam_edge *e = &p->e[count];
            size_t v0[3] = {i - ghost, j - ghost, k - ghost};
            size_t v1[3] = {v0[0], v0[1], v0[2]};
            ++v1[d];
            int64_t k0[4], k1[4];
            vertex_key(f, n, v0, k0);
            vertex_key(f, n, v1, k1);
            const int o = key_cmp(k0, k1, 4);
            if (o == 0) goto fail;
            e->sign = o < 0 ? 1 : -1;
            for (int a = 0; a < 4; ++a) {
              e->key[a] = o < 0 ? k0[a] : k1[a];
              e->key[a + 4] = o < 0 ? k1[a] : k0[a];
            }
            e->face = f;
            e->direction = d;
            e->index = orc_idx(grid[f].edge[d].extent, i, j, k);
            e->order = count;
            ++count;
          }
    }
  qsort(p->e, count, sizeof(seam_edge), edge_cmp);
  p->count = count;
  for (size_t a = 0; a < count;) {
    size_t b = a + 1;
    while (b < count && key_cmp(p->e[a].key, p->e[b].key, 8) == 0) ++b;
    if (b - a < 2 || b - a > 3) goto fail;
    a = b;
  }
  return p;
fail:
  orc_edge_plan_free(p);
  return NULL;
}

/* cubedsphere_interface.c:217-242: one signed mean per coincident edge. */
int orc_sync_edges(const orc_edge_plan *p, gamera_no_storage *st) {
  if (p == NULL) return -1;
  for (size_t a = 0; a < p->count;) {
    size_t b = a + 1;
    while (b < p->count && key_cmp(p->e[a].key, p->e[b].key, 8) == 0) ++b;
    double c = 0.0;
    for (size_t t = a; t < b; ++t)
      c += (double)p->e[t].sign *
           st[p->e[t].face].edge_emf[p->e[t].direction][p->e[t].index];
    c /= (double)(b - a);
    for (size_t t = a; t < b; ++t)
      st[p->e[t].face].edge_emf[p->e[t].direction][p->e[t].index] =
          (double)p->e[t].sign * c;
    a = b;
  }
  return 0;
}
