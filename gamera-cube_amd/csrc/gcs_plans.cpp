// Host-side construction of the immutable six-face plans: tangential ghost
// cell and ghost face donor blocks, seam face pairs, coincident seam-edge
// groups and the list of cells that touch a seam face.
//
// The plans depend on cell/face centroids only. They are built here with the
// C library's atan2 so that donor faces and weights are exactly the
// reference's (cubedsphere/cubedsphere_halo.c:401-559,
// cubedsphere/cubedsphere_interface.c:7-209).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

#include "gcs_device.hpp"

namespace gcs {

namespace {

// Cube-face frames {normal, tangent_a, tangent_b} (cubedsphere_mesh.c:8-14).
const int kFrame[6][3][3] = {
    {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},   {{-1, 0, 0}, {0, -1, 0}, {0, 0, 1}},
    {{0, 1, 0}, {-1, 0, 0}, {0, 0, 1}},  {{0, -1, 0}, {1, 0, 0}, {0, 0, 1}},
    {{0, 0, 1}, {1, 0, 0}, {0, 1, 0}},   {{0, 0, -1}, {1, 0, 0}, {0, -1, 0}}};

double project(const double v[3], const int f[3]) {
  return v[0] * (double)f[0] + v[1] * (double)f[1] + v[2] * (double)f[2];
}

}  // namespace

// cubedsphere_mesh.c:77-103
bool cube_inverse(const double x[3], int &face, double &a, double &b) {
  for (int d = 0; d < 3; ++d)
    if (!std::isfinite(x[d])) return false;
  int sel = -1;
  double dom = 0.0;
  for (int f = 0; f < 6; ++f) {
    const double p = project(x, kFrame[f][0]);
    if (p > dom) {
      dom = p;
      sel = f;
    }
  }
  if (sel < 0 || !std::isfinite(dom)) return false;
  face = sel;
  a = atan2(project(x, kFrame[sel][1]), dom);
  b = atan2(project(x, kFrame[sel][2]), dom);
  return true;
}

namespace {


// cubedsphere_mesh.c:58-75
bool cube_direction(int face, double a, double b, double out[3]) {
  const double qp = 0.25 * acos(-1.0);
  if (!std::isfinite(a) || !std::isfinite(b) || fabs(a) >= 2.0 * qp ||
      fabs(b) >= 2.0 * qp)
    return false;
  const double u = tan(a), v = tan(b);
  const double il = 1.0 / sqrt(1.0 + u * u + v * v);
  for (int d = 0; d < 3; ++d)
    out[d] = il * ((double)kFrame[face][0][d] + u * (double)kFrame[face][1][d] +
                   v * (double)kFrame[face][2][d]);
  return true;
}

// cubedsphere_mesh.c:105-162
bool cube_neighbor(int face, int edge, int &nface, int &nedge, bool &reversed) {
  const double qp = 0.25 * acos(-1.0);
  const double eps = 1.0e-7;
  const bool a_edge = edge == GAMERA_CS_A_LOWER || edge == GAMERA_CS_A_UPPER;
  double a = 0.0, b = 0.0;
  if (a_edge) a = (edge == GAMERA_CS_A_LOWER ? -1.0 : 1.0) * (qp + eps);
  else b = (edge == GAMERA_CS_B_LOWER ? -1.0 : 1.0) * (qp + eps);
  double out[3], ta, tb;
  int target;
  if (!cube_direction(face, a, b, out) || !cube_inverse(out, target, ta, tb) ||
      target == face)
    return false;
  const int dedge = fabs(ta) > fabs(tb)
                        ? (ta < 0.0 ? GAMERA_CS_A_LOWER : GAMERA_CS_A_UPPER)
                        : (tb < 0.0 ? GAMERA_CS_B_LOWER : GAMERA_CS_B_UPPER);
  double along[2];
  for (int n = 0; n < 2; ++n) {
    const double pos = n == 0 ? -0.2 : 0.2;
    const double sa = a_edge ? (edge == GAMERA_CS_A_LOWER ? -qp : qp) : pos;
    const double sb = a_edge ? pos : (edge == GAMERA_CS_B_LOWER ? -qp : qp);
    if (!cube_direction(face, sa, sb, out)) return false;
    const double p = project(out, kFrame[target][0]);
    const double da = atan2(project(out, kFrame[target][1]), p);
    const double db = atan2(project(out, kFrame[target][2]), p);
    along[n] = dedge < GAMERA_CS_B_LOWER ? db : da;
  }
  nface = target;
  nedge = dedge;
  reversed = along[1] < along[0];
  return true;
}

// cubedsphere_halo.c:25-30
int stencil_start(double x, int n) {
  long s = (long)floor(x) - 3L;
  if (s < 0) s = 0;
  if (s > (long)n - 8L) s = (long)n - 8L;
  return (int)s;
}

// cubedsphere_halo.c:32-57: the nonzero pair (base, base + 1) of the
// 8-point linear weights and their values.
void linear_pair(double x, int start, int &base, double w[2]) {
  const double l = x - (double)start;
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
  w[0] = 1.0 - frac;
  w[1] = frac;
}

struct Donor {
  int face;
  double x, y;
};

bool donor_of(const double point[3], int n, bool magnetic, Donor &out) {
  double a, b;
  if (!cube_inverse(point, out.face, a, b)) return false;
  const double qp = 0.25 * acos(-1.0);
  const double da = 2.0 * qp / (double)n;
  // Both operation orders of cubedsphere_halo.c:417-426 are kept.
  out.x = magnetic ? (a + qp) * (double)n / (2.0 * qp) - 0.5 : (a + qp) / da - 0.5;
  out.y = magnetic ? (b + qp) * (double)n / (2.0 * qp) - 0.5 : (b + qp) / da - 0.5;
  return true;
}

void block(const Layout &L, const Donor &d, int source_i, int &corner,
           double w[4]) {
  const int sx = stencil_start(d.x, L.n), sy = stencil_start(d.y, L.n);
  int bx, by;
  double wx[2], wy[2];
  linear_pair(d.x, sx, bx, wx);
  linear_pair(d.y, sy, by, wy);
  corner = (int)L.at(d.face, source_i, L.ghost + sx + bx, L.ghost + sy + by);
  w[0] = wx[0] * wy[0];
  w[1] = wx[0] * wy[1];
  w[2] = wx[1] * wy[0];
  w[3] = wx[1] * wy[1];
}

template <typename T>
int upload(const std::vector<T> &v, T **out) {
  *out = nullptr;
  if (v.empty()) return 0;
  GCS_HIP(hipMalloc(out, v.size() * sizeof(T)));
  GCS_HIP(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

struct SeamEdge {
  std::array<int64_t, 8> key;
  int order;
  int index, dir;
  double sign;
};

}  // namespace

int build_plans(Problem &P) {
  const Layout &L = P.L;
  const int g = L.ghost;
  std::vector<double> cen(3 * L.total), fcen(9 * L.total);
  for (int e = 0; e < 3; ++e)
    GCS_HIP(hipMemcpy(cen.data() + e * L.total, P.plane(P_CEN + e),
                      L.total * sizeof(double), hipMemcpyDeviceToHost));
  for (int d = 0; d < 3; ++d)
    for (int e = 0; e < 3; ++e)
      GCS_HIP(hipMemcpy(fcen.data() + (3 * d + e) * L.total,
                        P.plane(P_FACE + FACE_PLANES * d + FO_CEN + e),
                        L.total * sizeof(double), hipMemcpyDeviceToHost));
  // Tangential ghost cells of the active radial range
  // (cubedsphere_halo.c:502-514).
  std::vector<HaloCell> cells;
  for (int f = 0; f < 6; ++f)
    for (int i = g; i < g + L.nr; ++i)
      for (int j = 0; j < L.NJ; ++j)
        for (int k = 0; k < L.NK; ++k) {
          if (j >= g && j < g + L.n && k >= g && k < g + L.n) continue;
          const long at = L.at(f, i, j, k);
          const double x[3] = {cen[at], cen[L.total + at], cen[2 * L.total + at]};
          Donor d;
          if (!donor_of(x, L.n, false, d) || d.face == f) {
            set_error("cell halo donor search failed");
            return -1;
          }
          HaloCell h{};
          h.target = (int)at;
          block(L, d, i, h.donor, h.w);
          const int nx = (int)fmax(0.0, fmin((double)(L.n - 1), nearbyint(d.x)));
          const int ny = (int)fmax(0.0, fmin((double)(L.n - 1), nearbyint(d.y)));
          h.nearest = (int)L.at(d.face, i, g + nx, g + ny);
          cells.push_back(h);
        }
  // Non-physical magnetic faces (cubedsphere_halo.c:515-535). With the Earth
  // boundary the faces it rewrites afterwards are left out
  // (cubedsphere_earth_boundary.c:189-193).
  std::vector<HaloFace> faces;
  const int outer = g + L.nr;
  for (int f = 0; f < 6; ++f)
    for (int dir = 0; dir < 3; ++dir)
      for (int i = 0; i < L.NI + (dir == 0); ++i)
        for (int j = 0; j < L.NJ + (dir == 1); ++j)
          for (int k = 0; k < L.NK + (dir == 2); ++k) {
            const int c3[3] = {i, j, k};
            const int act[3] = {L.nr, L.n, L.n};
            bool physical = true;
            for (int a = 0; a < 3; ++a)
              if (c3[a] < g || c3[a] >= g + act[a] + (dir == a)) physical = false;
            if (physical) continue;
            if (P.bnd.earth_cells && (i < g || i >= outer + (dir == 0))) continue;
            const long at = L.at(f, i, j, k);
            const double x[3] = {fcen[(3 * dir) * L.total + at],
                                 fcen[(3 * dir + 1) * L.total + at],
                                 fcen[(3 * dir + 2) * L.total + at]};
            Donor d;
            if (!donor_of(x, L.n, true, d)) {
              set_error("face halo donor search failed");
              return -1;
            }
            HaloFace h{};
            h.target = (int)at;
            h.dir = dir;
            const int si = i < g ? g : i >= outer ? outer - 1 : i;
            block(L, d, si, h.donor, h.w);
            faces.push_back(h);
          }
  // Seam face pairs (cubedsphere_interface.c:7-81).
  std::vector<SeamPair> seams;
  for (int f = 0; f < 6; ++f)
    for (int edge = 0; edge < 4; ++edge) {
      int nf, ne;
      bool rev;
      if (!cube_neighbor(f, edge, nf, ne, rev)) {
        set_error("cube neighbour failed");
        return -1;
      }
      if (f >= nf) continue;
      const int sa = edge < GAMERA_CS_B_LOWER ? 1 : 2;
      const int da = ne < GAMERA_CS_B_LOWER ? 1 : 2;
      const int sb = edge % 2 == 0 ? g : g + L.n;
      const int db = ne % 2 == 0 ? g : g + L.n;
      const double sign = edge % 2 == ne % 2 ? -1.0 : 1.0;
      for (int i = g; i < g + L.nr; ++i)
        for (int along = 0; along < L.n; ++along) {
          const int dal = rev ? L.n - 1 - along : along;
          int s3[3] = {i, g + along, g + along};
          int t3[3] = {i, g + dal, g + dal};
          s3[sa] = sb;
          t3[da] = db;
          seams.push_back({(int)L.at(f, s3[0], s3[1], s3[2]),
                           (int)L.at(nf, t3[0], t3[1], t3[2]), sa, da, sign});
        }
    }
  // Coincident seam edges (cubedsphere_interface.c:126-209), grouped by
  // their integer endpoint keys (cubedsphere_mesh.c:36-56); equal keys keep
  // insertion order.
  std::vector<SeamEdge> edges;
  const int64_t nn = L.n;
  auto key = [&](int f, const int v[3], int64_t out[4]) {
    const int64_t a = 2 * (int64_t)v[1] - nn, b = 2 * (int64_t)v[2] - nn;
    out[0] = v[0];
    for (int d = 0; d < 3; ++d)
      out[d + 1] = (int64_t)kFrame[f][0][d] * nn + (int64_t)kFrame[f][1][d] * a +
                   (int64_t)kFrame[f][2][d] * b;
  };
  for (int f = 0; f < 6; ++f)
    for (int d = 0; d < 3; ++d)
      for (int i = g; i < g + L.nr + (d != 0); ++i)
        for (int j = g; j < g + L.n + (d != 1); ++j)
          for (int k = g; k < g + L.n + (d != 2); ++k) {
            const bool on_a = j == g || j == g + L.n;
            const bool on_b = k == g || k == g + L.n;
            const bool shared = d == 0 ? (on_a || on_b) : d == 1 ? on_b : on_a;
            if (!shared) continue;
            int v0[3] = {i - g, j - g, k - g};
            int v1[3] = {v0[0], v0[1], v0[2]};
            ++v1[d];
            int64_t k0[4], k1[4];
            key(f, v0, k0);
            key(f, v1, k1);
            const bool less = std::lexicographical_compare(k0, k0 + 4, k1, k1 + 4);
            const bool equal = std::equal(k0, k0 + 4, k1);
            if (equal) {
              set_error("degenerate seam edge");
              return -1;
            }
            SeamEdge s;
            for (int a = 0; a < 4; ++a) {
              s.key[a] = less ? k0[a] : k1[a];
              s.key[a + 4] = less ? k1[a] : k0[a];
            }
            s.sign = less ? 1.0 : -1.0;
            s.index = (int)L.at(f, i, j, k);
            s.dir = d;
            s.order = (int)edges.size();
            edges.push_back(s);
          }
  std::stable_sort(edges.begin(), edges.end(),
                   [](const SeamEdge &a, const SeamEdge &b) { return a.key < b.key; });
  std::vector<int> groups{0};
  std::vector<EdgeMember> members;
  for (size_t a = 0; a < edges.size();) {
    size_t b = a + 1;
    while (b < edges.size() && edges[b].key == edges[a].key) ++b;
    if (b - a < 2 || b - a > 3) {
      set_error("seam edge group is not a pair or a cube corner");
      return -1;
    }
    for (size_t t = a; t < b; ++t)
      members.push_back({edges[t].index, edges[t].dir, edges[t].sign});
    groups.push_back((int)members.size());
    a = b;
  }
  // Active cells with a seam face.
  std::vector<int> seam_cells;
  for (int f = 0; f < 6; ++f)
    for (int i = g; i < g + L.nr; ++i)
      for (int j = g; j < g + L.n; ++j)
        for (int k = g; k < g + L.n; ++k)
          if (j == g || j == g + L.n - 1 || k == g || k == g + L.n - 1)
            seam_cells.push_back((int)L.at(f, i, j, k));
  if (upload(cells, &P.halo_cells) || upload(faces, &P.halo_faces) ||
      upload(seams, &P.seams) || upload(members, &P.edge_members) ||
      upload(groups, &P.edge_groups) || upload(seam_cells, &P.seam_cells))
    return -1;
  P.n_halo_cells = (int)cells.size();
  P.n_halo_faces = (int)faces.size();
  P.n_seams = (int)seams.size();
  P.n_edge_groups = (int)groups.size() - 1;
  P.n_edge_members = (int)members.size();
  P.n_seam_cells = (int)seam_cells.size();
  return gather_seam_geometry(P);
}

}  // namespace gcs
