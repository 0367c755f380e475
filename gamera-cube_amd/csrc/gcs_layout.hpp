// Device-resident layout of the six-face state.
//
// Every quantity is a structure-of-arrays plane of doubles covering all six
// cube faces. Cells, faces and edges share one padded index space per cube
// face: element (i, j, k) of face f lives at
//     f * face_stride + (i * SJ + j) * SK + k,   SI = NI + 1, SJ = NJ + 1,
// with SK = NK + 1 rounded up to a multiple of 8 so every k-row starts on a
// 64-byte boundary. NI x NJ x NK is the cell extent (active + 2 ghosts);
// the +1 in each axis holds the staggered face/edge extents, so a face or
// edge in any direction indexes exactly like the cell at the same (i, j, k)
// (nonorthogonal_grid.c:7-10 gives the reference's per-array k-fastest
// order, which this layout keeps within each row).
#pragma once

#include <cstddef>
#include <cstdint>

namespace gcs {

struct Layout {
  int NI, NJ, NK;   // cell extent per cube face
  int SJ, SK;       // row strides (SI = NI + 1 is implicit)
  int ghost;
  int nr, n;        // active radial / angular cells
  long face_stride; // (NI + 1) * SJ * SK
  long total;       // 6 * face_stride

  __host__ __device__ long at(int f, int i, int j, int k) const {
    return f * face_stride + ((long)i * SJ + j) * SK + k;
  }
  __host__ __device__ long di() const { return (long)SJ * SK; }
};

// Offset, in doubles, of every FP64 plane from the start of its allocation:
// rows are a multiple of 16 doubles (SK) apart and planes a multiple of
// L.total, so with this offset the first active cell of every row (k =
// ghost) starts a 128-byte cache line, and a wave's run of active cells
// along k touches no partial line at its start.
inline int plane_shift(const Layout &L) { return (16 - L.ghost % 16) % 16; }
// The same for the FP32 planes (32 floats a line): rows 7 lines of 16
// doubles apart are 3.5 lines of floats apart, so the first active cell
// starts a line in every other row.
inline int plane_shift32(const Layout &L) { return (32 - L.ghost % 32) % 32; }

inline Layout make_layout(int nr, int n, int ghost) {
  Layout L;
  L.nr = nr;
  L.n = n;
  L.ghost = ghost;
  L.NI = nr + 2 * ghost;
  L.NJ = L.NK = n + 2 * ghost;
  L.SJ = L.NJ + 1;
  L.SK = ((L.NK + 1) + 7) / 8 * 8;
  L.face_stride = (long)(L.NI + 1) * L.SJ * L.SK;
  L.total = 6 * L.face_stride;
  return L;
}

// Extents of the reference arrays of one cube face (nonorthogonal_grid.c:
// 212-245): faces add one along their direction, edges along the other two.
inline void face_extent(const Layout &L, int d, size_t e[3]) {
  e[0] = L.NI + (d == 0);
  e[1] = L.NJ + (d == 1);
  e[2] = L.NK + (d == 2);
}
inline void edge_extent(const Layout &L, int d, size_t e[3]) {
  e[0] = L.NI + (d != 0);
  e[1] = L.NJ + (d != 1);
  e[2] = L.NK + (d != 2);
}

}  // namespace gcs
