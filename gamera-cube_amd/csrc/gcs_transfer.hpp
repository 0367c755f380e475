// Copies between device planes and reference-layout host arrays.
#pragma once

#include <cstddef>

namespace gcs {

struct Problem;

// One plane of one cube face <-> a strided host array of extent ext
// (element e at host[e * stride]).
int plane_to_host(Problem &P, int plane, int face, const size_t ext[3],
                  double *host, size_t stride);
int host_to_plane(Problem &P, int plane, int face, const size_t ext[3],
                  const double *host, size_t stride);
int host_to_byte_plane(Problem &P, int plane, int face, const size_t ext[3],
                       const unsigned char *host);
// ncomp consecutive planes <-> an interleaved host array [element][ncomp].
int planes_to_host(Problem &P, int first, int ncomp, int face,
                   const size_t ext[3], double *host);
int host_to_planes(Problem &P, int first, int ncomp, int face,
                   const size_t ext[3], const double *host);
// planes_to_host for the elements of the box lo <= (i, j, k) < hi only; the
// host array's other elements are left as they are.
int box_planes_to_host(Problem &P, int first, int ncomp, int face, const size_t ext[3],
                       const size_t lo[3], const size_t hi[3], double *host);

}  // namespace gcs
