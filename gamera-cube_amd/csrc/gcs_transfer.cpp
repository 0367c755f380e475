// Copies between device planes and the reference's array-of-structs host
// arrays (gamera_no_grid / gamera_no_storage / gamera_no_background_field,
// nonorthogonal_grid.h:12-30, nonorthogonal_storage.h:12-47): element e of an
// array with extent ext is at host[e * stride], k fastest
// (nonorthogonal_grid.c:7-10), and maps to plane position (i * SJ + j) * SK + k
// of its cube face.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "gcs_device.hpp"
#include "gcs_transfer.hpp"

namespace gcs {

int plane_to_host(Problem &P, int plane, int face, const size_t ext[3],
                  double *host, size_t stride) {
  const Layout &L = P.L;
  std::vector<double> &buf = P.staging;
  buf.resize(L.face_stride);
  GCS_HIP(hipMemcpy(buf.data(), P.plane(plane) + face * L.face_stride,
                    L.face_stride * sizeof(double), hipMemcpyDeviceToHost));
  size_t e = 0;
  for (size_t i = 0; i < ext[0]; ++i)
    for (size_t j = 0; j < ext[1]; ++j) {
      const double *row = buf.data() + (i * L.SJ + j) * L.SK;
      for (size_t k = 0; k < ext[2]; ++k, ++e) host[e * stride] = row[k];
    }
  return 0;
}

int box_planes_to_host(Problem &P, int first, int ncomp, int face, const size_t ext[3],
                       const size_t lo[3], const size_t hi[3], double *host) {
  const Layout &L = P.L;
  std::vector<double> &buf = P.staging;
  buf.resize(L.face_stride);
  for (int c = 0; c < ncomp; ++c) {
    GCS_HIP(hipMemcpy(buf.data(), P.plane(first + c) + face * L.face_stride,
                      L.face_stride * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = lo[0]; i < hi[0]; ++i)
      for (size_t j = lo[1]; j < hi[1]; ++j) {
        const double *row = buf.data() + (i * L.SJ + j) * L.SK;
        for (size_t k = lo[2]; k < hi[2]; ++k)
          host[((i * ext[1] + j) * ext[2] + k) * ncomp + c] = row[k];
      }
  }
  return 0;
}

int host_to_plane(Problem &P, int plane, int face, const size_t ext[3],
                  const double *host, size_t stride) {
  const Layout &L = P.L;
  std::vector<double> &buf = P.staging;
  // Positions outside the array's extent stay zero, as after allocation.
  buf.assign(L.face_stride, 0.0);
  size_t e = 0;
  for (size_t i = 0; i < ext[0]; ++i)
    for (size_t j = 0; j < ext[1]; ++j) {
      double *row = buf.data() + (i * L.SJ + j) * L.SK;
      for (size_t k = 0; k < ext[2]; ++k, ++e) row[k] = host[e * stride];
    }
  GCS_HIP(hipMemcpy(P.plane(plane) + face * L.face_stride, buf.data(),
                    L.face_stride * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int host_to_byte_plane(Problem &P, int plane, int face, const size_t ext[3],
                       const unsigned char *host) {
  const Layout &L = P.L;
  std::vector<uint8_t> buf(L.face_stride, 0);
  size_t e = 0;
  for (size_t i = 0; i < ext[0]; ++i)
    for (size_t j = 0; j < ext[1]; ++j)
      for (size_t k = 0; k < ext[2]; ++k, ++e)
        buf[(i * L.SJ + j) * L.SK + k] = host[e] ? 1 : 0;
  GCS_HIP(hipMemcpy(P.byte_plane(plane) + face * L.face_stride, buf.data(),
                    L.face_stride, hipMemcpyHostToDevice));
  return 0;
}

int planes_to_host(Problem &P, int first, int ncomp, int face,
                   const size_t ext[3], double *host) {
  for (int c = 0; c < ncomp; ++c)
    if (plane_to_host(P, first + c, face, ext, host + c, (size_t)ncomp) != 0)
      return -1;
  return 0;
}

int host_to_planes(Problem &P, int first, int ncomp, int face,
                   const size_t ext[3], const double *host) {
  for (int c = 0; c < ncomp; ++c)
    if (host_to_plane(P, first + c, face, ext, host + c, (size_t)ncomp) != 0)
      return -1;
  return 0;
}

}  // namespace gcs
