// Copies between device planes and the reference's array-of-structs host
// arrays (gamera_no_grid / gamera_no_storage / gamera_no_background_field,
// nonorthogonal_grid.h:12-30, nonorthogonal_storage.h:12-47): element e of an
// array with extent ext is at host[e * stride], k fastest
// (nonorthogonal_grid.c:7-10), and maps to plane position (i * SJ + j) * SK + k
// of its cube face.
//
// The planes of one array cross PCIe together through a pinned staging
// buffer (one strided copy for all its components), and the conversion to
// or from the interleaved host layout runs on a pool of host threads, each
// taking whole radial planes, so that every host element's components are
// written together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "gcs_device.hpp"
#include "gcs_transfer.hpp"

namespace gcs {

namespace {

// A persistent pool of host threads: run(n, fn) calls fn(begin, end) on
// disjoint ranges covering [0, n) and returns when all are done. Calls are
// serialised.
class HostPool {
 public:
  static HostPool &get() {
    static HostPool pool;
    return pool;
  }

  void run(long n, const std::function<void(long, long)> &fn) {
    std::lock_guard<std::mutex> serial(run_lock_);
    const long parts = std::min<long>(n, (long)workers_.size() + 1);
    if (parts <= 1) {
      if (n > 0) fn(0, n);
      return;
    }
    {
      std::unique_lock<std::mutex> lk(m_);
      fn_ = &fn;
      n_ = n;
      parts_ = parts;
      next_ = 1;  // part 0 runs on the calling thread
      done_ = 0;
      ++generation_;
    }
    cv_.notify_all();
    fn(0, n / parts);
    std::unique_lock<std::mutex> lk(m_);
    done_cv_.wait(lk, [&] { return done_ == parts_ - 1; });
    fn_ = nullptr;
  }

 private:
  HostPool() {
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned count = std::max(1u, std::min(16u, hw == 0 ? 1u : hw)) - 1;
    for (unsigned t = 0; t < count; ++t) workers_.emplace_back([this] { loop(); });
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread &t : workers_) t.join();
  }

  void loop() {
    unsigned long seen = 0;
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || (generation_ != seen && next_ < parts_); });
      if (stop_) return;
      const long part = next_++;
      if (next_ >= parts_) seen = generation_;
      const std::function<void(long, long)> *fn = fn_;
      const long n = n_, parts = parts_;
      lk.unlock();
      (*fn)(n * part / parts, n * (part + 1) / parts);
      lk.lock();
      if (++done_ == parts_ - 1) done_cv_.notify_one();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex run_lock_, m_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(long, long)> *fn_ = nullptr;
  long n_ = 0, parts_ = 0, next_ = 0, done_ = 0;
  unsigned long generation_ = 0;
  bool stop_ = false;
};

// The problem's pinned staging buffer, at least `count` doubles.
double *pinned(Problem &P, size_t count) {
  if (P.pinned_count < count) {
    if (P.pinned != nullptr) (void)hipHostFree(P.pinned);
    P.pinned = nullptr;
    P.pinned_count = 0;
    if (hipHostMalloc((void **)&P.pinned, count * sizeof(double)) != hipSuccess) {
      P.pinned = nullptr;
      set_error("pinned host allocation failed");
      return nullptr;
    }
    P.pinned_count = count;
  }
  return P.pinned;
}

// ncomp consecutive planes of one cube face <-> the pinned buffer
// [ncomp][face_stride].
int planes_down(Problem &P, int first, int ncomp, int face, double *buf) {
  const Layout &L = P.L;
  GCS_HIP(hipMemcpy2D(buf, L.face_stride * sizeof(double),
                      P.plane(first) + face * L.face_stride, L.total * sizeof(double),
                      L.face_stride * sizeof(double), ncomp, hipMemcpyDeviceToHost));
  return 0;
}

int planes_up(Problem &P, int first, int ncomp, int face, const double *buf) {
  const Layout &L = P.L;
  GCS_HIP(hipMemcpy2D(P.plane(first) + face * L.face_stride, L.total * sizeof(double), buf,
                      L.face_stride * sizeof(double), L.face_stride * sizeof(double), ncomp,
                      hipMemcpyHostToDevice));
  return 0;
}

// Component c of element e at host[e * stride + c].
int to_host(Problem &P, int first, int ncomp, int face, const size_t ext[3], double *host,
            size_t stride) {
  const Layout &L = P.L;
  double *buf = pinned(P, (size_t)ncomp * L.face_stride);
  if (buf == nullptr || planes_down(P, first, ncomp, face, buf) != 0) return -1;
  HostPool::get().run((long)ext[0], [&](long i0, long i1) {
    for (size_t i = (size_t)i0; i < (size_t)i1; ++i)
      for (size_t j = 0; j < ext[1]; ++j) {
        const size_t at = (i * L.SJ + j) * L.SK;
        double *h = host + ((i * ext[1] + j) * ext[2]) * stride;
        for (size_t k = 0; k < ext[2]; ++k, h += stride)
          for (int c = 0; c < ncomp; ++c) h[c] = buf[c * L.face_stride + at + k];
      }
  });
  return 0;
}

int from_host(Problem &P, int first, int ncomp, int face, const size_t ext[3],
              const double *host, size_t stride) {
  const Layout &L = P.L;
  double *buf = pinned(P, (size_t)ncomp * L.face_stride);
  if (buf == nullptr) return -1;
  // Positions outside the array's extent are zero, as after allocation.
  HostPool::get().run((long)(L.NI + 1), [&](long i0, long i1) {
    for (int c = 0; c < ncomp; ++c)
      std::memset(buf + c * L.face_stride + (size_t)i0 * L.SJ * L.SK, 0,
                  (size_t)(i1 - i0) * L.SJ * L.SK * sizeof(double));
    for (size_t i = (size_t)i0; i < std::min((size_t)i1, ext[0]); ++i)
      for (size_t j = 0; j < ext[1]; ++j) {
        const size_t at = (i * L.SJ + j) * L.SK;
        const double *h = host + ((i * ext[1] + j) * ext[2]) * stride;
        for (size_t k = 0; k < ext[2]; ++k, h += stride)
          for (int c = 0; c < ncomp; ++c) buf[c * L.face_stride + at + k] = h[c];
      }
  });
  return planes_up(P, first, ncomp, face, buf);
}

}  // namespace

int plane_to_host(Problem &P, int plane, int face, const size_t ext[3],
                  double *host, size_t stride) {
  return to_host(P, plane, 1, face, ext, host, stride);
}

int box_planes_to_host(Problem &P, int first, int ncomp, int face, const size_t ext[3],
                       const size_t lo[3], const size_t hi[3], double *host) {
  const Layout &L = P.L;
  double *buf = pinned(P, (size_t)ncomp * L.face_stride);
  if (buf == nullptr || planes_down(P, first, ncomp, face, buf) != 0) return -1;
  HostPool::get().run((long)(hi[0] - lo[0]), [&](long i0, long i1) {
    for (size_t i = lo[0] + (size_t)i0; i < lo[0] + (size_t)i1; ++i)
      for (size_t j = lo[1]; j < hi[1]; ++j) {
        const size_t at = (i * L.SJ + j) * L.SK;
        for (size_t k = lo[2]; k < hi[2]; ++k)
          for (int c = 0; c < ncomp; ++c)
            host[((i * ext[1] + j) * ext[2] + k) * ncomp + c] = buf[c * L.face_stride + at + k];
      }
  });
  return 0;
}

int host_to_plane(Problem &P, int plane, int face, const size_t ext[3],
                  const double *host, size_t stride) {
  return from_host(P, plane, 1, face, ext, host, stride);
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
  return to_host(P, first, ncomp, face, ext, host, (size_t)ncomp);
}

int host_to_planes(Problem &P, int first, int ncomp, int face,
                   const size_t ext[3], const double *host) {
  return from_host(P, first, ncomp, face, ext, host, (size_t)ncomp);
}

}  // namespace gcs
