// The drivers' CSV output of a device-resident problem: the six-face state
// snapshot and the output-frame cadence of the serial Earth driver
// (cubedsphere/tests/test_cubedsphere_dipole_background.c:120-193), and the
// ionosphere's CSV files (cubedsphere/cubedsphere_mi_dipole.c:620-690),
// which the snapshot appends. The values are computed on the device; the
// host only formats them (%.17g, as the reference prints them), with the
// rows of a face split over threads and written in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gcs_device.hpp"

namespace gcs {

int snapshot_face(Problem &P, int face, double gamma, std::vector<double> &rows,
                  long *first_bad);
int load_snapshot_rows(Problem &P, const std::vector<double> &rows, const std::vector<int> &idx,
                       double gamma, long *first_bad);

namespace {

// Rows [begin, end) of one face as the reference's fprintf formats them.
void format_rows(const std::vector<double> &rows, int face, int n, long begin, long end,
                 std::string &out) {
  char line[512];
  for (long t = begin; t < end; ++t) {
    const double *r = &rows[11 * (size_t)t];
    const size_t k = (size_t)(t % n), j = (size_t)((t / n) % n), i = (size_t)(t / ((long)n * n));
    const int len = std::snprintf(line, sizeof(line),
                                  "%d,%zu,%zu,%zu,%.17g,%.17g,%.17g,%.17g,%.17g,"
                                  "%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n",
                                  face, i, j, k, r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7],
                                  r[8], r[9], r[10]);
    out.append(line, (size_t)len);
  }
}

}  // namespace

// cubedsphere_mi_dipole.c:620-648
int write_mi_csv(Problem &P, const char *prefix) {
  MiInfo info;
  mi_info(P, info);
  if (!info.ready || prefix == nullptr || std::strlen(prefix) > 4000U) return -1;
  std::vector<double> fac, pot;
  if (mi_download(P, 1, fac) != 0 || mi_download(P, 0, pot) != 0) return -1;
  const size_t np = P.cfg.mi.azimuth_points, nt = P.cfg.mi.colatitude_points;
  for (int hemi = 0; hemi < 2; ++hemi) {
    const char *name = hemi == 0 ? "north" : "south";
    char path[4096];
    if (std::snprintf(path, sizeof(path), "%s.mi_%s.csv", prefix, name) >= (int)sizeof(path))
      return -1;
    FILE *file = std::fopen(path, "w");
    if (file == nullptr) return -1;
    std::fputs("time_s,hemisphere,theta_deg,phi_deg,fac_A_m2,potential_V\n", file);
    for (size_t t = 0; t < nt; ++t)
      for (size_t p = 0; p < np; ++p) {
        const size_t q = hemi * np * nt + t * np + p;
        std::fprintf(file, "%.17g,%s,%.17g,%.17g,%.17g,%.17g\n", info.epoch_s, name,
                     info.max_colat * (double)t / (double)(nt - 1U) * 180.0 / std::acos(-1.0),
                     360.0 * (double)p / (double)np, fac[q], pot[q]);
      }
    if (std::fclose(file) != 0) return -1;
  }
  return 0;
}

// cubedsphere_mi_dipole.c:650-690: J-parallel, magnetic latitude and
// longitude of the sample-shell cells at the last update.
int write_shell_current_csv(Problem &P, const char *prefix) {
  MiInfo info;
  mi_info(P, info);
  if (!info.ready || prefix == nullptr) return -1;
  char path[4096];
  if (std::snprintf(path, sizeof(path), "%s.shell_j.csv", prefix) >= (int)sizeof(path))
    return -1;
  std::vector<double> shell;
  if (mi_shell_current(P, shell) != 0) return -1;
  FILE *file = std::fopen(path, "w");
  if (file == nullptr) return -1;
  std::fputs("face,j,k,mlat_deg,phi_deg,jparallel_A_m2\n", file);
  const size_t nj = (size_t)P.L.n, nk = (size_t)P.L.n, per = 6 * nj * nk;
  for (int face = 0; face < 6; ++face)
    for (size_t j = 0; j < nj; ++j)
      for (size_t k = 0; k < nk; ++k) {
        const size_t q = ((size_t)face * nj + j) * nk + k;
        std::fprintf(file, "%d,%zu,%zu,%.17g,%.17g,%.17g\n", face, j, k, shell[per + q],
                     shell[2 * per + q], shell[q]);
      }
  return std::fclose(file) == 0 ? 0 : -1;
}

// test_cubedsphere_dipole_background.c:120-161 (write_earth_snapshot).
int write_snapshot(Problem &P, const char *path) {
  if (path == nullptr) return -1;
  FILE *out = std::fopen(path, "w");
  if (out == nullptr) {
    set_error(std::string("snapshot: cannot open ") + path);
    return -1;
  }
  std::fputs("face,i,j,k,x,y,z,rho,p,vx,vy,vz,bx,by,bz\n", out);
  const int n = P.L.n;
  const long cells = (long)P.L.nr * n * n;
  const unsigned hw = std::thread::hardware_concurrency();
  const int threads = (int)std::max(1u, std::min(16u, hw == 0 ? 1u : hw));
  std::vector<double> rows;
  for (int face = 0; face < 6; ++face) {
    long bad = -1;
    if (snapshot_face(P, face, P.cfg.startup.gamma, rows, &bad) != 0) {
      std::fclose(out);
      return -1;
    }
    // the reference writes the rows before the first cell that fails
    const long end = bad >= 0 ? bad : cells;
    std::vector<std::string> parts((size_t)threads);
    std::vector<std::thread> pool;
    const long chunk = (end + threads - 1) / threads;
    for (int w = 0; w < threads; ++w) {
      const long b = std::min(end, (long)w * chunk), e = std::min(end, b + chunk);
      pool.emplace_back([&, w, b, e] { format_rows(rows, face, n, b, e, parts[(size_t)w]); });
    }
    for (std::thread &t : pool) t.join();
    for (const std::string &s : parts)
      if (!s.empty() && std::fwrite(s.data(), 1, s.size(), out) != s.size()) {
        std::fclose(out);
        set_error("snapshot: write failed");
        return -1;
      }
    if (bad >= 0) {
      std::fclose(out);
      set_error("snapshot: a cell's state does not convert to primitives");
      return -1;
    }
  }
  if (std::fclose(out) != 0) return -1;
  if (P.cfg.mi_enabled && P.mi != nullptr)
    return write_mi_csv(P, path) == 0 && write_shell_current_csv(P, path) == 0 ? 0 : -1;
  return 0;
}

// test_cubedsphere_dipole_background.c:177-193 (write_drive_frame): the
// snapshot prefix_NNNNNN.csv and its line in the index file, whose header
// (:561) is written when the file is new or empty.
int write_drive_frame(Problem &P, const char *prefix, long frame, long step, double seconds,
                      const char *index_path) {
  if (prefix == nullptr || index_path == nullptr || frame < 0) return -1;
  char path[4096];
  const int length = std::snprintf(path, sizeof(path), "%s_%06ld.csv", prefix, frame);
  if (length < 0 || (size_t)length >= sizeof(path)) return -1;
  if (write_snapshot(P, path) != 0) return -1;
  FILE *index = std::fopen(index_path, "a");
  if (index == nullptr) return -1;
  int ok = std::fseek(index, 0, SEEK_END) == 0;
  if (ok && std::ftell(index) == 0)
    ok = std::fputs("frame,step,simulation_seconds,mhd_csv\n", index) >= 0;
  ok = ok && std::fprintf(index, "%ld,%ld,%.17g,%s\n", frame, step, seconds, path) >= 0;
  ok = (std::fclose(index) == 0) && ok;
  return ok ? 0 : -1;
}

// test_cubedsphere_dipole_background.c:61-118 (load_mhd_snapshot): the
// active cells' conserved state and residual cell B, both time levels, from
// a snapshot of the format above. Every line must parse as the reference's
// sscanf reads it, name an active cell and carry that cell's centroid
// (within 1e-8), and the line count must equal the active cell count.
// Unlike the reference, which writes as it reads, nothing is changed on
// the device when any check fails. The face fluxes are not part of a
// snapshot: the loaded cell B is for the MI coupling (the reference's MI
// replay, :526-536); a step recovers cell B from the face fluxes again.
int load_snapshot(Problem &P, const char *path) {
  if (path == nullptr) return -1;
  FILE *file = std::fopen(path, "rb");
  if (file == nullptr) {
    set_error(std::string("snapshot load: cannot open ") + path);
    return -1;
  }
  std::string text;
  char buffer[1 << 16];
  size_t got;
  while ((got = std::fread(buffer, 1, sizeof(buffer), file)) != 0) text.append(buffer, got);
  std::fclose(file);
  const char *header = "face,i,j,k,x,y,z,rho,p,vx,vy,vz,bx,by,bz\n";
  if (text.compare(0, std::strlen(header), header) != 0) {
    set_error("snapshot load: header");
    return -1;
  }
  // line starts (fgets reads lines of at most 1023 characters)
  std::vector<size_t> starts;
  for (size_t pos = std::strlen(header); pos < text.size();) {
    const size_t end = text.find('\n', pos);
    starts.push_back(pos);
    if (end == std::string::npos) break;
    pos = end + 1;
  }
  const long n = (long)starts.size();
  const long cells = 6L * P.L.nr * P.L.n * P.L.n;
  if (n != cells) {
    set_error("snapshot load: the line count is not the active cell count");
    return -1;
  }
  std::vector<double> rows((size_t)n * 11);
  std::vector<int> idx((size_t)n * 4);
  const unsigned hw = std::thread::hardware_concurrency();
  const int threads = (int)std::max(1u, std::min(16u, hw == 0 ? 1u : hw));
  std::vector<int> bad((size_t)threads, 0);
  std::vector<std::thread> pool;
  const long chunk = (n + threads - 1) / threads;
  for (int w = 0; w < threads; ++w)
    pool.emplace_back([&, w] {
      char line[1024];
      for (long r = (long)w * chunk; r < std::min(n, (long)(w + 1) * chunk) && !bad[w]; ++r) {
        const size_t b = starts[(size_t)r];
        const size_t nl = text.find('\n', b);
        const size_t e = nl == std::string::npos ? text.size() : nl + 1;
        // fgets(line, 1024) would split a longer line, and the reference
        // then fails on its remainder
        if (e - b > sizeof(line) - 1) {
          bad[w] = 1;
          break;
        }
        std::memcpy(line, text.data() + b, e - b);
        line[e - b] = '\0';
        int face;
        size_t i, j, k;
        double *x = &rows[11 * (size_t)r];
        if (std::sscanf(line, "%d,%zu,%zu,%zu,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf,%lf",
                        &face, &i, &j, &k, &x[0], &x[1], &x[2], &x[3], &x[4], &x[5], &x[6],
                        &x[7], &x[8], &x[9], &x[10]) != 15 ||
            face < 0 || face >= 6 || i >= (size_t)P.L.nr || j >= (size_t)P.L.n ||
            k >= (size_t)P.L.n) {
          bad[w] = 1;
          break;
        }
        int *q = &idx[4 * (size_t)r];
        q[0] = face;
        q[1] = (int)i;
        q[2] = (int)j;
        q[3] = (int)k;
      }
    });
  for (std::thread &t : pool) t.join();
  for (int b : bad)
    if (b) {
      set_error("snapshot load: a line does not parse or names no active cell");
      return -1;
    }
  long first_bad = -1;
  if (load_snapshot_rows(P, rows, idx, P.cfg.startup.gamma, &first_bad) != 0) return -1;
  if (first_bad >= 0) {
    set_error("snapshot load: line " + std::to_string(first_bad + 2) +
              ": centroid differs by more than 1e-8 or the state does not convert");
    return -1;
  }
  P.host_dirty = true;
  P.cfl_folded = false;
  P.b_recovered = false;
  return 0;
}

}  // namespace gcs
