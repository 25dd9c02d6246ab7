#include "fluxfile.h"

#include <cmath>
#include <cstdio>
#include <limits>
#include <memory>
#include <unordered_map>

#include "hip_solver.h"

namespace moc {

namespace {

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) std::fclose(f);
  }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

[[noreturn]] void io_error(const std::string& path, const std::string& what) {
  throw DeviceError(MOC_ERR_IO, "flux file " + path + ": " + what);
}

template <typename T>
void put(FILE* f, const T* v, size_t n, const std::string& path) {
  if (n && std::fwrite(v, sizeof(T), n, f) != n) io_error(path, "write failed");
}

template <typename T>
void get(FILE* f, T* v, size_t n, const std::string& path) {
  if (n && std::fread(v, sizeof(T), n, f) != n) io_error(path, "file is shorter than its header says");
}

// cell index of a coordinate: truncation toward zero of x / tolerance
int64_t cell_of(double x, double tolerance) {
  const double c = x / tolerance;
  if (!(std::fabs(c) < 9.0e15)) return c < 0 ? -(int64_t)9e15 : (int64_t)9e15;
  return (int64_t)c;
}

// 21 bits per axis: cells within +-2^20 of zero map without collision; beyond
// that keys may alias, which only adds candidates (distances decide)
uint64_t cell_key(int64_t x, int64_t y, int64_t z) {
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)x & m) | (((uint64_t)y & m) << 21) | (((uint64_t)z & m) << 42);
}

}  // namespace

void write_flux_file(const std::string& path, double keff, int32_t num_groups, int64_t num_fsrs,
                     const double* xyz, const double* flux) {
  if (num_groups <= 0 || num_fsrs <= 0 || !xyz || !flux)
    throw DeviceError(MOC_ERR_INVALID, "flux file needs groups, FSRs, centroids and fluxes");
  FilePtr f(std::fopen(path.c_str(), "wb"));
  if (!f) io_error(path, "could not be opened for writing");
  put(f.get(), &keff, 1, path);
  put(f.get(), &num_groups, 1, path);
  put(f.get(), &num_fsrs, 1, path);
  for (int64_t r = 0; r < num_fsrs; r++) {
    put(f.get(), xyz + 3 * r, 3, path);
    put(f.get(), flux + r * num_groups, (size_t)num_groups, path);
  }
  if (std::fflush(f.get()) != 0) io_error(path, "write failed");
}

FluxFile read_flux_file(const std::string& path) {
  FilePtr f(std::fopen(path.c_str(), "rb"));
  if (!f) io_error(path, "could not be opened");
  FluxFile out;
  get(f.get(), &out.keff, 1, path);
  get(f.get(), &out.num_groups, 1, path);
  get(f.get(), &out.num_fsrs, 1, path);
  if (out.num_groups <= 0 || out.num_groups > (1 << 16) || out.num_fsrs <= 0)
    io_error(path, "bad header (groups " + std::to_string(out.num_groups) + ", FSRs " +
                       std::to_string(out.num_fsrs) + ")");
  // the header must agree with the file size before anything is allocated
  if (std::fseek(f.get(), 0, SEEK_END) != 0) io_error(path, "not seekable");
  const long size = std::ftell(f.get());
  const long head = (long)(sizeof(double) + sizeof(int32_t) + sizeof(int64_t));
  const double want = (double)head + (double)out.num_fsrs * (3.0 + out.num_groups) * sizeof(double);
  if (size < 0 || (double)size != want)
    io_error(path, "size " + std::to_string(size) + " does not match its header");
  if (std::fseek(f.get(), head, SEEK_SET) != 0) io_error(path, "not seekable");
  const int G = out.num_groups;
  out.xyz.resize((size_t)out.num_fsrs * 3);
  out.flux.resize((size_t)out.num_fsrs * G);
  for (int64_t r = 0; r < out.num_fsrs; r++) {
    get(f.get(), out.xyz.data() + 3 * r, 3, path);
    get(f.get(), out.flux.data() + r * G, (size_t)G, path);
  }
  return out;
}

std::vector<int64_t> match_centroids(const double* file, int64_t m, const double* own, int64_t n,
                                     double tolerance, double* max_error) {
  if (!(tolerance > 0.0)) throw DeviceError(MOC_ERR_INVALID, "centroid tolerance must be > 0");
  if (m <= 0 || n < 0 || !file || (n > 0 && !own))
    throw DeviceError(MOC_ERR_INVALID, "centroid arrays are empty");
  // file centroids bucketed by cell, in file order
  std::unordered_map<uint64_t, std::vector<int64_t>> cells;
  cells.reserve((size_t)m);
  for (int64_t j = 0; j < m; j++)
    cells[cell_key(cell_of(file[3 * j], tolerance), cell_of(file[3 * j + 1], tolerance),
                   cell_of(file[3 * j + 2], tolerance))]
        .push_back(j);
  std::vector<int64_t> idx((size_t)n, -1);
  double worst = 0.0;
  for (int64_t r = 0; r < n; r++) {
    const double* p = own + 3 * r;
    const int64_t cx = cell_of(p[0], tolerance), cy = cell_of(p[1], tolerance),
                  cz = cell_of(p[2], tolerance);
    double best = std::numeric_limits<double>::infinity();
    int64_t pick = -1;
    for (int dx = -1; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dz = -1; dz <= 1; dz++) {
          auto it = cells.find(cell_key(cx + dx, cy + dy, cz + dz));
          if (it == cells.end()) continue;
          for (int64_t j : it->second) {
            const double ex = file[3 * j] - p[0], ey = file[3 * j + 1] - p[1],
                         ez = file[3 * j + 2] - p[2];
            const double d = std::sqrt(ex * ex + ey * ey + ez * ez);
            if (d < best || (d == best && j < pick)) {
              best = d;
              pick = j;
            }
          }
        }
    if (pick < 0)
      throw DeviceError(MOC_ERR_INVALID,
                        "no FSR of the flux file lies near FSR " + std::to_string(r) + " (" +
                            std::to_string(p[0]) + ", " + std::to_string(p[1]) + ", " +
                            std::to_string(p[2]) + ")");
    idx[(size_t)r] = pick;
    worst = std::max(worst, best);
  }
  if (max_error) *max_error = worst;
  return idx;
}

}  // namespace moc
