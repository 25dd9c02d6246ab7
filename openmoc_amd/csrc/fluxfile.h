// FSR scalar-flux files: the binary layout of the reference's
// Solver::dumpFSRFluxes / loadFSRFluxes (src/Solver.cpp:2066-2112,
// 2135-2318), host only.
//
//   f64 k_eff | i32 num_groups | i64 num_fsrs |
//   num_fsrs x { f64 x, f64 y, f64 z, f64 flux[num_groups] }
//
// native (little-endian) byte order, no padding.  Fluxes are written as
// doubles; the reference's float build writes 8*G bytes from its float flux
// array there, so only files of its double build hold meaningful fluxes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace moc {

struct FluxFile {
  double keff = 0.0;
  int32_t num_groups = 0;
  int64_t num_fsrs = 0;
  std::vector<double> xyz;   // [3N] FSR centroids
  std::vector<double> flux;  // [N*G]
};

// Throws DeviceError(MOC_ERR_IO) if the file cannot be written.
void write_flux_file(const std::string& path, double keff, int32_t num_groups, int64_t num_fsrs,
                     const double* xyz, const double* flux);

// Throws DeviceError(MOC_ERR_IO) on a missing, short or malformed file.
FluxFile read_flux_file(const std::string& path);

// For each of the n centroids `own` [3n], the index of the nearest centroid
// of `file` [3m] among those in the 3x3x3 cells of width `tolerance` around
// it (cell index = coordinate / tolerance truncated toward zero, as in the
// reference's lookup, src/Solver.cpp:2198-2290).  Ties go to the lower file
// index.  Throws DeviceError(MOC_ERR_INVALID) if some centroid has no
// candidate.  *max_error receives the largest matched distance.
std::vector<int64_t> match_centroids(const double* file, int64_t m, const double* own, int64_t n,
                                     double tolerance, double* max_error);

}  // namespace moc
