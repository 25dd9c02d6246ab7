#include "segfile.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <type_traits>
#include <vector>

#include "hip_solver.h"

namespace moc {

namespace {

struct FileCloser {
  void operator()(FILE* f) const {
    if (f) std::fclose(f);
  }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

constexpr char kLaydownMagic[4] = {'M', 'O', 'C', 'L'};
constexpr int32_t kLaydownVersion = 1;

[[noreturn]] void io_error(const std::string& path, const std::string& what) {
  throw DeviceError(MOC_ERR_IO, "segment file " + path + ": " + what);
}

class Out {
 public:
  Out(const std::string& path) : path_(path), f_(std::fopen(path.c_str(), "wb")) {
    if (!f_) io_error(path, "could not be opened for writing");
  }
  template <typename T>
  void put(const T& v) {
    static_assert(std::is_trivially_copyable<T>::value, "plain data only");
    raw(&v, sizeof(T));
  }
  template <typename T>
  void vec(const std::vector<T>& v) {
    put<int64_t>((int64_t)v.size());
    if (!v.empty()) raw(v.data(), v.size() * sizeof(T));
  }
  void str(const std::string& s) {
    put<int32_t>((int32_t)s.size() + 1);
    raw(s.c_str(), s.size() + 1);
  }
  void raw(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_.get()) != n) io_error(path_, "write failed");
  }
  void close() {
    if (std::fflush(f_.get()) != 0) io_error(path_, "write failed");
    f_.reset();
  }

 private:
  std::string path_;
  FilePtr f_;
};

class In {
 public:
  In(const std::string& path) : path_(path), f_(std::fopen(path.c_str(), "rb")) {
    if (!f_) io_error(path, "could not be opened");
    if (std::fseek(f_.get(), 0, SEEK_END) != 0) io_error(path, "not seekable");
    const long n = std::ftell(f_.get());
    if (n < 0 || std::fseek(f_.get(), 0, SEEK_SET) != 0) io_error(path, "not seekable");
    left_ = (uint64_t)n;
  }
  template <typename T>
  T get() {
    T v;
    raw(&v, sizeof(T));
    return v;
  }
  template <typename T>
  void vec(std::vector<T>& v) {
    const int64_t n = get<int64_t>();
    if (n < 0 || (uint64_t)n > left_ / sizeof(T)) io_error(path_, "array length beyond the file");
    v.resize((size_t)n);
    if (n) raw(v.data(), (size_t)n * sizeof(T));
  }
  std::string str() {
    const int32_t n = get<int32_t>();
    if (n <= 0 || (uint64_t)n > left_) io_error(path_, "string length beyond the file");
    std::string s((size_t)n, '\0');
    raw(&s[0], (size_t)n);
    if (s.back() != '\0') io_error(path_, "string is not terminated");
    s.pop_back();
    return s;
  }
  void raw(void* p, size_t n) {
    if (n > left_ || std::fread(p, 1, n, f_.get()) != n) io_error(path_, "file is truncated");
    left_ -= n;
  }
  uint64_t left() const { return left_; }
  // the next n bytes without consuming them; false if fewer are left
  bool peek(void* p, size_t n) {
    if (n > left_) return false;
    const long at = std::ftell(f_.get());
    const bool ok = std::fread(p, 1, n, f_.get()) == n;
    if (at < 0 || std::fseek(f_.get(), at, SEEK_SET) != 0) io_error(path_, "not seekable");
    return ok;
  }

 private:
  std::string path_;
  FilePtr f_;
  uint64_t left_ = 0;
};

// A 2D file in the reference's own layout (TrackGenerator::dumpSegmentsToFile,
// src/TrackGenerator.cpp:1388-1501, EXPLICIT_2D without CMFD): identity string,
// per track (uid order, DumpSegments::onTrack, src/TrackTraversingAlgorithms.cpp:
// 1101-1144) i32 count and 48-byte records, then the FSR table.  The identity
// is the reference's Geometry::toString, which names its CSG objects, so it
// cannot be compared with the problem text: as the reference lays its tracks
// down before reading (src/TrackGenerator.cpp:818-836), the tracks here are
// this problem's laydown, and the file is taken for this problem when its
// records fit that laydown exactly.  Each FSR's material and key come from the
// point the file gives for it (the reference writes the point where the FSR was
// first found); volumes and centroids from the segments.
bool read_reference_layout(In& in, const std::string& problem_text, TrackSet& ts) {
  const Problem p = parse_problem(problem_text);
  if (p.is_3d()) return false;
  TrackSet out = laydown_tracks_2d(p);
  const int64_t T = out.num_tracks;
  constexpr uint64_t kRecordBytes = 8 + 4 + 8 + 4 + 3 * 8;
  out.seg_offset.assign((size_t)T + 1, 0);
  std::vector<int32_t> seg_ref_mat;
  for (int64_t t = 0; t < T; t++) {
    int32_t n;
    if (in.left() < 4) return false;
    n = in.get<int32_t>();
    if (n < 0 || (uint64_t)n > in.left() / kRecordBytes) return false;
    for (int32_t k = 0; k < n; k++) {
      const double l = in.get<double>();
      const int32_t m = in.get<int32_t>();
      const int64_t f = in.get<int64_t>();
      (void)in.get<int32_t>();  // track_idx
      const double x = in.get<double>(), y = in.get<double>();
      (void)in.get<double>();   // z
      if (f < 0 || f >= ((int64_t)1 << 31) || !(l >= 0.0)) return false;
      out.seg_length.push_back(l);
      out.seg_fsr.push_back((int32_t)f);
      out.seg_start.push_back(x);
      out.seg_start.push_back(y);
      seg_ref_mat.push_back(m);
    }
    out.seg_offset[(size_t)t + 1] = (int64_t)out.seg_length.size();
  }
  if (in.left() < 8) return false;
  const int64_t N = in.get<int64_t>();
  if (N <= 0 || (uint64_t)N > in.left() / 40) return false;
  std::vector<double> point((size_t)(2 * N), 0.0);
  std::vector<int32_t> ref_mat((size_t)N, 0);
  std::vector<char> seen((size_t)N, 0);
  for (int64_t i = 0; i < N; i++) {
    (void)in.str();                      // key (FSR_keys_map order)
    const int64_t id = in.get<int64_t>();
    const double x = in.get<double>(), y = in.get<double>();
    (void)in.get<double>();
    ref_mat[(size_t)i] = in.get<int32_t>();  // material of FSR i (FSRs_to_material_IDs)
    (void)in.str();                      // key of FSR i
    if (id < 0 || id >= N || seen[(size_t)id]) return false;
    seen[(size_t)id] = 1;
    point[(size_t)(2 * id)] = x;
    point[(size_t)(2 * id + 1)] = y;
  }
  if (in.left() != 0) return false;
  if (!complete_fsrs_from_points(p, out, point)) return false;
  // the file's material ids name the problem's materials one to one
  std::map<int32_t, int32_t> to_ours, to_ref;
  auto same = [&](int32_t ref, int32_t ours) {
    auto a = to_ours.emplace(ref, ours).first;
    auto b = to_ref.emplace(ours, ref).first;
    return a->second == ours && b->second == ref;
  };
  for (int64_t r = 0; r < N; r++)
    if (!same(ref_mat[(size_t)r], out.fsr_material[(size_t)r])) return false;
  for (size_t k = 0; k < seg_ref_mat.size(); k++)
    if (!same(seg_ref_mat[k], out.fsr_material[(size_t)out.seg_fsr[k]])) return false;
  ts = std::move(out);
  return true;
}

std::string fsr_key_string(const TrackSet& ts, int64_t r) {
  return ts.fsr_key.empty() ? std::to_string(r) : std::to_string(ts.fsr_key[(size_t)r]);
}

// the laydown: everything but the segment arrays, which the records carry
void put_laydown(Out& o, const TrackSet& ts) {
  o.raw(kLaydownMagic, 4);
  o.put<int32_t>(kLaydownVersion);
  const Quadrature& q = ts.quad;
  o.put<int32_t>(ts.num_groups);
  o.put<int32_t>(q.num_azim);
  o.put<int32_t>(q.num_polar);
  for (const auto* v : {&q.phi, &q.azim_spacing, &q.azim_weight, &q.sin_theta, &q.polar_weight,
                        &q.total_weight, &q.inv_sin_theta, &q.theta, &q.polar_spacing})
    o.vec(*v);
  o.put<int64_t>(ts.num_tracks);
  o.vec(ts.num_x);
  o.vec(ts.num_y);
  o.vec(ts.azim);
  o.vec(ts.xy);
  for (const auto* v : {&ts.phi, &ts.start_x, &ts.start_y, &ts.end_x, &ts.end_y}) o.vec(*v);
  o.vec(ts.next_fwd);
  o.vec(ts.next_bwd);
  for (const auto* v : {&ts.next_fwd_is_fwd, &ts.next_bwd_is_fwd, &ts.bc_fwd, &ts.bc_bwd}) o.vec(*v);
  o.vec(ts.fsr_volume);
  o.vec(ts.fsr_centroid);
  o.vec(ts.fsr_key);
  o.put<double>(ts.max_optical_length);
  o.put<int32_t>(ts.is_3d ? 1 : 0);
  o.vec(ts.polar);
  o.vec(ts.start_z);
  o.vec(ts.end_z);
  o.vec(ts.fsr_centroid_z);
  o.put<int32_t>(ts.num_layers);
}

void get_laydown(In& in, const std::string& path, TrackSet& ts) {
  char magic[4];
  in.raw(magic, 4);
  if (std::memcmp(magic, kLaydownMagic, 4) != 0) io_error(path, "no laydown section");
  if (in.get<int32_t>() != kLaydownVersion) io_error(path, "unknown laydown version");
  Quadrature& q = ts.quad;
  ts.num_groups = in.get<int32_t>();
  q.num_azim = in.get<int32_t>();
  q.num_polar = in.get<int32_t>();
  for (auto* v : {&q.phi, &q.azim_spacing, &q.azim_weight, &q.sin_theta, &q.polar_weight,
                  &q.total_weight, &q.inv_sin_theta, &q.theta, &q.polar_spacing})
    in.vec(*v);
  ts.num_tracks = (long)in.get<int64_t>();
  in.vec(ts.num_x);
  in.vec(ts.num_y);
  in.vec(ts.azim);
  in.vec(ts.xy);
  for (auto* v : {&ts.phi, &ts.start_x, &ts.start_y, &ts.end_x, &ts.end_y}) in.vec(*v);
  in.vec(ts.next_fwd);
  in.vec(ts.next_bwd);
  for (auto* v : {&ts.next_fwd_is_fwd, &ts.next_bwd_is_fwd, &ts.bc_fwd, &ts.bc_bwd}) in.vec(*v);
  in.vec(ts.fsr_volume);
  in.vec(ts.fsr_centroid);
  in.vec(ts.fsr_key);
  ts.max_optical_length = in.get<double>();
  ts.is_3d = in.get<int32_t>() != 0;
  in.vec(ts.polar);
  in.vec(ts.start_z);
  in.vec(ts.end_z);
  in.vec(ts.fsr_centroid_z);
  ts.num_layers = in.get<int32_t>();
}

}  // namespace

void dump_segment_file(const std::string& path, const std::string& problem_text,
                       const TrackSet& ts) {
  const int64_t T = ts.num_tracks, N = ts.num_fsrs;
  if (T <= 0 || (int64_t)ts.seg_offset.size() != T + 1)
    throw DeviceError(MOC_ERR_STATE, "no segments to dump");
  Out o(path);
  o.str(problem_text);
  put_laydown(o, ts);
  // DumpSegments::onTrack records (src/TrackTraversingAlgorithms.cpp:1101-1144)
  const bool has_start2d = ts.seg_start.size() == 2 * ts.seg_length.size();
  for (int64_t t = 0; t < T; t++) {
    const int64_t b = ts.seg_offset[(size_t)t], e = ts.seg_offset[(size_t)t + 1];
    o.put<int32_t>((int32_t)(e - b));
    // 3D tracks carry no per-segment start points: walk from the track's
    // start along its direction
    double ux = 0, uy = 0, uz = 0, x = 0, y = 0, z = 0;
    if (!has_start2d) {
      x = ts.start_x[(size_t)t];
      y = ts.start_y[(size_t)t];
      z = ts.is_3d ? ts.start_z[(size_t)t] : 0.0;
      ux = ts.end_x[(size_t)t] - x;
      uy = ts.end_y[(size_t)t] - y;
      uz = ts.is_3d ? ts.end_z[(size_t)t] - z : 0.0;
      const double len = std::sqrt(ux * ux + uy * uy + uz * uz);
      if (len > 0) {
        ux /= len;
        uy /= len;
        uz /= len;
      }
    }
    for (int64_t k = b; k < e; k++) {
      const int32_t f = ts.seg_fsr[(size_t)k];
      o.put<double>(ts.seg_length[(size_t)k]);
      o.put<int32_t>(ts.fsr_material[(size_t)f]);
      o.put<int64_t>(f);
      o.put<int32_t>(0);
      if (has_start2d) {
        o.put<double>(ts.seg_start[(size_t)(2 * k)]);
        o.put<double>(ts.seg_start[(size_t)(2 * k + 1)]);
        o.put<double>(0.0);
      } else {
        o.put<double>(x);
        o.put<double>(y);
        o.put<double>(z);
        x += ux * ts.seg_length[(size_t)k];
        y += uy * ts.seg_length[(size_t)k];
        z += uz * ts.seg_length[(size_t)k];
      }
    }
  }
  // FSR table (src/TrackGenerator.cpp:1432-1469)
  o.put<int64_t>(N);
  for (int64_t r = 0; r < N; r++) {
    const std::string key = fsr_key_string(ts, r);
    o.str(key);
    o.put<int64_t>(r);
    o.put<double>(ts.fsr_centroid[(size_t)(2 * r)]);
    o.put<double>(ts.fsr_centroid[(size_t)(2 * r + 1)]);
    o.put<double>(ts.fsr_centroid_z.empty() ? 0.0 : ts.fsr_centroid_z[(size_t)r]);
    o.put<int32_t>(ts.fsr_material[(size_t)r]);
    o.str(key);
  }
  o.close();
}

bool read_segment_file(const std::string& path, const std::string& problem_text, TrackSet& ts) {
  In in(path);
  const std::string identity = in.str();
  char magic[4] = {0, 0, 0, 0};
  if (!in.peek(magic, 4) || std::memcmp(magic, kLaydownMagic, 4) != 0)
    return read_reference_layout(in, problem_text, ts);
  if (identity != problem_text) return false;
  TrackSet out;
  get_laydown(in, path, out);
  const int64_t T = out.num_tracks;
  if (T <= 0 || (int64_t)out.azim.size() != T) io_error(path, "inconsistent track count");
  // segment records (ReadSegments::onTrack, src/TrackTraversingAlgorithms.cpp:1176-1233)
  constexpr uint64_t kRecordBytes = 8 + 4 + 8 + 4 + 3 * 8;
  out.seg_offset.assign((size_t)T + 1, 0);
  std::vector<double> x2;
  for (int64_t t = 0; t < T; t++) {
    const int32_t n = in.get<int32_t>();
    if (n < 0 || (uint64_t)n > in.left() / kRecordBytes) io_error(path, "segment count beyond the file");
    for (int32_t k = 0; k < n; k++) {
      const double l = in.get<double>();
      (void)in.get<int32_t>();  // material: the FSR table's is authoritative
      const int64_t f = in.get<int64_t>();
      (void)in.get<int32_t>();  // track_idx
      const double x = in.get<double>(), y = in.get<double>();
      (void)in.get<double>();   // z
      if (f < 0 || f >= ((int64_t)1 << 31)) io_error(path, "region id out of range");
      out.seg_length.push_back(l);
      out.seg_fsr.push_back((int32_t)f);
      if (!out.is_3d) {
        x2.push_back(x);
        x2.push_back(y);
      }
    }
    out.seg_offset[(size_t)t + 1] = (int64_t)out.seg_length.size();
  }
  out.seg_start.swap(x2);
  // FSR table
  const int64_t N = in.get<int64_t>();
  if (N <= 0 || (uint64_t)N > in.left() / 40) io_error(path, "FSR count beyond the file");
  out.num_fsrs = (long)N;
  out.fsr_material.assign((size_t)N, 0);
  for (int64_t i = 0; i < N; i++) {
    (void)in.str();
    const int64_t id = in.get<int64_t>();
    (void)in.get<double>();
    (void)in.get<double>();
    (void)in.get<double>();
    const int32_t m = in.get<int32_t>();
    (void)in.str();
    if (id < 0 || id >= N) io_error(path, "FSR id out of range");
    out.fsr_material[(size_t)id] = m;
  }
  for (int32_t f : out.seg_fsr)
    if (f >= N) io_error(path, "segment region beyond the FSR table");
  if ((int64_t)out.fsr_volume.size() != N || (int64_t)out.fsr_centroid.size() != 2 * N)
    io_error(path, "FSR arrays do not match the FSR count");
  if (in.left() != 0) io_error(path, "trailing bytes");
  ts = std::move(out);
  return true;
}

}  // namespace moc
