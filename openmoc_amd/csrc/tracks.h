// CPU producer of the MOC hot path's input: problem description, polar /
// azimuthal quadrature, cyclic 2D track laydown with reflective/periodic
// linking, and analytic ray tracing of a Cartesian pin lattice into SoA
// segment streams with first-encounter FSR numbering.
//
// Behaviour follows the reference TrackGenerator / Geometry so that the
// tracks handed to the solver are the reference's tracks:
//   laydown + effective angles      src/TrackGenerator.cpp:950-1077
//   reflective / periodic links     src/TrackGenerator.cpp:1086-1220
//   uid order (azim-major)          src/TrackGenerator.cpp:1795-1818
//   azimuthal + polar weights       src/Quadrature.cpp:646-742, 822-894
//   segment marching (TINY_MOVE)    src/Geometry.cpp:1107-1114, 1159-1271,
//                                   2085-2220
//   FSR ids by first encounter      src/Geometry.cpp:1281-1338
//   FSR volumes                     src/MOCKernel.cpp:86-160
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace moc {

enum BoundaryType : uint8_t { BC_VACUUM = 0, BC_REFLECTIVE = 1, BC_PERIODIC = 2 };

struct MaterialData {
  std::string name;
  std::vector<double> sigma_t;     // [G]
  std::vector<double> sigma_s;     // [G*G] dest-major: s[dest*G + orig]
  std::vector<double> nu_sigma_f;  // [G]
  std::vector<double> chi;         // [G], normalised to sum 1 when nonzero
  bool fissionable = false;
};

struct PinCylinder { double radius; int material; int rings; int sectors; };

struct PinTemplate {
  std::vector<PinCylinder> cylinders;  // innermost first
  int outer_material = 0;
  int outer_sectors = 0;
};

// A problem: materials + Cartesian grid of pin templates + BCs + laydown and
// solver parameters ("moc-problem 1" text format).
struct Problem {
  std::string name;
  int num_groups = 0;
  std::vector<MaterialData> materials;
  std::vector<PinTemplate> pins;
  int nx = 0, ny = 0;
  double pitch_x = 0, pitch_y = 0;
  std::vector<int> cells;              // [ny*nx], row 0 = bottom (min y)
  BoundaryType bc[4] = {BC_REFLECTIVE, BC_REFLECTIVE, BC_REFLECTIVE, BC_REFLECTIVE};  // xmin xmax ymin ymax
  int num_azim = 4;
  double azim_spacing = 0.1;
  int num_polar = 6;
  std::string quadrature = "TY";
  // 3D (axially extruded lattice; height > 0 selects explicit 3D tracks):
  // `axial <height> <layers> <z_spacing> <bc_zmin> <bc_zmax>`
  double height = 0;
  int num_layers = 1;
  double z_spacing = 0;
  BoundaryType bc_z[2] = {BC_REFLECTIVE, BC_REFLECTIVE};  // zmin zmax
  double tolerance = 1e-5;
  int max_iters = 1000;
  std::string res_type = "FISSION_SOURCE";

  double width_x() const { return nx * pitch_x; }
  double width_y() const { return ny * pitch_y; }
  bool is_3d() const { return height > 0; }
};

// Parses "moc-problem 1" text. Throws std::runtime_error on malformed input.
Problem parse_problem(const std::string& text);

struct Quadrature {
  int num_azim = 0;                   // full circle
  int num_polar = 0;                  // full sphere (2D uses num_polar/2)
  std::vector<double> phi;            // [num_azim/2] effective angles
  std::vector<double> azim_spacing;   // [num_azim/2] effective spacing
  std::vector<double> azim_weight;    // [num_azim/2]
  std::vector<double> sin_theta;      // [num_polar]
  std::vector<double> polar_weight;   // [num_polar]
  std::vector<double> total_weight;   // [num_azim/2][num_polar] w(a,p)
  std::vector<double> inv_sin_theta;  // [num_polar]
  // 3D only: corrected polar angle and polar spacing per (azim, polar)
  std::vector<double> theta;          // [num_azim/2][num_polar]
  std::vector<double> polar_spacing;  // [num_azim/2][num_polar]
};

// Tracks in uid order plus their segments (SoA) and the FSR table.
struct TrackSet {
  int num_groups = 0;
  Quadrature quad;
  long num_tracks = 0;
  std::vector<int> num_x, num_y;            // per azimuthal index
  std::vector<int32_t> azim, xy;
  std::vector<double> phi, start_x, start_y, end_x, end_y;
  std::vector<int64_t> next_fwd, next_bwd;
  std::vector<uint8_t> next_fwd_is_fwd, next_bwd_is_fwd, bc_fwd, bc_bwd;
  std::vector<int64_t> seg_offset;          // [T+1]
  std::vector<double> seg_length;           // [S]
  std::vector<int32_t> seg_fsr;             // [S]
  std::vector<double> seg_start;            // [2S] start point (x, y) of each segment
  long num_fsrs = 0;
  std::vector<int32_t> fsr_material;        // [N]
  std::vector<double> fsr_volume;           // [N]
  std::vector<double> fsr_centroid;         // [2N]
  std::vector<int64_t> fsr_key;             // [N] cell*KEY_STRIDE + local region
                                            //     (3D: + layer * nx*ny*KEY_STRIDE)
  double max_optical_length = 0;            // max sigma_t * length over segments
  // 3D tracks (explicit Track3D): polar index and z of the end points
  bool is_3d = false;
  std::vector<int32_t> polar;               // [T]
  std::vector<double> start_z, end_z;       // [T]
  std::vector<double> fsr_centroid_z;       // [N]
  int num_layers = 1;
};

constexpr int64_t KEY_STRIDE = 4096;

// Builds the quadrature and the track laydown, traces all tracks and fills
// FSR volumes / centroids. `threads` <= 0 uses the hardware concurrency.
// A problem with a height builds explicit 3D tracks (generate_tracks_3d).
TrackSet generate_tracks(const Problem& p, int threads = 0);

// Explicit 3D tracks of an axially extruded lattice
// (TrackGenerator3D with EXPLICIT_3D segmentation):
//   Gauss-Legendre polar set              src/Quadrature.cpp:1014-1290
//   polar correction, z-stacks, chains    src/TrackGenerator3D.cpp:683-1250
//   3D ray tracing                        src/Geometry.cpp:2232-2350
//   3D weights and FSR volumes            src/Quadrature.cpp:674-742,
//                                         src/MOCKernel.cpp:86-103
// Links are found geometrically: the track that starts (or, traversed
// backward, ends) where a track leaves a face, in the reflected direction.
TrackSet generate_tracks_3d(const Problem& p, int threads = 0);

// The 2D quadrature and track laydown with links, without ray tracing (what
// the reference's TrackGenerator holds before readSegmentsFromFile,
// src/TrackGenerator.cpp:818-836).
TrackSet laydown_tracks_2d(const Problem& p);

// FSR data of a laid-down 2D TrackSet whose segments (seg_offset, seg_length,
// seg_fsr, seg_start) came from elsewhere, given one point inside each FSR
// [2N]: each FSR's key and material from its point, volumes and centroids from
// the segments, the maximum optical length.  False if the data do not fit the
// problem (a point outside the domain, two FSRs in one region, an FSR no
// segment crosses, a region id out of range).
bool complete_fsrs_from_points(const Problem& p, TrackSet& ts, const std::vector<double>& point);

// Splits every segment whose optical length exceeds `max_tau` into equal
// pieces (reference TrackGenerator::splitSegments semantics,
// src/TrackTraversingAlgorithms.cpp:195-268).
void split_segments(TrackSet& ts, const Problem& p, double max_tau);

}  // namespace moc
