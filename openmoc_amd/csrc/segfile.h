// Segment files: TrackGenerator::dumpSegmentsToFile / readSegmentsFromFile
// (src/TrackGenerator.cpp:1388-1501, 1518-1638; records of DumpSegments::onTrack,
// src/TrackTraversingAlgorithms.cpp:1101-1144), host only.
//
// Layout (native byte order):
//   i32 n | char text[n]          geometry identity: the problem text + NUL
//                                 (the reference writes Geometry::toString)
//   "MOCL" i32 version, then the laydown (quadrature, tracks, links, FSR
//   volumes, 3D track data) as length-prefixed arrays: the reference lays
//   tracks down before reading segments, this reader takes them from here,
//   so reading needs neither the laydown nor the ray tracing
//   per track, uid order:         i32 num_segments, then per segment
//                                 f64 length | i32 material | i64 region |
//                                 i32 track_idx (0) | f64 x, y, z (start)
//   i64 num_fsrs, per FSR:        i32 n | key[n] | i64 fsr_id | f64 x, y, z
//                                 (centroid) | i32 material | i32 n | key[n]
#pragma once

#include <string>

#include "tracks.h"

namespace moc {

// Throws DeviceError(MOC_ERR_IO) if the file cannot be written.
void dump_segment_file(const std::string& path, const std::string& problem_text,
                       const TrackSet& ts);

// Fills `ts` from a segment file: this layout, or a 2D file the reference's
// TrackGenerator::dumpSegmentsToFile wrote (no laydown section; the tracks
// are then this problem's laydown and the records must fit it exactly).
// Returns false, leaving `ts` untouched, if the file was written for another
// problem (the reference's geometry check); throws DeviceError(MOC_ERR_IO) on
// a missing, short or malformed file of this layout.
bool read_segment_file(const std::string& path, const std::string& problem_text, TrackSet& ts);

}  // namespace moc
