// CPU producer: problem parsing, quadrature, 2D track laydown, linking and
// analytic ray tracing of a Cartesian pin lattice.  See tracks.h for the
// reference functions each part follows.
#include "tracks.h"

#include <algorithm>
#include <cmath>
#include <sstream>
#include <stdexcept>
#include <thread>

namespace moc {

namespace {

constexpr double kTinyMove = 1e-8;       // reference TINY_MOVE (src/constants.h)
constexpr double kZeroSigmaT = 1e-6;     // reference ZERO_SIGMA_T
constexpr double kFltEpsilon = 1e-12;    // reference FLT_EPSILON

std::vector<double> read_doubles(std::istringstream& ss) {
  std::vector<double> v;
  std::string tok;
  while (ss >> tok) v.push_back(std::stod(tok));
  return v;
}

BoundaryType parse_bc(const std::string& s) {
  if (s == "V") return BC_VACUUM;
  if (s == "R") return BC_REFLECTIVE;
  if (s == "P") return BC_PERIODIC;
  throw std::runtime_error("unknown boundary condition '" + s + "'");
}

}  // namespace

Problem parse_problem(const std::string& text) {
  Problem p;
  std::istringstream in(text);
  std::string line;
  bool header = false;
  std::vector<int> cells_top_first;
  MaterialData cur;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    if (key == "moc-problem") {
      int ver = 0;
      ss >> ver;
      if (ver != 1) throw std::runtime_error("unsupported moc-problem version");
      header = true;
    } else if (!header) {
      throw std::runtime_error("missing 'moc-problem 1' header");
    } else if (key == "name") {
      ss >> p.name;
    } else if (key == "groups") {
      ss >> p.num_groups;
    } else if (key == "material") {
      int id;
      cur = MaterialData();
      ss >> id >> cur.name;
    } else if (key == "sigma_t") {
      cur.sigma_t = read_doubles(ss);
    } else if (key == "sigma_s") {
      cur.sigma_s = read_doubles(ss);
    } else if (key == "nu_sigma_f") {
      cur.nu_sigma_f = read_doubles(ss);
    } else if (key == "chi") {
      cur.chi = read_doubles(ss);
    } else if (key == "endmaterial") {
      const int G = p.num_groups;
      if (G <= 0) throw std::runtime_error("'groups' must precede materials");
      if ((int)cur.sigma_t.size() != G || (int)cur.sigma_s.size() != G * G ||
          (int)cur.nu_sigma_f.size() != G || (int)cur.chi.size() != G)
        throw std::runtime_error("material '" + cur.name + "' has wrong cross-section sizes");
      // sigma_t floor (src/Material.cpp:628-650)
      for (double& s : cur.sigma_t)
        if (std::fabs(s) < kZeroSigmaT) s = kZeroSigmaT;
      // scattering given origin-major, stored dest-major (src/Material.cpp:721-732)
      std::vector<double> sd(G * G);
      for (int d = 0; d < G; d++)
        for (int o = 0; o < G; o++) sd[d * G + o] = cur.sigma_s[o * G + d];
      cur.sigma_s = sd;
      // chi normalised to sum 1 (src/Material.cpp:894-910)
      double cs = 0;
      for (double c : cur.chi) cs += c;
      if (std::fabs(cs) >= kFltEpsilon)
        for (double& c : cur.chi) c /= cs;
      cur.fissionable = false;
      for (double v : cur.nu_sigma_f)
        if (v != 0.0) cur.fissionable = true;
      p.materials.push_back(cur);
    } else if (key == "pin") {
      int id, n;
      ss >> id >> n;
      PinTemplate t;
      for (int i = 0; i < n; i++) {
        PinCylinder c;
        ss >> c.radius >> c.material >> c.rings >> c.sectors;
        t.cylinders.push_back(c);
      }
      ss >> t.outer_material >> t.outer_sectors;
      if (!ss) throw std::runtime_error("malformed pin line");
      p.pins.push_back(t);
    } else if (key == "grid") {
      ss >> p.nx >> p.ny >> p.pitch_x >> p.pitch_y;
    } else if (key == "cells") {
      int v;
      while (ss >> v) cells_top_first.push_back(v);
    } else if (key == "bc") {
      std::string b[4];
      ss >> b[0] >> b[1] >> b[2] >> b[3];
      for (int i = 0; i < 4; i++) p.bc[i] = parse_bc(b[i]);
    } else if (key == "tracks") {
      ss >> p.num_azim >> p.azim_spacing >> p.num_polar >> p.quadrature;
    } else if (key == "axial") {
      std::string b0, b1;
      ss >> p.height >> p.num_layers >> p.z_spacing >> b0 >> b1;
      if (!ss) throw std::runtime_error("malformed axial line");
      p.bc_z[0] = parse_bc(b0);
      p.bc_z[1] = parse_bc(b1);
    } else if (key == "solver") {
      ss >> p.tolerance >> p.max_iters >> p.res_type;
    } else if (key == "end") {
      break;
    } else {
      throw std::runtime_error("unknown key '" + key + "'");
    }
  }
  if (p.nx <= 0 || p.ny <= 0 || p.pitch_x <= 0 || p.pitch_y <= 0)
    throw std::runtime_error("grid must be positive");
  if ((int)cells_top_first.size() != p.nx * p.ny)
    throw std::runtime_error("cells count does not match grid");
  if (p.num_azim < 4 || p.num_azim % 4 != 0)
    throw std::runtime_error("num_azim must be a positive multiple of 4");
  if (p.num_polar < 2 || p.num_polar % 2 != 0)
    throw std::runtime_error("num_polar must be a positive even number");
  if (p.azim_spacing <= 0) throw std::runtime_error("azim_spacing must be > 0");
  if (p.height < 0) throw std::runtime_error("height must be >= 0");
  if (p.is_3d()) {
    if (p.num_layers < 1 || p.z_spacing <= 0)
      throw std::runtime_error("3D problems need >= 1 axial layer and z_spacing > 0");
    if (p.quadrature != "GL") throw std::runtime_error("3D problems use the GL polar quadrature");
    if (p.bc_z[0] == BC_PERIODIC || p.bc_z[1] == BC_PERIODIC)
      throw std::runtime_error("periodic top/bottom boundaries are not supported");
  }
  // rows given top row first (src/Universe.cpp:1604-1619); stored bottom first
  p.cells.resize(p.nx * p.ny);
  for (int j = 0; j < p.ny; j++)
    for (int i = 0; i < p.nx; i++)
      p.cells[j * p.nx + i] = cells_top_first[(p.ny - 1 - j) * p.nx + i];
  for (int c : p.cells)
    if (c < 0 || c >= (int)p.pins.size()) throw std::runtime_error("cell refers to unknown pin");
  const int M = (int)p.materials.size();
  for (const PinTemplate& t : p.pins) {
    if (t.outer_material < 0 || t.outer_material >= M) throw std::runtime_error("bad material id");
    double prev = 0;
    for (const PinCylinder& c : t.cylinders) {
      if (c.material < 0 || c.material >= M) throw std::runtime_error("bad material id");
      if (c.radius <= prev) throw std::runtime_error("cylinder radii must increase");
      if (2 * c.radius > std::min(p.pitch_x, p.pitch_y))
        throw std::runtime_error("cylinder larger than its lattice cell");
      prev = c.radius;
    }
  }
  return p;
}

// ---------------------------------------------------------------------------
// Quadrature (TY polar set; azimuthal weights from effective angles)
// ---------------------------------------------------------------------------
namespace {

void polar_set(const std::string& kind, int num_polar, std::vector<double>& sin_half,
               std::vector<double>& w_half) {
  // Tabuchi-Yamamoto sines and weights (src/Quadrature.cpp:822-894)
  if (kind != "TY") throw std::runtime_error("only the TY polar quadrature is supported in 2D");
  if (num_polar == 2) {
    sin_half = {0.798184};
    w_half = {0.5};
  } else if (num_polar == 4) {
    sin_half = {0.363900, 0.899900};
    w_half = {0.212854 / 2.0, 0.787146 / 2.0};
  } else if (num_polar == 6) {
    sin_half = {0.166648, 0.537707, 0.932954};
    w_half = {0.046233 / 2.0, 0.283619 / 2.0, 0.670148 / 2.0};
  } else {
    throw std::runtime_error("TY quadrature supports 2, 4 or 6 polar angles");
  }
}

// Sector boundary rays of a region with `ns` sectors: the reference builds
// planes with normals at i*2pi/ns + pi/4 (src/Cell.cpp:1335-1350); sector i is
// the wedge between the rays at angles i*2pi/ns - pi/4 and (i+1)*2pi/ns - pi/4.
inline double sector_theta0() { return -M_PI / 4.0; }

struct Zone {           // radial zone of a pin template (ring or outer region)
  double r2_outer;      // squared outer radius (inf for the outer region)
  int material;
  int sectors;          // 0/1 => unsectored
};

struct PinZones {
  std::vector<Zone> zones;          // inner to outer; last = outer region
  std::vector<double> radii;        // all circle radii (zone boundaries), ascending
};

PinZones build_zones(const PinTemplate& t) {
  PinZones z;
  double r_in = 0;
  for (const PinCylinder& c : t.cylinders) {
    int nr = std::max(1, c.rings);
    // equal-area rings from the outer radius inward (src/Cell.cpp:1483-1517)
    std::vector<double> rr;
    double r1 = c.radius;
    double inc = M_PI * std::fabs(c.radius * c.radius - r_in * r_in) / nr;
    rr.push_back(r1);
    for (int i = 0; i < nr - 1; i++) {
      double r2 = std::sqrt(r1 * r1 - inc / M_PI);
      rr.push_back(r2);
      r1 = r2;
    }
    std::reverse(rr.begin(), rr.end());  // ascending
    for (double r : rr) {
      z.zones.push_back({r * r, c.material, c.sectors > 1 ? c.sectors : 0});
      z.radii.push_back(r);
    }
    r_in = c.radius;
  }
  z.zones.push_back({INFINITY, t.outer_material, t.outer_sectors > 1 ? t.outer_sectors : 0});
  return z;
}

inline int zone_of(const PinZones& z, double r2) {
  int n = (int)z.zones.size();
  for (int i = 0; i < n - 1; i++)
    if (r2 < z.zones[i].r2_outer) return i;
  return n - 1;
}

inline int sector_of(int ns, double dx, double dy) {
  if (ns <= 1) return 0;
  double a = std::atan2(dy, dx) - sector_theta0();
  const double two_pi = 2.0 * M_PI;
  a = std::fmod(a, two_pi);
  if (a < 0) a += two_pi;
  int s = (int)std::floor(a / (two_pi / ns));
  return std::min(std::max(s, 0), ns - 1);
}

struct RawSeg { double length; int64_t key; int material; double mx, my, sx, sy; };

struct Tracer {
  const Problem& p;
  std::vector<PinZones> pz;
  double xmin, ymin, xmax, ymax;

  explicit Tracer(const Problem& prob) : p(prob) {
    for (const PinTemplate& t : p.pins) pz.push_back(build_zones(t));
    xmin = -0.5 * p.width_x();
    xmax = +0.5 * p.width_x();
    ymin = -0.5 * p.width_y();
    ymax = +0.5 * p.width_y();
  }

  // Classify a point (inside the domain) -> (FSR key, material).
  void classify(double x, double y, int64_t& key, int& mat) const {
    int ix = (int)std::floor((x - xmin) / p.pitch_x);
    int iy = (int)std::floor((y - ymin) / p.pitch_y);
    ix = std::min(std::max(ix, 0), p.nx - 1);
    iy = std::min(std::max(iy, 0), p.ny - 1);
    int cell = iy * p.nx + ix;
    const PinZones& z = pz[p.cells[cell]];
    double cx = xmin + (ix + 0.5) * p.pitch_x;
    double cy = ymin + (iy + 0.5) * p.pitch_y;
    double dx = x - cx, dy = y - cy;
    int zi = zone_of(z, dx * dx + dy * dy);
    int si = sector_of(z.zones[zi].sectors, dx, dy);
    key = (int64_t)cell * KEY_STRIDE + zi * 64 + si;
    mat = z.zones[zi].material;
  }

  // All surface crossings (parameters along the track) inside (0, L).
  void breakpoints(double x0, double y0, double ux, double uy, double L,
                   std::vector<double>& bp) const {
    bp.clear();
    // lattice lines
    if (std::fabs(ux) > 0) {
      for (int i = 1; i < p.nx; i++) {
        double t = (xmin + i * p.pitch_x - x0) / ux;
        if (t > 0 && t < L) bp.push_back(t);
      }
    }
    if (std::fabs(uy) > 0) {
      for (int j = 1; j < p.ny; j++) {
        double t = (ymin + j * p.pitch_y - y0) / uy;
        if (t > 0 && t < L) bp.push_back(t);
      }
    }
    // walk lattice cells along the track and add pin surfaces of each
    std::vector<double> cells_t = bp;
    cells_t.push_back(0.0);
    cells_t.push_back(L);
    std::sort(cells_t.begin(), cells_t.end());
    for (size_t k = 0; k + 1 < cells_t.size(); k++) {
      double ta = cells_t[k], tb = cells_t[k + 1];
      if (tb - ta <= 0) continue;
      double tm = 0.5 * (ta + tb);
      double xm = x0 + tm * ux, ym = y0 + tm * uy;
      int ix = std::min(std::max((int)std::floor((xm - xmin) / p.pitch_x), 0), p.nx - 1);
      int iy = std::min(std::max((int)std::floor((ym - ymin) / p.pitch_y), 0), p.ny - 1);
      const PinZones& z = pz[p.cells[iy * p.nx + ix]];
      double cx = xmin + (ix + 0.5) * p.pitch_x;
      double cy = ymin + (iy + 0.5) * p.pitch_y;
      // local frame: point(t) - centre = (dx0 + t ux, dy0 + t uy)
      double dx0 = x0 - cx, dy0 = y0 - cy;
      double b = dx0 * ux + dy0 * uy;          // half linear coefficient
      double c0 = dx0 * dx0 + dy0 * dy0;
      for (double r : z.radii) {
        double disc = b * b - (c0 - r * r);
        if (disc <= 0) continue;
        double sq = std::sqrt(disc);
        for (double t : {-b - sq, -b + sq})
          if (t > ta && t < tb) bp.push_back(t);
      }
      // sector rays: only where the crossing lies in a sectored zone
      int max_ns = 0;
      for (const Zone& zz : z.zones) max_ns = std::max(max_ns, zz.sectors);
      if (max_ns > 1) {
        // collect distinct sector counts
        std::vector<int> counts;
        for (const Zone& zz : z.zones)
          if (zz.sectors > 1 && std::find(counts.begin(), counts.end(), zz.sectors) == counts.end())
            counts.push_back(zz.sectors);
        for (int ns : counts) {
          for (int i = 0; i < ns; i++) {
            double th = sector_theta0() + i * 2.0 * M_PI / ns;
            double ex = std::cos(th), ey = std::sin(th);
            // crossing of line through centre with direction e:  (d0 + t u) x e = 0
            double den = ux * ey - uy * ex;
            if (std::fabs(den) < 1e-300) continue;
            double t = -(dx0 * ey - dy0 * ex) / den;
            if (!(t > ta && t < tb)) continue;
            double px = dx0 + t * ux, py = dy0 + t * uy;
            if (px * ex + py * ey < 0) continue;       // opposite half of the line
            int zi = zone_of(z, px * px + py * py);
            if (z.zones[zi].sectors != ns) continue;   // not a boundary there
            bp.push_back(t);
          }
        }
      }
    }
    std::sort(bp.begin(), bp.end());
  }

  // Segments of one track, following the reference marching: the point is
  // nudged TINY_MOVE past each crossing and the segment's FSR is the one
  // containing its (nudged) start point.
  void trace(double x0, double y0, double x1, double y1, double phi,
             std::vector<double>& bp, std::vector<RawSeg>& out) const {
    out.clear();
    double ux = std::cos(phi), uy = std::sin(phi);
    double L = std::hypot(x1 - x0, y1 - y0);
    breakpoints(x0, y0, ux, uy, L, bp);
    double u = kTinyMove;
    size_t k = 0;
    while (u < L) {
      while (k < bp.size() && bp[k] <= u) k++;
      double next = (k < bp.size()) ? bp[k] + kTinyMove : L + kTinyMove;
      double xs = x0 + u * ux, ys = y0 + u * uy;
      RawSeg s;
      classify(xs, ys, s.key, s.material);
      s.length = (u == kTinyMove) ? next - kTinyMove : next - u;
      // segment start point: the marching point (Geometry::segmentize2D
      // records the nudged start, src/Geometry.cpp:2196-2202)
      s.sx = xs;
      s.sy = ys;
      double um = 0.5 * (u + std::min(next, L));
      s.mx = x0 + um * ux;
      s.my = y0 + um * uy;
      out.push_back(s);
      u = next;
    }
  }
};

}  // namespace

// Azimuthal quadrature, 2D laydown, boundary conditions and links
// (src/Quadrature.cpp:646-718, src/TrackGenerator.cpp:970-1220), shared by
// the 2D and the 3D generators.
static void laydown_2d(const Problem& p, TrackSet& ts) {
  ts.num_groups = p.num_groups;
  const int na = p.num_azim, na2 = na / 2, na4 = na / 4;
  const int np = p.num_polar;
  const double W = p.width_x(), H = p.width_y();
  const double xmin = -0.5 * W, ymin = -0.5 * H;

  // ---- quadrature + laydown (src/Quadrature.cpp:646-668, TrackGenerator.cpp:970-1013)
  Quadrature& q = ts.quad;
  q.num_azim = na;
  q.num_polar = np;
  q.phi.resize(na2);
  q.azim_spacing.resize(na2);
  q.azim_weight.resize(na2);
  for (int a = 0; a < na2; a++) q.phi[a] = 2.0 * M_PI / na * (0.5 + a);
  ts.num_x.assign(na2, 0);
  ts.num_y.assign(na2, 0);
  std::vector<double> dx_eff(na2), dy_eff(na2);
  for (int a = 0; a < na4; a++) {
    double phi = q.phi[a];
    int nxa = (int)(std::fabs(W / p.azim_spacing * std::sin(phi))) + 1;
    int nya = (int)(std::fabs(H / p.azim_spacing * std::cos(phi))) + 1;
    ts.num_x[a] = ts.num_x[na2 - a - 1] = nxa;
    ts.num_y[a] = ts.num_y[na2 - a - 1] = nya;
    double phi_eff = std::atan((H * nxa) / (W * nya));
    q.phi[a] = phi_eff;
    q.phi[na2 - a - 1] = M_PI - phi_eff;
    dx_eff[a] = dx_eff[na2 - a - 1] = W / nxa;
    dy_eff[a] = dy_eff[na2 - a - 1] = H / nya;
    double sp = dx_eff[a] * std::sin(phi_eff);
    q.azim_spacing[a] = q.azim_spacing[na2 - a - 1] = sp;
  }
  // azimuthal weights (src/Quadrature.cpp:700-718)
  for (int a = 0; a < na4; a++) {
    double x1 = (a < na4 - 1) ? 0.5 * (q.phi[a + 1] - q.phi[a]) : M_PI_2 - q.phi[a];
    double x2 = (a >= 1) ? 0.5 * (q.phi[a] - q.phi[a - 1]) : q.phi[a];
    q.azim_weight[a] = q.azim_weight[na2 - a - 1] = (x1 + x2) / M_PI;
  }
  // ---- tracks: start/end points (src/TrackGenerator.cpp:1016-1062)
  long T = 0;
  std::vector<long> first(na2 + 1, 0);
  for (int a = 0; a < na2; a++) {
    first[a] = T;
    T += ts.num_x[a] + ts.num_y[a];
  }
  first[na2] = T;
  ts.num_tracks = T;
  ts.azim.resize(T);
  ts.xy.resize(T);
  ts.phi.resize(T);
  ts.start_x.resize(T);
  ts.start_y.resize(T);
  ts.end_x.resize(T);
  ts.end_y.resize(T);
  ts.next_fwd.resize(T);
  ts.next_bwd.resize(T);
  ts.next_fwd_is_fwd.resize(T);
  ts.next_bwd_is_fwd.resize(T);
  ts.bc_fwd.resize(T);
  ts.bc_bwd.resize(T);
  auto id2d = [&](int a, int i) -> long { return first[a] + i; };
  const BoundaryType bxmin = p.bc[0], bxmax = p.bc[1], bymin = p.bc[2], bymax = p.bc[3];
  for (int a = 0; a < na2; a++) {
    const int nxa = ts.num_x[a], nya = ts.num_y[a];
    const int ac = na2 - a - 1;
    for (int i = 0; i < nxa + nya; i++) {
      long t = id2d(a, i);
      ts.azim[t] = a;
      ts.xy[t] = i;
      ts.phi[t] = q.phi[a];
      double sx, sy, ex, ey;
      if (a < na4) {
        if (i < nxa) { sx = xmin + W - dx_eff[a] * (i + 0.5); sy = ymin; }
        else { sx = xmin; sy = ymin + dy_eff[a] * (i - nxa + 0.5); }
        if (i < nya) { ex = xmin + W; ey = ymin + dy_eff[a] * (i + 0.5); }
        else { ex = xmin + W - dx_eff[a] * ((i - nya) + 0.5); ey = ymin + H; }
      } else {
        if (i < nxa) { sx = xmin + dx_eff[a] * (i + 0.5); sy = ymin; }
        else { sx = xmin + W; sy = ymin + dy_eff[a] * (i - nxa + 0.5); }
        if (i < nya) { ex = xmin; ey = ymin + dy_eff[a] * (i + 0.5); }
        else { ex = xmin + dx_eff[a] * (i - nya + 0.5); ey = ymin + H; }
      }
      ts.start_x[t] = sx;
      ts.start_y[t] = sy;
      ts.end_x[t] = ex;
      ts.end_y[t] = ey;

      // boundary conditions (src/TrackGenerator.cpp:1086-1160)
      if (a < na4) {
        ts.bc_fwd[t] = (i < nya) ? bxmax : bymax;
        ts.bc_bwd[t] = (i < nxa) ? bymin : bxmin;
      } else {
        ts.bc_fwd[t] = (i < nya) ? bxmin : bymax;
        ts.bc_bwd[t] = (i < nxa) ? bymin : bxmax;
      }
      // connecting tracks (src/TrackGenerator.cpp:1162-1218)
      if (i < nya) {
        ts.next_fwd_is_fwd[t] = 1;
        ts.next_fwd[t] = (ts.bc_fwd[t] == BC_PERIODIC) ? id2d(a, i + nxa) : id2d(ac, i + nxa);
      } else {
        if (bymax == BC_PERIODIC) {
          ts.next_fwd_is_fwd[t] = 1;
          ts.next_fwd[t] = id2d(a, i - nya);
        } else {
          ts.next_fwd_is_fwd[t] = 0;
          ts.next_fwd[t] = id2d(ac, (nxa + nya) - (i - nya) - 1);
        }
      }
      if (i < nxa) {
        if (bymin == BC_PERIODIC) {
          ts.next_bwd_is_fwd[t] = 0;
          ts.next_bwd[t] = id2d(a, i + nya);
        } else {
          ts.next_bwd_is_fwd[t] = 1;
          ts.next_bwd[t] = id2d(ac, nxa - i - 1);
        }
      } else {
        ts.next_bwd_is_fwd[t] = 0;
        ts.next_bwd[t] = (ts.bc_bwd[t] == BC_PERIODIC) ? id2d(a, i - nxa) : id2d(ac, i - nxa);
      }
    }
  }

}

// Polar set and total weights of the 2D quadrature (src/Quadrature.cpp:
// 674-742, 822-894): w(a,p) = 2pi * w_a * delta_a * w_p * 2 sin(theta_p).
static void quadrature_2d(const Problem& p, TrackSet& ts) {
  const int na = p.num_azim, na2 = na / 2, na4 = na / 4;
  const int np = p.num_polar;
  Quadrature& q = ts.quad;
  std::vector<double> s_half, w_half;
  polar_set(p.quadrature, np, s_half, w_half);
  q.sin_theta.assign(np, 0);
  q.polar_weight.assign(np, 0);
  q.inv_sin_theta.assign(np, 0);
  for (int k = 0; k < np / 2; k++) {
    double theta = std::asin(s_half[k]);
    double st = std::sin(theta);
    q.sin_theta[k] = q.sin_theta[np - k - 1] = st;
    q.polar_weight[k] = q.polar_weight[np - k - 1] = w_half[k];
    q.inv_sin_theta[k] = q.inv_sin_theta[np - k - 1] = 1.0 / st;
  }
  q.total_weight.assign((size_t)na2 * np, 0);
  for (int a = 0; a < na4; a++)
    for (int k = 0; k < np / 2; k++) {
      double w = 2.0 * M_PI * q.azim_weight[a] * q.azim_spacing[a] * q.polar_weight[k] *
                 2.0 * q.sin_theta[k];
      for (int aa : {a, na2 - a - 1})
        for (int kk : {k, np - k - 1}) q.total_weight[(size_t)aa * np + kk] = w;
    }
}

TrackSet laydown_tracks_2d(const Problem& p) {
  if (p.is_3d()) throw std::invalid_argument("the 2D laydown needs a 2D problem");
  TrackSet ts;
  laydown_2d(p, ts);
  quadrature_2d(p, ts);
  return ts;
}

bool complete_fsrs_from_points(const Problem& p, TrackSet& ts, const std::vector<double>& point) {
  const long N = (long)point.size() / 2, T = ts.num_tracks;
  const long S = (long)ts.seg_length.size();
  if (N <= 0 || (long)ts.seg_offset.size() != T + 1 || ts.seg_offset[T] != S ||
      (long)ts.seg_start.size() != 2 * S)
    return false;
  for (long k = 0; k < S; k++)
    if (ts.seg_fsr[k] < 0 || ts.seg_fsr[k] >= N) return false;
  Tracer tracer(p);
  ts.num_fsrs = N;
  ts.fsr_material.assign(N, 0);
  ts.fsr_key.assign(N, 0);
  std::vector<int64_t> keys;
  for (long r = 0; r < N; r++) {
    int64_t key;
    int mat;
    const double x = point[2 * r], y = point[2 * r + 1];
    if (!(x > tracer.xmin && x < tracer.xmax && y > tracer.ymin && y < tracer.ymax)) return false;
    tracer.classify(x, y, key, mat);
    ts.fsr_key[r] = key;
    ts.fsr_material[r] = mat;
    keys.push_back(key);
  }
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) return false;  // one FSR per region
  // volumes and centroids from the segments (src/MOCKernel.cpp:86-160), the
  // midpoint of a segment from its start point along the track
  const Quadrature& q = ts.quad;
  ts.fsr_volume.assign(N, 0.0);
  ts.fsr_centroid.assign(2 * N, 0.0);
  for (long t = 0; t < T; t++) {
    const int a = ts.azim[t];
    const double w = q.azim_spacing[a] * q.azim_weight[a];
    const double ux = std::cos(ts.phi[t]), uy = std::sin(ts.phi[t]);
    for (long k = ts.seg_offset[t]; k < ts.seg_offset[t + 1]; k++) {
      const int32_t f = ts.seg_fsr[k];
      const double l = ts.seg_length[k];
      ts.fsr_volume[f] += w * l;
      ts.fsr_centroid[2 * f] += w * l * (ts.seg_start[2 * k] + 0.5 * l * ux);
      ts.fsr_centroid[2 * f + 1] += w * l * (ts.seg_start[2 * k + 1] + 0.5 * l * uy);
    }
  }
  for (long f = 0; f < N; f++) {
    if (!(ts.fsr_volume[f] > 0)) return false;  // every FSR is crossed
    ts.fsr_centroid[2 * f] /= ts.fsr_volume[f];
    ts.fsr_centroid[2 * f + 1] /= ts.fsr_volume[f];
  }
  double max_tau = 0;
  std::vector<double> max_sig(p.materials.size(), 0);
  for (size_t m = 0; m < p.materials.size(); m++)
    for (double v : p.materials[m].sigma_t) max_sig[m] = std::max(max_sig[m], v);
  for (long k = 0; k < S; k++)
    max_tau = std::max(max_tau, ts.seg_length[k] * max_sig[ts.fsr_material[ts.seg_fsr[k]]]);
  ts.max_optical_length = max_tau;
  return true;
}

TrackSet generate_tracks(const Problem& p, int threads) {
  if (p.is_3d()) return generate_tracks_3d(p, threads);
  TrackSet ts;
  laydown_2d(p, ts);
  quadrature_2d(p, ts);
  Quadrature& q = ts.quad;
  const long T = ts.num_tracks;

  // ---- ray tracing (parallel over tracks), then first-encounter FSR ids
  Tracer tracer(p);
  std::vector<std::vector<RawSeg>> per_track(T);
  int nthreads = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nthreads = (int)std::min<long>(nthreads, std::max<long>(1, T));
  auto work = [&](int tid) {
    std::vector<double> bp;
    for (long t = tid; t < T; t += nthreads)
      tracer.trace(ts.start_x[t], ts.start_y[t], ts.end_x[t], ts.end_y[t], ts.phi[t], bp,
                   per_track[t]);
  };
  if (nthreads == 1) {
    work(0);
  } else {
    std::vector<std::thread> pool;
    for (int i = 0; i < nthreads; i++) pool.emplace_back(work, i);
    for (auto& th : pool) th.join();
  }

  const int64_t max_key = (int64_t)p.nx * p.ny * KEY_STRIDE;
  std::vector<int32_t> key_to_fsr(max_key, -1);
  ts.seg_offset.assign(T + 1, 0);
  long S = 0;
  for (long t = 0; t < T; t++) S += (long)per_track[t].size();
  ts.seg_length.resize(S);
  ts.seg_fsr.resize(S);
  ts.seg_start.resize(2 * S);
  long s = 0;
  for (long t = 0; t < T; t++) {
    ts.seg_offset[t] = s;
    for (const RawSeg& r : per_track[t]) {
      int32_t& id = key_to_fsr[r.key];
      if (id < 0) {
        id = (int32_t)ts.num_fsrs++;
        ts.fsr_material.push_back(r.material);
        ts.fsr_key.push_back(r.key);
      }
      ts.seg_length[s] = r.length;
      ts.seg_fsr[s] = id;
      ts.seg_start[2 * s] = r.sx;
      ts.seg_start[2 * s + 1] = r.sy;
      s++;
    }
  }
  ts.seg_offset[T] = S;

  // ---- FSR volumes and centroids (src/MOCKernel.cpp:86-160)
  const long N = ts.num_fsrs;
  ts.fsr_volume.assign(N, 0.0);
  ts.fsr_centroid.assign(2 * N, 0.0);
  for (long t = 0; t < T; t++) {
    int a = ts.azim[t];
    double w = q.azim_spacing[a] * q.azim_weight[a];
    long off = ts.seg_offset[t];
    for (size_t k = 0; k < per_track[t].size(); k++) {
      const RawSeg& r = per_track[t][k];
      int32_t f = ts.seg_fsr[off + k];
      ts.fsr_volume[f] += w * r.length;
      ts.fsr_centroid[2 * f] += w * r.length * r.mx;
      ts.fsr_centroid[2 * f + 1] += w * r.length * r.my;
    }
  }
  for (long f = 0; f < N; f++)
    if (ts.fsr_volume[f] > 0) {
      ts.fsr_centroid[2 * f] /= ts.fsr_volume[f];
      ts.fsr_centroid[2 * f + 1] /= ts.fsr_volume[f];
    }

  // max optical length (src/TrackTraversingAlgorithms.cpp:80-110)
  double max_tau = 0;
  std::vector<double> max_sig(p.materials.size(), 0);
  for (size_t m = 0; m < p.materials.size(); m++)
    for (double v : p.materials[m].sigma_t) max_sig[m] = std::max(max_sig[m], v);
  for (long k = 0; k < S; k++)
    max_tau = std::max(max_tau, ts.seg_length[k] * max_sig[ts.fsr_material[ts.seg_fsr[k]]]);
  ts.max_optical_length = max_tau;
  return ts;
}

void split_segments(TrackSet& ts, const Problem& p, double max_tau) {
  if (ts.max_optical_length <= max_tau) return;
  std::vector<double> max_sig(p.materials.size(), 0);
  for (size_t m = 0; m < p.materials.size(); m++)
    for (double v : p.materials[m].sigma_t) max_sig[m] = std::max(max_sig[m], v);
  std::vector<int64_t> off(ts.num_tracks + 1, 0);
  std::vector<double> len, start;
  std::vector<int32_t> fsr;
  for (long t = 0; t < ts.num_tracks; t++) {
    off[t] = (int64_t)len.size();
    const double ux = std::cos(ts.phi[t]), uy = std::sin(ts.phi[t]);
    for (int64_t k = ts.seg_offset[t]; k < ts.seg_offset[t + 1]; k++) {
      double tau = ts.seg_length[k] * max_sig[ts.fsr_material[ts.seg_fsr[k]]];
      int cuts = std::max(1, (int)std::ceil(tau / max_tau));
      // sub-segments start where the previous one ended
      // (src/TrackTraversingAlgorithms.cpp:238-262)
      double x = ts.seg_start[2 * k], y = ts.seg_start[2 * k + 1];
      for (int c = 0; c < cuts; c++) {
        const double l = ts.seg_length[k] / cuts;
        len.push_back(l);
        fsr.push_back(ts.seg_fsr[k]);
        start.push_back(x);
        start.push_back(y);
        x += l * ux;
        y += l * uy;
      }
    }
  }
  off[ts.num_tracks] = (int64_t)len.size();
  ts.seg_offset.swap(off);
  ts.seg_length.swap(len);
  ts.seg_fsr.swap(fsr);
  ts.seg_start.swap(start);
  ts.max_optical_length = std::min(ts.max_optical_length, max_tau);
}



// ===========================================================================
// Explicit 3D tracks
// ===========================================================================
namespace {

// Positive roots of the Legendre polynomial P_n, ascending, by Newton's
// method from the usual cosine initial guesses; Gauss-Legendre weights
// w = 2 (1 - x^2) / (n P_{n-1}(x))^2 (the reference's getGLWeights form).
void gauss_legendre_half(int n, std::vector<double>& roots, std::vector<double>& weights) {
  auto legendre2 = [n](double x, double& pn, double& pn1) {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; k++) {
      const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    pn = p1;
    pn1 = p0;
  };
  roots.clear();
  weights.clear();
  for (int i = 0; i < n / 2; i++) {
    double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double pn, pn1;
      legendre2(x, pn, pn1);
      const double dp = n * (x * pn - pn1) / (x * x - 1.0);
      const double dx = pn / dp;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    roots.push_back(x);
  }
  std::sort(roots.begin(), roots.end());
  for (double x : roots) {
    double pn, pn1;
    legendre2(x, pn, pn1);
    weights.push_back(2.0 * (1.0 - x * x) / ((double)n * n * pn1 * pn1));
  }
}

struct Track3DRaw {
  int azim, xy, polar;
  double x1, y1, z1, x2, y2, z2, theta, phi;
};

// Position of a boundary point of a z-stack's (s, z) rectangle along its
// perimeter (s: distance along the 2D track from its start, length L; z in
// [zmin, zmax]): left side, top, right side, bottom, counter-clockwise.
inline double perimeter_coord(double s, double z, double L, double zmin, double zmax, double tol) {
  const double H = zmax - zmin;
  if (std::fabs(s) <= tol) return z - zmin;
  if (std::fabs(z - zmax) <= tol) return H + s;
  if (std::fabs(s - L) <= tol) return H + L + (zmax - z);
  return 2.0 * H + L + (L - s);  // bottom
}

}  // namespace

TrackSet generate_tracks_3d(const Problem& p, int threads) {
  TrackSet ts;
  laydown_2d(p, ts);
  ts.is_3d = true;
  ts.num_layers = p.num_layers;
  Quadrature& q = ts.quad;
  const int na = p.num_azim, na2 = na / 2, na4 = na / 4;
  const int np = p.num_polar, nph = np / 2;
  const double W = p.width_x(), H = p.width_y(), Z = p.height;
  const double xmin = -0.5 * W, xmax = 0.5 * W, ymin = -0.5 * H, ymax = 0.5 * H;
  const double zmin = -0.5 * Z, zmax = 0.5 * Z;
  const long T2 = ts.num_tracks;

  // ---- Gauss-Legendre polar set, corrected per azimuthal angle
  // (src/TrackGenerator3D.cpp:750-808, src/Quadrature.cpp:1014-1080)
  std::vector<double> roots, glw;
  gauss_legendre_half(np, roots, glw);
  q.theta.assign((size_t)na2 * np, 0);
  q.polar_spacing.assign((size_t)na2 * np, 0);
  q.sin_theta.assign(np, 0);
  q.inv_sin_theta.assign(np, 0);
  q.polar_weight.assign(np, 0);
  for (int k = 0; k < nph; k++) q.polar_weight[k] = q.polar_weight[np - k - 1] = glw[k] / 2.0;
  std::vector<int> num_l((size_t)na2 * np), num_z((size_t)na2 * np);
  std::vector<double> dl_eff((size_t)na2 * np), dz_eff((size_t)na2 * np);
  auto at = [np](int a, int k) { return (size_t)a * np + k; };
  for (int a = 0; a < na4; a++) {
    const double phi = q.phi[a];
    for (int k = 0; k < nph; k++) {
      const double theta0 = std::acos(roots[k]);
      const double length = H / std::sin(phi);
      int nl = (int)std::ceil(length * std::tan(M_PI_2 - theta0) / p.z_spacing);
      int nz = (int)std::ceil(Z * nl * std::tan(theta0) / length);
      const double dl = H / (std::sin(phi) * nl);
      const double dz = Z / nz;
      const double theta = std::atan(dl / dz);
      const int as = na2 - a - 1, kc = np - k - 1;
      for (int aa : {a, as}) {
        for (int kk : {k, kc}) {
          num_l[at(aa, kk)] = nl;
          num_z[at(aa, kk)] = nz;
          dl_eff[at(aa, kk)] = dl;
          dz_eff[at(aa, kk)] = dz;
          q.polar_spacing[at(aa, kk)] = dz * std::sin(theta);
        }
        q.theta[at(aa, k)] = theta;
        q.theta[at(aa, kc)] = M_PI - theta;
      }
    }
  }
  // w(a,p) = 2 pi * w_a * delta_a * w_p * delta_p (src/Quadrature.cpp:720-741)
  q.total_weight.assign((size_t)na2 * np, 0);
  for (int a = 0; a < na2; a++)
    for (int k = 0; k < np; k++)
      q.total_weight[at(a, k)] = 2.0 * M_PI * q.azim_weight[a] * q.azim_spacing[a] *
                                 q.polar_weight[k] * q.polar_spacing[at(a, k)];
  for (int k = 0; k < np; k++) {
    q.sin_theta[k] = std::sin(q.theta[at(0, k)]);
    q.inv_sin_theta[k] = 1.0 / q.sin_theta[k];
  }

  // ---- 2D track chains along the periodic links (src/TrackGenerator3D.cpp:886-925)
  std::vector<long> first(na2 + 1, 0);
  for (int a = 0; a < na2; a++) first[a + 1] = first[a] + ts.num_x[a] + ts.num_y[a];
  auto id2d = [&](int a, int i) -> long { return first[a] + i; };
  std::vector<double> len2d(T2);
  for (long t = 0; t < T2; t++)
    len2d[t] = std::hypot(ts.end_x[t] - ts.start_x[t], ts.end_y[t] - ts.start_y[t]);
  std::vector<std::vector<std::vector<long>>> chains(na2);
  for (int a = 0; a < na2; a++) {
    chains[a].resize(ts.num_x[a]);
    for (int x = 0; x < ts.num_x[a]; x++) {
      int i = x;
      chains[a][x].push_back(id2d(a, i));
      while (i < ts.num_y[a]) {
        i += ts.num_x[a];
        chains[a][x].push_back(id2d(a, i));
      }
    }
  }

  // ---- 3D tracks along the chains, into z-stacks (azim, xy, polar) in
  // creation order (src/TrackGenerator3D.cpp:934-1250)
  std::vector<std::vector<Track3DRaw>> stacks((size_t)T2 * np);
  auto stack_of = [&](long t2, int k) -> std::vector<Track3DRaw>& { return stacks[(size_t)t2 * np + k]; };
  const double flt_eps = 1.1920928955078125e-07;  // FLT_EPSILON
  for (int a = 0; a < na2; a++)
    for (int x = 0; x < ts.num_x[a]; x++)
      for (int k = 0; k < np; k++) {
        const std::vector<long>& ch = chains[a][x];
        const int nl = num_l[at(a, k)], nz = num_z[at(a, k)];
        const double dz = dz_eff[at(a, k)], dl = dl_eff[at(a, k)];
        const double theta = q.theta[at(a, k)];
        for (int lz = 0; lz < nl + nz; lz++) {
          // first link and start point (getFirst2DTrackLinkIndex)
          const long t0 = ch[0];
          const double phi = ts.phi[t0];
          const double cph = std::cos(phi), sph = std::sin(phi);
          double l_start = 0.0;
          if (k < nph && lz < nl) l_start = H / sph - (lz + 0.5) * dl;
          else if (k >= nph && lz >= nz) l_start = dl * (lz - nz + 0.5);
          double x_ext = ts.start_x[t0] - xmin + l_start * cph;
          double y_ext = ts.start_y[t0] - ymin + l_start * sph;
          bool nudged = false;
          if (std::fabs(l_start) > flt_eps) {
            if (std::fabs(std::round(x_ext / W) * W - x_ext) < kTinyMove ||
                std::fabs(std::round(y_ext / H) * H - y_ext) < kTinyMove) {
              l_start += 10 * kTinyMove;
              x_ext = ts.start_x[t0] - xmin + l_start * cph;
              y_ext = ts.start_y[t0] - ymin + l_start * sph;
              nudged = true;
            }
          }
          int link = std::abs((int)std::floor(x_ext / W));
          double xs = (x_ext < 0.0) ? std::fmod(x_ext, W) + xmax : std::fmod(x_ext, W) + xmin;
          double ys = std::fmod(y_ext, H) + ymin;
          double zs, ze;
          if (k < nph) {
            zs = zmin + std::max(0., (lz - nl + 0.5)) * dz;
            ze = zmax + std::min(0., (lz - nz + 0.5)) * dz;
          } else {
            zs = zmax + std::min(0., (lz - nz + 0.5)) * dz;
            ze = zmin + std::max(0., (lz - nl + 0.5)) * dz;
          }
          if (nudged) {
            xs -= 10 * kTinyMove * cph;
            ys -= 10 * kTinyMove * sph;
          }
          // split at the chain's links (set3DTrackData)
          const int first_link = link;
          double x2 = xs, y2 = ys, z2 = zs;
          bool end_of_chain = false;
          while (!end_of_chain) {
            if (link >= (int)ch.size()) break;
            const long t2 = ch[link];
            const double ph = ts.phi[t2];
            const double x1 = x2, y1 = y2, z1 = z2;
            double dl_xy;
            if (link == first_link)
              dl_xy = std::hypot(ts.end_x[t2] - x1, ts.end_y[t2] - y1);
            else
              dl_xy = len2d[t2];
            double dl_z;
            if (k < nph) dl_z = (ze - z1) * std::tan(theta);
            else dl_z = (z1 - ze) / std::tan(theta - M_PI_2);
            const double d = std::min(dl_z, dl_xy);
            x2 = x1 + d * std::cos(ph);
            y2 = y1 + d * std::sin(ph);
            if (k < nph) z2 = z1 + d / std::tan(theta);
            else z2 = z1 - d * std::tan(theta - M_PI_2);
            if (std::fabs(x2 - x1) < kTinyMove || std::fabs(y2 - y1) < kTinyMove ||
                std::fabs(z2 - z1) < kTinyMove)
              break;
            if (dl_z < dl_xy || ts.xy[t2] >= ts.num_y[a]) end_of_chain = true;
            Track3DRaw r;
            r.azim = ts.azim[t2];
            r.xy = ts.xy[t2];
            r.polar = k;
            r.x1 = std::max(xmin, std::min(xmax, x1));
            r.y1 = std::max(ymin, std::min(ymax, y1));
            r.z1 = std::max(zmin, std::min(zmax, z1));
            r.x2 = std::max(xmin, std::min(xmax, x2));
            r.y2 = std::max(ymin, std::min(ymax, y2));
            r.z2 = std::max(zmin, std::min(zmax, z2));
            r.theta = theta;
            r.phi = ph;
            stack_of(t2, k).push_back(r);
            link++;
            if (!end_of_chain) x2 = (a < na4) ? xmin : xmax;
          }
        }
      }

  // ---- uid order: stacks by (azim, xy, polar), tracks by z index
  std::vector<Track3DRaw> trk;
  std::vector<int64_t> stack_first((size_t)T2 * np + 1, 0);
  for (long t2 = 0; t2 < T2; t2++)
    for (int k = 0; k < np; k++) {
      stack_first[(size_t)t2 * np + k] = (int64_t)trk.size();
      for (const Track3DRaw& r : stack_of(t2, k)) trk.push_back(r);
    }
  stack_first[(size_t)T2 * np] = (int64_t)trk.size();
  const long T = (long)trk.size();

  // ---- links: the track leaving a face continues as the track that starts
  // at the exit point in the reflected direction (or, traversed backward,
  // ends there).  Points are matched within each candidate stack through a
  // sorted perimeter coordinate of the stack's (s, z) rectangle.
  const double tol = 1e-6;
  struct Key { double c; long t; };
  std::vector<std::vector<Key>> starts((size_t)T2 * np), ends((size_t)T2 * np);
  // 2D track tables of the laydown (ts is refilled with the 3D tracks below)
  const std::vector<double> sx2 = ts.start_x, sy2 = ts.start_y, ex2 = ts.end_x, ey2 = ts.end_y;
  const std::vector<double> phi2 = ts.phi;
  auto s_along = [&](long t2, double x, double y) {
    return (x - sx2[t2]) * std::cos(phi2[t2]) + (y - sy2[t2]) * std::sin(phi2[t2]);
  };
  for (long t = 0; t < T; t++) {
    const Track3DRaw& r = trk[t];
    const long t2 = id2d(r.azim, r.xy);
    const size_t st = (size_t)t2 * np + r.polar;
    starts[st].push_back({perimeter_coord(s_along(t2, r.x1, r.y1), r.z1, len2d[t2], zmin, zmax, tol), t});
    ends[st].push_back({perimeter_coord(s_along(t2, r.x2, r.y2), r.z2, len2d[t2], zmin, zmax, tol), t});
  }
  for (auto* v : {&starts, &ends})
    for (auto& s : *v) std::sort(s.begin(), s.end(), [](const Key& a, const Key& b) { return a.c < b.c; });
  // the track of stack (t2, k) whose start (use_end = false) or end point is (x, y, z)
  auto find = [&](long t2, int k, bool use_end, double x, double y, double z) -> long {
    const std::vector<Key>& keys = (use_end ? ends : starts)[(size_t)t2 * np + k];
    const double L = len2d[t2];
    const double c = perimeter_coord(s_along(t2, x, y), z, L, zmin, zmax, tol);
    const double per = 2.0 * (L + (zmax - zmin));
    for (double cc : {c, c - per, c + per}) {
      auto it = std::lower_bound(keys.begin(), keys.end(), cc - 1e-5,
                                 [](const Key& a, double v) { return a.c < v; });
      for (; it != keys.end() && it->c <= cc + 1e-5; ++it) {
        const Track3DRaw& o = trk[it->t];
        const double ox = use_end ? o.x2 : o.x1, oy = use_end ? o.y2 : o.y1, oz = use_end ? o.z2 : o.z1;
        if (std::fabs(ox - x) < 1e-5 && std::fabs(oy - y) < 1e-5 && std::fabs(oz - z) < 1e-5)
          return it->t;
      }
    }
    return -1;
  };
  ts.num_tracks = T;
  ts.azim.assign(T, 0);
  ts.xy.assign(T, 0);
  ts.polar.assign(T, 0);
  ts.phi.assign(T, 0);
  ts.start_x.assign(T, 0);
  ts.start_y.assign(T, 0);
  ts.end_x.assign(T, 0);
  ts.end_y.assign(T, 0);
  ts.start_z.assign(T, 0);
  ts.end_z.assign(T, 0);
  std::vector<int64_t> nf(T, -1), nb(T, -1);
  std::vector<uint8_t> nff(T, 1), nbf(T, 0), bcf(T, BC_VACUUM), bcb(T, BC_VACUUM);
  // 2D link tables of the laydown (kept before ts is overwritten below)
  const std::vector<int64_t> nf2 = ts.next_fwd, nb2 = ts.next_bwd;
  const std::vector<uint8_t> nff2 = ts.next_fwd_is_fwd, nbf2 = ts.next_bwd_is_fwd;
  const std::vector<uint8_t> bcf2 = ts.bc_fwd, bcb2 = ts.bc_bwd;
  long unmatched = 0;
  for (long t = 0; t < T; t++) {
    const Track3DRaw& r = trk[t];
    const long t2 = id2d(r.azim, r.xy);
    const int k = r.polar, kc = np - 1 - k;
    const bool up = k < nph;
    ts.azim[t] = r.azim;
    ts.xy[t] = r.xy;
    ts.polar[t] = k;
    ts.phi[t] = r.phi;
    ts.start_x[t] = r.x1;
    ts.start_y[t] = r.y1;
    ts.start_z[t] = r.z1;
    ts.end_x[t] = r.x2;
    ts.end_y[t] = r.y2;
    ts.end_z[t] = r.z2;
    // Links are set whatever the boundary condition (as the reference does;
    // a vacuum end transfers nothing): across a top/bottom face into the
    // complementary polar angle of the same z-stack, across a side face
    // along the 2D link, and at an edge (both at once) by a double
    // reflection into the 2D-linked stack with z reversed.
    // forward end
    {
      const bool on_z = up ? std::fabs(r.z2 - zmax) < tol : std::fabs(r.z2 - zmin) < tol;
      const long n2 = nf2[t2];
      const bool f = nff2[t2] != 0;
      const bool periodic_xy = bcf2[t2] == BC_PERIODIC;
      if (on_z) {
        bcf[t] = p.bc_z[up ? 1 : 0];
        nf[t] = find(t2, kc, false, r.x2, r.y2, r.z2);
        nff[t] = 1;
        if (nf[t] < 0 && !periodic_xy) {
          nf[t] = f ? find(n2, kc, false, r.x2, r.y2, r.z2) : find(n2, k, true, r.x2, r.y2, r.z2);
          nff[t] = f ? 1 : 0;
          if (bcf2[t2] == BC_VACUUM) bcf[t] = BC_VACUUM;
        }
      } else {
        bcf[t] = bcf2[t2];
        // continuing forward keeps the polar index; traversing the next
        // track backward needs the complementary one (same z direction)
        const double ex = periodic_xy ? (f ? sx2[n2] : ex2[n2]) : r.x2;
        const double ey = periodic_xy ? (f ? sy2[n2] : ey2[n2]) : r.y2;
        nf[t] = f ? find(n2, k, false, ex, ey, r.z2) : find(n2, kc, true, ex, ey, r.z2);
        nff[t] = f ? 1 : 0;
      }
      if (nf[t] < 0 && bcf[t] != BC_VACUUM) unmatched++;
    }
    // backward end (start point, travelling against the track)
    {
      const bool on_z = up ? std::fabs(r.z1 - zmin) < tol : std::fabs(r.z1 - zmax) < tol;
      const long n2 = nb2[t2];
      const bool f = nbf2[t2] != 0;
      const bool periodic_xy = bcb2[t2] == BC_PERIODIC;
      if (on_z) {
        bcb[t] = p.bc_z[up ? 0 : 1];
        nb[t] = find(t2, kc, true, r.x1, r.y1, r.z1);
        nbf[t] = 0;
        if (nb[t] < 0 && !periodic_xy) {
          nb[t] = f ? find(n2, k, false, r.x1, r.y1, r.z1) : find(n2, kc, true, r.x1, r.y1, r.z1);
          nbf[t] = f ? 1 : 0;
          if (bcb2[t2] == BC_VACUUM) bcb[t] = BC_VACUUM;
        }
      } else {
        bcb[t] = bcb2[t2];
        const double sx = periodic_xy ? (f ? sx2[n2] : ex2[n2]) : r.x1;
        const double sy = periodic_xy ? (f ? sy2[n2] : ey2[n2]) : r.y1;
        nb[t] = f ? find(n2, kc, false, sx, sy, r.z1) : find(n2, k, true, sx, sy, r.z1);
        nbf[t] = f ? 1 : 0;
      }
      if (nb[t] < 0 && bcb[t] != BC_VACUUM) unmatched++;
    }
  }
  if (unmatched) throw std::runtime_error("3D laydown: " + std::to_string(unmatched) + " unlinked track ends");
  ts.next_fwd = nf;
  ts.next_bwd = nb;
  ts.next_fwd_is_fwd = nff;
  ts.next_bwd_is_fwd = nbf;
  ts.bc_fwd = bcf;
  ts.bc_bwd = bcb;

  // ---- 3D ray tracing (Geometry::segmentize3D, src/Geometry.cpp:2232-2350):
  // crossings of the radial surfaces (from the 2D crossings of the track's
  // projection, t3 = t2 / sin(theta)) and of the axial layer planes, marched
  // with the reference's TINY_MOVE nudges; the FSR of a segment is the one
  // containing its nudged start point: (2D region, axial layer).
  Tracer tracer(p);
  const double dzl = Z / p.num_layers;
  const int64_t key2 = (int64_t)p.nx * p.ny * KEY_STRIDE;
  int nthreads = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  auto parallel = [&](long n, auto&& fn) {
    const int nt = (int)std::min<long>(nthreads, std::max<long>(1, n));
    if (nt == 1) {
      fn(0, 1);
      return;
    }
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; i++) pool.emplace_back([&, i] { fn(i, nt); });
    for (auto& th : pool) th.join();
  };
  // Every 3D track lies over one 2D track: the radial crossings along each
  // 2D track (distance from its start) and the region of each interval
  // between them are found once.
  struct Line2 { std::vector<double> b; std::vector<int64_t> key; std::vector<int> mat; };
  std::vector<Line2> lines(T2);
  parallel(T2, [&](int tid, int nt) {
    std::vector<double> bp;
    for (long t2 = tid; t2 < T2; t2 += nt) {
      const double c = std::cos(phi2[t2]), sn = std::sin(phi2[t2]);
      tracer.breakpoints(sx2[t2], sy2[t2], c, sn, len2d[t2], bp);
      Line2& ln = lines[t2];
      ln.b.push_back(0.0);
      for (double v : bp) ln.b.push_back(v);
      ln.b.push_back(len2d[t2]);
      for (size_t j = 0; j + 1 < ln.b.size(); j++) {
        const double sm = 0.5 * (ln.b[j] + ln.b[j + 1]);
        int64_t k2;
        int m;
        tracer.classify(sx2[t2] + sm * c, sy2[t2] + sm * sn, k2, m);
        ln.key.push_back(k2);
        ln.mat.push_back(m);
      }
    }
  });
  // The march of one 3D track; emit(length, packed key << 8 | material,
  // midpoint distance) per segment.  Run twice: to count, then to fill.
  auto march = [&](long t, auto&& emit) {
    const Track3DRaw& r = trk[t];
    const long t2 = id2d(r.azim, r.xy);
    const Line2& ln = lines[t2];
    const double st = std::sin(r.theta), uz = std::cos(r.theta);
    const double L3 = std::sqrt((r.x2 - r.x1) * (r.x2 - r.x1) + (r.y2 - r.y1) * (r.y2 - r.y1) +
                                (r.z2 - r.z1) * (r.z2 - r.z1));
    const double sa = s_along(t2, r.x1, r.y1);
    const double sb = sa + L3 * st;
    thread_local std::vector<double> bp, bz;
    bp.clear();
    const size_t j0 = std::upper_bound(ln.b.begin(), ln.b.end(), sa) - ln.b.begin();
    for (size_t j = j0; j + 1 < ln.b.size() && ln.b[j] < sb; j++) bp.push_back((ln.b[j] - sa) / st);
    bz.clear();
    if (std::fabs(uz) > 0)
      for (int l = 1; l < p.num_layers; l++) {
        const double tz = (zmin + l * dzl - r.z1) / uz;
        if (tz > 0 && tz < L3) bz.push_back(tz);
      }
    if (uz < 0) std::reverse(bz.begin(), bz.end());
    const size_t nrad = bp.size();
    bp.insert(bp.end(), bz.begin(), bz.end());
    std::inplace_merge(bp.begin(), bp.begin() + nrad, bp.end());
    double u = kTinyMove;
    size_t kk = 0;
    size_t iv = j0 > 0 ? j0 - 1 : 0;  // 2D interval of the current point
    while (u < L3) {
      while (kk < bp.size() && bp[kk] <= u) kk++;
      const double next = (kk < bp.size()) ? bp[kk] + kTinyMove : L3 + kTinyMove;
      const double sp = sa + u * st;
      while (iv + 1 < ln.key.size() && ln.b[iv + 1] <= sp) iv++;
      int layer = (int)std::floor((r.z1 + u * uz - zmin) / dzl);
      layer = std::min(std::max(layer, 0), p.num_layers - 1);
      const int64_t key = (int64_t)layer * key2 + ln.key[iv];
      const double length = (u == kTinyMove) ? next - kTinyMove : next - u;
      emit(length, (key << 8) | (int64_t)ln.mat[iv], 0.5 * (u + std::min(next, L3)));
      u = next;
    }
  };
  std::vector<int64_t> count(T, 0);
  parallel(T, [&](int tid, int nt) {
    for (long t = tid; t < T; t += nt) {
      int64_t c = 0;
      march(t, [&](double, int64_t, double) { c++; });
      count[t] = c;
    }
  });
  ts.seg_offset.assign(T + 1, 0);
  for (long t = 0; t < T; t++) ts.seg_offset[t + 1] = ts.seg_offset[t] + count[t];
  const long S = (long)ts.seg_offset[T];
  ts.seg_length.resize(S);
  ts.seg_fsr.resize(S);
  ts.seg_start.clear();
  std::vector<int64_t> packed(S);
  parallel(T, [&](int tid, int nt) {
    for (long t = tid; t < T; t += nt) {
      int64_t k = ts.seg_offset[t];
      march(t, [&](double length, int64_t pk, double) {
        ts.seg_length[k] = length;
        packed[k] = pk;
        k++;
      });
    }
  });

  // FSR ids by first encounter in uid order
  std::vector<int32_t> key_to_fsr((size_t)key2 * p.num_layers, -1);
  ts.num_fsrs = 0;
  ts.fsr_material.clear();
  ts.fsr_key.clear();
  for (long k = 0; k < S; k++) {
    const int64_t key = packed[k] >> 8;
    int32_t& id = key_to_fsr[key];
    if (id < 0) {
      id = (int32_t)ts.num_fsrs++;
      ts.fsr_material.push_back((int)(packed[k] & 0xff));
      ts.fsr_key.push_back(key);
    }
    ts.seg_fsr[k] = id;
  }
  packed.clear();
  packed.shrink_to_fit();
  // volumes and centroids with the 3D volume weight delta_a w_a delta_p w_p
  // (src/MOCKernel.cpp:86-103), per-thread sums added in thread order
  const long N = ts.num_fsrs;
  const int ntv = (int)std::min<long>(nthreads, std::max<long>(1, T));
  std::vector<std::vector<double>> part(ntv, std::vector<double>(4 * N, 0.0));
  parallel(T, [&](int tid, int nt) {
    std::vector<double>& acc = part[tid];
    for (long t = tid; t < T; t += nt) {
      const Track3DRaw& r = trk[t];
      const int a = r.azim, k = r.polar;
      const double w = q.azim_spacing[a] * q.azim_weight[a] * q.polar_spacing[at(a, k)] * q.polar_weight[k];
      const double st = std::sin(r.theta), uz = std::cos(r.theta);
      const double ux = std::cos(r.phi) * st, uy = std::sin(r.phi) * st;
      int64_t j = ts.seg_offset[t];
      march(t, [&](double length, int64_t, double um) {
        const int32_t f = ts.seg_fsr[j++];
        acc[4 * f] += w * length;
        acc[4 * f + 1] += w * length * (r.x1 + um * ux);
        acc[4 * f + 2] += w * length * (r.y1 + um * uy);
        acc[4 * f + 3] += w * length * (r.z1 + um * uz);
      });
    }
  });
  ts.fsr_volume.assign(N, 0.0);
  ts.fsr_centroid.assign(2 * N, 0.0);
  ts.fsr_centroid_z.assign(N, 0.0);
  for (int i = 0; i < ntv; i++)
    for (long f = 0; f < N; f++) {
      ts.fsr_volume[f] += part[i][4 * f];
      ts.fsr_centroid[2 * f] += part[i][4 * f + 1];
      ts.fsr_centroid[2 * f + 1] += part[i][4 * f + 2];
      ts.fsr_centroid_z[f] += part[i][4 * f + 3];
    }
  for (long f = 0; f < N; f++)
    if (ts.fsr_volume[f] > 0) {
      ts.fsr_centroid[2 * f] /= ts.fsr_volume[f];
      ts.fsr_centroid[2 * f + 1] /= ts.fsr_volume[f];
      ts.fsr_centroid_z[f] /= ts.fsr_volume[f];
    }
  double max_tau = 0;
  std::vector<double> max_sig(p.materials.size(), 0);
  for (size_t m = 0; m < p.materials.size(); m++)
    for (double v : p.materials[m].sigma_t) max_sig[m] = std::max(max_sig[m], v);
  for (long j = 0; j < S; j++)
    max_tau = std::max(max_tau, ts.seg_length[j] * max_sig[ts.fsr_material[ts.seg_fsr[j]]]);
  ts.max_optical_length = max_tau;
  return ts;
}

}  // namespace moc
