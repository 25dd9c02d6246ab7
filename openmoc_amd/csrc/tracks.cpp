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

TrackSet generate_tracks(const Problem& p, int threads) {
  TrackSet ts;
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
  // polar set + total weights w(a,p) = 2pi * w_a * delta_a * w_p * 2 sin(theta_p)
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

}  // namespace moc
