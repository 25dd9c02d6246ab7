/*
 * ref_driver: runs the *reference* OpenMOC CPUSolver on a "moc-problem 1"
 * text file and writes everything a checker needs into a binary record file.
 *
 * TEST INFRASTRUCTURE ONLY.  This driver is compiled together with the
 * reference's own C++ sources (src/*.cpp under the reference tree) by
 * oracle/Makefile into oracle/_ref/.  It is used to:
 *   - pin the oracle restatement (oracle/moc_oracle.cpp) against the real
 *     CPUSolver on the reference's own tracks,
 *   - pin this package's tracer against the reference TrackGenerator
 *     (laydown, links, segment lengths, FSR numbering and volumes),
 *   - time the reference CPUSolver sweep as the CPU baseline.
 * It is never linked into, or called by, the product library.
 *
 * Model construction mirrors how the reference's own drivers build a pin
 * lattice (profile/models/pin-cell/pin-cell-3d.cpp:88-150): materials via
 * Material::setSigma*, ZCylinder-bounded fuel cells with setNumRings /
 * setNumSectors, a Lattice of pin universes inside a root cell bounded by
 * X/Y planes carrying the boundary conditions, TrackGenerator(geometry,
 * num_azim, spacing) with a TY polar quadrature, then
 * CPUSolver::computeEigenvalue (src/Solver.cpp:1542-1689).
 *
 * Output ("MOCDUMP1"): a sequence of records
 *   u32 name_len, name bytes, u8 dtype ('d' f64, 'f' f32, 'i' i32, 'l' i64),
 *   u64 count, count * sizeof(dtype) bytes.
 */
#include "CPULSSolver.h"
#include "CPUSolver.h"
#include "Geometry.h"
#include "Quadrature.h"
#include "TrackGenerator.h"
#include "TrackGenerator3D.h"
#include "log.h"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

struct Cyl { double r; int mat; int rings; int sectors; };
struct Pin { std::vector<Cyl> cyl; int outer_mat; int outer_sectors; };

struct Spec {
  std::string name;
  int groups = 0;
  std::vector<std::vector<double>> sigma_t, sigma_s, nu_sigma_f, chi;
  std::vector<std::string> mat_names;
  std::vector<Pin> pins;
  int nx = 0, ny = 0;
  double px = 0, py = 0;
  std::vector<int> cells;
  std::string bc[4];
  int num_azim = 4;
  double spacing = 0.1;
  int num_polar = 6;
  std::string quad = "TY";
  double tol = 1e-5;
  int max_iters = 1000;
  std::string res_type = "FISSION_SOURCE";
  double height = 0;        // > 0: axially extruded 3D problem
  int layers = 1;
  double z_spacing = 0;
  std::string bcz[2] = {"R", "R"};
};

std::vector<double> read_doubles(std::istringstream& ss) {
  std::vector<double> v;
  double x;
  while (ss >> x) v.push_back(x);
  return v;
}

Spec parse(const char* path) {
  std::ifstream in(path);
  if (!in) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  Spec s;
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    if (key == "moc-problem") continue;
    if (key == "name") ss >> s.name;
    else if (key == "groups") ss >> s.groups;
    else if (key == "material") { int id; std::string nm; ss >> id >> nm; s.mat_names.push_back(nm); }
    else if (key == "sigma_t") s.sigma_t.push_back(read_doubles(ss));
    else if (key == "sigma_s") s.sigma_s.push_back(read_doubles(ss));
    else if (key == "nu_sigma_f") s.nu_sigma_f.push_back(read_doubles(ss));
    else if (key == "chi") s.chi.push_back(read_doubles(ss));
    else if (key == "endmaterial") continue;
    else if (key == "pin") {
      int id, n; ss >> id >> n;
      Pin p;
      for (int i = 0; i < n; i++) { Cyl c; ss >> c.r >> c.mat >> c.rings >> c.sectors; p.cyl.push_back(c); }
      ss >> p.outer_mat >> p.outer_sectors;
      s.pins.push_back(p);
    }
    else if (key == "grid") ss >> s.nx >> s.ny >> s.px >> s.py;
    else if (key == "cells") { int v; while (ss >> v) s.cells.push_back(v); }
    else if (key == "bc") ss >> s.bc[0] >> s.bc[1] >> s.bc[2] >> s.bc[3];
    else if (key == "tracks") ss >> s.num_azim >> s.spacing >> s.num_polar >> s.quad;
    else if (key == "solver") ss >> s.tol >> s.max_iters >> s.res_type;
    else if (key == "axial") ss >> s.height >> s.layers >> s.z_spacing >> s.bcz[0] >> s.bcz[1];
    else if (key == "end") break;
    else { fprintf(stderr, "unknown key %s\n", key.c_str()); exit(2); }
  }
  return s;
}

boundaryType bc_of(const std::string& c) {
  if (c == "V") return VACUUM;
  if (c == "P") return PERIODIC;
  return REFLECTIVE;
}

struct Writer {
  FILE* f;
  explicit Writer(const char* path) {
    f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    fwrite("MOCDUMP1", 1, 8, f);
  }
  ~Writer() { fclose(f); }
  template <typename T> void put(const char* name, char dt, const T* data, uint64_t n) {
    uint32_t len = (uint32_t)strlen(name);
    fwrite(&len, 4, 1, f);
    fwrite(name, 1, len, f);
    fwrite(&dt, 1, 1, f);
    fwrite(&n, 8, 1, f);
    if (n) fwrite(data, sizeof(T), n, f);
  }
  void d(const char* n, const std::vector<double>& v) { put(n, 'd', v.data(), v.size()); }
  void i(const char* n, const std::vector<int32_t>& v) { put(n, 'i', v.data(), v.size()); }
  void l(const char* n, const std::vector<int64_t>& v) { put(n, 'l', v.data(), v.size()); }
  void d1(const char* n, double x) { put(n, 'd', &x, 1); }
  void l1(const char* n, int64_t x) { put(n, 'l', &x, 1); }
};

/* Subclasses only to read the protected timer and linear-source moments. */
class TimedCPUSolver : public CPUSolver {
 public:
  explicit TimedCPUSolver(TrackGenerator* tg) : CPUSolver(tg) {}
  double split(const char* name) { return _timer->getSplit(name); }
  void sweep_once() { transportSweep(); }
  void dump_ls(Writer&) {}
};

class TimedCPULSSolver : public CPULSSolver {
 public:
  explicit TimedCPULSSolver(TrackGenerator* tg) : CPULSSolver(tg) {}
  double split(const char* name) { return _timer->getSplit(name); }
  void sweep_once() { transportSweep(); }
  /* scalar-flux moments [r][x|y|z][g] and the FSR centroids */
  void dump_ls(Writer& w) {
    long n = _num_FSRs * _NUM_GROUPS * 3;
    std::vector<double> m(_scalar_flux_xyz, _scalar_flux_xyz + n);
    w.d("flux_xyz", m);
    std::vector<double> c(2 * _num_FSRs);
    for (long r = 0; r < _num_FSRs; r++) {
      Point* p = _geometry->getFSRCentroid(r);
      c[2 * r] = p->getX();
      c[2 * r + 1] = p->getY();
    }
    w.d("fsr_centroid_ls", c);
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s problem.txt out.bin [--threads N] [--no-solve] "
                    "[--max-iters N] [--tol X] [--sweeps N] [--balance] "
                    "[--stabilize FACTOR TYPE] [--krylov] [--linear-source] "
                    "[--dump-fluxes F] [--restart-from F] [--residual-reference F] "
                    "[--load-fluxes F] [--fixed-source-cell PIN CELL G V] "
                    "[--fixed-source-fsr R G V] [--fixed-source-material M G V] "
                    "[--sequence STEPS]\n", argv[0]);
    return 2;
  }
  int threads = 1;
  bool solve = true;
  int max_iters_override = -1;
  double tol_override = -1;
  int timed_sweeps = 0;
  bool balance = false, krylov = false, linear_source = false;
  double stab_factor = 0.0;
  int stab_type = -1;
  /* flux files (Solver::dumpFSRFluxes / loadInitialFSRFluxes /
   * setResidualByReference / loadFSRFluxes, src/Solver.cpp:2056-2318) */
  const char* dump_fluxes = nullptr;
  const char* restart_from = nullptr;
  const char* residual_reference = nullptr;
  const char* load_fluxes = nullptr;
  int restart_after = 0;  /* --restart-after N: N iterations, then setRestartStatus(true) */
  bool adjoint = false;   /* --adjoint: setSolverMode(ADJOINT) */
  int chi_material = -1;  /* --chi-spectrum M: setChiSpectrumMaterial(material M) */
  /* Fixed sources (Solver::setFixedSourceByCell / ByFSR / ByMaterial,
   * src/Solver.cpp:479-543), applied in the order given.  A cell is named by
   * its pin universe and its index in that pin (cylinders innermost first,
   * then the outer cell). */
  struct FixedSrc { char kind; int a, b, group; double value; };
  std::vector<FixedSrc> fixed_sources;
  /* --sequence STEPS: comma-separated solves run in order instead of the one
   * computeEigenvalue: "eig", "flux" (computeFlux(max_iters, true)),
   * "flux0" (computeFlux(max_iters, false)), "source:K" or "source:K:RT"
   * (computeSource(max_iters, K, res_type or RT)).  Step i writes step<i>_keff,
   * step<i>_num_iterations and step<i>_flux; the last step also writes the
   * usual keff / num_iterations / flux records. */
  const char* sequence = nullptr;
  /* --dump-segments DIR: TrackGenerator::dumpSegmentsToFile (2D) into
   * DIR/tracks/ under the reference's own file name (set_output_directory +
   * initializeTrackFileDirectory, src/TrackGenerator.cpp:1326-1346, 1388) */
  const char* dump_segments_dir = nullptr;
  for (int a = 3; a < argc; a++) {
    if (!strcmp(argv[a], "--threads")) threads = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--no-solve")) solve = false;
    else if (!strcmp(argv[a], "--max-iters")) max_iters_override = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--tol")) tol_override = atof(argv[++a]);
    else if (!strcmp(argv[a], "--sweeps")) timed_sweeps = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--balance")) balance = true;
    else if (!strcmp(argv[a], "--stabilize")) {
      stab_factor = atof(argv[++a]);
      stab_type = atoi(argv[++a]);
    }
    else if (!strcmp(argv[a], "--krylov")) krylov = true;
    else if (!strcmp(argv[a], "--linear-source")) linear_source = true;
    else if (!strcmp(argv[a], "--dump-fluxes")) dump_fluxes = argv[++a];
    else if (!strcmp(argv[a], "--restart-from")) restart_from = argv[++a];
    else if (!strcmp(argv[a], "--residual-reference")) residual_reference = argv[++a];
    else if (!strcmp(argv[a], "--load-fluxes")) load_fluxes = argv[++a];
    else if (!strcmp(argv[a], "--restart-after")) restart_after = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--adjoint")) adjoint = true;
    else if (!strcmp(argv[a], "--chi-spectrum")) chi_material = atoi(argv[++a]);
    else if (!strcmp(argv[a], "--fixed-source-cell")) {
      FixedSrc f{'c', atoi(argv[a + 1]), atoi(argv[a + 2]), atoi(argv[a + 3]), atof(argv[a + 4])};
      fixed_sources.push_back(f);
      a += 4;
    }
    else if (!strcmp(argv[a], "--fixed-source-fsr")) {
      FixedSrc f{'r', atoi(argv[a + 1]), 0, atoi(argv[a + 2]), atof(argv[a + 3])};
      fixed_sources.push_back(f);
      a += 3;
    }
    else if (!strcmp(argv[a], "--fixed-source-material")) {
      FixedSrc f{'m', atoi(argv[a + 1]), 0, atoi(argv[a + 2]), atof(argv[a + 3])};
      fixed_sources.push_back(f);
      a += 3;
    }
    else if (!strcmp(argv[a], "--sequence")) sequence = argv[++a];
    else if (!strcmp(argv[a], "--dump-segments")) dump_segments_dir = argv[++a];
  }
  Spec s = parse(argv[1]);
  if (max_iters_override > 0) s.max_iters = max_iters_override;
  if (tol_override > 0) s.tol = tol_override;
  set_log_level("WARNING");

  const int G = s.groups;
  std::vector<Material*> mats;
  for (size_t m = 0; m < s.sigma_t.size(); m++) {
    Material* mat = new Material(0, s.mat_names[m].c_str());
    mat->setNumEnergyGroups(G);
    mat->setSigmaT(s.sigma_t[m].data(), G);
    mat->setSigmaS(s.sigma_s[m].data(), G * G);
    mat->setNuSigmaF(s.nu_sigma_f[m].data(), G);
    mat->setChi(s.chi[m].data(), G);
    mats.push_back(mat);
  }

  std::vector<Universe*> pin_univ;
  std::vector<std::vector<Cell*>> pin_cells;  /* cylinders innermost first, then the outer cell */
  for (const Pin& p : s.pins) {
    Universe* u = new Universe();
    pin_cells.emplace_back();
    std::vector<ZCylinder*> zc;
    for (const Cyl& c : p.cyl) zc.push_back(new ZCylinder(0.0, 0.0, c.r));
    for (size_t i = 0; i < p.cyl.size(); i++) {
      Cell* c = new Cell();
      c->setFill(mats[p.cyl[i].mat]);
      c->addSurface(-1, zc[i]);
      if (i > 0) c->addSurface(+1, zc[i - 1]);
      if (p.cyl[i].rings > 1) c->setNumRings(p.cyl[i].rings);
      if (p.cyl[i].sectors > 1) c->setNumSectors(p.cyl[i].sectors);
      u->addCell(c);
      pin_cells.back().push_back(c);
    }
    Cell* outer = new Cell();
    outer->setFill(mats[p.outer_mat]);
    if (!zc.empty()) outer->addSurface(+1, zc.back());
    if (p.outer_sectors > 1) outer->setNumSectors(p.outer_sectors);
    u->addCell(outer);
    pin_cells.back().push_back(outer);
    pin_univ.push_back(u);
  }

  if (s.height > 0) {
    /* Axially extruded 3D model: the lattice repeated over `layers` equal z
     * levels inside a root cell bounded by X/Y/Z planes; TrackGenerator3D
     * with the Gauss-Legendre polar set and EXPLICIT_3D segmentation
     * (the construction of profile/models/pin-cell/pin-cell-3d.cpp). */
    const int L = s.layers;
    Lattice* lat3 = new Lattice();
    lat3->setWidth(s.px, s.py, s.height / L);
    std::vector<Universe*> cu((size_t)L * s.nx * s.ny);
    for (int l = 0; l < L; l++)
      for (int k = 0; k < s.nx * s.ny; k++) cu[(size_t)l * s.nx * s.ny + k] = pin_univ[s.cells[k]];
    lat3->setUniverses(L, s.ny, s.nx, cu.data());
    double hx = 0.5 * s.nx * s.px, hy = 0.5 * s.ny * s.py, hz = 0.5 * s.height;
    XPlane* xmin = new XPlane(-hx);
    XPlane* xmax = new XPlane(+hx);
    YPlane* ymin = new YPlane(-hy);
    YPlane* ymax = new YPlane(+hy);
    ZPlane* zmin = new ZPlane(-hz);
    ZPlane* zmax = new ZPlane(+hz);
    xmin->setBoundaryType(bc_of(s.bc[0]));
    xmax->setBoundaryType(bc_of(s.bc[1]));
    ymin->setBoundaryType(bc_of(s.bc[2]));
    ymax->setBoundaryType(bc_of(s.bc[3]));
    zmin->setBoundaryType(bc_of(s.bcz[0]));
    zmax->setBoundaryType(bc_of(s.bcz[1]));
    Cell* root_cell = new Cell();
    root_cell->addSurface(+1, xmin);
    root_cell->addSurface(-1, xmax);
    root_cell->addSurface(+1, ymin);
    root_cell->addSurface(-1, ymax);
    root_cell->addSurface(+1, zmin);
    root_cell->addSurface(-1, zmax);
    root_cell->setFill(lat3);
    Universe* root = new Universe();
    root->addCell(root_cell);
    Geometry* geometry = new Geometry();
    geometry->setRootUniverse(root);
    geometry->initializeFlatSourceRegions();
    TrackGenerator3D* tg = new TrackGenerator3D(geometry, s.num_azim, s.num_polar, s.spacing,
                                                s.z_spacing);
    Quadrature* quad = new GLPolarQuad();
    quad->setNumAzimAngles(s.num_azim);
    quad->setNumPolarAngles(s.num_polar);
    tg->setQuadrature(quad);
    tg->setSegmentFormation(EXPLICIT_3D);
    tg->setNumThreads(1);
    tg->generateTracks();

    Writer w(argv[2]);
    w.l1("num_groups", G);
    w.l1("num_azim", s.num_azim);
    w.l1("num_polar", s.num_polar);
    const int na2 = s.num_azim / 2, np = s.num_polar;
    {
      std::vector<double> phi(na2), spc(na2), aw(na2), th(na2 * np), tw(na2 * np), ps(na2 * np),
          pw(na2 * np);
      for (int a = 0; a < na2; a++) {
        phi[a] = quad->getPhi(a);
        spc[a] = quad->getAzimSpacing(a);
        aw[a] = quad->getAzimWeight(a);
        for (int p = 0; p < np; p++) {
          th[a * np + p] = quad->getTheta(a, p);
          tw[a * np + p] = quad->getWeight(a, p);
          ps[a * np + p] = quad->getPolarSpacing(a, p);
          pw[a * np + p] = quad->getPolarWeight(a, p);
        }
      }
      w.d("quad_phi", phi);
      w.d("quad_azim_spacing", spc);
      w.d("quad_azim_weight", aw);
      w.d("quad_theta", th);
      w.d("quad_total_weight", tw);
      w.d("quad_polar_spacing", ps);
      w.d("quad_polar_weight", pw);
    }
    /* 3D tracks in uid order (stacks by azim, xy, polar; z in the stack) */
    Track3D**** t3 = tg->get3DTracks();
    int*** tps = tg->getTracksPerStack();
    std::vector<int32_t> t_azim, t_xy, t_pol, t_nseg, t_nff, t_nbf, t_bcf, t_bcb;
    std::vector<int64_t> t_uid, t_nf, t_nb, s_fsr;
    std::vector<double> t_start, t_end, s_len;
    for (int a = 0; a < na2; a++)
      for (int i = 0; i < tg->getNumX(a) + tg->getNumY(a); i++)
        for (int p = 0; p < np; p++)
          for (int z = 0; z < tps[a][i][p]; z++) {
            Track3D* tr = &t3[a][i][p][z];
            t_uid.push_back(tr->getUid());
            t_azim.push_back(tr->getAzimIndex());
            t_xy.push_back(tr->getXYIndex());
            t_pol.push_back(tr->getPolarIndex());
            t_start.push_back(tr->getStart()->getX());
            t_start.push_back(tr->getStart()->getY());
            t_start.push_back(tr->getStart()->getZ());
            t_end.push_back(tr->getEnd()->getX());
            t_end.push_back(tr->getEnd()->getY());
            t_end.push_back(tr->getEnd()->getZ());
            t_nf.push_back(tr->getTrackNextFwd());
            t_nb.push_back(tr->getTrackNextBwd());
            t_nff.push_back(tr->getNextFwdFwd());
            t_nbf.push_back(tr->getNextBwdFwd());
            t_bcf.push_back(tr->getBCFwd());
            t_bcb.push_back(tr->getBCBwd());
            t_nseg.push_back(tr->getNumSegments());
            segment* seg = tr->getSegments();
            for (int k = 0; k < tr->getNumSegments(); k++) {
              s_len.push_back(seg[k]._length);
              s_fsr.push_back(seg[k]._region_id);
            }
          }
    w.l("trk_uid", t_uid);
    w.i("trk_azim", t_azim);
    w.i("trk_xy", t_xy);
    w.i("trk_polar", t_pol);
    w.d("trk_start", t_start);
    w.d("trk_end", t_end);
    w.l("trk_next_fwd", t_nf);
    w.l("trk_next_bwd", t_nb);
    w.i("trk_next_fwd_is_fwd", t_nff);
    w.i("trk_next_bwd_is_fwd", t_nbf);
    w.i("trk_bc_fwd", t_bcf);
    w.i("trk_bc_bwd", t_bcb);
    w.i("trk_num_segments", t_nseg);
    w.d("seg_length", s_len);
    w.l("seg_fsr", s_fsr);
    const long N = geometry->getNumFSRs();
    w.l1("num_fsrs", N);
    TimedCPUSolver solver(tg);
    solver.setNumThreads(threads);
    solver.setConvergenceThreshold(s.tol);
    if (solve) {
      residualType rt = FISSION_SOURCE;
      if (s.res_type == "SCALAR_FLUX") rt = SCALAR_FLUX;
      else if (s.res_type == "TOTAL_SOURCE") rt = TOTAL_SOURCE;
      if (restart_from) solver.loadInitialFSRFluxes(restart_from);
      if (residual_reference) solver.setResidualByReference(residual_reference);
      auto t0 = std::chrono::steady_clock::now();
      solver.computeEigenvalue(s.max_iters, rt);
      auto t1 = std::chrono::steady_clock::now();
      if (dump_fluxes) solver.dumpFSRFluxes(dump_fluxes);
      std::vector<double> flux(N * G);
      for (long r = 0; r < N; r++)
        for (int g = 0; g < G; g++) flux[r * G + g] = solver.getFlux(r, g + 1);
      w.d1("keff", solver.getKeff());
      w.l1("num_iterations", solver.getNumIterations());
      w.d("flux", flux);
      w.d1("time_sweep", solver.split("Transport Sweep"));
      w.d1("time_total", std::chrono::duration<double>(t1 - t0).count());
      long S3 = (long)s_len.size();
      w.l1("num_segments", S3);
      w.l1("num_integrations", 2L * G * S3 * solver.getNumIterations());
      fprintf(stdout, "keff %.10f iterations %d sweep_s %.6f total_s %.6f tracks %zu segments %ld\n",
              solver.getKeff(), solver.getNumIterations(), solver.split("Transport Sweep"),
              std::chrono::duration<double>(t1 - t0).count(), t_uid.size(), S3);
      FP_PRECISION* vol = tg->getFSRVolumes();
      std::vector<double> v(N), pt(3 * N);
      std::vector<int32_t> fm(N);
      for (long r = 0; r < N; r++) {
        v[r] = vol[r];
        Point* q = geometry->getFSRPoint(r);
        pt[3 * r] = q->getX();
        pt[3 * r + 1] = q->getY();
        pt[3 * r + 2] = q->getZ();
        Material* m = geometry->findFSRMaterial(r);
        fm[r] = -1;
        for (size_t k = 0; k < mats.size(); k++) if (mats[k] == m) fm[r] = (int)k;
      }
      w.d("fsr_volume", v);
      w.d("fsr_point", pt);
      w.i("fsr_material", fm);
      if (timed_sweeps > 0) {
        double before = solver.split("Transport Sweep");
        for (int k = 0; k < timed_sweeps; k++) solver.sweep_once();
        double dt = solver.split("Transport Sweep") - before;
        w.d1("timed_sweeps_s", dt);
        w.l1("timed_sweeps", timed_sweeps);
        fprintf(stdout, "timed_sweeps %d seconds %.6f\n", timed_sweeps, dt);
      }
    }
    return 0;
  }

  Lattice* lat = new Lattice();
  lat->setWidth(s.px, s.py);
  std::vector<Universe*> cell_univ(s.nx * s.ny);
  for (int k = 0; k < s.nx * s.ny; k++) cell_univ[k] = pin_univ[s.cells[k]];
  lat->setUniverses(1, s.ny, s.nx, cell_univ.data());

  double hx = 0.5 * s.nx * s.px, hy = 0.5 * s.ny * s.py;
  XPlane* xmin = new XPlane(-hx);
  XPlane* xmax = new XPlane(+hx);
  YPlane* ymin = new YPlane(-hy);
  YPlane* ymax = new YPlane(+hy);
  xmin->setBoundaryType(bc_of(s.bc[0]));
  xmax->setBoundaryType(bc_of(s.bc[1]));
  ymin->setBoundaryType(bc_of(s.bc[2]));
  ymax->setBoundaryType(bc_of(s.bc[3]));
  Cell* root_cell = new Cell();
  root_cell->addSurface(+1, xmin);
  root_cell->addSurface(-1, xmax);
  root_cell->addSurface(+1, ymin);
  root_cell->addSurface(-1, ymax);
  root_cell->setFill(lat);
  Universe* root = new Universe();
  root->addCell(root_cell);

  Geometry* geometry = new Geometry();
  geometry->setRootUniverse(root);
  geometry->initializeFlatSourceRegions();

  TrackGenerator* tg = new TrackGenerator(geometry, s.num_azim, s.spacing);
  Quadrature* quad = new TYPolarQuad();
  quad->setNumAzimAngles(s.num_azim);
  quad->setNumPolarAngles(s.num_polar);
  tg->setQuadrature(quad);
  tg->setNumThreads(1);
  tg->generateTracks();
  if (dump_segments_dir) {
    set_output_directory(const_cast<char*>(dump_segments_dir));
    tg->initializeTrackFileDirectory();
    tg->dumpSegmentsToFile();
  }

  Writer w(argv[2]);
  w.l1("num_groups", G);
  w.l1("num_azim", s.num_azim);
  w.l1("num_polar", s.num_polar);
  w.l1("threads", threads);

  /* Quadrature after laydown (effective angles and spacings). */
  {
    int na2 = s.num_azim / 2;
    std::vector<double> phi(na2), spc(na2), aw(na2), tw(na2 * s.num_polar), st(na2 * s.num_polar);
    for (int a = 0; a < na2; a++) {
      phi[a] = quad->getPhi(a);
      spc[a] = quad->getAzimSpacing(a);
      aw[a] = quad->getAzimWeight(a);
      for (int p = 0; p < s.num_polar; p++) {
        tw[a * s.num_polar + p] = quad->getWeight(a, p);
        st[a * s.num_polar + p] = quad->getSinTheta(a, p);
      }
    }
    w.d("quad_phi", phi);
    w.d("quad_azim_spacing", spc);
    w.d("quad_azim_weight", aw);
    w.d("quad_total_weight", tw);
    w.d("quad_sin_theta", st);
  }

  /* Tracks and segments in uid order. */
  long T = tg->getNum2DTracks();
  Track** tracks = tg->get2DTracksArray();
  std::vector<int32_t> t_azim(T), t_xy(T), t_nseg(T), t_nff(T), t_nbf(T), t_bcf(T), t_bcb(T);
  std::vector<int64_t> t_nf(T), t_nb(T), t_uid(T);
  std::vector<double> t_start(2 * T), t_end(2 * T), t_phi(T);
  std::vector<double> s_len;
  std::vector<int64_t> s_fsr;
  std::vector<int32_t> s_mat;
  for (long t = 0; t < T; t++) {
    Track* tr = tracks[t];
    t_uid[t] = tr->getUid();
    t_azim[t] = tr->getAzimIndex();
    t_xy[t] = tr->getXYIndex();
    t_phi[t] = tr->getPhi();
    t_start[2 * t] = tr->getStart()->getX();
    t_start[2 * t + 1] = tr->getStart()->getY();
    t_end[2 * t] = tr->getEnd()->getX();
    t_end[2 * t + 1] = tr->getEnd()->getY();
    t_nf[t] = tr->getTrackNextFwd();
    t_nb[t] = tr->getTrackNextBwd();
    t_nff[t] = tr->getNextFwdFwd();
    t_nbf[t] = tr->getNextBwdFwd();
    t_bcf[t] = tr->getBCFwd();
    t_bcb[t] = tr->getBCBwd();
    t_nseg[t] = tr->getNumSegments();
    segment* seg = tr->getSegments();
    for (int k = 0; k < tr->getNumSegments(); k++) {
      s_len.push_back(seg[k]._length);
      s_fsr.push_back(seg[k]._region_id);
      int mid = -1;
      for (size_t m = 0; m < mats.size(); m++)
        if (mats[m] == seg[k]._material) mid = (int)m;
      s_mat.push_back(mid);
    }
  }
  w.l("trk_uid", t_uid);
  w.i("trk_azim", t_azim);
  w.i("trk_xy", t_xy);
  w.d("trk_phi", t_phi);
  w.d("trk_start", t_start);
  w.d("trk_end", t_end);
  w.l("trk_next_fwd", t_nf);
  w.l("trk_next_bwd", t_nb);
  w.i("trk_next_fwd_is_fwd", t_nff);
  w.i("trk_next_bwd_is_fwd", t_nbf);
  w.i("trk_bc_fwd", t_bcf);
  w.i("trk_bc_bwd", t_bcb);
  w.i("trk_num_segments", t_nseg);
  w.d("seg_length", s_len);
  w.l("seg_fsr", s_fsr);
  w.i("seg_material", s_mat);

  auto body = [&](auto& solver) -> int {
  solver.setNumThreads(threads);
  solver.setConvergenceThreshold(s.tol);
  if (balance) solver.setKeffFromNeutronBalance();
  if (stab_type >= 0) solver.stabilizeTransport(stab_factor, (stabilizationType)stab_type);

  long N = geometry->getNumFSRs();
  w.l1("num_fsrs", N);

  for (const FixedSrc& f : fixed_sources) {
    if (f.kind == 'c') solver.setFixedSourceByCell(pin_cells.at(f.a).at(f.b), f.group, f.value);
    else if (f.kind == 'r') solver.setFixedSourceByFSR(f.a, f.group, f.value);
    else solver.setFixedSourceByMaterial(mats.at(f.a), f.group, f.value);
  }

  if (solve) {
    residualType rt = FISSION_SOURCE;
    if (s.res_type == "SCALAR_FLUX") rt = SCALAR_FLUX;
    else if (s.res_type == "TOTAL_SOURCE") rt = TOTAL_SOURCE;
    if (restart_from) solver.loadInitialFSRFluxes(restart_from);
    if (residual_reference) solver.setResidualByReference(residual_reference);
    if (adjoint) solver.setSolverMode(ADJOINT);
    if (chi_material >= 0) solver.setChiSpectrumMaterial(mats.at(chi_material));
    if (restart_after > 0) {
      solver.computeEigenvalue(restart_after, rt);
      solver.setRestartStatus(true);
    }
    auto t0 = std::chrono::steady_clock::now();
    if (sequence) {
      std::string steps(sequence);
      size_t pos = 0;
      int i = 0;
      while (pos <= steps.size()) {
        size_t end = steps.find(',', pos);
        if (end == std::string::npos) end = steps.size();
        const std::string st = steps.substr(pos, end - pos);
        if (st == "eig") solver.computeEigenvalue(s.max_iters, rt);
        else if (st == "flux") solver.computeFlux(s.max_iters, true);
        else if (st == "flux0") solver.computeFlux(s.max_iters, false);
        else if (st.rfind("source:", 0) == 0) {
          /* source:K or source:K:RESIDUAL_TYPE */
          residualType srt = rt;
          const size_t c2 = st.find(':', 7);
          if (c2 != std::string::npos) {
            const std::string r = st.substr(c2 + 1);
            srt = r == "SCALAR_FLUX" ? SCALAR_FLUX : r == "TOTAL_SOURCE" ? TOTAL_SOURCE : FISSION_SOURCE;
          }
          solver.computeSource(s.max_iters, atof(st.substr(7, c2 == std::string::npos ? std::string::npos : c2 - 7).c_str()), srt);
        }
        else { fprintf(stderr, "unknown sequence step %s\n", st.c_str()); exit(2); }
        std::vector<double> fl(N * G);
        for (long r = 0; r < N; r++)
          for (int g = 0; g < G; g++) fl[r * G + g] = solver.getFlux(r, g + 1);
        const std::string pre = "step" + std::to_string(i);
        w.d1((pre + "_keff").c_str(), solver.getKeff());
        w.l1((pre + "_num_iterations").c_str(), solver.getNumIterations());
        w.d((pre + "_flux").c_str(), fl);
        pos = end + 1;
        i++;
      }
    } else {
      solver.computeEigenvalue(s.max_iters, rt);
    }
    auto t1 = std::chrono::steady_clock::now();
    if (dump_fluxes) solver.dumpFSRFluxes(dump_fluxes);
    std::vector<double> flux(N * G);
    for (long r = 0; r < N; r++)
      for (int g = 0; g < G; g++) flux[r * G + g] = solver.getFlux(r, g + 1);
    w.d1("keff", solver.getKeff());
    w.l1("num_iterations", solver.getNumIterations());
    w.d("flux", flux);
    w.d1("time_sweep", solver.split("Transport Sweep"));
    w.d1("time_total", std::chrono::duration<double>(t1 - t0).count());
    long S = tg->getNumSegments();
    w.l1("num_segments", S);
    w.l1("num_integrations", 2L * G * (s.num_polar / 2) * S * solver.getNumIterations());
    fprintf(stdout, "keff %.10f iterations %d sweep_s %.6f total_s %.6f\n",
            solver.getKeff(), solver.getNumIterations(),
            solver.split("Transport Sweep"),
            std::chrono::duration<double>(t1 - t0).count());
  }

  /* FSR data exists once the solver initialised the FSRs (volumes and
   * centroids are computed in Solver::initializeFSRs, src/Solver.cpp:820-873). */
  if (solve) {
    FP_PRECISION* vol = tg->getFSRVolumes();
    std::vector<double> v(N), cen(2 * N), pt(2 * N);
    std::vector<int32_t> fm(N);
    for (long r = 0; r < N; r++) {
      v[r] = vol[r];
      Point* c = geometry->getFSRCentroid(r);
      cen[2 * r] = c->getX();
      cen[2 * r + 1] = c->getY();
      Point* p = geometry->getFSRPoint(r);
      pt[2 * r] = p->getX();
      pt[2 * r + 1] = p->getY();
      Material* m = geometry->findFSRMaterial(r);
      fm[r] = -1;
      for (size_t k = 0; k < mats.size(); k++) if (mats[k] == m) fm[r] = (int)k;
    }
    w.d("fsr_volume", v);
    w.d("fsr_centroid", cen);
    w.d("fsr_point", pt);
    w.i("fsr_material", fm);
  }

  /* Optional: the fluxes Solver::loadFSRFluxes reads from a flux file onto
   * this geometry's FSRs (k_eff not assigned, tolerance 0.01 cm). */
  if (solve && load_fluxes) {
    solver.loadFSRFluxes(load_fluxes, false, 0.01);
    std::vector<double> lf(N * G);
    for (long r = 0; r < N; r++)
      for (int g = 0; g < G; g++) lf[r * G + g] = solver.getFlux(r, g + 1);
    w.d("loaded_flux", lf);
  }

  /* Optional: Krylov operator sweeps on the converged flux (the pattern of
   * openmoc/krylov.py: initializeSolver zeroes the angular fluxes, then
   * setFluxes + fission/scatterTransportSweep + getFluxes). */
  if (solve && krylov) {
    std::vector<double> phi0(N * G);
    for (long r = 0; r < N; r++)
      for (int g = 0; g < G; g++) phi0[r * G + g] = solver.getFlux(r, g + 1);
    solver.initializeSolver(FORWARD);
    std::vector<FP_PRECISION> b1(phi0.begin(), phi0.end()), b2(b1), b3(b1), out(N * G);
    auto dump = [&](const char* name) {
      std::vector<double> v(out.begin(), out.end());
      w.d(name, v);
    };
    w.d("krylov_phi0", phi0);
    solver.setFluxes(b1.data(), (int)(N * G));
    solver.fissionTransportSweep();
    solver.getFluxes(out.data(), (int)(N * G));
    dump("krylov_fission_1");
    solver.setFluxes(b2.data(), (int)(N * G));
    solver.scatterTransportSweep();
    solver.getFluxes(out.data(), (int)(N * G));
    dump("krylov_scatter");
    solver.setFluxes(b3.data(), (int)(N * G));
    solver.fissionTransportSweep();
    solver.getFluxes(out.data(), (int)(N * G));
    dump("krylov_fission_2");
  }

  /* Optional: time a fixed number of extra sweeps on the converged state
   * (CPU-baseline sample); integrations per sweep = 2*G*P/2*S. */
  if (solve && timed_sweeps > 0) {
    double before = solver.split("Transport Sweep");
    for (int k = 0; k < timed_sweeps; k++) solver.sweep_once();
    double dt = solver.split("Transport Sweep") - before;
    w.d1("timed_sweeps_s", dt);
    w.l1("timed_sweeps", timed_sweeps);
    fprintf(stdout, "timed_sweeps %d seconds %.6f\n", timed_sweeps, dt);
  }
  if (solve && linear_source) solver.dump_ls(w);
  return 0;
  };
  if (linear_source) {
    TimedCPULSSolver solver(tg);
    return body(solver);
  }
  TimedCPUSolver solver(tg);
  return body(solver);
}
