// sayram_gpu_run -- the reference's driver (source/main.cc) on the gfx950 step.
//
//   sayram_gpu_run [--case albert_young|albert_young_lc] [--dump-case <file>] [<p.ini>]
//   sayram_gpu_run [--dump-case <file>] <p.ini> <problem.bin>
// Solver options: --precision 0|1|2 (sg_precision), --rtol x, --max-iters n.
//
// With no parameter file it reads p.ini of the working directory
// (source/Parameters.cc:20-23), and it copies the parameter file to
// ./output/<run_id>/<run_id>.ini (source/Parameters.cc:10-16).
//
// Without problem.bin the Equation is the reference's Albert & Young chorus
// case (source/main.cc:54 constructs Albert_Young; albert_young_lc is the
// loss-cone variant of source/Cases/Albert_Young_LC.cc), built from the
// diffusion table D/<dID>.h5 of the working directory as
// source/Cases/Albert_Young_IO.cc:19 opens it ([diffusion_coefficients] dID).
// --dump-case writes the Equation's inputs in problem.bin format and exits
// without touching the GPU.
//
// p.ini is read with the reference's keys and derived values
// (source/Parameters.cc:35-65): [basic] run_id, nalpha0, nE, alpha0min,
// alpha0max (deg), Emin, Emax (MeV), T (days), nsteps; [diagnostics] nplots;
// nsteps is rounded down to a multiple of nplots and dt = T / nsteps
// (source/Parameters.h:44). The grid is uniform in (alpha0 [rad], ln E) as in
// make_uniform (source/main.cc:20-37).
//
// problem.bin holds the Equation's per-cell inputs and boundary data for that
// grid (written by sayram_cases.write_problem); an Equation subclass serves
// them through the reference's accessors.
//
// Output, when built with HDF5: ./output/<run_id>/<run_id>_data.h5 with
// /alpha0 (deg), /logEN (ln E - ln E0), /f/<k> (nx, ny) every nsteps/nplots
// steps and /t = linspace(0, T, nplots + 1), as source/main.cc:57-89 writes.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "GpuSolver.h"
#include "albert_young.h"
#include "sayram_host.h"

#ifdef SAYRAM_HDF5
#include <hdf5.h>
#endif

namespace {

const double kPi = 3.141592653589793238462;   // source/common.h:39
const double kE0 = 0.511875;                  // MeV, source/common.h:42

// INI file with mINI's rules as the reference uses them: [sections], key =
// value, ';' comments, lines without '=' ignored, keys case-insensitive.
std::map<std::string, std::map<std::string, std::string>> read_ini(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::map<std::string, std::map<std::string, std::string>> out;
  std::string line, sec;
  auto trim = [](std::string s) {
    const char *ws = " \t\r\n";
    s.erase(0, s.find_first_not_of(ws));
    s.erase(s.find_last_not_of(ws) + 1);
    return s;
  };
  auto lower = [](std::string s) {
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
  };
  while (std::getline(in, line)) {
    line = trim(line);
    if (line.empty() || line[0] == ';') continue;
    if (line[0] == '[') {
      const auto e = line.find(']');
      if (e != std::string::npos) sec = lower(trim(line.substr(1, e - 1)));
      continue;
    }
    const auto eq = line.find('=');
    if (eq == std::string::npos) continue;
    out[sec][lower(trim(line.substr(0, eq)))] = trim(line.substr(eq + 1));
  }
  return out;
}

struct Params {
  std::string run_id;
  std::size_t nalpha0 = 0, nE = 0;
  double a0min = 0, a0max = 0, Emin = 0, Emax = 0, T = 0, nsteps = 0;
  int nplots = 1, save_every = 1;
  std::string dID;
  double dt() const { return T / nsteps; }
};

Params read_params(const std::string &path) {
  auto ini = read_ini(path);
  auto get = [&](const std::string &s, const std::string &k) {
    auto it = ini.find(s);
    if (it == ini.end()) throw std::runtime_error("section not found: " + s);
    auto kv = it->second.find(k);
    if (kv == it->second.end()) throw std::runtime_error("key not found: " + k);
    return kv->second;
  };
  Params p;
  p.run_id = get("basic", "run_id");
  p.nalpha0 = std::stoul(get("basic", "nalpha0"));
  p.nE = std::stoul(get("basic", "ne"));
  p.a0min = std::stod(get("basic", "alpha0min"));
  p.a0max = std::stod(get("basic", "alpha0max"));
  p.Emin = std::stod(get("basic", "emin"));
  p.Emax = std::stod(get("basic", "emax"));
  p.T = std::stod(get("basic", "t"));
  p.nsteps = std::stod(get("basic", "nsteps"));
  p.nplots = std::stoi(get("diagnostics", "nplots"));
  auto dc = ini.find("diffusion_coefficients");
  if (dc != ini.end() && dc->second.count("did")) p.dID = dc->second.at("did");
  // source/Parameters.cc:59-60. With nsteps < nplots the reference rounds
  // nsteps to 0 and dt = T / 0: there is no step to take, so that is an error.
  if (p.nplots < 1) throw std::runtime_error("[diagnostics] nplots must be >= 1");
  p.save_every = (int)(p.nsteps / p.nplots);
  if (p.save_every < 1)
    throw std::runtime_error("nsteps (" + get("basic", "nsteps") + ") is less than nplots (" +
                             std::to_string(p.nplots) +
                             "): nsteps rounds down to a multiple of nplots, here 0 steps");
  p.nsteps = p.save_every * p.nplots;
  return p;
}

Grid2D make_uniform(const Params &p) {   // source/main.cc:20-37
  std::vector<double> xe(p.nalpha0 + 1), ye(p.nE + 1);
  const double lo_x = p.a0min * kPi / 180, hi_x = p.a0max * kPi / 180;
  const double lo_y = std::log(p.Emin), hi_y = std::log(p.Emax);
  const double dx = (hi_x - lo_x) / (double)p.nalpha0, dy = (hi_y - lo_y) / (double)p.nE;
  for (std::size_t i = 0; i <= p.nalpha0; ++i) xe[i] = lo_x + dx * (double)i;
  for (std::size_t j = 0; j <= p.nE; ++j) ye[j] = lo_y + dy * (double)j;
  return Grid2D(std::move(xe), std::move(ye));
}

// Equation served from arrays (problem.bin)
class ArrayCase : public Equation {
 public:
  ArrayCase(const Mesh &m, const std::string &path) : Equation(m), nx_(m.nx()), ny_(m.ny()) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    char magic[8];
    in.read(magic, 8);
    if (std::memcmp(magic, "SGPROB1", 7) != 0) throw std::runtime_error("not a problem file");
    std::int32_t n[2];
    in.read(reinterpret_cast<char *>(n), sizeof(n));
    if ((std::size_t)n[0] != nx_ || (std::size_t)n[1] != ny_)
      throw std::runtime_error("problem file grid does not match the ini file");
    auto rd = [&](std::vector<double> &v, std::size_t k) {
      v.resize(k);
      in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(k * sizeof(double)));
    };
    std::vector<double> tmp;
    Array2D *dst[6] = {&G_, &Dxx_, &Dxy_, &Dyy_, &inv_tau_, &f0_};
    f0_.resize({nx_, ny_});
    for (Array2D *a : dst) {
      rd(tmp, nx_ * ny_);
      std::copy(tmp.begin(), tmp.end(), a->data());
    }
    std::int32_t bt[4];
    in.read(reinterpret_cast<char *>(bt), sizeof(bt));
    for (int s = 0; s < 4; ++s) {
      type_[s] = bt[s] == 0 ? BCType::Dirichlet : BCType::ZeroFlux;
      rd(val_[s], s < 2 ? ny_ + 1 : nx_ + 1);
    }
    if (!in) throw std::runtime_error("truncated problem file");
    // optional: rates of a time-dependent Equation (sayram_cases.time_dependent)
    in.read(magic, 8);
    if (in.gcount() == 8) {
      if (std::memcmp(magic, "SGTIME1", 7) != 0) throw std::runtime_error("bad problem file trailer");
      in.read(reinterpret_cast<char *>(rate_), sizeof(rate_));
      if (!in) throw std::runtime_error("truncated problem file");
      timed_ = true;
      base_[0] = G_; base_[1] = Dxx_; base_[2] = Dxy_; base_[3] = Dyy_; base_[4] = inv_tau_;
    }
  }
  BCType bc_type(BoundaryID side) const override { return type_[(int)side]; }
  double init_f(const Ind &ind) const override { return f0_(ind.i, ind.j); }
  // At time t: G (1 + rG t), D (1 + rD t), 1/tau (1 + rtau t) of the base
  // arrays; the Dirichlet values scale by (1 + rbc t) in dirichlet_vertex_value.
  void update(double t) override {
    if (!timed_) return;
    const double sG = 1.0 + rate_[0] * t, sD = 1.0 + rate_[1] * t, sT = 1.0 + rate_[2] * t;
    Array2D *dst[5] = {&G_, &Dxx_, &Dxy_, &Dyy_, &inv_tau_};
    const double sc[5] = {sG, sD, sD, sD, sT};
    for (int a = 0; a < 5; ++a)
      for (std::size_t q = 0; q < nx_ * ny_; ++q) dst[a]->data()[q] = base_[a].data()[q] * sc[a];
  }
  bool dirichlet_vertex_value(BoundaryID side, std::size_t i, std::size_t j, double t,
                              double *out) const override {
    const int s = (int)side;
    if (type_[s] != BCType::Dirichlet) return false;
    const double v = val_[s][s < 2 ? j : i];
    *out = timed_ ? v * (1.0 + rate_[3] * t) : v;
    return true;
  }

 private:
  std::size_t nx_, ny_;
  Array2D f0_;
  BCType type_[4];
  std::vector<double> val_[4];
  bool timed_ = false;
  double rate_[4] = {0, 0, 0, 0};
  Array2D base_[5];
};

#ifdef SAYRAM_HDF5
std::vector<double> h5_read(hid_t file, const char *name, std::vector<hsize_t> &dims) {
  hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
  if (ds < 0) throw std::runtime_error(std::string("HDF5: no dataset ") + name);
  hid_t space = H5Dget_space(ds);
  const int rank = H5Sget_simple_extent_ndims(space);
  dims.assign(rank > 0 ? rank : 0, 0);
  if (rank > 0) H5Sget_simple_extent_dims(space, dims.data(), nullptr);
  hsize_t n = 1;
  for (hsize_t d : dims) n *= d;
  std::vector<double> v(n);
  const herr_t rc = H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
  H5Sclose(space);
  H5Dclose(ds);
  if (rc < 0) throw std::runtime_error(std::string("HDF5: cannot read ") + name);
  return v;
}

// D/<dID>.h5 as source/Cases/Albert_Young_IO.cc:17-37 reads it
ay::DTable read_d_table(const std::string &path) {
  hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) throw std::runtime_error("cannot open " + path);
  ay::DTable t;
  std::vector<hsize_t> da, de, d2;
  t.x = h5_read(f, "/alpha0", da);
  for (double &v : t.x) v = v * ay::kPi / 180;    // degrees to radians
  t.y = h5_read(f, "/E", de);
  t.daa = h5_read(f, "/Daa", d2);
  const auto check = [&](const std::vector<hsize_t> &d, const char *nm) {
    if (d.size() != 2 || d[0] != t.x.size() || d[1] != t.y.size())
      throw std::runtime_error(std::string("D table: ") + nm + " is not (n_alpha0, n_E)");
  };
  check(d2, "/Daa");
  t.dap = h5_read(f, "/Dap", d2);
  check(d2, "/Dap");
  t.dpp = h5_read(f, "/Dpp", d2);
  check(d2, "/Dpp");
  H5Fclose(f);
  return t;
}

void h5_write(hid_t file, const std::string &name, const double *v, int rank, const hsize_t *dims) {
  hid_t space = H5Screate_simple(rank, dims, nullptr);
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  H5Pset_create_intermediate_group(lcpl, 1);
  hid_t ds = H5Dcreate2(file, name.c_str(), H5T_NATIVE_DOUBLE, space, lcpl, H5P_DEFAULT, H5P_DEFAULT);
  if (ds < 0) throw std::runtime_error("HDF5: cannot create " + name);
  H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v);
  H5Dclose(ds);
  H5Pclose(lcpl);
  H5Sclose(space);
}
#endif

// the Equation's inputs in problem.bin format (sayram_cases.write_problem)
void dump_case(const Mesh &m, const Equation &eq, const std::string &path) {
  std::ofstream fo(path, std::ios::binary);
  if (!fo) throw std::runtime_error("cannot write " + path);
  const std::size_t nx = m.nx(), ny = m.ny();
  fo.write("SGPROB1", 8);
  const std::int32_t n[2] = {(std::int32_t)nx, (std::int32_t)ny};
  fo.write(reinterpret_cast<const char *>(n), sizeof(n));
  std::vector<double> a(nx * ny);
  auto put = [&](auto get) {
    for (std::size_t i = 0; i < nx; ++i)
      for (std::size_t j = 0; j < ny; ++j) a[i * ny + j] = get(Ind{i, j});
    fo.write(reinterpret_cast<const char *>(a.data()), (std::streamsize)(a.size() * sizeof(double)));
  };
  put([&](const Ind &k) { return eq.G(k); });
  put([&](const Ind &k) { return eq.Dxx(k); });
  put([&](const Ind &k) { return eq.Dxy(k); });
  put([&](const Ind &k) { return eq.Dyy(k); });
  put([&](const Ind &k) { return eq.inv_tau(k); });
  put([&](const Ind &k) { return eq.init_f(k); });
  const BoundaryID sides[4] = {BoundaryID::XMIN, BoundaryID::XMAX, BoundaryID::YMIN, BoundaryID::YMAX};
  std::int32_t bt[4];
  for (int s = 0; s < 4; ++s) bt[s] = eq.bc_type(sides[s]) == BCType::Dirichlet ? 0 : 1;
  fo.write(reinterpret_cast<const char *>(bt), sizeof(bt));
  for (int s = 0; s < 4; ++s) {
    const std::size_t nv = s < 2 ? ny + 1 : nx + 1;
    std::vector<double> v(nv, 0.0);
    for (std::size_t k = 0; k < nv; ++k) {
      double out = 0.0;
      if (bt[s] == 0 && eq.dirichlet_vertex_value(sides[s], s < 2 ? (s == 0 ? 0 : nx) : k,
                                                  s < 2 ? k : (s == 2 ? 0 : ny), 0.0, &out))
        v[k] = out;
    }
    fo.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(nv * sizeof(double)));
  }
}

}  // namespace

int main(int argc, char **argv) {
  std::string case_name = "albert_young", dump;
  std::vector<std::string> pos;
  sg_opts opts{};   // zero: the library defaults
  for (int k = 1; k < argc; ++k) {
    const std::string a = argv[k];
    if (a == "--case" && k + 1 < argc) case_name = argv[++k];
    else if (a == "--dump-case" && k + 1 < argc) dump = argv[++k];
    else if (a == "--precision" && k + 1 < argc) opts.precision = std::atoi(argv[++k]);
    else if (a == "--rtol" && k + 1 < argc) opts.rtol = std::atof(argv[++k]);
    else if (a == "--max-iters" && k + 1 < argc) opts.max_iters = std::atoi(argv[++k]);
    else pos.push_back(a);
  }
  // no parameter file: p.ini of the working directory (source/Parameters.cc:20-23)
  if (pos.empty()) pos.push_back("p.ini");
  if (pos.size() > 2 || (case_name != "albert_young" && case_name != "albert_young_lc")) {
    std::cerr << "usage: sayram_gpu_run [--case albert_young|albert_young_lc] [--dump-case <file>] "
                 "[--precision p] [--rtol x] [--max-iters n] [<p.ini> [<problem.bin>]]"
              << std::endl;
    return 1;
  }
  try {
    const Params paras = read_params(pos[0]);
    // the output directory and a copy of the parameter file in it, as the
    // reference's Parameters constructor makes them (source/Parameters.cc:7-16)
    const std::string out_dir = "./output/" + paras.run_id + "/";
    if (dump.empty()) {
      std::filesystem::create_directories(out_dir);
      std::filesystem::copy_file(pos[0], out_dir + paras.run_id + ".ini",
                                 std::filesystem::copy_options::overwrite_existing);
    }
    Grid2D grid = make_uniform(paras);
    Mesh m(grid, paras.dt());
    std::unique_ptr<Equation> eqp;
    if (pos.size() == 2) {
      eqp.reset(new ArrayCase(m, pos[1]));
    } else {
#ifdef SAYRAM_HDF5
      if (paras.dID.empty()) throw std::runtime_error("[diffusion_coefficients] dID not set");
      const ay::DTable tab = read_d_table("D/" + paras.dID + ".h5");
      eqp.reset(new ay::AlbertYoung(m, tab, std::log(paras.Emin), case_name == "albert_young_lc"));
#else
      throw std::runtime_error("the Albert_Young case reads its HDF5 table: build with HDF5");
#endif
    }
    Equation &eq = *eqp;
    if (!dump.empty()) {
      dump_case(m, eq, dump);
      return 0;
    }
    GpuSolver solver(m, &eq, &opts);
    const std::size_t nx = m.nx(), ny = m.ny();
#ifdef SAYRAM_HDF5
    const std::string fname = out_dir + paras.run_id + "_data.h5";
    hid_t file = H5Fcreate(fname.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) throw std::runtime_error("HDF5: cannot create " + fname);
    std::vector<double> a0(nx), le(ny);
    for (std::size_t i = 0; i < nx; ++i) a0[i] = m.x(i) / kPi * 180.0;
    for (std::size_t j = 0; j < ny; ++j) le[j] = m.y(j) - std::log(kE0);
    const hsize_t dnx[1] = {nx}, dny[1] = {ny}, dnf[2] = {nx, ny};
    h5_write(file, "/alpha0", a0.data(), 1, dnx);
    h5_write(file, "/logEN", le.data(), 1, dny);
    h5_write(file, "/f/0", solver.f().data(), 2, dnf);
#endif
    const std::clock_t c0 = std::clock();
    const auto w0 = std::chrono::steady_clock::now();
    const int nsteps = (int)paras.nsteps;
    for (int tstep = 1; tstep <= nsteps; ++tstep) {
      solver.update();
      if (tstep % paras.save_every == 0) {
#ifdef SAYRAM_HDF5
        h5_write(file, "/f/" + std::to_string(tstep / paras.save_every), solver.f().data(), 2, dnf);
#endif
      }
    }
    (void)solver.f();
    const double wall =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
#ifdef SAYRAM_HDF5
    std::vector<double> t(paras.nplots + 1);
    for (int k = 0; k <= paras.nplots; ++k) t[k] = paras.T * k / paras.nplots;
    const hsize_t dt_[1] = {t.size()};
    h5_write(file, "/t", t.data(), 1, dt_);
    H5Fclose(file);
#endif
    const sg_step_stats st = solver.stats();
    std::cout << "CPU time used " << (double)(std::clock() - c0) / CLOCKS_PER_SEC << " seconds"
              << std::endl;
    std::cout << "wall time " << wall << " s, " << nsteps << " steps, "
              << (double)nx * ny * nsteps / wall << " cell-steps/s, last iterations "
              << st.last_iters << ", min f " << st.min_f << std::endl;
    // the final state, for comparison by the caller
    std::ofstream fo(out_dir + paras.run_id + "_final.bin", std::ios::binary);
    fo.write(reinterpret_cast<const char *>(solver.f().data()),
             (std::streamsize)(nx * ny * sizeof(double)));
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
