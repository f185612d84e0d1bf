// mfsv_drive.cpp -- command-line driver of the C++ mirror, in the shape of
// the reference's main (mfcm.cpp:19-49): build a RungeKuttaSolver over a
// SaintVenantSolver, call solve(), and write the outputs.
//
//   mfsv_drive [options] <config-file>
//
// reads a configuration in the reference's mf format (GlobalConfig::init,
// Config.cpp:102-174: the simulation base path is --base-path or the
// file's directory) and runs it as mfcm would, writing the configured
// output files, check files and measures under the base path.  Options:
//   --base-path P        simulation base path
//   --accel-device N     device index (also "device" in the device section)
//   --precision f64|f32  value precision (default f64, the plain method)
//   --host-loop          one host round trip per timestep (the reference's
//                        loop shape) instead of each outer step's inner loop
//                        on the device
//   --eager              launch kernels one by one instead of graph replay
//   --flags N            MFSV_E* extension bits (default 0)
//   --quiet              no progress lines
//   --dump-state DIR     also write the final h, u, v (raw float64) and
//                        stats.txt to DIR, as the array mode below does (with
//                        the device loop also volume0 and source_volume, the
//                        terms of the FP64 mass balance)
// and prints "solved: N timesteps, M repeated" at the end.
//
//   mfsv_drive <input-dir> <output-dir>
//
// <input-dir> holds scenario.txt ("key = value" lines, '!' starts a comment
// as in the reference's mf files) and raw little-endian float64 arrays of
// nx*ny values, row-major: zb.f64, h.f64, u.f64, v.f64 and optionally
// manning_n.f64.  Keys: nx ny dx dy precision (f64|f32) method (a named RK
// set, default heun2) start end step initial_dt flags cn_target n_shallow
// n_deep d_shallow d_deep head_cols h_dry u_max tile_rows rain_t rain_v
// (comma-separated series, m/s) storm_peak storm_sigma storm_cx0 storm_cy0
// storm_vx storm_vy verbose.
// <output-dir> receives h.f64, u.f64, v.f64 and stats.txt (timesteps,
// repeated timesteps, volume, precision).
//
// The reference reads its inputs through a Boost-based config system and
// field generators; this driver reads arrays instead (the fields of the
// Saint-Venant path are what it exercises).  Errors print a message and
// exit with status 1, as an uncaught exception would end the reference.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "ConfiguredModel.hpp"
#include "RungeKutta.hpp"
#include "SaintVenantSolver.hpp"

using namespace mfsv;

namespace {

std::map<std::string, std::string> read_keys(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::map<std::string, std::string> kv;
  std::string line;
  while (std::getline(f, line)) {
    const size_t bang = line.find('!');
    if (bang != std::string::npos) line.resize(bang);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    auto trim = [](std::string s) {
      const size_t a = s.find_first_not_of(" \t\r");
      const size_t b = s.find_last_not_of(" \t\r");
      return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    };
    kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  return kv;
}

std::vector<double> read_f64(const std::string& path, size_t n, bool required = true) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    if (required) throw std::runtime_error("cannot open " + path);
    return {};
  }
  std::vector<double> v(n);
  f.read(reinterpret_cast<char*>(v.data()), std::streamsize(n * sizeof(double)));
  if (size_t(f.gcount()) != n * sizeof(double))
    throw std::runtime_error(path + ": expected " + std::to_string(n) + " float64 values");
  return v;
}

void write_f64(const std::string& path, const double* p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(p), std::streamsize(n * sizeof(double)));
  if (!f) throw std::runtime_error("cannot write " + path);
}

std::vector<double> parse_list(const std::string& s) {
  std::vector<double> v;
  std::stringstream ss(s);
  std::string item;
  while (std::getline(ss, item, ',')) v.push_back(std::stod(item));
  return v;
}

template <typename Solver>
int run(const std::map<std::string, std::string>& kv, const std::string& in, const std::string& out,
        int precision) {
  auto num = [&](const char* k, double d) { return kv.count(k) ? std::stod(kv.at(k)) : d; };
  const int64_t nx = int64_t(num("nx", 0)), ny = int64_t(num("ny", 0));
  const Cartesian2DMesh mesh(nx, ny, num("dx", 1.0), num("dy", num("dx", 1.0)));
  const size_t n = mesh.cell_count();
  auto tparams = std::make_shared<TimeParameters<double>>(num("start", 0.0), num("end", 3600.0),
                                                          num("step", 60.0), num("initial_dt", 1.0));
  SaintVenantOptions opt;
  opt.precision = precision;
  opt.flags = int32_t(num("flags", 0));
  opt.cn_target = num("cn_target", 1.0);
  opt.n_shallow = num("n_shallow", 0.03);
  opt.n_deep = num("n_deep", opt.n_shallow);
  opt.d_shallow = num("d_shallow", 0.1);
  opt.d_deep = num("d_deep", 0.3);
  opt.head_cols = int64_t(num("head_cols", 0));
  opt.h_dry = num("h_dry", 1e-3);
  opt.u_max = num("u_max", 10.0);
  opt.tile_rows = int(num("tile_rows", 0));
  opt.manning_n = read_f64(in + "/manning_n.f64", n, false);
  if (kv.count("rain_t")) opt.rain = TimeSeries(parse_list(kv.at("rain_t")), parse_list(kv.at("rain_v")));
  opt.storm.storm_peak = num("storm_peak", 0.0);
  opt.storm.storm_sigma = num("storm_sigma", 1.0);
  opt.storm.storm_cx0 = num("storm_cx0", 0.0);
  opt.storm.storm_cy0 = num("storm_cy0", 0.0);
  opt.storm.storm_vx = num("storm_vx", 0.0);
  opt.storm.storm_vy = num("storm_vy", 0.0);
  const std::string method = kv.count("method") ? kv.at("method") : "heun2";

  auto scheme = std::make_shared<RungeKuttaSolver<Solver>>(
      tparams, method, mesh, read_f64(in + "/zb.f64", n), read_f64(in + "/h.f64", n),
      read_f64(in + "/u.f64", n), read_f64(in + "/v.f64", n), opt);
  scheme->set_verbose(num("verbose", 0) != 0);
  std::shared_ptr<TemporalScheme> base = scheme;
  base->solve();

  auto& U = scheme->solver()->state();
  write_f64(out + "/h.f64", U.h.data().data(), n);
  write_f64(out + "/u.f64", U.u.data().data(), n);
  write_f64(out + "/v.f64", U.v.data().data(), n);
  std::ofstream st(out + "/stats.txt");
  st.precision(17);
  st << "timesteps = " << scheme->total_timesteps() << "\n"
     << "repeated_timesteps = " << scheme->total_repeated_timesteps() << "\n"
     << "volume = " << scheme->solver()->volume() << "\n"
     << "precision = " << (precision == MFSV_F64 ? "f64" : "f32") << "\n";
  if (!st) throw std::runtime_error("cannot write stats.txt");
  std::cout << "solved: " << scheme->total_timesteps() << " timesteps, "
            << scheme->total_repeated_timesteps() << " repeated" << std::endl;
  return 0;
}

template <typename T>
int run_config(GlobalConfig& gc, const RunOptions& ro, const std::string& dump) {
  ConfiguredModel<T> model(gc, ro);
  model.solve();
  auto& s = model.scheme();
  if (!dump.empty()) {
    auto& U = model.solver().state();
    const size_t n = model.mesh().cell_count();
    write_f64(dump + "/h.f64", U.h.data().data(), n);
    write_f64(dump + "/u.f64", U.u.data().data(), n);
    write_f64(dump + "/v.f64", U.v.data().data(), n);
    std::ofstream st(dump + "/stats.txt");
    st.precision(17);
    st << "timesteps = " << s.total_timesteps() << "\n"
       << "repeated_timesteps = " << s.total_repeated_timesteps() << "\n"
       << "volume = " << model.solver().volume() << "\n"
       << "precision = " << (ro.precision == MFSV_F64 ? "f64" : "f32") << "\n";
    if (model.solver().device_loop()) {  // the device controller's mass balance
      const mfsv_stats ds = model.solver().device_stats();
      st << "volume0 = " << ds.volume0 << "\n"
         << "source_volume = " << ds.source_volume << "\n";
    }
    if (!st) throw std::runtime_error("cannot write stats.txt");
  }
  std::cout << "solved: " << s.total_timesteps() << " timesteps, " << s.total_repeated_timesteps()
            << " repeated" << std::endl;
  return 0;
}

void usage(const char* prog) {
  std::cerr << "usage: " << prog << " [--base-path P] [--accel-device N] [--precision f64|f32]"
            << " [--host-loop] [--eager] [--flags N] [--quiet] [--dump-state DIR] <config-file>\n"
            << "       " << prog << " <input-dir> <output-dir>" << std::endl;
}

int main_config(int argc, char* argv[]) {
  std::string base, file, prec = "f64", dump;
  RunOptions ro;
  bool device_given = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) throw std::runtime_error("option " + a + " needs a value");
      return argv[++i];
    };
    if (a == "--base-path") base = next();
    else if (a == "--accel-device") { ro.device = std::stoi(next()); device_given = true; }
    else if (a == "--accel-platform") (void)next();
    else if (a == "--precision") prec = next();
    else if (a == "--host-loop") ro.device_loop = false;
    else if (a == "--eager") ro.graph = 0;
    else if (a == "--flags") ro.flags = int32_t(std::stol(next(), nullptr, 0));
    else if (a == "--quiet") ro.quiet = true;
    else if (a == "--dump-state") dump = next();
    else if (a == "--help") { usage(argv[0]); return 2; }
    else if (!a.empty() && a[0] == '-') throw std::runtime_error("unknown option " + a);
    else if (file.empty()) file = a;
    else throw std::runtime_error("Multiple input files not yet supported.");
  }
  if (file.empty()) throw std::runtime_error("No input configuration specified.");
  GlobalConfig gc = GlobalConfig::from_file(file, base);
  if (!device_given) {
    if (const Config* d = gc.root().find("device")) ro.device = d->get<int>("device", 0);
  }
  if (prec == "f64") {
    ro.precision = MFSV_F64;
    return run_config<double>(gc, ro, dump);
  }
  if (prec == "f32") {
    ro.precision = MFSV_F32;
    return run_config<float>(gc, ro, dump);
  }
  throw std::runtime_error("precision must be f64 or f32");
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 2) {
    usage(argv[0]);
    return 2;
  }
  // two plain arguments: the array mode; otherwise a configuration file
  if (!(argc == 3 && argv[1][0] != '-' && argv[2][0] != '-')) {
    try {
      return main_config(argc, argv);
    } catch (const std::exception& e) {
      std::cerr << "mfsv_drive: " << e.what() << std::endl;
      return 1;
    }
  }
  try {
    const std::string in = argv[1], out = argv[2];
    const auto kv = read_keys(in + "/scenario.txt");
    const std::string prec = kv.count("precision") ? kv.at("precision") : "f64";
    if (prec == "f64") return run<SaintVenantSolver<double, double>>(kv, in, out, MFSV_F64);
    if (prec == "f32") return run<SaintVenantSolver<double, float>>(kv, in, out, MFSV_F32);
    throw std::runtime_error("precision must be f64 or f32");
  } catch (const std::exception& e) {
    std::cerr << "mfsv_drive: " << e.what() << std::endl;
    return 1;
  }
}
