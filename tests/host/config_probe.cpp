// config_probe.cpp -- prints what the C++ mirror's input layer makes of a
// configuration, for the CPU tests of mfcm_amd/host/{Config,Input,Output}.hpp.
//
//   config_probe tree <file>                 the parsed tree, one node a line
//   config_probe time <file> <text>          scheme time parser: seconds
//   config_probe tparams <file>              start end step dt0 num_steps
//   config_probe series <file> <name>        times and values of a series
//   config_probe select <file> <path>        ids of the selection at path
//   config_probe field <file> <name> <init>  the generated cell field
//   config_probe mesh <file>                 dx dy and locations
//   config_probe csv <out>                   a small table in CSV form
//   config_probe output <file> <name> <t>    a timed output of h = cell index
//   config_probe checkdefaults <file>        default check/output sections
#include <cstdio>
#include <iostream>

#include "../../mfcm_amd/host/Output.hpp"

using namespace mfsv;

static void dump(const Config& c, int depth) {
  for (const auto& e : c) {
    std::printf("%d|%s|%s\n", depth, e.first.c_str(), e.second.data().c_str());
    dump(e.second, depth + 1);
  }
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string cmd = argv[1];
  try {
    if (cmd == "tree") {
      dump(MfReader::read_file(argv[2]), 0);
      return 0;
    }
    if (cmd == "csv") {
      std::vector<double> a = {1.5, -2.25}, b = {1e-7, 12345.678901234};
      OutputTable t;
      t.names = {"h", "u v"};
      t.columns = {&a, &b};
      t.nrows = 2;
      MeshGeometry m(2, 1, 100.0, 200.0, 2.5, 2.5);
      t.location = [&](size_t i) { return m.cell_location(i); };
      write_csv(argv[2], t);
      return 0;
    }
    GlobalConfig gc = GlobalConfig::from_file(argv[2]);
    if (cmd == "time") {
      TimeParser p(gc.scheme_configuration());
      std::printf("%.17g\n", p.parse(argv[3]));
    } else if (cmd == "tparams") {
      TimeParser p(gc.scheme_configuration());
      auto tp = time_parameters_from(gc.scheme_configuration(), p);
      std::printf("%.17g %.17g %.17g %.17g %zu\n", tp->start_time(), tp->end_time(),
                  tp->step_duration(), tp->initial_timestep(), tp->num_steps());
    } else if (cmd == "series") {
      TimeParser p(gc.scheme_configuration());
      TimeSeriesDatabase db;
      const TimeSeries& ts = db.get(gc, p, argv[3]);
      for (size_t i = 0; i < ts.times().size(); ++i)
        std::printf("%.17g %.17g\n", ts.times()[i], ts.values()[i]);
    } else if (cmd == "select") {
      const MeshGeometry m = MeshGeometry::from_config(gc.mesh_configuration());
      const MeshSelection s = MeshSelection::from_config(m, gc.root().get_child(argv[3]));
      std::printf("%s", s.is_global() ? "global" : "ids");
      for (int64_t i : s.ids()) std::printf(" %lld", (long long)i);
      std::printf("\n");
    } else if (cmd == "field") {
      const MeshGeometry m = MeshGeometry::from_config(gc.mesh_configuration());
      RasterDatabase rdb;
      FieldGenerator fg(gc, rdb, m);
      for (double v : fg(argv[3], std::atof(argv[4]))) std::printf("%.17g\n", v);
    } else if (cmd == "mesh") {
      const MeshGeometry m = MeshGeometry::from_config(gc.mesh_configuration());
      std::printf("%.17g %.17g\n", m.dx(), m.dy());
      for (size_t i = 0; i < m.cell_count(); ++i) {
        auto l = m.cell_location(i);
        std::printf("c %.17g %.17g\n", l[0], l[1]);
      }
      for (size_t i = 0; i < m.face_count(); ++i) {
        auto l = m.face_location(i);
        std::printf("f %.17g %.17g\n", l[0], l[1]);
      }
      for (size_t i = 0; i < m.vertex_count(); ++i) {
        auto l = m.vertex_location(i);
        std::printf("v %.17g %.17g\n", l[0], l[1]);
      }
    } else if (cmd == "output") {
      const MeshGeometry m = MeshGeometry::from_config(gc.mesh_configuration());
      TimeParser p(gc.scheme_configuration());
      auto tp = time_parameters_from(gc.scheme_configuration(), p);
      MultiFieldOutputFile of(gc, argv[3], p, *tp);
      for (int k = 4; k < argc; ++k) {
        const std::string fn = of.timed_output(std::atof(argv[k]), m, [&](const std::string& f, MeshPart part)
                                                   -> std::optional<std::vector<double>> {
          const size_t n = part == MeshPart::Cells ? m.cell_count()
                           : part == MeshPart::Faces ? m.face_count() : m.vertex_count();
          if (f == "none") return std::nullopt;
          std::vector<double> v(n);
          for (size_t i = 0; i < n; ++i) v[i] = double(i) + (f == "u" ? 0.5 : 0.0);
          return v;
        });
        std::printf("%s\n", fn.c_str());
      }
    } else if (cmd == "checkdefaults") {
      const Config& c = gc.check_file_configuration("constants");
      const Config& d = gc.check_file_configuration("qbdy");
      std::printf("%s %s %s\n", c.get<std::string>("enabled").c_str(),
                  c.get<std::string>("filename").c_str(), d.get<std::string>("enabled").c_str());
    } else {
      return 2;
    }
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  return 0;
}
