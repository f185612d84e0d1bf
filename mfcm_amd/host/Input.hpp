// Input.hpp -- the reference's input layer over the configuration tree:
// the global configuration (Config/Config.cpp:102-370), the mesh geometry
// (Mesh/Cartesian2DMesh.cpp:25-61, .hpp:115-188), times (Config/TimeParser.cpp,
// Config/TimeParameters.cpp:22-33), mesh selections (Mesh/MeshSelection.cpp),
// time series (Input/TimeSeries.cpp:179-296), raster fields
// (Raster/Raster.cpp:36-70, Raster/RasterFormats/NIMROD.cpp:313-566) and the
// field generator (Input/FieldGenerator.cpp:105-177).
//
// All of it is host-side preparation of the arrays and per-step values that
// the C ABI receives; nothing here computes on the solver's state.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Config.hpp"
#include "SaintVenantSolver.hpp"
#include "TemporalScheme.hpp"

namespace mfsv {

inline bool path_is_absolute(const std::string& p) { return !p.empty() && p[0] == '/'; }
inline std::string path_join(const std::string& a, const std::string& b) {
  if (a.empty() || a == ".") return b;
  return a.back() == '/' ? a + b : a + "/" + b;
}
inline std::string path_parent(const std::string& p) {
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? std::string() : p.substr(0, s);
}
// A path value: the whole text, or the text between quotes.
inline std::string path_value(const std::string& s) {
  if (s.size() >= 2 && s.front() == '"' && s.back() == '"') return s.substr(1, s.size() - 2);
  return s;
}

// ---- global configuration (Config/Config.cpp:102-370).  Not a singleton:
// the mirror passes it to what reads it.  Lookups that the reference answers
// with a default section add that section, as the reference does.
class GlobalConfig {
 public:
  GlobalConfig(std::string base_path, Config config)
      : base_(std::move(base_path)), conf_(std::move(config)) {}

  // Config.cpp:102-174: the configuration file and the simulation base path
  // (--base-path, else the file's directory; a relative file name is taken
  // from the base path when one is given).
  static GlobalConfig from_file(const std::string& file, const std::string& base_path = "") {
    std::string f = file, base = base_path;
    if (!base.empty()) {
      if (!path_is_absolute(f)) f = path_join(base, f);
    } else {
      base = path_parent(file);
      if (base.empty()) base = ".";
    }
    return GlobalConfig(base, MfReader::read_file(f));
  }

  const std::string& base_path() const { return base_; }
  std::string resolve(const std::string& p) const {
    return path_is_absolute(p) ? p : path_join(base_, p);
  }
  const Config& root() const { return conf_; }

  const Config& device_configuration() const { return conf_.get_child("device"); }
  const Config& scheme_configuration() const { return conf_.get_child("scheme"); }
  const Config& mesh_configuration() const { return conf_.get_child("mesh"); }
  const Config& time_series_configuration(const std::string& name) const {
    return named("time series", name, "Could not find time series '" + name + "'.");
  }
  const Config& raster_configuration(const std::string& name) const {
    return named("raster field", name, "Could not find raster field '" + name + "'.");
  }
  // Config.cpp:219-234: a field without a section gets an empty one.
  const Config& field_configuration(const std::string& name) {
    for (const Config* c : conf_.equal_range("field"))
      if (c->data() == name) return *c;
    return conf_.add("field", name);
  }
  const Config& boundary_configuration(const std::string& name) {
    for (const Config* c : conf_.equal_range("boundary"))
      if (c->data() == name) return *c;
    return conf_.add("boundary", name);
  }
  // Config.cpp:248-274: unknown check files are disabled, except
  // "constants", written to check/constants.csv by default.
  const Config& check_file_configuration(const std::string& name) {
    for (const Config* c : conf_.equal_range("check"))
      if (c->data() == name) return *c;
    Config& c = conf_.add("check", name);
    c.put("enabled", "false");
    if (name == "constants") {
      c.put("enabled", "true");
      c.put("filename", "check/constants.csv");
      c.put("delimiter", ",");
    }
    return c;
  }
  std::vector<std::string> output_files_list() const {
    std::vector<std::string> out;
    for (const Config* c : conf_.equal_range("output")) out.push_back(c->data());
    return out;
  }
  const Config& output_file_configuration(const std::string& name) {
    for (const Config* c : conf_.equal_range("output"))
      if (c->data() == name) return *c;
    Config& c = conf_.add("output", name);
    c.put("enabled", "false");
    if (name == "constants") {
      c.put("enabled", "true");
      c.put("every seconds", "1e39");
      c.put("filename", "check/constants.csv");
      c.put("delimiter", ",");
    }
    return c;
  }
  const Config& measure_configuration() {
    if (const Config* c = conf_.find("measures")) return *c;
    return conf_.add("measures", "");
  }
  std::vector<const Config*> source_term_configurations() const {
    return conf_.equal_range("source term");
  }
  std::vector<const Config*> boundary_configurations() const {
    return conf_.equal_range("boundary");
  }

 private:
  const Config& named(const char* key, const std::string& name, const std::string& err) const {
    for (const Config* c : conf_.equal_range(key))
      if (c->data() == name) return *c;
    throw std::runtime_error(err);
  }
  std::string base_;
  Config conf_;
};

// ---- mesh geometry (Cartesian2DMesh.cpp:25-61): cell counts, origin, cell
// size and rotation angle; the 12 transform coefficients as the reference
// keeps them; locations of cells, faces (x-faces first) and vertices.
class MeshGeometry {
 public:
  MeshGeometry(int64_t nx, int64_t ny, double ox, double oy, double sx, double sy, double angle = 0.0)
      : nx_(nx), ny_(ny) {
    if (nx < 1 || ny < 1) throw std::runtime_error("mesh: cell counts must be positive");
    g_[0] = ox;
    g_[1] = sx * std::cos(angle);
    g_[2] = sx * std::sin(angle);
    g_[3] = oy;
    g_[4] = sy * std::sin(angle);
    g_[5] = sy * std::cos(angle);
    finish();
  }
  // From a GDAL-style transform (origin x, b, c, origin y, e, f).
  static MeshGeometry from_transform(int64_t nx, int64_t ny, const double t[6]) {
    MeshGeometry m(nx, ny, 0, 0, 1, 1);
    for (int i = 0; i < 6; ++i) m.g_[i] = t[i];
    m.finish();
    return m;
  }
  // "mesh" section: cell count, origin, cell size, angle.
  static MeshGeometry from_config(const Config& c) {
    const auto n = split_fixed<size_t, 2>(c.get<std::string>("cell count"));
    const auto o = split_fixed<double, 2>(c.get<std::string>("origin", "0.0, 0.0"));
    const auto s = split_fixed<double, 2>(c.get<std::string>("cell size"));
    return MeshGeometry(int64_t(n[0]), int64_t(n[1]), o[0], o[1], s[0], s[1], c.get<double>("angle", 0.0));
  }

  int64_t nx() const { return nx_; }
  int64_t ny() const { return ny_; }
  size_t cell_count() const { return size_t(nx_ * ny_); }
  size_t face_count() const { return size_t((nx_ + 1) * ny_ + nx_ * (ny_ + 1)); }
  size_t vertex_count() const { return size_t((nx_ + 1) * (ny_ + 1)); }
  // the spacings the solver uses (Cartesian2DMesh.hpp:242-243)
  double dx() const { return g_[10]; }
  double dy() const { return g_[11]; }
  double origin_x() const { return g_[0]; }
  double origin_y() const { return g_[3]; }
  const double* transform() const { return g_; }

  // sub-cell coordinate -> location (Cartesian2DMesh.hpp:115-119)
  std::array<double, 2> location(double xi, double yi) const {
    return {g_[0] + xi * g_[1] + yi * g_[2], g_[3] + xi * g_[4] + yi * g_[5]};
  }
  // location -> sub-cell coordinate (Cartesian2DMesh.hpp:125-134)
  std::array<double, 2> coordinate(double x, double y) const {
    return {((x - g_[0]) * g_[6] - (y - g_[3]) * g_[2] * g_[8]) * g_[9],
            ((y - g_[3]) * g_[7] - (x - g_[0]) * g_[4] * g_[8]) * g_[9]};
  }
  std::array<double, 2> cell_location(size_t i) const {
    return location(double(i % size_t(nx_)) + 0.5, double(i / size_t(nx_)) + 0.5);
  }
  std::array<double, 2> face_location(size_t i) const {
    const size_t nxf = size_t(nx_ + 1) * size_t(ny_);
    if (i < nxf) return location(double(i % size_t(nx_ + 1)), double(i / size_t(nx_ + 1)) + 0.5);
    const size_t j = i - nxf;
    return location(double(j % size_t(nx_)) + 0.5, double(j / size_t(nx_)));
  }
  std::array<double, 2> vertex_location(size_t i) const {
    return location(double(i % size_t(nx_ + 1)), double(i / size_t(nx_ + 1)));
  }
  // nearest cell, or cell_count() outside the mesh (Cartesian2DMesh.hpp:178-188)
  size_t nearest_cell(double x, double y) const {
    const auto c = coordinate(x, y);
    if (c[0] >= 0.0 && c[0] < double(nx_) && c[1] >= 0.0 && c[1] < double(ny_))
      return size_t(c[1]) * size_t(nx_) + size_t(c[0]);
    return cell_count();
  }

 private:
  void finish() {
    g_[6] = 1.0 / g_[1];
    g_[7] = 1.0 / g_[5];
    g_[8] = 1.0 / (g_[1] * g_[5]);
    g_[9] = 1.0 - (g_[2] * g_[4]) * g_[8];
    g_[10] = std::sqrt(g_[1] * g_[1] + g_[2] * g_[2]);
    g_[11] = std::sqrt(g_[4] * g_[4] + g_[5] * g_[5]);
  }
  int64_t nx_, ny_;
  double g_[12] = {};
};

// ---- times (Config/TimeParser.cpp): a number in the time unit (hours
// unless "time unit" says otherwise), or, with "time format", a formatted
// time point measured from "time zero".
class TimeParser {
 public:
  explicit TimeParser(const Config& c) { configure(c); }
  TimeParser(const TimeParser& base, const Config& c) : TimeParser(base) { configure(c); }

  double parse(const std::string& s) const {
    if (formatted_) {
      std::istringstream ss(s);
      std::tm t = {};
      ss >> std::get_time(&t, fmt_.c_str());
      if (!ss.fail()) return double(std::mktime(&t) - zero_);
    }
    return config_detail::lexical<double>(s) * unit_;
  }
  double unit_factor() const { return unit_; }
  bool formatted() const { return formatted_; }

 private:
  void configure(const Config& c) {
    fmt_ = c.get<std::string>("time format", fmt_);
    formatted_ = !fmt_.empty();
    if (formatted_) {
      std::istringstream ss(c.get<std::string>("time zero"));
      std::tm t = {};
      ss >> std::get_time(&t, fmt_.c_str());
      zero_ = std::mktime(&t);
    }
    const std::string u = c.get<std::string>("time unit", "default");
    static const std::map<std::string, double> units = {
        {"seconds", 1.0}, {"second", 1.0}, {"secs", 1.0}, {"sec", 1.0}, {"s", 1.0},
        {"minutes", 60.0}, {"minute", 60.0}, {"mins", 60.0}, {"min", 60.0}, {"m", 60.0},
        {"hours", 3600.0}, {"hour", 3600.0}, {"hrs", 3600.0}, {"hr", 3600.0}, {"h", 3600.0}};
    const auto it = units.find(u);
    if (it != units.end()) {
      unit_ = it->second;
      return;
    }
    // a leading number is a factor to seconds
    char* end = nullptr;
    const double f = std::strtod(u.c_str(), &end);
    if (end != u.c_str() && !std::isspace(static_cast<unsigned char>(u[0]))) {
      unit_ = f;
    } else if (u != "default") {
      throw std::runtime_error("Unknown time unit: \"" + u + "\"");
    }
  }
  std::string fmt_;
  bool formatted_ = false;
  std::time_t zero_ = 0;
  double unit_ = 3600.0;
};

// RungeKuttaTimeParameters(conf) (TimeParameters.cpp:22-33) on the "scheme"
// section.
inline std::shared_ptr<TimeParameters<double>> time_parameters_from(const Config& scheme,
                                                                    const TimeParser& p) {
  const double start = p.parse(scheme.get<std::string>("start time", "0"));
  const double end = p.parse(scheme.get<std::string>("end time"));
  double step = p.parse(scheme.get<std::string>("step duration", "0"));
  if (step == 0.0) step = scheme.get<double>("step duration seconds", 60.0);
  const double dt0 = scheme.get<double>("initial timestep seconds", 1.0);
  return std::make_shared<TimeParameters<double>>(start, end, step, dt0);
}

// ---- mesh selections of cells (Mesh/MeshSelection.cpp:146-197): global,
// "id list" (id / ids entries, sorted, duplicates dropped) or "at" (the cells
// nearest the WKT points).
class MeshSelection {
 public:
  MeshSelection() = default;
  static MeshSelection global_selection() { return MeshSelection(); }
  static MeshSelection from_config(const MeshGeometry& mesh, const Config& c) {
    MeshSelection s;
    const std::string kind = c.get_value<std::string>("global");
    if (kind == "global" || kind.empty()) return s;
    s.global_ = false;
    const size_t n = mesh.cell_count();
    if (kind == "id list") {
      std::vector<size_t> ids;
      for (const Config* e : c.equal_range("id")) ids.push_back(e->get_value<size_t>());
      for (const Config* e : c.equal_range("ids"))
        for (size_t v : split_values<size_t>(e->data())) ids.push_back(v);
      std::sort(ids.begin(), ids.end());
      ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
      for (size_t v : ids)
        if (v < n) s.ids_.push_back(int64_t(v));
      return s;
    }
    if (kind == "at") {
      if (c.count("wkt") != 1)
        throw std::runtime_error("an \"at\" selection needs one wkt entry (GDAL sources are not "
                                 "supported in this build)");
      std::vector<uint8_t> mask(n, 0);
      for (const auto& p : wkt_points(c.get<std::string>("wkt"))) {
        const size_t i = mesh.nearest_cell(p[0], p[1]);
        if (i < n) mask[i] = 1;
      }
      for (size_t i = 0; i < n; ++i)
        if (mask[i]) s.ids_.push_back(int64_t(i));
      return s;
    }
    throw std::runtime_error("Unknown mesh selection method: " + kind);
  }
  bool is_global() const { return global_; }
  const std::vector<int64_t>& ids() const { return ids_; }
  // f(i) for each selected cell, ascending
  template <typename F>
  void for_each(size_t ncells, F f) const {
    if (global_) {
      for (size_t i = 0; i < ncells; ++i) f(i);
    } else {
      for (int64_t i : ids_) f(size_t(i));
    }
  }

  // WKT of POINT geometries (Geometry/GeometryParser.hpp:93-221); a mask is
  // only updated from points (MeshSelection.cpp:100-112), so other tagged
  // geometries are rejected.
  static std::vector<std::array<double, 2>> wkt_points(const std::string& w) {
    std::vector<std::array<double, 2>> pts;
    size_t i = 0;
    auto ws = [&] { while (i < w.size() && std::isspace(static_cast<unsigned char>(w[i]))) ++i; };
    auto tag = [&]() {
      size_t j = i;
      while (j < w.size() && std::isalpha(static_cast<unsigned char>(w[j]))) ++j;
      std::string t = config_detail::lower(w.substr(i, j - i));
      i = j;
      return t;
    };
    while (true) {
      ws();
      std::string t = tag();
      if (t != "point") {
        if (t.empty()) throw std::runtime_error("Failed to parse WKT string: " + w.substr(i));
        throw std::runtime_error("Cannot update mask with geometry of type: " + t);
      }
      ws();
      // optional Z / M markers
      for (int k = 0; k < 2; ++k) {
        if (i < w.size() && (w[i] == 'z' || w[i] == 'Z' || w[i] == 'm' || w[i] == 'M')) { ++i; ws(); }
      }
      if (i >= w.size() || w[i] != '(') throw std::runtime_error("Failed to parse WKT string: " + w);
      ++i;
      std::vector<double> xs;
      while (true) {
        ws();
        const char* b = w.c_str() + i;
        char* e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) break;
        xs.push_back(v);
        i += size_t(e - b);
      }
      ws();
      if (i >= w.size() || w[i] != ')' || xs.size() < 2)
        throw std::runtime_error("Failed to parse WKT string: " + w);
      ++i;
      pts.push_back({xs[0], xs[1]});
      ws();
      if (i == w.size()) break;
      if (w[i] != ',') throw std::runtime_error("Failed to parse WKT string: " + w.substr(i));
      ++i;
    }
    return pts;
  }

 private:
  bool global_ = true;
  std::vector<int64_t> ids_;
};

// ---- time series (Input/TimeSeries.cpp:179-296 and :351-373) -------------
namespace ts_detail {

// A CSV reader with the settings the reference gives its CSV library
// (TimeSeries.cpp:225-229): a separator, cells trimmed of white space and of
// enclosing quotes (a doubled quote inside stands for one), lines whose first
// cell starts with the comment character and empty lines skipped, row
// `label_row` holding the column names (-1: none) and the rows after it the
// data, column `label_col` holding row names (-1: none).
struct Csv {
  std::vector<std::vector<std::string>> rows;
  std::map<std::string, long> names;
  long label_row = -1, label_col = -1;

  static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
  }
  static std::string unquote(const std::string& s) {
    if (s.size() >= 2 && s.front() == '"' && s.back() == '"') {
      std::string r, in = s.substr(1, s.size() - 2);
      for (size_t i = 0; i < in.size(); ++i) {
        r += in[i];
        if (in[i] == '"' && i + 1 < in.size() && in[i + 1] == '"') ++i;
      }
      return r;
    }
    return s;
  }
  void parse(const std::string& text, char sep, char comment, long lrow, long lcol) {
    label_row = lrow;
    label_col = lcol;
    std::vector<std::string> row;
    std::string cell;
    bool quoted = false;
    auto end_row = [&](bool last) {
      row.push_back(unquote(trim(cell)));
      if (last || row.empty() || row[0].empty() || row[0][0] != comment) rows.push_back(row);
      cell.clear();
      row.clear();
      quoted = false;
    };
    for (char ch : text) {
      if (ch == '"') {
        if (cell.empty() || cell[0] == '"') quoted = !quoted;
        cell += ch;
      } else if (ch == sep) {
        if (!quoted) {
          row.push_back(unquote(trim(cell)));
          cell.clear();
        } else {
          cell += ch;
        }
      } else if (ch == '\r') {
        // line ends are '\n'; a CR is dropped
      } else if (ch == '\n') {
        if (row.empty() && cell.empty()) continue;  // empty line
        end_row(false);
      } else {
        cell += ch;
      }
    }
    // a last line without a line end is kept as it is
    if (!cell.empty() || !row.empty()) {
      row.push_back(unquote(trim(cell)));
      rows.push_back(row);
    }
    if (label_row >= 0 && long(rows.size()) > label_row) {
      long i = 0;
      for (const auto& nm : rows[size_t(label_row)]) names[nm] = i++;
    }
  }
  long column_index(const std::string& name) const {
    const auto it = names.find(name);
    if (label_row < 0 || it == names.end()) return -1;
    return it->second - (label_col + 1);
  }
  std::vector<std::string> column(long idx) const {
    const long c = idx + (label_col + 1);
    std::vector<std::string> out;
    for (size_t r = 0; r < rows.size(); ++r) {
      if (long(r) <= label_row) continue;
      if (c < 0 || size_t(c) >= rows[r].size())
        throw std::runtime_error("CSV row " + std::to_string(r + 1) + " has no column " +
                                 std::to_string(c + 1));
      out.push_back(rows[r][size_t(c)]);
    }
    return out;
  }
};

// A number at the start of the cell, NaN when there is none (the CSV
// library's default converter).
inline double cell_number(const std::string& s) {
  const char* b = s.c_str();
  char* e = nullptr;
  const double v = std::strtod(b, &e);
  return e == b ? std::numeric_limits<double>::quiet_NaN() : v;
}

}  // namespace ts_detail

inline TimeSeries load_time_series(const GlobalConfig& gc, const TimeParser& parser,
                                   const Config& conf) {
  const std::string src = config_detail::lower(conf.get<std::string>("source", "not specified"));
  std::vector<double> t, v;
  if (src == "inline") {
    const Config& s = conf.get_child("source");
    const TimeParser tp(parser, s);
    for (const auto& e : s.get_child("data")) {
      const double ti = tp.parse(config_detail::lower(e.first));
      if (!t.empty() && ti <= t.back())
        throw std::runtime_error("Times in time series must increase.");
      t.push_back(ti);
      v.push_back(e.second.get_value<double>());
    }
  } else if (src == "csv") {
    const std::string file = gc.resolve(path_value(conf.get<std::string>("filename")));
    const char sep = conf.get<char>("separator", ',');
    const char comment = conf.get<char>("comment character", '#');
    const bool headers = conf.get<bool>("headers", true);
    long skip_rows = conf.get<long>("skip rows", -1);
    const long skip_cols = conf.get<long>("skip cols", -1);
    if (headers) skip_rows += 1;
    const TimeParser tp(parser, conf);
    std::ifstream f(file, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open time series file " + file);
    std::stringstream buf;
    buf << f.rdbuf();
    ts_detail::Csv csv;
    csv.parse(buf.str(), sep, comment, skip_rows, skip_cols);
    long tc, vc;
    if (headers) {
      tc = csv.column_index(conf.get<std::string>("time column"));
      vc = csv.column_index(conf.get<std::string>("value column"));
      if (tc < 0 || vc < 0) throw std::runtime_error("time series " + file + ": no such column");
    } else {
      tc = long(conf.get<size_t>("time column", 1)) - 1;
      vc = long(conf.get<size_t>("value column", 2)) - 1;
    }
    for (const auto& s : csv.column(tc)) t.push_back(tp.parse(s));
    for (const auto& s : csv.column(vc)) v.push_back(ts_detail::cell_number(s));
  } else {
    throw std::runtime_error("Unknown source type '" + src + "' for time series.");
  }
  // TimeSeries.cpp:64-99: (x + offset) * factor
  const Config empty;
  const Config& tr = conf.get_child("transform", empty);
  const double tf = tr.get<double>("time factor", 1.0), to = tr.get<double>("time offset", 0.0);
  const double vf = tr.get<double>("value factor", 1.0), vo = tr.get<double>("value offset", 0.0);
  for (double& x : t) x = (x + to) * tf;
  for (double& x : v) x = (x + vo) * vf;
  // sanity checks (TimeSeries.cpp:121-160)
  if (t.size() != v.size()) throw std::logic_error("Mismatched time and value counts in time series.");
  if (t.size() < 2) throw std::runtime_error("Time series does not have at least two data.");
  for (size_t i = 1; i < t.size(); ++i)
    if (!(t[i] > t[i - 1])) throw std::logic_error("Time points do not increase in a time series.");
  return TimeSeries(std::move(t), std::move(v));
}

// Loaded once per name (TimeSeries.cpp:351-373).
class TimeSeriesDatabase {
 public:
  const TimeSeries& get(const GlobalConfig& gc, const TimeParser& parser, const std::string& name) {
    auto it = db_.find(name);
    if (it == db_.end())
      it = db_.emplace(name, load_time_series(gc, parser, gc.time_series_configuration(name))).first;
    return it->second;
  }
  const std::map<std::string, TimeSeries>& all() const { return db_; }

 private:
  std::map<std::string, TimeSeries> db_;
};

// ---- raster fields ----------------------------------------------------------
// A north-up raster: ncols x nrows values, row 0 the northernmost, with the
// lower-left corner (x0, y0) and pixel sizes (px, py).
struct Raster {
  int64_t ncols = 0, nrows = 0;
  double x0 = 0, y0 = 0, px = 1, py = 1;
  double nodata = std::numeric_limits<double>::quiet_NaN();
  std::vector<double> values;  // row-major, row 0 = top

  // value of the pixel containing (x, y), or false outside
  bool sample(double x, double y, double* out) const {
    const double cx = (x - x0) / px;
    const double cy = (y0 + double(nrows) * py - y) / py;  // rows from the top
    if (!(cx >= 0.0 && cx < double(ncols) && cy >= 0.0 && cy < double(nrows))) return false;
    *out = values[size_t(cy) * size_t(ncols) + size_t(cx)];
    return true;
  }
};

namespace nimrod_detail {
inline uint32_t be32(const unsigned char* p) {
  return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}
inline int16_t be16(const unsigned char* p) { return int16_t((uint16_t(p[0]) << 8) | uint16_t(p[1])); }
inline float bef(const unsigned char* p) {
  const uint32_t u = be32(p);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
}  // namespace nimrod_detail

// Met Office NIMROD file (NIMROD.cpp:313-566): a 512-byte big-endian header
// record (31 int16, 28 float32, 45 float32, 56 characters, 51 int16) and one
// data record, each framed by 4-byte lengths.  Data types: 0 float32, 1 int16
// or int32, 2 char.  Only the National Grid with a top-left origin is
// supported, as in the reference.  An optional "bbox" (xmin, ymin, xmax,
// ymax) crops the raster to the pixels it covers.  Values are the stored
// numbers (the reference applies neither the scaling factor nor the
// missing-data value).
inline Raster read_nimrod(const std::string& path, const Config& conf) {
  using namespace nimrod_detail;
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) throw std::runtime_error("Could not open NIMROD data file " + path);
  std::vector<unsigned char> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  auto need = [&](size_t off, size_t n) {
    if (off + n > b.size()) throw std::runtime_error("NIMROD file " + path + " is truncated");
  };
  need(0, 4 + 512 + 8);
  if (be32(&b[0]) != 512) throw std::runtime_error("NIMROD: expected header size indicator 512");
  const unsigned char* h = &b[4];
  int16_t h1[31];
  float h2[28];
  for (int i = 0; i < 31; ++i) h1[i] = be16(h + 2 * i);
  for (int i = 0; i < 28; ++i) h2[i] = bef(h + 62 + 4 * i);
  if (be32(&b[4 + 512]) != 512) throw std::runtime_error("NIMROD: expected header size indicator 512");
  const uint32_t block = be32(&b[4 + 512 + 4]);
  const size_t data_off = 4 + 512 + 8;
  const int64_t nx = h1[16], ny = h1[15];
  if (nx <= 0 || ny <= 0) throw std::runtime_error("NIMROD: bad grid size");
  const int type = h1[11], bpp = h1[12];
  const size_t n = size_t(nx) * size_t(ny);
  std::vector<double> all(n);
  if (type == 0) {
    if (bpp != 4) throw std::runtime_error("Unsupported bpp for float data.");
    need(data_off, 4 * n);
    for (size_t i = 0; i < n; ++i) all[i] = bef(&b[data_off + 4 * i]);
  } else if (type == 1) {
    if (bpp == 2) {
      need(data_off, 2 * n);
      for (size_t i = 0; i < n; ++i) all[i] = be16(&b[data_off + 2 * i]);
    } else if (bpp == 4) {
      need(data_off, 4 * n);
      for (size_t i = 0; i < n; ++i) all[i] = double(int32_t(be32(&b[data_off + 4 * i])));
    } else {
      throw std::runtime_error("Unsupported bpp for integer data.");
    }
  } else if (type == 2) {
    if (bpp != 1) throw std::runtime_error("Unsupported bpp for char data.");
    need(data_off, n);
    for (size_t i = 0; i < n; ++i) all[i] = double(static_cast<char>(b[data_off + i]));
  } else {
    throw std::runtime_error("Unsupported NIMROD data type");
  }
  const size_t tail = data_off + n * size_t(bpp);
  need(tail, 4);
  if (be32(&b[tail]) != block) throw std::runtime_error("NIMROD: data block length mismatch");
  switch (h1[23]) {
    case 0: break;
    case 1: throw std::runtime_error("Bottom-left grid origin location.");
    case 2: throw std::runtime_error("Top-right grid origin location.");
    case 3: throw std::runtime_error("Bottom-right grid origin location.");
    default: throw std::runtime_error("Unknown grid origin location.");
  }
  // h2[2]: y of the first (top) row's centres, h2[3]: row spacing,
  // h2[4]: x of the first column's centres, h2[5]: column spacing
  const double px = h2[5], py = h2[3];
  const double llc_x = double(h2[4]) - 0.5 * px;
  const double urc_y = double(h2[2]) + 0.5 * py;
  Raster r;
  r.px = px;
  r.py = py;
  int64_t c0 = 0, r0 = 0, c1 = nx - 1, r1 = ny - 1;
  if (conf.count("bbox") > 1) throw std::runtime_error("Too many bounding boxes.");
  if (conf.count("bbox") == 1) {
    const auto bb = split_fixed<double, 4>(conf.get<std::string>("bbox"));
    if (bb[2] <= bb[0] || bb[3] <= bb[1]) throw std::runtime_error("Invalid bounding box.");
    auto clampi = [](double v, int64_t hi) {
      return v < 0 ? int64_t(0) : (v >= double(hi) ? hi - 1 : int64_t(v));
    };
    c0 = clampi((bb[0] - llc_x) / px, nx);
    c1 = clampi((bb[2] - llc_x) / px, nx);
    r0 = clampi((urc_y - bb[3]) / py, ny);
    r1 = clampi((urc_y - bb[1]) / py, ny);
  }
  r.ncols = c1 - c0 + 1;
  r.nrows = r1 - r0 + 1;
  r.x0 = llc_x + double(c0) * px;
  r.y0 = urc_y - double(r1 + 1) * py;
  r.values.resize(size_t(r.ncols * r.nrows));
  for (int64_t i = 0; i < r.nrows; ++i)
    for (int64_t j = 0; j < r.ncols; ++j)
      r.values[size_t(i * r.ncols + j)] = all[size_t((r0 + i) * nx + (c0 + j))];
  r.nodata = type == 0 ? double(h2[6]) : double(h1[24]);
  return r;
}

// Loaded once per name (Raster/Raster.cpp:36-70).
class RasterDatabase {
 public:
  const Raster& get(const GlobalConfig& gc, const std::string& name) {
    auto it = db_.find(name);
    if (it != db_.end()) return it->second;
    const Config& c = gc.raster_configuration(name);
    const std::string file = gc.resolve(path_value(c.get<std::string>("filename", ".")));
    const std::string src = config_detail::lower(c.get<std::string>("source"));
    if (src == "gdal") throw std::runtime_error("GDAL not supported in this build.");
    if (src != "nimrod")
      throw std::runtime_error("Unknown source type '" + src + "' for raster field: " + name);
    return db_.emplace(name, read_nimrod(file, c)).first->second;
  }

 private:
  std::map<std::string, Raster> db_;
};

// ---- field generator (Input/FieldGenerator.cpp:105-177): the field starts
// at its default value; each "set", "add" or "multiply" entry of its section
// loads an input field and applies it.  Input fields: "constant" (value),
// a number, "composite" (constant models on cell selections), "slope"
// (origin x, y, z and slopes), "raster" (a raster field by name, each cell
// taking the pixel that contains its centre).
class FieldGenerator {
 public:
  FieldGenerator(GlobalConfig& gc, RasterDatabase& rasters, const MeshGeometry& mesh)
      : gc_(gc), rasters_(rasters), mesh_(mesh) {}

  std::vector<double> operator()(const std::string& name, double init) {
    std::vector<double> f(mesh_.cell_count(), init);
    const Config& conf = gc_.field_configuration(name);
    for (const auto& e : conf) {
      const std::string op = e.first;
      if (op == "set") {
        load(f, e.second);
      } else if (op == "add") {
        std::vector<double> in(f.size(), 0.0);
        load(in, e.second);
        for (size_t i = 0; i < f.size(); ++i) f[i] += in[i];
      } else if (op == "multiply") {
        std::vector<double> in(f.size(), 1.0);
        load(in, e.second);
        for (size_t i = 0; i < f.size(); ++i) f[i] *= in[i];
      } else {
        throw std::runtime_error("Unknown field operation '" + op + "'.");
      }
    }
    return f;
  }
  // Was the field given a section of its own?
  bool configured(const std::string& name) const {
    for (const Config* c : gc_.root().equal_range("field"))
      if (c->data() == name && !c->empty()) return true;
    return false;
  }

 private:
  // (x * 0) + v on the selection (FieldGenerator.cpp:42-54)
  void constant(std::vector<double>& f, const MeshSelection& sel, double v) {
    sel.for_each(f.size(), [&](size_t i) { f[i] = f[i] * 0.0 + v; });
  }
  void load(std::vector<double>& f, const Config& c) {
    const Config empty;
    const MeshSelection sel = MeshSelection::from_config(mesh_, c.get_child("within", empty));
    const std::string type = config_detail::lower(c.get_value<std::string>());
    if (type == "constant") {
      constant(f, sel, c.get<double>("value"));
    } else if (type == "composite") {
      if (!sel.is_global())
        throw std::runtime_error("Composite input field only supported with global selection.");
      for (const Config* m : c.equal_range("model")) {
        if (m->data() != "constant")
          throw std::runtime_error("Unknown field model type: \"" + m->data() + "\"");
        const MeshSelection ms = MeshSelection::from_config(mesh_, m->get_child("cells"));
        constant(f, ms, m->get<double>("value"));
      }
    } else if (type == "slope") {
      if (!sel.is_global())
        throw std::runtime_error("Sloped input field only supported with global selection.");
      const auto o = split_fixed<double, 3>(c.get<std::string>("origin"));
      const auto s = split_fixed<double, 2>(c.get<std::string>("slope"));
      for (size_t i = 0; i < f.size(); ++i) {
        const auto loc = mesh_.cell_location(i);
        f[i] = o[2] + s[0] * (loc[0] - o[0]) + s[1] * (loc[1] - o[1]);
      }
    } else if (type == "raster") {
      const Raster& r = rasters_.get(gc_, c.get<std::string>("name"));
      sel.for_each(f.size(), [&](size_t i) {
        const auto loc = mesh_.cell_location(i);
        double v;
        if (r.sample(loc[0], loc[1], &v)) f[i] = v;
      });
    } else {
      char* end = nullptr;
      const double v = std::strtod(type.c_str(), &end);
      if (end == type.c_str())
        throw std::runtime_error("\"" + type + "\" is not a known input field specifier and "
                                 "cannot be interpreted as a constant value.");
      constant(f, sel, v);
    }
  }

  GlobalConfig& gc_;
  RasterDatabase& rasters_;
  const MeshGeometry& mesh_;
};

}  // namespace mfsv
