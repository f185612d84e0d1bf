// Output.hpp -- the reference's output layer: CSV formatting
// (Output/OutputFormat.hpp:83-121), timed multi-field output files
// (Output/OutputFile.hpp:31-256), check files (Output/CheckFile.hpp:27-172)
// and measures (Output/Measure.hpp, SaintVenant/Measure.hpp).
//
// Values reach these classes as host arrays that the caller downloads from
// the device (fields of the current state, the output-only fields of the last
// stage evaluation); the classes only format and write them.
#pragma once

#include <array>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "Input.hpp"

namespace mfsv {

// std::to_string of a value (the reference's TypedOutputFunction::at_str).
inline std::string value_text(double v) { return std::to_string(v); }

// One table to write: column names and values, with or without a location
// (X, Y) per row.
struct OutputTable {
  std::vector<std::string> names;
  std::vector<const std::vector<double>*> columns;  // each nrows long
  std::function<std::array<double, 2>(size_t)> location;  // empty: no X, Y
  size_t nrows = 0;
  // integer-valued columns (selection ids) are written as integers
  std::vector<bool> integral;
};

// CSVOutput::output (OutputFormat.hpp:93-121): a header of "X", "Y" and the
// quoted column names, each followed by the delimiter; per row the location
// (default stream formatting) and the values (std::to_string), each followed
// by the delimiter.
inline void write_csv(const std::string& filename, const OutputTable& t,
                      const std::string& delim = ", ") {
  std::ofstream ofs(filename);
  if (!ofs) throw std::runtime_error("cannot write " + filename);
  if (t.location) ofs << "X" << delim << "Y" << delim;
  for (const auto& n : t.names) ofs << std::quoted(n) << delim;
  ofs << std::endl;
  for (size_t r = 0; r < t.nrows; ++r) {
    if (t.location)
      for (double c : t.location(r)) ofs << c << delim;
    for (size_t c = 0; c < t.columns.size(); ++c) {
      const double v = (*t.columns[c])[r];
      if (c < t.integral.size() && t.integral[c]) ofs << std::to_string(static_cast<size_t>(v));
      else ofs << value_text(v);
      ofs << delim;
    }
    ofs << std::endl;
  }
  if (!ofs) throw std::runtime_error("cannot write " + filename);
}

inline void make_parent_dirs(const std::string& file) {
  const std::filesystem::path p(file);
  if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path());
}

// ---- timed multi-field output files (OutputFile.hpp:31-256) ---------------
enum class MeshPart { Cells, Faces, Vertices };

class MultiFieldOutputFile {
 public:
  // fields: name -> values at the mesh part, or nullopt when the solver has
  // no such field there (the column is left out, as the reference does)
  using FieldSource = std::function<std::optional<std::vector<double>>(const std::string&, MeshPart)>;

  MultiFieldOutputFile(GlobalConfig& gc, const std::string& name, const TimeParser& parser,
                       const TimeParameters<double>& tp)
      : name_(name) {
    const Config& c = gc.output_file_configuration(name);
    enabled_ = c.get<bool>("enabled", true);
    filename_ = gc.resolve(path_value(c.get<std::string>("filename", "output/" + name)));
    start_ = c.find("start time") ? parser.parse(c.get<std::string>("start time")) : tp.start_time();
    end_ = c.find("end time") ? parser.parse(c.get<std::string>("end time")) : tp.end_time();
    every_ = c.find("every") ? parser.parse(c.get<std::string>("every")) : c.get<double>("every seconds");
    if (c.count("field") > 0) {
      for (const Config* f : c.equal_range("field")) fields_.push_back(f->data());
    } else {
      fields_.push_back(name);
    }
    const std::string at = c.get<std::string>("at", "cells");
    if (at == "cells") part_ = MeshPart::Cells;
    else if (at == "faces") part_ = MeshPart::Faces;
    else if (at == "vertices") part_ = MeshPart::Vertices;
    else throw std::runtime_error("Invalid value for output.at");
  }
  const std::string& name() const { return name_; }
  MeshPart part() const { return part_; }
  const std::vector<std::string>& fields() const { return fields_; }
  bool due(double time_now) const {
    return enabled_ && time_now >= start_ + every_ * double(output_no_);
  }

  // MultiFieldOutputFile::timed_output (OutputFile.hpp:185-215): the file
  // <filename>_<time>.csv; returns its name, or "" when nothing was due.
  std::string timed_output(double time_now, const MeshGeometry& mesh, const FieldSource& src) {
    if (!due(time_now)) return std::string();
    const std::string fn = filename_ + "_" + std::to_string(time_now) + ".csv";
    make_parent_dirs(fn);
    std::vector<std::vector<double>> cols;
    OutputTable t;
    for (const auto& f : fields_) {
      auto v = src(f, part_);
      if (v) {
        cols.push_back(std::move(*v));
        t.names.push_back(f);
      }
    }
    for (const auto& c : cols) t.columns.push_back(&c);
    t.nrows = cols.empty() ? 0 : cols[0].size();
    switch (part_) {
      case MeshPart::Cells: t.location = [&](size_t i) { return mesh.cell_location(i); }; break;
      case MeshPart::Faces: t.location = [&](size_t i) { return mesh.face_location(i); }; break;
      case MeshPart::Vertices: t.location = [&](size_t i) { return mesh.vertex_location(i); }; break;
    }
    ++output_no_;
    write_csv(fn, t);
    return fn;
  }

 private:
  std::string name_, filename_;
  bool enabled_ = true;
  double start_ = 0, end_ = 0, every_ = 0;
  size_t output_no_ = 0;
  std::vector<std::string> fields_;
  MeshPart part_ = MeshPart::Cells;
};

// ---- check files (CheckFile.hpp:27-172) ---------------------------------
class CheckFile {
 public:
  CheckFile(GlobalConfig& gc, const std::string& name) : name_(name), conf_(&gc.check_file_configuration(name)) {
    enabled_ = conf_->get<bool>("enabled", true);
    filename_ = gc.resolve(path_value(conf_->get<std::string>("filename", "check/" + name + ".csv")));
  }
  bool enabled() const { return enabled_; }
  const std::string& filename() const { return filename_; }
  void set_filename(const std::string& f) { filename_ = f; }

  // CheckFile::output: csv (the only format), delimiter from the "format"
  // section (", " by default)
  void output(const OutputTable& t) const {
    if (!enabled_) return;
    const std::string fmt = conf_->get<std::string>("format", "csv");
    if (fmt != "csv") throw std::runtime_error("Unknown output format: " + fmt);
    const Config empty;
    const std::string delim = conf_->get_child("format", empty).get<std::string>("delimiter", ", ");
    make_parent_dirs(filename_);
    write_csv(filename_, t, delim);
  }
  // FieldCheckFile: named cell fields with their locations.
  void output_fields(const MeshGeometry& mesh, const std::vector<std::string>& names,
                     const std::vector<const std::vector<double>*>& cols) const {
    if (!enabled_) return;
    OutputTable t;
    t.names = names;
    t.columns = cols;
    t.nrows = cols.empty() ? 0 : cols[0]->size();
    t.location = [&](size_t i) { return mesh.cell_location(i); };
    output(t);
  }
  // MeshSelectionCheckFile: the selection's ids (all cells when global).
  void output_selection(const MeshSelection& sel, size_t ncells) const {
    if (!enabled_) return;
    std::vector<double> ids;
    sel.for_each(ncells, [&](size_t i) { ids.push_back(double(i)); });
    OutputTable t;
    t.names = {"ID"};
    t.columns = {&ids};
    t.integral = {true};
    t.nrows = ids.size();
    output(t);
  }
  // TimeSeriesCheckFile: <stem>_<series name><ext>, columns "ID" (times)
  // and the series' name (values).
  void output_series(const std::string& ts_name, const TimeSeries& ts) {
    if (!enabled_) return;
    const std::filesystem::path p(filename_);
    filename_ = (p.parent_path() / p.stem()).string() + "_" + ts_name + p.extension().string();
    std::vector<double> t = ts.times(), v = ts.values();
    OutputTable tab;
    tab.names = {"ID", ts_name};
    tab.columns = {&t, &v};
    tab.nrows = t.size();
    output(tab);
  }

 private:
  std::string name_, filename_;
  const Config* conf_;
  bool enabled_ = true;
};

// ---- measures (SaintVenant/Measure.hpp:97-226, Output/Measure.hpp:26-100):
// "h-point" entries of the "measures" section record the depth of the cell
// nearest "location" at each step start; output() writes the series as
// output/measures/<name>.csv (columns "ID" = time and the measure's name).
class HPointMeasure {
 public:
  HPointMeasure(GlobalConfig& gc, const Config& c, const MeshGeometry& mesh, size_t nsteps)
      : name_(c.data()),
        t_(nsteps + 1, std::numeric_limits<double>::quiet_NaN()),
        v_(nsteps + 1, std::numeric_limits<double>::quiet_NaN()) {
    filename_ = gc.resolve(path_value(c.get<std::string>("filename", "output/measures/" + name_ + ".csv")));
    const auto loc = split_fixed<double, 2>(c.get<std::string>("location"));
    cell_ = mesh.nearest_cell(loc[0], loc[1]);
    if (cell_ >= mesh.cell_count())
      throw std::runtime_error("measure \"" + name_ + "\": location outside the mesh");
  }
  const std::string& name() const { return name_; }
  size_t cell() const { return cell_; }
  void record(double time_now, double h) {
    if (next_ >= t_.size()) throw std::runtime_error("measure \"" + name_ + "\": too many points");
    t_[next_] = time_now;
    v_[next_] = h;
    ++next_;
  }
  const std::vector<double>& times() const { return t_; }
  const std::vector<double>& values() const { return v_; }
  void output() const {
    make_parent_dirs(filename_);
    OutputTable t;
    t.names = {"ID", name_};
    t.columns = {&t_, &v_};
    t.nrows = t_.size();
    write_csv(filename_, t);
  }

 private:
  std::string name_, filename_;
  size_t cell_ = 0;
  std::vector<double> t_, v_;
  size_t next_ = 0;
};

}  // namespace mfsv
