// ConfiguredModel.hpp -- a Saint-Venant run set up from the reference's own
// configuration files: what mfcm's main (mfcm.cpp:19-49) builds through
// GlobalConfig -- the scheme and its time parameters (create_scheme), the
// mesh (Cartesian2DMesh.cpp:25-61), the bed and initial state from field
// generators (Constants.cpp:27-44, State.cpp:24-38), the source terms
// (SourceTerm.cpp:30-50 and each term's create_source_term), the boundaries
// (BoundarySourceTerm.cpp:28-45 and each start_new_step), the outputs
// (RungeKutta.hpp:286-315, OutputFile.hpp:222-256), the measures
// (SaintVenant/Measure.hpp:210-224) and the check files.
//
// Where the solver's kernels fix an order the configuration could change
// (source terms: Manning, energy loss, infiltration, eddy viscosity;
// boundaries: discharge, head, stage), a configuration in another order is
// rejected for boundaries (their order changes results) and reported for
// source terms (their order changes only the rounding of a sum).
#pragma once

#include <iostream>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "Input.hpp"
#include "Output.hpp"
#include "RungeKutta.hpp"
#include "SaintVenantSolver.hpp"

namespace mfsv {

struct RunOptions {
  int precision = MFSV_F64;
  int device = 0;
  bool device_loop = true;  // inner loop on the device (false: one host round trip per timestep)
  int graph = 1;
  bool write_measures = true;
  bool quiet = false;
  int tile_rows = 0;
  int32_t flags = 0;        // MFSV_E* extensions (0: the reference's semantics)
};

template <typename T>
class ConfiguredModel {
 public:
  using Solver = SaintVenantSolver<double, T>;
  using Scheme = RungeKuttaSolver<Solver>;

  ConfiguredModel(GlobalConfig& gc, const RunOptions& ro)
      : gc_(gc), ro_(ro), mesh_(MeshGeometry::from_config(gc.mesh_configuration())),
        parser_(gc.scheme_configuration()) {
    const Config& sc = gc.scheme_configuration();
    const std::string type = sc.get_value<std::string>("runge-kutta");
    if (type != "runge-kutta") throw std::runtime_error("Unknown scheme type: \"" + type + "\"");
    method_ = sc.get<std::string>("method", "classic");
    tparams_ = time_parameters_from(sc, parser_);

    FieldGenerator fg(gc_, rasters_, mesh_);
    zb_ = fg("z_bed", 0.0);
    h0_ = fg("h", 0.0);
    u0_ = fg("u", 0.0);
    v0_ = fg("v", 0.0);

    SaintVenantOptions opt;
    opt.precision = ro.precision;
    opt.device = ro.device;
    opt.flags = ro.flags;
    opt.tile_rows = ro.tile_rows;
    opt.device_loop = ro.device_loop;
    opt.graph = ro.graph;
    opt.manning = false;
    source_terms(fg, opt);
    boundaries(opt);

    for (const auto& name : gc_.output_files_list())
      outputs_.emplace_back(gc_, name, parser_, *tparams_);
    const Config& mc = gc_.measure_configuration();
    for (const Config* m : mc.equal_range("h-point"))
      measures_.emplace_back(gc_, *m, mesh_, tparams_->num_steps());

    // the configuration is complete: the device context comes last
    const Cartesian2DMesh cm(mesh_.nx(), mesh_.ny(), mesh_.dx(), mesh_.dy(), mesh_.origin_x(),
                             mesh_.origin_y());
    scheme_ = std::make_shared<Scheme>(tparams_, method_, cm, zb_, h0_, u0_, v0_, opt);
    solver_ = scheme_->solver();

    solver_->set_end_of_step_hook([this](double t) { end_of_step(t); });
    solver_->set_output_hook([this](double t) { do_outputs(t); });
    check_files();
  }

  void solve() {
    std::shared_ptr<TemporalScheme> base = scheme_;
    base->solve();
    if (ro_.write_measures)
      for (const auto& m : measures_) m.output();
  }

  Scheme& scheme() { return *scheme_; }
  Solver& solver() { return *solver_; }
  const MeshGeometry& mesh() const { return mesh_; }
  const std::vector<std::string>& written() const { return written_; }
  const std::vector<HPointMeasure>& measures() const { return measures_; }
  const std::string& method() const { return method_; }

 private:
  // ---- source terms (SourceTerm.cpp:30-50)
  void source_terms(FieldGenerator& fg, SaintVenantOptions& opt) {
    const size_t n = mesh_.cell_count();
    int last_rank = -1;
    std::map<std::string, int> seen;
    for (const Config* c : gc_.source_term_configurations()) {
      const std::string type = c->data();
      int rank;
      if (++seen[type] > 1)
        throw std::runtime_error("source term \"" + type + "\" is configured twice (the reference "
                                 "would apply it twice from the same fields)");
      if (type == "manning roughness") {
        rank = 0;
        // ManningRoughnessSourceTerm.cpp:95-109 and .hpp:129-140
        const auto ns = fg("n_shallow", c->get<double>("default shallow n", 0.03));
        const auto nd = fg("n_deep", c->get<double>("default deep n", 0.3));
        const auto ds = fg("d_shallow", c->get<double>("default shallow depth", 0.1));
        const auto dd = fg("d_deep", c->get<double>("default deep depth", 0.3));
        manning_fields_ = {ns, nd, ds, dd};
        opt.manning = true;
        auto uniform = [](const std::vector<double>& a) {
          for (double x : a)
            if (!(x == a[0])) return false;
          return true;
        };
        opt.n_shallow = ns[0]; opt.n_deep = nd[0]; opt.d_shallow = ds[0]; opt.d_deep = dd[0];
        if (uniform(ns) && uniform(nd) && uniform(ds) && uniform(dd)) {
          // scalars
        } else if (ns == nd && uniform(ds) && uniform(dd)) {
          opt.manning_n = ns;  // n(h) == n wherever n_shallow == n_deep
        } else {
          opt.manning_n = ns;
          opt.manning_n_deep = nd;
          opt.manning_d_shallow = ds;
          opt.manning_d_deep = dd;
        }
      } else if (type == "energy loss") {
        rank = 1;
        eloss_on_ = true;
        // EnergyLossSourceTerm.cpp:92-109: constant models on cell
        // selections, or the fields fdx / fdy
        if (c->count("model") > 0) {
          std::vector<double> fx(n, 0.0), fy(n, 0.0);
          for (const Config* m : c->equal_range("model")) {
            if (m->data() != "constant")
              throw std::runtime_error("Unknown energy loss model type: \"" + m->data() + "\"");
            const MeshSelection sel = MeshSelection::from_config(mesh_, m->get_child("cells"));
            const double ax = m->get<double>("fx", 0.0), ay = m->get<double>("fy", 0.0);
            sel.for_each(n, [&](size_t i) { fx[i] += ax; fy[i] += ay; });
          }
          opt.eloss_fx = fx;
          opt.eloss_fy = fy;
        } else {
          opt.eloss_fx = fg("fdx", c->get<double>("default fx/m", 0.0));
          opt.eloss_fy = fg("fdy", c->get<double>("default fy/m", 0.0));
          eloss_check_ = {opt.eloss_fx, opt.eloss_fy};
        }
      } else if (type == "infiltration") {
        rank = 2;
        // InfiltrationSourceTerm.cpp:68-80
        opt.infil_rate = fg("IR", c->get<double>("default infiltration rate", 1e-6));
        opt.infil_capacity = fg("IC", c->get<double>("default infiltration capacity", 0.1));
        infil_check_ = {opt.infil_rate, opt.infil_capacity};
      } else if (type == "eddy viscosity") {
        rank = 3;
        // EddyViscositySourceTerm.cpp:112-135
        const std::string vt = c->get<std::string>("viscosity type", "constant");
        if (vt == "constant") {
          opt.eddy = 1;
          opt.eddy_mu = fg("mu", c->get<double>("default mu", 1.0));
        } else if (vt == "smagorinsky") {
          opt.eddy = 2;
          opt.mu_const = c->get<double>("constant coefficient", 0.05);
          opt.mu_smag = c->get<double>("dynamic coefficient", 0.5);
        } else {
          throw std::runtime_error("Unknown viscosity type: \"" + vt + "\"");
        }
      } else {
        throw std::runtime_error("Unknown source term type: \"" + type + "\"");
      }
      if (rank < last_rank && !ro_.quiet)
        std::cerr << "note: source terms apply in the order manning roughness, energy loss, "
                     "infiltration, eddy viscosity; the configured order changes only the "
                     "rounding of the velocity-source sum" << std::endl;
      last_rank = std::max(last_rank, rank);
    }
  }

  // ---- boundaries (BoundarySourceTerm.cpp:28-45, *BoundarySourceTerm.hpp:
  // 114-148)
  void boundaries(SaintVenantOptions& opt) {
    int last_rank = -1;
    std::map<std::string, int> seen;
    for (const Config* c : gc_.boundary_configurations()) {
      const std::string type = c->data();
      int rank;
      BoundaryEntry::Kind kind;
      if (type == "discharge") { rank = 0; kind = BoundaryEntry::Discharge; }
      else if (type == "head") { rank = 1; kind = BoundaryEntry::Head; }
      else if (type == "stage") { rank = 2; kind = BoundaryEntry::Stage; }
      else throw std::runtime_error("Unknown boundary type: \"" + type + "\"");
      if (++seen[type] > 1)
        throw std::runtime_error("more than one \"boundary == " + type + "\" section: the reference "
                                 "reads the first section once per section; put all entries in one");
      if (rank < last_rank)
        throw std::runtime_error("boundary sections must come in the order discharge, head, stage "
                                 "(the order in which they act on dh/dt)");
      last_rank = rank;
      for (const auto& kv : *c) {
        BoundaryEntry b;
        b.kind = kind;
        b.name = kv.second.data();
        if (kv.first == "constant") {
          b.constant = true;
          b.value = kv.second.get<double>("value");
        } else if (kv.first == "time series") {
          b.constant = false;
          const std::string sname = kv.second.get<std::string>("values");
          b.series = series_.get(gc_, parser_, sname);
          series_used_.push_back(sname);
        } else {
          continue;  // other keys are not boundary entries
        }
        const MeshSelection sel = MeshSelection::from_config(mesh_, kv.second.get_child("cells"));
        b.global = sel.is_global();
        b.cells = sel.ids();
        if (b.constant) constant_sels_.push_back({kind, sel});
        opt.boundaries.push_back(std::move(b));
      }
    }
  }

  // ---- check files written at set-up
  void check_files() {
    {  // Constants.cpp:42-43: z_bed and its Minmod3 slopes
      CheckFile cf(gc_, "constants");
      if (cf.enabled()) {
        mfsv_ctx* c = solver_->context();
        check(mfsv_diagnose(c, 0, 0.0, 1.0, MFSV_DIAG_BED_SLOPES), c, "mfsv_diagnose");
        const size_t n = mesh_.cell_count();
        std::vector<double> zx(n), zy(n);
        check(mfsv_download(c, MFSV_FIELD_DZDX_BED, zx.data(), int64_t(n)), c, "mfsv_download");
        check(mfsv_download(c, MFSV_FIELD_DZDY_BED, zy.data(), int64_t(n)), c, "mfsv_download");
        cf.output_fields(mesh_, {"z_bed", "dzdx_bed", "dzdy_bed"}, {&zb_, &zx, &zy});
      }
    }
    if (!manning_fields_.empty()) {
      CheckFile cf(gc_, "manning");
      cf.output_fields(mesh_, {"n_shallow", "n_deep", "d_shallow", "d_deep"},
                       {&manning_fields_[0], &manning_fields_[1], &manning_fields_[2],
                        &manning_fields_[3]});
    }
    if (!eloss_check_.empty()) {
      CheckFile cf(gc_, "energy-loss-coefficients");
      cf.output_fields(mesh_, {"fdx", "fdy"}, {&eloss_check_[0], &eloss_check_[1]});
    }
    if (!infil_check_.empty()) {
      CheckFile cf(gc_, "infiltration");
      cf.output_fields(mesh_, {"IR", "IC"}, {&infil_check_[0], &infil_check_[1]});
    }
    for (const auto& s : series_used_) {
      CheckFile cf(gc_, "time series");
      cf.output_series(s, series_.get(gc_, parser_, s));
    }
  }

  // qbdy / hbdy check files of a step start (the boundaries' endpoint fields
  // as they are set for the step) and the constant entries' selections
  void boundary_check_files(double t) {
    CheckFile q(gc_, "qbdy"), h(gc_, "hbdy"), qs(gc_, "qbdy_sel"), hs(gc_, "hbdy_sel");
    if (!q.enabled() && !h.enabled() && !qs.enabled() && !hs.enabled()) return;
    const auto e = solver_->endpoints(t);
    const size_t n = mesh_.cell_count();
    auto cells = [&](const std::vector<double>& f, double scalar) {
      return f.empty() ? std::vector<double>(n, scalar) : f;
    };
    bool any_q = false, any_h = false, any_s = false;
    for (const auto& cs : constant_sels_) {
      any_q |= cs.kind == BoundaryEntry::Discharge;
      any_h |= cs.kind == BoundaryEntry::Head;
      any_s |= cs.kind == BoundaryEntry::Stage;
      (cs.kind == BoundaryEntry::Discharge ? qs : hs).output_selection(cs.sel, n);
    }
    (void)any_q;
    (void)any_h;
    (void)any_s;
    if (q.enabled()) {
      const auto a = cells(e.qc0, e.q0), b = cells(e.qc1, e.q1);
      q.output_fields(mesh_, {"qbdy0", "qbdy1"}, {&a, &b});
    }
    if (h.enabled()) {
      const bool stage = !e.sc0.empty();
      const auto a = stage ? e.sc0 : cells(e.hc0, e.hb0), b = stage ? e.sc1 : cells(e.hc1, e.hb1);
      h.output_fields(mesh_, {"hbdy0", "hbdy1"}, {&a, &b});
    }
  }

  // ---- end of step: measures (Solver.hpp:90-98)
  void end_of_step(double t) {
    ++steps_started_;
    boundary_check_files(t);
    if (measures_.empty()) return;
    const size_t n = mesh_.cell_count();
    std::vector<double> h(n);
    mfsv_ctx* c = solver_->context();
    check(mfsv_download(c, MFSV_FIELD_H, h.data(), int64_t(n)), c, "mfsv_download");
    for (auto& m : measures_) m.record(t, h[m.cell()]);
  }

  // ---- outputs (RungeKuttaSolver::do_outputs, RungeKutta.hpp:446-452)
  void do_outputs(double t) {
    diag_valid_ = false;
    for (auto& of : outputs_) {
      const std::string fn = of.timed_output(
          t, mesh_, [this](const std::string& f, MeshPart p) { return field(f, p); });
      if (!fn.empty()) {
        written_.push_back(fn);
        if (!ro_.quiet) std::cout << "Writing output file \"" << of.name() << "\" to \"" << fn << "\"" << std::endl;
      }
    }
  }

  // The fields the reference's solver offers by name (Solver.hpp:139-161):
  // stage, the state, the source terms' fields, the face fluxes.
  std::optional<std::vector<double>> field(const std::string& name, MeshPart part) {
    mfsv_ctx* c = solver_->context();
    const size_t n = mesh_.cell_count(), nf = mesh_.face_count();
    auto get = [&](int id, size_t count) {
      std::vector<double> v(count);
      check(mfsv_download(c, id, v.data(), int64_t(count)), c, "mfsv_download");
      return v;
    };
    if (part == MeshPart::Cells) {
      if (name == "stage") return get(MFSV_FIELD_STAGE, n);
      if (name == "h") return get(MFSV_FIELD_H, n);
      if (name == "u") return get(MFSV_FIELD_U, n);
      if (name == "v") return get(MFSV_FIELD_V, n);
      if (!manning_fields_.empty() && (name == "mannings_n" || name == "friction_slope")) {
        if (!stage_evaluated()) return std::vector<double>(n, 0.0);
        diagnose();
        return get(name == "mannings_n" ? MFSV_FIELD_MANNINGS_N : MFSV_FIELD_FRICTION_SLOPE, n);
      }
      if (has_eloss() && (name == "fx" || name == "fy"))
        return get(name == "fx" ? MFSV_FIELD_ELOSS_FX : MFSV_FIELD_ELOSS_FY, n);
      return std::nullopt;
    }
    if (part == MeshPart::Faces) {
      static const std::map<std::string, int> ff = {{"flux-h", MFSV_FIELD_FLUX_H},
                                                    {"flux-u", MFSV_FIELD_FLUX_U},
                                                    {"flux-v", MFSV_FIELD_FLUX_V},
                                                    {"flux-z", MFSV_FIELD_FLUX_Z}};
      const auto it = ff.find(name);
      if (it == ff.end()) return std::nullopt;
      if (!stage_evaluated()) return std::vector<double>(nf, 0.0);
      diagnose();
      return get(it->second, nf);
    }
    return std::nullopt;
  }
  bool stage_evaluated() const {
    return scheme_->total_timesteps() + scheme_->total_repeated_timesteps() > 0;
  }
  bool has_eloss() const { return eloss_on_; }
  void diagnose() {
    if (diag_valid_) return;
    mfsv_ctx* c = solver_->context();
    check(mfsv_diagnose(c, 1, 0.0, 1.0, MFSV_DIAG_FLUXES | MFSV_DIAG_MANNING), c, "mfsv_diagnose");
    diag_valid_ = true;
  }

  GlobalConfig& gc_;
  RunOptions ro_;
  MeshGeometry mesh_;
  TimeParser parser_;
  std::string method_;
  std::shared_ptr<TimeParameters<double>> tparams_;
  RasterDatabase rasters_;
  TimeSeriesDatabase series_;
  std::vector<std::string> series_used_;
  std::vector<double> zb_, h0_, u0_, v0_;
  std::vector<std::vector<double>> manning_fields_;
  std::vector<std::vector<double>> eloss_check_, infil_check_;
  bool eloss_on_ = false;
  struct ConstSel {
    BoundaryEntry::Kind kind;
    MeshSelection sel;
  };
  std::vector<ConstSel> constant_sels_;
  std::shared_ptr<Scheme> scheme_;
  std::shared_ptr<Solver> solver_;
  std::vector<MultiFieldOutputFile> outputs_;
  std::vector<HPointMeasure> measures_;
  std::vector<std::string> written_;
  size_t steps_started_ = 0;
  bool diag_valid_ = false;
};

}  // namespace mfsv
