// SaintVenantSolver.hpp -- the Saint-Venant solver of the C++ mirror, over
// the C ABI of libmfsv_hip.so.
//
// Mirrors SaintVenantSolver<TT, T, Mesh> (SaintVenant/Solver.hpp:30-172,
// Solver.cpp:36-129) with the state of SaintVenantState (State.hpp:24-156:
// primitive h, u, v), the Manning source (ManningRoughnessSourceTerm.cpp:49-109)
// and the discharge / head boundaries (DischargeBoundarySourceTerm.hpp:114-148,
// HeadBoundarySourceTerm.cpp:43-66) with their time series
// (Input/TimeSeries.cpp:27-46).
//
// The device executes every RK stage of a timestep as one fused kernel
// launch per stage (slopes, faces, cell pass, sources, boundaries and the RK
// accumulation together).  The reference's per-stage calls keep their
// meaning: update_dUdt(s, t, dt) with s = 0 opens a timestep at local time t,
// update_y(n, dt) for the last row runs the fused stages of that timestep,
// and control_number() returns the control number of its result
// (State.cpp:199-229).  accept() makes that result the current state
// (RungeKutta.hpp:428-431, a buffer flip on the device).  dUdt_of_state()
// evaluates dU/dt of the current state alone, for inspection.
#pragma once

#include <cmath>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "DataArray.hpp"
#include "RungeKutta.hpp"
#include "TemporalScheme.hpp"

namespace mfsv {

// Piecewise-linear time series (Input/TimeSeries.cpp:27-46, 109-115).  Past
// the last sample the reference reads beyond its arrays; here that is NaN,
// so a series must cover [start, end + step].
class TimeSeries {
 public:
  TimeSeries() = default;
  TimeSeries(std::vector<double> t, std::vector<double> v) : t_(std::move(t)), v_(std::move(v)) {
    if (t_.size() != v_.size() || t_.size() < 2)
      throw std::runtime_error("TimeSeries: need matching times and values, at least 2 samples");
  }
  bool empty() const { return t_.empty(); }
  const std::vector<double>& times() const { return t_; }
  const std::vector<double>& values() const { return v_; }
  double at(double time) const {
    size_t i1 = 1;
    while (i1 < t_.size() && !(t_[i1] > time)) ++i1;
    if (i1 >= t_.size()) return std::numeric_limits<double>::quiet_NaN();
    const size_t i0 = i1 - 1;
    const double dvdt = (v_[i1] - v_[i0]) / (t_[i1] - t_[i0]);
    return v_[i0] + dvdt * (time - t_[i0]);
  }

 private:
  std::vector<double> t_, v_;
};

// One entry of a boundary section (DischargeBoundarySourceTerm.hpp:114-148,
// HeadBoundarySourceTerm.hpp, StageBoundarySourceTerm.hpp): a constant or a
// time series on a set of cells (ids ascending; `global` = every cell).  Per
// outer step a discharge entry adds its endpoints to the cells' endpoint
// fields, a head or stage entry sets them.
struct BoundaryEntry {
  enum Kind { Discharge, Head, Stage } kind = Discharge;
  std::string name;
  bool constant = true;
  double value = 0.0;
  TimeSeries series;
  bool global = true;
  std::vector<int64_t> cells;
};

// What the reference reads from its config sections (source terms,
// boundaries, scheme), as plain values; arrays are injected.
struct SaintVenantOptions {
  int precision = MFSV_F64;
  int device = 0;
  int32_t flags = 0;            // MFSV_E* extension bits (0 = the reference's semantics)
  double cn_target = 1.0;
  bool manning = true;
  double n_shallow = 0.03, n_deep = 0.03, d_shallow = 0.1, d_deep = 0.3;
  std::vector<double> manning_n;  // per-cell n (empty = uniform)
  TimeSeries rain;                // discharge boundary, global selection (m/s)
  int64_t head_cols = 0;          // head boundary on columns [0, head_cols), target depth 0
  mfsv_params storm{};            // storm_* fields used when flags has MFSV_E4_STORM
  double h_dry = 1e-3, u_max = 10.0;
  int tile_rows = 0;              // 0 = library default
  // EnergyLossSourceTerm (per-cell fx, fy; empty = off)
  std::vector<double> eloss_fx, eloss_fy;
  // InfiltrationSourceTerm (per-cell rate and initial capacity; empty = off)
  std::vector<double> infil_rate, infil_capacity;
  // EddyViscositySourceTerm: 0 off, 1 per-cell eddy_mu, 2 Smagorinsky
  int eddy = 0;
  std::vector<double> eddy_mu;
  double mu_const = 0.05, mu_smag = 0.5;
  // StageBoundarySourceTerm: water-level series on the cells where
  // stage_cells is non-zero (empty series = off)
  TimeSeries stage;
  std::vector<uint8_t> stage_cells;
  // per-cell n_deep, d_shallow, d_deep (ManningRoughnessSourceTerm.hpp:
  // 129-140); with them manning_n is n_shallow
  std::vector<double> manning_n_deep, manning_d_shallow, manning_d_deep;
  // boundary sections, in the order they apply
  std::vector<BoundaryEntry> boundaries;
  // inner loop of each outer step on the device (mfsv_run_step) instead of
  // one host round trip per timestep; graph: 1 = hipGraph replay
  bool device_loop = false;
  int graph = 1;
};

template <typename TT, typename T>
class SaintVenantSolver {
 public:
  using TimeType = TT;
  using ValueType = T;
  using MeshType = Cartesian2DMesh;

  // State (h, u, v) with host mirrors of the device fields.
  struct State {
    Field<double> h, u, v;  // host copies are FP64 in both precisions
    explicit State(const Cartesian2DMesh& m) : h("h", m), u("u", m), v("v", m) {}
  };

  // Solver.cpp:36-71: mesh, bed, initial state, sources and boundaries.
  SaintVenantSolver(const Cartesian2DMesh& mesh, const std::vector<double>& z_bed,
                    const std::vector<double>& h0, const std::vector<double>& u0,
                    const std::vector<double>& v0, const SaintVenantOptions& opt,
                    size_t n_states, const std::shared_ptr<TimeParameters<TT>>& tparams,
                    const RungeKuttaCoefficientSet& coeffs)
      : mesh_(mesh), opt_(opt), tparams_(tparams), nstages_(n_states - 1),
        ctx_(std::make_shared<Context>(mesh.grid(), opt.precision, opt.device)),
        U_(mesh), zb_("z_bed", mesh), k_(mesh) {
    const size_t n = mesh.cell_count();
    if (z_bed.size() != n || h0.size() != n || u0.size() != n || v0.size() != n)
      throw std::runtime_error("SaintVenantSolver: field sizes must equal the cell count");
    if (n_states != coeffs.nsteps() + 1)
      throw std::logic_error("SaintVenantSolver: n_states must be the scheme's stages + 1");
    mfsv_ctx* c = ctx_->get();
    const std::vector<double> tab = coeffs.matrix();
    check(mfsv_set_scheme(c, int(coeffs.nsteps()), tab.data()), c, "mfsv_set_scheme");
    mfsv_params p = opt.storm;
    p.flags = opt.flags;
    p.cn_target = opt.cn_target;
    p.manning = opt.manning ? 1 : 0;
    p.n_shallow = opt.n_shallow; p.n_deep = opt.n_deep;
    p.d_shallow = opt.d_shallow; p.d_deep = opt.d_deep;
    bool q_any = !opt.rain.empty(), q_cells = false, h_any = false, h_cells = false, s_any = false;
    for (const auto& b : opt.boundaries) {
      if (b.kind == BoundaryEntry::Discharge) { q_any = true; q_cells |= !b.global; }
      if (b.kind == BoundaryEntry::Head) { h_any = true; h_cells |= !b.global; }
      if (b.kind == BoundaryEntry::Stage) s_any = true;
    }
    if (h_any && opt.head_cols > 0)
      throw std::runtime_error("SaintVenantOptions: head_cols and head boundary entries together");
    p.discharge = q_any ? (q_cells ? 2 : 1) : 0;
    p.head = opt.head_cols > 0 ? 1 : (h_any ? (h_cells ? 2 : 1) : 0);
    p.head_x0 = 0; p.head_x1 = opt.head_cols > 0 ? opt.head_cols : (h_any ? mesh.nx() : 0);
    p.h_dry = opt.h_dry; p.u_max = opt.u_max;
    p.energy_loss = opt.eloss_fx.empty() ? 0 : 1;
    p.infiltration = opt.infil_rate.empty() ? 0 : 1;
    p.stage = (opt.stage.empty() && !s_any) ? 0 : 1;
    p.eddy = opt.eddy;
    p.mu_const = opt.mu_const; p.mu_smag = opt.mu_smag;
    check(mfsv_set_params(c, &p), c, "mfsv_set_params");
    params_ = p;
    check(mfsv_set_time(c, tparams->start_time(), tparams->end_time(), tparams->step_duration(),
                        tparams->initial_timestep()), c, "mfsv_set_time");
    if (opt.tile_rows > 0) check(mfsv_set_tile_rows(c, opt.tile_rows), c, "mfsv_set_tile_rows");
    bind_and_upload(zb_.data(), MFSV_FIELD_ZB, z_bed);
    if (!opt.manning_n.empty()) {
      if (opt.manning_n.size() != n) throw std::runtime_error("manning_n size != cell count");
      DataArray<double> nf(n);
      bind_and_upload(nf, MFSV_FIELD_MANNING_N, opt.manning_n);
    }
    const int mfields = int(!opt.manning_n_deep.empty()) + int(!opt.manning_d_shallow.empty()) +
                        int(!opt.manning_d_deep.empty());
    if (mfields != 0 && (mfields != 3 || opt.manning_n.empty()))
      throw std::runtime_error("per-cell Manning needs manning_n and all of n_deep, d_shallow, d_deep");
    auto extra = [&](const std::vector<double>& f, int id) {
      if (f.empty()) return;
      if (f.size() != n) throw std::runtime_error("source-term field size != cell count");
      DataArray<double> a(n);
      bind_and_upload(a, id, f);
    };
    extra(opt.eloss_fx, MFSV_FIELD_ELOSS_FX);
    extra(opt.eloss_fy, MFSV_FIELD_ELOSS_FY);
    extra(opt.infil_rate, MFSV_FIELD_INFIL_RATE);
    extra(opt.infil_capacity, MFSV_FIELD_INFIL_CAP);
    extra(opt.eddy_mu, MFSV_FIELD_EDDY_MU);
    extra(opt.manning_n_deep, MFSV_FIELD_N_DEEP);
    extra(opt.manning_d_shallow, MFSV_FIELD_D_SHALLOW);
    extra(opt.manning_d_deep, MFSV_FIELD_D_DEEP);
    bind_and_upload(U_.h.data(), MFSV_FIELD_H, h0);
    bind_and_upload(U_.u.data(), MFSV_FIELD_U, u0);
    bind_and_upload(U_.v.data(), MFSV_FIELD_V, v0);
    k_.h.data().bind(ctx_, MFSV_FIELD_DHDT);
    k_.u.data().bind(ctx_, MFSV_FIELD_DUDT);
    k_.v.data().bind(ctx_, MFSV_FIELD_DVDT);
  }

  double cn_target() const { return opt_.cn_target; }
  const Cartesian2DMesh& mesh() const { return mesh_; }
  mfsv_ctx* context() const { return ctx_->get(); }

  // Solver.cpp:77-114 for stage s; see the file comment for the fused mapping.
  void update_dUdt(size_t s, TT substep_time, TT timestep) {
    if (s == 0) {
      t_local_ = substep_time;  // c_1 = 0 for every explicit tableau
      dt_ = timestep;
      open_ = true;
    } else if (!open_ || timestep != dt_) {
      throw std::logic_error("update_dUdt: stage " + std::to_string(s) + " outside a timestep");
    }
  }
  // RungeKutta.hpp:371-405: rows 1..n-1 are formed inside the fused stage
  // launches; the last row launches them.
  void update_y(size_t step, TT timestep) {
    if (step < nstages_) return;
    if (!open_ || timestep != dt_) throw std::logic_error("update_y outside a timestep");
    double cn = 0.0;
    check(mfsv_timestep(ctx_->get(), double(t_local_), double(dt_), &cn), ctx_->get(),
          "mfsv_timestep");
    cn_ = cn;
    open_ = false;
  }
  // State.cpp:199-229.  The fused timestep returns the result's control
  // number with it; other states (and other timesteps) are evaluated on the
  // device.
  double control_number(size_t state, TT dt) const {
    if (state == nstages_ && double(dt) == double(dt_) && !open_) return cn_;
    double cn = 0.0;
    check(mfsv_control_number(ctx_->get(), int(state), double(dt), &cn), ctx_->get(),
          "mfsv_control_number");
    return cn;
  }
  void accept() { check(mfsv_accept(ctx_->get()), ctx_->get(), "mfsv_accept"); }

  // Solver.cpp:119-129: boundary endpoints for [time_now, time_now + step).
  void start_new_step(TT time_now) {
    const Endpoints e = endpoints(double(time_now));
    const size_t nc = mesh_.cell_count();
    mfsv_ctx* c = ctx_->get();
    auto up = [&](int field, const std::vector<double>& v, std::vector<double>& last) {
      if (v.empty() || v == last) return;
      check(mfsv_upload(c, field, v.data(), int64_t(nc)), c, "mfsv_upload");
      last = v;
    };
    up(MFSV_FIELD_QBDY0, e.qc0, up_[0]);
    up(MFSV_FIELD_QBDY1, e.qc1, up_[1]);
    up(MFSV_FIELD_HBDY0, e.hc0, up_[2]);
    up(MFSV_FIELD_HBDY1, e.hc1, up_[3]);
    up(MFSV_FIELD_SBDY0, e.sc0, up_[4]);
    up(MFSV_FIELD_SBDY1, e.sc1, up_[5]);
    last_ = e;
    // the device loop starts once the step's per-cell fields are in place
    if (opt_.device_loop && !begun_) begin_device_run();
    check(mfsv_start_new_step(c, double(time_now), e.q0, e.q1, e.hb0, e.hb1), c,
          "mfsv_start_new_step");
  }
  // Endpoint values of one outer step.  Scalars: global discharge and head;
  // fields: per-cell discharge (summed), head and stage (set; NaN = none).
  struct Endpoints {
    double q0 = 0.0, q1 = 0.0;
    double hb0 = std::numeric_limits<double>::quiet_NaN(), hb1 = std::numeric_limits<double>::quiet_NaN();
    std::vector<double> qc0, qc1, hc0, hc1, sc0, sc1;
  };
  Endpoints endpoints(double time_now) const {
    Endpoints e;
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const double t1 = time_now + double(tparams_->step_duration());
    const size_t nc = mesh_.cell_count();
    if (!opt_.rain.empty()) {
      e.q0 = 0.0 + opt_.rain.at(time_now);
      e.q1 = 0.0 + opt_.rain.at(t1);
    }
    if (opt_.head_cols > 0) e.hb0 = e.hb1 = 0.0;
    const mfsv_params& p = params_;
    if (p.discharge == 2) { e.qc0.assign(nc, 0.0); e.qc1.assign(nc, 0.0); }
    if (p.head == 2) { e.hc0.assign(nc, nan); e.hc1.assign(nc, nan); }
    if (p.stage && opt_.stage.empty()) { e.sc0.assign(nc, nan); e.sc1.assign(nc, nan); }
    for (const auto& b : opt_.boundaries) {
      const double v0 = b.constant ? b.value : b.series.at(time_now);
      const double v1 = b.constant ? b.value : b.series.at(t1);
      auto each = [&](auto f) {
        if (b.global) for (size_t i = 0; i < nc; ++i) f(i);
        else for (int64_t i : b.cells) f(size_t(i));
      };
      switch (b.kind) {
        case BoundaryEntry::Discharge:
          if (p.discharge == 2) each([&](size_t i) { e.qc0[i] += v0; e.qc1[i] += v1; });
          else { e.q0 += v0; e.q1 += v1; }
          break;
        case BoundaryEntry::Head:
          if (p.head == 2) each([&](size_t i) { e.hc0[i] = v0; e.hc1[i] = v1; });
          else { e.hb0 = v0; e.hb1 = v1; }
          break;
        case BoundaryEntry::Stage:
          each([&](size_t i) { e.sc0[i] = v0; e.sc1[i] = v1; });
          break;
      }
    }
    if (!opt_.stage.empty()) {  // StageBoundarySourceTerm.hpp:114-142
      const double a0 = opt_.stage.at(time_now), a1 = opt_.stage.at(t1);
      e.sc0.assign(nc, nan);
      e.sc1.assign(nc, nan);
      for (size_t i = 0; i < nc; ++i) {
        const bool on = opt_.stage_cells.empty() || opt_.stage_cells[i];
        e.sc0[i] = on ? a0 : nan;
        e.sc1[i] = on ? a1 : nan;
      }
    }
    return e;
  }
  const Endpoints& last_endpoints() const { return last_; }

  // The inner loop of the current outer step on the device; returns the
  // accepted and repeated timesteps it took.
  std::pair<size_t, size_t> run_device_step() {
    mfsv_ctx* c = ctx_->get();
    mfsv_stats a{}, b{};
    check(mfsv_get_stats(c, &a), c, "mfsv_get_stats");
    check(mfsv_run_step(c, opt_.graph), c, "mfsv_run_step");
    check(mfsv_get_stats(c, &b), c, "mfsv_get_stats");
    return {size_t(b.accepted - a.accepted), size_t(b.rejected - a.rejected)};
  }
  bool device_loop() const { return opt_.device_loop; }
  void set_end_of_step_hook(std::function<void(double)> f) { eos_hook_ = std::move(f); }
  void set_output_hook(std::function<void(double)> f) { out_hook_ = std::move(f); }

  // Solver.hpp:90-98: the stage field z_bed + h, then the measures.
  void end_of_step(TT time_now) {
    check(mfsv_update_stage_field(ctx_->get()), ctx_->get(), "mfsv_update_stage_field");
    if (eos_hook_) eos_hook_(double(time_now));
  }
  void do_outputs(TT time_now) {
    if (out_hook_) out_hook_(double(time_now));
  }

  // The current state U0, downloaded into the host mirrors.
  State& state() {
    U_.h.data().move_to_host();
    U_.u.data().move_to_host();
    U_.v.data().move_to_host();
    return U_;
  }
  // dU/dt of the current state at stage time t with timestep dt.
  State& dUdt_of_state(TT t, TT dt) {
    check(mfsv_dudt(ctx_->get(), double(t), double(dt), 0), ctx_->get(), "mfsv_dudt");
    k_.h.data().move_to_host();
    k_.u.data().move_to_host();
    k_.v.data().move_to_host();
    return k_;
  }
  // FP64 volume sum(h) dx dy of the current state (Output/Measure.hpp:137-153).
  double volume() const {
    double v = 0.0;
    check(mfsv_volume(ctx_->get(), &v), ctx_->get(), "mfsv_volume");
    return v;
  }
  // The device controller's counters and FP64 volume integrals (the volume
  // when the device run began and the integrated net source rate since).
  mfsv_stats device_stats() const {
    mfsv_stats s{};
    check(mfsv_get_stats(ctx_->get(), &s), ctx_->get(), "mfsv_get_stats");
    return s;
  }

 private:
  void bind_and_upload(DataArray<double>& a, int field, const std::vector<double>& values) {
    a = DataArray<double>(values.size());
    for (size_t i = 0; i < values.size(); ++i) a[i] = values[i];
    a.bind(ctx_, field);
    a.move_to_device();
  }

  // The device-controlled loop needs the scalar endpoints of every outer step
  // up front (a forcing table) and the controller initialised once.
  void begin_device_run() {
    mfsv_ctx* c = ctx_->get();
    const size_t ns = tparams_->num_steps();
    std::vector<double> table(4 * ns);
    for (size_t i = 0; i < ns; ++i) {
      const double t = double(tparams_->start_time()) + double(i) * double(tparams_->step_duration());
      const Endpoints e = endpoints(t);
      table[4 * i] = e.q0; table[4 * i + 1] = e.q1;
      table[4 * i + 2] = e.hb0; table[4 * i + 3] = e.hb1;
    }
    check(mfsv_set_forcing_table(c, int64_t(ns), table.data()), c, "mfsv_set_forcing_table");
    check(mfsv_run_begin(c), c, "mfsv_run_begin");
    begun_ = true;
  }

  Cartesian2DMesh mesh_;
  SaintVenantOptions opt_;
  mfsv_params params_{};
  bool begun_ = false;
  std::vector<double> up_[6];
  Endpoints last_;
  std::function<void(double)> eos_hook_, out_hook_;
  std::shared_ptr<TimeParameters<TT>> tparams_;
  size_t nstages_;
  std::shared_ptr<Context> ctx_;
  State U_;
  Field<double> zb_;
  State k_;
  TT t_local_ = 0, dt_ = 0;
  bool open_ = false;
  double cn_ = 0.0;
};

}  // namespace mfsv
