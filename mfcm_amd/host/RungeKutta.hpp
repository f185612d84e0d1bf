// RungeKutta.hpp -- explicit Runge-Kutta schemes of the C++ mirror.
//
// RungeKuttaCoefficientSet: Butcher rows as in TemporalScheme/RungeKutta.hpp:59-123
// (row i-1 = {c_i, a_i1, ...}, last row = {0, b_1, ...}), the named sets of
// TemporalScheme/RungeKutta.cpp:54-116 and the generic order-2/3 families of
// RungeKutta.cpp:33-52.
// RungeKuttaTemporalScheme::do_timestep: RungeKutta.cpp:143-167 (substeps,
// then the a-posteriori controller on the control number).
// RungeKuttaSolver<Solver>: RungeKutta.hpp:268-463, the glue between the
// scheme and a solver exposing update_dUdt / update_y / control_number /
// accept / end_of_step / start_new_step.
#pragma once

#include <cmath>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "TemporalScheme.hpp"

namespace mfsv {

class RungeKuttaCoefficientSet {
 public:
  RungeKuttaCoefficientSet() = default;
  explicit RungeKuttaCoefficientSet(std::vector<std::vector<double>> a) : a_(std::move(a)) {
    if (a_.size() < 2) throw std::logic_error("RungeKuttaCoefficientSet: need at least 2 rows");
    for (size_t row = 0; row + 1 < a_.size(); ++row)
      if (a_[row].size() != row + 1)
        throw std::logic_error("RungeKuttaCoefficientSet: row lengths must be 1, 2, ...");
  }
  // Generic second- and third-order families in the parameter alpha.
  RungeKuttaCoefficientSet(size_t order, double alpha) {
    a_.push_back({0.0});
    a_.push_back({alpha, alpha});
    if (order == 2) {
      const double w2 = 1.0 / (2.0 * alpha);
      a_.push_back({0.0, 1.0 - w2, w2});
    } else if (order == 3) {
      const double r = (1.0 - alpha) / (alpha * (3.0 * alpha - 2.0));
      a_.push_back({1.0, 1.0 + r, -r});
      a_.push_back({0.0, 0.5 - 1.0 / (6.0 * alpha), 1.0 / (6.0 * alpha * (1.0 - alpha)),
                    (2.0 - 3.0 * alpha) / (6.0 * (1.0 - alpha))});
    } else {
      std::cerr << "No generic Runge Kutta parameters for order " << order << std::endl;
      throw std::runtime_error("No generic Runge Kutta parameters available.");
    }
  }

  size_t nsteps() const { return a_.size() - 1; }
  double a(size_t i, size_t j) const { return a_[i - 1][j]; }
  double b(size_t i) const { return a_.back()[i]; }
  double c(size_t i) const { return a_[i - 1][0]; }

  // (nsteps+1)^2 row-major matrix, the layout mfsv_set_scheme takes.
  std::vector<double> matrix() const {
    const size_t n = nsteps() + 1;
    std::vector<double> m(n * n, 0.0);
    for (size_t i = 0; i < a_.size(); ++i)
      for (size_t j = 0; j < a_[i].size(); ++j) m[i * n + j] = a_[i][j];
    return m;
  }

  std::ostream& butcher_tableau(std::ostream& os) const {
    for (size_t i = 1; i <= nsteps(); ++i) {
      os << std::setw(10) << c(i) << " |";
      for (size_t j = 1; j < i; ++j) os << std::setw(10) << a(i + 1, j);
      os << "\n";
    }
    os << std::setw(10) << "" << " |";
    for (size_t j = 1; j <= nsteps(); ++j) os << std::setw(10) << b(j);
    return os << "\n";
  }

 private:
  std::vector<std::vector<double>> a_;
};

inline const std::map<std::string, RungeKuttaCoefficientSet>& runge_kutta_named_coefficient_sets() {
  static const std::map<std::string, RungeKuttaCoefficientSet> sets{
      {"euler", RungeKuttaCoefficientSet({{0.0}, {0.0, 1.0}})},
      {"midpoint", RungeKuttaCoefficientSet({{0.0}, {0.5, 0.5}, {0.0, 0.0, 1.0}})},
      {"heun2", RungeKuttaCoefficientSet({{0.0}, {1.0, 1.0}, {0.0, 0.5, 0.5}})},
      {"ralston2", RungeKuttaCoefficientSet({{0.0}, {2.0 / 3.0, 2.0 / 3.0}, {0.0, 0.25, 0.75}})},
      {"kutta3", RungeKuttaCoefficientSet({{0.0}, {0.5, 0.5}, {1.0, -1.0, 2.0},
                                           {0.0, 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}})},
      {"heun3", RungeKuttaCoefficientSet({{0.0}, {1.0 / 3.0, 1.0 / 3.0}, {2.0 / 3.0, 0.0, 2.0 / 3.0},
                                          {0.0, 0.25, 0.0, 0.75}})},
      {"wray3", RungeKuttaCoefficientSet({{0.0}, {8.0 / 15.0, 8.0 / 15.0},
                                          {2.0 / 3.0, 0.25, 5.0 / 12.0}, {0.0, 0.25, 0.0, 0.75}})},
      {"ralston3", RungeKuttaCoefficientSet({{0.0}, {0.5, 0.5}, {0.75, 0.0, 0.75},
                                             {0.0, 2.0 / 9.0, 1.0 / 3.0, 4.0 / 9.0}})},
      {"ssprk3", RungeKuttaCoefficientSet({{0.0}, {1.0, 1.0}, {0.5, 0.25, 0.25},
                                           {0.0, 1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0}})},
      {"classic", RungeKuttaCoefficientSet({{0.0}, {0.5, 0.5}, {0.5, 0.0, 0.5}, {1.0, 0.0, 0.0, 1.0},
                                            {0.0, 1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0}})},
      {"ralston4", RungeKuttaCoefficientSet({{0.0}, {0.4, 0.4}, {0.45573725, 0.29697761, 0.15875964},
                                             {1.0, 0.21810040, -3.05096516, 3.83286476},
                                             {0.0, 0.17476028, -0.55148066, 1.20553560, 0.17118478}})},
      {"3/8rule", RungeKuttaCoefficientSet({{0.0}, {1.0 / 3.0, 1.0 / 3.0}, {2.0 / 3.0, -1.0 / 3.0, 1.0},
                                            {1.0, 1.0, -1.0, 1.0},
                                            {0.0, 1.0 / 8.0, 3.0 / 8.0, 3.0 / 8.0, 1.0 / 8.0}})},
  };
  return sets;
}

inline RungeKuttaCoefficientSet named_coefficients(const std::string& name) {
  const auto& sets = runge_kutta_named_coefficient_sets();
  auto it = sets.find(name);
  if (it == sets.end()) {
    std::cerr << "Unknown Runge Kutta scheme name: " << std::quoted(name) << std::endl;
    throw std::runtime_error("Unknown name for Runge Kutta scheme.");
  }
  return it->second;
}

template <typename T>
class RungeKuttaTemporalScheme : public TypedTemporalScheme<T> {
 public:
  using TimeType = T;
  RungeKuttaTemporalScheme(const std::shared_ptr<TimeParameters<T>>& tparams,
                           const RungeKuttaCoefficientSet& coeffs)
      : TypedTemporalScheme<T>(tparams), coeffs_(coeffs) {}
  RungeKuttaTemporalScheme(const std::shared_ptr<TimeParameters<T>>& tparams,
                           const std::string& named_coeffs)
      : TypedTemporalScheme<T>(tparams), coeffs_(named_coefficients(named_coeffs)) {}

  const RungeKuttaCoefficientSet& coeffs() const { return coeffs_; }
  // Control-number target: 1.0 is the reference's hard-coded value
  // (RungeKutta.cpp:154); other targets scale the control number first.
  void set_cn_target(double t) {
    if (!(t > 0)) throw std::logic_error("cn target must be positive");
    cn_target_ = t;
  }

  virtual double get_latest_control_number(const TimeType& dt) = 0;

 protected:
  using timestep_result = typename TypedTemporalScheme<T>::timestep_result;
  virtual void update_k(const size_t& step, const TimeType& substep_time, const TimeType& timestep) = 0;
  virtual void update_y(const size_t& step, const TimeType& timestep) = 0;

  // RungeKutta.cpp:143-167
  timestep_result do_timestep(const TimeType& local_time, const TimeType& dt) override {
    for (size_t substep = 1; substep <= coeffs_.nsteps(); ++substep) {
      const TimeType substep_time = local_time + coeffs_.c(substep) * dt;
      this->update_k(substep, substep_time, dt);
      this->update_y(substep, dt);
    }
    const double raw = this->get_latest_control_number(dt);
    const double cn = cn_target_ == 1.0 ? raw : raw / cn_target_;
    if (cn > 1.0) {
      if (cn > 5.0) return {TimeType(dt / 5.0), true};
      return {TimeType(dt / (cn * 1.1)), true};
    } else if (cn > 0.9) {
      return {TimeType(dt), false};
    }
    return {TimeType(dt / cn), false};
  }

  RungeKuttaCoefficientSet coeffs_;
  double cn_target_ = 1.0;
};

// The glue of RungeKutta.hpp:268-463.  Solver must provide TimeType,
// a constructor (args..., n_states, tparams, coeffs), update_dUdt(s, t, dt),
// update_y(i, dt), control_number(i, dt), accept(), end_of_step(t),
// do_outputs(t) and start_new_step(t).
template <typename Solver>
class RungeKuttaSolver : public RungeKuttaTemporalScheme<typename Solver::TimeType> {
 public:
  using TimeType = typename Solver::TimeType;
  using SolverType = Solver;

  template <typename... Args>
  RungeKuttaSolver(const std::shared_ptr<TimeParameters<TimeType>>& tparams,
                   const RungeKuttaCoefficientSet& coeffs, Args&&... solver_args)
      : RungeKuttaTemporalScheme<TimeType>(tparams, coeffs),
        solver_(std::make_shared<Solver>(std::forward<Args>(solver_args)..., coeffs.nsteps() + 1,
                                         tparams, coeffs)) {
    this->set_cn_target(solver_->cn_target());
  }
  template <typename... Args>
  RungeKuttaSolver(const std::shared_ptr<TimeParameters<TimeType>>& tparams,
                   const std::string& named_coeffs, Args&&... solver_args)
      : RungeKuttaSolver(tparams, named_coefficients(named_coeffs),
                         std::forward<Args>(solver_args)...) {}

  std::shared_ptr<Solver> solver() const { return solver_; }

  // The inner loop on the device when the solver runs it there (the same
  // controller and landing rules, evaluated by the device).
  typename TypedTemporalScheme<TimeType>::step_result step() override {
    if (!solver_->device_loop()) return TypedTemporalScheme<TimeType>::step();
    const auto r = solver_->run_device_step();
    return {r.first, r.second};
  }

  double get_latest_control_number(const TimeType& dt) override {
    return solver_->control_number(this->coeffs().nsteps(), dt);
  }

 protected:
  void update_k(const size_t& step, const TimeType& substep_time, const TimeType& timestep) override {
    solver_->update_dUdt(step - 1, substep_time, timestep);
  }
  void update_y(const size_t& step, const TimeType& timestep) override {
    solver_->update_y(step, timestep);
  }
  void accept_timestep() override { solver_->accept(); }
  void end_of_step(const TimeType& time_now) override { solver_->end_of_step(time_now); }
  void do_outputs(const TimeType& time_now) override { solver_->do_outputs(time_now); }
  void start_new_step(const TimeType& time_now) override { solver_->start_new_step(time_now); }

  std::shared_ptr<Solver> solver_;
};

}  // namespace mfsv
