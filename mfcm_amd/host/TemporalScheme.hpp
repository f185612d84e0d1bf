// TemporalScheme.hpp -- the outer stepping loop of the C++ mirror.
//
// TimeParameters follows Config/TimeParameters.cpp:22-58 (times in seconds;
// the number of steps rounds up unless the duration is a whole number of
// steps within 1e-6).  TemporalScheme / TypedTemporalScheme follow
// TemporalScheme/TemporalScheme.hpp:29-278: fixed-length outer steps, each a
// sequence of adaptive timesteps with repeat-on-reject and the end-of-step
// landing rules (hit the end exactly, or split the last two timesteps 40/60).
#pragma once

#include <cstddef>
#include <iostream>
#include <memory>
#include <stdexcept>

namespace mfsv {

template <typename TT>
class TimeParameters {
 public:
  TimeParameters(TT start, TT end, TT step_duration, TT initial_timestep)
      : start_(start), end_(end), step_(step_duration), dt0_(initial_timestep) {
    if (!(step_duration > 0) || !(initial_timestep > 0) || !(end >= start))
      throw std::runtime_error("TimeParameters: invalid times");
  }
  TT start_time() const { return start_; }
  TT end_time() const { return end_; }
  TT step_duration() const { return step_; }
  TT initial_timestep() const { return dt0_; }
  // TimeParameters.cpp:40-58
  size_t num_steps() const {
    const TT eps = TT(1e-6);
    const TT dur = end_ - start_;
    const size_t n = size_t(dur / step_);
    if (step_ * TT(n) >= dur - eps && step_ * TT(n) <= dur + eps) return n;
    return n + 1;
  }

 private:
  TT start_, end_, step_, dt0_;
};

class TemporalScheme {
 public:
  virtual ~TemporalScheme() = default;
  virtual void solve() = 0;
};

template <typename T>
class TypedTemporalScheme : public TemporalScheme {
 public:
  using TimeType = T;

  struct step_result {
    size_t num_timesteps;
    size_t num_repeated_timesteps;
  };

  explicit TypedTemporalScheme(const std::shared_ptr<TimeParameters<T>>& tparams)
      : time_params_(tparams), dt_(tparams->initial_timestep()) {}

  std::shared_ptr<TimeParameters<T>> time_parameters() const { return time_params_; }
  TimeType current_timestep() const { return dt_; }
  size_t total_timesteps() const { return total_.num_timesteps; }
  size_t total_repeated_timesteps() const { return total_.num_repeated_timesteps; }
  void set_verbose(bool v) { verbose_ = v; }

  // TemporalScheme.hpp:112-149 (a solver may run this loop on the device)
  virtual step_result step() {
    step_result result{0, 0};
    TimeType t_local = 0.0;
    const TimeType step_duration = time_params_->step_duration();
    while (t_local < step_duration) {
      const timestep_result r = this->do_timestep(t_local, dt_);
      if (r.repeat) {
        if (verbose_) std::cout << "Repeating timestep at local time " << t_local << std::endl;
        result.num_repeated_timesteps++;
      } else {
        result.num_timesteps++;
        t_local += dt_;
        this->accept_timestep();
      }
      dt_ = r.new_dt;
      if (t_local >= step_duration) {
        return result;
      } else if (t_local + dt_ > step_duration) {
        dt_ = step_duration - t_local;
      } else if (t_local + 2.0 * dt_ > step_duration) {
        dt_ = (step_duration - t_local) * 0.4;
      }
    }
    throw std::runtime_error("Should not be possible to reach here.");
  }

  // TemporalScheme.hpp:244-276
  void solve() override {
    const TimeType start_time = time_params_->start_time();
    const size_t num_steps = time_params_->num_steps();
    const TimeType step_duration = time_params_->step_duration();
    dt_ = time_params_->initial_timestep();
    total_ = {0, 0};
    for (size_t i = 0; i < num_steps; ++i) {
      const TimeType time_now = start_time + TimeType(i) * step_duration;
      if (verbose_) std::cout << "Starting step " << i << " at time " << time_now << std::endl;
      this->end_of_step(time_now);
      this->do_outputs(time_now);
      this->start_new_step(time_now);
      const step_result r = this->step();
      total_.num_timesteps += r.num_timesteps;
      total_.num_repeated_timesteps += r.num_repeated_timesteps;
    }
    this->do_outputs(start_time + TimeType(num_steps) * step_duration);
  }

 protected:
  struct timestep_result {
    TimeType new_dt;
    bool repeat;
  };
  virtual timestep_result do_timestep(const TimeType& local_time, const TimeType& timestep) = 0;
  virtual void accept_timestep() = 0;
  virtual void end_of_step(const TimeType& time_now) = 0;
  virtual void do_outputs(const TimeType& time_now) = 0;
  virtual void start_new_step(const TimeType& time_now) = 0;

  std::shared_ptr<TimeParameters<T>> time_params_;
  TimeType dt_;
  step_result total_{0, 0};
  bool verbose_ = false;
};

}  // namespace mfsv
