#include "stepper.hh"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace dcopasi {

namespace {

// Dune::FloatCmp with its default relative-weak tolerance (8 DBL_EPSILON of
// the larger magnitude), as the reference compares times
constexpr double kEps = 8.0 * DBL_EPSILON;
bool feq(double a, double b) { return std::fabs(a - b) <= kEps * std::max(std::fabs(a), std::fabs(b)); }
bool flt(double a, double b) { return a < b && !feq(a, b); }
bool fle(double a, double b) { return a < b || feq(a, b); }

constexpr int kStepSizeLimit = 5;

struct Runner {
  Model& m;
  const EvolveOptions& opt;
  EvolveStats& st;
  const std::function<void(double)>& on_step;

  // SimpleAdaptiveStepper::check_dt (stepper.hh:352-363)
  bool dt_allowed(double dt) const {
    if (opt.min_step > 0.0 && flt(std::fabs(dt), std::fabs(opt.min_step))) return false;
    if (opt.max_step > 0.0 && flt(std::fabs(opt.max_step), std::fabs(dt))) return false;
    return true;
  }

  // TimeStepper::do_step (stepper.hh:95-125): one attempt from the model's time
  int attempt(double dt) {
    StepStats s;
    const int rc = m.step(m.time(), dt, &s);
    st.newton_iterations += s.newton_iterations;
    st.linear_iterations += s.linear_iterations;
    return rc;
  }

  // SimpleAdaptiveStepper::do_step (stepper.hh:330-350): shrink until a step
  // succeeds or leaves the allowed range, then suggest a larger next step
  int adaptive_step(double& dt) {
    if (!dt_allowed(dt)) return kStepSizeLimit;
    int rc = attempt(dt);
    while (rc) {
      ++st.failed_steps;
      dt *= opt.decrease_factor;
      if (!dt_allowed(dt)) return kStepSizeLimit;
      rc = attempt(dt);
    }
    ++st.steps;
    st.accepted_dt.push_back(dt);
    double next = dt * opt.increase_factor;
    if (opt.max_step > 0.0) next = std::clamp(next, -std::fabs(opt.max_step), std::fabs(opt.max_step));
    dt = next;
    if (on_step) on_step(m.time());
    return 0;
  }

  // TimeStepper::snap_to_time (stepper.hh:196-240)
  int snap(double& dt, double snap_time) {
    int count = 0;
    int rc = 0;
    while (flt(m.time(), snap_time)) {
      const double ratio = (snap_time - m.time()) / dt;
      const int n = static_cast<int>(std::ceil(ratio));
      if (n <= 0) throw Error("Timestep doesn't make advances towards snap step");
      dt = (snap_time - m.time()) / n;
      rc = adaptive_step(dt);
      if (rc) {
        if (count++ == 100) throw Error("Snapping time exceeded maximum iteration count");
        dt *= 0.5;
      }
    }
    return rc;
  }
};

}  // namespace

EvolveOptions EvolveOptions::from_model(const Model& m) {
  const auto& c = m.config();
  const std::string p = "model.time_step_operator.";
  EvolveOptions o;
  o.t_end = c.get(p + "time_end", 1.0);
  o.decrease_factor = c.get(p + "time_step_decrease_factor", 0.5);
  o.increase_factor = c.get(p + "time_step_increase_factor", 1.1);
  if (c.has_key(p + "time_step_min")) o.min_step = c.get(p + "time_step_min", 0.0);
  if (c.has_key(p + "time_step_max")) o.max_step = c.get(p + "time_step_max", 0.0);
  o.dt = c.get(p + "time_step_initial", o.max_step > 0.0 ? o.max_step : 0.1);
  return o;
}

EvolveStats evolve(Model& m, const EvolveOptions& opt, const std::function<void(double)>& on_step) {
  // constructor checks of SimpleAdaptiveStepper (stepper.hh:283-295)
  if (opt.decrease_factor > 1.0 || opt.decrease_factor < 0.0)
    throw IOError("Decrease factor " + std::to_string(opt.decrease_factor) + " must be in the range of (0,1]");
  if (opt.increase_factor < 1.0)
    throw IOError("Increase factor " + std::to_string(opt.increase_factor) + " must be in the range of (1,inf)");
  if (!(opt.dt > 0.0)) throw IOError("the initial time step must be positive");
  EvolveStats st;
  Runner r{m, opt, st, on_step};
  double dt = opt.dt;
  // TimeStepper::evolve (stepper.hh:145-175): stop 2 dt before the end when
  // snapping, 1 dt otherwise
  const double stop_dt = opt.snap_to_end_time ? 2.0 : 1.0;
  int rc = 0;
  while (fle(m.time() + stop_dt * dt, opt.t_end)) {
    rc = r.adaptive_step(dt);
    if (rc) break;
  }
  if (!rc && opt.snap_to_end_time) rc = r.snap(dt, opt.t_end);
  st.status = rc;
  st.time = m.time();
  st.dt = dt;
  return st;
}

}  // namespace dcopasi
