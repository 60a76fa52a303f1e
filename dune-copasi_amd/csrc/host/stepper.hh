// Time loop over a device model: the reference's SimpleAdaptiveStepper and
// TimeStepper::evolve / snap_to_time (dune/copasi/common/stepper.hh:95-240,
// 330-366), configured from the same keys the reference's CLI reads
// (src/dune_copasi.cc:384-401):
//   model.time_step_operator.{time_begin, time_end, time_step_initial,
//     time_step_min, time_step_max, time_step_decrease_factor,
//     time_step_increase_factor}
// A failed step leaves the model's state and time untouched (Model::step
// updates them only on success), so a retry restarts from the same state.
#pragma once
#include <functional>
#include <vector>

#include "model.hh"

namespace dcopasi {

struct EvolveOptions {
  double t_end = 1.0;
  double dt = 0.1;               // first step size tried
  double decrease_factor = 0.5;  // on failure, in (0, 1]
  double increase_factor = 1.1;  // on success, >= 1
  double min_step = 0.0;         // 0: no lower limit
  double max_step = 0.0;         // 0: no upper limit
  bool snap_to_end_time = true;
  static EvolveOptions from_model(const Model& m);
};

struct EvolveStats {
  int status = 0;  // 0 ok, 5 step size outside [min_step, max_step], else the failing step's status
  int steps = 0, failed_steps = 0;
  double time = 0.0, dt = 0.0;  // reached time, suggested next step size
  long newton_iterations = 0, linear_iterations = 0;
  std::vector<double> accepted_dt;  // sizes of the accepted steps, in order
};

// Advances the model from its current time towards opt.t_end; on_step is
// called after every accepted step with the new time.
EvolveStats evolve(Model& m, const EvolveOptions& opt, const std::function<void(double)>& on_step = {});

}  // namespace dcopasi
