// kinematics.hpp — host-side knife motion: pose at time t (double) and the
// per-step state ObjectSpatula holds during a step (object_spatula.hpp:164-233):
// pose at t_n, velocity (p(t_n) - p(t_{n-1}))/dt, angular velocity from the
// rotation between the two poses divided by dt.  The reference sets these at
// the end of the previous step; step 0 uses the same difference rule (the
// reference leaves it uninitialised).
#pragma once
#include <vector>
#include <stdint.h>
#include "../../../include/mpmx.h"

namespace mpmx {

struct Pose {
  double R[3][3];
  double c[3];
};

Pose knife_pose(const mpmx_knife& k, const std::vector<float>& frames7, float fps, double t);
mpmx_knife_state knife_state_at(const mpmx_knife& k, const std::vector<float>& frames7, float fps, double dt,
                                int64_t step);
mpmx_knife_state knife_state_at_time(const mpmx_knife& k, const std::vector<float>& frames7, float fps, double t1,
                                     double t0, double dt);

}  // namespace mpmx
