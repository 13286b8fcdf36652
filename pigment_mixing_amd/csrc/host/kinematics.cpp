// kinematics.cpp — knife poses for the analytic sweeps and the keyframe
// animation (ObjectSpatula::move, object_spatula.hpp:164-233: lerp of the
// position and slerp of the quaternion between 60 fps frames).
#include "kinematics.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mpmx {
namespace {

void rot_axis_angle(const double axis[3], double ang, double R[3][3]) {
  double n = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  double a0 = 1, a1 = 0, a2 = 0;
  if (n > 0) { a0 = axis[0] / n; a1 = axis[1] / n; a2 = axis[2] / n; }
  double c = std::cos(ang), s = std::sin(ang), C = 1 - c;
  R[0][0] = c + a0 * a0 * C;      R[0][1] = a0 * a1 * C - a2 * s; R[0][2] = a0 * a2 * C + a1 * s;
  R[1][0] = a1 * a0 * C + a2 * s; R[1][1] = c + a1 * a1 * C;      R[1][2] = a1 * a2 * C - a0 * s;
  R[2][0] = a2 * a0 * C - a1 * s; R[2][1] = a2 * a1 * C + a0 * s; R[2][2] = c + a2 * a2 * C;
}

void mul33(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}

void quat_xyzw_to_R(double x, double y, double z, double w, double R[3][3]) {
  double n = std::sqrt(x * x + y * y + z * z + w * w);
  if (n > 0) { x /= n; y /= n; z /= n; w /= n; }
  R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - z * w);     R[0][2] = 2 * (x * z + y * w);
  R[1][0] = 2 * (x * y + z * w);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - x * w);
  R[2][0] = 2 * (x * z - y * w);     R[2][1] = 2 * (y * z + x * w);     R[2][2] = 1 - 2 * (x * x + y * y);
}

// spherical linear interpolation, shortest arc
void quat_slerp(const double a[4], const double b_in[4], double t, double out[4]) {
  double b[4] = {b_in[0], b_in[1], b_in[2], b_in[3]};
  double d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  if (d < 0) { for (double& v : b) v = -v; d = -d; }
  double wa, wb;
  if (d > 1.0 - 1e-12) {
    wa = 1 - t; wb = t;
  } else {
    double th = std::acos(d), st = std::sin(th);
    wa = std::sin((1 - t) * th) / st;
    wb = std::sin(t * th) / st;
  }
  for (int i = 0; i < 4; ++i) out[i] = wa * a[i] + wb * b[i];
}

void circle_pose(const mpmx_knife& k, const double centre[3], double tt, const double Rb[3][3], Pose& P) {
  const double om = k.speed / k.radius;
  const double phi = k.phase0 + om * tt;
  P.c[0] = centre[0] + k.radius * std::cos(phi);
  P.c[1] = centre[1];
  P.c[2] = centre[2] + k.radius * std::sin(phi);
  // heading: local x follows the tangent (-sin phi, 0, cos phi) — a rotation about y
  const double yax[3] = {0, 1, 0};
  double Rh[3][3];
  rot_axis_angle(yax, -(phi + M_PI / 2), Rh);
  mul33(Rh, Rb, P.R);
}

}  // namespace

Pose knife_pose(const mpmx_knife& k, const std::vector<float>& frames7, float fps, double t) {
  Pose P;
  const double tilt[3] = {k.tilt_axis[0], k.tilt_axis[1], k.tilt_axis[2]};
  double Rb[3][3];
  rot_axis_angle(tilt, k.tilt_angle, Rb);
  switch (k.motion) {
    case MPMX_MOTION_LINEAR:
      for (int i = 0; i < 3; ++i) P.c[i] = k.p0[i] + (double)k.velocity[i] * t;
      std::memcpy(P.R, Rb, sizeof Rb);
      return P;
    case MPMX_MOTION_CIRCULAR: {
      const double ctr[3] = {k.p0[0], k.p0[1], k.p0[2]};
      circle_pose(k, ctr, t, Rb, P);
      return P;
    }
    case MPMX_MOTION_LINEAR_THEN_CIRCULAR: {
      if (t < k.t_switch) {
        for (int i = 0; i < 3; ++i) P.c[i] = k.p0[i] + (double)k.velocity[i] * t;
        std::memcpy(P.R, Rb, sizeof Rb);
        return P;
      }
      double ps[3];
      for (int i = 0; i < 3; ++i) ps[i] = k.p0[i] + (double)k.velocity[i] * k.t_switch;
      const double ctr[3] = {ps[0] - k.radius * std::cos((double)k.phase0), ps[1],
                             ps[2] - k.radius * std::sin((double)k.phase0)};
      circle_pose(k, ctr, t - k.t_switch, Rb, P);
      return P;
    }
    case MPMX_MOTION_KEYFRAMES: {
      const int nf = (int)(frames7.size() / 7);
      if (nf > 0) {
        const double fi = t * fps;
        const int i0u = (int)std::floor(fi);
        const int i0 = std::min(std::max(i0u, 0), nf - 1), i1 = std::min(std::max(i0u + 1, 0), nf - 1);
        const double u = fi - i0u;
        const float* f0 = &frames7[7 * (size_t)i0];
        const float* f1 = &frames7[7 * (size_t)i1];
        for (int i = 0; i < 3; ++i) P.c[i] = f0[i] + u * ((double)f1[i] - f0[i]);
        const double q0[4] = {f0[3], f0[4], f0[5], f0[6]}, q1[4] = {f1[3], f1[4], f1[5], f1[6]};
        double q[4];
        quat_slerp(q0, q1, u, q);
        quat_xyzw_to_R(q[0], q[1], q[2], q[3], P.R);
        return P;
      }
      break;
    }
    default:
      break;
  }
  for (int i = 0; i < 3; ++i) P.c[i] = k.p0[i];
  std::memcpy(P.R, Rb, sizeof Rb);
  return P;
}

mpmx_knife_state knife_state_at(const mpmx_knife& k, const std::vector<float>& frames7, float fps, double dt,
                                int64_t step) {
  return knife_state_at_time(k, frames7, fps, (double)step * dt, (double)(step - 1) * dt, dt);
}

// The state during the step that ends at time t1 (pose at t1, velocities from
// the pose at t0 = t1 - dt).
mpmx_knife_state knife_state_at_time(const mpmx_knife& k, const std::vector<float>& frames7, float fps, double t1,
                                     double t0, double dt) {
  const Pose P1 = knife_pose(k, frames7, fps, t1);
  const Pose P0 = knife_pose(k, frames7, fps, t0);
  mpmx_knife_state s;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      s.T[i * 4 + j] = (float)P1.R[i][j];
      s.Tinv[i * 4 + j] = (float)P1.R[j][i];
      s.Nmat[i * 3 + j] = (float)P1.R[i][j];  // (R^-1)^T = R for a rotation
    }
    s.T[i * 4 + 3] = (float)P1.c[i];
    s.Tinv[i * 4 + 3] = (float)(-(P1.R[0][i] * P1.c[0] + P1.R[1][i] * P1.c[1] + P1.R[2][i] * P1.c[2]));
    s.v[i] = (float)((P1.c[i] - P0.c[i]) / dt);
  }
  // relative rotation R1 R0^T -> axis * angle
  double R0t[3][3], Rr[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R0t[i][j] = P0.R[j][i];
  mul33(P1.R, R0t, Rr);
  const double tr = Rr[0][0] + Rr[1][1] + Rr[2][2];
  const double ang = std::acos(std::max(-1.0, std::min(1.0, (tr - 1) / 2)));
  const double ax[3] = {Rr[2][1] - Rr[1][2], Rr[0][2] - Rr[2][0], Rr[1][0] - Rr[0][1]};
  const double f = (ang < 1e-8) ? 0.5 : ang / (2 * std::sin(ang));
  for (int i = 0; i < 3; ++i) s.w[i] = (float)(ax[i] * f / dt);
  return s;
}

}  // namespace mpmx

extern "C" int mpmx_knife_state_eval(const mpmx_knife* k, const float* frames7, int n_frames, float fps, double dt,
                                     int64_t step, mpmx_knife_state* out) {
  if (!k || !out || !(dt > 0) || n_frames < 0 || (n_frames > 0 && (!frames7 || !(fps > 0)))) return MPMX_ERR_ARG;
  std::vector<float> fr;
  if (n_frames > 0) fr.assign(frames7, frames7 + 7 * (size_t)n_frames);
  *out = mpmx::knife_state_at(*k, fr, fps, dt, step);
  return MPMX_OK;
}
