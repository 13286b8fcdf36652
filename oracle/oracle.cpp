// oracle.cpp — CPU restatement of the reference's extended Matter MPM step.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library, and only as the checker (or the timed CPU
// baseline).  The product (libmpmx) never links or calls it.
//
// It follows the reference phase by phase (file:line cited at each function):
//   Simulation::advanceStep          matter/src/simulation/simulation.cpp:225-280
//   precomputeWeights (64-node cache) precompute_weight.cpp:3-43, tools.hpp:153-177,243-266
//   P2G (thread-local grids + merge) p2g.cpp:6-126
//   explicitEulerUpdate + HenckyPiola explicit_euler_update.cpp:7-112, simulation.hpp:287-294
//   boundaryCollision (objects, plates) boundary_collision.cpp:5-316
//   ObjectSpatula sdf/inside/normal/velocity object_spatula.hpp:40-162
//   ObjectPlate::inside              object_plate.hpp:30-78
//   G2P + smoothStep                 g2p.cpp:7-135
//   MUSL                             musl.cpp:6-98
//   deformationUpdate                deformation_update.cpp:7-52
//   plasticity (VMVisc branch)       plasticity.cpp:10-27, 163-238
//   positionUpdate                   position_update.cpp:6-91
//   moveObjects (velocity rule)      simulation.cpp:284-291, object_spatula.hpp:164-233
//   derived constants                simulation_units.cpp:167-183
//   lattice origin rule              remesh.cpp:45-180 (fixed lattice, see below)
//
// Deliberate, documented differences from the reference binary:
//   * Lattice is fixed at prepare time (origin = bbox_min - 2dx as remeshFixedInit,
//     node coordinates by the reference's float arange).  The reference re-meshes
//     each step (remeshFixedCont) on the same integer-dx lattice; node float values
//     can differ by rounding there.  Particles whose 3x3x3 stencil leaves the
//     fixed lattice are deactivated ("domain" rule).
//   * Fixed dt (updateDt's CFL rule is not applied).
//   * Knife: slab (exact box SDF) or the reference trapezoid, with analytic or
//     keyframed kinematics; the step-0 velocity is defined by the same backward
//     difference as later steps (the reference leaves it uninitialised).
//   * Principal frame of F: the reference's JacobiSVD of F followed by log sigma
//     is computed as the symmetric eigenproblem of F^T F - I (formed from F - I)
//     with log sigma = log1p(lam) / 2 (oracle_math.hpp principal_of): the same
//     function of F, without the float rounding of sigma ~ 1 that costs the small
//     Hencky strains 1e-5..1e-3 of their relative precision.  The plastic F is
//     rebuilt as F_trial + F_trial V diag(expm1(h_new - h_old)) V^T, equal to
//     U diag(e^h_new) V^T.
//
// n_threads == 1 gives the deterministic sequential order used for parity
// (the reference itself is only reproducible single-threaded: its threads'
// grid replicas are merged, and its force atomics land, in timing order).
// n_threads > 1 keeps the reference's OpenMP cost structure (per-thread full
// grid replicas merged under a critical section, atomic force scatter) for the
// CPU baseline.
#include "oracle_math.hpp"
#include "../include/mpmx.h"

#include <array>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <string>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {

typedef std::array<float, 7> L7;

struct Nbhd {  // ParticleNeighborhood (data_structures.hpp:75-79)
  int base[3];
  float w[64];
  V3 g[64];
};

struct KState {
  float T[12], Tinv[12], N[9], v[3], w[3];
};

// ---------------------------------------------------------------- splines
// tools.hpp:153-177.  The constants are double literals in the reference, so
// the polynomial is evaluated in double and rounded to float.
inline float spline_N(float u) {
  float uabs = std::fabs(u);
  if (uabs < 0.5f) return (float)(0.75 - (double)(uabs * uabs));
  if (uabs < 1.5f) return (float)(0.5 * (1.5 - (double)uabs) * (1.5 - (double)uabs));
  return 0.0f;
}
inline float spline_dN(float u) {
  float uabs = std::fabs(u);
  if (uabs < 0.5f) return -2.0f * u;
  if (uabs < 1.5f) {
    int sg = (0.0f < u) - (u < 0.0f);
    return (float)((double)u - 1.5 * sg);
  }
  return 0.0f;
}

// --------------------------------------------------------------- knife SDFs
// Exact box SDF (textbook): q = |p| - h;  |max(q,0)| + min(max(q.x,q.y,q.z),0).
inline float sdf_box(const float p[3], const float h[3]) {
  float q0 = std::fabs(p[0]) - h[0], q1 = std::fabs(p[1]) - h[1], q2 = std::fabs(p[2]) - h[2];
  float m0 = std::max(q0, 0.0f), m1 = std::max(q1, 0.0f), m2 = std::max(q2, 0.0f);
  float outside = std::sqrt(m0 * m0 + m1 * m1 + m2 * m2);
  float inside = std::min(std::max(q0, std::max(q1, q2)), 0.0f);
  return outside + inside;
}

// ObjectSpatula::sdSpatula (object_spatula.hpp:40-82): trapezoid in (|x|, z),
// rounded ends, extruded by halfThickness along y.
inline float sdf_spatula(const float p[3], float halfWidth, float halfThickness,
                         float halfLength, float offset) {
  float px = std::fabs(p[0]), pz = p[2];
  float b1 = halfWidth + offset;
  float b2 = (halfWidth + offset) * 0.25f;
  float he = halfLength + offset;
  float p1x = b1, p1y = -he, p2x = b2, p2y = he;
  float bax = p2x - p1x, bay = p2y - p1y;
  float pax = px - p1x, pay = pz - p1y;
  float h = std::max(0.0f, std::min(1.0f, (pax * bax + pay * bay) / (bax * bax + bay * bay)));
  float dxs = pax - bax * h, dys = pay - bay * h;
  float d_side = std::sqrt(dxs * dxs + dys * dys);
  float inside_x = b1 - (b1 - b2) * (pz + he) / (2.0f * he);
  float d_inside = px - inside_x;
  float d2 = (px < inside_x) ? d_inside : d_side;
  if (pz > he) {
    d2 = std::sqrt(px * px + (pz - he) * (pz - he)) - b2;
  } else if (pz < -he) {
    float rx = halfWidth + offset, rz = 1.5f * (halfWidth + offset);
    float prx = px, prz = pz + he;
    float k0 = std::sqrt((prx / rx) * (prx / rx) + (prz / rz) * (prz / rz));
    float k1 = std::sqrt((prx / (rx * rx)) * (prx / (rx * rx)) + (prz / (rz * rz)) * (prz / (rz * rz)));
    d2 = k0 * (k0 - 1.0f) / k1;
  }
  float dy = std::fabs(p[1]) - halfThickness;
  float mx = std::max(d2, dy);
  float ex = std::max(d2, 0.0f), ey = std::max(dy, 0.0f);
  return std::min(mx, 0.0f) + std::sqrt(ex * ex + ey * ey);
}

// ------------------------------------------------------------ kinematics (double)
struct DPose { double R[3][3]; double c[3]; };

void axis_angle(const double ax_in[3], double ang, double R[3][3]) {
  double n = std::sqrt(ax_in[0] * ax_in[0] + ax_in[1] * ax_in[1] + ax_in[2] * ax_in[2]);
  double a[3] = {1, 0, 0};
  if (n > 0) { a[0] = ax_in[0] / n; a[1] = ax_in[1] / n; a[2] = ax_in[2] / n; }
  double c = std::cos(ang), s = std::sin(ang), C = 1 - c;
  R[0][0] = c + a[0] * a[0] * C;        R[0][1] = a[0] * a[1] * C - a[2] * s; R[0][2] = a[0] * a[2] * C + a[1] * s;
  R[1][0] = a[1] * a[0] * C + a[2] * s; R[1][1] = c + a[1] * a[1] * C;        R[1][2] = a[1] * a[2] * C - a[0] * s;
  R[2][0] = a[2] * a[0] * C - a[1] * s; R[2][1] = a[2] * a[1] * C + a[0] * s; R[2][2] = c + a[2] * a[2] * C;
}
void mat_mul_d(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = A[i][0] * B[0][j] + A[i][1] * B[1][j] + A[i][2] * B[2][j];
}
void quat_to_R(const double q[4] /*x,y,z,w*/, double R[3][3]) {
  double x = q[0], y = q[1], z = q[2], w = q[3];
  double n = std::sqrt(x * x + y * y + z * z + w * w);
  if (n > 0) { x /= n; y /= n; z /= n; w /= n; }
  R[0][0] = 1 - 2 * (y * y + z * z); R[0][1] = 2 * (x * y - z * w);     R[0][2] = 2 * (x * z + y * w);
  R[1][0] = 2 * (x * y + z * w);     R[1][1] = 1 - 2 * (x * x + z * z); R[1][2] = 2 * (y * z - x * w);
  R[2][0] = 2 * (x * z - y * w);     R[2][1] = 2 * (y * z + x * w);     R[2][2] = 1 - 2 * (x * x + y * y);
}
void slerp(const double a_in[4], const double b_in[4], double t, double out[4]) {
  double a[4], b[4];
  for (int i = 0; i < 4; ++i) { a[i] = a_in[i]; b[i] = b_in[i]; }
  double d = a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
  if (d < 0) { for (int i = 0; i < 4; ++i) b[i] = -b[i]; d = -d; }
  double s0, s1;
  if (d > 1.0 - 1e-12) { s0 = 1 - t; s1 = t; }
  else {
    double th = std::acos(d), st = std::sin(th);
    s0 = std::sin((1 - t) * th) / st;
    s1 = std::sin(t * th) / st;
  }
  for (int i = 0; i < 4; ++i) out[i] = s0 * a[i] + s1 * b[i];
}

DPose knife_pose(const mpmx_knife& k, const std::vector<float>& frames, float frame_fps, double t) {
  DPose P;
  double tilt_ax[3] = {k.tilt_axis[0], k.tilt_axis[1], k.tilt_axis[2]};
  double Rb[3][3];
  axis_angle(tilt_ax, k.tilt_angle, Rb);
  double yax[3] = {0, 1, 0};
  auto circ = [&](double cx, double cy, double cz, double tt, double phase) {
    double om = k.speed / k.radius;
    double phi = phase + om * tt;
    P.c[0] = cx + k.radius * std::cos(phi);
    P.c[1] = cy;
    P.c[2] = cz + k.radius * std::sin(phi);
    double Rh[3][3];
    axis_angle(yax, -(phi + M_PI / 2), Rh);
    mat_mul_d(Rh, Rb, P.R);
  };
  switch (k.motion) {
    case MPMX_MOTION_LINEAR:
      for (int i = 0; i < 3; ++i) P.c[i] = k.p0[i] + k.velocity[i] * t;
      std::memcpy(P.R, Rb, sizeof(Rb));
      break;
    case MPMX_MOTION_CIRCULAR:
      circ(k.p0[0], k.p0[1], k.p0[2], t, k.phase0);
      break;
    case MPMX_MOTION_LINEAR_THEN_CIRCULAR:
      if (t < k.t_switch) {
        for (int i = 0; i < 3; ++i) P.c[i] = k.p0[i] + k.velocity[i] * t;
        std::memcpy(P.R, Rb, sizeof(Rb));
      } else {
        double ps[3];
        for (int i = 0; i < 3; ++i) ps[i] = k.p0[i] + k.velocity[i] * k.t_switch;
        double cc[3] = {ps[0] - k.radius * std::cos(k.phase0), ps[1], ps[2] - k.radius * std::sin(k.phase0)};
        circ(cc[0], cc[1], cc[2], t - k.t_switch, k.phase0);
      }
      break;
    case MPMX_MOTION_KEYFRAMES: {
      int nf = (int)(frames.size() / 7);
      if (nf == 0) { for (int i = 0; i < 3; ++i) P.c[i] = k.p0[i]; std::memcpy(P.R, Rb, sizeof(Rb)); break; }
      double fi = t * frame_fps;
      int i0u = (int)std::floor(fi);
      int i0 = std::min(std::max(i0u, 0), nf - 1), i1 = std::min(std::max(i0u + 1, 0), nf - 1);
      double tt = fi - i0u;
      const float* f0 = &frames[7 * i0];
      const float* f1 = &frames[7 * i1];
      for (int i = 0; i < 3; ++i) P.c[i] = f0[i] + tt * (f1[i] - f0[i]);
      double q0[4] = {f0[3], f0[4], f0[5], f0[6]}, q1[4] = {f1[3], f1[4], f1[5], f1[6]}, q[4];
      slerp(q0, q1, tt, q);
      quat_to_R(q, P.R);
      break;
    }
    default:
      for (int i = 0; i < 3; ++i) P.c[i] = k.p0[i];
      std::memcpy(P.R, Rb, sizeof(Rb));
  }
  return P;
}

// State used during step n: pose at t_n, velocity by the backward difference
// (p(t_n) - p(t_{n-1}))/dt and omega from R(t_n) R(t_{n-1})^T (object_spatula.hpp:183-197).
KState knife_state(const mpmx_knife& k, const std::vector<float>& frames, float frame_fps,
                   double dt, int64_t n) {
  DPose P1 = knife_pose(k, frames, frame_fps, n * dt);
  DPose P0 = knife_pose(k, frames, frame_fps, (n - 1) * dt);
  KState s;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      s.T[i * 4 + j] = (float)P1.R[i][j];
      s.Tinv[i * 4 + j] = (float)P1.R[j][i];
      s.N[i * 3 + j] = (float)P1.R[i][j];
    }
    s.T[i * 4 + 3] = (float)P1.c[i];
    s.Tinv[i * 4 + 3] = (float)(-(P1.R[0][i] * P1.c[0] + P1.R[1][i] * P1.c[1] + P1.R[2][i] * P1.c[2]));
    s.v[i] = (float)((P1.c[i] - P0.c[i]) / dt);
  }
  double Rr[3][3], R0t[3][3];
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) R0t[i][j] = P0.R[j][i];
  mat_mul_d(P1.R, R0t, Rr);
  double tr = Rr[0][0] + Rr[1][1] + Rr[2][2];
  double ca = std::max(-1.0, std::min(1.0, (tr - 1) / 2));
  double ang = std::acos(ca);
  double ax[3] = {Rr[2][1] - Rr[1][2], Rr[0][2] - Rr[2][0], Rr[1][0] - Rr[0][1]};
  double sfac = (ang < 1e-8) ? 0.5 : ang / (2 * std::sin(ang));
  for (int i = 0; i < 3; ++i) s.w[i] = (float)(ax[i] * sfac / dt);
  return s;
}

inline V3 knife_local(const KState& s, const V3& X) {
  // invTransform.matrix() * (X,1)   (object_spatula.hpp:90-92)
  V3 r;
  for (int i = 0; i < 3; ++i)
    r[i] = ((s.Tinv[i * 4 + 0] * X[0] + s.Tinv[i * 4 + 1] * X[1]) + s.Tinv[i * 4 + 2] * X[2]) + s.Tinv[i * 4 + 3];
  return r;
}

struct Oracle {
  mpmx_config cfg;
  int64_t Np = 0;
  std::vector<V3> x, v, pic, flip;
  std::vector<M3> F, Bm;
  std::vector<L7> pig;
  std::vector<float> diff, eps_pl_dev, dgamma;
  std::vector<float> eps_pl_vol, eps_prad;  // eps_pl_vol, eps_pl_vol_pradhana (data_structures.hpp:30-39)
  std::vector<uint8_t> active;
  float particle_mass = 0, particle_volume = 0, dx = 0;
  float lambda = 0, mu = 0, K = 0, d_prefac = 0, e_mu_prefac = 0, f_mu_prefac = 0;
  float one_over_dx = 0, apicDinverse = 0;
  int Nx = 0, Ny = 0, Nz = 0;
  std::vector<float> gx, gy, gz;
  float xc = 0, yc = 0, zc = 0;
  std::vector<V3> gv, gflip;
  std::vector<float> gmass, gshear;
  std::vector<L7> gpig;
  std::vector<Nbhd> nb;
  std::vector<mpmx_plate> plates;       // state during the current step
  std::vector<mpmx_plate> plates0;      // as set (construction velocities)
  bool has_knife = false;
  mpmx_knife knife;
  std::vector<float> frames;
  float frame_fps = 60.0f;
  KState ks;
  float time = 0;
  int64_t step = 0;
  int exit_flag = 0;
  int64_t plastic_count = 0;
  bool prepared = false;
  int nt = 1;

  inline int64_t ind(int i, int j, int k) const { return ((int64_t)i * Ny + j) * Nz + k; }
  size_t nodes() const { return (size_t)Nx * Ny * Nz; }

  // remesh.cpp:45-180 origin rule and tools.hpp:325-330 arange, on a fixed lattice.
  void make_lattice() {
    float mn[3] = {1e20f, 1e20f, 1e20f}, mx[3] = {-1e20f, -1e20f, -1e20f};
    bool any = false;
    for (int64_t p = 0; p < Np; ++p) {
      if (!active[p]) continue;
      any = true;
      for (int d = 0; d < 3; ++d) { mn[d] = std::min(mn[d], x[p][d]); mx[d] = std::max(mx[d], x[p][d]); }
    }
    if (!any) for (int d = 0; d < 3; ++d) mn[d] = mx[d] = 0;
    if (cfg.grid_reference_point[0] < 1e10f) {
      for (int d = 0; d < 3; ++d) {
        if (cfg.grid_reference_point[d] < mn[d]) mn[d] = cfg.grid_reference_point[d];
        else if (cfg.grid_reference_point[d] > mx[d]) mx[d] = cfg.grid_reference_point[d];
      }
    }
    float low[3], high[3];
    for (int d = 0; d < 3; ++d) {
      low[d] = mn[d] - dx * 2u;
      high[d] = mx[d] + dx * 2u;
      if (cfg.origin_mode == 1) low[d] = cfg.grid_origin[d];
    }
    std::vector<float>* ax[3] = {&gx, &gy, &gz};
    int dims[3];
    for (int d = 0; d < 3; ++d) {
      ax[d]->clear();
      if (cfg.grid_dims[d] > 0) {
        float val = low[d];
        for (int i = 0; i < cfg.grid_dims[d]; ++i) { ax[d]->push_back(val); val += dx; }
      } else {
        for (float val = low[d]; val < high[d] + dx; val += dx) ax[d]->push_back(val);
      }
      dims[d] = (int)ax[d]->size();
    }
    Nx = dims[0]; Ny = dims[1]; Nz = dims[2];
    xc = gx[0]; yc = gy[0]; zc = gz[0];
  }

  void derived() {
    // simulation_units.cpp:151-183 (constants computed in double, stored as float)
    apicDinverse = 4.0 / (dx * dx);
    lambda = cfg.nu * cfg.E / ((1.0 + cfg.nu) * (1.0 - 2.0 * cfg.nu));
    mu = cfg.E / (2.0 * (1.0 + cfg.nu));
    K = lambda + 2.0 * mu / 3;
    one_over_dx = 1.0 / dx;
    d_prefac = 1 / cfg.q_prefac;
    e_mu_prefac = 2 * cfg.q_prefac * mu;
    f_mu_prefac = 2 * cfg.q_prefac * cfg.q_prefac * mu;
    particle_volume = dx * dx * dx / cfg.ppc;
    particle_mass = cfg.rho * particle_volume;
  }

  bool stencil_inside(const V3& xp) const {
    float f[3] = {(xp[0] - xc) * one_over_dx, (xp[1] - yc) * one_over_dx, (xp[2] - zc) * one_over_dx};
    int n[3] = {Nx, Ny, Nz};
    for (int d = 0; d < 3; ++d) {
      if (!std::isfinite(f[d])) return false;
      float b = std::floor(f[d] - 0.5f);
      if (b < 0.0f || b + 2.0f > (float)(n[d] - 1)) return false;
    }
    return true;
  }

  // precompute_weight.cpp:3-43
  void precompute_weights() {
    nb.resize(Np);
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (int64_t p = 0; p < Np; ++p) {
      if (!active[p]) continue;
      const V3& xp = x[p];
      int ib = (int)std::floor((xp[0] - xc) * one_over_dx) - 1;
      int jb = (int)std::floor((xp[1] - yc) * one_over_dx) - 1;
      int kb = (int)std::floor((xp[2] - zc) * one_over_dx) - 1;
      Nbhd& n = nb[p];
      n.base[0] = std::max(0, ib); n.base[1] = std::max(0, jb); n.base[2] = std::max(0, kb);
      int count = 0;
      for (int i = n.base[0]; i < n.base[0] + 4; ++i)
        for (int j = n.base[1]; j < n.base[1] + 4; ++j)
          for (int k = n.base[2]; k < n.base[2] + 4; ++k) {
            float xi = i < Nx ? gx[i] : 1e30f, yi = j < Ny ? gy[j] : 1e30f, zi = k < Nz ? gz[k] : 1e30f;
            float ux = (xp[0] - xi) * one_over_dx, uy = (xp[1] - yi) * one_over_dx, uz = (xp[2] - zi) * one_over_dx;
            float Nx_ = spline_N(ux), Ny_ = spline_N(uy), Nz_ = spline_N(uz);
            float dNx = spline_dN(ux), dNy = spline_dN(uy), dNz = spline_dN(uz);
            n.w[count] = Nx_ * Ny_ * Nz_;
            n.g[count] = v3(dNx * Ny_ * Nz_ * one_over_dx, Nx_ * dNy * Nz_ * one_over_dx, Nx_ * Ny_ * dNz * one_over_dx);
            ++count;
          }
    }
  }

  void resize_grid() {  // remesh.cpp:5-13
    size_t n = nodes();
    gv.assign(n, zero3());
    gflip.assign(n, zero3());
    gmass.assign(n, 0.0f);
    L7 z; z.fill(0.0f);
    gpig.assign(n, z);
    gshear.assign(n, 0.0f);
  }

  // p2g.cpp:6-126
  void p2g() {
    size_t n = nodes();
#pragma omp parallel num_threads(nt) if (nt > 1)
    {
      std::vector<V3> lv(n, zero3());
      std::vector<float> lm(n, 0.0f);
      L7 z; z.fill(0.0f);
      std::vector<L7> lp(n, z);
      std::vector<float> ls(n, 0.0f);
#pragma omp for nowait
      for (int64_t p = 0; p < Np; ++p) {
        if (!active[p]) continue;
        const V3& xp = x[p];
        const Nbhd& nbp = nb[p];
        int count = 0;
        for (int i = nbp.base[0]; i < nbp.base[0] + 4; ++i) {
          for (int j = nbp.base[1]; j < nbp.base[1] + 4; ++j) {
            for (int k = nbp.base[2]; k < nbp.base[2] + 4; ++k) {
              float w = nbp.w[count];
              ++count;
              if (w > 1e-25f) {
                int64_t idx = ind(i, j, k);
                float xi = gx[i], yi = gy[j], zi = gz[k];
                lm[idx] += w;
                lv[idx] = add(lv[idx], scale(v[p], w));
                for (int c = 0; c < 7; ++c) lp[idx][c] += pig[p][c] * w;
                ls[idx] += diff[p] * w;
                if (cfg.flip_ratio < 0) {
                  V3 pd = v3(xi - xp[0], yi - xp[1], zi - xp[2]);
                  V3 bd = matvec(Bm[p], pd);
                  lv[idx] = add(lv[idx], scale(scale(bd, apicDinverse), w));
                }
              }
            }
          }
        }
      }
#pragma omp critical
      {
        for (size_t l = 0; l < n; ++l) {
          gmass[l] += lm[l];
          gv[l] = add(gv[l], lv[l]);
          for (int c = 0; c < 7; ++c) gpig[l][c] += lp[l][c];
          gshear[l] += ls[l];
        }
      }
    }
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (int64_t l = 0; l < (int64_t)n; ++l) {
      float mi = gmass[l];
      if (mi > 0) {
        gv[l] = v3(gv[l][0] / mi, gv[l][1] / mi, gv[l][2] / mi);
        for (int c = 0; c < 7; ++c) gpig[l][c] /= mi;
        gshear[l] /= mi;
        float s = ((gpig[l][0] + gpig[l][1]) + gpig[l][2]) + gpig[l][3];
        if (s > 1e-6f) for (int c = 0; c < 4; ++c) gpig[l][c] /= s;
      } else {
        gv[l] = zero3();
        gpig[l].fill(0.0f);
        gshear[l] = 0.0f;
      }
    }
    for (size_t l = 0; l < n; ++l) gmass[l] *= particle_mass;
  }

  // simulation.hpp:287-294 HenckyPiola, then tau = P F^T (explicit_euler_update.cpp:20):
  // P = U diag((2 mu ln s_i + lambda sum ln s) / s_i) V^T, so tau = P F^T =
  // U diag(2 mu ln s_i + lambda sum ln s) U^T, with ln s_i from the principal
  // frame of F (principal_of: log1p of the eigenvalues of F^T F - I).
  M3 kirchhoff(const M3& Fe) const {
    const Principal P = principal_of(Fe);
    float ls[3] = {log_sigma(P.lam[0]), log_sigma(P.lam[1]), log_sigma(P.lam[2])};
    float tr = (ls[0] + ls[1]) + ls[2];
    float d[3];
    for (int i = 0; i < 3; ++i) d[i] = 2 * mu * ls[i] + lambda * tr;
    M3 t;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        t(r, c) = (P.U(r, 0) * d[0] * P.U(c, 0) + P.U(r, 1) * d[1] * P.U(c, 1)) + P.U(r, 2) * d[2] * P.U(c, 2);
    return t;
  }

  // ObjectSpatula / slab knife pieces (object_spatula.hpp:40-162)
  float knife_sd(const V3& local) const {
    float p[3] = {local[0], local[1], local[2]};
    if (knife.shape == MPMX_KNIFE_SPATULA)
      return sdf_spatula(p, knife.half[0], knife.half[1], knife.half[2], knife.offset);
    return sdf_box(p, knife.half);
  }
  V3 knife_normal(const V3& X) const {
    V3 p = knife_local(ks, X);
    const float e = 0.0001f;
    V3 n;
    for (int d = 0; d < 3; ++d) {
      V3 a = p, b = p;
      a[d] += e; b[d] -= e;
      n[d] = knife_sd(a) - knife_sd(b);
    }
    float nn = norm(n);
    if (nn > 0) n = scale(n, 1.0f / nn);
    V3 w;
    for (int i = 0; i < 3; ++i) w[i] = ks.N[i * 3 + 0] * n[0] + ks.N[i * 3 + 1] * n[1] + ks.N[i * 3 + 2] * n[2];
    float wn = norm(w);
    if (wn > 0) w = scale(w, 1.0f / wn);
    return w;
  }
  V3 knife_velocity(const V3& X) const {
    V3 c = v3(ks.T[3], ks.T[7], ks.T[11]);
    V3 r = sub(X, c);
    V3 om = v3(ks.w[0], ks.w[1], ks.w[2]);
    return add(v3(ks.v[0], ks.v[1], ks.v[2]), cross(om, r));
  }

  // moveObjects (simulation.cpp:284-291) -> ObjectPlate::move
  // (object_plate.hpp:136-162): at the end of the step, velocity = construction
  // velocity (divided by vmin_factor before load_factor frames), then one float
  // step of dt v along the normal (pos_object) and tangential (lower / upper) axis.
  void move_plates() {
    const float frame_dt = (float)(1.0 / cfg.fps);
    for (size_t i = 0; i < plates.size(); ++i) {
      const mpmx_plate& o = plates0[i];
      mpmx_plate& p = plates[i];
      float vmin = o.vmin_factor > 0 ? o.vmin_factor : 1.0f;
      float factor = (time < o.load_factor * frame_dt) ? (float)(1.0 / vmin) : 1.0f;
      for (int c = 0; c < 3; ++c) p.v[c] = o.v[c] * factor;
      switch (p.plate_type) {
        case MPMX_PLATE_LEFT: case MPMX_PLATE_RIGHT:
          p.pos_object += cfg.dt * p.v[0]; p.pos_upper += cfg.dt * p.v[1]; p.pos_lower += cfg.dt * p.v[1]; break;
        case MPMX_PLATE_BOTTOM: case MPMX_PLATE_TOP:
          p.pos_object += cfg.dt * p.v[1]; p.pos_upper += cfg.dt * p.v[0]; p.pos_lower += cfg.dt * p.v[0]; break;
        default:
          p.pos_object += cfg.dt * p.v[2]; break;
      }
    }
  }

  // boundary_collision.cpp:5-316 (3-D branch; no MIBF)
  void boundary_collision(V3 Xi, V3& vi) {
    V3 vi_orig = vi;
    Xi = add(Xi, scale(vi, cfg.dt));
    bool influenced = false;
    if (has_knife) {
      V3 lp = knife_local(ks, Xi);
      if (knife_sd(lp) < (float)(2.0 * dx)) influenced = true;
    }
    if (has_knife) {
      V3 lp = knife_local(ks, Xi);
      bool colliding = knife_sd(lp) < 0.0f;
      if (colliding) {
        V3 v_obj = knife_velocity(Xi);
        V3 v_rel = sub(vi_orig, v_obj);
        if (knife.bc == MPMX_BC_NOSLIP) {
          v_rel = zero3();
        } else if (knife.bc == MPMX_BC_SLIPFREE) {
          V3 n = knife_normal(Xi);
          float d = dot(v_rel, n);
          if (d < 0) {
            float fr = knife.friction;
            V3 vt = sub(v_rel, scale(n, d));
            if (fr > 0) {
              float vtn = norm(vt);
              if (-d * fr < vtn) v_rel = add(vt, scale(scale(vt, 1.0f / vtn), d * fr));
              else v_rel = zero3();
            } else {
              v_rel = vt;
            }
          }
        } else if (knife.bc == MPMX_BC_SLIPSTICK) {
          V3 n = knife_normal(Xi);
          float d = dot(v_rel, n);
          v_rel = sub(v_rel, scale(n, d));
          if (d < 0) {
            float fr = knife.friction;
            if (fr > 0) {
              float vtn = norm(v_rel);
              if (-d * fr < vtn) v_rel = add(v_rel, scale(scale(v_rel, 1.0f / vtn), d * fr));
              else v_rel = zero3();
            }
          }
        } else {
          exit_flag = 1;
          return;
        }
        vi = add(v_rel, v_obj);
        influenced = true;
        if (Xi[1] < 0.8 * dx) {  // thin-film rule (boundary_collision.cpp:93-111)
          V3 n = knife_normal(Xi);
          V3 up = v3(0, 1, 0);
          V3 t = sub(up, scale(n, dot(up, n)));
          float tn = norm(t);
          if (tn > 1e-6f) t = scale(t, 1.0f / tn);
          else t = zero3();
          vi = add(v_obj, scale(t, (float)(0.2 * norm(v_obj))));
        }
        vi_orig = vi;
      }
    }
    for (const mpmx_plate& pl : plates) {
      bool inside = false;
      switch (pl.plate_type) {
        case MPMX_PLATE_LEFT:   inside = Xi[1] < pl.pos_upper && Xi[1] > pl.pos_lower && (Xi[0] - pl.pos_object) <= 0; break;
        case MPMX_PLATE_RIGHT:  inside = Xi[1] < pl.pos_upper && Xi[1] > pl.pos_lower && (pl.pos_object - Xi[0]) <= 0; break;
        case MPMX_PLATE_BOTTOM: inside = Xi[0] < pl.pos_upper && Xi[0] > pl.pos_lower && (Xi[1] - pl.pos_object) <= 0; break;
        case MPMX_PLATE_TOP:    inside = Xi[0] < pl.pos_upper && Xi[0] > pl.pos_lower && (pl.pos_object - Xi[1]) <= 0; break;
        case MPMX_PLATE_BACK:   inside = Xi[2] - pl.pos_object <= 0; break;
        case MPMX_PLATE_FRONT:  inside = pl.pos_object - Xi[2] <= 0; break;
      }
      if (!inside) continue;
      if (influenced && pl.plate_type == MPMX_PLATE_BOTTOM) {
        if (vi[1] < pl.v[1]) { vi[1] = pl.v[1]; vi_orig = vi; }
        continue;
      }
      float r[3] = {vi_orig[0] - pl.v[0], vi_orig[1] - pl.v[1], vi_orig[2] - pl.v[2]};
      int nax;  // normal axis
      if (pl.plate_type == MPMX_PLATE_TOP || pl.plate_type == MPMX_PLATE_BOTTOM) nax = 1;
      else if (pl.plate_type == MPMX_PLATE_LEFT || pl.plate_type == MPMX_PLATE_RIGHT) nax = 0;
      else nax = 2;
      int ta = (nax == 0) ? 1 : 0, tb = (nax == 2) ? 1 : 2;
      if (pl.bc == MPMX_BC_NOSLIP) {
        r[0] = r[1] = r[2] = 0;
      } else if (pl.bc == MPMX_BC_SLIPSTICK || pl.bc == MPMX_BC_SLIPFREE) {
        bool apply = true;
        if (pl.bc == MPMX_BC_SLIPFREE) {
          apply = (pl.plate_type == MPMX_PLATE_TOP && r[1] > 0) || (pl.plate_type == MPMX_PLATE_BOTTOM && r[1] < 0) ||
                  (pl.plate_type == MPMX_PLATE_LEFT && r[0] < 0) || (pl.plate_type == MPMX_PLATE_RIGHT && r[0] > 0) ||
                  (pl.plate_type == MPMX_PLATE_FRONT && r[2] > 0) || (pl.plate_type == MPMX_PLATE_BACK && r[2] < 0);
        }
        if (apply) {
          float fr = pl.friction;
          float vel_t = std::sqrt(r[ta] * r[ta] + r[tb] * r[tb]);
          float fvn = fr * std::fabs(r[nax]);
          if (fr == 0) {
          } else if (vel_t <= fvn) {
            r[ta] = 0; r[tb] = 0;
          } else {
            float sc = (vel_t - fvn) / vel_t;
            r[ta] *= sc; r[tb] *= sc;
          }
          r[nax] = 0;
        }
      } else {
        exit_flag = 1;
        return;
      }
      vi = v3(r[0] + pl.v[0], r[1] + pl.v[1], r[2] + pl.v[2]);
      vi_orig = vi;
    }
  }

  // explicit_euler_update.cpp:7-112
  void euler() {
    size_t n = nodes();
    std::vector<V3> force(n, zero3());
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (int64_t p = 0; p < Np; ++p) {
      if (!active[p]) continue;
      const Nbhd& nbp = nb[p];
      M3 tau = kirchhoff(F[p]);
      int count = 0;
      for (int i = nbp.base[0]; i < nbp.base[0] + 4; ++i)
        for (int j = nbp.base[1]; j < nbp.base[1] + 4; ++j)
          for (int k = nbp.base[2]; k < nbp.base[2] + 4; ++k) {
            const V3& g = nbp.g[count++];
            if (i >= Nx || j >= Ny || k >= Nz) continue;
            int64_t idx = ind(i, j, k);
            if (gmass[idx] > 0) {
              V3 fp = matvec(tau, g);
#pragma omp atomic
              force[idx][0] += fp[0];
#pragma omp atomic
              force[idx][1] += fp[1];
#pragma omp atomic
              force[idx][2] += fp[2];
            }
          }
    }
    float dt_pv = cfg.dt * particle_volume;
    V3 dtg = v3(cfg.dt * cfg.gravity[0], cfg.dt * cfg.gravity[1], cfg.dt * cfg.gravity[2]);
#pragma omp parallel for collapse(3) num_threads(nt) if (nt > 1)
    for (int i = 0; i < Nx; ++i)
      for (int j = 0; j < Ny; ++j)
        for (int k = 0; k < Nz; ++k) {
          int64_t idx = ind(i, j, k);
          float mi = gmass[idx];
          if (mi > 0) {
            V3 inc;
            for (int d = 0; d < 3; ++d) inc[d] = ((-dt_pv) * force[idx][d]) / mi + dtg[d];
            V3 old = gv[idx];
            V3 nv = add(old, inc);
            boundary_collision(v3(gx[i], gy[j], gz[k]), nv);
            gv[idx] = nv;
            gflip[idx] = sub(nv, old);
          }
        }
  }

  // g2p.cpp:7-135
  void g2p() {
    float cur_time = time;
    float tf = 1.0f;
    if (cur_time > cfg.start_boost_time && cfg.end_boost_time > cfg.start_boost_time) {
      float t = (cur_time - cfg.start_boost_time) / (cfg.end_boost_time - cfg.start_boost_time);
      t = std::min(std::max(t, 0.0f), 1.0f);
      tf = 1.0f + t * (cfg.boost_factor - 1.0f);
    }
    float dynD = cfg.D_max * tf;
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (int64_t p = 0; p < Np; ++p) {
      pic[p] = zero3(); flip[p] = zero3(); Bm[p] = zero33();
      if (!active[p]) continue;
      const Nbhd& nbp = nb[p];
      const V3& xp = x[p];
      V3 vp = zero3(), fp = zero3();
      M3 Bp = zero33(), L = zero33();
      L7 gp; gp.fill(0.0f);
      float gs = 0.0f;
      int count = 0;
      for (int i = nbp.base[0]; i < nbp.base[0] + 4; ++i)
        for (int j = nbp.base[1]; j < nbp.base[1] + 4; ++j)
          for (int k = nbp.base[2]; k < nbp.base[2] + 4; ++k) {
            float w = nbp.w[count];
            const V3& g = nbp.g[count];
            ++count;
            if (i >= Nx || j >= Ny || k >= Nz) continue;
            int64_t idx = ind(i, j, k);
            const V3& gvv = gv[idx];
            vp = add(vp, scale(gvv, w));
            if (cfg.flip_ratio < 0) {
              V3 pd = v3(gx[i] - xp[0], gy[j] - xp[1], gz[k] - xp[2]);
              for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) Bp(r, c) += (gvv[r] * pd[c]) * w;
            }
            if (cfg.flip_ratio >= -1) fp = add(fp, scale(gflip[idx], w));
            for (int r = 0; r < 3; ++r)
              for (int c = 0; c < 3; ++c) L(r, c) += gvv[r] * g[c];
            for (int c = 0; c < 7; ++c) gp[c] += gpig[idx][c] * w;
            gs += gshear[idx] * w;
          }
      pic[p] = vp;
      if (cfg.flip_ratio < 0) Bm[p] = Bp;
      if (cfg.flip_ratio >= -1) flip[p] = fp;
      float sn = 0.0f;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          float s = 0.5f * (L(r, c) + L(c, r));
          sn += s * s;
        }
      diff[p] = std::sqrt(sn);
      float tt = std::max(0.0f, std::min(1.0f, (gs - cfg.D_edge0) / (cfg.D_edge1 - cfg.D_edge0)));
      float mixf = tt * tt * (3.0f - 2.0f * tt);
      float smf = std::min(std::max(dynD * mixf * cfg.dt, 0.0f), 1.0f);
      for (int c = 0; c < 7; ++c) pig[p][c] = (1.0f - smf) * pig[p][c] + smf * gp[c];
      float s = ((pig[p][0] + pig[p][1]) + pig[p][2]) + pig[p][3];
      if (s > 1e-6f) for (int c = 0; c < 4; ++c) pig[p][c] /= s;
    }
  }

  // musl.cpp:6-98
  void musl() {
    float a = -cfg.flip_ratio;
    for (int64_t p = 0; p < Np; ++p) {
      if (!active[p]) continue;
      if (cfg.flip_ratio < -1) v[p] = pic[p];
      else if (cfg.flip_ratio < 0) v[p] = add(scale(add(v[p], flip[p]), a), scale(pic[p], 1 - a));
      else v[p] = add(scale(add(v[p], flip[p]), cfg.flip_ratio), scale(pic[p], 1 - cfg.flip_ratio));
    }
    size_t n = nodes();
    std::fill(gv.begin(), gv.end(), zero3());
#pragma omp parallel num_threads(nt) if (nt > 1)
    {
      std::vector<V3> lv(n, zero3());
#pragma omp for nowait
      for (int64_t p = 0; p < Np; ++p) {
        if (!active[p]) continue;
        const V3& xp = x[p];
        const Nbhd& nbp = nb[p];
        int count = 0;
        for (int i = nbp.base[0]; i < nbp.base[0] + 4; ++i)
          for (int j = nbp.base[1]; j < nbp.base[1] + 4; ++j)
            for (int k = nbp.base[2]; k < nbp.base[2] + 4; ++k) {
              float w = nbp.w[count++];
              if (w > 1e-25f) {
                int64_t idx = ind(i, j, k);
                lv[idx] = add(lv[idx], scale(v[p], w));
                if (cfg.flip_ratio < 0) {
                  V3 pd = v3(gx[i] - xp[0], gy[j] - xp[1], gz[k] - xp[2]);
                  lv[idx] = add(lv[idx], scale(scale(matvec(Bm[p], pd), apicDinverse), w));
                }
              }
            }
      }
#pragma omp critical
      for (size_t l = 0; l < n; ++l) gv[l] = add(gv[l], lv[l]);
    }
    for (size_t l = 0; l < n; ++l) {
      float mi = gmass[l];
      if (mi > 0) {
        float r = mi / particle_mass;
        gv[l] = v3(gv[l][0] / r, gv[l][1] / r, gv[l][2] / r);
      } else {
        gv[l] = zero3();
      }
    }
  }

 public:
  // Plastic state of one particle carried by the return maps.
  struct PState {
    float eps_dev = 0, eps_vol = 0, prad = 0, dg = 0;  // eps_pl_dev, eps_pl_vol, eps_pl_vol_pradhana, delta_gamma
  };

  // Perzyna overstress increment (plasticity.cpp:187-229 and 276-320): the
  // closed form for perzyna_exp == 1 (with xi == 0 where hardening applies),
  // else Newton on r(dgam) = (q_tr - f dgam)(dt/(visc dgam + dt))^n - q_y(dgam).
  // `harden` selects VMVisc's softening yield q_y(eps + dgam/d); DPVisc's q_y
  // is the fixed q_yield.
  float perzyna_dgamma(float q_trial, float q_yield, bool closed, bool harden, float eps_dev_in,
                       int& err) const {
    float dgam;
    if (closed) {
      dgam = (q_trial - q_yield) / (f_mu_prefac + q_yield * cfg.perzyna_visc / cfg.dt);
      if (dgam < 0) err = 1;
      return dgam;
    }
    dgam = 0.01f * (q_trial - q_yield) / f_mu_prefac;
    const int max_iter = 60;
    for (int it = 0; it < max_iter; ++it) {
      if (it == max_iter - 1) err = 1;
      float tm = cfg.perzyna_visc * dgam + cfg.dt;
      float tmp = cfg.dt / tm;
      float tmp1 = std::pow(tmp, cfg.perzyna_exp);
      float qyd = 0.0f;
      if (harden) {
        q_yield = cfg.q_min + (cfg.q_max - cfg.q_min) * std::exp(-cfg.xi * (eps_dev_in + (1.0f / d_prefac) * dgam));
        qyd = -cfg.xi / d_prefac * (cfg.q_max - cfg.q_min) * std::exp(-cfg.xi * (eps_dev_in + (1.0f / d_prefac) * dgam));
      }
      float res = (q_trial - f_mu_prefac * dgam) * tmp1 - q_yield;
      if (std::fabs(res) < 1e-2f) break;
      float rd = -f_mu_prefac * tmp1 + (q_trial - f_mu_prefac * dgam) * cfg.perzyna_exp *
                                           std::pow(tmp, cfg.perzyna_exp - 1) * (-cfg.perzyna_visc * cfg.dt) / (tm * tm) - qyd;
      if (std::fabs(rd) < 1e-14f) err = 1;
      dgam -= res / rd;
      if (dgam < 0) dgam = 1e-10f;
    }
    return dgam;
  }

  // F = U diag(e^h_new) V^T (plasticity.cpp:233).  While no singular value was
  // clamped (e^h_old = sigma), this equals F_trial V diag(e^(h_new - h_old)) V^T,
  // evaluated as F_trial + F_trial V diag(expm1(h_new - h_old)) V^T so that the
  // small correction keeps its relative precision.
  static M3 f_from_hencky(const M3& Ft, const Principal& P, const V3& h_old, const V3& h_new) {
    bool clamped = false;
    for (int d = 0; d < 3; ++d) clamped |= !(1.0f + P.lam[d] >= 1e-8f);
    M3 out;
    if (clamped) {
      M3 D = zero33();
      for (int d = 0; d < 3; ++d) D(d, d) = std::exp(h_new[d]);
      return matmul(matmul(P.U, D), transpose(P.V));
    }
    float e[3];
    for (int d = 0; d < 3; ++d) e[d] = std::expm1(h_new[d] - h_old[d]);
    M3 W;  // V diag(e) V^T
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        W(r, c) = (P.V(r, 0) * e[0] * P.V(c, 0) + P.V(r, 1) * e[1] * P.V(c, 1)) + P.V(r, 2) * e[2] * P.V(c, 2);
    M3 FW = matmul(Ft, W);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) out(r, c) = Ft(r, c) + FW(r, c);
    return out;
  }

  // plasticity.cpp:17-27: principal Hencky strain log(max(sigma, 1e-4)) (sigma
  // < 1e-4 is 1 + lam < 1e-8), its trace and unit deviator.
  static void hencky_split(const Principal& P, V3& h, float& htr, V3& dev, float& dn) {
    for (int d = 0; d < 3; ++d) h[d] = (1.0f + P.lam[d] >= 1e-8f) ? log_sigma(P.lam[d]) : std::log(1e-4f);
    htr = (h[0] + h[1]) + h[2];
    dev = v3(h[0] - htr / 3, h[1] - htr / 3, h[2] - htr / 3);
    dn = norm(dev);
    if (dn > 0) dev = scale(dev, 1.0f / dn);
  }

  // plasticity.cpp:163-238 (VMVisc) on the trial state's principal frame.  Sets `changed`
  // when F is replaced; returns 1 when the reference counts a projection
  // (plastic_count++: not for the capped-pressure branch, :172-180).
  int vmvisc(const M3& Ft, const Principal& P, const PState& in, M3& Fout, PState& out, bool& changed,
             int& err) const {
    V3 h, dev;
    float htr, dn;
    hencky_split(P, h, htr, dev, dn);
    float q_trial = e_mu_prefac * dn;
    float q_yield = cfg.q_min + (cfg.q_max - cfg.q_min) * std::exp(-cfg.xi * in.eps_dev);
    float p_trial = -K * htr;
    out = in;
    out.dg = 0.0f;
    changed = false;
    if (p_trial < cfg.p_min * std::exp(-cfg.xi * in.eps_vol)) {
      // capped von Mises: F = U V^T
      float dgam = q_trial / f_mu_prefac;
      float inst = -p_trial / K;
      Fout = matmul(P.U, transpose(P.V));
      changed = true;
      out.eps_vol = in.eps_vol + inst;
      out.eps_dev = (float)((double)in.eps_dev + (1.0 / (double)d_prefac) * (double)dgam);
      out.dg = dgam / cfg.dt;
      return 0;
    }
    if (q_trial > q_yield) {
      float dgam = perzyna_dgamma(q_trial, q_yield, cfg.perzyna_exp == 1 && cfg.xi == 0, true, in.eps_dev, err);
      // (1.0/d_prefac) is a double in the reference; Eigen converts the product
      // to the float scalar type before scaling the vector (plasticity.cpp:232-234).
      float fac = (float)((1.0 / (double)d_prefac) * (double)dgam);
      Fout = f_from_hencky(Ft, P, h, sub(h, scale(dev, fac)));
      changed = true;
      out.eps_dev = (float)((double)in.eps_dev + (1.0 / (double)d_prefac) * (double)dgam);
      out.dg = dgam / cfg.dt;
      return 1;
    }
    return 0;
  }

  // plasticity.cpp:240-340 (DPVisc; use_mibf is off, as prepareSimulation sets
  // it, simulation_units.cpp:185).  Pradhana's correction shifts the yield
  // surface by the volume gained at the tip (p_shift = -K eps_pl_vol_pradhana),
  // resets the accumulation on every state right of the shifted tip.
  int dpvisc(const M3& Ft, const Principal& P, const PState& in, M3& Fout, PState& out, bool& changed,
             int& err) const {
    V3 h, dev;
    float htr, dn;
    hencky_split(P, h, htr, dev, dn);
    float p_trial = -K * htr;
    float q_trial = e_mu_prefac * dn;
    float p_tip = -cfg.q_cohesion / cfg.M;
    float p_shift = 0.0f;
    if (cfg.use_pradhana) p_shift = -K * in.prad;
    float q_yield = cfg.M * (p_trial + p_shift) + cfg.q_cohesion;
    out = in;
    changed = false;
    int count = 0;
    if ((p_trial + p_shift) < p_tip) {
      // left of the shifted tip: project to the cohesion-only tip
      float dgam = d_prefac * dn;
      float p_proj = p_tip;
      float inst = (p_proj - p_trial) / K;
      count = 1;
      float hv = -p_proj / (K * 3);
      Fout = f_from_hencky(Ft, P, h, v3(hv, hv, hv));
      changed = true;
      out.dg = dgam / cfg.dt;
      out.eps_dev = (float)((double)in.eps_dev + (1.0 / (double)d_prefac) * (double)dgam);
      out.eps_vol = in.eps_vol + inst;
      if (cfg.use_pradhana) out.prad = in.prad + inst;
    } else {
      if (cfg.use_pradhana) out.prad = 0.0f;
      out.dg = 0.0f;
    }
    if ((p_trial + p_shift) > p_tip && q_trial > q_yield) {
      count = 1;
      float dgam = perzyna_dgamma(q_trial, q_yield, cfg.perzyna_exp == 1, false, in.eps_dev, err);
      float fac = (float)((1.0 / (double)d_prefac) * (double)dgam);
      Fout = f_from_hencky(Ft, P, h, sub(h, scale(dev, fac)));
      changed = true;
      out.eps_dev = (float)((double)in.eps_dev + (1.0 / (double)d_prefac) * (double)dgam);
      out.dg = dgam / cfg.dt;
    }
    return count;
  }

  // plasticity.cpp:10-27 dispatch on the model.
  int return_map(const M3& Ft, const Principal& P, const PState& in, M3& Fout, PState& out, bool& changed,
                 int& err) const {
    if (cfg.plastic_model == MPMX_PLASTIC_VMVISC) return vmvisc(Ft, P, in, Fout, out, changed, err);
    if (cfg.plastic_model == MPMX_PLASTIC_DPVISC) return dpvisc(Ft, P, in, Fout, out, changed, err);
    out = in;
    out.dg = 0.0f;
    changed = false;
    return 0;
  }

  // deformation_update.cpp:7-52 + plasticity.cpp:10-27
  void deformation_update() {
    int64_t pc = 0;
    int err_any = 0;
#pragma omp parallel for reduction(+ : pc) reduction(| : err_any) num_threads(nt) if (nt > 1)
    for (int64_t p = 0; p < Np; ++p) {
      dgamma[p] = 0.0f;
      if (!active[p]) continue;
      const Nbhd& nbp = nb[p];
      M3 sum = zero33();
      int count = 0;
      for (int i = nbp.base[0]; i < nbp.base[0] + 4; ++i)
        for (int j = nbp.base[1]; j < nbp.base[1] + 4; ++j)
          for (int k = nbp.base[2]; k < nbp.base[2] + 4; ++k) {
            const V3& g = nbp.g[count++];
            if (i >= Nx || j >= Ny || k >= Nz) continue;
            const V3& gvv = gv[ind(i, j, k)];
            for (int r = 0; r < 3; ++r)
              for (int c = 0; c < 3; ++c) sum(r, c) += gvv[r] * g[c];
          }
      M3 dL;
      for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) dL(r, c) = cfg.dt * sum(r, c);
      M3 prod = matmul(dL, F[p]);
      M3 Ft;
      for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) Ft(r, c) = F[p](r, c) + prod(r, c);
      F[p] = Ft;
      if (cfg.plastic_model != MPMX_PLASTIC_NONE) {
        const Principal P = principal_of(Ft);
        M3 Fo;
        PState in, out;
        in.eps_dev = eps_pl_dev[p]; in.eps_vol = eps_pl_vol[p]; in.prad = eps_prad[p];
        bool changed = false;
        int err = 0;
        pc += return_map(Ft, P, in, Fo, out, changed, err);
        if (changed) F[p] = Fo;
        eps_pl_dev[p] = out.eps_dev; eps_pl_vol[p] = out.eps_vol; eps_prad[p] = out.prad; dgamma[p] = out.dg;
        err_any |= err;
      }
    }
    plastic_count = pc;
    if (err_any) exit_flag = 1;
  }

  // position_update.cpp:6-91 (+ the fixed-lattice domain rule)
  void position_update() {
    float a = -cfg.flip_ratio;
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
    for (int64_t p = 0; p < Np; ++p) {
      if (!active[p]) continue;
      x[p] = add(x[p], scale(pic[p], cfg.dt));
      if (!cfg.use_musl) {
        if (cfg.flip_ratio < -1) v[p] = pic[p];
        else if (cfg.flip_ratio < 0) v[p] = add(scale(add(v[p], flip[p]), a), scale(pic[p], 1 - a));
        else v[p] = add(scale(add(v[p], flip[p]), cfg.flip_ratio), scale(pic[p], 1 - cfg.flip_ratio));
      }
      bool bad = false;
      for (int d = 0; d < 3; ++d)
        if (!std::isfinite(x[p][d]) || !std::isfinite(v[p][d])) bad = true;
      if (bad) {
        active[p] = 0;
        for (int d = 0; d < 3; ++d) {
          v[p][d] = 0.0f;
          x[p][d] = cfg.use_particle_boundaries ? cfg.particle_boundary_min[d] : 0.0f;
        }
        continue;
      }
      if (x[p][1] < 0.0f) {
        x[p][1] = 0.0f;
        if (v[p][1] < 0.0f) v[p][1] = 0.0f;
      }
      if (cfg.use_particle_boundaries) {
        for (int d = 0; d < 3; ++d) {
          if (x[p][d] < cfg.particle_boundary_min[d]) { x[p][d] = cfg.particle_boundary_min[d]; active[p] = 0; v[p][d] = 0; }
          else if (x[p][d] > cfg.particle_boundary_max[d]) { x[p][d] = cfg.particle_boundary_max[d]; active[p] = 0; v[p][d] = 0; }
        }
        if (!active[p]) { v[p] = zero3(); continue; }
      }
      if (!stencil_inside(x[p])) { active[p] = 0; v[p] = zero3(); }
    }
  }

  void advance_step() {
    if (has_knife) ks = knife_state(knife, frames, frame_fps, cfg.dt, step);
    resize_grid();
    precompute_weights();
    p2g();
    euler();
    g2p();
    if (cfg.use_musl) musl();
    deformation_update();
    position_update();
    move_plates();
    time += cfg.dt;
    ++step;
  }
};

}  // namespace

// ======================================================================= C API
extern "C" {

int orc_return_map_state(const mpmx_config* cfg, int64_t n, const float* Ft9, const float* in3, float* F9,
                         float* out4, int32_t* changed, int32_t* counted);

void* orc_create(const mpmx_config* cfg) {
  Oracle* o = new Oracle();
  o->cfg = *cfg;
  o->dx = cfg->dx;
  o->nt = cfg->n_threads > 0 ? cfg->n_threads : 1;
  return o;
}
void orc_destroy(void* h) { delete (Oracle*)h; }

int orc_upload(void* h, int64_t n, const float* x, const float* v, const float* F, const float* C,
               const float* pig, const float* diff, const float* eps, const uint8_t* active) {
  Oracle* o = (Oracle*)h;
  o->Np = n;
  o->x.resize(n); o->v.resize(n); o->pic.assign(n, zero3()); o->flip.assign(n, zero3());
  o->F.resize(n); o->Bm.resize(n); o->pig.resize(n); o->diff.resize(n); o->eps_pl_dev.resize(n);
  o->dgamma.assign(n, 0.0f); o->active.resize(n);
  o->eps_pl_vol.assign(n, 0.0f); o->eps_prad.assign(n, 0.0f);
  for (int64_t p = 0; p < n; ++p) {
    for (int d = 0; d < 3; ++d) {
      o->x[p][d] = x[3 * p + d];
      o->v[p][d] = v ? v[3 * p + d] : 0.0f;
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        o->F[p](r, c) = F ? F[9 * p + 3 * r + c] : (r == c ? 1.0f : 0.0f);
        o->Bm[p](r, c) = C ? C[9 * p + 3 * r + c] : 0.0f;
      }
    for (int c = 0; c < 7; ++c) o->pig[p][c] = pig ? pig[7 * p + c] : 0.0f;
    o->diff[p] = diff ? diff[p] : 0.0f;
    o->eps_pl_dev[p] = eps ? eps[p] : 0.0f;
    o->active[p] = active ? active[p] : 1;
  }
  return 0;
}

int orc_set_objects(void* h, const mpmx_plate* plates, int np, const mpmx_knife* knives, int nk) {
  Oracle* o = (Oracle*)h;
  o->plates.assign(plates, plates + np);
  o->plates0 = o->plates;
  o->has_knife = nk > 0;
  if (nk > 0) o->knife = knives[0];
  return 0;
}

int orc_set_keyframes(void* h, const float* frames7, int n, float fps) {
  Oracle* o = (Oracle*)h;
  o->frames.assign(frames7, frames7 + 7 * (size_t)n);
  o->frame_fps = fps;
  return 0;
}

int orc_prepare(void* h) {
  Oracle* o = (Oracle*)h;
  o->derived();
  o->make_lattice();
  for (int64_t p = 0; p < o->Np; ++p)
    if (o->active[p] && !o->stencil_inside(o->x[p])) { o->active[p] = 0; o->v[p] = zero3(); }
  o->prepared = true;
  o->time = 0;
  o->step = 0;
  return 0;
}

// Step of the following steps (the caller's updateDt, update_dt.cpp:5-71).
// Only for scenes without a moving knife or moving plates: this restatement
// evaluates their kinematics at step * dt.
int orc_set_dt(void* h, float dt) {
  Oracle* o = (Oracle*)h;
  if (!(dt > 0.0f)) return MPMX_ERR_ARG;
  o->cfg.dt = dt;
  if (o->prepared) o->derived();
  return 0;
}

int orc_step(void* h, int n) {
  Oracle* o = (Oracle*)h;
  if (!o->prepared) return MPMX_ERR_STATE;
  for (int i = 0; i < n; ++i) o->advance_step();
  return o->exit_flag ? MPMX_ERR_DEVICE : 0;
}

int orc_download(void* h, float* x, float* v, float* F, float* C, float* pig, float* diff, float* eps,
                 float* dg, uint8_t* active) {
  Oracle* o = (Oracle*)h;
  for (int64_t p = 0; p < o->Np; ++p) {
    for (int d = 0; d < 3; ++d) {
      if (x) x[3 * p + d] = o->x[p][d];
      if (v) v[3 * p + d] = o->v[p][d];
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        if (F) F[9 * p + 3 * r + c] = o->F[p](r, c);
        if (C) C[9 * p + 3 * r + c] = o->Bm[p](r, c);
      }
    for (int c = 0; c < 7; ++c) if (pig) pig[7 * p + c] = o->pig[p][c];
    if (diff) diff[p] = o->diff[p];
    if (eps) eps[p] = o->eps_pl_dev[p];
    if (dg) dg[p] = o->dgamma[p];
    if (active) active[p] = o->active[p];
  }
  return 0;
}

int orc_get_lattice(void* h, int* dims, float* origin, float* xs, float* ys, float* zs) {
  Oracle* o = (Oracle*)h;
  dims[0] = o->Nx; dims[1] = o->Ny; dims[2] = o->Nz;
  origin[0] = o->xc; origin[1] = o->yc; origin[2] = o->zc;
  if (xs) std::copy(o->gx.begin(), o->gx.end(), xs);
  if (ys) std::copy(o->gy.begin(), o->gy.end(), ys);
  if (zs) std::copy(o->gz.begin(), o->gz.end(), zs);
  return 0;
}

// Grid of the last completed step (after explicitEulerUpdate / MUSL).
int orc_get_grid(void* h, float* mass, float* v, float* flip, float* pig, float* shear) {
  Oracle* o = (Oracle*)h;
  size_t n = o->gmass.size();
  for (size_t l = 0; l < n; ++l) {
    if (mass) mass[l] = o->gmass[l];
    for (int d = 0; d < 3; ++d) {
      if (v) v[3 * l + d] = o->gv[l][d];
      if (flip) flip[3 * l + d] = o->gflip[l][d];
    }
    for (int c = 0; c < 7; ++c) if (pig) pig[7 * l + c] = o->gpig[l][c];
    if (shear) shear[l] = o->gshear[l];
  }
  return 0;
}

int orc_stats(void* h, int64_t* plastic_count, int* exit_flag, int64_t* n_active, double* time) {
  Oracle* o = (Oracle*)h;
  if (plastic_count) *plastic_count = o->plastic_count;
  if (exit_flag) *exit_flag = o->exit_flag;
  if (n_active) {
    int64_t c = 0;
    for (auto a : o->active) c += a;
    *n_active = c;
  }
  if (time) *time = o->time;
  return 0;
}

int orc_knife_state_at(const mpmx_knife* k, const float* frames7, int nframes, float fps, double dt,
                       int64_t step, float* T12, float* Tinv12, float* N9, float* v3out, float* w3out) {
  std::vector<float> fr;
  if (frames7 && nframes > 0) fr.assign(frames7, frames7 + 7 * (size_t)nframes);
  KState s = knife_state(*k, fr, fps, dt, step);
  std::memcpy(T12, s.T, sizeof(s.T));
  std::memcpy(Tinv12, s.Tinv, sizeof(s.Tinv));
  std::memcpy(N9, s.N, sizeof(s.N));
  std::memcpy(v3out, s.v, sizeof(s.v));
  std::memcpy(w3out, s.w, sizeof(s.w));
  return 0;
}

// ---- unit-level entry points for tests
// Principal frame of F (principal_of): U, sigma, V (row-major) and lam = sigma^2 - 1.
void orc_principal(const float* A9, float* U9, float* s3, float* V9, float* lam3) {
  M3 A;
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A(r, c) = A9[3 * r + c];
  const Principal P = principal_of(A);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) { U9[3 * r + c] = P.U(r, c); V9[3 * r + c] = P.V(r, c); }
  for (int d = 0; d < 3; ++d) { s3[d] = P.sig[d]; lam3[d] = P.lam[d]; }
}

// One grid node's boundary collision (boundary_collision.cpp:5-316) against the
// given plates and knife (its state at `step`; NULL: no knife), for tests:
// out3 = the node velocity after the collision rules.
int orc_boundary_collision(const mpmx_config* cfg, const mpmx_plate* plates, int np, const mpmx_knife* knife,
                           int64_t step, const float* X3, const float* v3in, float* out3) {
  Oracle o;
  o.cfg = *cfg;
  o.dx = cfg->dx;
  o.derived();
  if (np > 0) o.plates.assign(plates, plates + np);
  o.plates0 = o.plates;
  for (int64_t k = 0; k < step; ++k) {   // the plates' state during `step`
    o.move_plates();
    o.time += cfg->dt;
  }
  o.has_knife = knife != nullptr;
  if (knife) {
    o.knife = *knife;
    o.ks = knife_state(*knife, o.frames, o.frame_fps, cfg->dt, step);
  }
  V3 X = v3(X3[0], X3[1], X3[2]), v = v3(v3in[0], v3in[1], v3in[2]);
  o.boundary_collision(X, v);
  for (int d = 0; d < 3; ++d) out3[d] = v[d];
  return o.exit_flag ? MPMX_ERR_DEVICE : 0;
}

float orc_spline_N(float u) { return spline_N(u); }
float orc_spline_dN(float u) { return spline_dN(u); }

float orc_knife_sdf(int shape, const float* half, float offset, const float* p) {
  if (shape == MPMX_KNIFE_SPATULA) return sdf_spatula(p, half[0], half[1], half[2], offset);
  return sdf_box(p, half);
}

// The configured return map (VMVisc or DPVisc) on F_trial (n particles):
// writes F, eps_pl_dev, dgamma and a flag set where F was replaced.
int orc_return_map(const mpmx_config* cfg, int64_t n, const float* Ft9, const float* eps_in, float* F9,
                   float* eps_out, float* dg_out, int32_t* plastic) {
  std::vector<float> in3(3 * (size_t)n, 0.0f), out4(4 * (size_t)n);
  for (int64_t p = 0; p < n; ++p) in3[3 * p] = eps_in ? eps_in[p] : 0.0f;
  int rc = orc_return_map_state(cfg, n, Ft9, in3.data(), F9, out4.data(), plastic, nullptr);
  for (int64_t p = 0; p < n; ++p) { eps_out[p] = out4[4 * p]; dg_out[p] = out4[4 * p + 3]; }
  return rc;
}

// Same with the whole plastic state: in3 = (eps_pl_dev, eps_pl_vol,
// eps_pl_vol_pradhana) per particle, out4 = the same three after the map and
// delta_gamma.  changed[p]: F replaced; counted[p] (may be NULL): the
// reference's plastic_count increment.
int orc_return_map_state(const mpmx_config* cfg, int64_t n, const float* Ft9, const float* in3, float* F9,
                         float* out4, int32_t* changed, int32_t* counted) {
  Oracle o;
  o.cfg = *cfg;
  o.dx = cfg->dx;
  o.derived();
  int err_any = 0;
  for (int64_t p = 0; p < n; ++p) {
    M3 A, Fo;
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A(r, c) = Ft9[9 * p + 3 * r + c];
    const Principal P = principal_of(A);
    Oracle::PState in, out;
    in.eps_dev = in3[3 * p]; in.eps_vol = in3[3 * p + 1]; in.prad = in3[3 * p + 2];
    bool ch = false;
    int err = 0;
    int cnt = o.return_map(A, P, in, Fo, out, ch, err);
    if (!ch) Fo = A;
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) F9[9 * p + 3 * r + c] = Fo(r, c);
    out4[4 * p] = out.eps_dev; out4[4 * p + 1] = out.eps_vol; out4[4 * p + 2] = out.prad; out4[4 * p + 3] = out.dg;
    changed[p] = ch ? 1 : 0;
    if (counted) counted[p] = cnt;
    err_any |= err;
  }
  return err_any ? MPMX_ERR_DEVICE : 0;
}

// Set eps_pl_vol / eps_pl_vol_pradhana after orc_upload (NULL: zeros).
int orc_upload_vol(void* h, const float* eps_vol, const float* prad) {
  Oracle* o = (Oracle*)h;
  for (int64_t p = 0; p < o->Np; ++p) {
    o->eps_pl_vol[p] = eps_vol ? eps_vol[p] : 0.0f;
    o->eps_prad[p] = prad ? prad[p] : 0.0f;
  }
  return 0;
}

// eps_pl_vol and eps_pl_vol_pradhana per particle (either may be NULL).
int orc_download_vol(void* h, float* eps_vol, float* prad) {
  Oracle* o = (Oracle*)h;
  for (int64_t p = 0; p < o->Np; ++p) {
    if (eps_vol) eps_vol[p] = o->eps_pl_vol[p];
    if (prad) prad[p] = o->eps_prad[p];
  }
  return 0;
}

// Kirchhoff stress tau = P(F) F^T (HenckyPiola), for tests.
void orc_kirchhoff(const mpmx_config* cfg, const float* F9, float* tau9) {
  Oracle o;
  o.cfg = *cfg;
  o.dx = cfg->dx;
  o.derived();
  M3 A;
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A(r, c) = F9[3 * r + c];
  M3 t = o.kirchhoff(A);
  for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) tau9[3 * r + c] = t(r, c);
}

int orc_omp_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
