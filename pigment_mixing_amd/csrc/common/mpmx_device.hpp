// mpmx_device.hpp — device-side data layout and per-particle / per-node math of
// the MI355X-native MPM step (host-compilable for CPU unit checks).
//
// Layout in HBM
//   Node blocks: the lattice is tiled in 4x4x4-node blocks.  A scene's blocks are
//   numbered (bi*BY + bj)*BZ + bk from the scene's `node_block_base` in a pool
//   shared by all scenes of a batch.  Inside a block the 64 nodes are ordered
//   ((li*4 + lj)*4 + lk), and every channel is a contiguous 64-float row:
//   ACC[(blk*ACC_CH + ch)*64 + local], GV[(blk*GV_CH + ch)*64 + local],
//   MASSFX[blk*64 + local] (u64 fixed point).  A wave of 64 lanes touching one
//   block channel reads/writes 256 contiguous bytes.
//   Particles: structure of arrays, physically sorted by particle key (the node
//   block holding the particle's stencil base), ping-ponged every step.
//
// Reference math restated here: quadratic B-spline (tools.hpp:153-177), knife
// SDFs (object_spatula.hpp:40-82, exact box), boundaryCollision
// (boundary_collision.cpp:5-316), VMVisc return map (plasticity.cpp:163-238),
// Hencky Kirchhoff stress (simulation.hpp:287-294).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mpmx.h"

#ifndef MPMX_HD
#define MPMX_HD __host__ __device__ __forceinline__
#endif

namespace mpmx {

constexpr int BLK = 4;            // nodes per block edge
constexpr int BLK3 = 64;          // nodes per block
constexpr int ACC_CH = 15;        // mom3, force3, pig7, shear1, mass1 (float accumulators)
constexpr int GV_CH = 15;         // v3, flip3, pig7, shear1, mass1
constexpr int NPIG = 7;
constexpr int MAX_PLATES = 4;

// channel indices
enum { A_MOM = 0, A_FRC = 3, A_PIG = 6, A_SHR = 13, A_MASS = 14 };
enum { G_V = 0, G_FLIP = 3, G_PIG = 6, G_SHR = 13, G_MASS = 14 };

// device error bits (the reference's `exit = 1` causes)
enum : uint32_t {
  ERR_NEG_DGAMMA = 1u,       // plasticity.cpp:189-192
  ERR_NEWTON = 2u,           // plasticity.cpp:200-203, 219-222
  ERR_BAD_BC = 4u,           // boundary_collision.cpp:84-88
  ERR_TILE_OVERFLOW = 8u,    // internal: a particle moved > 1 cell in one step
  ERR_QUEUE_OVERFLOW = 16u,  // internal: a closure-queue segment ran past its particle range
};

struct PlateDev {
  float pos_object, friction, pos_lower, pos_upper;
  float v[3];
  int plate_type, bc, pad;
};

struct KnifeShapeDev {
  int present, shape, bc, pad;
  float half[3];
  float offset, friction, pad2, pad3;
};

struct KnifeStateDev {   // per scene, per step: the moving objects' state
  float T[12];
  float Tinv[12];
  float N[9];
  float v[3];
  float w[3];
  float time;            // simulated time at the step start (for pigment boost)
  // plates during this step (ObjectPlate::move, object_plate.hpp:136-162):
  // pos_object, pos_lower, pos_upper, vx, vy, vz
  float plate[MAX_PLATES][6];
};

struct SceneDev {
  // lattice
  int dims[3];
  int bdims[3];
  int node_block_base;
  int axis_off;              // offset of this scene's xs/ys/zs in the node-coordinate table
  float origin[3];
  float dx, one_over_dx;
  // material / step constants (simulation_units.cpp:167-183)
  float E, nu, mu, lambda, K;
  float inv_d_prefac;        // 1/d_prefac (= q_prefac), the factor of plasticity.cpp:232-234
  float e_mu_prefac, f_mu_prefac;
  float q_min, q_max, xi, p_min, perzyna_exp, perzyna_visc;
  float flip_ratio, apicDinv, particle_mass, particle_volume, dt;
  float gravity[3];
  float D_max, D_edge0, D_edge1, start_boost, end_boost, boost_factor;
  int plastic_model, use_musl, use_pb, closure;
  // DPVisc (plasticity.cpp:240-340): M = tan(friction angle), cohesion, Pradhana
  float M, q_cohesion, d_prefac;
  int use_pradhana;
  int track_vol;             // carry eps_pl_vol / eps_pl_vol_pradhana rows (DPVisc or a capped VMVisc)
  float pb_min[3], pb_max[3];
  float closure_tol;
  int n_plates;
  PlateDev plates[MAX_PLATES];
  KnifeShapeDev knife;
};

// ------------------------------------------------------------------- helpers
MPMX_HD int node_block_of(const SceneDev& s, int i, int j, int k) {
  return s.node_block_base + ((i >> 2) * s.bdims[1] + (j >> 2)) * s.bdims[2] + (k >> 2);
}
MPMX_HD int node_local(int i, int j, int k) { return (((i & 3) << 2) | (j & 3)) << 2 | (k & 3); }

// ---------------------------------------------------------- quadratic spline
// tools.hpp:153-177.  The reference evaluates with double constants and rounds
// to float; for |u| < 0.5, 0.75 - fl(u*u) is exact in double, and for
// 0.5 <= |u| < 1.5, t = 1.5 - |u| is exact in float and 0.5*fl(t*t) equals the
// rounded double 0.5*t*t, so the unfused float forms below are bit-identical.
// The products must not be contracted into FMAs: HIP's __fmul_rn / __fsub_rn
// are plain operators here, so these helpers switch contraction off.
//
// u = (x - X_node) / dx as the reference rounds it (one product, no FMA with
// what follows).
MPMX_HD float spline_u(float x, float X, float inv_dx) {
#pragma clang fp contract(off)
  return (x - X) * inv_dx;
}
MPMX_HD float spline_N(float u) {
#pragma clang fp contract(off)
  float a = fabsf(u);
#ifdef __HIP_DEVICE_COMPILE__
  if (a < 0.5f) return __fsub_rn(0.75f, __fmul_rn(a, a));
  if (a < 1.5f) { float t = __fsub_rn(1.5f, a); return __fmul_rn(0.5f, __fmul_rn(t, t)); }
#else
  if (a < 0.5f) return (float)(0.75 - (double)(a * a));
  if (a < 1.5f) return (float)(0.5 * (1.5 - (double)a) * (1.5 - (double)a));
#endif
  return 0.0f;
}
MPMX_HD float spline_dN(float u) {
#pragma clang fp contract(off)
  float a = fabsf(u);
  if (a < 0.5f) return -2.0f * u;
  if (a < 1.5f) return u > 0.0f ? u - 1.5f : (u < 0.0f ? u + 1.5f : u);
  return 0.0f;
}

// The same weights for stencil offset O of the particle's base node, where u
// lies in [0.5, 1.5) (O = 0), [-0.5, 0.5) (O = 1) or [-1.5, -0.5) (O = 2) up to
// the node coordinates' rounding: the branch of spline_N / spline_dN is known,
// so these are its unfused float forms with the same operations (bit-identical
// inside the interval).  Offsets 0 and 2 keep the zero beyond |u| = 1.5; a |u|
// a rounding past 0.5 at offset 1 changes the value by O(u - 0.5)^2, below
// float resolution.
template <int O>
MPMX_HD float spline_N_at(float u) {
#pragma clang fp contract(off)
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (O == 1) return __fsub_rn(0.75f, __fmul_rn(u, u));
  const float t = O == 0 ? __fsub_rn(1.5f, u) : __fadd_rn(1.5f, u);
  return t > 0.0f ? __fmul_rn(0.5f, __fmul_rn(t, t)) : 0.0f;
#else
  return spline_N(u);
#endif
}
template <int O>
MPMX_HD float spline_dN_at(float u) {
#pragma clang fp contract(off)
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (O == 1) return -2.0f * u;
  const float t = O == 0 ? u - 1.5f : u + 1.5f;
  return (O == 0 ? t < 0.0f : t > 0.0f) ? t : 0.0f;
#else
  return spline_dN(u);
#endif
}

// Per-axis weights of the 3-node stencil starting at node b.
struct AxisW { float N[3]; float dN[3]; };

MPMX_HD void axis_weights(float xp, const float* coords, int b, float inv_dx, AxisW& a) {
  float u[3];
#pragma unroll
  for (int o = 0; o < 3; ++o) u[o] = spline_u(xp, coords[b + o], inv_dx);
  a.N[0] = spline_N_at<0>(u[0]); a.dN[0] = spline_dN_at<0>(u[0]);
  a.N[1] = spline_N_at<1>(u[1]); a.dN[1] = spline_dN_at<1>(u[1]);
  a.N[2] = spline_N_at<2>(u[2]); a.dN[2] = spline_dN_at<2>(u[2]);
}

// stencil base node (the particle "bin"): floor((x - origin)/dx - 0.5)
MPMX_HD int stencil_base(float xp, float origin, float inv_dx) {
#pragma clang fp contract(off)
#ifdef __HIP_DEVICE_COMPILE__
  float f = __fmul_rn(__fsub_rn(xp, origin), inv_dx);
  return (int)floorf(__fsub_rn(f, 0.5f));
#else
  float f = (xp - origin) * inv_dx;
  return (int)floorf(f - 0.5f);
#endif
}

// ----------------------------------------------------------------- knife SDF
MPMX_HD float sdf_box(float px, float py, float pz, const float* h) {
  float q0 = fabsf(px) - h[0], q1 = fabsf(py) - h[1], q2 = fabsf(pz) - h[2];
  float m0 = fmaxf(q0, 0.0f), m1 = fmaxf(q1, 0.0f), m2 = fmaxf(q2, 0.0f);
  return sqrtf(m0 * m0 + m1 * m1 + m2 * m2) + fminf(fmaxf(q0, fmaxf(q1, q2)), 0.0f);
}

// object_spatula.hpp:40-82
MPMX_HD float sdf_spatula(float px_in, float py, float pz, float hw, float ht, float hl, float off) {
  float px = fabsf(px_in);
  float b1 = hw + off, b2 = (hw + off) * 0.25f, he = hl + off;
  float bax = b2 - b1, bay = 2.0f * he;
  float pax = px - b1, pay = pz + he;
  float h = fmaxf(0.0f, fminf(1.0f, (pax * bax + pay * bay) / (bax * bax + bay * bay)));
  float dxs = pax - bax * h, dys = pay - bay * h;
  float d_side = sqrtf(dxs * dxs + dys * dys);
  float inside_x = b1 - (b1 - b2) * (pz + he) / (2.0f * he);
  float d2 = (px < inside_x) ? (px - inside_x) : d_side;
  if (pz > he) {
    d2 = sqrtf(px * px + (pz - he) * (pz - he)) - b2;
  } else if (pz < -he) {
    float rx = b1, rz = 1.5f * b1;
    float prz = pz + he;
    float k0 = sqrtf((px / rx) * (px / rx) + (prz / rz) * (prz / rz));
    float k1 = sqrtf((px / (rx * rx)) * (px / (rx * rx)) + (prz / (rz * rz)) * (prz / (rz * rz)));
    d2 = k0 * (k0 - 1.0f) / k1;
  }
  float dy = fabsf(py) - ht;
  float ex = fmaxf(d2, 0.0f), ey = fmaxf(dy, 0.0f);
  return fminf(fmaxf(d2, dy), 0.0f) + sqrtf(ex * ex + ey * ey);
}

MPMX_HD float knife_sd_local(const KnifeShapeDev& k, float px, float py, float pz) {
  if (k.shape == 1) return sdf_spatula(px, py, pz, k.half[0], k.half[1], k.half[2], k.offset);
  return sdf_box(px, py, pz, k.half);
}

MPMX_HD void knife_to_local(const KnifeStateDev& s, const float X[3], float p[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
    p[i] = ((s.Tinv[i * 4 + 0] * X[0] + s.Tinv[i * 4 + 1] * X[1]) + s.Tinv[i * 4 + 2] * X[2]) + s.Tinv[i * 4 + 3];
}

// ObjectSpatula::normal: central differences with e = 1e-4 in local space,
// normalised, mapped by (T.linear^-1)^T and normalised (object_spatula.hpp:99-129).
MPMX_HD void knife_normal(const KnifeShapeDev& k, const KnifeStateDev& s, const float X[3], float n[3]) {
  float p[3];
  knife_to_local(s, X, p);
  const float e = 0.0001f;
  float nl[3];
  nl[0] = knife_sd_local(k, p[0] + e, p[1], p[2]) - knife_sd_local(k, p[0] - e, p[1], p[2]);
  nl[1] = knife_sd_local(k, p[0], p[1] + e, p[2]) - knife_sd_local(k, p[0], p[1] - e, p[2]);
  nl[2] = knife_sd_local(k, p[0], p[1], p[2] + e) - knife_sd_local(k, p[0], p[1], p[2] - e);
  float nn = sqrtf(nl[0] * nl[0] + nl[1] * nl[1] + nl[2] * nl[2]);
  if (nn > 0.0f) { nl[0] /= nn; nl[1] /= nn; nl[2] /= nn; }
  float w[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) w[i] = s.N[i * 3 + 0] * nl[0] + s.N[i * 3 + 1] * nl[1] + s.N[i * 3 + 2] * nl[2];
  float wn = sqrtf(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  if (wn > 0.0f) { w[0] /= wn; w[1] /= wn; w[2] /= wn; }
  n[0] = w[0]; n[1] = w[1]; n[2] = w[2];
}

// ObjectSpatula::velocity: v + omega x (X - centre) (object_spatula.hpp:136-162)
MPMX_HD void knife_velocity(const KnifeStateDev& s, const float X[3], float v[3]) {
  float r0 = X[0] - s.T[3], r1 = X[1] - s.T[7], r2 = X[2] - s.T[11];
  v[0] = s.v[0] + (s.w[1] * r2 - s.w[2] * r1);
  v[1] = s.v[1] + (s.w[2] * r0 - s.w[0] * r2);
  v[2] = s.v[2] + (s.w[0] * r1 - s.w[1] * r0);
}

// Simulation::boundaryCollision (boundary_collision.cpp:5-316), 3-D, no MIBF.
// Xnode: node position; vi: post-force velocity (in/out).  Returns error bits.
MPMX_HD uint32_t boundary_collision(const SceneDev& sc, const KnifeStateDev& ks, const float Xnode[3],
                                    float vi[3]) {
  float vo[3] = {vi[0], vi[1], vi[2]};
  float X[3] = {Xnode[0] + sc.dt * vi[0], Xnode[1] + sc.dt * vi[1], Xnode[2] + sc.dt * vi[2]};
  bool influenced = false;
  if (sc.knife.present) {
    float p[3];
    knife_to_local(ks, X, p);
    float sd = knife_sd_local(sc.knife, p[0], p[1], p[2]);
    if (sd < 2.0f * sc.dx) influenced = true;
    if (sd < 0.0f) {
      float vob[3];
      knife_velocity(ks, X, vob);
      float vr[3] = {vo[0] - vob[0], vo[1] - vob[1], vo[2] - vob[2]};
      int bc = sc.knife.bc;
      if (bc == 0) {
        vr[0] = vr[1] = vr[2] = 0.0f;
      } else if (bc == 2 || bc == 1) {
        float n[3];
        knife_normal(sc.knife, ks, X, n);
        float d = vr[0] * n[0] + vr[1] * n[1] + vr[2] * n[2];
        float fr = sc.knife.friction;
        if (bc == 2) {  // SlipFree
          if (d < 0.0f) {
            float vt[3] = {vr[0] - d * n[0], vr[1] - d * n[1], vr[2] - d * n[2]};
            if (fr > 0.0f) {
              float vtn = sqrtf(vt[0] * vt[0] + vt[1] * vt[1] + vt[2] * vt[2]);
              if (-d * fr < vtn) {
                float sc2 = d * fr / vtn;
                vr[0] = vt[0] + vt[0] * sc2; vr[1] = vt[1] + vt[1] * sc2; vr[2] = vt[2] + vt[2] * sc2;
              } else {
                vr[0] = vr[1] = vr[2] = 0.0f;
              }
            } else {
              vr[0] = vt[0]; vr[1] = vt[1]; vr[2] = vt[2];
            }
          }
        } else {  // SlipStick
          vr[0] -= d * n[0]; vr[1] -= d * n[1]; vr[2] -= d * n[2];
          if (d < 0.0f && fr > 0.0f) {
            float vtn = sqrtf(vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2]);
            if (-d * fr < vtn) {
              float sc2 = d * fr / vtn;
              vr[0] += vr[0] * sc2; vr[1] += vr[1] * sc2; vr[2] += vr[2] * sc2;
            } else {
              vr[0] = vr[1] = vr[2] = 0.0f;
            }
          }
        }
      } else {
        return ERR_BAD_BC;
      }
      vi[0] = vr[0] + vob[0]; vi[1] = vr[1] + vob[1]; vi[2] = vr[2] + vob[2];
      if (X[1] < 0.8f * sc.dx) {  // thin-film rule (boundary_collision.cpp:93-111)
        float n[3];
        knife_normal(sc.knife, ks, X, n);
        float t[3] = {-n[1] * n[0], 1.0f - n[1] * n[1], -n[1] * n[2]};
        float tn = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (tn > 1e-6f) { t[0] /= tn; t[1] /= tn; t[2] /= tn; } else { t[0] = t[1] = t[2] = 0.0f; }
        float s2 = 0.2f * sqrtf(vob[0] * vob[0] + vob[1] * vob[1] + vob[2] * vob[2]);
        vi[0] = vob[0] + t[0] * s2; vi[1] = vob[1] + t[1] * s2; vi[2] = vob[2] + t[2] * s2;
      }
      vo[0] = vi[0]; vo[1] = vi[1]; vo[2] = vi[2];
    }
  }
  for (int pi = 0; pi < sc.n_plates; ++pi) {
    const PlateDev& pl = sc.plates[pi];
    const float pos = ks.plate[pi][0], lower = ks.plate[pi][1], upper = ks.plate[pi][2];
    const float pv[3] = {ks.plate[pi][3], ks.plate[pi][4], ks.plate[pi][5]};
    bool inside;
    switch (pl.plate_type) {
      case 2: inside = X[1] < upper && X[1] > lower && (X[0] - pos) <= 0.0f; break;
      case 3: inside = X[1] < upper && X[1] > lower && (pos - X[0]) <= 0.0f; break;
      case 1: inside = X[0] < upper && X[0] > lower && (X[1] - pos) <= 0.0f; break;
      case 0: inside = X[0] < upper && X[0] > lower && (pos - X[1]) <= 0.0f; break;
      case 5: inside = X[2] - pos <= 0.0f; break;
      default: inside = pos - X[2] <= 0.0f; break;
    }
    if (!inside) continue;
    if (influenced && pl.plate_type == 1) {  // knife priority over the floor
      if (vi[1] < pv[1]) { vi[1] = pv[1]; vo[0] = vi[0]; vo[1] = vi[1]; vo[2] = vi[2]; }
      continue;
    }
    float r[3] = {vo[0] - pv[0], vo[1] - pv[1], vo[2] - pv[2]};
    int nax = (pl.plate_type == 0 || pl.plate_type == 1) ? 1 : ((pl.plate_type == 2 || pl.plate_type == 3) ? 0 : 2);
    int ta = (nax == 0) ? 1 : 0, tb = (nax == 2) ? 1 : 2;
    if (pl.bc == 0) {
      r[0] = r[1] = r[2] = 0.0f;
    } else if (pl.bc == 1 || pl.bc == 2) {
      bool apply = true;
      if (pl.bc == 2)
        apply = (pl.plate_type == 0 && r[1] > 0.0f) || (pl.plate_type == 1 && r[1] < 0.0f) ||
                (pl.plate_type == 2 && r[0] < 0.0f) || (pl.plate_type == 3 && r[0] > 0.0f) ||
                (pl.plate_type == 4 && r[2] > 0.0f) || (pl.plate_type == 5 && r[2] < 0.0f);
      if (apply) {
        float fr = pl.friction;
        float vel_t = sqrtf(r[ta] * r[ta] + r[tb] * r[tb]);
        float fvn = fr * fabsf(r[nax]);
        if (fr == 0.0f) {
        } else if (vel_t <= fvn) {
          r[ta] = 0.0f; r[tb] = 0.0f;
        } else {
          float s2 = (vel_t - fvn) / vel_t;
          r[ta] *= s2; r[tb] *= s2;
        }
        r[nax] = 0.0f;
      }
    } else {
      return ERR_BAD_BC;
    }
    vi[0] = r[0] + pv[0]; vi[1] = r[1] + pv[1]; vi[2] = r[2] + pv[2];
    vo[0] = vi[0]; vo[1] = vi[1]; vo[2] = vi[2];
  }
  return 0u;
}

// Hardware reciprocal, square root and reciprocal square root on the device
// (v_rcp_f32, v_sqrt_f32, v_rsq_f32, about 1 ulp); IEEE forms on the host.
MPMX_HD float rcp_approx(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
MPMX_HD float sqrt_approx(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
// a / b as a times the hardware reciprocal on the device (within ~2 ulp of the
// IEEE quotient, a fraction of the instructions), IEEE on the host.
MPMX_HD float fdiv(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
  return a * __builtin_amdgcn_rcpf(b);
#else
  return a / b;
#endif
}
// log1p and expm1 of the principal strains.  The device forms replace the
// libm routines (compensated arithmetic, ~110 and ~60 instructions) by short
// ones that keep the relative precision of small arguments:
//   log1p(x) = log(u) x / (u - 1) with u = fl(1 + x) (the quotient corrects the
//     rounding of u; D. Goldberg, "What every computer scientist should know
//     about floating-point arithmetic", 1991, theorem 4), log from the
//     hardware base-2 logarithm; x itself where u rounds to 1;
//   expm1(x) = its Taylor polynomial to x^10 for |x| < 1/2 (remainder below
//     2^-11 / 11! of x), e^x - 1 from the hardware base-2 exponential beyond.
// Measured on gfx950 against float64 over 2M arguments in +-[1e-9, 3]: at most
// 4.7 (log1p) and 5.5 (expm1) units of 2^-24 relative.
MPMX_HD float log1p_fast(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  const float u = 1.0f + x;
  const float d = u - 1.0f;
  const float lg = __builtin_amdgcn_logf(u) * 0.693147180559945309f;
  return d == 0.0f ? x : lg * (x * __builtin_amdgcn_rcpf(d));
#else
  return log1pf(x);
#endif
}
MPMX_HD float expm1_fast(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  float p = 1.0f / 3628800.0f;
  p = fmaf(p, x, 1.0f / 362880.0f);
  p = fmaf(p, x, 1.0f / 40320.0f);
  p = fmaf(p, x, 1.0f / 5040.0f);
  p = fmaf(p, x, 1.0f / 720.0f);
  p = fmaf(p, x, 1.0f / 120.0f);
  p = fmaf(p, x, 1.0f / 24.0f);
  p = fmaf(p, x, 1.0f / 6.0f);
  p = fmaf(p, x, 0.5f);
  const float small = fmaf(p * x, x, x);
  return fabsf(x) < 0.5f ? small : __builtin_amdgcn_exp2f(x * 1.44269504088896341f) - 1.0f;
#else
  return expm1f(x);
#endif
}
MPMX_HD float rsq_approx(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

// ------------------------------------------------------ principal frame of F
// The return map and the stress need F = U diag(sigma) V^T and the principal
// Hencky strains log sigma (plasticity.cpp:18-20, simulation.hpp:288-290).
// Near F = I a float sigma carries ~1 ulp of 1 (6e-8) of absolute error, which
// is 1e-5..1e-3 of the small strains it encodes, so two correct FP32 SVDs
// disagree there by that much.  Instead the strains come from the symmetric
// eigenproblem of C = F^T F - I, formed in double so that its entries keep the
// strain's relative precision: C = V diag(lam) V^T, sigma_i^2 = 1 + lam_i,
// log sigma_i = log1p(lam_i) / 2, U = F V diag(1 / sigma).
// Cyclic Jacobi, three sweeps over the pairs (0,1), (0,2), (1,2) (fixed count:
// branch-uniform across a wavefront; on 2M random F = R (I + E), |E| from 1e-6
// to 0.5, the eigenvalues after three sweeps are within 5e-10 of max |lam| of
// their converged values); each rotation zeroes C_pq with
// t = tan(angle), the smaller root of t^2 + 2 theta t - 1 = 0,
// theta = (C_qq - C_pp) / (2 C_pq), and updates the diagonal as
// C_pp - t C_pq, C_qq + t C_pq.  The eigenvalues depend on t alone; the
// cosine 1 / sqrt(1 + t^2) comes from the hardware reciprocal square root
// (1 ulp), so each rotation is orthogonal to ~1e-7 and V's columns stay unit
// to ~1e-6 after the nine rotations (V enters F only through the small
// plastic correction, U is normalised on its own).  lam sorted descending (as
// Eigen sorts sigma).
MPMX_HD void jacobi_rot(float& app, float& aqq, float& apq, float& arp, float& arq, float* V, int p, int q) {
  float t = 0.0f;
  if (fabsf(apq) > 1e-30f) {   // (a smaller C_pq is already zero at this scale)
    const float theta = (aqq - app) * rcp_approx(2.0f * apq);
    t = copysignf(rcp_approx(fabsf(theta) + sqrt_approx(fmaf(theta, theta, 1.0f))), theta);
  }
  const float c = rsq_approx(fmaf(t, t, 1.0f));
  const float sn = t * c;
  app = fmaf(-t, apq, app);
  aqq = fmaf(t, apq, aqq);
  apq = 0.0f;
  const float rp = arp, rq = arq;
  arp = c * rp - sn * rq;
  arq = sn * rp + c * rq;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float vp = V[3 * r + p], vq = V[3 * r + q];
    V[3 * r + p] = c * vp - sn * vq;
    V[3 * r + q] = sn * vp + c * vq;
  }
}

// U_i = F v_i / |F v_i| (= F v_i / sigma_i; the norm keeps U_i unit also where
// 1 + lam_i has lost relative precision; zero for a null column).
MPMX_HD void principal_u(const float F[9], const float V[9], float U[9]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float w[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) w[r] = F[3 * r] * V[c] + F[3 * r + 1] * V[3 + c] + F[3 * r + 2] * V[6 + c];
    const float n2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const float inv = n2 > 1e-30f ? rsq_approx(n2) : 0.0f;
#pragma unroll
    for (int r = 0; r < 3; ++r) U[3 * r + c] = w[r] * inv;
  }
}

MPMX_HD void principal3(const float F[9], float U[9], float lam[3], float V[9]) {
  // C = F^T F - I in double: the float products are exact, the sums and the
  // - 1 round at 2^-53, so C keeps the strain's relative precision (also when
  // F carries a rotation) until it is rounded to float
#define MPMX_FF(i, j)                                                             \
  fma((double)F[i], (double)F[j], fma((double)F[3 + (i)], (double)F[3 + (j)],      \
                                      (double)F[6 + (i)] * (double)F[6 + (j)]))
  float c00 = (float)(MPMX_FF(0, 0) - 1.0);
  float c11 = (float)(MPMX_FF(1, 1) - 1.0);
  float c22 = (float)(MPMX_FF(2, 2) - 1.0);
  float c01 = (float)MPMX_FF(0, 1);
  float c02 = (float)MPMX_FF(0, 2);
  float c12 = (float)MPMX_FF(1, 2);
#undef MPMX_FF
#pragma unroll
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0f : 0.0f;
#pragma unroll
  for (int sweep = 0; sweep < 3; ++sweep) {
    jacobi_rot(c00, c11, c01, c02, c12, V, 0, 1);   // r = 2
    jacobi_rot(c00, c22, c02, c01, c12, V, 0, 2);   // r = 1
    jacobi_rot(c11, c22, c12, c01, c02, V, 1, 2);   // r = 0
  }
  float l[3] = {c00, c11, c22};
#define MPMX_SWAPEIG(i, j)                                                     \
  if (l[i] < l[j]) {                                                           \
    float tt = l[i]; l[i] = l[j]; l[j] = tt;                                   \
    for (int r = 0; r < 3; ++r) {                                              \
      float t2 = V[3 * r + i]; V[3 * r + i] = V[3 * r + j]; V[3 * r + j] = t2; \
    }                                                                          \
  }
  MPMX_SWAPEIG(0, 1)
  MPMX_SWAPEIG(0, 2)
  MPMX_SWAPEIG(1, 2)
#undef MPMX_SWAPEIG
#pragma unroll
  for (int c = 0; c < 3; ++c) lam[c] = l[c];
  principal_u(F, V, U);
}

// Principal Hencky strain of the return map, log(max(sigma, 1e-4))
// (plasticity.cpp:20; sigma < 1e-4 is 1 + lam < 1e-8).
MPMX_HD float hencky_of(float lam) { return 1.0f + lam >= 1e-8f ? 0.5f * log1p_fast(lam) : -9.21034037f; }
// log |sigma| of the stress (simulation.hpp:290, no clamp).
MPMX_HD float log_sigma_of(float lam) { return 0.5f * log1p_fast(lam); }

// F = U diag(e^h_new) V^T after the return map (plasticity.cpp:233).  While no
// singular value was clamped (e^h_old = sigma), this is
// Ft V diag(e^(h_new - h_old)) V^T, evaluated as Ft + Ft V diag(expm1(h_new -
// h_old)) V^T so that the (small) correction keeps its relative precision.
MPMX_HD void f_from_principal(const float Ft[9], const float U[9], const float V[9], const float lam[3],
                              const float h_old[3], const float h_new[3], float Fn[9]) {
  const bool clamped = !(1.0f + lam[0] >= 1e-8f) || !(1.0f + lam[1] >= 1e-8f) || !(1.0f + lam[2] >= 1e-8f);
  if (clamped) {
    const float e0 = expf(h_new[0]), e1 = expf(h_new[1]), e2 = expf(h_new[2]);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        Fn[3 * r + c] = (U[3 * r + 0] * e0) * V[3 * c + 0] + (U[3 * r + 1] * e1) * V[3 * c + 1] +
                        (U[3 * r + 2] * e2) * V[3 * c + 2];
    return;
  }
  const float e[3] = {expm1_fast(h_new[0] - h_old[0]), expm1_fast(h_new[1] - h_old[1]),
                      expm1_fast(h_new[2] - h_old[2])};
  float W[6];  // V diag(e) V^T, symmetric: 00 11 22 01 02 12
  W[0] = V[0] * e[0] * V[0] + V[1] * e[1] * V[1] + V[2] * e[2] * V[2];
  W[1] = V[3] * e[0] * V[3] + V[4] * e[1] * V[4] + V[5] * e[2] * V[5];
  W[2] = V[6] * e[0] * V[6] + V[7] * e[1] * V[7] + V[8] * e[2] * V[8];
  W[3] = V[0] * e[0] * V[3] + V[1] * e[1] * V[4] + V[2] * e[2] * V[5];
  W[4] = V[0] * e[0] * V[6] + V[1] * e[1] * V[7] + V[2] * e[2] * V[8];
  W[5] = V[3] * e[0] * V[6] + V[4] * e[1] * V[7] + V[5] * e[2] * V[8];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const float a = Ft[3 * r], b = Ft[3 * r + 1], c = Ft[3 * r + 2];
    Fn[3 * r + 0] = a + (a * W[0] + b * W[3] + c * W[4]);
    Fn[3 * r + 1] = b + (a * W[3] + b * W[1] + c * W[5]);
    Fn[3 * r + 2] = c + (a * W[4] + b * W[5] + c * W[2]);
  }
}

// ------------------------------------------------------------ return map
// VMVisc (plasticity.cpp:163-238) on the principal Hencky strains.
// In: h[3] = log(max(sigma,1e-4)), eps (eps_pl_dev).  Out: h (corrected), eps,
// dgamma_rate (= delta_gamma/dt, plasticity.cpp:235).  Returns 1 if projected,
// sets err bits on the reference's fatal conditions.
// Perzyna consistency residual r = q (dt / (visc dgam + dt))^n - q_y(eps + dgam/d)
// at equivalent stress q (plasticity.cpp:196-229), and its derivative along the
// flow rule q = q_tr - f dgam.
// x^n of the Perzyna rate factor (plasticity.cpp:187-229 std::pow(tmp, n)):
// exact for the exponents 0, 1 and 2 (x^2 rounded once, as a correctly
// rounded pow returns it), powf otherwise.
MPMX_HD float perzyna_pow(float x, float n) {
  if (n == 0.0f) return 1.0f;
  if (n == 1.0f) return x;
  if (n == 2.0f) return x * x;
  return powf(x, n);
}

MPMX_HD float perzyna_residual(const SceneDev& sc, float q, float dgam, float eps_in) {
  const float tm = sc.perzyna_visc * dgam + sc.dt;
  const float tmp = fdiv(sc.dt, tm);
  const float tmp1 = perzyna_pow(tmp, sc.perzyna_exp);
  const float qy = sc.q_min + (sc.q_max - sc.q_min) * expf(-sc.xi * (eps_in + sc.inv_d_prefac * dgam));
  return q * tmp1 - qy;
}
MPMX_HD float perzyna_residual_slope(const SceneDev& sc, float q_trial, float dgam, float eps_in) {
  const float tm = sc.perzyna_visc * dgam + sc.dt;
  const float tmp = fdiv(sc.dt, tm);
  const float tmp1 = perzyna_pow(tmp, sc.perzyna_exp);
  const float ex = expf(-sc.xi * (eps_in + sc.inv_d_prefac * dgam));
  const float qyd = -sc.xi * sc.inv_d_prefac * (sc.q_max - sc.q_min) * ex;
  return -sc.f_mu_prefac * tmp1 +
         (q_trial - sc.f_mu_prefac * dgam) * sc.perzyna_exp *
             fdiv(perzyna_pow(tmp, sc.perzyna_exp - 1.0f) * (-sc.perzyna_visc * sc.dt), tm * tm) -
         qyd;
}

// Perzyna overstress increment dgam for trial stress q_trial against yield
// q_yield (plasticity.cpp:187-229, 276-320): the closed form when `closed`, else
// Newton on r(dgam) = (q_tr - f dgam)(dt/(visc dgam + dt))^n - q_y(dgam), with
// VMVisc's softening q_y(eps + dgam/d) when `harden` (DPVisc's q_y is fixed).
MPMX_HD float perzyna_dgamma(const SceneDev& sc, float q_trial, float q_yield, bool closed, bool harden, float eps,
                             uint32_t& err, int* newton_iters) {
  float dgam;
  if (closed) {
    dgam = fdiv(q_trial - q_yield, sc.f_mu_prefac + fdiv(q_yield * sc.perzyna_visc, sc.dt));
    if (dgam < 0.0f) err |= ERR_NEG_DGAMMA;
    return dgam;
  }
  dgam = fdiv(0.01f * (q_trial - q_yield), sc.f_mu_prefac);
  const int max_iter = 60;
  for (int it = 0; it < max_iter; ++it) {
    if (newton_iters) ++*newton_iters;
    if (it == max_iter - 1) err |= ERR_NEWTON;
    float tm = sc.perzyna_visc * dgam + sc.dt;
    float tmp = fdiv(sc.dt, tm);
    float tmp1 = perzyna_pow(tmp, sc.perzyna_exp);
    float qyd = 0.0f;
    if (harden) {
      float ex = expf(-sc.xi * (eps + sc.inv_d_prefac * dgam));
      q_yield = sc.q_min + (sc.q_max - sc.q_min) * ex;
      qyd = -sc.xi * sc.inv_d_prefac * (sc.q_max - sc.q_min) * ex;
    }
    float res = (q_trial - sc.f_mu_prefac * dgam) * tmp1 - q_yield;
    if (fabsf(res) < 1e-2f) break;
    float rd = -sc.f_mu_prefac * tmp1 +
               (q_trial - sc.f_mu_prefac * dgam) * sc.perzyna_exp * perzyna_pow(tmp, sc.perzyna_exp - 1.0f) *
                   fdiv(-sc.perzyna_visc * sc.dt, tm * tm) - qyd;
    if (fabsf(rd) < 1e-14f) err |= ERR_NEWTON;
    dgam -= fdiv(res, rd);
    if (dgam < 0.0f) dgam = 1e-10f;
  }
  return dgam;
}

// Unit deviatoric direction of the principal Hencky strain h (plasticity.cpp:20-27).
MPMX_HD void hencky_deviator(const float h[3], float& htr, float dv[3], float& dn) {
  htr = (h[0] + h[1]) + h[2];
  float m = fdiv(htr, 3.0f);
  dv[0] = h[0] - m; dv[1] = h[1] - m; dv[2] = h[2] - m;
  dn = sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  if (dn > 0.0f) { float inv = rcp_approx(dn); dv[0] *= inv; dv[1] *= inv; dv[2] *= inv; }
}

// VMVisc (plasticity.cpp:163-238).  Returns 1 when projected (counted as the
// reference's plastic_count), 2 for the capped-pressure branch (F = U V^T,
// not counted, :172-180), 0 when elastic.  eps_vol (eps_pl_vol, may be null:
// taken as 0 and not updated) enters the cap p_min e^{-xi eps_vol}.
MPMX_HD int vmvisc_principal(const SceneDev& sc, float h[3], float& eps, float& dg_rate, uint32_t& err,
                             int* newton_iters = nullptr, float* eps_vol = nullptr) {
  float htr, dv[3], dn;
  hencky_deviator(h, htr, dv, dn);
  float q_trial = sc.e_mu_prefac * dn;
  float q_yield = sc.q_min + (sc.q_max - sc.q_min) * expf(-sc.xi * eps);
  float p_trial = -sc.K * htr;
  dg_rate = 0.0f;
  const float cap = eps_vol ? sc.p_min * expf(-sc.xi * *eps_vol) : sc.p_min;
  if (p_trial < cap) {
    float dgam = fdiv(q_trial, sc.f_mu_prefac);
    if (eps_vol) *eps_vol = *eps_vol + fdiv(-p_trial, sc.K);
    h[0] = h[1] = h[2] = 0.0f;
    eps = eps + sc.inv_d_prefac * dgam;
    dg_rate = fdiv(dgam, sc.dt);
    return 2;
  }
  if (!(q_trial > q_yield)) return 0;
  float dgam = perzyna_dgamma(sc, q_trial, q_yield, sc.perzyna_exp == 1.0f && sc.xi == 0.0f, true, eps, err,
                              newton_iters);
  float fac = sc.inv_d_prefac * dgam;
  h[0] -= fac * dv[0]; h[1] -= fac * dv[1]; h[2] -= fac * dv[2];
  eps = eps + fac;
  dg_rate = fdiv(dgam, sc.dt);
  return 1;
}

// DPVisc with Pradhana's correction (plasticity.cpp:240-340; use_mibf off as
// the paint app sets it).  The trial pressure is shifted by the volume gained
// at the tip, p_shift = -K eps_pl_vol_pradhana; left of the cohesion tip the
// strain projects to the tip (isotropic h), right of it the Perzyna
// overstress returns along the deviator and the accumulation resets.  Returns
// 1 when projected (F = U e^h V^T with the returned h), else 0.
MPMX_HD int dpvisc_principal(const SceneDev& sc, float h[3], float& eps, float& eps_vol, float& prad,
                             float& dg_rate, uint32_t& err, int* newton_iters = nullptr) {
  float htr, dv[3], dn;
  hencky_deviator(h, htr, dv, dn);
  const float p_trial = -sc.K * htr;
  const float q_trial = sc.e_mu_prefac * dn;
  const float p_tip = -sc.q_cohesion / sc.M;
  const float p_shift = sc.use_pradhana ? -sc.K * prad : 0.0f;
  const float ps = p_trial + p_shift;
  const float q_yield = sc.M * ps + sc.q_cohesion;
  dg_rate = 0.0f;
  if (ps < p_tip) {
    const float dgam = sc.d_prefac * dn;
    const float inst = (p_tip - p_trial) / sc.K;
    const float hv = -p_tip / (sc.K * 3.0f);
    h[0] = h[1] = h[2] = hv;
    eps = eps + sc.inv_d_prefac * dgam;
    eps_vol = eps_vol + inst;
    if (sc.use_pradhana) prad = prad + inst;
    dg_rate = fdiv(dgam, sc.dt);
    return 1;
  }
  if (sc.use_pradhana) prad = 0.0f;
  if (!(ps > p_tip && q_trial > q_yield)) return 0;
  float dgam = perzyna_dgamma(sc, q_trial, q_yield, sc.perzyna_exp == 1.0f, false, eps, err, newton_iters);
  float fac = sc.inv_d_prefac * dgam;
  h[0] -= fac * dv[0]; h[1] -= fac * dv[1]; h[2] -= fac * dv[2];
  eps = eps + fac;
  dg_rate = fdiv(dgam, sc.dt);
  return 1;
}

// Kirchhoff stress from the principal frame: tau = U diag(2 mu h + lambda tr h) U^T
// (simulation.hpp:287-294 with tau = P F^T, F = U diag(e^h) V^T); returns the
// symmetric tau as (xx, yy, zz, xy, xz, yz).
MPMX_HD void kirchhoff_from_principal(const SceneDev& sc, const float U[9], const float h[3], float tau[6]) {
  float tr = (h[0] + h[1]) + h[2];
  float d0 = 2.0f * sc.mu * h[0] + sc.lambda * tr;
  float d1 = 2.0f * sc.mu * h[1] + sc.lambda * tr;
  float d2 = 2.0f * sc.mu * h[2] + sc.lambda * tr;
  // tau_rc = sum_k U_rk d_k U_ck
  tau[0] = U[0] * U[0] * d0 + U[1] * U[1] * d1 + U[2] * U[2] * d2;
  tau[1] = U[3] * U[3] * d0 + U[4] * U[4] * d1 + U[5] * U[5] * d2;
  tau[2] = U[6] * U[6] * d0 + U[7] * U[7] * d1 + U[8] * U[8] * d2;
  tau[3] = U[0] * U[3] * d0 + U[1] * U[4] * d1 + U[2] * U[5] * d2;
  tau[4] = U[0] * U[6] * d0 + U[1] * U[7] * d1 + U[2] * U[8] * d2;
  tau[5] = U[3] * U[6] * d0 + U[4] * U[7] * d1 + U[5] * U[8] * d2;
}

}  // namespace mpmx
