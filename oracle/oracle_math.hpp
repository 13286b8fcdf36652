// oracle_math.hpp — small FP32 3-vector / 3x3 helpers and a two-sided Jacobi
// SVD for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  This file belongs to the oracle (a CPU restatement
// of the reference step used to check the HIP path); nothing in the product
// includes it.
//
// The reference computes every SVD with Eigen's JacobiSVD
// (matter/src/simulation/simulation.hpp:288, plasticity.cpp:18).  Eigen is not
// available here, so this is a restatement of Eigen's published algorithm
// (two-sided Jacobi R-SVD: per (p,q) pair a 2x2 real Jacobi SVD — one rotation
// making the 2x2 block symmetric, one diagonalising it — sweeps until every
// off-diagonal is below max(tiny, 2*eps*max|diag|), then |diag| with the sign
// folded into U, sorted descending).  Eigen's version is unpinned (survey §8c):
// only isotropic functions of the SVD enter the step, so any accurate SVD gives
// the same results to rounding; tests pin this one against LAPACK (numpy.linalg.svd).
#pragma once
#include <cmath>
#include <algorithm>

namespace orc {

struct V3 {
  float a[3];
  float& operator[](int i) { return a[i]; }
  float operator[](int i) const { return a[i]; }
};
struct M3 {  // row-major a[r][c]
  float a[3][3];
  float& operator()(int r, int c) { return a[r][c]; }
  float operator()(int r, int c) const { return a[r][c]; }
};

inline V3 v3(float x, float y, float z) { V3 r; r[0] = x; r[1] = y; r[2] = z; return r; }
inline V3 zero3() { return v3(0, 0, 0); }
inline M3 zero33() { M3 m; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m(i, j) = 0; return m; }
inline M3 eye33() { M3 m = zero33(); m(0, 0) = m(1, 1) = m(2, 2) = 1; return m; }
inline V3 add(const V3& a, const V3& b) { return v3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline V3 sub(const V3& a, const V3& b) { return v3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline V3 scale(const V3& a, float s) { return v3(a[0] * s, a[1] * s, a[2] * s); }
inline float dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline float norm(const V3& a) { return std::sqrt(dot(a, a)); }
inline V3 cross(const V3& a, const V3& b) {
  return v3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline V3 matvec(const M3& m, const V3& x) {
  V3 r;
  for (int i = 0; i < 3; ++i) r[i] = m(i, 0) * x[0] + m(i, 1) * x[1] + m(i, 2) * x[2];
  return r;
}
inline M3 matmul(const M3& a, const M3& b) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
  return r;
}
inline M3 transpose(const M3& a) {
  M3 r;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r(i, j) = a(j, i);
  return r;
}

// 2x2 real Jacobi rotation (c, s) with J = [c s; -s c], applied as in Eigen's
// JacobiRotation conventions: "left" multiplies rows p,q, "right" columns p,q.
struct Rot { float c, s; };

inline Rot rot_transpose(Rot r) { return Rot{r.c, -r.s}; }
inline Rot rot_mul(Rot a, Rot b) {  // a * b
  return Rot{a.c * b.c - a.s * b.s, a.c * b.s + a.s * b.c};
}

// Rotations are G(c,s) = [c s; -s c] acting on the index pair (p,q).
// Symmetric Schur of the 2x2 symmetric block N = [x y; y z]: R = G(c,s) with
// R^T N R diagonal.  The off-diagonal of R^T N R is cs(x-z) + y(c^2-s^2); with
// t = s/c this is t^2 - 2 tau t - 1 = 0, tau = (x-z)/(2y); the smaller root is taken.
inline Rot make_jacobi(float x, float y, float z) {
  if (std::fabs(y) < 1e-37f) return Rot{1.0f, 0.0f};
  float tau = (x - z) / (2.0f * y);
  float w = std::sqrt(1.0f + tau * tau);
  float t = (tau >= 0.0f) ? -1.0f / (tau + w) : 1.0f / (w - tau);
  float c = 1.0f / std::sqrt(1.0f + t * t);
  return Rot{c, t * c};
}

// Apply on the left: rows p,q of M <- J^T-style rotation as Eigen applyOnTheLeft(p,q,J):
//   x = M(p,:), y = M(q,:):  M(p,:) = c*x + s*y ... with Eigen's convention
//   x' = c*x + conj(s)*y ; y' = -s*x + c*y   (for real s)
inline void apply_left(M3& m, int p, int q, Rot j) {
  for (int k = 0; k < 3; ++k) {
    float x = m(p, k), y = m(q, k);
    m(p, k) = j.c * x + j.s * y;
    m(q, k) = -j.s * x + j.c * y;
  }
}
// Apply on the right: columns p,q of M: x = M(:,p), y = M(:,q):
//   x' = c*x - s*y ; y' = s*x + c*y
inline void apply_right(M3& m, int p, int q, Rot j) {
  for (int k = 0; k < 3; ++k) {
    float x = m(k, p), y = m(k, q);
    m(k, p) = j.c * x - j.s * y;
    m(k, q) = j.s * x + j.c * y;
  }
}

// 2x2 real Jacobi SVD of the (p,q) block of a: returns left J1 and right J2 such that
// J1 * B * J2 is diagonal (B = a restricted to p,q).
inline void real_2x2_jacobi_svd(const M3& a, int p, int q, Rot& j_left, Rot& j_right) {
  float m00 = a(p, p), m01 = a(p, q), m10 = a(q, p), m11 = a(q, q);
  Rot rot1;
  float t = m00 + m11;
  float d = m10 - m01;
  if (std::fabs(d) < 1e-37f) {
    rot1 = Rot{1.0f, 0.0f};
  } else {
    float u = t / d;
    float tmp = std::sqrt(1.0f + u * u);
    rot1 = Rot{u / tmp, 1.0f / tmp};
  }
  // apply rot1 on the left of the 2x2 block
  float n00 = rot1.c * m00 + rot1.s * m10, n01 = rot1.c * m01 + rot1.s * m11;
  float n10 = -rot1.s * m00 + rot1.c * m10, n11 = -rot1.s * m01 + rot1.c * m11;
  (void)n10;
  // R^T (rot1 B) R = diag  =>  left = R^T rot1, right = R (2-D rotations commute).
  j_right = make_jacobi(n00, n01, n11);
  j_left = rot_mul(rot_transpose(j_right), rot1);
}

// A = U diag(s) V^T, s >= 0 sorted descending, U and V orthogonal (det may be -1).
inline void svd3_jacobi(const M3& A, M3& U, V3& s, M3& V) {
  M3 w = A;
  float scalev = 0.0f;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) scalev = std::max(scalev, std::fabs(w(i, j)));
  if (!(scalev > 0.0f) || !std::isfinite(scalev)) scalev = 1.0f;
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) w(i, j) /= scalev;
  U = eye33();
  V = eye33();
  const float consider_zero = 1.1754944e-38f * 2.0f;  // (min normal) * 2
  const float precision = 2.0f * 1.1920929e-7f;      // 2 * epsilon
  float max_diag = std::max(std::fabs(w(0, 0)), std::max(std::fabs(w(1, 1)), std::fabs(w(2, 2))));
  bool finished = false;
  int sweeps = 0;
  while (!finished && sweeps < 64) {
    finished = true;
    ++sweeps;
    for (int p = 1; p < 3; ++p) {
      for (int q = 0; q < p; ++q) {
        float threshold = std::max(consider_zero, precision * max_diag);
        if (std::fabs(w(p, q)) > threshold || std::fabs(w(q, p)) > threshold) {
          finished = false;
          Rot jl, jr;
          real_2x2_jacobi_svd(w, p, q, jl, jr);
          apply_left(w, p, q, jl);
          apply_right(U, p, q, rot_transpose(jl));
          apply_right(w, p, q, jr);
          apply_right(V, p, q, jr);
          max_diag = std::max(max_diag, std::max(std::fabs(w(p, p)), std::fabs(w(q, q))));
        }
      }
    }
  }
  for (int i = 0; i < 3; ++i) {
    float a = w(i, i);
    s[i] = std::fabs(a);
    if (a < 0.0f) for (int k = 0; k < 3; ++k) U(k, i) = -U(k, i);
  }
  for (int i = 0; i < 3; ++i) s[i] *= scalev;
  // sort descending (selection, swapping columns of U and V)
  for (int i = 0; i < 3; ++i) {
    int imax = i;
    for (int j = i + 1; j < 3; ++j) if (s[j] > s[imax]) imax = j;
    if (imax != i) {
      std::swap(s[i], s[imax]);
      for (int k = 0; k < 3; ++k) { std::swap(U(k, i), U(k, imax)); std::swap(V(k, i), V(k, imax)); }
    }
  }
}

}  // namespace orc
