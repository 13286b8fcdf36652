// oracle_math.hpp — small FP32 3-vector / 3x3 helpers and the principal frame of
// F (symmetric Jacobi eigen-decomposition of F^T F - I) for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  This file belongs to the oracle (a CPU restatement
// of the reference step used to check the HIP path); nothing in the product
// includes it.
//
// The reference computes every SVD with Eigen's JacobiSVD
// (matter/src/simulation/simulation.hpp:288, plasticity.cpp:18) and uses only
// isotropic functions of it: log of the singular values, U diag(.) U^T and
// U diag(.) V^T.  Those are computed here from the eigenproblem of
// F^T F - I (principal_of below); tests pin it against LAPACK (numpy) in
// float64 and against the Hencky closed forms.
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

// ---------------------------------------------------------------- principal frame of F
// The reference takes the principal Hencky strain as log of the singular values
// of F (plasticity.cpp:18-20, simulation.hpp:288-290).  Near F = I a float
// singular value carries ~1 ulp of 1 (6e-8) of absolute error, a relative error
// of 1e-5..1e-3 in the small strains it encodes.  This computes the same
// quantities from the symmetric eigenproblem of C = F^T F - I instead, with C
// formed in double from the float F, so that its entries keep the relative
// precision of the strain, also when F carries a rotation:
// C = V diag(lam) V^T,  sigma_i = sqrt(1 + lam_i),  log sigma_i = log1p(lam_i) / 2,
// U = F V diag(1 / sigma) (columns normalised).  For strongly compressed directions (sigma_i < 0.1)
// the precision of log sigma_i falls to ~6e-8 / sigma_i^2 (that of 1 + lam_i).
struct Principal {
  M3 U, V;
  V3 sig, lam;  // sigma_i, lam_i = sigma_i^2 - 1 (descending)
};

// Cyclic Jacobi on a symmetric 3x3 (upper triangle used), until the off-diagonal
// mass is below 1e-9 of the diagonal's (or zero); V accumulates the rotations.
inline void sym_eig3_jacobi(float C[3][3], M3& V) {
  V = eye33();
  for (int sweep = 0; sweep < 64; ++sweep) {
    float off = std::fabs(C[0][1]) + std::fabs(C[0][2]) + std::fabs(C[1][2]);
    float dia = std::fabs(C[0][0]) + std::fabs(C[1][1]) + std::fabs(C[2][2]);
    if (!(off > 1e-9f * dia) || off < 1e-35f) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        float apq = C[p][q];
        if (apq == 0.0f) continue;
        // rotation zeroing C[p][q]: t = tan(angle), the smaller root of
        // t^2 + 2 theta t - 1 = 0 with theta = (C_qq - C_pp) / (2 C_pq)
        float theta = (C[q][q] - C[p][p]) / (2.0f * apq);
        float t = (theta >= 0.0f ? 1.0f : -1.0f) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0f));
        float c = 1.0f / std::sqrt(t * t + 1.0f), sn = t * c;
        C[p][p] -= t * apq;
        C[q][q] += t * apq;
        C[p][q] = C[q][p] = 0.0f;
        for (int r = 0; r < 3; ++r) {
          if (r == p || r == q) continue;
          float arp = C[r][p], arq = C[r][q];
          C[r][p] = C[p][r] = c * arp - sn * arq;
          C[r][q] = C[q][r] = sn * arp + c * arq;
        }
        for (int r = 0; r < 3; ++r) {
          float vp = V(r, p), vq = V(r, q);
          V(r, p) = c * vp - sn * vq;
          V(r, q) = sn * vp + c * vq;
        }
      }
  }
}

inline Principal principal_of(const M3& F) {
  float C[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      // products of floats are exact in double; the sum and the - 1 round at
      // 2^-53, so C holds F^T F - I of the float F to ~1e-16 before rounding
      // to float (relative precision 6e-8 of C itself, i.e. of the strain)
      double d = ((double)F(0, i) * F(0, j) + (double)F(1, i) * F(1, j)) + (double)F(2, i) * F(2, j);
      C[i][j] = (float)(d - (i == j ? 1.0 : 0.0));
    }
  Principal P;
  sym_eig3_jacobi(C, P.V);
  for (int i = 0; i < 3; ++i) P.lam[i] = C[i][i];
  for (int i = 0; i < 3; ++i) {  // descending
    int m = i;
    for (int j = i + 1; j < 3; ++j)
      if (P.lam[j] > P.lam[m]) m = j;
    if (m != i) {
      std::swap(P.lam[i], P.lam[m]);
      for (int r = 0; r < 3; ++r) std::swap(P.V(r, i), P.V(r, m));
    }
  }
  for (int i = 0; i < 3; ++i) {
    P.sig[i] = std::sqrt(std::max(1.0f + P.lam[i], 0.0f));
    // U_i = F v_i / |F v_i| (= F v_i / sigma_i; the norm keeps U_i unit also
    // where 1 + lam_i has lost relative precision)
    float w[3];
    for (int r = 0; r < 3; ++r) w[r] = (F(r, 0) * P.V(0, i) + F(r, 1) * P.V(1, i)) + F(r, 2) * P.V(2, i);
    float nw = std::sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
    float inv = nw > 0.0f ? 1.0f / nw : 0.0f;
    for (int r = 0; r < 3; ++r) P.U(r, i) = w[r] * inv;
  }
  return P;
}

// log sigma_i (the stress's unclamped log |sigma|, simulation.hpp:290)
inline float log_sigma(float lam) { return 0.5f * std::log1p(lam); }

}  // namespace orc
