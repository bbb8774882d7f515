// Device math for the FEM substep: 3-vectors, 3x3 helpers, the signed-SVD
// principal-stretch clamp, the scale-invariant block inverse and the Tet4 /
// Hex8 Stable Neo-Hookean element routines.  All float32.
#pragma once

#include <hip/hip_runtime.h>

#define FEM_DI __device__ __forceinline__

namespace fem {

struct v3 {
  float x, y, z;
};
FEM_DI v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
FEM_DI v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
FEM_DI v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
FEM_DI float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
FEM_DI v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
FEM_DI v3 ld3(const float* __restrict__ p, int i) {
  const float* q = p + 3 * (size_t)i;
  return {q[0], q[1], q[2]};
}
FEM_DI void st3(float* __restrict__ p, int i, v3 v) {
  float* q = p + 3 * (size_t)i;
  q[0] = v.x;
  q[1] = v.y;
  q[2] = v.z;
}
FEM_DI float det3(const float m[9]) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}
// y = M x for row-major M.
FEM_DI v3 mv(const float m[9], v3 x) {
  return {m[0] * x.x + m[1] * x.y + m[2] * x.z, m[3] * x.x + m[4] * x.y + m[5] * x.z,
          m[6] * x.x + m[7] * x.y + m[8] * x.z};
}

// ---------------------------------------------------------------------------
// Signed 3x3 SVD for the principal-stretch clamp.  Method of McAdams, Selle,
// Tamstorf, Teran & Sifakis (2011), "Computing the Singular Value
// Decomposition of 3x3 matrices with minimal branching and elementary floating
// point operations": cyclic Jacobi eigen-solve of S = F^T F gives V and the
// squared singular values; Gram-Schmidt (QR) of B = F V gives a rotation U and
// signed singular values with the sign of det F on the smallest one.
// ---------------------------------------------------------------------------
FEM_DI void jacobi_rotate(float s[3][3], float v[3][3], int p, int q) {
  float apq = s[p][q];
  if (apq == 0.f) return;
  float theta = (s[q][q] - s[p][p]) / (2.f * apq);
  float t = copysignf(1.f, theta) / (fabsf(theta) + sqrtf(theta * theta + 1.f));
  float c = rsqrtf(t * t + 1.f), sn = t * c;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float akp = s[k][p], akq = s[k][q];
    s[k][p] = c * akp - sn * akq;
    s[k][q] = sn * akp + c * akq;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float apk = s[p][k], aqk = s[q][k];
    s[p][k] = c * apk - sn * aqk;
    s[q][k] = sn * apk + c * aqk;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float vkp = v[k][p], vkq = v[k][q];
    v[k][p] = c * vkp - sn * vkq;
    v[k][q] = sn * vkp + c * vkq;
  }
}

__device__ __noinline__ void svd_clamp_slow(float F[9], float mn, float mx) {
  float s[3][3], v[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      s[i][j] = F[i] * F[j] + F[3 + i] * F[3 + j] + F[6 + i] * F[6 + j];
      v[i][j] = (i == j) ? 1.f : 0.f;
    }
  for (int sweep = 0; sweep < 8; ++sweep) {
    jacobi_rotate(s, v, 0, 1);
    jacobi_rotate(s, v, 0, 2);
    jacobi_rotate(s, v, 1, 2);
  }
  float w[3] = {s[0][0], s[1][1], s[2][2]};
  // Sort eigenpairs descending (columns of v).
#define FEM_SWAP_COL(a, b)                    \
  {                                           \
    float tw = w[a];                          \
    w[a] = w[b];                              \
    w[b] = tw;                                \
    for (int k = 0; k < 3; ++k) {             \
      float tv = v[k][a];                     \
      v[k][a] = v[k][b];                      \
      v[k][b] = tv;                           \
    }                                         \
  }
  if (w[0] < w[1]) FEM_SWAP_COL(0, 1);
  if (w[0] < w[2]) FEM_SWAP_COL(0, 2);
  if (w[1] < w[2]) FEM_SWAP_COL(1, 2);
#undef FEM_SWAP_COL
  float dv_ = v[0][0] * (v[1][1] * v[2][2] - v[1][2] * v[2][1]) - v[0][1] * (v[1][0] * v[2][2] - v[1][2] * v[2][0]) +
              v[0][2] * (v[1][0] * v[2][1] - v[1][1] * v[2][0]);
  if (dv_ < 0.f) {
    v[0][2] = -v[0][2];
    v[1][2] = -v[1][2];
    v[2][2] = -v[2][2];
  }
  v3 b[3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
    b[j] = {F[0] * v[0][j] + F[1] * v[1][j] + F[2] * v[2][j], F[3] * v[0][j] + F[4] * v[1][j] + F[5] * v[2][j],
            F[6] * v[0][j] + F[7] * v[1][j] + F[8] * v[2][j]};
  v3 u0, u1;
  float n0 = sqrtf(dot(b[0], b[0]));
  u0 = (n0 > 1e-30f) ? b[0] * (1.f / n0) : v3{1.f, 0.f, 0.f};
  v3 t = b[1] - u0 * dot(u0, b[1]);
  float n1 = sqrtf(dot(t, t));
  if (n1 > 1e-30f) {
    u1 = t * (1.f / n1);
  } else {
    float a0 = fabsf(u0.x), a1 = fabsf(u0.y), a2 = fabsf(u0.z);
    v3 e = (a0 <= a1 && a0 <= a2) ? v3{1.f, 0.f, 0.f} : (a1 <= a2 ? v3{0.f, 1.f, 0.f} : v3{0.f, 0.f, 1.f});
    v3 r = e - u0 * dot(u0, e);
    u1 = r * rsqrtf(dot(r, r));
  }
  v3 u2 = cross(u0, u1);
  float sg[3] = {dot(u0, b[0]), dot(u1, b[1]), dot(u2, b[2])};
#pragma unroll
  for (int k = 0; k < 3; ++k) sg[k] = fmaxf(fminf(sg[k], mx), mn);
  v3 u[3] = {u0, u1, u2};
  // F = U diag(sg) V^T
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float ui0 = (i == 0) ? u[0].x : (i == 1 ? u[0].y : u[0].z);
      float ui1 = (i == 0) ? u[1].x : (i == 1 ? u[1].y : u[1].z);
      float ui2 = (i == 0) ? u[2].x : (i == 1 ? u[2].y : u[2].z);
      F[3 * i + j] = ui0 * sg[0] * v[j][0] + ui1 * sg[1] * v[j][1] + ui2 * sg[2] * v[j][2];
    }
  }
}

// Principal-stretch clamp F <- U clamp(S) V^T (reference kernels.py:421-444),
// disabled when either bound is negative.  When det F > 0, ||F||_F^2 <= mx^2
// and 2 det F / ||F||_F^2 >= mn, every singular value already lies in
// [mn, mx] (sigma_max^2 <= ||F||^2 and sigma_min >= det / (sigma_1 sigma_2)
// >= 2 det / ||F||^2), so F is returned unchanged without an SVD.
FEM_DI void clamp_stretch(float F[9], float mn, float mx) {
  if (mn < 0.f || mx < 0.f) return;
  float det = det3(F);
  float ic = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) ic += F[k] * F[k];
  if (det > 0.f && ic <= mx * mx && 2.f * det >= mn * ic) return;
  svd_clamp_slow(F, mn, mx);
}

// Scale-invariant 3x3 inverse with point-Jacobi fallback
// (reference inverse_diagonal_block, kernels.py:1817-1833).
FEM_DI void inverse_diagonal_block(const float B[9], float out[9]) {
  float scale = fmaxf(fabsf(B[0]), fmaxf(fabsf(B[4]), fabsf(B[8])));
  if (scale > 0.f) {
    float n[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) n[i] = B[i] / scale;
    float trace = (n[0] + n[4] + n[8]) / 3.f;
    float det = det3(n);
    if (trace > 0.f && fabsf(det) > 1e-9f * trace * trace * trace) {
      float id = 1.f / det;
      out[0] = (n[4] * n[8] - n[5] * n[7]) * id / scale;
      out[1] = (n[2] * n[7] - n[1] * n[8]) * id / scale;
      out[2] = (n[1] * n[5] - n[2] * n[4]) * id / scale;
      out[3] = (n[5] * n[6] - n[3] * n[8]) * id / scale;
      out[4] = (n[0] * n[8] - n[2] * n[6]) * id / scale;
      out[5] = (n[2] * n[3] - n[0] * n[5]) * id / scale;
      out[6] = (n[3] * n[7] - n[4] * n[6]) * id / scale;
      out[7] = (n[1] * n[6] - n[0] * n[7]) * id / scale;
      out[8] = (n[0] * n[4] - n[1] * n[3]) * id / scale;
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i] = 0.f;
  out[0] = B[0] > 0.f ? 1.f / B[0] : 1.f;
  out[4] = B[4] > 0.f ? 1.f / B[4] : 1.f;
  out[8] = B[8] > 0.f ? 1.f / B[8] : 1.f;
}

// ---------------------------------------------------------------------------
// Tet4 Stable Neo-Hookean (reference eval_tetrahedra kernels.py:511-632 and
// build_system_matrix_tet_kernel :635-850).
// ---------------------------------------------------------------------------
struct TetGeo {
  v3 e1, e2, e3;  // x1-x0, x2-x0, x3-x0
  float Dm[9];
  float F[9];     // clamped deformation gradient
  float inv_rest_volume, rest_volume;
};

FEM_DI void tet_geo(const v3 x[4], const float* __restrict__ pose, float mn, float mx, TetGeo& g) {
  g.e1 = x[1] - x[0];
  g.e2 = x[2] - x[0];
  g.e3 = x[3] - x[0];
#pragma unroll
  for (int k = 0; k < 9; ++k) g.Dm[k] = pose[k];
  g.inv_rest_volume = det3(g.Dm) * 6.f;
  g.rest_volume = 1.f / g.inv_rest_volume;
  const float ex[3] = {g.e1.x, g.e1.y, g.e1.z}, ey[3] = {g.e2.x, g.e2.y, g.e2.z}, ez[3] = {g.e3.x, g.e3.y, g.e3.z};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) g.F[3 * r + c] = ex[r] * g.Dm[c] + ey[r] * g.Dm[3 + c] + ez[r] * g.Dm[6 + c];
  clamp_stretch(g.F, mn, mx);
}

// Force added to node `a` of the element (the reference subtracts f_a from
// the force buffer).  Includes Rayleigh damping.
FEM_DI v3 tet_node_force(const TetGeo& g, const v3 v[4], float mu, float lam, float kd, int a) {
  float alpha = 1.f + mu / lam - mu / (4.f * lam);
  float k_mu = mu * g.rest_volume, k_lambda = lam * g.rest_volume, k_damp = kd * g.rest_volume;
  v3 v10 = v[1] - v[0], v20 = v[2] - v[0], v30 = v[3] - v[0];
  float dF[9];
  const float vx[3] = {v10.x, v10.y, v10.z}, vy[3] = {v20.x, v20.y, v20.z}, vz[3] = {v30.x, v30.y, v30.z};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) dF[3 * r + c] = vx[r] * g.Dm[c] + vy[r] * g.Dm[3 + c] + vz[r] * g.Dm[6 + c];
  float ic = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) ic += g.F[k] * g.F[k];
  float sdev = k_mu * (1.f - 1.f / (ic + 1.f));
  float P[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) P[k] = g.F[k] * sdev + dF[k] * k_damp;
  // H = P Dm^T; f_{1..3} are its columns.
  v3 f[4];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v3 col = {P[0] * g.Dm[3 * c] + P[1] * g.Dm[3 * c + 1] + P[2] * g.Dm[3 * c + 2],
              P[3] * g.Dm[3 * c] + P[4] * g.Dm[3 * c + 1] + P[5] * g.Dm[3 * c + 2],
              P[6] * g.Dm[3 * c] + P[7] * g.Dm[3 * c + 1] + P[8] * g.Dm[3 * c + 2]};
    f[c + 1] = col;
  }
  float J = fmaxf(det3(g.F), 0.01f);
  float s = g.inv_rest_volume / 6.f;
  v3 d1 = cross(g.e2, g.e3) * s, d2 = cross(g.e3, g.e1) * s, d3 = cross(g.e1, g.e2) * s;
  float f_volume = (J - alpha) * k_lambda;
  float f_damp = (dot(d1, v10) + dot(d2, v20) + dot(d3, v30)) * k_damp;
  float ft = f_volume + f_damp;
  f[1] = f[1] + d1 * ft;
  f[2] = f[2] + d2 * ft;
  f[3] = f[3] + d3 * ft;
  f[0] = (f[1] + f[2] + f[3]) * -1.f;
  v3 fa = (a == 0) ? f[0] : (a == 1 ? f[1] : (a == 2 ? f[2] : f[3]));
  return fa * -1.f;
}

// Volumetric activation force on node a (eval_tet_activations kernels.py:484-508).
FEM_DI v3 tet_activation_force(const v3 x[4], float lam, float act, int a) {
  float scale = lam * act / 6.f;
  v3 fb = cross(x[2] - x[0], x[3] - x[0]) * scale;
  v3 fc = cross(x[3] - x[0], x[1] - x[0]) * scale;
  v3 fd = cross(x[1] - x[0], x[2] - x[0]) * scale;
  if (a == 0) return fb + fc + fd;
  if (a == 1) return fb * -1.f;
  if (a == 2) return fc * -1.f;
  return fd * -1.f;
}

// Row `a` of the 12x12 element tangent: blocks K_ab, b = 0..3, scaled by h^2.
FEM_DI void tet_tangent_row(const TetGeo& g, float mu, float lam, float h2, int a, float K[4][9]) {
  if (g.inv_rest_volume <= 0.f) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int k = 0; k < 9; ++k) K[b][k] = 0.f;
    return;
  }
  float k_mu = mu * g.rest_volume, k_lambda = lam * g.rest_volume;
  float ic = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) ic += g.F[k] * g.F[k];
  float ic1 = fmaxf(ic + 1.f, 1e-6f);
  float s = 1.f - 1.f / ic1;
  float coef = k_mu * 2.f / (ic1 * ic1);
  float sv = g.inv_rest_volume / 6.f;
  v3 dJ[4], gr[4], Fg[4];
  dJ[1] = cross(g.e2, g.e3) * sv;
  dJ[2] = cross(g.e3, g.e1) * sv;
  dJ[3] = cross(g.e1, g.e2) * sv;
  dJ[0] = (dJ[1] + dJ[2] + dJ[3]) * -1.f;
  gr[1] = {g.Dm[0], g.Dm[1], g.Dm[2]};
  gr[2] = {g.Dm[3], g.Dm[4], g.Dm[5]};
  gr[3] = {g.Dm[6], g.Dm[7], g.Dm[8]};
  gr[0] = (gr[1] + gr[2] + gr[3]) * -1.f;
#pragma unroll
  for (int b = 0; b < 4; ++b) Fg[b] = mv(g.F, gr[b]);
  v3 ga = (a == 0) ? gr[0] : (a == 1 ? gr[1] : (a == 2 ? gr[2] : gr[3]));
  v3 fa = (a == 0) ? Fg[0] : (a == 1 ? Fg[1] : (a == 2 ? Fg[2] : Fg[3]));
  v3 ja = (a == 0) ? dJ[0] : (a == 1 ? dJ[1] : (a == 2 ? dJ[2] : dJ[3]));
  const float fa3[3] = {fa.x, fa.y, fa.z}, ja3[3] = {ja.x, ja.y, ja.z};
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    float dab = dot(ga, gr[b]);
    const float fb3[3] = {Fg[b].x, Fg[b].y, Fg[b].z}, jb3[3] = {dJ[b].x, dJ[b].y, dJ[b].z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float kdev = ((r == c) ? k_mu * s * dab : 0.f) + fa3[r] * fb3[c] * coef;
        float kvol = ja3[r] * jb3[c] * k_lambda;
        K[b][3 * r + c] = (kdev + kvol) * h2;
      }
  }
}

// ---------------------------------------------------------------------------
// Hex8 Q1 element with 2x2x2 Gauss quadrature (reference kernels.py:1356-1747).
// Corner / Gauss-point signs: xi+ for {1,2,5,6}, eta+ for {2,3,6,7}, zeta+ for {4..7}.
// ---------------------------------------------------------------------------
FEM_DI float hex_sx(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1.f : -1.f; }
FEM_DI float hex_sy(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1.f : -1.f; }
FEM_DI float hex_sz(int a) { return (a >= 4) ? 1.f : -1.f; }

// Reference-configuration gradients dN[b] = inv_J0^T grad_xi N_b at Gauss point g.
FEM_DI void hex_grads(const float* __restrict__ invJ, int g, v3 dN[8]) {
  const float gp = 0.5773502691896258f;
  float xi = hex_sx(g) * gp, eta = hex_sy(g) * gp, zeta = hex_sz(g) * gp;
  float J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = invJ[k];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    float xa = hex_sx(b), ya = hex_sy(b), za = hex_sz(b);
    v3 d = {xa * (1.f + ya * eta) * (1.f + za * zeta) * 0.125f, (1.f + xa * xi) * ya * (1.f + za * zeta) * 0.125f,
            (1.f + xa * xi) * (1.f + ya * eta) * za * 0.125f};
    dN[b] = {J[0] * d.x + J[3] * d.y + J[6] * d.z, J[1] * d.x + J[4] * d.y + J[7] * d.z,
             J[2] * d.x + J[5] * d.y + J[8] * d.z};
  }
}

FEM_DI void hex_F(const v3 x[8], const v3 dN[8], float F[9]) {
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = 0.f;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const float xs[3] = {x[b].x, x[b].y, x[b].z}, ds[3] = {dN[b].x, dN[b].y, dN[b].z};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) F[3 * r + c] += xs[r] * ds[c];
  }
}

FEM_DI void cofactor(const float F[9], float cof[9]) {
  v3 c0 = {F[0], F[3], F[6]}, c1 = {F[1], F[4], F[7]}, c2 = {F[2], F[5], F[8]};
  v3 a = cross(c1, c2), b = cross(c2, c0), c = cross(c0, c1);
  cof[0] = a.x; cof[1] = b.x; cof[2] = c.x;
  cof[3] = a.y; cof[4] = b.y; cof[5] = c.y;
  cof[6] = a.z; cof[7] = b.z; cof[8] = c.z;
}

}  // namespace fem
