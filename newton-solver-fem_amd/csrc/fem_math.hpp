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
// Selection of one of several values by a runtime index, component by
// component on scalars so that it lowers to v_cndmask selects (a ?: over
// structs or array elements makes clang select pointers into a private,
// scratch-resident array).
FEM_DI float sel4f(int a, float a0, float a1, float a2, float a3) {
  float r = a3;
  r = (a == 2) ? a2 : r;
  r = (a == 1) ? a1 : r;
  r = (a == 0) ? a0 : r;
  return r;
}
FEM_DI v3 sel4(int a, v3 a0, v3 a1, v3 a2, v3 a3) {
  return {sel4f(a, a0.x, a1.x, a2.x, a3.x), sel4f(a, a0.y, a1.y, a2.y, a3.y), sel4f(a, a0.z, a1.z, a2.z, a3.z)};
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

struct M9 {
  float m[9];
};

// Out of line and by value: the rarely taken SVD path neither inflates the
// register budget of the element kernels nor forces F into scratch memory.
__device__ __noinline__ M9 svd_clamp_slow(M9 Fin, float mn, float mx) {
  float* F = Fin.m;
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
  return Fin;
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
  M9 in;
#pragma unroll
  for (int k = 0; k < 9; ++k) in.m[k] = F[k];
  const M9 out = svd_clamp_slow(in, mn, mx);
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = out.m[k];
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
// build_system_matrix_tet_kernel :635-850).  Only scalars, float arrays with
// compile-time indices and by-value vectors, so everything stays in VGPRs.
// ---------------------------------------------------------------------------
struct TetGeo {
  v3 e1, e2, e3;  // x1-x0, x2-x0, x3-x0
  float Dm[9];
  float F[9];     // clamped deformation gradient
  float inv_rest_volume, rest_volume;
};

FEM_DI void tet_geo(v3 x0, v3 x1, v3 x2, v3 x3, const float* __restrict__ pose, float mn, float mx, TetGeo& g) {
  g.e1 = x1 - x0;
  g.e2 = x2 - x0;
  g.e3 = x3 - x0;
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

// Row c of Dm as a vector.
FEM_DI v3 dm_row(const float Dm[9], int c) { return {Dm[3 * c], Dm[3 * c + 1], Dm[3 * c + 2]}; }

// Force added to node `a` of the element (the reference subtracts f_a from
// the force buffer).  Includes Rayleigh damping.
FEM_DI v3 tet_node_force(const TetGeo& g, v3 v0, v3 v1, v3 v2, v3 v3_, float mu, float lam, float kd, int a) {
  const float alpha = 1.f + mu / lam - mu / (4.f * lam);
  const float k_mu = mu * g.rest_volume, k_lambda = lam * g.rest_volume, k_damp = kd * g.rest_volume;
  const v3 v10 = v1 - v0, v20 = v2 - v0, v30 = v3_ - v0;
  const float vx[3] = {v10.x, v10.y, v10.z}, vy[3] = {v20.x, v20.y, v20.z}, vz[3] = {v30.x, v30.y, v30.z};
  float ic = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) ic += g.F[k] * g.F[k];
  const float sdev = k_mu * (1.f - 1.f / (ic + 1.f));
  float P[9];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float dF = vx[r] * g.Dm[c] + vy[r] * g.Dm[3 + c] + vz[r] * g.Dm[6 + c];
      P[3 * r + c] = g.F[3 * r + c] * sdev + dF * k_damp;
    }
  // H = P Dm^T; f_{1..3} are its columns: f_c = P * (row c of Dm).
  v3 f1 = mv(P, dm_row(g.Dm, 0)), f2 = mv(P, dm_row(g.Dm, 1)), f3 = mv(P, dm_row(g.Dm, 2));
  const float J = fmaxf(det3(g.F), 0.01f);
  const float s = g.inv_rest_volume / 6.f;
  const v3 d1 = cross(g.e2, g.e3) * s, d2 = cross(g.e3, g.e1) * s, d3 = cross(g.e1, g.e2) * s;
  const float f_volume = (J - alpha) * k_lambda;
  const float f_damp = (dot(d1, v10) + dot(d2, v20) + dot(d3, v30)) * k_damp;
  const float ft = f_volume + f_damp;
  f1 = f1 + d1 * ft;
  f2 = f2 + d2 * ft;
  f3 = f3 + d3 * ft;
  const v3 f0 = (f1 + f2 + f3) * -1.f;
  return sel4(a, f0, f1, f2, f3) * -1.f;
}

// Linear spring between particles p (first) and q (second): the force on the
// endpoint `end` (0: p receives -fs, 1: q receives +fs), fs = dir (ke c +
// kd dc/dt), dir = (x_p - x_q) / |x_p - x_q| (eval_springs kernels.py:176-219).
FEM_DI v3 spring_end_force(v3 xp, v3 xq, v3 vp, v3 vq, float rest, float ke, float kd, int end) {
  const v3 xpq = xp - xq, vpq = vp - vq;
  const float len = sqrtf(dot(xpq, xpq));
  if (len < 1.0e-6f) return {0.f, 0.f, 0.f};
  const v3 dir = xpq * (1.f / len);
  const float c = len - rest;
  const v3 fs = dir * (ke * c + kd * dot(dir, vpq));
  return end == 0 ? fs * -1.f : fs;
}

// Bend torque of a spring toward its rest direction on end `end` (0: p, 1: q)
// (eval_springs_linear_and_torque kernels.py:222-292, torque part): with
// d = (x_p - x_q)/len and d0 = -rest_dir (|d0| < 0.5: no torque), theta =
// acos(d.d0), axis = d0 x d / |d0 x d|, omega = (d x d_dot).axis with
// d_dot = (v_pq - d (d.v_pq))/len, tau = -(k_t theta + k_w omega) len,
// F = (axis tau) x d / len, applied as -F at p and +F at q.
FEM_DI v3 spring_torque_end_force(v3 xp, v3 xq, v3 vp, v3 vq, v3 rest_dir, float kt, float kw, int end) {
  const v3 xpq = xp - xq, vpq = vp - vq;
  const float len = sqrtf(dot(xpq, xpq));
  if (len < 1.0e-6f) return {0.f, 0.f, 0.f};
  const v3 d = {xpq.x / len, xpq.y / len, xpq.z / len};
  const v3 d0 = rest_dir * -1.f;
  if (sqrtf(dot(d0, d0)) < 0.5f) return {0.f, 0.f, 0.f};
  const float theta = acosf(fminf(fmaxf(dot(d, d0), -1.f), 1.f));
  v3 axis = cross(d0, d);
  const float rl = sqrtf(dot(axis, axis));
  if (rl < 1.0e-6f) return {0.f, 0.f, 0.f};
  axis = {axis.x / rl, axis.y / rl, axis.z / rl};
  const v3 dd = vpq - d * dot(d, vpq);
  const v3 d_dot = {dd.x / len, dd.y / len, dd.z / len};
  const float omega = dot(cross(d, d_dot), axis);
  const float tau = -(kt * theta + kw * omega) * len;
  const v3 t = cross(axis * tau, d);
  const v3 F = {t.x / len, t.y / len, t.z / len};
  return end == 0 ? F * -1.f : F;
}

FEM_DI v3 normalized(v3 a) {
  const float l = sqrtf(dot(a, a));
  return l > 0.f ? a * (1.f / l) : v3{0.f, 0.f, 0.f};
}

// Triangle membrane (eval_triangles kernels.py:876-970): the force on corner c
// (0: +f0, 1: -f1, 2: -f2) from the deviatoric and area-preservation terms,
// their damping, and aerodynamic drag and lift.  pose = inverse rest edge
// matrix [[D00, D01], [D10, D11]]; mat = (k_mu, k_lambda, k_damp, k_drag, k_lift)
// per unit rest area.
FEM_DI v3 tri_corner_force(v3 x0, v3 x1, v3 x2, v3 v0, v3 v1, v3 v2, const float* __restrict__ pose,
                           const float* __restrict__ mat, int c) {
  const float D00 = pose[0], D01 = pose[1], D10 = pose[2], D11 = pose[3];
  const float inv_rest_area = (D00 * D11 - D01 * D10) * 2.f;
  const float rest_area = 1.f / inv_rest_area;
  const float k_mu = mat[0] * rest_area, k_lambda = mat[1] * rest_area, k_damp = mat[2] * rest_area;
  const float k_drag = mat[3], k_lift = mat[4];
  const v3 x10 = x1 - x0, x20 = x2 - x0, v10 = v1 - v0, v20 = v2 - v0;
  const v3 F1 = x10 * D00 + x20 * D10, F2 = x10 * D01 + x20 * D11;
  const v3 dF1 = v10 * D00 + v20 * D10, dF2 = v10 * D01 + v20 * D11;
  const float Ic = dot(F1, F1) + dot(F2, F2);
  const float dev = (Ic - 2.f) / Ic;
  const v3 P1 = F1 * (k_mu * dev) + dF1 * k_damp, P2 = F2 * (k_mu * dev) + dF2 * k_damp;
  v3 f1 = P1 * D00 + P2 * D01, f2 = P1 * D10 + P2 * D11;
  v3 n = cross(x10, x20);
  const float area = sqrtf(dot(n, n)) * 0.5f;
  const float cst = area * inv_rest_area - 1.f;
  n = normalized(n);
  const v3 dcdq = cross(x20, n) * (inv_rest_area * 0.5f);
  const v3 dcdr = cross(n, x10) * (inv_rest_area * 0.5f);
  const float f_area = k_lambda * cst;
  const float dcdt = dot(dcdq, v1) + dot(dcdr, v2) - dot(dcdq + dcdr, v0);
  const float f_damp = k_damp * dcdt;
  f1 = f1 + dcdq * (f_area + f_damp);
  f2 = f2 + dcdr * (f_area + f_damp);
  v3 f0 = f1 + f2;
  const v3 vmid = (v0 + v1 + v2) * 0.3333f;
  const v3 vdir = normalized(vmid);
  const v3 f_drag = vmid * (k_drag * area * fabsf(dot(n, vmid)));
  const float cosang = fminf(fmaxf(dot(n, vdir), -1.f), 1.f);
  const v3 f_lift = n * (k_lift * area * (1.5707963267948966f - acosf(cosang))) * dot(vmid, vmid);
  f0 = f0 - f_drag - f_lift;
  f1 = f1 + f_drag + f_lift;
  f2 = f2 + f_drag + f_lift;
  return c == 0 ? f0 : (c == 1 ? f1 * -1.f : f2 * -1.f);
}

// Reverse of tri_corner_force without drag and lift: for weights w_c of the
// corner forces (L = sum_c w_c . F_c), the gradients with respect to the edge
// vectors x10, x20 and their velocities v10, v20 (node 0 receives minus their
// sum).  With a1 = w0 - w1, a2 = w0 - w2: L = <P, [a1 a2] D> + g (a1 . dc/dx10
// + a2 . dc/dx20), P = k_mu (1 - 2/Ic) F + k_damp dF, g = k_lambda c + k_damp dc/dt.
FEM_DI void tri_edge_vjp(v3 x0, v3 x1, v3 x2, v3 v0, v3 v1, v3 v2, v3 w0, v3 w1, v3 w2,
                         const float* __restrict__ pose, const float* __restrict__ mat, v3& gx10, v3& gx20, v3& gv10,
                         v3& gv20) {
  const float D00 = pose[0], D01 = pose[1], D10 = pose[2], D11 = pose[3];
  const float inv_area = (D00 * D11 - D01 * D10) * 2.f;
  const float A0 = 1.f / inv_area;
  const float kmu = mat[0] * A0, klam = mat[1] * A0, kd = mat[2] * A0;
  const v3 x10 = x1 - x0, x20 = x2 - x0, v10 = v1 - v0, v20 = v2 - v0;
  const v3 a1 = w0 - w1, a2 = w0 - w2;
  const v3 F1 = x10 * D00 + x20 * D10, F2 = x10 * D01 + x20 * D11;
  const float Ic = dot(F1, F1) + dot(F2, F2);
  const float dev = (Ic - 2.f) / Ic;
  const v3 b1 = a1 * D00 + a2 * D10, b2 = a1 * D01 + a2 * D11;
  const float cI = 4.f * (dot(F1, b1) + dot(F2, b2)) / (Ic * Ic);
  const v3 G1 = (b1 * dev + F1 * cI) * kmu, G2 = (b2 * dev + F2 * cI) * kmu;
  gx10 = G1 * D00 + G2 * D01;
  gx20 = G1 * D10 + G2 * D11;
  gv10 = (b1 * D00 + b2 * D01) * kd;
  gv20 = (b1 * D10 + b2 * D11) * kd;
  const v3 nr = cross(x10, x20);
  const float ln = sqrtf(dot(nr, nr));
  if (!(ln > 0.f)) return;
  const v3 nh = nr * (1.f / ln);
  const float kap = inv_area * 0.5f;
  const v3 cq = cross(x20, nh) * kap, cr = cross(nh, x10) * kap;  // dc/dx10, dc/dx20
  const float c = ln * kap - 1.f;
  const float cdot = dot(cq, v10) + dot(cr, v20), hA = dot(cq, a1) + dot(cr, a2);
  const float g = klam * c + kd * cdot;
  // Gradients of dc/dx10 . y1 + dc/dx20 . y2 with respect to (x10, x20).
  const v3 mv_ = cross(v10, x20) + cross(x10, v20), ma = cross(a1, x20) + cross(x10, a2);
  const v3 uv = (mv_ - nh * dot(nh, mv_)) * (1.f / ln), ua = (ma - nh * dot(nh, ma)) * (1.f / ln);
  const v3 dv1 = (cross(x20, uv) + cross(v20, nh)) * kap, dv2 = (cross(uv, x10) + cross(nh, v10)) * kap;
  const v3 da1 = (cross(x20, ua) + cross(a2, nh)) * kap, da2 = (cross(ua, x10) + cross(nh, a1)) * kap;
  gx10 = gx10 + (cq * klam + dv1 * kd) * hA + da1 * g;
  gx20 = gx20 + (cr * klam + dv2 * kd) * hA + da2 * g;
  gv10 = gv10 + cq * (kd * hA);
  gv20 = gv20 + cr * (kd * hA);
}

// Reverse of the drag and lift terms of tri_corner_force: every corner
// receives -(f_drag + f_lift) with f_drag = k_drag A |n.u| u and f_lift =
// k_lift A asin(clamp(n.u/|u|)) (u.u) n, u = 0.3333 (v0 + v1 + v2), so for
// corner weights w_c the term is L = -W.(f_drag + f_lift), W = w0 + w1 + w2.
// Adds the gradients with respect to the edges x10, x20 (node 0 receives
// minus their sum) and returns the gradient with respect to each corner's
// velocity (the same for all three).  The clamp passes no gradient where it
// binds, like the acos adjoint at |c| = 1.
FEM_DI v3 tri_aero_vjp(v3 x0, v3 x1, v3 x2, v3 v0, v3 v1, v3 v2, v3 W, const float* __restrict__ mat, v3& gx10,
                       v3& gx20) {
  const float k_drag = mat[3], k_lift = mat[4];
  const v3 zero = {0.f, 0.f, 0.f};
  if (k_drag == 0.f && k_lift == 0.f) return zero;
  const v3 x10 = x1 - x0, x20 = x2 - x0;
  const v3 N = cross(x10, x20);
  const float ln = sqrtf(dot(N, N));
  if (!(ln > 0.f)) return zero;  // degenerate triangle: no aerodynamic force
  const v3 n = N * (1.f / ln);
  const float area = ln * 0.5f;
  const v3 u = (v0 + v1 + v2) * 0.3333f;
  const float sn = dot(n, u), Wu = dot(W, u), Wn = dot(W, n);
  const float sg = sn > 0.f ? 1.f : (sn < 0.f ? -1.f : 0.f);
  v3 gu = (n * (sg * Wu) + W * fabsf(sn)) * (-k_drag * area);
  v3 gn = u * (sg * Wu * -k_drag * area);
  float garea = -k_drag * fabsf(sn) * Wu;
  const float uu = dot(u, u), ul = sqrtf(uu);
  if (k_lift != 0.f && ul > 0.f) {
    const v3 uh = u * (1.f / ul);
    const float craw = dot(n, uh);
    const float c = fminf(fmaxf(craw, -1.f), 1.f);
    const float th = 1.5707963267948966f - acosf(c);
    const float dd = 1.f - c * c;
    const float dth = (dd > 0.f && fabsf(craw) < 1.f) ? 1.f / sqrtf(dd) : 0.f;
    gu = gu + ((n - uh * c) * (uu * dth / ul) + u * (2.f * th)) * (-k_lift * area * Wn);
    gn = gn + (uh * (dth * Wn) + W * th) * (-k_lift * area * uu);
    garea += -k_lift * th * uu * Wn;
  }
  const v3 gN = n * (garea * 0.5f) + (gn - n * dot(n, gn)) * (1.f / ln);
  gx10 = gx10 + cross(x20, gN);
  gx20 = gx20 + cross(gN, x10);
  return gu * 0.3333f;
}

// Volumetric activation force on node a (eval_tet_activations kernels.py:484-508).
FEM_DI v3 tet_activation_force(v3 x0, v3 x1, v3 x2, v3 x3, float lam, float act, int a) {
  const float scale = lam * act / 6.f;
  const v3 fb = cross(x2 - x0, x3 - x0) * scale;
  const v3 fc = cross(x3 - x0, x1 - x0) * scale;
  const v3 fd = cross(x1 - x0, x2 - x0) * scale;
  return sel4(a, fb + fc + fd, fb * -1.f, fc * -1.f, fd * -1.f);
}

// One 3x3 block h2*[c1 (ga.gb) I + c2 (Fga)(Fgb)^T + c3 (Ja)(Jb)^T].
FEM_DI void tangent_block(float out[9], float c1, float c2, float c3, float h2, v3 ga, v3 fa, v3 ja, v3 gb, v3 fb,
                          v3 jb) {
  const float dab = c1 * dot(ga, gb);
  const float fa3[3] = {fa.x, fa.y, fa.z}, ja3[3] = {ja.x, ja.y, ja.z};
  const float fb3[3] = {fb.x, fb.y, fb.z}, jb3[3] = {jb.x, jb.y, jb.z};
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float kdev = ((r == c) ? dab : 0.f) + fa3[r] * fb3[c] * c2;
      const float kvol = ja3[r] * jb3[c] * c3;
      out[3 * r + c] = (kdev + kvol) * h2;
    }
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
  const float k_mu = mu * g.rest_volume, k_lambda = lam * g.rest_volume;
  float ic = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) ic += g.F[k] * g.F[k];
  const float ic1 = fmaxf(ic + 1.f, 1e-6f);
  const float s = 1.f - 1.f / ic1;
  const float coef = k_mu * 2.f / (ic1 * ic1);
  const float sv = g.inv_rest_volume / 6.f;
  const v3 j1 = cross(g.e2, g.e3) * sv, j2 = cross(g.e3, g.e1) * sv, j3 = cross(g.e1, g.e2) * sv;
  const v3 j0 = (j1 + j2 + j3) * -1.f;
  const v3 g1 = dm_row(g.Dm, 0), g2 = dm_row(g.Dm, 1), g3 = dm_row(g.Dm, 2);
  const v3 g0 = (g1 + g2 + g3) * -1.f;
  const v3 f0 = mv(g.F, g0), f1 = mv(g.F, g1), f2 = mv(g.F, g2), f3 = mv(g.F, g3);
  const v3 ga = sel4(a, g0, g1, g2, g3), fa = sel4(a, f0, f1, f2, f3), ja = sel4(a, j0, j1, j2, j3);
  const float c1 = k_mu * s;
  tangent_block(K[0], c1, coef, k_lambda, h2, ga, fa, ja, g0, f0, j0);
  tangent_block(K[1], c1, coef, k_lambda, h2, ga, fa, ja, g1, f1, j1);
  tangent_block(K[2], c1, coef, k_lambda, h2, ga, fa, ja, g2, f2, j2);
  tangent_block(K[3], c1, coef, k_lambda, h2, ga, fa, ja, g3, f3, j3);
}

// ---------------------------------------------------------------------------
// Hex8 Q1 element with 2x2x2 Gauss quadrature (reference kernels.py:1356-1747).
// Corner / Gauss-point signs: xi+ for {1,2,5,6}, eta+ for {2,3,6,7}, zeta+ for {4..7}.
// Node data are held as scalar arrays (gx[8], gy[8], gz[8]).
// ---------------------------------------------------------------------------
FEM_DI float hex_sx(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1.f : -1.f; }
FEM_DI float hex_sy(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1.f : -1.f; }
FEM_DI float hex_sz(int a) { return (a >= 4) ? 1.f : -1.f; }

// Reference-configuration gradients dN_b = inv_J0^T grad_xi N_b at Gauss point g.
FEM_DI void hex_grads(const float* __restrict__ invJ, int g, float gx[8], float gy[8], float gz[8]) {
  const float gp = 0.5773502691896258f;
  const float xi = hex_sx(g) * gp, eta = hex_sy(g) * gp, zeta = hex_sz(g) * gp;
  float J[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) J[k] = invJ[k];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const float xa = hex_sx(b), ya = hex_sy(b), za = hex_sz(b);
    const float dx = xa * (1.f + ya * eta) * (1.f + za * zeta) * 0.125f;
    const float dy = (1.f + xa * xi) * ya * (1.f + za * zeta) * 0.125f;
    const float dz = (1.f + xa * xi) * (1.f + ya * eta) * za * 0.125f;
    gx[b] = J[0] * dx + J[3] * dy + J[6] * dz;
    gy[b] = J[1] * dx + J[4] * dy + J[7] * dz;
    gz[b] = J[2] * dx + J[5] * dy + J[8] * dz;
  }
}

// F = sum_b x_b (outer) dN_b.
FEM_DI void hex_F(const float px[8], const float py[8], const float pz[8], const float gx[8], const float gy[8],
                  const float gz[8], float F[9]) {
#pragma unroll
  for (int k = 0; k < 9; ++k) F[k] = 0.f;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    F[0] += px[b] * gx[b]; F[1] += px[b] * gy[b]; F[2] += px[b] * gz[b];
    F[3] += py[b] * gx[b]; F[4] += py[b] * gy[b]; F[5] += py[b] * gz[b];
    F[6] += pz[b] * gx[b]; F[7] += pz[b] * gy[b]; F[8] += pz[b] * gz[b];
  }
}

// dN_a at Gauss point g for a runtime corner a, from its sign pattern
// (no indexing into per-corner arrays).
FEM_DI v3 hex_grad_corner(const float* __restrict__ invJ, int g, int a) {
  const float gp = 0.5773502691896258f;
  const float xi = hex_sx(g) * gp, eta = hex_sy(g) * gp, zeta = hex_sz(g) * gp;
  const float xa = hex_sx(a), ya = hex_sy(a), za = hex_sz(a);
  const float dx = xa * (1.f + ya * eta) * (1.f + za * zeta) * 0.125f;
  const float dy = (1.f + xa * xi) * ya * (1.f + za * zeta) * 0.125f;
  const float dz = (1.f + xa * xi) * (1.f + ya * eta) * za * 0.125f;
  return {invJ[0] * dx + invJ[3] * dy + invJ[6] * dz, invJ[1] * dx + invJ[4] * dy + invJ[7] * dz,
          invJ[2] * dx + invJ[5] * dy + invJ[8] * dz};
}

FEM_DI void cofactor(const float F[9], float cof[9]) {
  v3 c0 = {F[0], F[3], F[6]}, c1 = {F[1], F[4], F[7]}, c2 = {F[2], F[5], F[8]};
  v3 a = cross(c1, c2), b = cross(c2, c0), c = cross(c0, c1);
  cof[0] = a.x; cof[1] = b.x; cof[2] = c.x;
  cof[3] = a.y; cof[4] = b.y; cof[5] = c.y;
  cof[6] = a.z; cof[7] = b.z; cof[8] = c.z;
}

}  // namespace fem
