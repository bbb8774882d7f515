// Reverse-mode element math for the differentiable substep (device, fp32).
//
// The reference differentiates its force kernels with Warp's tape
// (differentiation.py:289-394 over eval_tetrahedra kernels.py:511-632,
// eval_tet_activations :484-508, eval_hexahedra :1481-1601,
// eval_particle_ground_contacts :1132-1195, eval_linear_damping_kernel
// :1289-1307) and the stretch clamp with its custom adjoint
// clamp_deformation_stretch_grad (:447-481).  These functions are the same
// products written out by hand: for a term buffer t (restoring forces) and
// per-node weights zw, they return d(sum_i zw_i . t_i) / d(inputs).
#pragma once

#include "fem_math.hpp"

namespace fem {

struct SVD3 {
  float U[9];  // row-major, columns are the left singular vectors
  float S[3];  // signed singular values, descending magnitude, sign of det F on S[2]
  float V[9];  // row-major, columns are the right singular vectors
};

// Signed SVD factors (the decomposition behind clamp_stretch).
__device__ __forceinline__ SVD3 svd3_factors(M9 Fin) {
  const float* F = Fin.m;
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
  SVD3 out;
  out.S[0] = dot(u0, b[0]);
  out.S[1] = dot(u1, b[1]);
  out.S[2] = dot(u2, b[2]);
  out.U[0] = u0.x; out.U[1] = u1.x; out.U[2] = u2.x;
  out.U[3] = u0.y; out.U[4] = u1.y; out.U[5] = u2.y;
  out.U[6] = u0.z; out.U[7] = u1.z; out.U[8] = u2.z;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) out.V[3 * i + j] = v[i][j];
  return out;
}

// Adjoint of the clamp away from the interior: spectral derivative of
// U clamp(S) V^T (divided differences on the symmetric part, (c_i+c_j)/(s_i+s_j)
// on the skew part), as clamp_deformation_stretch_grad states it.
// (The adjoint travels as nine scalars: a second by-value M9 would be passed
// through scratch memory.)
__device__ __noinline__ M9 clamp_grad_slow(M9 Fin, float a0, float a1, float a2, float a3, float a4, float a5,
                                           float a6, float a7, float a8, float mn, float mx) {
  M9 adj;
  adj.m[0] = a0; adj.m[1] = a1; adj.m[2] = a2;
  adj.m[3] = a3; adj.m[4] = a4; adj.m[5] = a5;
  adj.m[6] = a6; adj.m[7] = a7; adj.m[8] = a8;
  const SVD3 d = svd3_factors(Fin);
  const float smin = fminf(d.S[0], fminf(d.S[1], d.S[2])), smax = fmaxf(d.S[0], fmaxf(d.S[1], d.S[2]));
  if (smin > mn && smax < mx) return adj;
  float H[9], R[9], c[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) c[i] = fmaxf(fminf(d.S[i], mx), mn);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float s = 0.f;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) s += d.U[3 * a + i] * adj.m[3 * a + b] * d.V[3 * b + j];
      H[3 * i + j] = s;
    }
#pragma unroll
  for (int k = 0; k < 9; ++k) R[k] = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float slope = (d.S[i] > mn && d.S[i] < mx) ? 1.f : 0.f;
    R[3 * i + i] = slope * H[3 * i + i];
#pragma unroll
    for (int j = i + 1; j < 3; ++j) {
      const float difference = d.S[i] - d.S[j];
      float divided = slope;
      if (fabsf(difference) > 1.0e-6f) divided = (c[i] - c[j]) / difference;
      const float total = d.S[i] + d.S[j];
      float safe_total = fmaxf(total, 1.0e-6f);
      if (total < 0.f) safe_total = fminf(total, -1.0e-6f);
      const float summed = (c[i] + c[j]) / safe_total;
      const float symmetric = 0.5f * divided * (H[3 * i + j] + H[3 * j + i]);
      const float skew = 0.5f * summed * (H[3 * i + j] - H[3 * j + i]);
      R[3 * i + j] = symmetric + skew;
      R[3 * j + i] = symmetric - skew;
    }
  }
  M9 out;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      float s = 0.f;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) s += d.U[3 * i + a] * R[3 * a + b] * d.V[3 * j + b];
      out.m[3 * i + j] = s;
    }
  return out;
}

// out = adjoint of F given the adjoint `adj` of clamp(F).  The interior test
// is clamp_stretch's SVD-free bound (all stretches inside [mn, mx]).
FEM_DI void clamp_stretch_grad(const float F[9], float mn, float mx, const float adj[9], float out[9]) {
  bool interior = (mn < 0.f || mx < 0.f);
  if (!interior) {
    const float det = det3(F);
    float ic = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) ic += F[k] * F[k];
    interior = det > 0.f && ic <= mx * mx && 2.f * det >= mn * ic;
  }
  if (interior) {
#pragma unroll
    for (int k = 0; k < 9; ++k) out[k] = adj[k];
    return;
  }
  M9 f;
#pragma unroll
  for (int k = 0; k < 9; ++k) f.m[k] = F[k];
  const M9 r = clamp_grad_slow(f, adj[0], adj[1], adj[2], adj[3], adj[4], adj[5], adj[6], adj[7], adj[8], mn, mx);
#pragma unroll
  for (int k = 0; k < 9; ++k) out[k] = r.m[k];
}

// d det(F) / dF.
FEM_DI void det_grad(const float F[9], float C[9]) {
  C[0] = F[4] * F[8] - F[5] * F[7];
  C[1] = F[5] * F[6] - F[3] * F[8];
  C[2] = F[3] * F[7] - F[4] * F[6];
  C[3] = F[2] * F[7] - F[1] * F[8];
  C[4] = F[0] * F[8] - F[2] * F[6];
  C[5] = F[1] * F[6] - F[0] * F[7];
  C[6] = F[1] * F[5] - F[2] * F[4];
  C[7] = F[2] * F[3] - F[0] * F[5];
  C[8] = F[0] * F[4] - F[1] * F[3];
}

// M = E Dm for E with columns e1, e2, e3 (row-major result).
FEM_DI void cols_times_dm(v3 e1, v3 e2, v3 e3, const float Dm[9], float M[9]) {
  const float r0[3] = {e1.x, e2.x, e3.x}, r1[3] = {e1.y, e2.y, e3.y}, r2[3] = {e1.z, e2.z, e3.z};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    M[c] = r0[0] * Dm[c] + r0[1] * Dm[3 + c] + r0[2] * Dm[6 + c];
    M[3 + c] = r1[0] * Dm[c] + r1[1] * Dm[3 + c] + r1[2] * Dm[6 + c];
    M[6 + c] = r2[0] * Dm[c] + r2[1] * Dm[3 + c] + r2[2] * Dm[6 + c];
  }
}

// Column a of M Dm^T.
FEM_DI v3 times_dm_t_col(const float M[9], const float Dm[9], int a) {
  return {M[0] * Dm[3 * a] + M[1] * Dm[3 * a + 1] + M[2] * Dm[3 * a + 2],
          M[3] * Dm[3 * a] + M[4] * Dm[3 * a + 1] + M[5] * Dm[3 * a + 2],
          M[6] * Dm[3 * a] + M[7] * Dm[3 * a + 1] + M[8] * Dm[3 * a + 2]};
}

struct TetVjp {
  v3 xe[3];     // dL/d(x_a - x_0), a = 1..3
  v3 ve[3];     // dL/d(v_a - v_0)
  float act;    // dL/d activation
  float mat[3]; // dL/d(mu, lambda, k_damp)
};

// Reverse of one Tet4's restoring force (+ activation force when use_act):
// L = sum_a zw_a . t_a.  Nodes follow the element order (x0..x3).
FEM_DI TetVjp tet_vjp(v3 x0, v3 x1, v3 x2, v3 x3, v3 v0, v3 v1, v3 v2, v3 v3_, v3 z0, v3 z1, v3 z2, v3 z3,
                      const float* __restrict__ pose, float mu, float lam, float kd, bool use_act, float act, float mn,
                      float mx) {
  TetVjp o;
  float Dm[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Dm[k] = pose[k];
  const float s = det3(Dm);
  const v3 x10 = x1 - x0, x20 = x2 - x0, x30 = x3 - x0;
  const v3 w1 = z1 - z0, w2 = z2 - z0, w3 = z3 - z0;
  const v3 dJ1 = cross(x20, x30) * s, dJ2 = cross(x30, x10) * s, dJ3 = cross(x10, x20) * s;
  const float cw = dot(w1, dJ1) + dot(w2, dJ2) + dot(w3, dJ3);
  v3 dJb1 = {0.f, 0.f, 0.f}, dJb2 = dJb1, dJb3 = dJb1;
  float cwb = 0.f;
  float Fb[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Fb[k] = 0.f;
  o.ve[0] = o.ve[1] = o.ve[2] = v3{0.f, 0.f, 0.f};
  o.act = 0.f;
  o.mat[0] = o.mat[1] = o.mat[2] = 0.f;
  if (!(mu == 0.f && lam == 0.f && kd == 0.f)) {
    const v3 u1 = v1 - v0, u2 = v2 - v0, u3 = v3_ - v0;
    const float V = 1.f / (s * 6.f);
    const float alpha = 1.f + mu / lam - mu / (4.f * lam);
    const float muV = mu * V, lamV = lam * V, kdV = kd * V;
    float F[9], Fc[9], dF[9], G[9];
    cols_times_dm(x10, x20, x30, Dm, F);
    cols_times_dm(u1, u2, u3, Dm, dF);
    cols_times_dm(w1, w2, w3, Dm, G);
#pragma unroll
    for (int k = 0; k < 9; ++k) Fc[k] = F[k];
    clamp_stretch(Fc, mn, mx);
    float ic = 0.f, GF = 0.f, GdF = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      ic += Fc[k] * Fc[k];
      GF += G[k] * Fc[k];
      GdF += G[k] * dF[k];
    }
    const float phi = 1.f - 1.f / (ic + 1.f);
    const float dphi = 1.f / ((ic + 1.f) * (ic + 1.f));
    const float detF = det3(Fc);
    const float J = fmaxf(detF, 0.01f);
    const float cv = dot(u1, dJ1) + dot(u2, dJ2) + dot(u3, dJ3);
    const float ft = (J - alpha) * lamV + kdV * cv;
    const float ftb = -cw;
    cwb = -ft;
    float Fcb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fcb[k] = -muV * phi * G[k] - 2.f * muV * dphi * GF * Fc[k];
    if (detF > 0.01f) {
      float C[9];
      det_grad(Fc, C);
#pragma unroll
      for (int k = 0; k < 9; ++k) Fcb[k] += ftb * lamV * C[k];
    }
    clamp_stretch_grad(F, mn, mx, Fcb, Fb);
    float Gk[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Gk[k] = -kdV * G[k];
    o.ve[0] = times_dm_t_col(Gk, Dm, 0) + dJ1 * (ftb * kdV);
    o.ve[1] = times_dm_t_col(Gk, Dm, 1) + dJ2 * (ftb * kdV);
    o.ve[2] = times_dm_t_col(Gk, Dm, 2) + dJ3 * (ftb * kdV);
    dJb1 = u1 * (ftb * kdV);
    dJb2 = u2 * (ftb * kdV);
    dJb3 = u3 * (ftb * kdV);
    const float lamVb = ftb * (J - alpha);
    const float alphab = -ftb * lamV;
    const float muVb = -phi * GF;
    const float kdVb = -GdF + ftb * cv;
    o.mat[0] = muVb * V + alphab * (0.75f / lam);
    o.mat[1] = lamVb * V + alphab * (-0.75f * mu / (lam * lam));
    o.mat[2] = kdVb * V;
  }
  if (use_act) {
    const float sc = lam * act / 6.f;
    cwb += -sc / s;
    o.act = -(lam / 6.f) * cw / s;
    o.mat[1] += -(act / 6.f) * cw / s;
  }
  dJb1 = dJb1 + w1 * cwb;
  dJb2 = dJb2 + w2 * cwb;
  dJb3 = dJb3 + w3 * cwb;
  o.xe[0] = (cross(dJb2, x30) + cross(x20, dJb3)) * s + times_dm_t_col(Fb, Dm, 0);
  o.xe[1] = (cross(dJb3, x10) + cross(x30, dJb1)) * s + times_dm_t_col(Fb, Dm, 1);
  o.xe[2] = (cross(dJb1, x20) + cross(x10, dJb2)) * s + times_dm_t_col(Fb, Dm, 2);
  return o;
}

// -z^T (dA/dmu) y and -z^T (dA/dlambda) y of one Tet4 (A assembled at x).
FEM_DI void tet_matrix_material_terms(v3 x0, v3 x1, v3 x2, v3 x3, v3 wz1, v3 wz2, v3 wz3, v3 wy1, v3 wy2, v3 wy3,
                                      const float* __restrict__ pose, float mn, float mx, float h2, float& mub,
                                      float& lamb) {
  float Dm[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) Dm[k] = pose[k];
  const float s = det3(Dm);
  mub = 0.f;
  lamb = 0.f;
  if (s * 6.f <= 0.f) return;
  const float V = 1.f / (s * 6.f);
  const v3 x10 = x1 - x0, x20 = x2 - x0, x30 = x3 - x0;
  float F[9], Gz[9], Gy[9];
  cols_times_dm(x10, x20, x30, Dm, F);
  cols_times_dm(wz1, wz2, wz3, Dm, Gz);
  cols_times_dm(wy1, wy2, wy3, Dm, Gy);
  clamp_stretch(F, mn, mx);
  float ic = 0.f, zy = 0.f, zF = 0.f, yF = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    ic += F[k] * F[k];
    zy += Gz[k] * Gy[k];
    zF += Gz[k] * F[k];
    yF += Gy[k] * F[k];
  }
  const float ic1 = fmaxf(ic + 1.f, 1e-6f);
  const float sdev = 1.f - 1.f / ic1;
  const v3 dJ1 = cross(x20, x30) * s, dJ2 = cross(x30, x10) * s, dJ3 = cross(x10, x20) * s;
  const float cz = dot(wz1, dJ1) + dot(wz2, dJ2) + dot(wz3, dJ3);
  const float cy = dot(wy1, dJ1) + dot(wy2, dJ2) + dot(wy3, dJ3);
  mub = -h2 * V * (sdev * zy + 2.f / (ic1 * ic1) * zF * yF);
  lamb = -h2 * V * cz * cy;
}

// Reverse of the ground penalty + friction force on one node.
FEM_DI void ground_vjp(v3 x, v3 v, v3 z, float radius, float inv_m, const float gn[4], float ke, float kd, float kf,
                       float gmu, float gmag, v3& xb, v3& vb) {
  xb = v3{0.f, 0.f, 0.f};
  vb = xb;
  const v3 n = {gn[0], gn[1], gn[2]};
  const float c = fminf(dot(n, x) + gn[3] - radius, 0.f);
  if (c >= 0.f) return;
  const float vn = dot(n, v);
  const float fn = c * ke + fminf(vn, 0.f) * kd;
  const v3 vt = v - n * vn;
  const float vs = sqrtf(dot(vt, vt));
  const v3 that = vs > 0.f ? vt * (1.f / vs) : vt;
  const float m = inv_m > 0.f ? 1.f / fmaxf(inv_m, 1e-9f) : 0.f;
  const float n_cap = m * gmag;
  const float afn = fabsf(fn);
  const float fn_eff = fminf(afn, n_cap);
  const float ftf = fminf(vs * kf, gmu * fn_eff);
  float fnb = -dot(z, n);
  const float ftb = -dot(z, that);
  const v3 thatb = z * (-ftf);
  float vsb = 0.f, fneb = 0.f;
  if (vs * kf < gmu * fn_eff) vsb += ftb * kf;
  else fneb += ftb * gmu;
  if (afn < n_cap) fnb += fneb * (fn < 0.f ? -1.f : 1.f);
  v3 vtb;
  if (vs > 0.f) {
    const float tt = dot(that, thatb);
    vtb = (thatb - that * tt) * (1.f / vs) + that * vsb;
  } else {
    vtb = thatb;
  }
  float vnb = -dot(n, vtb);
  if (vn < 0.f) vnb += fnb * kd;
  vb = vtb + n * vnb;
  xb = n * (fnb * ke);
}

}  // namespace fem
