// BiCGSTAB, conjugate residual and restarted GMRES for the step's linear
// solve, and the point-Jacobi preconditioners (see fem_krylov.hpp).
//
// Semantics restated from warp.optim.linear (warp-lang 1.17, the reference's
// pinned dependency; called from solver.py:1219-1252):
//   * tolerance: atol = max(tol ||b||, eps32^(9/4)) with atol = 0 given;
//     every update is masked by the latest residual test rr > atol^2;
//   * bicgstab: right-preconditioned BiCGSTAB with shadow residual r0 = r(0),
//     rho = <r0, r>, alpha = rho / <r0, v>, omega = <r, t> / <t, t>,
//     beta = (rho_new / <r0, v>) / omega, p = r + beta (p - omega v);
//   * cr: preconditioned conjugate residual with z = M r, y = M Ap,
//     alpha = <z, Az> / <Ap, y>, beta = <z, Az>_new / <z, Az>_old;
//   * gmres: restarted right-preconditioned GMRES(restart), classical
//     Gram-Schmidt Arnoldi on M, Givens least squares with pivot tolerance
//     eps32^2, x += M V y, true residual recomputed after every cycle.
// Breakdown guards (zero denominators give a zero step) are ours.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fem_krylov.hpp"
#include "fem_math.hpp"

namespace fem {
namespace {

constexpr int NT = KR_BLOCK;
constexpr int OPG = 4;            // lanes per row of the operator kernel
constexpr int OPR = NT / OPG;     // rows per operator block

enum Slot {
  S_SCALE = 0, S_ZERO, S_RR, S_BB, S_ATOL2, S_ITERS, S_RHO, S_R0V, S_ALPHA, S_OMEGA, S_BETA, S_ZAZ, S_LIVE, S_GBETA,
  S_H = 32  // GMRES: H [(restart+1) x restart] then y [restart]
};

enum Step {
  ST_NONE = 0, ST_INIT, ST_RR, ST_BI_ALPHA, ST_BI_OMEGA, ST_BI_BETA, ST_CR_ZAZ, ST_CR_ALPHA, ST_CR_BETA,
  ST_GM_BEGIN, ST_GM_H, ST_GM_NORM, ST_GM_SOLVE
};

__device__ __forceinline__ float* vec(const KrylovArgs& a, int k) { return a.work + 3 * (size_t)a.P * k; }
__device__ __forceinline__ v3 precond(const KrylovArgs& a, int i, v3 u) {
  return a.M ? mv(a.M + 9 * (size_t)i, u) : u;
}
__device__ __forceinline__ double dotd(v3 a, v3 b) {
  return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z;
}

// Per-block sums of n values (n <= pstride) into part[block * pstride + k];
// fixed order: waves by shuffles, then the block's waves in order.
__device__ __forceinline__ void block_partial(double v, double* part, int pstride, int k) {
  __shared__ double red[NT / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = red[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) s += red[w];
    part[(size_t)blockIdx.x * pstride + k] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Setup: scale of the normalised right-hand side (kernels.py:1779-1814) and
// the starting iterate.
// ---------------------------------------------------------------------------
__global__ void k_kr_scale(KrylovArgs a) {
  float m = 0.f;
  for (int k = threadIdx.x; k < a.rhs_max_slots; k += 64) m = fmaxf(m, __uint_as_float(a.rhs_max[k]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (threadIdx.x == 0) {
    const float s = (m > 0.f && m < 1e-12f) ? fminf(1.f / m, 1e30f) : 1.f;
    a.sc[S_SCALE] = s;
    a.sc[S_ZERO] = (m == 0.f) ? 1.0 : 0.0;
    a.sc[S_ITERS] = 0.0;
  }
}

__global__ __launch_bounds__(NT) void k_kr_setup(KrylovArgs a) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.P) return;
  const float s = (float)a.sc[S_SCALE];
  const bool zero = a.sc[S_ZERO] != 0.0 || a.row_fixed[i];
  st3(vec(a, 1), i, ld3(a.rhs, i) * s);
  st3(vec(a, 0), i, zero ? v3{0.f, 0.f, 0.f} : ld3(a.dv, i) * s);
}

// y = Op u (Op = A, or PAP + I - P over the fixed rows, fixed_particles.py:29-53);
// with bsub, y = bsub - Op u.  Partial sums: <d0, y>, <d1, y>, then <y, y> if yy.
__global__ __launch_bounds__(NT) void k_kr_op(KrylovArgs a, const float* __restrict__ u, float* __restrict__ y,
                                              const float* __restrict__ bsub, const float* __restrict__ d0,
                                              const float* __restrict__ d1, int yy, int pstride) {
  const int grp = threadIdx.x / OPG, lane = threadIdx.x % OPG;
  const int i = blockIdx.x * OPR + grp;
  const bool active = i < a.P;
  const bool fixed_row = active && a.row_fixed[i] != 0;
  v3 acc = {0.f, 0.f, 0.f};
  if (active && !fixed_row) {
    const int k1 = a.offsets[i + 1];
    for (int k = a.offsets[i] + lane; k < k1; k += OPG) {
      const int c = a.cols[k];
      if (a.row_fixed[c]) continue;
      acc = acc + mv(a.vals + 9 * (size_t)k, ld3(u, c));
    }
  }
#pragma unroll
  for (int o = OPG / 2; o > 0; o >>= 1) {
    acc.x += __shfl_xor(acc.x, o);
    acc.y += __shfl_xor(acc.y, o);
    acc.z += __shfl_xor(acc.z, o);
  }
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  if (active && lane == 0) {
    v3 r = fixed_row ? ld3(u, i) : acc;
    if (bsub) r = ld3(bsub, i) - r;
    st3(y, i, r);
    if (d0) p0 = dotd(ld3(d0, i), r);
    if (d1) p1 = dotd(ld3(d1, i), r);
    if (yy) p2 = dotd(r, r);
  }
  int k = 0;
  if (d0) block_partial(p0, a.part, pstride, k++);
  if (d1) block_partial(p1, a.part, pstride, k++);
  if (yy) block_partial(p2, a.part, pstride, k++);
}

// <u, w> partials (slot 0).
__global__ __launch_bounds__(NT) void k_kr_dot(KrylovArgs a, const float* __restrict__ u, const float* __restrict__ w,
                                               int pstride) {
  const int i = blockIdx.x * NT + threadIdx.x;
  const double v = i < a.P ? dotd(ld3(u, i), ld3(w, i)) : 0.0;
  block_partial(v, a.part, pstride, 0);
}

// r0 = r, p = r (BiCGSTAB start).
__global__ __launch_bounds__(NT) void k_bi_start(KrylovArgs a) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.P) return;
  const v3 r = ld3(vec(a, 2), i);
  st3(vec(a, 3), i, r);
  st3(vec(a, 4), i, r);
}

// BiCGSTAB vectors: x0 b1 r2 r0_3 p4 v5 t6 y7 z8.
// p = r + beta (p - omega v) (not on the first iteration); y = M p.
__global__ __launch_bounds__(NT) void k_bi_p(KrylovArgs a, int first) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.P) return;
  v3 p = ld3(vec(a, 4), i);
  if (!first) {
    const float beta = (float)a.sc[S_BETA], omega = (float)a.sc[S_OMEGA];
    p = ld3(vec(a, 2), i) + (p - ld3(vec(a, 5), i) * omega) * beta;
    st3(vec(a, 4), i, p);
  }
  st3(vec(a, 7), i, precond(a, i, p));
}

// x += alpha y, r -= alpha v, z = M r; partial <r, r>.
__global__ __launch_bounds__(NT) void k_bi_upd1(KrylovArgs a, int pstride) {
  const int i = blockIdx.x * NT + threadIdx.x;
  double rr = 0.0;
  if (i < a.P) {
    const float alpha = (float)a.sc[S_ALPHA];
    st3(vec(a, 0), i, ld3(vec(a, 0), i) + ld3(vec(a, 7), i) * alpha);
    const v3 r = ld3(vec(a, 2), i) - ld3(vec(a, 5), i) * alpha;
    st3(vec(a, 2), i, r);
    st3(vec(a, 8), i, precond(a, i, r));
    rr = dotd(r, r);
  }
  block_partial(rr, a.part, pstride, 0);
}

// x += omega z, r -= omega t; partials <r, r>, <r0, r>.
__global__ __launch_bounds__(NT) void k_bi_upd2(KrylovArgs a, int pstride) {
  const int i = blockIdx.x * NT + threadIdx.x;
  double rr = 0.0, rho = 0.0;
  if (i < a.P) {
    const float omega = (float)a.sc[S_OMEGA];
    st3(vec(a, 0), i, ld3(vec(a, 0), i) + ld3(vec(a, 8), i) * omega);
    const v3 r = ld3(vec(a, 2), i) - ld3(vec(a, 6), i) * omega;
    st3(vec(a, 2), i, r);
    rr = dotd(r, r);
    rho = dotd(ld3(vec(a, 3), i), r);
  }
  block_partial(rr, a.part, pstride, 0);
  block_partial(rho, a.part, pstride, 1);
}

// CR vectors: x0 b1 r2 z3 p4 Ap5 Az6 y7.  z = M r.
__global__ __launch_bounds__(NT) void k_cr_z(KrylovArgs a) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < a.P) st3(vec(a, 3), i, precond(a, i, ld3(vec(a, 2), i)));
}

// p = z + beta p, Ap = Az + beta Ap (first: p = z, Ap = Az); y = M Ap; partial <Ap, y>.
__global__ __launch_bounds__(NT) void k_cr_p(KrylovArgs a, int first, int pstride) {
  const int i = blockIdx.x * NT + threadIdx.x;
  double yap = 0.0;
  if (i < a.P) {
    v3 p = ld3(vec(a, 3), i), ap = ld3(vec(a, 6), i);
    if (!first) {
      const float beta = (float)a.sc[S_BETA];
      p = p + ld3(vec(a, 4), i) * beta;
      ap = ap + ld3(vec(a, 5), i) * beta;
    }
    st3(vec(a, 4), i, p);
    st3(vec(a, 5), i, ap);
    const v3 y = precond(a, i, ap);
    st3(vec(a, 7), i, y);
    yap = dotd(ap, y);
  }
  block_partial(yap, a.part, pstride, 0);
}

// x += alpha p, r -= alpha Ap, z -= alpha y; partial <r, r>.
__global__ __launch_bounds__(NT) void k_cr_upd(KrylovArgs a, int pstride) {
  const int i = blockIdx.x * NT + threadIdx.x;
  double rr = 0.0;
  if (i < a.P) {
    const float alpha = (float)a.sc[S_ALPHA];
    st3(vec(a, 0), i, ld3(vec(a, 0), i) + ld3(vec(a, 4), i) * alpha);
    const v3 r = ld3(vec(a, 2), i) - ld3(vec(a, 5), i) * alpha;
    st3(vec(a, 2), i, r);
    st3(vec(a, 3), i, ld3(vec(a, 3), i) - ld3(vec(a, 7), i) * alpha);
    rr = dotd(r, r);
  }
  block_partial(rr, a.part, pstride, 0);
}

// GMRES vectors: x0 b1 r2 w3, V_k = 4 + k (k = 0 .. restart).  Every kernel
// of a cycle does nothing when the cycle started converged (S_LIVE = 0).
// V0 = r / beta; t = M V0 into V1.
__global__ __launch_bounds__(NT) void k_gm_start(KrylovArgs a) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.P || a.sc[S_LIVE] == 0.0) return;
  const float beta = (float)a.sc[S_GBETA];
  const v3 v = ld3(vec(a, 2), i) * (1.f / beta);
  st3(vec(a, 4), i, v);
  st3(vec(a, 5), i, precond(a, i, v));
}

// Classical Gram-Schmidt: partials <w, V_k>, k = 0 .. j.
__global__ __launch_bounds__(NT) void k_gm_dots(KrylovArgs a, int j, int pstride) {
  const int i = blockIdx.x * NT + threadIdx.x;
  const bool on = i < a.P && a.sc[S_LIVE] != 0.0;
  const v3 w = on ? ld3(vec(a, 3), i) : v3{0.f, 0.f, 0.f};
  for (int k = 0; k <= j; ++k) block_partial(on ? dotd(w, ld3(vec(a, 4 + k), i)) : 0.0, a.part, pstride, k);
}

// w -= sum_k H[k][j] V_k; partial <w, w>.
__global__ __launch_bounds__(NT) void k_gm_orth(KrylovArgs a, int j, int pstride) {
  const int i = blockIdx.x * NT + threadIdx.x;
  double ww = 0.0;
  if (i < a.P && a.sc[S_LIVE] != 0.0) {
    const double* H = a.sc + S_H;
    v3 w = ld3(vec(a, 3), i);
    for (int k = 0; k <= j; ++k) w = w - ld3(vec(a, 4 + k), i) * (float)H[k * a.restart + j];
    st3(vec(a, 3), i, w);
    ww = dotd(w, w);
  }
  block_partial(ww, a.part, pstride, 0);
}

// V_{j+1} = w / H[j+1][j] (w itself when that norm is 0); t = M V_{j+1} into
// V_{j+2} unless the cycle ends.
__global__ __launch_bounds__(NT) void k_gm_next(KrylovArgs a, int j) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.P || a.sc[S_LIVE] == 0.0) return;
  const float hn = (float)a.sc[S_H + (j + 1) * a.restart + j];
  v3 v = ld3(vec(a, 3), i);
  if (hn != 0.f) v = v * (1.f / hn);
  st3(vec(a, 4 + j + 1), i, v);
  if (j + 1 < a.restart) st3(vec(a, 4 + j + 2), i, precond(a, i, v));
}

// x += M (V y).
__global__ __launch_bounds__(NT) void k_gm_x(KrylovArgs a) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= a.P || a.sc[S_LIVE] == 0.0) return;
  const double* y = a.sc + S_H + (a.restart + 1) * a.restart;
  v3 w = {0.f, 0.f, 0.f};
  for (int k = 0; k < a.restart; ++k) w = w + ld3(vec(a, 4 + k), i) * (float)y[k];
  st3(vec(a, 0), i, ld3(vec(a, 0), i) + precond(a, i, w));
}

// ---------------------------------------------------------------------------
// Reduction of the block partials (fixed order) and the scalar step.
// ---------------------------------------------------------------------------
__device__ void scalar_step(const KrylovArgs& a, int step, int j, const double* s) {
  double* sc = a.sc;
  switch (step) {
    case ST_INIT: {  // s = <b, b>; S_RR holds <r, r> of the starting iterate
      sc[S_BB] = s[0];
      const double atol = fmax(fmax((double)a.tolerance * sqrt(s[0]), 0.0), 2.6442191937624087e-16);
      sc[S_ATOL2] = atol * atol;
      sc[S_RHO] = sc[S_RR];  // BiCGSTAB: rho = <r0, r> with r0 = r
      break;
    }
    case ST_RR:
      sc[S_RR] = s[0];
      break;
    case ST_BI_ALPHA: {  // s = <r0, v>
      const bool live = sc[S_RR] > sc[S_ATOL2];
      if (live) sc[S_ITERS] += 1.0;
      sc[S_R0V] = s[0];
      sc[S_ALPHA] = (live && s[0] != 0.0) ? sc[S_RHO] / s[0] : 0.0;
      break;
    }
    case ST_BI_OMEGA:  // s = (<r, t>, <t, t>)
      sc[S_OMEGA] = (sc[S_RR] > sc[S_ATOL2] && s[1] > 0.0) ? s[0] / s[1] : 0.0;
      break;
    case ST_BI_BETA: {  // s = (<r, r>, <r0, r>)
      sc[S_RR] = s[0];
      const bool ok = s[0] > sc[S_ATOL2] && sc[S_R0V] != 0.0 && sc[S_OMEGA] != 0.0;
      sc[S_BETA] = ok ? (s[1] / sc[S_R0V]) / sc[S_OMEGA] : 0.0;
      sc[S_RHO] = s[1];
      break;
    }
    case ST_CR_ZAZ:
      sc[S_ZAZ] = s[0];
      break;
    case ST_CR_ALPHA: {  // s = <Ap, y>
      const bool live = sc[S_RR] > sc[S_ATOL2];
      if (live) sc[S_ITERS] += 1.0;
      sc[S_ALPHA] = (live && s[0] > 0.0) ? sc[S_ZAZ] / s[0] : 0.0;
      break;
    }
    case ST_CR_BETA:  // s = <z, Az> of the new z
      sc[S_BETA] = (sc[S_RR] > sc[S_ATOL2] && sc[S_ZAZ] > 0.0) ? s[0] / sc[S_ZAZ] : 0.0;
      sc[S_ZAZ] = s[0];
      break;
    case ST_GM_BEGIN: {
      const bool live = sc[S_RR] > sc[S_ATOL2];
      sc[S_LIVE] = live ? 1.0 : 0.0;
      sc[S_GBETA] = sqrt(sc[S_RR]);
      if (live) sc[S_ITERS] += a.restart;
      break;
    }
    case ST_GM_H:  // s = <w, V_k>, k = 0 .. j
      for (int k = 0; k <= j; ++k) sc[S_H + k * a.restart + j] = s[k];
      break;
    case ST_GM_NORM:
      sc[S_H + (j + 1) * a.restart + j] = sqrt(s[0]);
      break;
    case ST_GM_SOLVE: {
      // min ||beta e1 - H y||: Givens rotations of the Hessenberg matrix and
      // the right-hand side, then back substitution over the columns whose
      // pivot exceeds eps32^2.
      const int k = a.restart;
      double* H = sc + S_H;
      double* y = H + (k + 1) * k;
      double g = sc[S_GBETA];
      int kmax = k;
      for (int i = 0; i < k; ++i) {
        const double ha = H[i * k + i], hb = H[(i + 1) * k + i];
        const double nn = ha * ha + hb * hb;
        if (nn < 1.4210854715202004e-14) {
          kmax = i;
          break;
        }
        const double n = sqrt(nn), c = ha / n, sn = hb / n;
        for (int c2 = i; c2 < k; ++c2) {
          const double u = H[i * k + c2], w = H[(i + 1) * k + c2];
          H[i * k + c2] = c * u + sn * w;
          H[(i + 1) * k + c2] = c * w - sn * u;
        }
        y[i] = c * g;
        g = -sn * g;
      }
      for (int i = kmax; i < k; ++i) y[i] = 0.0;
      for (int i = kmax - 1; i >= 0; --i) {
        double yi = y[i];
        for (int c2 = i + 1; c2 < kmax; ++c2) yi -= H[i * k + c2] * y[c2];
        y[i] = yi / H[i * k + i];
      }
      if (sc[S_LIVE] == 0.0)
        for (int i = 0; i < k; ++i) y[i] = 0.0;
      break;
    }
    default:
      break;
  }
}

__global__ __launch_bounds__(NT) void k_kr_fin(KrylovArgs a, int nblk, int nv, int pstride, int step, int j) {
  __shared__ double red[NT / 64];
  __shared__ double sums[KR_MAX_RESTART + 1];
  for (int k = 0; k < nv; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += NT) v += a.part[(size_t)b * pstride + k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = red[0];
#pragma unroll
      for (int w = 1; w < NT / 64; ++w) s += red[w];
      sums[k] = s;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) scalar_step(a, step, j, sums);
}

// restore_linear_increment kernels.py:1808-1811, update_state :115-128,
// FixedParticles.restore(state_out); iteration statistics.
__global__ __launch_bounds__(NT) void k_kr_finish(KrylovArgs a) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i == 0) {
    a.stats[0] = (int)a.sc[S_ITERS];
    a.stats[1] = __float_as_int((float)a.sc[S_RR]);
    a.stats[2] = __float_as_int((float)a.sc[S_ATOL2]);
  }
  if (i >= a.P) return;
  const float inv = 1.f / (float)a.sc[S_SCALE];
  const v3 dv = a.row_fixed[i] ? v3{0.f, 0.f, 0.f} : ld3(vec(a, 0), i) * inv;
  st3(a.dv, i, dv);
  if (!a.q_out) return;
  v3 vel = ld3(a.qd_in, i) + dv;
  v3 q = ld3(a.q_in, i) + vel * a.h;
  if (a.row_fixed[i]) {
    q = ld3(a.anchor, i);
    vel = v3{0.f, 0.f, 0.f};
  }
  st3(a.q_out, i, q);
  st3(a.qd_out, i, vel);
}

__global__ __launch_bounds__(NT) void k_point_jacobi(int P, const int* __restrict__ offsets,
                                                     const int* __restrict__ diag_rel, const float* __restrict__ vals,
                                                     int kind, float* __restrict__ out) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= P) return;
  float d[3] = {0.f, 0.f, 0.f};
  if (kind != KP_ID && diag_rel[i] >= 0) {
    const float* b = vals + 9 * (size_t)(offsets[i] + diag_rel[i]);
    d[0] = b[0];
    d[1] = b[4];
    d[2] = b[8];
  }
  float* o = out + 9 * (size_t)i;
#pragma unroll
  for (int c = 0; c < 9; ++c) o[c] = 0.f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float e = kind == KP_DIAG_ABS ? fabsf(d[k]) : d[k];
    o[4 * k] = (kind == KP_ID || e == 0.f) ? 1.f : 1.f / e;
  }
}

}  // namespace

size_t krylov_work_floats(int P, int method, int restart) {
  const int nvec = method == KR_BICGSTAB ? 9 : method == KR_CR ? 8 : method == KR_GMRES ? 5 + restart : 0;
  return (size_t)nvec * 3 * (size_t)P;
}

hipError_t krylov_point_jacobi(int P, const int* offsets, const int* diag_rel, const float* vals, int kind, float* out,
                               hipStream_t st) {
  if (P == 0) return hipSuccess;
  hipLaunchKernelGGL(k_point_jacobi, dim3((P + NT - 1) / NT), dim3(NT), 0, st, P, offsets, diag_rel, vals, kind, out);
  return hipGetLastError();
}

hipError_t krylov_solve(const KrylovArgs& a, int method, hipStream_t st) {
  if (a.P == 0) return hipSuccess;
  const int nb = (a.P + NT - 1) / NT, nbo = (a.P + OPR - 1) / OPR;
  const int ps = krylov_part_stride(a.restart);
  const dim3 g(nb), go(nbo), t(NT);
  auto V = [&](int k) { return a.work + 3 * (size_t)a.P * k; };
  float *x = V(0), *b = V(1), *r = V(2);
  auto fin = [&](int nblk, int nv, int step, int j) {
    hipLaunchKernelGGL(k_kr_fin, dim3(1), t, 0, st, a, nblk, nv, ps, step, j);
  };
  auto op = [&](const float* u, float* y, const float* bsub, const float* d0, int yy) {
    hipLaunchKernelGGL(k_kr_op, go, t, 0, st, a, u, y, bsub, d0, (const float*)nullptr, yy, ps);
  };
  hipLaunchKernelGGL(k_kr_scale, dim3(1), dim3(64), 0, st, a);
  hipLaunchKernelGGL(k_kr_setup, g, t, 0, st, a);
  op(x, r, b, nullptr, 1);  // r = b - Op x; <r, r>
  fin(nbo, 1, ST_RR, 0);
  hipLaunchKernelGGL(k_kr_dot, g, t, 0, st, a, (const float*)b, (const float*)b, ps);  // <b, b>: tolerance
  fin(nb, 1, ST_INIT, 0);
  if (method == KR_BICGSTAB) {
    float *r0 = V(3), *v = V(5), *tv = V(6), *y = V(7), *z = V(8);
    hipLaunchKernelGGL(k_bi_start, g, t, 0, st, a);
    for (int it = 0; it < a.maxiter; ++it) {
      hipLaunchKernelGGL(k_bi_p, g, t, 0, st, a, (int)(it == 0));  // p update, y = M p
      op(y, v, nullptr, r0, 0);                                      // v = Op y; <r0, v>
      fin(nbo, 1, ST_BI_ALPHA, 0);
      hipLaunchKernelGGL(k_bi_upd1, g, t, 0, st, a, ps);            // x, r, z = M r; <r, r>
      fin(nb, 1, ST_RR, 0);
      op(z, tv, nullptr, r, 1);                                      // t = Op z; <r, t>, <t, t>
      fin(nbo, 2, ST_BI_OMEGA, 0);
      hipLaunchKernelGGL(k_bi_upd2, g, t, 0, st, a, ps);            // x, r; <r, r>, <r0, r>
      fin(nb, 2, ST_BI_BETA, 0);
    }
  } else if (method == KR_CR) {
    float *z = V(3), *az = V(6);
    hipLaunchKernelGGL(k_cr_z, g, t, 0, st, a);  // z = M r
    op(z, az, nullptr, z, 0);                     // Az = Op z; <z, Az>
    fin(nbo, 1, ST_CR_ZAZ, 0);
    for (int it = 0; it < a.maxiter; ++it) {
      hipLaunchKernelGGL(k_cr_p, g, t, 0, st, a, (int)(it == 0), ps);  // p, Ap, y = M Ap; <Ap, y>
      fin(nb, 1, ST_CR_ALPHA, 0);
      hipLaunchKernelGGL(k_cr_upd, g, t, 0, st, a, ps);  // x, r, z; <r, r>
      fin(nb, 1, ST_RR, 0);
      op(z, az, nullptr, z, 0);  // Az = Op z; <z, Az>
      fin(nbo, 1, ST_CR_BETA, 0);
    }
  } else if (method == KR_GMRES) {
    const int k = a.restart;
    float* w = V(3);
    const int cycles = (a.maxiter + k - 1) / k;
    for (int c = 0; c < cycles; ++c) {
      fin(0, 0, ST_GM_BEGIN, 0);
      hipLaunchKernelGGL(k_gm_start, g, t, 0, st, a);  // V0 = r / beta, V1 = M V0
      for (int j = 0; j < k; ++j) {
        op(V(4 + j + 1), w, nullptr, nullptr, 0);  // w = Op M V_j
        hipLaunchKernelGGL(k_gm_dots, g, t, 0, st, a, j, ps);
        fin(nb, j + 1, ST_GM_H, j);
        hipLaunchKernelGGL(k_gm_orth, g, t, 0, st, a, j, ps);
        fin(nb, 1, ST_GM_NORM, j);
        hipLaunchKernelGGL(k_gm_next, g, t, 0, st, a, j);
      }
      fin(0, 0, ST_GM_SOLVE, 0);
      hipLaunchKernelGGL(k_gm_x, g, t, 0, st, a);  // x += M V y
      op(x, r, b, nullptr, 1);                      // r = b - Op x; <r, r>
      fin(nbo, 1, ST_RR, 0);
    }
  }
  hipLaunchKernelGGL(k_kr_finish, g, t, 0, st, a);
  return hipGetLastError();
}

}  // namespace fem
