// CPU restatement of the SolverFEM substep, used ONLY as test infrastructure:
// the parity checker for the HIP library and the timed CPU baseline.  Nothing
// in the shipped package links, loads or calls this file.
//
// Every function follows the reference source (nvidia-beer/newton-solver-fem,
// src/newton_solver_fem/_src/solvers/fem/...) loop by loop in the reference's
// thread order, templated on the real type: `float` mirrors the reference's
// fp32 Warp CPU path, `double` is the FP64 oracle used to state errors.
//
// Deliberate differences (documented where they occur):
//  * Dot products of the Krylov solve accumulate in double for both real types.
//  * The signed 3x3 SVD of the stretch clamp is computed in double (Jacobi
//    eigen-solve of F^T F + Gram-Schmidt of F V, sign carried by the smallest
//    singular value -- the convention of McAdams et al. 2011, which the
//    reference's wp.svd3 follows); the clamped matrix is then rounded to Real.
//
// Parity pinning of each piece is done by tests/test_oracle_*.py (finite
// differences of an independent energy, dense solves, closed forms, numpy SVD).
// Parity unpinned (Warp and Newton cannot be imported to compare with): the
// bitwise conventions of Warp's svd3 for degenerate F (only the clamp result
// is pinned, which does not depend on them for distinct or clamped-equal
// singular values); the exact iterate sequence of Warp's cg and its fp32 dot
// order; the host-check loop count for check_every > 0 (restated as rounding up
// to the next check); wp.inverse rounding; the eps32^(9/4) residual floor
// constant; the blocks prune_numerical_zeros would drop when an element block
// sums to exactly zero (every valid triplet is kept here).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace orc {

struct V3d {
  double x, y, z;
};

// ---------------------------------------------------------------------------
// Signed SVD and principal-stretch clamp (reference kernels.py:421-444).
// ---------------------------------------------------------------------------
static void jacobi_eigen_sym3(double a[3][3], double v[3][3], double w[3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    double diag = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off <= 1e-36 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- J^T A J
          double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < 3; ++i) w[i] = a[i][i];
}

// F = U diag(S) V^T with U, V rotations and S[2] carrying sign(det F).
void svd3_signed(const double F[9], double U[9], double S[3], double Vout[9]) {
  double c[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += F[k * 3 + i] * F[k * 3 + j];
      c[i][j] = s;
    }
  double v[3][3], w[3];
  jacobi_eigen_sym3(c, v, w);
  int order[3] = {0, 1, 2};
  std::sort(order, order + 3, [&](int a, int b) { return w[a] > w[b]; });
  double V[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = v[i][order[j]];
  double detv = V[0][0] * (V[1][1] * V[2][2] - V[1][2] * V[2][1]) - V[0][1] * (V[1][0] * V[2][2] - V[1][2] * V[2][0]) +
                V[0][2] * (V[1][0] * V[2][1] - V[1][1] * V[2][0]);
  if (detv < 0)
    for (int i = 0; i < 3; ++i) V[i][2] = -V[i][2];
  double b[3][3];  // columns b_j = F v_j
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) b[j][i] = F[i * 3 + 0] * V[0][j] + F[i * 3 + 1] * V[1][j] + F[i * 3 + 2] * V[2][j];
  double u[3][3];
  double n0 = std::sqrt(b[0][0] * b[0][0] + b[0][1] * b[0][1] + b[0][2] * b[0][2]);
  if (n0 > 1e-300) {
    for (int i = 0; i < 3; ++i) u[0][i] = b[0][i] / n0;
  } else {
    u[0][0] = 1; u[0][1] = 0; u[0][2] = 0;
  }
  double d = u[0][0] * b[1][0] + u[0][1] * b[1][1] + u[0][2] * b[1][2];
  double t[3] = {b[1][0] - d * u[0][0], b[1][1] - d * u[0][1], b[1][2] - d * u[0][2]};
  double n1 = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  if (n1 > 1e-300 * (1.0 + n0)) {
    for (int i = 0; i < 3; ++i) u[1][i] = t[i] / n1;
  } else {  // any unit vector orthogonal to u0
    double a0 = std::fabs(u[0][0]), a1 = std::fabs(u[0][1]), a2 = std::fabs(u[0][2]);
    double e[3] = {0, 0, 0};
    e[(a0 <= a1 && a0 <= a2) ? 0 : (a1 <= a2 ? 1 : 2)] = 1.0;
    double dd = u[0][0] * e[0] + u[0][1] * e[1] + u[0][2] * e[2];
    double r[3] = {e[0] - dd * u[0][0], e[1] - dd * u[0][1], e[2] - dd * u[0][2]};
    double nr = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int i = 0; i < 3; ++i) u[1][i] = r[i] / nr;
  }
  u[2][0] = u[0][1] * u[1][2] - u[0][2] * u[1][1];
  u[2][1] = u[0][2] * u[1][0] - u[0][0] * u[1][2];
  u[2][2] = u[0][0] * u[1][1] - u[0][1] * u[1][0];
  for (int j = 0; j < 3; ++j) S[j] = u[j][0] * b[j][0] + u[j][1] * b[j][1] + u[j][2] * b[j][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      U[i * 3 + j] = u[j][i];
      Vout[i * 3 + j] = V[i][j];
    }
}

template <typename Real>
void clamp_stretch(Real F[9], Real min_s, Real max_s) {
  if (min_s < Real(0) || max_s < Real(0)) return;
  double Fd[9], U[9], S[3], V[9];
  for (int i = 0; i < 9; ++i) Fd[i] = (double)F[i];
  svd3_signed(Fd, U, S, V);
  for (int k = 0; k < 3; ++k) S[k] = std::max(std::min(S[k], (double)max_s), (double)min_s);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += U[i * 3 + k] * S[k] * V[j * 3 + k];
      F[i * 3 + j] = (Real)s;
    }
}

// clamp_deformation_stretch_grad, kernels.py:447-481: adjoint of the
// principal-stretch clamp (identity at interior stretches, spectral
// derivative of U clamp(S) V^T otherwise).
inline void clamp_stretch_grad(const double F[9], double mn, double mx, const double adj[9], double out[9]) {
  if (mn < 0.0 || mx < 0.0) {
    for (int k = 0; k < 9; ++k) out[k] = adj[k];
    return;
  }
  double U[9], S[3], V[9];
  svd3_signed(F, U, S, V);
  if (std::min(S[0], std::min(S[1], S[2])) > mn && std::max(S[0], std::max(S[1], S[2])) < mx) {
    for (int k = 0; k < 9; ++k) out[k] = adj[k];
    return;
  }
  double clipped[3], H[9], R[9];
  for (int i = 0; i < 3; ++i) clipped[i] = std::max(std::min(S[i], mx), mn);
  for (int i = 0; i < 3; ++i)  // H = U^T adj V
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) s += U[3 * a + i] * adj[3 * a + b] * V[3 * b + j];
      H[3 * i + j] = s;
    }
  for (int k = 0; k < 9; ++k) R[k] = 0;
  for (int i = 0; i < 3; ++i) {
    double slope = (S[i] > mn && S[i] < mx) ? 1.0 : 0.0;
    R[3 * i + i] = slope * H[3 * i + i];
    for (int j = i + 1; j < 3; ++j) {
      double difference = S[i] - S[j];
      double divided = slope;
      if (std::fabs(difference) > 1.0e-6) divided = (clipped[i] - clipped[j]) / difference;
      double total = S[i] + S[j];
      double safe_total = std::max(total, 1.0e-6);
      if (total < 0.0) safe_total = std::min(total, -1.0e-6);
      double summed = (clipped[i] + clipped[j]) / safe_total;
      double symmetric = 0.5 * divided * (H[3 * i + j] + H[3 * j + i]);
      double skew = 0.5 * summed * (H[3 * i + j] - H[3 * j + i]);
      R[3 * i + j] = symmetric + skew;
      R[3 * j + i] = symmetric - skew;
    }
  }
  for (int i = 0; i < 3; ++i)  // out = U R V^T
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) s += U[3 * i + a] * R[3 * a + b] * V[3 * j + b];
      out[3 * i + j] = s;
    }
}

// ---------------------------------------------------------------------------
// Small fixed-size helpers.
// ---------------------------------------------------------------------------
template <typename Real>
struct V3 {
  Real x, y, z;
};
template <typename R>
inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename R>
inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename R>
inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename R>
inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R>
inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename R>
inline R det3(const R m[9]) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
}

// Scale-invariant 3x3 inverse with point-Jacobi fallback (kernels.py:1817-1833).
template <typename Real>
void inverse_diagonal_block(const Real B[9], Real out[9]) {
  Real scale = std::max(std::fabs(B[0]), std::max(std::fabs(B[4]), std::fabs(B[8])));
  if (scale > Real(0)) {
    Real n[9];
    for (int i = 0; i < 9; ++i) n[i] = B[i] / scale;
    Real trace = (n[0] + n[4] + n[8]) / Real(3);
    Real det = det3(n);
    if (trace > Real(0) && std::fabs(det) > Real(1e-9) * trace * trace * trace) {
      Real inv[9];
      inv[0] = (n[4] * n[8] - n[5] * n[7]) / det;
      inv[1] = (n[2] * n[7] - n[1] * n[8]) / det;
      inv[2] = (n[1] * n[5] - n[2] * n[4]) / det;
      inv[3] = (n[5] * n[6] - n[3] * n[8]) / det;
      inv[4] = (n[0] * n[8] - n[2] * n[6]) / det;
      inv[5] = (n[2] * n[3] - n[0] * n[5]) / det;
      inv[6] = (n[3] * n[7] - n[4] * n[6]) / det;
      inv[7] = (n[1] * n[6] - n[0] * n[7]) / det;
      inv[8] = (n[0] * n[4] - n[1] * n[3]) / det;
      for (int i = 0; i < 9; ++i) out[i] = inv[i] / scale;
      return;
    }
  }
  for (int i = 0; i < 9; ++i) out[i] = Real(0);
  for (int a = 0; a < 3; ++a) out[a * 4] = (B[a * 4] > Real(0)) ? Real(1) / B[a * 4] : Real(1);
}

// Hex8 shape-function gradient in the reference cell (kernels.py:1357-1419).
static const double HEX_GP = 0.5773502691896258;
inline double hex_sign_xi(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1.0 : -1.0; }
inline double hex_sign_eta(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1.0 : -1.0; }
inline double hex_sign_zeta(int a) { return (a >= 4) ? 1.0 : -1.0; }

template <typename Real>
V3<Real> hex_dN_dxi(int a, int g) {
  Real xi = (Real)(hex_sign_xi(g) * HEX_GP), eta = (Real)(hex_sign_eta(g) * HEX_GP),
       zeta = (Real)(hex_sign_zeta(g) * HEX_GP);
  Real xa = (Real)hex_sign_xi(a), ya = (Real)hex_sign_eta(a), za = (Real)hex_sign_zeta(a);
  Real one = 1, e = (Real)0.125;
  return {xa * (one + ya * eta) * (one + za * zeta) * e, (one + xa * xi) * ya * (one + za * zeta) * e,
          (one + xa * xi) * (one + ya * eta) * za * e};
}

// ---------------------------------------------------------------------------
// The solver restatement.
// ---------------------------------------------------------------------------
struct Params {
  double min_stretch = 0.05, max_stretch = 20.0;
  int has_ground = 0;
  double ground[4] = {0, 0, 1, 0};
  double ke = 1e5, kd = 1e2, kf = 1e3, mu = 0.5;
  double linear_damping = 0.0;
  double gravity[3] = {0, 0, -9.81};
  int enable_activations = 0;
  double tol = 1e-2;
  int maxiter = 50;
};

enum { TERM_TET = 1, TERM_HEX = 2, TERM_GROUND = 4, TERM_GRAVITY = 8, TERM_DAMPING = 16, TERM_SPRING = 32, TERM_TRI = 64 };

template <typename Real>
struct Solver {
  using v3 = V3<Real>;
  int P = 0, T = 0, H = 0, S = 0, R = 0;
  std::vector<v3> anchor;  // construction-time positions (FixedParticles.reference)
  std::vector<Real> mass, inv_mass, radius;
  std::vector<int> flags;
  std::vector<int> tet_idx;  // T*4
  std::vector<Real> tet_pose, tet_mat;  // T*9, T*3
  std::vector<int> hex_idx;  // H*8
  std::vector<Real> hex_invJ0, hex_detJ0w, hex_mat;  // H*72, H*8, H*3
  std::vector<int> fixed_idx, fixed_mask;
  std::vector<int> spring_idx;  // S*2
  std::vector<Real> spring_rest, spring_ke, spring_kd;
  std::vector<int> tri_idx;  // R*3
  std::vector<Real> tri_pose, tri_mat;  // R*4 (mat22 row-major), R*5 (mu, lambda, damp, drag, lift)
  Params prm;

  // Triplets [P diag | 2S springs | 3R tris | 16T tets | 64H hexes] (solver.py:312-321, 1008-1136).
  std::vector<int> trow, tcol;
  std::vector<Real> tval;  // 9 per triplet
  // BSR.
  std::vector<int> offsets, columns, slots;
  std::vector<Real> values, minv;
  bool have_slots = false;
  std::vector<v3> dv;
  int last_iters = 0;
  int nthreads = 1;  // host threads of the CG loops (1: serial, the order the tests pin)
  double last_rr = 0, last_atol_sq = 0, last_scale = 1;

  // ---- element kernels ---------------------------------------------------
  // eval_tetrahedra, kernels.py:511-632.
  void eval_tets(const v3* x, const v3* v, std::vector<v3>& f) const {
    for (int e = 0; e < T; ++e) {
      int i = tet_idx[4 * e], j = tet_idx[4 * e + 1], k = tet_idx[4 * e + 2], l = tet_idx[4 * e + 3];
      Real k_mu = tet_mat[3 * e], k_lambda = tet_mat[3 * e + 1], k_damp = tet_mat[3 * e + 2];
      if (k_mu == Real(0) && k_lambda == Real(0) && k_damp == Real(0)) continue;
      v3 x0 = x[i], x1 = x[j], x2 = x[k], x3 = x[l];
      v3 v0 = v[i], v1 = v[j], v2 = v[k], v3_ = v[l];
      v3 x10 = x1 - x0, x20 = x2 - x0, x30 = x3 - x0;
      v3 v10 = v1 - v0, v20 = v2 - v0, v30 = v3_ - v0;
      const Real* Dm = &tet_pose[9 * e];
      Real inv_rest_volume = det3(Dm) * Real(6);
      Real rest_volume = Real(1) / inv_rest_volume;
      Real alpha = Real(1) + k_mu / k_lambda - k_mu / (Real(4) * k_lambda);
      k_mu *= rest_volume;
      k_lambda *= rest_volume;
      k_damp *= rest_volume;
      Real F[9], dF[9];
      for (int r = 0; r < 3; ++r) {
        Real ds[3] = {r == 0 ? x10.x : (r == 1 ? x10.y : x10.z), r == 0 ? x20.x : (r == 1 ? x20.y : x20.z),
                      r == 0 ? x30.x : (r == 1 ? x30.y : x30.z)};
        Real vs[3] = {r == 0 ? v10.x : (r == 1 ? v10.y : v10.z), r == 0 ? v20.x : (r == 1 ? v20.y : v20.z),
                      r == 0 ? v30.x : (r == 1 ? v30.y : v30.z)};
        for (int c = 0; c < 3; ++c) {
          F[3 * r + c] = ds[0] * Dm[c] + ds[1] * Dm[3 + c] + ds[2] * Dm[6 + c];
          dF[3 * r + c] = vs[0] * Dm[c] + vs[1] * Dm[3 + c] + vs[2] * Dm[6 + c];
        }
      }
      clamp_stretch(F, (Real)prm.min_stretch, (Real)prm.max_stretch);
      Real Ic = 0;
      for (int q = 0; q < 9; ++q) Ic += F[q] * F[q];
      Real Pm[9], Hm[9];
      Real sdev = k_mu * (Real(1) - Real(1) / (Ic + Real(1)));
      for (int q = 0; q < 9; ++q) Pm[q] = F[q] * sdev + dF[q] * k_damp;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Hm[3 * r + c] = Pm[3 * r] * Dm[3 * c] + Pm[3 * r + 1] * Dm[3 * c + 1] + Pm[3 * r + 2] * Dm[3 * c + 2];
      v3 f1 = {Hm[0], Hm[3], Hm[6]}, f2 = {Hm[1], Hm[4], Hm[7]}, f3 = {Hm[2], Hm[5], Hm[8]};
      Real J = std::max(det3(F), Real(0.01));
      Real s = inv_rest_volume / Real(6);
      v3 dJ1 = cross(x20, x30) * s, dJ2 = cross(x30, x10) * s, dJ3 = cross(x10, x20) * s;
      Real f_volume = (J - alpha) * k_lambda;
      Real f_damp = (dot(dJ1, v10) + dot(dJ2, v20) + dot(dJ3, v30)) * k_damp;
      Real ft = f_volume + f_damp;
      f1 = f1 + dJ1 * ft;
      f2 = f2 + dJ2 * ft;
      f3 = f3 + dJ3 * ft;
      v3 f0 = (f1 + f2 + f3) * Real(-1);
      f[i] = f[i] - f0;
      f[j] = f[j] - f1;
      f[k] = f[k] - f2;
      f[l] = f[l] - f3;
    }
  }

  // eval_tet_activations, kernels.py:484-508.
  void eval_tet_activations(const v3* x, const Real* act, std::vector<v3>& f) const {
    for (int e = 0; e < T; ++e) {
      int a = tet_idx[4 * e], b = tet_idx[4 * e + 1], c = tet_idx[4 * e + 2], d = tet_idx[4 * e + 3];
      Real scale = tet_mat[3 * e + 1] * act[e] / Real(6);
      v3 fb = cross(x[c] - x[a], x[d] - x[a]) * scale;
      v3 fc = cross(x[d] - x[a], x[b] - x[a]) * scale;
      v3 fd = cross(x[b] - x[a], x[c] - x[a]) * scale;
      f[a] = f[a] + (fb + fc + fd);
      f[b] = f[b] - fb;
      f[c] = f[c] - fc;
      f[d] = f[d] - fd;
    }
  }

  // build_system_matrix_tet_kernel, kernels.py:635-850 (block_offset = P).
  void tet_blocks(const v3* x, Real dt) {  // each element writes only its own 16 triplets
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
    for (int e = 0; e < T; ++e) {
      int n[4] = {tet_idx[4 * e], tet_idx[4 * e + 1], tet_idx[4 * e + 2], tet_idx[4 * e + 3]};
      Real k_mu = tet_mat[3 * e], k_lambda = tet_mat[3 * e + 1];
      size_t base = tet_off() + 16 * (size_t)e;
      if (k_mu == Real(0) && k_lambda == Real(0) && tet_mat[3 * e + 2] == Real(0)) {
        for (int q = 0; q < 16; ++q) {
          trow[base + q] = -1;
          tcol[base + q] = -1;
          std::fill(&tval[9 * (base + q)], &tval[9 * (base + q)] + 9, Real(0));
        }
        continue;
      }
      v3 x0 = x[n[0]], x1 = x[n[1]], x2 = x[n[2]], x3 = x[n[3]];
      v3 x10 = x1 - x0, x20 = x2 - x0, x30 = x3 - x0;
      const Real* Dm = &tet_pose[9 * e];
      Real det_Dm = det3(Dm);
      Real inv_rest_volume = det_Dm * Real(6);
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          trow[base + 4 * a + b] = n[a];
          tcol[base + 4 * a + b] = n[b];
        }
      if (inv_rest_volume <= Real(0)) {
        std::fill(&tval[9 * base], &tval[9 * base] + 9 * 16, Real(0));
        continue;
      }
      Real rest_volume = Real(1) / inv_rest_volume;
      k_mu *= rest_volume;
      k_lambda *= rest_volume;
      Real F[9];
      for (int r = 0; r < 3; ++r) {
        Real ds[3] = {r == 0 ? x10.x : (r == 1 ? x10.y : x10.z), r == 0 ? x20.x : (r == 1 ? x20.y : x20.z),
                      r == 0 ? x30.x : (r == 1 ? x30.y : x30.z)};
        for (int c = 0; c < 3; ++c) F[3 * r + c] = ds[0] * Dm[c] + ds[1] * Dm[3 + c] + ds[2] * Dm[6 + c];
      }
      clamp_stretch(F, (Real)prm.min_stretch, (Real)prm.max_stretch);
      Real Ic = 0;
      for (int q = 0; q < 9; ++q) Ic += F[q] * F[q];
      Real Ic1 = std::max(Ic + Real(1), Real(1e-6));
      Real s = Real(1) - Real(1) / Ic1;
      Real coef = k_mu * Real(2) / (Ic1 * Ic1);
      Real s_vol = inv_rest_volume / Real(6);
      v3 dJ[4];
      dJ[1] = cross(x20, x30) * s_vol;
      dJ[2] = cross(x30, x10) * s_vol;
      dJ[3] = cross(x10, x20) * s_vol;
      dJ[0] = (dJ[1] + dJ[2] + dJ[3]) * Real(-1);
      v3 g[4];
      g[1] = {Dm[0], Dm[1], Dm[2]};
      g[2] = {Dm[3], Dm[4], Dm[5]};
      g[3] = {Dm[6], Dm[7], Dm[8]};
      g[0] = (g[1] + g[2] + g[3]) * Real(-1);
      v3 Fg[4];
      for (int a = 0; a < 4; ++a)
        Fg[a] = {F[0] * g[a].x + F[1] * g[a].y + F[2] * g[a].z, F[3] * g[a].x + F[4] * g[a].y + F[5] * g[a].z,
                 F[6] * g[a].x + F[7] * g[a].y + F[8] * g[a].z};
      Real scale = dt * dt;
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
          Real dab = dot(g[a], g[b]);
          Real fa[3] = {Fg[a].x, Fg[a].y, Fg[a].z}, fb[3] = {Fg[b].x, Fg[b].y, Fg[b].z};
          Real ja[3] = {dJ[a].x, dJ[a].y, dJ[a].z}, jb[3] = {dJ[b].x, dJ[b].y, dJ[b].z};
          Real* out = &tval[9 * (base + 4 * a + b)];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
              Real kd = ((r == c) ? k_mu * s * dab : Real(0)) + fa[r] * fb[c] * coef;
              Real kv = ja[r] * jb[c] * k_lambda;
              out[3 * r + c] = (kd + kv) * scale;
            }
        }
    }
  }

  // Deformation gradient at one Hex8 Gauss point and its shape gradients.
  void hex_gp(int e, int g, const v3* x, v3 dN[8], Real F[9]) const {
    const Real* J = &hex_invJ0[72 * e + 9 * g];
    for (int a = 0; a < 8; ++a) {
      v3 d = hex_dN_dxi<Real>(a, g);  // inv_J0^T * d
      dN[a] = {J[0] * d.x + J[3] * d.y + J[6] * d.z, J[1] * d.x + J[4] * d.y + J[7] * d.z,
               J[2] * d.x + J[5] * d.y + J[8] * d.z};
    }
    for (int q = 0; q < 9; ++q) F[q] = 0;
    for (int a = 0; a < 8; ++a) {
      v3 xa = x[hex_idx[8 * e + a]];
      Real xs[3] = {xa.x, xa.y, xa.z}, ds[3] = {dN[a].x, dN[a].y, dN[a].z};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) F[3 * r + c] += xs[r] * ds[c];
    }
  }

  // eval_hexahedra, kernels.py:1481-1601.
  void eval_hexes(const v3* x, const v3* v, std::vector<v3>& f) const {
    for (int e = 0; e < H; ++e) {
      Real k_mu = hex_mat[3 * e], k_lambda = hex_mat[3 * e + 1], k_damp = hex_mat[3 * e + 2];
      Real alpha = Real(1) + k_mu / k_lambda - k_mu / (Real(4) * k_lambda);
      for (int g = 0; g < 8; ++g) {
        v3 dN[8];
        Real F[9], dF[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        hex_gp(e, g, x, dN, F);
        clamp_stretch(F, (Real)prm.min_stretch, (Real)prm.max_stretch);
        for (int a = 0; a < 8; ++a) {
          v3 va = v[hex_idx[8 * e + a]];
          Real vs[3] = {va.x, va.y, va.z}, ds[3] = {dN[a].x, dN[a].y, dN[a].z};
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) dF[3 * r + c] += vs[r] * ds[c];
        }
        v3 c0 = {F[0], F[3], F[6]}, c1 = {F[1], F[4], F[7]}, c2 = {F[2], F[5], F[8]};
        Real Ic = dot(c0, c0) + dot(c1, c1) + dot(c2, c2);
        Real Pm[9];
        Real sdev = k_mu * (Real(1) - Real(1) / (Ic + Real(1)));
        for (int q = 0; q < 9; ++q) Pm[q] = F[q] * sdev + dF[q] * k_damp;
        Real J = std::max(det3(F), Real(0.01));
        v3 cc0 = cross(c1, c2), cc1 = cross(c2, c0), cc2 = cross(c0, c1);
        Real cof[9] = {cc0.x, cc1.x, cc2.x, cc0.y, cc1.y, cc2.y, cc0.z, cc1.z, cc2.z};
        Real vol = k_lambda * (J - alpha);
        for (int q = 0; q < 9; ++q) Pm[q] = Pm[q] + cof[q] * vol;
        Real w = hex_detJ0w[8 * e + g];
        for (int a = 0; a < 8; ++a) {
          v3 pa = {Pm[0] * dN[a].x + Pm[1] * dN[a].y + Pm[2] * dN[a].z,
                   Pm[3] * dN[a].x + Pm[4] * dN[a].y + Pm[5] * dN[a].z,
                   Pm[6] * dN[a].x + Pm[7] * dN[a].y + Pm[8] * dN[a].z};
          int n = hex_idx[8 * e + a];
          f[n] = f[n] - pa * w;
        }
      }
    }
  }

  // build_system_matrix_hex_kernel, kernels.py:1609-1747 (block_offset = P+16T).
  void hex_blocks(const v3* x, Real dt) {
    // each hex writes only its own 64 triplets
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
    for (int e = 0; e < H; ++e) {
      Real k_mu = hex_mat[3 * e], k_lambda = hex_mat[3 * e + 1];
      size_t base = tet_off() + 16 * (size_t)T + 64 * (size_t)e;
      Real K[64][9];
      for (int q = 0; q < 64; ++q)
        for (int r = 0; r < 9; ++r) K[q][r] = 0;
      for (int g = 0; g < 8; ++g) {
        v3 dN[8];
        Real F[9];
        hex_gp(e, g, x, dN, F);
        clamp_stretch(F, (Real)prm.min_stretch, (Real)prm.max_stretch);
        v3 c0 = {F[0], F[3], F[6]}, c1 = {F[1], F[4], F[7]}, c2 = {F[2], F[5], F[8]};
        Real Ic = dot(c0, c0) + dot(c1, c1) + dot(c2, c2);
        Real Ic1 = std::max(Ic + Real(1), Real(1e-6));
        Real s = Real(1) - Real(1) / Ic1;
        Real coef = k_mu * Real(2) / (Ic1 * Ic1);
        v3 cc0 = cross(c1, c2), cc1 = cross(c2, c0), cc2 = cross(c0, c1);
        Real cof[9] = {cc0.x, cc1.x, cc2.x, cc0.y, cc1.y, cc2.y, cc0.z, cc1.z, cc2.z};
        Real w = hex_detJ0w[8 * e + g];
        Real Fd[8][3], Jd[8][3];
        for (int a = 0; a < 8; ++a)
          for (int r = 0; r < 3; ++r) {
            Fd[a][r] = F[3 * r] * dN[a].x + F[3 * r + 1] * dN[a].y + F[3 * r + 2] * dN[a].z;
            Jd[a][r] = cof[3 * r] * dN[a].x + cof[3 * r + 1] * dN[a].y + cof[3 * r + 2] * dN[a].z;
          }
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) {
            Real dab = dot(dN[a], dN[b]);
            for (int r = 0; r < 3; ++r)
              for (int c = 0; c < 3; ++c) {
                Real kdev = ((r == c) ? k_mu * s * dab : Real(0)) + Fd[a][r] * Fd[b][c] * coef;
                Real kvol = Jd[a][r] * Jd[b][c] * k_lambda;
                K[8 * a + b][3 * r + c] = K[8 * a + b][3 * r + c] + (kdev + kvol) * w;
              }
          }
      }
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) {
          trow[base + 8 * a + b] = hex_idx[8 * e + a];
          tcol[base + 8 * a + b] = hex_idx[8 * e + b];
          for (int r = 0; r < 9; ++r) tval[9 * (base + 8 * a + b) + r] = K[8 * a + b][r] * (dt * dt);
        }
    }
  }

  // Ground penalty + Coulomb friction, kernels.py:1132-1195.
  void eval_ground(const v3* x, const v3* v, std::vector<v3>& f) const {
    if (!prm.has_ground) return;
    v3 n = {(Real)prm.ground[0], (Real)prm.ground[1], (Real)prm.ground[2]};
    Real gx = (Real)prm.gravity[0], gy = (Real)prm.gravity[1], gz = (Real)prm.gravity[2];
    Real gmag = std::sqrt(gx * gx + gy * gy + gz * gz);
    for (int i = 0; i < P; ++i) {
      if ((flags[i] & 1) == 0) continue;
      Real c = std::min(dot(n, x[i]) + (Real)prm.ground[3] - radius[i], Real(0));
      Real vn = dot(n, v[i]);
      Real jn = c * (Real)prm.ke;
      if (c >= Real(0)) continue;
      Real jd = std::min(vn, Real(0)) * (Real)prm.kd;
      Real fn = jn + jd;
      v3 vt = v[i] - n * vn;
      Real vs = std::sqrt(dot(vt, vt));
      if (vs > Real(0)) vt = vt * (Real(1) / vs);
      Real inv_m = inv_mass[i];
      Real inv_m_safe = std::max(inv_m, Real(1e-9));
      Real m = inv_m > Real(0) ? Real(1) / inv_m_safe : Real(0);
      Real n_cap = m * gmag;
      Real fn_eff = std::min(std::fabs(fn), n_cap);
      Real ft = std::min(vs * (Real)prm.kf, (Real)prm.mu * fn_eff);
      f[i] = f[i] - n * fn - vt * ft;
    }
  }

  // eval_gravity_from_array kernels.py:1273-1286, eval_linear_damping_kernel :1289-1307.
  void eval_gravity(std::vector<v3>& f) const {
    for (int i = 0; i < P; ++i) {
      if ((flags[i] & 1) == 0) continue;
      Real m = mass[i];
      f[i] = {(Real)prm.gravity[0] * m, (Real)prm.gravity[1] * m, (Real)prm.gravity[2] * m};
    }
  }
  void eval_damping(const v3* v, std::vector<v3>& f) const {
    if (!(prm.linear_damping > 0.0)) return;
    for (int i = 0; i < P; ++i) {
      if ((flags[i] & 1) == 0) continue;
      f[i] = v[i] * (-(Real)prm.linear_damping * mass[i]);
    }
  }

  // ---- springs (kernels.py:176-219) and triangle membranes (:876-970) -----
  // eval_springs: f_i -= fs, f_j += fs, fs = dir (ke c + kd dc/dt).
  void eval_springs(const v3* x, const v3* v, std::vector<v3>& f) const {
    for (int s = 0; s < S; ++s) {
      int i = spring_idx[2 * s], j = spring_idx[2 * s + 1];
      if (i == -1 || j == -1) continue;
      v3 xij = x[i] - x[j], vij = v[i] - v[j];
      Real len = std::sqrt(dot(xij, xij));
      if (len < Real(1e-6)) continue;
      v3 dir = xij * (Real(1) / len);
      Real c = len - spring_rest[s];
      Real dcdt = dot(dir, vij);
      v3 fs = dir * (spring_ke[s] * c + spring_kd[s] * dcdt);
      f[i] = f[i] - fs;
      f[j] = f[j] + fs;
    }
  }
  static v3 normalized(v3 a) {
    Real l = std::sqrt(dot(a, a));
    return l > Real(0) ? a * (Real(1) / l) : v3{0, 0, 0};
  }
  // eval_triangles: deviatoric + area-preservation membrane, damping, drag and lift.
  void eval_tris(const v3* x, const v3* v, std::vector<v3>& f) const {
    for (int t = 0; t < R; ++t) {
      Real k_mu = tri_mat[5 * t], k_lambda = tri_mat[5 * t + 1], k_damp = tri_mat[5 * t + 2];
      Real k_drag = tri_mat[5 * t + 3], k_lift = tri_mat[5 * t + 4];
      int i = tri_idx[3 * t], j = tri_idx[3 * t + 1], k = tri_idx[3 * t + 2];
      v3 x10 = x[j] - x[i], x20 = x[k] - x[i];
      v3 v10 = v[j] - v[i], v20 = v[k] - v[i];
      const Real* Dm = &tri_pose[4 * t];  // [[Dm00, Dm01], [Dm10, Dm11]]
      Real inv_rest_area = (Dm[0] * Dm[3] - Dm[1] * Dm[2]) * Real(2);
      Real rest_area = Real(1) / inv_rest_area;
      k_mu *= rest_area;
      k_lambda *= rest_area;
      k_damp *= rest_area;
      v3 F1 = x10 * Dm[0] + x20 * Dm[2], F2 = x10 * Dm[1] + x20 * Dm[3];
      v3 dF1 = v10 * Dm[0] + v20 * Dm[2], dF2 = v10 * Dm[1] + v20 * Dm[3];
      Real Ic = dot(F1, F1) + dot(F2, F2);
      Real dev = (Ic - Real(2)) / Ic;
      v3 P1 = F1 * (k_mu * dev) + dF1 * k_damp, P2 = F2 * (k_mu * dev) + dF2 * k_damp;
      v3 f1 = P1 * Dm[0] + P2 * Dm[1], f2 = P1 * Dm[2] + P2 * Dm[3];
      v3 n = cross(x10, x20);
      Real area = std::sqrt(dot(n, n)) * Real(0.5);
      Real c = area * inv_rest_area - Real(1);
      n = normalized(n);
      v3 dcdq = cross(x20, n) * (inv_rest_area * Real(0.5));
      v3 dcdr = cross(n, x10) * (inv_rest_area * Real(0.5));
      Real f_area = k_lambda * c;
      Real dcdt = dot(dcdq, v[j]) + dot(dcdr, v[k]) - dot(dcdq + dcdr, v[i]);
      Real f_damp = k_damp * dcdt;
      f1 = f1 + dcdq * (f_area + f_damp);
      f2 = f2 + dcdr * (f_area + f_damp);
      v3 f0 = f1 + f2;
      v3 vmid = (v[i] + v[j] + v[k]) * Real(0.3333);
      v3 vdir = normalized(vmid);
      v3 f_drag = vmid * (k_drag * area * std::fabs(dot(n, vmid)));
      Real cosang = std::min(std::max(dot(n, vdir), Real(-1)), Real(1));
      v3 f_lift = n * (k_lift * area * (Real(1.5707963267948966) - std::acos(cosang))) * dot(vmid, vmid);
      f0 = f0 - f_drag - f_lift;
      f1 = f1 + f_drag + f_lift;
      f2 = f2 + f_drag + f_lift;
      f[i] = f[i] + f0;
      f[j] = f[j] - f1;
      f[k] = f[k] - f2;
    }
  }

  // ---- assembly ----------------------------------------------------------
  size_t tet_off() const { return (size_t)P + 2 * (size_t)S + 3 * (size_t)R; }

  void write_triplets(const v3* x, Real dt) {
    size_t n = tet_off() + 16 * (size_t)T + 64 * (size_t)H;
    trow.assign(n, 0);
    tcol.assign(n, 0);
    tval.assign(9 * n, Real(0));
    const Real dt2 = dt * dt;
    for (int i = 0; i < P; ++i) {
      // build_system_matrix_diagonal_mass_kernel kernels.py:141-153, or with
      // springs build_system_matrix_diagonal_kernel :294-323 (O(P S) scan)
      Real diag = mass[i];
      for (int s = 0; s < S; ++s)
        if (spring_idx[2 * s] == i || spring_idx[2 * s + 1] == i) diag = diag + dt * spring_kd[s] + dt2 * spring_ke[s];
      trow[i] = i;
      tcol[i] = i;
      tval[9 * i] = tval[9 * i + 4] = tval[9 * i + 8] = diag;
    }
    for (int s = 0; s < S; ++s) {  // build_system_matrix_sparse_kernel kernels.py:326-366
      size_t b = (size_t)P + 2 * (size_t)s;
      int i = spring_idx[2 * s], j = spring_idx[2 * s + 1];
      Real neg = -(dt * spring_kd[s] + dt * dt * spring_ke[s]);
      trow[b] = i;
      tcol[b] = j;
      trow[b + 1] = j;
      tcol[b + 1] = i;
      for (int q = 0; q < 2; ++q) tval[9 * (b + q)] = tval[9 * (b + q) + 4] = tval[9 * (b + q) + 8] = neg;
    }
    for (int t = 0; t < R; ++t) {  // build_system_matrix_tri_kernel kernels.py:972-1030
      size_t b = (size_t)P + 2 * (size_t)S + 3 * (size_t)t;
      const Real* Dm = &tri_pose[4 * t];
      Real inv_rest_area = (Dm[0] * Dm[3] - Dm[1] * Dm[2]) * Real(2);
      Real val = 0;
      if (inv_rest_area > Real(0)) {
        Real rest_area = Real(1) / inv_rest_area;
        val = dt * dt * (tri_mat[5 * t] * rest_area + tri_mat[5 * t + 1] * rest_area) / Real(3);
      }
      for (int q = 0; q < 3; ++q) {
        trow[b + q] = tcol[b + q] = tri_idx[3 * t + q];
        tval[9 * (b + q)] = tval[9 * (b + q) + 4] = tval[9 * (b + q) + 8] = val;
      }
    }
    tet_blocks(x, dt);
    hex_blocks(x, dt);
  }

  // bsr_set_from_triplets: stable sort by (row, col), sum duplicates in triplet
  // order, drop negative indices and optionally all-zero input blocks.
  void set_from_triplets(bool prune) {
    size_t n = trow.size();
    std::vector<uint32_t> keep;
    keep.reserve(n);
    for (size_t t = 0; t < n; ++t) {
      if (trow[t] < 0 || tcol[t] < 0) continue;
      if (prune) {
        bool zero = true;
        for (int r = 0; r < 9; ++r) zero = zero && tval[9 * t + r] == Real(0);
        if (zero) continue;
      }
      keep.push_back((uint32_t)t);
    }
    std::stable_sort(keep.begin(), keep.end(), [&](uint32_t a, uint32_t b) {
      return trow[a] != trow[b] ? trow[a] < trow[b] : tcol[a] < tcol[b];
    });
    offsets.assign(P + 1, 0);
    columns.clear();
    values.clear();
    int prev_r = -1, prev_c = -1;
    for (uint32_t t : keep) {
      if (trow[t] != prev_r || tcol[t] != prev_c) {
        columns.push_back(tcol[t]);
        values.insert(values.end(), 9, Real(0));
        offsets[trow[t] + 1]++;
        prev_r = trow[t];
        prev_c = tcol[t];
      }
      Real* dst = &values[values.size() - 9];
      for (int r = 0; r < 9; ++r) dst[r] += tval[9 * t + r];
    }
    for (int i = 0; i < P; ++i) offsets[i + 1] += offsets[i];
  }

  // map_triplet_blocks sparse_update.py:9-30.
  void map_slots() {
    size_t n = trow.size();
    slots.assign(n, -1);
    for (size_t t = 0; t < n; ++t) {
      int r = trow[t], c = tcol[t];
      if (r < 0 || c < 0) continue;
      int first = offsets[r], last = offsets[r + 1];
      while (first < last) {
        int mid = (first + last) / 2;
        if (columns[mid] < c) first = mid + 1;
        else last = mid;
      }
      slots[t] = first;
    }
  }

  // accumulate_triplet_blocks sparse_update.py:33-59 (serial launch order).
  void accumulate_slots() {
    std::fill(values.begin(), values.end(), Real(0));
    for (size_t t = 0; t < trow.size(); ++t) {
      int s = slots[t];
      if (s < 0) continue;
      bool zero = true;
      for (int r = 0; r < 9; ++r) zero = zero && tval[9 * t + r] == Real(0);
      if (zero) continue;
      for (int r = 0; r < 9; ++r) values[9 * (size_t)s + r] += tval[9 * t + r];
    }
  }

  // invert_block_diagonal_kernel kernels.py:1836-1849.
  void refresh_preconditioner() {
    minv.assign(9 * (size_t)P, Real(0));
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
    for (int i = 0; i < P; ++i) {
      Real d[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
      for (int k = offsets[i]; k < offsets[i + 1]; ++k)
        if (columns[k] == i)
          for (int r = 0; r < 9; ++r) d[r] = values[9 * (size_t)k + r];
      inverse_diagonal_block(d, &minv[9 * (size_t)i]);
    }
  }

  // Construction (solver.py:318-341): rest-pose tangent, pruned triplets.
  void construct(Real dt) {
    dv.assign(P, v3{0, 0, 0});
    q_asm = anchor;
    write_triplets(anchor.data(), dt);
    set_from_triplets(true);
    refresh_preconditioner();
  }

  // refresh_tangent solver.py:769-857 (fixed positions restored first).
  void refresh(v3* q, Real dt) {
    for (int i : fixed_idx) q[i] = anchor[i];
    q_asm.assign(q, q + P);
    write_triplets(q, dt);
    if (!have_slots) {
      set_from_triplets(false);
      map_slots();
      have_slots = true;
    } else {
      accumulate_slots();
    }
    refresh_preconditioner();
  }

  // ---- sparse operators --------------------------------------------------
  void bsr_mv(const v3* x, v3* y) const {
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
    for (int i = 0; i < P; ++i) {
      Real acc[3] = {0, 0, 0};
      for (int k = offsets[i]; k < offsets[i + 1]; ++k) {
        const Real* b = &values[9 * (size_t)k];
        v3 xv = x[columns[k]];
        for (int r = 0; r < 3; ++r) acc[r] += b[3 * r] * xv.x + b[3 * r + 1] * xv.y + b[3 * r + 2] * xv.z;
      }
      y[i] = {acc[0], acc[1], acc[2]};
    }
  }
  // Operator of the CG solve: A, or PAP+I-P with fixed nodes (fixed_particles.py:29-53).
  void op_mv(const v3* x, v3* y) const {
    if (fixed_idx.empty()) {
      bsr_mv(x, y);
      return;
    }
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
    for (int i = 0; i < P; ++i) {
      if (fixed_mask[i]) {
        y[i] = x[i];
        continue;
      }
      Real acc[3] = {0, 0, 0};
      for (int k = offsets[i]; k < offsets[i + 1]; ++k) {
        int c = columns[k];
        if (fixed_mask[c]) continue;
        const Real* b = &values[9 * (size_t)k];
        v3 xv = x[c];
        for (int r = 0; r < 3; ++r) acc[r] += b[3 * r] * xv.x + b[3 * r + 1] * xv.y + b[3 * r + 2] * xv.z;
      }
      y[i] = {acc[0], acc[1], acc[2]};
    }
  }
  void precond(const v3* r, v3* z) const {
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
    for (int i = 0; i < P; ++i) {
      const Real* m = &minv[9 * (size_t)i];
      z[i] = {m[0] * r[i].x + m[1] * r[i].y + m[2] * r[i].z, m[3] * r[i].x + m[4] * r[i].y + m[5] * r[i].z,
              m[6] * r[i].x + m[7] * r[i].y + m[8] * r[i].z};
    }
  }
  double vdot(const std::vector<v3>& a, const std::vector<v3>& b) const {
    double s = 0;
    const long n = (long)a.size();
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static) reduction(+ : s)
    for (long i = 0; i < n; ++i)
      s += (double)a[i].x * b[i].x + (double)a[i].y * b[i].y + (double)a[i].z * b[i].z;
    return s;
  }

  // Warp cg with tol (relative), atol=0, fixed iteration budget and the
  // residual-masked alpha/beta updates (equivalent to early exit).
  void cg(const std::vector<v3>& b, std::vector<v3>& x) {
    const double eps32 = 1.1920928955078125e-07;
    double bb = vdot(b, b);
    double atol = std::max(0.0, prm.tol * std::sqrt(bb));
    atol = std::max(atol, std::pow(eps32, 9.0 / 4.0));
    double atol_sq = atol * atol;
    std::vector<v3> r(P), z(P), p(P), Ap(P);
    op_mv(x.data(), Ap.data());
    for (int i = 0; i < P; ++i) r[i] = b[i] - Ap[i];
    double rr = vdot(r, r);
    precond(r.data(), z.data());
    double rz = vdot(r, z);
    p = z;
    int iters = 0;
    for (int it = 0; it < prm.maxiter; ++it) {
      if (!(rr > atol_sq)) break;
      op_mv(p.data(), Ap.data());
      double pAp = vdot(p, Ap);
      Real alpha = pAp != 0.0 ? (Real)(rz / pAp) : Real(0);
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
      for (int i = 0; i < P; ++i) {
        x[i] = x[i] + p[i] * alpha;
        r[i] = r[i] - Ap[i] * alpha;
      }
      rr = vdot(r, r);
      precond(r.data(), z.data());
      double rz_new = vdot(r, z);
      Real beta = rr > atol_sq ? (Real)(rz_new / rz) : Real(0);
      rz = rz_new;
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
      for (int i = 0; i < P; ++i) p[i] = z[i] + p[i] * beta;
      iters = it + 1;
    }
    last_iters = iters;
    last_rr = rr;
    last_atol_sq = atol_sq;
  }

  // SolverFEM.step fast path (solver.py:942-1002, 1254-1351).
  void step(v3* q_in, v3* qd_in, v3* f_out, const v3* ext, const Real* act, v3* q_out, v3* qd_out, Real dt) {
    for (int i : fixed_idx) {  // FixedParticles.restore(state_in)
      q_in[i] = anchor[i];
      qd_in[i] = v3{0, 0, 0};
    }
    forces_scaled(q_in, qd_in, ext, act, dt, f_out);
    // subtract_elastic_predictor implicit_predictor.py:9-23
    if (T || H || S) {
#pragma omp parallel for num_threads(nthreads) if (nthreads > 1) schedule(static)
      for (int i = 0; i < P; ++i) {
        Real c[3] = {0, 0, 0};
        for (int k = offsets[i]; k < offsets[i + 1]; ++k) {
          const Real* b = &values[9 * (size_t)k];
          v3 d = qd_in[columns[k]] - qd_in[i];
          for (int r = 0; r < 3; ++r) c[r] += b[3 * r] * d.x + b[3 * r + 1] * d.y + b[3 * r + 2] * d.z;
        }
        f_out[i] = f_out[i] - v3{c[0], c[1], c[2]};
      }
      // restore_spring_damping_predictor implicit_predictor.py:26-41
      for (int s = 0; s < S; ++s) {
        int a = spring_idx[2 * s], b = spring_idx[2 * s + 1];
        if (a >= 0 && b >= 0) {
          v3 corr = (qd_in[a] - qd_in[b]) * (dt * spring_kd[s]);
          f_out[a] = f_out[a] + corr;
          f_out[b] = f_out[b] - corr;
        }
      }
    }
    for (int i : fixed_idx) f_out[i] = v3{0, 0, 0};  // project_rhs
    // maximum_rhs_component / normalize_linear_rhs kernels.py:1779-1805.
    Real maximum = 0;
    for (int i = 0; i < P; ++i)
      maximum = std::max(maximum, std::max(std::fabs(f_out[i].x), std::max(std::fabs(f_out[i].y), std::fabs(f_out[i].z))));
    Real scale = 1;
    if (maximum > Real(0) && maximum < Real(1e-12)) scale = std::min(Real(1) / maximum, Real(1e30));
    std::vector<v3> b(P);
    for (int i = 0; i < P; ++i) {
      b[i] = f_out[i] * scale;
      dv[i] = (maximum == Real(0)) ? v3{0, 0, 0} : dv[i] * scale;
    }
    cg(b, dv);
    last_scale = scale;
    for (int i = 0; i < P; ++i) dv[i] = dv[i] * (Real(1) / scale);  // restore_linear_increment (x /= scale)
    for (int i = 0; i < P; ++i) {  // update_state kernels.py:115-128
      v3 vel = qd_in[i] + dv[i];
      q_out[i] = q_in[i] + vel * dt;
      qd_out[i] = vel;
    }
    for (int i : fixed_idx) {
      q_out[i] = anchor[i];
      qd_out[i] = v3{0, 0, 0};
    }
  }
  // _accumulate_forces solver.py:542-589 + external add :993-999:
  // particle_f = sum_term h*term in the reference's term order.
  void forces_scaled(const v3* q, const v3* qd, const v3* ext, const Real* act, Real dt, v3* out) {
    std::vector<v3> spr(P, v3{0, 0, 0}), tet(P, v3{0, 0, 0}), hex(P, v3{0, 0, 0}), tri(P, v3{0, 0, 0}),
        gr(P, v3{0, 0, 0}), grav(P, v3{0, 0, 0}), damp(P, v3{0, 0, 0});
    if (S) eval_springs(q, qd, spr);
    if (T) {
      eval_tets(q, qd, tet);
      if (prm.enable_activations && act) eval_tet_activations(q, act, tet);
    }
    if (H) eval_hexes(q, qd, hex);
    if (R) eval_tris(q, qd, tri);
    eval_ground(q, qd, gr);
    eval_gravity(grav);
    eval_damping(qd, damp);
    for (int i = 0; i < P; ++i) {
      v3 f = {0, 0, 0};
      f = f + spr[i] * dt;
      f = f + tet[i] * dt;
      f = f + hex[i] * dt;
      f = f + tri[i] * dt;
      f = f + gr[i] * dt;
      f = f + grav[i] * dt;
      f = f + damp[i] * dt;
      if (ext) f = f + ext[i] * dt;
      out[i] = f;
    }
  }

  // ---- differentiation (differentiation.py:289-394) ----------------------
  // The matrix of a recorded step is a fixed input: it is the one assembled
  // at q_asm by the last refresh.  Reverse-mode products of every force term
  // with per-node adjoint weights zw (weights of the unscaled term buffers).
  std::vector<v3> q_asm;

  static void cofactor(const Real F[9], Real C[9]) {  // C = d det(F) / dF
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
  void clamp_grad(const Real F[9], const Real adj[9], Real out[9]) const {
    double Fd[9], ad[9], od[9];
    for (int k = 0; k < 9; ++k) {
      Fd[k] = F[k];
      ad[k] = adj[k];
    }
    clamp_stretch_grad(Fd, prm.min_stretch, prm.max_stretch, ad, od);
    for (int k = 0; k < 9; ++k) out[k] = (Real)od[k];
  }

  // Reverse of eval_tets (+ eval_tet_activations): with t the tet term
  // buffer (restoring forces), L = sum_i zw_i . t_i; accumulates dL/dx,
  // dL/dv, dL/dact and dL/d(mu, lambda, k_damp) per tet.
  void vjp_tets(const v3* x, const v3* v, const Real* act, const v3* zw, v3* xb, v3* vb, Real* actb, Real* matb) const {
    for (int e = 0; e < T; ++e) {
      int n[4] = {tet_idx[4 * e], tet_idx[4 * e + 1], tet_idx[4 * e + 2], tet_idx[4 * e + 3]};
      const Real mu = tet_mat[3 * e], lam = tet_mat[3 * e + 1], kd = tet_mat[3 * e + 2];
      const Real* Dm = &tet_pose[9 * e];
      const Real s = det3(Dm);
      const v3 x10 = x[n[1]] - x[n[0]], x20 = x[n[2]] - x[n[0]], x30 = x[n[3]] - x[n[0]];
      const v3 w[3] = {zw[n[1]] - zw[n[0]], zw[n[2]] - zw[n[0]], zw[n[3]] - zw[n[0]]};
      v3 dJb[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      Real cwb = 0;
      const v3 dJ[3] = {cross(x20, x30) * s, cross(x30, x10) * s, cross(x10, x20) * s};
      const Real cw = dot(w[0], dJ[0]) + dot(w[1], dJ[1]) + dot(w[2], dJ[2]);
      v3 ub[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      Real Fb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      if (!(mu == Real(0) && lam == Real(0) && kd == Real(0))) {
        const v3 u[3] = {v[n[1]] - v[n[0]], v[n[2]] - v[n[0]], v[n[3]] - v[n[0]]};
        const Real V = Real(1) / (s * Real(6));
        const Real alpha = Real(1) + mu / lam - mu / (Real(4) * lam);
        const Real muV = mu * V, lamV = lam * V, kdV = kd * V;
        Real F[9], Fc[9], dF[9], G[9];
        const Real ds[3][3] = {{x10.x, x20.x, x30.x}, {x10.y, x20.y, x30.y}, {x10.z, x20.z, x30.z}};
        const Real vs[3][3] = {{u[0].x, u[1].x, u[2].x}, {u[0].y, u[1].y, u[2].y}, {u[0].z, u[1].z, u[2].z}};
        const Real ws[3][3] = {{w[0].x, w[1].x, w[2].x}, {w[0].y, w[1].y, w[2].y}, {w[0].z, w[1].z, w[2].z}};
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            F[3 * r + c] = ds[r][0] * Dm[c] + ds[r][1] * Dm[3 + c] + ds[r][2] * Dm[6 + c];
            dF[3 * r + c] = vs[r][0] * Dm[c] + vs[r][1] * Dm[3 + c] + vs[r][2] * Dm[6 + c];
            G[3 * r + c] = ws[r][0] * Dm[c] + ws[r][1] * Dm[3 + c] + ws[r][2] * Dm[6 + c];
          }
        for (int k = 0; k < 9; ++k) Fc[k] = F[k];
        clamp_stretch(Fc, (Real)prm.min_stretch, (Real)prm.max_stretch);
        Real Ic = 0, GF = 0, GdF = 0;
        for (int k = 0; k < 9; ++k) {
          Ic += Fc[k] * Fc[k];
          GF += G[k] * Fc[k];
          GdF += G[k] * dF[k];
        }
        const Real phi = Real(1) - Real(1) / (Ic + Real(1));
        const Real dphi = Real(1) / ((Ic + Real(1)) * (Ic + Real(1)));
        const Real detF = det3(Fc);
        const Real J = std::max(detF, Real(0.01));
        const Real cv = dot(u[0], dJ[0]) + dot(u[1], dJ[1]) + dot(u[2], dJ[2]);
        const Real ft = (J - alpha) * lamV + kdV * cv;
        // L = -<G, P> - ft * cw,  P = muV phi Fc + kdV dF.
        const Real ftb = -cw;
        cwb += -ft;
        Real Fcb[9];
        for (int k = 0; k < 9; ++k) Fcb[k] = -muV * phi * G[k] - Real(2) * muV * dphi * GF * Fc[k];
        if (detF > Real(0.01)) {
          Real C[9];
          cofactor(Fc, C);
          for (int k = 0; k < 9; ++k) Fcb[k] += ftb * lamV * C[k];
        }
        clamp_grad(F, Fcb, Fb);
        // dF/dt = Dv Dm: dL/dDv = -kdV G Dm^T; damping part of ft through cv.
        for (int a = 0; a < 3; ++a) {
          v3 col = {0, 0, 0};
          Real* cc = &col.x;
          for (int r = 0; r < 3; ++r) {
            Real acc = 0;
            for (int c = 0; c < 3; ++c) acc += -kdV * G[3 * r + c] * Dm[3 * a + c];
            cc[r] = acc;
          }
          ub[a] = col + dJ[a] * (ftb * kdV);
          dJb[a] = dJb[a] + u[a] * (ftb * kdV);
        }
        const Real phib = -muV * GF;  // dL/dphi... folded below into mu
        (void)phib;
        // Materials (raw, per tet).
        const Real lamVb = ftb * (J - alpha);
        const Real alphab = -ftb * lamV;
        const Real muVb = -phi * GF;
        const Real kdVb = -GdF + ftb * cv;
        matb[3 * e] += muVb * V + alphab * (Real(0.75) / lam);
        matb[3 * e + 1] += lamVb * V + alphab * (Real(-0.75) * mu / (lam * lam));
        matb[3 * e + 2] += kdVb * V;
      }
      // Activation term t_act = -(lam act / 6) (cw / s) in adjoint form.
      if (prm.enable_activations && act) {
        const Real sc = lam * act[e] / Real(6);
        cwb += -sc / s;
        actb[e] += -(lam / Real(6)) * cw / s;
        matb[3 * e + 1] += -(act[e] / Real(6)) * cw / s;
      }
      for (int a = 0; a < 3; ++a) dJb[a] = dJb[a] + w[a] * cwb;
      // dJ1 = s x20 x x30, dJ2 = s x30 x x10, dJ3 = s x10 x x20.
      v3 e10 = cross(dJb[1], x30) + cross(x20, dJb[2]);
      v3 e20 = cross(dJb[2], x10) + cross(x30, dJb[0]);
      v3 e30 = cross(dJb[0], x20) + cross(x10, dJb[1]);
      e10 = e10 * s;
      e20 = e20 * s;
      e30 = e30 * s;
      // F = Ds Dm: dL/dDs = Fb Dm^T.
      v3 dsb[3];
      for (int a = 0; a < 3; ++a) {
        Real* cc = &dsb[a].x;
        for (int r = 0; r < 3; ++r) cc[r] = Fb[3 * r] * Dm[3 * a] + Fb[3 * r + 1] * Dm[3 * a + 1] + Fb[3 * r + 2] * Dm[3 * a + 2];
      }
      const v3 eb[3] = {e10 + dsb[0], e20 + dsb[1], e30 + dsb[2]};
      for (int a = 0; a < 3; ++a) {
        xb[n[a + 1]] = xb[n[a + 1]] + eb[a];
        xb[n[0]] = xb[n[0]] - eb[a];
        vb[n[a + 1]] = vb[n[a + 1]] + ub[a];
        vb[n[0]] = vb[n[0]] - ub[a];
      }
    }
  }

  // Reverse of eval_hexes (hex materials are fixed inputs, as in the reference).
  void vjp_hexes(const v3* x, const v3* v, const v3* zw, v3* xb, v3* vb) const {
    for (int e = 0; e < H; ++e) {
      const Real mu = hex_mat[3 * e], lam = hex_mat[3 * e + 1], kd = hex_mat[3 * e + 2];
      const Real alpha = Real(1) + mu / lam - mu / (Real(4) * lam);
      for (int g = 0; g < 8; ++g) {
        v3 dN[8];
        Real F[9], Fc[9];
        hex_gp(e, g, x, dN, F);
        for (int k = 0; k < 9; ++k) Fc[k] = F[k];
        clamp_stretch(Fc, (Real)prm.min_stretch, (Real)prm.max_stretch);
        const Real wg = hex_detJ0w[8 * e + g];
        Real Pb[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // -w G, G = sum_a zw_a dN_a^T
        for (int a = 0; a < 8; ++a) {
          const v3 z = zw[hex_idx[8 * e + a]];
          const Real zs[3] = {z.x, z.y, z.z}, d[3] = {dN[a].x, dN[a].y, dN[a].z};
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) Pb[3 * r + c] -= wg * zs[r] * d[c];
        }
        Real Ic = 0, PF = 0;
        for (int k = 0; k < 9; ++k) {
          Ic += Fc[k] * Fc[k];
          PF += Pb[k] * Fc[k];
        }
        const Real phi = Real(1) - Real(1) / (Ic + Real(1));
        const Real dphi = Real(1) / ((Ic + Real(1)) * (Ic + Real(1)));
        const Real detF = det3(Fc);
        const Real J = std::max(detF, Real(0.01));
        Real C[9];
        const v3 f0 = {Fc[0], Fc[3], Fc[6]}, f1 = {Fc[1], Fc[4], Fc[7]}, f2 = {Fc[2], Fc[5], Fc[8]};
        const v3 cc0 = cross(f1, f2), cc1 = cross(f2, f0), cc2 = cross(f0, f1);
        const Real cof[9] = {cc0.x, cc1.x, cc2.x, cc0.y, cc1.y, cc2.y, cc0.z, cc1.z, cc2.z};
        Real Pcof = 0;
        for (int k = 0; k < 9; ++k) Pcof += Pb[k] * cof[k];
        Real Fcb[9];
        for (int k = 0; k < 9; ++k) Fcb[k] = mu * phi * Pb[k] + Real(2) * mu * dphi * PF * Fc[k];
        if (detF > Real(0.01)) {
          cofactor(Fc, C);
          for (int k = 0; k < 9; ++k) Fcb[k] += lam * Pcof * C[k];
        }
        // d<Pb, cof(F)>/dF by columns.
        const v3 p0 = {Pb[0], Pb[3], Pb[6]}, p1 = {Pb[1], Pb[4], Pb[7]}, p2 = {Pb[2], Pb[5], Pb[8]};
        const Real sc = lam * (J - alpha);
        const v3 g0 = (cross(p1, f2) + cross(f1, p2)) * sc, g1 = (cross(p2, f0) + cross(f2, p0)) * sc,
                 g2 = (cross(p0, f1) + cross(f0, p1)) * sc;
        Fcb[0] += g0.x; Fcb[3] += g0.y; Fcb[6] += g0.z;
        Fcb[1] += g1.x; Fcb[4] += g1.y; Fcb[7] += g1.z;
        Fcb[2] += g2.x; Fcb[5] += g2.y; Fcb[8] += g2.z;
        Real Fb[9];
        clamp_grad(F, Fcb, Fb);
        for (int a = 0; a < 8; ++a) {
          const int n = hex_idx[8 * e + a];
          const v3 xa = {Fb[0] * dN[a].x + Fb[1] * dN[a].y + Fb[2] * dN[a].z,
                         Fb[3] * dN[a].x + Fb[4] * dN[a].y + Fb[5] * dN[a].z,
                         Fb[6] * dN[a].x + Fb[7] * dN[a].y + Fb[8] * dN[a].z};
          const v3 va = {Pb[0] * dN[a].x + Pb[1] * dN[a].y + Pb[2] * dN[a].z,
                         Pb[3] * dN[a].x + Pb[4] * dN[a].y + Pb[5] * dN[a].z,
                         Pb[6] * dN[a].x + Pb[7] * dN[a].y + Pb[8] * dN[a].z};
          xb[n] = xb[n] + xa;
          vb[n] = vb[n] + va * kd;
        }
      }
    }
  }

  // Reverse of eval_ground (branches follow the forward: min/max/abs give
  // their derivative to the first argument only when it is selected
  // strictly, as Warp's adjoints do).
  void vjp_ground(const v3* x, const v3* v, const v3* zw, v3* xb, v3* vb) const {
    if (!prm.has_ground) return;
    const v3 n = {(Real)prm.ground[0], (Real)prm.ground[1], (Real)prm.ground[2]};
    const Real gx = (Real)prm.gravity[0], gy = (Real)prm.gravity[1], gz = (Real)prm.gravity[2];
    const Real gmag = std::sqrt(gx * gx + gy * gy + gz * gz);
    for (int i = 0; i < P; ++i) {
      if ((flags[i] & 1) == 0) continue;
      const Real c = std::min(dot(n, x[i]) + (Real)prm.ground[3] - radius[i], Real(0));
      if (c >= Real(0)) continue;
      const Real vn = dot(n, v[i]);
      const Real fn = c * (Real)prm.ke + std::min(vn, Real(0)) * (Real)prm.kd;
      v3 vt = v[i] - n * vn;
      const Real vs = std::sqrt(dot(vt, vt));
      const v3 that = vs > Real(0) ? vt * (Real(1) / vs) : vt;
      const Real inv_m = inv_mass[i];
      const Real m = inv_m > Real(0) ? Real(1) / std::max(inv_m, Real(1e-9)) : Real(0);
      const Real n_cap = m * gmag;
      const Real afn = std::fabs(fn);
      const Real fn_eff = std::min(afn, n_cap);
      const Real ftf = std::min(vs * (Real)prm.kf, (Real)prm.mu * fn_eff);
      // t = -n fn - that ft
      const v3 z = zw[i];
      Real fnb = -dot(z, n);
      const Real ftb = -dot(z, that);
      const v3 thatb = z * (-ftf);
      Real vsb = 0, fneb = 0;
      if (vs * (Real)prm.kf < (Real)prm.mu * fn_eff) vsb += ftb * (Real)prm.kf;
      else fneb += ftb * (Real)prm.mu;
      if (afn < n_cap) fnb += fneb * (fn < Real(0) ? Real(-1) : Real(1));
      v3 vtb;
      if (vs > Real(0)) {
        const Real tt = dot(that, thatb);
        vtb = (thatb - that * tt) * (Real(1) / vs) + that * vsb;
      } else {
        vtb = thatb;
      }
      Real vnb = -dot(n, vtb);
      v3 vbi = vtb;
      if (vn < Real(0)) vnb += fnb * (Real)prm.kd;
      vbi = vbi + n * vnb;
      xb[i] = xb[i] + n * (fnb * (Real)prm.ke);
      vb[i] = vb[i] + vbi;
    }
  }

  void vjp_damping(const v3* zw, v3* vb) const {
    if (!(prm.linear_damping > 0.0)) return;
    for (int i = 0; i < P; ++i) {
      if ((flags[i] & 1) == 0) continue;
      vb[i] = vb[i] + zw[i] * (-(Real)prm.linear_damping * mass[i]);
    }
  }

  // VJP of eval_terms(mask) with per-node weights zw.
  // Reverse of eval_springs: L = sum_i zw_i . f_i = sum_s (zw_q - zw_p) . fs,
  // fs = dir (ke (len - rest) + kd dir . v_pq), dir = x_pq / len.
  void vjp_springs(const v3* x, const v3* v, const v3* zw, v3* xb, v3* vb) const {
    for (int s = 0; s < S; ++s) {
      int p = spring_idx[2 * s], q = spring_idx[2 * s + 1];
      if (p == -1 || q == -1) continue;
      v3 xpq = x[p] - x[q], vpq = v[p] - v[q];
      Real len = std::sqrt(dot(xpq, xpq));
      if (len < Real(1e-6)) continue;
      v3 d = xpq * (Real(1) / len);
      Real phi = spring_ke[s] * (len - spring_rest[s]) + spring_kd[s] * dot(d, vpq);
      v3 w = zw[q] - zw[p];
      Real dw = dot(d, w), dvel = dot(d, vpq);
      v3 gx = (w - d * dw) * (phi / len) + (d * spring_ke[s] + (vpq - d * dvel) * (spring_kd[s] / len)) * dw;
      v3 gv = d * (dw * spring_kd[s]);
      xb[p] = xb[p] + gx;
      xb[q] = xb[q] - gx;
      vb[p] = vb[p] + gv;
      vb[q] = vb[q] - gv;
    }
  }

  // Reverse of eval_triangles without drag and lift (callers reject those):
  // L = sum_n zw_n . F_n with F_0 = f0, F_1 = -f1, F_2 = -f2.  With a1 = zw_0 - zw_1,
  // a2 = zw_0 - zw_2: L = <P, [a1 a2] D> + g (a1 . dc/dx10 + a2 . dc/dx20),
  // P = k_mu (1 - 2/Ic) F + k_damp dF, g = k_lambda c + k_damp dc/dt.
  void vjp_tris(const v3* x, const v3* v, const v3* zw, v3* xb, v3* vb) const {
    for (int t = 0; t < R; ++t) {
      const int i = tri_idx[3 * t], j = tri_idx[3 * t + 1], k = tri_idx[3 * t + 2];
      const Real* D = &tri_pose[4 * t];
      const Real inv_area = (D[0] * D[3] - D[1] * D[2]) * Real(2);
      const Real A0 = Real(1) / inv_area;
      const Real kmu = tri_mat[5 * t] * A0, klam = tri_mat[5 * t + 1] * A0, kd = tri_mat[5 * t + 2] * A0;
      const v3 x10 = x[j] - x[i], x20 = x[k] - x[i], v10 = v[j] - v[i], v20 = v[k] - v[i];
      const v3 a1 = zw[i] - zw[j], a2 = zw[i] - zw[k];
      const v3 F1 = x10 * D[0] + x20 * D[2], F2 = x10 * D[1] + x20 * D[3];
      const Real Ic = dot(F1, F1) + dot(F2, F2);
      const Real dev = (Ic - Real(2)) / Ic;
      const v3 b1 = a1 * D[0] + a2 * D[2], b2 = a1 * D[1] + a2 * D[3];
      const Real FB = dot(F1, b1) + dot(F2, b2);
      const Real cI = Real(4) * FB / (Ic * Ic);
      const v3 G1 = (b1 * dev + F1 * cI) * kmu, G2 = (b2 * dev + F2 * cI) * kmu;
      v3 gx10 = G1 * D[0] + G2 * D[1], gx20 = G1 * D[2] + G2 * D[3];
      v3 gv10 = (b1 * D[0] + b2 * D[1]) * kd, gv20 = (b1 * D[2] + b2 * D[3]) * kd;
      const v3 nr = cross(x10, x20);
      const Real ln = std::sqrt(dot(nr, nr));
      if (ln > Real(0)) {
        const v3 nh = nr * (Real(1) / ln);
        const Real kap = inv_area * Real(0.5);
        const v3 cq = cross(x20, nh) * kap, cr = cross(nh, x10) * kap;  // dc/dx10, dc/dx20
        const Real c = ln * kap - Real(1);
        const Real cdot = dot(cq, v10) + dot(cr, v20), hA = dot(cq, a1) + dot(cr, a2);
        const Real g = klam * c + kd * cdot;
        // gradient of Dc[Y] = dc/dx10 . y1 + dc/dx20 . y2 with respect to (x10, x20)
        auto dgrad = [&](v3 y1, v3 y2, v3& o1, v3& o2) {
          const v3 m = cross(y1, x20) + cross(x10, y2);
          const v3 u = (m - nh * dot(nh, m)) * (Real(1) / ln);
          o1 = (cross(x20, u) + cross(y2, nh)) * kap;
          o2 = (cross(u, x10) + cross(nh, y1)) * kap;
        };
        v3 dv1, dv2, da1, da2;
        dgrad(v10, v20, dv1, dv2);
        dgrad(a1, a2, da1, da2);
        gx10 = gx10 + (cq * klam + dv1 * kd) * hA + da1 * g;
        gx20 = gx20 + (cr * klam + dv2 * kd) * hA + da2 * g;
        gv10 = gv10 + cq * (kd * hA);
        gv20 = gv20 + cr * (kd * hA);
      }
      xb[j] = xb[j] + gx10;
      xb[k] = xb[k] + gx20;
      xb[i] = xb[i] - gx10 - gx20;
      vb[j] = vb[j] + gv10;
      vb[k] = vb[k] + gv20;
      vb[i] = vb[i] - gv10 - gv20;
    }
  }

  void vjp_terms(int mask, const v3* x, const v3* v, const Real* act, const v3* zw, v3* xb, v3* vb, Real* actb,
                 Real* matb) const {
    if ((mask & TERM_SPRING) && S) vjp_springs(x, v, zw, xb, vb);
    if ((mask & TERM_TRI) && R) vjp_tris(x, v, zw, xb, vb);
    if ((mask & TERM_TET) && T) vjp_tets(x, v, (prm.enable_activations ? act : nullptr), zw, xb, vb, actb, matb);
    if ((mask & TERM_HEX) && H) vjp_hexes(x, v, zw, xb, vb);
    if (mask & TERM_GROUND) vjp_ground(x, v, zw, xb, vb);
    if (mask & TERM_DAMPING) vjp_damping(zw, vb);
  }

  // -z^T (dA/d mu_e) y and -z^T (dA/d lambda_e) y per tet, with A assembled
  // at q_asm (the element tangent's block rows sum to zero, so the
  // predictor's -(dC/dmu) v folds into y = dv + v).
  void matrix_material_terms(const v3* z, const v3* y, Real dt, Real* matb) const {
    const v3* x = q_asm.data();
    for (int e = 0; e < T; ++e) {
      int n[4] = {tet_idx[4 * e], tet_idx[4 * e + 1], tet_idx[4 * e + 2], tet_idx[4 * e + 3]};
      if (tet_mat[3 * e] == Real(0) && tet_mat[3 * e + 1] == Real(0) && tet_mat[3 * e + 2] == Real(0)) continue;
      const Real* Dm = &tet_pose[9 * e];
      const Real s = det3(Dm);
      if (s * Real(6) <= Real(0)) continue;
      const Real V = Real(1) / (s * Real(6));
      const v3 x10 = x[n[1]] - x[n[0]], x20 = x[n[2]] - x[n[0]], x30 = x[n[3]] - x[n[0]];
      Real F[9], Gz[9], Gy[9];
      const v3 wz[3] = {z[n[1]] - z[n[0]], z[n[2]] - z[n[0]], z[n[3]] - z[n[0]]};
      const v3 wy[3] = {y[n[1]] - y[n[0]], y[n[2]] - y[n[0]], y[n[3]] - y[n[0]]};
      const Real ds[3][3] = {{x10.x, x20.x, x30.x}, {x10.y, x20.y, x30.y}, {x10.z, x20.z, x30.z}};
      const Real zs[3][3] = {{wz[0].x, wz[1].x, wz[2].x}, {wz[0].y, wz[1].y, wz[2].y}, {wz[0].z, wz[1].z, wz[2].z}};
      const Real ys[3][3] = {{wy[0].x, wy[1].x, wy[2].x}, {wy[0].y, wy[1].y, wy[2].y}, {wy[0].z, wy[1].z, wy[2].z}};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          F[3 * r + c] = ds[r][0] * Dm[c] + ds[r][1] * Dm[3 + c] + ds[r][2] * Dm[6 + c];
          Gz[3 * r + c] = zs[r][0] * Dm[c] + zs[r][1] * Dm[3 + c] + zs[r][2] * Dm[6 + c];
          Gy[3 * r + c] = ys[r][0] * Dm[c] + ys[r][1] * Dm[3 + c] + ys[r][2] * Dm[6 + c];
        }
      clamp_stretch(F, (Real)prm.min_stretch, (Real)prm.max_stretch);
      Real Ic = 0, zy = 0, zF = 0, yF = 0;
      for (int k = 0; k < 9; ++k) {
        Ic += F[k] * F[k];
        zy += Gz[k] * Gy[k];
        zF += Gz[k] * F[k];
        yF += Gy[k] * F[k];
      }
      const Real Ic1 = std::max(Ic + Real(1), Real(1e-6));
      const Real sdev = Real(1) - Real(1) / Ic1;
      const v3 dJ[3] = {cross(x20, x30) * s, cross(x30, x10) * s, cross(x10, x20) * s};
      const Real cz = dot(wz[0], dJ[0]) + dot(wz[1], dJ[1]) + dot(wz[2], dJ[2]);
      const Real cy = dot(wy[0], dJ[0]) + dot(wy[1], dJ[1]) + dot(wy[2], dJ[2]);
      const Real h2 = dt * dt;
      matb[3 * e] += -h2 * V * (sdev * zy + Real(2) / (Ic1 * Ic1) * zF * yF);
      matb[3 * e + 1] += -h2 * V * cz * cy;
    }
  }

  // Differentiable step forward (differentiation.py step): rhs = h sum terms
  // + h ext - sum_j A_ij (v_j - v_i), A dv = rhs solved from zero to `tol`.
  // With fixed particles (an extension: the reference rejects them when
  // differentiable) the step restores their inputs, projects the right-hand
  // side and solves with the projected operator PAP + I - P, as the fast path.
  void diff_forward(const v3* q_in, const v3* qd_in, const v3* ext, const Real* act, Real dt, double tol,
                    v3* q_out, v3* qd_out, v3* dv_out) {
    std::vector<v3> qr(q_in, q_in + P), vr(qd_in, qd_in + P);
    for (int i : fixed_idx) {
      qr[i] = anchor[i];
      vr[i] = v3{0, 0, 0};
    }
    const v3* q = qr.data();
    const v3* qd = vr.data();
    std::vector<v3> f(P);
    forces_scaled(q, qd, ext, act, dt, f.data());
    if (T || H || S) {
      std::vector<v3> c(P);
      predictor(qd, c.data());
      for (int i = 0; i < P; ++i) f[i] = f[i] - c[i];
      spring_restore(qd, dt, f.data());
    }
    for (int i : fixed_idx) f[i] = v3{0, 0, 0};
    std::vector<v3> x(P, v3{0, 0, 0});
    const double keep_tol = prm.tol;
    const int keep_max = prm.maxiter;
    prm.tol = tol;
    prm.maxiter = 100000;
    cg(f, x);
    prm.tol = keep_tol;
    prm.maxiter = keep_max;
    for (int i = 0; i < P; ++i) {
      dv_out[i] = x[i];
      qd_out[i] = qd[i] + x[i];
      q_out[i] = q[i] + qd_out[i] * dt;
    }
    for (int i : fixed_idx) {
      q_out[i] = anchor[i];
      qd_out[i] = v3{0, 0, 0};
    }
  }
  // restore_spring_damping_predictor implicit_predictor.py:26-41
  void spring_restore(const v3* qd, Real dt, v3* f) const {
    for (int s = 0; s < S; ++s) {
      int a = spring_idx[2 * s], b = spring_idx[2 * s + 1];
      if (a >= 0 && b >= 0) {
        v3 corr = (qd[a] - qd[b]) * (dt * spring_kd[s]);
        f[a] = f[a] + corr;
        f[b] = f[b] - corr;
      }
    }
  }
  void predictor(const v3* qd, v3* c) const {
    for (int i = 0; i < P; ++i) {
      Real acc[3] = {0, 0, 0};
      for (int k = offsets[i]; k < offsets[i + 1]; ++k) {
        const Real* b = &values[9 * (size_t)k];
        v3 d = qd[columns[k]] - qd[i];
        for (int r = 0; r < 3; ++r) acc[r] += b[3 * r] * d.x + b[3 * r + 1] * d.y + b[3 * r + 2] * d.z;
      }
      c[i] = {acc[0], acc[1], acc[2]};
    }
  }

  // Reverse of diff_forward for output adjoints (qb_out, vb_out).
  void diff_backward(const v3* q_in, const v3* qd_in, const Real* act, const v3* dv, const v3* qb_out,
                     const v3* vb_out, Real dt, double tol, v3* qb_in, v3* vb_in, Real* actb, v3* extb, Real* matb) {
    std::vector<v3> qr(q_in, q_in + P), vr(qd_in, qd_in + P);
    for (int i : fixed_idx) {
      qr[i] = anchor[i];
      vr[i] = v3{0, 0, 0};
    }
    const v3* q = qr.data();
    const v3* qd = vr.data();
    std::vector<v3> u(P), z(P, v3{0, 0, 0}), zw(P);
    for (int i = 0; i < P; ++i) u[i] = vb_out[i] + qb_out[i] * dt;
    for (int i : fixed_idx) u[i] = v3{0, 0, 0};  // fixed outputs are constants
    const double keep_tol = prm.tol;
    const int keep_max = prm.maxiter;
    prm.tol = tol;
    prm.maxiter = 100000;
    cg(u, z);  // A symmetric: the adjoint operator is A
    prm.tol = keep_tol;
    prm.maxiter = keep_max;
    for (int i = 0; i < P; ++i) {
      zw[i] = z[i] * dt;
      qb_in[i] = qb_out[i];
      vb_in[i] = u[i];
      extb[i] = z[i] * dt;
    }
    vjp_terms(TERM_SPRING | TERM_TET | TERM_HEX | TERM_TRI | TERM_GROUND | TERM_DAMPING, q, qd, act, zw.data(), qb_in,
              vb_in, actb, matb);
    for (int s = 0; s < S; ++s) {  // spring-damping restore: dt kd (z_p - z_q) . (v_p - v_q)
      int a = spring_idx[2 * s], b = spring_idx[2 * s + 1];
      if (a >= 0 && b >= 0) {
        v3 g = (z[a] - z[b]) * (dt * spring_kd[s]);
        vb_in[a] = vb_in[a] + g;
        vb_in[b] = vb_in[b] - g;
      }
    }
    if (T || H || S) {  // predictor: vb_j += -(A z)_j + (sum_k A_jk)^T z_j
      for (int i = 0; i < P; ++i) {
        Real az[3] = {0, 0, 0}, rs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = offsets[i]; k < offsets[i + 1]; ++k) {
          const Real* b = &values[9 * (size_t)k];
          v3 zc = z[columns[k]];
          for (int r = 0; r < 3; ++r) az[r] += b[3 * r] * zc.x + b[3 * r + 1] * zc.y + b[3 * r + 2] * zc.z;
          for (int q9 = 0; q9 < 9; ++q9) rs[q9] += b[q9];
        }
        const v3 zi = z[i];
        const v3 rt = {rs[0] * zi.x + rs[3] * zi.y + rs[6] * zi.z, rs[1] * zi.x + rs[4] * zi.y + rs[7] * zi.z,
                       rs[2] * zi.x + rs[5] * zi.y + rs[8] * zi.z};
        vb_in[i] = vb_in[i] - v3{az[0], az[1], az[2]} + rt;
      }
    }
    std::vector<v3> y(P);
    for (int i = 0; i < P; ++i) y[i] = dv[i] + qd[i];
    matrix_material_terms(z.data(), y.data(), dt, matb);
    for (int i : fixed_idx) {  // fixed inputs are overwritten by their anchors
      qb_in[i] = v3{0, 0, 0};
      vb_in[i] = v3{0, 0, 0};
    }
  }

  void eval_terms(int mask, const v3* q, const v3* qd, const Real* act, v3* out) {
    std::vector<v3> acc(P, v3{0, 0, 0});
    if ((mask & TERM_SPRING) && S) eval_springs(q, qd, acc);
    if ((mask & TERM_TRI) && R) eval_tris(q, qd, acc);
    if ((mask & TERM_TET) && T) {
      eval_tets(q, qd, acc);
      if (prm.enable_activations && act) eval_tet_activations(q, act, acc);
    }
    if ((mask & TERM_HEX) && H) eval_hexes(q, qd, acc);
    if (mask & TERM_GROUND) eval_ground(q, qd, acc);
    if (mask & TERM_GRAVITY) {
      std::vector<v3> g(P, v3{0, 0, 0});
      eval_gravity(g);
      for (int i = 0; i < P; ++i) acc[i] = acc[i] + g[i];
    }
    if (mask & TERM_DAMPING) {
      std::vector<v3> d(P, v3{0, 0, 0});
      eval_damping(qd, d);
      for (int i = 0; i < P; ++i) acc[i] = acc[i] + d[i];
    }
    for (int i = 0; i < P; ++i) out[i] = acc[i];
  }
};

}  // namespace orc

// ---------------------------------------------------------------------------
// C ABI (ctypes).  Arrays cross the boundary as double; `bits` selects the
// real type of the restatement (32 or 64).
// ---------------------------------------------------------------------------
using namespace orc;

struct Handle {
  int bits;
  Solver<float>* s32;
  Solver<double>* s64;
};

template <typename Real>
static void fill(Solver<Real>& s, int P, const double* q, const double* mass, const double* inv_mass,
                 const double* radius, const int* flags, int T, const int* ti, const double* tp, const double* tm,
                 int H, const int* hi, const double* hJ, const double* hw, const double* hm, int nf, const int* fx) {
  s.P = P;
  s.T = T;
  s.H = H;
  s.anchor.resize(P);
  for (int i = 0; i < P; ++i) s.anchor[i] = {(Real)q[3 * i], (Real)q[3 * i + 1], (Real)q[3 * i + 2]};
  s.mass.assign(mass, mass + P);
  s.inv_mass.assign(inv_mass, inv_mass + P);
  s.radius.assign(radius, radius + P);
  s.flags.assign(flags, flags + P);
  s.tet_idx.assign(ti, ti + 4 * (size_t)T);
  s.tet_pose.assign(tp, tp + 9 * (size_t)T);
  s.tet_mat.assign(tm, tm + 3 * (size_t)T);
  s.hex_idx.assign(hi, hi + 8 * (size_t)H);
  s.hex_invJ0.assign(hJ, hJ + 72 * (size_t)H);
  s.hex_detJ0w.assign(hw, hw + 8 * (size_t)H);
  s.hex_mat.assign(hm, hm + 3 * (size_t)H);
  s.fixed_idx.assign(fx, fx + nf);
  s.fixed_mask.assign(P, 0);
  for (int i = 0; i < nf; ++i) s.fixed_mask[fx[i]] = 1;
}

template <typename Real>
static std::vector<V3<Real>> to_v3(const double* a, int P) {
  std::vector<V3<Real>> v(P);
  for (int i = 0; i < P; ++i) v[i] = {(Real)a[3 * i], (Real)a[3 * i + 1], (Real)a[3 * i + 2]};
  return v;
}
template <typename Real>
static void from_v3(const std::vector<V3<Real>>& v, double* a) {
  for (size_t i = 0; i < v.size(); ++i) {
    a[3 * i] = v[i].x;
    a[3 * i + 1] = v[i].y;
    a[3 * i + 2] = v[i].z;
  }
}

#define DISPATCH(h, ...)         \
  do {                           \
    if ((h)->bits == 32) {       \
      auto& S = *(h)->s32;       \
      using Real = float;        \
      (void)sizeof(Real);        \
      __VA_ARGS__;               \
    } else {                     \
      auto& S = *(h)->s64;       \
      using Real = double;       \
      (void)sizeof(Real);        \
      __VA_ARGS__;               \
    }                            \
  } while (0)

extern "C" {

void* orc_create(int bits, int P, const double* q, const double* mass, const double* inv_mass, const double* radius,
                 const int* flags, int T, const int* ti, const double* tp, const double* tm, int H, const int* hi,
                 const double* hJ, const double* hw, const double* hm, int nf, const int* fx) {
  Handle* h = new Handle{bits, nullptr, nullptr};
  if (bits == 32) {
    h->s32 = new Solver<float>();
    fill(*h->s32, P, q, mass, inv_mass, radius, flags, T, ti, tp, tm, H, hi, hJ, hw, hm, nf, fx);
  } else {
    h->s64 = new Solver<double>();
    fill(*h->s64, P, q, mass, inv_mass, radius, flags, T, ti, tp, tm, H, hi, hJ, hw, hm, nf, fx);
  }
  return h;
}

void orc_destroy(void* p) {
  Handle* h = (Handle*)p;
  delete h->s32;
  delete h->s64;
  delete h;
}

// params: min_s, max_s, has_ground, g0..g3, ke, kd, kf, mu, lin_damp, gx, gy, gz, enable_act, tol, maxiter
// Host threads of the CG loops of later solves (1, the default, keeps the
// serial summation order; more threads only for timing).
void orc_set_threads(void* p, int n) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S.nthreads = n < 1 ? 1 : n);
}

void orc_set_params(void* p, const double* a) {
  Handle* h = (Handle*)p;
  Params prm;
  prm.min_stretch = a[0];
  prm.max_stretch = a[1];
  prm.has_ground = (int)a[2];
  for (int i = 0; i < 4; ++i) prm.ground[i] = a[3 + i];
  prm.ke = a[7];
  prm.kd = a[8];
  prm.kf = a[9];
  prm.mu = a[10];
  prm.linear_damping = a[11];
  for (int i = 0; i < 3; ++i) prm.gravity[i] = a[12 + i];
  prm.enable_activations = (int)a[15];
  prm.tol = a[16];
  prm.maxiter = (int)a[17];
  DISPATCH(h, S.prm = prm);
}

// Springs (indices S x 2, rest length, stiffness, damping) and triangle
// membranes (indices R x 3, inverse rest pose R x 2 x 2, materials R x 5);
// call before orc_construct.
void orc_set_springs_tris(void* p, int ns, const int* si, const double* rest, const double* ke, const double* kd,
                          int nr, const int* ti, const double* tp, const double* tm) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    S.S = ns;
    S.spring_idx.assign(si, si + 2 * (size_t)ns);
    S.spring_rest.assign(rest, rest + ns);
    S.spring_ke.assign(ke, ke + ns);
    S.spring_kd.assign(kd, kd + ns);
    S.R = nr;
    S.tri_idx.assign(ti, ti + 3 * (size_t)nr);
    S.tri_pose.assign(tp, tp + 4 * (size_t)nr);
    S.tri_mat.assign(tm, tm + 5 * (size_t)nr);
  });
}

void orc_construct(void* p, double dt) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S.construct((Real)dt));
}

void orc_refresh(void* p, double* q, double dt) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    auto v = to_v3<Real>(q, S.P);
    S.refresh(v.data(), (Real)dt);
    from_v3(v, q);
  });
}

int orc_nnz(void* p) {
  Handle* h = (Handle*)p;
  int n = 0;
  DISPATCH(h, n = (int)S.columns.size());
  return n;
}

void orc_get_matrix(void* p, int* offsets, int* columns, double* values, double* minv) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    std::copy(S.offsets.begin(), S.offsets.end(), offsets);
    std::copy(S.columns.begin(), S.columns.end(), columns);
    for (size_t i = 0; i < S.values.size(); ++i) values[i] = S.values[i];
    for (size_t i = 0; i < S.minv.size(); ++i) minv[i] = S.minv[i];
  });
}

void orc_get_dv(void* p, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, from_v3(S.dv, out));
}

void orc_set_dv(void* p, const double* in) {
  Handle* h = (Handle*)p;
  DISPATCH(h, S.dv = to_v3<Real>(in, S.P));
}

// Returns the number of CG iterations performed.
int orc_step(void* p, double* q_in, double* qd_in, double* f_out, const double* ext, const double* act, double* q_out,
             double* qd_out, double dt) {
  Handle* h = (Handle*)p;
  int it = 0;
  DISPATCH(h, {
    int P = S.P;
    auto q = to_v3<Real>(q_in, P), qd = to_v3<Real>(qd_in, P);
    std::vector<V3<Real>> f(P), qo(P), qdo(P), e;
    if (ext) e = to_v3<Real>(ext, P);
    std::vector<Real> a;
    if (act) a.assign(act, act + S.T);
    S.step(q.data(), qd.data(), f.data(), ext ? e.data() : nullptr, act ? a.data() : nullptr, qo.data(), qdo.data(),
           (Real)dt);
    from_v3(q, q_in);
    from_v3(qd, qd_in);
    from_v3(f, f_out);
    from_v3(qo, q_out);
    from_v3(qdo, qd_out);
    it = S.last_iters;
  });
  return it;
}

void orc_last_stats(void* p, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    out[0] = S.last_iters;
    out[1] = S.last_rr;
    out[2] = S.last_atol_sq;
    out[3] = S.last_scale;
  });
}

void orc_eval_terms(void* p, int mask, const double* q, const double* qd, const double* act, double* out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    auto qv = to_v3<Real>(q, S.P), qdv = to_v3<Real>(qd, S.P);
    std::vector<Real> a;
    if (act) a.assign(act, act + S.T);
    std::vector<V3<Real>> o(S.P);
    S.eval_terms(mask, qv.data(), qdv.data(), act ? a.data() : nullptr, o.data());
    from_v3(o, out);
  });
}

// Stand-alone pieces for unit checks.
void orc_clamp_stretch(int bits, const double* F, double min_s, double max_s, double* out) {
  if (bits == 32) {
    float f[9];
    for (int i = 0; i < 9; ++i) f[i] = (float)F[i];
    clamp_stretch<float>(f, (float)min_s, (float)max_s);
    for (int i = 0; i < 9; ++i) out[i] = f[i];
  } else {
    double f[9];
    std::memcpy(f, F, sizeof f);
    clamp_stretch<double>(f, min_s, max_s);
    std::memcpy(out, f, sizeof f);
  }
}

void orc_svd3(const double* F, double* U, double* S, double* V) { svd3_signed(F, U, S, V); }

void orc_block_inverse(int bits, const double* B, double* out) {
  if (bits == 32) {
    float b[9], o[9];
    for (int i = 0; i < 9; ++i) b[i] = (float)B[i];
    inverse_diagonal_block<float>(b, o);
    for (int i = 0; i < 9; ++i) out[i] = o[i];
  } else {
    inverse_diagonal_block<double>(B, out);
  }
}

// ---- differentiation (test infrastructure for the adjoint path) ----------
void orc_vjp_terms(void* p, int mask, const double* q, const double* qd, const double* act, const double* zw,
                   double* xb, double* vb, double* actb, double* matb) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    int P = S.P;
    auto x = to_v3<Real>(q, P), v = to_v3<Real>(qd, P), z = to_v3<Real>(zw, P);
    std::vector<V3<Real>> xo(P, V3<Real>{0, 0, 0}), vo(P, V3<Real>{0, 0, 0});
    std::vector<Real> a, ab(S.T, Real(0)), mb(3 * (size_t)S.T, Real(0));
    if (act) a.assign(act, act + S.T);
    S.vjp_terms(mask, x.data(), v.data(), act ? a.data() : nullptr, z.data(), xo.data(), vo.data(), ab.data(),
                mb.data());
    from_v3(xo, xb);
    from_v3(vo, vb);
    for (int e = 0; e < S.T; ++e) actb[e] = ab[e];
    for (size_t k = 0; k < mb.size(); ++k) matb[k] = mb[k];
  });
}

void orc_diff_forward(void* p, const double* q, const double* qd, const double* ext, const double* act, double dt,
                      double tol, double* q_out, double* qd_out, double* dv_out) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    int P = S.P;
    auto x = to_v3<Real>(q, P), v = to_v3<Real>(qd, P);
    std::vector<V3<Real>> e, qo(P), vo(P), d(P);
    if (ext) e = to_v3<Real>(ext, P);
    std::vector<Real> a;
    if (act) a.assign(act, act + S.T);
    S.diff_forward(x.data(), v.data(), ext ? e.data() : nullptr, act ? a.data() : nullptr, (Real)dt, tol, qo.data(),
                   vo.data(), d.data());
    from_v3(qo, q_out);
    from_v3(vo, qd_out);
    from_v3(d, dv_out);
  });
}

void orc_diff_backward(void* p, const double* q, const double* qd, const double* act, const double* dv,
                       const double* qb_out, const double* vb_out, double dt, double tol, double* qb_in, double* vb_in,
                       double* actb, double* extb, double* matb) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    int P = S.P;
    auto x = to_v3<Real>(q, P), v = to_v3<Real>(qd, P), d = to_v3<Real>(dv, P);
    auto qo = to_v3<Real>(qb_out, P), vo = to_v3<Real>(vb_out, P);
    std::vector<V3<Real>> qi(P), vi(P), eb(P);
    std::vector<Real> a, ab(S.T, Real(0)), mb(3 * (size_t)S.T, Real(0));
    if (act) a.assign(act, act + S.T);
    S.diff_backward(x.data(), v.data(), act ? a.data() : nullptr, d.data(), qo.data(), vo.data(), (Real)dt, tol,
                    qi.data(), vi.data(), ab.data(), eb.data(), mb.data());
    from_v3(qi, qb_in);
    from_v3(vi, vb_in);
    from_v3(eb, extb);
    for (int e = 0; e < S.T; ++e) actb[e] = ab[e];
    for (size_t k = 0; k < mb.size(); ++k) matb[k] = mb[k];
  });
}

// Replace the tet materials (mu, lambda, k_damp per tet); callers re-assemble.
void orc_set_tet_materials(void* p, const double* mat) {
  Handle* h = (Handle*)p;
  DISPATCH(h, {
    for (size_t k = 0; k < S.tet_mat.size(); ++k) S.tet_mat[k] = (Real)mat[k];
  });
}

}  // extern "C"
