// MI355X (gfx950) kernels and C-ABI of the SolverFEM substep.
//
// Kernels (one row = one node = one 3x3 block row of A):
//   k_prepare    fixed-node restore of the input state, per-step scalar reset
//   k_assemble   tangent assembly: G lanes per row walk the row's element
//                incidences, add the element blocks into an LDS copy of the
//                row (no global atomics), then stream the row out coalesced
//                and invert its diagonal block (block-Jacobi)
//   k_rhs        G lanes per row: elastic + damping forces from the row's
//                incidences, node forces, h-scaling, elastic predictor
//                (block row of A), fixed-row projection, per-instance max|b|
//   k_cg         persistent block-Jacobi PCG: every workgroup owns a
//                contiguous row range for the whole solve (x, r, z, p, q and
//                its row state stays in registers), tiled SpMV, one
//                instance-wide barrier per iteration, device-side residual
//                test, fused update_state + fixed restore epilogue
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "fem_hip.h"
#include "fem_adjoint.hpp"
#include "fem_contact.hpp"
#include "fem_krylov.hpp"
#include "fem_math.hpp"
#include "fem_pattern.hpp"

using namespace fem;

namespace {

constexpr int BLOCK = 256;
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
constexpr long BARRIER_SPIN_LIMIT = 1L << 22;
// Per-instance words holding max|b| (reduced by the solve).
constexpr int RHS_MAX_SLOTS = 64;
// CG partial sums per chunk and phase.
constexpr int NPART = 8;
// Slot 7 holds b.b after the prologue and r.p in every phase.
enum { PT_PAP = 0, PT_QZ, PT_QW, PT_QR, PT_QQ, PT_RZ, PT_RR, PT_BB, PT_RP = PT_BB };

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(FEM_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// Workgroup b of n runs on XCD b % 8; give each XCD a contiguous 1/8 of the
// logical blocks so that neighbouring rows share that XCD's L2.
__device__ __forceinline__ int xcd_swizzle(int b, int n) {
  if (n < 8 || (n & 7)) return b;
  return (b & 7) * (n >> 3) + (b >> 3);
}

struct RowArgs {
  int P;
  const int* offsets;
  const int* cols;
  const int* diag_rel;
  const int* tinc_ptr;
  const int* tinc;
  const uint32_t* tslots;
  const int* hinc_ptr;
  const int* hinc;
  const uint32_t* hslots;
  const int* slot_ptr;
  const uint16_t* contrib;
  const int* tet_idx;
  const float* tet_pose;
  const float* tet_mat;
  const int* hex_idx;
  const float* hex_invJ;
  const float* hex_detw;
  const float* hex_mat;
  // Per (hex, Gauss point) records of k_hex_gauss: the clamped F and the
  // tangent coefficients (read by the assembly), the first Piola stress
  // (read by the force evaluation).
  const float* hex_gp_t;
  const float* hex_gp_f;
  // Particle-particle contact force per node (fem_contact), NULL when the
  // system has no particle contact.
  const float* fpart;
  // Soft (particle-shape) contact force per node, NULL without a contact list.
  const float* fsoft;
  const float* mass;
  const float* inv_mass;
  const float* radius;
  const int* flags;
  const unsigned char* row_fixed;
  const int* row_inst;
  // springs and triangle membranes (NULL / empty when the model has none)
  const int* sinc_ptr;
  const int* sinc;       // (spring << 1) | end
  const int* sp_idx;     // [S][2]
  const float* sp_rest;
  const float* sp_ke;
  const float* sp_kd;
  const float* sp_dir;   // [S][3] bend-torque rest directions (NULL: no bend torque)
  float sp_kt, sp_kw;    // bend-torque stiffness and damping
  const int* rinc_ptr;
  const int* rinc;       // (tri << 2) | corner
  const int* tri_idx;    // [R][3]
  const float* tri_pose; // [R][4]
  const float* tri_mat;  // [R][5]
  const float2* slot_spring;  // per BSR slot signed (sum ke, sum kd) of the spring triplets on it
  const float* row_tri;       // per row sum of (k_mu + k_lambda) * rest area of its triangles
  float mn, mx;
  int has_ground;
  float gn[4];
  float ke, kd, kf, gmu;
  float lin_damp;
  float grav[3];
  float gmag;
  // Device gravity vector read at run time (Model.gravity is a device array
  // in the reference, so a captured graph sees set_gravity); NULL: grav/gmag.
  const float* gravity;
  int use_act;
};

// ---------------------------------------------------------------------------
// k_prepare: FixedParticles.restore(state_in) and per-step resets.
// ---------------------------------------------------------------------------
__global__ void k_prepare(const int* __restrict__ fixed, int nfixed, const float* __restrict__ anchor, float* q,
                          float* qd, unsigned* rhs_max, int ninst) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nfixed) {
    int i = fixed[t];
    st3(q, i, ld3(anchor, i));
    if (qd) st3(qd, i, v3{0.f, 0.f, 0.f});
  }
  if (t < ninst * RHS_MAX_SLOTS) rhs_max[t] = 0u;
}

// Per-solve reset of the CG barrier counters, XCC epochs and registration
// and of the device error word (the flat barrier's generation word keeps
// counting across launches).
// Clears the per-chunk publication flags of the persistent CG, the chunk
// queue word after them and the error word.
__global__ void k_cg_reset(unsigned* flags, int nflags, int* err) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nflags) flags[t] = 0u;
  if (t == 0) *err = 0;
}

// ---------------------------------------------------------------------------
// Hex8 helpers operating on one incidence (element e, local corner la).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hex_gather(const int* __restrict__ idx, const float* __restrict__ x, float px[8],
                                           float py[8], float pz[8]) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int n = idx[b];
    px[b] = x[3 * (size_t)n];
    py[b] = x[3 * (size_t)n + 1];
    pz[b] = x[3 * (size_t)n + 2];
  }
}

// Per (hex, Gauss point) records, one thread each (eval_hexahedra
// kernels.py:1481-1601 and build_system_matrix_hex_kernel :1609-1747 evaluate
// F and the stretch clamp for every corner of every Gauss point; here each
// is evaluated once and the element's eight corners read it).
constexpr int HEX_GT = 12;  // tangent record: clamped F (9), k_mu s, coef, padding
constexpr int HEX_GF = 9;   // force record: first Piola stress P

template <bool FORCE>
__global__ __launch_bounds__(BLOCK) void k_hex_gauss(RowArgs a, int H, const float* __restrict__ x,
                                                     const float* __restrict__ v, float* __restrict__ out) {
  const int t = blockIdx.x * BLOCK + threadIdx.x;
  if (t >= 8 * H) return;
  const int e = t >> 3, g = t & 7;
  float px[8], py[8], pz[8];
  hex_gather(a.hex_idx + 8 * (size_t)e, x, px, py, pz);
  float gx[8], gy[8], gz[8];
  hex_grads(a.hex_invJ + 72 * (size_t)e + 9 * g, g, gx, gy, gz);
  float F[9];
  hex_F(px, py, pz, gx, gy, gz, F);
  clamp_stretch(F, a.mn, a.mx);
  const float k_mu = a.hex_mat[3 * e], k_lambda = a.hex_mat[3 * e + 1];
  float ic = 0.f;
#pragma unroll
  for (int k = 0; k < 9; ++k) ic += F[k] * F[k];
  if (!FORCE) {
    const float ic1 = fmaxf(ic + 1.f, 1e-6f);
    const float s = 1.f - 1.f / ic1;
    const float coef = k_mu * 2.f / (ic1 * ic1);
    f4v* o = reinterpret_cast<f4v*>(out + HEX_GT * (size_t)t);
    o[0] = f4v{F[0], F[1], F[2], F[3]};
    o[1] = f4v{F[4], F[5], F[6], F[7]};
    o[2] = f4v{F[8], k_mu * s, coef, 0.f};
  } else {
    float vx[8], vy[8], vz[8];
    hex_gather(a.hex_idx + 8 * (size_t)e, v, vx, vy, vz);
    float dF[9];
    hex_F(vx, vy, vz, gx, gy, gz, dF);
    const float k_damp = a.hex_mat[3 * e + 2];
    const float alpha = 1.f + k_mu / k_lambda - k_mu / (4.f * k_lambda);
    const float sdev = k_mu * (1.f - 1.f / (ic + 1.f));
    float cof[9];
    cofactor(F, cof);
    const float J = fmaxf(det3(F), 0.01f);
    const float vol = k_lambda * (J - alpha);
    float* o = out + HEX_GF * (size_t)t;
#pragma unroll
    for (int k = 0; k < 9; ++k) o[k] = F[k] * sdev + dF[k] * k_damp + cof[k] * vol;
  }
}

// Restoring force on corner la of hex e from its Gauss-point stresses.
__device__ __forceinline__ v3 hex_node_force(const RowArgs& a, int e, int la) {
  v3 f = {0.f, 0.f, 0.f};
#pragma unroll 1
  for (int g = 0; g < 8; ++g) {
    const float* r = a.hex_gp_f + HEX_GF * (8 * (size_t)e + g);
    float P[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) P[k] = r[k];
    const float w = a.hex_detw[8 * (size_t)e + g];
    const v3 dNa = hex_grad_corner(a.hex_invJ + 72 * (size_t)e + 9 * g, g, la);
    f = f - mv(P, dNa) * w;
  }
  return f;
}

// Row `la` of the Hex8 tangent: K_ab, b = 0..7, summed over Gauss points, x h^2,
// from the Gauss-point records.
__device__ __forceinline__ void hex_tangent_row(const RowArgs& a, int e, int la, float h2, float K[8][9]) {
  const float k_lambda = a.hex_mat[3 * e + 1];
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int k = 0; k < 9; ++k) K[b][k] = 0.f;
  for (int g = 0; g < 8; ++g) {
    float gx[8], gy[8], gz[8];
    hex_grads(a.hex_invJ + 72 * (size_t)e + 9 * g, g, gx, gy, gz);
    const f4v* r = reinterpret_cast<const f4v*>(a.hex_gp_t + HEX_GT * (8 * (size_t)e + g));
    const f4v r0 = r[0], r1 = r[1], r2 = r[2];
    const float F[9] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w, r2.x};
    const float c1 = r2.y, coef = r2.z;
    float cof[9];
    cofactor(F, cof);
    const float w = a.hex_detw[8 * (size_t)e + g];
    const v3 dNa = hex_grad_corner(a.hex_invJ + 72 * (size_t)e + 9 * g, g, la);
    const v3 fa = mv(F, dNa), ja = mv(cof, dNa);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const v3 dNb = {gx[b], gy[b], gz[b]};
      float blk[9];
      tangent_block(blk, c1, coef, k_lambda, w, dNa, fa, ja, dNb, mv(F, dNb), mv(cof, dNb));
#pragma unroll
      for (int k = 0; k < 9; ++k) K[b][k] += blk[k];
    }
  }
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int k = 0; k < 9; ++k) K[b][k] = K[b][k] * h2;
}

// Reverse of the Hex8 force for corner la: dL/dx_la and dL/dv_la for
// L = sum_b zw_b . t_b over the element's restoring forces
// (eval_hexahedra kernels.py:1481-1601 under the tape).
__device__ __forceinline__ void hex_node_vjp(const RowArgs& a, const float* __restrict__ x,
                                             const float* __restrict__ zw, float zscale, int e, int la, v3& xb,
                                             v3& vb) {
  float px[8], py[8], pz[8], zx[8], zy[8], zz[8];
  hex_gather(a.hex_idx + 8 * (size_t)e, x, px, py, pz);
  hex_gather(a.hex_idx + 8 * (size_t)e, zw, zx, zy, zz);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    zx[b] *= zscale;
    zy[b] *= zscale;
    zz[b] *= zscale;
  }
  const float k_mu = a.hex_mat[3 * e], k_lambda = a.hex_mat[3 * e + 1], k_damp = a.hex_mat[3 * e + 2];
  const float alpha = 1.f + k_mu / k_lambda - k_mu / (4.f * k_lambda);
  xb = v3{0.f, 0.f, 0.f};
  vb = xb;
  for (int g = 0; g < 8; ++g) {
    float gx[8], gy[8], gz[8];
    hex_grads(a.hex_invJ + 72 * (size_t)e + 9 * g, g, gx, gy, gz);
    float F[9], Fc[9], G[9];
    hex_F(px, py, pz, gx, gy, gz, F);
    hex_F(zx, zy, zz, gx, gy, gz, G);
#pragma unroll
    for (int k = 0; k < 9; ++k) Fc[k] = F[k];
    clamp_stretch(Fc, a.mn, a.mx);
    const float w = a.hex_detw[8 * (size_t)e + g];
    float Pb[9], cof[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Pb[k] = -w * G[k];
    cofactor(Fc, cof);
    float ic = 0.f, PF = 0.f, Pcof = 0.f;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      ic += Fc[k] * Fc[k];
      PF += Pb[k] * Fc[k];
      Pcof += Pb[k] * cof[k];
    }
    const float phi = 1.f - 1.f / (ic + 1.f);
    const float dphi = 1.f / ((ic + 1.f) * (ic + 1.f));
    const float detF = det3(Fc);
    const float J = fmaxf(detF, 0.01f);
    float Fcb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) Fcb[k] = k_mu * phi * Pb[k] + 2.f * k_mu * dphi * PF * Fc[k];
    if (detF > 0.01f) {
#pragma unroll
      for (int k = 0; k < 9; ++k) Fcb[k] += k_lambda * Pcof * cof[k];
    }
    const v3 f0 = {Fc[0], Fc[3], Fc[6]}, f1 = {Fc[1], Fc[4], Fc[7]}, f2 = {Fc[2], Fc[5], Fc[8]};
    const v3 p0 = {Pb[0], Pb[3], Pb[6]}, p1 = {Pb[1], Pb[4], Pb[7]}, p2 = {Pb[2], Pb[5], Pb[8]};
    const float sc = k_lambda * (J - alpha);
    const v3 g0 = (cross(p1, f2) + cross(f1, p2)) * sc, g1 = (cross(p2, f0) + cross(f2, p0)) * sc,
             g2 = (cross(p0, f1) + cross(f0, p1)) * sc;
    Fcb[0] += g0.x; Fcb[3] += g0.y; Fcb[6] += g0.z;
    Fcb[1] += g1.x; Fcb[4] += g1.y; Fcb[7] += g1.z;
    Fcb[2] += g2.x; Fcb[5] += g2.y; Fcb[8] += g2.z;
    float Fb[9];
    clamp_stretch_grad(F, a.mn, a.mx, Fcb, Fb);
    const v3 dNa = hex_grad_corner(a.hex_invJ + 72 * (size_t)e + 9 * g, g, la);
    xb = xb + mv(Fb, dNa);
    vb = vb + mv(Pb, dNa) * k_damp;
  }
}

template <int G>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// ---------------------------------------------------------------------------
// k_assemble: A = M + h^2 K at positions x, and M^-1 = inv(diag blocks).
// Atomic-free and deterministic.  A workgroup owns the consecutive rows
// groups[blk] .. groups[blk + 1] - 1 (at most RB rows, packed on the host so
// that their incidences fit one pass of the workgroup's threads and the
// staging fits ASM_LDS_BUDGET bytes).
// Phase 1: every incidence (element, corner) of those rows computes its
// element-tangent row (4 blocks for a Tet4, 8 for a Hex8) into LDS staging
// with plain stores.  Phase 2: one thread per BSR slot of the rows sums, in
// incidence order, the staged blocks whose precomputed slot byte matches,
// adds the lumped mass on the diagonal, writes the slot (consecutive threads
// -> consecutive 36-byte blocks) and inverts the diagonal block.
// ---------------------------------------------------------------------------
// Staging strides (floats) per incidence: odd so that consecutive incidences
// fall in different LDS banks.
constexpr int ST_TET = 37, ST_HEX = 73;
// Dynamic LDS of one assembly workgroup: four fit on a CU beside the static
// row arrays.
constexpr size_t ASM_LDS_BUDGET = 38 * 1024;

// EL_TET kernels carry no Hex8, spring or triangle code (and so none of its
// register pressure); EL_ANY kernels handle every element kind.
enum { EL_TET = 1, EL_ANY = 7 };

// Per-group record of the assembly (host-built, one scalar load per
// workgroup instead of a chain of dependent index loads): first/end row,
// first/end tet incidence, first/end hex incidence, first/end slot, first/end
// contribution.
constexpr int ASM_HDR = 12;

template <int RB, int EL>
__global__ __launch_bounds__(BLOCK) void k_assemble(RowArgs a, const int* __restrict__ groups,
                                                    const float* __restrict__ x, float h, float h2,
                                                    float* __restrict__ vals, float* __restrict__ minv) {
  extern __shared__ float lds[];
  __shared__ int roff[RB + 1], rtb[RB], rhb[RB], rdiag[RB];
  __shared__ float rmass[RB];
  const int blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int4* hdr = reinterpret_cast<const int4*>(groups + ASM_HDR * (size_t)blk);
  const int4 g0 = hdr[0], g1 = hdr[1], g2 = hdr[2];
  const int r0 = g0.x, r1 = g0.y;
  if (r0 >= r1) return;  // uniform over the workgroup (padding groups are empty)
  const int nr = r1 - r0;
  const int tb = g0.z, nt = g0.w - tb;
  const int hb = g1.x, nh = g1.y - hb;
  const int kb = g1.z, ke = g1.w;
  const int cb = g2.x, nc = g2.y - cb;
  float* st_t = lds;                                                  // [nt][ST_TET]: 4 blocks
  float* st_h = st_t + ST_TET * (size_t)nt;                           // [nh][ST_HEX]: 8 blocks
  int* sptr = reinterpret_cast<int*>(st_h + ST_HEX * (size_t)nh);     // [ke-kb+1] slot -> first contribution
  uint16_t* codes = reinterpret_cast<uint16_t*>(sptr + (ke - kb + 1));  // [nc]
  uint8_t* srow = reinterpret_cast<uint8_t*>(codes + nc);             // [ke-kb] slot -> local row
  // The first incidence of every thread is fetched ahead of the row set-up
  // (its code, then its nodes and element data), so that the element pass's
  // chain of dependent loads overlaps the set-up's loads.
  const bool first = (int)threadIdx.x < nt;
  const int code_f = first ? a.tinc[tb + (int)threadIdx.x] : 0;
  // Row metadata: loads with clamped indices (no divergent branch before
  // they are issued), stores for the group's rows only.
  const int rl = min((int)threadIdx.x, nr), rr = min((int)threadIdx.x, nr - 1);
  const int off_l = a.offsets[r0 + rl], off_n = a.offsets[r0 + rr + 1];
  const int tptr = a.tinc_ptr[r0 + rr], hptr = a.hinc_ptr[r0 + rr], drel = a.diag_rel[r0 + rr];
  const float mrow = a.mass[r0 + rr];
  const int e_f = code_f >> 2;
  const int4 nd_f = first ? *reinterpret_cast<const int4*>(a.tet_idx + 4 * (size_t)e_f) : make_int4(0, 0, 0, 0);
  const v3 zf = {0.f, 0.f, 0.f};
  const v3 x0_f = first ? ld3(x, nd_f.x) : zf, x1_f = first ? ld3(x, nd_f.y) : zf;
  const v3 x2_f = first ? ld3(x, nd_f.z) : zf, x3_f = first ? ld3(x, nd_f.w) : zf;
  if (threadIdx.x <= nr) roff[threadIdx.x] = off_l - kb;
  if (threadIdx.x < nr) {
    rtb[threadIdx.x] = tptr - tb;
    rhb[threadIdx.x] = hptr - hb;
    rdiag[threadIdx.x] = drel;
    rmass[threadIdx.x] = mrow;
    const int s0 = off_l - kb, s1 = off_n - kb;
    for (int k = s0; k < s1; ++k) srow[k] = (uint8_t)threadIdx.x;
  }
  // Phase 2's slot pointers and contribution codes: the first ASM_PF per
  // thread are loaded into registers now, in flight with phase 1's element
  // loads, and stored to LDS after phase 1; the rest (large groups) after.
  constexpr int ASM_PF = 4;
  int pf_ptr = 0;
  uint16_t pf_code[ASM_PF];
  if ((int)threadIdx.x <= ke - kb) pf_ptr = a.slot_ptr[kb + threadIdx.x];
#pragma unroll
  for (int u = 0; u < ASM_PF; ++u) {
    const int j = threadIdx.x + u * BLOCK;
    pf_code[u] = j < nc ? a.contrib[cb + j] : (uint16_t)0;
  }
  for (int t = threadIdx.x; t < nt; t += BLOCK) {
    const bool pre = t == (int)threadIdx.x;
    const int code = pre ? code_f : a.tinc[tb + t];
    const int e = code >> 2, la = code & 3;
    TetGeo g;
    if (pre) {
      tet_geo(x0_f, x1_f, x2_f, x3_f, a.tet_pose + 9 * (size_t)e, a.mn, a.mx, g);
    } else {
      const int4 nd = *reinterpret_cast<const int4*>(a.tet_idx + 4 * (size_t)e);
      tet_geo(ld3(x, nd.x), ld3(x, nd.y), ld3(x, nd.z), ld3(x, nd.w), a.tet_pose + 9 * (size_t)e, a.mn, a.mx, g);
    }
    float K[4][9];
    tet_tangent_row(g, a.tet_mat[3 * e], a.tet_mat[3 * e + 1], h2, la, K);
    float* o = st_t + ST_TET * t;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int c = 0; c < 9; ++c) o[9 * b + c] = K[b][c];
  }
  if (EL & 2) for (int t = threadIdx.x; t < nh; t += BLOCK) {
    const int code = a.hinc[hb + t];
    float K[8][9];
    hex_tangent_row(a, code >> 3, code & 7, h2, K);
    float* o = st_h + ST_HEX * t;
#pragma unroll
    for (int b = 0; b < 8; ++b)
#pragma unroll
      for (int c = 0; c < 9; ++c) o[9 * b + c] = K[b][c];
  }
  if ((int)threadIdx.x <= ke - kb) sptr[threadIdx.x] = pf_ptr - cb;
  for (int k = threadIdx.x + BLOCK; k <= ke - kb; k += BLOCK) sptr[k] = a.slot_ptr[kb + k] - cb;
#pragma unroll
  for (int u = 0; u < ASM_PF; ++u) {
    const int j = threadIdx.x + u * BLOCK;
    if (j < nc) codes[j] = pf_code[u];
  }
  for (int j = threadIdx.x + ASM_PF * BLOCK; j < nc; j += BLOCK) codes[j] = a.contrib[cb + j];
  __syncthreads();
  for (int kk = threadIdx.x; kk < ke - kb; kk += BLOCK) {
    const int lr = srow[kk];
    const bool diag = (kk - roff[lr]) == rdiag[lr];
    float acc[9];
    const float m = diag ? rmass[lr] : 0.f;
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[c] = (c == 0 || c == 4 || c == 8) ? m : 0.f;
    if (EL & 4) {
      // Constant blocks in the reference's triplet order after the mass:
      // springs (diagonal m + sum(h kd + h^2 ke), off-diagonal -(h kd + h^2 ke),
      // kernels.py:294-366), then the lumped triangle diagonal
      // h^2 (k_mu + k_lambda) A / 3 (kernels.py:972-1030).
      float cst = 0.f;
      if (a.slot_spring) {
        const float2 sk = a.slot_spring[kb + kk];  // signed sums
        cst = h * sk.y + h2 * sk.x;
      }
      if (diag && a.row_tri) cst += h2 * a.row_tri[r0 + lr] / 3.f;
      acc[0] += cst;
      acc[4] += cst;
      acc[8] += cst;
    }
    const int tbase = rtb[lr], hbase = rhb[lr];
    auto src_of = [&](uint32_t code) -> const float* {
      const uint32_t t = (code >> 3) & 4095u, b = code & 7u;
      return (code & 0x8000u) ? st_h + ST_HEX * (hbase + t) + 9 * b : st_t + ST_TET * (tbase + t) + 9 * b;
    };
    // Contributions in list order, four at a time: their codes and blocks are
    // read together (independent LDS reads), the sums stay sequential.
    const int j1 = sptr[kk + 1];
    int j = sptr[kk];
    for (; j + 4 <= j1; j += 4) {
      const float* s0 = src_of(codes[j]);
      const float* s1 = src_of(codes[j + 1]);
      const float* s2 = src_of(codes[j + 2]);
      const float* s3 = src_of(codes[j + 3]);
      float v[4][9];
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        v[0][c] = s0[c];
        v[1][c] = s1[c];
        v[2][c] = s2[c];
        v[3][c] = s3[c];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int c = 0; c < 9; ++c) acc[c] += v[u][c];
    }
    for (; j < j1; ++j) {
      const float* src = src_of(codes[j]);
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[c] += src[c];
    }
    float* out = vals + 9 * (size_t)(kb + kk);
#pragma unroll
    for (int c = 0; c < 9; ++c) out[c] = acc[c];
    if (diag) {
      float inv[9];
      inverse_diagonal_block(acc, inv);
#pragma unroll
      for (int c = 0; c < 9; ++c) minv[9 * (size_t)(r0 + lr) + c] = inv[c];
    }
  }
  // Rows without a stored diagonal use the identity (kernels.py:1840-1843).
  for (int lr = threadIdx.x; lr < nr; lr += BLOCK) {
    if (rdiag[lr] < 0) {
      const float eye[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
      float inv[9];
      inverse_diagonal_block(eye, inv);
#pragma unroll
      for (int c = 0; c < 9; ++c) minv[9 * (size_t)(r0 + lr) + c] = inv[c];
    }
  }
}

// ---------------------------------------------------------------------------
// k_rhs: step right-hand side (STEP) or plain force evaluation (EVAL).
// ---------------------------------------------------------------------------
// MODE_LOAD: the right-hand side of a step whose h-scaled force sum the
// caller provides in `ext` (implicit_integration): ext - predictor terms.
enum { MODE_EVAL = 0, MODE_STEP = 1, MODE_LOAD = 2 };

template <int G, int EL>
__global__ __launch_bounds__(BLOCK) void k_rhs(RowArgs a, const float* __restrict__ x, const float* __restrict__ v,
                                               const float* __restrict__ act, const float* __restrict__ ext,
                                               const float* __restrict__ vals, float* __restrict__ out,
                                               unsigned* rhs_max, int terms, int mode, float h) {
  constexpr int NG = BLOCK / G;
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  const int blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int i = blk * NG + grp;
  const bool active = i < a.P;
  v3 ft = {0.f, 0.f, 0.f}, fh = {0.f, 0.f, 0.f}, corr = {0.f, 0.f, 0.f};
  v3 fsp = {0.f, 0.f, 0.f}, ftri = {0.f, 0.f, 0.f}, srest = {0.f, 0.f, 0.f};
  if (active) {
    if ((EL & 4) && a.sinc_ptr) {
      // eval_springs kernels.py:176-219 and restore_spring_damping_predictor
      // implicit_predictor.py:26-41 (h kd (v_node - v_other) per spring).
      const int s1 = a.sinc_ptr[i + 1];
      for (int t = a.sinc_ptr[i] + lane; t < s1; t += G) {
        const int code = a.sinc[t];
        const int sp = code >> 1, end = code & 1;
        const int p = a.sp_idx[2 * sp], q = a.sp_idx[2 * sp + 1];
        const v3 vp = ld3(v, p), vq = ld3(v, q);
        const float kd = a.sp_kd[sp];
        if (terms & FEM_TERM_SPRING) {
          const v3 xp = ld3(x, p), xq = ld3(x, q);
          fsp = fsp + spring_end_force(xp, xq, vp, vq, a.sp_rest[sp], a.sp_ke[sp], kd, end);
          if (a.sp_dir)
            fsp = fsp + spring_torque_end_force(xp, xq, vp, vq, ld3(a.sp_dir, sp), a.sp_kt, a.sp_kw, end);
        }
        srest = srest + (end == 0 ? vp - vq : vq - vp) * (h * kd);
      }
    }
    if ((EL & 4) && a.rinc_ptr && (terms & FEM_TERM_TRI)) {
      const int r1 = a.rinc_ptr[i + 1];
      for (int t = a.rinc_ptr[i] + lane; t < r1; t += G) {
        const int code = a.rinc[t];
        const int e = code >> 2, c = code & 3;
        const int n0 = a.tri_idx[3 * e], n1 = a.tri_idx[3 * e + 1], n2 = a.tri_idx[3 * e + 2];
        ftri = ftri + tri_corner_force(ld3(x, n0), ld3(x, n1), ld3(x, n2), ld3(v, n0), ld3(v, n1), ld3(v, n2),
                                       a.tri_pose + 4 * (size_t)e, a.tri_mat + 5 * (size_t)e, c);
      }
    }
    if (terms & FEM_TERM_TET) {
      const int t1 = a.tinc_ptr[i + 1];
      for (int t = a.tinc_ptr[i] + lane; t < t1; t += G) {
        const int code = a.tinc[t];
        const int e = code >> 2, la = code & 3;
        const int4 nd = *reinterpret_cast<const int4*>(a.tet_idx + 4 * (size_t)e);
        const v3 x0 = ld3(x, nd.x), x1 = ld3(x, nd.y), x2 = ld3(x, nd.z), x3 = ld3(x, nd.w);
        const float mu = a.tet_mat[3 * e], lam = a.tet_mat[3 * e + 1], kd = a.tet_mat[3 * e + 2];
        TetGeo g;
        tet_geo(x0, x1, x2, x3, a.tet_pose + 9 * (size_t)e, a.mn, a.mx, g);
        ft = ft + tet_node_force(g, ld3(v, nd.x), ld3(v, nd.y), ld3(v, nd.z), ld3(v, nd.w), mu, lam, kd, la);
        if (a.use_act && act) ft = ft + tet_activation_force(x0, x1, x2, x3, lam, act[e], la);
      }
    }
    if ((EL & 2) && (terms & FEM_TERM_HEX)) {
      const int h1 = a.hinc_ptr[i + 1];
      for (int t = a.hinc_ptr[i] + lane; t < h1; t += G) {
        const int code = a.hinc[t];
        fh = fh + hex_node_force(a, code >> 3, code & 7);
      }
    }
    if (mode != MODE_EVAL) {
      const v3 vi = ld3(v, i);
      const int k1 = a.offsets[i + 1];
      for (int k = a.offsets[i] + lane; k < k1; k += G) {
        const float* b = vals + 9 * (size_t)k;
        v3 d = ld3(v, a.cols[k]) - vi;
        corr = corr + v3{b[0] * d.x + b[1] * d.y + b[2] * d.z, b[3] * d.x + b[4] * d.y + b[5] * d.z,
                         b[6] * d.x + b[7] * d.y + b[8] * d.z};
      }
    }
  }
  ft = {group_sum<G>(ft.x), group_sum<G>(ft.y), group_sum<G>(ft.z)};
  fh = {group_sum<G>(fh.x), group_sum<G>(fh.y), group_sum<G>(fh.z)};
  if (EL & 4) {
    fsp = {group_sum<G>(fsp.x), group_sum<G>(fsp.y), group_sum<G>(fsp.z)};
    ftri = {group_sum<G>(ftri.x), group_sum<G>(ftri.y), group_sum<G>(ftri.z)};
    srest = {group_sum<G>(srest.x), group_sum<G>(srest.y), group_sum<G>(srest.z)};
  }
  if (mode != MODE_EVAL) corr = {group_sum<G>(corr.x), group_sum<G>(corr.y), group_sum<G>(corr.z)};
  float mval = 0.f;
  if (active && lane == 0) {
    const bool on = (a.flags[i] & 1) != 0;
    const v3 xi = ld3(x, i), vi = ld3(v, i);
    v3 fg = {0.f, 0.f, 0.f}, fgr = {0.f, 0.f, 0.f}, fdm = {0.f, 0.f, 0.f};
    // Particle-particle contact, evaluated over the hash grid before this
    // kernel (zero for inactive particles).
    const v3 fpc = (a.fpart && (terms & FEM_TERM_PARTICLE)) ? ld3(a.fpart, i) : v3{0.f, 0.f, 0.f};
    const v3 fsc = (a.fsoft && (terms & FEM_TERM_SOFT)) ? ld3(a.fsoft, i) : v3{0.f, 0.f, 0.f};
    if ((terms & FEM_TERM_GROUND) && a.has_ground && on) {
      // eval_particle_ground_contacts kernels.py:1132-1195
      const v3 n = {a.gn[0], a.gn[1], a.gn[2]};
      const float c = fminf(dot(n, xi) + a.gn[3] - a.radius[i], 0.f);
      const float vn = dot(n, vi);
      const float jn = c * a.ke;
      if (c < 0.f) {
        const float jd = fminf(vn, 0.f) * a.kd;
        const float fn = jn + jd;
        v3 vt = vi - n * vn;
        const float vs = sqrtf(dot(vt, vt));
        if (vs > 0.f) vt = vt * (1.f / vs);
        const float inv_m = a.inv_mass[i];
        const float m = inv_m > 0.f ? 1.f / fmaxf(inv_m, 1e-9f) : 0.f;
        const float gmag = a.gravity ? sqrtf(a.gravity[0] * a.gravity[0] + a.gravity[1] * a.gravity[1] +
                                             a.gravity[2] * a.gravity[2])
                                     : a.gmag;
        const float fn_eff = fminf(fabsf(fn), m * gmag);
        const float ftan = fminf(vs * a.kf, a.gmu * fn_eff);
        fg = fg - n * fn - vt * ftan;
      }
    }
    if ((terms & FEM_TERM_GRAVITY) && on) {
      const float m = a.mass[i];
      if (a.gravity) fgr = {a.gravity[0] * m, a.gravity[1] * m, a.gravity[2] * m};
      else fgr = {a.grav[0] * m, a.grav[1] * m, a.grav[2] * m};
    }
    if ((terms & FEM_TERM_DAMPING) && on && a.lin_damp > 0.f) fdm = vi * (-a.lin_damp * a.mass[i]);
    v3 f;
    if (mode == MODE_STEP) {
      // _accumulate_forces order solver.py:570-589 (spring, tet, hex,
      // triangle, particle, ground, soft contact, gravity, damping), then h*ext
      // (:993-999), then subtract_elastic_predictor and
      // restore_spring_damping_predictor implicit_predictor.py:9-41.
      f = {0.f, 0.f, 0.f};
      if (EL & 4) f = f + fsp * h;
      f = f + ft * h;
      f = f + fh * h;
      if (EL & 4) f = f + ftri * h;
      if (a.fpart) f = f + fpc * h;
      f = f + fg * h;
      if (a.fsoft) f = f + fsc * h;
      f = f + fgr * h;
      f = f + fdm * h;
      if (ext) f = f + ld3(ext, i) * h;
      f = f - corr;
      if (EL & 4) f = f + srest;
      if (a.row_fixed[i]) f = {0.f, 0.f, 0.f};
      mval = fmaxf(fabsf(f.x), fmaxf(fabsf(f.y), fabsf(f.z)));
    } else if (mode == MODE_LOAD) {
      // implicit_integration solver.py:1254-1307: the caller's h-scaled force
      // sum, minus the elastic predictor, plus the spring-damping restore, with
      // fixed rows projected out (FixedParticles.project_rhs).
      f = ld3(ext, i) - corr;
      if (EL & 4) f = f + srest;
      if (a.row_fixed[i]) f = {0.f, 0.f, 0.f};
      mval = fmaxf(fabsf(f.x), fmaxf(fabsf(f.y), fabsf(f.z)));
    } else {
      f = fsp + ft + fh + ftri + fpc + fg + fsc + fgr + fdm;
    }
    st3(out, i, f);
  }
  if (mode != MODE_EVAL) {
    // maximum_rhs_component kernels.py:1779-1782, per instance, spread over
    // RHS_MAX_SLOTS words so that no address sees more than a few atomics.
    if (a.row_inst == nullptr) {
      __shared__ float wmax[BLOCK / 64];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) mval = fmaxf(mval, __shfl_xor(mval, o));
      if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mval;
      __syncthreads();
      if (threadIdx.x == 0) {
        float m = wmax[0];
#pragma unroll
        for (int w = 1; w < BLOCK / 64; ++w) m = fmaxf(m, wmax[w]);
        if (m > 0.f) atomicMax(rhs_max + (blk & (RHS_MAX_SLOTS - 1)), __float_as_uint(m));
      }
    } else if (active && lane == 0 && mval > 0.f) {
      atomicMax(rhs_max + RHS_MAX_SLOTS * a.row_inst[i] + (i & (RHS_MAX_SLOTS - 1)), __float_as_uint(mval));
    }
  }
}

// ---------------------------------------------------------------------------
// Adjoint of the differentiable step (differentiation.py:289-394): for the
// solved adjoint z = A^-1 dv_bar, the reverse products of
// rhs = h * sum_terms(x, v, act) + h * ext - sum_j A_ij (v_j - v_i).
// ---------------------------------------------------------------------------
// Per node i (G lanes over its element incidences): xbar_i = h dT/dx^T z and
// vbar_i = h dT/dv^T z - (A z)_i + (sum_k A_ik)^T z_i.  Written, not added.
template <int G, int EL>
__global__ __launch_bounds__(BLOCK) void k_vjp_rows(RowArgs a, const float* __restrict__ x,
                                                    const float* __restrict__ v, const float* __restrict__ act,
                                                    const float* __restrict__ z, const float* __restrict__ vals,
                                                    float h, float* __restrict__ xbar, float* __restrict__ vbar) {
  constexpr int NG = BLOCK / G;
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  const int blk = xcd_swizzle(blockIdx.x, gridDim.x);
  const int i = blk * NG + grp;
  const bool active = i < a.P;
  v3 xb = {0.f, 0.f, 0.f}, vb = {0.f, 0.f, 0.f}, az = {0.f, 0.f, 0.f};
  float rs[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) rs[k] = 0.f;
  if (active) {
    const int t1 = a.tinc_ptr[i + 1];
    for (int t = a.tinc_ptr[i] + lane; t < t1; t += G) {
      const int code = a.tinc[t];
      const int e = code >> 2, la = code & 3;
      const int4 nd = *reinterpret_cast<const int4*>(a.tet_idx + 4 * (size_t)e);
      const TetVjp o = tet_vjp(ld3(x, nd.x), ld3(x, nd.y), ld3(x, nd.z), ld3(x, nd.w), ld3(v, nd.x), ld3(v, nd.y),
                               ld3(v, nd.z), ld3(v, nd.w), ld3(z, nd.x) * h, ld3(z, nd.y) * h, ld3(z, nd.z) * h,
                               ld3(z, nd.w) * h, a.tet_pose + 9 * (size_t)e, a.tet_mat[3 * e], a.tet_mat[3 * e + 1],
                               a.tet_mat[3 * e + 2], a.use_act && act, act ? act[e] : 0.f, a.mn, a.mx);
      const v3 x0 = (o.xe[0] + o.xe[1] + o.xe[2]) * -1.f, v0 = (o.ve[0] + o.ve[1] + o.ve[2]) * -1.f;
      xb = xb + sel4(la, x0, o.xe[0], o.xe[1], o.xe[2]);
      vb = vb + sel4(la, v0, o.ve[0], o.ve[1], o.ve[2]);
    }
    if (EL & 2) {
      const int h1 = a.hinc_ptr[i + 1];
      for (int t = a.hinc_ptr[i] + lane; t < h1; t += G) {
        const int code = a.hinc[t];
        v3 hx, hv;
        hex_node_vjp(a, x, z, h, code >> 3, code & 7, hx, hv);
        xb = xb + hx;
        vb = vb + hv;
      }
    }
    if ((EL & 4) && a.sinc_ptr) {
      // Reverse of eval_springs for this endpoint: L = sum_s w . fs with
      // w = h (z_q - z_p), fs = dir (ke (len - rest) + kd dir . v_pq); and of
      // the spring-damping predictor restore, h kd (z_p - z_q) . (v_p - v_q).
      const int s1 = a.sinc_ptr[i + 1];
      for (int t = a.sinc_ptr[i] + lane; t < s1; t += G) {
        const int code = a.sinc[t];
        const int sp = code >> 1;
        const float sg = (code & 1) ? -1.f : 1.f;  // +1 at the first endpoint p, -1 at q
        const int p = a.sp_idx[2 * sp], q = a.sp_idx[2 * sp + 1];
        const float kd = a.sp_kd[sp];
        const v3 zp = ld3(z, p), zq = ld3(z, q);
        vb = vb + (zp - zq) * (sg * h * kd);
        const v3 xpq = ld3(x, p) - ld3(x, q), vpq = ld3(v, p) - ld3(v, q);
        const float len = sqrtf(dot(xpq, xpq));
        if (len < 1.0e-6f) continue;
        const v3 d = xpq * (1.f / len);
        const float ke = a.sp_ke[sp];
        const float dvel = dot(d, vpq);
        const float phi = ke * (len - a.sp_rest[sp]) + kd * dvel;
        const v3 w = (zq - zp) * h;
        const float dw = dot(d, w);
        const v3 gx = (w - d * dw) * (phi / len) + (d * ke + (vpq - d * dvel) * (kd / len)) * dw;
        xb = xb + gx * sg;
        vb = vb + d * (sg * dw * kd);
      }
    }
    if ((EL & 4) && a.rinc_ptr) {
      // Reverse of the triangle membranes (weights h z of the corner forces).
      const int r1 = a.rinc_ptr[i + 1];
      for (int t = a.rinc_ptr[i] + lane; t < r1; t += G) {
        const int code = a.rinc[t];
        const int e = code >> 2, c = code & 3;
        const int n0 = a.tri_idx[3 * e], n1 = a.tri_idx[3 * e + 1], n2 = a.tri_idx[3 * e + 2];
        v3 gx10, gx20, gv10, gv20;
        const v3 x0 = ld3(x, n0), x1 = ld3(x, n1), x2 = ld3(x, n2), v0 = ld3(v, n0), v1 = ld3(v, n1),
                 v2 = ld3(v, n2), w0 = ld3(z, n0) * h, w1 = ld3(z, n1) * h, w2 = ld3(z, n2) * h;
        const float* mat = a.tri_mat + 5 * (size_t)e;
        tri_edge_vjp(x0, x1, x2, v0, v1, v2, w0, w1, w2, a.tri_pose + 4 * (size_t)e, mat, gx10, gx20, gv10, gv20);
        // drag and lift: the same velocity gradient for every corner
        vb = vb + tri_aero_vjp(x0, x1, x2, v0, v1, v2, w0 + w1 + w2, mat, gx10, gx20);
        if (c == 0) {
          xb = xb - gx10 - gx20;
          vb = vb - gv10 - gv20;
        } else if (c == 1) {
          xb = xb + gx10;
          vb = vb + gv10;
        } else {
          xb = xb + gx20;
          vb = vb + gv20;
        }
      }
    }
    const int k1 = a.offsets[i + 1];
    for (int k = a.offsets[i] + lane; k < k1; k += G) {
      const float* b = vals + 9 * (size_t)k;
      const v3 zc = ld3(z, a.cols[k]);
      az = az + v3{b[0] * zc.x + b[1] * zc.y + b[2] * zc.z, b[3] * zc.x + b[4] * zc.y + b[5] * zc.z,
                   b[6] * zc.x + b[7] * zc.y + b[8] * zc.z};
#pragma unroll
      for (int q = 0; q < 9; ++q) rs[q] += b[q];
    }
  }
  xb = {group_sum<G>(xb.x), group_sum<G>(xb.y), group_sum<G>(xb.z)};
  vb = {group_sum<G>(vb.x), group_sum<G>(vb.y), group_sum<G>(vb.z)};
  az = {group_sum<G>(az.x), group_sum<G>(az.y), group_sum<G>(az.z)};
#pragma unroll
  for (int q = 0; q < 9; ++q) rs[q] = group_sum<G>(rs[q]);
  if (active && lane == 0) {
    const v3 zi = ld3(z, i);
    if ((a.flags[i] & 1) != 0) {
      if (a.has_ground) {
        v3 gx, gv;
        const float gmag = a.gravity ? sqrtf(a.gravity[0] * a.gravity[0] + a.gravity[1] * a.gravity[1] +
                                             a.gravity[2] * a.gravity[2])
                                     : a.gmag;
        ground_vjp(ld3(x, i), ld3(v, i), zi * h, a.radius[i], a.inv_mass[i], a.gn, a.ke, a.kd, a.kf, a.gmu, gmag,
                   gx, gv);
        xb = xb + gx;
        vb = vb + gv;
      }
      if (a.lin_damp > 0.f) vb = vb + zi * (-a.lin_damp * a.mass[i] * h);
    }
    const v3 rt = {rs[0] * zi.x + rs[3] * zi.y + rs[6] * zi.z, rs[1] * zi.x + rs[4] * zi.y + rs[7] * zi.z,
                   rs[2] * zi.x + rs[5] * zi.y + rs[8] * zi.z};
    vb = vb - az + rt;
    st3(xbar, i, xb);
    st3(vbar, i, vb);
  }
}

// Per Tet4: dL/d activation and dL/d(mu, lambda, k_damp), the latter with the
// matrix terms -z^T (dA/dmu) (dv + v) of the fixed matrix assembled at xa.
__global__ __launch_bounds__(BLOCK) void k_vjp_tets(RowArgs a, int T, const float* __restrict__ x,
                                                    const float* __restrict__ v, const float* __restrict__ act,
                                                    const float* __restrict__ z, const float* __restrict__ dv,
                                                    const float* __restrict__ xa, float h,
                                                    float* __restrict__ act_bar, float* __restrict__ mat_bar) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= T) return;
  const int4 nd = *reinterpret_cast<const int4*>(a.tet_idx + 4 * (size_t)e);
  const float mu = a.tet_mat[3 * e], lam = a.tet_mat[3 * e + 1], kd = a.tet_mat[3 * e + 2];
  const v3 z0 = ld3(z, nd.x), z1 = ld3(z, nd.y), z2 = ld3(z, nd.z), z3 = ld3(z, nd.w);
  const v3 v0 = ld3(v, nd.x), v1 = ld3(v, nd.y), v2 = ld3(v, nd.z), v3_ = ld3(v, nd.w);
  const TetVjp o = tet_vjp(ld3(x, nd.x), ld3(x, nd.y), ld3(x, nd.z), ld3(x, nd.w), v0, v1, v2, v3_, z0 * h, z1 * h,
                           z2 * h, z3 * h, a.tet_pose + 9 * (size_t)e, mu, lam, kd, a.use_act && act,
                           act ? act[e] : 0.f, a.mn, a.mx);
  if (act_bar) act_bar[e] = o.act;
  if (mat_bar) {
    float mub = 0.f, lamb = 0.f;
    if (!(mu == 0.f && lam == 0.f && kd == 0.f)) {
      const v3 y0 = ld3(dv, nd.x) + v0, y1 = ld3(dv, nd.y) + v1, y2 = ld3(dv, nd.z) + v2, y3 = ld3(dv, nd.w) + v3_;
      tet_matrix_material_terms(ld3(xa, nd.x), ld3(xa, nd.y), ld3(xa, nd.z), ld3(xa, nd.w), z1 - z0, z2 - z0, z3 - z0,
                                y1 - y0, y2 - y0, y3 - y0, a.tet_pose + 9 * (size_t)e, a.mn, a.mx, h * h, mub, lamb);
    }
    mat_bar[3 * (size_t)e] = o.mat[0] + mub;
    mat_bar[3 * (size_t)e + 1] = o.mat[1] + lamb;
    mat_bar[3 * (size_t)e + 2] = o.mat[2];
  }
}

// max |b| per instance into the 64 spread words (maximum_rhs_component).
__global__ __launch_bounds__(BLOCK) void k_absmax(const float* __restrict__ b, int P, const int* __restrict__ row_inst,
                                                  unsigned* rhs_max) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  float m = 0.f;
  if (i < P) {
    const v3 bi = ld3(b, i);
    m = fmaxf(fabsf(bi.x), fmaxf(fabsf(bi.y), fabsf(bi.z)));
  }
  if (row_inst) {
    if (i < P && m > 0.f) atomicMax(rhs_max + RHS_MAX_SLOTS * row_inst[i] + (i & (RHS_MAX_SLOTS - 1)), __float_as_uint(m));
    return;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0 && m > 0.f)
    atomicMax(rhs_max + ((blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)) & (RHS_MAX_SLOTS - 1)), __float_as_uint(m));
}

__global__ void k_zero_doubles(double* p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

// FP64 ||A x - b||^2 and ||b||^2 (residual_norms differentiation.py:118-133),
// with the solve's operator: PAP + I - P when there are fixed rows.
__global__ __launch_bounds__(BLOCK) void k_resnorm(const int* __restrict__ offsets, const int* __restrict__ cols,
                                                   const float* __restrict__ vals,
                                                   const unsigned char* __restrict__ row_fixed,
                                                   const float* __restrict__ x, const float* __restrict__ b, int P,
                                                   double* norms) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double rr = 0.0, bb = 0.0;
  if (i < P) {
    v3 ax = {0.f, 0.f, 0.f};
    if (row_fixed[i]) {
      ax = ld3(x, i);
    } else {
      for (int k = offsets[i]; k < offsets[i + 1]; ++k)
        if (!row_fixed[cols[k]]) ax = ax + mv(vals + 9 * (size_t)k, ld3(x, cols[k]));
    }
    const v3 bi = ld3(b, i);
    const double r0 = (double)ax.x - bi.x, r1 = (double)ax.y - bi.y, r2 = (double)ax.z - bi.z;
    rr = r0 * r0 + r1 * r1 + r2 * r2;
    bb = (double)bi.x * bi.x + (double)bi.y * bi.y + (double)bi.z * bi.z;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    rr += __shfl_xor(rr, o);
    bb += __shfl_xor(bb, o);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(norms, rr);
    atomicAdd(norms + 1, bb);
  }
}

// record_residual differentiation.py:136-146: running max of the relative
// residual (stats[kind]) and of residual / (2e-5 ||b|| + 1e-10) (stats[2+kind]).
__global__ void k_record(const double* norms, int kind, double* stats) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double r = sqrt(norms[0]), sc = sqrt(norms[1]);
  double ratio = r / fmax(sc, 1.0e-30);
  double violation = r / (2.0e-5 * sc + 1.0e-10);
  if (!isfinite(r) || !isfinite(sc)) ratio = violation = 1.0e30;
  // Non-negative doubles order like their bit patterns.
  atomicMax(reinterpret_cast<unsigned long long*>(stats) + kind, (unsigned long long)__double_as_longlong(ratio));
  atomicMax(reinterpret_cast<unsigned long long*>(stats) + 2 + kind,
            (unsigned long long)__double_as_longlong(violation));
}

// ---------------------------------------------------------------------------
// Persistent PCG.
// ---------------------------------------------------------------------------
struct CgArgs {
  int P;
  const int* offsets;
  const int* cols;
  const float* vals;
  const float* minv;
  const unsigned char* row_fixed;
  const float* rhs;
  float* dv;
  const float* q_in;
  const float* qd_in;
  float* q_out;
  float* qd_out;
  float* x_out;  // solve mode: x = A^-1 rhs from a zero start, no state update
  const float* anchor;
  float h;
  float* xw;  // 2 x P exchange records (double buffered)
  const int* chunk_rows;
  const int* chunk_inst;
  const int* inst_chunks;
  const int* chunk_tiles;  // [nchunks+1] first tile of each chunk
  const int* tile_rows;    // [ntiles] first row of each tile (row-aligned, <= TILE blocks)
  const int* tile_cptr;    // [ntiles+1] offsets into tile_cols
  const int* tile_cols;    // distinct columns of each tile, ascending
  const uint16_t* lidx;    // [nnz] tile-local column number of each block
  // Resident variant (one tile per chunk): per thread {first block within the
  // chunk, block count, first row run} and its run-end mask; per row {first
  // run within the chunk, run count}.
  const int4* res_thr;
  const unsigned long long* res_mask;
  const int2* res_row;
  const unsigned* rhs_max;
  int chunk_base;          // first chunk of this launch group
  int nchunks_total;       // chunks of all groups (stride of the double-buffered partials)
  const int* chunk_order;  // queued launch: chunks in the order the queue hands them out (else null)
  float tolerance;
  int maxiter;
  int early_exit;
  double* part;
  unsigned* flags;  // [nchunks_total] last phase each chunk has published, then the chunk queue word
  int* stats;
  int* err;
  long spin_limit;  // sleep-polls a wait may take before it gives up
};

__device__ __forceinline__ unsigned ld_acq(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Spin until the monotonic word *p reaches target; bounded so that a launch
// whose workgroups are not all resident ends with an error flag instead of
// hanging.
__device__ __forceinline__ bool spin_until(const unsigned* p, unsigned target, int* err, long limit) {
  long spins = 0;
  while ((int)(ld_acq(p) - target) < 0) {
    __builtin_amdgcn_s_sleep(1);
    if (++spins > limit ||
        ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
      __hip_atomic_fetch_or(err, FEM_STATUS_BARRIER_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
  }
  return true;
}

// Synchronisation of the persistent CG: publish-and-wait, no cache maintenance.
// Everything a workgroup shares during a solve (exchange records, partial
// sums) is written with write-through stores (`sc1`: the store is
// acknowledged once it has reached memory, past the XCC's L2) and read by
// the others with non-temporal loads, which on gfx950 are not served from a
// stale L1 or L2 line of the reading XCC (checked with a cross-XCC probe:
// write-through store + non-temporal load 0/2000 stale; a plain or `nt`
// store stays invisible to other XCCs until an L2 write-back).  A phase
// ends with every wave waiting for its own stores, then one thread raising
// the chunk's flag to the phase's epoch; the next phase starts once every
// chunk of the instance has raised its flag.  No L2 write-back or invalidate
// is needed, so read-only data (inverse diagonal, streamed blocks, column
// lists) stays cached across iterations.
typedef float f3v __attribute__((ext_vector_type(3), aligned(4)));
// Four floats at 4-byte alignment (exchange records are packed at 36 bytes).
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ f4v ld_shared(const f4v* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ f4u ld_shared_u(const f4u* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ld_shared(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ double ld_shared(const double* p) { return __builtin_nontemporal_load(p); }
// The stores carry no memory clobber, so that independent loads (inverse
// diagonal blocks) may be scheduled across them; they stay ordered before
// publish() because both are volatile asm, and nothing in the kernel reads
// the stored data through a plain load.  The compiler's hazard recognizer
// does not see into the asm: a store of more than 64 bits needs two wait
// states before a VALU may overwrite its data registers (gfx940 family), so
// the 128-bit stores end with s_nop 2.
__device__ __forceinline__ void st_wt(float* p, f4v v) {
  asm volatile("global_store_dwordx4 %0, %1, off sc1\n s_nop 2" ::"v"(p), "v"(v));
}
__device__ __forceinline__ void st_wt(float* p, float v) {
  asm volatile("global_store_dword %0, %1, off sc1" ::"v"(p), "v"(v));
}
__device__ __forceinline__ void st_wt(double* p, d2v v) {
  asm volatile("global_store_dwordx4 %0, %1, off sc1\n s_nop 2" ::"v"(p), "v"(v));
}

// End of a phase: wait for this workgroup's shared stores, then raise its flag.
__device__ __forceinline__ void publish(unsigned* flag, unsigned epoch) {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0);
  asm volatile("" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of N floats over the workgroup (fp32 within waves, fp64 across waves,
// fixed order); thread 0 receives the result (`out` is unset elsewhere).
// One barrier: the LDS words `red` are written again only by wait_sum, which
// every thread reaches through publish()'s barrier, after thread 0 has read
// them here.
template <int NT, int N>
__device__ __forceinline__ void wg_sum(float (&v)[N], double (&out)[N], double* red) {
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[n] += __shfl_xor(v[n], o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int n = 0; n < N; ++n) red[w * N + n] = (double)v[n];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      double t = red[n];
#pragma unroll
      for (int w = 1; w < NT / 64; ++w) t += red[w * N + n];
      out[n] = t;
    }
  }
}

// Start of a phase: wait until every chunk of the instance has published
// `epoch`, and sum part[c*NPART + slot[n]] over those chunks, identical in
// every workgroup (same data, same reduction tree).  Thread t waits for
// chunks cb+t, cb+t+NT, ... and reads their sums as soon as their flags are
// up.  A single-chunk instance reads its own sums without waiting.  Returns
// false (every thread) when a wait timed out.  No trailing barrier: `red` is
// written again only by wg_sum, which every thread reaches through the
// SpMV's barriers, after all have read it here.
template <int NT, int N>
__device__ __forceinline__ bool wait_sum(const unsigned* flags, const double* part, int cb, int ce, unsigned epoch,
                                         const int (&slot)[N], double (&out)[N], double* red, int* err,
                                         long limit) {
  double acc[N];
#pragma unroll
  for (int n = 0; n < N; ++n) acc[n] = 0.0;
  int bad = 0;
  for (int c = cb + (int)threadIdx.x; c < ce; c += NT) {
    if (ce - cb > 1 && !spin_until(flags + c, epoch, err, limit)) {
      bad = 1;
      break;
    }
#pragma unroll
    for (int n = 0; n < N; ++n) acc[n] += ld_shared(part + NPART * (size_t)c + slot[n]);
  }
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc[n] += __shfl_xor(acc[n], o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int n = 0; n < N; ++n) red[w * N + n] = acc[n];
  bad = __syncthreads_or(bad);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double t = red[n];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) t += red[w * N + n];
    out[n] = t;
  }
  return bad == 0;
}

// Scratch of fem_step_solve_static (callers that cannot allocate): the
// persistent CG's partial sums, publication flags and exchange records live
// in this file-scope array instead of library-allocated memory.  Every GPU
// the library runs on holds its own copy until the process ends.  The host
// never uses the array's name as an address: its device address is taken
// once per device with hipGetSymbolAddress (static_scratch_ptr).
constexpr size_t STATIC_SCRATCH_BYTES = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

// From here to the end of the CG kernel no multiply-add is fused implicitly:
// every fused one is an explicit fmaf, so each kernel shape (rows per
// thread, queued or not) rounds every chunk alike and a chunk's solve does
// not depend on the launch it runs in.
#pragma clang fp contract(off)
__device__ __forceinline__ float dot_cg(v3 a, v3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ v3 mv_cg(const float m[9], v3 x) {
  return {fmaf(m[2], x.z, fmaf(m[1], x.y, m[0] * x.x)), fmaf(m[5], x.z, fmaf(m[4], x.y, m[3] * x.x)),
          fmaf(m[8], x.z, fmaf(m[7], x.y, m[6] * x.x))};
}
__device__ __forceinline__ v3 scale_cg(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ v3 sub_cg(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ v3 add_cg(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ v3 fma_cg(v3 a, float s, v3 b) { return {fmaf(a.x, s, b.x), fmaf(a.y, s, b.y), fmaf(a.z, s, b.z)}; }

constexpr int CG_J = 8;  // blocks per thread per tile (a tile is CG_J * NT blocks)
// Exchange record of one row: z(3) w(3) p(3), packed at 36 bytes (two 16-byte
// loads and one 4-byte load per gathered column; the halo gather moves about
// 2.5 records per row every iteration, so no padding is carried).
constexpr int XW = 9;

// q[k] = sum_j A_ij u_j for the thread's own rows i = R0 + t + NT k.
// Row-aligned tiles of <= (CG_J * NT) blocks.  Per tile, the operand u of every
// distinct column the tile references is formed once into LDS (u_l, indexed
// by the tile-local column number lidx), then CG_J coalesced blocks per
// thread are multiplied against it; block products go to LDS and each row is
// summed by its owner thread.  PROLOGUE: u = s*dv with fixed columns masked.
// Otherwise u_j = p_j' = (z_j - alpha w_j) + beta p_j from the previous
// phase's exchange records.  Fixed rows give q = 0.
// Distinct columns of `tile` held by this thread (-1 past the end).
template <int NT>
__device__ __forceinline__ void load_tile_cols(const CgArgs& a, int tile, int (&ucol)[CG_J]) {
  const int c0 = a.tile_cptr[tile], nu = a.tile_cptr[tile + 1] - c0;
#pragma unroll
  for (int j = 0; j < CG_J; ++j) {
    const int u = threadIdx.x + NT * j;
    ucol[j] = u < nu ? a.tile_cols[c0 + u] : -1;
  }
}

// ucol holds the first tile's distinct columns on entry and again on return:
// each tile prefetches the next one's list (the first tile's after the last)
// so that the operand loads of a tile never wait for its column list.
template <int K, bool PROLOGUE, int NT>
__device__ __forceinline__ void cg_spmv(const CgArgs& a, int R0, int nr, int t0, int t1, const int* off_l,
                                        float s, float alpha, float beta, const float* __restrict__ xw,
                                        float* part_l, float* u_l, int (&ucol)[CG_J], const bool (&fx)[K],
                                        float (&q)[K][3]) {
#pragma unroll
  for (int k = 0; k < K; ++k) q[k][0] = q[k][1] = q[k][2] = 0.f;
  for (int tile = t0; tile < t1; ++tile) {
    const int ta = a.tile_rows[tile] - R0;
    const int tb = (tile + 1 < t1 ? a.tile_rows[tile + 1] : R0 + nr) - R0;
    const int kb = off_l[ta], ke = off_l[tb];
    // This thread's blocks: local column numbers and values (issued before
    // the staging so that their latency overlaps it).
    int li[CG_J];
    float bv[CG_J][9];
#pragma unroll
    for (int j = 0; j < CG_J; ++j) {
      const int k = kb + threadIdx.x + NT * j;
      li[j] = k < ke ? (int)a.lidx[k] : -1;
      if (k < ke) {
        // Three 12-byte loads per block: a wave's load then touches each
        // cache line of its 64 blocks 3 times instead of 9.
        // (f3v is 16 bytes in size: address each row explicitly.)
        const float* b = a.vals + 9 * (size_t)k;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const f3v v = *reinterpret_cast<const f3v*>(b + 3 * r);
          bv[j][3 * r] = v.x;
          bv[j][3 * r + 1] = v.y;
          bv[j][3 * r + 2] = v.z;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < CG_J; ++j) {
      if (ucol[j] >= 0) {
        const int c = ucol[j];
        v3 u;
        if (PROLOGUE) {
          u = a.row_fixed[c] ? v3{0.f, 0.f, 0.f} : scale_cg(ld3(a.dv, c), s);
        } else {
          const f4u* e = reinterpret_cast<const f4u*>(xw + XW * (size_t)c);
          const f4u e0 = ld_shared_u(e), e1 = ld_shared_u(e + 1);
          const float p2 = ld_shared(xw + XW * (size_t)c + 8);
          u = {fmaf(beta, e1.z, fmaf(-alpha, e0.w, e0.x)), fmaf(beta, e1.w, fmaf(-alpha, e1.x, e0.y)),
               fmaf(beta, p2, fmaf(-alpha, e1.y, e0.z))};
        }
        float* o = u_l + 3 * (threadIdx.x + NT * j);
        o[0] = u.x;
        o[1] = u.y;
        o[2] = u.z;
      }
    }
    load_tile_cols<NT>(a, tile + 1 < t1 ? tile + 1 : t0, ucol);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CG_J; ++j) {
      if (li[j] >= 0) {
        const float* uu = u_l + 3 * li[j];
        const float ux = uu[0], uy = uu[1], uz = uu[2];
        float* o = part_l + 3 * (threadIdx.x + NT * j);
        o[0] = fmaf(bv[j][2], uz, fmaf(bv[j][1], uy, bv[j][0] * ux));
        o[1] = fmaf(bv[j][5], uz, fmaf(bv[j][4], uy, bv[j][3] * ux));
        o[2] = fmaf(bv[j][8], uz, fmaf(bv[j][7], uy, bv[j][6] * ux));
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int lr = threadIdx.x + NT * k;
      if (lr >= ta && lr < tb && !fx[k]) {
        float s0 = 0.f, s1 = 0.f, s2 = 0.f;
        const int e = off_l[lr + 1] - kb;
        for (int m = off_l[lr] - kb; m < e; ++m) {
          s0 += part_l[3 * m];
          s1 += part_l[3 * m + 1];
          s2 += part_l[3 * m + 2];
        }
        q[k][0] = s0;
        q[k][1] = s1;
        q[k][2] = s2;
      }
    }
    // The next tile writes u_l only (its products follow its own barrier),
    // so no barrier is needed between these row sums and its staging.
  }
}

// Resident variant of the SpMV: the chunk's blocks stay on chip for the whole
// solve (RES_JR per thread in registers, up to RES_JL more in LDS; at most
// RES_JS more per thread are streamed from HBM every iteration, loaded before
// the operands are formed so that their latency overlaps that step).  Thread
// t holds a contiguous run of the chunk's blocks; it accumulates their
// products row by row and writes one partial per row run (a row split
// between threads has two or more runs), and the owner thread of each row
// adds its runs in order.
constexpr int RES_NT = 256;  // one wave per SIMD: 512 registers per lane
constexpr int RES_K = 4;     // rows per thread (<= 1024 rows per chunk)
constexpr int RES_JR = 28;   // blocks per thread held in registers
constexpr int RES_JL = 10;   // blocks per thread held in LDS
constexpr int RES_JS = 4;    // blocks per thread streamed every iteration
constexpr int RES_UC = 11;   // distinct columns per thread (< 2816 per chunk)
constexpr int RES_SEG = RES_K * RES_NT + RES_NT;  // row runs per chunk
// Chunk-local column number of padding blocks: an operand slot that stays zero.
constexpr int RES_ZERO = RES_UC * RES_NT - 1;

struct ResThread {
  int k0;   // first block of the thread (global block index)
  int cnt;  // blocks of the thread
  int run0;
  int own_lo;  // operand slot of the chunk's first row (its rows take slots own_lo .. own_lo + nr - 1)
  int nr;
  int nhalo;   // distinct columns outside the chunk's rows
  int R0;      // first row of the chunk
  unsigned long long mask;  // bit j: block j ends a row run
};
static_assert(RES_JR <= 32 && RES_JR % 4 == 0 && RES_JR + RES_JL + RES_JS <= 64, "run-end mask layout");

template <int K, bool PROLOGUE>
__device__ __forceinline__ void res_spmv(const CgArgs& a, const ResThread& rt, const float (&A)[RES_JR][9],
                                         const uint32_t (&L)[RES_JR / 2], const int (&ucol)[RES_UC],
                                         const float* mat_l, const uint16_t* lidx_l, f4v* u_l, f4v* seg_l,
                                         float s, float alpha, float beta, const float* __restrict__ xw,
                                         const int (&rf)[K], const int (&rn)[K], const bool (&fx)[K],
                                         const bool (&own)[K], const float (&p)[K][3], float (&q)[K][3]) {
  constexpr int NT = RES_NT;
  // Operands: owners write their own rows' slots (p, or the masked scaled
  // warm start in the prologue; outside the prologue p is exactly what the
  // record rebuild gives); the halo columns are gathered four per thread at a
  // time (bounds the registers the records occupy), loads unconditional.
  // ucol[i] < 0 marks a slot past the halo: it stores a zero operand into
  // the zero slot RES_ZERO.
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (own[k]) {
      f4v u = f4v{p[k][0], p[k][1], p[k][2], 0.f};
      if (PROLOGUE) {
        const v3 d = fx[k] ? v3{0.f, 0.f, 0.f} : scale_cg(ld3(a.dv, rt.R0 + (int)threadIdx.x + NT * k), s);
        u = f4v{d.x, d.y, d.z, 0.f};
      }
      u_l[rt.own_lo + threadIdx.x + NT * k] = u;
    }
  }
  // Opaque copies (as for the block metadata below): the per-column record
  // offsets and operand slots are recomputed here instead of being hoisted
  // out of the solve loop into registers.
  int hc[RES_UC];
  int own_lo = rt.own_lo, nrow = rt.nr;
  asm volatile("" : "+v"(own_lo), "+v"(nrow));
#pragma unroll
  for (int i = 0; i < RES_UC; ++i) {
    hc[i] = ucol[i];
    asm volatile("" : "+v"(hc[i]));
  }
#pragma unroll
  for (int i0 = 0; i0 < RES_UC; i0 += 4) {
    if (i0 * NT >= rt.nhalo) break;
#pragma unroll
    for (int i = i0; i < i0 + 4 && i < RES_UC; ++i) {
      const int h = (int)threadIdx.x + NT * i;
      const bool live = hc[i] >= 0;
      const uint32_t c = live ? (uint32_t)hc[i] : 0u;
      v3 u;
      if (PROLOGUE) {
        u = a.row_fixed[c] ? v3{0.f, 0.f, 0.f} : scale_cg(ld3(a.dv, (int)c), s);
      } else {
        const float* rec = xw + (size_t)(c * (uint32_t)XW);
        const f4u* e = reinterpret_cast<const f4u*>(rec);
        const f4u e0 = ld_shared_u(e), e1 = ld_shared_u(e + 1);
        const float p2 = ld_shared(rec + 8);
        u = {fmaf(beta, e1.z, fmaf(-alpha, e0.w, e0.x)), fmaf(beta, e1.w, fmaf(-alpha, e1.x, e0.y)),
             fmaf(beta, p2, fmaf(-alpha, e1.y, e0.z))};
      }
      const int slot = !live ? RES_ZERO : h < own_lo ? h : h + nrow;
      u_l[slot] = live ? f4v{u.x, u.y, u.z, 0.f} : f4v{0.f, 0.f, 0.f, 0.f};
    }
    asm volatile("" ::: "memory");
  }
  // Streamed tier: loaded once the operands are formed, so that these
  // registers are free during the gathers; the loads complete while the
  // register and LDS tiers compute.
  float B[RES_JS][9];
  uint32_t bl[RES_JS];
#pragma unroll
  for (int i = 0; i < RES_JS; ++i) {
    const int j = RES_JR + RES_JL + i;
    bl[i] = RES_ZERO;
#pragma unroll
    for (int e = 0; e < 9; ++e) B[i][e] = 0.f;
    if (j < rt.cnt) {
      const float* b = a.vals + 9 * (size_t)(rt.k0 + j);
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const f3v v = *reinterpret_cast<const f3v*>(b + 3 * r);
        B[i][3 * r] = v.x;
        B[i][3 * r + 1] = v.y;
        B[i][3 * r + 2] = v.z;
      }
      bl[i] = a.lidx[rt.k0 + j];
    }
  }
  __syncthreads();
  // Block products accumulated per row run.  Every block stores the running
  // partial of its run (no divergent branches); the slot advances after a
  // run's last block, so each slot ends up holding its run's full partial.
  // Padding blocks (past the thread's count) store to the unused slot RES_SEG.
  f4v acc = {0.f, 0.f, 0.f, 0.f};
  int run = rt.run0;
  uint32_t mlo = (uint32_t)rt.mask, mhi = (uint32_t)(rt.mask >> 32);
  int cnt = rt.cnt;
  // Opaque copies: keeps the compiler from hoisting per-block addresses and
  // run-end flags out of the solve loop (they would occupy registers for the
  // whole solve); recomputing them is a few ALU operations per block.
  uint32_t Lw[RES_JR / 2];
#pragma unroll
  for (int i = 0; i < RES_JR / 2; ++i) {
    Lw[i] = L[i];
    asm volatile("" : "+v"(Lw[i]));
  }
  asm volatile("" : "+v"(mlo), "+v"(mhi), "+v"(run), "+v"(cnt));
  auto block = [&](int j, const f4v& u, float a0, float a1, float a2, float a3, float a4, float a5, float a6,
                   float a7, float a8) {
    acc.x = fmaf(a2, u.z, fmaf(a1, u.y, fmaf(a0, u.x, acc.x)));
    acc.y = fmaf(a5, u.z, fmaf(a4, u.y, fmaf(a3, u.x, acc.y)));
    acc.z = fmaf(a8, u.z, fmaf(a7, u.y, fmaf(a6, u.x, acc.z)));
    seg_l[j < cnt ? run : RES_SEG] = acc;
    const bool e = ((((uint64_t)mhi << 32) | mlo) >> j) & 1u;
    run += e;
    acc = e ? f4v{0.f, 0.f, 0.f, 0.f} : acc;
  };
  // Register tier; operand loads are issued four blocks ahead of their products.
  constexpr int G = 4;
#pragma unroll
  for (int g = 0; g < RES_JR; g += G) {
    f4v u[G];
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int j = g + i;
      u[i] = u_l[(Lw[j >> 1] >> (16 * (j & 1))) & 0xffffu];
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int j = g + i;
      block(j, u[i], A[j][0], A[j][1], A[j][2], A[j][3], A[j][4], A[j][5], A[j][6], A[j][7], A[j][8]);
    }
  }
  // LDS tier.
  constexpr int S = RES_JL * NT;  // element e of an LDS block at m[e * S]
  const int jle = min(cnt, RES_JR + RES_JL);
  for (int j = RES_JR; j < jle; ++j) {
    const int jl = j - RES_JR;
    const f4v u = u_l[lidx_l[jl * NT + threadIdx.x]];
    const float* m = mat_l + jl * NT + threadIdx.x;
    block(j, u, m[0], m[S], m[2 * S], m[3 * S], m[4 * S], m[5 * S], m[6 * S], m[7 * S], m[8 * S]);
  }
  // Streamed tier.
#pragma unroll
  for (int i = 0; i < RES_JS; ++i) {
    const f4v u = u_l[bl[i]];
    block(RES_JR + RES_JL + i, u, B[i][0], B[i][1], B[i][2], B[i][3], B[i][4], B[i][5], B[i][6], B[i][7], B[i][8]);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    f4v t = {0.f, 0.f, 0.f, 0.f};
    if (!fx[k])
      for (int m = rf[k]; m < rf[k] + rn[k]; ++m) t += seg_l[m];
    q[k][0] = t.x;
    q[k][1] = t.y;
    q[k][2] = t.z;
  }
  // The next call writes u_l first and seg_l only after its own barrier.
}

// Exchange records of the resident kernel, written in two steps so that the
// stores of every row fill whole cache lines.  For each k, a wave's rows are
// 64 consecutive rows of the chunk (local rows 256 k + 64 w .. + 63), whose
// records form one contiguous region.  xw_stage: each lane puts its 36-byte
// record into the wave's LDS area for k (576 floats at stride 9: no bank
// conflicts).  xw_flush: the wave writes the region as consecutive 16-byte
// pieces, instead of three partial writes of every 64-byte segment.  LDS
// operations of one wave complete in issue order and the compiler barriers
// keep that order, so no workgroup barrier is needed; every k is staged
// before the first is flushed, so one LDS round trip serves all four.
// Wave-uniform values are kept in scalar registers; lanes past the chunk's
// rows (own == false) stage nothing.
__device__ __forceinline__ float* xw_area(float* stage, int k) {
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  return stage + 64 * XW * (4 * k + wv);
}

__device__ __forceinline__ void xw_stage(float* stage, int k, bool own, const float (&z)[3], const float (&w)[3],
                                         const float (&p)[3]) {
  if (!own) return;
  float* e = xw_area(stage, k) + XW * ((int)threadIdx.x & 63);
  e[0] = z[0];
  e[1] = z[1];
  e[2] = z[2];
  e[3] = w[0];
  e[4] = w[1];
  e[5] = w[2];
  e[6] = p[0];
  e[7] = p[1];
  e[8] = p[2];
}

__device__ __forceinline__ void xw_flush(float* xw, int R0, int k, int nr, float* stage) {
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  const int first = RES_NT * k + 64 * wv;
  const int valid = min(64, nr - first);
  if (valid <= 0) return;
  const float* sw = xw_area(stage, k);
  const int lane = (int)threadIdx.x & 63;
  const int nf = XW * valid;
  float* dst = xw + XW * (size_t)(R0 + first);
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int f = 4 * lane + 256 * u;
    if (f + 4 <= nf) st_wt(dst + f, *reinterpret_cast<const f4v*>(sw + f));
  }
  if ((nf & 3) && lane == ((nf >> 2) & 63))  // a partial wave's last 1-3 floats
    for (int t = nf & ~3; t < nf; ++t) st_wt(dst + t, sw[t]);
}
constexpr int XW_STAGE = 64 * XW;  // LDS floats per wave and k for xw_stage

// Exchange record of one row written in place (three write-through stores);
// the streamed kernels use this (staging costs them registers that decide
// how many workgroups fit on a CU).
__device__ __forceinline__ void st_xw(float* xw, int i, const float (&z)[3], const float (&w)[3],
                                      const float (&p)[3]) {
  float* e = xw + XW * (size_t)i;
  st_wt(e, f4v{z[0], z[1], z[2], w[0]});
  st_wt(e + 4, f4v{w[1], w[2], p[0], p[1]});
  st_wt(e + 8, p[2]);
}

// Warp's absolute floor eps32^(9/4) on the residual norm (tol given, atol=0).
__device__ __forceinline__ double residual_floor() { return 2.6442191937624087e-16; }

// Block-Jacobi PCG with one instance-wide barrier per iteration.
//
// The iterates are those of the standard PCG of warp.optim.linear.cg
// (alpha_k = rz_k / p_k.Ap_k, beta_k = rz_{k+1} / rz_k, residual-masked
// updates), reorganised so that the reduction and the exchange of the SpMV
// operand share one synchronisation point.  Phase k (after the barrier):
//   alpha_{k-1} = r_{k-1}.p_{k-1} / p_{k-1}.q_{k-1} (the exact line search;
//   r.p = r.z in exact arithmetic);
//   rz_k = rz_{k-1} - 2 alpha q.z + alpha^2 q.w      (w = M^-1 q; M^-1 is
//   rr_k = rr_{k-1} - 2 alpha q.r + alpha^2 q.q       symmetric) give beta_{k-1};
//   x += alpha p, r -= alpha q, z -= alpha w, p = z + beta p   (own rows);
//   q = A p, gathering neighbour p from the exchange records (z, w, p) of
//   phase k-1; w = M^-1 q and z = M^-1 r afresh; publish (z, w, p) and the partial sums
//   p.q, q.z, q.w, q.r, q.q, r.z, r.r, r.p; barrier.
// The direct r.z and r.r of phase k replace the recurrence values at phase
// k+1, so recurrence error does not accumulate.  Exchange records and
// partial sums are double buffered: phase k writes buffer (k+1)&1, which no
// workgroup reads after the barrier that ended phase k-1.
//
// RES selects the resident SpMV (res_spmv: blocks kept on chip, RES_NT
// threads, RES_K rows per thread) instead of the streamed tiles (cg_spmv).
template <int K, int NT, bool RES>
constexpr size_t cg_smem_bytes() {
  return RES ? sizeof(f4v) * (RES_UC * NT + RES_SEG + 1) + sizeof(float) * 9 * RES_JL * NT +
                   sizeof(uint16_t) * RES_JL * NT
             : sizeof(float) * (6 * CG_J * NT + K * NT + 1);
}

// The solve of chunk c by one workgroup (smem: cg_smem_bytes, red: the
// workgroup reduction slots).
template <int K, int NT, bool RES>
__device__ __forceinline__ void cg_chunk(const CgArgs& a, const int c, unsigned char* smem, double* red) {
  // streamed tiles
  float* part_l = reinterpret_cast<float*>(smem);
  float* u_l = part_l + 3 * CG_J * NT;
  int* off_l = reinterpret_cast<int*>(u_l + 3 * CG_J * NT);
  // resident blocks
  f4v* ures_l = reinterpret_cast<f4v*>(smem);
  f4v* seg_l = ures_l + RES_UC * NT;
  float* mat_l = reinterpret_cast<float*>(seg_l + RES_SEG + 1);
  uint16_t* lidx_l = reinterpret_cast<uint16_t*>(mat_l + 9 * RES_JL * NT);
  // Record staging of the resident kernel (xw_stage): the operand table,
  // which nothing reads between the end of an SpMV and the next phase's
  // barrier.
  float* stage = reinterpret_cast<float*>(ures_l);
  static_assert(!RES || XW_STAGE * (NT / 64) * K <= 4 * RES_UC * NT, "record staging fits");
  const int inst = a.chunk_inst[c];
  const int R0 = a.chunk_rows[c], nr = a.chunk_rows[c + 1] - R0;
  const int cb = a.inst_chunks[inst], ce = a.inst_chunks[inst + 1], nci = ce - cb;
  const int t0 = a.chunk_tiles[c], t1 = a.chunk_tiles[c + 1];
  const size_t nch = (size_t)a.nchunks_total;
  float m = 0.f;
  for (int k = 0; k < RHS_MAX_SLOTS; ++k) m = fmaxf(m, __uint_as_float(a.rhs_max[RHS_MAX_SLOTS * inst + k]));
  const float s = (m > 0.f && m < 1e-12f) ? fminf(1.f / m, 1e30f) : 1.f;  // _linear_rhs_scale kernels.py:1785-1793
  const bool zero_start = (m == 0.f) || a.x_out != nullptr;

  if (!RES)
    for (int r = threadIdx.x; r <= nr; r += NT) off_l[r] = a.offsets[R0 + r];

  bool ok = true;
  __syncthreads();

  float x[K][3], r[K][3], z[K][3], p[K][3], qv[K][3], w[K][3];
  // M^-1 blocks of the own rows: resident for the whole solve when the
  // register budget allows (K <= 2), else reloaded where used.
  constexpr bool MRES = K <= 2;
  float Mk[MRES ? K : 1][9];
  bool own[K], fx[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int lr = threadIdx.x + NT * k;
    own[k] = lr < nr;
    fx[k] = own[k] ? (a.row_fixed[R0 + lr] != 0) : true;
    if (MRES) {
#pragma unroll
      for (int e = 0; e < 9; ++e) Mk[MRES ? k : 0][e] = own[k] ? a.minv[9 * (size_t)(R0 + lr) + e] : 0.f;
    }
  }
  auto minv_mul = [&](int k, int i, v3 v) -> v3 {
    if (MRES) return mv_cg(Mk[MRES ? k : 0], v);
    float m[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) m[e] = a.minv[9 * (size_t)i + e];
    return mv_cg(m, v);
  };
  int ucol[CG_J];
  // Resident blocks, their chunk-local column numbers (two per word), the
  // distinct columns this thread forms operands for, and the row runs of the
  // own rows.
  constexpr int JR = RES ? RES_JR : 1;
  float A[RES_JR][9];
  uint32_t L[RES_JR / 2];
  int rcol[RES_UC];
  int rf[K], rn[K];
  ResThread rt{0, 0, 0, 0, 0, 0, 0, 0ull};
  if constexpr (RES) {
    const int4 th = a.res_thr[(size_t)c * NT + threadIdx.x];
    rt.k0 = a.offsets[R0] + th.x;
    rt.cnt = th.y;
    rt.run0 = th.z;
    rt.own_lo = th.w;
    rt.nr = nr;
    rt.R0 = R0;
    rt.mask = a.res_mask[(size_t)c * NT + threadIdx.x];
    const size_t k0 = (size_t)rt.k0;
#pragma unroll
    for (int j = 0; j < JR; ++j) {
      uint32_t lid = RES_ZERO;  // padding: zero block against the zero operand
      if (j < rt.cnt) {
        const float* b = a.vals + 9 * (k0 + j);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const f3v v = *reinterpret_cast<const f3v*>(b + 3 * r);
          A[j][3 * r] = v.x;
          A[j][3 * r + 1] = v.y;
          A[j][3 * r + 2] = v.z;
        }
        lid = a.lidx[k0 + j];
      } else {
#pragma unroll
        for (int e = 0; e < 9; ++e) A[j][e] = 0.f;
      }
      if (j & 1) L[j >> 1] |= lid << 16;
      else L[j >> 1] = lid;
    }
    for (int j = RES_JR; j < min(rt.cnt, RES_JR + RES_JL); ++j) {
      const int jl = j - RES_JR;
      const float* b = a.vals + 9 * (k0 + j);
#pragma unroll
      for (int e = 0; e < 9; ++e) mat_l[(e * RES_JL + jl) * NT + threadIdx.x] = b[e];
      lidx_l[jl * NT + threadIdx.x] = a.lidx[k0 + j];
    }
    // Halo columns only (the chunk's own rows are formed by their owners):
    // halo number h maps to operand slot h below the own rows, h + nr above.
    const int c0 = a.tile_cptr[c], nu = a.tile_cptr[c + 1] - c0;
    rt.nhalo = nu - nr;
#pragma unroll
    for (int i = 0; i < RES_UC; ++i) {
      const int h = threadIdx.x + NT * i;
      rcol[i] = h < rt.nhalo ? a.tile_cols[c0 + (h < rt.own_lo ? h : h + nr)] : -1;
    }
    if (threadIdx.x == 0) ures_l[RES_ZERO] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int lr = threadIdx.x + NT * k;
      const int2 rr2 = own[k] ? a.res_row[R0 + lr] : make_int2(0, 0);
      rf[k] = rr2.x;
      rn[k] = rr2.y;
    }
    __syncthreads();
  } else {
    load_tile_cols<NT>(a, t0, ucol);
  }
  auto spmv = [&](auto prologue, float alpha, float beta, const float* xw) {
    constexpr bool PRO = decltype(prologue)::value;
    if constexpr (RES)
      res_spmv<K, PRO>(a, rt, A, L, rcol, mat_l, lidx_l, ures_l, seg_l, s, alpha, beta, xw, rf, rn, fx, own, p, qv);
    else
      cg_spmv<K, PRO, NT>(a, R0, nr, t0, t1, off_l, s, alpha, beta, xw, part_l, u_l, ucol, fx, qv);
  };

  // Prologue: r0 = b - A x0, z0 = M^-1 r0; records (z0, 0, 0) so that the
  // first phase gathers p0 = z0.
  if (!zero_start) {
    spmv(std::true_type{}, 0.f, 0.f, a.xw);
  } else {
#pragma unroll
    for (int k = 0; k < K; ++k) qv[k][0] = qv[k][1] = qv[k][2] = 0.f;
  }
  float pb[3] = {0.f, 0.f, 0.f};  // b.b, r.r, r.z
#pragma unroll
  for (int k = 0; k < K; ++k) {
#pragma unroll
    for (int d = 0; d < 3; ++d) x[k][d] = r[k][d] = z[k][d] = p[k][d] = w[k][d] = 0.f;
    if (own[k]) {
      const int i = R0 + threadIdx.x + NT * k;
      const v3 b = scale_cg(ld3(a.rhs, i), s);
      const v3 x0 = (zero_start || fx[k]) ? v3{0.f, 0.f, 0.f} : scale_cg(ld3(a.dv, i), s);
      const v3 rr0 = sub_cg(b, v3{qv[k][0], qv[k][1], qv[k][2]});
      const v3 zz = minv_mul(k, i, rr0);
      x[k][0] = x0.x; x[k][1] = x0.y; x[k][2] = x0.z;
      r[k][0] = rr0.x; r[k][1] = rr0.y; r[k][2] = rr0.z;
      z[k][0] = zz.x; z[k][1] = zz.y; z[k][2] = zz.z;
      pb[0] += dot_cg(b, b);
      pb[1] += dot_cg(rr0, rr0);
      pb[2] += dot_cg(rr0, zz);
    }
    if constexpr (RES)
      xw_stage(stage, k, own[k], z[k], w[k], p[k]);
    else if (own[k])
      st_xw(a.xw, R0 + threadIdx.x + NT * k, z[k], w[k], p[k]);
#pragma unroll
    for (int d = 0; d < 3; ++d) qv[k][d] = 0.f;
  }
  if constexpr (RES) {
    asm volatile("" ::: "memory");
#pragma unroll
    for (int k = 0; k < K; ++k) xw_flush(a.xw, R0, k, nr, stage);
    asm volatile("" ::: "memory");
  }
  {
    double wsum[3];
    wg_sum<NT, 3>(pb, wsum, red);
    static_assert(PT_QQ == 4 && PT_RZ == 5 && PT_RR == 6 && PT_BB == 7, "partial-sum slots");
    if (threadIdx.x == 0) {
      st_wt(a.part + NPART * (size_t)c + 4, d2v{0.0, wsum[2]});
      st_wt(a.part + NPART * (size_t)c + 6, d2v{wsum[1], wsum[0]});
    }
  }
  unsigned epoch = 1;
  publish(a.flags + c, epoch);
  double tot[3];
  {
    const int sl[3] = {PT_BB, PT_RR, PT_RZ};
    ok = wait_sum<NT, 3>(a.flags, a.part, cb, ce, epoch, sl, tot, red, a.err, a.spin_limit);
  }
  bool failed = !ok;
  double rr = tot[1], rz = tot[2];
  double atol = fmax(0.0, (double)a.tolerance * sqrt(tot[0]));
  atol = fmax(atol, residual_floor());
  const double atol_sq = atol * atol;
  int iters = 0;
  for (int ph = 0; ph <= a.maxiter && ok; ++ph) {
    const int cur = ph & 1;
    float alpha = 0.f, beta = 0.f;
    if (ph > 0) {
      double g[8];
      {
        const int sl[8] = {PT_PAP, PT_QZ, PT_QW, PT_QR, PT_QQ, PT_RZ, PT_RR, PT_RP};
        if (!wait_sum<NT, 8>(a.flags, a.part + NPART * nch * cur, cb, ce, epoch, sl, g, red, a.err, a.spin_limit)) {
          failed = true;
          break;
        }
      }
      rz = g[5];
      rr = g[6];
      const bool live = rr > atol_sq;
      if (!live && a.early_exit) break;
      if (live) ++iters;
      // Exact line search along p: alpha = r.p / p.Ap (= r.z / p.Ap in exact
      // arithmetic).  With beta from the r.z recurrence, r.z / p.Ap lets a
      // solve that reached the fp32 noise level drift away again.
      alpha = (live && g[0] != 0.0) ? (float)(g[7] / g[0]) : 0.f;
      const double al = (double)alpha;
      const double rr_next = rr - 2.0 * al * g[3] + al * al * g[4];
      const double rz_next = rz - 2.0 * al * g[1] + al * al * g[2];
      beta = (rr_next > atol_sq && rz != 0.0) ? (float)(rz_next / rz) : 0.f;
      rr = rr_next;
    }
    // x += alpha p, r -= alpha q, z -= alpha w, p = z + beta p (own rows,
    // bit-identical to what the neighbours rebuild in cg_spmv).
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        x[k][d] = fmaf(alpha, p[k][d], x[k][d]);
        r[k][d] = fmaf(-alpha, qv[k][d], r[k][d]);
        z[k][d] = fmaf(-alpha, w[k][d], z[k][d]);
        p[k][d] = fmaf(beta, p[k][d], z[k][d]);
      }
    }
    if (ph == a.maxiter) break;
    spmv(std::false_type{}, alpha, beta, a.xw + XW * (size_t)a.P * cur);
    float pp[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float* xw_next = a.xw + XW * (size_t)a.P * (cur ^ 1);
    // Inverse diagonal blocks not held in registers: all rows' loads are
    // issued together (rows past the chunk load its first row, unused).
    float Mi[MRES ? 1 : K][9];
    if constexpr (!MRES) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const size_t i = own[k] ? (size_t)(R0 + threadIdx.x + NT * k) : (size_t)R0;
#pragma unroll
        for (int e = 0; e < 9; ++e) Mi[MRES ? 0 : k][e] = a.minv[9 * i + e];
      }
    }
    // w = M^-1 q for every row before any record store: a load waited for
    // after a write-through store would also wait for that store (vector
    // memory operations complete in issue order).
    // z = M^-1 r is applied afresh too (as warp's cg preconditions the
    // residual in every iteration), so the recurrence z -= alpha w of the
    // next phase carries one step of rounding instead of accumulating drift
    // that stalls a tight solve; the record publishes this z.
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (own[k]) {
        const v3 qq = {qv[k][0], qv[k][1], qv[k][2]}, rv = {r[k][0], r[k][1], r[k][2]};
        const int i = R0 + threadIdx.x + NT * k;
        const v3 ww = MRES ? minv_mul(k, i, qq) : mv_cg(Mi[MRES ? 0 : k], qq);
        const v3 zz = MRES ? minv_mul(k, i, rv) : mv_cg(Mi[MRES ? 0 : k], rv);
        w[k][0] = ww.x; w[k][1] = ww.y; w[k][2] = ww.z;
        z[k][0] = zz.x; z[k][1] = zz.y; z[k][2] = zz.z;
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if constexpr (RES)
        xw_stage(stage, k, own[k], z[k], w[k], p[k]);
      else if (own[k])
        st_xw(xw_next, R0 + threadIdx.x + NT * k, z[k], w[k], p[k]);
      if (own[k]) {
        const v3 qq = {qv[k][0], qv[k][1], qv[k][2]}, ww = {w[k][0], w[k][1], w[k][2]};
        const v3 pv = {p[k][0], p[k][1], p[k][2]}, zv = {z[k][0], z[k][1], z[k][2]},
                 rv = {r[k][0], r[k][1], r[k][2]};
        pp[0] += dot_cg(pv, qq);
        pp[1] += dot_cg(qq, zv);
        pp[2] += dot_cg(qq, ww);
        pp[3] += dot_cg(qq, rv);
        pp[4] += dot_cg(qq, qq);
        pp[5] += dot_cg(rv, zv);
        pp[6] += dot_cg(rv, rv);
        pp[7] += dot_cg(rv, pv);
      }
    }
    if constexpr (RES) {
      asm volatile("" ::: "memory");
#pragma unroll
      for (int k = 0; k < K; ++k) xw_flush(xw_next, R0, k, nr, stage);
      asm volatile("" ::: "memory");
    }
    {
      double ws[8];
      wg_sum<NT, 8>(pp, ws, red);
      static_assert(PT_PAP == 0 && PT_QZ == 1 && PT_QW == 2 && PT_QR == 3 && PT_RR == 6, "partial-sum slots");
      if (threadIdx.x == 0) {
        double* dst = a.part + NPART * (nch * (cur ^ 1) + c);
        st_wt(dst, d2v{ws[0], ws[1]});
        st_wt(dst + 2, d2v{ws[2], ws[3]});
        st_wt(dst + 4, d2v{ws[4], ws[5]});
        st_wt(dst + 6, d2v{ws[6], ws[7]});
      }
    }
    publish(a.flags + c, ++epoch);
  }
  // A workgroup whose waits timed out (or that sees another's timeout) has no
  // valid iterate: it writes NaN instead of a plausible partial solve, so a
  // caller that does not read the status word still cannot take the result
  // for a state (the increment stays NaN until the next reset).
  if (!failed) failed = __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
  // restore_linear_increment kernels.py:1808-1811, update_state :115-128,
  // FixedParticles.restore(state_out) solver.py:1350-1351.
  const float nan = __builtin_nanf("");
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (own[k] && a.x_out) {
      const v3 xo = failed ? v3{nan, nan, nan} : scale_cg(v3{x[k][0], x[k][1], x[k][2]}, 1.f / s);
      st3(a.x_out, R0 + threadIdx.x + NT * k, xo);
    } else if (own[k]) {
      const int i = R0 + threadIdx.x + NT * k;
      const v3 dvi = failed ? v3{nan, nan, nan} : scale_cg(v3{x[k][0], x[k][1], x[k][2]}, 1.f / s);
      st3(a.dv, i, dvi);
      if (a.q_out) {
        v3 vel = add_cg(ld3(a.qd_in, i), dvi);
        v3 qo = fma_cg(vel, a.h, ld3(a.q_in, i));
        if (fx[k] && !failed) {
          qo = ld3(a.anchor, i);
          vel = v3{0.f, 0.f, 0.f};
        }
        st3(a.q_out, i, qo);
        st3(a.qd_out, i, vel);
      }
    }
  }
  if (c == cb && threadIdx.x == 0) {
    a.stats[4 * inst] = iters;
    a.stats[4 * inst + 1] = __float_as_int((float)rr);
    a.stats[4 * inst + 2] = __float_as_int((float)atol_sq);
  }
}

// One persistent launch.  Without QUEUE every chunk has a workgroup of its
// own: workgroup b solves chunk xcd_swizzle(b, n) (an XCC holds consecutive
// chunks).  With QUEUE the launch holds as many workgroups as fit at once and
// they take chunks from a queue (the word after the chunk flags), in the
// order chunk_order gives: whole instances, largest first.  A workgroup
// moves to the next chunk when its instance has converged.  The order keeps
// the queue from deadlocking: every instance before the one at the queue's
// head has all its chunks running, so it converges and frees workgroups for
// the rest of the head's chunks (an instance never needs more workgroups
// than fit, which the planner checks).  The queue loop keeps a few more
// values live, so it is a kernel of its own rather than a branch.
template <int K, int NT, bool RES, bool QUEUE>
__global__ __launch_bounds__(NT) void k_cg(CgArgs a) {
  static_assert(!RES || (NT == RES_NT && K == RES_K), "resident CG shape");
  __shared__ __align__(16) unsigned char smem[cg_smem_bytes<K, NT, RES>()];
  __shared__ double red[(NT / 64) * NPART];
  if constexpr (!QUEUE) {
    cg_chunk<K, NT, RES>(a, a.chunk_base + xcd_swizzle(blockIdx.x, gridDim.x), smem, red);
  } else {
    __shared__ int next_l;
    for (int item = blockIdx.x; item < a.nchunks_total;) {
      cg_chunk<K, NT, RES>(a, a.chunk_order[item], smem, red);
      __syncthreads();  // this chunk's shared memory and the previous next_l are no longer read
      if (threadIdx.x == 0)
        next_l = (int)gridDim.x + (int)__hip_atomic_fetch_add(a.flags + a.nchunks_total, 1u, __ATOMIC_RELAXED,
                                                               __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      item = next_l;
    }
  }
}

#pragma clang fp contract(fast)

// ---------------------------------------------------------------------------
// Prepared-step actions (PreparedFEMStep, prepared_step.py:64-302).
// ---------------------------------------------------------------------------
// y = Op x with Op = A, or PAP + I - P with fixed rows (fixed_particles.py:29-53);
// G lanes per row, fixed summation order.
template <int G>
__global__ __launch_bounds__(BLOCK) void k_apply_op(RowArgs a, const float* __restrict__ vals,
                                                    const float* __restrict__ x, float* __restrict__ y) {
  constexpr int NG = BLOCK / G;
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  const int i = blockIdx.x * NG + grp;
  const bool active = i < a.P;
  const bool fixed_row = active && a.row_fixed[i] != 0;
  v3 acc = {0.f, 0.f, 0.f};
  if (active && !fixed_row) {
    const int k1 = a.offsets[i + 1];
    for (int k = a.offsets[i] + lane; k < k1; k += G) {
      const int c = a.cols[k];
      if (a.row_fixed[c]) continue;
      const float* b = vals + 9 * (size_t)k;
      const v3 u = ld3(x, c);
      acc = acc + v3{b[0] * u.x + b[1] * u.y + b[2] * u.z, b[3] * u.x + b[4] * u.y + b[5] * u.z,
                     b[6] * u.x + b[7] * u.y + b[8] * u.z};
    }
  }
  acc = {group_sum<G>(acc.x), group_sum<G>(acc.y), group_sum<G>(acc.z)};
  if (active && lane == 0) st3(y, i, fixed_row ? ld3(x, i) : acc);
}

// y = M^-1 x (block_jacobi_mv_kernel kernels.py:1852-1868 with alpha 1, beta 0).
__global__ void k_apply_minv(const float* __restrict__ minv, const float* __restrict__ x, float* __restrict__ y,
                             int P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) st3(y, i, mv(minv + 9 * (size_t)i, ld3(x, i)));
}

// z = alpha M^-1 x + beta y (block_jacobi_mv_kernel kernels.py:1852-1868);
// y is not read when beta is 0; z may alias x or y (one row per thread).
__global__ void k_block_jacobi_mv(const float* __restrict__ minv, const float* x, const float* y, float* z,
                                  float alpha, float beta, int P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  v3 out = mv(minv + 9 * (size_t)i, ld3(x, i)) * alpha;
  if (beta != 0.f) out = out + ld3(y, i) * beta;
  st3(z, i, out);
}

// Unconstrained diagonal blocks, identity on fixed rows (prepared_step.py:46-61).
__global__ void k_diag_blocks(RowArgs a, const float* __restrict__ vals, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.P) return;
  float d[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (a.row_fixed[i]) {
    d[0] = d[4] = d[8] = 1.f;
  } else if (a.diag_rel[i] >= 0) {
    const float* b = vals + 9 * (size_t)(a.offsets[i] + a.diag_rel[i]);
#pragma unroll
    for (int c = 0; c < 9; ++c) d[c] = b[c];
  }
#pragma unroll
  for (int c = 0; c < 9; ++c) out[9 * (size_t)i + c] = d[c];
}

// out = base + h ext, fixed rows zero (the load term of the prepared right-hand side).
__global__ void k_add_load(const float* __restrict__ base, const float* __restrict__ ext, float h,
                           const unsigned char* __restrict__ row_fixed, float* __restrict__ out, int P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  v3 f = ld3(base, i);
  if (ext) f = f + ld3(ext, i) * h;
  if (row_fixed[i]) f = {0.f, 0.f, 0.f};
  st3(out, i, f);
}

// update_state kernels.py:115-128 + FixedParticles.restore(state_out).
__global__ void k_reconstruct(const float* __restrict__ dv, const float* __restrict__ q_in,
                              const float* __restrict__ qd_in, const unsigned char* __restrict__ row_fixed,
                              const float* __restrict__ anchor, float h, float* __restrict__ q_out,
                              float* __restrict__ qd_out, int P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  v3 vel = ld3(qd_in, i) + ld3(dv, i);
  v3 qo = ld3(q_in, i) + vel * h;
  if (row_fixed[i]) {
    qo = ld3(anchor, i);
    vel = {0.f, 0.f, 0.f};
  }
  st3(q_out, i, qo);
  st3(qd_out, i, vel);
}

__global__ void k_restore(const int* __restrict__ fixed, int nfixed, const float* __restrict__ anchor, float* q,
                          float* qd) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nfixed) {
    int i = fixed[t];
    st3(q, i, ld3(anchor, i));
    if (qd) st3(qd, i, v3{0.f, 0.f, 0.f});
  }
}

// Compact values: block k of the compact pattern is reserved block map[k];
// the live count is read from the device (the compact offsets' last entry), so
// one launch shape serves every topology.
__global__ __launch_bounds__(BLOCK) void k_gather_blocks(const int* __restrict__ map, const int* __restrict__ count,
                                                         const float* __restrict__ src, float* __restrict__ dst) {
  const size_t n = 9 * (size_t)*count;
  for (size_t t = (size_t)blockIdx.x * BLOCK + threadIdx.x; t < n; t += (size_t)gridDim.x * BLOCK) {
    const size_t k = t / 9;
    dst[t] = src[9 * (size_t)map[k] + (t - 9 * k)];
  }
}

__global__ void k_copy_floats(const float* __restrict__ src, float* __restrict__ dst, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = src[t];
}

// ---------------------------------------------------------------------------
// Component scaling (component_scaling.py:14-49): every equation of a
// mechanically disconnected component is divided by the component's total
// mass.  k_component_mass sums the masses per label in fp64, k_component_scale
// forms the per-row factor (1 for a row whose label is outside [0, capacity)
// or whose component has no positive mass), k_scale_system writes the scaled
// copies of the matrix rows, the inverse diagonal blocks and the right-hand
// side that the solve then uses.
// ---------------------------------------------------------------------------
__global__ void k_component_mass(const int* __restrict__ labels, const float* __restrict__ mass, int P, int cap,
                                 double* __restrict__ total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int c = labels[i];
  if (c >= 0 && c < cap) atomicAdd(total + c, (double)mass[i]);
}

__global__ void k_component_scale(const int* __restrict__ labels, const double* __restrict__ total, int P, int cap,
                                  float* __restrict__ scale) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int c = labels[i];
  float f = 1.f;
  if (c >= 0 && c < cap && total[c] > 0.0) f = (float)(1.0 / total[c]);
  scale[i] = f;
}

// One wave per row: lanes stride over the row's 9 * nnz values.
__global__ __launch_bounds__(BLOCK) void k_scale_system(const int* __restrict__ offsets, const float* __restrict__ vals,
                                                        const float* __restrict__ minv, const float* __restrict__ rhs,
                                                        const float* __restrict__ scale, int P, float* __restrict__ svals,
                                                        float* __restrict__ sminv, float* __restrict__ srhs) {
  const int row = blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= P) return;
  const float f = scale[row];
  const size_t b0 = 9 * (size_t)offsets[row], b1 = 9 * (size_t)offsets[row + 1];
  for (size_t k = b0 + lane; k < b1; k += 64) svals[k] = f * vals[k];
  if (lane < 9) sminv[9 * (size_t)row + lane] = minv[9 * (size_t)row + lane] / f;
  if (lane < 3) srhs[3 * (size_t)row + lane] = f * rhs[3 * (size_t)row + lane];
}

// ---------------------------------------------------------------------------
// Launch dispatch over the template parameters.
// ---------------------------------------------------------------------------
template <template <int> class F, typename... Args>
int dispatch_g(int g, Args&&... args) {
  switch (g) {
    case 4: return F<4>::run(args...);
    case 8: return F<8>::run(args...);
    case 16: return F<16>::run(args...);
    case 32: return F<32>::run(args...);
    case 64: return F<64>::run(args...);
  }
  return fail(FEM_ERR_ARG, "lanes per row must be 4, 8, 16, 32 or 64");
}

int padded_blocks(int rows, int rows_per_block) {
  int n = (rows + rows_per_block - 1) / rows_per_block;
  if (n > 8) n = (n + 7) & ~7;
  return std::max(n, 1);
}

template <int RB>
struct LaunchAssemble {
  static int run(bool hex, const RowArgs& ra, const int* groups, int ngroups, const float* x, float h, float* vals,
                 float* minv, size_t lds, hipStream_t st) {
    // The static row arrays share the 160 KiB with the dynamic staging.
    const size_t cap = 160 * 1024 - 2048;
    if (lds > cap) return fail(FEM_ERR_UNSUPPORTED, "assembly staging of one row exceeds the LDS");
    auto kern = hex ? k_assemble<RB, EL_ANY> : k_assemble<RB, EL_TET>;
    static size_t attr[2] = {0, 0};
    if (lds > 48 * 1024 && lds > attr[hex]) {
      HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)cap));
      attr[hex] = cap;
    }
    hipLaunchKernelGGL(kern, dim3(ngroups), dim3(BLOCK), lds, st, ra, groups, x, h, h * h, vals, minv);
    HIP_TRY(hipGetLastError());
    return FEM_OK;
  }
};

template <int G>
struct LaunchApplyOp {
  static int run(const RowArgs& ra, const float* vals, const float* x, float* y, hipStream_t st) {
    hipLaunchKernelGGL(k_apply_op<G>, dim3((ra.P + BLOCK / G - 1) / (BLOCK / G)), dim3(BLOCK), 0, st, ra, vals, x,
                       y);
    HIP_TRY(hipGetLastError());
    return FEM_OK;
  }
};

template <int G>
struct LaunchRhs {
  static int run(bool hex, const RowArgs& ra, const float* x, const float* v, const float* act, const float* ext,
                 const float* vals, float* out, unsigned* rmax, int terms, int mode, float h, hipStream_t st) {
    int nb = padded_blocks(ra.P, BLOCK / G);
    auto kern = hex ? k_rhs<G, EL_ANY> : k_rhs<G, EL_TET>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(BLOCK), 0, st, ra, x, v, act, ext, vals, out, rmax, terms, mode, h);
    HIP_TRY(hipGetLastError());
    return FEM_OK;
  }
};

template <int G>
struct LaunchVjpRows {
  static int run(bool hex, const RowArgs& ra, const float* x, const float* v, const float* act, const float* z,
                 const float* vals, float h, float* xbar, float* vbar, hipStream_t st) {
    int nb = padded_blocks(ra.P, BLOCK / G);
    auto kern = hex ? k_vjp_rows<G, EL_ANY> : k_vjp_rows<G, EL_TET>;
    hipLaunchKernelGGL(kern, dim3(nb), dim3(BLOCK), 0, st, ra, x, v, act, z, vals, h, xbar, vbar);
    HIP_TRY(hipGetLastError());
    return FEM_OK;
  }
};

// Rows per thread K and threads per workgroup NT of the persistent CG
// (res: the resident variant, RES_K rows and RES_NT threads; queue: the
// workgroups take chunks from the device queue).
template <bool Q>
const void* cg_kernel_ptr_q(int k, int nt, bool res) {
  if (res) return (k == RES_K && nt == RES_NT) ? (const void*)k_cg<RES_K, RES_NT, true, Q> : nullptr;
  if (nt == 512) {
    switch (k) {
      case 1: return (const void*)k_cg<1, 512, false, Q>;
      case 2: return (const void*)k_cg<2, 512, false, Q>;
    }
    return nullptr;
  }
  switch (k) {
    case 1: return (const void*)k_cg<1, 256, false, Q>;
    case 2: return (const void*)k_cg<2, 256, false, Q>;
    case 4: return (const void*)k_cg<4, 256, false, Q>;
    case 8: return (const void*)k_cg<8, 256, false, Q>;
  }
  return nullptr;
}
const void* cg_kernel_ptr(int k, int nt, bool res, bool queue) {
  return queue ? cg_kernel_ptr_q<true>(k, nt, res) : cg_kernel_ptr_q<false>(k, nt, res);
}

}  // namespace

// ---------------------------------------------------------------------------
// System object.
// ---------------------------------------------------------------------------
struct fem_system {
  int P = 0, T = 0, H = 0, S = 0, R = 0, nfixed = 0, ninst = 1;
  Pattern pt;
  std::vector<int32_t> inst_off;
  // device copies (library-owned)
  int *d_offsets = nullptr, *d_cols = nullptr, *d_diag = nullptr;
  int *d_tinc_ptr = nullptr, *d_tinc = nullptr, *d_hinc_ptr = nullptr, *d_hinc = nullptr;
  uint32_t *d_tslots = nullptr, *d_hslots = nullptr;
  int* d_slot_ptr = nullptr;
  uint16_t* d_contrib = nullptr;
  int *d_tet_idx = nullptr, *d_hex_idx = nullptr;
  float *d_tet_pose = nullptr, *d_tet_mat = nullptr, *d_hex_invJ = nullptr, *d_hex_detw = nullptr,
        *d_hex_mat = nullptr;
  float *d_hex_gp_t = nullptr, *d_hex_gp_f = nullptr;  // k_hex_gauss records (H > 0)
  // Particle-particle contact (fem_set_particle_contact): grid storage,
  // coefficients, and whether the step evaluates the term.
  ContactGrid contact;
  ContactParams cparams{0.f, 0.f, 0.f, 0.f, 0.f};
  bool contact_on = false;
  SoftContacts soft;  // particle-shape contact list (fem_set_soft_contacts)
  bool soft_on = false;
  size_t soft_temp_cap = 0;
  float max_radius = 0.f;
  // Spring bend torque (fem_set_spring_torque).
  float* d_sp_dir = nullptr;
  bool torque_on = false;
  float sp_kt = 0.f, sp_kw = 0.f;
  // Linear solver of the step (fem_set_solver): Krylov method, point-Jacobi
  // preconditioner blocks and the work storage of the methods besides PCG.
  int kr_method = KR_CG, kr_precond = KP_BLOCK, kr_restart = 31;
  float* d_pjac = nullptr;  // [P][9] when kr_precond != KP_BLOCK
  float* d_kr_work = nullptr;
  double *d_kr_part = nullptr, *d_kr_sc = nullptr;
  // Blocks of the configured preconditioner (M in z = M r).
  const float* precond_blocks() const { return kr_precond == KP_BLOCK ? b_minv : d_pjac; }
  float *d_mass = nullptr, *d_inv_mass = nullptr, *d_radius = nullptr, *d_anchor = nullptr;
  int* d_flags = nullptr;
  unsigned char* d_row_fixed = nullptr;
  int* d_fixed = nullptr;
  int* d_row_inst = nullptr;
  int *d_sinc_ptr = nullptr, *d_sinc = nullptr, *d_sp_idx = nullptr, *d_rinc_ptr = nullptr, *d_rinc = nullptr,
      *d_tri_idx = nullptr;
  float *d_sp_rest = nullptr, *d_sp_ke = nullptr, *d_sp_kd = nullptr, *d_tri_pose = nullptr, *d_tri_mat = nullptr,
        *d_row_tri = nullptr;
  float2* d_slot_spring = nullptr;
  const float* d_gravity = nullptr;  // caller-owned device gravity (fem_bind_gravity)
  // Kernels with Hex8, spring or triangle code (EL_ANY) are needed.
  bool full() const { return H > 0 || S > 0 || R > 0; }
  // solve scratch
  float* d_xw = nullptr;
  unsigned *d_rhs_max = nullptr, *d_cgflags = nullptr;
  double *d_norms = nullptr, *d_dstats = nullptr;  // checked-solve residual norms and running maxima
  double* d_part = nullptr;
  int *d_stats = nullptr, *d_err = nullptr;
  long spin_limit = BARRIER_SPIN_LIMIT;  // bound of every persistent-CG wait
  unsigned char* static_scratch = nullptr;  // this device's `scratch` (device address)
  int *d_chunk_rows = nullptr, *d_chunk_inst = nullptr, *d_inst_chunks = nullptr, *d_chunk_tiles = nullptr,
      *d_tile_rows = nullptr, *d_tile_cptr = nullptr, *d_tile_cols = nullptr;
  uint16_t* d_lidx = nullptr;
  int *d_res_thr = nullptr, *d_res_row = nullptr;
  uint64_t* d_res_mask = nullptr;
  int nchunks = 0, cg_k = 1, cg_nt = 256;
  std::vector<int32_t> h_chunk_rows;  // [nchunks + 1] rows of the current CG plan
  size_t cap_tile_cols = 0, cap_lidx = 0;  // storage of the current CG plan
  bool cg_res = false;
  int cg_variant_request = 0;  // 0 any, 1 streamed, 2 resident
  std::vector<int32_t> cg_groups;  // chunk boundaries of the CG launch groups (one group)
  int* d_chunk_order = nullptr;    // queued CG launch: order of the chunk queue (null: one workgroup per chunk)
  int cg_grid = 0;                 // workgroups of the CG launch
  bool cg_queue = false;           // the launch takes chunks from the device queue
  // bound
  int* b_offsets = nullptr;
  int* b_cols = nullptr;
  float *b_vals = nullptr, *b_minv = nullptr, *b_dv = nullptr;
  fem_params prm{};
  int elem_g = 4, asm_rb = 16, cg_wg_request = 0;
  size_t asm_lds = 0;
  int* d_asm_groups = nullptr;  // [asm_ngroups][ASM_HDR] record of each assembly workgroup
  int asm_ngroups = 0;
  int asm_cap = 0;  // fixed assembly group count of a reserved topology pool (0: none)
  int max_resident = 0;
  std::vector<void*> allocs;
  std::unordered_map<const void*, size_t> alloc_bytes;  // size of every live allocation
  bool keep_replaced = false;
  int64_t b_cap = 0;  // block capacity of the bound offsets/columns/values storage
  // Host copies of the element data a topology rebuild needs besides the tets.
  std::vector<int32_t> h_hex_idx, h_sp_idx, h_tri_idx;
  std::vector<float> h_sp_ke, h_sp_kd, h_tri_pose, h_tri_mat;
  // Reserved topology pool (streaming.py): original node count, the last
  // published elements (host) and the positions of the last assembly (device).
  int topo_count = 0;
  Pattern spare;  // incidence storage of the previous edit
  std::vector<int32_t> old_nodes;
  std::vector<float> old_poses, old_mats;
  std::vector<unsigned char> old_active;
  float* d_lin_q = nullptr;
  // Compact Krylov pattern of a reserved pool (streaming.py:209-384): the live
  // slots -- diagonals and slots with element contributions -- of the reserved
  // rows, their reserved slot numbers, and their values, gathered after every
  // assembly.  The CG runs over it; A itself stays in the reserved pattern.
  bool compact = false;
  std::vector<int64_t> c_offsets;
  std::vector<int32_t> c_columns, c_map;
  int *d_coffsets = nullptr, *d_ccols = nullptr, *d_cmap = nullptr;
  float* d_cvals = nullptr;
  const std::vector<int64_t>& kp_offsets() const { return compact ? c_offsets : pt.offsets; }
  const std::vector<int32_t>& kp_columns() const { return compact ? c_columns : pt.columns; }
  // Live slots of the reserved rows: the diagonal and every slot with an
  // element contribution (_compact_rows streaming.py:209-231).
  void build_compact() {
    auto live = [&](int i, int64_t k) { return pt.slot_ptr[k + 1] > pt.slot_ptr[k] || pt.columns[k] == i; };
    c_offsets.assign(P + 1, 0);
    parallel_pieces(P, 2048, [&](int r0, int r1) {  // live count per row
      for (int i = r0; i < r1; ++i) {
        int64_t n = 0;
        for (int64_t k = pt.offsets[i]; k < pt.offsets[i + 1]; ++k) n += live(i, k);
        c_offsets[i + 1] = n;
      }
    });
    for (int i = 0; i < P; ++i) c_offsets[i + 1] += c_offsets[i];
    c_columns.resize(c_offsets[P]);
    c_map.resize(c_offsets[P]);
    parallel_pieces(P, 2048, [&](int r0, int r1) {
      for (int i = r0; i < r1; ++i) {
        int64_t o = c_offsets[i];
        for (int64_t k = pt.offsets[i]; k < pt.offsets[i + 1]; ++k)
          if (live(i, k)) {
            c_columns[o] = pt.columns[k];
            c_map[o++] = (int32_t)k;
          }
      }
    });
  }
  int upload_compact() {
    std::vector<int32_t> off32(P + 1);
    for (int i = 0; i <= P; ++i) off32[i] = (int32_t)c_offsets[i];
    HIP_TRY(hipMemcpy(d_coffsets, off32.data(), off32.size() * 4, hipMemcpyHostToDevice));
    if (!c_columns.empty()) {
      HIP_TRY(hipMemcpy(d_ccols, c_columns.data(), c_columns.size() * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d_cmap, c_map.data(), c_map.size() * 4, hipMemcpyHostToDevice));
    }
    return FEM_OK;
  }
  // Component scaling (component_scaling.py), caller-owned storage.
  float *cs_scale = nullptr, *cs_vals = nullptr, *cs_minv = nullptr, *cs_rhs = nullptr;
  double* cs_total = nullptr;
  int cs_cap = 0;

  template <typename T_>
  int upload(T_** dst, const T_* src, size_t n) {
    *dst = nullptr;
    if (n == 0) n = 1;
    HIP_TRY(hipMalloc((void**)dst, n * sizeof(T_)));
    allocs.push_back(*dst);
    alloc_bytes[*dst] = n * sizeof(T_);
    if (src) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T_), hipMemcpyHostToDevice));
    else HIP_TRY(hipMemset(*dst, 0, n * sizeof(T_)));
    return FEM_OK;
  }
  template <typename T_>
  int alloc(T_** dst, size_t n) {
    return upload<T_>(dst, nullptr, n);
  }
  // A replaced allocation.  While keep_replaced is set it stays allocated
  // until the system is destroyed (a graph captured earlier may still name
  // it); otherwise it is freed now.
  void drop(void* p) {
    if (!p || keep_replaced) return;
    (void)hipFree(p);
    allocs.erase(std::remove(allocs.begin(), allocs.end(), p), allocs.end());
    alloc_bytes.erase(p);
  }
  // Whether allocation p holds at least `bytes`.
  bool holds(const void* p, size_t bytes) const {
    if (!p) return bytes == 0;
    const auto it = alloc_bytes.find(p);
    return it != alloc_bytes.end() && it->second >= bytes;
  }
  // n elements of src into the existing allocation *dst (which holds them).
  template <typename T_>
  int write(T_* dst, const T_* src, size_t n) {
    if (n) HIP_TRY(hipMemcpy(dst, src, n * sizeof(T_), hipMemcpyHostToDevice));
    return FEM_OK;
  }
  // Device copy of n elements into a new allocation of at least cap elements.
  template <typename T_>
  int upload_cap(T_** dst, const T_* src, size_t n, size_t cap) {
    if (int rc = upload(dst, (const T_*)nullptr, std::max(n, cap))) return rc;
    return write(*dst, src, n);
  }
  // Device copy of n elements of src into a new allocation of at least cap
  // elements; the previous allocation *dst (if any) is dropped.
  template <typename T_>
  int replace(T_** dst, const T_* src, size_t n, size_t cap = 0) {
    T_* old = *dst;
    *dst = nullptr;
    if (int rc = alloc(dst, std::max(n, cap))) return rc;
    if (n) HIP_TRY(hipMemcpy(*dst, src, n * sizeof(T_), hipMemcpyHostToDevice));
    drop(old);
    return FEM_OK;
  }

  RowArgs row_args() const {
    RowArgs a;
    a.P = P;
    a.offsets = d_offsets;
    a.cols = d_cols;
    a.diag_rel = d_diag;
    a.tinc_ptr = d_tinc_ptr;
    a.tinc = d_tinc;
    a.tslots = d_tslots;
    a.hinc_ptr = d_hinc_ptr;
    a.hinc = d_hinc;
    a.hslots = d_hslots;
    a.slot_ptr = d_slot_ptr;
    a.contrib = d_contrib;
    a.tet_idx = d_tet_idx;
    a.tet_pose = d_tet_pose;
    a.tet_mat = d_tet_mat;
    a.hex_idx = d_hex_idx;
    a.hex_invJ = d_hex_invJ;
    a.hex_detw = d_hex_detw;
    a.hex_mat = d_hex_mat;
    a.hex_gp_t = d_hex_gp_t;
    a.hex_gp_f = d_hex_gp_f;
    a.fpart = contact_on ? contact.force : nullptr;
    a.fsoft = soft_on ? soft.force : nullptr;
    a.mass = d_mass;
    a.inv_mass = d_inv_mass;
    a.radius = d_radius;
    a.flags = d_flags;
    a.row_fixed = d_row_fixed;
    a.row_inst = ninst > 1 ? d_row_inst : nullptr;
    a.sinc_ptr = S > 0 ? d_sinc_ptr : nullptr;
    a.sinc = d_sinc;
    a.sp_idx = d_sp_idx;
    a.sp_rest = d_sp_rest;
    a.sp_ke = d_sp_ke;
    a.sp_kd = d_sp_kd;
    a.sp_dir = torque_on ? d_sp_dir : nullptr;
    a.sp_kt = sp_kt;
    a.sp_kw = sp_kw;
    a.rinc_ptr = R > 0 ? d_rinc_ptr : nullptr;
    a.rinc = d_rinc;
    a.tri_idx = d_tri_idx;
    a.tri_pose = d_tri_pose;
    a.tri_mat = d_tri_mat;
    a.slot_spring = S > 0 ? d_slot_spring : nullptr;
    a.row_tri = R > 0 ? d_row_tri : nullptr;
    a.mn = prm.min_stretch;
    a.mx = prm.max_stretch;
    a.has_ground = prm.has_ground;
    for (int k = 0; k < 4; ++k) a.gn[k] = prm.ground[k];
    a.ke = prm.ground_ke;
    a.kd = prm.ground_kd;
    a.kf = prm.ground_kf;
    a.gmu = prm.ground_mu;
    a.lin_damp = prm.linear_damping;
    for (int k = 0; k < 3; ++k) a.grav[k] = prm.gravity[k];
    a.gmag = sqrtf(prm.gravity[0] * prm.gravity[0] + prm.gravity[1] * prm.gravity[1] +
                   prm.gravity[2] * prm.gravity[2]);
    a.gravity = d_gravity;
    a.use_act = prm.enable_tet_activations;
    return a;
  }

  // Constant spring blocks: per slot the signed (ke, kd) sums of the spring
  // triplets landing on it -- +1 per spring on its endpoints' diagonals
  // (build_system_matrix_diagonal_kernel kernels.py:294-323 counts a spring
  // once per endpoint row), -1 per (i, j) and (j, i) block
  // (build_system_matrix_sparse_kernel :326-366), so a spring whose two
  // endpoints coincide nets -1 on its diagonal as in the reference.  A spring
  // with a -1 endpoint keeps its diagonal term at the other endpoint (the
  // diagonal kernel tests each endpoint); its off-diagonal triplets carry row
  // or column -1 and are dropped.
  std::vector<float2> spring_slots() const {
    std::vector<float2> slot_spring(pt.columns.size(), make_float2(0.f, 0.f));
    auto slot_of = [&](int r, int c) {
      const int32_t* row = pt.columns.data() + pt.offsets[r];
      return (size_t)(pt.offsets[r] + (std::lower_bound(row, row + (pt.offsets[r + 1] - pt.offsets[r]), c) - row));
    };
    for (int k = 0; k < S; ++k) {
      const int p = h_sp_idx[2 * k], q = h_sp_idx[2 * k + 1];
      const float ke = h_sp_ke[k], kd = h_sp_kd[k];
      if (p >= 0 && q >= 0)
        for (size_t sl : {slot_of(p, q), slot_of(q, p)}) {
          slot_spring[sl].x -= ke;
          slot_spring[sl].y -= kd;
        }
      if (p >= 0) {
        slot_spring[slot_of(p, p)].x += ke;
        slot_spring[slot_of(p, p)].y += kd;
      }
      if (q >= 0 && q != p) {
        slot_spring[slot_of(q, q)].x += ke;
        slot_spring[slot_of(q, q)].y += kd;
      }
    }
    return slot_spring;
  }

  // Device copies of everything derived from the pattern: rows, incidences,
  // their slots, per-slot contribution lists and the constant spring blocks.
  // tinc_cap / contrib_cap reserve room for later incidence updates.
  int upload_pattern(size_t tinc_cap = 0, size_t contrib_cap = 0, size_t block_cap = 0) {
    std::vector<int32_t> off32(P + 1), tptr(P + 1), hptr(P + 1), sptr(P + 1), rptr(P + 1);
    for (int i = 0; i <= P; ++i) {
      off32[i] = (int32_t)pt.offsets[i];
      tptr[i] = (int32_t)pt.tinc_ptr[i];
      hptr[i] = (int32_t)pt.hinc_ptr[i];
      sptr[i] = (int32_t)pt.sinc_ptr[i];
      rptr[i] = (int32_t)pt.rinc_ptr[i];
    }
    int rc;
    if ((rc = replace(&d_offsets, off32.data(), off32.size())) ||
        (rc = replace(&d_cols, pt.columns.data(), pt.columns.size(), block_cap)) ||
        (rc = replace(&d_diag, pt.diag_rel.data(), pt.diag_rel.size())) ||
        (rc = replace(&d_tinc_ptr, tptr.data(), tptr.size())) ||
        (rc = replace(&d_tinc, pt.tinc.data(), pt.tinc.size(), tinc_cap)) ||
        (rc = replace(&d_tslots, pt.tslots.data(), pt.tslots.size(), tinc_cap)) ||
        (rc = replace(&d_hinc_ptr, hptr.data(), hptr.size())) ||
        (rc = replace(&d_hinc, pt.hinc.data(), pt.hinc.size())) ||
        (rc = replace(&d_hslots, pt.hslots.data(), pt.hslots.size())) ||
        (rc = replace(&d_slot_ptr, pt.slot_ptr.data(), pt.slot_ptr.size(), block_cap ? block_cap + 1 : 0)) ||
        (rc = replace(&d_contrib, pt.contrib.data(), pt.contrib.size(), contrib_cap)))
      return rc;
    if (S > 0 || R > 0) {
      const std::vector<float2> slot_spring = S > 0 ? spring_slots() : std::vector<float2>();
      if ((rc = replace(&d_sinc_ptr, sptr.data(), sptr.size())) ||
          (rc = replace(&d_sinc, pt.sinc.data(), pt.sinc.size())) ||
          (rc = replace(&d_rinc_ptr, rptr.data(), rptr.size())) ||
          (rc = replace(&d_rinc, pt.rinc.data(), pt.rinc.size())) ||
          (rc = replace(&d_slot_spring, slot_spring.data(), slot_spring.size(), S > 0 ? block_cap : 0)))
        return rc;
    }
    return FEM_OK;
  }

  // Dynamic LDS bytes of the assembly of rows r0 .. r1 - 1.
  size_t assemble_lds(int r0, int r1) const {
    const size_t nt = (size_t)(pt.tinc_ptr[r1] - pt.tinc_ptr[r0]);
    const size_t nh = (size_t)(pt.hinc_ptr[r1] - pt.hinc_ptr[r0]);
    const size_t ns = (size_t)(pt.offsets[r1] - pt.offsets[r0]);
    const size_t nc = (size_t)(pt.slot_ptr[pt.offsets[r1]] - pt.slot_ptr[pt.offsets[r0]]);
    return ((size_t)ST_TET * nt + (size_t)ST_HEX * nh + ns + 1) * sizeof(float) + nc * sizeof(uint16_t) + ns;
  }

  // Group records of the assembly (ASM_HDR ints per group, padding groups
  // empty at row P) for the row boundaries g.
  std::vector<int32_t> assemble_headers(const std::vector<int32_t>& g, size_t ngroups) const {
    std::vector<int32_t> hd(ASM_HDR * ngroups, 0);
    for (size_t k = 0; k < ngroups; ++k) {
      const int r0 = k + 1 < g.size() ? g[k] : P, r1 = k + 1 < g.size() ? g[k + 1] : P;
      int32_t* o = &hd[ASM_HDR * k];
      o[0] = r0;
      o[1] = r1;
      o[2] = pt.tinc_ptr[r0];
      o[3] = pt.tinc_ptr[r1];
      o[4] = pt.hinc_ptr[r0];
      o[5] = pt.hinc_ptr[r1];
      o[6] = (int32_t)pt.offsets[r0];
      o[7] = (int32_t)pt.offsets[r1];
      o[8] = pt.slot_ptr[pt.offsets[r0]];
      o[9] = pt.slot_ptr[pt.offsets[r1]];
    }
    return hd;
  }

  // Row groups of the assembly: consecutive rows packed greedily, at most rb
  // rows, at most BLOCK element incidences (one pass of phase 1) and at most
  // ASM_LDS_BUDGET bytes of staging per group; a row that alone exceeds the
  // limits forms its own group.  `worst` receives the largest staging.
  void assemble_groups(int rb, std::vector<int32_t>& g, size_t& worst) const {
    g.assign(1, 0);
    worst = 16;
    int r0 = 0;
    while (r0 < P) {
      int r1 = r0 + 1;
      while (r1 < P && r1 - r0 < rb) {
        const int inc = (pt.tinc_ptr[r1 + 1] - pt.tinc_ptr[r0]) + (pt.hinc_ptr[r1 + 1] - pt.hinc_ptr[r0]);
        if (inc > BLOCK || assemble_lds(r0, r1 + 1) > ASM_LDS_BUDGET) break;
        ++r1;
      }
      worst = std::max(worst, assemble_lds(r0, r1));
      g.push_back(r1);
      r0 = r1;
    }
  }

  // Installs the assembly row groups.  The group count is padded to a
  // multiple of 8 with empty groups (XCD swizzle).  With a reserved topology
  // pool the launch keeps a fixed group count and staging size (asm_cap,
  // asm_lds) and the table is rewritten in place, so that a captured launch
  // stays valid across topology updates; otherwise the table is replaced.
  int plan_assemble(int rb) {
    std::vector<int32_t> g;
    size_t worst = 0;
    assemble_groups(rb, g, worst);
    const int n = (int)g.size() - 1;
    if (asm_cap > 0) {
      if (n > asm_cap || worst > asm_lds)
        return fail(FEM_ERR_UNSUPPORTED, "the edited topology exceeds the reserved assembly launch");
      g.resize(asm_cap + 1, P);
      const std::vector<int32_t> hd = assemble_headers(g, asm_cap);
      HIP_TRY(hipMemcpy(d_asm_groups, hd.data(), hd.size() * sizeof(int32_t), hipMemcpyHostToDevice));
      asm_rb = rb;
      return FEM_OK;
    }
    // A few spare (empty) groups and, when the bound storage leaves room for
    // more blocks, the full staging budget: a topology edit without a pool
    // that still fits this launch is then written in place.
    const bool room = b_cap > (int64_t)pt.offsets[P];
    const int spare = std::max(8, n / 256);
    const int padded = n > 8 ? (n + spare + 7) & ~7 : std::max(n, 1) + spare;
    while ((int)g.size() < padded + 1) g.push_back(P);
    const std::vector<int32_t> hd = assemble_headers(g, padded);
    int* dg = nullptr;
    if (int rc = upload(&dg, hd.data(), hd.size())) return rc;
    drop(d_asm_groups);
    d_asm_groups = dg;
    asm_ngroups = padded;
    asm_rb = rb;
    asm_lds = room ? std::max(worst, ASM_LDS_BUDGET) : worst;
    return FEM_OK;
  }

  // Fixed assembly launch for a reserved topology pool: a group count that
  // bounds the greedy packing of any element set within the capacities (each
  // closed group either holds rb rows or, together with the next row, exceeds
  // an incidence or staging limit) and the full staging budget.
  int reserve_assemble(int rb) {
    const double inc_total = 4.0 * T;
    const double lds_total = (double)(ST_TET * inc_total + pt.offsets[P] + 2.0 * P) * sizeof(float) +
                             16.0 * T * sizeof(uint16_t) + (double)pt.offsets[P];
    std::vector<int32_t> g;
    size_t worst = 0;
    assemble_groups(rb, g, worst);
    long cap = (P + rb - 1) / rb + (long)std::ceil(2.0 * inc_total / BLOCK) +
               (long)std::ceil(2.0 * lds_total / ASM_LDS_BUDGET) + 8;
    cap = std::max<long>(cap, (long)g.size() - 1);
    cap = (cap + 7) & ~7L;
    int* dg = nullptr;
    if (int rc = alloc(&dg, (size_t)ASM_HDR * cap)) return rc;
    drop(d_asm_groups);
    d_asm_groups = dg;
    asm_cap = (int)cap;
    asm_ngroups = (int)cap;
    asm_lds = std::max(worst, ASM_LDS_BUDGET);
    return plan_assemble(rb);
  }

  // Row partition of the persistent CG: per instance, contiguous chunks,
  // row-aligned tiles inside every chunk.  Preference order: the resident
  // kernel (every chunk's blocks fit on chip, one workgroup per CU), the
  // streamed 512-thread kernel (<= 2 rows per thread, one workgroup per CU),
  // the streamed 256-thread kernel (<= 8 rows per thread).  cg_variant_request
  // 1 / 2 restricts the choice to streamed / resident.  All state changes are
  // committed only after every check has passed.
  int plan_cg() {
    std::string why;
    if (cg_variant_request != 1) {
      const int rc = plan_cg_nt(RES_NT, true);
      if (rc == FEM_OK || cg_variant_request == 2) return rc;
      why = g_error + "; ";
    }
    const int rc = plan_cg_nt(512, false);
    if (rc == FEM_OK) return rc;
    why += g_error;
    const int rc2 = plan_cg_nt(256, false);
    if (rc2 != FEM_OK) g_error = why + "; " + g_error;
    return rc2;
  }

  // Topology edit without a reserved pool, in place: the edited pattern `np`
  // goes into the existing pattern arrays, the assembly keeps its launch
  // (group count and staging) with the groups rewritten, and the resident CG
  // is re-planned over its chunk rows within its storage -- the same device
  // arrays and launches as before, so a graph captured before the edit
  // replays the edited topology (as the reference re-sorts into the same
  // A_bsr storage, solver.py:920-935).  Returns FEM_ERR_UNSUPPORTED, with
  // nothing changed, when the edit does not fit (streamed CG plan, arrays or
  // launches too small); the caller then reallocates.
  int edit_pattern_in_place(Pattern& np) {
    const size_t nnz = (size_t)np.offsets[P];
    if (!cg_res || compact || h_chunk_rows.empty())
      return fail(FEM_ERR_UNSUPPORTED, "in-place topology edits need the resident CG plan");
    if (!holds(d_cols, 4 * nnz) || !holds(d_slot_ptr, 4 * (nnz + 1)) || !holds(d_tinc, 4 * np.tinc.size()) ||
        !holds(d_tslots, 4 * np.tslots.size()) || !holds(d_hinc, 4 * np.hinc.size()) ||
        !holds(d_hslots, 4 * np.hslots.size()) || !holds(d_contrib, 2 * np.contrib.size()) ||
        !holds(d_sinc, 4 * np.sinc.size()) || !holds(d_rinc, 4 * np.rinc.size()) ||
        (S > 0 && !holds(d_slot_spring, sizeof(float2) * nnz)))
      return fail(FEM_ERR_UNSUPPORTED, "the edited pattern exceeds the pattern storage");
    std::swap(pt, np);  // plan over the edited pattern; swapped back on failure
    std::vector<int32_t> g;
    size_t worst = 0;
    assemble_groups(asm_rb, g, worst);
    int rc = FEM_OK;
    if ((int)g.size() - 1 > asm_ngroups || worst > asm_lds)
      rc = fail(FEM_ERR_UNSUPPORTED, "the edited topology exceeds the assembly launch");
    else
      rc = plan_cg_nt(cg_nt, true, true);  // checks everything before it writes
    if (rc != FEM_OK) {
      std::swap(pt, np);
      return rc == FEM_ERR_HIP ? rc : fail(FEM_ERR_UNSUPPORTED, g_error);
    }
    std::vector<int32_t> off32(P + 1), tptr(P + 1), hptr(P + 1), sptr(P + 1), rptr(P + 1);
    for (int i = 0; i <= P; ++i) {
      off32[i] = (int32_t)pt.offsets[i];
      tptr[i] = (int32_t)pt.tinc_ptr[i];
      hptr[i] = (int32_t)pt.hinc_ptr[i];
      sptr[i] = (int32_t)pt.sinc_ptr[i];
      rptr[i] = (int32_t)pt.rinc_ptr[i];
    }
    if ((rc = write(d_offsets, off32.data(), off32.size())) || (rc = write(d_cols, pt.columns.data(), nnz)) ||
        (rc = write(d_diag, pt.diag_rel.data(), pt.diag_rel.size())) ||
        (rc = write(d_tinc_ptr, tptr.data(), tptr.size())) ||
        (rc = write(d_tinc, pt.tinc.data(), pt.tinc.size())) ||
        (rc = write(d_tslots, pt.tslots.data(), pt.tslots.size())) ||
        (rc = write(d_hinc_ptr, hptr.data(), hptr.size())) || (rc = write(d_hinc, pt.hinc.data(), pt.hinc.size())) ||
        (rc = write(d_hslots, pt.hslots.data(), pt.hslots.size())) ||
        (rc = write(d_slot_ptr, pt.slot_ptr.data(), pt.slot_ptr.size())) ||
        (rc = write(d_contrib, pt.contrib.data(), pt.contrib.size())))
      return rc;
    if (S > 0 || R > 0) {
      const std::vector<float2> slot_spring = S > 0 ? spring_slots() : std::vector<float2>();
      if ((rc = write(d_sinc_ptr, sptr.data(), sptr.size())) || (rc = write(d_sinc, pt.sinc.data(), pt.sinc.size())) ||
          (rc = write(d_rinc_ptr, rptr.data(), rptr.size())) || (rc = write(d_rinc, pt.rinc.data(), pt.rinc.size())) ||
          (rc = write(d_slot_spring, slot_spring.data(), slot_spring.size())))
        return rc;
    }
    g.resize(asm_ngroups + 1, P);
    const std::vector<int32_t> hd = assemble_headers(g, asm_ngroups);
    return write(d_asm_groups, hd.data(), hd.size());
  }

  // keep_rows: re-plan the resident kernel for the current Krylov pattern over
  // the previous plan's chunk rows, written into the previous plan's storage
  // (same launch, so a recorded graph stays valid); fails if a chunk no longer
  // fits on chip or the storage is too small.
  int plan_cg_nt(const int nt, const bool res, const bool keep_rows = false) {
    const std::vector<int64_t>& koff = kp_offsets();
    const std::vector<int32_t>& kcol = kp_columns();
    if (keep_rows && (!res || !cg_res || nt != cg_nt || h_chunk_rows.empty()))
      return fail(FEM_ERR_STATE, "a kept-row CG plan needs a previous resident plan");
    int dev = 0, ncu = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int kmax = res ? RES_K : nt == 512 ? 2 : 8;
    const long max_rows_chunk = (long)kmax * nt;
    std::vector<int> per(ninst, 1);
    std::vector<double> cost(ninst);
    long need = 0;
    double total = 0;
    for (int k = 0; k < ninst; ++k) {
      long rows = inst_off[k + 1] - inst_off[k];
      per[k] = (int)std::max(1L, (rows + max_rows_chunk - 1) / max_rows_chunk);
      need += per[k];
      cost[k] = (double)(koff[inst_off[k + 1]] - koff[inst_off[k]]) + 8.0 * rows;
      total += cost[k];
    }
    // Default chunk count: about 512 rows per workgroup for small meshes, at
    // most one 512-thread workgroup per CU for large ones (2 rows per thread
    // of the 256-thread kernel); the resident kernel also needs enough chunks
    // for every chunk's blocks to fit on chip.  cg_wg_request overrides.
    // A single resident system takes about 160 rows per workgroup: below one
    // workgroup per CU the iteration is bound by each chunk's serial block
    // products, not by the barrier (measured on an 18,785-node Hex8 mesh:
    // 13.8 us per iteration at 48 chunks, 11.4 us at 96-128, 12.0 us at
    // 256).  Batches keep one chunk per instance where that suffices, so an
    // instance is split as a looped one-workgroup solve of it would be.
    const long res_blocks = (long)RES_NT * (RES_JR + RES_JL + RES_JS);  // blocks one resident chunk holds
    const long rows_per_wg = (res && ninst == 1) ? 160 : 512;
    // (The 256-thread streamed kernel holds two workgroups per CU.)
    long natural = (nt == 512 || res) ? std::min<long>(P / rows_per_wg, ncu) : std::min<long>(P / 512, 2L * ncu);
    if (res)
      natural = std::min<long>(ncu, std::max<long>(natural, (long)std::ceil((double)koff[P] / (0.95 * res_blocks))));
    long target = cg_wg_request > 0 ? cg_wg_request : std::max<long>(need, natural);
    target = std::max(target, need);
    if (target > need && total > 0) {
      long extra = target - need;
      for (int k = 0; k < ninst; ++k) {
        long rows = inst_off[k + 1] - inst_off[k];
        long add = (long)std::floor(cost[k] / total * extra);
        per[k] = (int)std::min<long>(per[k] + add, std::max<long>(rows, 1));
      }
    }
    auto count = [&]() {
      long n = 0;
      for (int v : per) n += v;
      return n;
    };
    std::vector<int32_t> chunk_rows, chunk_inst, inst_chunks;
    std::vector<int32_t> stamp(res ? P : 0, -1);
    int maxrows = 0;
    for (int attempt = 0; !keep_rows; ++attempt) {
      long n = count();
      if (cg_wg_request == 0 && n > 8 && (n & 7)) {  // automatic count: pad to a multiple of 8 (XCD swizzle)
        int big = (int)(std::max_element(cost.begin(), cost.end()) - cost.begin());
        long rows = inst_off[big + 1] - inst_off[big];
        long pad = 8 - (n & 7);
        if (per[big] + pad <= rows) per[big] += (int)pad;
      }
      chunk_rows.clear();
      chunk_inst.clear();
      inst_chunks.assign(1, 0);
      maxrows = 0;
      int worst = -1;
      std::vector<int> over;  // instances with a chunk the kernel cannot hold
      for (int k = 0; k < ninst; ++k) {
        std::vector<int32_t> bnd;
        split_rows(koff, inst_off[k], inst_off[k + 1], per[k], bnd);
        bool bad = false;
        for (int j = 0; j < per[k]; ++j) {
          chunk_rows.push_back(bnd[j]);
          chunk_inst.push_back(k);
          if (bnd[j + 1] - bnd[j] > maxrows) {
            maxrows = bnd[j + 1] - bnd[j];
            worst = k;
          }
          if (res && !bad) {  // on-chip capacity: blocks and distinct columns of the chunk
            const int64_t kb = koff[bnd[j]], ke = koff[bnd[j + 1]];
            int ncols = 0;
            const int id = (int)chunk_rows.size();
            for (int64_t q = kb; q < ke; ++q)
              if (stamp[kcol[q]] != id) {
                stamp[kcol[q]] = id;
                ++ncols;
              }
            bad = ke - kb > res_blocks || ncols > RES_ZERO || bnd[j + 1] - bnd[j] > max_rows_chunk;
          }
        }
        if (bad) over.push_back(k);
        inst_chunks.push_back((int)chunk_rows.size());
      }
      if (maxrows <= max_rows_chunk && over.empty()) break;
      if (attempt > (res ? 256 : 64)) return fail(FEM_ERR_UNSUPPORTED, "cannot split the CG rows into small enough chunks");
      if (over.empty()) over.push_back(worst);
      for (int k : over) per[k] += res ? std::max(1, per[k] / 64) : 1;
      std::fill(stamp.begin(), stamp.end(), -1);
    }
    if (keep_rows) {
      chunk_rows.assign(h_chunk_rows.begin(), h_chunk_rows.end() - 1);
      chunk_inst.assign(chunk_rows.size(), 0);
      inst_chunks = {0, (int32_t)chunk_rows.size()};
      maxrows = 0;
      for (size_t c = 0; c < chunk_rows.size(); ++c) {
        const int r0 = chunk_rows[c], r1 = h_chunk_rows[c + 1];
        maxrows = std::max(maxrows, r1 - r0);
        const int64_t kb = koff[r0], ke = koff[r1];
        int ncols = 0;
        for (int64_t q = kb; q < ke; ++q)
          if (stamp[kcol[q]] != (int)c) {
            stamp[kcol[q]] = (int)c;
            ++ncols;
          }
        if (ke - kb > res_blocks || ncols > RES_ZERO)
          return fail(FEM_ERR_UNSUPPORTED, "the edited topology no longer fits the resident CG launch");
      }
    }
    chunk_rows.push_back(P);
    const int n_chunks = (int)chunk_inst.size();
    // Row-aligned tiles of at most CG_J * nt blocks inside every chunk (the
    // resident kernel: one tile per chunk).
    std::vector<int32_t> chunk_tiles(1, 0), tile_rows;
    for (int c = 0; c < n_chunks; ++c) {
      int r = chunk_rows[c];
      const int r1 = chunk_rows[c + 1];
      if (res) {
        tile_rows.push_back(r);
        r = r1;
      }
      while (r < r1) {
        tile_rows.push_back(r);
        int e = r;
        while (e < r1 && koff[e + 1] - koff[r] <= (int64_t)CG_J * nt) ++e;
        if (e == r) return fail(FEM_ERR_UNSUPPORTED, "row longer than a CG tile");
        r = e;
      }
      chunk_tiles.push_back((int)tile_rows.size());
    }
    // Distinct columns of every tile and the tile-local column number of
    // every block (the SpMV forms each distinct operand once per tile).
    const int ntiles = (int)tile_rows.size();
    std::vector<int32_t> tile_cptr(1, 0), tile_cols;
    std::vector<uint16_t> lidx(koff[P]);
    {
      // Per tile: mark its columns in a bitmap, read the marks back in
      // ascending order (the tile's distinct columns and their ranks), clear.
      std::vector<uint64_t> mark(((size_t)P + 63) / 64, 0);
      std::vector<int32_t> rank(P);
      for (int t = 0; t < ntiles; ++t) {
        const int ra = tile_rows[t], rb = t + 1 < ntiles ? tile_rows[t + 1] : P;
        const int64_t ka = koff[ra], kb = koff[rb];
        size_t lo = mark.size(), hi = 0;
        for (int64_t k = ka; k < kb; ++k) {
          const size_t w = (size_t)kcol[k] >> 6;
          mark[w] |= 1ull << (kcol[k] & 63);
          lo = std::min(lo, w);
          hi = std::max(hi, w + 1);
        }
        const size_t base = tile_cols.size();
        for (size_t w = lo; w < hi; ++w)
          for (uint64_t bits = mark[w]; bits; bits &= bits - 1) {
            const int32_t c = (int32_t)(64 * w + (size_t)__builtin_ctzll(bits));
            rank[c] = (int32_t)(tile_cols.size() - base);
            tile_cols.push_back(c);
          }
        for (size_t w = lo; w < hi; ++w) mark[w] = 0;
        for (int64_t k = ka; k < kb; ++k) lidx[k] = (uint16_t)rank[kcol[k]];
        tile_cptr.push_back((int32_t)tile_cols.size());
      }
    }
    // Resident kernel: thread t of a chunk holds blocks [t nb / NT, (t+1) nb / NT)
    // of the chunk's nb; a row run ends at a row's last block or at the
    // thread's last block.  Runs are numbered in (thread, block) order.
    std::vector<int32_t> res_thr, res_row;
    std::vector<uint64_t> res_mask;
    if (res) {
      res_thr.assign((size_t)n_chunks * nt * 4, 0);
      res_mask.assign((size_t)n_chunks * nt, 0);
      res_row.assign((size_t)2 * P, 0);
      // Chunks are independent (each writes its own threads' entries and its
      // own rows'), so they are mapped on host threads.
      std::atomic<int> why{0};
      parallel_pieces(n_chunks, 8, [&](int c0, int c1) {
      for (int c = c0; c < c1; ++c) {
        const int r0 = chunk_rows[c], r1 = chunk_rows[c + 1];
        const int64_t kb = koff[r0], nb = koff[r1] - kb;
        const int ncols = tile_cptr[c + 1] - tile_cptr[c];
        if (ncols > RES_ZERO) {
          why.store(1);
          return;
        }
        int run = 0, row = r0;
        for (int t = 0; t < nt; ++t) {
          const int64_t b0 = nb * t / nt, b1 = nb * (t + 1) / nt;
          if (b1 - b0 > RES_JR + RES_JL + RES_JS) {
            why.store(2);
            return;
          }
          int32_t* th = &res_thr[((size_t)c * nt + t) * 4];
          th[0] = (int32_t)b0;
          th[1] = (int32_t)(b1 - b0);
          th[2] = run;
          th[3] = (int32_t)(std::lower_bound(tile_cols.begin() + tile_cptr[c], tile_cols.begin() + tile_cptr[c + 1], r0) -
                            (tile_cols.begin() + tile_cptr[c]));
          uint64_t mask = 0;
          for (int64_t b = b0; b < b1; ++b) {
            while (koff[row + 1] <= kb + b) ++row;
            if (b == b0 || koff[row] == kb + b) {  // a run of `row` starts here
              if (res_row[2 * row + 1] == 0) res_row[2 * row] = run;
              ++res_row[2 * row + 1];
            }
            if (koff[row + 1] == kb + b + 1 || b + 1 == b1) {
              mask |= 1ull << (b - b0);
              ++run;
            }
          }
          res_mask[(size_t)c * nt + t] = mask;
        }
        if (run > RES_SEG) {
          why.store(3);
          return;
        }
      }
      });
      static const char* msg[4] = {"", "resident CG: too many distinct columns in a chunk",
                                   "resident CG: chunk exceeds the on-chip capacity",
                                   "resident CG: too many row runs in a chunk"};
      if (why.load()) return fail(FEM_ERR_UNSUPPORTED, msg[why.load()]);
    }
    if (keep_rows) {
      if (n_chunks != nchunks || tile_cols.size() > cap_tile_cols || lidx.size() > cap_lidx)
        return fail(FEM_ERR_UNSUPPORTED, "the edited topology no longer fits the reserved CG plan storage");
      HIP_TRY(hipMemcpy(d_res_thr, res_thr.data(), res_thr.size() * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d_res_mask, res_mask.data(), res_mask.size() * 8, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d_res_row, res_row.data(), res_row.size() * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d_tile_cptr, tile_cptr.data(), tile_cptr.size() * 4, hipMemcpyHostToDevice));
      if (!tile_cols.empty())
        HIP_TRY(hipMemcpy(d_tile_cols, tile_cols.data(), tile_cols.size() * 4, hipMemcpyHostToDevice));
      if (!lidx.empty()) HIP_TRY(hipMemcpy(d_lidx, lidx.data(), lidx.size() * 2, hipMemcpyHostToDevice));
      return FEM_OK;
    }
    const int k_rows = res ? RES_K : maxrows <= nt ? 1 : maxrows <= 2 * nt ? 2 : maxrows <= 4 * nt ? 4 : 8;
    const void* kp = cg_kernel_ptr(k_rows, nt, res, false);
    if (!kp) return fail(FEM_ERR_ARG, "unsupported CG rows per thread");
    int nb = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kp, nt, 0));
    // One persistent launch.  With more chunks than workgroups fit at once
    // the queue kernel runs with `cap` workgroups (as many of its own as fit)
    // that take chunks from a queue: whole instances, the most expensive
    // first (they bound the solve's length), each instance's chunks
    // consecutive.  An instance's chunks must all be co-resident.
    int cap = nb * ncu;
    const bool queue = n_chunks > cap;
    if (queue) {
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, cg_kernel_ptr(k_rows, nt, res, true), nt, 0));
      cap = std::min(cap, nb * ncu);
    }
    for (int k = 0; k < ninst; ++k) {
      const int nk = inst_chunks[k + 1] - inst_chunks[k];
      if (nk > cap) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "CG needs %d co-resident workgroups for one instance, device holds %d", nk,
                      cap);
        return fail(FEM_ERR_UNSUPPORTED, buf);
      }
    }
    const std::vector<int32_t> groups = {0, n_chunks};
    std::vector<int32_t> order;
    if (queue) {
      std::vector<int32_t> inst_order(ninst);
      for (int k = 0; k < ninst; ++k) inst_order[k] = k;
      std::stable_sort(inst_order.begin(), inst_order.end(), [&](int u, int v) { return cost[u] > cost[v]; });
      order.reserve(n_chunks);
      for (const int k : inst_order)
        for (int c = inst_chunks[k]; c < inst_chunks[k + 1]; ++c) order.push_back(c);
    }
    int *n_rows = nullptr, *n_inst = nullptr, *n_ichunks = nullptr, *n_tiles = nullptr, *n_trows = nullptr,
        *n_tcptr = nullptr, *n_tcols = nullptr, *n_rthr = nullptr, *n_rrow = nullptr;
    uint64_t* n_rmask = nullptr;
    uint16_t* n_lidx = nullptr;
    int* n_order = nullptr;
    double* n_part = nullptr;
    unsigned* n_flags = nullptr;
    int rc;
    if (res && ((rc = upload(&n_rthr, res_thr.data(), res_thr.size())) ||
                (rc = upload(&n_rmask, res_mask.data(), res_mask.size())) ||
                (rc = upload(&n_rrow, res_row.data(), res_row.size()))))
      return rc;
    // Column lists and local column numbers with room for an edit of the
    // pattern that keeps the chunk rows (written in place then): every block
    // the bound storage can still take adds at most one column to a chunk.
    const size_t room = (size_t)std::max<int64_t>(0, b_cap - (int64_t)koff[P]);
    const size_t tcols_cap = tile_cols.size() + tile_cols.size() / 16 + 64 + room;
    const size_t lidx_cap = lidx.size() + room;
    if ((rc = upload(&n_tcptr, tile_cptr.data(), tile_cptr.size())) ||
        (rc = upload_cap(&n_tcols, tile_cols.data(), tile_cols.size(), tcols_cap)) ||
        (rc = upload_cap(&n_lidx, lidx.data(), lidx.size(), lidx_cap)) ||
        (rc = upload(&n_rows, chunk_rows.data(), chunk_rows.size())) ||
        (rc = upload(&n_inst, chunk_inst.data(), chunk_inst.size())) ||
        (rc = upload(&n_ichunks, inst_chunks.data(), inst_chunks.size())) ||
        (rc = upload(&n_tiles, chunk_tiles.data(), chunk_tiles.size())) ||
        (rc = upload(&n_trows, tile_rows.data(), tile_rows.size())) || (rc = alloc(&n_part, 2 * (size_t)NPART * n_chunks)) ||
        (rc = alloc(&n_flags, (size_t)n_chunks + 1)) || (!order.empty() && (rc = upload(&n_order, order.data(), order.size()))))
      return rc;
    for (void* p : {(void*)d_chunk_rows, (void*)d_chunk_inst, (void*)d_inst_chunks, (void*)d_part,
                    (void*)d_chunk_tiles, (void*)d_tile_rows, (void*)d_tile_cptr, (void*)d_tile_cols,
                    (void*)d_lidx, (void*)d_res_thr, (void*)d_res_mask, (void*)d_res_row, (void*)d_cgflags,
                    (void*)d_chunk_order})
      drop(p);
    d_chunk_rows = n_rows;
    d_chunk_inst = n_inst;
    d_inst_chunks = n_ichunks;
    d_chunk_tiles = n_tiles;
    d_tile_rows = n_trows;
    d_tile_cptr = n_tcptr;
    d_tile_cols = n_tcols;
    d_lidx = n_lidx;
    d_res_thr = n_rthr;
    d_res_mask = n_rmask;
    d_res_row = n_rrow;
    cg_res = res;
    d_part = n_part;
    d_cgflags = n_flags;
    d_chunk_order = n_order;
    cg_queue = !order.empty();
    cg_grid = cg_queue ? cap : n_chunks;
    nchunks = n_chunks;
    h_chunk_rows = chunk_rows;
    cap_tile_cols = tcols_cap;
    cap_lidx = lidx_cap;
    cg_groups = groups;
    cg_k = k_rows;
    cg_nt = nt;
    max_resident = nb * ncu;
    return FEM_OK;
  }

  ~fem_system() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace {
// Device address of `scratch` on the current device, taken once per device
// (hipGetSymbolAddress) and kept for the life of the process.
int static_scratch_ptr(unsigned char** out) {
  constexpr int MAX_DEVICES = 64;
  static std::mutex mu;
  static unsigned char* addr[MAX_DEVICES] = {};
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(FEM_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(FEM_ERR_HIP, "hipGetSymbolAddress(scratch) returned no address");
    addr[dev] = static_cast<unsigned char*>(p);
  }
  *out = addr[dev];
  return FEM_OK;
}

int pick_lanes(double avg) {
  int g = 4;
  while (g < 64 && g < avg) g *= 2;
  return g;
}
}  // namespace

extern "C" {

const char* fem_last_error(void) { return g_error.c_str(); }

int fem_create(const fem_model_desc* d, fem_system** out) {
  if (!d || !out) return fail(FEM_ERR_ARG, "null argument");
  *out = nullptr;
  if (d->particle_count < 0 || d->tet_count < 0 || d->hex_count < 0 || d->fixed_count < 0 || d->spring_count < 0 ||
      d->tri_count < 0)
    return fail(FEM_ERR_ARG, "negative count");
  auto* s = new fem_system();
  auto cleanup = [&](int rc) {
    delete s;
    return rc;
  };
  s->P = d->particle_count;
  s->T = d->tet_count;
  s->H = d->hex_count;
  s->S = d->spring_count;
  s->R = d->tri_count;
  s->nfixed = d->fixed_count;
  const int P = s->P;
  if (d->instance_count > 1) {
    s->ninst = d->instance_count;
    s->inst_off.assign(d->instance_offsets, d->instance_offsets + s->ninst + 1);
    if (s->inst_off[0] != 0 || s->inst_off[s->ninst] != P) return cleanup(fail(FEM_ERR_ARG, "instance_offsets"));
    for (int k = 0; k < s->ninst; ++k)
      if (s->inst_off[k + 1] <= s->inst_off[k]) return cleanup(fail(FEM_ERR_ARG, "empty or unsorted instance"));
  } else {
    s->ninst = 1;
    s->inst_off = {0, P};
  }
  std::string err;
  int rc = build_pattern(P, s->T, d->tet_indices, d->tet_materials, s->H, d->hex_indices, s->S, d->spring_indices,
                         s->R, d->tri_indices, s->pt, err);
  if (rc) return cleanup(fail(rc, err));
  s->h_hex_idx.assign(d->hex_indices, d->hex_indices + 8 * (size_t)s->H);
  s->h_sp_idx.assign(d->spring_indices, d->spring_indices + 2 * (size_t)s->S);
  s->h_sp_ke.assign(d->spring_stiffness, d->spring_stiffness + s->S);
  s->h_sp_kd.assign(d->spring_damping, d->spring_damping + s->S);
  s->h_tri_idx.assign(d->tri_indices, d->tri_indices + 3 * (size_t)s->R);
  s->h_tri_pose.assign(d->tri_poses, d->tri_poses + 4 * (size_t)s->R);
  s->h_tri_mat.assign(d->tri_materials, d->tri_materials + 5 * (size_t)s->R);
  if (s->S > 0 || s->R > 0) {
    // Per row the triangles' (k_mu + k_lambda) * rest area; degenerate
    // triangles contribute nothing (build_system_matrix_tri_kernel
    // kernels.py:994-1008).
    std::vector<float> row_tri(s->R > 0 ? P : 0, 0.f);
    for (int t = 0; t < s->R; ++t) {
      const float* D = d->tri_poses + 4 * (size_t)t;
      const float inv_area = (D[0] * D[3] - D[1] * D[2]) * 2.f;
      if (!(inv_area > 0.f)) continue;
      const float area = 1.f / inv_area;
      const float c = d->tri_materials[5 * (size_t)t] * area + d->tri_materials[5 * (size_t)t + 1] * area;
      for (int a = 0; a < 3; ++a) row_tri[d->tri_indices[3 * (size_t)t + a]] += c;
    }
    if ((rc = s->upload(&s->d_sp_idx, d->spring_indices, 2 * (size_t)s->S)) ||
        (rc = s->upload(&s->d_sp_rest, d->spring_rest_length, (size_t)s->S)) ||
        (rc = s->upload(&s->d_sp_ke, d->spring_stiffness, (size_t)s->S)) ||
        (rc = s->upload(&s->d_sp_kd, d->spring_damping, (size_t)s->S)) ||
        (rc = s->upload(&s->d_tri_idx, d->tri_indices, 3 * (size_t)s->R)) ||
        (rc = s->upload(&s->d_tri_pose, d->tri_poses, 4 * (size_t)s->R)) ||
        (rc = s->upload(&s->d_tri_mat, d->tri_materials, 5 * (size_t)s->R)) ||
        (rc = s->upload(&s->d_row_tri, row_tri.data(), row_tri.size())))
      return cleanup(rc);
  }
  // Elements must not couple different instances.
  if (s->ninst > 1) {
    std::vector<int32_t> inst_of(P);
    for (int k = 0; k < s->ninst; ++k)
      for (int i = s->inst_off[k]; i < s->inst_off[k + 1]; ++i) inst_of[i] = k;
    for (int i = 0; i < P; ++i)
      for (int64_t k = s->pt.offsets[i]; k < s->pt.offsets[i + 1]; ++k)
        if (inst_of[s->pt.columns[k]] != inst_of[i]) return cleanup(fail(FEM_ERR_ARG, "element spans two instances"));
    if ((rc = s->upload(&s->d_row_inst, inst_of.data(), P))) return cleanup(rc);
  }
  if ((rc = s->upload_pattern(4 * (size_t)s->T, 16 * (size_t)s->T + 64 * (size_t)s->H))) return cleanup(rc);
  std::vector<unsigned char> fixed_mask(std::max(P, 1), 0);
  for (int k = 0; k < s->nfixed; ++k) {
    int i = d->fixed_indices[k];
    if (i < 0 || i >= P) return cleanup(fail(FEM_ERR_ARG, "fixed index out of range"));
    fixed_mask[i] = 1;
  }
  if ((rc = s->upload(&s->d_tet_idx, d->tet_indices, 4 * (size_t)s->T)) ||
      (rc = s->upload(&s->d_tet_pose, d->tet_poses, 9 * (size_t)s->T)) ||
      (rc = s->upload(&s->d_tet_mat, d->tet_materials, 3 * (size_t)s->T)) ||
      (rc = s->upload(&s->d_hex_idx, d->hex_indices, 8 * (size_t)s->H)) ||
      (rc = s->upload(&s->d_hex_invJ, d->hex_inv_J0, 72 * (size_t)s->H)) ||
      (rc = s->upload(&s->d_hex_detw, d->hex_det_J0_w, 8 * (size_t)s->H)) ||
      (rc = s->upload(&s->d_hex_mat, d->hex_materials, 3 * (size_t)s->H)) ||
      (s->H > 0 && ((rc = s->alloc(&s->d_hex_gp_t, HEX_GT * 8 * (size_t)s->H)) ||
                    (rc = s->alloc(&s->d_hex_gp_f, HEX_GF * 8 * (size_t)s->H)))) ||
      (rc = s->upload(&s->d_mass, d->particle_mass, P)) ||
      (rc = s->upload(&s->d_inv_mass, d->particle_inv_mass, P)) ||
      (rc = s->upload(&s->d_radius, d->particle_radius, P)) ||
      (rc = s->upload(&s->d_flags, d->particle_flags, P)) ||
      (rc = s->upload(&s->d_anchor, d->particle_q, 3 * (size_t)P)) ||
      (rc = s->upload(&s->d_row_fixed, fixed_mask.data(), fixed_mask.size())) ||
      (rc = s->upload(&s->d_fixed, d->fixed_indices, s->nfixed)) ||
      (rc = s->alloc(&s->d_xw, 2 * (size_t)XW * P)) ||
      (rc = s->alloc(&s->d_rhs_max, (size_t)RHS_MAX_SLOTS * s->ninst)) || 
      (rc = s->alloc(&s->d_stats, 4 * (size_t)s->ninst)) || (rc = s->alloc(&s->d_err, 1)) ||
      (rc = s->alloc(&s->d_norms, 2)) || (rc = s->alloc(&s->d_dstats, 4)))
    return cleanup(rc);
  for (int i = 0; i < P; ++i)  // Model.particle_max_radius
    if (std::isfinite(d->particle_radius[i])) s->max_radius = std::max(s->max_radius, d->particle_radius[i]);
  double avg_inc = P ? (double)(s->pt.tinc_ptr[P] + s->pt.hinc_ptr[P]) / P : 1.0;
  double avg_nnz = P ? (double)s->pt.offsets[P] / P : 1.0;
  (void)avg_nnz;
  s->elem_g = std::min(8, pick_lanes(avg_inc / 4.0));
  // Assembly row groups: the incidence and LDS budgets bind before 64 rows.
  if (P > 0 && (rc = s->plan_assemble(64))) return cleanup(rc);
  s->prm.min_stretch = 0.05f;
  s->prm.max_stretch = 20.f;
  s->prm.tolerance = 1e-2f;
  s->prm.maxiter = 50;
  s->prm.early_exit = 1;
  s->prm.gravity[2] = -9.81f;
  if (P > 0 && (rc = s->plan_cg())) return cleanup(rc);
  if ((rc = static_scratch_ptr(&s->static_scratch))) return cleanup(rc);
  *out = s;
  return FEM_OK;
}

int fem_destroy(fem_system* s) {
  delete s;
  return FEM_OK;
}

int fem_memory_stats(const fem_system* s, int64_t* allocations, int64_t* bytes) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  int64_t total = 0;
  for (const void* p : s->allocs) {
    const auto it = s->alloc_bytes.find(p);
    if (it != s->alloc_bytes.end()) total += (int64_t)it->second;
  }
  if (allocations) *allocations = (int64_t)s->allocs.size();
  if (bytes) *bytes = total;
  return FEM_OK;
}

int fem_set_particle_contact(fem_system* s, float ke, float kd, float kf, float mu, float cohesion) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (!std::isfinite(ke) || !std::isfinite(kd) || !std::isfinite(kf) || !std::isfinite(mu) || !std::isfinite(cohesion) ||
      cohesion < 0.f)
    return fail(FEM_ERR_ARG, "particle contact coefficients must be finite and the cohesion distance nonnegative");
  // The reference skips the term for ke <= 0, fewer than two particles or no
  // positive radius (solver.py:1481-1486).
  if (ke <= 0.f || s->P <= 1 || !(s->max_radius > 0.f)) {
    s->contact_on = false;
    return FEM_OK;
  }
  ContactGrid& g = s->contact;
  if (!g.force) {
    int bits = 6;
    while (bits < 20 && (1 << bits) < 2 * s->P) ++bits;
    const size_t tb = contact_temp_bytes(bits);
    int rc;
    unsigned char* temp = nullptr;
    if ((rc = s->alloc(&g.keys, (size_t)s->P)) || (rc = s->alloc(&g.ids, (size_t)s->P)) ||
        (rc = s->alloc(&temp, tb)) || (rc = s->alloc(&g.force, 3 * (size_t)s->P)))
      return rc;
    g.temp = temp;
    g.P = s->P;
    g.bits = bits;
  }
  // Cell width: the search radius 2 max_radius + cohesion (solver.py:1487).
  g.max_radius = s->max_radius;
  g.cell = 2.f * s->max_radius + cohesion;
  s->cparams = ContactParams{ke, kd, kf, mu, cohesion};
  s->contact_on = true;
  return FEM_OK;
}

int fem_set_soft_contacts(fem_system* s, const int* count, const int* particle, const float* body_pos,
                          const float* body_vel, const float* normal, int max_contacts, float ke, float kd, float kf,
                          float mu) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (!count || max_contacts <= 0) {  // no contact list: the term is zero
    s->soft_on = false;
    return FEM_OK;
  }
  if (!particle || !body_pos || !body_vel || !normal) return fail(FEM_ERR_ARG, "null soft-contact array");
  if (!std::isfinite(ke) || !std::isfinite(kd) || !std::isfinite(kf) || !std::isfinite(mu))
    return fail(FEM_ERR_ARG, "soft contact coefficients must be finite");
  if (s->P > 1024 * 1024) return fail(FEM_ERR_UNSUPPORTED, "soft contacts support at most 1048576 particles");
  if (s->P == 0) {
    s->soft_on = false;
    return FEM_OK;
  }
  SoftContacts& c = s->soft;
  const size_t tb = soft_contact_temp_bytes(s->P, max_contacts);
  if (!c.force || tb > s->soft_temp_cap) {  // grows with the largest contact list seen
    unsigned char* temp = nullptr;
    int rc;
    if ((rc = s->alloc(&temp, tb))) return rc;
    if (!c.force && (rc = s->alloc(&c.force, 3 * (size_t)s->P))) return rc;
    if (c.temp) s->drop(c.temp);
    c.temp = temp;
    s->soft_temp_cap = tb;
  }
  c.P = s->P;
  c.max_contacts = max_contacts;
  c.count = count;
  c.particle = particle;
  c.body_pos = body_pos;
  c.body_vel = body_vel;
  c.normal = normal;
  c.ke = ke;
  c.kd = kd;
  c.kf = kf;
  c.mu = mu;
  s->soft_on = true;
  return FEM_OK;
}

int fem_set_spring_torque(fem_system* s, const float* rest_direction, float stiffness, float damping) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (!std::isfinite(stiffness) || !std::isfinite(damping))
    return fail(FEM_ERR_ARG, "torque stiffness and damping must be finite");
  // The reference applies the torque only for stiffness > 0 (solver.py:1442).
  if (!(stiffness > 0.f) || !rest_direction || s->S == 0) {
    s->torque_on = false;
    return FEM_OK;
  }
  for (int k = 0; k < 3 * s->S; ++k)
    if (!std::isfinite(rest_direction[k])) return fail(FEM_ERR_ARG, "spring_rest_direction must be finite");
  int rc;
  if (!s->d_sp_dir) {
    if ((rc = s->upload(&s->d_sp_dir, rest_direction, 3 * (size_t)s->S))) return rc;
  } else if ((rc = s->write(s->d_sp_dir, rest_direction, 3 * (size_t)s->S))) {
    return rc;
  }
  s->sp_kt = stiffness;
  s->sp_kw = damping;
  s->torque_on = true;
  return FEM_OK;
}

int fem_set_solver(fem_system* s, int32_t method, int32_t preconditioner, int32_t restart) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (method < KR_CG || method > KR_GMRES) return fail(FEM_ERR_ARG, "unknown Krylov method");
  if (preconditioner < KP_BLOCK || preconditioner > KP_ID) return fail(FEM_ERR_ARG, "unknown preconditioner");
  if (restart < 1 || restart > KR_MAX_RESTART) return fail(FEM_ERR_ARG, "GMRES restart must be in [1, 64]");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if ((method != KR_CG || preconditioner != KP_BLOCK) && (s->ninst > 1 || s->topo_count || s->cs_scale))
    return fail(FEM_ERR_UNSUPPORTED,
                "batched instances, reserved pools and component scaling run the block-Jacobi PCG only");
  int rc;
  if (preconditioner != KP_BLOCK && !s->d_pjac && (rc = s->alloc(&s->d_pjac, 9 * (size_t)std::max(s->P, 1))))
    return rc;
  if (method != KR_CG) {
    const size_t wf = krylov_work_floats(s->P, method, restart);
    if (!s->holds(s->d_kr_work, sizeof(float) * wf)) {
      float* w = nullptr;
      if ((rc = s->alloc(&w, std::max<size_t>(wf, 1)))) return rc;  // zero-filled
      s->drop(s->d_kr_work);
      s->d_kr_work = w;
    }
    const size_t pd = krylov_part_doubles(s->P, restart);
    if (!s->holds(s->d_kr_part, sizeof(double) * pd)) {
      double* p = nullptr;
      if ((rc = s->alloc(&p, pd))) return rc;
      s->drop(s->d_kr_part);
      s->d_kr_part = p;
    }
    if (!s->d_kr_sc && (rc = s->alloc(&s->d_kr_sc, (size_t)KR_NSC))) return rc;
  }
  s->kr_method = method;
  s->kr_precond = preconditioner;
  s->kr_restart = restart;
  if (preconditioner != KP_BLOCK && s->P > 0) {
    const hipError_t e = krylov_point_jacobi(s->P, s->d_offsets, s->d_diag, s->b_vals, preconditioner, s->d_pjac,
                                             nullptr);
    if (e != hipSuccess) return fail(FEM_ERR_HIP, std::string("point-Jacobi preconditioner: ") + hipGetErrorString(e));
    HIP_TRY(hipDeviceSynchronize());
  }
  return FEM_OK;
}

int fem_pattern_size(const fem_system* s, int64_t* nnz, int32_t* max_row_blocks) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (nnz) *nnz = s->pt.offsets[s->P];
  if (max_row_blocks) *max_row_blocks = s->pt.max_row_nnz;
  return FEM_OK;
}

int fem_bind(fem_system* s, int32_t* offsets, int32_t* columns, float* values, float* inv_diag, float* increment,
             int64_t block_capacity) {
  if (!s || !offsets || !columns || !values || !inv_diag || !increment) return fail(FEM_ERR_ARG, "null buffer");
  const int64_t nnz = s->pt.offsets[s->P];
  if (block_capacity == 0) block_capacity = nnz;
  if (block_capacity < nnz) return fail(FEM_ERR_ARG, "block_capacity is smaller than the pattern");
  if (s->cs_scale && block_capacity != s->b_cap)
    return fail(FEM_ERR_STATE, "component scaling storage is bound to the previous block capacity");
  const bool grow = block_capacity > s->b_cap && block_capacity > nnz && !s->topo_count && s->P > 0 && s->ninst == 1;
  s->b_cap = block_capacity;
  // Storage that leaves room for more blocks: the pattern arrays, the CG
  // plan storage and the assembly launch are sized for it now (before any
  // launch is captured), so that a later topology edit without a reserved
  // pool that fits is written in place.
  if (grow) {
    int rc;
    if ((rc = s->upload_pattern(4 * (size_t)s->T, 16 * (size_t)s->T + 64 * (size_t)s->H, (size_t)block_capacity)) ||
        (rc = s->plan_assemble(s->asm_rb)) || (rc = s->plan_cg()))
      return rc;
  }
  s->b_offsets = offsets;
  s->b_cols = columns;
  s->b_vals = values;
  s->b_minv = inv_diag;
  s->b_dv = increment;
  std::vector<int32_t> off32(s->P + 1);
  for (int i = 0; i <= s->P; ++i) off32[i] = (int32_t)s->pt.offsets[i];
  HIP_TRY(hipMemcpy(offsets, off32.data(), off32.size() * 4, hipMemcpyHostToDevice));
  if (!s->pt.columns.empty())
    HIP_TRY(hipMemcpy(columns, s->pt.columns.data(), s->pt.columns.size() * 4, hipMemcpyHostToDevice));
  return FEM_OK;
}

int fem_bind_gravity(fem_system* s, const float* gravity) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  s->d_gravity = gravity;
  return FEM_OK;
}

int fem_set_params(fem_system* s, const fem_params* p) {
  if (!s || !p) return fail(FEM_ERR_ARG, "null argument");
  if (!(p->tolerance > 0.f) || !std::isfinite(p->tolerance)) return fail(FEM_ERR_ARG, "tolerance must be finite and positive");
  if (p->maxiter < 0) return fail(FEM_ERR_ARG, "maxiter must be nonnegative");
  s->prm = *p;
  return FEM_OK;
}

int fem_set_launch(fem_system* s, int32_t elem_lanes, int32_t assemble_rows, int32_t cg_workgroups) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (s->topo_count) return fail(FEM_ERR_STATE, "launch shapes are fixed once tetrahedral updates are prepared");
  auto ok = [](int g) { return g == 4 || g == 8 || g == 16 || g == 32 || g == 64; };
  if (elem_lanes && !ok(elem_lanes)) return fail(FEM_ERR_ARG, "elem_lanes must be a power of two in [4, 64]");
  if (assemble_rows && !ok(assemble_rows)) return fail(FEM_ERR_ARG, "assemble_rows must be a power of two in [4, 64]");
  if (cg_workgroups < 0) return fail(FEM_ERR_ARG, "cg_workgroups must be >= 0");
  const int old_request = s->cg_wg_request;
  s->cg_wg_request = cg_workgroups;
  if (s->P > 0) {
    int rc = s->plan_cg();
    if (rc) {
      s->cg_wg_request = old_request;
      return rc;
    }
  }
  if (elem_lanes) s->elem_g = elem_lanes;
  if (assemble_rows && s->P > 0) {
    if (int rc = s->plan_assemble(assemble_rows)) return rc;
  } else if (assemble_rows) {
    s->asm_rb = assemble_rows;
  }
  return FEM_OK;
}

int fem_get_launch(const fem_system* s, int32_t* elem_lanes, int32_t* assemble_rows, int32_t* cg_workgroups,
                   int32_t* cg_rows_per_thread, int32_t* cg_launch_workgroups, int32_t* cg_variant) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (cg_variant) *cg_variant = s->cg_res ? FEM_CG_RESIDENT : FEM_CG_STREAMED;
  if (cg_launch_workgroups) *cg_launch_workgroups = s->cg_grid;
  if (elem_lanes) *elem_lanes = s->elem_g;
  if (assemble_rows) *assemble_rows = s->asm_rb;
  if (cg_workgroups) *cg_workgroups = s->nchunks;
  if (cg_rows_per_thread) *cg_rows_per_thread = s->cg_k;
  return FEM_OK;
}

int fem_set_cg_variant(fem_system* s, int32_t variant) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (s->topo_count) return fail(FEM_ERR_STATE, "launch shapes are fixed once tetrahedral updates are prepared");
  if (variant < FEM_CG_AUTO || variant > FEM_CG_RESIDENT) return fail(FEM_ERR_ARG, "unknown CG variant");
  const int old = s->cg_variant_request;
  s->cg_variant_request = variant;
  if (s->P > 0) {
    const int rc = s->plan_cg();
    if (rc) {  // a failed plan leaves the previous one in place
      s->cg_variant_request = old;
      return rc;
    }
  }
  return FEM_OK;
}

}  // extern "C"

namespace {
// Gauss-point records of every hex at positions x (and velocities v for the
// stresses), ahead of the kernels that read them on the same stream.
template <bool FORCE>
int hex_gauss(const fem_system* s, const RowArgs& ra, const float* x, const float* v, hipStream_t st) {
  if (s->H == 0) return FEM_OK;
  const int n = 8 * s->H;
  hipLaunchKernelGGL(k_hex_gauss<FORCE>, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, ra, s->H, x, v,
                     FORCE ? s->d_hex_gp_f : s->d_hex_gp_t);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

// Particle-particle contact forces at (x, v) ahead of the force kernel.
int particle_contact(const fem_system* s, const float* x, const float* v, hipStream_t st) {
  if (!s->contact_on) return FEM_OK;
  const hipError_t e = contact_forces(s->contact, s->cparams, x, v, s->d_radius, s->d_flags, st);
  if (e != hipSuccess) return fail(FEM_ERR_HIP, std::string("particle contact: ") + hipGetErrorString(e));
  return FEM_OK;
}

// Soft-contact forces of the bound contact list at (x, v) ahead of the force kernel.
int soft_contact(const fem_system* s, const float* x, const float* v, hipStream_t st) {
  if (!s->soft_on) return FEM_OK;
  const hipError_t e = soft_contact_forces(s->soft, x, v, s->d_radius, st);
  if (e != hipSuccess) return fail(FEM_ERR_HIP, std::string("soft contact: ") + hipGetErrorString(e));
  return FEM_OK;
}
}  // namespace

extern "C" {

int fem_restore_fixed(fem_system* s, float* q, float* qd, void* stream) {
  if (!s || !q) return fail(FEM_ERR_ARG, "null argument");
  if (s->nfixed == 0) return FEM_OK;
  int nb = (s->nfixed + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL(k_restore, dim3(nb), dim3(BLOCK), 0, (hipStream_t)stream, s->d_fixed, s->nfixed, s->d_anchor,
                     q, qd);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_assemble(fem_system* s, const float* q, float dt, void* stream) {
  if (!s || !q) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (s->P == 0) return FEM_OK;
  RowArgs ra = s->row_args();
  if (int rc = hex_gauss<false>(s, ra, q, nullptr, (hipStream_t)stream)) return rc;
  int rc = dispatch_g<LaunchAssemble>(s->asm_rb, s->full(), ra, s->d_asm_groups, s->asm_ngroups, q, dt, s->b_vals,
                                      s->b_minv, s->asm_lds, (hipStream_t)stream);
  if (!rc && s->kr_precond != KP_BLOCK) {
    // preconditioner(A, "diag" / "diag_abs" / "id") of the refreshed matrix (solver.py:1183-1185)
    const hipError_t e = krylov_point_jacobi(s->P, s->d_offsets, s->d_diag, s->b_vals, s->kr_precond, s->d_pjac,
                                             (hipStream_t)stream);
    if (e != hipSuccess) return fail(FEM_ERR_HIP, std::string("point-Jacobi preconditioner: ") + hipGetErrorString(e));
  }
  if (rc || !s->topo_count) return rc;
  // Positions of the current tangent (streaming.py:603, :716), compared with
  // by the next topology update.
  hipLaunchKernelGGL(k_copy_floats, dim3((3 * s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream, q,
                     s->d_lin_q, 3 * s->P);
  if (s->compact)
    hipLaunchKernelGGL(k_gather_blocks, dim3(1024), dim3(BLOCK), 0, (hipStream_t)stream, (const int*)s->d_cmap,
                       (const int*)(s->d_coffsets + s->P), (const float*)s->b_vals, s->d_cvals);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_eval_forces(fem_system* s, int32_t terms, const float* q, const float* qd, const float* act, float* out,
                    void* stream) {
  if (!s || !q || !qd || !out) return fail(FEM_ERR_ARG, "null argument");
  if (s->P == 0) return FEM_OK;
  int rc_hex = FEM_OK;
  RowArgs ra = s->row_args();
  if ((terms & FEM_TERM_HEX) && (rc_hex = hex_gauss<true>(s, ra, q, qd, (hipStream_t)stream))) return rc_hex;
  if ((terms & FEM_TERM_PARTICLE) && (rc_hex = particle_contact(s, q, qd, (hipStream_t)stream))) return rc_hex;
  if ((terms & FEM_TERM_SOFT) && (rc_hex = soft_contact(s, q, qd, (hipStream_t)stream))) return rc_hex;
  return dispatch_g<LaunchRhs>(s->elem_g, s->full(), ra, q, qd, act, (const float*)nullptr, (const float*)s->b_vals, out,
                               s->d_rhs_max, (int)terms, (int)MODE_EVAL, 1.f, (hipStream_t)stream);
}

int fem_step_rhs(fem_system* s, float* q_in, float* qd_in, float* f_out, const float* ext, const float* act, float dt,
                 void* stream) {
  if (!s || !q_in || !qd_in || !f_out) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (s->P == 0) return FEM_OK;
  hipStream_t st = (hipStream_t)stream;
  {
    int n = std::max(std::max(s->nfixed, s->ninst * RHS_MAX_SLOTS), 1);
    hipLaunchKernelGGL(k_prepare, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, s->d_fixed, s->nfixed,
                       s->d_anchor, q_in, qd_in, s->d_rhs_max, s->ninst);
    HIP_TRY(hipGetLastError());
  }
  RowArgs ra = s->row_args();
  const int terms = FEM_TERM_SPRING | FEM_TERM_TET | FEM_TERM_HEX | FEM_TERM_TRI | FEM_TERM_PARTICLE |
                    FEM_TERM_GROUND | FEM_TERM_SOFT | FEM_TERM_GRAVITY | FEM_TERM_DAMPING;
  if (int rc = hex_gauss<true>(s, ra, q_in, qd_in, st)) return rc;
  if (int rc = particle_contact(s, q_in, qd_in, st)) return rc;
  if (int rc = soft_contact(s, q_in, qd_in, st)) return rc;
  return dispatch_g<LaunchRhs>(s->elem_g, s->full(), ra, (const float*)q_in, (const float*)qd_in, act, ext,
                               (const float*)s->b_vals, f_out, s->d_rhs_max, terms, (int)MODE_STEP, dt, st);
}

// Right-hand side from a caller-provided h-scaled force sum, in place.
int fem_integrate_rhs(fem_system* s, const float* qd_in, float* f, float dt, void* stream) {
  if (!s || !qd_in || !f) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (s->P == 0) return FEM_OK;
  hipStream_t st = (hipStream_t)stream;
  {
    int n = std::max(s->ninst * RHS_MAX_SLOTS, 1);
    hipLaunchKernelGGL(k_prepare, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, (const int*)nullptr, 0,
                       (const float*)nullptr, (float*)nullptr, (float*)nullptr, s->d_rhs_max, s->ninst);
    HIP_TRY(hipGetLastError());
  }
  RowArgs ra = s->row_args();
  // No force term is evaluated: positions are not read (qd stands in for x).
  return dispatch_g<LaunchRhs>(s->elem_g, s->full(), ra, qd_in, qd_in, (const float*)nullptr, (const float*)f,
                               (const float*)s->b_vals, f, s->d_rhs_max, 0, (int)MODE_LOAD, dt, st);
}

}  // extern "C"

namespace {
CgArgs cg_args(const fem_system* s) {
  CgArgs ca;
  std::memset(&ca, 0, sizeof ca);
  ca.P = s->P;
  ca.offsets = s->compact ? s->d_coffsets : s->d_offsets;
  ca.cols = s->compact ? s->d_ccols : s->d_cols;
  ca.vals = s->compact ? s->d_cvals : s->b_vals;
  ca.minv = s->precond_blocks();
  ca.row_fixed = s->d_row_fixed;
  ca.dv = s->b_dv;
  ca.anchor = s->d_anchor;
  ca.xw = s->d_xw;
  ca.chunk_rows = s->d_chunk_rows;
  ca.chunk_tiles = s->d_chunk_tiles;
  ca.tile_rows = s->d_tile_rows;
  ca.tile_cptr = s->d_tile_cptr;
  ca.tile_cols = s->d_tile_cols;
  ca.lidx = s->d_lidx;
  ca.res_thr = reinterpret_cast<const int4*>(s->d_res_thr);
  ca.res_mask = reinterpret_cast<const unsigned long long*>(s->d_res_mask);
  ca.res_row = reinterpret_cast<const int2*>(s->d_res_row);
  ca.chunk_inst = s->d_chunk_inst;
  ca.inst_chunks = s->d_inst_chunks;
  ca.rhs_max = s->d_rhs_max;
  ca.tolerance = s->prm.tolerance;
  ca.maxiter = s->prm.maxiter;
  ca.early_exit = s->prm.early_exit;
  ca.part = s->d_part;
  ca.flags = s->d_cgflags;
  ca.stats = s->d_stats;
  ca.err = s->d_err;
  ca.spin_limit = s->spin_limit;
  ca.nchunks_total = s->nchunks;
  return ca;
}

// Flag and queue reset + the persistent launch.
int launch_cg(const fem_system* s, CgArgs ca, hipStream_t st) {
  const void* kp = cg_kernel_ptr(s->cg_k, s->cg_nt, s->cg_res, s->cg_queue);
  if (!kp) return fail(FEM_ERR_ARG, "unsupported CG launch shape");
  hipLaunchKernelGGL(k_cg_reset, dim3((s->nchunks + 1 + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, ca.flags,
                     s->nchunks + 1, s->d_err);
  HIP_TRY(hipGetLastError());
  ca.chunk_base = 0;
  ca.chunk_order = s->d_chunk_order;
  void* args[] = {&ca};
  HIP_TRY(hipLaunchKernel(kp, dim3(s->cg_grid), dim3(s->cg_nt), args, 0, st));
  return FEM_OK;
}
}  // namespace

namespace {
// Bytes of static scratch a solve of `s` needs: partial sums (2 x NPART
// doubles per chunk), then the publication flags (one word per chunk and the
// chunk queue word), then
// the exchange records (2 x XW floats per row), each part 256-byte aligned.
size_t static_scratch_layout(const fem_system* s, size_t* flags_off, size_t* xw_off) {
  auto up = [](size_t n) { return (n + 255) & ~size_t(255); };
  const size_t part = up(sizeof(double) * 2 * NPART * (size_t)s->nchunks);
  const size_t flags = up(sizeof(unsigned) * ((size_t)s->nchunks + 1));
  *flags_off = part;
  *xw_off = part + flags;
  return part + flags + sizeof(float) * 2 * XW * (size_t)s->P;
}

// The configured Krylov method besides PCG on the step's system with
// right-hand side `rhs` and warm start / increment `dv` (no state update
// unless the caller sets q_out).
KrylovArgs krylov_args(const fem_system* s, const float* rhs, float* dv) {
  KrylovArgs ka;
  std::memset(&ka, 0, sizeof ka);
  ka.P = s->P;
  ka.offsets = s->d_offsets;
  ka.cols = s->d_cols;
  ka.vals = s->b_vals;
  ka.row_fixed = s->d_row_fixed;
  ka.M = s->kr_precond == KP_ID ? nullptr : s->precond_blocks();
  ka.rhs = rhs;
  ka.rhs_max = s->d_rhs_max;
  ka.rhs_max_slots = RHS_MAX_SLOTS;
  ka.dv = dv;
  ka.anchor = s->d_anchor;
  ka.tolerance = s->prm.tolerance;
  ka.maxiter = s->prm.maxiter;
  ka.restart = std::max(1, std::min(s->kr_restart, s->prm.maxiter));
  ka.work = s->d_kr_work;
  ka.part = s->d_kr_part;
  ka.sc = s->d_kr_sc;
  ka.stats = s->d_stats;
  return ka;
}

int krylov_launch(const fem_system* s, const KrylovArgs& ka, hipStream_t st) {
  const hipError_t e = krylov_solve(ka, s->kr_method, st);
  if (e != hipSuccess) return fail(FEM_ERR_HIP, std::string("krylov solve: ") + hipGetErrorString(e));
  return FEM_OK;
}

int step_solve(fem_system* s, const float* q_in, const float* qd_in, const float* f_in, float* q_out, float* qd_out,
               float dt, void* stream, bool use_static) {
  if (!s || !q_in || !qd_in || !f_in || !q_out || !qd_out) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (s->P == 0) return FEM_OK;
  if (s->kr_method != KR_CG) {
    if (use_static) return fail(FEM_ERR_UNSUPPORTED, "the static-scratch solve runs the block-Jacobi PCG only");
    KrylovArgs ka = krylov_args(s, f_in, s->b_dv);
    ka.q_in = q_in;
    ka.qd_in = qd_in;
    ka.q_out = q_out;
    ka.qd_out = qd_out;
    ka.h = dt;
    return krylov_launch(s, ka, (hipStream_t)stream);
  }
  CgArgs ca = cg_args(s);
  if (use_static) {
    size_t flags_off = 0, xw_off = 0;
    const size_t need = static_scratch_layout(s, &flags_off, &xw_off);
    if (need > STATIC_SCRATCH_BYTES) {
      char buf[200];
      std::snprintf(buf, sizeof buf,
                    "the solve needs %zu bytes of static scratch (%d rows, %d CG chunks); the limit is %zu bytes",
                    need, s->P, s->nchunks, STATIC_SCRATCH_BYTES);
      return fail(FEM_ERR_UNSUPPORTED, buf);
    }
    if (!s->static_scratch) return fail(FEM_ERR_STATE, "static scratch address not initialised");
    ca.part = reinterpret_cast<double*>(s->static_scratch);
    ca.flags = reinterpret_cast<unsigned*>(s->static_scratch + flags_off);
    ca.xw = reinterpret_cast<float*>(s->static_scratch + xw_off);
  }
  ca.rhs = f_in;
  if (s->cs_scale) {
    // Scaled copies (ComponentScaling.prepare, solver.py:1300-1304), then the
    // maximum of the scaled right-hand side (:1306-1312).
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_scale_system, dim3((s->P + BLOCK / 64 - 1) / (BLOCK / 64)), dim3(BLOCK), 0, st,
                       (const int*)s->d_offsets, (const float*)s->b_vals, (const float*)s->b_minv, f_in,
                       (const float*)s->cs_scale, s->P, s->cs_vals, s->cs_minv, s->cs_rhs);
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(BLOCK), 0, st, (const int*)nullptr, 0, (const float*)nullptr,
                       (float*)nullptr, (float*)nullptr, s->d_rhs_max, 1);
    hipLaunchKernelGGL(k_absmax, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, (const float*)s->cs_rhs, s->P,
                       (const int*)nullptr, s->d_rhs_max);
    if (s->compact)
      hipLaunchKernelGGL(k_gather_blocks, dim3(1024), dim3(BLOCK), 0, st, (const int*)s->d_cmap,
                         (const int*)(s->d_coffsets + s->P), (const float*)s->cs_vals, s->d_cvals);
    HIP_TRY(hipGetLastError());
    ca.vals = s->compact ? s->d_cvals : s->cs_vals;
    ca.minv = s->cs_minv;
    ca.rhs = s->cs_rhs;
  }
  ca.q_in = q_in;
  ca.qd_in = qd_in;
  ca.q_out = q_out;
  ca.qd_out = qd_out;
  ca.h = dt;
  return launch_cg(s, ca, (hipStream_t)stream);
}
}  // namespace

extern "C" {

int fem_step_solve(fem_system* s, const float* q_in, const float* qd_in, const float* f_in, float* q_out,
                   float* qd_out, float dt, void* stream) {
  return step_solve(s, q_in, qd_in, f_in, q_out, qd_out, dt, stream, false);
}

int fem_step_solve_static(fem_system* s, const float* q_in, const float* qd_in, const float* f_in, float* q_out,
                          float* qd_out, float dt, void* stream) {
  return step_solve(s, q_in, qd_in, f_in, q_out, qd_out, dt, stream, true);
}

int fem_solve(fem_system* s, const float* b, float* x, float tolerance, int32_t maxiter, int32_t kind,
              void* stream) {
  if (!s || !b || !x) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (!(tolerance > 0.f) || !std::isfinite(tolerance)) return fail(FEM_ERR_ARG, "tolerance must be finite and positive");
  if (maxiter < 0) return fail(FEM_ERR_ARG, "maxiter must be >= 0");
  if (kind != 0 && kind != 1) return fail(FEM_ERR_ARG, "kind must be 0 (forward) or 1 (adjoint)");
  if (s->P == 0) return FEM_OK;
  hipStream_t st = (hipStream_t)stream;
  const int nz = std::max(s->ninst * RHS_MAX_SLOTS, 1);
  hipLaunchKernelGGL(k_prepare, dim3((nz + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, (const int*)nullptr, 0,
                     (const float*)nullptr, (float*)nullptr, (float*)nullptr, s->d_rhs_max, s->ninst);
  hipLaunchKernelGGL(k_absmax, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, b, s->P,
                     s->ninst > 1 ? (const int*)s->d_row_inst : (const int*)nullptr, s->d_rhs_max);
  HIP_TRY(hipGetLastError());
  CgArgs ca = cg_args(s);
  ca.rhs = b;
  ca.x_out = x;
  ca.tolerance = tolerance;
  ca.maxiter = maxiter;
  ca.early_exit = 1;
  int rc = launch_cg(s, ca, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_doubles, dim3(1), dim3(64), 0, st, s->d_norms, 2);
  hipLaunchKernelGGL(k_resnorm, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, (const int*)s->d_offsets,
                     (const int*)s->d_cols, (const float*)s->b_vals, (const unsigned char*)s->d_row_fixed,
                     (const float*)x, b, s->P, s->d_norms);
  hipLaunchKernelGGL(k_record, dim3(1), dim3(64), 0, st, (const double*)s->d_norms, (int)kind, s->d_dstats);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_apply_operator(fem_system* s, const float* x, float* y, void* stream) {
  if (!s || !x || !y) return fail(FEM_ERR_ARG, "null argument");
  if (x == y) return fail(FEM_ERR_ARG, "operator action requires distinct input and output storage");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (s->P == 0) return FEM_OK;
  RowArgs ra = s->row_args();
  return dispatch_g<LaunchApplyOp>(s->elem_g, ra, (const float*)s->b_vals, x, y, (hipStream_t)stream);
}

int fem_apply_preconditioner(fem_system* s, const float* x, float* y, void* stream) {
  if (!s || !x || !y) return fail(FEM_ERR_ARG, "null argument");
  if (x == y) return fail(FEM_ERR_ARG, "preconditioner action requires distinct input and output storage");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (s->P == 0) return FEM_OK;
  hipLaunchKernelGGL(k_apply_minv, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream,
                     s->precond_blocks(), x, y, s->P);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_set_barrier_spin_limit(fem_system* s, int64_t spins) {
  if (!s) return fail(FEM_ERR_ARG, "null argument");
  if (spins < 0) return fail(FEM_ERR_ARG, "the spin limit must be nonnegative");
  s->spin_limit = spins == 0 ? BARRIER_SPIN_LIMIT : (long)spins;
  return FEM_OK;
}

int fem_block_jacobi_mv(fem_system* s, const float* x, const float* y, float* z, float alpha, float beta,
                        void* stream) {
  if (!s || !x || !z || (beta != 0.f && !y)) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_minv) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (s->P == 0) return FEM_OK;
  hipLaunchKernelGGL(k_block_jacobi_mv, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream,
                     s->precond_blocks(), x, y, z, alpha, beta, s->P);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_diagonal_blocks(fem_system* s, float* out, void* stream) {
  if (!s || !out) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (s->P == 0) return FEM_OK;
  hipLaunchKernelGGL(k_diag_blocks, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream,
                     s->row_args(), (const float*)s->b_vals, out);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_add_load(fem_system* s, const float* base, const float* ext, float dt, float* out, void* stream) {
  if (!s || !base || !out) return fail(FEM_ERR_ARG, "null argument");
  if (ext && ext == out) return fail(FEM_ERR_ARG, "the output must not alias the external-force input");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (s->P == 0) return FEM_OK;
  hipLaunchKernelGGL(k_add_load, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream, base, ext,
                     dt, (const unsigned char*)s->d_row_fixed, out, s->P);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_solve_increment(fem_system* s, const float* b, float* dv, void* stream) {
  if (!s || !b || !dv) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (s->P == 0) return FEM_OK;
  hipStream_t st = (hipStream_t)stream;
  const int nz = std::max(s->ninst * RHS_MAX_SLOTS, 1);
  hipLaunchKernelGGL(k_prepare, dim3((nz + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, (const int*)nullptr, 0,
                     (const float*)nullptr, (float*)nullptr, (float*)nullptr, s->d_rhs_max, s->ninst);
  hipLaunchKernelGGL(k_absmax, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, b, s->P,
                     s->ninst > 1 ? (const int*)s->d_row_inst : (const int*)nullptr, s->d_rhs_max);
  HIP_TRY(hipGetLastError());
  if (s->kr_method != KR_CG) return krylov_launch(s, krylov_args(s, b, dv), st);
  CgArgs ca = cg_args(s);
  ca.rhs = b;
  ca.dv = dv;
  return launch_cg(s, ca, st);
}

int fem_reconstruct(fem_system* s, const float* dv, const float* q_in, const float* qd_in, float* q_out,
                    float* qd_out, float dt, void* stream) {
  if (!s || !dv || !q_in || !qd_in || !q_out || !qd_out) return fail(FEM_ERR_ARG, "null argument");
  if (q_out == q_in || qd_out == qd_in) return fail(FEM_ERR_ARG, "reconstruction must not overwrite its input state");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (s->P == 0) return FEM_OK;
  hipLaunchKernelGGL(k_reconstruct, dim3((s->P + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, (hipStream_t)stream, dv,
                     q_in, qd_in, (const unsigned char*)s->d_row_fixed, (const float*)s->d_anchor, dt, q_out, qd_out,
                     s->P);
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

int fem_vjp(fem_system* s, const float* q, const float* qd, const float* act, const float* z, const float* dv,
            const float* q_asm, float dt, float* q_bar, float* qd_bar, float* act_bar, float* mat_bar,
            void* stream) {
  if (!s || !q || !qd || !z || !q_bar || !qd_bar) return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (mat_bar && (!dv || !q_asm)) return fail(FEM_ERR_ARG, "material gradients need dv and the assembly positions");
  if (s->P == 0) return FEM_OK;
  hipStream_t st = (hipStream_t)stream;
  RowArgs ra = s->row_args();
  int rc = dispatch_g<LaunchVjpRows>(s->elem_g, s->full(), ra, q, qd, act, z, (const float*)s->b_vals, dt, q_bar, qd_bar, st);
  if (rc) return rc;
  if (s->T > 0 && (act_bar || mat_bar)) {
    hipLaunchKernelGGL(k_vjp_tets, dim3((s->T + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, ra, s->T, q, qd, act, z, dv,
                       q_asm, dt, act_bar, mat_bar);
    HIP_TRY(hipGetLastError());
  }
  return FEM_OK;
}

int fem_set_tet_materials(fem_system* s, const float* mat, void* stream) {
  if (!s || !mat) return fail(FEM_ERR_ARG, "null argument");
  if (s->T == 0) return FEM_OK;
  HIP_TRY(hipMemcpyAsync(s->d_tet_mat, mat, 3 * (size_t)s->T * sizeof(float), hipMemcpyDeviceToDevice,
                         (hipStream_t)stream));
  return FEM_OK;
}

int fem_diff_stats(fem_system* s, void* stream, double* out, int32_t reset) {
  if (!s || !out) return fail(FEM_ERR_ARG, "null argument");
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  HIP_TRY(hipMemcpy(out, s->d_dstats, 4 * sizeof(double), hipMemcpyDeviceToHost));
  if (reset) {
    hipLaunchKernelGGL(k_zero_doubles, dim3(1), dim3(64), 0, (hipStream_t)stream, s->d_dstats, 4);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  }
  return FEM_OK;
}

int fem_step(fem_system* s, float* q_in, float* qd_in, float* f_out, const float* ext, const float* act, float* q_out,
             float* qd_out, float dt, void* stream) {
  if (!q_out || !qd_out) return fail(FEM_ERR_ARG, "null argument");
  int rc = fem_step_rhs(s, q_in, qd_in, f_out, ext, act, dt, stream);
  if (rc) return rc;
  return fem_step_solve(s, q_in, qd_in, f_out, q_out, qd_out, dt, stream);
}

int fem_solve_stats(fem_system* s, void* stream, int32_t* iterations, float* residual_sq, float* atol_sq,
                    int32_t* error_flags) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  std::vector<int32_t> st(4 * (size_t)s->ninst);
  int err = 0;
  HIP_TRY(hipMemcpy(st.data(), s->d_stats, st.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(&err, s->d_err, 4, hipMemcpyDeviceToHost));
  for (int k = 0; k < s->ninst; ++k) {
    if (iterations) iterations[k] = st[4 * k];
    if (residual_sq) std::memcpy(&residual_sq[k], &st[4 * k + 1], 4);
    if (atol_sq) std::memcpy(&atol_sq[k], &st[4 * k + 2], 4);
    if (error_flags) error_flags[k] = err;
  }
  return FEM_OK;
}


// ---------------------------------------------------------------------------
// Topology updates (solver.py:695-720, :870-935; streaming.py).
// ---------------------------------------------------------------------------
int fem_prepare_topology(fem_system* s, int32_t original_node_count) {
  if (!s) return fail(FEM_ERR_ARG, "null system");
  if (s->topo_count) return fail(FEM_ERR_STATE, "Tetrahedral updates are already prepared");
  if (s->cs_scale) return fail(FEM_ERR_STATE, "Prepare tetrahedral updates before component scaling");
  if (s->nfixed) return fail(FEM_ERR_ARG, "Streaming topology is not supported with fixed_particle_indices");
  if (s->kr_precond != KP_BLOCK) return fail(FEM_ERR_ARG, "Persistent topology updates require the block_diag preconditioner");
  if (s->kr_method != KR_CG) return fail(FEM_ERR_UNSUPPORTED, "Persistent topology updates run the block-Jacobi PCG only");
  if (s->T < 1 || s->H || s->S || s->R) return fail(FEM_ERR_ARG, "Streaming assembly requires tetrahedral elements only");
  if (s->ninst > 1) return fail(FEM_ERR_UNSUPPORTED, "topology updates of a batched system are not supported");
  const int T = s->T, P = s->P;
  HIP_TRY(hipDeviceSynchronize());
  std::vector<int32_t> nodes(4 * (size_t)T);
  std::vector<float> poses(9 * (size_t)T), mats(3 * (size_t)T);
  HIP_TRY(hipMemcpy(nodes.data(), s->d_tet_idx, nodes.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(poses.data(), s->d_tet_pose, poses.size() * 4, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(mats.data(), s->d_tet_mat, mats.size() * 4, hipMemcpyDeviceToHost));
  Pattern np;
  std::string err;
  if (int rc = build_reserved_pattern(P, original_node_count, T, nodes.data(), mats.data(), np, err))
    return fail(rc == 1 ? FEM_ERR_ARG : FEM_ERR_UNSUPPORTED, err);
  s->pt = std::move(np);
  int rc;
  if ((rc = s->upload_pattern(4 * (size_t)T, 16 * (size_t)T))) return rc;
  if (!s->d_lin_q && (rc = s->alloc(&s->d_lin_q, 3 * (size_t)P))) return rc;
  s->old_nodes = nodes;
  s->old_poses = poses;
  s->old_mats = mats;
  s->old_active.assign(T, 0);
  for (int e = 0; e < T; ++e)
    s->old_active[e] = !(mats[3 * e] == 0.f && mats[3 * e + 1] == 0.f && mats[3 * e + 2] == 0.f);
  if ((rc = s->reserve_assemble(s->asm_rb)) || (rc = s->plan_cg())) return rc;
  // Krylov products over the live slots when the resident kernel holds the
  // reserved rows (the compact plan then reuses the reserved plan's rows and
  // storage); otherwise over the reserved pattern, whose dead slots are zero.
  if (s->cg_res) {
    const size_t cap = s->pt.columns.size();
    if ((rc = s->alloc(&s->d_coffsets, (size_t)P + 1)) || (rc = s->alloc(&s->d_ccols, cap)) ||
        (rc = s->alloc(&s->d_cmap, cap)) || (rc = s->alloc(&s->d_cvals, 9 * cap)))
      return rc;
    s->build_compact();
    s->compact = true;
    if (s->plan_cg_nt(RES_NT, true, true) != FEM_OK) {
      s->compact = false;
      if ((rc = s->plan_cg())) return rc;
    } else if ((rc = s->upload_compact())) {
      return rc;
    }
  }
  s->topo_count = original_node_count;
  // The pattern changed: the matrix storage must be bound again.
  s->b_offsets = s->b_cols = nullptr;
  s->b_vals = nullptr;
  s->b_cap = 0;
  return FEM_OK;
}

namespace {
bool tet_enabled(const float* m) { return m[0] != 0.f || m[1] != 0.f || m[2] != 0.f; }

// Elements of a reserved pool against the last published ones
// (_stage_elements / _remove_elements / _add_elements / _compact_rows,
// streaming.py:34-245): validation, lineage and the update counters
// [dirty, removed, reused, active, touched rows, topology-inserted,
// topology-removed].  Returns the largest error code (1 source out of range,
// 2 node outside the pool, 3 edge outside the reserved adjacency, 4 repeated
// original node; streaming.py:654-662), 0 when valid.  Elements are visited
// in ascending order, so when several children name one source the first
// claims its retained topology and cached blocks.
int stage_elements(const fem_system* s, const std::vector<int32_t>& nodes, const std::vector<float>& poses,
                   const std::vector<float>& mats, const std::vector<int32_t>& src, const std::vector<float>& q_now,
                   const std::vector<float>& q_lin, int32_t counters[8]) {
  const int T = s->T, P = s->P, count = s->topo_count;
  const Pattern& pt = s->pt;
  auto has_slot = [&](int r, int c) {
    const int32_t* row = pt.columns.data() + pt.offsets[r];
    const int32_t* end = pt.columns.data() + pt.offsets[r + 1];
    const int32_t* at = std::lower_bound(row, end, c);
    return at != end && *at == c;
  };
  std::vector<unsigned char> kept(T, 0), tkept(T, 0), dirty(T, 0);
  std::vector<int32_t> retained(T, -1);
  for (int k = 0; k < 8; ++k) counters[k] = 0;
  int code = 0;
  for (int e = 0; e < T; ++e) {
    if (!tet_enabled(&mats[3 * (size_t)e])) continue;
    const int source = src[e];
    if (source < -1 || source >= T) {
      code = std::max(code, 1);
      continue;
    }
    const int32_t* nd = &nodes[4 * (size_t)e];
    bool bad = false;
    for (int a = 0; a < 4 && !bad; ++a) {
      if (nd[a] < 0 || nd[a] >= P) {
        code = std::max(code, 2);
        bad = true;
        break;
      }
      for (int b = 0; b < a; ++b)
        if (nd[a] % count == nd[b] % count) {
          code = std::max(code, 4);
          bad = true;
          break;
        }
    }
    if (bad) continue;
    bool same = source >= 0 && s->old_active[source];
    for (int a = 0; a < 4 && same; ++a) same = nd[a] == s->old_nodes[4 * (size_t)source + a];
    if (same) {
      if (!tkept[source]) {
        tkept[source] = 1;
        retained[e] = source;
      }
    } else {
      for (int a = 0; a < 4 && !bad; ++a)
        for (int b = 0; b < 4 && !bad; ++b)
          if (!has_slot(nd[a], nd[b])) bad = true;
      if (bad) {
        code = std::max(code, 3);
        continue;
      }
    }
    if (retained[e] < 0) {
      ++counters[5];
      same = false;
    }
    if (same) {
      for (int a = 0; a < 4 && same; ++a)
        for (int d = 0; d < 3; ++d)
          if (q_now[3 * (size_t)nd[a] + d] != q_lin[3 * (size_t)nd[a] + d]) same = false;
      for (int k = 0; k < 3 && same; ++k) same = mats[3 * (size_t)e + k] == s->old_mats[3 * (size_t)source + k];
      for (int k = 0; k < 9 && same; ++k) same = poses[9 * (size_t)e + k] == s->old_poses[9 * (size_t)source + k];
    }
    if (same) {
      same = !kept[source];
      kept[source] = 1;
    }
    ++counters[3];
    if (same) {
      ++counters[2];
    } else {
      dirty[e] = 1;
      ++counters[0];
    }
  }
  if (code) return code;
  // Old contributions that are not kept are removed; rows lose or gain
  // connectivity where an old element's topology was not retained or a new
  // element's topology was inserted.
  std::vector<unsigned char> touched(P, 0);
  int removed = 0, topo_removed = 0;
  for (int e = 0; e < T; ++e) {
    if (!s->old_active[e] || kept[e]) continue;
    ++removed;
    if (!tkept[e]) {
      ++topo_removed;
      for (int a = 0; a < 4; ++a) touched[s->old_nodes[4 * (size_t)e + a]] = 1;
    }
  }
  for (int e = 0; e < T; ++e)
    if (dirty[e] && retained[e] < 0)
      for (int a = 0; a < 4; ++a) touched[nodes[4 * (size_t)e + a]] = 1;
  int ntouched = 0;
  for (int i = 0; i < P; ++i) ntouched += touched[i];
  // counters so far: [0] dirty, [2] reused, [3] active, [5] inserted
  const int dirty_n = counters[0], reused = counters[2], active = counters[3], inserted = counters[5];
  counters[0] = dirty_n;
  counters[1] = removed;
  counters[2] = reused;
  counters[3] = active;
  counters[4] = ntouched;
  counters[5] = inserted;
  counters[6] = topo_removed;
  counters[7] = dirty_n > active / 2 ? 1 : 0;
  return 0;
}
}  // namespace

int fem_update_topology(fem_system* s, const int32_t* tet_indices, const float* tet_poses, const float* tet_materials,
                        const float* particle_mass, const float* particle_inv_mass, const int32_t* particle_flags,
                        const float* q, const int32_t* element_sources, float dt, int32_t* stats, void* stream) {
  if (!s || !q || !particle_mass || !particle_inv_mass || !particle_flags ||
      (s->T > 0 && (!tet_indices || !tet_poses || !tet_materials)))
    return fail(FEM_ERR_ARG, "null argument");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  if (!(dt > 0.f) || !std::isfinite(dt)) return fail(FEM_ERR_ARG, "dt must be finite and positive");
  if (s->ninst > 1) return fail(FEM_ERR_UNSUPPORTED, "topology updates of a batched system are not supported");
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipStreamSynchronize(st));
  const int T = s->T, P = s->P;
  std::vector<int32_t> nodes(4 * (size_t)T), src(T);
  std::vector<float> poses(9 * (size_t)T), mats(3 * (size_t)T);
  if (T > 0) {
    HIP_TRY(hipMemcpy(nodes.data(), tet_indices, nodes.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(poses.data(), tet_poses, poses.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(mats.data(), tet_materials, mats.size() * 4, hipMemcpyDeviceToHost));
  }
  Pattern np;
  std::string err;
  int32_t counters[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (s->topo_count) {
    if (element_sources) {
      HIP_TRY(hipMemcpy(src.data(), element_sources, src.size() * 4, hipMemcpyDeviceToHost));
    } else {
      for (int e = 0; e < T; ++e) src[e] = e;
    }
    std::vector<float> q_now(3 * (size_t)P), q_lin(3 * (size_t)P);
    HIP_TRY(hipMemcpy(q_now.data(), q, q_now.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(q_lin.data(), s->d_lin_q, q_lin.size() * 4, hipMemcpyDeviceToHost));
    static const char* msg[5] = {"", "Element source is outside the previous element capacity",
                                 "Tetrahedral node is outside the reserved particle pool",
                                 "Cut introduces an edge outside the reserved original-node adjacency",
                                 "A tetrahedron must have four distinct original basis nodes"};
    if (int code = stage_elements(s, nodes, poses, mats, src, q_now, q_lin, counters)) return fail(FEM_ERR_ARG, msg[code]);
    // New incidences into the spare pattern (its storage is reused from edit
    // to edit), then exchanged with the current ones; the previous ones
    // become the spare.
    if (int rc = assign_tet_incidences(T, nodes.data(), mats.data(), s->pt, s->spare, err))
      return fail(FEM_ERR_ARG, err);
    swap_incidences(s->pt, s->spare);
    // Assembly groups and, over live slots, the CG plan of the new
    // incidences within the reserved launches; on failure the previous ones
    // are planned again (they fitted before).
    int rc = s->plan_assemble(s->asm_rb);
    if (rc == FEM_OK && s->compact) {
      s->build_compact();
      rc = s->plan_cg_nt(RES_NT, true, true);
    }
    if (rc != FEM_OK) {
      const std::string why = g_error;
      swap_incidences(s->pt, s->spare);
      if (s->plan_assemble(s->asm_rb) == FEM_OK && s->compact) {
        s->build_compact();
        (void)s->plan_cg_nt(RES_NT, true, true);
      }
      g_error = why;
      return rc;
    }
    if (s->compact && (rc = s->upload_compact())) return rc;
    // Commit: incidences into their reserved storage, element and particle data.
    const Pattern& pt = s->pt;
    std::vector<int32_t> tptr(P + 1);
    for (int i = 0; i <= P; ++i) tptr[i] = (int32_t)pt.tinc_ptr[i];
    HIP_TRY(hipMemcpy(s->d_tinc_ptr, tptr.data(), tptr.size() * 4, hipMemcpyHostToDevice));
    if (!pt.tinc.empty()) {
      HIP_TRY(hipMemcpy(s->d_tinc, pt.tinc.data(), pt.tinc.size() * 4, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(s->d_tslots, pt.tslots.data(), pt.tslots.size() * 4, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(s->d_slot_ptr, pt.slot_ptr.data(), pt.slot_ptr.size() * 4, hipMemcpyHostToDevice));
    if (!pt.contrib.empty())
      HIP_TRY(hipMemcpy(s->d_contrib, pt.contrib.data(), pt.contrib.size() * 2, hipMemcpyHostToDevice));
    s->old_nodes = nodes;
    s->old_poses = poses;
    s->old_mats = mats;
    for (int e = 0; e < T; ++e) s->old_active[e] = tet_enabled(&mats[3 * (size_t)e]);
  } else {
    // No reserved pool: the exact pattern of the edited model, sorted again
    // (solver.py:920-935), within the bound matrix storage.
    if (int rc = build_pattern(P, T, nodes.data(), mats.data(), s->H, s->h_hex_idx.data(), s->S, s->h_sp_idx.data(), s->R,
                               s->h_tri_idx.data(), np, err))
      return fail(rc, err);
    if (np.offsets[P] > s->b_cap)
      return fail(FEM_ERR_UNSUPPORTED, "the edited topology needs more blocks than the bound matrix storage holds");
    const Pattern& cur = s->pt;
    const bool same = np.offsets == cur.offsets && np.columns == cur.columns && np.tinc == cur.tinc &&
                      np.tinc_ptr == cur.tinc_ptr && np.hinc == cur.hinc && np.hinc_ptr == cur.hinc_ptr &&
                      np.sinc == cur.sinc && np.sinc_ptr == cur.sinc_ptr;
    if (!same) {  // an unchanged topology keeps every array (and allocates nothing)
      int rc = s->edit_pattern_in_place(np);
      if (rc == FEM_ERR_UNSUPPORTED) {
        // Does not fit the current arrays and launches: new ones.  Graphs
        // captured before this edit name the old ones, which therefore stay
        // allocated (and describe the old topology) until the system is
        // destroyed; such graphs must be captured again (counters[0] = 1).
        s->keep_replaced = true;
        s->pt = std::move(np);
        if ((rc = s->upload_pattern(4 * (size_t)T, 16 * (size_t)T + 64 * (size_t)s->H, (size_t)s->b_cap)))
          return rc;
        if ((rc = s->plan_assemble(s->asm_rb)) || (rc = s->plan_cg())) return rc;
        counters[0] = 1;
      } else if (rc != FEM_OK) {
        return rc;
      }
      std::vector<int32_t> off32(P + 1);
      for (int i = 0; i <= P; ++i) off32[i] = (int32_t)s->pt.offsets[i];
      HIP_TRY(hipMemcpy(s->b_offsets, off32.data(), off32.size() * 4, hipMemcpyHostToDevice));
      if (!s->pt.columns.empty())
        HIP_TRY(hipMemcpy(s->b_cols, s->pt.columns.data(), s->pt.columns.size() * 4, hipMemcpyHostToDevice));
    }
  }
  if (T > 0) {
    HIP_TRY(hipMemcpyAsync(s->d_tet_idx, tet_indices, 16 * (size_t)T, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(s->d_tet_pose, tet_poses, 36 * (size_t)T, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(s->d_tet_mat, tet_materials, 12 * (size_t)T, hipMemcpyDeviceToDevice, st));
  }
  HIP_TRY(hipMemcpyAsync(s->d_mass, particle_mass, 4 * (size_t)P, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_inv_mass, particle_inv_mass, 4 * (size_t)P, hipMemcpyDeviceToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_flags, particle_flags, 4 * (size_t)P, hipMemcpyDeviceToDevice, st));
  if (int rc = fem_assemble(s, q, dt, stream)) return rc;
  HIP_TRY(hipStreamSynchronize(st));
  if (stats)
    for (int k = 0; k < 8; ++k) stats[k] = counters[k];
  return FEM_OK;
}

int fem_bind_component_scaling(fem_system* s, int32_t capacity, double* totals, float* scale, float* values,
                               float* inv_diag, float* rhs) {
  if (!s || !totals || !scale || !values || !inv_diag || !rhs) return fail(FEM_ERR_ARG, "null buffer");
  if (s->cs_scale) return fail(FEM_ERR_STATE, "Component scaling is already prepared");
  if (s->nfixed) return fail(FEM_ERR_ARG, "Component scaling is not supported with fixed_particle_indices");
  if (s->kr_precond != KP_BLOCK) return fail(FEM_ERR_ARG, "Component scaling requires block_diag preconditioning");
  if (s->kr_method != KR_CG) return fail(FEM_ERR_UNSUPPORTED, "Component scaling runs the block-Jacobi PCG only");
  if (capacity < 1) return fail(FEM_ERR_ARG, "component_capacity must be a positive integer");
  if (s->ninst > 1) return fail(FEM_ERR_UNSUPPORTED, "component scaling of a batched system is not supported");
  if (!s->b_vals) return fail(FEM_ERR_STATE, "fem_bind must be called first");
  s->cs_cap = capacity;
  s->cs_total = totals;
  s->cs_scale = scale;
  s->cs_vals = values;
  s->cs_minv = inv_diag;
  s->cs_rhs = rhs;
  return FEM_OK;
}

int fem_refresh_component_scaling(fem_system* s, const int32_t* labels, const float* particle_mass, void* stream) {
  if (!s || !labels || !particle_mass) return fail(FEM_ERR_ARG, "null argument");
  if (!s->cs_scale) return fail(FEM_ERR_STATE, "Prepare component scaling before refreshing it");
  hipStream_t st = (hipStream_t)stream;
  const int nb = (std::max(s->P, s->cs_cap) + BLOCK - 1) / BLOCK;
  hipLaunchKernelGGL(k_zero_doubles, dim3((s->cs_cap + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, s->cs_total, s->cs_cap);
  if (s->P > 0) {
    hipLaunchKernelGGL(k_component_mass, dim3(nb), dim3(BLOCK), 0, st, labels, particle_mass, s->P, s->cs_cap,
                       s->cs_total);
    hipLaunchKernelGGL(k_component_scale, dim3(nb), dim3(BLOCK), 0, st, labels, (const double*)s->cs_total, s->P,
                       s->cs_cap, s->cs_scale);
  }
  HIP_TRY(hipGetLastError());
  return FEM_OK;
}

}  // extern "C"
