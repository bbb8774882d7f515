// Particle-particle contact through a hash grid (reference
// eval_particle_forces kernels.py:1069-1114, particle_force :1043-1066;
// eval_particle_particle_forces solver.py:1477-1510).  See fem_contact.hpp.
//
// Grid build per evaluation, all with fixed launch shapes:
//   k_pc_clear    bucket counts = 0
//   k_pc_count    bucket of every particle (cell hash), counts by atomics
//   k_pc_scan*    exclusive prefix sum of the counts -> bucket starts
//   k_pc_scatter  particle ids into their bucket (arrival order)
//   k_pc_order    each bucket's ids sorted ascending (so that the order of
//                 every force sum is fixed: results are deterministic)
//   k_pc_forces   one thread per particle over the 27 neighbouring cells
#include <hip/hip_runtime.h>

#include <cmath>

#include "fem_contact.hpp"
#include "fem_math.hpp"

namespace fem {
namespace {

constexpr int NT = 256;
constexpr int SCAN_NT = 1024;  // one scan block covers 1024 buckets; at most 1024 blocks

__device__ __forceinline__ int cell_of(float x, float inv_cell) {
  // Far-away (or non-finite) coordinates still map to some cell; pairs are
  // always confirmed by their distance.
  const float c = fminf(fmaxf(floorf(x * inv_cell), -1.0e9f), 1.0e9f);
  return (int)c;
}

__device__ __forceinline__ uint32_t bucket_of(int ix, int iy, int iz, uint32_t mask) {
  return ((uint32_t)ix * 73856093u ^ (uint32_t)iy * 19349663u ^ (uint32_t)iz * 83492791u) & mask;
}

__global__ __launch_bounds__(NT) void k_pc_clear(int* count, int n) {
  const int t = blockIdx.x * NT + threadIdx.x;
  if (t < n) count[t] = 0;
}

__global__ __launch_bounds__(NT) void k_pc_count(const float* __restrict__ x, int P, float inv_cell, uint32_t mask,
                                                 uint32_t* __restrict__ keys, int* count) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= P) return;
  const v3 p = ld3(x, i);
  const uint32_t h = bucket_of(cell_of(p.x, inv_cell), cell_of(p.y, inv_cell), cell_of(p.z, inv_cell), mask);
  keys[i] = h;
  atomicAdd(count + h, 1);
}

// Exclusive scan of n counts (n <= SCAN_NT * SCAN_NT) in three launches.
__device__ __forceinline__ int block_scan_inclusive(int v, int* lds) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  if (lane == 63) lds[w] = v;
  __syncthreads();
  if (t < 64) {
    int s = t < SCAN_NT / 64 ? lds[t] : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(s, o);
      if (lane >= o) s += u;
    }
    if (t < SCAN_NT / 64) lds[t] = s;
  }
  __syncthreads();
  return v + (w > 0 ? lds[w - 1] : 0);
}

__global__ __launch_bounds__(SCAN_NT) void k_pc_scan_blocks(const int* __restrict__ count, int n,
                                                            int* __restrict__ start, int* __restrict__ bsum) {
  __shared__ int lds[SCAN_NT / 64];
  const int i = blockIdx.x * SCAN_NT + threadIdx.x;
  const int v = i < n ? count[i] : 0;
  const int inc = block_scan_inclusive(v, lds);
  if (i < n) start[i] = inc - v;
  if (threadIdx.x == SCAN_NT - 1) bsum[blockIdx.x] = inc;
}

__global__ __launch_bounds__(SCAN_NT) void k_pc_scan_sums(int* bsum, int nb) {
  __shared__ int lds[SCAN_NT / 64];
  const int t = threadIdx.x;
  const int v = t < nb ? bsum[t] : 0;
  const int inc = block_scan_inclusive(v, lds);
  if (t < nb) bsum[t] = inc - v;
}

__global__ __launch_bounds__(SCAN_NT) void k_pc_scan_add(int* __restrict__ start, int n, const int* __restrict__ bsum,
                                                         int* __restrict__ cursor, int total) {
  const int i = blockIdx.x * SCAN_NT + threadIdx.x;
  if (i < n) {
    const int s = start[i] + bsum[blockIdx.x];
    start[i] = s;
    cursor[i] = s;
  }
  if (i == 0) start[n] = total;
}

__global__ __launch_bounds__(NT) void k_pc_scatter(const uint32_t* __restrict__ keys, int P, int* cursor,
                                                   int* __restrict__ ids) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= P) return;
  ids[atomicAdd(cursor + keys[i], 1)] = i;
}

// Each bucket's ids in ascending order (insertion sort; buckets hold a few
// particles each).
__global__ __launch_bounds__(NT) void k_pc_order(const int* __restrict__ start, int n, int* __restrict__ ids) {
  const int h = blockIdx.x * NT + threadIdx.x;
  if (h >= n) return;
  const int b = start[h], e = start[h + 1];
  for (int k = b + 1; k < e; ++k) {
    const int v = ids[k];
    int m = k - 1;
    while (m >= b && ids[m] > v) {
      ids[m + 1] = ids[m];
      --m;
    }
    ids[m + 1] = v;
  }
}

// particle_force kernels.py:1043-1066: penalty + damping + Coulomb friction on
// particle i; n points from j to i, v is the relative velocity, c the gap.
__device__ __forceinline__ v3 pair_force(v3 n, v3 v, float c, const ContactParams& p) {
  const float vn = dot(n, v);
  const float jn = c * p.ke;
  const float jd = fminf(vn, 0.f) * p.kd;
  const float fn = jn + jd;
  v3 vt = v - n * vn;
  const float vs = sqrtf(dot(vt, vt));
  if (vs > 0.f) vt = v3{vt.x / vs, vt.y / vs, vt.z / vs};
  const float ft = fminf(vs * p.kf, p.mu * fabsf(fn));
  return n * (-fn) - vt * ft;
}

__global__ __launch_bounds__(NT) void k_pc_forces(const float* __restrict__ x, const float* __restrict__ v,
                                                  const float* __restrict__ radius, const int* __restrict__ flags,
                                                  int P, float inv_cell, uint32_t mask, const int* __restrict__ start,
                                                  const int* __restrict__ ids, ContactParams p,
                                                  float* __restrict__ out) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= P) return;
  v3 f = {0.f, 0.f, 0.f};
  if (flags[i] & 1) {
    const v3 xi = ld3(x, i), vi = ld3(v, i);
    const float ri = radius[i];
    const int cx = cell_of(xi.x, inv_cell), cy = cell_of(xi.y, inv_cell), cz = cell_of(xi.z, inv_cell);
    for (int s = 0; s < 27; ++s) {
      const int dx = s % 3 - 1, dy = (s / 3) % 3 - 1, dz = s / 9 - 1;
      const uint32_t h = bucket_of(cx + dx, cy + dy, cz + dz, mask);
      // A bucket reached from an earlier neighbouring cell is walked once.
      bool seen = false;
      for (int u = 0; u < s && !seen; ++u)
        seen = bucket_of(cx + u % 3 - 1, cy + (u / 3) % 3 - 1, cz + u / 9 - 1, mask) == h;
      if (seen) continue;
      const int e = start[h + 1];
      for (int k = start[h]; k < e; ++k) {
        const int j = ids[k];
        if (j == i || (flags[j] & 1) == 0) continue;
        v3 n = xi - ld3(x, j);
        const float d = sqrtf(dot(n, n));
        if (d < 1.0e-10f) continue;  // co-located vertices have no contact direction
        const float err = d - ri - radius[j];
        if (err <= p.cohesion) {
          n = v3{n.x / d, n.y / d, n.z / d};
          f = f + pair_force(n, vi - ld3(v, j), err, p);
        }
      }
    }
  }
  st3(out, i, f);
}

}  // namespace

size_t contact_temp_bytes(int bits) {
  // counts + starts (n + 1) + cursors + block sums, ints
  const size_t n = size_t(1) << bits;
  return sizeof(int) * (3 * n + 1 + SCAN_NT);
}

namespace {
// Contacts per particle (entries past the device count or with a particle
// outside 0 .. P-1 take no part).
__global__ __launch_bounds__(NT) void k_sc_count(const int* __restrict__ count, int max_contacts,
                                                 const int* __restrict__ particle, int P, int* per) {
  const int k = blockIdx.x * NT + threadIdx.x;
  if (k >= max_contacts || k >= *count) return;
  const int p = particle[k];
  if (p >= 0 && p < P) atomicAdd(per + p, 1);
}

__global__ __launch_bounds__(NT) void k_sc_scatter(const int* __restrict__ count, int max_contacts,
                                                   const int* __restrict__ particle, int P, int* cursor,
                                                   int* __restrict__ ids) {
  const int k = blockIdx.x * NT + threadIdx.x;
  if (k >= max_contacts || k >= *count) return;
  const int p = particle[k];
  if (p >= 0 && p < P) ids[atomicAdd(cursor + p, 1)] = k;
}

// One thread per particle: its contacts put in ascending order, then their
// forces summed in that order (eval_soft_contacts kernels.py:1222-1252).
__global__ __launch_bounds__(NT) void k_sc_forces(const float* __restrict__ x, const float* __restrict__ v,
                                                  const float* __restrict__ radius, int P,
                                                  const int* __restrict__ start, const int* __restrict__ per,
                                                  int* __restrict__ ids, const float* __restrict__ body_pos,
                                                  const float* __restrict__ body_vel,
                                                  const float* __restrict__ normal, float ke, float kd, float kf,
                                                  float mu, float* __restrict__ force) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i >= P) return;
  const int b = start[i], e = b + per[i];
  for (int k = b + 1; k < e; ++k) {
    const int c = ids[k];
    int m = k - 1;
    while (m >= b && ids[m] > c) {
      ids[m + 1] = ids[m];
      --m;
    }
    ids[m + 1] = c;
  }
  v3 f = {0.f, 0.f, 0.f};
  if (e > b) {
    const v3 xi = ld3(x, i), vi = ld3(v, i);
    const float r = radius[i];
    for (int k = b; k < e; ++k) {
      const int c = ids[k];
      const v3 n = ld3(normal, c);
      // Penetration along the normal (body -> particle).
      const float pen = r - dot(xi - ld3(body_pos, c), n);
      if (pen > 0.f) {
        const v3 rel = vi - ld3(body_vel, c);
        const float vn = dot(rel, n);
        const v3 vt = rel - n * vn;
        const float fn = ke * pen - kd * fminf(vn, 0.f);
        const float vt_mag = sqrtf(dot(vt, vt));
        if (vt_mag > 1.0e-6f) {
          const float ft = fminf(kf * vt_mag, mu * fmaxf(0.f, fn));
          f = f + (n * fn - vt * (ft / vt_mag));
        } else {
          f = f + n * fn;
        }
      }
    }
  }
  st3(force, i, f);
}
}  // namespace

size_t soft_contact_temp_bytes(int P, int max_contacts) {
  return sizeof(int) * (3 * (size_t)P + 1 + SCAN_NT + (size_t)std::max(max_contacts, 1));
}

hipError_t soft_contact_forces(const SoftContacts& c, const float* x, const float* v, const float* radius,
                               hipStream_t st) {
  if (c.P == 0) return hipSuccess;
  const int P = c.P;
  int* per = static_cast<int*>(c.temp);
  int* start = per + P;        // [P + 1]
  int* cursor = start + P + 1; // [P]
  int* bsum = cursor + P;      // [SCAN_NT]
  int* ids = bsum + SCAN_NT;   // [max_contacts]
  const int nbk = (P + SCAN_NT - 1) / SCAN_NT;
  const int gp = (P + NT - 1) / NT, gc = (std::max(c.max_contacts, 1) + NT - 1) / NT;
  hipLaunchKernelGGL(k_pc_clear, dim3(gp), dim3(NT), 0, st, per, P);
  if (c.max_contacts > 0)
    hipLaunchKernelGGL(k_sc_count, dim3(gc), dim3(NT), 0, st, c.count, c.max_contacts, c.particle, P, per);
  hipLaunchKernelGGL(k_pc_scan_blocks, dim3(nbk), dim3(SCAN_NT), 0, st, (const int*)per, P, start, bsum);
  hipLaunchKernelGGL(k_pc_scan_sums, dim3(1), dim3(SCAN_NT), 0, st, bsum, nbk);
  hipLaunchKernelGGL(k_pc_scan_add, dim3(nbk), dim3(SCAN_NT), 0, st, start, P, (const int*)bsum, cursor, 0);
  if (c.max_contacts > 0)
    hipLaunchKernelGGL(k_sc_scatter, dim3(gc), dim3(NT), 0, st, c.count, c.max_contacts, c.particle, P, cursor, ids);
  hipLaunchKernelGGL(k_sc_forces, dim3(gp), dim3(NT), 0, st, x, v, radius, P, (const int*)start, (const int*)per, ids,
                     c.body_pos, c.body_vel, c.normal, c.ke, c.kd, c.kf, c.mu, c.force);
  return hipGetLastError();
}

hipError_t contact_forces(const ContactGrid& g, const ContactParams& p, const float* x, const float* v,
                          const float* radius, const int* flags, hipStream_t st) {
  if (g.P == 0) return hipSuccess;
  const int n = 1 << g.bits;
  const uint32_t mask = (uint32_t)n - 1u;
  const float inv_cell = 1.f / g.cell;
  int* count = static_cast<int*>(g.temp);
  int* start = count + n;          // [n + 1]
  int* cursor = start + n + 1;     // [n]
  int* bsum = cursor + n;          // [SCAN_NT]
  const int nbk = (n + SCAN_NT - 1) / SCAN_NT;
  const int gp = (g.P + NT - 1) / NT, gn = (n + NT - 1) / NT;
  hipLaunchKernelGGL(k_pc_clear, dim3(gn), dim3(NT), 0, st, count, n);
  hipLaunchKernelGGL(k_pc_count, dim3(gp), dim3(NT), 0, st, x, g.P, inv_cell, mask, g.keys, count);
  hipLaunchKernelGGL(k_pc_scan_blocks, dim3(nbk), dim3(SCAN_NT), 0, st, (const int*)count, n, start, bsum);
  hipLaunchKernelGGL(k_pc_scan_sums, dim3(1), dim3(SCAN_NT), 0, st, bsum, nbk);
  hipLaunchKernelGGL(k_pc_scan_add, dim3(nbk), dim3(SCAN_NT), 0, st, start, n, (const int*)bsum, cursor, g.P);
  hipLaunchKernelGGL(k_pc_scatter, dim3(gp), dim3(NT), 0, st, (const uint32_t*)g.keys, g.P, cursor, g.ids);
  hipLaunchKernelGGL(k_pc_order, dim3(gn), dim3(NT), 0, st, (const int*)start, n, g.ids);
  hipLaunchKernelGGL(k_pc_forces, dim3(gp), dim3(NT), 0, st, x, v, radius, flags, g.P, inv_cell, mask,
                     (const int*)start, (const int*)g.ids, p, g.force);
  return hipGetLastError();
}

}  // namespace fem
