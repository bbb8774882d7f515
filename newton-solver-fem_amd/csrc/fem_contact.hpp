// Particle-particle contact through a hash grid (reference
// eval_particle_forces kernels.py:1069-1114, driven by
// SolverFEM.eval_particle_particle_forces solver.py:1477-1510).
//
// Per evaluation: every particle's cell (width = 2 max radius + cohesion, so
// that any pair that can touch lies in neighbouring cells) is hashed into a
// power-of-two bucket table; a counting sort groups the particle ids by
// bucket, each bucket's ids are put in ascending order, and one thread per
// particle walks the (deduplicated) buckets of its 27 neighbouring cells in
// a fixed order and sums the pair forces.  The order of every sum is fixed,
// so results are deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fem {

struct ContactGrid {
  int P = 0;
  int bits = 0;            // log2 of the bucket count (<= 20)
  float cell = 0.f;        // cell width
  float max_radius = 0.f;  // largest particle radius
  uint32_t* keys = nullptr;  // [P] bucket of each particle
  int* ids = nullptr;        // [P] particle ids grouped by bucket
  void* temp = nullptr;      // contact_temp_bytes: counts, bucket starts, cursors, scan sums
  float* force = nullptr;    // [P][3] contact force of the last evaluation
};

struct ContactParams {
  float ke, kd, kf, mu, cohesion;
};

// Scratch bytes of the grid build for `bits` bucket bits.
size_t contact_temp_bytes(int bits);

// Contact force of every particle at (x, v) into g.force (zero for inactive
// particles: flags bit 0 clear).  Allocates nothing; graph-capturable.
hipError_t contact_forces(const ContactGrid& g, const ContactParams& p, const float* x, const float* v,
                          const float* radius, const int* flags, hipStream_t st);

}  // namespace fem
