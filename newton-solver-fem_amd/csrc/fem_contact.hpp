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

// Soft contacts between particles and rigid shapes, from a collision
// pipeline's contact list (reference eval_soft_contacts kernels.py:1199-1252,
// driven by SolverFEM.eval_soft_contact_forces solver.py:1540-1566): contact
// k < min(*count, max_contacts) pushes particle particle[k] out of the body
// surface point body_pos[k] along normal[k] (body -> particle) with penalty
// ke, damping kd on approach and Coulomb friction (kf, mu).  The reference
// adds every contact's force with an atomic; here the contacts are grouped by
// particle (counting sort) and each particle sums its own in contact order,
// so results are deterministic.
struct SoftContacts {
  int P = 0;
  int max_contacts = 0;
  const int* count = nullptr;       // [1] device count of valid entries
  const int* particle = nullptr;    // [max_contacts] (negative: skipped)
  const float* body_pos = nullptr;  // [max_contacts][3]
  const float* body_vel = nullptr;  // [max_contacts][3]
  const float* normal = nullptr;    // [max_contacts][3]
  float ke = 0.f, kd = 0.f, kf = 0.f, mu = 0.f;
  void* temp = nullptr;    // soft_contact_temp_bytes(P, max_contacts)
  float* force = nullptr;  // [P][3] soft-contact force of the last evaluation
};

// Scratch bytes of the grouping for P particles and max_contacts entries.
size_t soft_contact_temp_bytes(int P, int max_contacts);

// Soft-contact force of every particle at (x, v) into c.force.  Allocates
// nothing; graph-capturable (the count is read on the device).
hipError_t soft_contact_forces(const SoftContacts& c, const float* x, const float* v, const float* radius,
                               hipStream_t st);

}  // namespace fem
