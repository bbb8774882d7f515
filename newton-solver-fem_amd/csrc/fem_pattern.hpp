// Host-side BSR pattern and element incidence construction.
#pragma once

#include <algorithm>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace fem {

struct Pattern {
  int P = 0;
  std::vector<int64_t> offsets;   // [P+1]
  std::vector<int32_t> columns;   // [nnz]
  std::vector<int32_t> diag_rel;  // [P] position of (i,i) inside row i
  std::vector<int64_t> tinc_ptr;  // [P+1]
  std::vector<int32_t> tinc;      // (element << 2) | corner
  std::vector<uint32_t> tslots;   // 4 relative slots, 8 bits each
  std::vector<int64_t> hinc_ptr;  // [P+1]
  std::vector<int32_t> hinc;      // (element << 3) | corner
  std::vector<uint32_t> hslots;   // 8 relative slots, two words per incidence
  // Per slot, the element blocks that sum into it, in fixed order (tets by
  // element, then hexes): (is_hex << 15) | (incidence index in row << 3) | corner b.
  std::vector<int32_t> slot_ptr;  // [nnz+1]
  std::vector<uint16_t> contrib;
  // Springs with both endpoints valid: per node, (spring << 1) | end (end 0:
  // the node is the spring's first particle).  Triangles: (tri << 2) | corner.
  std::vector<int64_t> sinc_ptr;  // [P+1]
  std::vector<int32_t> sinc;
  std::vector<int64_t> rinc_ptr;  // [P+1]
  std::vector<int32_t> rinc;
  int max_row_nnz = 0;
};

int build_pattern(int P, int T, const int32_t* tet_idx, const float* tet_mat, int H, const int32_t* hex_idx, int S,
                  const int32_t* spring_idx, int R, const int32_t* tri_idx, Pattern& pt, std::string& err);

// Rows reserved for a pool of P / count copies of `count` original nodes
// (every copy of every original neighbour), with the valid tets' incidences.
int build_reserved_pattern(int P, int count, int T, const int32_t* tet_idx, const float* tet_mat, Pattern& pt,
                           std::string& err);

// Tet4 incidences, slots and contributions (tets only) of new elements
// against the rows of `rows`, written into the incidence fields of `out` (which
// may be `rows` itself; its storage is reused); returns 2 when a valid tet has a
// block outside the rows.
int assign_tet_incidences(int T, const int32_t* tet_idx, const float* tet_mat, const Pattern& rows, Pattern& out,
                          std::string& err);

// Exchanges the incidence, slot and contribution fields of two patterns.
void swap_incidences(Pattern& a, Pattern& b);

// Runs fn(begin, end) over contiguous pieces of [0, n) of about min_per items
// or more, on up to 16 host threads (the calling thread takes the first piece).
template <typename F>
void parallel_pieces(int n, int min_per, F fn) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int parts = std::max(1, std::min({16, hw, n / std::max(1, min_per)}));
  if (parts <= 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> threads;
  for (int j = 1; j < parts; ++j)
    threads.emplace_back(fn, (int)((int64_t)n * j / parts), (int)((int64_t)n * (j + 1) / parts));
  fn(0, (int)((int64_t)n / parts));
  for (auto& t : threads) t.join();
}

void split_rows(const std::vector<int64_t>& offsets, int r0, int r1, int nchunks, std::vector<int32_t>& bounds);

}  // namespace fem
