// Host-side construction of the BSR pattern of A = M + h^2 K and of the
// per-node element incidence lists used by the row-centric HIP kernels.
//
// The pattern equals what the reference obtains from bsr_set_from_triplets
// over the triplet layout [P diagonal | 2 per spring | 3 diagonal per
// triangle | 16 per Tet4 | 64 per Hex8] (solver.py:312-341, 1008-1136,
// kernels.py:141-153, :294-366, :972-1030, :635-850, :1609-1747): per row the
// sorted unique column indices of every valid triplet.  Tets whose three
// material coefficients are all zero write row/column -1 and are excluded
// (kernels.py:679-692), as are springs with a -1 endpoint; every Hex8 and every
// particle diagonal is present (triangles only add diagonal blocks).
//
// Incidences list, for each node i, the (element, local corner) pairs with
// n_a == i, in ascending element order.  For each incidence the relative
// slot (position inside row i) of every block K_ab is precomputed, replacing
// the triplet->slot binary search of sparse_update.map_triplet_blocks
// (sparse_update.py:9-30).
#include "fem_pattern.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <thread>

namespace fem {

static bool tet_valid(const float* mat) { return !(mat[0] == 0.f && mat[1] == 0.f && mat[2] == 0.f); }

static bool spring_valid(const int32_t* idx) { return idx[0] >= 0 && idx[1] >= 0; }

// Tet4 and Hex8 incidence lists of every node (valid tets only), ascending
// element order.
static int element_incidences(int P, int T, const int32_t* tet_idx, const float* tet_mat, int H, const int32_t* hex_idx,
                              Pattern& pt, std::string& err) {
  // Incidence counts.
  pt.tinc_ptr.assign(P + 1, 0);
  pt.hinc_ptr.assign(P + 1, 0);
  for (int e = 0; e < T; ++e) {
    if (!tet_valid(tet_mat + 3 * (size_t)e)) continue;
    for (int a = 0; a < 4; ++a) {
      int n = tet_idx[4 * (size_t)e + a];
      if (n < 0 || n >= P) {
        err = "tet_indices contains an out-of-range node";
        return 1;
      }
      pt.tinc_ptr[n + 1]++;
    }
  }
  for (int e = 0; e < H; ++e)
    for (int a = 0; a < 8; ++a) {
      int n = hex_idx[8 * (size_t)e + a];
      if (n < 0 || n >= P) {
        err = "hex_indices contains an out-of-range node";
        return 1;
      }
      pt.hinc_ptr[n + 1]++;
    }
  for (int i = 0; i < P; ++i) {
    pt.tinc_ptr[i + 1] += pt.tinc_ptr[i];
    pt.hinc_ptr[i + 1] += pt.hinc_ptr[i];
  }
  pt.tinc.resize(pt.tinc_ptr[P]);
  pt.hinc.resize(pt.hinc_ptr[P]);
  {
    std::vector<int64_t> fill(pt.tinc_ptr.begin(), pt.tinc_ptr.end() - 1);
    for (int e = 0; e < T; ++e) {
      if (!tet_valid(tet_mat + 3 * (size_t)e)) continue;
      for (int a = 0; a < 4; ++a) pt.tinc[fill[tet_idx[4 * (size_t)e + a]]++] = (int32_t)(((uint32_t)e << 2) | a);
    }
    std::vector<int64_t> hfill(pt.hinc_ptr.begin(), pt.hinc_ptr.end() - 1);
    for (int e = 0; e < H; ++e)
      for (int a = 0; a < 8; ++a) pt.hinc[hfill[hex_idx[8 * (size_t)e + a]]++] = (int32_t)(((uint32_t)e << 3) | a);
  }
  if (T >= (1 << 29) || H >= (1 << 28)) {
    err = "element count exceeds the packed incidence limit";
    return 3;
  }
  return 0;
}

// Rows [r0, r1) of element_slots: relative slot bytes of the rows'
// incidences, the per-slot counts and their prefix inside each row (the
// slot_ptr entries of a row's slots are private to that row), and the row's
// total.  Returns 2 when an element block has no slot in the row.
static int slots_rows(int r0, int r1, const int32_t* tet_idx, const int32_t* hex_idx, const Pattern& rows,
                      Pattern& pt, int64_t* row_total) {
  for (int i = r0; i < r1; ++i) {
    const int32_t* row = rows.columns.data() + rows.offsets[i];
    const int nnz = (int)(rows.offsets[i + 1] - rows.offsets[i]);
    int32_t* cnt = pt.slot_ptr.data() + rows.offsets[i] + 1;  // cnt[s] counts slot s of row i
    std::fill(cnt, cnt + nnz, 0);
    for (int64_t k = pt.tinc_ptr[i]; k < pt.tinc_ptr[i + 1]; ++k) {
      const int e = pt.tinc[k] >> 2;
      uint32_t packed = 0;
      for (int b = 0; b < 4; ++b) {
        const int c = tet_idx[4 * (size_t)e + b];
        const int32_t* at = std::lower_bound(row, row + nnz, c);
        if (at == row + nnz || *at != c) return 2;
        packed |= (uint32_t)(at - row) << (8 * b);
        ++cnt[at - row];
      }
      pt.tslots[k] = packed;
    }
    for (int64_t k = pt.hinc_ptr[i]; k < pt.hinc_ptr[i + 1]; ++k) {
      const int e = pt.hinc[k] >> 3;
      uint32_t lo = 0, hi = 0;
      for (int b = 0; b < 8; ++b) {
        const int c = hex_idx[8 * (size_t)e + b];
        const int32_t* at = std::lower_bound(row, row + nnz, c);
        if (at == row + nnz || *at != c) return 2;
        const uint32_t rel = (uint32_t)(at - row);
        if (b < 4) lo |= rel << (8 * b);
        else hi |= rel << (8 * (b - 4));
        ++cnt[rel];
      }
      pt.hslots[2 * k] = lo;
      pt.hslots[2 * k + 1] = hi;
    }
    int64_t run = 0;
    for (int s = 0; s < nnz; ++s) {
      run += cnt[s];
      cnt[s] = (int32_t)run;
    }
    row_total[i] = run;
  }
  return 0;
}

// Rows [r0, r1): the row base added to the row's slot_ptr entries, then each
// slot's contributions in incidence order (tets, then hexes).
static void contrib_rows(int r0, int r1, const Pattern& rows, Pattern& pt, const int64_t* row_base) {
  int32_t cur[256];
  for (int i = r0; i < r1; ++i) {
    const int64_t o = rows.offsets[i];
    const int nnz = (int)(rows.offsets[i + 1] - o);
    int32_t* sp = pt.slot_ptr.data() + o;  // sp[s] = start of slot s once the base is added
    const int32_t base = (int32_t)row_base[i];
    for (int s = 1; s <= nnz; ++s) sp[s] += base;
    for (int s = 0; s < nnz; ++s) cur[s] = s == 0 ? base : sp[s];
    for (int64_t k = pt.tinc_ptr[i]; k < pt.tinc_ptr[i + 1]; ++k) {
      const uint32_t t = (uint32_t)(k - pt.tinc_ptr[i]);
      for (int b = 0; b < 4; ++b)
        pt.contrib[cur[(pt.tslots[k] >> (8 * b)) & 255u]++] = (uint16_t)((t << 3) | (uint32_t)b);
    }
    for (int64_t k = pt.hinc_ptr[i]; k < pt.hinc_ptr[i + 1]; ++k) {
      const uint32_t t = (uint32_t)(k - pt.hinc_ptr[i]);
      for (int b = 0; b < 8; ++b) {
        const uint32_t w = pt.hslots[2 * k + (b >> 2)];
        pt.contrib[cur[(w >> (8 * (b & 3))) & 255u]++] = (uint16_t)(0x8000u | (t << 3) | (uint32_t)b);
      }
    }
  }
}

// Runs fn(r0, r1) over `parts` contiguous row ranges of about equal work
// (incidences and blocks), on that many threads.
template <typename F>
static void for_row_ranges(const Pattern& rows, const Pattern& pt, int parts, F fn) {
  const int P = pt.P;
  if (parts <= 1 || P < 2048) {
    fn(0, P);
    return;
  }
  auto work = [&](int r) { return (double)(4 * pt.tinc_ptr[r] + 8 * pt.hinc_ptr[r] + rows.offsets[r]) + r; };
  const double total = work(P);
  std::vector<int> cut(parts + 1, P);
  cut[0] = 0;
  for (int j = 1; j < parts; ++j) {
    int lo = cut[j - 1], hi = P;
    const double target = total * j / parts;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (work(mid) < target) lo = mid + 1;
      else hi = mid;
    }
    cut[j] = lo;
  }
  std::vector<std::thread> threads;
  for (int j = 1; j < parts; ++j) threads.emplace_back(fn, cut[j], cut[j + 1]);
  fn(cut[0], cut[1]);
  for (auto& t : threads) t.join();
}

// Relative slots of every incidence's blocks and the per-slot contribution
// lists of pt's incidences, against the rows of `rows` (offsets/columns; may be
// pt itself), in fixed order whatever the thread count.  Returns 2 when an
// element block has no slot in those rows.
static int element_slots(const int32_t* tet_idx, const int32_t* hex_idx, const Pattern& rows, Pattern& pt,
                         std::string& err) {
  const int P = pt.P;
  for (int i = 0; i < P; ++i)
    if (pt.tinc_ptr[i + 1] - pt.tinc_ptr[i] > 4096 || pt.hinc_ptr[i + 1] - pt.hinc_ptr[i] > 4096) {
      err = "a node belongs to more than 4096 elements of one kind";
      return 3;
    }
  const size_t nnz = (size_t)rows.offsets[P];
  pt.tslots.resize(pt.tinc.size());
  pt.hslots.resize(2 * pt.hinc.size());
  pt.slot_ptr.resize(nnz + 1);  // every entry but the first is written per row
  std::vector<int64_t> row_total(P + 1, 0);
  const int parts = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  std::atomic<int> failed{0};
  for_row_ranges(rows, pt, parts, [&](int r0, int r1) {
    if (slots_rows(r0, r1, tet_idx, hex_idx, rows, pt, row_total.data())) failed.store(1);
  });
  if (failed.load()) {
    err = "an element block lies outside the sparse pattern";
    return 2;
  }
  // Exclusive scan of the row totals: the first contribution of each row.
  int64_t run = 0;
  for (int i = 0; i < P; ++i) {
    const int64_t t = row_total[i];
    row_total[i] = run;
    run += t;
  }
  pt.contrib.resize((size_t)run);
  for_row_ranges(rows, pt, parts, [&](int r0, int r1) { contrib_rows(r0, r1, rows, pt, row_total.data()); });
  pt.slot_ptr[0] = 0;
  return 0;
}

int build_pattern(int P, int T, const int32_t* tet_idx, const float* tet_mat, int H, const int32_t* hex_idx, int S,
                  const int32_t* spring_idx, int R, const int32_t* tri_idx, Pattern& pt, std::string& err) {
  pt = Pattern();
  pt.P = P;
  // Spring and triangle incidences.
  pt.sinc_ptr.assign(P + 1, 0);
  pt.rinc_ptr.assign(P + 1, 0);
  for (int s = 0; s < S; ++s) {
    const int32_t* e = spring_idx + 2 * (size_t)s;
    if (e[0] < -1 || e[1] < -1 || e[0] >= P || e[1] >= P) {
      err = "spring_indices contains an out-of-range particle";
      return 1;
    }
    if (!spring_valid(e)) continue;
    pt.sinc_ptr[e[0] + 1]++;
    pt.sinc_ptr[e[1] + 1]++;
  }
  for (int t = 0; t < R; ++t)
    for (int a = 0; a < 3; ++a) {
      int n = tri_idx[3 * (size_t)t + a];
      if (n < 0 || n >= P) {
        err = "tri_indices contains an out-of-range particle";
        return 1;
      }
      pt.rinc_ptr[n + 1]++;
    }
  for (int i = 0; i < P; ++i) {
    pt.sinc_ptr[i + 1] += pt.sinc_ptr[i];
    pt.rinc_ptr[i + 1] += pt.rinc_ptr[i];
  }
  if (S >= (1 << 30) || R >= (1 << 29)) {
    err = "spring or triangle count exceeds the packed incidence limit";
    return 3;
  }
  pt.sinc.resize(pt.sinc_ptr[P]);
  pt.rinc.resize(pt.rinc_ptr[P]);
  {
    std::vector<int64_t> fill(pt.sinc_ptr.begin(), pt.sinc_ptr.end() - 1);
    for (int s = 0; s < S; ++s) {
      const int32_t* e = spring_idx + 2 * (size_t)s;
      if (!spring_valid(e)) continue;
      pt.sinc[fill[e[0]]++] = (int32_t)((uint32_t)s << 1);
      pt.sinc[fill[e[1]]++] = (int32_t)(((uint32_t)s << 1) | 1u);
    }
    std::vector<int64_t> rfill(pt.rinc_ptr.begin(), pt.rinc_ptr.end() - 1);
    for (int t = 0; t < R; ++t)
      for (int a = 0; a < 3; ++a) pt.rinc[rfill[tri_idx[3 * (size_t)t + a]]++] = (int32_t)(((uint32_t)t << 2) | a);
  }
  if (int rc = element_incidences(P, T, tet_idx, tet_mat, H, hex_idx, pt, err)) return rc;
  // Rows: sorted unique columns.
  pt.offsets.assign(P + 1, 0);
  pt.columns.clear();
  pt.columns.reserve((size_t)pt.tinc_ptr[P] * 3 + (size_t)pt.hinc_ptr[P] * 7 + P);
  std::vector<int32_t> cand;
  pt.diag_rel.assign(P, -1);
  int maxnnz = 0;
  for (int i = 0; i < P; ++i) {
    cand.clear();
    cand.push_back(i);
    for (int64_t k = pt.tinc_ptr[i]; k < pt.tinc_ptr[i + 1]; ++k) {
      int e = pt.tinc[k] >> 2;
      for (int b = 0; b < 4; ++b) cand.push_back(tet_idx[4 * (size_t)e + b]);
    }
    for (int64_t k = pt.hinc_ptr[i]; k < pt.hinc_ptr[i + 1]; ++k) {
      int e = pt.hinc[k] >> 3;
      for (int b = 0; b < 8; ++b) cand.push_back(hex_idx[8 * (size_t)e + b]);
    }
    for (int64_t k = pt.sinc_ptr[i]; k < pt.sinc_ptr[i + 1]; ++k) {
      const int s = pt.sinc[k] >> 1;
      cand.push_back(spring_idx[2 * (size_t)s + (1 - (pt.sinc[k] & 1))]);
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    int nnz = (int)cand.size();
    maxnnz = std::max(maxnnz, nnz);
    if (nnz > 255) {
      char buf[128];
      std::snprintf(buf, sizeof buf, "row %d has %d blocks; the assembly kernels support at most 255", i, nnz);
      err = buf;
      return 3;
    }
    pt.diag_rel[i] = (int32_t)(std::lower_bound(cand.begin(), cand.end(), i) - cand.begin());
    pt.columns.insert(pt.columns.end(), cand.begin(), cand.end());
    pt.offsets[i + 1] = (int64_t)pt.columns.size();
    if (pt.offsets[i + 1] >= ((int64_t)1 << 31)) {
      err = "more than 2^31 blocks";
      return 3;
    }
  }
  pt.max_row_nnz = maxnnz;
  return element_slots(tet_idx, hex_idx, pt, pt, err);
}

// Reserved rows of a duplicated-node pool (streaming.py:487-516): the pool
// holds P / count copies of `count` original nodes, pool node i standing for
// original node i % count.  Row i holds i and every copy of every original
// neighbour of i % count, the neighbours taken from the valid tets' original
// connectivity; a later cut may connect any copies of those pairs without
// changing the rows.  The current tets' incidences are assigned as well.
int build_reserved_pattern(int P, int count, int T, const int32_t* tet_idx, const float* tet_mat, Pattern& pt,
                           std::string& err) {
  pt = Pattern();
  pt.P = P;
  if (count < 1 || P % count) {
    err = "The particle pool must contain whole copies of the original node set";
    return 1;
  }
  const int copies = P / count;
  std::vector<std::vector<int32_t>> adj(count);
  for (int e = 0; e < T; ++e) {
    if (!tet_valid(tet_mat + 3 * (size_t)e)) continue;
    int orig[4];
    for (int a = 0; a < 4; ++a) {
      const int n = tet_idx[4 * (size_t)e + a];
      if (n < 0 || n >= P) {
        err = "Tetrahedral node indices are outside the reserved particle pool";
        return 1;
      }
      orig[a] = n % count;
      for (int b = 0; b < a; ++b)
        if (orig[b] == orig[a]) {
          err = "Each tetrahedron must have four distinct original basis nodes";
          return 1;
        }
    }
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        if (a != b) adj[orig[a]].push_back(orig[b]);
  }
  for (auto& v : adj) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  pt.offsets.assign(P + 1, 0);
  pt.diag_rel.assign(P, -1);
  std::vector<int32_t> cand;
  int maxnnz = 0;
  for (int i = 0; i < P; ++i) {
    cand.assign(1, i);
    for (int c = 0; c < copies; ++c)
      for (int32_t nb : adj[i % count]) cand.push_back(nb + c * count);
    std::sort(cand.begin(), cand.end());
    const int nnz = (int)cand.size();
    maxnnz = std::max(maxnnz, nnz);
    if (nnz > 255) {
      char buf[128];
      std::snprintf(buf, sizeof buf, "row %d has %d blocks; the assembly kernels support at most 255", i, nnz);
      err = buf;
      return 3;
    }
    pt.diag_rel[i] = (int32_t)(std::lower_bound(cand.begin(), cand.end(), i) - cand.begin());
    pt.columns.insert(pt.columns.end(), cand.begin(), cand.end());
    pt.offsets[i + 1] = (int64_t)pt.columns.size();
    if (pt.offsets[i + 1] >= ((int64_t)1 << 31)) {
      err = "more than 2^31 blocks";
      return 3;
    }
  }
  pt.max_row_nnz = maxnnz;
  return assign_tet_incidences(T, tet_idx, tet_mat, pt, pt, err);
}

int assign_tet_incidences(int T, const int32_t* tet_idx, const float* tet_mat, const Pattern& rows, Pattern& out,
                          std::string& err) {
  out.P = rows.P;
  out.sinc_ptr.assign(out.P + 1, 0);
  out.sinc.clear();
  out.rinc_ptr.assign(out.P + 1, 0);
  out.rinc.clear();
  out.hslots.clear();
  if (int rc = element_incidences(out.P, T, tet_idx, tet_mat, 0, nullptr, out, err)) return rc;
  return element_slots(tet_idx, nullptr, rows, out, err);
}

void swap_incidences(Pattern& a, Pattern& b) {
  a.tinc_ptr.swap(b.tinc_ptr);
  a.tinc.swap(b.tinc);
  a.tslots.swap(b.tslots);
  a.hinc_ptr.swap(b.hinc_ptr);
  a.hinc.swap(b.hinc);
  a.hslots.swap(b.hslots);
  a.slot_ptr.swap(b.slot_ptr);
  a.contrib.swap(b.contrib);
  a.sinc_ptr.swap(b.sinc_ptr);
  a.sinc.swap(b.sinc);
  a.rinc_ptr.swap(b.rinc_ptr);
  a.rinc.swap(b.rinc);
}

// Balanced contiguous row ranges: `nchunks` pieces of [r0, r1) with roughly
// equal cost = blocks + 8 * rows.
void split_rows(const std::vector<int64_t>& offsets, int r0, int r1, int nchunks, std::vector<int32_t>& bounds) {
  bounds.clear();
  bounds.push_back(r0);
  auto cost = [&](int r) { return (double)(offsets[r] - offsets[r0]) + 8.0 * (r - r0); };
  double total = cost(r1);
  int r = r0;
  for (int c = 1; c < nchunks; ++c) {
    double target = total * c / nchunks;
    int lo = r, hi = r1;
    while (lo < hi) {
      int mid = lo + (hi - lo) / 2;
      if (cost(mid) < target) lo = mid + 1;
      else hi = mid;
    }
    r = lo;
    bounds.push_back(r);
  }
  bounds.push_back(r1);
}

}  // namespace fem
