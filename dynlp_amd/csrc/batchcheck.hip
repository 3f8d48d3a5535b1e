// Host-side check of an edge batch before any device work (no GPU calls): ids in [0, n),
// insertion weights finite and positive, and whether an unordered pair repeats within the
// deletions or within the insertions.  One pass with an open-addressing table per list, instead
// of the sorts a NumPy check needs; the caller collapses repeats only when this finds some.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

template <typename T>
int64_t first_out_of_range(int64_t m, const T* a, int64_t n) {
  for (int64_t i = 0; i < m; ++i)
    if ((int64_t)a[i] < 0 || (int64_t)a[i] >= n) return i;
  return -1;
}

// true when some unordered pair {u[i], v[i]} occurs twice (ids already in range)
template <typename T>
bool has_repeat(int64_t m, const T* u, const T* v, int64_t n, std::vector<uint64_t>& table) {
  if (m < 2) return false;
  int lg = 4;
  while (((size_t)1 << lg) < 2 * (size_t)m) ++lg;
  const size_t cap = (size_t)1 << lg;
  table.assign(cap, ~0ull);   // ~0: empty (no key reaches it: key < n^2 <= 2^62)
  const size_t mask = cap - 1;
  for (int64_t i = 0; i < m; ++i) {
    const uint64_t a = (uint64_t)(int64_t)u[i], b = (uint64_t)(int64_t)v[i];
    const uint64_t key = (a < b ? a : b) * (uint64_t)n + (a < b ? b : a);
    size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - lg));   // Fibonacci hashing
    while (true) {
      if (table[h] == ~0ull) {
        table[h] = key;
        break;
      }
      if (table[h] == key) return true;
      h = (h + 1) & mask;
    }
  }
  return false;
}

template <typename T>
int check(int64_t n, int64_t nd, const T* du, const T* dv, int64_t ni, const T* iu, const T* iv, const float* iw,
          int64_t* where) {
  const T* lists[4] = {du, dv, iu, iv};
  const int64_t lens[4] = {nd, nd, ni, ni};
  int64_t base = 0;
  for (int k = 0; k < 4; ++k) {
    const int64_t i = first_out_of_range(lens[k], lists[k], n);
    if (i >= 0) {
      if (where) *where = base + i;
      return -1;
    }
    base += lens[k];
  }
  for (int64_t i = 0; i < ni; ++i) {
    if (!(std::isfinite(iw[i]) && iw[i] > 0.f)) {
      if (where) *where = i;
      return -2;
    }
  }
  thread_local std::vector<uint64_t> table;
  int rc = 0;
  if (has_repeat(nd, du, dv, n, table)) rc |= 1;
  if (has_repeat(ni, iu, iv, n, table)) rc |= 2;
  return rc;
}

}  // namespace

// ids: id_bytes 4 (int32) or 8 (int64) per entry.  Returns 0 (valid, no repeated pair), a mask of
// 1 (a pair repeats within the deletions) and 2 (within the insertions), -1 (an id outside
// [0, n); *where = its position in del_u, del_v, ins_u, ins_v taken one after the other), -2 (an
// insertion weight that is not finite and positive; *where = its index) or -3 (bad arguments).
extern "C" int dlp_check_edge_batch(int64_t n, int64_t nd, const void* del_u, const void* del_v, int64_t ni,
                                    const void* ins_u, const void* ins_v, const float* ins_w, int id_bytes,
                                    int64_t* where) {
  if (n < 1 || n > (int64_t)1 << 31 || nd < 0 || ni < 0 || (id_bytes != 4 && id_bytes != 8) ||
      (nd && (!del_u || !del_v)) || (ni && (!ins_u || !ins_v || !ins_w)))
    return -3;
  if (id_bytes == 4)
    return check<int32_t>(n, nd, (const int32_t*)del_u, (const int32_t*)del_v, ni, (const int32_t*)ins_u,
                          (const int32_t*)ins_v, ins_w, where);
  return check<int64_t>(n, nd, (const int64_t*)del_u, (const int64_t*)del_v, ni, (const int64_t*)ins_u,
                        (const int64_t*)ins_v, ins_w, where);
}
