// Exact multi-index machinery on the host.
//
// Restates the numba kernels of src/smolyax/indices.py with the same DFS push order and the same
// sequence of IEEE double operations (remaining / used budgets are formed exactly as there), so
// every membership test, cardinality and Smolyak coefficient is bit-identical.  This file must be
// compiled with -ffp-contract=off: a fused multiply-add in `rem - j * k[i]` would change results.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/smolyax_amd.h"
#include "common.h"

namespace {

// indices.py:121-168
int64_t smolyak_coefficient_impl(const double* k, int64_t d, double rem_t, int32_t parity,
                                 std::vector<std::pair<int64_t, std::pair<double, int32_t>>>& st) {
  int64_t total = 0;
  st.clear();
  st.push_back({0, {rem_t, parity}});
  while (!st.empty()) {
    auto e = st.back();
    st.pop_back();
    const int64_t i = e.first;
    const double rt = e.second.first;
    const int32_t p = e.second.second;
    if (i >= d) {
      total += 1 - (p << 1);
      continue;
    }
    if (i + 1 < d && k[i + 1] < rt) {
      st.push_back({i + 1, {rt, p}});
    } else {
      total += 1 - (p << 1);
    }
    const double cost = k[i];
    if (cost < rt) st.push_back({i + 1, {rt - cost, p ^ 1}});
  }
  return total;
}

// indices.py:72-118
int64_t indexset_cardinality_impl(const double* k, int64_t d, double t) {
  std::vector<std::pair<int64_t, double>> st;
  st.push_back({0, 0.0});
  int64_t count = 0;
  while (!st.empty()) {
    auto e = st.back();
    st.pop_back();
    const int64_t dim_i = e.first;
    const double used_t = e.second;
    if (dim_i >= d) {
      count += 1;
      continue;
    }
    const double remaining_t = t - used_t;
    if (dim_i + 1 < d && k[dim_i + 1] < remaining_t) {
      st.push_back({dim_i + 1, used_t});
    } else {
      count += 1;
    }
    int64_t j = 1;
    while (used_t + (double)j * k[dim_i] < t) {
      const double new_used = used_t + (double)j * k[dim_i];
      st.push_back({dim_i + 1, new_used});
      j += 1;
    }
  }
  return count;
}

// indices.py:175-211
int64_t nodeset_cardinality_non_nested_impl(const double* k, int64_t d, double t) {
  struct E {
    int64_t dim;
    double rem;
    int32_t parity;
    int64_t prod;
  };
  std::vector<E> st;
  std::vector<std::pair<int64_t, std::pair<double, int32_t>>> scratch;
  st.push_back({0, t, 0, 1});
  int64_t total = 0;
  while (!st.empty()) {
    E e = st.back();
    st.pop_back();
    if (e.dim >= d || !(e.dim + 1 < d && k[e.dim + 1] < e.rem)) {
      if (smolyak_coefficient_impl(k, d, e.rem, e.parity, scratch) != 0) total += e.prod;
    }
    if (e.dim < d) {
      if (e.dim + 1 < d && k[e.dim + 1] < e.rem) st.push_back({e.dim + 1, e.rem, e.parity, e.prod});
      const double cost = k[e.dim];
      int64_t j = 1;
      while (cost * (double)j < e.rem) {
        st.push_back({e.dim + 1, e.rem - cost * (double)j, e.parity ^ (int32_t)(j & 1),
                      e.prod * (j + 1)});
        j += 1;
      }
    }
  }
  return total;
}

int64_t nodeset_cardinality_impl(const double* k, int64_t d, double t, int32_t nested) {
  return nested ? indexset_cardinality_impl(k, d, t) : nodeset_cardinality_non_nested_impl(k, d, t);
}

// Sparse multi-index on the DFS stack: CSR slices into a shared pool avoid per-node allocation.
struct Pool {
  std::vector<int32_t> dims, degs;
};

}  // namespace

// k must hold d >= 0 values (k may be NULL only when d == 0)
static bool bad_k(const double* k, int64_t d, const char* fn) {
  if (d < 0 || (d > 0 && !k)) {
    sbi::set_error(std::string(fn) + ": null k or negative d");
    return true;
  }
  return false;
}

extern "C" {

int64_t sbi_indexset_cardinality(const double* k, int64_t d, double t) {
  if (bad_k(k, d, "sbi_indexset_cardinality")) return -1;
  return indexset_cardinality_impl(k, d, t);
}

int64_t sbi_smolyak_coefficient(const double* k, int64_t d, double rem_t, int32_t parity) {
  if (bad_k(k, d, "sbi_smolyak_coefficient")) return -1;
  std::vector<std::pair<int64_t, std::pair<double, int32_t>>> st;
  return smolyak_coefficient_impl(k, d, rem_t, parity, st);
}

int64_t sbi_nodeset_cardinality(const double* k, int64_t d, double t, int32_t nested) {
  if (bad_k(k, d, "sbi_nodeset_cardinality")) return -1;
  return nodeset_cardinality_impl(k, d, t, nested);
}

// indices.py:292-355
double sbi_find_approximate_threshold(const double* k, int64_t d, int64_t m, int32_t nested,
                                      int32_t max_iter, double accuracy) {
  if (bad_k(k, d, "sbi_find_approximate_threshold")) return NAN;
  if (m <= 0) {
    sbi::set_error("find_approximate_threshold: m must be > 0");
    return NAN;
  }
  if (m == 1) return 1.0;
  double lo = 1.0, hi = 2.0;
  while (nodeset_cardinality_impl(k, d, lo, nested) > m) {
    hi = lo;
    lo = lo / 1.2;
  }
  while (nodeset_cardinality_impl(k, d, hi, nested) < m) {
    lo = hi;
    hi = hi * 1.2;
  }
  double t_cand = lo + (hi - lo) / 2.0;
  int64_t m_cand = nodeset_cardinality_impl(k, d, t_cand, nested);
  for (int32_t it = 0; it < max_iter; ++it) {
    if (m_cand > m) {
      hi = t_cand;
    } else {
      lo = t_cand;
    }
    t_cand = lo + (hi - lo) / 2.0;
    m_cand = nodeset_cardinality_impl(k, d, t_cand, nested);
    // np.abs(m_cand - m) / m < accuracy: integer difference, then true division in double
    if (std::fabs((double)(m_cand - m)) / (double)m < accuracy) break;
  }
  return t_cand;
}

// indices.py:20-69
int sbi_indexset(const double* k, int64_t d, double t, int64_t* n_nu, int64_t* n_entries,
                 int64_t* offsets, int32_t* dims, int32_t* degs) {
  if (!k || d < 0 || !n_nu || !n_entries) {
    sbi::set_error("sbi_indexset: null argument");
    return SBI_ERR_INVALID;
  }
  struct E {
    int64_t i;
    double rem;
    int64_t head, len;  // slice of the pool
  };
  Pool pool;
  std::vector<E> st;
  st.push_back({0, t, 0, 0});
  int64_t cnt = 0, ent = 0;
  const bool fill = dims != nullptr;
  if (fill) offsets[0] = 0;
  while (!st.empty()) {
    E e = st.back();
    st.pop_back();
    if (e.i >= d || k[e.i] >= e.rem) {
      if (fill) {
        for (int64_t q = 0; q < e.len; ++q) {
          dims[ent + q] = pool.dims[e.head + q];
          degs[ent + q] = pool.degs[e.head + q];
        }
        offsets[cnt + 1] = ent + e.len;
      }
      cnt += 1;
      ent += e.len;
      continue;
    }
    st.push_back({e.i + 1, e.rem, e.head, e.len});
    int64_t j = 1;
    const double k_i = k[e.i];
    while ((double)j * k_i < e.rem) {
      const int64_t h = (int64_t)pool.dims.size();
      for (int64_t q = 0; q < e.len; ++q) {
        pool.dims.push_back(pool.dims[e.head + q]);
        pool.degs.push_back(pool.degs[e.head + q]);
      }
      pool.dims.push_back((int32_t)e.i);
      pool.degs.push_back((int32_t)j);
      st.push_back({e.i + 1, e.rem - (double)j * k_i, h, e.len + 1});
      j += 1;
    }
  }
  *n_nu = cnt;
  *n_entries = ent;
  return SBI_OK;
}

// indices.py:214-259
int sbi_nonzero_indices_and_zetas(const double* k, int64_t d, double t, int64_t* n_nu,
                                  int64_t* n_entries, int64_t* offsets, int32_t* dims,
                                  int32_t* degs, int64_t* zetas) {
  if (!k || d < 0 || !n_nu || !n_entries) {
    sbi::set_error("sbi_nonzero_indices_and_zetas: null argument");
    return SBI_ERR_INVALID;
  }
  struct E {
    int64_t i;
    double rem;
    int64_t head, len;
  };
  Pool pool;
  std::vector<E> st;
  std::vector<std::pair<int64_t, std::pair<double, int32_t>>> scratch;
  st.push_back({0, t, 0, 0});
  int64_t cnt = 0, ent = 0;
  const bool fill = dims != nullptr;
  if (fill) offsets[0] = 0;
  while (!st.empty()) {
    E e = st.back();
    st.pop_back();
    if (e.i >= d || !(e.i + 1 < d && k[e.i + 1] < e.rem)) {
      const int64_t zeta = smolyak_coefficient_impl(k, d, e.rem, 0, scratch);
      if (zeta != 0) {
        if (fill) {
          for (int64_t q = 0; q < e.len; ++q) {
            dims[ent + q] = pool.dims[e.head + q];
            degs[ent + q] = pool.degs[e.head + q];
          }
          offsets[cnt + 1] = ent + e.len;
          zetas[cnt] = zeta;
        }
        cnt += 1;
        ent += e.len;
      }
    }
    if (e.i < d) {
      if (e.i + 1 < d && k[e.i + 1] < e.rem) st.push_back({e.i + 1, e.rem, e.head, e.len});
      int64_t j = 1;
      while ((double)j * k[e.i] < e.rem) {
        const int64_t h = (int64_t)pool.dims.size();
        for (int64_t q = 0; q < e.len; ++q) {
          pool.dims.push_back(pool.dims[e.head + q]);
          pool.degs.push_back(pool.degs[e.head + q]);
        }
        pool.dims.push_back((int32_t)e.i);
        pool.degs.push_back((int32_t)j);
        st.push_back({e.i + 1, e.rem - (double)j * k[e.i], h, e.len + 1});
        j += 1;
      }
    }
  }
  *n_nu = cnt;
  *n_entries = ent;
  return SBI_OK;
}

}  // extern "C"
