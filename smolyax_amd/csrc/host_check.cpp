// Host-side exercise of the C ABI for a build under AddressSanitizer and UndefinedBehaviorSanitizer
// (CPU only: no HIP call, no device memory).  It runs the index-set entry points with exactly
// sized buffers and through their error paths, and the plan entry points (creation, validation,
// destruction) for barycentric and Newton descriptors whose pointers only have to be non-null.
// It prints one line per index-set case,
//   case <d> <m> <t> <|Lambda|> <entries> <dims checksum> <nonzero blocks> <sum of zetas> <nodes>
// for comparison with the production library, then "ok"; any sanitizer report or failed check
// ends it with a non-zero status.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/smolyax_amd.h"

namespace {

int failures = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "host_check: failed: %s (%s)\n", what, sbi_last_error());
    ++failures;
  }
}

void index_case(int d, long m) {
  std::vector<double> k(d);
  for (int j = 0; j < d; ++j) k[j] = std::log((2.0 + j) / std::log(2.0));
  const double t = sbi_find_approximate_threshold(k.data(), d, m, 1, 32, 0.001);
  expect(std::isfinite(t), "threshold");
  const int64_t card = sbi_indexset_cardinality(k.data(), d, t);
  int64_t n_nu = -1, n_ent = -1;
  expect(sbi_indexset(k.data(), d, t, &n_nu, &n_ent, nullptr, nullptr, nullptr) == SBI_OK,
         "indexset sizes");
  expect(n_nu == card, "indexset size equals cardinality");
  // exactly sized buffers (one element at least, as the bindings allocate)
  std::vector<int64_t> off(n_nu + 1);
  std::vector<int32_t> dims(n_ent > 0 ? n_ent : 1), degs(n_ent > 0 ? n_ent : 1);
  expect(sbi_indexset(k.data(), d, t, &n_nu, &n_ent, off.data(), dims.data(), degs.data()) ==
             SBI_OK,
         "indexset fill");
  long long chk = 0;
  for (int64_t i = 0; i < n_ent; ++i) chk += (long long)(dims[i] + 1) * (degs[i] + 3) * (i % 7 + 1);
  int64_t z_nu = -1, z_ent = -1;
  expect(sbi_nonzero_indices_and_zetas(k.data(), d, t, &z_nu, &z_ent, nullptr, nullptr, nullptr,
                                       nullptr) == SBI_OK,
         "nonzero sizes");
  std::vector<int64_t> zoff(z_nu + 1), zetas(z_nu > 0 ? z_nu : 1);
  std::vector<int32_t> zdims(z_ent > 0 ? z_ent : 1), zdegs(z_ent > 0 ? z_ent : 1);
  expect(sbi_nonzero_indices_and_zetas(k.data(), d, t, &z_nu, &z_ent, zoff.data(), zdims.data(),
                                       zdegs.data(), zetas.data()) == SBI_OK,
         "nonzero fill");
  long long zsum = 0;
  for (int64_t i = 0; i < z_nu; ++i) zsum += zetas[i];
  const int64_t nodes = sbi_nodeset_cardinality(k.data(), d, t, 1);
  std::printf("case %d %ld %.17g %lld %lld %lld %lld %lld %lld\n", d, m, t, (long long)n_nu,
              (long long)n_ent, chk, (long long)z_nu, zsum, (long long)nodes);
}

void error_paths() {
  int64_t a = 0, b = 0;
  expect(sbi_indexset_cardinality(nullptr, 3, 1.0) == -1, "cardinality rejects a null k");
  double k1[1] = {1.0};
  expect(sbi_indexset_cardinality(k1, -1, 1.0) == -1, "cardinality rejects a negative d");
  expect(std::isnan(sbi_find_approximate_threshold(nullptr, 2, 10, 1, 32, 0.001)),
         "threshold rejects a null k");
  expect(sbi_indexset(nullptr, 2, 1.0, &a, &b, nullptr, nullptr, nullptr) == SBI_ERR_INVALID,
         "indexset rejects a null k");
  expect(sbi_nonzero_indices_and_zetas(nullptr, 2, 1.0, &a, &b, nullptr, nullptr, nullptr,
                                       nullptr) == SBI_ERR_INVALID,
         "nonzero rejects a null k");
  expect(std::strlen(sbi_last_error()) > 0, "an error message is set");
}

void plan_paths() {
  static double dummy[64];
  void* p = dummy;
  sbi_bary_desc d;
  std::memset(&d, 0, sizeof(d));
  d.d_in = 3;
  d.d_out = 1;
  d.n_slots = 2;
  d.dmax1 = 12;
  d.n_prefix_vals = 4;
  d.n_prefix_slots = 1;
  d.n_runs = 1;
  d.n_blocks = 2;
  d.max_n = 2;
  d.F = (const double*)p;
  d.offset = (const double*)p;
  d.blk_desc = (const int32_t*)p;
  d.seg_off = (const int32_t*)p;
  d.run_start = (const int32_t*)p;
  d.blk_fac = (const int32_t*)p;
  d.blk_foff = (const int64_t*)p;
  d.n_split = 16;
  sbi_plan* plan = nullptr;
  expect(sbi_bary_plan_create(&d, &plan) == SBI_OK && plan, "bary plan");
  const sbi_bary_desc* back = nullptr;
  expect(sbi_bary_plan_desc(plan, &back) == SBI_OK && back && back->d_in == 3, "bary plan desc");
  expect(sbi_bary_integral_workspace(plan) >= 0, "integral workspace size");
  expect(sbi_plan_destroy(plan) == SBI_OK, "bary plan destroy");
  // record program fields must be complete and consistent
  d.pseg = (const int32_t*)p;
  expect(sbi_bary_plan_create(&d, &plan) == SBI_ERR_INVALID, "program without its arrays");
  d.pseg_off = (const int32_t*)p;
  d.pbasis = (const double*)p;
  d.prow = (const int32_t*)p;
  d.pF = (const double*)p;
  d.n_pseg = 1;
  d.ones_row = d.n_prefix_vals;
  expect(sbi_bary_plan_create(&d, &plan) == SBI_OK && plan, "bary plan with program");
  expect(sbi_plan_destroy(plan) == SBI_OK, "bary plan with program destroy");
  d.n_prefix_vals = d.ones_row = 400;
  expect(sbi_bary_plan_create(&d, &plan) == SBI_ERR_CAPACITY, "program table beyond LDS");
  d.n_split = 8;
  expect(sbi_bary_plan_create(&d, &plan) == SBI_ERR_INVALID, "bary plan with 8 waves");
  expect(sbi_bary_plan_create(nullptr, &plan) == SBI_ERR_INVALID, "bary plan without desc");

  sbi_newton_desc n;
  std::memset(&n, 0, sizeof(n));
  n.d_in = 3;
  n.d_out = 1;
  n.n_nodes = 8;
  n.n_internal = 4;
  n.n_groups = 2;
  n.n_split = 16;
  n.kmax1 = 4;
  n.s_root = (const double*)p;
  n.wave_grp = (const int32_t*)p;
  n.grp = (const int32_t*)p;
  n.node_q = (const int32_t*)p;
  n.node_s32 = (const float*)p;
  n.node_s64 = (const double*)p;
  sbi_newton_plan* np = nullptr;
  expect(sbi_newton_plan_create(&n, &np) == SBI_OK && np, "newton plan");
  expect(sbi_newton_mixed_staged(np) == 0, "newton plan without staged tables streams");
  expect(sbi_newton_eval_mixed_gemm_workspace(np, 64) == -1, "no GEMM fields: no workspace size");
  expect(sbi_newton_plan_destroy(np) == SBI_OK, "newton plan destroy");
  n.gemm_kq = 16;  // basis_n = 0: no FP32 rows to pad
  expect(sbi_newton_plan_create(&n, &np) == SBI_ERR_INVALID, "newton plan with stray GEMM rows");
  n.basis_n = 20;
  n.gemm_head = 4;
  n.head_fac = (const int32_t*)p;
  n.gemm_s64 = (const double*)p;
  n.ld_s64 = 1;
  n.gemm_sq = (const float*)p;
  expect(sbi_newton_plan_create(&n, &np) == SBI_OK && np, "newton plan with GEMM fields");
  expect(sbi_newton_eval_mixed_gemm_workspace(np, 64) >= 16 * 64 * 4 + 4 * 64 * 8,
         "GEMM workspace holds both basis parts");
  expect(sbi_newton_plan_destroy(np) == SBI_OK, "newton plan with GEMM fields destroy");
  n.gemm_kq = n.gemm_head = 0;
  n.n_split = 4;
  expect(sbi_newton_plan_create(&n, &np) == SBI_ERR_INVALID, "newton plan with 4 waves");
}

}  // namespace

int main() {
  index_case(1, 9);
  index_case(4, 120);
  index_case(10, 1000);
  index_case(30, 600);
  error_paths();
  plan_paths();
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
