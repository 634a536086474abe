// Continuous-batching scheduler for many resident interpolants on one GPU (the serving form of
// __call__, interpolation.py:264-304, for many interpolants at once), in C++ so that the host
// work per batch is a handful of HIP calls.
//
// Requests (interpolant, points on the device, output on the device) are admitted in arrival
// order: the oldest pending request picks the interpolant and the following requests of that
// interpolant join its batch (whole) up to the interpolant's point cap and kMaxBatchReqs
// requests; a request above the cap is cut into pieces of at most the cap, one batch each.  Up to
// n_streams batches are in flight, one per HIP stream.  A batch of one piece reads its points and
// writes its results in place; a batch of several requests gathers their points into the slot's
// staging buffer with one kernel, evaluates, and scatters the rows back with one kernel.  Every
// evaluation kernel computes each point on its own, so a request's result is the same bits alone
// and in any batch.
//
// Ordering: submit() records a pooled event on the caller's stream (after the points were
// produced there); each batch's stream waits for the events of its requests before reading.  A
// request is complete when the events of all its batches have been observed by step().

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <exception>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/smolyax_amd.h"
#include "common.h"

namespace {

constexpr int kMaxBatchReqs = 64;

struct RowCopy {
  int32_t n;                         // requests in the batch
  int32_t width;                     // doubles per row (d_in or d_out)
  const double* src[kMaxBatchReqs];  // gather: request rows; scatter: staging rows
  double* dst[kMaxBatchReqs];
  int64_t src_ld[kMaxBatchReqs];
  int64_t dst_ld[kMaxBatchReqs];
  int64_t rows[kMaxBatchReqs];
};

// One request per blockIdx.y: rows x width doubles from src (row stride src_ld) to dst (dst_ld).
__global__ __launch_bounds__(256) void copy_rows_kernel(RowCopy c) {
  const int i = blockIdx.y;
  if (i >= c.n) return;
  const int64_t w = c.width, total = c.rows[i] * w;
  const double* __restrict__ s = c.src[i];
  double* __restrict__ d = c.dst[i];
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * 256) {
    const int64_t r = e / w, col = e - r * w;
    d[r * c.dst_ld[i] + col] = s[r * c.src_ld[i] + col];
  }
}

struct Interp {
  const sbi_newton_plan* plan;
  int64_t d_in, d_out, cap;
  bool stream_ok;  // d_out == 1 and the streaming kernel takes the plan
  bool gemm_ok;    // the plan has the mixed GEMM fields
};

struct Req {
  int32_t ip;
  const double* x;
  int64_t P, ldx;
  double* y;
  int64_t ldy;
  int64_t pos = 0;    // points already in launched batches
  int32_t parts = 0;  // launched batches not observed complete
  bool taken = false;
  hipEvent_t ready = nullptr;
};

struct Piece {
  int64_t rid, a0, n;
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Piece> pieces;
  char* buf = nullptr;  // staging x | staging y | evaluation workspace
  int64_t buf_bytes = 0;
};

int64_t align256(int64_t v) { return (v + 255) / 256 * 256; }

}  // namespace

struct sbi_server {
  std::vector<Interp> ips;
  std::vector<Slot> slots;
  std::unordered_map<int64_t, Req> reqs;
  std::deque<int64_t> pending;                // arrival order (taken requests skipped lazily)
  std::vector<std::deque<int64_t>> by_ip;     // per interpolant, arrival order
  std::vector<hipEvent_t> ev_pool;
  int64_t next = 0;
  int64_t n_pending = 0;
  int64_t xs_bytes = 0, ys_bytes = 0, ws_bytes = 0;  // per-slot buffer parts
  bool broken = false;  // a call failed: every later call but destroy is refused
};

namespace {

int set_invalid(const char* msg) {
  sbi::set_error(msg);
  return SBI_ERR_INVALID;
}

int eval(sbi_server* s, const Interp& ip, Slot& sl, const double* x, int64_t P, int64_t ldx,
         double* y, int64_t ldy) {
  if (ip.stream_ok && ldy == 1) {
    const int rc = sbi_newton_eval_mixed(ip.plan, x, P, ldx, y, sl.st);
    if (rc != SBI_ERR_CAPACITY) return rc;
  }
  if (!ip.gemm_ok) return set_invalid("sbi_server_step: the interpolant has no mixed GEMM fields");
  return sbi_newton_eval_mixed_gemm(ip.plan, x, P, ldx, y, ldy,
                                    sl.buf + s->xs_bytes + s->ys_bytes, s->ws_bytes, sl.st);
}

int launch(sbi_server* s, Slot& sl, int32_t ipi, const std::vector<Piece>& pieces, int64_t pts) {
  const Interp& ip = s->ips[ipi];
  for (const Piece& pc : pieces) {
    Req& r = s->reqs.at(pc.rid);
    if (r.ready) SBI_HIP_CHECK(hipStreamWaitEvent(sl.st, r.ready, 0));
    if (r.taken && r.ready) {  // no later batch waits for it: the event may be recorded again
      s->ev_pool.push_back(r.ready);
      r.ready = nullptr;
    }
  }
  int rc;
  if (pieces.size() == 1) {
    const Req& r = s->reqs.at(pieces[0].rid);
    const int64_t a0 = pieces[0].a0;
    rc = eval(s, ip, sl, r.x + a0 * r.ldx, pieces[0].n, r.ldx, r.y + a0 * r.ldy, r.ldy);
  } else {
    double* xs = (double*)sl.buf;
    double* ys = (double*)(sl.buf + s->xs_bytes);
    RowCopy g{}, c{};
    g.n = c.n = (int32_t)pieces.size();
    g.width = (int32_t)ip.d_in;
    c.width = (int32_t)ip.d_out;
    int64_t row = 0, most = 0;
    for (size_t i = 0; i < pieces.size(); ++i) {
      const Req& r = s->reqs.at(pieces[i].rid);
      const int64_t a0 = pieces[i].a0, n = pieces[i].n;
      g.src[i] = r.x + a0 * r.ldx;
      g.src_ld[i] = r.ldx;
      g.dst[i] = xs + row * ip.d_in;
      g.dst_ld[i] = ip.d_in;
      g.rows[i] = n;
      c.src[i] = ys + row * ip.d_out;
      c.src_ld[i] = ip.d_out;
      c.dst[i] = r.y + a0 * r.ldy;
      c.dst_ld[i] = r.ldy;
      c.rows[i] = n;
      row += n;
      if (n > most) most = n;
    }
    auto grid_for = [&](int64_t width) {
      int64_t gx = (most * width + 255) / 256;
      if (gx > 1024) gx = 1024;
      if (gx < 1) gx = 1;
      return dim3((unsigned)gx, (unsigned)pieces.size());
    };
    hipLaunchKernelGGL(copy_rows_kernel, grid_for(ip.d_in), dim3(256), 0, sl.st, g);
    SBI_HIP_CHECK(hipGetLastError());
    rc = eval(s, ip, sl, xs, pts, ip.d_in, ys, ip.d_out);
    if (rc == SBI_OK) {
      hipLaunchKernelGGL(copy_rows_kernel, grid_for(ip.d_out), dim3(256), 0, sl.st, c);
      SBI_HIP_CHECK(hipGetLastError());
    }
  }
  if (rc != SBI_OK) return rc;
  SBI_HIP_CHECK(hipEventRecord(sl.done, sl.st));
  sl.busy = true;
  sl.pieces = pieces;
  return SBI_OK;
}

int admit(sbi_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.busy) continue;
    while (!s->pending.empty()) {
      auto it = s->reqs.find(s->pending.front());
      if (it != s->reqs.end() && !it->second.taken) break;
      s->pending.pop_front();
    }
    if (s->pending.empty()) return SBI_OK;
    const int32_t ipi = s->reqs.at(s->pending.front()).ip;
    std::deque<int64_t>& q = s->by_ip[ipi];
    const int64_t cap = s->ips[ipi].cap;
    std::vector<Piece> pieces;
    int64_t pts = 0;
    while (!q.empty() && pts < cap && (int)pieces.size() < kMaxBatchReqs) {
      Req& r = s->reqs.at(q.front());
      const int64_t rest = r.P - r.pos;
      const int64_t n = rest < cap - pts ? rest : cap - pts;
      if (!pieces.empty() && (r.pos > 0 || n < r.P)) break;  // joins a batch whole, or starts one
      pieces.push_back({q.front(), r.pos, n});
      r.pos += n;
      r.parts += 1;
      pts += n;
      if (r.pos == r.P) {
        r.taken = true;
        q.pop_front();
        s->n_pending -= 1;
      }
    }
    if (pieces.empty()) return SBI_OK;
    const int rc = launch(s, sl, ipi, pieces, pts);
    if (rc != SBI_OK) return rc;
  }
  return SBI_OK;
}

void free_slot_buffers(sbi_server* s) {
  for (Slot& sl : s->slots) {
    if (sl.buf) (void)hipFree(sl.buf);
    sl.buf = nullptr;
    sl.buf_bytes = 0;
  }
}

}  // namespace

extern "C" int sbi_server_create(int32_t n_streams, sbi_server** out) {
  if (!out || n_streams < 1 || n_streams > 16)
    return set_invalid("sbi_server_create: invalid argument (1 <= n_streams <= 16)");
  sbi_server* s = new sbi_server();
  s->slots.resize(n_streams);
  for (Slot& sl : s->slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      sbi::set_error(std::string("sbi_server_create: ") + hipGetErrorString(e));
      sbi_server_destroy(s);
      return SBI_ERR_HIP;
    }
  }
  *out = s;
  return SBI_OK;
}

extern "C" int sbi_server_add(sbi_server* s, const sbi_newton_plan* plan, int64_t cap_points,
                              int32_t* id) {
  const sbi_newton_desc* d = nullptr;
  if (!s || !id || sbi_newton_plan_desc(plan, &d) != SBI_OK || cap_points < 1 ||
      cap_points > (int64_t)65535 * 64)
    return set_invalid("sbi_server_add: invalid argument (1 <= cap_points <= 65535 * 64)");
  if (s->broken) return set_invalid("sbi_server_add: the server failed earlier; destroy it");
  if (s->n_pending > 0 || !s->reqs.empty())
    return set_invalid("sbi_server_add: interpolants are added before requests are submitted");
  Interp ip{plan, d->d_in, d->d_out, cap_points, d->d_out == 1, d->gemm_head + d->gemm_kq > 0};
  if (!ip.stream_ok && !ip.gemm_ok)
    return set_invalid("sbi_server_add: the plan has neither a streaming kernel (d_out = 1) nor "
                       "the mixed GEMM fields");
  // per-slot buffers: staging for batches of several requests, and the GEMM workspace
  const int64_t xs = align256(cap_points * ip.d_in * 8), ys = align256(cap_points * ip.d_out * 8);
  int64_t ws = 0;
  if (ip.gemm_ok) {  // the largest over batch sizes (small batches split k: partial tiles)
    for (int64_t q = 64;; q += 64) {
      const int64_t w = sbi_newton_eval_mixed_gemm_workspace(plan, q < cap_points ? q : cap_points);
      if (w < 0) return SBI_ERR_INVALID;
      if (w > ws) ws = w;
      if (q >= cap_points) break;
    }
  }
  const int64_t nx = xs > s->xs_bytes ? xs : s->xs_bytes, ny = ys > s->ys_bytes ? ys : s->ys_bytes,
                nw = align256(ws) > s->ws_bytes ? align256(ws) : s->ws_bytes;
  if (nx + ny + nw > s->xs_bytes + s->ys_bytes + s->ws_bytes) {
    free_slot_buffers(s);
    for (Slot& sl : s->slots) {
      const hipError_t e = hipMalloc((void**)&sl.buf, (size_t)(nx + ny + nw));
      if (e != hipSuccess) {
        sl.buf = nullptr;
        free_slot_buffers(s);
        s->xs_bytes = s->ys_bytes = s->ws_bytes = 0;
        s->broken = true;
        sbi::set_error(std::string("sbi_server_add: staging buffers: ") + hipGetErrorString(e));
        return SBI_ERR_HIP;
      }
      sl.buf_bytes = nx + ny + nw;
    }
  }
  s->xs_bytes = nx;
  s->ys_bytes = ny;
  s->ws_bytes = nw;
  *id = (int32_t)s->ips.size();
  s->ips.push_back(ip);
  s->by_ip.emplace_back();
  return SBI_OK;
}

extern "C" int sbi_server_submit(sbi_server* s, int32_t id, const double* x, int64_t P,
                                 int64_t ldx, double* y, int64_t ldy, void* stream, int64_t* rid) {
  if (!s || !rid || id < 0 || id >= (int32_t)s->ips.size() || !x || !y || P < 1)
    return set_invalid("sbi_server_submit: invalid argument");
  if (s->broken) return set_invalid("sbi_server_submit: the server failed earlier; destroy it");
  const Interp& ip = s->ips[id];
  if (ldx < ip.d_in || ldy < ip.d_out)
    return set_invalid("sbi_server_submit: row strides below d_in / d_out");
  Req r;
  r.ip = id;
  r.x = x;
  r.P = P;
  r.ldx = ldx;
  r.y = y;
  r.ldy = ldy;
  if (!s->ev_pool.empty()) {
    r.ready = s->ev_pool.back();
    s->ev_pool.pop_back();
  } else {
    SBI_HIP_CHECK(hipEventCreateWithFlags(&r.ready, hipEventDisableTiming));
  }
  const hipError_t er = hipEventRecord(r.ready, (hipStream_t)stream);
  if (er != hipSuccess) {
    s->ev_pool.push_back(r.ready);
    sbi::set_error(std::string("sbi_server_submit: ") + hipGetErrorString(er));
    return SBI_ERR_HIP;
  }
  const int64_t k = s->next++;
  try {
    s->reqs.emplace(k, r);
    s->pending.push_back(k);
    s->by_ip[id].push_back(k);
  } catch (const std::exception& e) {
    sbi::set_error(std::string("sbi_server_submit: ") + e.what());
    return SBI_ERR_INVALID;
  }
  s->n_pending += 1;
  *rid = k;
  return SBI_OK;
}

static int server_step(sbi_server* s, int64_t* done, int32_t max_done, int32_t* n_done,
                       int32_t* busy);

extern "C" int sbi_server_step(sbi_server* s, int64_t* done, int32_t max_done, int32_t* n_done,
                               int32_t* busy) {
  if (!s || !n_done || !busy || max_done < kMaxBatchReqs || !done)
    return set_invalid("sbi_server_step: invalid argument (done[] holds at least 64 entries)");
  if (s->broken) return set_invalid("sbi_server_step: the server failed earlier; destroy it");
  int rc;
  try {
    rc = server_step(s, done, max_done, n_done, busy);
  } catch (const std::exception& e) {
    sbi::set_error(std::string("sbi_server_step: ") + e.what());
    rc = SBI_ERR_INVALID;
  }
  if (rc != SBI_OK) s->broken = true;
  return rc;
}

static int server_step(sbi_server* s, int64_t* done, int32_t max_done, int32_t* n_done,
                       int32_t* busy) {
  *n_done = 0;
  for (Slot& sl : s->slots) {
    if (!sl.busy) continue;
    // a slot's completions are reported together: keep it busy while the buffer is too small
    int32_t fin = 0;
    for (const Piece& pc : sl.pieces) {
      const auto it = s->reqs.find(pc.rid);
      if (it != s->reqs.end() && it->second.taken && it->second.parts == 1) ++fin;
    }
    if (*n_done + fin > max_done) continue;
    const hipError_t e = hipEventQuery(sl.done);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) {
      sbi::set_error(std::string("sbi_server_step: ") + hipGetErrorString(e));
      return SBI_ERR_HIP;
    }
    for (const Piece& pc : sl.pieces) {
      auto it = s->reqs.find(pc.rid);
      if (it == s->reqs.end()) continue;  // (every piece's request is live until its last piece)
      Req& r = it->second;
      r.parts -= 1;
      if (r.taken && r.parts == 0) {
        done[(*n_done)++] = pc.rid;
        s->reqs.erase(it);
      }
    }
    sl.busy = false;
    sl.pieces.clear();
  }
  const int rc = admit(s);
  bool any = s->n_pending > 0;
  for (const Slot& sl : s->slots) any = any || sl.busy;
  *busy = any ? 1 : 0;
  return rc;
}

extern "C" int sbi_server_destroy(sbi_server* s) {
  if (!s) return set_invalid("sbi_server_destroy: null server");
  for (Slot& sl : s->slots)
    if (sl.st) (void)hipStreamSynchronize(sl.st);
  free_slot_buffers(s);
  for (Slot& sl : s->slots) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.st) (void)hipStreamDestroy(sl.st);
  }
  for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
  for (auto& kv : s->reqs)
    if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
  delete s;
  return SBI_OK;
}
