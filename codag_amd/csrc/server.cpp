// server.cpp -- continuous-batching decode server over the C ABI.
//
// Requests (RLE v1 / RLE v2 / Deflate containers in host or device memory)
// enter a queue. Each of max_inflight worker slots owns a codag context (its
// own HIP stream, default priority) and stream-ordered device buffers from
// the server's memory pool. A free slot admits the first queued request
// whose device footprint fits the remaining HBM budget, so requests that
// arrive while others run start as soon as a slot and budget free up and
// their kernels overlap on the device. A request that does not fit may be
// passed by later, smaller ones at most kMaxPass times; after that nothing
// passes it until it is admitted (the aging bound). Latency is measured from
// submit to the result being in host memory (device-resident requests: to
// the decode and its status check being complete).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "codag.h"
#include "internal.hpp"

using namespace codag;

namespace {

// A request blocked on the budget lets at most this many later requests be
// admitted ahead of it.
constexpr uint32_t kMaxPass = 8;

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pending {
  uint64_t id;
  codag_request req;
  double t_submit;
  uint64_t bytes;       // device footprint charged against the budget
  uint32_t passed = 0;  // later requests admitted ahead of this one
};

// A stream-ordered device buffer of one worker slot. It is taken from the
// server's memory pool on the slot's stream and returned to it when a larger
// one is needed, so a slot never calls hipMalloc/hipFree (which synchronize
// the device and would stall the other slots) once the pool holds its sizes.
struct Grow {
  void *p = nullptr;
  uint64_t cap = 0;
  hipMemPool_t pool = nullptr;
  hipStream_t st = nullptr;
  hipError_t ensure(uint64_t n) {
    if (n <= cap && p) return hipSuccess;
    if (p) (void)hipFreeAsync(p, st);
    p = nullptr;
    cap = 0;
    const uint64_t want = n ? n : 1;
    hipError_t e = pool ? hipMallocFromPoolAsync(&p, want, pool, st) : hipMallocAsync(&p, want, st);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFreeAsync(p, st);
    p = nullptr;
    cap = 0;
  }
};

uint64_t footprint(const codag_request &r) {
  const uint64_t meta = r.codec == CODAG_DEFLATE ? (r.n_chunks + 1) * 8 : r.n_chunks * 16;
  return r.comp_bytes + meta + r.n_chunks * r.chunk_bytes + r.n_chunks * 4;
}

}  // namespace

struct codag_server {
  codag_server_config cfg{};
  std::vector<codag_ctx *> ctxs;  // one per worker slot
  hipMemPool_t pool = nullptr;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Pending> queue;
  std::deque<codag_completion> done;
  uint64_t admitted = 0;
  uint32_t inflight = 0;
  bool stop = false;
  uint64_t next_id = 1;
  std::vector<std::thread> workers;

  // Index of the queued request to admit next, or -1: the first one in
  // queue order that fits the budget, not passing a request that has been
  // passed kMaxPass times. Called with mu held.
  long admissible() const {
    for (size_t i = 0; i < queue.size(); i++) {
      if (admitted + queue[i].bytes <= cfg.hbm_budget_bytes) return (long)i;
      if (queue[i].passed >= kMaxPass) return -1;
    }
    return -1;
  }
};

static codag_status run_one(codag_ctx *ctx, hipStream_t st, Grow *b, const Pending &p,
                            uint32_t *failed) {
  const codag_request &r = p.req;
  const uint64_t n = r.n_chunks;
  const uint64_t out_bytes = n * r.chunk_bytes;
  const uint64_t noff = r.codec == CODAG_DEFLATE ? n + 1 : n;
  if (r.flags & CODAG_REQ_DEVICE_IO) {  // data already in HBM: decode in place
    hipError_t e = b[4].ensure(n * 4);
    if (e != hipSuccess) return fail(CODAG_ERR_NOMEM, std::string("server: ") + hipGetErrorString(e));
    int32_t *stat = static_cast<int32_t *>(b[4].p);
    codag_status cs;
    if (r.codec == CODAG_RLEV1)
      cs = codag_rlev1_decode(ctx, r.comp, r.comp_bytes, r.blk_off, r.col_len, n, r.elem_bytes,
                              r.chunk_bytes, r.out, stat, st);
    else if (r.codec == CODAG_RLEV2)
      cs = codag_rlev2_decode(ctx, r.comp, r.comp_bytes, r.blk_off, r.col_len, n, r.elem_bytes,
                              r.chunk_bytes, r.out, stat, st);
    else
      cs = codag_inflate_chunks(ctx, r.comp, r.comp_bytes, r.blk_off, n, r.chunk_bytes, r.out,
                                nullptr, stat, 0, st);
    if (cs) return cs;
    std::vector<int32_t> hs(n);
    if (n) e = hipMemcpyAsync(hs.data(), stat, n * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(CODAG_ERR_HIP, std::string("server: ") + hipGetErrorString(e));
    uint32_t bad = 0;
    for (auto v : hs) bad += (v != 0);
    *failed = bad;
    return bad ? CODAG_ERR_DECODE : CODAG_OK;
  }
  hipError_t e = b[0].ensure(r.comp_bytes);
  if (e == hipSuccess) e = b[1].ensure(noff * 8);
  if (e == hipSuccess) e = b[2].ensure(n * 8);
  if (e == hipSuccess) e = b[3].ensure(out_bytes);
  if (e == hipSuccess) e = b[4].ensure(n * 4);
  if (e != hipSuccess) return fail(CODAG_ERR_NOMEM, std::string("server: ") + hipGetErrorString(e));
  if (r.comp_bytes) e = hipMemcpyAsync(b[0].p, r.comp, r.comp_bytes, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && noff)
    e = hipMemcpyAsync(b[1].p, r.blk_off, noff * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && r.codec != CODAG_DEFLATE && n)
    e = hipMemcpyAsync(b[2].p, r.col_len, n * 8, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return fail(CODAG_ERR_HIP, std::string("server H2D: ") + hipGetErrorString(e));
  codag_status cs;
  uint8_t *comp = static_cast<uint8_t *>(b[0].p);
  uint64_t *off = static_cast<uint64_t *>(b[1].p);
  uint64_t *len = static_cast<uint64_t *>(b[2].p);
  uint8_t *out = static_cast<uint8_t *>(b[3].p);
  int32_t *stat = static_cast<int32_t *>(b[4].p);
  if (r.codec == CODAG_RLEV1)
    cs = codag_rlev1_decode(ctx, comp, r.comp_bytes, off, len, n, r.elem_bytes, r.chunk_bytes,
                            out, stat, st);
  else if (r.codec == CODAG_RLEV2)
    cs = codag_rlev2_decode(ctx, comp, r.comp_bytes, off, len, n, r.elem_bytes, r.chunk_bytes,
                            out, stat, st);
  else
    cs = codag_inflate_chunks(ctx, comp, r.comp_bytes, off, n, r.chunk_bytes, out, nullptr,
                              stat, 0, st);
  if (cs) return cs;
  std::vector<int32_t> hs(n);
  if (out_bytes) e = hipMemcpyAsync(r.out, out, out_bytes, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && n) e = hipMemcpyAsync(hs.data(), stat, n * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fail(CODAG_ERR_HIP, std::string("server D2H: ") + hipGetErrorString(e));
  uint32_t bad = 0;
  for (auto v : hs) bad += (v != 0);
  *failed = bad;
  return bad ? CODAG_ERR_DECODE : CODAG_OK;
}

static void worker(codag_server *s, uint32_t slot) {
  codag_ctx *ctx = s->ctxs[slot];
  (void)hipSetDevice(s->cfg.device);
  hipStream_t st = static_cast<hipStream_t>(codag_ctx_stream(ctx));
  Grow bufs[5];
  for (auto &g : bufs) {
    g.pool = s->pool;
    g.st = st;
  }
  for (;;) {
    std::unique_lock<std::mutex> lk(s->mu);
    long k = -1;
    s->cv_work.wait(lk, [&] { return s->stop || (k = s->admissible()) >= 0; });
    if (s->stop) break;
    for (long i = 0; i < k; i++) s->queue[i].passed++;
    Pending p = s->queue[k];
    s->queue.erase(s->queue.begin() + k);
    s->admitted += p.bytes;
    s->inflight++;
    lk.unlock();
    const double t_adm = now_s();
    uint32_t failed = 0;
    codag_status cs = run_one(ctx, st, bufs, p, &failed);
    const double t_end = now_s();
    lk.lock();
    s->admitted -= p.bytes;
    s->inflight--;
    codag_completion c{};
    c.request_id = p.id;
    c.user_tag = p.req.user_tag;
    c.status = cs;
    c.failed_chunks = failed;
    c.queue_seconds = t_adm - p.t_submit;
    c.latency_seconds = t_end - p.t_submit;
    s->done.push_back(c);
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
  for (auto &g : bufs) g.release();
  (void)hipStreamSynchronize(st);
}

extern "C" {

codag_status codag_server_create(const codag_server_config *cfg, codag_server **out) {
  if (!cfg || !out) return fail(CODAG_ERR_INVALID, "server_create: null argument");
  if (cfg->max_inflight == 0 || cfg->max_inflight > 64)
    return fail(CODAG_ERR_INVALID, "server_create: max_inflight must be 1..64");
  if (cfg->hbm_budget_bytes == 0) return fail(CODAG_ERR_INVALID, "server_create: zero budget");
  auto *s = new (std::nothrow) codag_server();
  if (!s) return fail(CODAG_ERR_NOMEM, "server_create");
  s->cfg = *cfg;
  for (uint32_t i = 0; i < cfg->max_inflight; i++) {
    codag_ctx *c = nullptr;
    codag_status cs = codag_ctx_create(cfg->device, &c);
    if (cs) {
      for (auto *x : s->ctxs) codag_ctx_destroy(x);
      delete s;
      return cs;
    }
    s->ctxs.push_back(c);
  }
  {
    // The server's own stream-ordered pool: freed blocks stay in it (release
    // threshold: never), so steady-state admissions allocate nothing.
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(cfg->device);
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = cfg->device;
    if (hipMemPoolCreate(&s->pool, &props) == hipSuccess) {
      uint64_t thr = ~0ull;
      (void)hipMemPoolSetAttribute(s->pool, hipMemPoolAttrReleaseThreshold, &thr);
    } else {
      (void)hipGetLastError();
      s->pool = nullptr;  // fall back to the device's default pool
    }
    (void)hipSetDevice(prev);
  }
  for (uint32_t i = 0; i < cfg->max_inflight; i++) s->workers.emplace_back(worker, s, i);
  *out = s;
  return CODAG_OK;
}

codag_status codag_server_submit(codag_server *s, const codag_request *req, uint64_t *id) {
  if (!s || !req) return fail(CODAG_ERR_INVALID, "server_submit: null argument");
  if (req->codec != CODAG_RLEV1 && req->codec != CODAG_RLEV2 && req->codec != CODAG_DEFLATE)
    return fail(CODAG_ERR_INVALID, "server_submit: unknown codec");
  if (req->n_chunks && (!req->blk_off || !req->out || (!req->comp && req->comp_bytes)))
    return fail(CODAG_ERR_INVALID, "server_submit: null buffer");
  if (req->codec != CODAG_DEFLATE && req->n_chunks && !req->col_len)
    return fail(CODAG_ERR_INVALID, "server_submit: null col_len");
  // validate host metadata so a malformed request never reaches the GPU
  // (device-resident metadata is bounded by comp_bytes inside the kernels)
  const bool host_meta = !(req->flags & CODAG_REQ_DEVICE_IO);
  for (uint64_t i = 0; host_meta && i < req->n_chunks; i++) {
    if (req->codec == CODAG_DEFLATE) {
      if (req->blk_off[i] > req->blk_off[i + 1] || req->blk_off[i + 1] > req->comp_bytes)
        return fail(CODAG_ERR_FORMAT, "server_submit: chunk outside the stream");
    } else if (req->blk_off[i] > req->comp_bytes ||
               req->col_len[i] > req->comp_bytes - req->blk_off[i]) {
      return fail(CODAG_ERR_FORMAT, "server_submit: chunk outside the stream");
    }
  }
  Pending p{0, *req, now_s(), footprint(*req)};
  if (p.bytes > s->cfg.hbm_budget_bytes)
    return fail(CODAG_ERR_INVALID, "server_submit: request larger than the HBM budget");
  std::lock_guard<std::mutex> lk(s->mu);
  p.id = s->next_id++;
  if (id) *id = p.id;
  s->queue.push_back(p);
  s->cv_work.notify_all();
  return CODAG_OK;
}

codag_status codag_server_poll(codag_server *s, codag_completion *out, uint32_t max,
                               uint32_t *n_out, int timeout_ms) {
  if (!s || !n_out || (max && !out)) return fail(CODAG_ERR_INVALID, "server_poll: bad argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] { return !s->done.empty(); };
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
  uint32_t k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  *n_out = k;
  return k ? CODAG_OK : CODAG_ERR_TIMEOUT;
}

codag_status codag_server_stats(codag_server *s, uint64_t *admitted, uint32_t *inflight,
                                uint32_t *queued) {
  if (!s) return fail(CODAG_ERR_INVALID, "server_stats: null server");
  std::lock_guard<std::mutex> lk(s->mu);
  if (admitted) *admitted = s->admitted;
  if (inflight) *inflight = s->inflight;
  if (queued) *queued = (uint32_t)s->queue.size();
  return CODAG_OK;
}

codag_status codag_server_destroy(codag_server *s) {
  if (!s) return CODAG_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto *c : s->ctxs) codag_ctx_destroy(c);
  if (s->pool) (void)hipMemPoolDestroy(s->pool);
  delete s;
  return CODAG_OK;
}

}  // extern "C"
