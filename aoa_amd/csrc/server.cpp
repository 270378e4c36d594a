// Serving loop: continuous batching of encrypted-op requests on one GPU.
//
// Requests (a ciphertext plus a chain of HomMul+relin+rescale and rotation
// ops) are admitted while the HBM pool can hold their two working
// ciphertexts; evaluation and rotation keys stay resident, registered once per
// parameter set.  Each step takes the next op of every active request whose
// workspace fits the step budget, plans one stage-fused batch for all of them
// (a forward-NTT layer for multiplications whose operand is in the coefficient
// domain, then the op layer) and enqueues it.  The serving thread plans and
// enqueues step t+1 while step t runs and retires step t's finished requests
// after that, so host planning overlaps device work.  Requests submitted
// while a step runs join at the next step boundary.
// No counterpart in the reference, which evaluates one call at a time on one
// stream (aoa-ntt/include/newckks/tool/context.h:16-49).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "internal.hpp"

using namespace aoa;

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device region, 256-byte granules.
class Pool {
 public:
  void init(char* base, size_t bytes) {
    base_ = base;
    size_ = bytes;
    free_.clear();
    free_[0] = bytes;
    used_ = 0;
  }
  char* alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_ += bytes;
        sizes_[off] = bytes;
        return base_ + off;
      }
    }
    return nullptr;
  }
  void release(char* p) {
    if (!p) return;
    size_t off = (size_t)(p - base_);
    auto s = sizes_.find(off);
    if (s == sizes_.end()) return;
    size_t len = s->second;
    sizes_.erase(s);
    used_ -= len;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
      len += next->second;
      free_.erase(next);
    }
    auto prev = free_.lower_bound(off);
    if (prev != free_.begin()) {
      --prev;
      if (prev->first + prev->second == off) {
        prev->second += len;
        return;
      }
    }
    free_[off] = len;
  }
  size_t used() const { return used_; }
  size_t size() const { return size_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_, sizes_;
};

struct KeySet {
  const aoa_params* params;
  const uint64_t* evk_b;
  const uint64_t* evk_a;
  std::map<int, std::pair<const uint64_t*, const uint64_t*>> rtk;
  size_t key_bytes;
};

enum { PENDING = 0, ACTIVE = 1, DONE = 2, FAILED = 3 };

struct Request {
  int64_t id = 0;
  const KeySet* ks = nullptr;  // taken under the lock at submit; deque elements never move
  const uint64_t* in[2] = {nullptr, nullptr};
  const uint64_t* m[2] = {nullptr, nullptr};
  uint64_t* out[2] = {nullptr, nullptr};
  std::vector<aoa_op> ops;
  size_t next = 0;  // next op to plan
  int level = 0;    // level after the ops planned so far
  bool ntt = false; // domain after the ops planned so far
  char* mem = nullptr;  // pool block: two working ciphertexts
  uint64_t* buf[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
  int cur = 0;
  double t_submit = 0, t_done = 0;
  int state = PENDING;
};

// What one enqueued step will finish, and whether its enqueue failed.
struct Step {
  std::vector<Request*> finish;  // requests whose last op (or final copy) is in this step
  std::vector<Request*> failed;  // requests whose step could not be enqueued
  int ops = 0;
  int rc = AOA_OK;
  bool enqueued = false;
};

}  // namespace

// Finished requests kept for aoa_server_result before the oldest are dropped.
constexpr size_t kRetainFinished = (size_t)1 << 16;

struct aoa_server {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<KeySet> keysets;  // push_back keeps references to existing elements valid
  std::deque<Request*> pending;
  std::vector<Request*> active;
  std::map<int64_t, Request*> all;
  std::deque<int64_t> finished;  // ids in completion order, for bounded retention
  int64_t next_id = 1;
  int max_active = 64;
  Pool pool;
  void* pool_mem = nullptr;
  void* step_ws = nullptr;
  size_t step_ws_bytes = 0;
  void* desc_dev = nullptr;
  char* desc_host[2] = {nullptr, nullptr};  // one pinned staging buffer per step in flight
  size_t desc_cap = 0;
  aoa_batch* plan = nullptr;
  hipStream_t own_stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::thread worker;
  bool running = false, stop = false;
  int last_rc = AOA_OK;
  std::string last_msg;
  // statistics
  int64_t completed = 0, ops_done = 0, steps = 0;
  double first_submit = -1, last_done = 0;
};

static size_t ct_bytes(const KeySet& ks, int level) {
  return 2 * (size_t)(level + 1) * ks.params->host.N * sizeof(uint64_t);
}

static size_t op_ws_bytes(const KeySet& ks, const aoa_op& op, int level) {
  const size_t N = ks.params->host.N, K = ks.params->host.K, L = level + 1;
  const size_t limbs = op.kind == AOA_OP_MUL ? 6 * L + 3 * K : 5 * L + 3 * K;
  return limbs * N * 8 + 256;
}

extern "C" int aoa_server_create(size_t pool_bytes, size_t step_ws_bytes, int max_active, aoa_server** out) {
  if (!out || pool_bytes == 0 || step_ws_bytes == 0 || max_active <= 0) return set_error(AOA_ERR_ARG, "bad argument");
  aoa_server* s = new aoa_server();
  s->max_active = max_active;
  s->step_ws_bytes = step_ws_bytes;
  s->desc_cap = (size_t)64 << 20;
  if (hipMalloc(&s->pool_mem, pool_bytes) != hipSuccess || hipMalloc(&s->step_ws, step_ws_bytes) != hipSuccess ||
      hipMalloc(&s->desc_dev, s->desc_cap) != hipSuccess ||
      hipHostMalloc((void**)&s->desc_host[0], s->desc_cap, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->desc_host[1], s->desc_cap, hipHostMallocDefault) != hipSuccess ||
      hipStreamCreateWithFlags(&s->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&s->ev[1], hipEventDisableTiming) != hipSuccess) {
    if (s->pool_mem) (void)hipFree(s->pool_mem);
    if (s->step_ws) (void)hipFree(s->step_ws);
    if (s->desc_dev) (void)hipFree(s->desc_dev);
    for (char* h : s->desc_host)
      if (h) (void)hipHostFree(h);
    if (s->own_stream) (void)hipStreamDestroy(s->own_stream);
    for (hipEvent_t e : s->ev)
      if (e) (void)hipEventDestroy(e);
    delete s;
    return set_error(AOA_ERR_HIP, "server allocation failed");
  }
  s->pool.init((char*)s->pool_mem, pool_bytes);
  aoa_batch_create(&s->plan);
  *out = s;
  return AOA_OK;
}

extern "C" int aoa_server_register(aoa_server* s, const aoa_params* p, const uint64_t* evk_b, const uint64_t* evk_a,
                                   int nrot, const int* rot_steps, const uint64_t* const* rtk_b,
                                   const uint64_t* const* rtk_a) {
  if (!s || !p || !evk_b || !evk_a || (nrot > 0 && (!rot_steps || !rtk_b || !rtk_a)))
    return -set_error(AOA_ERR_ARG, "null argument");
  KeySet ks;
  ks.params = p;
  ks.evk_b = evk_b;
  ks.evk_a = evk_a;
  const size_t key = 2 * (size_t)(p->host.k + p->host.K) * p->host.N * 8;
  ks.key_bytes = key * (1 + nrot);
  for (int i = 0; i < nrot; ++i) {
    if (!rtk_b[i] || !rtk_a[i]) return -set_error(AOA_ERR_ARG, "null rotation key");
    ks.rtk[rot_steps[i]] = {rtk_b[i], rtk_a[i]};
  }
  std::lock_guard<std::mutex> g(s->mu);
  s->keysets.push_back(ks);
  return (int)s->keysets.size() - 1;
}

extern "C" int aoa_server_submit(aoa_server* s, int keyset, const uint64_t* c0, const uint64_t* c1, int level,
                                 int in_ntt, const uint64_t* m0, const uint64_t* m1, const aoa_op* ops, int nops,
                                 uint64_t* out0, uint64_t* out1, int64_t* req_id) {
  if (!s || !c0 || !c1 || !out0 || !out1 || !req_id || nops < 0 || (nops && !ops))
    return set_error(AOA_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  if (keyset < 0 || keyset >= (int)s->keysets.size()) return set_error(AOA_ERR_ARG, "unknown key set");
  const KeySet& ks = s->keysets[keyset];
  if (level < 0 || level >= ks.params->host.k) return set_error(AOA_ERR_ARG, "level out of range");
  int lv = level;
  for (int i = 0; i < nops; ++i) {
    if (ops[i].kind == AOA_OP_MUL) {
      if (!m0 || !m1) return set_error(AOA_ERR_ARG, "multiplication without operand");
      if (lv == 0) return set_error(AOA_ERR_LEVEL, "The ciphertext level is 0. No more rescaling is possible.");
      --lv;
    } else if (ops[i].kind == AOA_OP_ROTATE) {
      if (!ks.rtk.count(ops[i].rot_slots)) return set_error(AOA_ERR_ARG, "no rotation key for this step");
    } else {
      return set_error(AOA_ERR_ARG, "unknown op");
    }
  }
  // a request that can never fit is rejected now rather than starved
  size_t need_ws = 0;
  lv = level;
  for (int i = 0; i < nops; ++i) {
    need_ws = std::max(need_ws, op_ws_bytes(ks, ops[i], lv));
    if (ops[i].kind == AOA_OP_MUL) --lv;
  }
  if (2 * ct_bytes(ks, level) + 1024 > s->pool.size() || need_ws > s->step_ws_bytes)
    return set_error(AOA_ERR_WORKSPACE, "request exceeds the HBM budget");
  Request* r = new Request();
  r->id = s->next_id++;
  r->ks = &ks;
  r->in[0] = c0;
  r->in[1] = c1;
  r->m[0] = m0;
  r->m[1] = m1;
  r->out[0] = out0;
  r->out[1] = out1;
  r->ops.assign(ops, ops + nops);
  r->level = level;
  r->ntt = in_ntt != 0;
  r->t_submit = now_ms();
  if (s->first_submit < 0) s->first_submit = r->t_submit;
  s->pending.push_back(r);
  s->all[r->id] = r;
  *req_id = r->id;
  s->cv.notify_all();
  return AOA_OK;
}

// Plans one step and enqueues it on `st`: admits pending requests while the
// pool has room (copying their inputs in), takes the next op of every active
// request whose workspace fits the step budget, and appends the final copy of
// every request whose chain ends with this step.  The requests' host state
// (next op, level, domain, current buffer) advances at planning time: the
// device runs the steps in stream order, so step t+1 may be planned and
// enqueued while step t still runs, and a pool block released here is reused
// only by work enqueued after this step's final copies.  The descriptors are
// staged in desc_host[slot], which the caller keeps untouched until this
// step's event has completed.
static Step enqueue_step(aoa_server* s, hipStream_t st, int slot) {
  Step out;
  std::vector<Request*> run;
  {
    std::lock_guard<std::mutex> g(s->mu);
    while (!s->pending.empty() && (int)s->active.size() < s->max_active) {
      Request* r = s->pending.front();
      const size_t cb = ct_bytes(*r->ks, r->level);
      char* mem = s->pool.alloc(2 * cb + 1024);
      if (!mem) break;
      s->pending.pop_front();
      r->mem = mem;
      const size_t pb = cb / 2;  // one polynomial
      const size_t pb_al = (pb + 255) & ~(size_t)255;
      r->buf[0][0] = (uint64_t*)mem;
      r->buf[0][1] = (uint64_t*)(mem + pb_al);
      r->buf[1][0] = (uint64_t*)(mem + 2 * pb_al);
      r->buf[1][1] = (uint64_t*)(mem + 3 * pb_al);
      r->cur = 0;
      r->state = ACTIVE;
      for (int y = 0; y < 2; ++y)
        (void)hipMemcpyAsync(r->buf[0][y], r->in[y], pb, hipMemcpyDeviceToDevice, st);
      s->active.push_back(r);
    }
    size_t ws = 0;
    for (Request* r : s->active) {
      if (r->next >= r->ops.size()) {
        out.finish.push_back(r);  // no ops: only the copies
        continue;
      }
      const size_t need = op_ws_bytes(*r->ks, r->ops[r->next], r->level);
      if (ws + need > s->step_ws_bytes) continue;  // waits for the next step
      ws += need;
      run.push_back(r);
    }
  }
  int rc = AOA_OK;
  if (!run.empty()) {
    aoa_batch* b = s->plan;
    batch_clear(b);
    // layer 1: bring multiplication operands into the NTT domain, rotation
    // inputs into the coefficient domain
    bool any = false;
    for (Request* r : run) {
      const int kind = r->ops[r->next].kind;
      if ((kind == AOA_OP_MUL) != r->ntt) {
        aoa_ntt_item it = {{r->buf[r->cur][0], r->buf[r->cur][1]}, r->level,
                           kind == AOA_OP_MUL ? AOA_FORWARD : AOA_INVERSE};
        batch_add_ntt(b, r->ks->params, &it, 1);
        any = true;
      }
    }
    if (any) batch_barrier(b);
    // layer 2: the ops
    for (Request* r : run) {
      const KeySet& ks = *r->ks;
      const aoa_op& op = r->ops[r->next];
      uint64_t* const* src = r->buf[r->cur];
      uint64_t* const* dst = r->buf[1 - r->cur];
      if (op.kind == AOA_OP_MUL) {
        aoa_mul_item it;
        it.c1[0] = src[0];
        it.c1[1] = src[1];
        it.c2[0] = r->m[0];
        it.c2[1] = r->m[1];
        it.out[0] = dst[0];
        it.out[1] = dst[1];
        it.level = r->level;
        rc = batch_add_mul(b, ks.params, ks.evk_b, ks.evk_a, &it, 1);
      } else {
        auto key = ks.rtk.at(op.rot_slots);
        aoa_rot_item it;
        it.in[0] = src[0];
        it.in[1] = src[1];
        it.out[0] = dst[0];
        it.out[1] = dst[1];
        it.rtk_b = key.first;
        it.rtk_a = key.second;
        it.rot_slots = op.rot_slots;
        it.level = r->level;
        it.accumulate = 0;
        rc = batch_add_rotate(b, ks.params, &it, 1);
      }
      if (rc) break;
    }
    size_t wsb = 0, descb = 0;
    if (!rc) {
      batch_sizes(b, &wsb, &descb);
      if (wsb > s->step_ws_bytes || descb > s->desc_cap) rc = set_error(AOA_ERR_WORKSPACE, "step exceeds its budget");
    }
    if (!rc) {
      size_t used = batch_fill(b, (uint64_t*)s->step_ws, s->desc_host[slot], (char*)s->desc_dev, s->desc_cap);
      if (!used) rc = set_error(AOA_ERR_WORKSPACE, "descriptor buffer too small");
      else if (hipMemcpyAsync(s->desc_dev, s->desc_host[slot], used, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = set_error(AOA_ERR_HIP, "descriptor upload failed");
    }
    if (!rc) rc = batch_enqueue(b, st);
  }
  // advance the host state of the planned requests
  std::vector<Request*> ended;
  for (Request* r : run) {
    if (rc) {
      out.failed.push_back(r);
      continue;
    }
    const aoa_op& op = r->ops[r->next];
    r->cur = 1 - r->cur;
    if (op.kind == AOA_OP_MUL) --r->level;
    r->ntt = false;
    ++r->next;
    ++out.ops;
    if (r->next >= r->ops.size()) out.finish.push_back(r);
  }
  // final copies to the caller's buffers (coefficient domain at the final level)
  for (Request* r : out.finish) {
    const size_t pb = (size_t)(r->level + 1) * r->ks->params->host.N * 8;
    for (int y = 0; y < 2; ++y) (void)hipMemcpyAsync(r->out[y], r->buf[r->cur][y], pb, hipMemcpyDeviceToDevice, st);
  }
  out.rc = rc;
  out.enqueued = !run.empty() || !out.finish.empty() || !out.failed.empty();
  if (out.enqueued && hipEventRecord(s->ev[slot], st) != hipSuccess && !rc) {
    out.rc = set_error(AOA_ERR_HIP, "step event record failed");
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->steps += run.empty() ? 0 : 1;
    // leave the active set now: later steps never see these requests, and
    // their pool blocks are reused only by work enqueued after this step
    for (auto* list : {&out.finish, &out.failed}) {
      for (Request* r : *list) {
        s->pool.release(r->mem);
        r->mem = nullptr;
        s->active.erase(std::remove(s->active.begin(), s->active.end(), r), s->active.end());
      }
    }
  }
  return out;
}

// Waits for a step's event and reports its requests.
static int retire_step(aoa_server* s, Step& step, int slot) {
  int rc = step.rc;
  if (step.enqueued && hipEventSynchronize(s->ev[slot]) != hipSuccess)
    rc = rc ? rc : set_error(AOA_ERR_HIP, "step synchronisation failed");
  std::lock_guard<std::mutex> g(s->mu);
  const double t = now_ms();
  auto finish = [&](Request* r, int state) {
    r->state = state;
    r->t_done = t;
    s->last_done = t;
    s->finished.push_back(r->id);
    if (state == DONE) ++s->completed;
  };
  for (Request* r : step.failed) finish(r, FAILED);
  for (Request* r : step.finish) finish(r, rc ? FAILED : DONE);
  s->ops_done += rc ? 0 : step.ops;
  // bounded retention: drop the oldest finished records
  while (s->finished.size() > kRetainFinished) {
    auto it = s->all.find(s->finished.front());
    s->finished.pop_front();
    if (it != s->all.end()) {
      delete it->second;
      s->all.erase(it);
    }
  }
  if (rc) {
    s->last_rc = rc;
    s->last_msg = aoa_last_error();
  }
  step = Step();
  return rc;
}

extern "C" int aoa_server_step(aoa_server* s, void* stream) {
  if (!s) return -set_error(AOA_ERR_ARG, "null server");
  if (s->running) return -set_error(AOA_ERR_ARG, "server thread is running");
  Step step = enqueue_step(s, stream ? (hipStream_t)stream : s->own_stream, 0);
  const int rc = retire_step(s, step, 0);
  if (rc) return -rc;
  std::lock_guard<std::mutex> g(s->mu);
  return (int)(s->pending.size() + s->active.size());
}

// The serving thread keeps up to two steps in flight: it enqueues step t,
// then waits for step t-1 and reports its requests while step t runs, so the
// host's planning and descriptor upload overlap the device's work.
extern "C" int aoa_server_start(aoa_server* s) {
  if (!s) return set_error(AOA_ERR_ARG, "null server");
  if (s->running) return AOA_OK;
  s->stop = false;
  s->running = true;
  s->worker = std::thread([s]() {
    Step prev;
    int slot = 0;
    bool have_prev = false;
    while (true) {
      {
        std::unique_lock<std::mutex> g(s->mu);
        if (s->stop) break;
        if (!have_prev && s->pending.empty() && s->active.empty()) {
          s->cv.wait_for(g, std::chrono::milliseconds(1));
          continue;
        }
      }
      // desc_host[slot] was last used by step t-2, whose event has completed
      Step cur = enqueue_step(s, s->own_stream, slot);
      if (have_prev) {
        retire_step(s, prev, 1 - slot);  // failed requests are marked; keep serving the others
        have_prev = false;
      }
      if (cur.enqueued) {
        prev = std::move(cur);
        have_prev = true;
        slot = 1 - slot;
      }
    }
    if (have_prev) retire_step(s, prev, 1 - slot);
  });
  return AOA_OK;
}

extern "C" int aoa_server_stop(aoa_server* s) {
  if (!s) return set_error(AOA_ERR_ARG, "null server");
  if (!s->running) return AOA_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  s->worker.join();
  s->running = false;
  return AOA_OK;
}

extern "C" int aoa_server_result(aoa_server* s, int64_t req_id, int* state, int* level, double* latency_ms) {
  if (!s) return set_error(AOA_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->all.find(req_id);
  if (it == s->all.end()) return set_error(AOA_ERR_ARG, "unknown or released request");
  Request* r = it->second;
  const bool fin = r->state >= DONE;
  if (state) *state = r->state;
  // the level is final only once the request has finished
  if (level) *level = r->level;
  if (latency_ms) *latency_ms = fin ? r->t_done - r->t_submit : -1.0;
  return AOA_OK;
}

extern "C" int aoa_server_release(aoa_server* s, int64_t req_id) {
  if (!s) return set_error(AOA_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->all.find(req_id);
  if (it == s->all.end()) return set_error(AOA_ERR_ARG, "unknown or released request");
  if (it->second->state < DONE) return set_error(AOA_ERR_ARG, "request has not finished");
  delete it->second;
  s->all.erase(it);
  return AOA_OK;
}

extern "C" int aoa_server_stats(aoa_server* s, int64_t* completed, int64_t* ops, int64_t* steps, double* span_ms,
                                size_t* pool_used, size_t* key_bytes) {
  if (!s) return set_error(AOA_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  if (completed) *completed = s->completed;
  if (ops) *ops = s->ops_done;
  if (steps) *steps = s->steps;
  if (span_ms) *span_ms = s->first_submit < 0 ? 0 : s->last_done - s->first_submit;
  if (pool_used) *pool_used = s->pool.used();
  if (key_bytes) {
    size_t kb = 0;
    for (auto& k : s->keysets) kb += k.key_bytes;
    *key_bytes = kb;
  }
  return AOA_OK;
}

extern "C" void aoa_server_destroy(aoa_server* s) {
  if (!s) return;
  aoa_server_stop(s);
  (void)hipStreamSynchronize(s->own_stream);
  for (auto& kv : s->all) delete kv.second;
  if (s->plan) aoa_batch_destroy(s->plan);
  (void)hipFree(s->pool_mem);
  (void)hipFree(s->step_ws);
  (void)hipFree(s->desc_dev);
  for (char* h : s->desc_host) (void)hipHostFree(h);
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(s->own_stream);
  delete s;
}
