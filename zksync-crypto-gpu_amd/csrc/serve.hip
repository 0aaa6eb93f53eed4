// serve.hip -- serve mode: a request loop for coset-LDE jobs with memory-budget admission and
// continuous batching.
//
// No reference counterpart: shivini runs one prover's LDEs in program order
// (crates/shivini/src/data_structures/storage.rs:154-260, cache.rs:526-676).  Here clients
// submit LDE requests from any thread; a scheduler thread owned by the server
//   1. admits queued requests in arrival order while the device arena (one allocation of the
//      memory budget, carved by a first-fit allocator) has room for the request's working
//      buffer (evaluations -> monomials) and its outputs; while the oldest request waits for
//      memory, later requests that fit may pass it, at most max_bypass of them (then the queue
//      waits for the oldest), so a large request neither blocks small ones nor starves,
//   2. turns everything admitted since the last launch into ONE batch plan (columns merged by
//      transform, grouped pass launches) and enqueues it on a free slot stream, so a batch
//      that arrives while others run overlaps with them on the device,
//   3. polls each slot's completion event, returns working buffers to the arena and wakes the
//      waiters; outputs stay resident in HBM until the client releases them.
// The twiddle and coset tables stay resident in the context for the server's lifetime.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "api_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over [0, size) in 256-byte units of a device arena.
class Arena {
 public:
  explicit Arena(size_t size) : size_(size) { free_[0] = size; }
  static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
  bool alloc(size_t bytes, size_t &off) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      off = it->first;
      const size_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round(bytes);
    used_ -= bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + bytes == next->first) {  // merge with the block after
      bytes += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {  // merge with the block before
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        prev->second += bytes;
        return;
      }
    }
    free_[off] = bytes;
  }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t size() const { return size_; }

 private:
  size_t size_, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> length
};

enum class State { Queued, Admitted, Running, Done, Failed };

struct Request {
  ba_serve_request req;
  State state = State::Queued;
  int status = 0;
  size_t work_off = 0, out_off = 0, work_bytes = 0, out_bytes = 0;
  Clock::time_point submit_t, admit_t, launch_t, done_t;
  uint64_t batch = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  void *staging = nullptr;  // pinned host copy of the plan image
  void *desc = nullptr;     // its device copy
  bool busy = false;
  uint64_t batch = 0;
  std::vector<uint64_t> tickets;
};

constexpr size_t kDescBytes = 16u << 20;  // plan image capacity per slot

uint64_t work_elems(const ba_serve_request &r) { return (uint64_t)r.num_cols << r.log_n; }
uint64_t out_elems(const ba_serve_request &r) { return work_elems(r) << r.log_blowup; }

// Size class of a request's output: 0 below 2^26 elements, 1 below 2^30, 2 from 2^30 up.  A batch
// holds one class: merged into one plan with a request ~10^4 times its size, a small request would
// complete only when that one does.
int size_class(const ba_serve_request &r) {
  const uint64_t e = out_elems(r);
  return e < (1ull << 26) ? 0 : (e < (1ull << 30) ? 1 : 2);
}

}  // namespace

struct ba_server {
  ba_context *ctx = nullptr;
  int device = 0;
  void *base = nullptr;
  Arena arena;
  std::vector<Slot> slots;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queued;
  // requests admitted ahead of the current oldest queued request (head_ticket), and the bound
  uint64_t head_ticket = 0;
  uint32_t head_passed = 0, max_bypass = 32;
  std::vector<uint64_t> admitted;
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1, batches = 0, completed = 0, failed = 0, out_total = 0, batched_jobs = 0;
  std::vector<double> latencies;
  bool have_first = false;
  Clock::time_point first_submit, last_done;
  bool stop = false;
  std::thread worker;

  explicit ba_server(size_t budget) : arena(budget) {}

  void loop();
  void finish(Slot &s, int status);
  void fail(Request &r, int status, Clock::time_point now);
  bool launch(Slot &s, std::unique_lock<std::mutex> &lk);
};

void ba_server::finish(Slot &s, int status) {
  const auto now = Clock::now();
  for (uint64_t t : s.tickets) {
    auto it = reqs.find(t);
    if (it == reqs.end()) continue;
    Request &r = it->second;
    arena.release(r.work_off, r.work_bytes);
    r.work_bytes = 0;
    r.done_t = now;
    if (status == 0) {
      r.state = State::Done;
      completed++;
      out_total += out_elems(r.req);
      latencies.push_back(ms_between(r.submit_t, now));
    } else {
      fail(r, status, now);
    }
  }
  last_done = now;
  s.tickets.clear();
  s.busy = false;
  cv_done.notify_all();
}

// The batch-plan job of one admitted request, with its buffers in the arena at `base`.
static ba_job request_job(const Request &r, void *base) {
  const ba_serve_request &q = r.req;
  const uint64_t n = 1ull << q.log_n;
  uint64_t *work = (uint64_t *)((char *)base + r.work_off);
  ba_job j{};
  j.kind = BA_JOB_LDE;
  j.log_n = q.log_n;
  j.num_cols = q.num_cols;
  j.log_blowup = q.log_blowup;
  j.outputs = (uint64_t *)((char *)base + r.out_off);
  j.stride_out = n << q.log_blowup;
  j.out_coset_stride = n;
  if (q.input_kind == BA_SERVE_INPUT_DEVICE) {  // the caller's evaluations stay untouched
    j.inputs = const_cast<uint64_t *>(q.evals);
    j.stride_in = q.in_stride;
    j.monomials = work;
  } else {
    j.inputs = work;
    j.stride_in = n;
  }
  return j;
}

void ba_server::fail(Request &r, int status, Clock::time_point now) {
  if (r.work_bytes) arena.release(r.work_off, r.work_bytes);
  if (r.out_bytes) arena.release(r.out_off, r.out_bytes);
  r.work_bytes = r.out_bytes = 0;
  r.state = State::Failed;
  r.status = status;
  r.done_t = now;
  failed++;
}

// Enqueues the admitted requests of the oldest admitted request's size class (as many as the plan
// image capacity allows) on slot `s`; the other classes wait for the next free slot.  Called with
// the lock held; releases it while planning and enqueueing.  Each request is first validated on
// its own, so that one client's invalid request (say, device evaluations that overlap the arena)
// fails only itself and not the requests batched with it.
bool ba_server::launch(Slot &s, std::unique_lock<std::mutex> &lk) {
  const int cls = size_class(reqs.at(admitted.front()).req);
  std::vector<uint64_t> take;
  for (uint64_t t : admitted)
    if (size_class(reqs.at(t).req) == cls) take.push_back(t);
  std::vector<Request> rs;
  for (uint64_t t : take) rs.push_back(reqs.at(t));
  const uint64_t batch = batches++;
  s.busy = true;
  lk.unlock();
  std::vector<size_t> good;
  std::vector<std::pair<uint64_t, int>> bad;
  for (size_t i = 0; i < rs.size(); i++) {
    const ba_job j = request_job(rs[i], base);
    const int v = ba_validate_lde(j.inputs, j.stride_in, j.monomials, j.outputs, j.stride_out, j.out_coset_stride,
                                  j.log_n, j.num_cols, j.log_blowup);
    if (v)
      bad.push_back({take[i], v});
    else
      good.push_back(i);
  }
  BaPlanImage img;
  std::vector<ba_job> jobs;
  size_t count = good.size();
  int rc = BA_SUCCESS;
  while (count) {  // the largest prefix of the valid requests whose plan fits the slot
    jobs.clear();
    for (size_t k = 0; k < count; k++) jobs.push_back(request_job(rs[good[k]], base));
    // batches run on several slot streams at once: three-waves strided builds leave them room
    rc = ba_plan_image_build(ctx, jobs.data(), (uint32_t)jobs.size(), img, true);
    if (rc || img.bytes <= kDescBytes || count == 1) break;
    count = (count + 1) / 2;
  }
  if (rc == BA_SUCCESS && img.bytes > kDescBytes) rc = BA_ERROR_INVALID_VALUE;
  hipError_t err = hipSuccess;
  if (rc == BA_SUCCESS && count) {
    for (size_t k = 0; err == hipSuccess && k < count; k++) {
      const ba_serve_request &q = rs[good[k]].req;
      const uint64_t n = 1ull << q.log_n;
      uint64_t *work = (uint64_t *)((char *)base + rs[good[k]].work_off);
      if (q.input_kind == BA_SERVE_INPUT_HOST)
        err = hipMemcpy2DAsync(work, n * 8, q.evals, q.in_stride * 8, n * 8, q.num_cols, hipMemcpyHostToDevice,
                               s.stream);
      else if (q.input_kind == BA_SERVE_INPUT_SYNTHETIC)
        err = (hipError_t)ba_fill_uniform(work, n, n, q.num_cols, q.seed, s.stream);
    }
    if (err == hipSuccess && img.bytes) {
      ba_plan_image_write(img, s.staging, s.desc);
      err = hipMemcpyAsync(s.desc, s.staging, img.bytes, hipMemcpyHostToDevice, s.stream);
    }
    for (size_t g = 0; err == hipSuccess && g < img.groups.size(); g++)
      err = ba_plan_launch_group(img.groups[g], s.desc, img, s.stream);
    if (err == hipSuccess) err = hipEventRecord(s.done, s.stream);
    rc = (int)err;
  }
  lk.lock();
  const auto now = Clock::now();
  std::vector<uint64_t> gone;  // leave the admitted list: the invalid ones and the launched prefix
  for (const auto &b : bad) {
    fail(reqs.at(b.first), b.second, now);
    gone.push_back(b.first);
  }
  s.batch = batch;
  s.tickets.clear();
  for (size_t k = 0; k < count; k++) s.tickets.push_back(take[good[k]]);
  gone.insert(gone.end(), s.tickets.begin(), s.tickets.end());
  admitted.erase(std::remove_if(admitted.begin(), admitted.end(),
                                [&](uint64_t t) { return std::find(gone.begin(), gone.end(), t) != gone.end(); }),
                 admitted.end());
  for (uint64_t t : s.tickets) {
    Request &r = reqs.at(t);
    r.state = State::Running;
    r.batch = batch;
    r.launch_t = now;
  }
  batched_jobs += count;
  if (!bad.empty()) cv_done.notify_all();
  if (count == 0) {
    s.busy = false;
    return false;
  }
  if (rc != BA_SUCCESS) {
    // nothing of this batch may still run when its buffers go back to the arena
    (void)hipStreamSynchronize(s.stream);
    finish(s, rc);
    return false;
  }
  return true;
}

void ba_server::loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(m);
  for (;;) {
    bool busy = false;
    for (Slot &s : slots) {
      if (!s.busy || s.tickets.empty()) continue;
      const hipError_t q = hipEventQuery(s.done);
      if (q == hipSuccess)
        finish(s, 0);
      else if (q != hipErrorNotReady)
        finish(s, (int)q);
      else
        busy = true;
    }
    // admission in arrival order; while the head (oldest queued request) waits for memory, later
    // requests that fit pass it, at most max_bypass per head, after which the queue waits for it
    auto try_admit = [&](Request &r) {
      size_t w, o;
      if (!arena.alloc(r.work_bytes, w)) return false;
      if (!arena.alloc(r.out_bytes, o)) {
        arena.release(w, r.work_bytes);
        return false;
      }
      r.work_off = w;
      r.out_off = o;
      r.state = State::Admitted;
      r.admit_t = Clock::now();
      return true;
    };
    for (auto it = queued.begin(); it != queued.end();) {
      if (queued.front() != head_ticket) head_ticket = queued.front(), head_passed = 0;
      const bool is_head = it == queued.begin();
      if (!is_head && head_passed >= max_bypass) break;
      if (try_admit(reqs.at(*it))) {
        admitted.push_back(*it);
        if (!is_head) head_passed++;
        it = queued.erase(it);
      } else {
        ++it;
      }
    }
    // one batch per free slot, oldest size class first, while admitted requests remain
    for (Slot &s : slots) {
      if (admitted.empty()) break;
      if (s.busy) continue;
      launch(s, lk);
      busy = true;
    }
    if (stop && !busy && admitted.empty() && !queued.empty()) {
      // nothing runs and the head still does not fit: only releases could make room, and a
      // server being destroyed gets none, so the rest of the queue fails
      const auto now = Clock::now();
      for (uint64_t t : queued) {
        Request &r = reqs.at(t);
        r.state = State::Failed;
        r.status = BA_ERROR_OUT_OF_MEMORY;
        r.admit_t = r.done_t = now;
        r.work_bytes = r.out_bytes = 0;
        failed++;
      }
      queued.clear();
      cv_done.notify_all();
    }
    if (stop && queued.empty() && admitted.empty() && !busy) break;
    if (busy)
      cv_work.wait_for(lk, std::chrono::microseconds(50));
    else
      cv_work.wait(lk);
  }
}

extern "C" int ba_server_create(ba_context_t ctx, size_t memory_budget_bytes, uint32_t max_inflight_batches,
                                ba_server_t *out) {
  return ba_guarded([&]() -> int {
    if (!out) return BA_ERROR_INVALID_VALUE;
    *out = nullptr;
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    if (memory_budget_bytes < 4096 || max_inflight_batches < 1 || max_inflight_batches > 8)
      return BA_ERROR_INVALID_VALUE;
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device != ctx->device) return BA_ERROR_INVALID_VALUE;
    ba_server *s = new ba_server(Arena::round(memory_budget_bytes));
    s->ctx = ctx;
    ctx->users.fetch_add(1);
    s->device = device;
    hipError_t err = hipMalloc(&s->base, s->arena.size());
    if (err != hipSuccess) s->base = nullptr;
    s->slots.resize(max_inflight_batches);
    for (Slot &sl : s->slots) {
      if (err == hipSuccess) err = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
      if (err == hipSuccess) err = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
      if (err == hipSuccess) err = hipHostMalloc(&sl.staging, kDescBytes, hipHostMallocDefault);
      if (err == hipSuccess) err = hipMalloc(&sl.desc, kDescBytes);
    }
    if (err != hipSuccess) {
      ba_server_destroy(s);
      return err == hipErrorOutOfMemory ? BA_ERROR_OUT_OF_MEMORY : (int)err;
    }
    s->worker = std::thread([s] { s->loop(); });
    *out = s;
    return BA_SUCCESS;
  });
}

extern "C" int ba_server_set_max_bypass(ba_server_t s, uint32_t max_bypass) {
  return ba_guarded([&]() -> int {
    if (!s) return BA_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> g(s->m);
    s->max_bypass = max_bypass;
    s->cv_work.notify_one();
    return BA_SUCCESS;
  });
}

extern "C" int ba_server_submit(ba_server_t s, const ba_serve_request *req, uint64_t *ticket) {
  return ba_guarded([&]() -> int {
    if (!s || !req || !ticket) return BA_ERROR_INVALID_VALUE;
    const ba_serve_request &q = *req;
    const uint64_t n = q.log_n <= BA_MAX_LOG_N ? 1ull << q.log_n : 0;
    if (q.log_n < 1 || q.log_n > BA_MAX_LOG_N || q.num_cols == 0 || q.log_blowup < 1 ||
        q.log_n + q.log_blowup > BA_OMEGA_LOG_ORDER || q.input_kind > BA_SERVE_INPUT_SYNTHETIC)
      return BA_ERROR_INVALID_VALUE;
    ba_serve_request r = q;
    // one column has no stride to speak of; several need columns at least n apart
    if (r.in_stride == 0 || r.num_cols == 1) r.in_stride = n;
    if (r.input_kind != BA_SERVE_INPUT_SYNTHETIC && (!r.evals || r.in_stride < n)) return BA_ERROR_INVALID_VALUE;
    Request rec;
    rec.req = r;
    rec.work_bytes = work_elems(r) * 8;
    rec.out_bytes = out_elems(r) * 8;
    if (Arena::round(rec.work_bytes) + Arena::round(rec.out_bytes) > s->arena.size()) return BA_ERROR_OUT_OF_MEMORY;
    std::lock_guard<std::mutex> g(s->m);
    if (s->stop) return BA_ERROR_INVALID_VALUE;
    rec.submit_t = Clock::now();
    if (!s->have_first) s->have_first = true, s->first_submit = rec.submit_t;
    const uint64_t t = s->next_ticket++;
    s->reqs.emplace(t, rec);
    s->queued.push_back(t);
    *ticket = t;
    s->cv_work.notify_one();
    return BA_SUCCESS;
  });
}

static void fill_result(ba_server *s, const Request &r, ba_serve_result *res) {
  if (!res) return;
  *res = ba_serve_result{};
  res->status = r.state == State::Done ? 0 : r.status;
  res->outputs = r.state == State::Done ? (const uint64_t *)((char *)s->base + r.out_off) : nullptr;
  res->out_col_stride = (uint64_t)1 << (r.req.log_n + r.req.log_blowup);
  res->queue_ms = ms_between(r.submit_t, r.admit_t);
  res->latency_ms = ms_between(r.submit_t, r.done_t);
  res->batch = r.batch;
}

extern "C" int ba_server_wait(ba_server_t s, uint64_t ticket, ba_serve_result *res) {
  return ba_guarded([&]() -> int {
    if (!s) return BA_ERROR_INVALID_VALUE;
    std::unique_lock<std::mutex> lk(s->m);
    auto done = [&] {
      auto it = s->reqs.find(ticket);
      return it == s->reqs.end() || it->second.state == State::Done || it->second.state == State::Failed;
    };
    s->cv_done.wait(lk, done);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BA_ERROR_INVALID_VALUE;  // unknown or already released
    fill_result(s, it->second, res);
    return BA_SUCCESS;
  });
}

extern "C" int ba_server_poll(ba_server_t s, uint64_t ticket, int *done, ba_serve_result *res) {
  return ba_guarded([&]() -> int {
    if (!s || !done) return BA_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> g(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return BA_ERROR_INVALID_VALUE;
    *done = it->second.state == State::Done || it->second.state == State::Failed;
    if (*done) fill_result(s, it->second, res);
    return BA_SUCCESS;
  });
}

extern "C" int ba_server_release(ba_server_t s, uint64_t ticket) {
  return ba_guarded([&]() -> int {
    if (!s) return BA_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> g(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || (it->second.state != State::Done && it->second.state != State::Failed))
      return BA_ERROR_INVALID_VALUE;
    if (it->second.out_bytes) s->arena.release(it->second.out_off, it->second.out_bytes);
    s->reqs.erase(it);
    s->cv_work.notify_one();
    return BA_SUCCESS;
  });
}

extern "C" int ba_server_stats(ba_server_t s, ba_serve_stats *st) {
  return ba_guarded([&]() -> int {
    if (!s || !st) return BA_ERROR_INVALID_VALUE;
    std::lock_guard<std::mutex> g(s->m);
    *st = ba_serve_stats{};
    st->completed = s->completed;
    st->failed = s->failed;
    st->batches = s->batches;
    st->output_elems = s->out_total;
    st->budget_bytes = s->arena.size();
    st->peak_bytes = s->arena.peak();
    st->in_use_bytes = s->arena.used();
    st->mean_batch_jobs = s->batches ? (double)s->batched_jobs / (double)s->batches : 0.0;
    if (!s->latencies.empty()) {
      std::vector<double> v = s->latencies;
      auto q = [&](double f) {
        const size_t k = std::min(v.size() - 1, (size_t)(f * (double)(v.size() - 1) + 0.5));
        std::nth_element(v.begin(), v.begin() + k, v.end());
        return v[k];
      };
      st->p50_ms = q(0.50);
      st->p99_ms = q(0.99);
      double sum = 0, mx = 0;
      for (double x : s->latencies) sum += x, mx = std::max(mx, x);
      st->mean_ms = sum / (double)s->latencies.size();
      st->max_ms = mx;
    }
    if (s->have_first && s->completed) st->elapsed_s = ms_between(s->first_submit, s->last_done) / 1e3;
    return BA_SUCCESS;
  });
}

extern "C" int ba_server_destroy(ba_server_t s) {
  return ba_guarded([&]() -> int {
    if (!s) return BA_ERROR_INVALID_VALUE;
    if (s->worker.joinable()) {
      {
        std::lock_guard<std::mutex> g(s->m);
        s->stop = true;
        s->cv_work.notify_one();
      }
      s->worker.join();  // every submitted request has completed
    }
    for (Slot &sl : s->slots) {
      if (sl.stream) (void)hipStreamSynchronize(sl.stream);
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
      if (sl.done) (void)hipEventDestroy(sl.done);
      if (sl.staging) (void)hipHostFree(sl.staging);
      if (sl.desc) (void)hipFree(sl.desc);
    }
    if (s->base) (void)hipFree(s->base);
    if (s->ctx) s->ctx->users.fetch_sub(1);
    delete s;
    return BA_SUCCESS;
  });
}
