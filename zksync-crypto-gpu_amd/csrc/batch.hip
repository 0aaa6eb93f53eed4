// batch.hip -- uneven batches: many NTT / LDE jobs of different sizes and directions, planned
// once and launched as a few grouped kernels.
//
// No reference counterpart: boojum-cuda launches one transform at a time (ntt.rs:179-291) and
// shivini loops over them (primitives/ntt.rs:146-244).  Here the columns of all jobs that share
// a transform (kind, size, direction, input order, coset; for an LDE: size and blowup) are
// merged into one multi-column transform whose columns are given by per-column pointer lists,
// so workgroups repeat tiles over columns of different jobs and amortise twiddles and scale
// factors as a strided batch does.  Every merged transform is expanded into its pass launches,
// each tagged with a dependency level (pass p of an NTT at level p; an LDE's inverse NTT at
// levels 0..P-1 and its coset NTTs at levels P..2P-1).  Launches of the same level and pass
// shape (kernel template) form a group that runs as ONE grouped launch: workgroup b finds its
// launch by binary search in a prefix table of workgroup counts.  Levels run in order on the
// stream, so the whole batch is (#levels x #shapes per level) launches, capturable as one
// hipGraph, with no synchronisation between workgroups.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "api_internal.h"

static_assert(sizeof(ba_job) == 80, "ba_job layout is part of the ABI");

struct ba_batch_plan {
  ba_context *ctx = nullptr;  // its tables are read by the plan's launches
  int device = 0;
  void *mem = nullptr;  // the uploaded image: PassArgs table, column pointer lists, prefix table
  BaPlanImage image;    // groups ordered by level, largest group first within a level
  // groups of one level are independent: they fork onto auxiliary streams (all at the default
  // priority) and join back, so a level runs as concurrent branches, in a graph too
  static constexpr int kAux = 4;
  hipStream_t aux[kAux] = {};
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> done;  // one per group
  ~ba_batch_plan() {
    if (ctx) ctx->users.fetch_sub(1);
    for (hipStream_t s : aux)
      if (s) (void)hipStreamDestroy(s);
    if (fork) (void)hipEventDestroy(fork);
    for (hipEvent_t e : done) (void)hipEventDestroy(e);
    if (mem) (void)hipFree(mem);
  }
};

namespace {

uint32_t bitrev_bits(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

// what makes two columns share every pass launch
struct Signature {
  uint32_t kind, log_n, bitrev, inverse, log_ext, coset, log_blowup;
  bool operator<(const Signature &o) const {
    return std::tie(kind, log_n, bitrev, inverse, log_ext, coset, log_blowup) <
           std::tie(o.kind, o.log_n, o.bitrev, o.inverse, o.log_ext, o.coset, o.log_blowup);
  }
};
struct Columns {
  std::vector<uint64_t *> in, out;             // NTT: inputs, outputs; LDE: evaluations
  std::vector<uint64_t *> mono;                // LDE: bit-reversed monomials
  std::vector<std::vector<uint64_t *>> coset;  // LDE: per coset, per column
};

int validate_job(const ba_job &j) {
  if (j.kind == BA_JOB_NTT) {
    const int rc = ba_validate_ntt(j.log_n, j.stride_in, j.stride_out, j.log_extension_degree, j.coset_index);
    if (rc) return rc;
    if (j.num_cols && (!j.inputs || !j.outputs)) return BA_ERROR_INVALID_VALUE;
    return BA_SUCCESS;
  }
  if (j.kind == BA_JOB_LDE)
    return ba_validate_lde(j.inputs, j.stride_in, j.monomials, j.outputs, j.stride_out, j.out_coset_stride,
                           j.log_n, j.num_cols, j.log_blowup);
  return BA_ERROR_INVALID_VALUE;
}

// a launch whose column pointers sit in the plan's pointer table, at offsets known before upload
struct PendingLaunch {
  uint32_t level;
  BaPassLaunch launch;
  size_t in_list, out_list;
};

int add_transform(const ba_context *ctx, uint32_t log_n, bool bitrev, bool inverse, uint32_t log_ext,
                  uint32_t coset, bool gather, BaLatPolicy lat, const std::vector<uint64_t *> &in,
                  const std::vector<uint64_t *> &out, uint32_t first_level, std::vector<const void *> &table,
                  std::vector<PendingLaunch> &pending, uint32_t *depth, bool shared) {
  std::vector<BaPassLaunch> ls;
  // merged NTT jobs take the latency-shaped passes (grouped like the others) where tuned to
  const int rc = ba_build_ntt(ctx, nullptr, nullptr, log_n, (uint32_t)in.size(), 0, 0, bitrev, inverse, log_ext,
                              coset, gather, ls, lat, shared);
  if (rc) return rc;
  const size_t li = table.size();
  table.insert(table.end(), in.begin(), in.end());
  const size_t lo = table.size();
  table.insert(table.end(), out.begin(), out.end());
  uint32_t d = 0;
  for (const BaPassLaunch &l : ls) {
    // pass 0 reads the inputs, later passes work in place on the outputs
    pending.push_back({first_level + (uint32_t)l.pass, l, (l.pass == 0 ? li : lo) + l.col0, lo + l.col0});
    d = std::max<uint32_t>(d, (uint32_t)l.pass + 1);
  }
  if (depth) *depth = d;
  return BA_SUCCESS;
}

}  // namespace

int ba_plan_image_build(const ba_context *ctx, const ba_job *jobs, uint32_t num_jobs, BaPlanImage &img,
                        bool shared) {
  img = BaPlanImage{};
  if (num_jobs && !jobs) return BA_ERROR_INVALID_VALUE;
  // 1. columns of all jobs, merged by signature
  std::map<Signature, Columns> merged;
  for (uint32_t i = 0; i < num_jobs; i++) {
    const ba_job &j = jobs[i];
    const int rc = validate_job(j);
    if (rc) return rc;
    if (j.num_cols == 0) continue;
    if (j.kind == BA_JOB_NTT) {
      Columns &c = merged[{BA_JOB_NTT, j.log_n, j.bitrev_inputs != 0, j.inverse != 0, j.log_extension_degree,
                           j.coset_index, 0}];
      for (uint64_t k = 0; k < j.num_cols; k++) {
        c.in.push_back(j.inputs + k * j.stride_in);
        c.out.push_back(j.outputs + k * j.stride_out);
      }
    } else {
      const uint64_t n = 1ull << j.log_n;
      const uint32_t blowup = 1u << j.log_blowup;
      Columns &c = merged[{BA_JOB_LDE, j.log_n, 0, 0, 0, 0, j.log_blowup}];
      c.coset.resize(blowup);
      for (uint64_t k = 0; k < j.num_cols; k++) {
        uint64_t *ev = j.inputs + k * j.stride_in;
        c.in.push_back(ev);
        c.mono.push_back(j.monomials ? j.monomials + k * n : ev);
        for (uint32_t cs = 0; cs < blowup; cs++)
          c.coset[cs].push_back(j.outputs + k * j.stride_out + cs * j.out_coset_stride);
      }
    }
  }
  // 2. pass launches of every merged transform
  std::vector<PendingLaunch> pending;
  for (const auto &kv : merged) {
    const Signature &sg = kv.first;
    const Columns &c = kv.second;
    int rc;
    if (sg.kind == BA_JOB_NTT) {
      rc = add_transform(ctx, sg.log_n, sg.bitrev, sg.inverse, sg.log_ext, sg.coset, false, BA_LAT_TUNED, c.in, c.out, 0,
                         img.table, pending, nullptr, shared);
    } else {
      // n2b inverse NTT to bit-reversed monomials, then every coset's forward NTT gathers them
      uint32_t depth = 0;
      rc = add_transform(ctx, sg.log_n, false, true, 0, 0, false, BA_LAT_NEVER, c.in, c.mono, 0, img.table, pending,
                         &depth, shared);
      for (uint32_t cs = 0; rc == BA_SUCCESS && cs < c.coset.size(); cs++)
        rc = add_transform(ctx, sg.log_n, false, false, sg.log_blowup, bitrev_bits(cs, sg.log_blowup), true, BA_LAT_NEVER,
                           c.mono, c.coset[cs], depth, img.table, pending, nullptr, shared);
    }
    if (rc) return rc;
  }
  // 3. group by (level, shape); std::map keeps levels ascending
  std::map<std::pair<uint32_t, void *>, std::vector<const PendingLaunch *>> by_group;
  for (const PendingLaunch &l : pending) by_group[{l.level, (void *)l.launch.kernels.group}].push_back(&l);
  for (auto &kv : by_group) {
    // split a group whose workgroup count would overflow the 32-bit prefix table or grid
    size_t i = 0;
    const auto &ls = kv.second;
    while (i < ls.size()) {
      BaGroup g{};
      g.kernel = ls[i]->launch.kernels.group;
      g.name = ls[i]->launch.name;
      g.level = kv.first.first;
      g.args_off = img.args.size();
      g.prefix_off = img.prefix.size();
      uint64_t wg = 0;
      while (i < ls.size() && wg + ls[i]->launch.grid <= 0x7FFFFFFFull) {
        img.prefix.push_back((uint32_t)wg);
        img.args.push_back(ls[i]->launch.args);
        img.lists.push_back({ls[i]->in_list, ls[i]->out_list});
        wg += ls[i]->launch.grid;
        i++;
      }
      if (img.args.size() == g.args_off) return BA_ERROR_INVALID_VALUE;  // one launch larger than a grid
      g.njobs = (uint32_t)(img.args.size() - g.args_off);
      g.workgroups = wg;
      img.workgroups += wg;
      img.levels = std::max(img.levels, g.level + 1);
      img.groups.push_back(g);
    }
  }
  std::stable_sort(img.groups.begin(), img.groups.end(), [](const BaGroup &a, const BaGroup &b) {
    return a.level != b.level ? a.level < b.level : a.workgroups > b.workgroups;
  });
  // device layout: PassArgs table, pointer table, prefix table, each 256-byte aligned
  const size_t args_bytes = img.args.size() * sizeof(gl::ntt::PassArgs);
  img.table_at = (args_bytes + 255) & ~(size_t)255;
  img.prefix_at = (img.table_at + img.table.size() * sizeof(void *) + 255) & ~(size_t)255;
  img.bytes = img.args.empty() ? 0 : img.prefix_at + img.prefix.size() * sizeof(uint32_t);
  return BA_SUCCESS;
}

void ba_plan_image_write(const BaPlanImage &img, void *host, void *device) {
  char *h = (char *)host;
  const void **d_table = (const void **)((char *)device + img.table_at);
  gl::ntt::PassArgs *ha = (gl::ntt::PassArgs *)h;
  for (size_t k = 0; k < img.args.size(); k++) {
    ha[k] = img.args[k];
    ha[k].in_cols = (const uint64_t *const *)(d_table + img.lists[k].first);
    ha[k].out_cols = (uint64_t *const *)(d_table + img.lists[k].second);
  }
  if (!img.table.empty()) memcpy(h + img.table_at, img.table.data(), img.table.size() * sizeof(void *));
  if (!img.prefix.empty()) memcpy(h + img.prefix_at, img.prefix.data(), img.prefix.size() * sizeof(uint32_t));
}

hipError_t ba_plan_launch_group(const BaGroup &g, const void *device, const BaPlanImage &img, hipStream_t s) {
  const gl::ntt::PassArgs *d_args = (const gl::ntt::PassArgs *)device + g.args_off;
  const uint32_t *d_prefix = (const uint32_t *)((const char *)device + img.prefix_at) + g.prefix_off;
  ba_trace_before(s, g.name);
  hipLaunchKernelGGL(g.kernel, dim3((uint32_t)g.workgroups), dim3(gl::ntt::THREADS), 0, s, d_args, d_prefix,
                     g.njobs);
  ba_trace_after(s);
  return hipGetLastError();
}

extern "C" int ba_batch_plan_create(ba_context_t ctx, const ba_job *jobs, uint32_t num_jobs,
                                    ba_batch_plan_t *plan) {
  return ba_guarded([&]() -> int {
    if (!plan) return BA_ERROR_INVALID_VALUE;
    *plan = nullptr;
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    ba_batch_plan *p = new ba_batch_plan();
    p->ctx = ctx;
    ctx->users.fetch_add(1);
    (void)hipGetDevice(&p->device);
    int rc = ba_plan_image_build(ctx, jobs, num_jobs, p->image);
    if (rc) {
      delete p;
      return rc;
    }
    hipError_t err = hipEventCreateWithFlags(&p->fork, hipEventDisableTiming);
    for (int i = 0; err == hipSuccess && i < ba_batch_plan::kAux; i++)
      err = hipStreamCreateWithFlags(&p->aux[i], hipStreamNonBlocking);
    p->done.resize(p->image.groups.size(), nullptr);
    for (size_t i = 0; err == hipSuccess && i < p->done.size(); i++)
      err = hipEventCreateWithFlags(&p->done[i], hipEventDisableTiming);
    if (err == hipSuccess && p->image.bytes) {
      err = hipMalloc(&p->mem, p->image.bytes);
      if (err != hipSuccess) p->mem = nullptr;
      if (err == hipSuccess) {
        std::vector<char> host(p->image.bytes);
        ba_plan_image_write(p->image, host.data(), p->mem);
        err = hipMemcpy(p->mem, host.data(), host.size(), hipMemcpyHostToDevice);
      }
    }
    if (err != hipSuccess) {
      delete p;
      return (int)err;
    }
    *plan = p;
    return BA_SUCCESS;
  });
}

extern "C" int ba_batch_plan_launch(ba_batch_plan_t plan, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    if (!plan) return BA_ERROR_INVALID_VALUE;
    hipStream_t s = (hipStream_t)stream;
    const std::vector<BaGroup> &groups = plan->image.groups;
    size_t i = 0;
    while (i < groups.size()) {
      size_t e = i;
      while (e < groups.size() && groups[e].level == groups[i].level) e++;
      hipError_t err = hipSuccess;
      if (e - i == 1) {
        err = ba_plan_launch_group(groups[i], plan->mem, plan->image, s);
      } else {
        // the largest group stays on the caller's stream, the others fork and join
        err = hipEventRecord(plan->fork, s);
        for (size_t k = i + 1; err == hipSuccess && k < e; k++) {
          hipStream_t a = plan->aux[(k - i - 1) % ba_batch_plan::kAux];
          err = hipStreamWaitEvent(a, plan->fork, 0);
          if (err == hipSuccess) err = ba_plan_launch_group(groups[k], plan->mem, plan->image, a);
          if (err == hipSuccess) err = hipEventRecord(plan->done[k], a);
        }
        if (err == hipSuccess) err = ba_plan_launch_group(groups[i], plan->mem, plan->image, s);
        for (size_t k = i + 1; err == hipSuccess && k < e; k++) err = hipStreamWaitEvent(s, plan->done[k], 0);
      }
      if (err != hipSuccess) return (int)err;
      i = e;
    }
    return BA_SUCCESS;
  });
}

extern "C" int ba_batch_plan_stats(ba_batch_plan_t plan, uint32_t *launches, uint32_t *levels,
                                   uint64_t *workgroups) {
  return ba_guarded([&]() -> int {
    if (!plan) return BA_ERROR_INVALID_VALUE;
    if (launches) *launches = (uint32_t)plan->image.groups.size();
    if (levels) *levels = plan->image.levels;
    if (workgroups) *workgroups = plan->image.workgroups;
    return BA_SUCCESS;
  });
}

extern "C" int ba_batch_plan_destroy(ba_batch_plan_t plan) {
  return ba_guarded([&]() -> int {
    if (!plan) return BA_ERROR_INVALID_VALUE;
    delete plan;  // releases streams, events and the descriptor tables
    return BA_SUCCESS;
  });
}
