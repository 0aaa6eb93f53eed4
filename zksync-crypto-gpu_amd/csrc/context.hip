// context.hip -- resident power tables (boojum_cuda::context::Context equivalent).
//
// Reference: crates/boojum-cuda/src/context.rs:157-296 builds powers of omega_{2^30}
// (domain_generator_for_size(2^30)), of g = 7 and of g^-1 in fine/coarse halves, plus
// inv_sizes[k] = 2^-k.  Here they are generated on the host from the same closed forms, copied
// into one HBM allocation, and described by a gl::Tables struct that kernels take by value.
#include <mutex>
#include <vector>

#include "api_internal.h"

namespace {

typedef unsigned __int128 u128;
constexpr uint64_t P = gl::P;

uint64_t hmul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % P); }
uint64_t hpow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  a %= P;
  while (e) {
    if (e & 1) r = hmul(r, a);
    a = hmul(a, a);
    e >>= 1;
  }
  return r;
}

// omega_{2^k}: the 2^32-th root 7^((p-1)/2^32) squared (32 - k) times
uint64_t omega(unsigned log_size) {
  uint64_t w = hpow(7, (P - 1) >> 32);
  for (unsigned i = log_size; i < 32; i++) w = hmul(w, w);
  return w;
}

void fill_powers(uint64_t *dst, size_t count, uint64_t base) {
  uint64_t acc = 1;
  for (size_t i = 0; i < count; i++) {
    dst[i] = acc;
    acc = hmul(acc, base);
  }
}

std::mutex g_registry_mutex;
std::vector<ba_context *> g_live;  // creation order; the last live one per device is current

}  // namespace

ba_context *ba_current_context(int device) {
  std::lock_guard<std::mutex> lock(g_registry_mutex);
  for (auto it = g_live.rbegin(); it != g_live.rend(); ++it)
    if ((*it)->device == device) return *it;
  return nullptr;
}

extern "C" int ba_context_create(uint32_t w_coarse_log, uint32_t g_coarse_log, ba_context_t *out) {
  return ba_guarded([&]() -> int {
    if (!out || w_coarse_log > gl::OMEGA_LOG_ORDER || g_coarse_log > gl::OMEGA_LOG_ORDER)
      return BA_ERROR_INVALID_VALUE;
    *out = nullptr;
    int device = 0;
    hipError_t err = hipGetDevice(&device);
    if (err != hipSuccess) return (int)err;
    const uint32_t wf = gl::OMEGA_LOG_ORDER - w_coarse_log, gf = gl::OMEGA_LOG_ORDER - g_coarse_log;
    const size_t n_wf = (size_t)1 << wf, n_wc = (size_t)1 << w_coarse_log;
    const size_t n_gf = (size_t)1 << gf, n_gc = (size_t)1 << g_coarse_log;
    const size_t n_inv = gl::OMEGA_LOG_ORDER + 1;
    const size_t total = n_wf + n_wc + 2 * (n_gf + n_gc) + n_inv;
    std::vector<uint64_t> host(total);
    uint64_t *h = host.data();
    const uint64_t w30 = omega(gl::OMEGA_LOG_ORDER);
    const uint64_t g = 7, g_inv = hpow(7, P - 2);
    size_t off_wf = 0, off_wc = off_wf + n_wf, off_gf = off_wc + n_wc, off_gc = off_gf + n_gf;
    size_t off_if = off_gc + n_gc, off_ic = off_if + n_gf, off_inv = off_ic + n_gc;
    fill_powers(h + off_wf, n_wf, w30);
    fill_powers(h + off_wc, n_wc, hpow(w30, n_wf));
    fill_powers(h + off_gf, n_gf, g);
    fill_powers(h + off_gc, n_gc, hpow(g, n_gf));
    fill_powers(h + off_if, n_gf, g_inv);
    fill_powers(h + off_ic, n_gc, hpow(g_inv, n_gf));
    fill_powers(h + off_inv, n_inv, hpow(2, P - 2));

    ba_context *ctx = new ba_context();
    ctx->device = device;
    ctx->w_coarse_log = w_coarse_log;
    ctx->g_coarse_log = g_coarse_log;
    void *mem = nullptr;
    err = hipMalloc(&mem, total * sizeof(uint64_t));
    if (err != hipSuccess) {
      delete ctx;
      return (int)err;
    }
    err = hipMemcpy(mem, h, total * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (err != hipSuccess) {
      (void)hipFree(mem);
      delete ctx;
      return (int)err;
    }
    const uint64_t *d = (const uint64_t *)mem;
    ctx->mem = mem;
    ctx->bytes = total * sizeof(uint64_t);
    ctx->tab.w = {d + off_wf, d + off_wc, wf, (uint32_t)(n_wf - 1)};
    ctx->tab.g = {d + off_gf, d + off_gc, gf, (uint32_t)(n_gf - 1)};
    ctx->tab.g_inv = {d + off_if, d + off_ic, gf, (uint32_t)(n_gf - 1)};
    ctx->tab.inv_sizes = d + off_inv;
    // the LDE's auxiliary streams and events, created here so that no LDE call creates any (it may
    // run under stream capture); without them ba_lde_batch keeps every coset on the caller's stream
    bool aux_ok = hipEventCreateWithFlags(&ctx->lde_fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; aux_ok && i < ba_context::kLdeAux; i++)
      aux_ok = hipStreamCreateWithFlags(&ctx->lde_aux[i], hipStreamNonBlocking) == hipSuccess &&
               hipEventCreateWithFlags(&ctx->lde_join[i], hipEventDisableTiming) == hipSuccess;
    if (!aux_ok) {
      for (int i = 0; i < ba_context::kLdeAux; i++) {
        if (ctx->lde_aux[i]) (void)hipStreamDestroy(ctx->lde_aux[i]);
        if (ctx->lde_join[i]) (void)hipEventDestroy(ctx->lde_join[i]);
        ctx->lde_aux[i] = nullptr;
        ctx->lde_join[i] = nullptr;
      }
      if (ctx->lde_fork) (void)hipEventDestroy(ctx->lde_fork);
      ctx->lde_fork = nullptr;
      (void)hipGetLastError();  // the context works without them
    }
    {
      std::lock_guard<std::mutex> lock(g_registry_mutex);
      g_live.push_back(ctx);
    }
    *out = ctx;
    return BA_SUCCESS;
  });
}

extern "C" int ba_context_destroy(ba_context_t ctx) {
  return ba_guarded([&]() -> int {
    if (!ctx) return BA_ERROR_INVALID_VALUE;
    {
      std::lock_guard<std::mutex> lock(g_registry_mutex);
      // refused while a batch plan or server still uses its tables (checked before the lookup so
      // a refused destroy leaves the context fully usable)
      for (ba_context *c : g_live)
        if (c == ctx && ctx->users.load() > 0) return BA_ERROR_INVALID_VALUE;
      bool found = false;
      for (auto it = g_live.begin(); it != g_live.end(); ++it)
        if (*it == ctx) {
          g_live.erase(it);
          found = true;
          break;
        }
      if (!found) return BA_ERROR_INVALID_VALUE;
    }
    for (int i = 0; i < ba_context::kLdeAux; i++) {
      if (ctx->lde_aux[i]) {
        (void)hipStreamSynchronize(ctx->lde_aux[i]);
        (void)hipStreamDestroy(ctx->lde_aux[i]);
      }
      if (ctx->lde_join[i]) (void)hipEventDestroy(ctx->lde_join[i]);
    }
    if (ctx->lde_fork) (void)hipEventDestroy(ctx->lde_fork);
    hipError_t err = hipFree(ctx->mem);
    if (ctx->p2_round_constants) {
      const hipError_t e2 = hipFree(ctx->p2_round_constants);
      if (err == hipSuccess) err = e2;
    }
    delete ctx;
    return (int)err;
  });
}
