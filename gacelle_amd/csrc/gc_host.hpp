// gc_host.hpp -- host-side infrastructure shared by the translation units of
// libgacelle_hip.so: the context, error reporting, device and pinned
// buffers, model checks, and the table of per-model operations.  Each
// register-tiled model variant is compiled in a translation unit of its own
// (ops_*.hip) and reached through its ModelOps table.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/gacelle.h"
#include "gc_gwglobal.hpp"
#include "gc_epgx.hpp"

struct gc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  gc_model* d_model = nullptr;  // device copy of the last model description
  gc_model h_model{};           // its host staging copy (with the derived fields)
  int use_graphs = 1;           // GC_OPT_GRAPHS: step-form blocks replayed from hipGraphs
  int mcmc_block = 0;           // GC_OPT_MCMC_BLOCK: MCMC threads per block (0 = autotuned)
  // sampler block sizes picked by the autotuner, per model variant and
  // chain layout
  std::map<std::array<int64_t, 9>, int> mcmc_tuned;
  int ask_block = 0;            // GC_OPT_ASK_BLOCK: blocked askadam threads per block (0 = autotuned)
  // the blocked askadam kernel timed by the autotuner, per model variant:
  // its fastest block size, the fixed cost of one launch cycle (launch,
  // partial reduction, copy back, synchronisation) and its time per
  // voxel-iteration
  struct AskTune {
    int threads;
    double launch_ms, ms_per_voxel_iter;
  };
  std::map<std::array<int64_t, 8>, AskTune> ask_tuned;
};


namespace gc {


inline int set_err(gc_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

#define GC_HIP(ctx, call)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return set_err((ctx), GC_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                                      \
  } while (0)

// ---- caching allocators ------------------------------------------------------
// A fit allocates its working set (state sets, iteration partials, pinned
// descriptors) on every call.  hipMalloc / hipHostMalloc of hundreds of MB
// cost milliseconds, so released blocks are kept per device (device memory)
// or per process (pinned host memory) and handed out again to requests of
// the same size class.  Blocks are returned only when the stream that used
// them has drained; cached blocks are freed when an allocation fails, when
// the cache exceeds its cap and when the last context is destroyed.
class BlockCache {
 public:
  using AllocFn = hipError_t (*)(void**, size_t);
  using FreeFn = hipError_t (*)(void*);
  BlockCache(AllocFn a, FreeFn f, size_t cap) : alloc_(a), free_(f), cap_(cap) {}
  static size_t size_class(size_t bytes) {
    const size_t g = bytes <= (size_t(1) << 20) ? 4096 : (size_t(2) << 20);
    return (bytes + g - 1) / g * g;
  }
  hipError_t get(size_t bytes, void** out) {
    const size_t sz = size_class(bytes);
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = blocks_.lower_bound(sz);
      if (it != blocks_.end() && it->first <= sz + sz / 4) {
        *out = it->second.p;
        cached_ -= it->first;
        blocks_.erase(it);
        return hipSuccess;
      }
    }
    hipError_t e = alloc_(out, sz);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      trim();
      e = alloc_(out, sz);
    }
    return e;
  }
  void put(void* p, size_t bytes) {
    const size_t sz = size_class(bytes);
    std::lock_guard<std::mutex> lk(mu_);
    if (cached_ + sz > cap_) {
      (void)free_(p);
      return;
    }
    blocks_.emplace(sz, Block{p});
    cached_ += sz;
  }
  void trim() {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& b : blocks_) (void)free_(b.second.p);
    blocks_.clear();
    cached_ = 0;
  }

 private:
  struct Block {
    void* p;
  };
  AllocFn alloc_;
  FreeFn free_;
  size_t cap_;
  size_t cached_ = 0;
  std::mutex mu_;
  std::multimap<size_t, Block> blocks_;
};

inline hipError_t dev_malloc(void** p, size_t n) { return hipMalloc(p, n); }
inline hipError_t dev_free(void* p) { return hipFree(p); }
inline hipError_t pinned_malloc(void** p, size_t n) { return hipHostMalloc(p, n, hipHostMallocDefault); }
inline hipError_t pinned_free(void* p) { return hipHostFree(p); }

// One cache per device (never destroyed: freeing at process exit would run
// after the HIP runtime has shut down).
inline BlockCache& device_cache(int dev) {
  static std::mutex mu;
  static std::map<int, BlockCache*>* caches = new std::map<int, BlockCache*>();
  std::lock_guard<std::mutex> lk(mu);
  BlockCache*& c = (*caches)[dev];
  if (!c) c = new BlockCache(&dev_malloc, &dev_free, size_t(16) << 30);
  return *c;
}
inline BlockCache& pinned_cache() {
  static BlockCache* c = new BlockCache(&pinned_malloc, &pinned_free, size_t(256) << 20);
  return *c;
}
inline void trim_caches() {
  int n = 0;
  if (hipGetDeviceCount(&n) == hipSuccess)
    for (int d = 0; d < n; ++d) device_cache(d).trim();
  pinned_cache().trim();
}

// The stream the calling entry point issues its work on, for the buffers it
// allocates (set by StreamScope; null stream when unset).
inline hipStream_t& active_stream() {
  static thread_local hipStream_t s = nullptr;
  return s;
}
struct StreamScope {
  hipStream_t prev;
  explicit StreamScope(hipStream_t s) : prev(active_stream()) { active_stream() = s; }
  ~StreamScope() { active_stream() = prev; }
};
// Wait until work on `s` that may still use a block has finished.
inline void drain(hipStream_t s) {
  if (hipStreamQuery(s) != hipSuccess) (void)hipStreamSynchronize(s);
  (void)hipGetLastError();
}

// RAII device buffer from the device's block cache
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  int dev = 0;
  hipStream_t stream = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (!p) return;
    drain(stream);
    device_cache(dev).put(p, n);
    p = nullptr;
  }
  hipError_t alloc(size_t bytes) {
    release();
    n = bytes;
    if (bytes == 0) return hipSuccess;
    (void)hipGetDevice(&dev);
    stream = active_stream();
    return device_cache(dev).get(bytes, &p);
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// RAII pinned host buffer from the process's block cache
struct HostBuf {
  void* p = nullptr;
  size_t n = 0;
  hipStream_t stream = nullptr;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { release(); }
  void release() {
    if (!p) return;
    drain(stream);
    pinned_cache().put(p, n);
    p = nullptr;
  }
  hipError_t alloc(size_t bytes) {
    release();
    n = std::max<size_t>(bytes, 64);
    stream = active_stream();
    return pinned_cache().get(n, &p);
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline unsigned nblocks(int64_t n, int b = kBlock) { return unsigned((n + b - 1) / b); }

// Model descriptions are read by kernels from device memory (uniform
// addresses: scalar loads), not from the kernel-argument segment, so that
// table lookups with a loop index never force a private copy.
//
// Derived fields of the device copy, for R2* and joint R1-R2* with equally
// spaced echo times: when the spacing dTE = k TE_0 for k = 1 or 2 (relative
// 1e-6), i[7] = k and f[9] = -log2(e) TE_0, and both decay factors come from
// one exponential; otherwise i[7] = 0.  Any caller values in these slots are
// replaced.
inline void derive_model(gc_model& m) {
  if (m.model_id != GC_MODEL_R2STAR && m.model_id != GC_MODEL_JOINT_R1R2STAR) return;
  m.i[7] = 0;
  m.f[9] = 0.f;
  if (!(m.flags & GC_FLAG_UNIFORM_TE)) return;
  const double te0 = m.table[0], dte = m.f[8];
  if (!(te0 > 0.0) || !(dte > 0.0)) return;
  for (int k = 1; k <= 2; ++k) {
    if (std::fabs(te0 * k - dte) <= 1e-6 * dte) {
      m.i[7] = k;
      m.f[9] = float(-1.4426950408889634 * te0);
      return;
    }
  }
}

inline int upload_model(gc_ctx* ctx, const gc_model* md, const gc_model** out) {
  if (!ctx->d_model) GC_HIP(ctx, hipMalloc(&ctx->d_model, sizeof(gc_model)));
  // (a pageable source: the copy has read it when the call returns)
  ctx->h_model = *md;
  derive_model(ctx->h_model);
  GC_HIP(ctx, hipMemcpyAsync(ctx->d_model, &ctx->h_model, sizeof(gc_model), hipMemcpyHostToDevice, ctx->stream));
  *out = ctx->d_model;
  return GC_OK;
}
template <class F>
int with_loss(gc_ctx* ctx, int loss, F&& f) {
  switch (loss) {
    case GC_LOSS_L1: return f(std::integral_constant<int, GC_LOSS_L1>{});
    case GC_LOSS_L2: return f(std::integral_constant<int, GC_LOSS_L2>{});
    case GC_LOSS_HUBER: return f(std::integral_constant<int, GC_LOSS_HUBER>{});
    case GC_LOSS_MSE: return f(std::integral_constant<int, GC_LOSS_MSE>{});
    default: return set_err(ctx, GC_ERR_ARG, "unknown loss function %d", loss);
  }
}

template <class F>
int with_bool(bool b, F&& f) {
  if (b) return f(std::true_type{});
  return f(std::false_type{});
}

inline int model_np(const gc_model& md) {
  switch (md.model_id) {
    case GC_MODEL_R2STAR: return 2;
    case GC_MODEL_JOINT_R1R2STAR: return 3;
    case GC_MODEL_NEXI: return md.i[1] == 2 ? 5 : 4;
    case GC_MODEL_MCRMWI:
      return 9 + ((md.flags & GC_FLAG_MCRMWI_EXCHANGE) ? 1 : 0) +
             ((md.flags & GC_FLAG_MCRMWI_COMPLEX) ? md.i[1] + 1 : 0);
    default: return -1;
  }
}

inline int check_model(gc_ctx* ctx, const gc_model* md) {
  if (!md) return set_err(ctx, GC_ERR_ARG, "model is NULL");
  if (md->n_par < 1 || md->n_par > GC_MAX_PAR) return set_err(ctx, GC_ERR_ARG, "n_par out of range");
  if (md->n_extra < 0 || md->n_extra > kMaxExtra)
    return set_err(ctx, GC_ERR_ARG, "n_extra out of range (0..%d)", kMaxExtra);
  const int np = model_np(*md);
  if (np < 0) return set_err(ctx, GC_ERR_UNSUPPORTED, "model %d is not supported", md->model_id);
  if (md->n_par < np) return set_err(ctx, GC_ERR_ARG, "n_par (%d) below the model's %d parameters", md->n_par, np);
  int m_expect = 0;
  switch (md->model_id) {
    case GC_MODEL_R2STAR: m_expect = md->i[0]; break;
    case GC_MODEL_JOINT_R1R2STAR: m_expect = md->i[0] * md->i[1]; break;
    case GC_MODEL_NEXI: m_expect = md->i[0] * (md->i[1] == 2 ? 2 : 1); break;
    case GC_MODEL_MCRMWI: m_expect = md->i[0] * md->i[1] * ((md->flags & GC_FLAG_MCRMWI_COMPLEX) ? 2 : 1); break;
  }
  if (md->n_meas != m_expect)
    return set_err(ctx, GC_ERR_INPUT_MISMATCH, "n_meas (%d) does not match the acquisition (%d)", md->n_meas, m_expect);
  if (md->model_id == GC_MODEL_JOINT_R1R2STAR) {
    const int need = (md->flags & GC_FLAG_R1R2S_TRUE_FA) ? md->i[1] : 1;
    if (md->n_extra < need) return set_err(ctx, GC_ERR_ARG, "joint R1-R2* needs %d extra rows", need);
  }
  if (md->model_id == GC_MODEL_MCRMWI) {
    if (md->i[0] < 1 || md->i[0] > 32 || md->i[1] < 1) return set_err(ctx, GC_ERR_ARG, "MCRMWI: 1..32 echoes");
    const int need = 3 + md->i[1] * ((md->flags & GC_FLAG_MCRMWI_PINI_FA) ? 2 : 1);
    if (md->n_extra < need)
      return set_err(ctx, GC_ERR_ARG,
                     "MCRMWI needs %d extra rows (B1, pini, ff, freqBKG per flip angle[, pini per flip angle])", need);
  }
  if (md->model_id == GC_MODEL_NEXI) {
    const int ns = md->i[0], nn = md->i[2];
    if (nn < 2 || 2 * ns + 2 * nn + 34 > GC_MAX_TABLE) return set_err(ctx, GC_ERR_ARG, "NEXI table too large");
    if ((md->flags & GC_FLAG_NEXI_RICIAN) && md->n_extra < 1)
      return set_err(ctx, GC_ERR_ARG, "Rician NEXI needs sigma as extra row 0");
  }
  return GC_OK;
}

// Operations that depend on the register-tiled model type.
struct ModelOps {
  // forward model: s = FWD(theta) (model_forward of each model class)
  // (aux: the per-voxel rows of models that take them, else null)
  int (*fwd)(gc_ctx* ctx, const gc_model* dmd, const float* theta, const float* extra, float* s, int64_t nv,
             const float* aux);
  // loss, adjoint and residuals at x (askadam.model_gradient); part gets
  // one FP64 loss partial per block of kBlock voxels when not null
  int (*fwd_adj)(gc_ctx* ctx, const gc_model* dmd, int loss, int transform, const float* x, const float* y,
                 const float* w, const float* extra, int64_t nv, float scale, float* grad, double* part,
                 float* residual, const float* aux);
  // the askadam loop over a batch of volumes
  int (*askadam)(gc_ctx* ctx, const gc_model* md, const gc_askadam_cfg* cfg, gc_volume* vols, int n_vol);
  // one MH or GW launch (kernel-only timing through ev0/ev1)
  // one MH / GW sampler launch with bs threads per block
  int (*mcmc)(gc_ctx* ctx, const gc_model* dmd, bool gw, bool hw, const McmcParams& P, int bs,
              hipEvent_t ev0, hipEvent_t ev1);
  // one launch of the GW sampler with global parameters: kind 0 = initial
  // ensemble, 1 = iteration step (S), 2 = final statistics
  int (*gwg)(gc_ctx* ctx, const gc_model* dmd, int kind, bool hw, const GwgParams& P, const GwgStep& S,
             unsigned grid);
};

// ---- EPG-X networks (MCR-MWI isEPG) ------------------------------------------
// The two networks of a gc_model as the matrix-core weight tables the k_epgx
// kernels read (gc_epgx.hpp), built once per call.
struct EpgxDev {
  DevBuf wbuf;
  EpgxNets nets{};
  bool ready = false;
};
int epgx_prepare(gc_ctx* ctx, const gc_model& md, EpgxDev& out);
// aux rows from z_in of every active volume of a step launch (epgx.hip)
int epgx_step(gc_ctx* ctx, const EpgxDev& e, const gc_model* dmd, const StepParams& sp, unsigned grid, int transform,
              hipStream_t stream);
// aux rows from parameters x [P][nv] in the given transform space
int epgx_eval(gc_ctx* ctx, const EpgxDev& e, const gc_model* dmd, const float* x, const float* ex, int64_t nv,
              int transform, float* aux, hipStream_t stream);

// one getter per model variant, each defined in its own ops_*.hip
const ModelOps* ops_r2s12_uniform();
const ModelOps* ops_r2s12();
const ModelOps* ops_r2s16();
const ModelOps* ops_r2s32();
const ModelOps* ops_joint12x4();
const ModelOps* ops_joint16x4();
const ModelOps* ops_joint16x8();
const ModelOps* ops_nexi16();
const ModelOps* ops_nexi16_l2();
const ModelOps* ops_mcrmwi4_exch_cplx();
const ModelOps* ops_mcrmwi4_exch_mag();
const ModelOps* ops_mcrmwi4_cplx();
const ModelOps* ops_mcrmwi4_mag();
const ModelOps* ops_mcrmwi4_epg_cplx();
const ModelOps* ops_mcrmwi4_epg_mag();

}  // namespace gc
