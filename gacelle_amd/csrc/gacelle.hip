// gacelle.hip -- the C ABI of include/gacelle.h: argument checks, model
// dispatch to the per-model operation tables, layout helpers, TV and the
// MCMC set-up.  The askadam driver is in gc_askadam.hpp; model-dependent
// launches are compiled per model variant in ops_*.hip.
#include <atomic>
#include <mutex>

#include "gc_dict.hpp"
#include "gc_host.hpp"

using namespace gc;

namespace {

// ---- model shape dispatch ------------------------------------------------
// The register-tiled model variant matching the description, or
// GC_ERR_UNSUPPORTED.
int select_ops(gc_ctx* ctx, const gc_model& md, const ModelOps** out) {
  *out = nullptr;
  switch (md.model_id) {
    case GC_MODEL_R2STAR: {
      const int nte = md.i[0];
      if (nte == 12 && (md.flags & GC_FLAG_UNIFORM_TE)) *out = ops_r2s12_uniform();
      else if (nte == 12) *out = ops_r2s12();
      else if (nte <= 16) *out = ops_r2s16();
      else if (nte <= 32) *out = ops_r2s32();
      else return set_err(ctx, GC_ERR_UNSUPPORTED, "R2star: at most 32 echoes supported (got %d)", nte);
      return GC_OK;
    }
    case GC_MODEL_JOINT_R1R2STAR: {
      const int nte = md.i[0], nfa = md.i[1];
      if (nte == 12 && nfa == 4) *out = ops_joint12x4();
      else if (nte <= 16 && nfa <= 4) *out = ops_joint16x4();
      else if (nte <= 16 && nfa <= 8) *out = ops_joint16x8();
      else
        return set_err(ctx, GC_ERR_UNSUPPORTED, "joint R1-R2*: at most 16 echoes x 8 flip angles (got %d x %d)",
                       nte, nfa);
      return GC_OK;
    }
    case GC_MODEL_NEXI: {
      const int ns = md.i[0];
      if (ns > 16) return set_err(ctx, GC_ERR_UNSUPPORTED, "NEXI: at most 16 shells (got %d)", ns);
      *out = (md.i[1] == 2) ? ops_nexi16_l2() : ops_nexi16();
      return GC_OK;
    }
    case GC_MODEL_MCRMWI: {
      if (md.i[1] != 4)
        return set_err(ctx, GC_ERR_UNSUPPORTED, "MCRMWI: 4 flip angles supported (got %d)", md.i[1]);
      const bool ex = (md.flags & GC_FLAG_MCRMWI_EXCHANGE) != 0, cx = (md.flags & GC_FLAG_MCRMWI_COMPLEX) != 0;
      if (md.flags & GC_FLAG_MCRMWI_EPG) {
        if (!ex) return set_err(ctx, GC_ERR_ARG, "MCRMWI: the EPG-X path needs the exchange model (isFitExchange)");
        *out = cx ? ops_mcrmwi4_epg_cplx() : ops_mcrmwi4_epg_mag();
        return GC_OK;
      }
      *out = ex ? (cx ? ops_mcrmwi4_exch_cplx() : ops_mcrmwi4_exch_mag())
                : (cx ? ops_mcrmwi4_cplx() : ops_mcrmwi4_mag());
      return GC_OK;
    }
    default:
      return set_err(ctx, GC_ERR_UNSUPPORTED, "model %d is not supported by this entry point", md.model_id);
  }
}

}  // namespace

// ---- scratch of gc_fwd_adj_noalloc -----------------------------------------
// What gc_fwd_adj allocates per call, in one fixed array per device: the
// model description's device copy (first kNoAllocModelBytes) and the FP64
// loss partial of every 256-voxel block.  In host code `scratch` names a
// host-side stand-in, not device memory: its device address on each device
// is looked up once (scratch_setup, with that device current) and only that
// address is used.
constexpr size_t kNoAllocModelBytes = 4096;
static_assert(sizeof(gc_model) <= kNoAllocModelBytes, "gc_model outgrew the scratch model slot");
static_assert(GC_NOALLOC_MAX_NV % kBlock == 0, "whole blocks");
constexpr size_t kNoAllocScratchBytes = kNoAllocModelBytes + sizeof(double) * size_t(GC_NOALLOC_MAX_NV / kBlock);
static_assert(kNoAllocScratchBytes == size_t(GC_NOALLOC_SCRATCH_BYTES), "header and library disagree");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device
hipError_t g_scratch_err[kMaxDevices] = {};       // the lookup's error, per device

// Look up the device address of `scratch` on `device` (the current device)
// once; later calls return the recorded address or error.
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device] || g_scratch_err[device] != hipSuccess) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) g_scratch_err[device] = (e != hipSuccess) ? e : hipErrorInvalidSymbol;
  else g_scratch_dev[device] = static_cast<unsigned char*>(p);
}

int scratch_of(gc_ctx* ctx, unsigned char** out) {
  *out = nullptr;
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return set_err(ctx, GC_ERR_UNSUPPORTED, "gc_fwd_adj_noalloc: device %d beyond %d", ctx->device, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[ctx->device])
    return set_err(ctx, GC_ERR_HIP, "gc_fwd_adj_noalloc: no scratch address on device %d (%s)", ctx->device,
                   hipGetErrorString(g_scratch_err[ctx->device]));
  *out = g_scratch_dev[ctx->device];
  return GC_OK;
}
}  // namespace

// ===========================================================================
// C ABI
extern "C" {

int gc_version(void) { return GC_ABI_VERSION; }

// live contexts: the block caches are released with the last one
static std::atomic<int> g_live_contexts{0};

int gc_ctx_create(int32_t device, gc_ctx** out) {
  if (!out) return GC_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GC_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return GC_ERR_ARG;
  gc_ctx* c = new gc_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return GC_ERR_HIP; }
  c->stream = nullptr;  // the device's default stream until gc_ctx_set_stream
  c->own_stream = false;
  scratch_setup(device);  // once per device, with the device current
  g_live_contexts.fetch_add(1);
  *out = c;
  return GC_OK;
}

int gc_ctx_destroy(gc_ctx* ctx) {
  if (!ctx) return GC_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_model) (void)hipFree(ctx->d_model);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  if (g_live_contexts.fetch_sub(1) == 1) trim_caches();
  return GC_OK;
}

int gc_ctx_set_stream(gc_ctx* ctx, void* stream) {
  if (!ctx) return GC_ERR_ARG;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  ctx->own_stream = false;
  ctx->stream = static_cast<hipStream_t>(stream);
  return GC_OK;
}

int gc_ctx_synchronize(gc_ctx* ctx) {
  if (!ctx) return GC_ERR_ARG;
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GC_OK;
}

const char* gc_last_error(const gc_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int gc_ctx_set_option(gc_ctx* ctx, int32_t option, int64_t value) {
  if (!ctx) return GC_ERR_ARG;
  switch (option) {
    case GC_OPT_GRAPHS:
      if (value != 0 && value != 1) return set_err(ctx, GC_ERR_ARG, "GC_OPT_GRAPHS takes 0 or 1");
      ctx->use_graphs = int(value);
      return GC_OK;
    case GC_OPT_MCMC_BLOCK:
      if (value != 0 && value != 64 && value != 128 && value != 256)
        return set_err(ctx, GC_ERR_ARG, "GC_OPT_MCMC_BLOCK takes 0 (autotune), 64, 128 or 256");
      ctx->mcmc_block = int(value);
      return GC_OK;
    case GC_OPT_ASK_BLOCK:
      if (value != 0 && value != 64 && value != 128 && value != 256)
        return set_err(ctx, GC_ERR_ARG, "GC_OPT_ASK_BLOCK takes 0 (autotune), 64, 128 or 256");
      ctx->ask_block = int(value);
      return GC_OK;
    default:
      return set_err(ctx, GC_ERR_ARG, "unknown context option %d", option);
  }
}

int gc_nd_to_gd(gc_ctx* ctx, const float* nd, int64_t nvox, int32_t M, const int64_t* vox, int64_t nv, float* gd) {
  if (!ctx || !nd || !vox || !gd || nv < 0 || M < 0) return set_err(ctx, GC_ERR_ARG, "gc_nd_to_gd: bad argument");
  if (nv * M == 0) return GC_OK;
  hipLaunchKernelGGL(k_nd_to_gd, dim3(nblocks(nv * M)), dim3(kBlock), 0, ctx->stream, nd, nvox, M, vox, nv, gd);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

int gc_gd_to_nd(gc_ctx* ctx, const float* gd, int64_t nv, int32_t M, const int64_t* vox, int64_t nvox, float* nd) {
  if (!ctx || !nd || !vox || !gd || nv < 0 || M < 0) return set_err(ctx, GC_ERR_ARG, "gc_gd_to_nd: bad argument");
  if (nvox * M > 0)
    hipLaunchKernelGGL(k_fill, dim3(nblocks(nvox * M)), dim3(kBlock), 0, ctx->stream, nd, nvox * M, 0.0f);
  if (nv * M > 0)
    hipLaunchKernelGGL(k_gd_to_nd, dim3(nblocks(nv * M)), dim3(kBlock), 0, ctx->stream, gd, nv, M, vox, nvox, nd);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

int gc_rescale_parameters(gc_ctx* ctx, const gc_model* md, int32_t transform, const float* th, float* z, int64_t nv) {
  if (!ctx || !md || !th || !z || nv < 0) return set_err(ctx, GC_ERR_ARG, "gc_rescale_parameters: bad argument");
  const int64_t n = nv * md->n_par;
  if (n == 0) return GC_OK;
  const gc_model* dmd = nullptr;
  int rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  hipLaunchKernelGGL(k_rescale, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, dmd, transform, th, z, nv);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

int gc_unscale_parameters(gc_ctx* ctx, const gc_model* md, int32_t transform, const float* z, float* th, int64_t nv) {
  if (!ctx || !md || !th || !z || nv < 0) return set_err(ctx, GC_ERR_ARG, "gc_unscale_parameters: bad argument");
  const int64_t n = nv * md->n_par;
  if (n == 0) return GC_OK;
  const gc_model* dmd = nullptr;
  int rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  hipLaunchKernelGGL(k_unscale, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, dmd, transform, z, th, nv);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

// The EPG-X rows of MCR-MWI at parameters x (nothing for other models)
static int epgx_aux(gc_ctx* ctx, const gc_model& md, const gc_model* dmd, const float* x, const float* extra,
                    int64_t nv, int transform, EpgxDev& e, DevBuf& aux) {
  if (md.model_id != GC_MODEL_MCRMWI || !(md.flags & GC_FLAG_MCRMWI_EPG)) return GC_OK;
  int rc = epgx_prepare(ctx, md, e);
  if (rc) return rc;
  GC_HIP(ctx, aux.alloc(sizeof(float) * size_t(kEpgxRows) * md.i[1] * nv));
  return epgx_eval(ctx, e, dmd, x, extra, nv, transform, aux.as<float>(), ctx->stream);
}

int gc_fwd(gc_ctx* ctx, const gc_model* md, const float* theta, const float* extra, float* s, int64_t nv) {
  if (!ctx) return GC_ERR_ARG;
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!theta || !s || nv < 0 || (md->n_extra > 0 && !extra)) return set_err(ctx, GC_ERR_ARG, "gc_fwd: bad argument");
  if (nv == 0) return GC_OK;
  const ModelOps* ops = nullptr;
  rc = select_ops(ctx, *md, &ops);
  if (rc) return rc;
  const gc_model* dmd = nullptr;
  rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  StreamScope scope_(ctx->stream);  // buffers of this call drain on its stream
  EpgxDev epgx;
  DevBuf aux;
  rc = epgx_aux(ctx, *md, dmd, theta, extra, nv, GC_TRANSFORM_NONE, epgx, aux);
  if (rc) return rc;
  rc = ops->fwd(ctx, dmd, theta, extra, s, nv, aux.as<float>());
  if (rc) return rc;
  if (aux.p) GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // scratch freed on return
  return GC_OK;
}

int gc_fwd_adj(gc_ctx* ctx, const gc_model* md, int32_t loss, int32_t transform, const float* x, const float* y,
               const float* w, const float* extra, int64_t nv, float scale, float* grad, double* loss_sum,
               float* residual) {
  if (!ctx) return GC_ERR_ARG;
  StreamScope scope_(ctx->stream);  // buffers of this call drain on its stream
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!x || !y || nv <= 0 || (md->n_extra > 0 && !extra)) return set_err(ctx, GC_ERR_ARG, "gc_fwd_adj: bad argument");
  const ModelOps* ops = nullptr;
  rc = select_ops(ctx, *md, &ops);
  if (rc) return rc;
  DevBuf part;
  const unsigned nb = nblocks(nv);
  if (loss_sum) GC_HIP(ctx, part.alloc(sizeof(double) * nb));
  const gc_model* dmd = nullptr;
  rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  EpgxDev epgx;
  DevBuf aux;
  rc = epgx_aux(ctx, *md, dmd, x, extra, nv, transform, epgx, aux);
  if (rc) return rc;
  rc = ops->fwd_adj(ctx, dmd, loss, transform, x, y, w, extra, nv, scale, grad,
                    loss_sum ? part.as<double>() : nullptr, residual, aux.as<float>());
  if (rc) return rc;
  if (loss_sum) {
    hipLaunchKernelGGL(k_reduce_rows_d, dim3(1), dim3(kBlock), 0, ctx->stream, part.as<double>(), int64_t(nb),
                       loss_sum);
    GC_HIP(ctx, hipGetLastError());
  }
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // scratch freed on return
  return GC_OK;
}

int gc_fwd_adj_noalloc(gc_ctx* ctx, const gc_model* md, int32_t loss, int32_t transform, const float* x,
                       const float* y, const float* w, const float* extra, int64_t nv, float scale, float* grad,
                       double* loss_sum, float* residual) {
  if (!ctx) return GC_ERR_ARG;
  if (nv > GC_NOALLOC_MAX_NV)
    return set_err(ctx, GC_ERR_ARG, "gc_fwd_adj_noalloc: at most %lld voxels per call (got %lld)",
                   (long long)GC_NOALLOC_MAX_NV, (long long)nv);
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!x || !y || nv <= 0 || (md->n_extra > 0 && !extra))
    return set_err(ctx, GC_ERR_ARG, "gc_fwd_adj_noalloc: bad argument");
  if (md->model_id == GC_MODEL_MCRMWI && (md->flags & GC_FLAG_MCRMWI_EPG))
    return set_err(ctx, GC_ERR_UNSUPPORTED, "gc_fwd_adj_noalloc: the EPG-X model needs per-call memory (gc_fwd_adj)");
  const ModelOps* ops = nullptr;
  rc = select_ops(ctx, *md, &ops);
  if (rc) return rc;
  unsigned char* base = nullptr;
  rc = scratch_of(ctx, &base);
  if (rc) return rc;
  gc_model* dmd = reinterpret_cast<gc_model*>(base);
  double* part = reinterpret_cast<double*>(base + kNoAllocModelBytes);
  const unsigned nb = nblocks(nv);
  // every block writes its own partial, and the reduction reads only those
  // the model with its derived fields, as upload_model writes it
  ctx->h_model = *md;
  derive_model(ctx->h_model);
  GC_HIP(ctx, hipMemcpyAsync(dmd, &ctx->h_model, sizeof(gc_model), hipMemcpyHostToDevice, ctx->stream));
  rc = ops->fwd_adj(ctx, dmd, loss, transform, x, y, w, extra, nv, scale, grad, loss_sum ? part : nullptr,
                    residual, nullptr);
  if (rc) return rc;
  if (loss_sum) {
    hipLaunchKernelGGL(k_reduce_rows_d, dim3(1), dim3(kBlock), 0, ctx->stream, part, int64_t(nb), loss_sum);
    GC_HIP(ctx, hipGetLastError());
  }
  return GC_OK;
}

int gc_tv_fwd_adj(gc_ctx* ctx, const float* img, const uint8_t* mask, const int32_t dims[3], int32_t mode,
                  const double vsize[3], double lambda, int64_t n_masked, double* loss, float* grad) {
  if (!ctx || !img || !mask || !dims || !vsize || !loss || n_masked <= 0)
    return set_err(ctx, GC_ERR_ARG, "gc_tv_fwd_adj: bad argument");
  StreamScope scope_(ctx->stream);  // buffers of this call drain on its stream
  const int64_t n = int64_t(dims[0]) * dims[1] * dims[2];
  if (n <= 0) return set_err(ctx, GC_ERR_ARG, "gc_tv_fwd_adj: empty grid");
  const unsigned nb = nblocks(n);
  DevBuf part, tot;
  GC_HIP(ctx, part.alloc(sizeof(double) * nb));
  const int nd = (mode == GC_TV_3D) ? 3 : 2;
  const float scale = float(lambda / double(n_masked));
  hipLaunchKernelGGL(k_tv, dim3(nb), dim3(kBlock), 0, ctx->stream, img, mask, dims[0], dims[1], dims[2], nd,
                     float(1.0 / vsize[0]), float(1.0 / vsize[1]), float(1.0 / vsize[2]), scale, part.as<double>(), grad);
  GC_HIP(ctx, hipGetLastError());
  GC_HIP(ctx, tot.alloc(sizeof(double)));
  hipLaunchKernelGGL(k_reduce_rows_d, dim3(1), dim3(kBlock), 0, ctx->stream, part.as<double>(), int64_t(nb),
                     tot.as<double>());
  double h = 0;
  GC_HIP(ctx, hipMemcpyAsync(&h, tot.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *loss = h * lambda / double(n_masked);
  return GC_OK;
}

int gc_posterior_stats(gc_ctx* ctx, const float* post, int32_t Q, int32_t n, int64_t nv, const float* lb,
                       const float* ub, int32_t nbin, float* median, float* iqr, float* mode) {
  if (!ctx) return GC_ERR_ARG;
  StreamScope scope_(ctx->stream);  // buffers of this call drain on its stream
  if (!post || Q < 1 || Q > GC_MAX_PAR || n < 1 || n > 16384 || nv < 1)
    return set_err(ctx, GC_ERR_ARG, "gc_posterior_stats: bad argument");
  if (mode && (!lb || !ub || nbin < 2)) return set_err(ctx, GC_ERR_ARG, "gc_posterior_stats: mode needs lb, ub, nbin");
  int n2 = 1;
  while (n2 < n) n2 <<= 1;
  const int vpb = std::max(1, std::min(16, 8192 / n2));
  int max_lds = 0;
  GC_HIP(ctx, hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  const size_t lds = size_t(vpb) * n2 * sizeof(float) + 2 * kBlock * sizeof(int);
  if (lds > size_t(max_lds))
    return set_err(ctx, GC_ERR_UNSUPPORTED, "gc_posterior_stats: %d samples per voxel need %zu B of LDS (max %d)", n, lds,
                   max_lds);
  DevBuf edges;
  std::vector<double> he;
  if (mode) {
    // linspace(lb - 1e-8, ub + 1e-8, nbin) in double (MATLAB single lb/ub promoted)
    he.resize(size_t(Q) * nbin);
    for (int q = 0; q < Q; ++q) {
      const double lo = double(lb[q]) - 1e-8, hi = double(ub[q]) + 1e-8;
      for (int i = 0; i < nbin; ++i)
        he[size_t(q) * nbin + i] = (i == nbin - 1) ? hi : lo + (hi - lo) * double(i) / double(nbin - 1);
    }
    GC_HIP(ctx, edges.alloc(sizeof(double) * he.size()));
    GC_HIP(ctx, hipMemcpyAsync(edges.p, he.data(), sizeof(double) * he.size(), hipMemcpyHostToDevice, ctx->stream));
  }
  PostStatParams sp{};
  sp.post = post;
  sp.edges = edges.as<double>();
  sp.median = median;
  sp.iqr = iqr;
  sp.mode = mode;
  sp.nv = nv;
  sp.n = n;
  sp.n2 = n2;
  sp.vpb = vpb;
  sp.nbin = nbin;
  const unsigned gx = unsigned((nv + vpb - 1) / vpb);
  hipLaunchKernelGGL(k_posterior_stats, dim3(gx, Q), dim3(kBlock), size_t(vpb) * n2 * sizeof(float), ctx->stream, sp);
  GC_HIP(ctx, hipGetLastError());
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // edges freed on return
  return GC_OK;
}

int gc_dict_search(gc_ctx* ctx, const float* y, int64_t nv, int32_t m, const float* dict, int64_t nd, const float* w,
                   int32_t* best, double* sse) {
  if (!ctx) return GC_ERR_ARG;
  StreamScope scope_(ctx->stream);  // buffers of this call drain on its stream
  return dict_search(ctx, y, nv, m, dict, nd, w, best, sse);
}

int gc_askadam_run(gc_ctx* ctx, const gc_model* md, const gc_askadam_cfg* cfg, gc_volume* vols, int32_t n_vol) {
  if (!ctx) return GC_ERR_ARG;
  StreamScope scope_(ctx->stream);  // buffers of this call drain on its stream
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!cfg || !vols || n_vol < 1 || n_vol > GC_MAX_VOL) return set_err(ctx, GC_ERR_ARG, "gc_askadam_run: bad argument");
  if (cfg->optimiser < 0 || cfg->optimiser > 2) return set_err(ctx, GC_ERR_ARG, "unknown optimiser");
  if (cfg->transform != GC_TRANSFORM_SIGMOID && cfg->transform != GC_TRANSFORM_LINEAR)
    return set_err(ctx, GC_ERR_ARG, "unknown parameter transform");
  if (cfg->loss < GC_LOSS_L1 || cfg->loss > GC_LOSS_MSE) return set_err(ctx, GC_ERR_ARG, "unknown loss function");
  const ModelOps* ops = nullptr;
  rc = select_ops(ctx, *md, &ops);
  if (rc) return rc;
  return ops->askadam(ctx, md, cfg, vols, n_vol);
}

// Native partner table of the GW sampler: same definition as the oracle's
// (Fisher-Yates derangement by rejection; red-black anchors with replacement),
// drawn from Philox4x32-10 on the host.
static void host_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t o[4]) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(0xD2511F53u) * c0;
    const uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
    const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
    c1 = uint32_t(p1);
    c3 = uint32_t(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

// `tag` selects the stream: 0xFFFF for the voxelwise blocks, 0xFFFC for the
// global-parameter block of goodman_weare_wglobal_constant.
static void host_partners(uint64_t seed, int rep, int it, int W, int mode, int32_t* out, uint32_t tag = 0xFFFFu) {
  const uint64_t stream = (uint64_t(tag) << 48) | (uint64_t(rep) << 32);
  auto draw = [&](uint64_t chain, uint32_t purpose, uint32_t o[4]) {
    host_philox(uint32_t(it), uint32_t(chain), uint32_t(chain >> 32), purpose, uint32_t(seed), uint32_t(seed >> 32), o);
  };
  if (mode == GC_ENSEMBLE_SIMULTANEOUS) {
    uint32_t attempt = 0;
    for (;;) {
      for (int i = 0; i < W; ++i) out[i] = i;
      for (int i = W - 1; i > 0; --i) {
        uint32_t o[4];
        draw(stream + attempt, uint32_t(i), o);
        const int j = int((uint64_t(o[0]) * uint64_t(i + 1)) >> 32);
        std::swap(out[i], out[j]);
      }
      bool ok = true;
      for (int i = 0; i < W; ++i) ok = ok && (out[i] != i);
      if (ok || W < 2) break;
      ++attempt;
    }
  } else {
    const int h = W / 2;
    const int na[2] = {h, W - h};
    int pos = 0;
    for (int s = 0; s < 2; ++s) {
      const int fro0 = (s == 0) ? h : 0;
      const int nfro = (s == 0) ? W - h : h;
      for (int j = 0; j < na[s]; ++j) {
        uint32_t o[4];
        draw(stream, uint32_t(1000 + pos), o);
        out[pos++] = fro0 + int((uint64_t(o[0]) * uint64_t(nfro)) >> 32);
      }
    }
  }
}

// Partner table [rep][steps][W]: the native stream, or the replayed one
// after a range check (red-black entries index the other half).
static int partner_table(gc_ctx* ctx, const gc_mcmc_cfg* cfg, int W, int steps, uint32_t tag,
                         const int32_t* replay_dev, DevBuf& buf, const int32_t** out) {
  const size_t n = size_t(cfg->repetition) * std::max(steps, 1) * W;
  std::vector<int32_t> tab(n, 0);
  if (cfg->rng_mode == GC_RNG_NATIVE) {
    for (int r = 0; r < cfg->repetition; ++r)
      for (int k = 1; k <= steps; ++k)
        host_partners(cfg->seed, r, k, W, cfg->ensemble_update, &tab[(size_t(r) * steps + (k - 1)) * W], tag);
    GC_HIP(ctx, buf.alloc(n * sizeof(int32_t)));
    GC_HIP(ctx, hipMemcpyAsync(buf.p, tab.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // tab is freed on return
    *out = buf.as<int32_t>();
    return GC_OK;
  }
  if (steps > 0) {
    GC_HIP(ctx, hipMemcpyAsync(tab.data(), replay_dev, n * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  const int h = W / 2;
  for (size_t i = 0; i < n && steps > 0; ++i) {
    const int w = int(i % W);
    int lo = 0, hi = W;
    if (cfg->ensemble_update == GC_ENSEMBLE_REDBLACK) {
      lo = (w < h) ? h : 0;
      hi = (w < h) ? W : h;
    }
    if (tab[i] < lo || tab[i] >= hi)
      return set_err(ctx, GC_ERR_ARG, "replayed partner index %d out of range at entry %zu", tab[i], i);
  }
  *out = replay_dev;
  return GC_OK;
}

// Zero posterior slots [n_stored, ns) of a [nq][rep][ns][inner] buffer.
static int zero_unfilled(gc_ctx* ctx, float* post, int nq, int rep, int ns, int n_stored, size_t inner) {
  if (!post || n_stored >= ns) return GC_OK;
  for (int q = 0; q < nq; ++q)
    for (int r = 0; r < rep; ++r) {
      float* p = post + ((size_t(q) * rep + r) * ns + n_stored) * inner;
      GC_HIP(ctx, hipMemsetAsync(p, 0, size_t(ns - n_stored) * inner * sizeof(float), ctx->stream));
    }
  return GC_OK;
}

// goodman_weare_wglobal_constant (utils/mcmc.m:385-578): the launch
// sequence of gc_gwglobal.hpp, one iteration at a time.
static int mcmc_global_run(gc_ctx* ctx, const ModelOps* ops, const gc_model* md, const gc_mcmc_cfg* cfg,
                           gc_mcmc_io* io, int ns, int n_stored) {
  const int Q = cfg->n_var, W = cfg->n_walker, R = cfg->repetition;
  const uint32_t all = (1u << Q) - 1u;
  const uint32_t gm = cfg->global_mask;
  const int NG = __builtin_popcount(gm & all), NND = Q - NG;
  if ((gm & ~all) != 0)
    return set_err(ctx, GC_ERR_ARG, "global_mask names variables beyond n_var (%d)", Q);
  if (NND == 0 || NG == 0)
    return set_err(ctx, GC_ERR_ARG,
                   "Need at least one local and one global parameter (NND=%d, NGlobal=%d). Use goodman_weare for "
                   "the all-local case.", NND, NG);
  if (cfg->algorithm != GC_MCMC_GW)
    return set_err(ctx, GC_ERR_UNSUPPORTED, "global parameters need the ensemble (Goodman-Weare) sampler");
  if (cfg->global_warmup < 0) return set_err(ctx, GC_ERR_ARG, "globalWarmup must be >= 0");
  if (cfg->rng_mode == GC_RNG_REPLAY && (!io->g_uniforms_dev || !io->g_partners_dev))
    return set_err(ctx, GC_ERR_ARG, "replay mode needs the global-block draw buffers");
  const int C = W * R;
  const int64_t nv = io->nv;
  const int T = cfg->global_warmup + cfg->iteration;
  const int gpb = kBlock / C;
  const int64_t ntile = (nv + gpb - 1) / gpb;
  const int nblk = int(std::min<int64_t>(ntile, 2048));

  DevBuf dlpart, dgpart, dxl, dlp, dlpg, dnacc, dg, dgprop, dgoob, dpend, dpartial, ds, dgs;
  GwgParams P{};
  int rc = partner_table(ctx, cfg, W, cfg->iteration, 0xFFFFu, io->partners_dev, dlpart, &P.partners);
  if (rc) return rc;
  rc = partner_table(ctx, cfg, W, T, 0xFFFCu, io->g_partners_dev, dgpart, &P.g_partners);
  if (rc) return rc;
  GC_HIP(ctx, dxl.alloc(size_t(NND) * nv * C * sizeof(float)));
  GC_HIP(ctx, dlp.alloc(size_t(nv) * C * sizeof(float)));
  GC_HIP(ctx, dlpg.alloc(size_t(nv) * C * sizeof(float)));
  GC_HIP(ctx, dnacc.alloc(size_t(nv) * C * sizeof(uint32_t)));
  GC_HIP(ctx, dg.alloc(size_t(C) * NG * sizeof(float)));
  GC_HIP(ctx, dgprop.alloc(size_t(C) * NG * sizeof(float)));
  GC_HIP(ctx, dgoob.alloc(size_t(C) * sizeof(int32_t)));
  GC_HIP(ctx, dpend.alloc(size_t(C) * sizeof(int32_t)));
  GC_HIP(ctx, dpartial.alloc(size_t(nblk) * C * sizeof(double)));
  GC_HIP(ctx, ds.alloc(size_t(2) * NND * nv * sizeof(double)));
  GC_HIP(ctx, dgs.alloc(size_t(2) * C * NG * sizeof(double)));
  GC_HIP(ctx, hipMemsetAsync(dpend.p, 0, size_t(C) * sizeof(int32_t), ctx->stream));
  GC_HIP(ctx, hipMemsetAsync(dgoob.p, 0, size_t(C) * sizeof(int32_t), ctx->stream));
  GC_HIP(ctx, hipMemsetAsync(ds.p, 0, size_t(2) * NND * nv * sizeof(double), ctx->stream));
  GC_HIP(ctx, hipMemsetAsync(dgs.p, 0, size_t(2) * C * NG * sizeof(double), ctx->stream));
  P.y = io->y_dev;
  P.w = io->w_dev;
  P.ex = io->extra_dev;
  P.x0 = io->x0_dev;
  P.normals = io->normals_dev;
  P.uniforms = io->uniforms_dev;
  P.g_uniforms = io->g_uniforms_dev;
  P.xl = dxl.as<float>();
  P.lp = dlp.as<float>();
  P.lpg = dlpg.as<float>();
  P.nacc = dnacc.as<uint32_t>();
  P.g = dg.as<float>();
  P.gprop = dgprop.as<float>();
  P.goob = dgoob.as<int32_t>();
  P.pend = dpend.as<int32_t>();
  P.part = dpartial.as<double>();
  P.s1 = ds.as<double>();
  P.s2 = P.s1 + size_t(NND) * nv;
  P.gs1 = dgs.as<double>();
  P.gs2 = P.gs1 + size_t(C) * NG;
  P.posterior = io->posterior_dev;
  P.g_posterior = io->g_posterior_dev;
  P.mean = io->mean_dev;
  P.stdv = io->std_dev;
  P.accept = io->accept_dev;
  P.xfinal = io->xfinal_dev;
  P.g_xfinal = io->g_xfinal_dev;
  P.trace = io->trace_dev;
  P.g_trace = io->g_trace_dev;
  P.nv = nv;
  P.seed = cfg->seed;
  P.iteration = cfg->iteration;
  P.T = T;
  P.rep = R;
  P.n_walker = W;
  P.nchain = C;
  P.n_sample = ns;
  P.n_stored = n_stored;
  P.rng = cfg->rng_mode;
  P.ens = cfg->ensemble_update;
  P.nnd = NND;
  P.ng = NG;
  P.nblk = nblk;
  P.gmask = gm;
  P.step_size = float(cfg->step_size);
  P.start_range = float(cfg->start_range);
  const gc_model* dmd = nullptr;
  rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  const bool hw = io->w_dev != nullptr;
  const bool rb = cfg->ensemble_update == GC_ENSEMBLE_REDBLACK;
  const unsigned grid = unsigned(nblk);
  const unsigned dgrid = unsigned((C + kBlock / 64 - 1) / (kBlock / 64));
  hipEvent_t ev0, ev1;
  GC_HIP(ctx, hipEventCreate(&ev0));
  GC_HIP(ctx, hipEventCreate(&ev1));
  auto run = [&]() -> int {
    GwgStep S{0, 0, 0, -1, -1};
    int r = ops->gwg(ctx, dmd, 0, hw, P, S, grid);
    if (r) return r;
    GC_HIP(ctx, hipEventRecord(ev0, ctx->stream));
    // one global block: red-black runs its halves one after the other
    auto global_block = [&](int k, int t, bool with_local, int keep) -> int {
      const int nsub = rb ? 2 : 1;
      for (int sub = 0; sub < nsub; ++sub) {
        const GwgStep st{k, t, (with_local && sub == 0) ? 1 : 0, sub, (sub == 0) ? keep : -1};
        int e = ops->gwg(ctx, dmd, 1, hw, P, st, grid);
        if (e) return e;
        const GwgStep sd{k, t, 0, sub, (sub == nsub - 1) ? keep : -1};
        hipLaunchKernelGGL(k_gwg_decide, dim3(dgrid), dim3(kBlock), 0, ctx->stream, dmd, P, sd);
        GC_HIP(ctx, hipGetLastError());
      }
      return GC_OK;
    };
    // global-only warm-up with the locals frozen (:503-507)
    for (int t = 1; t <= cfg->global_warmup; ++t) {
      r = global_block(0, t, false, -1);
      if (r) return r;
    }
    int counter = 0;
    for (int k = 1; k <= cfg->iteration; ++k) {
      int keep = -1;
      if (k > cfg->n_burnin && ((k - cfg->n_burnin + 1) % cfg->thinning) == 0) {  // :554
        keep = (counter < ns) ? counter : -1;
        ++counter;
      }
      r = global_block(k, cfg->global_warmup + k, true, keep);
      if (r) return r;
    }
    r = ops->gwg(ctx, dmd, 2, hw, P, S, grid);
    if (r) return r;
    GC_HIP(ctx, hipEventRecord(ev1, ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    io->kernel_ms = ms;
    return GC_OK;
  };
  rc = run();
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  return rc;
}

// Threads per block of the MH / GW sampler for this model variant and chain
// layout: each candidate (64, 128, 256 threads, at least one voxel's chains
// per block) runs a shortened copy of the call, with every output pointer
// null, on the first voxels, and the fastest is kept in the context.  The
// trial length doubles from 4 iterations until the 256-thread run takes a
// millisecond (or the call's own iteration count is reached).  Samples do
// not depend on the block size: the draws are indexed by (iteration, chain,
// voxel) and the per-voxel sums run in a fixed order.
static int tune_mcmc_block(gc_ctx* ctx, const ModelOps* ops, const gc_model* md, const gc_model* dmd, bool gw,
                           bool hw, const McmcParams& P, hipEvent_t ev0, hipEvent_t ev1, int* bs_out) {
  const int gsz = P.rep * (gw ? P.n_walker : 1);
  const std::array<int64_t, 9> key{md->model_id, md->flags, md->n_meas, md->n_par, gw ? 1 : 0, gsz,
                                   hw ? 1 : 0,    P.rng,     P.ens};
  auto it = ctx->mcmc_tuned.find(key);
  if (it != ctx->mcmc_tuned.end()) {
    *bs_out = it->second;
    return GC_OK;
  }
  std::vector<int> cand;
  for (int b : {64, 128, 256})
    if (b >= gsz) cand.push_back(b);
  if (cand.empty()) return set_err(ctx, GC_ERR_UNSUPPORTED, "Nwalker*repetition must be <= %d", kBlock);
  int best = cand.back();
  if (cand.size() > 1 && P.iteration > 0) {
    McmcParams T = P;
    T.posterior = T.mean = T.stdv = T.accept = T.xfinal = nullptr;
    T.trace = nullptr;
    // about 2^20 chains: enough blocks to fill the GPU at every candidate;
    // strides by T.nv keep every read inside the call's buffers
    T.nv = std::min<int64_t>(P.nv, std::max<int64_t>(1, (int64_t(1) << 20) / gsz));
    auto timed = [&](int b, float* ms) -> int {
      int e = ops->mcmc(ctx, dmd, gw, hw, T, b, ev0, ev1);
      if (e) return e;
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      GC_HIP(ctx, hipEventElapsedTime(ms, ev0, ev1));
      return GC_OK;
    };
    float ms = 0.f;
    T.iteration = std::min(P.iteration, 4);
    int e = timed(cand.back(), &ms);  // also loads the kernel
    if (e) return e;
    while (ms < 1.0f && T.iteration < P.iteration) {
      T.iteration = std::min(P.iteration, T.iteration * 2);
      e = timed(cand.back(), &ms);
      if (e) return e;
    }
    float best_ms = ms;
    for (int b : cand) {
      if (b == cand.back()) continue;
      e = timed(b, &ms);  // warm-up of this launch shape
      if (e) return e;
      e = timed(b, &ms);
      if (e) return e;
      if (ms < best_ms) {
        best_ms = ms;
        best = b;
      }
    }
  }
  ctx->mcmc_tuned[key] = best;
  *bs_out = best;
  return GC_OK;
}

int gc_mcmc_run(gc_ctx* ctx, const gc_model* md, const gc_mcmc_cfg* cfg, gc_mcmc_io* io) {
  if (!ctx) return GC_ERR_ARG;
  StreamScope scope_(ctx->stream);  // buffers of this call drain on its stream
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!cfg || !io || io->nv <= 0 || !io->y_dev || !io->x0_dev) return set_err(ctx, GC_ERR_ARG, "gc_mcmc_run: bad argument");
  if (md->n_extra > 0 && !io->extra_dev) return set_err(ctx, GC_ERR_ARG, "gc_mcmc_run: extra data missing");
  const ModelOps* ops = nullptr;
  rc = select_ops(ctx, *md, &ops);
  if (rc) return rc;
  const int np = model_np(*md);
  if (cfg->n_var != np + 1 || cfg->noise_slot != np || md->n_par != np + 1)
    return set_err(ctx, GC_ERR_UNSUPPORTED, "MCMC expects the model parameters followed by 'noise' (Q = %d)", np + 1);
  if (cfg->iteration < 0 || cfg->thinning < 1 || cfg->repetition < 1)
    return set_err(ctx, GC_ERR_ARG, "gc_mcmc_run: bad iteration/thinning/repetition");
  const bool gw = cfg->algorithm == GC_MCMC_GW;
  const int W = gw ? cfg->n_walker : 1;
  if (gw && cfg->ensemble_update == GC_ENSEMBLE_REDBLACK && W < 2)
    return set_err(ctx, GC_ERR_NWALKER, "redblack update needs Nwalker >= 2 (use even, >= 2*Nvar in practice).");
  if (gw && W < 2) return set_err(ctx, GC_ERR_NWALKER, "Goodman-Weare needs Nwalker >= 2");
  const int gsz = W * cfg->repetition;
  if (gsz > kBlock) return set_err(ctx, GC_ERR_UNSUPPORTED, "Nwalker*repetition must be <= %d", kBlock);
  if (cfg->global_mask != 0) {
    if (!gw) return set_err(ctx, GC_ERR_UNSUPPORTED, "global parameters need the ensemble (Goodman-Weare) sampler");
  }
  if (cfg->rng_mode == GC_RNG_REPLAY && (!io->normals_dev || !io->uniforms_dev || (gw && !io->partners_dev)))
    return set_err(ctx, GC_ERR_ARG, "replay mode needs the draw buffers");
  const int ns = (cfg->iteration < cfg->n_burnin + 1) ? 0 : (cfg->iteration - cfg->n_burnin - 1) / cfg->thinning + 1;
  io->n_sample = ns;
  {
    int cnt = 0;
    for (int k = cfg->n_burnin + 1; k <= cfg->iteration; ++k)
      if (((k - cfg->n_burnin + 1) % cfg->thinning) == 0) ++cnt;
    io->n_stored = std::min(cnt, ns);
  }
  {
    // Posterior slots that thinning leaves unfilled hold zeros, as in the
    // reference's preallocated posterior (utils/mcmc.m:121, :245, :481-482).
    const uint32_t all = (1u << cfg->n_var) - 1u;
    const int ng = cfg->global_mask ? __builtin_popcount(cfg->global_mask & all) : 0;
    rc = zero_unfilled(ctx, io->posterior_dev, cfg->n_var - ng, cfg->repetition, ns, io->n_stored,
                       size_t(W) * size_t(io->nv));
    if (rc) return rc;
    if (ng) rc = zero_unfilled(ctx, io->g_posterior_dev, ng, cfg->repetition, ns, io->n_stored, size_t(W));
    if (rc) return rc;
  }
  if (cfg->global_mask != 0) {
    io->block_threads = kBlock;
    return mcmc_global_run(ctx, ops, md, cfg, io, ns, io->n_stored);
  }
  McmcParams P{};
  P.y = io->y_dev;
  P.w = io->w_dev;
  P.ex = io->extra_dev;
  P.x0 = io->x0_dev;
  P.normals = io->normals_dev;
  P.uniforms = io->uniforms_dev;
  P.posterior = io->posterior_dev;
  P.mean = io->mean_dev;
  P.stdv = io->std_dev;
  P.accept = io->accept_dev;
  P.xfinal = io->xfinal_dev;
  P.trace = io->trace_dev;
  P.nv = io->nv;
  P.seed = cfg->seed;
  P.iteration = cfg->iteration;
  P.thinning = cfg->thinning;
  P.n_burnin = cfg->n_burnin;
  P.rep = cfg->repetition;
  P.n_walker = W;
  P.n_sample = ns;
  P.rng = cfg->rng_mode;
  P.ens = cfg->ensemble_update;
  P.q = cfg->n_var;
  P.noise = cfg->noise_slot;
  P.step_size = float(cfg->step_size);
  P.start_range = float(cfg->start_range);
  for (int q = 0; q < GC_MAX_PAR; ++q) P.x_step[q] = cfg->x_step[q];
  DevBuf dpart;
  const gc_model* dmd = nullptr;
  rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  if (gw) {
    if (cfg->rng_mode == GC_RNG_NATIVE) {
      const size_t n = size_t(cfg->repetition) * std::max(cfg->iteration, 1) * W;
      std::vector<int32_t> tab(n);
      for (int r = 0; r < cfg->repetition; ++r)
        for (int k = 1; k <= cfg->iteration; ++k)
          host_partners(cfg->seed, r, k, W, cfg->ensemble_update, &tab[(size_t(r) * cfg->iteration + (k - 1)) * W]);
      GC_HIP(ctx, dpart.alloc(n * sizeof(int32_t)));
      GC_HIP(ctx, hipMemcpyAsync(dpart.p, tab.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // tab is freed on return
      P.partners = dpart.as<int32_t>();
    } else {
      // replayed partner indices are checked before they index walker slots
      const size_t n = size_t(cfg->repetition) * std::max(cfg->iteration, 1) * W;
      std::vector<int32_t> tab(n);
      GC_HIP(ctx, hipMemcpyAsync(tab.data(), io->partners_dev, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                                 ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      const int h = W / 2;
      for (size_t i = 0; i < n; ++i) {
        const int w = int(i % W);
        int lo = 0, hi = W;
        if (cfg->ensemble_update == GC_ENSEMBLE_REDBLACK) {
          lo = (w < h) ? h : 0;
          hi = (w < h) ? W : h;
        }
        if (tab[i] < lo || tab[i] >= hi)
          return set_err(ctx, GC_ERR_ARG, "replayed partner index %d out of range at entry %zu", tab[i], i);
      }
      P.partners = io->partners_dev;
    }
  }
  hipEvent_t ev0, ev1;
  GC_HIP(ctx, hipEventCreate(&ev0));
  GC_HIP(ctx, hipEventCreate(&ev1));
  int bs = ctx->mcmc_block;
  if (bs == 0) rc = tune_mcmc_block(ctx, ops, md, dmd, gw, io->w_dev != nullptr, P, ev0, ev1, &bs);
  if (!rc && bs < gsz) rc = set_err(ctx, GC_ERR_ARG, "GC_OPT_MCMC_BLOCK %d is below Nwalker*repetition = %d", bs, gsz);
  io->block_threads = bs;
  if (!rc) rc = ops->mcmc(ctx, dmd, gw, io->w_dev != nullptr, P, bs, ev0, ev1);
  if (rc) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    return rc;
  }
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  io->kernel_ms = ms;
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  return GC_OK;
}

}  // extern "C"
