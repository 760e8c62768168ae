// gacelle.hip -- host side of libgacelle_hip.so: the C ABI of include/gacelle.h,
// kernel dispatch, and the askadam / mcmc drivers.
//
// The askadam driver reproduces the epoch loop of utils/askadam.m:261-299
// (min-loss tracking, convergence signals, stopping, update) in FP64 on the
// host, fed by per-iteration sums that the fused kernel produces K
// iterations at a time.  A stop or a new minimum that falls inside a block
// of K iterations is resolved exactly by replaying that volume from the
// block's starting state, which is kept (double-buffered state sets).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/gacelle.h"
#include "gc_kernels.hpp"

struct gc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err;
  gc_model* d_model = nullptr;  // device copy of the last model description
};

namespace {

using namespace gc;

int set_err(gc_ctx* ctx, int code, const char* fmt, ...) {
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

// RAII device buffer
struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = bytes;
    if (bytes == 0) return hipSuccess;
    return hipMalloc(&p, bytes);
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

struct HostBuf {
  void* p = nullptr;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() { if (p) (void)hipHostFree(p); }
  hipError_t alloc(size_t bytes) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    return hipHostMalloc(&p, std::max<size_t>(bytes, 64), hipHostMallocDefault);
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline unsigned nblocks(int64_t n, int b = kBlock) { return unsigned((n + b - 1) / b); }

// Model descriptions are read by kernels from device memory (uniform
// addresses: scalar loads), not from the kernel-argument segment, so that
// table lookups with a loop index never force a private copy.
int upload_model(gc_ctx* ctx, const gc_model* md, const gc_model** out) {
  if (!ctx->d_model) GC_HIP(ctx, hipMalloc(&ctx->d_model, sizeof(gc_model)));
  GC_HIP(ctx, hipMemcpyAsync(ctx->d_model, md, sizeof(gc_model), hipMemcpyHostToDevice, ctx->stream));
  *out = ctx->d_model;
  return GC_OK;
}

// ---- model shape dispatch ------------------------------------------------
// Calls f(Model{}) with the register-tiled model matching the description,
// or returns GC_ERR_UNSUPPORTED.
template <class F>
int with_model(gc_ctx* ctx, const gc_model& md, F&& f) {
  switch (md.model_id) {
    case GC_MODEL_R2STAR: {
      const int nte = md.i[0];
      if (nte == 12 && (md.flags & GC_FLAG_UNIFORM_TE)) return f(R2sModel<12, true, true>{});
      if (nte == 12) return f(R2sModel<12, true>{});
      if (nte <= 16) return f(R2sModel<16>{});
      if (nte <= 32) return f(R2sModel<32>{});
      return set_err(ctx, GC_ERR_UNSUPPORTED, "R2star: at most 32 echoes supported (got %d)", nte);
    }
    case GC_MODEL_JOINT_R1R2STAR: {
      const int nte = md.i[0], nfa = md.i[1];
      if (nte == 12 && nfa == 4) return f(JointModel<12, 4>{});
      if (nte <= 16 && nfa <= 4) return f(JointModel<16, 4>{});
      if (nte <= 16 && nfa <= 8) return f(JointModel<16, 8>{});
      return set_err(ctx, GC_ERR_UNSUPPORTED, "joint R1-R2*: at most 16 echoes x 8 flip angles (got %d x %d)",
                     nte, nfa);
    }
    case GC_MODEL_NEXI: {
      const int ns = md.i[0];
      if (ns > 16) return set_err(ctx, GC_ERR_UNSUPPORTED, "NEXI: at most 16 shells (got %d)", ns);
      if (md.i[1] == 2) return f(NexiModel<16, true>{});
      return f(NexiModel<16, false>{});
    }
    default:
      return set_err(ctx, GC_ERR_UNSUPPORTED, "model %d is not supported by this entry point", md.model_id);
  }
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

int model_np(const gc_model& md) {
  switch (md.model_id) {
    case GC_MODEL_R2STAR: return 2;
    case GC_MODEL_JOINT_R1R2STAR: return 3;
    case GC_MODEL_NEXI: return md.i[1] == 2 ? 5 : 4;
    default: return -1;
  }
}

int check_model(gc_ctx* ctx, const gc_model* md) {
  if (!md) return set_err(ctx, GC_ERR_ARG, "model is NULL");
  if (md->n_par < 1 || md->n_par > GC_MAX_PAR) return set_err(ctx, GC_ERR_ARG, "n_par out of range");
  if (md->n_extra < 0 || md->n_extra > 8) return set_err(ctx, GC_ERR_ARG, "n_extra out of range (0..8)");
  const int np = model_np(*md);
  if (np < 0) return set_err(ctx, GC_ERR_UNSUPPORTED, "model %d is not supported", md->model_id);
  if (md->n_par < np) return set_err(ctx, GC_ERR_ARG, "n_par (%d) below the model's %d parameters", md->n_par, np);
  int m_expect = 0;
  switch (md->model_id) {
    case GC_MODEL_R2STAR: m_expect = md->i[0]; break;
    case GC_MODEL_JOINT_R1R2STAR: m_expect = md->i[0] * md->i[1]; break;
    case GC_MODEL_NEXI: m_expect = md->i[0] * (md->i[1] == 2 ? 2 : 1); break;
  }
  if (md->n_meas != m_expect)
    return set_err(ctx, GC_ERR_INPUT_MISMATCH, "n_meas (%d) does not match the acquisition (%d)", md->n_meas, m_expect);
  if (md->model_id == GC_MODEL_JOINT_R1R2STAR) {
    const int need = (md->flags & GC_FLAG_R1R2S_TRUE_FA) ? md->i[1] : 1;
    if (md->n_extra < need) return set_err(ctx, GC_ERR_ARG, "joint R1-R2* needs %d extra rows", need);
  }
  if (md->model_id == GC_MODEL_NEXI) {
    const int ns = md->i[0], nn = md->i[2];
    if (nn < 2 || 2 * ns + 2 * nn + 34 > GC_MAX_TABLE) return set_err(ctx, GC_ERR_ARG, "NEXI table too large");
    if ((md->flags & GC_FLAG_NEXI_RICIAN) && md->n_extra < 1)
      return set_err(ctx, GC_ERR_ARG, "Rician NEXI needs sigma as extra row 0");
  }
  return GC_OK;
}

// ---- askadam driver --------------------------------------------------------
struct VolRun {
  int64_t nv = 0;
  int nblk = 0;
  DevBuf set[3][3];  // [state set][z, m, v]
  DevBuf zlast[2];
  DevBuf partials;
  DevBuf ones;       // weights when a volume has none but others do
  int cur = 0;          // input state set of the next launch
  int pinned = -1;      // state set kept for the min-loss replay
  int pin_e0 = 0, pin_r = 0;
  int min_zlast = -1;   // z_last buffer holding z at the min-loss iteration
  int zl_next = 0;      // z_last buffer the next launch writes
  bool active = true;
  int next_epoch = 1;
  // loop state of utils/askadam.m (FP64)
  double ema = 0, conv = 0;
  std::vector<double> cbuf;
  int pc = 0, ps = 0, pg = 0;
  double carry_step = 0;
  double loss = 0, fid = 0, reg = 0;
  double min_loss = 0, min_fid = 0, min_reg = 0;
  int min_iter = 0, last_epoch = 0, stop = GC_STOP_NONE;
  bool have_init = false;
  int final_from = -1;  // z_last buffer holding the final z, or -1 (replay)
  int final_set = -1, final_e0 = 0, final_r = 0;
  int launches = 0;
};

struct AskDriver {
  gc_ctx* ctx;
  const gc_model& md;
  const gc_askadam_cfg& cfg;
  gc_volume* vols;
  int n_vol;
  int K = 0;
  int total_blk = 0;
  const gc_model* dmd = nullptr;
  bool has_w = false;
  std::vector<VolRun> run;
  DevBuf d_vols, d_es, d_sums;
  HostBuf h_vols, h_es, h_sums;
  bool desc_in_flight = false;  // an async copy may still read h_vols/h_es
  // HIP events around every fused launch (kernel-only timing)
  std::vector<hipEvent_t> ev_start, ev_stop;
  double kernel_ms = 0.0;
  int kernel_launches = 0;

  ~AskDriver() {
    for (auto e : ev_start) (void)hipEventDestroy(e);
    for (auto e : ev_stop) (void)hipEventDestroy(e);
  }

  // accumulate the durations of recorded launches (after a stream sync)
  void collect_events() {
    for (size_t i = 0; i < ev_start.size(); ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev_start[i], ev_stop[i]) == hipSuccess) kernel_ms += ms;
      (void)hipEventDestroy(ev_start[i]);
      (void)hipEventDestroy(ev_stop[i]);
    }
    ev_start.clear();
    ev_stop.clear();
  }

  AskDriver(gc_ctx* c, const gc_model& m, const gc_askadam_cfg& a, gc_volume* v, int n)
      : ctx(c), md(m), cfg(a), vols(v), n_vol(n), run(n) {}

  bool stopping_possible() const {
    return cfg.tol > 0 || cfg.convergence_value >= 0 || cfg.step_tol > 0 || cfg.grad_tol > 0;
  }

  int setup() {
    int rcm = upload_model(ctx, &md, &dmd);
    if (rcm) return rcm;
    int64_t nv_total = 0;
    for (int i = 0; i < n_vol; ++i) {
      if (vols[i].nv <= 0) return set_err(ctx, GC_ERR_ARG, "volume %d has no voxels", i);
      if (!vols[i].y_dev || !vols[i].z_dev) return set_err(ctx, GC_ERR_ARG, "volume %d: y/z missing", i);
      if (md.n_extra > 0 && !vols[i].extra_dev) return set_err(ctx, GC_ERR_ARG, "volume %d: extra data missing", i);
      nv_total += vols[i].nv;
      if (vols[i].w_dev) has_w = true;
    }
    // iterations per launch: ~2.5e8 voxel-iterations per launch, fewer when a
    // stop can land inside the block (bounds the replay)
    if (cfg.block_iters > 0) K = cfg.block_iters;
    else {
      const double k = 2.5e8 / double(std::max<int64_t>(nv_total, 1));
      K = int(std::min(2000.0, std::max(16.0, k)));
      if (stopping_possible()) K = std::min(K, 256);
    }
    K = std::max(1, std::min(K, std::max(1, cfg.iteration)));
    const int P = md.n_par;
    total_blk = 0;
    for (int i = 0; i < n_vol; ++i) {
      VolRun& R = run[i];
      R.nv = vols[i].nv;
      R.nblk = int((R.nv + kBlock - 1) / kBlock);
      const size_t pbytes = size_t(P) * R.nv * sizeof(float);
      for (int s = 0; s < 3; ++s)
        for (int j = 0; j < 3; ++j) GC_HIP(ctx, R.set[s][j].alloc(pbytes));
      for (int s = 0; s < 2; ++s) GC_HIP(ctx, R.zlast[s].alloc(pbytes));
      GC_HIP(ctx, R.partials.alloc(size_t(K) * 4 * R.nblk * sizeof(float)));
      GC_HIP(ctx, hipMemcpyAsync(R.set[0][0].p, vols[i].z_dev, pbytes, hipMemcpyDeviceToDevice, ctx->stream));
      GC_HIP(ctx, hipMemsetAsync(R.set[0][1].p, 0, pbytes, ctx->stream));
      GC_HIP(ctx, hipMemsetAsync(R.set[0][2].p, 0, pbytes, ctx->stream));
      if (cfg.transform == GC_TRANSFORM_LINEAR) {
        // set_boundary01 of the starting point (utils/askadam.m:183-185)
        hipLaunchKernelGGL(k_clamp01, dim3(nblocks(int64_t(P) * R.nv)), dim3(kBlock), 0, ctx->stream,
                           R.set[0][0].as<float>(), int64_t(P) * R.nv);
      }
      if (has_w && !vols[i].w_dev) {
        const int64_t n = int64_t(md.n_meas) * R.nv;
        GC_HIP(ctx, R.ones.alloc(size_t(n) * sizeof(float)));
        hipLaunchKernelGGL(k_fill, dim3(nblocks(n)), dim3(kBlock), 0, ctx->stream, R.ones.as<float>(), n, 1.0f);
      }
      R.cbuf.assign(std::max(cfg.convergence_window, 1), 1.0);
      total_blk += R.nblk;
    }
    GC_HIP(ctx, d_vols.alloc(sizeof(VolDev) * n_vol));
    GC_HIP(ctx, h_vols.alloc(sizeof(VolDev) * n_vol));
    GC_HIP(ctx, d_es.alloc(sizeof(EpochScal) * K * n_vol));
    GC_HIP(ctx, h_es.alloc(sizeof(EpochScal) * K * n_vol));
    GC_HIP(ctx, d_sums.alloc(sizeof(double) * K * 4 * n_vol));
    GC_HIP(ctx, h_sums.alloc(sizeof(double) * K * 4 * n_vol));
    return GC_OK;
  }

  EpochScal epoch_scal(int e) const {
    EpochScal s{};
    const double lr = cfg.lr0 / (1.0 + cfg.decay * (e - 1));  // update_learn_rate, :1180-1182
    if (cfg.optimiser == GC_OPT_ADAM) {
      const double bc1 = 1.0 - std::pow(cfg.beta1, e), bc2 = 1.0 - std::pow(cfg.beta2, e);
      if (cfg.adam_form == GC_ADAM_KINGMA_BA) {
        s.lr = float(lr);
        s.c1 = float(1.0 / bc1);
        s.c2 = float(1.0 / bc2);
      } else {
        s.lr = float(lr * std::sqrt(bc2) / bc1);
        s.c1 = s.c2 = 1.f;
      }
    } else {
      s.lr = float(lr);
    }
    return s;
  }

  // Launch one block of iterations for the volumes with n_ep[i] > 0.
  // in_set/out_set select state sets; zl selects the z_last buffer (-1 none).
  template <class Model, int LOSS, bool HAS_W>
  int launch(const std::vector<int>& e0, const std::vector<int>& n_ep, const std::vector<int>& in_set,
             const std::vector<int>& out_set, const std::vector<int>& zl, int iteration_cap) {
    // the pinned descriptor buffers are reused: wait for the copies of the
    // previous launch before overwriting them
    if (desc_in_flight) GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    VolDev* hv = h_vols.as<VolDev>();
    EpochScal* hes = h_es.as<EpochScal>();
    int blk0 = 0;
    for (int i = 0; i < n_vol; ++i) {
      VolRun& R = run[i];
      VolDev d{};
      d.y = vols[i].y_dev;
      d.w = HAS_W ? (vols[i].w_dev ? vols[i].w_dev : R.ones.as<float>()) : nullptr;
      d.ex = vols[i].extra_dev;
      const int si = std::max(in_set[i], 0), so = std::max(out_set[i], 0);
      d.z_in = R.set[si][0].as<float>();
      d.m_in = R.set[si][1].as<float>();
      d.v_in = R.set[si][2].as<float>();
      d.z_out = R.set[so][0].as<float>();
      d.m_out = R.set[so][1].as<float>();
      d.v_out = R.set[so][2].as<float>();
      d.z_last = zl[i] >= 0 ? R.zlast[zl[i]].as<float>() : nullptr;
      d.partials = R.partials.as<float>();
      d.es = d_es.as<EpochScal>() + size_t(i) * K;
      d.nv = R.nv;
      d.blk0 = blk0;
      d.nblk = R.nblk;
      d.epoch0 = e0[i];
      d.n_epochs = n_ep[i];
      d.iteration = iteration_cap;
      d.inv_count = float(1.0 / (double(md.n_meas) * double(R.nv)));
      hv[i] = d;
      for (int k = 0; k < n_ep[i]; ++k) hes[size_t(i) * K + k] = epoch_scal(e0[i] + k);
      blk0 += R.nblk;
      if (n_ep[i] > 0) R.launches++;
    }
    GC_HIP(ctx, hipMemcpyAsync(d_vols.p, h_vols.p, sizeof(VolDev) * n_vol, hipMemcpyHostToDevice, ctx->stream));
    GC_HIP(ctx, hipMemcpyAsync(d_es.p, h_es.p, sizeof(EpochScal) * K * n_vol, hipMemcpyHostToDevice, ctx->stream));
    desc_in_flight = true;
    AskParams ap{};
    ap.vols = d_vols.as<VolDev>();
    ap.n_vol = n_vol;
    ap.opt = cfg.optimiser;
    ap.transform = cfg.transform;
    ap.adam_form = cfg.adam_form;
    ap.beta1 = float(cfg.beta1);
    ap.beta2 = float(cfg.beta2);
    ap.eps = float(cfg.eps);
    ap.momentum = float(cfg.momentum);
    ap.rms_decay = float(cfg.rms_decay);
    ap.rms_eps = float(cfg.rms_eps);
    hipEvent_t ev_a, ev_b;
    GC_HIP(ctx, hipEventCreate(&ev_a));
    GC_HIP(ctx, hipEventCreate(&ev_b));
    ev_start.push_back(ev_a);
    ev_stop.push_back(ev_b);
    GC_HIP(ctx, hipEventRecord(ev_a, ctx->stream));
    hipLaunchKernelGGL((k_askadam<Model, LOSS, HAS_W>), dim3(total_blk), dim3(kBlock), 0, ctx->stream, dmd, ap);
    GC_HIP(ctx, hipGetLastError());
    GC_HIP(ctx, hipEventRecord(ev_b, ctx->stream));
    ++kernel_launches;
    for (int i = 0; i < n_vol; ++i) {
      if (n_ep[i] <= 0) continue;
      hipLaunchKernelGGL(k_reduce_rows, dim3(n_ep[i] * 4), dim3(kBlock), 0, ctx->stream,
                         run[i].partials.as<float>(), int64_t(run[i].nblk), d_sums.as<double>() + size_t(i) * K * 4);
      GC_HIP(ctx, hipGetLastError());
    }
    return GC_OK;
  }

  // One iteration of the host-side loop logic (utils/askadam.m:264-292 minus
  // the update itself).  Returns true when the volume stops at `epoch`.
  bool epoch_logic(VolRun& R, int epoch, double sum_l, double sum_step_prev, double sum_den, double sum_g,
                   double* trace) {
    const double count = double(md.n_meas) * double(R.nv);
    R.fid = sum_l / count;
    R.reg = 0.0;
    R.loss = R.fid + R.reg;
    if (trace) trace[epoch] = R.loss;
    R.last_epoch = epoch;
    // update_min_loss (:433-446)
    bool new_min = false;
    if (R.min_loss > R.loss) {
      R.min_loss = R.loss; R.min_fid = R.fid; R.min_reg = R.reg; R.min_iter = epoch;
      new_min = true;
    }
    (void)new_min;
    // update_convergence_signals (:512-632)
    if (cfg.convergence_model == GC_CONV_EMA) {
      const double e_new = cfg.ema_decay * R.ema + (1.0 - cfg.ema_decay) * R.loss;
      R.conv = std::fabs(e_new - R.ema) / (std::fabs(R.ema) + 1e-8);
      R.ema = e_new;
    } else {
      std::rotate(R.cbuf.begin(), R.cbuf.begin() + 1, R.cbuf.end());
      R.cbuf.back() = R.loss;
      const int n = int(R.cbuf.size());
      double sx = 0, sy = 0, sxx = 0, sxy = 0;
      for (int i = 0; i < n; ++i) {
        const double xi = i + 1;
        sx += xi; sy += R.cbuf[i]; sxx += xi * xi; sxy += xi * R.cbuf[i];
      }
      const double den = n * sxx - sx * sx;
      R.conv = -((den != 0) ? (n * sxy - sx * sy) / den : 0.0);
    }
    if (R.conv > cfg.convergence_value || epoch <= cfg.convergence_window) R.pc = 0;
    else if ((R.min_loss - R.loss) > cfg.convergence_value) R.pc = 0;
    else R.pc += 1;
    if (cfg.step_tol > 0) {
      const double num = (epoch == 1) ? 0.0 : sum_step_prev;
      const double sn = std::sqrt(num) / (1.0 + std::sqrt(sum_den));
      R.ps = (sn < cfg.step_tol) ? R.ps + 1 : 0;
    }
    if (cfg.grad_tol > 0) {
      const double gn = std::sqrt(sum_g);
      R.pg = (gn < cfg.grad_tol) ? R.pg + 1 : 0;
    }
    // check_stopping (:634-671), in order
    int stop = GC_STOP_NONE;
    if (R.pc > cfg.patience_conv) stop = GC_STOP_CONVERGENCE;
    else if (R.loss < cfg.tol) stop = GC_STOP_TOL;
    else if (cfg.step_tol > 0 && R.ps > cfg.patience_step) stop = GC_STOP_STEP;
    else if (cfg.grad_tol > 0 && R.pg > cfg.patience_grad) stop = GC_STOP_GRAD;
    R.stop = stop;
    return stop != GC_STOP_NONE;
  }

  template <class Model, int LOSS, bool HAS_W>
  int run_all() {
    const int P = md.n_par;
    const int iter = cfg.iteration;
    double* sums = h_sums.as<double>();
    std::vector<int> e0(n_vol), nep(n_vol), in(n_vol), out(n_vol), zl(n_vol);

    if (iter <= 0) {
      // evaluation only (epoch 0): run one non-updating iteration
      for (int i = 0; i < n_vol; ++i) { e0[i] = 1; nep[i] = 1; in[i] = 0; out[i] = 1; zl[i] = 0; }
      int rc = launch<Model, LOSS, HAS_W>(e0, nep, in, out, zl, 0);
      if (rc) return rc;
      GC_HIP(ctx, hipMemcpyAsync(sums, d_sums.p, sizeof(double) * K * 4 * n_vol, hipMemcpyDeviceToHost, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      for (int i = 0; i < n_vol; ++i) {
        VolRun& R = run[i];
        const double* s = sums + size_t(i) * K * 4;
        R.fid = s[0] / (double(md.n_meas) * double(R.nv));
        R.loss = R.fid;
        R.min_loss = R.loss; R.min_fid = R.fid; R.min_iter = 0; R.last_epoch = 0;
        if (vols[i].loss_trace_host) vols[i].loss_trace_host[0] = R.loss;
        R.final_from = 0;
        R.min_zlast = 0;
      }
      return GC_OK;
    }

    bool any = true;
    while (any) {
      for (int i = 0; i < n_vol; ++i) {
        VolRun& R = run[i];
        if (!R.active) { nep[i] = 0; in[i] = R.cur; out[i] = R.cur; zl[i] = -1; e0[i] = 0; continue; }
        e0[i] = R.next_epoch;
        nep[i] = std::min(K, iter - R.next_epoch + 1);
        in[i] = R.cur;
        int o = 0;
        while (o == R.cur || o == R.pinned) ++o;
        out[i] = o;
        zl[i] = R.zl_next;
      }
      int rc = launch<Model, LOSS, HAS_W>(e0, nep, in, out, zl, iter);
      if (rc) return rc;
      GC_HIP(ctx, hipMemcpyAsync(sums, d_sums.p, sizeof(double) * K * 4 * n_vol, hipMemcpyDeviceToHost, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      desc_in_flight = false;
      collect_events();
      any = false;
      for (int i = 0; i < n_vol; ++i) {
        VolRun& R = run[i];
        if (!R.active) continue;
        const double* s = sums + size_t(i) * K * 4;
        double* trace = vols[i].loss_trace_host;
        if (!R.have_init) {
          // starting-point evaluation (:197-219) equals the first iteration's
          const double l0 = s[0] / (double(md.n_meas) * double(R.nv));
          R.loss = R.fid = l0; R.reg = 0;
          R.min_loss = l0; R.min_fid = l0; R.min_reg = 0; R.min_iter = 0;
          R.pinned = in[i]; R.pin_e0 = e0[i]; R.pin_r = 0; R.min_zlast = -1;
          R.ema = l0;
          if (trace) trace[0] = l0;
          R.have_init = true;
        }
        bool stopped = false;
        int k = 0;
        for (; k < nep[i]; ++k) {
          const int epoch = e0[i] + k;
          const double prev_min = R.min_loss;
          const double step_prev = (k == 0) ? R.carry_step : s[(k - 1) * 4 + 1];
          stopped = epoch_logic(R, epoch, s[k * 4 + 0], step_prev, s[k * 4 + 2], s[k * 4 + 3], trace);
          if (R.min_loss < prev_min) {
            if (k == nep[i] - 1) {
              R.min_zlast = zl[i];
              R.pinned = -1;
            } else {
              R.min_zlast = -1;
              R.pinned = in[i]; R.pin_e0 = e0[i]; R.pin_r = k;
            }
          }
          if (stopped) break;
        }
        if (stopped) {
          R.active = false;
          if (k == nep[i] - 1) { R.final_from = zl[i]; }
          else { R.final_from = -1; R.final_set = in[i]; R.final_e0 = e0[i]; R.final_r = k; }
        } else {
          R.carry_step = s[(nep[i] - 1) * 4 + 1];
          R.cur = out[i];
          R.next_epoch = e0[i] + nep[i];
          if (R.next_epoch > iter) {
            R.active = false;
            R.final_from = zl[i];
          }
        }
        // the next launch must not overwrite the z_last buffer holding the min
        R.zl_next = (R.min_zlast == 0) ? 1 : ((R.min_zlast == 1) ? 0 : (zl[i] == 0 ? 1 : 0));
        if (R.final_from >= 0 && R.final_from == R.zl_next) R.zl_next ^= 1;
        any = any || R.active;
      }
    }
    (void)P;
    return GC_OK;
  }

  // Replay `r`+1 iterations of one volume from state set `set` starting at
  // epoch e0 and leave z (pre-update) of iteration e0+r in z_last[zb].
  template <class Model, int LOSS, bool HAS_W>
  int replay(int vi, int set, int e0v, int r, int zb) {
    std::vector<int> e0(n_vol, 0), nep(n_vol, 0), in(n_vol, 0), out(n_vol, 0), zl(n_vol, -1);
    VolRun& R = run[vi];
    int o = 0;
    while (o == set || o == R.pinned) ++o;
    if (o > 2) o = (set + 1) % 3;
    e0[vi] = e0v; nep[vi] = r + 1; in[vi] = set; out[vi] = o; zl[vi] = zb;
    for (int i = 0; i < n_vol; ++i) if (i != vi) { in[i] = run[i].cur; out[i] = run[i].cur; }
    return launch<Model, LOSS, HAS_W>(e0, nep, in, out, zl, cfg.iteration);
  }

  template <class Model, int LOSS, bool HAS_W>
  int finalize() {
    const int P = md.n_par;
    for (int i = 0; i < n_vol; ++i) {
      VolRun& R = run[i];
      const size_t pbytes = size_t(P) * R.nv * sizeof(float);
      // final parameters
      int zb = R.final_from;
      if (zb < 0) {
        zb = (R.min_zlast == 0) ? 1 : 0;
        int rc = replay<Model, LOSS, HAS_W>(i, R.final_set, R.final_e0, R.final_r, zb);
        if (rc) return rc;
      }
      GC_HIP(ctx, hipMemcpyAsync(vols[i].z_dev, R.zlast[zb].p, pbytes, hipMemcpyDeviceToDevice, ctx->stream));
      // parameters at minimum loss
      if (vols[i].zmin_dev) {
        if (R.min_zlast >= 0) {
          GC_HIP(ctx, hipMemcpyAsync(vols[i].zmin_dev, R.zlast[R.min_zlast].p, pbytes, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        } else if (R.pin_r == 0) {
          GC_HIP(ctx, hipMemcpyAsync(vols[i].zmin_dev, R.set[R.pinned][0].p, pbytes, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        } else {
          const int zb2 = 0;
          int rc = replay<Model, LOSS, HAS_W>(i, R.pinned, R.pin_e0, R.pin_r, zb2);
          if (rc) return rc;
          GC_HIP(ctx, hipMemcpyAsync(vols[i].zmin_dev, R.zlast[zb2].p, pbytes, hipMemcpyDeviceToDevice,
                                     ctx->stream));
        }
      }
      if (cfg.transform == GC_TRANSFORM_LINEAR) {
        // set_boundary01 on both results (:314-317)
        hipLaunchKernelGGL(k_clamp01, dim3(nblocks(int64_t(P) * R.nv)), dim3(kBlock), 0, ctx->stream,
                           vols[i].z_dev, int64_t(P) * R.nv);
        if (vols[i].zmin_dev)
          hipLaunchKernelGGL(k_clamp01, dim3(nblocks(int64_t(P) * R.nv)), dim3(kBlock), 0, ctx->stream,
                             vols[i].zmin_dev, int64_t(P) * R.nv);
      }
      // residual outputs (out.final.resloss / residual, out.min.*)
      for (int which = 0; which < 2; ++which) {
        float* rl = which == 0 ? vols[i].resloss_dev : vols[i].resloss_min_dev;
        float* rs = which == 0 ? vols[i].residual_dev : vols[i].residual_min_dev;
        const float* zsrc = which == 0 ? vols[i].z_dev : vols[i].zmin_dev;
        if ((!rl && !rs) || !zsrc) continue;
        EvalParams ep{};
        ep.x = zsrc;
        ep.y = vols[i].y_dev;
        ep.w = HAS_W ? (vols[i].w_dev ? vols[i].w_dev : R.ones.as<float>()) : nullptr;
        ep.ex = vols[i].extra_dev;
        ep.residual = rs;
        ep.resloss = rl;
        ep.nv = R.nv;
        ep.transform = cfg.transform;
        ep.mode = 1;
        ep.scale = 1.f;
        hipLaunchKernelGGL((k_eval<Model, LOSS, HAS_W>), dim3(nblocks(R.nv)), dim3(kBlock), 0, ctx->stream, dmd, ep);
        GC_HIP(ctx, hipGetLastError());
      }
      gc_volume& V = vols[i];
      V.loss = R.loss; V.loss_fidelity = R.fid; V.loss_reg = R.reg;
      V.min_loss = R.min_loss; V.min_loss_fidelity = R.min_fid; V.min_loss_reg = R.min_reg;
      V.convergence = R.conv;
      V.n_iteration = R.last_epoch;
      V.min_iteration = R.min_iter;
      V.stop_reason = R.stop;
      V.n_launch = R.launches;
    }
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    desc_in_flight = false;
    collect_events();
    for (int i = 0; i < n_vol; ++i) {
      vols[i].kernel_ms = kernel_ms;
      vols[i].kernel_launches = kernel_launches;
      vols[i].iters_per_launch = K;
    }
    return GC_OK;
  }
};

}  // namespace

// ===========================================================================
// C ABI
extern "C" {

int gc_version(void) { return GC_ABI_VERSION; }

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
  *out = c;
  return GC_OK;
}

int gc_ctx_destroy(gc_ctx* ctx) {
  if (!ctx) return GC_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_model) (void)hipFree(ctx->d_model);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
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

int gc_fwd(gc_ctx* ctx, const gc_model* md, const float* theta, const float* extra, float* s, int64_t nv) {
  if (!ctx) return GC_ERR_ARG;
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!theta || !s || nv < 0 || (md->n_extra > 0 && !extra)) return set_err(ctx, GC_ERR_ARG, "gc_fwd: bad argument");
  if (nv == 0) return GC_OK;
  const gc_model* dmd = nullptr;
  rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  return with_model(ctx, *md, [&](auto model) -> int {
    using Model = decltype(model);
    EvalParams ep{};
    ep.x = theta;
    ep.ex = extra;
    ep.s_out = s;
    ep.nv = nv;
    ep.transform = GC_TRANSFORM_NONE;
    ep.mode = 0;
    hipLaunchKernelGGL((k_eval<Model, GC_LOSS_L2, false>), dim3(nblocks(nv)), dim3(kBlock), 0, ctx->stream, dmd, ep);
    GC_HIP(ctx, hipGetLastError());
    return GC_OK;
  });
}

int gc_fwd_adj(gc_ctx* ctx, const gc_model* md, int32_t loss, int32_t transform, const float* x, const float* y,
               const float* w, const float* extra, int64_t nv, float scale, float* grad, double* loss_sum,
               float* residual) {
  if (!ctx) return GC_ERR_ARG;
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!x || !y || nv <= 0 || (md->n_extra > 0 && !extra)) return set_err(ctx, GC_ERR_ARG, "gc_fwd_adj: bad argument");
  DevBuf part;
  const unsigned nb = nblocks(nv);
  if (loss_sum) GC_HIP(ctx, part.alloc(sizeof(double) * nb));
  const gc_model* dmd = nullptr;
  rc = upload_model(ctx, md, &dmd);
  if (rc) return rc;
  rc = with_model(ctx, *md, [&](auto model) -> int {
    using Model = decltype(model);
    return with_loss(ctx, loss, [&](auto lc) -> int {
      constexpr int LOSS = decltype(lc)::value;
      return with_bool(w != nullptr, [&](auto hw) -> int {
        constexpr bool HW = decltype(hw)::value;
        EvalParams ep{};
        ep.x = x;
        ep.y = y;
        ep.w = w;
        ep.ex = extra;
        ep.grad = grad;
        ep.residual = residual;
        ep.partial = loss_sum ? part.as<double>() : nullptr;
        ep.nv = nv;
        ep.transform = transform;
        ep.mode = 1;
        ep.scale = scale;
        hipLaunchKernelGGL((k_eval<Model, LOSS, HW>), dim3(nb), dim3(kBlock), 0, ctx->stream, dmd, ep);
        GC_HIP(ctx, hipGetLastError());
        return GC_OK;
      });
    });
  });
  if (rc) return rc;
  if (loss_sum) {
    hipLaunchKernelGGL(k_reduce_rows_d, dim3(1), dim3(kBlock), 0, ctx->stream, part.as<double>(), int64_t(nb),
                       loss_sum);
    GC_HIP(ctx, hipGetLastError());
  }
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // scratch freed on return
  return GC_OK;
}

int gc_tv_fwd_adj(gc_ctx* ctx, const float* img, const uint8_t* mask, const int32_t dims[3], int32_t mode,
                  const double vsize[3], double lambda, int64_t n_masked, double* loss, float* grad) {
  if (!ctx || !img || !mask || !dims || !vsize || !loss || n_masked <= 0)
    return set_err(ctx, GC_ERR_ARG, "gc_tv_fwd_adj: bad argument");
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

int gc_askadam_run(gc_ctx* ctx, const gc_model* md, const gc_askadam_cfg* cfg, gc_volume* vols, int32_t n_vol) {
  if (!ctx) return GC_ERR_ARG;
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!cfg || !vols || n_vol < 1 || n_vol > GC_MAX_VOL) return set_err(ctx, GC_ERR_ARG, "gc_askadam_run: bad argument");
  if (cfg->robust) return set_err(ctx, GC_ERR_UNSUPPORTED, "robustConvergence is not supported");
  if (cfg->n_reg > 0 && cfg->lambda[0] > 0)
    return set_err(ctx, GC_ERR_UNSUPPORTED, "TV regularisation needs the per-iteration path (not available for this model)");
  if (cfg->optimiser < 0 || cfg->optimiser > 2) return set_err(ctx, GC_ERR_ARG, "unknown optimiser");
  if (cfg->transform != GC_TRANSFORM_SIGMOID && cfg->transform != GC_TRANSFORM_LINEAR)
    return set_err(ctx, GC_ERR_ARG, "unknown parameter transform");
  AskDriver drv(ctx, *md, *cfg, vols, n_vol);
  rc = drv.setup();
  if (rc) return rc;
  return with_model(ctx, *md, [&](auto model) -> int {
    using Model = decltype(model);
    return with_loss(ctx, cfg->loss, [&](auto lc) -> int {
      constexpr int LOSS = decltype(lc)::value;
      return with_bool(drv.has_w, [&](auto hw) -> int {
        constexpr bool HW = decltype(hw)::value;
        int r = drv.template run_all<Model, LOSS, HW>();
        if (r) return r;
        return drv.template finalize<Model, LOSS, HW>();
      });
    });
  });
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

static void host_partners(uint64_t seed, int rep, int it, int W, int mode, int32_t* out) {
  const uint64_t stream = (uint64_t(0xFFFFu) << 48) | (uint64_t(rep) << 32);
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

int gc_mcmc_run(gc_ctx* ctx, const gc_model* md, const gc_mcmc_cfg* cfg, gc_mcmc_io* io) {
  if (!ctx) return GC_ERR_ARG;
  int rc = check_model(ctx, md);
  if (rc) return rc;
  if (!cfg || !io || io->nv <= 0 || !io->y_dev || !io->x0_dev) return set_err(ctx, GC_ERR_ARG, "gc_mcmc_run: bad argument");
  if (md->n_extra > 0 && !io->extra_dev) return set_err(ctx, GC_ERR_ARG, "gc_mcmc_run: extra data missing");
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
      GC_HIP(ctx, hipMemcpy(dpart.p, tab.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
      P.partners = dpart.as<int32_t>();
    } else {
      // replayed partner indices are checked before they index walker slots
      const size_t n = size_t(cfg->repetition) * std::max(cfg->iteration, 1) * W;
      std::vector<int32_t> tab(n);
      GC_HIP(ctx, hipMemcpy(tab.data(), io->partners_dev, n * sizeof(int32_t), hipMemcpyDeviceToHost));
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
  const int gpb = kBlock / gsz;
  const unsigned grid = unsigned((io->nv + gpb - 1) / gpb);
  hipEvent_t ev0, ev1;
  GC_HIP(ctx, hipEventCreate(&ev0));
  GC_HIP(ctx, hipEventCreate(&ev1));
  GC_HIP(ctx, hipEventRecord(ev0, ctx->stream));
  rc = with_model(ctx, *md, [&](auto model) -> int {
    using Model = decltype(model);
    return with_bool(io->w_dev != nullptr, [&](auto hw) -> int {
      constexpr bool HW = decltype(hw)::value;
      if (gw)
        hipLaunchKernelGGL((k_gw<Model, HW>), dim3(grid), dim3(kBlock), 0, ctx->stream, dmd, P);
      else
        hipLaunchKernelGGL((k_mh<Model, HW>), dim3(grid), dim3(kBlock), 0, ctx->stream, dmd, P);
      GC_HIP(ctx, hipGetLastError());
      return GC_OK;
    });
  });
  if (rc) {
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    return rc;
  }
  GC_HIP(ctx, hipEventRecord(ev1, ctx->stream));
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  io->kernel_ms = ms;
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  return GC_OK;
}

}  // extern "C"
