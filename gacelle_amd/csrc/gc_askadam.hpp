// gc_askadam.hpp -- the askadam driver: the epoch loop of
// utils/askadam.m:261-299 (min-loss tracking, convergence signals,
// stopping, update) in FP64 on the host, fed by per-iteration sums that the
// fused kernel produces K iterations at a time.  A stop or a new minimum that
// falls inside a block of K iterations is resolved exactly by replaying that
// volume from the block's starting state, which is kept (state sets).
//
// Two launch forms share that logic.  The blocked form runs K iterations in
// one k_askadam launch with each voxel's data and state in registers.  The
// step form runs one k_askadam_step launch per iteration (state ping-pong
// through HBM): it carries the spatial TV regulariser, whose stencil needs
// the neighbours' previous parameters, and models whose data are streamed.
#pragma once
#include <cfloat>
#include <chrono>
#include <cstring>

#include "gc_host.hpp"

namespace gc {

// ---- askadam driver --------------------------------------------------------
struct VolRun {
  int64_t nv = 0;
  double nmean = 0;     // N of the loss mean (gc_volume.n_mean, else nv)
  int nblk = 0;      // blocks of kBlock voxels (the volume's padding)
  int nprod = 0;     // partial-sum rows: wavefronts (blocked form) or blocks (step form)
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
  double reg_offset = 0;  // the prior's out-of-mask loss (constant)
  int min_iter = 0, last_epoch = 0, stop = GC_STOP_NONE;
  bool have_init = false;
  int final_from = -1;  // z_last buffer holding the final z, or -1 (replay)
  int final_set = -1, final_e0 = 0, final_r = 0;
  int launches = 0;
  int chunks = 1;       // chunks per partial-sum row (k_reduce_rows)
  DevBuf pp[2][3];      // step form: intermediate states of a launch sequence
  DevBuf nbr;           // step form with TV: [2*nd][nv] grid neighbours
  DevBuf snb, sie, srv, srie;  // surface TV tables (StepVol)
  DevBuf aux;                  // EPG-X rows [15 n_fa][nv] of the current evaluation
  int snf = 0, snr = 0;
  // robustConvergence (utils/askadam.m:448-510, 541-566)
  std::vector<double> pvl_init, pvl, hist;  // per-voxel mean loss; history [W][nv]
  std::vector<uint8_t> main_mask, mask_conv;
  std::vector<int> flag_count;
  int64_t n_conv = 0;   // voxels in mask_conv
  int reprime = 0;
  DevBuf wmod, wmod_pin, mconv, pvl0_d, pvl_d;
  bool pin_mod = false; // the pinned block started with modulated weights
};

// per-iteration sums kept for the host loop: fidelity, step^2, |z|^2, |g|^2,
// regularisation, per-voxel mean loss over the convergence set
constexpr int kSum = kStepSums;

// step launches per captured graph (a chunk of an iteration block)
constexpr int kGraphSteps = 64;

// MATLAB median of a vector (mean of the two middle values for even n)
inline double median_of(std::vector<double>& v) {
  const size_t n = v.size();
  if (n == 0) return std::nan("");
  auto mid = v.begin() + n / 2;
  std::nth_element(v.begin(), mid, v.end());
  const double hi = *mid;
  if (n % 2) return hi;
  return 0.5 * (*std::max_element(v.begin(), mid) + hi);
}

struct AskDriver {
  gc_ctx* ctx;
  const gc_model& md;
  const gc_askadam_cfg& cfg;
  gc_volume* vols;
  int n_vol;
  int K = 0;
  int total_blk = 0;
  int blk_threads = kBlock;     // threads per block of the blocked kernel
  double tune_launch_ms = -1.0;  // measured cost of one launch cycle (ms)
  double tune_vox_ms = 0.0;     // measured kernel time per voxel-iteration (ms)
  // times the blocked kernel on the first call for a model variant
  // (tune_blocked, bound to the model by the operator table)
  int (*tuner)(AskDriver&) = nullptr;
  const gc_model* dmd = nullptr;
  bool has_w = false;
  std::vector<VolRun> run;
  DevBuf d_vols, d_es, d_sums, d_rows;
  HostBuf h_vols, h_es, h_rows;
  int max_chunks = 1;
  std::vector<double> row_sums;  // [vol][K*kSum] after fold_sums
  bool desc_in_flight = false;  // an async copy may still read h_vols/h_es
  bool step_mode = false;       // one launch per iteration (TV or streamed model)
  bool use_reg = false;  // a regulariser runs (n_reg > 0, lambda[0] > 0)
  bool prior = false;    // ... the normal prior
  bool use_tv = false;   // ... spatial total variation
  bool use_stv = false;  // ... surface (mesh) total variation
  bool use_epgx = false; // MCR-MWI EPG-X: k_epgx before every evaluation
  EpgxDev epgx;
  bool robust = false;
  bool replaying = false;       // a replay launch: no per-voxel loss outputs
  bool replay_pinned = false;   // ... from the pinned block (its weight modulator)
  DevBuf d_steps;               // step form: [K][n_vol] descriptors
  HostBuf h_steps;
  // start/stop events attached to every fused launch (kernel-only timing),
  // created once and reused
  std::vector<hipEvent_t> ev_start, ev_stop;
  size_t ev_used = 0;
  double kernel_ms = 0.0;
  int kernel_launches = 0;

  // Step form: graphs of n consecutive k_askadam_step launches, captured
  // once per call and size, then replayed for every chunk of n iterations.
  // Launch j of a graph reads descriptor row j of d_steps; chunks after the
  // first copy their rows to the front of d_steps (device to device, in
  // stream order) before their replay.  A graph holds kernel nodes only.
  std::vector<std::pair<int, hipGraphExec_t>> graphs;
  hipStream_t cap_stream = nullptr;  // capture only: nothing executes on it

  ~AskDriver() {
    for (auto e : ev_start) (void)hipEventDestroy(e);
    for (auto e : ev_stop) (void)hipEventDestroy(e);
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
  }

  // the graph of n launches, capturing it on first use; launch_k(j, stream)
  // issues launch j
  template <class LaunchK>
  int step_graph(int n, LaunchK&& launch_k, hipGraphExec_t* out) {
    for (auto& g : graphs)
      if (g.first == n) {
        *out = g.second;
        return GC_OK;
      }
    if (!cap_stream) GC_HIP(ctx, hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
    GC_HIP(ctx, hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < n; ++j) launch_k(j, cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cap_stream, &g);
    if (ec != hipSuccess || !g) return set_err(ctx, GC_ERR_HIP, "step graph capture: %s", hipGetErrorString(ec));
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return set_err(ctx, GC_ERR_HIP, "step graph instantiate: %s", hipGetErrorString(ei));
    graphs.push_back({n, ex});
    *out = ex;
    return GC_OK;
  }

  // accumulate the durations of recorded launches (after a stream sync)
  void collect_events() {
    for (size_t i = 0; i < ev_used; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, ev_start[i], ev_stop[i]) == hipSuccess) kernel_ms += ms;
    }
    ev_used = 0;
  }

  // the per-iteration sums of every volume, folded on the device by
  // k_fold_chunks (after a stream sync)
  void fold_sums() { std::memcpy(row_sums.data(), h_rows.p, sizeof(double) * K * kSum * n_vol); }

  // partial sums per iteration: all five in the step form, else loss,
  // step^2, |z|^2, |g|^2 or the loss alone
  int npv_count() const { return step_mode ? kStepSums : ((cfg.step_tol > 0 || cfg.grad_tol > 0) ? 4 : 1); }

  AskDriver(gc_ctx* c, const gc_model& m, const gc_askadam_cfg& a, gc_volume* v, int n)
      : ctx(c), md(m), cfg(a), vols(v), n_vol(n), run(n) {}

  bool stopping_possible() const {
    return cfg.tol > 0 || cfg.convergence_value >= 0 || cfg.step_tol > 0 || cfg.grad_tol > 0;
  }

  // stream_model: the model's data do not fit in registers (step form only)
  // epgx_model: the MCR-MWI EPG-X path (aux rows from the k_epgx pre-pass
  // before every model evaluation)
  int setup(bool stream_model, bool epgx_model = false) {
    use_reg = cfg.n_reg > 0 && cfg.lambda[0] > 0;
    prior = use_reg && cfg.reg_type == GC_REG_PRIOR;
    use_tv = use_reg && cfg.reg_type == GC_REG_TV;
    use_stv = use_reg && cfg.reg_type == GC_REG_SURFACE_TV;
    robust = cfg.robust != 0;
    step_mode = use_reg || stream_model || robust;
    if (use_reg && !prior && !use_tv && !use_stv)
      return set_err(ctx, GC_ERR_ARG, "unknown regulariser type %d", cfg.reg_type);
    if (robust && (cfg.weight_update_interval < 1 || cfg.outlier_check_window < 1 || cfg.convergence_window < 1))
      return set_err(ctx, GC_ERR_ARG, "robustConvergence needs weightUpdateInterval, outlierCheckWindow and "
                                      "convergenceWindow >= 1");
    if (use_reg) {
      if (cfg.n_reg > GC_MAX_REG) return set_err(ctx, GC_ERR_ARG, "at most %d regularised maps", GC_MAX_REG);
      for (int k = 0; k < cfg.n_reg; ++k)
        if (cfg.regmap[k] < 0 || cfg.regmap[k] >= model_np(md))
          return set_err(ctx, GC_ERR_ARG, "regmap %d names slot %d, not a model parameter", k, cfg.regmap[k]);
    }
    if (use_tv) {
      if (cfg.tv_mode != GC_TV_2D && cfg.tv_mode != GC_TV_3D) return set_err(ctx, GC_ERR_ARG, "unknown TVmode");
      for (int d = 0; d < 3; ++d)
        if (!(cfg.voxel_size[d] > 0)) return set_err(ctx, GC_ERR_ARG, "voxelSize must be positive");
    }
    if (prior)
      for (int i = 0; i < n_vol; ++i)
        if (!vols[i].prior_dev)
          return set_err(ctx, GC_ERR_ARG, "volume %d: the prior regulariser needs prior_dev (mu and sigma)", i);
    int rcm = upload_model(ctx, &md, &dmd);
    if (rcm) return rcm;
    use_epgx = epgx_model;
    if (use_epgx) {
      rcm = epgx_prepare(ctx, md, epgx);
      if (rcm) return rcm;
    }
    int64_t nv_total = 0;
    for (int i = 0; i < n_vol; ++i) {
      if (vols[i].nv <= 0) return set_err(ctx, GC_ERR_ARG, "volume %d has no voxels", i);
      if (!vols[i].y_dev || !vols[i].z_dev) return set_err(ctx, GC_ERR_ARG, "volume %d: y/z missing", i);
      if (md.n_extra > 0 && !vols[i].extra_dev) return set_err(ctx, GC_ERR_ARG, "volume %d: extra data missing", i);
      nv_total += vols[i].nv;
      if (vols[i].w_dev) has_w = true;
    }
    if (!step_mode && tuner) {
      const int rct = tuner(*this);
      if (rct) return rct;
    }
    // Iterations per launch.  When a stop can land inside a block, blocks of
    // ~2.5e8 voxel-iterations (at most 256 iterations) bound the replay that
    // resolves it.  Without stopping criteria the host has nothing to decide
    // between blocks: as many iterations per launch as the block partials
    // (4 bytes per block, iteration and sum) allow within 1 GiB.
    //
    // With a measured launch cost c (ms per launch cycle) and iteration time
    // t (ms per iteration of this call), blocks of K iterations cost about
    // (N / K) c in launches over N iterations, and a stop inside a block
    // about 1.5 K t (the rest of the block and the replay to the stop), so K
    // = sqrt(N c / (1.5 t)), within 16 .. 256.
    const int rows_per_blk = step_mode ? 1 : kBlock / 64;
    int64_t nblk_total = 0;
    for (int i = 0; i < n_vol; ++i) nblk_total += (vols[i].nv + kBlock - 1) / kBlock;
    const int64_t row_bytes = 4 * int64_t(npv_count()) * rows_per_blk;
    const int k_mem =
        int(std::max<int64_t>(16, (int64_t(1) << 30) / (row_bytes * std::max<int64_t>(nblk_total, 1))));
    if (cfg.block_iters > 0) K = cfg.block_iters;
    else if (stopping_possible()) {
      double k = 2.5e8 / double(std::max<int64_t>(nv_total, 1));
      if (!step_mode && tune_launch_ms > 0 && tune_vox_ms > 0) {
        const double t = tune_vox_ms * double(std::max<int64_t>(nv_total, 1));
        k = std::sqrt(double(std::max(cfg.iteration, 1)) * tune_launch_ms / (1.5 * t));
      }
      K = int(std::lround(std::min(256.0, std::max(16.0, k))));
    } else {
      K = std::min(k_mem, 8192);
    }
    if (robust) {
      // outlier checks and convergence-set refreshes fall on block ends
      int g = std::gcd(cfg.weight_update_interval, cfg.convergence_window);
      if (cfg.block_iters > 0) g = std::gcd(g, cfg.block_iters);
      K = g;
    }
    K = std::max(1, std::min(K, std::max(1, cfg.iteration)));
    if (K > 16383) return set_err(ctx, GC_ERR_ARG, "iterationsPerLaunch %d exceeds 16383", K);
    const int P = md.n_par;
    total_blk = 0;
    for (int i = 0; i < n_vol; ++i) {
      VolRun& R = run[i];
      R.nv = vols[i].nv;
      if (vols[i].n_mean < 0) return set_err(ctx, GC_ERR_ARG, "volume %d: n_mean must be >= 0", i);
      R.nmean = vols[i].n_mean > 0 ? double(vols[i].n_mean) : double(R.nv);
      if (vols[i].sign_trace_dev && (step_mode || cfg.loss != GC_LOSS_L1))
        return set_err(ctx, GC_ERR_UNSUPPORTED, "volume %d: the sign trace needs the L1 loss and the blocked form", i);
      R.reg_offset = prior ? vols[i].reg_offset : 0.0;
      R.nblk = int((R.nv + kBlock - 1) / kBlock);
      R.nprod = R.nblk * rows_per_blk;
      const size_t pbytes = size_t(P) * R.nv * sizeof(float);
      for (int s = 0; s < 3; ++s)
        for (int j = 0; j < 3; ++j) GC_HIP(ctx, R.set[s][j].alloc(pbytes));
      for (int s = 0; s < 2; ++s) GC_HIP(ctx, R.zlast[s].alloc(pbytes));
      GC_HIP(ctx, R.partials.alloc(size_t(K) * npv_count() * R.nprod * sizeof(float)));
      if (step_mode)
        for (int s = 0; s < 2; ++s)
          for (int j = 0; j < 3; ++j) GC_HIP(ctx, R.pp[s][j].alloc(pbytes));
      if (use_tv) {
        const int rc = build_neighbours(i);
        if (rc) return rc;
      }
      if (use_stv) {
        const int rc = build_surface(i);
        if (rc) return rc;
      }
      if (use_epgx) GC_HIP(ctx, R.aux.alloc(sizeof(float) * size_t(kEpgxRows) * md.i[1] * R.nv));
      if (robust) {
        GC_HIP(ctx, R.wmod.alloc(sizeof(float) * R.nv));
        GC_HIP(ctx, R.wmod_pin.alloc(sizeof(float) * R.nv));
        GC_HIP(ctx, R.mconv.alloc(R.nv));
        GC_HIP(ctx, R.pvl0_d.alloc(sizeof(float) * R.nv));
        GC_HIP(ctx, R.pvl_d.alloc(sizeof(float) * R.nv));
        hipLaunchKernelGGL(k_fill, dim3(nblocks(R.nv)), dim3(kBlock), 0, ctx->stream, R.wmod.as<float>(), R.nv, 1.0f);
        GC_HIP(ctx, hipMemsetAsync(R.mconv.p, 1, R.nv, ctx->stream));
        R.main_mask.assign(R.nv, 1);
        R.mask_conv.assign(R.nv, 1);
        R.flag_count.assign(R.nv, 0);
        R.n_conv = R.nv;
      }
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
      R.chunks = int((R.nprod + kReduceChunk - 1) / kReduceChunk);
      max_chunks = std::max(max_chunks, R.chunks);
      total_blk += R.nblk;
    }
    GC_HIP(ctx, d_vols.alloc(sizeof(VolDev) * n_vol));
    GC_HIP(ctx, h_vols.alloc(sizeof(VolDev) * n_vol));
    GC_HIP(ctx, d_es.alloc(sizeof(EpochScal) * K * n_vol));
    GC_HIP(ctx, h_es.alloc(sizeof(EpochScal) * K * n_vol));
    GC_HIP(ctx, d_sums.alloc(sizeof(double) * K * kSum * max_chunks * n_vol));
    GC_HIP(ctx, d_rows.alloc(sizeof(double) * K * kSum * n_vol));
    GC_HIP(ctx, h_rows.alloc(sizeof(double) * K * kSum * n_vol));
    row_sums.assign(size_t(K) * kSum * n_vol, 0.0);
    if (step_mode) {
      GC_HIP(ctx, d_steps.alloc(sizeof(StepVol) * K * n_vol));
      GC_HIP(ctx, h_steps.alloc(sizeof(StepVol) * K * n_vol));
    }
    return GC_OK;
  }

  // Grid neighbours of a volume's masked voxels for the TV stencil.  The
  // voxel indices are checked on the host first: they index the grid.
  int build_neighbours(int i) {
    const gc_volume& v = vols[i];
    const int64_t nv = v.nv;
    const int64_t X = v.dims[0], Y = v.dims[1], Z = v.dims[2];
    if (!v.vox_dev || X < 1 || Y < 1 || Z < 1)
      return set_err(ctx, GC_ERR_ARG, "volume %d: TV needs vox_dev and dims", i);
    const int64_t ngrid = X * Y * Z;
    if (ngrid >= (int64_t(1) << 31) || nv > ngrid)
      return set_err(ctx, GC_ERR_ARG, "volume %d: grid of %lld voxels unsupported", i, (long long)ngrid);
    std::vector<int64_t> hv(nv);
    GC_HIP(ctx, hipMemcpyAsync(hv.data(), v.vox_dev, sizeof(int64_t) * nv, hipMemcpyDeviceToHost, ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int64_t k = 0; k < nv; ++k)
      if (hv[k] < 0 || hv[k] >= ngrid || (k > 0 && hv[k] <= hv[k - 1]))
        return set_err(ctx, GC_ERR_ARG, "volume %d: vox must be increasing grid indices (entry %lld)", i,
                       (long long)k);
    const int nd = (cfg.tv_mode == GC_TV_3D) ? 3 : 2;
    DevBuf inv;
    GC_HIP(ctx, inv.alloc(sizeof(int32_t) * ngrid));
    GC_HIP(ctx, run[i].nbr.alloc(sizeof(int32_t) * 2 * nd * nv));
    hipLaunchKernelGGL(k_fill_i32, dim3(nblocks(ngrid)), dim3(kBlock), 0, ctx->stream, inv.as<int32_t>(), ngrid,
                       int32_t(-1));
    hipLaunchKernelGGL(k_scatter_index, dim3(nblocks(nv)), dim3(kBlock), 0, ctx->stream, v.vox_dev, nv,
                       inv.as<int32_t>());
    hipLaunchKernelGGL(k_neighbours, dim3(nblocks(nv)), dim3(kBlock), 0, ctx->stream, v.vox_dev, nv,
                       inv.as<int32_t>(), int32_t(X), int32_t(Y), int32_t(Z), int32_t(nd),
                       run[i].nbr.as<int32_t>());
    GC_HIP(ctx, hipGetLastError());
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // inv is freed on return
    return GC_OK;
  }

  // Mesh tables of a surface volume (GC_REG_SURFACE_TV), built on the host
  // from the caller's grid-form neighbours and edge lengths.  Entries whose
  // term is zero whatever theta is are left out: a vertex listing itself
  // (theta_i - theta_i = 0), and edge lengths whose reciprocal is zero or
  // not finite (0 -> the reference zeroes the non-finite quotient; inf -> 0).
  int build_surface(int i) {
    const gc_volume& v = vols[i];
    VolRun& R = run[i];
    const int64_t nv = v.nv;
    const int64_t ngrid = int64_t(v.dims[0]) * v.dims[1] * v.dims[2];
    const int nb = v.surf_max_nb;
    if (!v.vox_dev || !v.surf_nbr_dev || !v.surf_len_dev || nb < 1 || v.dims[0] < 1 || v.dims[1] < 1 ||
        v.dims[2] < 1)
      return set_err(ctx, GC_ERR_ARG, "volume %d: surface TV needs vox_dev, dims, surf_nbr_dev, surf_len_dev and "
                                      "surf_max_nb >= 1", i);
    if (ngrid >= (int64_t(1) << 31) || nv > ngrid)
      return set_err(ctx, GC_ERR_ARG, "volume %d: vertex grid of %lld unsupported", i, (long long)ngrid);
    std::vector<int64_t> hv(nv);
    std::vector<int32_t> hn(size_t(nb) * ngrid);
    std::vector<float> hl(size_t(nb) * ngrid);
    GC_HIP(ctx, hipMemcpyAsync(hv.data(), v.vox_dev, sizeof(int64_t) * nv, hipMemcpyDeviceToHost, ctx->stream));
    GC_HIP(ctx, hipMemcpyAsync(hn.data(), v.surf_nbr_dev, sizeof(int32_t) * hn.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    GC_HIP(ctx, hipMemcpyAsync(hl.data(), v.surf_len_dev, sizeof(float) * hl.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> inv(ngrid, -1);
    for (int64_t k = 0; k < nv; ++k) {
      if (hv[k] < 0 || hv[k] >= ngrid || (k > 0 && hv[k] <= hv[k - 1]))
        return set_err(ctx, GC_ERR_ARG, "volume %d: vox must be increasing grid indices (entry %lld)", i,
                       (long long)k);
      inv[hv[k]] = int32_t(k);
    }
    std::vector<std::vector<std::pair<int32_t, float>>> fw(nv), rv(nv);
    for (int64_t k = 0; k < nv; ++k) {
      const int64_t g = hv[k];
      for (int n = 0; n < nb; ++n) {
        const int32_t j = hn[size_t(n) * ngrid + g];
        if (j < 0 || j >= ngrid)
          return set_err(ctx, GC_ERR_ARG, "volume %d: surface neighbour %d of vertex %lld is outside the grid", i,
                         j, (long long)g);
        if (j == g) continue;
        const double ie = 1.0 / double(hl[size_t(n) * ngrid + g]);
        if (!std::isfinite(ie) || ie == 0.0) continue;
        const int32_t u = inv[j];
        fw[k].push_back({u, float(ie)});
        if (u >= 0) rv[u].push_back({int32_t(k), float(ie)});
      }
    }
    R.snf = R.snr = 0;
    for (int64_t k = 0; k < nv; ++k) {
      R.snf = std::max<int>(R.snf, int(fw[k].size()));
      R.snr = std::max<int>(R.snr, int(rv[k].size()));
    }
    auto upload = [&](const std::vector<std::vector<std::pair<int32_t, float>>>& L, int width, int32_t pad,
                      DevBuf& di, DevBuf& de) -> int {
      if (width == 0) return GC_OK;
      std::vector<int32_t> idx(size_t(width) * nv, pad);
      std::vector<float> ie(size_t(width) * nv, 0.f);
      for (int64_t k = 0; k < nv; ++k)
        for (size_t n = 0; n < L[k].size(); ++n) {
          idx[n * nv + k] = L[k][n].first;
          ie[n * nv + k] = L[k][n].second;
        }
      GC_HIP(ctx, di.alloc(sizeof(int32_t) * idx.size()));
      GC_HIP(ctx, de.alloc(sizeof(float) * ie.size()));
      GC_HIP(ctx, hipMemcpyAsync(di.p, idx.data(), sizeof(int32_t) * idx.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
      GC_HIP(ctx, hipMemcpyAsync(de.p, ie.data(), sizeof(float) * ie.size(), hipMemcpyHostToDevice, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the host vectors go out of scope
      return GC_OK;
    };
    int rc = upload(fw, R.snf, -2, R.snb, R.sie);
    if (rc) return rc;
    return upload(rv, R.snr, -1, R.srv, R.srie);
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
        // (a zero step gives 1 / ls = FLT_MAX: a zero update, not 0 * inf)
        const double ls = lr * std::sqrt(bc2) / bc1 * (1.0 - cfg.beta1) / std::sqrt(1.0 - cfg.beta2);
        const double fmax = double(FLT_MAX);
        s.inv_ls = float(ls != 0.0 ? std::max(-fmax, std::min(fmax, 1.0 / ls)) : fmax);
      }
    } else {
      s.lr = float(lr);
    }
    return s;
  }

  // Adam's folded effective form (adam_eff_den) for this call: the form,
  // beta2 < 1, at most kEpsSlots parameters and a positive finite
  // rs_p = (ub_p - lb_p) inv_count (as the kernel forms it) for every
  // parameter at the given inverse counts
  template <class Model>
  bool adam_fold(const float* inv_counts, int n) const {
    if (cfg.transform != GC_TRANSFORM_SIGMOID || cfg.optimiser != GC_OPT_ADAM ||
        cfg.adam_form == GC_ADAM_KINGMA_BA || !(float(cfg.beta2) < 1.f) || Model::NP > kEpsSlots)
      return false;
    for (int i = 0; i < n; ++i)
      for (int p = 0; p < Model::NP; ++p) {
        const float rs = (md.ub[p] - md.lb[p]) * inv_counts[i];
        if (!(rs > 0.f) || !std::isfinite(rs)) return false;
      }
    return true;
  }
  // eps_s / (ls rs_p) of the folded Adam step, per parameter
  void fold_eps(EpochScal& s, float inv_count, int np) const {
    const double eps_s = cfg.eps / std::sqrt(1.0 - cfg.beta2);
    for (int p = 0; p < np && p < kEpsSlots; ++p) {
      const float rs = (md.ub[p] - md.lb[p]) * inv_count;
      s.eps_ls[p] = float(std::min(double(FLT_MAX), eps_s * double(s.inv_ls) / double(rs)));
    }
  }

  OptParams opt_params() const {
    OptParams o{};
    o.opt = cfg.optimiser;
    o.adam_form = cfg.adam_form;
    o.beta1 = float(cfg.beta1);
    o.beta2 = float(cfg.beta2);
    o.eps = float(cfg.eps);
    o.momentum = float(cfg.momentum);
    o.rms_decay = float(cfg.rms_decay);
    o.rms_eps = float(cfg.rms_eps);
    return o;
  }

  // an event pair for the next timed dispatch (created once, reused)
  int next_events(hipEvent_t& a, hipEvent_t& b) {
    if (ev_used == ev_start.size()) {
      hipEvent_t ea, eb;
      GC_HIP(ctx, hipEventCreate(&ea));
      GC_HIP(ctx, hipEventCreate(&eb));
      ev_start.push_back(ea);
      ev_stop.push_back(eb);
    }
    a = ev_start[ev_used];
    b = ev_stop[ev_used];
    ++ev_used;
    return GC_OK;
  }

  // FP64 sums of the block partials of every volume that ran, folded into
  // d_rows ([vol][K*kSum], zero for the volumes that did not run)
  int reduce_partials(const std::vector<int>& n_ep) {
    const int npv = npv_count();
    for (int i = 0; i < n_vol; ++i) {
      const int nrows = n_ep[i] > 0 ? n_ep[i] * npv : 0;
      double* chunk_sums = d_sums.as<double>() + size_t(i) * K * kSum * max_chunks;
      if (nrows > 0) {
        launch_reduce_rows(run[i].partials.as<float>(), int64_t(run[i].nprod), nrows, run[i].chunks, chunk_sums,
                           ctx->stream);
        GC_HIP(ctx, hipGetLastError());
      }
      const int nout = K * kSum;
      hipLaunchKernelGGL(k_fold_chunks, dim3(nblocks(nout)), dim3(kBlock), 0, ctx->stream, chunk_sums,
                         run[i].chunks, nrows, npv, kSum, nout, d_rows.as<double>() + size_t(i) * K * kSum);
      GC_HIP(ctx, hipGetLastError());
    }
    return GC_OK;
  }

  // Launch one block of iterations for the volumes with n_ep[i] > 0.
  // in_set/out_set select state sets; zl selects the z_last buffer (-1 none).
  template <class Model, int LOSS, bool HAS_W>
  int launch(const std::vector<int>& e0, const std::vector<int>& n_ep, const std::vector<int>& in_set,
             const std::vector<int>& out_set, const std::vector<int>& zl, int iteration_cap) {
    if (step_mode) return launch_steps<Model, LOSS, HAS_W>(e0, n_ep, in_set, out_set, zl, iteration_cap);
    if constexpr (Model::kStreamOnly) {
      return set_err(ctx, GC_ERR_UNSUPPORTED, "this model runs in the per-iteration form only");
    } else {
      return launch_blocked<Model, LOSS, HAS_W>(e0, n_ep, in_set, out_set, zl, iteration_cap);
    }
  }

  template <class Model, int LOSS, bool HAS_W>
  int launch_blocked(const std::vector<int>& e0, const std::vector<int>& n_ep, const std::vector<int>& in_set,
                     const std::vector<int>& out_set, const std::vector<int>& zl, int iteration_cap) {
    // the pinned descriptor buffers are reused: wait for the copies of the
    // previous launch before overwriting them
    if (desc_in_flight) GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    VolDev* hv = h_vols.as<VolDev>();
    EpochScal* hes = h_es.as<EpochScal>();
    bool want_signs = false;
    int blk0 = 0;
    std::vector<float> invc(n_vol);
    for (int i = 0; i < n_vol; ++i) invc[i] = float(1.0 / (double(md.n_meas) * run[i].nmean));
    const bool fold = adam_fold<Model>(invc.data(), n_vol);
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
      d.inv_count = invc[i];
      d.signs = vols[i].sign_trace_dev;
      if (d.signs) want_signs = true;
      hv[i] = d;
      for (int k = 0; k < n_ep[i]; ++k) {
        EpochScal& es = hes[size_t(i) * K + k];
        es = epoch_scal(e0[i] + k);
        if (fold) fold_eps(es, d.inv_count, Model::NP);
      }
      blk0 += R.nblk;
      if (n_ep[i] > 0) R.launches++;
    }
    GC_HIP(ctx, hipMemcpyAsync(d_vols.p, h_vols.p, sizeof(VolDev) * n_vol, hipMemcpyHostToDevice, ctx->stream));
    GC_HIP(ctx, hipMemcpyAsync(d_es.p, h_es.p, sizeof(EpochScal) * K * n_vol, hipMemcpyHostToDevice, ctx->stream));
    desc_in_flight = true;
    AskParams ap{};
    ap.vols = d_vols.as<VolDev>();
    ap.n_vol = n_vol;
    ap.transform = cfg.transform;
    ap.need_norms = npv_count() == 4 ? 1 : 0;
    ap.adam_eff = fold ? 1 : 0;
    ap.o = opt_params();
    hipEvent_t ev_a, ev_b;
    int rce = next_events(ev_a, ev_b);
    if (rce) return rce;
    // events attached to the dispatch itself time the kernel alone
    const int rcl = launch_fused<Model, LOSS, HAS_W>(want_signs, total_blk, blk_threads, ap, ev_a, ev_b);
    if (rcl) return rcl;
    ++kernel_launches;
    return reduce_partials(n_ep);
  }

  // k_askadam over nblk256 blocks of kBlock voxel slots, in blocks of bs
  // threads (64, 128 or 256), with the wavefronts' LDS rings when only the
  // loss is summed
  template <class Model, int LOSS, bool HAS_W>
  int launch_fused(bool signs, int nblk256, int bs, const AskParams& ap, hipEvent_t ev_a, hipEvent_t ev_b) {
    if (bs != 64 && bs != 128 && bs != 256) return set_err(ctx, GC_ERR_ARG, "askadam block of %d threads", bs);
    const dim3 grid(unsigned(int64_t(nblk256) * (kBlock / bs)));
    const size_t lds = ap.need_norms ? 0 : size_t(bs / 64) * kRing * 64 * sizeof(float);
    if constexpr (LOSS == GC_LOSS_L1) {
      if (signs)
        hipExtLaunchKernelGGL((k_askadam<Model, LOSS, HAS_W, true>), grid, dim3(bs), lds, ctx->stream, ev_a, ev_b, 0,
                              dmd, ap);
      else
        hipExtLaunchKernelGGL((k_askadam<Model, LOSS, HAS_W>), grid, dim3(bs), lds, ctx->stream, ev_a, ev_b, 0, dmd,
                              ap);
    } else {
      (void)signs;
      hipExtLaunchKernelGGL((k_askadam<Model, LOSS, HAS_W>), grid, dim3(bs), lds, ctx->stream, ev_a, ev_b, 0, dmd,
                            ap);
    }
    GC_HIP(ctx, hipGetLastError());
    return GC_OK;
  }

  // Autotuner of the blocked kernel (run once per model variant and
  // context, before the call's buffers exist): times k_askadam at each block
  // size on the first min(nv, 2^20) voxels of volume 0, reading the call's
  // data and starting point and writing only buffers of its own, and the
  // fixed cost of one launch cycle (launch, k_reduce_rows, copy of the sums
  // to the host, synchronisation).  Strides by the trial's voxel count keep
  // every read inside the call's arrays.  Block size and K change no result:
  // every wavefront sums its own 64 voxels, and the host replays any block.
  template <class Model, int LOSS, bool HAS_W>
  int tune_blocked() {
    if constexpr (Model::kStreamOnly) {
      return GC_OK;
    } else {
      const int npv = npv_count();
      const std::array<int64_t, 8> key{md.model_id, md.flags, md.n_meas, md.n_par,
                                       LOSS,        HAS_W ? 1 : 0, npv, cfg.transform};
      const int fixed = ctx->ask_block;
      auto it = ctx->ask_tuned.find(key);
      if (it != ctx->ask_tuned.end()) {
        blk_threads = fixed ? fixed : it->second.threads;
        tune_launch_ms = it->second.launch_ms;
        tune_vox_ms = it->second.ms_per_voxel_iter;
        return GC_OK;
      }
      constexpr int kTrialMax = 64;  // iterations of the longest trial launch
      const int64_t nvt = std::min<int64_t>(vols[0].nv, int64_t(1) << 20);
      const int nbt = int((nvt + kBlock - 1) / kBlock);
      const int P = md.n_par;
      const size_t pbytes = sizeof(float) * size_t(P) * size_t(nvt);
      DevBuf zt, mt, vt, pt, est, dvt, dsum;
      GC_HIP(ctx, zt.alloc(pbytes));
      GC_HIP(ctx, mt.alloc(pbytes));
      GC_HIP(ctx, vt.alloc(pbytes));
      GC_HIP(ctx, pt.alloc(sizeof(float) * size_t(kTrialMax) * npv * nbt * (kBlock / 64)));
      GC_HIP(ctx, est.alloc(sizeof(EpochScal) * kTrialMax));
      GC_HIP(ctx, dvt.alloc(sizeof(VolDev)));
      const int chunks = int((int64_t(nbt) * (kBlock / 64) + kReduceChunk - 1) / kReduceChunk);
      GC_HIP(ctx, dsum.alloc(sizeof(double) * npv * chunks));
      GC_HIP(ctx, hipMemsetAsync(mt.p, 0, pbytes, ctx->stream));
      GC_HIP(ctx, hipMemsetAsync(vt.p, 0, pbytes, ctx->stream));
      const float inv_trial = float(1.0 / (double(md.n_meas) * double(nvt)));
      const bool fold = adam_fold<Model>(&inv_trial, 1);
      std::vector<EpochScal> hes(kTrialMax);
      for (int k = 0; k < kTrialMax; ++k) {
        hes[k] = epoch_scal(1 + k);
        if (fold) fold_eps(hes[k], inv_trial, Model::NP);
      }
      GC_HIP(ctx, hipMemcpyAsync(est.p, hes.data(), sizeof(EpochScal) * kTrialMax, hipMemcpyHostToDevice,
                                 ctx->stream));
      VolDev d{};
      d.y = vols[0].y_dev;
      d.w = HAS_W ? (vols[0].w_dev ? vols[0].w_dev : vols[0].y_dev) : nullptr;  // any finite weights
      d.ex = vols[0].extra_dev;
      d.z_in = vols[0].z_dev;
      d.m_in = mt.as<float>();
      d.v_in = vt.as<float>();
      d.z_out = zt.as<float>();
      d.m_out = mt.as<float>();
      d.v_out = vt.as<float>();
      d.partials = pt.as<float>();
      d.es = est.as<EpochScal>();
      d.nv = nvt;
      d.blk0 = 0;
      d.nblk = nbt;
      d.epoch0 = 1;
      d.iteration = INT32_MAX;
      d.inv_count = inv_trial;
      AskParams ap{};
      ap.vols = dvt.as<VolDev>();
      ap.n_vol = 1;
      ap.transform = cfg.transform;
      ap.need_norms = npv == 4 ? 1 : 0;
      ap.adam_eff = fold ? 1 : 0;
      ap.o = opt_params();
      hipEvent_t e0, e1;
      GC_HIP(ctx, hipEventCreate(&e0));
      GC_HIP(ctx, hipEventCreate(&e1));
      std::vector<double> hsum(size_t(npv) * chunks);
      // one trial launch of n iterations at bs threads: kernel ms, and with
      // cycle = true the wall time of the whole launch cycle
      auto trial = [&](int bs, int n, bool cycle, float* ms, double* wall) -> int {
        d.n_epochs = n;
        const auto t0 = std::chrono::steady_clock::now();
        GC_HIP(ctx, hipMemcpyAsync(dvt.p, &d, sizeof(VolDev), hipMemcpyHostToDevice, ctx->stream));
        int e = launch_fused<Model, LOSS, HAS_W>(false, nbt, bs, ap, e0, e1);
        if (e) return e;
        if (cycle) {
          launch_reduce_rows(pt.as<float>(), int64_t(nbt) * (kBlock / 64), n * npv, chunks, dsum.as<double>(),
                             ctx->stream);
          GC_HIP(ctx, hipGetLastError());
          GC_HIP(ctx, hipMemcpyAsync(hsum.data(), dsum.p, sizeof(double) * hsum.size(), hipMemcpyDeviceToHost,
                                     ctx->stream));
        }
        GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (wall) *wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        GC_HIP(ctx, hipEventElapsedTime(ms, e0, e1));
        return GC_OK;
      };
      int rc = GC_OK;
      std::vector<int> cand;
      if (fixed) cand.push_back(fixed);
      else cand = {64, 128, 256};
      float ms = 0.f;
      int n = 4;
      rc = trial(cand.back(), n, false, &ms, nullptr);  // also loads the kernel
      while (!rc && ms < 1.0f && n < kTrialMax) {
        n = std::min(kTrialMax, 2 * n);
        rc = trial(cand.back(), n, false, &ms, nullptr);
      }
      int best = cand.back();
      float best_ms = ms;
      for (size_t c = 0; !rc && c + 1 < cand.size(); ++c) {
        rc = trial(cand[c], n, false, &ms, nullptr);  // warm-up of this launch shape
        if (!rc) rc = trial(cand[c], n, false, &ms, nullptr);
        if (!rc && ms < best_ms) {
          best_ms = ms;
          best = cand[c];
        }
      }
      // the launch cycle of one iteration: wall time less kernel time,
      // median of three
      std::vector<double> cyc;
      for (int r = 0; !rc && r < 3; ++r) {
        double wall = 0.0;
        rc = trial(best, 1, true, &ms, &wall);
        cyc.push_back(std::max(0.0, wall - double(ms)));
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (rc) return rc;
      std::sort(cyc.begin(), cyc.end());
      gc_ctx::AskTune t{best, std::max(cyc[1], 1e-3), double(best_ms) / (double(n) * double(nvt))};
      ctx->ask_tuned[key] = t;
      blk_threads = fixed ? fixed : best;
      tune_launch_ms = t.launch_ms;
      tune_vox_ms = t.ms_per_voxel_iter;
      return GC_OK;
    }
  }

  // The step form: one k_askadam_step launch per iteration.  Iteration k of
  // volume i reads state S_k and writes S_k+1 with S_0 = in_set, S_n = out_set
  // and the states between in the ping-pong sets, so the input set survives
  // for a replay.
  template <class Model, int LOSS, bool HAS_W>
  int launch_steps(const std::vector<int>& e0, const std::vector<int>& n_ep, const std::vector<int>& in_set,
                   const std::vector<int>& out_set, const std::vector<int>& zl, int iteration_cap) {
    if (desc_in_flight) GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int nmax = 0;
    for (int i = 0; i < n_vol; ++i) nmax = std::max(nmax, n_ep[i]);
    StepVol* hs = h_steps.as<StepVol>();
    for (int k = 0; k < nmax; ++k) {
      int blk0 = 0;
      for (int i = 0; i < n_vol; ++i) {
        VolRun& R = run[i];
        StepVol d{};
        d.y = vols[i].y_dev;
        d.w = HAS_W ? (vols[i].w_dev ? vols[i].w_dev : R.ones.as<float>()) : nullptr;
        d.ex = vols[i].extra_dev;
        d.nv = R.nv;
        d.blk0 = blk0;
        d.nblk = R.nblk;
        d.active = k < n_ep[i] ? 1 : 0;
        if (d.active) {
          const int n = n_ep[i];
          DevBuf* src = (k == 0) ? R.set[std::max(in_set[i], 0)] : R.pp[(k - 1) & 1];
          DevBuf* dst = (k == n - 1) ? R.set[std::max(out_set[i], 0)] : R.pp[k & 1];
          d.z_in = src[0].as<float>();
          d.m_in = src[1].as<float>();
          d.v_in = src[2].as<float>();
          d.z_out = dst[0].as<float>();
          d.m_out = dst[1].as<float>();
          d.v_out = dst[2].as<float>();
          d.z_last = (k == n - 1 && zl[i] >= 0) ? R.zlast[zl[i]].as<float>() : nullptr;
          d.partials = R.partials.as<float>();
          d.nbr = use_tv ? R.nbr.as<int32_t>() : nullptr;
          d.aux = use_epgx ? R.aux.as<float>() : nullptr;
          if (use_stv) {
            d.snb = R.snb.as<int32_t>();
            d.sie = R.sie.as<float>();
            d.srv = R.srv.as<int32_t>();
            d.srie = R.srie.as<float>();
            d.snf = R.snf;
            d.snr = R.snr;
          }
          d.prior = prior ? vols[i].prior_dev : nullptr;
          if (robust) {
            const int e = e0[i] + k;
            // update_outlier_mask returns modulated weights only on a check epoch
            if (k == 0 && e - 1 >= 1 && (e - 1) % cfg.weight_update_interval == 0)
              d.wmod = (replaying && replay_pinned && R.pin_mod) ? R.wmod_pin.as<float>() : R.wmod.as<float>();
            d.mconv = R.mconv.as<uint8_t>();
            if (!replaying) d.pvl_out = (k == n - 1) ? R.pvl_d.as<float>() : (e == 1 ? R.pvl0_d.as<float>() : nullptr);
          }
          d.k = k;
          d.n_epochs = n;
          d.update = (e0[i] + k < iteration_cap) ? 1 : 0;
          d.inv_count = float(1.0 / (double(md.n_meas) * R.nmean));
          d.inv_nv = float(1.0 / R.nmean);
          d.es = epoch_scal(e0[i] + k);
        }
        hs[size_t(k) * n_vol + i] = d;
        blk0 += R.nblk;
      }
    }
    for (int i = 0; i < n_vol; ++i)
      if (n_ep[i] > 0) run[i].launches++;
    GC_HIP(ctx, hipMemcpyAsync(d_steps.p, h_steps.p, sizeof(StepVol) * size_t(nmax) * n_vol,
                               hipMemcpyHostToDevice, ctx->stream));
    desc_in_flight = true;
    StepParams sp{};
    sp.n_vol = n_vol;
    sp.n_reg = use_reg ? cfg.n_reg : 0;
    sp.nd = (cfg.tv_mode == GC_TV_3D) ? 3 : 2;
    sp.reg_type = prior ? GC_REG_PRIOR : (use_stv ? GC_REG_SURFACE_TV : GC_REG_TV);
    for (int k = 0; k < GC_MAX_REG; ++k) {
      sp.regmap[k] = k < cfg.n_reg ? cfg.regmap[k] : 0;
      sp.lam[k] = k < cfg.n_reg ? float(cfg.lambda[k]) : 0.f;
    }
    for (int d = 0; d < 3; ++d) sp.inv_vs[d] = float(1.0 / cfg.voxel_size[d]);
    sp.o = opt_params();
    if (ctx->use_graphs && nmax >= 2) {
      StepVol* const d0 = d_steps.as<StepVol>();
      auto launch_k = [&](int j, hipStream_t s) {
        StepParams spj = sp;
        spj.vols = d0 + size_t(j) * n_vol;
        if (use_epgx) (void)epgx_step(ctx, epgx, dmd, spj, unsigned(total_blk), cfg.transform, s);
        if (cfg.transform == GC_TRANSFORM_LINEAR)
          hipLaunchKernelGGL((k_askadam_step<Model, LOSS, HAS_W, GC_TRANSFORM_LINEAR>), dim3(total_blk),
                             dim3(kBlock), 0, s, dmd, spj);
        else
          hipLaunchKernelGGL((k_askadam_step<Model, LOSS, HAS_W, GC_TRANSFORM_SIGMOID>), dim3(total_blk),
                             dim3(kBlock), 0, s, dmd, spj);
      };
      const int G = std::min(nmax, kGraphSteps);
      hipEvent_t ev_a, ev_b;
      int rce = next_events(ev_a, ev_b);
      if (rce) return rce;
      GC_HIP(ctx, hipEventRecord(ev_a, ctx->stream));
      for (int c0 = 0; c0 < nmax; c0 += G) {
        const int n = std::min(G, nmax - c0);
        hipGraphExec_t ex = nullptr;
        rce = step_graph(n, launch_k, &ex);
        if (rce) return rce;
        if (c0 > 0)
          GC_HIP(ctx, hipMemcpyAsync(d0, d0 + size_t(c0) * n_vol, sizeof(StepVol) * size_t(n) * n_vol,
                                     hipMemcpyDeviceToDevice, ctx->stream));
        GC_HIP(ctx, hipGraphLaunch(ex, ctx->stream));
        kernel_launches += n;
      }
      GC_HIP(ctx, hipEventRecord(ev_b, ctx->stream));
      return reduce_partials(n_ep);
    }
    for (int k = 0; k < nmax; ++k) {
      sp.vols = d_steps.as<StepVol>() + size_t(k) * n_vol;
      if (use_epgx) {
        const int rcx = epgx_step(ctx, epgx, dmd, sp, unsigned(total_blk), cfg.transform, ctx->stream);
        if (rcx) return rcx;
      }
      hipEvent_t ev_a, ev_b;
      int rce = next_events(ev_a, ev_b);
      if (rce) return rce;
      if (cfg.transform == GC_TRANSFORM_LINEAR)
        hipExtLaunchKernelGGL((k_askadam_step<Model, LOSS, HAS_W, GC_TRANSFORM_LINEAR>), dim3(total_blk),
                              dim3(kBlock), 0, ctx->stream, ev_a, ev_b, 0, dmd, sp);
      else
        hipExtLaunchKernelGGL((k_askadam_step<Model, LOSS, HAS_W, GC_TRANSFORM_SIGMOID>), dim3(total_blk),
                              dim3(kBlock), 0, ctx->stream, ev_a, ev_b, 0, dmd, sp);
      GC_HIP(ctx, hipGetLastError());
      ++kernel_launches;
    }
    return reduce_partials(n_ep);
  }

  // One iteration of the host-side loop logic (utils/askadam.m:264-292 minus
  // the update itself).  Returns true when the volume stops at `epoch`.
  // loss_c: the loss the convergence signal follows (the full loss, or with
  // robustConvergence the mean per-voxel loss over the convergence set)
  bool epoch_logic(VolRun& R, int epoch, double sum_l, double sum_step_prev, double sum_den, double sum_g,
                   double sum_reg, double* trace, const double* loss_c_robust = nullptr) {
    const double count = double(md.n_meas) * R.nmean;
    R.fid = sum_l / count;
    R.reg = sum_reg / R.nmean + R.reg_offset;  // / Nmasked (+ the prior's out-of-mask constant)
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
    const double loss_c = loss_c_robust ? *loss_c_robust : R.loss;
    if (cfg.convergence_model == GC_CONV_EMA) {
      const double e_new = cfg.ema_decay * R.ema + (1.0 - cfg.ema_decay) * loss_c;
      R.conv = std::fabs(e_new - R.ema) / (std::fabs(R.ema) + 1e-8);
      R.ema = e_new;
    } else {
      std::rotate(R.cbuf.begin(), R.cbuf.begin() + 1, R.cbuf.end());
      R.cbuf.back() = loss_c;
      const int n = int(R.cbuf.size());
      double sx = 0, sy = 0, sxx = 0, sxy = 0;
      for (int i = 0; i < n; ++i) {
        const double xi = i + 1;
        sx += xi; sy += R.cbuf[i]; sxx += xi * xi; sxy += xi * R.cbuf[i];
      }
      const double den = n * sxx - sx * sx;
      R.conv = -((den != 0) ? (n * sxy - sx * sy) / den : 0.0);
    }
    if (R.reprime > 0) { --R.reprime; R.pc = 0; }
    else if (R.conv > cfg.convergence_value || epoch <= cfg.convergence_window) R.pc = 0;
    else if (!robust && (R.min_loss - R.loss) > cfg.convergence_value) R.pc = 0;
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

  // robustConvergence at a block end `epoch`: update_outlier_mask
  // (utils/askadam.m:448-510) on a check epoch, the convergence-set refresh
  // on a convergenceWindow boundary (:541-566), and the mean per-voxel loss
  // over the convergence set.  The new weight modulator and set are
  // uploaded for the next block.
  int robust_block_end(int i, int epoch, double& loss_c) {
    VolRun& R = run[i];
    const int64_t nv = R.nv;
    const int W = cfg.outlier_check_window;
    bool mod_changed = false;
    if (epoch % cfg.weight_update_interval == 0) {
      std::rotate(R.hist.begin(), R.hist.begin() + nv, R.hist.end());
      std::copy(R.pvl.begin(), R.pvl.end(), R.hist.end() - nv);
      std::vector<double> impA(nv), impB(nv);
      for (int64_t v = 0; v < nv; ++v) {
        const double h0 = R.hist[v], h1 = R.hist[size_t(W - 1) * nv + v];
        impA[v] = (h0 - h1) / (h0 + 1e-8);
        impB[v] = (R.pvl_init[v] - R.pvl[v]) / (R.pvl_init[v] + 1e-8);
      }
      std::vector<double> a = impA, b = impB;
      const double medA = median_of(a), medB = median_of(b);
      for (int64_t v = 0; v < nv; ++v) {
        const bool failA = impA[v] < cfg.outlier_voxel_thres && medA > cfg.outlier_pop_thres;
        const bool failB = impB[v] < cfg.outlier_init_thres && medB > cfg.outlier_init_pop_thres;
        if (failA && failB) R.flag_count[v] += 1;
        else R.flag_count[v] = std::max(R.flag_count[v] - 1, 0);
        R.main_mask[v] = R.flag_count[v] == 0;
      }
      mod_changed = true;
    }
    if (epoch % cfg.convergence_window == 0 && R.main_mask != R.mask_conv) {
      R.mask_conv = R.main_mask;
      R.reprime = cfg.convergence_window;
      R.n_conv = 0;
      for (int64_t v = 0; v < nv; ++v) R.n_conv += R.mask_conv[v];
      GC_HIP(ctx, hipMemcpyAsync(R.mconv.p, R.mask_conv.data(), nv, hipMemcpyHostToDevice, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    double acc = 0.0;
    for (int64_t v = 0; v < nv; ++v)
      if (R.mask_conv[v]) acc += R.pvl[v];
    loss_c = R.n_conv ? acc / double(R.n_conv) : std::nan("");
    if (mod_changed) {
      std::vector<float> m(nv);
      for (int64_t v = 0; v < nv; ++v) m[v] = R.main_mask[v] ? 1.0f : float(cfg.outlier_weight);
      GC_HIP(ctx, hipMemcpyAsync(R.wmod.p, m.data(), sizeof(float) * nv, hipMemcpyHostToDevice, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    }
    return GC_OK;
  }

  // per-voxel mean losses of a finished block (block end; epoch 1 as well)
  int fetch_pvl(int i, bool first) {
    VolRun& R = run[i];
    std::vector<float> h(R.nv);
    GC_HIP(ctx, hipMemcpyAsync(h.data(), R.pvl_d.p, sizeof(float) * R.nv, hipMemcpyDeviceToHost, ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    R.pvl.assign(h.begin(), h.end());
    if (first) {
      // starting-point evaluation = epoch 1 (written at the block end when K = 1)
      if (K > 1) {
        GC_HIP(ctx, hipMemcpyAsync(h.data(), R.pvl0_d.p, sizeof(float) * R.nv, hipMemcpyDeviceToHost, ctx->stream));
        GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      }
      R.pvl_init.assign(h.begin(), h.end());
      R.hist.resize(size_t(cfg.outlier_check_window) * R.nv);
      for (int k = 0; k < cfg.outlier_check_window; ++k)
        std::copy(R.pvl_init.begin(), R.pvl_init.end(), R.hist.begin() + size_t(k) * R.nv);
    }
    return GC_OK;
  }

  template <class Model, int LOSS, bool HAS_W>
  int run_all() {
    const int P = md.n_par;
    const int iter = cfg.iteration;
    const double* sums = row_sums.data();
    const size_t sum_bytes = sizeof(double) * K * kSum * n_vol;
    std::vector<int> e0(n_vol), nep(n_vol), in(n_vol), out(n_vol), zl(n_vol);

    if (iter <= 0) {
      // evaluation only (epoch 0): run one non-updating iteration
      for (int i = 0; i < n_vol; ++i) { e0[i] = 1; nep[i] = 1; in[i] = 0; out[i] = 1; zl[i] = 0; }
      int rc = launch<Model, LOSS, HAS_W>(e0, nep, in, out, zl, 0);
      if (rc) return rc;
      GC_HIP(ctx, hipMemcpyAsync(h_rows.p, d_rows.p, sum_bytes, hipMemcpyDeviceToHost, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      fold_sums();
      for (int i = 0; i < n_vol; ++i) {
        VolRun& R = run[i];
        const double* s = sums + size_t(i) * K * kSum;
        R.fid = s[0] / (double(md.n_meas) * R.nmean);
        R.reg = s[4] / R.nmean + R.reg_offset;
        R.loss = R.fid + R.reg;
        R.min_loss = R.loss; R.min_fid = R.fid; R.min_reg = R.reg; R.min_iter = 0; R.last_epoch = 0;
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
      GC_HIP(ctx, hipMemcpyAsync(h_rows.p, d_rows.p, sum_bytes, hipMemcpyDeviceToHost, ctx->stream));
      GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
      desc_in_flight = false;
      collect_events();
      fold_sums();
      any = false;
      for (int i = 0; i < n_vol; ++i) {
        VolRun& R = run[i];
        if (!R.active) continue;
        const double* s = sums + size_t(i) * K * kSum;
        double* trace = vols[i].loss_trace_host;
        if (robust) {
          const int rcp = fetch_pvl(i, !R.have_init);
          if (rcp) return rcp;
        }
        if (!R.have_init) {
          // starting-point evaluation (:197-219) equals the first iteration's
          R.fid = s[0] / (double(md.n_meas) * R.nmean);
          R.reg = s[4] / R.nmean + R.reg_offset;
          const double l0 = R.fid + R.reg;
          R.loss = l0;
          R.min_loss = l0; R.min_fid = R.fid; R.min_reg = R.reg; R.min_iter = 0;
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
          const double* sk = s + size_t(k) * kSum;
          const double step_prev = (k == 0) ? R.carry_step : sk[1 - kSum];
          double loss_c = 0.0;
          if (robust) {
            if (k == nep[i] - 1) {
              const int rcr = robust_block_end(i, epoch, loss_c);
              if (rcr) return rcr;
            } else {
              loss_c = R.n_conv ? sk[5] / double(R.n_conv) : std::nan("");
            }
          }
          stopped = epoch_logic(R, epoch, sk[0], step_prev, sk[2], sk[3], sk[4], trace, robust ? &loss_c : nullptr);
          if (R.min_loss < prev_min) {
            if (k == nep[i] - 1) {
              R.min_zlast = zl[i];
              R.pinned = -1;
            } else {
              R.min_zlast = -1;
              R.pinned = in[i]; R.pin_e0 = e0[i]; R.pin_r = k;
              if (robust) {
                // keep the weight modulator this block started with for its replay
                const int e = e0[i];
                R.pin_mod = e - 1 >= 1 && (e - 1) % cfg.weight_update_interval == 0;
                if (R.pin_mod)
                  GC_HIP(ctx, hipMemcpyAsync(R.wmod_pin.p, R.wmod.p, sizeof(float) * R.nv, hipMemcpyDeviceToDevice,
                                             ctx->stream));
              }
            }
          }
          if (stopped) break;
        }
        if (stopped) {
          R.active = false;
          if (k == nep[i] - 1) { R.final_from = zl[i]; }
          else { R.final_from = -1; R.final_set = in[i]; R.final_e0 = e0[i]; R.final_r = k; }
        } else {
          R.carry_step = s[size_t(nep[i] - 1) * kSum + 1];
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
    replaying = true;
    replay_pinned = (set == R.pinned);
    const int rc = launch<Model, LOSS, HAS_W>(e0, nep, in, out, zl, cfg.iteration);
    replaying = replay_pinned = false;
    return rc;
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
        if (use_epgx) {
          const int rcx = epgx_eval(ctx, epgx, dmd, zsrc, vols[i].extra_dev, R.nv, cfg.transform, R.aux.as<float>(),
                                    ctx->stream);
          if (rcx) return rcx;
          ep.aux = R.aux.as<float>();
        }
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
      vols[i].block_threads = step_mode ? kBlock : blk_threads;
    }
    return GC_OK;
  }
};


}  // namespace gc
