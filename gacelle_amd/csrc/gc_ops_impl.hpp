// gc_ops_impl.hpp -- the model-dependent operations behind ModelOps, as
// templates over the register-tiled model type.  Included only by the
// ops_*.hip translation units, each of which instantiates one model.
#pragma once
#include "gc_askadam.hpp"

namespace gc {

template <class Model>
struct OpsImpl {
  static int fwd(gc_ctx* ctx, const gc_model* dmd, const float* theta, const float* extra, float* s, int64_t nv,
                 const float* aux) {
    EvalParams ep{};
    ep.aux = aux;
    ep.x = theta;
    ep.ex = extra;
    ep.s_out = s;
    ep.nv = nv;
    ep.transform = GC_TRANSFORM_NONE;
    ep.mode = 0;
    hipLaunchKernelGGL((k_eval<Model, GC_LOSS_L2, false>), dim3(nblocks(nv)), dim3(kBlock), 0, ctx->stream, dmd, ep);
    GC_HIP(ctx, hipGetLastError());
    return GC_OK;
  }

  static int fwd_adj(gc_ctx* ctx, const gc_model* dmd, int loss, int transform, const float* x, const float* y,
                     const float* w, const float* extra, int64_t nv, float scale, float* grad, double* part,
                     float* residual, const float* aux) {
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
        ep.partial = part;
        ep.aux = aux;
        ep.nv = nv;
        ep.transform = transform;
        ep.mode = 1;
        ep.scale = scale;
        hipLaunchKernelGGL((k_eval<Model, LOSS, HW>), dim3(nblocks(nv)), dim3(kBlock), 0, ctx->stream, dmd, ep);
        GC_HIP(ctx, hipGetLastError());
        return GC_OK;
      });
    });
  }

  static int askadam(gc_ctx* ctx, const gc_model* md, const gc_askadam_cfg* cfg, gc_volume* vols, int n_vol) {
    AskDriver drv(ctx, *md, *cfg, vols, n_vol);
    drv.tuner = [](AskDriver& d) -> int {
      return with_loss(d.ctx, d.cfg.loss, [&](auto lc) -> int {
        constexpr int LOSS = decltype(lc)::value;
        return with_bool(d.has_w, [&](auto hw) -> int {
          constexpr bool HW = decltype(hw)::value;
          return d.template tune_blocked<Model, LOSS, HW>();
        });
      });
    };
    int rc = drv.setup(Model::kStreamOnly, has_aux<Model>::value);
    if (rc) return rc;
    return with_loss(ctx, cfg->loss, [&](auto lc) -> int {
      constexpr int LOSS = decltype(lc)::value;
      return with_bool(drv.has_w, [&](auto hw) -> int {
        constexpr bool HW = decltype(hw)::value;
        int r = drv.template run_all<Model, LOSS, HW>();
        if (r) return r;
        return drv.template finalize<Model, LOSS, HW>();
      });
    });
  }

  // one sampler launch with bs threads per block (bs / (rep * walkers)
  // voxels per block)
  static int mcmc(gc_ctx* ctx, const gc_model* dmd, bool gw, bool hw, const McmcParams& P, int bs,
                  hipEvent_t ev0, hipEvent_t ev1) {
    if constexpr (Model::kStreamOnly) {
      // the samplers keep a voxel's data in registers
      return set_err(ctx, GC_ERR_UNSUPPORTED, "MCMC is not available for this model");
    } else {
      return mcmc_launch(ctx, dmd, gw, hw, P, bs, ev0, ev1);
    }
  }

  static int mcmc_launch(gc_ctx* ctx, const gc_model* dmd, bool gw, bool hw, const McmcParams& P, int bs,
                         hipEvent_t ev0, hipEvent_t ev1) {
    constexpr int Q = Model::NP + 1;
    const int gsz = P.rep * (gw ? P.n_walker : 1);
    if (bs % 64 != 0 || bs > kBlock || gsz > bs)
      return set_err(ctx, GC_ERR_ARG, "sampler block of %d threads for groups of %d", bs, gsz);
    const int64_t gpb = bs / gsz;
    const unsigned grid = unsigned((P.nv + gpb - 1) / gpb);
    const size_t lds = mcmc_lds_bytes(Q, bs, gw && !gw_wave_local(gsz));
    return with_bool(hw, [&](auto hwc) -> int {
      constexpr bool HW = decltype(hwc)::value;
      if (gw)
        hipExtLaunchKernelGGL((k_gw<Model, HW>), dim3(grid), dim3(bs), lds, ctx->stream, ev0, ev1, 0, dmd, P);
      else
        hipExtLaunchKernelGGL((k_mh<Model, HW>), dim3(grid), dim3(bs), lds, ctx->stream, ev0, ev1, 0, dmd, P);
      GC_HIP(ctx, hipGetLastError());
      return GC_OK;
    });
  }

  static int gwg(gc_ctx* ctx, const gc_model* dmd, int kind, bool hw, const GwgParams& P, const GwgStep& S,
                 unsigned grid) {
    if constexpr (Model::kStreamOnly) {
      return set_err(ctx, GC_ERR_UNSUPPORTED, "MCMC is not available for this model");
    } else {
      return with_bool(hw, [&](auto hwc) -> int {
        constexpr bool HW = decltype(hwc)::value;
        if (kind == 0)
          hipLaunchKernelGGL((k_gwg_init<Model, HW>), dim3(grid), dim3(kBlock), 0, ctx->stream, dmd, P);
        else if (kind == 1)
          hipLaunchKernelGGL((k_gwg_step<Model, HW>), dim3(grid), dim3(kBlock), 0, ctx->stream, dmd, P, S);
        else
          hipLaunchKernelGGL((k_gwg_final<Model>), dim3(grid), dim3(kBlock), 0, ctx->stream, dmd, P);
        GC_HIP(ctx, hipGetLastError());
        return GC_OK;
      });
    }
  }
};

template <class Model>
const ModelOps* ops_table() {
  static const ModelOps t{&OpsImpl<Model>::fwd, &OpsImpl<Model>::fwd_adj, &OpsImpl<Model>::askadam,
                          &OpsImpl<Model>::mcmc, &OpsImpl<Model>::gwg};
  return &t;
}

}  // namespace gc
