// epgx.hip -- host side of the MCR-MWI EPG-X pre-pass (gc_epgx.hpp): the
// networks' padded device copy and the k_epgx launchers.
#include "gc_host.hpp"

namespace gc {

int epgx_prepare(gc_ctx* ctx, const gc_model& md, EpgxDev& out) {
  int wmax = 0;
  size_t count[2] = {0, 0};
  for (int n = 0; n < 2; ++n) {
    const gc_mlp& m = md.mlp[n];
    if (m.n_layer < 1 || m.n_layer > GC_MLP_MAX_LAYERS || !m.params)
      return set_err(ctx, GC_ERR_ARG, "EPG-X %s network: 1..%d layers and parameters needed", n ? "magnitude" : "phase",
                     GC_MLP_MAX_LAYERS);
    if (m.width[0] != 11 || m.width[m.n_layer] != (n ? 2 : 1))
      return set_err(ctx, GC_ERR_ARG, "EPG-X %s network: 11 inputs and %d output(s) expected",
                     n ? "magnitude" : "phase", n ? 2 : 1);
    for (int l = 0; l <= m.n_layer; ++l) {
      if (m.width[l] < 1 || m.width[l] > GC_MLP_MAX_WIDTH)
        return set_err(ctx, GC_ERR_UNSUPPORTED, "EPG-X network layer width %d outside 1..%d", m.width[l],
                       GC_MLP_MAX_WIDTH);
      wmax = std::max(wmax, m.width[l]);
      if (l > 0) count[n] += size_t(m.width[l]) * m.width[l - 1] + m.width[l];
    }
  }
  const int wm = wmax <= 16 ? 16 : (wmax <= 32 ? 32 : 64);
  const size_t per_layer = size_t(wm) * wm + wm;
  std::vector<float> pad(per_layer * (md.mlp[0].n_layer + md.mlp[1].n_layer), 0.f);
  size_t base = 0;
  for (int n = 0; n < 2; ++n) {
    const gc_mlp& m = md.mlp[n];
    std::vector<float> h(count[n]);
    // on the context's stream: the caller may have just written the parameters there
    GC_HIP(ctx, hipMemcpyAsync(h.data(), m.params, sizeof(float) * count[n], hipMemcpyDeviceToHost, ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    size_t src = 0;
    for (int l = 1; l <= m.n_layer; ++l) {
      const int ni = m.width[l - 1], no = m.width[l];
      float* dst = pad.data() + base + size_t(l - 1) * per_layer;
      // column-major [out, in] -> row j = output unit, padded to wm
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < no; ++j) dst[size_t(j) * wm + i] = h[src + j + size_t(no) * i];
      src += size_t(no) * ni;
      for (int j = 0; j < no; ++j) dst[size_t(wm) * wm + j] = h[src + j];
      src += no;
    }
    for (float v : h)
      if (!std::isfinite(v)) return set_err(ctx, GC_ERR_ARG, "EPG-X network parameters must be finite");
    base += per_layer * m.n_layer;
  }
  GC_HIP(ctx, out.wbuf.alloc(sizeof(float) * pad.size()));
  GC_HIP(ctx, hipMemcpyAsync(out.wbuf.p, pad.data(), sizeof(float) * pad.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // pad is freed on return
  out.wm = wm;
  const float* w0 = out.wbuf.as<float>();
  for (int n = 0; n < 2; ++n) {
    out.nets.w[n] = w0 + (n ? per_layer * md.mlp[0].n_layer : 0);
    out.nets.nl[n] = md.mlp[n].n_layer;
    for (int l = 0; l <= GC_MLP_MAX_LAYERS; ++l) out.nets.width[n][l] = l <= md.mlp[n].n_layer ? md.mlp[n].width[l] : 0;
    out.nets.alpha[n] = md.mlp[n].alpha;
  }
  return GC_OK;
}

template <int WM>
static void step_launch(const EpgxDev& e, const gc_model* dmd, const StepParams& sp, unsigned grid, int transform,
                        hipStream_t s) {
  if (transform == GC_TRANSFORM_LINEAR)
    hipLaunchKernelGGL((k_epgx_step<WM, GC_TRANSFORM_LINEAR>), dim3(grid), dim3(kBlock), 0, s, dmd, sp, e.nets);
  else
    hipLaunchKernelGGL((k_epgx_step<WM, GC_TRANSFORM_SIGMOID>), dim3(grid), dim3(kBlock), 0, s, dmd, sp, e.nets);
}

int epgx_step(gc_ctx* ctx, const EpgxDev& e, const gc_model* dmd, const StepParams& sp, unsigned grid, int transform,
              hipStream_t stream) {
  switch (e.wm) {
    case 16: step_launch<16>(e, dmd, sp, grid, transform, stream); break;
    case 32: step_launch<32>(e, dmd, sp, grid, transform, stream); break;
    case 64: step_launch<64>(e, dmd, sp, grid, transform, stream); break;
    default: return set_err(ctx, GC_ERR_ARG, "EPG-X networks not prepared");
  }
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

int epgx_eval(gc_ctx* ctx, const EpgxDev& e, const gc_model* dmd, const float* x, const float* ex, int64_t nv,
              int transform, float* aux, hipStream_t stream) {
  const dim3 grid(unsigned((nv + kBlock - 1) / kBlock));
  switch (e.wm) {
    case 16: hipLaunchKernelGGL((k_epgx_eval<16>), grid, dim3(kBlock), 0, stream, dmd, x, ex, nv, transform, aux, e.nets); break;
    case 32: hipLaunchKernelGGL((k_epgx_eval<32>), grid, dim3(kBlock), 0, stream, dmd, x, ex, nv, transform, aux, e.nets); break;
    case 64: hipLaunchKernelGGL((k_epgx_eval<64>), grid, dim3(kBlock), 0, stream, dmd, x, ex, nv, transform, aux, e.nets); break;
    default: return set_err(ctx, GC_ERR_ARG, "EPG-X networks not prepared");
  }
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

}  // namespace gc
