// epgx.hip -- host side of the MCR-MWI EPG-X pre-pass (gc_epgx.hpp): the
// networks' weight tables in the matrix-core operand order, and the k_epgx
// launchers.
#include "gc_host.hpp"

namespace gc {

int epgx_prepare(gc_ctx* ctx, const gc_model& md, EpgxDev& out) {
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
      if (l > 0) count[n] += size_t(m.width[l]) * m.width[l - 1] + m.width[l];
    }
  }
  const int nfa = md.i[1];
  if (nfa < 1 || nfa > kBlock / 64)
    return set_err(ctx, GC_ERR_UNSUPPORTED, "EPG-X pre-pass: 1..%d flip angles (got %d)", kBlock / 64, nfa);
  auto tiles = [](int w) { return (w + 15) / 16; };
  std::vector<float> tab;
  EpgxNets& N = out.nets;
  N = EpgxNets{};
  for (int n = 0; n < 2; ++n) {
    const gc_mlp& m = md.mlp[n];
    std::vector<float> h(count[n]);
    // on the context's stream: the caller may have just written the parameters there
    GC_HIP(ctx, hipMemcpyAsync(h.data(), m.params, sizeof(float) * count[n], hipMemcpyDeviceToHost, ctx->stream));
    GC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (float v : h)
      if (!std::isfinite(v)) return set_err(ctx, GC_ERR_ARG, "EPG-X network parameters must be finite");
    N.nl[n] = m.n_layer;
    N.nout[n] = m.width[m.n_layer];
    N.alpha[n] = m.alpha;
    size_t src = 0;
    for (int l = 1; l <= m.n_layer; ++l) {
      const int ni = m.width[l - 1], no = m.width[l];
      const int nti = tiles(ni), nto = tiles(no);
      // MATLAB's column-major [out, in] Weights, then Bias
      auto W = [&](int o, int i) -> float { return (o < no && i < ni) ? h[src + size_t(o) + size_t(no) * i] : 0.f; };
      const size_t bsrc = src + size_t(no) * ni;
      auto B = [&](int o) -> float { return o < no ? h[bsrc + o] : 0.f; };
      EpgxLayer& Ly = N.layer[n][l - 1];
      Ly.nti = nti;
      Ly.nto = nto;
      // every entry: the four floats (r = 0..3) of one lane, contiguous
      Ly.fwd = int32_t(tab.size());
      for (int to = 0; to < nto; ++to)
        for (int ti = 0; ti < nti; ++ti)
          for (int ln = 0; ln < 64; ++ln)
            for (int r = 0; r < 4; ++r) tab.push_back(W(16 * to + (ln & 15), 16 * ti + 4 * (ln >> 4) + r));
      Ly.rev = int32_t(tab.size());
      for (int ti = 0; ti < nti; ++ti)
        for (int to = 0; to < nto; ++to)
          for (int ln = 0; ln < 64; ++ln)
            for (int r = 0; r < 4; ++r) tab.push_back(W(16 * to + 4 * (ln >> 4) + r, 16 * ti + (ln & 15)));
      Ly.bias = int32_t(tab.size());
      for (int to = 0; to < nto; ++to)
        for (int ln = 0; ln < 64; ++ln)
          for (int r = 0; r < 4; ++r) tab.push_back(B(16 * to + 4 * (ln >> 4) + r));
      if (l == m.n_layer) {
        N.last[n] = int32_t(tab.size());
        for (int o = 0; o < no; ++o)
          for (int ti = 0; ti < nti; ++ti)
            for (int ln = 0; ln < 64; ++ln)
              for (int r = 0; r < 4; ++r) tab.push_back(W(o, 16 * ti + 4 * (ln >> 4) + r));
      }
      src = bsrc + no;
    }
  }
  // zero tail: epgx_reverse reads whole kEpgxMaxTiles rows of the last tables
  tab.resize(tab.size() + size_t(kEpgxMaxTiles) * 64 * 4, 0.f);
  // (at most 2 x 8 layers x 3 tables x 25 tile pairs x 1 KiB: far below the
  // buffer resource's 32-bit range)
  GC_HIP(ctx, out.wbuf.alloc(sizeof(float) * tab.size()));
  GC_HIP(ctx, hipMemcpyAsync(out.wbuf.p, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, ctx->stream));
  GC_HIP(ctx, hipStreamSynchronize(ctx->stream));  // tab is freed on return
  N.tab = out.wbuf.as<float>();
  N.tab_bytes = int32_t(tab.size() * sizeof(float));
  out.ready = true;
  return GC_OK;
}

int epgx_step(gc_ctx* ctx, const EpgxDev& e, const gc_model* dmd, const StepParams& sp, unsigned grid, int transform,
              hipStream_t stream) {
  if (!e.ready) return set_err(ctx, GC_ERR_ARG, "EPG-X networks not prepared");
  // grid counts the askadam kernels' 256-voxel blocks
  const dim3 g(grid * unsigned(kEpgxGroupsPerBlock));
  if (transform == GC_TRANSFORM_LINEAR)
    hipLaunchKernelGGL((k_epgx_step<GC_TRANSFORM_LINEAR>), g, dim3(kBlock), 0, stream, dmd, sp, e.nets);
  else
    hipLaunchKernelGGL((k_epgx_step<GC_TRANSFORM_SIGMOID>), g, dim3(kBlock), 0, stream, dmd, sp, e.nets);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

int epgx_eval(gc_ctx* ctx, const EpgxDev& e, const gc_model* dmd, const float* x, const float* ex, int64_t nv,
              int transform, float* aux, hipStream_t stream) {
  if (!e.ready) return set_err(ctx, GC_ERR_ARG, "EPG-X networks not prepared");
  const dim3 grid(unsigned((nv + kEpgxSamples - 1) / kEpgxSamples));
  hipLaunchKernelGGL(k_epgx_eval, grid, dim3(kBlock), 0, stream, dmd, x, ex, nv, transform, aux, e.nets);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

}  // namespace gc
