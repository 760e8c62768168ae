// gc_gwglobal.hpp -- Goodman-Weare with global (shared) parameters,
// mcmc.goodman_weare_wglobal_constant and gw_global_block
// (utils/mcmc.m:385-626), on the device.
//
// Every global decision needs a sum over all voxels, so the sampler runs one
// iteration per launch sequence instead of keeping chains in registers:
//
//   k_gwg_step    per (voxel, chain): applies the last global decision to
//                 logPCurr, runs the local GW block (globals fixed, exponent
//                 z^(NND-1)), stores the thinned local sample, then evaluates
//                 logP at (current local, proposed global) for the chains of
//                 the active global sub-step and leaves per-block FP64 sums of
//                 logP_g - logPCurr per chain;
//   k_gwg_decide  one wavefront per chain: reduces the block sums in a fixed
//                 order (dlogP in FP64, utils/mcmc.m:594) and accepts or
//                 rejects the chain's global proposal (exponent z^(NG-1)).
//
// Chains c = r*W + w (repetition r, walker w) are independent ensembles of
// the same initial state (the reference's repetitions).  State lives in HBM
// chain-fastest, so the C lanes of a voxel group touch one contiguous run:
//   xl   [NND][nv][C]  local variables        lp, lpg  [nv][C]
//   nacc [nv][C]       local acceptances      g, gprop [C][NG]
#pragma once
#include "gc_kernels.hpp"

namespace gc {

struct GwgParams {
  const float* y;
  const float* w;
  const float* ex;
  const float* x0;
  const float* normals;
  const float* uniforms;
  const int32_t* partners;
  const float* g_uniforms;
  const int32_t* g_partners;
  float* xl;
  float* lp;
  float* lpg;
  uint32_t* nacc;
  float* g;
  float* gprop;
  int32_t* goob;
  int32_t* pend;
  double* part;
  double* s1;
  double* s2;
  double* gs1;
  double* gs2;
  float* posterior;
  float* g_posterior;
  float* mean;
  float* stdv;
  float* accept;
  float* xfinal;
  float* g_xfinal;
  uint8_t* trace;
  uint8_t* g_trace;
  int64_t nv;
  uint64_t seed;
  int32_t iteration, T, rep, n_walker, nchain, n_sample, n_stored, rng, ens, nnd, ng, nblk;
  uint32_t gmask;
  float step_size, start_range;
};

// One launch of the sequence: iteration k (0 for warm-up), global step t,
// whether the local block runs, the active global sub-step (-1 none; 0 =
// all walkers, or the first red-black half; 1 = the second half), and the
// posterior slot of a thinned sample (-1 none).
struct GwgStep {
  int32_t k, t, local, gsub, keep_idx;
};

// chain ids of the per-walker global draws, shared by every voxel
GC_DEV uint64_t gwg_global_chain(int r, int wk) {
  return (uint64_t(0xFFFEu) << 48) | (uint64_t(uint32_t(r)) << 32) | uint64_t(uint32_t(wk));
}
GC_DEV uint64_t gwg_init_chain(int wk) { return (uint64_t(0xFFFDu) << 48) | uint64_t(uint32_t(wk)); }

GC_DEV bool gwg_isg(uint32_t mask, int q) { return (mask >> q) & 1u; }
// row of variable q among the global (or local) variables
GC_DEV int gwg_grow(uint32_t mask, int q) { return __builtin_popcount(mask & ((1u << q) - 1u)); }
GC_DEV int gwg_lrow(uint32_t mask, int q) { return __builtin_popcount(~mask & ((1u << q) - 1u)); }

GC_DEV float pick4(const uint32_t o[4], int i) {
  const uint32_t x = (i == 0) ? o[0] : (i == 1) ? o[1] : (i == 2) ? o[2] : o[3];
  return u01(x);
}

// stretch factor zz = ((a-1) u + 1)^2 / a (utils/mcmc.m:515, :587)
GC_DEV float gwg_zz(float u, float a) {
  const float t = (a - 1.f) * u + 1.f;
  return t * t / a;
}

GC_DEV bool gwg_active(const GwgParams& P, int gsub, int wk) {
  if (gsub < 0) return false;
  if (P.ens != GC_ENSEMBLE_REDBLACK) return true;
  const int h = P.n_walker / 2;
  return (gsub == 0) ? (wk < h) : (wk >= h);
}

// The two draws of chain (r, wk)'s global proposal at global step t.
GC_DEV void gwg_global_draws(const GwgParams& P, int r, int wk, int t, int gsub, float& u1, float& u2) {
  const bool rb = P.ens == GC_ENSEMBLE_REDBLACK;
  const int W = P.n_walker, h = W / 2, hm = W - h;
  if (P.rng == GC_RNG_REPLAY) {
    const int64_t step = int64_t(r) * P.T + (t - 1);
    const int j = rb ? ((gsub == 0) ? wk : wk - h) : wk;
    const int64_t base = int64_t(W) * P.ng + (rb ? (step * 2 + gsub) * 2 * hm : step * 2 * W);
    u1 = P.g_uniforms[base + j];
    u2 = P.g_uniforms[base + (rb ? hm : W) + j];
  } else {
    uint32_t o[4];
    draw4(P.seed, uint32_t(t), gwg_global_chain(r, wk), PURPOSE_ZZ + uint32_t(gsub), o);
    u1 = u01(o[0]);
    u2 = u01(o[1]);
  }
}

template <class Model, bool HAS_W>
GC_DEV void gwg_load_voxel(const gc_model& md, const GwgParams& P, int64_t vv, float* yr, float* wr, float* ex) {
#pragma unroll
  for (int rr = 0; rr < Model::MR; ++rr) {
    if (Model::valid(md, rr)) {
      const int64_t m = Model::meas(md, rr);
      yr[rr] = P.y[m * P.nv + vv];
      if (HAS_W) wr[HAS_W ? rr : 0] = P.w[m * P.nv + vv];
    } else {
      yr[rr] = 0.f;
      if (HAS_W) wr[HAS_W ? rr : 0] = 0.f;
    }
  }
#pragma unroll
  for (int e = 0; e < kMaxExtra; ++e) ex[e] = (e < md.n_extra) ? P.ex[int64_t(e) * P.nv + vv] : 0.f;
}

// Initial ensemble and logP0 (utils/mcmc.m:484-492): locals
// x0 + (ub-lb) startRange randn, clamped; globals lb + (ub-lb) rand.
template <class Model, bool HAS_W>
__global__ __launch_bounds__(kBlock) void k_gwg_init(const gc_model* __restrict__ mdp, const GwgParams P) {
  const gc_model& md = *mdp;
  constexpr int Q = Model::NP + 1;
  const int C = P.nchain, W = P.n_walker;
  const int gpb = kBlock / C;
  const int tl = threadIdx.x, gi = tl / C, c = tl % C;
  const int wk = c % W;
  const bool lane_ok = gi < gpb;
  const int64_t nv = P.nv;
  const int64_t ntile = (nv + gpb - 1) / gpb;
  for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int64_t v = tile * gpb + gi;
    const bool active = lane_ok && v < nv;
    const int64_t vv = active ? v : 0;
    float yr[Model::MR], wr[HAS_W ? Model::MR : 1], ex[kMaxExtra];
    gwg_load_voxel<Model, HAS_W>(md, P, vv, yr, wr, ex);
    float x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const float lb = md.lb[q], ub = md.ub[q];
      if (gwg_isg(P.gmask, q)) {
        const int i = gwg_grow(P.gmask, q);
        float u;
        if (P.rng == GC_RNG_REPLAY) {
          u = P.g_uniforms[int64_t(wk) * P.ng + i];
        } else {
          uint32_t o[4];
          draw4(P.seed, 0u, gwg_init_chain(wk), PURPOSE_INIT + uint32_t(i / 4), o);
          u = pick4(o, i % 4);
        }
        x[q] = lb + (ub - lb) * u;
      } else {
        const int j = gwg_lrow(P.gmask, q);
        float n;
        if (P.rng == GC_RNG_REPLAY) {
          n = P.normals[(int64_t(wk) * nv + vv) * P.nnd + j];
        } else {
          uint32_t o[4];
          draw4(P.seed, 0u, uint64_t(wk) * uint64_t(nv) + uint64_t(vv), PURPOSE_INIT + uint32_t(j / 4), o);
          const int pr = (j % 4) / 2;
          float n0, n1;
          box_muller(u01(pr ? o[2] : o[0]), u01(pr ? o[3] : o[1]), n0, n1);
          n = (j % 2 == 0) ? n0 : n1;
        }
        x[q] = fminf(fmaxf(P.x0[int64_t(q) * nv + vv] + (ub - lb) * P.start_range * n, lb), ub);
      }
    }
    const float lp0 = logp_gaussian(model_sos<Model, HAS_W>(md, x, ex, yr, wr), x[Q - 1], md.n_meas);
    if (active) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (!gwg_isg(P.gmask, q)) P.xl[(int64_t(gwg_lrow(P.gmask, q)) * nv + v) * C + c] = x[q];
      P.lp[v * C + c] = lp0;
      P.nacc[v * C + c] = 0u;
      if (v == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q)
          if (gwg_isg(P.gmask, q)) P.g[int64_t(c) * P.ng + gwg_grow(P.gmask, q)] = x[q];
      }
    }
  }
}

// One launch of the iteration sequence (see the file comment).
template <class Model, bool HAS_W>
__global__ __launch_bounds__(kBlock) void k_gwg_step(const gc_model* __restrict__ mdp, const GwgParams P,
                                                      const GwgStep S) {
  const gc_model& md = *mdp;
  constexpr int Q = Model::NP + 1;
  const int C = P.nchain, W = P.n_walker;
  const int gpb = kBlock / C;
  const int tl = threadIdx.x, gi = tl / C, c = tl % C;
  const int r = c / W, wk = c % W;
  const bool lane_ok = gi < gpb;
  const int64_t nv = P.nv;
  const bool rb = P.ens == GC_ENSEMBLE_REDBLACK;
  const int h = W / 2, hm = W - h;
  const int M = md.n_meas;
  const bool pend = lane_ok && P.pend[c] != 0;

  // this chain's global proposal, identical at every voxel (:586-590)
  const bool gmine = lane_ok && gwg_active(P, S.gsub, wk);
  float gp[Q];
  bool goob = false;
#pragma unroll
  for (int q = 0; q < Q; ++q) gp[q] = 0.f;
  if (gmine) {
    float u1, u2;
    gwg_global_draws(P, r, wk, S.t, S.gsub, u1, u2);
    const float zz = gwg_zz(u1, P.step_size);
    const int pw = P.g_partners ? P.g_partners[(int64_t(r) * P.T + (S.t - 1)) * W + wk] : 0;
    const int pc = r * W + pw;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (gwg_isg(P.gmask, q)) {
        const int i = gwg_grow(P.gmask, q);
        const float xa = P.g[int64_t(pc) * P.ng + i];
        const float val = xa + (P.g[int64_t(c) * P.ng + i] - xa) * zz;
        goob = goob || (val < md.lb[q]) || (val > md.ub[q]);
        gp[q] = fminf(fmaxf(val, md.lb[q]), md.ub[q]);
      }
    }
    if (blockIdx.x == 0 && gi == 0) {
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if (gwg_isg(P.gmask, q)) P.gprop[int64_t(c) * P.ng + gwg_grow(P.gmask, q)] = gp[q];
      P.goob[c] = goob ? 1 : 0;
    }
  }

  __shared__ float shx[Q][kBlock];
  __shared__ double shd[2][kBlock];
  double dsum = 0.0;
  const int base = gi * C + r * W;  // LDS slot of walker 0 of this (voxel, repetition)
  const int64_t ntile = (nv + gpb - 1) / gpb;
  for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int64_t v = tile * gpb + gi;
    const bool active = lane_ok && v < nv;
    const int64_t vv = active ? v : 0;
    float yr[Model::MR], wr[HAS_W ? Model::MR : 1], ex[kMaxExtra];
    gwg_load_voxel<Model, HAS_W>(md, P, vv, yr, wr, ex);
    const int cc = lane_ok ? c : 0;
    float x[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (gwg_isg(P.gmask, q)) x[q] = P.g[int64_t(cc) * P.ng + gwg_grow(P.gmask, q)];
      else x[q] = P.xl[(int64_t(gwg_lrow(P.gmask, q)) * nv + vv) * C + cc];
    }
    // a global acceptance moves the cache to logP at the accepted globals (:599)
    float lpv = pend ? P.lpg[vv * C + cc] : P.lp[vv * C + cc];
    uint32_t na = 0;

    if (S.local) {
      // ---- local block, globals fixed (:512-545) -------------------------
      const int64_t step = int64_t(r) * P.iteration + (S.k - 1);
      const int nsub = rb ? 2 : 1;
      for (int sub = 0; sub < nsub; ++sub) {
#pragma unroll
        for (int q = 0; q < Q; ++q) shx[q][tl] = x[q];
        __syncthreads();
        bool mine = lane_ok;
        int j = wk;
        if (rb) {
          mine = mine && ((sub == 0) ? (wk < h) : (wk >= h));
          j = (sub == 0) ? wk : wk - h;
        }
        const int pw = mine ? P.partners[step * W + wk] : wk;
        float u1 = 0.f, u2 = 0.f;
        if (mine) {
          if (P.rng == GC_RNG_REPLAY) {
            int64_t ub0, ust;
            if (rb) { ub0 = (step * 2 + sub) * 2 * hm; ust = hm; }
            else { ub0 = step * 2 * W; ust = W; }
            u1 = P.uniforms[(ub0 + j) * nv + vv];
            u2 = P.uniforms[(ub0 + ust + j) * nv + vv];
          } else {
            const uint64_t chain = (uint64_t(r) * W + wk) * uint64_t(nv) + uint64_t(vv);
            uint32_t o[4];
            draw4(P.seed, uint32_t(S.k), chain, PURPOSE_ZZ + sub, o);
            u1 = u01(o[0]);
            u2 = u01(o[1]);
          }
        }
        const float zz = gwg_zz(u1, P.step_size);
        float xp[Q];
        bool oob = false;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (gwg_isg(P.gmask, q)) {
            xp[q] = x[q];
          } else {
            const float xa = shx[q][base + pw];
            const float val = xa + (x[q] - xa) * zz;  // :516
            oob = oob || (val < md.lb[q]) || (val > md.ub[q]);
            xp[q] = fminf(fmaxf(val, md.lb[q]), md.ub[q]);
          }
        }
        __syncthreads();
        if (mine) {
          // logP at (new local, old global), :519-520
          const float lpp = logp_gaussian(model_sos<Model, HAS_W>(md, xp, ex, yr, wr), xp[Q - 1], M);
          float zpow = 1.f;
          for (int i = 0; i < P.nnd - 1; ++i) zpow *= zz;
          const float ratio = min1(zpow * __expf(lpp - lpv));  // :521
          const bool acc = (ratio > u2) && !oob;              // :522-523
          if (P.trace != nullptr && active) P.trace[(step * W + wk) * nv + v] = acc ? 1 : 0;
          if (acc) {
#pragma unroll
            for (int q = 0; q < Q; ++q) x[q] = xp[q];
            lpv = lpp;
            ++na;
          }
        }
      }
      // ---- thinned local sample (:554-558) and its running sums ------------
      if (S.keep_idx >= 0) {
        const int R = P.rep;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (gwg_isg(P.gmask, q)) continue;
          const int jr = gwg_lrow(P.gmask, q);
          if (P.posterior != nullptr && active)
            P.posterior[(((int64_t(jr) * R + r) * P.n_sample + S.keep_idx) * W + wk) * nv + v] = x[q];
          const double d = active ? double(x[q]) - double(P.x0[int64_t(q) * nv + v]) : 0.0;
          shd[0][tl] = d;
          shd[1][tl] = d * d;
          __syncthreads();
          if (active && c == 0) {
            double a = 0.0, b = 0.0;
            for (int i = 0; i < C; ++i) {
              a += shd[0][gi * C + i];
              b += shd[1][gi * C + i];
            }
            P.s1[int64_t(jr) * nv + v] += a;
            P.s2[int64_t(jr) * nv + v] += b;
          }
          __syncthreads();
        }
      }
    }

    // ---- logP at (current local, proposed global), :591-594 --------------
    if (gmine) {
      float xg[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) xg[q] = gwg_isg(P.gmask, q) ? gp[q] : x[q];
      const float lpg = logp_gaussian(model_sos<Model, HAS_W>(md, xg, ex, yr, wr), xg[Q - 1], M);
      if (active) {
        P.lpg[v * C + c] = lpg;
        dsum += double(lpg - lpv);  // double(logP_g - logPCurr)
      }
    }
    if (active) {
      if (S.local) {
#pragma unroll
        for (int q = 0; q < Q; ++q)
          if (!gwg_isg(P.gmask, q)) P.xl[(int64_t(gwg_lrow(P.gmask, q)) * nv + v) * C + c] = x[q];
        if (na) P.nacc[v * C + c] += na;
      }
      if (S.local || pend) P.lp[v * C + c] = lpv;
    }
  }
  // per-block sums of logP_g - logPCurr per chain, in a fixed order
  if (S.gsub >= 0) {
    shd[0][tl] = dsum;
    __syncthreads();
    if (gi == 0 && tl < C) {
      double a = 0.0;
      for (int i = 0; i < gpb; ++i) a += shd[0][i * C + tl];
      P.part[int64_t(blockIdx.x) * C + tl] = a;
    }
  }
}

// Global decision of every chain (:594-599), one wavefront per chain.
static __global__ __launch_bounds__(kBlock) void k_gwg_decide(const gc_model* __restrict__ mdp, const GwgParams P,
                                                              const GwgStep S) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (c >= P.nchain) return;
  const int W = P.n_walker, r = c / W, wk = c % W;
  const bool act = gwg_active(P, S.gsub, wk);
  double d = 0.0;
  if (act)
    for (int b = lane; b < P.nblk; b += 64) d += P.part[int64_t(b) * P.nchain + c];
  d = wave_sum_d(d);
  if (lane != 0) return;
  bool acc = false;
  if (act) {
    float u1, u2;
    gwg_global_draws(P, r, wk, S.t, S.gsub, u1, u2);
    const float zz = gwg_zz(u1, P.step_size);
    float zpow = 1.f;
    for (int i = 0; i < P.ng - 1; ++i) zpow *= zz;
    // zz.^(NGlobal-1) .* exp(dlogP) in single, exp of the FP64 sum
    const float ratio = min1(zpow * float(exp(d)));
    acc = (ratio > u2) && (P.goob[c] == 0);
    if (acc)
      for (int i = 0; i < P.ng; ++i) P.g[int64_t(c) * P.ng + i] = P.gprop[int64_t(c) * P.ng + i];
    if (P.g_trace != nullptr) P.g_trace[(int64_t(r) * P.T + (S.t - 1)) * W + wk] = acc ? 1 : 0;
  }
  P.pend[c] = acc ? 1 : 0;
  if (S.keep_idx >= 0) {
    const gc_model& md = *mdp;
    int i = 0;
    for (int q = 0; q < GC_MAX_PAR && i < P.ng; ++q) {
      if (!gwg_isg(P.gmask, q)) continue;
      const float gv = P.g[int64_t(c) * P.ng + i];
      if (P.g_posterior != nullptr)
        P.g_posterior[((int64_t(i) * P.rep + r) * P.n_sample + S.keep_idx) * W + wk] = gv;
      const double dd = double(gv) - 0.5 * (double(md.lb[q]) + double(md.ub[q]));
      P.gs1[int64_t(c) * P.ng + i] += dd;
      P.gs2[int64_t(c) * P.ng + i] += dd * dd;
      ++i;
    }
  }
}

// Final state, acceptance and res2out mean / std (unfilled posterior slots
// are zeros, utils/mcmc.m:481-482, and enter the statistics).
template <class Model>
__global__ __launch_bounds__(kBlock) void k_gwg_final(const gc_model* __restrict__ mdp, const GwgParams P) {
  const gc_model& md = *mdp;
  constexpr int Q = Model::NP + 1;
  const int C = P.nchain, W = P.n_walker;
  const int gpb = kBlock / C;
  const int tl = threadIdx.x, gi = tl / C, c = tl % C;
  const int r = c / W, wk = c % W;
  const bool lane_ok = gi < gpb;
  const int64_t nv = P.nv;
  const double n_total = double(P.n_sample) * C;
  const double unfilled = double(P.n_sample - P.n_stored) * C;
  __shared__ uint32_t sha[kBlock];
  const int64_t ntile = (nv + gpb - 1) / gpb;
  for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const int64_t v = tile * gpb + gi;
    const bool active = lane_ok && v < nv;
    sha[tl] = active ? P.nacc[v * C + c] : 0u;
    if (active && P.xfinal != nullptr) {
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (gwg_isg(P.gmask, q)) continue;
        const int j = gwg_lrow(P.gmask, q);
        P.xfinal[((int64_t(r) * W + wk) * P.nnd + j) * nv + v] = P.xl[(int64_t(j) * nv + v) * C + c];
      }
    }
    if (active && v == 0 && P.g_xfinal != nullptr)
      for (int i = 0; i < P.ng; ++i) P.g_xfinal[int64_t(c) * P.ng + i] = P.g[int64_t(c) * P.ng + i];
    __syncthreads();
    if (active && c == 0) {
      uint64_t t = 0;
      for (int i = 0; i < C; ++i) t += sha[gi * C + i];
      if (P.accept != nullptr) P.accept[v] = float(double(t) / (double(P.iteration) * C));
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double a, b, xref;
        if (gwg_isg(P.gmask, q)) {
          const int i = gwg_grow(P.gmask, q);
          xref = 0.5 * (double(md.lb[q]) + double(md.ub[q]));
          a = 0.0;
          b = 0.0;
          for (int k = 0; k < C; ++k) {
            a += P.gs1[int64_t(k) * P.ng + i];
            b += P.gs2[int64_t(k) * P.ng + i];
          }
        } else {
          const int j = gwg_lrow(P.gmask, q);
          xref = double(P.x0[int64_t(q) * nv + v]);
          a = P.s1[int64_t(j) * nv + v];
          b = P.s2[int64_t(j) * nv + v];
        }
        a += -xref * unfilled;
        b += xref * xref * unfilled;
        if (P.mean != nullptr) P.mean[int64_t(q) * nv + v] = float(xref + a / n_total);
        if (P.stdv != nullptr) {
          const double var = (n_total > 1) ? (b - a * a / n_total) / (n_total - 1) : 0.0;
          P.stdv[int64_t(q) * nv + v] = float(sqrt(fmax(var, 0.0)));
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace gc
