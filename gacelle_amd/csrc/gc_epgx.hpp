// gc_epgx.hpp -- the EPG-X MLP surrogate of MCR-MWI (isEPG,
// MCRMWI/gpuMCRMWI.m:664-681) as a pre-pass: for every (voxel, flip angle)
// sample the 11 features of
// MCRMWI/EPGXgen_net/feature_preprocess_MCRMWI_MLP_EPGX_leakyrelu.m:24-95,
// the phase and magnitude networks of mlp_model_leakyRelu.m:23-47, and the
// derivatives of their three outputs with respect to (MVF, R1IEW, kIEWM,
// R2sIW), written as the 15 aux rows that McrmwiModel<..., EPG> reads.
//
// One thread per voxel runs its flip angles one after the other.  A
// network's weights are the same for every lane, so they are read at
// wave-uniform addresses (scalar loads) from a padded copy the library
// builds: per layer a [WM][WM] weight block (row = output unit, zero
// beyond the layer's widths) and a [WM] bias.  Activations live in
// registers (WM each for the current and next layer); the signs of the
// hidden pre-activations are kept as 64-bit masks in LDS for the reverse
// passes, one per network output.
#pragma once
#include "gc_kernels.hpp"

namespace gc {

// the networks as the kernels read them
struct EpgxNets {
  const float* w[2];                         // padded weights of the phase / magnitude network
  int32_t nl[2];                             // layers
  int32_t width[2][GC_MLP_MAX_LAYERS + 1];   // units per layer (width[.][0] = 11)
  float alpha[2];                            // leaky ReLU scale
};

template <int WM>
constexpr size_t epgx_layer_floats() { return size_t(WM) * WM + WM; }

// Forward pass of one network on a[] (inputs, zero beyond width[0]); on
// return a[] holds the outputs.  The sign mask of hidden layer l goes to
// msk[l * kBlock] (this thread's column of an LDS table).  Zero rows beyond
// a layer's width and zero chunks of 16 inputs are skipped by uniform
// branches.  (Straight-line row groups, which let the scalar loads run
// further ahead, raise the 64-unit kernel above 256 registers: one wave per
// SIMD, measured no faster.)
template <int WM>
GC_DEV void mlp_fwd(const float* __restrict__ W, int nl, const int32_t* width, float alpha, float (&a)[WM],
                    uint64_t* msk) {
  for (int l = 0; l < nl; ++l) {
    const float* __restrict__ Wl = W + size_t(l) * epgx_layer_floats<WM>();
    const int wi = width[l], wo = width[l + 1];
    float z[WM];
#pragma unroll
    for (int j = 0; j < WM; ++j) {
      float t = 0.f;
      if (j < wo) {
        t = Wl[WM * WM + j];
#pragma unroll
        for (int i0 = 0; i0 < WM; i0 += 16) {
          if (i0 < wi) {
#pragma unroll
            for (int i = i0; i < i0 + 16; ++i) t = fmaf(Wl[j * WM + i], a[i], t);
          }
        }
      }
      z[j] = t;
    }
    if (l + 1 < nl) {
      uint64_t m = 0;
#pragma unroll
      for (int j = 0; j < WM; ++j) {
        const bool pos = z[j] > 0.f;
        m |= uint64_t(pos) << j;
        a[j] = pos ? z[j] : alpha * z[j];
      }
      msk[l * kBlock] = m;
    } else {
#pragma unroll
      for (int j = 0; j < WM; ++j) a[j] = z[j];
    }
  }
}

// Reverse pass for output o: g[0..width[0]) = d out_o / d input.
template <int WM>
GC_DEV void mlp_rev(const float* __restrict__ W, int nl, const int32_t* width, float alpha, int o,
                    const uint64_t* msk, float (&g)[WM]) {
#pragma unroll
  for (int j = 0; j < WM; ++j) g[j] = (j == o) ? 1.f : 0.f;
  for (int l = nl - 1; l >= 0; --l) {
    const float* __restrict__ Wl = W + size_t(l) * epgx_layer_floats<WM>();
    const int wi = width[l], wo = width[l + 1];
    float gi[WM];
#pragma unroll
    for (int i = 0; i < WM; ++i) gi[i] = 0.f;
#pragma unroll
    for (int j = 0; j < WM; ++j) {
      if (j < wo) {
#pragma unroll
        for (int i0 = 0; i0 < WM; i0 += 16) {
          if (i0 < wi) {
#pragma unroll
            for (int i = i0; i < i0 + 16; ++i) gi[i] = fmaf(Wl[j * WM + i], g[j], gi[i]);
          }
        }
      }
    }
    if (l > 0) {
      const uint64_t m = msk[(l - 1) * kBlock];
#pragma unroll
      for (int i = 0; i < WM; ++i) g[i] = ((m >> i) & 1ull) ? gi[i] : alpha * gi[i];
    } else {
#pragma unroll
      for (int i = 0; i < WM; ++i) g[i] = gi[i];
    }
  }
}

// One (voxel, flip angle) sample: rows k * nfa + fi of aux (stride nv) get
// {ph, dph/d(MVF, R1IEW, kIEWM, R2sIW), dIEW, d.../d(...), dMW, d.../d(...)}.
// fa = nominal flip angle x B1 [rad]; msk and xs point at this thread's
// columns of the LDS tables (sign masks; the 11 features and their 4
// derivatives, kept out of registers while the networks run).
template <int WM>
GC_DEV void epgx_sample(const EpgxNets& N, float mvf, float r1iew, float kfr, float r2siw, float fa, float tr,
                        float t1mw, uint64_t* msk, float* xs, float* __restrict__ aux, int64_t rstride) {
  {
    using D4 = DualF<4>;
    const D4 fm = D4::var(mvf, 0), r1 = D4::var(r1iew, 1), kf = D4::var(kfr, 2), r2 = D4::var(r2siw, 3);
    float sa, ca;
    sincosf(fa, &sa, &ca);
    const D4 kmw = (kf * (1.f - fm)) / fm;
    D4 x[11];
    x[0] = fm;
    x[1] = dexp(-tr * r1);                       // exp(-TR / T1iew)
    x[2] = dexp(-tr * kf);                       // exp(-TR kiewmw)
    x[3] = D4::cst(sa);
    x[4] = 1.f - ca * x[1];                      // 1 - E1 cos(alpha)
    x[5] = dexp(-tr * kmw);                      // exp(-TR kmwiew)
    x[6] = dexp(-tr * (r1 + kf));
    x[7] = dexp(-tr * ((1.f / t1mw) + kmw));
    x[8] = (sa * ((1.f - fm) * (1.f - x[1]))) / (1.f - ca * x[1]);   // Bloch steady state of IEW
    const float e1m = __expf(-tr / t1mw);
    x[9] = (sa * (1.f - e1m) / (1.f - ca * e1m)) * fm;                // ... of M
    x[10] = dexp(-tr * r2);                      // exp(-TR / T2iew)
#pragma unroll
    for (int f = 0; f < 11; ++f) {
      xs[(f * 5) * kBlock] = x[f].v;
#pragma unroll
      for (int q = 0; q < 4; ++q) xs[(f * 5 + 1 + q) * kBlock] = x[f].d[q];
    }
  }
  int k = 0;
#pragma unroll 1
  for (int net = 0; net < 2; ++net) {
    float a[WM];
#pragma unroll
    for (int i = 0; i < WM; ++i) a[i] = (i < 11) ? xs[((i < 11 ? i : 0) * 5) * kBlock] : 0.f;
    mlp_fwd<WM>(N.w[net], N.nl[net], N.width[net], N.alpha[net], a, msk);
    const int nout = N.width[net][N.nl[net]];
    const float val[2] = {a[0], a[1]};
#pragma unroll 1
    for (int o = 0; o < nout; ++o) {
      float g[WM];
      mlp_rev<WM>(N.w[net], N.nl[net], N.width[net], N.alpha[net], o, msk, g);
      aux[int64_t(k) * rstride] = val[o];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float s = 0.f;
#pragma unroll
        for (int f = 0; f < 11; ++f) s = fmaf(g[f], xs[(f * 5 + 1 + q) * kBlock], s);
        aux[int64_t(k + 1 + q) * rstride] = s;
      }
      k += 5;
    }
  }
}

// MVF and the network inputs of MCR-MWI's parameter slots
// (S0, MWF, IWF, R1IEW, kIEWM, R2sMW, R2sIW, ...), as McrmwiModel::amplitudes
GC_DEV float epgx_mvf(const float* th, float rho) {
  const float S0MW = th[0] * th[1];
  const float S0IW = th[0] * (1.f - th[1]) * th[2];
  const float S0EW = th[0] * (1.f - th[1]) * (1.f - th[2]);
  const float tv = fmaxf(S0IW + S0EW + (1.f / rho) * S0MW, 1e-8f);
  return fmaxf(((1.f / rho) * S0MW) / tv, 1e-8f);
}

// One voxel's flip angles.  The pointers are __restrict__ so that the
// weight loads are known not to be clobbered by the aux stores: the weights
// then come through scalar loads (a plain pointer to aux lets every store
// alias them, and the loads turn into per-lane vector loads).
template <int WM, class TF>
GC_DEV void epgx_voxel(const gc_model& md, const EpgxNets& N, const float* __restrict__ z,
                       const float* __restrict__ ex, float* __restrict__ aux, int64_t nv, int64_t v, TF&& transform,
                       uint64_t* msk, float* xs) {
  float th[7];
#pragma unroll
  for (int p = 0; p < 7; ++p) th[p] = transform(z[int64_t(p) * nv + v], md.lb[p], md.ub[p]);
  const float mvf = epgx_mvf(th, md.f[3]);
  const int nte = md.i[0], nfa = md.i[1];
  const float b1 = ex[v];
  for (int fi = 0; fi < nfa; ++fi)
    epgx_sample<WM>(N, mvf, th[3], th[4], th[6], md.table[nte + fi] * b1, md.f[0], md.f[4], msk, xs,
                    aux + int64_t(fi) * nv + v, int64_t(nfa) * nv);
}

// the step form: aux rows of every volume's voxels from z_in (same block
// mapping as k_askadam_step)
template <int WM, int TR>
__global__ __launch_bounds__(kBlock) void k_epgx_step(const gc_model* __restrict__ mdp, const StepParams sp,
                                                      const EpgxNets N) {
  int vi = 0;
  for (int i = 1; i < sp.n_vol; ++i)
    if (int(blockIdx.x) >= sp.vols[i].blk0) vi = i;
  const StepVol& V = sp.vols[vi];
  if (!V.active) return;
  const int64_t nv = V.nv;
  const int64_t v = int64_t(blockIdx.x - V.blk0) * kBlock + threadIdx.x;
  if (v >= nv) return;
  __shared__ uint64_t msk[GC_MLP_MAX_LAYERS][kBlock];
  __shared__ float xs[55][kBlock];
  epgx_voxel<WM>(*mdp, N, V.z_in, V.ex, V.aux, nv, v,
                 [](float z, float lb, float ub) {
                   float t, d;
                   transform_fwd_t<TR>(z, lb, ub, t, d);
                   return t;
                 },
                 &msk[0][threadIdx.x], &xs[0][threadIdx.x]);
}

// the one-shot form (gc_fwd, gc_fwd_adj, the askadam outputs): aux rows
// from parameters x in the given transform space
template <int WM>
__global__ __launch_bounds__(kBlock) void k_epgx_eval(const gc_model* __restrict__ mdp, const float* __restrict__ x,
                                                      const float* __restrict__ ex, int64_t nv, int32_t transform,
                                                      float* __restrict__ aux, const EpgxNets N) {
  const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (v >= nv) return;
  __shared__ uint64_t msk[GC_MLP_MAX_LAYERS][kBlock];
  __shared__ float xs[55][kBlock];
  epgx_voxel<WM>(*mdp, N, x, ex, aux, nv, v,
                 [transform](float z, float lb, float ub) {
                   float t, d;
                   transform_fwd(transform, z, lb, ub, t, d);
                   return t;
                 },
                 &msk[0][threadIdx.x], &xs[0][threadIdx.x]);
}

}  // namespace gc
