// gc_epgx.hpp -- the EPG-X MLP surrogate of MCR-MWI (isEPG,
// MCRMWI/gpuMCRMWI.m:664-681) as a pre-pass: for every (voxel, flip angle)
// sample the 11 features of
// MCRMWI/EPGXgen_net/feature_preprocess_MCRMWI_MLP_EPGX_leakyrelu.m:24-95,
// the phase and magnitude networks of mlp_model_leakyRelu.m:23-47, and the
// derivatives of their three outputs with respect to (MVF, R1IEW, kIEWM,
// R2sIW), written as the 15 aux rows that McrmwiModel<..., EPG> reads.
//
// Every layer of a network is a product [units x samples] = W [units x in] x
// A [in x samples], run on the FP32 matrix cores (v_mfma_f32_16x16x4_f32).
// A wavefront holds kEpgxNB column blocks of 16 samples.  For that
// instruction lane l holds D rows 4 (l/16) + r (r = 0..3) of column l % 16,
// and it supplies B row l/16 of column l % 16: so instruction r of an input
// tile t takes, as its B operand, register r of the previous layer's output
// tile t, with B row g standing for unit 16 t + 4 g + r.  Activations and the
// reverse-pass gradients therefore stay in registers from layer to layer;
// only the weights (the A operand) are read, one float per lane and
// instruction, from tables the library arranges in exactly that order.
#pragma once
#include "gc_kernels.hpp"

namespace gc {

constexpr int kEpgxMaxTiles = (GC_MLP_MAX_WIDTH + 15) / 16;  // 16-unit tiles of the widest layer
constexpr int kEpgxNB = 2;                                   // column blocks of 16 samples per wavefront
constexpr int kEpgxSamples = 16 * kEpgxNB;                   // samples (voxels) per wavefront
// workgroups of the pre-pass per 256-voxel block of the askadam kernels: a
// workgroup has one wavefront per flip angle over kEpgxSamples voxels
constexpr int kEpgxGroupsPerBlock = kBlock / kEpgxSamples;

// One layer of a network in the table buffer (offsets in floats; every entry
// is one f32x4 per lane, the four weights a lane feeds to the four
// instructions r = 0..3 of an input tile, so a lane fetches them with one
// 16-byte load):
//  fwd  [nto][nti][64]: W[16 to + l%16][16 ti + 4 (l/16) + r]
//  rev  [nti][nto][64]: W[16 to + 4 (l/16) + r][16 ti + l%16]
//  bias [nto][64]:      b[16 to + 4 (l/16) + r]
// (zero outside the layer's widths).  The buffer ends with kEpgxMaxTiles x 64
// zero entries, so reading a whole kEpgxMaxTiles row of the last table stays
// inside it.
struct EpgxLayer {
  int32_t fwd, rev, bias;
  int32_t nti, nto;
};

struct EpgxNets {
  const float* tab;                        // every table of both networks
  int32_t tab_bytes;                       // its size
  EpgxLayer layer[2][GC_MLP_MAX_LAYERS];
  int32_t last[2];   // the last layer's rows [nout][nti][64]: W_L[o][16 ti + 4 (l/16) + r]
  int32_t nl[2];     // layers
  int32_t nout[2];   // outputs (1 phase, 2 magnitude)
  float alpha[2];    // leaky ReLU scale
};

using f32x4 = __attribute__((ext_vector_type(4))) float;

GC_DEV f32x4 mma16(float a, float b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// LDS of one wavefront: the 11 features of each of its samples and their
// derivatives with respect to the 4 parameters ([f * 5 + 0] value,
// [f * 5 + 1 + q] d/dq), and the sign masks of the hidden pre-activations
// (bit 4 t + r of word [layer][block] of a lane: unit 16 t + 4 (l/16) + r of
// its column is positive).
struct EpgxLds {
  float xs[55][kEpgxSamples];
  uint32_t msk[GC_MLP_MAX_LAYERS][kEpgxNB][64];
};

// The 11 features of one sample with their derivatives, into column c of xs.
GC_DEV void epgx_features(float mvf, float r1iew, float kfr, float r2siw, float fa, float tr, float t1mw,
                          EpgxLds& L, int c) {
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
    L.xs[f * 5][c] = x[f].v;
#pragma unroll
    for (int q = 0; q < 4; ++q) L.xs[f * 5 + 1 + q][c] = x[f].d[q];
  }
}

using EpgxRsrc = __amdgpu_buffer_rsrc_t;

// The table buffer as a raw buffer resource (gfx9 word 3: 32-bit data
// format); reads past its end return zero.
GC_DEV EpgxRsrc epgx_rsrc(const EpgxNets& N) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(N.tab), 0, N.tab_bytes, 0x00020000);
}

// Entry e of the table that starts `base` floats into the buffer: one
// 16-byte buffer load with the lane's offset in a vector register shared by
// every load and the table offset in a scalar register.
GC_DEV f32x4 tab_entry(EpgxRsrc rs, int lane, int32_t base, int e) {
  return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(rs, lane * 16, base * 4 + e * 1024, 0));
}

// One forward layer with NTI input and NTO output tiles, known at compile
// time: the layer is one straight block, so the scheduler issues its weight
// loads well ahead of the matrix instructions that use them.  Hidden layers
// apply the leaky ReLU and record their sign masks in msk.
template <int NTI, int NTO>
GC_DEV void epgx_fwd_layer(EpgxRsrc rs, const EpgxLayer& Ly, bool hidden, float alpha,
                           f32x4 (&act)[kEpgxMaxTiles][kEpgxNB], uint32_t (&msk)[kEpgxNB][64], int lane) {
  f32x4 acc[NTO][kEpgxNB];
#pragma unroll
  for (int to = 0; to < NTO; ++to) {
    const f32x4 b = tab_entry(rs, lane, Ly.bias, to);
#pragma unroll
    for (int k = 0; k < kEpgxNB; ++k) acc[to][k] = b;
  }
#pragma unroll
  for (int ti = 0; ti < NTI; ++ti) {
    f32x4 a[NTO];
#pragma unroll
    for (int to = 0; to < NTO; ++to) a[to] = tab_entry(rs, lane, Ly.fwd, to * NTI + ti);
    // consecutive instructions go to different accumulators
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int to = 0; to < NTO; ++to)
#pragma unroll
        for (int k = 0; k < kEpgxNB; ++k) acc[to][k] = mma16(a[to][r], act[ti][k][r], acc[to][k]);
  }
  if (hidden) {
    // leaky ReLU; padded units stay zero (zero weights and bias)
#pragma unroll
    for (int k = 0; k < kEpgxNB; ++k) {
      uint32_t m = 0;
#pragma unroll
      for (int to = 0; to < NTO; ++to) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float z = acc[to][k][r];
          const bool pos = z > 0.f;
          m |= uint32_t(pos) << (to * 4 + r);
          act[to][k][r] = pos ? z : alpha * z;
        }
      }
      msk[k][lane] = m;
    }
  } else {
#pragma unroll
    for (int to = 0; to < NTO; ++to)
#pragma unroll
      for (int k = 0; k < kEpgxNB; ++k) act[to][k] = acc[to][k];
  }
}

// One reverse layer (the transpose of a forward layer with NTI input and NTO
// output tiles): g holds d out / d (output of the layer) on entry and
// d out / d (input of the layer) on return, through the derivative of the
// leaky ReLU whose signs msk holds (nullptr: the network's inputs).
template <int NTI, int NTO>
GC_DEV void epgx_rev_layer(EpgxRsrc rs, const EpgxLayer& Ly, float alpha,
                           f32x4 (&g)[kEpgxMaxTiles][kEpgxNB], const uint32_t (*msk)[64], int lane) {
  f32x4 acc[NTI][kEpgxNB];
#pragma unroll
  for (int ti = 0; ti < NTI; ++ti)
#pragma unroll
    for (int k = 0; k < kEpgxNB; ++k) acc[ti][k] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int to = 0; to < NTO; ++to) {
    f32x4 a[NTI];
#pragma unroll
    for (int ti = 0; ti < NTI; ++ti) a[ti] = tab_entry(rs, lane, Ly.rev, ti * NTO + to);
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int ti = 0; ti < NTI; ++ti)
#pragma unroll
        for (int k = 0; k < kEpgxNB; ++k) acc[ti][k] = mma16(a[ti][r], g[to][k][r], acc[ti][k]);
  }
#pragma unroll
  for (int k = 0; k < kEpgxNB; ++k) {
    const uint32_t m = msk ? msk[k][lane] : 0xffffffffu;
#pragma unroll
    for (int ti = 0; ti < NTI; ++ti)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float v = acc[ti][k][r];
        g[ti][k][r] = ((m >> (ti * 4 + r)) & 1u) ? v : alpha * v;
      }
  }
}

// The layer kernels above for the layer's tile counts (uniform branches over
// the kEpgxMaxTiles^2 shapes).
template <int S = 0>
GC_DEV void epgx_fwd_any(EpgxRsrc rs, const EpgxLayer& Ly, bool hidden, float alpha,
                         f32x4 (&act)[kEpgxMaxTiles][kEpgxNB], uint32_t (&msk)[kEpgxNB][64], int lane) {
  constexpr int NTI = S / kEpgxMaxTiles + 1, NTO = S % kEpgxMaxTiles + 1;
  if (Ly.nti == NTI && Ly.nto == NTO)
    epgx_fwd_layer<NTI, NTO>(rs, Ly, hidden, alpha, act, msk, lane);
  else if constexpr (S + 1 < kEpgxMaxTiles * kEpgxMaxTiles)
    epgx_fwd_any<S + 1>(rs, Ly, hidden, alpha, act, msk, lane);
}

template <int S = 0>
GC_DEV void epgx_rev_any(EpgxRsrc rs, const EpgxLayer& Ly, float alpha,
                         f32x4 (&g)[kEpgxMaxTiles][kEpgxNB], const uint32_t (*msk)[64], int lane) {
  constexpr int NTI = S / kEpgxMaxTiles + 1, NTO = S % kEpgxMaxTiles + 1;
  if (Ly.nti == NTI && Ly.nto == NTO)
    epgx_rev_layer<NTI, NTO>(rs, Ly, alpha, g, msk, lane);
  else if constexpr (S + 1 < kEpgxMaxTiles * kEpgxMaxTiles)
    epgx_rev_any<S + 1>(rs, Ly, alpha, g, msk, lane);
}

// Forward pass of network `net` over the wavefront's column blocks: act
// holds the input tiles on entry (features in tile 0) and the output tile on
// return.  Hidden layers record their sign masks in L.msk.
GC_DEV void epgx_forward(const EpgxNets& N, int net, f32x4 (&act)[kEpgxMaxTiles][kEpgxNB], EpgxLds& L, int lane) {
  const int nl = N.nl[net];
  const EpgxRsrc rs = epgx_rsrc(N);
  for (int l = 0; l < nl; ++l) epgx_fwd_any(rs, N.layer[net][l], l + 1 < nl, N.alpha[net], act, L.msk[l], lane);
}

// Reverse pass of network `net` for output o: on return g[0] holds
// d out_o / d feature for units 4 (l/16) + r of the lane's columns.
GC_DEV void epgx_reverse(const EpgxNets& N, int net, int o, f32x4 (&g)[kEpgxMaxTiles][kEpgxNB], const EpgxLds& L,
                         int lane) {
  const int nl = N.nl[net];
  const float alpha = N.alpha[net];
  // d out_o / d (input of the last layer) = row o of its weights, through the
  // derivative of the leaky ReLU that produced that input.  A whole row of
  // kEpgxMaxTiles entries is read (the table buffer's zero tail keeps it in
  // bounds) so that the loads need no branch.
  const int nti = N.layer[net][nl - 1].nti;
  const EpgxRsrc rs = epgx_rsrc(N);
  const int32_t row = N.last[net] + o * nti * 256;
#pragma unroll
  for (int k = 0; k < kEpgxNB; ++k) {
    const uint32_t m = (nl >= 2) ? L.msk[nl - 2][k][lane] : 0xffffffffu;
#pragma unroll
    for (int t = 0; t < kEpgxMaxTiles; ++t) {
      const f32x4 w = tab_entry(rs, lane, row, t);
#pragma unroll
      for (int r = 0; r < 4; ++r)
        g[t][k][r] = t < nti ? (((m >> (t * 4 + r)) & 1u) ? w[r] : alpha * w[r]) : 0.f;
    }
  }
  for (int l = nl - 2; l >= 0; --l) epgx_rev_any(rs, N.layer[net][l], alpha, g, l >= 1 ? L.msk[l - 1] : nullptr, lane);
}

// MVF of MCR-MWI's parameter slots (S0, MWF, IWF, ...), as McrmwiModel::amplitudes
GC_DEV float epgx_mvf(const float* th, float rho) {
  const float S0MW = th[0] * th[1];
  const float S0IW = th[0] * (1.f - th[1]) * th[2];
  const float S0EW = th[0] * (1.f - th[1]) * (1.f - th[2]);
  const float tv = fmaxf(S0IW + S0EW + (1.f / rho) * S0MW, 1e-8f);
  return fmaxf(((1.f / rho) * S0MW) / tv, 1e-8f);
}

// One wavefront: flip angle fi of voxels v0 .. v0 + kEpgxSamples - 1 of a
// volume.  Lanes l < kEpgxSamples compute the features of sample l (voxel
// v0 + l); the matrix passes treat sample 16 k + j as column j of block k.
// All lanes of the wavefront run every step (samples beyond nv are computed
// on the last voxel and not stored).
template <class TF>
GC_DEV void epgx_wave(const gc_model& md, const EpgxNets& N, const float* __restrict__ z,
                      const float* __restrict__ ex, float* __restrict__ aux, int64_t nv, int64_t v0, int fi,
                      TF&& transform, EpgxLds& L) {
  const int lane = threadIdx.x & 63;
  const int nfa = md.i[1];
  if (lane < kEpgxSamples) {
    const int64_t v = v0 + lane;
    const int64_t vv = v < nv ? v : nv - 1;
    float th[7];
#pragma unroll
    for (int p = 0; p < 7; ++p) th[p] = transform(z[int64_t(p) * nv + vv], md.lb[p], md.ub[p]);
    const float mvf = epgx_mvf(th, md.f[3]);
    const int f = fi < nfa ? fi : nfa - 1;
    epgx_features(mvf, th[3], th[4], th[6], md.table[md.i[0] + f] * ex[vv], md.f[0], md.f[4], L, lane);
  }
  __syncthreads();
  const int g = lane >> 4, j = lane & 15;
  int row = 0;  // aux row of the next output
#pragma unroll 1
  for (int net = 0; net < 2; ++net) {
    f32x4 act[kEpgxMaxTiles][kEpgxNB];
#pragma unroll
    for (int t = 0; t < kEpgxMaxTiles; ++t)
#pragma unroll
      for (int k = 0; k < kEpgxNB; ++k) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int u = 16 * t + 4 * g + r;
          act[t][k][r] = u < 11 ? L.xs[(u < 11 ? u : 0) * 5][16 * k + j] : 0.f;
        }
      }
    epgx_forward(N, net, act, L, lane);
    // outputs: unit o of the last layer sits in register o of lanes with g = 0
    float val[kEpgxNB][2];
#pragma unroll
    for (int k = 0; k < kEpgxNB; ++k) {
      val[k][0] = act[0][k][0];
      val[k][1] = act[0][k][1];
    }
#pragma unroll 1
    for (int o = 0; o < N.nout[net]; ++o) {
      f32x4 gr[kEpgxMaxTiles][kEpgxNB];
      epgx_reverse(N, net, o, gr, L, lane);
#pragma unroll
      for (int k = 0; k < kEpgxNB; ++k) {
        const int64_t v = v0 + 16 * k + j;
        float d[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          float s = 0.f;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int u = 4 * g + r;
            if (u < 11) s = fmaf(gr[0][k][r], L.xs[u * 5 + 1 + q][16 * k + j], s);
          }
          // the four lane groups hold units 0-3, 4-7, 8-11, 12-15 of one column
          s += __shfl_xor(s, 16);
          s += __shfl_xor(s, 32);
          d[q] = s;
        }
        if (g == 0 && v < nv && fi < nfa) {
          float* __restrict__ out = aux + int64_t(fi) * nv + v;
          const int64_t rs = int64_t(nfa) * nv;
          out[int64_t(row) * rs] = o == 0 ? val[k][0] : val[k][1];
#pragma unroll
          for (int q = 0; q < 4; ++q) out[int64_t(row + 1 + q) * rs] = d[q];
        }
      }
      row += 5;
    }
  }
}

// the step form: aux rows of every active volume's voxels from z_in.
// Workgroup b of volume i (counted from kEpgxGroupsPerBlock blk0,
// kEpgxSamples voxels each) has one wavefront per flip angle.
template <int TR>
__global__ __launch_bounds__(kBlock, 3) void k_epgx_step(const gc_model* __restrict__ mdp, const StepParams sp,
                                                      const EpgxNets N) {
  int vi = 0;
  for (int i = 1; i < sp.n_vol; ++i)
    if (int(blockIdx.x) >= kEpgxGroupsPerBlock * sp.vols[i].blk0) vi = i;
  const StepVol& V = sp.vols[vi];
  if (!V.active) return;  // uniform over the block
  const int64_t v0 = int64_t(blockIdx.x - kEpgxGroupsPerBlock * V.blk0) * kEpgxSamples;
  if (v0 >= V.nv) return;  // uniform over the block
  __shared__ EpgxLds L[kBlock / 64];
  const int w = threadIdx.x >> 6;
  epgx_wave(*mdp, N, V.z_in, V.ex, V.aux, V.nv, v0, w,
            [](float zz, float lb, float ub) {
              float t, d;
              transform_fwd_t<TR>(zz, lb, ub, t, d);
              return t;
            },
            L[w]);
}

// the one-shot form (gc_fwd, gc_fwd_adj, the askadam outputs): aux rows
// from parameters x in the given transform space
__global__ __launch_bounds__(kBlock, 3) static void k_epgx_eval(const gc_model* __restrict__ mdp,
                                                             const float* __restrict__ x,
                                                             const float* __restrict__ ex, int64_t nv,
                                                             int32_t transform, float* __restrict__ aux,
                                                             const EpgxNets N) {
  const int64_t v0 = int64_t(blockIdx.x) * kEpgxSamples;
  __shared__ EpgxLds L[kBlock / 64];
  const int w = threadIdx.x >> 6;
  epgx_wave(*mdp, N, x, ex, aux, nv, v0, w,
            [transform](float zz, float lb, float ub) {
              float t, d;
              transform_fwd(transform, zz, lb, ub, t, d);
              return t;
            },
            L[w]);
}

}  // namespace gc
