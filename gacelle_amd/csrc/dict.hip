// dict.hip -- dictionary search on the FP64 matrix cores.
//
// The reference's likelihood start (NEXI/gpuNEXI.m:822-849) scores every
// dictionary entry d_j against a voxel's signal y with the scaled
// projection  c_j = (y . d_j) / (d_j . d_j)  and keeps the entry of least
//   sse_j = sum_k w_k (y_k - c_j d_kj)^2,
// the first one on ties.  Expanded,
//   sse_j = Y2 - score_j,   score_j = t_j (2 b_j - t_j q_j),
//   t_j = a_j / n_j,  a_j = y . d_j,  b_j = sum_k w_k y_k d_kj,
//   n_j = d_j . d_j,  q_j = sum_k w_k d_kj^2,  Y2 = sum_k w_k y_k^2,
// and with unit weights score_j = a_j^2 / n_j.  a_j (and 2 b_j) are a
// [voxels x m] x [m x entries] product: v_mfma_f64_16x16x4f64 forms them
// exactly enough (FP64 products of FP32 inputs) that the cancellation in
// Y2 - score costs no ranking accuracy, and the per-entry epilogue (two or
// three FP64 operations and a compare) runs on the vector ALUs beside it.
//
// Operand layouts of v_mfma_f64_16x16x4f64 (wavefront of 64, lane l):
//   A[i][k]: i = l % 16, k = l / 16       (voxel rows, measurement k)
//   B[k][j]: j = l % 16, k = l / 16       (entry columns)
//   D[i][j]: j = l % 16, i = l / 16 + 4 r for the r-th of four values
//            (unlike the FP32 16x16x4 forms, whose r-th value is row
//            4 (l / 16) + r).
#include "gc_dict.hpp"

namespace gc {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kEntTile = 16;   // entries per MFMA column block
constexpr int kVoxTiles = 4;   // 16-voxel row blocks per wavefront
constexpr int kDictBlock = 256;

struct DictWeights {
  float w[kDictMaxMeas];
};

// Dictionary in B-operand order, dt[tile][ks][lane] = d[4 ks + lane/16][16 tile + lane%16]
// zero padded in k.  With weights: wdt the same for 2 w_k d_kj, and per
// entry inv_n = 1 / n_j and q_j.  Without: each column is stored
// normalised, d_kj / sqrt(n_j), so the product is a_j / sqrt(n_j) itself.
// Padding entries (j >= nd) get NaN, which never wins; so does a zero
// entry (0 * inf), in both forms.
template <bool W>
__global__ void k_dict_prep(const float* __restrict__ dict, int64_t nd, int m, int ks_n, int64_t ntiles, DictWeights wv,
                            double* __restrict__ dt, double* __restrict__ wdt, double* __restrict__ inv_n,
                            double* __restrict__ q) {
  const int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n_el = ntiles * ks_n * 64;
  if (e < n_el) {
    const int64_t tile = e / (ks_n * 64);
    const int rem = int(e - tile * ks_n * 64);
    const int ks = rem >> 6, lane = rem & 63;
    const int k = ks * 4 + (lane >> 4);
    const int64_t j = tile * kEntTile + (lane & 15);
    double v;
    if (j >= nd) {
      v = W ? 0.0 : __builtin_nan("");
    } else if (k >= m) {
      v = 0.0;
    } else {
      v = double(dict[int64_t(k) * nd + j]);
      if (!W) {
        double n = 0.0;
        for (int kk = 0; kk < m; ++kk) {
          const double u = dict[int64_t(kk) * nd + j];
          n = fma(u, u, n);
        }
        v *= 1.0 / sqrt(n);
      }
    }
    dt[e] = v;
    if (W) wdt[e] = (k < m) ? 2.0 * double(wv.w[k]) * v : 0.0;
  }
  if (W && e < ntiles * kEntTile) {
    double n = 0.0, qq = 0.0;
    if (e < nd) {
      for (int k = 0; k < m; ++k) {
        const double v = dict[int64_t(k) * nd + e];
        n = fma(v, v, n);
        qq = fma(double(wv.w[k]) * v, v, qq);
      }
    }
    inv_n[e] = (e < nd) ? 1.0 / n : __builtin_nan("");
    q[e] = qq;
  }
}

// max(|x|, best) as one v_max_f64 with an abs source modifier (fmax would
// add canonicalising maxes for signalling NaNs; every NaN here comes out of
// an arithmetic instruction and is quiet, and a quiet NaN operand returns
// the other one).  ALU-only inline asm.
GC_DEV double max_abs(double x, double best) {
  double r;
  asm("v_max_f64 %0, |%1|, %2" : "=v"(r) : "v"(x), "v"(best));
  return r;
}

// (mask & ins) | (~mask & base) as one v_bfi_b32 (the mask and the insert
// are both wave-uniform, and a VOP3 reads one scalar register at most, so
// the compiler would otherwise split it into an and and an or).
GC_DEV uint32_t bitfield_insert(uint32_t mask, uint32_t ins, uint32_t base) {
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(mask), "v"(ins), "v"(base));
  return r;
}

// The winner's sse, from scratch in FP64: Y2 - t (2 b - t q), t = a / n.
GC_DEV double exact_sse(const float* __restrict__ y, int64_t nv, int64_t v, const float* __restrict__ dict, int64_t nd,
                        int64_t j, int m, const DictWeights& wv, bool weighted) {
  double y2 = 0.0, a = 0.0, b = 0.0, n = 0.0, q = 0.0;
  for (int k = 0; k < m; ++k) {
    const double yk = y[int64_t(k) * nv + v];
    const double dk = dict[int64_t(k) * nd + j];
    const double wk = weighted ? double(wv.w[k]) : 1.0;
    y2 = fma(wk * yk, yk, y2);
    a = fma(yk, dk, a);
    b = fma(wk * yk, dk, b);
    n = fma(dk, dk, n);
    q = fma(wk * dk, dk, q);
  }
  const double t = a / n;
  return y2 - t * (2.0 * b - t * q);
}

// One wavefront scores 64 voxels (four 16-voxel row blocks) against every
// entry, 16 entries per step; the four row blocks accumulate independently
// (k-step outer) so their MFMA chains overlap.  Per lane, the best of each
// of its 16 slots (row block, value r = row l / 16 + 4 r) over the entries
// of its column class; then a 16-lane butterfly picks the best per voxel
// (lowest index on ties).
//
// Without weights the slot keeps max |a_j / sqrt(n_j)| (the product with
// the normalised column; v_max_f64 with an abs modifier; NaN never wins), and the step's tile index rides in the
// low kbits bits of the mantissa as (2^kbits - 1 - tile): equal values
// then favour the earlier tile, and the ranking is exact to 2^(kbits - 52)
// relative, far below the FP32 rounding of the inputs.  With weights the
// slot keeps the exact score t (2 b - t q) and the entry index beside it
// (scores may be negative, where the mantissa order would reverse).
template <int KS, bool W>
__global__ __launch_bounds__(kDictBlock) void k_dict_search(const float* __restrict__ y, int64_t nv, int m,
                                                           const double* __restrict__ dt,
                                                           const double* __restrict__ wdt,
                                                           const double* __restrict__ inv_n,
                                                           const double* __restrict__ q, int64_t ntiles,
                                                           uint32_t kmask, const float* __restrict__ dict, int64_t nd,
                                                           DictWeights wv, int32_t* __restrict__ best,
                                                           double* __restrict__ sse) {
  const int lane = threadIdx.x & 63;
  const int64_t v0 = ((int64_t(blockIdx.x) * kDictBlock + threadIdx.x) >> 6) * 64;
  if (v0 >= nv) return;  // uniform per wavefront
  const int col = lane & 15, grp = lane >> 4;

  double a[kVoxTiles][KS];
#pragma unroll
  for (int t = 0; t < kVoxTiles; ++t) {
    const int64_t v = v0 + t * 16 + col;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int k = ks * 4 + grp;
      a[t][ks] = (k < m && v < nv) ? double(y[int64_t(k) * nv + v]) : 0.0;
    }
  }

  double bs[kVoxTiles][4];
  int32_t bi[kVoxTiles][4];
#pragma unroll
  for (int t = 0; t < kVoxTiles; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bs[t][r] = -__builtin_inf();
      bi[t][r] = -1;
    }

  // the next step's dictionary operands load while this step computes
  double bn[KS], bwn[KS], inn = 0.0, qqn = 0.0;
  auto load_step = [&](int64_t T) {
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      bn[ks] = dt[T * (KS * 64) + ks * 64 + lane];
      if (W) bwn[ks] = wdt[T * (KS * 64) + ks * 64 + lane];
    }
    if (W) {
      inn = inv_n[T * kEntTile + col];
      qqn = q[T * kEntTile + col];
    }
  };
  load_step(0);
  for (int64_t T = 0; T < ntiles; ++T) {
    double b[KS], bw[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      b[ks] = bn[ks];
      if (W) bw[ks] = bwn[ks];
    }
    const double in = inn;
    const double qq = qqn;
    if (T + 1 < ntiles) load_step(T + 1);
    d4 acc[kVoxTiles], acc2[kVoxTiles];
#pragma unroll
    for (int t = 0; t < kVoxTiles; ++t) {
      acc[t] = d4{0.0, 0.0, 0.0, 0.0};
      if (W) acc2[t] = d4{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int t = 0; t < kVoxTiles; ++t) {
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][ks], b[ks], acc[t], 0, 0, 0);
        if (W) acc2[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][ks], bw[ks], acc2[t], 0, 0, 0);
      }
    if (W) {
      const int32_t jidx = int32_t(T * kEntTile + col);
#pragma unroll
      for (int t = 0; t < kVoxTiles; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double tt = acc[t][r] * in;
          const double sc = tt * fma(-tt, qq, acc2[t][r]);
          if (sc > bs[t][r]) {
            bs[t][r] = sc;
            bi[t][r] = jidx;
          }
        }
    } else {
      const uint32_t code = kmask - uint32_t(T);
#pragma unroll
      for (int t = 0; t < kVoxTiles; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const uint64_t bits = uint64_t(__double_as_longlong(acc[t][r]));
          const uint32_t lo = bitfield_insert(kmask, code, uint32_t(bits));
          const uint64_t tagged = (bits & 0xFFFFFFFF00000000ull) | lo;
          bs[t][r] = max_abs(__longlong_as_double(int64_t(tagged)), bs[t][r]);
        }
    }
  }

  if (!W) {  // decode the entry of each slot: tile from the mantissa, column from the lane
#pragma unroll
    for (int t = 0; t < kVoxTiles; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t lo = uint32_t(uint64_t(__double_as_longlong(bs[t][r])));
        bi[t][r] = bs[t][r] >= 0.0 ? int32_t((kmask - (lo & kmask)) * kEntTile + col) : -1;
      }
  }

  // best per voxel across the 16 column classes; lane (grp, col) then
  // writes the voxel of row block col / 4, row grp + 4 (col % 4).
  double my_s = -__builtin_inf();
  int32_t my_i = -1;
#pragma unroll
  for (int t = 0; t < kVoxTiles; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double s = bs[t][r];
      int32_t i = bi[t][r];
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) {
        const double so = __shfl_xor(s, off);
        const int32_t io = __shfl_xor(i, off);
        if (so > s || (so == s && uint32_t(io) < uint32_t(i))) {
          s = so;
          i = io;
        }
      }
      if (t * 4 + r == col) {
        my_s = s;
        my_i = i;
      }
    }
  (void)my_s;
  const int64_t vv = v0 + (col >> 2) * 16 + grp + 4 * (col & 3);
  if (vv < nv) {
    best[vv] = my_i;
    if (sse) sse[vv] = my_i >= 0 ? exact_sse(y, nv, vv, dict, nd, my_i, m, wv, W) : __builtin_nan("");
  }
}

struct SearchArgs {
  const float* y;
  int64_t nv;
  int m;
  const double *dt, *wdt, *inv_n, *q;
  int64_t ntiles;
  uint32_t kmask;
  const float* dict;
  int64_t nd;
  DictWeights wv;
  int32_t* best;
  double* sse;
};

template <int KS, bool W>
void launch_search(hipStream_t s, const SearchArgs& g) {
  const int64_t waves = (g.nv + 63) / 64;
  const unsigned grid = unsigned((waves + (kDictBlock / 64) - 1) / (kDictBlock / 64));
  hipLaunchKernelGGL((k_dict_search<KS, W>), dim3(grid), dim3(kDictBlock), 0, s, g.y, g.nv, g.m, g.dt, g.wdt,
                     g.inv_n, g.q, g.ntiles, g.kmask, g.dict, g.nd, g.wv, g.best, g.sse);
}

template <bool W>
void dispatch(int ks, hipStream_t s, const SearchArgs& g) {
  switch (ks) {
    case 1: launch_search<1, W>(s, g); break;
    case 2: launch_search<2, W>(s, g); break;
    case 3: launch_search<3, W>(s, g); break;
    case 4: launch_search<4, W>(s, g); break;
    case 5: launch_search<5, W>(s, g); break;
    case 6: launch_search<6, W>(s, g); break;
    case 7: launch_search<7, W>(s, g); break;
    default: launch_search<8, W>(s, g); break;
  }
}

}  // namespace

int dict_search(gc_ctx* ctx, const float* y, int64_t nv, int m, const float* dict, int64_t nd, const float* w,
                int32_t* best, double* sse) {
  if (m < 1 || m > kDictMaxMeas)
    return set_err(ctx, GC_ERR_ARG, "gc_dict_search: m = %d outside 1..%d", m, kDictMaxMeas);
  if (nv < 0 || nd < 1 || nd > (int64_t(1) << 24))
    return set_err(ctx, GC_ERR_ARG, "gc_dict_search: bad sizes (nv = %lld, nd = %lld)", (long long)nv, (long long)nd);
  if (!y || !dict || !best) return set_err(ctx, GC_ERR_ARG, "gc_dict_search: NULL buffer");
  if (nv == 0) return GC_OK;
  const bool weighted = w != nullptr;
  DictWeights wv{};
  for (int k = 0; k < m; ++k) {
    wv.w[k] = weighted ? w[k] : 1.0f;
    if (!(wv.w[k] >= 0.0f) || !std::isfinite(wv.w[k]))
      return set_err(ctx, GC_ERR_ARG, "gc_dict_search: weight %d is negative or not finite", k);
  }
  const int ks_n = (m + 3) / 4;
  const int64_t ntiles = (nd + kEntTile - 1) / kEntTile;
  const int64_t n_el = ntiles * ks_n * 64;
  DevBuf dt, wdt, inv_n, q;
  GC_HIP(ctx, dt.alloc(sizeof(double) * n_el));
  if (weighted) {
    GC_HIP(ctx, inv_n.alloc(sizeof(double) * ntiles * kEntTile));
    GC_HIP(ctx, wdt.alloc(sizeof(double) * n_el));
    GC_HIP(ctx, q.alloc(sizeof(double) * ntiles * kEntTile));
  }
  const int64_t n_prep = weighted ? std::max<int64_t>(n_el, ntiles * kEntTile) : n_el;
  if (weighted)
    hipLaunchKernelGGL(k_dict_prep<true>, dim3(nblocks(n_prep)), dim3(kBlock), 0, ctx->stream, dict, nd, m, ks_n,
                       ntiles, wv, dt.as<double>(), wdt.as<double>(), inv_n.as<double>(), q.as<double>());
  else
    hipLaunchKernelGGL(k_dict_prep<false>, dim3(nblocks(n_prep)), dim3(kBlock), 0, ctx->stream, dict, nd, m, ks_n,
                       ntiles, wv, dt.as<double>(), nullptr, nullptr, nullptr);
  GC_HIP(ctx, hipGetLastError());
  uint32_t kbits = 1;
  while ((int64_t(1) << kbits) < ntiles) ++kbits;
  SearchArgs g{y, nv, m, dt.as<double>(), wdt.as<double>(), inv_n.as<double>(), q.as<double>(), ntiles,
               (uint32_t(1) << kbits) - 1u, dict, nd, wv, best, sse};
  if (weighted)
    dispatch<true>(ks_n, ctx->stream, g);
  else
    dispatch<false>(ks_n, ctx->stream, g);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

}  // namespace gc
