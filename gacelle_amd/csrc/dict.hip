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
// (zero padded); wdt the same for 2 w_k d_kj; per entry inv_n = 1 / n_j
// (NaN for the padding entries, which therefore never win) and q_j.
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
    const double v = (k < m && j < nd) ? double(dict[int64_t(k) * nd + j]) : 0.0;
    dt[e] = v;
    if (W) wdt[e] = (k < m) ? 2.0 * double(wv.w[k]) * v : 0.0;
  }
  if (e < ntiles * kEntTile) {
    double n = 0.0, qq = 0.0;
    if (e < nd) {
      for (int k = 0; k < m; ++k) {
        const double v = dict[int64_t(k) * nd + e];
        n = fma(v, v, n);
        if (W) qq = fma(double(wv.w[k]) * v, v, qq);
      }
    }
    inv_n[e] = (e < nd) ? 1.0 / n : __builtin_nan("");
    if (W) q[e] = qq;
  }
}

// One wavefront scores 64 voxels (four 16-voxel row blocks) against every
// entry, 16 entries per step.  Per lane, the best score of each of its 16
// slots (row block, value r = row l / 16 + 4 r) over the entries of its
// column class; then a 16-lane butterfly picks the best per voxel (lowest
// index on ties).
template <int KS, bool W>
__global__ __launch_bounds__(kDictBlock) void k_dict_search(const float* __restrict__ y, int64_t nv, int m,
                                                           const double* __restrict__ dt,
                                                           const double* __restrict__ wdt,
                                                           const double* __restrict__ inv_n,
                                                           const double* __restrict__ q, int64_t ntiles,
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

  for (int64_t T = 0; T < ntiles; ++T) {
    const double* dtt = dt + T * (KS * 64) + lane;
    double b[KS], bw[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      b[ks] = dtt[ks * 64];
      if (W) bw[ks] = wdt[T * (KS * 64) + lane + ks * 64];
    }
    const double in = inv_n[T * kEntTile + col];
    const double qq = W ? q[T * kEntTile + col] : 0.0;
    const int32_t jidx = int32_t(T * kEntTile + col);
#pragma unroll
    for (int t = 0; t < kVoxTiles; ++t) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      d4 acc2 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][ks], b[ks], acc, 0, 0, 0);
        if (W) acc2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t][ks], bw[ks], acc2, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double tt = acc[r] * in;
        const double s = W ? tt * fma(-tt, qq, acc2[r]) : tt * acc[r];
        if (s > bs[t][r]) {
          bs[t][r] = s;
          bi[t][r] = jidx;
        }
      }
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
  const int64_t vv = v0 + (col >> 2) * 16 + grp + 4 * (col & 3);
  if (vv < nv) {
    best[vv] = my_i;
    if (sse) {
      double y2 = 0.0;
      for (int k = 0; k < m; ++k) {
        const double yk = y[int64_t(k) * nv + vv];
        y2 = fma(W ? double(wv.w[k]) * yk : yk, yk, y2);
      }
      sse[vv] = my_i >= 0 ? y2 - my_s : __builtin_nan("");
    }
  }
}

template <int KS, bool W>
void launch_search(hipStream_t s, const float* y, int64_t nv, int m, const double* dt, const double* wdt,
                   const double* inv_n, const double* q, int64_t ntiles, const DictWeights& wv, int32_t* best,
                   double* sse) {
  const int64_t waves = (nv + 63) / 64;
  const unsigned grid = unsigned((waves + (kDictBlock / 64) - 1) / (kDictBlock / 64));
  hipLaunchKernelGGL((k_dict_search<KS, W>), dim3(grid), dim3(kDictBlock), 0, s, y, nv, m, dt, wdt, inv_n, q,
                     ntiles, wv, best, sse);
}

template <bool W>
void dispatch(int ks, hipStream_t s, const float* y, int64_t nv, int m, const double* dt, const double* wdt,
              const double* inv_n, const double* q, int64_t ntiles, const DictWeights& wv, int32_t* best,
              double* sse) {
  switch (ks) {
    case 1: launch_search<1, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
    case 2: launch_search<2, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
    case 3: launch_search<3, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
    case 4: launch_search<4, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
    case 5: launch_search<5, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
    case 6: launch_search<6, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
    case 7: launch_search<7, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
    default: launch_search<8, W>(s, y, nv, m, dt, wdt, inv_n, q, ntiles, wv, best, sse); break;
  }
}

}  // namespace

int dict_search(gc_ctx* ctx, const float* y, int64_t nv, int m, const float* dict, int64_t nd, const float* w,
                int32_t* best, double* sse) {
  if (m < 1 || m > kDictMaxMeas)
    return set_err(ctx, GC_ERR_ARG, "gc_dict_search: m = %d outside 1..%d", m, kDictMaxMeas);
  if (nv < 0 || nd < 1 || nd > INT32_MAX - kEntTile)
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
  GC_HIP(ctx, inv_n.alloc(sizeof(double) * ntiles * kEntTile));
  if (weighted) {
    GC_HIP(ctx, wdt.alloc(sizeof(double) * n_el));
    GC_HIP(ctx, q.alloc(sizeof(double) * ntiles * kEntTile));
  }
  const int64_t n_prep = std::max<int64_t>(n_el, ntiles * kEntTile);
  if (weighted)
    hipLaunchKernelGGL(k_dict_prep<true>, dim3(nblocks(n_prep)), dim3(kBlock), 0, ctx->stream, dict, nd, m, ks_n,
                       ntiles, wv, dt.as<double>(), wdt.as<double>(), inv_n.as<double>(), q.as<double>());
  else
    hipLaunchKernelGGL(k_dict_prep<false>, dim3(nblocks(n_prep)), dim3(kBlock), 0, ctx->stream, dict, nd, m, ks_n,
                       ntiles, wv, dt.as<double>(), nullptr, inv_n.as<double>(), nullptr);
  GC_HIP(ctx, hipGetLastError());
  if (weighted)
    dispatch<true>(ks_n, ctx->stream, y, nv, m, dt.as<double>(), wdt.as<double>(), inv_n.as<double>(),
                   q.as<double>(), ntiles, wv, best, sse);
  else
    dispatch<false>(ks_n, ctx->stream, y, nv, m, dt.as<double>(), nullptr, inv_n.as<double>(), nullptr, ntiles, wv,
                    best, sse);
  GC_HIP(ctx, hipGetLastError());
  return GC_OK;
}

}  // namespace gc
