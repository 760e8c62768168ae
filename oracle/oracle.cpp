// oracle/oracle.cpp -- CPU restatement of GACELLE's voxelwise-estimation
// hot path: askadam (utils/askadam.m), mcmc MH / Goodman-Weare
// (utils/mcmc.m), logP_Gaussian (utils/logP_Gaussian.m) and
// spatial_total_variation (utils/spatial_total_variation.m).
//
// TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP library's parity
// tests and the CPU baseline timed beside it; nothing in the library links
// or calls it.  Arithmetic is templated: double for the parity oracle,
// float (+ OpenMP over voxels) for the CPU-path baseline.
//
// Every function cites the reference lines it restates.  Semantics of the
// closed MATLAB built-ins (adamupdate, l1loss, ...) are stated where used.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>

#include "../include/gacelle.h"
#include "dual.hpp"
#include "models.hpp"

namespace orc {

// ===========================================================================
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random numbers:
// as easy as 1, 2, 3"), written from the paper's round function.  The native
// RNG stream of the library is defined in terms of this generator.
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
  const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
  const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0);
  const uint32_t hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  c[0] = n0;
  c[1] = lo1;
  c[2] = n2;
  c[3] = lo0;
}

static inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
    philox_round(c, k);
  }
  for (int i = 0; i < 4; ++i) out[i] = c[i];
}

// uint32 -> uniform in (0,1): ((x >> 8) + 0.5) * 2^-24 (exact in float)
static inline double u01(uint32_t x) { return (double(x >> 8) + 0.5) * (1.0 / 16777216.0); }

// Native stream layout.  Counter = {iteration, chain_lo, chain_hi, purpose}.
enum : uint32_t { PURPOSE_NORMAL0 = 0, PURPOSE_ZZ = 16, PURPOSE_INIT = 32 };

static inline void draw(uint64_t seed, uint32_t it, uint64_t chain, uint32_t purpose, uint32_t out[4]) {
  const uint32_t ctr[4] = {it, uint32_t(chain), uint32_t(chain >> 32), purpose};
  const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
  philox4x32_10(ctr, key, out);
}

// Goodman-Weare native draws of walker `chain` at iteration `it` and
// (red-black) half `sub`: one Philox call gives the stretch uniform (o[0])
// and the acceptance uniform (o[1]).
static inline void native_gw_uniforms(uint64_t seed, uint32_t it, uint64_t chain, int sub, double* u_zz,
                                      double* u_acc) {
  uint32_t o[4];
  draw(seed, it, chain, PURPOSE_ZZ + uint32_t(sub), o);
  *u_zz = u01(o[0]);
  *u_acc = u01(o[1]);
}

// Box-Muller normal number q of a draw group (4 normals per Philox call).
static inline double native_normal(uint64_t seed, uint32_t it, uint64_t chain, uint32_t base, int q) {
  uint32_t o[4];
  draw(seed, it, chain, base + uint32_t(q / 4), o);
  const int pair = (q % 4) / 2;
  const double u1 = u01(o[2 * pair]), u2 = u01(o[2 * pair + 1]);
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double a = 2.0 * kPi * u2;
  return (q % 2 == 0) ? r * std::cos(a) : r * std::sin(a);
}

// Standard normal quantile sqrt(2) erfinv(2u - 1) with the single-precision
// erfinv of M. Giles, "Approximating the erfinv function" (GPU Computing
// Gems Jade Edition, 2011), evaluated in double with (1-x)(1+x) = 4u(1-u).
static inline double norm_inv(double u) {
  const double x = 2.0 * u - 1.0;
  double w = -std::log(4.0 * u * (1.0 - u));
  double p;
  if (w < 5.0) {
    w -= 2.5;
    p = 2.81022636e-08;
    p = 3.43273939e-07 + p * w;
    p = -3.5233877e-06 + p * w;
    p = -4.39150654e-06 + p * w;
    p = 0.00021858087 + p * w;
    p = -0.00125372503 + p * w;
    p = -0.00417768164 + p * w;
    p = 0.246640727 + p * w;
    p = 1.50140941 + p * w;
  } else {
    w = std::sqrt(w) - 3.0;
    p = -0.000200214257;
    p = 0.000100950558 + p * w;
    p = 0.00134934322 + p * w;
    p = -0.00367342844 + p * w;
    p = 0.00573950773 + p * w;
    p = -0.0076224613 + p * w;
    p = 0.00943887047 + p * w;
    p = 1.00167406 + p * w;
    p = 2.83297682 + p * w;
  }
  return std::sqrt(2.0) * p * x;
}

// The MH native stream at (iteration it, chain): uniforms u[0..Q] from
// Philox groups of four (purposes PURPOSE_NORMAL0 + g); normals n[q] in
// Box-Muller pairs (u[q], u[q+1]), an odd last normal from norm_inv(u[Q-1]);
// the acceptance uniform is u[Q].
static inline void native_mh_draws(uint64_t seed, uint32_t it, uint64_t chain, int Q, double* n, double* u_acc) {
  double u[4 * ((GC_MAX_PAR + 1 + 3) / 4)];
  for (int g = 0; g < (Q + 1 + 3) / 4; ++g) {
    uint32_t o[4];
    draw(seed, it, chain, PURPOSE_NORMAL0 + uint32_t(g), o);
    for (int i = 0; i < 4; ++i) u[4 * g + i] = u01(o[i]);
  }
  for (int q = 0; q + 1 < Q; q += 2) {
    const double r = std::sqrt(-2.0 * std::log(u[q]));
    n[q] = r * std::cos(2.0 * kPi * u[q + 1]);
    n[q + 1] = r * std::sin(2.0 * kPi * u[q + 1]);
  }
  if (Q % 2 == 1) n[Q - 1] = norm_inv(u[Q - 1]);
  *u_acc = u[Q];
}

// Partner tables for the native GW stream (shared by all voxels, as the
// reference's find_partner / randi draws are, utils/mcmc.m:297,329,797-803).
// Simultaneous: a uniform derangement of 0..W-1 by Fisher-Yates shuffles
// rejected until no fixed point (the randperm-until-derangement loop).
// Red-black: for each half, na anchors drawn uniformly with replacement
// from the other half.
// `tag` selects the stream: kTagLocal for the voxelwise blocks, kTagGlobal
// for the global-parameter block of goodman_weare_wglobal_constant.
enum : uint32_t { kTagLocal = 0xFFFFu, kTagGlobal = 0xFFFCu };
// chain ids of the per-walker global draws (shared by all voxels)
static inline uint64_t global_chain(int rep, int wk) { return (uint64_t(0xFFFEu) << 48) | (uint64_t(rep) << 32) | uint64_t(wk); }
static inline uint64_t global_init_chain(int wk) { return (uint64_t(0xFFFDu) << 48) | uint64_t(wk); }

static void native_partners(uint64_t seed, int rep, int it, int W, int mode, int32_t* out, uint32_t tag = kTagLocal) {
  const uint64_t stream = (uint64_t(tag) << 48) | (uint64_t(rep) << 32);
  uint32_t attempt = 0;
  if (mode == GC_ENSEMBLE_SIMULTANEOUS) {
    for (;;) {
      for (int i = 0; i < W; ++i) out[i] = i;
      for (int i = W - 1; i > 0; --i) {
        uint32_t o[4];
        draw(seed, uint32_t(it), stream + attempt, uint32_t(i), o);
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
        draw(seed, uint32_t(it), stream, uint32_t(1000 + pos), o);
        out[pos++] = fro0 + int((uint64_t(o[0]) * uint64_t(nfro)) >> 32);
      }
    }
  }
}

// MATLAB's min(x, 1) ignores NaN: min(NaN, 1) = 1 (acceptance ratios of
// utils/mcmc.m:160 and :313).
template <class S>
static inline S min1(S x) { return std::isnan(x) ? S(1) : std::min(x, S(1)); }

// ===========================================================================
// logP_Gaussian, utils/logP_Gaussian.m:18-22
template <class S>
static inline S logp_gaussian(S sos, S noise, int nm) {
  return -sos / (S(2) * noise * noise) - S(nm) / S(2) * std::log(S(2) * S(kPi) * noise * noise);
}

// ===========================================================================
// Loss functions of askadam.model_gradient, utils/askadam.m:86-95.  The
// MATLAB built-ins (l1loss / l2loss / huber / mse with Reduction="none")
// are taken elementwise: w|d|, w d^2, w*huber_1(d), d^2/2 (unweighted).
template <class V, class S>
static inline V loss_elem(int kind, const V& r, S w) {
  switch (kind) {
    case GC_LOSS_L1: return fabs(r) * w;
    case GC_LOSS_L2: return r * r * w;
    case GC_LOSS_HUBER: {
      const V a = fabs(r);
      if (value(a) <= S(1)) return r * r * S(0.5) * w;
      return (a - S(0.5)) * w;
    }
    default: return r * r * S(0.5);
  }
}

// parameter transform z -> theta, utils/askadam.m:1063-1073; the linear
// branch clamps a local copy first (forward_pass, :424-426).
template <class V, class S>
static inline V unscale(int transform, const V& z, S lb, S ub) {
  if (transform == GC_TRANSFORM_SIGMOID) return S(lb) + S(ub - lb) * (S(1) / (S(1) + exp(-z)));
  if (transform == GC_TRANSFORM_LINEAR) return vmin(vmax(z, S(0)), S(1)) * S(ub - lb) + S(lb);
  return z;
}

// One voxel of askadam.model_gradient without the regulariser: the loss
// elements and d(sum of loss elements)/dz.
// codes (optional, L1 only): the sign of each residual to differentiate
// with (1 = +, 2 = -, 0 = no gradient), e.g. the signs the device kernel
// took at this iteration; the loss value stays w|r| of this evaluation.
// The L1 subgradient w*sign(r) is discontinuous at r = 0, where an FP32 and
// an FP64 evaluation can disagree on the sign: following one sign sequence
// isolates the rest of the arithmetic.
template <class S, int NP>
static void eval_voxel(const gc_model& md, int loss_kind, int transform, const S* z, const float* y,
                       const float* w, const S* extra, double& lsum, S* grad, S* elem,
                       const uint8_t* codes = nullptr) {
  using D = Dual<S, NP>;
  const int P = md.n_par;
  D th[GC_MAX_PAR];
  for (int p = 0; p < P; ++p) {
    const S zval = z[p];
    D zp = (p < NP) ? D::var(zval, p) : D(zval);
    th[p] = unscale<D, S>(transform, zp, S(md.lb[p]), S(md.ub[p]));
  }
  static thread_local std::vector<D> sbuf;
  if (int(sbuf.size()) < md.n_meas) sbuf.resize(md.n_meas);
  D* s = sbuf.data();
  fwd_model<D, S>(md, th, extra, s);
  for (int p = 0; p < P; ++p) grad[p] = 0;
  lsum = 0;
  for (int m = 0; m < md.n_meas; ++m) {
    D sm = s[m];
    if (!finite(sm)) sm = D(S(0));  // utils.set_nan_inf_zero, utils/utils.m:347-350
    const D r = sm - S(y[m]);
    D l = loss_elem<D, S>(loss_kind, r, w ? S(w[m]) : S(1));
    if (codes && loss_kind == GC_LOSS_L1) {
      const S sg = codes[m] == 1 ? S(1) : (codes[m] == 2 ? S(-1) : S(0));
      const S ws = w ? S(w[m]) : S(1);
      for (int p = 0; p < NP; ++p) l.d[p] = r.d[p] * (sg * ws);
    }
    lsum += double(l.v);
    if (elem) elem[m] = l.v;
    for (int p = 0; p < P && p < NP; ++p) grad[p] += l.d[p];
  }
}

// d theta / d z of the transform (for chaining the regulariser gradient)
template <class S>
static inline S dtheta_dz(int transform, S z, S lb, S ub) {
  if (transform == GC_TRANSFORM_SIGMOID) {
    const S sg = S(1) / (S(1) + std::exp(-z));
    return (ub - lb) * sg * (S(1) - sg);
  }
  if (transform == GC_TRANSFORM_LINEAR) return (z >= S(0) && z <= S(1)) ? (ub - lb) : S(0);
  return S(1);
}

// ===========================================================================
// spatial_total_variation, utils/spatial_total_variation.m:24-66.  img is
// column-major [X,Y,Z] with zeros outside the mask (reshape_GD2ND).  Cost:
// sum_i mask_i sum_d |circshift(img,-1,d)_i - img_i| / vsize_d, times
// lambda / Nmasked.  grad (optional) accumulates the subgradient
// (sign(0) = 0) with respect to img.
template <class S>
static double tv_loss_grad(const S* img, const uint8_t* mask, const int32_t dims[3], int mode,
                           const double vsize[3], double lambda, int64_t nmasked, S* grad) {
  const int64_t X = dims[0], Y = dims[1], Z = dims[2];
  const int nd = (mode == GC_TV_3D) ? 3 : 2;
  double acc = 0;
  const double scale = lambda / double(nmasked);
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t yy = 0; yy < Y; ++yy)
      for (int64_t x = 0; x < X; ++x) {
        const int64_t i = x + X * (yy + Y * z);
        if (!mask[i]) continue;
        for (int d = 0; d < nd; ++d) {
          int64_t j;
          if (d == 0) j = ((x + 1) % X) + X * (yy + Y * z);
          else if (d == 1) j = x + X * (((yy + 1) % Y) + Y * z);
          else j = x + X * (yy + Y * ((z + 1) % Z));
          const double g = (double(img[j]) - double(img[i])) / vsize[d];
          acc += std::fabs(g);
          if (grad) {
            const double sg = g > 0 ? 1.0 : (g < 0 ? -1.0 : 0.0);
            grad[j] += S(scale * sg / vsize[d]);
            grad[i] -= S(scale * sg / vsize[d]);
          }
        }
      }
  return acc * scale;
}

// surface_total_variation, NEXI/gpuNEXI.m:1165-1254.  img is the vertex
// image [Ngrid] (reshape_GD2ND: zeros outside the mask; Ngrid = Nvertex *
// Nhemi); nbr is [max_nb][Ngrid] of global 0-based vertex indices
// (neighbours + hemiOffset, :1232-1233) and elen the matching edge lengths.
// dthetadr = (theta - theta(nbr)) ./ edgeLength (:1240), non-finite entries
// zeroed (:1245-1247), summed as |.| over the masked vertices and every
// slot (reshape_ND2GD, :1248) / Nsample * lambda (:1250).  grad (optional)
// accumulates the subgradient (sign(0) = 0) with respect to img.
template <class S>
static double surface_tv_loss_grad(const S* img, const uint8_t* mask, int64_t ngrid, const int32_t* nbr,
                                   const float* elen, int max_nb, double lambda, int64_t nmasked, S* grad) {
  double acc = 0;
  const double scale = lambda / double(nmasked);
  for (int n = 0; n < max_nb; ++n)
    for (int64_t i = 0; i < ngrid; ++i) {
      if (!mask[i]) continue;
      const int64_t j = nbr[size_t(n) * ngrid + i];
      const double e = double(elen[size_t(n) * ngrid + i]);
      const double d = (double(img[i]) - double(img[j])) / e;
      if (!std::isfinite(d)) continue;
      acc += std::fabs(d);
      if (grad) {
        const double sg = d > 0 ? 1.0 : (d < 0 ? -1.0 : 0.0);
        grad[i] += S(scale * sg / e);
        grad[j] -= S(scale * sg / e);
      }
    }
  return acc * scale;
}

// ===========================================================================
// MATLAB median of a vector (mean of the two middle values for even n)
static double median_of(std::vector<double> v) {
  const size_t n = v.size();
  if (n == 0) return std::nan("");
  std::sort(v.begin(), v.end());
  return (n % 2) ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// askadam().optimisation, utils/askadam.m:121-365, with robustConvergence
// (update_outlier_mask :448-510, the frozen convergence set :541-566).
struct AskOut {
  double loss, fid, reg, min_loss, min_fid, min_reg, conv;
  int32_t niter, min_iter, stop;
};

template <class S, int NP>
static int askadam_run(const gc_model& md, const gc_askadam_cfg& cfg, int64_t nv, const float* y,
                       const float* w, const float* extra, const int64_t* vox, const int32_t* dims,
                       const float* prior_grid, const int32_t* surf_nbr, const float* surf_len, int surf_max_nb,
                       const float* z0, double* zfinal, double* zmin, double* res_final,
                       double* res_min, double* trace, AskOut& out, int nthreads,
                       const uint32_t* sign_trace = nullptr, int64_t n_mean = 0) {
  const int P = md.n_par, M = md.n_meas, E = md.n_extra;
  // N of the mean (:96): the caller's count when this is a piece of a larger fit
  const double nmean = n_mean > 0 ? double(n_mean) : double(nv);
  const S inv_count = S(1.0 / (double(M) * nmean));
  const int nw = (M + 15) / 16;
  int sign_row = 0;  // row of sign_trace the next evaluation follows (epoch - 1)
  std::vector<S> z(size_t(P) * nv), g(size_t(P) * nv), m1(size_t(P) * nv, S(0)), m2(size_t(P) * nv, S(0)),
      zprev, zmn;
  for (size_t i = 0; i < z.size(); ++i) z[i] = S(z0[i]);
  std::vector<S> ex(size_t(std::max(E, 1)) * nv, S(0));
  for (int64_t i = 0; i < int64_t(E) * nv; ++i) ex[i] = S(extra[i]);

  // linear mode: set_boundary01 of the starting point (:183-185)
  if (cfg.transform == GC_TRANSFORM_LINEAR)
    for (auto& v : z) v = std::min(std::max(v, S(0)), S(1));

  const bool use_reg = cfg.n_reg > 0 && cfg.lambda[0] > 0;
  const bool use_prior = use_reg && cfg.reg_type == GC_REG_PRIOR;
  const bool use_stv = use_reg && cfg.reg_type == GC_REG_SURFACE_TV;
  const bool use_tv = use_reg && cfg.reg_type == GC_REG_TV;
  if (use_reg && !use_prior && !use_stv && !use_tv) return GC_ERR_ARG;
  if (use_prior && (!prior_grid || !vox || !dims)) return GC_ERR_ARG;
  if (use_stv && (!surf_nbr || !surf_len || surf_max_nb < 1 || !vox || !dims)) return GC_ERR_ARG;
  int64_t nvox = 1;
  std::vector<uint8_t> mask;
  std::vector<S> img, gimg;
  if (use_reg) {
    nvox = int64_t(dims[0]) * dims[1] * dims[2];
    mask.assign(nvox, 0);
    for (int64_t v = 0; v < nv; ++v) mask[vox[v]] = 1;
    img.assign(nvox, S(0));
    gimg.assign(nvox, S(0));
  }

  // robustConvergence: weights are the caller's times a per-voxel modulator
  // on the epoch after each outlier check (update_outlier_mask returns the
  // unmodulated weights on every other epoch)
  const bool robust = cfg.robust != 0;
  std::vector<float> w_orig, w_cur;
  if (robust) {
    w_orig.assign(size_t(M) * nv, 1.0f);
    if (w) for (size_t i = 0; i < w_orig.size(); ++i) w_orig[i] = w[i];
    w_cur = w_orig;
    w = w_cur.data();
  }
  std::vector<S> elem_cur(size_t(M) * nv);
  // model_gradient: returns (loss, fid, reg) and fills g, elem_cur
  auto evaluate = [&](double& loss, double& fid, double& reg) {
    double lsum_total = 0;
#pragma omp parallel for schedule(static) reduction(+ : lsum_total) num_threads(nthreads)
    for (int64_t v = 0; v < nv; ++v) {
      S zv[GC_MAX_PAR], gv[GC_MAX_PAR], exv[64], el[1024];
      float yv[1024], wv[1024];
      for (int p = 0; p < P; ++p) zv[p] = z[size_t(p) * nv + v];
      for (int e = 0; e < E; ++e) exv[e] = ex[size_t(e) * nv + v];
      for (int m = 0; m < M; ++m) {
        yv[m] = y[size_t(m) * nv + v];
        wv[m] = w ? w[size_t(m) * nv + v] : 1.0f;
      }
      double ls;
      uint8_t codes[1024];
      if (sign_trace) {
        const uint32_t* row = sign_trace + size_t(sign_row) * nw * nv;
        for (int m = 0; m < M; ++m) codes[m] = uint8_t((row[size_t(m >> 4) * nv + v] >> ((m & 15) * 2)) & 3u);
      }
      eval_voxel<S, NP>(md, cfg.loss, cfg.transform, zv, yv, w ? wv : nullptr, exv, ls, gv, el,
                        sign_trace ? codes : nullptr);
      lsum_total += ls;
      for (int p = 0; p < P; ++p) g[size_t(p) * nv + v] = gv[p] * inv_count;
      for (int m = 0; m < M; ++m) elem_cur[size_t(m) * nv + v] = el[m];
    }
    fid = lsum_total / (double(M) * nmean);  // mean(residuals), :96
    reg = 0;
    if (use_tv) {
      for (int k = 0; k < cfg.n_reg; ++k) {
        const int p = cfg.regmap[k];
        std::fill(img.begin(), img.end(), S(0));
        std::fill(gimg.begin(), gimg.end(), S(0));
        for (int64_t v = 0; v < nv; ++v)
          img[vox[v]] = unscale<S, S>(cfg.transform, z[size_t(p) * nv + v], S(md.lb[p]), S(md.ub[p]));
        reg += tv_loss_grad<S>(img.data(), mask.data(), dims, cfg.tv_mode, cfg.voxel_size, cfg.lambda[k], int64_t(nmean),
                               gimg.data());
        for (int64_t v = 0; v < nv; ++v)
          g[size_t(p) * nv + v] += gimg[vox[v]] * dtheta_dz<S>(cfg.transform, z[size_t(p) * nv + v],
                                                             S(md.lb[p]), S(md.ub[p]));
      }
    }
    if (use_stv) {
      for (int k = 0; k < cfg.n_reg; ++k) {
        const int p = cfg.regmap[k];
        std::fill(img.begin(), img.end(), S(0));
        std::fill(gimg.begin(), gimg.end(), S(0));
        for (int64_t v = 0; v < nv; ++v)
          img[vox[v]] = unscale<S, S>(cfg.transform, z[size_t(p) * nv + v], S(md.lb[p]), S(md.ub[p]));
        reg += surface_tv_loss_grad<S>(img.data(), mask.data(), nvox, surf_nbr, surf_len, surf_max_nb,
                                       cfg.lambda[k], int64_t(nmean), gimg.data());
        for (int64_t v = 0; v < nv; ++v)
          g[size_t(p) * nv + v] += gimg[vox[v]] * dtheta_dz<S>(cfg.transform, z[size_t(p) * nv + v],
                                                             S(md.lb[p]), S(md.ub[p]));
      }
    }
    // gpuNEXI.prior_distribution_normal (NEXI/gpuNEXI.m:1256-1303): theta as
    // an image (zeros outside the mask, reshape_GD2ND), zscore =
    // (theta - mu) ./ sigma over the whole grid, loss_reg +=
    // sum(abs(zscore(finite))) / Nsample * lambda; the gradient reaches the
    // masked voxels only (sign(0) = 0, as dlgradient of abs)
    if (use_prior) {
      for (int k = 0; k < cfg.n_reg; ++k) {
        const int p = cfg.regmap[k];
        const float* mu = prior_grid + size_t(k) * nvox;                 // mu rows, then
        const float* sd = prior_grid + size_t(cfg.n_reg + k) * nvox;     // sigma rows
        std::fill(img.begin(), img.end(), S(0));
        for (int64_t v = 0; v < nv; ++v)
          img[vox[v]] = unscale<S, S>(cfg.transform, z[size_t(p) * nv + v], S(md.lb[p]), S(md.ub[p]));
        double sum = 0;
        for (int64_t i = 0; i < nvox; ++i) {
          const S zs = (img[i] - S(mu[i])) / S(sd[i]);
          if (std::isfinite(double(zs))) sum += std::fabs(double(zs));
        }
        reg += sum / nmean * cfg.lambda[k];
        const S scale = S(cfg.lambda[k] / nmean);
        for (int64_t v = 0; v < nv; ++v) {
          const int64_t i = vox[v];
          const S zs = (img[i] - S(mu[i])) / S(sd[i]);
          if (!std::isfinite(double(zs))) continue;
          const S sg = zs > S(0) ? S(1) : (zs < S(0) ? S(-1) : S(0));
          g[size_t(p) * nv + v] += scale * sg / S(sd[i]) *
                                   dtheta_dz<S>(cfg.transform, z[size_t(p) * nv + v], S(md.lb[p]), S(md.ub[p]));
        }
      }
    }
    loss = fid + reg;  // :102
  };

  // per-voxel mean elementwise loss, mean(reshape(residuals, Nmeas, Nvol), 1)
  auto per_voxel_loss = [&](std::vector<double>& pvl) {
    pvl.assign(nv, 0.0);
    for (int64_t v = 0; v < nv; ++v) {
      double a = 0;
      for (int m = 0; m < M; ++m) a += double(elem_cur[size_t(m) * nv + v]);
      pvl[v] = a / M;
    }
  };

  // 2.5 starting point evaluation (:197-200)
  double loss, fid, reg;
  evaluate(loss, fid, reg);
  std::vector<double> pvl_init, pvl;
  if (robust) per_voxel_loss(pvl_init);
  // 2.8 robust state (:223-241)
  const int W = std::max(cfg.outlier_check_window, 1);
  std::vector<double> hist;  // [W][nv], oldest first
  std::vector<uint8_t> main_mask(nv, 1), mask_conv(nv, 1);
  std::vector<int> flag_count(nv, 0);
  int reprime = 0;
  if (robust) {
    hist.resize(size_t(W) * nv);
    for (int k = 0; k < W; ++k) std::copy(pvl_init.begin(), pvl_init.end(), hist.begin() + size_t(k) * nv);
  }
  if (trace) trace[0] = loss;
  double min_loss = loss, min_fid = fid, min_reg = reg;
  int min_iter = 0, epoch = 0;
  zmn = z;
  std::vector<S> elem_min = elem_cur;

  // 2.7 convergence state (:212-219)
  double ema = loss;
  std::vector<double> cbuf(std::max(cfg.convergence_window, 1), 1.0);
  int pc = 0, ps = 0, pg = 0;
  double conv = 0;
  zprev = z;
  int stop = GC_STOP_NONE;

  for (epoch = 1; epoch <= cfg.iteration; ++epoch) {
    sign_row = epoch - 1;
    if (epoch > 1) evaluate(loss, fid, reg);  // epoch 1 evaluates the unchanged start
    if (trace) trace[epoch] = loss;
    // 3.2 update_min_loss (:433-446): strict improvement, pre-update params
    if (min_loss > loss) {
      min_loss = loss; min_fid = fid; min_reg = reg; zmn = z; min_iter = epoch; elem_min = elem_cur;
    }
    // 3.3 update_outlier_mask (:448-510)
    double loss_c = loss;
    if (robust) {
      per_voxel_loss(pvl);
      w_cur = w_orig;
      if (cfg.weight_update_interval > 0 && epoch % cfg.weight_update_interval == 0) {
        std::rotate(hist.begin(), hist.begin() + nv, hist.end());
        std::copy(pvl.begin(), pvl.end(), hist.end() - nv);
        std::vector<double> impA(nv), impB(nv);
        for (int64_t v = 0; v < nv; ++v) {
          const double h0 = hist[v], h1 = hist[size_t(W - 1) * nv + v];
          impA[v] = (h0 - h1) / (h0 + 1e-8);
          impB[v] = (pvl_init[v] - pvl[v]) / (pvl_init[v] + 1e-8);
        }
        const double medA = median_of(impA), medB = median_of(impB);
        for (int64_t v = 0; v < nv; ++v) {
          const bool failA = impA[v] < cfg.outlier_voxel_thres && medA > cfg.outlier_pop_thres;
          const bool failB = impB[v] < cfg.outlier_init_thres && medB > cfg.outlier_init_pop_thres;
          if (failA && failB) flag_count[v] += 1;
          else flag_count[v] = std::max(flag_count[v] - 1, 0);
          main_mask[v] = flag_count[v] == 0;
          if (!main_mask[v])
            for (int m = 0; m < M; ++m) w_cur[size_t(m) * nv + v] = float(w_orig[size_t(m) * nv + v] * cfg.outlier_weight);
        }
      }
      // frozen convergence set, refreshed on convergenceWindow boundaries (:541-566)
      if (cfg.convergence_window > 0 && epoch % cfg.convergence_window == 0 && main_mask != mask_conv) {
        mask_conv = main_mask;
        reprime = cfg.convergence_window;
      }
      double a = 0;
      int64_t cnt = 0;
      for (int64_t v = 0; v < nv; ++v)
        if (mask_conv[v]) { a += pvl[v]; ++cnt; }
      loss_c = cnt ? a / double(cnt) : std::nan("");
    }
    // 3.4 update_convergence_signals (:512-632)
    if (cfg.convergence_model == GC_CONV_EMA) {
      // update_convergence_ema, :1192-1204
      const double e_new = cfg.ema_decay * ema + (1 - cfg.ema_decay) * loss_c;
      conv = std::fabs(e_new - ema) / (std::fabs(ema) + 1e-8);
      ema = e_new;
    } else {
      // update_convergence, :1185-1190: minus the least-squares slope
      std::rotate(cbuf.begin(), cbuf.begin() + 1, cbuf.end());
      cbuf.back() = loss_c;
      const int n = int(cbuf.size());
      double sx = 0, sy = 0, sxx = 0, sxy = 0;
      for (int i = 0; i < n; ++i) {
        const double xi = i + 1;
        sx += xi; sy += cbuf[i]; sxx += xi * xi; sxy += xi * cbuf[i];
      }
      const double den = n * sxx - sx * sx;
      const double slope = (den != 0) ? (n * sxy - sx * sy) / den : 0.0;
      conv = -slope;
    }
    if (reprime > 0) { --reprime; pc = 0; }
    else if (conv > cfg.convergence_value || epoch <= cfg.convergence_window) pc = 0;
    else if (!robust && (min_loss - loss) > cfg.convergence_value) pc = 0;
    else pc += 1;
    if (cfg.step_tol > 0) {
      double num = 0, den = 0;
      for (size_t i = 0; i < z.size(); ++i) {
        const double dz = double(z[i]) - double(zprev[i]);
        num += dz * dz;
        den += double(z[i]) * double(z[i]);
      }
      const double sn = std::sqrt(num) / (1 + std::sqrt(den));
      zprev = z;
      ps = (sn < cfg.step_tol) ? ps + 1 : 0;
    }
    if (cfg.grad_tol > 0) {
      double gn = 0;
      for (size_t i = 0; i < g.size(); ++i) gn += double(g[i]) * double(g[i]);
      gn = std::sqrt(gn);
      pg = (gn < cfg.grad_tol) ? pg + 1 : 0;
    }
    // 3.5 check_stopping (:634-671), in order
    if (pc > cfg.patience_conv) stop = GC_STOP_CONVERGENCE;
    else if (loss < cfg.tol) stop = GC_STOP_TOL;
    else if (cfg.step_tol > 0 && ps > cfg.patience_step) stop = GC_STOP_STEP;
    else if (cfg.grad_tol > 0 && pg > cfg.patience_grad) stop = GC_STOP_GRAD;
    if (stop != GC_STOP_NONE) break;
    // 3.6 update_parameters (:673-691), not on the last iteration
    if (epoch < cfg.iteration) {
      const double lr = cfg.lr0 / (1 + cfg.decay * (epoch - 1));  // update_learn_rate, :1180-1182
      if (cfg.optimiser == GC_OPT_ADAM) {
        // adamupdate(params, grad, avg, avgsq, t=epoch, lr, b1, b2, eps)
        const double bc1 = 1 - std::pow(cfg.beta1, epoch), bc2 = 1 - std::pow(cfg.beta2, epoch);
        const S b1 = S(cfg.beta1), b2 = S(cfg.beta2), eps = S(cfg.eps);
        const S lr_eff = S(lr * std::sqrt(bc2) / bc1);
        for (size_t i = 0; i < z.size(); ++i) {
          m1[i] = b1 * m1[i] + (S(1) - b1) * g[i];
          m2[i] = b2 * m2[i] + (S(1) - b2) * g[i] * g[i];
          if (cfg.adam_form == GC_ADAM_KINGMA_BA)
            z[i] -= S(lr) * (m1[i] / S(bc1)) / (std::sqrt(m2[i] / S(bc2)) + eps);
          else
            z[i] -= lr_eff * m1[i] / (std::sqrt(m2[i]) + eps);
        }
      } else if (cfg.optimiser == GC_OPT_SGDM) {
        // sgdmupdate: vel = momentum*vel - lr*grad; params = params + vel
        for (size_t i = 0; i < z.size(); ++i) {
          m1[i] = S(cfg.momentum) * m1[i] - S(lr) * g[i];
          z[i] += m1[i];
        }
      } else {
        // rmspropupdate: avgsq = d*avgsq + (1-d)*g^2; params -= lr*g/(sqrt(avgsq)+eps)
        for (size_t i = 0; i < z.size(); ++i) {
          m2[i] = S(cfg.rms_decay) * m2[i] + (S(1) - S(cfg.rms_decay)) * g[i] * g[i];
          z[i] -= S(lr) * g[i] / (std::sqrt(m2[i]) + S(cfg.rms_eps));
        }
      }
    }
  }
  if (epoch > cfg.iteration) epoch = cfg.iteration;
  // 4. finalisation (:312-362): linear mode clamps both results to [0,1]
  if (cfg.transform == GC_TRANSFORM_LINEAR) {
    for (auto& v : z) v = std::min(std::max(v, S(0)), S(1));
    for (auto& v : zmn) v = std::min(std::max(v, S(0)), S(1));
  }
  for (size_t i = 0; i < z.size(); ++i) { zfinal[i] = double(z[i]); zmin[i] = double(zmn[i]); }
  if (res_final) for (size_t i = 0; i < elem_cur.size(); ++i) res_final[i] = double(elem_cur[i]);
  if (res_min) for (size_t i = 0; i < elem_min.size(); ++i) res_min[i] = double(elem_min[i]);
  out.loss = loss; out.fid = fid; out.reg = reg;
  out.min_loss = min_loss; out.min_fid = min_fid; out.min_reg = min_reg;
  out.conv = conv; out.niter = epoch; out.min_iter = min_iter; out.stop = stop;
  return 0;
}

// ===========================================================================
// MCMC: utils/mcmc.m:82-190 (MH) and :192-383 (GW, simultaneous and
// red-black).  Chains of different repetitions are independent here; the
// reference runs them one after another from the same start (:136, :281).
struct McmcArgs {
  const gc_model* md;
  const gc_mcmc_cfg* cfg;
  int64_t nv;
  const float *y, *w, *extra, *x0, *normals, *uniforms;
  const int32_t* partners;
  double *posterior, *mean, *stdv, *accept, *xfinal, *margin;
  uint8_t* trace;
  // goodman_weare_wglobal_constant only
  const float* g_uniforms;
  const int32_t* g_partners;
  double *g_posterior, *g_xfinal, *g_margin;
  uint8_t* g_trace;
};

template <class S>
static S sos_of(const gc_model& md, const S* x, const float* y, const float* w, const S* ex, int64_t nv, int64_t v) {
  S s[1024];
  fwd_model<S, S>(md, x, ex, s);
  S acc = 0;
  for (int m = 0; m < md.n_meas; ++m) {
    S sm = s[m];
    const S r = sm - S(y[size_t(m) * nv + v]);
    acc += (w ? S(w[size_t(m) * nv + v]) : S(1)) * r * r;
  }
  return acc;
}

static inline int mcmc_ns(const gc_mcmc_cfg& c) {
  // Ns = numel(Nburnin+1:thinning:iteration), utils/mcmc.m:110,235
  if (c.iteration < c.n_burnin + 1) return 0;
  return (c.iteration - c.n_burnin - 1) / c.thinning + 1;
}

static inline bool keep_sample(const gc_mcmc_cfg& c, int k) {
  // (k > Nburnin) && mod(k-Nburnin+1, thinning) == 0, utils/mcmc.m:170,363
  return k > c.n_burnin && ((k - c.n_burnin + 1) % c.thinning) == 0;
}

template <class S>
static void mh_run(const McmcArgs& a, int nthreads) {
  const gc_model& md = *a.md;
  const gc_mcmc_cfg& c = *a.cfg;
  const int Q = c.n_var, M = md.n_meas, E = md.n_extra, R = c.repetition;
  const int64_t nv = a.nv;
  const int Ns = mcmc_ns(c);
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t v = 0; v < nv; ++v) {
    S ex[64], x0[GC_MAX_PAR];
    for (int e = 0; e < E; ++e) ex[e] = S(a.extra[size_t(e) * nv + v]);
    for (int q = 0; q < Q; ++q) {
      // xCurr = max(xCurr,lb); min(.,ub) (:126-128)
      x0[q] = std::min(std::max(S(a.x0[size_t(q) * nv + v]), S(md.lb[q])), S(md.ub[q]));
    }
    const S lp0 = logp_gaussian<S>(sos_of<S>(md, x0, a.y, a.w, ex, nv, v), x0[c.noise_slot], M);  // :129
    double s1[GC_MAX_PAR] = {0}, s2[GC_MAX_PAR] = {0};
    int64_t nacc = 0;
    for (int r = 0; r < R; ++r) {
      S x[GC_MAX_PAR], xp[GC_MAX_PAR];
      for (int q = 0; q < Q; ++q) x[q] = x0[q];
      S lp = lp0;
      int counter = 0;
      const uint64_t chain = uint64_t(r) * uint64_t(nv) + uint64_t(v);
      for (int k = 1; k <= c.iteration; ++k) {
        const size_t step = size_t(r) * c.iteration + (k - 1);
        bool oob = false;
        double nat[GC_MAX_PAR], u_nat = 0.0;
        if (c.rng_mode != GC_RNG_REPLAY) native_mh_draws(c.seed, uint32_t(k), chain, Q, nat, &u_nat);
        for (int q = 0; q < Q; ++q) {
          const double n = (c.rng_mode == GC_RNG_REPLAY) ? double(a.normals[(step * nv + v) * Q + q]) : nat[q];
          // xProposed = xCurr + xStepsize.*randn (:147)
          xp[q] = x[q] + S(c.x_step[q]) * S(n);
          oob = oob || (xp[q] < S(md.lb[q])) || (xp[q] > S(md.ub[q]));  // :149
          xp[q] = std::min(std::max(xp[q], S(md.lb[q])), S(md.ub[q]));  // :151
        }
        const S lpp = logp_gaussian<S>(sos_of<S>(md, xp, a.y, a.w, ex, nv, v), xp[c.noise_slot], M);  // :158
        const double u = (c.rng_mode == GC_RNG_REPLAY) ? double(a.uniforms[step * nv + v]) : u_nat;
        const S ratio = min1<S>(S(std::exp(lpp - lp)));  // :160
        const bool acc = (ratio > S(u)) && !oob;              // :161-162
        if (a.trace) a.trace[step * nv + v] = acc ? 1 : 0;
        if (a.margin) a.margin[step * nv + v] = oob ? -INFINITY : double(lpp - lp) - std::log(u);
        if (acc) {
          for (int q = 0; q < Q; ++q) x[q] = xp[q];
          lp = lpp;
          ++nacc;
        }
        if (keep_sample(c, k)) {
          if (counter < Ns) {
            for (int q = 0; q < Q; ++q) {
              if (a.posterior)
                a.posterior[((size_t(q) * R + r) * Ns + counter) * nv + v] = double(x[q]);
              const double d = double(x[q]) - double(x0[q]);
              s1[q] += d; s2[q] += d * d;
            }
          }
          ++counter;
        }
      }
      const int filled = std::min(counter, Ns);
      for (int q = 0; q < Q; ++q) {
        // unfilled posterior slots stay zero (utils/mcmc.m:121) and enter res2out
        const double d = -double(x0[q]);
        s1[q] += d * (Ns - filled); s2[q] += d * d * (Ns - filled);
        if (a.xfinal) a.xfinal[(size_t(r) * Q + q) * nv + v] = double(x[q]);
      }
    }
    const double n = double(Ns) * R;
    for (int q = 0; q < Q; ++q) {
      if (a.mean) a.mean[size_t(q) * nv + v] = double(x0[q]) + s1[q] / n;
      if (a.stdv) a.stdv[size_t(q) * nv + v] = (n > 1) ? std::sqrt(std::max(0.0, (s2[q] - s1[q] * s1[q] / n) / (n - 1))) : 0.0;
    }
    if (a.accept) a.accept[v] = double(nacc) / (double(c.iteration) * R);
  }
}

template <class S>
static void gw_run(const McmcArgs& a, int nthreads) {
  const gc_model& md = *a.md;
  const gc_mcmc_cfg& c = *a.cfg;
  const int Q = c.n_var, M = md.n_meas, E = md.n_extra, R = c.repetition, W = c.n_walker;
  const int64_t nv = a.nv;
  const int Ns = mcmc_ns(c);
  const S aa = S(c.step_size);
  const bool rb = c.ensemble_update == GC_ENSEMBLE_REDBLACK;
  const int h = W / 2;
  // partner tables for the native stream (shared by all voxels)
  std::vector<int32_t> ptab;
  const int ptab_w = W;
  if (c.rng_mode == GC_RNG_NATIVE) {
    ptab.resize(size_t(R) * c.iteration * ptab_w);
    for (int r = 0; r < R; ++r)
      for (int k = 1; k <= c.iteration; ++k)
        native_partners(c.seed, r, k, W, c.ensemble_update, &ptab[(size_t(r) * c.iteration + (k - 1)) * ptab_w]);
  }
  const int32_t* partners = (c.rng_mode == GC_RNG_NATIVE) ? ptab.data() : a.partners;

#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t v = 0; v < nv; ++v) {
    S ex[64], xc0[GC_MAX_PAR];
    for (int e = 0; e < E; ++e) ex[e] = S(a.extra[size_t(e) * nv + v]);
    for (int q = 0; q < Q; ++q) xc0[q] = S(a.x0[size_t(q) * nv + v]);
    // walker initialisation: x0 + (ub-lb)*startRange.*randn, clamped (:270-273)
    std::vector<S> xinit(size_t(W) * Q), lpinit(W);
    for (int wk = 0; wk < W; ++wk) {
      for (int q = 0; q < Q; ++q) {
        double n;
        if (c.rng_mode == GC_RNG_REPLAY) n = a.normals[(size_t(wk) * nv + v) * Q + q];
        else n = native_normal(c.seed, 0u, uint64_t(wk) * uint64_t(nv) + uint64_t(v), PURPOSE_INIT, q);
        S xi = xc0[q] + S(md.ub[q] - md.lb[q]) * S(c.start_range) * S(n);
        xi = std::min(std::max(xi, S(md.lb[q])), S(md.ub[q]));
        xinit[size_t(wk) * Q + q] = xi;
      }
      lpinit[wk] = logp_gaussian<S>(sos_of<S>(md, &xinit[size_t(wk) * Q], a.y, a.w, ex, nv, v),
                                   xinit[size_t(wk) * Q + c.noise_slot], M);  // :275
    }
    double s1[GC_MAX_PAR] = {0}, s2[GC_MAX_PAR] = {0};
    int64_t nacc = 0;
    std::vector<S> x(size_t(W) * Q), lp(W), xp(size_t(W) * Q), lpp(W), zz(W);
    std::vector<uint8_t> oob(W);
    for (int r = 0; r < R; ++r) {
      x = xinit;
      lp = lpinit;
      int counter = 0;
      for (int k = 1; k <= c.iteration; ++k) {
        const size_t step = size_t(r) * c.iteration + (k - 1);
        const int32_t* part = partners + step * W;
        const int nsub = rb ? 2 : 1;
        for (int sub = 0; sub < nsub; ++sub) {
          int a0 = 0, na = W;
          if (rb) { a0 = (sub == 0) ? 0 : h; na = (sub == 0) ? h : W - h; }
          // uniforms for this (sub-)step: [2][na][nv] at offset
          const size_t ubase = rb ? (step * 2 + sub) * 2 * size_t(h > W - h ? h : W - h) : step * 2 * size_t(W);
          const size_t ustride = rb ? size_t(h > W - h ? h : W - h) : size_t(W);
          for (int j = 0; j < na; ++j) {
            const int wk = a0 + j;
            const int pw = rb ? part[(sub == 0 ? 0 : h) + j] : part[wk];
            double u1;
            const uint64_t chain = (uint64_t(r) * W + wk) * uint64_t(nv) + uint64_t(v);
            // native: one draw gives the stretch (o[0]) and acceptance (o[1]) uniforms
            if (c.rng_mode == GC_RNG_REPLAY) u1 = a.uniforms[(ubase + j) * nv + v];
            else { double ua; native_gw_uniforms(c.seed, uint32_t(k), chain, sub, &u1, &ua); }
            // zz = ((a-1)*rand + 1).^2 / a (:299, :332)
            const S t = (aa - S(1)) * S(u1) + S(1);
            zz[wk] = t * t / aa;
            bool ob = false;
            for (int q = 0; q < Q; ++q) {
              const S xa = x[size_t(pw) * Q + q];
              S val = xa + (x[size_t(wk) * Q + q] - xa) * zz[wk];  // :300, :333
              ob = ob || val < S(md.lb[q]) || val > S(md.ub[q]);
              val = std::min(std::max(val, S(md.lb[q])), S(md.ub[q]));
              xp[size_t(wk) * Q + q] = val;
            }
            oob[wk] = ob;
            lpp[wk] = logp_gaussian<S>(sos_of<S>(md, &xp[size_t(wk) * Q], a.y, a.w, ex, nv, v),
                                       xp[size_t(wk) * Q + c.noise_slot], M);
          }
          for (int j = 0; j < na; ++j) {
            const int wk = a0 + j;
            const uint64_t chain = (uint64_t(r) * W + wk) * uint64_t(nv) + uint64_t(v);
            double u2;
            if (c.rng_mode == GC_RNG_REPLAY) u2 = a.uniforms[(ubase + ustride + j) * nv + v];
            else { double uz; native_gw_uniforms(c.seed, uint32_t(k), chain, sub, &uz, &u2); }
            // min(zz.^(Nvar-1).*exp(logPProposed-logPCurr),1) > rand (:313-315)
            const S ratio = min1<S>(S(std::pow(zz[wk], S(Q - 1)) * std::exp(lpp[wk] - lp[wk])));
            const bool acc = ratio > S(u2) && !oob[wk];
            if (a.trace) a.trace[(step * W + wk) * nv + v] = acc ? 1 : 0;
            if (a.margin)
              a.margin[(step * W + wk) * nv + v] =
                  oob[wk] ? -INFINITY : double(Q - 1) * std::log(double(zz[wk])) + double(lpp[wk] - lp[wk]) - std::log(u2);
            if (acc) {
              for (int q = 0; q < Q; ++q) x[size_t(wk) * Q + q] = xp[size_t(wk) * Q + q];
              lp[wk] = lpp[wk];
              ++nacc;
            }
          }
        }
        if (keep_sample(c, k)) {
          if (counter < Ns) {
            for (int wk = 0; wk < W; ++wk)
              for (int q = 0; q < Q; ++q) {
                const double xv = double(x[size_t(wk) * Q + q]);
                if (a.posterior)
                  a.posterior[(((size_t(q) * R + r) * Ns + counter) * W + wk) * nv + v] = xv;
                const double d = xv - double(xc0[q]);
                s1[q] += d; s2[q] += d * d;
              }
          }
          ++counter;
        }
      }
      const int filled = std::min(counter, Ns);
      for (int q = 0; q < Q; ++q) {
        const double d = -double(xc0[q]);
        s1[q] += d * double(Ns - filled) * W;
        s2[q] += d * d * double(Ns - filled) * W;
      }
      if (a.xfinal)
        for (int wk = 0; wk < W; ++wk)
          for (int q = 0; q < Q; ++q)
            a.xfinal[((size_t(r) * W + wk) * Q + q) * nv + v] = double(x[size_t(wk) * Q + q]);
    }
    const double n = double(Ns) * W * R;
    for (int q = 0; q < Q; ++q) {
      if (a.mean) a.mean[size_t(q) * nv + v] = double(xc0[q]) + s1[q] / n;
      if (a.stdv) a.stdv[size_t(q) * nv + v] = (n > 1) ? std::sqrt(std::max(0.0, (s2[q] - s1[q] * s1[q] / n) / (n - 1))) : 0.0;
    }
    if (a.accept) a.accept[v] = double(nacc) / (double(c.iteration) * R * W);
  }
}

// Goodman-Weare with global parameters: utils/mcmc.m:385-578
// (goodman_weare_wglobal_constant) and :581-626 (gw_global_block).  Walker
// wk of repetition r carries NND local values per voxel and NG global values
// shared by every voxel.  Each iteration runs a local GW block with the
// globals fixed (exponent z^(NND-1)), then a global GW block with the locals
// fixed, accepted on the FP64 sum over voxels of logP(proposal) - logPCurr
// (exponent z^(NG-1)).  Repetitions restart from the same initial ensemble
// and its logP0 (:497-501).
template <class S>
static void gwg_run(const McmcArgs& a, int nthreads) {
  const gc_model& md = *a.md;
  const gc_mcmc_cfg& c = *a.cfg;
  const int Q = c.n_var, M = md.n_meas, E = md.n_extra, R = c.repetition, W = c.n_walker;
  const int64_t nv = a.nv;
  const int Ns = mcmc_ns(c);
  const S aa = S(c.step_size);
  const bool rb = c.ensemble_update == GC_ENSEMBLE_REDBLACK;
  const int h = W / 2, hm = W - W / 2;
  const int T = c.global_warmup + c.iteration;
  std::vector<int> lq, gq;  // model index of each local / global variable
  for (int q = 0; q < Q; ++q) (((c.global_mask >> q) & 1u) ? gq : lq).push_back(q);
  const int NND = int(lq.size()), NG = int(gq.size());
  const bool native = c.rng_mode == GC_RNG_NATIVE;

  // partner tables (shared by all voxels): local block per iteration,
  // global block per global step
  std::vector<int32_t> ltab, gtab;
  if (native) {
    ltab.resize(size_t(R) * c.iteration * W);
    gtab.resize(size_t(R) * T * W);
    for (int r = 0; r < R; ++r) {
      for (int k = 1; k <= c.iteration; ++k)
        native_partners(c.seed, r, k, W, c.ensemble_update, &ltab[(size_t(r) * c.iteration + (k - 1)) * W], kTagLocal);
      for (int t = 1; t <= T; ++t)
        native_partners(c.seed, r, t, W, c.ensemble_update, &gtab[(size_t(r) * T + (t - 1)) * W], kTagGlobal);
    }
  }
  const int32_t* lpart = native ? ltab.data() : a.partners;
  const int32_t* gpart = native ? gtab.data() : a.g_partners;

  std::vector<S> ex(size_t(nv) * (E > 0 ? E : 1));
  for (int64_t v = 0; v < nv; ++v)
    for (int e = 0; e < E; ++e) ex[size_t(v) * E + e] = S(a.extra[size_t(e) * nv + v]);
  // full parameter vector of walker wk at voxel v from its local and global parts
  auto assemble = [&](const S* xl, const S* g, S* x) {
    for (int j = 0; j < NND; ++j) x[lq[j]] = xl[j];
    for (int i = 0; i < NG; ++i) x[gq[i]] = g[i];
  };
  auto logp_at = [&](const S* xl, const S* g, int64_t v) {
    S x[GC_MAX_PAR];
    assemble(xl, g, x);
    return logp_gaussian<S>(sos_of<S>(md, x, a.y, a.w, E > 0 ? &ex[size_t(v) * E] : nullptr, nv, v),
                            x[c.noise_slot], M);
  };

  // ---- initial ensemble (:484-492) ----------------------------------------
  // local: x0 + (ub-lb)*startRange.*randn([NND,Nv,W]), clamped
  std::vector<S> xl0(size_t(W) * nv * NND), g0(size_t(W) * NG), lp0(size_t(W) * nv);
  for (int wk = 0; wk < W; ++wk)
    for (int64_t v = 0; v < nv; ++v)
      for (int j = 0; j < NND; ++j) {
        const int q = lq[j];
        double n;
        if (!native) n = a.normals[(size_t(wk) * nv + v) * NND + j];
        else n = native_normal(c.seed, 0u, uint64_t(wk) * uint64_t(nv) + uint64_t(v), PURPOSE_INIT, j);
        S xi = S(a.x0[size_t(q) * nv + v]) + S(md.ub[q] - md.lb[q]) * S(c.start_range) * S(n);
        xl0[(size_t(wk) * nv + v) * NND + j] = std::min(std::max(xi, S(md.lb[q])), S(md.ub[q]));
      }
  // global: over-dispersed, lb + (ub-lb).*rand([NG,1,W]) (:489)
  for (int wk = 0; wk < W; ++wk)
    for (int i = 0; i < NG; ++i) {
      const int q = gq[i];
      double u;
      if (!native) u = a.g_uniforms[size_t(wk) * NG + i];
      else { uint32_t o[4]; draw(c.seed, 0u, global_init_chain(wk), PURPOSE_INIT + uint32_t(i / 4), o); u = u01(o[i % 4]); }
      g0[size_t(wk) * NG + i] = S(md.lb[q]) + S(md.ub[q] - md.lb[q]) * S(u);
    }
  // logP0 at (local, global) of every walker (:492)
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t v = 0; v < nv; ++v)
    for (int wk = 0; wk < W; ++wk)
      lp0[size_t(wk) * nv + v] = logp_at(&xl0[(size_t(wk) * nv + v) * NND], &g0[size_t(wk) * NG], v);

  std::vector<double> s1(size_t(NND) * nv, 0.0), s2(size_t(NND) * nv, 0.0), gs1(NG, 0.0), gs2(NG, 0.0);
  std::vector<int64_t> nacc(nv, 0);
  std::vector<S> xl, g, lp, xp(size_t(W) * nv * NND), lpp(size_t(W) * nv), zz(size_t(W) * nv);
  std::vector<S> gp(size_t(W) * NG), lpg(size_t(W) * nv), zzg(W);
  std::vector<uint8_t> oob(size_t(W) * nv), oobg(W);
  std::vector<double> dlogp(W);
  const size_t ubase_n = rb ? 2 * size_t(hm) : 2 * size_t(W);  // uniform rows per (sub-)step

  for (int r = 0; r < R; ++r) {
    xl = xl0;
    g = g0;
    lp = lp0;

    // ---- global block, gw_global_block (:581-626) --------------------------
    auto global_block = [&](int t) {
      const size_t step = size_t(r) * T + (t - 1);
      const int32_t* part = gpart + step * W;
      const int nsub = rb ? 2 : 1;
      for (int sub = 0; sub < nsub; ++sub) {
        int a0 = 0, na = W, pofs = 0;
        if (rb) { a0 = (sub == 0) ? 0 : h; na = (sub == 0) ? h : W - h; pofs = (sub == 0) ? 0 : h; }
        // draws: zz_g then acceptance, per active walker
        std::vector<double> uacc(na);
        for (int j = 0; j < na; ++j) {
          const int wk = a0 + j;
          double u1, u2;
          if (!native) {
            const size_t b = size_t(W) * NG + (rb ? (step * 2 + sub) * 2 * size_t(hm) : step * 2 * size_t(W));
            const size_t st = rb ? size_t(hm) : size_t(W);
            u1 = a.g_uniforms[b + j];
            u2 = a.g_uniforms[b + st + j];
          } else {
            uint32_t o[4];
            draw(c.seed, uint32_t(t), global_chain(r, wk), PURPOSE_ZZ + sub, o);
            u1 = u01(o[0]);
            u2 = u01(o[1]);
          }
          uacc[j] = u2;
          const S tt = (aa - S(1)) * S(u1) + S(1);
          zzg[wk] = tt * tt / aa;  // :587
          const int pw = rb ? part[pofs + j] : part[wk];
          bool ob = false;
          for (int i = 0; i < NG; ++i) {
            const int q = gq[i];
            const S xa = g[size_t(pw) * NG + i];
            S val = xa + (g[size_t(wk) * NG + i] - xa) * zzg[wk];  // :588
            ob = ob || val < S(md.lb[q]) || val > S(md.ub[q]);      // :589
            gp[size_t(wk) * NG + i] = std::min(std::max(val, S(md.lb[q])), S(md.ub[q]));
          }
          oobg[wk] = ob;
        }
        // logP at (current local, proposed global) of every voxel (:592)
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int64_t v = 0; v < nv; ++v)
          for (int j = 0; j < na; ++j) {
            const int wk = a0 + j;
            lpg[size_t(wk) * nv + v] = logp_at(&xl[(size_t(wk) * nv + v) * NND], &gp[size_t(wk) * NG], v);
          }
        for (int j = 0; j < na; ++j) {
          const int wk = a0 + j;
          // dlogP = sum(double(logP_g - logPCurr), 2), in voxel order (:594)
          double d = 0.0;
          for (int64_t v = 0; v < nv; ++v) d += double(S(lpg[size_t(wk) * nv + v] - lp[size_t(wk) * nv + v]));
          dlogp[wk] = d;
          S zpow = S(1);
          for (int i = 0; i < NG - 1; ++i) zpow *= zzg[wk];
          const S ratio = min1<S>(zpow * S(std::exp(d)));  // :595
          const bool acc = ratio > S(uacc[j]) && !oobg[wk];  // :596-597
          if (a.g_trace) a.g_trace[step * W + wk] = acc ? 1 : 0;
          if (a.g_margin)
            a.g_margin[step * W + wk] = oobg[wk] ? -INFINITY : double(NG - 1) * std::log(double(zzg[wk])) + d - std::log(uacc[j]);
          if (acc) {  // :598-599
            for (int i = 0; i < NG; ++i) g[size_t(wk) * NG + i] = gp[size_t(wk) * NG + i];
            for (int64_t v = 0; v < nv; ++v) lp[size_t(wk) * nv + v] = lpg[size_t(wk) * nv + v];
          }
        }
      }
    };

    // optional global-only warm-up, locals frozen (:503-507)
    for (int t = 1; t <= c.global_warmup; ++t) global_block(t);

    int counter = 0;
    for (int k = 1; k <= c.iteration; ++k) {
      // ---- local block, globals fixed (:512-545) ---------------------------
      const size_t step = size_t(r) * c.iteration + (k - 1);
      const int32_t* part = lpart + step * W;
      const int nsub = rb ? 2 : 1;
      for (int sub = 0; sub < nsub; ++sub) {
        int a0 = 0, na = W, pofs = 0;
        if (rb) { a0 = (sub == 0) ? 0 : h; na = (sub == 0) ? h : W - h; pofs = (sub == 0) ? 0 : h; }
        const size_t ub0 = rb ? (step * 2 + sub) * ubase_n : step * ubase_n;
        const size_t ust = rb ? size_t(hm) : size_t(W);
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int64_t v = 0; v < nv; ++v) {
          for (int j = 0; j < na; ++j) {
            const int wk = a0 + j;
            const int pw = rb ? part[pofs + j] : part[wk];
            double u1;
            if (!native) u1 = a.uniforms[(ub0 + j) * nv + v];
            else {
              uint32_t o[4];
              draw(c.seed, uint32_t(k), (uint64_t(r) * W + wk) * uint64_t(nv) + uint64_t(v), PURPOSE_ZZ + sub, o);
              u1 = u01(o[0]);
            }
            const S tt = (aa - S(1)) * S(u1) + S(1);
            const size_t wv = size_t(wk) * nv + v;
            zz[wv] = tt * tt / aa;  // :515
            bool ob = false;
            for (int jj = 0; jj < NND; ++jj) {
              const int q = lq[jj];
              const S xa = xl[(size_t(pw) * nv + v) * NND + jj];
              S val = xa + (xl[wv * NND + jj] - xa) * zz[wv];  // :516
              ob = ob || val < S(md.lb[q]) || val > S(md.ub[q]);  // :517
              xp[wv * NND + jj] = std::min(std::max(val, S(md.lb[q])), S(md.ub[q]));
            }
            oob[wv] = ob;
            lpp[wv] = logp_at(&xp[wv * NND], &g[size_t(wk) * NG], v);  // (new local, old global), :519-520
          }
          for (int j = 0; j < na; ++j) {
            const int wk = a0 + j;
            const size_t wv = size_t(wk) * nv + v;
            double u2;
            if (!native) u2 = a.uniforms[(ub0 + ust + j) * nv + v];
            else {
              uint32_t o[4];
              draw(c.seed, uint32_t(k), (uint64_t(r) * W + wk) * uint64_t(nv) + uint64_t(v), PURPOSE_ZZ + sub, o);
              u2 = u01(o[1]);
            }
            S zpow = S(1);
            for (int i = 0; i < NND - 1; ++i) zpow *= zz[wv];
            const S ratio = min1<S>(zpow * std::exp(lpp[wv] - lp[wv]));  // :521
            const bool acc = ratio > S(u2) && !oob[wv];                  // :522-523
            if (a.trace) a.trace[(step * W + wk) * nv + v] = acc ? 1 : 0;
            if (a.margin)
              a.margin[(step * W + wk) * nv + v] =
                  oob[wv] ? -INFINITY : double(NND - 1) * std::log(double(zz[wv])) + double(lpp[wv] - lp[wv]) - std::log(u2);
            if (acc) {  // :524-525
              for (int jj = 0; jj < NND; ++jj) xl[wv * NND + jj] = xp[wv * NND + jj];
              lp[wv] = lpp[wv];
              ++nacc[v];
            }
          }
        }
      }
      // ---- global block (:549-551) ------------------------------------------
      global_block(c.global_warmup + k);

      // thinning / burn-in (:554-558)
      if (keep_sample(c, k)) {
        if (counter < Ns) {
          for (int wk = 0; wk < W; ++wk) {
            for (int64_t v = 0; v < nv; ++v)
              for (int j = 0; j < NND; ++j) {
                const double xv = double(xl[(size_t(wk) * nv + v) * NND + j]);
                if (a.posterior) a.posterior[(((size_t(j) * R + r) * Ns + counter) * W + wk) * nv + v] = xv;
                const double d = xv - double(a.x0[size_t(lq[j]) * nv + v]);
                s1[size_t(j) * nv + v] += d;
                s2[size_t(j) * nv + v] += d * d;
              }
            for (int i = 0; i < NG; ++i) {
              const double gv = double(g[size_t(wk) * NG + i]);
              if (a.g_posterior) a.g_posterior[((size_t(i) * R + r) * Ns + counter) * W + wk] = gv;
              const double d = gv - 0.5 * (md.lb[gq[i]] + md.ub[gq[i]]);
              gs1[i] += d;
              gs2[i] += d * d;
            }
          }
        }
        ++counter;
      }
    }
    // unfilled posterior slots are zeros (:481-482) and enter res2out
    const int unfilled = Ns - std::min(counter, Ns);
    for (int j = 0; j < NND; ++j)
      for (int64_t v = 0; v < nv; ++v) {
        const double d = -double(a.x0[size_t(lq[j]) * nv + v]);
        s1[size_t(j) * nv + v] += d * double(unfilled) * W;
        s2[size_t(j) * nv + v] += d * d * double(unfilled) * W;
      }
    for (int i = 0; i < NG; ++i) {
      const double d = -0.5 * (md.lb[gq[i]] + md.ub[gq[i]]);
      gs1[i] += d * double(unfilled) * W;
      gs2[i] += d * d * double(unfilled) * W;
    }
    if (a.xfinal)
      for (int wk = 0; wk < W; ++wk)
        for (int j = 0; j < NND; ++j)
          for (int64_t v = 0; v < nv; ++v)
            a.xfinal[((size_t(r) * W + wk) * NND + j) * nv + v] = double(xl[(size_t(wk) * nv + v) * NND + j]);
    if (a.g_xfinal)
      for (int wk = 0; wk < W; ++wk)
        for (int i = 0; i < NG; ++i) a.g_xfinal[(size_t(r) * W + wk) * NG + i] = double(g[size_t(wk) * NG + i]);
  }
  const double n = double(Ns) * W * R;
  auto stdv = [&](double a1, double a2) {
    return (n > 1) ? std::sqrt(std::max(0.0, (a2 - a1 * a1 / n) / (n - 1))) : 0.0;
  };
  for (int j = 0; j < NND; ++j)
    for (int64_t v = 0; v < nv; ++v) {
      const size_t qv = size_t(lq[j]) * nv + v;
      if (a.mean) a.mean[qv] = double(a.x0[qv]) + s1[size_t(j) * nv + v] / n;
      if (a.stdv) a.stdv[qv] = stdv(s1[size_t(j) * nv + v], s2[size_t(j) * nv + v]);
    }
  for (int i = 0; i < NG; ++i)
    for (int64_t v = 0; v < nv; ++v) {
      const size_t qv = size_t(gq[i]) * nv + v;
      if (a.mean) a.mean[qv] = 0.5 * (md.lb[gq[i]] + md.ub[gq[i]]) + gs1[i] / n;
      if (a.stdv) a.stdv[qv] = stdv(gs1[i], gs2[i]);
    }
  if (a.accept)
    for (int64_t v = 0; v < nv; ++v) a.accept[v] = double(nacc[v]) / (double(c.iteration) * R * W);
}

template <class F>
static int with_np(int P, F&& f) {
  if (P <= 2) return f(std::integral_constant<int, 2>{});
  if (P <= 3) return f(std::integral_constant<int, 3>{});
  if (P <= 4) return f(std::integral_constant<int, 4>{});
  if (P <= 5) return f(std::integral_constant<int, 5>{});
  if (P <= 8) return f(std::integral_constant<int, 8>{});
  if (P <= 16) return f(std::integral_constant<int, 16>{});
  return f(std::integral_constant<int, GC_MAX_PAR>{});
}

}  // namespace orc

// ===========================================================================
// C API of the oracle (called through ctypes by tests/ and bench.py's CPU
// baseline).
extern "C" {

int orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  orc::philox4x32_10(ctr, key, out);
  return 0;
}

int orc_native_partners(uint64_t seed, int32_t rep, int32_t it, int32_t W, int32_t mode, int32_t* out) {
  orc::native_partners(seed, rep, it, W, mode, out);
  return 0;
}

// partner table of the global block of goodman_weare_wglobal_constant
int orc_native_partners_global(uint64_t seed, int32_t rep, int32_t it, int32_t W, int32_t mode, int32_t* out) {
  orc::native_partners(seed, rep, it, W, mode, out, orc::kTagGlobal);
  return 0;
}

double orc_native_normal(uint64_t seed, uint32_t it, uint64_t chain, uint32_t base, int32_t q) {
  return orc::native_normal(seed, it, chain, base, q);
}

double orc_logp_gaussian(double sos, double noise, int32_t nm) { return orc::logp_gaussian<double>(sos, noise, nm); }

// gpuNEXI.likelihood (NEXI/gpuNEXI.m:822-849), the dictionary search of the
// likelihood start: for each voxel column y of y[m][nv], over the entries
// d_j of dict[m][nd] in order, c = (y' d_j) / dot(d_j, d_j) and
// sse = sum(wt .* (y - c d_j).^2); MATLAB's min skips NaN and returns the
// first minimum of a batch, and a batch's minimum replaces the running
// best only when strictly smaller, so the first global minimum wins.  w
// NULL: wt = 1.  best[v] = -1 when every sse is NaN (pars_best stays 0).
// Threads over voxels (each voxel's scan is sequential).
int orc_dict_search(const float* y, int64_t nv, int32_t m, const float* dict, int64_t nd, const double* w,
                    int32_t* best, double* sse_out, int32_t threads) {
  if (m < 1 || nd < 1 || nv < 0) return GC_ERR_ARG;
  std::vector<double> dd(size_t(m) * nd), nn(nd);
  for (int64_t j = 0; j < nd; ++j) {
    double n = 0;
    for (int k = 0; k < m; ++k) {
      const double v = dict[size_t(k) * nd + j];
      dd[size_t(j) * m + k] = v;
      n += v * v;
    }
    nn[j] = n;
  }
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(static)
  for (int64_t v = 0; v < nv; ++v) {
    double yy[64];
    for (int k = 0; k < m; ++k) yy[k] = y[size_t(k) * nv + v];
    double bs = std::numeric_limits<double>::infinity();
    int32_t bi = -1;
    for (int64_t j = 0; j < nd; ++j) {
      const double* d = &dd[size_t(j) * m];
      double a = 0;
      for (int k = 0; k < m; ++k) a += yy[k] * d[k];
      const double c = a / nn[j];
      double s = 0;
      for (int k = 0; k < m; ++k) {
        const double r = yy[k] - c * d[k];
        s += (w ? w[k] : 1.0) * r * r;
      }
      if (s < bs) {  // false for NaN
        bs = s;
        bi = int32_t(j);
      }
    }
    best[v] = bi;
    if (sse_out) sse_out[v] = bi >= 0 ? bs : std::numeric_limits<double>::quiet_NaN();
  }
  return 0;
}

double orc_norm_inv(double u) { return orc::norm_inv(u); }

// Goodman-Weare native draws (stretch uniform, acceptance uniform)
int orc_native_gw_uniforms(uint64_t seed, uint32_t it, uint64_t chain, int32_t sub, double* u_zz, double* u_acc) {
  orc::native_gw_uniforms(seed, it, chain, sub, u_zz, u_acc);
  return 0;
}

// MH native draws: n[Q] normals and the acceptance uniform
int orc_native_mh_draws(uint64_t seed, uint32_t it, uint64_t chain, int32_t Q, double* n, double* u_acc) {
  if (Q < 1 || Q > GC_MAX_PAR) return GC_ERR_ARG;
  orc::native_mh_draws(seed, it, chain, Q, n, u_acc);
  return 0;
}

// s[M x nv] = FWD(theta[P x nv]); jac (optional) [M x P x nv] = ds/dtheta.
int orc_fwd(const gc_model* md, int64_t nv, const double* theta, const double* extra, double* s, double* jac) {
  using D = orc::Dual<double, GC_MAX_PAR>;
  const int P = md->n_par, M = md->n_meas, E = md->n_extra;
  if (M > 1024 || P > GC_MAX_PAR || E > 64) return GC_ERR_ARG;
  for (int64_t v = 0; v < nv; ++v) {
    double ex[64];
    for (int e = 0; e < E; ++e) ex[e] = extra[size_t(e) * nv + v];
    D th[GC_MAX_PAR];
    for (int p = 0; p < P; ++p) th[p] = D::var(theta[size_t(p) * nv + v], p);
    static thread_local D out[1024];
    orc::fwd_model<D, double>(*md, th, ex, out);
    for (int m = 0; m < M; ++m) {
      s[size_t(m) * nv + v] = out[m].v;
      if (jac)
        for (int p = 0; p < P; ++p) jac[(size_t(m) * P + p) * nv + v] = out[m].d[p];
    }
  }
  return 0;
}

// Plain-value forward in double (no derivatives) -- used for finite differences.
int orc_fwd_value(const gc_model* md, int64_t nv, const double* theta, const double* extra, double* s) {
  const int P = md->n_par, M = md->n_meas, E = md->n_extra;
  if (M > 1024 || P > GC_MAX_PAR || E > 64) return GC_ERR_ARG;
  for (int64_t v = 0; v < nv; ++v) {
    double ex[64], th[GC_MAX_PAR], out[1024];
    for (int e = 0; e < E; ++e) ex[e] = extra[size_t(e) * nv + v];
    for (int p = 0; p < P; ++p) th[p] = theta[size_t(p) * nv + v];
    orc::fwd_model<double, double>(*md, th, ex, out);
    for (int m = 0; m < M; ++m) s[size_t(m) * nv + v] = out[m];
  }
  return 0;
}

int orc_tv(const double* img, const uint8_t* mask, const int32_t dims[3], int32_t mode, const double vsize[3],
           double lambda, int64_t nmasked, double* loss, double* grad) {
  *loss = orc::tv_loss_grad<double>(img, mask, dims, mode, vsize, lambda, nmasked, grad);
  return 0;
}

// The EPG-X pre-pass of one (voxel, flip angle) sample: out[15] = phase,
// IEW and MW corrections, each followed by its derivatives with respect to
// (MVF, R1IEW, kIEWM, R2sIW), from the features and networks of
// fwd_mcrmwi's EPG path, differentiated with duals.
int orc_epgx_sample(const gc_model* md, double mvf, double r1iew, double kfr, double r2siw, double fa, double* out) {
  using D = orc::Dual<double, 4>;
  const D fm = D::var(mvf, 0), r1 = D::var(r1iew, 1), kf = D::var(kfr, 2), r2 = D::var(r2siw, 3);
  D x[11], ph, sd[2];
  orc::epgx_features<D, double>(fm, 1.0 / r1, kf, fa, double(md->f[0]), 1.0 / r2, double(md->f[4]), x);
  orc::mlp_eval<D, double>(md->mlp[0], x, &ph);
  orc::mlp_eval<D, double>(md->mlp[1], x, sd);
  const D* o[3] = {&ph, &sd[0], &sd[1]};
  for (int k = 0; k < 3; ++k) {
    out[5 * k] = o[k]->v;
    for (int q = 0; q < 4; ++q) out[5 * k + 1 + q] = o[k]->d[q];
  }
  return 0;
}

int orc_surface_tv(const double* img, const uint8_t* mask, int64_t ngrid, const int32_t* nbr, const float* elen,
                   int32_t max_nb, double lambda, int64_t nmasked, double* loss, double* grad) {
  for (int64_t i = 0; i < int64_t(max_nb) * ngrid; ++i)
    if (nbr[i] < 0 || nbr[i] >= ngrid) return GC_ERR_ARG;
  *loss = orc::surface_tv_loss_grad<double>(img, mask, ngrid, nbr, elen, max_nb, lambda, nmasked, grad);
  return 0;
}

// askadam.rescale_parameters (utils/askadam.m:1077-1089) on theta [P x nv]:
// sigmoid: clamp to [lb + 1e-4 (ub-lb), ub - 1e-4 (ub-lb)], rescale01, logit;
// linear: rescale01 (:1092-1094).  FP64.
int orc_rescale_parameters(const gc_model* md, int32_t transform, const double* theta, double* z, int64_t nv) {
  for (int p = 0; p < md->n_par; ++p) {
    const double lb = md->lb[p], ub = md->ub[p];
    for (int64_t v = 0; v < nv; ++v) {
      const double x = theta[size_t(p) * nv + v];
      double out;
      if (transform == GC_TRANSFORM_SIGMOID) {
        const double eb = 1e-4 * (ub - lb);
        const double c = std::max(std::min(x, ub - eb), lb + eb);
        const double t = (c - lb) / (ub - lb);
        out = std::log(t / (1 - t));
      } else if (transform == GC_TRANSFORM_LINEAR) {
        out = (x - lb) / (ub - lb);
      } else {
        out = x;
      }
      z[size_t(p) * nv + v] = out;
    }
  }
  return 0;
}

// askadam.unscale_parameters (utils/askadam.m:1063-1073): lb + (ub-lb)
// sigmoid(z), or unscale01 z (ub-lb) + lb (:1097-1099).  FP64.
int orc_unscale_parameters(const gc_model* md, int32_t transform, const double* z, double* theta, int64_t nv) {
  for (int p = 0; p < md->n_par; ++p) {
    const double lb = md->lb[p], ub = md->ub[p];
    for (int64_t v = 0; v < nv; ++v) {
      const double x = z[size_t(p) * nv + v];
      theta[size_t(p) * nv + v] = transform == GC_TRANSFORM_SIGMOID ? lb + (ub - lb) * (1 / (1 + std::exp(-x)))
                                  : transform == GC_TRANSFORM_LINEAR ? x * (ub - lb) + lb
                                                                     : x;
    }
  }
  return 0;
}

// askadam; precision 0 = double (oracle), 1 = float (CPU-path baseline).
// sign_trace (optional, L1): [iteration][ceil(M/16)][nv] residual sign codes
// (gc_volume.sign_trace_dev layout) that replace the oracle's own L1 signs;
// n_mean: N of the loss mean (0 = nv), as gc_volume.n_mean.
// scal_out[10] = {loss, fid, reg, min_loss, min_fid, min_reg, conv, niter, min_iter, stop}
int orc_askadam(const gc_model* md, const gc_askadam_cfg* cfg, int64_t nv, const float* y, const float* w,
                const float* extra, const int64_t* vox, const int32_t* dims, const float* prior_grid,
                const int32_t* surf_nbr, const float* surf_len, int32_t surf_max_nb, const float* z0, double* zfinal,
                double* zmin, double* res_final, double* res_min, double* trace, double* scal_out, int32_t precision,
                int32_t nthreads, const uint32_t* sign_trace, int64_t n_mean) {
  if (md->n_meas > 1024 || md->n_par > GC_MAX_PAR || md->n_extra > 64) return GC_ERR_ARG;
  orc::AskOut o{};
  int rc;
  if (precision == 0) {
    rc = orc::with_np(md->n_par, [&](auto npc) {
      constexpr int NP = decltype(npc)::value;
      return orc::askadam_run<double, NP>(*md, *cfg, nv, y, w, extra, vox, dims, prior_grid, surf_nbr, surf_len,
                                          surf_max_nb, z0, zfinal, zmin, res_final,
                                          res_min, trace, o, std::max(1, nthreads), sign_trace, n_mean);
    });
  } else {
    rc = orc::with_np(md->n_par, [&](auto npc) {
      constexpr int NP = decltype(npc)::value;
      return orc::askadam_run<float, NP>(*md, *cfg, nv, y, w, extra, vox, dims, prior_grid, surf_nbr, surf_len,
                                         surf_max_nb, z0, zfinal, zmin, res_final,
                                         res_min, trace, o, std::max(1, nthreads), sign_trace, n_mean);
    });
  }
  scal_out[0] = o.loss; scal_out[1] = o.fid; scal_out[2] = o.reg;
  scal_out[3] = o.min_loss; scal_out[4] = o.min_fid; scal_out[5] = o.min_reg;
  scal_out[6] = o.conv; scal_out[7] = o.niter; scal_out[8] = o.min_iter; scal_out[9] = o.stop;
  return rc;
}

// MCMC (MH or GW).  Outputs are double arrays in the gc_mcmc_io layouts;
// margin (optional) receives log(acceptance ratio) - log(u) per decision.
int orc_mcmc(const gc_model* md, const gc_mcmc_cfg* cfg, int64_t nv, const float* y, const float* w,
             const float* extra, const float* x0, const float* normals, const float* uniforms,
             const int32_t* partners, double* posterior, double* mean, double* stdv, double* accept, double* xfinal,
             uint8_t* trace, double* margin, int32_t precision, int32_t nthreads) {
  if (md->n_meas > 1024 || cfg->n_var > GC_MAX_PAR || md->n_extra > 64) return GC_ERR_ARG;
  if (cfg->algorithm == GC_MCMC_GW && cfg->ensemble_update == GC_ENSEMBLE_REDBLACK && cfg->n_walker < 2)
    return GC_ERR_NWALKER;
  orc::McmcArgs a{md, cfg, nv, y, w, extra, x0, normals, uniforms, partners, posterior, mean, stdv, accept,
                  xfinal, margin, trace, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  const int nt = std::max(1, nthreads);
  if (cfg->algorithm == GC_MCMC_MH) {
    if (precision == 0) orc::mh_run<double>(a, nt); else orc::mh_run<float>(a, nt);
  } else {
    if (precision == 0) orc::gw_run<double>(a, nt); else orc::gw_run<float>(a, nt);
  }
  return 0;
}

// goodman_weare_wglobal_constant (cfg->global_mask != 0).  Layouts of the
// draws and outputs as in gc_mcmc_io; g_margin receives the global block's
// log(acceptance ratio) - log(u) per decision, [rep][T][W].
int orc_mcmc_global(const gc_model* md, const gc_mcmc_cfg* cfg, int64_t nv, const float* y, const float* w,
                    const float* extra, const float* x0, const float* normals, const float* uniforms,
                    const int32_t* partners, const float* g_uniforms, const int32_t* g_partners, double* posterior,
                    double* g_posterior, double* mean, double* stdv, double* accept, double* xfinal,
                    double* g_xfinal, uint8_t* trace, uint8_t* g_trace, double* margin, double* g_margin,
                    int32_t precision, int32_t nthreads) {
  if (md->n_meas > 1024 || cfg->n_var > GC_MAX_PAR || md->n_extra > 64) return GC_ERR_ARG;
  if (cfg->algorithm != GC_MCMC_GW || cfg->n_walker < 2) return GC_ERR_NWALKER;
  const uint32_t all = (cfg->n_var >= 32) ? 0xFFFFFFFFu : ((1u << cfg->n_var) - 1u);
  if ((cfg->global_mask & all) == 0 || (cfg->global_mask & all) == all || (cfg->global_mask & ~all) != 0)
    return GC_ERR_ARG;
  orc::McmcArgs a{md, cfg, nv, y, w, extra, x0, normals, uniforms, partners, posterior, mean, stdv, accept,
                  xfinal, margin, trace, g_uniforms, g_partners, g_posterior, g_xfinal, g_margin, g_trace};
  const int nt = std::max(1, nthreads);
  if (precision == 0) orc::gwg_run<double>(a, nt); else orc::gwg_run<float>(a, nt);
  return 0;
}

}  // extern "C"
