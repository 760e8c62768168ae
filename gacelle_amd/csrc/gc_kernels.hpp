// gc_kernels.hpp -- gfx950 kernels of the voxelwise-estimation path.
//
//  k_askadam  fused forward + adjoint + transform chain + loss + optimiser
//             update, K iterations per launch with a voxel's data and
//             optimiser state held in registers; per-iteration block
//             partial sums (loss, step, |z|^2, |g|^2) for the host-side
//             stopping logic (utils/askadam.m:261-299).
//  k_eval     one evaluation of model_gradient per voxel (FWD, adjoint,
//             residual outputs) for the operator pair and finalisation.
//  k_mh       Metropolis-Hastings chains (utils/mcmc.m:82-190).
//  k_gw       Goodman-Weare ensembles, simultaneous and red-black
//             (utils/mcmc.m:192-383).
//  k_reduce   deterministic FP64 reduction of block partials.
#pragma once
#include <type_traits>

#include "gc_device.hpp"
#include "gc_models.hpp"

namespace gc {

constexpr int kBlock = 256;
// rows of per-voxel auxiliary data a kernel holds in registers (rows a model
// never reads are dead code)
constexpr int kMaxExtra = 12;

// Kernels that are not templates have internal linkage (static): several
// translation units include this header and launch them.

constexpr int kRing = 16;  // iterations buffered in LDS before a partial flush
constexpr int kReduceChunk = 1024;  // block partials per k_reduce_rows block

// ---- loss functions of askadam.model_gradient (utils/askadam.m:86-95) -----
template <int LOSS>
GC_DEV void loss_fn(float r, float w, float& l, float& dl) {
  if (LOSS == GC_LOSS_L1) {
    l = w * fabsf(r);
    dl = (r == 0.f) ? 0.f : copysignf(w, r);
  } else if (LOSS == GC_LOSS_L2) {
    l = w * r * r;
    dl = 2.f * w * r;
  } else if (LOSS == GC_LOSS_HUBER) {
    const float a = fabsf(r);
    if (a <= 1.f) { l = 0.5f * w * r * r; dl = w * r; }
    else { l = w * (a - 0.5f); dl = w * sgnf(r); }
  } else {  // mse: half squared error, unweighted
    l = 0.5f * r * r;
    dl = r;
  }
}

// ---- parameter transform (utils/askadam.m:1063-1073, forward_pass :424) ---
// theta = lb + (ub - lb) * sigmoid(z)  or  lb + (ub - lb) * clamp(z, 0, 1),
// with d theta / d z.
template <int TR>
GC_DEV void transform_fwd_t(float z, float lb, float ub, float& th, float& dth) {
  const float r = ub - lb;
  if (TR == GC_TRANSFORM_SIGMOID) {
    const float sg = rcpf(1.f + __expf(-z));
    th = lb + r * sg;
    dth = r * sg * (1.f - sg);
  } else if (TR == GC_TRANSFORM_LINEAR) {
    th = fminf(fmaxf(z, 0.f), 1.f) * r + lb;
    dth = (z >= 0.f && z <= 1.f) ? r : 0.f;
  } else {
    th = z;
    dth = 1.f;
  }
}

// The same theta with d theta / d z scaled by the loss mean's 1/count c, for
// the blocked loop: rs = (ub - lb) c is launch-uniform, and the sigmoid's
// derivative sg (1 - sg) is formed as sg - sg^2 (one FMA), so the scaled
// derivative costs an FMA and a multiply.
template <int TR>
GC_DEV void transform_fwd_scaled(float z, float lb, float ub, float rs, float c, float& th, float& dths) {
  if (TR == GC_TRANSFORM_SIGMOID) {
    const float sg = rcpf(1.f + __expf(-z));
    th = fmaf(ub - lb, sg, lb);
    dths = fmaf(-sg, sg, sg) * rs;
  } else if (TR == GC_TRANSFORM_LINEAR) {
    th = fminf(fmaxf(z, 0.f), 1.f) * (ub - lb) + lb;
    dths = (z >= 0.f && z <= 1.f) ? rs : 0.f;
  } else {
    th = z;
    dths = c;
  }
}

// Run-time selection for the single-evaluation kernels.  Branch-free: the
// selection is a select, never control flow around values that are later
// combined.
GC_DEV void transform_fwd(int transform, float z, float lb, float ub, float& th, float& dth) {
  float ts, ds, tl, dl;
  transform_fwd_t<GC_TRANSFORM_SIGMOID>(z, lb, ub, ts, ds);
  transform_fwd_t<GC_TRANSFORM_LINEAR>(z, lb, ub, tl, dl);
  const bool sig = transform == GC_TRANSFORM_SIGMOID;
  const bool lin = transform == GC_TRANSFORM_LINEAR;
  th = sig ? ts : (lin ? tl : z);
  dth = sig ? ds : (lin ? dl : 1.f);
}

// ---- d loss / d theta from one measurement --------------------------------
// A model with kRelJ hands J[0] = d s / d theta_0 >= 0 and, for p > 0, the
// ratio (d s / d theta_p) / J[0] (its post_grad restores any common factor).
// With the L1 fast-path derivative dl = copysign(w, r) and J[0] >= 0,
// dl * J[0] = copysign(w * J[0], r): one bit operation.
template <class Model, bool L1FAST, bool HAS_W>
GC_DEV void accum_grad(float dl, float w, float res, const float* J, float* gth) {
  constexpr int NP = Model::NP;
  if constexpr (Model::kRelJ) {
    const float a = L1FAST ? copysignf(HAS_W ? w * J[0] : J[0], res) : dl * J[0];
    gth[0] += a;
#pragma unroll
    for (int p = 1; p < NP; ++p) gth[p] = fmaf(a, J[p], gth[p]);
  } else {
#pragma unroll
    for (int p = 0; p < NP; ++p) gth[p] = fmaf(dl, J[p], gth[p]);
  }
}

// ---- forward + loss + adjoint of one voxel -----------------------------------
// lossfn(r, s) evaluates the loss of measurement r at model value s and
// returns d loss / d s (0 for a measurement that carries no gradient).
// Models with a fused adjoint (kFusedAdj) accumulate the gradient
// themselves; the others hand a Jacobian per measurement.
template <class Model>
struct has_aux {
  template <class M>
  static constexpr bool f(decltype(M::kAux)*) { return M::kAux; }
  template <class M>
  static constexpr bool f(...) { return false; }
  static constexpr bool value = f<Model>(nullptr);
};

template <class Model, class LF>
GC_DEV void model_adjoint(const gc_model& md, const float* th, const float* ex, LF&& lossfn, float* gth,
                          AuxRef aux = {}) {
  if constexpr (has_aux<Model>::value) {
    Model::eval_adj(md, th, ex, lossfn, gth, aux);
  } else if constexpr (Model::kFusedAdj) {
    Model::eval_adj(md, th, ex, lossfn, gth);
  } else {
    Model::template eval<true>(md, th, ex, [&](int r, float s, const float* J) {
      const float dl = lossfn(r, s);
      if (dl != 0.f) accum_grad<Model, false, false>(dl, 1.f, 0.f, J, gth);
    });
  }
  Model::post_grad(th, gth);
}

// Models whose gradient factors per group of measurements (kBaseAdj): for
// measurement r of a group, d s_r / d theta_p = base_r (alpha_p + beta_p tw_r)
// with base_r >= 0 and group constants alpha, beta, so the model accumulates
// sum_r dl_r base_r and sum_r dl_r base_r tw_r and applies alpha and beta
// once per group.  eval_base_adj(md, th, vp, f, gth) calls
// f(r, s_r, base_r) -> dl_r base_r (0 for a measurement without gradient)
// for every used slot and adds d loss / d theta to gth; vp holds the
// model's per-voxel constants (voxel_prep(md, ex, vp), kVoxPrep floats).
template <class Model>
struct has_base_adj {
  template <class M>
  static constexpr bool f(decltype(M::kBaseAdj)*) { return M::kBaseAdj; }
  template <class M>
  static constexpr bool f(...) { return false; }
  static constexpr bool value = f<Model>(nullptr);
};

// per-voxel constants of a grouped-adjoint model (Model::voxel_prep fills
// kVoxPrep floats that eval_base_adj reads)
template <class Model>
struct vox_prep {
  template <class M>
  static constexpr int f(decltype(M::kVoxPrep)*) { return M::kVoxPrep; }
  template <class M>
  static constexpr int f(...) { return 1; }
  static constexpr int value = f<Model>(nullptr);
};

// ---- askadam block kernel -------------------------------------------------
// parameters of the folded Adam step (the blocked models have at most five)
constexpr int kEpsSlots = 8;

struct EpochScal {
  float lr;    // adam: lr*sqrt(1-b2^t)/(1-b1^t) (effective form) or lr; sgdm/rmsprop: lr
  float c1;    // 1/(1-b1^t) (Kingma-Ba form)
  float c2;    // 1/(1-b2^t) (Kingma-Ba form)
  // adam, effective form (adam_eff_den): 1 / ls for the step
  // ls = lr (1-b1) / sqrt(1-b2), and per parameter eps_s / (ls rs_p)
  float inv_ls;
  float eps_ls[kEpsSlots];
};

struct VolDev {
  const float* y;
  const float* w;
  const float* ex;
  const float* z_in;
  const float* m_in;
  const float* v_in;
  float* z_out;
  float* m_out;
  float* v_out;
  float* z_last;     // z (pre-update) at the launch's last iteration, or null
  float* partials;   // [4 nblk][n_epochs][npv], npv = 4 or 1: the sums of each
                     // wavefront (64 voxels) contiguous
  const EpochScal* es;
  int64_t nv;
  int32_t blk0, nblk;  // in blocks of kBlock voxels (the volumes' padding)
  int32_t epoch0, n_epochs, iteration, pad;
  float inv_count;   // 1/(M*nv)
  float pad2[3];
  // test hook (k_askadam<..., SIGNS = true>): residual signs, two bits per
  // measurement, [iteration][ceil(M/16)][nv] words, row e-1 for epoch e
  uint32_t* signs;
};

// optimiser of askadam.update_parameters (utils/askadam.m:673-691)
struct OptParams {
  int32_t opt;
  int32_t adam_form;
  float beta1, beta2, eps, momentum, rms_decay, rms_eps;
};

// the parameter step dz of one slot; m1 / m2 are its moment estimates
GC_DEV float opt_step(const OptParams& o, const EpochScal& es, float g, float& m1, float& m2) {
  float dz;
  if (o.opt == GC_OPT_ADAM) {
    m1 = o.beta1 * m1 + (1.f - o.beta1) * g;
    m2 = o.beta2 * m2 + (1.f - o.beta2) * g * g;
    if (o.adam_form == GC_ADAM_KINGMA_BA)
      dz = -es.lr * (m1 * es.c1) * rcpf(sqrtv(m2 * es.c2) + o.eps);
    else
      dz = -es.lr * m1 * rcpf(sqrtv(m2) + o.eps);
  } else if (o.opt == GC_OPT_SGDM) {
    m1 = o.momentum * m1 - es.lr * g;
    dz = m1;
  } else {
    m2 = o.rms_decay * m2 + (1.f - o.rms_decay) * g * g;
    dz = -es.lr * g * rcpf(sqrtv(m2) + o.rms_eps);
  }
  return dz;
}

// Adam in its default (effective learning-rate) form for loops instantiated
// with the optimiser known, on scaled moments m1 = m / (1 - beta1) and
// m2 = v / (1 - beta2): m1 <- beta1 m1 + g and m2 <- beta2 m2 + g^2 drop the
// (1 - beta) factors, and the step
//   lr m / (sqrt(v) + eps) = [lr (1 - beta1) / sqrt(1 - beta2)] m1 / (sqrt(m2) + eps / sqrt(1 - beta2)).
// The blocked kernel also drops the positive factor rs_p = (ub_p - lb_p) / N
// of the sigmoid's derivative, which is uniform over the launch: it runs on
// g / rs_p, with moments m1 / rs_p and m2 / rs_p^2, and Adam's step is
// invariant under that scaling once eps is divided by rs_p too.  With the
// bracket ls and the scaled epsilon eps_s = eps / sqrt(1 - beta2) the
// denominator is (sqrt(m2) + eps_s / rs_p) / ls = fma(sqrt(m2), 1 / ls,
// eps_s / (ls rs_p)), whose constants come from the host per iteration.
// The blocked kernel keeps its Adam state in this scaled form whenever it
// runs this update (every launch of such a call does).  adam_eff_den updates
// the moments and returns the denominator; the step is -m1 / den (the
// kernel takes the reciprocals of all slots together, rcp_batch).
GC_DEV float adam_eff_den(const OptParams& o, float inv_ls, float eps_ls, float g, float& m1, float& m2) {
  // m1 <- beta1 m1 + g as a three-address v_fma_f32 (the compiler's
  // accumulator form would put the result in g's register and copy it
  // back to m1's every iteration)
  float m1n;
  asm("v_fma_f32 %0, %1, %2, %3" : "=v"(m1n) : "s"(o.beta1), "v"(m1), "v"(g));
  m1 = m1n;
  m2 = fmaf(g, g, o.beta2 * m2);
  // (sqrt(m2) + eps_s / rs_p) / ls: the step folded into the denominator
  return fmaf(sqrtv(m2), inv_ls, eps_ls);
}

struct AskParams {
  const VolDev* vols;
  int32_t n_vol;
  int32_t transform;
  int32_t need_norms;  // step / |z|^2 / |g|^2 sums wanted (step or gradient tolerance active)
  int32_t adam_eff;    // Adam in the folded effective form (adam_eff_den); sigmoid transform
  OptParams o;
};

GC_DEV int find_volume(const VolDev* vols, int n_vol, int blk) {
  int vi = 0;
  for (int i = 1; i < n_vol; ++i)
    if (blk >= vols[i].blk0) vi = i;
  return vi;
}

// minimum waves per SIMD the blocked askadam kernel is compiled for (a
// register budget)
template <class Model>
struct ask_waves {
  template <class M>
  static constexpr int f(decltype(M::kAskWaves)*) { return M::kAskWaves; }
  template <class M>
  static constexpr int f(...) { return 1; }
  static constexpr int value = f<Model>(nullptr);
};

// Launched with 64, 128 or 256 threads per block (picked by a timed trial or
// set with GC_OPT_ASK_BLOCK); the volumes are padded to kBlock voxels, so a
// block never straddles two.  Every wavefront writes its own per-iteration
// sums (a wavefront covers the same 64 voxels whatever the block size), so
// the results do not depend on the block size.  The loss-only sums wait in
// a wavefront's own ring of dynamic LDS (kRing x 64 floats) and are reduced
// once per kRing iterations; no block barrier is needed.
//
// SIGNS (L1 only): also write the sign code of every differentiated residual
// to V.signs (1 = positive, 2 = negative, 0 = zero or a non-finite model
// signal), so that an FP64 replay can follow the same subgradient path.  The
// arithmetic is that of the SIGNS = false kernel; only the stores differ.
template <class Model, int LOSS, bool HAS_W, bool SIGNS = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(ask_waves<Model>::value)))
void k_askadam(const gc_model* __restrict__ mdp, const AskParams ap) {
  const gc_model& md = *mdp;
  constexpr int NP = Model::NP;
  constexpr int MR = Model::MR;
  const int64_t g0 = int64_t(blockIdx.x) * blockDim.x;  // first voxel slot of the block
  const int vi = find_volume(ap.vols, ap.n_vol, int(g0 / kBlock));
  const VolDev V = ap.vols[vi];
  if (V.n_epochs <= 0) return;
  const int64_t nv = V.nv;
  const int64_t v = g0 - int64_t(V.blk0) * kBlock + threadIdx.x;
  const bool valid = v < nv;
  const int64_t vv = valid ? v : (nv - 1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t lw = v >> 6;  // the wavefront's partial row within the volume

  float yr[MR];
  float wr[HAS_W ? MR : 1];
#pragma unroll
  for (int r = 0; r < MR; ++r) {
    if (Model::valid(md, r)) {
      const int64_t m = Model::meas(md, r);
      yr[r] = V.y[m * nv + vv];
      if (HAS_W) wr[HAS_W ? r : 0] = V.w[m * nv + vv];
    } else {
      yr[r] = 0.f;
      if (HAS_W) wr[HAS_W ? r : 0] = 0.f;
    }
  }
  float ex[kMaxExtra];
#pragma unroll
  for (int e = 0; e < kMaxExtra; ++e) ex[e] = (e < md.n_extra) ? V.ex[int64_t(e) * nv + vv] : 0.f;
  // the grouped-adjoint models' per-voxel constants, formed once
  float vp[has_base_adj<Model>::value ? vox_prep<Model>::value : 1];
  if constexpr (has_base_adj<Model>::value) Model::voxel_prep(md, ex, vp);

  float z[NP], m1[NP], m2[NP], lbp[NP], ubp[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    z[p] = V.z_in[int64_t(p) * nv + vv];
    m1[p] = V.m_in[int64_t(p) * nv + vv];
    m2[p] = V.v_in[int64_t(p) * nv + vv];
    lbp[p] = md.lb[p];
    ubp[p] = md.ub[p];
  }
  // slots the model never reads: constant, counted in |z|^2 only
  float dead_z2 = 0.f;
  for (int p = NP; p < md.n_par; ++p) {
    const float zd = V.z_in[int64_t(p) * nv + vv];
    dead_z2 += zd * zd;
  }

  // The epoch table is read-only for the whole launch: reading it through
  // the constant address space lets the wave-uniform loads go scalar.
  using ConstEpochScal = __attribute__((address_space(4))) const EpochScal;
  ConstEpochScal* es_table = reinterpret_cast<ConstEpochScal*>(reinterpret_cast<uintptr_t>(V.es));
  // Per-iteration wavefront sums.  With the norms (step and gradient
  // stopping criteria) each wavefront reduces its four sums with DPP and
  // cross-row permutes and stores them; for the loss alone every thread parks
  // its value in the wavefront's LDS ring, reduced once per kRing iterations
  // (four lanes per iteration), which keeps the reduction off the
  // per-iteration VALU path.
  extern __shared__ float gc_ask_ring[];
  float* const ring = gc_ask_ring + wave * (kRing * 64);
  // voxels of this wavefront inside the volume (wave-uniform)
  const int64_t nvw = nv - (v - lane);
  const bool wave_full = nvw >= 64;
  const float inv_count = V.inv_count;
  float rsp[NP];  // (ub - lb) / count
#pragma unroll
  for (int p = 0; p < NP; ++p) rsp[p] = (ubp[p] - lbp[p]) * inv_count;
  const XRow xr = xrow_addr(lane);

  // The epoch loop, instantiated per parameter transform and per norms flag
  // so that neither costs a per-iteration selection, and with the sigmoid
  // transform also for the default optimiser (Adam, effective form), whose
  // update then runs without the optimiser switch.
  auto epochs = [&](auto tr, auto nm, auto oa) {
    constexpr int TR = decltype(tr)::value;
    constexpr bool norms = decltype(nm)::value;
    constexpr bool adam_eff = decltype(oa)::value;
    for (int k = 0; k < V.n_epochs; ++k) {
      const int e = V.epoch0 + k;
      float th[NP], dths[NP];
      if constexpr (TR == GC_TRANSFORM_SIGMOID) {
        // transform_fwd_scaled's sigmoid with the reciprocals of all slots
        // taken together
        float den1[NP], sg[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) den1[p] = 1.f + __expf(-z[p]);
        rcp_batch<NP>(den1, sg);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          th[p] = fmaf(ubp[p] - lbp[p], sg[p], lbp[p]);
          dths[p] = fmaf(-sg[p], sg[p], sg[p]) * (adam_eff ? 1.f : rsp[p]);
        }
      } else {
#pragma unroll
        for (int p = 0; p < NP; ++p)
          transform_fwd_scaled<TR>(z[p], lbp[p], ubp[p], rsp[p], inv_count, th[p], dths[p]);
      }
      float lsum = 0.f;
      float gth[NP];
      // Fast pass without per-element checks; a non-finite model signal makes
      // the loss sum non-finite, and only then is the voxel re-evaluated with
      // utils.set_nan_inf_zero semantics (zero signal, no gradient).  The fast
      // pass takes the L1 derivative as copysign(w, r) and keeps min |r|: an
      // exact zero residual (sign 0) sends the voxel to the checked pass too.
      float amin = 1.f;
      // sign words of this voxel and epoch (SIGNS): word j at sw[j * nv]
      uint32_t* const sw = SIGNS ? V.signs + int64_t(e - 1) * ((md.n_meas + 15) >> 4) * nv + v : nullptr;
      auto accumulate = [&](auto checked) {
        constexpr bool CHECK = decltype(checked)::value;
        constexpr bool L1FAST = !CHECK && LOSS == GC_LOSS_L1;
        lsum = 0.f;
        // -0 is the additive identity (-0 + x = x for every x, +0 + -0 is
        // +0), so the first gradient term needs no add
#pragma unroll
        for (int p = 0; p < NP; ++p) gth[p] = -0.f;
        if constexpr (SIGNS) {
          if (valid)
            for (int j = 0; j < ((md.n_meas + 15) >> 4); ++j) sw[int64_t(j) * nv] = 0u;
        }
        // loss of measurement r at model value s: residual, loss sum, the
        // sign record (SIGNS); returns the residual and, through dl and
        // fin, the loss derivative and whether the measurement carries one
        auto measure = [&](int r, float s, float& dl, bool& fin) {
          fin = true;
          float sv = s;
          if (CHECK) {
            fin = isfinite(s);
            sv = fin ? s : 0.f;
          }
          const float w = HAS_W ? wr[HAS_W ? r : 0] : 1.f;
          const float res = sv - yr[r];
          float l;
          dl = 0.f;
          if (L1FAST) {
            l = w * fabsf(res);
            amin = fminf(amin, fabsf(res));
          } else {
            loss_fn<LOSS>(res, w, l, dl);
          }
          lsum += l;
          if constexpr (SIGNS) {
            if (valid) {
              const int m = Model::meas(md, r);
              const uint32_t code = (!CHECK || fin) ? (res > 0.f ? 1u : (res < 0.f ? 2u : 0u)) : 0u;
              uint32_t* const p = sw + int64_t(m >> 4) * nv;
              *p = *p | (code << ((m & 15) * 2));
            }
          }
          return res;
        };
        if constexpr (has_base_adj<Model>::value) {
          // the model accumulates its gradient from d loss / d s times the
          // measurement's base signal (>= 0): with the L1 fast-path
          // derivative copysign(w, r) that product is one bit operation
          Model::eval_base_adj(md, th, vp, [&](int r, float s, float base) {
            float dl;
            bool fin;
            const float res = measure(r, s, dl, fin);
            const float w = HAS_W ? wr[HAS_W ? r : 0] : 1.f;
            if (L1FAST) return copysignf(HAS_W ? w * base : base, res);
            return fin ? dl * base : 0.f;
          }, gth);
        } else {
          Model::template eval<true>(md, th, ex, [&](int r, float s, const float* J) {
            float dl;
            bool fin;
            const float res = measure(r, s, dl, fin);
            const float w = HAS_W ? wr[HAS_W ? r : 0] : 1.f;
            if (!CHECK || fin) accum_grad<Model, L1FAST, HAS_W>(L1FAST ? copysignf(w, res) : dl, w, res, J, gth);
          });
        }
      };
      accumulate(std::false_type{});
      // one test for both: lsum * 0 is NaN exactly when lsum is not finite
      // (lsum >= 0), and otherwise the sum is amin itself
      if (!(fmaf(lsum, 0.f, amin) > 0.f)) accumulate(std::true_type{});
      Model::post_grad(th, gth);
      float den = dead_z2, gn = 0.f, step = 0.f;
      float g[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p) g[p] = gth[p] * dths[p];
      if constexpr (norms) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          den += z[p] * z[p];
          const float gt = adam_eff ? g[p] * rsp[p] : g[p];  // d loss / d z
          gn += gt * gt;
        }
      }
      if (k == V.n_epochs - 1 && V.z_last != nullptr && valid) {
#pragma unroll
        for (int p = 0; p < NP; ++p) V.z_last[int64_t(p) * nv + v] = z[p];
      }
      if (e < V.iteration) {  // update_parameters skips the last iteration (:677)
        const EpochScal es{es_table[k].lr, es_table[k].c1, es_table[k].c2};
        // 1 / ls in a VGPR, so that each parameter's fma reads one SGPR
        // (its eps) instead of copying it to a VGPR
        float inv_ls = 0.f;
        float rden[NP];
        if constexpr (adam_eff) {
          asm("v_mov_b32 %0, %1" : "=v"(inv_ls) : "s"(es_table[k].inv_ls));
          float dn[NP];
#pragma unroll
          for (int p = 0; p < NP; ++p)
            dn[p] = adam_eff_den(ap.o, inv_ls, es_table[k].eps_ls[p < kEpsSlots ? p : 0], g[p], m1[p], m2[p]);
          rcp_batch<NP>(dn, rden);
        }
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          float dz;
          if constexpr (adam_eff)
            dz = -m1[p] * rden[p];
          else
            dz = opt_step(ap.o, es, g[p], m1[p], m2[p]);
          const float zn = z[p] + dz;
          if constexpr (norms) {
            const float d = zn - z[p];
            step += d * d;
          }
          z[p] = zn;
        }
      }
      if constexpr (norms) {
        if (!valid) { lsum = 0.f; step = 0.f; den = 0.f; gn = 0.f; }
        // four sums per iteration: loss, step^2, |z|^2, |g|^2
        const float tot = wave_sum4(lsum, step, den, gn, lane, xr);
        if (lane < 4) V.partials[((lw * V.n_epochs + k) << 2) + wave_sum4_index(lane)] = tot;
      } else {
        // every lane parks its loss; the lanes past the volume's end (only
        // in its last wavefront) are dropped when the ring is summed
        const int slot = k & (kRing - 1);
        ring[slot * 64 + lane] = lsum;
        if (slot == kRing - 1 || k == V.n_epochs - 1) {
          // lanes 4 kk .. 4 kk + 3 sum ring slot kk, 16 values each; LDS
          // operations of one wavefront complete in order, so the slot
          // stores above are visible to these loads
          static_assert(kRing * 4 == 64, "four lanes per ring slot");
          __builtin_amdgcn_wave_barrier();
          const int cnt = slot + 1;
          const int kbase = k - slot;
          const int kk = lane >> 2, part = lane & 3;
          float s = 0.f;
          if (kk < cnt) {
            const float4* p4 = reinterpret_cast<const float4*>(&ring[kk * 64 + part * 16]);
            if (wave_full) {
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const float4 q = p4[i];
                s += (q.x + q.y) + (q.z + q.w);
              }
            } else {
#pragma unroll
              for (int i = 0; i < 4; ++i) {
                const float4 q = p4[i];
                const int l0 = part * 16 + 4 * i;
                s += ((l0 < nvw ? q.x : 0.f) + (l0 + 1 < nvw ? q.y : 0.f)) +
                     ((l0 + 2 < nvw ? q.z : 0.f) + (l0 + 3 < nvw ? q.w : 0.f));
              }
            }
          }
          s += dpp<kDppXor1>(s);
          s += dpp<kDppXor2>(s);
          if (part == 0 && kk < cnt) V.partials[lw * V.n_epochs + kbase + kk] = s;
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
  };
  const bool with_norms = ap.need_norms != 0;
  // (the host sets it for Adam's effective form with beta2 < 1, at most
  // kEpsSlots parameters and ub > lb)
  const bool adam_eff = NP <= kEpsSlots && ap.adam_eff != 0;
  using Lin = std::integral_constant<int, GC_TRANSFORM_LINEAR>;
  using Sig = std::integral_constant<int, GC_TRANSFORM_SIGMOID>;
  if (ap.transform == GC_TRANSFORM_LINEAR) {
    if (with_norms) epochs(Lin{}, std::true_type{}, std::false_type{});
    else epochs(Lin{}, std::false_type{}, std::false_type{});
  } else if (adam_eff) {
    if (with_norms) epochs(Sig{}, std::true_type{}, std::true_type{});
    else epochs(Sig{}, std::false_type{}, std::true_type{});
  } else {
    if (with_norms) epochs(Sig{}, std::true_type{}, std::false_type{});
    else epochs(Sig{}, std::false_type{}, std::false_type{});
  }

  if (valid) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      V.z_out[int64_t(p) * nv + v] = z[p];
      V.m_out[int64_t(p) * nv + v] = m1[p];
      V.v_out[int64_t(p) * nv + v] = m2[p];
    }
    for (int p = NP; p < md.n_par; ++p) {
      const float zd = V.z_in[int64_t(p) * nv + v];
      V.z_out[int64_t(p) * nv + v] = zd;
      V.m_out[int64_t(p) * nv + v] = 0.f;
      V.v_out[int64_t(p) * nv + v] = 0.f;
      if (V.z_last != nullptr) V.z_last[int64_t(p) * nv + v] = zd;
    }
  }
}

// ---- askadam, one iteration per launch --------------------------------------
// For what the K-iteration kernel cannot hold: the spatial TV term of
// askadam.model_gradient (utils/askadam.m:99, utils/spatial_total_variation.m:
// 24-66), which couples each voxel to its grid neighbours' previous
// parameters, and models whose measurements do not fit in registers
// (streamed from HBM every iteration).  Each launch advances every active
// volume by one iteration from (z, m, v)_in to (z, m, v)_out.
// fidelity, step^2, |z|^2, |g|^2, regularisation, per-voxel mean loss over
// the convergence set (robustConvergence)
constexpr int kStepSums = 6;

struct StepVol {
  const float* y;
  const float* w;
  const float* ex;
  const float* z_in;
  const float* m_in;
  const float* v_in;
  float* z_out;
  float* m_out;
  float* v_out;
  float* z_last;       // z (pre-update) of this iteration, or null
  float* partials;     // [nblk][n_epochs][kStepSums]
  const int32_t* nbr;  // [2*nd][nv]: +axis, -axis neighbour G-D index per axis, -1 outside the mask
  const float* prior;  // GC_REG_PRIOR: [2*n_reg][nv], mu rows then sigma rows, or null
  const float* wmod;   // robustConvergence: per-voxel weight factor of this iteration, or null
  const uint8_t* mconv;  // robustConvergence: convergence voxel set, or null
  float* pvl_out;      // robustConvergence: per-voxel mean loss of this iteration, or null
  float* aux;          // models with per-iteration aux rows (MCR-MWI EPG-X): [rows][nv]
  // GC_REG_SURFACE_TV: per masked vertex, its mesh neighbours (G-D index,
  // -1 outside the mask, -2 ends the list) with 1/edgeLength, [snf][nv],
  // and the masked vertices listing it as a neighbour (-1 ends the list)
  // with their 1/edgeLength, [snr][nv]
  const int32_t* snb;
  const float* sie;
  const int32_t* srv;
  const float* srie;
  int32_t snf, snr;
  int64_t nv;
  int32_t blk0, nblk;
  int32_t active;      // the volume runs this iteration
  int32_t k, n_epochs; // this iteration's row among the launch sequence's partials
  int32_t update;      // update_parameters runs (epoch < iteration, :677)
  float inv_count;     // 1/(M nv)
  float inv_nv;        // 1/Nmasked of spatial_total_variation
  EpochScal es;
};

struct StepParams {
  const StepVol* vols;
  int32_t n_vol;
  int32_t n_reg;       // regularised maps (0: none)
  int32_t nd;          // TV axes: 2 ('2D') or 3 ('3D')
  int32_t reg_type;    // GC_REG_TV | GC_REG_PRIOR | GC_REG_SURFACE_TV
  int32_t regmap[4];   // parameter slot of each regularised map
  float lam[4];
  float inv_vs[3];     // 1/voxelSize
  OptParams o;
};

// minimum waves per SIMD the step kernel is compiled for (a register
// budget): models may ask for more than the default
template <class Model>
struct step_waves {
  template <class M>
  static constexpr int f(decltype(M::kStepWaves)*) { return M::kStepWaves; }
  template <class M>
  static constexpr int f(...) { return 1; }
  static constexpr int value = f<Model>(nullptr);
};

// measurement streams of a model's loss calls (MeasAhead); 0: the model
// declares no call order
template <class Model>
struct meas_streams {
  template <class M>
  static constexpr int f(decltype(M::kMeasStreams)*) { return M::kMeasStreams; }
  template <class M>
  static constexpr int f(...) { return 0; }
  static constexpr int value = f<Model>(nullptr);
};

// The step kernel's reads of one voxel's measurements (and weights).  A
// model that declares kMeasStreams = S calls the loss in a fixed order: S
// streams of len = n_meas / S calls (stream c: c len + k, k ascending),
// interleaved call by call (k of every stream, then k + 1, ...).  The reader
// then keeps the values of the next kAhead calls in flight, with no branch:
// each call takes the oldest and loads the one kAhead calls later (kAhead /
// S steps further in the same stream).  Models without a declared order
// read each value when called.
template <class Model, bool HAS_W, int S = meas_streams<Model>::value>
struct MeasAhead {
  static constexpr int kAhead = 4;
  static_assert(S == 1 || S == 2, "one or two interleaved measurement streams");
  const gc_model& md;
  gfloat* y;
  gfloat* w;
  int64_t nv, vv;
  int32_t n;
  float yq[kAhead], wq[kAhead];

  GC_DEV MeasAhead(const gc_model& m, const float* yp, const float* wp, int64_t nv_, int64_t vv_)
      : md(m), y(as_global(yp)), w(as_global(wp)), nv(nv_), vv(vv_), n(m.n_meas) {
    const int32_t len = n / S;
#pragma unroll
    for (int i = 0; i < kAhead; ++i) {
      const int32_t r = (i % S) * len + i / S;  // the i-th call
      yq[i] = ld(y, r);
      if constexpr (HAS_W) wq[i] = ld(w, r);
    }
  }
  GC_DEV float ld(gfloat* p, int32_t r) const {
    r = r < n ? r : n - 1;  // reads past the last call stay in the array
    return p[int64_t(Model::meas(md, r)) * nv + vv];
  }
  GC_DEV void get(int32_t r, float& yv, float& wv) {
    yv = yq[0];
    wv = HAS_W ? wq[0] : 1.f;
#pragma unroll
    for (int i = 0; i + 1 < kAhead; ++i) {
      yq[i] = yq[i + 1];
      if constexpr (HAS_W) wq[i] = wq[i + 1];
    }
    yq[kAhead - 1] = ld(y, r + kAhead / S);
    if constexpr (HAS_W) wq[kAhead - 1] = ld(w, r + kAhead / S);
  }
};

template <class Model, bool HAS_W>
struct MeasAhead<Model, HAS_W, 0> {
  const gc_model& md;
  gfloat* y;
  gfloat* w;
  int64_t nv, vv;
  GC_DEV MeasAhead(const gc_model& m, const float* yp, const float* wp, int64_t nv_, int64_t vv_)
      : md(m), y(as_global(yp)), w(as_global(wp)), nv(nv_), vv(vv_) {}
  GC_DEV void get(int32_t r, float& yv, float& wv) {
    const int64_t m = Model::meas(md, r);
    yv = y[m * nv + vv];
    wv = HAS_W ? w[m * nv + vv] : 1.f;
  }
};

template <class Model, int LOSS, bool HAS_W, int TR>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(step_waves<Model>::value)))
void k_askadam_step(const gc_model* __restrict__ mdp, const StepParams sp) {
  const gc_model& md = *mdp;
  constexpr int NP = Model::NP;
  int vi = 0;
  for (int i = 1; i < sp.n_vol; ++i)
    if (int(blockIdx.x) >= sp.vols[i].blk0) vi = i;
  const StepVol& V = sp.vols[vi];
  if (!V.active) return;
  const int lblk = blockIdx.x - V.blk0;
  const int64_t nv = V.nv;
  const int64_t v = int64_t(lblk) * kBlock + threadIdx.x;
  const bool valid = v < nv;
  const int64_t vv = valid ? v : (nv - 1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // the first measurements are read while the parameters are prepared
  MeasAhead<Model, HAS_W> meas(md, V.y, V.w, nv, vv);

  float ex[kMaxExtra];
#pragma unroll
  for (int e = 0; e < kMaxExtra; ++e) ex[e] = (e < md.n_extra) ? V.ex[int64_t(e) * nv + vv] : 0.f;
  // only theta lives through the model evaluation: z is read again after it
  // and the transform's derivative formed again from it (the same
  // arithmetic, so the same bits), and the optimiser state is read there
  float th[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    float d;
    transform_fwd_t<TR>(V.z_in[int64_t(p) * nv + vv], md.lb[p], md.ub[p], th[p], d);
  }
  float dead_z2 = 0.f;
  for (int p = NP; p < md.n_par; ++p) {
    const float zd = V.z_in[int64_t(p) * nv + vv];
    dead_z2 += zd * zd;
  }

  // data term (set_nan_inf_zero semantics on the model signal)
  float lsum = 0.f;
  float gth[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) gth[p] = 0.f;
  const float wm = (V.wmod != nullptr) ? V.wmod[vv] : 1.f;
  model_adjoint<Model>(
      md, th, ex,
      [&](int r, float s) -> float {
        float yv, wr;
        meas.get(r, yv, wr);
        const bool fin = isfinite(s);
        const float sv = fin ? s : 0.f;
        const float wv = wr * wm;
        float l, dl;
        loss_fn<LOSS>(sv - yv, wv, l, dl);
        lsum += l;
        return fin ? dl : 0.f;
      },
      gth, AuxRef{V.aux ? V.aux + vv : nullptr, nv});
  // a compiler-only barrier: without it the loads below would be merged with
  // the ones above and z held in registers through the model
  asm volatile("" ::: "memory");
  float z[NP], g[NP], dth[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    z[p] = V.z_in[int64_t(p) * nv + vv];
    float t;
    transform_fwd_t<TR>(z[p], md.lb[p], md.ub[p], t, dth[p]);
    g[p] = gth[p] * dth[p] * V.inv_count;
  }
  const float pvl = lsum / float(md.n_meas);
  if (V.pvl_out != nullptr && valid) V.pvl_out[v] = pvl;
  float msum = (V.mconv != nullptr && V.mconv[vv]) ? pvl : 0.f;

  // spatial TV: cost sum_d |theta(i + e_d) - theta(i)| / vs_d over masked i
  // (forward differences with periodic wrap, zeros outside the mask), and
  // its subgradient (sign(0) = 0) at i from i's own and its -e_d
  // neighbour's terms
  //
  // normal prior (gpuNEXI.prior_distribution_normal): cost
  // |(theta - mu) / sigma| where finite, subgradient sign(.) / sigma; the
  // grid voxels outside the mask add a constant the host carries
  float reg = 0.f;
  for (int kr = 0; kr < sp.n_reg; ++kr) {
    const int pr = sp.regmap[kr];
    const float lam = sp.lam[kr];
    const float lb = md.lb[pr], ub = md.ub[pr];
    float ti = 0.f, dti = 0.f;
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (p == pr) { ti = th[p]; dti = dth[p]; }
    float gt = 0.f;
    if (sp.reg_type == GC_REG_PRIOR) {
      const float mu = V.prior[int64_t(kr) * nv + vv];
      const float sd = V.prior[int64_t(sp.n_reg + kr) * nv + vv];
      const float zs = (ti - mu) / sd;
      if (isfinite(zs)) {
        reg += lam * fabsf(zs);
        gt = lam * sgnf(zs) / sd;
      }
    }
    if (sp.reg_type == GC_REG_SURFACE_TV) {
      // mesh TV (gpuNEXI.surface_total_variation): cost sum over i's
      // neighbours u of |theta_i - theta_u| / len (theta_u = 0 outside the
      // mask); i's subgradient from its own terms and from the terms of the
      // masked vertices that list i
      for (int n = 0; n < V.snf; ++n) {
        const int32_t u = V.snb[int64_t(n) * nv + vv];
        if (u == -2) break;
        const float ie = V.sie[int64_t(n) * nv + vv];
        float tu = 0.f, dd;
        if (u >= 0) transform_fwd_t<TR>(V.z_in[int64_t(pr) * nv + u], lb, ub, tu, dd);
        const float dv = (ti - tu) * ie;
        reg += lam * fabsf(dv);
        gt += lam * sgnf(dv) * ie;
      }
      for (int n = 0; n < V.snr; ++n) {
        const int32_t wv = V.srv[int64_t(n) * nv + vv];
        if (wv < 0) break;
        const float ie = V.srie[int64_t(n) * nv + vv];
        float tw, dd;
        transform_fwd_t<TR>(V.z_in[int64_t(pr) * nv + wv], lb, ub, tw, dd);
        gt -= lam * sgnf((tw - ti) * ie) * ie;
      }
    }
    for (int d = 0; d < (sp.reg_type == GC_REG_TV ? sp.nd : 0); ++d) {
      const float ivs = sp.inv_vs[d];
      const int32_t jf = V.nbr[int64_t(2 * d) * nv + vv];
      float tf = 0.f, dd;
      if (jf >= 0) transform_fwd_t<TR>(V.z_in[int64_t(pr) * nv + jf], lb, ub, tf, dd);
      const float Di = (tf - ti) * ivs;
      reg += lam * fabsf(Di);
      gt -= sgnf(Di) * ivs * lam;
      const int32_t jb = V.nbr[int64_t(2 * d + 1) * nv + vv];
      if (jb >= 0) {
        float tb;
        transform_fwd_t<TR>(V.z_in[int64_t(pr) * nv + jb], lb, ub, tb, dd);
        gt += sgnf((ti - tb) * ivs) * ivs * lam;
      }
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
      if (p == pr) g[p] += gt * V.inv_nv * dti;
  }

  float den = dead_z2, gn = 0.f, step = 0.f;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    den += z[p] * z[p];
    gn += g[p] * g[p];
  }
  if (V.z_last != nullptr && valid) {
#pragma unroll
    for (int p = 0; p < NP; ++p) V.z_last[int64_t(p) * nv + v] = z[p];
  }
  float m1[NP], m2[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    m1[p] = V.m_in[int64_t(p) * nv + vv];
    m2[p] = V.v_in[int64_t(p) * nv + vv];
  }
  if (V.update) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const float dz = opt_step(sp.o, V.es, g[p], m1[p], m2[p]);
      const float zn = z[p] + dz;
      const float dlt = zn - z[p];
      step += dlt * dlt;
      z[p] = zn;
    }
  }
  if (valid) {
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      V.z_out[int64_t(p) * nv + v] = z[p];
      V.m_out[int64_t(p) * nv + v] = m1[p];
      V.v_out[int64_t(p) * nv + v] = m2[p];
    }
    for (int p = NP; p < md.n_par; ++p) {
      const float zd = V.z_in[int64_t(p) * nv + v];
      V.z_out[int64_t(p) * nv + v] = zd;
      V.m_out[int64_t(p) * nv + v] = 0.f;
      V.v_out[int64_t(p) * nv + v] = 0.f;
      if (V.z_last != nullptr) V.z_last[int64_t(p) * nv + v] = zd;
    }
  }

  // block sums of this iteration
  if (!valid) { lsum = 0.f; step = 0.f; den = 0.f; gn = 0.f; reg = 0.f; msum = 0.f; }
  const float t4 = wave_sum4(lsum, step, den, gn, lane, xrow_addr(lane));
  const float tr = wave_sum(reg);
  const float tm = wave_sum(msum);
  __shared__ float red[kBlock / 64][kStepSums];
  if (lane < 4) red[wave][wave_sum4_index(lane)] = t4;
  if (lane == 0) {
    red[wave][4] = tr;
    red[wave][5] = tm;
  }
  __syncthreads();
  if (threadIdx.x < kStepSums) {
    float t = 0.f;
#pragma unroll
    for (int wv = 0; wv < kBlock / 64; ++wv) t += red[wv][threadIdx.x];
    V.partials[(int64_t(lblk) * V.n_epochs + V.k) * kStepSums + threadIdx.x] = t;
  }
}

// ---- grid neighbours of masked voxels (TV stencil) --------------------------
static __global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t val) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = val;
}

// inv[vox[v]] = v (vox: column-major grid index of masked voxel v, in range)
static __global__ void k_scatter_index(const int64_t* __restrict__ vox, int64_t nv, int32_t* __restrict__ inv) {
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < nv) inv[vox[v]] = int32_t(v);
}

// nbr[2d][v] / nbr[2d+1][v]: G-D index of the +1 / -1 neighbour of voxel v
// along axis d with periodic wrap (circshift), or -1 outside the mask
static __global__ void k_neighbours(const int64_t* __restrict__ vox, int64_t nv, const int32_t* __restrict__ inv,
                                    int32_t X, int32_t Y, int32_t Z, int32_t nd, int32_t* __restrict__ nbr) {
  const int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const int64_t lin = vox[v];
  const int64_t x = lin % X, y = (lin / X) % Y, z = lin / (int64_t(X) * Y);
  const int64_t XY = int64_t(X) * Y;
  for (int d = 0; d < nd; ++d) {
    int64_t fw, bw;
    if (d == 0) {
      fw = ((x + 1) % X) + X * y + XY * z;
      bw = ((x + X - 1) % X) + X * y + XY * z;
    } else if (d == 1) {
      fw = x + X * ((y + 1) % Y) + XY * z;
      bw = x + X * ((y + Y - 1) % Y) + XY * z;
    } else {
      fw = x + X * y + XY * ((z + 1) % Z);
      bw = x + X * y + XY * ((z + Z - 1) % Z);
    }
    nbr[int64_t(2 * d) * nv + v] = inv[fw];
    nbr[int64_t(2 * d + 1) * nv + v] = inv[bw];
  }
}

// ---- deterministic FP64 reduction of block partials ------------------------
// in is [nb][nrows] (each block's row sums contiguous).  The blocks are split
// into `gridDim.x` chunks of `clen`; block (c, t) sums rows
// 64 RL t .. 64 RL (t+1) - 1 over chunk c, RL consecutive rows per lane (one
// 4 RL-byte load) and four wavefronts per row set, and writes
// out[row * gridDim.x + c].  Every row is added in the same order whatever
// RL: wavefront q takes the partials b0 + q, b0 + q + 4, ... and the four
// wavefront sums are added in order.  k_fold_chunks adds the chunk sums of a
// row in order, so the result does not depend on timing.
template <int RL>
__global__ __launch_bounds__(kBlock) void k_reduce_rows(const float* __restrict__ in, int64_t nb, int32_t nrows,
                                                         int64_t clen, double* __restrict__ out) {
  static_assert(RL == 1 || RL == 4, "one or four rows per lane");
  const int c = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int row = (blockIdx.y * 64 + lane) * RL;  // first row of the lane
  const int sub = threadIdx.x >> 6;
  const int64_t b0 = int64_t(c) * clen, b1 = min(nb, b0 + clen);
  double s[RL];
#pragma unroll
  for (int j = 0; j < RL; ++j) s[j] = 0.0;
  if (row < nrows) {  // with RL = 4, nrows is a multiple of 4
    for (int64_t b = b0 + sub; b < b1; b += 4) {
      if constexpr (RL == 4) {
        const float4 q = *reinterpret_cast<const float4*>(in + b * nrows + row);
        s[0] += double(q.x);
        s[1] += double(q.y);
        s[2] += double(q.z);
        s[3] += double(q.w);
      } else {
        s[0] += double(in[b * nrows + row]);
      }
    }
  }
  __shared__ double sh[4][64 * RL];
#pragma unroll
  for (int j = 0; j < RL; ++j) sh[sub][lane * RL + j] = s[j];
  __syncthreads();
  if (sub == 0 && row < nrows) {
#pragma unroll
    for (int j = 0; j < RL; ++j) {
      const int l = lane * RL + j;
      out[int64_t(row + j) * gridDim.x + c] = ((sh[0][l] + sh[1][l]) + sh[2][l]) + sh[3][l];
    }
  }
}

// The chunk sums of rows [0, nrows) added in chunk order, written in the
// per-iteration layout dst[k * kSlots + j] (row k * npv + j; every other
// slot of the K kSlots outputs is 0).
static __global__ __launch_bounds__(kBlock) void k_fold_chunks(const double* __restrict__ sums, int32_t chunks,
                                                                int32_t nrows, int32_t npv, int32_t kslots,
                                                                int32_t nout, double* __restrict__ dst) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout) return;
  const int k = t / kslots, j = t - k * kslots;
  const int r = k * npv + j;
  double acc = 0.0;
  if (j < npv && r < nrows) {
    const double* src = sums + int64_t(r) * chunks;
    for (int c = 0; c < chunks; ++c) acc += src[c];
  }
  dst[t] = acc;
}

// k_reduce_rows with four rows per lane where the row count allows
static inline void launch_reduce_rows(const float* in, int64_t nb, int32_t nrows, int32_t chunks,
                                       double* out, hipStream_t st) {
  if (nrows % 4 == 0)
    hipLaunchKernelGGL(k_reduce_rows<4>, dim3(chunks, (nrows / 4 + 63) / 64), dim3(kBlock), 0, st, in, nb, nrows,
                       int64_t(kReduceChunk), out);
  else
    hipLaunchKernelGGL(k_reduce_rows<1>, dim3(chunks, (nrows + 63) / 64), dim3(kBlock), 0, st, in, nb, nrows,
                       int64_t(kReduceChunk), out);
}

static __global__ __launch_bounds__(kBlock) void k_reduce_rows_d(const double* __restrict__ in, int64_t nb,
                                                           double* __restrict__ out) {
  const int64_t row = blockIdx.x;
  double s = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += kBlock) s += in[row * nb + b];
  s = wave_sum_d(s);
  __shared__ double sh[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlock / 64; ++i) t += sh[i];
    out[row] = t;
  }
}

// ---- single evaluation (operator pair / finalisation) ---------------------
struct EvalParams {
  const float* x;       // [P][nv] theta or z
  const float* y;       // [M][nv] (adjoint / residual modes)
  const float* w;
  const float* ex;
  float* s_out;         // [M][nv] model signal (forward mode)
  float* grad;          // [P][nv] d(scale*sum loss)/dx
  float* residual;      // [M][nv] elementwise loss
  float* resloss;       // [nv] mean elementwise loss per voxel
  double* partial;      // [nblk] block loss sums
  const float* aux;     // [rows][nv] aux rows of models that take them, or null
  int64_t nv;
  int32_t transform;
  int32_t mode;         // 0 forward, 1 loss (+ optional outputs)
  float scale;
};

template <class Model, int LOSS, bool HAS_W>
__global__ __launch_bounds__(kBlock) void k_eval(const gc_model* __restrict__ mdp, const EvalParams ep) {
  const gc_model& md = *mdp;
  constexpr int NP = Model::NP;
  constexpr int MR = Model::MR;
  const int64_t nv = ep.nv;
  const int64_t v = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const bool valid = v < nv;
  const int64_t vv = valid ? v : (nv - 1);
  float ex[kMaxExtra];
#pragma unroll
  for (int e = 0; e < kMaxExtra; ++e) ex[e] = (e < md.n_extra) ? ep.ex[int64_t(e) * nv + vv] : 0.f;
  float th[NP], dth[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p)
    transform_fwd(ep.transform, ep.x[int64_t(p) * nv + vv], md.lb[p], md.ub[p], th[p], dth[p]);
  double lsum = 0.0;
  const AuxRef aux{ep.aux ? ep.aux + vv : nullptr, nv};
  if (ep.mode == 0) {
    auto put = [&](int r, float s, const float*) {
      if (valid) ep.s_out[int64_t(Model::meas(md, r)) * nv + v] = isfinite(s) ? s : 0.f;
    };
    if constexpr (has_aux<Model>::value) Model::template eval<false>(md, th, ex, put, aux);
    else Model::template eval<false>(md, th, ex, put);
  } else {
    float gth[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) gth[p] = 0.f;
    float lvox = 0.f;
    model_adjoint<Model>(
        md, th, ex,
        [&](int r, float s) -> float {
          const int64_t m = Model::meas(md, r);
          const bool fin = isfinite(s);
          const float sv = fin ? s : 0.f;
          const float wv = HAS_W ? ep.w[m * nv + vv] : 1.f;
          float l, dl;
          loss_fn<LOSS>(sv - ep.y[m * nv + vv], wv, l, dl);
          lvox += l;
          if (ep.residual != nullptr && valid) ep.residual[m * nv + v] = l;
          return fin ? dl : 0.f;
        },
        gth, aux);
    lsum = valid ? double(lvox) : 0.0;
    if (valid && ep.resloss != nullptr) ep.resloss[v] = lvox / float(md.n_meas);
    if (valid && ep.grad != nullptr) {
#pragma unroll
      for (int p = 0; p < NP; ++p) ep.grad[int64_t(p) * nv + v] = gth[p] * dth[p] * ep.scale;
      for (int p = NP; p < md.n_par; ++p) ep.grad[int64_t(p) * nv + v] = 0.f;
    }
    if (ep.partial != nullptr) {
      lsum = wave_sum_d(lsum);
      __shared__ double sh[kBlock / 64];
      if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = lsum;
      __syncthreads();
      if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kBlock / 64; ++i) t += sh[i];
        ep.partial[blockIdx.x] = t;
      }
    }
  }
}

// ---- posterior statistics of res2out (utils/mcmc.m:823-857) ---------------
// One workgroup sorts the n samples of `vpb` consecutive voxels of one
// variable in LDS (bitonic network over n2 = next power of two >= n, padded
// with +inf; the voxels' samples are read as vpb-wide segments), then takes
// MATLAB's prctile order statistics (median = prctile 50; iqr = prctile 75
// - prctile 25; sample k sits at 100 (k - 0.5) / n) and the mode of
// discretize(x, edges) (bins [e_i, e_i+1), the last closed; the smallest bin
// among the most frequent) from binary searches of the sorted samples.
struct PostStatParams {
  const float* post;     // [Q][n][nv]
  const double* edges;   // [Q][nbin]
  float* median;         // [Q][nv] or null
  float* iqr;
  float* mode;
  int64_t nv;
  int32_t n, n2, vpb, nbin;
};

GC_DEV double prctile_sorted(const float* s, int n, double p) {
  const double k = p / 100.0 * n + 0.5;  // 1-based position
  if (k <= 1.0) return double(s[0]);
  if (k >= double(n)) return double(s[n - 1]);
  const int i = int(floor(k));
  const double f = k - double(i);
  return double(s[i - 1]) + f * (double(s[i]) - double(s[i - 1]));
}

// first index in s[0, n) with s[i] >= x (or > x when upper)
GC_DEV int bsearch_sorted(const float* s, int n, double x, bool upper) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const double v = double(s[mid]);
    if (upper ? (v <= x) : (v < x)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

static __global__ __launch_bounds__(kBlock) void k_posterior_stats(const PostStatParams sp) {
  extern __shared__ float buf[];  // [vpb][n2]
  const int q = blockIdx.y;
  const int64_t v0 = int64_t(blockIdx.x) * sp.vpb;
  const int n = sp.n, n2 = sp.n2, vpb = sp.vpb;
  const int total = vpb * n2;
  for (int t = threadIdx.x; t < total; t += kBlock) {
    const int j = t % vpb, i = t / vpb;  // voxel fastest: coalesced reads
    const int64_t v = v0 + j;
    float x = INFINITY;
    if (i < n && v < sp.nv) x = sp.post[(int64_t(q) * n + i) * sp.nv + v];
    buf[j * n2 + i] = x;
  }
  __syncthreads();
  for (int k = 2; k <= n2; k <<= 1) {
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int t = threadIdx.x; t < total; t += kBlock) {
        const int vj = t / n2, i = t % n2;
        const int l = i ^ jj;
        if (l > i) {
          float* a = buf + vj * n2;
          const float x = a[i], y = a[l];
          const bool up = (i & k) == 0;
          if ((x > y) == up) { a[i] = y; a[l] = x; }
        }
      }
      __syncthreads();
    }
  }
  // order statistics: one thread per voxel
  if (threadIdx.x < vpb) {
    const int64_t v = v0 + threadIdx.x;
    if (v < sp.nv) {
      const float* a = buf + threadIdx.x * n2;
      if (sp.median) sp.median[int64_t(q) * sp.nv + v] = float(prctile_sorted(a, n, 50.0));
      if (sp.iqr) sp.iqr[int64_t(q) * sp.nv + v] = float(prctile_sorted(a, n, 75.0) - prctile_sorted(a, n, 25.0));
    }
  }
  if (sp.mode == nullptr) return;
  // mode: threads of a voxel split its bins; best (count, first index)
  const int tpv = kBlock / vpb;
  const int vj = threadIdx.x / tpv, lane_v = threadIdx.x % tpv;
  const double* e = sp.edges + int64_t(q) * sp.nbin;
  const int nb = sp.nbin - 1;
  int best = -1, bestc = -1;
  if (vj < vpb) {
    const float* a = buf + vj * n2;
    for (int b = lane_v; b < nb; b += tpv) {
      const int lo = bsearch_sorted(a, n, e[b], false);
      const int hi = (b == nb - 1) ? bsearch_sorted(a, n, e[b + 1], true) : bsearch_sorted(a, n, e[b + 1], false);
      const int c = hi - lo;
      if (c > bestc) { bestc = c; best = b; }
    }
  }
  __shared__ int rc[kBlock], rb[kBlock];
  rc[threadIdx.x] = bestc;
  rb[threadIdx.x] = best;
  __syncthreads();
  if (vj < vpb && lane_v == 0) {
    for (int t = 1; t < tpv; ++t) {
      const int c = rc[threadIdx.x + t], b = rb[threadIdx.x + t];
      if (c > bestc || (c == bestc && b < best)) { bestc = c; best = b; }
    }
    const int64_t v = v0 + vj;
    if (v < sp.nv) sp.mode[int64_t(q) * sp.nv + v] = (bestc > 0) ? float(0.5 * (e[best] + e[best + 1])) : NAN;
  }
}

// ---- MCMC -------------------------------------------------------------------
struct McmcParams {
  const float* y;
  const float* w;
  const float* ex;
  const float* x0;
  const float* normals;
  const float* uniforms;
  const int32_t* partners;
  float* posterior;
  float* mean;
  float* stdv;
  float* accept;
  float* xfinal;
  uint8_t* trace;
  int64_t nv;
  uint64_t seed;
  int32_t iteration, thinning, n_burnin, rep, n_walker, n_sample, rng, ens, q, noise;
  float step_size, start_range;
  float x_step[GC_MAX_PAR];
};

GC_DEV bool keep_sample(const McmcParams& p, int k) {
  return k > p.n_burnin && ((k - p.n_burnin + 1) % p.thinning) == 0;  // utils/mcmc.m:170
}

// MATLAB min(x,1) ignoring NaN (utils/mcmc.m:160)
GC_DEV float min1(float x) { return isnan(x) ? 1.f : fminf(x, 1.f); }

// logP_Gaussian, utils/logP_Gaussian.m:18-22
// (hardware reciprocal and log2: -sos / (2 n^2) - nm/2 ln(2) log2(2 pi n^2))
GC_DEV float logp_gaussian(float sos, float noise, int nm) {
  const float n2 = noise * noise;
  return (-0.5f * sos) * rcpf(n2) - (0.5f * kLn2F * float(nm)) * log2v(2.f * kPiF * n2);
}

template <class Model, bool HAS_W>
GC_DEV float model_sos(const gc_model& md, const float* th, const float* ex, const float* yr, const float* wr) {
  float acc = 0.f;
  Model::template eval<false>(md, th, ex, [&](int r, float s, const float*) {
    const float d = s - yr[r];
    acc += (HAS_W ? wr[HAS_W ? r : 0] : 1.f) * d * d;
  });
  return acc;
}

// Group reduction of per-chain (sum, sumsq) into posterior mean / std for
// the voxel owning `gsz` consecutive threads of the block.
template <int Q>
GC_DEV void group_stats(const McmcParams& P, const double* s1, const double* s2, int gsz, int64_t v, bool active,
                        const float* xref, double n_total, double* sh, int bs) {
  // sh: [Q][2][bs]
  for (int q = 0; q < Q; ++q) {
    sh[(q * 2 + 0) * bs + threadIdx.x] = active ? s1[q] : 0.0;
    sh[(q * 2 + 1) * bs + threadIdx.x] = active ? s2[q] : 0.0;
  }
  __syncthreads();
  const int g0 = (threadIdx.x / gsz) * gsz;
  if (active && threadIdx.x == g0) {
    for (int q = 0; q < Q; ++q) {
      double a = 0.0, b = 0.0;
      for (int i = 0; i < gsz; ++i) {
        a += sh[(q * 2 + 0) * bs + g0 + i];
        b += sh[(q * 2 + 1) * bs + g0 + i];
      }
      if (P.mean) P.mean[int64_t(q) * P.nv + v] = float(double(xref[q]) + a / n_total);
      if (P.stdv) {
        const double var = (n_total > 1) ? (b - a * a / n_total) / (n_total - 1) : 0.0;
        P.stdv[int64_t(q) * P.nv + v] = float(sqrt(fmax(var, 0.0)));
      }
    }
  }
}

// Dynamic LDS of the samplers for a block of bs threads: the per-lane FP64
// sums of group_stats [2Q][bs], the acceptance counts [bs] and, for a
// Goodman-Weare ensemble that spans waves, the walker exchange [Q][bs].
constexpr size_t mcmc_lds_bytes(int Q, int bs, bool exchange) {
  return size_t(2) * Q * bs * sizeof(double) + size_t(bs) * sizeof(int) +
         (exchange ? size_t(Q) * bs * sizeof(float) : 0);
}

// A voxel's walker group fits in one wavefront (and never straddles two):
// walker positions are exchanged by cross-lane permutes, with no LDS and no
// block barrier.
__host__ __device__ inline bool gw_wave_local(int gsz) { return gsz <= 64 && (64 % gsz) == 0; }

// Metropolis-Hastings, utils/mcmc.m:82-190.  Thread layout: voxel-major,
// the `rep` chains (repetitions) of a voxel in consecutive lanes.  The
// block size (64, 128 or 256 threads) is a launch choice; results do not
// depend on it.
template <class Model, bool HAS_W>
__global__ __launch_bounds__(kBlock) void k_mh(const gc_model* __restrict__ mdp, const McmcParams P) {
  const gc_model& md = *mdp;
  constexpr int NP = Model::NP;
  constexpr int Q = NP + 1;
  constexpr int MR = Model::MR;
  const int R = P.rep;
  const int bs = blockDim.x;
  const int gpb = bs / R;  // voxels per block
  const int tl = threadIdx.x;
  const int gi = tl / R, r = tl % R;
  const int64_t v = int64_t(blockIdx.x) * gpb + gi;
  const bool active = (gi < gpb) && (v < P.nv);
  const int64_t nv = P.nv;
  const int64_t vv = active ? v : 0;

  float yr[MR];
  float wr[HAS_W ? MR : 1];
#pragma unroll
  for (int rr = 0; rr < MR; ++rr) {
    if (Model::valid(md, rr)) {
      const int64_t m = Model::meas(md, rr);
      yr[rr] = P.y[m * nv + vv];
      if (HAS_W) wr[HAS_W ? rr : 0] = P.w[m * nv + vv];
    } else {
      yr[rr] = 0.f;
      if (HAS_W) wr[HAS_W ? rr : 0] = 0.f;
    }
  }
  float ex[kMaxExtra];
#pragma unroll
  for (int e = 0; e < kMaxExtra; ++e) ex[e] = (e < md.n_extra) ? P.ex[int64_t(e) * nv + vv] : 0.f;
  float x0[Q], x[Q], lbq[Q], ubq[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    lbq[q] = md.lb[q];
    ubq[q] = md.ub[q];
    x0[q] = fminf(fmaxf(P.x0[int64_t(q) * nv + vv], lbq[q]), ubq[q]);  // :126-128
    x[q] = x0[q];
  }
  const int M = md.n_meas;
  const float lp0 = logp_gaussian(model_sos<Model, HAS_W>(md, x0, ex, yr, wr), x0[Q - 1], M);  // :129
  float lp = lp0;
  double s1[Q], s2[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) { s1[q] = 0.0; s2[q] = 0.0; }
  int counter = 0;
  int nacc = 0;
  const uint64_t chain = uint64_t(r) * uint64_t(nv) + uint64_t(vv);
  for (int k = 1; k <= P.iteration; ++k) {
    const int64_t step = int64_t(r) * P.iteration + (k - 1);
    float n[Q];
    float u_acc = 0.f;
    if (P.rng == GC_RNG_REPLAY) {
#pragma unroll
      for (int q = 0; q < Q; ++q) n[q] = P.normals[(step * nv + vv) * Q + q];
    } else {
      // native stream: uniforms u[0..Q] from Philox groups of four; normals
      // in Box-Muller pairs, an odd last one from the normal quantile, and
      // the acceptance uniform u[Q]
      constexpr int NU = Q + 1;
      float uu[4 * ((NU + 3) / 4)];
#pragma unroll
      for (int g4 = 0; g4 < (NU + 3) / 4; ++g4) {
        uint32_t o[4];
        draw4(P.seed, uint32_t(k), chain, PURPOSE_NORMAL0 + g4, o);
#pragma unroll
        for (int i = 0; i < 4; ++i) uu[4 * g4 + i] = u01(o[i]);
      }
#pragma unroll
      for (int q = 0; q + 1 < Q; q += 2) box_muller(uu[q], uu[q + 1], n[q], n[q + 1]);
      if constexpr (Q % 2 == 1) n[Q - 1] = norm_inv(uu[Q - 1]);
      u_acc = uu[Q];
    }
    float xp[Q];
    bool oob = false;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      xp[q] = x[q] + P.x_step[q] * n[q];                   // :147
      oob = oob || (xp[q] < lbq[q]) || (xp[q] > ubq[q]);  // :149
      xp[q] = fminf(fmaxf(xp[q], lbq[q]), ubq[q]);         // :151
    }
    const float lpp = logp_gaussian(model_sos<Model, HAS_W>(md, xp, ex, yr, wr), xp[Q - 1], M);  // :158
    const float u = (P.rng == GC_RNG_REPLAY) ? P.uniforms[step * nv + vv] : u_acc;
    const float ratio = min1(__expf(lpp - lp));  // :160
    const bool acc = (ratio > u) && !oob;        // :161-162
    if (P.trace != nullptr && active) P.trace[step * nv + v] = acc ? 1 : 0;
    if (acc) {
#pragma unroll
      for (int q = 0; q < Q; ++q) x[q] = xp[q];
      lp = lpp;
      ++nacc;
    }
    if (keep_sample(P, k)) {  // :170-173
      if (counter < P.n_sample) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (P.posterior != nullptr && active)
            P.posterior[((int64_t(q) * R + r) * P.n_sample + counter) * nv + v] = x[q];
          const double d = double(x[q]) - double(x0[q]);
          s1[q] += d;
          s2[q] += d * d;
        }
      }
      ++counter;
    }
  }
  const int filled = min(counter, P.n_sample);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    // unfilled posterior slots are zeros in the reference buffer (:121)
    const double d = -double(x0[q]);
    s1[q] += d * double(P.n_sample - filled);
    s2[q] += d * d * double(P.n_sample - filled);
    if (P.xfinal != nullptr && active) P.xfinal[(int64_t(r) * Q + q) * nv + v] = x[q];
  }
  extern __shared__ double mh_lds[];  // mcmc_lds_bytes(Q, bs, false)
  double* sh = mh_lds;
  int* shacc = reinterpret_cast<int*>(mh_lds + 2 * Q * bs);
  shacc[tl] = active ? nacc : 0;
  group_stats<Q>(P, s1, s2, R, v, active && gi < gpb, x0, double(P.n_sample) * R, sh, bs);
  if (active && r == 0 && P.accept != nullptr) {
    int t = 0;
    for (int i = 0; i < R; ++i) t += shacc[tl + i];
    P.accept[v] = float(double(t) / (double(P.iteration) * R));
  }
}

// Goodman-Weare affine-invariant ensemble, utils/mcmc.m:192-383.  A group of
// rep*W consecutive threads owns one voxel: walker w of repetition r at
// group lane r*W + w.  Walker positions are exchanged through LDS.
// minimum waves per SIMD the GW sampler is compiled for (a register budget)
template <class Model>
struct gw_waves {
  template <class M>
  static constexpr int f(decltype(M::kGwWaves)*) { return M::kGwWaves; }
  template <class M>
  static constexpr int f(...) { return 1; }
  static constexpr int value = f<Model>(nullptr);
};

template <class Model, bool HAS_W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(gw_waves<Model>::value)))
void k_gw(const gc_model* __restrict__ mdp, const McmcParams P) {
  const gc_model& md = *mdp;
  constexpr int NP = Model::NP;
  constexpr int Q = NP + 1;
  constexpr int MR = Model::MR;
  const int W = P.n_walker, R = P.rep;
  const int gsz = W * R;
  const int bs = blockDim.x;
  const int gpb = bs / gsz;
  const bool wave_local = gw_wave_local(gsz);
  const int tl = threadIdx.x;
  const int gi = tl / gsz, gl = tl % gsz;
  const int r = gl / W, wk = gl % W;
  const int64_t v = int64_t(blockIdx.x) * gpb + gi;
  const bool active = (gi < gpb) && (v < P.nv);
  const int64_t nv = P.nv;
  const int64_t vv = active ? v : 0;
  const bool rb = P.ens == GC_ENSEMBLE_REDBLACK;
  const int h = W / 2;
  const float inv_a = 1.f / P.step_size;  // the stretch z = t^2 / a (:301-304)

  float yr[MR];
  float wr[HAS_W ? MR : 1];
#pragma unroll
  for (int rr = 0; rr < MR; ++rr) {
    if (Model::valid(md, rr)) {
      const int64_t m = Model::meas(md, rr);
      yr[rr] = P.y[m * nv + vv];
      if (HAS_W) wr[HAS_W ? rr : 0] = P.w[m * nv + vv];
    } else {
      yr[rr] = 0.f;
      if (HAS_W) wr[HAS_W ? rr : 0] = 0.f;
    }
  }
  float ex[kMaxExtra];
#pragma unroll
  for (int e = 0; e < kMaxExtra; ++e) ex[e] = (e < md.n_extra) ? P.ex[int64_t(e) * nv + vv] : 0.f;
  float xc0[Q], x[Q], lbq[Q], ubq[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    lbq[q] = md.lb[q];
    ubq[q] = md.ub[q];
    xc0[q] = P.x0[int64_t(q) * nv + vv];
  }
  // walker initialisation x0 + (ub-lb)*startRange*randn, clamped (:270-273)
  {
    const uint64_t ichain = uint64_t(wk) * uint64_t(nv) + uint64_t(vv);
    float n[Q];
    if (P.rng == GC_RNG_REPLAY) {
#pragma unroll
      for (int q = 0; q < Q; ++q) n[q] = P.normals[(int64_t(wk) * nv + vv) * Q + q];
    } else {
#pragma unroll
      for (int g4 = 0; g4 < (Q + 3) / 4; ++g4) {
        uint32_t o[4];
        draw4(P.seed, 0u, ichain, PURPOSE_INIT + g4, o);
        float a0, a1, a2, a3;
        box_muller(u01(o[0]), u01(o[1]), a0, a1);
        if (4 * g4 + 2 < Q) box_muller(u01(o[2]), u01(o[3]), a2, a3);
        if (4 * g4 + 0 < Q) n[4 * g4 + 0] = a0;
        if (4 * g4 + 1 < Q) n[4 * g4 + 1] = a1;
        if (4 * g4 + 2 < Q) n[4 * g4 + 2] = a2;
        if (4 * g4 + 3 < Q) n[4 * g4 + 3] = a3;
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
      x[q] = fminf(fmaxf(xc0[q] + (ubq[q] - lbq[q]) * P.start_range * n[q], lbq[q]), ubq[q]);
  }
  const int M = md.n_meas;
  float lp = logp_gaussian(model_sos<Model, HAS_W>(md, x, ex, yr, wr), x[Q - 1], M);  // :275

  extern __shared__ double gw_lds[];  // mcmc_lds_bytes(Q, bs, !wave_local)
  double* sh = gw_lds;
  int* shacc = reinterpret_cast<int*>(gw_lds + 2 * Q * bs);
  float* shx = reinterpret_cast<float*>(shacc + bs);  // [Q][bs], exchange across waves
  double s1[Q], s2[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) { s1[q] = 0.0; s2[q] = 0.0; }
  int counter = 0, nacc = 0;
  const uint64_t chain = (uint64_t(r) * W + wk) * uint64_t(nv) + uint64_t(vv);
  const int base = gi * gsz + r * W;  // LDS index of walker 0 of this (voxel, rep)
  const int hm = (h > W - h) ? h : (W - h);
  for (int k = 1; k <= P.iteration; ++k) {
    const int64_t step = int64_t(r) * P.iteration + (k - 1);
    const int32_t* part = P.partners + step * W;
    const int nsub = rb ? 2 : 1;
    for (int sub = 0; sub < nsub; ++sub) {
      if (!wave_local) {
#pragma unroll
        for (int q = 0; q < Q; ++q) shx[q * bs + tl] = x[q];
        __syncthreads();
      }
      bool mine = true;
      int j = wk, pw;
      if (rb) {
        mine = (sub == 0) ? (wk < h) : (wk >= h);
        j = (sub == 0) ? wk : wk - h;
        pw = mine ? part[((sub == 0) ? 0 : h) + j] : 0;
      } else {
        pw = part[wk];
      }
      float u1 = 0.f, u2 = 0.f;
      if (mine) {
        if (P.rng == GC_RNG_REPLAY) {
          int64_t ub0, ust;
          if (rb) { ub0 = (step * 2 + sub) * 2 * hm; ust = hm; }
          else { ub0 = step * 2 * W; ust = W; }
          u1 = P.uniforms[(ub0 + j) * nv + vv];
          u2 = P.uniforms[(ub0 + ust + j) * nv + vv];
        } else {
          // one draw: stretch and acceptance uniforms
          uint32_t o[4];
          draw4(P.seed, uint32_t(k), chain, PURPOSE_ZZ + sub, o);
          u1 = u01(o[0]);
          u2 = u01(o[1]);
        }
      }
      // stretch move (:299-304)
      const float t = (P.step_size - 1.f) * u1 + 1.f;
      const float zz = t * t * inv_a;
      float xp[Q];
      bool oob = false;
      // partner walker pw of this (voxel, repetition): lane base + pw
      const int src = wave_local ? ((tl & 63) - gl + r * W + pw) : 0;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const float xa = wave_local ? __shfl(x[q], src, 64) : shx[q * bs + base + pw];
        float val = xa + (x[q] - xa) * zz;
        oob = oob || (val < lbq[q]) || (val > ubq[q]);
        xp[q] = fminf(fmaxf(val, lbq[q]), ubq[q]);
      }
      if (!wave_local) __syncthreads();
      if (mine) {
        const float lpp = logp_gaussian(model_sos<Model, HAS_W>(md, xp, ex, yr, wr), xp[Q - 1], M);
        // min(zz.^(Nvar-1).*exp(logPProposed-logPCurr),1) > rand (:313-315)
        float zpow = 1.f;
#pragma unroll
        for (int q = 0; q < Q - 1; ++q) zpow *= zz;
        const float ratio = min1(zpow * __expf(lpp - lp));
        const bool acc = (ratio > u2) && !oob;
        if (P.trace != nullptr && active) P.trace[(step * W + wk) * nv + v] = acc ? 1 : 0;
        if (acc) {
#pragma unroll
          for (int q = 0; q < Q; ++q) x[q] = xp[q];
          lp = lpp;
          ++nacc;
        }
      }
    }
    if (keep_sample(P, k)) {  // :363-366
      if (counter < P.n_sample) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          if (P.posterior != nullptr && active)
            P.posterior[(((int64_t(q) * R + r) * P.n_sample + counter) * W + wk) * nv + v] = x[q];
          const double d = double(x[q]) - double(xc0[q]);
          s1[q] += d;
          s2[q] += d * d;
        }
      }
      ++counter;
    }
  }
  const int filled = min(counter, P.n_sample);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const double d = -double(xc0[q]);
    s1[q] += d * double(P.n_sample - filled);
    s2[q] += d * d * double(P.n_sample - filled);
    if (P.xfinal != nullptr && active) P.xfinal[((int64_t(r) * W + wk) * Q + q) * nv + v] = x[q];
  }
  shacc[tl] = active ? nacc : 0;
  group_stats<Q>(P, s1, s2, gsz, v, active, xc0, double(P.n_sample) * R * W, sh, bs);
  if (active && gl == 0 && P.accept != nullptr) {
    int t = 0;
    for (int i = 0; i < gsz; ++i) t += shacc[tl + i];
    P.accept[v] = float(double(t) / (double(P.iteration) * R * W));
  }
}

// ---- layout / transform kernels ---------------------------------------------
static __global__ void k_nd_to_gd(const float* __restrict__ nd, int64_t nvox, int32_t M, const int64_t* __restrict__ vox,
                           int64_t nv, float* __restrict__ gd) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nv * M) return;
  const int64_t m = i / nv, v = i % nv;
  gd[i] = nd[vox[v] + nvox * m];
}

static __global__ void k_clamp01(float* __restrict__ p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = fminf(fmaxf(p[i], 0.f), 1.f);
}

static __global__ void k_fill(float* __restrict__ p, int64_t n, float val) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = val;
}

static __global__ void k_gd_to_nd(const float* __restrict__ gd, int64_t nv, int32_t M, const int64_t* __restrict__ vox,
                           int64_t nvox, float* __restrict__ nd) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nv * M) return;
  const int64_t m = i / nv, v = i % nv;
  nd[vox[v] + nvox * m] = gd[i];
}

// theta -> z (askadam.initialise_parameter / rescale_parameters, :368-412, :1077-1089)
static __global__ void k_rescale(const gc_model* __restrict__ mdp, int32_t transform, const float* __restrict__ th,
                          float* __restrict__ z, int64_t nv) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const gc_model& md = *mdp;
  if (i >= nv * md.n_par) return;
  const int p = int(i / nv);
  const float lb = md.lb[p], ub = md.ub[p];
  const float x = th[i];
  if (transform == GC_TRANSFORM_SIGMOID) {
    const float eb = 1e-4f * (ub - lb);
    const float c = fmaxf(fminf(x, ub - eb), lb + eb);
    const float t = (c - lb) / (ub - lb);
    z[i] = logf(t / (1.f - t));
  } else if (transform == GC_TRANSFORM_LINEAR) {
    z[i] = (x - lb) / (ub - lb);
  } else {
    z[i] = x;
  }
}

// z -> theta (unscale_parameters, :1063-1073)
static __global__ void k_unscale(const gc_model* __restrict__ mdp, int32_t transform, const float* __restrict__ z,
                          float* __restrict__ th, int64_t nv) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const gc_model& md = *mdp;
  if (i >= nv * md.n_par) return;
  const int p = int(i / nv);
  const float lb = md.lb[p], ub = md.ub[p];
  if (transform == GC_TRANSFORM_SIGMOID) th[i] = lb + (ub - lb) * (1.f / (1.f + expf(-z[i])));
  else if (transform == GC_TRANSFORM_LINEAR) th[i] = z[i] * (ub - lb) + lb;
  else th[i] = z[i];
}

// ---- spatial total variation (utils/spatial_total_variation.m:24-66) --------
// One thread per grid voxel; masked voxels add their forward-difference
// terms and scatter-free gather the adjoint: the gradient at voxel i is
// sum_d (1/v_d)[m(i-e_d) sgn(D_d(i-e_d)) - m(i) sgn(D_d(i))] * lambda/N.
static __global__ __launch_bounds__(kBlock) void k_tv(const float* __restrict__ img, const uint8_t* __restrict__ mask,
                                                int32_t X, int32_t Y, int32_t Z, int32_t nd, float ivx, float ivy,
                                                float ivz, float scale, double* __restrict__ partial,
                                                float* __restrict__ grad) {
  const int64_t n = int64_t(X) * Y * Z;
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  double acc = 0.0;
  if (i < n) {
    const int x = int(i % X);
    const int yy = int((i / X) % Y);
    const int zz = int(i / (int64_t(X) * Y));
    const float c = img[i];
    const float iv[3] = {ivx, ivy, ivz};
    float g = 0.f;
    for (int d = 0; d < nd; ++d) {
      int64_t jp, jm;
      if (d == 0) {
        jp = ((x + 1) % X) + int64_t(X) * (yy + int64_t(Y) * zz);
        jm = ((x + X - 1) % X) + int64_t(X) * (yy + int64_t(Y) * zz);
      } else if (d == 1) {
        jp = x + int64_t(X) * (((yy + 1) % Y) + int64_t(Y) * zz);
        jm = x + int64_t(X) * (((yy + Y - 1) % Y) + int64_t(Y) * zz);
      } else {
        jp = x + int64_t(X) * (yy + int64_t(Y) * ((zz + 1) % Z));
        jm = x + int64_t(X) * (yy + int64_t(Y) * ((zz + Z - 1) % Z));
      }
      const float dp = (img[jp] - c) * iv[d];
      if (mask[i]) {
        acc += fabs(double(dp));
        g -= sgnf(dp) * iv[d];
      }
      if (mask[jm]) {
        const float dm = (c - img[jm]) * iv[d];
        g += sgnf(dm) * iv[d];
      }
    }
    if (grad != nullptr) grad[i] = g * scale;
  }
  acc = wave_sum_d(acc);
  __shared__ double sh[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) t += sh[k];
    partial[blockIdx.x] = t;
  }
}

}  // namespace gc
