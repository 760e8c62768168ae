// gc_models.hpp -- device forward models with hand-written Jacobians.
//
// Every model exposes the same interface to the solver kernels:
//   NP                     parameters the model reads (slots 0..NP-1)
//   MR                     register slots for a voxel's measurements
//   valid(md, r), meas(md, r)
//                          whether register slot r is used and which G-D
//                          measurement row it holds
//   eval<JAC>(md, th, ex, f)
//                          calls f(r, s, J) for every used slot r, with the
//                          model signal s and (JAC) J[p] = ds/dtheta_p.
// Loops run over compile-time bounds with runtime guards so that, after
// inlining, every slot index is a constant and a voxel's measurements stay
// in registers.  Acquisition tables are read from the device copy of the
// gc_model at wave-uniform addresses (scalar loads into SGPRs).
#pragma once
#include "gc_device.hpp"
#include "../../include/gacelle.h"

namespace gc {

// ---------------------------------------------------------------------------
// exp(-TE_0 R2*) and exp(-dTE R2*) of equally spaced echo times (TE_0 = first
// echo, dTE = f[8]).  When dTE = k TE_0 for k = 1 or 2 (i[7] = k, set on
// upload with f[9] = -log2(e) TE_0) the spacing's factor is a power of the
// first echo's, else it is a second exponential.  The branch is wave-uniform.
GC_DEV void uniform_decays(const gc_model& md, float r2s, float& e0, float& q) {
  constexpr float kNegLog2e = -1.4426950408889634f;
  if (md.i[7] > 0) {
    e0 = __builtin_amdgcn_exp2f(r2s * md.f[9]);
    q = md.i[7] == 2 ? e0 * e0 : e0;
  } else {
    e0 = __builtin_amdgcn_exp2f(r2s * (kNegLog2e * md.table[0]));
    q = __builtin_amdgcn_exp2f(r2s * (kNegLog2e * md.f[8]));
  }
}

// ---------------------------------------------------------------------------
// s = m0 * exp(-te * r2s): R2star/gpuR2starMapping.m:427-434 (FWD :321-363)
// theta = {M0, R2star}.  EXACT: the echo count equals MT (no guards).
// With GC_FLAG_UNIFORM_TE the decays exp(-TE_k R2*) are formed as
// exp(-TE_0 R2*) * exp(-dTE R2*)^k (two exponentials per voxel instead of
// one per echo; relative error <= k ulp).
template <int MT, bool EXACT = false, bool UNI = false>
struct R2sModel {
  static constexpr int NP = 2, MR = MT, NE = 0;
  static constexpr bool kRelJ = true;  // J[1] is relative to J[0] (see accum_grad)
  static constexpr bool kFusedAdj = false, kStreamOnly = false;
  GC_DEV static bool valid(const gc_model& md, int r) { return EXACT ? (r < MT) : (r < md.i[0]); }
  GC_DEV static int meas(const gc_model&, int r) { return r; }
  template <bool JAC, class F>
  GC_DEV static void eval(const gc_model& md, const float* th, const float*, F&& f) {
    const int nte = EXACT ? MT : md.i[0];
    const float m0 = th[0], r2s = th[1];
    constexpr float kNegLog2e = -1.4426950408889634f;
    const float kr = kNegLog2e * r2s;  // -log2(e) * R2*
    float e = 0.f, q = 0.f;
    if (UNI) uniform_decays(md, r2s, e, q);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      if (EXACT || t < nte) {
        const float te = md.table[t];
        float et;
        if (UNI) {
          et = e;
          e *= q;
        } else {
          et = __builtin_amdgcn_exp2f(kr * te);
        }
        const float s = m0 * et;
        float J[NP];
        if (JAC) {
          J[0] = et;   // d s / d M0 >= 0
          J[1] = -te;  // (d s / d R2*) / J[0] without the M0 factor (post_grad)
        }
        f(t, s, J);
      }
    }
  }
  // the M0 factor common to every echo's d s / d R2*, applied once
  GC_DEV static void post_grad(const float* th, float* g) { g[1] *= th[0]; }
};

// ---------------------------------------------------------------------------
// Joint VFA-SPGR x multi-echo, R1R2s/gpuJointR1R2starMapping.m:513-525:
//   s = m0*sin(a)*(1-E1)/(1-cos(a)*E1)*exp(-te*r2s), E1 = exp(-TR*R1),
//   a = B1*FA (:384-391) or the supplied true flip angle.
// theta = {M0, R1, R2star}; register slot t + MT*f holds measurement t + nte*f.
// EXACT: the echo and flip-angle counts equal MT and MF (no guards).
template <int MT, int MF, bool EXACT = false>
struct JointModel {
  // blocked askadam kernel within 128 VGPRs (some bytes per lane spilled);
  // three waves without spill measured 44.0 and five waves (more spill)
  // 49.3 against 42.5 ms on 2.9M voxels
  static constexpr int kAskWaves = 4;
  static constexpr int NP = 3, MR = MT * MF, NE = 1;
  static constexpr bool kRelJ = false;
  static constexpr bool kFusedAdj = false, kStreamOnly = false;
  // the blocked kernel's adjoint: one group per flip angle, base = exp(-te R2*)
  static constexpr bool kBaseAdj = true;
  GC_DEV static void post_grad(const float*, float*) {}
  GC_DEV static bool valid(const gc_model& md, int r) { return EXACT || ((r % MT) < md.i[0] && (r / MT) < md.i[1]); }
  GC_DEV static int meas(const gc_model& md, int r) { return (r % MT) + (EXACT ? MT : md.i[0]) * (r / MT); }

  // exp(-te R2*) of every echo; with GC_FLAG_UNIFORM_TE by recurrence from
  // two exponentials (relative error <= t ulp at echo t)
  GC_DEV static void decays(const gc_model& md, float r2s, float* ete) {
    const int nte = EXACT ? MT : md.i[0];
    if (md.flags & GC_FLAG_UNIFORM_TE) {
      float e, q;
      uniform_decays(md, r2s, e, q);
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        ete[t] = (EXACT || t < nte) ? e : 0.f;
        e *= q;
      }
    } else {
#pragma unroll
      for (int t = 0; t < MT; ++t) ete[t] = (EXACT || t < nte) ? __expf(-md.table[t] * r2s) : 0.f;
    }
  }
  // steady-state amplitude per unit M0, g = sin(a)(1 - E1)/(1 - cos(a) E1),
  // and its R1 derivative per unit M0, at flip angle fi
  GC_DEV static void fa_terms(const gc_model& md, const float* ex, int fi, float e1, float& g, float& dg_r1) {
    const int nte = EXACT ? MT : md.i[0];
    const float tr = md.f[0];
    const float alpha = (md.flags & GC_FLAG_R1R2S_TRUE_FA) ? ex[fi] : ex[0] * md.table[nte + fi];
    float sa, ca;
    __sincosf(alpha, &sa, &ca);
    const float den = 1.f - ca * e1;
    const float rden = rcpf(den);
    g = sa * (1.f - e1) * rden;
    // d g / d R1 = sa*TR*E1*(1-ca)/den^2
    dg_r1 = sa * tr * e1 * (1.f - ca) * rden * rden;
  }

  template <bool JAC, class F>
  GC_DEV static void eval(const gc_model& md, const float* th, const float* ex, F&& f) {
    const int nte = EXACT ? MT : md.i[0], nfa = EXACT ? MF : md.i[1];
    const float tr = md.f[0];
    const float m0 = th[0], r1 = th[1], r2s = th[2];
    float ete[MT];
    decays(md, r2s, ete);
    const float e1 = __expf(-tr * r1);
#pragma unroll
    for (int fi = 0; fi < MF; ++fi) {
      if (EXACT || fi < nfa) {
        float g, dg_r1;
        fa_terms(md, ex, fi, e1, g, dg_r1);
        const float amp = m0 * g;
        const float damp_r1 = m0 * dg_r1;
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          if (EXACT || t < nte) {
            const float s = amp * ete[t];
            float J[NP];
            if (JAC) {
              J[0] = g * ete[t];
              J[1] = damp_r1 * ete[t];
              J[2] = -md.table[t] * s;
            }
            f(t + MT * fi, s, J);
          }
        }
      }
    }
  }

  // The blocked kernel's per-voxel constants (voxel_prep, formed once per
  // launch): sin and cos of each flip angle, which depend on the voxel's B1
  // or true flip angle only.
  static constexpr int kVoxPrep = 2 * MF;
  GC_DEV static void voxel_prep(const gc_model& md, const float* ex, float* vp) {
    const int nte = EXACT ? MT : md.i[0], nfa = EXACT ? MF : md.i[1];
#pragma unroll
    for (int fi = 0; fi < MF; ++fi) {
      vp[2 * fi] = 0.f;
      vp[2 * fi + 1] = 0.f;
      if (EXACT || fi < nfa) {
        const float alpha = (md.flags & GC_FLAG_R1R2S_TRUE_FA) ? ex[fi] : ex[0] * md.table[nte + fi];
        __sincosf(alpha, &vp[2 * fi], &vp[2 * fi + 1]);
      }
    }
  }

  // d s / d M0 = g ete, d s / d R1 = M0 dg_r1 ete, d s / d R2* = -te M0 g ete:
  // per flip angle the sums A = sum dl ete and T = sum dl ete te carry the
  // whole gradient (alpha = {g, M0 dg_r1, 0}, beta = {0, 0, -M0 g}).
  // vp = voxel_prep's sines and cosines; fa_terms's g and dg_r1 with the
  // reciprocals of the flip angles taken together.
  template <class F>
  GC_DEV static void eval_base_adj(const gc_model& md, const float* th, const float* vp, F&& f, float* gth) {
    const int nte = EXACT ? MT : md.i[0], nfa = EXACT ? MF : md.i[1];
    const float tr = md.f[0];
    const float m0 = th[0], r1 = th[1], r2s = th[2];
    float ete[MT];
    decays(md, r2s, ete);
    const float e1 = __expf(-tr * r1);
    float den[MF], rden[MF];
#pragma unroll
    for (int fi = 0; fi < MF; ++fi) den[fi] = (EXACT || fi < nfa) ? 1.f - vp[2 * fi + 1] * e1 : 1.f;
    rcp_batch<MF>(den, rden);
#pragma unroll
    for (int fi = 0; fi < MF; ++fi) {
      if (EXACT || fi < nfa) {
        const float sa = vp[2 * fi], ca = vp[2 * fi + 1];
        const float g = sa * (1.f - e1) * rden[fi];
        const float dg_r1 = sa * tr * e1 * (1.f - ca) * rden[fi] * rden[fi];
        const float amp = m0 * g;
        float A = -0.f, T = -0.f;
#pragma unroll
        for (int t = 0; t < MT; ++t) {
          if (EXACT || t < nte) {
            const float a = f(t + MT * fi, amp * ete[t], ete[t]);
            A += a;
            T = fmaf(a, md.table[t], T);
          }
        }
        gth[0] = fmaf(g, A, gth[0]);
        gth[1] = fmaf(m0 * dg_r1, A, gth[1]);
        gth[2] = fmaf(-amp, T, gth[2]);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// NEXI spherical mean (NEXI/gpuNEXI.m:902-996, integrand :1090-1097), with
// optional Rician first-moment correction (NEXI/gpuNEXIrice.m:635-660,
// 726-755).  Quadrature S = sum_j w_j M(x_j) with nodes/weights from the
// table (the trapezoid of gacelle_trapz.m:217-226 is a fixed weight set).
// theta = {fa, Da, De, ra[, p2]}; slot k = Sl0 of shell k, MS + k = Sl2.
template <int MS, bool L2>
struct NexiModel {
  static constexpr int NP = L2 ? 5 : 4, MR = L2 ? 2 * MS : MS, NE = 1;
  // GW sampler within 128 VGPRs (104 bytes per lane spilled); three waves
  // without spill measured 17.9 against 16.5 s on the 500k-voxel ensemble
  static constexpr int kGwWaves = 4;
  static constexpr bool kRelJ = false;
  static constexpr bool kFusedAdj = false, kStreamOnly = false;
  GC_DEV static void post_grad(const float*, float*) {}
  GC_DEV static bool valid(const gc_model& md, int r) { return (r % MS) < md.i[0]; }
  GC_DEV static int meas(const gc_model& md, int r) { return (r < MS) ? r : md.i[0] + (r - MS); }

  // L_1/2(x) and dL/dx with the 34-node Bessel trapezoid (utils/rician.m:44-56)
  GC_DEV static void laguerre_half(const gc_model& md, float x, float& L, float& dL) {
    const int ns = md.i[0], nn = md.i[2];
    const float* cost = md.table + 2 * ns + 2 * nn;
    bool clamped = false;
    if (x < -170.f) { x = -170.f; clamped = true; }
    // exp(z c) = 2^(z log2(e) c), z = -x/2
    const float z2 = -0.7213475204444817f * x;
    float i0 = 0.f, i1 = 0.f, d1 = 0.f;
#pragma unroll
    for (int k = 0; k < 34; ++k) {
      const float c = cost[k];
      const float wk = (k == 0 || k == 33) ? 0.5f : 1.f;
      const float e = __builtin_amdgcn_exp2f(z2 * c) * wk;
      const float ec = e * c;
      i0 += e;
      i1 += ec;
      d1 = fmaf(ec, c, d1);
    }
    const float h = 1.f / 33.f;  // (pi/33) / pi
    i0 *= h; i1 *= h; d1 *= h;
    const float ex2 = __builtin_amdgcn_exp2f(0.7213475204444817f * x);  // exp(x/2)
    const float core = (1.f - x) * i0 - x * i1;
    L = ex2 * core;
    // dI/dx = dI/dz * (-1/2); the derivative of the c-weighted integral is i1 again
    const float dcore = -i0 + (1.f - x) * i1 * (-0.5f) - i1 - x * d1 * (-0.5f);
    dL = clamped ? 0.f : (0.5f * L + ex2 * dcore);
  }

  // The Karger integrand of one shell at node x (NEXI/gpuNEXI.m:1090-1097):
  //   M(x) = Pp E1 + (1 - Pp) E2,  E1 = exp(-(l1 + l2)),  E2 = exp(-(l1 - l2)),
  //   l1 = (r1 + r2 + d1 + d2)/2,  l2 = sqrt((r1 - r2 + d1 - d2)^2 + 4 r1 r2)/2
  //   (at least 1e-8),  Pp = (fa d1 + (1 - fa) d2 - (l1 - l2)) / (2 l2),
  // with d1 = b Da x^2, d2 = b De, r1 = Delta ra, r2 = r1 fa / (1 - fa).
  // The x-independent parts are folded once per shell, so a node costs 14
  // VALU operations besides one rsq and two exp2.
  struct Shell {
    float dab;  // d1 / x^2 = b Da
    float cu;   // r1 - r2 - d2
    float c4;   // 4 r1 r2
    float fam;  // (fa - 1/2) b Da
    float cn;   // (1 - fa) d2 - (r1 + r2 + d2) / 2
    float as0;  // -log2(e) (r1 + r2 + d2) / 2
    float asd;  // -log2(e) b Da / 2
  };
  static constexpr float kHalfLog2e = 0.7213475204444817f;
  GC_DEV static Shell shell(float fa, float dab, float d2, float r1, float r2) {
    const float cs = r1 + r2 + d2;
    Shell K;
    K.dab = dab;
    K.cu = r1 - r2 - d2;
    K.c4 = 4.f * r1 * r2;
    K.fam = (fa - 0.5f) * dab;
    K.cn = (1.f - fa) * d2 - 0.5f * cs;
    K.as0 = -kHalfLog2e * cs;
    K.asd = -kHalfLog2e * dab;
    return K;
  }
  // node values: 2 l2 (clamped), 1/(2 l2), Pp, E1, E2, M and the u = r1 - r2 + d1 - d2
  // and clamp flag the derivatives need
  struct Node {
    float u, sqh, rq, Pp, E1, E2, M;
    bool clamp;
  };
  GC_DEV static Node node(const Shell& K, float x2) {
    Node n;
    n.u = fmaf(K.dab, x2, K.cu);
    // 2 l2 = sqrt(v) and its reciprocal from one v_rsq_f32: sqrt(v) = v rsq(v)
    const float v = fmaf(n.u, n.u, K.c4);
    const float t = __builtin_amdgcn_rsqf(v);
    const float sq = v * t;
    n.clamp = !(sq >= 2e-8f);             // l2 < 1e-8 (v = 0 gives 0 * inf)
    n.sqh = n.clamp ? 2e-8f : sq;         // 2 l2
    n.rq = n.clamp ? (1.f / 2e-8f) : t;   // 1 / (2 l2)
    n.Pp = fmaf(0.5f, n.sqh, fmaf(K.fam, x2, K.cn)) * n.rq;
    const float as = fmaf(K.asd, x2, K.as0);  // -log2(e) l1
    n.E1 = __builtin_amdgcn_exp2f(fmaf(-kHalfLog2e, n.sqh, as));
    n.E2 = __builtin_amdgcn_exp2f(fmaf(kHalfLog2e, n.sqh, as));
    n.M = fmaf(n.Pp, n.E1 - n.E2, n.E2);
    return n;
  }

  template <bool JAC, class F>
  GC_DEV static void eval(const gc_model& md, const float* th, const float* ex, F&& f) {
    const int ns = md.i[0], nn = md.i[2];
    const bool guard = (md.flags & GC_FLAG_NEXI_RE_GUARD) != 0;
    const float fa_raw = th[0];
    const float fa = fmaxf(fa_raw, md.f[0]);
    const float dfa = (fa_raw >= md.f[0]) ? 1.f : 0.f;
    const float Da = th[1], De = th[2], ra = th[3];
    const float omf = 1.f - fa;
    const float den_re = guard ? fmaxf(omf, 1e-8f) : omf;
    const float rden_re = 1.f / den_re;
    const float* xs = md.table + 2 * ns;
    const float* ws = md.table + 2 * ns + nn;
    // Rician mean: -1/(2 sigma^2) and sigma sqrt(pi/2) of this voxel
    float nis2 = 0.f, c0 = 0.f;
    if (md.flags & GC_FLAG_NEXI_RICIAN) {
      const float sigma = ex[0];
      nis2 = -0.5f / (sigma * sigma);
      c0 = sigma * 1.2533141373155003f;
    }
#pragma unroll
    for (int k = 0; k < MS; ++k) {
      if (k < ns) {
        const float b = md.table[k], D = md.table[ns + k];
        const float dab = b * Da, deb = b * De, rab = D * ra;
        const float re = rab * fa * rden_re;
        // d re / d fa (through both numerator and 1/(1-fa))
        const float dre_dfa = (guard && omf < 1e-8f) ? rab * rden_re : rab * rden_re * rden_re;
        const float dre_dra = D * fa * rden_re;
        const Shell K = shell(fa, dab, deb, rab, re);
        float s0 = 0.f, s2 = 0.f;
        float gA0 = 0, gD10 = 0, gD20 = 0, gR10 = 0, gR20 = 0;  // Sl0 partial sums
        float gA2 = 0, gD12 = 0, gD22 = 0, gR12 = 0, gR22 = 0;  // Sl2 partial sums
#pragma unroll 2
        for (int j = 0; j < nn; ++j) {
          const float x = xs[j], w = ws[j];
          const float x2 = x * x;
          const Node n = node(K, x2);
          const float w2 = w * (1.5f * x2 - 0.5f);
          s0 = fmaf(w, n.M, s0);
          if (L2) s2 = fmaf(w2, n.M, s2);
          if (JAC) {
            const float d1 = dab * x2, d2 = deb, r1 = rab, r2 = re;
            const float rl2 = 2.f * n.rq;  // 1 / l2
            // partials of l2 with respect to d1, d2, r1, r2
            const float c = n.clamp ? 0.f : 0.5f * n.rq;
            const float u = n.u;
            const float l2_d1 = c * u, l2_d2 = -c * u, l2_r1 = c * (u + 2.f * r2), l2_r2 = c * (2.f * r1 - u);
            const float Pp = n.Pp, E1 = n.E1, E2 = n.E2;
            const float EmE = E1 - E2;
            // dM/dz = dPp/dz*(E1-E2) - Pp*E1*(dl1+dl2) - (1-Pp)*E2*(dl1-dl2)
            // dPp/dz = (dnum/dz)/(2 l2) - Pp*dl2/dz / l2 ; dnum = da*(d1-d2) + a dd1 + (1-a) dd2 - dlm
            auto dM = [&](float dnum_extra, float dl2) {
              const float dl1 = 0.5f;
              const float dlm = dl1 - dl2;
              const float dPp = 0.5f * (dnum_extra - dlm) * rl2 - Pp * dl2 * rl2;
              return dPp * EmE - Pp * E1 * (dl1 + dl2) - (1.f - Pp) * E2 * dlm;
            };
            const float mA = 0.5f * (d1 - d2) * rl2 * EmE;
            const float mD1 = dM(fa, l2_d1);
            const float mD2 = dM(1.f - fa, l2_d2);
            const float mR1 = dM(0.f, l2_r1);
            const float mR2 = dM(0.f, l2_r2);
            gA0 += w * mA; gD10 += w * mD1 * x2; gD20 += w * mD2; gR10 += w * mR1; gR20 += w * mR2;
            if (L2) { gA2 += w2 * mA; gD12 += w2 * mD1 * x2; gD22 += w2 * mD2; gR12 += w2 * mR1; gR22 += w2 * mR2; }
          }
        }
        // chain to theta = {fa, Da, De, ra}
        float J0[NP], J2[NP];
        if (JAC) {
          J0[0] = (gA0 + gR20 * dre_dfa) * dfa;
          J0[1] = gD10 * b;
          J0[2] = gD20 * b;
          J0[3] = gR10 * D + gR20 * dre_dra;
          if (L2) J0[4] = 0.f;
        }
        if (md.flags & GC_FLAG_NEXI_RICIAN) {
          const float xr = (s0 * s0) * nis2;  // -s0^2 / (2 sigma^2)
          float L, dL;
          laguerre_half(md, xr, L, dL);
          const float m1 = c0 * L;
          if (m1 >= s0) {
            const float ds = c0 * dL * (2.f * s0 * nis2);  // d xr / d s0 = -s0 / sigma^2
            if (JAC) {
#pragma unroll
              for (int p = 0; p < NP; ++p) J0[p] *= ds;
            }
            s0 = m1;
          }
        }
        if (s0 > 1.f) {
          s0 = 1.f;
          if (JAC) {
#pragma unroll
            for (int p = 0; p < NP; ++p) J0[p] = 0.f;
          }
        }
        f(k, s0, J0);
        if (L2) {
          const float p2 = th[4];
          const float sg = sgnf(s2);
          float sl2 = p2 * fabsf(s2);
          if (JAC) {
            J2[0] = p2 * sg * (gA2 + gR22 * dre_dfa) * dfa;
            J2[1] = p2 * sg * gD12 * b;
            J2[2] = p2 * sg * gD22 * b;
            J2[3] = p2 * sg * (gR12 * D + gR22 * dre_dra);
            J2[4] = fabsf(s2);
          }
          if (sl2 > 1.f) {
            sl2 = 1.f;
            if (JAC) {
#pragma unroll
              for (int p = 0; p < NP; ++p) J2[p] = 0.f;
            }
          }
          f(MS + k, sl2, J2);
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Forward-mode dual number with N derivative slots (device side).
template <int N>
struct DualF {
  float v;
  float d[N];
  GC_DEV static DualF cst(float x) {
    DualF r;
    r.v = x;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = 0.f;
    return r;
  }
  GC_DEV static DualF var(float x, int k) {
    DualF r = cst(x);
    r.d[k] = 1.f;
    return r;
  }
};
template <int N>
GC_DEV DualF<N> operator+(const DualF<N>& a, const DualF<N>& b) {
  DualF<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
GC_DEV DualF<N> operator-(const DualF<N>& a, const DualF<N>& b) {
  DualF<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
GC_DEV DualF<N> operator*(const DualF<N>& a, const DualF<N>& b) {
  DualF<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
GC_DEV DualF<N> operator/(const DualF<N>& a, const DualF<N>& b) {
  DualF<N> r;
  const float ib = rcpf(b.v);
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N>
GC_DEV DualF<N> operator*(float a, const DualF<N>& b) {
  DualF<N> r;
  r.v = a * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a * b.d[i];
  return r;
}
template <int N>
GC_DEV DualF<N> operator+(float a, const DualF<N>& b) {
  DualF<N> r = b;
  r.v = a + b.v;
  return r;
}
template <int N>
GC_DEV DualF<N> operator-(float a, const DualF<N>& b) {
  DualF<N> r;
  r.v = a - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -b.d[i];
  return r;
}
template <int N>
GC_DEV DualF<N> operator-(const DualF<N>& a) {
  return 0.f - a;
}
template <int N>
GC_DEV DualF<N> dsqrt(const DualF<N>& a) {
  DualF<N> r;
  r.v = sqrtv(a.v);
  const float h = (r.v > 0.f) ? 0.5f * rcpf(r.v) : 0.f;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h;
  return r;
}
template <int N>
GC_DEV DualF<N> dexp(const DualF<N>& a) {
  DualF<N> r;
  r.v = __expf(a.v);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * r.v;
  return r;
}
// max(a, c) with the derivative of the selected branch
template <int N>
GC_DEV DualF<N> dmax(const DualF<N>& a, float c) {
  return (a.v >= c) ? a : DualF<N>::cst(c);
}

// ---------------------------------------------------------------------------
// MCR-MWI three-pool GRE model, MCRMWI/gpuMCRMWI.m:566-714 with the DIMWI
// terms fitted (no HCFM), isEPG = false, and compute_gremwi_signal
// (MCRMWI/private/compute_gremwi_signal.m:1-15).  T1 steady state per flip
// angle: Bloch-McConnell two-pool analytical solution (model_BM_2T1_analytical,
// :992-1023) when EXCH, else the non-exchanging Bloch model (:1026-1032).
//
// theta = {S0, MWF, IWF, R1IEW, [kIEWM], R2sMW, R2sIW, R2sEW, freqMW, freqIW,
//          [dfreqBKG x MF, dpini]}; NA = 4 + EXCH amplitude parameters first.
// Measurement r = t + nte*fa (real part or magnitude), + nte*nfa for the
// imaginary part when CPLX.  The adjoint is fused: the amplitude parameters'
// derivatives come from duals over the per-flip-angle steady state (NA
// slots), every per-echo term has closed-form derivatives, and the gradient
// is accumulated from the loss derivatives the caller returns per
// measurement.  Data are streamed (kStreamOnly): M = 2 x 32 x MF is too
// large to keep in registers.
// Per-voxel rows computed by a pre-pass (stride = voxels of the volume)
struct AuxRef {
  const float* p = nullptr;
  int64_t stride = 0;
  GC_DEV float operator[](int k) const { return as_global(p)[int64_t(k) * stride]; }
};

// EPG-X path (EPG, needs EXCH; MCRMWI/gpuMCRMWI.m:664-681): the MLP
// surrogate's outputs at each flip angle come from the k_epgx pre-pass as
// 15 aux rows per flip angle, [k * n_fa + fi]: the IEW phase, the IEW and MW
// steady-state corrections, each followed by its derivatives with respect
// to (MVF, R1IEW, kIEWM, R2sIW).
constexpr int kEpgxRows = 15;

template <int MF, bool EXCH, bool CPLX, bool EPG = false>
struct McrmwiModel {
  static_assert(!EPG || EXCH, "the EPG-X path needs the exchange model");
  static constexpr bool kAux = EPG;
  static constexpr int kStepWaves = 3;  // step kernel within 168 VGPRs
  static constexpr int NA = EXCH ? 5 : 4;
  static constexpr int NP = NA + 5 + (CPLX ? MF + 1 : 0);
  static constexpr int MT = 32;
  static constexpr int MR = MT * MF * (CPLX ? 2 : 1);
  static constexpr int NE = 3 + MF;
  static constexpr bool kRelJ = false;
  static constexpr bool kFusedAdj = true, kStreamOnly = true;
  // The order of eval_adj's loss calls, which the step kernel reads ahead
  // by: flip angle by flip angle, echo by echo, the real part (call
  // t + nte fi) then, with complex data, the imaginary part (call
  // t + nte fi + nte nfa); every call made, none skipped.
  static constexpr int kMeasStreams = CPLX ? 2 : 1;
  using D = DualF<NA>;
  GC_DEV static void post_grad(const float*, float*) {}
  GC_DEV static bool valid(const gc_model& md, int r) { return r < md.n_meas; }
  GC_DEV static int meas(const gc_model&, int r) { return r; }

  // compartment amplitudes (MW, IW, EW) at flip angle index fi; with EPG
  // also the IEW phase ph and the R2sIW derivatives (r2[0..2] of the
  // amplitudes, r2[3] of ph), which the NA duals do not carry
  GC_DEV static void amplitudes(const gc_model& md, const float* th, const float* ex, int fi, D& aMW, D& aIW,
                                D& aEW, AuxRef aux = {}, D* ph = nullptr, float* r2 = nullptr) {
    const int nte = md.i[0];
    const float TR = md.f[0], rho = md.f[3], t1mw = md.f[4];
    const D S0 = D::var(th[0], 0), mwf = D::var(th[1], 1), iwf = D::var(th[2], 2), r1iew = D::var(th[3], 3);
    const D S0MW = S0 * mwf;
    const D S0IW = S0 * (1.f - mwf) * iwf;
    const D S0EW = S0 * (1.f - mwf) * (1.f - iwf);
    const D tv = dmax(S0IW + S0EW + (1.f / rho) * S0MW, 1e-8f);
    const float famp = md.table[nte + fi] * ex[0];
    float sa, ca;
    sincos_rev(famp * kInv2PiF, sa, ca);
    D aIEW;
    if constexpr (EXCH) {
      const D kfr = D::var(th[4], 4);
      const D f = dmax(((1.f / rho) * S0MW) / tv, 1e-8f);  // MVF
      const float R1r = 1.f / t1mw;
      const D a = -r1iew - kfr;
      const D b = (1.f - f) * kfr / f;
      const D c = kfr;
      const D d = -R1r - b;
      const D apd = a + d;
      const D det = a * d - b * c;
      const D disc = dsqrt(apd * apd - 4.f * det);
      const D l1 = 0.5f * (apd + disc), l2 = 0.5f * (apd - disc);
      const D E1 = dexp(TR * l1), E2 = dexp(TR * l2);
      const D dl = l1 - l2;
      const D e11 = (E1 * (a - l2) - E2 * (a - l1)) / dl;
      const D e12 = b * (E1 - E2) / dl;
      const D e21 = c * (E1 - E2) / dl;
      const D e22 = (E1 * (d - l2) - E2 * (d - l1)) / dl;
      const D C = D::cst(sa) / ((1.f - ca * (e22 + e11) + (ca * ca) * (e11 * e22 - e21 * e12)) * det);
      const D ap = (1.f - ca * e22) * (e11 - D::cst(1.f)) + ca * (e12 * e21);
      const D bp = (1.f - ca * e22) * e12 + ca * e12 * (e22 - D::cst(1.f));
      const D cp = ca * e21 * (e11 - D::cst(1.f)) + (1.f - ca * e11) * e21;
      const D dp = ca * (e21 * e12) + (1.f - ca * e11) * (e22 - D::cst(1.f));
      const D siew = ((ap * d - bp * c) * r1iew * (1.f - f) + R1r * (bp * a - ap * b) * f) * C;
      const D sm = ((cp * d - c * dp) * r1iew * (1.f - f) + R1r * (a * dp - cp * b) * f) * C;
      if constexpr (EPG) {
        // MLP corrections, chained to the NA directions through MVF (f),
        // R1IEW (slot 3) and kIEWM (slot 4); clamped at askadam.epsilon
        auto row = [&](int k) { return aux[k * md.i[1] + fi]; };
        auto lift = [&](int k) {
          D o = D::cst(row(k));
#pragma unroll
          for (int q = 0; q < NA; ++q) o.d[q] = row(k + 1) * f.d[q];
          o.d[3] += row(k + 2);
          o.d[4] += row(k + 3);
          return o;
        };
        *ph = lift(0);
        const D dIEW = lift(5), dMW = lift(10);
        const D mw = rho * ((sm + dMW) * tv);
        const D iew = (siew + dIEW) * tv;
        const bool mw_on = mw.v >= 1e-8f, iew_on = iew.v >= 1e-8f;
        aMW = dmax(mw, 1e-8f);
        aIEW = dmax(iew, 1e-8f);
        const float tvr = tv.v;
        r2[0] = mw_on ? rho * tvr * row(14) : 0.f;
        const float riew = iew_on ? tvr * row(9) : 0.f;
        r2[1] = riew * th[2];
        r2[2] = riew * (1.f - th[2]);
        r2[3] = row(4);
      } else {
        aMW = rho * (sm * tv);
        aIEW = siew * tv;
      }
    } else {
      const float e1mw = __expf(-TR / t1mw);
      aMW = (sa * (1.f - e1mw) / (1.f - ca * e1mw)) * S0MW;
      const D e1iew = dexp(-TR * r1iew);
      aIEW = (S0IW + S0EW) * (sa * (1.f - e1iew)) / (1.f - ca * e1iew);
    }
    aIW = aIEW * iwf;
    aEW = aIEW * (1.f - iwf);
  }

  // Visit the echoes of flip angle fi with each pool's complex factor
  // w_p = exp(-te r2s_p) (cos + i sin)(phase_p(te)); g(t, te, wr[3], wi[3]).
  // With equally spaced echoes w_p is advanced by recurrence (one complex
  // multiply per pool and echo).
  template <class G>
  GC_DEV static void per_echo(const gc_model& md, const float* th, const float* ex, int fi, G&& g, float ph = 0.f) {
    const int nte = md.i[0];
    const float B0g = md.f[1] * md.f[2];
    const float freqEW = md.f[5];
    const float r2s[3] = {th[NA], th[NA + 1], th[NA + 2]};
    const float freq[3] = {th[NA + 3], th[NA + 4], freqEW};
    float fbkg = 0.f, pini = 0.f;
    if constexpr (CPLX) {
      // one initial phase per flip angle (rows 3+MF.., n_fa == MF) or one per
      // voxel; uniform selects over the MF slots keep the indices static
      const bool pfa = (md.flags & GC_FLAG_MCRMWI_PINI_FA) != 0;
      float pex = ex[1];
#pragma unroll
      for (int j = 0; j < MF; ++j)
        if (j == fi) {
          fbkg = th[NA + 5 + j] + ex[3 + j];
          if (pfa) pex = ex[3 + MF + j];
        }
      pini = th[NA + 5 + MF] + pex;
    }
    // phases in revolutions: frequencies (freq + fbkg) B0 gamma in Hz, the
    // initial and IEW phases divided by 2 pi
    float om[3], wr[3], wi[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) om[p] = (freq[p] + fbkg) * B0g;
    const float pinr = pini * kInv2PiF, phr = ph * kInv2PiF;
    if (md.flags & GC_FLAG_UNIFORM_TE) {
      // a loop of its own: two echoes per trip, so the step kernel's
      // read-ahead queue (two values per stream) returns to the same
      // registers every trip
      const float te0 = md.table[0], dte = md.f[8];
      float qr[3], qi[3];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        float e, c, sn;
        e = __expf(-te0 * r2s[p]);
        sincos_rev(te0 * om[p] + pinr + (p > 0 ? phr : 0.f), sn, c);
        wr[p] = e * c;
        wi[p] = e * sn;
        e = __expf(-dte * r2s[p]);
        sincos_rev(dte * om[p], sn, c);
        qr[p] = e * c;
        qi[p] = e * sn;
      }
#pragma unroll 2
      for (int t = 0; t < nte; ++t) {
        g(t, md.table[t], wr, wi);
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const float r2 = wr[p] * qr[p] - wi[p] * qi[p];
          const float i2 = wr[p] * qi[p] + wi[p] * qr[p];
          wr[p] = r2;
          wi[p] = i2;
        }
      }
    } else {
      for (int t = 0; t < nte; ++t) {
        const float te = md.table[t];
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          float c, sn;
          const float e = __expf(-te * r2s[p]);
          sincos_rev(te * om[p] + pinr + (p > 0 ? phr : 0.f), sn, c);
          wr[p] = e * c;
          wi[p] = e * sn;
        }
        g(t, te, wr, wi);
      }
    }
  }

  template <bool JAC, class F>
  GC_DEV static void eval(const gc_model& md, const float* th, const float* ex, F&& f, AuxRef aux = {}) {
    static_assert(!JAC, "McrmwiModel provides its derivatives through eval_adj");
    const int nte = md.i[0], nfa = md.i[1];
    const float ff = ex[2];
    for (int fi = 0; fi < nfa; ++fi) {
      D aMW, aIW, aEW, ph = D::cst(0.f);
      float r2[4];
      amplitudes(md, th, ex, fi, aMW, aIW, aEW, aux, &ph, r2);
      // the per-voxel scale folded into the amplitudes
      const float A[3] = {ff * aMW.v, ff * aIW.v, ff * aEW.v};
      per_echo(md, th, ex, fi, [&](int t, float, const float* wr, const float* wi) {
        const float re = A[0] * wr[0] + A[1] * wr[1] + A[2] * wr[2];
        const float im = A[0] * wi[0] + A[1] * wi[1] + A[2] * wi[2];
        if constexpr (CPLX) {
          f(t + nte * fi, re, nullptr);
          f(t + nte * fi + nte * nfa, im, nullptr);
        } else {
          f(t + nte * fi, sqrtv(re * re + im * im), nullptr);
        }
      }, ph.v);
    }
  }

  // dloss(r, s) returns d loss / d s_r (0 when the measurement carries no
  // gradient); g[NP] accumulates d loss / d theta.
  //
  // With S = sum_p A_p w_p (A_p = ff a_p) and R = dre + i dim the loss
  // gradient at one echo: d/dA_p = Re(conj(R) w_p) =: u_p, d/dr2s_p =
  // -te A_p u_p, and d/dphase_p = A_p Im(conj(R) w_p) =: A_p v_p.  The echo
  // loop accumulates only the amplitude-free sums of u_p, te u_p, v_p and
  // te v_p; the amplitudes multiply them once per flip angle.
  template <class F>
  GC_DEV static void eval_adj(const gc_model& md, const float* th, const float* ex, F&& dloss, float* g,
                              AuxRef aux = {}) {
    const int nte = md.i[0], nfa = md.i[1];
    const float ff = ex[2];
    const float B0g = md.f[1] * md.f[2];
    constexpr float k2pi = 6.283185307179586f;
    for (int fi = 0; fi < nfa; ++fi) {
      D aMW, aIW, aEW, ph = D::cst(0.f);
      float r2[4] = {0.f, 0.f, 0.f, 0.f};
      amplitudes(md, th, ex, fi, aMW, aIW, aEW, aux, &ph, r2);
      const float A[3] = {ff * aMW.v, ff * aIW.v, ff * aEW.v};
      float su[3] = {0.f, 0.f, 0.f}, stu[3] = {0.f, 0.f, 0.f};  // sum u_p, sum te u_p
      float stv[3] = {0.f, 0.f, 0.f};                           // sum te v_p
      // sum A_p v_p over all pools (pini) and over IW and EW (the EPG phase)
      float sav = 0.f, sav12 = 0.f;
      per_echo(md, th, ex, fi, [&](int t, float te, const float* wr, const float* wi) {
        const float re = A[0] * wr[0] + A[1] * wr[1] + A[2] * wr[2];
        const float im = A[0] * wi[0] + A[1] * wi[1] + A[2] * wi[2];
        float dre, dim;
        if constexpr (CPLX) {
          dre = dloss(t + nte * fi, re);
          dim = dloss(t + nte * fi + nte * nfa, im);
        } else {
          const float mag = sqrtv(re * re + im * im);
          const float dm = dloss(t + nte * fi, mag);
          const float im_ = (mag > 0.f) ? rcpf(mag) : 0.f;
          dre = dm * re * im_;
          dim = dm * im * im_;
        }
#pragma unroll
        for (int p = 0; p < 3; ++p) {
          const float u = dre * wr[p] + dim * wi[p];
          const float v = dim * wr[p] - dre * wi[p];
          su[p] += u;
          stu[p] += te * u;
          stv[p] += te * v;
          if (EPG && p > 0)
            sav12 += A[p] * v;
          else
            sav += A[p] * v;
        }
      }, ph.v);
      // a non-finite amplitude made the signal non-finite at every echo, so
      // every loss derivative was zero and the pool adds nothing (the sums
      // weighted by A_p inside the loop are then NaN: they are dropped)
      const float w0 = k2pi * B0g;
      const bool allfin = isfinite(A[0]) && isfinite(A[1]) && isfinite(A[2]);
      float gamp[3], gbkg = 0.f;
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const bool fin = isfinite(A[p]);
        const float Ap = fin ? A[p] : 0.f;
        gamp[p] = fin ? ff * su[p] : 0.f;                 // d loss / d a_p
        g[NA + p] -= Ap * stu[p];                          // R2s of pool p
        if (p < 2) g[NA + 3 + p] += Ap * w0 * stv[p];      // freqMW, freqIW
        gbkg += Ap * stv[p];
      }
      const float gph = (EPG && allfin) ? sav12 : 0.f;     // IEW phase
      const float gpini = allfin ? sav + sav12 : 0.f;
#pragma unroll
      for (int q = 0; q < NA; ++q)
        g[q] += gamp[0] * aMW.d[q] + gamp[1] * aIW.d[q] + gamp[2] * aEW.d[q] + (EPG ? gph * ph.d[q] : 0.f);
      if constexpr (EPG) g[NA + 1] += gamp[0] * r2[0] + gamp[1] * r2[1] + gamp[2] * r2[2] + gph * r2[3];  // R2sIW
      if constexpr (CPLX) {
        g[NA + 5 + MF] += gpini;                           // dpini
        // (a uniform select over the MF slots: a runtime index would put g in scratch)
#pragma unroll
        for (int j = 0; j < MF; ++j)
          if (j == fi) g[NA + 5 + j] += w0 * gbkg;
      }
    }
  }
};

}  // namespace gc
