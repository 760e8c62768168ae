// oracle/models.hpp -- CPU restatement of GACELLE's forward models.
//
// TEST INFRASTRUCTURE ONLY (the oracle).  Every model is written once as a
// template over the value type V (double, float or a Dual for exact
// derivatives) and follows the cited MATLAB lines term by term.
#pragma once
#include <cmath>
#include <cstdint>

#include "../include/gacelle.h"
#include "dual.hpp"

namespace orc {

using std::cos;
using std::exp;
using std::fabs;
using std::log;
using std::sin;
using std::sqrt;

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Mono-exponential R2*: s = m0 .* exp(-te .* r2s)
// R2star/gpuR2starMapping.m:427-434, private/model_R2s_singlecompartment.m:2-9
template <class V, class S>
void fwd_r2star(const gc_model& md, const V* th, const S* /*extra*/, V* s) {
  const int nte = md.i[0];
  for (int t = 0; t < nte; ++t) s[t] = th[0] * exp(-(S(md.table[t]) * th[1]));
}

// ---------------------------------------------------------------------------
// Joint VFA-SPGR x multi-echo:
// s = m0 .* sin(fa) .* (1-exp(-TR.*R1)) ./ (1-cos(fa).*exp(-TR.*R1)) .* exp(-te.*r2s)
// R1R2s/gpuJointR1R2starMapping.m:513-525; FA = b1 .* deg2rad(fa) (:384-391);
// parameter order {M0,R1,R2star} (:25); measurement order TE fastest (:384-385).
template <class V, class S>
void fwd_joint(const gc_model& md, const V* th, const S* extra, V* s) {
  const int nte = md.i[0], nfa = md.i[1];
  const S tr = S(md.f[0]);
  const V& m0 = th[0];
  const V& r1 = th[1];
  const V& r2s = th[2];
  for (int f = 0; f < nfa; ++f) {
    S alpha;
    if (md.flags & GC_FLAG_R1R2S_TRUE_FA)
      alpha = extra[f];
    else
      alpha = extra[0] * S(md.table[nte + f]);
    const V e1 = exp(-(tr * r1));
    const V amp = m0 * S(std::sin(alpha)) * (S(1) - e1) / (S(1) - S(std::cos(alpha)) * e1);
    for (int t = 0; t < nte; ++t) s[t + nte * f] = amp * exp(-(S(md.table[t]) * r2s));
  }
}

// ---------------------------------------------------------------------------
// NEXI integrand M(x), NEXI/gpuNEXI.m:1090-1097 (== private/NEXI_M.m:12-19)
template <class V, class S>
V nexi_M(S x, const V& fa, const V& Da, const V& d2, const V& r1, const V& r2) {
  const V d1 = Da * (x * x);
  const V l1 = (r1 + r2 + d1 + d2) / S(2);
  V l2 = sqrt((r1 - r2 + d1 - d2) * (r1 - r2 + d1 - d2) + S(4) * r1 * r2) / S(2);
  l2 = vmax(l2, S(1e-8));
  const V lm = l1 - l2;
  const V Pp = (fa * d1 + (S(1) - fa) * d2 - lm) / (l2 * S(2));
  return Pp * exp(-(l1 + l2)) + (S(1) - Pp) * exp(-lm);
}

// Bessel I_nu by the 34-node trapezoid on [0, pi]:
// utils/rician.m:52-56 == utils/gacelle_besseli.m:3-9 == NEXI/gpuNEXIrice.m:751-755
template <class V, class S>
V besseli_trapz(int nu, const V& z) {
  const int nx = 34;
  V acc = V(S(0));
  V prev = V(S(0));
  double tprev = 0;
  for (int k = 0; k < nx; ++k) {
    const double t = kPi * double(k) / double(nx - 1);
    const V f = exp(z * S(std::cos(t))) * S(std::cos(nu * t));
    if (k > 0) acc = acc + (prev + f) * S((t - tprev) / 2.0);
    prev = f;
    tprev = t;
  }
  return acc / S(kPi);
}

// Laguerre L_1/2 with the -170 clamp: utils/rician.m:44-50
template <class V, class S>
V laguerre_half(V x) {
  x = vmax(x, S(-170));
  return exp(x / S(2)) * ((S(1) - x) * besseli_trapz<V, S>(0, -x / S(2)) -
                          x * besseli_trapz<V, S>(1, -x / S(2)));
}

// Rician first moment: max(sigma*sqrt(pi/2)*L12(-S^2/2/sigma^2), S)
// utils/rician.m:40-42 == NEXI/gpuNEXIrice.m:726-737
template <class V, class S>
V rician_mean(const V& sig, S sigma) {
  const V x = -(sig * sig) / S(2) / (sigma * sigma);
  const V m1 = S(sigma * std::sqrt(kPi / 2)) * laguerre_half<V, S>(x);
  return vmax(m1, sig);
}

// NEXI spherical-mean signal, NEXI/gpuNEXI.m:902-996 (FWD -> Slmax2) with
// the Rician variant of NEXI/gpuNEXIrice.m:635-687.  Quadrature: the
// reference's trapezoid over x = linspace(0,1,Nx) (gacelle_trapz.m:217-226)
// or a supplied node/weight table.
template <class V, class S>
void fwd_nexi(const gc_model& md, const V* th, const S* extra, V* s) {
  const int ns = md.i[0], lmax = md.i[1], nn = md.i[2], quad = md.i[3];
  const V fa = vmax(th[0], S(md.f[0]));
  const V& Da = th[1];
  const V& De = th[2];
  const V& ra = th[3];
  for (int k = 0; k < ns; ++k) {
    const S b = S(md.table[k]);
    const S D = S(md.table[ns + k]);
    const V dab = b * Da, deb = b * De, rab = D * ra;
    V re;
    if (md.flags & GC_FLAG_NEXI_RE_GUARD)
      re = rab * fa / vmax(S(1) - fa, S(1e-8));
    else
      re = rab * fa / (S(1) - fa);
    V s0 = V(S(0)), s2 = V(S(0));
    if (quad == GC_QUAD_TRAPZ) {
      V m_prev = V(S(0)), m2_prev = V(S(0));
      S x_prev = 0;
      for (int j = 0; j < nn; ++j) {
        const S x = S(md.table[2 * ns + j]);
        const V m = nexi_M<V, S>(x, fa, dab, deb, rab, re);
        const V m2 = m * ((S(3) * x * x - S(1)) / S(2));
        if (j > 0) {
          s0 = s0 + (m_prev + m) * ((x - x_prev) / S(2));
          s2 = s2 + (m2_prev + m2) * ((x - x_prev) / S(2));
        }
        m_prev = m;
        m2_prev = m2;
        x_prev = x;
      }
    } else {
      for (int j = 0; j < nn; ++j) {
        const S x = S(md.table[2 * ns + j]);
        const S w = S(md.table[2 * ns + nn + j]);
        const V m = nexi_M<V, S>(x, fa, dab, deb, rab, re);
        s0 = s0 + m * w;
        s2 = s2 + m * ((S(3) * x * x - S(1)) / S(2) * w);
      }
    }
    if (md.flags & GC_FLAG_NEXI_RICIAN) s0 = rician_mean<V, S>(s0, extra[0]);
    s[k] = vmin(s0, S(1));
    if (lmax == 2) s[ns + k] = vmin(th[4] * fabs(s2), S(1));
  }
}

// ---------------------------------------------------------------------------
// MCR-MWI two-pool Bloch-McConnell analytical steady state,
// MCRMWI/gpuMCRMWI.m:992-1023 (model_BM_2T1_analytical)
template <class V, class S>
void bm_2t1(S TR, S famp, const V& f, const V& R1f, S R1r, const V& kfr, V& S0M, V& S0IEW) {
  const V a = -R1f - kfr;
  const V b = (S(1) - f) * kfr / f;
  const V c = kfr;
  const V d = -R1r - b;
  const V disc = sqrt((a + d) * (a + d) - S(4) * (a * d - b * c));
  const V l1 = ((a + d) + disc) / S(2);
  const V l2 = ((a + d) - disc) / S(2);
  const V E1 = exp(l1 * TR), E2 = exp(l2 * TR);
  const V dl = l1 - l2;
  const V e11 = (E1 * (a - l2) - E2 * (a - l1)) / dl;
  const V e12 = b * (E1 - E2) / dl;
  const V e21 = c * (E1 - E2) / dl;
  const V e22 = (E1 * (d - l2) - E2 * (d - l1)) / dl;
  const S ca = S(std::cos(famp)), sa = S(std::sin(famp));
  const V C = sa / ((S(1) - (e22 + e11) * ca + (e11 * e22 - e21 * e12) * (ca * ca)) * (a * d - b * c));
  const V ap = (S(1) - e22 * ca) * (e11 - S(1)) + e12 * e21 * ca;
  const V bp = (S(1) - e22 * ca) * e12 + (e12 * ca) * (e22 - S(1));
  const V cp = e21 * ca * (e11 - S(1)) + (S(1) - e11 * ca) * e21;
  const V dp = e21 * ca * e12 + (S(1) - e11 * ca) * (e22 - S(1));
  S0IEW = ((ap * d - bp * c) * R1f * (S(1) - f) + (-ap * b + bp * a) * R1r * f) * C;
  S0M = ((cp * d - c * dp) * R1f * (S(1) - f) + (-cp * b + a * dp) * R1r * f) * C;
}

// feature_preprocess_MCRMWI_MLP_EPGX_leakyrelu.m:24-95: the 11 features of a
// (voxel, flip angle) sample
template <class V, class S>
void epgx_features(const V& fm, const V& t1iew, const V& kiewmw, S fa, S tr, const V& t2iew, S t1mw, V* x) {
  const V kmwiew = (vval(fm) == S(0)) ? V(S(0)) : kiewmw * (S(1) - fm) / fm;   // :59-60
  const S sa = S(std::sin(fa)), ca = S(std::cos(fa));
  x[0] = fm;                                                      // :65
  x[1] = exp(-(tr / t1iew));                                      // :66 E1
  x[2] = exp(-(tr * kiewmw));                                     // :67
  x[3] = V(sa);                                                   // :68
  x[4] = S(1) - x[1] * ca;                                        // :71
  x[5] = exp(-(tr * kmwiew));                                     // :72
  x[6] = exp(-(tr * (S(1) / t1iew + kiewmw)));                    // :73
  x[7] = exp(-(tr * (S(1) / t1mw + kmwiew)));                     // :74
  // model_GRE_steadystate_Bloch (:97-100): m0 (1-E1) sin(a) / (1 - E1 cos(a))
  const V e1 = exp(-(tr / t1iew));
  x[8] = ((S(1) - fm) * (S(1) - e1) * sa) / (S(1) - e1 * ca);    // :77
  const S e1m = S(std::exp(-tr / t1mw));
  x[9] = (fm * (S(1) - e1m) * sa) / (S(1) - e1m * ca);           // :78
  x[10] = exp(-(tr * (S(1) / t2iew)));                            // :81
}

// mlp_model_leakyRelu.m:23-47: fc1, then leaky ReLU and fc_l for the other
// layers.  Weights are column-major [out, in] (MATLAB's own storage),
// followed by the bias, layer after layer.
template <class V, class S>
void mlp_eval(const gc_mlp& net, const V* x, V* out) {
  V a[GC_MLP_MAX_WIDTH], z[GC_MLP_MAX_WIDTH];
  for (int i = 0; i < net.width[0]; ++i) a[i] = x[i];
  const float* p = net.params;
  for (int l = 1; l <= net.n_layer; ++l) {
    const int ni = net.width[l - 1], no = net.width[l];
    if (l > 1)
      for (int i = 0; i < ni; ++i) a[i] = (vval(a[i]) > S(0)) ? a[i] : S(net.alpha) * a[i];  // leakyrelu
    const float* W = p;
    const float* b = p + size_t(no) * ni;
    for (int j = 0; j < no; ++j) {
      V t = V(S(b[j]));
      for (int i = 0; i < ni; ++i) t = t + S(W[j + size_t(no) * i]) * a[i];
      z[j] = t;
    }
    for (int j = 0; j < no; ++j) a[j] = z[j];
    p = b + no;
  }
  for (int j = 0; j < net.width[net.n_layer]; ++j) out[j] = a[j];
}

// Number of model parameters of the MCRMWI slot layout for the given flags.
inline int mcrmwi_npar(const gc_model& md) {
  int p = 10;
  if (!(md.flags & GC_FLAG_MCRMWI_EXCHANGE)) p -= 1;
  if (md.flags & GC_FLAG_MCRMWI_COMPLEX) p += md.i[1] + 1;
  return p;
}

// MCR-MWI forward model, MCRMWI/gpuMCRMWI.m:566-714 with the DIMWI terms
// fitted (no HCFM), the EPG-X MLP, BM or Bloch T1 path, and compute_gremwi_signal
// (private/compute_gremwi_signal.m:1-15).  The flip-angle axis is kept in
// dimension 5 for both T1 paths (the Bloch path's shiftdim at :691 is taken
// to mean the same axis).
template <class V, class S>
void fwd_mcrmwi(const gc_model& md, const V* th, const S* extra, V* s) {
  const int nte = md.i[0], nfa = md.i[1];
  const bool cplx = (md.flags & GC_FLAG_MCRMWI_COMPLEX) != 0;
  const bool exch = (md.flags & GC_FLAG_MCRMWI_EXCHANGE) != 0;
  const S TR = S(md.f[0]), B0 = S(md.f[1]), gyro = S(md.f[2]), rho = S(md.f[3]), t1mw = S(md.f[4]),
          freqEW = S(md.f[5]);
  int k = 0;
  const V S0 = th[k++];
  const V mwf = th[k++];
  const V iwf = th[k++];
  const V r1iew = th[k++];
  V kiewm = V(S(0));
  if (exch) kiewm = th[k++];
  const V r2sMW = th[k++];
  const V r2sIW = th[k++];
  const V r2sEW = th[k++];
  const V freqMW = th[k++];
  const V freqIW = th[k++];
  const int kbkg = k;
  const int kpini = k + nfa;

  const V S0MW = S0 * mwf;
  const V S0IW = S0 * (S(1) - mwf) * iwf;
  const V S0EW = S0 * (S(1) - mwf) * (S(1) - iwf);
  const V tv = vmax(S0IW + S0EW + S0MW / rho, S(1e-8));
  const V mvf = vmax((S0MW / rho) / tv, S(1e-8));

  const S b1 = extra[0];
  const S ff = extra[2];
  for (int fi = 0; fi < nfa; ++fi) {
    const S famp = S(md.table[nte + fi]) * b1;
    V aMW, aIEW, ph = V(S(0));
    if (exch && (md.flags & GC_FLAG_MCRMWI_EPG)) {
      // EPG-X (MCRMWI/gpuMCRMWI.m:664-681): the MLPs' phase of the long-T2
      // pools and their corrections of the BM steady state
      V sm, siew, x[11], sd[2];
      bm_2t1<V, S>(TR, famp, mvf, r1iew, S(1) / t1mw, kiewm, sm, siew);
      epgx_features<V, S>(mvf, S(1) / r1iew, kiewm, famp, TR, S(1) / r2sIW, t1mw, x);
      mlp_eval<V, S>(md.mlp[0], x, &ph);
      mlp_eval<V, S>(md.mlp[1], x, sd);
      aMW = vmax((sm + sd[1]) * tv * rho, S(1e-8));   // :677
      aIEW = vmax((siew + sd[0]) * tv, S(1e-8));      // :678
    } else if (exch) {
      V sm, siew;
      bm_2t1<V, S>(TR, famp, mvf, r1iew, S(1) / t1mw, kiewm, sm, siew);
      aMW = sm * tv * rho;
      aIEW = siew * tv;
    } else {
      // model_Bloch_2T1, MCRMWI/gpuMCRMWI.m:1026-1032
      const S ca = S(std::cos(famp)), sa = S(std::sin(famp));
      const S e1mw = S(std::exp(-TR / t1mw));
      aMW = S0MW * sa * (S(1) - e1mw) / (S(1) - ca * e1mw);
      const V e1iew = exp(-TR / (S(1) / r1iew));
      aIEW = (S0IW + S0EW) * sa * (S(1) - e1iew) / (S(1) - ca * e1iew);
    }
    const V aIW = aIEW * iwf;
    const V aEW = aIEW * (S(1) - iwf);
    V fbkg = V(S(0)), pini = V(S(0));
    if (cplx) {
      fbkg = th[kbkg + fi] + extra[3 + fi];
      // pini per flip angle when freqBKG is (MCRMWI/gpuMCRMWI.m:890-893, :580)
      pini = th[kpini] + ((md.flags & GC_FLAG_MCRMWI_PINI_FA) ? extra[3 + nfa + fi] : extra[1]);
    }
    for (int t = 0; t < nte; ++t) {
      const S te = S(md.table[t]);
      const S w0 = te * S(2) * S(kPi);
      const V pMW = w0 * (freqMW + fbkg) * B0 * gyro + pini;
      // compute_gremwi_signal: the IEW phase offsets IW and EW (:8-9)
      const V pIW = w0 * (freqIW + fbkg) * B0 * gyro + pini + ph;
      const V pEW = w0 * (freqEW + fbkg) * B0 * gyro + pini + ph;
      const V dMW = aMW * exp(-(te * r2sMW));
      const V dIW = aIW * exp(-(te * r2sIW));
      const V dEW = aEW * exp(-(te * r2sEW));
      const V re = (dMW * cos(pMW) + dIW * cos(pIW) + dEW * cos(pEW)) * ff;
      const V im = (dMW * sin(pMW) + dIW * sin(pIW) + dEW * sin(pEW)) * ff;
      if (cplx) {
        s[t + nte * fi] = re;
        s[t + nte * fi + nte * nfa] = im;
      } else {
        s[t + nte * fi] = sqrt(re * re + im * im);
      }
    }
  }
}

// Dispatch
template <class V, class S>
void fwd_model(const gc_model& md, const V* th, const S* extra, V* s) {
  switch (md.model_id) {
    case GC_MODEL_R2STAR: fwd_r2star<V, S>(md, th, extra, s); break;
    case GC_MODEL_JOINT_R1R2STAR: fwd_joint<V, S>(md, th, extra, s); break;
    case GC_MODEL_NEXI: fwd_nexi<V, S>(md, th, extra, s); break;
    case GC_MODEL_MCRMWI: fwd_mcrmwi<V, S>(md, th, extra, s); break;
    default: break;
  }
}

}  // namespace orc
