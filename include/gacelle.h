/*
 * gacelle.h -- C ABI of the MI355X (gfx950) voxelwise-estimation library
 * (libgacelle_hip.so).
 *
 * The library replaces GACELLE's MATLAB hot path: the askadam minimisation
 * loop, the mcmc Metropolis-Hastings and Goodman-Weare samplers, and the
 * forward models they drive.  Every entry point below names the reference
 * interface it stands in for (paths relative to the GACELLE source tree).
 *
 * Conventions
 *  - Plain C11, no C++ or framework types.  All functions return an int
 *    status: GC_OK (0) or a negative GC_ERR_* code; gc_last_error() gives
 *    the message.  Nothing throws across the ABI.
 *  - Array arguments named *_dev are DEVICE pointers (hipMalloc'ed or any
 *    device-accessible allocation) on the context's device.  Small
 *    descriptor structs (gc_model, gc_askadam_cfg, gc_mcmc_cfg) are host
 *    memory and are copied by value.
 *  - Voxel data use the "G-D" layout of utils/utils.m:157-163 (measurement
 *    by voxel), stored structure-of-arrays with the voxel index fastest:
 *    element (m, v) of an [M x Nv] matrix lives at m*Nv + v.  Measurement
 *    order is that of utils/utils.m:215-238: echo time fastest, then flip
 *    angle (or shell), then real/imaginary.
 *  - Parameter maps are [P x Nv] in the same layout, in the order of the
 *    model class's modelParams list.
 *  - A context is used by one host thread at a time; calls are ordered on
 *    the context's HIP stream.
 */
#ifndef GACELLE_H
#define GACELLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GC_ABI_VERSION 9
#define GC_MAX_PAR 24   /* parameter slots per voxel            */
#define GC_MAX_TABLE 256 /* acquisition table entries             */
#define GC_MAX_VOL 16   /* volumes per batched askadam launch     */
#define GC_MAX_REG 4    /* regularised parameter maps             */
#define GC_MLP_MAX_LAYERS 8  /* fully connected layers per network     */
#define GC_MLP_MAX_WIDTH 80  /* units per layer (inputs included)      */

/* ---- status codes (mirroring the MATLAB error identifiers) ------------- */
enum {
  GC_OK = 0,
  GC_ERR_ARG = -1,            /* invalid argument / shape                */
  GC_ERR_HIP = -2,            /* HIP runtime failure                     */
  GC_ERR_UNSUPPORTED = -3,    /* option not implemented by this library  */
  GC_ERR_NOMEM = -4,          /* 'GACELLE:memoryError' (utils/utils.m:762) */
  GC_ERR_NWALKER = -5,        /* 'goodman_weare:Nwalker' (utils/mcmc.m:243) */
  GC_ERR_INPUT_MISMATCH = -6, /* 'GACELLE:inputMismatch'                  */
  GC_ERR_NO_DEVICE = -7       /* no usable gfx950 device                  */
};

/* ---- forward models ----------------------------------------------------- */
enum {
  GC_MODEL_R2STAR = 1,         /* R2star/gpuR2starMapping.m                 */
  GC_MODEL_JOINT_R1R2STAR = 2, /* R1R2s/gpuJointR1R2starMapping.m           */
  GC_MODEL_NEXI = 3,           /* NEXI/gpuNEXI.m, NEXI/gpuNEXIrice.m        */
  GC_MODEL_MCRMWI = 4          /* MCRMWI/gpuMCRMWI.m (BM / Bloch / EPG-X T1) */
};

/* model option flags (gc_model.flags) */
enum {
  GC_FLAG_R1R2S_TRUE_FA = 1 << 0,  /* extra rows hold trueFlipAngle per FA    */
  GC_FLAG_NEXI_RICIAN = 1 << 1,    /* Rician mean correction, sigma = extra row 0 */
  GC_FLAG_NEXI_RE_GUARD = 1 << 2,  /* re = ra*fa/max(1-fa,eps) (gpuNEXIrice.m:678) */
  GC_FLAG_MCRMWI_COMPLEX = 1 << 3, /* complex data: re/im stacked, dfreqBKG/dpini fitted */
  GC_FLAG_MCRMWI_EXCHANGE = 1 << 4, /* BM two-pool exchange steady state (kIEWM fitted) */
  GC_FLAG_MCRMWI_PINI_FA = 1 << 6,  /* one initial phase per flip angle: extra
                                     * rows 3+n_fa..3+2n_fa replace row 1    */
  GC_FLAG_UNIFORM_TE = 1 << 5,      /* echo times equally spaced: f[8] = TE spacing;
                                       exp(-TE_k R) is formed by recurrence */
  GC_FLAG_MCRMWI_EPG = 1 << 7       /* EPG-X MLP T1 path (isEPG, needs EXCHANGE):
                                     * gc_model.mlp[0] = phase network,
                                     * mlp[1] = magnitude network           */
};

/* A multilayer perceptron of the MCR-MWI EPG-X path
 * (MCRMWI/EPGXgen_net/mlp_model_leakyRelu.m:23-47): fc1, then leaky ReLU
 * and fc_l for l = 2..n_layer.  params points at, per layer l = 1..n_layer,
 * Weights [width[l] x width[l-1]] row-major (MATLAB dlnet.parameters.fc<l>
 * .Weights, [out, in]) followed by Bias [width[l]], FP32.  width[0] = 11
 * (the features of feature_preprocess_MCRMWI_MLP_EPGX_leakyrelu.m:24-95);
 * width[n_layer] = 1 (phase network) or 2 (magnitude network: the IEW, then
 * the MW steady-state difference).  The library reads params as device
 * memory; the oracle reads it as host memory. */
typedef struct gc_mlp {
  int32_t n_layer;
  int32_t width[GC_MLP_MAX_LAYERS + 1];
  float alpha;        /* leaky ReLU scale (dlnet.alpha = 0.01, gpuMCRMWI.m:386-387) */
  int32_t reserved;
  const float* params;
} gc_mlp;

/* NEXI quadrature kinds (gc_model.i[3]) */
enum { GC_QUAD_TRAPZ = 0, GC_QUAD_TABLE = 1 };

/*
 * Model description.  Acquisition tables are uniform across voxels; the
 * library keeps a device copy that kernels read at wave-uniform addresses
 * (scalar loads through the scalar cache on gfx950).
 *
 *   R2STAR  : i[0]=n_te; table[0:n_te]=TE [s].              P=2 {M0,R2star}
 *   JOINT   : i[0]=n_te, i[1]=n_fa; f[0]=TR [s];
 *             table[0:n_te]=TE [s], table[n_te:n_te+n_fa]=FA [rad];
 *             extra row 0 = B1 (or, with GC_FLAG_R1R2S_TRUE_FA, n_fa rows
 *             of true flip angle [rad]); f[8] = TE spacing with
 *             GC_FLAG_UNIFORM_TE.
 *                                                        P=3 {M0,R1,R2star}
 *   NEXI    : i[0]=n_shell, i[1]=lmax (0|2), i[2]=n_nodes, i[3]=quad kind;
 *             table[0:ns]=b [ms/um^2], table[ns:2ns]=Delta [ms],
 *             table[2ns:2ns+nn]=nodes x in [0,1],
 *             table[2ns+nn:2ns+2nn]=quadrature weights (the trapezoid's
 *             weights for GC_QUAD_TRAPZ), table[2ns+2nn:+34]=cos of the 34
 *             Bessel trapezoid nodes on [0,pi] (Rician correction);
 *             f[0]=fa floor (askadam.epsilon, 1e-8); extra row 0 = sigma
 *             when GC_FLAG_NEXI_RICIAN.
 *                                        P=4 {fa,Da,De,ra} (+p2 when lmax=2)
 *   MCRMWI  : i[0]=n_te, i[1]=n_fa; f[0]=TR [s], f[1]=B0 [T], f[2]=gyro
 *             [MHz/T], f[3]=rho_mw, f[4]=t1_mw [s], f[5]=freq_EW [ppm];
 *             table[0:n_te]=TE [s], table[n_te:n_te+n_fa]=FA [rad];
 *             extra rows: 0=B1, 1=pini [rad], 2=ff, 3..3+n_fa=freqBKG [ppm],
 *             and with GC_FLAG_MCRMWI_PINI_FA 3+n_fa..3+2n_fa = pini per
 *             flip angle (the default pini of a per-FA freqBKG,
 *             MCRMWI/gpuMCRMWI.m:890-893).  At most 12 extra rows.
 *             P = {S0,MWF,IWF,R1IEW,[kIEWM],R2sMW,R2sIW,R2sEW,freqMW,freqIW,
 *                  [dfreqBKG x n_fa, dpini]}  ([] depends on flags)
 *
 * n_par may exceed the model's own parameter count: trailing slots are
 * carried (and optimised with zero gradient) exactly like a modelParams
 * entry that the forward model never reads.
 *
 * i[7] and f[9] of R2STAR and JOINT belong to the library: with
 * GC_FLAG_UNIFORM_TE and a spacing of one or two first-echo times it
 * writes there (in its device copy) the multiple and -log2(e) TE_0, so that
 * one exponential gives both decay factors; caller values are replaced.
 */
typedef struct gc_model {
  int32_t model_id;
  int32_t n_meas;  /* M: real measurements per voxel              */
  int32_t n_par;   /* P: parameter slots per voxel                */
  int32_t n_extra; /* rows of per-voxel auxiliary data            */
  int32_t flags;
  int32_t i[8];
  float f[16];
  float table[GC_MAX_TABLE];
  float lb[GC_MAX_PAR]; /* fitting.lb, one entry per slot          */
  float ub[GC_MAX_PAR]; /* fitting.ub                              */
  gc_mlp mlp[2];        /* GC_FLAG_MCRMWI_EPG: phase and magnitude networks */
} gc_model;

/* ---- askadam configuration: POD mirror of askadam.check_set_default_basic
 *      (utils/askadam.m:735-878) ------------------------------------------ */
enum { GC_OPT_ADAM = 0, GC_OPT_SGDM = 1, GC_OPT_RMSPROP = 2 };
enum { GC_LOSS_L1 = 0, GC_LOSS_L2 = 1, GC_LOSS_HUBER = 2, GC_LOSS_MSE = 3 };
enum { GC_TRANSFORM_SIGMOID = 0, GC_TRANSFORM_LINEAR = 1, GC_TRANSFORM_NONE = 2 };
enum { GC_CONV_EMA = 0, GC_CONV_LINEAR = 1 };
/* bias-correction form of adamupdate: lr*sqrt(1-b2^t)/(1-b1^t) * m/(sqrt(v)+eps)
 * (effective learning rate) or Kingma & Ba Alg. 1: lr * mhat/(sqrt(vhat)+eps) */
enum { GC_ADAM_EFFECTIVE_LR = 0, GC_ADAM_KINGMA_BA = 1 };
enum { GC_TV_2D = 0, GC_TV_3D = 1 };
/* regulariser of the regmap parameters: spatial total variation
 * (utils/spatial_total_variation.m:24-66), the normal prior of gpuNEXI
 * (prior_distribution_normal, NEXI/gpuNEXI.m:1256-1303) or the mesh total
 * variation of surface data (surface_total_variation, NEXI/gpuNEXI.m:1165-1254) */
enum { GC_REG_TV = 0, GC_REG_PRIOR = 1, GC_REG_SURFACE_TV = 2 };
enum {
  GC_STOP_NONE = 0,        /* ran all iterations                      */
  GC_STOP_CONVERGENCE = 1, /* loss convergence below tolerance        */
  GC_STOP_TOL = 2,         /* loss below fitting.tol                  */
  GC_STOP_STEP = 3,        /* step norm below convergenceStepTol      */
  GC_STOP_GRAD = 4         /* gradient norm below convergenceGradTol  */
};

typedef struct gc_askadam_cfg {
  int32_t iteration;          /* fitting.iteration                      */
  int32_t optimiser;          /* GC_OPT_*                               */
  int32_t loss;               /* GC_LOSS_* (fitting.lossFunction)       */
  int32_t transform;          /* GC_TRANSFORM_SIGMOID | _LINEAR         */
  int32_t adam_form;          /* GC_ADAM_*                              */
  int32_t convergence_model;  /* GC_CONV_*                              */
  int32_t convergence_window; /* fitting.convergenceWindow              */
  int32_t patience_conv;      /* fitting.patienceConvergence            */
  int32_t patience_step;      /* fitting.patienceStep                   */
  int32_t patience_grad;      /* fitting.patienceGrad                   */
  int32_t block_iters;        /* iterations per launch (0 = automatic)  */
  int32_t tv_mode;            /* GC_TV_2D | GC_TV_3D                    */
  int32_t n_reg;              /* number of regularised maps (0 = none)  */
  int32_t regmap[GC_MAX_REG]; /* parameter slot of each regularised map */
  int32_t robust;             /* fitting.robustConvergence              */
  int32_t reg_type;           /* GC_REG_* (fitting.regularisationType,
                                 fitting.dataType 'surface' with 'TV')   */
  double lr0;                 /* fitting.initialLearnRate               */
  double decay;               /* fitting.decayRate                      */
  double beta1, beta2, eps;   /* adamupdate decays and epsilon          */
  double momentum;            /* sgdmupdateMomentum                     */
  double rms_decay, rms_eps;  /* rmspropupdateSqGradDecay / Epsilon     */
  double tol;                 /* fitting.tol                            */
  double convergence_value;   /* fitting.convergenceValue               */
  double ema_decay;           /* fitting.emaDecay                       */
  double step_tol;            /* fitting.convergenceStepTol (0 = off)   */
  double grad_tol;            /* fitting.convergenceGradTol (0 = off)   */
  double lambda[GC_MAX_REG];  /* fitting.lambda                         */
  double voxel_size[3];       /* fitting.voxelSize                      */
  /* robustConvergence (askadam.update_outlier_mask, utils/askadam.m:448-510,
   * and the frozen convergence set of :541-566) */
  int32_t weight_update_interval; /* fitting.weightUpdateInterval        */
  int32_t outlier_check_window;   /* fitting.outlierCheckWindow          */
  double outlier_weight;          /* fitting.outlierWeight               */
  double outlier_voxel_thres;     /* fitting.outlierVoxelThres           */
  double outlier_pop_thres;       /* fitting.outlierPopThres             */
  double outlier_init_thres;      /* fitting.outlierInitThres            */
  double outlier_init_pop_thres;  /* fitting.outlierInitPopThres         */
} gc_askadam_cfg;

/* One masked volume of a batched askadam run.  Inputs are device arrays;
 * scalar outputs are written back into this struct. */
typedef struct gc_volume {
  int64_t nv;                /* masked voxels                            */
  const float* y_dev;        /* [M x nv] data                            */
  const float* w_dev;        /* [M x nv] weights, NULL = ones            */
  const float* extra_dev;    /* [n_extra x nv], NULL when n_extra = 0    */
  const int64_t* vox_dev;    /* [nv] linear grid index of each voxel
                                (column-major find(mask)); TV only       */
  int32_t dims[3];           /* image grid (TV only)                     */
  int32_t reserved0;
  float* z_dev;              /* in: initial optimisation variables [P x nv]
                                (askadam.initialise_parameter space);
                                out: final variables (out.final)         */
  float* zmin_dev;           /* out: variables at minimum loss (out.min) */
  float* resloss_dev;        /* out: [nv] per-voxel mean residual, final; NULL ok */
  float* resloss_min_dev;    /* out: [nv] same at minimum loss; NULL ok  */
  float* residual_dev;       /* out: [M x nv] residuals, final; NULL ok  */
  float* residual_min_dev;   /* out: [M x nv] residuals at min; NULL ok  */
  double* loss_trace_host;   /* optional HOST array [iteration+1]: loss at
                                the starting point and at every evaluated
                                iteration (NULL = not recorded)         */
  /* scalar outputs */
  double loss, loss_fidelity, loss_reg;
  double min_loss, min_loss_fidelity, min_loss_reg;
  double convergence;        /* final convergence signal                 */
  int32_t n_iteration;       /* out.final.Niteration                     */
  int32_t min_iteration;     /* out.min.Niteration                       */
  int32_t stop_reason;       /* GC_STOP_*                                */
  int32_t n_launch;          /* fused launches that advanced this volume */
  /* timing of the fused iteration kernel over the whole batched run
   * (HIP events on the context stream; same values in every volume) */
  double kernel_ms;          /* summed duration of all fused launches    */
  int32_t kernel_launches;   /* number of fused launches                 */
  int32_t iters_per_launch;  /* K, iterations per launch                 */
  /* GC_REG_PRIOR: loss_reg = sum_k lambda_k / nv * sum over the image grid
   * of the finite |(theta_k - mu_k) / sigma_k|, theta_k zero outside the
   * mask.  prior_dev holds mu then sigma of each regularised map at the
   * masked voxels, [2*n_reg x nv]; reg_offset is the part the grid voxels
   * outside the mask contribute (lambda_k / nv * sum |mu_k / sigma_k|), a
   * constant added to loss_reg every iteration. */
  const float* prior_dev;
  double reg_offset;
  /* GC_REG_SURFACE_TV: the vertex grid is dims = {1, Nvertex, Nhemi} and
   * vox_dev the grid index of each masked vertex.  surf_nbr_dev is
   * [surf_max_nb x Nvertex*Nhemi] of global 0-based vertex indices
   * (the reference's neighbours + hemiOffset - 1, same layout) and
   * surf_len_dev the matching edge lengths.  loss_reg = sum_k lambda_k / nv
   * * sum over masked vertices i and slots n of the finite
   * |(theta_k(i) - theta_k(nbr)) / len|, theta_k zero outside the mask. */
  const int32_t* surf_nbr_dev;
  const float* surf_len_dev;
  int32_t surf_max_nb;
  int32_t reserved1;
  /* The voxel count N of the mean in askadam.model_gradient (loss =
   * mean over N*M residuals, utils/askadam.m:96) and of the regulariser's
   * 1/Nmasked; 0 = nv.  Set it to the whole problem's count when this
   * volume is one piece (a shard) of a larger fit, so that its gradient
   * scale, and with stopping off its trajectory, are those of the whole
   * fit.  Scalar losses are reported with the same N. */
  int64_t n_mean;
  /* Test hook (L1 loss, blocked form only): when not NULL, receives the
   * sign of every residual the fused kernel differentiates, two bits per
   * measurement m in word m/16 at bit 2*(m%16): 1 = positive, 2 =
   * negative, 0 = zero or non-finite model signal (no gradient).  Layout
   * [iteration][ceil(M/16)][nv] of uint32, row e-1 for epoch e.  Lets an
   * FP64 replay follow the same L1 subgradient path. */
  uint32_t* sign_trace_dev;
  int32_t block_threads;     /* out: threads per block of the fused launches
                                (GC_OPT_ASK_BLOCK or the autotuner's pick;
                                256 in the per-iteration form)            */
} gc_volume;

/* ---- MCMC configuration: POD mirror of mcmc.check_set_default_basic
 *      (utils/mcmc.m:633-679) ---------------------------------------------- */
enum { GC_MCMC_MH = 0, GC_MCMC_GW = 1 };
enum { GC_ENSEMBLE_SIMULTANEOUS = 0, GC_ENSEMBLE_REDBLACK = 1 };
enum { GC_RNG_NATIVE = 0, GC_RNG_REPLAY = 1 };

typedef struct gc_mcmc_cfg {
  int32_t algorithm;       /* GC_MCMC_MH | GC_MCMC_GW                   */
  int32_t iteration;       /* fitting.iteration                         */
  int32_t thinning;        /* fitting.thinning                          */
  int32_t n_burnin;        /* mcmc.get_number_burnin (utils/mcmc.m:788) */
  int32_t repetition;      /* fitting.repetition (independent chains)   */
  int32_t n_walker;        /* fitting.Nwalker (GW)                      */
  int32_t ensemble_update; /* GC_ENSEMBLE_* (fitting.Ensembleupdate)    */
  int32_t rng_mode;        /* GC_RNG_NATIVE | GC_RNG_REPLAY             */
  int32_t noise_slot;      /* index of 'noise' among the Q variables    */
  int32_t n_var;           /* Q: variables per voxel incl. noise        */
  uint64_t seed;           /* native-RNG seed                           */
  double step_size;        /* fitting.StepSize (GW stretch scale a)     */
  double start_range;      /* fitting.startRange (GW walker spread)     */
  float x_step[GC_MAX_PAR];/* fitting.xStepSize (MH proposal sd)        */
  /* Goodman-Weare with global (shared) parameters,
   * goodman_weare_wglobal_constant (utils/mcmc.m:385-578): bit q set =
   * variable q is one value per walker shared by all voxels (a field of
   * pars0 with one element, utils/mcmc.m:54-64, :760-770).  0 = none. */
  uint32_t global_mask;
  int32_t global_warmup;   /* fitting.globalWarmup: global-only steps   */
} gc_mcmc_cfg;

typedef struct gc_mcmc_io {
  int64_t nv;
  const float* y_dev;        /* [M x nv]                                  */
  const float* w_dev;        /* [M x nv] or NULL (ones)                   */
  const float* extra_dev;    /* [n_extra x nv] or NULL                    */
  const float* x0_dev;       /* [Q x nv] starting point                   */
  /* replay draws (GC_RNG_REPLAY), in the reference's draw order:
   *  MH : normals  [rep][iter][nv][Q]   (randn([Q,Nv]) of utils/mcmc.m:147)
   *       uniforms [rep][iter][nv]      (rand(1,Nv) of :161)
   *  GW : normals  [W][nv][Q]           (walker initialisation, :271)
   *       uniforms [rep][iter][2][W][nv] (zz draw then acceptance draw,
   *                 :299 and :314; red-black: per half [rep][iter][2 halves]
   *                 [2][W/2][nv])
   *       partners [rep][iter][W]       (find_partner derangement, 0-based;
   *                 red-black: [rep][iter][2][W/2] anchor walker indices)  */
  const float* normals_dev;
  const float* uniforms_dev;
  const int32_t* partners_dev;
  /* With global parameters (cfg.global_mask != 0) the local block keeps
   * the GW layouts above over the NND local variables (normals
   * [W][nv][NND]; uniforms and partners per local block, :507-545), and
   * the global block (gw_global_block, :581-626) draws from
   *  g_uniforms [W][NG]             (rand(size(ub_global)), :489)
   *             then [rep][T][2][W]  (zz_g then acceptance, :587 and :597;
   *             red-black: [rep][T][2 halves][2][hm])
   *  g_partners [rep][T][W]          (find_partner; red-black anchors)
   * where T = global_warmup + iteration, warm-up steps first, and
   * hm = W - W/2.                                                      */
  const float* g_uniforms_dev;
  const int32_t* g_partners_dev;
  /* outputs */
  float* posterior_dev;      /* [Q][rep][Ns][W][nv] (MH: W = 1) or NULL  */
  float* mean_dev;           /* [Q x nv] posterior mean (res2out 'mean')  */
  float* std_dev;            /* [Q x nv] posterior std, N-1 ('std')       */
  float* accept_dev;         /* [nv] acceptance fraction, or NULL          */
  float* xfinal_dev;         /* [rep][W][Q][nv] final chain state, or NULL */
  uint8_t* trace_dev;        /* test hook: [rep][iter][W][nv] accept flags */
  /* global-parameter outputs (cfg.global_mask != 0).  posterior_dev and
   * xfinal_dev then hold the NND local variables only, in model order
   * ([NND][rep][Ns][W][nv], [rep][W][NND][nv]); mean_dev / std_dev rows
   * of global variables hold the scalar posterior statistic at every
   * voxel.                                                             */
  float* g_posterior_dev;    /* [NG][rep][Ns][W] or NULL                  */
  float* g_xfinal_dev;       /* [rep][W][NG] or NULL                      */
  uint8_t* g_trace_dev;      /* test hook: [rep][T][W] global accept flags */
  int32_t n_sample;          /* out: Ns = numel(Nburnin+1:thinning:iteration) */
  int32_t n_stored;          /* out: samples actually stored per chain    */
  double kernel_ms;          /* out: sampler kernel duration (HIP events) */
  int32_t block_threads;     /* out: threads per block of the MH / GW launch
                              * (GC_OPT_MCMC_BLOCK or the autotuner's pick) */
} gc_mcmc_io;

/* ---- context ------------------------------------------------------------ */
typedef struct gc_ctx gc_ctx;

/* ABI version of the loaded library (GC_ABI_VERSION). */
int gc_version(void);

/* Create a context on HIP device `device` (replaces gpuDevice,
 * utils/askadam.m:149 and utils/mcmc.m:113-118).  Work is issued on the
 * device's default stream until gc_ctx_set_stream. */
int gc_ctx_create(int32_t device, gc_ctx** out);
int gc_ctx_destroy(gc_ctx* ctx);
/* Issue work on an existing hipStream_t (NULL = the default stream). */
int gc_ctx_set_stream(gc_ctx* ctx, void* hip_stream);
int gc_ctx_synchronize(gc_ctx* ctx);
const char* gc_last_error(const gc_ctx* ctx);

/* Context options (no counterpart in the reference: MATLAB's dlaccelerate
 * trace cache, utils/askadam.m:192, is the nearest analogue of
 * GC_OPT_GRAPHS).
 *  GC_OPT_GRAPHS      1 (default): the per-iteration launches of the step
 *                     form (TV, prior, robustConvergence, streamed models)
 *                     are captured in hipGraphs of up to 64 launches and
 *                     replayed per iteration block; 0: launched one by one.
 *  GC_OPT_MCMC_BLOCK  threads per block of the MCMC samplers: 0 (default)
 *                     = picked per model and chain layout by a short timed
 *                     trial on the first call (cached in the context), or
 *                     64, 128 or 256.  Results do not depend on it.
 *  GC_OPT_ASK_BLOCK   threads per block of the blocked askadam kernel (the
 *                     form without a regulariser): 0 (default) = picked per
 *                     model variant by a short timed trial on the first call
 *                     (cached in the context), or 64, 128 or 256.  The same
 *                     trial measures the cost of one launch cycle and of one
 *                     voxel-iteration, from which the iterations per launch
 *                     (K) are set when stopping criteria are on and
 *                     gc_askadam_cfg.block_iters is 0.  Results depend on
 *                     neither. */
enum { GC_OPT_GRAPHS = 1, GC_OPT_MCMC_BLOCK = 2, GC_OPT_ASK_BLOCK = 3 };
int gc_ctx_set_option(gc_ctx* ctx, int32_t option, int64_t value);

/* ---- layout helpers (utils/utils.m:157-163, 194-198, 299-315) ----------- */
/* G-D gather: gd[m*nv+v] = nd[vox[v] + nvox_grid*m]; nd is column-major
 * [X,Y,Z,M] (reshape_ND2GD with find(mask) order). */
int gc_nd_to_gd(gc_ctx* ctx, const float* nd_dev, int64_t nvox_grid, int32_t n_meas,
                const int64_t* vox_dev, int64_t nv, float* gd_dev);
/* Inverse scatter with zeros outside the mask (reshape_GD2ND). */
int gc_gd_to_nd(gc_ctx* ctx, const float* gd_dev, int64_t nv, int32_t n_meas,
                const int64_t* vox_dev, int64_t nvox_grid, float* nd_dev);

/* ---- parameter transforms (utils/askadam.m:368-412, 1063-1099) ---------- */
/* theta -> z: clamp to [lb+1e-4(ub-lb), ub-1e-4(ub-lb)] then logit
 * (sigmoid), or rescale01 (linear). */
int gc_rescale_parameters(gc_ctx* ctx, const gc_model* model, int32_t transform,
                          const float* theta_dev, float* z_dev, int64_t nv);
/* z -> theta: lb + (ub-lb)*sigmoid(z) or z*(ub-lb)+lb. */
int gc_unscale_parameters(gc_ctx* ctx, const gc_model* model, int32_t transform,
                          const float* z_dev, float* theta_dev, int64_t nv);

/* ---- operator pair (the FWD plugin + dlgradient of askadam.model_gradient,
 *      utils/askadam.m:23-117) ------------------------------------------- */
/* s = FWD(theta): R2star/gpuR2starMapping.m:321-363,
 * R1R2s/gpuJointR1R2starMapping.m:375-422, NEXI/gpuNEXI.m:902-996,
 * MCRMWI/gpuMCRMWI.m:566-714.  NaN/Inf outputs are zeroed
 * (utils/utils.m:347-350). */
int gc_fwd(gc_ctx* ctx, const gc_model* model, const float* theta_dev,
           const float* extra_dev, float* s_dev, int64_t nv);
/* Fused forward + adjoint of loss = scale * sum(lossfn(w, FWD(x), y)).
 * x is theta (transform = GC_TRANSFORM_NONE) or z (sigmoid/linear, the
 * gradient then includes the transform).  grad_dev [P x nv] receives
 * d loss / d x; loss_sum_dev (one double) receives sum(lossfn) (unscaled);
 * residual_dev (optional, [M x nv]) receives the elementwise loss. */
int gc_fwd_adj(gc_ctx* ctx, const gc_model* model, int32_t loss, int32_t transform,
               const float* x_dev, const float* y_dev, const float* w_dev,
               const float* extra_dev, int64_t nv, float scale, float* grad_dev,
               double* loss_sum_dev, float* residual_dev);

/* ---- the operator pair without allocation --------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same arguments, kernels and results as gc_fwd_adj, but
 * what gc_fwd_adj allocates per call (the device copy of the model
 * description and one FP64 loss partial per 256-voxel block) lives in a
 * fixed library array of GC_NOALLOC_SCRATCH_BYTES on each device, held
 * until the process ends.  Limits: nv <= GC_NOALLOC_MAX_NV (else
 * GC_ERR_ARG); the EPG-X MCR-MWI model (GC_FLAG_MCRMWI_EPG), whose network
 * repacking needs per-call memory, is rejected with GC_ERR_UNSUPPORTED.
 * Every call on a device uses that device's one array: calls ordered on one
 * stream are safe, calls that overlap on one device (other streams or
 * threads) race.  Asynchronous on the context's stream (gc_fwd_adj
 * synchronises). */
#define GC_NOALLOC_MAX_NV (1 << 25)
#define GC_NOALLOC_SCRATCH_BYTES (4096 + 8 * (GC_NOALLOC_MAX_NV / 256))
int gc_fwd_adj_noalloc(gc_ctx* ctx, const gc_model* model, int32_t loss, int32_t transform,
                       const float* x_dev, const float* y_dev, const float* w_dev, const float* extra_dev,
                       int64_t nv, float scale, float* grad_dev, double* loss_sum_dev, float* residual_dev);

/* ---- regulariser (utils/spatial_total_variation.m:24-66) ---------------- */
/* loss = lambda/n_masked * sum_i mask_i * sum_d |D_d img_i| / vsize_d with
 * periodic forward differences; grad_img_dev (optional) receives
 * d loss / d img (subgradient, sign(0) = 0).  img is column-major [X,Y,Z];
 * loss is written to a HOST double (the call synchronises). */
int gc_tv_fwd_adj(gc_ctx* ctx, const float* img_dev, const uint8_t* mask_dev,
                  const int32_t dims[3], int32_t tv_mode, const double voxel_size[3],
                  double lambda, int64_t n_masked, double* loss, float* grad_img_dev);

/* ---- posterior statistics ------------------------------------------------ */
/* res2out 'median', 'iqr' and 'mode' (utils/mcmc.m:823-857) over the n
 * samples of each voxel and variable of a posterior [Q][n][nv] (voxel
 * fastest: gc_mcmc_io.posterior_dev with n = rep * Ns * W).  lb / ub are
 * HOST arrays [Q]; mode bins are discretize(x, linspace(lb-1e-8, ub+1e-8,
 * nbin)).  Outputs [Q][nv], each optional (NULL skips it).  n <= 16384. */
int gc_posterior_stats(gc_ctx* ctx, const float* posterior_dev, int32_t n_var, int32_t n, int64_t nv,
                       const float* lb, const float* ub, int32_t nbin, float* median_dev, float* iqr_dev,
                       float* mode_dev);

/* ---- dictionary search ------------------------------------------------- */
/* The likelihood start of NEXI (gpuNEXI.likelihood, NEXI/gpuNEXI.m:822-849;
 * gpuNEXIrice.likelihood, NEXI/gpuNEXIrice.m:602-625): for each voxel v of
 * y_dev [m][nv] (voxel fastest) the entry j of dict_dev [m][nd] of least
 *   sse_j = sum_k w_k (y_kv - c_j d_kj)^2,  c_j = (y_v . d_j) / (d_j . d_j),
 * the first one on ties; entries with a NaN score never win (MATLAB's min
 * skips NaN).  w is a HOST array [m] of non-negative weights (NULL: ones).
 * best_dev [nv] receives the 0-based entry (-1 when no entry scores), and
 * sse_dev [nv] (FP64, NULL skips it) its sse.  Scores are formed in FP64
 * from the FP32 inputs.  1 <= m <= 32.  Synchronous on the context's
 * stream. */
int gc_dict_search(gc_ctx* ctx, const float* y_dev, int64_t nv, int32_t m, const float* dict_dev, int64_t nd,
                   const float* w, int32_t* best_dev, double* sse_dev);

/* ---- solvers ------------------------------------------------------------ */
/* askadam().optimisation (utils/askadam.m:121-365) for one model over a
 * batch of independent volumes, each with its own N, stopping state and
 * outputs. */
int gc_askadam_run(gc_ctx* ctx, const gc_model* model, const gc_askadam_cfg* cfg,
                   gc_volume* vols, int32_t n_vol);

/* mcmc().metropolis_hastings (utils/mcmc.m:82-190),
 * mcmc().goodman_weare (utils/mcmc.m:192-383) and, when cfg.global_mask
 * is set, mcmc().goodman_weare_wglobal_constant with gw_global_block
 * (utils/mcmc.m:385-626), with res2out mean/std (utils/mcmc.m:806-862)
 * computed on the device. */
int gc_mcmc_run(gc_ctx* ctx, const gc_model* model, const gc_mcmc_cfg* cfg,
                gc_mcmc_io* io);

#ifdef __cplusplus
}
#endif
#endif /* GACELLE_H */
