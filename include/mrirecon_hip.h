/* mrirecon_hip.h -- C ABI of the MI355X (gfx950) NUFFT + NLCG hot path.
 *
 * Plain pointers and sizes only.  "device" pointers are HIP device memory
 * (e.g. torch.Tensor.data_ptr() of a CUDA tensor on ROCm), "host" pointers are
 * ordinary CPU memory.  complex64 is interleaved (re, im) float pairs.  Every
 * kernel is enqueued on the `stream` argument (a hipStream_t; pass the stream
 * that is current in the caller, never a cached one) and nothing blocks the
 * host except plan / engine creation and the explicit *_read calls.
 *
 * Every entry point returns 0 on success or an MRH_E* code; mrh_last_error()
 * returns the message of the last failure on the calling thread.
 *
 * Reference file:line citations name the function of timodelgado/mrirecon
 * (package graspcg) that each entry point replaces.
 */
#ifndef MRIRECON_HIP_H
#define MRIRECON_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRH_DCF_NONE 0
#define MRH_DCF_BALANCED 1
#define MRH_DCF_STANDARD 2

/* ------------------------------------------------------------------ errors */

/* Message of the last failed call on this thread ("" if none). */
const char* mrh_last_error(void);

/* ABI version of the library (bumped on any signature change). */
int mrh_abi_version(void);

/* ------------------------------------------------------------------ NUFFT */

typedef struct mrh_nufft_plan mrh_nufft_plan;

typedef struct {
  int ndim;                 /* 2 (3-D volumes use mrh_nufft3_desc below)              */
  int S0, S1;               /* image size (maps.shape[1:])                            */
  int G0, G1;               /* oversampled grid size                                  */
  int W;                    /* Kaiser-Bessel width in grid cells (2..8)                */
  double beta;              /* KB shape; <= 0: Beatty's formula at osf = G0 / S0       */
  int C;                    /* coils                                                   */
  int B;                    /* frames (batch of trajectories)                          */
  int like;                 /* largest Like-fold used at call time (>= 1)              */
  const int64_t* k_offsets; /* host (B+1): frame b owns samples [k_offsets[b], k_offsets[b+1]) */
  const float* traj;        /* host: per frame a (2, K_b) block in radians, blocks in frame order */
  const float* dcf;         /* host (sum K_b) or NULL                                  */
  int dcf_mode;             /* MRH_DCF_*                                               */
  int dcf_in_fwd;           /* 'standard' mode: multiply by dcf in A                   */
  int dcf_in_adj;           /* 'standard' mode: multiply by dcf in A^H                 */
  const void* maps;         /* device complex64 (C, S0, S1); copied into the plan      */
  int strip_cols;           /* grid columns per strip workgroup: 0 (= 8), 8, 16 (G0 <= 512)
                               or 4 (W = 4); the Python adapter autotunes it            */
} mrh_nufft_desc;

typedef struct {
  int G0, G1, W, strip_cols, n_strips, kb_terms;
  double beta;
  int64_t n_samples, n_entries, max_strip_entries, workspace_bytes;
  int coil_chunk;           /* coils per strip workgroup (mrh_nufft_set_coil_chunk)    */
} mrh_nufft_info;

/* Build a plan: footprints, strip binning, KB coefficients, deapodisation,
 * twiddles and the intermediate workspace.  Replaces
 * TorchKbNUFFTAdapter.prepare (graspcg/nufft/backends/torchkb.py:98-144). */
int mrh_nufft_plan_create(const mrh_nufft_desc* desc, mrh_nufft_plan** plan);
int mrh_nufft_plan_destroy(mrh_nufft_plan* plan);
int mrh_nufft_plan_info(const mrh_nufft_plan* plan, mrh_nufft_info* info);

/* Work split of the strip kernels: each workgroup handles one (frame, strip)
 * and `coils_per_block` coils (default C: all coils, one workgroup per
 * (frame, strip)).  Smaller chunks give more workgroups for batches of few
 * frames; results are bit-identical for every value.  Chosen per plan by the
 * autotuner (NufftPlan.autotune, which times each candidate).  No reference
 * counterpart (the reference NUFFT is a third-party call, torchkb.py:249,309). */
int mrh_nufft_set_coil_chunk(mrh_nufft_plan* plan, int coils_per_block);

/* Forward A: x (B*like, S0, S1) coil-combined image (or (B*like, C, S0, S1)
 * coil images when x_has_coils) -> y, frame b / like l block at
 * C*(k_offsets[b]*like + l*K_b) laid out (C, K_b); for uniform K this is
 * (B*like, C, K).  Replaces TorchKbNUFFTAdapter.A
 * (graspcg/nufft/backends/torchkb.py:203-270) and NUFFT.A (graspcg/nufft/api.py:167-220). */
int mrh_nufft_forward(mrh_nufft_plan* plan, const void* x, void* y, int like, int x_has_coils,
                      void* stream);

/* Adjoint A^H: y -> x (B*like, S0, S1), coil-combined with conj(maps).
 * Replaces TorchKbNUFFTAdapter.AH (graspcg/nufft/backends/torchkb.py:273-323)
 * and NUFFT.AH (graspcg/nufft/api.py:223-270). */
int mrh_nufft_adjoint(mrh_nufft_plan* plan, const void* y, void* x, int like, void* stream);

/* Forward / adjoint restricted to the frames of active problems: frame b
 * belongs to problem b / frames_per_problem, and its workgroups run only when
 * problem_active[problem] != 0 (device int array, e.g. mrh_cg_active_flags);
 * the k-space (forward) or image (adjoint) entries of skipped frames keep their
 * previous values.  Lets a batch of slices that stop at different iterations
 * stop paying for the NUFFT of the finished ones.  No reference counterpart
 * (the reference has no batched stop). */
int mrh_nufft_forward_masked(mrh_nufft_plan* plan, const void* x, void* y, int like, int x_has_coils,
                             const int* problem_active, int frames_per_problem, void* stream);
int mrh_nufft_adjoint_masked(mrh_nufft_plan* plan, const void* y, void* x, int like,
                             const int* problem_active, int frames_per_problem, void* stream);

/* The two kernels of each direction, launched separately so that a caller can
 * time them with events on `stream`.  Forward: stage 0 = row pass
 * (maps * deapod * zero-pad * row FFT, clears y), stage 1 = strip pass (column
 * FFT + interpolation); stage 0 must precede stage 1.  Adjoint: stage 0 =
 * strip pass (spreading + inverse column FFT), stage 1 = row pass (inverse row
 * FFT, crop, deapod, conj(maps) coil sum).  Running both stages in order is
 * exactly mrh_nufft_forward / mrh_nufft_adjoint.  No counterpart in the
 * reference (whose NUFFT is a single third-party call, torchkb.py:249,309). */
int mrh_nufft_forward_stage(mrh_nufft_plan* plan, int stage, const void* x, void* y, int like,
                            int x_has_coils, void* stream);
int mrh_nufft_adjoint_stage(mrh_nufft_plan* plan, int stage, const void* y, void* x, int like,
                            void* stream);

/* ------------------------------------------------------------------ 3-D NUFFT
 * Volumes (S0, S1, S2) and trajectories (3, K_b) per frame: the ndim=3 case of the
 * reference front end and its default backend (graspcg/nufft/api.py:85-89,
 * graspcg/nufft/backends/torchkb.py:98-323 with ndim=3), same operator contract as the
 * 2-D plan (maps, DCF policy, centred placement, KB window and deapodisation per axis).
 * The oversampled grid lives in the plan's HBM workspace (C G0 G1 G2 complex64); one
 * like image per frame. */
typedef struct mrh_nufft3_plan mrh_nufft3_plan;

typedef struct {
  int S0, S1, S2;           /* volume size (maps.shape[1:])                          */
  int G0, G1, G2;           /* grid size per axis (32 ... 1024, 2^a 3^b)               */
  int W;                    /* KB width (2 ... 8)                                      */
  double beta;              /* KB beta; <= 0: Beatty's formula at osf = G0 / S0         */
  int C;                    /* coils                                                   */
  int B;                    /* frames                                                  */
  const int64_t* k_offsets; /* host (B+1): sample offsets of each frame                 */
  const float* traj;        /* host: per frame a (3, K_b) block in radians             */
  const float* dcf;         /* host (sum K_b) or NULL                                  */
  int dcf_mode;             /* MRH_DCF_*                                               */
  int dcf_in_fwd;           /* 'standard' mode: multiply by dcf in A                   */
  int dcf_in_adj;           /* 'standard' mode: multiply by dcf in A^H                 */
  const void* maps;         /* device complex64 (C, S0, S1, S2); copied into the plan  */
} mrh_nufft3_desc;

int mrh_nufft3_plan_create(const mrh_nufft3_desc* desc, mrh_nufft3_plan** plan);
int mrh_nufft3_plan_destroy(mrh_nufft3_plan* plan);
/* x (B, S0, S1, S2) complex64 -> y: per frame a (C, K_b) block (torchkb.py:203-270). */
int mrh_nufft3_forward(mrh_nufft3_plan* plan, const void* x, void* y, void* stream);
/* y -> x (B, S0, S1, S2), coil-combined with conj(maps) (torchkb.py:273-323).  The spread
 * is a gather over footprint entries sorted by grid point at plan time (deterministic). */
int mrh_nufft3_adjoint(mrh_nufft3_plan* plan, const void* y, void* x, void* stream);

/* ------------------------------------------------------------------ reductions / objective */

/* Number of float64 workspace slots the reduction entry points below need. */
int64_t mrh_reduce_workspace_len(void);

/* r = t*Ad + Ax - y (in place into r, that order) and out[0] = sum |r|^2 in
 * float64.  Replaces _ker_residual_energy_ (graspcg/ops/objective.py:31-45).
 * `work`: device float64[mrh_reduce_workspace_len()]; `out`: device float64. */
int mrh_residual_energy(void* r, const void* Ax, const void* Ad, const void* y, double t,
                        int64_t n, double* work, double* out, void* stream);

/* out[0] = Re sum conj(a) * b over n complex64 elements, float64 accumulation.
 * Replaces dot_real0 / dot_chunked (graspcg/numerics/utils.py:94-101,
 * graspcg/ops/dot.py:5-71). */
int mrh_dot_real(const void* a, const void* b, int64_t n, double* work, double* out, void* stream);

/* ------------------------------------------------------------------ TV */

#define MRH_TV_AXIS_T 1  /* temporal (frame) axis, within each problem */
#define MRH_TV_AXIS_X 2  /* first image axis (S0)                       */
#define MRH_TV_AXIS_Y 4  /* second image axis (S1)                      */
#define MRH_TV_MAX_TERMS 3

typedef struct {
  float weight;      /* lambda                        */
  float eps;         /* Huber knee                    */
  int axes;          /* OR of MRH_TV_AXIS_*           */
  int isotropic;     /* 1: l2 over axes, 0: l1        */
  float axis_w[3];   /* per-axis weights (t, x, y)    */
} mrh_tv_term;

typedef struct {
  int P;             /* independent problems (slices)              */
  int F;             /* frames per problem                         */
  int S0, S1;        /* image size                                 */
  int n_terms;
  mrh_tv_term term[MRH_TV_MAX_TERMS];
} mrh_tv_desc;

/* g += sum_terms -lambda * div(flux) at x (when g != NULL) and
 * energy[p] = sum_terms lambda * sum e for every problem p (device float64[P]).
 * Image layout (P*F, S0, S1); temporal differences never cross problems.
 * Replaces TVND.energy_grad_fixed_axes (graspcg/regularization/tv_nd.py:54-88)
 * with the stencils of graspcg/ops/ndops.py:28-149. */
int mrh_tv_energy_grad(const mrh_tv_desc* desc, const void* x, void* g, double* work,
                       double* energy, void* stream);

/* mrh_tv_energy_grad without the workspace argument and without any allocation: the block
 * partials live in a static device array of 1 MiB per GPU (131072 doubles, the length
 * mrh_reduce_workspace_len() gives), whose address is looked up once per device on the first
 * call.  Same kernels and partition as mrh_tv_energy_grad, so the results are identical.
 * 1 <= desc->P <= 65535, else MRH_EINVAL.  Calls that overlap on one GPU (other streams or
 * threads) share the array and race.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int mrh_tv_energy_grad_noalloc(const mrh_tv_desc* desc, const void* x, void* g, double* energy,
                               void* stream);

/* ------------------------------------------------------------------ device NLCG engine */

typedef struct mrh_cg mrh_cg;

#define MRH_CG_DIR_FR 0
#define MRH_CG_DIR_PRPLUS 1
#define MRH_CG_DIR_DY 2       /* Dai-Yuan: beta = |g|^2 / Re<d, g - g_prev> */

#define MRH_CG_LS_ARMIJO 0    /* backtracking Armijo (graspcg/numerics/line_search.py:24-79)        */
#define MRH_CG_LS_WOLFE 1     /* strong Wolfe with bisection zoom (line_search.py:89-204)          */

#define MRH_CG_ACTIVE 0
#define MRH_CG_LS_FAILED 1
#define MRH_CG_NOT_DESCENT 2
#define MRH_CG_REL_TOL 3
#define MRH_CG_TARGET 4       /* objective reached the target set by mrh_cg_set_targets */

typedef struct {
  mrh_tv_desc tv;           /* image layout + regularisers                      */
  const int64_t* k_prob;    /* host (P+1): complex64 offsets of each problem's k-space */
  int direction;            /* MRH_CG_DIR_*                                     */
  double c1;                /* Armijo constant (1e-4)                           */
  double ls_beta;           /* backtracking factor (0.5)                        */
  int ls_max_iter;          /* trials (20)                                      */
  double rel_tol;           /* stop when |f_k - f_{k-1}| <= rel_tol |f_{k-1}|; 0 = never */
  int max_hist;             /* iterations of history kept on the device          */
  double restart_c;         /* < 0: reference behaviour (a non-descent direction
                               stops the problem).  >= 0: restart with beta = 0
                               whenever the new direction -g + beta d would have
                               Re<g, d_new> > -restart_c |g|^2 (not a reference
                               rule; the descent safeguard of PR+ with restarts) */
  int ls_kind;              /* MRH_CG_LS_* (CGConfig.ls_name, graspcg/solvers/cg.py:37) */
  double c2;                /* strong-Wolfe curvature constant (0.9, cg.py:39)   */
  float* diag;              /* device, P*F*S0*S1 floats: the diagonal preconditioner D of the
                               direction rules (the workspace's "diag", cg.py:95), read by
                               mrh_cg_gradient / mrh_cg_direction; NULL: D = 1 */
  double gl_weight;         /* graph-Laplacian term along frames (graspcg/regularization/
                               graph_laplacian.py:22-321): lambda; 0 = no term          */
  int gl_F;                 /* frames of its graph (= tv.F, at most 48)               */
  const float* gl_L;        /* host F*F row-major Laplacian (plain / rw / sym, from the
                               symmetrised graph)                                      */
  const float* gl_prof;     /* host F: its diagonal majoriser per frame (lambda deg or
                               lambda [deg > 0], graph_laplacian.py:283-321)            */
  int fp32_reductions;      /* 1: every inner product, norm, energy, beta and line-search
                               value in FP32, as the reference computes them
                               (graspcg/solvers/cg.py:89-90, graspcg/ops/dot.py:54-71) -- for
                               precision studies; Armijo search without a graph term only.
                               0: FP64 (default)                                          */
} mrh_cg_desc;

/* Per-problem, per-iteration record copied out by mrh_cg_read_history. */
typedef struct {
  double f0;        /* objective at the start of the iteration (f_base)      */
  double g0d;       /* Re<g, d> at t = 0                                     */
  double t;         /* accepted step                                         */
  double ft;        /* objective at the accepted step                        */
  double rho;       /* |g|^2 at the new iterate                              */
  double beta;      /* CG beta used for the next direction                   */
  double freg;      /* regulariser energy at the new iterate                 */
  double gdot_step; /* Re<g_new, d> (slope at the accepted step)              */
  double xnorm2;    /* |x_new|^2                                             */
  double dnorm2;    /* |d|^2 of the direction that was searched               */
} mrh_cg_hist;

/* Allocate the engine's per-problem scalars, partial sums and history.
 * Replaces the solver state of CGSolver.__init__ (graspcg/solvers/cg.py:65-157). */
int mrh_cg_create(const mrh_cg_desc* desc, mrh_cg** eng);
int mrh_cg_destroy(mrh_cg* eng);
/* Reset iteration counters / status (before a new run on the same shapes). */
int mrh_cg_reset(mrh_cg* eng, void* stream);

/* r = Ax - y over all problems (prime step, t = 0). */
int mrh_cg_residual(mrh_cg* eng, const void* Ax, const void* y, void* r, void* stream);

/* Device-side Armijo line search for every active problem:
 * f(t) = 0.5 (a + 2 b t + c t^2) + TV(x + t d) with a = |Ax - y|^2,
 * b = Re<Ax - y, Ad>, c = |Ad|^2, TV at t in {0, 1, beta, beta^2, ...} from one
 * pass over x and d, first t with f(t) <= f(0) + c1 t Re<g, d> accepted.
 * Accepts the same t as sequential backtracking
 * (graspcg/numerics/line_search.py:24-79 with Objective.f_g_tensor,
 * graspcg/ops/objective.py:184-291) without evaluating gradients at trials.
 * With ls_kind = MRH_CG_LS_WOLFE the strong-Wolfe search of line_search.py:89-204
 * (doubling from t = 1, bisection zoom, lo end accepted after ls_max_iter bisections):
 * f'(t) = b + c t + Re<g_TV(x + t d), d>, with TV values and slopes at seven candidate
 * steps per image pass; the sequential search is replayed over the evaluated steps, so it
 * accepts the same t as the reference's one-trial-at-a-time search.
 * y may be null: Ax then holds the residual r = Ax - y (as mrh_cg_residual and mrh_cg_accept
 * leave it), which gives the same a and b bits without reading Ax and y. */
int mrh_cg_linesearch(mrh_cg* eng, const void* x, const void* d, const void* g, const void* Ax,
                      const void* Ad, const void* y, void* stream);

/* x += t d, Ax += t Ad, r = Ax - y with each problem's accepted t
 * (graspcg/solvers/cg.py:205-207; objective residual of the accepted step). */
int mrh_cg_accept(mrh_cg* eng, void* x, const void* d, void* Ax, const void* Ad, const void* y,
                  void* r, void* stream);

/* g += TV gradient at x (g holds A^H r on entry); reduces |g|^2, Re<g_prev, g>,
 * Re<g, d> and the TV energy, then computes beta (FR with the first beta = 1,
 * or PR+) and the relative-change stop rule.  prime != 0 marks the prime step.
 * Replaces RegManager.energy_and_grad (graspcg/regularization/manager.py:38-142)
 * and the reductions of DirFR / DirPRPlus (graspcg/numerics/directions.py:75-162). */
int mrh_cg_gradient(mrh_cg* eng, const void* x, void* g, const void* d, const void* g_prev,
                    int prime, void* stream);

/* d = beta d - g (prime: d = -g); PR+ also copies g into g_prev.  Frozen for
 * problems whose status is not ACTIVE.  (graspcg/numerics/directions.py:158-161,
 * graspcg/solvers/cg.py:183-184) */
int mrh_cg_direction(mrh_cg* eng, void* d, const void* g, void* g_prev, int prime, void* stream);

/* Per-problem objective targets (host double[P]; NULL or NaN entries clear them):
 * a problem stops with MRH_CG_TARGET once an accepted step has f <= target.
 * Used to time how long a run takes to reach a given objective (e.g. another
 * run's final value); not a reference stop rule.  Synchronises `stream`. */
int mrh_cg_set_targets(mrh_cg* eng, const double* targets, void* stream);

/* Fill desc.diag with the regularisers' diagonal majoriser at x: 1 plus, per TV term, its
 * frame-axis profile 2 lambda w_0^2 x degree (a term over frames) or the lagged-diffusivity
 * weights lambda w_a^2 / sqrt(|D_a x|^2 + eps^2) x degree per axis (any other term).
 * Replaces RegManager.add_diag (graspcg/regularization/manager.py:145-218, TVND
 * majorizer_profile / add_diag at graspcg/regularization/tv_nd.py:113-215). */
int mrh_cg_build_diag(mrh_cg* eng, const void* x, void* stream);

/* Device int[P]: 1 while problem p iterates, 0 once it stopped (updated by the
 * engine's kernels; reset to 1 by mrh_cg_reset).  Owned by the engine. */
int mrh_cg_active_flags(mrh_cg* eng, const int** flags);

/* Copy status (int[P]) and iteration counts (int[P]) to host (synchronising `stream`). */
int mrh_cg_read_status(mrh_cg* eng, int* status, int* iters, void* stream);
/* Copy history (host mrh_cg_hist[n_iter * P], iteration-major) to host. */
int mrh_cg_read_history(mrh_cg* eng, mrh_cg_hist* hist, int n_iter, void* stream);
/* Status (int[P]), iteration counts (int[P]) and each problem's latest complete history
 * record (host mrh_cg_hist[P]; zeros before the first iteration) in one copy and one stream
 * synchronisation: what the reference's per-iteration convergence test reads
 * (graspcg/solvers/cg.py:228-244 with graspcg/solvers/utils.py:40-43).  Any pointer may be NULL. */
int mrh_cg_read_latest(mrh_cg* eng, int* status, int* iters, mrh_cg_hist* last, void* stream);

/* ------------------------------------------------------------------ U-Net warm start
 *
 * No reference code: the reference's only warm-start hook is init_mode="keep",
 * which keeps a caller-filled x as the initial point (graspcg/solvers/cg.py:48,142).
 * These kernels train and run the 2D+t U-Net that fills it.
 *
 * Activations are BF16 (16-bit storage, pointers typed void*) channels-last
 * (nvol * F, H, W, C).  Forward weights are BF16 [ntap][Cout_pad][Cin]; the
 * backward-data copy is BF16 [ntap][Cin][Cout] with taps in reverse order.
 * Weight and bias gradients, master weights and Adam moments are FP32.
 * Temporal taps never cross a volume's first or last frame; every tap set
 * zero-pads at the borders. */
#define MRH_UNET_TAP_1X1 0   /* one tap */
#define MRH_UNET_TAP_3X3 1   /* 9 spatial taps, t = 3 (dy + 1) + (dx + 1) */
#define MRH_UNET_TAP_T3 2    /* 3 temporal taps, t = df + 1 */

/* out[p][co] = act(bias[co] + sum_(t, ci) w[t][co][ci] x[p + off(t)][ci]),
 * act = ReLU if relu != 0.  Cin % 8 == 0, Cout <= Cout_pad, Cout_pad % 32 == 0,
 * bias (FP32, may be NULL) holds Cout values.  With w = the backward-data copy,
 * x = dL/dout (Cin := Cout), Cout := Cin, bias NULL and relu 0 it computes dL/dx. */
int mrh_unet_conv_fwd(const void* x, const void* w, const float* bias, void* out, int kind, int nvol,
                      int F, int H, int W, int Cin, int Cout, int Cout_pad, int relu, void* stream);
/* dw[t][co][ci] += sum_p gout[p][co] x[p + off(t)][ci]; db[co] += sum_p gout[p][co]
 * (db may be NULL).  FP32 atomics: the result is reproducible to rounding only. */
int mrh_unet_conv_bwd_weight(const void* gout, const void* x, float* dw, float* db, int kind, int nvol,
                             int F, int H, int W, int Cin, int Cout, int Cout_pad, void* stream);
/* g[i] = 0 where y[i] <= 0 (y = the ReLU output), n elements (n % 8 == 0). */
int mrh_unet_relu_bwd(void* g, const void* y, int64_t n, void* stream);
/* 2 x 2 average pooling (nf, H, W, C) -> (nf, H/2, W/2, C) and its adjoint
 * gx = pool^T gy (+ add, which may be NULL).  C % 8 == 0 (as for C1, C2 below). */
int mrh_unet_pool2(const void* x, void* y, int nf, int H, int W, int C, void* stream);
int mrh_unet_pool2_bwd(const void* gy, const void* add, void* gx, int nf, int H, int W, int C, void* stream);
/* dst (nf, H, W, C1 + C2) = concat(nearest 2x upsampling of src (nf, H/2, W/2, C1),
 * skip (nf, H, W, C2)), and its adjoint (gsrc = up^T gdst[..., :C1], gskip = gdst[..., C1:]). */
int mrh_unet_up2_concat(const void* src, const void* skip, void* dst, int nf, int H, int W, int C1,
                        int C2, void* stream);
int mrh_unet_up2_concat_bwd(const void* gdst, void* gsrc, void* gskip, int nf, int H, int W, int C1,
                            int C2, void* stream);
/* 3x3 convolution of the concatenation above without storing it: out = conv_fwd of
 * concat(up2(src), skip) with Cin = C1 + C2 (same weights layout, bias and ReLU as
 * mrh_unet_conv_fwd, identical result).  For inference; MRH_EUNSUP unless
 * mrh_unet_upcat_supported(H, W, C1, C2) returns 1 (C1, C2 multiples of 32 and an
 * image that the 2-D tiles cover). */
int mrh_unet_conv3x3_upcat(const void* src, const void* skip, const void* w, const float* bias, void* out,
                           int nvol, int F, int H, int W, int C1, int C2, int Cout, int Cout_pad, int relu,
                           void* stream);
int mrh_unet_upcat_supported(int H, int W, int C1, int C2);
/* complex64 image (npix) * scale -> BF16 (npix, C) channels (re, im, 0, ...). */
int mrh_unet_pack_input(const void* img, void* out, int64_t npix, int C, float scale, void* stream);
/* xout = xin + (net[:, 0] + i net[:, 1]) * inv_scale (complex64; net is BF16 (npix, Cn)). */
int mrh_unet_unpack_residual(const void* xin, const void* net, void* xout, int64_t npix, int Cn,
                             float inv_scale, void* stream);
/* Inference with the full-resolution level fused (no reference counterpart; one volume of F
 * frames per call, 32 channels at that level; MRH_EUNSUP unless
 * mrh_unet_level0_fused_supported(H, W) returns 1, i.e. H % 8 == 0 and W % 32 == 0).
 * mrh_unet_enc0_fused: e0 = relu(convT(relu(conv3x3(pack(x))))) and i1 = pool2(e0), where
 *   pack(x) = the 8-channel input of mrh_unet_pack_input (x complex64 (F, H, W), scale),
 *   ws/bs the 3x3 layer (BF16 [9][32][8], FP32 bias[32]), wt/bt the temporal layer
 *   (BF16 [3][32][32], bias[32]); e0 BF16 (F, H, W, 32), i1 BF16 (F, H/2, W/2, 32).  The
 *   3x3 layer's output never reaches memory.
 * mrh_unet_dec0_tail: xout = xin + (out[:, 0] + i out[:, 1]) * inv_scale with
 *   out = conv1x1(relu(convT(b0))): b0 BF16 (F, H, W, 32), wt/bt the temporal layer,
 *   wo/bo the 1x1 layer (BF16 [1][32][32] with rows past Cout zero, bias >= 2 values); the
 *   temporal layer's output and out never reach memory.
 * Each equals the chain of mrh_unet_conv_fwd / pool2 / pack / unpack calls it replaces up
 * to the FP32 summation order inside the MFMA tiles (activations are rounded to BF16 at the
 * same points). */
int mrh_unet_level0_fused_supported(int H, int W);
int mrh_unet_enc0_fused(const void* x, float scale, const void* ws, const float* bs, const void* wt,
                        const float* bt, void* e0, void* i1, int F, int H, int W, void* stream);
int mrh_unet_dec0_tail(const void* b0, const void* wt, const float* bt, const void* wo, const float* bo,
                       const void* xin, float inv_scale, void* xout, int F, int H, int W, void* stream);
/* loss += weight * sum |(xout - xt) scale|^2 (FP64, device scalar) and
 * gnet[:, 0:2] = 2 weight (xout - xt) scale + gx_coef (Re gx, Im gx),
 * gnet[:, 2:] = 0 (BF16 (npix, Cn)).  gx (complex64, may be NULL) is the
 * gradient of an extra loss term with respect to xout, e.g. the NLCG
 * objective's A^H (A xout - y) + grad TV(xout). */
int mrh_unet_mse_grad(const void* xout, const void* xt, const void* gx, float gx_coef, void* gnet,
                      double* loss, int64_t npix, int Cn, float scale, float weight, void* stream);
/* p[0:n] = value (FP32); used to clear gradient accumulators. */
int mrh_unet_fill(float* p, int64_t n, float value, void* stream);
/* Adam step (bias-corrected, step >= 1) on the FP32 master weights wm and bias,
 * then refresh the BF16 forward copy wf and, if wbd != NULL, the backward-data
 * copy.  lr == 0 only refreshes the copies (gw, m, v may then be NULL). */
int mrh_unet_adam(float* wm, const float* gw, float* m, float* v, void* wf, void* wbd, float* bias,
                  const float* gb, float* mb, float* vb, int ntap, int Cout, int Cout_pad, int Cin,
                  float lr, float beta1, float beta2, float eps, int step, void* stream);

/* ---------------------------------------------------------------- low-rank temporal basis
 * u = U V with U (T, K) complex64 row-major (frame t's weights on K basis images), V (K, N)
 * complex64 coefficient images and u (T, N) complex64 frames; N counts every non-frame element.
 * Replaces TemporalBasisOp (graspcg/regularization/mapping.py:112-231).  Fixed-order FP32 sums
 * (bitwise repeatable); 1 <= K <= MRH_TBASIS_MAX_K, T >= 1, N >= 0. */
#define MRH_TBASIS_MAX_K 16
/* u = U V.  Replaces TemporalBasisOp.forward_apply (mapping.py:129-147). */
int mrh_tbasis_forward(const void* U, int T, int K, const void* V, int64_t N, void* u, void* stream);
/* gV = U^H gu, or gV += U^H gu with accumulate != 0 (the reference adds into g_V).
 * Replaces TemporalBasisOp.adjoint_apply (mapping.py:149-164). */
int mrh_tbasis_adjoint(const void* U, int T, int K, const void* gu, int64_t N, void* gV, int accumulate,
                       void* stream);
/* diag_V[k, n] += sum_t |U[t, k]|^2 ku[t, n] (FP32).  ku is a (T, N) field (profile = 0,
 * TemporalBasisOp.diag_push, mapping.py:170-199) or a per-frame profile of T values broadcast
 * over n (profile != 0, TemporalBasisOp.diag_push_profile, mapping.py:200-231). */
int mrh_tbasis_diag(const void* U, int T, int K, const float* ku, int profile, int64_t N, float* diag_V,
                    void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MRIRECON_HIP_H */
