// Objective, Huber-TV and device-side NLCG kernels for MI355X / gfx950.
//
// All reductions accumulate in float64 per thread, reduce per block in a fixed
// order and are finished by a second kernel that sums the block partials in a
// fixed order: results are bit-reproducible run to run.
//
// TV follows the reference stencils (graspcg/ops/ndops.py:28-149): forward
// difference with the last slice zero, divergence = backward difference with
// the first slice kept, isotropic energy sqrt(sum |w_a D_a x|^2 + eps^2) - eps,
// flux w_a D_a x / max(den, 1e-12), g += -lambda div(flux)
// (graspcg/regularization/tv_nd.py:70-88).  Temporal differences stay inside
// each problem (slice) of F frames.
#include "mrh_common.h"
#include "mrirecon_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#define RT 256            // threads per reduction block
#define RED_BLOCKS 2048   // blocks of the standalone reductions (8 per CU)
#define NT 21             // line-search candidates: t = 0, 1, beta, ..., beta^19
#define NT_A 6            // candidates of the first image pass: t = 0, 1, ..., beta^4
#define TG_T 16           // side of the pixel tiles of the gradient and line-search image passes
#define PI_STRIDE 32      // doubles per block partial (image pass)
#define PK_STRIDE 4       // doubles per block partial (k-space pass)
#define NW 7              // strong-Wolfe candidates per image pass
#define PQ_STRIDE 4       // doubles per block partial (graph-Laplacian passes)
#define GL_MAXF 48        // frames per problem of a graph-Laplacian term (LDS of k_gl_ls <= 64 KiB)
#define GL_TILE 64        // pixels per workgroup tile of the graph-Laplacian passes

// ---------------------------------------------------------------- block reduction
// R is the accumulation type: double (the engine's FP64 reductions) or float (the all-FP32
// variant of the precision study, whose partials are stored as doubles holding float values).
template <typename R>
__device__ __forceinline__ R warp_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NV, typename R = double>
__device__ __forceinline__ void block_reduce_store(R (&v)[NV], double* out) {
  __shared__ R sh[NV][RT / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    R w = warp_sum<R>(v[k]);
    if (lane == 0) sh[k][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    R s = R(0);
#pragma unroll
    for (int w = 0; w < RT / 64; ++w) s += sh[threadIdx.x][w];
    out[threadIdx.x] = (double)s;
  }
}

// Sum n partials (stride `stride`, value offset `v`) in a fixed order with one block.
__device__ double block_sum_partials(const double* part, int n, int stride, int v) {
  __shared__ double sh[RT];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += RT) s += part[(int64_t)i * stride + v];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int h = RT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// Sums of values [v0, v0 + NV) over n block partials (stride doubles apart) in one
// sweep: each thread adds its partials for every value, then one block reduction.
// Thread 0 receives the sums in out.
template <int NV, typename R = double>
__device__ void block_sum_partials_n(const double* part, int n, int stride, int v0, double* out) {
  __shared__ R sh[NV][RT / 64];
  R s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = R(0);
  for (int i = threadIdx.x; i < n; i += RT) {
    const double* q = part + (int64_t)i * stride + v0;
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] += (R)q[k];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const R w = warp_sum<R>(s[k]);
    if (lane == 0) sh[k][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      R t = R(0);
#pragma unroll
      for (int w = 0; w < RT / 64; ++w) t += sh[k][w];
      out[k] = (double)t;
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(RT) void k_finish_sum(const double* part, int n, double* out) {
  double s = block_sum_partials(part, n, 1, 0);
  if (threadIdx.x == 0) out[0] = s;
}

extern "C" int64_t mrh_reduce_workspace_len(void) { return (int64_t)RED_BLOCKS * 64; }

static int red_blocks(int64_t n) {
  int64_t b = (n + RT * 2 - 1) / (RT * 2);
  return (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, b));
}

// ---------------------------------------------------------------- standalone ops
__global__ __launch_bounds__(RT) void k_residual_energy(float2* r, const float2* Ax, const float2* Ad,
                                                        const float2* y, float t, int64_t n, double* part) {
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    float2 a = Ad[i];
    a.x *= t; a.y *= t;
    const float2 ax = Ax[i], yy = y[i];
    a.x += ax.x; a.y += ax.y;
    a.x -= yy.x; a.y -= yy.y;
    r[i] = a;
    v[0] += (double)a.x * a.x + (double)a.y * a.y;
  }
  block_reduce_store<1>(v, part + blockIdx.x);
}

extern "C" int mrh_residual_energy(void* r, const void* Ax, const void* Ad, const void* y, double t,
                                   int64_t n, double* work, double* out, void* stream) {
  MRH_REQUIRE(r && Ax && Ad && y && work && out, MRH_EINVAL, "residual_energy: null argument");
  MRH_REQUIRE(n >= 0, MRH_EINVAL, "residual_energy: n < 0");
  hipStream_t st = (hipStream_t)stream;
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_residual_energy, dim3(nb), dim3(RT), 0, st, (float2*)r, (const float2*)Ax,
                     (const float2*)Ad, (const float2*)y, (float)t, n, work);
  MRH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(RT), 0, st, (const double*)work, nb, out);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

__global__ __launch_bounds__(RT) void k_dot_real(const float2* a, const float2* b, int64_t n, double* part) {
  double v[1] = {0.0};
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
    const float2 p = a[i], q = b[i];
    v[0] += (double)p.x * q.x + (double)p.y * q.y;
  }
  block_reduce_store<1>(v, part + blockIdx.x);
}

extern "C" int mrh_dot_real(const void* a, const void* b, int64_t n, double* work, double* out, void* stream) {
  MRH_REQUIRE(a && b && work && out, MRH_EINVAL, "dot_real: null argument");
  hipStream_t st = (hipStream_t)stream;
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot_real, dim3(nb), dim3(RT), 0, st, (const float2*)a, (const float2*)b, n, work);
  MRH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(RT), 0, st, (const double*)work, nb, out);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

// ---------------------------------------------------------------- TV device helpers
struct TVTermDev {
  float lam, eps, w[3];
  int axes, iso;
};
struct TVDev {
  int n, F, S0, S1;
  TVTermDev t[MRH_TV_MAX_TERMS];
};

static int make_tvdev(const mrh_tv_desc* d, TVDev* o) {
  MRH_REQUIRE(d->P >= 1 && d->F >= 1 && d->S0 >= 1 && d->S1 >= 1, MRH_EINVAL, "tv: bad layout");
  MRH_REQUIRE(d->n_terms >= 0 && d->n_terms <= MRH_TV_MAX_TERMS, MRH_EINVAL, "tv: n_terms outside [0,3]");
  std::memset(o, 0, sizeof(*o));
  o->n = d->n_terms; o->F = d->F; o->S0 = d->S0; o->S1 = d->S1;
  for (int i = 0; i < d->n_terms; ++i) {
    const mrh_tv_term& t = d->term[i];
    MRH_REQUIRE((t.axes & ~7) == 0, MRH_EINVAL, "tv: bad axes mask %d", t.axes);
    o->t[i].lam = t.weight; o->t[i].eps = t.eps; o->t[i].axes = t.axes; o->t[i].iso = t.isotropic;
    for (int a = 0; a < 3; ++a) o->t[i].w[a] = t.axis_w[a];
  }
  return MRH_OK;
}

// Weighted forward difference along axis a (0: frame, 1: X, 2: Y) at (f, ix, iy), n = flat index.
__device__ __forceinline__ float2 fdiff(const float2* __restrict__ v, int64_t n, int f, int ix, int iy,
                                        int a, const TVDev& tv, float w) {
  int64_t nn;
  if (a == 0) { if (f + 1 >= tv.F) return make_float2(0.f, 0.f); nn = n + (int64_t)tv.S0 * tv.S1; }
  else if (a == 1) { if (ix + 1 >= tv.S0) return make_float2(0.f, 0.f); nn = n + tv.S1; }
  else { if (iy + 1 >= tv.S1) return make_float2(0.f, 0.f); nn = n + 1; }
  const float2 p = v[nn], q = v[n];
  return make_float2(w * (p.x - q.x), w * (p.y - q.y));
}

// Flux components p_a and energy density of one term from its weighted forward differences dv.
// Contraction is off and the fused multiply-adds are written out: a flux enters the divergence of
// its own pixel and of its forward neighbours, and every caller (the standalone TV kernel, which
// evaluates it again at the neighbour, and the engine's tiled gradient, which shares it through
// LDS) then gets the same rounded values.
// The denominators sqrt(|z|^2 + eps^2) and their reciprocals come from the hardware square root and
// reciprocal square root (v_sqrt_f32, v_rsq_f32: 1 ulp each) instead of the correctly rounded
// sqrtf and division, about 25 instructions fewer per flux; 1 / max(den, 1e-12) as
// rsq(max(den^2, 1e-24)).
__device__ __forceinline__ float abs2_plus(float2 v, float s) { return fmaf(v.x, v.x, fmaf(v.y, v.y, s)); }
__device__ __forceinline__ float hw_sqrt(float t) { return __builtin_amdgcn_sqrtf(t); }
__device__ __forceinline__ float hw_rsq_guarded(float t) { return __builtin_amdgcn_rsqf(fmaxf(t, 1e-24f)); }

__device__ __forceinline__ float flux_of(const float2 (&dv)[3], const TVTermDev& tm, float2 (&p)[3]) {
#pragma clang fp contract(off)
  float ss = 0.f;
  const float e2 = tm.eps * tm.eps;
  float e = 0.f;
  if (tm.iso) {
#pragma unroll
    for (int a = 0; a < 3; ++a) ss = abs2_plus(dv[a], ss);
    const float t = ss + e2;
    e = hw_sqrt(t) - tm.eps;
    const float inv = hw_rsq_guarded(t);
#pragma unroll
    for (int a = 0; a < 3; ++a) p[a] = make_float2(dv[a].x * inv, dv[a].y * inv);
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (tm.axes & (1 << a)) {
        const float t = abs2_plus(dv[a], e2);
        e += hw_sqrt(t) - tm.eps;
        const float inv = hw_rsq_guarded(t);
        p[a] = make_float2(dv[a].x * inv, dv[a].y * inv);
      } else {
        p[a] = make_float2(0.f, 0.f);
      }
    }
  }
  return e;
}

// Flux components p_a and energy density of one term at (f, ix, iy).
__device__ __forceinline__ float term_flux(const float2* __restrict__ x, int64_t n, int f, int ix, int iy,
                                           const TVTermDev& tm, const TVDev& tv, float2 (&p)[3]) {
  float2 dv[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    dv[a] = make_float2(0.f, 0.f);
    if (tm.axes & (1 << a)) dv[a] = fdiff(x, n, f, ix, iy, a, tv, tm.w[a]);
  }
  return flux_of(dv, tm, p);
}

// TV gradient contribution (-lambda div flux) and lambda * energy density of all terms at pixel n.
__device__ __forceinline__ float2 tv_grad_at(const float2* __restrict__ x, int64_t n, int f, int ix, int iy,
                                             const TVDev& tv, float& energy) {
  float2 gsum = make_float2(0.f, 0.f);
  energy = 0.f;
  for (int k = 0; k < tv.n; ++k) {
    const TVTermDev& tm = tv.t[k];
    if (tm.lam == 0.f || tm.axes == 0) continue;
    float2 p[3];
    const float e = term_flux(x, n, f, ix, iy, tm, tv, p);
    float2 div = make_float2(0.f, 0.f);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (!(tm.axes & (1 << a))) continue;
      div = cadd(div, p[a]);
      int fb = f, xb = ix, yb = iy;
      int64_t nb;
      bool has;
      if (a == 0) { has = f > 0; fb = f - 1; nb = n - (int64_t)tv.S0 * tv.S1; }
      else if (a == 1) { has = ix > 0; xb = ix - 1; nb = n - tv.S1; }
      else { has = iy > 0; yb = iy - 1; nb = n - 1; }
      if (has) {
        float2 pb[3];
        (void)term_flux(x, nb, fb, xb, yb, tm, tv, pb);
        div = csub(div, pb[a]);
      }
    }
    gsum.x = fmaf(-tm.lam, div.x, gsum.x);
    gsum.y = fmaf(-tm.lam, div.y, gsum.y);
    energy = fmaf(tm.lam, e, energy);
  }
  return gsum;
}

// ---------------------------------------------------------------- standalone TV
__global__ __launch_bounds__(RT) void k_tv_energy_grad(const float2* __restrict__ x, float2* __restrict__ g,
                                                       TVDev tv, int64_t npix_p, double* part) {
  const int p = blockIdx.y;
  double v[1] = {0.0};
  const int plane = tv.S0 * tv.S1;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < npix_p; i += (int64_t)gridDim.x * RT) {
    const int f = (int)(i / plane), rem = (int)(i - (int64_t)f * plane);
    const int ix = rem / tv.S1, iy = rem - ix * tv.S1;
    const int64_t n = (int64_t)p * npix_p + i;
    float e;
    const float2 gt = tv_grad_at(x, n, f, ix, iy, tv, e);
    if (g) { float2 gv = g[n]; g[n] = cadd(gv, gt); }
    v[0] += (double)e;
  }
  block_reduce_store<1>(v, part + (int64_t)p * gridDim.x + blockIdx.x);
}

__global__ __launch_bounds__(RT) void k_finish_per_problem(const double* part, int nb, double* out) {
  const int p = blockIdx.x;
  double s = block_sum_partials(part + (int64_t)p * nb, nb, 1, 0);
  if (threadIdx.x == 0) out[p] = s;
}

extern "C" int mrh_tv_energy_grad(const mrh_tv_desc* desc, const void* x, void* g, double* work,
                                  double* energy, void* stream) {
  MRH_REQUIRE(desc && x && work && energy, MRH_EINVAL, "tv_energy_grad: null argument");
  TVDev tv;
  int rc = make_tvdev(desc, &tv);
  if (rc) return rc;
  const int64_t npix_p = (int64_t)desc->F * desc->S0 * desc->S1;
  int nb = red_blocks(npix_p);
  nb = std::max(1, std::min<int>(nb, (int)(mrh_reduce_workspace_len() / desc->P)));
  MRH_REQUIRE(desc->P <= 65535, MRH_EINVAL, "tv: too many problems");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_tv_energy_grad, dim3(nb, desc->P), dim3(RT), 0, st, (const float2*)x, (float2*)g,
                     tv, npix_p, work);
  MRH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finish_per_problem, dim3(desc->P), dim3(RT), 0, st, (const double*)work, nb, energy);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

// ---------------------------------------------------------------- TV without allocation
// mrh_tv_energy_grad_noalloc keeps its block partials in this file-scope array instead of a
// caller's workspace.  Host code must not pass `tv_scratch` itself to a kernel (in host code the
// name is a stand-in, not a device address): each device's address comes from
// hipGetSymbolAddress, fetched once per device on its first call and kept below.
#define TV_SCRATCH_DOUBLES (RED_BLOCKS * 64)   // 1 MiB per device, as mrh_reduce_workspace_len()
#define TV_SCRATCH_MAX_DEVICES 64
static __device__ __align__(256) double tv_scratch[TV_SCRATCH_DOUBLES];
static double* g_tv_scratch_ptr[TV_SCRATCH_MAX_DEVICES];
static std::mutex g_tv_scratch_mu;

static int tv_scratch_for_current_device(double** out) {
  int dev = -1;
  MRH_HIP(hipGetDevice(&dev));
  MRH_REQUIRE(dev >= 0 && dev < TV_SCRATCH_MAX_DEVICES, MRH_EUNSUP, "tv_energy_grad_noalloc: device %d beyond %d",
              dev, TV_SCRATCH_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_tv_scratch_mu);
  if (!g_tv_scratch_ptr[dev]) {
    void* p = nullptr;
    MRH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(tv_scratch)));
    g_tv_scratch_ptr[dev] = (double*)p;
  }
  *out = g_tv_scratch_ptr[dev];
  return MRH_OK;
}

extern "C" int mrh_tv_energy_grad_noalloc(const mrh_tv_desc* desc, const void* x, void* g, double* energy,
                                          void* stream) {
  MRH_REQUIRE(desc && x && energy, MRH_EINVAL, "tv_energy_grad_noalloc: null argument");
  MRH_REQUIRE(desc->P >= 1 && desc->P <= 65535, MRH_EINVAL,
              "tv_energy_grad_noalloc: %d problems; the static scratch serves 1 to 65535", desc->P);
  TVDev tv;
  int rc = make_tvdev(desc, &tv);
  if (rc) return rc;
  double* work = nullptr;
  rc = tv_scratch_for_current_device(&work);
  if (rc) return rc;
  // the same partition as mrh_tv_energy_grad (whose workspace has the same length), so both
  // entry points return identical sums; every partial is written before it is read
  const int64_t npix_p = (int64_t)desc->F * desc->S0 * desc->S1;
  int nb = red_blocks(npix_p);
  nb = std::max(1, std::min<int>(nb, (int)(TV_SCRATCH_DOUBLES / desc->P)));
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_tv_energy_grad, dim3(nb, desc->P), dim3(RT), 0, st, (const float2*)x, (float2*)g, tv,
                     npix_p, work);
  MRH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finish_per_problem, dim3(desc->P), dim3(RT), 0, st, (const double*)work, nb, energy);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

// ---------------------------------------------------------------- graph Laplacian
// Graph-Laplacian term along the frame axis (graspcg/regularization/graph_laplacian.py:22-321):
// E = (lambda / 2) sum_p Re<x[:, p], (L x)[:, p]>, gradient lambda L x, L (F x F) the plain,
// random-walk or symmetric Laplacian of the term's symmetrised graph (built on the host).
// A workgroup stages L and a tile of GL_TILE pixels x F frames in LDS (dynamic shared memory);
// wavefront w computes rows w, w + 4, ... of L x for its lane's pixel.  Along d the energy is
// the quadratic (lambda / 2)(a + t (b1 + b2) + t^2 c), so a line search needs one pass.
struct GLDev {
  int F;
  float lam;          // 0: no term
  const float* L;     // F x F, row-major
  const float* prof;  // F: the term's diagonal majoriser per frame
};

__host__ __device__ constexpr int gl_lds_floats(int F, int tiles) { return ((F * F + 1) & ~1) + tiles * 2 * F * GL_TILE; }

// ---------------------------------------------------------------- NLCG engine
struct ProbState {
  double a, b, c, g0d, dd;
  double gq_a, gq_b1, gq_b2, gq_c;   // graph term along d: Re<x,Lx>, Re<d,Lx>, Re<x,Ld>, Re<d,Ld>
  double tv[NT];
  double f0, t, ft, rho, rho_prev, beta, gpg, freg, gdot_step, f_prev;
  // strong-Wolfe search: candidates of the next image pass, the expansion's next step and
  // previous trial (t, f, f'), the zoom bracket (lo with f, f'; hi)
  double wt[NW];
  double w_t, w_tp, w_fp, w_gp, w_lo, w_flo, w_glo, w_hi;
  int status, iter, have_rho_prev, have_f_prev;
  int ls_pending;   // Armijo: the first pass found no step; Wolfe: the next pass has candidates
  int wn, wzoom, wi; // Wolfe: candidates in wt, zooming, trials (expansion) or bisections (zoom) done
};

struct mrh_cg {
  int P, F, S0, S1, direction, ls_max_iter, max_hist, ls_kind;
  double c1, ls_beta, rel_tol, restart_c, c2;
  int64_t npix_p, kmax_p;
  int nb_i, nb_k;
  int tg_tiles_x, tg_ntiles;   // tiled passes: tiles per row, per frame
  int tg_fg, nb_g;             // k_cg_grad_tiled: frames per group, blocks per problem
  int tl_fg, nb_l;             // k_ls_image: frames per group, blocks per problem
  TVDev tv;
  int64_t* d_kprob;
  double* d_part_i;   // [P][nb_i][PI_STRIDE]
  double* d_part_k;   // [P][nb_k][PK_STRIDE]
  ProbState* d_st;
  ProbState* h_st;      // pinned host copy for status reads (one async copy + one stream sync)
  mrh_cg_hist* d_hist;  // [max_hist][P]
  mrh_cg_hist* d_last;  // [P] each problem's latest complete record (mrh_cg_read_latest)
  mrh_cg_hist* h_last;  // pinned host copy of d_last
  double* d_target;     // [P] objective targets (-inf: none)
  int* d_active;        // [P] 1 while the problem iterates (read by the masked NUFFT entry points)
  float* d_diag;        // caller's diagonal preconditioner (P F S0 S1 floats) or null
  GLDev gl;             // graph-Laplacian term (gl.lam == 0: none); L and its profile in d_gl
  float* d_gl;
  double* d_part_q;     // [P][nb_q][PQ_STRIDE] graph-term partials
  int nb_q;
  size_t gl_lds_grad, gl_lds_ls;   // dynamic LDS bytes of k_gl_grad / k_gl_ls
  int fp32;             // all-FP32 reductions (precision study), else FP64
};

// Launch kernel template K<R> with R = float for an engine with FP32 reductions, else double.
#define LAUNCH_R(e, K, ...)                                                            \
  do {                                                                                 \
    if ((e)->fp32) hipLaunchKernelGGL(K<float>, __VA_ARGS__);                          \
    else hipLaunchKernelGGL(K<double>, __VA_ARGS__);                                   \
  } while (0)
// Same with a second (bool) template argument.
#define LAUNCH_R2(e, K, B, ...)                                                        \
  do {                                                                                 \
    if ((e)->fp32) hipLaunchKernelGGL((K<float, B>), __VA_ARGS__);                     \
    else hipLaunchKernelGGL((K<double, B>), __VA_ARGS__);                              \
  } while (0)

// g += lambda L x; block partial: E.
__global__ __launch_bounds__(256) void k_gl_grad(const float2* __restrict__ x, float2* __restrict__ g, GLDev gl,
                                                 int plane, int64_t npix_p, const ProbState* __restrict__ st,
                                                 int prime, double* part) {
  extern __shared__ float gl_lds[];
  const int F = gl.F, p = blockIdx.y;
  float* Ls = gl_lds;
  float2* Xs = reinterpret_cast<float2*>(gl_lds + ((F * F + 1) & ~1));
  double v[1] = {0.0};
  if (prime || st[p].status == MRH_CG_ACTIVE) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < F * F; i += 256) Ls[i] = gl.L[i];
    const float2* xp = x + (int64_t)p * npix_p;
    float2* gp = g + (int64_t)p * npix_p;
    for (int t0 = blockIdx.x * GL_TILE; t0 < plane; t0 += gridDim.x * GL_TILE) {
      const int pix = t0 + lane;
      const bool ok = pix < plane;
      __syncthreads();                       // L staged / the previous tile consumed
      for (int f = wv; f < F; f += 4)
        Xs[f * GL_TILE + lane] = ok ? xp[(int64_t)f * plane + pix] : make_float2(0.f, 0.f);
      __syncthreads();
      if (ok) {
        for (int i = wv; i < F; i += 4) {
          const float* Li = Ls + i * F;
          float yr = 0.f, yi = 0.f;
          for (int j = 0; j < F; ++j) {
            const float l = Li[j];
            const float2 xj = Xs[j * GL_TILE + lane];
            yr = fmaf(l, xj.x, yr);
            yi = fmaf(l, xj.y, yi);
          }
          const float2 xi = Xs[i * GL_TILE + lane];
          v[0] += 0.5 * (double)gl.lam * ((double)xi.x * yr + (double)xi.y * yi);
          const int64_t n = (int64_t)i * plane + pix;
          float2 gv = gp[n];
          gv.x += gl.lam * yr;
          gv.y += gl.lam * yi;
          gp[n] = gv;
        }
      }
    }
  }
  block_reduce_store<1>(v, part + ((int64_t)p * gridDim.x + blockIdx.x) * PQ_STRIDE);
}

// Line-search pass: block partials of Re<x, Lx>, Re<d, Lx>, Re<x, Ld>, Re<d, Ld>.
__global__ __launch_bounds__(256) void k_gl_ls(const float2* __restrict__ x, const float2* __restrict__ d,
                                               GLDev gl, int plane, int64_t npix_p,
                                               const ProbState* __restrict__ st, double* part) {
  extern __shared__ float gl_lds[];
  const int F = gl.F, p = blockIdx.y;
  float* Ls = gl_lds;
  float2* Xs = reinterpret_cast<float2*>(gl_lds + ((F * F + 1) & ~1));
  float2* Ds = Xs + F * GL_TILE;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (st[p].status == MRH_CG_ACTIVE) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < F * F; i += 256) Ls[i] = gl.L[i];
    const float2* xp = x + (int64_t)p * npix_p;
    const float2* dp = d + (int64_t)p * npix_p;
    for (int t0 = blockIdx.x * GL_TILE; t0 < plane; t0 += gridDim.x * GL_TILE) {
      const int pix = t0 + lane;
      const bool ok = pix < plane;
      __syncthreads();
      for (int f = wv; f < F; f += 4) {
        Xs[f * GL_TILE + lane] = ok ? xp[(int64_t)f * plane + pix] : make_float2(0.f, 0.f);
        Ds[f * GL_TILE + lane] = ok ? dp[(int64_t)f * plane + pix] : make_float2(0.f, 0.f);
      }
      __syncthreads();
      if (ok) {
        for (int i = wv; i < F; i += 4) {
          const float* Li = Ls + i * F;
          float xr = 0.f, xim = 0.f, dr = 0.f, dim = 0.f;
          for (int j = 0; j < F; ++j) {
            const float l = Li[j];
            const float2 xj = Xs[j * GL_TILE + lane], dj = Ds[j * GL_TILE + lane];
            xr = fmaf(l, xj.x, xr); xim = fmaf(l, xj.y, xim);
            dr = fmaf(l, dj.x, dr); dim = fmaf(l, dj.y, dim);
          }
          const float2 xi = Xs[i * GL_TILE + lane], di = Ds[i * GL_TILE + lane];
          v[0] += (double)xi.x * xr + (double)xi.y * xim;
          v[1] += (double)di.x * xr + (double)di.y * xim;
          v[2] += (double)xi.x * dr + (double)xi.y * dim;
          v[3] += (double)di.x * dr + (double)di.y * dim;
        }
      }
    }
  }
  block_reduce_store<4>(v, part + ((int64_t)p * gridDim.x + blockIdx.x) * PQ_STRIDE);
}

// The graph term's value and slope along d at step t (zero without a term).
__device__ __forceinline__ double gl_value(const ProbState& s, double lam, double t) {
  return 0.5 * lam * (s.gq_a + t * (s.gq_b1 + s.gq_b2) + t * t * s.gq_c);
}
__device__ __forceinline__ double gl_slope(const ProbState& s, double lam, double t) {
  return lam * (s.gq_b1 + t * s.gq_c);
}

// Sum the line-search pass's graph partials into the state (zeros without a term).
__device__ __forceinline__ void gl_collect(ProbState& s, const double* part_q, int nb_q, int p) {
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  if (part_q) block_sum_partials_n<4>(part_q + (int64_t)p * nb_q * PQ_STRIDE, nb_q, PQ_STRIDE, 0, q);
  if (threadIdx.x == 0) { s.gq_a = q[0]; s.gq_b1 = q[1]; s.gq_b2 = q[2]; s.gq_c = q[3]; }
}

__global__ void k_cg_reset(ProbState* st, int* active, int P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  active[p] = 1;
  ProbState s;
  double* z = (double*)&s;
  for (size_t i = 0; i < sizeof(ProbState) / sizeof(double); ++i) z[i] = 0.0;
  s.status = MRH_CG_ACTIVE; s.iter = 0; s.have_rho_prev = 0; s.have_f_prev = 0;
  st[p] = s;
}

__global__ __launch_bounds__(RT) void k_cg_residual(const float2* __restrict__ Ax, const float2* __restrict__ y,
                                                    float2* __restrict__ r, const int64_t* __restrict__ kprob) {
  const int p = blockIdx.y;
  const int64_t k0 = kprob[p], k1 = kprob[p + 1];
  for (int64_t i = k0 + blockIdx.x * (int64_t)RT + threadIdx.x; i < k1; i += (int64_t)gridDim.x * RT) {
    const float2 a = Ax[i], b = y[i];
    r[i] = make_float2(a.x - b.x, a.y - b.y);
  }
}

// Pass 1 of the line search: a = |Ax - y|^2, b = Re<Ax - y, Ad>, c = |Ad|^2 per problem.  RES: Ax
// already holds the residual Ax - y (as k_cg_residual and k_accept_k leave it, with the same FP32
// subtraction), so y is not read and the sums are the same bits.
template <typename R, bool RES>
__global__ __launch_bounds__(RT) void k_ls_kspace(const float2* __restrict__ Ax, const float2* __restrict__ Ad,
                                                  const float2* __restrict__ y, const int64_t* __restrict__ kprob,
                                                  const ProbState* __restrict__ st, double* part) {
  const int p = blockIdx.y;
  R v[3] = {R(0), R(0), R(0)};
  if (st[p].status == MRH_CG_ACTIVE) {
    const int64_t k0 = kprob[p], k1 = kprob[p + 1];
    for (int64_t i = k0 + blockIdx.x * (int64_t)RT + threadIdx.x; i < k1; i += (int64_t)gridDim.x * RT) {
      const float2 ax = Ax[i], ad = Ad[i];
      float rx = ax.x, ry = ax.y;
      if (!RES) {
        const float2 yy = y[i];
        rx = ax.x - yy.x;
        ry = ax.y - yy.y;
      }
      v[0] += (R)rx * rx + (R)ry * ry;
      v[1] += (R)rx * ad.x + (R)ry * ad.y;
      v[2] += (R)ad.x * ad.x + (R)ad.y * ad.y;
    }
  }
  block_reduce_store<3, R>(v, part + ((int64_t)p * gridDim.x + blockIdx.x) * PK_STRIDE);
}

// Pass 2 of the line search: Re<g, d>, |d|^2 (first pass) and TV(x + t_j d) for the
// candidates j in [J0, J1).  The first pass covers t = 0 and the NT_A - 1 largest
// steps, which is where Armijo almost always stops; the second pass covers the rest
// and does work only for problems whose first pass found no step.  Both passes cut the
// frames the same way, so each candidate's sum is accumulated exactly as in a single
// pass over all candidates.
// 16 x 16-pixel tiles walk their frames in order: x and d of frames f and f + 1 over the
// tile and one row / column after it (the forward differences) are staged in LDS, the
// next frame's stage (and, first pass, g of the next frame) in flight while the current
// one is evaluated.  Grid (ntiles * ngroups, P): block b runs tile b % ntiles over
// frames [fg (b / ntiles), + fg).
#define TL_H (TG_T + 1)          // staged side: the tile and one after
template <int J0, int J1, typename R>
__global__ __launch_bounds__(RT) void k_ls_image(const float2* __restrict__ x, const float2* __restrict__ d,
                                                 const float2* __restrict__ g, TVDev tv, int64_t npix_p,
                                                 double ls_beta, const ProbState* __restrict__ st, int tiles_x,
                                                 int ntiles, int fg, double* part) {
  constexpr bool FIRST = J0 == 0;
  constexpr int NV = (FIRST ? 2 : 0) + (J1 - J0), V0 = FIRST ? 2 : 0;
  __shared__ float2 sv[2][2][TL_H * TL_H];     // [x, d][frame parity][staged square]
  const int p = blockIdx.y, tid = threadIdx.x;
  if (!FIRST && !(st[p].status == MRH_CG_ACTIVE && st[p].ls_pending)) return;   // partials unread: skip
  R v[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = R(0);
  float tj[NT];
  tj[0] = 0.f;
  {
    double t = 1.0;
#pragma unroll
    for (int j = 1; j < NT; ++j) { tj[j] = (float)t; t *= ls_beta; }
  }
  if (st[p].status == MRH_CG_ACTIVE && (FIRST || st[p].ls_pending)) {
    const int tile = blockIdx.x % ntiles, grp = blockIdx.x / ntiles;
    const int X0 = (tile / tiles_x) * TG_T, Y0 = (tile % tiles_x) * TG_T;
    const int F = tv.F, S0 = tv.S0, S1 = tv.S1, plane = S0 * S1;
    const int fa = grp * fg, fb = min(F, fa + fg);
    const int lx = tid / TG_T, ly = tid % TG_T, ix = X0 + lx, iy = Y0 + ly;
    const bool mine = ix < S0 && iy < S1;
    const int c = lx * TL_H + ly;                // this pixel in the staged square
    const int64_t base = (int64_t)p * npix_p;
    // x and d of frame f at elements tid and tid + RT of the staged square (zero outside the volume)
    auto fetch = [&](int f, float2 (&r)[2][2]) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int i = tid + j * RT, row = i / TL_H, col = i - row * TL_H;
        const int gx = X0 + row, gy = Y0 + col;
        const bool in = i < TL_H * TL_H && f < F && gx < S0 && gy < S1;
        const int64_t n = base + (int64_t)f * plane + (int64_t)gx * S1 + gy;
        r[0][j] = in ? x[n] : make_float2(0.f, 0.f);
        r[1][j] = in ? d[n] : make_float2(0.f, 0.f);
      }
    };
    auto put = [&](int par, const float2 (&r)[2][2]) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        sv[q][par][tid] = r[q][0];
        if (tid + RT < TL_H * TL_H) sv[q][par][tid + RT] = r[q][1];
      }
    };
    float2 nx[2][2];
    fetch(fa, nx);
    put(0, nx);
    fetch(fa + 1, nx);
    float2 ng = make_float2(0.f, 0.f);
    const int64_t n0 = base + (int64_t)ix * S1 + iy;
    if (FIRST && mine && fa < fb) ng = g[n0 + (int64_t)fa * plane];
    int cur = 0;
    for (int f = fa; f < fb; ++f) {
      const float2 gn = ng;
      __syncthreads();                   // frame f - 1's stage has been read
      put(cur ^ 1, nx);                  // frame f + 1
      if (f + 2 <= fb) fetch(f + 2, nx);
      if (FIRST && mine && f + 1 < fb) ng = g[n0 + (int64_t)(f + 1) * plane];
      __syncthreads();
      if (mine) {
        const float2 *x0 = sv[0][cur], *x1 = sv[0][cur ^ 1], *d0 = sv[1][cur], *d1 = sv[1][cur ^ 1];
        if (FIRST) {
          const float2 dn = d0[c];
          v[0] += (R)gn.x * dn.x + (R)gn.y * dn.y;
          v[1] += (R)dn.x * dn.x + (R)dn.y * dn.y;
        }
        const float2 xq = x0[c], dq = d0[c];
        float ej[J1 - J0];               // sum over terms of lambda * energy density, per candidate
#pragma unroll
        for (int j = 0; j < J1 - J0; ++j) ej[j] = 0.f;
#pragma unroll
        for (int k = 0; k < MRH_TV_MAX_TERMS; ++k) {
          if (k >= tv.n) break;
          const TVTermDev& tm = tv.t[k];
          if (tm.lam == 0.f || tm.axes == 0) continue;
          // weighted forward differences of x and d (fdiff's arithmetic)
          float2 dx[3], dd[3];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            dx[a] = make_float2(0.f, 0.f); dd[a] = make_float2(0.f, 0.f);
            if (!(tm.axes & (1 << a))) continue;
            int o;
            const float2 *xs, *ds;
            if (a == 0) { if (f + 1 >= F) continue; o = c; xs = x1; ds = d1; }
            else if (a == 1) { if (ix + 1 >= S0) continue; o = c + TL_H; xs = x0; ds = d0; }
            else { if (iy + 1 >= S1) continue; o = c + 1; xs = x0; ds = d0; }
            const float w = tm.w[a];
            const float2 xu = xs[o], du = ds[o];
            dx[a] = make_float2(w * (xu.x - xq.x), w * (xu.y - xq.y));
            dd[a] = make_float2(w * (du.x - dq.x), w * (du.y - dq.y));
          }
          const float e2 = tm.eps * tm.eps;
#pragma unroll
          for (int j = J0; j < J1; ++j) {
            const float t = tj[j];
            float e = 0.f;
            if (tm.iso) {
              float ss = 0.f;
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                const float zx = fmaf(t, dd[a].x, dx[a].x), zy = fmaf(t, dd[a].y, dx[a].y);
                ss = fmaf(zx, zx, fmaf(zy, zy, ss));
              }
              e = hw_sqrt(ss + e2) - tm.eps;
            } else {
#pragma unroll
              for (int a = 0; a < 3; ++a) {
                if (tm.axes & (1 << a)) {
                  const float zx = fmaf(t, dd[a].x, dx[a].x), zy = fmaf(t, dd[a].y, dx[a].y);
                  e += hw_sqrt(fmaf(zx, zx, zy * zy) + e2) - tm.eps;
                }
              }
            }
            ej[j - J0] = fmaf(tm.lam, e, ej[j - J0]);
          }
        }
#pragma unroll
        for (int j = 0; j < J1 - J0; ++j) v[V0 + j] += (R)ej[j];
      }
      cur ^= 1;
    }
  }
  block_reduce_store<NV, R>(v, part + ((int64_t)p * gridDim.x + blockIdx.x) * PI_STRIDE);
}

// f(t) of the Armijo candidates: 0.5 (a + 2 b t + c t^2) + TV(t), evaluated in R.
template <typename R>
__device__ __forceinline__ double ls_value(const ProbState& s, double t, double tvj) {
  const R tt = (R)t;
  return (double)((R)0.5 * ((R)s.a + (R)2 * tt * (R)s.b + tt * tt * (R)s.c) + (R)tvj);
}

__device__ __forceinline__ void ls_record(ProbState& s, mrh_cg_hist* hist, int max_hist, int P, int p) {
  if (s.iter < max_hist) {
    mrh_cg_hist& h = hist[(int64_t)s.iter * P + p];
    h.f0 = s.f0; h.g0d = s.g0d; h.t = s.t; h.ft = s.ft;
  }
}

// Finish the first pass's reductions and try its steps (one block per problem):
// Armijo f(t) <= f(0) + c1 t <g, d> for t = 1, beta, ... in that order.
template <typename R>
__global__ __launch_bounds__(RT) void k_ls_select(ProbState* st, const double* part_i, const double* part_k,
                                                  int nb_i, int nb_k, double c1, double ls_beta, int ls_max,
                                                  mrh_cg_hist* hist, int max_hist, int P, int* active,
                                                  const double* part_q, int nb_q, double gl_lam) {
  const int p = blockIdx.x;
  double vals[3 + 2 + NT_A];
  block_sum_partials_n<3, R>(part_k + (int64_t)p * nb_k * PK_STRIDE, nb_k, PK_STRIDE, 0, vals);
  block_sum_partials_n<2 + NT_A, R>(part_i + (int64_t)p * nb_i * PI_STRIDE, nb_i, PI_STRIDE, 0, vals + 3);
  gl_collect(st[p], part_q, nb_q, p);
  if (threadIdx.x != 0) return;
  ProbState& s = st[p];
  s.ls_pending = 0;
  if (s.status != MRH_CG_ACTIVE) { s.t = 0.0; return; }
  s.a = vals[0]; s.b = vals[1]; s.c = vals[2]; s.g0d = vals[3]; s.dd = vals[4];
  for (int j = 0; j < NT_A; ++j) s.tv[j] = vals[5 + j];
  const double f0 = (double)((R)0.5 * (R)s.a + (R)s.tv[0]) + gl_value(s, gl_lam, 0.0);
  s.f0 = f0;
  s.t = 0.0;
  s.ft = f0;
  if (!(s.g0d < 0.0) || !isfinite(s.g0d)) {
    s.status = MRH_CG_NOT_DESCENT;
    active[p] = 0;
    return;
  }
  double t = 1.0;
  for (int j = 1; j <= ls_max && j < NT_A; ++j) {
    const double ft = ls_value<R>(s, t, s.tv[j]) + gl_value(s, gl_lam, t);
    if ((R)ft <= (R)f0 + (R)c1 * (R)t * (R)s.g0d) {
      s.t = t; s.ft = ft;
      ls_record(s, hist, max_hist, P, p);
      return;
    }
    t *= ls_beta;
  }
  if (ls_max >= NT_A) { s.ls_pending = 1; return; }
  s.status = MRH_CG_LS_FAILED; s.t = 0.0; active[p] = 0;
}

// Second pass: the remaining steps t = beta^(NT_A - 1), ..., for problems still pending.
template <typename R>
__global__ __launch_bounds__(RT) void k_ls_select_rest(ProbState* st, const double* part_i, int nb_i, double c1,
                                                       double ls_beta, int ls_max, mrh_cg_hist* hist,
                                                       int max_hist, int P, int* active, double gl_lam) {
  const int p = blockIdx.x;
  ProbState& s = st[p];
  if (!(s.status == MRH_CG_ACTIVE && s.ls_pending)) return;   // uniform over the block
  double vals[NT - NT_A];
  block_sum_partials_n<NT - NT_A, R>(part_i + (int64_t)p * nb_i * PI_STRIDE, nb_i, PI_STRIDE, 0, vals);
  if (threadIdx.x != 0) return;
  s.ls_pending = 0;
  for (int j = NT_A; j < NT; ++j) s.tv[j] = vals[j - NT_A];
  double t = 1.0;
  for (int j = 1; j < NT_A; ++j) t *= ls_beta;
  for (int j = NT_A; j <= ls_max && j < NT; ++j) {
    const double ft = ls_value<R>(s, t, s.tv[j]) + gl_value(s, gl_lam, t);
    if ((R)ft <= (R)s.f0 + (R)c1 * (R)t * (R)s.g0d) {
      s.t = t; s.ft = ft;
      ls_record(s, hist, max_hist, P, p);
      return;
    }
    t *= ls_beta;
  }
  s.status = MRH_CG_LS_FAILED; s.t = 0.0; active[p] = 0;
}

// ---------------------------------------------------------------- strong-Wolfe search
// graspcg/numerics/line_search.py:89-204 needs f(t) and f'(t) = Re<g(x + t d), d> at every
// trial.  Along d the data term is 0.5 (a + 2 b t + c t^2) with slope b + c t; the
// regulariser's value and slope come from image passes that each evaluate NW candidate
// steps.  A step kernel then replays the sequential search over the evaluated candidates
// (looked up by their exact t) until it needs a step that was not evaluated; that step
// opens the next pass's candidates: the expansion's next NW doublings, or the three
// bisection levels (7 midpoints) of the zoom bracket, which contain the next three
// midpoints whichever end each bisection moves.

// Image pass: TV(x + t d) and Re<g_TV(x + t d), d> for the candidates.  g_TV is the
// regulariser gradient the reference evaluates (-lambda times the divergence of the flux,
// without the axis weight, graspcg/regularization/tv_nd.py:54-88), so the slope is the one
// the reference computes from its gradient at the trial.  The first pass also sums
// Re<g, d> and |d|^2 and takes the candidates t = 0, 1, 2, 4, ..., 32; later passes take the
// state's candidates and skip problems with nothing pending.
template <bool FIRST>
__global__ __launch_bounds__(RT) void k_wolfe_image(const float2* __restrict__ x, const float2* __restrict__ d,
                                                    const float2* __restrict__ g, TVDev tv, int64_t npix_p,
                                                    const ProbState* __restrict__ st, double* part) {
  constexpr int V0 = FIRST ? 2 : 0, NV = V0 + 2 * NW;
  const int p = blockIdx.y;
  const bool live = st[p].status == MRH_CG_ACTIVE && (FIRST || st[p].ls_pending);
  if (!FIRST && !live) return;   // partials unread: skip
  double v[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) v[j] = 0.0;
  float tj[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) tj[j] = FIRST ? (j == 0 ? 0.f : (float)(1 << (j - 1))) : (float)st[p].wt[j];
  if (live) {
    const int plane = tv.S0 * tv.S1;
    for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < npix_p; i += (int64_t)gridDim.x * RT) {
      const int f = (int)(i / plane), rem = (int)(i - (int64_t)f * plane);
      const int ix = rem / tv.S1, iy = rem - ix * tv.S1;
      const int64_t n = (int64_t)p * npix_p + i;
      if (FIRST) {
        const float2 dn = d[n], gn = g[n];
        v[0] += (double)gn.x * dn.x + (double)gn.y * dn.y;
        v[1] += (double)dn.x * dn.x + (double)dn.y * dn.y;
      }
      for (int k = 0; k < tv.n; ++k) {
        const TVTermDev& tm = tv.t[k];
        if (tm.lam == 0.f || tm.axes == 0) continue;
        float2 dx[3], du[3], dw[3];   // weighted differences of x and d, unweighted of d
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          dx[a] = make_float2(0.f, 0.f); du[a] = make_float2(0.f, 0.f);
          if (tm.axes & (1 << a)) {
            dx[a] = fdiff(x, n, f, ix, iy, a, tv, tm.w[a]);
            du[a] = fdiff(d, n, f, ix, iy, a, tv, 1.f);
          }
          dw[a] = make_float2(tm.w[a] * du[a].x, tm.w[a] * du[a].y);
        }
        const float e2 = tm.eps * tm.eps;
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          const float t = tj[j];
          float e = 0.f, sl = 0.f;
          if (tm.iso) {
            float ss = 0.f, zd = 0.f;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              const float zx = fmaf(t, dw[a].x, dx[a].x), zy = fmaf(t, dw[a].y, dx[a].y);
              ss = fmaf(zx, zx, fmaf(zy, zy, ss));
              zd = fmaf(zx, du[a].x, fmaf(zy, du[a].y, zd));
            }
            const float q = ss + e2;
            e = hw_sqrt(q) - tm.eps;
            sl = zd * hw_rsq_guarded(q);
          } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
              if (tm.axes & (1 << a)) {
                const float zx = fmaf(t, dw[a].x, dx[a].x), zy = fmaf(t, dw[a].y, dx[a].y);
                const float q = fmaf(zx, zx, zy * zy) + e2;
                e += hw_sqrt(q) - tm.eps;
                sl += fmaf(zx, du[a].x, zy * du[a].y) * hw_rsq_guarded(q);
              }
            }
          }
          v[V0 + j] += (double)(tm.lam * e);
          v[V0 + NW + j] += (double)(tm.lam * sl);
        }
      }
    }
  }
  block_reduce_store<NV>(v, part + ((int64_t)p * gridDim.x + blockIdx.x) * PI_STRIDE);
}

// Index of step t among the first wn candidates (-1: not evaluated).  Unrolled compares and
// selects keep the candidate arrays in registers.
__device__ __forceinline__ int wolfe_find(const double (&wt)[NW], int wn, double t) {
  int r = -1;
#pragma unroll
  for (int j = NW - 1; j >= 0; --j)
    if (j < wn && wt[j] == t) r = j;
  return r;
}

__device__ __forceinline__ double wolfe_pick(const double (&a)[NW], int j) {
  double r = 0.0;
#pragma unroll
  for (int k = 0; k < NW; ++k)
    if (k == j) r = a[k];
  return r;
}

// Step kernel (one block per problem): finish the pass's sums and continue the
// sequential search of line_search.py:89-204 over the evaluated candidates.
//   expansion (:121-152): trial i at t (t = 1, 2, 4, ...): Armijo break, or f >= f(previous
//     trial) for i > 0, zooms on (previous, t); |f'| <= c2 |f'(0)| accepts; f' >= 0 zooms on
//     (t, previous); else t = min(2 t, 1e6).  ls_max trials without a result fail the search.
//   zoom (:161-204): mid = (lo + hi) / 2; Armijo break or f(mid) >= f(lo) sets hi = mid;
//     otherwise the curvature test accepts, else hi = lo when f'(mid) (hi - lo) >= 0, and
//     lo = mid.  After ls_max bisections lo is accepted (t = 0 included), as the reference does.
// `last`: the final pass of the graph; a search still pending there fails (the pass count
// covers the longest search, so this does not occur).
template <bool FIRST>
__global__ __launch_bounds__(RT) void k_wolfe_step(ProbState* st, const double* part_i, const double* part_k,
                                                   int nb_i, int nb_k, double c1, double c2, int ls_max,
                                                   int last, mrh_cg_hist* hist, int max_hist, int P,
                                                   int* active, const double* part_q, int nb_q, double gl_lam) {
  const int p = blockIdx.x;
  ProbState& s = st[p];
  if (!FIRST && !(s.status == MRH_CG_ACTIVE && s.ls_pending)) return;   // uniform over the block
  double vals[5 + 2 * NW];
  if (FIRST) {
    block_sum_partials_n<3>(part_k + (int64_t)p * nb_k * PK_STRIDE, nb_k, PK_STRIDE, 0, vals);
    block_sum_partials_n<2 + 2 * NW>(part_i + (int64_t)p * nb_i * PI_STRIDE, nb_i, PI_STRIDE, 0, vals + 3);
    gl_collect(s, part_q, nb_q, p);
  } else {
    block_sum_partials_n<2 * NW>(part_i + (int64_t)p * nb_i * PI_STRIDE, nb_i, PI_STRIDE, 0, vals + 5);
  }
  if (threadIdx.x != 0) return;
  double tvv[NW], tvs[NW];             // TV(x + t_j d) and its slope
#pragma unroll
  for (int j = 0; j < NW; ++j) { tvv[j] = vals[5 + j]; tvs[j] = vals[5 + NW + j]; }
  double wt[NW];
  int wn;
  s.ls_pending = 0;
  if (FIRST) {
    if (s.status != MRH_CG_ACTIVE) { s.t = 0.0; return; }
    s.a = vals[0]; s.b = vals[1]; s.c = vals[2]; s.g0d = vals[3]; s.dd = vals[4];
    s.f0 = 0.5 * s.a + tvv[0] + gl_value(s, gl_lam, 0.0);
    s.t = 0.0;
    s.ft = s.f0;
    if (!(s.g0d < 0.0) || !isfinite(s.g0d)) {
      s.status = MRH_CG_NOT_DESCENT;
      active[p] = 0;
      return;
    }
#pragma unroll
    for (int j = 0; j < NW; ++j) wt[j] = j == 0 ? 0.0 : (double)(1 << (j - 1));
    wn = NW;
    s.wzoom = 0; s.wi = 0;
    s.w_t = 1.0; s.w_tp = 0.0; s.w_fp = s.f0; s.w_gp = s.g0d;
  } else {
#pragma unroll
    for (int j = 0; j < NW; ++j) wt[j] = s.wt[j];
    wn = s.wn;
  }
  const double f0 = s.f0, g0d = s.g0d, gcurv = c2 * fabs(g0d);
  auto accept = [&](double t, double ft) {
    s.t = t; s.ft = ft;
    ls_record(s, hist, max_hist, P, p);
  };
  auto zoom = [&](double lo, double flo, double glo, double hi) {
    s.wzoom = 1; s.wi = 0;
    s.w_lo = lo; s.w_flo = flo; s.w_glo = glo; s.w_hi = hi;
  };
  for (;;) {
    if (!s.wzoom) {
      if (s.wi >= ls_max) { s.status = MRH_CG_LS_FAILED; s.t = 0.0; active[p] = 0; return; }
      const double t = s.w_t;
      const int j = wolfe_find(wt, wn, t);
      if (j < 0) {   // next pass: this step and its doublings, as many as trials remain
        const int n = min(NW, ls_max - s.wi);
        double u = t;
        for (int k = 0; k < NW; ++k) {
          s.wt[k] = k < n ? u : s.wt[0];
          u = fmin(2.0 * u, 1e6);
        }
        s.wn = n;
        s.ls_pending = 1;
        break;
      }
      const double ft = 0.5 * (s.a + 2.0 * t * s.b + t * t * s.c) + wolfe_pick(tvv, j) + gl_value(s, gl_lam, t);
      const double gt = s.b + t * s.c + wolfe_pick(tvs, j) + gl_slope(s, gl_lam, t);
      if (ft > f0 + c1 * t * g0d || (s.wi > 0 && ft >= s.w_fp)) { zoom(s.w_tp, s.w_fp, s.w_gp, t); continue; }
      if (fabs(gt) <= gcurv) { accept(t, ft); return; }
      if (gt >= 0.0) { zoom(t, ft, gt, s.w_tp); continue; }
      s.w_tp = t; s.w_fp = ft; s.w_gp = gt;
      s.w_t = fmin(2.0 * t, 1e6);
      s.wi += 1;
    } else {
      if (s.wi >= ls_max) { accept(s.w_lo, s.w_flo); return; }
      const double lo = s.w_lo, hi = s.w_hi, m = 0.5 * (lo + hi);
      const int j = wolfe_find(wt, wn, m);
      if (j < 0) {   // next pass: the bracket's midpoints to depth three
        const double mL = 0.5 * (lo + m), mR = 0.5 * (m + hi);
        s.wt[0] = m; s.wt[1] = mL; s.wt[2] = mR;
        s.wt[3] = 0.5 * (lo + mL); s.wt[4] = 0.5 * (mL + m);
        s.wt[5] = 0.5 * (m + mR); s.wt[6] = 0.5 * (mR + hi);
        s.wn = NW;
        s.ls_pending = 1;
        break;
      }
      const double fm = 0.5 * (s.a + 2.0 * m * s.b + m * m * s.c) + wolfe_pick(tvv, j) + gl_value(s, gl_lam, m);
      const double gm = s.b + m * s.c + wolfe_pick(tvs, j) + gl_slope(s, gl_lam, m);
      if (fm > f0 + c1 * m * g0d || fm >= s.w_flo) {
        s.w_hi = m;
      } else {
        if (fabs(gm) <= gcurv) { accept(m, fm); return; }
        if (gm * (hi - lo) >= 0.0) s.w_hi = lo;
        s.w_lo = m; s.w_flo = fm; s.w_glo = gm;
      }
      s.wi += 1;
    }
  }
  if (last) { s.ls_pending = 0; s.status = MRH_CG_LS_FAILED; s.t = 0.0; active[p] = 0; }
}

// Image passes the strong-Wolfe search may need after the first: ls_max expansion trials
// (NW - 1 in the first pass, NW per later pass) and ls_max bisections (three per pass), the
// zoom starting on the pass after the one that opened it.
static int wolfe_rest_passes(int ls_max) {
  const int exp_passes = 1 + (std::max(0, ls_max - (NW - 1)) + NW - 1) / NW;
  const int zoom_passes = (ls_max + 2) / 3;
  return exp_passes + zoom_passes - 1;
}

// x += t d (image) for every problem.
__global__ __launch_bounds__(RT) void k_accept_img(float2* __restrict__ x, const float2* __restrict__ d,
                                                   int64_t npix_p, const ProbState* __restrict__ st) {
  const int p = blockIdx.y;
  const float t = (float)st[p].t;
  if (t == 0.f) return;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < npix_p; i += (int64_t)gridDim.x * RT) {
    const int64_t n = (int64_t)p * npix_p + i;
    const float2 dv = d[n];
    float2 xv = x[n];
    xv.x = fmaf(t, dv.x, xv.x);
    xv.y = fmaf(t, dv.y, xv.y);
    x[n] = xv;
  }
}

// Ax += t Ad, r = Ax - y (k-space) for every problem.
__global__ __launch_bounds__(RT) void k_accept_k(float2* __restrict__ Ax, const float2* __restrict__ Ad,
                                                 const float2* __restrict__ y, float2* __restrict__ r,
                                                 const int64_t* __restrict__ kprob, const ProbState* __restrict__ st) {
  const int p = blockIdx.y;
  const float t = (float)st[p].t;
  const int64_t k0 = kprob[p], k1 = kprob[p + 1];
  for (int64_t i = k0 + blockIdx.x * (int64_t)RT + threadIdx.x; i < k1; i += (int64_t)gridDim.x * RT) {
    float2 a = Ax[i];
    if (t != 0.f) {
      const float2 ad = Ad[i];
      a.x = fmaf(t, ad.x, a.x);
      a.y = fmaf(t, ad.y, a.y);
      Ax[i] = a;
    }
    const float2 yy = y[i];
    r[i] = make_float2(a.x - yy.x, a.y - yy.y);
  }
}

// g += TV gradient at x; partials: TV energy, |g|^2, Re<g_prev, g>, Re<g, d>.  With a diagonal
// preconditioner D (null: identity) the first two are weighted by 1 / D (graspcg/ops/dot.py:5-71
// with diag, as the direction rules use them, graspcg/numerics/directions.py:75-162).
// 16 x 16-pixel tiles walk their frames in order: each tile stages x of the frames f and f + 1 over
// its pixels, one row / column before them (the backward neighbours) and one after (the forward
// differences) in LDS; every flux is evaluated once (the tile's pixels, plus the 32 neighbours above
// and to the left, which hold the spatial backward fluxes), and a term's temporal flux at f - 1 stays
// in registers.  The fluxes are those of term_flux on the same values (same arithmetic, flux_of),
// summed into the divergence in tv_grad_at's order, so g equals k_tv_energy_grad's to the bit; only
// the per-thread order of the block partials differs.  Frame f + 2's stage and frame f + 1's g, g_prev,
// d and D are loaded while frame f is computed, so a frame waits for loads issued one frame earlier.
// Grid (ntiles * ngroups, P): block b runs tile b % ntiles over frames [fg (b / ntiles), + fg) (the
// group re-reads x at its first frame - 1 for that frame's temporal flux); the groups are as many as
// fit on the device at once.
#define TG_H (TG_T + 2)          // staged side: one before, one after
#define TG_P (TG_T + 1)          // flux side: one before
template <typename R>
__global__ __launch_bounds__(RT) void k_cg_grad_tiled(const float2* __restrict__ x, float2* __restrict__ g,
                                                      const float2* __restrict__ d, const float2* __restrict__ gp,
                                                      const float* __restrict__ D, TVDev tv, int64_t npix_p,
                                                      const ProbState* __restrict__ st, int prime, int tiles_x,
                                                      int ntiles, int fg, double* part) {
  __shared__ float2 sx[2][TG_H * TG_H];
  __shared__ float2 sp[MRH_TV_MAX_TERMS][2][TG_P * TG_P];
  const int p = blockIdx.y, tid = threadIdx.x;
  R v[5] = {R(0), R(0), R(0), R(0), R(0)};
  const bool on = prime || st[p].status == MRH_CG_ACTIVE;
  const int tile = blockIdx.x % ntiles, grp = blockIdx.x / ntiles;
  const int X0 = (tile / tiles_x) * TG_T, Y0 = (tile % tiles_x) * TG_T;   // first row (axis 1), column (axis 2)
  const int F = tv.F, S0 = tv.S0, S1 = tv.S1;
  const int fa = grp * fg, fb = min(F, fa + fg);
  const int lx = tid / TG_T, ly = tid % TG_T;          // this thread's pixel in the tile
  const int ix = X0 + lx, iy = Y0 + ly;
  const bool mine = ix < S0 && iy < S1;
  // halo flux position of threads 0..31: above the tile (row X0 - 1) or left of it (column Y0 - 1)
  const bool hal = tid < 2 * TG_T;
  const int hx = tid < TG_T ? -1 : tid - TG_T, hy = tid < TG_T ? tid : -1;
  const int plane = S0 * S1;
  const float2* xp = x + (int64_t)p * npix_p;
  // x of frame f over rows X0-1 .. X0+16, columns Y0-1 .. Y0+16: element tid and tid + RT of the
  // staged square, fetched into registers and put into LDS later (zero outside the volume)
  auto fetch = [&](int f, float2 (&r)[2]) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int i = tid + j * RT, row = i / TG_H, c = i - row * TG_H;
      const int gx = X0 - 1 + row, gy = Y0 - 1 + c;
      r[j] = (i < TG_H * TG_H && f < F && (unsigned)gx < (unsigned)S0 && (unsigned)gy < (unsigned)S1)
                 ? xp[(int64_t)f * plane + (int64_t)gx * S1 + gy] : make_float2(0.f, 0.f);
    }
  };
  auto put = [&](float2* s, const float2 (&r)[2]) {
    s[tid] = r[0];
    if (tid + RT < TG_H * TG_H) s[tid + RT] = r[1];
  };
  // flux of term tm at tile position (rx, ry) (staged coordinates minus one) of frame f
  auto flux_at = [&](const float2* s0, const float2* s1, int rx, int ry, int f, const TVTermDev& tm,
                     float2 (&pf)[3]) {
    const int gx = X0 + rx, gy = Y0 + ry, i = (rx + 1) * TG_H + (ry + 1);
    const float2 q = s0[i];
    float2 dv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) dv[a] = make_float2(0.f, 0.f);
    if ((tm.axes & 1) && f + 1 < F) { const float2 u = s1[i]; dv[0] = make_float2(tm.w[0] * (u.x - q.x), tm.w[0] * (u.y - q.y)); }
    if ((tm.axes & 2) && gx + 1 < S0) { const float2 u = s0[i + TG_H]; dv[1] = make_float2(tm.w[1] * (u.x - q.x), tm.w[1] * (u.y - q.y)); }
    if ((tm.axes & 4) && gy + 1 < S1) { const float2 u = s0[i + 1]; dv[2] = make_float2(tm.w[2] * (u.x - q.x), tm.w[2] * (u.y - q.y)); }
    return flux_of(dv, tm, pf);
  };
  float2 pt[MRH_TV_MAX_TERMS];          // each term's temporal flux at (f - 1, this pixel)
#pragma unroll
  for (int k = 0; k < MRH_TV_MAX_TERMS; ++k) pt[k] = make_float2(0.f, 0.f);
  if (on && tv.n > 0) {
    int cur = 0;
    float2 nx[2];                        // the next frame to stage, in flight during the current one
    if (fa > 0) {                        // temporal fluxes of frame fa - 1
      fetch(fa - 1, nx);
      put(sx[0], nx);
      fetch(fa, nx);
      put(sx[1], nx);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < MRH_TV_MAX_TERMS; ++k) {
        if (k >= tv.n) break;
        const TVTermDev& tm = tv.t[k];
        if (tm.lam == 0.f || tm.axes == 0 || !(tm.axes & 1)) continue;
        float2 pf[3];
        (void)flux_at(sx[0], sx[1], lx, ly, fa - 1, tm, pf);
        pt[k] = pf[0];
      }
      cur = 1;
    } else {
      fetch(fa, nx);
      put(sx[0], nx);
    }
    fetch(fa + 1, nx);
    // g, g_prev, d and D of the next frame at this pixel, in flight during the current frame (each
    // pixel's g is read and written by its own thread only)
    const int64_t n0 = (int64_t)p * npix_p + (int64_t)ix * S1 + iy;
    float2 ng = make_float2(0.f, 0.f), np = ng, nd = ng;
    float nD = 1.f;
    auto fetch_g = [&](int f) {
      const int64_t n = n0 + (int64_t)f * plane;
      if (mine) {
        ng = g[n];
        if (gp) np = gp[n];
        if (d) nd = d[n];
        if (D) nD = D[n];
      }
    };
    fetch_g(fa);
    for (int f = fa; f < fb; ++f) {
      const int64_t n = n0 + (int64_t)f * plane;
      float2 gv = ng;
      const float2 qp = np, qd = nd;
      const float dn = nD;
      __syncthreads();                   // frame f - 1's stage and fluxes have been read
      put(sx[cur ^ 1], nx);              // frame f + 1
      if (f + 2 <= fb) fetch(f + 2, nx);
      if (f + 1 < fb) fetch_g(f + 1);
      __syncthreads();
      const float2* s0 = sx[cur];
      const float2* s1 = sx[cur ^ 1];
      float2 pq[MRH_TV_MAX_TERMS][3];
      float eq[MRH_TV_MAX_TERMS];
#pragma unroll
      for (int k = 0; k < MRH_TV_MAX_TERMS; ++k) {
        eq[k] = 0.f;
        if (k >= tv.n) break;
        const TVTermDev& tm = tv.t[k];
        if (tm.lam == 0.f || tm.axes == 0) continue;
        eq[k] = flux_at(s0, s1, lx, ly, f, tm, pq[k]);
        const int o = (lx + 1) * TG_P + (ly + 1);
        sp[k][0][o] = pq[k][1];
        sp[k][1][o] = pq[k][2];
        if (hal) {
          float2 ph[3];
          (void)flux_at(s0, s1, hx, hy, f, tm, ph);
          const int oh = (hx + 1) * TG_P + (hy + 1);
          sp[k][0][oh] = ph[1];
          sp[k][1][oh] = ph[2];
        }
      }
      __syncthreads();
      if (mine) {
        // tv_grad_at's sums, in its order
        float2 gsum = make_float2(0.f, 0.f);
        float energy = 0.f;
#pragma unroll
        for (int k = 0; k < MRH_TV_MAX_TERMS; ++k) {
          if (k >= tv.n) break;
          const TVTermDev& tm = tv.t[k];
          if (tm.lam == 0.f || tm.axes == 0) continue;
          float2 div = make_float2(0.f, 0.f);
          if (tm.axes & 1) {
            div = cadd(div, pq[k][0]);
            if (f > 0) div = csub(div, pt[k]);
            pt[k] = pq[k][0];
          }
          if (tm.axes & 2) {
            div = cadd(div, pq[k][1]);
            if (ix > 0) div = csub(div, sp[k][0][lx * TG_P + (ly + 1)]);
          }
          if (tm.axes & 4) {
            div = cadd(div, pq[k][2]);
            if (iy > 0) div = csub(div, sp[k][1][(lx + 1) * TG_P + ly]);
          }
          gsum.x = fmaf(-tm.lam, div.x, gsum.x);
          gsum.y = fmaf(-tm.lam, div.y, gsum.y);
          energy = fmaf(tm.lam, eq[k], energy);
        }
        gv = cadd(gv, gsum);
        g[n] = gv;
        const float2 xv = s0[(lx + 1) * TG_H + (ly + 1)];
        v[4] += (R)xv.x * xv.x + (R)xv.y * xv.y;
        v[0] += (R)energy;
        const R wd = D ? (R)1 / (R)dn : (R)1;
        v[1] += ((R)gv.x * gv.x + (R)gv.y * gv.y) * wd;
        if (gp) v[2] += ((R)qp.x * gv.x + (R)qp.y * gv.y) * wd;
        if (d) v[3] += (R)qd.x * gv.x + (R)qd.y * gv.y;
      }
      cur ^= 1;
    }
  } else if (on) {                       // no TV term: the norms and dots of g alone
    for (int f = fa; f < fb; ++f) {
      if (!mine) continue;
      const int64_t n = (int64_t)p * npix_p + (int64_t)f * plane + (int64_t)ix * S1 + iy;
      const float2 gv = g[n], xv = x[n];
      v[4] += (R)xv.x * xv.x + (R)xv.y * xv.y;
      const R wd = D ? (R)1 / (R)D[n] : (R)1;
      v[1] += ((R)gv.x * gv.x + (R)gv.y * gv.y) * wd;
      if (gp) { const float2 q = gp[n]; v[2] += ((R)q.x * gv.x + (R)q.y * gv.y) * wd; }
      if (d) { const float2 q = d[n]; v[3] += (R)q.x * gv.x + (R)q.y * gv.y; }
    }
  }
  block_reduce_store<5, R>(v, part + ((int64_t)p * gridDim.x + blockIdx.x) * PI_STRIDE);
}

// Finish gradient reductions: beta, stop rule, history, iteration count.
template <typename R>
__global__ __launch_bounds__(RT) void k_cg_beta(ProbState* st, const double* part_i, int nb_i, int prime,
                                                int direction, double rel_tol, double restart_c,
                                                const double* __restrict__ target, mrh_cg_hist* hist,
                                                int max_hist, int P, int* active, const double* part_q,
                                                int nb_q) {
  const int p = blockIdx.x;
  double vals[5], eq[1] = {0.0};
  block_sum_partials_n<5, R>(part_i + (int64_t)p * nb_i * PI_STRIDE, nb_i, PI_STRIDE, 0, vals);
  if (part_q) block_sum_partials_n<1>(part_q + (int64_t)p * nb_q * PQ_STRIDE, nb_q, PQ_STRIDE, 0, eq);
  if (threadIdx.x != 0) return;
  ProbState& s = st[p];
  if (!prime && s.status != MRH_CG_ACTIVE) return;
  s.freg = vals[0] + eq[0];
  s.rho = vals[1];
  s.gpg = vals[2];
  s.gdot_step = vals[3];
  if (prime) {
    s.beta = 0.0;
    if (direction == MRH_CG_DIR_PRPLUS) { s.rho_prev = s.rho; s.have_rho_prev = 1; }
    return;
  }
  R beta;
  const R rho = (R)s.rho, tiny = (R)1e-20;
  if (direction == MRH_CG_DIR_PRPLUS) {
    beta = (rho - (R)s.gpg) / fmax((R)s.rho_prev, tiny);
    beta = fmax(beta, (R)0);
  } else if (direction == MRH_CG_DIR_DY) {
    // directions.py:111-139: Re<d, g> is this pass's gdot_step, Re<d, g_prev> the line search's g0d
    beta = rho / fmax((R)s.gdot_step - (R)s.g0d, tiny);
  } else {
    if (!s.have_rho_prev) { s.rho_prev = s.rho; s.have_rho_prev = 1; }  // first beta = 1 (directions.py:145-149)
    beta = rho / fmax((R)s.rho_prev, tiny);
  }
  // descent safeguard: Re<g, -g + beta d> = -rho + beta Re<g, d>
  if (restart_c >= 0.0 && -rho + beta * (R)s.gdot_step > -(R)restart_c * rho) beta = (R)0;
  s.rho_prev = s.rho;
  s.beta = beta;
  if (s.iter < max_hist) {
    mrh_cg_hist& h = hist[(int64_t)s.iter * P + p];
    h.rho = s.rho; h.beta = beta; h.freg = s.freg; h.gdot_step = s.gdot_step;
    h.xnorm2 = vals[4]; h.dnorm2 = s.dd;
  }
  s.iter += 1;
  if (rel_tol > 0.0 && s.have_f_prev && fabs((R)s.ft - (R)s.f_prev) <= (R)rel_tol * fabs((R)s.f_prev))
    s.status = MRH_CG_REL_TOL;
  if (s.ft <= target[p]) s.status = MRH_CG_TARGET;
  if (s.status != MRH_CG_ACTIVE) active[p] = 0;
  s.f_prev = s.ft;
  s.have_f_prev = 1;
}

// d = beta d - g / D (prime: d = -g / D; D null: identity); PR+ keeps g in g_prev.
__global__ __launch_bounds__(RT) void k_cg_direction(float2* __restrict__ d, const float2* __restrict__ g,
                                                     float2* __restrict__ gp, const float* __restrict__ D,
                                                     int64_t npix_p, const ProbState* __restrict__ st,
                                                     int prime) {
  const int p = blockIdx.y;
  if (!prime && st[p].status != MRH_CG_ACTIVE) return;
  const float beta = prime ? 0.f : (float)st[p].beta;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < npix_p; i += (int64_t)gridDim.x * RT) {
    const int64_t n = (int64_t)p * npix_p + i;
    const float2 gv = g[n];
    float2 dv = prime ? make_float2(0.f, 0.f) : d[n];
    if (D) {
      const float dn = D[n];
      dv.x = beta * dv.x - gv.x / dn;
      dv.y = beta * dv.y - gv.y / dn;
    } else {
      dv.x = beta * dv.x - gv.x;
      dv.y = beta * dv.y - gv.y;
    }
    d[n] = dv;
    if (gp) gp[n] = gv;
  }
}

// Diagonal majoriser of the regularisers at x (graspcg/regularization/manager.py:145-218), the
// preconditioner D of the direction rules.  D starts at 1 (the workspace's ones,
// graspcg/solvers/cg.py:95); per TV term, in order:
//   a term over the frame axis adds its axis profile only (tv_nd.py:181-215): 2 lambda w_0^2,
//     twice that on interior frames (the 1-D Laplacian degree);
//   any other term adds the lagged-diffusivity weights (tv_nd.py:113-162): per axis a,
//     lambda w_a^2 / sqrt(|D_a x|^2 + eps^2) at the pixel (D_a the unweighted forward
//     difference, zero on the last slice), twice that away from the axis' two ends.
__device__ __forceinline__ float axis_degree(int pos, int n) { return (n > 2 && pos > 0 && pos < n - 1) ? 2.f : 1.f; }

__global__ __launch_bounds__(RT) void k_build_diag(const float2* __restrict__ x, float* __restrict__ D, TVDev tv,
                                                   GLDev gl, int64_t npix_p) {
  const int p = blockIdx.y;
  const int plane = tv.S0 * tv.S1;
  for (int64_t i = blockIdx.x * (int64_t)RT + threadIdx.x; i < npix_p; i += (int64_t)gridDim.x * RT) {
    const int f = (int)(i / plane), rem = (int)(i - (int64_t)f * plane);
    const int ix = rem / tv.S1, iy = rem - ix * tv.S1;
    const int64_t n = (int64_t)p * npix_p + i;
    float dv = 1.f;
    for (int k = 0; k < tv.n; ++k) {
      const TVTermDev& tm = tv.t[k];
      if (tm.lam == 0.f || tm.axes == 0) continue;
      if (tm.axes & 1) {
        dv += 2.f * tm.lam * tm.w[0] * tm.w[0] * axis_degree(f, tv.F);
        continue;
      }
      for (int a = 1; a < 3; ++a) {
        if (!(tm.axes & (1 << a))) continue;
        const float2 gi = fdiff(x, n, f, ix, iy, a, tv, 1.f);
        const float wi = 1.f / sqrtf(gi.x * gi.x + gi.y * gi.y + tm.eps * tm.eps);
        dv += tm.lam * (tm.w[a] * tm.w[a]) * wi * axis_degree(a == 1 ? ix : iy, a == 1 ? tv.S0 : tv.S1);
      }
    }
    if (gl.lam != 0.f) dv += gl.prof[f];   // graph term: its frame profile (graph_laplacian.py:283-321)
    D[n] = dv;
  }
}

// ---------------------------------------------------------------- engine C ABI
// Blocks per problem for the grid-stride image / k-space kernels: about two elements per
// thread, at most 2048 blocks (8 per CU on 256 CUs, so a problem's sweep keeps enough
// wavefronts in flight to hide its loads); each block leaves one FP64 partial per value.
static int blocks_for(int64_t n) {
  int64_t b = (n + RT * 2 - 1) / (RT * 2);
  return (int)std::max<int64_t>(1, std::min<int64_t>(2048, b));
}

// Frame groups of a tiled pass: as many as fit on the device at once (one round of workgroups; a
// partly filled second round costs a whole group's frames again).  fg = frames per group,
// nb = workgroups per problem.
template <typename K>
static void tile_frame_groups(K kernel, int ntiles, int P, int F, int* fg, int* nb) {
  int dev = 0, ncu = 0, per_cu = 0;
  int64_t resident = 2048;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, RT, 0) == hipSuccess && ncu > 0 && per_cu > 0)
    resident = (int64_t)ncu * per_cu;
  const int64_t want = std::max<int64_t>(1, resident / ((int64_t)ntiles * P));
  const int ng = (int)std::min<int64_t>(F, want);
  *fg = (F + ng - 1) / ng;
  *nb = ntiles * ((F + *fg - 1) / *fg);
}

extern "C" int mrh_cg_create(const mrh_cg_desc* d, mrh_cg** out) {
  MRH_REQUIRE(d && out && d->k_prob, MRH_EINVAL, "cg_create: null argument");
  *out = nullptr;
  MRH_REQUIRE(d->ls_max_iter >= 1 && d->ls_max_iter <= NT - 1, MRH_EUNSUP,
              "ls_max_iter must be in [1, %d]", NT - 1);
  MRH_REQUIRE(d->direction == MRH_CG_DIR_FR || d->direction == MRH_CG_DIR_PRPLUS || d->direction == MRH_CG_DIR_DY,
              MRH_EINVAL, "direction must be FR, PR+ or DY");
  MRH_REQUIRE(d->tv.P >= 1 && d->tv.P <= 65535, MRH_EINVAL, "P outside [1, 65535]");
  MRH_REQUIRE(d->ls_kind == MRH_CG_LS_ARMIJO || d->ls_kind == MRH_CG_LS_WOLFE, MRH_EINVAL,
              "ls_kind must be Armijo or strong Wolfe");
  MRH_REQUIRE(!(d->ls_kind == MRH_CG_LS_WOLFE) || (d->c2 >= 0.0 && std::isfinite(d->c2)), MRH_EINVAL,
              "c2 must be a finite value >= 0");
  mrh_cg* e = new mrh_cg();
  std::memset(e, 0, sizeof(*e));
  int rc = make_tvdev(&d->tv, &e->tv);
  if (rc) { delete e; return rc; }
  e->P = d->tv.P; e->F = d->tv.F; e->S0 = d->tv.S0; e->S1 = d->tv.S1;
  e->direction = d->direction; e->ls_max_iter = d->ls_max_iter;
  e->c1 = d->c1; e->ls_beta = d->ls_beta; e->rel_tol = d->rel_tol; e->restart_c = d->restart_c;
  e->ls_kind = d->ls_kind; e->c2 = d->c2;
  e->fp32 = d->fp32_reductions ? 1 : 0;
  if (e->fp32 && (d->ls_kind != MRH_CG_LS_ARMIJO || d->gl_weight != 0.0)) {
    delete e;
    mrh_set_error("FP32 reductions are built for the Armijo search without a graph-Laplacian term");
    return MRH_EUNSUP;
  }
  e->d_diag = d->diag;
  if (d->gl_weight != 0.0) {
    if (d->gl_F != e->F || d->gl_F < 1 || d->gl_F > GL_MAXF || !d->gl_L || !d->gl_prof) {
      delete e;
      mrh_set_error("graph term: F = %d must equal the frames per problem (%d), be <= %d, with L and profile",
                    d->gl_F, d->tv.F, GL_MAXF);
      return MRH_EINVAL;
    }
    e->gl.F = d->gl_F;
    e->gl.lam = (float)d->gl_weight;
  }
  e->max_hist = std::max(1, d->max_hist);
  e->npix_p = (int64_t)e->F * e->S0 * e->S1;
  int64_t kmax = 0;
  for (int p = 0; p < e->P; ++p) {
    int64_t n = d->k_prob[p + 1] - d->k_prob[p];
    if (n < 0) { delete e; mrh_set_error("k_prob must be non-decreasing"); return MRH_EINVAL; }
    kmax = std::max(kmax, n);
  }
  e->kmax_p = kmax;
  e->nb_i = blocks_for(e->npix_p);
  e->nb_k = blocks_for(kmax);
  // tiled passes (gradient, line-search image): 16 x 16 tiles, frames split into as many groups as
  // fit on the device at once
  e->tg_tiles_x = (e->S1 + TG_T - 1) / TG_T;
  e->tg_ntiles = ((e->S0 + TG_T - 1) / TG_T) * e->tg_tiles_x;
  if (e->fp32) {
    tile_frame_groups(k_cg_grad_tiled<float>, e->tg_ntiles, e->P, e->F, &e->tg_fg, &e->nb_g);
    tile_frame_groups(k_ls_image<0, NT_A, float>, e->tg_ntiles, e->P, e->F, &e->tl_fg, &e->nb_l);
  } else {
    tile_frame_groups(k_cg_grad_tiled<double>, e->tg_ntiles, e->P, e->F, &e->tg_fg, &e->nb_g);
    tile_frame_groups(k_ls_image<0, NT_A, double>, e->tg_ntiles, e->P, e->F, &e->tl_fg, &e->nb_l);
  }
  hipError_t err;
#define ALLOC(ptr, bytes)                                                         \
  err = hipMalloc((void**)&(ptr), (bytes));                                      \
  if (err != hipSuccess) { mrh_cg_destroy(e); mrh_set_error("cg_create: hipMalloc failed: %s", hipGetErrorString(err)); return MRH_ENOMEM; }
  ALLOC(e->d_kprob, sizeof(int64_t) * (e->P + 1));
  ALLOC(e->d_part_i, sizeof(double) * (size_t)e->P * std::max({e->nb_i, e->nb_g, e->nb_l}) * PI_STRIDE);
  ALLOC(e->d_part_k, sizeof(double) * (size_t)e->P * e->nb_k * PK_STRIDE);
  ALLOC(e->d_st, sizeof(ProbState) * e->P);
  ALLOC(e->d_hist, sizeof(mrh_cg_hist) * (size_t)e->P * e->max_hist);
  ALLOC(e->d_target, sizeof(double) * e->P);
  ALLOC(e->d_last, sizeof(mrh_cg_hist) * e->P);
  ALLOC(e->d_active, sizeof(int) * e->P);
  err = hipHostMalloc((void**)&e->h_st, sizeof(ProbState) * e->P, hipHostMallocDefault);
  if (err != hipSuccess) { e->h_st = nullptr; mrh_cg_destroy(e); mrh_set_error("cg_create: hipHostMalloc failed: %s", hipGetErrorString(err)); return MRH_ENOMEM; }
  err = hipHostMalloc((void**)&e->h_last, sizeof(mrh_cg_hist) * e->P, hipHostMallocDefault);
  if (err != hipSuccess) { e->h_last = nullptr; mrh_cg_destroy(e); mrh_set_error("cg_create: hipHostMalloc failed: %s", hipGetErrorString(err)); return MRH_ENOMEM; }
  if (e->gl.lam != 0.f) {
    const int plane = e->S0 * e->S1;
    e->nb_q = std::max(1, std::min(2048, (plane + GL_TILE - 1) / GL_TILE));
    e->gl_lds_grad = sizeof(float) * gl_lds_floats(e->gl.F, 1);
    e->gl_lds_ls = sizeof(float) * gl_lds_floats(e->gl.F, 2);
    ALLOC(e->d_gl, sizeof(float) * (size_t)e->gl.F * (e->gl.F + 1));
    ALLOC(e->d_part_q, sizeof(double) * (size_t)e->P * e->nb_q * PQ_STRIDE);
    err = hipMemcpy(e->d_gl, d->gl_L, sizeof(float) * e->gl.F * e->gl.F, hipMemcpyHostToDevice);
    if (err == hipSuccess)
      err = hipMemcpy(e->d_gl + e->gl.F * e->gl.F, d->gl_prof, sizeof(float) * e->gl.F, hipMemcpyHostToDevice);
    if (err != hipSuccess) { mrh_cg_destroy(e); mrh_set_error("cg_create: copy failed: %s", hipGetErrorString(err)); return MRH_EHIP; }
    e->gl.L = e->d_gl;
    e->gl.prof = e->d_gl + e->gl.F * e->gl.F;
  }
#undef ALLOC
  {
    std::vector<double> none(e->P, -HUGE_VAL);
    err = hipMemcpy(e->d_target, none.data(), sizeof(double) * e->P, hipMemcpyHostToDevice);
    if (err != hipSuccess) { mrh_cg_destroy(e); mrh_set_error("cg_create: copy failed: %s", hipGetErrorString(err)); return MRH_EHIP; }
  }
  err = hipMemcpy(e->d_kprob, d->k_prob, sizeof(int64_t) * (e->P + 1), hipMemcpyHostToDevice);
  if (err != hipSuccess) { mrh_cg_destroy(e); mrh_set_error("cg_create: copy failed: %s", hipGetErrorString(err)); return MRH_EHIP; }
  hipLaunchKernelGGL(k_cg_reset, dim3((e->P + 63) / 64), dim3(64), 0, 0, e->d_st, e->d_active, e->P);
  err = hipDeviceSynchronize();
  if (err != hipSuccess) { mrh_cg_destroy(e); mrh_set_error("cg_create: reset failed: %s", hipGetErrorString(err)); return MRH_EHIP; }
  *out = e;
  return MRH_OK;
}

extern "C" int mrh_cg_destroy(mrh_cg* e) {
  if (!e) return MRH_OK;
  void* bufs[] = {e->d_kprob, e->d_part_i, e->d_part_k, e->d_st, e->d_hist, e->d_target, e->d_active, e->d_gl,
                  e->d_part_q, e->d_last};
  for (void* b : bufs) if (b) (void)hipFree(b);
  if (e->h_st) (void)hipHostFree(e->h_st);
  if (e->h_last) (void)hipHostFree(e->h_last);
  delete e;
  return MRH_OK;
}

extern "C" int mrh_cg_reset(mrh_cg* e, void* stream) {
  MRH_REQUIRE(e, MRH_EINVAL, "cg_reset: null engine");
  hipLaunchKernelGGL(k_cg_reset, dim3((e->P + 63) / 64), dim3(64), 0, (hipStream_t)stream, e->d_st, e->d_active,
                     e->P);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_cg_residual(mrh_cg* e, const void* Ax, const void* y, void* r, void* stream) {
  MRH_REQUIRE(e && Ax && y && r, MRH_EINVAL, "cg_residual: null argument");
  hipLaunchKernelGGL(k_cg_residual, dim3(e->nb_k, e->P), dim3(RT), 0, (hipStream_t)stream, (const float2*)Ax,
                     (const float2*)y, (float2*)r, e->d_kprob);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_cg_linesearch(mrh_cg* e, const void* x, const void* d, const void* g, const void* Ax,
                                 const void* Ad, const void* y, void* stream) {
  MRH_REQUIRE(e && x && d && g && Ax && Ad, MRH_EINVAL, "cg_linesearch: null argument");
  hipStream_t st = (hipStream_t)stream;
  if (y) {
    LAUNCH_R2(e, k_ls_kspace, false, dim3(e->nb_k, e->P), dim3(RT), 0, st, (const float2*)Ax, (const float2*)Ad,
              (const float2*)y, e->d_kprob, e->d_st, e->d_part_k);
  } else {                                 // Ax holds the residual r = Ax - y
    LAUNCH_R2(e, k_ls_kspace, true, dim3(e->nb_k, e->P), dim3(RT), 0, st, (const float2*)Ax, (const float2*)Ad,
              (const float2*)nullptr, e->d_kprob, e->d_st, e->d_part_k);
  }
  MRH_CHECK_LAUNCH();
  const double gl_lam = e->gl.lam;
  const double* part_q = e->gl.lam != 0.f ? e->d_part_q : nullptr;
  if (part_q) {
    hipLaunchKernelGGL(k_gl_ls, dim3(e->nb_q, e->P), dim3(256), e->gl_lds_ls, st, (const float2*)x, (const float2*)d,
                       e->gl, e->S0 * e->S1, e->npix_p, e->d_st, e->d_part_q);
    MRH_CHECK_LAUNCH();
  }
  if (e->ls_kind == MRH_CG_LS_WOLFE) {
    const int rest = wolfe_rest_passes(e->ls_max_iter);
    hipLaunchKernelGGL(k_wolfe_image<true>, dim3(e->nb_i, e->P), dim3(RT), 0, st, (const float2*)x,
                       (const float2*)d, (const float2*)g, e->tv, e->npix_p, e->d_st, e->d_part_i);
    MRH_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_wolfe_step<true>, dim3(e->P), dim3(RT), 0, st, e->d_st, e->d_part_i, e->d_part_k, e->nb_i,
                       e->nb_k, e->c1, e->c2, e->ls_max_iter, rest == 0 ? 1 : 0, e->d_hist, e->max_hist, e->P,
                       e->d_active, part_q, e->nb_q, gl_lam);
    MRH_CHECK_LAUNCH();
    for (int q = 0; q < rest; ++q) {
      hipLaunchKernelGGL(k_wolfe_image<false>, dim3(e->nb_i, e->P), dim3(RT), 0, st, (const float2*)x,
                         (const float2*)d, (const float2*)g, e->tv, e->npix_p, e->d_st, e->d_part_i);
      MRH_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_wolfe_step<false>, dim3(e->P), dim3(RT), 0, st, e->d_st, e->d_part_i, e->d_part_k,
                         e->nb_i, e->nb_k, e->c1, e->c2, e->ls_max_iter, q + 1 == rest ? 1 : 0, e->d_hist,
                         e->max_hist, e->P, e->d_active, part_q, e->nb_q, gl_lam);
      MRH_CHECK_LAUNCH();
    }
    return MRH_OK;
  }
  if (e->fp32) hipLaunchKernelGGL((k_ls_image<0, NT_A, float>), dim3(e->nb_l, e->P), dim3(RT), 0, st, (const float2*)x,
                     (const float2*)d, (const float2*)g, e->tv, e->npix_p, e->ls_beta, e->d_st, e->tg_tiles_x,
                     e->tg_ntiles, e->tl_fg, e->d_part_i);
  else hipLaunchKernelGGL((k_ls_image<0, NT_A, double>), dim3(e->nb_l, e->P), dim3(RT), 0, st, (const float2*)x,
                     (const float2*)d, (const float2*)g, e->tv, e->npix_p, e->ls_beta, e->d_st, e->tg_tiles_x,
                     e->tg_ntiles, e->tl_fg, e->d_part_i);
  MRH_CHECK_LAUNCH();
  LAUNCH_R(e, k_ls_select, dim3(e->P), dim3(RT), 0, st, e->d_st, e->d_part_i, e->d_part_k, e->nb_l,
                     e->nb_k, e->c1, e->ls_beta, e->ls_max_iter, e->d_hist, e->max_hist, e->P, e->d_active,
                     part_q, e->nb_q, gl_lam);
  MRH_CHECK_LAUNCH();
  if (e->fp32) hipLaunchKernelGGL((k_ls_image<NT_A, NT, float>), dim3(e->nb_l, e->P), dim3(RT), 0, st, (const float2*)x,
                     (const float2*)d, (const float2*)g, e->tv, e->npix_p, e->ls_beta, e->d_st, e->tg_tiles_x,
                     e->tg_ntiles, e->tl_fg, e->d_part_i);
  else hipLaunchKernelGGL((k_ls_image<NT_A, NT, double>), dim3(e->nb_l, e->P), dim3(RT), 0, st, (const float2*)x,
                     (const float2*)d, (const float2*)g, e->tv, e->npix_p, e->ls_beta, e->d_st, e->tg_tiles_x,
                     e->tg_ntiles, e->tl_fg, e->d_part_i);
  MRH_CHECK_LAUNCH();
  LAUNCH_R(e, k_ls_select_rest, dim3(e->P), dim3(RT), 0, st, e->d_st, e->d_part_i, e->nb_l, e->c1,
                     e->ls_beta, e->ls_max_iter, e->d_hist, e->max_hist, e->P, e->d_active, gl_lam);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_cg_accept(mrh_cg* e, void* x, const void* d, void* Ax, const void* Ad, const void* y,
                             void* r, void* stream) {
  MRH_REQUIRE(e && x && d && Ax && Ad && y && r, MRH_EINVAL, "cg_accept: null argument");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_accept_img, dim3(e->nb_i, e->P), dim3(RT), 0, st, (float2*)x, (const float2*)d,
                     e->npix_p, e->d_st);
  MRH_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_accept_k, dim3(e->nb_k, e->P), dim3(RT), 0, st, (float2*)Ax, (const float2*)Ad,
                     (const float2*)y, (float2*)r, e->d_kprob, e->d_st);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_cg_gradient(mrh_cg* e, const void* x, void* g, const void* d, const void* g_prev,
                               int prime, void* stream) {
  MRH_REQUIRE(e && x && g, MRH_EINVAL, "cg_gradient: null argument");
  hipStream_t st = (hipStream_t)stream;
  const double* part_q = e->gl.lam != 0.f ? e->d_part_q : nullptr;
  if (part_q) {
    hipLaunchKernelGGL(k_gl_grad, dim3(e->nb_q, e->P), dim3(256), e->gl_lds_grad, st, (const float2*)x, (float2*)g,
                       e->gl, e->S0 * e->S1, e->npix_p, e->d_st, prime, e->d_part_q);
    MRH_CHECK_LAUNCH();
  }
  LAUNCH_R(e, k_cg_grad_tiled, dim3(e->nb_g, e->P), dim3(RT), 0, st, (const float2*)x, (float2*)g,
                     (const float2*)d, (const float2*)g_prev, (const float*)e->d_diag, e->tv, e->npix_p, e->d_st,
                     prime, e->tg_tiles_x, e->tg_ntiles, e->tg_fg, e->d_part_i);
  MRH_CHECK_LAUNCH();
  LAUNCH_R(e, k_cg_beta, dim3(e->P), dim3(RT), 0, st, e->d_st, e->d_part_i, e->nb_g, prime,
                     e->direction, e->rel_tol, e->restart_c, e->d_target, e->d_hist, e->max_hist, e->P,
                     e->d_active, part_q, e->nb_q);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_cg_direction(mrh_cg* e, void* d, const void* g, void* g_prev, int prime, void* stream) {
  MRH_REQUIRE(e && d && g, MRH_EINVAL, "cg_direction: null argument");
  hipLaunchKernelGGL(k_cg_direction, dim3(e->nb_i, e->P), dim3(RT), 0, (hipStream_t)stream, (float2*)d,
                     (const float2*)g, (float2*)g_prev, (const float*)e->d_diag, e->npix_p, e->d_st, prime);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_cg_build_diag(mrh_cg* e, const void* x, void* stream) {
  MRH_REQUIRE(e && x, MRH_EINVAL, "cg_build_diag: null argument");
  MRH_REQUIRE(e->d_diag, MRH_EINVAL, "cg_build_diag: the engine was created without a diag buffer");
  hipLaunchKernelGGL(k_build_diag, dim3(e->nb_i, e->P), dim3(RT), 0, (hipStream_t)stream, (const float2*)x,
                     e->d_diag, e->tv, e->gl, e->npix_p);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_cg_active_flags(mrh_cg* e, const int** flags) {
  MRH_REQUIRE(e && flags, MRH_EINVAL, "cg_active_flags: null argument");
  *flags = e->d_active;
  return MRH_OK;
}

extern "C" int mrh_cg_set_targets(mrh_cg* e, const double* targets, void* stream) {
  MRH_REQUIRE(e, MRH_EINVAL, "cg_set_targets: null engine");
  std::vector<double> t(e->P, -HUGE_VAL);
  if (targets)
    for (int p = 0; p < e->P; ++p) t[p] = std::isnan(targets[p]) ? -HUGE_VAL : targets[p];
  MRH_HIP(hipStreamSynchronize((hipStream_t)stream));
  MRH_HIP(hipMemcpy(e->d_target, t.data(), sizeof(double) * e->P, hipMemcpyHostToDevice));
  return MRH_OK;
}

extern "C" int mrh_cg_read_status(mrh_cg* e, int* status, int* iters, void* stream) {
  MRH_REQUIRE(e, MRH_EINVAL, "cg_read_status: null engine");
  // stream order puts the copy after the caller's work; one sync then covers both
  MRH_HIP(hipMemcpyAsync(e->h_st, e->d_st, sizeof(ProbState) * e->P, hipMemcpyDeviceToHost, (hipStream_t)stream));
  MRH_HIP(hipStreamSynchronize((hipStream_t)stream));
  for (int p = 0; p < e->P; ++p) {
    if (status) status[p] = e->h_st[p].status;
    if (iters) iters[p] = e->h_st[p].iter;
  }
  return MRH_OK;
}

extern "C" int mrh_cg_read_history(mrh_cg* e, mrh_cg_hist* hist, int n_iter, void* stream) {
  MRH_REQUIRE(e && hist, MRH_EINVAL, "cg_read_history: null argument");
  MRH_REQUIRE(n_iter >= 0 && n_iter <= e->max_hist, MRH_EINVAL, "n_iter outside [0, %d]", e->max_hist);
  MRH_HIP(hipStreamSynchronize((hipStream_t)stream));
  if (n_iter == 0) return MRH_OK;
  MRH_HIP(hipMemcpy(hist, e->d_hist, sizeof(mrh_cg_hist) * (size_t)n_iter * e->P, hipMemcpyDeviceToHost));
  return MRH_OK;
}

// last[p] = problem p's latest complete history record (zeros before the first one).
__global__ void k_gather_latest(const ProbState* __restrict__ st, const mrh_cg_hist* __restrict__ hist,
                                int max_hist, int P, mrh_cg_hist* __restrict__ last) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int n = min(st[p].iter, max_hist);
  mrh_cg_hist h;
  if (n > 0) {
    h = hist[(int64_t)(n - 1) * P + p];
  } else {
    double* z = (double*)&h;
    for (size_t i = 0; i < sizeof(mrh_cg_hist) / sizeof(double); ++i) z[i] = 0.0;
  }
  last[p] = h;
}

extern "C" int mrh_cg_read_latest(mrh_cg* e, int* status, int* iters, mrh_cg_hist* last, void* stream) {
  MRH_REQUIRE(e, MRH_EINVAL, "cg_read_latest: null engine");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_gather_latest, dim3((e->P + 63) / 64), dim3(64), 0, st, e->d_st, e->d_hist, e->max_hist, e->P,
                     e->d_last);
  MRH_CHECK_LAUNCH();
  MRH_HIP(hipMemcpyAsync(e->h_st, e->d_st, sizeof(ProbState) * e->P, hipMemcpyDeviceToHost, st));
  MRH_HIP(hipMemcpyAsync(e->h_last, e->d_last, sizeof(mrh_cg_hist) * e->P, hipMemcpyDeviceToHost, st));
  MRH_HIP(hipStreamSynchronize(st));
  for (int p = 0; p < e->P; ++p) {
    if (status) status[p] = e->h_st[p].status;
    if (iters) iters[p] = e->h_st[p].iter;
    if (last) last[p] = e->h_last[p];
  }
  return MRH_OK;
}
