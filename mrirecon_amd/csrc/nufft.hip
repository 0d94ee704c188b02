// Kaiser-Bessel gridding NUFFT (2-D SENSE operator) for MI355X / gfx950.
//
// Forward  A: x (nbatch, S0, S1) complex64 -> y (flat k-space)
//   1. k_fwd_rows   : x * map_c * deapod, centred zero-pad to G1, length-G1 FFT
//                     along the second image axis for every image row (LDS
//                     Stockham), written to the strip-major intermediate T
//                     (cf, strip, S0, Tc).  Also clears this coil-frame's y.
//   2. k_fwd_strips : one workgroup per (frame, strip of Tc grid columns,
//                     chunk of coils): loads the S0 x Tc block of T, FFTs the
//                     Tc columns in LDS (the first pass skips the zero
//                     padding to G0 rows), then interpolates every
//                     sample whose KB footprint touches the strip straight out
//                     of LDS.  A footprint spans at most two strips, so each
//                     sample receives at most two partial sums (atomic adds of
//                     two terms onto zero commute exactly).
// Adjoint  A^H: the same two passes reversed: k_adj_strips spreads the strip's
//   samples into an LDS-resident strip (exact int64 fixed-point LDS atomics),
//   inverse-FFTs the columns and writes only the S0 cropped rows; k_adj_rows inverse-FFTs each
//   row of every coil, crops, deapodises and accumulates conj(map_c) over the
//   coils in registers.
// The G0 x G1 oversampled grid never exists in HBM.
//
// Operator (pinned by the FP64 oracle):
//   y_k = wA_k sum_{i,j<W} phi(u_k - m0x - i) phi(v_k - m0y - j)
//         * FFT2[ pad( x * S_c * dpx dpy ) ][(m0x+i) mod G0, (m0y+j) mod G1]
//   u = w0 G0 / 2pi, m0x = floor(u - W/2) + 1, phi(d) = I0(beta sqrt(1-(2d/W)^2)) / I0(beta),
//   image index i placed at grid index (i - S/2) mod G.
#include "mrh_common.h"
#include "fft_lds.h"
#include "mrirecon_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

// ---------------------------------------------------------------- KB window
#define MRH_KB_MAXNP 32
struct KBPoly {
  float qscale;      // beta^2 / 4
  float two_over_w;  // 2 / W
  float a[MRH_KB_MAXNP];  // a_n = 1 / ((n!)^2 I0(beta)), zero-padded
};

// phi(d) = sum_n a_n q^n with q = (beta^2/4)(1 - (2d/W)^2): the power series of
// I0(beta sqrt(1 - (2d/W)^2)) / I0(beta) in Horner form (all terms positive).
template <int NP>
__device__ __forceinline__ float kb_weight(const KBPoly& kb, float d) {
  const float s = d * kb.two_over_w;
  const float q = kb.qscale * fmaxf(1.f - s * s, 0.f);
  float p = kb.a[NP - 1];
#pragma unroll
  for (int n = NP - 2; n >= 0; --n) p = fmaf(p, q, kb.a[n]);
  return p;
}

#define MRH_WMAX 8
#define STRIP_TC 8                     // default grid columns per strip workgroup (4 and 16 also built)
#define STRIP_EPT 1                    // sample footprints cached in registers per thread

// The run of footprint columns my ... my + W - 1 (mod G1) that falls in the strip
// [c0, c0 + TC): first footprint column jlo, first strip column col, length ncol
// (ncol <= 0: the footprint misses the strip).
__host__ __device__ __forceinline__ void strip_run(int my, int W, int c0, int TC, int G1, int& jlo, int& col,
                                                   int& ncol) {
  int rel = my - c0;
  if (rel < 0) rel += G1;
  if (rel < TC) {
    jlo = 0; col = rel; ncol = W < TC - rel ? W : TC - rel;
  } else {
    jlo = G1 - rel; col = 0; ncol = W - jlo < TC ? W - jlo : TC;
  }
}

// ---------------------------------------------------------------- plan
struct mrh_nufft_plan {
  int S0, S1, G0, G1, W, C, B, like_max, Tc, nstrips, np;
  int cchunk;   // coils per strip workgroup (autotuned; C = one workgroup per (frame, strip))
  double beta;
  int64_t Ktot, n_entries;
  int max_strip_entries;
  KBPoly kb;
  int device;
  int64_t* d_koff;
  int2* d_em;        // (m0x | m0y << 16, k)
  float4* d_ef;      // (u - m0x, v - m0y, wA, wAH)
  int* d_soff;       // B * nstrips + 1
  int* d_srank;      // B * nstrips: frame b's strips in decreasing order of entries (dispatch order)
  float* d_dx;
  float* d_dy;
  float2* d_tw0;
  float2* d_tw1;
  float2* d_maps;
  float2* d_T;
  size_t T_elems;
};

static double bessel_i0(double x) {
  double q = 0.25 * x * x, term = 1.0, sum = 1.0;
  for (int n = 1; n < 200; ++n) {
    term *= q / ((double)n * n);
    sum += term;
    if (term < 1e-18 * sum) break;
  }
  return sum;
}

// Fourier transform of the normalised KB window (Jackson et al. 1991).
static double kb_ft(double xi, int W, double beta) {
  double z = beta * beta - (M_PI * W * xi) * (M_PI * W * xi);
  double v;
  if (z > 0) { double s = std::sqrt(z); v = std::sinh(s) / s; }
  else if (z < 0) { double s = std::sqrt(-z); v = std::sin(s) / s; }
  else v = 1.0;
  return W * v / bessel_i0(beta);
}

static bool fft_size_ok(int n) {
  switch (n) {
    case 32: case 64: case 96: case 128: case 192: case 256: case 384:
    case 512: case 768: case 1024: return true;
    default: return false;
  }
}

static int free_plan(mrh_nufft_plan* p) {
  if (!p) return MRH_OK;
  void* bufs[] = {p->d_koff, p->d_em, p->d_ef, p->d_soff, p->d_srank, p->d_dx, p->d_dy,
                  p->d_tw0, p->d_tw1, p->d_maps, p->d_T};
  for (void* b : bufs) if (b) (void)hipFree(b);
  delete p;
  return MRH_OK;
}

template <typename T>
static int upload(T** dst, const T* src, size_t n) {
  *dst = nullptr;
  if (n == 0) n = 1;
  MRH_HIP(hipMalloc((void**)dst, n * sizeof(T)));
  if (src) MRH_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return MRH_OK;
}

// Pass-ordered twiddles of a length-n transform (the layout fft_pass reads,
// fft_tw_index); n entries, the last one unused.
static void twiddles(std::vector<float2>& tw, int n) {
  tw.assign(n, make_float2(1.f, 0.f));
  for (int ns = 1; ns < n;) {
    const int R = fft_next_radix(n / ns), tstep = n / (ns * R);
    for (int r = 1; r < R; ++r)
      for (int m = 0; m < ns; ++m) {
        const double a = -2.0 * M_PI * (double)(m * tstep * r) / (double)n;
        tw[fft_tw_index(ns, r, m)] = make_float2((float)std::cos(a), (float)std::sin(a));
      }
    ns *= R;
  }
}

// Runs fn(b) for every frame b, on up to 16 host threads when the plan is large
// (frames interleaved over the threads); fn must only touch frame b's data.
template <typename Fn>
static void for_each_frame(int B, int64_t nsamples, Fn fn) {
  int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (nsamples < (1 << 18) || B < 2) nt = 1;
  nt = std::min(nt, B);
  if (nt == 1) {
    for (int b = 0; b < B; ++b) fn(b);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] { for (int b = t; b < B; b += nt) fn(b); });
  for (auto& x : th) x.join();
}

extern "C" int mrh_nufft_plan_create(const mrh_nufft_desc* d, mrh_nufft_plan** out) {
  MRH_REQUIRE(d && out, MRH_EINVAL, "plan_create: null argument");
  *out = nullptr;
  MRH_REQUIRE(d->ndim == 2, MRH_EUNSUP, "this plan is 2-D (got ndim=%d; 3-D: mrh_nufft3_plan_create)", d->ndim);
  MRH_REQUIRE(d->S0 >= 2 && d->S1 >= 2, MRH_EINVAL, "image size must be >= 2");
  MRH_REQUIRE(d->G0 >= d->S0 && d->G1 >= d->S1, MRH_EINVAL, "grid must be at least the image size");
  MRH_REQUIRE(fft_size_ok(d->G0) && fft_size_ok(d->G1), MRH_EUNSUP,
              "grid sizes %d x %d not in the compiled FFT set {32..1024}", d->G0, d->G1);
  MRH_REQUIRE(d->W >= 2 && d->W <= MRH_WMAX, MRH_EUNSUP, "KB width %d outside [2, %d]", d->W, MRH_WMAX);
  MRH_REQUIRE(d->C >= 1 && d->B >= 1 && d->like >= 1, MRH_EINVAL, "C, B, like must be >= 1");
  MRH_REQUIRE(d->k_offsets && d->traj, MRH_EINVAL, "k_offsets and traj are required (host pointers)");
  MRH_REQUIRE(d->maps, MRH_EINVAL, "maps (device pointer) is required");
  int Tc = d->strip_cols > 0 ? d->strip_cols : STRIP_TC;
  MRH_REQUIRE(Tc == 4 || Tc == 8 || Tc == 16, MRH_EUNSUP, "strip width must be 4, 8 or 16 (got %d)", Tc);
  MRH_REQUIRE(Tc != 16 || d->G0 <= 512, MRH_EUNSUP, "16-column strips need G0 <= 512 (LDS), got %d", d->G0);
  MRH_REQUIRE(d->G1 % Tc == 0, MRH_EINVAL, "G1=%d not a multiple of the strip width %d", d->G1, Tc);
  MRH_REQUIRE(Tc >= d->W, MRH_EINVAL, "strip width %d must be >= KB width %d", Tc, d->W);
  MRH_REQUIRE(d->G0 < 65536 && d->G1 < 65536, MRH_EINVAL, "grid too large");

  mrh_nufft_plan* p = new mrh_nufft_plan();
  std::memset(p, 0, sizeof(*p));
  p->S0 = d->S0; p->S1 = d->S1; p->G0 = d->G0; p->G1 = d->G1; p->W = d->W;
  p->C = d->C; p->B = d->B; p->like_max = d->like; p->Tc = Tc; p->nstrips = d->G1 / Tc;
  p->cchunk = d->C;
  (void)hipGetDevice(&p->device);

  // KB parameters
  double osf = (double)d->G0 / (double)d->S0;
  double beta = d->beta > 0 ? d->beta
                            : M_PI * std::sqrt(std::pow((d->W / osf) * (osf - 0.5), 2) - 0.8);
  p->beta = beta;
  double i0b = bessel_i0(beta), q = beta * beta / 4.0;
  double coef[64];
  {
    double fact = 1.0;
    for (int n = 0; n < 64; ++n) {
      if (n > 0) fact *= n;
      coef[n] = 1.0 / (fact * fact * i0b);
    }
  }
  int np = 1;
  while (np < MRH_KB_MAXNP && coef[np] * std::pow(q, np) > 1e-10) ++np;
  MRH_REQUIRE(coef[np] * std::pow(q, np) < 1e-9, MRH_EUNSUP, "KB beta %.3f too large for the series", beta);
  p->np = np;
  if (Tc == 4 && !(d->W == 4 && np <= 20)) {
    free_plan(p);
    mrh_set_error("4-column strips need W = 4 with at most 20 KB series terms (W=%d, %d terms)", d->W, np);
    return MRH_EUNSUP;
  }
  p->kb.qscale = (float)q;
  p->kb.two_over_w = (float)(2.0 / d->W);
  for (int n = 0; n < MRH_KB_MAXNP; ++n) p->kb.a[n] = n < np ? (float)coef[n] : 0.f;

  // frame offsets
  const int B = d->B;
  std::vector<int64_t> koff(d->k_offsets, d->k_offsets + B + 1);
  for (int b = 0; b < B; ++b)
    if (koff[b + 1] < koff[b]) { free_plan(p); mrh_set_error("k_offsets must be non-decreasing"); return MRH_EINVAL; }
  p->Ktot = koff[B];
  if (p->Ktot >= (int64_t)1 << 31) { free_plan(p); mrh_set_error("too many samples for 32-bit entry indices"); return MRH_EUNSUP; }

  // footprints and strip binning: a stable counting sort by (frame, strip), two
  // passes over the samples, frames spread over host threads (each frame's
  // entries fill their own range, in sample order)
  const int W = d->W, G0 = d->G0, G1 = d->G1, ns = p->nstrips;
  const double sc0 = G0 / (2.0 * M_PI), sc1 = G1 / (2.0 * M_PI);
  struct FootH { int mx, my, s0, s1; float ux, uy; };
  auto footprint = [&](int b, int64_t k) {
    const int64_t Kb = koff[b + 1] - koff[b];
    const float* t0 = d->traj + 2 * koff[b];
    const double u = (double)t0[k] * sc0, v = (double)t0[Kb + k] * sc1;
    const double fu = std::floor(u - 0.5 * W) + 1.0, fv = std::floor(v - 0.5 * W) + 1.0;
    FootH f;
    f.mx = (int)((((int64_t)fu % G0) + G0) % G0);
    f.my = (int)((((int64_t)fv % G1) + G1) % G1);
    f.ux = (float)(u - fu);
    f.uy = (float)(v - fv);
    f.s0 = f.my / Tc;
    f.s1 = ((f.my + W - 1) % G1) / Tc;
    return f;
  };
  std::vector<int> cnt((size_t)B * ns, 0);
  for_each_frame(B, p->Ktot, [&](int b) {
    int* cb = cnt.data() + (size_t)b * ns;
    for (int64_t k = 0, Kb = koff[b + 1] - koff[b]; k < Kb; ++k) {
      const FootH f = footprint(b, k);
      cb[f.s0]++;
      if (f.s1 != f.s0) cb[f.s1]++;
    }
  });
  std::vector<int> soff((size_t)B * ns + 1, 0);
  int64_t acc = 0;
  int maxe = 0;
  for (size_t i = 0; i < (size_t)B * ns; ++i) {
    soff[i] = (int)acc;
    acc += cnt[i];
    maxe = std::max(maxe, cnt[i]);
  }
  soff[(size_t)B * ns] = (int)acc;
  // Strip kernels run a frame's strips heaviest first (longest-processing-time order): the
  // strips around the k-space centre hold up to twenty times the entries of the outer ones, and
  // started last they would be the launch's tail.
  std::vector<int> srank((size_t)B * ns);
  for (int b = 0; b < B; ++b) {
    int* rk = srank.data() + (size_t)b * ns;
    for (int st = 0; st < ns; ++st) rk[st] = st;
    const int* cb = cnt.data() + (size_t)b * ns;
    std::stable_sort(rk, rk + ns, [cb](int u, int v) { return cb[u] > cb[v]; });
  }
  p->n_entries = acc;
  p->max_strip_entries = maxe;
  std::unique_ptr<int2[]> em(new int2[(size_t)acc]);
  std::unique_ptr<float4[]> ef(new float4[(size_t)acc]);
  for_each_frame(B, p->Ktot, [&](int b) {
    std::vector<int> cur(soff.begin() + (size_t)b * ns, soff.begin() + (size_t)(b + 1) * ns);
    for (int64_t k = 0, Kb = koff[b + 1] - koff[b]; k < Kb; ++k) {
      const FootH f = footprint(b, k);
      float dval = 1.f;
      if (d->dcf) dval = d->dcf[koff[b] + k];
      float wA = 1.f, wAH = 1.f;
      if (d->dcf && d->dcf_mode == MRH_DCF_BALANCED) {
        wA = wAH = (float)std::sqrt(std::max((double)dval, 0.0));
      } else if (d->dcf && d->dcf_mode == MRH_DCF_STANDARD) {
        wA = d->dcf_in_fwd ? dval : 1.f;
        wAH = d->dcf_in_adj ? dval : 1.f;
      }
      const int2 m = make_int2(f.mx | (f.my << 16), (int)k);
      const float4 fv = make_float4(f.ux, f.uy, wA, wAH);
      int e = cur[f.s0]++;
      em[e] = m; ef[e] = fv;
      if (f.s1 != f.s0) {
        e = cur[f.s1]++;
        em[e] = m; ef[e] = fv;
      }
    }
  });

  // deapodisation 1 / Phi((i - S/2) / G)
  std::vector<float> dx(d->S0), dy(d->S1);
  for (int i = 0; i < d->S0; ++i) dx[i] = (float)(1.0 / kb_ft((double)(i - d->S0 / 2) / G0, W, beta));
  for (int i = 0; i < d->S1; ++i) dy[i] = (float)(1.0 / kb_ft((double)(i - d->S1 / 2) / G1, W, beta));
  std::vector<float2> tw0, tw1;
  twiddles(tw0, G0);
  twiddles(tw1, G1);

  int rc;
#define UP(dst, src, n) if ((rc = upload(&(dst), (src), (n))) != MRH_OK) { free_plan(p); return rc; }
  UP(p->d_koff, koff.data(), koff.size());
  UP(p->d_em, em.get(), (size_t)acc);
  UP(p->d_ef, ef.get(), (size_t)acc);
  UP(p->d_soff, soff.data(), soff.size());
  UP(p->d_srank, srank.data(), srank.size());
  UP(p->d_dx, dx.data(), dx.size());
  UP(p->d_dy, dy.data(), dy.size());
  UP(p->d_tw0, tw0.data(), tw0.size());
  UP(p->d_tw1, tw1.data(), tw1.size());
  size_t nmaps = (size_t)d->C * d->S0 * d->S1;
  UP(p->d_maps, (const float2*)nullptr, nmaps);
#undef UP
  hipError_t e = hipMemcpy(p->d_maps, d->maps, nmaps * sizeof(float2), hipMemcpyDeviceToDevice);
  if (e != hipSuccess) { free_plan(p); mrh_set_error("maps copy failed: %s", hipGetErrorString(e)); return MRH_EHIP; }
  p->T_elems = (size_t)B * d->like * d->C * d->S0 * G1;
  e = hipMalloc((void**)&p->d_T, p->T_elems * sizeof(float2));
  if (e != hipSuccess) {
    free_plan(p);
    mrh_set_error("workspace allocation of %zu bytes failed: %s", p->T_elems * sizeof(float2), hipGetErrorString(e));
    return MRH_ENOMEM;
  }
  *out = p;
  return MRH_OK;
}

extern "C" int mrh_nufft_plan_destroy(mrh_nufft_plan* p) { return free_plan(p); }

extern "C" int mrh_nufft_set_coil_chunk(mrh_nufft_plan* p, int coils_per_block) {
  MRH_REQUIRE(p, MRH_EINVAL, "set_coil_chunk: null plan");
  MRH_REQUIRE(coils_per_block >= 1 && coils_per_block <= p->C, MRH_EINVAL,
              "coils per strip workgroup must be in [1, %d] (got %d)", p->C, coils_per_block);
  p->cchunk = coils_per_block;
  return MRH_OK;
}

extern "C" int mrh_nufft_plan_info(const mrh_nufft_plan* p, mrh_nufft_info* info) {
  MRH_REQUIRE(p && info, MRH_EINVAL, "plan_info: null argument");
  info->G0 = p->G0; info->G1 = p->G1; info->W = p->W; info->strip_cols = p->Tc;
  info->n_strips = p->nstrips; info->kb_terms = p->np; info->beta = p->beta;
  info->n_samples = p->Ktot; info->n_entries = p->n_entries;
  info->max_strip_entries = p->max_strip_entries;
  info->coil_chunk = p->cchunk;
  info->workspace_bytes = (int64_t)(p->T_elems * sizeof(float2));
  return MRH_OK;
}

// ---------------------------------------------------------------- kernels
#define ROWS_PER_BLOCK 4

// Frames of stopped problems can be skipped: problem p = (bb / like) / fpp is
// processed only when m[p] != 0 (m == nullptr: every frame).  Workgroups of
// skipped frames return before touching memory, so their k-space / image
// entries keep their previous values.
struct FrameMask {
  const int* m;
  int fpp, like;
};
__device__ __forceinline__ bool frame_skipped(const FrameMask& fm, int bb) {
  return fm.m != nullptr && fm.m[(bb / fm.like) / fm.fpp] == 0;
}

// Forward row pass.  Grid (ceil(S0/4), nbatch), 256 threads: one wavefront per
// image row, looping over the coils (the deapodised image row is loaded once;
// the next coil's map row is fetched while the current row is transformed).
template <int G1>
__global__ __launch_bounds__(256) void k_fwd_rows(
    const float2* __restrict__ x, const float2* __restrict__ maps, const float* __restrict__ dx,
    const float* __restrict__ dy, const float2* __restrict__ tw1, float2* __restrict__ T,
    float2* __restrict__ y, const int64_t* __restrict__ koff, int S0, int S1, int C, int like,
    int Tc, int x_has_coils, FrameMask fm) {
  constexpr int LP = fft_padded(G1);
  constexpr int NQ = (G1 + 63) / 64;
  __shared__ float2 sbuf[ROWS_PER_BLOCK * LP];
  __shared__ float2 stw[G1];
  const int tid = threadIdx.x, lane = tid & 63, wrow = tid >> 6;
  const int bb = blockIdx.y;
  if (frame_skipped(fm, bb)) return;
  const int ix = blockIdx.x * ROWS_PER_BLOCK + wrow;
  for (int i = tid; i < G1; i += blockDim.x) stw[i] = tw1[i];
  {  // clear this block's share of the frame's k-space (the strip pass accumulates into it)
    const int b = bb / like, l = bb - b * like;
    const int64_t Kb = koff[b + 1] - koff[b];
    const int64_t base = (int64_t)C * (koff[b] * like + (int64_t)l * Kb);
    const int64_t n = (int64_t)C * Kb;
    const int64_t lo = n * blockIdx.x / gridDim.x, hi = n * (blockIdx.x + 1) / gridDim.x;
    for (int64_t k = lo + tid; k < hi; k += blockDim.x) y[base + k] = make_float2(0.f, 0.f);
  }
  float2* buf = sbuf + wrow * LP;
  const int h1 = S1 / 2;
  const bool row_ok = ix < S0;
  // grid column pp holds image column iy with (iy - S1/2) mod G1 == pp
  float2 basev[NQ], fac[NQ];
  int iyq[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int pp = lane + 64 * q;
    const bool lo = pp < S1 - h1, hi = pp >= G1 - h1 && pp < G1;
    iyq[q] = (row_ok && (lo || hi)) ? (lo ? pp + h1 : pp - G1 + h1) : -1;
    basev[q] = make_float2(0.f, 0.f);
    if (iyq[q] >= 0) {
      const float dd = dx[ix] * dy[iyq[q]];
      basev[q] = x_has_coils ? make_float2(dd, 0.f)
                             : cscale(x[((int64_t)bb * S0 + ix) * S1 + iyq[q]], dd);
    }
  }
  auto load_fac = [&](int c) {
    const float2* src = x_has_coils ? x + ((int64_t)(bb * C + c) * S0 + ix) * S1
                                    : maps + ((int64_t)c * S0 + ix) * S1;
#pragma unroll
    for (int q = 0; q < NQ; ++q) fac[q] = iyq[q] >= 0 ? src[iyq[q]] : make_float2(0.f, 0.f);
  };
  load_fac(0);
  __syncthreads();
  const int nstr = G1 / Tc;
  for (int c = 0; c < C; ++c) {
    if constexpr (G1 == 512) {   // the lane's grid columns are its FFT rows: passes 1 and 3 in registers
      float2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = cmul(basev[q], fac[q]);
      if (c + 1 < C) load_fac(c + 1);
      fft512_wave_regs<false>(v, buf, stw, lane);
      if (row_ok) {
        float2* Tcf = T + (int64_t)(bb * C + c) * nstr * S0 * Tc;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int pp = lane + 64 * q;
          const int st = pp / Tc, j = pp - st * Tc;
          Tcf[((int64_t)st * S0 + ix) * Tc + j] = v[q];
        }
      }
      continue;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int pp = lane + 64 * q;
      if (pp < G1) buf[PIDX(pp)] = cmul(basev[q], fac[q]);
    }
    if (c + 1 < C) load_fac(c + 1);
    fft_lds_wave<G1, false>(buf, stw, lane);
    if (row_ok) {
      float2* Tcf = T + (int64_t)(bb * C + c) * nstr * S0 * Tc;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int pp = lane + 64 * q;
        if (pp < G1) {
          const int st = pp / Tc, j = pp - st * Tc;
          Tcf[((int64_t)st * S0 + ix) * Tc + j] = buf[PIDX(pp)];
        }
      }
    }
    fft_sync<true>();
  }
}

// ---------------------------------------------------------------- strip kernels
// One workgroup per (batch frame, strip of TC grid columns); it loops
// over the coils of the frame.  Each thread evaluates the KB footprint of up to
// STRIP_EPT of the strip's samples once and keeps it in registers, so the
// weights are shared by all coils; samples beyond that spill to an
// evaluate-on-the-fly path.
//
// A footprint's columns inside the strip are a contiguous run: ncol columns
// starting at strip column col, footprint columns jlo ... jlo + ncol - 1; only
// their weights are kept.
template <int WM>
struct Foot {
  float wx[WM], wy[WM];  // wy[jj]: weight of footprint column jlo + jj (jj < ncol), else 0
  int r0, col, ncol, k;  // first footprint row (mod G), first strip column, columns inside, sample (-1: empty)
  float wA, wAH;
};

// Horner terms evaluated per footprint width: enough for every plan that the
// dispatch sends to that width (W=4 needs at most 20 terms, W=6 at most 24);
// the remaining coefficients are zero, so the value equals the full series.
__host__ __device__ constexpr int kb_terms(int wm) { return wm == 4 ? 20 : wm == 6 ? 24 : MRH_KB_MAXNP; }

template <int WM, int TC>
__device__ __forceinline__ void load_foot(Foot<WM>& F, int2 m, float4 f, int W, int c0, int G1,
                                          const KBPoly& kb) {
  F.r0 = m.x & 0xffff;
  const int my = (int)((unsigned)m.x >> 16);
  F.k = m.y;
  F.wA = f.z;
  F.wAH = f.w;
  int jlo;
  strip_run(my, W, c0, TC, G1, jlo, F.col, F.ncol);
#pragma unroll
  for (int i = 0; i < WM; ++i) F.wx[i] = (i < W) ? kb_weight<kb_terms(WM)>(kb, f.x - (float)i) : 0.f;
#pragma unroll
  for (int jj = 0; jj < WM; ++jj)
    F.wy[jj] = (jj < F.ncol) ? kb_weight<kb_terms(WM)>(kb, f.y - (float)(jlo + jj)) : 0.f;
}

// Forward: sum of the footprint over the LDS strip (column-major, padded rows),
// column by column in footprint order; the W reads of a column are issued together.
template <int WM, int CS>
__device__ __forceinline__ float2 interp_one(const Foot<WM>& F, const float2* smem, int W, int G0) {
  int ro[WM];
#pragma unroll
  for (int i = 0; i < WM; ++i) {
    int rr = F.r0 + i;
    if (rr >= G0) rr -= G0;
    ro[i] = PIDX(rr);
  }
  float2 acc = make_float2(0.f, 0.f);
#pragma unroll
  for (int jj = 0; jj < WM; ++jj) {
    if (jj < F.ncol) {
      const float2* colp = smem + (F.col + jj) * CS;
      float2 g[WM];
#pragma unroll
      for (int i = 0; i < WM; ++i) g[i] = colp[ro[i]];
      float2 rs = make_float2(0.f, 0.f);
#pragma unroll
      for (int i = 0; i < WM; ++i) {
        if (i < W) {
          rs.x = fmaf(g[i].x, F.wx[i], rs.x);
          rs.y = fmaf(g[i].y, F.wx[i], rs.y);
        }
      }
      acc.x = fmaf(rs.x, F.wy[jj], acc.x);
      acc.y = fmaf(rs.y, F.wy[jj], acc.y);
    }
  }
  return acc;
}

// Forward strip pass.  Grid (coil chunks, nbatch, nstrips): blockIdx.z is the rank of the strip in
// its frame's heaviest-first order (srank); 64 TC threads (one wavefront per column).
template <int G0, int WT, int TC>
__global__ __launch_bounds__(64 * TC, (G0 <= 512 ? 4 : 2)) void k_fwd_strips(
    const float2* __restrict__ T, float2* __restrict__ y, const int2* __restrict__ em,
    const float4* __restrict__ ef, const int* __restrict__ soff, const int* __restrict__ srank,
    const int64_t* __restrict__ koff, const float2* __restrict__ tw0, KBPoly kb, int S0, int G1, int C, int like,
    int Wrt, int cchunk, FrameMask fm) {
  constexpr int NTH = 64 * TC, CS = fft_padded(G0) + 1;
  constexpr int WM = WT ? WT : MRH_WMAX;
  const int W = WT ? WT : Wrt;
  if (frame_skipped(fm, blockIdx.y)) return;
  const int cb = blockIdx.x * cchunk, ce = min(C, cb + cchunk);   // this workgroup's coils
  __shared__ float2 smem2[2 * TC * CS];   // double buffer: coil c fills one half while c-1 is read
  __shared__ float2 stw[G0];
  const int tid = threadIdx.x, lane = tid & 63, col = tid >> 6;
  const int bb = blockIdx.y, nstr = gridDim.z;
  const int b = bb / like, l = bb - b * like;
  const int s = srank[b * nstr + blockIdx.z];
  for (int i = tid; i < G0; i += NTH) stw[i] = tw0[i];
  const int e0 = soff[b * nstr + s], e1 = soff[b * nstr + s + 1];
  const int64_t Kb = koff[b + 1] - koff[b];
  float2* ybase = y + (int64_t)C * (koff[b] * like + (int64_t)l * Kb);
  const int c0 = s * TC;
  Foot<WM> F[STRIP_EPT];
#pragma unroll
  for (int q = 0; q < STRIP_EPT; ++q) {
    const int e = e0 + tid + q * NTH;
    if (e < e1) load_foot<WM, TC>(F[q], em[e], ef[e], W, c0, G1, kb);
    else F[q].k = -1;
  }
  const int h0 = S0 / 2;
  const int64_t cstride = (int64_t)nstr * S0 * TC;
  const float2* Tbb = T + ((int64_t)bb * C * nstr + s) * S0 * TC;
  // the S0 x TC data block of T for coil c is fetched into registers one coil ahead
  constexpr int DPT = ((G0 / 2) * TC + NTH - 1) / NTH;
  const int ndata = S0 * TC;
  float2 pre[DPT];
  auto fetch = [&](int c) {
    const float2* Tb = Tbb + c * cstride;
#pragma unroll
    for (int t = 0; t < DPT; ++t) {
      const int e = tid + t * NTH;
      pre[t] = e < ndata ? Tb[e] : make_float2(0.f, 0.f);
    }
  };
  fetch(cb);
  for (int c = cb; c < ce; ++c) {
    float2* smem = smem2 + (c & 1) * TC * CS;
#pragma unroll
    for (int t = 0; t < DPT; ++t) {
      const int e = tid + t * NTH;
      if (e < ndata) {
        const int ix = e / TC, j = e - ix * TC;
        const int r = ix >= h0 ? ix - h0 : ix - h0 + G0;
        smem[j * CS + PIDX(r)] = pre[t];
      }
    }
    for (int e = DPT * NTH + tid; e < ndata; e += NTH) {   // images wider than G0/2 (osf < 2)
      const int ix = e / TC, j = e - ix * TC;
      const int r = ix >= h0 ? ix - h0 : ix - h0 + G0;
      smem[j * CS + PIDX(r)] = Tbb[c * cstride + e];
    }
    if (c + 1 < ce) fetch(c + 1);
    __syncthreads();                       // strip filled (and every thread is past coil c-1)
    // rows [S0 - h0, G0 - h0) are the zero padding: never written, skipped by the first pass
    fft_lds_wave_padded<G0, false>(smem + col * CS, stw, lane, S0 - h0, G0 - h0);
    __syncthreads();                       // all columns transformed
    float2* yc = ybase + (int64_t)c * Kb;
#pragma unroll
    for (int q = 0; q < STRIP_EPT; ++q) {
      if (F[q].k >= 0) {
        const float2 a = interp_one<WM, CS>(F[q], smem, W, G0);
        unsafeAtomicAdd(&yc[F[q].k].x, a.x * F[q].wA);
        unsafeAtomicAdd(&yc[F[q].k].y, a.y * F[q].wA);
      }
    }
    for (int e = e0 + STRIP_EPT * NTH + tid; e < e1; e += NTH) {
      Foot<WM> Fo;
      load_foot<WM, TC>(Fo, em[e], ef[e], W, c0, G1, kb);
      const float2 a = interp_one<WM, CS>(Fo, smem, W, G0);
      unsafeAtomicAdd(&yc[Fo.k].x, a.x * Fo.wA);
      unsafeAtomicAdd(&yc[Fo.k].y, a.y * Fo.wA);
    }
  }
}

// Running max of |re|, |im| over the weighted samples; a non-finite sample makes it +inf
// (fmaxf alone would drop a NaN).
__device__ __forceinline__ float fmax_finite(float m, float2 v) {
  return (isfinite(v.x) && isfinite(v.y)) ? fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y))) : INFINITY;
}

// Deterministic fixed-point accumulation for the adjoint.  Per strip and coil
// the weighted samples are scaled by 2^sc so that the largest component is
// below 2^30; every contribution is rounded once to an integer (a quantum of
// 2^-30 of that largest component, far below FP32 rounding of the large
// terms) and summed exactly as int64 with ds_add_u64, which runs at full LDS
// rate on gfx950 where float LDS atomics do not.  The sum is independent of
// the order of the adds.
__device__ __forceinline__ unsigned long long fixed_of(float scaled) {
  const int q = (int)rintf(scaled);
  return (unsigned long long)(long long)q;
}

// Back to float, times 2^-sc, in 32-bit arithmetic: q = hi 2^32 + lo with lo the
// low word read as signed (|lo| <= 2^31) and hi adjusted to match, so q < 2^31
// converts exactly rounded and larger q within one unit in the last place of
// the correctly rounded value (the fma's product is exact, so only the rounding
// of lo and the fma's own rounding enter); six 32-bit instructions in place of
// a double-precision path.
__device__ __forceinline__ float from_fixed(unsigned long long u, int sc) {
  const int lo = (int)(unsigned)u;
  const int hi = (int)(unsigned)(u >> 32) - (lo >> 31);
  return ldexpf(fmaf((float)hi, 4294967296.f, (float)lo), -sc);
}

// Inverse length-N FFT of one column held as fixed-point planes (re at word
// PIDX(row) of acc, im at word CS + PIDX(row)), by one wavefront: the first
// Stockham pass converts while it loads, writes its floats to buf[PIDX(.)] (buf
// aliases the re plane) and the remaining passes run in place.  The result
// equals converting every row first and calling fft_lds_wave.
template <int N, int CS>
__device__ __forceinline__ void fft_wave_from_fixed(float2* buf, const unsigned long long* acc, const float2* tw,
                                                    int lane, int sc) {
  constexpr int R = fft_next_radix(N), NBF = N / R, NB = (NBF + 63) / 64;
  fft_sync<true>();
  float2 v[NB][R];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int j = lane + q * 64;
    if (NBF % 64 == 0 || j < NBF) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int w = PIDX(j + r * NBF);
        v[q][r] = make_float2(from_fixed(acc[w], sc), from_fixed(acc[CS + w], sc));
      }
      DFT<R, true>::run(v[q]);
    }
  }
  fft_sync<true>();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int j = lane + q * 64;
    if (NBF % 64 == 0 || j < NBF) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[PIDX(j * R + r)] = v[q][r];
    }
  }
  fft_sync<true>();
  FFTStages<N, R, true, 64, true>::run(buf, tw, lane);
}

// Adjoint footprint: the weights of all W columns, zero for the columns
// outside the strip (the spread runs column j only where its weight is
// non-zero).  Measured on the GRASP workload, this form spreads 3 % faster than
// the forward's compact run of columns, which in turn interpolates 16 % faster.
template <int WM>
struct FootA {
  float wx[WM], wy[WM];  // wy[j] = 0 when column (my + j) mod G1 lies outside the strip
  int r0, my, k;         // first footprint row / column (mod G), sample index in the frame (-1: empty)
  float wA, wAH;
};

template <int WM, int TC>
__device__ __forceinline__ void load_foot_adj(FootA<WM>& F, int2 m, float4 f, int W, int c0, int G1,
                                              const KBPoly& kb) {
  F.r0 = m.x & 0xffff;
  F.my = (int)((unsigned)m.x >> 16);
  F.k = m.y;
  F.wA = f.z;
  F.wAH = f.w;
#pragma unroll
  for (int i = 0; i < WM; ++i) F.wx[i] = (i < W) ? kb_weight<kb_terms(WM)>(kb, f.x - (float)i) : 0.f;
#pragma unroll
  for (int j = 0; j < WM; ++j) {
    int cc = F.my + j;
    if (cc >= G1) cc -= G1;
    F.wy[j] = (j < W && (unsigned)(cc - c0) < (unsigned)TC) ? kb_weight<kb_terms(WM)>(kb, f.y - (float)j) : 0.f;
  }
}

// vs: the weighted sample already scaled by 2^sc.  Column j of the strip owns
// words [2 j CS, 2 j CS + 2 CS) of acc: row rr's re at PIDX(rr), its im at CS +
// PIDX(rr) (two planes: lanes on consecutive rows hit consecutive 8-byte words,
// half the bank conflicts of interleaved (re, im) pairs).
template <int WM, int CS>
__device__ __forceinline__ void spread_one(const FootA<WM>& F, float2 vs, unsigned long long* acc, int W, int c0,
                                           int G0, int G1) {
#pragma unroll
  for (int j = 0; j < WM; ++j) {
    if (F.wy[j] != 0.f) {
      int cc = F.my + j;
      if (cc >= G1) cc -= G1;
      unsigned long long* colp = acc + 2 * (cc - c0) * CS;
      const float2 vj = cscale(vs, F.wy[j]);
#pragma unroll
      for (int i = 0; i < WM; ++i) {
        if (i < W) {
          int rr = F.r0 + i;
          if (rr >= G0) rr -= G0;
          unsigned long long* dst = colp + PIDX(rr);
          atomicAdd(dst, fixed_of(vj.x * F.wx[i]));
          atomicAdd(dst + CS, fixed_of(vj.y * F.wx[i]));
        }
      }
    }
  }
}

// Adjoint strip pass: fixed-point spreading into the LDS strip, inverse column
// FFT, crop to the S0 image rows.  Grid (coil chunks, nbatch, nstrips) as in the forward, 64 TC threads.
template <int G0, int WT, int TC>
__global__ __launch_bounds__(64 * TC, (G0 <= 512 ? 4 : 2)) void k_adj_strips(
    const float2* __restrict__ y, float2* __restrict__ T, const int2* __restrict__ em,
    const float4* __restrict__ ef, const int* __restrict__ soff, const int* __restrict__ srank,
    const int64_t* __restrict__ koff, const float2* __restrict__ tw0, KBPoly kb, int S0, int G1, int C, int like,
    int Wrt, int cchunk, FrameMask fm) {
  constexpr int NTH = 64 * TC, CS = fft_padded(G0) + 1;
  constexpr int WM = WT ? WT : MRH_WMAX;
  const int W = WT ? WT : Wrt;
  if (frame_skipped(fm, blockIdx.y)) return;
  const int cb = blockIdx.x * cchunk, ce = min(C, cb + cchunk);   // this workgroup's coils
  __shared__ unsigned long long acc[2 * TC * CS];
  __shared__ float2 stw[G0];
  __shared__ float wmax[NTH / 64];
  float2* fbuf = reinterpret_cast<float2*>(acc);
  const int tid = threadIdx.x, lane = tid & 63, col = tid >> 6;
  const int bb = blockIdx.y, nstr = gridDim.z;
  const int b = bb / like, l = bb - b * like;
  const int s = srank[b * nstr + blockIdx.z];
  for (int i = tid; i < G0; i += NTH) stw[i] = tw0[i];
  const int e0 = soff[b * nstr + s], e1 = soff[b * nstr + s + 1];
  const int64_t Kb = koff[b + 1] - koff[b];
  const float2* ybase = y + (int64_t)C * (koff[b] * like + (int64_t)l * Kb);
  const int c0 = s * TC;
  FootA<WM> F[STRIP_EPT];
#pragma unroll
  for (int q = 0; q < STRIP_EPT; ++q) {
    const int e = e0 + tid + q * NTH;
    if (e < e1) load_foot_adj<WM, TC>(F[q], em[e], ef[e], W, c0, G1, kb);
    else F[q].k = -1;
  }
  const int h0 = S0 / 2;
  const int64_t cstride = (int64_t)nstr * S0 * TC;
  float2* Tbb = T + ((int64_t)bb * C * nstr + s) * S0 * TC;
  // weighted samples of the cached footprints, fetched one coil ahead
  float2 vn[STRIP_EPT];
  auto fetch = [&](int c) {
    const float2* yc = ybase + (int64_t)c * Kb;
#pragma unroll
    for (int q = 0; q < STRIP_EPT; ++q)
      vn[q] = F[q].k >= 0 ? cscale(yc[F[q].k], F[q].wAH) : make_float2(0.f, 0.f);
  };
  fetch(cb);
  for (int i = tid; i < 2 * TC * CS; i += NTH) acc[i] = 0ull;
  for (int c = cb; c < ce; ++c) {
    const float2* yc = ybase + (int64_t)c * Kb;
    float2 v[STRIP_EPT];
    float m = 0.f;
#pragma unroll
    for (int q = 0; q < STRIP_EPT; ++q) {
      v[q] = vn[q];
      m = fmax_finite(m, v[q]);
    }
    if (c + 1 < ce) fetch(c + 1);
    for (int e = e0 + STRIP_EPT * NTH + tid; e < e1; e += NTH) {
      const float2 vo = cscale(yc[em[e].y], ef[e].w);
      m = fmax_finite(m, vo);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (lane == 0) wmax[col] = m;
    __syncthreads();
    float vmax = 0.f;
#pragma unroll
    for (int w = 0; w < NTH / 64; ++w) vmax = fmaxf(vmax, wmax[w]);
    // vmax = +inf: a weighted sample of the strip is not finite; nothing is spread and the
    // strip's output for this coil is NaN (as the NaN or inf would reach every grid point of
    // the strip's columns through the inverse FFT)
    const bool bad = !(vmax < INFINITY);
    int ex = 0;
    if (!bad) frexpf(vmax, &ex);   // vmax < 2^ex
    const int sc = 30 - ex;
    if (vmax > 0.f && !bad) {
#pragma unroll
      for (int q = 0; q < STRIP_EPT; ++q)
        if (F[q].k >= 0)
          spread_one<WM, CS>(F[q], make_float2(ldexpf(v[q].x, sc), ldexpf(v[q].y, sc)), acc, W, c0, G0, G1);
      for (int e = e0 + STRIP_EPT * NTH + tid; e < e1; e += NTH) {
        FootA<WM> Fo;
        load_foot_adj<WM, TC>(Fo, em[e], ef[e], W, c0, G1, kb);
        const float2 vo = cscale(yc[Fo.k], Fo.wAH);
        spread_one<WM, CS>(Fo, make_float2(ldexpf(vo.x, sc), ldexpf(vo.y, sc)), acc, W, c0, G0, G1);
      }
    }
    __syncthreads();                       // every contribution is in
    // Inverse column FFT by the wavefront that owns the column; its first pass
    // reads the fixed-point accumulators and writes floats into the column's own
    // int64 region (float slot p is the region's word p).  The wavefront then
    // zeroes the rows that the read-out below does not visit (the read-out
    // zeroes the ones it reads) and the im plane.  Slots that hold no row (the
    // padding holes of PIDX and the tail) are never written after the initial
    // clear, so they stay zero.
    {
      unsigned long long* ac = acc + 2 * col * CS;
      float2* fc = fbuf + 2 * col * CS;
      fft_wave_from_fixed<G0, CS>(fc, ac, stw, lane, sc);
#pragma unroll
      for (int k = 0; k < (CS + 63) / 64; ++k) {   // the im plane
        const int q = lane + 64 * k;
        if (CS % 64 == 0 || q < CS) ac[CS + q] = 0ull;
      }
      for (int r = S0 - h0 + lane; r < G0 - h0; r += 64) fc[PIDX(r)] = make_float2(0.f, 0.f);
    }
    __syncthreads();                       // all columns transformed
    float2* Tb = Tbb + c * cstride;
    for (int e = tid; e < S0 * TC; e += NTH) {
      const int ix = e / TC, j = e - ix * TC;
      const int r = ix >= h0 ? ix - h0 : ix - h0 + G0;
      float2* sp = fbuf + 2 * j * CS + PIDX(r);
      Tb[e] = bad ? make_float2(__int_as_float(0x7fc00000), __int_as_float(0x7fc00000)) : *sp;
      *sp = make_float2(0.f, 0.f);
    }
  }
}

// Adjoint row pass: per image row, inverse row FFT of every coil, crop,
// conj(map) combine in registers, deapodise.  The next coil's row of T and of
// the map are fetched while the current row is transformed.  Grid (ceil(S0/4), nbatch).
template <int G1>
__global__ __launch_bounds__(256) void k_adj_rows(
    const float2* __restrict__ T, const float2* __restrict__ maps, const float* __restrict__ dx,
    const float* __restrict__ dy, const float2* __restrict__ tw1, float2* __restrict__ x,
    int S0, int S1, int C, int Tc, FrameMask fm) {
  constexpr int LP = fft_padded(G1);
  constexpr int NA = (G1 + 63) / 64;
  __shared__ float2 sbuf[ROWS_PER_BLOCK * LP];
  __shared__ float2 stw[G1];
  const int tid = threadIdx.x, lane = tid & 63, wrow = tid >> 6;
  const int bb = blockIdx.y;
  if (frame_skipped(fm, bb)) return;
  const int ix = blockIdx.x * ROWS_PER_BLOCK + wrow;
  const bool row_ok = ix < S0;
  for (int i = tid; i < G1; i += blockDim.x) stw[i] = tw1[i];
  float2* buf = sbuf + wrow * LP;
  const int nstr = G1 / Tc, h1 = S1 / 2;
  float2 acc[NA], tp[NA], mp[NA];
#pragma unroll
  for (int q = 0; q < NA; ++q) acc[q] = make_float2(0.f, 0.f);
  auto load = [&](int c) {
    const float2* Tcf = T + (int64_t)(bb * C + c) * nstr * S0 * Tc;
    const float2* mr = maps + ((int64_t)c * S0 + ix) * S1;
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int pp = lane + 64 * q;
      tp[q] = make_float2(0.f, 0.f);
      mp[q] = make_float2(0.f, 0.f);
      if (row_ok && pp < G1) {
        const int st = pp / Tc, j = pp - st * Tc;
        tp[q] = Tcf[((int64_t)st * S0 + ix) * Tc + j];
      }
      if (row_ok && pp < S1) mp[q] = mr[pp];
    }
  };
  if constexpr (G1 == 512) {
    auto load_t = [&](int c) {
      const float2* Tcf = T + (int64_t)(bb * C + c) * nstr * S0 * Tc;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int pp = lane + 64 * q;
        const int st = pp / Tc, j = pp - st * Tc;
        tp[q] = row_ok ? Tcf[((int64_t)st * S0 + ix) * Tc + j] : make_float2(0.f, 0.f);
      }
    };
    load_t(0);
    __syncthreads();
    // The lane's T columns are its FFT rows and, when S1 / 2 is a multiple of 64, the image
    // columns it combines are its own output rows too (image column lane + 64 q is grid column
    // lane + 64 ((q - S1 / 128) mod 8)): passes 1 and 3 stay in registers.
    if (h1 % 64 == 0 && h1 <= 256) {
      // only T is fetched a coil ahead; the map row is fetched during the transform
      auto run = [&](auto shift) {
        constexpr int SH = decltype(shift)::value;
        for (int c = 0; c < C; ++c) {
          float2 v[8], mc[NA];
          const float2* mr = maps + ((int64_t)c * S0 + ix) * S1;
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            v[q] = tp[q];
            const int iy = lane + 64 * q;
            mc[q] = (row_ok && iy < S1) ? mr[iy] : make_float2(0.f, 0.f);
          }
          if (c + 1 < C) load_t(c + 1);
          fft512_wave_regs<true>(v, buf, stw, lane);
#pragma unroll
          for (int q = 0; q < NA; ++q) {
            const int iy = lane + 64 * q;
            if (row_ok && iy < S1) acc[q] = cadd(acc[q], cmulc(v[(q - SH + 8) & 7], mc[q]));
          }
        }
      };
      switch (h1 / 64) {
        case 0: run(std::integral_constant<int, 0>{}); break;
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        case 3: run(std::integral_constant<int, 3>{}); break;
        default: run(std::integral_constant<int, 4>{}); break;
      }
      if (row_ok) {
        const float dxi = dx[ix];
        float2* xr = x + ((int64_t)bb * S0 + ix) * S1;
#pragma unroll
        for (int q = 0; q < NA; ++q) {
          const int iy = lane + 64 * q;
          if (iy < S1) xr[iy] = cscale(acc[q], dxi * dy[iy]);
        }
      }
      return;
    }
  }
  load(0);
  __syncthreads();
  for (int c = 0; c < C; ++c) {
    float2 mc[NA];
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int pp = lane + 64 * q;
      if (pp < G1) buf[PIDX(pp)] = tp[q];
      mc[q] = mp[q];
    }
    if (c + 1 < C) load(c + 1);
    fft_lds_wave<G1, true>(buf, stw, lane);
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int iy = lane + 64 * q;
      if (row_ok && iy < S1) {
        const int pp = iy >= h1 ? iy - h1 : iy - h1 + G1;
        acc[q] = cadd(acc[q], cmulc(buf[PIDX(pp)], mc[q]));
      }
    }
    fft_sync<true>();
  }
  if (row_ok) {
    const float dxi = dx[ix];
    float2* xr = x + ((int64_t)bb * S0 + ix) * S1;
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int iy = lane + 64 * q;
      if (iy < S1) xr[iy] = cscale(acc[q], dxi * dy[iy]);
    }
  }
}

// ---------------------------------------------------------------- dispatch
template <int G>
static int launch_rows_fwd(const mrh_nufft_plan* p, const float2* x, float2* y, int nb, int like,
                           int x_has_coils, FrameMask fm, hipStream_t st) {
  dim3 grid((p->S0 + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK, nb);
  hipLaunchKernelGGL(k_fwd_rows<G>, grid, dim3(256), 0, st, x, p->d_maps, p->d_dx, p->d_dy,
                     p->d_tw1, p->d_T, y, p->d_koff, p->S0, p->S1, p->C, like, p->Tc, x_has_coils, fm);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

template <int G>
static int launch_rows_adj(const mrh_nufft_plan* p, float2* x, int nb, FrameMask fm, hipStream_t st) {
  dim3 grid((p->S0 + ROWS_PER_BLOCK - 1) / ROWS_PER_BLOCK, nb);
  hipLaunchKernelGGL(k_adj_rows<G>, grid, dim3(256), 0, st, p->d_T, p->d_maps, p->d_dx, p->d_dy,
                     p->d_tw1, x, p->S0, p->S1, p->C, p->Tc, fm);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

template <int G, int WT, int TC>
static int launch_strips_tc(const mrh_nufft_plan* p, bool adjoint, const float2* yin, float2* yout,
                            int nb, int like, FrameMask fm, hipStream_t st) {
  // grid (coil chunk, frame, strip rank): the dispatch order runs every frame's heaviest strips first
  dim3 grid((p->C + p->cchunk - 1) / p->cchunk, nb, p->nstrips), block(64 * TC);
  if (adjoint)
    hipLaunchKernelGGL((k_adj_strips<G, WT, TC>), grid, block, 0, st, yin, p->d_T, p->d_em, p->d_ef,
                       p->d_soff, p->d_srank, p->d_koff, p->d_tw0, p->kb, p->S0, p->G1, p->C, like, p->W,
                       p->cchunk, fm);
  else
    hipLaunchKernelGGL((k_fwd_strips<G, WT, TC>), grid, block, 0, st, p->d_T, yout, p->d_em, p->d_ef,
                       p->d_soff, p->d_srank, p->d_koff, p->d_tw0, p->kb, p->S0, p->G1, p->C, like, p->W,
                       p->cchunk, fm);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

// Strip widths built: 8 for every grid and width, 16 for G0 <= 512 (two 16-column
// strips need more LDS than a CU has beyond that), 4 for the W = 4 kernels (a
// footprint must span at most two strips).  plan_create admits only these.
template <int G, int WT>
static int launch_strips(const mrh_nufft_plan* p, bool adjoint, const float2* yin, float2* yout,
                         int nb, int like, FrameMask fm, hipStream_t st) {
  if constexpr (G <= 512)
    if (p->Tc == 16) return launch_strips_tc<G, WT, 16>(p, adjoint, yin, yout, nb, like, fm, st);
  if constexpr (WT == 4)
    if (p->Tc == 4) return launch_strips_tc<G, WT, 4>(p, adjoint, yin, yout, nb, like, fm, st);
  if (p->Tc != 8) {
    mrh_set_error("no %d-column strip kernels for G0=%d, W=%d", p->Tc, G, p->W);
    return MRH_EUNSUP;
  }
  return launch_strips_tc<G, WT, 8>(p, adjoint, yin, yout, nb, like, fm, st);
}

#define MRH_FOR_EACH_G(X) X(32) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024)

static int dispatch_rows_fwd(const mrh_nufft_plan* p, const float2* x, float2* y, int nb, int like,
                             int xc, FrameMask fm, hipStream_t st) {
  switch (p->G1) {
#define X(G) case G: return launch_rows_fwd<G>(p, x, y, nb, like, xc, fm, st);
    MRH_FOR_EACH_G(X)
#undef X
  }
  mrh_set_error("unsupported G1=%d", p->G1);
  return MRH_EUNSUP;
}

static int dispatch_rows_adj(const mrh_nufft_plan* p, float2* x, int nb, FrameMask fm, hipStream_t st) {
  switch (p->G1) {
#define X(G) case G: return launch_rows_adj<G>(p, x, nb, fm, st);
    MRH_FOR_EACH_G(X)
#undef X
  }
  mrh_set_error("unsupported G1=%d", p->G1);
  return MRH_EUNSUP;
}

template <int WT>
static int dispatch_strips_w(const mrh_nufft_plan* p, bool adj, const float2* yin, float2* yout,
                             int nb, int like, FrameMask fm, hipStream_t st) {
  switch (p->G0) {
#define X(G) case G: return launch_strips<G, WT>(p, adj, yin, yout, nb, like, fm, st);
    MRH_FOR_EACH_G(X)
#undef X
  }
  mrh_set_error("unsupported G0=%d", p->G0);
  return MRH_EUNSUP;
}

static int dispatch_strips(const mrh_nufft_plan* p, bool adj, const float2* yin, float2* yout,
                           int nb, int like, FrameMask fm, hipStream_t st) {
  switch (p->W) {
    case 4:
      if (p->np <= kb_terms(4)) return dispatch_strips_w<4>(p, adj, yin, yout, nb, like, fm, st);
      break;
    case 6:
      if (p->np <= kb_terms(6)) return dispatch_strips_w<6>(p, adj, yin, yout, nb, like, fm, st);
      break;
    case 8: return dispatch_strips_w<8>(p, adj, yin, yout, nb, like, fm, st);
    default: break;
  }
  return dispatch_strips_w<0>(p, adj, yin, yout, nb, like, fm, st);   // any W <= 8, full series
}

extern "C" int mrh_nufft_forward(mrh_nufft_plan* p, const void* x, void* y, int like,
                                 int x_has_coils, void* stream) {
  MRH_REQUIRE(p && x && y, MRH_EINVAL, "nufft_forward: null argument");
  MRH_REQUIRE(like >= 1 && like <= p->like_max, MRH_EINVAL, "like=%d outside [1, %d]", like, p->like_max);
  hipStream_t st = (hipStream_t)stream;
  const int nb = p->B * like;
  const FrameMask fm{nullptr, 1, like};
  int rc = dispatch_rows_fwd(p, (const float2*)x, (float2*)y, nb, like, x_has_coils ? 1 : 0, fm, st);
  if (rc) return rc;
  return dispatch_strips(p, false, nullptr, (float2*)y, nb, like, fm, st);
}

extern "C" int mrh_nufft_adjoint(mrh_nufft_plan* p, const void* y, void* x, int like, void* stream) {
  MRH_REQUIRE(p && x && y, MRH_EINVAL, "nufft_adjoint: null argument");
  MRH_REQUIRE(like >= 1 && like <= p->like_max, MRH_EINVAL, "like=%d outside [1, %d]", like, p->like_max);
  hipStream_t st = (hipStream_t)stream;
  const int nb = p->B * like;
  const FrameMask fm{nullptr, 1, like};
  int rc = dispatch_strips(p, true, (const float2*)y, nullptr, nb, like, fm, st);
  if (rc) return rc;
  return dispatch_rows_adj(p, (float2*)x, nb, fm, st);
}

extern "C" int mrh_nufft_forward_masked(mrh_nufft_plan* p, const void* x, void* y, int like, int x_has_coils,
                                        const int* problem_active, int frames_per_problem, void* stream) {
  MRH_REQUIRE(p && x && y && problem_active, MRH_EINVAL, "nufft_forward_masked: null argument");
  MRH_REQUIRE(like >= 1 && like <= p->like_max, MRH_EINVAL, "like=%d outside [1, %d]", like, p->like_max);
  MRH_REQUIRE(frames_per_problem >= 1 && p->B % frames_per_problem == 0, MRH_EINVAL,
              "frames_per_problem=%d does not divide the %d frames", frames_per_problem, p->B);
  hipStream_t st = (hipStream_t)stream;
  const int nb = p->B * like;
  const FrameMask fm{problem_active, frames_per_problem, like};
  int rc = dispatch_rows_fwd(p, (const float2*)x, (float2*)y, nb, like, x_has_coils ? 1 : 0, fm, st);
  if (rc) return rc;
  return dispatch_strips(p, false, nullptr, (float2*)y, nb, like, fm, st);
}

extern "C" int mrh_nufft_adjoint_masked(mrh_nufft_plan* p, const void* y, void* x, int like,
                                        const int* problem_active, int frames_per_problem, void* stream) {
  MRH_REQUIRE(p && x && y && problem_active, MRH_EINVAL, "nufft_adjoint_masked: null argument");
  MRH_REQUIRE(like >= 1 && like <= p->like_max, MRH_EINVAL, "like=%d outside [1, %d]", like, p->like_max);
  MRH_REQUIRE(frames_per_problem >= 1 && p->B % frames_per_problem == 0, MRH_EINVAL,
              "frames_per_problem=%d does not divide the %d frames", frames_per_problem, p->B);
  hipStream_t st = (hipStream_t)stream;
  const int nb = p->B * like;
  const FrameMask fm{problem_active, frames_per_problem, like};
  int rc = dispatch_strips(p, true, (const float2*)y, nullptr, nb, like, fm, st);
  if (rc) return rc;
  return dispatch_rows_adj(p, (float2*)x, nb, fm, st);
}

extern "C" int mrh_nufft_forward_stage(mrh_nufft_plan* p, int stage, const void* x, void* y, int like,
                                       int x_has_coils, void* stream) {
  MRH_REQUIRE(p && x && y, MRH_EINVAL, "nufft_forward_stage: null argument");
  MRH_REQUIRE(like >= 1 && like <= p->like_max, MRH_EINVAL, "like=%d outside [1, %d]", like, p->like_max);
  MRH_REQUIRE(stage == 0 || stage == 1, MRH_EINVAL, "stage must be 0 (rows) or 1 (strips)");
  hipStream_t st = (hipStream_t)stream;
  const int nb = p->B * like;
  const FrameMask fm{nullptr, 1, like};
  if (stage == 0) return dispatch_rows_fwd(p, (const float2*)x, (float2*)y, nb, like, x_has_coils ? 1 : 0, fm, st);
  return dispatch_strips(p, false, nullptr, (float2*)y, nb, like, fm, st);
}

extern "C" int mrh_nufft_adjoint_stage(mrh_nufft_plan* p, int stage, const void* y, void* x, int like,
                                       void* stream) {
  MRH_REQUIRE(p && x && y, MRH_EINVAL, "nufft_adjoint_stage: null argument");
  MRH_REQUIRE(like >= 1 && like <= p->like_max, MRH_EINVAL, "like=%d outside [1, %d]", like, p->like_max);
  MRH_REQUIRE(stage == 0 || stage == 1, MRH_EINVAL, "stage must be 0 (strips) or 1 (rows)");
  hipStream_t st = (hipStream_t)stream;
  const int nb = p->B * like;
  const FrameMask fm{nullptr, 1, like};
  if (stage == 0) return dispatch_strips(p, true, (const float2*)y, nullptr, nb, like, fm, st);
  return dispatch_rows_adj(p, (float2*)x, nb, fm, st);
}

// ================================================================ 3-D NUFFT
// Volumes (S0, S1, S2), trajectories (3, K_b) per frame.  Forward: every coil's
// x * map * deapodisation placed centred in its zero grid (k3_pad writes every grid
// point), one LDS FFT pass per axis (k3_fft_lines: one wavefront per grid line), KB
// interpolation of W^3 grid points per sample (k3_interp).  Adjoint: the spreading
// as its transpose, a gather per grid point over the frame's footprint entries sorted by
// grid point at plan time (k3_gather: no atomics, deterministic, writes every grid point),
// inverse FFT passes, crop, conj(map) combine and deapodisation (k3_crop).  Same conventions as the 2-D plan: DC at grid
// index 0, image index i at (i - S/2) mod G, footprint nodes m0 = floor(u - W/2) + 1 ...

struct Samp3 {
  int m[3];      // first footprint node per axis, reduced mod G
  float u[3];    // distance from the sample to that node (grid units)
  float wA, wAH;
};

struct mrh_nufft3_plan {
  int S[3], G[3], W, C, B, np;
  double beta;
  int64_t Ktot;
  KBPoly kb;
  int64_t* h_koff;     // host copy (launches loop over frames)
  Samp3* d_samp;
  float* d_deap[3];
  float2* d_tw[3];
  float2* d_maps;
  float2* d_grid;      // C G0 G1 G2
  // adjoint spreading as a gather: per frame, the footprint entries of every grid point
  int* d_goff;         // B (G0 G1 G2 + 1): entry range of grid point g of frame b (absolute)
  int2* d_gkw;         // (sample index within the frame, KB weight x wAH as float bits)
  float2* d_yt;        // one frame's k-space sample-major (K_b, C), read by the gather
  // grid points with more than GATHER_PART entries: their lists in parts of at most that many
  // (summed into d_part, then added in part order), so no thread walks a long list
  int4* d_heavy;       // per heavy point: (g, first part, end part, 0), all frames
  int2* d_parts;       // per part: entry range [begin, end)
  float2* d_part;      // partial sums (max parts of a frame) x C
  std::vector<int64_t> h_hoff, h_poff;   // per frame: heavy-point / part offsets (B + 1)
};

static int free_plan3(mrh_nufft3_plan* p) {
  if (!p) return MRH_OK;
  void* bufs[] = {p->d_samp, p->d_deap[0], p->d_deap[1], p->d_deap[2], p->d_tw[0], p->d_tw[1], p->d_tw[2],
                  p->d_maps, p->d_grid, p->d_goff, p->d_gkw, p->d_yt, p->d_heavy, p->d_parts,
                  p->d_part};
  for (void* b : bufs) if (b) (void)hipFree(b);
  delete[] p->h_koff;
  delete p;
  return MRH_OK;
}

// The grid points that hold image samples along an axis: image index r sits at grid index
// (r - S/2) mod G, i.e. [0, S - S/2) and [G - S/2, G).  Rows outside are the zero padding.
struct AxisRows {
  int e, b;   // grid index g is a row iff g < e || g >= b (e = G: every row)
  __host__ __device__ bool has(int g) const { return g < e || g >= b; }
  __host__ __device__ int nth(int i) const { return i < e ? i : b + (i - e); }   // i-th row in order
};
static AxisRows image_rows(int S, int G) { return AxisRows{S - S / 2, G - S / 2}; }
static AxisRows all_rows(int G) { return AxisRows{G, G}; }

// grid[c][g] = x[r] map_c[r] deap(r) at the grid points g = (r - S/2) mod G of the image samples r;
// the zero padding is not written (the forward FFT passes never read it).
__global__ __launch_bounds__(256) void k3_pad(const float2* __restrict__ x, const float2* __restrict__ maps,
                                              const float* __restrict__ d0, const float* __restrict__ d1,
                                              const float* __restrict__ d2, float2* __restrict__ grid, int S0,
                                              int S1, int S2, int G0, int G1, int G2, int C) {
  const int64_t svol = (int64_t)S0 * S1 * S2, n = (int64_t)C * svol;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int c = (int)(t / svol);
    const int64_t r = t - (int64_t)c * svol;
    int64_t q = r;
    const int r2 = (int)(q % S2); q /= S2;
    const int r1 = (int)(q % S1);
    const int r0 = (int)(q / S1);
    const int g0 = (r0 - S0 / 2 + G0) % G0, g1 = (r1 - S1 / 2 + G1) % G1, g2 = (r2 - S2 / 2 + G2) % G2;
    const float2 xv = x[r], m = maps[t];
    grid[(int64_t)c * G0 * G1 * G2 + ((int64_t)g0 * G1 + g1) * G2 + g2] = cscale(cmul(xv, m), d0[r0] * d1[r1] * d2[r2]);
  }
}

// Which lines / tiles of a pass are transformed: the volume is (n_outer, N, n_inner); an outer index
// o = c Ga + ga (tiles) or a line index o = (c Gb + gb) Ga + ga (lines along the last axis) is taken
// for the rows ra x rb of (ga, gb) only (nA = ra's count; nB = rb's count, nB = Gb = 1: no second
// filter), enumerated compactly.  in / out: the rows along the transformed axis that are read (the
// others are zero) and written.
struct Pass3 {
  AxisRows ra, rb, in, out;
  int Ga, Gb, nA, nB;
  __device__ int64_t index(int64_t k) const {   // k-th taken outer (or line) index
    const int64_t ia = k % nA, q = k / nA, ib = q % nB, c = q / nB;
    return (c * Gb + rb.nth((int)ib)) * Ga + ra.nth((int)ia);
  }
};

// In-place length-N DFT of lines along the last axis (n_inner = 1): one wavefront per line (no
// block barriers), elements staged in a padded LDS buffer.
template <int N, bool INV>
__global__ __launch_bounds__(256) void k3_fft_lines(float2* __restrict__ v, const float2* __restrict__ tw0,
                                                    Pass3 ps, int64_t n_lines) {
  __shared__ float2 buf[4][fft_padded(N)];
  __shared__ float2 tw[N];
  for (int i = threadIdx.x; i < N; i += 256) tw[i] = tw0[i];
  __syncthreads();
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float2* b = buf[wv];
  for (int64_t k = (int64_t)blockIdx.x * 4 + wv; k < n_lines; k += (int64_t)gridDim.x * 4) {
    float2* base = v + ps.index(k) * N;
    for (int i = lane; i < N; i += 64) b[PIDX(i)] = ps.in.has(i) ? base[i] : make_float2(0.f, 0.f);
    fft_sync<true>();
    fft_lds_wave<N, INV>(b, tw, lane);
    fft_sync<true>();
    for (int i = lane; i < N; i += 64)
      if (ps.out.has(i)) base[i] = b[PIDX(i)];
    fft_sync<true>();
  }
}

// The same transform for a middle axis whose inner extent is a multiple of TI: a workgroup stages
// an N x TI tile of TI neighbouring lines (rows of TI contiguous elements, so the global reads and
// writes are whole 8 TI-byte segments) in LDS, line-major with an odd line stride (conflict-free
// staging), and its wavefronts transform the tile's lines in place.
template <int N, bool INV, int TI>
__global__ __launch_bounds__(256) void k3_fft_tiles(float2* __restrict__ v, const float2* __restrict__ tw0,
                                                    Pass3 ps, int64_t n_inner, int64_t n_tiles) {
  constexpr int LS = fft_padded(N) + 1;
  __shared__ float2 tile[TI * LS];
  __shared__ float2 tw[N];
  for (int i = threadIdx.x; i < N; i += 256) tw[i] = tw0[i];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t per_outer = n_inner / TI;
  for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int64_t ko = t / per_outer, inner0 = (t - ko * per_outer) * TI;
    float2* base = v + ps.index(ko) * N * n_inner + inner0;
    __syncthreads();                       // twiddles staged / the previous tile written back
    for (int e = threadIdx.x; e < N * TI; e += 256) {
      const int i = e / TI, j = e - i * TI;
      tile[j * LS + PIDX(i)] = ps.in.has(i) ? base[(int64_t)i * n_inner + j] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int j = wv; j < TI; j += 4) fft_lds_wave<N, INV>(tile + j * LS, tw, lane);
    __syncthreads();
    for (int e = threadIdx.x; e < N * TI; e += 256) {
      const int i = e / TI, j = e - i * TI;
      if (ps.out.has(i)) base[(int64_t)i * n_inner + j] = tile[j * LS + PIDX(i)];
    }
  }
}

// y[c][k] = wA sum over the W^3 footprint of w_x w_y w_z grid[c][...], one thread per (coil, sample).
__global__ __launch_bounds__(256) void k3_interp(const float2* __restrict__ grid, const Samp3* __restrict__ sp,
                                                 float2* __restrict__ y, KBPoly kb, int W, int G0, int G1, int G2,
                                                 int C, int K) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)C * K) return;
  const int c = (int)(t / K), k = (int)(t - (int64_t)c * K);
  const Samp3 s = sp[k];
  float w0[MRH_WMAX], w1[MRH_WMAX], w2[MRH_WMAX];
#pragma unroll
  for (int i = 0; i < MRH_WMAX; ++i) {
    w0[i] = i < W ? kb_weight<MRH_KB_MAXNP>(kb, s.u[0] - (float)i) : 0.f;
    w1[i] = i < W ? kb_weight<MRH_KB_MAXNP>(kb, s.u[1] - (float)i) : 0.f;
    w2[i] = i < W ? kb_weight<MRH_KB_MAXNP>(kb, s.u[2] - (float)i) : 0.f;
  }
  const float2* gc = grid + (int64_t)c * G0 * G1 * G2;
  float2 acc = make_float2(0.f, 0.f);
  // A footprint row's W points are contiguous along axis 2 unless the row wraps round the grid.
  // With G2 even (rows start 16-byte aligned) such a run is read as 16-byte pairs from the even
  // index at or below its first point: ceil((W + sh) / 2) loads instead of W.  The last pair ends
  // at most at index m2 + W, which is inside the row (m2 + W is then odd and at most G2 - 1).
  const int m2 = s.m[2], a2 = m2 & ~1, sh = m2 - a2;
  const bool run = (G2 & 1) == 0 && m2 + W <= G2;
  for (int i = 0; i < W; ++i) {
    int i0 = s.m[0] + i; if (i0 >= G0) i0 -= G0;
    for (int j = 0; j < W; ++j) {
      int i1 = s.m[1] + j; if (i1 >= G1) i1 -= G1;
      const float2* row = gc + ((int64_t)i0 * G1 + i1) * G2;
      const float wij = w0[i] * w1[j];
      float2 r = make_float2(0.f, 0.f);
      if (run) {
        float2 gv[MRH_WMAX + 2];
        const float4* rp = reinterpret_cast<const float4*>(row + a2);
#pragma unroll
        for (int q = 0; q < (MRH_WMAX + 2) / 2; ++q) {
          if (2 * q < W + sh) {
            const float4 v = rp[q];
            gv[2 * q] = make_float2(v.x, v.y);
            gv[2 * q + 1] = make_float2(v.z, v.w);
          }
        }
#pragma unroll
        for (int l = 0; l < MRH_WMAX; ++l) {
          if (l < W) {
            const float2 g = sh ? gv[l + 1] : gv[l];
            r.x = fmaf(g.x, w2[l], r.x);
            r.y = fmaf(g.y, w2[l], r.y);
          }
        }
      } else {
        for (int l = 0; l < W; ++l) {
          int i2 = m2 + l; if (i2 >= G2) i2 -= G2;
          const float2 gv = row[i2];
          r.x = fmaf(gv.x, w2[l], r.x);
          r.y = fmaf(gv.y, w2[l], r.y);
        }
      }
      acc.x = fmaf(r.x, wij, acc.x);
      acc.y = fmaf(r.y, wij, acc.y);
    }
  }
  y[(int64_t)c * K + k] = cscale(acc, s.wA);
}

// yt[k][c] = y[c][k]: one frame's k-space sample-major, so that the gather reads a sample's
// coils from one cache line.
__global__ __launch_bounds__(256) void k3_transpose_y(const float2* __restrict__ y, float2* __restrict__ yt, int C,
                                                      int K) {
  const int64_t n = (int64_t)C * K;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int k = (int)(t / C), c = (int)(t - (int64_t)k * C);
    yt[t] = y[(int64_t)c * K + k];
  }
}

#define GATHER_PART 64   // longest entry list one gather thread walks

// out[q * ostride] = sum over entries [e0, e1) of w_e yt[k_e][c0 + q] for the coils of one group,
// (k_e, w_e) = gkw[e] read as one 8-byte pair.
__device__ __forceinline__ void gather_range(const float2* __restrict__ yt, const int2* __restrict__ gkw, int e0,
                                             int e1, int c0, int C, float2* __restrict__ out, int64_t ostride) {
  float2 acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = make_float2(0.f, 0.f);
  // a full group of 8 coils with C even starts 16 bytes aligned: read as four 16-byte pairs
  const bool pairs = (C & 1) == 0 && c0 + 8 <= C;
  for (int e = e0; e < e1; ++e) {
    const int2 kw = gkw[e];
    const float2* ys = yt + (int64_t)kw.x * C + c0;
    const float w = __int_as_float(kw.y);
    float2 v[8];
    if (pairs) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 u = reinterpret_cast<const float4*>(ys)[q];
        v[2 * q] = make_float2(u.x, u.y);
        v[2 * q + 1] = make_float2(u.z, u.w);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = c0 + q < C ? ys[q] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (c0 + q < C) {
        acc[q].x = fmaf(w, v[q].x, acc[q].x);
        acc[q].y = fmaf(w, v[q].y, acc[q].y);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (c0 + q < C) out[q * ostride] = acc[q];
}

// grid[c][g] = sum over the entries e of grid point g of w_e yt[k_e][c]; coils in groups of 8
// held in registers, so each entry is read once per group.  Points with more than GATHER_PART
// entries are left to k3_gather_parts / k3_gather_heavy.
__global__ __launch_bounds__(256) void k3_gather(const float2* __restrict__ yt, const int* __restrict__ goff,
                                                 const int2* __restrict__ gkw, float2* __restrict__ grid,
                                                 int64_t gvol, int C) {
  for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < gvol; g += (int64_t)gridDim.x * 256) {
    const int e0 = goff[g], e1 = goff[g + 1];
    if (e1 - e0 > GATHER_PART) continue;
    for (int c0 = 0; c0 < C; c0 += 8) gather_range(yt, gkw, e0, e1, c0, C, grid + (int64_t)c0 * gvol + g, gvol);
  }
}

// One part of a long list: part[i][c] = its entries' sum (parts of a frame, in list order).
__global__ __launch_bounds__(256) void k3_gather_parts(const float2* __restrict__ yt, const int2* __restrict__ parts,
                                                       const int2* __restrict__ gkw, float2* __restrict__ part,
                                                       int n_parts, int C) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_parts) return;
  const int2 r = parts[i];
  for (int c0 = 0; c0 < C; c0 += 8) gather_range(yt, gkw, r.x, r.y, c0, C, part + (int64_t)i * C + c0, 1);
}

// A long list's grid point: the sum of its parts in order (deterministic).
__global__ __launch_bounds__(256) void k3_gather_heavy(const int4* __restrict__ heavy, const float2* __restrict__ part,
                                                       float2* __restrict__ grid, int64_t gvol, int n_heavy, int p_base,
                                                       int C) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_heavy) return;
  const int4 h = heavy[i];
  for (int c = 0; c < C; ++c) {
    float2 acc = make_float2(0.f, 0.f);
    for (int q = h.y; q < h.z; ++q) acc = cadd(acc, part[(int64_t)(q - p_base) * C + c]);
    grid[(int64_t)c * gvol + h.x] = acc;
  }
}

// x[r] = deap(r) sum_c conj(map_c[r]) grid[c][(r - S/2) mod G].
__global__ __launch_bounds__(256) void k3_crop(const float2* __restrict__ grid, const float2* __restrict__ maps,
                                               const float* __restrict__ d0, const float* __restrict__ d1,
                                               const float* __restrict__ d2, float2* __restrict__ x, int S0,
                                               int S1, int S2, int G0, int G1, int G2, int C) {
  const int64_t svol = (int64_t)S0 * S1 * S2, gvol = (int64_t)G0 * G1 * G2;
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < svol; r += (int64_t)gridDim.x * 256) {
    int64_t q = r;
    const int r2 = (int)(q % S2); q /= S2;
    const int r1 = (int)(q % S1);
    const int r0 = (int)(q / S1);
    const int g0 = (r0 - S0 / 2 + G0) % G0, g1 = (r1 - S1 / 2 + G1) % G1, g2 = (r2 - S2 / 2 + G2) % G2;
    const int64_t g = ((int64_t)g0 * G1 + g1) * G2 + g2;
    float2 acc = make_float2(0.f, 0.f);
    for (int c = 0; c < C; ++c) acc = cadd(acc, cmulc(grid[(int64_t)c * gvol + g], maps[(int64_t)c * svol + r]));
    x[r] = cscale(acc, d0[r0] * d1[r1] * d2[r2]);
  }
}

template <bool INV>
static int launch_fft_lines3(int N, float2* v, const float2* tw, const Pass3& ps, int64_t n_lines, hipStream_t st) {
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(65535, (n_lines + 3) / 4));
#define MRH_FFT3_CASE(n)                                                                                   \
  case n: hipLaunchKernelGGL((k3_fft_lines<n, INV>), dim3(blocks), dim3(256), 0, st, v, tw, ps, n_lines); break;
  switch (N) {
    MRH_FFT3_CASE(32) MRH_FFT3_CASE(64) MRH_FFT3_CASE(96) MRH_FFT3_CASE(128) MRH_FFT3_CASE(192)
    MRH_FFT3_CASE(256) MRH_FFT3_CASE(384) MRH_FFT3_CASE(512) MRH_FFT3_CASE(768) MRH_FFT3_CASE(1024)
    default: mrh_set_error("3-D grid size %d not built", N); return MRH_EUNSUP;
  }
#undef MRH_FFT3_CASE
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

// Passes along axes 0 and 1 run on 16-line tiles (every built grid size is a multiple of 16).
template <bool INV>
static int launch_fft_tiles3(int N, float2* v, const float2* tw, const Pass3& ps, int64_t n_inner, int64_t n_outer,
                             hipStream_t st) {
  constexpr int TI = 16;
  MRH_REQUIRE(n_inner % TI == 0, MRH_EUNSUP, "3-D FFT pass: inner extent %lld is not a multiple of %d",
              (long long)n_inner, TI);
  const int64_t n_tiles = n_outer * (n_inner / TI);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(65535, n_tiles));
#define MRH_FFT3_CASE(n)                                                                                      \
  case n:                                                                                                    \
    hipLaunchKernelGGL((k3_fft_tiles<n, INV, TI>), dim3(blocks), dim3(256), 0, st, v, tw, ps, n_inner, n_tiles); \
    break;
  switch (N) {
    MRH_FFT3_CASE(32) MRH_FFT3_CASE(64) MRH_FFT3_CASE(96) MRH_FFT3_CASE(128) MRH_FFT3_CASE(192)
    MRH_FFT3_CASE(256) MRH_FFT3_CASE(384) MRH_FFT3_CASE(512) MRH_FFT3_CASE(768) MRH_FFT3_CASE(1024)
    default: mrh_set_error("3-D grid size %d not built", N); return MRH_EUNSUP;
  }
#undef MRH_FFT3_CASE
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

static int rows_count(const AxisRows& r, int G) { return r.e + (G - r.b); }

// FFT of every coil's grid along all three axes (grid layout (C, G0, G1, G2), G2 fastest), pruned to
// the rows that hold data.  Forward (k3_pad wrote only the image points): axis 2 over the lines whose
// (g0, g1) are image rows, reading their image rows; axis 1 over the planes whose g0 is an image row,
// reading image rows; axis 0 over everything, reading image rows; the zero padding is never read.
// Inverse (the gather wrote every point), axes in the opposite order: axis 0 writes only the image
// rows, axis 1 transforms those planes and writes image rows, axis 2 transforms the lines (g0, g1)
// of image rows and writes the image rows that the crop reads.
template <bool INV>
static int fft3(const mrh_nufft3_plan* p, hipStream_t st) {
  const int G0 = p->G[0], G1 = p->G[1], G2 = p->G[2], C = p->C;
  const AxisRows r0 = image_rows(p->S[0], G0), r1 = image_rows(p->S[1], G1), r2 = image_rows(p->S[2], G2);
  const AxisRows one = all_rows(1);
  const int n0 = rows_count(r0, G0), n1 = rows_count(r1, G1);
  // axis 2: lines (c, g0 in r0, g1 in r1)
  const Pass3 pa{r1, r0, INV ? all_rows(G2) : r2, INV ? r2 : all_rows(G2), G1, G0, n1, n0};
  // axis 1: planes (c, g0 in r0), inner extent G2
  const Pass3 pb{r0, one, INV ? all_rows(G1) : r1, INV ? r1 : all_rows(G1), G0, 1, n0, 1};
  // axis 0: every coil, inner extent G1 G2
  const Pass3 pc{one, one, INV ? all_rows(G0) : r0, INV ? r0 : all_rows(G0), 1, 1, 1, 1};
  int rc;
  if (!INV) {
    rc = launch_fft_lines3<INV>(G2, p->d_grid, p->d_tw[2], pa, (int64_t)C * n0 * n1, st);
    if (!rc) rc = launch_fft_tiles3<INV>(G1, p->d_grid, p->d_tw[1], pb, G2, (int64_t)C * n0, st);
    if (!rc) rc = launch_fft_tiles3<INV>(G0, p->d_grid, p->d_tw[0], pc, (int64_t)G1 * G2, C, st);
  } else {
    rc = launch_fft_tiles3<INV>(G0, p->d_grid, p->d_tw[0], pc, (int64_t)G1 * G2, C, st);
    if (!rc) rc = launch_fft_tiles3<INV>(G1, p->d_grid, p->d_tw[1], pb, G2, (int64_t)C * n0, st);
    if (!rc) rc = launch_fft_lines3<INV>(G2, p->d_grid, p->d_tw[2], pa, (int64_t)C * n0 * n1, st);
  }
  return rc;
}

static int grid_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256)); }

extern "C" int mrh_nufft3_plan_create(const mrh_nufft3_desc* d, mrh_nufft3_plan** out) {
  MRH_REQUIRE(d && out, MRH_EINVAL, "nufft3_plan_create: null argument");
  *out = nullptr;
  for (int a = 0; a < 3; ++a) {
    const int S = a == 0 ? d->S0 : a == 1 ? d->S1 : d->S2, G = a == 0 ? d->G0 : a == 1 ? d->G1 : d->G2;
    MRH_REQUIRE(S >= 2 && G >= S, MRH_EINVAL, "axis %d: image size %d, grid %d", a, S, G);
    MRH_REQUIRE(fft_size_ok(G), MRH_EUNSUP, "axis %d: grid size %d not built (32 ... 1024, 2^a 3^b)", a, G);
  }
  MRH_REQUIRE(d->W >= 2 && d->W <= MRH_WMAX, MRH_EUNSUP, "KB width %d outside [2, %d]", d->W, MRH_WMAX);
  MRH_REQUIRE(d->C >= 1 && d->B >= 1, MRH_EINVAL, "C, B must be >= 1");
  MRH_REQUIRE(d->k_offsets && d->traj && d->maps, MRH_EINVAL, "k_offsets, traj (host) and maps (device) are required");
  mrh_nufft3_plan* p = new mrh_nufft3_plan();   // value-initialised: pointers null, vectors empty
  p->S[0] = d->S0; p->S[1] = d->S1; p->S[2] = d->S2;
  p->G[0] = d->G0; p->G[1] = d->G1; p->G[2] = d->G2;
  p->W = d->W; p->C = d->C; p->B = d->B;
  // KB series (as the 2-D plan: beta from the first axis' oversampling unless given)
  const double osf = (double)d->G0 / (double)d->S0;
  const double beta = d->beta > 0 ? d->beta : M_PI * std::sqrt(std::pow((d->W / osf) * (osf - 0.5), 2) - 0.8);
  p->beta = beta;
  const double i0b = bessel_i0(beta), q = beta * beta / 4.0;
  double coef[64];
  {
    double fact = 1.0;
    for (int n = 0; n < 64; ++n) {
      if (n > 0) fact *= n;
      coef[n] = 1.0 / (fact * fact * i0b);
    }
  }
  int np = 1;
  while (np < MRH_KB_MAXNP && coef[np] * std::pow(q, np) > 1e-10) ++np;
  if (!(coef[np] * std::pow(q, np) < 1e-9)) { free_plan3(p); mrh_set_error("KB beta %.3f too large for the series", beta); return MRH_EUNSUP; }
  p->np = np;
  p->kb.qscale = (float)q;
  p->kb.two_over_w = (float)(2.0 / d->W);
  for (int n = 0; n < MRH_KB_MAXNP; ++n) p->kb.a[n] = n < np ? (float)coef[n] : 0.f;
  // samples
  const int B = d->B;
  p->h_koff = new int64_t[B + 1];
  for (int b = 0; b <= B; ++b) p->h_koff[b] = d->k_offsets[b];
  for (int b = 0; b < B; ++b)
    if (p->h_koff[b + 1] < p->h_koff[b]) { free_plan3(p); mrh_set_error("k_offsets must be non-decreasing"); return MRH_EINVAL; }
  p->Ktot = p->h_koff[B];
  MRH_REQUIRE(p->Ktot * (int64_t)d->C < ((int64_t)1 << 31), MRH_EUNSUP, "too many coil-samples for one plan");
  std::vector<Samp3> samp((size_t)std::max<int64_t>(1, p->Ktot));
  for (int b = 0; b < B; ++b) {
    const int64_t k0 = p->h_koff[b], Kb = p->h_koff[b + 1] - k0;
    const float* tb = d->traj + 3 * k0;
    for (int64_t k = 0; k < Kb; ++k) {
      Samp3 s;
      for (int a = 0; a < 3; ++a) {
        const double u = (double)tb[a * Kb + k] * p->G[a] / (2.0 * M_PI);
        const double f = std::floor(u - 0.5 * d->W) + 1.0;
        s.m[a] = (int)((((int64_t)f % p->G[a]) + p->G[a]) % p->G[a]);
        s.u[a] = (float)(u - f);
      }
      const float dval = d->dcf ? d->dcf[k0 + k] : 1.f;
      s.wA = 1.f; s.wAH = 1.f;
      if (d->dcf && d->dcf_mode == MRH_DCF_BALANCED) {
        s.wA = s.wAH = (float)std::sqrt(std::max((double)dval, 0.0));
      } else if (d->dcf && d->dcf_mode == MRH_DCF_STANDARD) {
        s.wA = d->dcf_in_fwd ? dval : 1.f;
        s.wAH = d->dcf_in_adj ? dval : 1.f;
      }
      samp[(size_t)(k0 + k)] = s;
    }
  }
  int rc = upload(&p->d_samp, samp.data(), samp.size());
  // gather lists: every footprint entry (grid point, sample, weight x wAH), counting-sorted by
  // grid point within each frame; the weights are the KB series in FP64
  {
    const int W = d->W;
    const int64_t gvol = (int64_t)d->G0 * d->G1 * d->G2;
    const int64_t n_ent = p->Ktot * W * W * W;
    if (n_ent >= ((int64_t)1 << 31) || (int64_t)B * (gvol + 1) >= ((int64_t)1 << 31)) {
      free_plan3(p);
      mrh_set_error("nufft3: %lld footprint entries / %lld grid offsets exceed 32-bit indexing",
                    (long long)n_ent, (long long)B * (gvol + 1));
      return MRH_EUNSUP;
    }
    auto kbw = [&](double dd) {
      const double t = 2.0 * dd / W, qq = q * std::max(1.0 - t * t, 0.0);
      double v = coef[np - 1];
      for (int n = np - 2; n >= 0; --n) v = v * qq + coef[n];
      return v;
    };
    std::vector<int> goff((size_t)(B * (gvol + 1)), 0);
    std::vector<int2> gkw((size_t)std::max<int64_t>(1, n_ent));
    int64_t base = 0;
    std::vector<int> fill;
    for (int b = 0; b < B && !rc; ++b) {
      const int64_t k0 = p->h_koff[b], Kb = p->h_koff[b + 1] - k0;
      int* off = goff.data() + (size_t)b * (gvol + 1);
      auto gidx = [&](const Samp3& s, int i, int j, int l) {
        int i0 = s.m[0] + i; if (i0 >= d->G0) i0 -= d->G0;
        int i1 = s.m[1] + j; if (i1 >= d->G1) i1 -= d->G1;
        int i2 = s.m[2] + l; if (i2 >= d->G2) i2 -= d->G2;
        return ((int64_t)i0 * d->G1 + i1) * d->G2 + i2;
      };
      for (int64_t k = 0; k < Kb; ++k)
        for (int i = 0; i < W; ++i)
          for (int j = 0; j < W; ++j)
            for (int l = 0; l < W; ++l) off[gidx(samp[(size_t)(k0 + k)], i, j, l) + 1]++;
      off[0] = (int)base;
      for (int64_t g = 0; g < gvol; ++g) off[g + 1] += off[g];
      fill.assign(off, off + gvol);
      for (int64_t k = 0; k < Kb; ++k) {
        const Samp3& s = samp[(size_t)(k0 + k)];
        for (int i = 0; i < W; ++i) {
          const double wi = kbw(s.u[0] - i);
          for (int j = 0; j < W; ++j) {
            const double wij = wi * kbw(s.u[1] - j);
            for (int l = 0; l < W; ++l) {
              const int64_t g = gidx(s, i, j, l);
              const int e = fill[(size_t)g]++;
              const float w = (float)(wij * kbw(s.u[2] - l) * s.wAH);
              int wb;
              std::memcpy(&wb, &w, sizeof(wb));
              gkw[(size_t)e] = make_int2((int)k, wb);
            }
          }
        }
      }
      base = off[gvol];
    }
    // long lists, split into parts
    std::vector<int4> heavy;
    std::vector<int2> parts;
    int64_t max_parts = 1;
    p->h_hoff.assign(1, 0);
    p->h_poff.assign(1, 0);
    for (int b = 0; b < B; ++b) {
      const int* off = goff.data() + (size_t)b * (gvol + 1);
      const int64_t p_first = (int64_t)parts.size();
      for (int64_t g = 0; g < gvol; ++g) {
        const int e0 = off[g], e1 = off[g + 1];
        if (e1 - e0 <= GATHER_PART) continue;
        const int q0 = (int)parts.size();
        for (int e = e0; e < e1; e += GATHER_PART) parts.push_back(make_int2(e, std::min(e1, e + GATHER_PART)));
        heavy.push_back(make_int4((int)g, q0, (int)parts.size(), 0));
      }
      max_parts = std::max<int64_t>(max_parts, (int64_t)parts.size() - p_first);
      p->h_hoff.push_back((int64_t)heavy.size());
      p->h_poff.push_back((int64_t)parts.size());
    }
    if (!rc) rc = upload(&p->d_heavy, heavy.empty() ? nullptr : heavy.data(), std::max<size_t>(1, heavy.size()));
    if (!rc) rc = upload(&p->d_parts, parts.empty() ? nullptr : parts.data(), std::max<size_t>(1, parts.size()));
    if (!rc) rc = upload<float2>(&p->d_part, nullptr, (size_t)(max_parts * d->C));
    int64_t kmax = 1;
    for (int b = 0; b < B; ++b) kmax = std::max<int64_t>(kmax, p->h_koff[b + 1] - p->h_koff[b]);
    if (!rc) rc = upload<float2>(&p->d_yt, nullptr, (size_t)(kmax * d->C));
    if (!rc) rc = upload(&p->d_goff, goff.data(), goff.size());
    if (!rc) rc = upload(&p->d_gkw, gkw.data(), gkw.size());
  }
  for (int a = 0; a < 3 && !rc; ++a) {
    std::vector<float> dv(p->S[a]);
    for (int i = 0; i < p->S[a]; ++i) dv[i] = (float)(1.0 / kb_ft((double)(i - p->S[a] / 2) / p->G[a], d->W, beta));
    rc = upload(&p->d_deap[a], dv.data(), dv.size());
    std::vector<float2> tw;
    twiddles(tw, p->G[a]);
    if (!rc) rc = upload(&p->d_tw[a], tw.data(), tw.size());
  }
  const int64_t svol = (int64_t)d->S0 * d->S1 * d->S2, gvol = (int64_t)d->G0 * d->G1 * d->G2;
  if (!rc) rc = upload<float2>(&p->d_maps, nullptr, (size_t)(d->C * svol));
  if (!rc) rc = upload<float2>(&p->d_grid, nullptr, (size_t)(d->C * gvol));
  if (rc) { free_plan3(p); return rc; }
  hipError_t err = hipMemcpy(p->d_maps, d->maps, sizeof(float2) * d->C * svol, hipMemcpyDeviceToDevice);
  if (err != hipSuccess) { free_plan3(p); mrh_set_error("nufft3: maps copy failed: %s", hipGetErrorString(err)); return MRH_EHIP; }
  *out = p;
  return MRH_OK;
}

extern "C" int mrh_nufft3_plan_destroy(mrh_nufft3_plan* p) { return free_plan3(p); }

extern "C" int mrh_nufft3_forward(mrh_nufft3_plan* p, const void* x, void* y, void* stream) {
  MRH_REQUIRE(p && x && y, MRH_EINVAL, "nufft3_forward: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t svol = (int64_t)p->S[0] * p->S[1] * p->S[2];
  for (int b = 0; b < p->B; ++b) {
    const int64_t k0 = p->h_koff[b], Kb = p->h_koff[b + 1] - k0;
    hipLaunchKernelGGL(k3_pad, dim3(grid_blocks((int64_t)p->C * svol)), dim3(256), 0, st, (const float2*)x + b * svol, p->d_maps,
                       p->d_deap[0], p->d_deap[1], p->d_deap[2], p->d_grid, p->S[0], p->S[1], p->S[2], p->G[0],
                       p->G[1], p->G[2], p->C);
    MRH_CHECK_LAUNCH();
    int rc = fft3<false>(p, st);
    if (rc) return rc;
    if (Kb == 0) continue;
    const int64_t n = (int64_t)p->C * Kb;
    hipLaunchKernelGGL(k3_interp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p->d_grid, p->d_samp + k0,
                       (float2*)y + (int64_t)p->C * k0, p->kb, p->W, p->G[0], p->G[1], p->G[2], p->C, (int)Kb);
    MRH_CHECK_LAUNCH();
  }
  return MRH_OK;
}

extern "C" int mrh_nufft3_adjoint(mrh_nufft3_plan* p, const void* y, void* x, void* stream) {
  MRH_REQUIRE(p && x && y, MRH_EINVAL, "nufft3_adjoint: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int64_t svol = (int64_t)p->S[0] * p->S[1] * p->S[2], gtot = (int64_t)p->C * p->G[0] * p->G[1] * p->G[2];
  for (int b = 0; b < p->B; ++b) {
    const int64_t k0 = p->h_koff[b], Kb = p->h_koff[b + 1] - k0;
    const int64_t gvol = gtot / p->C;
    if (Kb > 0) {
      hipLaunchKernelGGL(k3_transpose_y, dim3(grid_blocks((int64_t)p->C * Kb)), dim3(256), 0, st,
                         (const float2*)y + (int64_t)p->C * k0, p->d_yt, p->C, (int)Kb);
      MRH_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k3_gather, dim3(grid_blocks(gvol)), dim3(256), 0, st, (const float2*)p->d_yt,
                       p->d_goff + (int64_t)b * (gvol + 1), p->d_gkw, p->d_grid, gvol, p->C);
    MRH_CHECK_LAUNCH();
    const int n_parts = (int)(p->h_poff[b + 1] - p->h_poff[b]), n_heavy = (int)(p->h_hoff[b + 1] - p->h_hoff[b]);
    if (n_heavy > 0) {
      hipLaunchKernelGGL(k3_gather_parts, dim3((n_parts + 255) / 256), dim3(256), 0, st, (const float2*)p->d_yt,
                         p->d_parts + p->h_poff[b], p->d_gkw, p->d_part, n_parts, p->C);
      MRH_CHECK_LAUNCH();
      hipLaunchKernelGGL(k3_gather_heavy, dim3((n_heavy + 255) / 256), dim3(256), 0, st, p->d_heavy + p->h_hoff[b],
                         (const float2*)p->d_part, p->d_grid, gvol, n_heavy, (int)p->h_poff[b], p->C);
      MRH_CHECK_LAUNCH();
    }
    int rc = fft3<true>(p, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k3_crop, dim3(grid_blocks(svol)), dim3(256), 0, st, p->d_grid, p->d_maps, p->d_deap[0],
                       p->d_deap[1], p->d_deap[2], (float2*)x + b * svol, p->S[0], p->S[1], p->S[2], p->G[0],
                       p->G[1], p->G[2], p->C);
    MRH_CHECK_LAUNCH();
  }
  return MRH_OK;
}
