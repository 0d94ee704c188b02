// Low-rank temporal basis: u = U V, the parameter-space map of the reference's
// TemporalBasisOp (graspcg/regularization/mapping.py:112-231).
//
//   U   (T, K)  complex64, row-major: frame t's weights on the K basis images
//   V   (K, N)  complex64: the K coefficient images, N = every non-frame element
//   u   (T, N)  complex64: the frames
//
// forward   u[t, n]  = sum_k U[t, k] V[k, n]                 (mapping.py:141-147)
// adjoint   gV[k, n] (+)= sum_t conj(U[t, k]) gu[t, n]        (mapping.py:149-164)
// diag      DV[k, n] += sum_t |U[t, k]|^2 ku[t, n]            (mapping.py:170-199)
//           with ku a field (T, N) or a per-frame profile (T)  (mapping.py:200-231)
//
// All three are HBM-bound streams over n (K + T complex values per n for the
// forward and adjoint): a thread per n (and, in the forward, per group of frames)
// keeps its accumulators in registers, reads U through uniform (scalar) loads and
// walks t in order, so every output is a fixed-order FP32 sum and repeat calls are
// bitwise equal.
#include "mrh_common.h"
#include "mrirecon_hip.h"

#define TB_THREADS 256
#define TB_FRAMES 8     // frames per thread in the forward map

namespace {

int tb_blocks(int64_t n) {
  const int64_t b = (n + TB_THREADS - 1) / TB_THREADS;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20));
}

// blockIdx.y picks frames [y * TB_FRAMES, min(T, (y + 1) * TB_FRAMES)): more threads in flight
// than one per element (V's re-reads per frame group hit L2).
template <int K>
__global__ __launch_bounds__(TB_THREADS) void k_tb_forward(const float2* __restrict__ U, int T,
                                                           const float2* __restrict__ V, int64_t N,
                                                           float2* __restrict__ u) {
  const int t0 = blockIdx.y * TB_FRAMES, t1 = min(T, t0 + TB_FRAMES);
  for (int64_t n = blockIdx.x * (int64_t)TB_THREADS + threadIdx.x; n < N; n += (int64_t)gridDim.x * TB_THREADS) {
    float2 v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = V[(int64_t)k * N + n];
    for (int t = t0; t < t1; ++t) {
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < K; ++k) acc = cadd(acc, cmul(U[t * K + k], v[k]));
      u[(int64_t)t * N + n] = acc;
    }
  }
}

template <int K>
__global__ __launch_bounds__(TB_THREADS) void k_tb_adjoint(const float2* __restrict__ U, int T,
                                                           const float2* __restrict__ gu, int64_t N,
                                                           float2* __restrict__ gV, int accumulate) {
  for (int64_t n = blockIdx.x * (int64_t)TB_THREADS + threadIdx.x; n < N; n += (int64_t)gridDim.x * TB_THREADS) {
    float2 acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = make_float2(0.f, 0.f);
    for (int t = 0; t < T; ++t) {
      const float2 g = gu[(int64_t)t * N + n];
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] = cadd(acc[k], cmulc(g, U[t * K + k]));   // conj(U) g
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      float2* o = gV + (int64_t)k * N + n;
      *o = accumulate ? cadd(*o, acc[k]) : acc[k];
    }
  }
}

// ku[t, n] read at ku[t * ld_t + n * ld_n]: a field has (ld_t, ld_n) = (N, 1), a profile (1, 0).
template <int K>
__global__ __launch_bounds__(TB_THREADS) void k_tb_diag(const float2* __restrict__ U, int T,
                                                        const float* __restrict__ ku, int64_t ld_t, int64_t ld_n,
                                                        int64_t N, float* __restrict__ DV) {
  for (int64_t n = blockIdx.x * (int64_t)TB_THREADS + threadIdx.x; n < N; n += (int64_t)gridDim.x * TB_THREADS) {
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.f;
    for (int t = 0; t < T; ++t) {
      const float w = ku[(int64_t)t * ld_t + n * ld_n];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float2 a = U[t * K + k];
        acc[k] = fmaf(a.x * a.x + a.y * a.y, w, acc[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) DV[(int64_t)k * N + n] += acc[k];
  }
}

// One instantiation per basis size 1..MRH_TBASIS_MAX_K (the accumulators stay in registers).
template <template <int> class Launch, typename... A>
int tb_dispatch(int K, A... a) {
  switch (K) {
#define TB_CASE(k) case k: return Launch<k>::run(a...);
    TB_CASE(1) TB_CASE(2) TB_CASE(3) TB_CASE(4) TB_CASE(5) TB_CASE(6) TB_CASE(7) TB_CASE(8)
    TB_CASE(9) TB_CASE(10) TB_CASE(11) TB_CASE(12) TB_CASE(13) TB_CASE(14) TB_CASE(15) TB_CASE(16)
#undef TB_CASE
    default:
      mrh_set_error("tbasis: K = %d basis images, need 1..%d", K, MRH_TBASIS_MAX_K);
      return MRH_EINVAL;
  }
}

template <int K>
struct LaunchFwd {
  static int run(const float2* U, int T, const float2* V, int64_t N, float2* u, hipStream_t st) {
    const dim3 grid(tb_blocks(N), (T + TB_FRAMES - 1) / TB_FRAMES);
    hipLaunchKernelGGL(k_tb_forward<K>, grid, dim3(TB_THREADS), 0, st, U, T, V, N, u);
    MRH_CHECK_LAUNCH();
    return MRH_OK;
  }
};

template <int K>
struct LaunchAdj {
  static int run(const float2* U, int T, const float2* gu, int64_t N, float2* gV, int acc, hipStream_t st) {
    hipLaunchKernelGGL(k_tb_adjoint<K>, dim3(tb_blocks(N)), dim3(TB_THREADS), 0, st, U, T, gu, N, gV, acc);
    MRH_CHECK_LAUNCH();
    return MRH_OK;
  }
};

template <int K>
struct LaunchDiag {
  static int run(const float2* U, int T, const float* ku, int64_t ld_t, int64_t ld_n, int64_t N, float* DV,
                 hipStream_t st) {
    hipLaunchKernelGGL(k_tb_diag<K>, dim3(tb_blocks(N)), dim3(TB_THREADS), 0, st, U, T, ku, ld_t, ld_n, N, DV);
    MRH_CHECK_LAUNCH();
    return MRH_OK;
  }
};

int tb_check(const char* what, int T, int K, int64_t N) {
  MRH_REQUIRE(T >= 1 && T <= 65535 * TB_FRAMES, MRH_EINVAL, "%s: T = %d frames, need 1..%d", what, T,
              65535 * TB_FRAMES);
  MRH_REQUIRE(K >= 1 && K <= MRH_TBASIS_MAX_K, MRH_EINVAL, "%s: K = %d basis images, need 1..%d", what, K,
              MRH_TBASIS_MAX_K);
  MRH_REQUIRE(N >= 0, MRH_EINVAL, "%s: N = %lld elements per image, need >= 0", what, (long long)N);
  return MRH_OK;
}

}  // namespace

extern "C" int mrh_tbasis_forward(const void* U, int T, int K, const void* V, int64_t N, void* u, void* stream) {
  int rc = tb_check("tbasis_forward", T, K, N);
  if (rc != MRH_OK) return rc;
  if (N == 0) return MRH_OK;
  MRH_REQUIRE(U && V && u, MRH_EINVAL, "tbasis_forward: null argument");
  return tb_dispatch<LaunchFwd>(K, (const float2*)U, T, (const float2*)V, N, (float2*)u, (hipStream_t)stream);
}

extern "C" int mrh_tbasis_adjoint(const void* U, int T, int K, const void* gu, int64_t N, void* gV, int accumulate,
                                  void* stream) {
  int rc = tb_check("tbasis_adjoint", T, K, N);
  if (rc != MRH_OK) return rc;
  if (N == 0) return MRH_OK;
  MRH_REQUIRE(U && gu && gV, MRH_EINVAL, "tbasis_adjoint: null argument");
  return tb_dispatch<LaunchAdj>(K, (const float2*)U, T, (const float2*)gu, N, (float2*)gV, accumulate ? 1 : 0,
                                (hipStream_t)stream);
}

extern "C" int mrh_tbasis_diag(const void* U, int T, int K, const float* ku, int profile, int64_t N, float* diag_V,
                               void* stream) {
  int rc = tb_check("tbasis_diag", T, K, N);
  if (rc != MRH_OK) return rc;
  if (N == 0) return MRH_OK;
  MRH_REQUIRE(U && ku && diag_V, MRH_EINVAL, "tbasis_diag: null argument");
  const int64_t ld_t = profile ? 1 : N, ld_n = profile ? 0 : 1;
  return tb_dispatch<LaunchDiag>(K, (const float2*)U, T, ku, ld_t, ld_n, N, diag_V, (hipStream_t)stream);
}
