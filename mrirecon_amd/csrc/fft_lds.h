// LDS-resident Stockham FFT for CDNA4, one transform per group of NTHR lanes.
//
// Formulation (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC'08): pass p with radix R and
// accumulated length Ns reads v[r] = buf[j + r N/R], multiplies by the twiddle
// W_N^{(j mod Ns) r N/(Ns R)}, applies an R-point DFT and writes
// buf[(j - j mod Ns) R + j mod Ns + r Ns].  The output is in natural order, so
// no bit reversal is needed.  Every pass holds its butterflies in registers
// across a barrier, so one LDS buffer per transform is enough.
//
// Sizes are compile-time (N = 2^a 3^b 5^c); radices are chosen 8 first, then
// 4, 2, 3, 5.  LDS indices are padded by one element per 32 (PIDX) so that the
// stride-R writes of the early passes spread over the banks (measured on the
// NUFFT pair: one per 32 beats one per 16 by 1.6 %, one per 8 by 6 % and no
// padding by 13 %).
#pragma once
#include "mrh_common.h"

#define PIDX(i) ((i) + ((i) >> 5))

__host__ __device__ constexpr int fft_padded(int n) { return n + (n >> 5); }

template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 z) {  // z * (-i) forward, z * (+i) inverse
  return INV ? make_float2(-z.y, z.x) : make_float2(z.y, -z.x);
}

template <int R, bool INV> struct DFT;

template <bool INV> struct DFT<2, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
  }
};

template <bool INV> struct DFT<4, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
  }
};

template <bool INV> struct DFT<8, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    DFT<4, INV>::run(e);
    DFT<4, INV>::run(o);
    const float h = 0.70710678118654752440f;
    // w^1 = (1 -/+ i)/sqrt2, w^2 = -/+ i, w^3 = (-1 -/+ i)/sqrt2
    float2 o1 = INV ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                    : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    float2 o2 = mul_mi<INV>(o[2]);
    float2 o3 = INV ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                    : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  }
};

template <bool INV> struct DFT<3, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    const float c = -0.5f, s = 0.86602540378443864676f;
    float2 sm = cadd(v[1], v[2]), df = csub(v[1], v[2]);
    float2 a = make_float2(fmaf(c, sm.x, v[0].x), fmaf(c, sm.y, v[0].y));
    float2 b = mul_mi<INV>(cscale(df, s));
    v[0] = cadd(v[0], sm);
    v[1] = cadd(a, b);
    v[2] = csub(a, b);
  }
};

template <bool INV> struct DFT<5, INV> {
  __device__ __forceinline__ static void run(float2* v) {
    // y_q = sum_r v_r w^{rq}, w = exp(-/+ 2 pi i / 5), via the symmetric pairs
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    float2 r1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    float2 r2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    float2 i1 = mul_mi<INV>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    float2 i2 = mul_mi<INV>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = make_float2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    v[1] = cadd(r1, i1); v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2); v[3] = csub(r2, i2);
  }
};

__host__ __device__ constexpr int fft_next_radix(int rem) {
  return (rem % 8 == 0) ? 8 : (rem % 4 == 0) ? 4 : (rem % 2 == 0) ? 2
       : (rem % 3 == 0) ? 3 : (rem % 5 == 0) ? 5 : 0;
}

__host__ __device__ constexpr bool fft_supported(int n) {
  while (n > 1) {
    int r = fft_next_radix(n);
    if (r == 0) return false;
    n /= r;
  }
  return true;
}

// Synchronisation between the read and write halves of a pass: the whole
// block (the transform's lanes span several wavefronts) or one wavefront (the
// transform is owned by a single wavefront, whose LDS operations complete in
// order; the fences order them for the compiler and the memory model).
template <bool WAVE>
__device__ __forceinline__ void fft_sync() {
  if (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Twiddle table in pass order (N - 1 entries): the pass with accumulated
// length NS and radix R finds W_N^{m (N / (NS R)) r}, r in [1, R), m in [0, NS),
// at (NS - 1) + (r - 1) NS + m.  Consecutive butterflies (m = j mod NS) read
// consecutive entries, so a wavefront's twiddle loads are free of LDS bank
// conflicts (a table indexed by the exponent is read at a stride of r).
__host__ __device__ constexpr int fft_tw_index(int ns, int r, int m) { return (ns - 1) + (r - 1) * ns + m; }

// One Stockham pass.  `tw` holds the pass-ordered twiddle table (LDS).
template <int N, int R, int NS, bool INV, int NTHR, bool WAVE = false>
__device__ __forceinline__ void fft_pass(float2* buf, const float2* tw, int t) {
  constexpr int NBF = N / R;
  constexpr int NB = (NBF + NTHR - 1) / NTHR;
  float2 v[NB][R];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int j = t + q * NTHR;
    if (NBF % NTHR == 0 || j < NBF) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[PIDX(j + r * NBF)];
      if (NS > 1) {
        const int m = j % NS;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[fft_tw_index(NS, r, m)];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      DFT<R, INV>::run(v[q]);
    }
  }
  fft_sync<WAVE>();
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int j = t + q * NTHR;
    if (NBF % NTHR == 0 || j < NBF) {
      const int jm = j % NS;
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[PIDX(base + r * NS)] = v[q][r];
    }
  }
  fft_sync<WAVE>();
}

template <int N, int NS, bool INV, int NTHR, bool WAVE>
struct FFTStages {
  __device__ __forceinline__ static void run(float2* buf, const float2* tw, int t) {
    if constexpr (NS < N) {
      constexpr int R = fft_next_radix(N / NS);
      static_assert(R != 0, "unsupported FFT size");
      fft_pass<N, R, NS, INV, NTHR, WAVE>(buf, tw, t);
      FFTStages<N, NS * R, INV, NTHR, WAVE>::run(buf, tw, t);
    }
  }
};

// In-place length-N DFT of buf (padded layout); every thread of the block must
// call it (it contains barriers).  Forward: exp(-i), inverse: exp(+i), unscaled.
template <int N, bool INV, int NTHR>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* tw, int t) {
  FFTStages<N, 1, INV, NTHR, false>::run(buf, tw, t);
}

// Same transform owned by one wavefront (lane = t in [0, 64)): no block
// barriers, so the other wavefronts of the block are never held up.  The
// caller must make the input visible to the wavefront and must not touch
// `buf` from other wavefronts during the call.
template <int N, bool INV>
__device__ __forceinline__ void fft_lds_wave(float2* buf, const float2* tw, int lane) {
  fft_sync<true>();
  FFTStages<N, 1, INV, 64, true>::run(buf, tw, lane);
}

// Length-512 transform owned by one wavefront whose lane holds rows lane + 64 r
// (r < 8) in v, in and out.  With radices 8 x 8 x 8 the first pass's butterfly
// `lane` reads exactly those rows and the last pass's butterfly `lane` writes
// them, so the first pass runs on the registers and the last pass leaves its
// results there: only the middle pass goes through LDS (one write and one read
// of the row instead of five of each).  The operations are those of fft_lds_wave;
// where the caller consumes the outputs in registers the compiler may contract
// the last pass's multiplies into different FMAs, so results can differ from
// fft_lds_wave's in the last bit (measured: 3e-8 relative on a GRASP adjoint).
// The caller must not touch `buf` from other wavefronts meanwhile.
template <bool INV>
__device__ __forceinline__ void fft512_wave_regs(float2 (&v)[8], float2* buf, const float2* tw, int lane) {
  constexpr int N = 512;
  DFT<8, INV>::run(v);                       // pass 1 (NS = 1: no twiddles)
  fft_sync<true>();                          // earlier reads of buf are done
#pragma unroll
  for (int r = 0; r < 8; ++r) buf[PIDX(lane * 8 + r)] = v[r];
  fft_sync<true>();
  fft_pass<N, 8, 8, INV, 64, true>(buf, tw, lane);   // pass 2 (NS = 8)
#pragma unroll
  for (int r = 0; r < 8; ++r) v[r] = buf[PIDX(lane + r * 64)];   // pass 3 (NS = 64), m = lane
#pragma unroll
  for (int r = 1; r < 8; ++r) {
    float2 w = tw[fft_tw_index(64, r, lane)];
    if (INV) w.y = -w.y;
    v[r] = cmul(v[r], w);
  }
  DFT<8, INV>::run(v);
}

// Same transform of a column whose rows [a, b) are zero (the zero padding of a
// centred, zero-padded image column): the first pass reads only rows outside
// [a, b), so the caller neither writes nor clears the padding.
template <int N, bool INV>
__device__ __forceinline__ void fft_lds_wave_padded(float2* buf, const float2* tw, int lane, int a, int b) {
  constexpr int R = fft_next_radix(N), NBF = N / R, NB = (NBF + 63) / 64;
  fft_sync<true>();
  float2 v[NB][R];
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int j = lane + q * 64;
    if (NBF % 64 == 0 || j < NBF) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int row = j + r * NBF;
        v[q][r] = (row < a || row >= b) ? buf[PIDX(row)] : make_float2(0.f, 0.f);
      }
      DFT<R, INV>::run(v[q]);
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
  FFTStages<N, R, INV, 64, true>::run(buf, tw, lane);
}
