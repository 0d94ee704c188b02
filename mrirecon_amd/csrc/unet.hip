// BF16 MFMA kernels of the 2D+t U-Net warm start (gfx950).
//
// Activations: BF16, channels-last (volumes * frames, H, W, C), C a multiple of
// 8.  Forward weights: BF16 [tap][Cout_pad][Cin] (Cin contiguous, Cout_pad a
// multiple of 32, padded rows zero).  Tap sets: 3x3 spatial (t = 3 (dy+1) +
// (dx+1)), 3-tap temporal (t = df+1, never crossing a volume's first or last
// frame) and 1x1.  Out-of-range neighbours read as zero (zero padding).
//
// Convolutions are implicit GEMMs on BF16 MFMA with FP32 accumulation:
// v_mfma_f32_16x16x32_bf16 (lane l holds A row l % 16 and B column l % 16 with
// the eight K indices 8 (l/16) .. +7) and v_mfma_f32_16x16x16_bf16 (four K
// indices 4 (l/16) .. +3); D rows 4 (l/16) .. +3 of column l % 16.
//   forward          D[cout][pixel] = sum_(tap, cin) W[tap][cout][cin] x[pixel + off(tap)][cin]
//                    epilogue: + bias, optional ReLU, round to BF16.
//   backward-data    the same kernel applied to the output gradient with the
//                    weights transposed (cin <-> cout) and tap-flipped
//                    (off(T-1-t) = -off(t) for every tap set used here).
//   backward-weight  dW[tap][cout][cin] = sum_pixel g[pixel][cout] x[pixel + off(tap)][cin]:
//                    pixel chunks (split-K) staged transposed in LDS, FP32
//                    atomics per 32 x 32 block.
#include <algorithm>
#include "mrh_common.h"
#include "mrirecon_hip.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(u16 b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ u16 f2bf(float f) {   // round to nearest even (finite inputs)
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

struct TapGeom {
  int kind;      // MRH_UNET_TAP_1X1, _3X3 or _T3
  int ntap;      // 1, 9 or 3
  int F, H, W;   // frames per volume, image size
};

__device__ __forceinline__ void tap_offset(const TapGeom& g, int t, int& df, int& dy, int& dx) {
  df = dy = dx = 0;
  if (g.kind == MRH_UNET_TAP_3X3) { dy = t / 3 - 1; dx = t - 3 * (t / 3) - 1; }
  else if (g.kind == MRH_UNET_TAP_T3) { df = t - 1; }
}

__device__ __forceinline__ bf16x8 load8(const u16* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

// Workgroups are dealt to the 8 XCDs round-robin in dispatch order.  Remap the
// linear workgroup index so that each XCD runs a contiguous range of tiles
// (pixel tiles fastest): the halo rows a tile shares with its neighbours are
// then fetched into one XCD's L2 instead of several.  A bijection of
// [0, gx*gy) for any grid size.
__device__ __forceinline__ void xcd_tile(int& bx, int& by) {
  const int gx = gridDim.x, n = gx * gridDim.y;
  const int b = blockIdx.x + blockIdx.y * gx;
  const int q = n / 8, rem = n % 8, x = b % 8, k = b / 8;
  const int t = x * q + min(x, rem) + k;
  bx = t % gx;
  by = t / gx;
}

// ------------------------------------------------------------------ forward / backward-data
// Grid (ceil(npix / 128), Cout_pad / 32), 256 threads; wave w computes couts
// [32 by, +32) x pixels [128 bx + 32 w, +32) as 2 x 2 MFMA tiles.
__global__ __launch_bounds__(256) void k_conv_fwd(const u16* __restrict__ x, const u16* __restrict__ w,
                                                  const float* __restrict__ bias, u16* __restrict__ out,
                                                  TapGeom g, int npix, int Cin, int Cout, int Cout_pad,
                                                  int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int co0 = blockIdx.y * 32;
  const int px0 = blockIdx.x * 128 + wave * 32;
  const int r = lane & 15, kg = lane >> 4;
  const int HW = g.H * g.W;
  int pfr[2], py[2], pxx[2];
  bool pv[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = px0 + 16 * h + r;
    pv[h] = p < npix;
    const int pp = pv[h] ? p : 0;
    pxx[h] = pp % g.W;
    const int t1 = pp / g.W;
    py[h] = t1 % g.H;
    pfr[h] = (t1 / g.H) % g.F;   // frame within its volume
  }
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int h = 0; h < 2; ++h) acc[a][h] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int cgroups = Cin >> 3;            // K = ntap * cgroups groups of 8 channels
  const int ngroups = g.ntap * cgroups;
  int t = kg / cgroups, cg = kg - t * cgroups;   // this lane's group (g0 + kg) as (tap, channel group)
  const u16* wrow0 = w + (size_t)(co0 + r) * Cin;
  const u16* wrow1 = w + (size_t)(co0 + 16 + r) * Cin;
  const size_t wtap = (size_t)Cout_pad * Cin;
  for (int g0 = 0; g0 < ngroups; g0 += 4) {
    bf16x8 av[2], bv[2];
    if (t < g.ntap) {
      const int ci = cg * 8;
      av[0] = load8(wrow0 + t * wtap + ci);
      av[1] = load8(wrow1 + t * wtap + ci);
      int df, dy, dx;
      tap_offset(g, t, df, dy, dx);
      const int off = df * HW + dy * g.W + dx;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const bool ok = pv[h] && (unsigned)(pfr[h] + df) < (unsigned)g.F &&
                        (unsigned)(py[h] + dy) < (unsigned)g.H && (unsigned)(pxx[h] + dx) < (unsigned)g.W;
        bv[h] = ok ? load8(x + (size_t)(px0 + 16 * h + r + off) * Cin + ci) : zero8();
      }
    } else {
      av[0] = av[1] = bv[0] = bv[1] = zero8();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        acc[a][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[h], acc[a][h], 0, 0, 0);
    cg += 4;
    while (cg >= cgroups) { cg -= cgroups; ++t; }
  }
  // lane holds couts co0 + 16a + 4kg + j of pixel px0 + 16h + r
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int p = px0 + 16 * h + r;
    if (p >= npix) continue;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int cb = co0 + 16 * a + 4 * kg;
      if (cb >= Cout) continue;
      u16 o[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float v = acc[a][h][j] + ((bias && cb + j < Cout) ? bias[cb + j] : 0.f);
        if (relu) v = fmaxf(v, 0.f);
        o[j] = f2bf(v);
      }
      if (cb + 3 < Cout) {
        *reinterpret_cast<uint2*>(out + (size_t)p * Cout + cb) =
            make_uint2((unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16));
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (cb + j < Cout) out[(size_t)p * Cout + cb + j] = o[j];
      }
    }
  }
}

// ------------------------------------------------------------------ forward, LDS-tiled (Cin % 16 == 0)
// Block: 4 waves along pixels; wave tile (16 MT) couts x (16 NT) pixels, block
// tile BC = 16 MT couts x BP = 64 NT consecutive (flattened) pixels.  Per chunk
// of 16 input channels the block stages into LDS
//   * the input rows the taps touch: for each tap row offset r (dy for 3x3, df
//     for temporal, none for 1x1) the pixels [p0 + r*RS - 1, p0 + r*RS + BP + 1)
//     (RS = W or H*W), 16 channels each, out-of-range pixels as zero;
//   * the weights [tap][BC couts][16 channels];
// then runs one 16x16x16 MFMA K-step per tap.  Each staged input element serves
// every tap and every cout tile of the block (the register-only kernel above
// re-reads it per tap); boundary masks are applied per (pixel, tap) at read time.
template <int MT, int NT, int KC>
__global__ __launch_bounds__(256) void k_conv_fwd_lds(const u16* __restrict__ x, const u16* __restrict__ w,
                                                      const float* __restrict__ bias, u16* __restrict__ out,
                                                      TapGeom g, int npix, int Cin, int Cout, int Cout_pad,
                                                      int relu) {
  typedef short s4 __attribute__((ext_vector_type(4)));
  constexpr int BP = 64 * NT, BC = 16 * MT, RW = BP + 2, NCH = KC / 8;   // NCH 16-byte groups per pixel
  __shared__ __align__(16) u16 sIn[3 * RW * KC];
  __shared__ __align__(16) u16 sW[9 * BC * KC];
  static_assert(KC == 16 || KC % 32 == 0, "channel chunk");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  // 3x3 taps: neighbouring tiles share halo rows, so each XCD takes a contiguous run of
  // tiles.  Temporal taps reuse rows a whole frame away, which the plain dispatch order
  // already keeps on one XCD (and the contiguous runs would evict before reuse).
  int bx = blockIdx.x, by = blockIdx.y;
  if (g.kind == MRH_UNET_TAP_3X3) xcd_tile(bx, by);
  const int p0 = bx * BP, co0 = by * BC;
  const int HW = g.H * g.W;
  const int nrows = g.kind == MRH_UNET_TAP_1X1 ? 1 : 3;
  const int RS = g.kind == MRH_UNET_TAP_3X3 ? g.W : (g.kind == MRH_UNET_TAP_T3 ? HW : 0);
  int pxx[NT], py[NT], pfr[NT];
  bool pv[NT];
#pragma unroll
  for (int h = 0; h < NT; ++h) {
    const int p = p0 + wave * 16 * NT + 16 * h + r;
    pv[h] = p < npix;
    const int pp = pv[h] ? p : 0;
    pxx[h] = pp % g.W;
    const int t1 = pp / g.W;
    py[h] = t1 % g.H;
    pfr[h] = (t1 / g.H) % g.F;
  }
  f32x4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int h = 0; h < NT; ++h) acc[m][h] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int ci0 = 0; ci0 < Cin; ci0 += KC) {
    __syncthreads();   // the previous chunk's fragments have been read
    for (int i = tid; i < nrows * RW * NCH; i += 256) {
      const int row = i / (RW * NCH), rem = i - row * (RW * NCH);
      const int q = rem / NCH, half = rem - q * NCH;
      const int rr = nrows == 1 ? 0 : row - 1;
      const long long src = (long long)p0 + (long long)rr * RS - 1 + q;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (src >= 0 && src < npix) v = *reinterpret_cast<const uint4*>(x + src * Cin + ci0 + half * 8);
      *reinterpret_cast<uint4*>(sIn + (row * RW + q) * KC + half * 8) = v;
    }
    for (int i = tid; i < g.ntap * BC * NCH; i += 256) {
      const int t = i / (BC * NCH), rem = i - t * (BC * NCH);
      const int co = rem / NCH, half = rem - co * NCH;
      *reinterpret_cast<uint4*>(sW + (t * BC + co) * KC + half * 8) =
          *reinterpret_cast<const uint4*>(w + ((size_t)t * Cout_pad + co0 + co) * Cin + ci0 + half * 8);
    }
    __syncthreads();
    for (int t = 0; t < g.ntap; ++t) {
      int df, dy, dx;
      tap_offset(g, t, df, dy, dx);
      const int row = nrows == 1 ? 0 : (g.kind == MRH_UNET_TAP_3X3 ? dy : df) + 1;
      bool ok[NT];
      int lp[NT];
#pragma unroll
      for (int h = 0; h < NT; ++h) {
        ok[h] = pv[h] && (unsigned)(pfr[h] + df) < (unsigned)g.F && (unsigned)(py[h] + dy) < (unsigned)g.H &&
                (unsigned)(pxx[h] + dx) < (unsigned)g.W;
        lp[h] = (row * RW + wave * 16 * NT + 16 * h + r + 1 + dx) * KC;
      }
      if constexpr (KC == 16) {            // one 16x16x16 step: lane holds channels 4 kg .. +3
        s4 a[MT], b[NT];
#pragma unroll
        for (int m = 0; m < MT; ++m) a[m] = *reinterpret_cast<const s4*>(sW + (t * BC + 16 * m + r) * KC + 4 * kg);
#pragma unroll
        for (int h = 0; h < NT; ++h) b[h] = ok[h] ? *reinterpret_cast<const s4*>(sIn + lp[h] + 4 * kg) : s4{0, 0, 0, 0};
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int h = 0; h < NT; ++h)
            acc[m][h] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b[h], acc[m][h], 0, 0, 0);
      } else {                             // 16x16x32 steps: lane holds channels 32 s + 8 kg .. +7
#pragma unroll
        for (int s = 0; s < KC / 32; ++s) {
          bf16x8 a[MT], b[NT];
#pragma unroll
          for (int m = 0; m < MT; ++m) a[m] = load8(sW + (t * BC + 16 * m + r) * KC + 32 * s + 8 * kg);
#pragma unroll
          for (int h = 0; h < NT; ++h) b[h] = ok[h] ? load8(sIn + lp[h] + 32 * s + 8 * kg) : zero8();
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int h = 0; h < NT; ++h)
              acc[m][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[h], acc[m][h], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int h = 0; h < NT; ++h) {
    const int p = p0 + wave * 16 * NT + 16 * h + r;
    if (p >= npix) continue;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int cb = co0 + 16 * m + 4 * kg;
      if (cb >= Cout) continue;
      u16 o[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float v = acc[m][h][j] + ((bias && cb + j < Cout) ? bias[cb + j] : 0.f);
        if (relu) v = fmaxf(v, 0.f);
        o[j] = f2bf(v);
      }
      if (cb + 3 < Cout) {
        *reinterpret_cast<uint2*>(out + (size_t)p * Cout + cb) =
            make_uint2((unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16));
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (cb + j < Cout) out[(size_t)p * Cout + cb + j] = o[j];
      }
    }
  }
}

// ------------------------------------------------------------------ 3x3 / temporal forward on tiles
// Block: 256 output pixels of one image x BC = 16 MT couts; Cin in chunks of
// 32 (one 16x16x32 K-step per tap).
//   3x3 (TEMP = false): the tile is TR = 256 / TW rows x TW columns; per chunk
//     the block stages the (TR + 2) x (TW + 2) halo (zero outside the image).
//   temporal (TEMP = true): the tile is 256 consecutive pixels of one frame;
//     per chunk the block stages those pixels in the previous, the same and
//     the next frame of the volume (zero beyond its first / last frame).
// The chunk's [tap][BC] weights are staged beside them.  Pixel and cout rows
// of the LDS tiles are 40 halves (80 B) apart, so the 16 lanes of a fragment
// read hit distinct banks.  The next chunk's input is fetched into registers
// while the current one is multiplied.  Wave w owns tile pixels
// [64 w, 64 w + 64) as four 16-pixel MFMA columns.
// Needs Cin % 32 == 0 and, for 3x3, W % TW == 0 and H % TR == 0 (temporal:
// H W % 256 == 0).
// UPC (3x3 only): the input is the decoder's concatenation [up2(x), x2] read in
// place -- channels [0, C1) from the half-resolution x (nearest neighbour),
// channels [C1, Cin) from the full-resolution skip x2 -- so it is never stored
// (C1 % 32 == 0: a chunk never straddles the two).
template <int MT, int TW, bool TEMP, bool UPC = false>
__global__ __launch_bounds__(256, 2) void k_conv_tile(const u16* __restrict__ x, const u16* __restrict__ w,
                                                     const float* __restrict__ bias, u16* __restrict__ out, int F,
                                                     int H, int W, int Cin, int Cout, int Cout_pad, int relu,
                                                     const u16* __restrict__ x2 = nullptr, int C1 = 0) {
  constexpr int TR = TEMP ? 1 : 256 / TW, HC = TEMP ? 256 : TW + 2, NHP = (TR + 2) * HC;
  constexpr int NTAP = TEMP ? 3 : 9, KC = 32, PS = KC + 8, BC = 16 * MT;
  constexpr int NIN = NHP * 4, IPT = (NIN + 255) / 256;        // 16-byte pieces of the halo, per thread
  constexpr int NWT = NTAP * BC * 4, WPT = (NWT + 255) / 256;   // and of the weights
  __shared__ __align__(16) u16 sIn[NHP * PS];
  __shared__ __align__(16) u16 sW[NTAP * BC * PS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  int bx, by;
  xcd_tile(bx, by);
  const int HW = H * W;
  const int tx_n = TEMP ? 1 : W / TW, tpi = TEMP ? HW / 256 : tx_n * (H / TR);   // tiles per image
  // temporal: frames fastest, so the three blocks that read a frame's tile run back to back
  int img, trem;
  if (TEMP) {
    const int vol = bx / (F * tpi), vr = bx - vol * (F * tpi);
    trem = vr / F;
    img = vol * F + (vr - trem * F);
  } else {
    img = bx / tpi;
    trem = bx - img * tpi;
  }
  const int y0 = TEMP ? 0 : (trem / tx_n) * TR, x0 = TEMP ? 0 : (trem % tx_n) * TW;
  const int p0 = TEMP ? trem * 256 : 0, fv = img % F, co0 = by * BC;
  const long long ibase = (long long)img * HW;
  u32x4 pin[IPT], pw[WPT];
  auto gload = [&](int ci0) {
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int idx = tid + 256 * i;
      pin[i] = u32x4{0u, 0u, 0u, 0u};
      if (idx < NIN) {
        const int hp = idx >> 2, q = idx & 3;
        const int hy = hp / HC, hx = hp - hy * HC;
        long long src = -1;
        if (TEMP) {
          if ((unsigned)(fv + hy - 1) < (unsigned)F) src = ibase + (long long)(hy - 1) * HW + p0 + hx;
        } else {
          const int gy = y0 + hy - 1, gx = x0 + hx - 1;
          if ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W) {
            src = ibase + (long long)gy * W + gx;
            if (UPC && ci0 < C1) src = ((long long)img * (H / 2) + gy / 2) * (W / 2) + gx / 2;
          }
        }
        if (src >= 0) {
          const u16* base = x + src * Cin + ci0;
          if (UPC) base = ci0 < C1 ? x + src * C1 + ci0 : x2 + src * (Cin - C1) + (ci0 - C1);
          pin[i] = *reinterpret_cast<const u32x4*>(base + 8 * q);
        }
      }
    }
  };
  // halo index of this lane's pixel in each of the wave's four MFMA columns (centre tap)
  int hp0[4];
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int tp = 64 * wave + 16 * h + r;
    hp0[h] = TEMP ? HC + tp : (tp / TW + 1) * HC + (tp % TW) + 1;
  }
  f32x4 acc[MT][4];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int h = 0; h < 4; ++h) acc[m][h] = f32x4{0.f, 0.f, 0.f, 0.f};
  gload(0);
  for (int ci0 = 0; ci0 < Cin; ci0 += KC) {
    __syncthreads();                       // the previous chunk's fragments have been read
#pragma unroll
    for (int i = 0; i < WPT; ++i) {
      const int idx = tid + 256 * i;
      if (idx < NWT) {
        const int t = idx / (BC * 4), rem = idx - t * (BC * 4);
        pw[i] = *reinterpret_cast<const u32x4*>(w + ((size_t)t * Cout_pad + co0 + (rem >> 2)) * Cin + ci0 +
                                                8 * (rem & 3));
      }
    }
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int idx = tid + 256 * i;
      if (idx < NIN) *reinterpret_cast<u32x4*>(sIn + (idx >> 2) * PS + 8 * (idx & 3)) = pin[i];
    }
#pragma unroll
    for (int i = 0; i < WPT; ++i) {
      const int idx = tid + 256 * i;
      if (idx < NWT) *reinterpret_cast<u32x4*>(sW + (idx >> 2) * PS + 8 * (idx & 3)) = pw[i];
    }
    __syncthreads();
    if (ci0 + KC < Cin) gload(ci0 + KC);
#pragma unroll
    for (int t = 0; t < NTAP; ++t) {
      const int off = TEMP ? (t - 1) * HC : (t / 3 - 1) * HC + (t % 3 - 1);
      bf16x8 a[MT], b[4];
#pragma unroll
      for (int m = 0; m < MT; ++m) a[m] = load8(sW + (t * BC + 16 * m + r) * PS + 8 * kg);
#pragma unroll
      for (int h = 0; h < 4; ++h) b[h] = load8(sIn + (hp0[h] + off) * PS + 8 * kg);
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int h = 0; h < 4; ++h) acc[m][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b[h], acc[m][h], 0, 0, 0);
    }
  }
  // lane holds couts co0 + 16 m + 4 kg + j of tile pixel 64 w + 16 h + r
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int tp = 64 * wave + 16 * h + r;
    const long long p = ibase + (TEMP ? p0 + tp : (long long)(y0 + tp / TW) * W + x0 + tp % TW);
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const int cb = co0 + 16 * m + 4 * kg;
      if (cb >= Cout) continue;
      u16 o[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        float v = acc[m][h][j] + ((bias && cb + j < Cout) ? bias[cb + j] : 0.f);
        if (relu) v = fmaxf(v, 0.f);
        o[j] = f2bf(v);
      }
      if (cb + 3 < Cout) {
        *reinterpret_cast<uint2*>(out + p * Cout + cb) =
            make_uint2((unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16));
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (cb + j < Cout) out[p * Cout + cb + j] = o[j];
      }
    }
  }
}

// ------------------------------------------------------------------ temporal layers walking frames
// 3-tap temporal convolution with Cin = 32 NCH and 16 MT-cout blocks.  A temporal tap reads the
// same pixel in the previous, current and next frame, so there is no spatial halo: every MFMA B
// fragment (16 bytes of one pixel's channels per lane) is loaded straight from the channels-last
// input into a ring of registers that holds frames f - 1 .. f + PD, and each wave walks a group of
// frames of its 16 NT pixels: an input frame is loaded once per wave instead of once per tap, and
// the next PD frames are in flight while frame f is multiplied.  The block's [tap][16 MT couts][Cin]
// weights are staged in LDS once.  Per output the MFMAs run chunk by chunk, taps inside a chunk,
// with zero fragments beyond the volume's first / last frame and the same epilogue (+ bias,
// optional ReLU, round to BF16), as k_conv_tile<.., TEMP = true>: the results are the same bits.
// Grid (nvol * ngroups * npb * ncb), WPB waves: block b runs cout block b % ncb of pixel block
// (b / ncb) % npb (16 NT WPB pixels, 16 NT per wave) over frames [fg grp, + fg) of its volume.
template <int MT, int NT, int NCH, int PD, int WPB>
__global__ __launch_bounds__(64 * WPB) void k_conv_t3_walk(const u16* __restrict__ x, const u16* __restrict__ w,
                                                      const float* __restrict__ bias, u16* __restrict__ out, int F,
                                                      int HW, int Cout, int Cout_pad, int relu, int fg, int npb,
                                                      int ncb) {
  constexpr int BC = 16 * MT, CIN = 32 * NCH, CP = CIN + 8, RS = 2 + PD;
  __shared__ __align__(16) u16 sW[3 * BC * CP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, kg = lane >> 4;
  const int b = blockIdx.x;
  const int cb = b % ncb, pb = (b / ncb) % npb, vg = b / (ncb * npb);
  const int ng = (F + fg - 1) / fg, vol = vg / ng, grp = vg - vol * ng;
  const int co0 = cb * BC, fa = grp * fg, fb = min(F, fa + fg);
  for (int i = tid; i < 3 * BC * (CIN / 8); i += 64 * WPB) {
    const int row = i / (CIN / 8), q = i - row * (CIN / 8);   // row = tap * BC + cout
    const int t = row / BC, c = row - t * BC;
    *reinterpret_cast<u32x4*>(sW + row * CP + 8 * q) =
        *reinterpret_cast<const u32x4*>(w + ((size_t)t * Cout_pad + co0 + c) * CIN + 8 * q);
  }
  const int px = (pb * WPB + wave) * 16 * NT + r;            // this lane's pixel in MFMA column 0
  const u16* xv = x + ((size_t)vol * F * HW + px) * CIN + 8 * kg;
  bf16x8 ring[RS][NT][NCH];
  auto fetch = [&](int f, bf16x8 (&v)[NT][NCH]) {
    const bool in = (unsigned)f < (unsigned)F;
#pragma unroll
    for (int h = 0; h < NT; ++h)
#pragma unroll
      for (int c = 0; c < NCH; ++c)
        v[h][c] = in ? load8(xv + ((size_t)f * HW + 16 * h) * CIN + 32 * c) : zero8();
  };
#pragma unroll
  for (int s = 0; s < RS - 1; ++s) fetch(fa - 1 + s, ring[s]);
  __syncthreads();
  for (int f = fa; f < fb; ++f) {
    if (f + PD <= fb) fetch(f + PD, ring[RS - 1]);
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int h = 0; h < NT; ++h) acc[m][h] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < NCH; ++c)
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        bf16x8 a[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) a[m] = load8(sW + (t * BC + 16 * m + r) * CP + 32 * c + 8 * kg);
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int h = 0; h < NT; ++h)
            acc[m][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], ring[t][h][c], acc[m][h], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);   // keep each step's weight loads beside its MFMAs (registers)
      }
    // lane holds couts co0 + 16 m + 4 kg + j of pixel px + 16 h
#pragma unroll
    for (int h = 0; h < NT; ++h) {
      const size_t p = ((size_t)vol * F + f) * HW + px + 16 * h;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int cbase = co0 + 16 * m + 4 * kg;
        if (cbase >= Cout) continue;
        u16 o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          float v = acc[m][h][j] + ((bias && cbase + j < Cout) ? bias[cbase + j] : 0.f);
          if (relu) v = fmaxf(v, 0.f);
          o[j] = f2bf(v);
        }
        if (cbase + 3 < Cout) {
          *reinterpret_cast<uint2*>(out + p * Cout + cbase) =
              make_uint2((unsigned)o[0] | ((unsigned)o[1] << 16), (unsigned)o[2] | ((unsigned)o[3] << 16));
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (cbase + j < Cout) out[p * Cout + cbase + j] = o[j];
        }
      }
    }
#pragma unroll
    for (int s = 0; s < RS - 1; ++s)
#pragma unroll
      for (int h = 0; h < NT; ++h)
#pragma unroll
        for (int c = 0; c < NCH; ++c) ring[s][h][c] = ring[s + 1][h][c];
  }
}

// ------------------------------------------------------------------ backward-weight
// Grid (nchunks, Cout_pad / 32, ntap * ceil(Cin / 32)), 256 threads.  Block
// computes dW[tap][32 couts][32 cins] over pixels [chunk * bx, +chunk) in
// steps of 128 pixels; waves form a 2 x 2 grid of 16 x 16 MFMA tiles.
#define BW_STEP 128
__global__ __launch_bounds__(256) void k_conv_bw(const u16* __restrict__ gout, const u16* __restrict__ x,
                                                 float* __restrict__ dw, TapGeom g, int npix, int chunk,
                                                 int Cin, int Cout, int Cout_pad) {
  __shared__ __align__(16) u16 sA[32][BW_STEP + 8];   // [cout][pixel]
  __shared__ __align__(16) u16 sB[32][BW_STEP + 8];   // [cin][pixel]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int co0 = blockIdx.y * 32;
  const int cblocks = (Cin + 31) / 32;
  const int t = blockIdx.z / cblocks, ci0 = (blockIdx.z - t * cblocks) * 32;
  const int p_begin = blockIdx.x * chunk;
  const int p_end = min(npix, p_begin + chunk);
  const int HW = g.H * g.W;
  const int r = lane & 15, kg = lane >> 4;
  const int wa = wave >> 1, wb = wave & 1;
  int df, dy, dx;
  tap_offset(g, t, df, dy, dx);
  const int off = df * HW + dy * g.W + dx;
  const int sc = (tid & 7) * 4;               // staged channel quad
  const bool a_full = co0 + sc + 3 < Cout, b_ok = ci0 + sc < Cin;   // Cin % 8 == 0: quads never straddle
  f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int p0 = p_begin; p0 < p_end; p0 += BW_STEP) {
    uint2 va[4], vb[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int sp = (tid >> 3) + 32 * s;
      const int p = p0 + sp;
      va[s] = vb[s] = make_uint2(0, 0);
      if (p < p_end) {
        if (a_full) {
          va[s] = *reinterpret_cast<const uint2*>(gout + (size_t)p * Cout + co0 + sc);
        } else {
          unsigned q[4] = {0, 0, 0, 0};
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (co0 + sc + j < Cout) q[j] = gout[(size_t)p * Cout + co0 + sc + j];
          va[s] = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
        }
        if (b_ok) {
          const int xx = p % g.W, t1 = p / g.W, yy = t1 % g.H, f = (t1 / g.H) % g.F;
          if ((unsigned)(f + df) < (unsigned)g.F && (unsigned)(yy + dy) < (unsigned)g.H &&
              (unsigned)(xx + dx) < (unsigned)g.W)
            vb[s] = *reinterpret_cast<const uint2*>(x + (size_t)(p + off) * Cin + ci0 + sc);
        }
      }
    }
    __syncthreads();   // every wave has read the previous step's fragments
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int sp = (tid >> 3) + 32 * s;
      sA[sc + 0][sp] = (u16)(va[s].x & 0xffffu); sA[sc + 1][sp] = (u16)(va[s].x >> 16);
      sA[sc + 2][sp] = (u16)(va[s].y & 0xffffu); sA[sc + 3][sp] = (u16)(va[s].y >> 16);
      sB[sc + 0][sp] = (u16)(vb[s].x & 0xffffu); sB[sc + 1][sp] = (u16)(vb[s].x >> 16);
      sB[sc + 2][sp] = (u16)(vb[s].y & 0xffffu); sB[sc + 3][sp] = (u16)(vb[s].y >> 16);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BW_STEP / 32; ++kk) {
      const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sA[16 * wa + r][32 * kk + 8 * kg]));
      const bf16x8 b = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&sB[16 * wb + r][32 * kk + 8 * kg]));
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
  }
  const int ci = ci0 + 16 * wb + r;
  if (ci < Cin) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int co = co0 + 16 * wa + 4 * kg + j;
      if (co < Cout) unsafeAtomicAdd(&dw[((size_t)t * Cout_pad + co) * Cin + ci], acc[j]);
    }
  }
}

// db[c] += sum over a 4096-pixel chunk of g[p][c]; 256 threads = (256 / C) pixel lanes x C channels
__global__ __launch_bounds__(256) void k_bias_grad(const u16* __restrict__ gout, float* __restrict__ db, int npix,
                                                   int C) {
  __shared__ float part[256];
  const int lanes = 256 / C;
  const int c = threadIdx.x % C, pl = threadIdx.x / C;
  float s = 0.f;
  if (pl < lanes) {
    const int p1 = min(npix, (int)(blockIdx.x + 1) * 4096);
    for (int p = blockIdx.x * 4096 + pl; p < p1; p += lanes) s += bf2f(gout[(size_t)p * C + c]);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  if (pl == 0) {
    for (int q = 1; q < lanes; ++q) s += part[q * C + c];
    unsafeAtomicAdd(&db[c], s);
  }
}

// ------------------------------------------------------------------ elementwise
#define GRID_STRIDE(i, n) \
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// Elementwise kernels move 8 channels (16 bytes) per thread; every channel count is a multiple of 8.
__device__ __forceinline__ void unpack8(uint4 v, float* f) {
  const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[2 * k] = __uint_as_float(w[k] << 16);
    f[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}
__device__ __forceinline__ uint4 pack8(const float* f) {
  unsigned w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = (unsigned)f2bf(f[2 * k]) | ((unsigned)f2bf(f[2 * k + 1]) << 16);
  return make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ uint4 ld8(const u16* p) { return *reinterpret_cast<const uint4*>(p); }
__device__ __forceinline__ void st8(u16* p, uint4 v) { *reinterpret_cast<uint4*>(p) = v; }

// g = 0 where the ReLU output y <= 0 (n elements, n % 8 == 0)
__global__ void k_relu_bwd(u16* __restrict__ g, const u16* __restrict__ y, int64_t n8) {
  GRID_STRIDE(i, n8) {
    float fy[8], fg[8];
    unpack8(ld8(y + 8 * i), fy);
    unpack8(ld8(g + 8 * i), fg);
#pragma unroll
    for (int k = 0; k < 8; ++k) if (!(fy[k] > 0.f)) fg[k] = 0.f;
    st8(g + 8 * i, pack8(fg));
  }
}

__global__ void k_pool_fwd(const u16* __restrict__ x, u16* __restrict__ y, int nf, int H, int W, int C) {
  const int Ho = H / 2, Wo = W / 2, CG = C / 8;
  GRID_STRIDE(i, (int64_t)nf * Ho * Wo * CG) {
    const int c = (int)(i % CG) * 8;
    int64_t q = i / CG;
    const int xo = q % Wo; q /= Wo;
    const int yo = q % Ho;
    const int64_t f = q / Ho;
    const size_t b = (((size_t)f * H + 2 * yo) * W + 2 * xo) * C + c;
    const size_t row = (size_t)W * C;
    float a[8], t[8];
    unpack8(ld8(x + b), a);
    unpack8(ld8(x + b + C), t);
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] += t[k];
    unpack8(ld8(x + b + row), t);
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] += t[k];
    unpack8(ld8(x + b + row + C), t);
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = 0.25f * (a[k] + t[k]);
    st8(y + 8 * i, pack8(a));
  }
}

// gx = pool^T gy (+ add)
__global__ void k_pool_bwd(const u16* __restrict__ gy, const u16* __restrict__ add, u16* __restrict__ gx, int nf,
                           int H, int W, int C) {
  const int CG = C / 8;
  GRID_STRIDE(i, (int64_t)nf * H * W * CG) {
    const int c = (int)(i % CG) * 8;
    int64_t q = i / CG;
    const int xx = q % W; q /= W;
    const int yy = q % H;
    const int64_t f = q / H;
    float v[8];
    unpack8(ld8(gy + (((size_t)f * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c), v);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] *= 0.25f;
    if (add) {
      float a[8];
      unpack8(ld8(add + 8 * i), a);
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] += a[k];
    }
    st8(gx + 8 * i, pack8(v));
  }
}

// dst = concat(nearest_up2(src) [C1 channels], skip [C2 channels]) at H x W
__global__ void k_up_concat(const u16* __restrict__ src, const u16* __restrict__ skip, u16* __restrict__ dst,
                            int nf, int H, int W, int C1, int C2) {
  const int CG1 = C1 / 8, CG = (C1 + C2) / 8;
  GRID_STRIDE(i, (int64_t)nf * H * W * CG) {
    const int cg = (int)(i % CG);
    const int64_t pix = i / CG;
    uint4 v;
    if (cg < CG1) {
      const int xx = pix % W;
      const int64_t q = pix / W;
      const int yy = q % H;
      const int64_t f = q / H;
      v = ld8(src + (((size_t)f * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C1 + 8 * cg);
    } else {
      v = ld8(skip + (size_t)pix * C2 + 8 * (cg - CG1));
    }
    st8(dst + 8 * i, v);
  }
}

// gsrc = up2^T gdst[:, :C1] (H/2 x W/2), gskip = gdst[:, C1:]
__global__ void k_up_concat_bwd(const u16* __restrict__ gdst, u16* __restrict__ gsrc, u16* __restrict__ gskip,
                                int nf, int H, int W, int C1, int C2) {
  const int C = C1 + C2, Ho = H / 2, Wo = W / 2, CG1 = C1 / 8, CG2 = C2 / 8;
  const int64_t nsrc = (int64_t)nf * Ho * Wo * CG1, nsk = (int64_t)nf * H * W * CG2;
  GRID_STRIDE(i, nsrc + nsk) {
    if (i < nsrc) {
      const int c = (int)(i % CG1) * 8;
      int64_t q = i / CG1;
      const int xo = q % Wo; q /= Wo;
      const int yo = q % Ho;
      const int64_t f = q / Ho;
      const size_t b = (((size_t)f * H + 2 * yo) * W + 2 * xo) * C + c;
      const size_t row = (size_t)W * C;
      float a[8], t[8];
      unpack8(ld8(gdst + b), a);
      unpack8(ld8(gdst + b + C), t);
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] += t[k];
      unpack8(ld8(gdst + b + row), t);
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] += t[k];
      unpack8(ld8(gdst + b + row + C), t);
#pragma unroll
      for (int k = 0; k < 8; ++k) a[k] += t[k];
      st8(gsrc + 8 * i, pack8(a));
    } else {
      const int64_t j = i - nsrc;
      st8(gskip + 8 * j, ld8(gdst + (j / CG2) * C + C1 + 8 * (j % CG2)));
    }
  }
}

__global__ void k_pack_input(const float2* __restrict__ img, u16* __restrict__ out, int64_t npix, int C,
                             float scale) {
  GRID_STRIDE(i, npix) {
    const float2 v = img[i];
    u16* o = out + i * C;
    o[0] = f2bf(v.x * scale);
    o[1] = f2bf(v.y * scale);
    for (int c = 2; c < C; ++c) o[c] = 0;
  }
}

__global__ void k_unpack_residual(const float2* __restrict__ xin, const u16* __restrict__ net,
                                  float2* __restrict__ xout, int64_t npix, int Cn, float inv_scale) {
  GRID_STRIDE(i, npix) {
    const float2 a = xin[i];
    xout[i] = make_float2(a.x + bf2f(net[i * Cn]) * inv_scale, a.y + bf2f(net[i * Cn + 1]) * inv_scale);
  }
}

// ------------------------------------------------------------------ fused full-resolution level (inference)
// The 32-channel activations of the full-resolution level that only feed the next
// layer (a_0, h_0, the network output and the packed input) never reach HBM: each
// kernel fixes a pixel tile and visits the frames of its frame range in order, so a
// temporal tap's three frames are on chip.  Same operations, roundings and epilogues
// as the separate layers (BF16 storage of every activation, FP32 accumulation); only
// the MFMA accumulation order of the sums differs.

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ unsigned bf_pair(float a, float b) { return (unsigned)f2bf(a) | ((unsigned)f2bf(b) << 16); }

#define L0_C 32        // channels of the full-resolution level
#define L0_RS 40       // LDS stride (halves) of one pixel's 32 channels: 16-lane fragment reads hit distinct banks
#define E0_TH 8        // enc0 tile: 8 rows x 32 columns, wave w owns rows 2w and 2w + 1
#define E0_TW 32
#define E0_HX (E0_TW + 2)
#define E0_NH ((E0_TH + 2) * E0_HX)   // halo pixels of the 3x3 taps

// e_0 = relu(convT(relu(conv3x3(pack(x))))) and i_1 = avgpool2x2(e_0) for one volume's
// frames [fg by, fg by + fg).  pack(x) = (re, im, 0, ..., 0) * scale as BF16 (8 channels).
// The MFMA operands are those of k_conv_fwd for these layers (3x3: K steps of four taps x 8
// channels; temporal: one K step per tap), so every sum and rounding is the same and the
// result equals the separate layers bit for bit.  Grid (H/8 * W/32, ceil(F/fg)).
__global__ __launch_bounds__(256) void k_enc0_fused(const float2* __restrict__ x, const u16* __restrict__ ws,
                                                    const float* __restrict__ bs, const u16* __restrict__ wt,
                                                    const float* __restrict__ bt, u16* __restrict__ e0,
                                                    u16* __restrict__ i1, int F, int H, int W, int fg,
                                                    float scale) {
  __shared__ unsigned sX[2][E0_NH];
  __shared__ __align__(16) u16 sR[4][3][64 * L0_RS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, kg = lane >> 4;
  const int tx_n = W / E0_TW;
  const int y0 = (blockIdx.x / tx_n) * E0_TH, x0 = (blockIdx.x % tx_n) * E0_TW;
  const int fa = blockIdx.y * fg, fb = min(F, fa + fg);
  const size_t HW = (size_t)H * W;
  // A fragments (weights) for the block's lifetime; 3x3 K step s holds tap 4 s + kg in this lane
  bf16x8 As[3][2], At[3][2];
  float bsv[2][4], btv[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int co = 16 * a + r;
#pragma unroll
    for (int st = 0; st < 3; ++st) {
      const int t = 4 * st + kg;
      As[st][a] = t < 9 ? load8(ws + ((size_t)t * L0_C + co) * 8) : zero8();
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) At[d][a] = load8(wt + ((size_t)d * L0_C + co) * L0_C + 8 * kg);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bsv[a][j] = bs[16 * a + 4 * kg + j];
      btv[a][j] = bt[16 * a + 4 * kg + j];
    }
  }
  // halo offsets of this lane's taps 4 s + kg, s = 0, 1, 2
  int toff[3];
  bool tok[3];
#pragma unroll
  for (int st = 0; st < 3; ++st) {
    const int t = 4 * st + kg;
    tok[st] = t < 9;
    toff[st] = tok[st] ? (t / 3) * E0_HX + (t % 3) : 0;
  }
  // halo pixels tid and tid + 256 of a frame, prefetched one frame ahead
  float2 pre[2];
  auto fetch = [&](int f) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int hp = tid + 256 * q;
      pre[q] = make_float2(0.f, 0.f);
      if (hp < E0_NH) {
        const int gy = y0 + hp / E0_HX - 1, gx = x0 + hp % E0_HX - 1;
        if ((unsigned)gy < (unsigned)H && (unsigned)gx < (unsigned)W) pre[q] = x[(size_t)f * HW + (size_t)gy * W + gx];
      }
    }
  };
  u16* ring = sR[wave][0];
  auto temporal = [&](int fo) {   // e_0 and i_1 of frame fo from the ring
    f32x4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int h = 0; h < 4; ++h) acc[a][h] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int f = fo - 1 + d;
      if (f < 0 || f >= F) continue;
      const u16* sl = ring + (f % 3) * 64 * L0_RS;
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const bf16x8 bv = load8(sl + (16 * h + r) * L0_RS + 8 * kg);
#pragma unroll
        for (int a = 0; a < 2; ++a) acc[a][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(At[d][a], bv, acc[a][h], 0, 0, 0);
      }
    }
    // epilogue: bias, ReLU, BF16; the 2 x 2 pool of rows 2w, 2w+1 (columns h and h + 2 of this
    // lane) with the horizontal neighbour in lane r ^ 1, summed in the order of k_pool_fwd
    float ev[2][4][4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int ty = 2 * wave + (h >> 1), tx = 16 * (h & 1) + r;
      u16* dst = e0 + (((size_t)fo * H + y0 + ty) * W + x0 + tx) * L0_C;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int j = 0; j < 4; ++j) ev[a][h][j] = bf2f(f2bf(fmaxf(acc[a][h][j] + btv[a][j], 0.f)));
        *reinterpret_cast<uint2*>(dst + 16 * a + 4 * kg) =
            make_uint2(bf_pair(ev[a][h][0], ev[a][h][1]), bf_pair(ev[a][h][2], ev[a][h][3]));
      }
    }
    const int Ho = H / 2, Wo = W / 2;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int tx = 16 * h + r;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        float pv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float p00 = ev[a][h][j], p10 = ev[a][h + 2][j];
          const float p01 = __shfl_xor(p00, 1, 64), p11 = __shfl_xor(p10, 1, 64);
          float s = p00;
          s += p01;
          s += p10;
          pv[j] = 0.25f * (s + p11);
        }
        if ((r & 1) == 0) {
          u16* dst = i1 + (((size_t)fo * Ho + y0 / 2 + wave) * Wo + (x0 + tx) / 2) * L0_C + 16 * a + 4 * kg;
          *reinterpret_cast<uint2*>(dst) = make_uint2(bf_pair(pv[0], pv[1]), bf_pair(pv[2], pv[3]));
        }
      }
    }
  };
  const int s_lo = max(fa - 1, 0), s_hi = min(fb, F - 1);
  fetch(s_lo);
  for (int fs = s_lo; fs <= s_hi; ++fs) {
    unsigned* xb = sX[fs & 1];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int hp = tid + 256 * q;
      if (hp < E0_NH) xb[hp] = bf_pair(pre[q].x * scale, pre[q].y * scale);
    }
    if (fs < s_hi) fetch(fs + 1);
    __syncthreads();   // the halo tile of frame fs is in; every wave is past the tile of frame fs - 2
    // a_0 of frame fs for this wave's 64 pixels -> ring slot fs % 3
    f32x4 acc[2][4];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int ty = 2 * wave + (h >> 1), tx = 16 * (h & 1) + r;
      const int base = ty * E0_HX + tx;
#pragma unroll
      for (int a = 0; a < 2; ++a) acc[a][h] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int st = 0; st < 3; ++st) {
        // channels 0..7 of tap 4 st + kg: (re, im) then six zeros
        const bf16x8 bv = __builtin_bit_cast(bf16x8, make_uint4(tok[st] ? xb[base + toff[st]] : 0u, 0u, 0u, 0u));
#pragma unroll
        for (int a = 0; a < 2; ++a) acc[a][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(As[st][a], bv, acc[a][h], 0, 0, 0);
      }
    }
    wave_lds_sync();   // the ring slot being replaced (frame fs - 3) has been read
    u16* sl = ring + (fs % 3) * 64 * L0_RS;
#pragma unroll
    for (int h = 0; h < 4; ++h)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = fmaxf(acc[a][h][j] + bsv[a][j], 0.f);
        *reinterpret_cast<uint2*>(sl + (16 * h + r) * L0_RS + 16 * a + 4 * kg) =
            make_uint2(bf_pair(v[0], v[1]), bf_pair(v[2], v[3]));
      }
    wave_lds_sync();
    if (fs - 1 >= fa && fs - 1 < fb) temporal(fs - 1);
  }
  if (s_hi == F - 1 && F - 1 >= fa && F - 1 < fb) temporal(F - 1);
}

// x_out = x_in + (out_0 + i out_1) / scale with out = conv1x1(h_0), h_0 = relu(convT(b_0)), for one
// volume's frames [fg by, fg by + fg): the temporal layer, the 1x1 output layer and the residual
// of k_unpack_residual in one pass.  Each wave owns 64 consecutive pixels of a frame; its three
// b_0 frames are registers (the MFMA B fragment is 16 bytes of one pixel's channels, read
// straight from memory); h_0 goes through a wave-private LDS tile into the 1x1 layer's B
// layout.  Grid (H W / 256, ceil(F/fg)); needs H W % 256 == 0, out channels 0, 1 = (re, im).
__global__ __launch_bounds__(256) void k_dec0_tail(const u16* __restrict__ b0, const u16* __restrict__ wt,
                                                   const float* __restrict__ bt, const u16* __restrict__ wo,
                                                   const float* __restrict__ bo, const float2* __restrict__ xin,
                                                   float2* __restrict__ xout, int F, int HW, int fg,
                                                   float inv_scale) {
  __shared__ __align__(16) u16 sH[4][64 * L0_RS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, kg = lane >> 4;
  const int p0 = blockIdx.x * 256 + 64 * wave;
  const int fa = blockIdx.y * fg, fb = min(F, fa + fg);
  bf16x8 At[3][2], Ao;
  float btv[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
#pragma unroll
    for (int d = 0; d < 3; ++d) At[d][a] = load8(wt + ((size_t)d * L0_C + 16 * a + r) * L0_C + 8 * kg);
#pragma unroll
    for (int j = 0; j < 4; ++j) btv[a][j] = bt[16 * a + 4 * kg + j];
  }
  Ao = load8(wo + (size_t)r * L0_C + 8 * kg);   // output channels 0..15 (rows past Cout are zero)
  const float bo0 = bo[0], bo1 = bo[1];
  auto frame = [&](int f, bf16x8* v) {
#pragma unroll
    for (int h = 0; h < 4; ++h)
      v[h] = (f >= 0 && f < F) ? load8(b0 + ((size_t)f * HW + p0 + 16 * h + r) * L0_C + 8 * kg) : zero8();
  };
  bf16x8 cm[4], c0[4], cp[4], nx[4];
  frame(fa - 1, cm);
  frame(fa, c0);
  frame(fa + 1, cp);
  u16* sh = sH[wave];
  for (int fo = fa; fo < fb; ++fo) {
    frame(fo + 1 < fb ? fo + 2 : F, nx);   // frame F reads as zero: nothing to fetch past the group
    f32x4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        f32x4 s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(At[0][a], cm[h], f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(At[1][a], c0[h], s, 0, 0, 0);
        acc[a][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(At[2][a], cp[h], s, 0, 0, 0);
      }
    wave_lds_sync();   // the previous frame's h_0 tile has been read
#pragma unroll
    for (int h = 0; h < 4; ++h)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = fmaxf(acc[a][h][j] + btv[a][j], 0.f);
        *reinterpret_cast<uint2*>(sh + (16 * h + r) * L0_RS + 16 * a + 4 * kg) =
            make_uint2(bf_pair(v[0], v[1]), bf_pair(v[2], v[3]));
      }
    wave_lds_sync();
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const bf16x8 bv = load8(sh + (16 * h + r) * L0_RS + 8 * kg);
      const f32x4 o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(Ao, bv, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
      if (kg == 0) {   // rows 0, 1 of the output tile: channels re, im of pixel 16 h + r
        const size_t p = (size_t)fo * HW + p0 + 16 * h + r;
        const float2 a = xin[p];
        xout[p] = make_float2(a.x + bf2f(f2bf(o[0] + bo0)) * inv_scale, a.y + bf2f(f2bf(o[1] + bo1)) * inv_scale);
      }
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      cm[h] = c0[h];
      c0[h] = cp[h];
      cp[h] = nx[h];
    }
  }
}

__global__ __launch_bounds__(256) void k_mse_grad(const float2* __restrict__ xout, const float2* __restrict__ xt,
                                                  const float2* __restrict__ gx, float gx_coef,
                                                  u16* __restrict__ gnet, int64_t npix, int Cn, float scale,
                                                  float w, double* __restrict__ loss) {
  double s = 0.0;
  GRID_STRIDE(i, npix) {
    const float2 a = xout[i], b = xt[i];
    const float ex = (a.x - b.x) * scale, ey = (a.y - b.y) * scale;
    s += (double)(ex * ex + ey * ey) * w;
    float gr = 2.f * w * ex, gi = 2.f * w * ey;
    if (gx) {
      const float2 q = gx[i];
      gr = fmaf(gx_coef, q.x, gr);
      gi = fmaf(gx_coef, q.y, gi);
    }
    u16* o = gnet + i * Cn;
    o[0] = f2bf(gr);
    o[1] = f2bf(gi);
    for (int c = 2; c < Cn; ++c) o[c] = 0;
  }
  s = warp_sum_d(s);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(loss, s);
}

__global__ void k_fill_f32(float* __restrict__ p, int64_t n, float v) { GRID_STRIDE(i, n) p[i] = v; }

// Adam on the FP32 master weights [tap][Cout_pad][Cin]; refreshes the BF16
// forward copy (same layout, padded rows 0) and the backward-data copy
// [T-1-tap][Cin][Cout].  lr == 0 only refreshes the copies.
__global__ void k_adam(float* __restrict__ wm, const float* __restrict__ gw, float* __restrict__ m,
                       float* __restrict__ v, u16* __restrict__ wf, u16* __restrict__ wbd, int ntap, int Cout,
                       int Cout_pad, int Cin, float lr, float b1, float b2, float eps, float bc1, float bc2) {
  GRID_STRIDE(i, (int64_t)ntap * Cout_pad * Cin) {
    const int ci = i % Cin;
    const int64_t q = i / Cin;
    const int co = q % Cout_pad;
    const int t = q / Cout_pad;
    float wv = wm[i];
    if (co < Cout && lr > 0.f) {
      const float gg = gw[i];
      const float mm = b1 * m[i] + (1.f - b1) * gg;
      const float vv = b2 * v[i] + (1.f - b2) * gg * gg;
      m[i] = mm;
      v[i] = vv;
      wv -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
      wm[i] = wv;
    }
    const u16 wb = co < Cout ? f2bf(wv) : (u16)0;
    wf[i] = wb;
    if (wbd && co < Cout) wbd[((size_t)(ntap - 1 - t) * Cin + ci) * Cout + co] = wb;
  }
}

__global__ void k_adam_bias(float* __restrict__ b, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float bc1,
                            float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float mm = b1 * m[i] + (1.f - b1) * g[i];
  const float vv = b2 * v[i] + (1.f - b2) * g[i] * g[i];
  m[i] = mm;
  v[i] = vv;
  b[i] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// ------------------------------------------------------------------ C ABI
static unsigned grid_for(int64_t n) {
  return (unsigned)std::min<int64_t>(262144, std::max<int64_t>(1, (n + 255) / 256));
}

static int make_geom(int kind, int F, int H, int W, TapGeom* g) {
  MRH_REQUIRE(kind == MRH_UNET_TAP_1X1 || kind == MRH_UNET_TAP_3X3 || kind == MRH_UNET_TAP_T3, MRH_EINVAL,
              "unet conv: tap kind %d is not 1x1 (0), 3x3 (1) or temporal-3 (2)", kind);
  MRH_REQUIRE(F >= 1 && H >= 1 && W >= 1, MRH_EINVAL, "unet conv: bad geometry F=%d H=%d W=%d", F, H, W);
  g->kind = kind;
  g->ntap = kind == MRH_UNET_TAP_1X1 ? 1 : (kind == MRH_UNET_TAP_3X3 ? 9 : 3);
  g->F = F; g->H = H; g->W = W;
  return MRH_OK;
}

#define MRH_TRY(call) do { int _rc = (call); if (_rc) return _rc; } while (0)

// Temporal layers with Cin = 64, 128 or 256 on k_conv_t3_walk: 8 waves per workgroup (64-cout blocks,
// 32 for Cin = 256 so that the register ring fits), frame groups of at least 5 frames (a group loads
// its first frame - 1 and last frame + 1 again) and about 512 workgroups, two per CU.  Measured on
// the 34 x 256^2 volume against 4 waves and 3-frame groups, 16 waves, and 256 - 2048 workgroups.
// Returns false for shapes it does not cover.
static bool launch_t3_walk(const u16* x, const u16* w, const float* bias, u16* out, int nvol, int F, int HW,
                           int Cin, int Cout, int Cout_pad, int relu, hipStream_t st) {
  constexpr int WPB = 8, TARGET_BLOCKS = 512, FG_MIN = 5;
  int mt, nt;
  if (Cin == 64) { mt = 4; nt = 2; }
  else if (Cin == 128) { mt = 4; nt = 1; }
  else if (Cin == 256) { mt = 2; nt = 1; }
  else return false;
  if (Cout_pad % (16 * mt) != 0 || HW % (16 * nt * WPB) != 0) return false;
  const int npb = HW / (16 * nt * WPB), ncb = Cout_pad / (16 * mt);
  const int64_t per_group = (int64_t)nvol * npb * ncb;
  const int ng_max = std::max(1, (F + FG_MIN - 1) / FG_MIN);
  const int ng = (int)std::max<int64_t>(1, std::min<int64_t>(ng_max, (TARGET_BLOCKS + per_group - 1) / per_group));
  const int fg = (F + ng - 1) / ng, ngroups = (F + fg - 1) / fg;
  const int64_t nblocks = per_group * ngroups;
  if (nblocks > INT32_MAX) return false;
  const dim3 grid((unsigned)nblocks);
#define MRH_T3W(MT, NT, NCH, PD)                                                                               \
  hipLaunchKernelGGL((k_conv_t3_walk<MT, NT, NCH, PD, WPB>), grid, dim3(64 * WPB), 0, st, x, w, bias, out, F, HW, \
                     Cout, Cout_pad, relu, fg, npb, ncb)
  if (Cin == 64) MRH_T3W(4, 2, 2, 2);
  else if (Cin == 128) MRH_T3W(4, 1, 4, 2);
  else MRH_T3W(2, 1, 8, 1);
#undef MRH_T3W
  return true;
}

extern "C" int mrh_unet_conv_fwd(const void* x, const void* w, const float* bias, void* out, int kind, int nvol,
                                 int F, int H, int W, int Cin, int Cout, int Cout_pad, int relu, void* stream) {
  MRH_REQUIRE(x && w && out, MRH_EINVAL, "unet_conv_fwd: null buffer");
  TapGeom g;
  MRH_TRY(make_geom(kind, F, H, W, &g));
  MRH_REQUIRE(Cin >= 8 && Cin % 8 == 0, MRH_EINVAL, "unet_conv_fwd: Cin=%d must be a positive multiple of 8", Cin);
  MRH_REQUIRE(Cout >= 1 && Cout <= Cout_pad && Cout_pad % 32 == 0, MRH_EINVAL,
              "unet_conv_fwd: need 1 <= Cout=%d <= Cout_pad=%d, Cout_pad %% 32 == 0", Cout, Cout_pad);
  const int64_t npix = (int64_t)nvol * F * H * W;
  MRH_REQUIRE(nvol >= 1 && npix < (1ll << 31) - 256, MRH_EINVAL, "unet_conv_fwd: %lld pixels out of range",
              (long long)npix);
  hipStream_t st = (hipStream_t)stream;
  // Tile choice measured per layer shape on MI355X (34 x 256^2 volume, every layer of the network):
  // 3x3 layers with Cin % 32 == 0 on 2-D tiles (4 x 64 or 8 x 32 pixels, 1.7-2.8x the 1-D LDS
  // tiles), temporal layers with >= 64 couts on 256-pixel tiles; 1x1 convolutions, Cin = 8 and
  // 32-cout temporal layers on the register tiles; other shapes on the 1-D LDS tiles.
  const u16 *xs = (const u16*)x, *ws = (const u16*)w;
  u16* os = (u16*)out;
#define MRH_LDS_CONV(MT, NT, KC)                                                                              \
  hipLaunchKernelGGL((k_conv_fwd_lds<MT, NT, KC>),                                                        \
                     dim3((unsigned)((npix + 64 * NT - 1) / (64 * NT)), Cout_pad / (16 * MT)), dim3(256), 0, st, \
                     xs, ws, bias, os, g, (int)npix, Cin, Cout, Cout_pad, relu)
  if (kind == MRH_UNET_TAP_T3 && launch_t3_walk(xs, ws, bias, os, nvol, F, H * W, Cin, Cout, Cout_pad, relu, st)) {
    MRH_CHECK_LAUNCH();
    return MRH_OK;
  }
  const int tw = W >= 64 && W % 64 == 0 ? 64 : (W == 32 ? 32 : 0);
  const bool tile3 = kind == MRH_UNET_TAP_3X3 && tw && H % (256 / tw) == 0;
  // temporal tiles pay off with 64-cout blocks; 32-cout temporal layers stay on the register tiles
  const bool tilet = kind == MRH_UNET_TAP_T3 && (H * W) % 256 == 0 && Cout_pad % 64 == 0;
  if (Cin % 32 == 0 && (tile3 || tilet)) {
    const int mt = Cout_pad % 64 == 0 ? 4 : 2;
    dim3 grid((unsigned)(npix / 256), Cout_pad / (16 * mt));
#define MRH_TILE_CONV(MT, TW, TEMP) \
  hipLaunchKernelGGL((k_conv_tile<MT, TW, TEMP>), grid, dim3(256), 0, st, xs, ws, bias, os, F, H, W, Cin, Cout, Cout_pad, relu)
    if (tilet) MRH_TILE_CONV(4, 256, true);
    else if (mt == 4) { if (tw == 64) MRH_TILE_CONV(4, 64, false); else MRH_TILE_CONV(4, 32, false); }
    else { if (tw == 64) MRH_TILE_CONV(2, 64, false); else MRH_TILE_CONV(2, 32, false); }
#undef MRH_TILE_CONV
  } else if (Cin % 16 != 0 || kind == MRH_UNET_TAP_1X1 || (kind == MRH_UNET_TAP_T3 && Cout_pad == 32)) {
    hipLaunchKernelGGL(k_conv_fwd, dim3((unsigned)((npix + 127) / 128), Cout_pad / 32), dim3(256), 0, st, xs, ws,
                       bias, os, g, (int)npix, Cin, Cout, Cout_pad, relu);
  } else if (Cout_pad % 64 != 0) {
    if (Cin % 32 == 0) MRH_LDS_CONV(2, 2, 32);
    else MRH_LDS_CONV(2, 4, 16);
  } else if (kind == MRH_UNET_TAP_3X3 && Cin >= 128) {
    MRH_LDS_CONV(4, 4, 16);
  } else {
    MRH_LDS_CONV(4, 2, 16);
  }
#undef MRH_LDS_CONV
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

static bool upcat_tiles(int H, int W, int C1, int C2) {
  const int tw = W >= 64 && W % 64 == 0 ? 64 : (W == 32 ? 32 : 0);
  return tw && H % (256 / tw) == 0 && H % 2 == 0 && W % 2 == 0 && C1 % 32 == 0 && C2 % 32 == 0 && C1 > 0 &&
         C2 > 0;
}

extern "C" int mrh_unet_upcat_supported(int H, int W, int C1, int C2) { return upcat_tiles(H, W, C1, C2) ? 1 : 0; }

extern "C" int mrh_unet_conv3x3_upcat(const void* s, const void* skip, const void* w, const float* bias, void* out,
                                      int nvol, int F, int H, int W, int C1, int C2, int Cout, int Cout_pad,
                                      int relu, void* stream) {
  MRH_REQUIRE(s && skip && w && out, MRH_EINVAL, "unet_conv3x3_upcat: null buffer");
  TapGeom g;
  MRH_TRY(make_geom(MRH_UNET_TAP_3X3, F, H, W, &g));
  MRH_REQUIRE(upcat_tiles(H, W, C1, C2), MRH_EUNSUP,
              "unet_conv3x3_upcat: %dx%d image with %d + %d channels has no tile (mrh_unet_upcat_supported)", H, W,
              C1, C2);
  MRH_REQUIRE(Cout >= 1 && Cout <= Cout_pad && Cout_pad % 32 == 0, MRH_EINVAL,
              "unet_conv3x3_upcat: need 1 <= Cout=%d <= Cout_pad=%d, Cout_pad %% 32 == 0", Cout, Cout_pad);
  const int64_t npix = (int64_t)nvol * F * H * W;
  MRH_REQUIRE(nvol >= 1 && npix < (1ll << 31) - 256, MRH_EINVAL, "unet_conv3x3_upcat: %lld pixels out of range",
              (long long)npix);
  hipStream_t st = (hipStream_t)stream;
  const u16 *xs = (const u16*)s, *x2 = (const u16*)skip, *ws = (const u16*)w;
  u16* os = (u16*)out;
  const int tw = W >= 64 ? 64 : 32, mt = Cout_pad % 64 == 0 ? 4 : 2, Cin = C1 + C2;
  dim3 grid((unsigned)(npix / 256), Cout_pad / (16 * mt));
#define MRH_UPC_CONV(MT, TW)                                                                                   \
  hipLaunchKernelGGL((k_conv_tile<MT, TW, false, true>), grid, dim3(256), 0, st, xs, ws, bias, os, F, H, W, Cin, \
                     Cout, Cout_pad, relu, x2, C1)
  if (mt == 4) { if (tw == 64) MRH_UPC_CONV(4, 64); else MRH_UPC_CONV(4, 32); }
  else { if (tw == 64) MRH_UPC_CONV(2, 64); else MRH_UPC_CONV(2, 32); }
#undef MRH_UPC_CONV
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_conv_bwd_weight(const void* gout, const void* x, float* dw, float* db, int kind, int nvol,
                                        int F, int H, int W, int Cin, int Cout, int Cout_pad, void* stream) {
  MRH_REQUIRE(gout && x && dw, MRH_EINVAL, "unet_conv_bwd_weight: null buffer");
  TapGeom g;
  MRH_TRY(make_geom(kind, F, H, W, &g));
  MRH_REQUIRE(Cin >= 8 && Cin % 8 == 0, MRH_EINVAL, "unet_conv_bwd_weight: Cin=%d must be a multiple of 8", Cin);
  MRH_REQUIRE(Cout >= 1 && Cout <= Cout_pad && Cout_pad % 32 == 0 && Cout <= 256, MRH_EINVAL,
              "unet_conv_bwd_weight: need 1 <= Cout=%d <= min(Cout_pad=%d, 256), Cout_pad %% 32 == 0", Cout,
              Cout_pad);
  const int64_t npix = (int64_t)nvol * F * H * W;
  MRH_REQUIRE(nvol >= 1 && npix < (1ll << 31) - 8192, MRH_EINVAL, "unet_conv_bwd_weight: %lld pixels out of range",
              (long long)npix);
  hipStream_t st = (hipStream_t)stream;
  // split-K over pixels: about 2048 blocks in total, at least one 128-pixel step each
  const int64_t tiles = (int64_t)(Cout_pad / 32) * g.ntap * ((Cin + 31) / 32);
  int64_t nchunks = std::max<int64_t>(1, (2048 + tiles - 1) / tiles);
  int64_t chunk = (npix + nchunks - 1) / nchunks;
  chunk = std::max<int64_t>(BW_STEP, (chunk + BW_STEP - 1) / BW_STEP * BW_STEP);
  nchunks = (npix + chunk - 1) / chunk;
  dim3 grid((unsigned)nchunks, Cout_pad / 32, g.ntap * ((Cin + 31) / 32));
  hipLaunchKernelGGL(k_conv_bw, grid, dim3(256), 0, st, (const u16*)gout, (const u16*)x, dw, g, (int)npix,
                     (int)chunk, Cin, Cout, Cout_pad);
  MRH_CHECK_LAUNCH();
  if (db) {
    hipLaunchKernelGGL(k_bias_grad, dim3((unsigned)((npix + 4095) / 4096)), dim3(256), 0, st, (const u16*)gout, db,
                       (int)npix, Cout);
    MRH_CHECK_LAUNCH();
  }
  return MRH_OK;
}

extern "C" int mrh_unet_relu_bwd(void* g, const void* y, int64_t n, void* stream) {
  MRH_REQUIRE(g && y && n >= 0 && n % 8 == 0, MRH_EINVAL, "unet_relu_bwd: n must be a multiple of 8");
  if (n == 0) return MRH_OK;
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_for(n / 8)), dim3(256), 0, (hipStream_t)stream, (u16*)g, (const u16*)y,
                     n / 8);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_pool2(const void* x, void* y, int nf, int H, int W, int C, void* stream) {
  MRH_REQUIRE(x && y && nf >= 1 && C >= 8 && C % 8 == 0 && H >= 2 && W >= 2 && H % 2 == 0 && W % 2 == 0,
              MRH_EINVAL, "unet_pool2: need even H=%d, W=%d and C=%d a multiple of 8", H, W, C);
  hipLaunchKernelGGL(k_pool_fwd, dim3(grid_for((int64_t)nf * H * W * C / 32)), dim3(256), 0, (hipStream_t)stream,
                     (const u16*)x, (u16*)y, nf, H, W, C);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_pool2_bwd(const void* gy, const void* add, void* gx, int nf, int H, int W, int C,
                                  void* stream) {
  MRH_REQUIRE(gy && gx && nf >= 1 && C >= 8 && C % 8 == 0 && H % 2 == 0 && W % 2 == 0 && H >= 2 && W >= 2,
              MRH_EINVAL, "unet_pool2_bwd: need even H=%d, W=%d and C=%d a multiple of 8", H, W, C);
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for((int64_t)nf * H * W * C / 8)), dim3(256), 0, (hipStream_t)stream,
                     (const u16*)gy, (const u16*)add, (u16*)gx, nf, H, W, C);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_up2_concat(const void* src, const void* skip, void* dst, int nf, int H, int W, int C1,
                                   int C2, void* stream) {
  MRH_REQUIRE(src && skip && dst && nf >= 1 && C1 >= 8 && C2 >= 8 && C1 % 8 == 0 && C2 % 8 == 0 && H % 2 == 0 &&
                  W % 2 == 0 && H >= 2 && W >= 2,
              MRH_EINVAL, "unet_up2_concat: bad arguments (H=%d W=%d C1=%d C2=%d)", H, W, C1, C2);
  hipLaunchKernelGGL(k_up_concat, dim3(grid_for((int64_t)nf * H * W * (C1 + C2) / 8)), dim3(256), 0,
                     (hipStream_t)stream, (const u16*)src, (const u16*)skip, (u16*)dst, nf, H, W, C1, C2);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_up2_concat_bwd(const void* gdst, void* gsrc, void* gskip, int nf, int H, int W, int C1,
                                       int C2, void* stream) {
  MRH_REQUIRE(gdst && gsrc && gskip && nf >= 1 && C1 >= 8 && C2 >= 8 && C1 % 8 == 0 && C2 % 8 == 0 && H % 2 == 0 &&
                  W % 2 == 0 && H >= 2 && W >= 2,
              MRH_EINVAL, "unet_up2_concat_bwd: bad arguments (H=%d W=%d C1=%d C2=%d)", H, W, C1, C2);
  const int64_t n = ((int64_t)nf * (H / 2) * (W / 2) * C1 + (int64_t)nf * H * W * C2) / 8;
  hipLaunchKernelGGL(k_up_concat_bwd, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (const u16*)gdst,
                     (u16*)gsrc, (u16*)gskip, nf, H, W, C1, C2);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_pack_input(const void* img, void* out, int64_t npix, int C, float scale, void* stream) {
  MRH_REQUIRE(img && out && npix >= 1 && C >= 2, MRH_EINVAL, "unet_pack_input: bad arguments");
  hipLaunchKernelGGL(k_pack_input, dim3(grid_for(npix)), dim3(256), 0, (hipStream_t)stream, (const float2*)img,
                     (u16*)out, npix, C, scale);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_unpack_residual(const void* xin, const void* net, void* xout, int64_t npix, int Cn,
                                        float inv_scale, void* stream) {
  MRH_REQUIRE(xin && net && xout && npix >= 1 && Cn >= 2, MRH_EINVAL, "unet_unpack_residual: bad arguments");
  hipLaunchKernelGGL(k_unpack_residual, dim3(grid_for(npix)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)xin, (const u16*)net, (float2*)xout, npix, Cn, inv_scale);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

// Frame groups per volume: enough workgroups for two per CU (a group re-reads its two edge frames).
static int level0_groups(int tiles, int F) { return std::max(1, std::min(F, (512 + tiles - 1) / tiles)); }

extern "C" int mrh_unet_level0_fused_supported(int H, int W) {
  return (H >= 8 && W >= 32 && H % 8 == 0 && W % 32 == 0 && ((int64_t)H * W) % 256 == 0) ? 1 : 0;
}

extern "C" int mrh_unet_enc0_fused(const void* x, float scale, const void* ws, const float* bs, const void* wt,
                                   const float* bt, void* e0, void* i1, int F, int H, int W, void* stream) {
  MRH_REQUIRE(x && ws && bs && wt && bt && e0 && i1, MRH_EINVAL, "unet_enc0_fused: null buffer");
  MRH_REQUIRE(F >= 1 && mrh_unet_level0_fused_supported(H, W), MRH_EUNSUP,
              "unet_enc0_fused: %d frames of %dx%d (needs H %% 8 == 0, W %% 32 == 0)", F, H, W);
  MRH_REQUIRE((int64_t)F * H * W < (1ll << 31), MRH_EINVAL, "unet_enc0_fused: volume too large");
  const int tiles = (H / E0_TH) * (W / E0_TW), ng = level0_groups(tiles, F), fg = (F + ng - 1) / ng;
  hipLaunchKernelGGL(k_enc0_fused, dim3(tiles, (F + fg - 1) / fg), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)x, (const u16*)ws, bs, (const u16*)wt, bt, (u16*)e0, (u16*)i1, F, H, W, fg,
                     scale);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_dec0_tail(const void* b0, const void* wt, const float* bt, const void* wo, const float* bo,
                                  const void* xin, float inv_scale, void* xout, int F, int H, int W, void* stream) {
  MRH_REQUIRE(b0 && wt && bt && wo && bo && xin && xout, MRH_EINVAL, "unet_dec0_tail: null buffer");
  MRH_REQUIRE(F >= 1 && mrh_unet_level0_fused_supported(H, W), MRH_EUNSUP,
              "unet_dec0_tail: %d frames of %dx%d (needs H %% 8 == 0, W %% 32 == 0)", F, H, W);
  MRH_REQUIRE((int64_t)F * H * W < (1ll << 31), MRH_EINVAL, "unet_dec0_tail: volume too large");
  const int HW = H * W, tiles = HW / 256, ng = level0_groups(tiles, F), fg = (F + ng - 1) / ng;
  hipLaunchKernelGGL(k_dec0_tail, dim3(tiles, (F + fg - 1) / fg), dim3(256), 0, (hipStream_t)stream,
                     (const u16*)b0, (const u16*)wt, bt, (const u16*)wo, bo, (const float2*)xin, (float2*)xout, F,
                     HW, fg, inv_scale);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_mse_grad(const void* xout, const void* xt, const void* gx, float gx_coef, void* gnet,
                                 double* loss, int64_t npix, int Cn, float scale, float weight, void* stream) {
  MRH_REQUIRE(xout && xt && gnet && loss && npix >= 1 && Cn >= 2, MRH_EINVAL, "unet_mse_grad: bad arguments");
  hipLaunchKernelGGL(k_mse_grad, dim3(std::min<unsigned>(grid_for(npix), 2048)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)xout, (const float2*)xt, (const float2*)gx, gx_coef, (u16*)gnet, npix, Cn,
                     scale, weight, loss);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_fill(float* p, int64_t n, float value, void* stream) {
  MRH_REQUIRE(p && n >= 0, MRH_EINVAL, "unet_fill: bad arguments");
  if (n == 0) return MRH_OK;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, p, n, value);
  MRH_CHECK_LAUNCH();
  return MRH_OK;
}

extern "C" int mrh_unet_adam(float* wm, const float* gw, float* m, float* v, void* wf, void* wbd, float* bias,
                             const float* gb, float* mb, float* vb, int ntap, int Cout, int Cout_pad, int Cin,
                             float lr, float beta1, float beta2, float eps, int step, void* stream) {
  MRH_REQUIRE(wm && wf && ntap >= 1 && Cin >= 1 && Cout >= 1 && Cout <= Cout_pad, MRH_EINVAL,
              "unet_adam: bad arguments");
  MRH_REQUIRE(lr == 0.f || (gw && m && v && step >= 1), MRH_EINVAL,
              "unet_adam: an update (lr > 0) needs gradients, moments and step >= 1");
  hipStream_t st = (hipStream_t)stream;
  const float bc1 = step >= 1 ? 1.f - powf(beta1, (float)step) : 1.f;
  const float bc2 = step >= 1 ? 1.f - powf(beta2, (float)step) : 1.f;
  hipLaunchKernelGGL(k_adam, dim3(grid_for((int64_t)ntap * Cout_pad * Cin)), dim3(256), 0, st, wm, gw, m, v,
                     (u16*)wf, (u16*)wbd, ntap, Cout, Cout_pad, Cin, lr, beta1, beta2, eps, bc1, bc2);
  MRH_CHECK_LAUNCH();
  if (lr != 0.f && bias && gb && mb && vb) {
    hipLaunchKernelGGL(k_adam_bias, dim3((Cout + 255) / 256), dim3(256), 0, st, bias, gb, mb, vb, Cout, lr, beta1,
                       beta2, eps, bc1, bc2);
    MRH_CHECK_LAUNCH();
  }
  return MRH_OK;
}
