// ub_unet.hip -- the learned denoiser's kernels and host driver (see
// ub_unet.h for the network). Convolutions are implicit GEMMs on the gfx950
// MFMA unit (v_mfma_f32_16x16x32_bf16: a wave multiplies a 16 x 32 bf16 tile
// by a 32 x 16 one into a 16 x 16 fp32 accumulator):
//   forward / input gradient: rows = 16 pixels of one image row, k = 32
//     input channels of one 3x3 tap, columns = 16 output channels;
//   weight gradient: rows = 16 output channels, k = 32 pixels, columns = 16
//     input channels, split over pixel chunks whose partial sums are added in
//     chunk order, so every gradient is the same on every run.
// Operand lane layout (checked on the device): lane l holds row l % 16 of A
// and column l % 16 of B with k = 8 (l / 16) .. + 7; the result has column
// l % 16 and rows 4 (l / 16) .. + 3.
#include "ub_unet.h"

#include <algorithm>
#include <cmath>
#include <random>

namespace ubg {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }
// round to nearest even (finite values)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

// A convolution operand: a bf16 NHWC tensor of C channels and H x W pixels,
// read directly (mode 0), 2x2-averaged (1: its pixels are twice the
// convolution's) or nearest-upsampled (2: half the convolution's); mode 3:
// the fp32 input features (kUnetFeatures per pixel at p) read as the bf16
// input padded with zeros to C channels.
struct SrcDev {
  const uint16_t* p;
  int C, mode, H, W;
};

__device__ __forceinline__ uint4 avg4(uint4 a, uint4 b, uint4 c, uint4 d) {
  const uint32_t* pa = reinterpret_cast<const uint32_t*>(&a);
  const uint32_t* pb = reinterpret_cast<const uint32_t*>(&b);
  const uint32_t* pc = reinterpret_cast<const uint32_t*>(&c);
  const uint32_t* pd = reinterpret_cast<const uint32_t*>(&d);
  uint4 r;
  uint32_t* pr = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t o = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int sh = 16 * h;
      const float s = ((bf2f(static_cast<uint16_t>(pa[i] >> sh)) + bf2f(static_cast<uint16_t>(pb[i] >> sh))) +
                       bf2f(static_cast<uint16_t>(pc[i] >> sh))) +
                      bf2f(static_cast<uint16_t>(pd[i] >> sh));
      o |= static_cast<uint32_t>(f2bf(s * 0.25f)) << sh;
    }
    pr[i] = o;
  }
  return r;
}

// 8 consecutive channels c .. c + 7 of the operand at (n, y, x) in the
// convolution's pixel grid
__device__ __forceinline__ uint4 load8(const SrcDev& s, int n, int y, int x, int c) {
  if (s.mode == 3) {  // channels c .. c + 7 of 12 (c = 0, 8, 16, 24); 8-byte aligned pairs
    static_assert(kUnetFeatures == 12, "the feature loads below read 12 channels");
    const float2* f = reinterpret_cast<const float2*>(reinterpret_cast<const float*>(s.p) +
                                                      ((static_cast<size_t>(n) * s.H + y) * s.W + x) * kUnetFeatures);
    if (c >= 16) return make_uint4(0u, 0u, 0u, 0u);
    if (c == 8) {
      const float2 a = f[4], b = f[5];
      return make_uint4(f2bf(a.x) | static_cast<uint32_t>(f2bf(a.y)) << 16,
                        f2bf(b.x) | static_cast<uint32_t>(f2bf(b.y)) << 16, 0u, 0u);
    }
    const float2 a = f[0], b = f[1], d = f[2], e = f[3];
    return make_uint4(f2bf(a.x) | static_cast<uint32_t>(f2bf(a.y)) << 16, f2bf(b.x) | static_cast<uint32_t>(f2bf(b.y)) << 16,
                      f2bf(d.x) | static_cast<uint32_t>(f2bf(d.y)) << 16, f2bf(e.x) | static_cast<uint32_t>(f2bf(e.y)) << 16);
  }
  if (s.mode == 0)
    return *reinterpret_cast<const uint4*>(s.p + ((static_cast<size_t>(n) * s.H + y) * s.W + x) * s.C + c);
  if (s.mode == 2)
    return *reinterpret_cast<const uint4*>(s.p + ((static_cast<size_t>(n) * s.H + (y >> 1)) * s.W + (x >> 1)) * s.C + c);
  const uint16_t* b = s.p + ((static_cast<size_t>(n) * s.H + 2 * y) * s.W + 2 * x) * s.C + c;
  const size_t rowStride = static_cast<size_t>(s.W) * s.C;
  return avg4(*reinterpret_cast<const uint4*>(b), *reinterpret_cast<const uint4*>(b + s.C),
              *reinterpret_cast<const uint4*>(b + rowStride), *reinterpret_cast<const uint4*>(b + rowStride + s.C));
}
// one channel of the same
__device__ __forceinline__ uint16_t load1(const SrcDev& s, int n, int y, int x, int c) {
  if (s.mode == 0) return s.p[((static_cast<size_t>(n) * s.H + y) * s.W + x) * s.C + c];
  if (s.mode == 2) return s.p[((static_cast<size_t>(n) * s.H + (y >> 1)) * s.W + (x >> 1)) * s.C + c];
  const uint16_t* b = s.p + ((static_cast<size_t>(n) * s.H + 2 * y) * s.W + 2 * x) * s.C + c;
  const size_t rowStride = static_cast<size_t>(s.W) * s.C;
  const float sum = ((bf2f(b[0]) + bf2f(b[s.C])) + bf2f(b[rowStride])) + bf2f(b[rowStride + s.C]);
  return f2bf(sum * 0.25f);
}

// Direct global -> LDS loads (global_load_lds_dwordx4: lane l of the wave
// writes 16 bytes at the wave-uniform LDS base + 16 l) of a unit array:
// dst[u] for u = threadIdx.x + 256 k < count, with src(u) the global
// address of unit u, or zero where src(u) is null. Nothing waits for a
// register, so every load of the staging is in flight at once. The LDS
// writes of these loads complete asynchronously and are counted by vmcnt, so
// a workgroup barrier that publishes them must be preceded by dmaDone().
template <int COUNT, typename Src>
__device__ __forceinline__ void stageDma(uint4* dst, Src src) {
  const int wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < (COUNT + 255) / 256; ++k) {
    const int u = static_cast<int>(threadIdx.x) + 256 * k;
    if (u < COUNT) {
      const uint16_t* g = src(u);
      if (g)
        __builtin_amdgcn_global_load_lds(g, dst + 256 * k + 64 * wv, 16, 0, 0);
      else
        dst[u] = make_uint4(0u, 0u, 0u, 0u);
    }
  }
}

// Wait until every direct-to-LDS load this wave issued has written LDS
// (vmcnt 0; expcnt and lgkmcnt left at their no-wait maxima, gfx9 encoding).
// A plain workgroup barrier does not imply it: the memory model's release
// fence at a barrier waits on lgkmcnt only.
__device__ __forceinline__ void dmaDone() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// 3x3 convolution, zero padding: out[p, co] = sum over taps and input
// channels of in[p + tap, ci] * w[tap][co][ci] (flip: w[8 - tap]), plus bias
// and ReLU when asked; output bf16 or fp32. Input channels [0, s0.C) come from
// s0, the rest from s1.
// A workgroup (4 waves) covers kConvRows consecutive rows x 16 NP pixels of
// one image and the NT * 16 output channels from coBase; wave w takes row
// y0 + w. Per 32 input channels, the workgroup stages in LDS the patch those
// rows read ((rows + 2) x (16 NP + 2) pixels, pooling / upsampling applied,
// zero outside the image) and, one tap row (3 taps) at a time, the weight
// fragments of those taps; every wave then runs the taps on them, each
// weight fragment feeding NP matrix products. A bf16 result goes back
// through LDS so each lane stores 16 consecutive bytes.
constexpr int kConvRows = 4;
// Tap rows of weights staged at once: all three where patch + weights stay
// within 53 KB (3 workgroups per CU), else one.
constexpr int convTapRows(int nt, int np) {
  return (kConvRows + 2) * (16 * np + 2) * 64 + 9 * 16 * nt * 64 <= 53 * 1024 ? 3 : 1;
}
// LDS of one instantiation: patch + staged weights, or the bf16 output
// tiles, whichever is larger
constexpr int convLdsBytes(int nt, int np) {
  return (kConvRows + 2) * (16 * np + 2) * 64 + 3 * convTapRows(nt, np) * 16 * nt * 64 > kConvRows * 16 * np * nt * 32
             ? (kConvRows + 2) * (16 * np + 2) * 64 + 3 * convTapRows(nt, np) * 16 * nt * 64
             : kConvRows * 16 * np * nt * 32;
}

// LDS slot of 16-byte unit k (8 channels) of row r of a staged array (a patch
// pixel, a weight row): 4 r + (k ^ swz(r)); slot u holds unit
// (u & 3) ^ swz(u >> 2) of row u >> 2. An operand read (ds_read_b128) has
// 16 lanes read rows r .. r + 15 at one k, served 8 lanes (128 bytes, 32
// banks) at a time; unswizzled, rows r and r + 2 fall on the same banks
// (4-way conflicts), with swz(r) = (r >> 1) & 3 the 8 rows of a pass cover
// all 32 banks from any first row (measured: the network 3.6 % faster;
// (r >> 2) & 3, which would suit 16-lane passes, gained nothing).
__device__ __forceinline__ int swz(int r) { return (r >> 1) & 3; }

template <int NT, int NP>
__global__ __launch_bounds__(256) void k_conv3x3(SrcDev s0, SrcDev s1, int N, int H, int W, int Cin, int Cout,
                                                 int coBase, const uint16_t* __restrict__ w, int flip,
                                                 const float* __restrict__ bias, int relu,
                                                 uint16_t* __restrict__ outBf, float* __restrict__ outF,
                                                 const float* __restrict__ resid, float* __restrict__ pred) {
  constexpr int S = 16 * NP;        // pixels per row of the tile
  constexpr int PW = S + 2;         // patch width
  constexpr int PU = (kConvRows + 2) * PW * 4;  // patch size in 16-byte units (32 channels)
  constexpr int TR = convTapRows(NT, NP);       // tap rows per weight stage
  constexpr int WU = 3 * TR * 16 * NT * 4;      // one weight stage, 16-byte units
  constexpr int kLdsBytes = convLdsBytes(NT, NP);
  static_assert((PU + WU) * 16 <= kLdsBytes && kConvRows * S * NT * 16 * 2 <= kLdsBytes, "LDS buffer too small");
  __shared__ uint4 lds[kLdsBytes / 16];
  uint4* wl = lds + PU;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tilesX = (W + S - 1) / S, tilesY = (H + kConvRows - 1) / kConvRows;
  const int tx = blockIdx.x % tilesX, rest = blockIdx.x / tilesX;
  const int ty = rest % tilesY, n = rest / tilesY;
  const int x0 = tx * S, y0 = ty * kConvRows, y = y0 + wv;
  const int row = lane & 15, grp = lane >> 4;
  f32x4 acc[NP][NT];
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[q][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
    const bool first = ci0 < s0.C;
    const SrcDev& s = first ? s0 : s1;
    const int cBase = first ? ci0 : ci0 - s0.C;
    if (ci0 > 0) __syncthreads();  // the previous chunk's patch and weights have been read
    if (s.mode == 0 || s.mode == 2) {  // plain loads: straight into LDS
      const int sh = s.mode == 2 ? 1 : 0;
      stageDma<PU>(lds, [&](int u) -> const uint16_t* {
        const int pix = u >> 2, c8 = (u & 3) ^ swz(pix);
        const int px = pix % PW, pr = pix / PW;
        const int yy = y0 + pr - 1, xx = x0 + px - 1;
        if (yy < 0 || yy >= H || xx < 0 || xx >= W) return nullptr;
        return s.p + ((static_cast<size_t>(n) * s.H + (yy >> sh)) * s.W + (xx >> sh)) * s.C + cBase + 8 * c8;
      });
    } else {  // pooled or converted: through registers
      for (int u = threadIdx.x; u < PU; u += 256) {
        const int pix = u >> 2, c8 = (u & 3) ^ swz(pix);
        const int px = pix % PW, pr = pix / PW;
        const int yy = y0 + pr - 1, xx = x0 + px - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = load8(s, n, yy, xx, cBase + 8 * c8);
        lds[u] = v;
      }
    }
#pragma unroll
    for (int d0 = 0; d0 < 3; d0 += TR) {
      if (d0 > 0) __syncthreads();  // the previous stage's weights have been read
      // wl[4 (t * 16 NT + c) + (k8 ^ swz(t * 16 NT + c))]
      //   = w[tap 3 d0 + t (flipped)][coBase + c][ci0 + 8 k8 ..]
      stageDma<WU>(wl, [&](int u) -> const uint16_t* {
        const int v = u >> 2, k8 = (u & 3) ^ swz(v), c = v % (16 * NT), t = v / (16 * NT);
        const int tap = 3 * d0 + t;
        return w + (static_cast<size_t>(flip ? 8 - tap : tap) * Cout + coBase + c) * Cin + ci0 + 8 * k8;
      });
      dmaDone();  // this wave's patch and weight loads have landed in LDS
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 3 * TR; ++t) {
        const int dy = d0 + t / 3, dx = t % 3;
        bf16x8 a[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          const int pr = (wv + dy) * PW + 16 * q + row + dx;
          a[q] = __builtin_bit_cast(bf16x8, lds[4 * pr + (grp ^ swz(pr))]);
        }
#pragma unroll
        for (int j = 0; j < NT; ++j) {
          const int wr = t * 16 * NT + 16 * j + row;
          const bf16x8 b = __builtin_bit_cast(bf16x8, wl[4 * wr + (grp ^ swz(wr))]);
#pragma unroll
          for (int q = 0; q < NP; ++q) acc[q][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b, acc[q][j], 0, 0, 0);
        }
      }
    }
  }
  if (pred) {  // the prediction: log(1 + E) of the input + output channels 0..2
    if (y >= H || coBase != 0 || row >= 3) return;
    const float bb = bias ? bias[row] : 0.0f;
    float rv[NP][4];
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = min(x0 + 16 * q + 4 * grp + r, W - 1);
        rv[q][r] = resid[((static_cast<size_t>(n) * H + y) * W + x) * kUnetFeatures + row];
      }
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int x = x0 + 16 * q + 4 * grp + r;
        float v = acc[q][0][r] + bb;
        if (relu) v = fmaxf(v, 0.0f);
        if (x < W) pred[((static_cast<size_t>(n) * H + y) * W + x) * 3 + row] = v + rv[q][r];
      }
    return;
  }
  if (outF) {
    if (y >= H) return;
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int co = coBase + 16 * j + row;
        const float bb = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int x = x0 + 16 * q + 4 * grp + r;
          if (x >= W) continue;
          float v = acc[q][j][r] + bb;
          if (relu) v = fmaxf(v, 0.0f);
          outF[((static_cast<size_t>(n) * H + y) * W + x) * Cout + co] = v;
        }
      }
    return;
  }
  // bf16: the wave's tile [S pixels][16 NT channels] through LDS
  __syncthreads();  // every wave is done with the patch
  uint16_t* tile = reinterpret_cast<uint16_t*>(lds) + wv * (S * NT * 16);
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const float bb = bias ? bias[coBase + 16 * j + row] : 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float v = acc[q][j][r] + bb;
        if (relu) v = fmaxf(v, 0.0f);
        tile[(16 * q + 4 * grp + r) * (NT * 16) + 16 * j + row] = f2bf(v);
      }
    }
  __syncthreads();
  if (y >= H) return;
  const uint4* t4 = reinterpret_cast<const uint4*>(tile);
  for (int u = lane; u < S * NT * 2; u += 64) {
    const int px = u / (NT * 2), c8 = u % (NT * 2);
    const int x = x0 + px;
    if (x < W)
      *reinterpret_cast<uint4*>(outBf + ((static_cast<size_t>(n) * H + y) * W + x) * Cout + coBase + 8 * c8) = t4[u];
  }
}

// Weight gradient partials: part[g][tap][co][ci] = sum over the pixels p of
// chunk g of dz[p, co] * in[p + tap, ci]. One wave = one (chunk, tap, 16
// output channels, 16 input channels).
__global__ __launch_bounds__(256) void k_conv_dw(const uint16_t* __restrict__ dz, int Cout, SrcDev s0, SrcDev s1,
                                                 int Cin, int N, int H, int W, int chunk, int nChunks,
                                                 float* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  int wave = static_cast<int>((static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
  const int ciT = Cin / 16, coT = Cout / 16;
  if (wave >= nChunks * 9 * coT * ciT) return;
  const int cit = wave % ciT;
  wave /= ciT;
  const int cot = wave % coT;
  wave /= coT;
  const int tap = wave % 9, g = wave / 9;
  const int row = lane & 15, grp = lane >> 4;
  const int co = cot * 16 + row, ci = cit * 16 + row;
  const bool first = cit * 16 < s0.C;
  const SrcDev& s = first ? s0 : s1;
  const int cl = first ? ci : ci - s0.C;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int64_t P = static_cast<int64_t>(N) * H * W;
  const int64_t pBeg = static_cast<int64_t>(g) * chunk;
  const int64_t pEnd = std::min<int64_t>(P, pBeg + chunk);
  f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int64_t p0 = pBeg; p0 < pEnd; p0 += 32) {
    bf16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t p = p0 + 8 * grp + j;
      uint16_t av = 0, bv = 0;
      if (p < pEnd) {
        av = dz[p * Cout + co];
        const int x = static_cast<int>(p % W);
        const int64_t q = p / W;
        const int y = static_cast<int>(q % H), n = static_cast<int>(q / H);
        const int yy = y + dy, xx = x + dx;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) bv = load1(s, n, yy, xx, cl);
      }
      a[j] = __builtin_bit_cast(__bf16, av);
      b[j] = __builtin_bit_cast(__bf16, bv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  // result column = input channel ci, rows = output channels
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int cor = cot * 16 + 4 * grp + r;
    part[((static_cast<int64_t>(g) * 9 + tap) * Cout + cor) * Cin + ci] = acc[r];
  }
}

// out[i] = sum over g of part[g][i], in chunk order
__global__ void k_reduce_parts(const float* __restrict__ part, int64_t n, int nChunks, float* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int g = 0; g < nChunks; ++g) s += part[static_cast<int64_t>(g) * n + i];
  out[i] = s;
}

// bias gradient partials: part[g][c] = sum over the pixels of chunk g of the
// masked output gradient (fp32)
__global__ void k_bias_part(const float* __restrict__ dy, const uint16_t* __restrict__ y, int relu, int C,
                            int64_t P, int64_t chunk, float* __restrict__ part) {
  const int c = threadIdx.x;
  const int g = blockIdx.x;
  if (c >= C) return;
  const int64_t pBeg = g * chunk, pEnd = std::min<int64_t>(P, pBeg + chunk);
  float s = 0.0f;
  for (int64_t p = pBeg; p < pEnd; ++p) {
    const float d = dy[p * C + c];
    if (!relu || bf2f(y[p * C + c]) > 0.0f) s += d;
  }
  part[static_cast<int64_t>(g) * C + c] = s;
}

// dz = bf16(dy * relu'(y))
__global__ void k_mask_grad(int64_t n, const float* __restrict__ dy, const uint16_t* __restrict__ y, int relu,
                            uint16_t* __restrict__ dz) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = (!relu || bf2f(y[i]) > 0.0f) ? dy[i] : 0.0f;
  dz[i] = f2bf(d);
}

// The gradient of a convolution's operand back onto its source tensor:
// dS += dx[.., off .. off + Cs) through the operand's read (direct, 2x2
// average, nearest upsample).
__global__ void k_scatter(const float* __restrict__ dx, int Cin, int off, int N, int H, int W, int mode,
                          float* __restrict__ dS, int Cs, int Hs, int Ws) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t total = static_cast<int64_t>(N) * Hs * Ws * Cs;
  if (i >= total) return;
  const int c = static_cast<int>(i % Cs);
  const int64_t q = i / Cs;
  const int xs = static_cast<int>(q % Ws);
  const int64_t q2 = q / Ws;
  const int ys = static_cast<int>(q2 % Hs), n = static_cast<int>(q2 / Hs);
  auto at = [&](int y, int x) { return dx[((static_cast<int64_t>(n) * H + y) * W + x) * Cin + off + c]; };
  float v;
  if (mode == 0) v = at(ys, xs);
  else if (mode == 1) v = 0.25f * at(ys >> 1, xs >> 1);
  else v = ((at(2 * ys, 2 * xs) + at(2 * ys, 2 * xs + 1)) + at(2 * ys + 1, 2 * xs)) + at(2 * ys + 1, 2 * xs + 1);
  dS[i] += v;
}

// input features (fp32, kUnetFeatures per pixel) -> bf16, padded to kUnetInPad
__global__ void k_unet_input(int64_t P, const float* __restrict__ feat, uint16_t* __restrict__ act) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= P * kUnetInPad) return;
  const int64_t p = i / kUnetInPad;
  const int c = static_cast<int>(i % kUnetInPad);
  act[i] = c < kUnetFeatures ? f2bf(feat[p * kUnetFeatures + c]) : static_cast<uint16_t>(0);
}

// Loss: per-block partial sums of |pred - target| (l2 == 0: mean L1) or
// (pred - target)^2 (l2 == 1: mean squared error) and the output gradient
// sign(pred - target) / count or 2 (pred - target) / count on the 3 used
// channels
__global__ void k_unet_loss(int64_t P, const float* __restrict__ outF, const float* __restrict__ feat,
                            const float* __restrict__ target, float invCount, int l2, float* __restrict__ dOut,
                            float* __restrict__ lossPart) {
  __shared__ float red[256];
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < P * kUnetOutPad) {
    const int64_t p = i / kUnetOutPad;
    const int c = static_cast<int>(i % kUnetOutPad);
    float g = 0.0f;
    if (c < 3) {
      const float d = (outF[i] + feat[p * kUnetFeatures + c]) - target[p * 3 + c];
      if (l2) {
        l = d * d;
        g = 2.0f * d * invCount;
      } else {
        l = fabsf(d);
        g = d > 0.0f ? invCount : (d < 0.0f ? -invCount : 0.0f);
      }
    }
    dOut[i] = g;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lossPart[blockIdx.x] = red[0];
}

__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float c1, float c2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// bf16 copies of one convolution's weights: W[tap][co][ci] and the
// transpose WT[tap][ci][co]
__global__ void k_weights_bf16(const float* __restrict__ w, int cout, int cin, uint16_t* __restrict__ wb,
                               uint16_t* __restrict__ wtb) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = 9LL * cout * cin;
  if (i >= n) return;
  const int ci = static_cast<int>(i % cin);
  const int64_t q = i / cin;
  const int co = static_cast<int>(q % cout), tap = static_cast<int>(q / cout);
  const uint16_t b = f2bf(w[i]);
  wb[i] = b;
  wtb[(static_cast<int64_t>(tap) * cin + ci) * cout + co] = b;
}

inline unsigned blocks(int64_t n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

// Output channels in passes of at most 128 (Cout a multiple of 32). A wave
// takes 4 pixel tiles for 32 output channels, 2 for more, so each weight
// fragment feeds 2 to 4 products; fewer where rows are narrower than that
// or where the launch would have fewer than minWgs workgroups (1,024 by
// default: on the small deep levels parallelism matters more than weight
// reuse).
// resid / pred: write only the prediction (fp32 output, see k_conv3x3).
void launchConv(SrcDev s0, SrcDev s1, int N, int H, int W, int Cin, int Cout, const uint16_t* w, int flip,
                const float* bias, int relu, uint16_t* outBf, float* outF, hipStream_t st, int64_t minWgs,
                const float* resid = nullptr, float* pred = nullptr) {
  const int64_t rowTiles = static_cast<int64_t>(N) * ((H + kConvRows - 1) / kConvRows);
  for (int co = 0; co < Cout; co += 128) {
    const int nt = std::min(8, (Cout - co) / 16);
    int np = nt <= 2 ? 4 : 2;
    while (np > 1 && (W < 16 * np || rowTiles * ((W + 16 * np - 1) / (16 * np)) < minWgs)) np /= 2;
    const int64_t wgs = rowTiles * ((W + 16 * np - 1) / (16 * np));
    const dim3 g(static_cast<unsigned>(wgs)), b(256);
#define UBG_CONV(K, P) \
  hipLaunchKernelGGL((k_conv3x3<K, P>), g, b, 0, st, s0, s1, N, H, W, Cin, Cout, co, w, flip, bias, relu, outBf, outF, \
                     resid, pred)
    if (np == 4) {
      UBG_CONV(2, 4);
    } else if (np == 2) {
      switch (nt) {
        case 2: UBG_CONV(2, 2); break;
        case 4: UBG_CONV(4, 2); break;
        case 6: UBG_CONV(6, 2); break;
        default: UBG_CONV(8, 2); break;
      }
    } else {
      switch (nt) {
        case 2: UBG_CONV(2, 1); break;
        case 4: UBG_CONV(4, 1); break;
        case 6: UBG_CONV(6, 1); break;
        default: UBG_CONV(8, 1); break;
      }
    }
#undef UBG_CONV
  }
}

}  // namespace

Unet::~Unet() {
  for (void* p : scratch_) (void)hipFree(p);
  for (void* p : {static_cast<void*>(params_), static_cast<void*>(grads_), static_cast<void*>(m_),
                  static_cast<void*>(v_), static_cast<void*>(wBf_), static_cast<void*>(wtBf_)})
    if (p) (void)hipFree(p);
}

int Unet::init(int levels, const int* channels, uint64_t seed, std::string& err) {
  if (levels < 2 || levels > kUnetMaxLevels) {
    err = "unet: levels must be 2..6";
    return 1;
  }
  L_ = levels;
  for (int l = 0; l < L_; ++l) {
    ch_[l] = channels[l];
    if (ch_[l] <= 0 || ch_[l] % 32 != 0 || ch_[l] > 128 || (ch_[l] / 16 != 2 && ch_[l] / 16 != 4 &&
                                                              ch_[l] / 16 != 6 && ch_[l] / 16 != 8)) {
      err = "unet: channels must be 32, 64, 96 or 128";
      return 1;
    }
  }
  acts_.clear();
  convs_.clear();
  acts_.push_back({kUnetInPad, 0});
  auto addAct = [&](int C, int level) {
    acts_.push_back({C, level});
    return static_cast<int>(acts_.size()) - 1;
  };
  auto addConv = [&](int level, Src s0, Src s1, int cout, int relu, int out) {
    Conv c;
    c.level = level;
    c.s0 = s0;
    c.s1 = s1;
    c.cin = acts_[s0.act].C + (s1.act >= 0 ? acts_[s1.act].C : 0);
    c.cout = cout;
    c.relu = relu;
    c.out = out;
    convs_.push_back(c);
  };
  std::vector<int> e(L_);
  int prev = 0;
  for (int l = 0; l < L_; ++l) {
    const int a = addAct(ch_[l], l);
    addConv(l, Src{prev, l == 0 ? 0 : 1}, Src{}, ch_[l], 1, a);
    e[l] = addAct(ch_[l], l);
    addConv(l, Src{a, 0}, Src{}, ch_[l], 1, e[l]);
    prev = e[l];
  }
  int d = e[L_ - 1];
  for (int l = L_ - 2; l >= 0; --l) {
    const int a = addAct(ch_[l], l);
    addConv(l, Src{e[l], 0}, Src{d, 2}, ch_[l], 1, a);
    const int b = addAct(ch_[l], l);
    addConv(l, Src{a, 0}, Src{}, ch_[l], 1, b);
    d = b;
  }
  addConv(0, Src{d, 0}, Src{}, kUnetOutPad, 0, -1);
  // parameter layout and He-normal initialisation (the output convolution
  // scaled down, biases zero, padded input / output channels zero)
  nParams_ = 0;
  wCount_ = 0;
  wBfOff_.clear();
  for (Conv& c : convs_) {
    c.wOff = nParams_;
    nParams_ += 9LL * c.cout * c.cin;
    c.bOff = nParams_;
    nParams_ += c.cout;
    wBfOff_.push_back(wCount_);
    wCount_ += 9LL * c.cout * c.cin;
  }
  std::vector<float> host(static_cast<size_t>(nParams_), 0.0f);
  std::mt19937_64 rng(seed);
  for (size_t k = 0; k < convs_.size(); ++k) {
    const Conv& c = convs_[k];
    const int cinReal = (k == 0) ? kUnetFeatures : c.cin;
    const int coutReal = c.out < 0 ? 3 : c.cout;
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * cinReal)) * (c.out < 0 ? 0.1f : 1.0f));
    for (int tap = 0; tap < 9; ++tap)
      for (int co = 0; co < c.cout; ++co)
        for (int ci = 0; ci < c.cin; ++ci) {
          const float v = nd(rng);
          if (ci < cinReal && co < coutReal) host[static_cast<size_t>(c.wOff + (static_cast<int64_t>(tap) * c.cout + co) * c.cin + ci)] = v;
        }
  }
  const size_t pb = static_cast<size_t>(nParams_) * 4;
  if (hipMalloc(&params_, pb) != hipSuccess || hipMalloc(&grads_, pb) != hipSuccess ||
      hipMalloc(&m_, pb) != hipSuccess || hipMalloc(&v_, pb) != hipSuccess ||
      hipMalloc(&wBf_, static_cast<size_t>(wCount_) * 2) != hipSuccess ||
      hipMalloc(&wtBf_, static_cast<size_t>(wCount_) * 2) != hipSuccess) {
    err = "unet: device allocation failed";
    return 1;
  }
  if (hipMemcpy(params_, host.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(grads_, 0, pb) != hipSuccess || hipMemset(m_, 0, pb) != hipSuccess ||
      hipMemset(v_, 0, pb) != hipSuccess) {
    err = "unet: device initialisation failed";
    return 1;
  }
  step_ = 0;
  bf16Dirty_ = true;
  return 0;
}

hipError_t Unet::getParams(float* host, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(host, params_, static_cast<size_t>(nParams_) * 4, hipMemcpyDeviceToHost, st);
  return e != hipSuccess ? e : hipStreamSynchronize(st);
}

hipError_t Unet::setParams(const float* host, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(params_, host, static_cast<size_t>(nParams_) * 4, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  bf16Dirty_ = true;
  return hipStreamSynchronize(st);
}

hipError_t Unet::copyGrads(float* host, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(host, grads_, static_cast<size_t>(nParams_) * 4, hipMemcpyDeviceToHost, st);
  return e != hipSuccess ? e : hipStreamSynchronize(st);
}

void Unet::syncBf16(hipStream_t st) {
  if (!bf16Dirty_) return;
  for (size_t k = 0; k < convs_.size(); ++k) {
    const Conv& c = convs_[k];
    const int64_t n = 9LL * c.cout * c.cin;
    hipLaunchKernelGGL(k_weights_bf16, dim3(blocks(n, 256)), dim3(256), 0, st, params_ + c.wOff, c.cout, c.cin,
                       wBf_ + wBfOff_[k], wtBf_ + wBfOff_[k]);
  }
  bf16Dirty_ = false;
}

hipError_t Unet::ensure(int n, int h, int w) {
  // every buffer's size is a function of the pixel count n h w alone, so a
  // shape of no more pixels (a transposed image, a smaller batch) reuses them
  const int64_t P0 = static_cast<int64_t>(n) * h * w;
  if (P0 <= capP0_) return hipSuccess;
  for (void* p : scratch_) (void)hipFree(p);
  scratch_.clear();
  actBf_.assign(acts_.size(), nullptr);
  actGrad_.assign(acts_.size(), nullptr);
  auto alloc = [&](size_t bytes, void** p) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 256));
    if (e == hipSuccess) scratch_.push_back(*p);
    return e;
  };
  capP0_ = 0;
  hipError_t e = hipSuccess;
  int64_t maxDz = 0, maxDx = 0, maxPart = 0;
  for (size_t a = 0; a < acts_.size() && e == hipSuccess; ++a) {
    const int64_t P = P0 >> (2 * acts_[a].level);
    e = alloc(static_cast<size_t>(P * acts_[a].C) * 2, reinterpret_cast<void**>(&actBf_[a]));
    if (e == hipSuccess && a > 0) e = alloc(static_cast<size_t>(P * acts_[a].C) * 4, reinterpret_cast<void**>(&actGrad_[a]));
  }
  for (const Conv& c : convs_) {
    const int64_t P = P0 >> (2 * c.level);
    maxDz = std::max(maxDz, P * c.cout);
    maxDx = std::max(maxDx, P * c.cin);
    const int64_t nChunks = std::min<int64_t>(64, std::max<int64_t>(1, P / 1024));
    maxPart = std::max<int64_t>(maxPart, nChunks * 9LL * c.cout * c.cin);
    maxPart = std::max<int64_t>(maxPart, 4096LL * c.cout);
  }
  if (e == hipSuccess) e = alloc(static_cast<size_t>(P0) * kUnetOutPad * 4, reinterpret_cast<void**>(&outF_));
  if (e == hipSuccess) e = alloc(static_cast<size_t>(P0) * kUnetOutPad * 4, reinterpret_cast<void**>(&dOut_));
  if (e == hipSuccess) e = alloc(static_cast<size_t>(maxDz) * 2, reinterpret_cast<void**>(&dz_));
  if (e == hipSuccess) e = alloc(static_cast<size_t>(maxDx) * 4, reinterpret_cast<void**>(&dx_));
  if (e == hipSuccess) e = alloc(static_cast<size_t>(maxPart) * 4, reinterpret_cast<void**>(&part_));
  if (e == hipSuccess) e = alloc(static_cast<size_t>(blocks(P0 * kUnetOutPad, 256)) * 4, reinterpret_cast<void**>(&lossPart_));
  if (e != hipSuccess) return e;
  capP0_ = P0;
  return hipSuccess;
}

hipError_t Unet::forward(int n, int h, int w, const float* feat, float* out, hipStream_t st) {
  hipError_t e = ensure(n, h, w);
  if (e != hipSuccess) return e;
  syncBf16(st);
  const int64_t P0 = static_cast<int64_t>(n) * h * w;
  // inference reads the features directly in the first convolution and
  // writes the prediction from the last; training keeps the padded input
  // (the weight gradient reads it) and the full output (the loss reads it)
  if (!out) hipLaunchKernelGGL(k_unet_input, dim3(blocks(P0 * kUnetInPad, 256)), dim3(256), 0, st, P0, feat, actBf_[0]);
  for (size_t k = 0; k < convs_.size(); ++k) {
    const Conv& c = convs_[k];
    const int H = h >> c.level, W = w >> c.level;
    auto src = [&](const Src& s) {
      SrcDev d{nullptr, 0, 0, 0, 0};
      if (s.act < 0) return d;
      if (s.act == 0 && out) return SrcDev{reinterpret_cast<const uint16_t*>(feat), kUnetInPad, 3, h, w};
      const int lv = acts_[s.act].level;
      d.p = actBf_[s.act];
      d.C = acts_[s.act].C;
      d.mode = s.mode;
      d.H = h >> lv;
      d.W = w >> lv;
      return d;
    };
    const bool last = c.out < 0;
    launchConv(src(c.s0), src(c.s1), n, H, W, c.cin, c.cout, wBf_ + wBfOff_[k], 0, params_ + c.bOff, c.relu,
               last ? nullptr : actBf_[c.out], last && !out ? outF_ : nullptr, st, minWgs_,
               last ? feat : nullptr, last ? out : nullptr);
  }
  return hipGetLastError();
}

hipError_t Unet::gradients(int n, int h, int w, const float* feat, const float* target, float* loss,
                           hipStream_t st) {
  hipError_t e = forward(n, h, w, feat, nullptr, st);
  if (e != hipSuccess) return e;
  const int64_t P0 = static_cast<int64_t>(n) * h * w;
  const unsigned lb = blocks(P0 * kUnetOutPad, 256);
  hipLaunchKernelGGL(k_unet_loss, dim3(lb), dim3(256), 0, st, P0, outF_, feat, target,
                     1.0f / static_cast<float>(P0 * 3), lossL2_ ? 1 : 0, dOut_, lossPart_);
  for (size_t a = 1; a < acts_.size(); ++a) {
    const int64_t P = P0 >> (2 * acts_[a].level);
    e = hipMemsetAsync(actGrad_[a], 0, static_cast<size_t>(P * acts_[a].C) * 4, st);
    if (e != hipSuccess) return e;
  }
  for (int k = static_cast<int>(convs_.size()) - 1; k >= 0; --k) {
    const Conv& c = convs_[static_cast<size_t>(k)];
    const int H = h >> c.level, W = w >> c.level;
    const int64_t P = static_cast<int64_t>(n) * H * W;
    const float* dy = c.out >= 0 ? actGrad_[c.out] : dOut_;
    const uint16_t* y = c.out >= 0 ? actBf_[c.out] : nullptr;
    const int relu = c.out >= 0 ? c.relu : 0;
    hipLaunchKernelGGL(k_mask_grad, dim3(blocks(P * c.cout, 256)), dim3(256), 0, st, P * c.cout, dy, y, relu, dz_);
    // bias gradient
    {
      const int64_t g = std::min<int64_t>(4096, std::max<int64_t>(1, P / 64));
      const int64_t chunk = (P + g - 1) / g;
      hipLaunchKernelGGL(k_bias_part, dim3(static_cast<unsigned>(g)), dim3(128), 0, st, dy, y, relu, c.cout, P,
                         chunk, part_);
      hipLaunchKernelGGL(k_reduce_parts, dim3(blocks(c.cout, 128)), dim3(128), 0, st, part_,
                         static_cast<int64_t>(c.cout), static_cast<int>(g), grads_ + c.bOff);
    }
    auto src = [&](const Src& s) {
      SrcDev d{nullptr, 0, 0, 0, 0};
      if (s.act < 0) return d;
      const int lv = acts_[s.act].level;
      d.p = actBf_[s.act];
      d.C = acts_[s.act].C;
      d.mode = s.mode;
      d.H = h >> lv;
      d.W = w >> lv;
      return d;
    };
    const SrcDev s0 = src(c.s0), s1 = src(c.s1);
    // weight gradient (split over pixel chunks)
    {
      const int64_t nChunks = std::min<int64_t>(64, std::max<int64_t>(1, P / 1024));
      int64_t chunk = (P + nChunks - 1) / nChunks;
      chunk = (chunk + 31) / 32 * 32;
      const int64_t waves = nChunks * 9LL * (c.cout / 16) * (c.cin / 16);
      hipLaunchKernelGGL(k_conv_dw, dim3(blocks(waves * 64, 256)), dim3(256), 0, st, dz_, c.cout, s0, s1, c.cin, n,
                         H, W, static_cast<int>(chunk), static_cast<int>(nChunks), part_);
      const int64_t nw = 9LL * c.cout * c.cin;
      hipLaunchKernelGGL(k_reduce_parts, dim3(blocks(nw, 256)), dim3(256), 0, st, part_, nw,
                         static_cast<int>(nChunks), grads_ + c.wOff);
    }
    // operand gradient, then back onto the sources (not onto the input features)
    const bool need0 = c.s0.act > 0, need1 = c.s1.act > 0;
    if (!need0 && !need1) continue;
    SrcDev dzs{dz_, c.cout, 0, H, W};
    SrcDev none{nullptr, 0, 0, 0, 0};
    launchConv(dzs, none, n, H, W, c.cout, c.cin, wtBf_ + wBfOff_[static_cast<size_t>(k)], 1, nullptr, 0,
               nullptr, dx_, st, minWgs_);
    auto scatter = [&](const Src& s, int off) {
      const int lv = acts_[s.act].level;
      const int Hs = h >> lv, Ws = w >> lv, Cs = acts_[s.act].C;
      const int64_t total = static_cast<int64_t>(n) * Hs * Ws * Cs;
      hipLaunchKernelGGL(k_scatter, dim3(blocks(total, 256)), dim3(256), 0, st, dx_, c.cin, off, n, H, W, s.mode,
                         actGrad_[s.act], Cs, Hs, Ws);
    };
    if (need0) scatter(c.s0, 0);
    if (need1) scatter(c.s1, acts_[c.s0.act].C);
  }
  if (loss) {
    std::vector<float> lp(lb);
    e = hipMemcpyAsync(lp.data(), lossPart_, static_cast<size_t>(lb) * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    double s = 0.0;
    for (float v : lp) s += v;
    *loss = static_cast<float>(s / static_cast<double>(P0 * 3));
  }
  return hipGetLastError();
}

hipError_t Unet::adam(float lr, hipStream_t st) {
  ++step_;
  const float c1 = 1.0f - std::pow(0.9f, static_cast<float>(step_));
  const float c2 = 1.0f - std::pow(0.999f, static_cast<float>(step_));
  hipLaunchKernelGGL(k_adam, dim3(blocks(nParams_, 256)), dim3(256), 0, st, nParams_, params_, grads_, m_, v_, lr,
                     c1, c2);
  bf16Dirty_ = true;
  return hipGetLastError();
}

}  // namespace ubg
