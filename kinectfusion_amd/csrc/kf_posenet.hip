// Pose-proposal network for the learned ICP warm start: BF16 MFMA training
// and inference on gfx950.  The reference has no counterpart; its hook is
// the tracker's initial pose (src/app/reconstruction.cpp:218-250), and the
// fallback reads IcpNormalEquations::mean_squared_residual
// (icp_correspondence.hpp:133-139).
//
// Network: input 2 x 120 x 160 (previous and current coarsest-level depth,
// NHWC, bf16) -> five 3x3 stride-2 pad-1 convolutions with ReLU
// (16, 32, 64, 64, 128 channels; 60x80, 30x40, 15x20, 8x10, 4x5) -> flatten
// (4 * 5 * 128 = 2560, NHWC order) -> 256 with ReLU -> 6.  Input depths are
// metres minus 1.5 m (0 where invalid), which keeps bf16's 8-bit mantissa
// fine around the working range.  The output times (0.1, 0.1, 0.1, 0.03,
// 0.03, 0.03) is the twist (omega rad, t m) of the motion from the previous
// camera to the current one: T_cur = T_prev * [R(omega) | t].
//
// Every layer is one GEMM on the matrix cores (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation): convolutions as im2col (K = 9 * Cin, k = (ky * 3 + kx)
// * Cin + ci) times W^T, their weight gradients as split-K GEMMs over the
// batch-pixel axis with fixed-order partial sums, input gradients as GEMMs
// followed by a gather col2im.  Master weights and Adam moments are FP32;
// the GEMM operands are bf16 copies.  Every reduction has a fixed order, so
// training is deterministic.
#include <cmath>
#include <cstring>
#include <vector>

#include "kf_internal.hpp"

namespace {

using kf::fail;

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kInH = 120, kInW = 160;
constexpr int kConv = 5;
constexpr int kCh[kConv + 1] = {2, 16, 32, 64, 64, 128};
constexpr int kHidden = 256, kOut = 6;
constexpr int kLayers = kConv + 2;
constexpr float kDepthCentre = 1.5F;  // metres
__host__ __device__ inline float out_scale(int component) { return component < 3 ? 0.1F : 0.03F; }

__host__ __device__ inline uint16_t to_bf16(float f) {  // round to nearest even
  const uint32_t b = __builtin_bit_cast(uint32_t, f);
  return static_cast<uint16_t>((b + 0x7FFFU + ((b >> 16U) & 1U)) >> 16U);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
  return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16U);
}

struct Shape {
  bool conv;
  int cin, cout, hin, win, hout, wout;
  int fan_in;         // K of the layer's GEMM
  size_t w_off, b_off;  // into the packed parameter vector
};

std::vector<Shape> network_shapes(size_t* total) {
  std::vector<Shape> s;
  int h = kInH, w = kInW;
  size_t off = 0;
  for (int l = 0; l < kConv; ++l) {
    Shape x{true, kCh[l], kCh[l + 1], h, w, (h + 1) / 2, (w + 1) / 2, 9 * kCh[l], 0, 0};
    x.w_off = off;
    off += static_cast<size_t>(x.cout) * x.fan_in;
    x.b_off = off;
    off += x.cout;
    s.push_back(x);
    h = x.hout;
    w = x.wout;
  }
  const int flat = h * w * kCh[kConv];
  Shape f1{false, flat, kHidden, 1, 1, 1, 1, flat, 0, 0};
  f1.w_off = off;
  off += static_cast<size_t>(kHidden) * flat;
  f1.b_off = off;
  off += kHidden;
  s.push_back(f1);
  Shape f2{false, kHidden, kOut, 1, 1, 1, 1, kHidden, 0, 0};
  f2.w_off = off;
  off += static_cast<size_t>(kOut) * kHidden;
  f2.b_off = off;
  off += kOut;
  s.push_back(f2);
  if (total != nullptr) *total = off;
  return s;
}

// ---------------------------------------------------------------------------
// GEMM: C(m, n) = sum_k A(m, k) B(k, n), bf16 operands with arbitrary
// strides, FP32 accumulation.  64x64 tile per 256-thread workgroup, each of
// the four waves a 32x32 quadrant of 2x2 16x16x32 MFMA blocks, K staged 32
// at a time through LDS.  gridDim.z splits K: every split writes its FP32
// partial tile and a second kernel adds the partials in split order 0, 1,
// ... and applies the epilogue, so the sum does not depend on timing.
// A 3x3 stride-2 pad-1 convolution read as its im2col matrix, element
// (m, k) = x[b][2 yo - 1 + ky][2 xo - 1 + kx][c] with m = (b Ho + yo) Wo + xo,
// k = (ky 3 + kx) C + c (0 outside the image).  With `raw_prev` set the input
// is the depth pair itself (u16, channel 0 previous, 1 current), converted as
// the input layer does.
struct ConvView {
  const uint16_t* x;
  const uint16_t* raw_prev;  // pair b at raw_prev + b * raw_stride
  const uint16_t* raw_cur;
  long long raw_stride;
  float inv_scale;
  int H, W, C, Ho, Wo;
};

struct Gemm {
  const uint16_t* a;  // null: A is the im2col matrix of `conv`
  ConvView conv;
  long long sam, sak;
  const uint16_t* b;
  long long sbk, sbn;
  int M, N, K, kchunk;
  float* partial;  // [splits][M][N] when gridDim.z > 1
  // epilogue: v (+ bias[n]) (ReLU) -> outf and/or outh
  const float* bias;
  int relu;
  float* outf;
  long long ofm, ofn;
  uint16_t* outh;
  long long ohm, ohn;
  int scale_out;  // outf receives v * out_scale(n) (the twist)
};

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // 80-byte rows: 16-byte fragment reads stay aligned
constexpr int kStageRows = 256 / BK;           // rows one pass of the workgroup stages
constexpr int kRowsPerThread = BM / kStageRows;  // 8
constexpr int kMaxTiles = 256;                   // at most this many tiles split K

__device__ inline uint16_t raw_to_input(uint16_t r, float inv_scale) {
  return to_bf16(r == 0 ? 0.0F : static_cast<float>(r) * inv_scale - kDepthCentre);
}

__device__ inline void gemm_store(const Gemm& g, int m, int n, float v) {
  if (g.bias != nullptr) v += g.bias[n];
  if (g.relu != 0) v = fmaxf(v, 0.0F);
  if (g.outf != nullptr) g.outf[m * g.ofm + n * g.ofn] = g.scale_out != 0 ? v * out_scale(n) : v;
  if (g.outh != nullptr) g.outh[m * g.ohm + n * g.ohn] = to_bf16(v);
}

__global__ void __launch_bounds__(256) gemm_kernel(Gemm g) {
  __shared__ __align__(16) uint16_t As[BM][LDK];
  __shared__ __align__(16) uint16_t Bs[BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int kb = blockIdx.z * g.kchunk;
  const int ke = min(g.K, kb + g.kchunk);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int li = lane & 15, lq = lane >> 4;
  // This thread stages column `sc` of rows sr + kStageRows * i of both tiles.
  const int sc = tid & (BK - 1), sr = tid >> 5;
  const bool implicit = g.a == nullptr;
  // Implicit im2col: the receptive-field origin of each staged row is fixed
  // over the K loop (rows past M get an origin no tap reaches).
  int oy[kRowsPerThread], ox[kRowsPerThread], ob[kRowsPerThread];
#pragma unroll
  for (int i = 0; i < kRowsPerThread; ++i) {
    oy[i] = -4;
    ox[i] = -4;
    ob[i] = 0;
    const int m = m0 + sr + kStageRows * i;
    if (implicit && m < g.M) {
      const int xo = m % g.conv.Wo, t = m / g.conv.Wo;
      oy[i] = 2 * (t % g.conv.Ho) - 1;
      ox[i] = 2 * xo - 1;
      ob[i] = t / g.conv.Ho;
    }
  }
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
  for (int k0 = kb; k0 < ke; k0 += BK) {
    const int k = k0 + sc;
    const bool kin = k < ke;
    int ky = 0, kx = 0, ch = 0;
    if (implicit) {
      const int tap = k / g.conv.C;
      ch = k - tap * g.conv.C;
      ky = tap / 3;
      kx = tap - ky * 3;
    }
#pragma unroll
    for (int i = 0; i < kRowsPerThread; ++i) {
      const int r = sr + kStageRows * i;
      const int m = m0 + r, n = n0 + r;
      uint16_t av = 0;
      if (kin) {
        if (!implicit) {
          if (m < g.M) av = g.a[m * g.sam + k * g.sak];
        } else {
          const int yi = oy[i] + ky, xi = ox[i] + kx;
          if (yi >= 0 && yi < g.conv.H && xi >= 0 && xi < g.conv.W) {
            if (g.conv.raw_prev != nullptr) {
              const uint16_t* src = ch == 0 ? g.conv.raw_prev : g.conv.raw_cur;
              av = raw_to_input(src[ob[i] * g.conv.raw_stride + yi * g.conv.W + xi], g.conv.inv_scale);
            } else {
              av = g.conv.x[((static_cast<long long>(ob[i]) * g.conv.H + yi) * g.conv.W + xi) * g.conv.C + ch];
            }
          }
        }
      }
      As[r][sc] = av;
      Bs[r][sc] = (n < g.N && kin) ? g.b[k * g.sbk + n * g.sbn] : static_cast<uint16_t>(0);
    }
    __syncthreads();
    {  // the staged 32 k as one 16x16x32 step
      s16x8 av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        av[i] = *reinterpret_cast<const s16x8*>(&As[wm + i * 16 + li][lq * 8]);
        bv[i] = *reinterpret_cast<const s16x8*>(&Bs[wn + i * 16 + li][lq * 8]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, av[i]),
                                                              __builtin_bit_cast(bf16x8, bv[j]), acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // lane holds D[4 * lq + r][li] of each 16x16 block
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + lq * 4 + r, n = n0 + wn + j * 16 + li;
        if (m >= g.M || n >= g.N) continue;
        if (split) {
          g.partial[(static_cast<size_t>(blockIdx.z) * g.M + m) * g.N + n] = acc[i][j][r];
        } else {
          gemm_store(g, m, n, acc[i][j][r]);
        }
      }
}

// Sum of the split partials in split order 0, 1, ... (eight reads in
// flight per thread), then the epilogue.
__global__ void gemm_finish_kernel(Gemm g, int splits) {
  const long long total = static_cast<long long>(g.M) * g.N;
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float* p = g.partial + i;
  float s = 0.0F;
  int z = 0;
  for (; z + 8 <= splits; z += 8) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[(z + j) * total];
#pragma unroll
    for (int j = 0; j < 8; ++j) s += v[j];
  }
  for (; z < splits; ++z) s += p[z * total];
  gemm_store(g, static_cast<int>(i / g.N), static_cast<int>(i % g.N), s);
}

// ---------------------------------------------------------------------------
// Small GEMM (few output tiles, K up to a few thousand: the layers of a pose
// proposal, and the forward layers of small batches).  One workgroup per
// 16 x 16 output tile; its Waves waves (8, or 16 for long reductions) split
// the K steps (32 wide: v_mfma_f32_16x16x32_bf16, gfx950's bf16 shape)
// between them and load their operand fragments straight from global memory
// into registers, kSmallChunk steps at a time, so that a layer costs one or
// two memory round trips instead of one per K step.  Lane l holds
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16] (j = 0..7) of a
// step and D[4 (l / 16) + r][l % 16] of the tile.  The per-wave tiles are
// added in wave order through LDS, so the result does not depend on timing.
constexpr int kSmallChunk = 5;

// A(m, k) of the im2col matrix for a row whose receptive-field origin is
// (oy, ox) of pair ob (oy < -2 marks a row past M).
__device__ inline uint16_t conv_a(const ConvView& v, int oy, int ox, int ob, int k) {
  const int tap = k / v.C, ch = k - tap * v.C;
  const int ky = tap / 3, kx = tap - ky * 3;
  const int yi = oy + ky, xi = ox + kx;
  if (yi < 0 || yi >= v.H || xi < 0 || xi >= v.W) return 0;
  if (v.raw_prev != nullptr) {
    const uint16_t* src = ch == 0 ? v.raw_prev : v.raw_cur;
    return raw_to_input(src[ob * v.raw_stride + yi * v.W + xi], v.inv_scale);
  }
  return v.x[((static_cast<long long>(ob) * v.H + yi) * v.W + xi) * v.C + ch];
}

__device__ inline short as_short(uint16_t h) { return static_cast<short>(h); }

template <int Waves>
__global__ void __launch_bounds__(Waves * 64) gemm_small_kernel(Gemm g, int vec_a, int vec_b) {
  __shared__ float red[Waves][16][17];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lq = lane >> 4;
  const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
  const int m = m0 + li, n = n0 + li;  // this lane's A row and B column
  const bool mok = m < g.M, nok = n < g.N;
  const bool implicit = g.a == nullptr;
  int oy = -4, ox = -4, ob = 0;
  if (implicit && mok) {
    const int xo = m % g.conv.Wo, t = m / g.conv.Wo;
    oy = 2 * (t % g.conv.Ho) - 1;
    ox = 2 * xo - 1;
    ob = t / g.conv.Ho;
  }
  const int steps = (g.K + 31) / 32;
  const int per = (steps + Waves - 1) / Waves;
  const int s_begin = wave * per, s_end = min(steps, s_begin + per);
  f32x4 acc = f32x4{0.0F, 0.0F, 0.0F, 0.0F};
  for (int c0 = s_begin; c0 < s_end; c0 += kSmallChunk) {
    s16x8 af[kSmallChunk], bfr[kSmallChunk];
#pragma unroll
    for (int c = 0; c < kSmallChunk; ++c) {
      af[c] = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
      bfr[c] = s16x8{0, 0, 0, 0, 0, 0, 0, 0};
      const int k = (c0 + c) * 32 + lq * 8;
      if (c0 + c >= s_end || k >= g.K) continue;
      const bool full = k + 7 < g.K;
      if (mok) {
        if (implicit) {
          if (vec_a != 0 && full) {
            // C % 8 == 0: the eight k are channels ch..ch+7 of one tap
            const int tap = k / g.conv.C, ch = k - tap * g.conv.C;
            const int ky = tap / 3, kx = tap - ky * 3;
            const int yi = oy + ky, xi = ox + kx;
            if (yi >= 0 && yi < g.conv.H && xi >= 0 && xi < g.conv.W) {
              af[c] = *reinterpret_cast<const s16x8*>(
                  g.conv.x + ((static_cast<long long>(ob) * g.conv.H + yi) * g.conv.W + xi) * g.conv.C + ch);
            }
          } else {
#pragma unroll
            for (int j = 0; j < 8; ++j)
              if (k + j < g.K) af[c][j] = as_short(conv_a(g.conv, oy, ox, ob, k + j));
          }
        } else if (vec_a != 0 && full) {
          af[c] = *reinterpret_cast<const s16x8*>(g.a + m * g.sam + k);
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (k + j < g.K) af[c][j] = as_short(g.a[m * g.sam + (k + j) * g.sak]);
        }
      }
      if (nok) {
        if (vec_b != 0 && full) {
          bfr[c] = *reinterpret_cast<const s16x8*>(g.b + n * g.sbn + k);
        } else {
#pragma unroll
          for (int j = 0; j < 8; ++j)
            if (k + j < g.K) bfr[c][j] = as_short(g.b[(k + j) * g.sbk + n * g.sbn]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kSmallChunk; ++c)
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, af[c]), __builtin_bit_cast(bf16x8, bfr[c]),
                                                    acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) red[wave][lq * 4 + r][li] = acc[r];
  __syncthreads();
  if (threadIdx.x < 256) {
    const int row = threadIdx.x >> 4, col = threadIdx.x & 15;
    const int om = m0 + row, on = n0 + col;
    if (om < g.M && on < g.N) {
      float s = 0.0F;
#pragma unroll
      for (int w = 0; w < Waves; ++w) s += red[w][row][col];
      gemm_store(g, om, on, s);
    }
  }
}

// ---------------------------------------------------------------------------
// Depth pair -> NHWC bf16 centred metres.  pair b: prev = prev0 + b * stride,
// cur = cur0 + b * stride (u16 raw, 160 x 120 each).
__global__ void input_kernel(const uint16_t* prev0, const uint16_t* cur0, long long stride, int batch,
                             float inv_scale, uint16_t* x0) {
  const int px = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (px >= kInH * kInW || b >= batch) return;
  const uint16_t rp = prev0[b * stride + px], rc = cur0[b * stride + px];
  const float p = rp == 0 ? 0.0F : static_cast<float>(rp) * inv_scale - kDepthCentre;
  const float c = rc == 0 ? 0.0F : static_cast<float>(rc) * inv_scale - kDepthCentre;
  const size_t o = (static_cast<size_t>(b) * kInH * kInW + px) * 2;
  x0[o] = to_bf16(p);
  x0[o + 1] = to_bf16(c);
}

// col[m][k]: m = (b * Ho + yo) * Wo + xo, k = (ky * 3 + kx) * C + c.
__global__ void im2col_kernel(const uint16_t* x, uint16_t* col, int batch, int H, int W, int C, int Ho,
                              int Wo) {
  const int K = 9 * C;
  const long long total = static_cast<long long>(batch) * Ho * Wo * K;
  for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long m = e / K;
    const int k = static_cast<int>(e % K);
    const int tap = k / C, c = k % C;
    const int ky = tap / 3, kx = tap % 3;
    const int xo = static_cast<int>(m % Wo);
    const long long t = m / Wo;
    const int yo = static_cast<int>(t % Ho);
    const int b = static_cast<int>(t / Ho);
    const int yi = 2 * yo - 1 + ky, xi = 2 * xo - 1 + kx;
    col[e] = (yi >= 0 && yi < H && xi >= 0 && xi < W)
                 ? x[((static_cast<long long>(b) * H + yi) * W + xi) * C + c]
                 : static_cast<uint16_t>(0);
  }
}

// dx[b][yi][xi][c] = sum over the taps that read it of dcol, in (ky, kx)
// order, times the ReLU mask of the layer input; bf16 out.
__global__ void col2im_kernel(const float* dcol, const uint16_t* xin, uint16_t* dx, int batch, int H, int W,
                              int C, int Ho, int Wo) {
  const long long total = static_cast<long long>(batch) * H * W * C;
  const int K = 9 * C;
  for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    float s = 0.0F;
    if (from_bf16(xin[e]) > 0.0F) {
      const int c = static_cast<int>(e % C);
      const long long p = e / C;
      const int xi = static_cast<int>(p % W);
      const long long t = p / W;
      const int yi = static_cast<int>(t % H);
      const int b = static_cast<int>(t / H);
      for (int ky = 0; ky < 3; ++ky) {
        const int ty = yi + 1 - ky;
        if (ty < 0 || (ty & 1) != 0 || (ty >> 1) >= Ho) continue;
        for (int kx = 0; kx < 3; ++kx) {
          const int tx = xi + 1 - kx;
          if (tx < 0 || (tx & 1) != 0 || (tx >> 1) >= Wo) continue;
          const long long m = (static_cast<long long>(b) * Ho + (ty >> 1)) * Wo + (tx >> 1);
          s += dcol[m * K + (ky * 3 + kx) * C + c];
        }
      }
    }
    dx[e] = to_bf16(s);
  }
}

// out = act > 0 ? bf16(g) : 0
__global__ void relu_mask_kernel(const float* g, const uint16_t* act, uint16_t* out, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = from_bf16(act[i]) > 0.0F ? to_bf16(g[i]) : static_cast<uint16_t>(0);
}

// Mean squared error against the targets divided by the output scale; its
// gradient as bf16.
// One workgroup, fixed-order tree.
__global__ void __launch_bounds__(256) loss_kernel(const float* out, const float* target, int batch,
                                                   uint16_t* dout, float* loss) {
  __shared__ float red[256];
  const int n = batch * kOut;
  const float inv = 1.0F / static_cast<float>(n);
  float s = 0.0F;
  for (int i = threadIdx.x; i < n; i += 256) {
    const float d = out[i] - target[i] / out_scale(i % kOut);
    s += d * d;
    dout[i] = to_bf16(2.0F * d * inv);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] * inv;
}

// Column sums of a bf16 [M][N] matrix (bias gradients), two passes with a
// fixed order: chunks of 256 rows, then the chunks.
constexpr int kColRows = 256;
__global__ void __launch_bounds__(256) colsum_chunk_kernel(const uint16_t* x, int M, int N, float* part) {
  __shared__ float red[256];
  const int r0 = blockIdx.x * kColRows, r1 = min(M, r0 + kColRows);
  for (int nb = 0; nb < N; nb += 256) {
    const int cols = min(256, N - nb);
    const int groups = 256 / cols;  // row lanes per column
    const int col = threadIdx.x % cols, grp = threadIdx.x / cols;
    float s = 0.0F;
    if (grp < groups)
      for (int r = r0 + grp; r < r1; r += groups) s += from_bf16(x[static_cast<long long>(r) * N + nb + col]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < cols) {
      float t = 0.0F;
      for (int q = 0; q < groups; ++q) t += red[q * cols + threadIdx.x];
      part[static_cast<long long>(blockIdx.x) * N + nb + threadIdx.x] = t;
    }
    __syncthreads();
  }
}
__global__ void colsum_final_kernel(const float* part, int chunks, int N, float* out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.0F;
  for (int c = 0; c < chunks; ++c) s += part[static_cast<long long>(c) * N + n];
  out[n] = s;
}

__global__ void adam_kernel(float* w, const float* g, float* m, float* v, uint16_t* wh, long long n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0F - b1) * gi;
  const float vi = b2 * v[i] + (1.0F - b2) * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  wh[i] = to_bf16(wi);
}

__global__ void to_bf16_kernel(const float* w, uint16_t* wh, long long n) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) wh[i] = to_bf16(w[i]);
}


int blocks_for(long long n, int per = 256) { return static_cast<int>((n + per - 1) / per); }

}  // namespace

// ---------------------------------------------------------------------------
struct kf_posenet {
  kf_posenet_config cfg{};
  std::string error;
  hipStream_t stream = nullptr;
  std::vector<Shape> shapes;
  size_t n_params = 0;
  float* w = nullptr;      // FP32 master
  float* grad = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  uint16_t* wh = nullptr;  // bf16 copy used by the GEMMs
  uint16_t* act[kLayers] = {};  // act[0] input, act[l + 1] output of layer l (bf16); fc2 output is `out`
  uint16_t* col[kConv] = {};
  float* out = nullptr;    // [capacity][6] network output (normalised twist)
  float* tmp = nullptr;    // FP32 GEMM outputs of the backward pass
  uint16_t* dz[2] = {};    // bf16 gradients of layer outputs
  float* partial = nullptr;
  size_t partial_floats = 0;
  float* colpart = nullptr;
  float* loss = nullptr;   // device scalar
  float* loss_host = nullptr;  // pinned
  long long step = 0;
};

namespace {

template <typename T>
T* nalloc(size_t count) {
  void* p = nullptr;
  KF_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
  return static_cast<T*>(p);
}

void gemm(kf_posenet* net, hipStream_t s, Gemm g) {
  const int tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
  const int tiles = tm * tn;
  if (tiles <= 128 && g.K <= 4096) {
    // vector fragment loads need eight consecutive k in memory, 16-byte aligned
    auto aligned16 = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15U) == 0U; };
    int vec_a = 0;
    if (g.a == nullptr) {
      vec_a = g.conv.raw_prev == nullptr && g.conv.C % 8 == 0 && aligned16(g.conv.x) ? 1 : 0;
    } else {
      vec_a = g.sak == 1 && g.sam % 8 == 0 && aligned16(g.a) ? 1 : 0;
    }
    const int vec_b = g.sbk == 1 && g.sbn % 8 == 0 && aligned16(g.b) ? 1 : 0;
    g.partial = nullptr;
    // a reduction longer than 8 waves x kSmallChunk steps (fc1: K = 2560) takes 16 waves: one read round
    const dim3 grid((g.M + 15) / 16, (g.N + 15) / 16);
    if ((g.K + 31) / 32 > 8 * kSmallChunk) {
      hipLaunchKernelGGL(gemm_small_kernel<16>, grid, dim3(16 * 64), 0, s, g, vec_a, vec_b);
    } else {
      hipLaunchKernelGGL(gemm_small_kernel<8>, grid, dim3(8 * 64), 0, s, g, vec_a, vec_b);
    }
    KF_HIP(hipGetLastError());
    return;
  }
  // Few tiles and a long reduction: split K until about 512 workgroups run,
  // each over as little as one K step (the loop is latency bound at these
  // sizes: one pose proposal is a handful of tiles per layer).
  int splits = 1;
  if (tiles <= kMaxTiles && g.K > 2 * BK) {
    splits = std::min((g.K + BK - 1) / BK, (512 + tiles - 1) / tiles);
    const size_t cap = net->partial_floats / (static_cast<size_t>(g.M) * g.N);
    splits = static_cast<int>(std::min<size_t>(static_cast<size_t>(splits), std::max<size_t>(cap, 1)));
  }
  g.kchunk = ((g.K + splits - 1) / splits + BK - 1) / BK * BK;
  splits = (g.K + g.kchunk - 1) / g.kchunk;
  g.partial = splits > 1 ? net->partial : nullptr;
  hipLaunchKernelGGL(gemm_kernel, dim3(tm, tn, splits), dim3(256), 0, s, g);
  if (splits > 1) {
    const long long total = static_cast<long long>(g.M) * g.N;
    hipLaunchKernelGGL(gemm_finish_kernel, dim3(blocks_for(total)), dim3(256), 0, s, g, splits);
  }
  KF_HIP(hipGetLastError());
}

Gemm make_gemm(const uint16_t* a, long long sam, long long sak, const uint16_t* b, long long sbk, long long sbn,
               int M, int N, int K) {
  Gemm g{};
  g.a = a;
  g.sam = sam;
  g.sak = sak;
  g.b = b;
  g.sbk = sbk;
  g.sbn = sbn;
  g.M = M;
  g.N = N;
  g.K = K;
  return g;
}

void colsum(kf_posenet* net, hipStream_t s, const uint16_t* x, int M, int N, float* out) {
  const int chunks = (M + kColRows - 1) / kColRows;
  hipLaunchKernelGGL(colsum_chunk_kernel, dim3(chunks), dim3(256), 0, s, x, M, N, net->colpart);
  hipLaunchKernelGGL(colsum_final_kernel, dim3(blocks_for(N)), dim3(256), 0, s, net->colpart, chunks, N, out);
  KF_HIP(hipGetLastError());
}

int pixels(const Shape& sh) { return sh.hout * sh.wout; }

// Forward pass over act[0] (already filled) for `batch` pairs.
void forward(kf_posenet* net, hipStream_t s, int batch) {
  for (int l = 0; l < kConv; ++l) {
    const Shape& sh = net->shapes[l];
    const int M = batch * pixels(sh), K = sh.fan_in;
    const long long total = static_cast<long long>(M) * K;
    hipLaunchKernelGGL(im2col_kernel, dim3(std::min(blocks_for(total), 8192)), dim3(256), 0, s, net->act[l],
                       net->col[l], batch, sh.hin, sh.win, sh.cin, sh.hout, sh.wout);
    Gemm g = make_gemm(net->col[l], K, 1, net->wh + sh.w_off, 1, K, M, sh.cout, K);
    g.bias = net->w + sh.b_off;
    g.relu = 1;
    g.outh = net->act[l + 1];
    g.ohm = sh.cout;
    g.ohn = 1;
    gemm(net, s, g);
  }
  const Shape& f1 = net->shapes[kConv];
  Gemm g1 = make_gemm(net->act[kConv], f1.fan_in, 1, net->wh + f1.w_off, 1, f1.fan_in, batch, f1.cout, f1.fan_in);
  g1.bias = net->w + f1.b_off;
  g1.relu = 1;
  g1.outh = net->act[kConv + 1];
  g1.ohm = f1.cout;
  g1.ohn = 1;
  gemm(net, s, g1);
  const Shape& f2 = net->shapes[kConv + 1];
  Gemm g2 = make_gemm(net->act[kConv + 1], f2.fan_in, 1, net->wh + f2.w_off, 1, f2.fan_in, batch, f2.cout,
                      f2.fan_in);
  g2.bias = net->w + f2.b_off;
  g2.outf = net->out;
  g2.ofm = kOut;
  g2.ofn = 1;
  gemm(net, s, g2);
}

// Loss and gradients of every parameter into net->grad.
void backward(kf_posenet* net, hipStream_t s, const float* targets, int batch) {
  hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, s, net->out, targets, batch, net->dz[0], net->loss);
  KF_HIP(hipGetLastError());
  // fc2: dW = dout^T h, db = colsum(dout), dh = dout W masked by h > 0
  const Shape& f2 = net->shapes[kConv + 1];
  const uint16_t* h = net->act[kConv + 1];
  {
    Gemm g = make_gemm(net->dz[0], 1, kOut, h, f2.fan_in, 1, kOut, f2.fan_in, batch);
    g.outf = net->grad + f2.w_off;
    g.ofm = f2.fan_in;
    g.ofn = 1;
    gemm(net, s, g);
    colsum(net, s, net->dz[0], batch, kOut, net->grad + f2.b_off);
    Gemm d = make_gemm(net->dz[0], kOut, 1, net->wh + f2.w_off, f2.fan_in, 1, batch, f2.fan_in, kOut);
    d.outf = net->tmp;
    d.ofm = f2.fan_in;
    d.ofn = 1;
    gemm(net, s, d);
    const long long n = static_cast<long long>(batch) * f2.fan_in;
    hipLaunchKernelGGL(relu_mask_kernel, dim3(blocks_for(n)), dim3(256), 0, s, net->tmp, h, net->dz[1], n);
  }
  // fc1
  const Shape& f1 = net->shapes[kConv];
  const uint16_t* x5 = net->act[kConv];
  {
    Gemm g = make_gemm(net->dz[1], 1, f1.cout, x5, f1.fan_in, 1, f1.cout, f1.fan_in, batch);
    g.outf = net->grad + f1.w_off;
    g.ofm = f1.fan_in;
    g.ofn = 1;
    gemm(net, s, g);
    colsum(net, s, net->dz[1], batch, f1.cout, net->grad + f1.b_off);
    Gemm d = make_gemm(net->dz[1], f1.cout, 1, net->wh + f1.w_off, f1.fan_in, 1, batch, f1.fan_in, f1.cout);
    d.outf = net->tmp;
    d.ofm = f1.fan_in;
    d.ofn = 1;
    gemm(net, s, d);
    const long long n = static_cast<long long>(batch) * f1.fan_in;
    hipLaunchKernelGGL(relu_mask_kernel, dim3(blocks_for(n)), dim3(256), 0, s, net->tmp, x5, net->dz[0], n);
  }
  // convolutions, last to first; dz[cur] holds the gradient of layer l's output
  int cur = 0;
  for (int l = kConv - 1; l >= 0; --l) {
    const Shape& sh = net->shapes[l];
    const int M = batch * pixels(sh), N = sh.cout, K = sh.fan_in;
    Gemm g = make_gemm(net->dz[cur], 1, N, net->col[l], K, 1, N, K, M);
    g.outf = net->grad + sh.w_off;
    g.ofm = K;
    g.ofn = 1;
    gemm(net, s, g);
    colsum(net, s, net->dz[cur], M, N, net->grad + sh.b_off);
    if (l == 0) break;
    Gemm d = make_gemm(net->dz[cur], N, 1, net->wh + sh.w_off, K, 1, M, K, N);
    d.outf = net->tmp;
    d.ofm = K;
    d.ofn = 1;
    gemm(net, s, d);
    const long long total = static_cast<long long>(batch) * sh.hin * sh.win * sh.cin;
    hipLaunchKernelGGL(col2im_kernel, dim3(std::min(blocks_for(total), 8192)), dim3(256), 0, s, net->tmp,
                       net->act[l], net->dz[1 - cur], batch, sh.hin, sh.win, sh.cin, sh.hout, sh.wout);
    KF_HIP(hipGetLastError());
    cur = 1 - cur;
  }
}

void check_batch(const kf_posenet* net, int batch) {
  if (batch < 1 || batch > net->cfg.batch_capacity) {
    fail(KF_ERR_INVALID_ARGUMENT, "posenet: batch must be in [1, batch_capacity]");
  }
}

void prepare_pairs(kf_posenet* net, hipStream_t s, const uint16_t* pairs, int batch) {
  const long long hw = static_cast<long long>(kInH) * kInW;
  hipLaunchKernelGGL(input_kernel, dim3(blocks_for(hw), batch), dim3(256), 0, s, pairs, pairs + hw, 2 * hw, batch,
                     1.0F / net->cfg.depth_scale, net->act[0]);
  KF_HIP(hipGetLastError());
}

template <typename F>
kf_status net_guard(kf_posenet* net, F&& body) {
  try {
    body();
    return KF_OK;
  } catch (const kf::Failure& f) {
    if (net != nullptr) net->error = f.message;
    return f.status;
  } catch (const std::exception& e) {
    if (net != nullptr) net->error = e.what();
    return KF_ERR_DEVICE;
  }
}

// Seeded He-uniform initialisation (SplitMix64), biases zero, the output
// layer scaled by 0.1.
void init_weights(const std::vector<Shape>& shapes, uint64_t seed, std::vector<float>& w) {
  uint64_t state = seed * 0x9E3779B97F4A7C15ULL + 0x2545F4914F6CDD1DULL;
  auto next = [&]() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (size_t l = 0; l < shapes.size(); ++l) {
    const Shape& sh = shapes[l];
    const double bound = std::sqrt(6.0 / sh.fan_in) * (l + 1 == shapes.size() ? 0.1 : 1.0);
    for (size_t i = 0; i < static_cast<size_t>(sh.cout) * sh.fan_in; ++i) {
      w[sh.w_off + i] = static_cast<float>((2.0 * next() - 1.0) * bound);
    }
    for (int i = 0; i < sh.cout; ++i) w[sh.b_off + i] = 0.0F;
  }
}

// Inference forward pass for `batch` depth pairs (pair b: previous at
// prev + b * stride, current at cur + b * stride, 160 x 120 u16 each) ->
// twist[batch][6] (omega rad, t m).  Eight launches: the convolutions read
// their input as an implicit im2col matrix (the first one straight from the
// depth pairs) and the last layer writes the scaled twist.
void infer(kf_posenet* net, hipStream_t s, const uint16_t* prev, const uint16_t* cur, long long stride, int batch,
           float* twist) {
  for (int l = 0; l < kConv; ++l) {
    const Shape& sh = net->shapes[l];
    const int M = batch * pixels(sh), K = sh.fan_in;
    Gemm g = make_gemm(nullptr, 0, 0, net->wh + sh.w_off, 1, K, M, sh.cout, K);
    g.conv = ConvView{net->act[l],   l == 0 ? prev : nullptr, l == 0 ? cur : nullptr, stride,
                      1.0F / net->cfg.depth_scale, sh.hin, sh.win, sh.cin, sh.hout, sh.wout};
    g.bias = net->w + sh.b_off;
    g.relu = 1;
    g.outh = net->act[l + 1];
    g.ohm = sh.cout;
    g.ohn = 1;
    gemm(net, s, g);
  }
  const Shape& f1 = net->shapes[kConv];
  Gemm g1 = make_gemm(net->act[kConv], f1.fan_in, 1, net->wh + f1.w_off, 1, f1.fan_in, batch, f1.cout, f1.fan_in);
  g1.bias = net->w + f1.b_off;
  g1.relu = 1;
  g1.outh = net->act[kConv + 1];
  g1.ohm = f1.cout;
  g1.ohn = 1;
  gemm(net, s, g1);
  const Shape& f2 = net->shapes[kConv + 1];
  Gemm g2 = make_gemm(net->act[kConv + 1], f2.fan_in, 1, net->wh + f2.w_off, 1, f2.fan_in, batch, f2.cout,
                      f2.fan_in);
  g2.bias = net->w + f2.b_off;
  g2.outf = twist;
  g2.ofm = kOut;
  g2.ofn = 1;
  g2.scale_out = 1;
  gemm(net, s, g2);
}

}  // namespace

namespace kf {

// The pose proposal on the pipeline's stream for one pair.
void posenet_propose(kf_posenet* net, hipStream_t s, const uint16_t* prev, const uint16_t* cur, float* twist) {
  infer(net, s, prev, cur, 0LL, 1, twist);
}

bool posenet_input_matches(int width, int height) { return width == kInW && height == kInH; }

void posenet_prepare_attach(kf_posenet* net, float depth_scale) {
  if (net->cfg.depth_scale != depth_scale)
    fail(KF_ERR_INVALID_ARGUMENT, "kf_attach_posenet: the network's depth scale differs from the context's");
  KF_HIP(hipStreamSynchronize(net->stream));
}

}  // namespace kf

extern "C" {

void kf_posenet_config_defaults(kf_posenet_config* cfg) {
  if (cfg == nullptr) return;
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->batch_capacity = 64;
  cfg->seed = 0;
  cfg->learning_rate = 1.0e-3F;
  cfg->beta1 = 0.9F;
  cfg->beta2 = 0.999F;
  cfg->epsilon = 1.0e-8F;
  cfg->depth_scale = 1000.0F;
}

int64_t kf_posenet_parameter_count(void) {
  size_t total = 0;
  network_shapes(&total);
  return static_cast<int64_t>(total);
}

kf_status kf_posenet_create(const kf_posenet_config* cfg, kf_posenet** out) {
  if (cfg == nullptr || out == nullptr) return KF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* net = new kf_posenet();
  const kf_status st = net_guard(net, [&] {
    if (cfg->batch_capacity < 1 || cfg->batch_capacity > 4096) fail(KF_ERR_INVALID_ARGUMENT, "batch_capacity");
    if (!(cfg->depth_scale > 0.0F)) fail(KF_ERR_INVALID_ARGUMENT, "depth_scale must be > 0");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) fail(KF_ERR_UNSUPPORTED, "no HIP device");
    net->cfg = *cfg;
    net->shapes = network_shapes(&net->n_params);
    KF_HIP(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
    const size_t P = net->n_params;
    net->w = nalloc<float>(P);
    net->grad = nalloc<float>(P);
    net->m = nalloc<float>(P);
    net->v = nalloc<float>(P);
    net->wh = nalloc<uint16_t>(P);
    const int B = cfg->batch_capacity;
    size_t max_tmp = 0, max_dz = 0, max_m = 0;
    net->act[0] = nalloc<uint16_t>(static_cast<size_t>(B) * kInH * kInW * 2);
    for (int l = 0; l < kLayers - 1; ++l) {
      const Shape& sh = net->shapes[l];
      const size_t M = static_cast<size_t>(B) * pixels(sh);
      net->act[l + 1] = nalloc<uint16_t>(M * sh.cout);
      if (l < kConv) net->col[l] = nalloc<uint16_t>(M * sh.fan_in);
      max_tmp = std::max(max_tmp, M * sh.fan_in);  // dcol / dx of any layer
      max_dz = std::max(max_dz, M * sh.cout);
      max_m = std::max(max_m, M);
    }
    max_dz = std::max(max_dz, static_cast<size_t>(B) * kOut);
    net->out = nalloc<float>(static_cast<size_t>(B) * kOut);
    net->tmp = nalloc<float>(max_tmp);
    net->dz[0] = nalloc<uint16_t>(max_dz);
    net->dz[1] = nalloc<uint16_t>(max_dz);
    net->partial_floats = static_cast<size_t>(8) << 20;
    net->partial = nalloc<float>(net->partial_floats);
    net->colpart = nalloc<float>(((max_m + kColRows - 1) / kColRows) * 256);
    net->loss = nalloc<float>(1);
    KF_HIP(hipHostMalloc(reinterpret_cast<void**>(&net->loss_host), sizeof(float), hipHostMallocDefault));
    std::vector<float> w(P);
    init_weights(net->shapes, cfg->seed, w);
    KF_HIP(hipMemcpyAsync(net->w, w.data(), P * sizeof(float), hipMemcpyHostToDevice, net->stream));
    KF_HIP(hipMemsetAsync(net->m, 0, P * sizeof(float), net->stream));
    KF_HIP(hipMemsetAsync(net->v, 0, P * sizeof(float), net->stream));
    KF_HIP(hipMemsetAsync(net->grad, 0, P * sizeof(float), net->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(P)), dim3(256), 0, net->stream, net->w, net->wh,
                       static_cast<long long>(P));
    KF_HIP(hipGetLastError());
    KF_HIP(hipStreamSynchronize(net->stream));
  });
  if (st != KF_OK) {
    kf_posenet_destroy(net);
    return st;
  }
  *out = net;
  return KF_OK;
}

void kf_posenet_destroy(kf_posenet* net) {
  if (net == nullptr) return;
  if (net->stream != nullptr) (void)hipStreamSynchronize(net->stream);
  for (void* p : {static_cast<void*>(net->w), static_cast<void*>(net->grad), static_cast<void*>(net->m),
                  static_cast<void*>(net->v), static_cast<void*>(net->wh), static_cast<void*>(net->out),
                  static_cast<void*>(net->tmp), static_cast<void*>(net->dz[0]), static_cast<void*>(net->dz[1]),
                  static_cast<void*>(net->partial), static_cast<void*>(net->colpart),
                  static_cast<void*>(net->loss)}) {
    if (p != nullptr) (void)hipFree(p);
  }
  for (auto* p : net->act)
    if (p != nullptr) (void)hipFree(p);
  for (auto* p : net->col)
    if (p != nullptr) (void)hipFree(p);
  if (net->loss_host != nullptr) (void)hipHostFree(net->loss_host);
  if (net->stream != nullptr) (void)hipStreamDestroy(net->stream);
  delete net;
}

const char* kf_posenet_last_error(const kf_posenet* net) {
  return net == nullptr ? "null posenet" : net->error.c_str();
}

kf_status kf_posenet_get_parameters(kf_posenet* net, float* params) {
  return net_guard(net, [&] {
    if (net == nullptr || params == nullptr) fail(KF_ERR_INVALID_ARGUMENT, "null argument");
    KF_HIP(hipMemcpyAsync(params, net->w, net->n_params * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    KF_HIP(hipStreamSynchronize(net->stream));
  });
}

kf_status kf_posenet_set_parameters(kf_posenet* net, const float* params) {
  return net_guard(net, [&] {
    if (net == nullptr || params == nullptr) fail(KF_ERR_INVALID_ARGUMENT, "null argument");
    KF_HIP(hipMemcpyAsync(net->w, params, net->n_params * sizeof(float), hipMemcpyHostToDevice, net->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(net->n_params)), dim3(256), 0, net->stream, net->w, net->wh,
                       static_cast<long long>(net->n_params));
    KF_HIP(hipGetLastError());
    KF_HIP(hipMemsetAsync(net->m, 0, net->n_params * sizeof(float), net->stream));
    KF_HIP(hipMemsetAsync(net->v, 0, net->n_params * sizeof(float), net->stream));
    net->step = 0;
    KF_HIP(hipStreamSynchronize(net->stream));
  });
}

kf_status kf_posenet_forward(kf_posenet* net, const uint16_t* pairs, int32_t batch, float* twists) {
  return net_guard(net, [&] {
    if (net == nullptr || pairs == nullptr || twists == nullptr) fail(KF_ERR_INVALID_ARGUMENT, "null argument");
    check_batch(net, batch);
    const long long hw = static_cast<long long>(kInH) * kInW;
    infer(net, net->stream, pairs, pairs + hw, 2 * hw, batch, twists);
    KF_HIP(hipGetLastError());
    KF_HIP(hipStreamSynchronize(net->stream));
  });
}

kf_status kf_posenet_gradients(kf_posenet* net, const uint16_t* pairs, const float* targets, int32_t batch,
                               float* loss, float* grads) {
  return net_guard(net, [&] {
    if (net == nullptr || pairs == nullptr || targets == nullptr) fail(KF_ERR_INVALID_ARGUMENT, "null argument");
    check_batch(net, batch);
    prepare_pairs(net, net->stream, pairs, batch);
    forward(net, net->stream, batch);
    backward(net, net->stream, targets, batch);
    KF_HIP(hipMemcpyAsync(net->loss_host, net->loss, sizeof(float), hipMemcpyDeviceToHost, net->stream));
    if (grads != nullptr) {
      KF_HIP(hipMemcpyAsync(grads, net->grad, net->n_params * sizeof(float), hipMemcpyDeviceToHost, net->stream));
    }
    KF_HIP(hipStreamSynchronize(net->stream));
    if (loss != nullptr) *loss = *net->loss_host;
  });
}

kf_status kf_posenet_train_step(kf_posenet* net, const uint16_t* pairs, const float* targets, int32_t batch,
                                float* loss) {
  return net_guard(net, [&] {
    if (net == nullptr || pairs == nullptr || targets == nullptr) fail(KF_ERR_INVALID_ARGUMENT, "null argument");
    check_batch(net, batch);
    prepare_pairs(net, net->stream, pairs, batch);
    forward(net, net->stream, batch);
    backward(net, net->stream, targets, batch);
    net->step += 1;
    const double t = static_cast<double>(net->step);
    const float c1 = static_cast<float>(1.0 / (1.0 - std::pow(static_cast<double>(net->cfg.beta1), t)));
    const float c2 = static_cast<float>(1.0 / (1.0 - std::pow(static_cast<double>(net->cfg.beta2), t)));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(net->n_params)), dim3(256), 0, net->stream, net->w, net->grad,
                       net->m, net->v, net->wh, static_cast<long long>(net->n_params), net->cfg.learning_rate,
                       net->cfg.beta1, net->cfg.beta2, net->cfg.epsilon, c1, c2);
    KF_HIP(hipGetLastError());
    if (loss != nullptr) {
      KF_HIP(hipMemcpyAsync(net->loss_host, net->loss, sizeof(float), hipMemcpyDeviceToHost, net->stream));
      KF_HIP(hipStreamSynchronize(net->stream));
      *loss = *net->loss_host;
    }
  });
}

kf_status kf_posenet_set_learning_rate(kf_posenet* net, float learning_rate) {
  return net_guard(net, [&] {
    if (net == nullptr || !(learning_rate > 0.0F)) fail(KF_ERR_INVALID_ARGUMENT, "learning rate must be > 0");
    net->cfg.learning_rate = learning_rate;
  });
}

kf_status kf_posenet_synchronize(kf_posenet* net) {
  return net_guard(net, [&] {
    if (net == nullptr) fail(KF_ERR_INVALID_ARGUMENT, "null posenet");
    KF_HIP(hipStreamSynchronize(net->stream));
  });
}

}  // extern "C"
