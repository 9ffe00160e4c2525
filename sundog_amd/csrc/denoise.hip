// Learned denoiser: a 5-level BF16 U-Net trained by this library's own MFMA
// kernels. It stands in for the reference's vendor denoiser
// (src/denoise.cpp:18-53: HDR beauty + albedo + normal guides in, denoised
// HDR out).
//
// Network (channels 32/64/96/128/160, 3x3 convolutions, ReLU):
//   in(9 -> padded 16) -> e0(32,32) -> pool -> e1(64,64) -> pool -> e2(96,96)
//   -> pool -> e3(128,128) -> pool -> e4(160,160) -> up+cat(e3) -> d3(128,128)
//   -> up+cat(e2) -> d2(96,96) -> up+cat(e1) -> d1(64,64) -> up+cat(e0)
//   -> d0(32,32) -> out(3, padded 32), linear
// Inputs: log(1+max(beauty,0)) RGB, albedo RGB, normal XYZ. Output: the
// log(1+x) of the denoised radiance, predicted as a residual added to the
// noisy input's log(1+x) (so the untrained network starts at the identity).
// Loss: L1 in that log-HDR space.
//
// Activations are NHWC BF16, gradients NHWC FP32, master weights FP32 with
// Adam; BF16 weight copies are kept in two layouts, [co][tap][ci] for the
// forward pass and [ci][8 - tap][co] for backward-data, so every MFMA operand
// a lane needs is 16 contiguous bytes.
//
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16 (FP32
// accumulate): forward M = pixels, N = Cout, K = 9*Cin; backward-data M =
// pixels, N = Cin, K = 9*Cout; backward-weights M = Cout, N = Cin, K = pixels
// per tap, split over pixel chunks whose FP32 partials are reduced in a fixed
// order (deterministic).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sundog_api.h"
#include "session.h"

namespace sdn {

using sdk::DevBuf;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
__host__ __device__ inline uint16_t f2bf(float f) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline bf16x8 asBf(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

constexpr int kTile = 32;  // MFMA 32x32x16

// ---------------------------------------------------------------------------
// forward / backward-data convolution (implicit GEMM)
// ---------------------------------------------------------------------------
// out[p][co] = bias[co] + sum_t sum_ci in[p + off(t)][ci] * w[co][t][ci],
// off(t) = (t / 3 - 1, t % 3 - 1). The backward-data pass is the same
// kernel over the FP32 gradient (HS_F32) with the transposed weights, taps
// mirrored ([ci][8 - t][co], as off(8 - t) = -off(t)).
//
// Tiled in space: a work item is 8 x 16
// output pixels of one image (wave w: rows 2w and 2w+1, 32 pixels) by one
// block of kNT x 32 output channels; a layer has Cout / (32 kNT) channel
// blocks, and the channel blocks of one tile are consecutive items, so they
// run together and read the same halo from L2. Per chunk of up to 32 input
// channels the block stages the 10 x 18 input halo (zero outside the image; each halo row is one
// contiguous run of pixels in memory) and all nine taps of the chunk's
// weights in LDS, then runs the chunk's 9 x (chunk / 16) K steps from LDS
// only: the nine taps of a pixel read the same halo instead of issuing nine
// scattered 16-byte global loads per fragment. K runs chunk-major (chunk,
// tap, channel); each output is the same sum as the direct one up to FP32
// rounding order.
//
// Blocks are persistent: block b takes items b, b + gridDim.x, ... (so the
// blocks resident at one time cover a contiguous band of tiles and share
// halo rows in L2) and walks the steps (item, chunk) in order. The global
// loads of step s + 1 are issued into registers before step s's MFMAs and
// stored to LDS after them, so memory latency overlaps the matrix work. With
// a single chunk and a single channel block the weights are the same for
// every step and are staged once per block.
//
// A halo pixel holds 4 16-byte vectors (64 bytes), a weight row 36; vector v
// of pixel / row i is stored at v ^ ((i >> 2) & 3), so 16 consecutive pixels
// or rows reading the same vector cover all 16 16-byte bank groups.
constexpr int kHaloTH = 8, kHaloTW = 16, kHaloCi = 32;
constexpr int kHaloPix = (kHaloTH + 2) * (kHaloTW + 2);  // 180
constexpr int kHaloRowVecs = 9 * kHaloCi / 8;            // 36 weight vectors per output channel
constexpr int kHaloAV = (kHaloPix * 4 + 255) / 256;      // halo vectors per thread (3)

template <int kNT>
struct HaloStage {
  static constexpr int kBV = (kNT * 32 * kHaloRowVecs + 255) / 256;  // weight vectors per thread
  u16x8 a[kHaloAV];
  u16x8 b[kBV];
};

// Where a halo kernel's input pixels come from, and what its epilogue
// writes. The inference forward reads the pooled and the concatenated
// layer inputs straight from their sources instead of materialising them
// (training keeps them: its backward pass reads them).
enum HaloSrc {
  HS_PLAIN = 0,  // in[P][Cin]
  HS_POOL = 1,   // 2x2 average of in[N][2H][2W][Cin], rounded to BF16 as k_pool_fwd does
  HS_UPCAT = 2,  // channels < ca: in[N][H/2][W/2][ca] (nearest up-sample); the rest: in2[P][Cin - ca]
  HS_F32 = 3,    // inF[P][Cin] FP32 (gradients), rounded to BF16 as it is staged
};
enum HaloOut {
  HO_BF16 = 0,   // out[P][Cout] BF16
  HO_F32 = 1,    // out[P][Cout] FP32
  HO_FINAL = 2,  // out[P][3] FP32 = conv + log1p(max(raw[P][9] beauty, 0)), as k_final does
};

struct HaloArgs {
  const uint16_t* in;
  const uint16_t* in2;
  const float* inF;
  const uint16_t* w;
  const float* bias;
  void* out;
  const float* raw;  // HO_FINAL: the network's FP32 input [P][9]
  int H, W, Cin, Cout, ca, numItems, relu;
};

__device__ __forceinline__ float4 bf4lo(u16x8 v) { return make_float4(bf2f(v[0]), bf2f(v[1]), bf2f(v[2]), bf2f(v[3])); }
__device__ __forceinline__ float4 bf4hi(u16x8 v) { return make_float4(bf2f(v[4]), bf2f(v[5]), bf2f(v[6]), bf2f(v[7])); }

// 8 channels (16 bytes) of input pixel (n, gy, gx), channels c .. c + 7
template <int kSrc>
__device__ __forceinline__ u16x8 haloPixel(const HaloArgs& A, int n, int gy, int gx, int c) {
  if (kSrc == HS_POOL) {
    // the same sum order as k_pool_fwd: (y, x), (y, x+1), (y+1, x), (y+1, x+1)
    const size_t W2 = (size_t)A.W * 2;
    const uint16_t* b = A.in + (((size_t)n * A.H * 2 + 2 * gy) * W2 + 2 * gx) * A.Cin + c;
    const u16x8 p0 = *(const u16x8*)b, p1 = *(const u16x8*)(b + A.Cin);
    const u16x8 p2 = *(const u16x8*)(b + W2 * A.Cin), p3 = *(const u16x8*)(b + W2 * A.Cin + A.Cin);
    u16x8 r;
#pragma unroll
    for (int e = 0; e < 8; e++) r[e] = f2bf(0.25f * (bf2f(p0[e]) + bf2f(p1[e]) + bf2f(p2[e]) + bf2f(p3[e])));
    return r;
  } else if (kSrc == HS_UPCAT) {
    if (c < A.ca)
      return *(const u16x8*)(A.in + (((size_t)n * (A.H / 2) + gy / 2) * (A.W / 2) + gx / 2) * A.ca + c);
    const int cb = A.Cin - A.ca;
    return *(const u16x8*)(A.in2 + (((size_t)n * A.H + gy) * A.W + gx) * cb + (c - A.ca));
  } else if (kSrc == HS_F32) {
    const float* b = A.inF + (((size_t)n * A.H + gy) * A.W + gx) * A.Cin + c;
    const float4 lo = *(const float4*)b, hi = *(const float4*)(b + 4);
    return u16x8{f2bf(lo.x), f2bf(lo.y), f2bf(lo.z), f2bf(lo.w), f2bf(hi.x), f2bf(hi.y), f2bf(hi.z), f2bf(hi.w)};
  } else {
    return *(const u16x8*)(A.in + (((size_t)n * A.H + gy) * A.W + gx) * A.Cin + c);
  }
}

template <int kNT, int kSrc>
__device__ __forceinline__ void haloLoad(HaloStage<kNT>& st, const HaloArgs& A, int n, int y0, int x0, int co0,
                                         int ci0, int cn, bool weights) {
#pragma unroll
  for (int k = 0; k < kHaloAV; k++) {
    const int i = (int)threadIdx.x + k * 256;
    const int pix = i >> 2, v = i & 3;
    const int gy = y0 + pix / (kHaloTW + 2) - 1, gx = x0 + pix % (kHaloTW + 2) - 1;
    st.a[k] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    if (i < kHaloPix * 4 && v * 8 < cn && gy >= 0 && gy < A.H && gx >= 0 && gx < A.W)
      st.a[k] = haloPixel<kSrc>(A, n, gy, gx, ci0 + v * 8);
  }
  if (!weights) return;
#pragma unroll
  for (int k = 0; k < HaloStage<kNT>::kBV; k++) {
    const int i = (int)threadIdx.x + k * 256;
    const int r = i / kHaloRowVecs, q = i - r * kHaloRowVecs;  // q = tap * 4 + vector
    const int t = q >> 2, v = q & 3;
    st.b[k] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    if (i < kNT * 32 * kHaloRowVecs && v * 8 < cn)
      st.b[k] = *(const u16x8*)(A.w + ((size_t)(co0 + r) * 9 + t) * A.Cin + ci0 + v * 8);
  }
}

template <int kNT>
__device__ __forceinline__ void haloStore(const HaloStage<kNT>& st, u16x8* sA, u16x8* sB, bool weights) {
#pragma unroll
  for (int k = 0; k < kHaloAV; k++) {
    const int i = (int)threadIdx.x + k * 256;
    const int pix = i >> 2, v = i & 3;
    if (i < kHaloPix * 4) sA[pix * 4 + (v ^ ((pix >> 2) & 3))] = st.a[k];
  }
  if (!weights) return;
#pragma unroll
  for (int k = 0; k < HaloStage<kNT>::kBV; k++) {
    const int i = (int)threadIdx.x + k * 256;
    const int r = i / kHaloRowVecs, q = i - r * kHaloRowVecs;
    if (i < kNT * 32 * kHaloRowVecs) sB[r * kHaloRowVecs + (q ^ ((r >> 2) & 3))] = st.b[k];
  }
}

// Cout % (32 * kNT) == 0; Cin % 16 == 0 (HS_UPCAT: ca % 32 == 0);
// gridDim.x <= numItems = N x tilesY x tilesX x Cout / (32 * kNT)
template <int kOut, int kNT, int kSrc>
__global__ __launch_bounds__(256, kNT == 1 ? 4 : 3) void k_conv_halo(HaloArgs A) {
  const int H = A.H, W = A.W, Cin = A.Cin, Cout = A.Cout;
  const int coBlocks = Cout / (32 * kNT);
  __shared__ u16x8 sA[kHaloPix * 4];
  __shared__ u16x8 sB[kNT * 32 * kHaloRowVecs];
  __shared__ float sF[kOut == HO_FINAL ? 4 * 96 : 1];  // HO_FINAL: the output channels 0..2 per wave
  const int tilesX = (W + kHaloTW - 1) / kHaloTW, tilesY = (H + kHaloTH - 1) / kHaloTH;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 31, half = lane >> 5;
  const int py = 2 * wave + (row >> 4), px = row & 15;  // this lane's A row (tile coordinates)
  const int nc = (Cin + kHaloCi - 1) / kHaloCi;
  const int myItems = (A.numItems - (int)blockIdx.x + (int)gridDim.x - 1) / (int)gridDim.x;
  const int steps = myItems * nc;
  const bool fixedW = nc == 1 && coBlocks == 1;  // the same weights for every step
  // (image, tile origin, first output channel) of step s
  auto origin = [&](int s, int& n, int& y0, int& x0, int& co0) {
    const int item = (int)blockIdx.x + (s / nc) * (int)gridDim.x;
    const int tile = item / coBlocks;
    co0 = (item - tile * coBlocks) * 32 * kNT;
    n = tile / (tilesX * tilesY);
    const int ty = tile - n * tilesX * tilesY;
    y0 = (ty / tilesX) * kHaloTH;
    x0 = (ty % tilesX) * kHaloTW;
  };
  HaloStage<kNT> st;
  {
    int n, y0, x0, co0;
    origin(0, n, y0, x0, co0);
    haloLoad<kNT, kSrc>(st, A, n, y0, x0, co0, 0, min(kHaloCi, Cin), true);
  }
  f32x16 acc[kNT];
#pragma unroll
  for (int j = 0; j < kNT; j++) acc[j] = f32x16{};
  for (int s = 0; s < steps; s++) {
    const int c = s % nc, ci0 = c * kHaloCi, cn = min(kHaloCi, Cin - ci0);
    __syncthreads();  // step s - 1's MFMAs are done with the LDS
    haloStore<kNT>(st, sA, sB, !fixedW || s == 0);
    __syncthreads();
    if (s + 1 < steps) {
      int n1, y1, x1, co1;
      origin(s + 1, n1, y1, x1, co1);
      const int c1 = (s + 1) % nc;
      haloLoad<kNT, kSrc>(st, A, n1, y1, x1, co1, c1 * kHaloCi, min(kHaloCi, Cin - c1 * kHaloCi), !fixedW);
    }
#pragma unroll 3
    for (int t = 0; t < 9; t++) {
      const int hp = (py + t / 3) * (kHaloTW + 2) + px + t % 3;
      for (int kk = 0; kk < cn; kk += 16) {
        const int va = (kk >> 3) + half;
        const u16x8 a = sA[hp * 4 + (va ^ ((hp >> 2) & 3))];
        const int q = t * 4 + va;
#pragma unroll
        for (int j = 0; j < kNT; j++) {
          const int r = j * 32 + row;
          const u16x8 b = sB[r * kHaloRowVecs + (q ^ ((r >> 2) & 3))];
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(asBf(a), asBf(b), acc[j], 0, 0, 0);
        }
      }
    }
    if (c == nc - 1) {
      int n, y0, x0, co0;
      origin(s, n, y0, x0, co0);
#pragma unroll
      for (int j = 0; j < kNT; j++) {
        const int co = co0 + j * kTile + row;  // D column = lane % 32
        const float bv = A.bias ? A.bias[co] : 0.0f;
        if (kOut == HO_FINAL) {
          // channels 0..2 of the wave's 32 pixels go through LDS so that all
          // 64 lanes share the 96 log1p evaluations and stores
          if (co < 3) {
#pragma unroll
            for (int r = 0; r < 16; r++) {
              const int i = 8 * (r / 4) + 4 * half + (r % 4);
              float v = acc[j][r] + bv;
              if (A.relu) v = fmaxf(v, 0.0f);
              sF[wave * 96 + i * 3 + co] = v;
            }
          }
          __syncthreads();
          for (int e = lane; e < 96; e += 64) {
            const int i = e / 3, ch = e - (e / 3) * 3;
            const int oy = y0 + 2 * wave + (i >> 4), ox = x0 + (i & 15);
            if (oy < H && ox < W) {
              const size_t pp = ((size_t)n * H + oy) * W + ox;
              ((float*)A.out)[pp * 3 + ch] = sF[wave * 96 + e] + log1pf(fmaxf(A.raw[pp * 9 + ch], 0.0f));
            }
          }
        } else if (kOut == HO_BF16) {
          // through LDS (the halo buffer, free once every wave is past this
          // step's MFMAs) so that each lane stores 16-byte vectors: the
          // wave's 32 pixels x 32 channels are 128 vectors, 2 per lane,
          // instead of 16 two-byte stores per lane
          __syncthreads();
          uint16_t* tb = reinterpret_cast<uint16_t*>(sA) + wave * (32 * 32);
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int i = 8 * (r / 4) + 4 * half + (r % 4);  // D row: this wave's i-th pixel
            float v = acc[j][r] + bv;
            if (A.relu) v = fmaxf(v, 0.0f);
            tb[i * 32 + row] = f2bf(v);
          }
          __syncthreads();
#pragma unroll
          for (int k = 0; k < 2; k++) {
            const int e = lane + 64 * k, i = e >> 2, v = e & 3;
            const int oy = y0 + 2 * wave + (i >> 4), ox = x0 + (i & 15);
            if (oy < H && ox < W)
              *(u16x8*)((uint16_t*)A.out + (((size_t)n * H + oy) * W + ox) * Cout + co0 + j * kTile + v * 8) =
                  reinterpret_cast<const u16x8*>(tb)[e];
          }
        } else {
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int i = 8 * (r / 4) + 4 * half + (r % 4);  // D row: this wave's i-th pixel
            const int oy = y0 + 2 * wave + (i >> 4), ox = x0 + (i & 15);
            if (oy < H && ox < W) {
              const size_t pp = ((size_t)n * H + oy) * W + ox;
              float v = acc[j][r] + bv;
              if (A.relu) v = fmaxf(v, 0.0f);
              ((float*)A.out)[pp * Cout + co] = v;
            }
          }
        }
        acc[j] = f32x16{};
      }
    }
  }
}

template <int kOut, int kNT, int kSrc>
static int haloOccupancy() {
  int per = 0;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_conv_halo<kOut, kNT, kSrc>, 256, 0));
  return per;
}

// Blocks of k_conv_halo<kOut, kNT, kSrc> resident on the current device at
// once (the persistent grid size), cached per device (devices 0..63; others
// are queried on every call).
template <int kOut, int kNT, int kSrc>
static int haloResident() {
  static std::atomic<int> cache[64];
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev >= 0 && dev < 64) {
    const int v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
  }
  int cus = 0;
  HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const int v = std::max(1, cus * std::max(1, haloOccupancy<kOut, kNT, kSrc>()));
  if (dev >= 0 && dev < 64) cache[dev].store(v, std::memory_order_relaxed);
  return v;
}

template <int kOut, int kNT, int kSrc>
static void launchHalo(const HaloArgs& A, hipStream_t s) {
  const int g = std::min(A.numItems, haloResident<kOut, kNT, kSrc>());
  hipLaunchKernelGGL((k_conv_halo<kOut, kNT, kSrc>), dim3(g), dim3(256), 0, s, A);
}

// ---------------------------------------------------------------------------
// backward-weights (implicit GEMM: M = Cout, N = 3 taps x 32 input channels
// per wave, K = output pixels):
//   dW[co][t][ci] = sum_p dy[p][co] * x[p + off(t)][ci],  db[co] = sum_p dy[p][co]
//
// Output pixels are walked in tiles of kTC columns x (64 / kTC) rows of one
// image; columns and rows past the image edge carry dy = 0, so any H x W
// works. A block owns one 32-channel input slice and a run of consecutive
// tiles (its K chunk) and writes FP32 partial sums, one row of
// [Cout][9][Cin] weights followed by [Cout] biases per chunk; the chunks are
// summed afterwards in a fixed order (deterministic).
//
// Per tile the block stages in LDS
//   dys[co][64 px]               dy rounded to BF16, pixel-contiguous per co
//   xs[dx + 1][ci][halo row][c]  the input slice shifted by dx = -1, 0, +1
// so that each MFMA operand a lane needs (8 pixels of one channel, for both
// A = dy^T and B = x) is one aligned 16-byte LDS read. Both strides carry 16
// bytes of padding, which spreads every 8 lanes of a 16-byte access over
// distinct banks. Wave w = 3 * coTile + tapRow forms the 32x32 tiles of its
// co tile for the three taps of its tap row (dy = tapRow - 1): per 16-pixel
// K step one A read and three B reads feed three MFMAs. The next tile's
// global loads (16-byte loads of four dy channels or eight input channels
// of one pixel, consecutive lanes at consecutive pixels) are issued into
// registers before the current tile's MFMAs and transposed into the LDS
// layouts with 2-byte stores. Each staging slot of a thread always holds the
// same pixel of the tile and the same four dy channels, so it sums them in
// FP32 for the bias in tile order; a fixed butterfly over a wave's 64 pixel
// slots then gives each channel's sum.
// ---------------------------------------------------------------------------
template <int kTC, int kNco>
struct DwTile {
  static constexpr int kRows = 64 / kTC;            // output rows per tile
  static constexpr int kHalo = kRows + 2;           // input rows per tile
  static constexpr int kXStride = kHalo * kTC + 8;  // BF16 per (dx, ci) plane
  static constexpr int kDyStride = 64 + 8;          // BF16 per co row
  static constexpr int kCout = 32 * kNco;
  static constexpr int kThreads = 192 * kNco;       // 3 waves per co tile = 6 * kCout
  // staging: 16-byte loads, pixel-fastest across lanes
  static constexpr int kCo4 = kCout / 4;            // float4 groups of a pixel's dy
  static constexpr int kDyVec = 64 * kCo4;          // dy float4 loads per tile
  static constexpr int kDyTasks = (kDyVec + kThreads - 1) / kThreads;
  static constexpr int kXCols = kTC + 2;            // input columns x0 - 1 .. x0 + kTC
  static constexpr int kXVec = kHalo * kXCols * 4;  // x loads of 8 channels per tile
  static constexpr int kXTasks = (kXVec + kThreads - 1) / kThreads;
  static_assert(kThreads % 64 == 0, "a wave's slots each hold one channel group");
};

struct DwArgs {
  const float* dy;     // [N][H][W][Cout] FP32
  const uint16_t* x;   // [N][H][W][Cin] BF16, Cin a multiple of 8
  float* partial;      // [chunks][Cout * 9 * Cin + Cout]
  int N, H, W, Cin, Cout;
  int tilesX, tilesY, tiles, tilesPerChunk;
};

template <int kTC, int kNco>
struct DwRegs {
  using T = DwTile<kTC, kNco>;
  float4 dy[T::kDyTasks];  // four channels of one pixel
  uint4 x[T::kXTasks];     // eight BF16 channels of one input pixel
};

// Loads one tile's operands into registers with 16-byte loads. dy task k
// (of kDyVec): pixel k % 64 of the tile, channels 4 (k / 64) .. +3, so a
// wave's lanes read 64 pixels of one channel group; x task k (of kXVec):
// column k % kXCols of halo row k / (4 kXCols), channels ci0 + 8 ((k /
// kXCols) % 4) .. +7. Out-of-image pixels and channels past Cin load zeros.
template <int kTC, int kNco>
__device__ __forceinline__ void dwLoad(const DwArgs& A, int tile, int ci0, DwRegs<kTC, kNco>& R) {
  using T = DwTile<kTC, kNco>;
  const int tid = threadIdx.x;
  const int perImage = A.tilesX * A.tilesY;
  const int n = tile / perImage, rem = tile - n * perImage;
  const int y0 = (rem / A.tilesX) * T::kRows, x0 = (rem % A.tilesX) * kTC;
#pragma unroll
  for (int i = 0; i < T::kDyTasks; i++) {
    const int task = tid + i * T::kThreads;
    const int px = task & 63, co4 = task >> 6;
    const int y = y0 + px / kTC, c = x0 + px % kTC;
    const bool ok = task < T::kDyVec && y < A.H && c < A.W;
    R.dy[i] = ok ? *reinterpret_cast<const float4*>(A.dy + ((size_t)(n * A.H + y) * A.W + c) * T::kCout + 4 * co4)
                 : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
#pragma unroll
  for (int i = 0; i < T::kXTasks; i++) {
    const int task = tid + i * T::kThreads;
    const int col = task % T::kXCols, rest = task / T::kXCols;
    const int cg = rest & 3, hr = rest >> 2;
    const int y = y0 - 1 + hr, c = x0 - 1 + col, ci = ci0 + 8 * cg;
    const bool ok = task < T::kXVec && ci < A.Cin && y >= 0 && y < A.H && c >= 0 && c < A.W;
    R.x[i] = ok ? *reinterpret_cast<const uint4*>(A.x + ((size_t)(n * A.H + y) * A.W + c) * A.Cin + ci)
                : make_uint4(0u, 0u, 0u, 0u);
  }
}

// Writes the registers into the LDS layouts the MFMAs read (dys[co][px] in
// BF16; xs[dx][ci][halo row][col], the column shifted by dx = -1, 0, +1 in
// the three copies) with 2-byte stores, consecutive lanes at consecutive
// pixels, and adds each dy value to its slot's FP32 bias sum (a slot always
// holds the same pixel and channels).
template <int kTC, int kNco>
__device__ __forceinline__ void dwStore(const DwRegs<kTC, kNco>& R, uint16_t* dys, uint16_t* xs,
                                        float (&bias)[DwTile<kTC, kNco>::kDyTasks][4]) {
  using T = DwTile<kTC, kNco>;
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < T::kDyTasks; i++) {
    const int task = tid + i * T::kThreads;
    if (task >= T::kDyVec) continue;
    const int px = task & 63, co4 = task >> 6;
    const float v[4] = {R.dy[i].x, R.dy[i].y, R.dy[i].z, R.dy[i].w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      dys[(4 * co4 + k) * T::kDyStride + px] = f2bf(v[k]);
      bias[i][k] += v[k];
    }
  }
#pragma unroll
  for (int i = 0; i < T::kXTasks; i++) {
    const int task = tid + i * T::kThreads;
    if (task >= T::kXVec) continue;
    const int col = task % T::kXCols, rest = task / T::kXCols;
    const int cg = rest & 3, hr = rest >> 2;
    const uint32_t w[4] = {R.x[i].x, R.x[i].y, R.x[i].z, R.x[i].w};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int cOut = col - d;  // copy d holds input column x0 - 1 + cOut + d at column cOut
      if (cOut < 0 || cOut >= kTC) continue;
      uint16_t* dst = xs + (d * 32 + 8 * cg) * T::kXStride + hr * kTC + cOut;
#pragma unroll
      for (int k = 0; k < 8; k++) dst[k * T::kXStride] = (uint16_t)(w[k >> 1] >> (16 * (k & 1)));
    }
  }
}

template <int kTC, int kNco>
__global__ __launch_bounds__(192 * kNco) void k_conv_dw(DwArgs A) {
  using T = DwTile<kTC, kNco>;
  __shared__ __align__(16) uint16_t dys[T::kCout * T::kDyStride];
  __shared__ __align__(16) uint16_t xs[3 * 32 * T::kXStride];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = lane & 31, half = lane >> 5;
  const int coTile = wave / 3, tapRow = wave % 3;
  const int chunk = blockIdx.x, ci0 = blockIdx.y * 32;
  const int t0 = chunk * A.tilesPerChunk, t1 = min(A.tiles, t0 + A.tilesPerChunk);
  f32x16 acc[3] = {};
  float bias[T::kDyTasks][4] = {};
  DwRegs<kTC, kNco> R;
  dwLoad<kTC, kNco>(A, t0, ci0, R);
  for (int tile = t0; tile < t1; tile++) {
    if (tile > t0) __syncthreads();  // the previous tile's operands are consumed
    dwStore<kTC, kNco>(R, dys, xs, bias);
    __syncthreads();
    if (tile + 1 < t1) dwLoad<kTC, kNco>(A, tile + 1, ci0, R);
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
      const int s = 2 * ks + half;  // this half-wave's 8-pixel segment
      const int r = s / (kTC / 8), cs = s % (kTC / 8);
      const u16x8 a = *(const u16x8*)(dys + (coTile * 32 + row) * T::kDyStride + s * 8);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const u16x8 b = *(const u16x8*)(xs + (d * 32 + row) * T::kXStride + (r + tapRow) * kTC + 8 * cs);
        acc[d] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(asBf(a), asBf(b), acc[d], 0, 0, 0);
      }
    }
  }
  // D[i = co offset][j = ci offset]; lane holds column j = row
  const size_t nW = (size_t)T::kCout * 9 * A.Cin;
  float* dst = A.partial + (size_t)chunk * (nW + T::kCout);
  const int ci = ci0 + row;
  if (ci < A.Cin) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int t = 3 * tapRow + d;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int i = 8 * (r / 4) + 4 * half + (r % 4);
        dst[((size_t)(coTile * 32 + i) * 9 + t) * A.Cin + ci] = acc[d][r];
      }
    }
  }
  if (blockIdx.y == 0) {
    // A wave's lanes hold the 64 pixels of one channel group in each slot
    // (slot i of wave w: group w + i * kThreads / 64): a fixed-order
    // butterfly over the lanes gives the group's four sums.
#pragma unroll
    for (int i = 0; i < T::kDyTasks; i++) {
      const int co4 = wave + i * (T::kThreads / 64);
      if (co4 >= T::kCo4) continue;  // uniform per wave
#pragma unroll
      for (int k = 0; k < 4; k++) {
        float v = bias[i][k];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) dst[nW + 4 * co4 + k] = v;
      }
    }
  }
}

// out[e] = sum over chunks k of partial[k][e] (rows of n floats) in a fixed
// order: eight interleaved chunk groups summed in ascending chunk order, then
// the group sums in ascending group order. A block covers 32 elements.
__global__ __launch_bounds__(256) void k_reduce_chunks(const float* __restrict__ partial, float* __restrict__ out,
                                                       int chunks, size_t n) {
  __shared__ float part[8][32];
  const int lane = threadIdx.x & 31, g = threadIdx.x >> 5;
  const size_t e = (size_t)blockIdx.x * 32 + lane;
  float s = 0.0f;
  if (e < n) {
#pragma unroll 4
    for (int k = g; k < chunks; k += 8) s += partial[(size_t)k * n + e];
  }
  part[g][lane] = s;
  __syncthreads();
  if (g == 0 && e < n) {
    float t = part[0][lane];
#pragma unroll
    for (int q = 1; q < 8; q++) t += part[q][lane];
    out[e] = t;
  }
}

// ---------------------------------------------------------------------------
// elementwise / resampling kernels
// ---------------------------------------------------------------------------
__global__ void k_relu_bwd(float* __restrict__ g, const uint16_t* __restrict__ y, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(bf2f(y[i]) > 0.0f)) g[i] = 0.0f;
}

__global__ void k_pool_fwd(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int N, int H, int W, int C) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  int Ho = H / 2, Wo = W / 2;
  size_t total = (size_t)N * Ho * Wo * C;
  if (i >= total) return;
  int c = i % C;
  size_t p = i / C;
  int xo = p % Wo, yo = (p / Wo) % Ho, n = p / ((size_t)Wo * Ho);
  size_t b = (((size_t)n * H + 2 * yo) * W + 2 * xo) * C + c;
  float s = bf2f(in[b]) + bf2f(in[b + C]) + bf2f(in[b + (size_t)W * C]) + bf2f(in[b + (size_t)W * C + C]);
  out[i] = f2bf(0.25f * s);
}

// dIn[y][x] = 0.25 * dOut[y/2][x/2]  (accumulates into dIn)
__global__ void k_pool_bwd(const float* __restrict__ gout, float* __restrict__ gin, int N, int H, int W, int C) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)N * H * W * C;
  if (i >= total) return;
  int c = i % C;
  size_t p = i / C;
  int x = p % W, y = (p / W) % H, n = p / ((size_t)W * H);
  gin[i] += 0.25f * gout[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

// out = concat(up2(a) [Ca channels], skip [Cb channels]) at the skip's size
// One thread per 8 channels (16 bytes) of the output; Ca % 8 == Cb % 8 == 0
// and N * H * W < 2^31.
__global__ void k_upcat_fwd(const uint16_t* __restrict__ a, const uint16_t* __restrict__ skip,
                            uint16_t* __restrict__ out, int N, int H, int W, int Ca, int Cb) {
  const int Ca8 = Ca / 8, Cb8 = Cb / 8, C8 = Ca8 + Cb8;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)N * H * W * C8) return;
  const int p = (int)(i / C8), c8 = (int)(i - (size_t)p * C8);
  const int x = p % W, yn = p / W, y = yn % H, n = yn / H;
  u16x8 v;
  if (c8 < Ca8) v = ((const u16x8*)a)[((size_t)(n * (H / 2) + y / 2) * (W / 2) + x / 2) * Ca8 + c8];
  else v = ((const u16x8*)skip)[(size_t)p * Cb8 + (c8 - Ca8)];
  ((u16x8*)out)[i] = v;
}

// gA[y/2][x/2] += sum over the 4 children (gather form, no atomics);
// gSkip[p] += the skip part
__global__ void k_upcat_bwd_a(const float* __restrict__ g, float* __restrict__ ga, int N, int H, int W, int Ca,
                              int Cb) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  int Ho = H / 2, Wo = W / 2, C = Ca + Cb;
  size_t total = (size_t)N * Ho * Wo * Ca;
  if (i >= total) return;
  int c = i % Ca;
  size_t p = i / Ca;
  int xo = p % Wo, yo = (p / Wo) % Ho, n = p / ((size_t)Wo * Ho);
  size_t b = (((size_t)n * H + 2 * yo) * W + 2 * xo) * C + c;
  ga[i] += g[b] + g[b + C] + g[b + (size_t)W * C] + g[b + (size_t)W * C + C];
}
__global__ void k_upcat_bwd_skip(const float* __restrict__ g, float* __restrict__ gs, int N, int H, int W, int Ca,
                                 int Cb) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t total = (size_t)N * H * W * Cb;
  if (i >= total) return;
  int c = i % Cb;
  size_t p = i / Cb;
  gs[i] += g[p * (Ca + Cb) + Ca + c];
}

// network input: log1p(beauty), albedo, normal, zero pad to 16 channels
__global__ void k_pack_input(const float* __restrict__ src, uint16_t* __restrict__ dst, size_t P) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * 16) return;
  size_t p = i / 16;
  int c = i % 16;
  float v = 0.0f;
  if (c < 3) v = log1pf(fmaxf(src[p * 9 + c], 0.0f));
  else if (c < 9) v = src[p * 9 + c];
  dst[i] = f2bf(v);
}

// Network input from device films (float4 beauty / albedo / normal of a
// w x h frame) at the padded size W x H (multiples of 16); padded pixels
// repeat the nearest edge pixel.
__global__ void k_pack_films(const float4* __restrict__ beauty, const float4* __restrict__ albedo,
                             const float4* __restrict__ normal, int w, int h, int W, int H,
                             float* __restrict__ dst) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)W * H) return;
  int x = min((int)(i % W), w - 1), y = min((int)(i / W), h - 1);
  size_t s = (size_t)y * w + x;
  float4 b = beauty[s], a = albedo[s], n = normal[s];
  float* d = dst + i * 9;
  d[0] = b.x; d[1] = b.y; d[2] = b.z;
  d[3] = a.x; d[4] = a.y; d[5] = a.z;
  d[6] = n.x; d[7] = n.y; d[8] = n.z;
}

// Denoised linear HDR of the valid w x h region as float4 (alpha 1).
__global__ void k_unpack_films(const float* __restrict__ fin, int w, int h, int W, float4* __restrict__ out) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)w * h) return;
  int x = (int)(i % w), y = (int)(i / w);
  const float* f = fin + ((size_t)y * W + x) * 3;
  out[i] = make_float4(expm1f(f[0]), expm1f(f[1]), expm1f(f[2]), 1.0f);
}

// Final log-space output: the network predicts a residual on top of the
// noisy input, out[p][c] = net[p][c] + log1p(max(beauty[p][c], 0)), c < 3.
__global__ void k_final(const float* __restrict__ net, const float* __restrict__ src, float* __restrict__ out,
                        size_t P) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * 3) return;
  size_t p = i / 3;
  int c = i % 3;
  out[i] = net[p * 32 + c] + log1pf(fmaxf(src[p * 9 + c], 0.0f));
}

// L1 loss in log space on the final output [P][3]; writes the gradient
// w.r.t. the network output [P][32] (sign / count on the 3 used channels,
// zero elsewhere) and per-block partial loss sums
__global__ void k_l1(const float* __restrict__ fin, const float* __restrict__ target, float* __restrict__ grad,
                     float* __restrict__ partial, size_t P) {
  __shared__ float s[256];
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < P * 32) {
    size_t p = i / 32;
    int c = i % 32;
    float g = 0.0f;
    if (c < 3) {
      float d = fin[p * 3 + c] - log1pf(fmaxf(target[p * 3 + c], 0.0f));
      l = fabsf(d);
      g = (d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f)) / (float)(P * 3);
    }
    grad[i] = g;
  }
  s[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void k_unpack_output(const float* __restrict__ fin, float* __restrict__ out, size_t P) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * 3) return;
  out[i] = expm1f(fin[i]);
}

// Where each convolution's weights sit in the flat parameter vector
// ([co][t][ci], Cin padded, followed by bias[co]) and its two BF16 copies.
struct WeightLayout {
  size_t wOff, wEnd;  // weights: [wOff, wEnd)
  int cin, cout;
  uint16_t* wf;       // [co][t][ci]
  uint16_t* wt;       // [ci][8 - t][co]
};

// the BF16 copies of parameter i (with value w) if it is a weight; the
// table is in parameter order
__device__ __forceinline__ void storeLayouts(const WeightLayout* __restrict__ L, int count, size_t i, float w) {
  for (int k = 0; k < count; k++) {
    if (i < L[k].wOff) return;  // the previous convolution's bias
    if (i < L[k].wEnd) {
      const size_t e = i - L[k].wOff;
      const int cin = L[k].cin;
      const int ci = (int)(e % cin), t = (int)((e / cin) % 9), co = (int)(e / ((size_t)9 * cin));
      const uint16_t b = f2bf(w);
      L[k].wf[e] = b;
      L[k].wt[((size_t)ci * 9 + (8 - t)) * L[k].cout + co] = b;  // tap 8 - t has the opposite offset
      return;
    }
  }
}

// Adam on the FP32 masters, then the BF16 layouts of the updated weights
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, size_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2, const WeightLayout* __restrict__ L, int count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  storeLayouts(L, count, i, pi);
}

__global__ void k_to_bf16_layouts(const float* __restrict__ p, size_t n, const WeightLayout* __restrict__ L,
                                  int count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) storeLayouts(L, count, i, p[i]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

#define DN_OK(expr) HIP_OK(expr)

static inline unsigned blocks(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// target block count of a backward-weights launch (split over pixel chunks)
constexpr int kDwBlocks = 512;
constexpr size_t kDwPartialBytes = 16u << 20;

struct Conv {
  int cin, cout;  // cin padded to a multiple of 16, cout to a multiple of 32
  bool relu;
  size_t wOff, bOff;  // into the flat parameter vector
  DevBuf wf, wt;      // BF16 layouts
};

struct Net {
  std::vector<Conv> convs;
  size_t nParams = 0;
  DevBuf params, grads, adamM, adamV;
  DevBuf layouts;  // WeightLayout per conv, in parameter order
  int step = 0;

  void addConv(int cin, int cout, bool relu) {
    Conv c;
    c.cin = cin;
    c.cout = cout;
    c.relu = relu;
    c.wOff = nParams;
    nParams += (size_t)cout * 9 * cin;
    c.bOff = nParams;
    nParams += cout;
    convs.push_back(std::move(c));
  }
};

// channels per level and the conv list in execution order
static const int kCh[5] = {32, 64, 96, 128, 160};

}  // namespace sdn

using namespace sdn;

struct sundog_denoiser {
  Net net;
  hipStream_t stream = nullptr;
  // per-resolution scratch
  int N = 0, H = 0, W = 0;
  std::vector<DevBuf> y;           // conv outputs (BF16)
  std::vector<int> yH, yW;         // spatial size per conv
  DevBuf input, pooled[4], cat[4], pred, fin;
  std::vector<DevBuf> gy;          // gradients w.r.t. conv outputs (FP32)
  DevBuf gPool[4], gCat[4], gInput, gPred, partial, lossPartial, target, devIn;
  size_t partialCap = 0;

  ~sundog_denoiser() {
    if (stream) (void)hipStreamDestroy(stream);
  }

  void init(uint64_t seed) {
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    Net& n = net;
    int prev = 16;
    for (int l = 0; l < 5; l++) {
      n.addConv(prev, kCh[l], true);
      n.addConv(kCh[l], kCh[l], true);
      prev = kCh[l];
    }
    for (int l = 3; l >= 0; l--) {
      n.addConv(kCh[l + 1] + kCh[l], kCh[l], true);
      n.addConv(kCh[l], kCh[l], true);
    }
    n.addConv(kCh[0], 32, false);  // output: first 3 channels used
    // He-normal initialisation from a PCG32 stream (Box-Muller)
    std::vector<float> p(n.nParams, 0.0f);
    uint64_t state = 0, inc = (54u << 1) | 1u;
    auto next = [&]() {
      uint64_t old = state;
      state = old * 6364136223846793005ULL + inc;
      uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u), rot = (uint32_t)(old >> 59u);
      return (xs >> rot) | (xs << ((32u - rot) & 31u));
    };
    next();
    state += seed;
    next();
    for (size_t ci = 0; ci < n.convs.size(); ci++) {
      const Conv& c = n.convs[ci];
      float stdv = sqrtf(2.0f / (9.0f * c.cin));
      if (ci == n.convs.size() - 1) stdv *= 0.1f;
      for (size_t k = 0; k < (size_t)c.cout * 9 * c.cin; k += 2) {
        float u1 = ((next() >> 8) + 1) * 0x1p-24f, u2 = (next() >> 8) * 0x1p-24f;
        float r = sqrtf(-2.0f * logf(u1));
        p[c.wOff + k] = stdv * r * cosf(6.2831853f * u2);
        if (k + 1 < (size_t)c.cout * 9 * c.cin) p[c.wOff + k + 1] = stdv * r * sinf(6.2831853f * u2);
      }
    }
    n.params.upload(p, stream);
    n.grads.alloc(n.nParams * 4);
    n.adamM.alloc(n.nParams * 4);
    n.adamV.alloc(n.nParams * 4);
    HIP_OK(hipMemsetAsync(n.adamM.p, 0, n.nParams * 4, stream));
    HIP_OK(hipMemsetAsync(n.adamV.p, 0, n.nParams * 4, stream));
    std::vector<WeightLayout> table;
    for (Conv& c : n.convs) {
      c.wf.alloc((size_t)c.cout * 9 * c.cin * 2);
      c.wt.alloc((size_t)c.cout * 9 * c.cin * 2);
      table.push_back({c.wOff, c.wOff + (size_t)c.cout * 9 * c.cin, c.cin, c.cout, c.wf.as<uint16_t>(),
                       c.wt.as<uint16_t>()});
    }
    n.layouts.upload(table, stream);
    syncWeights();
    HIP_OK(hipStreamSynchronize(stream));
  }

  void syncWeights() {
    hipLaunchKernelGGL(k_to_bf16_layouts, dim3(blocks(net.nParams)), dim3(256), 0, stream, net.params.as<float>(),
                       net.nParams, net.layouts.as<WeightLayout>(), (int)net.convs.size());
    HIP_OK(hipGetLastError());
  }

  void resize(int n, int h, int w) {
    if (n == N && h == H && w == W) return;
    if (h % 16 || w % 16) throw std::runtime_error("denoiser: height and width must be multiples of 16");
    N = n;
    H = h;
    W = w;
    size_t P0 = (size_t)n * h * w;
    size_t nc = net.convs.size();
    y.clear();
    y.resize(nc);
    gy.clear();
    gy.resize(nc);
    yH.assign(nc, 0);
    yW.assign(nc, 0);
    int ci = 0;
    for (int l = 0; l < 5; l++)
      for (int k = 0; k < 2; k++, ci++) {
        yH[ci] = h >> l;
        yW[ci] = w >> l;
      }
    for (int l = 3; l >= 0; l--)
      for (int k = 0; k < 2; k++, ci++) {
        yH[ci] = h >> l;
        yW[ci] = w >> l;
      }
    yH[ci] = h;
    yW[ci] = w;
    for (size_t k = 0; k < nc; k++) {
      size_t px = (size_t)n * yH[k] * yW[k];
      if (k + 1 < nc) y[k].alloc(px * net.convs[k].cout * 2);
      gy[k].alloc(px * net.convs[k].cout * 4);
    }
    input.alloc(P0 * 16 * 2);
    gInput.alloc(P0 * 16 * 4);
    pred.alloc(P0 * 32 * 4);
    fin.alloc(P0 * 3 * 4);
    for (int l = 0; l < 4; l++) {
      size_t pl = (size_t)n * (h >> (l + 1)) * (w >> (l + 1));
      pooled[l].alloc(pl * kCh[l] * 2);
      gPool[l].alloc(pl * kCh[l] * 4);
      size_t pc = (size_t)n * (h >> l) * (w >> l);
      cat[l].alloc(pc * (kCh[l + 1] + kCh[l]) * 2);
      gCat[l].alloc(pc * (kCh[l + 1] + kCh[l]) * 4);
    }
    target.alloc(P0 * 3 * 4);
    devIn.alloc(P0 * 9 * 4);
    gPred.alloc(P0 * 32 * 4);
    lossPartial.alloc(blocks(P0 * 32) * 4);
  }

  // forward convolution k: the spatially tiled kernel with persistent
  // blocks, two output tiles per item where the channel count allows, else
  // one. src / in2 / ca describe the input (HaloSrc), kOut the output.
  void conv(int k, const uint16_t* in, void* out, int outKind, int h, int w, int src = HS_PLAIN,
            const uint16_t* in2 = nullptr, int ca = 0) {
    const Conv& c = net.convs[k];
    HaloArgs A;
    A.in = in;
    A.in2 = in2;
    A.inF = nullptr;
    A.w = c.wf.as<uint16_t>();
    A.bias = net.params.as<float>() + c.bOff;
    A.out = out;
    A.raw = devIn.as<float>();
    A.H = h;
    A.W = w;
    A.Cin = c.cin;
    A.Cout = c.cout;
    A.ca = ca;
    A.relu = c.relu ? 1 : 0;
    const int ht = (c.cout / 32) % 2 == 0 ? 2 : 1;
    A.numItems = N * ((w + kHaloTW - 1) / kHaloTW) * ((h + kHaloTH - 1) / kHaloTH) * (c.cout / (32 * ht));
    if (src == HS_UPCAT && (ca % kHaloCi || (c.cin - ca) % 8))
      throw std::runtime_error("denoiser: concatenated input needs 32-channel-aligned sources");
#define SD_HALO(O, T, S)                                       \
  if (outKind == O && ht == T && src == S) {                   \
    launchHalo<O, T, S>(A, stream);                            \
    HIP_OK(hipGetLastError());                                 \
    return;                                                    \
  }
    SD_HALO(HO_BF16, 1, HS_PLAIN) SD_HALO(HO_BF16, 2, HS_PLAIN)
    SD_HALO(HO_F32, 1, HS_PLAIN) SD_HALO(HO_F32, 2, HS_PLAIN)
    SD_HALO(HO_BF16, 1, HS_POOL) SD_HALO(HO_BF16, 2, HS_POOL)
    SD_HALO(HO_BF16, 1, HS_UPCAT) SD_HALO(HO_BF16, 2, HS_UPCAT)
    SD_HALO(HO_FINAL, 1, HS_PLAIN)
#undef SD_HALO
    throw std::runtime_error("denoiser: unsupported convolution variant");
  }

  // forward; in = packed BF16 input (already in `input`). Training keeps
  // the pooled and concatenated layer inputs and the network output `pred`
  // (its backward pass reads them); inference (`infer`) has the
  // convolutions read the pooled and concatenated inputs from their sources
  // and the output layer write the final log-space image directly. Both
  // perform the same floating-point operations, so `fin` is bit-identical.
  void forward(bool infer = false) {
    const uint16_t* cur = input.as<uint16_t>();
    int k = 0;
    for (int l = 0; l < 5; l++) {
      int h = H >> l, w = W >> l;
      if (l > 0 && infer) {
        conv(k, y[k - 1].as<uint16_t>(), y[k].p, HO_BF16, h, w, HS_POOL);
      } else {
        if (l > 0) {
          size_t tot = (size_t)N * h * w * kCh[l - 1];
          hipLaunchKernelGGL(k_pool_fwd, dim3(blocks(tot)), dim3(256), 0, stream, y[k - 1].as<uint16_t>(),
                             pooled[l - 1].as<uint16_t>(), N, h * 2, w * 2, kCh[l - 1]);
          cur = pooled[l - 1].as<uint16_t>();
        }
        conv(k, cur, y[k].p, HO_BF16, h, w);
      }
      conv(k + 1, y[k].as<uint16_t>(), y[k + 1].p, HO_BF16, h, w);
      cur = y[k + 1].as<uint16_t>();
      k += 2;
    }
    for (int l = 3; l >= 0; l--) {
      int h = H >> l, w = W >> l;
      int skipConv = 2 * l + 1;
      if (infer) {
        conv(k, cur, y[k].p, HO_BF16, h, w, HS_UPCAT, y[skipConv].as<uint16_t>(), kCh[l + 1]);
      } else {
        size_t tot = (size_t)N * h * w * (kCh[l + 1] + kCh[l]) / 8;  // 16-byte vectors
        hipLaunchKernelGGL(k_upcat_fwd, dim3(blocks(tot)), dim3(256), 0, stream, cur, y[skipConv].as<uint16_t>(),
                           cat[l].as<uint16_t>(), N, h, w, kCh[l + 1], kCh[l]);
        conv(k, cat[l].as<uint16_t>(), y[k].p, HO_BF16, h, w);
      }
      conv(k + 1, y[k].as<uint16_t>(), y[k + 1].p, HO_BF16, h, w);
      cur = y[k + 1].as<uint16_t>();
      k += 2;
    }
    if (infer) {
      conv(k, cur, fin.p, HO_FINAL, H, W);
    } else {
      conv(k, cur, pred.p, HO_F32, H, W);
      size_t P = (size_t)N * H * W;
      hipLaunchKernelGGL(k_final, dim3(blocks(P * 3)), dim3(256), 0, stream, pred.as<float>(), devIn.as<float>(),
                         fin.as<float>(), P);
    }
    HIP_OK(hipGetLastError());
  }

  // dW, db for conv k given dy (FP32, already relu-masked) and its input x
  void convGradW(int k, const float* dy, const uint16_t* x, int h, int w) {
    const Conv& c = net.convs[k];
    if (c.cout % 32 || c.cout > 160) throw std::runtime_error("denoiser: backward-weights needs Cout % 32 == 0, <= 160");
    // tile width: the least padding past the image edge, wider on ties
    int tc = 32;
    for (int cand : {16, 8}) {
      auto waste = [&](int t) { return (w + t - 1) / t * t - w; };
      if (waste(cand) < waste(tc)) tc = cand;
    }
    DwArgs A;
    A.dy = dy;
    A.x = x;
    A.N = N;
    A.H = h;
    A.W = w;
    A.Cin = c.cin;
    A.Cout = c.cout;
    A.tilesX = (w + tc - 1) / tc;
    A.tilesY = (h + 64 / tc - 1) / (64 / tc);
    A.tiles = N * A.tilesX * A.tilesY;
    if (c.cin % 8) throw std::runtime_error("denoiser: backward-weights needs Cin a multiple of 8");
    const int ciTiles = (c.cin + 31) / 32;
    // about kDwBlocks blocks, each over at least one tile, with at most
    // kDwPartialBytes of partial sums
    const size_t row = (size_t)c.cout * 9 * c.cin + c.cout;
    const int maxChunks = (int)std::max<size_t>(1, kDwPartialBytes / (row * 4));
    A.tilesPerChunk = std::max({1, (A.tiles * ciTiles + kDwBlocks - 1) / kDwBlocks,
                                (A.tiles + maxChunks - 1) / maxChunks});
    const int chunks = (A.tiles + A.tilesPerChunk - 1) / A.tilesPerChunk;
    const size_t need = (size_t)chunks * row * 4;
    if (need > partialCap) {
      partial.alloc(need);
      partialCap = need;
    }
    A.partial = partial.as<float>();
    const dim3 grid(chunks, ciTiles);
    const int nco = c.cout / 32;
#define SD_DW(TC, NCO) \
  if (tc == TC && nco == NCO) hipLaunchKernelGGL((k_conv_dw<TC, NCO>), grid, dim3(192 * NCO), 0, stream, A); else
#define SD_DW_TC(TC) SD_DW(TC, 1) SD_DW(TC, 2) SD_DW(TC, 3) SD_DW(TC, 4) SD_DW(TC, 5)
    SD_DW_TC(32) SD_DW_TC(16) SD_DW_TC(8) throw std::runtime_error("denoiser: no backward-weights variant");
#undef SD_DW_TC
#undef SD_DW
    hipLaunchKernelGGL(k_reduce_chunks, dim3(blocks(row, 32)), dim3(256), 0, stream, partial.as<float>(),
                       net.grads.as<float>() + c.wOff, chunks, row);
    HIP_OK(hipGetLastError());
  }

  // dx = conv^T(dy): the forward kernel over dy (FP32, rounded to BF16 as
  // it is staged) with the transposed, tap-mirrored weights [ci][8 - t][co]
  void convGradX(int k, const float* dy, float* dx, int h, int w) {
    const Conv& c = net.convs[k];
    if (c.cin % 32) throw std::runtime_error("denoiser: backward-data needs Cin % 32 == 0");
    HaloArgs A;
    A.in = nullptr;
    A.in2 = nullptr;
    A.inF = dy;
    A.w = c.wt.as<uint16_t>();
    A.bias = nullptr;
    A.out = dx;
    A.raw = nullptr;
    A.H = h;
    A.W = w;
    A.Cin = c.cout;
    A.Cout = c.cin;
    A.ca = 0;
    A.relu = 0;
    const int ht = (c.cin / 32) % 2 == 0 ? 2 : 1;
    A.numItems = N * ((w + kHaloTW - 1) / kHaloTW) * ((h + kHaloTH - 1) / kHaloTH) * (c.cin / (32 * ht));
    if (ht == 2)
      launchHalo<HO_F32, 2, HS_F32>(A, stream);
    else
      launchHalo<HO_F32, 1, HS_F32>(A, stream);
    HIP_OK(hipGetLastError());
  }

  void reluBwd(int k, int h, int w) {
    size_t n = (size_t)N * h * w * net.convs[k].cout;
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks(n)), dim3(256), 0, stream, gy[k].as<float>(), y[k].as<uint16_t>(), n);
  }

  // backward from gPred (FP32 [P][32]) to parameter gradients
  void backward() {
    int nc = (int)net.convs.size();
    int k = nc - 1;
    // output conv (linear)
    const uint16_t* dIn = y[k - 1].as<uint16_t>();
    convGradW(k, gPred.as<float>(), dIn, H, W);
    convGradX(k, gPred.as<float>(), gy[k - 1].as<float>(), H, W);
    k--;
    for (int l = 0; l <= 3; l++) {
      int h = H >> l, w = W >> l;
      // second conv of decoder level l: k
      reluBwd(k, h, w);
      convGradW(k, gy[k].as<float>(), y[k - 1].as<uint16_t>(), h, w);
      convGradX(k, gy[k].as<float>(), gy[k - 1].as<float>(), h, w);
      k--;
      reluBwd(k, h, w);
      convGradW(k, gy[k].as<float>(), cat[l].as<uint16_t>(), h, w);
      convGradX(k, gy[k].as<float>(), gCat[l].as<float>(), h, w);
      k--;
      // split the concat gradient: upsampled part -> producer of the lower
      // level (conv k), skip part -> encoder output 2l+1
      int ca = kCh[l + 1], cb = kCh[l];
      size_t na = (size_t)N * (h / 2) * (w / 2) * ca;
      HIP_OK(hipMemsetAsync(gy[k].p, 0, na * 4, stream));
      hipLaunchKernelGGL(k_upcat_bwd_a, dim3(blocks(na)), dim3(256), 0, stream, gCat[l].as<float>(),
                         gy[k].as<float>(), N, h, w, ca, cb);
      // skip gradient accumulates into the encoder output's gradient, which
      // is zeroed before the decoder pass starts (see trainStep)
      size_t ns = (size_t)N * h * w * cb;
      hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(blocks(ns)), dim3(256), 0, stream, gCat[l].as<float>(),
                         gy[2 * l + 1].as<float>(), N, h, w, ca, cb);
    }
    // encoder, deepest first; k == 9 (second conv of level 4)
    for (int l = 4; l >= 0; l--) {
      int h = H >> l, w = W >> l;
      reluBwd(k, h, w);
      convGradW(k, gy[k].as<float>(), y[k - 1].as<uint16_t>(), h, w);
      convGradX(k, gy[k].as<float>(), gy[k - 1].as<float>(), h, w);
      k--;
      reluBwd(k, h, w);
      const uint16_t* xin = l == 0 ? input.as<uint16_t>() : pooled[l - 1].as<uint16_t>();
      convGradW(k, gy[k].as<float>(), xin, h, w);
      if (l > 0) {
        convGradX(k, gy[k].as<float>(), gPool[l - 1].as<float>(), h, w);
        // pool backward accumulates into the previous encoder output's gradient
        size_t nb = (size_t)N * (h * 2) * (w * 2) * kCh[l - 1];
        hipLaunchKernelGGL(k_pool_bwd, dim3(blocks(nb)), dim3(256), 0, stream, gPool[l - 1].as<float>(),
                           gy[k - 1].as<float>(), N, h * 2, w * 2, kCh[l - 1]);
      }
      k--;
    }
    HIP_OK(hipGetLastError());
  }

  void uploadBatch(const float* in, const float* tgt, int n, int h, int w) {
    resize(n, h, w);
    size_t P = (size_t)n * h * w;
    HIP_OK(hipMemcpyAsync(devIn.p, in, P * 9 * 4, hipMemcpyHostToDevice, stream));
    if (tgt) HIP_OK(hipMemcpyAsync(target.p, tgt, P * 3 * 4, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_pack_input, dim3(blocks(P * 16)), dim3(256), 0, stream, devIn.as<float>(),
                       input.as<uint16_t>(), P);
    HIP_OK(hipGetLastError());
  }

  float lossAndGrad() {
    size_t P = (size_t)N * H * W;
    unsigned nb = blocks(P * 32);
    hipLaunchKernelGGL(k_l1, dim3(nb), dim3(256), 0, stream, fin.as<float>(), target.as<float>(),
                       gPred.as<float>(), lossPartial.as<float>(), P);
    std::vector<float> lp(nb);
    HIP_OK(hipMemcpyAsync(lp.data(), lossPartial.p, nb * 4, hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    double s = 0.0;
    for (float v : lp) s += v;
    return (float)(s / (double)(P * 3));
  }

  void zeroEncoderSkipGrads() {
    // encoder outputs receive their gradient from two places (pool of the
    // next level and the decoder skip); zero them before accumulating
    for (int l = 0; l < 4; l++) {
      int k = 2 * l + 1;
      size_t n = (size_t)N * (H >> l) * (W >> l) * kCh[l];
      HIP_OK(hipMemsetAsync(gy[k].p, 0, n * 4, stream));
    }
  }

  float forwardBackward() {
    forward();
    float loss = lossAndGrad();
    zeroEncoderSkipGrads();
    backward();
    return loss;
  }

  void adam(float lr) {
    net.step++;
    float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - powf(b1, (float)net.step), c2 = 1.0f - powf(b2, (float)net.step);
    hipLaunchKernelGGL(k_adam, dim3(blocks(net.nParams)), dim3(256), 0, stream, net.params.as<float>(),
                       net.adamM.as<float>(), net.adamV.as<float>(), net.grads.as<float>(), net.nParams, lr, b1, b2,
                       1e-8f, c1, c2, net.layouts.as<WeightLayout>(), (int)net.convs.size());
    HIP_OK(hipGetLastError());
  }
};

namespace {
template <class F>
int guardedD(F&& f) {
  try {
    return f();
  } catch (const std::exception& e) {
    sdk::setLastError(e.what());
    return SUNDOG_ERROR;
  }
}
}  // namespace

namespace sdn {

// Denoises device films of a w x h frame with the weights in `weightsPath`;
// writes float4 RGBA (w x h) to outHost. Used by sundog_render on request.
double denoiseFilms(const char* weightsPath, const float4* beauty, const float4* albedo, const float4* normal,
                    int w, int h, float* outHost) {
  std::unique_ptr<sundog_denoiser> d(new sundog_denoiser());
  d->init(1);
  if (sundog_denoiser_load(d.get(), weightsPath) != SUNDOG_OK) throw std::runtime_error(sundog_last_error());
  const int W = (w + 15) / 16 * 16, H = (h + 15) / 16 * 16;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  d->resize(1, H, W);
  size_t P = (size_t)W * H;
  HIP_OK(hipEventRecord(e0, d->stream));
  hipLaunchKernelGGL(k_pack_films, dim3(blocks(P)), dim3(256), 0, d->stream, beauty, albedo, normal, w, h, W, H,
                     d->devIn.as<float>());
  hipLaunchKernelGGL(k_pack_input, dim3(blocks(P * 16)), dim3(256), 0, d->stream, d->devIn.as<float>(),
                     d->input.as<uint16_t>(), P);
  d->forward(true);
  DevBuf o;
  o.alloc((size_t)w * h * 16);
  hipLaunchKernelGGL(k_unpack_films, dim3(blocks((size_t)w * h)), dim3(256), 0, d->stream, d->fin.as<float>(), w, h,
                     W, o.as<float4>());
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(e1, d->stream));
  HIP_OK(hipMemcpyAsync(outHost, o.p, (size_t)w * h * 16, hipMemcpyDeviceToHost, d->stream));
  HIP_OK(hipStreamSynchronize(d->stream));
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

}  // namespace sdn

extern "C" sundog_denoiser* sundog_denoiser_create(uint64_t seed) {
  try {
    std::unique_ptr<sundog_denoiser> d(new sundog_denoiser());
    d->init(seed);
    return d.release();
  } catch (const std::exception& e) {
    sdk::setLastError(e.what());
    return nullptr;
  }
}

extern "C" void sundog_denoiser_destroy(sundog_denoiser* d) { delete d; }

extern "C" int64_t sundog_denoiser_param_count(sundog_denoiser* d) { return d ? (int64_t)d->net.nParams : -1; }

extern "C" int sundog_denoiser_get_params(sundog_denoiser* d, float* out) {
  return guardedD([&] {
    if (!d || !out) throw std::runtime_error("denoiser: null argument");
    HIP_OK(hipStreamSynchronize(d->stream));
    HIP_OK(hipMemcpy(out, d->net.params.p, d->net.nParams * 4, hipMemcpyDeviceToHost));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_denoiser_set_params(sundog_denoiser* d, const float* in) {
  return guardedD([&] {
    if (!d || !in) throw std::runtime_error("denoiser: null argument");
    HIP_OK(hipMemcpyAsync(d->net.params.p, in, d->net.nParams * 4, hipMemcpyHostToDevice, d->stream));
    d->syncWeights();
    HIP_OK(hipStreamSynchronize(d->stream));
    return SUNDOG_OK;
  });
}

extern "C" int sundog_denoiser_forward_backward(sundog_denoiser* d, const float* input, const float* target, int n,
                                                int h, int w, float* loss, float* grads, float* pred_out) {
  return guardedD([&] {
    if (!d || !input || !target) throw std::runtime_error("denoiser: null argument");
    d->uploadBatch(input, target, n, h, w);
    float l = d->forwardBackward();
    if (loss) *loss = l;
    if (grads) {
      HIP_OK(hipStreamSynchronize(d->stream));
      HIP_OK(hipMemcpy(grads, d->net.grads.p, d->net.nParams * 4, hipMemcpyDeviceToHost));
    }
    if (pred_out) {
      size_t P = (size_t)n * h * w;
      HIP_OK(hipStreamSynchronize(d->stream));
      HIP_OK(hipMemcpy(pred_out, d->fin.p, P * 3 * 4, hipMemcpyDeviceToHost));
    }
    return SUNDOG_OK;
  });
}

extern "C" int sundog_denoiser_train_step(sundog_denoiser* d, const float* input, const float* target, int n, int h,
                                          int w, float lr, float* loss) {
  return guardedD([&] {
    if (!d || !input || !target) throw std::runtime_error("denoiser: null argument");
    d->uploadBatch(input, target, n, h, w);
    float l = d->forwardBackward();
    d->adam(lr);
    if (loss) *loss = l;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_denoiser_infer(sundog_denoiser* d, const float* input, int n, int h, int w, float* out) {
  return guardedD([&] {
    if (!d || !input || !out) throw std::runtime_error("denoiser: null argument");
    d->uploadBatch(input, nullptr, n, h, w);
    d->forward(true);
    size_t P = (size_t)n * h * w;
    DevBuf o;
    o.alloc(P * 3 * 4);
    hipLaunchKernelGGL(k_unpack_output, dim3(blocks(P * 3)), dim3(256), 0, d->stream, d->fin.as<float>(),
                       o.as<float>(), P);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(out, o.p, P * 3 * 4, hipMemcpyDeviceToHost, d->stream));
    HIP_OK(hipStreamSynchronize(d->stream));
    return SUNDOG_OK;
  });
}

// Timed forward pass at the current resolution (for MFMA utilisation); ms
// of `iters` forwards measured with HIP events on the denoiser's stream.
extern "C" int sundog_denoiser_time_forward(sundog_denoiser* d, int n, int h, int w, int iters, double* ms_out,
                                            double* flops_out) {
  return guardedD([&] {
    if (!d) throw std::runtime_error("denoiser: null handle");
    d->resize(n, h, w);
    HIP_OK(hipMemsetAsync(d->input.p, 0, (size_t)n * h * w * 16 * 2, d->stream));
    HIP_OK(hipMemsetAsync(d->devIn.p, 0, (size_t)n * h * w * 9 * 4, d->stream));
    d->forward(true);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, d->stream));
    for (int i = 0; i < iters; i++) d->forward(true);
    HIP_OK(hipEventRecord(e1, d->stream));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    double fl = 0;
    for (size_t k = 0; k < d->net.convs.size(); k++) {
      const Conv& c = d->net.convs[k];
      fl += 2.0 * 9.0 * c.cin * c.cout * (double)n * d->yH[k] * d->yW[k];
    }
    if (ms_out) *ms_out = ms / iters;
    if (flops_out) *flops_out = fl;
    return SUNDOG_OK;
  });
}

extern "C" int sundog_denoiser_save(sundog_denoiser* d, const char* path) {
  return guardedD([&] {
    if (!d || !path) throw std::runtime_error("denoiser: null argument");
    std::vector<float> p(d->net.nParams);
    HIP_OK(hipStreamSynchronize(d->stream));
    HIP_OK(hipMemcpy(p.data(), d->net.params.p, p.size() * 4, hipMemcpyDeviceToHost));
    FILE* f = fopen(path, "wb");
    if (!f) throw std::runtime_error(std::string("denoiser: cannot write ") + path);
    const char magic[8] = {'S', 'D', 'U', 'N', 'E', 'T', '0', '1'};
    uint64_t n = p.size();
    fwrite(magic, 1, 8, f);
    fwrite(&n, 8, 1, f);
    fwrite(p.data(), 4, p.size(), f);
    fclose(f);
    return SUNDOG_OK;
  });
}

extern "C" int sundog_denoiser_load(sundog_denoiser* d, const char* path) {
  return guardedD([&] {
    if (!d || !path) throw std::runtime_error("denoiser: null argument");
    FILE* f = fopen(path, "rb");
    if (!f) throw std::runtime_error(std::string("denoiser: cannot read ") + path);
    char magic[8];
    uint64_t n = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "SDUNET01", 8) == 0 && fread(&n, 8, 1, f) == 1 &&
              n == d->net.nParams;
    std::vector<float> p(ok ? n : 0);
    ok = ok && fread(p.data(), 4, n, f) == n;
    fclose(f);
    if (!ok) throw std::runtime_error(std::string("denoiser: bad weights file ") + path);
    return sundog_denoiser_set_params(d, p.data());
  });
}
