// kernels_cnn.hip -- the learned nonlinear-term closure: a 3D CNN over the
// coarse channel volume, trained and run with BF16 MFMA kernels.
//
// Volume: v[y][z][x][c], channels last, periodic in x and z, zero outside
// the y range (the planes of the u grid).  Network: four 3x3x3 convolutions,
// 4 -> 32 -> 32 -> 32 -> 3 channels, GELU after the first three; BF16 weights
// and activations, FP32 accumulation, FP32 master weights (Adam).
//
// Every convolution is an implicit GEMM on v_mfma_f32_32x32x16_bf16: a wave
// computes 32 output points (one x-run of a z row) by 32 output channels,
// K = 27 taps x input channels in steps of 16.  Lane layout of that MFMA on
// gfx950 (checked on the device): A[m][k] in lane m % 32, elements
// k = 8 (l / 32) + i; B[k][n] in lane n % 32, elements k = 8 (l / 32) + i;
// C[m][n] in lane n % 32, element v at m = 8 (v / 4) + 4 (l / 32) + v % 4.
// The forward and backward-data blocks (4 waves, 8 z rows, two per wave)
// keep the layer's weights and a ring of three halo planes (10 rows x 34
// points) in LDS; the weight-gradient blocks (4 z rows) stage a 3-plane halo
// of 6 rows.  The LDS point stride is padded (48 B for 16 channels, 80 B for
// 32) so the 16-byte reads of one MFMA operand hit distinct banks.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace cfd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kC = 32;      // hidden channels
constexpr int kTX = 32;     // points per wave (x)
constexpr int kTZ = 4;      // z rows per block (one per wave)
constexpr int kHX = kTX + 2, kHZ = kTZ + 2, kHY = 3;
constexpr int kHalo = kHX * kHZ * kHY;

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  // d/dz of z Phi(z) = Phi(z) + z phi(z)
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// LDS channel stride (elements) of a point for CIN channels
template <int CIN>
struct Lds {
  static constexpr int CP = CIN == 16 ? 24 : 40;
};

// stage the halo of block (x0, z0, y) into LDS: hal[(hy * kHZ + hz) * kHX + hx][c]
// from a global volume of CG channels (CG <= CIN; the rest zero)
template <int CIN, int CG>
__device__ __forceinline__ void stage_halo(__bf16* hal, const __bf16* __restrict__ in, int x0, int z0, int y, int NY,
                                           int NZ, int NX) {
  constexpr int CP = Lds<CIN>::CP;
  if constexpr (CG == CIN) {
    // 8-channel (16 B) chunks
    constexpr int NCH = CIN / 8;
    for (int w = threadIdx.x; w < kHalo * NCH; w += blockDim.x) {
      const int pnt = w / NCH, ch = w - pnt * NCH;
      const int hx = pnt % kHX, hz = (pnt / kHX) % kHZ, hy = pnt / (kHX * kHZ);
      const int yy = y + hy - 1;
      bf16x8 v;
      if (yy < 0 || yy >= NY) {
        for (int i = 0; i < 8; ++i) v[i] = (__bf16)0.0f;
      } else {
        const int xx = wrap(x0 + hx - 1, NX), zz = wrap(z0 + hz - 1, NZ);
        v = *(const bf16x8*)(in + (((size_t)yy * NZ + zz) * NX + xx) * CG + 8 * ch);
      }
      *(bf16x8*)(hal + (size_t)pnt * CP + 8 * ch) = v;
    }
  } else {
    // CG (4) channels per point, padded with zeros to CIN
    for (int pnt = threadIdx.x; pnt < kHalo; pnt += blockDim.x) {
      const int hx = pnt % kHX, hz = (pnt / kHX) % kHZ, hy = pnt / (kHX * kHZ);
      const int yy = y + hy - 1;
      bf16x8 v0, v1;
      for (int i = 0; i < 8; ++i) {
        v0[i] = (__bf16)0.0f;
        v1[i] = (__bf16)0.0f;
      }
      if (yy >= 0 && yy < NY) {
        const int xx = wrap(x0 + hx - 1, NX), zz = wrap(z0 + hz - 1, NZ);
        const __bf16* s = in + (((size_t)yy * NZ + zz) * NX + xx) * CG;
        for (int c = 0; c < CG; ++c) v0[c] = s[c];
      }
      *(bf16x8*)(hal + (size_t)pnt * CP) = v0;
      *(bf16x8*)(hal + (size_t)pnt * CP + 8) = v1;
    }
  }
}

}  // namespace

// Tiles of the forward / backward-data convolution: 8 z rows per block, two
// per wave, so every weight fragment feeds two independent accumulators.
constexpr int kTZF = 8;
constexpr int kHZF = kTZF + 2;
constexpr int kYB = 8;  // planes walked by one block

// One plane of a block's halo (kHZF rows x kHX points), loaded into
// registers by all threads and stored to an LDS ring slot later, so the
// global loads of the next plane are in flight while the MFMAs of the
// current plane run.  Slot layout [hz * kHX + hx][c], point stride CP.
template <int CIN, int CG>
struct PlaneLoad {
  static constexpr int CP = Lds<CIN>::CP;
  static constexpr int NPL = kHX * kHZF;
  static constexpr int NCH = CG == CIN ? CIN / 8 : 1;
  static constexpr int NITEM = NPL * NCH;
  static constexpr int PER = (NITEM + 255) / 256;
  bf16x8 r[PER];

  __device__ __forceinline__ void load(const __bf16* __restrict__ in, int x0, int z0, int yy, int NY, int NZ,
                                       int NX) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int w = threadIdx.x + 256 * k;
      for (int i = 0; i < 8; ++i) r[k][i] = (__bf16)0.0f;
      if (w >= NITEM || yy < 0 || yy >= NY) continue;
      const int pnt = w / NCH, ch = w - pnt * NCH;
      const int hx = pnt % kHX, hz = pnt / kHX;
      const int xx = wrap(x0 + hx - 1, NX), zz = wrap(z0 + hz - 1, NZ);
      const __bf16* src = in + (((size_t)yy * NZ + zz) * NX + xx) * CG;
      if constexpr (CG == CIN) {
        r[k] = *(const bf16x8*)(src + 8 * ch);
      } else {
        for (int c = 0; c < CG; ++c) r[k][c] = src[c];
      }
    }
  }
  __device__ __forceinline__ void store(__bf16* slot) const {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int w = threadIdx.x + 256 * k;
      if (w >= NITEM) continue;
      const int pnt = w / NCH, ch = w - pnt * NCH;
      *(bf16x8*)(slot + (size_t)pnt * CP + 8 * ch) = r[k];
      if constexpr (CG != CIN) {
        bf16x8 z;
        for (int i = 0; i < 8; ++i) z[i] = (__bf16)0.0f;
        for (int c = 8; c < CIN; c += 8) *(bf16x8*)(slot + (size_t)pnt * CP + c) = z;
      }
    }
  }
};

// Convolution of one layer.  MODE 0: z = conv + b -> zout (BF16), a = gelu(z)
// -> aout (BF16).  MODE 1: y = conv + b -> yout (FP32, channels < nout).
// MODE 2 (backward data): taps flipped, weights W'[t][ci][co] given as the
// operand layout, d = conv * gelu'(zprev) -> aout (BF16); no bias.
// Weights w: [27][32 (n)][CIN (k)] BF16, staged in LDS once per block.  A
// block walks kYB planes of one (x, 8 z rows) tile with its halo in a ring
// of three plane slots; the plane two ahead is loaded into registers during
// the MFMAs of the current one.  Grid (NX/32, NZ/8, ceil(NY/kYB)), 256
// threads.
template <int CIN, int CG, int MODE>
__global__ void __launch_bounds__(256) k_conv(const __bf16* __restrict__ in, const __bf16* __restrict__ w,
                                              const float* __restrict__ bias, const __bf16* __restrict__ zprev,
                                              __bf16* __restrict__ zout, __bf16* __restrict__ aout,
                                              float* __restrict__ yout, int nout, int NY, int NZ, int NX) {
  constexpr int CP = Lds<CIN>::CP;
  constexpr int SLOT = kHX * kHZF * CP;
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* ring = (__bf16*)smem;                // 3 * SLOT
  __bf16* wl = ring + (size_t)3 * SLOT;        // 27 * 32 * CP
  const int x0 = blockIdx.x * kTX, z0 = blockIdx.y * kTZF;
  const int y0 = blockIdx.z * kYB, y1 = min(y0 + kYB, NY);
  for (int e = threadIdx.x; e < 27 * kC * (CIN / 8); e += blockDim.x) {
    const int ch = e % (CIN / 8), row = e / (CIN / 8);  // row = t * 32 + n
    *(bf16x8*)(wl + (size_t)row * CP + 8 * ch) = *(const bf16x8*)(w + (size_t)row * CIN + 8 * ch);
  }
  PlaneLoad<CIN, CG> pl;
  for (int d = -1; d <= 1; ++d) {
    pl.load(in, x0, z0, y0 + d, NY, NZ, NX);
    pl.store(ring + (size_t)((y0 + d + 3) % 3) * SLOT);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 31, kh = lane >> 5;
  const int n = m;
  const float bn = MODE == 2 ? 0.0f : bias[n];
  for (int y = y0; y < y1; ++y) {
    const bool more = y + 1 < y1;
    if (more) pl.load(in, x0, z0, y + 2, NY, NZ, NX);
    f32x16 acc0 = {}, acc1 = {};  // z rows 2 wv and 2 wv + 1 of the tile
#pragma unroll
    for (int t = 0; t < 27; ++t) {
      int dy = t / 9 - 1, dz = (t / 3) % 3 - 1, dx = t % 3 - 1;
      if (MODE == 2) {
        dy = -dy;
        dz = -dz;
        dx = -dx;
      }
      const __bf16* arow = ring + (size_t)((y + dy + 3) % 3) * SLOT +
                           (size_t)((1 + 2 * wv + dz) * kHX + (1 + m + dx)) * CP;
      const __bf16* brow = wl + (size_t)(t * kC + m) * CP;
#pragma unroll
      for (int s = 0; s < CIN / 16; ++s) {
        const bf16x8 b = *(const bf16x8*)(brow + 16 * s + 8 * kh);
        const bf16x8 a0 = *(const bf16x8*)(arow + 16 * s + 8 * kh);
        const bf16x8 a1 = *(const bf16x8*)(arow + (size_t)kHX * CP + 16 * s + 8 * kh);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b, acc1, 0, 0, 0);
      }
    }
    // epilogue: lane holds channel n = m of points 8 (v/4) + 4 kh + v % 4
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int z = z0 + 2 * wv + r;
      const size_t rowp = ((size_t)y * NZ + z) * NX + x0;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int pm = 8 * (v / 4) + 4 * kh + (v & 3);
        const size_t p = rowp + pm;
        const float acc = r == 0 ? acc0[v] : acc1[v];
        if (MODE == 0) {
          const float zz = acc + bn;
          zout[p * kC + n] = (__bf16)zz;
          aout[p * kC + n] = (__bf16)gelu(zz);
        } else if (MODE == 1) {
          if (n < nout) yout[p * nout + n] = acc + bn;
        } else {
          aout[p * kC + n] = (__bf16)(acc * gelu_d((float)zprev[p * kC + n]));
        }
      }
    }
    if (more) {
      __syncthreads();  // every wave is done with plane y - 1
      pl.store(ring + (size_t)((y + 2) % 3) * SLOT);
      __syncthreads();
    }
  }
}

// weight gradients: dW[t][ci][co] = sum_p x[p + off(t)][ci] dz[p][co] and
// db[co] = sum_p dz[p][co], per block into partial[blk][27 * 32 * 32 + 32]
// (ci padded to 32).  Wave w accumulates the taps t = w, w + 4, ... over the
// 4 z rows of the block tile and the blocks of its grid-stride loop.
// Grid: nblk blocks of 256 threads looping over the (NX/32, NZ/4, NY) tiles.
template <int CIN, int CG>
__global__ void __launch_bounds__(256) k_conv_dw(const __bf16* __restrict__ in, const __bf16* __restrict__ dz,
                                                 float* __restrict__ partial, int NY, int NZ, int NX) {
  constexpr int CP = Lds<CIN>::CP;
  constexpr int DP = 40;  // LDS stride of a dz point
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* hal = (__bf16*)smem;
  __bf16* dzl = hal + (size_t)kHalo * CP;  // kTZ * kTX * DP
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 31, kh = lane >> 5;
  constexpr int NT = 7;  // taps per wave (27 = 4 * 7 - 1)
  f32x16 acc[NT];
  for (int q = 0; q < NT; ++q) acc[q] = f32x16{};
  float dbs = 0.0f;
  const int tx = NX / kTX, tz = NZ / kTZ, ntiles = tx * tz * NY;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int bx = tile % tx, bz = (tile / tx) % tz, y = tile / (tx * tz);
    const int x0 = bx * kTX, z0 = bz * kTZ;
    __syncthreads();
    stage_halo<CIN, CG>(hal, in, x0, z0, y, NY, NZ, NX);
    for (int e = threadIdx.x; e < kTZ * kTX * (kC / 8); e += blockDim.x) {
      const int ch = e % (kC / 8), pnt = e / (kC / 8);
      const int zr = pnt / kTX, xr = pnt % kTX;
      *(bf16x8*)(dzl + (size_t)pnt * DP + 8 * ch) =
          *(const bf16x8*)(dz + ((((size_t)y * NZ + z0 + zr) * NX + x0 + xr) * kC + 8 * ch));
    }
    __syncthreads();
    for (int zr = 0; zr < kTZ; ++zr) {
      // bias gradient: lanes 0..31 of wave 0 own a channel each
      if (wv == 0 && kh == 0)
        for (int xr = 0; xr < kTX; ++xr) dbs += (float)dzl[(size_t)(zr * kTX + xr) * DP + m];
      for (int s = 0; s < 2; ++s) {  // 16 points per MFMA
        // B[k = point][n = co]: points 16 s + 8 kh + i, channel m
        bf16x8 b;
        for (int i = 0; i < 8; ++i) b[i] = dzl[(size_t)(zr * kTX + 16 * s + 8 * kh + i) * DP + m];
#pragma unroll
        for (int q = 0; q < NT; ++q) {
          const int t = wv + 4 * q;
          if (t < 27) {
            const int dy = t / 9 - 1, dz2 = (t / 3) % 3 - 1, dx = t % 3 - 1;
            // A[m = ci][k = point]: x[point + off][ci], ci = m (zero above CIN)
            bf16x8 a;
            const __bf16* base =
                hal + (size_t)(((1 + dy) * kHZ + (1 + zr + dz2)) * kHX + (1 + 16 * s + 8 * kh + dx)) * CP;
            for (int i = 0; i < 8; ++i) a[i] = m < CIN ? base[(size_t)i * CP + m] : (__bf16)0.0f;
            acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
          }
        }
      }
    }
  }
  float* out = partial + (size_t)blockIdx.x * (27 * kC * kC + kC);
#pragma unroll
  for (int q = 0; q < NT; ++q) {
    const int t = wv + 4 * q;
    if (t >= 27) continue;
    for (int v = 0; v < 16; ++v) {
      const int ci = 8 * (v / 4) + 4 * kh + (v & 3);
      out[((size_t)t * kC + ci) * kC + m] = acc[q][v];
    }
  }
  if (wv == 0 && kh == 0) out[27 * kC * kC + m] = dbs;
}

// sum of the per-block partials: g[e] = sum_b partial[b][e]; block y sums
// the partials b = y, y + gridDim.y, ... and adds its sum atomically (the
// order of the additions, and so the last bits of g, vary between runs)
__global__ void k_reduce_partials(const float* __restrict__ partial, int nblk, int n, float* __restrict__ g) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.0f;
  for (int b = blockIdx.y; b < nblk; b += gridDim.y) s += partial[(size_t)b * n + e];
  atomicAdd(g + e, s);
}

// loss and output gradient: L = sum over points and the nout channels of
// (y - t)^2 / (npts nout); dy = 2 (y - t) / (npts nout) into a 32-channel
// BF16 volume (zero in the other channels); lsum[blockIdx] = block sum
__global__ void __launch_bounds__(256) k_loss(const float* __restrict__ y, const __bf16* __restrict__ tgt, int tc,
                                              int nout, size_t npts, __bf16* __restrict__ dy,
                                              float* __restrict__ lsum) {
  const float sc = 1.0f / (float)((double)npts * nout);
  float s = 0.0f;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < npts; p += (size_t)gridDim.x * blockDim.x) {
    bf16x8 g0, g1, g2, g3;
    for (int i = 0; i < 8; ++i) {
      g0[i] = (__bf16)0.0f;
      g1[i] = (__bf16)0.0f;
      g2[i] = (__bf16)0.0f;
      g3[i] = (__bf16)0.0f;
    }
    for (int c = 0; c < nout; ++c) {
      const float d = y[p * nout + c] - (float)tgt[p * tc + c];
      s += d * d;
      g0[c] = (__bf16)(2.0f * d * sc);
    }
    bf16x8* o = (bf16x8*)(dy + p * kC);
    o[0] = g0;
    o[1] = g1;
    o[2] = g2;
    o[3] = g3;
  }
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) lsum[blockIdx.x] = red[0] * sc;
}

// Adam on the FP32 master parameters (bias-corrected, Kingma & Ba 2015)
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float a = b1 * m1[e] + (1.0f - b1) * gg;
  const float b = b2 * m2[e] + (1.0f - b2) * gg * gg;
  m1[e] = a;
  m2[e] = b;
  w[e] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 operand layouts of one layer from the FP32 master [t][ci 32][co 32]:
// forward wf[t][co][ci (cin)], backward-data wb[t][ci][co]
__global__ void k_wlayouts(const float* __restrict__ w, int cin, __bf16* __restrict__ wf, __bf16* __restrict__ wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 27 * kC * kC) return;
  const int co = e % kC, ci = (e / kC) % kC, t = e / (kC * kC);
  const float v = w[e];
  if (ci < cin) wf[((size_t)t * kC + co) * cin + ci] = (__bf16)v;
  wb[((size_t)t * kC + ci) * kC + co] = (__bf16)v;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace {

template <int CIN>
constexpr size_t conv_lds() {
  return ((size_t)3 * kHX * kHZF * Lds<CIN>::CP + (size_t)27 * kC * Lds<CIN>::CP) * sizeof(__bf16);
}
template <int CIN>
constexpr size_t dw_lds() {
  return ((size_t)kHalo * Lds<CIN>::CP + (size_t)kTZ * kTX * 40) * sizeof(__bf16);
}

// raise the dynamic-LDS limit of a kernel once (kernels above 64 KB)
template <typename K>
void lds_attr(K kern, size_t bytes) {
  static std::vector<const void*> done;
  const void* p = (const void*)kern;
  for (const void* q : done)
    if (q == p) return;
  (void)hipFuncSetAttribute(p, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  done.push_back(p);
}

}  // namespace

bool cnn_dims_ok(int NY, int NZ, int NX) { return NY >= 1 && NZ % kTZF == 0 && NX % kTX == 0 && NZ > 0 && NX > 0; }

void launch_conv_fwd(hipStream_t st, int layer, const __bf16* in, const __bf16* wf, const float* bias, __bf16* zout,
                     __bf16* aout, float* yout, int nout, int NY, int NZ, int NX) {
  const dim3 grid(NX / kTX, NZ / kTZF, (NY + kYB - 1) / kYB);
  if (layer == 0) {
    auto k = k_conv<16, 4, 0>;
    lds_attr(k, conv_lds<16>());
    hipLaunchKernelGGL(k, grid, dim3(256), conv_lds<16>(), st, in, wf, bias, nullptr, zout, aout, nullptr, 0, NY, NZ,
                       NX);
  } else if (yout) {
    auto k = k_conv<32, 32, 1>;
    lds_attr(k, conv_lds<32>());
    hipLaunchKernelGGL(k, grid, dim3(256), conv_lds<32>(), st, in, wf, bias, nullptr, nullptr, nullptr, yout, nout,
                       NY, NZ, NX);
  } else {
    auto k = k_conv<32, 32, 0>;
    lds_attr(k, conv_lds<32>());
    hipLaunchKernelGGL(k, grid, dim3(256), conv_lds<32>(), st, in, wf, bias, nullptr, zout, aout, nullptr, 0, NY, NZ,
                       NX);
  }
}

void launch_conv_bwd_data(hipStream_t st, const __bf16* dz, const __bf16* wb, const __bf16* zprev, __bf16* dprev,
                          int NY, int NZ, int NX) {
  const dim3 grid(NX / kTX, NZ / kTZF, (NY + kYB - 1) / kYB);
  auto k = k_conv<32, 32, 2>;
  lds_attr(k, conv_lds<32>());
  hipLaunchKernelGGL(k, grid, dim3(256), conv_lds<32>(), st, dz, wb, nullptr, zprev, nullptr, dprev, nullptr, 0, NY,
                     NZ, NX);
}

void launch_conv_dw(hipStream_t st, int layer, const __bf16* in, const __bf16* dz, float* partial, int nblk,
                    float* grad, int NY, int NZ, int NX) {
  if (layer == 0) {
    auto k = k_conv_dw<16, 4>;
    lds_attr(k, dw_lds<16>());
    hipLaunchKernelGGL(k, dim3(nblk), dim3(256), dw_lds<16>(), st, in, dz, partial, NY, NZ, NX);
  } else {
    auto k = k_conv_dw<32, 32>;
    lds_attr(k, dw_lds<32>());
    hipLaunchKernelGGL(k, dim3(nblk), dim3(256), dw_lds<32>(), st, in, dz, partial, NY, NZ, NX);
  }
  const int n = 27 * kC * kC + kC;
  (void)hipMemsetAsync(grad, 0, (size_t)n * sizeof(float), st);
  hipLaunchKernelGGL(k_reduce_partials, dim3((n + 255) / 256, 16), dim3(256), 0, st, partial, nblk, n, grad);
}

void launch_loss(hipStream_t st, const float* y, const __bf16* tgt, int tc, int nout, size_t npts, __bf16* dy,
                 float* lsum, int nblk) {
  hipLaunchKernelGGL(k_loss, dim3(nblk), dim3(256), 0, st, y, tgt, tc, nout, npts, dy, lsum);
}

void launch_adam(hipStream_t st, float* w, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2,
                 float eps, int step) {
  const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, w, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
}

void launch_wlayouts(hipStream_t st, const float* w, int cin, __bf16* wf, __bf16* wb) {
  hipLaunchKernelGGL(k_wlayouts, dim3((27 * kC * kC + 255) / 256), dim3(256), 0, st, w, cin, wf, wb);
}

}  // namespace cfd

// ---------------------------------------------------------------------------
// closure data movement between the solver and the network volume
// ---------------------------------------------------------------------------
namespace cfd {

namespace {

__device__ __forceinline__ double sinc(double a) { return a == 0.0 ? 1.0 : sin(a) / a; }

// out[q][k Mxc + i] = a * Hx(i) Hz(kz) fine[q][kf Mxf + i] + b * other[q][...]
// (box filter of one coarse cell width, spectral form, on the coarse modes)
__global__ void k_filter(const double2* __restrict__ fine, const double2* __restrict__ other, double2* __restrict__ out,
                         int Mxf, int N2f, int Mxc, int N2c, int ngxc, int ngzc, double a, double b, int nplanes) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;
  const int ncol = Mxc * N2c;
  if (col >= ncol || q >= nplanes) return;
  const int k = col / Mxc, i = col - k * Mxc;
  const int kz = k < N2c / 2 ? k : k - N2c;
  const int kf = kz >= 0 ? kz : N2f + kz;
  const double pi = 3.14159265358979323846;
  const double h = sinc(pi * i / ngxc) * sinc(pi * (kz < 0 ? -kz : kz) / ngzc);
  const double2 f = fine[(size_t)q * Mxf * N2f + (size_t)kf * Mxf + i];
  double2 v = make_double2(a * h * f.x, a * h * f.y);
  if (other) {
    const double2 o = other[(size_t)q * ncol + col];
    v = make_double2(v.x + b * o.x, v.y + b * o.y);
  }
  out[(size_t)q * ncol + col] = v;
}

// volume point (y, z, x) <- physical planes: u1, u3 on u plane y, u2 on v
// plane y (zero on the last plane), the mask; scaled by 1 / s
__global__ void k_pack(const double* __restrict__ p1, const double* __restrict__ p2, const double* __restrict__ p3,
                       const unsigned char* __restrict__ mask, int nyu, int nyv, int ngz, int ngx, float s1, float s2,
                       float s3, int withmask, __bf16* __restrict__ out) {
  const size_t pl = (size_t)ngz * ngx;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pl * nyu) return;
  const int y = (int)(e / pl);
  const float a = (float)p1[e] / s1;
  const float b = y < nyv ? (float)p2[e] / s2 : 0.0f;
  const float c = (float)p3[e] / s3;
  const float m = withmask && mask ? (float)mask[e] : 0.0f;
  __bf16* o = out + e * 4;
  o[0] = (__bf16)a;
  o[1] = (__bf16)b;
  o[2] = (__bf16)c;
  o[3] = (__bf16)m;
}

// physical corrections <- network output y[p][3] * s on the interior planes
// (u planes 1..nyu-2 for components 1 and 3, v planes 1..nyv-2 for 2)
// With u1..u3 given (the physical velocities of the same planes; all three
// or none), a
// component whose correction c would feed energy into the resolved field
// (u c < 0: the correction enters the momentum equation as -c) is set to
// zero there, so the correction only removes energy (backscatter clipping).
__global__ void k_unpack(const float* __restrict__ y, int nyu, int nyv, int ngz, int ngx, float s1, float s2, float s3,
                         double* __restrict__ p1, double* __restrict__ p2, double* __restrict__ p3,
                         const double* __restrict__ u1, const double* __restrict__ u2, const double* __restrict__ u3) {
  const size_t pl = (size_t)ngz * ngx;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pl * nyu) return;
  const int q = (int)(e / pl);
  const bool iu = q >= 1 && q <= nyu - 2, iv = q >= 1 && q <= nyv - 2;
  double c1 = iu ? (double)(y[e * 3 + 0] * s1) : 0.0;
  double c3 = iu ? (double)(y[e * 3 + 2] * s3) : 0.0;
  if (u1) {
    if (u1[e] * c1 < 0.0) c1 = 0.0;
    if (u3[e] * c3 < 0.0) c3 = 0.0;
  }
  p1[e] = c1;
  p3[e] = c3;
  if (q < nyv) {
    double c2 = iv ? (double)(y[e * 3 + 1] * s2) : 0.0;
    if (u1 && u2[e] * c2 < 0.0) c2 = 0.0;
    p2[e] = c2;
  }
}

// a[q] += c[q] on planes q0 .. q1 - 1 of Ncol complex values; a in FP64 or,
// on the mixed path, FP32 (the sum is formed in FP64 and rounded once)
template <typename C2>
__global__ void k_add_planes(C2* __restrict__ a, const double2* __restrict__ c, int ncol, int q0, int q1) {
  using R = decltype(a->x);
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = (size_t)(q1 - q0) * ncol;
  if (e >= n) return;
  const size_t o = (size_t)q0 * ncol + e;
  const C2 x = a[o];
  a[o].x = (R)((double)x.x + c[o].x);
  a[o].y = (R)((double)x.y + c[o].y);
}

// The (0,0) column of a correction on u planes q0 .. q1 - 1 loses its
// mean over y weighted by the plane widths yv[q] - yv[q - 1] (u plane q lies
// between v faces q - 1 and q): the correction then exerts no net force on
// the flow, as the divergence of a stress that vanishes at both walls does.
// One thread: a column of at most a few hundred planes.
__global__ void k_zero_net_force(double2* __restrict__ c, int ncol, int q0, int q1, const double* __restrict__ yv) {
  double F = 0.0, H = 0.0;
  for (int q = q0; q < q1; ++q) {
    const double w = yv[q] - yv[q - 1];
    F += w * c[(size_t)q * ncol].x;
    H += w;
  }
  const double m = F / H;
  for (int q = q0; q < q1; ++q) c[(size_t)q * ncol].x -= m;
}

// sum of squares of n values into *acc (one atomic per block)
__global__ void k_sumsq(const double* __restrict__ v, size_t n, double* __restrict__ acc) {
  double s = 0.0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    s += v[e] * v[e];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(acc, red[0]);
}

}  // namespace

void launch_filter(hipStream_t st, const double2* fine, const double2* other, double2* out, int Mxf, int N2f, int Mxc,
                   int N2c, int ngxc, int ngzc, double a, double b, int nplanes) {
  if (nplanes <= 0) return;
  hipLaunchKernelGGL(k_filter, dim3((Mxc * N2c + 255) / 256, nplanes), dim3(256), 0, st, fine, other, out, Mxf, N2f,
                     Mxc, N2c, ngxc, ngzc, a, b, nplanes);
}

void launch_pack(hipStream_t st, const double* p1, const double* p2, const double* p3, const unsigned char* mask,
                 int nyu, int nyv, int ngz, int ngx, const float* s, int withmask, __bf16* out) {
  const size_t n = (size_t)nyu * ngz * ngx;
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p1, p2, p3, mask, nyu, nyv, ngz, ngx,
                     s[0], s[1], s[2], withmask, out);
}

void launch_unpack(hipStream_t st, const float* y, int nyu, int nyv, int ngz, int ngx, const float* s, double* p1,
                   double* p2, double* p3, const double* u1, const double* u2, const double* u3) {
  const size_t n = (size_t)nyu * ngz * ngx;
  hipLaunchKernelGGL(k_unpack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, y, nyu, nyv, ngz, ngx, s[0], s[1],
                     s[2], p1, p2, p3, u1, u2, u3);
}

void launch_add_planes(hipStream_t st, double2* a, const double2* c, int ncol, int q0, int q1) {
  if (q1 <= q0) return;
  const size_t n = (size_t)(q1 - q0) * ncol;
  hipLaunchKernelGGL(k_add_planes<double2>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, c, ncol, q0, q1);
}
void launch_add_planes(hipStream_t st, float2* a, const double2* c, int ncol, int q0, int q1) {
  if (q1 <= q0) return;
  const size_t n = (size_t)(q1 - q0) * ncol;
  hipLaunchKernelGGL(k_add_planes<float2>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, c, ncol, q0, q1);
}

void launch_zero_net_force(hipStream_t st, double2* c, int ncol, int q0, int q1, const double* yv) {
  if (q1 <= q0) return;
  hipLaunchKernelGGL(k_zero_net_force, dim3(1), dim3(1), 0, st, c, ncol, q0, q1, yv);
}

void launch_sumsq(hipStream_t st, const double* v, size_t n, double* acc) {
  hipLaunchKernelGGL(k_sumsq, dim3(512), dim3(256), 0, st, v, n, acc);
}

}  // namespace cfd
