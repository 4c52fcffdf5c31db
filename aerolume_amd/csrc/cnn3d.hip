// Small 3D convolutional network on a periodic grid (the learned eddy-viscosity closure):
// layers of 3x3x3 convolutions with BF16 weights and FP32 accumulation, forward, backward
// (data and weight gradients) and Adam, as hand-written HIP kernels for gfx950.
//
//   forward    y[co][p] = act(b[co] + sum_ci sum_k w[co][ci][k] x[ci][p + d_k])
//   backward   dx[ci][p] = sum_co sum_k w[co][ci][k] g[co][p - d_k]  (the forward kernel on
//              flipped, transposed weights), dw[co][ci][k] = sum_p g[co][p] x[ci][p + d_k],
//              db[co] = sum_p g[co][p], with g = dy * act'(pre-activation)
//
// Layout: channel-major fields (C, nz, ny, nx), x fastest; positions wrap periodically.
// Parameters: per layer w[co][ci][27] then b[co], float32 masters; the convolutions read a
// bfloat16 copy of w (round to nearest even), so forward and backward use the same rounded
// weights. No atomics: weight gradients go through per-block partials reduced in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolume_amd.h"
#include "common.h"

namespace aero {
namespace cnn {

constexpr int kMaxLayers = 8;
constexpr int TX = 8, TY = 8, TZ = 4;       // output positions per block (one per thread)
constexpr int kThreads = TX * TY * TZ;      // 256
constexpr int CI_CH = 8;                    // input channels staged per pass
constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
constexpr int kWgBricks = 16;              // 8x8x4 bricks per weight-gradient block

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));  // inf / nan
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;                        // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

__device__ __forceinline__ float act_f(int a, float x) {
  if (a == 1) return x > 0.f ? x : 0.f;
  if (a == 2) return x > 20.f ? x : log1pf(expf(x));      // softplus
  return x;
}
__device__ __forceinline__ float act_df(int a, float x) {
  if (a == 1) return x > 0.f ? 1.f : 0.f;
  if (a == 2) return x > 20.f ? 1.f : 1.f / (1.f + expf(-x));
  return 1.f;
}

// Forward convolution of one layer (COUT output channels per thread held in registers).
// w16: bf16 weights [co][ci][27]; b: bias (null: none); pre: pre-activation output (null: not
// kept); y: activated output. One thread per output position of an 8x8x4 block tile.
template <int COUT>
__global__ void __launch_bounds__(kThreads) k_conv_fwd(const float *__restrict__ x, const uint16_t *__restrict__ w16,
                                                      const float *__restrict__ b, float *__restrict__ pre,
                                                      float *__restrict__ y, int cin, int nz, int ny, int nx,
                                                      int act) {
  __shared__ float sx[CI_CH][HZ][HY][HX];
  __shared__ float sw[CI_CH][27][COUT];
  const int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  const int tx = threadIdx.x % TX, ty = (threadIdx.x / TX) % TY, tz = threadIdx.x / (TX * TY);
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
  const size_t nxyz = (size_t)nx * ny * nz;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = 0.f;
  for (int c0 = 0; c0 < cin; c0 += CI_CH) {
    __syncthreads();
    for (int e = threadIdx.x; e < CI_CH * HZ * HY * HX; e += kThreads) {
      const int hx = e % HX, t1 = e / HX, hy = t1 % HY, t2 = t1 / HY, hz = t2 % HZ, ci = t2 / HZ;
      float v = 0.f;
      if (c0 + ci < cin) {
        const int gx = wrapi(x0 - 1 + hx, nx), gy = wrapi(y0 - 1 + hy, ny), gz = wrapi(z0 - 1 + hz, nz);
        v = x[(size_t)(c0 + ci) * nxyz + ((size_t)gz * ny + gy) * nx + gx];
      }
      sx[ci][hz][hy][hx] = v;
    }
    for (int e = threadIdx.x; e < CI_CH * 27 * COUT; e += kThreads) {
      const int co = e % COUT, t1 = e / COUT, k = t1 % 27, ci = t1 / 27;
      sw[ci][k][co] = c0 + ci < cin ? bf16_to_f32(w16[((size_t)co * cin + c0 + ci) * 27 + k]) : 0.f;
    }
    __syncthreads();
    const int cn = min(CI_CH, cin - c0);
    for (int ci = 0; ci < cn; ++ci) {
#pragma unroll
      for (int k = 0; k < 27; ++k) {
        const int dz = k / 9, dy = (k / 3) % 3, dx = k % 3;
        const float xv = sx[ci][tz + dz][ty + dy][tx + dx];
#pragma unroll
        for (int co = 0; co < COUT; ++co) acc[co] += sw[ci][k][co] * xv;
      }
    }
  }
  const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
  if (gx >= nx || gy >= ny || gz >= nz) return;
  const size_t p = ((size_t)gz * ny + gy) * nx + gx;
#pragma unroll
  for (int co = 0; co < COUT; ++co) {
    const float v = acc[co] + (b ? b[co] : 0.f);
    if (pre) pre[(size_t)co * nxyz + p] = v;
    y[(size_t)co * nxyz + p] = act_f(act, v);
  }
}


// ------------------------------------------------------------------ MFMA forward convolution
// The same convolution for 32 output channels as an implicit GEMM on the matrix cores
// (v_mfma_f32_32x32x16_bf16): D[position][co] = sum_K A[position][K] B[K][co] with K = (tap,
// input channel). The weights are bf16 already; each float32 input value is split exactly into
// three bf16 parts (x = hi + mid + lo: truncations to 8 significant bits, then to 8 of the
// remainder, which leaves at most 8), so the three products with a bf16 weight are exact in
// float32 and the MFMA sums them in float32: the arithmetic of the scalar kernel (float32
// accumulation of exact products), in another order.
//
// Block: a 16x8x4 output brick, 512 threads; wave w takes rows y 4 (w & 1) .. + 3 of z-plane
// w / 2 as two 32-row M blocks of two 16-wide x rows each (so the 16 lanes of an LDS read group
// read 256 contiguous bytes: no bank conflicts). Input channels go in chunks of 8: the chunk's
// input brick with its one-cell halo (18x10x6 positions) is staged in LDS as the three bf16
// parts, position-major with the 8 channels of a position in 16 bytes, and the chunk's weights
// as [tap][co][8]. One K step of 16 is a pair of taps: lanes 0-31 read tap 2q, lanes 32-63 tap
// 2q + 1 (tap 27 has zero weights), 16 bytes each from LDS, so a K step costs 7 LDS reads and
// 6 MFMAs per lane.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int MX = 16, MY = 8, MZ = 4;                 // output brick of the MFMA kernel
constexpr int kMThreads = 512;                         // one wave per 16x4 half plane
constexpr int MHX = MX + 2, MHY = MY + 2, MHZ = MZ + 2, MHALO = MHX * MHY * MHZ;
constexpr int kMTaps = 28;                             // 27 taps and a zero one (pairs of taps)

// LDS position offset of tap k = (dz, dy, dx) in the staged halo brick (tap 27: the centre)
#define TAP_OFF(dz, dy, dx) (((dz) * MHY + (dy)) * MHX + (dx))
__constant__ int kTapOff[kMTaps] = {
    TAP_OFF(0, 0, 0), TAP_OFF(0, 0, 1), TAP_OFF(0, 0, 2), TAP_OFF(0, 1, 0), TAP_OFF(0, 1, 1), TAP_OFF(0, 1, 2),
    TAP_OFF(0, 2, 0), TAP_OFF(0, 2, 1), TAP_OFF(0, 2, 2), TAP_OFF(1, 0, 0), TAP_OFF(1, 0, 1), TAP_OFF(1, 0, 2),
    TAP_OFF(1, 1, 0), TAP_OFF(1, 1, 1), TAP_OFF(1, 1, 2), TAP_OFF(1, 2, 0), TAP_OFF(1, 2, 1), TAP_OFF(1, 2, 2),
    TAP_OFF(2, 0, 0), TAP_OFF(2, 0, 1), TAP_OFF(2, 0, 2), TAP_OFF(2, 1, 0), TAP_OFF(2, 1, 1), TAP_OFF(2, 1, 2),
    TAP_OFF(2, 2, 0), TAP_OFF(2, 2, 1), TAP_OFF(2, 2, 2), TAP_OFF(1, 1, 1)};
#undef TAP_OFF

struct MfmaLds {
  struct {
    uint4 x[3][MHALO];                                 // bf16 parts hi / mid / lo, 8 channels each
    uint4 w[kMTaps][32];                               // bf16 weights [tap][co][8 channels]
  } in;
};

__device__ __forceinline__ uint32_t hi16(float v) { return __float_as_uint(v) >> 16; }

// exact three-way bf16 split of v: the upper 16 bits of hi, mid, lo
__device__ __forceinline__ void split3(float v, uint32_t &h, uint32_t &m, uint32_t &l) {
  const float fh = __uint_as_float(__float_as_uint(v) & 0xffff0000u);
  const float r1 = v - fh;
  const float fm = __uint_as_float(__float_as_uint(r1) & 0xffff0000u);
  const float r2 = r1 - fm;
  h = hi16(fh);
  m = hi16(fm);
  l = hi16(r2);
}

__global__ void __launch_bounds__(kMThreads) k_conv_fwd_mfma(const float *__restrict__ x,
                                                           const uint16_t *__restrict__ w16,
                                                           const float *__restrict__ b, float *__restrict__ pre,
                                                           float *__restrict__ y, int cin, int nz, int ny, int nx,
                                                           int act) {
  extern __shared__ unsigned char smem[];
  MfmaLds &L = *reinterpret_cast<MfmaLds *>(smem);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x0 = blockIdx.x * MX, y0 = blockIdx.y * MY, z0 = blockIdx.z * MZ;
  const size_t nxyz = (size_t)nx * ny * nz;
  // this lane's rows: M block mb holds positions (x = m % 16, y = yb + 2 mb + m / 16) of plane wz
  const int m = lane & 31, half = lane >> 5;
  const int ax = m & 15, ay = m >> 4;
  const int wz = wave >> 1, yb = 4 * (wave & 1);
  f32x16_t acc[2];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[0][r] = acc[1][r] = 0.f;
  // the staging of a chunk: every thread loads its halo positions (e = tid, tid + 512, ...) and
  // weight slots (tid, tid + 512) into registers, then splits and stores them; their global
  // offsets are the same for every chunk and worked out once
  constexpr int kPos = (MHALO + kMThreads - 1) / kMThreads, kWs = (kMTaps * 32 + kMThreads - 1) / kMThreads;
  size_t gpos[kPos];
  int wsrc[kWs];
#pragma unroll
  for (int i = 0; i < kPos; ++i) {
    const int e = tid + kMThreads * i;
    const int ee = e < MHALO ? e : 0;
    const int hx = ee % MHX, t1 = ee / MHX, hy = t1 % MHY, hz = t1 / MHY;
    const int gx = wrapi(x0 - 1 + hx, nx), gy = wrapi(y0 - 1 + hy, ny), gz = wrapi(z0 - 1 + hz, nz);
    gpos[i] = ((size_t)gz * ny + gy) * nx + gx;
  }
#pragma unroll
  for (int i = 0; i < kWs; ++i) {
    const int e = tid + kMThreads * i;
    const int co = e & 31, k = e >> 5;
    wsrc[i] = k < 27 ? co * cin * 27 + k : -1;           // -1: the zero tap (or past the table)
  }
  float xr[kPos][8];
  uint32_t wr[kWs][4];
  auto fetch = [&](int c0) {
#pragma unroll
    for (int i = 0; i < kPos; ++i)
#pragma unroll
      for (int c = 0; c < 8; ++c) xr[i][c] = c0 + c < cin ? x[(size_t)(c0 + c) * nxyz + gpos[i]] : 0.f;
#pragma unroll
    for (int i = 0; i < kWs; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ca = c0 + 2 * q, cb = ca + 1;
        const uint32_t wa = (wsrc[i] >= 0 && ca < cin) ? w16[wsrc[i] + ca * 27] : 0u;
        const uint32_t wb = (wsrc[i] >= 0 && cb < cin) ? w16[wsrc[i] + cb * 27] : 0u;
        wr[i][q] = wa | (wb << 16);
      }
  };
  // LDS position of this lane's row in each M block; tap offsets are compile-time constants
  int base[2];
#pragma unroll
  for (int mb = 0; mb < 2; ++mb) base[mb] = (wz * MHY + (yb + 2 * mb + ay)) * MHX + ax;
  for (int c0 = 0; c0 < cin; c0 += 8) {
    fetch(c0);
    __syncthreads();
    // the chunk's input brick + halo, split into three bf16 parts, and its weights
#pragma unroll
    for (int i = 0; i < kPos; ++i) {
      const int e = tid + kMThreads * i;
      if (e >= MHALO) continue;
      uint32_t ph[4], pm[4], pl[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        uint32_t h0, m0, l0, h1, m1, l1;
        split3(xr[i][2 * q], h0, m0, l0);
        split3(xr[i][2 * q + 1], h1, m1, l1);
        ph[q] = h0 | (h1 << 16);
        pm[q] = m0 | (m1 << 16);
        pl[q] = l0 | (l1 << 16);
      }
      L.in.x[0][e] = make_uint4(ph[0], ph[1], ph[2], ph[3]);
      L.in.x[1][e] = make_uint4(pm[0], pm[1], pm[2], pm[3]);
      L.in.x[2][e] = make_uint4(pl[0], pl[1], pl[2], pl[3]);
    }
#pragma unroll
    for (int i = 0; i < kWs; ++i) {
      const int e = tid + kMThreads * i;
      if (e < kMTaps * 32) L.in.w[e >> 5][e & 31] = make_uint4(wr[i][0], wr[i][1], wr[i][2], wr[i][3]);
    }
    __syncthreads();
#pragma unroll 2
    for (int tp = 0; tp < kMTaps / 2; ++tp) {
      // taps 2 tp (lanes 0-31) and 2 tp + 1 (lanes 32-63); the zero tap 27 reads the centre
      const int off = half ? kTapOff[2 * tp + 1] : kTapOff[2 * tp];
      const uint4 wv = L.in.w[2 * tp + half][m];
      bf16x8_t bw;
      __builtin_memcpy(&bw, &wv, 16);
#pragma unroll
      for (int mb = 0; mb < 2; ++mb) {
        const int pos = base[mb] + off;
#pragma unroll
        for (int part = 0; part < 3; ++part) {
          const uint4 av = L.in.x[part][pos];
          bf16x8_t a;
          __builtin_memcpy(&a, &av, 16);
          acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw, acc[mb], 0, 0, 0);
        }
      }
    }
  }
  // results: registers 4g .. 4g + 3 of a lane hold four x-neighbours (M rows 8g + 4 half + 0..3,
  // x = row % 16, y = yb + 2 mb + row / 16) of channel m, stored as one 16-byte write per array
  const float bias = b ? b[m] : 0.f;
  const bool full = x0 + MX <= nx && (nx & 3) == 0;
#pragma unroll
  for (int mb = 0; mb < 2; ++mb)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int row = 8 * g + 4 * half;
      const int gx = x0 + (row & 15), gy = y0 + yb + 2 * mb + (row >> 4), gz = z0 + wz;
      if (gy >= ny || gz >= nz) continue;
      const size_t p = (size_t)m * nxyz + ((size_t)gz * ny + gy) * nx + gx;
      float v[4], a[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[q] = acc[mb][4 * g + q] + bias;
        a[q] = act_f(act, v[q]);
      }
      if (full) {
        if (pre) *reinterpret_cast<float4 *>(pre + p) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(y + p) = make_float4(a[0], a[1], a[2], a[3]);
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (gx + q >= nx) continue;
          if (pre) pre[p + q] = v[q];
          y[p + q] = a[q];
        }
      }
    }
}


// ------------------------------------------------------------------ one output channel
// The convolution to a single output channel (the network's last layer) in two passes:
// z[k][p] = sum_ci w[ci][k] x[ci][p] for every position p and tap k, a plain GEMM (positions x
// input channels) x (input channels x 27 taps) on the matrix cores with the same exact bf16
// split of x; then y[p] = act(b + sum_k z[k][p + d_k]), a 27-point gather. Each wave takes M
// blocks of 32 consecutive positions; a K step is 16 input channels.
__global__ void __launch_bounds__(256) k_conv_taps_mfma(const float *__restrict__ x, const uint16_t *__restrict__ w16,
                                                       float *__restrict__ z, int cin, size_t nxyz) {
  const int lane = threadIdx.x & 63, half = lane >> 5, col = lane & 31;
  const int nks = (cin + 15) / 16;
  // B fragments (weights of tap `col`, input channels 16 s + 8 half + i), the same for every block
  bf16x8_t bw[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    uint32_t pw[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ca = 16 * s + 8 * half + 2 * q, cb = ca + 1;
      const uint32_t wa = (s < nks && col < 27 && ca < cin) ? w16[ca * 27 + col] : 0u;
      const uint32_t wb = (s < nks && col < 27 && cb < cin) ? w16[cb * 27 + col] : 0u;
      pw[q] = wa | (wb << 16);
    }
    __builtin_memcpy(&bw[s], pw, 16);
  }
  const size_t nblk = (nxyz + 31) / 32;
  const size_t wstride = (size_t)gridDim.x * (blockDim.x >> 6);
  for (size_t blk = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); blk < nblk; blk += wstride) {
    const size_t p0 = blk * 32;
    const size_t p = p0 + col < nxyz ? p0 + col : nxyz - 1;   // row of this lane (clamped: stores skip it)
    f32x16_t acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (s >= nks) break;
      uint32_t ph[4], pm[4], pl[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int ca = 16 * s + 8 * half + 2 * q, cb = ca + 1;
        uint32_t h0, m0, l0, h1, m1, l1;
        split3(ca < cin ? x[(size_t)ca * nxyz + p] : 0.f, h0, m0, l0);
        split3(cb < cin ? x[(size_t)cb * nxyz + p] : 0.f, h1, m1, l1);
        ph[q] = h0 | (h1 << 16);
        pm[q] = m0 | (m1 << 16);
        pl[q] = l0 | (l1 << 16);
      }
      bf16x8_t a;
      __builtin_memcpy(&a, ph, 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[s], acc, 0, 0, 0);
      __builtin_memcpy(&a, pm, 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[s], acc, 0, 0, 0);
      __builtin_memcpy(&a, pl, 16);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bw[s], acc, 0, 0, 0);
    }
    // register 4g + q: position p0 + 8g + 4 half + q, tap col
    if (col < 27) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const size_t pr = p0 + 8 * g + 4 * half;
        float *dst = z + (size_t)col * nxyz + pr;
        if (pr + 3 < nxyz && (nxyz & 3) == 0) {
          *reinterpret_cast<float4 *>(dst) = make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (pr + q < nxyz) dst[q] = acc[4 * g + q];
        }
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_conv_taps_sum(const float *__restrict__ z, const float *__restrict__ b,
                                                      float *__restrict__ pre, float *__restrict__ y, int nz, int ny,
                                                      int nx, int act) {
  const size_t nxyz = (size_t)nx * ny * nz;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < nxyz; p += (size_t)gridDim.x * blockDim.x) {
    const int gx = (int)(p % nx), gy = (int)((p / nx) % ny), gz = (int)(p / ((size_t)nx * ny));
    float v = 0.f;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
      const int dz = k / 9 - 1, dy = (k / 3) % 3 - 1, dx = k % 3 - 1;
      const size_t q = ((size_t)wrapi(gz + dz, nz) * ny + wrapi(gy + dy, ny)) * nx + wrapi(gx + dx, nx);
      v += z[(size_t)k * nxyz + q];
    }
    v = v + (b ? b[0] : 0.f);
    if (pre) pre[p] = v;
    y[p] = act_f(act, v);
  }
}

// g = dy * act'(pre) in place into g (elementwise).
__global__ void k_act_grad(const float *__restrict__ dy, const float *__restrict__ pre, float *__restrict__ g, size_t n,
                           int act) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
    g[e] = dy[e] * act_df(act, pre[e]);
}

// Weights for the data gradient: wt[ci][co][k] = w[co][ci][26 - k] (flipped, transposed), bf16.
__global__ void k_flip(const uint16_t *__restrict__ w16, uint16_t *__restrict__ wt, int cout, int cin) {
  const int n = cout * cin * 27;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int k = e % 27, t = e / 27, co = t % cout, ci = t / cout;
    wt[e] = w16[((size_t)co * cin + ci) * 27 + (26 - k)];
  }
}

// Weight-gradient partials: block (input-channel chunk, run of bricks) stages, per 8x8x4
// brick, the output gradient g of every co and the input brick with its one-cell halo; thread
// (ci, k) accumulates dw[co][ci][k] for all co over the bricks. Blocks of chunk 0 also sum
// db[co]. Partials: [position chunk][co][cin][27], then [position chunk][co] for db.
template <int COUT>
__global__ void __launch_bounds__(256) k_conv_wgrad(const float *__restrict__ x, const float *__restrict__ g,
                                                   float *__restrict__ part, int cin, int nz, int ny, int nx,
                                                   int bricks_per_block) {
  __shared__ float sg[COUT][kThreads];
  __shared__ float sx[CI_CH][HZ][HY][HX];
  const int c0 = blockIdx.x * CI_CH;
  const int nbx = (nx + TX - 1) / TX, nby = (ny + TY - 1) / TY, nbz = (nz + TZ - 1) / TZ;
  const int nbr = nbx * nby * nbz;
  const int br0 = blockIdx.y * bricks_per_block, br1 = min(nbr, br0 + bricks_per_block);
  const size_t nxyz = (size_t)nx * ny * nz;
  const int t = threadIdx.x;
  const int ci = t / 27, k = t % 27;
  const bool work = t < CI_CH * 27 && c0 + ci < cin;
  const int dz = k / 9, dy = (k / 3) % 3, dx = k % 3;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = 0.f;
  float bacc = 0.f;
  for (int br = br0; br < br1; ++br) {
    const int bx = br % nbx, by = (br / nbx) % nby, bz = br / (nbx * nby);
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
    __syncthreads();
    for (int e = t; e < COUT * kThreads; e += 256) {
      const int co = e / kThreads, q = e % kThreads;
      const int gx = x0 + q % TX, gy = y0 + (q / TX) % TY, gz = z0 + q / (TX * TY);
      sg[co][q] = (gx < nx && gy < ny && gz < nz) ? g[(size_t)co * nxyz + ((size_t)gz * ny + gy) * nx + gx] : 0.f;
    }
    for (int e = t; e < CI_CH * HZ * HY * HX; e += 256) {
      const int hx = e % HX, t1 = e / HX, hy = t1 % HY, t2 = t1 / HY, hz = t2 % HZ, cc = t2 / HZ;
      float v = 0.f;
      if (c0 + cc < cin) {
        const int gx = wrapi(x0 - 1 + hx, nx), gy = wrapi(y0 - 1 + hy, ny), gz = wrapi(z0 - 1 + hz, nz);
        v = x[(size_t)(c0 + cc) * nxyz + ((size_t)gz * ny + gy) * nx + gx];
      }
      sx[cc][hz][hy][hx] = v;
    }
    __syncthreads();
    if (work) {
      for (int q = 0; q < kThreads; ++q) {
        const int qx = q % TX, qy = (q / TX) % TY, qz = q / (TX * TY);
        const float xv = sx[ci][qz + dz][qy + dy][qx + dx];
#pragma unroll
        for (int co = 0; co < COUT; ++co) acc[co] += sg[co][q] * xv;
      }
    }
    if (blockIdx.x == 0 && t < COUT)
      for (int q = 0; q < kThreads; ++q) bacc += sg[t][q];
  }
  const int npc = gridDim.y;
  if (work) {
#pragma unroll
    for (int co = 0; co < COUT; ++co) part[(((size_t)blockIdx.y * COUT + co) * cin + c0 + ci) * 27 + k] = acc[co];
  }
  if (blockIdx.x == 0 && t < COUT) part[(size_t)npc * COUT * cin * 27 + (size_t)blockIdx.y * COUT + t] = bacc;
}


// ------------------------------------------------------------------ MFMA weight gradient
// dw[co][ci][k] = sum_p g[co][p] x[ci][p + d_k] for 32 output channels as GEMMs over positions
// on the matrix cores: D[co][(ci, k)] with K = positions. g and x are split into two bf16 parts
// each (hi + lo, 16 significant bits) and the three products hi*hi + hi*lo + lo*hi are summed in
// float32 (relative error of a product about 2^-16). db[co] = sum_p g[co][p] in float32.
//
// Block: 512 threads, bricks of 16x8x4 positions (one CU's worth; a block walks `bricks`
// consecutive bricks). Per brick, g of all 32 channels is staged as [part][co][position]; per
// chunk of 8 input channels, x with its halo is staged three times, shifted by dx = 0, 1, 2
// along x, as [dx][part][ci][z][y][16], so that the 8 positions of a K slice start on a 16-byte
// boundary for every tap. A K step is one 16-wide x row of the brick; wave w < 7 accumulates the
// columns (ci, k) = (n % 8, 4 w + n / 8) of each chunk (the eighth wave only stages).
constexpr int WX = 16, WY = 8, WZ = 4, WPOS = WX * WY * WZ;
constexpr int WHY = WY + 2, WHZ = WZ + 2;
constexpr int kWChunks = 4;                            // up to 32 input channels

// Rows are padded by 16 bytes (g: per channel, x: per input channel of a copy) so that the 16
// lanes of an LDS read group, which read different channels, hit different banks.
constexpr int kWgRow = WPOS / 8 + 1, kWxCi = WHZ * WHY * (WX / 8) + 1;
struct WgradLds {
  uint4 g[2][32][kWgRow];                              // bf16 g parts, 8 positions per 16 bytes
  uint4 x[3][2][8][kWxCi];                             // bf16 x parts per dx shift, [z][y][16] per ci
};

__device__ __forceinline__ void split2(float v, uint32_t &h, uint32_t &l) {
  const float fh = __uint_as_float(__float_as_uint(v) & 0xffff0000u);
  h = hi16(fh);
  l = hi16(v - fh);
}

__global__ void __launch_bounds__(512) k_conv_wgrad_mfma(const float *__restrict__ x, const float *__restrict__ g,
                                                        float *__restrict__ part, int cin, int nz, int ny, int nx,
                                                        int bricks) {
  extern __shared__ unsigned char smem[];
  WgradLds &L = *reinterpret_cast<WgradLds *>(smem);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, half = lane >> 5, n = lane & 31;
  const size_t nxyz = (size_t)nx * ny * nz;
  const int nbx = (nx + WX - 1) / WX, nby = (ny + WY - 1) / WY, nbz = (nz + WZ - 1) / WZ;
  const int nbr = nbx * nby * nbz;
  const int nch = (cin + 7) / 8;
  // this lane's column: input channel n % 8 of the chunk, tap 4 wave + n / 8 (27: none)
  const int tap = 4 * wave + (n >> 3), cil = n & 7;
  const int tk = tap < 27 ? tap : 13;
  const int tdz = tk / 9, tdy = (tk / 3) % 3, tdx = tk % 3;
  f32x16_t acc[kWChunks];
#pragma unroll
  for (int c = 0; c < kWChunks; ++c)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[c][r] = 0.f;
  float dbs[32];                                        // this thread's db shares per channel
#pragma unroll
  for (int co = 0; co < 32; ++co) dbs[co] = 0.f;
  const int br0 = blockIdx.x * bricks, br1 = min(nbr, br0 + bricks);
  for (int br = br0; br < br1; ++br) {
    const int bx = br % nbx, by = (br / nbx) % nby, bz = br / (nbx * nby);
    const int x0 = bx * WX, y0 = by * WY, z0 = bz * WZ;
    __syncthreads();
    // g of the brick (zero outside the domain) in two parts, one position per thread slot
    // (consecutive threads: consecutive x), and the thread's db shares: slot e = tid + 512 i
    // holds channel e / 512 = i
    {
      uint16_t *gh = reinterpret_cast<uint16_t *>(&L.g[0][0][0]), *gl = reinterpret_cast<uint16_t *>(&L.g[1][0][0]);
      const int px = tid % WX, py = (tid / WX) % WY, pz = tid / (WX * WY);
      const int gx = x0 + px, gy = y0 + py, gz = z0 + pz;
      const bool in = gx < nx && gy < ny && gz < nz;
      const size_t pp = in ? ((size_t)gz * ny + gy) * nx + gx : 0;
#pragma unroll 8
      for (int co = 0; co < 32; ++co) {
        const float v = in ? g[(size_t)co * nxyz + pp] : 0.f;
        dbs[co] += v;
        uint32_t h, l;
        split2(v, h, l);
        gh[co * kWgRow * 8 + tid] = (uint16_t)h;
        gl[co * kWgRow * 8 + tid] = (uint16_t)l;
      }
    }
    for (int c = 0; c < nch; ++c) {
      const int c0 = 8 * c;
      __syncthreads();
      // x of the chunk with its halo (18 positions per row), two parts, each value loaded once
      // (consecutive threads: consecutive x) and written to the copies shifted by dx = 0, 1, 2
      {
        constexpr int HXW = WX + 2;
        uint16_t *xb = reinterpret_cast<uint16_t *>(&L.x[0][0][0][0]);
        constexpr int kCi = kWxCi * 8, kPart = 8 * kCi, kShift = 2 * kPart;   // uint16 strides
        for (int e = tid; e < 8 * WHZ * WHY * HXW; e += 512) {
          const int hx = e % HXW, t1 = e / HXW, hy = t1 % WHY, t2 = t1 / WHY, hz = t2 % WHZ, ci = t2 / WHZ;
          const int gx = wrapi(x0 - 1 + hx, nx), gy = wrapi(y0 - 1 + hy, ny), gz = wrapi(z0 - 1 + hz, nz);
          const float v = c0 + ci < cin ? x[(size_t)(c0 + ci) * nxyz + ((size_t)gz * ny + gy) * nx + gx] : 0.f;
          uint32_t h, l;
          split2(v, h, l);
          const int rowbase = ci * kCi + (hz * WHY + hy) * WX;
#pragma unroll
          for (int dx = 0; dx < 3; ++dx) {
            const int j = hx - dx;                     // copy dx holds x0 + j + dx - 1 at j
            if (j < 0 || j >= WX) continue;
            xb[dx * kShift + rowbase + j] = (uint16_t)h;
            xb[dx * kShift + kPart + rowbase + j] = (uint16_t)l;
          }
        }
      }
      __syncthreads();
      if (wave < 7) {
#pragma unroll 2
        for (int ks = 0; ks < WY * WZ; ++ks) {
          const int ry = ks % WY, rz = ks / WY;          // the K step's row of the brick
          bf16x8_t ah, al, bh, bl;
          const uint4 gh = L.g[0][n][(rz * WY + ry) * 2 + half], gl = L.g[1][n][(rz * WY + ry) * 2 + half];
          const int xo = ((rz + tdz) * WHY + (ry + tdy)) * 2 + half;
          const uint4 xh = L.x[tdx][0][cil][xo];
          const uint4 xl = L.x[tdx][1][cil][xo];
          __builtin_memcpy(&ah, &gh, 16);
          __builtin_memcpy(&al, &gl, 16);
          __builtin_memcpy(&bh, &xh, 16);
          __builtin_memcpy(&bl, &xl, 16);
#pragma unroll
          for (int cc = 0; cc < kWChunks; ++cc)
            if (cc == c) {
              acc[cc] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[cc], 0, 0, 0);
              acc[cc] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[cc], 0, 0, 0);
              acc[cc] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[cc], 0, 0, 0);
            }
        }
      }
    }
  }
  // partials of this block, in the layout of the scalar kernel: [block][co][cin][27], then db
  const int npc = gridDim.x;
  if (wave < 7 && tap < 27) {
#pragma unroll
    for (int cc = 0; cc < kWChunks; ++cc) {
      const int ci = 8 * cc + cil;
      if (ci >= cin) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int co = 8 * (r >> 2) + 4 * half + (r & 3);
        part[(((size_t)blockIdx.x * 32 + co) * cin + ci) * 27 + tap] = acc[cc][r];
      }
    }
  }
  // db: per channel, the shares of the block's threads summed in a fixed order (waves, then
  // the eight wave sums through LDS, which the staging no longer needs)
  __syncthreads();
  float *red = reinterpret_cast<float *>(&L.g[0][0][0]);
#pragma unroll
  for (int co = 0; co < 32; ++co) {
    const float sw = wave_sum(dbs[co]);
    if (lane == 0) red[co * 8 + wave] = sw;
  }
  __syncthreads();
  if (tid < 32) {
    float sdb = 0.f;
    for (int w = 0; w < 8; ++w) sdb += red[tid * 8 + w];
    part[(size_t)npc * 32 * cin * 27 + (size_t)blockIdx.x * 32 + tid] = sdb;
  }
}

// Sum the partials over position chunks into grad (w then b of the layer): a block takes 64
// values, each of its four thread rows a quarter of the chunks, and the quarter sums are added
// in order (a fixed order: the result does not depend on timing).
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float *__restrict__ part, float *__restrict__ grad, int nw,
                                                      int nb, int npc) {
  __shared__ float qs[4][64];
  const int n = nw + nb;
  const int el = threadIdx.x & 63, qr = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + el;
  const int cq = (npc + 3) / 4, ca = qr * cq, cb = min(npc, ca + cq);
  float s = 0.f;
  if (e < nw) {
    for (int c = ca; c < cb; ++c) s += part[(size_t)c * nw + e];
  } else if (e < n) {
    const int co = e - nw;
    for (int c = ca; c < cb; ++c) s += part[(size_t)npc * nw + (size_t)c * nb + co];
  }
  qs[qr][el] = s;
  __syncthreads();
  if (qr == 0 && e < n) grad[e] = ((qs[0][el] + qs[1][el]) + qs[2][el]) + qs[3][el];
}

// Adam on the float32 masters, then the bf16 copy of every weight (biases stay float32).
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const float ge = g[e];
    const float me = b1 * m[e] + (1.f - b1) * ge;
    const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
  }
}

__global__ void k_to_bf16(const float *__restrict__ src, uint16_t *__restrict__ dst, int n) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) dst[e] = f32_to_bf16(src[e]);
}

}  // namespace cnn
}  // namespace aero

using namespace aero;
using namespace aero::cnn;

struct aero_cnn {
  aero_cnn_desc desc;
  size_t nxyz = 0;
  int nparam = 0;
  int woff[kMaxLayers], boff[kMaxLayers];
  std::vector<void *> allocs;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr, *part = nullptr;
  uint16_t *w16 = nullptr, *wt16 = nullptr;
  float *act[kMaxLayers + 1];   // act[0] = the input copy, act[l + 1] = layer l output
  float *pre[kMaxLayers];
  float *g0 = nullptr, *g1 = nullptr;
  size_t bytes = 0;
};

namespace {

int cnn_alloc(aero_cnn *c, void **ptr, size_t n) {
  *ptr = nullptr;
  if (!n) return AERO_OK;
  AERO_HIP_CHECK(hipMalloc(ptr, n));
  c->allocs.push_back(*ptr);
  c->bytes += n;
  AERO_HIP_CHECK(hipMemset(*ptr, 0, n));
  return AERO_OK;
}

int max_ch(const aero_cnn_desc &d) {
  int m = 0;
  for (int l = 0; l <= d.n_layers; ++l) m = std::max(m, d.channels[l]);
  return m;
}

// zbuf: 27 x cells of scratch for the single-output-channel path (null: the scalar kernel)
void launch_fwd(hipStream_t st, const float *x, const uint16_t *w, const float *b, float *pre, float *y, int cin,
                int cout, int nz, int ny, int nx, int act, float *zbuf = nullptr) {
  if (cout == 1 && zbuf && cin <= 32) {
    const size_t nxyz = (size_t)nx * ny * nz;
    const int nb = (int)std::min<size_t>(2048, ((nxyz + 31) / 32 + 3) / 4);
    hipLaunchKernelGGL(k_conv_taps_mfma, dim3(nb), dim3(256), 0, st, x, w, zbuf, cin, nxyz);
    const int ns = (int)std::min<size_t>(4096, (nxyz + 255) / 256);
    hipLaunchKernelGGL(k_conv_taps_sum, dim3(ns), dim3(256), 0, st, zbuf, b, pre, y, nz, ny, nx, act);
    return;
  }
  if (cout == 32) {
    // 32 output channels: the matrix-core kernel
    const dim3 mgrid((nx + MX - 1) / MX, (ny + MY - 1) / MY, (nz + MZ - 1) / MZ);
    hipLaunchKernelGGL(k_conv_fwd_mfma, mgrid, dim3(kMThreads), sizeof(MfmaLds), st, x, w, b, pre, y, cin, nz, ny,
                       nx, act);
    return;
  }
  const dim3 grid((nx + TX - 1) / TX, (ny + TY - 1) / TY, (nz + TZ - 1) / TZ);
  if (cout == 1)
    hipLaunchKernelGGL(k_conv_fwd<1>, grid, dim3(kThreads), 0, st, x, w, b, pre, y, cin, nz, ny, nx, act);
  else if (cout == 8)
    hipLaunchKernelGGL(k_conv_fwd<8>, grid, dim3(kThreads), 0, st, x, w, b, pre, y, cin, nz, ny, nx, act);
  else if (cout == 9)
    hipLaunchKernelGGL(k_conv_fwd<9>, grid, dim3(kThreads), 0, st, x, w, b, pre, y, cin, nz, ny, nx, act);
  else if (cout == 16)
    hipLaunchKernelGGL(k_conv_fwd<16>, grid, dim3(kThreads), 0, st, x, w, b, pre, y, cin, nz, ny, nx, act);
  else
    hipLaunchKernelGGL(k_conv_fwd<32>, grid, dim3(kThreads), 0, st, x, w, b, pre, y, cin, nz, ny, nx, act);
}

bool cout_ok(int c) { return c == 1 || c == 8 || c == 9 || c == 16 || c == 32; }

// blocks of the matrix-core weight gradient (one per CU's worth of bricks, at most 256) and
// the bricks each walks
int wgrad_bricks_total(const aero_cnn_desc &d) {
  return ((d.nx + WX - 1) / WX) * ((d.ny + WY - 1) / WY) * ((d.nz + WZ - 1) / WZ);
}
int wgrad_blocks(const aero_cnn_desc &d) {
  const int nbr = wgrad_bricks_total(d);
  const int per = (nbr + 255) / 256;
  return (nbr + per - 1) / per;
}

}  // namespace

extern "C" {

int aero_cnn_create(const aero_cnn_desc *d, aero_cnn **out) {
  if (!out || !d) { set_error("null argument"); return AERO_EINVAL; }
  *out = nullptr;
  if (d->n_layers < 1 || d->n_layers > kMaxLayers) { set_error("1..8 layers"); return AERO_EINVAL; }
  if (d->nx < 3 || d->ny < 3 || d->nz < 3) { set_error("grid of at least 3 cells per axis"); return AERO_EINVAL; }
  for (int l = 0; l < d->n_layers; ++l) {
    if (d->channels[l] < 1 || d->channels[l] > 32) { set_error("1..32 input channels per layer"); return AERO_EINVAL; }
    if (!cout_ok(d->channels[l + 1])) {
      set_error("output channels per layer must be 1, 8, 9, 16 or 32");
      return AERO_EUNSUPPORTED;
    }
    if (d->act[l] < 0 || d->act[l] > 2) { set_error("activation 0 (none), 1 (relu) or 2 (softplus)"); return AERO_EINVAL; }
  }
  // the matrix-core convolution stages more than the default 64 KB of LDS
  AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_conv_fwd_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(MfmaLds)));
  AERO_HIP_CHECK(hipFuncSetAttribute((const void *)k_conv_wgrad_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)sizeof(WgradLds)));
  aero_cnn *c = new aero_cnn();
  c->desc = *d;
  c->nxyz = (size_t)d->nx * d->ny * d->nz;
  int off = 0, nw_max = 0;
  for (int l = 0; l < d->n_layers; ++l) {
    c->woff[l] = off;
    off += d->channels[l + 1] * d->channels[l] * 27;
    c->boff[l] = off;
    off += d->channels[l + 1];
    nw_max = std::max(nw_max, d->channels[l + 1] * d->channels[l] * 27);
  }
  c->nparam = off;
  const int mc = max_ch(*d);
  const size_t nbr = (size_t)((d->nx + TX - 1) / TX) * ((d->ny + TY - 1) / TY) * ((d->nz + TZ - 1) / TZ);
  // weight-gradient partials: the scalar kernel's position chunks or the matrix-core kernel's blocks
  const size_t npc = std::max((nbr + kWgBricks - 1) / kWgBricks, (size_t)wgrad_blocks(*d));
  int rc = AERO_OK;
  auto A = [&](void **p, size_t n) { if (!rc) rc = cnn_alloc(c, p, n); };
  A((void **)&c->params, off * 4);
  A((void **)&c->grads, off * 4);
  A((void **)&c->adam_m, off * 4);
  A((void **)&c->adam_v, off * 4);
  A((void **)&c->w16, off * 2);
  A((void **)&c->wt16, (size_t)nw_max * 2);
  A((void **)&c->part, (npc * (size_t)(nw_max + 32)) * 4);
  for (int l = 0; l <= d->n_layers; ++l) A((void **)&c->act[l], (size_t)d->channels[l] * c->nxyz * 4);
  for (int l = 0; l < d->n_layers; ++l) A((void **)&c->pre[l], (size_t)d->channels[l + 1] * c->nxyz * 4);
  // gradient buffers; between backward passes they also hold the 27 tap sums of a
  // single-output-channel layer
  const size_t gch = (size_t)std::max(mc, 27);
  A((void **)&c->g0, gch * c->nxyz * 4);
  A((void **)&c->g1, gch * c->nxyz * 4);
  if (rc) {
    aero_cnn_destroy(c);
    return rc;
  }
  *out = c;
  return AERO_OK;
}

int aero_cnn_destroy(aero_cnn *c) {
  if (!c) return AERO_OK;
  for (void *a : c->allocs) (void)hipFree(a);
  delete c;
  return AERO_OK;
}

int aero_cnn_param_count(const aero_cnn *c) { return c ? c->nparam : 0; }
size_t aero_cnn_workspace_bytes(const aero_cnn *c) { return c ? c->bytes : 0; }

int aero_cnn_set_params(aero_cnn *c, void *stream, const float *params) {
  if (!c || !params) { set_error("null argument"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  AERO_HIP_CHECK(hipMemcpyAsync(c->params, params, (size_t)c->nparam * 4, hipMemcpyDeviceToDevice, st));
  AERO_HIP_CHECK(hipMemsetAsync(c->adam_m, 0, (size_t)c->nparam * 4, st));
  AERO_HIP_CHECK(hipMemsetAsync(c->adam_v, 0, (size_t)c->nparam * 4, st));
  hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, st, c->params, c->w16, c->nparam);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

int aero_cnn_get_params(aero_cnn *c, void *stream, float *params) {
  if (!c || !params) { set_error("null argument"); return AERO_EINVAL; }
  AERO_HIP_CHECK(hipMemcpyAsync(params, c->params, (size_t)c->nparam * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return AERO_OK;
}

int aero_cnn_get_grads(aero_cnn *c, void *stream, float *grads) {
  if (!c || !grads) { set_error("null argument"); return AERO_EINVAL; }
  AERO_HIP_CHECK(hipMemcpyAsync(grads, c->grads, (size_t)c->nparam * 4, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return AERO_OK;
}

int aero_cnn_forward(aero_cnn *c, void *stream, const float *x, float *y) {
  if (!c || !x || !y) { set_error("null argument"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  const aero_cnn_desc &d = c->desc;
  AERO_HIP_CHECK(hipMemcpyAsync(c->act[0], x, (size_t)d.channels[0] * c->nxyz * 4, hipMemcpyDeviceToDevice, st));
  for (int l = 0; l < d.n_layers; ++l) {
    float *dst = l == d.n_layers - 1 ? y : c->act[l + 1];
    launch_fwd(st, c->act[l], c->w16 + c->woff[l], c->params + c->boff[l], c->pre[l], dst, d.channels[l],
               d.channels[l + 1], d.nz, d.ny, d.nx, d.act[l], c->g0);
    if (l == d.n_layers - 1)
      AERO_HIP_CHECK(hipMemcpyAsync(c->act[l + 1], y, (size_t)d.channels[l + 1] * c->nxyz * 4,
                                    hipMemcpyDeviceToDevice, st));
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

int aero_cnn_infer(aero_cnn *c, void *stream, const float *x, float *y) {
  if (!c || !x || !y) { set_error("null argument"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  const aero_cnn_desc &d = c->desc;
  const float *src = x;
  for (int l = 0; l < d.n_layers; ++l) {
    float *dst = l == d.n_layers - 1 ? y : (l % 2 == 0 ? c->g0 : c->g1);
    // the buffer src is not in holds the taps of a single-output-channel layer
    float *zb = src == c->g0 ? c->g1 : c->g0;
    launch_fwd(st, src, c->w16 + c->woff[l], c->params + c->boff[l], nullptr, dst, d.channels[l], d.channels[l + 1],
               d.nz, d.ny, d.nx, d.act[l], zb);
    src = dst;
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

int aero_cnn_backward(aero_cnn *c, void *stream, const float *dy, float *dx) {
  if (!c || !dy) { set_error("null argument"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  const aero_cnn_desc &d = c->desc;
  const size_t n = c->nxyz;
  const int nbr = ((d.nx + TX - 1) / TX) * ((d.ny + TY - 1) / TY) * ((d.nz + TZ - 1) / TZ);
  const int npc = (nbr + kWgBricks - 1) / kWgBricks;
  const float *up = dy;       // gradient w.r.t. the layer's activated output
  for (int l = d.n_layers - 1; l >= 0; --l) {
    const int ci = d.channels[l], co = d.channels[l + 1];
    float *g = (l % 2 == 0) ? c->g0 : c->g1;            // gradient w.r.t. the pre-activation
    hipLaunchKernelGGL(k_act_grad, dim3(1024), dim3(256), 0, st, up, c->pre[l], g, (size_t)co * n, d.act[l]);
    const dim3 wg((ci + CI_CH - 1) / CI_CH, npc);
    int npart = npc;
    if (co == 32) {
      const int nbr_m = wgrad_bricks_total(d), nblk = wgrad_blocks(d);
      const int per = (nbr_m + nblk - 1) / nblk;
      hipLaunchKernelGGL(k_conv_wgrad_mfma, dim3(nblk), dim3(512), sizeof(WgradLds), st, c->act[l], g, c->part, ci,
                         d.nz, d.ny, d.nx, per);
      npart = nblk;
    } else if (co == 1)
      hipLaunchKernelGGL(k_conv_wgrad<1>, wg, dim3(256), 0, st, c->act[l], g, c->part, ci, d.nz, d.ny, d.nx,
                         kWgBricks);
    else if (co == 8)
      hipLaunchKernelGGL(k_conv_wgrad<8>, wg, dim3(256), 0, st, c->act[l], g, c->part, ci, d.nz, d.ny, d.nx,
                         kWgBricks);
    else if (co == 9)
      hipLaunchKernelGGL(k_conv_wgrad<9>, wg, dim3(256), 0, st, c->act[l], g, c->part, ci, d.nz, d.ny, d.nx,
                         kWgBricks);
    else if (co == 16)
      hipLaunchKernelGGL(k_conv_wgrad<16>, wg, dim3(256), 0, st, c->act[l], g, c->part, ci, d.nz, d.ny, d.nx,
                         kWgBricks);
    else
      hipLaunchKernelGGL(k_conv_wgrad<32>, wg, dim3(256), 0, st, c->act[l], g, c->part, ci, d.nz, d.ny, d.nx,
                         kWgBricks);
    const int nw = co * ci * 27;
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((nw + co + 63) / 64), dim3(256), 0, st, c->part, c->grads + c->woff[l],
                       nw, co, npart);
    if (l > 0 || dx) {
      // data gradient: the forward convolution with flipped, transposed weights, no bias
      hipLaunchKernelGGL(k_flip, dim3(64), dim3(256), 0, st, c->w16 + c->woff[l], c->wt16, co, ci);
      float *dst = l > 0 ? ((l % 2 == 0) ? c->g1 : c->g0) : dx;
      if (!cout_ok(ci)) {
        set_error("input channels of a layer must be 1, 8, 9, 16 or 32 for its data gradient");
        return AERO_EUNSUPPORTED;
      }
      launch_fwd(st, g, c->wt16, nullptr, nullptr, dst, co, ci, d.nz, d.ny, d.nx, 0);
      up = dst;
    }
  }
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

int aero_cnn_adam_step(aero_cnn *c, void *stream, float lr, float beta1, float beta2, float eps, int t) {
  if (!c || t < 1) { set_error("bad Adam arguments"); return AERO_EINVAL; }
  hipStream_t st = (hipStream_t)stream;
  const float c1 = 1.f - std::pow(beta1, (float)t), c2 = 1.f - std::pow(beta2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, st, c->params, c->grads, c->adam_m, c->adam_v, c->nparam, lr,
                     beta1, beta2, eps, c1, c2);
  hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, st, c->params, c->w16, c->nparam);
  AERO_HIP_CHECK(hipGetLastError());
  return AERO_OK;
}

}  // extern "C"
