// Shared device helpers of the MI355X Navier-Stokes kernels: complex arithmetic,
// LDS Stockham FFTs of arbitrary 2^a 3^b 5^c length, real transforms built on them
// (DCT-II by Makhoul's reordering, discrete Hartley), deterministic block reductions
// and order-independent max atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace aero {

// ------------------------------------------------------------------ errors (host)
void set_error(const std::string &msg);

#define AERO_HIP_CHECK(expr)                                                        \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::aero::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));         \
      return AERO_EHIP;                                                             \
    }                                                                               \
  } while (0)

// ------------------------------------------------------------------ complex
template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }

// LDS a workgroup may allocate on gfx950 (160 KB per CU, all of it to one workgroup);
// dynamic allocations above 64 KB need hipFuncAttributeMaxDynamicSharedMemorySize.
constexpr size_t kLdsBlockMax = 163840;

// ------------------------------------------------------------------ FFT plan
// Radix sequence of an N-point transform and the twiddle table exp(-2 pi i m / N).
constexpr int kMaxRadices = 24;
struct FftPlan {
  int n;
  int nrad;
  int radix[kMaxRadices];
};

// Twiddle table of an n-point transform in the order its Stockham stages read it: the stage
// of span p (the product of the earlier radices) and radix R holds w^(r i n / (p R)) for
// r = 1 .. R-1, i = 0 .. p-1 at (p - 1) + (r - 1) p + i, w = exp(-2 pi i / n) (the stages'
// blocks add up to n - 1 entries). Consecutive butterflies of a stage read consecutive
// entries (no LDS bank conflicts); the values are those of the natural table tw[r i n / (p R)].
template <typename T, typename Plan>
inline void stage_twiddles(const Plan &pl, const cplx<T> *natural, cplx<T> *out) {
  const int n = pl.n;
  int p = 1;
  for (int s = 0; s < pl.nrad; ++s) {
    const int R = pl.radix[s], stride = n / (p * R);
    for (int r = 1; r < R; ++r)
      for (int i = 0; i < p; ++i) out[(p - 1) + (r - 1) * p + i] = natural[r * i * stride];
    p *= R;
  }
  out[n - 1] = natural[0];
}

// Radix sequence of an n-point transform (4s first, then 2, 3, 5); false when n has
// another prime factor.
inline bool factor_radices(int n, FftPlan &pl) {
  pl.n = n;
  pl.nrad = 0;
  int m = n;
  while (m % 4 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 4; m /= 4; }
  while (m % 2 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 2; m /= 2; }
  while (m % 3 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 3; m /= 3; }
  while (m % 5 == 0 && pl.nrad < kMaxRadices) { pl.radix[pl.nrad++] = 5; m /= 5; }
  return m == 1;
}

// Forward radix-R butterflies (sign -1).
template <typename T>
__device__ __forceinline__ void dft2(cplx<T> *v) {
  cplx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <typename T>
__device__ __forceinline__ void dft3(cplx<T> *v) {
  const T s = (T)0.86602540378443864676;  // sin(2 pi / 3)
  cplx<T> a = v[0], b = v[1], c = v[2];
  cplx<T> p = cadd(b, c), z = csub(b, c);
  cplx<T> t = {a.x - (T)0.5 * p.x, a.y - (T)0.5 * p.y};
  v[0] = cadd(a, p);
  v[1] = {t.x + s * z.y, t.y - s * z.x};
  v[2] = {t.x - s * z.y, t.y + s * z.x};
}
template <typename T>
__device__ __forceinline__ void dft4(cplx<T> *v) {
  cplx<T> s0 = cadd(v[0], v[2]), s1 = csub(v[0], v[2]);
  cplx<T> s2 = cadd(v[1], v[3]), s3 = csub(v[1], v[3]);
  v[0] = cadd(s0, s2);
  v[2] = csub(s0, s2);
  v[1] = {s1.x + s3.y, s1.y - s3.x};
  v[3] = {s1.x - s3.y, s1.y + s3.x};
}
template <typename T>
__device__ __forceinline__ void dft5(cplx<T> *v) {
  const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
  const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
  cplx<T> a0 = v[0];
  cplx<T> b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
  cplx<T> d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
  cplx<T> t1 = {a0.x + c1 * b1.x + c2 * b2.x, a0.y + c1 * b1.y + c2 * b2.y};
  cplx<T> t2 = {a0.x + c2 * b1.x + c1 * b2.x, a0.y + c2 * b1.y + c1 * b2.y};
  cplx<T> u1 = {s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y};
  cplx<T> u2 = {s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y};
  v[0] = {a0.x + b1.x + b2.x, a0.y + b1.y + b2.y};
  v[1] = {t1.x + u1.y, t1.y - u1.x};
  v[4] = {t1.x - u1.y, t1.y + u1.x};
  v[2] = {t2.x + u2.y, t2.y - u2.x};
  v[3] = {t2.x - u2.y, t2.y + u2.x};
}

// log2 of a power of two, -1 otherwise (block-uniform stage constants: the index split of a
// butterfly then costs a shift and a mask instead of an integer division)
__device__ __forceinline__ int log2_or_neg(int v) { return (v & (v - 1)) == 0 ? __builtin_ctz((unsigned)v) : -1; }

// a / n for a >= 0: a shift when n is a power of two (lg = log2_or_neg(n) >= 0)
__device__ __forceinline__ int fdiv(int a, int n, int lg) { return lg >= 0 ? (a >> lg) : a / n; }

template <typename T, int R>
__device__ __forceinline__ void stockham_stage(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst,
                                               int n, int p, const cplx<T> *__restrict__ tw, int nslot,
                                               int slot_stride, int tid, int nthr) {
  const int nr = n / R;
  const int lnr = log2_or_neg(nr), lp = log2_or_neg(p);
  for (int w = tid; w < nslot * nr; w += nthr) {
    const int f = lnr >= 0 ? (w >> lnr) : w / nr;
    const int k = w - f * nr;
    const cplx<T> *s = src + f * slot_stride;
    cplx<T> *d = dst + f * slot_stride;
    const int i = lp >= 0 ? (k & (p - 1)) : k % p;
    cplx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[k + r * nr];
    if (p > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[(p - 1) + (r - 1) * p + i]);
    }
    if constexpr (R == 2) dft2(v);
    if constexpr (R == 3) dft3(v);
    if constexpr (R == 4) dft4(v);
    if constexpr (R == 5) dft5(v);
    const int j = (k - i) * R + i;
#pragma unroll
    for (int r = 0; r < R; ++r) d[j + r * p] = v[r];
  }
}

// Pointers typed as global memory: loads and stores through them compile to global instead of
// flat instructions. For device allocations only: the batched kernels read their plans from a
// device table, which leaves the compiler unsure where the plans' pointers point.
// (the host pass of hipcc keeps the plain pointer: it compiles these bodies but never runs them,
// and its address space 1 is not the device's global memory)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename X>
using gptr = __attribute__((address_space(1))) X *;
#else
template <typename X>
using gptr = X *;
#endif
// structs holding gptr members are filled on the host and read on the device: both passes must
// lay them out alike
static_assert(sizeof(gptr<float>) == sizeof(float *) && alignof(gptr<float>) == alignof(float *),
              "global pointers must have the size and alignment of plain pointers");
// X in a parameter that takes no part in template argument deduction (a global pointer then
// converts to the parameter's plain pointer type instead of failing the deduction)
template <typename X>
struct nd_ident {
  using type = X;
};
template <typename X>
using nd = typename nd_ident<X>::type;

template <typename X>
__host__ __device__ __forceinline__ gptr<X> global_ptr(X *p) {
  return (gptr<X>)p;
}

// Forward complex FFT of `nslot` length-n sequences stored slot_stride apart in LDS
// buffer a (b is a scratch buffer of the same size). Returns the buffer holding the
// result in natural order. All threads of the block must call it.
template <typename T>
__device__ __forceinline__ cplx<T> *fft_lds(cplx<T> *a, cplx<T> *b, const FftPlan &pl, const cplx<T> *__restrict__ tw,
                            int nslot, int slot_stride) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  cplx<T> *src = a, *dst = b;
  int p = 1;
  for (int s = 0; s < pl.nrad; ++s) {
    const int R = pl.radix[s];
    switch (R) {
      case 4: stockham_stage<T, 4>(src, dst, pl.n, p, tw, nslot, slot_stride, tid, nthr); break;
      case 2: stockham_stage<T, 2>(src, dst, pl.n, p, tw, nslot, slot_stride, tid, nthr); break;
      case 3: stockham_stage<T, 3>(src, dst, pl.n, p, tw, nslot, slot_stride, tid, nthr); break;
      default: stockham_stage<T, 5>(src, dst, pl.n, p, tw, nslot, slot_stride, tid, nthr); break;
    }
    __syncthreads();
    cplx<T> *t = src;
    src = dst;
    dst = t;
    p *= R;
  }
  return src;
}

// Power-of-two lengths with the stage structure fixed at compile time: radix-4 stages of span
// 1, 4, 16, ... and, for odd log2 n, a last radix-2 stage (the radix order factor_radices
// gives, so the twiddle table and the results are those of fft_lds). Every index split is a
// shift or a mask and every loop bound a constant, which leaves each stage its loads, the
// butterfly and its stores.
template <int V>
struct Log2 { static constexpr int value = 1 + Log2<V / 2>::value; };
template <>
struct Log2<1> { static constexpr int value = 0; };

template <typename T, int N, int R, int P>
__device__ __forceinline__ void pow2_stage(const cplx<T> *__restrict__ src, cplx<T> *__restrict__ dst,
                                           const cplx<T> *__restrict__ tw, int nslot, int slot_stride, int tid,
                                           int nthr) {
  constexpr int NR = N / R, LNR = Log2<NR>::value, LR = Log2<R>::value;
  for (int w = tid; w < (nslot << LNR); w += nthr) {
    const int f = w >> LNR, k = w & (NR - 1);
    const cplx<T> *s = src + f * slot_stride;
    cplx<T> *d = dst + f * slot_stride;
    const int i = k & (P - 1);
    cplx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[k + r * NR];
    if constexpr (P > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[(P - 1) + (r - 1) * P + i]);
    }
    if constexpr (R == 2) dft2(v);
    if constexpr (R == 4) dft4(v);
    const int j = ((k - i) << LR) + i;
#pragma unroll
    for (int r = 0; r < R; ++r) d[j + r * P] = v[r];
  }
}

template <typename T, int N, int S>
__device__ __forceinline__ cplx<T> *pow2_stages(cplx<T> *src, cplx<T> *dst, const cplx<T> *__restrict__ tw, int nslot,
                                                int slot_stride, int tid, int nthr) {
  constexpr int LG = Log2<N>::value;
  if constexpr (2 * S + 2 <= LG) {
    pow2_stage<T, N, 4, (1 << (2 * S))>(src, dst, tw, nslot, slot_stride, tid, nthr);
    __syncthreads();
    return pow2_stages<T, N, S + 1>(dst, src, tw, nslot, slot_stride, tid, nthr);
  } else if constexpr (2 * S + 1 == LG) {
    pow2_stage<T, N, 2, (1 << (2 * S))>(src, dst, tw, nslot, slot_stride, tid, nthr);
    __syncthreads();
    return dst;
  } else {
    return src;
  }
}

// Forward transform of nslot sequences, as fft_lds: the power-of-two lengths 64 .. 2048 take
// their compile-time stage sequence, other lengths the runtime plan.
template <typename T>
__device__ __forceinline__ cplx<T> *fft_run(cplx<T> *a, cplx<T> *b, const FftPlan &pl, const cplx<T> *__restrict__ tw, int nslot,
                            int slot_stride) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  switch (pl.n) {
    case 64: return pow2_stages<T, 64, 0>(a, b, tw, nslot, slot_stride, tid, nthr);
    case 128: return pow2_stages<T, 128, 0>(a, b, tw, nslot, slot_stride, tid, nthr);
    case 256: return pow2_stages<T, 256, 0>(a, b, tw, nslot, slot_stride, tid, nthr);
    case 512: return pow2_stages<T, 512, 0>(a, b, tw, nslot, slot_stride, tid, nthr);
    case 1024: return pow2_stages<T, 1024, 0>(a, b, tw, nslot, slot_stride, tid, nthr);
    case 2048: return pow2_stages<T, 2048, 0>(a, b, tw, nslot, slot_stride, tid, nthr);
    default: return fft_lds(a, b, pl, tw, nslot, slot_stride);
  }
}

// ------------------------------------------------------------------ real transforms of pairs
// Two real sequences xa, xb of length n are packed into one complex sequence
// z = xa + i xb; one complex FFT gives both spectra (Hermitian split).
//
// DCT-II (no factor 2):  X_k = sum_n x_n cos(pi k (2n+1) / 2n)
//   Makhoul: v_m = x_{2m}, v_{n-1-m} = x_{2m+1}; V = FFT(v); X_k = Re(e^{-i pi k / 2n} V_k).
// DHT:  H_k = sum_n x_n cas(2 pi k n / n) = Re F_k - Im F_k.
// Inverses: DCT-III scaled so that idct(dct(x)) = x; DHT is its own inverse up to 1/n.
//
// Index of the reordered position m of the natural index q (Makhoul order).
__device__ __forceinline__ int makhoul_pos(int q, int n) { return (q & 1) ? (n - 1 - (q >> 1)) : (q >> 1); }

// After fft_lds of packed (xa + i xb): spectra of the two real inputs at k.
template <typename T>
__device__ __forceinline__ void split_pair(const cplx<T> *Z, int n, int k, cplx<T> &va, cplx<T> &vb) {
  const cplx<T> z = Z[k];
  const cplx<T> zc = Z[k == 0 ? 0 : n - k];
  va = {(T)0.5 * (z.x + zc.x), (T)0.5 * (z.y - zc.y)};
  vb = {(T)0.5 * (z.y + zc.y), (T)0.5 * (zc.x - z.x)};
}

// ------------------------------------------------------------------ reductions
// Block-wide sum (deterministic order for a fixed block size); every thread gets the result.
template <typename T>
__device__ T block_sum(T v, T *lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  __syncthreads();
  lds[tid] = v;
  __syncthreads();
  for (int s = 1; s < nthr; s <<= 1) {
    // pairwise tree over the power-of-two block
    if ((tid % (2 * s)) == 0 && tid + s < nthr) lds[tid] = lds[tid] + lds[tid + s];
    __syncthreads();
  }
  T r = lds[0];
  __syncthreads();
  return r;
}

template <typename T>
__device__ T block_max(T v, T *lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthr = blockDim.x * blockDim.y;
  __syncthreads();
  lds[tid] = v;
  __syncthreads();
  for (int s = nthr / 2; s > 0; s >>= 1) {
    if (tid < s) lds[tid] = lds[tid] > lds[tid + s] ? lds[tid] : lds[tid + s];
    __syncthreads();
  }
  T r = lds[0];
  __syncthreads();
  return r;
}

// Wavefront (64-lane) reductions through DPP lane moves (register-to-register, no LDS round
// trips): pairs, quads, half rows and rows by quad permutes and mirrors, then the row
// broadcasts 15 and 31 carry rows 0 and 2 into rows 1 and 3 and row 1 into rows 2 and 3;
// lane 63 ends with the total, which is returned to every lane. The combination order is
// fixed, so results are deterministic. Call from uniform control flow (all 64 lanes active).
namespace dpp {
constexpr int kPairs = 0xb1;      // quad_perm [1, 0, 3, 2]
constexpr int kQuads = 0x4e;      // quad_perm [2, 3, 0, 1]
constexpr int kRowMirror = 0x140;
constexpr int kHalfMirror = 0x141;
constexpr int kBcast15 = 0x142;
constexpr int kBcast31 = 0x143;
// v moved by CTRL; lanes outside ROWS (a 4-bit row mask) get `old`
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float mov(float v, float old) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double mov(double v, double old) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)(o & 0xffffffffLL), (int)(unsigned)(b & 0xffffffffLL), CTRL,
                                             ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ __forceinline__ double lane63(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
}  // namespace dpp

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v = v + dpp::mov<dpp::kPairs>(v, (T)0);
  v = v + dpp::mov<dpp::kQuads>(v, (T)0);
  v = v + dpp::mov<dpp::kHalfMirror>(v, (T)0);
  v = v + dpp::mov<dpp::kRowMirror>(v, (T)0);
  v = v + dpp::mov<dpp::kBcast15, 0xa>(v, (T)0);
  v = v + dpp::mov<dpp::kBcast31, 0xc>(v, (T)0);
  return dpp::lane63(v);
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  T w = dpp::mov<dpp::kPairs>(v, v);
  v = w > v ? w : v;
  w = dpp::mov<dpp::kQuads>(v, v);
  v = w > v ? w : v;
  w = dpp::mov<dpp::kHalfMirror>(v, v);
  v = w > v ? w : v;
  w = dpp::mov<dpp::kRowMirror>(v, v);
  v = w > v ? w : v;
  w = dpp::mov<dpp::kBcast15, 0xa>(v, v);
  v = w > v ? w : v;
  w = dpp::mov<dpp::kBcast31, 0xc>(v, v);
  v = w > v ? w : v;
  return dpp::lane63(v);
}
// Block sum of `n` values per thread (blocks of 64 to 1024 threads): wave sums, then the wave
// partials added in order. `lds` holds (waves) * n doubles. Thread 0 gets the result.
template <int N>
__device__ __forceinline__ void block_sum4(double (&v)[N], double *lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int w = tid >> 6, lane = tid & 63;
  const int nw = (int)(blockDim.x * blockDim.y) >> 6;
#pragma unroll
  for (int q = 0; q < N; ++q) v[q] = wave_sum(v[q]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) lds[w * N + q] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < N; ++q) {
      double r = lds[q];
      for (int w2 = 1; w2 < nw; ++w2) r += lds[w2 * N + q];   // waves in order
      v[q] = r;
    }
  }
}

// Max of non-negative values as an unsigned-integer atomic (order independent).
__device__ __forceinline__ void atomic_max_nonneg(float *addr, float v) {
  atomicMax(reinterpret_cast<unsigned int *>(addr), __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

template <typename T>
__device__ __forceinline__ T load_relaxed(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace aero
