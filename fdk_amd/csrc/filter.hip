// Cosine pre-weighting and ramp/apodised row filtering on gfx950.
//
// Reference behaviour (peng-ac/fdk):
//   FilterCosCorrect / kernelCorrect    Filterlib/FilterLib.h:8-13,
//                                       FilterlibCufft/cudaFftFilter.cu:270-296
//   FilterImageRow (mean4 pad, IPP FFT) Filterlib/FilterLib.h:15-58
//   kernelPad/DotScale/Unpad + cuFFT    FilterlibCufft/cudaFftFilter.cu:14-61, 298-370
//   FilterProjImage                     FilterlibCufft/cudaFftFilter.cu:394-452
//
// Each detector row of nu samples is centred in N = 2^m samples at offset
// (N - nu) / 2, the pad takes the mean of the first / last 3 or 4 samples (or
// zero), the row is multiplied in the frequency domain by the real, even
// filter H and the centre nu samples of the inverse transform (1/N) are kept.
//
// MI355X design: one workgroup filters two rows at once as one complex
// transform (row a in the real part, row b in the imaginary part).  Because H
// is real and even, H * DFT(a + i b) inverts to filt(a) + i filt(b) exactly, so
// no real-to-complex unpacking is needed.  For N = 2^9 .. 2^13 the transform is
// a Stockham plan of radix-16 (and radix-2/4/8) passes with compile-time
// strides, the inverse running the forward radices in reverse order: each
// thread keeps the points it loaded for the first pass, multiplies by H in
// registers between the last forward and first inverse pass, and stores the
// points its last inverse pass produced, so only the inner passes go through
// LDS (N complex floats, 64 KiB at N = 8192).  Other lengths run an in-place
// radix-2^2 FFT entered in bit-reversed order, the filter multiply doubling as
// the inverse's bit-reversal.  Loads apply the cos table, stores crop,
// optionally transpose (v fastest, what the backprojector reads) and convert to
// the stored dtype.  Twiddles are an FP64-built table.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace {

struct FilterArgs {
  const float* src;       // rows of nu samples
  long long src_view;     // elements between views
  int src_pitch;          // elements between rows
  float* dst;             // FP32 output (dtype handled by a second pass)
  long long dst_view;
  int dst_pitch;          // row pitch when not transposed
  int transpose;          // 1: dst[view][u][row] with row fastest, pitch = nrows
  const float* cos_table; // nrows x nu (pitch nu) or null
  const float* H;         // N
  const float2* tw;       // N twiddles e^{-2 pi i k / N}
  int nu, nrows, N, logN, pad;
  unsigned* minmax;       // per view {min, max} as ordered uints, or null
  // Output mode: kStoreF32 writes dst as floats; kStoreNone writes nothing
  // (the min/max pass of a two-pass 16-bit filter); kStoreF16 / kStoreBF16
  // write (p - min) / (max - min) of view {min, max} words qmm as 16-bit.
  int qmode;
  const unsigned* qmm;
};

constexpr int kStoreF32 = 0, kStoreNone = 1, kStoreF16 = 2, kStoreBF16 = 3;

__device__ inline unsigned brev_n(unsigned k, int logN) { return __brev(k) >> (32 - logN); }

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-place FFT over buf (bit-reversed input -> natural output); sign = -1
// forward, +1 inverse (conjugate twiddles).  Radix-2 first stage when log2 N is
// odd, then fused radix-2^2 stages.
__device__ void fft_lds(float2* buf, const float2* __restrict__ tw, int N, int logN, float sign) {
  const int tid = threadIdx.x, nt = blockDim.x;
  int L = 1;
  if (logN & 1) {
    for (int q = tid; q < N / 2; q += nt) {
      const float2 x0 = buf[2 * q], x1 = buf[2 * q + 1];
      buf[2 * q] = make_float2(x0.x + x1.x, x0.y + x1.y);
      buf[2 * q + 1] = make_float2(x0.x - x1.x, x0.y - x1.y);
    }
    L = 2;
    __syncthreads();
  }
  for (; L < N; L *= 4) {
    const int s1 = N / (2 * L), s2 = N / (4 * L);
    for (int q = tid; q < N / 4; q += nt) {
      const int j = q % L;
      const int g = (q / L) * 4 * L + j;
      float2 w1 = tw[j * s1], w2 = tw[j * s2];
      w1.y *= -sign;  // table holds e^{-i...}; inverse conjugates
      w2.y *= -sign;
      const float2 x0 = buf[g], x1 = cmul(buf[g + L], w1);
      const float2 x2 = buf[g + 2 * L], x3 = cmul(buf[g + 3 * L], w1);
      const float2 y0 = make_float2(x0.x + x1.x, x0.y + x1.y);
      const float2 y1 = make_float2(x0.x - x1.x, x0.y - x1.y);
      const float2 y2 = make_float2(x2.x + x3.x, x2.y + x3.y);
      const float2 y3 = make_float2(x2.x - x3.x, x2.y - x3.y);
      const float2 t2 = cmul(y2, w2);
      float2 t3 = cmul(y3, w2);
      // forward: multiply by -i; inverse: by +i
      t3 = sign < 0 ? make_float2(t3.y, -t3.x) : make_float2(-t3.y, t3.x);
      buf[g] = make_float2(y0.x + t2.x, y0.y + t2.y);
      buf[g + 2 * L] = make_float2(y0.x - t2.x, y0.y - t2.y);
      buf[g + L] = make_float2(y1.x + t3.x, y1.y + t3.y);
      buf[g + 3 * L] = make_float2(y1.x - t3.x, y1.y - t3.y);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- Stockham FFT
// Self-sorting mixed-radix FFT in LDS (natural order in and out): each pass
// reads R values per butterfly (stride N/R), twiddles them, runs a radix-R DFT
// in registers and writes them back at stride Ns, the product of the radices
// before it.  Radix-16/8 passes cut the LDS round trips per transform to 3-4
// (against log4 N in-place passes with a bit-reversed entry), and the filter
// multiply rides on the first inverse pass's reads.

// v * (sign * i): sign -1 multiplies by -i, +1 by +i.
__device__ __forceinline__ float2 mul_si(float2 v, float sign) {
  return sign < 0 ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// X[k] = sum_n v[n] e^{sign 2 pi i n k / R}
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  const float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3, float sign) {
  const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3);
  const float2 a3 = mul_si(csub(v1, v3), sign);
  v0 = cadd(a0, a2);
  v2 = csub(a0, a2);
  v1 = cadd(a1, a3);
  v3 = csub(a1, a3);
}

// e^{sign 2 pi i p / 16}, p = 0..9
__device__ __forceinline__ float2 w16(int p, float sign) {
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508977f, c2 = 0.70710678118654752f;
  float2 w;
  switch (p) {
    case 0: w = make_float2(1.f, 0.f); break;
    case 1: w = make_float2(c1, s1); break;
    case 2: w = make_float2(c2, c2); break;
    case 3: w = make_float2(s1, c1); break;
    case 4: w = make_float2(0.f, 1.f); break;
    case 5: w = make_float2(-s1, c1); break;
    case 6: w = make_float2(-c2, c2); break;
    case 7: w = make_float2(-c1, s1); break;
    case 8: w = make_float2(-1.f, 0.f); break;
    default: w = make_float2(-c1, -s1); break;  // 9
  }
  w.y *= sign;
  return w;
}

template <int R>
__device__ __forceinline__ void dft(float2 (&v)[R], float sign);

template <>
__device__ __forceinline__ void dft<2>(float2 (&v)[2], float) {
  dft2(v[0], v[1]);
}

template <>
__device__ __forceinline__ void dft<4>(float2 (&v)[4], float sign) {
  dft4(v[0], v[1], v[2], v[3], sign);
}

template <>
__device__ __forceinline__ void dft<8>(float2 (&v)[8], float sign) {
  // even / odd halves, then X[k] = E[k] + W8^k O[k], X[k+4] = E[k] - W8^k O[k]
  dft4(v[0], v[2], v[4], v[6], sign);
  dft4(v[1], v[3], v[5], v[7], sign);
  float2 o[4] = {v[1], cmul(v[3], w16(2, sign)), mul_si(v[5], sign), cmul(v[7], w16(6, sign))};
  const float2 e[4] = {v[0], v[2], v[4], v[6]};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 4] = csub(e[k], o[k]);
  }
}

template <>
__device__ __forceinline__ void dft<16>(float2 (&v)[16], float sign) {
  // F_e = DFT4(v[e], v[e+4], v[e+8], v[e+12]); X[k'+4q] = DFT4_e(W16^{e k'} F_e[k'])[q]
#pragma unroll
  for (int e = 0; e < 4; ++e) dft4(v[e], v[e + 4], v[e + 8], v[e + 12], sign);
  // v[e + 4k'] now holds F_e[k']
#pragma unroll
  for (int e = 1; e < 4; ++e)
#pragma unroll
    for (int k = 1; k < 4; ++k) v[e + 4 * k] = cmul(v[e + 4 * k], w16(e * k, sign));
  float2 x[16];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    float2 g0 = v[4 * k], g1 = v[1 + 4 * k], g2 = v[2 + 4 * k], g3 = v[3 + 4 * k];
    dft4(g0, g1, g2, g3, sign);
    x[k] = g0;
    x[k + 4] = g1;
    x[k + 8] = g2;
    x[k + 12] = g3;
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = x[k];
}

// LDS slot of point i: with PAD, one spare slot after every 16 points, so the
// stride-16 writes of the first pass (and the 2 KiB-apart runs of the second)
// spread over the banks instead of landing on one.
template <bool PAD>
__device__ __forceinline__ int sk_slot(int i) {
  return PAD ? i + (i >> 4) : i;
}

// Slot distance of a run of d points, d a multiple of 16: the pad term of
// sk_slot(i + d) then carries no bit from i, so sk_slot(i + d) = sk_slot(i) + sk_span(d).
template <bool PAD>
constexpr int sk_span(int d) {
  return PAD ? d + d / 16 : d;
}

// Stockham passes over N points in buf with BPT butterflies per thread: a pass
// of radix R reads the R values of butterfly j at stride NB = N / R, twiddles
// them by e^{-2 pi i (j mod NS) r / (NS R)} (NS = the product of the radices
// before it), runs a radix-R DFT in registers and writes them at stride NS.
// N, NS and the radix are compile-time, so every LDS access is one per-thread
// base plus an immediate offset.  Butterfly q of a thread is j = tid + 256 q;
// a pass with fewer than 256 butterflies leaves the upper threads idle.
template <int N, int R, int BPT>
__device__ __forceinline__ bool sk_active(int q) {
  constexpr int NB = N / R;
  return NB >= 256 * (q + 1) || (int)threadIdx.x + q * 256 < NB;
}

template <int R, int BPT, bool PAD, int N>
__device__ __forceinline__ void sk_read(const float2* buf, float2 (&v)[BPT][R]) {
  constexpr int NB = N / R;
  static_assert(NB % 16 == 0, "a read stride keeps the pad term separable");
  constexpr int kReadSpan = sk_span<PAD>(NB);
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    if (sk_active<N, R, BPT>(q)) {
      const float2* src = buf + sk_slot<PAD>(threadIdx.x + q * 256);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[r * kReadSpan];
    }
  }
}

template <int R, int BPT, int N, int NS>
__device__ __forceinline__ void sk_dft(float2 (&v)[BPT][R], float sign, const float2* __restrict__ tw) {
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    if (sk_active<N, R, BPT>(q)) {
      if constexpr (NS > 1) {
        // twiddle r of this butterfly: tw[step r], e^{-2 pi i step r / N}
        const int j = threadIdx.x + q * 256;
        const unsigned step8 = (unsigned)((j & (NS - 1)) * (N / (NS * R))) * 8u;
        const __amdgpu_buffer_rsrc_t twr =
            __builtin_amdgcn_make_buffer_rsrc((void*)tw, 0, N * 8, fdk::kRawRsrcWord3);
        unsigned off = 0;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          off += step8;
          float2 w = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(twr, off, 0, 0));
          w.y *= -sign;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft<R>(v[q], sign);
    }
  }
}

template <int R, int BPT, bool PAD, int N, int NS>
__device__ __forceinline__ void sk_write(float2* buf, const float2 (&v)[BPT][R]) {
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    if (sk_active<N, R, BPT>(q)) {
      const int j = threadIdx.x + q * 256;
      const int k = j & (NS - 1);
      const int base = (j - k) * R + k;
      if constexpr (NS == 1) {
        // base = j R: its R outputs stay inside one 16-point run (R <= 16, base % 16 + R <= 16)
        float2* dst = buf + sk_slot<PAD>(base);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r] = v[q][r];
      } else if constexpr (NS % 16 == 0) {
        float2* dst = buf + sk_slot<PAD>(base);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * sk_span<PAD>(NS)] = v[q][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) buf[sk_slot<PAD>(base + r * NS)] = v[q][r];
      }
    }
  }
}

// One pass LDS -> LDS.
template <int R, int BPT, bool PAD, int N, int NS>
__device__ __forceinline__ void sk_pass(float2* buf, float sign, const float2* __restrict__ tw) {
  float2 v[BPT][R];
  sk_read<R, BPT, PAD, N>(buf, v);
  sk_dft<R, BPT, N, NS>(v, sign, tw);
  __syncthreads();  // every thread has read its inputs
  sk_write<R, BPT, PAD, N, NS>(buf, v);
  __syncthreads();
}

template <int N, int R>
constexpr int sk_bpt() {
  return N / R >= 256 ? N / R / 256 : 1;
}

// The whole filter on a thread's N/256 points, x[i] = point tid + 256 i on the
// way in and out: forward passes of radices R1, R2, R3 (, R4), the filter
// multiply, and inverse passes in the reverse radix order.  With that order
// the points a thread loads are its first pass's inputs, its last forward
// pass's outputs are its first inverse pass's inputs (the multiply runs in
// registers), and its last inverse pass's outputs are the points it stores, so
// of the 2 x passes LDS round trips only 2 x (passes - 1) remain.  R4 = 1:
// three passes per direction.
template <int R1, int R2, int R3, int R4, bool PAD>
__device__ __forceinline__ void sk_filter(float2* buf, const float2* __restrict__ tw,
                                          const float* __restrict__ H,
                                          float2 (&x)[R1 * R2 * R3 * R4 / 256]) {
  constexpr int N = R1 * R2 * R3 * R4;
  constexpr int RL = R4 > 1 ? R4 : R3;      // last forward radix = first inverse radix
  constexpr int NSL = N / RL;               // its stride in both directions
  constexpr int B1 = sk_bpt<N, R1>(), BL = sk_bpt<N, RL>();
  static_assert(B1 * R1 * 256 == N, "the first pass's inputs are the loaded points");
  {
    float2 v[B1][R1];
#pragma unroll
    for (int q = 0; q < B1; ++q)
#pragma unroll
      for (int r = 0; r < R1; ++r) v[q][r] = x[q + B1 * r];
    sk_dft<R1, B1, N, 1>(v, -1.f, tw);
    sk_write<R1, B1, PAD, N, 1>(buf, v);
    __syncthreads();
  }
  sk_pass<R2, sk_bpt<N, R2>(), PAD, N, R1>(buf, -1.f, tw);
  if constexpr (R4 > 1) sk_pass<R3, sk_bpt<N, R3>(), PAD, N, R1 * R2>(buf, -1.f, tw);
  {
    float2 v[BL][RL];
    sk_read<RL, BL, PAD, N>(buf, v);
    sk_dft<RL, BL, N, NSL>(v, -1.f, tw);
#pragma unroll
    for (int q = 0; q < BL; ++q) {
      if (sk_active<N, RL, BL>(q)) {
        const int j = threadIdx.x + q * 256;
#pragma unroll
        for (int r = 0; r < RL; ++r) {
          const float h = H[j + r * NSL];
          v[q][r] = make_float2(v[q][r].x * h, v[q][r].y * h);
        }
      }
    }
    sk_dft<RL, BL, N, 1>(v, 1.f, tw);
    __syncthreads();  // every thread has read the last forward pass's inputs
    sk_write<RL, BL, PAD, N, 1>(buf, v);
    __syncthreads();
  }
  if constexpr (R4 > 1) {
    sk_pass<R3, sk_bpt<N, R3>(), PAD, N, R4>(buf, 1.f, tw);
    sk_pass<R2, sk_bpt<N, R2>(), PAD, N, R4 * R3>(buf, 1.f, tw);
  } else {
    sk_pass<R2, sk_bpt<N, R2>(), PAD, N, R3>(buf, 1.f, tw);
  }
  {
    float2 v[B1][R1];
    sk_read<R1, B1, PAD, N>(buf, v);
    sk_dft<R1, B1, N, N / R1>(v, 1.f, tw);
#pragma unroll
    for (int q = 0; q < B1; ++q)
#pragma unroll
      for (int r = 0; r < R1; ++r) x[q + B1 * r] = v[q][r];
  }
}

// Radix plans for N = 2^M, 9 <= M <= 13 (256 threads): the first radix is
// N / 256 (16 at M = 13, two butterflies per thread), so every thread starts
// with the points it loaded.  The LDS is padded for M <= 12 (N + N/16 slots; at
// M = 13 that would pass 64 KiB).
template <int M>
constexpr bool sk_pad() {
  return M <= 12;
}

template <int M>
__device__ __forceinline__ void filter_sk(float2* buf, const float2* tw, const float* H,
                                          float2 (&x)[(1 << M) / 256]) {
  constexpr bool P = sk_pad<M>();
  if constexpr (M == 9) sk_filter<2, 16, 16, 1, P>(buf, tw, H, x);
  else if constexpr (M == 10) sk_filter<4, 16, 16, 1, P>(buf, tw, H, x);
  else if constexpr (M == 11) sk_filter<8, 16, 16, 1, P>(buf, tw, H, x);
  else if constexpr (M == 12) sk_filter<16, 16, 16, 1, P>(buf, tw, H, x);
  else {
    static_assert(M == 13, "plans cover 2^9 .. 2^13");
    sk_filter<16, 16, 16, 2, P>(buf, tw, H, x);
  }
}

__device__ inline unsigned f2ord(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float ord2f(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// q = (p - min) * (1 / (max - min)) as FP16 / BF16 bits (round to nearest even).
template <int Q>
__device__ inline unsigned short quantize16(float p, float lo, float inv) {
  const float q = (p - lo) * inv;
  if (Q == kStoreF16) return __half_as_ushort(__float2half_rn(q));
  const unsigned b = __float_as_uint(q);
  return (unsigned short)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

// Crop, scale by 1/N and store one output sample pair (rows ra, ra+1 of the
// view) of the 16-bit modes.
template <int Q>
__device__ inline void store16(const FilterArgs& a, int view, int u, int ra, bool has_b, float fa,
                               float fb, float lo, float inv) {
  unsigned short* d = static_cast<unsigned short*>(static_cast<void*>(a.dst)) + view * a.dst_view;
  const unsigned short qa = quantize16<Q>(fa, lo, inv);
  const unsigned short qb = has_b ? quantize16<Q>(fb, lo, inv) : 0;
  if (a.transpose) {
    unsigned short* o = d + (long long)u * a.nrows + ra;
    if (has_b && (a.nrows & 1) == 0) {
      *reinterpret_cast<unsigned*>(o) = (unsigned)qa | ((unsigned)qb << 16);  // 4-byte aligned
    } else {
      o[0] = qa;
      if (has_b) o[1] = qb;
    }
  } else {
    d[(long long)ra * a.dst_pitch + u] = qa;
    if (has_b) d[(long long)(ra + 1) * a.dst_pitch + u] = qb;
  }
}

// Output side of the filter kernel: the 16-bit quantisation constants of the
// view and the running min/max of the values stored.
struct StoreCtx {
  float qlo, qinv, lo, hi;
};

__device__ __forceinline__ StoreCtx store_ctx(const FilterArgs& a, int view) {
  StoreCtx c{0.f, 0.f, INFINITY, -INFINITY};
  if (a.qmode == kStoreF16 || a.qmode == kStoreBF16) {
    c.qlo = ord2f(a.qmm[2 * view]);
    const float range = ord2f(a.qmm[2 * view + 1]) - c.qlo;
    c.qinv = range > 0.f ? 1.0f / range : 0.f;
  }
  return c;
}

// Per-view min/max: a wave-level reduction, then one atomic per wave.
__device__ __forceinline__ void store_minmax(const FilterArgs& a, StoreCtx& c, int view) {
  if (!a.minmax) return;
  float lo = c.lo, hi = c.hi;
  for (int o = 32; o > 0; o >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, o));
    hi = fmaxf(hi, __shfl_xor(hi, o));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&a.minmax[2 * view], f2ord(lo));
    atomicMax(&a.minmax[2 * view + 1], f2ord(hi));
  }
}

// Stores output sample u of rows ra (fa) and ra+1 (fb) in store mode Q's form
// and layout, and with MM tracks their min/max.
template <int Q, bool MM>
__device__ __forceinline__ void store_pair_q(const FilterArgs& a, StoreCtx& c, int view, int u, int ra,
                                             bool has_b, float fa, float fb) {
  if constexpr (Q == kStoreF16 || Q == kStoreBF16) {
    store16<Q>(a, view, u, ra, has_b, fa, fb, c.qlo, c.qinv);
  } else if constexpr (Q == kStoreF32) {
    float* d = a.dst + view * a.dst_view;
    if (a.transpose) {
      float* o = d + (long long)u * a.nrows + ra;
      if (has_b && (a.nrows & 1) == 0) {
        *reinterpret_cast<float2*>(o) = make_float2(fa, fb);  // 8-byte aligned: u*nrows + ra even
      } else {
        o[0] = fa;
        if (has_b) o[1] = fb;
      }
    } else {
      d[(long long)ra * a.dst_pitch + u] = fa;
      if (has_b) d[(long long)(ra + 1) * a.dst_pitch + u] = fb;
    }
  }
  if constexpr (MM) {
    c.lo = fminf(c.lo, has_b ? fminf(fa, fb) : fa);
    c.hi = fmaxf(c.hi, has_b ? fmaxf(fa, fb) : fa);
  }
}

// The same with the mode and min/max tracking chosen at run time.
__device__ __forceinline__ void store_pair(const FilterArgs& a, StoreCtx& c, int view, int u, int ra,
                                           bool has_b, float fa, float fb) {
  if (a.qmode == kStoreNone)
    store_pair_q<kStoreNone, true>(a, c, view, u, ra, has_b, fa, fb);
  else if (a.qmode == kStoreF16)
    store_pair_q<kStoreF16, true>(a, c, view, u, ra, has_b, fa, fb);
  else if (a.qmode == kStoreBF16)
    store_pair_q<kStoreBF16, true>(a, c, view, u, ra, has_b, fa, fb);
  else
    store_pair_q<kStoreF32, true>(a, c, view, u, ra, has_b, fa, fb);
}

// A thread's E output points x[i] (point tid + 256 i) cropped, scaled and
// stored: one loop per store mode, so no per-point mode branches remain.
template <int Q, bool MM, int E>
__device__ __forceinline__ void store_points(const FilterArgs& a, int view, int ra, bool has_b,
                                             int off, const float2 (&x)[E]) {
  const float inv = 1.0f / (float)a.N;
  StoreCtx sc = store_ctx(a, view);
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int u = (int)threadIdx.x + 256 * i - off;
    if (u >= 0 && u < a.nu) store_pair_q<Q, MM>(a, sc, view, u, ra, has_b, x[i].x * inv, x[i].y * inv);
  }
  if constexpr (MM) store_minmax(a, sc, view);
}


// M = log2 N selects the Stockham plan (9..13); M = 0 runs the in-place
// bit-reversed radix-2^2 FFT for the other lengths.
template <int M>
__global__ __launch_bounds__(256) void filter_rows_kernel(FilterArgs a) {
  extern __shared__ float2 buf[];
  const int view = blockIdx.y;
  // Row pairs remapped so that each XCD (hardware block b -> XCD b % 8) filters
  // a contiguous run of rows: the 8-byte transposed stores of neighbouring
  // pairs then meet in one L2 and leave it as whole lines.
  const unsigned gx = gridDim.x;
  const unsigned bx = (gx % 8u == 0u) ? (blockIdx.x % 8u) * (gx / 8u) + blockIdx.x / 8u : blockIdx.x;
  const int ra = 2 * (int)bx, rb = ra + 1;
  const bool has_b = rb < a.nrows;
  const float* sa = a.src + view * a.src_view + (long long)ra * a.src_pitch;
  const float* sb = sa + a.src_pitch;
  const float* ca = a.cos_table ? a.cos_table + (long long)ra * a.nu : nullptr;
  const float* cb = ca ? ca + a.nu : nullptr;
  const int off = (a.N - a.nu) >> 1;
  const int nu = a.nu;

  // Edge means of the cos-weighted rows (FilterLib.h:37-46 forms mean4 in
  // float; the mean3 variants of type.h:4202-4205 in double).
  float la = 0.f, lb = 0.f, rra = 0.f, rrb = 0.f;
  if (a.pad == FDK_PAD_MEAN4 || a.pad == FDK_PAD_MEAN3) {
    const int k = a.pad;
    float xa[8], xb[8];
    for (int q = 0; q < k; ++q) {
      const int il = q, ir = nu - 1 - q;
      xa[q] = sa[il] * (ca ? ca[il] : 1.f);
      xa[4 + q] = sa[ir] * (ca ? ca[ir] : 1.f);
      xb[q] = has_b ? sb[il] * (cb ? cb[il] : 1.f) : 0.f;
      xb[4 + q] = has_b ? sb[ir] * (cb ? cb[ir] : 1.f) : 0.f;
    }
    if (k == 4) {
      la = (xa[0] + xa[1] + xa[2] + xa[3]) / 4;
      rra = (xa[4] + xa[5] + xa[6] + xa[7]) / 4;
      lb = (xb[0] + xb[1] + xb[2] + xb[3]) / 4;
      rrb = (xb[4] + xb[5] + xb[6] + xb[7]) / 4;
    } else {
      la = (float)(((double)xa[0] + xa[1] + xa[2]) / 3);
      rra = (float)(((double)xa[4] + xa[5] + xa[6]) / 3);
      lb = (float)(((double)xb[0] + xb[1] + xb[2]) / 3);
      rrb = (float)(((double)xb[4] + xb[5] + xb[6]) / 3);
    }
  }
  if constexpr (M != 0) {
    // Every load of the thread's E points issued before any is used (clamped
    // addresses, pads selected afterwards), so one memory latency covers them.
    // Row and cos-table loads go through buffer resources (one 32-bit offset
    // per point serves all four); a missing cos table is a zero-size resource
    // whose loads return 0 and are replaced by 1.
    constexpr int E = (1 << M) / 256;
    const float* sbb = has_b ? sb : sa;
    const float* cbb = has_b ? cb : ca;
    const int row_bytes = nu * 4;
    const __amdgpu_buffer_rsrc_t rxa = __builtin_amdgcn_make_buffer_rsrc((void*)sa, 0, row_bytes, fdk::kRawRsrcWord3);
    const __amdgpu_buffer_rsrc_t rxb = __builtin_amdgcn_make_buffer_rsrc((void*)sbb, 0, row_bytes, fdk::kRawRsrcWord3);
    const __amdgpu_buffer_rsrc_t rwa =
        __builtin_amdgcn_make_buffer_rsrc((void*)ca, 0, ca ? row_bytes : 0, fdk::kRawRsrcWord3);
    const __amdgpu_buffer_rsrc_t rwb =
        __builtin_amdgcn_make_buffer_rsrc((void*)cbb, 0, cbb ? row_bytes : 0, fdk::kRawRsrcWord3);
    float xa[E], xb[E], wa[E], wb[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const unsigned po = (unsigned)min(max((int)threadIdx.x + 256 * i - off, 0), nu - 1) * 4u;
      xa[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rxa, po, 0, 0));
      xb[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rxb, po, 0, 0));
      wa[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rwa, po, 0, 0));
      wb[i] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rwb, po, 0, 0));
    }
    if (!ca) {
#pragma unroll
      for (int i = 0; i < E; ++i) wa[i] = wb[i] = 1.f;
    }
    float2 x[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int n = threadIdx.x + 256 * i;
      const int p = n - off;
      float va = xa[i] * wa[i], vb = has_b ? xb[i] * wb[i] : 0.f;
      if (p < 0) {
        va = la;
        vb = lb;
      } else if (p >= nu) {
        va = rra;
        vb = rrb;
      }
      x[i] = make_float2(va, vb);
    }
    filter_sk<M>(buf, a.tw, a.H, x);
    // x[i] is now output point tid + 256 i: crop, scale and store it
    if (a.qmode == kStoreNone)
      store_points<kStoreNone, true>(a, view, ra, has_b, off, x);
    else if (a.qmode == kStoreF16)
      store_points<kStoreF16, false>(a, view, ra, has_b, off, x);
    else if (a.qmode == kStoreBF16)
      store_points<kStoreBF16, false>(a, view, ra, has_b, off, x);
    else if (a.minmax)
      store_points<kStoreF32, true>(a, view, ra, has_b, off, x);
    else
      store_points<kStoreF32, false>(a, view, ra, has_b, off, x);
    return;
  } else {
    for (int n = threadIdx.x; n < a.N; n += blockDim.x) {
      const int p = n - off;
      float va, vb;
      if (p < 0) { va = la; vb = lb; }
      else if (p >= nu) { va = rra; vb = rrb; }
      else {
        va = sa[p] * (ca ? ca[p] : 1.f);
        vb = has_b ? sb[p] * (cb ? cb[p] : 1.f) : 0.f;
      }
      buf[brev_n(n, a.logN)] = make_float2(va, vb);
    }
  }
  __syncthreads();
  {
    fft_lds(buf, a.tw, a.N, a.logN, -1.f);
    // Filter multiply fused with the bit-reversal permutation of the inverse pass.
    for (int k = threadIdx.x; k < a.N; k += blockDim.x) {
      const unsigned r = brev_n(k, a.logN);
      if ((unsigned)k < r) {
        const float2 xk = buf[k], xr = buf[r];
        const float hk = a.H[k], hr = a.H[r];
        buf[k] = make_float2(xr.x * hr, xr.y * hr);
        buf[r] = make_float2(xk.x * hk, xk.y * hk);
      } else if ((unsigned)k == r) {
        const float hk = a.H[k];
        buf[k] = make_float2(buf[k].x * hk, buf[k].y * hk);
      }
    }
    __syncthreads();
    fft_lds(buf, a.tw, a.N, a.logN, 1.f);
  }
  const float inv = 1.0f / (float)a.N;
  StoreCtx sc = store_ctx(a, view);
  for (int u = threadIdx.x; u < nu; u += blockDim.x) {
    const float2 c = buf[off + u];
    store_pair(a, sc, view, u, ra, has_b, c.x * inv, c.y * inv);
  }
  store_minmax(a, sc, view);
}

// FP32 staging -> FP16/BF16 with q = (p - min) / (max - min) per view, four
// samples per thread per step (16-byte loads, 8-byte stores) when the view
// size and both view strides are multiples of 4 (VEC), else one.
template <int DT, bool VEC>
__global__ __launch_bounds__(256) void quantize_kernel(const float* __restrict__ src, long long sview,
                                                       void* dst, long long dview, long long n,
                                                       const unsigned* minmax) {
  constexpr int Q = DT == FDK_F16 ? kStoreF16 : kStoreBF16;
  const int view = blockIdx.y;
  const float lo = ord2f(minmax[2 * view]), hi = ord2f(minmax[2 * view + 1]);
  const float range = hi - lo;
  const float inv = range > 0.f ? 1.0f / range : 0.f;
  const float* s = src + view * sview;
  unsigned short* d = static_cast<unsigned short*>(dst) + view * dview;
  const long long stride = (long long)gridDim.x * 256;
  if (VEC) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; 4 * i < n; i += stride) {
      const float4 p = reinterpret_cast<const float4*>(s)[i];
      const unsigned w0 = (unsigned)quantize16<Q>(p.x, lo, inv) | ((unsigned)quantize16<Q>(p.y, lo, inv) << 16);
      const unsigned w1 = (unsigned)quantize16<Q>(p.z, lo, inv) | ((unsigned)quantize16<Q>(p.w, lo, inv) << 16);
      reinterpret_cast<uint2*>(d)[i] = make_uint2(w0, w1);
    }
  } else {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += stride)
      d[i] = quantize16<Q>(s[i], lo, inv);
  }
}

__global__ void minmax_init_kernel(unsigned* mm, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    mm[2 * i] = 0xffffffffu;
    mm[2 * i + 1] = 0u;
  }
}

__global__ void correct_kernel(const float* __restrict__ src, const float* __restrict__ table,
                               float* dst, int width, int height, int pitch, int tpitch) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int y = blockIdx.y * 4 + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= width || y >= height) return;
  const long long o = ((long long)z * height + y) * pitch + x;
  dst[o] = src[o] * table[(long long)y * tpitch + x];
}

// Twiddles e^{-2 pi i k / N}, k < N (the Stockham passes index the whole
// circle), evaluated in FP64 and rounded to float.
__global__ void twiddle_kernel(float2* tw, int N) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < N) {
    const double a = -2.0 * 3.141592653589793238462 * k / N;
    tw[k] = make_float2((float)cos(a), (float)sin(a));
  }
}

void launch_twiddles(float2* tw, int N, hipStream_t st) {
  hipLaunchKernelGGL(twiddle_kernel, dim3((N + 255) / 256), dim3(256), 0, st, tw, N);
}

// Device twiddle tables, one per (device, N), built on the first use and kept
// for the life of the process.
std::mutex g_tw_mutex;
std::map<std::pair<int, int>, float2*> g_tw;

int twiddles(int N, const float2** out) {
  int dev = 0;
  FDK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  auto it = g_tw.find({dev, N});
  if (it != g_tw.end()) {
    *out = it->second;
    return FDK_OK;
  }
  float2* d = nullptr;
  FDK_HIP(hipMalloc(&d, sizeof(float2) * N));
  hipStream_t st = nullptr;
  FDK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  launch_twiddles(d, N, st);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipStreamDestroy(st);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return fdk::fail(FDK_ERR_HIP, "twiddle table: %s", hipGetErrorString(e));
  }
  g_tw[{dev, N}] = d;
  *out = d;
  return FDK_OK;
}

float host_ord2f(unsigned o) {
  const unsigned u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  float f;
  std::memcpy(&f, &u, sizeof(f));
  return f;
}

int log2_exact(int N) {
  int m = 0;
  while ((1 << m) < N) ++m;
  return (1 << m) == N ? m : -1;
}

// 1: Stockham FFT where a plan exists (default); FDK_FILTER_FFT=inplace keeps
// the in-place radix-2^2 FFT for every length.
int fft_choice() {
  static const int k = [] {
    const char* e = std::getenv("FDK_FILTER_FFT");
    return (e && std::strcmp(e, "inplace") == 0) ? 0 : 1;
  }();
  return k;
}

int launch_filter(FilterArgs& a, int nviews, hipStream_t st) {
  FDK_REQUIRE(a.N >= 16 && a.N <= 16384 && log2_exact(a.N) > 0, "filter length %d must be a power of two in 16..16384", a.N);
  FDK_REQUIRE(a.nu >= 4 && a.nu <= a.N, "row length %d must be in 4..N (%d)", a.nu, a.N);
  FDK_REQUIRE(a.pad == FDK_PAD_ZERO || a.pad == FDK_PAD_MEAN3 || a.pad == FDK_PAD_MEAN4,
              "pad_mode %d not in {0, 3, 4}", a.pad);
  FDK_REQUIRE(a.nrows >= 1 && nviews >= 0, "bad row / view count");
  if (nviews == 0) return FDK_OK;
  a.logN = log2_exact(a.N);
  if (!a.tw)
    if (int r = twiddles(a.N, &a.tw)) return r;
  const int plan = (fft_choice() && a.logN >= 9 && a.logN <= 13) ? a.logN : 0;
  const size_t lds = sizeof(float2) * (plan >= 9 && plan <= 12 ? a.N + a.N / 16 : a.N);
  dim3 grid((a.nrows + 1) / 2, nviews);
  switch (plan) {
    case 9: hipLaunchKernelGGL(filter_rows_kernel<9>, grid, dim3(256), lds, st, a); break;
    case 10: hipLaunchKernelGGL(filter_rows_kernel<10>, grid, dim3(256), lds, st, a); break;
    case 11: hipLaunchKernelGGL(filter_rows_kernel<11>, grid, dim3(256), lds, st, a); break;
    case 12: hipLaunchKernelGGL(filter_rows_kernel<12>, grid, dim3(256), lds, st, a); break;
    case 13: hipLaunchKernelGGL(filter_rows_kernel<13>, grid, dim3(256), lds, st, a); break;
    default: hipLaunchKernelGGL(filter_rows_kernel<0>, grid, dim3(256), lds, st, a); break;
  }
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

}  // namespace

extern "C" int fdk_correct_image(const float* dev_src, const float* dev_table, float* dev_dst,
                                 int width, int height, int depth, int pitch, int table_pitch,
                                 void* stream) {
  FDK_REQUIRE(dev_src && dev_table && dev_dst, "NULL pointer");
  FDK_REQUIRE(width > 0 && height > 0 && depth >= 0 && pitch >= width && table_pitch >= width,
              "bad image size");
  if (depth == 0) return FDK_OK;
  dim3 grid((width + 63) / 64, (height + 3) / 4, depth);
  hipLaunchKernelGGL(correct_kernel, grid, dim3(64, 4), 0, static_cast<hipStream_t>(stream),
                     dev_src, dev_table, dev_dst, width, height, pitch, table_pitch);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_filter_rows(const float* dev_src, float* dev_dst, int width, int height,
                               int depth, int src_pitch, int dst_pitch, const float* dev_filter,
                               int N, int pad_mode, void* stream) {
  FDK_REQUIRE(dev_src && dev_dst && dev_filter, "NULL pointer");
  FDK_REQUIRE(src_pitch >= width && dst_pitch >= width, "pitch smaller than width");
  FilterArgs a{};
  a.src = dev_src;
  a.src_pitch = src_pitch;
  a.src_view = (long long)src_pitch * height;
  a.dst = dev_dst;
  a.dst_pitch = dst_pitch;
  a.dst_view = (long long)dst_pitch * height;
  a.transpose = 0;
  a.cos_table = nullptr;
  a.H = dev_filter;
  a.nu = width;
  a.nrows = height;
  a.N = N;
  a.pad = pad_mode;
  a.minmax = nullptr;
  // In place is safe: every row pair is read into LDS before it is written.
  return launch_filter(a, depth, static_cast<hipStream_t>(stream));
}

namespace {

FilterArgs view_args(const float* dev_src, int64_t view_stride_in, int nu, int nv,
                     const float* dev_cos, const float* dev_filter, int N, int pad_mode,
                     int transpose) {
  FilterArgs a{};
  a.src = dev_src;
  a.src_view = view_stride_in;
  a.src_pitch = nu;
  a.transpose = transpose;
  a.dst_pitch = nu;
  a.cos_table = dev_cos;
  a.H = dev_filter;
  a.nu = nu;
  a.nrows = nv;
  a.N = N;
  a.pad = pad_mode;
  return a;
}

// 16-bit output: FP32 filtered batch into `stage`, per-view {min, max} into
// `mm`, then (p - min) / (max - min) quantized into dst.  All on `st`.
int filter_views_q(FilterArgs a, int nviews, void* dev_dst, int64_t view_stride_out, int dtype,
                   float* stage, unsigned* mm, hipStream_t st) {
  const size_t view = (size_t)a.nu * a.nrows;
  hipLaunchKernelGGL(minmax_init_kernel, dim3((nviews + 255) / 256), dim3(256), 0, st, mm, nviews);
  a.dst = stage;
  a.dst_view = (long long)view;
  a.minmax = mm;
  if (int r = launch_filter(a, nviews, st)) return r;
  // 16-byte aligned FP32 rows and 8-byte aligned 16-bit rows: the staging
  // batch is a pool allocation, dev_dst's alignment is checked
  const bool vec = view % 4 == 0 && view_stride_out % 4 == 0 &&
                   reinterpret_cast<uintptr_t>(dev_dst) % 8 == 0;
  const size_t items = vec ? view / 4 : view;
  const int blocks = (int)std::min<size_t>(1024, (items + 255) / 256);
  const dim3 g(blocks, nviews);
  const long long n = (long long)view, sv = (long long)view, dv = (long long)view_stride_out;
  if (dtype == FDK_F16) {
    if (vec) hipLaunchKernelGGL((quantize_kernel<FDK_F16, true>), g, dim3(256), 0, st, stage, sv, dev_dst, dv, n, mm);
    else hipLaunchKernelGGL((quantize_kernel<FDK_F16, false>), g, dim3(256), 0, st, stage, sv, dev_dst, dv, n, mm);
  } else {
    if (vec) hipLaunchKernelGGL((quantize_kernel<FDK_BF16, true>), g, dim3(256), 0, st, stage, sv, dev_dst, dv, n, mm);
    else hipLaunchKernelGGL((quantize_kernel<FDK_BF16, false>), g, dim3(256), 0, st, stage, sv, dev_dst, dv, n, mm);
  }
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

int check_views(const float* dev_src, void* dev_dst, const float* dev_filter, int transpose,
                int dtype, int nu, int nv, int64_t view_stride_in, int64_t view_stride_out) {
  FDK_REQUIRE(dev_src && dev_dst && dev_filter, "NULL pointer");
  FDK_REQUIRE(transpose == 0 || transpose == 1, "transpose must be 0 or 1");
  FDK_REQUIRE(dtype == FDK_F32 || dtype == FDK_F16 || dtype == FDK_BF16, "unknown dtype %d", dtype);
  FDK_REQUIRE(nu > 0 && nv > 0, "bad view size %d x %d", nu, nv);
  FDK_REQUIRE(view_stride_in >= (int64_t)nu * nv && view_stride_out >= (int64_t)nu * nv,
              "view strides smaller than a view");
  return FDK_OK;
}

}  // namespace

extern "C" int fdk_filter_views(const float* dev_src, int64_t view_stride_in, void* dev_dst,
                                int64_t view_stride_out, int dtype, int nu, int nv, int nviews,
                                const float* dev_cos, const float* dev_filter, int N, int pad_mode,
                                int transpose, float* host_scale_offset, void* stream) {
  if (int r = check_views(dev_src, dev_dst, dev_filter, transpose, dtype, nu, nv, view_stride_in,
                          view_stride_out))
    return r;
  hipStream_t st = static_cast<hipStream_t>(stream);
  FilterArgs a = view_args(dev_src, view_stride_in, nu, nv, dev_cos, dev_filter, N, pad_mode,
                           transpose);
  if (dtype == FDK_F32) {
    a.dst = static_cast<float*>(dev_dst);
    a.dst_view = view_stride_out;
    a.minmax = nullptr;
    return launch_filter(a, nviews, st);
  }
  FDK_REQUIRE(host_scale_offset != nullptr, "FP16/BF16 output needs host_scale_offset");
  if (nviews <= 0) return FDK_OK;
  // FP32 staging of the filtered batch, per-view min/max, then quantize.
  const size_t view = (size_t)nu * nv;
  float* tmp = nullptr;
  unsigned* mm = nullptr;
  if (int r = fdk::stage_alloc((void**)&tmp, view * nviews * sizeof(float), st)) return r;
  if (int r = fdk::stage_alloc((void**)&mm, 2 * sizeof(unsigned) * nviews, st)) {
    (void)hipFreeAsync(tmp, st);
    return r;
  }
  int r = filter_views_q(a, nviews, dev_dst, view_stride_out, dtype, tmp, mm, st);
  std::vector<unsigned> hmm(2 * nviews);
  if (r == FDK_OK) {
    hipError_t e = hipMemcpyAsync(hmm.data(), mm, hmm.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "min/max readback: %s", hipGetErrorString(e));
  }
  (void)hipFreeAsync(tmp, st);
  (void)hipFreeAsync(mm, st);
  if (r != FDK_OK) return r;
  for (int v = 0; v < nviews; ++v) {
    const unsigned lo_o = hmm[2 * v], hi_o = hmm[2 * v + 1];
    const float lo = host_ord2f(lo_o), hi = host_ord2f(hi_o);
    host_scale_offset[2 * v] = hi - lo;
    host_scale_offset[2 * v + 1] = lo;
  }
  return FDK_OK;
}

extern "C" int fdk_filter_views_async(const float* dev_src, int64_t view_stride_in, void* dev_dst,
                                      int64_t view_stride_out, int dtype, int nu, int nv,
                                      int nviews, const float* dev_cos, const float* dev_filter,
                                      int N, int pad_mode, int transpose, float* dev_workspace,
                                      unsigned* dev_minmax, void* stream) {
  if (int r = check_views(dev_src, dev_dst, dev_filter, transpose, dtype, nu, nv, view_stride_in,
                          view_stride_out))
    return r;
  FDK_REQUIRE(dtype == FDK_F16 || dtype == FDK_BF16, "dtype %d is not a 16-bit storage type", dtype);
  FDK_REQUIRE(dev_workspace && dev_minmax, "workspace or min/max buffer is NULL");
  if (nviews <= 0) return FDK_OK;
  FilterArgs a = view_args(dev_src, view_stride_in, nu, nv, dev_cos, dev_filter, N, pad_mode,
                           transpose);
  return filter_views_q(a, nviews, dev_dst, view_stride_out, dtype, dev_workspace, dev_minmax,
                        static_cast<hipStream_t>(stream));
}

extern "C" int fdk_filter_proj_image(float* host_proj, int nu, int nv, int nviews,
                                     const float* host_cos, const float* host_filter, int N,
                                     int pad_mode, int depth, float* elapsed_ms) {
  FDK_REQUIRE(host_proj && host_filter, "NULL pointer");
  FDK_REQUIRE(nu > 0 && nv > 0 && nviews >= 0, "bad size");
  if (elapsed_ms) *elapsed_ms = 0.f;
  if (depth <= 0) depth = 32;
  const size_t view = (size_t)nu * nv;
  hipStream_t st = nullptr;
  FDK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float *d = nullptr, *dcos = nullptr, *dH = nullptr;
  int r = FDK_OK;
  hipError_t e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  if (e == hipSuccess) e = hipMalloc(&d, view * depth * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dH, sizeof(float) * N);
  if (e == hipSuccess) e = hipMemcpy(dH, host_filter, sizeof(float) * N, hipMemcpyHostToDevice);
  if (e == hipSuccess && host_cos) e = hipMalloc(&dcos, view * sizeof(float));
  if (e == hipSuccess && host_cos) e = hipMemcpy(dcos, host_cos, view * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) r = fdk::fail(e == hipErrorOutOfMemory ? FDK_ERR_OOM : FDK_ERR_HIP, "FilterProjImage setup: %s", hipGetErrorString(e));
  // Each batch -- upload, cos weighting + filter, download -- is timed between
  // two events and the batch times are summed (the reference's per-batch
  // timeGetTime sum, without the set-up).
  for (int v0 = 0; v0 < nviews && r == FDK_OK; v0 += depth) {
    const int cnt = std::min(depth, nviews - v0);
    e = hipEventRecord(ev0, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d, host_proj + view * v0, view * cnt * sizeof(float), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) { r = fdk::fail(FDK_ERR_HIP, "upload: %s", hipGetErrorString(e)); break; }
    r = fdk_filter_views(d, (int64_t)view, d, (int64_t)view, FDK_F32, nu, nv, cnt, dcos, dH, N,
                         pad_mode, 0, nullptr, st);
    if (r != FDK_OK) break;
    e = hipMemcpyAsync(host_proj + view * v0, d, view * cnt * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(ev1, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
    if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "download: %s", hipGetErrorString(e));
    else if (elapsed_ms) *elapsed_ms += ms;
  }
  (void)hipStreamSynchronize(st);
  if (d) (void)hipFree(d);
  if (dH) (void)hipFree(dH);
  if (dcos) (void)hipFree(dcos);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  (void)hipStreamDestroy(st);
  return r;
}

extern "C" int fdk_filter_frame_host(float* host_src, float* host_dst, int nu, int nv,
                                     const float* host_cos, const float* host_filter, int N,
                                     int pad_mode, int transpose) {
  FDK_REQUIRE(host_src && host_dst && host_cos && host_filter, "NULL pointer");
  FDK_REQUIRE(nu > 0 && nv > 0, "bad image size");
  const size_t view = (size_t)nu * nv;
  float *s = nullptr, *d = nullptr, *c = nullptr, *h = nullptr;
  hipError_t e = hipMalloc(&s, view * 4);
  if (e == hipSuccess) e = hipMalloc(&d, view * 4);
  if (e == hipSuccess) e = hipMalloc(&c, view * 4);
  if (e == hipSuccess) e = hipMalloc(&h, sizeof(float) * std::max(N, 1));
  if (e == hipSuccess) e = hipMemcpy(s, host_src, view * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c, host_cos, view * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h, host_filter, sizeof(float) * N, hipMemcpyHostToDevice);
  int r = e == hipSuccess ? FDK_OK : fdk::fail(FDK_ERR_HIP, "FilterEngine upload: %s", hipGetErrorString(e));
  if (r == FDK_OK)
    r = fdk_filter_views(s, (int64_t)view, d, (int64_t)view, FDK_F32, nu, nv, 1, c, h, N, pad_mode,
                         transpose, nullptr, nullptr);
  if (r == FDK_OK) r = fdk_correct_image(s, c, s, nu, nv, 1, nu, nu, nullptr);
  if (r == FDK_OK) {
    e = hipMemcpy(host_dst, d, view * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(host_src, s, view * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "FilterEngine download: %s", hipGetErrorString(e));
  }
  if (s) (void)hipFree(s);
  if (d) (void)hipFree(d);
  if (c) (void)hipFree(c);
  if (h) (void)hipFree(h);
  return r;
}

// ------------------------------------------------------------ no-allocation filter
// fdk_filter_views without any allocation: the 16-bit forms run the filter
// twice (the first pass only forms each view's min/max, the second quantizes
// as it stores) instead of staging the FP32 batch, and the min/max words and
// the twiddle table live in `scratch`, a fixed device array of
// kScratchBytes per device.
namespace {
constexpr int kNoAllocViews = 1024;  // views per call
constexpr int kNoAllocN = 16384;     // FFT length
constexpr size_t kScratchMinmax = 2 * kNoAllocViews * sizeof(unsigned);
constexpr size_t kScratchBytes = kScratchMinmax + kNoAllocN * sizeof(float2);  // 136 KiB
constexpr int kMaxDevices = 64;
}  // namespace

// Per-view {min, max} words, then the twiddle table of the call's N.  Each
// device holds its own copy for the life of the process, shared by every call
// on that device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxDevices] = {};

// The device address of the current device's `scratch`, looked up on the
// first call on that device (the array's name is not a device address in host
// code).
int scratch_ptr(unsigned char** out) {
  int dev = 0;
  FDK_HIP(hipGetDevice(&dev));
  FDK_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fdk::fail(FDK_ERR_HIP, "address of the filter scratch: %s", hipGetErrorString(e));
    g_scratch[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch[dev];
  return FDK_OK;
}
}  // namespace

extern "C" int fdk_filter_views_noalloc(const float* dev_src, int64_t view_stride_in,
                                        void* dev_dst, int64_t view_stride_out, int dtype, int nu,
                                        int nv, int nviews, const float* dev_cos,
                                        const float* dev_filter, int N, int pad_mode,
                                        int transpose, float* host_scale_offset, void* stream) {
  if (int r = check_views(dev_src, dev_dst, dev_filter, transpose, dtype, nu, nv, view_stride_in,
                          view_stride_out))
    return r;
  FDK_REQUIRE(nviews >= 0 && nviews <= kNoAllocViews,
              "%d views outside the no-allocation filter's 0..%d per call", nviews, kNoAllocViews);
  FDK_REQUIRE(N >= 16 && N <= kNoAllocN, "filter length %d outside the no-allocation filter's "
              "16..%d", N, kNoAllocN);
  FDK_REQUIRE(dtype == FDK_F32 || host_scale_offset != nullptr,
              "FP16/BF16 output needs host_scale_offset");
  if (dtype != FDK_F32 && nviews > 0) {
    // the second pass reads the source again: the 16-bit output must not overlap it
    const char* s0 = reinterpret_cast<const char*>(dev_src);
    const char* s1 = s0 + ((size_t)(nviews - 1) * view_stride_in + (size_t)nu * nv) * 4;
    const char* d0 = static_cast<const char*>(dev_dst);
    const char* d1 = d0 + ((size_t)(nviews - 1) * view_stride_out + (size_t)nu * nv) * 2;
    FDK_REQUIRE(d1 <= s0 || d0 >= s1, "16-bit output overlaps the source views");
  }
  if (nviews == 0) return FDK_OK;
  unsigned char* base = nullptr;
  if (int r = scratch_ptr(&base)) return r;
  unsigned* mm = reinterpret_cast<unsigned*>(base);
  float2* tw = reinterpret_cast<float2*>(base + kScratchMinmax);
  hipStream_t st = static_cast<hipStream_t>(stream);
  launch_twiddles(tw, N, st);  // the last call's table may be of another N
  FilterArgs a = view_args(dev_src, view_stride_in, nu, nv, dev_cos, dev_filter, N, pad_mode,
                           transpose);
  a.tw = tw;
  a.dst = static_cast<float*>(dev_dst);
  a.dst_view = view_stride_out;
  if (dtype == FDK_F32) return launch_filter(a, nviews, st);
  hipLaunchKernelGGL(minmax_init_kernel, dim3((nviews + 255) / 256), dim3(256), 0, st, mm, nviews);
  a.qmode = kStoreNone;
  a.minmax = mm;
  if (int r = launch_filter(a, nviews, st)) return r;
  a.qmode = dtype == FDK_F16 ? kStoreF16 : kStoreBF16;
  a.qmm = mm;
  a.minmax = nullptr;
  if (int r = launch_filter(a, nviews, st)) return r;
  // the words land in the caller's array and are turned into {scale, offset} there
  hipError_t e = hipMemcpyAsync(host_scale_offset, mm, 2 * sizeof(unsigned) * nviews,
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return fdk::fail(FDK_ERR_HIP, "min/max readback: %s", hipGetErrorString(e));
  for (int v = 0; v < nviews; ++v) {
    unsigned w[2];
    std::memcpy(w, host_scale_offset + 2 * v, sizeof(w));
    const float lo = host_ord2f(w[0]), hi = host_ord2f(w[1]);
    host_scale_offset[2 * v] = hi - lo;
    host_scale_offset[2 * v + 1] = lo;
  }
  return FDK_OK;
}
