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
// no real-to-complex unpacking is needed.  The transform is an in-place
// radix-4 (radix-2^2) decimation-in-time FFT in LDS (N complex floats, 64 KiB at
// N = 8192), entered in bit-reversed order; the filter multiply doubles as the
// bit-reversal permutation for the inverse pass.  Loads apply the cos table,
// stores crop, optionally transpose (v fastest, what the backprojector reads)
// and convert to the stored dtype.  Twiddles are an FP64-built table.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
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
  const float2* tw;       // N/2 twiddles e^{-2 pi i k / N}
  int nu, nrows, N, logN, pad;
  unsigned* minmax;       // per view {min, max} as ordered uints, or null
};

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

__device__ inline unsigned f2ord(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void filter_rows_kernel(FilterArgs a) {
  extern __shared__ float2 buf[];
  const int view = blockIdx.y;
  const int ra = 2 * blockIdx.x, rb = ra + 1;
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
  __syncthreads();
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
  const float inv = 1.0f / (float)a.N;
  float* d = a.dst + view * a.dst_view;
  float lo = INFINITY, hi = -INFINITY;
  for (int u = threadIdx.x; u < nu; u += blockDim.x) {
    const float2 c = buf[off + u];
    const float fa = c.x * inv, fb = c.y * inv;
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
      if (has_b) d[(long long)rb * a.dst_pitch + u] = fb;
    }
    lo = fminf(lo, has_b ? fminf(fa, fb) : fa);
    hi = fmaxf(hi, has_b ? fmaxf(fa, fb) : fa);
  }
  if (a.minmax) {
    // wave-level reduction, then one atomic per wave
    for (int o = 32; o > 0; o >>= 1) {
      lo = fminf(lo, __shfl_xor(lo, o));
      hi = fmaxf(hi, __shfl_xor(hi, o));
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin(&a.minmax[2 * view], f2ord(lo));
      atomicMax(&a.minmax[2 * view + 1], f2ord(hi));
    }
  }
}

__device__ inline float ord2f(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// FP32 staging -> FP16/BF16 with q = (p - min) / (max - min) per view.
template <int DT>
__global__ __launch_bounds__(256) void quantize_kernel(const float* __restrict__ src, long long sview,
                                                       void* dst, long long dview, long long n,
                                                       const unsigned* minmax) {
  const int view = blockIdx.y;
  const float lo = ord2f(minmax[2 * view]), hi = ord2f(minmax[2 * view + 1]);
  const float range = hi - lo;
  const float inv = range > 0.f ? 1.0f / range : 0.f;
  const float* s = src + view * sview;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float q = (s[i] - lo) * inv;
    if (DT == FDK_F16) {
      static_cast<__half*>(dst)[view * dview + i] = __float2half_rn(q);
    } else {
      const unsigned b = __float_as_uint(q);
      const unsigned r = (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;  // round to nearest even
      static_cast<unsigned short*>(dst)[view * dview + i] = (unsigned short)r;
    }
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

// Device twiddle tables, one per (device, N), built in FP64 and kept for the
// life of the process.
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
  std::vector<float2> h(N / 2);
  for (int k = 0; k < N / 2; ++k) {
    const double a = -2.0 * 3.141592653589793238462 * k / N;
    h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  FDK_HIP(hipMalloc(&d, sizeof(float2) * h.size()));
  FDK_HIP(hipMemcpy(d, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice));
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

int launch_filter(FilterArgs& a, int nviews, hipStream_t st) {
  FDK_REQUIRE(a.N >= 16 && a.N <= 16384 && log2_exact(a.N) > 0, "filter length %d must be a power of two in 16..16384", a.N);
  FDK_REQUIRE(a.nu >= 4 && a.nu <= a.N, "row length %d must be in 4..N (%d)", a.nu, a.N);
  FDK_REQUIRE(a.pad == FDK_PAD_ZERO || a.pad == FDK_PAD_MEAN3 || a.pad == FDK_PAD_MEAN4,
              "pad_mode %d not in {0, 3, 4}", a.pad);
  FDK_REQUIRE(a.nrows >= 1 && nviews >= 0, "bad row / view count");
  if (nviews == 0) return FDK_OK;
  a.logN = log2_exact(a.N);
  if (int r = twiddles(a.N, &a.tw)) return r;
  const size_t lds = sizeof(float2) * a.N;
  dim3 grid((a.nrows + 1) / 2, nviews);
  hipLaunchKernelGGL(filter_rows_kernel, grid, dim3(256), lds, st, a);
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

extern "C" int fdk_filter_views(const float* dev_src, int64_t view_stride_in, void* dev_dst,
                                int64_t view_stride_out, int dtype, int nu, int nv, int nviews,
                                const float* dev_cos, const float* dev_filter, int N, int pad_mode,
                                int transpose, float* host_scale_offset, void* stream) {
  FDK_REQUIRE(dev_src && dev_dst && dev_filter, "NULL pointer");
  FDK_REQUIRE(transpose == 0 || transpose == 1, "transpose must be 0 or 1");
  FDK_REQUIRE(dtype == FDK_F32 || dtype == FDK_F16 || dtype == FDK_BF16, "unknown dtype %d", dtype);
  FDK_REQUIRE(view_stride_in >= (int64_t)nu * nv && view_stride_out >= (int64_t)nu * nv,
              "view strides smaller than a view");
  hipStream_t st = static_cast<hipStream_t>(stream);
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
  FDK_HIP(hipMallocAsync((void**)&tmp, view * nviews * sizeof(float), st));
  FDK_HIP(hipMallocAsync((void**)&mm, 2 * sizeof(unsigned) * nviews, st));
  hipLaunchKernelGGL(minmax_init_kernel, dim3((nviews + 255) / 256), dim3(256), 0, st, mm, nviews);
  a.dst = tmp;
  a.dst_view = (long long)view;
  a.minmax = mm;
  int r = launch_filter(a, nviews, st);
  if (r == FDK_OK) {
    const int blocks = (int)std::min<size_t>(1024, (view + 255) / 256);
    if (dtype == FDK_F16)
      hipLaunchKernelGGL(quantize_kernel<FDK_F16>, dim3(blocks, nviews), dim3(256), 0, st, tmp,
                         (long long)view, dev_dst, (long long)view_stride_out, (long long)view, mm);
    else
      hipLaunchKernelGGL(quantize_kernel<FDK_BF16>, dim3(blocks, nviews), dim3(256), 0, st, tmp,
                         (long long)view, dev_dst, (long long)view_stride_out, (long long)view, mm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "quantize: %s", hipGetErrorString(e));
  }
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

extern "C" int fdk_filter_proj_image(float* host_proj, int nu, int nv, int nviews,
                                     const float* host_cos, const float* host_filter, int N,
                                     int pad_mode, int depth) {
  FDK_REQUIRE(host_proj && host_filter, "NULL pointer");
  FDK_REQUIRE(nu > 0 && nv > 0 && nviews >= 0, "bad size");
  if (depth <= 0) depth = 32;
  const size_t view = (size_t)nu * nv;
  hipStream_t st = nullptr;
  FDK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  float *d = nullptr, *dcos = nullptr, *dH = nullptr;
  int r = FDK_OK;
  hipError_t e = hipMalloc(&d, view * depth * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dH, sizeof(float) * N);
  if (e == hipSuccess) e = hipMemcpy(dH, host_filter, sizeof(float) * N, hipMemcpyHostToDevice);
  if (e == hipSuccess && host_cos) e = hipMalloc(&dcos, view * sizeof(float));
  if (e == hipSuccess && host_cos) e = hipMemcpy(dcos, host_cos, view * sizeof(float), hipMemcpyHostToDevice);
  if (e != hipSuccess) r = fdk::fail(e == hipErrorOutOfMemory ? FDK_ERR_OOM : FDK_ERR_HIP, "FilterProjImage setup: %s", hipGetErrorString(e));
  for (int v0 = 0; v0 < nviews && r == FDK_OK; v0 += depth) {
    const int cnt = std::min(depth, nviews - v0);
    e = hipMemcpyAsync(d, host_proj + view * v0, view * cnt * sizeof(float), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) { r = fdk::fail(FDK_ERR_HIP, "upload: %s", hipGetErrorString(e)); break; }
    r = fdk_filter_views(d, (int64_t)view, d, (int64_t)view, FDK_F32, nu, nv, cnt, dcos, dH, N,
                         pad_mode, 0, nullptr, st);
    if (r != FDK_OK) break;
    e = hipMemcpyAsync(host_proj + view * v0, d, view * cnt * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "download: %s", hipGetErrorString(e));
  }
  (void)hipStreamSynchronize(st);
  if (d) (void)hipFree(d);
  if (dH) (void)hipFree(dH);
  if (dcos) (void)hipFree(dcos);
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
