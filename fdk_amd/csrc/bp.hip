// Voxel-driven FDK backprojection for gfx950 (CDNA4), 32 views per pass.
//
// Reference behaviour (peng-ac/fdk):
//   cudaFdkBP + SUBPIXEL      cudaBP/backprojectionFDK.cu:15-129
//   BackProjectionFDK         cudaBP/backprojectionFDK.cu:161-302
//   kernel_fdk_3Dgrid         cudaBP/rtkCudaFDKBackProjectionImageFilter.cu:156-191
//   BackProjectionRTK         cudaBP/rtkCudaFDKBackProjectionImageFilter.cu:200-400
//
// For a voxel (x, y, z) and view s the 3x4 matrix M_s gives (u w, v w, w);
// the sample is the bilinear value of the filtered view at (u, v), zero unless
// 0 <= u < nu-1 and 0 <= v < nv-1, weighted by 1/w^2 and summed over views.
//
// MI355X design.
//  * One wave64 owns one voxel column (x, y) and a run of 64*ZPT voxels along z
//    (lane = z), plus their mirror voxels nz-1-z when the sub-volume has a
//    mirror slab.  For the circular-orbit matrices rows 0 and 2 do not depend on
//    z, so u, w and the u-direction bilinear weights are per (column, view)
//    values, and v is affine in z: each voxel update is one FMA for v, two 8-byte
//    gathers (rows u and u+1 of the transposed view, taps v and v+1) and five
//    FMAs.  Lanes along z read one contiguous strip of each detector row, so a
//    wave's gathers touch only a few 128-byte lines.
//  * Matrices (and per-view scale/offset for FP16/BF16 storage) travel in the
//    kernel argument block: uniform scalar loads, nothing global per handle, so
//    two handles on one device never share state.
//  * Volume in HBM is [x][y][zz] with zz fastest (zz = top slab then bottom
//    slab), so the once-per-pass read-modify-write of a wave is one coalesced
//    256-byte run.  GetVolumeData transposes to [zz][y][x] on the way out.
//  * Blocks (4 waves = 2x2 columns) walk 32x32-column tiles, tiles walk the
//    plane, and block ids are remapped so that each of the 8 XCDs works on a
//    contiguous run of tiles: neighbouring columns share detector rows in the
//    same L2.
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace {

constexpr int kMaxProj = FDK_BP_MAX_PROJ;
constexpr int kWave = 64;
constexpr int kTile = 16;  // tile edge in 2x2-column patches (32 columns)

struct BpMats {
  float m[kMaxProj][12];
};
struct BpScale {
  float so[kMaxProj][2];  // p = q * so[0] + so[1]
};

struct BpArgs {
  float* vol;            // [xs][ys][zz], zz fastest
  const void* proj;      // views, transposed: H rows (u) of W samples (v)
  long long view_stride; // elements between views
  int W, H;              // stored view width (v extent) and height (u extent)
  int nproj;
  int x0, y0, z0;        // global index of the sub-volume origin (top slab)
  int nxs, nys;          // sub-volume extent in x and y
  int zext;              // z extent handled per column: top depth (mirror) or nzz
  int nzz;               // total zz extent of the buffer
  int nz;                // global nz (mirror voxel z = nz - 1 - z)
  int tiles_x;           // tiles along x
  long long view_bytes;  // view_stride in bytes
  int view_bytes32;      // W * H * sample bytes (< 2^31), the range of a view's resource
  int row_bytes;         // W * sample bytes
  int buf16;             // 16-bit samples with 4-byte aligned rows and views (buffer path)
  const unsigned* mm;    // per-view {min, max} words on the device (16-bit views), or null
};

// The filter's per-view {min, max} of a 16-bit batch, stored as order-preserving
// uints (negative floats bit-inverted, others with the sign bit set).
__device__ __forceinline__ float ord_to_float(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// Dequantisation of view s, p = q * sc + of: from the kernel arguments, or,
// when the batch's min/max stayed on the device, sc = max - min and of = min
// (the same float arithmetic the host applies to the words it reads back).
__device__ __forceinline__ float2 view_so(const BpArgs& a, const BpScale& S, int s) {
  if (a.mm) {
    const float lo = ord_to_float(a.mm[2 * s]), hi = ord_to_float(a.mm[2 * s + 1]);
    return make_float2(hi - lo, lo);
  }
  return make_float2(S.so[s][0], S.so[s][1]);
}

typedef float f32x2_a4 __attribute__((ext_vector_type(2), aligned(4)));

template <int DT>
struct Taps;

template <>
struct Taps<FDK_F32> {
  // Taps (v, v+1) of rows u and u+1 as two unaligned 8-byte loads.
  __device__ static inline void load(const void* base, long long off, int W, float& a0, float& a1,
                                     float& b0, float& b1) {
    const float* p = static_cast<const float*>(base) + off;
    const f32x2_a4 a = *reinterpret_cast<const f32x2_a4*>(p);
    const f32x2_a4 b = *reinterpret_cast<const f32x2_a4*>(p + W);
    a0 = a.x; a1 = a.y; b0 = b.x; b1 = b.y;
  }
};

template <>
struct Taps<FDK_F16> {
  __device__ static inline void load(const void* base, long long off, int W, float& a0, float& a1,
                                     float& b0, float& b1) {
    const __half* p = static_cast<const __half*>(base) + off;
    a0 = __half2float(p[0]); a1 = __half2float(p[1]);
    b0 = __half2float(p[W]); b1 = __half2float(p[W + 1]);
  }
};

template <>
struct Taps<FDK_BF16> {
  __device__ static inline float bf(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
  __device__ static inline void load(const void* base, long long off, int W, float& a0, float& a1,
                                     float& b0, float& b1) {
    const unsigned short* p = static_cast<const unsigned short*>(base) + off;
    a0 = bf(p[0]); a1 = bf(p[1]); b0 = bf(p[W]); b1 = bf(p[W + 1]);
  }
};

// One gathered, weighted bilinear sample at (u fixed per row pair, v):
// returns w0 * lerp(row u) + w1 * lerp(row u+1), or 0 outside 0 <= v < W-1.
template <int DT>
__device__ inline float sample_v(const void* rows, float v, int W, float w0, float w1) {
  const float fv = floorf(v);
  const int iv = (int)fv;
  const bool ok = (unsigned)iv <= (unsigned)(W - 2);  // 0 <= v < W-1
  const float t = v - fv;
  float a0, a1, b0, b1;
  Taps<DT>::load(rows, ok ? iv : 0, W, a0, a1, b0, b1);
  const float ra = fmaf(t, a1 - a0, a0);
  const float rb = fmaf(t, b1 - b0, b0);
  const float s = fmaf(w0, ra, w1 * rb);
  return ok ? s : 0.0f;
}

// FP32 sample through the row-u / row-u+1 resources: one 32-bit offset, two
// 8-byte loads (taps v, v+1 of each row).  0 outside 0 <= v < W-1.
__device__ inline float sample_buf(__amdgpu_buffer_rsrc_t r0, __amdgpu_buffer_rsrc_t r1, float v,
                                   int rowbase, int W, float w0, float w1) {
  const float fv = floorf(v);
  const int iv = (int)fv;
  const bool ok = (unsigned)iv <= (unsigned)(W - 2);
  const float t = v - fv;
  const int off = (int)(((unsigned)rowbase + (unsigned)iv) << 2);  // wraps harmlessly when !ok
  const float2 a = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r0, off, 0, 0));
  const float2 b = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r1, off, 0, 0));
  const float ra = fmaf(t, a.y - a.x, a.x);
  const float rb = fmaf(t, b.y - b.x, b.x);
  return ok ? fmaf(w0, ra, w1 * rb) : 0.0f;
}

// 16-bit sample (FP16 or BF16 stored, FP32 math) through the same two
// resources.  The taps e = rowbase + iv and e + 1 sit in the 8-byte window that
// starts at the 4-byte aligned element e & ~1; alignbyte picks the 32 bits that
// hold them.  With an even row width W the window of every valid sample lies
// inside its resource (a window that crosses the end only drops taps that the
// sample does not use).  Returns the interpolated stored value q (weighted by
// w0, w1) and sets *ok; the caller applies the per-view affine scale.
template <int DT>
__device__ inline float sample_buf16(__amdgpu_buffer_rsrc_t r0, __amdgpu_buffer_rsrc_t r1, float v,
                                     int rowbase, int W, float w0, float w1, bool* okp) {
  const float fv = floorf(v);
  const int iv = (int)fv;
  const bool ok = (unsigned)iv <= (unsigned)(W - 2);
  const float t = v - fv;
  const unsigned e = (unsigned)rowbase + (unsigned)iv;
  const int off = (int)((e & ~1u) << 1);
  const unsigned sh = (e & 1u) << 1;
  const auto a = __builtin_amdgcn_raw_buffer_load_b64(r0, off, 0, 0);
  const auto b = __builtin_amdgcn_raw_buffer_load_b64(r1, off, 0, 0);
  const unsigned wa = __builtin_amdgcn_alignbyte(a[1], a[0], sh);
  const unsigned wb = __builtin_amdgcn_alignbyte(b[1], b[0], sh);
  float a0, a1, b0, b1;
  if constexpr (DT == FDK_F16) {
    a0 = __half2float(__ushort_as_half((unsigned short)(wa & 0xffffu)));
    a1 = __half2float(__ushort_as_half((unsigned short)(wa >> 16)));
    b0 = __half2float(__ushort_as_half((unsigned short)(wb & 0xffffu)));
    b1 = __half2float(__ushort_as_half((unsigned short)(wb >> 16)));
  } else {
    a0 = __uint_as_float(wa << 16);
    a1 = __uint_as_float(wa & 0xffff0000u);
    b0 = __uint_as_float(wb << 16);
    b1 = __uint_as_float(wb & 0xffff0000u);
  }
  const float ra = fmaf(t, a1 - a0, a0);
  const float rb = fmaf(t, b1 - b0, b0);
  *okp = ok;
  return ok ? fmaf(w0, ra, w1 * rb) : 0.0f;
}

// Block id remap so that consecutive logical blocks land on one XCD: hardware
// hands block b to XCD b % 8, logical block = (b % 8) * (n / 8) + b / 8.
__device__ inline unsigned xcd_remap(unsigned b, unsigned n) {
  return (n % 8u == 0u) ? (b % 8u) * (n / 8u) + b / 8u : b;
}

template <int DT, bool SYM, int ZPT>
__global__ __launch_bounds__(256) void bp_column_kernel(BpArgs a, BpMats M, BpScale S) {
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int tile = lb / (kTile * kTile);
  const int within = lb % (kTile * kTile);
  const int px = (tile % a.tiles_x) * kTile + within % kTile;
  const int py = (tile / a.tiles_x) * kTile + within / kTile;
  const int xs = 2 * px + (wave & 1);
  const int ys = 2 * py + (wave >> 1);
  if (xs >= a.nxs || ys >= a.nys) return;  // whole wave leaves together
  const int zbase = blockIdx.y * (kWave * ZPT) + lane;
  if (zbase >= a.zext) return;

  const float x = (float)(a.x0 + xs);
  const float y = (float)(a.y0 + ys);
  float acc[ZPT], accm[ZPT];
  float zt[ZPT], zm[ZPT];
#pragma unroll
  for (int j = 0; j < ZPT; ++j) {
    acc[j] = 0.f;
    accm[j] = 0.f;
    const int zl = zbase + j * kWave;
    zt[j] = (float)(a.z0 + zl);
    zm[j] = (float)(a.nz - 1 - (a.z0 + zl));
  }

  for (int s = 0; s < a.nproj; ++s) {
    const float* m = M.m[s];
    const float wd = fmaf(m[8], x, fmaf(m[9], y, m[11]));
    const float Z = 1.0f / wd;
    const float U = fmaf(m[0], x, fmaf(m[1], y, m[3])) * Z;
    const float Vb = fmaf(m[4], x, fmaf(m[5], y, m[7])) * Z;
    const float dV = m[6] * Z;
    const float Z2 = Z * Z;
    const float fu0 = floorf(U);
    const int iu = (int)fu0;
    const bool uok = (unsigned)iu <= (unsigned)(a.H - 2);  // 0 <= u < H-1
    const float fu = U - fu0;
    float w0 = uok ? (1.0f - fu) * Z2 : 0.0f;
    float w1 = uok ? fu * Z2 : 0.0f;
    float off = 0.0f;
    if (DT != FDK_F32) {
      const float2 so = view_so(a, S, s);
      w0 *= so.x;
      w1 *= so.x;
      off = so.y * Z2;  // bilinear weights sum to 1 inside the view
    }
    if constexpr (DT == FDK_F32) {
      // Two buffer resources per view: rows u (from the view start) and u+1
      // (from one row later), each bounded by the view's end, so every lane
      // offset is range-checked by the hardware and an out-of-view offset
      // reads 0 instead of touching memory.  Lane offsets are 32-bit.
      const char* vbase = static_cast<const char*>(a.proj) + (size_t)s * a.view_bytes;
      const __amdgpu_buffer_rsrc_t r0 =
          __builtin_amdgcn_make_buffer_rsrc((void*)vbase, 0, a.view_bytes32, fdk::kRawRsrcWord3);
      const __amdgpu_buffer_rsrc_t r1 = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(vbase + a.row_bytes), 0, a.view_bytes32 - a.row_bytes, fdk::kRawRsrcWord3);
      const int rowbase = uok ? iu * a.W : 0;
#pragma unroll
      for (int j = 0; j < ZPT; ++j) {
        acc[j] += sample_buf(r0, r1, fmaf(zt[j], dV, Vb), rowbase, a.W, w0, w1);
        if (SYM) accm[j] += sample_buf(r0, r1, fmaf(zm[j], dV, Vb), rowbase, a.W, w0, w1);
      }
    } else if (a.buf16) {
      const char* vbase = static_cast<const char*>(a.proj) + (size_t)s * a.view_bytes;
      const __amdgpu_buffer_rsrc_t r0 =
          __builtin_amdgcn_make_buffer_rsrc((void*)vbase, 0, a.view_bytes32, fdk::kRawRsrcWord3);
      const __amdgpu_buffer_rsrc_t r1 = __builtin_amdgcn_make_buffer_rsrc(
          (void*)(vbase + a.row_bytes), 0, a.view_bytes32 - a.row_bytes, fdk::kRawRsrcWord3);
      const int rowbase = uok ? iu * a.W : 0;
#pragma unroll
      for (int j = 0; j < ZPT; ++j) {
        bool ok;
        const float c = sample_buf16<DT>(r0, r1, fmaf(zt[j], dV, Vb), rowbase, a.W, w0, w1, &ok);
        acc[j] += (uok && ok) ? c + off : 0.f;
        if (SYM) {
          const float cm = sample_buf16<DT>(r0, r1, fmaf(zm[j], dV, Vb), rowbase, a.W, w0, w1, &ok);
          accm[j] += (uok && ok) ? cm + off : 0.f;
        }
      }
    } else {
      const long long rowoff = (long long)s * a.view_stride + (long long)(uok ? iu : 0) * a.W;
      const void* rows = static_cast<const unsigned short*>(a.proj) + rowoff;
#pragma unroll
      for (int j = 0; j < ZPT; ++j) {
        const float v = fmaf(zt[j], dV, Vb);
        float c = sample_v<DT>(rows, v, a.W, w0, w1);
        c += (uok && v >= 0.f && v < (float)(a.W - 1)) ? off : 0.f;
        acc[j] += c;
        if (SYM) {
          const float vm = fmaf(zm[j], dV, Vb);
          float cm = sample_v<DT>(rows, vm, a.W, w0, w1);
          cm += (uok && vm >= 0.f && vm < (float)(a.W - 1)) ? off : 0.f;
          accm[j] += cm;
        }
      }
    }
  }

  float* col = a.vol + ((long long)xs * a.nys + ys) * (long long)a.nzz;
#pragma unroll
  for (int j = 0; j < ZPT; ++j) {
    const int zl = zbase + j * kWave;
    if (zl < a.zext) {
      col[zl] += acc[j];
      if (SYM) col[a.nzz - 1 - zl] += accm[j];
    }
  }
}

// ------------------------------------------------------------- LDS-staged kernels
// Per view, a block of bp_lds_kernel / bp_lds_pf_kernel (same blocks, columns
// and lanes as bp_column_kernel) first agrees on the detector strip its four
// columns need -- rows iu..iu+1 of the columns whose u is inside the view, and
// the v range of each wave's whole top (and mirror) z-run -- and stages it into
// LDS as FP32, one wave per strip row, with 16-byte (FP32) or 8-byte
// (FP16/BF16, dequantized on the way) range-checked buffer loads; the taps are
// then LDS reads.  This trades the two vector-memory gathers per voxel update
// for about a quarter of a 16-byte load.  R x CT (and R x CT for the mirror
// run) is sized on the host from the matrices with 4 spare columns; a view
// whose strip does not fit is gathered straight from memory instead.  Samples
// outside the view (and every sample of a column whose u is outside it) read a
// zero slot, the last two columns of rows R-2 and R-1, which staging never
// writes.

// One column's view: v(z) = Vb + z*dV, the u-direction weights times 1/w^2
// (zero when u is outside the view) and the row iu.
struct ColView {
  float Vb, dV, w0, w1, Z2;
  int iu;
  bool uok;
};

__device__ __forceinline__ ColView col_view(const float* m, float x, float y, int H, bool active) {
  ColView c;
  const float wd = fmaf(m[8], x, fmaf(m[9], y, m[11]));
  const float Z = 1.0f / wd;
  const float U = fmaf(m[0], x, fmaf(m[1], y, m[3])) * Z;
  c.Vb = fmaf(m[4], x, fmaf(m[5], y, m[7])) * Z;
  c.dV = m[6] * Z;
  c.Z2 = Z * Z;
  const float fu0 = floorf(U);
  c.iu = (int)fu0;
  c.uok = active && (unsigned)c.iu <= (unsigned)(H - 2);
  const float fu = U - fu0;
  c.w0 = c.uok ? (1.0f - fu) * c.Z2 : 0.0f;
  c.w1 = c.uok ? fu * c.Z2 : 0.0f;
  return c;
}

// Lane 0 of each wave publishes its footprint: info[k] = value k of waves 0..3
// (row min/max, top-run column min/max, mirror-run column min/max).
template <bool SYM>
__device__ __forceinline__ void footprint_put(int4* info, int wave, const ColView& c, float zt0,
                                              float zt1, float zm0, float zm1) {
  int* inf = reinterpret_cast<int*>(info);
  const float va = fmaf(zt0, c.dV, c.Vb), vb = fmaf(zt1, c.dV, c.Vb);
  inf[0 * 4 + wave] = c.uok ? c.iu : INT_MAX;
  inf[1 * 4 + wave] = c.uok ? c.iu : INT_MIN;
  inf[2 * 4 + wave] = c.uok ? (int)floorf(fminf(va, vb)) - 1 : INT_MAX;
  inf[3 * 4 + wave] = c.uok ? (int)floorf(fmaxf(va, vb)) + 2 : INT_MIN;
  if (SYM) {
    const float ma = fmaf(zm0, c.dV, c.Vb), mb = fmaf(zm1, c.dV, c.Vb);
    inf[4 * 4 + wave] = c.uok ? (int)floorf(fminf(ma, mb)) - 1 : INT_MAX;
    inf[5 * 4 + wave] = c.uok ? (int)floorf(fmaxf(ma, mb)) + 2 : INT_MIN;
  }
}

// The block's strip for one view (wave-uniform values).
struct Strip {
  int r0, nrow, c0, q4, d0, q4m;
  bool fits;
};

__device__ __forceinline__ int min4(int4 v) { return min(min(v.x, v.y), min(v.z, v.w)); }
__device__ __forceinline__ int max4(int4 v) { return max(max(v.x, v.y), max(v.z, v.w)); }

template <bool SYM>
__device__ __forceinline__ Strip strip_get(const int4* info, int W, int R, int CT) {
  Strip st;
  st.r0 = __builtin_amdgcn_readfirstlane(min4(info[0]));
  const int r1 = __builtin_amdgcn_readfirstlane(max4(info[1]));
  int c0 = __builtin_amdgcn_readfirstlane(min4(info[2]));
  int c1 = __builtin_amdgcn_readfirstlane(max4(info[3]));
  int d0 = 0, d1 = -1;
  if (SYM) {
    d0 = __builtin_amdgcn_readfirstlane(min4(info[4]));
    d1 = __builtin_amdgcn_readfirstlane(max4(info[5]));
  }
  st.nrow = (r1 >= st.r0) ? r1 + 2 - st.r0 : 0;
  if (st.nrow == 0) {  // no column of the block is inside the view
    c0 = 0;
    c1 = -1;
    d0 = 0;
    d1 = -1;
  }
  st.c0 = max(c0, 0) & ~3;
  st.q4 = max((min(c1, W - 1) - st.c0 + 4) >> 2, 0);
  st.d0 = 0;
  st.q4m = 0;
  if (SYM) {
    st.d0 = max(d0, 0) & ~3;
    st.q4m = max((min(d1, W - 1) - st.d0 + 4) >> 2, 0);
  }
  st.fits = st.nrow <= R && 4 * st.q4 <= CT - 4 && 4 * st.q4m <= CT - 4;
  return st;
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Four consecutive stored samples: raw bits (FP32: float4; 16-bit: 8 bytes).
template <int DT>
struct Quad {
  typedef float4 raw;
  __device__ static __forceinline__ raw load(__amdgpu_buffer_rsrc_t rv, int byte_off) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rv, byte_off, 0, 0));
  }
  __device__ static __forceinline__ float4 cvt(raw v, float, float) { return v; }
};

template <int DT>
struct Quad16 {
  typedef u32x2 raw;
  __device__ static __forceinline__ raw load(__amdgpu_buffer_rsrc_t rv, int byte_off) {
    return __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(rv, byte_off, 0, 0));
  }
  __device__ static __forceinline__ float4 cvt(raw h, float sc, float of) {
    float f[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const unsigned short bits = (unsigned short)(h[e >> 1] >> (16 * (e & 1)));
      const float qv = DT == FDK_F16 ? __half2float(__ushort_as_half(bits))
                                     : __uint_as_float((unsigned)bits << 16);
      f[e] = fmaf(qv, sc, of);
    }
    return make_float4(f[0], f[1], f[2], f[3]);
  }
};
template <>
struct Quad<FDK_F16> : Quad16<FDK_F16> {};
template <>
struct Quad<FDK_BF16> : Quad16<FDK_BF16> {};

// Strip row rr (top rows first, then mirror rows) of a view: its source byte
// offset in the view, its LDS row and its quad count.
template <bool SYM>
__device__ __forceinline__ void strip_row(const Strip& st, int rr, int W, int esize, int CT,
                                          float* stage, float* stm, int* src, float** dst,
                                          int* qq) {
  const bool mir = SYM && rr >= st.nrow;
  const int r = mir ? rr - st.nrow : rr;
  *qq = mir ? st.q4m : st.q4;
  *src = ((st.r0 + r) * W + (mir ? st.d0 : st.c0)) * esize;
  *dst = (mir ? stm : stage) + r * CT;
}

// Stages a view's strip directly (loads, then LDS stores).
template <int DT, bool SYM>
__device__ __forceinline__ void strip_stage(const Strip& st, __amdgpu_buffer_rsrc_t rv, int W,
                                            int CT, float sc, float of, float* stage, float* stm,
                                            int wave, int lane) {
  const int esize = DT == FDK_F32 ? 4 : 2;
  const int nr_all = SYM ? 2 * st.nrow : st.nrow;
  for (int rr = wave; rr < nr_all; rr += 4) {
    int src, qq;
    float* dst;
    strip_row<SYM>(st, rr, W, esize, CT, stage, stm, &src, &dst, &qq);
    for (int q = lane; q < qq; q += kWave)
      *reinterpret_cast<float4*>(dst + 4 * q) =
          Quad<DT>::cvt(Quad<DT>::load(rv, src + 4 * esize * q), sc, of);
  }
}

// Register-prefetched staging: at most kPfRows strip rows per wave and kPfQuads
// quads per lane and row (the host checks R and CT against these).
constexpr int kPfRows = 3;
constexpr int kPfQuads = 3;

template <int DT, bool SYM>
__device__ __forceinline__ void strip_issue(typename Quad<DT>::raw (&pre)[kPfRows][kPfQuads],
                                            const Strip& st, __amdgpu_buffer_rsrc_t rv, int W,
                                            int CT, float* stage, float* stm, int wave, int lane) {
  const int esize = DT == FDK_F32 ? 4 : 2;
  const int nr_all = SYM ? 2 * st.nrow : st.nrow;
#pragma unroll
  for (int i = 0; i < kPfRows; ++i) {
    const int rr = wave + 4 * i;
    if (rr < nr_all) {
      int src, qq;
      float* dst;
      strip_row<SYM>(st, rr, W, esize, CT, stage, stm, &src, &dst, &qq);
#pragma unroll
      for (int k = 0; k < kPfQuads; ++k) {
        const int q = lane + kWave * k;
        if (q < qq) pre[i][k] = Quad<DT>::load(rv, src + 4 * esize * q);
      }
    }
  }
}

template <int DT, bool SYM>
__device__ __forceinline__ void strip_commit(const typename Quad<DT>::raw (&pre)[kPfRows][kPfQuads],
                                             const Strip& st, int W, int CT, float sc, float of,
                                             float* stage, float* stm, int wave, int lane) {
  const int esize = DT == FDK_F32 ? 4 : 2;
  const int nr_all = SYM ? 2 * st.nrow : st.nrow;
#pragma unroll
  for (int i = 0; i < kPfRows; ++i) {
    const int rr = wave + 4 * i;
    if (rr < nr_all) {
      int src, qq;
      float* dst;
      strip_row<SYM>(st, rr, W, esize, CT, stage, stm, &src, &dst, &qq);
#pragma unroll
      for (int k = 0; k < kPfQuads; ++k) {
        const int q = lane + kWave * k;
        if (q < qq) *reinterpret_cast<float4*>(dst + 4 * q) = Quad<DT>::cvt(pre[i][k], sc, of);
      }
    }
  }
}

// floor(v) as an int and v - floor(v), one instruction each.
__device__ __forceinline__ int floor_int(float v) {
  int iv;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(iv) : "v"(v));
  return iv;
}

// The taps of one view from the staged strip: a tap index iv reads
// row[iv - c0] (a sample inside the view always lies in the staged columns),
// any other sample the zero slot.
template <bool SYM, int ZPT>
__device__ __forceinline__ void compute_lds(float (&acc)[ZPT], float (&accm)[ZPT],
                                            const float (&zt)[ZPT], const float (&zm)[ZPT],
                                            const ColView& c, const Strip& st, const float* stage,
                                            const float* stm, int W, int R, int CT) {
  const int lr = c.uok ? c.iu - st.r0 : 0;
  const int zs = (R - 2 - lr) * CT + CT - 2;  // zero slot, relative to row lr
  const float* rowt = stage + lr * CT - st.c0;
  const float* rowm = stm + lr * CT - st.d0;
  const int zt_idx = zs + st.c0, zm_idx = zs + st.d0;
#pragma unroll
  for (int j = 0; j < ZPT; ++j) {
    {
      const float v = fmaf(zt[j], c.dV, c.Vb);
      const int iv = floor_int(v);
      const bool ok = c.uok && (unsigned)iv <= (unsigned)(W - 2);
      const float t = __builtin_amdgcn_fractf(v);
      const float* p = rowt + (ok ? iv : zt_idx);
      const float ra = fmaf(t, p[1] - p[0], p[0]);
      const float rb = fmaf(t, p[CT + 1] - p[CT], p[CT]);
      acc[j] += fmaf(c.w0, ra, c.w1 * rb);
    }
    if (SYM) {
      const float v = fmaf(zm[j], c.dV, c.Vb);
      const int iv = floor_int(v);
      const bool ok = c.uok && (unsigned)iv <= (unsigned)(W - 2);
      const float t = __builtin_amdgcn_fractf(v);
      const float* p = rowm + (ok ? iv : zm_idx);
      const float ra = fmaf(t, p[1] - p[0], p[0]);
      const float rb = fmaf(t, p[CT + 1] - p[CT], p[CT]);
      accm[j] += fmaf(c.w0, ra, c.w1 * rb);
    }
  }
}

// The taps of one view gathered straight from memory (a strip that does not fit).
template <int DT, bool SYM, int ZPT>
__device__ __forceinline__ void compute_direct(float (&acc)[ZPT], float (&accm)[ZPT],
                                               const float (&zt)[ZPT], const float (&zm)[ZPT],
                                               const ColView& c, const BpArgs& a,
                                               const char* vbase, float sc, float of) {
  const __amdgpu_buffer_rsrc_t g0 =
      __builtin_amdgcn_make_buffer_rsrc((void*)vbase, 0, a.view_bytes32, fdk::kRawRsrcWord3);
  const __amdgpu_buffer_rsrc_t g1 = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(vbase + a.row_bytes), 0, a.view_bytes32 - a.row_bytes, fdk::kRawRsrcWord3);
  const int rowbase = c.iu * a.W;
  const float off = of * c.Z2;
#pragma unroll
  for (int j = 0; j < ZPT; ++j) {
    if constexpr (DT == FDK_F32) {
      acc[j] += sample_buf(g0, g1, fmaf(zt[j], c.dV, c.Vb), rowbase, a.W, c.w0, c.w1);
      if (SYM) accm[j] += sample_buf(g0, g1, fmaf(zm[j], c.dV, c.Vb), rowbase, a.W, c.w0, c.w1);
    } else {
      bool ok;
      const float q = sample_buf16<DT>(g0, g1, fmaf(zt[j], c.dV, c.Vb), rowbase, a.W, c.w0 * sc,
                                       c.w1 * sc, &ok);
      acc[j] += ok ? q + off : 0.f;
      if (SYM) {
        const float qm = sample_buf16<DT>(g0, g1, fmaf(zm[j], c.dV, c.Vb), rowbase, a.W,
                                          c.w0 * sc, c.w1 * sc, &ok);
        accm[j] += ok ? qm + off : 0.f;
      }
    }
  }
}

// Shared prologue of the LDS kernels: block/column indices, z coordinates,
// zero slot(s).
#define FDK_LDS_PROLOGUE                                                              \
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);                              \
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);              \
  const int lane = threadIdx.x % kWave;                                              \
  const int tile = lb / (kTile * kTile);                                             \
  const int within = lb % (kTile * kTile);                                           \
  const int px = (tile % a.tiles_x) * kTile + within % kTile;                        \
  const int py = (tile / a.tiles_x) * kTile + within / kTile;                        \
  const int xs = 2 * px + (wave & 1);                                                \
  const int ys = 2 * py + (wave >> 1);                                               \
  const bool active = xs < a.nxs && ys < a.nys; /* inactive waves join barriers */   \
  const int zb0 = blockIdx.y * (kWave * ZPT);                                        \
  const int zb1 = zb0 + kWave * ZPT - 1; /* the whole run, also past zext */         \
  const int zbase = zb0 + lane;                                                      \
  float* stm = stage + R * CT;                                                       \
  if (threadIdx.x < (SYM ? 8 : 4)) {                                                 \
    const int k = threadIdx.x;                                                       \
    float* zb = (k >= 4 ? stm : stage) + (R - 2 + ((k >> 1) & 1)) * CT + CT - 2;     \
    zb[k & 1] = 0.0f;                                                                \
  }                                                                                  \
  const float x = (float)(a.x0 + xs);                                                \
  const float y = (float)(a.y0 + ys);                                                \
  float acc[ZPT], accm[ZPT], zt[ZPT], zm[ZPT];                                       \
  _Pragma("unroll") for (int j = 0; j < ZPT; ++j) {                                  \
    acc[j] = 0.f;                                                                    \
    accm[j] = 0.f;                                                                   \
    const int zl = zbase + j * kWave;                                                \
    zt[j] = (float)(a.z0 + zl);                                                      \
    zm[j] = (float)(a.nz - 1 - (a.z0 + zl));                                         \
  }                                                                                  \
  const float zt0 = (float)(a.z0 + zb0), zt1 = (float)(a.z0 + zb1);                  \
  const float zm0 = (float)(a.nz - 1 - (a.z0 + zb0));                                \
  const float zm1 = (float)(a.nz - 1 - (a.z0 + zb1));

template <bool SYM, int ZPT>
__device__ __forceinline__ void column_write(const BpArgs& a, int xs, int ys, int zbase,
                                             const float (&acc)[ZPT], const float (&accm)[ZPT]) {
  float* col = a.vol + ((long long)xs * a.nys + ys) * (long long)a.nzz;
#pragma unroll
  for (int j = 0; j < ZPT; ++j) {
    const int zl = zbase + j * kWave;
    if (zl < a.zext) {
      col[zl] += acc[j];
      if (SYM) col[a.nzz - 1 - zl] += accm[j];
    }
  }
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t view_rsrc(const BpArgs& a, int s) {
  const char* vbase = static_cast<const char*>(a.proj) + (size_t)s * a.view_bytes;
  return __builtin_amdgcn_make_buffer_rsrc((void*)vbase, 0, a.view_bytes32, fdk::kRawRsrcWord3);
}

// Stage, then compute, every view (two block barriers per view).
template <int DT, bool SYM, int ZPT>
__global__ __launch_bounds__(256) void bp_lds_kernel(BpArgs a, BpMats M, BpScale S, int R, int CT) {
  extern __shared__ float stage[];  // [R][CT] top run, then [R][CT] mirror run
  __shared__ int4 info[6];
  FDK_LDS_PROLOGUE
  for (int s = 0; s < a.nproj; ++s) {
    const ColView c = col_view(M.m[s], x, y, a.H, active);
    if (lane == 0) footprint_put<SYM>(info, wave, c, zt0, zt1, zm0, zm1);
    __syncthreads();
    const Strip st = strip_get<SYM>(info, a.W, R, CT);
    const float2 so = view_so(a, S, s);
    if (st.fits)
      strip_stage<DT, SYM>(st, view_rsrc(a, s), a.W, CT, so.x, so.y, stage, stm, wave, lane);
    __syncthreads();  // every wave has read info; the next view writes it after its compute
    if (st.fits)
      compute_lds<SYM, ZPT>(acc, accm, zt, zm, c, st, stage, stm, a.W, R, CT);
    else if (c.uok)
      compute_direct<DT, SYM, ZPT>(acc, accm, zt, zm, c, a,
                                   static_cast<const char*>(a.proj) + (size_t)s * a.view_bytes,
                                   so.x, so.y);
  }
  if (active) column_write<SYM, ZPT>(a, xs, ys, zbase, acc, accm);
}

// As bp_lds_kernel, with view s+1's strip loaded into registers while view s
// is computed from LDS, and written to LDS after the block's last read of
// view s: the staging loads' latency hides under the compute.  Needs at most
// kPfRows strip rows per wave and kPfQuads quads per lane (checked on the host).
template <int DT, bool SYM, int ZPT>
__global__ __launch_bounds__(256) void bp_lds_pf_kernel(BpArgs a, BpMats M, BpScale S, int R,
                                                        int CT) {
  extern __shared__ float stage[];
  __shared__ int4 info[2][6];  // footprints of alternate views
  FDK_LDS_PROLOGUE
  if (a.nproj <= 0) return;
  {
    const ColView c0v = col_view(M.m[0], x, y, a.H, active);
    if (lane == 0) footprint_put<SYM>(info[0], wave, c0v, zt0, zt1, zm0, zm1);
  }
  __syncthreads();
  Strip cur = strip_get<SYM>(info[0], a.W, R, CT);
  if (cur.fits) {
    const float2 so0 = view_so(a, S, 0);
    strip_stage<DT, SYM>(cur, view_rsrc(a, 0), a.W, CT, so0.x, so0.y, stage, stm, wave, lane);
  }
  for (int s = 0; s < a.nproj; ++s) {
    const bool more = s + 1 < a.nproj;
    if (more && lane == 0) {
      const ColView cn = col_view(M.m[s + 1], x, y, a.H, active);
      footprint_put<SYM>(info[(s + 1) & 1], wave, cn, zt0, zt1, zm0, zm1);
    }
    __syncthreads();  // strip s is in LDS; the footprint of view s+1 is visible
    Strip nxt;
    nxt.fits = false;
    typename Quad<DT>::raw pre[kPfRows][kPfQuads];
    if (more) {
      nxt = strip_get<SYM>(info[(s + 1) & 1], a.W, R, CT);
      if (nxt.fits) strip_issue<DT, SYM>(pre, nxt, view_rsrc(a, s + 1), a.W, CT, stage, stm, wave, lane);
    }
    const ColView cv = col_view(M.m[s], x, y, a.H, active);
    if (cur.fits)
      compute_lds<SYM, ZPT>(acc, accm, zt, zm, cv, cur, stage, stm, a.W, R, CT);
    else if (cv.uok) {
      const float2 so = view_so(a, S, s);
      compute_direct<DT, SYM, ZPT>(acc, accm, zt, zm, cv, a,
                                   static_cast<const char*>(a.proj) + (size_t)s * a.view_bytes,
                                   so.x, so.y);
    }
    __syncthreads();  // every wave is done with strip s
    if (nxt.fits) {
      const float2 so = view_so(a, S, s + 1);
      strip_commit<DT, SYM>(pre, nxt, a.W, CT, so.x, so.y, stage, stm, wave, lane);
    }
    cur = nxt;
  }
  if (active) column_write<SYM, ZPT>(a, xs, ys, zbase, acc, accm);
}

// ------------------------------------------------------------ wave-strip kernel
// bp_wave_kernel: the same blocks, columns and lanes, for a mirror slab pair
// whose matrices are mirror-symmetric (v(z) + v(nz-1-z) = W-1 for every
// column, checked on the host), with no block-level synchronisation:
//  * each wave stages, per view, ITS OWN u-interpolated detector strips in LDS:
//    r(c) = w0 P[iu][c] + w1 P[iu+1][c] (w0, w1 the u weights times 1/w^2) over
//    the v columns of its top z-run, and the same for the mirror run, which by
//    symmetry reads columns W-1-c, stored reversed; top and mirror samples are
//    interleaved, so one 16-byte LDS read brings r_top(iv), r_mir(iv),
//    r_top(iv+1), r_mir(iv+1);
//  * the mirror voxel has v' = W-1-v, so iv' = W-2-iv and t' = 1-t: its
//    sample is r_mir(iv) + t (r_mir(iv+1) - r_mir(iv)) with the top voxel's iv
//    and t, and the same validity (0 <= iv <= W-2);
//  * the per-(column, view) values (u weights, v = Vb + z dV, strip bounds and
//    flags) are computed once, lane s for view s, and broadcast with readlane;
//  * the next view's strip loads are issued into registers before the current
//    view is computed from LDS.
// Per voxel pair: v, floor, fraction, 1-t, the LDS address (and the validity
// test where the run crosses the detector edge) and four FMAs.
constexpr int kWsQuads = 3;  // strip quads per lane and run (the host checks CT against it)
// Quad slots a lane stages at a given ZPT: a 256-voxel run's strip fits 2 x 64
// quads, and an always-empty third slot would still cost its (range-checked,
// zero) loads and offsets every view.
template <int ZPT>
constexpr int ws_quads() { return ZPT == 4 ? 2 : kWsQuads; }

// Per-view values of one column, kept in lane s for view s.
struct WaveView {
  float Vb, dV, w0, w1, Z2;
  int iu;
  int c0;     // first strip column (multiple of 4)
  int quads;  // strip quads of 4 columns
  int flags;  // 1: some voxel of the runs inside the view, 2: all inside, 4: the strip fits
};

template <int DT>
__device__ __forceinline__ float4 ws_combine(typename Quad<DT>::raw A, typename Quad<DT>::raw B,
                                             float w0, float w1, float zof) {
  if constexpr (DT == FDK_F32) {
    return make_float4(fmaf(w1, B.x, w0 * A.x), fmaf(w1, B.y, w0 * A.y), fmaf(w1, B.z, w0 * A.z),
                       fmaf(w1, B.w, w0 * A.w));
  } else {
    // w0, w1 already carry the view's scale; zof = (w0 + w1) * offset
    const float4 qa = Quad<DT>::cvt(A, 1.0f, 0.0f), qb = Quad<DT>::cvt(B, 1.0f, 0.0f);
    return make_float4(fmaf(w1, qb.x, fmaf(w0, qa.x, zof)), fmaf(w1, qb.y, fmaf(w0, qa.y, zof)),
                       fmaf(w1, qb.z, fmaf(w0, qa.z, zof)), fmaf(w1, qb.w, fmaf(w0, qa.w, zof)));
  }
}

template <int DT>
struct WsPre {
  typename Quad<DT>::raw ta[kWsQuads], tb[kWsQuads], ma[kWsQuads], mb[kWsQuads];  // NQ <= kWsQuads used
};

// Strip loads of one view: quad q of the top run reads columns c0+4q..c0+4q+3
// of rows iu and iu+1; of the mirror run, columns W-4-c0-4q..W-1-c0-4q.
// Columns outside the row read other samples of the view or (negative or past
// the view) zero through the buffer range check; only columns of valid taps
// are ever used.
// FP32: branch-free, every offset computed before the first load and quads past
// the strip given an offset past the view (zeros, never used), so the loads of
// one view issue back to back and no register reuse has to wait on them.
// 16-bit views (8-byte loads, half the registers in flight) keep a branch per
// quad slot, which measured faster for them (2577 against 2463 GUPS in FP16).
template <int DT, int NQ = kWsQuads>
__device__ __forceinline__ void ws_issue(WsPre<DT>& pre, __amdgpu_buffer_rsrc_t rv, int W, int iu,
                                         int c0, int quads, int lane) {
  const int es = DT == FDK_F32 ? 4 : 2;
  const int rowa = iu * W * es, rowb = rowa + W * es;
  if constexpr (DT != FDK_F32) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const int q = lane + kWave * k;
      if (q < quads) {
        const int ct = (c0 + 4 * q) * es, cm = (W - 4 - c0 - 4 * q) * es;
        pre.ta[k] = Quad<DT>::load(rv, rowa + ct);
        pre.tb[k] = Quad<DT>::load(rv, rowb + ct);
        pre.ma[k] = Quad<DT>::load(rv, rowa + cm);
        pre.mb[k] = Quad<DT>::load(rv, rowb + cm);
      }
    }
    return;
  }
  constexpr int kOut = (int)0xFFFFFFF0u;  // past any view's buffer range
  int ta[NQ], tb[NQ], ma[NQ], mb[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int q = lane + kWave * k;
    const bool ok = q < quads;
    const int ct = (c0 + 4 * q) * es, cm = (W - 4 - c0 - 4 * q) * es;
    ta[k] = ok ? rowa + ct : kOut;
    tb[k] = ok ? rowb + ct : kOut;
    ma[k] = ok ? rowa + cm : kOut;
    mb[k] = ok ? rowb + cm : kOut;
  }
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    pre.ta[k] = Quad<DT>::load(rv, ta[k]);
    pre.tb[k] = Quad<DT>::load(rv, tb[k]);
    pre.ma[k] = Quad<DT>::load(rv, ma[k]);
    pre.mb[k] = Quad<DT>::load(rv, mb[k]);
  }
}

template <int DT, int NQ = kWsQuads>
__device__ __forceinline__ void ws_commit(const WsPre<DT>& pre, float2* strip, float w0, float w1,
                                          float zof, int quads, int lane) {
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int q = lane + kWave * k;
    if (q < quads) {
      const float4 t = ws_combine<DT>(pre.ta[k], pre.tb[k], w0, w1, zof);
      const float4 m = ws_combine<DT>(pre.ma[k], pre.mb[k], w0, w1, zof);
      float4* d = reinterpret_cast<float4*>(strip + 4 * q);
      d[0] = make_float4(t.x, m.w, t.y, m.z);  // mirror quad reversed
      d[1] = make_float4(t.z, m.y, t.w, m.x);
    }
  }
}

template <bool CHECK, int ZPT>
__device__ __forceinline__ void ws_compute(float (&acc)[ZPT], float (&accm)[ZPT],
                                           const float (&zt)[ZPT], float Vb, float dV,
                                           const float2* base, const float2* zslot, int W) {
#pragma unroll
  for (int j = 0; j < ZPT; ++j) {
    const float v = fmaf(zt[j], dV, Vb);
    const int iv = floor_int(v);
    const float t = __builtin_amdgcn_fractf(v);
    const float2* p = base + iv;
    if (CHECK) p = (unsigned)iv <= (unsigned)(W - 2) ? p : zslot;
    const float2 a = p[0], b = p[1];
    const float omt = 1.0f - t;
    acc[j] = fmaf(t, b.x, fmaf(omt, a.x, acc[j]));
    accm[j] = fmaf(t, b.y, fmaf(omt, a.y, accm[j]));
  }
}

__device__ __forceinline__ float lane_f(float v, int s) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), s));
}
__device__ __forceinline__ int lane_i(int v, int s) { return __builtin_amdgcn_readlane(v, s); }

template <int DT, int ZPT>
__global__ __launch_bounds__(256) void bp_wave_kernel(BpArgs a, BpMats M, BpScale S, int CT) {
  extern __shared__ float2 wstrip[];  // per wave: CT float2, the last two the zero slot
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int lane = threadIdx.x % kWave;
  const int tile = lb / (kTile * kTile);
  const int within = lb % (kTile * kTile);
  const int px = (tile % a.tiles_x) * kTile + within % kTile;
  const int py = (tile / a.tiles_x) * kTile + within / kTile;
  const int xs = 2 * px + (wave & 1);
  const int ys = 2 * py + (wave >> 1);
  if (xs >= a.nxs || ys >= a.nys) return;  // no block barrier: the wave just leaves
  float2* strip = wstrip + wave * CT;
  const float2* zslot = strip + CT - 2;
  if (lane < 2) strip[CT - 2 + lane] = make_float2(0.f, 0.f);
  const int zb0 = blockIdx.y * (kWave * ZPT);
  const int zbase = zb0 + lane;
  const float x = (float)(a.x0 + xs);
  const float y = (float)(a.y0 + ys);
  float acc[ZPT], accm[ZPT], zt[ZPT];
#pragma unroll
  for (int j = 0; j < ZPT; ++j) {
    acc[j] = 0.f;
    accm[j] = 0.f;
    zt[j] = (float)(a.z0 + zbase + j * kWave);
  }

  // lane s: the values of view s for this column
  WaveView wv{};
  if (lane < a.nproj) {
    const float* m = M.m[lane];
    const float wd = fmaf(m[8], x, fmaf(m[9], y, m[11]));
    const float Z = 1.0f / wd;
    const float U = fmaf(m[0], x, fmaf(m[1], y, m[3])) * Z;
    wv.Vb = fmaf(m[4], x, fmaf(m[5], y, m[7])) * Z;
    wv.dV = m[6] * Z;
    wv.Z2 = Z * Z;
    const float fu0 = floorf(U);
    wv.iu = (int)fu0;
    const bool uok = (unsigned)wv.iu <= (unsigned)(a.H - 2);
    const float fu = U - fu0;
    wv.w0 = uok ? (1.0f - fu) * wv.Z2 : 0.0f;
    wv.w1 = uok ? fu * wv.Z2 : 0.0f;
    // v over the top run, as the compute evaluates it (monotone in z)
    const float va = fmaf((float)(a.z0 + zb0), wv.dV, wv.Vb);
    const float vb = fmaf((float)(a.z0 + zb0 + kWave * ZPT - 1), wv.dV, wv.Vb);
    const int ilo = floor_int(fminf(va, vb)), ihi = floor_int(fmaxf(va, vb));
    const bool any = uok && ihi >= 0 && ilo <= a.W - 2;
    const bool all = uok && ilo >= 0 && ihi <= a.W - 2;
    wv.c0 = max(ilo, 0) & ~3;
    const int c1 = min(ihi + 1, a.W - 1);
    wv.quads = any ? (c1 - wv.c0 + 4) >> 2 : 0;
    const bool fits = 4 * wv.quads <= CT - 2 && wv.quads <= kWave * ws_quads<ZPT>();
    wv.flags = (any ? 1 : 0) | (all ? 2 : 0) | (fits ? 4 : 0);
  }

  WsPre<DT> pre;
  int f_n = lane_i(wv.flags, 0);
  if ((f_n & 5) == 5)
    ws_issue<DT, ws_quads<ZPT>()>(pre, view_rsrc(a, 0), a.W, lane_i(wv.iu, 0), lane_i(wv.c0, 0),
                                  lane_i(wv.quads, 0), lane);
  for (int s = 0; s < a.nproj; ++s) {
    const int f = f_n;
    const float Vb = lane_f(wv.Vb, s), dV = lane_f(wv.dV, s);
    float w0 = lane_f(wv.w0, s), w1 = lane_f(wv.w1, s);
    const int c0 = lane_i(wv.c0, s), quads = lane_i(wv.quads, s);
    float zof = 0.f;
    float2 so = make_float2(1.f, 0.f);
    if (DT != FDK_F32) {
      so = view_so(a, S, s);
      zof = (w0 + w1) * so.y;
    }
    if ((f & 5) == 5)
      ws_commit<DT, ws_quads<ZPT>()>(pre, strip, DT == FDK_F32 ? w0 : w0 * so.x,
                                     DT == FDK_F32 ? w1 : w1 * so.x, zof, quads, lane);
    if (s + 1 < a.nproj) {
      f_n = lane_i(wv.flags, s + 1);
      if ((f_n & 5) == 5)
        ws_issue<DT, ws_quads<ZPT>()>(pre, view_rsrc(a, s + 1), a.W, lane_i(wv.iu, s + 1),
                                      lane_i(wv.c0, s + 1), lane_i(wv.quads, s + 1), lane);
    }
    if (!(f & 1)) continue;
    if (f & 4) {
      const float2* base = strip - c0;
      if (f & 2)
        ws_compute<false, ZPT>(acc, accm, zt, Vb, dV, base, zslot, a.W);
      else
        ws_compute<true, ZPT>(acc, accm, zt, Vb, dV, base, zslot, a.W);
    } else {
      // strip too wide for the LDS slot: the taps straight from memory
      ColView c;
      c.Vb = Vb;
      c.dV = dV;
      c.w0 = w0;
      c.w1 = w1;
      c.Z2 = lane_f(wv.Z2, s);
      c.iu = lane_i(wv.iu, s);
      c.uok = true;
      float zm[ZPT];
#pragma unroll
      for (int j = 0; j < ZPT; ++j) zm[j] = (float)(a.nz - 1) - zt[j];
      compute_direct<DT, true, ZPT>(acc, accm, zt, zm, c, a,
                                    static_cast<const char*>(a.proj) + (size_t)s * a.view_bytes,
                                    so.x, so.y);
    }
  }
  column_write<true, ZPT>(a, xs, ys, zbase, acc, accm);
}

// General matrices (a z term in rows 0 or 2): every voxel does the full
// projective map.  Same layout and sampling as bp_column_kernel, ZPT = 1.
template <int DT, bool SYM>
__global__ __launch_bounds__(256) void bp_general_kernel(BpArgs a, BpMats M, BpScale S) {
  const unsigned lb = xcd_remap(blockIdx.x, gridDim.x);
  const int wave = threadIdx.x / kWave;
  const int lane = threadIdx.x % kWave;
  const int tile = lb / (kTile * kTile);
  const int within = lb % (kTile * kTile);
  const int px = (tile % a.tiles_x) * kTile + within % kTile;
  const int py = (tile / a.tiles_x) * kTile + within / kTile;
  const int xs = 2 * px + (wave & 1);
  const int ys = 2 * py + (wave >> 1);
  if (xs >= a.nxs || ys >= a.nys) return;
  const int zl = blockIdx.y * kWave + lane;
  if (zl >= a.zext) return;
  const float x = (float)(a.x0 + xs), y = (float)(a.y0 + ys);
  const float zs[2] = {(float)(a.z0 + zl), (float)(a.nz - 1 - (a.z0 + zl))};
  float acc[2] = {0.f, 0.f};
  for (int s = 0; s < a.nproj; ++s) {
    const float* m = M.m[s];
    for (int q = 0; q < (SYM ? 2 : 1); ++q) {
      const float z = zs[q];
      const float Z = 1.0f / fmaf(m[8], x, fmaf(m[9], y, fmaf(m[10], z, m[11])));
      const float U = fmaf(m[0], x, fmaf(m[1], y, fmaf(m[2], z, m[3]))) * Z;
      const float V = fmaf(m[4], x, fmaf(m[5], y, fmaf(m[6], z, m[7]))) * Z;
      const float fu0 = floorf(U);
      const int iu = (int)fu0;
      const bool uok = (unsigned)iu <= (unsigned)(a.H - 2);
      const float fu = U - fu0, Z2 = Z * Z;
      float w0 = uok ? (1.f - fu) * Z2 : 0.f, w1 = uok ? fu * Z2 : 0.f;
      const float2 so = view_so(a, S, s);
      if (DT != FDK_F32) { w0 *= so.x; w1 *= so.x; }
      const long long rowoff = (long long)s * a.view_stride + (long long)(uok ? iu : 0) * a.W;
      const void* rows = (DT == FDK_F32) ? (const void*)(static_cast<const float*>(a.proj) + rowoff)
                                         : (const void*)(static_cast<const unsigned short*>(a.proj) + rowoff);
      float c = sample_v<DT>(rows, V, a.W, w0, w1);
      if (DT != FDK_F32) c += (uok && V >= 0.f && V < (float)(a.W - 1)) ? so.y * Z2 : 0.f;
      acc[q] += c;
    }
  }
  float* col = a.vol + ((long long)xs * a.nys + ys) * (long long)a.nzz;
  col[zl] += acc[0];
  if (SYM) col[a.nzz - 1 - zl] += acc[1];
}

// [n][rows][cols] -> [n][cols][rows] through a 32x33 LDS tile.
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ src,
                                                        float* __restrict__ dst, int rows, int cols,
                                                        long long src_stride, long long dst_stride) {
  __shared__ float t[32][33];
  const int n = blockIdx.z;
  const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  const float* s = src + n * src_stride;
  float* d = dst + n * dst_stride;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int r = r0 + k, c = c0 + threadIdx.x;
    if (r < rows && c < cols) t[k][threadIdx.x] = s[(long long)r * cols + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int c = c0 + k, r = r0 + threadIdx.x;
    if (r < rows && c < cols) d[(long long)c * rows + r] = t[threadIdx.x][k];
  }
}

int launch_transpose(const float* src, float* dst, int rows, int cols, int n, long long ss,
                     long long ds, hipStream_t st) {
  dim3 grid((cols + 31) / 32, (rows + 31) / 32, n);
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, st, src, dst, rows, cols, ss, ds);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

// Volume [xs][ys][zz] -> [zz][ys][xs] for zz in [z_begin, z_begin + nzc).
__global__ __launch_bounds__(256) void volume_out_kernel(const float* __restrict__ vol,
                                                         float* __restrict__ out, int nxs, int nys,
                                                         int nzz, int z_begin, int nzc) {
  __shared__ float t[32][33];
  const int ys = blockIdx.z;
  const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int xs = x0 + k, zc = z0 + threadIdx.x;
    if (xs < nxs && zc < nzc)
      t[k][threadIdx.x] = vol[((long long)xs * nys + ys) * nzz + z_begin + zc];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += 8) {
    const int zc = z0 + k, xs = x0 + threadIdx.x;
    if (xs < nxs && zc < nzc) out[((long long)zc * nys + ys) * nxs + xs] = t[threadIdx.x][k];
  }
}

struct Box {
  int64_t x, y, z, w, h, d;
};

}  // namespace

struct fdk_bp {
  int proj_w = 0, proj_h = 0, proj_count = 0;
  int nx = 0, ny = 0, nz = 0;
  Box top{}, bot{};
  int nzz = 0;
  bool sym = false;
  float* vol = nullptr;
  float* staging = nullptr;  // proj_count views, host uploads land here
  float* staging_t = nullptr;
  size_t staging_elems = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int count = 0;
  float kernel_ms = 0.f;
  int device = 0;
};

namespace {

int fill_args(fdk_bp* h, int W, int H, long long view_stride, int esize, int nproj,
              const void* proj, BpArgs& a) {
  const long long view_bytes = (long long)W * H * esize;
  FDK_REQUIRE(view_bytes < (1LL << 31), "a view of %d x %d samples exceeds 2 GiB", W, H);
  a.vol = h->vol;
  a.proj = proj;
  a.view_stride = view_stride;
  a.W = W;
  a.H = H;
  a.nproj = nproj;
  a.x0 = (int)h->top.x;
  a.y0 = (int)h->top.y;
  a.z0 = (int)h->top.z;
  a.nxs = (int)h->top.w;
  a.nys = (int)h->top.h;
  a.zext = h->sym ? (int)h->top.d : h->nzz;
  a.nzz = h->nzz;
  a.nz = h->nz;
  a.tiles_x = ((a.nxs + 1) / 2 + kTile - 1) / kTile;
  a.view_bytes = view_stride * esize;
  a.view_bytes32 = (int)view_bytes;
  a.row_bytes = W * esize;
  a.buf16 = esize == 2 && (W % 2 == 0) && (view_stride % 2 == 0) &&
            (reinterpret_cast<uintptr_t>(proj) % 4 == 0);
  a.mm = nullptr;
  return FDK_OK;
}

// Voxels per lane along z: 8, or 4 when that pads the column's z extent to
// fewer wave runs (a 256-deep slab fills a ZPT 4 wave exactly and half of a
// ZPT 8 one).  FDK_BP_ZPT = 4, 8 or 16 forces a value.
int zpt_choice(int zext) {
  static const int forced = [] {
    const char* e = std::getenv("FDK_BP_ZPT");
    const int v = e ? std::atoi(e) : 0;
    return (v == 4 || v == 8 || v == 16) ? v : 0;
  }();
  if (forced) return forced;
  const long long pad8 = (zext + 511LL) / 512 * 512, pad4 = (zext + 255LL) / 256 * 256;
  return pad4 < pad8 ? 4 : 8;
}

// Rows R and columns CT of the detector strip one block of bp_lds_kernel
// stages per view, bounded from the matrices at the sub-volume's corners: the
// u spread of a 2x2 column tile (one voxel step in x and in y), and the v run
// of one wave (64 * ZPT voxels along z) plus the tile's v spread, with margins
// for floor rounding and 4-column alignment.  false when a corner is behind
// the source.
bool stage_dims(const BpArgs& a, const BpMats& M, int zpt, int* R, int* CT) {
  const double xs[2] = {(double)a.x0, (double)(a.x0 + a.nxs - 1)};
  const double ys[2] = {(double)a.y0, (double)(a.y0 + a.nys - 1)};
  const double zs[4] = {(double)a.z0, (double)(a.z0 + a.zext - 1), (double)(a.nz - 1 - a.z0),
                        (double)(a.nz - 1 - (a.z0 + a.zext - 1))};
  double su = 0.0, dvz = 0.0, sv = 0.0;
  for (int s = 0; s < a.nproj; ++s) {
    const float* m = M.m[s];
    for (double x : xs)
      for (double y : ys) {
        const double w = m[8] * x + m[9] * y + m[11];
        if (!(w > 0.0)) return false;
        const double u = (m[0] * x + m[1] * y + m[3]) / w;
        su = std::max(su, std::fabs((m[0] - u * m[8]) / w) + std::fabs((m[1] - u * m[9]) / w));
        dvz = std::max(dvz, std::fabs(m[6] / w));
        for (double z : zs) {
          const double v = (m[4] * x + m[5] * y + m[6] * z + m[7]) / w;
          sv = std::max(sv, std::fabs((m[4] - v * m[8]) / w) + std::fabs((m[5] - v * m[9]) / w));
        }
      }
  }
  *R = (int)std::ceil(1.25 * su) + 3;
  const double span = dvz * (kWave * zpt - 1) + 1.25 * sv;
  *CT = (((int)std::ceil(span) + 11) & ~3) + 4;  // + the zero slot's spare columns
  return true;
}

// Kernel choice: FDK_BP_KERNEL=v2 selects the direct-gather bp_column_kernel,
// v3 the LDS-staged bp_lds_kernel, v4 bp_lds_pf_kernel where its register
// prefetch covers the strip; by default bp_wave_kernel runs for mirror-symmetric
// slab pairs whose strips fit, then the v4 / v3 / v2 chain.
int kernel_choice() {
  static const int k = [] {
    const char* e = std::getenv("FDK_BP_KERNEL");
    if (e && std::strcmp(e, "v2") == 0) return 2;
    if (e && std::strcmp(e, "v3") == 0) return 3;
    if (e && std::strcmp(e, "v4") == 0) return 4;
    return 5;
  }();
  return k;
}

// v(z) + v(nz-1-z) = W-1 for every column of the sub-volume and every view, to
// 1e-3 pixel (FP64 evaluation of the FP32 matrices at the corners; the error is
// affine in x and y): the mirror voxel's sample position is then W-1-v.
bool mirror_symmetric(const BpArgs& a, const BpMats& M) {
  const double xs[2] = {(double)a.x0, (double)(a.x0 + a.nxs - 1)};
  const double ys[2] = {(double)a.y0, (double)(a.y0 + a.nys - 1)};
  for (int s = 0; s < a.nproj; ++s) {
    const float* m = M.m[s];
    for (double x : xs)
      for (double y : ys) {
        const double w = (double)m[8] * x + (double)m[9] * y + (double)m[11];
        if (!(w > 0.0)) return false;
        const double e = (2.0 * ((double)m[4] * x + (double)m[5] * y + (double)m[7]) +
                          (double)m[6] * (a.nz - 1)) / w - (a.W - 1);
        if (!(std::fabs(e) <= 1e-3)) return false;
      }
  }
  return true;
}

// Strip slot CT (float2 per column, plus the two-entry zero slot) of
// bp_wave_kernel: the v extent of one wave's z-run (64 * ZPT voxels) over the
// views and the sub-volume's corners, with room for floor rounding and the
// 4-column alignment of its start.  0 when a corner is behind the source.
int wave_strip_ct(const BpArgs& a, const BpMats& M, int zpt) {
  const double xs[2] = {(double)a.x0, (double)(a.x0 + a.nxs - 1)};
  const double ys[2] = {(double)a.y0, (double)(a.y0 + a.nys - 1)};
  double dvz = 0.0;
  for (int s = 0; s < a.nproj; ++s) {
    const float* m = M.m[s];
    for (double x : xs)
      for (double y : ys) {
        const double w = (double)m[8] * x + (double)m[9] * y + (double)m[11];
        if (!(w > 0.0)) return 0;
        dvz = std::max(dvz, std::fabs((double)m[6] / w));
      }
  }
  const double span = dvz * (kWave * zpt - 1);
  const int quads = ((int)std::ceil(span) + 9) / 4 + 1;
  return 4 * quads + 2;
}

template <int DT, int ZPT>
bool launch_lds(fdk_bp* h, const BpArgs& a, const BpMats& M, const BpScale& S, unsigned nblk) {
  if (kernel_choice() < 3 || a.W % 4 != 0) return false;
  if (DT != FDK_F32 && ((reinterpret_cast<uintptr_t>(a.proj) % 8) != 0 || a.view_stride % 4 != 0))
    return false;
  int R = 0, CT = 0;
  if (!stage_dims(a, M, ZPT, &R, &CT)) return false;
  const size_t lds = (size_t)R * CT * sizeof(float) * (h->sym ? 2 : 1);
  if (lds > 64 * 1024) return false;
  dim3 grid(nblk, (a.zext + kWave * ZPT - 1) / (kWave * ZPT));
  const bool pf = DT == FDK_F32 && kernel_choice() >= 4 && (h->sym ? 2 * R : R) <= 4 * kPfRows &&
                  CT <= 4 * kWave * kPfQuads;
  if (pf) {
    if (h->sym)
      hipLaunchKernelGGL((bp_lds_pf_kernel<DT, true, ZPT>), grid, dim3(256), lds, h->stream, a, M, S,
                         R, CT);
    else
      hipLaunchKernelGGL((bp_lds_pf_kernel<DT, false, ZPT>), grid, dim3(256), lds, h->stream, a, M,
                         S, R, CT);
  } else if (h->sym) {
    hipLaunchKernelGGL((bp_lds_kernel<DT, true, ZPT>), grid, dim3(256), lds, h->stream, a, M, S, R, CT);
  } else {
    hipLaunchKernelGGL((bp_lds_kernel<DT, false, ZPT>), grid, dim3(256), lds, h->stream, a, M, S, R, CT);
  }
  return true;
}

template <int DT, int ZPT>
bool launch_wave(fdk_bp* h, const BpArgs& a, const BpMats& M, const BpScale& S, unsigned nblk) {
  if (kernel_choice() < 5 || !h->sym || a.W % 4 != 0) return false;
  if (DT != FDK_F32 && ((reinterpret_cast<uintptr_t>(a.proj) % 8) != 0 || a.view_stride % 4 != 0))
    return false;
  if (!mirror_symmetric(a, M)) return false;
  const int CT = wave_strip_ct(a, M, ZPT);
  if (CT <= 0 || (CT - 2) / 4 > kWave * ws_quads<ZPT>()) return false;
  const size_t lds = (size_t)4 * CT * sizeof(float2);
  if (lds > 64 * 1024) return false;
  dim3 grid(nblk, (a.zext + kWave * ZPT - 1) / (kWave * ZPT));
  hipLaunchKernelGGL((bp_wave_kernel<DT, ZPT>), grid, dim3(256), lds, h->stream, a, M, S, CT);
  return true;
}

template <int DT, int ZPT>
void launch_column(fdk_bp* h, const BpArgs& a, const BpMats& M, const BpScale& S, unsigned nblk) {
  if (launch_wave<DT, ZPT>(h, a, M, S, nblk)) return;
  if (launch_lds<DT, ZPT>(h, a, M, S, nblk)) return;
  dim3 grid(nblk, (a.zext + kWave * ZPT - 1) / (kWave * ZPT));
  if (h->sym) hipLaunchKernelGGL((bp_column_kernel<DT, true, ZPT>), grid, dim3(256), 0, h->stream, a, M, S);
  else hipLaunchKernelGGL((bp_column_kernel<DT, false, ZPT>), grid, dim3(256), 0, h->stream, a, M, S);
}

template <int DT>
int launch_bp_dt(fdk_bp* h, const BpArgs& a, const BpMats& M, const BpScale& S, bool general) {
  const int tiles_y = ((a.nys + 1) / 2 + kTile - 1) / kTile;
  const unsigned nblk = (unsigned)(a.tiles_x * tiles_y * kTile * kTile);
  if (general) {
    dim3 grid(nblk, (a.zext + kWave - 1) / kWave);
    if (h->sym) hipLaunchKernelGGL((bp_general_kernel<DT, true>), grid, dim3(256), 0, h->stream, a, M, S);
    else hipLaunchKernelGGL((bp_general_kernel<DT, false>), grid, dim3(256), 0, h->stream, a, M, S);
  } else {
    switch (zpt_choice(a.zext)) {
      case 4: launch_column<DT, 4>(h, a, M, S, nblk); break;
      case 16: launch_column<DT, 16>(h, a, M, S, nblk); break;
      default: launch_column<DT, 8>(h, a, M, S, nblk); break;
    }
  }
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

int launch_bp(fdk_bp* h, int dtype, const void* proj, int W, int H, long long view_stride,
              const float* mats, const float* scale_offset, int nproj,
              const unsigned* dev_minmax = nullptr) {
  FDK_REQUIRE(nproj >= 0 && nproj <= kMaxProj, "nProj %d exceeds %d views per pass", nproj, kMaxProj);
  FDK_REQUIRE(mats != nullptr || nproj == 0, "matrices are NULL");
  if (nproj == 0) return FDK_OK;
  FDK_REQUIRE(W >= 2 && H >= 2, "views must be at least 2 x 2");
  BpMats M;
  BpScale S;
  std::memset(&M, 0, sizeof(M));
  std::memset(&S, 0, sizeof(S));
  bool general = false;
  for (int s = 0; s < nproj; ++s) {
    for (int e = 0; e < 12; ++e) M.m[s][e] = mats[12 * s + e];
    if (M.m[s][2] != 0.0f || M.m[s][10] != 0.0f) general = true;
    S.so[s][0] = scale_offset ? scale_offset[2 * s] : 1.0f;
    S.so[s][1] = scale_offset ? scale_offset[2 * s + 1] : 0.0f;
  }
  BpArgs a;
  if (int r = fill_args(h, W, H, view_stride, dtype == FDK_F32 ? 4 : 2, nproj, proj, a)) return r;
  a.mm = dtype == FDK_F32 ? nullptr : dev_minmax;
  h->count += nproj;
  switch (dtype) {
    case FDK_F32: return launch_bp_dt<FDK_F32>(h, a, M, S, general);
    case FDK_F16: return launch_bp_dt<FDK_F16>(h, a, M, S, general);
    case FDK_BF16: return launch_bp_dt<FDK_BF16>(h, a, M, S, general);
    default: return fdk::fail(FDK_ERR_INVALID, "unknown dtype %d", dtype);
  }
}

int check_box(const int64_t* b, int nx, int ny, int nz, const char* name) {
  FDK_REQUIRE(b != nullptr, "%s box is NULL", name);
  FDK_REQUIRE(b[0] >= 0 && b[1] >= 0 && b[2] >= 0, "%s box origin is negative", name);
  FDK_REQUIRE(b[3] >= 0 && b[4] >= 0 && b[5] >= 0, "%s box size is negative", name);
  FDK_REQUIRE(b[0] + b[3] <= nx && b[1] + b[4] <= ny && b[2] + b[5] <= nz,
              "%s box exceeds the %d x %d x %d volume", name, nx, ny, nz);
  return FDK_OK;
}

}  // namespace

extern "C" int fdk_bp_create(int proj_width, int proj_height, int proj_count, int nx, int ny,
                             int nz, const int64_t box_top[6], const int64_t box_bottom[6],
                             fdk_bp_t* out) {
  FDK_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  FDK_REQUIRE(proj_width >= 2 && proj_height >= 2 && proj_count >= 1, "bad projection size");
  FDK_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad volume size");
  if (int r = check_box(box_top, nx, ny, nz, "top")) return r;
  const int64_t zero_box[6] = {0, 0, 0, 0, 0, 0};
  const int64_t* bb = box_bottom ? box_bottom : zero_box;
  if (bb[5] > 0) {
    if (int r = check_box(bb, nx, ny, nz, "bottom")) return r;
    // backprojectionFDK.cu:173-175: the bottom slab mirrors the top one.
    FDK_REQUIRE(box_top[2] + bb[2] + bb[5] - 1 == nz - 1,
                "bottom slab must mirror the top slab: top.z + bottom.maxZ == nz - 1");
    FDK_REQUIRE(bb[5] == box_top[5] && bb[0] == box_top[0] && bb[1] == box_top[1] &&
                    bb[3] == box_top[3] && bb[4] == box_top[4],
                "bottom slab must have the top slab's x/y extent and depth");
  }
  FDK_REQUIRE(box_top[3] > 0 && box_top[4] > 0 && box_top[5] > 0, "top box is empty");
  fdk_bp* h = new fdk_bp();
  h->proj_w = proj_width;
  h->proj_h = proj_height;
  h->proj_count = proj_count;
  h->nx = nx; h->ny = ny; h->nz = nz;
  h->top = Box{box_top[0], box_top[1], box_top[2], box_top[3], box_top[4], box_top[5]};
  h->bot = Box{bb[0], bb[1], bb[2], bb[3], bb[4], bb[5]};
  h->sym = bb[5] > 0;
  h->nzz = (int)(h->top.d + (h->sym ? h->bot.d : 0));
  (void)hipGetDevice(&h->device);
  const size_t nvox = (size_t)h->top.w * h->top.h * h->nzz;
  hipError_t e = hipMalloc(&h->vol, nvox * sizeof(float));
  if (e == hipSuccess) e = hipMemset(h->vol, 0, nvox * sizeof(float));
  h->staging_elems = (size_t)proj_width * proj_height * proj_count;
  if (e == hipSuccess) e = hipMalloc(&h->staging, h->staging_elems * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) h->own_stream = true;
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e != hipSuccess) {
    fdk_bp_destroy(h);
    return fdk::fail(e == hipErrorOutOfMemory ? FDK_ERR_OOM : FDK_ERR_HIP,
                     "fdk_bp_create: %s (volume %zu voxels)", hipGetErrorString(e), nvox);
  }
  *out = h;
  return FDK_OK;
}

extern "C" int fdk_bp_destroy(fdk_bp_t h) {
  if (!h) return FDK_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->vol) (void)hipFree(h->vol);
  if (h->staging) (void)hipFree(h->staging);
  if (h->staging_t) (void)hipFree(h->staging_t);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FDK_OK;
}

extern "C" int fdk_bp_set_stream(fdk_bp_t h, void* stream) {
  FDK_REQUIRE(h, "handle is NULL");
  FDK_HIP(hipStreamSynchronize(h->stream));
  if (h->own_stream) (void)hipStreamDestroy(h->stream);
  h->stream = static_cast<hipStream_t>(stream);
  h->own_stream = false;
  return FDK_OK;
}

extern "C" int fdk_bp_batch(fdk_bp_t h, int transpose, const float* host_proj, int width,
                            int height, const float* mats, int nproj) {
  FDK_REQUIRE(h, "handle is NULL");
  FDK_REQUIRE(transpose == 0 || transpose == 1, "transpose must be 0 or 1");
  FDK_REQUIRE(host_proj != nullptr || nproj == 0, "projections are NULL");
  FDK_REQUIRE(nproj <= kMaxProj && nproj <= h->proj_count, "nProj %d exceeds the %d-view batch",
              nproj, std::min(kMaxProj, h->proj_count));
  if (nproj <= 0) return FDK_OK;
  // Stored layout the kernel reads: W = nv columns (v fastest) by H = nu rows.
  const int W = transpose ? width : height;
  const int H = transpose ? height : width;
  // The constructor's projWidth x projHeight x projCount sizes the staging
  // buffer (backprojectionFDK.cu:181); a call may use any view that fits it.
  FDK_REQUIRE((size_t)width * height <= (size_t)h->proj_w * h->proj_h,
              "view %d x %d exceeds the handle's %d x %d", width, height, h->proj_w, h->proj_h);
  const size_t view = (size_t)width * height;
  if (transpose) {
    FDK_HIP(hipMemcpyAsync(h->staging, host_proj, view * nproj * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
  } else {
    if (!h->staging_t) FDK_HIP(hipMalloc(&h->staging_t, h->staging_elems * sizeof(float)));
    FDK_HIP(hipMemcpyAsync(h->staging_t, host_proj, view * nproj * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
    if (int r = launch_transpose(h->staging_t, h->staging, height, width, nproj, view, view, h->stream))
      return r;
  }
  FDK_HIP(hipEventRecord(h->ev0, h->stream));
  if (int r = launch_bp(h, FDK_F32, h->staging, W, H, (long long)view, mats, nullptr, nproj)) return r;
  FDK_HIP(hipEventRecord(h->ev1, h->stream));
  FDK_HIP(hipStreamSynchronize(h->stream));
  float ms = 0.f;
  FDK_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->kernel_ms += ms;
  return FDK_OK;
}

extern "C" int fdk_bp_batch_device(fdk_bp_t h, int transpose, const void* dev_proj, int dtype,
                                   int width, int height, int64_t view_stride, const float* mats,
                                   const float* scale_offset, int nproj) {
  FDK_REQUIRE(h, "handle is NULL");
  FDK_REQUIRE(dev_proj != nullptr || nproj == 0, "projections are NULL");
  FDK_REQUIRE(transpose == 0 || transpose == 1, "transpose must be 0 or 1");
  FDK_REQUIRE(dtype == FDK_F32 || transpose == 1, "FP16/BF16 views must be stored transposed");
  if (nproj <= 0) return FDK_OK;
  FDK_REQUIRE(nproj <= kMaxProj, "nProj %d exceeds %d views per pass", nproj, kMaxProj);
  const int W = transpose ? width : height;
  const int H = transpose ? height : width;
  FDK_REQUIRE(view_stride >= (int64_t)width * height, "view_stride %lld is smaller than a view",
              (long long)view_stride);
  if (transpose)
    return launch_bp(h, dtype, dev_proj, W, H, (long long)view_stride, mats, scale_offset, nproj);
  FDK_REQUIRE(nproj <= h->proj_count, "nProj exceeds the staging batch");
  const size_t view = (size_t)width * height;
  if (int r = launch_transpose(static_cast<const float*>(dev_proj), h->staging, height, width,
                               nproj, view_stride, view, h->stream))
    return r;
  return launch_bp(h, FDK_F32, h->staging, W, H, (long long)view, mats, nullptr, nproj);
}

extern "C" int fdk_bp_batch_device_minmax(fdk_bp_t h, const void* dev_proj, int dtype, int width,
                                          int height, int64_t view_stride, const float* mats,
                                          const unsigned* dev_minmax, int nproj) {
  FDK_REQUIRE(h, "handle is NULL");
  FDK_REQUIRE(dtype == FDK_F16 || dtype == FDK_BF16, "dtype %d is not a 16-bit storage type", dtype);
  FDK_REQUIRE(dev_proj != nullptr || nproj == 0, "projections are NULL");
  FDK_REQUIRE(dev_minmax != nullptr || nproj == 0, "device min/max words are NULL");
  if (nproj <= 0) return FDK_OK;
  FDK_REQUIRE(nproj <= kMaxProj, "nProj %d exceeds %d views per pass", nproj, kMaxProj);
  FDK_REQUIRE(view_stride >= (int64_t)width * height, "view_stride %lld is smaller than a view",
              (long long)view_stride);
  return launch_bp(h, dtype, dev_proj, width, height, (long long)view_stride, mats, nullptr, nproj,
                   dev_minmax);
}

extern "C" int fdk_bp_sync(fdk_bp_t h) {
  FDK_REQUIRE(h, "handle is NULL");
  FDK_HIP(hipStreamSynchronize(h->stream));
  return FDK_OK;
}

extern "C" int fdk_bp_count(fdk_bp_t h) { return h ? h->count : fdk::fail(FDK_ERR_INVALID, "NULL"); }

extern "C" float fdk_bp_kernel_ms(fdk_bp_t h) { return h ? h->kernel_ms : 0.f; }

extern "C" int fdk_bp_zero(fdk_bp_t h) {
  FDK_REQUIRE(h, "handle is NULL");
  FDK_HIP(hipMemsetAsync(h->vol, 0, (size_t)h->top.w * h->top.h * h->nzz * sizeof(float), h->stream));
  h->count = 0;
  h->kernel_ms = 0.f;
  return FDK_OK;
}

namespace {

int volume_out(fdk_bp* h, float* dst, size_t count, bool device_dst) {
  const size_t plane = (size_t)h->top.w * h->top.h;
  FDK_REQUIRE(dst != nullptr, "destination is NULL");
  FDK_REQUIRE(count >= plane * h->nzz, "destination holds %zu floats, need %zu", count,
              plane * h->nzz);
  const int nxs = (int)h->top.w, nys = (int)h->top.h;
  if (device_dst) {
    dim3 grid((nxs + 31) / 32, (h->nzz + 31) / 32, nys);
    hipLaunchKernelGGL(volume_out_kernel, grid, dim3(32, 8), 0, h->stream, h->vol, dst, nxs, nys,
                       h->nzz, 0, h->nzz);
    FDK_HIP(hipGetLastError());
    FDK_HIP(hipStreamSynchronize(h->stream));
    return FDK_OK;
  }
  // Host: transpose through a bounded device buffer, chunk by chunk of z planes
  // (the reference halves its staging depth until it fits, backprojectionFDK.cu:206-212).
  int chunk = (int)std::max<size_t>(1, std::min<size_t>(h->nzz, (256u << 20) / (plane * 4)));
  float* tmp = nullptr;
  FDK_HIP(hipMalloc(&tmp, plane * chunk * sizeof(float)));
  int rc = FDK_OK;
  for (int z = 0; z < h->nzz && rc == FDK_OK; z += chunk) {
    const int nzc = std::min(chunk, h->nzz - z);
    dim3 grid((nxs + 31) / 32, (nzc + 31) / 32, nys);
    hipLaunchKernelGGL(volume_out_kernel, grid, dim3(32, 8), 0, h->stream, h->vol, tmp, nxs, nys,
                       h->nzz, z, nzc);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(dst + plane * z, tmp, plane * nzc * sizeof(float), hipMemcpyDeviceToHost,
                         h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = fdk::fail(FDK_ERR_HIP, "GetVolumeData: %s", hipGetErrorString(e));
  }
  (void)hipFree(tmp);
  return rc;
}

}  // namespace

extern "C" int fdk_bp_get_volume(fdk_bp_t h, float* host_dst, size_t count) {
  FDK_REQUIRE(h, "handle is NULL");
  return volume_out(h, host_dst, count, false);
}

extern "C" int fdk_bp_get_volume_device(fdk_bp_t h, float* dev_dst, size_t count) {
  FDK_REQUIRE(h, "handle is NULL");
  return volume_out(h, dev_dst, count, true);
}

extern "C" int fdk_bp_volume_ptr(fdk_bp_t h, void** dev_ptr, int64_t dims[3]) {
  FDK_REQUIRE(h && dev_ptr && dims, "NULL argument");
  *dev_ptr = h->vol;
  dims[0] = h->top.w;
  dims[1] = h->top.h;
  dims[2] = h->nzz;
  return FDK_OK;
}

// ------------------------------------------------------------------ RTK path
namespace {

struct RtkArgs {
  const float* vin;
  float* vout;
  const float* proj;  // [s][v][u], u fastest (W = nu, H = nv)
  int W, H, nproj;
  int nx, ny, nz;
};

// kernel_fdk_3Dgrid: one thread per voxel, x fastest across lanes, full
// projective map per voxel, vol_out = vol_in + sum (rtk...cu:156-191).
__global__ __launch_bounds__(256) void rtk_kernel(RtkArgs a, BpMats M) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= a.nx || j >= a.ny || k >= a.nz) return;
  const float x = (float)i, y = (float)j, z = (float)k;
  float acc = 0.f;
  for (int s = 0; s < a.nproj; ++s) {
    const float* m = M.m[s];
    const float Z = 1.0f / fmaf(m[8], x, fmaf(m[9], y, fmaf(m[10], z, m[11])));
    const float U = fmaf(m[0], x, fmaf(m[1], y, fmaf(m[2], z, m[3]))) * Z;
    const float V = fmaf(m[4], x, fmaf(m[5], y, fmaf(m[6], z, m[7]))) * Z;
    if (U >= 0.f && U < (float)(a.W - 1) && V >= 0.f && V < (float)(a.H - 1)) {
      const float fu0 = floorf(U), fv0 = floorf(V);
      const float fu = U - fu0, fv = V - fv0;
      const float* p = a.proj + (long long)s * a.W * a.H + (long long)(int)fv0 * a.W + (int)fu0;
      const float r0 = fmaf(fu, p[1] - p[0], p[0]);
      const float r1 = fmaf(fu, p[a.W + 1] - p[a.W], p[a.W]);
      acc = fmaf(fmaf(fv, r1 - r0, r0), Z * Z, acc);
    }
  }
  const long long idx = ((long long)k * a.ny + j) * a.nx + i;
  a.vout[idx] = a.vin[idx] + acc;
}

int launch_rtk(const int proj_size[3], const int vol_size[3], const float* mats,
               const float* vin, float* vout, const float* dproj, hipStream_t st) {
  FDK_REQUIRE(proj_size[2] >= 0 && proj_size[2] <= kMaxProj, "nProj %d exceeds %d", proj_size[2], kMaxProj);
  FDK_REQUIRE(proj_size[0] >= 2 && proj_size[1] >= 2, "views must be at least 2 x 2");
  FDK_REQUIRE(vol_size[0] > 0 && vol_size[1] > 0 && vol_size[2] > 0, "bad volume size");
  BpMats M;
  std::memset(&M, 0, sizeof(M));
  for (int s = 0; s < proj_size[2]; ++s)
    for (int e = 0; e < 12; ++e) M.m[s][e] = mats[12 * s + e];
  RtkArgs a{vin, vout, dproj, proj_size[0], proj_size[1], proj_size[2], vol_size[0], vol_size[1], vol_size[2]};
  dim3 grid((vol_size[0] + 63) / 64, (vol_size[1] + 3) / 4, vol_size[2]);
  hipLaunchKernelGGL(rtk_kernel, grid, dim3(64, 4), 0, st, a, M);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

}  // namespace

struct fdk_rtk {
  int proj_w, proj_h, proj_count, nx, ny, nz;
  float* vin = nullptr;
  float* vout = nullptr;
  float* staging = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float kernel_ms = 0.f;
};

extern "C" int fdk_rtk_create(int proj_width, int proj_height, int proj_count, int nx, int ny,
                              int nz, fdk_rtk_t* out) {
  FDK_REQUIRE(out, "out is NULL");
  *out = nullptr;
  FDK_REQUIRE(proj_width >= 2 && proj_height >= 2 && proj_count >= 1, "bad projection size");
  FDK_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad volume size");
  fdk_rtk* h = new fdk_rtk{proj_width, proj_height, proj_count, nx, ny, nz};
  const size_t nvox = (size_t)nx * ny * nz;
  hipError_t e = hipMalloc(&h->vin, nvox * 4);
  if (e == hipSuccess) e = hipMalloc(&h->vout, nvox * 4);
  if (e == hipSuccess) e = hipMemset(h->vin, 0, nvox * 4);
  if (e == hipSuccess) e = hipMemset(h->vout, 0, nvox * 4);
  if (e == hipSuccess) e = hipMalloc(&h->staging, (size_t)proj_width * proj_height * proj_count * 4);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e != hipSuccess) {
    fdk_rtk_destroy(h);
    return fdk::fail(e == hipErrorOutOfMemory ? FDK_ERR_OOM : FDK_ERR_HIP, "fdk_rtk_create: %s",
                     hipGetErrorString(e));
  }
  *out = h;
  return FDK_OK;
}

extern "C" int fdk_rtk_destroy(fdk_rtk_t h) {
  if (!h) return FDK_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->vin) (void)hipFree(h->vin);
  if (h->vout) (void)hipFree(h->vout);
  if (h->staging) (void)hipFree(h->staging);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return FDK_OK;
}

extern "C" int fdk_rtk_bp(fdk_rtk_t h, int transpose, const float* host_proj, int width,
                          int height, const float* mats, int nproj) {
  FDK_REQUIRE(h, "handle is NULL");
  FDK_REQUIRE(transpose == 0, "BackProjectionRTK takes untransposed views (TRANSPOSE == 0)");
  FDK_REQUIRE(width == h->proj_w && height == h->proj_h, "view size does not match the handle");
  FDK_REQUIRE(nproj >= 0 && nproj <= std::min(kMaxProj, h->proj_count), "bad nProj %d", nproj);
  if (nproj == 0) return FDK_OK;
  FDK_HIP(hipMemcpyAsync(h->staging, host_proj, (size_t)width * height * nproj * 4,
                         hipMemcpyHostToDevice, h->stream));
  std::swap(h->vin, h->vout);  // SwapInOutBuffer (rtk...cu:391)
  const int ps[3] = {width, height, nproj};
  const int vs[3] = {h->nx, h->ny, h->nz};
  FDK_HIP(hipEventRecord(h->ev0, h->stream));
  if (int r = launch_rtk(ps, vs, mats, h->vin, h->vout, h->staging, h->stream)) return r;
  FDK_HIP(hipEventRecord(h->ev1, h->stream));
  FDK_HIP(hipStreamSynchronize(h->stream));
  float ms = 0.f;
  FDK_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->kernel_ms += ms;
  return FDK_OK;
}

extern "C" int fdk_rtk_get_volume(fdk_rtk_t h, float* host_dst, size_t count) {
  FDK_REQUIRE(h && host_dst, "NULL argument");
  const size_t nvox = (size_t)h->nx * h->ny * h->nz;
  FDK_REQUIRE(count >= nvox, "destination holds %zu floats, need %zu", count, nvox);
  FDK_HIP(hipMemcpyAsync(host_dst, h->vout, nvox * 4, hipMemcpyDeviceToHost, h->stream));
  FDK_HIP(hipStreamSynchronize(h->stream));
  return FDK_OK;
}

extern "C" float fdk_rtk_kernel_ms(fdk_rtk_t h) { return h ? h->kernel_ms : 0.f; }

extern "C" int fdk_reconstruct_conebeam(const int proj_size[3], const int vol_size[3],
                                        const float* mats, const float* dev_vol_in,
                                        float* dev_vol_out, const float* proj, int proj_on_device,
                                        void* stream) {
  FDK_REQUIRE(proj_size && vol_size && mats && dev_vol_in && dev_vol_out && proj, "NULL argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (proj_on_device) return launch_rtk(proj_size, vol_size, mats, dev_vol_in, dev_vol_out, proj, st);
  const size_t bytes = (size_t)proj_size[0] * proj_size[1] * std::max(proj_size[2], 0) * 4;
  float* tmp = nullptr;
  FDK_HIP(hipMalloc(&tmp, std::max<size_t>(bytes, 4)));
  hipError_t e = hipMemcpyAsync(tmp, proj, bytes, hipMemcpyHostToDevice, st);
  int r = e == hipSuccess ? launch_rtk(proj_size, vol_size, mats, dev_vol_in, dev_vol_out, tmp, st)
                          : fdk::fail(FDK_ERR_HIP, "upload: %s", hipGetErrorString(e));
  (void)hipStreamSynchronize(st);
  (void)hipFree(tmp);
  return r;
}
