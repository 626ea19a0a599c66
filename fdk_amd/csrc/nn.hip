// Kernels of the learned sparse-view correction: a 3-D residual U-Net on
// BF16 MFMA (gfx950 v_mfma_f32_16x16x32_bf16), FP32 accumulation.  No reference
// counterpart: the reference has no learned component.
//
// Activations are channels-last, [n][z][y][x][c] BF16.  A 3x3x3 'same'
// convolution is an implicit GEMM: M = output voxels, N = output channels,
// K = 27 taps x input channels.  With channels innermost, the 8 consecutive K
// values a lane feeds one MFMA are 16 contiguous bytes of one voxel (A) or of
// one weight row (B), so operands come straight from L1/L2 as 16-byte loads.
//
// MFMA 16x16x32 BF16 lane layout (checked by fdk_nn_mfma_probe's test):
//   A (16 x 32): lane l holds A[l % 16][8 (l / 16) + j], j = 0..7
//   B (32 x 16): lane l holds B[8 (l / 16) + j][l % 16]
//   D (16 x 16): lane l holds D[4 (l / 16) + i][l % 16], i = 0..3
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>

#include "common.hpp"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

__device__ inline float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ inline u16 f2bf(float f) {  // round to nearest even
  const unsigned b = __float_as_uint(f);
  if ((b & 0x7f800000u) == 0x7f800000u) return (u16)(b >> 16);  // inf / nan
  return (u16)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

// Activation of a convolution's output: 0 none, 1 ReLU, 2 leaky ReLU (slope
// kLeak for negative inputs).
constexpr float kLeak = 0.01f;
__device__ inline float activate(float v, int mode) {
  return mode == 1 ? fmaxf(v, 0.0f) : mode == 2 ? (v > 0.0f ? v : kLeak * v) : v;
}

__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 load8(const u16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline bf16x8 zero8() {
  bf16x8 z;
  for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
  return z;
}

struct Conv {
  const u16* x;      // [n][d][h][w][cin], or [n][d][h][w][cin - cup] with `up`
  const u16* wt;     // [27][cout][cin] (cin == 1: [cout][32] taps, zero padded)
  const float* bias; // [cout] or null
  const float* resid;// [n][d][h][w][cout] FP32 added to the output, or null
  void* y;           // [n][d][h][w][cout], BF16 or FP32
  int n, d, h, w, cin, cout;
  int relu, out_f32;
  // Decoder form (conv3d_fwd4_kernel only): the input is the channel
  // concatenation [nearest 2x upsampling of up ([n][d/2][h/2][w/2][cup]), x],
  // read from the two tensors without forming it.  up == null: plain input.
  const u16* up = nullptr;
  int cup = 0;
  // conv3d_fwd4_kernel computes output channels [co0, co0 + 16 NT) of the cout.
  int co0 = 0;
};

template <int NT>
__device__ inline void conv_epilogue(const Conv& a, f32x4 (&acc)[2][NT], int nzy, int x0,
                                     int lane) {
  const int m = lane % 16, kq = lane / 16;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + mt * 16 + 4 * kq + i;
        const int co = nt * 16 + m;
        const long long o = ((long long)nzy * a.w + x) * a.cout + co;
        float v = acc[mt][nt][i];
        if (a.bias) v += a.bias[co];
        v = activate(v, a.relu);
        if (a.resid) v += a.resid[o];
        if (a.out_f32) static_cast<float*>(a.y)[o] = v;
        else static_cast<u16*>(a.y)[o] = f2bf(v);
      }
}

// One wave: 32 output voxels of one x-row (two 16-voxel M tiles) x all cout
// (NT tiles of 16).  Block: 4 waves = 4 consecutive rows.  cin % 32 == 0.
template <int NT>
__global__ __launch_bounds__(256) void conv3d_fwd_kernel(Conv a) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int x0 = blockIdx.x * 32;
  const int y = blockIdx.y * 4 + wave;
  const int nz = blockIdx.z;  // n * d + z
  const int z = nz % a.d, n = nz / a.d;
  if (y >= a.h) return;
  f32x4 acc[2][NT];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int zz = z + dz, yy = y + dy;
    const bool zy = zz >= 0 && zz < a.d && yy >= 0 && yy < a.h;
    const long long rowbase = (((long long)n * a.d + zz) * a.h + yy) * a.w;
    for (int c0 = 0; c0 < a.cin; c0 += 32) {
      bf16x8 af[2];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        const int xx = x0 + mt * 16 + m + dx;
        af[mt] = (zy && xx >= 0 && xx < a.w) ? load8(a.x + (rowbase + xx) * a.cin + c0 + kq * 8)
                                            : zero8();
      }
      const u16* wb = a.wt + (long long)tap * a.cout * a.cin + c0 + kq * 8;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const bf16x8 bf = load8(wb + (long long)(nt * 16 + m) * a.cin);
        acc[0][nt] = mfma(af[0], bf, acc[0][nt]);
        acc[1][nt] = mfma(af[1], bf, acc[1][nt]);
      }
    }
  }
  conv_epilogue<NT>(a, acc, nz * a.h + y, x0, lane);
}

// As conv3d_fwd_kernel with a 2x larger wave tile and a software pipeline:
// one wave owns 2 rows x 32 voxels (four 16-voxel M tiles) x all cout, a block
// 8 consecutive rows; the A and B fragments of the next (tap, channel chunk)
// step are loaded while the MFMAs of the current one run.  H % 8 == 0,
// cin % 32 == 0.
template <int NT>
struct Frags {
  bf16x8 a[4];
  bf16x8 b[NT];
};

template <int NT>
__device__ __forceinline__ void conv2_load(const Conv& a, Frags<NT>& f, int it, int n, int z,
                                           int y0, int x0, int m, int kq) {
  const int tap = it % 27, c0 = (it / 27) * 32;
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int zz = z + dz;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int yy = y0 + (mt >> 1) + dy;
    const int xx = x0 + 16 * (mt & 1) + m + dx;
    const bool ok = zz >= 0 && zz < a.d && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
    const long long v = ((((long long)n * a.d + zz) * a.h + yy) * a.w + xx);
    f.a[mt] = ok ? load8(a.x + v * a.cin + c0 + kq * 8) : zero8();
  }
  const u16* wb = a.wt + (long long)tap * a.cout * a.cin + c0 + kq * 8;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) f.b[nt] = load8(wb + (long long)(nt * 16 + m) * a.cin);
}

template <int NT>
__device__ __forceinline__ void conv2_mma(f32x4 (&acc)[4][NT], const Frags<NT>& f) {
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) acc[mt][nt] = mfma(f.a[mt], f.b[nt], acc[mt][nt]);
}

template <int NT>
__global__ __launch_bounds__(256) void conv3d_fwd2_kernel(Conv a) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int x0 = blockIdx.x * 32;
  const int y0 = blockIdx.y * 8 + 2 * wave;
  const int nz = blockIdx.z;
  const int z = nz % a.d, n = nz / a.d;
  f32x4 acc[4][NT];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int niters = 27 * (a.cin / 32);
  Frags<NT> f0, f1;
  conv2_load<NT>(a, f0, 0, n, z, y0, x0, m, kq);
  for (int it = 0; it < niters; it += 2) {
    if (it + 1 < niters) conv2_load<NT>(a, f1, it + 1, n, z, y0, x0, m, kq);
    conv2_mma<NT>(acc, f0);
    if (it + 1 < niters) {
      if (it + 2 < niters) conv2_load<NT>(a, f0, it + 2, n, z, y0, x0, m, kq);
      conv2_mma<NT>(acc, f1);
    }
  }
  // epilogue: M tile mt is row y0 + (mt >> 1), voxels x0 + 16 (mt & 1) + 4 kq + i
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + 16 * (mt & 1) + 4 * kq + i;
        const int co = nt * 16 + m;
        const long long o = (((long long)nz * a.h + y0 + (mt >> 1)) * a.w + x) * a.cout + co;
        float v = acc[mt][nt][i];
        if (a.bias) v += a.bias[co];
        v = activate(v, a.relu);
        if (a.resid) v += a.resid[o];
        if (a.out_f32) static_cast<float*>(a.y)[o] = v;
        else static_cast<u16*>(a.y)[o] = f2bf(v);
      }
}

// LDS-staged form for cin % 32 == 0: a block owns 2 planes x 4 rows x 32
// voxels (one wave: 1 plane x 2 rows, four 16-voxel M tiles, all cout) and,
// per 32-channel chunk, stages the input halo (4 planes x 6 rows x 34 voxels)
// in LDS as 16-byte (8-channel) chunks laid out [plane][row][chunk][x], so
// the 16 lanes of an A fragment read 256 contiguous bytes.  The 27 taps then
// take A from LDS and B (weights, L2-resident) from memory.  Needs W % 32,
// H % 4 and D % 2 == 0.
constexpr int kHaloX = 34, kHaloRows = 6, kHaloPlanes = 4;
constexpr int kHaloChunks = kHaloPlanes * kHaloRows * 4 * kHaloX;  // 16-byte chunks

template <int NT>
__device__ __forceinline__ void fwd3_load_b(bf16x8 (&b)[NT], const u16* wc, int tap,
                                            long long tap_stride, int cin) {
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) b[nt] = load8(wc + tap * tap_stride + (long long)nt * 16 * cin);
}

__device__ __forceinline__ void fwd3_load_a(bf16x8 (&af)[4], const uint4* halo, int tap, int pz,
                                            int py, int m, int kq) {
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const int hz = pz + 1 + dz, hy = py + (mt >> 1) + 1 + dy;
    const int hx = 16 * (mt & 1) + m + 1 + dx;
    af[mt] = __builtin_bit_cast(bf16x8, halo[((hz * kHaloRows + hy) * 4 + kq) * kHaloX + hx]);
  }
}

template <int NT>
__device__ __forceinline__ void fwd3_mma(f32x4 (&acc)[4][NT], const bf16x8 (&af)[4],
                                         const bf16x8 (&b)[NT]) {
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) acc[mt][nt] = mfma(af[mt], b[nt], acc[mt][nt]);
}

template <int NT>
__global__ __launch_bounds__(256) void conv3d_fwd3_kernel(Conv a) {
  __shared__ uint4 halo[kHaloChunks];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 4;
  const int zb = blockIdx.z * 2;  // n * d + z of the block's first plane
  const int n = zb / a.d, z0 = zb % a.d;
  const int pz = wave >> 1, py = 2 * (wave & 1);  // the wave's plane and first row
  f32x4 acc[4][NT];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int c0 = 0; c0 < a.cin; c0 += 32) {
    __syncthreads();  // the previous chunk's reads are done
    for (int i = threadIdx.x; i < kHaloChunks; i += 256) {
      const int q = i & 3;
      int r = i >> 2;
      const int hx = r % kHaloX;
      r /= kHaloX;
      const int hy = r % kHaloRows, hz = r / kHaloRows;
      const int zz = z0 - 1 + hz, yy = y0 - 1 + hy, xx = x0 - 1 + hx;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (zz >= 0 && zz < a.d && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
        v = *reinterpret_cast<const uint4*>(
            a.x + ((((long long)n * a.d + zz) * a.h + yy) * a.w + xx) * a.cin + c0 + 8 * q);
      halo[((hz * kHaloRows + hy) * 4 + q) * kHaloX + hx] = v;
    }
    __syncthreads();
    // weight fragments run two taps ahead of the MFMAs in a 3-deep register ring
    const u16* wc = a.wt + c0 + kq * 8 + (long long)m * a.cin;
    const long long tap_stride = (long long)a.cout * a.cin;
    bf16x8 b0[NT], b1[NT], b2[NT];
    fwd3_load_b<NT>(b0, wc, 0, tap_stride, a.cin);
    fwd3_load_b<NT>(b1, wc, 1, tap_stride, a.cin);
    // Scheduling barriers keep each weight load where it is issued, two taps
    // ahead of its MFMAs, so its latency runs under the taps between.
    bf16x8 af[4];
    for (int t3 = 0; t3 < 27; t3 += 3) {
      fwd3_load_b<NT>(b2, wc, t3 + 2, tap_stride, a.cin);
      __builtin_amdgcn_sched_barrier(0);
      fwd3_load_a(af, halo, t3, pz, py, m, kq);
      fwd3_mma<NT>(acc, af, b0);
      if (t3 + 3 < 27) fwd3_load_b<NT>(b0, wc, t3 + 3, tap_stride, a.cin);
      __builtin_amdgcn_sched_barrier(0);
      fwd3_load_a(af, halo, t3 + 1, pz, py, m, kq);
      fwd3_mma<NT>(acc, af, b1);
      if (t3 + 4 < 27) fwd3_load_b<NT>(b1, wc, t3 + 4, tap_stride, a.cin);
      __builtin_amdgcn_sched_barrier(0);
      fwd3_load_a(af, halo, t3 + 2, pz, py, m, kq);
      fwd3_mma<NT>(acc, af, b2);
    }
  }
  const int zo = blockIdx.z * 2 + pz;  // n * d + z of the wave's plane
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int x = x0 + 16 * (mt & 1) + 4 * kq + i;
        const int co = nt * 16 + m;
        const long long o =
            (((long long)zo * a.h + y0 + py + (mt >> 1)) * a.w + x) * a.cout + co;
        float v = acc[mt][nt][i];
        if (a.bias) v += a.bias[co];
        v = activate(v, a.relu);
        if (a.resid) v += a.resid[o];
        if (a.out_f32) static_cast<float*>(a.y)[o] = v;
        else static_cast<u16*>(a.y)[o] = f2bf(v);
      }
}

// Lean form of the LDS-halo kernel for cout = 32 (NT = 2, PF) and 64 (NT = 4,
// no PF: its registers cannot also hold the next halo; cout 128 and 192 run
// as 64-channel passes of it, output channels co0 .. co0 + 63), on the same
// tiles (2 planes x 4 rows x 32 voxels per block, 1 plane x 2 rows per wave)
// and halo layout as conv3d_fwd3_kernel:
//  * the 27 taps are unrolled, so every A-fragment read is one ds_read_b128
//    at an immediate offset from a per-lane base;
//  * weights come through a buffer resource (lane offset in a VGPR, tap and
//    chunk offset in an SGPR), P taps ahead of their MFMAs;
//  * the MFMA takes the weights as its A operand and the voxels as B, so lane
//    (m, kq) ends with output channels 4 kq .. 4 kq + 3 of voxel m: one 8-byte
//    BF16 store (16-byte FP32) per lane and tile instead of four 2-byte ones;
//  * a block walks `tpb` consecutive tiles; with PF the next (tile, chunk) halo
//    is loaded into registers while the current one's taps run, without it
//    each halo is loaded right before it is staged.
// Staging offsets are 32-bit (4 planes of the input fit in 2^31 elements).
constexpr int kHaloPer = (kHaloChunks + 255) / 256;  // halo slots per thread
constexpr int kHaloRowSlots = 4 * kHaloX;            // 16-byte slots per halo row

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

__device__ inline uint2 pack4_bf16(float a, float b, float c, float d) {
  bf16x2 lo, hi;
  lo[0] = (__bf16)a;  // v_cvt_pk_bf16_f32: round to nearest even
  lo[1] = (__bf16)b;
  hi[0] = (__bf16)c;
  hi[1] = (__bf16)d;
  return make_uint2(__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi));
}

struct Fwd4Tile {
  int x0, y0, z0, n, zb;  // zb: n * d + z0
  bool interior;          // the whole halo lies inside the volume
};

// Tile `tile` of TP planes x TR rows x 32 voxels (x fastest, then rows, then
// the flattened n * d planes; d % TP == 0).
template <int TP = 2, int TR = 4>
__device__ __forceinline__ Fwd4Tile fwd4_tile(const Conv& a, int tile, int tiles_x, int tiles_y) {
  Fwd4Tile f;
  const int tx = tile % tiles_x, r = tile / tiles_x;
  const int ty = r % tiles_y;
  f.zb = (r / tiles_y) * TP;
  f.x0 = tx * 32;
  f.y0 = ty * TR;
  f.n = f.zb / a.d;
  f.z0 = f.zb - f.n * a.d;
  f.interior = f.z0 >= 1 && f.z0 + TP + 1 <= a.d && f.y0 >= 1 && f.y0 + TR + 1 <= a.h &&
               f.x0 >= 1 && f.x0 + 33 <= a.w;
  return f;
}

// Halo of a TP x TR x 32 tile: (TP + 2) planes x (TR + 2) rows x 34 voxels,
// 16-byte chunks of 8 channels, handed out to 256 threads PER at a time.
template <int TP, int TR>
struct HaloShape {
  static constexpr int rows = TR + 2;                         // halo rows per plane
  static constexpr int chunks = (TP + 2) * rows * 4 * kHaloX;  // 16-byte chunks
  static constexpr int per = (chunks + 255) / 256;
};
template <int TP, int TR>
using HaloRegs = uint4[HaloShape<TP, TR>::per];  // one thread's share of the halo

// Loads the halo of (tile, 32-channel chunk c0) into registers: slot s = t + 256 j
// is halo row s / 136 (plane-major), x s % 136 / 4 and 8-channel group s % 4.
// Offsets are signed and relative to the tile's first voxel (z0, y0, x0), which
// is always inside the volume; slots outside it are never addressed.  Interior
// tiles (uniform) skip the range checks.  A chunk of a decoder input's
// upsampled half (uniform) reads the coarse tensor `up`: full-resolution voxel
// z0 - 1 + hz maps to coarse (z0 - 1 + hz) >> 1 = z0 / 2 + ((hz - 1) >> 1) (z0,
// y0 and x0 are even), the same offset formula with every step halved.
template <int TP = 2, int TR = 4>
__device__ __forceinline__ void fwd4_fetch(HaloRegs<TP, TR>& pre, const Conv& a, const Fwd4Tile& f,
                                           int c0) {
  constexpr int kRows = HaloShape<TP, TR>::rows, kPer = HaloShape<TP, TR>::per,
                kChunks = HaloShape<TP, TR>::chunks;
  const int t = threadIdx.x;
  const bool coarse = a.up != nullptr && c0 < a.cup;
  const int sh = coarse ? 1 : 0;
  const int cc = coarse ? a.cup : a.cin - a.cup;  // channels per voxel of the source
  const int hh = a.h >> sh, ww = a.w >> sh;
  const u16* org = (coarse ? a.up : a.x) +
                   ((((long long)f.n * (a.d >> sh) + (f.z0 >> sh)) * hh + (f.y0 >> sh)) * ww +
                    (f.x0 >> sh)) * cc + (coarse ? c0 : c0 - a.cup);
  const int plane = hh * ww * cc, rowe = ww * cc;
  int row = t / kHaloRowSlots, w = t - row * kHaloRowSlots;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int hz = row / kRows, hy = row - hz * kRows, hx = w >> 2;
    bool ok = j + 1 < kPer || t + 256 * j < kChunks;
    if (!f.interior) {
      const int zz = f.z0 - 1 + hz, yy = f.y0 - 1 + hy, xx = f.x0 - 1 + hx;
      ok = ok && zz >= 0 && zz < a.d && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
    }
    const int off = ((hz - 1) >> sh) * plane + ((hy - 1) >> sh) * rowe + ((hx - 1) >> sh) * cc + 8 * (w & 3);
    pre[j] = ok ? *reinterpret_cast<const uint4*>(org + off) : make_uint4(0u, 0u, 0u, 0u);
    w += 256 - kHaloRowSlots;  // 256 = one row (136 slots) + 120
    row += 1;
    if (w >= kHaloRowSlots) {
      w -= kHaloRowSlots;
      row += 1;
    }
  }
}

template <int TP = 2, int TR = 4>
__device__ __forceinline__ void fwd4_commit(uint4* halo, const HaloRegs<TP, TR>& pre) {
  constexpr int kPer = HaloShape<TP, TR>::per, kChunks = HaloShape<TP, TR>::chunks;
  const int t = threadIdx.x;
  int row = t / kHaloRowSlots, w = t - row * kHaloRowSlots;
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    if (t + 256 * j < kChunks) halo[(row * 4 + (w & 3)) * kHaloX + (w >> 2)] = pre[j];
    w += 256 - kHaloRowSlots;
    row += 1;
    if (w >= kHaloRowSlots) {
      w -= kHaloRowSlots;
      row += 1;
    }
  }
}

// fwd4_fetch through a buffer resource based at the tile's lowest halo element
// (hz, hy, hx = 0): one 32-bit offset per slot instead of a 64-bit address, and
// slots outside the volume take an offset past the resource's end, which the
// buffer range check turns into zeros, instead of a branch.  The resource
// spans TP + 2 planes and 2 rows of the source (< 2^32 - 16 bytes: checked by
// the launcher).
template <int TP, int TR>
__device__ __forceinline__ void fwd5_fetch(HaloRegs<TP, TR>& pre, const Conv& a, const Fwd4Tile& f,
                                           int c0) {
  constexpr int kRows = HaloShape<TP, TR>::rows, kPer = HaloShape<TP, TR>::per,
                kChunks = HaloShape<TP, TR>::chunks;
  const int t = threadIdx.x;
  const bool coarse = a.up != nullptr && c0 < a.cup;
  const int sh = coarse ? 1 : 0;
  const int cc = coarse ? a.cup : a.cin - a.cup;
  const int hh = a.h >> sh, ww = a.w >> sh;
  const int plane = hh * ww * cc, rowe = ww * cc;
  const long long lo = (((long long)f.n * (a.d >> sh) + (f.z0 >> sh)) * hh + (f.y0 >> sh)) * ww +
                       (f.x0 >> sh) - ((long long)hh * ww + ww + 1);  // voxel (hz, hy, hx) = 0
  const u16* base = (coarse ? a.up : a.x) + lo * cc + (coarse ? c0 : c0 - a.cup);
  // in-volume slots lie within TP + 2 planes + 2 rows of the base
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)base, 0, (int)(((unsigned)(TP + 2) * (unsigned)plane + 2u * (unsigned)rowe) * 2u),
      fdk::kRawRsrcWord3);
  int row = t / kHaloRowSlots, w = t - row * kHaloRowSlots;
  // opaque: the slot coordinates are recomputed per call, not held in registers
  // across the tap loop
  asm volatile("" : "+v"(row), "+v"(w));
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int hz = row / kRows, hy = row - hz * kRows, hx = w >> 2;
    bool ok = j + 1 < kPer || t + 256 * j < kChunks;
    if (!f.interior) {
      const int zz = f.z0 - 1 + hz, yy = f.y0 - 1 + hy, xx = f.x0 - 1 + hx;
      ok = ok && zz >= 0 && zz < a.d && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
    }
    // fwd4_fetch's ((h - 1) >> sh) + 1 == (h + sh) >> sh for h >= 0, sh in {0, 1}
    const unsigned off = ((unsigned)((hz + sh) >> sh) * (unsigned)plane +
                          (unsigned)(((hy + sh) >> sh) * rowe + ((hx + sh) >> sh) * cc + 8 * (w & 3))) *
                         2u;
    pre[j] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, ok ? off : 0xFFFFFFF0u, 0, 0));
    w += 256 - kHaloRowSlots;
    row += 1;
    if (w >= kHaloRowSlots) {
      w -= kHaloRowSlots;
      row += 1;
    }
  }
}

// Weight fragments of one tap at the running offset `soff` (advanced by one tap;
// the empty asm keeps it a running SGPR instead of 27 hoisted products).
template <int NT>
__device__ __forceinline__ void fwd4_load_w(bf16x8 (&b)[NT], __amdgpu_buffer_rsrc_t wr,
                                            unsigned lane_off, unsigned& soff, unsigned nt_bytes,
                                            unsigned tap_bytes) {
  unsigned o = soff;
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    b[nt] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(wr, lane_off, o, 0));
    o += nt_bytes;
    asm volatile("" : "+s"(o));
  }
  soff += tap_bytes;
  asm volatile("" : "+s"(soff));
}

// The 4 A fragments (voxels) of one tap from the halo, at immediate offsets.
__device__ __forceinline__ void fwd4_read_a(bf16x8 (&af)[4], const uint4* hb, int tap) {
  constexpr int kRowStride = 4 * kHaloX;
  const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;  // 0..2: offset + 1
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
    af[mt] = __builtin_bit_cast(
        bf16x8, hb[(dz * kHaloRows + (mt >> 1) + dy) * kRowStride + 16 * (mt & 1) + dx]);
}

template <int NT, bool PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void conv3d_fwd4_kernel(
    Conv a, int tiles_x, int tiles_y, int ntiles, int tpb) {
  constexpr int P = NT == 2 ? 3 : 2;  // weight prefetch distance in taps
  constexpr int R = P + 1;            // weight ring slots
  __shared__ uint4 halo[kHaloChunks];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int pz = wave >> 1, py = 2 * (wave & 1);  // the wave's plane and first row in the tile
  const uint4* hb = halo + ((pz * kHaloRows + py) * 4 + kq) * kHaloX + m;
  const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc(
      (void*)a.wt, 0, (int)(27u * a.cout * a.cin * 2u), fdk::kRawRsrcWord3);
  const unsigned lane_off = (unsigned)(kq * 8 + (a.co0 + m) * a.cin) * 2u;
  const unsigned tap_bytes = (unsigned)(a.cout * a.cin) * 2u, nt_bytes = (unsigned)(16 * a.cin) * 2u;
  const int nchunks = a.cin / 32;
  const int first = blockIdx.x * tpb, last = min(first + tpb, ntiles);
  uint4 pre[kHaloPer];
  if (PF && first < last) fwd5_fetch<2, 4>(pre, a, fwd4_tile(a, first, tiles_x, tiles_y), 0);
  for (int tile = first; tile < last; ++tile) {
    const Fwd4Tile f = fwd4_tile(a, tile, tiles_x, tiles_y);
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ch = 0; ch < nchunks; ++ch) {
      if (!PF) fwd5_fetch<2, 4>(pre, a, f, ch * 32);  // staged synchronously: no registers held
      __syncthreads();  // every wave is done reading the previous halo
      fwd4_commit(halo, pre);
      __syncthreads();
      unsigned soff = (unsigned)ch * 64u;  // 32 channels x 2 bytes; then + tap x tap_bytes
      bf16x8 b[R][NT];
#pragma unroll
      for (int tp = 0; tp < P; ++tp) fwd4_load_w<NT>(b[tp], wr, lane_off, soff, nt_bytes, tap_bytes);
      // the next halo's loads run under this chunk's taps
      if (PF) {
        int nt_ = tile, nc = ch + 1;
        if (nc == nchunks) {
          nc = 0;
          ++nt_;
        }
        if (nt_ < last) fwd5_fetch<2, 4>(pre, a, nt_ == tile ? f : fwd4_tile(a, nt_, tiles_x, tiles_y), nc * 32);
      }
      // NT == 2: A fragments are read one tap ahead of their MFMAs (two register
      // sets); NT == 4 has no registers to spare and reads them in the tap.
      constexpr int AB = NT == 2 ? 2 : 1;
      bf16x8 af[AB][4];
      if (AB == 2) fwd4_read_a(af[0], hb, 0);
#pragma unroll
      for (int tap = 0; tap < 27; ++tap) {
        if (tap + P < 27)
          fwd4_load_w<NT>(b[(tap + P) % R], wr, lane_off, soff, nt_bytes, tap_bytes);
        if (AB == 2 && tap + 1 < 27) fwd4_read_a(af[(tap + 1) % AB], hb, tap + 1);
        __builtin_amdgcn_sched_barrier(0);
        if (AB == 1) fwd4_read_a(af[0], hb, tap);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt)
#pragma unroll
          for (int mt = 0; mt < 4; ++mt)
            acc[mt][nt] = mfma(b[tap % R][nt], af[tap % AB][mt], acc[mt][nt]);
      }
    }
    // lane (m, kq) holds channels nt * 16 + 4 kq + i of voxel m of M tile mt
    const int zo = f.zb + pz;
    float4 bias[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      bias[nt] = a.bias ? *reinterpret_cast<const float4*>(a.bias + a.co0 + nt * 16 + 4 * kq)
                        : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      const long long vox =
          ((long long)zo * a.h + f.y0 + py + (mt >> 1)) * a.w + f.x0 + 16 * (mt & 1) + m;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const long long o = vox * a.cout + a.co0 + nt * 16 + 4 * kq;
        float v[4] = {acc[mt][nt][0] + bias[nt].x, acc[mt][nt][1] + bias[nt].y,
                      acc[mt][nt][2] + bias[nt].z, acc[mt][nt][3] + bias[nt].w};
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = activate(v[i], a.relu);
        if (a.resid) {
          const float4 r = *reinterpret_cast<const float4*>(a.resid + o);
          v[0] += r.x;
          v[1] += r.y;
          v[2] += r.z;
          v[3] += r.w;
        }
        if (a.out_f32)
          *reinterpret_cast<float4*>(static_cast<float*>(a.y) + o) = make_float4(v[0], v[1], v[2], v[3]);
        else
          *reinterpret_cast<uint2*>(static_cast<u16*>(a.y) + o) = pack4_bf16(v[0], v[1], v[2], v[3]);
      }
    }
  }
}

// conv3d_fwd5_kernel: the LDS-halo convolution with each A fragment (16 input
// voxels x 32 channels of one halo row at one x shift) read once per weight
// slice and applied to every output row that uses it.
//  * tile: TP planes x Q rows x 32 voxels x 64 / TP output channels; wave w
//    owns output plane w % TP, all Q rows, both 16-voxel halves and NT = 2
//    16-channel output tiles from channel co0 + 32 (w / TP) (cout 32: TP = 4;
//    64-channel passes: TP = 2, the two waves of a plane sharing its A
//    fragments); halo ((TP + 2) planes x (Q + 2) rows x 34: 78 KB at TP = 4,
//    52 KB at TP = 2) in dynamic LDS, the wave reading planes w % TP + 0..2;
//  * the 27 taps run as 9 slices (dx, dz), each 3 dy taps x NT weight
//    fragments held in registers (one slice loaded ahead); in a slice, input
//    row yi (0..Q+1) of plane w + dz feeds output rows yi - dy, 0 <= yi - dy < Q;
//  * per 32-channel chunk and wave: 18 (Q + 2) A fragments and 27 NT weight
//    fragments for 54 Q NT MFMAs: 0.25 LDS and 0.125 weight fragment reads per
//    MFMA (conv3d_fwd4_kernel: 0.5 and 0.25 at cout 32, 0.25 and 0.25 at 64).
// Needs d % TP == 0 and h % Q == 0 besides conv3d_fwd4_kernel's rules.
template <int NT>
__device__ __forceinline__ void fwd5_load_slice(bf16x8 (&b)[3][NT], __amdgpu_buffer_rsrc_t wr,
                                                unsigned lane_off, unsigned o, unsigned nt_bytes,
                                                unsigned dy_bytes) {
#pragma unroll
  for (int dy = 0; dy < 3; ++dy) {
    unsigned p = o;
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      b[dy][nt] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(wr, lane_off, p, 0));
      p += nt_bytes;
      asm volatile("" : "+s"(p));
    }
    o += dy_bytes;
    asm volatile("" : "+s"(o));
  }
}

// A fragment of step i (slice i / K: dx = slice / 3, dz = slice % 3; input row
// (i % K) / 2, half i % 2) from the wave's halo window.
template <int Q>
__device__ __forceinline__ bf16x8 fwd5_read_a(const uint4* hb, int i) {
  constexpr int K = 2 * (Q + 2);
  const int sl = i / K, k = i % K;
  const int dx = sl / 3, dz = sl % 3, yi = k / 2, mt = k % 2;
  return __builtin_bit_cast(bf16x8, hb[(dz * (Q + 2) + yi) * (4 * kHaloX) + 16 * mt + dx]);
}

template <int NT, int Q, int TP, bool PF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void conv3d_fwd5_kernel(
    Conv a, int tiles_x, int tiles_y, int ntiles, int tpb) {
  using HS = HaloShape<TP, Q>;
  extern __shared__ uint4 halo5[];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int pz = wave % TP;                          // the wave's output plane in the tile
  const int cw = a.co0 + (wave / TP) * (16 * NT);   // and its first output channel
  const uint4* hb = halo5 + (pz * HS::rows * 4 + kq) * kHaloX + m;
  const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc(
      (void*)a.wt, 0, (int)(27u * a.cout * a.cin * 2u), fdk::kRawRsrcWord3);
  const unsigned lane_off = (unsigned)(kq * 8 + (cw + m) * a.cin) * 2u;
  const unsigned tap_bytes = (unsigned)(a.cout * a.cin) * 2u, nt_bytes = (unsigned)(16 * a.cin) * 2u;
  const int nchunks = a.cin / 32;
  const int first = blockIdx.x * tpb, last = min(first + tpb, ntiles);
  uint4 pre[HS::per];
  if (PF && first < last) fwd5_fetch<TP, Q>(pre, a, fwd4_tile<TP, Q>(a, first, tiles_x, tiles_y), 0);
  for (int tile = first; tile < last; ++tile) {
    const Fwd4Tile f = fwd4_tile<TP, Q>(a, tile, tiles_x, tiles_y);
    f32x4 acc[Q][2][NT];
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[q][mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int ch = 0; ch < nchunks; ++ch) {
      if (!PF) fwd5_fetch<TP, Q>(pre, a, f, ch * 32);
      __syncthreads();  // every wave is done reading the previous halo
      fwd4_commit<TP, Q>(halo5, pre);
      __syncthreads();
      const unsigned cb = (unsigned)ch * 64u;  // 32 channels x 2 bytes
      bf16x8 b[2][3][NT];
      fwd5_load_slice<NT>(b[0], wr, lane_off, cb, nt_bytes, 3u * tap_bytes);
      if (PF) {
        int nt_ = tile, nc = ch + 1;
        if (nc == nchunks) {
          nc = 0;
          ++nt_;
        }
        if (nt_ < last)
          fwd5_fetch<TP, Q>(pre, a, nt_ == tile ? f : fwd4_tile<TP, Q>(a, nt_, tiles_x, tiles_y), nc * 32);
      }
      // steps i = (slice, input row, half): A fragments kAhead steps ahead of
      // their MFMAs, the next slice's weights at the first step of a slice
      constexpr int K = 2 * (Q + 2), kSteps = 9 * K, kAhead = 2;
      bf16x8 ar[kAhead + 1];
#pragma unroll
      for (int i = 0; i < kAhead; ++i) ar[i] = fwd5_read_a<Q>(hb, i);
#pragma unroll
      for (int sl = 0; sl < 9; ++sl) {
        if (sl + 1 < 9) {
          const int dx1 = (sl + 1) / 3, dz1 = (sl + 1) % 3;
          unsigned o = (unsigned)(dz1 * 9 + dx1) * tap_bytes;
          asm volatile("" : "+s"(o));
          fwd5_load_slice<NT>(b[(sl + 1) & 1], wr, lane_off, cb + o, nt_bytes, 3u * tap_bytes);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int i = sl * K + k, yi = k / 2, mt = k % 2;
          if (i + kAhead < kSteps) ar[(i + kAhead) % (kAhead + 1)] = fwd5_read_a<Q>(hb, i + kAhead);
          __builtin_amdgcn_sched_barrier(0);
          const bf16x8 af = ar[i % (kAhead + 1)];
#pragma unroll
          for (int dy = 0; dy < 3; ++dy) {
            const int q = yi - dy;
            if (q < 0 || q >= Q) continue;
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) acc[q][mt][nt] = mfma(b[sl & 1][dy][nt], af, acc[q][mt][nt]);
          }
        }
      }
    }
    // lane (m, kq) holds channels nt * 16 + 4 kq + i of voxel m of half mt of row q
    const int zo = f.zb + pz;
    float4 bias[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      bias[nt] = a.bias ? *reinterpret_cast<const float4*>(a.bias + cw + nt * 16 + 4 * kq)
                        : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        const long long vox = ((long long)zo * a.h + f.y0 + q) * a.w + f.x0 + 16 * mt + m;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
          const long long o = vox * a.cout + cw + nt * 16 + 4 * kq;
          float v[4] = {acc[q][mt][nt][0] + bias[nt].x, acc[q][mt][nt][1] + bias[nt].y,
                        acc[q][mt][nt][2] + bias[nt].z, acc[q][mt][nt][3] + bias[nt].w};
#pragma unroll
          for (int i = 0; i < 4; ++i) v[i] = activate(v[i], a.relu);
          if (a.resid) {
            const float4 r = *reinterpret_cast<const float4*>(a.resid + o);
            v[0] += r.x;
            v[1] += r.y;
            v[2] += r.z;
            v[3] += r.w;
          }
          if (a.out_f32)
            *reinterpret_cast<float4*>(static_cast<float*>(a.y) + o) = make_float4(v[0], v[1], v[2], v[3]);
          else
            *reinterpret_cast<uint2*>(static_cast<u16*>(a.y) + o) = pack4_bf16(v[0], v[1], v[2], v[3]);
        }
      }
  }
}

// cin == 1 (the network input): the 27 taps of the single channel form K (one
// MFMA K step, padded to 32); weights [cout][32].
template <int NT>
__global__ __launch_bounds__(256) void conv3d_fwd_c1_kernel(Conv a) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int x0 = blockIdx.x * 32;
  const int y = blockIdx.y * 4 + wave;
  const int nz = blockIdx.z;
  const int z = nz % a.d, n = nz / a.d;
  if (y >= a.h) return;
  f32x4 acc[2][NT];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  bf16x8 af[2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    const int x = x0 + mt * 16 + m;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int tap = kq * 8 + j;
      float v = 0.f;
      if (tap < 27) {
        const int zz = z + tap / 9 - 1, yy = y + (tap / 3) % 3 - 1, xx = x + tap % 3 - 1;
        if (zz >= 0 && zz < a.d && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
          v = bf2f(a.x[(((long long)n * a.d + zz) * a.h + yy) * a.w + xx]);
      }
      af[mt][j] = (__bf16)v;
    }
  }
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) {
    const bf16x8 bf = load8(a.wt + (nt * 16 + m) * 32 + kq * 8);
    acc[0][nt] = mfma(af[0], bf, acc[0][nt]);
    acc[1][nt] = mfma(af[1], bf, acc[1][nt]);
  }
  conv_epilogue<NT>(a, acc, nz * a.h + y, x0, lane);
}

// cin == 1, row form: one wave per (n, z, y) row, stepping along x 32 voxels
// (two 16-voxel tiles) at a time, so a launch has h / 4 x n d blocks instead of
// one per 32 x 4 voxels.  The MFMA takes the weights as A and the voxels' taps
// as B, so lane (m, kq) ends with channels 4 kq .. 4 kq + 3 of voxel m: one
// 8-byte BF16 (16-byte FP32) store per tile and 16 channels.  The row offsets
// and z / y validity of a lane's 8 taps are fixed for the whole row; x is
// checked only in the first and last steps.
template <int NT>
__global__ __launch_bounds__(256) void conv3d_fwd_c1r_kernel(Conv a) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int y = blockIdx.x * 4 + wave;
  const int nz = blockIdx.y;
  const int z = nz % a.d, n = nz / a.d;
  bf16x8 wf[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) wf[nt] = load8(a.wt + (nt * 16 + m) * 32 + kq * 8);
  long long rowb[8];  // element offset of tap j's input row, + dx
  int dxs[8];
  bool tap_ok[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int tap = kq * 8 + j;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int zz = z + dz, yy = y + dy;
    tap_ok[j] = tap < 27 && zz >= 0 && zz < a.d && yy >= 0 && yy < a.h;
    dxs[j] = dx;
    rowb[j] = (((long long)n * a.d + zz) * a.h + yy) * a.w + dx;
  }
  float4 bias[NT];
#pragma unroll
  for (int nt = 0; nt < NT; ++nt)
    bias[nt] = a.bias ? *reinterpret_cast<const float4*>(a.bias + nt * 16 + 4 * kq)
                      : make_float4(0.f, 0.f, 0.f, 0.f);
  const long long out_row = ((long long)nz * a.h + y) * a.w;
  for (int x0 = 0; x0 < a.w; x0 += 32) {
    const bool inner = x0 >= 1 && x0 + 33 <= a.w;  // uniform
    bf16x8 vf[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int x = x0 + mt * 16 + m;
      u16 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool ok = tap_ok[j] && (inner || ((unsigned)(x + dxs[j]) < (unsigned)a.w));
        v[j] = ok ? a.x[rowb[j] + x] : (u16)0;
      }
      const uint4 packed = make_uint4(v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16),
                                      v[4] | ((unsigned)v[5] << 16), v[6] | ((unsigned)v[7] << 16));
      vf[mt] = __builtin_bit_cast(bf16x8, packed);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const long long vox = out_row + x0 + mt * 16 + m;
#pragma unroll
      for (int nt = 0; nt < NT; ++nt) {
        const f32x4 acc = mfma(wf[nt], vf[mt], f32x4{0.f, 0.f, 0.f, 0.f});
        const long long o = vox * a.cout + nt * 16 + 4 * kq;
        float r[4] = {acc[0] + bias[nt].x, acc[1] + bias[nt].y, acc[2] + bias[nt].z,
                      acc[3] + bias[nt].w};
#pragma unroll
        for (int i = 0; i < 4; ++i) r[i] = activate(r[i], a.relu);
        if (a.resid) {
          const float4 q = *reinterpret_cast<const float4*>(a.resid + o);
          r[0] += q.x;
          r[1] += q.y;
          r[2] += q.z;
          r[3] += q.w;
        }
        if (a.out_f32)
          *reinterpret_cast<float4*>(static_cast<float*>(a.y) + o) = make_float4(r[0], r[1], r[2], r[3]);
        else
          *reinterpret_cast<uint2*>(static_cast<u16*>(a.y) + o) = pack4_bf16(r[0], r[1], r[2], r[3]);
      }
    }
  }
}

// Weight gradient dW[tap][co][ci] += sum over voxels dY[v][co] X[v + tap][ci]
// (K = voxels).  A block owns one tap and a run of `rows` output rows (n, z, y);
// per row it stages dY and the tap-shifted X row in LDS transposed to [c][x] so
// each lane's 8 consecutive voxels are one 16-byte LDS read, accumulates all
// (co, ci) 16x16 tiles split over its 4 waves, and adds them to dW once.
constexpr int kWgradRows = 64;

struct Wgrad {
  const u16* x;   // [n][d][h][w][cin], or [n][d][h][w][cin - cup] with `up`
  const u16* dy;  // [n][d][h][w][cout]
  float* dw;      // [27][cout][cin], accumulated with atomics
  int n, d, h, w, cin, cout;
  // Decoder form (conv3d_wgrad_tr_kernel only): the input is [nearest 2x
  // upsampling of up ([n][d/2][h/2][w/2][cup]), x], read in place.
  const u16* up = nullptr;
  int cup = 0;
};

template <int MAXT>
__global__ __launch_bounds__(256) void conv3d_wgrad_kernel(Wgrad a) {
  extern __shared__ u16 lds[];
  const int W = a.w, WP = W + 8;   // padded row length in LDS
  u16* dyT = lds;                  // [cout][WP]
  u16* xT = lds + a.cout * WP;     // [cin][WP]
  const int tap = blockIdx.y;
  const int dz = tap / 9 - 1, dy_ = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int m = lane % 16, kq = lane / 16;
  const int tco = a.cout / 16, tci = a.cin / 16, ntiles = tco * tci;
  f32x4 acc[MAXT];
#pragma unroll
  for (int t = 0; t < MAXT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int total_rows = a.n * a.d * a.h;
  const int r0 = blockIdx.x * kWgradRows;
  for (int r = r0; r < min(r0 + kWgradRows, total_rows); ++r) {
    const int y = r % a.h, z = (r / a.h) % a.d, n = r / (a.h * a.d);
    const int zz = z + dz, yy = y + dy_;
    const bool zy = zz >= 0 && zz < a.d && yy >= 0 && yy < a.h;
    __syncthreads();
    // 16-byte loads of 8 channels of one voxel, spread into the transposed
    // [c][x] rows (consecutive lanes take consecutive x)
    const u16* dyrow = a.dy + (long long)r * W * a.cout;
    for (int i = threadIdx.x; i < W * (a.cout / 8); i += 256) {
      const int q = i / W, x = i - q * W;
      const uint4 v = *reinterpret_cast<const uint4*>(dyrow + (long long)x * a.cout + 8 * q);
      const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        dyT[(8 * q + 2 * e) * WP + x] = (u16)(u[e] & 0xffffu);
        dyT[(8 * q + 2 * e + 1) * WP + x] = (u16)(u[e] >> 16);
      }
    }
    const u16* xrow = a.x + ((((long long)n * a.d + zz) * a.h + yy) * W) * a.cin;
    for (int i = threadIdx.x; i < W * (a.cin / 8); i += 256) {
      const int q = i / W, x = i - q * W;
      const int xs = x + dx;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (zy && xs >= 0 && xs < W)
        v = *reinterpret_cast<const uint4*>(xrow + (long long)xs * a.cin + 8 * q);
      const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        xT[(8 * q + 2 * e) * WP + x] = (u16)(u[e] & 0xffffu);
        xT[(8 * q + 2 * e + 1) * WP + x] = (u16)(u[e] >> 16);
      }
    }
    __syncthreads();
    for (int k0 = 0; k0 < W; k0 += 32) {
      int t = 0;
      for (int tile = wave; tile < ntiles; tile += 4, ++t) {
        const int co = (tile / tci) * 16 + m, ci = (tile % tci) * 16 + m;
        const bf16x8 av = *reinterpret_cast<const bf16x8*>(dyT + co * WP + k0 + kq * 8);
        const bf16x8 bv = *reinterpret_cast<const bf16x8*>(xT + ci * WP + k0 + kq * 8);
#pragma unroll
        for (int q = 0; q < MAXT; ++q)
          if (q == t) acc[q] = mfma(av, bv, acc[q]);
      }
    }
  }
  int t = 0;
  for (int tile = wave; tile < ntiles; tile += 4, ++t) {
    const int co0 = (tile / tci) * 16, ci = (tile % tci) * 16 + m;
#pragma unroll
    for (int q = 0; q < MAXT; ++q)
      if (q == t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int co = co0 + 4 * kq + i;
          atomicAdd(&a.dw[((long long)tap * a.cout + co) * a.cin + ci], acc[q][i]);
        }
  }
}

// Weight gradient on the forward kernel's tiles (2 planes x 4 rows x 32 voxels):
// a block owns one 32-channel chunk of cout and of cin and walks tiles; per
// tile it stages the dY chunk (256 voxels x 32 channels) and the input halo
// chunk (4 planes x 6 rows x 34 voxels x 32 channels) in LDS in their natural
// channels-last order -- plain 16-byte copies, no transposition -- and then
// accumulates all 27 taps: dW[tap][co][ci] += sum over the tile's voxels v of
// dY[v][co] X[v + tap][ci], with K = 32 voxels of one tile row per MFMA.
//
// The operands come out of LDS through gfx950's transposing read
// ds_read_b64_tr_b16: within each 16-lane group, source lane s loads 4
// consecutive 16-bit values R[s][0..3] and lane l receives R[4 j + l / 4][l % 4],
// j = 0..3 (lane layout checked on the device).  Pointing source lane s at
// voxel x0 + s / 4, channels c0 + 4 (s % 4) .. + 3 of a [voxel][channel] tile
// hands lane l the values of channel c0 + l for voxels x0 .. x0 + 3: the
// 8-voxel K run of an MFMA operand is two such reads.  A tap's x shift is only
// a different voxel row in those addresses.
//
// Wave w computes co tile (w & 1) x ci tile (w >> 1) of the chunk pair for all
// 27 taps (27 accumulators); per tile it holds the dY fragments of its 8 rows
// and reads each of the 72 input fragments (24 halo rows x 3 x shifts) once,
// applying it to each of the 1..4 (tap, row) pairs it serves (216 MFMAs); each
// block adds its partial sums to dW with one FP32 atomic per element at the end.  Needs cin % 32 == cout % 32 == 0,
// W % 32, H % 4 and D % 2 == 0.
typedef short s16x4 __attribute__((ext_vector_type(4)));
constexpr int kWgDy = 256 * 4;  // 16-byte chunks of the dY tile (256 voxels x 32 channels)

// Byte offset, from the lane's own, of the input fragment of step i of the
// weight gradient's loop: halo row i / 3 (plane-major), x shift i % 3.
__host__ __device__ constexpr unsigned wg_b_off(int i) {
  return (unsigned)(((i / 3) * kHaloX + i % 3) * (32 * 2));
}

__device__ __forceinline__ bf16x8 tr_frag(const unsigned char* lds, unsigned off_lo, unsigned off_hi) {
  typedef __attribute__((address_space(3))) s16x4 lds_s16x4;
  const s16x4 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4*)(lds + off_lo));
  const s16x4 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s16x4*)(lds + off_hi));
  typedef short s16x8 __attribute__((ext_vector_type(8)));
  const s16x8 v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  return __builtin_bit_cast(bf16x8, v);
}

__global__ __launch_bounds__(256) void conv3d_wgrad_tr_kernel(Wgrad g, int tiles_x, int tiles_y,
                                                              int ntiles) {
  __shared__ uint4 halo[kHaloChunks];
  __shared__ uint4 dyt[kWgDy];
  const int t = threadIdx.x, wave = t / 64, lane = t % 64;
  const int kq = lane / 16, sl = lane % 16;  // 16-lane group and the lane within it
  const int ncoc = g.cout / 32;
  const int co0 = (blockIdx.y % ncoc) * 32, ci0 = (blockIdx.y / ncoc) * 32;
  const int ct = wave & 1, it = wave >> 1;
  // the forward kernel's tile and halo helpers, on the input tensor (or the
  // decoder input's two tensors)
  Conv a{g.x, nullptr, nullptr, nullptr, nullptr, g.n, g.d, g.h, g.w, g.cin, 32, 0, 0, g.up, g.cup};
  const unsigned char* hb = reinterpret_cast<const unsigned char*>(halo);
  const unsigned char* db = reinterpret_cast<const unsigned char*>(dyt);
  // per-lane byte offsets of the transposing reads: voxel x0 + sl / 4 (+ 4 for
  // the second half), channel 4 (sl % 4) of the wave's 16-channel tile
  const unsigned a_lane = ((8 * kq + sl / 4) * 32 + ct * 16 + 4 * (sl % 4)) * 2;
  const unsigned b_lane = ((8 * kq + sl / 4) * 32 + it * 16 + 4 * (sl % 4)) * 2;
  f32x4 acc[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const Fwd4Tile f = fwd4_tile(a, tile, tiles_x, tiles_y);
    uint4 pre[kHaloPer];
    fwd5_fetch<2, 4>(pre, a, f, ci0);
    // dY chunk: slot s = t + 256 j is tile row s / 128 (plane (s / 128) / 4, row
    // (s / 128) % 4), voxel (s % 128) / 4, 8-channel group s % 4; a tile row
    // is 32 voxels x 64 bytes, so the slots of one row are contiguous
    const u16* dyo = g.dy + (((long long)f.zb * g.h + f.y0) * g.w + f.x0) * g.cout + co0;
    const long long plane_e = (long long)g.h * g.w * g.cout, row_e = (long long)g.w * g.cout;
    const int pr0 = t >> 7, off0 = ((t & 127) >> 2) * g.cout + 8 * (t & 3);
    const uint4 d0 = *reinterpret_cast<const uint4*>(dyo + (pr0 >> 2) * plane_e + (pr0 & 3) * row_e + off0);
    const uint4 d1 = *reinterpret_cast<const uint4*>(dyo + ((pr0 + 2) >> 2) * plane_e + ((pr0 + 2) & 3) * row_e + off0);
    const uint4 d2 = *reinterpret_cast<const uint4*>(dyo + ((pr0 + 4) >> 2) * plane_e + ((pr0 + 4) & 3) * row_e + off0);
    const uint4 d3 = *reinterpret_cast<const uint4*>(dyo + ((pr0 + 6) >> 2) * plane_e + ((pr0 + 6) & 3) * row_e + off0);
    __syncthreads();  // the previous tile's reads are done
#pragma unroll
    for (int j = 0; j < kHaloPer; ++j)
      if (j + 1 < kHaloPer || t + 256 * j < kHaloChunks) halo[t + 256 * j] = pre[j];
    dyt[t] = d0;
    dyt[t + 256] = d1;
    dyt[t + 512] = d2;
    dyt[t + 768] = d3;
    __syncthreads();
    // the tile's 8 dY fragments (row kc: plane kc / 4, row kc % 4) stay in
    // registers; each input fragment (halo row r = (hz, hy), x shift dx) is
    // read once, two steps ahead, and feeds tap (hz - pz, hy - py, dx) of every
    // tile row (pz, py) it lies under
    bf16x8 af[8];
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) {
      const unsigned ao = a_lane + (unsigned)kc * (32 * 32 * 2);
      af[kc] = tr_frag(db, ao, ao + 4 * 32 * 2);
    }
    constexpr int kSteps = kHaloPlanes * kHaloRows * 3, kAhead = 4;
    bf16x8 br[kAhead + 1];
#pragma unroll
    for (int i = 0; i < kAhead; ++i) br[i] = tr_frag(hb, b_lane + wg_b_off(i), b_lane + wg_b_off(i) + 256);
#pragma unroll
    for (int hz = 0; hz < kHaloPlanes; ++hz)
#pragma unroll
      for (int hy = 0; hy < kHaloRows; ++hy)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          const int i = (hz * kHaloRows + hy) * 3 + dx;
          if (i + kAhead < kSteps) {
            const unsigned bo = b_lane + wg_b_off(i + kAhead);
            br[(i + kAhead) % (kAhead + 1)] = tr_frag(hb, bo, bo + 4 * 32 * 2);
          }
          __builtin_amdgcn_sched_barrier(0);
#pragma unroll
          for (int dz = 0; dz < 3; ++dz)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
              const int pz = hz - dz, py = hy - dy;
              if (pz < 0 || pz > 1 || py < 0 || py > 3) continue;
              const int tap = (dz * 3 + dy) * 3 + dx;
              acc[tap] = mfma(af[pz * 4 + py], br[i % (kAhead + 1)], acc[tap]);
            }
        }
  }
  // lane (sl, kq) holds co = co0 + 16 ct + 4 kq + i, ci = ci0 + 16 it + sl
#pragma unroll
  for (int tap = 0; tap < 27; ++tap)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int co = co0 + ct * 16 + 4 * kq + i, ci = ci0 + it * 16 + sl;
      atomicAdd(&g.dw[((long long)tap * g.cout + co) * g.cin + ci], acc[tap][i]);
    }
}

// Weight gradient of the cin == 1 convolution: dW[co][tap] (tap < 27, row
// length 32) += sum_v dY[v][co] X[v + tap].  FP32 vector math.
__global__ __launch_bounds__(256) void conv3d_wgrad_c1_kernel(Wgrad a) {
  const int co = threadIdx.x % 32;
  const int sub = threadIdx.x / 32;  // 8 voxel lanes per channel
  float acc[27];
  for (int t = 0; t < 27; ++t) acc[t] = 0.f;
  const long long nvox = (long long)a.n * a.d * a.h * a.w;
  for (long long v = (long long)blockIdx.x * 8 + sub; v < nvox; v += (long long)gridDim.x * 8) {
    const int x = v % a.w, y = (v / a.w) % a.h, z = (v / ((long long)a.w * a.h)) % a.d;
    const long long n = v / ((long long)a.w * a.h * a.d);
    const float g = bf2f(a.dy[v * a.cout + co]);
    if (g == 0.f) continue;
    for (int t = 0; t < 27; ++t) {
      const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
      if (zz >= 0 && zz < a.d && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w)
        acc[t] += g * bf2f(a.x[((n * a.d + zz) * a.h + yy) * a.w + xx]);
    }
  }
  for (int t = 0; t < 27; ++t) atomicAdd(&a.dw[co * 32 + t], acc[t]);
}

// The same weight gradient on 2-plane x 4-row x 64-voxel tiles: the block
// stages the tile's input halo (4 x 6 x 66 values, zero outside the volume) in
// LDS as FP32; thread (co = t % 32, tile row t / 32) walks its row's 64 voxels
// with the 3 x 3 x 3 input window in registers (9 new values per voxel, the
// unrolled walk renames the rest) and its 27 sums, one dY value per voxel
// (32 channels of one voxel = 64 contiguous bytes per 32 lanes).  Each block
// reduces its rows' sums in LDS and adds them with 27 x 32 FP32 atomics at the
// end.  W % 64, H % 4, D % 2.
constexpr int kC1X = 64;  // tile row length
__global__ __launch_bounds__(256) void conv3d_wgrad_c1t_kernel(Wgrad a, int tiles_x, int tiles_y,
                                                               int ntiles) {
  __shared__ float xs[4][6][kC1X + 2];
  const int t = threadIdx.x, co = t & 31, r = t >> 5, pz = r >> 2, py = r & 3;
  float acc[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = 0.f;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int tx = tile % tiles_x, rr = tile / tiles_x;
    const int y0 = (rr % tiles_y) * 4, zb = (rr / tiles_y) * 2, x0 = tx * kC1X;
    const int n = zb / a.d, z0 = zb - n * a.d;
    __syncthreads();  // the previous tile's window reads are done
    for (int i = t; i < 4 * 6 * (kC1X + 2); i += 256) {
      const int hz = i / (6 * (kC1X + 2)), hy = (i / (kC1X + 2)) % 6, hx = i % (kC1X + 2);
      const int zz = z0 - 1 + hz, yy = y0 - 1 + hy, xx = x0 - 1 + hx;
      const bool in = zz >= 0 && zz < a.d && yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
      xs[hz][hy][hx] = in ? bf2f(a.x[(((long long)n * a.d + zz) * a.h + yy) * a.w + xx]) : 0.f;
    }
    __syncthreads();
    const u16* dyr = a.dy + (((long long)(zb + pz) * a.h + y0 + py) * a.w + x0) * 32 + co;
    float win[3][3][3];  // win[dz][dy][dx] = X at (pz + dz - 1, py + dy - 1, x + dx - 1)
#pragma unroll
    for (int dz = 0; dz < 3; ++dz)
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        win[dz][dy][1] = xs[pz + dz][py + dy][0];
        win[dz][dy][2] = xs[pz + dz][py + dy][1];
      }
#pragma unroll
    for (int x = 0; x < kC1X; ++x) {
      const float g = bf2f(dyr[(long long)x * 32]);
#pragma unroll
      for (int dz = 0; dz < 3; ++dz)
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
          win[dz][dy][0] = win[dz][dy][1];
          win[dz][dy][1] = win[dz][dy][2];
          win[dz][dy][2] = xs[pz + dz][py + dy][x + 2];
#pragma unroll
          for (int dx = 0; dx < 3; ++dx) acc[(dz * 3 + dy) * 3 + dx] = fmaf(g, win[dz][dy][dx], acc[(dz * 3 + dy) * 3 + dx]);
        }
    }
  }
  // the 8 tile rows' sums meet in LDS first: 27 x 32 global atomics per block
  __shared__ float red[27 * 32];
  for (int i = t; i < 27 * 32; i += 256) red[i] = 0.f;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 27; ++k) atomicAdd(&red[k * 32 + co], acc[k]);
  __syncthreads();
  for (int i = t; i < 27 * 32; i += 256) atomicAdd(&a.dw[(i & 31) * 32 + (i >> 5)], red[i]);
}

// dW of a [27][cout][cin] forward weight -> backward-data weight
// Wb[t][ci][co] = W[26 - t][co][ci] (spatial flip, channel transpose).
__global__ void flip_weights_kernel(const u16* w, u16* wb, int cin, int cout) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= 27LL * cin * cout) return;
  const int co = i % cout, ci = (i / cout) % cin, t = i / ((long long)cout * cin);
  wb[i] = w[((26 - t) * (long long)cout + co) * cin + ci];
}

__global__ void bias_grad_kernel(const u16* dy, float* db, long long nvox, int c) {
  const int ch = threadIdx.x % c;
  const int lanes = 256 / c;
  const int sub = threadIdx.x / c;
  if (sub >= lanes) return;
  float s = 0.f;
  for (long long v = (long long)blockIdx.x * lanes + sub; v < nvox; v += (long long)gridDim.x * lanes)
    s += bf2f(dy[v * c + ch]);
  atomicAdd(&db[ch], s);
}

// Adjoint of the activation at output y (mode 1 ReLU, 2 leaky ReLU: y > 0
// exactly where the input was).
__device__ __forceinline__ u16 relu_bwd1(u16 dy, u16 y, int mode) {
  return (bf2f(y) > 0.f) ? dy : mode == 2 ? f2bf(kLeak * bf2f(dy)) : (u16)0;
}

__global__ void relu_bwd_kernel(const u16* dy, const u16* y, u16* dz, long long n, int mode) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    dz[i] = relu_bwd1(dy[i], y[i], mode);
}

// The same, 8 values per 16-byte vector (n % 8 == 0, 16-byte aligned).
__global__ void relu_bwd_v8_kernel(const uint4* dy, const uint4* y, uint4* dz, long long n8, int mode) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n8; i += (long long)gridDim.x * 256) {
    const uint4 g = dy[i], v = y[i];
    const unsigned gw[4] = {g.x, g.y, g.z, g.w}, vw[4] = {v.x, v.y, v.z, v.w};
    unsigned o[4];
#pragma unroll
    for (int e = 0; e < 4; ++e)
      o[e] = (unsigned)relu_bwd1((u16)(gw[e] & 0xFFFF), (u16)(vw[e] & 0xFFFF), mode) |
             ((unsigned)relu_bwd1((u16)(gw[e] >> 16), (u16)(vw[e] >> 16), mode) << 16);
    dz[i] = make_uint4(o[0], o[1], o[2], o[3]);
  }
}

// 2x2x2 average pooling (and its adjoint) on [n][d][h][w][c].
__global__ void avgpool_kernel(const u16* x, u16* y, int n, int d, int h, int w, int c) {
  const long long total = (long long)n * (d / 2) * (h / 2) * (w / 2) * c;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int ch = i % c;
    long long v = i / c;
    const int xo = v % (w / 2); v /= (w / 2);
    const int yo = v % (h / 2); v /= (h / 2);
    const int zo = v % (d / 2);
    const long long nn = v / (d / 2);
    float s = 0.f;
    for (int q = 0; q < 8; ++q) {
      const int zz = 2 * zo + (q >> 2), yy = 2 * yo + ((q >> 1) & 1), xx = 2 * xo + (q & 1);
      s += bf2f(x[(((nn * d + zz) * h + yy) * w + xx) * c + ch]);
    }
    y[i] = f2bf(0.125f * s);
  }
}

// 16-byte form of avgpool_kernel (c % 8 == 0): block (x-run, yo, n*Do+zo),
// one thread per output voxel and 8-channel chunk.
__global__ void avgpool_v8_kernel(const uint4* x, uint4* y, int d, int h, int w, int cq) {
  const int Do = d / 2, Ho = h / 2, Wo = w / 2;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= Wo * cq) return;
  const int xo = idx / cq, ch = idx - xo * cq;
  const int yo = blockIdx.y;
  const int nz = blockIdx.z, zo = nz % Do, nn = nz / Do;
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int zz = 2 * zo + (q >> 2), yy = 2 * yo + ((q >> 1) & 1), xx = 2 * xo + (q & 1);
    const uint4 v = x[((((long long)nn * d + zz) * h + yy) * w + xx) * cq + ch];
    const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      acc[2 * e] += __uint_as_float(u[e] << 16);
      acc[2 * e + 1] += __uint_as_float(u[e] & 0xffff0000u);
    }
  }
  unsigned o[4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
    o[e] = (unsigned)f2bf(0.125f * acc[2 * e]) | ((unsigned)f2bf(0.125f * acc[2 * e + 1]) << 16);
  y[(((long long)nz * Ho + yo) * Wo + xo) * cq + ch] = make_uint4(o[0], o[1], o[2], o[3]);
}

__global__ void avgpool_bwd_kernel(const u16* dy, u16* dx, int n, int d, int h, int w, int c) {
  const long long total = (long long)n * d * h * w * c;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int ch = i % c;
    long long v = i / c;
    const int x = v % w; v /= w;
    const int y = v % h; v /= h;
    const int z = v % d;
    const long long nn = v / d;
    const long long o = (((nn * (d / 2) + z / 2) * (h / 2) + y / 2) * (w / 2) + x / 2) * c + ch;
    dx[i] = f2bf(0.125f * bf2f(dy[o]));
  }
}

// 8 BF16 values of a 16-byte vector as floats, and back (round to nearest even).
__device__ __forceinline__ void unpack8(const uint4& q, float (&f)[8]) {
  const unsigned u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    f[2 * e] = __uint_as_float(u[e] << 16);
    f[2 * e + 1] = __uint_as_float(u[e] & 0xFFFF0000u);
  }
}
__device__ __forceinline__ uint4 pack8(const float (&f)[8]) {
  unsigned u[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) u[e] = (unsigned)f2bf(f[2 * e]) | ((unsigned)f2bf(f[2 * e + 1]) << 16);
  return make_uint4(u[0], u[1], u[2], u[3]);
}

// avgpool_bwd_kernel with 8 channels per 16-byte vector (the same values):
// thread (X, 8-channel group) of fine row (blockIdx.z = n d + Z, blockIdx.y = Y).
__global__ void avgpool_bwd_v8_kernel(const uint4* dy, uint4* dx, int d, int h, int w, int cq) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= w * cq) return;
  const int X = idx / cq, ch = idx - X * cq;
  const int Y = blockIdx.y, nz = blockIdx.z, Z = nz % d, nn = nz / d;
  const long long src = (((long long)(nn * (d / 2) + Z / 2) * (h / 2) + Y / 2) * (w / 2) + X / 2) * cq + ch;
  float f[8];
  unpack8(dy[src], f);
#pragma unroll
  for (int e = 0; e < 8; ++e) f[e] *= 0.125f;
  dx[(((long long)nz * h + Y) * w + X) * cq + ch] = pack8(f);
}

// Nearest 2x upsampling of x ([n][d][h][w][cx]) into channels [c0, c0+cx) of
// y ([n][2d][2h][2w][cy]); the adjoint sums the 8 children.
__global__ void upsample_kernel(const u16* x, u16* y, int n, int d, int h, int w, int cx, int cy,
                                int c0) {
  const long long total = (long long)n * 8 * d * h * w * cx;
  const int D = 2 * d, H = 2 * h, Wd = 2 * w;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int ch = i % cx;
    long long v = i / cx;
    const int X = v % Wd; v /= Wd;
    const int Y = v % H; v /= H;
    const int Z = v % D;
    const long long nn = v / D;
    y[(((nn * D + Z) * H + Y) * Wd + X) * cy + c0 + ch] =
        x[(((nn * d + Z / 2) * h + Y / 2) * w + X / 2) * cx + ch];
  }
}

// 16-byte form of upsample_kernel (cx, cy, c0 multiples of 8): block
// (X-run, Y, n*D+Z), one thread per output voxel and 8-channel chunk.
__global__ void upsample_v8_kernel(const uint4* x, uint4* y, int d, int h, int w, int cq, int cyq,
                                   int c0q) {
  const int D = 2 * d, H = 2 * h, Wd = 2 * w;
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= Wd * cq) return;
  const int X = idx / cq, ch = idx - X * cq;
  const int Y = blockIdx.y;
  const int nz = blockIdx.z, Z = nz % D, nn = nz / D;
  const long long src = (((long long)(nn * d + Z / 2) * h + Y / 2) * w + X / 2) * cq + ch;
  const long long dst = (((long long)nz * H + Y) * Wd + X) * cyq + c0q + ch;
  y[dst] = x[src];
}

// upsample_bwd_kernel with 8 channels per 16-byte vector (the same sums, in the
// same order): thread (x, 8-channel group) of coarse row (blockIdx.z = n d + z,
// blockIdx.y = y).
__global__ void upsample_bwd_v8_kernel(const uint4* dy, uint4* dx, int d, int h, int w, int cq,
                                       int cyq, int c0q) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= w * cq) return;
  const int x = idx / cq, ch = idx - x * cq;
  const int y = blockIdx.y, nz = blockIdx.z, z = nz % d, nn = nz / d;
  const int D = 2 * d, H = 2 * h, Wd = 2 * w;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int Z = 2 * z + (q >> 2), Y = 2 * y + ((q >> 1) & 1), X = 2 * x + (q & 1);
    float f[8];
    unpack8(dy[(((long long)(nn * D + Z) * H + Y) * Wd + X) * cyq + c0q + ch], f);
#pragma unroll
    for (int e = 0; e < 8; ++e) s[e] += f[e];
  }
  dx[(((long long)nz * h + y) * w + x) * cq + ch] = pack8(s);
}

__global__ void upsample_bwd_kernel(const u16* dy, u16* dx, int n, int d, int h, int w, int cx,
                                    int cy, int c0) {
  const long long total = (long long)n * d * h * w * cx;
  const int D = 2 * d, H = 2 * h, Wd = 2 * w;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int ch = i % cx;
    long long v = i / cx;
    const int x = v % w; v /= w;
    const int y = v % h; v /= h;
    const int z = v % d;
    const long long nn = v / d;
    float s = 0.f;
    for (int q = 0; q < 8; ++q) {
      const int Z = 2 * z + (q >> 2), Y = 2 * y + ((q >> 1) & 1), X = 2 * x + (q & 1);
      s += bf2f(dy[(((nn * D + Z) * H + Y) * Wd + X) * cy + c0 + ch]);
    }
    dx[i] = f2bf(s);
  }
}

// Copy channels [src_c0, src_c0+cc) of a [nvox][cs] tensor into channels
// [dst_c0, dst_c0+cc) of a [nvox][cd] tensor (concatenation and its adjoint).
__global__ void channel_copy_kernel(const u16* src, int cs, int src_c0, u16* dst, int cd,
                                    int dst_c0, int cc, long long nvox) {
  const long long total = nvox * cc;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const long long v = i / cc;
    const int ch = i % cc;
    dst[v * cd + dst_c0 + ch] = src[v * cs + src_c0 + ch];
  }
}

// 16-byte form of channel_copy_kernel (all channel counts and offsets
// multiples of 8, nvox * cc / 8 < 2^31).
__global__ void channel_copy_v8_kernel(const uint4* src, int csq, int s0q, uint4* dst, int cdq,
                                       int d0q, int ccq, int total) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const int v = i / ccq, ch = i - v * ccq;
  dst[(long long)v * cdq + d0q + ch] = src[(long long)v * csq + s0q + ch];
}

// Output head: out = in + sum_c w[c] h[c] + b (FP32), and its adjoints.
__global__ void head_fwd_kernel(const u16* h, const float* w, const float* b, const float* in,
                                float* out, long long nvox, int c) {
  for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < nvox; v += (long long)gridDim.x * 256) {
    float s = b[0];
    for (int ch = 0; ch < c; ++ch) s = fmaf(w[ch], bf2f(h[v * c + ch]), s);
    out[v] = in[v] + s;
  }
}

// 16-byte form of head_fwd_kernel for C channels (C % 8 == 0): one voxel per
// thread, its C BF16 activations in C / 8 loads; w and b are wave-uniform.
template <int C>
__global__ __launch_bounds__(256) void head_fwd_v8_kernel(const uint4* h, const float* w,
                                                          const float* b, const float* in,
                                                          float* out, long long nvox) {
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvox) return;
  float s = b[0];
#pragma unroll
  for (int q = 0; q < C / 8; ++q) {
    const uint4 x = h[v * (C / 8) + q];
    const unsigned u[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      s = fmaf(w[8 * q + 2 * e], __uint_as_float(u[e] << 16), s);
      s = fmaf(w[8 * q + 2 * e + 1], __uint_as_float(u[e] & 0xffff0000u), s);
    }
  }
  out[v] = in[v] + s;
}

__global__ void head_bwd_kernel(const float* dout, const u16* h, const float* w, u16* dh,
                                float* dw, float* db, long long nvox, int c) {
  __shared__ float red[64];
  for (int i = threadIdx.x; i < 64; i += 256) red[i] = 0.f;
  __syncthreads();
  float acc[64];
  for (int ch = 0; ch < c; ++ch) acc[ch] = 0.f;
  float accb = 0.f;
  for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < nvox; v += (long long)gridDim.x * 256) {
    const float g = dout[v];
    accb += g;
    for (int ch = 0; ch < c; ++ch) {
      acc[ch] = fmaf(g, bf2f(h[v * c + ch]), acc[ch]);
      dh[v * c + ch] = f2bf(g * w[ch]);
    }
  }
  for (int ch = 0; ch < c; ++ch) atomicAdd(&red[ch], acc[ch]);
  atomicAdd(&red[c], accb);
  __syncthreads();
  for (int i = threadIdx.x; i < c; i += 256) atomicAdd(&dw[i], red[i]);
  if (threadIdx.x == 0) atomicAdd(db, red[c]);
}

// head_bwd_kernel for C channels with 16-byte vectors: one 16-byte chunk (8
// features of one voxel) per thread, consecutive threads on consecutive
// chunks; the weight-gradient sums go through LDS atomics per block and one
// global atomic per channel and block.
template <int C>
__global__ __launch_bounds__(256) void head_bwd_v8_kernel(const float* dout, const uint4* h,
                                                          const float* w, uint4* dh, float* dw,
                                                          float* db, long long nvox) {
  constexpr int G = C / 8;  // 16-byte chunks per voxel; thread i: chunk i % G of voxel i / G
  __shared__ float red[C + 1];
  for (int i = threadIdx.x; i < C + 1; i += 256) red[i] = 0.f;
  __syncthreads();
  const int q = threadIdx.x % G;  // fixed: 256 and the grid stride are multiples of G
  float wq[8], acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    wq[e] = w[8 * q + e];
    acc[e] = 0.f;
  }
  float accb = 0.f;
  const long long nch = nvox * G;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nch; i += (long long)gridDim.x * 256) {
    const float g = dout[i / G];
    if (q == 0) accb += g;
    float f[8], o[8];
    unpack8(h[i], f);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      acc[e] = fmaf(g, f[e], acc[e]);
      o[e] = g * wq[e];
    }
    dh[i] = pack8(o);
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) atomicAdd(&red[8 * q + e], acc[e]);
  if (q == 0) atomicAdd(&red[C], accb);
  __syncthreads();
  for (int i = threadIdx.x; i < C; i += 256) atomicAdd(&dw[i], red[i]);
  if (threadIdx.x == 0) atomicAdd(db, red[C]);
}

// bias_grad_kernel for C % 8 == 0 channels with 16-byte loads: thread t sums
// channels 8 (t % (C / 8)) .. + 7 over every (256 / (C / 8))-th voxel of the
// block's share; partial sums meet in LDS, then one atomic per channel.
template <int C>
__global__ __launch_bounds__(256) void bias_grad_v8_kernel(const uint4* dy, float* db, long long nvox) {
  constexpr int G = C / 8, L = 256 / G;  // channel groups, voxel lanes
  __shared__ float red[C];
  for (int i = threadIdx.x; i < C; i += 256) red[i] = 0.f;
  __syncthreads();
  const int g = threadIdx.x % G, l = threadIdx.x / G;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (long long v = (long long)blockIdx.x * L + l; v < nvox; v += (long long)gridDim.x * L) {
    const uint4 q = dy[v * G + g];
    const unsigned u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      s[2 * e] += __uint_as_float(u[e] << 16);
      s[2 * e + 1] += __uint_as_float(u[e] & 0xFFFF0000u);
    }
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) atomicAdd(&red[8 * g + e], s[e]);
  __syncthreads();
  for (int i = threadIdx.x; i < C; i += 256) atomicAdd(&db[i], red[i]);
}

// Space-to-depth at the network's ends, block edge F (2 or 4).  An FP32 volume
// [n][F d][F h][F w] maps to the coarse grid [n][d][h][w][cy] BF16: voxel
// (z, y, x) holds the F^3 block (F z + iz, F y + iy, F x + ix) as channels
// (iz F + iy) F + ix, and channels F^3..cy-1 are zero.  One thread per coarse
// voxel; each block row of F samples is one F x 4-byte load.
__device__ inline unsigned pack_bf2(float lo, float hi) {
  return (unsigned)f2bf(lo) | ((unsigned)f2bf(hi) << 16);
}

// Index of the fine sample (iz, iy, 0) of coarse voxel v, and the fine row and
// plane strides.
struct S2dMap {
  long long o0, W, HW;
};

template <int F>
__device__ inline S2dMap s2d_map(long long v, int hc, int wc) {
  const int xx = (int)(v % wc);
  const long long r = v / wc;  // (n d + z) h + y of the coarse voxel
  const int yy = (int)(r % hc);
  const long long nz = r / hc;
  S2dMap m;
  m.W = (long long)F * wc;
  m.HW = (long long)F * hc * m.W;
  m.o0 = (F * nz) * m.HW + (long long)F * yy * m.W + (long long)F * xx;
  return m;
}

template <int F>
__device__ inline void row_load(const float* p, float (&r)[F]) {
  if constexpr (F == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    r[0] = t.x;
    r[1] = t.y;
  } else {
    const float4 t = *reinterpret_cast<const float4*>(p);
    r[0] = t.x;
    r[1] = t.y;
    r[2] = t.z;
    r[3] = t.w;
  }
}

template <int F>
__device__ inline void row_store(float* p, const float (&r)[F]) {
  if constexpr (F == 2) *reinterpret_cast<float2*>(p) = make_float2(r[0], r[1]);
  else *reinterpret_cast<float4*>(p) = make_float4(r[0], r[1], r[2], r[3]);
}

template <int F>
__global__ __launch_bounds__(256) void unshuffle_kernel(const float* x, uint4* y, int hc, int wc,
                                                        int cyq, long long nvox) {
  constexpr int K = F * F * F;
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvox) return;
  const S2dMap mp = s2d_map<F>(v, hc, wc);
  float q[K];
#pragma unroll
  for (int r = 0; r < F * F; ++r) {  // r = iz F + iy
    float row[F];
    row_load<F>(x + mp.o0 + (r / F) * mp.HW + (r % F) * mp.W, row);
#pragma unroll
    for (int i = 0; i < F; ++i) q[r * F + i] = row[i];
  }
  uint4* o = y + v * cyq;
#pragma unroll
  for (int j = 0; j < K / 8; ++j)
    o[j] = make_uint4(pack_bf2(q[8 * j], q[8 * j + 1]), pack_bf2(q[8 * j + 2], q[8 * j + 3]),
                      pack_bf2(q[8 * j + 4], q[8 * j + 5]), pack_bf2(q[8 * j + 6], q[8 * j + 7]));
  for (int j = K / 8; j < cyq; ++j) o[j] = make_uint4(0u, 0u, 0u, 0u);
}

template <int C>
__device__ inline void load_bf_row(const uint4* h, long long v, float (&hv)[C]) {
#pragma unroll
  for (int q = 0; q < C / 8; ++q) {
    const uint4 t = h[v * (C / 8) + q];
    const unsigned u[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      hv[8 * q + 2 * e] = __uint_as_float(u[e] << 16);
      hv[8 * q + 2 * e + 1] = __uint_as_float(u[e] & 0xffff0000u);
    }
  }
}

// Depth-to-space output head: for coarse voxel v, K = F^3 outputs
// s_k = b[k] + sum_c w[k][c] h[v][c] land on its block, k = (iz F + iy) F + ix:
// out = in + s.  C == 32; w, b staged in LDS.
template <int F, int C>
__global__ __launch_bounds__(256) void head_shuffle_kernel(const uint4* h, const float* w,
                                                           const float* b, const float* in,
                                                           float* out, int hc, int wc,
                                                           long long nvox) {
  constexpr int K = F * F * F;
  __shared__ float ws[K * C + K];
  for (int i = threadIdx.x; i < K * C; i += 256) ws[i] = w[i];
  if (threadIdx.x < K) ws[K * C + threadIdx.x] = b[threadIdx.x];
  __syncthreads();
  const long long v = (long long)blockIdx.x * 256 + threadIdx.x;
  if (v >= nvox) return;
  float hv[C];
  load_bf_row<C>(h, v, hv);
  const S2dMap mp = s2d_map<F>(v, hc, wc);
#pragma unroll
  for (int r = 0; r < F * F; ++r) {
    const long long o = mp.o0 + (r / F) * mp.HW + (r % F) * mp.W;
    float row[F];
    row_load<F>(in + o, row);
#pragma unroll
    for (int i = 0; i < F; ++i) {
      const int k = r * F + i;
      float acc = ws[K * C + k];
#pragma unroll
      for (int c = 0; c < C; ++c) acc = fmaf(ws[k * C + c], hv[c], acc);
      row[i] += acc;
    }
    row_store<F>(out + o, row);
  }
}

// Adjoint of head_shuffle_kernel: dh[v][c] = sum_k w[k][c] g_k (BF16),
// dw[k][c] += sum_v g_k h[v][c], db[k] += sum_v g_k, with g_k = dout at block
// entry k of voxel v.  A block walks tiles of TV voxels; each tile's g and h go
// through LDS and thread t accumulates dw[k][t % C] for k = t / C + 8 e,
// e < K / 8.  C == 32.
template <int F, int C>
__global__ __launch_bounds__(256) void head_shuffle_bwd_kernel(const float* dout, const uint4* h,
                                                               const float* w, uint4* dh,
                                                               float* dw, float* db, int hc,
                                                               int wc, long long nvox) {
  constexpr int K = F * F * F, E = K / 8, TV = F == 2 ? 256 : 64;
  static_assert(8 * C == 256, "8 k-groups of C threads");
  __shared__ float ws[K * C];
  __shared__ float gs[TV][K + 1];
  __shared__ float hs[TV][C + 1];
  for (int i = threadIdx.x; i < K * C; i += 256) ws[i] = w[i];
  __syncthreads();
  const int t = threadIdx.x, kg = t / C, cc = t % C;
  float acc[E], accb = 0.f;
#pragma unroll
  for (int e = 0; e < E; ++e) acc[e] = 0.f;
  for (long long tile = (long long)blockIdx.x * TV; tile < nvox; tile += (long long)gridDim.x * TV) {
    if (t < TV) {
      const long long v = tile + t;
      float g[K], hv[C];
#pragma unroll
      for (int k = 0; k < K; ++k) g[k] = 0.f;
#pragma unroll
      for (int c = 0; c < C; ++c) hv[c] = 0.f;
      if (v < nvox) {
        const S2dMap mp = s2d_map<F>(v, hc, wc);
#pragma unroll
        for (int r = 0; r < F * F; ++r) {
          float row[F];
          row_load<F>(dout + mp.o0 + (r / F) * mp.HW + (r % F) * mp.W, row);
#pragma unroll
          for (int i = 0; i < F; ++i) g[r * F + i] = row[i];
        }
        load_bf_row<C>(h, v, hv);
#pragma unroll
        for (int q = 0; q < C / 8; ++q) {
          float d8[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            float a = 0.f;
#pragma unroll
            for (int k = 0; k < K; ++k) a = fmaf(ws[k * C + 8 * q + e], g[k], a);
            d8[e] = a;
          }
          dh[v * (C / 8) + q] = make_uint4(pack_bf2(d8[0], d8[1]), pack_bf2(d8[2], d8[3]),
                                           pack_bf2(d8[4], d8[5]), pack_bf2(d8[6], d8[7]));
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) gs[t][k] = g[k];
#pragma unroll
      for (int c = 0; c < C; ++c) hs[t][c] = hv[c];
    }
    __syncthreads();
    for (int j = 0; j < TV; ++j) {
      const float hj = hs[j][cc];
#pragma unroll
      for (int e = 0; e < E; ++e) acc[e] = fmaf(gs[j][kg + 8 * e], hj, acc[e]);
    }
    if (t < K)
      for (int j = 0; j < TV; ++j) accb += gs[j][t];
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < E; ++e) atomicAdd(&dw[(kg + 8 * e) * C + cc], acc[e]);
  if (t < K) atomicAdd(&db[t], accb);
}

// head_shuffle_bwd_kernel<4, 32> with every thread on the feature gradients:
// per 64-voxel tile the block stages the 64 output gradients (16 rows of 4 fine
// samples each) and the 32 features of every voxel in LDS, then thread t
// computes the 8 feature gradients 8 (t % 4) .. + 7 of voxel t / 4 (64 FMAs
// each, k ascending as in the one-thread-per-voxel form, so the same values)
// before the weight-gradient sums of head_shuffle_bwd_kernel.
template <int C>
__global__ __launch_bounds__(256) void head_shuffle4_bwd_kernel(const float* dout, const uint4* h,
                                                                const float* w, uint4* dh,
                                                                float* dw, float* db, int hc,
                                                                int wc, long long nvox) {
  constexpr int F = 4, K = 64, TV = 64, E = K / 8, G = C / 8;
  static_assert(8 * C == 256 && TV * G == 256, "8 k-groups of C threads; one chunk per thread");
  __shared__ float ws[K * C];
  __shared__ float gs[TV][K + 1];
  __shared__ float hs[TV][C + 1];
  for (int i = threadIdx.x; i < K * C; i += 256) ws[i] = w[i];
  const int t = threadIdx.x, kg = t / C, cc = t % C;
  const int jv = t / G, q = t % G;  // feature-gradient phase: voxel jv, channel chunk q
  float acc[E], accb = 0.f;
#pragma unroll
  for (int e = 0; e < E; ++e) acc[e] = 0.f;
  for (long long tile = (long long)blockIdx.x * TV; tile < nvox; tile += (long long)gridDim.x * TV) {
    __syncthreads();  // the previous tile's reads are done (and ws is written)
#pragma unroll
    for (int i = 0; i < 4; ++i) {  // rows r = t % 16 of voxels t / 16 + 16 i
      const int j = (t >> 4) + 16 * i, r = t & 15;
      const long long v = tile + j;
      float row[F] = {0.f, 0.f, 0.f, 0.f};
      if (v < nvox) {
        const S2dMap mp = s2d_map<F>(v, hc, wc);
        row_load<F>(dout + mp.o0 + (r / F) * mp.HW + (r % F) * mp.W, row);
      }
#pragma unroll
      for (int e = 0; e < F; ++e) gs[j][r * F + e] = row[e];
    }
    {
      const long long v = tile + jv;
      float f[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (v < nvox) unpack8(h[v * G + q], f);
#pragma unroll
      for (int e = 0; e < 8; ++e) hs[jv][8 * q + e] = f[e];
    }
    __syncthreads();
    {
      float d8[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 8
      for (int k = 0; k < K; ++k) {
        const float g = gs[jv][k];
#pragma unroll
        for (int e = 0; e < 8; ++e) d8[e] = fmaf(ws[k * C + 8 * q + e], g, d8[e]);
      }
      if (tile + jv < nvox) dh[(tile + jv) * G + q] = pack8(d8);
    }
    for (int j = 0; j < TV; ++j) {
      const float hj = hs[j][cc];
#pragma unroll
      for (int e = 0; e < E; ++e) acc[e] = fmaf(gs[j][kg + 8 * e], hj, acc[e]);
    }
    if (t < K)
      for (int j = 0; j < TV; ++j) accb += gs[j][t];
  }
#pragma unroll
  for (int e = 0; e < E; ++e) atomicAdd(&dw[(kg + 8 * e) * C + cc], acc[e]);
  if (t < K) atomicAdd(&db[t], accb);
}

// Mean-squared error: loss_sum += sum (out - t)^2, dout = 2 (out - t) / nvox.
__global__ void mse_kernel(const float* out, const float* t, float* dout, float* loss_sum,
                           long long nvox) {
  float s = 0.f;
  for (long long v = (long long)blockIdx.x * 256 + threadIdx.x; v < nvox; v += (long long)gridDim.x * 256) {
    const float e = out[v] - t[v];
    s += e * e;
    if (dout) dout[v] = 2.0f * e / (float)nvox;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss_sum, s);
}

// Adam on FP32 master weights; also refreshes the BF16 copy the MFMA reads
// (layout given by `bf_stride`: the BF16 copy may pad rows, e.g. the cin == 1
// conv's 27 taps padded to 32; index i maps to (i / row) * bf_row + i % row).
__global__ void adam_kernel(float* w, u16* wbf, const float* g, float* m, float* v, long long n,
                            float lr, float b1, float b2, float eps, float c1, float c2, int row,
                            int bf_row) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wbf) wbf[(i / row) * bf_row + i % row] = f2bf(wi);
  }
}

__global__ void f32_to_bf16_kernel(const float* s, u16* d, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    d[i] = f2bf(s[i]);
}

__global__ void bf16_to_f32_kernel(const u16* s, float* d, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    d[i] = bf2f(s[i]);
}

// Box copy between [D][H][W] FP32 volumes: dst[z][y][x] = src[z+sz][y+sy][x+sx]
// for the (nz, ny, nx) box; a channel-1 activation is the same array.
__global__ void box_copy_kernel(const float* src, int sD, int sH, int sW, int sz, int sy, int sx,
                                float* dst, int dD, int dH, int dW, int dz, int dy, int dx, int nz,
                                int ny, int nx, float scale) {
  const long long total = (long long)nz * ny * nx;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int x = i % nx, y = (i / nx) % ny, z = i / ((long long)nx * ny);
    dst[((long long)(z + dz) * dH + (y + dy)) * dW + (x + dx)] =
        scale * src[((long long)(z + sz) * sH + (y + sy)) * sW + (x + sx)];
  }
}

// Photon noise on line integrals: counts ~ Poisson(I0 exp(-mu p)),
// p <- -ln(max(counts, 1) / I0) / mu.  Counter-based hash RNG per (seed, i, draw),
// so a seed reproduces every pixel whatever the launch shape.  The counts are
// exact Poisson variates: below 10 expected counts by multiplying uniforms
// until the product falls under e^-lambda (Knuth's method), from 10 up by
// W. Hoermann's transformed rejection with squeeze, PTRS ("The transformed
// rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12, 1993), with the acceptance test in FP64 (at
// lambda = 1e5 its log terms are ~1e6 and cancel to O(1)).
__device__ inline unsigned hash32(unsigned x) {
  x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
  return x;
}
// Uniform in (0, 1) from 32 hashed bits: draw k of pixel i.
__device__ inline double unif(unsigned seed, unsigned long long i, unsigned k) {
  const unsigned h = hash32(hash32(seed ^ (unsigned)(i * 0x9E3779B97F4A7C15ULL >> 32)) ^
                            hash32((unsigned)i + 0x632BE5ABu * k));
  return ((double)h + 0.5) * (1.0 / 4294967296.0);
}

__device__ double poisson_small(double lam, unsigned seed, unsigned long long i) {
  const double L = exp(-lam);
  double prod = unif(seed, i, 0);
  int c = 0;
  while (prod > L && c < 1000) {  // P(k > 1000) is 0 in double for lambda < 10
    ++c;
    prod *= unif(seed, i, (unsigned)c);
  }
  return (double)c;
}

__device__ double poisson_ptrs(double lam, unsigned seed, unsigned long long i) {
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double log_inv_alpha = log(1.1239 + 1.1328 / (b - 3.4));
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  double k = floor(lam);
  for (unsigned t = 0; t < 128; ++t) {  // each try accepts with probability > 0.9
    const double U = unif(seed, i, 2 * t + 1) - 0.5;
    const double V = unif(seed, i, 2 * t + 2);
    const double us = 0.5 - fabs(U);
    k = floor((2.0 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;  // squeeze: inside the hat's core
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log_inv_alpha - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
      return k;
  }
  return fmax(k, 0.0);
}

__global__ void poisson_kernel(float* p, long long n, double i0, double mu, unsigned seed) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double lam = i0 * exp(-mu * (double)p[i]);
    if (!(lam >= 0.0 && lam < 1e15)) continue;  // NaN or absurd counts: left as they are
    const double k = lam < 10.0 ? poisson_small(lam, seed, (unsigned long long)i)
                                : poisson_ptrs(lam, seed, (unsigned long long)i);
    p[i] = (float)(-log(fmax(k, 1.0) / i0) / mu);
  }
}

// Layout probe of v_mfma_f32_16x16x32_bf16: D = A (16x32) * B (32x16) with
// A, B given row-major in FP32 (rounded to BF16), D row-major.
__global__ void mfma_probe_kernel(const float* A, const float* B, float* D) {
  const int l = threadIdx.x, m = l % 16, kq = l / 16;
  bf16x8 a, b;
  for (int j = 0; j < 8; ++j) {
    a[j] = (__bf16)A[m * 32 + kq * 8 + j];
    b[j] = (__bf16)B[(kq * 8 + j) * 16 + m];
  }
  const f32x4 d = mfma(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
  for (int i = 0; i < 4; ++i) D[(4 * kq + i) * 16 + m] = d[i];
}

int blocks_for(long long n) { return (int)std::min<long long>(4096, std::max<long long>(1, (n + 255) / 256)); }

}  // namespace

extern "C" int fdk_nn_mfma_probe(const float* A, const float* B, float* D, void* stream) {
  FDK_REQUIRE(A && B && D, "NULL pointer");
  hipLaunchKernelGGL(mfma_probe_kernel, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), A, B, D);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

// Forward convolution kernel: the row-reuse LDS-halo kernel (5, default), the
// lean LDS-halo kernel (4; both cout 32 / 64 passes) or the LDS-halo kernel
// where their shape rules hold, else the pipelined 2-row tile, else the 1-row
// kernel; FDK_NN_CONV=4, 3, 2 or 1 caps the choice.
static int conv_kernel_choice() {
  static const int k = [] {
    const char* e = std::getenv("FDK_NN_CONV");
    if (e && e[0] == '1') return 1;
    if (e && e[0] == '2') return 2;
    if (e && e[0] == '3') return 3;
    if (e && e[0] == '4') return 4;
    return 5;
  }();
  return k;
}

static bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

// Shapes and alignments the LDS-halo kernels take (cout a multiple of 32, whole
// 32-channel chunks, even depth, 16-byte vectors, 32-bit staging offsets).
static bool lean_ok(const Conv& a) {
  const int cx = a.cin - a.cup;
  return a.cin % 32 == 0 && a.cup % 32 == 0 && a.d % 2 == 0 && a.cout % 32 == 0 &&
         aligned16(a.x) && aligned16(a.wt) && aligned16(a.y) && (!a.bias || aligned16(a.bias)) &&
         (!a.resid || aligned16(a.resid)) && (!a.up || aligned16(a.up)) &&
         ((4LL * a.h * a.w + 2LL * a.w) * std::max(cx, a.cup) < (1LL << 31) - 8) &&
         27LL * a.cout * a.cin * 2 < (1LL << 31);
}

template <int TP, bool PF>
static int launch_fwd5_pass(const Conv& a, hipStream_t st) {
  const void* fn = reinterpret_cast<const void*>(&conv3d_fwd5_kernel<2, 4, TP, PF>);
  const int lds = HaloShape<TP, 4>::chunks * (int)sizeof(uint4);
  FDK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
  const int tiles_x = a.w / 32, tiles_y = a.h / 4, ntiles = tiles_x * tiles_y * (a.n * a.d / TP);
  const int tpb = 2;
  hipLaunchKernelGGL((conv3d_fwd5_kernel<2, 4, TP, PF>), dim3((ntiles + tpb - 1) / tpb), dim3(256), lds,
                     st, a, tiles_x, tiles_y, ntiles, tpb);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

// conv3d_fwd5_kernel where its tile fits: 64-channel passes on 2-plane tiles
// for the first 64 floor(cout / 64) channels, a 32-channel pass on 4-plane tiles
// (d % 4 == 0) for the rest (cout 32, 96, ...); h % 4 == 0; its halo resource
// ((TP + 2) planes + 2 rows) below 2^32 - 16 bytes.
static bool fwd5_ok(const Conv& a) {
  const int cx = a.cin - a.cup;
  return conv_kernel_choice() >= 5 && a.cout % 32 == 0 && (a.cout % 64 == 0 || a.d % 4 == 0) &&
         a.h % 4 == 0 && (6LL * a.h * a.w + 2LL * a.w) * std::max(cx, a.cup) < (1LL << 31) - 8;
}

static int launch_fwd5(const Conv& a, hipStream_t st) {
  Conv p = a;
  for (p.co0 = 0; p.co0 + 64 <= a.cout; p.co0 += 64)
    if (int r = launch_fwd5_pass<2, true>(p, st)) return r;
  if (p.co0 < a.cout) return launch_fwd5_pass<4, true>(p, st);  // the last 32 channels
  return FDK_OK;
}

static int launch_fwd4(const Conv& a, hipStream_t st) {
  if (fwd5_ok(a)) return launch_fwd5(a, st);
  const int tiles_x = a.w / 32, tiles_y = a.h / 4, ntiles = tiles_x * tiles_y * (a.n * a.d / 2);
  const int tpb = 2;
  const dim3 g4((ntiles + tpb - 1) / tpb);
  // 64 output channels per pass (cout 128 and 192 re-stage the halo per pass,
  // which costs less than the wider accumulator set would in registers), then
  // a 32-channel pass for cout 32, 96, ...
  Conv p = a;
  for (p.co0 = 0; p.co0 + 64 <= a.cout; p.co0 += 64)
    hipLaunchKernelGGL((conv3d_fwd4_kernel<4, false>), g4, dim3(256), 0, st, p, tiles_x, tiles_y, ntiles, tpb);
  if (p.co0 < a.cout)
    hipLaunchKernelGGL((conv3d_fwd4_kernel<2, true>), g4, dim3(256), 0, st, p, tiles_x, tiles_y, ntiles, tpb);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_conv3d_fwd(const void* x, const void* w, const float* bias, const float* resid,
                                 void* y, int n, int d, int h, int wd, int cin, int cout, int relu,
                                 int out_f32, void* stream) {
  FDK_REQUIRE(x && w && y, "NULL pointer");
  FDK_REQUIRE(n > 0 && d > 0 && h > 0 && wd > 0, "bad tensor size");
  FDK_REQUIRE(wd % 32 == 0 && h % 4 == 0, "conv3d needs W %% 32 == 0 and H %% 4 == 0 (got %d, %d)", wd, h);
  FDK_REQUIRE(cin == 1 || cin % 32 == 0, "cin must be 1 or a multiple of 32 (got %d)", cin);
  FDK_REQUIRE((long long)n * d <= 65535, "n * d exceeds the grid");
  FDK_REQUIRE(cout % 16 == 0, "cout %d not a multiple of 16", cout);
  FDK_REQUIRE(relu >= 0 && relu <= 2, "activation mode %d not in {0, 1, 2}", relu);
  Conv a{static_cast<const u16*>(x), static_cast<const u16*>(w), bias, resid, y, n, d, h, wd,
         cin, cout, relu, out_f32};
  dim3 grid(wd / 32, h / 4, n * d);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int ck = conv_kernel_choice();
  if (ck >= 4 && lean_ok(a)) return launch_fwd4(a, st);
  if (ck >= 4 && cin == 1 && aligned16(y) && (!bias || aligned16(bias)) &&
      (!resid || aligned16(resid))) {
    const dim3 gr(h / 4, n * d);
    switch (cout / 16) {
      case 2: hipLaunchKernelGGL(conv3d_fwd_c1r_kernel<2>, gr, dim3(256), 0, st, a); break;
      case 4: hipLaunchKernelGGL(conv3d_fwd_c1r_kernel<4>, gr, dim3(256), 0, st, a); break;
      default: return fdk::fail(FDK_ERR_UNSUPPORTED, "cin 1: cout %d not 32 or 64", cout);
    }
    FDK_HIP(hipGetLastError());
    return FDK_OK;
  }
  const bool staged = cin % 32 == 0 && d % 2 == 0 && cout <= 128 && ck >= 3;
  const bool tiled = !staged && cin % 32 == 0 && h % 8 == 0 && cout <= 128 && ck >= 2;
  dim3 grid2(wd / 32, h / 8, n * d);
  dim3 grid3(wd / 32, h / 4, n * d / 2);
#define FDK_CONV_CASE(NT)                                                                   \
  case NT:                                                                                  \
    if (cin == 1) hipLaunchKernelGGL(conv3d_fwd_c1_kernel<NT>, grid, dim3(256), 0, st, a);  \
    else if (staged) hipLaunchKernelGGL(conv3d_fwd3_kernel<NT>, grid3, dim3(256), 0, st, a); \
    else if (tiled) hipLaunchKernelGGL(conv3d_fwd2_kernel<NT>, grid2, dim3(256), 0, st, a); \
    else hipLaunchKernelGGL(conv3d_fwd_kernel<NT>, grid, dim3(256), 0, st, a);              \
    break;
  switch (cout / 16) {
    FDK_CONV_CASE(2)
    FDK_CONV_CASE(4)
    FDK_CONV_CASE(6)
    FDK_CONV_CASE(8)
    FDK_CONV_CASE(12)
    default:
      return fdk::fail(FDK_ERR_UNSUPPORTED, "cout %d not in {32, 64, 96, 128, 192}", cout);
  }
#undef FDK_CONV_CASE
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_conv3d_fwd_cat(const void* up, int cup, const void* x, const void* w,
                                     const float* bias, const float* resid, void* y, int n, int d,
                                     int h, int wd, int cin, int cout, int relu, int out_f32,
                                     void* stream) {
  FDK_REQUIRE(up && x && w && y, "NULL pointer");
  FDK_REQUIRE(n > 0 && d > 0 && h > 0 && wd > 0, "bad tensor size");
  FDK_REQUIRE(wd % 32 == 0 && h % 4 == 0 && d % 2 == 0,
              "conv3d_fwd_cat needs W %% 32 == 0, H %% 4 == 0 and even D (got %d, %d, %d)", wd, h, d);
  FDK_REQUIRE(cup > 0 && cup < cin && cup % 32 == 0 && cin % 32 == 0,
              "conv3d_fwd_cat needs 0 < cup < cin, both multiples of 32 (got %d, %d)", cup, cin);
  FDK_REQUIRE(relu >= 0 && relu <= 2, "activation mode %d not in {0, 1, 2}", relu);
  Conv a{static_cast<const u16*>(x), static_cast<const u16*>(w), bias, resid, y, n, d, h, wd,
         cin, cout, relu, out_f32, static_cast<const u16*>(up), cup};
  if (!lean_ok(a))
    return fdk::fail(FDK_ERR_UNSUPPORTED, "conv3d_fwd_cat: cout %d (a multiple of 32), 16-byte aligned "
                     "tensors and 32-bit halo offsets required", cout);
  return launch_fwd4(a, static_cast<hipStream_t>(stream));
}

// Shapes conv3d_wgrad_tr_kernel takes: 32-channel chunks of cin (and of cup),
// cout % 32, H % 4, even D, 16-byte vectors, its halo resource (4 planes + 2
// rows) below 2^32 - 16 bytes.
static bool wgrad_tr_ok(const Wgrad& a) {
  const int cx = a.cin - a.cup;
  return conv_kernel_choice() >= 4 && a.cin % 32 == 0 && a.cup % 32 == 0 && a.cout % 32 == 0 &&
         a.h % 4 == 0 && a.d % 2 == 0 && a.w % 32 == 0 && aligned16(a.x) && aligned16(a.dy) &&
         (!a.up || aligned16(a.up)) &&
         (4LL * a.h * a.w + 2LL * a.w) * std::max(cx, a.cup) < (1LL << 31) - 8;
}

static int launch_wgrad_tr(const Wgrad& a, hipStream_t st) {
  const int tiles_x = a.w / 32, tiles_y = a.h / 4, vt = tiles_x * tiles_y * (a.n * a.d / 2);
  const int chunks = (a.cout / 32) * (a.cin / 32);
  // about 1024 blocks: every CU busy several times over, few partial sums to add
  const int gx = std::max(1, std::min(vt, 1024 / chunks));
  hipLaunchKernelGGL(conv3d_wgrad_tr_kernel, dim3(gx, chunks), dim3(256), 0, st, a, tiles_x, tiles_y, vt);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_conv3d_wgrad_cat(const void* up, int cup, const void* x, const void* dy,
                                       float* dw, int n, int d, int h, int wd, int cin, int cout,
                                       void* stream) {
  FDK_REQUIRE(up && x && dy && dw, "NULL pointer");
  FDK_REQUIRE(n > 0 && d > 0 && h > 0 && wd > 0, "bad tensor size");
  FDK_REQUIRE(cup > 0 && cup < cin && cup % 32 == 0 && cin % 32 == 0,
              "conv3d_wgrad_cat needs 0 < cup < cin, both multiples of 32 (got %d, %d)", cup, cin);
  Wgrad a{static_cast<const u16*>(x), static_cast<const u16*>(dy), dw, n, d, h, wd, cin, cout,
          static_cast<const u16*>(up), cup};
  if (!wgrad_tr_ok(a))
    return fdk::fail(FDK_ERR_UNSUPPORTED, "conv3d_wgrad_cat: cout %% 32, W %% 32, H %% 4, even D, "
                     "16-byte aligned tensors and 32-bit halo offsets required");
  return launch_wgrad_tr(a, static_cast<hipStream_t>(stream));
}

extern "C" int fdk_nn_conv3d_wgrad(const void* x, const void* dy, float* dw, int n, int d, int h,
                                   int wd, int cin, int cout, void* stream) {
  FDK_REQUIRE(x && dy && dw, "NULL pointer");
  FDK_REQUIRE(cout % 16 == 0 && cout <= 256, "cout must be a multiple of 16 up to 256");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Wgrad a{static_cast<const u16*>(x), static_cast<const u16*>(dy), dw, n, d, h, wd, cin, cout};
  if (cin == 1) {
    FDK_REQUIRE(cout == 32, "the cin == 1 weight gradient takes cout == 32");
    if (wd % kC1X == 0 && h % 4 == 0 && d % 2 == 0 && aligned16(dy)) {
      const int tiles_x = wd / kC1X, tiles_y = h / 4, ntiles = tiles_x * tiles_y * (n * d / 2);
      hipLaunchKernelGGL(conv3d_wgrad_c1t_kernel, dim3(std::min(ntiles, 2048)), dim3(256), 0, st, a,
                         tiles_x, tiles_y, ntiles);
    } else {
      hipLaunchKernelGGL(conv3d_wgrad_c1_kernel, dim3(2048), dim3(256), 0, st, a);
    }
    FDK_HIP(hipGetLastError());
    return FDK_OK;
  }
  FDK_REQUIRE(cin % 16 == 0 && wd % 32 == 0, "cin %% 16 and W %% 32 must be 0");
  if (wgrad_tr_ok(a)) return launch_wgrad_tr(a, st);
  const int ntiles = (cout / 16) * (cin / 16);
  const int per_wave = (ntiles + 3) / 4;
  const size_t lds = (size_t)(cout + cin) * (wd + 8) * sizeof(u16);
  FDK_REQUIRE(lds <= 160 * 1024, "weight-gradient tile needs %zu bytes of LDS", lds);
  const int rows = n * d * h;
  dim3 grid((rows + kWgradRows - 1) / kWgradRows, 27);
  if (per_wave <= 4) hipLaunchKernelGGL(conv3d_wgrad_kernel<4>, grid, dim3(256), lds, st, a);
  else if (per_wave <= 8) hipLaunchKernelGGL(conv3d_wgrad_kernel<8>, grid, dim3(256), lds, st, a);
  else if (per_wave <= 16) hipLaunchKernelGGL(conv3d_wgrad_kernel<16>, grid, dim3(256), lds, st, a);
  else if (per_wave <= 24) hipLaunchKernelGGL(conv3d_wgrad_kernel<24>, grid, dim3(256), lds, st, a);
  else return fdk::fail(FDK_ERR_UNSUPPORTED, "too many weight tiles (%d)", ntiles);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_flip_weights(const void* w, void* wb, int cin, int cout, void* stream) {
  FDK_REQUIRE(w && wb, "NULL pointer");
  const long long n = 27LL * cin * cout;
  hipLaunchKernelGGL(flip_weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const u16*>(w), static_cast<u16*>(wb), cin, cout);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_bias_grad(const void* dy, float* db, int64_t nvox, int c, void* stream) {
  FDK_REQUIRE(dy && db && c > 0 && c <= 256 && 256 % c == 0, "bias_grad needs c dividing 256");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 g(1024), b(256);
  const uint4* v = static_cast<const uint4*>(dy);
  if (!aligned16(dy)) c = -c;  // the scalar kernel below
  switch (c) {
    case 32: hipLaunchKernelGGL(bias_grad_v8_kernel<32>, g, b, 0, st, v, db, (long long)nvox); break;
    case 64: hipLaunchKernelGGL(bias_grad_v8_kernel<64>, g, b, 0, st, v, db, (long long)nvox); break;
    case 128: hipLaunchKernelGGL(bias_grad_v8_kernel<128>, g, b, 0, st, v, db, (long long)nvox); break;
    default:
      hipLaunchKernelGGL(bias_grad_kernel, g, b, 0, st, static_cast<const u16*>(dy), db, (long long)nvox,
                         c < 0 ? -c : c);
  }
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_relu_bwd(const void* dy, const void* y, void* dz, int64_t n, int mode,
                               void* stream) {
  FDK_REQUIRE(dy && y && dz, "NULL pointer");
  FDK_REQUIRE(mode == 1 || mode == 2, "activation mode must be 1 (ReLU) or 2 (leaky ReLU)");
  if (n % 8 == 0 && aligned16(dy) && aligned16(y) && aligned16(dz))
    hipLaunchKernelGGL(relu_bwd_v8_kernel, dim3(blocks_for(n / 8)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const uint4*>(dy),
                       static_cast<const uint4*>(y), static_cast<uint4*>(dz), (long long)(n / 8), mode);
  else
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(blocks_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const u16*>(dy), static_cast<const u16*>(y), static_cast<u16*>(dz),
                       (long long)n, mode);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_avgpool2(const void* x, void* y, int n, int d, int h, int w, int c, int adjoint,
                               void* stream) {
  FDK_REQUIRE(x && y && d % 2 == 0 && h % 2 == 0 && w % 2 == 0, "avgpool needs even sizes");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long full = (long long)n * d * h * w * c;
  if (!adjoint && c % 8 == 0 && (long long)n * (d / 2) <= 65535 && h / 2 <= 65535 &&
      reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0) {
    const int cq = c / 8;
    dim3 grid(((w / 2) * cq + 255) / 256, h / 2, n * (d / 2));
    hipLaunchKernelGGL(avgpool_v8_kernel, grid, dim3(256), 0, st, static_cast<const uint4*>(x),
                       static_cast<uint4*>(y), d, h, w, cq);
  } else if (!adjoint) {
    hipLaunchKernelGGL(avgpool_kernel, dim3(blocks_for(full / 8)), dim3(256), 0, st,
                       static_cast<const u16*>(x), static_cast<u16*>(y), n, d, h, w, c);
  } else if (c % 8 == 0 && (long long)n * d <= 65535 && h <= 65535 &&
             reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0) {
    const int cq = c / 8;
    dim3 grid((w * cq + 255) / 256, h, n * d);
    hipLaunchKernelGGL(avgpool_bwd_v8_kernel, grid, dim3(256), 0, st, static_cast<const uint4*>(x),
                       static_cast<uint4*>(y), d, h, w, cq);
  } else {
    hipLaunchKernelGGL(avgpool_bwd_kernel, dim3(blocks_for(full)), dim3(256), 0, st,
                       static_cast<const u16*>(x), static_cast<u16*>(y), n, d, h, w, c);
  }
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_upsample2(const void* x, void* y, int n, int d, int h, int w, int cx, int cy,
                                int c0, int adjoint, void* stream) {
  FDK_REQUIRE(x && y && c0 >= 0 && c0 + cx <= cy, "bad channel range");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const long long small = (long long)n * d * h * w * cx;
  const bool v8 = cx % 8 == 0 && cy % 8 == 0 && c0 % 8 == 0 && (long long)n * 2 * d <= 65535 &&
                  2 * h <= 65535 && reinterpret_cast<uintptr_t>(x) % 16 == 0 &&
                  reinterpret_cast<uintptr_t>(y) % 16 == 0;
  if (!adjoint && v8) {
    const int cq = cx / 8;
    dim3 grid((2 * w * cq + 255) / 256, 2 * h, n * 2 * d);
    hipLaunchKernelGGL(upsample_v8_kernel, grid, dim3(256), 0, st, static_cast<const uint4*>(x),
                       static_cast<uint4*>(y), d, h, w, cq, cy / 8, c0 / 8);
  } else if (!adjoint) {
    hipLaunchKernelGGL(upsample_kernel, dim3(blocks_for(small * 8)), dim3(256), 0, st,
                       static_cast<const u16*>(x), static_cast<u16*>(y), n, d, h, w, cx, cy, c0);
  } else if (v8) {
    const int cq = cx / 8;
    dim3 grid((w * cq + 255) / 256, h, n * d);
    hipLaunchKernelGGL(upsample_bwd_v8_kernel, grid, dim3(256), 0, st, static_cast<const uint4*>(x),
                       static_cast<uint4*>(y), d, h, w, cq, cy / 8, c0 / 8);
  } else {
    hipLaunchKernelGGL(upsample_bwd_kernel, dim3(blocks_for(small)), dim3(256), 0, st,
                       static_cast<const u16*>(x), static_cast<u16*>(y), n, d, h, w, cx, cy, c0);
  }
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_channel_copy(const void* src, int cs, int src_c0, void* dst, int cd,
                                   int dst_c0, int cc, int64_t nvox, void* stream) {
  FDK_REQUIRE(src && dst && src_c0 + cc <= cs && dst_c0 + cc <= cd, "bad channel range");
  if (cs % 8 == 0 && src_c0 % 8 == 0 && cd % 8 == 0 && dst_c0 % 8 == 0 && cc % 8 == 0 &&
      nvox * (cc / 8) < (1LL << 31) && reinterpret_cast<uintptr_t>(src) % 16 == 0 &&
      reinterpret_cast<uintptr_t>(dst) % 16 == 0) {
    const int total = (int)(nvox * (cc / 8));
    hipLaunchKernelGGL(channel_copy_v8_kernel, dim3((total + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const uint4*>(src), cs / 8,
                       src_c0 / 8, static_cast<uint4*>(dst), cd / 8, dst_c0 / 8, cc / 8, total);
    FDK_HIP(hipGetLastError());
    return FDK_OK;
  }
  hipLaunchKernelGGL(channel_copy_kernel, dim3(blocks_for(nvox * cc)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const u16*>(src), cs, src_c0,
                     static_cast<u16*>(dst), cd, dst_c0, cc, (long long)nvox);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_head(const void* h, const float* w, const float* b, const float* in,
                           float* out, int64_t nvox, int c, void* stream) {
  FDK_REQUIRE(h && w && b && in && out && c <= 63, "bad head arguments");
  if (c == 32 && reinterpret_cast<uintptr_t>(h) % 16 == 0) {
    hipLaunchKernelGGL(head_fwd_v8_kernel<32>, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const uint4*>(h), w, b, in,
                       out, (long long)nvox);
    FDK_HIP(hipGetLastError());
    return FDK_OK;
  }
  hipLaunchKernelGGL(head_fwd_kernel, dim3(blocks_for(nvox)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const u16*>(h), w, b, in, out,
                     (long long)nvox, c);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_head_bwd(const float* dout, const void* h, const float* w, void* dh,
                               float* dw, float* db, int64_t nvox, int c, void* stream) {
  FDK_REQUIRE(dout && h && w && dh && dw && db && c <= 63, "bad head arguments");
  const dim3 g(std::min<long long>(1024, blocks_for(nvox)));
  if (c == 32 && aligned16(h) && aligned16(dh))
    hipLaunchKernelGGL(head_bwd_v8_kernel<32>, g, dim3(256), 0, static_cast<hipStream_t>(stream), dout,
                       static_cast<const uint4*>(h), w, static_cast<uint4*>(dh), dw, db, (long long)nvox);
  else
    hipLaunchKernelGGL(head_bwd_kernel, g, dim3(256), 0, static_cast<hipStream_t>(stream), dout,
                       static_cast<const u16*>(h), w, static_cast<u16*>(dh), dw, db, (long long)nvox, c);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

static bool s2d_sizes_ok(int n, int d, int h, int w, int f) {
  return (f == 2 || f == 4) && n > 0 && d > 0 && h > 0 && w > 0 && d % f == 0 && h % f == 0 &&
         w % f == 0;
}

extern "C" int fdk_nn_unshuffle(const float* x, void* y, int n, int d, int h, int w, int f, int cy,
                                void* stream) {
  FDK_REQUIRE(x && y, "NULL pointer");
  FDK_REQUIRE(s2d_sizes_ok(n, d, h, w, f),
              "unshuffle needs f in {2, 4} and d, h, w positive multiples of f (got f %d, %d, %d, %d)",
              f, d, h, w);
  FDK_REQUIRE(cy >= f * f * f && cy % 8 == 0, "cy must be a multiple of 8, at least f^3 (got %d)", cy);
  FDK_REQUIRE(reinterpret_cast<uintptr_t>(x) % (4 * f) == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0,
              "unshuffle needs 4f-byte aligned x and 16-byte aligned y");
  const long long nvox = (long long)n * (d / f) * (h / f) * (w / f);
  const dim3 g((unsigned)((nvox + 255) / 256));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (f == 2)
    hipLaunchKernelGGL(unshuffle_kernel<2>, g, dim3(256), 0, st, x, static_cast<uint4*>(y), h / 2,
                       w / 2, cy / 8, nvox);
  else
    hipLaunchKernelGGL(unshuffle_kernel<4>, g, dim3(256), 0, st, x, static_cast<uint4*>(y), h / 4,
                       w / 4, cy / 8, nvox);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_head_shuffle(const void* h, const float* w, const float* b, const float* in,
                                   float* out, int n, int d, int hh, int ww, int f, int c,
                                   void* stream) {
  FDK_REQUIRE(h && w && b && in && out, "NULL pointer");
  FDK_REQUIRE(c == 32, "the depth-to-space head takes 32 channels (got %d)", c);
  FDK_REQUIRE(s2d_sizes_ok(n, d, hh, ww, f),
              "head_shuffle needs f in {2, 4} and d, h, w positive multiples of f (got f %d, %d, %d, %d)",
              f, d, hh, ww);
  FDK_REQUIRE(reinterpret_cast<uintptr_t>(h) % 16 == 0 && reinterpret_cast<uintptr_t>(in) % (4 * f) == 0 &&
                  reinterpret_cast<uintptr_t>(out) % (4 * f) == 0,
              "head_shuffle needs 16-byte aligned h and 4f-byte aligned in / out");
  const long long nvox = (long long)n * (d / f) * (hh / f) * (ww / f);
  const dim3 g((unsigned)((nvox + 255) / 256));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (f == 2)
    hipLaunchKernelGGL((head_shuffle_kernel<2, 32>), g, dim3(256), 0, st,
                       static_cast<const uint4*>(h), w, b, in, out, hh / 2, ww / 2, nvox);
  else
    hipLaunchKernelGGL((head_shuffle_kernel<4, 32>), g, dim3(256), 0, st,
                       static_cast<const uint4*>(h), w, b, in, out, hh / 4, ww / 4, nvox);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_head_shuffle_bwd(const float* dout, const void* h, const float* w, void* dh,
                                       float* dw, float* db, int n, int d, int hh, int ww, int f,
                                       int c, void* stream) {
  FDK_REQUIRE(dout && h && w && dh && dw && db, "NULL pointer");
  FDK_REQUIRE(c == 32, "the depth-to-space head takes 32 channels (got %d)", c);
  FDK_REQUIRE(s2d_sizes_ok(n, d, hh, ww, f),
              "head_shuffle needs f in {2, 4} and d, h, w positive multiples of f (got f %d, %d, %d, %d)",
              f, d, hh, ww);
  FDK_REQUIRE(reinterpret_cast<uintptr_t>(h) % 16 == 0 && reinterpret_cast<uintptr_t>(dh) % 16 == 0 &&
                  reinterpret_cast<uintptr_t>(dout) % (4 * f) == 0,
              "head_shuffle_bwd needs 16-byte aligned h, dh and 4f-byte aligned dout");
  const long long nvox = (long long)n * (d / f) * (hh / f) * (ww / f);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (f == 2)
    hipLaunchKernelGGL((head_shuffle_bwd_kernel<2, 32>), dim3(std::min<long long>(1024, (nvox + 255) / 256)),
                       dim3(256), 0, st, dout, static_cast<const uint4*>(h), w,
                       static_cast<uint4*>(dh), dw, db, hh / 2, ww / 2, nvox);
  else
    hipLaunchKernelGGL((head_shuffle4_bwd_kernel<32>), dim3(std::min<long long>(1024, (nvox + 63) / 64)),
                       dim3(256), 0, st, dout, static_cast<const uint4*>(h), w,
                       static_cast<uint4*>(dh), dw, db, hh / 4, ww / 4, nvox);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_mse(const float* out, const float* target, float* dout, float* loss_sum,
                          int64_t nvox, void* stream) {
  FDK_REQUIRE(out && target && loss_sum, "NULL pointer");
  hipLaunchKernelGGL(mse_kernel, dim3(std::min<long long>(2048, blocks_for(nvox))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), out, target, dout, loss_sum, (long long)nvox);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_adam(float* w, void* wbf, const float* g, float* m, float* v, int64_t n,
                           float lr, float b1, float b2, float eps, int step, int row, int bf_row,
                           void* stream) {
  FDK_REQUIRE(w && g && m && v && step >= 1 && row >= 1 && bf_row >= row, "bad Adam arguments");
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     w, static_cast<u16*>(wbf), g, m, v, (long long)n, lr, b1, b2, eps, c1, c2, row, bf_row);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_convert(const void* src, void* dst, int64_t n, int to_bf16, void* stream) {
  FDK_REQUIRE(src && dst, "NULL pointer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (to_bf16)
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, st,
                       static_cast<const float*>(src), static_cast<u16*>(dst), (long long)n);
  else
    hipLaunchKernelGGL(bf16_to_f32_kernel, dim3(blocks_for(n)), dim3(256), 0, st,
                       static_cast<const u16*>(src), static_cast<float*>(dst), (long long)n);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_nn_box_copy(const float* src, const int src_dims[3], const int src_org[3],
                               float* dst, const int dst_dims[3], const int dst_org[3],
                               const int box[3], float scale, void* stream) {
  FDK_REQUIRE(src && dst && src_dims && src_org && dst_dims && dst_org && box, "NULL pointer");
  for (int e = 0; e < 3; ++e) {
    FDK_REQUIRE(box[e] >= 0 && src_org[e] >= 0 && dst_org[e] >= 0, "negative box");
    FDK_REQUIRE(src_org[e] + box[e] <= src_dims[e] && dst_org[e] + box[e] <= dst_dims[e],
                "box leaves a volume on axis %d", e);
  }
  const long long n = (long long)box[0] * box[1] * box[2];
  if (n == 0) return FDK_OK;
  hipLaunchKernelGGL(box_copy_kernel, dim3(blocks_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     src, src_dims[0], src_dims[1], src_dims[2], src_org[0], src_org[1], src_org[2],
                     dst, dst_dims[0], dst_dims[1], dst_dims[2], dst_org[0], dst_org[1], dst_org[2],
                     box[0], box[1], box[2], scale);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_poisson_noise(float* dev_proj, int64_t n, double i0, double mu, uint32_t seed,
                                 void* stream) {
  FDK_REQUIRE(n >= 0, "negative count");
  FDK_REQUIRE(i0 > 0 && mu > 0, "bad noise arguments (i0 %g, mu %g)", i0, mu);
  if (n == 0) return FDK_OK;
  FDK_REQUIRE(dev_proj, "NULL projections");
  hipLaunchKernelGGL(poisson_kernel, dim3(blocks_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     dev_proj, (long long)n, i0, mu, seed);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

namespace {
__global__ void add_bf16_kernel(const u16* a, const u16* b, u16* y, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    y[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

__global__ void add_bf16_v8_kernel(const uint4* a, const uint4* b, uint4* y, long long n8) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n8; i += (long long)gridDim.x * 256) {
    float fa[8], fb[8];
    unpack8(a[i], fa);
    unpack8(b[i], fb);
#pragma unroll
    for (int e = 0; e < 8; ++e) fa[e] += fb[e];
    y[i] = pack8(fa);
  }
}
}  // namespace

extern "C" int fdk_nn_add(const void* a, const void* b, void* y, int64_t n, void* stream) {
  FDK_REQUIRE(a && b && y, "NULL pointer");
  if (n % 8 == 0 && aligned16(a) && aligned16(b) && aligned16(y))
    hipLaunchKernelGGL(add_bf16_v8_kernel, dim3(blocks_for(n / 8)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), static_cast<const uint4*>(a),
                       static_cast<const uint4*>(b), static_cast<uint4*>(y), (long long)(n / 8));
  else
    hipLaunchKernelGGL(add_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const u16*>(a), static_cast<const u16*>(b), static_cast<u16*>(y),
                       (long long)n);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}
