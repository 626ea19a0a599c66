// Ray-marching forward projection and random-ellipsoid phantoms on gfx950.
//
// Reference behaviour (peng-ac/fdk):
//   kernel_forwardProject (single volume)  genProj/cudaForwardProjection.cu:174-254
//   cuIntersectBox                          genProj/cudaForwardProjection.cu:28-48
//   Text3DLinear (clamp-to-edge trilinear)  genProj/cudaIterType.cuh:231-278
//   ForwardProjection / cudaCTForwardProjection
//                                           genProj/cudaForwardProjection.cu:338-487
//   CPU ForwardProjection (zero outside)    3dct/3dct.cpp:430-573, common/type.h:2658-2712
//
// One thread per detector pixel, 64 lanes along u so a wave's rays stay
// together through the volume.  The ray set-up (source, pixel, direction, slab
// test) runs in FP64 once per ray; the march runs in FP32 and reads the eight
// trilinear taps straight from HBM/L2 (no texture units): x-adjacent tap pairs
// are 8-byte loads.  The ellipsoid phantom kernels (no reference counterpart;
// the reference loads external phantoms) use FP64 for the chord lengths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace {

constexpr double kPi = 3.141592653589793238462;

struct RayGeom {
  int nu, nv;
  double du, dv, cu, cv;  // detector pixel pitch and centre (with offsets)
  double sod, sid;
  double vx, vy, vz;      // voxel size
  double cx, cy, cz;      // volume centre index
  int nx, ny, nz;
  int direction;
  double scan_deg, vpa;
  int nproj, first;
};

// Source and unit direction of the ray through pixel (i, j) of view s:
// source R(a)(0,-SOD,0), pixel R(a)(u_mm, SID-SOD, v_mm), a = -(dir*deg + vpa).
__device__ inline void pixel_ray(const RayGeom& g, int s, int i, int j, double o[3], double d[3]) {
  const double deg = s * g.scan_deg / (double)g.nproj;
  const double a = -(g.direction * deg / 180.0 * kPi + g.vpa);
  double sn, cs;
  sincos(a, &sn, &cs);
  const double umm = (i - g.cu) * g.du, vmm = (g.cv - j) * g.dv;
  const double dety = g.sid - g.sod;
  o[0] = g.sod * sn;
  o[1] = -g.sod * cs;
  o[2] = 0.0;
  d[0] = umm * cs - dety * sn - o[0];
  d[1] = umm * sn + dety * cs - o[1];
  d[2] = vmm;
  const double n = rsqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  d[0] *= n; d[1] *= n; d[2] *= n;
}

template <int MODE>
__device__ inline float tap_sample(const float* __restrict__ vol, int nx, int ny, int nz, float px,
                                   float py, float pz) {
  const int ax = (int)px, ay = (int)py, az = (int)pz;  // truncation, as the reference
  const float fx = px - ax, fy = py - ay, fz = pz - az;
  int x0 = ax, x1 = ax + 1, y0 = ay, y1 = ay + 1, z0 = az, z1 = az + 1;
  float mx0 = 1.f, mx1 = 1.f, my0 = 1.f, my1 = 1.f, mz0 = 1.f, mz1 = 1.f;
  if (MODE == FDK_FP_CLAMP) {
    x0 = min(max(x0, 0), nx - 1); x1 = min(max(x1, 0), nx - 1);
    y0 = min(max(y0, 0), ny - 1); y1 = min(max(y1, 0), ny - 1);
    z0 = min(max(z0, 0), nz - 1); z1 = min(max(z1, 0), nz - 1);
  } else {
    // zero outside: mask taps and clamp their addresses into the volume
    mx0 = (x0 >= 0 && x0 < nx) ? 1.f : 0.f; mx1 = (x1 >= 0 && x1 < nx) ? 1.f : 0.f;
    my0 = (y0 >= 0 && y0 < ny) ? 1.f : 0.f; my1 = (y1 >= 0 && y1 < ny) ? 1.f : 0.f;
    mz0 = (z0 >= 0 && z0 < nz) ? 1.f : 0.f; mz1 = (z1 >= 0 && z1 < nz) ? 1.f : 0.f;
    x0 = min(max(x0, 0), nx - 1); x1 = min(max(x1, 0), nx - 1);
    y0 = min(max(y0, 0), ny - 1); y1 = min(max(y1, 0), ny - 1);
    z0 = min(max(z0, 0), nz - 1); z1 = min(max(z1, 0), nz - 1);
  }
  const long long pz0 = (long long)z0 * ny * nx, pz1 = (long long)z1 * ny * nx;
  const long long r00 = pz0 + (long long)y0 * nx, r01 = pz0 + (long long)y1 * nx;
  const long long r10 = pz1 + (long long)y0 * nx, r11 = pz1 + (long long)y1 * nx;
  const float t000 = vol[r00 + x0] * (mx0 * my0 * mz0), t001 = vol[r00 + x1] * (mx1 * my0 * mz0);
  const float t010 = vol[r01 + x0] * (mx0 * my1 * mz0), t011 = vol[r01 + x1] * (mx1 * my1 * mz0);
  const float t100 = vol[r10 + x0] * (mx0 * my0 * mz1), t101 = vol[r10 + x1] * (mx1 * my0 * mz1);
  const float t110 = vol[r11 + x0] * (mx0 * my1 * mz1), t111 = vol[r11 + x1] * (mx1 * my1 * mz1);
  const float gx = 1.f - fx, gy = 1.f - fy;
  const float v0 = (t000 * gx + t001 * fx) * gy + (t010 * gx + t011 * fx) * fy;
  const float v1 = (t100 * gx + t101 * fx) * gy + (t110 * gx + t111 * fx) * fy;
  return v0 * (1.f - fz) + v1 * fz;
}

// tap_sample through a buffer resource over the volume (32-bit offsets, for
// volumes under 2 GiB): the taps x0 and x1 of each (y, z) row come from one
// 8-byte load at xb = clamp(ax, 0, nx - 2), each tap picking the element it
// clamps to, so a sample costs 4 loads instead of 8.  Same values and
// arithmetic as tap_sample.  It comes in two halves, so that several samples'
// loads can be in flight before any is combined: tap_fetch_buf issues the four
// loads and keeps what the blend needs; tap_blend forms the sample.
struct TapBuf {
  float2 q[4];              // rows (z0,y0), (z0,y1), (z1,y0), (z1,y1)
  float fx, fy, fz;
  float mx0, mx1, my0, my1, mz0, mz1;
  bool s0, s1;              // tap x0 / x1 is the pair's second element
};

template <int MODE>
__device__ inline TapBuf tap_fetch_buf(__amdgpu_buffer_rsrc_t rv, int nx, int ny, int nz, float px,
                                       float py, float pz) {
  TapBuf b;
  const int ax = (int)px, ay = (int)py, az = (int)pz;  // truncation, as the reference
  b.fx = px - ax;
  b.fy = py - ay;
  b.fz = pz - az;
  int y0 = ay, y1 = ay + 1, z0 = az, z1 = az + 1;
  b.mx0 = b.mx1 = b.my0 = b.my1 = b.mz0 = b.mz1 = 1.f;
  if (MODE != FDK_FP_CLAMP) {
    b.mx0 = (ax >= 0 && ax < nx) ? 1.f : 0.f; b.mx1 = (ax + 1 >= 0 && ax + 1 < nx) ? 1.f : 0.f;
    b.my0 = (y0 >= 0 && y0 < ny) ? 1.f : 0.f; b.my1 = (y1 >= 0 && y1 < ny) ? 1.f : 0.f;
    b.mz0 = (z0 >= 0 && z0 < nz) ? 1.f : 0.f; b.mz1 = (z1 >= 0 && z1 < nz) ? 1.f : 0.f;
  }
  y0 = min(max(y0, 0), ny - 1); y1 = min(max(y1, 0), ny - 1);
  z0 = min(max(z0, 0), nz - 1); z1 = min(max(z1, 0), nz - 1);
  const int x0 = min(max(ax, 0), nx - 1), x1 = min(max(ax + 1, 0), nx - 1);
  const int xb = min(max(ax, 0), nx - 2);
  b.s0 = x0 != xb;
  b.s1 = x1 != xb;
  const int pz0 = z0 * ny * nx, pz1 = z1 * ny * nx;
  const int o[4] = {(pz0 + y0 * nx + xb) * 4, (pz0 + y1 * nx + xb) * 4, (pz1 + y0 * nx + xb) * 4,
                    (pz1 + y1 * nx + xb) * 4};
#pragma unroll
  for (int k = 0; k < 4; ++k)
    b.q[k] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(rv, o[k], 0, 0));
  return b;
}

__device__ inline float tap_blend(const TapBuf& b) {
  const float t000 = (b.s0 ? b.q[0].y : b.q[0].x) * (b.mx0 * b.my0 * b.mz0);
  const float t001 = (b.s1 ? b.q[0].y : b.q[0].x) * (b.mx1 * b.my0 * b.mz0);
  const float t010 = (b.s0 ? b.q[1].y : b.q[1].x) * (b.mx0 * b.my1 * b.mz0);
  const float t011 = (b.s1 ? b.q[1].y : b.q[1].x) * (b.mx1 * b.my1 * b.mz0);
  const float t100 = (b.s0 ? b.q[2].y : b.q[2].x) * (b.mx0 * b.my0 * b.mz1);
  const float t101 = (b.s1 ? b.q[2].y : b.q[2].x) * (b.mx1 * b.my0 * b.mz1);
  const float t110 = (b.s0 ? b.q[3].y : b.q[3].x) * (b.mx0 * b.my1 * b.mz1);
  const float t111 = (b.s1 ? b.q[3].y : b.q[3].x) * (b.mx1 * b.my1 * b.mz1);
  const float gx = 1.f - b.fx, gy = 1.f - b.fy;
  const float v0 = (t000 * gx + t001 * b.fx) * gy + (t010 * gx + t011 * b.fx) * b.fy;
  const float v1 = (t100 * gx + t101 * b.fx) * gy + (t110 * gx + t111 * b.fx) * b.fy;
  return v0 * (1.f - b.fz) + v1 * b.fz;
}

template <int MODE, bool BUF>
__global__ __launch_bounds__(256) void fp_kernel(RayGeom g, const float* __restrict__ vol,
                                                 float step, float* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int sv = blockIdx.z;
  if (i >= g.nu || j >= g.nv) return;
  double o[3], d[3];
  pixel_ray(g, g.first + sv, i, j, o, d);
  // slab test against the voxel-centre box
  const double bmax[3] = {g.cx * g.vx, g.cy * g.vy, g.cz * g.vz};
  double tn = -1e300, tf = 1e300;
  for (int q = 0; q < 3; ++q) {
    const double inv = 1.0 / d[q];
    const double t0 = inv * (-bmax[q] - o[q]), t1 = inv * (bmax[q] - o[q]);
    tn = fmax(tn, fmin(t0, t1));
    tf = fmin(tf, fmax(t0, t1));
  }
  float sum = 0.f;
  if (tf > tn) {
    const float ox = (float)(o[0] / g.vx + g.cx), oy = (float)(g.cy - o[1] / g.vy),
                oz = (float)(g.cz - o[2] / g.vz);
    const float sx = (float)(d[0] / g.vx), sy = (float)(-d[1] / g.vy), sz = (float)(-d[2] / g.vz);
    const float tfar = (float)tf;
    // Four samples per iteration: their gathers are issued together, so one
    // memory latency covers four steps.  t advances by the same float additions
    // and the samples are summed in the same order as a one-step loop, and a
    // sample past tfar adds nothing, so the result is the one-step loop's.
    const __amdgpu_buffer_rsrc_t rv = __builtin_amdgcn_make_buffer_rsrc(
        (void*)vol, 0, BUF ? g.nx * g.ny * g.nz * 4 : 0, fdk::kRawRsrcWord3);
    float t = (float)tn;
    bool more = MODE == FDK_FP_CLAMP ? t < tfar : t <= tfar;
    while (more) {
      float tk[4], sk[4];
      bool ok[4];
      tk[0] = t;
#pragma unroll
      for (int k = 1; k < 4; ++k) tk[k] = tk[k - 1] + step;
      if constexpr (BUF) {
        TapBuf tb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          tb[k] = tap_fetch_buf<MODE>(rv, g.nx, g.ny, g.nz, fmaf(sx, tk[k], ox),
                                      fmaf(sy, tk[k], oy), fmaf(sz, tk[k], oz));
#pragma unroll
        for (int k = 0; k < 4; ++k) sk[k] = tap_blend(tb[k]);
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          sk[k] = tap_sample<MODE>(vol, g.nx, g.ny, g.nz, fmaf(sx, tk[k], ox),
                                   fmaf(sy, tk[k], oy), fmaf(sz, tk[k], oz));
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) ok[k] = MODE == FDK_FP_CLAMP ? tk[k] < tfar : tk[k] <= tfar;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (ok[k]) sum += sk[k];
      t = tk[3] + step;
      more = ok[3] && (MODE == FDK_FP_CLAMP ? t < tfar : t <= tfar);
    }
  }
  out[((long long)sv * g.nv + j) * g.nu + i] = sum * step;
}

// Ellipsoid in the form used on the device: local = Q (p - c), Q = diag(1/a) R^T.
struct EllDev {
  double c[3];
  double q[9];
  double density;
};

std::vector<EllDev> prepare(const fdk_ellipsoid* e, int n) {
  std::vector<EllDev> out(n);
  for (int k = 0; k < n; ++k) {
    const double a[3] = {e[k].ax, e[k].ay, e[k].az};
    out[k].c[0] = e[k].cx; out[k].c[1] = e[k].cy; out[k].c[2] = e[k].cz;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) out[k].q[3 * r + c] = e[k].rot[3 * c + r] / a[r];
    out[k].density = e[k].density;
  }
  return out;
}

__global__ __launch_bounds__(256) void phantom_raster_kernel(const EllDev* __restrict__ e, int n,
                                                             RayGeom g, float* __restrict__ vol) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= g.nx || j >= g.ny) return;
  const double p[3] = {(i - g.cx) * g.vx, (g.cy - j) * g.vy, (g.cz - k) * g.vz};
  double acc = 0.0;
  for (int q = 0; q < n; ++q) {
    const double d0 = p[0] - e[q].c[0], d1 = p[1] - e[q].c[1], d2 = p[2] - e[q].c[2];
    double r2 = 0.0;
    for (int r = 0; r < 3; ++r) {
      const double l = e[q].q[3 * r] * d0 + e[q].q[3 * r + 1] * d1 + e[q].q[3 * r + 2] * d2;
      r2 += l * l;
    }
    if (r2 <= 1.0) acc += e[q].density;
  }
  vol[((long long)k * g.ny + j) * g.nx + i] = (float)acc;
}

__global__ __launch_bounds__(256) void phantom_project_kernel(const EllDev* __restrict__ e, int n,
                                                              RayGeom g, float* __restrict__ out) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y * 4 + threadIdx.y;
  const int sv = blockIdx.z;
  if (i >= g.nu || j >= g.nv) return;
  double o[3], d[3];
  pixel_ray(g, g.first + sv, i, j, o, d);
  double acc = 0.0;
  for (int q = 0; q < n; ++q) {
    // move the origin to the point of the ray nearest the centre first, which
    // keeps the quadratic's coefficients O(1)
    const double rc[3] = {o[0] - e[q].c[0], o[1] - e[q].c[1], o[2] - e[q].c[2]};
    const double t0 = -(rc[0] * d[0] + rc[1] * d[1] + rc[2] * d[2]);
    const double pc[3] = {rc[0] + t0 * d[0], rc[1] + t0 * d[1], rc[2] + t0 * d[2]};
    double A = 0.0, B = 0.0, C = -1.0;
    for (int r = 0; r < 3; ++r) {
      const double lo = e[q].q[3 * r] * pc[0] + e[q].q[3 * r + 1] * pc[1] + e[q].q[3 * r + 2] * pc[2];
      const double ld = e[q].q[3 * r] * d[0] + e[q].q[3 * r + 1] * d[1] + e[q].q[3 * r + 2] * d[2];
      A += ld * ld;
      B += 2.0 * lo * ld;
      C += lo * lo;
    }
    const double disc = B * B - 4.0 * A * C;
    if (disc > 0.0) acc += e[q].density * sqrt(disc) / A;
  }
  out[((long long)sv * g.nv + j) * g.nu + i] = (float)acc;
}

int make_geom(const fdk_geometry* g, int first, RayGeom& r) {
  FDK_REQUIRE(g != nullptr, "geometry is NULL");
  FDK_REQUIRE(g->nx > 0 && g->ny > 0 && g->nz > 0 && g->nu > 0 && g->nv > 0 && g->nproj > 0,
              "geometry sizes must be positive");
  FDK_REQUIRE(g->dx > 0 && g->dy > 0 && g->dz > 0 && g->du > 0 && g->dv > 0, "spacings must be positive");
  FDK_REQUIRE(g->sod > 0 && g->sid > g->sod, "need 0 < sod < sid");
  r.nu = g->nu; r.nv = g->nv;
  r.du = g->du; r.dv = g->dv;
  r.cu = (g->nu - 1) / 2.0 + g->offset_nu;
  r.cv = (g->nv - 1) / 2.0 + g->offset_nv;
  r.sod = g->sod; r.sid = g->sid;
  r.vx = g->dx; r.vy = g->dy; r.vz = g->dz;
  r.nx = g->nx; r.ny = g->ny; r.nz = g->nz;
  r.cx = (g->nx - 1) / 2.0; r.cy = (g->ny - 1) / 2.0; r.cz = (g->nz - 1) / 2.0;
  r.direction = g->direction;
  r.scan_deg = g->scan_angle_deg;
  r.vpa = g->volume_position_angle_rad;
  r.nproj = g->nproj;
  r.first = first;
  return FDK_OK;
}

int upload_ell(const fdk_ellipsoid* host, int n, hipStream_t st, EllDev** out) {
  FDK_REQUIRE(host != nullptr || n == 0, "ellipsoids are NULL");
  FDK_REQUIRE(n >= 0 && n <= 100000, "bad ellipsoid count %d", n);
  for (int k = 0; k < n; ++k)
    FDK_REQUIRE(host[k].ax > 0 && host[k].ay > 0 && host[k].az > 0, "ellipsoid %d has a non-positive axis", k);
  std::vector<EllDev> d = prepare(host, n);
  FDK_HIP(hipMallocAsync((void**)out, sizeof(EllDev) * std::max(n, 1), st));
  FDK_HIP(hipMemcpyAsync(*out, d.data(), sizeof(EllDev) * n, hipMemcpyHostToDevice, st));
  FDK_HIP(hipStreamSynchronize(st));  // d goes out of scope
  return FDK_OK;
}

}  // namespace

extern "C" int fdk_fp_project(const fdk_geometry* g, const float* dev_vol, int first, int count,
                              double step_mm, int mode, float* dev_out, void* stream) {
  RayGeom r;
  if (int rc = make_geom(g, first, r)) return rc;
  FDK_REQUIRE(dev_vol && dev_out, "NULL pointer");
  FDK_REQUIRE(first >= 0 && count >= 0, "bad view range");
  FDK_REQUIRE(mode == FDK_FP_CLAMP || mode == FDK_FP_ZERO, "unknown mode %d", mode);
  if (step_mm <= 0) step_mm = std::min(g->dx, std::min(g->dy, g->dz));
  if (count == 0) return FDK_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  dim3 grid((g->nu + 63) / 64, (g->nv + 3) / 4, count);
  // paired-tap buffer path when the volume's byte offsets fit 32 bits
  const bool buf = g->nx >= 2 && (long long)g->nx * g->ny * g->nz * 4 < (1LL << 31);
  if (mode == FDK_FP_CLAMP && buf)
    hipLaunchKernelGGL((fp_kernel<FDK_FP_CLAMP, true>), grid, dim3(64, 4), 0, st, r, dev_vol, (float)step_mm, dev_out);
  else if (mode == FDK_FP_CLAMP)
    hipLaunchKernelGGL((fp_kernel<FDK_FP_CLAMP, false>), grid, dim3(64, 4), 0, st, r, dev_vol, (float)step_mm, dev_out);
  else if (buf)
    hipLaunchKernelGGL((fp_kernel<FDK_FP_ZERO, true>), grid, dim3(64, 4), 0, st, r, dev_vol, (float)step_mm, dev_out);
  else
    hipLaunchKernelGGL((fp_kernel<FDK_FP_ZERO, false>), grid, dim3(64, 4), 0, st, r, dev_vol, (float)step_mm, dev_out);
  FDK_HIP(hipGetLastError());
  return FDK_OK;
}

extern "C" int fdk_phantom_rasterize(const fdk_ellipsoid* host_ell, int n, const fdk_geometry* g,
                                     float* dev_vol, void* stream) {
  RayGeom r;
  if (int rc = make_geom(g, 0, r)) return rc;
  FDK_REQUIRE(dev_vol, "volume is NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  EllDev* d = nullptr;
  if (int rc = upload_ell(host_ell, n, st, &d)) return rc;
  dim3 grid((g->nx + 63) / 64, (g->ny + 3) / 4, g->nz);
  hipLaunchKernelGGL(phantom_raster_kernel, grid, dim3(64, 4), 0, st, d, n, r, dev_vol);
  hipError_t e = hipGetLastError();
  (void)hipFreeAsync(d, st);
  if (e != hipSuccess) return fdk::fail(FDK_ERR_HIP, "phantom_rasterize: %s", hipGetErrorString(e));
  return FDK_OK;
}

extern "C" int fdk_phantom_project(const fdk_ellipsoid* host_ell, int n, const fdk_geometry* g,
                                   int first, int count, float* dev_out, void* stream) {
  RayGeom r;
  if (int rc = make_geom(g, first, r)) return rc;
  FDK_REQUIRE(dev_out, "output is NULL");
  FDK_REQUIRE(first >= 0 && count >= 0, "bad view range");
  if (count == 0) return FDK_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  EllDev* d = nullptr;
  if (int rc = upload_ell(host_ell, n, st, &d)) return rc;
  dim3 grid((g->nu + 63) / 64, (g->nv + 3) / 4, count);
  hipLaunchKernelGGL(phantom_project_kernel, grid, dim3(64, 4), 0, st, d, n, r, dev_out);
  hipError_t e = hipGetLastError();
  (void)hipFreeAsync(d, st);
  if (e != hipSuccess) return fdk::fail(FDK_ERR_HIP, "phantom_project: %s", hipGetErrorString(e));
  return FDK_OK;
}

extern "C" int fdk_fp_project_host(const fdk_geometry* g, const float* host_vol, int first,
                                   int count, double step_mm, int mode, float* host_out) {
  FDK_REQUIRE(g && host_vol && host_out, "NULL pointer");
  FDK_REQUIRE(count >= 0, "bad view count");
  const size_t nvox = (size_t)g->nx * g->ny * g->nz;
  const size_t nout = (size_t)g->nu * g->nv * count;
  float *v = nullptr, *o = nullptr;
  hipError_t e = hipMalloc(&v, std::max<size_t>(nvox, 1) * 4);
  if (e == hipSuccess) e = hipMalloc(&o, std::max<size_t>(nout, 1) * 4);
  if (e == hipSuccess) e = hipMemcpy(v, host_vol, nvox * 4, hipMemcpyHostToDevice);
  int r = e == hipSuccess ? FDK_OK : fdk::fail(FDK_ERR_HIP, "forward projection upload: %s", hipGetErrorString(e));
  if (r == FDK_OK) r = fdk_fp_project(g, v, first, count, step_mm, mode, o, nullptr);
  if (r == FDK_OK) {
    e = hipMemcpy(host_out, o, nout * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "forward projection download: %s", hipGetErrorString(e));
  }
  if (v) (void)hipFree(v);
  if (o) (void)hipFree(o);
  return r;
}
