/* TEST INFRASTRUCTURE ONLY -- the oracle (CPU checker) for the FDK hot path.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only to check results; the product library never does.
 *
 * Plain C (C99 + OpenMP), FP64 arithmetic unless a function says otherwise.
 * Each function restates the reference algorithm it cites.  How each is pinned
 * (tests/test_oracle.py):
 *   orc_nextpow2, orc_filter_design      golden fixture from the compiled reference
 *                                        (all six window types, several lengths)
 *   orc_angular_ramp_factor, orc_cos_table, orc_proj_matrix(es)
 *                                        golden fixture + closed forms (centre ray,
 *                                        z-independence of rows 0/2, mirror symmetry)
 *   orc_fft                              numpy.fft + Parseval + golden from the reference FFT
 *   orc_filter_rows                      golden from the reference filterImageRow (mean3),
 *                                        delta response == IDFT(H), linearity, constant rows
 *   orc_bp (cpu_trunc)                   golden from the reference 3dct BackProjection
 *   orc_bp (gpu_strict)                  brute-force pure-Python bilinear on tiny inputs,
 *                                        constant-projection closed form, mirror symmetry
 *   orc_fp (cpu_zero)                    golden from the reference 3dct ForwardProjection
 *   orc_fp (gpu_clamp)                   analytic ellipsoid line integrals, constant
 *                                        volume chord lengths, agreement with cpu_zero
 *                                        inside the box
 *   orc_phantom_*                        closed-form chord lengths, volume of an ellipsoid
 */
#include "fdk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double kPi = 3.141592653589793238462;

/* type.h:4418-4425 -- the power of two strictly greater than n (bit count). */
int orc_nextpow2(int n) {
  unsigned m = (unsigned)n;
  int bits = 0;
  while (m > 0) { m >>= 1; ++bits; }
  return 1 << bits;
}

/* type.h:4427-4475 -- ramp 2k/N times a window, mirrored to length N, entries
 * above pi*d zeroed.  The shepp-logan and hamming windows are stated exactly as
 * the reference evaluates them (sin(w)/w and 0.54 + 0.46*cos(w)/d). */
int orc_filter_design(int type, int len, int d, double* out, int cap) {
  const int order = orc_nextpow2(len);
  const int half = order / 2;
  if (order > cap) return -1;
  for (int i = 0; i <= half; ++i) {
    double f = i * 2.0 / order;
    const double w = 2.0 * kPi * i / order;
    if (i > 0) {
      switch (type) {
        case 0: case 1: break;                                   /* none, ram-lak */
        case 2: f *= sin(w) / w; break;                          /* shepp-logan */
        case 3: f *= cos(w / (2.0 * d)); break;                  /* cosine */
        case 4: f *= 0.54 + 0.46 * cos(w) / d; break;            /* hamming */
        case 5: f *= (1.0 + cos(w / d)) / 2.0; break;            /* hann */
        default: return -2;
      }
    }
    out[i] = f;
  }
  for (int i = half + 1; i < order; ++i) out[i] = out[order - i];
  for (int i = 0; i < order; ++i)
    if (out[i] > kPi * d) out[i] = 0.0;
  return order;
}

/* type.h:4736 -- 1/2/du * (2 pi / nProj) / 2 * (SID/SOD), reference names. */
double orc_angular_ramp_factor(const RefGeom* g) {
  return 1.0 / 2.0 / g->du * (2.0 * kPi / g->nproj) / 2.0 * (g->sid / g->sod);
}

/* type.h:4756-4785 with DetectorCoor<float> (type.h:4019-4051): the reference
 * evaluates this table in float with a double square root; the same roundings
 * are applied here so the table is reproduced to the last bit. */
void orc_cos_table(const RefGeom* g, float* out) {
  const float du = (float)g->du, dv = (float)g->dv;
  const float cx = (float)(((float)g->nu - 1) / 2.0 + (float)g->offset_nu);
  const float cy = (float)(((float)g->nv - 1) / 2.0 + (float)g->offset_nv);
  const double af = orc_angular_ramp_factor(g);
  const double scale = g->sod * g->sod * af;
  for (int y = 0; y < g->nv; ++y) {
    for (int x = 0; x < g->nu; ++x) {
      const float u = (float)x - cx;
      const float v = cy - (float)y;
      const float umm = u * du, vmm = v * dv;
      const float dis = (float)sqrt((double)(umm * umm + vmm * vmm) + g->sid * g->sid);
      float w = (float)(g->sid / dis);
      w = (float)(w * scale);
      out[(size_t)y * g->nu + x] = w;
    }
  }
}

/* type.h:4732-4733, 4824: direction * (s * scan_angle / nProj) deg + vpa, in rad. */
double orc_view_angle(const RefGeom* g, int view) {
  const double deg = view * g->scan_angle_deg / (double)g->nproj;
  return g->direction * deg / 180.0 * kPi + g->volume_position_angle_rad;
}

static void mat4_mul(const double* a, const double* b, double* c) {
  double t[16];
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k) {
      double s = 0.0;
      for (int q = 0; q < 4; ++q) s += a[r * 4 + q] * b[q * 4 + k];
      t[r * 4 + k] = s;
    }
  memcpy(c, t, sizeof(t));
}

/* type.h:4898-4944 -- M = S_det^-1 * P(SID) * A(SOD) * R(theta) * D(dx,dy,dz) * C,
 * acting on (i, j, k, 1) voxel indices; rows give (u*w, v*w, w), u, v in detector
 * pixels and w in mm.  Built in FP64 (the reference uses float). */
void orc_proj_matrix(const RefGeom* g, double a, double* m12) {
  const double c = cos(a), s = sin(a);
  const double m0[16] = {1, 0, 0, -(g->nx - 1) / 2.0, 0, -1, 0, (g->ny - 1) / 2.0,
                         0, 0, -1, (g->nz - 1) / 2.0, 0, 0, 0, 1};
  const double m1[16] = {g->dx, 0, 0, 0, 0, g->dy, 0, 0, 0, 0, g->dz, 0, 0, 0, 0, 1};
  const double m2[16] = {c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  const double m3[16] = {1, 0, 0, 0, 0, 0, -1, 0, 0, 1, 0, g->sod, 0, 0, 0, 1};
  const double m4[16] = {g->sid, 0, (g->nu - 1) / 2.0 * g->du, 0,
                         0, g->sid, (g->nv - 1) / 2.0 * g->dv, 0,
                         0, 0, 1, 0, 0, 0, 0, 1};
  const double m5[16] = {1.0 / g->du, 0, 0, 0, 0, 1.0 / g->dv, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  double t[16];
  mat4_mul(m5, m4, t);
  mat4_mul(t, m3, t);
  mat4_mul(t, m2, t);
  mat4_mul(t, m1, t);
  mat4_mul(t, m0, t);
  memcpy(m12, t, 12 * sizeof(double));
}

/* type.h:4821-4830 -- one matrix per view. */
void orc_proj_matrices(const RefGeom* g, double* out) {
  for (int s = 0; s < g->nproj; ++s) orc_proj_matrix(g, orc_view_angle(g, s), out + 12 * s);
}

/* Radix-2 complex DFT, n = 2^m points, in place.  dir = +1: X[k] = sum x[n]
 * e^{-2 pi i nk/N}; dir = -1: inverse including the 1/N (the convention of the
 * reference FFT, type.h:4099-4166, and of IPP_FFT_DIV_INV_BY_N, IppLib.h:76). */
void orc_fft(int dir, int m, double* re, double* im) {
  const int n = 1 << m;
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  double* twr = (double*)malloc(sizeof(double) * (n / 2 + 1));
  double* twi = (double*)malloc(sizeof(double) * (n / 2 + 1));
  for (int k = 0; k < n / 2; ++k) {
    twr[k] = cos(2.0 * kPi * k / n);
    twi[k] = -dir * sin(2.0 * kPi * k / n);
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int h = len / 2, stride = n / len;
    for (int b = 0; b < n; b += len) {
      for (int j = 0; j < h; ++j) {
        const double wr = twr[j * stride], wi = twi[j * stride];
        const int p = b + j, q = p + h;
        const double tr = wr * re[q] - wi * im[q];
        const double ti = wr * im[q] + wi * re[q];
        re[q] = re[p] - tr; im[q] = im[p] - ti;
        re[p] += tr; im[p] += ti;
      }
    }
  }
  free(twr);
  free(twi);
  if (dir == -1)
    for (int i = 0; i < n; ++i) { re[i] /= n; im[i] /= n; }
}

/* FilterLib.h:8-13 / cudaFftFilter.cu:270-281 -- dst = src * table per view
 * (one float multiply, so any correct implementation is bit-exact). */
void orc_cos_weight(const float* src, const float* table, float* dst, size_t n, int nviews) {
  for (int v = 0; v < nviews; ++v)
    for (size_t i = 0; i < n; ++i) dst[(size_t)v * n + i] = src[(size_t)v * n + i] * table[i];
}

/* FilterLib.h:15-58 (mean4), type.h:4169-4239 (mean3), zero padding.
 * Each row of nu samples is centred in N at offset (N-nu)/2, the pad regions take
 * the mean of the first / last pad_mode samples (zero for pad_mode 0), the row is
 * transformed, multiplied by the real filter H, inverse transformed with 1/N and
 * the centre nu samples are kept.  FP64 throughout (the reference runs IPP FP32;
 * mean4 is formed in float as FilterLib.h:38,44 does). */
void orc_filter_rows(const float* src, float* dst, int nu, long nrows, const double* H, int N,
                     int pad_mode) {
  int m = 0;
  while ((1 << m) < N) ++m;
  const int off = (N - nu) >> 1;
#pragma omp parallel
  {
    double* re = (double*)malloc(sizeof(double) * N);
    double* im = (double*)malloc(sizeof(double) * N);
#pragma omp for schedule(static)
    for (long r = 0; r < nrows; ++r) {
      const float* s = src + (size_t)r * nu;
      double left = 0.0, right = 0.0;
      if (pad_mode == ORC_PAD_MEAN4) {
        left = (s[0] + s[1] + s[2] + s[3]) / 4;
        right = (s[nu - 1] + s[nu - 2] + s[nu - 3] + s[nu - 4]) / 4;
      } else if (pad_mode == ORC_PAD_MEAN3) {
        left = ((double)s[0] + (double)s[1] + (double)s[2]) / 3;
        right = ((double)s[nu - 1] + (double)s[nu - 2] + (double)s[nu - 3]) / 3;
      }
      for (int i = 0; i < N; ++i) {
        re[i] = i < off ? left : (i < off + nu ? (double)s[i - off] : right);
        im[i] = 0.0;
      }
      orc_fft(1, m, re, im);
      for (int i = 0; i < N; ++i) { re[i] *= H[i]; im[i] *= H[i]; }
      orc_fft(-1, m, re, im);
      for (int i = 0; i < nu; ++i) dst[(size_t)r * nu + i] = (float)re[off + i];
    }
    free(re);
    free(im);
  }
}

/* One backprojected sample of view image P (nv rows of nu, u fastest) at
 * detector position (u, v), in one of the reference's two conventions:
 *   gpu_strict  backprojectionFDK.cu:15-48 -- floor, zero unless 0<=u<nu-1 and
 *               0<=v<nv-1 (the same test as RtkSUBPIXEL, rtk...cu:83-112);
 *   cpu_trunc   type.h:2160-2207 -- truncation toward zero, valid if the
 *               truncated indices are inside the image, out-of-image taps are 0. */
static double sample2d(const float* P, int nu, int nv, double u, double v, int mode) {
  if (mode == ORC_BP_GPU_STRICT) {
    if (!(u >= 0 && u < nu - 1 && v >= 0 && v < nv - 1)) return 0.0;
    const double fu0 = floor(u), fv0 = floor(v);
    const int iu = (int)fu0, iv = (int)fv0;
    const double fu = u - fu0, fv = v - fv0;
    const float* p0 = P + (size_t)iv * nu + iu;
    const float* p1 = p0 + nu;
    return (1 - fv) * ((1 - fu) * p0[0] + fu * p0[1]) + fv * ((1 - fu) * p1[0] + fu * p1[1]);
  }
  const int au = (int)u, av = (int)v;
  if (!(au >= 0 && au < nu && av >= 0 && av < nv)) return 0.0;
  const double cu = u - au, cv = v - av;
  double d[4] = {0, 0, 0, 0};
  for (int q = 0; q < 4; ++q) {
    const int x = au + (q & 1), y = av + (q >> 1);
    if (x >= 0 && x < nu && y >= 0 && y < nv) d[q] = P[(size_t)y * nu + x];
  }
  return (d[0] * (1 - cu) + d[1] * cu) * (1 - cv) + (d[2] * (1 - cu) + d[3] * cu) * cv;
}

/* Voxel-driven FDK backprojection (backprojectionFDK.cu:50-129, 3dct.cpp:66-99):
 * for each view s, (uw, vw, w) = M_s (i, j, k, 1), sample at (uw/w, vw/w), weight
 * 1/w^2, summed over views.  proj: nviews x nv x nu (u fastest, not transposed).
 * vol: nzs x ny x nx doubles for k in [z0, z0+nzs), accumulated into. */
void orc_bp(const RefGeom* g, const float* proj, const double* mats, int nviews, int mode,
            int z0, int nzs, double* vol) {
  const size_t frame = (size_t)g->nu * g->nv;
#pragma omp parallel for schedule(dynamic, 1)
  for (int kk = 0; kk < nzs; ++kk) {
    const int k = z0 + kk;
    for (int j = 0; j < g->ny; ++j) {
      double* row = vol + ((size_t)kk * g->ny + j) * g->nx;
      for (int i = 0; i < g->nx; ++i) {
        double acc = 0.0;
        for (int s = 0; s < nviews; ++s) {
          const double* m = mats + 12 * s;
          const double w = m[8] * i + m[9] * j + m[10] * k + m[11];
          const double z = 1.0 / w;
          const double u = (m[0] * i + m[1] * j + m[2] * k + m[3]) * z;
          const double v = (m[4] * i + m[5] * j + m[6] * k + m[7]) * z;
          acc += sample2d(proj + s * frame, g->nu, g->nv, u, v, mode) * z * z;
        }
        row[i] += acc;
      }
    }
  }
}

/* orc_bp for an explicit list of voxels (ijk: nvox x 3 ints, i = x index);
 * lets tests check full-size volumes on a sample of voxels. */
void orc_bp_voxels(const RefGeom* g, const float* proj, const double* mats, int nviews, int mode,
                   const int* ijk, long nvox, double* out) {
  const size_t frame = (size_t)g->nu * g->nv;
#pragma omp parallel for schedule(static)
  for (long q = 0; q < nvox; ++q) {
    const int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
    double acc = 0.0;
    for (int s = 0; s < nviews; ++s) {
      const double* m = mats + 12 * s;
      const double w = m[8] * i + m[9] * j + m[10] * k + m[11];
      const double z = 1.0 / w;
      const double u = (m[0] * i + m[1] * j + m[2] * k + m[3]) * z;
      const double v = (m[4] * i + m[5] * j + m[6] * k + m[7]) * z;
      acc += sample2d(proj + s * frame, g->nu, g->nv, u, v, mode) * z * z;
    }
    out[q] += acc;
  }
}

/* Ray geometry shared by the forward projectors (cudaForwardProjection.cu:
 * 338-371, 174-210; 3dct.cpp:436-502): source R(a)(0,-SOD,0) and pixel
 * R(a)(u_mm, SID-SOD, v_mm) with a = -(direction*deg + vpa). */
typedef struct { double o[3], d[3]; } OrcRay;

static OrcRay pixel_ray(const RefGeom* g, int view, int i, int j) {
  const double a = -orc_view_angle(g, view);
  const double c = cos(a), s = sin(a);
  const double cxd = (g->nu - 1) / 2.0 + g->offset_nu, cyd = (g->nv - 1) / 2.0 + g->offset_nv;
  const double umm = (i - cxd) * g->du, vmm = (cyd - j) * g->dv;
  const double dety = g->sid - g->sod;
  OrcRay r;
  r.o[0] = g->sod * s; r.o[1] = -g->sod * c; r.o[2] = 0.0;
  const double p[3] = {umm * c - dety * s, umm * s + dety * c, vmm};
  double n = 0.0;
  for (int q = 0; q < 3; ++q) { r.d[q] = p[q] - r.o[q]; n += r.d[q] * r.d[q]; }
  n = sqrt(n);
  for (int q = 0; q < 3; ++q) r.d[q] /= n;
  return r;
}

/* Slab test (type.h:1597-1619, cudaForwardProjection.cu:28-48). */
static int intersect_box(const OrcRay* r, const double* bmin, const double* bmax, double* tn,
                         double* tf) {
  double tmin[3], tmax[3];
  for (int q = 0; q < 3; ++q) {
    const double inv = 1.0 / r->d[q];
    const double a = inv * (bmin[q] - r->o[q]), b = inv * (bmax[q] - r->o[q]);
    tmin[q] = a < b ? a : b;
    tmax[q] = a < b ? b : a;
  }
  double lo = tmin[0] > tmin[1] ? tmin[0] : tmin[1];
  if (tmin[2] > lo) lo = tmin[2];
  double hi = tmax[0] < tmax[1] ? tmax[0] : tmax[1];
  if (tmax[2] < hi) hi = tmax[2];
  *tn = lo;
  *tf = hi;
  return hi > lo;
}

/* Trilinear at index-space point p of vol (nz x ny x nx, x fastest), indices
 * truncated toward zero; taps outside are clamped to the edge (gpu_clamp,
 * cudaIterType.cuh:231-278 over a clamp-addressed point texture) or read as 0
 * (cpu_zero, type.h:2658-2712). */
static double trilinear(const float* vol, int nx, int ny, int nz, const double* p, int mode) {
  const int a[3] = {(int)p[0], (int)p[1], (int)p[2]};
  const double c[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  const int n[3] = {nx, ny, nz};
  double t[8];
  for (int q = 0; q < 8; ++q) {
    int idx[3] = {a[0] + (q & 1), a[1] + ((q >> 1) & 1), a[2] + (q >> 2)};
    int ok = 1;
    for (int e = 0; e < 3; ++e) {
      if (mode == ORC_FP_GPU_CLAMP) {
        if (idx[e] < 0) idx[e] = 0;
        if (idx[e] > n[e] - 1) idx[e] = n[e] - 1;
      } else if (idx[e] < 0 || idx[e] > n[e] - 1) {
        ok = 0;
      }
    }
    t[q] = ok ? vol[((size_t)idx[2] * ny + idx[1]) * nx + idx[0]] : 0.0;
  }
  const double v0 = (t[0] * (1 - c[0]) + t[1] * c[0]) * (1 - c[1]) + (t[2] * (1 - c[0]) + t[3] * c[0]) * c[1];
  const double v1 = (t[4] * (1 - c[0]) + t[5] * c[0]) * (1 - c[1]) + (t[6] * (1 - c[0]) + t[7] * c[0]) * c[1];
  return v0 * (1 - c[2]) + v1 * c[2];
}

/* Ray-marching forward projection of one view (out: nv x nu, u fastest).
 * Box = voxel-centre extents; march from tnear with step step_mm and sum the
 * trilinear samples, times step_mm.  gpu_clamp stops at t < tfar
 * (cudaForwardProjection.cu:218-245); cpu_zero at t <= tfar (3dct.cpp:534). */
void orc_fp(const RefGeom* g, const float* vol, int view, double step_mm, int mode, float* out) {
  const double cx = (g->nx - 1) / 2.0, cy = (g->ny - 1) / 2.0, cz = (g->nz - 1) / 2.0;
  const double bmin[3] = {-cx * g->dx, -cy * g->dy, -cz * g->dz};
  const double bmax[3] = {cx * g->dx, cy * g->dy, cz * g->dz};
#pragma omp parallel for schedule(dynamic, 4)
  for (int j = 0; j < g->nv; ++j) {
    for (int i = 0; i < g->nu; ++i) {
      const OrcRay r = pixel_ray(g, view, i, j);
      double tn, tf, sum = 0.0;
      if (intersect_box(&r, bmin, bmax, &tn, &tf)) {
        const double o[3] = {r.o[0] / g->dx + cx, cy - r.o[1] / g->dy, cz - r.o[2] / g->dz};
        const double st[3] = {r.d[0] / g->dx, -r.d[1] / g->dy, -r.d[2] / g->dz};
        for (double t = tn; mode == ORC_FP_GPU_CLAMP ? t < tf : t <= tf; t += step_mm) {
          const double p[3] = {o[0] + st[0] * t, o[1] + st[1] * t, o[2] + st[2] * t};
          sum += trilinear(vol, g->nx, g->ny, g->nz, p, mode);
        }
      }
      out[(size_t)j * g->nu + i] = (float)(sum * step_mm);
    }
  }
}

/* Additive ellipsoid phantom sampled at voxel centres (volume frame of
 * type.h:4053-4087: x = (i-cx)dx, y = (cy-j)dy, z = (cz-k)dz). */
void orc_phantom_rasterize(const OrcEllipsoid* e, int n, const RefGeom* g, float* vol) {
  const double cx = (g->nx - 1) / 2.0, cy = (g->ny - 1) / 2.0, cz = (g->nz - 1) / 2.0;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < g->nz; ++k)
    for (int j = 0; j < g->ny; ++j)
      for (int i = 0; i < g->nx; ++i) {
        const double p[3] = {(i - cx) * g->dx, (cy - j) * g->dy, (cz - k) * g->dz};
        double acc = 0.0;
        for (int q = 0; q < n; ++q) {
          const OrcEllipsoid* E = e + q;
          const double d[3] = {p[0] - E->cx, p[1] - E->cy, p[2] - E->cz};
          double r2 = 0.0;
          const double ax[3] = {E->ax, E->ay, E->az};
          for (int c = 0; c < 3; ++c) {
            const double l = (E->rot[c] * d[0] + E->rot[3 + c] * d[1] + E->rot[6 + c] * d[2]) / ax[c];
            r2 += l * l;
          }
          if (r2 <= 1.0) acc += E->density;
        }
        vol[((size_t)k * g->ny + j) * g->nx + i] = (float)acc;
      }
}

/* Exact line integrals of the ellipsoid phantom along the rays of one view. */
void orc_phantom_project(const OrcEllipsoid* e, int n, const RefGeom* g, int view, float* out) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < g->nv; ++j)
    for (int i = 0; i < g->nu; ++i) {
      const OrcRay r = pixel_ray(g, view, i, j);
      double acc = 0.0;
      for (int q = 0; q < n; ++q) {
        const OrcEllipsoid* E = e + q;
        const double d0[3] = {r.o[0] - E->cx, r.o[1] - E->cy, r.o[2] - E->cz};
        const double ax[3] = {E->ax, E->ay, E->az};
        double A = 0, B = 0, C = -1.0;
        for (int c = 0; c < 3; ++c) {
          const double oc = (E->rot[c] * d0[0] + E->rot[3 + c] * d0[1] + E->rot[6 + c] * d0[2]) / ax[c];
          const double dc = (E->rot[c] * r.d[0] + E->rot[3 + c] * r.d[1] + E->rot[6 + c] * r.d[2]) / ax[c];
          A += dc * dc; B += 2 * oc * dc; C += oc * oc;
        }
        const double disc = B * B - 4 * A * C;
        if (disc > 0) acc += E->density * sqrt(disc) / A;
      }
      out[(size_t)j * g->nu + i] = (float)acc;
    }
}
