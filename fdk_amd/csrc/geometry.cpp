// Scan-geometry tables on the host in FP64: the reference's Parameter::init
// (common/type.h:4719-4754) and the functions it calls.  Projection matrices are
// built in closed form rather than as the reference's six-matrix float product
// (type.h:4898-4944); both give the same map (x, y, z, 1) -> (u w, v w, w).
#include <cmath>

#include "common.hpp"

namespace {

constexpr double kPi = 3.141592653589793238462;

int check_geom(const fdk_geometry* g) {
  FDK_REQUIRE(g != nullptr, "geometry is NULL");
  FDK_REQUIRE(g->nu > 0 && g->nv > 0 && g->nproj > 0, "nu, nv and nproj must be positive");
  FDK_REQUIRE(g->du > 0 && g->dv > 0 && g->sod > 0 && g->sid > 0, "du, dv, sod, sid must be positive");
  return FDK_OK;
}

// type.h:4418-4425: the power of two strictly greater than n.
int next_pow2_strict(int n) {
  int bits = 0;
  for (unsigned m = (unsigned)n; m; m >>= 1) ++bits;
  return 1 << bits;
}

}  // namespace

extern "C" int fdk_geom_angular_ramp_factor(const fdk_geometry* g, double* out) {
  if (int r = check_geom(g)) return r;
  FDK_REQUIRE(out, "out is NULL");
  // type.h:4736: 1/(2 du) * (2 pi / nProj) / 2 * SID/SOD  (reference names)
  *out = 1.0 / 2.0 / g->du * (2.0 * kPi / g->nproj) / 2.0 * (g->sid / g->sod);
  return FDK_OK;
}

extern "C" int fdk_geom_cos_table(const fdk_geometry* g, float* out) {
  if (int r = check_geom(g)) return r;
  FDK_REQUIRE(out, "out is NULL");
  double af = 0;
  fdk_geom_angular_ramp_factor(g, &af);
  // type.h:4756-4785 evaluates this in float (DetectorCoor<float>) with a double
  // square root; the same roundings are kept so the table matches bit for bit.
  const float du = (float)g->du, dv = (float)g->dv;
  const float cu = (float)(((float)g->nu - 1) / 2.0 + (float)g->offset_nu);
  const float cv = (float)(((float)g->nv - 1) / 2.0 + (float)g->offset_nv);
  const double scale = g->sod * g->sod * af;
  const double sid2 = g->sid * g->sid;
  for (int y = 0; y < g->nv; ++y) {
    const float vmm = (cv - (float)y) * dv;
    for (int x = 0; x < g->nu; ++x) {
      const float umm = ((float)x - cu) * du;
      const float dist = (float)std::sqrt((double)(umm * umm + vmm * vmm) + sid2);
      const float w = (float)(g->sid / dist);
      out[(size_t)y * g->nu + x] = (float)(w * scale);
    }
  }
  return FDK_OK;
}

extern "C" int fdk_geom_filter_length(const fdk_geometry* g) {
  if (!g || g->nu <= 0) return fdk::fail(FDK_ERR_INVALID, "invalid geometry");
  return next_pow2_strict(g->nu);
}

extern "C" int fdk_geom_filter(const fdk_geometry* g, double* out, int cap) {
  if (int r = check_geom(g)) return r;
  FDK_REQUIRE(g->filter_type >= 0 && g->filter_type <= 5, "filter_type %d not in 0..5", g->filter_type);
  const int n = next_pow2_strict(g->nu);
  FDK_REQUIRE(out && cap >= n, "filter buffer holds %d, need %d", cap, n);
  const int d = 1;  // Parameter::init passes d = 1 (type.h:4745)
  // type.h:4427-4475: ramp 2k/N for k <= N/2 times the window, mirrored; the
  // shepp-logan and hamming windows are kept as the reference evaluates them.
  for (int k = 0; k <= n / 2; ++k) {
    const double w = 2.0 * kPi * k / n;
    double h = 2.0 * k / n;
    if (k > 0) {
      switch (g->filter_type) {
        case 2: h *= std::sin(w) / w; break;
        case 3: h *= std::cos(w / (2.0 * d)); break;
        case 4: h *= 0.54 + 0.46 * std::cos(w) / d; break;
        case 5: h *= 0.5 * (1.0 + std::cos(w / d)); break;
        default: break;  // none, ram-lak
      }
    }
    out[k] = h > kPi * d ? 0.0 : h;
    if (k > 0 && k < n / 2) out[n - k] = out[k];
  }
  return n;
}

extern "C" int fdk_geom_view_angle(const fdk_geometry* g, int view, double* out) {
  if (int r = check_geom(g)) return r;
  FDK_REQUIRE(out, "out is NULL");
  const double deg = view * g->scan_angle_deg / (double)g->nproj;  // type.h:4733
  *out = g->direction * deg / 180.0 * kPi + g->volume_position_angle_rad;  // type.h:4824
  return FDK_OK;
}

extern "C" int fdk_geom_proj_matrices(const fdk_geometry* g, int first, int count, double* out) {
  if (int r = check_geom(g)) return r;
  FDK_REQUIRE(out && count >= 0, "bad output");
  FDK_REQUIRE(g->nx > 0 && g->ny > 0 && g->nz > 0, "volume size must be positive");
  // Voxel (i, j, k) sits at x = dx (i - cx), y = -dy (j - cy), z = -dz (k - cz)
  // (type.h:4053-4087).  Rotating by theta about z gives x' = x c - y s,
  // y' = x s + y c; then w = y' + SOD is the depth along the central ray,
  // u = SID x' / (du w) + (nu-1)/2 and v = -SID z / (dv w) + (nv-1)/2.
  const double cx = (g->nx - 1) / 2.0, cy = (g->ny - 1) / 2.0, cz = (g->nz - 1) / 2.0;
  const double cu = (g->nu - 1) / 2.0, cv = (g->nv - 1) / 2.0;
  for (int q = 0; q < count; ++q) {
    double th;
    fdk_geom_view_angle(g, first + q, &th);
    const double c = std::cos(th), s = std::sin(th);
    const double w[4] = {s * g->dx, -c * g->dy, 0.0, g->sod - s * g->dx * cx + c * g->dy * cy};
    const double xu[4] = {c * g->dx, s * g->dy, 0.0, -c * g->dx * cx - s * g->dy * cy};
    const double zv[4] = {0.0, 0.0, g->dz, -g->dz * cz};
    double* m = out + 12 * (size_t)q;
    for (int e = 0; e < 4; ++e) {
      m[e] = g->sid / g->du * xu[e] + cu * w[e];
      m[4 + e] = g->sid / g->dv * zv[e] + cv * w[e];
      m[8 + e] = w[e];
    }
  }
  return FDK_OK;
}
