// Command-line drivers over the C++ drop-in layer (include/fdk_amd.hpp), in the
// shape of the reference's own drivers: cuFDK's filter thread + BP thread
// (cuFDK/mpiFDK.cpp:58-154, 167-317), cudaBP's TestFDK (cudaBP/cudaBP.cpp:85-153)
// and genProj (genProj/genProj.cpp:31-81).
//
//   fdk_drivers fdk <n> <views> <out_prefix>
//       n^3 volume, n^2 detector, ellipsoid phantom projected exactly on the
//       GPU, FilterEngine (cos + mean-4 pad + ramp/Hann, transposed) per view,
//       BackProjectionFDK::BP(1, ...) per 32 views, GetVolumeData.  Writes
//       <out_prefix>.proj.raw (views x n x n floats, as projected) and
//       <out_prefix>.vol.raw (n^3 floats, [z][y][x]); prints GUPS.
//   fdk_drivers fp <n> <views> <out_prefix>
//       cudaCTForwardProjection_f32 of a rasterized phantom, all views.
//   fdk_drivers rtk <n> <views> <out_prefix>
//       the fdk path with BackProjectionRTK (untransposed views).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fdk_amd.hpp"

namespace {

Parameter make_param(int n, int views) {
  Parameter p;
  p.nx = p.ny = p.nz = n;
  p.nu = p.nv = n;
  p.du = p.dv = 409.6 / n;
  p.dx = p.dy = p.dz = p.du * 1000.0 / 1500.0;
  p.SOD = 1000.0;
  p.SID = 1500.0;
  p.nProj = views;
  p.direction = -1;
  p.scan_angle = 360.0;
  p.volume_postion_angle = 0.0;
  p.str_filter = "hann";
  p.init();
  return p;
}

std::vector<fdk_ellipsoid> phantom(double fov) {
  // three fixed ellipsoids inside the field of view
  std::vector<fdk_ellipsoid> e(3);
  const double id[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double c = std::cos(0.5), s = std::sin(0.5);
  const double rz[9] = {c, -s, 0, s, c, 0, 0, 0, 1};
  e[0] = {0, 0, 0, 0.35 * fov, 0.3 * fov, 0.32 * fov, {0}, 1.0};
  e[1] = {0.1 * fov, -0.05 * fov, 0.05 * fov, 0.1 * fov, 0.05 * fov, 0.08 * fov, {0}, -0.4};
  e[2] = {-0.12 * fov, 0.08 * fov, -0.1 * fov, 0.06 * fov, 0.09 * fov, 0.04 * fov, {0}, 0.7};
  std::memcpy(e[0].rot, id, sizeof(id));
  std::memcpy(e[1].rot, rz, sizeof(rz));
  std::memcpy(e[2].rot, id, sizeof(id));
  return e;
}

void write_raw(const std::string& path, const float* d, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::string("cannot write ") + path;
  std::fwrite(d, sizeof(float), n, f);
  std::fclose(f);
}

int run_fdk(int n, int views, const std::string& out, bool rtk) {
  Parameter para = make_param(n, views);
  const fdk_geometry g = para.geometry();
  const size_t view = (size_t)n * n;
  ProjectionImage proj(n, n, views);
  {
    std::vector<fdk_ellipsoid> e = phantom(n * para.dx);
    float* d = nullptr;
    if (hipMalloc(&d, view * views * 4) != hipSuccess) throw std::string("hipMalloc");
    fdk_amd_detail::check(fdk_phantom_project(e.data(), (int)e.size(), &g, 0, views, d, nullptr),
                          "phantom_project");
    if (hipMemcpy(proj.buffer(), d, view * views * 4, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::string("hipMemcpy");
    (void)hipFree(d);
  }
  write_raw(out + ".proj.raw", proj.buffer(), view * views);
  // Filter thread: FilterEngine::Filtering(src, dst, transpose) per view.
  FrameT<float> cos((float*)para.cos_sigma.data(), (int)para.nu, (int)para.nv, 0);
  FilterEngine engine(cos, para.fft_kernel_real_f32);
  std::vector<float> filtered(view * views);
  std::vector<float> src(view);
  for (int s = 0; s < views; ++s) {
    std::memcpy(src.data(), proj.buffer() + view * s, view * 4);
    FrameT<float> fs(src.data(), n, n, s);
    if (rtk) {
      FrameT<float> fd(filtered.data() + view * s, n, n, s);
      engine.Filtering(fs, fd, false);
    } else {
      FrameT<float> fd(filtered.data() + view * s, n, n, s);  // transposed: nv wide, nu tall
      engine.Filtering(fs, fd, true);
    }
  }
  std::vector<float> vol(view * n);
  const auto t0 = std::chrono::steady_clock::now();
  float kernel_ms = 0.f;
  if (!rtk) {
    // BP thread: 32-view batches (PushBpData, cuFDK/mpiFDK.cpp:167-203).
    geo::BoxT<size_t> top(0, 0, 0, n, n, n / 2), bottom(0, 0, n / 2, n, n, n / 2);
    BackProjectionFDK bp((int)para.nv, (int)para.nu, cudaBackProjection::MAX_PROJ, n, n, n, top,
                         bottom);
    for (int s = 0; s < views; s += cudaBackProjection::MAX_PROJ) {
      const int cnt = std::min<int>(cudaBackProjection::MAX_PROJ, views - s);
      bp.BP(1, filtered.data() + view * s, (int)para.nv, (int)para.nu,
            para.vec_rot_mat_ijk3x4.data() + 12 * s, cnt);
    }
    bp.GetVolumeData(vol.data(), vol.size());
    kernel_ms = bp.fKernelTime;
    if (bp.nBpCount != views) throw std::string("nBpCount mismatch");
  } else {
    BackProjectionRTK bp((int)para.nu, (int)para.nv, cudaBackProjection::MAX_PROJ, n, n, n);
    for (int s = 0; s < views; s += cudaBackProjection::MAX_PROJ) {
      const int cnt = std::min<int>(cudaBackProjection::MAX_PROJ, views - s);
      bp.BP(0, filtered.data() + view * s, (int)para.nu, (int)para.nv,
            para.vec_rot_mat_ijk3x4.data() + 12 * s, cnt);
    }
    bp.GetVolumeData(vol.data(), vol.size());
    kernel_ms = bp.fKernelTime;
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  write_raw(out + ".vol.raw", vol.data(), vol.size());
  double sum = 0.0;
  for (float v : vol) sum += v;
  const double upd = (double)n * n * n * views;
  std::printf("%s: %d^3 from %d views, BP kernels %.3f ms (%.1f GUPS), BP calls incl. upload %.3f s, "
              "volume sum %.6e\n", rtk ? "rtk" : "fdk", n, views, kernel_ms,
              upd / (kernel_ms * 1e-3) / 1e9, wall, sum);
  return 0;
}

int run_fp(int n, int views, const std::string& out) {
  Parameter para = make_param(n, views);
  const fdk_geometry g = para.geometry();
  const size_t nvox = (size_t)n * n * n;
  std::vector<float> vol(nvox);
  {
    std::vector<fdk_ellipsoid> e = phantom(n * para.dx);
    float* d = nullptr;
    if (hipMalloc(&d, nvox * 4) != hipSuccess) throw std::string("hipMalloc");
    fdk_amd_detail::check(fdk_phantom_rasterize(e.data(), (int)e.size(), &g, d, nullptr), "raster");
    if (hipMemcpy(vol.data(), d, nvox * 4, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::string("hipMemcpy");
    (void)hipFree(d);
  }
  ProjectionImage proj(n, n, views);
  cudaCTForwardProjection_f32(para, vol, proj, -1, 0.5 * para.dx);
  write_raw(out + ".vol.raw", vol.data(), nvox);
  write_raw(out + ".fp.raw", proj.buffer(), (size_t)n * n * views);
  std::printf("fp: %d views of %d^2 through %d^3\n", views, n, n);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s fdk|rtk|fp <n> <views> <out_prefix>\n", argv[0]);
    return 2;
  }
  const std::string mode = argv[1];
  const int n = std::atoi(argv[2]), views = std::atoi(argv[3]);
  if (n < 8 || n % 2 || views < 1) {
    std::fprintf(stderr, "need an even n >= 8 and views >= 1\n");
    return 2;
  }
  try {
    if (mode == "fdk") return run_fdk(n, views, argv[4], false);
    if (mode == "rtk") return run_fdk(n, views, argv[4], true);
    if (mode == "fp") return run_fp(n, views, argv[4]);
  } catch (const std::string& e) {
    std::fprintf(stderr, "error: %s\n", e.c_str());
    return 1;
  }
  std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
  return 2;
}
