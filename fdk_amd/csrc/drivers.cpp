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
//       cudaCTForwardProjection_f32 of a rasterized phantom, all views, called
//       as genProj does (genProj/genProj.cpp:60-75): device volume, one device
//       image per view, the projection copied into an ImageT.
//   fdk_drivers rtk <n> <views> <out_prefix>
//       the fdk path with BackProjectionRTK (untransposed views).
//   fdk_drivers filter <n> <views> <out_prefix>
//       FilterlibCufft's test_filter (FilterlibCufft/FilterlibCufft.cpp:8-20):
//       LoadFolder from str_proj_img_folder, FilterProjImage (returns ms),
//       SaveToFolder into str_working_folder.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <memory>
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
    if (hipMemcpy(proj.Buffer(), d, view * views * 4, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::string("hipMemcpy");
    (void)hipFree(d);
  }
  write_raw(out + ".proj.raw", proj.Buffer(), view * views);
  // Filter thread (cuFDK/mpiFDK.cpp:104-127): per view an ImageT, FrameT views
  // of it and of the transposed destination, one FilterEngine built from the
  // parameter's cos table.
  std::shared_ptr<FilterEngine> engine;
  std::vector<float> filtered(view * views);
  for (int s = 0; s < views; ++s) {
    ImageT<float> img(para.nu, para.nv);
    std::memcpy(img.buffer, proj.GetFrame(s).buffer, view * 4);
    FrameT<float> src((float*)img.buffer, para.nu, para.nv, 0);
    if (!engine)
      engine = std::make_shared<FilterEngine>(para.cos_image_table.cos_sigma.GetFrame(),
                                              para.fft_kernel_real_f32);
    if (rtk) {
      FrameT<float> dst(filtered.data() + view * s, para.nu, para.nv, 0);
      engine->Filtering(src, dst, false);
    } else {
      FrameT<float> dst(filtered.data() + view * s, para.nv, para.nu, 0);  // transposed
      engine->Filtering(src, dst, true);
    }
  }
  VolumeT<float> vol;
  const auto t0 = std::chrono::steady_clock::now();
  float kernel_ms = 0.f;
  if (!rtk) {
    // BP thread (cuFDK/mpiFDK.cpp:167-317): size_t sizes, a mirror slab pair,
    // 32-view batches of matrices taken from vec_rot_mat_ijk3x4 (PushBpData).
    size_t _projWidth = para.nv, _projHeight = para.nu;
    size_t _projCount = cudaBackProjection::MAX_PROJ;
    size_t _nx = para.nx, _ny = para.ny, _nz = para.nz;
    geo::BoxT<size_t> _subTop(0, 0, 0, _nx, _ny, _nz / 2), _subBottom(0, 0, _nz / 2, _nx, _ny, _nz / 2);
    bool bDualBuffer = false;
    std::shared_ptr<BackProjectionFDK> pBP = std::make_shared<BackProjectionFDK>(
        _projWidth, _projHeight, _projCount, _nx, _ny, _nz, _subTop, _subBottom, bDualBuffer);
    for (int s = 0; s < views; s += cudaBackProjection::MAX_PROJ) {
      const int nNum = std::min<int>(cudaBackProjection::MAX_PROJ, views - s);
      pBP->BP(1, filtered.data() + view * s, para.nv, para.nu,
              &para.vec_rot_mat_ijk3x4[0] + cudaBackProjection::PROJ_MAT_SIZE * s, nNum);
    }
    vol.MallocBuffer(_subTop.depth + _subBottom.depth, _subTop.height, _subTop.width);
    pBP->GetVolumeData(vol.buffer, vol.BufferSize());
    kernel_ms = pBP->fKernelTime;
    if (pBP->nBpCount != views) throw std::string("nBpCount mismatch");
  } else {
    BackProjectionRTK bp((int)para.nu, (int)para.nv, cudaBackProjection::MAX_PROJ, n, n, n);
    for (int s = 0; s < views; s += cudaBackProjection::MAX_PROJ) {
      const int cnt = std::min<int>(cudaBackProjection::MAX_PROJ, views - s);
      bp.BP(0, filtered.data() + view * s, (int)para.nu, (int)para.nv,
            para.vec_rot_mat_ijk3x4.data() + 12 * s, cnt);
    }
    vol.MallocBuffer(n, n, n);
    bp.GetVolumeData(vol.buffer, vol.BufferSize());
    kernel_ms = bp.fKernelTime;
  }
  const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  write_raw(out + ".vol.raw", vol.buffer, vol.BufferSize());
  double sum = 0.0;
  for (float v : vol.data) sum += v;
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
  // genProj/genProj.cpp:60-75: host volume with its spacing, device volume,
  // one device image per view, the host image written out after each view.
  ProjectionImage proj_ori, proj_fp;
  Volume volume(para.nx, para.ny, para.nz);
  std::memcpy(volume.buffer, vol.data(), nvox * 4);
  volume.SetUint(para.dx, para.dy, para.dz);
  std::vector<DevData<float>> vec_dev_forward_proj(para.nProj);
  DevArray3D<float> dev_vol(volume.nx, volume.ny, volume.nz);
  dev_vol.CopyFromHost(volume.buffer, volume.nx, volume.nx, volume.ny, volume.nz);
  ImageT<float> image(para.nu, para.nv);
  DevData<float> dev_image(para.nu, para.nv);
  std::vector<float> views_out;
  for (int i = 0; i < para.nProj; i++) {
    if (!cudaCTForwardProjection_f32(proj_ori, para, volume, &dev_vol, NULL, image, dev_image,
                                     &vec_dev_forward_proj, NULL, i))
      throw std::string("cudaCTForwardProjection_f32 failed");
    views_out.insert(views_out.end(), image.buffer, image.buffer + image.width * image.height);
  }
  // The all-views form into a second stack must agree with the per-view calls.
  proj_fp.MallocBuffer(para.nu, para.nv, para.nProj);
  if (!cudaCTForwardProjection_f32(proj_ori, para, volume, &dev_vol, NULL, image, dev_image,
                                   &vec_dev_forward_proj, &proj_fp, -1))
    throw std::string("cudaCTForwardProjection_f32 (all views) failed");
  if (std::memcmp(proj_fp.Buffer(), views_out.data(), views_out.size() * 4) != 0)
    throw std::string("all-views projection differs from the per-view calls");
  write_raw(out + ".vol.raw", vol.data(), nvox);
  write_raw(out + ".fp.raw", views_out.data(), views_out.size());
  std::printf("fp: %d views of %d^2 through %d^3\n", views, n, n);
  return 0;
}

// FilterlibCufft/FilterlibCufft.cpp:8-20: load the scan's views from
// str_proj_img_folder, filter them in place on the GPU, save them into
// str_working_folder.  The folder is first filled with phantom projections.
int run_filter(int n, int views, const std::string& out) {
  Parameter para = make_param(n, views);
  para.str_proj_img_folder = out + "_raw/";
  para.str_working_folder = out + "_flt/";
  {
    const fdk_geometry g = para.geometry();
    std::vector<fdk_ellipsoid> e = phantom(n * para.dx);
    ProjectionImage raw(n, n, views);
    float* d = nullptr;
    if (hipMalloc(&d, (size_t)n * n * views * 4) != hipSuccess) throw std::string("hipMalloc");
    fdk_amd_detail::check(fdk_phantom_project(e.data(), (int)e.size(), &g, 0, views, d, nullptr),
                          "phantom_project");
    if (hipMemcpy(raw.Buffer(), d, (size_t)n * n * views * 4, hipMemcpyDeviceToHost) != hipSuccess)
      throw std::string("hipMemcpy");
    (void)hipFree(d);
    if (!raw.SaveToFolder(para.str_proj_img_folder)) throw std::string("SaveToFolder");
  }
  const int batchCount = 16;
  const char* pSrc = para.str_proj_img_folder.c_str();
  const char* pDst = para.str_working_folder.c_str();
  ProjectionImageT<float> proj;
  if (!proj.LoadFolder<float>(pSrc, para.nu, para.nv, para.nProj, para.step))
    throw std::string("LoadFolder");
  float tm = FilterProjImage(para, proj, batchCount);
  std::printf("filtering_%dx%d, count=%d, time=%f ms, avg=%f ms, %f fps\n", (int)para.nu,
              (int)para.nv, (int)para.nProj, tm, tm / para.nProj, para.nProj / (tm / 1000.f));
  if (!(tm > 0.f)) throw std::string("FilterProjImage returned no time");
  if (!proj.SaveToFolder(pDst)) throw std::string("SaveToFolder");
  if (para.GetProjPath(1) != para.str_proj_img_folder + "img0001.raw")
    throw std::string("GetProjPath");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) {
    std::fprintf(stderr, "usage: %s fdk|rtk|fp|filter <n> <views> <out_prefix>\n", argv[0]);
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
    if (mode == "filter") return run_filter(n, views, argv[4]);
  } catch (const std::string& e) {
    std::fprintf(stderr, "error: %s\n", e.c_str());
    return 1;
  }
  std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
  return 2;
}
