// fdk_amd C++ drop-in layer: the reference's C++ entry points for the FDK hot
// path, implemented over the C ABI of fdk_amd.h (header-only).
//
//   geo::BoxT, cudaBackProjection, BackProjectionFDK, BackProjectionRTK
//                               cudaBP/backprojection.cuh:49-119, 119-132, 180-198, 246-271
//   DataRef, FilterRow_Real32, CorrectImage_Real32, FilterProjImage
//                               FilterlibCufft/cudaBackProjection.cuh:121-229,
//                               FilterlibCufft/cudaFftFilter.cuh:7-9, cudaFftFilter.cu:394
//   FrameT, FilterEngine        common/type.h:1893-2005, Filterlib/FilterLib.h:148-203
//   Parameter                   common/type.h:4627-4895 (fields and the tables of init())
//   cudaCTForwardProjection_f32 genProj/cudaIterType.cuh:280 (core arguments only)
//   ProjectionImageT::LoadFolder/SaveToFolder, VolumeT::Save/Load
//                               common/type.h:1796-1828, 3831-3858, 3256-3310
//   CheckImageQuality, NCC      common/type.h:4991-5030, 4960-4987
//
// Failures throw std::string, as the reference's VERIFY_TRUE / CUDA_CHECK_ERROR
// do (common/type.h:194-196, common/cudaLib.cuh:49-56); the message is the
// library's fdk_last_error().  BP() and the filter calls are synchronous like
// the reference's; BackProjectionFDK owns one stream, so two objects on one
// device may run concurrently (the reference's __constant__ matrices forbid it).
#ifndef FDK_AMD_HPP
#define FDK_AMD_HPP

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "fdk_amd.h"

namespace fdk_amd_detail {
inline std::string last_error() {
  char buf[1024];
  fdk_last_error(buf, sizeof(buf));
  return std::string(buf);
}
inline void check(int rc, const char* what) {
  if (rc < 0) throw std::string(what) + ": " + last_error();
}
// Creates every directory on the way to the last '/' of path (the reference's
// MakePathDirIfNeed); existing directories are fine.
inline void make_dirs(const std::string& path) {
  for (size_t i = 1; i < path.size(); ++i)
    if (path[i] == '/') ::mkdir(path.substr(0, i).c_str(), 0755);
}
}  // namespace fdk_amd_detail

namespace geo {
template <typename T>
struct BoxT {
  BoxT() : x(0), y(0), z(0), width(0), height(0), depth(0) {}
  BoxT(T x_, T y_, T z_, T w, T h, T d) : x(x_), y(y_), z(z_), width(w), height(h), depth(d) {}
  bool IsValid() const { return width > 0 && height > 0 && depth > 0; }
  T MinX() const { return x; }
  T MinY() const { return y; }
  T MinZ() const { return z; }
  T MaxX() const { return x + width - 1; }
  T MaxY() const { return y + height - 1; }
  T MaxZ() const { return z + depth - 1; }
  T x, y, z, width, height, depth;
};
}  // namespace geo

struct cudaBackProjection {
  enum { MAX_PROJ = FDK_BP_MAX_PROJ, PROJ_MAT_ROWS = 3, PROJ_MAT_COLS = 4, PROJ_MAT_SIZE = 12 };
  float fKernelTime = 0.0f;
  bool bDualBuffer = false;
};

struct BackProjectionFDK : public cudaBackProjection {
  BackProjectionFDK(int projWidth_, int projHeight_, int projCount_, int nx_, int ny_, int nz_,
                    geo::BoxT<size_t> subTop_, geo::BoxT<size_t> subBottom_,
                    bool bDualBuffer_ = false)
      : projWidth(projWidth_), projHeight(projHeight_), projCount(projCount_), nx(nx_), ny(ny_),
        nz(nz_), subTop(subTop_), subBottom(subBottom_), nBpCount(0) {
    if (bDualBuffer_) throw std::string("BackProjectionFDK: dual buffer is not supported");
    const int64_t t[6] = {(int64_t)subTop.x, (int64_t)subTop.y, (int64_t)subTop.z,
                          (int64_t)subTop.width, (int64_t)subTop.height, (int64_t)subTop.depth};
    const int64_t b[6] = {(int64_t)subBottom.x, (int64_t)subBottom.y, (int64_t)subBottom.z,
                          (int64_t)subBottom.width, (int64_t)subBottom.height,
                          (int64_t)subBottom.depth};
    fdk_amd_detail::check(fdk_bp_create(projWidth_, projHeight_, projCount_, nx_, ny_, nz_, t, b, &h),
                          "BackProjectionFDK");
  }
  ~BackProjectionFDK() { fdk_bp_destroy(h); }
  BackProjectionFDK(const BackProjectionFDK&) = delete;
  BackProjectionFDK& operator=(const BackProjectionFDK&) = delete;

  bool BP(int TRANSPOSE, const float* pProjData, int width, int height, float* pProjMat,
          int nProj) {
    fdk_amd_detail::check(fdk_bp_batch(h, TRANSPOSE, pProjData, width, height, pProjMat, nProj),
                          "BackProjectionFDK::BP");
    nBpCount = fdk_bp_count(h);
    fKernelTime = fdk_bp_kernel_ms(h);
    return true;
  }
  // Device-resident variant (no reference counterpart): asynchronous.
  bool BPDevice(const void* devProj, int dtype, int width, int height, int64_t viewStride,
                const float* pProjMat, const float* scaleOffset, int nProj) {
    fdk_amd_detail::check(fdk_bp_batch_device(h, 1, devProj, dtype, width, height, viewStride,
                                              pProjMat, scaleOffset, nProj),
                          "BackProjectionFDK::BPDevice");
    nBpCount = fdk_bp_count(h);
    return true;
  }
  bool GetVolumeData(float* pVol, size_t size) {
    fdk_amd_detail::check(fdk_bp_get_volume(h, pVol, size), "BackProjectionFDK::GetVolumeData");
    return true;
  }
  fdk_bp_t handle() const { return h; }

  size_t projWidth, projHeight, projCount;
  size_t nx, ny, nz;
  geo::BoxT<size_t> subTop, subBottom;
  int nBpCount;

 private:
  fdk_bp_t h = nullptr;
};

struct BackProjectionRTK : public cudaBackProjection {
  BackProjectionRTK(int projWidth_, int projHeight_, int projCount_, int nx_, int ny_, int nz_,
                    bool bDualBuffer_ = false)
      : projWidth(projWidth_), projHeight(projHeight_), projCount(projCount_), nx(nx_), ny(ny_),
        nz(nz_) {
    bDualBuffer = bDualBuffer_;
    fdk_amd_detail::check(fdk_rtk_create(projWidth_, projHeight_, projCount_, nx_, ny_, nz_, &h),
                          "BackProjectionRTK");
  }
  ~BackProjectionRTK() { fdk_rtk_destroy(h); }
  BackProjectionRTK(const BackProjectionRTK&) = delete;
  BackProjectionRTK& operator=(const BackProjectionRTK&) = delete;
  bool BP(int TRANSPOSE, float* pProjData, int width, int height, float* pProjMat, int nProj) {
    fdk_amd_detail::check(fdk_rtk_bp(h, TRANSPOSE, pProjData, width, height, pProjMat, nProj),
                          "BackProjectionRTK::BP");
    fKernelTime = fdk_rtk_kernel_ms(h);
    return true;
  }
  bool GetVolumeData(float* pVol, size_t size) {
    fdk_amd_detail::check(fdk_rtk_get_volume(h, pVol, size), "BackProjectionRTK::GetVolumeData");
    return true;
  }
  // CUDA_reconstruct_conebeam: proj is host memory, as in the reference.
  void CUDA_reconstruct_conebeam(int proj_size[3], int vol_size[3], float* matrices,
                                 float* dev_vol_in, float* dev_vol_out, float* dev_proj) {
    fdk_amd_detail::check(fdk_reconstruct_conebeam(proj_size, vol_size, matrices, dev_vol_in,
                                                   dev_vol_out, dev_proj, 0, nullptr),
                          "CUDA_reconstruct_conebeam");
  }
  int projWidth, projHeight, projCount;
  int nx, ny, nz;

 private:
  fdk_rtk_t h = nullptr;
};

// ------------------------------------------------------------------- filtering
template <typename T>
struct DataRef {
  typedef T DataType;
  DataRef() : width(0), height(0), depth(0), pitch(0), channels(0), data(nullptr) {}
  DataRef(T* d, int w, int h, int dp, int p, int c)
      : width(w), height(h), depth(dp), pitch(p), channels(c), data(d) {}
  bool Is3D() const { return channels == 3; }
  bool Is2D() const { return channels == 2 || (Is3D() && depth == 1); }
  bool Is1D() const { return channels == 1; }
  int width, height, depth, pitch, channels;
  T* data;
};

// Device rows (src.depth images of src.height rows, pitch in elements) filtered
// by the device filter of length filter.width; the cuFFT path's 3-sample pad.
inline bool FilterRow_Real32(const DataRef<float>& src, DataRef<float>& dst,
                             const DataRef<float>& filter) {
  if (src.width != dst.width || src.height != dst.height || src.depth > dst.depth)
    throw std::string("FilterRow_Real32: src and dst sizes differ");
  fdk_amd_detail::check(fdk_filter_rows(src.data, dst.data, src.width, src.height,
                                        src.depth > 0 ? src.depth : 1, src.pitch, dst.pitch,
                                        filter.data, filter.width, FDK_PAD_MEAN3, nullptr),
                        "FilterRow_Real32");
  return true;
}

inline bool CorrectImage_Real32(const DataRef<float>& src, const DataRef<float>& table,
                                DataRef<float>& dst) {
  if (src.pitch != dst.pitch || src.width != dst.width || src.height != dst.height)
    throw std::string("CorrectImage_Real32: src and dst layouts differ");
  fdk_amd_detail::check(fdk_correct_image(src.data, table.data, dst.data, src.width, src.height,
                                          src.depth > 0 ? src.depth : 1, src.pitch, table.pitch,
                                          nullptr),
                        "CorrectImage_Real32");
  return true;
}

template <typename D>
struct FrameT {
  FrameT() : buffer(nullptr), width(0), height(0), index(-1) {}
  FrameT(D* data, int w, int h, int i) : buffer(data), width(w), height(h), index(i) {}
  D* buffer;
  int width, height, index;
};

// Scan geometry and the tables Parameter::init builds (type.h:4719-4754).
struct Parameter {
  Parameter() { init(); }
  int64_t nx = 512, ny = 512, nz = 512;
  double dx = 0.157, dy = 0.157, dz = 0.146;
  int64_t nu = 608, nv = 616;
  double du = 0.2, dv = 0.2;
  double offset_nu = 0.0, offset_nv = 0.0;
  double SOD = 468.0;  // source to object (reference naming)
  double SID = 700.0;  // source to detector (reference naming)
  int64_t direction = -1;
  double scan_angle = 360.0;
  int64_t nProj = 512;
  double volume_postion_angle = -3.141592653589793238462 * 48.8671875 / 180.0;
  int step = 1;
  std::string str_filter = "shepp-logan";

  double AngularAndrampFactor = 0.0;
  std::vector<float> cos_sigma;             // nv x nu, u fastest
  std::vector<double> fft_kernel_real;      // N
  std::vector<float> fft_kernel_real_f32;   // N
  std::vector<float> vec_rot_mat_ijk3x4;    // nProj x 12
  std::vector<double> vec_deg;

  fdk_geometry geometry() const {
    static const char* names[6] = {"none", "ram-lak", "shepp-logan", "cosine", "hamming", "hann"};
    int ft = -1;
    for (int i = 0; i < 6; ++i)
      if (str_filter == names[i]) ft = i;
    if (ft < 0) throw std::string("Parameter: unknown filter ") + str_filter;
    fdk_geometry g;
    g.nx = (int)nx; g.ny = (int)ny; g.nz = (int)nz;
    g.dx = dx; g.dy = dy; g.dz = dz;
    g.nu = (int)nu; g.nv = (int)nv;
    g.du = du; g.dv = dv;
    g.offset_nu = offset_nu; g.offset_nv = offset_nv;
    g.sod = SOD; g.sid = SID;
    g.direction = (int)direction;
    g.scan_angle_deg = scan_angle;
    g.nproj = (int)nProj;
    g.volume_position_angle_rad = volume_postion_angle;
    g.step = step;
    g.filter_type = ft;
    return g;
  }

  void init() {
    const fdk_geometry g = geometry();
    fdk_amd_detail::check(fdk_geom_angular_ramp_factor(&g, &AngularAndrampFactor), "Parameter");
    cos_sigma.resize((size_t)nu * nv);
    fdk_amd_detail::check(fdk_geom_cos_table(&g, cos_sigma.data()), "CalculateCosTable");
    const int n = fdk_geom_filter_length(&g);
    fdk_amd_detail::check(n, "filter length");
    fft_kernel_real.resize(n);
    fdk_amd_detail::check(fdk_geom_filter(&g, fft_kernel_real.data(), n), "filter");
    fft_kernel_real_f32.assign(fft_kernel_real.begin(), fft_kernel_real.end());
    std::vector<double> m((size_t)nProj * 12);
    fdk_amd_detail::check(fdk_geom_proj_matrices(&g, 0, (int)nProj, m.data()), "CalculateProjMat");
    vec_rot_mat_ijk3x4.assign(m.begin(), m.end());
    vec_deg.resize(nProj);
    for (int64_t i = 0; i < nProj; ++i) vec_deg[i] = i * scan_angle / double(nProj);
  }
};

// A stack of projection images in host memory: count frames of width x height.
template <typename T>
struct ProjectionImageT {
  ProjectionImageT() : width(0), height(0), count(0) {}
  ProjectionImageT(int w, int h, int c) : width(w), height(h), count(c), data((size_t)w * h * c) {}
  int GetCount() const { return count; }
  bool GetFrame(FrameT<T>& f, int i) {
    if (i < 0 || i >= count) return false;
    f = FrameT<T>(data.data() + (size_t)width * height * i, width, height, i);
    return true;
  }
  T* buffer() { return data.data(); }

  // LoadFolder (type.h:3831-3844 with LoadProjectionImageFolder, 1796-1828):
  // frame k*step <- <folder>img%04d.raw numbered k*step + start_index, each
  // file exactly w*h samples of D, u fastest; the other frames stay zero.
  // Every file's size is checked before any is read.
  template <typename D = T>
  bool LoadFolder(const std::string& folder, int w, int h, int c, int step = 1,
                  int start_index = 0) {
    if (w <= 0 || h <= 0 || c <= 0 || step <= 0) return false;
    *this = ProjectionImageT(w, h, c);
    const size_t n = (size_t)w * h;
    std::vector<std::string> paths;
    for (int i = 0; i < c; i += step) paths.push_back(frame_path(folder, i + start_index));
    for (const std::string& p : paths) {
      std::FILE* f = std::fopen(p.c_str(), "rb");
      if (!f) return false;
      std::fseek(f, 0, SEEK_END);
      const long size = std::ftell(f);
      std::fclose(f);
      if (size < 0 || (size_t)size != n * sizeof(D)) return false;
    }
    std::vector<D> tmp(n);
    for (size_t k = 0; k < paths.size(); ++k) {
      std::FILE* f = std::fopen(paths[k].c_str(), "rb");
      if (!f) return false;
      const size_t got = std::fread(tmp.data(), sizeof(D), n, f);
      std::fclose(f);
      if (got != n) return false;
      T* dst = data.data() + n * k * step;
      for (size_t e = 0; e < n; ++e) dst[e] = (T)tmp[e];
    }
    return true;
  }
  // SaveToFolder (type.h:3845-3858): frames 0, step, 2*step, ... as
  // <folder>img%04d.raw, raw samples of T, u fastest.
  bool SaveToFolder(const std::string& folder, int step = 1) const {
    fdk_amd_detail::make_dirs(folder);
    const size_t n = (size_t)width * height;
    for (int i = 0; i < count; i += step) {
      std::FILE* f = std::fopen(frame_path(folder, i).c_str(), "wb");
      if (!f) return false;
      const size_t put = std::fwrite(data.data() + n * i, sizeof(T), n, f);
      std::fclose(f);
      if (put != n) return false;
    }
    return true;
  }
  static std::string frame_path(const std::string& folder, int index) {
    char name[32];
    std::snprintf(name, sizeof(name), "img%04d.raw", index);
    return folder + name;
  }

  int width, height, count;
  std::vector<T> data;
};
typedef ProjectionImageT<float> ProjectionImage;

// A host volume nz x ny x nx (x fastest) with the reference's file format
// (VolumeT::Save / Load, type.h:3256-3310): three int32 (nx, ny, nz), then
// the samples slice by slice.
template <typename T>
struct VolumeT {
  VolumeT() : nx(0), ny(0), nz(0) {}
  VolumeT(int x, int y, int z) : nx(x), ny(y), nz(z), data((size_t)x * y * z) {}
  T* buffer() { return data.data(); }
  const T* buffer() const { return data.data(); }
  bool Save(const std::string& path) const {
    fdk_amd_detail::make_dirs(path);
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const int32_t hdr[3] = {nx, ny, nz};
    bool ok = std::fwrite(hdr, sizeof(int32_t), 3, f) == 3;
    ok = ok && std::fwrite(data.data(), sizeof(T), data.size(), f) == data.size();
    std::fclose(f);
    return ok;
  }
  bool Load(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    int32_t hdr[3];
    bool ok = std::fread(hdr, sizeof(int32_t), 3, f) == 3 && hdr[0] > 0 && hdr[1] > 0 &&
              hdr[2] > 0;
    if (ok) {
      *this = VolumeT(hdr[0], hdr[1], hdr[2]);
      ok = std::fread(data.data(), sizeof(T), data.size(), f) == data.size();
    }
    std::fclose(f);
    return ok;
  }
  // One raw nx*ny file per slice, <folder>%06d.raw numbered z0 + k (the
  // per-slice output of cuFDK/mpiFDK.cpp:548-575 without its MPI-grid suffix).
  bool SaveSlices(const std::string& folder, int z0 = 0) const {
    fdk_amd_detail::make_dirs(folder);
    const size_t n = (size_t)nx * ny;
    for (int k = 0; k < nz; ++k) {
      char name[32];
      std::snprintf(name, sizeof(name), "%06d.raw", z0 + k);
      std::FILE* f = std::fopen((folder + name).c_str(), "wb");
      if (!f) return false;
      const size_t put = std::fwrite(data.data() + n * k, sizeof(T), n, f);
      std::fclose(f);
      if (put != n) return false;
    }
    return true;
  }
  int32_t nx, ny, nz;
  std::vector<T> data;
};

// FilterProjImage: views filtered in place, DEPTH views per device batch;
// returns the elapsed milliseconds as the reference does.
inline float FilterProjImage(const Parameter& para, ProjectionImage& proj_img, int DEPTH = 32) {
  if (proj_img.width != para.nu || proj_img.height != para.nv)
    throw std::string("FilterProjImage: image size differs from the parameter");
  fdk_amd_detail::check(fdk_filter_proj_image(proj_img.buffer(), (int)para.nu, (int)para.nv,
                                              proj_img.GetCount(), para.cos_sigma.data(),
                                              para.fft_kernel_real_f32.data(),
                                              (int)para.fft_kernel_real_f32.size(), FDK_PAD_MEAN3,
                                              DEPTH),
                        "FilterProjImage");
  return 0.0f;
}

// FilterEngine (Filterlib/FilterLib.h:148-203): cos weighting (src is
// overwritten with the weighted image), mean-4 padded row filter, copy or
// transpose into dst; runs on the GPU one image at a time.
struct FilterEngine {
  FilterEngine(FrameT<float> cos_image_table, const std::vector<float>& fft_kernel_real_f32)
      : cos(cos_image_table), filt(fft_kernel_real_f32) {}
  FilterEngine& Filtering(FrameT<float>& src, FrameT<float>& dst, bool transpose = false) {
    if (src.width != cos.width || src.height != cos.height)
      throw std::string("FilterEngine::Filtering: image size differs from the cos table");
    const bool ok = transpose ? (dst.height == cos.width && dst.width == cos.height)
                              : (dst.width == cos.width && dst.height == cos.height);
    if (!ok) throw std::string("FilterEngine::Filtering: destination size");
    fdk_amd_detail::check(
        fdk_filter_frame_host(src.buffer, dst.buffer, src.width, src.height, cos.buffer,
                              filt.data(), (int)filt.size(), FDK_PAD_MEAN4, transpose ? 1 : 0),
        "FilterEngine::Filtering");
    return *this;
  }
  FrameT<float> cos;
  std::vector<float> filt;
};

// cudaCTForwardProjection_f32, core arguments: ray-march view proj_index (all
// views when -1) of para through the host volume (nz x ny x nx) into out.
inline bool cudaCTForwardProjection_f32(const Parameter& para, const std::vector<float>& volume,
                                        ProjectionImage& out, int proj_index = -1,
                                        double step_mm = 0.0) {
  const fdk_geometry g = para.geometry();
  const int first = proj_index >= 0 ? proj_index : 0;
  const int count = proj_index >= 0 ? 1 : (int)para.nProj;
  if (out.width != para.nu || out.height != para.nv || out.count < count)
    throw std::string("cudaCTForwardProjection_f32: output size");
  fdk_amd_detail::check(fdk_fp_project_host(&g, volume.data(), first, count, step_mm,
                                            FDK_FP_CLAMP, out.buffer()),
                        "cudaCTForwardProjection_f32");
  return true;
}

// CheckImageQuality (common/type.h:4991-5030): error per voxel, MSE, PSNR
// against a peak of 2.0, QI, and the pass thresholds (PSNR >= snrThreshold,
// error per voxel <= errorPerPixelThreshold).  Prints the reference's report
// to stdout and returns the figures; refLowerThreshold / refUpperThreshold are
// accepted and unused, as in the reference.
struct ImageQuality {
  double error_per_pixel, mse, psnr, qi;
  bool passed;
};

template <typename R, typename X>
inline ImageQuality CheckImageQuality(const VolumeT<R>& ref, const VolumeT<X>& recon,
                                      double refLowerThreshold = 0.0,
                                      double refUpperThreshold = 0.0, double snrThreshold = 26.0,
                                      double errorPerPixelThreshold = 0.03) {
  (void)refLowerThreshold;
  (void)refUpperThreshold;
  if (ref.nx != recon.nx || ref.ny != recon.ny || ref.nz != recon.nz || ref.data.empty())
    throw std::string("CheckImageQuality: volume sizes differ");
  double abs_sum = 0.0, sq_sum = 0.0;
  for (size_t i = 0; i < ref.data.size(); ++i) {
    const double d = (double)recon.data[i] - (double)ref.data[i];
    abs_sum += std::fabs(d);
    sq_sum += d * d;
  }
  ImageQuality q;
  q.error_per_pixel = abs_sum / ref.data.size();
  q.mse = sq_sum / ref.data.size();
  q.psnr = 20.0 * std::log10(2.0) - 10.0 * std::log10(q.mse);
  q.qi = (2.0 - q.error_per_pixel) / 2.0;
  q.passed = q.error_per_pixel <= errorPerPixelThreshold && q.psnr >= snrThreshold;
  std::printf("\nError per Pixel = %g\nMSE = %g\nPSNR = %gdB\nQI = %g\n", q.error_per_pixel, q.mse,
              q.psnr, q.qi);
  return q;
}

// NCC (common/type.h:4960-4987): normalised cross-correlation about the means.
template <typename A, typename B>
inline double NCC(const A* a, const B* b, size_t n) {
  if (n == 0) return 0.0;
  double m0 = 0.0, m1 = 0.0;
  for (size_t i = 0; i < n; ++i) {
    m0 += a[i];
    m1 += b[i];
  }
  m0 /= n;
  m1 /= n;
  double x = 0.0, y0 = 0.0, y1 = 0.0;
  for (size_t i = 0; i < n; ++i) {
    const double s0 = a[i] - m0, s1 = b[i] - m1;
    x += s0 * s1;
    y0 += s0 * s0;
    y1 += s1 * s1;
  }
  return x / (std::sqrt(y0) * std::sqrt(y1));
}

#endif  // FDK_AMD_HPP
