// fdk_amd C++ drop-in layer: the reference's C++ entry points for the FDK hot
// path, implemented over the C ABI of fdk_amd.h (header-only).
//
//   geo::BoxT, cudaBackProjection, BackProjectionFDK, BackProjectionRTK
//                               cudaBP/backprojection.cuh:49-119, 119-132, 180-198, 246-271
//   DataRef, FilterRow_Real32, CorrectImage_Real32, FilterProjImage
//                               FilterlibCufft/cudaBackProjection.cuh:121-229,
//                               FilterlibCufft/cudaFftFilter.cuh:7-9, cudaFftFilter.cu:394
//   FrameT, ImageT, FilterEngine common/type.h:1893-2005, 2010-2115, Filterlib/FilterLib.h:148-203
//   Parameter                   common/type.h:4627-4895 (fields, init() tables, GetProjPath)
//   DevData, DevArray3D         cudaBP/cudaLib.cuh:298-420, 552-640 (linear device memory)
//   cudaCTForwardProjection_f32 genProj/cudaIterType.cuh:280, cudaForwardProjection.cu:374-489
//   ProjectionImageT, VolumeT   common/type.h:3650-3960 (frames, LoadFolder/SaveToFolder,
//                               Transpose), 2440-2480, 3256-3325 (Save/Load, SetUint)
//   CheckImageQuality, NCC      common/type.h:4991-5030, 4960-4987
//
// Failures throw std::string, as the reference's VERIFY_TRUE / CUDA_CHECK_ERROR
// do (common/type.h:194-196, common/cudaLib.cuh:49-56); the message is the
// library's fdk_last_error().  BP() and the filter calls are synchronous like
// the reference's; BackProjectionFDK owns one stream, so two objects on one
// device may run concurrently (the reference's __constant__ matrices forbid it).
#ifndef FDK_AMD_HPP
#define FDK_AMD_HPP

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
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
  typedef D DataType;
  FrameT() : buffer(nullptr), width(0), height(0), index(-1) {}
  FrameT(D* data, int w, int h, int i) : buffer(data), width(w), height(h), index(i) {}
  D* buffer;
  int width, height, index;
};

// A host image of width x height samples, x fastest (ImageT, type.h:2010-2115).
// Copies are deep; `buffer` points at the image's own storage.
template <typename T>
struct ImageT {
  typedef T DataType;
  ImageT() : width(0), height(0), buffer(nullptr) {}
  ImageT(int w, int h) : ImageT() { MallocBuffer(w, h); }
  explicit ImageT(const FrameT<T>& f) : ImageT() {
    MallocBuffer(f.width, f.height);
    std::copy(f.buffer, f.buffer + mem.size(), buffer);
  }
  template <typename A>
  ImageT(const A* data, int w, int h) : ImageT() {
    MallocBuffer(w, h);
    for (size_t i = 0; i < mem.size(); ++i) buffer[i] = (T)data[i];
  }
  ImageT(const ImageT& o) : ImageT() { *this = o; }
  ImageT& operator=(const ImageT& o) {
    if (this != &o) {
      mem = o.mem;
      width = o.width;
      height = o.height;
      buffer = mem.empty() ? nullptr : mem.data();
    }
    return *this;
  }
  // Same size keeps the samples; another size reallocates zeroed storage.
  ImageT& MallocBuffer(int w, int h) {
    if (w < 0 || h < 0) throw std::string("ImageT::MallocBuffer: negative size");
    if (w != width || h != height || !buffer) {
      mem.assign((size_t)w * h, T());
      width = w;
      height = h;
      buffer = mem.empty() ? nullptr : mem.data();
    }
    return *this;
  }
  FrameT<T> GetFrame() const { return FrameT<T>(buffer, width, height, 0); }
  T& operator()(int x, int y) { return buffer[(size_t)y * width + x]; }
  const T& operator()(int x, int y) const { return buffer[(size_t)y * width + x]; }
  ImageT& Zero() {
    std::fill(mem.begin(), mem.end(), T());
    return *this;
  }
  ImageT& Clear() { return Zero(); }
  int width, height;
  T* buffer;

 private:
  std::vector<T> mem;
};

// Device buffer of width x height x depth samples, rows packed (DataPitch() ==
// width), on the current device (DevData, cudaBP/cudaLib.cuh:298-420).
// Copies between host and device are synchronous; copies of the object are deep.
template <typename T>
struct DevData {
  typedef T DataType;
  DevData() {}
  explicit DevData(size_t w, size_t h = 1, size_t d = 1) { MallocBuffer(w, h, d); }
  DevData(const DevData& o) { *this = o; }
  DevData(DevData&& o) noexcept { swap(o); }
  DevData& operator=(const DevData& o) {
    if (this != &o) {
      MallocBuffer(o.width, o.height, o.depth);
      fdk_amd_detail::check(fdk_memcpy(ptr, o.ptr, bytes()), "DevData copy");
    }
    return *this;
  }
  DevData& operator=(DevData&& o) noexcept {
    swap(o);
    return *this;
  }
  virtual ~DevData() { FreeBuffer(); }
  DevData& MallocBuffer(size_t w, size_t h, size_t d) {
    if (w == width && h == height && d == depth && (ptr || w * h * d == 0)) return *this;
    FreeBuffer();
    void* p = nullptr;
    fdk_amd_detail::check(fdk_dev_malloc(&p, sizeof(T) * w * h * d), "DevData::MallocBuffer");
    ptr = static_cast<T*>(p);
    width = w;
    height = h;
    depth = d;
    channels = d == 1 ? (h == 1 ? 1 : 2) : 3;
    return *this;
  }
  DevData& MallocBuffer(size_t w, size_t h) { return MallocBuffer(w, h, 1); }
  DevData& MallocBuffer(size_t w) { return MallocBuffer(w, 1, 1); }
  void FreeBuffer() {
    if (ptr) fdk_dev_free(ptr);
    ptr = nullptr;
    width = height = depth = channels = 0;
  }
  // Host rows of _w samples, _pitch samples apart, _h * _d rows.
  void CopyFromHost(const T* host, size_t _w, size_t _pitch, size_t _h, size_t _d = 1) {
    copy_rows(ptr, width, host, _pitch, _w, _h * _d, "DevData::CopyFromHost");
  }
  void CopyToHost(T* host, size_t _w, size_t _pitch, size_t _h, size_t _d = 1) const {
    copy_rows(host, _pitch, ptr, width, _w, _h * _d, "DevData::CopyToHost");
  }
  DevData& Zero() {
    fdk_amd_detail::check(fdk_dev_memset(ptr, 0, bytes()), "DevData::Zero");
    return *this;
  }
  size_t DataPitch() const { return width; }
  T* GetData() const { return ptr; }
  size_t width = 0, height = 0, depth = 0, channels = 0;

 protected:
  size_t bytes() const { return sizeof(T) * width * height * depth; }
  void copy_rows(T* dst, size_t dpitch, const T* src, size_t spitch, size_t w, size_t rows,
                 const char* what) const {
    if (w > width || rows > height * depth)
      throw std::string(what) + ": copy larger than the buffer";
    fdk_amd_detail::check(fdk_memcpy_2d(dst, dpitch * sizeof(T), src, spitch * sizeof(T),
                                        w * sizeof(T), rows),
                          what);
  }
  void swap(DevData& o) noexcept {
    std::swap(ptr, o.ptr);
    std::swap(width, o.width);
    std::swap(height, o.height);
    std::swap(depth, o.depth);
    std::swap(channels, o.channels);
  }
  T* ptr = nullptr;
};

// Device volume of width x height x depth samples, x fastest (DevArray3D,
// cudaBP/cudaLib.cuh:552-640).  Linear memory: the kernels gather through
// buffer loads, not texture units, so the cudaArray flags are accepted and
// ignored.
template <typename T>
struct DevArray3D : public DevData<T> {
  DevArray3D(unsigned int flags = 0) { (void)flags; }
  DevArray3D(size_t w, size_t h, size_t d, unsigned int flags = 0) {
    (void)flags;
    this->MallocBuffer(w, h, d);
  }
  DevArray3D& MallocBuffer(size_t w, size_t h, size_t d, unsigned int flags = 0) {
    (void)flags;
    DevData<T>::MallocBuffer(w, h, d);
    return *this;
  }
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
  int compute_device = 1;
  int64_t verbose = 0, nBatchCount = 0;
  std::string str_filter = "shepp-logan";
  std::string str_main_folder, str_proj_img_folder, str_working_folder, str_mask_folder;

  // Derived by init().
  double sx = 0, sy = 0, sz = 0, su = 0, sv = 0;  // extents in mm
  double offset_su = 0, offset_sv = 0;            // detector offsets in mm
  double AngularAndrampFactor = 0.0;
  struct {
    ImageT<float> cos_sigma, csc_sigma;  // nu x nv (u fastest): weight and 1 / weight
  } cos_image_table;
  std::vector<double> fft_kernel_real;      // N
  std::vector<float> fft_kernel_real_f32;   // N
  std::vector<float> vec_rot_mat_ijk3x4;    // nProj x 12
  std::vector<double> vec_deg;

  // Path of view `index` of the scan: <str_proj_img_folder>img%04d.raw numbered
  // index * step (type.h:4852-4856).
  std::string GetProjPath(int index) const {
    if (index < 0 || index >= nProj)
      throw std::string("Parameter::GetProjPath: index outside 0..nProj-1");
    char name[32];
    std::snprintf(name, sizeof(name), "img%04d.raw", index * step);
    return str_proj_img_folder + name;
  }

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
    sx = dx * nx;
    sy = dy * ny;
    sz = dz * nz;
    su = du * nu;
    sv = dv * nv;
    offset_su = offset_nu * du;
    offset_sv = offset_nv * dv;
    fdk_amd_detail::check(fdk_geom_angular_ramp_factor(&g, &AngularAndrampFactor), "Parameter");
    ImageT<float>& cs = cos_image_table.cos_sigma;
    ImageT<float>& csc = cos_image_table.csc_sigma;
    cs.MallocBuffer((int)nu, (int)nv);
    csc.MallocBuffer((int)nu, (int)nv);
    fdk_amd_detail::check(fdk_geom_cos_table(&g, cs.buffer), "CalculateCosTable");
    for (size_t i = 0; i < (size_t)nu * nv; ++i) csc.buffer[i] = (float)(1.0 / cs.buffer[i]);
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

// A stack of projection images in host memory: count frames of width x height
// (ProjectionImageT, type.h:3650-3960).  As in the reference, frames of a const
// stack are writable (FilterProjImage filters a const reference in place).
template <typename T>
struct ProjectionImageT {
  typedef T DataType;
  typedef FrameT<T> FrameType;
  ProjectionImageT() : width(0), height(0), count(0) {}
  ProjectionImageT(int w, int h, int c) : width(w), height(h), count(c), data((size_t)w * h * c) {}
  int GetCount() const { return count; }
  int GetStep() const { return step; }
  ProjectionImageT& SetStep(int s) {
    step = s;
    return *this;
  }
  // Same shape keeps the samples; another shape reallocates zeroed frames.
  ProjectionImageT& MallocBuffer(int64_t w, int64_t h, int64_t c, int _step = 1) {
    if (w < 0 || h < 0 || c < 0) throw std::string("ProjectionImageT::MallocBuffer: negative size");
    if (w != width || h != height || c != count) *this = ProjectionImageT((int)w, (int)h, (int)c);
    step = _step;
    return *this;
  }
  bool IsValid() const { return !data.empty() && count > 0; }
  T* Buffer() const { return const_cast<T*>(data.data()); }
  int64_t BufferSize() const { return (int64_t)width * height * count; }
  bool GetFrame(FrameT<T>& f, int i) const {
    if (i < 0 || i >= count) return false;
    f = FrameT<T>(Buffer() + (size_t)width * height * i, width, height, i);
    return true;
  }
  FrameT<T> GetFrame(int i) const {
    FrameT<T> f;
    if (!GetFrame(f, i)) throw std::string("ProjectionImageT::GetFrame: index outside the stack");
    return f;
  }
  bool SetFrame(const T* src, int w, int h, int i) {
    if (w != width || h != height || i < 0 || i >= count) return false;
    std::copy(src, src + (size_t)w * h, data.begin() + (size_t)w * h * i);
    return true;
  }
  // Every frame transposed in place; width and height swap (type.h:3936-3958).
  ProjectionImageT& Transpose() {
    std::vector<T> tmp((size_t)width * height);
    for (int i = 0; i < count; ++i) {
      T* f = data.data() + (size_t)width * height * i;
      for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) tmp[(size_t)height * x + y] = f[(size_t)y * width + x];
      std::copy(tmp.begin(), tmp.end(), f);
    }
    std::swap(width, height);
    return *this;
  }

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
    this->step = step;
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
  int step = 1;
  std::vector<T> data;
};
typedef ProjectionImageT<float> ProjectionImage;

// A host volume nz x ny x nx (x fastest) with the reference's file format
// (VolumeT::Save / Load, type.h:3256-3310): three int32 (nx, ny, nz), then
// the samples slice by slice.
template <typename T>
struct VolumeT {
  typedef T DataType;
  VolumeT() : nx(0), ny(0), nz(0), buffer(nullptr) {}
  VolumeT(int64_t x, int64_t y, int64_t z) : VolumeT() { MallocBuffer(x, y, z); }
  VolumeT(const VolumeT& o) : VolumeT() { *this = o; }
  VolumeT& operator=(const VolumeT& o) {
    if (this != &o) {
      data = o.data;
      nx = o.nx;
      ny = o.ny;
      nz = o.nz;
      dx = o.dx;
      dy = o.dy;
      dz = o.dz;
      buffer = data.empty() ? nullptr : data.data();
    }
    return *this;
  }
  // Same size keeps the samples; another size reallocates zeroed storage.
  VolumeT& MallocBuffer(int64_t x, int64_t y, int64_t z) {
    if (x < 0 || y < 0 || z < 0) throw std::string("VolumeT::MallocBuffer: negative size");
    if (x != nx || y != ny || z != nz || !buffer) {
      data.assign((size_t)(x * y * z), T());
      nx = x;
      ny = y;
      nz = z;
      buffer = data.empty() ? nullptr : data.data();
    }
    return *this;
  }
  int64_t BufferSize() const { return nx * ny * nz; }
  // Voxel spacing in mm (VolumeT::SetUint, type.h:3321-3325).
  void SetUint(float _dx, float _dy, float _dz) {
    dx = _dx;
    dy = _dy;
    dz = _dz;
  }
  bool Save(const std::string& path) const {
    fdk_amd_detail::make_dirs(path);
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const int32_t hdr[3] = {(int32_t)nx, (int32_t)ny, (int32_t)nz};
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
      MallocBuffer(hdr[0], hdr[1], hdr[2]);
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
  int64_t nx, ny, nz;
  float dx = 0.f, dy = 0.f, dz = 0.f;
  T* buffer;  // the samples, nz x ny x nx with x fastest
  std::vector<T> data;
};
typedef VolumeT<float> Volume;

// FilterProjImage (FilterlibCufft/cudaFftFilter.cu:394-452): views filtered in
// place, DEPTH views per device batch, with the cuFFT path's 3-sample pad;
// returns the summed milliseconds of the batches (upload, cos weighting and
// filter, download), as the reference does.  The reference takes the stack by
// const reference and writes its frames; so does this.
inline float FilterProjImage(const Parameter& para, const ProjectionImage& proj_img,
                             int DEPTH = 32) {
  if (proj_img.width != para.nu || proj_img.height != para.nv)
    throw std::string("FilterProjImage: image size differs from the parameter");
  float ms = 0.f;
  fdk_amd_detail::check(
      fdk_filter_proj_image(proj_img.Buffer(), (int)para.nu, (int)para.nv, proj_img.GetCount(),
                            para.cos_image_table.cos_sigma.buffer, para.fft_kernel_real_f32.data(),
                            (int)para.fft_kernel_real_f32.size(), FDK_PAD_MEAN3, DEPTH, &ms),
      "FilterProjImage");
  return ms;
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
                                            FDK_FP_CLAMP, out.Buffer()),
                        "cudaCTForwardProjection_f32");
  return true;
}

// cudaCTForwardProjection_f32 with the reference's arguments
// (genProj/cudaIterType.cuh:280, cudaForwardProjection.cu:374-489).
//  * The sampled volume is the device volume pDevVolRoi (the reference binds it
//    to the texture its projector reads); when it is NULL the host `volume` is
//    uploaded.  Sizes and spacing come from `volume` (nx, ny, nz, dx, dy, dz:
//    VolumeT::SetUint), as the reference's volume geometry does.
//  * pDevVolMask is accepted and not read: the reference binds it, and its
//    single-volume projector (kernel_forwardProject) never samples it.
//  * proj_index >= 0: that view is projected into dev_forward_proj (allocated
//    nu x nv by the caller) and copied into forward_proj.  proj_index < 0: views
//    0, step, 2 step, ... are projected into (*pvec_dev_forward_proj)[i] when it
//    is given, and into forward_proj and frame i of *pforward_proj_2th when that
//    is given.  forward_proj ends holding the last view projected.
//  * GPU sampling semantics of the reference projector: step min(dx, dy, dz),
//    t < tfar, clamp-to-edge trilinear.  `proj` only fed the reference's
//    on-screen comparison and is not read.
//  * Errors are printed and return false, as the reference's try/catch does.
inline bool cudaCTForwardProjection_f32(const ProjectionImageT<float>& proj, const Parameter& para,
                                        const VolumeT<float>& volume,
                                        DevArray3D<float>* pDevVolRoi,
                                        DevArray3D<float>* pDevVolMask,
                                        ImageT<float>& forward_proj,
                                        DevData<float>& dev_forward_proj,
                                        std::vector<DevData<float>>* pvec_dev_forward_proj,
                                        ProjectionImageT<float>* pforward_proj_2th = nullptr,
                                        int proj_index = -1) {
  (void)proj;
  (void)pDevVolMask;
  try {
    if (!(volume.dx > 0 && volume.dy > 0 && volume.dz > 0))
      throw std::string("volume spacing not set (VolumeT::SetUint)");
    if (forward_proj.width <= 0 || forward_proj.height <= 0)
      throw std::string("forward_proj has no size");
    fdk_geometry g = para.geometry();
    g.nx = (int)volume.nx;
    g.ny = (int)volume.ny;
    g.nz = (int)volume.nz;
    g.dx = volume.dx;
    g.dy = volume.dy;
    g.dz = volume.dz;
    DevArray3D<float> uploaded;
    const DevData<float>* vol = pDevVolRoi;
    if (!vol) {
      uploaded.MallocBuffer(volume.nx, volume.ny, volume.nz);
      uploaded.CopyFromHost(volume.buffer, volume.nx, volume.nx, volume.ny, volume.nz);
      vol = &uploaded;
    }
    if ((int64_t)vol->width != volume.nx || (int64_t)vol->height != volume.ny ||
        (int64_t)vol->depth != volume.nz)
      throw std::string("device volume size differs from the volume");
    auto project = [&](int i, DevData<float>& dst) {
      if ((int64_t)(dst.width * dst.height * dst.depth) < para.nu * para.nv)
        throw std::string("device projection buffer smaller than nu x nv");
      fdk_amd_detail::check(fdk_fp_project(&g, vol->GetData(), i, 1, 0.0, FDK_FP_CLAMP,
                                           dst.GetData(), nullptr),
                            "fdk_fp_project");
      forward_proj.MallocBuffer((int)para.nu, (int)para.nv);
      dst.CopyToHost(forward_proj.buffer, para.nu, para.nu, para.nv);
    };
    if (proj_index >= 0) {
      if (proj_index < para.nProj) project(proj_index, dev_forward_proj);  // else no view, as the reference
      return true;
    }
    if (pvec_dev_forward_proj && (int64_t)pvec_dev_forward_proj->size() < para.nProj)
      throw std::string("fewer device projections than views");
    for (int i = 0; i < para.nProj; i += para.step) {
      bool done = false;
      if (pvec_dev_forward_proj) {
        DevData<float>& d = (*pvec_dev_forward_proj)[i];
        if (!d.GetData()) d.MallocBuffer(para.nu, para.nv);
        project(i, d);
        done = true;
      }
      if (pforward_proj_2th) {
        if (!done) {
          if (!dev_forward_proj.GetData()) dev_forward_proj.MallocBuffer(para.nu, para.nv);
          project(i, dev_forward_proj);
        }
        if (!pforward_proj_2th->SetFrame(forward_proj.buffer, forward_proj.width,
                                         forward_proj.height, i))
          throw std::string("pforward_proj_2th frame size");
      }
    }
    return true;
  } catch (const std::string& e) {
    std::printf("error : cudaCTForwardProjection_f32: %s\n", e.c_str());
    return false;
  }
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
