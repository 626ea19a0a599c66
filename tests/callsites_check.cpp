// Compile-only check that the reference's callers of the FDK path build against
// include/fdk_amd.hpp unchanged (modulo the include).  Each function below holds
// the expressions of one call site, in the caller's own form:
//   cuFDK/mpiFDK.cpp:104-127      FilterThread: ImageT, FrameT, GetProjPath,
//                                 FilterEngine from para.cos_image_table.cos_sigma
//   cuFDK/mpiFDK.cpp:167-203, 245, 272-302
//                                 PushBpData / BpThread: BackProjectionFDK from
//                                 size_t sizes, BP(1, ...), nBpCount, GetVolumeData
//   cudaBP/cudaBP.cpp:85-153      TestFDK: ProjectionImageT::MallocBuffer /
//                                 LoadFolder / Transpose / GetFrame(i).buffer,
//                                 fKernelTime, VolumeT::MallocBuffer / BufferSize
//   FilterlibCufft/FilterlibCufft.cpp:8-20
//                                 test_filter: block-scope extern FilterProjImage
//                                 returning milliseconds, str_*_folder
//   genProj/genProj.cpp:31-81     ForwardProjection: Volume, SetUint, DevArray3D,
//                                 DevData vector, ImageT, the 10-argument
//                                 cudaCTForwardProjection_f32
// The callers' general utilities from common/type.h (VERIFY_TRUE, NoConstant,
// StringFormat) are not part of the path; small local forms stand in for them.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "fdk_amd.hpp"

#define VERIFY_TRUE(x) \
  do {                 \
    if (!(x)) throw std::string(#x); \
  } while (0)
template <typename T>
T* NoConstant(const T* p) {
  return const_cast<T*>(p);
}
static std::string StringFormat(const char* fmt, int a, int b, int c) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), fmt, a, b, c);
  return buf;
}

// cuFDK/mpiFDK.cpp:104-127
void filter_thread_view(const Parameter& para, std::vector<std::shared_ptr<FilterEngine>>& vecFlt,
                        float* header_data, int i, int tid) {
  ImageT<float> img(para.nu, para.nv);
  char szPath[1024] = "";
  std::strcpy(szPath, para.GetProjPath(i).c_str());
  (void)szPath;
  FrameT<float> src((float*)img.buffer, para.nu, para.nv, 0);
  FrameT<float> dst((float*)header_data, para.nv, para.nu, 0);
  if (!vecFlt[tid])
    vecFlt[tid] = std::make_shared<FilterEngine>(para.cos_image_table.cos_sigma.GetFrame(),
                                                 para.fft_kernel_real_f32);
  vecFlt[tid]->Filtering(src, dst, true);
}

// cuFDK/mpiFDK.cpp:167-203, 245-302
struct BpThreadLike {
  std::shared_ptr<BackProjectionFDK> pBP;
  std::vector<float> projImg, projMat;
  int nNum = 0;
  void PushBpData(const Parameter& para, int id, const float* pImg, const float* pMat) {
    bool bBp = false;
    if (pImg && pMat) {
      std::memcpy(&projImg.data()[para.nu * para.nv * nNum], pImg, para.nu * para.nv * sizeof(float));
      std::memcpy(&projMat.data()[cudaBackProjection::PROJ_MAT_SIZE * nNum], pMat,
                  sizeof(float) * cudaBackProjection::PROJ_MAT_SIZE);
      nNum++;
      if (nNum >= cudaBackProjection::MAX_PROJ) bBp = true;
    } else if (pImg == NULL && pMat == NULL) {
      VERIFY_TRUE(id < 0);
      if (nNum > 0) bBp = true;
    }
    if (bBp) {
      pBP->BP(1, projImg.data(), para.nv, para.nu, projMat.data(), nNum);
      nNum = 0;
    }
  }
  void OnProc(const Parameter& para, geo::BoxT<size_t>& _subTop, geo::BoxT<size_t>& _subBottom,
              float* buffer, size_t buffer_size, const char* data, int id) {
    size_t _projWidth = para.nv;
    size_t _projHeight = para.nu;
    size_t _projCount = cudaBackProjection::MAX_PROJ;
    size_t _nx = para.nx;
    size_t _ny = para.ny;
    size_t _nz = para.nz;
    bool bDualBuffer = false;
    pBP = std::make_shared<BackProjectionFDK>(_projWidth, _projHeight, _projCount, _nx, _ny, _nz,
                                              _subTop, _subBottom, bDualBuffer);
    projImg.resize(_projWidth * _projHeight * _projCount);
    projMat.resize(cudaBackProjection::PROJ_MAT_SIZE * _projCount);
    PushBpData(para, id, (float*)data,
               &para.vec_rot_mat_ijk3x4[0] + cudaBackProjection::PROJ_MAT_SIZE * id);
    PushBpData(para, -1, NULL, NULL);
    std::printf("nBpCount=%d\n", pBP->nBpCount);
    pBP->GetVolumeData(buffer, buffer_size);
  }
};

// cudaBP/cudaBP.cpp:85-153
void test_fdk(int TRANPOSE, const Parameter& para, bool bDualBuffer, bool bSaveVolume) {
  typedef BackProjectionFDK BackprojectionType;
  ProjectionImageT<float> proj;
  VERIFY_TRUE(proj.MallocBuffer(para.nu, para.nv, para.nProj).IsValid());
  VERIFY_TRUE(proj.LoadFolder<float>(para.str_proj_img_folder.c_str(), para.nu, para.nv, para.nProj));
  if (TRANPOSE == 1) proj.Transpose();
  std::shared_ptr<BackprojectionType> pBp;
  geo::BoxT<size_t> subTop(0, 0, 0, para.nx, para.ny, para.nz / 4);
  geo::BoxT<size_t> subBottom(0, 0, para.nz * 3 / 4, subTop.width, subTop.height, 0);
  pBp = std::make_shared<BackprojectionType>(proj.width, proj.height, cudaBackProjection::MAX_PROJ,
                                             para.nx, para.ny, para.nz, subTop, subBottom,
                                             bDualBuffer);
  BackprojectionType& bp = *pBp;
  for (int i = 0; i < para.nProj; i += cudaBackProjection::MAX_PROJ) {
    if (i + cudaBackProjection::MAX_PROJ > para.nProj) break;
    float* pProjData = proj.GetFrame(i).buffer;
    float* pProjMat = NoConstant(&para.vec_rot_mat_ijk3x4[0] + 12 * i);
    bp.BP(TRANPOSE, pProjData, proj.width, proj.height, pProjMat, cudaBackProjection::MAX_PROJ);
  }
  float gups = (para.nx / 1000.0) * (para.ny / 1000.0) * (para.nz / 1000.0) * para.nProj /
               (bp.fKernelTime / 1000.0);
  (void)gups;
  VolumeT<float> vol;
  vol.MallocBuffer(subTop.depth + subBottom.depth, subTop.height, subTop.width);
  bp.GetVolumeData(vol.buffer, vol.BufferSize());
  if (bSaveVolume)
    VERIFY_TRUE(vol.Save(para.str_working_folder +
                         StringFormat("%d_%d_%d.vol.raw", subTop.width, subTop.height,
                                      subTop.depth * 2)));
  VolumeT<float> vol2;
  vol2.MallocBuffer(para.nx, para.ny, para.nz);
  bp.GetVolumeData(vol2.buffer, vol2.nx * vol2.ny * vol2.nz);
}

// FilterlibCufft/FilterlibCufft.cpp:8-20
int test_filter(const Parameter& para, int batchCount) {
  extern float FilterProjImage(const Parameter& para, const ProjectionImage& proj_img,
                               int DEPTH = 32);
  const char* pSrc = para.str_proj_img_folder.c_str();
  const char* pDst = para.str_working_folder.c_str();
  int order = std::log(para.fft_kernel_real.size()) / std::log(2) + 0.5;
  (void)order;
  ProjectionImageT<float> proj;
  VERIFY_TRUE(proj.LoadFolder<float>(pSrc, para.nu, para.nv, para.nProj, para.step));
  float tm = FilterProjImage(para, proj, batchCount);
  std::printf("filtering_%dx%d, count=%d, time=%f ms, avg=%f ms, %f fps\n", (int)para.nu,
              (int)para.nv, (int)para.nProj, tm, tm / para.nProj, para.nProj / (tm / 1000.f));
  proj.SaveToFolder(pDst);
  return 0;
}

// genProj/genProj.cpp:31-81
void forward_projection(const char* szVolume, Parameter& para) {
  ProjectionImage proj_ori, proj_fp;
  Volume volume(para.nx, para.ny, para.nz);
  VERIFY_TRUE(volume.Load(szVolume));
  volume.SetUint(para.dx, para.dy, para.dz);
  std::vector<DevData<float>> vec_dev_proj(para.nProj), vec_dev_forward_proj(para.nProj);
  DevArray3D<float> dev_vol(volume.nx, volume.ny, volume.nz /*, cudaArraySurfaceLoadStore*/);
  dev_vol.CopyFromHost(volume.buffer, volume.nx, volume.nx, volume.ny, volume.nz);
  ImageT<float> image(para.nu, para.nv);
  DevData<float> dev_image(para.nu, para.nv);
  for (int i = 0; i < para.nProj; i++) {
    cudaCTForwardProjection_f32(proj_ori, para, volume, &dev_vol, NULL, image, dev_image,
                                &vec_dev_forward_proj, NULL, i);
    char szPath[1024];
    std::sprintf(szPath, "%simg%04d.raw", para.str_working_folder.c_str(), i);
    (void)szPath;
    (void)(image.width * image.height);
  }
}

// Parameter members the callers read (common/type.h:4859-4895).
double parameter_fields(const Parameter& para) {
  return para.sx + para.sy + para.sz + para.su + para.sv + para.offset_su + para.offset_sv +
         para.AngularAndrampFactor + para.cos_image_table.csc_sigma(0, 0) +
         para.str_main_folder.size() + para.str_mask_folder.size() + para.compute_device +
         para.vec_deg[0];
}
