// Host-only check of the C++ drop-in formats: reads the projection folder and
// volume file written by fdk_amd.formats, writes its own, and prints sums.
//   formats_check <dir>
#include <cstdio>
#include <string>

#include "fdk_amd.hpp"

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::string dir = std::string(argv[1]) + "/";
  ProjectionImageT<float> p;
  if (!p.LoadFolder(dir + "proj/", 6, 5, 7, 3, 0)) return 3;
  double s = 0.0;
  for (float v : p.data) s += v;
  std::printf("proj_sum %.6f frame3_0 %.6f frame1_0 %.6f\n", s, p.data[3 * 30], p.data[30]);
  ProjectionImageT<float> bad;
  if (bad.LoadFolder(dir + "proj/", 6, 4, 7, 3, 0)) return 4;  // wrong size must fail
  VolumeT<float> v;
  if (!v.Load(dir + "vol.raw")) return 5;
  double t = 0.0;
  for (float x : v.data) t += x;
  std::printf("vol %d %d %d sum %.6f\n", (int)v.nx, (int)v.ny, (int)v.nz, t);
  VolumeT<float> w(3, 2, 4);
  for (size_t i = 0; i < w.data.size(); ++i) w.data[i] = 0.5f * (float)i;
  if (!w.Save(dir + "vol_cpp.raw")) return 6;
  if (!w.SaveSlices(dir + "slices_cpp/", 10)) return 7;
  if (!p.SaveToFolder(dir + "proj_cpp/", 3)) return 8;
  VolumeT<float> shifted = w;
  for (float& x : shifted.data) x += 0.05f;
  const ImageQuality q = CheckImageQuality(w, shifted, -1024, 1024);
  std::printf("quality %.9f %.9f %.9f %.9f %d\n", q.error_per_pixel, q.mse, q.psnr, q.qi,
              (int)q.passed);
  std::printf("ncc %.12f\n", NCC(w.data.data(), shifted.data.data(), w.data.size()));
  return 0;
}
