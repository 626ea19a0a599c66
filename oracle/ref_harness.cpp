// Test infrastructure only: this file is never part of the product library.
//
// Builds the reference's own CPU FDK code, unchanged and where it lies
// (common/type.h and 3dct/3dct.cpp under the reference checkout), into a small
// shared library with a C entry point per function the oracle restates.  It is
// the checker's checker: tests compare the C oracle (oracle/fdk_oracle.c) with
// it, and tests/golden/make_golden.py writes its outputs as fixtures.
//
// Recipe: oracle/build_ref.sh (clang++, -fdelayed-template-parsing, OpenMP).
// Output: oracle/_ref/libfdkref.so (never committed).

#define main fdk_reference_3dct_main  // keep the reference driver's main() out of the way
#include "3dct/3dct.cpp"
#undef main

#include "ref_geom.h"

namespace {

const char* kFilterNames[6] = {"none", "ram-lak", "shepp-logan", "cosine", "hamming", "hann"};

// Fill the reference's Parameter (common/type.h:4627-4895) from the plain C
// record and let Parameter::init build its tables (type.h:4719-4754).
void Setup(Parameter& p, const RefGeom* g) {
  p.nx = g->nx; p.ny = g->ny; p.nz = g->nz;
  p.dx = g->dx; p.dy = g->dy; p.dz = g->dz;
  p.nu = g->nu; p.nv = g->nv;
  p.du = g->du; p.dv = g->dv;
  p.offset_nu = g->offset_nu; p.offset_nv = g->offset_nv;
  p.SOD = g->sod; p.SID = g->sid;
  p.direction = g->direction;
  p.scan_angle = g->scan_angle_deg;
  p.nProj = g->nproj;
  p.volume_postion_angle = g->volume_position_angle_rad;
  p.step = g->step;
  p.str_filter = kFilterNames[g->filter_type];
  p.init();
}

}  // namespace

extern "C" {

// Geometry tables: cos table (nu*nv floats, x fastest), filter (N doubles),
// 3x4 matrices (nproj*12 floats), AngularAndrampFactor.
int ref_tables(const RefGeom* g, float* cos_out, double* filt_out, int filt_cap,
               float* mats_out, double* af_out) {
  Parameter p;
  Setup(p, g);
  const int n = (int)p.fft_kernel_real.size();
  if (n > filt_cap) return -1;
  for (int i = 0; i < n; ++i) filt_out[i] = p.fft_kernel_real[i];
  for (int i = 0; i < g->nu * g->nv; ++i) cos_out[i] = p.cos_image_table.cos_sigma.buffer[i];
  for (int i = 0; i < g->nproj * 12; ++i) mats_out[i] = p.vec_rot_mat_ijk3x4[i];
  *af_out = p.AngularAndrampFactor;
  return n;
}

// filter() of type.h:4427 for an arbitrary length/type.
int ref_filter_design(int type, int len, int d, double* out, int cap) {
  std::vector<double> v = filter(kFilterNames[type], len, d);
  if ((int)v.size() > cap) return -1;
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
  return (int)v.size();
}

// filterImageRow of type.h:4169 (double FFT, 3-sample edge-mean pad).
int ref_filter_rows_3dct(const float* src, float* dst, int width, int height,
                         const double* filt, int n) {
  std::vector<double> f(filt, filt + n);
  filterImageRow<float, float>(f, src, dst, width, height);
  return 0;
}

// Bourke FFT of type.h:4099 in double, in place.
int ref_fft(int dir, int m, double* re, double* im) {
  return FFT<double>(dir, m, re, im) ? 0 : -1;
}

// 3dct CPU FDK: BackProjection<float> (3dct/3dct.cpp:3-241) over all views.
// frames: nproj x nv x nu floats (u fastest).  vol: nz x ny x nx (x fastest),
// accumulated into.  prefiltered != 0 feeds already-filtered projections through
// the str_filter=="none" branch, undoing its 1/(nProj/step) scale first.
int ref_bp_3dct_views(const RefGeom* g, const float* frames, int prefiltered, int nviews,
                      float* vol_out);

int ref_bp_3dct(const RefGeom* g, const float* frames, int prefiltered, float* vol_out) {
  return ref_bp_3dct_views(g, frames, prefiltered, g->nproj, vol_out);
}

// As ref_bp_3dct over views [0, nviews) only (the geometry keeps its nproj, so
// the angles and weights are those of the full scan).
int ref_bp_3dct_views(const RefGeom* g, const float* frames, int prefiltered, int nviews,
                      float* vol_out) {
  Parameter p;
  RefGeom gg = *g;
  if (prefiltered) gg.filter_type = 0;
  Setup(p, &gg);
  Volume vol(g->nx, g->ny, g->nz);
  vol.Zeros();
  VolumeT<ushort> cnt(g->nx, g->ny, g->nz);
  const size_t frame_size = (size_t)g->nu * g->nv;
  std::vector<float> buf(frame_size);
  const float undo = float(p.nProj) / p.step;
  for (int s = 0; s < nviews && s < g->nproj; s += g->step) {
    for (size_t i = 0; i < frame_size; ++i)
      buf[i] = prefiltered ? frames[s * frame_size + i] * undo : frames[s * frame_size + i];
    FrameT<float> frame(buf.data(), g->nu, g->nv, s);
    const double angle = p.direction * p.vec_deg[s] / 180.0 * PI + p.volume_postion_angle;
    BackProjection<float>(frame, p, angle, vol, cnt);
  }
  const size_t nvox = (size_t)g->nx * g->ny * g->nz;
  for (size_t i = 0; i < nvox; ++i) vol_out[i] += vol.buffer[i];
  return 0;
}

// 3dct CPU ray-march ForwardProjection<float> (3dct/3dct.cpp:430-573) of one view.
int ref_fp_3dct(const RefGeom* g, const float* vol_in, int view, float* out) {
  Parameter p;
  Setup(p, g);
  Volume vol(vol_in, g->nx, g->ny, g->nz);
  ImageT<float> image(g->nu, g->nv);
  ImageT<float> frm(g->nu, g->nv);
  const double angle = -(p.direction * p.vec_deg[view] / 180.0 * PI + p.volume_postion_angle);
  ForwardProjection<float>(p, angle, vol, frm.GetFrame(), image);
  for (int i = 0; i < g->nu * g->nv; ++i) out[i] = image.buffer[i];
  return 0;
}

}  // extern "C"
