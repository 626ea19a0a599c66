/* Test infrastructure only.  Plain-C geometry record shared by the oracle
 * (fdk_oracle.c) and the reference harness (ref_harness.cpp).  Field meaning
 * follows the reference's Parameter (common/type.h:4627-4895); note the naming:
 * sod = source-to-object (reference SOD), sid = source-to-detector (reference SID). */
#ifndef FDK_ORACLE_REF_GEOM_H
#define FDK_ORACLE_REF_GEOM_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct RefGeom {
  int nx, ny, nz;
  double dx, dy, dz;
  int nu, nv;
  double du, dv;
  double offset_nu, offset_nv;
  double sod, sid;
  int direction;
  double scan_angle_deg;
  int nproj;
  double volume_position_angle_rad;
  int step;
  int filter_type; /* 0 none, 1 ram-lak, 2 shepp-logan, 3 cosine, 4 hamming, 5 hann */
} RefGeom;

#ifdef __cplusplus
}
#endif
#endif
