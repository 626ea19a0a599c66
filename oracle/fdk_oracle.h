/* TEST INFRASTRUCTURE ONLY -- the oracle.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker.  The
 * product (fdk_amd/) never links, loads or calls it.
 *
 * Plain-C, FP64 restatement of the reference's FDK hot path.  Each function
 * cites the reference file:line it follows.  Parity of each function is pinned
 * by tests/test_oracle.py (see the header of fdk_oracle.c for the list). */
#ifndef FDK_ORACLE_H
#define FDK_ORACLE_H

#include <stddef.h>
#include "ref_geom.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_PAD_ZERO = 0, ORC_PAD_MEAN3 = 3, ORC_PAD_MEAN4 = 4 };
enum { ORC_BP_GPU_STRICT = 0, ORC_BP_CPU_TRUNC = 1 };
enum { ORC_FP_GPU_CLAMP = 0, ORC_FP_CPU_ZERO = 1 };

typedef struct OrcEllipsoid {
  double cx, cy, cz;   /* centre, mm, volume frame */
  double ax, ay, az;   /* semi-axes, mm */
  double rot[9];       /* row-major R; local = R^T (p - c) */
  double density;      /* additive */
} OrcEllipsoid;

int orc_nextpow2(int n);
int orc_filter_design(int type, int len, int d, double* out, int cap);
double orc_angular_ramp_factor(const RefGeom* g);
void orc_cos_table(const RefGeom* g, float* out);
double orc_view_angle(const RefGeom* g, int view);
void orc_proj_matrix(const RefGeom* g, double angle_rad, double* m12);
void orc_proj_matrices(const RefGeom* g, double* out);
void orc_fft(int dir, int m, double* re, double* im);
void orc_cos_weight(const float* src, const float* table, float* dst, size_t n_per_view, int nviews);
void orc_filter_rows(const float* src, float* dst, int nu, long nrows, const double* H, int N,
                     int pad_mode);
void orc_bp(const RefGeom* g, const float* proj, const double* mats, int nviews, int mode,
            int z0, int nzs, double* vol);
void orc_bp_voxels(const RefGeom* g, const float* proj, const double* mats, int nviews, int mode,
                   const int* ijk, long nvox, double* out);
void orc_fp(const RefGeom* g, const float* vol, int view, double step_mm, int mode, float* out);
void orc_phantom_rasterize(const OrcEllipsoid* e, int n, const RefGeom* g, float* vol);
void orc_phantom_project(const OrcEllipsoid* e, int n, const RefGeom* g, int view, float* out);

#ifdef __cplusplus
}
#endif
#endif
