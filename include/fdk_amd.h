/* fdk_amd -- MI355X (gfx950) FDK cone-beam reconstruction: C ABI.
 *
 * Plain C types only: pointers, sizes, ints, doubles.  Every entry point
 * returns an int status (FDK_OK == 0, negative on error) unless noted; the
 * message of the last error on the calling thread is fdk_last_error().
 * "dev" pointers are device (HBM) addresses, "host" pointers host memory.
 * "stream" is a hipStream_t passed as void* (NULL = the null stream).
 *
 * Reference interfaces replaced (peng-ac/fdk, file:line):
 *   fdk_geom_*           Parameter::init / CalculateCosTable / filter() /
 *                        CalculateProjMat        common/type.h:4719-4944, 4427-4475
 *   fdk_bp_*             BackProjectionFDK ctor / BP / GetVolumeData
 *                                                 cudaBP/backprojection.cuh:246-271,
 *                                                 cudaBP/backprojectionFDK.cu:161-302
 *   fdk_rtk_*            BackProjectionRTK ctor / BP / GetVolumeData
 *                                                 cudaBP/backprojection.cuh:180-198,
 *                                                 rtkCudaFDKBackProjectionImageFilter.cu:340-400
 *   fdk_reconstruct_conebeam
 *                        BackProjectionRTK::CUDA_reconstruct_conebeam
 *                                                 rtkCudaFDKBackProjectionImageFilter.cu:200-330
 *   fdk_correct_image    CorrectImage_Real32      FilterlibCufft/cudaFftFilter.cuh:9
 *   fdk_filter_rows      FilterRow_Real32         FilterlibCufft/cudaFftFilter.cuh:7
 *   fdk_filter_views     fused cos weight + row filter (+ transpose / FP16 / BF16 store),
 *                        the device half of FilterProjImage (cudaFftFilter.cu:394-452)
 *                        and FilterEngine::Filtering (Filterlib/FilterLib.h:167-200)
 *   fdk_filter_proj_image FilterProjImage (host in, host out, in place)
 *                                                 FilterlibCufft/cudaFftFilter.cu:394
 *   fdk_pipe_*           FilterThread / BpThread / PushBpData   cuFDK/mpiFDK.cpp:8-317
 *   fdk_fp_project       cudaCTForwardProjection_f32 / ForwardProjection
 *                                                 genProj/cudaIterType.cuh:280,
 *                                                 genProj/cudaForwardProjection.cu:338-487
 */
#ifndef FDK_AMD_H
#define FDK_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDK_OK 0
#define FDK_ERR_INVALID (-1)
#define FDK_ERR_OOM (-2)
#define FDK_ERR_HIP (-3)
#define FDK_ERR_UNSUPPORTED (-4)

/* Stored sample types of projections. */
#define FDK_F32 0
#define FDK_F16 1
#define FDK_BF16 2

/* Row padding before the FFT: zero, mean of the 3 or 4 edge samples. */
#define FDK_PAD_ZERO 0
#define FDK_PAD_MEAN3 3  /* FilterlibCufft kernelPad, type.h filterImageRow */
#define FDK_PAD_MEAN4 4  /* Filterlib FilterImageRow (the CPU reference) */

/* Forward-projector sampling modes. */
#define FDK_FP_CLAMP 0   /* genProj GPU: t < tfar, clamp-to-edge trilinear */
#define FDK_FP_ZERO 1    /* 3dct CPU:   t <= tfar, zero outside the volume */

/* Scan geometry, the reference's Parameter (common/type.h:4627-4895).
 * NOTE the reference naming kept here: sod = source-to-object distance
 * (reference SOD), sid = source-to-detector distance (reference SID), mm. */
typedef struct fdk_geometry {
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
} fdk_geometry;

/* Random-phantom ellipsoid (no reference counterpart): additive density inside
 * |R^T (p - c) / a| <= 1, p in the volume frame (mm). */
typedef struct fdk_ellipsoid {
  double cx, cy, cz;
  double ax, ay, az;
  double rot[9];
  double density;
} fdk_ellipsoid;

int fdk_last_error(char* buf, size_t len);
int fdk_version(void);
/* Number of HIP devices visible (0 when none); never fails. */
int fdk_device_count(void);

/* Device memory for the C++ layer's DevData / DevArray3D (cudaBP/cudaLib.cuh:298-420,
 * 552-640): linear allocations on the current device; copies use unified
 * addressing (host or device pointers on either side) and are synchronous. */
int fdk_dev_malloc(void** out, size_t bytes);
int fdk_dev_free(void* p);
int fdk_memcpy(void* dst, const void* src, size_t bytes);
int fdk_memcpy_2d(void* dst, size_t dst_pitch, const void* src, size_t src_pitch,
                  size_t width_bytes, size_t rows);
int fdk_dev_memset(void* p, int value, size_t bytes);

/* ------------------------------------------------------------ geometry (host, FP64) */
/* type.h:4736 */
int fdk_geom_angular_ramp_factor(const fdk_geometry* g, double* out);
/* type.h:4756-4785: nv x nu floats, u fastest. */
int fdk_geom_cos_table(const fdk_geometry* g, float* out);
/* type.h:4418-4425: FFT length N = nextpow2(nu) (strictly greater power of two). */
int fdk_geom_filter_length(const fdk_geometry* g);
/* type.h:4427-4475: N doubles. */
int fdk_geom_filter(const fdk_geometry* g, double* out, int cap);
/* type.h:4732, 4824: view angle in rad (direction * deg + volume_position_angle). */
int fdk_geom_view_angle(const fdk_geometry* g, int view, double* out);
/* type.h:4898-4944: count matrices (3x4 row-major doubles) for views first.. */
int fdk_geom_proj_matrices(const fdk_geometry* g, int first, int count, double* out);

/* ------------------------------------------------------------ backprojection (FDK) */
typedef struct fdk_bp* fdk_bp_t;
#define FDK_BP_MAX_PROJ 32

/* BackProjectionFDK(projWidth, projHeight, projCount, nx, ny, nz, subTop, subBottom).
 * box_top / box_bottom: {x, y, z, width, height, depth}; box_bottom depth 0 means
 * no mirror slab.  With a mirror slab box_top.z + box_bottom.z + box_bottom.depth
 * must equal nz (backprojectionFDK.cu:173-175).  Allocates the volume
 * (zeroed) and a projCount-view staging buffer on the current device. */
int fdk_bp_create(int proj_width, int proj_height, int proj_count, int nx, int ny, int nz,
                  const int64_t box_top[6], const int64_t box_bottom[6], fdk_bp_t* out);
int fdk_bp_destroy(fdk_bp_t h);
/* Stream for all later work of the handle (default: a stream the handle owns). */
int fdk_bp_set_stream(fdk_bp_t h, void* stream);
/* BackProjectionFDK::BP (backprojectionFDK.cu:229-276): synchronous.
 * transpose 1: each view is width (= nv) columns x height (= nu) rows, v fastest;
 * transpose 0: width (= nu) x height (= nv), u fastest.  mats: nproj x 12 floats
 * acting on (x, y, z, 1) voxel indices.  nproj <= FDK_BP_MAX_PROJ. */
int fdk_bp_batch(fdk_bp_t h, int transpose, const float* host_proj, int width, int height,
                 const float* mats, int nproj);
/* Device-resident batch, asynchronous on the handle's stream.  dev_proj holds
 * nproj views view_stride elements apart, transposed layout (width = nv
 * contiguous, height = nu rows) for FDK_F16/FDK_BF16; FDK_F32 also takes
 * transpose 0.  scale_offset (host, 2*nproj floats or NULL) maps stored samples
 * q to projection values p = q*scale + offset per view. */
int fdk_bp_batch_device(fdk_bp_t h, int transpose, const void* dev_proj, int dtype, int width,
                        int height, int64_t view_stride, const float* mats,
                        const float* scale_offset, int nproj);
/* As fdk_bp_batch_device for FDK_F16 / FDK_BF16 transposed views whose per-view
 * {min, max} words (2*nproj, order-preserving uints as fdk_filter_views_async
 * leaves them) stay on the device: scale = max - min and offset = min are
 * formed in the kernel, so no host round trip sits between the filter and the
 * backprojection (the reference's BpThread takes host batches,
 * cuFDK/mpiFDK.cpp:245; this is the device form of that hand-over). */
int fdk_bp_batch_device_minmax(fdk_bp_t h, const void* dev_proj, int dtype, int width, int height,
                               int64_t view_stride, const float* mats,
                               const unsigned* dev_minmax, int nproj);
int fdk_bp_sync(fdk_bp_t h);
/* Views backprojected so far (nBpCount) and summed kernel time (fKernelTime, ms;
 * only synchronous calls record it). */
int fdk_bp_count(fdk_bp_t h);
float fdk_bp_kernel_ms(fdk_bp_t h);
int fdk_bp_zero(fdk_bp_t h);
/* GetVolumeData: host [z over top then bottom][y][x], count floats. */
int fdk_bp_get_volume(fdk_bp_t h, float* host_dst, size_t count);
/* Same layout into device memory (dst must hold the whole sub-volume). */
int fdk_bp_get_volume_device(fdk_bp_t h, float* dev_dst, size_t count);
/* The working volume itself: device pointer and its extents {nxs, nys, nzz};
 * layout [x][y][zz] with zz fastest (zz: the top slab, then the bottom slab,
 * each in increasing z).  Valid until fdk_bp_destroy. */
int fdk_bp_volume_ptr(fdk_bp_t h, void** dev_ptr, int64_t dims[3]);

/* ------------------------------------------------------------ backprojection (RTK) */
typedef struct fdk_rtk* fdk_rtk_t;
/* BackProjectionRTK(projWidth, projHeight, projCount, nx, ny, nz): volume
 * [z][y][x] x fastest, in/out ping-pong as the reference. */
int fdk_rtk_create(int proj_width, int proj_height, int proj_count, int nx, int ny, int nz,
                   fdk_rtk_t* out);
int fdk_rtk_destroy(fdk_rtk_t h);
/* BackProjectionRTK::BP: transpose must be 0; host projections width (= nu) x
 * height (= nv); synchronous. */
int fdk_rtk_bp(fdk_rtk_t h, int transpose, const float* host_proj, int width, int height,
               const float* mats, int nproj);
int fdk_rtk_get_volume(fdk_rtk_t h, float* host_dst, size_t count);
float fdk_rtk_kernel_ms(fdk_rtk_t h);
/* CUDA_reconstruct_conebeam: vol_out = vol_in + backprojection of proj_size[2]
 * views (<= FDK_BP_MAX_PROJ) of proj_size[0] x proj_size[1] samples.  Volumes are
 * device pointers; proj is host memory when proj_on_device is 0. */
int fdk_reconstruct_conebeam(const int proj_size[3], const int vol_size[3], const float* mats,
                             const float* dev_vol_in, float* dev_vol_out, const float* proj,
                             int proj_on_device, void* stream);

/* ------------------------------------------------------------ filtering */
/* CorrectImage_Real32: dst = src * table over depth images (table broadcast);
 * pitch, table_pitch in elements. dst may equal src. */
int fdk_correct_image(const float* dev_src, const float* dev_table, float* dev_dst, int width,
                      int height, int depth, int pitch, int table_pitch, void* stream);
/* FilterRow_Real32: each of height*depth rows of width samples (pitch apart) is
 * padded into N (power of two, 16..16384) per pad_mode, filtered by the real
 * length-N filter H and cropped back; dst may equal src. */
int fdk_filter_rows(const float* dev_src, float* dev_dst, int width, int height, int depth,
                    int src_pitch, int dst_pitch, const float* dev_filter, int N, int pad_mode,
                    void* stream);
/* Fused cos weighting + row filter + optional transpose + stored dtype.
 * src: nviews x nv x nu floats (u fastest, view_stride_in elements apart).
 * dev_cos: nv x nu floats or NULL (no weighting).  dev_filter: N floats.
 * transpose 1 writes each view as nu rows of nv (v fastest, the layout
 * fdk_bp_batch_device reads); 0 keeps nv rows of nu.  dst may equal src only
 * when transpose is 0 and dtype is FDK_F32.  dtype FDK_F16/FDK_BF16
 * stores (p - min)/(max - min) per view and writes {max - min, min} per view to
 * host_scale_offset (2*nviews floats); the call then synchronizes the stream. */
int fdk_filter_views(const float* dev_src, int64_t view_stride_in, void* dev_dst,
                     int64_t view_stride_out, int dtype, int nu, int nv, int nviews,
                     const float* dev_cos, const float* dev_filter, int N, int pad_mode,
                     int transpose, float* host_scale_offset, void* stream);
/* Asynchronous form of fdk_filter_views for FDK_F16 / FDK_BF16 output (the
 * filter stage of the streaming pipeline, FilterThread in cuFDK/mpiFDK.cpp:58-154):
 * the FP32 filtered batch goes through dev_workspace (nviews * nu * nv floats)
 * and the per-view {min, max} stay in dev_minmax (2*nviews uints, for
 * fdk_bp_batch_device_minmax); nothing synchronizes, and nothing is allocated
 * except the FFT twiddle table of length N, built once per device and kept. */
int fdk_filter_views_async(const float* dev_src, int64_t view_stride_in, void* dev_dst,
                           int64_t view_stride_out, int dtype, int nu, int nv, int nviews,
                           const float* dev_cos, const float* dev_filter, int N, int pad_mode,
                           int transpose, float* dev_workspace, unsigned* dev_minmax,
                           void* stream);
/* fdk_filter_views without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc, code under stream capture for the FP32 form).  Same
 * arguments, results and errors as fdk_filter_views, with limits nviews <= 1024
 * and N <= 16384 per call (beyond them: FDK_ERR_INVALID).  It uses a fixed
 * 136 KiB device array held by each device the library runs on until the
 * process ends: the 16-bit forms filter twice (min/max, then quantized
 * stores) instead of staging the FP32 batch, and then synchronize the stream
 * to return host_scale_offset; the FP32 form only enqueues kernels.  Every call
 * on a device shares that array, so calls that overlap on one device race. */
int fdk_filter_views_noalloc(const float* dev_src, int64_t view_stride_in, void* dev_dst,
                             int64_t view_stride_out, int dtype, int nu, int nv, int nviews,
                             const float* dev_cos, const float* dev_filter, int N, int pad_mode,
                             int transpose, float* host_scale_offset, void* stream);
/* FilterProjImage: host views (nviews x nv x nu) filtered in place, in batches of
 * depth views (32 by default when depth <= 0), cos weighting by host_cos.
 * *elapsed_ms (when not NULL) receives the summed time of the batches' upload,
 * filter and download, as the reference's return value (cudaFftFilter.cu:416-441). */
int fdk_filter_proj_image(float* host_proj, int nu, int nv, int nviews, const float* host_cos,
                          const float* host_filter, int N, int pad_mode, int depth,
                          float* elapsed_ms);

/* FilterEngine::Filtering on one host image (nv rows of nu): src is overwritten
 * with the cos-weighted image (FilterLib.h:181), dst receives the filtered
 * image, transposed (nu rows of nv) when transpose is 1.  Synchronous. */
int fdk_filter_frame_host(float* host_src, float* host_dst, int nu, int nv, const float* host_cos,
                          const float* host_filter, int N, int pad_mode, int transpose);

/* ------------------------------------------------------------ streaming pipeline */
/* Host views in batches of <= FDK_BP_MAX_PROJ -> copy stream -> (cos weighting
 * + ramp filter) -> backprojection stream, with slots (2..4) device batch
 * buffers in a ring: the single-node form of the reference's FilterThread /
 * BpThread / SafeList pipeline and PushBpData batching
 * (cuFDK/mpiFDK.cpp:8-317, 167-203).  The volume is a slab pair as in
 * fdk_bp_create (nproj of g is the number of views of the scan; matrices come
 * from g).  Projections are stored on the device in store_dtype. */
typedef struct fdk_pipe* fdk_pipe_t;
int fdk_pipe_create(const fdk_geometry* g, const int64_t box_top[6], const int64_t box_bottom[6],
                    int store_dtype, int pad_mode, int slots, fdk_pipe_t* out);
/* Raw views first_view + k*view_step (k < count), host [count][nv][nu] FP32:
 * cos weighting, pad_mode row filter and transposition on the device, then
 * backprojection.  Returns once the batch is copied (host_views may be reused). */
int fdk_pipe_push_raw(fdk_pipe_t p, const float* host_views, int count, int first_view,
                      int view_step);
/* Views already filtered: host [count][nu][nv] in store_dtype, with 2*count
 * floats {scale, offset} per view for FDK_F16 / FDK_BF16 (value = q*scale + offset). */
int fdk_pipe_push_filtered(fdk_pipe_t p, const void* host_views, int count, int first_view,
                           int view_step, const float* scale_offset);
/* Waits for every pushed batch. */
int fdk_pipe_finish(fdk_pipe_t p);
/* out: views pushed, host-to-device bytes, filter ms, backprojection ms (device
 * time of finished batches). */
int fdk_pipe_stats(fdk_pipe_t p, double out[4]);
/* The pipeline's backprojection handle (fdk_bp_get_volume, fdk_bp_volume_ptr, ...). */
fdk_bp_t fdk_pipe_bp(fdk_pipe_t p);
int fdk_pipe_destroy(fdk_pipe_t p);

/* ------------------------------------------------------------ forward projection */
/* Ray-marching projection of views [first, first+count) of geometry g through
 * dev_vol (nz x ny x nx floats, x fastest; sizes and spacing from g), step_mm
 * (<= 0: min(dx,dy,dz), the reference's step).  dev_out: count x nv x nu floats. */
int fdk_fp_project(const fdk_geometry* g, const float* dev_vol, int first, int count,
                   double step_mm, int mode, float* dev_out, void* stream);

/* Host-memory form of fdk_fp_project (the volume is uploaded, views come back
 * to host_out); synchronous. */
int fdk_fp_project_host(const fdk_geometry* g, const float* host_vol, int first, int count,
                        double step_mm, int mode, float* host_out);

/* ------------------------------------------------------------ random phantoms */
/* Voxel-centre rasterization into dev_vol (nz x ny x nx floats, overwritten). */
int fdk_phantom_rasterize(const fdk_ellipsoid* host_ell, int n, const fdk_geometry* g,
                          float* dev_vol, void* stream);
/* Exact line integrals for views [first, first+count): dev_out count x nv x nu. */
int fdk_phantom_project(const fdk_ellipsoid* host_ell, int n, const fdk_geometry* g, int first,
                        int count, float* dev_out, void* stream);

/* ------------------------------------------------------------ learned correction */
/* Kernels of the 3-D residual U-Net (no reference counterpart: the reference has
 * no learned component).  Activations are channels-last [n][z][y][x][c] BF16
 * (uint16 storage); weights BF16 for the MFMA, FP32 masters for Adam.
 * A 3x3x3 'same' convolution with zero padding: weights [27][cout][cin]
 * (tap = 9 dz + 3 dy + dx over -1..1), or [cout][32] taps for cin == 1.
 * y = act(conv(x) + bias) + resid, stored BF16 or FP32 (out_f32); act (the
 * `relu` argument): 0 none, 1 ReLU, 2 leaky ReLU (slope 0.01 below zero).
 * Needs W % 32 == 0, H % 4 == 0, cin == 1 or cin % 32 == 0, cout in
 * {32, 64, 96, 128, 192}. */
int fdk_nn_conv3d_fwd(const void* x, const void* w, const float* bias, const float* resid, void* y,
                      int n, int d, int h, int w_, int cin, int cout, int relu, int out_f32,
                      void* stream);
/* The same convolution on a decoder input that is the channel concatenation
 * [nearest 2x upsampling of up, x] (up: [n][d/2][h/2][w/2][cup], x:
 * [n][d][h][w][cin - cup]), read from the two tensors without forming the
 * concatenation.  Bit-identical to fdk_nn_conv3d_fwd on the formed input.
 * Needs cout, cup and cin multiples of 32, even d, 16-byte aligned
 * tensors; FDK_ERR_UNSUPPORTED otherwise. */
int fdk_nn_conv3d_fwd_cat(const void* up, int cup, const void* x, const void* w, const float* bias,
                          const float* resid, void* y, int n, int d, int h, int w_, int cin,
                          int cout, int relu, int out_f32, void* stream);
/* dW[27][cout][cin] (FP32, accumulated) += sum_v dY[v][co] X[v + tap][ci];
 * cin == 1 accumulates into [cout][32]. */
int fdk_nn_conv3d_wgrad(const void* x, const void* dy, float* dw, int n, int d, int h, int w_,
                        int cin, int cout, void* stream);
/* The same weight gradient for a decoder convolution whose input is the
 * channel concatenation [nearest 2x upsampling of up, x] (up:
 * [n][d/2][h/2][w/2][cup], x: [n][d][h][w][cin - cup]), read from the two
 * tensors without forming it (the input of fdk_nn_conv3d_fwd_cat).  Needs cin,
 * cup and cout multiples of 32, W % 32, H % 4, even d, 16-byte aligned
 * tensors; FDK_ERR_UNSUPPORTED otherwise.  No counterpart in the reference (the
 * learned path's training). */
int fdk_nn_conv3d_wgrad_cat(const void* up, int cup, const void* x, const void* dy, float* dw,
                            int n, int d, int h, int w_, int cin, int cout, void* stream);
/* Backward-data weights: wb[t][ci][co] = w[26 - t][co][ci]; the input gradient
 * is then fdk_nn_conv3d_fwd(dy, wb, ..., cin = cout, cout = cin). */
int fdk_nn_flip_weights(const void* w, void* wb, int cin, int cout, void* stream);
int fdk_nn_bias_grad(const void* dy, float* db, int64_t nvox, int c, void* stream);
/* Activation adjoint at output y: dz = dy where y > 0, else 0 (mode 1, ReLU) or
 * 0.01 dy (mode 2, leaky ReLU). */
int fdk_nn_relu_bwd(const void* dy, const void* y, void* dz, int64_t n, int mode, void* stream);
/* 2x2x2 average pooling of [n][d][h][w][c] (adjoint 1: [n][d/2][h/2][w/2][c] -> full). */
int fdk_nn_avgpool2(const void* x, void* y, int n, int d, int h, int w, int c, int adjoint,
                    void* stream);
/* Nearest 2x upsampling of [n][d][h][w][cx] into channels [c0, c0+cx) of a
 * [n][2d][2h][2w][cy] tensor (adjoint 1: the reverse, summing the 8 children). */
int fdk_nn_upsample2(const void* x, void* y, int n, int d, int h, int w, int cx, int cy, int c0,
                     int adjoint, void* stream);
int fdk_nn_channel_copy(const void* src, int cs, int src_c0, void* dst, int cd, int dst_c0,
                        int cc, int64_t nvox, void* stream);
int fdk_nn_add(const void* a, const void* b, void* y, int64_t n, void* stream);
/* Output head: out = in + sum_c w[c] h[v][c] + b (FP32), and its adjoint. */
int fdk_nn_head(const void* h, const float* w, const float* b, const float* in, float* out,
                int64_t nvox, int c, void* stream);
int fdk_nn_head_bwd(const float* dout, const void* h, const float* w, void* dh, float* dw,
                    float* db, int64_t nvox, int c, void* stream);
/* Space-to-depth input, block edge f (2 or 4): FP32 x[n][d][h][w] -> BF16
 * y[n][d/f][h/f][w/f][cy], the f^3 block (f z + iz, f y + iy, f x + ix) as
 * channels (iz f + iy) f + ix, channels f^3..cy-1 zero.  d, h, w multiples of f;
 * cy % 8 == 0, cy >= f^3. */
int fdk_nn_unshuffle(const float* x, void* y, int n, int d, int h, int w, int f, int cy,
                     void* stream);
/* Depth-to-space head over the same block map: for coarse voxel v and block entry
 * k, out[entry k] = in[entry k] + b[k] + sum_c w[k][c] h[v][c] (w [f^3][c], b [f^3],
 * FP32; d, h, w the fine sizes; c == 32), and its adjoint (dh BF16; dw [f^3][c],
 * db [f^3] accumulated). */
int fdk_nn_head_shuffle(const void* h, const float* w, const float* b, const float* in, float* out,
                        int n, int d, int h_, int w_, int f, int c, void* stream);
int fdk_nn_head_shuffle_bwd(const float* dout, const void* h, const float* w, void* dh, float* dw,
                            float* db, int n, int d, int h_, int w_, int f, int c, void* stream);
/* loss_sum += sum (out - t)^2; dout = 2 (out - t) / nvox when dout != NULL. */
int fdk_nn_mse(const float* out, const float* target, float* dout, float* loss_sum, int64_t nvox,
               void* stream);
/* Adam step (bias-corrected) on FP32 weights; refreshes the BF16 copy when wbf != NULL
 * (element i -> (i / row) * bf_row + i % row). */
int fdk_nn_adam(float* w, void* wbf, const float* g, float* m, float* v, int64_t n, float lr,
                float b1, float b2, float eps, int step, int row, int bf_row, void* stream);
int fdk_nn_convert(const void* src, void* dst, int64_t n, int to_bf16, void* stream);
/* dst[dst_org + i] = scale * src[src_org + i] over a box of FP32 [D][H][W] volumes. */
int fdk_nn_box_copy(const float* src, const int src_dims[3], const int src_org[3], float* dst,
                    const int dst_dims[3], const int dst_org[3], const int box[3], float scale,
                    void* stream);
/* Photon noise on line integrals p: counts ~ Poisson(i0 exp(-mu p)) (exact
 * variates: multiplication method below 10 counts, Hoermann's PTRS above), then
 * p = -ln(max(counts, 1) / i0) / mu.  Deterministic per (seed, element index). */
int fdk_poisson_noise(float* dev_proj, int64_t n, double i0, double mu, uint32_t seed,
                      void* stream);
/* D (16x16) = A (16x32) B (32x16) through one v_mfma_f32_16x16x32_bf16 (row-major
 * FP32 in, BF16-rounded operands); a lane-layout check. */
int fdk_nn_mfma_probe(const float* A, const float* B, float* D, void* stream);

/* ---- Native U-Net inference (the learned sparse-view correction; no reference
 * counterpart).  A weight file is the .npz the trainer writes (UNet3D.save in
 * fdk_amd/unet.py: conv0..conv9 .w / .b, head.w / .b, arch.s2d, arch.act); the
 * network runs from the fdk_nn_* kernels above, bit-identical to the Python
 * sequencing of the same kernels.  A handle is bound to the device current at
 * fdk_unet_load. */
typedef struct fdk_unet fdk_unet;
/* Architecture of a weight file (base channels, space-to-depth edge 1/2/4,
 * activation 1 ReLU / 2 leaky ReLU) without touching a device. */
int fdk_unet_inspect(const char* path, int* base, int* s2d, int* act);
int fdk_unet_load(const char* path, fdk_unet** out);
int fdk_unet_info(const fdk_unet* u, int* base, int* s2d, int* act);
/* out = x + net(x) for FP32 device volumes [n][d][h][w]; d % 4f, h % 16f and
 * w % 128f == 0 (f the space-to-depth edge). */
int fdk_unet_forward(fdk_unet* u, const float* dev_x, float* dev_out, int n, int d, int h, int w,
                     void* stream);
/* Whole-volume inference of a [d][h][w] FP32 device volume in z-slabs of whole
 * slices, `slab` deep (<= 0: the whole volume for space-to-depth networks, 528
 * for full resolution), neighbours overlapping by `overlap` planes split at the
 * midpoint; dev_out must not alias dev_vol. */
int fdk_unet_infer(fdk_unet* u, const float* dev_vol, float* dev_out, int d, int h, int w, int slab,
                   int overlap, void* stream);
int fdk_unet_destroy(fdk_unet* u);

#ifdef __cplusplus
}
#endif
#endif /* FDK_AMD_H */
