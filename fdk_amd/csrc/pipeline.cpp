// Streaming reconstruction pipeline: host views in 32-view batches ->
// host-to-device copy on a copy stream -> (cos weighting + ramp filter) on a
// filter stream -> backprojection on a compute stream, with a ring of device
// batch slots so that the copy and the filtering of batch k+1 run under the
// backprojection of batch k.
//
// The single-node counterpart of the reference's FilterThread / BpThread pair
// joined by SafeList queues (cuFDK/mpiFDK.cpp:8-317, PushBpData 167-203): the
// reference filters on the host in one thread while another makes synchronous
// BP calls (cudaBP/backprojectionFDK.cu:229-276); here one host thread
// enqueues, three streams carry the copy, the filter and the backprojection,
// and HIP events order them per slot.  16-bit storage keeps each batch's
// per-view min/max on the device, where the backprojector reads its scale and
// offset, so nothing waits on the host between filter and backprojection.  A
// push returns once its batch has been copied, so the caller may refill its
// host buffer at once (page-locked buffers keep the copy asynchronous to the
// other streams' kernels).  Device buffers are allocated at creation, the
// raw-view buffers at the first raw push; later pushes allocate nothing.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "common.hpp"

struct fdk_pipe {
  int dev = 0;
  fdk_geometry g{};
  fdk_bp_t bp = nullptr;
  int store_dtype = FDK_F32;
  int pad_mode = FDK_PAD_MEAN4;
  int N = 0;
  hipStream_t copy = nullptr, filter = nullptr, compute = nullptr;
  float* d_cos = nullptr;
  float* d_filter = nullptr;
  float* d_stage = nullptr;  // FP32 filtered batch before 16-bit quantisation (filter stream only)
  struct Slot {
    void* raw = nullptr;       // [32][nv][nu] FP32, raw views (allocated on first raw push)
    void* store = nullptr;     // [32][nu][nv] store dtype, filtered and transposed
    unsigned* mm = nullptr;    // [32][2] per-view min/max words of a 16-bit filtered batch
    hipEvent_t copied = nullptr, flt_start = nullptr, flt_end = nullptr, bp_start = nullptr,
               bp_end = nullptr;
    bool pending = false;   // timing events recorded and not yet read
    bool filtered = false;  // this use ran the filter (flt_start .. flt_end)
  };
  std::vector<Slot> slots;
  int next = 0;
  size_t view_elems = 0;
  double bp_ms = 0.0, filter_ms = 0.0, h2d_bytes = 0.0;
  long long views = 0;
};

namespace {

int esize(int dt) { return dt == FDK_F32 ? 4 : 2; }

int collect(fdk_pipe* p, fdk_pipe::Slot& s) {
  if (!s.pending) return 0;
  FDK_HIP(hipEventSynchronize(s.bp_end));
  float ms = 0.f;
  FDK_HIP(hipEventElapsedTime(&ms, s.bp_start, s.bp_end));
  p->bp_ms += ms;
  if (s.filtered) {
    FDK_HIP(hipEventElapsedTime(&ms, s.flt_start, s.flt_end));
    p->filter_ms += ms;
  }
  s.pending = false;
  return 0;
}

int release(fdk_pipe* p) {
  if (!p) return 0;
  (void)hipSetDevice(p->dev);
  if (p->compute) (void)hipStreamSynchronize(p->compute);
  if (p->filter) (void)hipStreamSynchronize(p->filter);
  if (p->copy) (void)hipStreamSynchronize(p->copy);
  for (auto& s : p->slots) {
    if (s.raw) (void)hipFree(s.raw);
    if (s.store) (void)hipFree(s.store);
    if (s.mm) (void)hipFree(s.mm);
    for (hipEvent_t e : {s.copied, s.flt_start, s.flt_end, s.bp_start, s.bp_end})
      if (e) (void)hipEventDestroy(e);
  }
  if (p->bp) fdk_bp_destroy(p->bp);
  if (p->d_cos) (void)hipFree(p->d_cos);
  if (p->d_filter) (void)hipFree(p->d_filter);
  if (p->d_stage) (void)hipFree(p->d_stage);
  if (p->copy) (void)hipStreamDestroy(p->copy);
  if (p->filter) (void)hipStreamDestroy(p->filter);
  if (p->compute) (void)hipStreamDestroy(p->compute);
  delete p;
  return 0;
}

int setup(fdk_pipe* p, const fdk_geometry* g, const int64_t top[6], const int64_t bottom[6],
          int nslots) {
  FDK_HIP(hipGetDevice(&p->dev));
  p->g = *g;
  p->view_elems = (size_t)g->nu * g->nv;
  p->N = fdk_geom_filter_length(g);
  if (p->N < 0) return p->N;
  std::vector<float> cos(p->view_elems);
  int rc = fdk_geom_cos_table(g, cos.data());
  if (rc < 0) return rc;
  std::vector<double> h(p->N);
  rc = fdk_geom_filter(g, h.data(), p->N);
  if (rc < 0) return rc;
  std::vector<float> hf(h.begin(), h.end());
  rc = fdk_bp_create(g->nv, g->nu, FDK_BP_MAX_PROJ, g->nx, g->ny, g->nz, top, bottom, &p->bp);
  if (rc < 0) return rc;
  FDK_HIP(hipStreamCreateWithFlags(&p->copy, hipStreamNonBlocking));
  FDK_HIP(hipStreamCreateWithFlags(&p->filter, hipStreamNonBlocking));
  FDK_HIP(hipStreamCreateWithFlags(&p->compute, hipStreamNonBlocking));
  rc = fdk_bp_set_stream(p->bp, p->compute);
  if (rc < 0) return rc;
  FDK_HIP(hipMalloc(&p->d_cos, cos.size() * sizeof(float)));
  FDK_HIP(hipMalloc(&p->d_filter, hf.size() * sizeof(float)));
  FDK_HIP(hipMemcpy(p->d_cos, cos.data(), cos.size() * sizeof(float), hipMemcpyHostToDevice));
  FDK_HIP(hipMemcpy(p->d_filter, hf.data(), hf.size() * sizeof(float), hipMemcpyHostToDevice));
  p->slots.resize(nslots);
  const size_t store_bytes = FDK_BP_MAX_PROJ * p->view_elems * esize(p->store_dtype);
  for (auto& s : p->slots) {
    FDK_HIP(hipMalloc(&s.store, store_bytes));
    if (p->store_dtype != FDK_F32)
      FDK_HIP(hipMalloc((void**)&s.mm, 2 * FDK_BP_MAX_PROJ * sizeof(unsigned)));
    for (hipEvent_t* e : {&s.copied, &s.flt_start, &s.flt_end, &s.bp_start, &s.bp_end})
      FDK_HIP(hipEventCreate(e));
  }
  return 0;
}

int mats_for(const fdk_pipe* p, int first, int step, int count, float* out) {
  for (int k = 0; k < count; ++k) {
    double m[12];
    const int rc = fdk_geom_proj_matrices(&p->g, first + k * step, 1, m);
    if (rc < 0) return rc;
    for (int e = 0; e < 12; ++e) out[12 * k + e] = (float)m[e];
  }
  return 0;
}

// One batch: copy host views into the next slot, filter them when raw, and
// backproject them; returns after the copy has finished.
int push(fdk_pipe* p, const void* host, int count, int first, int step, bool raw,
         const float* scale_offset) {
  FDK_REQUIRE(p, "null pipeline");
  FDK_REQUIRE(host, "null host views");
  FDK_REQUIRE(count >= 1 && count <= FDK_BP_MAX_PROJ, "count %d outside 1..%d", count,
              FDK_BP_MAX_PROJ);
  FDK_REQUIRE(step >= 1, "step must be positive");
  FDK_REQUIRE(first >= 0 && (long long)first + (long long)(count - 1) * step < p->g.nproj,
              "views %d + k*%d (k < %d) outside 0..%d", first, step, count, p->g.nproj - 1);
  FDK_REQUIRE(raw || p->store_dtype == FDK_F32 || scale_offset,
              "16-bit filtered views need their scale_offset");
  FDK_HIP(hipSetDevice(p->dev));
  float mats[FDK_BP_MAX_PROJ * 12];
  int rc = mats_for(p, first, step, count, mats);
  if (rc < 0) return rc;

  fdk_pipe::Slot& s = p->slots[p->next];
  p->next = (p->next + 1) % (int)p->slots.size();
  // Waits until this slot's previous backprojection (and so its filtering)
  // has finished: the slot's buffers are then free to overwrite.
  rc = collect(p, s);
  if (rc < 0) return rc;
  if (raw && !s.raw)
    FDK_HIP(hipMalloc(&s.raw, FDK_BP_MAX_PROJ * p->view_elems * sizeof(float)));
  if (raw && p->store_dtype != FDK_F32 && !p->d_stage)
    FDK_HIP(hipMalloc((void**)&p->d_stage, FDK_BP_MAX_PROJ * p->view_elems * sizeof(float)));

  const size_t bytes = count * p->view_elems * (raw ? sizeof(float) : esize(p->store_dtype));
  FDK_HIP(hipMemcpyAsync(raw ? s.raw : s.store, host, bytes, hipMemcpyHostToDevice, p->copy));
  FDK_HIP(hipEventRecord(s.copied, p->copy));

  s.filtered = raw;
  const bool dev_scale = raw && p->store_dtype != FDK_F32;
  if (raw) {
    // Filter stream: cos weighting + ramp filter + transposition (+ per-view
    // min/max and quantisation for 16-bit storage) once the copy has landed.
    FDK_HIP(hipStreamWaitEvent(p->filter, s.copied, 0));
    FDK_HIP(hipEventRecord(s.flt_start, p->filter));
    if (p->store_dtype == FDK_F32)
      rc = fdk_filter_views(static_cast<const float*>(s.raw), (int64_t)p->view_elems, s.store,
                            (int64_t)p->view_elems, FDK_F32, p->g.nu, p->g.nv, count, p->d_cos,
                            p->d_filter, p->N, p->pad_mode, 1, nullptr, p->filter);
    else
      rc = fdk_filter_views_async(static_cast<const float*>(s.raw), (int64_t)p->view_elems,
                                  s.store, (int64_t)p->view_elems, p->store_dtype, p->g.nu,
                                  p->g.nv, count, p->d_cos, p->d_filter, p->N, p->pad_mode, 1,
                                  p->d_stage, s.mm, p->filter);
    if (rc < 0) return rc;
    FDK_HIP(hipEventRecord(s.flt_end, p->filter));
    FDK_HIP(hipStreamWaitEvent(p->compute, s.flt_end, 0));
  } else {
    FDK_HIP(hipStreamWaitEvent(p->compute, s.copied, 0));
  }
  FDK_HIP(hipEventRecord(s.bp_start, p->compute));
  if (dev_scale)
    rc = fdk_bp_batch_device_minmax(p->bp, s.store, p->store_dtype, p->g.nv, p->g.nu,
                                    (int64_t)p->view_elems, mats, s.mm, count);
  else
    rc = fdk_bp_batch_device(p->bp, 1, s.store, p->store_dtype, p->g.nv, p->g.nu,
                             (int64_t)p->view_elems, mats, raw ? nullptr : scale_offset, count);
  if (rc < 0) return rc;
  FDK_HIP(hipEventRecord(s.bp_end, p->compute));
  s.pending = true;
  p->views += count;
  p->h2d_bytes += (double)bytes;
  FDK_HIP(hipEventSynchronize(s.copied));
  return 0;
}

}  // namespace

extern "C" {

int fdk_pipe_create(const fdk_geometry* g, const int64_t box_top[6], const int64_t box_bottom[6],
                    int store_dtype, int pad_mode, int slots, fdk_pipe_t* out) {
  FDK_REQUIRE(out, "null output handle");
  *out = nullptr;
  FDK_REQUIRE(g && box_top, "null geometry or box");
  FDK_REQUIRE(store_dtype == FDK_F32 || store_dtype == FDK_F16 || store_dtype == FDK_BF16,
              "store dtype %d", store_dtype);
  FDK_REQUIRE(pad_mode >= FDK_PAD_ZERO && pad_mode <= FDK_PAD_MEAN4, "pad_mode %d", pad_mode);
  FDK_REQUIRE(slots >= 2 && slots <= 4, "slots %d outside 2..4", slots);
  FDK_REQUIRE(g->nu > 0 && g->nv > 0 && g->nproj > 0, "geometry sizes must be positive");
  fdk_pipe* p = new fdk_pipe();
  p->store_dtype = store_dtype;
  p->pad_mode = pad_mode;
  const int rc = setup(p, g, box_top, box_bottom, slots);
  if (rc < 0) {
    char msg[1024];
    fdk_last_error(msg, sizeof(msg));
    release(p);
    return ::fdk::fail(rc, "%s", msg);
  }
  *out = p;
  return 0;
}

int fdk_pipe_push_raw(fdk_pipe_t p, const float* host_views, int count, int first_view,
                      int view_step) {
  return push(p, host_views, count, first_view, view_step, true, nullptr);
}

int fdk_pipe_push_filtered(fdk_pipe_t p, const void* host_views, int count, int first_view,
                           int view_step, const float* scale_offset) {
  return push(p, host_views, count, first_view, view_step, false, scale_offset);
}

int fdk_pipe_finish(fdk_pipe_t p) {
  FDK_REQUIRE(p, "null pipeline");
  FDK_HIP(hipSetDevice(p->dev));
  FDK_HIP(hipStreamSynchronize(p->copy));
  FDK_HIP(hipStreamSynchronize(p->filter));
  FDK_HIP(hipStreamSynchronize(p->compute));
  for (auto& s : p->slots) {
    const int rc = collect(p, s);
    if (rc < 0) return rc;
  }
  return 0;
}

int fdk_pipe_stats(fdk_pipe_t p, double out[4]) {
  FDK_REQUIRE(p && out, "null argument");
  out[0] = (double)p->views;
  out[1] = p->h2d_bytes;
  out[2] = p->filter_ms;
  out[3] = p->bp_ms;
  return 0;
}

fdk_bp_t fdk_pipe_bp(fdk_pipe_t p) { return p ? p->bp : nullptr; }

int fdk_pipe_destroy(fdk_pipe_t p) { return release(p); }

}  // extern "C"
