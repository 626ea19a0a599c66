// Native inference runner of the residual U-Net of the learned sparse-view
// correction (no reference counterpart: the reference has no learned
// component).  Weights come from the .npz files the trainer writes
// (fdk_amd/unet.py UNet3D.save: conv0..conv9 .w/.b, head.w/.b, arch.s2d,
// arch.act); the forward pass is sequenced from the fdk_nn_* kernels with
// activations from the library's stream-ordered pool, each released once its
// last reader is queued; whole volumes run in overlapping z-slabs.
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

// ---------------------------------------------------------------- .npz reading
struct NpyArray {
  std::string descr;
  std::vector<int64_t> shape;
  std::vector<char> data;
  int64_t count() const {
    int64_t c = 1;
    for (int64_t s : shape) c *= s;
    return c;
  }
};

uint32_t le32(const unsigned char* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }
uint16_t le16(const unsigned char* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint64_t le64(const unsigned char* p) { return (uint64_t)le32(p) | ((uint64_t)le32(p + 4) << 32); }

// Parses a .npy blob: little-endian dtype, C order, shape from the header dict.
int parse_npy(const std::string& name, const std::vector<char>& raw, NpyArray& a) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(raw.data());
  if (raw.size() < 10 || std::memcmp(p, "\x93NUMPY", 6) != 0)
    return fdk::fail(FDK_ERR_INVALID, "%s: not a .npy array", name.c_str());
  const int major = p[6];
  size_t hl, h0;
  if (major == 1) {
    hl = le16(p + 8);
    h0 = 10;
  } else {
    if (raw.size() < 12) return fdk::fail(FDK_ERR_INVALID, "%s: truncated .npy header", name.c_str());
    hl = le32(p + 8);
    h0 = 12;
  }
  if (h0 + hl > raw.size()) return fdk::fail(FDK_ERR_INVALID, "%s: truncated .npy header", name.c_str());
  const std::string hdr(raw.data() + h0, hl);
  auto field = [&](const char* key) -> size_t {
    const size_t k = hdr.find(key);
    return k == std::string::npos ? k : hdr.find(':', k) + 1;
  };
  size_t k = field("'descr'");
  if (k == std::string::npos) return fdk::fail(FDK_ERR_INVALID, "%s: no descr", name.c_str());
  const size_t q0 = hdr.find('\'', k), q1 = hdr.find('\'', q0 + 1);
  a.descr = hdr.substr(q0 + 1, q1 - q0 - 1);
  k = field("'fortran_order'");
  if (k == std::string::npos || hdr.compare(hdr.find_first_not_of(' ', k), 5, "False") != 0)
    return fdk::fail(FDK_ERR_INVALID, "%s: Fortran-ordered or unspecified order", name.c_str());
  k = field("'shape'");
  if (k == std::string::npos) return fdk::fail(FDK_ERR_INVALID, "%s: no shape", name.c_str());
  const size_t s0 = hdr.find('(', k), s1 = hdr.find(')', s0);
  a.shape.clear();
  for (size_t i = s0 + 1; i < s1;) {
    while (i < s1 && (hdr[i] == ' ' || hdr[i] == ',')) ++i;
    if (i >= s1) break;
    a.shape.push_back(std::strtoll(hdr.c_str() + i, nullptr, 10));
    while (i < s1 && hdr[i] != ',') ++i;
  }
  const size_t esize = a.descr == "<f4" || a.descr == "<i4" ? 4 : a.descr == "<f8" || a.descr == "<i8" ? 8 : 0;
  if (!esize) return fdk::fail(FDK_ERR_UNSUPPORTED, "%s: dtype %s", name.c_str(), a.descr.c_str());
  const size_t bytes = (size_t)a.count() * esize;
  if (h0 + hl + bytes > raw.size()) return fdk::fail(FDK_ERR_INVALID, "%s: truncated data", name.c_str());
  a.data.assign(raw.data() + h0 + hl, raw.data() + h0 + hl + bytes);
  return FDK_OK;
}

// Reads every member of a .npz (a zip of .npy files, stored or deflated) by
// walking the local file headers.
int read_npz(const char* path, std::map<std::string, NpyArray>& out) {
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return fdk::fail(FDK_ERR_INVALID, "cannot open %s", path);
  std::vector<unsigned char> file;
  unsigned char chunk[1 << 16];
  size_t got;
  while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) file.insert(file.end(), chunk, chunk + got);
  std::fclose(fp);
  size_t off = 0;
  while (off + 30 <= file.size() && le32(&file[off]) == 0x04034b50u) {
    const unsigned char* h = &file[off];
    const int flags = le16(h + 6), method = le16(h + 8);
    uint64_t csize = le32(h + 18), usize = le32(h + 22);
    const int nl = le16(h + 26), el = le16(h + 28);
    if (off + 30 + nl + el > file.size()) break;
    std::string name(reinterpret_cast<const char*>(h + 30), nl);
    const unsigned char* ex = h + 30 + nl;
    for (int e = 0; e + 4 <= el;) {  // zip64 sizes when the 32-bit fields overflow
      const int id = le16(ex + e), sz = le16(ex + e + 2);
      if (id == 1 && sz >= 16 && (usize == 0xffffffffu || csize == 0xffffffffu)) {
        usize = le64(ex + e + 4);
        csize = le64(ex + e + 12);
      }
      e += 4 + sz;
    }
    if (flags & 8) return fdk::fail(FDK_ERR_UNSUPPORTED, "%s: streamed zip member %s", path, name.c_str());
    const size_t d0 = off + 30 + nl + el;
    if (d0 + csize > file.size()) return fdk::fail(FDK_ERR_INVALID, "%s: truncated member %s", path, name.c_str());
    std::vector<char> raw(usize);
    if (method == 0) {
      if (csize != usize) return fdk::fail(FDK_ERR_INVALID, "%s: bad stored member %s", path, name.c_str());
      std::memcpy(raw.data(), &file[d0], usize);
    } else if (method == 8) {
      z_stream zs{};
      if (inflateInit2(&zs, -15) != Z_OK) return fdk::fail(FDK_ERR_HIP, "zlib init failed");
      zs.next_in = const_cast<Bytef*>(&file[d0]);
      zs.avail_in = (uInt)csize;
      zs.next_out = reinterpret_cast<Bytef*>(raw.data());
      zs.avail_out = (uInt)usize;
      const int r = inflate(&zs, Z_FINISH);
      inflateEnd(&zs);
      if (r != Z_STREAM_END || zs.total_out != usize)
        return fdk::fail(FDK_ERR_INVALID, "%s: cannot inflate member %s", path, name.c_str());
    } else {
      return fdk::fail(FDK_ERR_UNSUPPORTED, "%s: compression method %d", path, method);
    }
    if (name.size() > 4 && name.compare(name.size() - 4, 4, ".npy") == 0) name.resize(name.size() - 4);
    if (int r = parse_npy(name, raw, out[name])) return r;
    off = d0 + csize;
  }
  if (out.empty()) return fdk::fail(FDK_ERR_INVALID, "%s: no arrays (not a .npz file?)", path);
  return FDK_OK;
}

int scalar_i(const std::map<std::string, NpyArray>& m, const char* key, int dflt) {
  auto it = m.find(key);
  if (it == m.end()) return dflt;
  const NpyArray& a = it->second;
  if (a.descr == "<i4") return *reinterpret_cast<const int32_t*>(a.data.data());
  if (a.descr == "<i8") return (int)*reinterpret_cast<const int64_t*>(a.data.data());
  return dflt;
}

// Architecture of a weight set: base channels, space-to-depth edge, activation.
int arch_of(const std::map<std::string, NpyArray>& m, int* base, int* s2d, int* act) {
  auto it = m.find("conv1.w");
  if (it == m.end() || it->second.shape.size() != 3)
    return fdk::fail(FDK_ERR_INVALID, "weights lack conv1.w [27][base][base]");
  *base = (int)it->second.shape[1];
  *s2d = scalar_i(m, "arch.s2d", 1);
  *act = scalar_i(m, "arch.act", 1);
  if (*s2d != 1 && *s2d != 2 && *s2d != 4) return fdk::fail(FDK_ERR_INVALID, "arch.s2d %d", *s2d);
  if (*act != 1 && *act != 2) return fdk::fail(FDK_ERR_INVALID, "arch.act %d", *act);
  if (*base % 16 != 0 || (*s2d > 1 && *base != 32))
    return fdk::fail(FDK_ERR_UNSUPPORTED, "base %d (a multiple of 16; 32 with space-to-depth)", *base);
  return FDK_OK;
}

struct Layer {
  int cin = 0, cout = 0;
  void* w = nullptr;     // BF16 [27][cout][cin], or [cout][32] for cin == 1
  float* b = nullptr;    // [cout]
};

}  // namespace

struct fdk_unet {
  int device = 0, base = 0, s2d = 1, act = 1;
  Layer conv[10];
  float* hw = nullptr;  // [f^3][base] (s2d > 1) or [base]
  float* hb = nullptr;  // [f^3] or [1]
};

namespace {

void release(fdk_unet* u) {
  for (Layer& l : u->conv) {
    if (l.w) (void)hipFree(l.w);
    if (l.b) (void)hipFree(l.b);
  }
  if (u->hw) (void)hipFree(u->hw);
  if (u->hb) (void)hipFree(u->hb);
  delete u;
}

int upload_f32(const NpyArray& a, const std::vector<int64_t>& shape, const char* name, float** dev) {
  if (a.descr != "<f4" || a.shape != shape) return fdk::fail(FDK_ERR_INVALID, "%s: unexpected dtype or shape", name);
  FDK_HIP(hipMalloc(dev, a.data.size()));
  FDK_HIP(hipMemcpy(*dev, a.data.data(), a.data.size(), hipMemcpyHostToDevice));
  return FDK_OK;
}

// Buffers of one forward pass, from the stream-ordered pool, each freed on
// the stream once its last reader is queued.
struct Pass {
  hipStream_t st;
  std::vector<void*> live;
  explicit Pass(hipStream_t s) : st(s) {}
  ~Pass() {
    for (void* p : live) (void)hipFreeAsync(p, st);
  }
  int get(size_t bytes, void** p) {
    if (int r = fdk::stage_alloc(p, bytes, st)) return r;
    live.push_back(*p);
    return FDK_OK;
  }
  void drop(void* p) {
    for (auto it = live.begin(); it != live.end(); ++it)
      if (*it == p) {
        (void)hipFreeAsync(p, st);
        live.erase(it);
        return;
      }
  }
};

int conv(const Layer& l, int act, const void* x, void* y, int n, int d, int h, int w, hipStream_t st) {
  return fdk_nn_conv3d_fwd(x, l.w, l.b, nullptr, y, n, d, h, w, l.cin, l.cout, act, 0, st);
}

// A decoder convolution on [upsample2(up), skip]: fused when the kernel serves
// the shape, else through a formed concatenation.
int conv_cat(Pass& P, const Layer& l, int act, const void* up, int cup, const void* skip, void* y,
             int n, int d, int h, int w) {
  const int cskip = l.cin - cup;
  if (l.cout % 32 == 0 && cup % 32 == 0 && cskip > 0 && l.cin % 32 == 0)
    return fdk_nn_conv3d_fwd_cat(up, cup, skip, l.w, l.b, nullptr, y, n, d, h, w, l.cin, l.cout,
                                 act, 0, P.st);
  void* cat = nullptr;
  const int64_t nvox = (int64_t)n * d * h * w;
  if (int r = P.get((size_t)nvox * l.cin * 2, &cat)) return r;
  if (int r = fdk_nn_upsample2(up, cat, n, d / 2, h / 2, w / 2, cup, l.cin, 0, 0, P.st)) return r;
  if (int r = fdk_nn_channel_copy(skip, cskip, 0, cat, l.cin, cup, cskip, nvox, P.st)) return r;
  const int r = conv(l, act, cat, y, n, d, h, w, P.st);
  P.drop(cat);
  return r;
}

int forward(fdk_unet* u, const float* x, float* out, int n, int D, int H, int W, hipStream_t st) {
  const int f = u->s2d;
  FDK_REQUIRE(n > 0 && D > 0 && H > 0 && W > 0, "bad volume size");
  FDK_REQUIRE(D % (4 * f) == 0 && H % (16 * f) == 0 && W % (128 * f) == 0,
              "U-Net input %d x %d x %d: needs D %% %d, H %% %d and W %% %d == 0", D, H, W, 4 * f,
              16 * f, 128 * f);
  const int c1 = u->base, c2 = 2 * c1, c3 = 4 * c1, act = u->act;
  const int d1 = D / f, h1 = H / f, w1 = W / f, d2 = d1 / 2, h2 = h1 / 2, w2 = w1 / 2;
  const int64_t v1 = (int64_t)n * d1 * h1 * w1, v2 = v1 / 8, v3 = v2 / 8;
  Pass P(st);
  void *xb, *e1a, *s1, *p1, *e2a, *s2, *p2, *ba, *bb, *d2a, *dd2, *d1a, *dd1;
  const int c0 = u->conv[0].cin;
  if (int r = P.get((size_t)v1 * c0 * 2, &xb)) return r;
  if (f > 1) {
    if (int r = fdk_nn_unshuffle(x, xb, n, D, H, W, f, c0, st)) return r;
  } else if (int r = fdk_nn_convert(x, xb, v1, 1, st)) {
    return r;
  }
  if (int r = P.get((size_t)v1 * c1 * 2, &e1a)) return r;
  if (int r = conv(u->conv[0], act, xb, e1a, n, d1, h1, w1, st)) return r;
  P.drop(xb);
  if (int r = P.get((size_t)v1 * c1 * 2, &s1)) return r;
  if (int r = conv(u->conv[1], act, e1a, s1, n, d1, h1, w1, st)) return r;
  P.drop(e1a);
  if (int r = P.get((size_t)v2 * c1 * 2, &p1)) return r;
  if (int r = fdk_nn_avgpool2(s1, p1, n, d1, h1, w1, c1, 0, st)) return r;
  if (int r = P.get((size_t)v2 * c2 * 2, &e2a)) return r;
  if (int r = conv(u->conv[2], act, p1, e2a, n, d2, h2, w2, st)) return r;
  P.drop(p1);
  if (int r = P.get((size_t)v2 * c2 * 2, &s2)) return r;
  if (int r = conv(u->conv[3], act, e2a, s2, n, d2, h2, w2, st)) return r;
  P.drop(e2a);
  if (int r = P.get((size_t)v3 * c2 * 2, &p2)) return r;
  if (int r = fdk_nn_avgpool2(s2, p2, n, d2, h2, w2, c2, 0, st)) return r;
  if (int r = P.get((size_t)v3 * c3 * 2, &ba)) return r;
  if (int r = conv(u->conv[4], act, p2, ba, n, d2 / 2, h2 / 2, w2 / 2, st)) return r;
  P.drop(p2);
  if (int r = P.get((size_t)v3 * c3 * 2, &bb)) return r;
  if (int r = conv(u->conv[5], act, ba, bb, n, d2 / 2, h2 / 2, w2 / 2, st)) return r;
  P.drop(ba);
  if (int r = P.get((size_t)v2 * c2 * 2, &d2a)) return r;
  if (int r = conv_cat(P, u->conv[6], act, bb, c3, s2, d2a, n, d2, h2, w2)) return r;
  P.drop(bb);
  P.drop(s2);
  if (int r = P.get((size_t)v2 * c2 * 2, &dd2)) return r;
  if (int r = conv(u->conv[7], act, d2a, dd2, n, d2, h2, w2, st)) return r;
  P.drop(d2a);
  if (int r = P.get((size_t)v1 * c1 * 2, &d1a)) return r;
  if (int r = conv_cat(P, u->conv[8], act, dd2, c2, s1, d1a, n, d1, h1, w1)) return r;
  P.drop(dd2);
  P.drop(s1);
  if (int r = P.get((size_t)v1 * c1 * 2, &dd1)) return r;
  if (int r = conv(u->conv[9], act, d1a, dd1, n, d1, h1, w1, st)) return r;
  P.drop(d1a);
  if (f > 1) return fdk_nn_head_shuffle(dd1, u->hw, u->hb, x, out, n, D, H, W, f, c1, st);
  return fdk_nn_head(dd1, u->hw, u->hb, x, out, v1, c1, st);
}

}  // namespace

extern "C" int fdk_unet_inspect(const char* path, int* base, int* s2d, int* act) {
  FDK_REQUIRE(path && base && s2d && act, "NULL pointer");
  std::map<std::string, NpyArray> m;
  if (int r = read_npz(path, m)) return r;
  return arch_of(m, base, s2d, act);
}

extern "C" int fdk_unet_load(const char* path, fdk_unet** out) {
  FDK_REQUIRE(path && out, "NULL pointer");
  *out = nullptr;
  std::map<std::string, NpyArray> m;
  if (int r = read_npz(path, m)) return r;
  int base, s2d, act;
  if (int r = arch_of(m, &base, &s2d, &act)) return r;
  fdk_unet* u = new fdk_unet;
  u->base = base;
  u->s2d = s2d;
  u->act = act;
  int r = FDK_OK;
  if (hipGetDevice(&u->device) != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "no current device");
  const int c1 = base, c2 = 2 * base, c3 = 4 * base, k = s2d * s2d * s2d;
  const int cin0 = s2d > 1 ? std::max(32, k) : 1;
  const int io[10][2] = {{cin0, c1}, {c1, c1}, {c1, c2}, {c2, c2}, {c2, c3},
                         {c3, c3}, {c3 + c2, c2}, {c2, c2}, {c2 + c1, c1}, {c1, c1}};
  float* tmp = nullptr;
  for (int i = 0; i < 10 && r == FDK_OK; ++i) {
    Layer& l = u->conv[i];
    l.cin = io[i][0];
    l.cout = io[i][1];
    const std::string wn = "conv" + std::to_string(i) + ".w", bn = "conv" + std::to_string(i) + ".b";
    auto wi = m.find(wn), bi = m.find(bn);
    if (wi == m.end() || bi == m.end()) {
      r = fdk::fail(FDK_ERR_INVALID, "%s: missing %s or %s", path, wn.c_str(), bn.c_str());
      break;
    }
    const std::vector<int64_t> ws = l.cin == 1 ? std::vector<int64_t>{l.cout, 32}
                                               : std::vector<int64_t>{27, l.cout, l.cin};
    r = upload_f32(wi->second, ws, wn.c_str(), &tmp);
    if (r == FDK_OK) {
      const int64_t nw = wi->second.count();
      hipError_t e = hipMalloc(&l.w, nw * 2);
      if (e == hipSuccess) {
        r = fdk_nn_convert(tmp, l.w, nw, 1, nullptr);  // round to nearest even, as the trainer
        if (r == FDK_OK) e = hipDeviceSynchronize();
      }
      if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "%s upload: %s", wn.c_str(), hipGetErrorString(e));
    }
    if (tmp) (void)hipFree(tmp);
    tmp = nullptr;
    if (r == FDK_OK) r = upload_f32(bi->second, {l.cout}, bn.c_str(), &l.b);
  }
  if (r == FDK_OK) {
    auto hw = m.find("head.w"), hb = m.find("head.b");
    if (hw == m.end() || hb == m.end()) {
      r = fdk::fail(FDK_ERR_INVALID, "%s: missing head.w or head.b", path);
    } else if (s2d > 1) {
      r = upload_f32(hw->second, {k, c1}, "head.w", &u->hw);
      if (r == FDK_OK) r = upload_f32(hb->second, {k}, "head.b", &u->hb);
    } else {
      r = upload_f32(hw->second, {c1}, "head.w", &u->hw);
      if (r == FDK_OK) r = upload_f32(hb->second, {1}, "head.b", &u->hb);
    }
  }
  if (r != FDK_OK) {
    release(u);
    return r;
  }
  *out = u;
  return FDK_OK;
}

extern "C" int fdk_unet_info(const fdk_unet* u, int* base, int* s2d, int* act) {
  FDK_REQUIRE(u && base && s2d && act, "NULL pointer");
  *base = u->base;
  *s2d = u->s2d;
  *act = u->act;
  return FDK_OK;
}

extern "C" int fdk_unet_forward(fdk_unet* u, const float* dev_x, float* dev_out, int n, int d,
                                int h, int w, void* stream) {
  FDK_REQUIRE(u && dev_x && dev_out, "NULL pointer");
  return forward(u, dev_x, dev_out, n, d, h, w, static_cast<hipStream_t>(stream));
}

extern "C" int fdk_unet_infer(fdk_unet* u, const float* dev_vol, float* dev_out, int d, int h,
                              int w, int slab, int overlap, void* stream) {
  FDK_REQUIRE(u && dev_vol && dev_out, "NULL pointer");
  FDK_REQUIRE(dev_vol != dev_out, "in-place inference is not supported");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (slab <= 0) slab = u->s2d > 1 ? d : 528;
  if (slab >= d) return forward(u, dev_vol, dev_out, 1, d, h, w, st);
  FDK_REQUIRE(overlap >= 0 && overlap < slab, "overlap %d not in [0, slab %d)", overlap, slab);
  // slab origins every slab - overlap planes (the last one flush with the end);
  // neighbours split their overlap at its midpoint
  std::vector<int> starts;
  for (int s = 0; s <= d - slab; s += slab - overlap) starts.push_back(s);
  if (starts.back() + slab < d) starts.push_back(d - slab);
  const size_t plane = (size_t)h * w;
  float* tmp = nullptr;
  if (int r = fdk::stage_alloc((void**)&tmp, plane * slab * sizeof(float), st)) return r;
  int r = FDK_OK;
  for (size_t k = 0; k < starts.size() && r == FDK_OK; ++k) {
    const int s = starts[k];
    const int z0 = k == 0 ? 0 : (starts[k] + starts[k - 1] + slab) / 2;
    const int z1 = k + 1 == starts.size() ? d : (starts[k + 1] + starts[k] + slab) / 2;
    r = forward(u, dev_vol + plane * s, tmp, 1, slab, h, w, st);
    if (r == FDK_OK) {
      const hipError_t e = hipMemcpyAsync(dev_out + plane * z0, tmp + plane * (z0 - s),
                                          plane * (z1 - z0) * sizeof(float), hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) r = fdk::fail(FDK_ERR_HIP, "slab copy: %s", hipGetErrorString(e));
    }
  }
  (void)hipFreeAsync(tmp, st);
  return r;
}

extern "C" int fdk_unet_destroy(fdk_unet* u) {
  if (!u) return FDK_OK;
  (void)hipDeviceSynchronize();  // no pass may still read the weights
  release(u);
  return FDK_OK;
}
