// Weight and map I/O: .npz archives (numpy.savez / numpy.savez_compressed) read without a third-party loader, for
//   mppi_npz_read            -- one array, converted to FP32 in C order;
//   mppi_fnn_load_npz        -- FNNHelper::loadParams (utils/nn_helpers/fnn_helper.cu:44-127): the network's layer
//                               sizes from dynamics_W1 / dynamics_b1, dynamics_b2, ... and its parameter vector in the
//                               engine's FNN layout (per layer W [out][in] row-major, then b [out]);
//   mppi_track_map_load_npz  -- ARStandardCost::loadTrackData (cost_functions/autorally/ar_standard_cost.cu:85-142):
//                               xBounds, yBounds, pixelsPerMeter and a channel as a cost map of the quadratic cost.
// An .npz is a zip archive of .npy members: stored or deflate-compressed (inflated with zlib), with zip64 extra fields
// (numpy writes them). Host code only.
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "host.hpp"

namespace mppi_amd {
namespace {

struct ZipEntry {
  int method = 0;
  uint64_t comp = 0, size = 0, local = 0;
};

uint16_t rd16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
uint64_t rd64(const uint8_t* p) { return (uint64_t)rd32(p) | ((uint64_t)rd32(p + 4) << 32); }

int read_file(const char* path, std::vector<uint8_t>& out)
{
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(MPPI_E_INVALID, std::string("npz: cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(n > 0 ? (size_t)n : 0);
  const size_t got = n > 0 ? std::fread(out.data(), 1, (size_t)n, f) : 0;
  std::fclose(f);
  if (n < 0 || got != (size_t)n) return fail(MPPI_E_INVALID, std::string("npz: cannot read ") + path);
  return MPPI_OK;
}

// The central directory: member name -> (method, sizes, local header offset), zip64 fields resolved.
int zip_directory(const std::vector<uint8_t>& z, std::map<std::string, ZipEntry>& dir)
{
  const size_t n = z.size();
  if (n < 22) return fail(MPPI_E_INVALID, "npz: not a zip archive");
  // end of central directory: the last occurrence of its signature (a comment may follow it)
  size_t eocd = std::string::npos;
  for (size_t i = n - 22 + 1; i-- > 0 && n - i <= 22 + 65535;)
    if (rd32(&z[i]) == 0x06054b50u) {
      eocd = i;
      break;
    }
  if (eocd == std::string::npos) return fail(MPPI_E_INVALID, "npz: no end-of-central-directory record");
  uint64_t count = rd16(&z[eocd + 10]), cd_size = rd32(&z[eocd + 12]), cd_off = rd32(&z[eocd + 16]);
  if ((count == 0xFFFF || cd_off == 0xFFFFFFFFu) && eocd >= 20 && rd32(&z[eocd - 20]) == 0x07064b50u) {
    const uint64_t e64 = rd64(&z[eocd - 20 + 8]);
    if (e64 + 56 > n || rd32(&z[e64]) != 0x06064b50u) return fail(MPPI_E_INVALID, "npz: bad zip64 directory record");
    count = rd64(&z[e64 + 32]);
    cd_size = rd64(&z[e64 + 40]);
    cd_off = rd64(&z[e64 + 48]);
  }
  if (cd_off + cd_size > n) return fail(MPPI_E_INVALID, "npz: central directory out of range");
  size_t p = cd_off;
  for (uint64_t i = 0; i < count; ++i) {
    if (p + 46 > n || rd32(&z[p]) != 0x02014b50u) return fail(MPPI_E_INVALID, "npz: bad central directory entry");
    ZipEntry e;
    e.method = rd16(&z[p + 10]);
    e.comp = rd32(&z[p + 20]);
    e.size = rd32(&z[p + 24]);
    const int name_len = rd16(&z[p + 28]), extra_len = rd16(&z[p + 30]), comment_len = rd16(&z[p + 32]);
    e.local = rd32(&z[p + 42]);
    if (p + 46 + (size_t)name_len + extra_len + comment_len > n) return fail(MPPI_E_INVALID, "npz: truncated entry");
    const std::string name(reinterpret_cast<const char*>(&z[p + 46]), name_len);
    // zip64 extended information: the 64-bit values of the fields saturated above, in their order
    size_t x = p + 46 + name_len;
    const size_t xend = x + extra_len;
    while (x + 4 <= xend) {
      const int id = rd16(&z[x]), len = rd16(&z[x + 2]);
      if (id == 0x0001) {
        size_t q = x + 4;
        if (e.size == 0xFFFFFFFFu && q + 8 <= x + 4 + len) e.size = rd64(&z[q]), q += 8;
        if (e.comp == 0xFFFFFFFFu && q + 8 <= x + 4 + len) e.comp = rd64(&z[q]), q += 8;
        if (e.local == 0xFFFFFFFFu && q + 8 <= x + 4 + len) e.local = rd64(&z[q]);
      }
      x += 4 + len;
    }
    dir[name] = e;
    p += 46 + name_len + extra_len + comment_len;
  }
  return MPPI_OK;
}

int zip_extract(const std::vector<uint8_t>& z, const ZipEntry& e, std::vector<uint8_t>& out)
{
  if (e.local + 30 > z.size() || rd32(&z[e.local]) != 0x04034b50u) return fail(MPPI_E_INVALID, "npz: bad local header");
  const uint64_t data = e.local + 30 + rd16(&z[e.local + 26]) + rd16(&z[e.local + 28]);
  if (data + e.comp > z.size()) return fail(MPPI_E_INVALID, "npz: member data out of range");
  out.resize(e.size);
  if (e.method == 0) {
    if (e.comp != e.size) return fail(MPPI_E_INVALID, "npz: stored member with differing sizes");
    std::memcpy(out.data(), &z[data], e.size);
    return MPPI_OK;
  }
  if (e.method != 8) return fail(MPPI_E_UNSUPPORTED, "npz: compression method " + std::to_string(e.method));
  z_stream s;
  std::memset(&s, 0, sizeof(s));
  if (inflateInit2(&s, -MAX_WBITS) != Z_OK) return fail(MPPI_E_INVALID, "npz: inflateInit2 failed");
  // inflated in pieces: zlib's counters are 32-bit
  uint64_t in_done = 0, out_done = 0;
  int rc = Z_OK;
  while (rc == Z_OK) {
    const uInt ci = (uInt)std::min<uint64_t>(e.comp - in_done, 1u << 30);
    const uInt co = (uInt)std::min<uint64_t>(e.size - out_done, 1u << 30);
    s.next_in = const_cast<Bytef*>(&z[data + in_done]);
    s.avail_in = ci;
    s.next_out = out.data() + out_done;
    s.avail_out = co;
    rc = inflate(&s, Z_NO_FLUSH);
    in_done += ci - s.avail_in;
    out_done += co - s.avail_out;
    if (rc == Z_BUF_ERROR && (ci == 0 || co == 0)) break;
  }
  inflateEnd(&s);
  if (rc != Z_STREAM_END || out_done != e.size) return fail(MPPI_E_INVALID, "npz: corrupt deflate member");
  return MPPI_OK;
}

struct NpyArray {
  std::vector<int64_t> shape;
  std::vector<float> data;  // C order
};

// The text value after `key` in a .npy header dictionary ({'descr': '<f8', 'fortran_order': False, 'shape': (2, 3), }).
std::string npy_field(const std::string& h, const std::string& key)
{
  const size_t k = h.find("'" + key + "'");
  if (k == std::string::npos) return "";
  size_t v = h.find(':', k);
  if (v == std::string::npos) return "";
  ++v;
  while (v < h.size() && h[v] == ' ') ++v;
  size_t e = v;
  if (h[v] == '(') e = h.find(')', v) + 1;
  else if (h[v] == '\'') e = h.find('\'', v + 1) + 1;
  else
    while (e < h.size() && h[e] != ',' && h[e] != '}') ++e;
  return h.substr(v, e - v);
}

int npy_parse(const std::vector<uint8_t>& b, NpyArray& a)
{
  static const uint8_t magic[6] = {0x93, 'N', 'U', 'M', 'P', 'Y'};
  if (b.size() < 10 || std::memcmp(b.data(), magic, 6) != 0) return fail(MPPI_E_INVALID, "npy: bad magic");
  const int major = b[6];
  size_t hlen, off;
  if (major == 1) {
    hlen = rd16(&b[8]);
    off = 10;
  } else if (major == 2 || major == 3) {
    if (b.size() < 12) return fail(MPPI_E_INVALID, "npy: truncated header");
    hlen = rd32(&b[8]);
    off = 12;
  } else {
    return fail(MPPI_E_UNSUPPORTED, "npy: format version " + std::to_string(major));
  }
  if (off + hlen > b.size()) return fail(MPPI_E_INVALID, "npy: truncated header");
  const std::string h(reinterpret_cast<const char*>(&b[off]), hlen);
  const std::string descr = npy_field(h, "descr"), fortran = npy_field(h, "fortran_order"),
                    shape = npy_field(h, "shape");
  if (descr.size() < 4 || shape.empty()) return fail(MPPI_E_INVALID, "npy: header without descr or shape");
  const char order = descr[1], kind = descr[2];
  const int width = std::atoi(descr.substr(3, descr.size() - 4).c_str());
  if (order == '>' && width > 1) return fail(MPPI_E_UNSUPPORTED, "npy: big-endian data");
  if (!((kind == 'f' && (width == 4 || width == 8)) || (kind == 'i' && (width == 4 || width == 8))))
    return fail(MPPI_E_UNSUPPORTED, "npy: dtype " + descr + " (float32/64 and int32/64 are read)");
  a.shape.clear();
  int64_t count = 1;
  for (size_t i = 1; i < shape.size();) {
    while (i < shape.size() && (shape[i] == ' ' || shape[i] == ',')) ++i;
    if (i >= shape.size() || shape[i] == ')') break;
    const int64_t d = std::atoll(&shape[i]);
    a.shape.push_back(d);
    count *= d;
    while (i < shape.size() && shape[i] >= '0' && shape[i] <= '9') ++i;
  }
  const size_t data_off = off + hlen;
  if (count < 0 || data_off + (size_t)count * width > b.size()) return fail(MPPI_E_INVALID, "npy: data out of range");
  std::vector<float> raw((size_t)count);
  const uint8_t* d = &b[data_off];
  for (int64_t i = 0; i < count; ++i) {
    const uint8_t* q = d + (size_t)i * width;
    if (kind == 'f' && width == 4) {
      float v;
      std::memcpy(&v, q, 4);
      raw[i] = v;
    } else if (kind == 'f') {
      double v;
      std::memcpy(&v, q, 8);
      raw[i] = (float)v;
    } else if (width == 4) {
      int32_t v;
      std::memcpy(&v, q, 4);
      raw[i] = (float)v;
    } else {
      int64_t v;
      std::memcpy(&v, q, 8);
      raw[i] = (float)v;
    }
  }
  // Fortran-ordered arrays are brought to C order (first index slowest)
  if (fortran.rfind("True", 0) == 0 && a.shape.size() >= 2) {
    const size_t nd = a.shape.size();
    std::vector<int64_t> cstride(nd, 1), fstride(nd, 1);
    for (size_t k = nd - 1; k-- > 0;) cstride[k] = cstride[k + 1] * a.shape[k + 1];
    for (size_t k = 1; k < nd; ++k) fstride[k] = fstride[k - 1] * a.shape[k - 1];
    a.data.resize(raw.size());
    for (int64_t i = 0; i < count; ++i) {
      int64_t rem = i, fi = 0;
      for (size_t k = 0; k < nd; ++k) {
        const int64_t ix = rem / cstride[k];
        rem -= ix * cstride[k];
        fi += ix * fstride[k];
      }
      a.data[(size_t)i] = raw[(size_t)fi];
    }
  } else {
    a.data = std::move(raw);
  }
  return MPPI_OK;
}

// An opened archive: its bytes and directory; arrays are extracted on request.
struct Npz {
  std::vector<uint8_t> bytes;
  std::map<std::string, ZipEntry> dir;
  int open(const char* path)
  {
    if (!path) return fail(MPPI_E_INVALID, "npz: null path");
    MPPI_CHECK(read_file(path, bytes));
    return zip_directory(bytes, dir);
  }
  bool has(const std::string& key) const { return dir.count(key + ".npy") != 0; }
  int get(const std::string& key, NpyArray& a) const
  {
    auto it = dir.find(key + ".npy");
    if (it == dir.end()) return fail(MPPI_E_INVALID, "npz: no array '" + key + "'");
    std::vector<uint8_t> member;
    MPPI_CHECK(zip_extract(bytes, it->second, member));
    return npy_parse(member, a);
  }
};

}  // namespace
}  // namespace mppi_amd

using namespace mppi_amd;

extern "C" {

int mppi_npz_read(const char* path, const char* key, float* out, int64_t max_count, int32_t* ndim, int64_t* shape,
                  int64_t* count)
{
  if (!key) return fail(MPPI_E_INVALID, "npz: null key");
  Npz z;
  MPPI_CHECK(z.open(path));
  NpyArray a;
  MPPI_CHECK(z.get(key, a));
  if (a.shape.size() > 4) return fail(MPPI_E_UNSUPPORTED, "npz: more than 4 dimensions");
  if (ndim) *ndim = (int32_t)a.shape.size();
  if (shape)
    for (size_t i = 0; i < 4; ++i) shape[i] = i < a.shape.size() ? a.shape[i] : 0;
  if (count) *count = (int64_t)a.data.size();
  if (out) {
    if ((int64_t)a.data.size() > max_count) return fail(MPPI_E_SIZE, "npz: output buffer too small");
    std::memcpy(out, a.data.data(), a.data.size() * sizeof(float));
  }
  return MPPI_OK;
}

int mppi_fnn_load_npz(const char* path, const char* prefix, int32_t max_layers, int32_t* num_layers, int32_t* layers,
                      float* theta, int32_t max_params, int32_t* num_params)
{
  Npz z;
  MPPI_CHECK(z.open(path));
  // loadParams(prefix, dict, add_slash = true): a nonempty prefix is followed by '/'
  std::string pre = prefix ? prefix : "";
  if (!pre.empty() && pre.back() != '/') pre.push_back('/');
  auto wname = [&](int i) { return pre + "dynamics_W" + std::to_string(i); };
  auto bname = [&](int i) { return pre + "dynamics_b" + std::to_string(i); };
  NpyArray w1, b1;
  MPPI_CHECK(z.get(wname(1), w1));
  MPPI_CHECK(z.get(bname(1), b1));
  if (b1.data.empty() || w1.data.size() % b1.data.size() != 0)
    return fail(MPPI_E_INVALID, "fnn npz: dynamics_W1 is not a multiple of dynamics_b1");
  std::vector<int32_t> ls = {(int32_t)(w1.data.size() / b1.data.size())};
  for (int i = 1; z.has(bname(i)); ++i) {
    NpyArray b;
    MPPI_CHECK(z.get(bname(i), b));
    ls.push_back((int32_t)b.data.size());
  }
  int64_t total = 0;
  for (size_t l = 0; l + 1 < ls.size(); ++l) total += (int64_t)ls[l] * ls[l + 1] + ls[l + 1];
  if (num_layers) *num_layers = (int32_t)ls.size();
  if (num_params) *num_params = (int32_t)total;
  if (layers) {
    if ((int32_t)ls.size() > max_layers) return fail(MPPI_E_SIZE, "fnn npz: more layers than max_layers");
    std::copy(ls.begin(), ls.end(), layers);
  }
  if (!theta) return MPPI_OK;
  if (total > max_params) return fail(MPPI_E_SIZE, "fnn npz: more parameters than max_params");
  int64_t off = 0;
  for (size_t l = 0; l + 1 < ls.size(); ++l) {
    NpyArray w, b;
    MPPI_CHECK(z.get(wname((int)l + 1), w));
    MPPI_CHECK(z.get(bname((int)l + 1), b));
    const int64_t nw = (int64_t)ls[l] * ls[l + 1];
    if ((int64_t)w.data.size() != nw || (int64_t)b.data.size() != ls[l + 1])
      return fail(MPPI_E_INVALID, "fnn npz: layer " + std::to_string(l + 1) + " sizes do not chain");
    for (float v : w.data)
      if (!std::isfinite(v)) return fail(MPPI_E_INVALID, "fnn npz: non-finite weight");
    for (float v : b.data)
      if (!std::isfinite(v)) return fail(MPPI_E_INVALID, "fnn npz: non-finite bias");
    std::memcpy(theta + off, w.data.data(), sizeof(float) * nw);
    std::memcpy(theta + off + nw, b.data.data(), sizeof(float) * ls[l + 1]);
    off += nw + ls[l + 1];
  }
  return MPPI_OK;
}

int mppi_track_map_load_npz(const char* path, const char* channel, float* values, int64_t max_count, int32_t* width,
                            int32_t* height, float* origin_x, float* origin_y, float* resolution)
{
  Npz z;
  MPPI_CHECK(z.open(path));
  NpyArray xb, yb, ppm, ch;
  MPPI_CHECK(z.get("xBounds", xb));
  MPPI_CHECK(z.get("yBounds", yb));
  MPPI_CHECK(z.get("pixelsPerMeter", ppm));
  if (xb.data.size() < 2 || yb.data.size() < 2 || ppm.data.empty() || !(ppm.data[0] > 0.0f))
    return fail(MPPI_E_INVALID, "track map: bad xBounds / yBounds / pixelsPerMeter");
  // the texture is (x_max - x_min) ppm texels wide and (y_max - y_min) ppm high, texel size 1 / ppm, origin at
  // (x_min, y_min) (the transform R, trs of ar_standard_cost.cu:131-138 in texel terms)
  const int w = (int)((xb.data[1] - xb.data[0]) * ppm.data[0]);
  const int h = (int)((yb.data[1] - yb.data[0]) * ppm.data[0]);
  if (w < 2 || h < 2) return fail(MPPI_E_INVALID, "track map: fewer than 2 x 2 texels");
  if (width) *width = w;
  if (height) *height = h;
  if (origin_x) *origin_x = xb.data[0];
  if (origin_y) *origin_y = yb.data[0];
  if (resolution) *resolution = 1.0f / ppm.data[0];
  if (!values) return MPPI_OK;
  MPPI_CHECK(z.get(channel ? channel : "channel0", ch));
  // the channel's values in their stored order make the [height][width] texel rows, as the texture upload takes them
  if ((int64_t)ch.data.size() < (int64_t)w * h) return fail(MPPI_E_INVALID, "track map: channel smaller than the map");
  if ((int64_t)w * h > max_count) return fail(MPPI_E_SIZE, "track map: output buffer too small");
  std::memcpy(values, ch.data.data(), sizeof(float) * (size_t)w * h);
  return MPPI_OK;
}

}  // extern "C"
