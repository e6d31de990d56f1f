// ncio.hip -- NetCDF classic files with fixed-size variables: the file format of
// Coupler::write_output_file and Coupler::overwrite_with_restart (model/core/coupler.h:533-723),
// which the reference writes through PnetCDF with NC_64BIT_DATA (CDF-5).
//
// Format (the published NetCDF classic / CDF-5 specification): magic "CDF" + version (1 classic,
// 2 64-bit offset, 5 64-bit data); numrecs; dimension list; global attributes (none here);
// variable list (name, dimension ids, attributes (none), type, vsize, begin); then each variable's
// values, big-endian, at its `begin`.  Integer fields of the header are 32-bit, except that CDF-5
// widens every count and size to 64 bits and CDF-2/5 widen `begin` to 64 bits.  No record
// dimension is used (the reference's "t" is a fixed dimension of length 1).
//
// The header is given a reserved region (1 MiB, the reference's nc_header_align_size) so that
// variables defined after the first enddef (the module output hooks do redef / create / enddef)
// only rewrite the header; every variable starts on a 4 KiB boundary after the previous one.
//
// Device data are encoded on the GPU: a kernel converts a chunk of doubles to the variable's type
// and byte-swaps it to big-endian into a device staging buffer, which goes to pinned host memory
// and then to the file; two staging buffers alternate so that the GPU encodes and copies chunk
// i+1 while the host writes chunk i.  Reading runs the same pipeline backwards.
#include "pu_common.h"

#include <fcntl.h>
#include <unistd.h>
#include <sys/stat.h>

#include <cstring>
#include <string>
#include <vector>

namespace pu {

constexpr long long NC_HEADER_RESERVE = 1 << 20, NC_VAR_ALIGN = 4096;
constexpr size_t NC_CHUNK = (size_t)32 << 20;     // staging chunk, bytes of encoded data

int nc_type_size(int t) {
  switch (t) {
    case PU_NC_BYTE: case PU_NC_CHAR: case PU_NC_UBYTE: return 1;
    case PU_NC_SHORT: return 2;
    case PU_NC_INT: case PU_NC_FLOAT: return 4;
    case PU_NC_DOUBLE: case PU_NC_INT64: return 8;
    default: return 0;
  }
}

template <class T> __device__ __forceinline__ void store_be(unsigned char *dst, long long i, T v);
template <> __device__ __forceinline__ void store_be<float>(unsigned char *dst, long long i, float v) {
  reinterpret_cast<unsigned int *>(dst)[i] = __builtin_bswap32(__builtin_bit_cast(unsigned int, v));
}
template <> __device__ __forceinline__ void store_be<double>(unsigned char *dst, long long i, double v) {
  reinterpret_cast<unsigned long long *>(dst)[i] = __builtin_bswap64(__builtin_bit_cast(unsigned long long, v));
}
template <> __device__ __forceinline__ void store_be<int>(unsigned char *dst, long long i, int v) {
  reinterpret_cast<unsigned int *>(dst)[i] = __builtin_bswap32((unsigned int)v);
}

// doubles -> big-endian values of the variable's type (the reference's .as<float>() is a
// round-to-nearest conversion; integers truncate like a C cast)
template <class T>
__global__ __launch_bounds__(256) void k_nc_encode(const double *__restrict__ src, long long n, unsigned char *dst) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  store_be<T>(dst, i, (T)src[i]);
}

template <class T>
__global__ __launch_bounds__(256) void k_nc_decode(const unsigned char *__restrict__ src, long long n, double *dst) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double v;
  if constexpr (sizeof(T) == 4) {
    const unsigned int b = __builtin_bswap32(reinterpret_cast<const unsigned int *>(src)[i]);
    if constexpr (std::is_same<T, float>::value) v = (double)__builtin_bit_cast(float, b);
    else v = (double)(int)b;
  } else {
    v = __builtin_bit_cast(double, __builtin_bswap64(reinterpret_cast<const unsigned long long *>(src)[i]));
  }
  dst[i] = v;
}

}  // namespace pu

using namespace pu;

struct pu_ncfile {
  struct Dim { std::string name; long long len; };
  struct Var {
    std::string name;
    int type = 0;
    std::vector<int> dimids;
    long long nelems = 1, nbytes = 0, begin = 0;
  };
  int fd = -1;
  bool writable = false, in_define = true;
  int version = 2;               // 1, 2 or 5
  bool version_fixed = false;    // set by the caller (PU_NC_CDF2 / PU_NC_CDF5) or read from a file
  std::vector<Dim> dims;
  std::vector<Var> vars;
  long long data_end = NC_HEADER_RESERVE;   // first free byte after the variables placed so far
  // staging for device transfers
  unsigned char *d_stage[2] = {nullptr, nullptr};
  unsigned char *h_stage[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace {

int find_var(const pu_ncfile *f, const char *name) {
  for (size_t i = 0; i < f->vars.size(); i++) if (f->vars[i].name == name) return (int)i;
  return -1;
}
int find_dim(const pu_ncfile *f, const char *name) {
  for (size_t i = 0; i < f->dims.size(); i++) if (f->dims[i].name == name) return (int)i;
  return -1;
}

// ---- header encoding --------------------------------------------------------------------------
struct Out {
  std::vector<unsigned char> b;
  int version;
  void u32(unsigned long long v) { for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
  void u64(unsigned long long v) { for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
  void count(unsigned long long v) { version == 5 ? u64(v) : u32(v); }          // NON_NEG
  void offset(unsigned long long v) { version == 1 ? u32(v) : u64(v); }         // OFFSET
  void name(std::string const &s) {
    count(s.size());
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
  }
  void absent() { u32(0); count(0); }
};

std::vector<unsigned char> encode_header(const pu_ncfile *f) {
  Out o;
  o.version = f->version;
  o.b = {'C', 'D', 'F', (unsigned char)f->version};
  o.count(0);                                   // numrecs: no record dimension
  if (f->dims.empty()) o.absent();
  else {
    o.u32(0x0A);                                // NC_DIMENSION
    o.count(f->dims.size());
    for (auto &d : f->dims) { o.name(d.name); o.count((unsigned long long)d.len); }
  }
  o.absent();                                   // no global attributes
  if (f->vars.empty()) o.absent();
  else {
    o.u32(0x0B);                                // NC_VARIABLE
    o.count(f->vars.size());
    for (auto &v : f->vars) {
      o.name(v.name);
      o.count(v.dimids.size());
      for (int id : v.dimids) o.count((unsigned long long)id);
      o.absent();                               // no variable attributes
      o.u32((unsigned long long)v.type);
      unsigned long long vsize = (unsigned long long)((v.nbytes + 3) / 4 * 4);
      if (f->version != 5 && vsize > 0xFFFFFFFCull) vsize = 0xFFFFFFFFull;
      o.count(vsize);
      o.offset((unsigned long long)v.begin);
    }
  }
  return o.b;
}

// ---- header decoding --------------------------------------------------------------------------
struct In {
  const unsigned char *p;
  size_t n, i = 0;
  int version = 1;
  bool ok = true;
  unsigned long long u32() {
    if (i + 4 > n) { ok = false; return 0; }
    unsigned long long v = 0;
    for (int k = 0; k < 4; k++) v = (v << 8) | p[i++];
    return v;
  }
  unsigned long long u64() {
    if (i + 8 > n) { ok = false; return 0; }
    unsigned long long v = 0;
    for (int k = 0; k < 8; k++) v = (v << 8) | p[i++];
    return v;
  }
  unsigned long long count() { return version == 5 ? u64() : u32(); }
  unsigned long long offset() { return version == 1 ? u32() : u64(); }
  std::string name() {
    unsigned long long len = count();
    if (!ok || i + len > n) { ok = false; return ""; }
    std::string s((const char *)p + i, (size_t)len);
    i += (size_t)((len + 3) / 4 * 4);
    return s;
  }
  // attribute list: skipped (name, type, nelems, values padded to 4 bytes)
  void skip_atts() {
    unsigned long long tag = u32(), cnt = count();
    if (tag == 0) return;
    for (unsigned long long a = 0; a < cnt && ok; a++) {
      name();
      int t = (int)u32();
      unsigned long long ne = count();
      size_t sz = (size_t)ne * nc_type_size(t);
      i += (sz + 3) / 4 * 4;
      if (i > n) ok = false;
    }
  }
};

int decode_header(pu_ncfile *f, const std::vector<unsigned char> &b) {
  In in{b.data(), b.size()};
  if (b.size() < 4 || b[0] != 'C' || b[1] != 'D' || b[2] != 'F' || (b[3] != 1 && b[3] != 2 && b[3] != 5))
    return fail(PU_ERR_ARG, "not a NetCDF classic (CDF-1/2/5) file");
  in.i = 4;
  in.version = b[3];
  f->version = b[3];
  unsigned long long numrecs = in.count();
  unsigned long long tag = in.u32(), cnt = in.count();
  if (tag == 0x0A) {
    for (unsigned long long d = 0; d < cnt && in.ok; d++) {
      std::string nm = in.name();
      long long len = (long long)in.count();
      if (len == 0) return fail(PU_ERR_UNSUPPORTED, "record dimensions are not supported (" + nm + ")");
      f->dims.push_back({nm, len});
    }
  }
  (void)numrecs;
  in.skip_atts();
  tag = in.u32();
  cnt = in.count();
  if (tag == 0x0B) {
    for (unsigned long long v = 0; v < cnt && in.ok; v++) {
      pu_ncfile::Var var;
      var.name = in.name();
      unsigned long long nd = in.count();
      for (unsigned long long d = 0; d < nd && in.ok; d++) {
        int id = (int)in.count();
        if (id < 0 || id >= (int)f->dims.size()) return fail(PU_ERR_ARG, "bad dimension id in " + var.name);
        var.dimids.push_back(id);
        var.nelems *= f->dims[id].len;
      }
      in.skip_atts();
      var.type = (int)in.u32();
      in.count();                                       // vsize (recomputed)
      var.begin = (long long)in.offset();
      if (nc_type_size(var.type) == 0) return fail(PU_ERR_UNSUPPORTED, "unsupported type of " + var.name);
      var.nbytes = var.nelems * nc_type_size(var.type);
      f->vars.push_back(var);
    }
  }
  if (!in.ok) return fail(PU_ERR_ARG, "truncated NetCDF header");
  return PU_OK;
}

int write_all(int fd, const void *buf, size_t n, long long off) {
  const char *p = (const char *)buf;
  while (n > 0) {
    ssize_t w = pwrite(fd, p, n, (off_t)off);
    if (w <= 0) return fail(PU_ERR_HIP, std::string("pwrite: ") + std::strerror(errno));
    p += w; n -= (size_t)w; off += w;
  }
  return PU_OK;
}
int read_all(int fd, void *buf, size_t n, long long off) {
  char *p = (char *)buf;
  while (n > 0) {
    ssize_t r = pread(fd, p, n, (off_t)off);
    if (r <= 0) return fail(PU_ERR_ARG, "NetCDF file shorter than its header says");
    p += r; n -= (size_t)r; off += r;
  }
  return PU_OK;
}

int ensure_staging(pu_ncfile *f) {
  if (f->d_stage[0]) return PU_OK;
  for (int b = 0; b < 2; b++) {
    PU_HIP_CHECK(hipMalloc(&f->d_stage[b], NC_CHUNK));
    PU_HIP_CHECK(hipHostMalloc(&f->h_stage[b], NC_CHUNK, hipHostMallocDefault));
    PU_HIP_CHECK(hipEventCreateWithFlags(&f->ev[b], hipEventDisableTiming));
  }
  return PU_OK;
}

int check_host_type(int t) {
  return (t == PU_NC_DOUBLE || t == PU_NC_FLOAT || t == PU_NC_INT || t == PU_NC_UBYTE)
             ? PU_OK
             : fail(PU_ERR_ARG, "host data type must be PU_NC_DOUBLE, PU_NC_FLOAT, PU_NC_INT or PU_NC_UBYTE");
}

// one host value as double
double host_get(const void *data, int t, long long i) {
  if (t == PU_NC_DOUBLE) return ((const double *)data)[i];
  if (t == PU_NC_FLOAT) return ((const float *)data)[i];
  if (t == PU_NC_UBYTE) return ((const unsigned char *)data)[i];
  return ((const int *)data)[i];
}

}  // namespace

extern "C" {

int pu_nc_create(const char *path, int format, pu_ncfile **out) {
  if (!path || !out) return fail(PU_ERR_ARG, "pu_nc_create: null argument");
  if (format != PU_NC_AUTO && format != PU_NC_CDF2 && format != PU_NC_CDF5)
    return fail(PU_ERR_ARG, "pu_nc_create: format must be PU_NC_AUTO, PU_NC_CDF2 or PU_NC_CDF5");
  int fd = open(path, O_RDWR | O_CREAT | O_TRUNC, 0644);
  if (fd < 0) return fail(PU_ERR_ARG, std::string("pu_nc_create: ") + path + ": " + std::strerror(errno));
  pu_ncfile *f = new pu_ncfile();
  f->fd = fd;
  f->writable = true;
  f->version = format == PU_NC_CDF5 ? 5 : 2;
  f->version_fixed = format != PU_NC_AUTO;
  *out = f;
  return PU_OK;
}

int pu_nc_open(const char *path, pu_ncfile **out) {
  if (!path || !out) return fail(PU_ERR_ARG, "pu_nc_open: null argument");
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fail(PU_ERR_ARG, std::string("pu_nc_open: ") + path + ": " + std::strerror(errno));
  struct stat st;
  fstat(fd, &st);
  size_t hn = (size_t)std::min<long long>((long long)st.st_size, NC_HEADER_RESERVE);
  std::vector<unsigned char> h(hn);
  pu_ncfile *f = new pu_ncfile();
  f->fd = fd;
  f->in_define = false;
  f->version_fixed = true;
  int rc = read_all(fd, h.data(), hn, 0);
  if (rc == PU_OK) rc = decode_header(f, h);
  if (rc != PU_OK) { close(fd); delete f; return rc; }
  *out = f;
  return PU_OK;
}

int pu_nc_def_dim(pu_ncfile *f, const char *name, long long len) {
  if (!f || !name || !f->writable) return fail(PU_ERR_ARG, "pu_nc_def_dim: null argument or read-only file");
  if (!f->in_define) return fail(PU_ERR_STATE, "pu_nc_def_dim: not in define mode (pu_nc_redef)");
  if (len < 1) return fail(PU_ERR_ARG, "pu_nc_def_dim: dimensions must have length >= 1");
  int d = find_dim(f, name);
  if (d >= 0) return f->dims[d].len == len ? PU_OK : fail(PU_ERR_ARG, std::string("dimension ") + name + " exists with another length");
  f->dims.push_back({name, len});
  return PU_OK;
}

int pu_nc_def_var(pu_ncfile *f, const char *name, int nc_type, int ndims, const char *const *dimnames) {
  if (!f || !name || !f->writable || (ndims > 0 && !dimnames)) return fail(PU_ERR_ARG, "pu_nc_def_var: null argument");
  if (!f->in_define) return fail(PU_ERR_STATE, "pu_nc_def_var: not in define mode (pu_nc_redef)");
  if (nc_type_size(nc_type) == 0) return fail(PU_ERR_ARG, "pu_nc_def_var: unknown type");
  if (find_var(f, name) >= 0) return fail(PU_ERR_ARG, std::string("variable ") + name + " already defined");
  pu_ncfile::Var v;
  v.name = name;
  v.type = nc_type;
  for (int d = 0; d < ndims; d++) {
    int id = find_dim(f, dimnames[d]);
    if (id < 0) return fail(PU_ERR_ARG, std::string("pu_nc_def_var: unknown dimension ") + dimnames[d]);
    v.dimids.push_back(id);
    v.nelems *= f->dims[id].len;
  }
  v.nbytes = v.nelems * nc_type_size(nc_type);
  v.begin = -1;
  f->vars.push_back(v);
  return PU_OK;
}

int pu_nc_redef(pu_ncfile *f) {
  if (!f || !f->writable) return fail(PU_ERR_ARG, "pu_nc_redef: null or read-only file");
  f->in_define = true;
  return PU_OK;
}

int pu_nc_enddef(pu_ncfile *f) {
  if (!f || !f->writable) return fail(PU_ERR_ARG, "pu_nc_enddef: null or read-only file");
  if (!f->in_define) return PU_OK;
  // CDF-2 holds variables up to 4 GiB - 4 bytes (only the last one may be larger) and no
  // unsigned bytes or 64-bit integers; the automatic format switches to CDF-5 when it must
  bool need5 = false;
  for (size_t i = 0; i < f->vars.size(); i++) {
    auto &v = f->vars[i];
    if (v.type == PU_NC_UBYTE || v.type == PU_NC_INT64) need5 = true;
    if (v.nbytes > 0xFFFFFFFCll) need5 = true;
  }
  if (need5) {
    if (f->version_fixed && f->version != 5)
      return fail(PU_ERR_ARG, "pu_nc_enddef: a variable needs CDF-5 (over 4 GiB or a 64-bit/unsigned type)");
    if (!f->version_fixed && f->data_end > NC_HEADER_RESERVE && f->version != 5)
      return fail(PU_ERR_STATE, "pu_nc_enddef: CDF-5 is needed after data were written as CDF-2");
    f->version = 5;
  }
  for (auto &v : f->vars) {
    if (v.begin >= 0) continue;
    v.begin = (f->data_end + NC_VAR_ALIGN - 1) / NC_VAR_ALIGN * NC_VAR_ALIGN;
    f->data_end = v.begin + (v.nbytes + 3) / 4 * 4;
  }
  std::vector<unsigned char> h = encode_header(f);
  if ((long long)h.size() > NC_HEADER_RESERVE) return fail(PU_ERR_ARG, "pu_nc_enddef: header exceeds 1 MiB");
  int rc = write_all(f->fd, h.data(), h.size(), 0);
  if (rc != PU_OK) return rc;
  // the file spans every defined variable even before its data are written
  if (ftruncate(f->fd, (off_t)f->data_end) != 0) return fail(PU_ERR_HIP, std::string("ftruncate: ") + std::strerror(errno));
  f->in_define = false;
  return PU_OK;
}

int pu_nc_inq_var(pu_ncfile *f, const char *name, int *nc_type, int *ndims, long long *dims) {
  if (!f || !name) return fail(PU_ERR_ARG, "pu_nc_inq_var: null argument");
  int i = find_var(f, name);
  if (i < 0) return fail(PU_ERR_ARG, std::string("no variable ") + name);
  auto &v = f->vars[i];
  if (nc_type) *nc_type = v.type;
  if (ndims) *ndims = (int)v.dimids.size();
  if (dims) for (size_t d = 0; d < v.dimids.size(); d++) dims[d] = f->dims[v.dimids[d]].len;
  return PU_OK;
}

int pu_nc_put_host(pu_ncfile *f, const char *name, const void *data, int host_type) {
  if (!f || !name || !data || !f->writable) return fail(PU_ERR_ARG, "pu_nc_put_host: null argument or read-only file");
  if (f->in_define) return fail(PU_ERR_STATE, "pu_nc_put_host: call pu_nc_enddef first");
  int rc = check_host_type(host_type);
  if (rc) return rc;
  int i = find_var(f, name);
  if (i < 0) return fail(PU_ERR_ARG, std::string("no variable ") + name);
  auto &v = f->vars[i];
  const int ts = nc_type_size(v.type);
  std::vector<unsigned char> buf((size_t)v.nelems * ts);
  for (long long e = 0; e < v.nelems; e++) {
    double x = host_get(data, host_type, e);
    unsigned long long bits;
    if (v.type == PU_NC_DOUBLE) std::memcpy(&bits, &x, 8);
    else if (v.type == PU_NC_FLOAT) { float y = (float)x; unsigned int b; std::memcpy(&b, &y, 4); bits = b; }
    else if (v.type == PU_NC_INT64) bits = (unsigned long long)(long long)x;
    else if (v.type == PU_NC_INT) bits = (unsigned int)(int)x;
    else if (v.type == PU_NC_SHORT) bits = (unsigned short)(short)x;
    else bits = (unsigned char)(int)x;
    for (int k = 0; k < ts; k++) buf[(size_t)e * ts + k] = (unsigned char)(bits >> (8 * (ts - 1 - k)));
  }
  return write_all(f->fd, buf.data(), buf.size(), v.begin);
}

int pu_nc_get_host(pu_ncfile *f, const char *name, void *data, int host_type) {
  if (!f || !name || !data) return fail(PU_ERR_ARG, "pu_nc_get_host: null argument");
  int rc = check_host_type(host_type);
  if (rc) return rc;
  int i = find_var(f, name);
  if (i < 0) return fail(PU_ERR_ARG, std::string("no variable ") + name);
  auto &v = f->vars[i];
  const int ts = nc_type_size(v.type);
  std::vector<unsigned char> buf((size_t)v.nelems * ts);
  rc = read_all(f->fd, buf.data(), buf.size(), v.begin);
  if (rc) return rc;
  for (long long e = 0; e < v.nelems; e++) {
    unsigned long long bits = 0;
    for (int k = 0; k < ts; k++) bits = (bits << 8) | buf[(size_t)e * ts + k];
    double x;
    if (v.type == PU_NC_DOUBLE) std::memcpy(&x, &bits, 8);
    else if (v.type == PU_NC_FLOAT) { unsigned int b = (unsigned int)bits; float y; std::memcpy(&y, &b, 4); x = y; }
    else if (v.type == PU_NC_INT64) x = (double)(long long)bits;
    else if (v.type == PU_NC_INT) x = (int)(unsigned int)bits;
    else if (v.type == PU_NC_SHORT) x = (short)(unsigned short)bits;
    else if (v.type == PU_NC_BYTE) x = (signed char)bits;
    else x = (unsigned char)bits;
    if (host_type == PU_NC_DOUBLE) ((double *)data)[e] = x;
    else if (host_type == PU_NC_FLOAT) ((float *)data)[e] = (float)x;
    else if (host_type == PU_NC_UBYTE) ((unsigned char *)data)[e] = (unsigned char)(int)x;
    else ((int *)data)[e] = (int)x;
  }
  return PU_OK;
}

int pu_nc_put_device(pu_ncfile *f, const char *name, const double *d_data, pu_ctx *ctx) {
  if (!f || !name || !d_data || !f->writable) return fail(PU_ERR_ARG, "pu_nc_put_device: null argument or read-only file");
  if (f->in_define) return fail(PU_ERR_STATE, "pu_nc_put_device: call pu_nc_enddef first");
  int i = find_var(f, name);
  if (i < 0) return fail(PU_ERR_ARG, std::string("no variable ") + name);
  auto &v = f->vars[i];
  if (v.type != PU_NC_FLOAT && v.type != PU_NC_DOUBLE && v.type != PU_NC_INT)
    return fail(PU_ERR_ARG, "pu_nc_put_device: variable type must be float, double or int");
  int rc = ensure_staging(f);
  if (rc) return rc;
  const int ts = nc_type_size(v.type);
  const long long per = (long long)(NC_CHUNK / ts);
  const long long nchunks = (v.nelems + per - 1) / per;
  hipStream_t s = ctx ? ctx->stream : nullptr;
  // chunk c: encode + copy to h_stage[c % 2] on the stream, then the host writes it; the copy of
  // chunk c+1 is queued before the host waits for chunk c
  auto enqueue = [&](long long c) -> int {
    const long long e0 = c * per, n = std::min(per, v.nelems - e0);
    const int b = (int)(c & 1);
    const dim3 grd((unsigned)((n + 255) / 256));
    if (v.type == PU_NC_FLOAT) hipLaunchKernelGGL(k_nc_encode<float>, grd, dim3(256), 0, s, d_data + e0, n, f->d_stage[b]);
    else if (v.type == PU_NC_DOUBLE) hipLaunchKernelGGL(k_nc_encode<double>, grd, dim3(256), 0, s, d_data + e0, n, f->d_stage[b]);
    else hipLaunchKernelGGL(k_nc_encode<int>, grd, dim3(256), 0, s, d_data + e0, n, f->d_stage[b]);
    PU_HIP_CHECK(hipGetLastError());
    PU_HIP_CHECK(hipMemcpyAsync(f->h_stage[b], f->d_stage[b], (size_t)n * ts, hipMemcpyDeviceToHost, s));
    PU_HIP_CHECK(hipEventRecord(f->ev[b], s));
    return PU_OK;
  };
  if (nchunks > 0 && (rc = enqueue(0))) return rc;
  for (long long c = 0; c < nchunks; c++) {
    const int b = (int)(c & 1);
    PU_HIP_CHECK(hipEventSynchronize(f->ev[b]));
    // h_stage[1-b] was written by the host in the previous iteration: free for chunk c+1
    if (c + 1 < nchunks && (rc = enqueue(c + 1))) return rc;
    const long long e0 = c * per, n = std::min(per, v.nelems - e0);
    if ((rc = write_all(f->fd, f->h_stage[b], (size_t)n * ts, v.begin + e0 * ts))) return rc;
  }
  return PU_OK;
}

int pu_nc_get_device(pu_ncfile *f, const char *name, double *d_data, pu_ctx *ctx) {
  if (!f || !name || !d_data) return fail(PU_ERR_ARG, "pu_nc_get_device: null argument");
  int i = find_var(f, name);
  if (i < 0) return fail(PU_ERR_ARG, std::string("no variable ") + name);
  auto &v = f->vars[i];
  if (v.type != PU_NC_FLOAT && v.type != PU_NC_DOUBLE && v.type != PU_NC_INT)
    return fail(PU_ERR_ARG, "pu_nc_get_device: variable type must be float, double or int");
  int rc = ensure_staging(f);
  if (rc) return rc;
  const int ts = nc_type_size(v.type);
  const long long per = (long long)(NC_CHUNK / ts);
  const long long nchunks = (v.nelems + per - 1) / per;
  hipStream_t s = ctx ? ctx->stream : nullptr;
  for (long long c = 0; c < nchunks; c++) {
    const int b = (int)(c & 1);
    const long long e0 = c * per, n = std::min(per, v.nelems - e0);
    // h_stage[b] was last used by chunk c-2, whose copy has completed once its event has
    if (c >= 2) PU_HIP_CHECK(hipEventSynchronize(f->ev[b]));
    if ((rc = read_all(f->fd, f->h_stage[b], (size_t)n * ts, v.begin + e0 * ts))) return rc;
    PU_HIP_CHECK(hipMemcpyAsync(f->d_stage[b], f->h_stage[b], (size_t)n * ts, hipMemcpyHostToDevice, s));
    const dim3 grd((unsigned)((n + 255) / 256));
    if (v.type == PU_NC_FLOAT) hipLaunchKernelGGL(k_nc_decode<float>, grd, dim3(256), 0, s, f->d_stage[b], n, d_data + e0);
    else if (v.type == PU_NC_DOUBLE) hipLaunchKernelGGL(k_nc_decode<double>, grd, dim3(256), 0, s, f->d_stage[b], n, d_data + e0);
    else hipLaunchKernelGGL(k_nc_decode<int>, grd, dim3(256), 0, s, f->d_stage[b], n, d_data + e0);
    PU_HIP_CHECK(hipGetLastError());
    PU_HIP_CHECK(hipEventRecord(f->ev[b], s));
  }
  PU_HIP_CHECK(hipStreamSynchronize(s));
  return PU_OK;
}

int pu_nc_format(pu_ncfile *f, int *version) {
  if (!f || !version) return fail(PU_ERR_ARG, "pu_nc_format: null argument");
  *version = f->version;
  return PU_OK;
}

int pu_nc_close(pu_ncfile *f) {
  if (!f) return PU_OK;
  int rc = PU_OK;
  if (f->writable && f->in_define) rc = pu_nc_enddef(f);
  for (int b = 0; b < 2; b++) {
    if (f->ev[b]) hipEventSynchronize(f->ev[b]);
    if (f->d_stage[b]) hipFree(f->d_stage[b]);
    if (f->h_stage[b]) hipHostFree(f->h_stage[b]);
    if (f->ev[b]) hipEventDestroy(f->ev[b]);
  }
  if (f->fd >= 0 && close(f->fd) != 0 && rc == PU_OK) rc = fail(PU_ERR_HIP, std::string("close: ") + std::strerror(errno));
  delete f;
  return rc;
}

}  // extern "C"
