// io.hip -- the on-disk formats either side of the path: the dycore's tracer
// restart (fv_tracer.res.tile<t>.nc, one file per tile, written and read by
// the dycore's fv_io through FMS; the physics restarts beside it are
// io/fv3atm_restart_io.F90:126-146) and the dycore history bundle in the
// layout of the write component's netCDF writer (io/module_write_netcdf.F90:
// 50-560: grid_xt, grid_yt, nchars, pfull, phalf, tile, time dimensions;
// lon/lat; fields (grid_xt, grid_yt, [pfull,] tile, time); BitRound
// quantisation of 3-D fields, :417-443).
//
// Files are netCDF classic with 64-bit offsets (CDF-2, the NF90_64BIT_OFFSET
// choice of module_write_netcdf.F90:35 and the FMS restart default), written
// and read here without libnetcdf.  The device does the per-cell work in one
// pass per level chunk: strip the halo, convert to the file's type, round
// (BitRound), accumulate the FMS checksum (the sum modulo 2^64 of the values'
// bit patterns read as two's-complement integers), and swap to big-endian;
// the host moves the bytes.  The read runs the same pass backwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fv3atm_amd.h"
#include "fv3_internal.h"
#include "ncfile.h"

namespace fv3 {

namespace {

using namespace nc;

// ---- device pass ---------------------------------------------------------------
// One cell of levels [k0, k0 + nk) of one tile's field: (isd:ied, jsd:jed, k)
// with compute domain (is:ie, js:je), lane order i fastest.
struct Slab {
  long long s2;       // one level of the device array
  int nid, ni, nj;    // device row length, compute-domain extents
  int ioff, joff;     // is - isd, js - jsd
};

// the value's bit pattern as a two's-complement integer, sign-extended
__device__ __forceinline__ unsigned long long ck_bits(double v) { return __builtin_bit_cast(unsigned long long, v); }
__device__ __forceinline__ unsigned long long ck_bits(float v) {
  return (unsigned long long)(long long)__builtin_bit_cast(int, v);
}

// BitRound: keep nsd mantissa bits, rounding the dropped ones to nearest
// (a half rounds away from zero); NaN and infinities pass through
__device__ __forceinline__ float bitround(float v, int nsd) {
  if (nsd <= 0 || nsd >= 23) return v;
  unsigned u = __builtin_bit_cast(unsigned, v);
  if ((u & 0x7F800000u) == 0x7F800000u) return v;
  const int sh = 23 - nsd;
  u = (u + (1u << (sh - 1))) & ~((1u << sh) - 1u);
  return __builtin_bit_cast(float, u);
}
__device__ __forceinline__ double bitround(double v, int nsd) {
  if (nsd <= 0 || nsd >= 52) return v;
  unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  if ((u & 0x7FF0000000000000ULL) == 0x7FF0000000000000ULL) return v;
  const int sh = 52 - nsd;
  u = (u + (1ULL << (sh - 1))) & ~((1ULL << sh) - 1ULL);
  return __builtin_bit_cast(double, u);
}

__device__ __forceinline__ unsigned long long to_be(double v) { return __builtin_bswap64(__builtin_bit_cast(unsigned long long, v)); }
__device__ __forceinline__ unsigned to_be(float v) { return __builtin_bswap32(__builtin_bit_cast(unsigned, v)); }

// block sum of the threads' checksum terms (wave shuffles, then the four
// waves' partials in LDS) and one 64-bit atomic per block: integer sums, so
// the result does not depend on the order
__device__ __forceinline__ void ck_add(unsigned long long *ck, unsigned long long v) {
  __shared__ unsigned long long part[4];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(ck, part[0] + part[1] + part[2] + part[3]);
}

// blocks of the grid-stride device passes: enough to fill the device, few
// enough that the per-block checksum atomics stay negligible
constexpr int IO_BLOCK = 256, IO_GRID = 2048;
dim3 io_grid(long long n) { return dim3((unsigned)std::max(1LL, std::min<long long>((n + IO_BLOCK - 1) / IO_BLOCK, IO_GRID))); }

template <typename Tin, typename Tout, typename Ubits>
__global__ void __launch_bounds__(IO_BLOCK) k_pack(const Tin *src, Slab s, int nk, Ubits *dst,
                                                   unsigned long long *ck, int nsd) {
  const long long n = (long long)s.ni * s.nj * nk;
  unsigned long long c = 0;
  for (long long idx = (long long)blockIdx.x * IO_BLOCK + threadIdx.x; idx < n; idx += (long long)gridDim.x * IO_BLOCK) {
    const int i = (int)(idx % s.ni);
    const long long r = idx / s.ni;
    const int j = (int)(r % s.nj), k = (int)(r / s.nj);
    const Tout v = bitround((Tout)src[(long long)k * s.s2 + (long long)(j + s.joff) * s.nid + (i + s.ioff)], nsd);
    c += ck_bits(v);
    dst[idx] = to_be(v);
  }
  ck_add(ck, c);
}

template <typename Tdev>
__global__ void __launch_bounds__(IO_BLOCK) k_unpack(const unsigned long long *src, Slab s, int nk, Tdev *dst,
                                                     unsigned long long *ck) {
  const long long n = (long long)s.ni * s.nj * nk;
  unsigned long long c = 0;
  for (long long idx = (long long)blockIdx.x * IO_BLOCK + threadIdx.x; idx < n; idx += (long long)gridDim.x * IO_BLOCK) {
    const int i = (int)(idx % s.ni);
    const long long r = idx / s.ni;
    const int j = (int)(r % s.nj), k = (int)(r / s.nj);
    const double v = __builtin_bit_cast(double, __builtin_bswap64(src[idx]));
    c += ck_bits(v);
    dst[(long long)k * s.s2 + (long long)(j + s.joff) * s.nid + (i + s.ioff)] = (Tdev)v;
  }
  ck_add(ck, c);
}

template <typename T>
__global__ void __launch_bounds__(IO_BLOCK) k_checksum(const T *src, Slab s, int nk, unsigned long long *ck) {
  const long long n = (long long)s.ni * s.nj * nk;
  unsigned long long c = 0;
  for (long long idx = (long long)blockIdx.x * IO_BLOCK + threadIdx.x; idx < n; idx += (long long)gridDim.x * IO_BLOCK) {
    const int i = (int)(idx % s.ni);
    const long long r = idx / s.ni;
    const int j = (int)(r % s.nj), k = (int)(r / s.nj);
    c += ck_bits(src[(long long)k * s.s2 + (long long)(j + s.joff) * s.nid + (i + s.ioff)]);
  }
  ck_add(ck, c);
}

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

Slab slab_of(const Dims &d) {
  Slab s;
  s.s2 = d.s2;
  s.nid = d.nid;
  s.ni = d.ie - d.is + 1;
  s.nj = d.je - d.js + 1;
  s.ioff = d.is - d.isd;
  s.joff = d.js - d.jsd;
  return s;
}

// levels per device pass: about 64 MiB of file bytes
int chunk_levels(const Slab &s, int nlev, int out_bytes) {
  const long long per = (long long)s.ni * s.nj * out_bytes;
  long long k = (64LL << 20) / std::max(per, 1LL);
  if (k < 1) k = 1;
  return (int)std::min<long long>(k, nlev);
}

// Device staging for one call: a chunk of packed bytes and the checksums.
struct Staging {
  void *dev = nullptr;          // packed bytes of one chunk (from ctx_stage)
  unsigned long long *ck = nullptr;
  void *host = nullptr;         // pinned
  size_t host_bytes = 0;
  ~Staging() {
    if (host) (void)hipHostFree(host);
  }
};

int staging(Ctx &c, Staging &st, size_t chunk_bytes, int nck, hipStream_t s) {
  const size_t off = (chunk_bytes + 255) & ~(size_t)255;
  char *base = (char *)ctx_stage(c, off + sizeof(unsigned long long) * std::max(nck, 1));
  if (!base) return 1;
  st.dev = base;
  st.ck = (unsigned long long *)(base + off);
  FV3_HIP(hipHostMalloc(&st.host, chunk_bytes, hipHostMallocDefault));
  st.host_bytes = chunk_bytes;
  hipLaunchKernelGGL(k_zero_u64, dim3((nck + 255) / 256), dim3(256), 0, s, st.ck, nck);
  FV3_HIP(hipGetLastError());
  return 0;
}

// packs levels of one tile's field and writes them at file offset off
template <typename Tin>
int pack_write(Ctx &c, Staging &st, FILE *fp, unsigned long long off, const Tin *field, int nlev, int out_bytes,
               int nsd, unsigned long long *ck, hipStream_t s) {
  const Slab sl = slab_of(c.d);
  const int kc = chunk_levels(sl, nlev, out_bytes);
  if (fseeko(fp, (off_t)off, SEEK_SET)) return set_error("seek in the output file failed: %s", strerror(errno));
  for (int k0 = 0; k0 < nlev; k0 += kc) {
    const int nk = std::min(kc, nlev - k0);
    const long long n = (long long)sl.ni * sl.nj * nk;
    const dim3 g = io_grid(n);
    if (out_bytes == 8)
      hipLaunchKernelGGL((k_pack<Tin, double, unsigned long long>), g, dim3(IO_BLOCK), 0, s, field + k0 * sl.s2, sl, nk,
                         (unsigned long long *)st.dev, ck, nsd);
    else
      hipLaunchKernelGGL((k_pack<Tin, float, unsigned>), g, dim3(IO_BLOCK), 0, s, field + k0 * sl.s2, sl, nk,
                         (unsigned *)st.dev, ck, nsd);
    FV3_HIP(hipGetLastError());
    const size_t nb = (size_t)n * out_bytes;
    FV3_HIP(hipMemcpyAsync(st.host, st.dev, nb, hipMemcpyDeviceToHost, s));
    FV3_HIP(hipStreamSynchronize(s));
    if (fwrite(st.host, 1, nb, fp) != nb) return set_error("write to the output file failed: %s", strerror(errno));
  }
  return 0;
}

// reads levels of one tile's double field at file offset off into the device
template <typename Tdev>
int read_unpack(Ctx &c, Staging &st, FILE *fp, unsigned long long off, Tdev *field, int nlev,
                unsigned long long *ck, hipStream_t s) {
  const Slab sl = slab_of(c.d);
  const int kc = chunk_levels(sl, nlev, 8);
  if (fseeko(fp, (off_t)off, SEEK_SET)) return set_error("seek in the restart file failed: %s", strerror(errno));
  for (int k0 = 0; k0 < nlev; k0 += kc) {
    const int nk = std::min(kc, nlev - k0);
    const long long n = (long long)sl.ni * sl.nj * nk;
    const size_t nb = (size_t)n * 8;
    if (fread(st.host, 1, nb, fp) != nb) return set_error("the restart file ends early");
    FV3_HIP(hipMemcpyAsync(st.dev, st.host, nb, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_unpack<Tdev>, io_grid(n), dim3(IO_BLOCK), 0, s,
                       (const unsigned long long *)st.dev, sl, nk, field + k0 * sl.s2, ck);
    FV3_HIP(hipGetLastError());
    FV3_HIP(hipStreamSynchronize(s));
  }
  return 0;
}

std::string tile_path(const char *prefix, int t) { return std::string(prefix) + ".tile" + std::to_string(t) + ".nc"; }

// FMS writes the checksum as a Fortran Z16 edit: upper-case hex, right-justified in 16 characters
std::string cksum_text(unsigned long long v) {
  char b[32];
  snprintf(b, sizeof b, "%16" PRIX64, (uint64_t)v);
  return b;
}

struct FileCloser {
  FILE *f;
  ~FileCloser() {
    if (f) fclose(f);
  }
};

int write_header(FILE *fp, const NcFile &f) {
  const std::vector<unsigned char> h = f.header();
  if (fseeko(fp, 0, SEEK_SET) || fwrite(h.data(), 1, h.size(), fp) != h.size())
    return set_error("writing the file header failed: %s", strerror(errno));
  return 0;
}

int write_host(FILE *fp, unsigned long long off, const std::vector<unsigned char> &b) {
  if (fseeko(fp, (off_t)off, SEEK_SET) || fwrite(b.data(), 1, b.size(), fp) != b.size())
    return set_error("write to the output file failed: %s", strerror(errno));
  return 0;
}

std::vector<unsigned char> be_doubles(const double *v, long long n) {
  std::vector<unsigned char> o;
  o.reserve(n * 8);
  for (long long i = 0; i < n; i++) {
    unsigned long long u;
    std::memcpy(&u, &v[i], 8);
    put_be(o, u, 8);
  }
  return o;
}
std::vector<unsigned char> be_floats(const float *v, long long n) {
  std::vector<unsigned char> o;
  for (long long i = 0; i < n; i++) {
    unsigned u;
    std::memcpy(&u, &v[i], 4);
    put_be(o, u, 4);
  }
  return o;
}

bool names_ok(int n, const char *const *names) {
  if (n < 1 || !names) return false;
  for (int i = 0; i < n; i++)
    if (!names[i] || !names[i][0]) return false;
  return true;
}

}  // namespace

// ---- tracer restart -------------------------------------------------------------
static int restart_write(Ctx &c, const char *prefix, int tile_base, int nq, const char *const *names, const void *q,
                         int elem, hipStream_t s) {
  const Dims &d = c.d;
  const int npz = c.npz, nx = d.ie - d.is + 1, ny = d.je - d.js + 1;
  const Slab sl = slab_of(d);
  const int kc = chunk_levels(sl, npz, 8);
  Staging st;
  if (staging(c, st, (size_t)sl.ni * sl.nj * kc * 8, nq * c.ntiles, s)) return 1;
  for (int t = 0; t < c.ntiles; t++) {
    NcFile f;
    f.numrecs = 1;
    const int dx = f.dim("xaxis_1", nx), dy = f.dim("yaxis_1", ny), dz = f.dim("zaxis_1", npz), dt = f.dim("Time", 0);
    const char *ax[4] = {"xaxis_1", "yaxis_1", "zaxis_1", "Time"};
    const char *cart[4] = {"X", "Y", "Z", "T"};
    const int adim[4] = {dx, dy, dz, dt};
    for (int a = 0; a < 4; a++) {
      Var &v = f.var(ax[a], NC_DOUBLE, {adim[a]});
      v.atts.push_back(att_text("long_name", ax[a]));
      v.atts.push_back(att_text("units", a == 3 ? "time level" : "none"));
      v.atts.push_back(att_text("cartesian_axis", cart[a]));
    }
    for (int iq = 0; iq < nq; iq++) {
      Var &v = f.var(names[iq], NC_DOUBLE, {dt, dz, dy, dx});
      v.atts.push_back(att_text("long_name", names[iq]));
      v.atts.push_back(att_text("units", "none"));
      v.atts.push_back(att_text("checksum", cksum_text(0)));
    }
    f.layout();
    const std::string path = tile_path(prefix, tile_base + t);
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) return set_error("cannot create %s: %s", path.c_str(), strerror(errno));
    FileCloser fc{fp};
    std::vector<double> axis;
    for (int a = 0; a < 4; a++) {
      const long long n = a == 3 ? 1 : f.dims[adim[a]].len;
      axis.resize(n);
      for (long long i = 0; i < n; i++) axis[i] = (double)(i + 1);
      if (write_host(fp, f.vars[a].begin, be_doubles(axis.data(), n))) return 1;
    }
    for (int iq = 0; iq < nq; iq++) {
      const long long fo = ((long long)t * nq + iq) * npz * d.s2;
      unsigned long long *ck = st.ck + t * nq + iq;
      const int rc = elem == 8 ? pack_write(c, st, fp, f.vars[4 + iq].begin, (const double *)q + fo, npz, 8, 0, ck, s)
                               : pack_write(c, st, fp, f.vars[4 + iq].begin, (const float *)q + fo, npz, 8, 0, ck, s);
      if (rc) return 1;
    }
    std::vector<unsigned long long> ck(nq);
    FV3_HIP(hipMemcpy(ck.data(), st.ck + t * nq, sizeof(unsigned long long) * nq, hipMemcpyDeviceToHost));
    for (int iq = 0; iq < nq; iq++) f.vars[4 + iq].atts[2] = att_text("checksum", cksum_text(ck[iq]));
    if (write_header(fp, f)) return 1;
  }
  return 0;
}

static int restart_read(Ctx &c, const char *prefix, int tile_base, int nq, const char *const *names, void *q,
                        int elem, int ignore_cksum, hipStream_t s) {
  const Dims &d = c.d;
  const int npz = c.npz, nx = d.ie - d.is + 1, ny = d.je - d.js + 1;
  const Slab sl = slab_of(d);
  const int kc = chunk_levels(sl, npz, 8);
  Staging st;
  if (staging(c, st, (size_t)sl.ni * sl.nj * kc * 8, nq * c.ntiles, s)) return 1;
  for (int t = 0; t < c.ntiles; t++) {
    const std::string path = tile_path(prefix, tile_base + t);
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) return set_error("cannot open %s: %s", path.c_str(), strerror(errno));
    FileCloser fc{fp};
    // the header: read until it parses (it is short next to the data)
    std::vector<unsigned char> hb;
    NcFile f;
    bool parsed = false;
    for (size_t want = 1 << 16; !parsed; want *= 4) {
      hb.resize(want);
      if (fseeko(fp, 0, SEEK_SET)) return set_error("seek in %s failed", path.c_str());
      const size_t got = fread(hb.data(), 1, want, fp);
      hb.resize(got);
      f = NcFile();
      Reader r;
      parsed = r.parse(hb, f);
      if (!parsed && (got < want || want > (1u << 28)))
        return set_error("%s is not a netCDF classic file or its header is damaged", path.c_str());
    }
    for (int iq = 0; iq < nq; iq++) {
      const Var *v = nullptr;
      for (const Var &w : f.vars)
        if (w.name == names[iq]) v = &w;
      if (!v) return set_error("%s: no variable %s", path.c_str(), names[iq]);
      if (v->type != NC_DOUBLE) return set_error("%s: %s is not double", path.c_str(), names[iq]);
      // (Time,) zaxis_1, yaxis_1, xaxis_1 with this context's extents
      std::vector<long long> ext;
      for (size_t k = v->rec ? 1 : 0; k < v->dims.size(); k++) ext.push_back(f.dims[v->dims[k]].len);
      if (ext.size() != 3 || ext[0] != npz || ext[1] != ny || ext[2] != nx)
        return set_error("%s: %s does not have the shape (%d, %d, %d)", path.c_str(), names[iq], npz, ny, nx);
      if (v->rec && f.numrecs < 1) return set_error("%s: %s holds no record", path.c_str(), names[iq]);
      // the last record of a record variable
      const unsigned long long off = v->begin + (v->rec ? (unsigned long long)(f.numrecs - 1) * f.recsize : 0);
      const long long fo = ((long long)t * nq + iq) * npz * d.s2;
      unsigned long long *ck = st.ck + t * nq + iq;
      const int rc = elem == 8 ? read_unpack(c, st, fp, off, (double *)q + fo, npz, ck, s)
                               : read_unpack(c, st, fp, off, (float *)q + fo, npz, ck, s);
      if (rc) return 1;
      if (!ignore_cksum) {
        const Att *a = nullptr;
        for (const Att &w : v->atts)
          if (w.name == "checksum") a = &w;
        if (!a || a->type != NC_CHAR) return set_error("%s: %s carries no checksum attribute", path.c_str(), names[iq]);
        std::string txt(a->be.begin(), a->be.end());
        const size_t b = txt.find_first_not_of(' ');
        txt = b == std::string::npos ? "" : txt.substr(b);
        unsigned long long want = 0;
        char *endp = nullptr;
        want = strtoull(txt.c_str(), &endp, 16);
        if (txt.empty() || (endp && *endp && *endp != ' '))
          return set_error("%s: %s has a malformed checksum '%s'", path.c_str(), names[iq], txt.c_str());
        unsigned long long got = 0;
        FV3_HIP(hipMemcpy(&got, ck, sizeof got, hipMemcpyDeviceToHost));
        if (got != want)
          return set_error("%s: checksum of %s is %s, the file says %s", path.c_str(), names[iq],
                           cksum_text(got).c_str(), cksum_text(want).c_str());
      }
    }
  }
  return 0;
}

// ---- history bundle -----------------------------------------------------------
static int history_write(Ctx &c, const char *path, const fv3_history_args &a, hipStream_t s) {
  const Dims &d = c.d;
  const int npz = c.npz, nx = d.ie - d.is + 1, ny = d.je - d.js + 1, nt = c.ntiles;
  const int ob = a.out_bytes;
  NcFile f;
  const int dim_x = f.dim("grid_xt", nx), dim_y = f.dim("grid_yt", ny), dim_c = f.dim("nchars", 20);
  const int dim_pf = f.dim("pfull", npz), dim_ph = f.dim("phalf", npz + 1), dim_t = f.dim("tile", nt);
  const int dim_tm = f.dim("time", 1);
  {
    Var &v = f.var("pfull", NC_FLOAT, {dim_pf});
    v.atts = {att_text("long_name", "ref full pressure level"), att_text("units", "mb"),
              att_text("cartesian_axis", "Z"), att_text("positive", "down"), att_text("edges", "phalf")};
  }
  {
    Var &v = f.var("phalf", NC_FLOAT, {dim_ph});
    v.atts = {att_text("long_name", "ref half pressure level"), att_text("units", "mb"),
              att_text("cartesian_axis", "Z"), att_text("positive", "down")};
  }
  {
    Var &v = f.var("time", NC_DOUBLE, {dim_tm});
    v.atts = {att_text("long_name", "time"), att_text("units", a.time_units ? a.time_units : "hours"),
              att_text("cartesian_axis", "T"), att_text("calendar_type", "JULIAN"), att_text("calendar", "JULIAN")};
  }
  f.var("grid_xt", NC_DOUBLE, {dim_x}).atts = {att_text("cartesian_axis", "X")};
  f.var("grid_yt", NC_DOUBLE, {dim_y}).atts = {att_text("cartesian_axis", "Y")};
  f.var("tile", NC_INT, {dim_t}).atts = {att_text("long_name", "cubed-sphere face")};
  f.var("time_iso", NC_CHAR, {dim_tm, dim_c}).atts = {att_text("long_name", "valid time"),
                                                      att_text("description", "ISO 8601 datetime string"),
                                                      att_text("_Encoding", "UTF-8")};
  f.var("lon", NC_DOUBLE, {dim_t, dim_y, dim_x}).atts = {att_text("long_name", "T-cell longitude"),
                                                         att_text("units", "degrees_E")};
  f.var("lat", NC_DOUBLE, {dim_t, dim_y, dim_x}).atts = {att_text("long_name", "T-cell latitude"),
                                                         att_text("units", "degrees_N")};
  const int v0 = (int)f.vars.size();
  for (int i = 0; i < a.nfields; i++) {
    const fv3_history_field &h = a.fields[i];
    std::vector<int> dd = h.nlev > 1 ? std::vector<int>{dim_tm, dim_t, dim_pf, dim_y, dim_x}
                                     : std::vector<int>{dim_tm, dim_t, dim_y, dim_x};
    Var &v = f.var(h.name, ob == 4 ? NC_FLOAT : NC_DOUBLE, dd);
    v.atts.push_back(att_text("long_name", h.long_name ? h.long_name : h.name));
    v.atts.push_back(att_text("units", h.units ? h.units : "none"));
    if (h.nlev > 1 && a.quantize_nsd > 0)
      v.atts.push_back(att_int("_QuantizeBitRoundNumberOfSignificantBits", a.quantize_nsd));
    v.atts.push_back(att_text("coordinates", "lon lat"));
    v.atts.push_back(att_text("grid_mapping", "cubed_sphere"));
  }
  f.gatts.push_back(att_text("grid", "cubed_sphere"));
  f.gatts.push_back(att_int("im", nx));
  f.gatts.push_back(att_int("jm", ny));
  if (a.ak) f.gatts.push_back(att_doubles("ak", a.ak, npz + 1));
  if (a.bk) f.gatts.push_back(att_doubles("bk", a.bk, npz + 1));
  f.layout();

  FILE *fp = fopen(path, "wb");
  if (!fp) return set_error("cannot create %s: %s", path, strerror(errno));
  FileCloser fc{fp};
  if (write_header(fp, f)) return 1;
  if (write_host(fp, f.vars[0].begin, be_floats(a.pfull, npz))) return 1;
  if (write_host(fp, f.vars[1].begin, be_floats(a.phalf, npz + 1))) return 1;
  if (write_host(fp, f.vars[2].begin, be_doubles(&a.time, 1))) return 1;
  std::vector<double> ax(std::max(nx, ny));
  for (int i = 0; i < nx; i++) ax[i] = i + 1;
  if (write_host(fp, f.vars[3].begin, be_doubles(ax.data(), nx))) return 1;
  for (int j = 0; j < ny; j++) ax[j] = j + 1;
  if (write_host(fp, f.vars[4].begin, be_doubles(ax.data(), ny))) return 1;
  {
    std::vector<unsigned char> tb;
    for (int t = 0; t < nt; t++) put_be(tb, (unsigned)(a.tile_base + t), 4);
    if (write_host(fp, f.vars[5].begin, tb)) return 1;
  }
  {
    std::vector<unsigned char> iso(20, 0);
    if (a.time_iso) {
      const size_t n = std::min<size_t>(20, strlen(a.time_iso));
      std::memcpy(iso.data(), a.time_iso, n);
    }
    if (write_host(fp, f.vars[6].begin, iso)) return 1;
  }
  const long long ncell = (long long)nx * ny * nt;
  std::vector<double> zero;
  for (int w = 0; w < 2; w++) {
    const double *src = w == 0 ? a.lon : a.lat;
    if (!src) {
      zero.assign(ncell, 0.0);
      src = zero.data();
    }
    if (write_host(fp, f.vars[7 + w].begin, be_doubles(src, ncell))) return 1;
  }
  const Slab sl = slab_of(d);
  int kmax = 1;
  for (int i = 0; i < a.nfields; i++) kmax = std::max(kmax, a.fields[i].nlev);
  const int kc = chunk_levels(sl, kmax, ob);
  Staging st;
  if (staging(c, st, (size_t)sl.ni * sl.nj * kc * ob, std::max(1, a.nfields * nt), s)) return 1;
  for (int i = 0; i < a.nfields; i++) {
    const fv3_history_field &h = a.fields[i];
    const int nsd = h.nlev > 1 ? a.quantize_nsd : 0;
    for (int t = 0; t < nt; t++) {
      const unsigned long long off = f.vars[v0 + i].begin + (unsigned long long)t * h.nlev * nx * ny * ob;
      const long long fo = (long long)t * h.nlev * d.s2;
      unsigned long long *ck = st.ck + i * nt + t;
      const int rc = h.elem_bytes == 8
                         ? pack_write(c, st, fp, off, (const double *)h.data + fo, h.nlev, ob, nsd, ck, s)
                         : pack_write(c, st, fp, off, (const float *)h.data + fo, h.nlev, ob, nsd, ck, s);
      if (rc) return 1;
    }
  }
  return 0;
}

// ---- checksum of a device field ----------------------------------------------
static int field_checksum(Ctx &c, const void *field, int nslab, int elem, unsigned long long *out, hipStream_t s) {
  const Slab sl = slab_of(c.d);
  unsigned long long *ck = (unsigned long long *)ctx_stage(c, sizeof(unsigned long long) * c.ntiles);
  if (!ck) return 1;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, s, ck, c.ntiles);
  const long long n = (long long)sl.ni * sl.nj * nslab;
  for (int t = 0; t < c.ntiles && n > 0; t++) {
    const long long fo = (long long)t * nslab * c.d.s2;
    const dim3 g = io_grid(n);
    if (elem == 8)
      hipLaunchKernelGGL(k_checksum<double>, g, dim3(IO_BLOCK), 0, s, (const double *)field + fo, sl, nslab, ck + t);
    else
      hipLaunchKernelGGL(k_checksum<float>, g, dim3(IO_BLOCK), 0, s, (const float *)field + fo, sl, nslab, ck + t);
    FV3_HIP(hipGetLastError());
  }
  FV3_HIP(hipMemcpyAsync(out, ck, sizeof(unsigned long long) * c.ntiles, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace fv3

using namespace fv3;

static int io_check(fv3_ctx *h, int elem, const void *p, const char *what) {
  if (!h || !p) return set_error("%s: NULL argument", what);
  if (elem != 4 && elem != 8) return set_error("%s: element size must be 4 or 8 bytes (got %d)", what, elem);
  return 0;
}

extern "C" {

int fv3_field_checksum(fv3_ctx *h, const void *field, int nslab, int elem_bytes, unsigned long long *out,
                       void *stream) {
  if (io_check(h, elem_bytes, field, "fv3_field_checksum") || !out) return out ? 1 : set_error("fv3_field_checksum: NULL out");
  if (nslab < 1) return set_error("fv3_field_checksum: nslab must be >= 1 (got %d)", nslab);
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  return field_checksum(c, field, nslab, elem_bytes, out, stream ? (hipStream_t)stream : c.stream);
}

int fv3_tracer_restart_write(fv3_ctx *h, const char *prefix, int tile_base, int nq, const char *const *names,
                             const void *q, int elem_bytes, void *stream) {
  if (io_check(h, elem_bytes, q, "fv3_tracer_restart_write")) return 1;
  if (!prefix || !prefix[0]) return set_error("fv3_tracer_restart_write: empty file prefix");
  if (!names_ok(nq, names)) return set_error("fv3_tracer_restart_write: nq >= 1 tracer names required");
  if (tile_base < 1) return set_error("tile_base must be >= 1 (got %d)", tile_base);
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  return restart_write(c, prefix, tile_base, nq, names, q, elem_bytes, stream ? (hipStream_t)stream : c.stream);
}

int fv3_tracer_restart_read(fv3_ctx *h, const char *prefix, int tile_base, int nq, const char *const *names,
                            void *q, int elem_bytes, int ignore_cksum, void *stream) {
  if (io_check(h, elem_bytes, q, "fv3_tracer_restart_read")) return 1;
  if (!prefix || !prefix[0]) return set_error("fv3_tracer_restart_read: empty file prefix");
  if (!names_ok(nq, names)) return set_error("fv3_tracer_restart_read: nq >= 1 tracer names required");
  if (tile_base < 1) return set_error("tile_base must be >= 1 (got %d)", tile_base);
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  return restart_read(c, prefix, tile_base, nq, names, q, elem_bytes, ignore_cksum,
                      stream ? (hipStream_t)stream : c.stream);
}

int fv3_history_write(fv3_ctx *h, const char *path, const fv3_history_args *a, void *stream) {
  if (!h || !path || !a) return set_error("fv3_history_write: NULL argument");
  if (a->out_bytes != 4 && a->out_bytes != 8) return set_error("out_bytes must be 4 or 8 (got %d)", a->out_bytes);
  if (a->nfields < 0 || (a->nfields > 0 && !a->fields)) return set_error("fv3_history_write: no field list");
  if (!a->pfull || !a->phalf) return set_error("fv3_history_write: pfull and phalf are required");
  if (a->quantize_nsd < 0 || a->quantize_nsd > (a->out_bytes == 4 ? 23 : 52))
    return set_error("quantize_nsd=%d is outside 0..%d", a->quantize_nsd, a->out_bytes == 4 ? 23 : 52);
  if (a->tile_base < 1) return set_error("tile_base must be >= 1 (got %d)", a->tile_base);
  Ctx &c = h->c;
  for (int i = 0; i < a->nfields; i++) {
    const fv3_history_field &f = a->fields[i];
    if (!f.name || !f.name[0] || !f.data) return set_error("history field %d: missing name or data", i);
    if (f.elem_bytes != 4 && f.elem_bytes != 8)
      return set_error("history field %s: element size must be 4 or 8 bytes", f.name);
    if (f.nlev != 1 && f.nlev != c.npz)
      return set_error("history field %s: nlev must be 1 or npz=%d (got %d)", f.name, c.npz, f.nlev);
  }
  std::lock_guard<std::mutex> lk(c.mu);
  return history_write(c, path, *a, stream ? (hipStream_t)stream : c.stream);
}

}  // extern "C"
