// ncfile.h -- netCDF classic files (CDF-1 read, CDF-2 read and write) in
// host code, following the published classic-format grammar: a big-endian
// header (magic, record count, dimension, global attribute and variable
// lists, each variable with its attributes, type, size and begin offset)
// followed by the fixed-size variables and then the records, each record
// holding one slice of every record variable (a lone record variable is not
// padded).  Enough of the format for the restart and history files of this
// library; no libnetcdf is needed.
#pragma once
#include <cstring>
#include <string>
#include <vector>

namespace fv3 {
namespace nc {

// ---- netCDF classic (CDF-1 / CDF-2) ------------------------------------------
enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };

inline int type_size(int t) {
  switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
  }
  return 0;
}

inline long long pad4(long long n) { return (n + 3) & ~3LL; }

struct Att {
  std::string name;
  int type = NC_CHAR;
  long long n = 0;
  std::vector<unsigned char> be;  // values, big-endian
};

struct Dim {
  std::string name;
  long long len;  // 0: the record dimension
};

struct Var {
  std::string name;
  std::vector<int> dims;
  std::vector<Att> atts;
  int type = NC_DOUBLE;
  long long vsize = 0;       // bytes of one record (record variables) or of the variable
  unsigned long long begin = 0;
  bool rec = false;
};

inline void put_be(std::vector<unsigned char> &o, unsigned long long v, int bytes) {
  for (int b = bytes - 1; b >= 0; b--) o.push_back((unsigned char)(v >> (8 * b)));
}

inline Att att_text(const std::string &name, const std::string &s) {
  Att a;
  a.name = name;
  a.type = NC_CHAR;
  a.n = (long long)s.size();
  a.be.assign(s.begin(), s.end());
  return a;
}
inline Att att_int(const std::string &name, int v) {
  Att a;
  a.name = name;
  a.type = NC_INT;
  a.n = 1;
  put_be(a.be, (unsigned)v, 4);
  return a;
}
inline Att att_doubles(const std::string &name, const double *v, long long n) {
  Att a;
  a.name = name;
  a.type = NC_DOUBLE;
  a.n = n;
  for (long long i = 0; i < n; i++) {
    unsigned long long u;
    std::memcpy(&u, &v[i], 8);
    put_be(a.be, u, 8);
  }
  return a;
}

struct NcFile {
  std::vector<Dim> dims;
  std::vector<Att> gatts;
  std::vector<Var> vars;
  long long numrecs = 0;
  long long recsize = 0;

  int dim(const std::string &name, long long len) {
    dims.push_back({name, len});
    return (int)dims.size() - 1;
  }
  Var &var(const std::string &name, int type, std::vector<int> d) {
    Var v;
    v.name = name;
    v.type = type;
    v.dims = std::move(d);
    vars.push_back(v);
    return vars.back();
  }

  static void put_name(std::vector<unsigned char> &o, const std::string &s) {
    put_be(o, s.size(), 4);
    o.insert(o.end(), s.begin(), s.end());
    while (o.size() % 4) o.push_back(0);
  }
  static void put_atts(std::vector<unsigned char> &o, const std::vector<Att> &atts) {
    if (atts.empty()) {
      put_be(o, 0, 4);
      put_be(o, 0, 4);
      return;
    }
    put_be(o, TAG_ATT, 4);
    put_be(o, atts.size(), 4);
    for (const Att &a : atts) {
      put_name(o, a.name);
      put_be(o, (unsigned)a.type, 4);
      put_be(o, (unsigned long long)a.n, 4);
      o.insert(o.end(), a.be.begin(), a.be.end());
      while (o.size() % 4) o.push_back(0);
    }
  }

  // sizes, record layout and offsets; the header is laid out twice, since
  // its own length decides where the data begins
  void layout() {
    recsize = 0;
    int nrec = 0;
    for (Var &v : vars) {
      long long n = type_size(v.type);
      v.rec = !v.dims.empty() && dims[v.dims[0]].len == 0;
      for (size_t i = v.rec ? 1 : 0; i < v.dims.size(); i++) n *= dims[v.dims[i]].len;
      v.vsize = pad4(n);
      if (v.rec) nrec++;
    }
    for (int pass = 0; pass < 2; pass++) {
      unsigned long long off = header().size();
      for (Var &v : vars)
        if (!v.rec) {
          v.begin = off;
          off += v.vsize;
        }
      for (Var &v : vars)
        if (v.rec) {
          v.begin = off;
          off += v.vsize;
        }
    }
    for (const Var &v : vars)
      if (v.rec) recsize += v.vsize;
    (void)nrec;
  }

  std::vector<unsigned char> header() const {
    std::vector<unsigned char> o = {'C', 'D', 'F', 2};
    put_be(o, (unsigned long long)numrecs, 4);
    if (dims.empty()) {
      put_be(o, 0, 4);
      put_be(o, 0, 4);
    } else {
      put_be(o, TAG_DIM, 4);
      put_be(o, dims.size(), 4);
      for (const Dim &d : dims) {
        put_name(o, d.name);
        put_be(o, (unsigned long long)d.len, 4);
      }
    }
    put_atts(o, gatts);
    if (vars.empty()) {
      put_be(o, 0, 4);
      put_be(o, 0, 4);
    } else {
      put_be(o, TAG_VAR, 4);
      put_be(o, vars.size(), 4);
      for (const Var &v : vars) {
        put_name(o, v.name);
        put_be(o, v.dims.size(), 4);
        for (int d : v.dims) put_be(o, (unsigned)d, 4);
        put_atts(o, v.atts);
        put_be(o, (unsigned)v.type, 4);
        put_be(o, v.vsize > 0xFFFFFFFCLL ? 0xFFFFFFFFULL : (unsigned long long)v.vsize, 4);
        put_be(o, v.begin, 8);
      }
    }
    return o;
  }
};

// a header parsed from a file (CDF-1 or CDF-2)
struct Reader {
  const unsigned char *p = nullptr, *end = nullptr;
  bool ok = true;
  int version = 0;
  unsigned long long u(int bytes) {
    if (end - p < bytes) {
      ok = false;
      return 0;
    }
    unsigned long long v = 0;
    for (int b = 0; b < bytes; b++) v = (v << 8) | *p++;
    return v;
  }
  std::string name() {
    const unsigned long long n = u(4);
    if (!ok || (unsigned long long)(end - p) < n) {
      ok = false;
      return "";
    }
    std::string s((const char *)p, (size_t)n);
    p += pad4((long long)n);
    if (p > end) ok = false;
    return s;
  }
  void atts(std::vector<Att> &out) {
    const unsigned long long tag = u(4), n = u(4);
    if (tag == 0) return;
    if (tag != TAG_ATT) {
      ok = false;
      return;
    }
    for (unsigned long long i = 0; i < n && ok; i++) {
      Att a;
      a.name = name();
      a.type = (int)u(4);
      a.n = (long long)u(4);
      const long long nb = a.n * type_size(a.type);
      if (type_size(a.type) == 0 || end - p < pad4(nb)) {
        ok = false;
        return;
      }
      a.be.assign(p, p + nb);
      p += pad4(nb);
      out.push_back(a);
    }
  }
  bool parse(const std::vector<unsigned char> &buf, NcFile &f) {
    p = buf.data();
    end = p + buf.size();
    if (buf.size() < 8 || p[0] != 'C' || p[1] != 'D' || p[2] != 'F' || (p[3] != 1 && p[3] != 2)) return false;
    version = p[3];
    p += 4;
    f.numrecs = (long long)u(4);
    unsigned long long tag = u(4), n = u(4);
    if (tag == TAG_DIM)
      for (unsigned long long i = 0; i < n && ok; i++) {
        std::string nm = name();
        f.dims.push_back({nm, (long long)u(4)});
      }
    else if (tag != 0)
      return false;
    atts(f.gatts);
    tag = u(4);
    n = u(4);
    if (tag == TAG_VAR)
      for (unsigned long long i = 0; i < n && ok; i++) {
        Var v;
        v.name = name();
        const unsigned long long nd = u(4);
        for (unsigned long long d = 0; d < nd && ok; d++) {
          const unsigned long long id = u(4);
          if (id >= f.dims.size()) ok = false;
          v.dims.push_back((int)id);
        }
        atts(v.atts);
        v.type = (int)u(4);
        v.vsize = (long long)u(4);
        v.begin = u(version == 2 ? 8 : 4);
        v.rec = !v.dims.empty() && f.dims[v.dims[0]].len == 0;
        f.vars.push_back(v);
      }
    else if (tag != 0)
      return false;
    for (const Var &v : f.vars)
      if (v.rec) f.recsize += v.vsize;
    // one record variable alone is not padded to 4 bytes
    int nrec = 0;
    for (const Var &v : f.vars) nrec += v.rec;
    if (nrec == 1)
      for (const Var &v : f.vars)
        if (v.rec) {
          long long n2 = type_size(v.type);
          for (size_t i = 1; i < v.dims.size(); i++) n2 *= f.dims[v.dims[i]].len;
          f.recsize = n2;
        }
    return ok;
  }
};

}  // namespace nc
}  // namespace fv3
