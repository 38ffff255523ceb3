// State files: the netCDF classic layout of io_save_state (io.f90:295-392)
// and the reading of io_load_state / io_load_coll (io.f90:98-232) with the
// resizing in K and M and the radial interpolation (io_interp_wts,
// io_interp, io.f90:238-289).
//
// The reference writes with netCDF-Fortran: nf90_create(..., nf90_clobber)
// gives the classic format (CDF-1).  The file holds
//   global attributes t, Re, alpha (double);
//   dimensions r = N, H = number of harmonics, ReIm = 2;
//   variables r(r), the scalars dt, dtcor, dtcfl, and Ur, Ut, Uz(ReIm, H, r)
//   (C order: the Fortran dims (r, H, ReIm) reversed), each with the int
//   attributes K, M, Mp.
// netCDF itself is not a dependency: the classic format is a documented
// big-endian header followed by the variables' data, restated here (reading
// accepts CDF-1 and CDF-2, writing produces CDF-1 as the reference does).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "opf_internal.hpp"

namespace opf {
namespace {

enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };
enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };

struct NcAtt {
  int type = 0;
  std::vector<double> num;   // numeric values (any numeric type widened)
};

struct NcVar {
  std::string name;
  std::vector<int> dimids;
  std::map<std::string, NcAtt> atts;
  int type = 0;
  uint64_t begin = 0;
};

struct NcFile {
  std::vector<std::pair<std::string, uint64_t>> dims;
  std::map<std::string, NcAtt> gatts;
  std::vector<NcVar> vars;
  std::vector<unsigned char> bytes;
};

int type_size(int t) {
  switch (t) {
    case NC_BYTE: case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT: case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: throw std::runtime_error("state file: unknown netCDF type");
  }
}

// ---- reading ------------------------------------------------------------
struct Reader {
  const std::vector<unsigned char>& b;
  size_t p = 0;
  int version = 1;
  explicit Reader(const std::vector<unsigned char>& bytes) : b(bytes) {}
  void need(size_t n) const {
    if (p + n > b.size()) throw std::runtime_error("state file: truncated header");
  }
  uint32_t u32() {
    need(4);
    const uint32_t v = (uint32_t)b[p] << 24 | (uint32_t)b[p + 1] << 16 | (uint32_t)b[p + 2] << 8 | b[p + 3];
    p += 4;
    return v;
  }
  uint64_t u64() {
    const uint64_t hi = u32();
    return hi << 32 | u32();
  }
  std::string name() {
    const uint32_t n = u32();
    need(n);
    std::string s((const char*)&b[p], n);
    p += (n + 3) / 4 * 4;
    return s;
  }
  double value(int type) {
    const int sz = type_size(type);
    need(sz);
    uint64_t v = 0;
    for (int i = 0; i < sz; ++i) v = v << 8 | b[p + i];
    p += sz;
    switch (type) {
      case NC_BYTE: return (double)(int8_t)v;
      case NC_CHAR: return (double)v;
      case NC_SHORT: return (double)(int16_t)v;
      case NC_INT: return (double)(int32_t)v;
      case NC_FLOAT: { uint32_t w = (uint32_t)v; float f; std::memcpy(&f, &w, 4); return f; }
      default: { double d; std::memcpy(&d, &v, 8); return d; }
    }
  }
  std::map<std::string, NcAtt> atts() {
    std::map<std::string, NcAtt> out;
    const uint32_t tag = u32(), n = u32();
    if (tag == 0 && n == 0) return out;
    if (tag != NC_ATTRIBUTE) throw std::runtime_error("state file: bad attribute list");
    for (uint32_t i = 0; i < n; ++i) {
      const std::string nm = name();
      NcAtt a;
      a.type = (int)u32();
      const uint32_t cnt = u32();
      const size_t start = p;
      for (uint32_t j = 0; j < cnt; ++j) a.num.push_back(value(a.type));
      p = start + ((size_t)cnt * type_size(a.type) + 3) / 4 * 4;
      out[nm] = a;
    }
    return out;
  }
};

NcFile nc_read(const std::string& path) {
  NcFile f;
  FILE* fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("state file not found: " + path);
  std::fseek(fp, 0, SEEK_END);
  const long sz = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  f.bytes.resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = f.bytes.empty() ? 0 : std::fread(f.bytes.data(), 1, f.bytes.size(), fp);
  std::fclose(fp);
  if (got != f.bytes.size() || f.bytes.size() < 8 || std::memcmp(f.bytes.data(), "CDF", 3) != 0)
    throw std::runtime_error("not a netCDF classic file: " + path);
  Reader r(f.bytes);
  r.p = 3;
  r.version = f.bytes[r.p++];
  if (r.version != 1 && r.version != 2) throw std::runtime_error("unsupported netCDF format version: " + path);
  r.u32();   // numrecs (no record variables in a state file)
  {
    const uint32_t tag = r.u32(), n = r.u32();
    if (!(tag == 0 && n == 0)) {
      if (tag != NC_DIMENSION) throw std::runtime_error("state file: bad dimension list");
      for (uint32_t i = 0; i < n; ++i) {
        const std::string nm = r.name();
        f.dims.emplace_back(nm, r.u32());
      }
    }
  }
  f.gatts = r.atts();
  const uint32_t tag = r.u32(), n = r.u32();
  if (!(tag == 0 && n == 0)) {
    if (tag != NC_VARIABLE) throw std::runtime_error("state file: bad variable list");
    for (uint32_t i = 0; i < n; ++i) {
      NcVar v;
      v.name = r.name();
      const uint32_t nd = r.u32();
      for (uint32_t j = 0; j < nd; ++j) v.dimids.push_back((int)r.u32());
      v.atts = r.atts();
      v.type = (int)r.u32();
      r.u32();   // vsize
      v.begin = r.version == 1 ? r.u32() : r.u64();
      f.vars.push_back(v);
    }
  }
  return f;
}

const NcVar* find_var(const NcFile& f, const char* nm) {
  for (const auto& v : f.vars)
    if (v.name == nm) return &v;
  return nullptr;
}

uint64_t var_count(const NcFile& f, const NcVar& v) {
  uint64_t c = 1;
  for (int d : v.dimids) c *= f.dims.at(d).second;
  return c;
}

std::vector<double> var_data(const NcFile& f, const NcVar& v) {
  const uint64_t n = var_count(f, v);
  const int sz = type_size(v.type);
  if (v.begin + n * sz > f.bytes.size()) throw std::runtime_error("state file: variable " + v.name + " truncated");
  Reader r(f.bytes);
  r.p = v.begin;
  std::vector<double> out(n);
  for (uint64_t i = 0; i < n; ++i) out[i] = r.value(v.type);
  return out;
}

double att_num(const std::map<std::string, NcAtt>& a, const char* nm, double dflt) {
  auto it = a.find(nm);
  return (it == a.end() || it->second.num.empty()) ? dflt : it->second.num[0];
}

// ---- writing (CDF-1) ----------------------------------------------------
struct Writer {
  std::vector<unsigned char> b;
  void u32(uint32_t v) {
    for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void f64(double d) {
    uint64_t v;
    std::memcpy(&v, &d, 8);
    for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s));
  }
  void name(const std::string& s) {
    u32((uint32_t)s.size());
    b.insert(b.end(), s.begin(), s.end());
    while (b.size() % 4) b.push_back(0);
  }
  void att_double(const std::string& nm, double v) {
    name(nm);
    u32(NC_DOUBLE);
    u32(1);
    f64(v);
  }
  void att_int(const std::string& nm, int v) {
    name(nm);
    u32(NC_INT);
    u32(1);
    u32((uint32_t)v);
  }
};

}  // namespace

// io_save_state: fields in coll layout (Re/Im planes of N x H, r fastest).
void write_state_file(const std::string& path, int N, int K, int M, int Mp, double t, double Re, double alpha,
                      double dt, double dtcor, double dtcfl, const double* r, const double* const f[6]) {
  const int H = (2 * K - 1) * M - (K - 1);
  struct V { const char* nm; std::vector<int> dims; bool coll; };
  const V vars[] = {{"r", {0}, false}, {"dt", {}, false}, {"dtcor", {}, false}, {"dtcfl", {}, false},
                    {"Ur", {2, 1, 0}, true}, {"Ut", {2, 1, 0}, true}, {"Uz", {2, 1, 0}, true}};
  auto vsize = [&](const V& v) -> uint64_t {
    if (v.dims.empty()) return 8;
    return v.coll ? (uint64_t)2 * H * N * 8 : (uint64_t)N * 8;
  };
  // the header, with begin offsets filled in a second pass once its size is known
  auto header = [&](uint32_t data0) {
    Writer w;
    w.b = {'C', 'D', 'F', 1};
    w.u32(0);   // numrecs
    w.u32(NC_DIMENSION);
    w.u32(3);
    w.name("r"); w.u32((uint32_t)N);
    w.name("H"); w.u32((uint32_t)H);
    w.name("ReIm"); w.u32(2);
    w.u32(NC_ATTRIBUTE);
    w.u32(3);
    w.att_double("t", t);
    w.att_double("Re", Re);
    w.att_double("alpha", alpha);
    w.u32(NC_VARIABLE);
    w.u32(7);
    uint32_t off = data0;
    for (const V& v : vars) {
      w.name(v.nm);
      w.u32((uint32_t)v.dims.size());
      for (int d : v.dims) w.u32((uint32_t)d);
      if (v.coll) {
        w.u32(NC_ATTRIBUTE);
        w.u32(3);
        w.att_int("K", K);
        w.att_int("M", M);
        w.att_int("Mp", Mp);
      } else {
        w.u32(0);
        w.u32(0);
      }
      w.u32(NC_DOUBLE);
      const uint64_t vs = vsize(v);
      if (off + vs > 0x7fffffffu) throw std::runtime_error("state too large for the classic netCDF format");
      w.u32((uint32_t)vs);
      w.u32(off);
      off += (uint32_t)vs;
    }
    return w;
  };
  Writer w = header(0);
  w = header((uint32_t)w.b.size());
  for (int n = 0; n < N; ++n) w.f64(r[n]);
  w.f64(dt);
  w.f64(dtcor);
  w.f64(dtcfl);
  for (int c = 0; c < 3; ++c)
    for (int part = 0; part < 2; ++part)
      for (size_t i = 0; i < (size_t)N * H; ++i) w.f64(f[2 * c + part][i]);
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp) throw std::runtime_error("cannot write state file: " + path);
  const size_t put = std::fwrite(w.b.data(), 1, w.b.size(), fp);
  const int rc = std::fclose(fp);
  if (put != w.b.size() || rc != 0) throw std::runtime_error("short write of state file: " + path);
}

StateInfo read_state_info(const std::string& path) {
  const NcFile f = nc_read(path);
  const NcVar* vr = find_var(f, "r");
  const NcVar* vu = find_var(f, "Ur");
  if (!vr || !vu) throw std::runtime_error("state file without r or Ur: " + path);
  StateInfo s{};
  s.N = (int)var_count(f, *vr);
  s.K = (int)att_num(vu->atts, "K", 0);
  s.M = (int)att_num(vu->atts, "M", 0);
  s.Mp = (int)att_num(vu->atts, "Mp", 1);
  s.t = att_num(f.gatts, "t", 0.0);
  s.Re = att_num(f.gatts, "Re", 0.0);
  s.alpha = att_num(f.gatts, "alpha", 0.0);
  const NcVar* vd = find_var(f, "dt");
  s.dt = vd ? var_data(f, *vd).at(0) : 0.0;
  return s;
}

// io_load_state onto the grid (N, K, M, Mp, r): the file's harmonics (k, m)
// that exist on the grid are copied, the others dropped, and missing ones
// left zero (io_load_coll); a file on other radial points is interpolated
// with the local polynomials of io_interp_wts / io_interp, the field
// continued across the axis with the parity of m*Mp + S (S = 1 for u_r and
// u_theta, 0 for u_z).
void read_state_file(const std::string& path, int N, int K, int M, int Mp, const double* r, double* t,
                     double* Re, double* alpha, double* dt, double* dtcor, double* dtcfl, double* const out[6]) {
  const NcFile f = nc_read(path);
  const NcVar* vr = find_var(f, "r");
  if (!vr) throw std::runtime_error("state file without variable r: " + path);
  const std::vector<double> rf = var_data(f, *vr);
  const int Nf = (int)rf.size();
  if (Nf <= KL) throw std::runtime_error("state file: fewer radial points than the interpolation stencil");
  if (t) *t = att_num(f.gatts, "t", 0.0);
  if (Re) *Re = att_num(f.gatts, "Re", 0.0);
  if (alpha) *alpha = att_num(f.gatts, "alpha", 0.0);
  auto scalar = [&](const char* nm, double dflt) {
    const NcVar* v = find_var(f, nm);
    return v ? var_data(f, *v).at(0) : dflt;
  };
  if (dt) *dt = scalar("dt", 0.0);
  if (dtcor) *dtcor = scalar("dtcor", 0.0);
  if (dtcfl) *dtcfl = scalar("dtcfl", 0.0);

  bool interp = Nf != N;
  if (!interp)
    for (int n = 0; n < N; ++n)
      if (std::fabs(r[n] - rf[n]) > 1e-8) interp = true;
  // xi = r(1-KL:Nf): the file's points continued across the axis
  const int ni = KL + Nf;
  std::vector<double> xi(ni);
  for (int i = 0; i < KL; ++i) xi[i] = -rf[KL - 1 - i];
  for (int n = 0; n < Nf; ++n) xi[KL + n] = rf[n];
  struct Stencil { int l, nn; std::vector<double> w; };
  std::vector<Stencil> st;
  if (interp) {
    st.resize(N);
    for (int n = 0; n < N; ++n) {
      int j = 0;   // io_interp_wts: first xi not below xo(n) - 1e-8
      while (xi[j] < r[n] - 1e-8 && j < ni - 1) ++j;
      const int l = std::max(0, j - KL), rr = std::min(j + KL, ni - 1), nn = rr - l + 1;
      std::vector<double> A((size_t)nn * nn);
      for (int i = 0; i < nn; ++i) A[(size_t)i * nn] = 1.0;
      for (int c = 1; c < nn; ++c)
        for (int i = 0; i < nn; ++i)
          A[(size_t)i * nn + c] = A[(size_t)i * nn + c - 1] * (xi[l + i] - r[n]) / (double)c;
      mat_invert(A, nn);
      st[n] = Stencil{l, nn, std::vector<double>(A.begin(), A.begin() + nn)};   // row 1
    }
  }

  const int H = (2 * K - 1) * M - (K - 1);
  const char* names[3] = {"Ur", "Ut", "Uz"};
  const int S[3] = {1, 1, 0};
  for (int c = 0; c < 3; ++c) {
    const NcVar* v = find_var(f, names[c]);
    if (!v) throw std::runtime_error(std::string("state file without field ") + names[c]);
    const int Kf = (int)att_num(v->atts, "K", 0), Mf = (int)att_num(v->atts, "M", 0);
    const int Mpf = (int)att_num(v->atts, "Mp", Mp);
    if (Mpf != Mp && Mf > 2) throw std::runtime_error("io_load_coll: Mp /= i_Mp");
    const int Hf = (2 * Kf - 1) * Mf - (Kf - 1);
    if (v->dimids.size() != 3 || var_count(f, *v) != (uint64_t)2 * Hf * Nf)
      throw std::runtime_error(std::string("state file: field ") + names[c] + " has the wrong shape");
    const std::vector<double> data = var_data(f, *v);   // [ReIm][Hf][Nf]
    double* re = out[2 * c];
    double* im = out[2 * c + 1];
    std::fill(re, re + (size_t)N * H, 0.0);
    std::fill(im, im + (size_t)N * H, 0.0);
    const int K1 = std::max(Kf, K) - 1, M1 = std::max(Mf, M) - 1;
    int nh = -1, nhf = -1;
    std::vector<double> fn(ni);
    // nhf counts the file's harmonics in the file's own order.  io.f90:205
    // also counts a harmonic that neither the grid nor the file holds, which
    // only happens when the file is larger in one of K, M and smaller in the
    // other; such files are read by their true order here.
    for (int m = 0; m <= M1; ++m)
      for (int k = -K1; k <= K1; ++k) {
        if (k < 0 && m == 0) continue;
        const bool in_file = std::abs(k) < Kf && m < Mf;
        if (in_file) ++nhf;
        if (std::abs(k) >= K || m >= M) continue;   // not on the grid: dropped
        ++nh;
        if (!in_file) continue;                     // not in the file: stays zero
        for (int part = 0; part < 2; ++part) {
          const double* src = &data[((size_t)part * Hf + nhf) * Nf];
          double* dst = (part ? im : re) + (size_t)nh * N;
          if (!interp) {
            for (int n = 0; n < N; ++n) dst[n] = src[n];
            continue;
          }
          for (int n = 0; n < Nf; ++n) fn[KL + n] = src[n];
          const bool odd = ((m * Mp + S[c]) % 2) == 1;
          for (int i = 0; i < KL; ++i) fn[i] = odd ? -src[KL - 1 - i] : src[KL - 1 - i];
          for (int n = 0; n < N; ++n) {
            double s = 0.0;   // dot_product(A(1,1:nn,n), fi(l:r))
            for (int i = 0; i < st[n].nn; ++i) s = s + st[n].w[i] * fn[st[n].l + i];
            dst[n] = s;
          }
        }
      }
  }
}

}  // namespace opf
