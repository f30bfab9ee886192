// les_io.hip -- the reference's run files on the host side of the library:
//   * Fortran sequential unformatted records (checkpoint, io.f90:1314-1338;
//     the initial-condition reads of initial.f90:43-175): each record is
//     framed by 4-byte little-endian byte counts; a record longer than
//     2^31 - 9 bytes is split into subrecords, the leading count negative
//     while more subrecords follow and the trailing one negative after the
//     first (the convention of the gfortran / flang runtimes);
//   * the state checkpoint and initial-condition reader built on them, with
//     the layouts the reference uses (including initial.f90's model-7 list,
//     which does not match what checkpoint writes);
//   * total_time.dat (param.f90:200-210, io.f90:1375-1379), the surface
//     files wts.dat / wqs.dat / zom.dat (bottombc.f90:23-57) and the
//     param_nml namelist of input.nml (param.f90:174-230).
// Nothing here launches a kernel; the state entry points copy device fields
// through host buffers.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/les_amd.h"
#include "les_internal.h"

namespace {

constexpr int64_t MAX_SUB = 2147483639;  // largest subrecord payload

bool put_i32(FILE* f, int32_t v) {
  unsigned char b[4] = {(unsigned char)(v & 0xff), (unsigned char)((v >> 8) & 0xff), (unsigned char)((v >> 16) & 0xff),
                        (unsigned char)((v >> 24) & 0xff)};
  return fwrite(b, 1, 4, f) == 4;
}
bool get_i32(FILE* f, int32_t* v) {
  unsigned char b[4];
  if (fread(b, 1, 4, f) != 4) return false;
  *v = (int32_t)((uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24));
  return true;
}

struct Seg {
  const void* p;
  uint64_t n;
};

// one logical record made of the concatenated segments
bool write_record(FILE* f, const std::vector<Seg>& segs) {
  uint64_t total = 0;
  for (const Seg& s : segs) total += s.n;
  size_t si = 0;
  uint64_t off = 0;  // offset inside segs[si]
  uint64_t left = total;
  bool first = true;
  do {
    const int64_t chunk = (int64_t)(left < (uint64_t)MAX_SUB ? left : (uint64_t)MAX_SUB);
    const bool more = (uint64_t)chunk < left;
    if (!put_i32(f, (int32_t)(more ? -chunk : chunk))) return false;
    int64_t need = chunk;
    while (need > 0) {
      const uint64_t have = segs[si].n - off, take = have < (uint64_t)need ? have : (uint64_t)need;
      if (take && fwrite((const char*)segs[si].p + off, 1, take, f) != take) return false;
      off += take;
      need -= (int64_t)take;
      if (off == segs[si].n) {
        ++si;
        off = 0;
      }
    }
    if (!put_i32(f, (int32_t)(first ? chunk : -chunk))) return false;
    left -= (uint64_t)chunk;
    first = false;
  } while (left > 0);
  return true;
}

// the next logical record; false at end of file or on a malformed frame
bool read_record(FILE* f, std::vector<unsigned char>& out, std::string* err) {
  out.clear();
  bool first = true;
  for (;;) {
    int32_t head;
    if (!get_i32(f, &head)) {
      if (!first && err) *err = "record: end of file inside a record";
      return false;
    }
    const bool more = head < 0;
    const int64_t n = more ? -(int64_t)head : head;
    const size_t at = out.size();
    out.resize(at + (size_t)n);
    if (n && fread(out.data() + at, 1, (size_t)n, f) != (size_t)n) {
      if (err) *err = "record: file ends before the record's byte count";
      return false;
    }
    int32_t tail;
    if (!get_i32(f, &tail) || (tail < 0 ? -(int64_t)tail : tail) != n) {
      if (err) *err = "record: trailing byte count does not match the leading one";
      return false;
    }
    if (!more) return true;
    first = false;
  }
}

// ---- state fields in the checkpoint / initial-condition layouts ----
enum Kind { F3, PLANE_LD, PLANE_NX, SKIP3 };  // (ld,ny,nz) | (ld,ny) | (nx,ny) | (ld,ny,nz) read and dropped
struct Item {
  const char* name;
  Kind kind;
};

uint64_t item_doubles(const les_ctx* c, Kind k) {
  const uint64_t ld = c->g.ld, ny = c->g.ny, nz = c->g.nz, nx = c->g.nx;
  return k == F3 || k == SKIP3 ? ld * ny * nz : (k == PLANE_LD ? ld * ny : nx * ny);
}

// io.f90:1314-1338
std::vector<Item> checkpoint_layout(bool scal) {
  if (scal)
    return {{"u", F3},        {"v", F3},        {"w", F3},        {"theta", F3},    {"qmix", F3},   {"RHSx", F3},
            {"RHSy", F3},     {"RHSz", F3},     {"RHS_T", F3},    {"RHS_q", F3},    {"sgs_t3", PLANE_LD},
            {"sgs_q3", PLANE_LD}, {"psi_m", PLANE_NX}, {"psi_m0", PLANE_NX}, {"psi_h", PLANE_NX},
            {"psi_h0", PLANE_NX}, {"Cs_opt2", F3}, {"F_LM", F3}, {"F_MM", F3}, {"F_QN", F3}, {"F_NN", F3},
            {"G_LM", F3},     {"G_MM", F3},     {"G_QN", F3},     {"G_NN", F3},     {"Pr_t", F3}};
  return {{"u", F3},    {"v", F3},    {"w", F3},    {"RHSx", F3}, {"RHSy", F3}, {"RHSz", F3},
          {"Cs_opt2", F3}, {"F_LM", F3}, {"F_MM", F3}, {"F_QN", F3}, {"F_NN", F3}};
}

// initial.f90:46-175 (initu true): the read list for model, initsc, inilag
bool initial_layout(int model, bool initsc, bool inilag, std::vector<Item>& L) {
  const std::vector<Item> uvw = {{"u", F3}, {"v", F3}, {"w", F3}};
  const std::vector<Item> rhs = {{"RHSx", F3}, {"RHSy", F3}, {"RHSz", F3}};
  const std::vector<Item> FF = {{"F_LM", F3}, {"F_MM", F3}, {"F_QN", F3}, {"F_NN", F3}};
  const std::vector<Item> GG = {{"G_LM", F3}, {"G_MM", F3}, {"G_QN", F3}, {"G_NN", F3}};
  auto add = [&](const std::vector<Item>& v) { L.insert(L.end(), v.begin(), v.end()); };
  L.clear();
  if (initsc) {
    if (model == 5) {  // :80-93
      add(uvw);
      add({{"theta", F3}, {"qmix", F3}});
      add(rhs);
      add({{"RHS_T", F3}, {"RHS_q", F3}, {"sgs_t3", PLANE_LD}, {"sgs_q3", PLANE_LD}, {"psi_m", PLANE_NX}});
      if (inilag) {
        add({{"Cs_opt2", F3}});
      } else {
        add({{"psi_m0", PLANE_NX}, {"psi_h", PLANE_NX}, {"psi_h0", PLANE_NX}, {"Cs_opt2", F3}});
        add(FF);
        add(GG);
        add({{"Pr_t", F3}});
      }
      return true;
    }
    if (model < 1 || model > 7) return false;
    add(uvw);
    add({{"theta", F3}});
    add(rhs);
    add({{"RHS_T", F3}, {"sgs_t3", PLANE_LD}, {"psi_m", PLANE_NX}});  // case 1, :55-57
    if (model == 1) return true;
    add({{"Cs_opt2", F3}});  // cases 2:3, 4 with inilag, 6, 7 with inilag
    if (model == 4 && !inilag) add({{"F_LM", F3}, {"F_MM", F3}, {"crap", SKIP3}, {"crap", SKIP3}});  // :70-73
    if (model == 7 && !inilag) {  // :104-114
      add(FF);
      add(GG);
      add({{"Pr_t", F3}});
    }
    return true;
  }
  if (model < 1 || model > 7) return false;
  add(uvw);  // :135-170
  add(rhs);
  if (model == 1) return true;
  add({{"Cs_opt2", F3}});
  if ((model == 5 || model == 7) && !inilag) add(FF);
  return true;
}

bool is_scalar_field(const char* n) { return !strcmp(n, "theta") || !strcmp(n, "qmix"); }

// ---- free-format (list-directed) reading ----
struct Lex {
  FILE* f;
  int pushed = -1;
  int get() {
    if (pushed >= 0) {
      const int c = pushed;
      pushed = -1;
      return c;
    }
    return fgetc(f);
  }
  void unget(int c) { pushed = c; }
  // next value token of a list-directed record, or "" at end of file
  std::string token(bool stop_at_newline = false) {
    int c;
    for (;;) {
      c = get();
      if (c == EOF) return "";
      if (c == '\n' && stop_at_newline) return "\n";
      if (!isspace(c) && c != ',') break;
    }
    std::string s(1, (char)c);
    for (;;) {
      c = get();
      if (c == EOF) break;
      if (isspace(c) || c == ',' || c == '/') {
        unget(c);
        break;
      }
      s += (char)c;
    }
    return s;
  }
  void skip_line() {
    int c;
    while ((c = get()) != EOF && c != '\n') {
    }
  }
};

// a Fortran real literal: D or d exponents as E
bool parse_real(std::string s, double* v) {
  for (char& ch : s)
    if (ch == 'd' || ch == 'D') ch = 'E';
  char* end = nullptr;
  *v = strtod(s.c_str(), &end);
  return end && *end == '\0' && !s.empty();
}

// n values of one list-directed read (r*c repeat counts allowed); the next
// read starts on a new line, as after a Fortran READ
bool read_values(Lex& lx, int n, std::vector<double>& out) {
  out.clear();
  while ((int)out.size() < n) {
    const std::string t = lx.token();
    if (t.empty()) return false;
    const size_t star = t.find('*');
    double v;
    if (star != std::string::npos) {
      const int r = atoi(t.substr(0, star).c_str());
      if (r <= 0 || !parse_real(t.substr(star + 1), &v)) return false;
      for (int i = 0; i < r; ++i) out.push_back(v);
    } else {
      if (!parse_real(t, &v)) return false;
      out.push_back(v);
    }
  }
  out.resize(n);
  lx.skip_line();
  return true;
}

std::string lower(std::string s) {
  for (char& ch : s) ch = (char)tolower((unsigned char)ch);
  return s;
}

// key = value pairs of namelist group `group` (case-insensitive); values are
// kept as their literal text, one per key (last one wins)
bool read_namelist(const char* path, const std::string& group, std::vector<std::pair<std::string, std::string>>& kv,
                   std::string* err) {
  FILE* f = fopen(path, "r");
  if (!f) {
    if (err) *err = std::string("namelist: cannot open ") + path;
    return false;
  }
  std::string text;
  int ch;
  while ((ch = fgetc(f)) != EOF) text += (char)ch;
  fclose(f);
  const std::string lt = lower(text), g = "&" + lower(group);
  size_t at = 0;
  for (;;) {
    at = lt.find(g, at);
    if (at == std::string::npos) {
      if (err) *err = "namelist: no group &" + group + " in " + path;
      return false;
    }
    const size_t after = at + g.size();
    if (after >= lt.size() || isspace((unsigned char)lt[after])) break;
    at = after;
  }
  size_t i = at + g.size();
  auto skip_sep = [&]() {
    while (i < text.size() && (isspace((unsigned char)text[i]) || text[i] == ',')) ++i;
  };
  for (;;) {
    skip_sep();
    if (i >= text.size() || text[i] == '/' || text[i] == '&') return true;  // '/' or &end
    if (text[i] == '!') {  // comment to end of line
      while (i < text.size() && text[i] != '\n') ++i;
      continue;
    }
    size_t k0 = i;
    while (i < text.size() && (isalnum((unsigned char)text[i]) || text[i] == '_')) ++i;
    const std::string key = lower(text.substr(k0, i - k0));
    while (i < text.size() && isspace((unsigned char)text[i])) ++i;
    if (key.empty() || i >= text.size() || text[i] != '=') {
      if (err) *err = "namelist: expected 'name =' in &" + group;
      return false;
    }
    ++i;
    while (i < text.size() && isspace((unsigned char)text[i])) ++i;
    size_t v0 = i;
    if (i < text.size() && (text[i] == '\'' || text[i] == '"')) {
      const char q = text[i++];
      while (i < text.size() && text[i] != q) ++i;
      ++i;
    } else {
      while (i < text.size() && !isspace((unsigned char)text[i]) && text[i] != ',' && text[i] != '/') ++i;
    }
    kv.emplace_back(key, text.substr(v0, i - v0));
  }
}

bool nml_logical(const std::string& v, int* out) {
  std::string s = lower(v);
  if (!s.empty() && s[0] == '.') s = s.substr(1);
  if (s.empty()) return false;
  if (s[0] == 't') *out = 1;
  else if (s[0] == 'f') *out = 0;
  else return false;
  return true;
}

}  // namespace

using namespace les;

extern "C" {

int les_record_write(const char* path, int append, int nseg, const void* const* seg, const uint64_t* bytes) {
  if (!path || nseg < 0 || (nseg && (!seg || !bytes))) return -1;
  FILE* f = fopen(path, append ? "ab" : "wb");
  if (!f) return 1;
  std::vector<Seg> s;
  for (int i = 0; i < nseg; ++i) s.push_back(Seg{seg[i], bytes[i]});
  const bool ok = write_record(f, s);
  return fclose(f) == 0 && ok ? 0 : 1;
}

int les_record_read(const char* path, int index, void* buf, uint64_t buf_bytes, uint64_t* record_bytes) {
  if (!path || index < 0) return -1;
  FILE* f = fopen(path, "rb");
  if (!f) return 1;
  std::vector<unsigned char> r;
  std::string err;
  bool ok = true;
  for (int i = 0; i <= index && ok; ++i) ok = read_record(f, r, &err);
  fclose(f);
  if (!ok) return 2;
  if (record_bytes) *record_bytes = r.size();
  if (buf) memcpy(buf, r.data(), r.size() < buf_bytes ? r.size() : buf_bytes);
  return 0;
}

int les_state_checkpoint(les_ctx* c, const char* path) {
  if (!c) return -1;
  c->err.clear();
  if (!c->st || !path) {
    c->err = "checkpoint: needs a state and a path";
    return 1;
  }
  const bool scal = les_state_field(c, "theta", nullptr, nullptr) != nullptr;
  const std::vector<Item> L = checkpoint_layout(scal);
  double tref = 0.0, qref = 0.0;
  les_scalar_ref(c, &tref, &qref);
  c->err.clear();
  std::vector<std::vector<double>> host(L.size());
  std::vector<Seg> segs;
  if (les_sync(c) != 0) return 1;
  for (size_t i = 0; i < L.size(); ++i) {
    int nlev = 0, es = 0;
    void* d = les_state_field(c, L[i].name, &nlev, &es);
    const uint64_t n = item_doubles(c, L[i].kind);
    if (!d) {
      c->err = std::string("checkpoint: the state has no ") + L[i].name;
      return 1;
    }
    host[i].resize(n);
    if (es == 8) {
      if (hipMemcpy(host[i].data(), d, n * 8, hipMemcpyDeviceToHost) != hipSuccess) {
        c->err = "checkpoint: device to host copy failed";
        return 1;
      }
    } else {  // FP32-bulk field: widened to the reference's real(8)
      std::vector<float> tmp(n);
      if (hipMemcpy(tmp.data(), d, n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        c->err = "checkpoint: device to host copy failed";
        return 1;
      }
      const double ref = !strcmp(L[i].name, "theta") ? tref : (!strcmp(L[i].name, "qmix") ? qref : 0.0);
      const uint64_t pl = (uint64_t)c->g.ld * c->g.ny;
      for (uint64_t j = 0; j < n; ++j) {
        const bool pad = (j % (uint64_t)c->g.ld) >= (uint64_t)c->g.nx;
        host[i][j] = (double)tmp[j] + (is_scalar_field(L[i].name) && !pad ? ref : 0.0);
      }
      (void)pl;
    }
    segs.push_back(Seg{host[i].data(), n * 8});
  }
  FILE* f = fopen(path, "wb");
  if (!f) {
    c->err = std::string("checkpoint: cannot open ") + path;
    return 1;
  }
  const bool ok = write_record(f, segs);
  if (fclose(f) != 0 || !ok) {
    c->err = std::string("checkpoint: write failed: ") + path;
    return 1;
  }
  return 0;
}

int les_state_read_initial(les_ctx* c, const char* path, int model, int initsc, int inilag) {
  if (!c) return -1;
  c->err.clear();
  if (!c->st || !path) {
    c->err = "read_initial: needs a state and a path";
    return 1;
  }
  std::vector<Item> L;
  if (!initial_layout(model, initsc != 0, inilag != 0, L)) {
    c->err = "read_initial: invalid model number";  // initial.f90:116,172
    return 2;
  }
  FILE* f = fopen(path, "rb");
  if (!f) {
    c->err = std::string("read_initial: cannot open ") + path;
    return 1;
  }
  std::vector<unsigned char> rec;
  std::string err;
  const bool ok = read_record(f, rec, &err);
  fclose(f);
  if (!ok) {
    c->err = "read_initial: " + (err.empty() ? std::string("empty file") : err);
    return 1;
  }
  uint64_t need = 0;
  for (const Item& it : L) need += 8 * item_doubles(c, it.kind);
  if (need > rec.size()) {
    c->err = "read_initial: the record holds " + std::to_string(rec.size()) + " bytes, the read list needs " +
             std::to_string(need);
    return 3;
  }
  double tref = 0.0, qref = 0.0;
  les_scalar_ref(c, &tref, &qref);
  c->err.clear();
  if (les_sync(c) != 0) return 1;
  uint64_t off = 0;
  for (const Item& it : L) {
    const uint64_t n = item_doubles(c, it.kind);
    const double* src = (const double*)(rec.data() + off);
    off += 8 * n;
    if (it.kind == SKIP3) continue;
    int nlev = 0, es = 0;
    void* d = les_state_field(c, it.name, &nlev, &es);
    if (!d) {
      c->err = std::string("read_initial: the layout reads ") + it.name + ", which this state does not hold";
      return 1;
    }
    hipError_t e;
    if (es == 8) {
      e = hipMemcpy(d, src, n * 8, hipMemcpyHostToDevice);
    } else {
      const double ref = !strcmp(it.name, "theta") ? tref : (!strcmp(it.name, "qmix") ? qref : 0.0);
      std::vector<float> tmp(n);
      for (uint64_t j = 0; j < n; ++j) {
        const bool pad = (j % (uint64_t)c->g.ld) >= (uint64_t)c->g.nx;
        tmp[j] = (float)(src[j] - (is_scalar_field(it.name) && !pad ? ref : 0.0));
      }
      e = hipMemcpy(d, tmp.data(), n * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
      c->err = "read_initial: host to device copy failed";
      return 1;
    }
  }
  return 0;
}

int les_read_total_time(const char* path, int* jt_total) {
  if (!path || !jt_total) return -1;
  FILE* f = fopen(path, "r");
  if (!f) return 1;  // the reference then assumes jt_total = 0
  const int n = fscanf(f, "%d", jt_total);
  fclose(f);
  return n == 1 ? 0 : 2;
}

int les_write_total_time(const char* path, int jt_total) {
  if (!path) return -1;
  FILE* f = fopen(path, "w");
  if (!f) return 1;
  fprintf(f, "%12d\n", jt_total);  // list-directed integer, read back free-format
  return fclose(f) == 0 ? 0 : 1;
}

int les_read_surface(const char* dir, int nx, int ny, double z_i, double* wt_s, double* wq_s, double* zo) {
  if (!dir || nx <= 0 || ny <= 0 || !wt_s || !wq_s || !zo) return -1;
  const char* names[3] = {"wts.dat", "wqs.dat", "zom.dat"};
  double* out[3] = {wt_s, wq_s, zo};
  for (int k = 0; k < 3; ++k) {
    const std::string p = std::string(dir) + "/" + names[k];
    FILE* f = fopen(p.c_str(), "r");
    if (!f) return 1;
    Lex lx{f};
    std::vector<double> row;
    for (int j = 0; j < ny; ++j) {  // one READ per row y, nx values (bottombc.f90:35-39)
      if (!read_values(lx, nx, row)) {
        fclose(f);
        return 2;
      }
      for (int i = 0; i < nx; ++i) out[k][(size_t)j * nx + i] = row[i];
    }
    fclose(f);
  }
  // bottombc.f90:50-57 (param.f90:90-94: rho_d, cp, Le)
  const double rho_d = 1.2923, cp = 1005.0, Le = (double)2.5008E06f;
  for (size_t i = 0; i < (size_t)nx * ny; ++i) {
    zo[i] = zo[i] / z_i;
    wt_s[i] = wt_s[i] / (rho_d * cp);
    wq_s[i] = 1000.0 * wq_s[i] / (rho_d * Le);
  }
  return 0;
}

int les_read_nml_value(const char* path, const char* group, const char* key, double* value) {
  if (!path || !group || !key || !value) return -1;
  std::vector<std::pair<std::string, std::string>> kv;
  std::string err;
  if (!read_namelist(path, group, kv, &err)) return 1;
  const std::string k = lower(key);
  for (auto it = kv.rbegin(); it != kv.rend(); ++it)
    if (it->first == k) {
      int b;
      if (parse_real(it->second, value)) return 0;
      if (nml_logical(it->second, &b)) {
        *value = b;
        return 0;
      }
      return 3;
    }
  return 2;
}

int les_read_param_nml(const char* nml_path, const char* total_time_path, les_param_nml* p) {
  if (!p) return -1;
  // param.f90:63-85 defaults
  les_param_nml d{};
  d.q_scale = 100.0;
  d.T_scale = 400.0;
  d.u_star = 0.0;
  d.ug0 = 10.0;
  d.vg0 = 1.0;
  d.inv_strength = 0.0050;
  d.inv_strength_q = -0.004;
  d.dTdz_top = 0.0050;
  d.dqdz_top = -0.004;
  d.T_init = 283.00;
  d.q_init = 1.00;
  d.dt_dim = 0.1;
  d.z_i = 1000.0;
  d.coriolis_forcing = 0;
  d.ubc = 0;
  d.damping_method = 1;
  d.initu = 1;
  d.initsc = 1;
  d.inilag = 0;
  *p = d;
  int rc = 0;
  if (nml_path) {
    std::vector<std::pair<std::string, std::string>> kv;
    std::string err;
    if (!read_namelist(nml_path, "param_nml", kv, &err)) return 1;
    struct R {
      const char* k;
      double* v;
    } reals[] = {{"q_scale", &p->q_scale}, {"t_scale", &p->T_scale}, {"u_star", &p->u_star}, {"ug0", &p->ug0},
                 {"vg0", &p->vg0}, {"inv_strength", &p->inv_strength}, {"inv_strength_q", &p->inv_strength_q},
                 {"dtdz_top", &p->dTdz_top}, {"dqdz_top", &p->dqdz_top}, {"t_init", &p->T_init},
                 {"q_init", &p->q_init}, {"dt_dim", &p->dt_dim}};
    struct I {
      const char* k;
      int* v;
      bool logical;
    } ints[] = {{"coriolis_forcing", &p->coriolis_forcing, true}, {"nsteps", &p->nsteps, false},
                {"c_count", &p->c_count, false}, {"p_count", &p->p_count, false}, {"ubc", &p->ubc, false},
                {"damping_method", &p->damping_method, false}, {"initu", &p->initu, true},
                {"initsc", &p->initsc, true}, {"inilag", &p->inilag, true}};
    for (const auto& e : kv) {
      bool known = false;
      for (auto& r : reals)
        if (e.first == r.k) {
          known = parse_real(e.second, r.v);
          break;
        }
      for (auto& q : ints)
        if (e.first == q.k) {
          if (q.logical) {
            known = nml_logical(e.second, q.v);
          } else {
            double x;
            known = parse_real(e.second, &x) && x == std::floor(x);
            if (known) *q.v = (int)x;
          }
          break;
        }
      if (!known) {  // a Fortran namelist READ stops at the first bad item, keeping what it read
        rc = 3;
        break;
      }
    }
  }
  p->jt_total = 0;
  if (total_time_path && les_read_total_time(total_time_path, &p->jt_total) != 0) p->jt_total = 0;
  if (p->jt_total > 0) {  // param.f90:222-231: a continued run reads the previous checkpoint
    p->initu = 1;
    p->initsc = 1;
    p->inilag = 0;
  }
  p->dt = p->dt_dim * p->u_star / p->z_i;  // param.f90:219
  return rc;
}

}  // extern "C"
