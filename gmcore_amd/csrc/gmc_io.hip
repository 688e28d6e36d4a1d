// gmc_io.hip -- h0 history and r0 restart files of the shallow-water path in
// the netCDF classic format, 64-bit offset variant (include/gmcore_io.h).
//
// The format (netCDF "classic" file format specification): a big-endian
// header -- magic "CDF\x02", the record count, the dimension list, the
// global attributes, the variable list (name, dimension ids, attributes,
// type, byte size, 64-bit offset) -- then the fixed-size variables one after
// another, then the records, each record holding the slice of every record
// variable in header order.  Only the pieces the h0 and r0 layouts use are
// written (char attributes and double attributes; float and double
// variables); the reader accepts both offset variants ("CDF\x01" with 32-bit
// offsets too) and any classic file that holds the named variables.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gmc_internal.hpp"
#include "../../include/gmcore_io.h"

namespace gmc {
namespace nc {

enum { NC_CHAR = 2, NC_FLOAT = 5, NC_DOUBLE = 6 };
enum { TAG_DIM = 0x0A, TAG_VAR = 0x0B, TAG_ATT = 0x0C };

struct Att {
  std::string name;
  int type = NC_CHAR;
  std::string text;           // NC_CHAR
  std::vector<double> vals;   // NC_DOUBLE
};
struct Dim {
  std::string name;
  uint64_t len = 0;           // 0: the record (unlimited) dimension
};
struct Var {
  std::string name;
  std::vector<int> dims;
  std::vector<Att> atts;
  int type = NC_DOUBLE;
  uint64_t vsize = 0;         // bytes of the variable, or of one record slice
  uint64_t begin = 0;
  bool record = false;
};
struct File {
  int version = 2;
  uint64_t numrecs = 0;
  std::vector<Dim> dims;
  std::vector<Att> gatts;
  std::vector<Var> vars;
  uint64_t recsize = 0;       // bytes of one record
  uint64_t rec0 = 0;          // offset of record 0
};

// ---- big-endian encoding ----------------------------------------------------
struct Out {
  std::vector<unsigned char> b;
  void u32(uint64_t v) { for (int s = 24; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
  void u64(uint64_t v) { for (int s = 56; s >= 0; s -= 8) b.push_back((unsigned char)(v >> s)); }
  void pad() { while (b.size() % 4) b.push_back(0); }
  void name(const std::string &s) {
    u32(s.size());
    b.insert(b.end(), s.begin(), s.end());
    pad();
  }
  void f64(double d) {
    uint64_t v;
    std::memcpy(&v, &d, 8);
    u64(v);
  }
  void f32(float f) {
    uint32_t v;
    std::memcpy(&v, &f, 4);
    u32(v);
  }
};

int type_size(int t) { return t == NC_DOUBLE ? 8 : (t == NC_FLOAT ? 4 : 1); }

void put_atts(Out &o, const std::vector<Att> &atts) {
  if (atts.empty()) { o.u32(0); o.u32(0); return; }
  o.u32(TAG_ATT);
  o.u32(atts.size());
  for (const Att &a : atts) {
    o.name(a.name);
    o.u32(a.type);
    if (a.type == NC_CHAR) {
      o.u32(a.text.size());
      o.b.insert(o.b.end(), a.text.begin(), a.text.end());
      o.pad();
    } else {
      o.u32(a.vals.size());
      for (double d : a.vals) o.f64(d);
    }
  }
}

// header bytes; the offsets are 64-bit (version 2) or 32-bit (version 1)
std::vector<unsigned char> header(const File &f) {
  Out o;
  o.b = {'C', 'D', 'F', (unsigned char)f.version};
  o.u32(f.numrecs);
  if (f.dims.empty()) { o.u32(0); o.u32(0); }
  else {
    o.u32(TAG_DIM);
    o.u32(f.dims.size());
    for (const Dim &d : f.dims) { o.name(d.name); o.u32(d.len); }
  }
  put_atts(o, f.gatts);
  if (f.vars.empty()) { o.u32(0); o.u32(0); }
  else {
    o.u32(TAG_VAR);
    o.u32(f.vars.size());
    for (const Var &v : f.vars) {
      o.name(v.name);
      o.u32(v.dims.size());
      for (int d : v.dims) o.u32(d);
      put_atts(o, v.atts);
      o.u32(v.type);
      o.u32(v.vsize);
      if (f.version == 2) o.u64(v.begin);
      else o.u32(v.begin);
    }
  }
  return o.b;
}

// sizes and offsets of every variable (the header length does not depend
// on the offsets' values)
void layout(File &f) {
  for (Var &v : f.vars) {
    uint64_t n = type_size(v.type);
    v.record = false;
    for (int d : v.dims) {
      if (f.dims[d].len == 0) v.record = true;
      else n *= f.dims[d].len;
    }
    v.vsize = (n + 3) / 4 * 4;
  }
  uint64_t off = header(f).size();
  for (Var &v : f.vars)
    if (!v.record) { v.begin = off; off += v.vsize; }
  f.rec0 = off;
  f.recsize = 0;
  for (Var &v : f.vars)
    if (v.record) { v.begin = off + f.recsize; f.recsize += v.vsize; }
}

// ---- decoding ----------------------------------------------------------------
struct In {
  const std::vector<unsigned char> &b;
  size_t p = 0;
  bool bad = false;
  explicit In(const std::vector<unsigned char> &buf) : b(buf) {}
  uint64_t u32() {
    if (p + 4 > b.size()) { bad = true; return 0; }
    uint64_t v = 0;
    for (int k = 0; k < 4; k++) v = v << 8 | b[p++];
    return v;
  }
  uint64_t u64() { const uint64_t hi = u32(); return hi << 32 | u32(); }
  std::string name() {
    const uint64_t n = u32();
    if (bad || p + n > b.size()) { bad = true; return ""; }
    std::string s(b.begin() + p, b.begin() + p + n);
    p += (n + 3) / 4 * 4;
    return s;
  }
};

bool get_atts(In &in, std::vector<Att> &atts) {
  const uint64_t tag = in.u32(), n = in.u32();
  if (tag == 0 && n == 0) return !in.bad;
  if (tag != TAG_ATT) return false;
  for (uint64_t k = 0; k < n && !in.bad; k++) {
    Att a;
    a.name = in.name();
    a.type = (int)in.u32();
    const uint64_t cnt = in.u32();
    const int sz = a.type == NC_DOUBLE ? 8 : a.type == NC_FLOAT || a.type == 4 ? 4 : a.type == 3 ? 2 : 1;
    if (in.p + cnt * sz > in.b.size()) return false;
    if (a.type == NC_CHAR) a.text.assign(in.b.begin() + in.p, in.b.begin() + in.p + cnt);
    in.p += (cnt * sz + 3) / 4 * 4;
    atts.push_back(a);
  }
  return !in.bad;
}

bool parse(const std::vector<unsigned char> &buf, File &f) {
  if (buf.size() < 8 || buf[0] != 'C' || buf[1] != 'D' || buf[2] != 'F' || (buf[3] != 1 && buf[3] != 2)) return false;
  In in(buf);
  in.p = 4;
  f.version = buf[3];
  f.numrecs = in.u32();
  uint64_t tag = in.u32(), n = in.u32();
  if (tag == TAG_DIM)
    for (uint64_t k = 0; k < n && !in.bad; k++) {
      Dim d;
      d.name = in.name();
      d.len = in.u32();
      f.dims.push_back(d);
    }
  else if (tag != 0 || n != 0) return false;
  if (!get_atts(in, f.gatts)) return false;
  tag = in.u32();
  n = in.u32();
  if (tag == TAG_VAR)
    for (uint64_t k = 0; k < n && !in.bad; k++) {
      Var v;
      v.name = in.name();
      const uint64_t nd = in.u32();
      for (uint64_t j = 0; j < nd; j++) {
        const uint64_t id = in.u32();
        if (id >= f.dims.size()) return false;
        v.dims.push_back((int)id);
        if (f.dims[id].len == 0) v.record = true;
      }
      if (!get_atts(in, v.atts)) return false;
      v.type = (int)in.u32();
      v.vsize = in.u32();
      v.begin = f.version == 2 ? in.u64() : in.u32();
      f.vars.push_back(v);
    }
  else if (tag != 0 || n != 0) return false;
  if (in.bad) return false;
  f.recsize = 0;
  int nrec = 0;
  for (const Var &v : f.vars)
    if (v.record) { f.recsize += v.vsize; nrec++; }
  // one record variable: its record slice is not padded
  if (nrec == 1)
    for (const Var &v : f.vars)
      if (v.record) {
        uint64_t s = type_size(v.type);
        for (int d : v.dims) if (f.dims[d].len) s *= f.dims[d].len;
        f.recsize = s;
      }
  return true;
}

// ---- values ------------------------------------------------------------------
void encode(Out &o, int type, const double *x, size_t n) {
  for (size_t k = 0; k < n; k++) {
    if (type == NC_DOUBLE) o.f64(x[k]);
    else o.f32((float)x[k]);
  }
  o.pad();
}

bool decode(const std::vector<unsigned char> &buf, uint64_t off, int type, double *x, size_t n) {
  const int sz = type_size(type);
  if ((type != NC_DOUBLE && type != NC_FLOAT) || off + (uint64_t)n * sz > buf.size()) return false;
  for (size_t k = 0; k < n; k++) {
    const unsigned char *p = buf.data() + off + k * sz;
    if (type == NC_DOUBLE) {
      uint64_t v = 0;
      for (int j = 0; j < 8; j++) v = v << 8 | p[j];
      std::memcpy(&x[k], &v, 8);
    } else {
      uint32_t v = 0;
      for (int j = 0; j < 4; j++) v = v << 8 | p[j];
      float fl;
      std::memcpy(&fl, &v, 4);
      x[k] = fl;
    }
  }
  return true;
}

bool read_file(const char *path, std::vector<unsigned char> &buf) {
  FILE *fp = std::fopen(path, "rb");
  if (!fp) return false;
  std::fseek(fp, 0, SEEK_END);
  const long len = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  buf.resize(len > 0 ? (size_t)len : 0);
  const bool ok = len >= 0 && std::fread(buf.data(), 1, buf.size(), fp) == buf.size();
  std::fclose(fp);
  return ok;
}

Att text(const char *name, const std::string &s) {
  Att a;
  a.name = name;
  a.type = NC_CHAR;
  a.text = s;
  return a;
}
Att num(const char *name, double v) {
  Att a;
  a.name = name;
  a.type = NC_DOUBLE;
  a.vals = {v};
  return a;
}
Var var(const char *name, std::vector<int> dims, int type, const char *long_name, const char *units) {
  Var v;
  v.name = name;
  v.dims = std::move(dims);
  v.type = type;
  v.atts = {text("long_name", long_name), text("units", units)};
  return v;
}

// the dimensions and coordinate variables of both layouts: time (record),
// lon, lat, ilon, ilat (fiona_add_dim(..., add_var=.true.))
enum { D_TIME, D_LON, D_LAT, D_ILON, D_ILAT };
void grid(File &f, int nlon, int nlat, const char *time_units) {
  f.dims = {{"time", 0}, {"lon", (uint64_t)nlon}, {"lat", (uint64_t)nlat}, {"ilon", (uint64_t)nlon},
            {"ilat", (uint64_t)(nlat - 1)}};
  f.vars.push_back(var("time", {D_TIME}, NC_DOUBLE, "time", time_units ? time_units : ""));
  f.vars.push_back(var("lon", {D_LON}, NC_DOUBLE, "longitude", "degrees_east"));
  f.vars.push_back(var("lat", {D_LAT}, NC_DOUBLE, "latitude", "degrees_north"));
  f.vars.push_back(var("ilon", {D_ILON}, NC_DOUBLE, "half longitude", "degrees_east"));
  f.vars.push_back(var("ilat", {D_ILAT}, NC_DOUBLE, "half latitude", "degrees_north"));
}

const Var *find(const File &f, const char *name) {
  for (const Var &v : f.vars)
    if (v.name == name) return &v;
  return nullptr;
}

// values of one variable, field order (record 0 for a record variable);
// fixed-size variables followed in file order, then the record
struct Values {
  const char *name;
  const double *x;
  size_t n;
};

int write_new(const char *path, File &f, const std::vector<Values> &vals) {
  layout(f);
  f.numrecs = 1;
  Out o;
  o.b = header(f);
  auto put = [&](const Var &v) -> bool {
    for (const Values &w : vals)
      if (v.name == w.name) {
        if (o.b.size() != v.begin) return false;
        encode(o, v.type, w.x, w.n);
        return true;
      }
    return false;
  };
  for (const Var &v : f.vars)
    if (!v.record && !put(v)) return GMC_EINVAL;
  for (const Var &v : f.vars)
    if (v.record && !put(v)) return GMC_EINVAL;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return GMC_EINVAL;
  const bool ok = std::fwrite(o.b.data(), 1, o.b.size(), fp) == o.b.size();
  return std::fclose(fp) == 0 && ok ? GMC_OK : GMC_EINVAL;
}

}  // namespace nc

namespace {
bool dims_ok(int nlon, int nlat) { return nlon >= 2 && nlat >= 3; }

// h0 definitions of history_setup_h0_swm (history_mod.F90:125-150)
nc::File h0_layout(const gmc_h0_record *r) {
  using namespace nc;
  File f;
  grid(f, r->nlon, r->nlat, r->time_units);
  f.gatts = {num("time_step_size", r->dt)};
  f.vars.push_back(var("tm", {D_TIME}, NC_FLOAT, "total mass", "m"));
  f.vars.push_back(var("te", {D_TIME}, NC_DOUBLE, "total energy", "m4 s-4"));
  f.vars.push_back(var("tpe", {D_TIME}, NC_DOUBLE, "total potential enstrophy", "m2 s-5"));
  f.vars.push_back(var("u", {D_TIME, D_LAT, D_ILON}, NC_DOUBLE, "u wind component", "m s-1"));
  f.vars.push_back(var("v", {D_TIME, D_ILAT, D_LON}, NC_DOUBLE, "v wind component", "m s-1"));
  f.vars.push_back(var("z", {D_TIME, D_LAT, D_LON}, NC_DOUBLE, "height", "m"));
  f.vars.push_back(var("pv", {D_TIME, D_ILAT, D_ILON}, NC_DOUBLE, "potential vorticity", "s-1 m-1"));
  f.vars.push_back(var("div", {D_TIME, D_LAT, D_LON}, NC_DOUBLE, "divergence", "s-1"));
  f.vars.push_back(var("zs", {D_LAT, D_LON}, NC_FLOAT, "surface height", "m"));
  return f;
}
}  // namespace

}  // namespace gmc

using namespace gmc;

extern "C" {

int gmc_h0_write(const char *path, const gmc_h0_record *r, int32_t new_file) {
  if (!path || !r || !dims_ok(r->nlon, r->nlat) || !r->u || !r->v || !r->z || !r->pv || !r->div) return GMC_EINVAL;
  const size_t nl = r->nlon, full = nl * r->nlat, half = nl * (r->nlat - 1);
  const double tm = r->tm, te = r->te, tpe = r->tpe, t = r->time_value;
  std::vector<nc::Values> rec = {{"time", &t, 1}, {"tm", &tm, 1}, {"te", &te, 1}, {"tpe", &tpe, 1},
                                 {"u", r->u, full}, {"v", r->v, half}, {"z", r->z, full},
                                 {"pv", r->pv, half}, {"div", r->div, full}};
  if (new_file) {
    if (!r->lon_deg || !r->lat_deg || !r->ilon_deg || !r->ilat_deg || !r->zs) return GMC_EINVAL;
    nc::File f = h0_layout(r);
    std::vector<nc::Values> vals = rec;
    vals.insert(vals.end(), {{"lon", r->lon_deg, nl}, {"lat", r->lat_deg, (size_t)r->nlat}, {"ilon", r->ilon_deg, nl},
                             {"ilat", r->ilat_deg, (size_t)r->nlat - 1}, {"zs", r->zs, full}});
    return nc::write_new(path, f, vals);
  }
  // append one record to a file of this layout
  std::vector<unsigned char> buf;
  nc::File f;
  if (!nc::read_file(path, buf) || !nc::parse(buf, f)) return GMC_EINVAL;
  const nc::Var *u = nc::find(f, "u");
  if (!u || f.dims[u->dims[1]].len != (uint64_t)r->nlat || f.dims[u->dims[2]].len != nl) return GMC_EINVAL;
  FILE *fp = std::fopen(path, "r+b");
  if (!fp) return GMC_EINVAL;
  bool ok = true;
  for (const nc::Var &v : f.vars) {
    if (!v.record) continue;
    const nc::Values *w = nullptr;
    for (const nc::Values &x : rec)
      if (v.name == x.name) w = &x;
    if (!w) { ok = false; break; }
    nc::Out o;
    nc::encode(o, v.type, w->x, w->n);
    ok = ok && o.b.size() == v.vsize && std::fseek(fp, (long)(v.begin + f.numrecs * f.recsize), SEEK_SET) == 0 &&
         std::fwrite(o.b.data(), 1, o.b.size(), fp) == o.b.size();
  }
  // the record count, after the records
  nc::Out cnt;
  cnt.u32(f.numrecs + 1);
  ok = ok && std::fseek(fp, 4, SEEK_SET) == 0 && std::fwrite(cnt.b.data(), 1, 4, fp) == 4;
  return std::fclose(fp) == 0 && ok ? GMC_OK : GMC_EINVAL;
}

int gmc_r0_write(const char *path, int32_t nlon, int32_t nlat, const double *lon_deg, const double *lat_deg,
                 const double *ilon_deg, const double *ilat_deg, double time_value, const char *time_units,
                 double dt_dyn, const char *restart_interval, const double *u, const double *v,
                 const double *gz, const double *gzs) {
  using namespace nc;
  if (!path || !dims_ok(nlon, nlat) || !lon_deg || !lat_deg || !ilon_deg || !ilat_deg || !u || !v || !gz || !gzs)
    return GMC_EINVAL;
  File f;
  grid(f, nlon, nlat, time_units);
  f.gatts = {num("time_step_size", dt_dyn), text("restart_interval", restart_interval ? restart_interval : "")};
  f.vars.push_back(var("u", {D_TIME, D_LAT, D_ILON}, NC_DOUBLE, "u wind component", "m s-1"));
  f.vars.push_back(var("v", {D_TIME, D_ILAT, D_LON}, NC_DOUBLE, "v wind component", "m s-1"));
  f.vars.push_back(var("gz", {D_TIME, D_LAT, D_LON}, NC_DOUBLE, "geopotential height", "m2 s-2"));
  f.vars.push_back(var("gzs", {D_TIME, D_LAT, D_LON}, NC_DOUBLE, "surface geopotential height", "m2 s-2"));
  const size_t nl = nlon, full = nl * nlat, half = nl * (nlat - 1);
  const double t = time_value;
  return write_new(path, f, {{"time", &t, 1}, {"lon", lon_deg, nl}, {"lat", lat_deg, (size_t)nlat},
                             {"ilon", ilon_deg, nl}, {"ilat", ilat_deg, (size_t)nlat - 1}, {"u", u, full},
                             {"v", v, half}, {"gz", gz, full}, {"gzs", gzs, full}});
}

int gmc_r0_read(const char *path, int32_t nlon, int32_t nlat, double *u, double *v, double *gz, double *gzs,
                double *time_value) {
  if (!path || !dims_ok(nlon, nlat) || !u || !v || !gz || !gzs) return GMC_EINVAL;
  std::vector<unsigned char> buf;
  nc::File f;
  if (!nc::read_file(path, buf) || !nc::parse(buf, f) || f.numrecs < 1) return GMC_EINVAL;
  const size_t nl = nlon, full = nl * nlat, half = nl * (nlat - 1);
  struct Want { const char *name; double *x; size_t n; uint64_t rows; } want[] = {
      {"u", u, full, (uint64_t)nlat}, {"v", v, half, (uint64_t)nlat - 1}, {"gz", gz, full, (uint64_t)nlat},
      {"gzs", gzs, full, (uint64_t)nlat}};
  for (const Want &w : want) {
    const nc::Var *var = nc::find(f, w.name);
    if (!var || var->dims.size() < 2) return GMC_EINVAL;
    const size_t nd = var->dims.size();
    if (f.dims[var->dims[nd - 1]].len != nl || f.dims[var->dims[nd - 2]].len != w.rows) return GMC_EINVAL;
    if (!nc::decode(buf, var->begin, var->type, w.x, w.n)) return GMC_EINVAL;   // record 0
  }
  if (time_value) {
    const nc::Var *t = nc::find(f, "time");
    if (!t || !nc::decode(buf, t->begin, t->type, time_value, 1)) return GMC_EINVAL;
  }
  return GMC_OK;
}

// ---- from a device context ---------------------------------------------------
namespace {
struct Coords {
  std::vector<double> lon, lat, ilon, ilat;
};
// full_lon_deg / half_lon_deg (mesh_mod.F90:180-185, start_lon = 0) and the
// latitude degrees of the metric table
int coords(gmc_ctx *c, Coords &x) {
  const Grid &g = c->grid;
  const double deg = 180.0 / M_PI;
  x.lon.resize(g.nlon);
  x.ilon.resize(g.nlon);
  for (int i = 0; i < g.nlon; i++) {
    const double fl = i * c->dlon;
    x.lon[i] = fl * deg;
    x.ilon[i] = (fl + 0.5 * c->dlon) * deg;
  }
  x.lat.resize(g.nlat);
  x.ilat.resize(g.nhalf);
  int rc = check_hip(c, hipMemcpy(x.lat.data(), c->d_mesh + GMC_FULL_LAT_DEG * g.nlat, sizeof(double) * g.nlat,
                                  hipMemcpyDeviceToHost), "coordinates");
  if (!rc) rc = check_hip(c, hipMemcpy(x.ilat.data(), c->d_mesh + GMC_HALF_LAT_DEG * g.nlat, sizeof(double) * g.nhalf,
                                       hipMemcpyDeviceToHost), "coordinates");
  return rc;
}
int fetch(gmc_ctx *c, int level, int field, std::vector<double> &a) {
  const int rows = (field == GMC_V_LAT || field == GMC_V_F || field == GMC_DV || field == GMC_PV) ? c->grid.nhalf
                                                                                                     : c->grid.nlat;
  a.resize((size_t)rows * c->grid.nlon);
  return gmc_download(c, level, field, a.data(), c->grid.nlon);
}
}  // namespace

int gmc_history_h0(gmc_ctx *c, const char *path, int32_t level, double time_value, const char *time_units,
                   double dt, int32_t new_file) {
  if (!c || !path) return GMC_EINVAL;
  if (level < 1 || level > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  double d[4];
  int rc = gmc_diagnose(c, level, d);     // tm, te, tpe; pv of the level
  if (!rc) rc = ops_calc_div(c, level);   // div of the level's filtered winds
  std::vector<double> u, v, gz, gzs, pv, div;
  if (!rc) rc = fetch(c, level, GMC_U_LON, u);
  if (!rc) rc = fetch(c, level, GMC_V_LAT, v);
  if (!rc) rc = fetch(c, level, GMC_GZ, gz);
  if (!rc) rc = fetch(c, level, GMC_GZS, gzs);
  if (!rc) rc = fetch(c, level, GMC_PV, pv);
  if (!rc) rc = fetch(c, level, GMC_DIV, div);
  Coords x;
  if (!rc) rc = coords(c, x);
  if (rc) return rc;
  // z = gz / g, zs = gzs / g (history_mod.F90:418-420)
  for (double &e : gz) e = e / c->cfg.g;
  for (double &e : gzs) e = e / c->cfg.g;
  gmc_h0_record r{c->grid.nlon, c->grid.nlat, x.lon.data(), x.lat.data(), x.ilon.data(), x.ilat.data(), time_value,
                  time_units, dt, d[0], d[1], d[2], u.data(), v.data(), gz.data(), pv.data(), div.data(), gzs.data()};
  rc = gmc_h0_write(path, &r, new_file);
  return rc ? set_error(c, rc, std::string("cannot write h0 file ") + path) : GMC_OK;
}

int gmc_restart_write(gmc_ctx *c, const char *path, int32_t level, double time_value, const char *time_units,
                      double dt_dyn, const char *restart_interval) {
  if (!c || !path) return GMC_EINVAL;
  if (level < 1 || level > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  std::vector<double> u, v, gz, gzs;
  int rc = fetch(c, level, GMC_U_LON, u);
  if (!rc) rc = fetch(c, level, GMC_V_LAT, v);
  if (!rc) rc = fetch(c, level, GMC_GZ, gz);
  if (!rc) rc = fetch(c, level, GMC_GZS, gzs);
  Coords x;
  if (!rc) rc = coords(c, x);
  if (rc) return rc;
  rc = gmc_r0_write(path, c->grid.nlon, c->grid.nlat, x.lon.data(), x.lat.data(), x.ilon.data(), x.ilat.data(),
                    time_value, time_units, dt_dyn, restart_interval, u.data(), v.data(), gz.data(), gzs.data());
  return rc ? set_error(c, rc, std::string("cannot write r0 file ") + path) : GMC_OK;
}

int gmc_restart_read(gmc_ctx *c, const char *path, int32_t level, double *time_value) {
  if (!c || !path) return GMC_EINVAL;
  if (level < 1 || level > 3) return set_error(c, GMC_EINVAL, "level must be 1..3");
  const Grid &g = c->grid;
  std::vector<double> u((size_t)g.nlat * g.nlon), v((size_t)g.nhalf * g.nlon), gz(u.size()), gzs(u.size());
  int rc = gmc_r0_read(path, g.nlon, g.nlat, u.data(), v.data(), gz.data(), gzs.data(), time_value);
  if (rc) return set_error(c, rc, std::string("cannot read r0 file ") + path);
  rc = gmc_upload(c, level, GMC_U_LON, u.data(), g.nlon);
  if (!rc) rc = gmc_upload(c, level, GMC_V_LAT, v.data(), g.nlon);
  if (!rc) rc = gmc_upload(c, level, GMC_GZ, gz.data(), g.nlon);
  if (!rc) rc = gmc_upload(c, level, GMC_GZS, gzs.data(), g.nlon);
  if (!rc) c->prepared = 0;   // gmc_prepare must run again
  return rc;
}

}  // extern "C"
