// NetCDFGrid3DWriter / NetCDFGrid3DReader (ocustorage/gridnetcdf.cpp) on a
// self-contained implementation of the NetCDF classic file format (the
// published "NetCDF Classic and 64-bit Offset Format" specification):
// big-endian header of dimensions, global attributes and variables, then
// the fixed-size variables, then one record per step of the unlimited
// dimension holding a slab of every record variable.
#include "ocustorage/gridnetcdf.h"

#include <sys/types.h>

#include <algorithm>
#include <cstring>

namespace ocu {

using netcdf_detail::Attr;
using netcdf_detail::Dim;
using netcdf_detail::Var;

namespace {

const uint32_t NC_DIMENSION_TAG = 0x0A, NC_VARIABLE_TAG = 0x0B, NC_ATTRIBUTE_TAG = 0x0C;
const uint32_t STREAMING = 0xFFFFFFFFu;

int type_size(nc_type t) {
  switch (t) {
    case NC_BYTE:
    case NC_CHAR: return 1;
    case NC_SHORT: return 2;
    case NC_INT:
    case NC_FLOAT: return 4;
    case NC_DOUBLE: return 8;
    default: return 0;
  }
}

uint64_t pad4(uint64_t n) { return (n + 3) & ~(uint64_t)3; }

// big-endian encoding of values of 1, 2, 4 or 8 bytes
void put_be(std::vector<unsigned char> &out, const void *v, int size) {
  const unsigned char *p = (const unsigned char *)v;
  for (int b = size - 1; b >= 0; b--) out.push_back(p[b]);
}
void put_u32(std::vector<unsigned char> &out, uint32_t v) { put_be(out, &v, 4); }
void put_u64(std::vector<unsigned char> &out, uint64_t v) { put_be(out, &v, 8); }
void put_name(std::vector<unsigned char> &out, const std::string &s) {
  put_u32(out, (uint32_t)s.size());
  out.insert(out.end(), s.begin(), s.end());
  while (out.size() % 4) out.push_back(0);
}

// the library's default fill values (netcdf.h NC_FILL_*)
void put_fill(std::vector<unsigned char> &out, nc_type t) {
  if (t == NC_FLOAT) {
    float f = 9.9692099683868690e+36f;
    put_be(out, &f, 4);
  } else if (t == NC_DOUBLE) {
    double d = 9.9692099683868690e+36;
    put_be(out, &d, 8);
  } else if (t == NC_INT) {
    int32_t i = -2147483647;
    put_be(out, &i, 4);
  } else {
    for (int b = 0; b < type_size(t); b++) out.push_back(0);
  }
}

// the interior box (nx, ny, nz) of a grid with the given strides, as
// big-endian values in x, y, z order (z fastest)
void encode_box(std::vector<unsigned char> &out, const void *data, int atom, int nx, int ny, int nz, int64_t xs,
                int64_t ys) {
  out.reserve(out.size() + (size_t)nx * ny * nz * atom);
  const unsigned char *base = (const unsigned char *)data;
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++) {
      const unsigned char *row = base + ((int64_t)i * xs + (int64_t)j * ys) * atom;
      for (int k = 0; k < nz; k++) put_be(out, row + (size_t)k * atom, atom);
    }
}

struct Cursor {
  const unsigned char *p, *end;
  bool ok = true;
  bool need(size_t n) {
    if ((size_t)(end - p) < n) ok = false;
    return ok;
  }
  uint64_t be(int size) {
    if (!need(size)) return 0;
    uint64_t v = 0;
    for (int b = 0; b < size; b++) v = (v << 8) | p[b];
    p += size;
    return v;
  }
  uint32_t u32() { return (uint32_t)be(4); }
  std::string name() {
    uint32_t n = u32();
    if (!need(pad4(n))) return std::string();
    std::string s((const char *)p, n);
    p += pad4(n);
    return s;
  }
  void skip(uint64_t n) {
    if (need(n)) p += n;
  }
};

// copy 4-byte-aligned big-endian values to host order
void decode_values(void *dst, const unsigned char *src, size_t count, int atom) {
  unsigned char *d = (unsigned char *)dst;
  for (size_t q = 0; q < count; q++)
    for (int b = 0; b < atom; b++) d[q * atom + b] = src[q * atom + (atom - 1 - b)];
}

}  // namespace

// ------------------------------------------------------------------ writer

NetCDFGrid3DWriter::NetCDFGrid3DWriter()
    : _side(nullptr), _side_size(0), _time_id(-1), _time_var_id(-1), _state(ST_CLOSED), _nx(0), _ny(0), _nz(0),
      _hx(0), _hy(0), _hz(0) {}

NetCDFGrid3DWriter::~NetCDFGrid3DWriter() {
  if (is_open()) close();
}

// gridnetcdf.cpp:22-136: a coordinate dimension and variable per distinct
// offset (XAxisCentered / XAxisUFace / XAxisOffset%.4f ...), one longer for
// the face offset 0, with values i*h + offset
int NetCDFGrid3DWriter::coord(int axis, float offset) {
  std::map<float, int> &m = axis == 0 ? _xcoordoffset_map : axis == 1 ? _ycoordoffset_map : _zcoordoffset_map;
  auto it = m.find(offset);
  if (it != m.end()) return it->second;
  const char *ax = axis == 0 ? "X" : axis == 1 ? "Y" : "Z";
  const char *face = axis == 0 ? "UFace" : axis == 1 ? "VFace" : "WFace";
  char name[1024];
  if (offset == .5f) snprintf(name, sizeof(name), "%sAxisCentered", ax);
  else if (offset == 0) snprintf(name, sizeof(name), "%sAxis%s", ax, face);
  else snprintf(name, sizeof(name), "%sAxisOffset%.4f", ax, offset);
  const int n = axis == 0 ? _nx : axis == 1 ? _ny : _nz;
  const float h = axis == 0 ? _hx : axis == 1 ? _hy : _hz;
  const int len = offset == 0 ? n + 1 : n;
  const int dimid = (int)_dims.size();
  _dims.push_back(Dim{name, (uint64_t)len});
  Var v;
  v.name = name;
  v.type = NC_FLOAT;
  v.dims = {dimid};
  for (int i = 0; i < len; i++) {
    float x = i * h + offset;
    put_be(v.fixed, &x, 4);
  }
  _vars.push_back(v);
  m[offset] = dimid;
  return dimid;
}

void NetCDFGrid3DWriter::add_zero_time_level(int time_level) {
  if (time_level == 0 && _time_steps.size() == 0) {
    size_t dummy;
    add_time_level(0, dummy);
  }
}

// gridnetcdf.cpp:152-177
bool NetCDFGrid3DWriter::add_time_level(float time, size_t &level) {
  if (!is_open()) {
    printf("[ERROR] NetCDFGrid3DWriter::add_time_level - file not open\n");
    return false;
  }
  if (_state == ST_OPEN_DEFINE_MODE) _state = ST_OPEN_DATA_MODE;
  _time_steps.push_back(time);
  level = _time_steps.size() - 1;
  return true;
}

// gridnetcdf.cpp:179-225
bool NetCDFGrid3DWriter::open(const char *filename, int nx, int ny, int nz, float hx, float hy, float hz) {
  if (is_open()) {
    printf("[ERROR] NetCDFGrid3DWriter::open - already open\n");
    return false;
  }
  if (!filename || nx <= 0 || ny <= 0 || nz <= 0) {
    printf("[ERROR] NetCDFGrid3DWriter::open - invalid arguments\n");
    return false;
  }
  // created (and truncated) now, as nc_create(NC_CLOBBER) does
  FILE *f = fopen(filename, "wb");
  if (!f) {
    printf("[ERROR] NetCDFGrid3DWriter::open - could not open file %s\n", filename);
    return false;
  }
  fclose(f);
  _sidename = std::string(filename) + ".ocu-records";
  _side = fopen(_sidename.c_str(), "w+b");
  if (!_side) {
    printf("[ERROR] NetCDFGrid3DWriter::open - could not open scratch file %s\n", _sidename.c_str());
    return false;
  }
  _side_size = 0;
  _filename = filename;
  _state = ST_OPEN_DEFINE_MODE;
  _nx = nx, _ny = ny, _nz = nz;
  _hx = hx, _hy = hy, _hz = hz;
  _dims.clear();
  _gatts.clear();
  _vars.clear();
  _slabs.clear();
  _varid_map.clear();
  _time_steps.clear();
  _xcoordoffset_map.clear();
  _ycoordoffset_map.clear();
  _zcoordoffset_map.clear();
  _time_id = 0;
  _dims.push_back(Dim{"Time", 0});
  auto iatt = [&](const char *name, int v) {
    Attr a{name, NC_INT, {}, 1};
    put_be(a.bytes, &v, 4);
    _gatts.push_back(a);
  };
  auto fatt = [&](const char *name, float v) {
    Attr a{name, NC_FLOAT, {}, 1};
    put_be(a.bytes, &v, 4);
    _gatts.push_back(a);
  };
  iatt("nx", _nx);
  iatt("ny", _ny);
  iatt("nz", _nz);
  fatt("hx", _hx);
  fatt("hy", _hy);
  fatt("hz", _hz);
  Var t;
  t.name = "Time";
  t.type = NC_FLOAT;
  t.dims = {_time_id};
  _time_var_id = (int)_vars.size();
  _vars.push_back(t);
  return true;
}

// gridnetcdf.cpp:227-276
bool NetCDFGrid3DWriter::define_variable(const char *name, nc_type var_type, GridStaggering stag, float offsetx,
                                         float offsety, float offsetz) {
  if (!is_open()) {
    printf("[ERROR] NetCDFGrid3DWriter::define_variable - file not open\n");
    return false;
  }
  if (var_type != NC_FLOAT && var_type != NC_DOUBLE && var_type != NC_INT) {
    printf("[ERROR] NetCDFGrid3DWriter::define_variable - invalid var_type %d\n", var_type);
    return false;
  }
  if (!name || _varid_map.count(name) != 0) {
    printf("[ERROR] NetCDFGrid3DWriter::define_variable - duplicate name %s\n", name ? name : "(null)");
    return false;
  }
  for (const Var &v : _vars)
    if (v.name == name) {
      printf("[ERROR] NetCDFGrid3DWriter::define_variable - name %s is a coordinate\n", name);
      return false;
    }
  if (_state == ST_OPEN_DATA_MODE) _state = ST_OPEN_DEFINE_MODE;
  // if staggering is GS_CUSTOM, the given offsets stand
  grid_staggering_to_offset(stag, offsetx, offsety, offsetz);
  Variable var;
  var.type = var_type;
  var.coords[0] = _time_id;
  var.coords[1] = xcoord(offsetx);
  var.coords[2] = ycoord(offsety);
  var.coords[3] = zcoord(offsetz);
  grid_staggering_to_dimension(stag, _nx, _ny, _nz, var.nx, var.ny, var.nz);
  Var v;
  v.name = name;
  v.type = var_type;
  v.dims = {var.coords[0], var.coords[1], var.coords[2], var.coords[3]};
  var.id = (int)_vars.size();
  _vars.push_back(v);
  _varid_map[name] = var;
  return true;
}

bool NetCDFGrid3DWriter::add_data_internal(const char *name, const Grid3DUntyped &grid, nc_type data_type,
                                           const void *data, int time_level) {
  if (!is_open()) {
    printf("[ERROR] NetCDFGrid3DWriter::add_data - file not open\n");
    return false;
  }
  if (_state == ST_OPEN_DEFINE_MODE) _state = ST_OPEN_DATA_MODE;
  auto it = name ? _varid_map.find(name) : _varid_map.end();
  if (it == _varid_map.end()) {
    printf("[ERROR] NetCDFGrid3DWriter::add_data - variable %s not found\n", name ? name : "(null)");
    return false;
  }
  const Variable &var = it->second;
  if (grid.nx() != var.nx || grid.ny() != var.ny || grid.nz() != var.nz) {
    printf("[ERROR] NetCDFGrid3DWriter::add_data - dimension mismatch (%d, %d, %d) != (%d, %d, %d)\n", grid.nx(),
           grid.ny(), grid.nz(), var.nx, var.ny, var.nz);
    return false;
  }
  if (var.type != data_type) {
    printf("[ERROR] NetCDFGrid3DWriter::add_data - type mismatch\n");
    return false;
  }
  if (time_level < 0) {
    printf("[ERROR] NetCDFGrid3DWriter::add_data - invalid time level %d\n", time_level);
    return false;
  }
  add_zero_time_level(time_level);
  // the slab with the file's extents (a custom staggering may give a
  // coordinate more points than the variable has: fill values there)
  const Var &v = _vars[var.id];
  const int lx = (int)_dims[v.dims[1]].len, ly = (int)_dims[v.dims[2]].len, lz = (int)_dims[v.dims[3]].len;
  std::vector<unsigned char> bytes;
  const int atom = type_size(data_type);
  if (lx == var.nx && ly == var.ny && lz == var.nz) {
    encode_box(bytes, data, atom, var.nx, var.ny, var.nz, grid.xstride(), grid.ystride());
  } else {
    std::vector<unsigned char> fill;
    put_fill(fill, data_type);
    for (int i = 0; i < lx; i++)
      for (int j = 0; j < ly; j++)
        for (int k = 0; k < lz; k++) {
          if (i < var.nx && j < var.ny && k < var.nz)
            put_be(bytes, (const unsigned char *)data + ((int64_t)i * grid.xstride() + (int64_t)j * grid.ystride() + k) * atom,
                   atom);
          else
            bytes.insert(bytes.end(), fill.begin(), fill.end());
        }
  }
  if (fseeko(_side, (off_t)_side_size, SEEK_SET) != 0 || fwrite(bytes.data(), 1, bytes.size(), _side) != bytes.size()) {
    printf("[ERROR] NetCDFGrid3DWriter::add_data - write failed\n");
    return false;
  }
  _slabs.push_back(Slab{var.id, time_level, _side_size});
  _side_size += bytes.size();
  return true;
}

bool NetCDFGrid3DWriter::add_data(const char *name, const Grid3DHostD &grid, int time_level) {
  return add_data_internal(name, grid, NC_DOUBLE, &grid.at(0, 0, 0), time_level);
}
bool NetCDFGrid3DWriter::add_data(const char *name, const Grid3DHostF &grid, int time_level) {
  return add_data_internal(name, grid, NC_FLOAT, &grid.at(0, 0, 0), time_level);
}
bool NetCDFGrid3DWriter::add_data(const char *name, const Grid3DHostI &grid, int time_level) {
  return add_data_internal(name, grid, NC_INT, &grid.at(0, 0, 0), time_level);
}

bool NetCDFGrid3DWriter::add_data_layout(const char *name, const Grid3DUntyped &layout, nc_type type,
                                         const void *interior, int time_level) {
  return add_data_internal(name, layout, type, interior, time_level);
}

// lay the file out: header, fixed-size variables, records
bool NetCDFGrid3DWriter::write_file() {
  int numrecs = (int)_time_steps.size();
  for (const Slab &s : _slabs) numrecs = std::max(numrecs, s.level + 1);
  // slab sizes and offsets
  uint64_t fixed_end = 0;
  for (Var &v : _vars) {
    v.record = !v.dims.empty() && _dims[v.dims[0]].len == 0;
    uint64_t n = type_size(v.type);
    for (size_t d = v.record ? 1 : 0; d < v.dims.size(); d++) n *= _dims[v.dims[d]].len;
    v.vsize = pad4(n);
  }
  int nrec_vars = 0;
  for (const Var &v : _vars) nrec_vars += v.record;
  // the header, with begin offsets of 8 bytes (CDF-2)
  auto header = [&](std::vector<unsigned char> &h) {
    h.clear();
    h.push_back('C');
    h.push_back('D');
    h.push_back('F');
    h.push_back(2);
    put_u32(h, (uint32_t)numrecs);
    put_u32(h, NC_DIMENSION_TAG);
    put_u32(h, (uint32_t)_dims.size());
    for (const Dim &d : _dims) {
      put_name(h, d.name);
      put_u32(h, (uint32_t)d.len);
    }
    put_u32(h, NC_ATTRIBUTE_TAG);
    put_u32(h, (uint32_t)_gatts.size());
    for (const Attr &a : _gatts) {
      put_name(h, a.name);
      put_u32(h, (uint32_t)a.type);
      put_u32(h, a.count);
      h.insert(h.end(), a.bytes.begin(), a.bytes.end());
      while (h.size() % 4) h.push_back(0);
    }
    put_u32(h, NC_VARIABLE_TAG);
    put_u32(h, (uint32_t)_vars.size());
    for (const Var &v : _vars) {
      put_name(h, v.name);
      put_u32(h, (uint32_t)v.dims.size());
      for (int d : v.dims) put_u32(h, (uint32_t)d);
      put_u32(h, 0);  // no variable attributes
      put_u32(h, 0);
      put_u32(h, (uint32_t)v.type);
      put_u32(h, v.vsize > 0xFFFFFFFCull ? 0xFFFFFFFFu : (uint32_t)v.vsize);
      put_u64(h, v.begin);
    }
  };
  std::vector<unsigned char> head;
  header(head);  // sizes only depend on counts, not on the offsets
  uint64_t off = head.size();
  for (Var &v : _vars)
    if (!v.record) {
      v.begin = off;
      off += v.vsize;
    }
  fixed_end = off;
  uint64_t recsize = 0;
  for (Var &v : _vars)
    if (v.record) {
      v.begin = fixed_end + recsize;
      recsize += v.vsize;
    }
  if (nrec_vars == 1)  // a single record variable is not padded
    for (Var &v : _vars)
      if (v.record) {
        uint64_t n = type_size(v.type);
        for (size_t d = 1; d < v.dims.size(); d++) n *= _dims[v.dims[d]].len;
        recsize = n;
      }
  header(head);
  FILE *f = fopen(_filename.c_str(), "wb");
  if (!f) {
    printf("[ERROR] NetCDFGrid3DWriter::close - could not write %s\n", _filename.c_str());
    return false;
  }
  bool ok = fwrite(head.data(), 1, head.size(), f) == head.size();
  for (const Var &v : _vars)
    if (!v.record && ok) {
      std::vector<unsigned char> d = v.fixed;
      d.resize(v.vsize, 0);
      ok = fwrite(d.data(), 1, d.size(), f) == d.size();
    }
  // the last slab written for each (variable, level)
  std::map<std::pair<int, int>, uint64_t> where;
  for (const Slab &s : _slabs) where[{s.var, s.level}] = s.offset;
  std::vector<unsigned char> buf;
  for (int r = 0; r < numrecs && ok; r++)
    for (size_t q = 0; q < _vars.size() && ok; q++) {
      const Var &v = _vars[q];
      if (!v.record) continue;
      const uint64_t n = (nrec_vars == 1) ? recsize : v.vsize;
      buf.clear();
      if ((int)q == _time_var_id) {
        if (r < (int)_time_steps.size()) put_be(buf, &_time_steps[r], 4);
        else put_fill(buf, NC_FLOAT);
      } else {
        auto it = where.find({(int)q, r});
        if (it != where.end()) {
          buf.resize(n);
          ok = fseeko(_side, (off_t)it->second, SEEK_SET) == 0 && fread(buf.data(), 1, n, _side) == n;
        } else {
          const uint64_t count = n / type_size(v.type);
          std::vector<unsigned char> fill;
          put_fill(fill, v.type);
          for (uint64_t c = 0; c < count; c++) buf.insert(buf.end(), fill.begin(), fill.end());
        }
      }
      buf.resize(n, 0);
      ok = ok && fwrite(buf.data(), 1, buf.size(), f) == buf.size();
    }
  ok = (fclose(f) == 0) && ok;
  if (!ok) printf("[ERROR] NetCDFGrid3DWriter::close - write to %s failed\n", _filename.c_str());
  return ok;
}

// gridnetcdf.cpp:324-340
bool NetCDFGrid3DWriter::close() {
  if (!is_open()) {
    printf("[ERROR] NetCDFGrid3DWriter::close - file not open\n");
    return false;
  }
  bool ok = write_file();
  fclose(_side);
  remove(_sidename.c_str());
  _side = nullptr;
  _state = ST_CLOSED;
  _varid_map.clear();
  _time_steps.clear();
  _slabs.clear();
  return ok;
}

// ------------------------------------------------------------------ reader

NetCDFGrid3DReader::NetCDFGrid3DReader()
    : _file(nullptr), _time_id(-1), _recsize(0), _nx(0), _ny(0), _nz(0), _hx(0), _hy(0), _hz(0), _is_open(false) {}

NetCDFGrid3DReader::~NetCDFGrid3DReader() {
  if (is_open()) close();
}

// gridnetcdf.cpp:432-530
bool NetCDFGrid3DReader::open(const char *filename) {
  if (is_open()) {
    printf("[ERROR] NetCDFGrid3DReader::open - already open\n");
    return false;
  }
  FILE *f = filename ? fopen(filename, "rb") : nullptr;
  if (!f) {
    printf("[ERROR] NetCDFGrid3DReader::open - could not open file %s\n", filename ? filename : "(null)");
    return false;
  }
  fseeko(f, 0, SEEK_END);
  const uint64_t file_size = (uint64_t)ftello(f);
  fseeko(f, 0, SEEK_SET);
  // the header is read in growing chunks until it parses
  std::vector<unsigned char> head;
  bool parsed = false;
  int version = 0;
  uint32_t numrecs = 0;
  std::vector<Dim> dims;
  std::vector<Var> vars;
  std::vector<Attr> gatts;
  for (uint64_t want = 4096;; want *= 4) {
    const uint64_t n = std::min(want, file_size);
    head.resize(n);
    fseeko(f, 0, SEEK_SET);
    if (fread(head.data(), 1, n, f) != n) break;
    Cursor c{head.data(), head.data() + head.size()};
    dims.clear(), vars.clear(), gatts.clear();
    if (!c.need(4) || memcmp(c.p, "CDF", 3) != 0 || (c.p[3] != 1 && c.p[3] != 2)) break;
    version = c.p[3];
    c.skip(4);
    numrecs = c.u32();
    auto read_atts = [&](std::vector<Attr> *out) {
      const uint32_t tag = c.u32(), count = c.u32();
      if (tag != NC_ATTRIBUTE_TAG && !(tag == 0 && count == 0)) c.ok = false;
      for (uint32_t a = 0; a < count && c.ok; a++) {
        Attr at;
        at.name = c.name();
        at.type = (nc_type)c.u32();
        at.count = c.u32();
        const uint64_t bytes = (uint64_t)at.count * type_size(at.type);
        if (!type_size(at.type) || !c.need(pad4(bytes))) {
          c.ok = false;
          break;
        }
        at.bytes.assign(c.p, c.p + bytes);
        c.skip(pad4(bytes));
        if (out) out->push_back(at);
      }
    };
    uint32_t tag = c.u32(), count = c.u32();
    if (tag != NC_DIMENSION_TAG && !(tag == 0 && count == 0)) break;
    for (uint32_t d = 0; d < count && c.ok; d++) {
      Dim dm;
      dm.name = c.name();
      dm.len = c.u32();
      dims.push_back(dm);
    }
    read_atts(&gatts);
    tag = c.u32(), count = c.u32();
    if (c.ok && tag != NC_VARIABLE_TAG && !(tag == 0 && count == 0)) break;
    for (uint32_t q = 0; q < count && c.ok; q++) {
      Var v;
      v.name = c.name();
      const uint32_t nd = c.u32();
      for (uint32_t d = 0; d < nd && c.ok; d++) v.dims.push_back((int)c.u32());
      read_atts(nullptr);
      v.type = (nc_type)c.u32();
      v.vsize = c.u32();
      v.begin = version == 2 ? c.be(8) : c.u32();
      for (int d : v.dims)
        if (d < 0 || d >= (int)dims.size()) c.ok = false;
      v.record = c.ok && !v.dims.empty() && dims[v.dims[0]].len == 0;
      vars.push_back(v);
    }
    if (c.ok) {
      parsed = true;
      break;
    }
    if (n == file_size) break;
  }
  if (!parsed) {
    fclose(f);
    printf("[ERROR] NetCDFGrid3DReader::open - %s is not a NetCDF classic file\n", filename);
    return false;
  }
  // the record size: padded slabs of every record variable (one record
  // variable: unpadded)
  uint64_t recsize = 0;
  int nrec = 0;
  for (Var &v : vars)
    if (v.record) {
      uint64_t n = type_size(v.type);
      for (size_t d = 1; d < v.dims.size(); d++) n *= dims[v.dims[d]].len;
      v.vsize = pad4(n);
      recsize += v.vsize;
      nrec++;
    }
  if (nrec == 1)
    for (Var &v : vars)
      if (v.record) {
        uint64_t n = type_size(v.type);
        for (size_t d = 1; d < v.dims.size(); d++) n *= dims[v.dims[d]].len;
        recsize = n;
      }
  if (numrecs == STREAMING && recsize > 0) {
    uint64_t first = UINT64_MAX;
    for (const Var &v : vars)
      if (v.record) first = std::min(first, v.begin);
    numrecs = (uint32_t)((file_size - first) / recsize);
  }
  auto geti = [&](const char *name, int &out) {
    for (const Attr &a : gatts)
      if (a.name == name && a.type == NC_INT && a.count >= 1) {
        int32_t v;
        decode_values(&v, a.bytes.data(), 1, 4);
        out = v;
      }
  };
  auto getf = [&](const char *name, float &out) {
    for (const Attr &a : gatts)
      if (a.name == name && a.type == NC_FLOAT && a.count >= 1) decode_values(&out, a.bytes.data(), 1, 4);
  };
  _nx = _ny = _nz = 0;
  _hx = _hy = _hz = 0;
  geti("nx", _nx), geti("ny", _ny), geti("nz", _nz);
  getf("hx", _hx), getf("hy", _hy), getf("hz", _hz);
  _time_id = -1;
  for (size_t d = 0; d < dims.size(); d++)
    if (dims[d].name == "Time") _time_id = (int)d;
  if (_time_id < 0) {
    fclose(f);
    printf("[ERROR] NetCDFGrid3DReader::open - no Time dimension in %s\n", filename);
    return false;
  }
  const uint64_t ntime = dims[_time_id].len == 0 ? numrecs : dims[_time_id].len;
  _time_steps.assign(ntime, 0.0f);
  _dims = dims;
  _vars = vars;
  _recsize = recsize;
  _file = f;
  _filename = filename;
  // time values from a 1-D Time variable over the Time dimension
  for (const Var &v : vars)
    if (v.name == "Time" && v.dims.size() == 1 && v.dims[0] == _time_id && v.type == NC_FLOAT) {
      for (uint64_t t = 0; t < ntime; t++) {
        unsigned char b[4];
        const uint64_t at = v.record ? v.begin + t * recsize : v.begin + 4 * t;
        if (fseeko(f, (off_t)at, SEEK_SET) == 0 && fread(b, 1, 4, f) == 4) decode_values(&_time_steps[t], b, 1, 4);
      }
    }
  // the 4-D variables (time, x, y, z)
  _varid_map.clear();
  for (size_t q = 0; q < vars.size(); q++) {
    const Var &v = vars[q];
    if (v.dims.size() != 4) continue;
    Variable var;
    var.id = (int)q;
    var.type = v.type;
    for (int d = 0; d < 4; d++) var.coords[d] = v.dims[d];
    var.nx = (int)dims[v.dims[1]].len;
    var.ny = (int)dims[v.dims[2]].len;
    var.nz = (int)dims[v.dims[3]].len;
    _varid_map[v.name] = var;
  }
  _is_open = true;
  return true;
}

std::vector<std::string> NetCDFGrid3DReader::list_variables() const {
  std::vector<std::string> result;
  for (auto &kv : _varid_map) result.push_back(kv.first);
  return result;
}

bool NetCDFGrid3DReader::variable_type(const char *name, nc_type &type) const {
  auto it = _varid_map.find(name);
  if (it == _varid_map.end()) {
    printf("[ERROR] NetCDFGrid3DReader::variable_type - variable %s not found\n", name);
    return false;
  }
  type = it->second.type;
  return true;
}

bool NetCDFGrid3DReader::variable_size(const char *name, int &nx, int &ny, int &nz) const {
  auto it = _varid_map.find(name);
  if (it == _varid_map.end()) {
    printf("[ERROR] NetCDFGrid3DReader::variable_size - variable %s not found\n", name);
    return false;
  }
  nx = it->second.nx, ny = it->second.ny, nz = it->second.nz;
  return true;
}

// gridnetcdf.cpp:584-626
bool NetCDFGrid3DReader::read_variable_internal(const char *name, const Grid3DUntyped &grid, nc_type data_type,
                                                void *data, int time_level) const {
  if (!is_open()) {
    printf("[ERROR] NetCDFGrid3DReader::read_variable - file not open\n");
    return false;
  }
  auto it = name ? _varid_map.find(name) : _varid_map.end();
  if (it == _varid_map.end()) {
    printf("[ERROR] NetCDFGrid3DReader::read_variable - variable %s not found\n", name ? name : "(null)");
    return false;
  }
  const Variable &var = it->second;
  if (grid.nx() != var.nx || grid.ny() != var.ny || grid.nz() != var.nz) {
    printf("[ERROR] NetCDFGrid3DReader::read_variable - dimension mismatch (%d, %d, %d) != (%d, %d, %d)\n", grid.nx(),
           grid.ny(), grid.nz(), var.nx, var.ny, var.nz);
    return false;
  }
  if (var.type != data_type) {
    printf("[ERROR] NetCDFGrid3DReader::read_variable - type mismatch\n");
    return false;
  }
  const Var &v = _vars[var.id];
  const uint64_t t0 = _dims[v.dims[0]].len;  // fixed first dimension: its length
  const uint64_t levels = v.record ? (uint64_t)_time_steps.size() : t0;
  if (time_level < 0 || (uint64_t)time_level >= levels) {
    printf("[ERROR] NetCDFGrid3DReader::read_variable - invalid time level %d\n", time_level);
    return false;
  }
  const int atom = type_size(data_type);
  const uint64_t slab = (uint64_t)var.nx * var.ny * var.nz * atom;
  const uint64_t at = v.record ? v.begin + (uint64_t)time_level * _recsize : v.begin + (uint64_t)time_level * slab;
  std::vector<unsigned char> buf(slab);
  if (fseeko(_file, (off_t)at, SEEK_SET) != 0 || fread(buf.data(), 1, slab, _file) != slab) {
    printf("[ERROR] NetCDFGrid3DReader::read_variable - read failed\n");
    return false;
  }
  unsigned char *base = (unsigned char *)data;
  const unsigned char *src = buf.data();
  for (int i = 0; i < var.nx; i++)
    for (int j = 0; j < var.ny; j++) {
      decode_values(base + ((int64_t)i * grid.xstride() + (int64_t)j * grid.ystride()) * atom, src, var.nz, atom);
      src += (size_t)var.nz * atom;
    }
  return true;
}

bool NetCDFGrid3DReader::read_variable(const char *name, Grid3DHostD &grid, int time_level) const {
  return read_variable_internal(name, grid, NC_DOUBLE, &grid.at(0, 0, 0), time_level);
}
bool NetCDFGrid3DReader::read_variable(const char *name, Grid3DHostF &grid, int time_level) const {
  return read_variable_internal(name, grid, NC_FLOAT, &grid.at(0, 0, 0), time_level);
}
bool NetCDFGrid3DReader::read_variable(const char *name, Grid3DHostI &grid, int time_level) const {
  return read_variable_internal(name, grid, NC_INT, &grid.at(0, 0, 0), time_level);
}

bool NetCDFGrid3DReader::read_variable_layout(const char *name, const Grid3DUntyped &layout, nc_type type,
                                              void *interior, int time_level) const {
  return read_variable_internal(name, layout, type, interior, time_level);
}

bool NetCDFGrid3DReader::close() {
  if (!is_open()) {
    printf("[ERROR] NetCDFGrid3DReader::close - file not open\n");
    return false;
  }
  fclose(_file);
  _file = nullptr;
  _is_open = false;
  return true;
}

}  // namespace ocu
