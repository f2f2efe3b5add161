// Interface after OpenCurrent (Copyright 2008-2009 NVIDIA Corporation, Apache License 2.0; see NOTICE).
// Grid files, the drop-in for ocustorage/gridnetcdf.h: NetCDFGrid3DWriter /
// NetCDFGrid3DReader write and read Grid3DHost fields as 4-D variables
// (Time, X, Y, Z) of a NetCDF file, with coordinate variables per
// staggering (cell centres, u/v/w faces), global attributes nx, ny, nz, hx,
// hy, hz and an unlimited Time dimension.
//
// The file is the NetCDF classic format with 64-bit offsets (CDF-2),
// written and read by this library itself (no libnetcdf): the definitions
// and the data are collected while the file is open - the data of each
// add_data in a side file, so memory stays bounded - and close() lays the
// file out: header, coordinate variables, then one record per time level.
// The reader accepts CDF-1 and CDF-2 files.
#ifndef __OCU_STORAGE_GRID_NETCFD_H__
#define __OCU_STORAGE_GRID_NETCFD_H__

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "ocustorage/grid3d.h"

#ifndef NC_INT
typedef int nc_type;
#define NC_BYTE 1
#define NC_CHAR 2
#define NC_SHORT 3
#define NC_INT 4
#define NC_FLOAT 5
#define NC_DOUBLE 6
#endif

namespace ocu {

// gridnetcdf.h:33-39
enum GridStaggering { GS_CENTER_POINT, GS_U_FACE, GS_V_FACE, GS_W_FACE, GS_CUSTOM };

// gridnetcdf.h:42-51
inline void grid_staggering_to_offset(GridStaggering stag, float &xoffset, float &yoffset, float &zoffset) {
  switch (stag) {
    case GS_CENTER_POINT: xoffset = 0.5f; yoffset = 0.5f; zoffset = 0.5f; break;
    case GS_U_FACE: xoffset = 0.0f; yoffset = 0.5f; zoffset = 0.5f; break;
    case GS_V_FACE: xoffset = 0.5f; yoffset = 0.0f; zoffset = 0.5f; break;
    case GS_W_FACE: xoffset = 0.5f; yoffset = 0.5f; zoffset = 0.0f; break;
    default: break;
  }
}

// gridnetcdf.h:53-66
inline void grid_staggering_to_dimension(GridStaggering stag, int nx, int ny, int nz, int &out_nx, int &out_ny,
                                         int &out_nz) {
  out_nx = nx;
  out_ny = ny;
  out_nz = nz;
  switch (stag) {
    case GS_U_FACE: out_nx = nx + 1; break;
    case GS_V_FACE: out_ny = ny + 1; break;
    case GS_W_FACE: out_nz = nz + 1; break;
    default: break;
  }
}

namespace netcdf_detail {
struct Dim {
  std::string name;
  uint64_t len;  // 0: the record (Time) dimension
};
struct Attr {
  std::string name;
  nc_type type;
  std::vector<unsigned char> bytes;  // big-endian values
  uint32_t count;
};
struct Var {
  std::string name;
  nc_type type;
  std::vector<int> dims;
  std::vector<unsigned char> fixed;  // big-endian data of a fixed-size variable
  uint64_t begin = 0, vsize = 0;     // filled by the layout
  bool record = false;
};
}  // namespace netcdf_detail

class NetCDFGrid3DWriter {
  enum State { ST_CLOSED, ST_OPEN_DEFINE_MODE, ST_OPEN_DATA_MODE };

  struct Variable {
    int id;
    nc_type type;
    int coords[4];  // time, x, y, z
    int nx, ny, nz;
  };
  struct Slab {
    int var, level;
    uint64_t offset;  // in the side file
  };

  std::string _filename, _sidename;
  FILE *_side;
  uint64_t _side_size;
  int _time_id, _time_var_id;
  std::map<float, int> _xcoordoffset_map, _ycoordoffset_map, _zcoordoffset_map;
  std::map<std::string, Variable> _varid_map;
  std::vector<float> _time_steps;
  std::vector<netcdf_detail::Dim> _dims;
  std::vector<netcdf_detail::Attr> _gatts;
  std::vector<netcdf_detail::Var> _vars;
  std::vector<Slab> _slabs;
  State _state;
  int _nx, _ny, _nz;
  float _hx, _hy, _hz;

  void add_zero_time_level(int time_level);
  int coord(int axis, float offset);
  int xcoord(float offset) { return coord(0, offset); }
  int ycoord(float offset) { return coord(1, offset); }
  int zcoord(float offset) { return coord(2, offset); }
  bool add_data_internal(const char *name, const Grid3DUntyped &grid, nc_type data_type, const void *data,
                         int time_level);
  bool write_file();

 public:
  NetCDFGrid3DWriter();
  ~NetCDFGrid3DWriter();

  bool is_open() const { return _state != ST_CLOSED; }
  bool add_time_level(float time, size_t &level);
  int get_time_level(float time) const {
    for (size_t i = 0; i < _time_steps.size(); i++)
      if (_time_steps[i] == time) return (int)i;
    return -1;
  }
  float get_time(int time_level) const { return _time_steps[time_level]; }

  bool open(const char *filename, int nx, int ny, int nz, float hx = 1.0f, float hy = 1.0f, float hz = 1.0f);
  bool define_variable(const char *name, nc_type var_type, GridStaggering stag, float offsetx = 0.5f,
                       float offsety = 0.5f, float offsetz = 0.5f);
  bool add_data(const char *name, const Grid3DHostD &grid, int time_level = 0);
  bool add_data(const char *name, const Grid3DHostF &grid, int time_level = 0);
  bool add_data(const char *name, const Grid3DHostI &grid, int time_level = 0);
  // not in the reference: the interior of any grid with this layout
  // (host memory at cell (0,0,0)), for callers without a Grid3DHost
  bool add_data_layout(const char *name, const Grid3DUntyped &layout, nc_type type, const void *interior,
                       int time_level = 0);
  bool close();
};

class NetCDFGrid3DReader {
  struct Variable {
    int id;
    nc_type type;
    int coords[4];  // time, x, y, z
    int nx, ny, nz;
  };

  std::string _filename;
  FILE *_file;
  int _time_id;
  std::map<std::string, Variable> _varid_map;
  std::vector<float> _time_steps;
  std::vector<netcdf_detail::Dim> _dims;
  std::vector<netcdf_detail::Var> _vars;
  uint64_t _recsize;
  int _nx, _ny, _nz;
  float _hx, _hy, _hz;
  bool _is_open;

  bool read_variable_internal(const char *name, const Grid3DUntyped &grid, nc_type data_type, void *data,
                              int time_level) const;

 public:
  NetCDFGrid3DReader();
  ~NetCDFGrid3DReader();

  bool is_open() const { return _is_open; }
  bool open(const char *filename);

  int nx() const { return _nx; }
  int ny() const { return _ny; }
  int nz() const { return _nz; }
  float hx() const { return _hx; }
  float hy() const { return _hy; }
  float hz() const { return _hz; }

  int num_time_levels() const { return (int)_time_steps.size(); }
  float get_time(int time_level) const { return _time_steps[time_level]; }

  std::vector<std::string> list_variables() const;
  bool variable_type(const char *name, nc_type &type) const;
  bool variable_size(const char *name, int &nx, int &ny, int &nz) const;

  bool read_variable(const char *name, Grid3DHostD &grid, int time_level = 0) const;
  bool read_variable(const char *name, Grid3DHostF &grid, int time_level = 0) const;
  bool read_variable(const char *name, Grid3DHostI &grid, int time_level = 0) const;
  // not in the reference: into any grid with this layout (host memory at
  // cell (0,0,0))
  bool read_variable_layout(const char *name, const Grid3DUntyped &layout, nc_type type, void *interior,
                            int time_level = 0) const;

  bool close();
};

}  // namespace ocu

#endif
