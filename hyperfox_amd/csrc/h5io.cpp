// HDF5 mesh and field files in the reference's layout (HDF5Io,
// src/io/HDF5Io.cpp:13-390), on the HDF5 C library:
//   /Mesh/Nodes      nNodes x dim, double
//   /Mesh/Cells      nCells x nodes per cell, int (the reference element's order)
//   /FieldData/NAME  nEntities x objects per entity x values per object, double,
//                    int attribute "ftype" (None -1, Node 0, Edge 1, Face 2, Cell 3)
// Built as its own library (libhfx_io.so) so that the GPU library does not
// depend on HDF5.
#include "../../include/hfx_io.h"

#include <hdf5.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

struct Quiet {  // HDF5 prints its error stack by default; errors are reported here
  Quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }
};

struct H {  // hid_t owner
  hid_t id;
  herr_t (*close)(hid_t);
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() {
    if (id >= 0) close(id);
  }
  H(const H&) = delete;
  H& operator=(const H&) = delete;
};

void need(bool c, const std::string& msg) {
  if (!c) throw std::runtime_error(msg);
}

hid_t open_file(const char* path) {
  need(path != nullptr, "HDF5Io: null file name");
  need(H5Fis_hdf5(path) > 0, std::string("HDF5Io: ") + path + " is not an HDF5 file");
  const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  need(f >= 0, std::string("HDF5Io: could not open ") + path);
  return f;
}

void dims_of(hid_t dset, int rank, hsize_t* d, const std::string& what) {
  H sp(H5Dget_space(dset), H5Sclose);
  need(sp.id >= 0, "HDF5Io: no dataspace for " + what);
  need(H5Sget_simple_extent_ndims(sp.id) == rank, "HDF5Io: " + what + " must have rank " + std::to_string(rank));
  need(H5Sget_simple_extent_dims(sp.id, d, nullptr) == rank, "HDF5Io: could not get the shape of " + what);
}

}  // namespace

#define IO_TRY \
  try {        \
    Quiet q_;
#define IO_END                    \
  }                               \
  catch (const std::exception& e) { \
    g_err = e.what();             \
    return 1;                     \
  }                               \
  return 0;

extern "C" const char* hfx_io_last_error(void) { return g_err.c_str(); }

extern "C" int hfx_h5_mesh_sizes(const char* path, long long* out) {
  IO_TRY
  H f(open_file(path), H5Fclose);
  need(H5Lexists(f.id, "Mesh", H5P_DEFAULT) > 0, "HDF5Io: no Mesh group in " + std::string(path));
  H g(H5Gopen(f.id, "Mesh", H5P_DEFAULT), H5Gclose);
  hsize_t dn[2], dc[2];
  {
    H ds(H5Dopen(g.id, "Nodes", H5P_DEFAULT), H5Dclose);
    need(ds.id >= 0, "HDF5Io: no Mesh/Nodes dataset");
    dims_of(ds.id, 2, dn, "Mesh/Nodes");
  }
  {
    H ds(H5Dopen(g.id, "Cells", H5P_DEFAULT), H5Dclose);
    need(ds.id >= 0, "HDF5Io: no Mesh/Cells dataset");
    dims_of(ds.id, 2, dc, "Mesh/Cells");
  }
  out[0] = (long long)dn[0];
  out[1] = (long long)dn[1];
  out[2] = (long long)dc[0];
  out[3] = (long long)dc[1];
  IO_END
}

extern "C" int hfx_h5_read_mesh(const char* path, double* nodes, int* cells) {
  IO_TRY
  H f(open_file(path), H5Fclose);
  need(H5Lexists(f.id, "Mesh", H5P_DEFAULT) > 0, "HDF5Io: no Mesh group in " + std::string(path));
  H g(H5Gopen(f.id, "Mesh", H5P_DEFAULT), H5Gclose);
  if (nodes) {
    H ds(H5Dopen(g.id, "Nodes", H5P_DEFAULT), H5Dclose);
    need(ds.id >= 0 && H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, nodes) >= 0,
         "HDF5Io: could not read Mesh/Nodes");
  }
  if (cells) {
    H ds(H5Dopen(g.id, "Cells", H5P_DEFAULT), H5Dclose);
    need(ds.id >= 0 && H5Dread(ds.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, cells) >= 0,
         "HDF5Io: could not read Mesh/Cells");
  }
  IO_END
}

extern "C" int hfx_h5_field_info(const char* path, const char* name, long long* shape, int* ftype) {
  IO_TRY
  need(name != nullptr, "HDF5Io: null field name");
  H f(open_file(path), H5Fclose);
  need(H5Lexists(f.id, "FieldData", H5P_DEFAULT) > 0, "HDF5Io: no FieldData group in " + std::string(path));
  H g(H5Gopen(f.id, "FieldData", H5P_DEFAULT), H5Gclose);
  need(H5Lexists(g.id, name, H5P_DEFAULT) > 0,
       std::string("HDF5Io: field with name ") + name + " was not found in FieldData");
  H ds(H5Dopen(g.id, name, H5P_DEFAULT), H5Dclose);
  hsize_t d[3];
  dims_of(ds.id, 3, d, std::string("field ") + name);
  need(H5Aexists(ds.id, "ftype") > 0, std::string("HDF5Io: attribute ftype could not be found for field: ") + name);
  H a(H5Aopen_by_name(ds.id, ".", "ftype", H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  int t = -1;
  need(H5Aread(a.id, H5T_NATIVE_INT, &t) >= 0, std::string("HDF5Io: could not read ftype of ") + name);
  for (int k = 0; k < 3; ++k) shape[k] = (long long)d[k];
  if (ftype) *ftype = t;
  IO_END
}

extern "C" int hfx_h5_read_field(const char* path, const char* name, double* values) {
  IO_TRY
  need(name != nullptr && values != nullptr, "HDF5Io: null argument");
  H f(open_file(path), H5Fclose);
  need(H5Lexists(f.id, "FieldData", H5P_DEFAULT) > 0, "HDF5Io: no FieldData group in " + std::string(path));
  H g(H5Gopen(f.id, "FieldData", H5P_DEFAULT), H5Gclose);
  need(H5Lexists(g.id, name, H5P_DEFAULT) > 0,
       std::string("HDF5Io: field with name ") + name + " was not found in FieldData");
  H ds(H5Dopen(g.id, name, H5P_DEFAULT), H5Dclose);
  need(H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, values) >= 0,
       std::string("HDF5Io: could not load field values of ") + name);
  IO_END
}

extern "C" int hfx_h5_write(const char* path, int nNodes, int dim, const double* nodes, int nCells, int nPerCell,
                            const int* cells, int nFields, const char* const* names, const long long* shapes,
                            const int* ftypes, const double* const* values) {
  IO_TRY
  need(path != nullptr, "HDF5Io: null file name");
  need(nFields >= 0 && (nFields == 0 || (names && shapes && ftypes && values)), "HDF5Io: null field argument");
  H f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  need(f.id >= 0, std::string("HDF5Io: could not create ") + path);
  if (nodes && cells) {
    H g(H5Gcreate2(f.id, "Mesh", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    const hsize_t dn[2] = {(hsize_t)nNodes, (hsize_t)dim}, dc[2] = {(hsize_t)nCells, (hsize_t)nPerCell};
    H sn(H5Screate_simple(2, dn, nullptr), H5Sclose);
    H n(H5Dcreate2(g.id, "Nodes", H5T_NATIVE_DOUBLE, sn.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    need(n.id >= 0 && H5Dwrite(n.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, nodes) >= 0,
         "HDF5Io: could not write Mesh/Nodes");
    H sc(H5Screate_simple(2, dc, nullptr), H5Sclose);
    H c(H5Dcreate2(g.id, "Cells", H5T_NATIVE_INT, sc.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    need(c.id >= 0 && H5Dwrite(c.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, cells) >= 0,
         "HDF5Io: could not write Mesh/Cells");
  }
  if (nFields > 0) {
    H g(H5Gcreate2(f.id, "FieldData", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    for (int i = 0; i < nFields; ++i) {
      need(names[i] != nullptr && values[i] != nullptr, "HDF5Io: null field");
      need(ftypes[i] >= -1 && ftypes[i] <= 3, std::string("HDF5Io: unknown field type for ") + names[i]);
      const hsize_t d[3] = {(hsize_t)shapes[3 * i], (hsize_t)shapes[3 * i + 1], (hsize_t)shapes[3 * i + 2]};
      H sp(H5Screate_simple(3, d, nullptr), H5Sclose);
      H ds(H5Dcreate2(g.id, names[i], H5T_NATIVE_DOUBLE, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
      need(ds.id >= 0, std::string("HDF5Io: could not create field ") + names[i]);
      const hsize_t one = 1;
      H as(H5Screate_simple(1, &one, nullptr), H5Sclose);
      H a(H5Acreate2(ds.id, "ftype", H5T_NATIVE_INT, as.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
      need(a.id >= 0 && H5Awrite(a.id, H5T_NATIVE_INT, &ftypes[i]) >= 0,
           std::string("HDF5Io: problem writing ftype attribute for field: ") + names[i]);
      need(H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, values[i]) >= 0,
           std::string("HDF5Io: could not write field ") + names[i]);
    }
  }
  IO_END
}
