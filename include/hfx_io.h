/* HDF5 mesh and field files in the reference's layout (libhfx_io.so, host
 * only; replaces HDF5Io, src/io/HDF5Io.cpp:13-390):
 *   /Mesh/Nodes      nNodes x dim doubles
 *   /Mesh/Cells      nCells x nodes-per-cell ints (reference element order)
 *   /FieldData/NAME  nEntities x objects-per-entity x values-per-object
 *                    doubles with an int attribute "ftype" (FieldType:
 *                    None -1, Node 0, Edge 1, Face 2, Cell 3)
 * Every function returns 0 on success, nonzero on error with the message in
 * hfx_io_last_error() (the texts of HDF5Io's ErrorHandle throws). */
#ifndef HFX_IO_H
#define HFX_IO_H

#ifdef __cplusplus
extern "C" {
#endif

const char* hfx_io_last_error(void);

/* HDF5Io::load, mesh part (HDF5Io.cpp:111-153): out[0..3] = nNodes, dim,
 * nCells, nodes per cell. */
int hfx_h5_mesh_sizes(const char* path, long long* out);
/* Read /Mesh/Nodes (nNodes x dim) and /Mesh/Cells; either may be NULL. */
int hfx_h5_read_mesh(const char* path, double* nodes, int* cells);

/* HDF5Io::loadFields (HDF5Io.cpp:154-192): shape[0..2] and the ftype
 * attribute of field `name`; errors when the field or its ftype is missing. */
int hfx_h5_field_info(const char* path, const char* name, long long* shape, int* ftype);
int hfx_h5_read_field(const char* path, const char* name, double* values);

/* HDF5Io::write (HDF5Io.cpp:55-108, 253-390): a new file (truncating) with
 * the mesh (skipped when nodes or cells is NULL) and nFields fields
 * (shapes nFields x 3, ftypes, values). */
int hfx_h5_write(const char* path, int nNodes, int dim, const double* nodes, int nCells, int nPerCell,
                 const int* cells, int nFields, const char* const* names, const long long* shapes, const int* ftypes,
                 const double* const* values);

#ifdef __cplusplus
}
#endif
#endif
