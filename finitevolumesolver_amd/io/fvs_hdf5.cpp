// HDF5 output of a composite uniform-grid state, in the layout the
// reference writes (src/AMReX/PatchHierarchy.cpp:442-600:
// CreateHdf5Database on the first call, OpenHdf5Database appending a frame
// afterwards; called from WriteToHDF5, :598-607).  The file holds
//   /data    float64 (frames, ncomp, extent0, extent1), chunked one frame,
//            unlimited in frames; each frame is the composite buffer as it
//            lies in memory (component-major, x fastest), attributes
//            xlower and cell_size (two float64 each);
//   /fields  the component names (variable-length strings);
//   /times   float64 per frame, /cycles int64 per frame (unlimited).
// Host code only: it links libhdf5 and is built apart from the device
// library.
#include "fvs_io.h"

#include <hdf5.h>
#include <sys/stat.h>

namespace {

struct Closer {
  hid_t id;
  herr_t (*close)(hid_t);
  ~Closer() {
    if (id >= 0) close(id);
  }
};

int AppendScalar(hid_t file, const char* name, hid_t type, const void* value) {
  Closer ds{H5Dopen2(file, name, H5P_DEFAULT), H5Dclose};
  if (ds.id < 0) return -1;
  hsize_t dims = 0, maxdims = 0;
  {
    Closer sp{H5Dget_space(ds.id), H5Sclose};
    H5Sget_simple_extent_dims(sp.id, &dims, &maxdims);
  }
  const hsize_t grown = dims + 1, count = 1, offset = dims;
  if (H5Dset_extent(ds.id, &grown) < 0) return -1;
  Closer fs{H5Dget_space(ds.id), H5Sclose};
  H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, &offset, nullptr, &count, nullptr);
  Closer ms{H5Screate_simple(1, &count, nullptr), H5Sclose};
  return H5Dwrite(ds.id, type, ms.id, fs.id, H5P_DEFAULT, value) < 0 ? -1 : 0;
}

int CreateScalar(hid_t file, const char* name, hid_t type, const void* value) {
  hsize_t dims = 1, maxdims = H5S_UNLIMITED;
  Closer prop{H5Pcreate(H5P_DATASET_CREATE), H5Pclose};
  H5Pset_chunk(prop.id, 1, &dims);
  Closer sp{H5Screate_simple(1, &dims, &maxdims), H5Sclose};
  Closer ds{H5Dcreate2(file, name, type, sp.id, H5P_DEFAULT, prop.id, H5P_DEFAULT), H5Dclose};
  if (ds.id < 0) return -1;
  return H5Dwrite(ds.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, value) < 0 ? -1 : 0;
}

int Create(const char* path, const double* data, int64_t ncomp, int64_t n0, int64_t n1,
           const double* xlower, const double* cell_size, const char* const* fields,
           int nfields, double time, int64_t cycle) {
  Closer file{H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
  if (file.id < 0) return -1;
  {
    hsize_t dims[4] = {1, hsize_t(ncomp), hsize_t(n0), hsize_t(n1)};
    hsize_t maxdims[4] = {H5S_UNLIMITED, hsize_t(ncomp), hsize_t(n0), hsize_t(n1)};
    Closer prop{H5Pcreate(H5P_DATASET_CREATE), H5Pclose};
    H5Pset_chunk(prop.id, 4, dims);
    H5Pset_alloc_time(prop.id, H5D_ALLOC_TIME_EARLY);
    Closer sp{H5Screate_simple(4, dims, maxdims), H5Sclose};
    Closer ds{H5Dcreate2(file.id, "/data", H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, prop.id,
                         H5P_DEFAULT), H5Dclose};
    if (ds.id < 0) return -1;
    if (H5Dwrite(ds.id, H5T_IEEE_F64LE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) return -1;
    const hsize_t two = 2;
    Closer s2{H5Screate_simple(1, &two, nullptr), H5Sclose};
    Closer xl{H5Acreate2(ds.id, "xlower", H5T_IEEE_F64LE, s2.id, H5P_DEFAULT, H5P_DEFAULT),
              H5Aclose};
    Closer cs{H5Acreate2(ds.id, "cell_size", H5T_IEEE_F64LE, s2.id, H5P_DEFAULT, H5P_DEFAULT),
              H5Aclose};
    if (H5Awrite(xl.id, H5T_IEEE_F64LE, xlower) < 0) return -1;
    if (H5Awrite(cs.id, H5T_IEEE_F64LE, cell_size) < 0) return -1;
  }
  {
    const hsize_t n = hsize_t(nfields);
    Closer sp{H5Screate_simple(1, &n, nullptr), H5Sclose};
    Closer type{H5Tcopy(H5T_C_S1), H5Tclose};
    H5Tset_size(type.id, H5T_VARIABLE);
    Closer ds{H5Dcreate2(file.id, "/fields", type.id, sp.id, H5P_DEFAULT, H5P_DEFAULT,
                         H5P_DEFAULT), H5Dclose};
    if (ds.id < 0) return -1;
    if (nfields > 0 &&
        H5Dwrite(ds.id, type.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, fields) < 0)
      return -1;
  }
  if (CreateScalar(file.id, "/times", H5T_IEEE_F64LE, &time)) return -1;
  return CreateScalar(file.id, "/cycles", H5T_STD_I64LE, &cycle);
}

int Append(const char* path, const double* data, int64_t ncomp, int64_t n0, int64_t n1,
           double time, int64_t cycle) {
  Closer file{H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose};
  if (file.id < 0) return -1;
  {
    Closer ds{H5Dopen2(file.id, "/data", H5P_DEFAULT), H5Dclose};
    if (ds.id < 0) return -1;
    hsize_t dims[4] = {}, maxdims[4] = {};
    {
      Closer sp{H5Dget_space(ds.id), H5Sclose};
      if (H5Sget_simple_extent_ndims(sp.id) != 4) return -1;
      H5Sget_simple_extent_dims(sp.id, dims, maxdims);
    }
    if (dims[1] != hsize_t(ncomp) || dims[2] != hsize_t(n0) || dims[3] != hsize_t(n1))
      return -2;
    hsize_t grown[4] = {dims[0] + 1, dims[1], dims[2], dims[3]};
    if (H5Dset_extent(ds.id, grown) < 0) return -1;
    Closer fs{H5Dget_space(ds.id), H5Sclose};
    hsize_t offset[4] = {dims[0], 0, 0, 0};
    hsize_t count[4] = {1, dims[1], dims[2], dims[3]};
    H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, offset, nullptr, count, nullptr);
    Closer ms{H5Screate_simple(4, count, nullptr), H5Sclose};
    if (H5Dwrite(ds.id, H5T_IEEE_F64LE, ms.id, fs.id, H5P_DEFAULT, data) < 0) return -1;
  }
  if (AppendScalar(file.id, "/times", H5T_IEEE_F64LE, &time)) return -1;
  return AppendScalar(file.id, "/cycles", H5T_STD_I64LE, &cycle);
}

} // namespace

extern "C" int fvsio_hdf5_write(const char* path, const double* data, int64_t ncomp,
                                int64_t n0, int64_t n1, const double* xlower,
                                const double* cell_size, const char* const* fields,
                                int nfields, double time, int64_t cycle) {
  if (!path || !data || ncomp < 1 || n0 < 1 || n1 < 1 || !xlower || !cell_size ||
      (nfields > 0 && !fields))
    return -3;
  struct stat st;
  if (stat(path, &st) != 0)
    return Create(path, data, ncomp, n0, n1, xlower, cell_size, fields, nfields, time, cycle);
  if (S_ISREG(st.st_mode)) return Append(path, data, ncomp, n0, n1, time, cycle);
  return -4;
}
