/* Host-side output of libfvs_hdf5.so (built apart from the device library,
 * against the HDF5 C library).  No device code. */
#ifndef FVS_IO_H
#define FVS_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Writes one frame of a composite uniform-grid state to an HDF5 file in the
 * reference's layout (WriteToHDF5 -> CreateHdf5Database / OpenHdf5Database,
 * src/AMReX/PatchHierarchy.cpp:442-607): a new file gets /data (frames x
 * ncomp x n0 x n1 float64, xlower and cell_size attributes), /fields,
 * /times and /cycles; an existing file gets the frame, time and cycle
 * appended.  data: host, ncomp * n0 * n1 doubles, component-major, x
 * fastest (n0 is the x extent).  Returns 0, -1 (HDF5 error), -2 (frame
 * shape differs from the file's), -3 (bad argument) or -4 (path is not a
 * regular file). */
int fvsio_hdf5_write(const char* path, const double* data, int64_t ncomp, int64_t n0,
                     int64_t n1, const double* xlower, const double* cell_size,
                     const char* const* fields, int nfields, double time, int64_t cycle);

#ifdef __cplusplus
}
#endif

#endif /* FVS_IO_H */
