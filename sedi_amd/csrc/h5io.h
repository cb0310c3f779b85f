// h5io.h -- validation archives: 3-D FP64 datasets in HDF5 files.
//
// The reference's validation modes read datasets "x_k", "v_k" (ValidateSolver)
// and "diag_inv" (ValidateEstimator) of shape [n1][n2][n3] with i3 fastest
// (Slice3d::pack_into / unpack_from, Slice3d.hpp:42-66; load_hdf5 :144-160,
// native double :169-171).  Arrays here are i1 fastest, as everywhere else in
// the library, and are transposed on the way in and out.
//
// libhdf5 is loaded at run time (dlopen), so the library has no link-time
// dependency on it; without it these calls fail with SEDI_ERR_UNSUPPORTED.
#pragma once

#include "sedi_internal.h"

namespace sedi {

bool h5_available();
// Reads dataset `name` (must be [n1][n2][n3]) into out (i1 fastest).
int h5_read(const char* path, const char* name, const Shape& s, double* out);
// Creates (truncates) `path` and writes the datasets, each an i1-fastest array.
int h5_write(const char* path, int count, const char* const* names, const double* const* arrays,
             const Shape& s);

}  // namespace sedi
