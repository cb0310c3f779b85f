// h5io.cpp -- HDF5 validation archives through a run-time loaded libhdf5.
//
// Only the HDF5 1.10 C API calls the reference's Slice3d::load_hdf5 makes
// (H5Fopen, H5Dopen2, H5Dget_space, H5Sget_simple_extent_*, H5Dread) plus
// their writing counterparts (H5Fcreate, H5Screate_simple, H5Dcreate2,
// H5Dwrite) are used.  The constants are the stable values of the 1.10 API.
#include "h5io.h"

#include <dlfcn.h>

#include <cstdint>
#include <mutex>
#include <vector>

namespace sedi {

namespace {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;
constexpr unsigned ACC_RDONLY = 0x0000u, ACC_TRUNC = 0x0002u;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0;

struct H5 {
    void* so = nullptr;
    herr_t (*open)();
    hid_t (*fopen)(const char*, unsigned, hid_t);
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*dopen2)(hid_t, const char*, hid_t);
    hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dget_space)(hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dclose)(hid_t);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t*, hsize_t*);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    herr_t (*sclose)(hid_t);
    herr_t (*eset_auto2)(hid_t, void*, void*);
    hid_t* native_double;
    bool ok = false;
};

template <class F>
bool sym(void* so, const char* name, F& f) {
    f = reinterpret_cast<F>(dlsym(so, name));
    return f != nullptr;
}

H5& h5() {
    static H5 h;
    static std::once_flag once;
    std::call_once(once, [] {
        const char* names[] = {"libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so.103",
                               "/opt/conda/lib/libhdf5.so"};
        for (const char* n : names)
            if ((h.so = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
        if (!h.so) return;
        bool ok = sym(h.so, "H5open", h.open) && sym(h.so, "H5Fopen", h.fopen) &&
                  sym(h.so, "H5Fcreate", h.fcreate) && sym(h.so, "H5Fclose", h.fclose) &&
                  sym(h.so, "H5Dopen2", h.dopen2) && sym(h.so, "H5Dcreate2", h.dcreate2) &&
                  sym(h.so, "H5Dget_space", h.dget_space) && sym(h.so, "H5Dread", h.dread) &&
                  sym(h.so, "H5Dwrite", h.dwrite) && sym(h.so, "H5Dclose", h.dclose) &&
                  sym(h.so, "H5Sget_simple_extent_ndims", h.sget_ndims) &&
                  sym(h.so, "H5Sget_simple_extent_dims", h.sget_dims) &&
                  sym(h.so, "H5Screate_simple", h.screate_simple) && sym(h.so, "H5Sclose", h.sclose) &&
                  sym(h.so, "H5Eset_auto2", h.eset_auto2) &&
                  sym(h.so, "H5T_NATIVE_DOUBLE_g", h.native_double);
        if (!ok || h.open() < 0) return;
        h.eset_auto2(0 /* H5E_DEFAULT */, nullptr, nullptr);  // errors are reported here instead
        h.ok = true;
    });
    return h;
}

// i1-fastest (library) <-> i3-fastest (archive) for an n1 x n2 x n3 grid.
void to_archive(const double* src, double* dst, const Shape& s) {
    long long idx = 0;
    for (long long i1 = 0; i1 < s.n1; i1++)
        for (long long i2 = 0; i2 < s.n2; i2++)
            for (long long i3 = 0; i3 < s.n3; i3++)
                dst[idx++] = src[i1 + s.n1 * (i2 + s.n2 * i3)];
}

void from_archive(const double* src, double* dst, const Shape& s) {
    long long idx = 0;
    for (long long i1 = 0; i1 < s.n1; i1++)
        for (long long i2 = 0; i2 < s.n2; i2++)
            for (long long i3 = 0; i3 < s.n3; i3++)
                dst[i1 + s.n1 * (i2 + s.n2 * i3)] = src[idx++];
}

}  // namespace

bool h5_available() { return h5().ok; }

int h5_read(const char* path, const char* name, const Shape& s, double* out) {
    H5& h = h5();
    SEDI_REQUIRE(h.ok, SEDI_ERR_UNSUPPORTED, "libhdf5 is not available");
    const hid_t f = h.fopen(path, ACC_RDONLY, P_DEFAULT);
    SEDI_REQUIRE(f >= 0, SEDI_ERR_ARG, "cannot open HDF5 file %s", path);
    const hid_t d = h.dopen2(f, name, P_DEFAULT);
    if (d < 0) {
        h.fclose(f);
        set_error("no dataset %s in %s", name, path);
        return SEDI_ERR_ARG;
    }
    const hid_t sp = h.dget_space(d);
    hsize_t dims[3] = {0, 0, 0};
    const int nd = h.sget_ndims(sp);
    if (nd == 3) h.sget_dims(sp, dims, nullptr);
    h.sclose(sp);
    int rc = SEDI_OK;
    if (nd != 3 || dims[0] != (hsize_t)s.n1 || dims[1] != (hsize_t)s.n2 || dims[2] != (hsize_t)s.n3) {
        set_error("dataset %s in %s is not a %d x %d x %d grid", name, path, s.n1, s.n2, s.n3);
        rc = SEDI_ERR_ARG;
    } else {
        std::vector<double> buf((size_t)s.n());
        if (h.dread(d, *h.native_double, S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0) {
            set_error("reading %s from %s failed", name, path);
            rc = SEDI_ERR_ARG;
        } else {
            from_archive(buf.data(), out, s);
        }
    }
    h.dclose(d);
    h.fclose(f);
    return rc;
}

int h5_write(const char* path, int count, const char* const* names, const double* const* arrays,
             const Shape& s) {
    H5& h = h5();
    SEDI_REQUIRE(h.ok, SEDI_ERR_UNSUPPORTED, "libhdf5 is not available");
    const hid_t f = h.fcreate(path, ACC_TRUNC, P_DEFAULT, P_DEFAULT);
    SEDI_REQUIRE(f >= 0, SEDI_ERR_ARG, "cannot create HDF5 file %s", path);
    const hsize_t dims[3] = {(hsize_t)s.n1, (hsize_t)s.n2, (hsize_t)s.n3};
    std::vector<double> buf((size_t)s.n());
    int rc = SEDI_OK;
    for (int k = 0; k < count && rc == SEDI_OK; k++) {
        const hid_t sp = h.screate_simple(3, dims, nullptr);
        const hid_t d = h.dcreate2(f, names[k], *h.native_double, sp, P_DEFAULT, P_DEFAULT, P_DEFAULT);
        if (d < 0) {
            set_error("cannot create dataset %s in %s", names[k], path);
            rc = SEDI_ERR_ARG;
        } else {
            to_archive(arrays[k], buf.data(), s);
            if (h.dwrite(d, *h.native_double, S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0) {
                set_error("writing %s to %s failed", names[k], path);
                rc = SEDI_ERR_ARG;
            }
            h.dclose(d);
        }
        h.sclose(sp);
    }
    h.fclose(f);
    return rc;
}

}  // namespace sedi
