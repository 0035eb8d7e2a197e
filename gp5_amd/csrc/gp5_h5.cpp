// HDF5 field files in the reference's on-disk format (src/io.cpp:32-45,
// 56-152, 154-254): one dataset "field" of type H5T_ARRAY double[2] (re, im)
// with dims (nC, Nz, Ny, Nx), i.e. the column-major (x, y, z, c) complex
// tensor written as-is.  One GPU means one writer, so the file is written
// serially (the reference uses MPI-IO hyperslabs of the same dataset); the
// bytes on disk are the same.
//
// libhdf5 is loaded at run time with dlopen, so libgp5.so has no link-time
// dependency on it: GP5_HDF5_LIB names the library, else the default search
// path and /opt/conda/lib are tried.  Only the HDF5 1.10 C API functions below
// are used.  Host code only.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "gp5.h"
#include "gp5_internal.h"

namespace {

using gp5::set_error;

using hid_t = int64_t;       // H5Ipublic.h (1.10)
using herr_t = int;          // H5public.h
using hsize_t = unsigned long long;

constexpr unsigned kAccRdonly = 0x0000u;  // H5F_ACC_RDONLY
constexpr unsigned kAccTrunc = 0x0002u;   // H5F_ACC_TRUNC
constexpr hid_t kDefault = 0;             // H5P_DEFAULT, H5S_ALL, H5E_DEFAULT
constexpr int kClassFloat = 1;            // H5T_FLOAT
constexpr int kClassArray = 10;           // H5T_ARRAY

struct H5 {
    void *so = nullptr;
    std::string error;
    herr_t (*open)();
    herr_t (*eset_auto)(hid_t, void *, void *);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*tarray_create)(hid_t, unsigned, const hsize_t *);
    herr_t (*tclose)(hid_t);
    int (*tget_class)(hid_t);
    int (*tget_array_ndims)(hid_t);
    int (*tget_array_dims)(hid_t, hsize_t *);
    hid_t (*tget_super)(hid_t);
    size_t (*tget_size)(hid_t);
    hid_t (*dcreate)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen)(hid_t, const char *, hid_t);
    herr_t (*dclose)(hid_t);
    hid_t (*dget_type)(hid_t);
    hid_t (*dget_space)(hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    hid_t *native_double;  // H5T_NATIVE_DOUBLE_g, valid after H5open()
};

template <class F>
bool sym(H5 &h, F &f, const char *name)
{
    f = reinterpret_cast<F>(dlsym(h.so, name));
    if (!f) h.error = std::string("libhdf5 lacks ") + name;
    return f != nullptr;
}

H5 *hdf5(std::string *why)
{
    static H5 h;
    static std::once_flag once;
    std::call_once(once, [] {
        const char *env = std::getenv("GP5_HDF5_LIB");
        const char *cands[] = {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"};
        for (const char *c : cands)
            if (c && *c && (h.so = dlopen(c, RTLD_NOW | RTLD_LOCAL))) break;
        if (!h.so) {
            h.error = "libhdf5 not found (set GP5_HDF5_LIB to its path)";
            return;
        }
        bool ok = sym(h, h.open, "H5open") && sym(h, h.eset_auto, "H5Eset_auto2") &&
                  sym(h, h.fcreate, "H5Fcreate") && sym(h, h.fopen, "H5Fopen") && sym(h, h.fclose, "H5Fclose") &&
                  sym(h, h.screate_simple, "H5Screate_simple") && sym(h, h.sclose, "H5Sclose") &&
                  sym(h, h.sget_ndims, "H5Sget_simple_extent_ndims") &&
                  sym(h, h.sget_dims, "H5Sget_simple_extent_dims") &&
                  sym(h, h.tarray_create, "H5Tarray_create2") && sym(h, h.tclose, "H5Tclose") &&
                  sym(h, h.tget_class, "H5Tget_class") && sym(h, h.tget_array_ndims, "H5Tget_array_ndims") &&
                  sym(h, h.tget_array_dims, "H5Tget_array_dims2") && sym(h, h.tget_super, "H5Tget_super") &&
                  sym(h, h.tget_size, "H5Tget_size") && sym(h, h.dcreate, "H5Dcreate2") &&
                  sym(h, h.dopen, "H5Dopen2") && sym(h, h.dclose, "H5Dclose") &&
                  sym(h, h.dget_type, "H5Dget_type") && sym(h, h.dget_space, "H5Dget_space") &&
                  sym(h, h.dwrite, "H5Dwrite") && sym(h, h.dread, "H5Dread") &&
                  sym(h, h.native_double, "H5T_NATIVE_DOUBLE_g");
        if (!ok || h.open() < 0) {
            if (h.error.empty()) h.error = "H5open failed";
            h.so = nullptr;
            return;
        }
        h.eset_auto(kDefault, nullptr, nullptr);  // errors come back as codes, not stderr dumps
    });
    if (!h.so && why) *why = h.error;
    return h.so ? &h : nullptr;
}

// closes every id it holds, in reverse order of acquisition
struct Ids {
    H5 *h;
    hid_t f = -1, s = -1, t = -1, d = -1, mt = -1, super = -1;
    ~Ids()
    {
        if (super >= 0) h->tclose(super);
        if (mt >= 0) h->tclose(mt);
        if (d >= 0) h->dclose(d);
        if (t >= 0) h->tclose(t);
        if (s >= 0) h->sclose(s);
        if (f >= 0) h->fclose(f);
    }
};

int need_lib(H5 *&h)
{
    std::string why;
    h = hdf5(&why);
    if (h) return GP5_OK;
    return set_error(GP5_ERR_UNSUPPORTED, ("HDF5 I/O unavailable: " + why).c_str());
}

// Reads the dataset "field" of an open file: its dims (nC, Nz, Ny, Nx) and
// whether it is the reference's array type (double[2] elements) or a plain
// double dataset with a trailing dimension 2 (as h5py writes complex pairs).
int inspect(Ids &I, const char *filename, int dims[4], bool &array_type)
{
    H5 *h = I.h;
    I.f = h->fopen(filename, kAccRdonly, kDefault);
    if (I.f < 0) return set_error(GP5_ERR_ARG, (std::string("cannot open HDF5 file ") + filename).c_str());
    I.d = h->dopen(I.f, "field", kDefault);
    if (I.d < 0) return set_error(GP5_ERR_ARG, (std::string(filename) + ": no dataset \"field\"").c_str());
    I.t = h->dget_type(I.d);
    I.s = h->dget_space(I.d);
    if (I.t < 0 || I.s < 0) return set_error(GP5_ERR_ARG, "cannot query dataset \"field\"");
    hsize_t sd[8] = {0};
    int rank = h->sget_ndims(I.s);
    if (rank < 1 || rank > 8 || h->sget_dims(I.s, sd, nullptr) < 0)
        return set_error(GP5_ERR_ARG, "dataset \"field\" has an unreadable dataspace");
    int cls = h->tget_class(I.t);
    if (cls == kClassArray) {
        hsize_t ad[4] = {0};
        if (h->tget_array_ndims(I.t) != 1 || h->tget_array_dims(I.t, ad) != 1 || ad[0] != 2)
            return set_error(GP5_ERR_ARG, "dataset \"field\": array element is not double[2]");
        I.super = h->tget_super(I.t);
        if (I.super < 0 || h->tget_class(I.super) != kClassFloat || h->tget_size(I.super) != 8)
            return set_error(GP5_ERR_ARG, "dataset \"field\": array element is not double[2]");
        if (rank != 4) return set_error(GP5_ERR_ARG, "dataset \"field\" must have rank 4 (nC, Nz, Ny, Nx)");
        array_type = true;
    } else if (cls == kClassFloat) {
        if (rank != 5 || sd[4] != 2)
            return set_error(GP5_ERR_ARG, "dataset \"field\" of doubles must have dims (nC, Nz, Ny, Nx, 2)");
        array_type = false;
    } else {
        return set_error(GP5_ERR_ARG, "dataset \"field\" is neither double[2] elements nor doubles");
    }
    for (int i = 0; i < 4; ++i) {
        if (sd[i] == 0 || sd[i] > (1u << 30)) return set_error(GP5_ERR_ARG, "dataset \"field\" has bad dims");
        dims[i] = (int)sd[i];
    }
    return GP5_OK;
}

}  // namespace

extern "C" {

int gp5_h5_write(const char *filename, const double *host, const int shape[3], int nc)
{
    if (!filename || !host || !shape || nc < 1) return set_error(GP5_ERR_ARG, "gp5_h5_write: null argument");
    for (int d = 0; d < 3; ++d)
        if (shape[d] < 1) return set_error(GP5_ERR_ARG, "gp5_h5_write: bad shape");
    H5 *h;
    if (int rc = need_lib(h)) return rc;
    Ids I{h};
    const hsize_t pair[1] = {2};
    const hsize_t dims[4] = {(hsize_t)nc, (hsize_t)shape[2], (hsize_t)shape[1], (hsize_t)shape[0]};
    I.t = h->tarray_create(*h->native_double, 1, pair);
    I.f = h->fcreate(filename, kAccTrunc, kDefault, kDefault);
    if (I.f < 0) return set_error(GP5_ERR_ARG, (std::string("cannot create HDF5 file ") + filename).c_str());
    I.s = h->screate_simple(4, dims, nullptr);
    if (I.t < 0 || I.s < 0) return set_error(GP5_ERR_ARG, "gp5_h5_write: cannot build type or dataspace");
    I.d = h->dcreate(I.f, "field", I.t, I.s, kDefault, kDefault, kDefault);
    if (I.d < 0) return set_error(GP5_ERR_ARG, "gp5_h5_write: cannot create dataset \"field\"");
    if (h->dwrite(I.d, I.t, kDefault, kDefault, kDefault, host) < 0)
        return set_error(GP5_ERR_ARG, "gp5_h5_write: write failed");
    return GP5_OK;
}

int gp5_h5_shape(const char *filename, int shape[3], int *nc)
{
    if (!filename || !shape || !nc) return set_error(GP5_ERR_ARG, "gp5_h5_shape: null argument");
    H5 *h;
    if (int rc = need_lib(h)) return rc;
    Ids I{h};
    int dims[4];
    bool array_type;
    if (int rc = inspect(I, filename, dims, array_type)) return rc;
    *nc = dims[0];
    shape[0] = dims[3];
    shape[1] = dims[2];
    shape[2] = dims[1];
    return GP5_OK;
}

int gp5_h5_read(const char *filename, double *host, const int shape[3], int nc)
{
    if (!filename || !host || !shape || nc < 1) return set_error(GP5_ERR_ARG, "gp5_h5_read: null argument");
    H5 *h;
    if (int rc = need_lib(h)) return rc;
    Ids I{h};
    int dims[4];
    bool array_type;
    if (int rc = inspect(I, filename, dims, array_type)) return rc;
    if (dims[0] != nc || dims[1] != shape[2] || dims[2] != shape[1] || dims[3] != shape[0]) {
        std::string m = std::string(filename) + ": dataset \"field\" has dims (" + std::to_string(dims[0]) + "," +
                        std::to_string(dims[1]) + "," + std::to_string(dims[2]) + "," + std::to_string(dims[3]) +
                        "), expected (" + std::to_string(nc) + "," + std::to_string(shape[2]) + "," +
                        std::to_string(shape[1]) + "," + std::to_string(shape[0]) + ")";
        return set_error(GP5_ERR_ARG, m.c_str());
    }
    // the memory type matches the file's element layout, so the bytes land as
    // interleaved (re, im) pairs in either case
    hid_t mem = *h->native_double;
    if (array_type) {
        const hsize_t pair[1] = {2};
        I.mt = h->tarray_create(*h->native_double, 1, pair);
        if (I.mt < 0) return set_error(GP5_ERR_ARG, "gp5_h5_read: cannot build memory type");
        mem = I.mt;
    }
    if (h->dread(I.d, mem, kDefault, kDefault, kDefault, host) < 0)
        return set_error(GP5_ERR_ARG, "gp5_h5_read: read failed");
    return GP5_OK;
}

}  // extern "C"
