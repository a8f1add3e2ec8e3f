// h5io.cpp -- see h5io.h.  Uses the HDF5 C API (1.10 ABI: hid_t is 64-bit)
// through dlopen/dlsym.
#include "h5io.h"

#include <dlfcn.h>
#include <sys/stat.h>

#include <cstdint>
#include <cerrno>
#include <cstdio>

namespace h5io {
namespace {

using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;

struct Api {
    bool ok = false;
    herr_t (*open)();
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t);
    hid_t (*fopen)(const char*, unsigned, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*);
    herr_t (*sclose)(hid_t);
    int (*sget_ndims)(hid_t);
    int (*sget_dims)(hid_t, hsize_t*, hsize_t*);
    hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t);
    hid_t (*dopen2)(hid_t, const char*, hid_t);
    hid_t (*dget_space)(hid_t);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*);
    herr_t (*dclose)(hid_t);
    herr_t (*eset_auto2)(hid_t, void*, void*);
    hid_t* native_int;
    hid_t* native_float;
};

Api& api()
{
    static Api a;
    static bool tried = false;
    if (tried) return a;
    tried = true;
    const char* names[] = {"libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "libhdf5.so",
                           "/opt/conda/lib/libhdf5.so", "libhdf5_serial.so"};
    void* h = nullptr;
    for (const char* n : names)
        if ((h = dlopen(n, RTLD_NOW | RTLD_LOCAL))) break;
    if (!h) return a;
#define SYM(field, name) *(void**)(&a.field) = dlsym(h, name)
    SYM(open, "H5open");
    SYM(fcreate, "H5Fcreate");
    SYM(fopen, "H5Fopen");
    SYM(fclose, "H5Fclose");
    SYM(screate_simple, "H5Screate_simple");
    SYM(sclose, "H5Sclose");
    SYM(sget_ndims, "H5Sget_simple_extent_ndims");
    SYM(sget_dims, "H5Sget_simple_extent_dims");
    SYM(dcreate2, "H5Dcreate2");
    SYM(dopen2, "H5Dopen2");
    SYM(dget_space, "H5Dget_space");
    SYM(dwrite, "H5Dwrite");
    SYM(dread, "H5Dread");
    SYM(dclose, "H5Dclose");
    SYM(eset_auto2, "H5Eset_auto2");
    SYM(native_int, "H5T_NATIVE_INT_g");
    SYM(native_float, "H5T_NATIVE_FLOAT_g");
#undef SYM
    a.ok = a.open && a.fcreate && a.fopen && a.fclose && a.screate_simple && a.sclose &&
           a.sget_ndims && a.sget_dims && a.dcreate2 && a.dopen2 && a.dget_space && a.dwrite &&
           a.dread && a.dclose && a.native_int && a.native_float;
    if (a.ok) {
        a.open();  // initialises the native type ids
        if (a.eset_auto2) a.eset_auto2(0, nullptr, nullptr);
    }
    return a;
}

constexpr unsigned kAccTrunc = 0x0002u, kAccRdonly = 0x0000u;
constexpr hid_t kDefault = 0, kAll = 0;

}  // namespace

bool available() { return api().ok; }

Writer::Writer(const std::string& path) : path_(path)
{
    Api& a = api();
    if (a.ok) {
        file_ = a.fcreate(path.c_str(), kAccTrunc, kDefault, kDefault);
        ok_ = file_ >= 0;
    } else {
        raw_ = true;
        ok_ = mkdir((path + ".d").c_str(), 0755) == 0 || errno == EEXIST;
    }
}

Writer::~Writer()
{
    if (!raw_ && file_ >= 0) api().fclose(file_);
}

static bool write_any(long long file, bool raw, const std::string& path, const std::string& name,
                      const void* data, size_t n, size_t elem, bool is_int)
{
    if (raw) {
        FILE* f = std::fopen((path + ".d/" + name + ".bin").c_str(), "wb");
        if (!f) return false;
        const bool good = std::fwrite(data, elem, n, f) == n;
        std::fclose(f);
        return good;
    }
    Api& a = api();
    hsize_t dim = n;
    hid_t space = a.screate_simple(1, &dim, nullptr);
    if (space < 0) return false;
    const hid_t type = is_int ? *a.native_int : *a.native_float;
    hid_t ds = a.dcreate2(file, name.c_str(), type, space, kDefault, kDefault, kDefault);
    bool good = ds >= 0 && (n == 0 || a.dwrite(ds, type, kAll, kAll, kDefault, data) >= 0);
    if (ds >= 0) a.dclose(ds);
    a.sclose(space);
    return good;
}

bool Writer::write_f32(const std::string& name, const float* data, size_t n)
{
    return ok_ && write_any(file_, raw_, path_, name, data, n, sizeof(float), false);
}

bool Writer::write_i32(const std::string& name, const int* data, size_t n)
{
    return ok_ && write_any(file_, raw_, path_, name, data, n, sizeof(int), true);
}

bool read_f32(const std::string& path, const std::string& name, std::vector<float>& out)
{
    Api& a = api();
    if (!a.ok) return false;
    struct stat st;
    if (stat(path.c_str(), &st) != 0) return false;
    hid_t f = a.fopen(path.c_str(), kAccRdonly, kDefault);
    if (f < 0) return false;
    bool good = false;
    hid_t ds = a.dopen2(f, name.c_str(), kDefault);
    if (ds >= 0) {
        hid_t sp = a.dget_space(ds);
        hsize_t dims[8] = {0};
        const int nd = sp >= 0 ? a.sget_ndims(sp) : -1;
        if (nd >= 1 && nd <= 8 && a.sget_dims(sp, dims, nullptr) >= 0) {
            size_t n = 1;
            for (int q = 0; q < nd; ++q) n *= dims[q];
            out.resize(n);
            good = n == 0 || a.dread(ds, *a.native_float, kAll, kAll, kDefault, out.data()) >= 0;
        }
        if (sp >= 0) a.sclose(sp);
        a.dclose(ds);
    }
    a.fclose(f);
    return good;
}

}  // namespace h5io
