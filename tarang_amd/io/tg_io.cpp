// tg_io.cpp -- field dumps in the reference's HDF5 layout (include/tarang_io.h).
//
// Host code over the HDF5 C library.  A dump goes through one host buffer in
// the file's shape: the array is copied from the device when it lives there,
// gathered into the plan's file layout, and written as one dataset; a read
// does the reverse and checks the dataset's dimensions first.
#include <hdf5.h>
#include <hip/hip_runtime_api.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/tarang_gpu.h"
#include "../../include/tarang_io.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

struct Shape {
    size_t mem[3];   // memory array extents
    size_t file[3];  // dataset extents
    bool complex_;
    size_t elem() const { return complex_ ? 2 : 1; }  // doubles per element
    size_t mem_count() const { return mem[0] * mem[1] * mem[2]; }
    size_t file_count() const { return file[0] * file[1] * file[2]; }
};

// file index -> memory index along one axis of a plan, -1 when the file entry
// has no memory counterpart (the reduced Nyquist index of a Fourier axis).
// BasicIO_SP.h:183-196: a Fourier axis keeps file [0, R/2) <-> memory
// [0, R/2) and file [R/2+1, R) <-> memory [N-R/2+1, N); other axes (and the
// half-complex z axis) keep the first R entries.
long axis_map(size_t f, size_t n_mem, size_t r_file, bool fourier, bool reduced)
{
    if (!reduced) return (long)f;
    if (!fourier) return (long)f;
    const size_t h = r_file / 2;
    if (f < h) return (long)f;
    if (f == h) return -1;
    return (long)(f + n_mem - r_file);
}

int plan_shape(const tgio_layout *L, int plan, Shape &s, bool fourier[3], bool &reduced)
{
    if (!L) return fail(TG_ERR_ARG, "null layout");
    if (L->nx < 1 || L->ny < 1 || L->nz < 2) return fail(TG_ERR_SHAPE, "bad grid %dx%dx%d", L->nx, L->ny, L->nz);
    const size_t nzc = (size_t)L->nz / 2 + 1;
    fourier[0] = L->fourier_x != 0;
    fourier[1] = true;
    fourier[2] = false;  // the half-complex z axis keeps its first entries
    reduced = plan == TGIO_REDUCED || plan == TGIO_KZ0_REDUCED;
    s.complex_ = plan != TGIO_REAL;
    s.mem[0] = L->nx;
    s.mem[1] = L->ny;
    s.mem[2] = s.complex_ ? nzc : (size_t)L->nz + 2;
    switch (plan) {
        case TGIO_FULL: s.file[0] = L->nx; s.file[1] = L->ny; s.file[2] = nzc; break;
        case TGIO_KZ0_FULL: s.file[0] = L->nx; s.file[1] = L->ny; s.file[2] = 1; break;
        case TGIO_REAL: s.file[0] = L->nx; s.file[1] = L->ny; s.file[2] = (size_t)L->nz + 2; break;
        case TGIO_REDUCED:
        case TGIO_KZ0_REDUCED: {
            const int *R = L->reduced;
            if (R[0] < 1 || R[1] < 1 || R[2] < 2 || R[0] > L->nx || R[1] > L->ny || R[2] > L->nz)
                return fail(TG_ERR_SHAPE, "reduced size %dx%dx%d must lie within the grid %dx%dx%d", R[0], R[1], R[2],
                            L->nx, L->ny, L->nz);
            s.file[0] = R[0];
            s.file[1] = R[1];
            s.file[2] = plan == TGIO_REDUCED ? (size_t)R[2] / 2 + 1 : 1;
            break;
        }
        default: return fail(TG_ERR_ARG, "unknown IO plan %d", plan);
    }
    return TG_OK;
}

// host copy of the memory array
int fetch(const double *data, const tgio_layout *L, const Shape &s, std::vector<double> &h)
{
    h.resize(s.mem_count() * s.elem());
    const size_t bytes = h.size() * sizeof(double);
    if (L->on_device) {
        hipError_t e = hipMemcpy(h.data(), data, bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(TG_ERR_HIP, "hipMemcpy D2H failed: %s", hipGetErrorString(e));
    } else {
        std::memcpy(h.data(), data, bytes);
    }
    return TG_OK;
}

int store(double *data, const tgio_layout *L, const std::vector<double> &h)
{
    const size_t bytes = h.size() * sizeof(double);
    if (L->on_device) {
        hipError_t e = hipMemcpy(data, h.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(TG_ERR_HIP, "hipMemcpy H2D failed: %s", hipGetErrorString(e));
    } else {
        std::memcpy(data, h.data(), bytes);
    }
    return TG_OK;
}

// visits every file element with its memory index (or -1)
template <typename F>
void for_each_entry(const Shape &s, const bool fourier[3], bool reduced, F fn)
{
    size_t fi = 0;
    for (size_t a = 0; a < s.file[0]; a++) {
        const long ma = axis_map(a, s.mem[0], s.file[0], fourier[0], reduced);
        for (size_t b = 0; b < s.file[1]; b++) {
            const long mb = axis_map(b, s.mem[1], s.file[1], fourier[1], reduced);
            for (size_t c = 0; c < s.file[2]; c++, fi++) {
                const long mc = (long)c;  // z: first entries
                if (ma < 0 || mb < 0) { fn(fi, -1L); continue; }
                fn(fi, (long)(((size_t)ma * s.mem[1] + (size_t)mb) * s.mem[2] + (size_t)mc));
            }
        }
    }
}

struct H5Types {
    hid_t mem = -1, file = -1;
    explicit H5Types(bool cplx)
    {
        if (cplx) {
            mem = H5Tcreate(H5T_COMPOUND, 2 * sizeof(double));
            H5Tinsert(mem, "r", 0, H5T_NATIVE_DOUBLE);
            H5Tinsert(mem, "i", sizeof(double), H5T_NATIVE_DOUBLE);
            file = H5Tcreate(H5T_COMPOUND, 16);
            H5Tinsert(file, "r", 0, H5T_IEEE_F64LE);
            H5Tinsert(file, "i", 8, H5T_IEEE_F64LE);
        } else {
            mem = H5Tcopy(H5T_NATIVE_DOUBLE);
            file = H5Tcopy(H5T_IEEE_F64LE);
        }
    }
    ~H5Types()
    {
        if (mem >= 0) H5Tclose(mem);
        if (file >= 0) H5Tclose(file);
    }
};

void quiet_hdf5() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

}  // namespace

extern "C" {

const char *tgio_last_error_string(void) { return g_err.c_str(); }

int tgio_write(const double *data, const tgio_layout *L, int plan, const char *out_dir, const char *folder,
               const char *file_name)
{
    if (!data || !out_dir || !folder || !file_name) return fail(TG_ERR_ARG, "null argument");
    Shape s{};
    bool fourier[3], reduced;
    int rc = plan_shape(L, plan, s, fourier, reduced);
    if (rc) return rc;
    std::vector<double> h;
    rc = fetch(data, L, s, h);
    if (rc) return rc;
    const size_t e = s.elem();
    std::vector<double> fbuf(s.file_count() * e, 0.0);
    for_each_entry(s, fourier, reduced, [&](size_t fi, long mi) {
        if (mi < 0) return;
        for (size_t k = 0; k < e; k++) fbuf[fi * e + k] = h[(size_t)mi * e + k];
    });
    quiet_hdf5();
    const std::string dir = std::string(out_dir) + "/" + folder;
    if (mkdir(dir.c_str(), S_IRWXU | S_IRWXG) != 0 && errno != EEXIST)
        return fail(TG_ERR_ARG, "cannot create folder '%s': %s", dir.c_str(), strerror(errno));
    const std::string path = dir + "/" + file_name + ".h5";
    hid_t f = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail(TG_ERR_ARG, "cannot create HDF5 file '%s'", path.c_str());
    hsize_t dims[3] = {s.file[0], s.file[1], s.file[2]};
    hid_t sp = H5Screate_simple(3, dims, nullptr);
    H5Types t(s.complex_);
    hid_t ds = H5Dcreate2(f, file_name, t.file, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t w = ds >= 0 ? H5Dwrite(ds, t.mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, fbuf.data()) : -1;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    herr_t c = H5Fclose(f);
    if (ds < 0 || w < 0 || c < 0) return fail(TG_ERR_ARG, "writing dataset '%s' to '%s' failed", file_name, path.c_str());
    return TG_OK;
}

int tgio_read(double *data, const tgio_layout *L, int plan, const char *in_dir, const char *file_name)
{
    if (!data || !in_dir || !file_name) return fail(TG_ERR_ARG, "null argument");
    Shape s{};
    bool fourier[3], reduced;
    int rc = plan_shape(L, plan, s, fourier, reduced);
    if (rc) return rc;
    quiet_hdf5();
    const std::string path = std::string(in_dir) + "/" + file_name + ".h5";
    struct stat st;
    if (stat(path.c_str(), &st) != 0) return fail(TG_ERR_ARG, "File does not exist: '%s'", path.c_str());
    hid_t f = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(TG_ERR_ARG, "cannot open HDF5 file '%s'", path.c_str());
    hid_t ds = H5Dopen2(f, file_name, H5P_DEFAULT);
    if (ds < 0) {
        H5Fclose(f);
        return fail(TG_ERR_ARG, "Dataset '%s' not found in file '%s'", file_name, path.c_str());
    }
    hid_t sp = H5Dget_space(ds);
    hsize_t dims[8] = {0};
    const int rank = H5Sget_simple_extent_ndims(sp);
    if (rank > 0 && rank <= 8) H5Sget_simple_extent_dims(sp, dims, nullptr);
    H5Sclose(sp);
    if (rank != 3 || dims[0] != s.file[0] || dims[1] != s.file[1] || dims[2] != s.file[2]) {
        H5Dclose(ds);
        H5Fclose(f);
        std::string got;
        for (int i = 0; i < rank && i < 8; i++) got += (i ? "x" : "") + std::to_string((unsigned long long)dims[i]);
        return fail(TG_ERR_SHAPE, "Dataset dimensions in input file '%s' [%s] differ from specified dimensions "
                    "[%zux%zux%zu]", file_name, got.c_str(), s.file[0], s.file[1], s.file[2]);
    }
    const size_t e = s.elem();
    std::vector<double> fbuf(s.file_count() * e);
    H5Types t(s.complex_);
    herr_t r = H5Dread(ds, t.mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, fbuf.data());
    H5Dclose(ds);
    H5Fclose(f);
    if (r < 0) return fail(TG_ERR_ARG, "reading dataset '%s' of '%s' failed (type mismatch?)", file_name, path.c_str());
    std::vector<double> h;
    if (plan == TGIO_KZ0_FULL) {
        rc = fetch(data, L, s, h);  // only the kz = 0 plane changes
        if (rc) return rc;
    } else {
        h.assign(s.mem_count() * e, 0.0);
    }
    for_each_entry(s, fourier, reduced, [&](size_t fi, long mi) {
        if (mi < 0) return;
        for (size_t k = 0; k < e; k++) h[(size_t)mi * e + k] = fbuf[fi * e + k];
    });
    return store(data, L, h);
}

int tgio_fill_vz(const double *V1, const double *V2, double *V3, const tgio_layout *L, const double kfactor[3])
{
    if (!V1 || !V2 || !V3 || !kfactor) return fail(TG_ERR_ARG, "null argument");
    if (!L->fourier_x) return fail(TG_ERR_ARG, "tgio_fill_vz: FFF layouts only");
    Shape s{};
    bool fourier[3], reduced;
    int rc = plan_shape(L, TGIO_FULL, s, fourier, reduced);
    if (rc) return rc;
    std::vector<double> a, b, c;
    if ((rc = fetch(V1, L, s, a)) || (rc = fetch(V2, L, s, b)) || (rc = fetch(V3, L, s, c))) return rc;
    const size_t nx = s.mem[0], ny = s.mem[1], nzc = s.mem[2];
    for (size_t lx = 0; lx < nx; lx++) {
        const long kx = lx <= nx / 2 ? (long)lx : (long)lx - (long)nx;
        for (size_t ly = 0; ly < ny; ly++) {
            const long ky = ly <= ny / 2 ? (long)ly : (long)ly - (long)ny;
            const double Kx = kx * kfactor[0], Ky = ky * kfactor[1];
            for (size_t lz = 1; lz < nzc; lz++) {
                const double Kz = lz * kfactor[2];
                const size_t i = 2 * ((lx * ny + ly) * nzc + lz);
                // (i Kx V1 + i Ky V2) / (-i Kz) = -(Kx V1 + Ky V2) / Kz
                c[i] = -(Kx * a[i] + Ky * b[i]) / Kz;
                c[i + 1] = -(Kx * a[i + 1] + Ky * b[i + 1]) / Kz;
            }
        }
    }
    return store(V3, L, c);
}

}  // extern "C"
