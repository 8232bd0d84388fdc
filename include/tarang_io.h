/*
 * tarang_io.h -- field dumps in the reference's HDF5 layout (C ABI).
 *
 * The reference writes each component of a spectral field to its own file,
 * <data_dir>/out/<folder>/<field>.<comp>.h5 holding one dataset named
 * <field>.<comp> (BasicIO::Write, trunk/lib/basis/basicfn/BasicIO.cc:289-312;
 * CVF::Write_complex_field, trunk/lib/fields/Cvf.cc:494-506, folder
 * "time_<t>"), and reads <data_dir>/in/<field>.<comp>.h5 back with a check of
 * the dataset's dimensions against the plan (BasicIO::Read, BasicIO.cc:222-
 * 285; CVF::Read_complex_field, Cvf.cc:528-539).  The plans are those of
 * BasicIO::Set_H5_plans (BasicIO_SP.h:46-231): full complex array, the kz = 0
 * plane, reduced arrays (low modes of each axis) and the full real array.
 *
 * Complex values are stored as an HDF5 compound of two IEEE doubles named
 * "r" and "i" (the reference's complex type comes from its h5si dependency,
 * which is not vendored; h5py reads this compound as complex128).
 *
 * This library is host code over the HDF5 C library; data pointers are
 * device memory (HIP) or host memory, as tgio_layout.on_device says.
 * Functions return TG_OK (0) or a negative TG_ERR_* code of tarang_gpu.h;
 * tgio_last_error_string() describes the last failure of the calling thread.
 */
#ifndef TARANG_IO_H
#define TARANG_IO_H

#ifdef __cplusplus
extern "C" {
#endif

/* IO plans (h5::Plan objects of BasicIO::Set_H5_plans) */
enum {
    TGIO_FULL = 0,         /* H5_full: complex (Nx, Ny, Nz/2+1) */
    TGIO_KZ0_FULL = 1,     /* H5_kz0_full: complex (Nx, Ny, 1), the kz = 0 plane */
    TGIO_REDUCED = 2,      /* H5_in_reduced / H5_out_reduced: complex (Rx, Ry, Rz/2+1) */
    TGIO_KZ0_REDUCED = 3,  /* H5_in_kz0_reduced / H5_out_kz0_reduced: complex (Rx, Ry, 1) */
    TGIO_REAL = 4          /* H5_real: double (Nx, Ny, Nz+2) */
};

typedef struct tgio_layout {
    int nx, ny, nz;       /* grid: complex arrays (nx, ny, nz/2+1), real arrays (nx, ny, nz+2) */
    int fourier_x;        /* 1: x is a Fourier axis (FFF); 0: sin/cos along x (SFF) */
    int reduced[3];       /* io.N_in_reduced / N_out_reduced (grid sizes), REDUCED plans only */
    int on_device;        /* 1: data pointers are HIP device memory; 0: host memory */
} tgio_layout;

/* BasicIO::Write: <out_dir>/<folder>/<file_name>.h5, dataset <file_name>;
 * the folder is created (mode 0770) when missing, the file is replaced. */
int tgio_write(const double *data, const tgio_layout *lay, int plan, const char *out_dir, const char *folder,
               const char *file_name);

/* BasicIO::Read: <in_dir>/<file_name>.h5, dataset <file_name>.  TG_ERR_SHAPE
 * when the dataset's dimensions differ from the plan's.  TGIO_FULL and
 * TGIO_REAL fill the whole array; TGIO_KZ0_FULL writes only the kz = 0 plane;
 * the reduced plans zero the array, then fill the low modes
 * (CVF::Read_reduced_complex_field, Cvf.cc:541-556). */
int tgio_read(double *data, const tgio_layout *lay, int plan, const char *in_dir, const char *file_name);

/* Universal::Fill_Vz with FFF_PENCIL::Last_component (universal_basic.cc:
 * 371-383, FFF_pencil_basic.cc:51-73): V3 = (i Kx V1 + i Ky V2)/(-i Kz) for
 * kz > 0, from incompressibility, after V3's kz = 0 plane was read alone
 * (input_vx_vy_switch).  FFF layouts only. */
int tgio_fill_vz(const double *V1, const double *V2, double *V3, const tgio_layout *lay, const double kfactor[3]);

const char *tgio_last_error_string(void);

#ifdef __cplusplus
}
#endif
#endif
