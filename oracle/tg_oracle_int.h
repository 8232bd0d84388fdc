/*
 * tg_oracle_int.h -- helpers shared between the FFF and SFF files of the
 * oracle (test infrastructure only).
 */
#ifndef TG_ORACLE_INT_H
#define TG_ORACLE_INT_H

#include <complex.h>

#include "tg_oracle.h"

typedef double complex or_cplx;

/* unnormalised transforms of a (Nx, Ny, Nz/2+1) complex array */
void or_int_fft_y(const or_grid *g, or_cplx *A, int sign);
void or_int_z_r2c(const or_grid *g, const double *real_in, or_cplx *A);
void or_int_z_c2r(const or_grid *g, const or_cplx *A, double *real_out);

#endif
