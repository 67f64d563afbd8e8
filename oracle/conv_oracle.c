/*****************************************************************************
 *
 *  conv_oracle.c -- FP64 CPU convolution checker for the surrogate network
 *
 *  TEST INFRASTRUCTURE ONLY (as ldw_oracle.h): loaded by tests/ to check
 *  the GPU library's MFMA convolutions and network kernels; never linked
 *  or called by the product.
 *
 *  There is no counterpart in the reference (the learned surrogate is an
 *  addition).  The definitions are the textbook ones, written as direct
 *  sums: a 3 x 3 x 3 convolution with periodic padding on channels-last
 *  arrays x[p][c], p = ((b nx + x) ny + y) nz + z, weights w[27][cout][cin]
 *  with tap t = 9 (dx+1) + 3 (dy+1) + (dz+1), offset o_t = (dx, dy, dz):
 *
 *    y[p][co]     = bias[co] + sum_t sum_ci w[t][co][ci] x[p + o_t][ci]
 *    dx[q][ci]    = sum_t sum_co w[t][co][ci] dy[q - o_t][co]
 *    dw[t][co][ci] = sum_p dy[p][co] x[p + o_t][ci]
 *
 *  The adjoint relations (dx, dw are the gradients of sum g.y) are checked
 *  against central finite differences in tests/test_conv_oracle.py.
 *
 *****************************************************************************/

#include <stdint.h>

static inline int wrap(int i, int n) { return (i % n + n) % n; }

static inline int64_t shift(int64_t p, int t, int sgn, int nx, int ny, int nz) {
  int z = (int) (p % nz); p /= nz;
  int y = (int) (p % ny); p /= ny;
  int x = (int) (p % nx); p /= nx;
  int b = (int) p;
  int dx = t/9 - 1, dy = (t/3)%3 - 1, dz = t%3 - 1;
  x = wrap(x + sgn*dx, nx); y = wrap(y + sgn*dy, ny); z = wrap(z + sgn*dz, nz);
  return (((int64_t) b*nx + x)*ny + y)*nz + z;
}

void ora_conv3_fwd(const double * x, const double * w, const double * bias, double * y,
		   int batch, int nx, int ny, int nz, int cin, int cout) {
  const int64_t np = (int64_t) batch*nx*ny*nz;
  #pragma omp parallel for
  for (int64_t p = 0; p < np; p++) {
    for (int co = 0; co < cout; co++) {
      double s = bias ? bias[co] : 0.0;
      for (int t = 0; t < 27; t++) {
	const double * xr = x + shift(p, t, +1, nx, ny, nz)*cin;
	const double * wr = w + ((int64_t) t*cout + co)*cin;
	for (int ci = 0; ci < cin; ci++) s += wr[ci]*xr[ci];
      }
      y[p*cout + co] = s;
    }
  }
}

void ora_conv3_bwd_data(const double * dy, const double * w, double * dx,
			int batch, int nx, int ny, int nz, int cin, int cout) {
  const int64_t np = (int64_t) batch*nx*ny*nz;
  #pragma omp parallel for
  for (int64_t q = 0; q < np; q++) {
    for (int ci = 0; ci < cin; ci++) {
      double s = 0.0;
      for (int t = 0; t < 27; t++) {
	const double * dr = dy + shift(q, t, -1, nx, ny, nz)*cout;
	for (int co = 0; co < cout; co++) s += w[((int64_t) t*cout + co)*cin + ci]*dr[co];
      }
      dx[q*cin + ci] = s;
    }
  }
}

void ora_conv3_bwd_weight(const double * x, const double * dy, double * dw,
			  int batch, int nx, int ny, int nz, int cin, int cout) {
  const int64_t np = (int64_t) batch*nx*ny*nz;
  #pragma omp parallel for collapse(2)
  for (int t = 0; t < 27; t++) {
    for (int co = 0; co < cout; co++) {
      for (int ci = 0; ci < cin; ci++) {
	double s = 0.0;
	for (int64_t p = 0; p < np; p++) s += dy[p*cout + co]*x[shift(p, t, +1, nx, ny, nz)*cin + ci];
	dw[((int64_t) t*cout + co)*cin + ci] = s;
      }
    }
  }
}
