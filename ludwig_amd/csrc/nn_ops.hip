// nn_ops.hip -- the elementwise and reshaping kernels of the surrogate
// network around the MFMA convolutions (conv3d.hip): activation, 2x
// average pooling and nearest upsampling on periodic lattices, channel
// concatenation, bias gradients, field encoding/decoding, the loss and
// the Adam update.  All arrays are channels-last (position-major), with
// p = ((b*nx + x)*ny + y)*nz + z; BF16 values are raw 16-bit patterns.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace ldw {

namespace {

__device__ __forceinline__ uint16_t bf16_of(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t) (u >> 16);
}

__device__ __forceinline__ float f32_of(uint16_t h) {
  return __uint_as_float(((uint32_t) h) << 16);
}

#define GRID_STRIDE(i, n) \
  for (int64_t i = blockIdx.x*(int64_t) blockDim.x + threadIdx.x; i < (n); i += (int64_t) gridDim.x*blockDim.x)

int grid_of(int64_t n) {
  int64_t g = (n + 255)/256;
  return (int) (g < 16384 ? (g > 0 ? g : 1) : 16384);
}

// SiLU: a = z sigma(z)
__global__ void silu_fwd_kernel(const float * __restrict__ z, uint16_t * __restrict__ a, int64_t n) {
  GRID_STRIDE(i, n) {
    const float v = z[i];
    a[i] = bf16_of(v/(1.0f + expf(-v)));
  }
}

// dz = da sigma(z) (1 + z (1 - sigma(z)))
__global__ void silu_bwd_kernel(const float * __restrict__ z, const float * __restrict__ da,
				float * __restrict__ dz, uint16_t * __restrict__ dzb, int64_t n) {
  GRID_STRIDE(i, n) {
    const float v = z[i];
    const float s = 1.0f/(1.0f + expf(-v));
    const float g = da[i]*s*(1.0f + v*(1.0f - s));
    if (dz) dz[i] = g;
    dzb[i] = bf16_of(g);
  }
}

// per-block partial sums of each channel over a slice of positions
__global__ void bias_grad_kernel(const float * __restrict__ dz, float * __restrict__ part,
				 int64_t npos, int c, int64_t chunk) {
  const int64_t p0 = blockIdx.x*chunk;
  const int64_t p1 = (p0 + chunk < npos) ? p0 + chunk : npos;
  for (int ch = threadIdx.x; ch < c; ch += blockDim.x) {
    float s = 0.0f;
    for (int64_t p = p0; p < p1; p++) s += dz[p*c + ch];
    part[(int64_t) blockIdx.x*c + ch] = s;
  }
}

__global__ void sum_rows_kernel(const float * __restrict__ part, float * __restrict__ out,
				int nrow, int c, int accumulate) {
  for (int ch = blockIdx.x*blockDim.x + threadIdx.x; ch < c; ch += gridDim.x*blockDim.x) {
    float s = 0.0f;
    for (int r = 0; r < nrow; r++) s += part[(int64_t) r*c + ch];
    out[ch] = accumulate ? out[ch] + s : s;
  }
}

// coarse (nx/2, ny/2, nz/2) = mean of the 8 fine children
__global__ void pool2_fwd_kernel(const uint16_t * __restrict__ in, uint16_t * __restrict__ out,
				 int batch, int nx, int ny, int nz, int c) {
  const int mx = nx/2, my = ny/2, mz = nz/2;
  const int64_t n = (int64_t) batch*mx*my*mz*c;
  GRID_STRIDE(i, n) {
    const int ch = (int) (i % c);
    int64_t q = i / c;
    const int z = (int) (q % mz); q /= mz;
    const int y = (int) (q % my); q /= my;
    const int x = (int) (q % mx);
    const int b = (int) (q / mx);
    float s = 0.0f;
    for (int d = 0; d < 8; d++) {
      const int64_t f = (((int64_t) b*nx + 2*x + (d >> 2))*ny + 2*y + ((d >> 1) & 1))*nz + 2*z + (d & 1);
      s += f32_of(in[f*c + ch]);
    }
    out[i] = bf16_of(0.125f*s);
  }
}

// fine gradient = coarse gradient / 8 (nx, ny, nz: fine sizes)
__global__ void pool2_bwd_kernel(const float * __restrict__ dout, float * __restrict__ din,
				 int batch, int nx, int ny, int nz, int c, int accumulate) {
  const int64_t n = (int64_t) batch*nx*ny*nz*c;
  GRID_STRIDE(i, n) {
    const int ch = (int) (i % c);
    int64_t q = i / c;
    const int z = (int) (q % nz); q /= nz;
    const int y = (int) (q % ny); q /= ny;
    const int x = (int) (q % nx);
    const int b = (int) (q / nx);
    const int64_t cp = (((int64_t) b*(nx/2) + x/2)*(ny/2) + y/2)*(nz/2) + z/2;
    const float v = 0.125f*dout[cp*c + ch];
    din[i] = accumulate ? din[i] + v : v;
  }
}

// fine (2nx, 2ny, 2nz) = coarse parent (nx, ny, nz: coarse sizes)
__global__ void up2_fwd_kernel(const uint16_t * __restrict__ in, uint16_t * __restrict__ out,
			       int batch, int nx, int ny, int nz, int c) {
  const int fx = 2*nx, fy = 2*ny, fz = 2*nz;
  const int64_t n = (int64_t) batch*fx*fy*fz*c;
  GRID_STRIDE(i, n) {
    const int ch = (int) (i % c);
    int64_t q = i / c;
    const int z = (int) (q % fz); q /= fz;
    const int y = (int) (q % fy); q /= fy;
    const int x = (int) (q % fx);
    const int b = (int) (q / fx);
    const int64_t cp = (((int64_t) b*nx + x/2)*ny + y/2)*nz + z/2;
    out[i] = in[cp*c + ch];
  }
}

// coarse gradient = sum of the 8 fine children's gradients
__global__ void up2_bwd_kernel(const float * __restrict__ dout, float * __restrict__ din,
			       int batch, int nx, int ny, int nz, int c) {
  const int64_t n = (int64_t) batch*nx*ny*nz*c;
  const int fx = 2*nx, fy = 2*ny, fz = 2*nz;
  GRID_STRIDE(i, n) {
    const int ch = (int) (i % c);
    int64_t q = i / c;
    const int z = (int) (q % nz); q /= nz;
    const int y = (int) (q % ny); q /= ny;
    const int x = (int) (q % nx);
    const int b = (int) (q / nx);
    float s = 0.0f;
    for (int d = 0; d < 8; d++) {
      const int64_t f = (((int64_t) b*fx + 2*x + (d >> 2))*fy + 2*y + ((d >> 1) & 1))*fz + 2*z + (d & 1);
      s += dout[f*c + ch];
    }
    din[i] = s;
  }
}

// fine output [P_f][cu + cs] = (coarse parent's cu channels, skip's cs
// channels): nearest upsampling and concatenation in one pass, 8 channels
// (16 bytes) per thread; cu and cs are multiples of 8
__global__ void up2_concat_kernel(const uint16_t * __restrict__ coarse, int cu,
				  const uint16_t * __restrict__ skip, int cs, uint16_t * __restrict__ out,
				  int batch, int fx, int fy, int fz) {
  const int c8 = (cu + cs)/8;
  const int64_t n = (int64_t) batch*fx*fy*fz*c8;
  GRID_STRIDE(i, n) {
    const int k = (int) (i % c8);
    const int64_t p = i / c8;
    int64_t q = p;
    const int z = (int) (q % fz); q /= fz;
    const int y = (int) (q % fy); q /= fy;
    const int x = (int) (q % fx);
    const int b = (int) (q / fx);
    uint4 v;
    if (8*k < cu) {
      const int64_t cp = (((int64_t) b*(fx/2) + x/2)*(fy/2) + y/2)*(fz/2) + z/2;
      v = *reinterpret_cast<const uint4 *>(coarse + cp*cu + 8*k);
    }
    else {
      v = *reinterpret_cast<const uint4 *>(skip + p*cs + 8*k - cu);
    }
    *reinterpret_cast<uint4 *>(out + p*(cu + cs) + 8*k) = v;
  }
}

__global__ void concat_kernel(const uint16_t * __restrict__ a, int ca, const uint16_t * __restrict__ b,
			      int cb, uint16_t * __restrict__ out, int64_t npos) {
  const int c = ca + cb;
  GRID_STRIDE(i, npos*c) {
    const int64_t p = i / c;
    const int ch = (int) (i - p*c);
    out[i] = (ch < ca) ? a[p*ca + ch] : b[p*cb + ch - ca];
  }
}

__global__ void split_kernel(const float * __restrict__ d, int ca, int cb, float * __restrict__ da,
			     float * __restrict__ db, int64_t npos, int acc_a, int acc_b) {
  const int c = ca + cb;
  GRID_STRIDE(i, npos*c) {
    const int64_t p = i / c;
    const int ch = (int) (i - p*c);
    const float v = d[i];
    if (ch < ca) { if (da) { float * o = da + p*ca + ch; *o = acc_a ? *o + v : v; } }
    else { if (db) { float * o = db + p*cb + ch - ca; *o = acc_b ? *o + v : v; } }
  }
}

__global__ void to_bf16_kernel(const float * __restrict__ a, uint16_t * __restrict__ b, int64_t n) {
  GRID_STRIDE(i, n) b[i] = bf16_of(a[i]);
}

// network input of one sample: channels (phi, ux, uy, uz, p0, p1, p2) scaled,
// the rest zero.  state: 4 planes of nsite (device-layout SoA).
__global__ void encode_kernel(const float * __restrict__ state, int64_t nsite, EncodeNorm e,
			      uint16_t * __restrict__ out, int c) {
  GRID_STRIDE(i, nsite*c) {
    const int64_t s = i / c;
    const int ch = (int) (i - s*c);
    float v = 0.0f;
    if (ch == 0) v = (state[s] - e.phi_shift)*e.phi_scale;
    else if (ch < 4) v = state[ch*nsite + s]*e.u_scale;
    else if (ch < 7) v = e.prm[ch - 4];
    out[i] = bf16_of(v);
  }
}

// the same, one thread per site, for c a multiple of 8 and out 16-byte
// aligned: the site's channels leave as 16-byte vectors (the element kernel
// divides a 64-bit index per channel and stores 2 bytes at a time); the
// values are the element kernel's
__global__ void encode_site_kernel(const float * __restrict__ state, int64_t nsite, EncodeNorm e,
				   uint4 * __restrict__ out, int c8) {
  GRID_STRIDE(s, nsite) {
    const uint32_t h0 = bf16_of((state[s] - e.phi_shift)*e.phi_scale);
    const uint32_t h1 = bf16_of(state[nsite + s]*e.u_scale);
    const uint32_t h2 = bf16_of(state[2*nsite + s]*e.u_scale);
    const uint32_t h3 = bf16_of(state[3*nsite + s]*e.u_scale);
    const uint32_t h4 = bf16_of(e.prm[0]), h5 = bf16_of(e.prm[1]), h6 = bf16_of(e.prm[2]);
    const uint32_t h7 = bf16_of(0.0f);
    uint4 * o = out + s*c8;
    o[0] = make_uint4(h0 | (h1 << 16), h2 | (h3 << 16), h4 | (h5 << 16), h6 | (h7 << 16));
    for (int k = 1; k < c8; k++) o[k] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// training target: scaled change of (phi, u) over the jump, channels-last [nsite][4]
__global__ void target_kernel(const float * __restrict__ now, const float * __restrict__ next,
			      int64_t nsite, float * __restrict__ tgt, float dphi_scale, float du_scale) {
  GRID_STRIDE(i, nsite*4) {
    const int64_t s = i / 4;
    const int ch = (int) (i - s*4);
    const float d = next[ch*nsite + s] - now[ch*nsite + s];
    tgt[i] = d*(ch == 0 ? dphi_scale : du_scale);
  }
}

// prediction: state + change, back to 4 SoA planes, clamped to
// |phi| <= phi_max and |u_a| <= u_max (a guard against a wild jump)
__global__ void decode_kernel(const float * __restrict__ now, const float * __restrict__ y, int cy,
			      int64_t nsite, float * __restrict__ out, float dphi_scale, float du_scale,
			      float phi_max, float u_max) {
  GRID_STRIDE(i, nsite*4) {
    const int ch = (int) (i / nsite);
    const int64_t s = i - (int64_t) ch*nsite;
    const float v = now[i] + y[s*cy + ch]/(ch == 0 ? dphi_scale : du_scale);
    const float m = (ch == 0) ? phi_max : u_max;
    out[i] = fminf(fmaxf(v, -m), m);
  }
}

// the same, one thread per site, for cy a multiple of 4 and y 16-byte
// aligned: the site's four outputs are one 16-byte load (the element
// kernel reads them 4 bytes at a time, cy floats apart); the values are
// the element kernel's
__global__ void decode_site_kernel(const float * __restrict__ now, const float4 * __restrict__ y, int cy4,
				   int64_t nsite, float * __restrict__ out, float dphi_scale, float du_scale,
				   float phi_max, float u_max) {
  GRID_STRIDE(s, nsite) {
    const float4 yv = y[s*cy4];
    const float d[4] = {yv.x, yv.y, yv.z, yv.w};
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
      const int64_t i = (int64_t) ch*nsite + s;
      const float v = now[i] + d[ch]/(ch == 0 ? dphi_scale : du_scale);
      const float m = (ch == 0) ? phi_max : u_max;
      out[i] = fminf(fmaxf(v, -m), m);
    }
  }
}

// weighted mean-square error over the first 4 channels of y: w_c (y - t)^2
// into err (for a compensated sum), gradient 2 w_c (y - t)/(4 npos) into dy
// (BF16) and, if given, dyf (FP32); zero for the other channels
__global__ void mse_kernel(const float * __restrict__ y, int cy, const float * __restrict__ tgt,
			   int64_t npos, float * __restrict__ err, uint16_t * __restrict__ dy,
			   float * __restrict__ dyf, ChannelWeights w) {
  const float scale = 2.0f/(4.0f*(float) npos);
  GRID_STRIDE(i, npos*cy) {
    const int64_t p = i / cy;
    const int ch = (int) (i - p*cy);
    float g = 0.0f;
    if (ch < 4) {
      const float e = y[i] - tgt[p*4 + ch];
      err[p*4 + ch] = w.w[ch]*e*e;
      g = scale*w.w[ch]*e;
    }
    dy[i] = bf16_of(g);
    if (dyf) dyf[i] = g;
  }
}

// Adam with bias correction folded into lr_t = lr sqrt(1 - b2^t)/(1 - b1^t);
// the gradient is first multiplied by gscale (global-norm clipping)
__global__ void adam_kernel(float * __restrict__ w, const float * __restrict__ g,
			    float * __restrict__ m, float * __restrict__ v, int64_t n,
			    float lr_t, float b1, float b2, float eps, float wd, float gscale) {
  GRID_STRIDE(i, n) {
    const float gi = g[i]*gscale;
    const float mi = b1*m[i] + (1.0f - b1)*gi;
    const float vi = b2*v[i] + (1.0f - b2)*gi*gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr_t*(mi/(sqrtf(vi) + eps) + wd*w[i]);
  }
}

__global__ void scale_kernel(float * __restrict__ a, int64_t n, float s) {
  GRID_STRIDE(i, n) a[i] *= s;
}

__global__ void shift_kernel(float * __restrict__ a, int64_t n, float s) {
  GRID_STRIDE(i, n) a[i] += s;
}

// An element of the cubic group on a (phi, ux, uy, uz) state of an n^3
// periodic lattice (4 SoA planes): output axis a takes input axis perm[a],
// reflected about the site n-1-t where flip[a] (a map of sites to sites);
// the velocity is a vector, u'_a = (flip[a] ? -1 : 1) u_perm[a].  One
// thread per output site, gathering.
struct CubicOp { int perm[3], flip[3]; };

__global__ void cubic_kernel(const float * __restrict__ in, float * __restrict__ out, int n, CubicOp op) {
  const int64_t ns = (int64_t) n*n*n;
  GRID_STRIDE(s, ns) {
    const int o[3] = {(int) (s/((int64_t) n*n)), (int) ((s/n) % n), (int) (s % n)};
    int r[3];
#pragma unroll
    for (int a = 0; a < 3; a++) r[op.perm[a]] = op.flip[a] ? n - 1 - o[a] : o[a];
    const int64_t src = ((int64_t) r[0]*n + r[1])*n + r[2];
    out[s] = in[src];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const float v = in[(int64_t) (1 + op.perm[a])*ns + src];
      out[(int64_t) (1 + a)*ns + s] = op.flip[a] ? -v : v;
    }
  }
}

// per-block sums (SQUARE: of squares) in FP64 (fixed reduction order)
template <bool SQUARE>
__global__ void sumsq_kernel(const float * __restrict__ a, int64_t n, double * __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  GRID_STRIDE(i, n) { const double v = a[i]; s += SQUARE ? v*v : v; }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int) threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Block partial sums of plane p of `now` (part[p*nb + b]) and of `out`
// (part[(nplane + p)*nb + b]); grid (nb, nplane, 2)
__global__ void plane_sums_kernel(const float * __restrict__ now, const float * __restrict__ out,
				  int64_t ns, int nplane, double * __restrict__ part) {
  __shared__ double sh[256];
  const float * a = (blockIdx.z == 0 ? now : out) + (int64_t) blockIdx.y*ns;
  double s = 0.0;
  for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < ns;
       i += (int64_t) gridDim.x*blockDim.x) s += a[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int) threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[((int64_t) blockIdx.z*nplane + blockIdx.y)*gridDim.x + blockIdx.x] = sh[0];
}

// out[p] += (sum now[p] - sum out[p])/ns, the partial sums added in a
// fixed order by every thread; grid (.., nplane)
__global__ void plane_shift_kernel(float * __restrict__ out, int64_t ns, int nplane, int nb,
				   const double * __restrict__ part) {
  const int p = blockIdx.y;
  double sa = 0.0, sb = 0.0;
  for (int b = 0; b < nb; b++) {
    sa += part[(int64_t) p*nb + b];
    sb += part[((int64_t) nplane + p)*nb + b];
  }
  const float d = (float) ((sa - sb)/(double) ns);
  float * a = out + (int64_t) p*ns;
  for (int64_t i = (int64_t) blockIdx.x*blockDim.x + threadIdx.x; i < ns;
       i += (int64_t) gridDim.x*blockDim.x) a[i] += d;
}

__global__ void sum_parts_kernel(const double * __restrict__ part, int nb, double * __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[b];
    out[0] = s;
  }
}

}  // namespace

#define LAUNCH(k, n, ...) do { hipLaunchKernelGGL(k, dim3(grid_of(n)), dim3(256), 0, s, __VA_ARGS__); } while (0)

hipError_t nn_silu_fwd(const float * z, uint16_t * a, int64_t n, hipStream_t s) {
  LAUNCH(silu_fwd_kernel, n, z, a, n);
  return hipGetLastError();
}

hipError_t nn_silu_bwd(const float * z, const float * da, float * dz, uint16_t * dzb, int64_t n,
		       hipStream_t s) {
  LAUNCH(silu_bwd_kernel, n, z, da, dz, dzb, n);
  return hipGetLastError();
}

int64_t nn_bias_grad_workspace(int64_t npos, int c) {
  (void) npos;
  return (int64_t) 1024*c;
}

hipError_t nn_bias_grad(const float * dz, float * db, float * work, int64_t npos, int c,
			int accumulate, hipStream_t s) {
  int64_t nblk = 1024;
  if (npos < nblk) nblk = npos > 0 ? npos : 1;
  const int64_t chunk = (npos + nblk - 1)/nblk;
  nblk = (npos + chunk - 1)/chunk;
  if (nblk < 1) nblk = 1;
  hipLaunchKernelGGL(bias_grad_kernel, dim3((unsigned) nblk), dim3(64), 0, s, dz, work, npos, c, chunk);
  hipLaunchKernelGGL(sum_rows_kernel, dim3((c + 255)/256), dim3(256), 0, s, work, db, (int) nblk, c, accumulate);
  return hipGetLastError();
}

hipError_t nn_pool2_fwd(const uint16_t * in, uint16_t * out, int batch, int nx, int ny, int nz, int c,
			hipStream_t s) {
  const int64_t n = (int64_t) batch*(nx/2)*(ny/2)*(nz/2)*c;
  LAUNCH(pool2_fwd_kernel, n, in, out, batch, nx, ny, nz, c);
  return hipGetLastError();
}

hipError_t nn_pool2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c,
			int accumulate, hipStream_t s) {
  const int64_t n = (int64_t) batch*nx*ny*nz*c;
  LAUNCH(pool2_bwd_kernel, n, dout, din, batch, nx, ny, nz, c, accumulate);
  return hipGetLastError();
}

hipError_t nn_up2_fwd(const uint16_t * in, uint16_t * out, int batch, int nx, int ny, int nz, int c,
		      hipStream_t s) {
  const int64_t n = (int64_t) batch*8*nx*ny*nz*c;
  LAUNCH(up2_fwd_kernel, n, in, out, batch, nx, ny, nz, c);
  return hipGetLastError();
}

hipError_t nn_up2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c,
		      hipStream_t s) {
  const int64_t n = (int64_t) batch*nx*ny*nz*c;
  LAUNCH(up2_bwd_kernel, n, dout, din, batch, nx, ny, nz, c);
  return hipGetLastError();
}

hipError_t nn_up2_concat(const uint16_t * coarse, int cu, const uint16_t * skip, int cs, uint16_t * out,
			int batch, int fx, int fy, int fz, hipStream_t s) {
  const int64_t n = (int64_t) batch*fx*fy*fz*((cu + cs)/8);
  LAUNCH(up2_concat_kernel, n, coarse, cu, skip, cs, out, batch, fx, fy, fz);
  return hipGetLastError();
}

hipError_t nn_concat(const uint16_t * a, int ca, const uint16_t * b, int cb, uint16_t * out,
		     int64_t npos, hipStream_t s) {
  LAUNCH(concat_kernel, npos*(ca + cb), a, ca, b, cb, out, npos);
  return hipGetLastError();
}

hipError_t nn_split(const float * d, int ca, int cb, float * da, float * db, int64_t npos,
		    int acc_a, int acc_b, hipStream_t s) {
  LAUNCH(split_kernel, npos*(ca + cb), d, ca, cb, da, db, npos, acc_a, acc_b);
  return hipGetLastError();
}

hipError_t nn_to_bf16(const float * a, uint16_t * b, int64_t n, hipStream_t s) {
  LAUNCH(to_bf16_kernel, n, a, b, n);
  return hipGetLastError();
}

hipError_t nn_encode(const float * state, int64_t nsite, const EncodeNorm & e, uint16_t * out, int c,
		     hipStream_t s) {
  if (c % 8 == 0 && ((uintptr_t) out % 16) == 0) {
    LAUNCH(encode_site_kernel, nsite, state, nsite, e, reinterpret_cast<uint4 *>(out), c/8);
  }
  else {
    LAUNCH(encode_kernel, nsite*c, state, nsite, e, out, c);
  }
  return hipGetLastError();
}

hipError_t nn_target(const float * now, const float * next, int64_t nsite, float * tgt,
		     float dphi_scale, float du_scale, hipStream_t s) {
  LAUNCH(target_kernel, nsite*4, now, next, nsite, tgt, dphi_scale, du_scale);
  return hipGetLastError();
}

hipError_t nn_decode(const float * now, const float * y, int cy, int64_t nsite, float * out,
		     float dphi_scale, float du_scale, float phi_max, float u_max, hipStream_t s) {
  if (cy % 4 == 0 && ((uintptr_t) y % 16) == 0) {
    LAUNCH(decode_site_kernel, nsite, now, reinterpret_cast<const float4 *>(y), cy/4, nsite, out, dphi_scale,
	   du_scale, phi_max, u_max);
  }
  else {
    LAUNCH(decode_kernel, nsite*4, now, y, cy, nsite, out, dphi_scale, du_scale, phi_max, u_max);
  }
  return hipGetLastError();
}

hipError_t nn_mse(const float * y, int cy, const float * tgt, int64_t npos, float * err, uint16_t * dy,
		  float * dyf, const ChannelWeights & w, hipStream_t s) {
  LAUNCH(mse_kernel, npos*cy, y, cy, tgt, npos, err, dy, dyf, w);
  return hipGetLastError();
}

hipError_t nn_adam(float * w, const float * g, float * m, float * v, int64_t n, float lr_t,
		   float b1, float b2, float eps, float wd, float gscale, hipStream_t s) {
  LAUNCH(adam_kernel, n, w, g, m, v, n, lr_t, b1, b2, eps, wd, gscale);
  return hipGetLastError();
}

hipError_t nn_scale(float * a, int64_t n, float f, hipStream_t s) {
  LAUNCH(scale_kernel, n, a, n, f);
  return hipGetLastError();
}

constexpr int SUMSQ_BLOCKS = 512;

hipError_t nn_sumsq(const float * a, int64_t n, double * work, hipStream_t s) {
  hipLaunchKernelGGL(sumsq_kernel<true>, dim3(SUMSQ_BLOCKS), dim3(256), 0, s, a, n, work + 1);
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(64), 0, s, work + 1, SUMSQ_BLOCKS, work);
  return hipGetLastError();
}

hipError_t nn_sum(const float * a, int64_t n, double * work, hipStream_t s) {
  hipLaunchKernelGGL(sumsq_kernel<false>, dim3(SUMSQ_BLOCKS), dim3(256), 0, s, a, n, work + 1);
  hipLaunchKernelGGL(sum_parts_kernel, dim3(1), dim3(64), 0, s, work + 1, SUMSQ_BLOCKS, work);
  return hipGetLastError();
}

constexpr int RESTORE_BLOCKS = 64;

int nn_restore_sums_scratch(int nplane) { return 2*nplane*RESTORE_BLOCKS; }

hipError_t nn_restore_sums(const float * now, float * out, int64_t ns, int nplane, double * work,
			   hipStream_t s) {
  hipLaunchKernelGGL(plane_sums_kernel, dim3(RESTORE_BLOCKS, nplane, 2), dim3(256), 0, s,
		     now, out, ns, nplane, work);
  hipLaunchKernelGGL(plane_shift_kernel, dim3(256, nplane), dim3(256), 0, s,
		     out, ns, nplane, RESTORE_BLOCKS, work);
  return hipGetLastError();
}

hipError_t nn_cubic(const float * in, float * out, int n, const int perm[3], const int flip[3],
		    hipStream_t s) {
  CubicOp op;
  for (int a = 0; a < 3; a++) { op.perm[a] = perm[a]; op.flip[a] = flip[a] ? 1 : 0; }
  LAUNCH(cubic_kernel, (int64_t) n*n*n, in, out, n, op);
  return hipGetLastError();
}

hipError_t nn_shift(float * a, int64_t n, float v, hipStream_t s) {
  LAUNCH(shift_kernel, n, a, n, v);
  return hipGetLastError();
}

}  // namespace ldw
