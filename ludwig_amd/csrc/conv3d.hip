// conv3d.hip -- 3 x 3 x 3 convolutions on periodic lattices as implicit
// GEMMs on the MFMA units (BF16 operands, FP32 accumulation), for the
// learned surrogate's network.
//
// Layout: activations are channels-last, x[p][c] with p = ((b*nx + x)*ny + y)*nz + z
// over the batch, so the 27 taps of a position are 27 contiguous channel
// rows.  Padding is periodic (the lattice is periodic), so no halo exists
// and the adjoint of a convolution is again a convolution.  Channel
// counts are multiples of 16.
//
// Forward, y = b + sum_t W_t x(p + o_t):  GEMM with M = positions,
//   N = Cout, K = 27 Cin.  A wave owns 32 positions x 16*NT16 output
//   channels: 2 x NT16 tiles of v_mfma_f32_16x16x16_bf16.  The A fragment
//   of a lane is 4 consecutive channels of one (shifted) position and the
//   B fragment 4 consecutive input channels of one output channel, so both
//   are single 8-byte loads with weights stored as W[t][cout][cin].
// Backward data, dx = sum_t W_{26-t}^T dy(p + o_t): the same kernel with the
//   weights flipped and transposed (conv3_weights_prep).
// Backward weight, dW_t = sum_p dy(p) x(p + o_t)^T:  GEMM with M = Cout,
//   N = Cin, K = positions; 64-position slices of dy and of the shifted x
//   are staged through LDS (contiguous rows) and read as fragments;
//   position chunks give partial sums that a second kernel adds in a fixed
//   order (deterministic).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "kernels.h"

namespace ldw {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int wrapi(int i, int n) {
  return (i < 0) ? i + n : (i >= n ? i - n : i);
}

struct Pos { int b, x, y, z; };

__device__ __forceinline__ uint16_t silu_bf16(float v) {
  uint32_t u = __float_as_uint(v/(1.0f + expf(-v)));
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t) (u >> 16);
}

// epilogue: y = acc + bias (+ y) if y is given; a = bf16(SiLU(that)) if a is given
__device__ __forceinline__ void store_out(float * __restrict__ y, uint16_t * __restrict__ a, int64_t idx,
					  float v, int accumulate) {
  if (y) {
    if (accumulate) v += y[idx];
    y[idx] = v;
  }
  if (a) a[idx] = silu_bf16(v);
}

__device__ __forceinline__ Pos pos_of(int64_t p, int nx, int ny, int nz) {
  Pos q;
  q.z = (int) (p % nz); p /= nz;
  q.y = (int) (p % ny); p /= ny;
  q.x = (int) (p % nx); p /= nx;
  q.b = (int) p;
  return q;
}

__device__ __forceinline__ int64_t shifted(const Pos & q, int t, int nx, int ny, int nz) {
  const int dx = t/9 - 1, dy = (t/3)%3 - 1, dz = t%3 - 1;
  return (((int64_t) q.b*nx + wrapi(q.x + dx, nx))*ny + wrapi(q.y + dy, ny))*nz + wrapi(q.z + dz, nz);
}

template <int NT16>
__global__ void __launch_bounds__(256)
conv3_fwd_kernel(const uint16_t * __restrict__ x, const uint16_t * __restrict__ w,
		 const float * __restrict__ bias, float * __restrict__ y, uint16_t * __restrict__ act,
		 int64_t npos, int nx, int ny, int nz, int cin, int cout, int accumulate) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int64_t m0 = (int64_t) blockIdx.x*(blockDim.x/2) + wave*32;   // 32 positions per wave
  const int n0 = blockIdx.y*16*NT16;

  Pos q[2];
  bool ok[2];
#pragma unroll
  for (int rb = 0; rb < 2; rb++) {
    const int64_t p = m0 + 16*rb + r;
    ok[rb] = (p < npos);
    q[rb] = pos_of(ok[rb] ? p : 0, nx, ny, nz);
  }
  f4 acc[2][NT16];
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int nb = 0; nb < NT16; nb++) acc[rb][nb] = f4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int t = 0; t < 27; t++) {
    const int64_t s0 = shifted(q[0], t, nx, ny, nz)*cin + 4*kq;
    const int64_t s1 = shifted(q[1], t, nx, ny, nz)*cin + 4*kq;
    const uint16_t * wt = w + ((int64_t) t*cout + n0 + r)*cin + 4*kq;
    for (int c0 = 0; c0 < cin; c0 += 16) {
      s4 a0 = ok[0] ? *reinterpret_cast<const s4 *>(x + s0 + c0) : s4{0, 0, 0, 0};
      s4 a1 = ok[1] ? *reinterpret_cast<const s4 *>(x + s1 + c0) : s4{0, 0, 0, 0};
#pragma unroll
      for (int nb = 0; nb < NT16; nb++) {
	const s4 b = *reinterpret_cast<const s4 *>(wt + (int64_t) 16*nb*cin + c0);
	acc[0][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b, acc[0][nb], 0, 0, 0);
	acc[1][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b, acc[1][nb], 0, 0, 0);
      }
    }
  }
  // D layout: lane holds rows 4*kq + i of column r
#pragma unroll
  for (int rb = 0; rb < 2; rb++)
#pragma unroll
    for (int nb = 0; nb < NT16; nb++) {
      const int ch = n0 + 16*nb + r;
      const float bv = bias ? bias[ch] : 0.0f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
	const int64_t p = m0 + 16*rb + 4*kq + i;
	if (p < npos) store_out(y, act, p*cout + ch, acc[rb][nb][i] + bv, accumulate);
      }
    }
}

// ---------------------------------------------------------------------------
// LDS-tiled forward pass for lattices at least one tile in size.
//
// A block owns a 4 x 4 x 16 (x, y, z) tile of one sample and 16*NT output
// channels; wave w owns the 4 z rows (x0 + w, y0 + j, z0..z0+15), j < 4,
// i.e. 4 row blocks of 16 positions.  Input channels go in chunks of 32:
// the chunk's 6 x 6 x 18 halo tile is staged in LDS once (periodic wrap on
// load), so the 27 taps read their A fragments from LDS instead of
// re-fetching them; the weights of 3 taps at a time are staged in LDS too,
// shared by the 4 waves, and the next 3 taps' weights are fetched into
// registers while the current ones are used.  v_mfma_f32_16x16x32_bf16:
// a lane's A fragment is 8 consecutive channels of one position (16 bytes),
// its B fragment 8 consecutive input channels of one output channel.

typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

constexpr int LT_X = 4, LT_Y = 4, LT_Z = 16;
constexpr int LH_X = LT_X + 2, LH_Y = LT_Y + 2, LH_Z = LT_Z + 2;
constexpr int LH_N = LH_X*LH_Y*LH_Z;          // 648 halo positions
constexpr int LK = 32;                        // input channels per chunk
constexpr int LNT_MAX = 6;                    // output channel blocks per block

// LDS rows are 4 x 16 bytes (32 channels).  The 16 lanes of a fragment
// read 16 consecutive rows at one 16-byte part, 64 bytes apart, which puts
// them on the same banks; the part index is XOR-swizzled with bits of the
// row index so that they spread over all banks.  swz_shift selects the bits.
__device__ __forceinline__ int lds_part(int row, int part, int swz_shift) {
  return part ^ ((row >> swz_shift) & 3);
}

template <int NT>
__global__ void __launch_bounds__(256)
conv3_fwd_lds_kernel(const uint16_t * __restrict__ x, const uint16_t * __restrict__ w,
		     const float * __restrict__ bias, float * __restrict__ y, uint16_t * __restrict__ act,
		     int nx, int ny, int nz, int cin, int cout, int accumulate,
		     int ntx, int nty, int ntz, int swz) {
  __shared__ uint4 sa[LH_N*(LK/8)];           // halo tile: [pos][4 x 8 channels]   41.5 KB
  __shared__ uint4 sb[3][16*NT][LK/8];        // 3 taps x 16 NT outputs x 32 inputs
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  int t = blockIdx.x;
  const int tz = t % ntz; t /= ntz;
  const int ty = t % nty; t /= nty;
  const int tx = t % ntx;
  const int b = t / ntx;
  const int x0 = tx*LT_X, y0 = ty*LT_Y, z0 = tz*LT_Z;
  const int n0 = blockIdx.y*16*NT;
  constexpr int NB = 3*16*NT*(LK/8);           // uint4 per weight stage
  constexpr int NBT = (NB + 255)/256;          // per thread

  f4 acc[4][NT];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int n = 0; n < NT; n++) acc[j][n] = f4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int c0 = 0; c0 < cin; c0 += LK) {
    const int kc = (cin - c0 < LK) ? cin - c0 : LK;          // 16 or 32
    for (int i = threadIdx.x; i < LH_N*4; i += 256) {
      const int h = i >> 2, part = i & 3;
      const int hz = h % LH_Z, hy = (h / LH_Z) % LH_Y, hx = h / (LH_Z*LH_Y);
      uint4 v = {0u, 0u, 0u, 0u};
      if (8*part < kc) {
	const int gx = wrapi(x0 + hx - 1, nx), gy = wrapi(y0 + hy - 1, ny), gz = wrapi(z0 + hz - 1, nz);
	v = *reinterpret_cast<const uint4 *>(x + ((((int64_t) b*nx + gx)*ny + gy)*nz + gz)*cin + c0 + 8*part);
      }
      sa[4*h + lds_part(h, part, swz)] = v;
    }
    // weights of taps 3s .. 3s+2: [tap][n][part]
    uint4 pre[NBT];
    auto fetch = [&](int st) {
#pragma unroll
      for (int k = 0; k < NBT; k++) {
	const int i = threadIdx.x + 256*k;
	uint4 v = {0u, 0u, 0u, 0u};
	if (i < NB) {
	  const int part = i & 3, n = (i >> 2) % (16*NT), tp = (i >> 2)/(16*NT);
	  if (8*part < kc)
	    v = *reinterpret_cast<const uint4 *>(w + ((int64_t) (3*st + tp)*cout + n0 + n)*cin + c0 + 8*part);
	}
	pre[k] = v;
      }
    };
    fetch(0);
    for (int st = 0; st < 9; st++) {
#pragma unroll
      for (int k = 0; k < NBT; k++) {
	const int i = threadIdx.x + 256*k;
	if (i < NB) (&sb[0][0][0])[4*(i >> 2) + lds_part(i >> 2, i & 3, swz)] = pre[k];
      }
      __syncthreads();                        // halo (first stage) and weights in place
      if (st + 1 < 9) fetch(st + 1);
      const int dx = st/3, dy = st%3;
#pragma unroll
      for (int dz = 0; dz < 3; dz++) {
	s8 bv[NT];
#pragma unroll
	for (int n = 0; n < NT; n++) {
	  const int row = dz*16*NT + 16*n + r;
	  bv[n] = *reinterpret_cast<const s8 *>(&sb[dz][16*n + r][lds_part(row, q, swz)]);
	}
#pragma unroll
	for (int j = 0; j < 4; j++) {
	  const int h = ((wave + dx)*LH_Y + (j + dy))*LH_Z + (r + dz);
	  const s8 a = *reinterpret_cast<const s8 *>(&sa[h*4 + lds_part(h, q, swz)]);
#pragma unroll
	  for (int n = 0; n < NT; n++)
	    acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
								 __builtin_bit_cast(b8, bv[n]), acc[j][n], 0, 0, 0);
	}
      }
      __syncthreads();                        // everyone done with this stage's weights
    }
  }
  const int gx = x0 + wave;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int gy = y0 + j;
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const int ch = n0 + 16*n + r;
      const float bv = bias ? bias[ch] : 0.0f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
	const int gz = z0 + 4*q + i;
	if (gx < nx && gy < ny && gz < nz) {
	  const int64_t p = (((int64_t) b*nx + gx)*ny + gy)*nz + gz;
	  store_out(y, act, p*cout + ch, acc[j][n][i] + bv, accumulate);
	}
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Pipelined variant of the LDS-tiled forward pass (the default up to 64
// output channels per block).  The tiling and the MFMA fragments are those of
// conv3_fwd_lds_kernel; what changes is where the memory latency goes.
// rocprof counters of the tiled kernel (MFMA busy 25 % of cycles, waves
// waiting half their cycles, LDS active a quarter of the time, no bank
// conflicts) show the halo fill of each input-channel chunk as the binding
// cost: every block stops for it with nothing to compute.  Here
//   - the halo of the NEXT input-channel chunk is loaded into registers
//     while the current chunk is computed, and written to LDS when the
//     current chunk is done;
//   - the weights run one stage ahead across chunk boundaries;
//   - a chunk of 16 channels (cin = 16 mod 32) runs on the K = 16 MFMA
//     (v_mfma_f32_16x16x16_bf16) instead of a K = 32 one half of zeros.
// The products and the FP32 accumulation bound are those of the tiled
// kernel; a 16-channel chunk's MFMA sums 16 products where the tiled
// kernel's sums them with 16 zero ones, so the last bits may differ.

constexpr int HPT = (LH_N*4 + 255)/256;       // halo uint4 per thread

// One stage (taps dx, dy, dz = 0..2) of the pipelined kernel on the LDS
// tiles: K = 32 MFMAs on 8-channel fragments.  For a 16-channel chunk taps
// dz = 0 and 1 share one K = 32 MFMA and tap dz = 2 takes a K = 16 MFMA,
// lane (r, q) holding channels 4q .. 4q+3 (the (q & 1) half of 16-byte
// part q >> 1): on gfx950 a K = 16 MFMA occupies the unit as long as a
// K = 32 one, so the pair saves a third of the chunk's MFMA time.
template <int NT, bool K16>
__device__ __forceinline__ void conv_stage(const uint4 * sa, const uint4 (*sb)[16*NT][LK/8], f4 (&acc)[4][NT],
					   int wave, int r, int q, int dx, int dy, int swz) {
#pragma unroll
  for (int dz = 0; dz < 3; dz++) {
    if constexpr (!K16) {
      s8 bv[NT];
#pragma unroll
      for (int n = 0; n < NT; n++) {
	const int row = dz*16*NT + 16*n + r;
	bv[n] = *reinterpret_cast<const s8 *>(&sb[dz][16*n + r][lds_part(row, q, swz)]);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
	const int h = ((wave + dx)*LH_Y + (j + dy))*LH_Z + (r + dz);
	const s8 a = *reinterpret_cast<const s8 *>(&sa[h*4 + lds_part(h, q, swz)]);
#pragma unroll
	for (int n = 0; n < NT; n++)
	  acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
							       __builtin_bit_cast(b8, bv[n]), acc[j][n], 0, 0, 0);
      }
    }
    else if (dz == 0) {
      // taps dz = 0 and 1 in one K = 32 MFMA: lanes q = 0, 1 hold tap 0's
      // 16 channels, lanes q = 2, 3 tap 1's (8 channels each, part q & 1)
      const int dq = q >> 1;
      s8 bv[NT];
#pragma unroll
      for (int n = 0; n < NT; n++) {
	const int row = dq*16*NT + 16*n + r;
	bv[n] = *reinterpret_cast<const s8 *>(&sb[dq][16*n + r][lds_part(row, q & 1, swz)]);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
	const int h = ((wave + dx)*LH_Y + (j + dy))*LH_Z + (r + dq);
	const s8 a = *reinterpret_cast<const s8 *>(&sa[h*4 + lds_part(h, q & 1, swz)]);
#pragma unroll
	for (int n = 0; n < NT; n++)
	  acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
							       __builtin_bit_cast(b8, bv[n]), acc[j][n], 0, 0, 0);
      }
    }
    else if (dz == 2) {
      s4 bv[NT];
#pragma unroll
      for (int n = 0; n < NT; n++) {
	const int row = dz*16*NT + 16*n + r;
	bv[n] = *reinterpret_cast<const s4 *>(reinterpret_cast<const char *>(&sb[dz][16*n + r][lds_part(row, q >> 1, swz)])
					      + 8*(q & 1));
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
	const int h = ((wave + dx)*LH_Y + (j + dy))*LH_Z + (r + dz);
	const s4 a = *reinterpret_cast<const s4 *>(reinterpret_cast<const char *>(&sa[h*4 + lds_part(h, q >> 1, swz)])
						   + 8*(q & 1));
#pragma unroll
	for (int n = 0; n < NT; n++)
	  acc[j][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv[n], acc[j][n], 0, 0, 0);
      }
    }
  }
}

// With 16 or 32 output channels per block the kernel is held to 168
// registers, i.e. 3 blocks (12 waves) per CU as the LDS allows; 48 and 64
// channels do not fit 168 without spilling and run at 2 blocks per CU;
// wider blocks take conv3_fwd_lds_kernel.
template <int NT>
__global__ void __launch_bounds__(256, (NT <= 2 ? 3 : 2))
conv3_fwd_pipe_kernel(const uint16_t * __restrict__ x, const uint16_t * __restrict__ w,
		      const float * __restrict__ bias, float * __restrict__ y, uint16_t * __restrict__ act,
		      int nx, int ny, int nz, int cin, int cout, int accumulate,
		      int ntx, int nty, int ntz, int swz) {
  __shared__ uint4 sa[LH_N*(LK/8)];
  __shared__ uint4 sb[3][16*NT][LK/8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  int t = blockIdx.x;
  const int tz = t % ntz; t /= ntz;
  const int ty = t % nty; t /= nty;
  const int tx = t % ntx;
  const int b = t / ntx;
  const int x0 = tx*LT_X, y0 = ty*LT_Y, z0 = tz*LT_Z;
  const int n0 = blockIdx.y*16*NT;
  constexpr int NB = 3*16*NT*(LK/8);
  constexpr int NBT = (NB + 255)/256;
  const uint16_t * xb = x + (int64_t) b*nx*ny*nz*cin;

  // halo of chunk c0 into registers
  uint4 hpre[HPT];
  auto fetch_halo = [&](int c0) {
    const int kc = (cin - c0 < LK) ? cin - c0 : LK;
#pragma unroll
    for (int k = 0; k < HPT; k++) {
      const int i = threadIdx.x + 256*k;
      const int h = i >> 2, part = i & 3;
      uint4 v = {0u, 0u, 0u, 0u};
      if (i < LH_N*4 && 8*part < kc) {
	const int hz = h % LH_Z, hy = (h / LH_Z) % LH_Y, hx = h / (LH_Z*LH_Y);
	const int gx = wrapi(x0 + hx - 1, nx), gy = wrapi(y0 + hy - 1, ny), gz = wrapi(z0 + hz - 1, nz);
	v = *reinterpret_cast<const uint4 *>(xb + (((int64_t) gx*ny + gy)*nz + gz)*cin + c0 + 8*part);
      }
      hpre[k] = v;
    }
  };
  auto store_halo = [&]() {
#pragma unroll
    for (int k = 0; k < HPT; k++) {
      const int i = threadIdx.x + 256*k;
      if (i < LH_N*4) sa[4*(i >> 2) + lds_part(i >> 2, i & 3, swz)] = hpre[k];
    }
  };
  // weights of taps 3st .. 3st+2 of chunk c0: [tap][n][part]
  uint4 wpre[NBT];
  auto fetch_w = [&](int c0, int st) {
    const int kc = (cin - c0 < LK) ? cin - c0 : LK;
#pragma unroll
    for (int k = 0; k < NBT; k++) {
      const int i = threadIdx.x + 256*k;
      uint4 v = {0u, 0u, 0u, 0u};
      if (i < NB) {
	const int part = i & 3, n = (i >> 2) % (16*NT), tp = (i >> 2)/(16*NT);
	if (8*part < kc)
	  v = *reinterpret_cast<const uint4 *>(w + ((int64_t) (3*st + tp)*cout + n0 + n)*cin + c0 + 8*part);
      }
      wpre[k] = v;
    }
  };
  auto store_w = [&]() {
#pragma unroll
    for (int k = 0; k < NBT; k++) {
      const int i = threadIdx.x + 256*k;
      if (i < NB) (&sb[0][0][0])[4*(i >> 2) + lds_part(i >> 2, i & 3, swz)] = wpre[k];
    }
  };

  f4 acc[4][NT];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int n = 0; n < NT; n++) acc[j][n] = f4{0.0f, 0.0f, 0.0f, 0.0f};

  // full 32-channel chunks; the halo of the next chunk (full or not) is
  // loaded while this one is computed, the weights one stage ahead
  const int cfull = cin - cin % LK;
  fetch_halo(0);
  fetch_w(0, 0);
#pragma unroll 1
  for (int c0 = 0; c0 < cfull; c0 += LK) {
    // every wave is done with the previous chunk's halo (barrier at the
    // end of its last stage)
    store_halo();
    if (c0 + LK < cin) fetch_halo(c0 + LK);
#pragma unroll 1
    for (int st = 0; st < 9; st++) {
      store_w();
      __syncthreads();                        // halo (first stage) and weights in place
      if (st + 1 < 9) fetch_w(c0, st + 1);
      else if (c0 + LK < cin) fetch_w(c0 + LK, 0);
      conv_stage<NT, false>(sa, sb, acc, wave, r, q, st/3, st%3, swz);
      __syncthreads();                        // everyone done with this stage's weights
    }
  }
  // a last chunk of 16 channels
  if (cfull < cin) {
    store_halo();
#pragma unroll 1
    for (int st = 0; st < 9; st++) {
      store_w();
      __syncthreads();
      if (st + 1 < 9) fetch_w(cfull, st + 1);
      conv_stage<NT, true>(sa, sb, acc, wave, r, q, st/3, st%3, swz);
      __syncthreads();
    }
  }
  const int gx = x0 + wave;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int gy = y0 + j;
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const int ch = n0 + 16*n + r;
      const float bv = bias ? bias[ch] : 0.0f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
	const int gz = z0 + 4*q + i;
	if (gx < nx && gy < ny && gz < nz) {
	  const int64_t p = (((int64_t) b*nx + gx)*ny + gy)*nz + gz;
	  store_out(y, act, p*cout + ch, acc[j][n][i] + bv, accumulate);
	}
      }
    }
  }
}

constexpr int WG_P = 64;          // positions per LDS slice
constexpr int WG_CMAX = 160;      // channel capacity of the LDS rows (host checks)

__global__ void __launch_bounds__(256)
conv3_wgrad_kernel(const uint16_t * __restrict__ x, const uint16_t * __restrict__ dy,
		   float * __restrict__ part, int64_t npos, int64_t chunk,
		   int nx, int ny, int nz, int cin, int cout) {
  __shared__ uint16_t sy[WG_P][WG_CMAX + 4];
  __shared__ uint16_t sx[WG_P][WG_CMAX + 4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 15, kq = lane >> 4;
  const int t = blockIdx.y;
  const int64_t pbeg = (int64_t) blockIdx.x*chunk;
  const int64_t pend = (pbeg + chunk < npos) ? pbeg + chunk : npos;
  const int nco = cout/16, nci = cin/16;
  const int npair = nco*nci;
  // this block's pairs: [16*blockIdx.z, 16*blockIdx.z + 16); wave: 4 of them
  int pco[4], pci[4];
  bool has[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int pr = 16*blockIdx.z + 4*wave + j;
    has[j] = (pr < npair);
    pco[j] = has[j] ? pr / nci : 0;
    pci[j] = has[j] ? pr % nci : 0;
  }
  f4 acc[4];
#pragma unroll
  for (int j = 0; j < 4; j++) acc[j] = f4{0.0f, 0.0f, 0.0f, 0.0f};

  for (int64_t p0 = pbeg; p0 < pend; p0 += WG_P) {
    // stage 64 rows of dy and of x shifted by tap t (4 channels per load)
    const int ly = cout/4, lx = cin/4;
    for (int i = threadIdx.x; i < WG_P*ly; i += 256) {
      const int row = i / ly, c = 4*(i - row*ly);
      const int64_t p = p0 + row;
      uint64_t v = 0;
      if (p < pend) v = *reinterpret_cast<const uint64_t *>(dy + p*cout + c);
      *reinterpret_cast<uint64_t *>(&sy[row][c]) = v;
    }
    for (int i = threadIdx.x; i < WG_P*lx; i += 256) {
      const int row = i / lx, c = 4*(i - row*lx);
      const int64_t p = p0 + row;
      uint64_t v = 0;
      if (p < pend) {
	const Pos q = pos_of(p, nx, ny, nz);
	v = *reinterpret_cast<const uint64_t *>(x + shifted(q, t, nx, ny, nz)*cin + c);
      }
      *reinterpret_cast<uint64_t *>(&sx[row][c]) = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < WG_P/16; ks++) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
	if (!has[j]) continue;
	s4 a, b;
#pragma unroll
	for (int i = 0; i < 4; i++) {
	  a[i] = (short) sy[16*ks + 4*kq + i][16*pco[j] + r];
	  b[i] = (short) sx[16*ks + 4*kq + i][16*pci[j] + r];
	}
	acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // partial[chunk][t][co][ci]
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (!has[j]) continue;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = 16*pco[j] + 4*kq + i, ci = 16*pci[j] + r;
      part[(((int64_t) blockIdx.x*27 + t)*cout + co)*cin + ci] = acc[j][i];
    }
  }
}

__global__ void wgrad_reduce_kernel(const float * __restrict__ part, float * __restrict__ dw,
				    int nchunk, int64_t n, int accumulate) {
  for (int64_t i = blockIdx.x*(int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x*blockDim.x) {
    float s = 0.0f;
    for (int c = 0; c < nchunk; c++) s += part[(int64_t) c*n + i];
    dw[i] = accumulate ? dw[i] + s : s;
  }
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t) ((u >> 16) | 0x40u);   // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                              // nearest even
  return (uint16_t) (u >> 16);
}

// wf[t][co][ci] = bf16(w[t][co][ci]); wd[t][ci][co] = bf16(w[26 - t][co][ci])
__global__ void weights_prep_kernel(const float * __restrict__ w, uint16_t * __restrict__ wf,
				    uint16_t * __restrict__ wd, int cin, int cout) {
  const int64_t n = (int64_t) 27*cout*cin;
  for (int64_t i = blockIdx.x*(int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x*blockDim.x) {
    const int ci = (int) (i % cin);
    const int co = (int) ((i / cin) % cout);
    const int t = (int) (i / ((int64_t) cin*cout));
    const uint16_t v = to_bf16(w[i]);
    if (wf) wf[i] = v;
    if (wd) wd[((int64_t) (26 - t)*cin + ci)*cout + co] = v;
  }
}

__global__ void to_bf16_kernel(const float * __restrict__ a, uint16_t * __restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x*(int64_t) blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x*blockDim.x)
    b[i] = to_bf16(a[i]);
}

// LDS swizzle shift of the tiled kernel (LDW_CONV_SWZ; 16 = no swizzle)
int conv_swz() {
  static int v = -1;
  if (v < 0) {
    const char * e = getenv("LDW_CONV_SWZ");
    v = e ? atoi(e) : 1;
  }
  return v;
}

// LDW_CONV_LDS=0 selects the fragment-direct kernel everywhere, =1 the
// tiled kernel without the pipeline (A/B timing); by default the
// pipelined kernel (up to 64 output channels per block)
int conv_lds_enabled() {
  static int on = -1;
  if (on < 0) {
    const char * e = getenv("LDW_CONV_LDS");
    on = e ? atoi(e) : 2;
  }
  return on;
}

int grid_for(int64_t n) {
  int64_t g = (n + 255)/256;
  return (int) (g < 8192 ? (g > 0 ? g : 1) : 8192);
}

}  // namespace

hipError_t launch_conv3_fwd(const uint16_t * x, const uint16_t * w, const float * bias, float * y,
			    uint16_t * act, int batch, int nx, int ny, int nz, int cin, int cout,
			    int accumulate, hipStream_t s) {
  const int64_t npos = (int64_t) batch*nx*ny*nz;
  const int nb16 = cout/16;
  if (nz >= LT_Z && ny >= LT_Y && nx >= LT_X && conv_lds_enabled()) {
    int nt = 1;
    for (int c = LNT_MAX; c >= 1; c--) if (nb16 % c == 0) { nt = c; break; }
    const int ntx = (nx + LT_X - 1)/LT_X, nty = (ny + LT_Y - 1)/LT_Y, ntz = (nz + LT_Z - 1)/LT_Z;
    dim3 grid((unsigned) ((int64_t) batch*ntx*nty*ntz), nb16/nt), block(256);
    if (conv_lds_enabled() == 2 && nt <= 4) {
#define PIPE_(NT) hipLaunchKernelGGL(conv3_fwd_pipe_kernel<NT>, grid, block, 0, s, x, w, bias, y, act, nx, ny, nz, \
				     cin, cout, accumulate, ntx, nty, ntz, conv_swz())
      switch (nt) {
      case 4: PIPE_(4); break;
      case 3: PIPE_(3); break;
      case 2: PIPE_(2); break;
      default: PIPE_(1); break;
      }
#undef PIPE_
      return hipGetLastError();
    }
#define LDS_(NT) hipLaunchKernelGGL(conv3_fwd_lds_kernel<NT>, grid, block, 0, s, x, w, bias, y, act, nx, ny, nz, \
				    cin, cout, accumulate, ntx, nty, ntz, conv_swz())
    switch (nt) {
    case 6: LDS_(6); break;
    case 5: LDS_(5); break;
    case 4: LDS_(4); break;
    case 3: LDS_(3); break;
    case 2: LDS_(2); break;
    default: LDS_(1); break;
    }
#undef LDS_
    return hipGetLastError();
  }
  // small lattices (the network's coarsest level): fewer output blocks
  // per wave and fewer waves per block, so that the grid still spreads
  // over the CUs (at 8 x 8^3 positions, 64 -> 96 channels: 768 one-wave
  // blocks instead of 64 four-wave ones)
  int nt = (nb16 % 4 == 0) ? 4 : (nb16 % 3 == 0) ? 3 : (nb16 % 2 == 0) ? 2 : 1;
  int wpb = 4;
  const int64_t want = 1024;
  while (wpb > 1 && ((npos + 32*wpb - 1)/(32*wpb))*(nb16/nt) < want) wpb /= 2;
  while (nt > 1 && ((npos + 32*wpb - 1)/(32*wpb))*(nb16/nt) < want) {
    int c = nt - 1;
    while (c > 1 && nb16 % c) c--;
    nt = c;
  }
  dim3 grid((unsigned) ((npos + 32*wpb - 1)/(32*wpb)), nb16/nt), block(64*wpb);
  switch (nt) {
  case 4: hipLaunchKernelGGL(conv3_fwd_kernel<4>, grid, block, 0, s, x, w, bias, y, act, npos, nx, ny, nz, cin, cout, accumulate); break;
  case 3: hipLaunchKernelGGL(conv3_fwd_kernel<3>, grid, block, 0, s, x, w, bias, y, act, npos, nx, ny, nz, cin, cout, accumulate); break;
  case 2: hipLaunchKernelGGL(conv3_fwd_kernel<2>, grid, block, 0, s, x, w, bias, y, act, npos, nx, ny, nz, cin, cout, accumulate); break;
  default: hipLaunchKernelGGL(conv3_fwd_kernel<1>, grid, block, 0, s, x, w, bias, y, act, npos, nx, ny, nz, cin, cout, accumulate); break;
  }
  return hipGetLastError();
}

int64_t conv3_wgrad_workspace(int batch, int nx, int ny, int nz, int cin, int cout, int * nchunk,
			      int64_t * chunk) {
  const int64_t npos = (int64_t) batch*nx*ny*nz;
  const int groups = (cout/16*(cin/16) + 15)/16;
  // about 2048 blocks in all; chunks are whole 64-position slices
  int64_t nc = 2048/(27*groups);
  if (nc < 1) nc = 1;
  int64_t ch = (npos + nc - 1)/nc;
  ch = ((ch + WG_P - 1)/WG_P)*WG_P;
  nc = (npos + ch - 1)/ch;
  *nchunk = (int) nc;
  *chunk = ch;
  return nc*27*(int64_t) cout*cin;
}

hipError_t launch_conv3_wgrad(const uint16_t * x, const uint16_t * dy, float * dw, float * work,
			      int batch, int nx, int ny, int nz, int cin, int cout, int accumulate,
			      hipStream_t s) {
  int nchunk;
  int64_t chunk;
  conv3_wgrad_workspace(batch, nx, ny, nz, cin, cout, &nchunk, &chunk);
  const int64_t npos = (int64_t) batch*nx*ny*nz;
  const int groups = (cout/16*(cin/16) + 15)/16;
  hipLaunchKernelGGL(conv3_wgrad_kernel, dim3(nchunk, 27, groups), dim3(256), 0, s,
		     x, dy, work, npos, chunk, nx, ny, nz, cin, cout);
  const int64_t n = (int64_t) 27*cout*cin;
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(grid_for(n)), dim3(256), 0, s, work, dw, nchunk, n, accumulate);
  return hipGetLastError();
}

hipError_t launch_conv3_weights_prep(const float * w, uint16_t * wf, uint16_t * wd, int cin, int cout,
				     hipStream_t s) {
  const int64_t n = (int64_t) 27*cout*cin;
  hipLaunchKernelGGL(weights_prep_kernel, dim3(grid_for(n)), dim3(256), 0, s, w, wf, wd, cin, cout);
  return hipGetLastError();
}

hipError_t launch_to_bf16(const float * a, uint16_t * b, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, s, a, b, n);
  return hipGetLastError();
}

}  // namespace ldw
