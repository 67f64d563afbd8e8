// binary_step.hip -- the whole binary-fluid time step in one launch.
//
// The reference performs, per step (ludwig.c:528-860): the thermodynamic
// force from phi_t, the Cahn-Hilliard update phi_t -> phi_t+1 with the
// velocity of the previous collision, the collision of f_t with that force
// (giving rho_t, u_t), and the propagation.  Here one block owns a tile of
// (y, z) columns (8 x 64 by default) and marches along x through a chunk of planes
// (step_tile.h): for each site it forms F and the new phi from the LDS
// tiles, then collides the site's 19 pulled distributions with that F.  The
// force never goes to memory.
//
// Traffic per site (FP64): f 19 read + 19 write, phi read + write, u of the
// previous step read, u and rho written: 376 B (against 448 B for three
// separate passes).  Buffers: phi and u are double-buffered so that no block
// reads a value another block has already replaced.  Arithmetic is that of
// fe_ch.hip and lb_collide.hip, so FP64 results are bit-identical to the
// reference.

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "step_tile.h"

namespace ldw {

namespace {

using namespace tile;

template <typename T, class S, int ORDER, int METHOD, bool GZ, bool CONS, bool GMU>
__global__ void __launch_bounds__(S::NT)
binary_step_kernel(Geom g, FeParams fe, LbParams prm, int xchunk, int nbz, int nby, int nbx, int remap,
		   const T * __restrict__ phi, T * __restrict__ phi_out,
		   const T * __restrict__ u_prev, T * __restrict__ u_next,
		   T * __restrict__ rho, const T * __restrict__ g_in,
		   T * __restrict__ g_out, T fz0, T fz1, T fz2, T * __restrict__ csum) {

  __shared__ Smem<T, S> s;

  int bzi, byi, bxi;
  if (!tile_of_block(nbz, nby, nbx, remap, bzi, byi, bxi)) return;
  const int xb = bxi*xchunk;
  const int xe = min(xb + xchunk, g.nx);
  step_tile<T, S, ORDER, METHOD, GZ, CONS, GMU>(s, g, fe, prm, byi*S::TY, bzi*S::TZ, xb, xe, phi, phi_out,
					   u_prev, u_next, rho, g_in, g_out, fz0, fz1, fz2, csum);
}

// Blocks of this tile that the whole GPU holds at once (resident blocks per
// CU from the occupancy of the order-1 kernel, times the CUs); cached.
template <typename T, class S>
int resident_slots() {
  static int slots = 0;
  if (slots == 0) {
    int per_cu = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess
	|| hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, binary_step_kernel<T, S, 1, 0, true, false, false>,
							 S::NT, 0) != hipSuccess || per_cu < 1) {
      (void) hipGetLastError();
      return 256;
    }
    slots = per_cu*prop.multiProcessorCount;
  }
  return slots;
}

// Planes per block.  A block that marches over xchunk planes first loads
// 4 phi planes and forms 2 Q planes (about 3 planes of work), so a launch
// costs about ceil(nblocks/slots)*(xchunk + 3) plane-times: the chunk that
// minimises it balances the prologue against the last, partly filled round
// of blocks (on MI355X 3.6 % faster than a fixed 2048-block grid at 256^3
// FP64, 3.4 % at 128^3 FP32).
inline int step_xchunk(int nx, int64_t ncol, int slots) {
  int best = nx;
  int64_t best_cost = -1;
  for (int xc = nx; ; xc = (xc + 1)/2) {
    const int64_t nblk = ncol*((nx + xc - 1)/xc);
    const int64_t cost = ((nblk + slots - 1)/slots)*(int64_t) (xc + 3);
    if (best_cost < 0 || cost <= best_cost) { best = xc; best_cost = cost; }
    if (xc <= 4) break;
  }
  return best;
}

template <typename T, class S, bool CONS, bool GMU = false>
void launch_shape(const Geom & g, const FeParams & fe, const LbParams & prm, const T * phi,
		  T * phi_out, const T * u_prev, T * u_next, T * rho, const T * g_in,
		  T * g_out, const double fzero[3], T * csum, hipStream_t st) {
  const int bz = (g.nz + S::TZ - 1)/S::TZ;
  const int by = (g.ny + S::TY - 1)/S::TY;
  int xchunk = g.nx;
  if (getenv("LDW_MIN_BLOCKS")) {
    // the fixed rule: halve the chunk until the grid has that many blocks
    const int want = tune_min_blocks();
    while (xchunk > 8 && (int64_t) bz*by*((g.nx + xchunk - 1)/xchunk) < want) xchunk = (xchunk + 1)/2;
  }
  else {
    xchunk = step_xchunk(g.nx, (int64_t) bz*by, resident_slots<T, S>());
  }
  const int bx = (g.nx + xchunk - 1)/xchunk;
  const int ntile = bz*by*bx;
  dim3 grid(((ntile + NXCD - 1)/NXCD)*NXCD), block(S::TZ, S::TY, 1);
  const T f0 = T(fzero[0]), f1 = T(fzero[1]), f2 = T(fzero[2]);
  const bool gz = (prm.relax == 0);
#define L_(O, M, G) hipLaunchKernelGGL((binary_step_kernel<T, S, O, M, G, CONS, GMU>), grid, block, 0, st, \
				       g, fe, prm, xchunk, bz, by, bx, tune_remap_step(bz*by, ntile), phi, phi_out, u_prev, u_next, rho, \
				       g_in, g_out, f0, f1, f2, csum)
#define L_G(O, M) do { if (gz) L_(O, M, true); else L_(O, M, false); } while (0)
#define L_M(O) do { if (fe.force_method == 0) L_G(O, 0); else L_G(O, 1); } while (0)
  if (fe.order == 4) {
    // advection order 4 is built for the 64 x 8 tile only (launch() picks it)
    if constexpr (S::TZ == 64 && S::TY == 8) L_M(4);
  }
  else if (fe.order == 1) L_M(1);
  else if (fe.order == 2) L_M(2);
  else L_M(3);
#undef L_M
#undef L_G
#undef L_
}

// The block tile: 64 x 4 (z, y) columns (256 threads), or the 512-thread
// 64 x 8 / 128 x 4 tiles, which halve the share of pulled rows that cross a
// tile edge in y or z; chosen by precision and lattice size
// (tune_step_tile).  The tile changes only which block computes a site,
// not the arithmetic.  With a compensation array (conserve 1), advection
// order 4 or an external chemical potential gradient the 64 x 8 tile is
// used, so those kernels are built for one tile only.  The kernels without
// the external gradient leave its code out, which keeps the FP32 kernels
// at 3 waves per SIMD.
template <typename T>
void launch(const Geom & g, const FeParams & fe, const LbParams & prm, const T * phi,
	    T * phi_out, const T * u_prev, T * u_next, T * rho, const T * g_in,
	    T * g_out, const double fzero[3], T * csum, hipStream_t st) {
  if (fe.gmu_on) {
    if (csum) launch_shape<T, Shape<64, 8>, true, true>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in,
							 g_out, fzero, csum, st);
    else launch_shape<T, Shape<64, 8>, false, true>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in,
						    g_out, fzero, nullptr, st);
    return;
  }
  if (csum) {
    launch_shape<T, Shape<64, 8>, true>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero,
					csum, st);
    return;
  }
  if (fe.order == 4) {
    launch_shape<T, Shape<64, 8>, false>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero,
					 nullptr, st);
    return;
  }
  switch (tune_step_tile(sizeof(T) == 8 ? 64 : 32, g.nsite)) {
  case 1: launch_shape<T, Shape<64, 8>, false>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero, nullptr, st); break;
  case 2: launch_shape<T, Shape<128, 4>, false>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero, nullptr, st); break;
  default: launch_shape<T, Shape64x4, false>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero, nullptr, st); break;
  }
}

}  // namespace

void launch_binary_step_f64(const Geom & g, const FeParams & fe, const LbParams & prm,
			    const double * phi, double * phi_out, const double * u_prev,
			    double * u_next, double * rho, const double * g_in,
			    double * g_out, const double fzero[3], double * csum, hipStream_t s) {
  launch<double>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero, csum, s);
}

void launch_binary_step_f32(const Geom & g, const FeParams & fe, const LbParams & prm,
			    const float * phi, float * phi_out, const float * u_prev,
			    float * u_next, float * rho, const float * g_in,
			    float * g_out, const double fzero[3], float * csum, hipStream_t s) {
  launch<float>(g, fe, prm, phi, phi_out, u_prev, u_next, rho, g_in, g_out, fzero, csum, s);
}

}  // namespace ldw
