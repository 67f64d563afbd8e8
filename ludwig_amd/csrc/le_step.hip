// le_step.hip -- the binary-fluid time step with Lees-Edwards planes in two
// launches of the tiled single-pass kernel (le_step_tile.h), FP64.
//
// The lattice is cut into nplane slabs between planes; slab s runs from
// the site above plane s to the site below plane s + 1 (the last slab
// wraps round x = 0).  Launch 1 forms every site except the two edge
// slices of each slab, in chunks of xchunk planes, while the per-plane
// force correction sums (a serial chain each, le_plane_sum.h) run beside
// it on a second stream in blocks small enough to share its CUs.  Launch 2
// forms the edge slices, whose force and Cahn-Hilliard x-faces take the
// plane values, and the LB plane reprojection of their post-collision
// distributions (le_lb.h).  Before them the small plane kernels of
// lees_edwards.hip fill the buffer planes (phi, gradient next to the
// planes, u) and the raw x-faces at the planes; after them one more
// displaces and interpolates the reprojections into the plane sites.

#include <hip/hip_runtime.h>

#include "kernels.h"
#include "le_step_tile.h"

namespace ldw {

namespace {

using namespace tile;

struct LeStepArgs {
  const double * phib;
  const double * ub;
  const double * fcor;         // the force correction sums (launch 2)
  const double * chface;       // raw Cahn-Hilliard x-faces (launch 2)
  double ra;
  int jw, je;
  double frw, fre;
  double * sbuf;               // LB plane reprojections (launch 2)
  double duy0, duy1;
};

__device__ __forceinline__ LeSlab slab_of(const Geom & g, const LeGeom & le, const LeStepArgs & a, int sl,
					  bool edges) {
  LeSlab L;
  L.phib = a.phib;
  L.ub = a.ub;
  L.nbuf = (int64_t) le.nxbuf*g.ny*g.nz;
  L.nh = le.nh;
  L.lo = sl;
  L.hi = (sl + 1 < le.nplane) ? sl + 1 : 0;
  L.slab_lo = le.loc[sl];
  L.slab_hi = (sl + 1 < le.nplane) ? le.loc[sl + 1] - 1 : le.loc[0] - 1 + g.nx;
  L.fcor = edges ? a.fcor : nullptr;
  L.ra = a.ra;
  L.chface = edges ? a.chface : nullptr;
  L.jw = a.jw; L.je = a.je;
  L.frw = a.frw; L.fre = a.fre;
  L.sbuf = edges ? a.sbuf : nullptr;
  L.duy0 = a.duy0; L.duy1 = a.duy1;
  return L;
}

// launch 1: the slab interiors, tiles z-fastest, then y, then (slab, chunk),
// in the step kernel's block-to-tile order (tile_of_block)
template <class S, int ORDER, bool GZ, bool GMU>
__global__ void __launch_bounds__(S::NT)
le_step_interior_kernel(Geom g, FeParams fe, LbParams prm, LeGeom le, LeStepArgs a, int xchunk,
			int nchunk, int nbz, int nby, int remap, const double * __restrict__ phi,
			double * __restrict__ phi_out, const double * __restrict__ u_prev,
			double * __restrict__ u_next, double * __restrict__ rho,
			const double * __restrict__ g_in, double * __restrict__ g_out,
			double * __restrict__ force, double * __restrict__ csum) {
  __shared__ Smem<double, S> s;
  int bzi, byi, ck;
  if (!tile_of_block(nbz, nby, le.nplane*nchunk, remap, bzi, byi, ck)) return;
  const int sl = ck/nchunk, k = ck % nchunk;
  const LeSlab L = slab_of(g, le, a, sl, false);
  const int xb = L.slab_lo + 1 + k*xchunk;
  const int xe = min(xb + xchunk, L.slab_hi);
  if (xb >= xe) return;
  le_step_tile<S, ORDER, GZ, GMU>(s, g, fe, prm, L, byi*S::TY, bzi*S::TZ, xb, xe, phi, phi_out, u_prev,
			     u_next, rho, g_in, g_out, force, csum, 0.0, 0.0, 0.0);
}

// launch 2: the edge slices, (slab, edge) after the tiles
template <class S, int ORDER, bool GZ, bool GMU>
__global__ void __launch_bounds__(S::NT)
le_step_edges_kernel(Geom g, FeParams fe, LbParams prm, LeGeom le, LeStepArgs a, int nbz, int nby,
		     const double * __restrict__ phi, double * __restrict__ phi_out,
		     const double * __restrict__ u_prev, double * __restrict__ u_next,
		     double * __restrict__ rho, const double * __restrict__ g_in,
		     double * __restrict__ g_out, double * __restrict__ force, double * __restrict__ csum) {
  __shared__ Smem<double, S> s;
  const int t = blockIdx.x;
  const int bzi = t % nbz;
  const int byi = (t/nbz) % nby;
  const int e = t/(nbz*nby);
  if (e >= 2*le.nplane) return;
  const LeSlab L = slab_of(g, le, a, e/2, true);
  const int x = (e % 2) ? L.slab_hi : L.slab_lo;
  le_step_tile<S, ORDER, GZ, GMU, true>(s, g, fe, prm, L, byi*S::TY, bzi*S::TZ, x, x + 1, phi, phi_out, u_prev,
				   u_next, rho, g_in, g_out, force, csum, 0.0, 0.0, 0.0);
}

template <class S>
int resident_slots_le() {
  static int slots = 0;
  if (slots == 0) {
    int per_cu = 0, dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess
	|| hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, le_step_interior_kernel<S, 1, true, false>,
							 S::NT, 0) != hipSuccess || per_cu < 1) {
      (void) hipGetLastError();
      return 256;
    }
    slots = per_cu*prop.multiProcessorCount;
  }
  return slots;
}

// planes per chunk for m interior planes per slab: the launch costs about
// ceil(blocks/slots) (xchunk + 3) plane-times, as for the plane-free step;
// every chunk length is tried (the slabs are short)
inline int le_xchunk(int m, int64_t ncol, int nplane, int slots) {
  int best = m;
  int64_t best_cost = -1;
  for (int xc = m; xc >= 1; xc--) {
    const int64_t nblk = ncol*nplane*((m + xc - 1)/xc);
    const int64_t cost = ((nblk + slots - 1)/slots)*(int64_t) (xc + 3);
    if (best_cost < 0 || cost < best_cost) { best = xc; best_cost = cost; }
  }
  return best;
}

template <class S, int ORDER, bool GZ, bool GMU>
void launch_both(const Geom & g, const FeParams & fe, const LbParams & prm, const LeGeom & le,
		 const LeStepArgs & a, const double * phi, double * phi_out, const double * u_prev,
		 double * u_next, double * rho, const double * g_in, double * g_out, double * force,
		 double * csum, hipStream_t st, hipEvent_t sums_done) {
  const int bz = (g.nz + S::TZ - 1)/S::TZ;
  const int by = (g.ny + S::TY - 1)/S::TY;
  const int m = g.nx/le.nplane - 2;                   // interior planes per slab
  const int xchunk = le_xchunk(m, (int64_t) bz*by, le.nplane, resident_slots_le<S>());
  const int nchunk = (m + xchunk - 1)/xchunk;
  dim3 block(S::TZ, S::TY, 1);
  // the tile order of the plane-free step: XCD-contiguous where every
  // XCD's range is whole (slab, chunk) x-ranges
  const int ntile = bz*by*le.nplane*nchunk;
  const int remap = tune_remap_step(bz*by, ntile);
  const unsigned n1 = (unsigned) (remap ? ((ntile + NXCD - 1)/NXCD)*NXCD : ntile);
  hipLaunchKernelGGL((le_step_interior_kernel<S, ORDER, GZ, GMU>), dim3(n1), block, 0, st, g, fe, prm, le, a,
		     xchunk, nchunk, bz, by, remap, phi, phi_out, u_prev, u_next, rho, g_in, g_out, force, csum);
  (void) hipStreamWaitEvent(st, sums_done, 0);
  const unsigned n2 = (unsigned) (bz*by*2*le.nplane);
  hipLaunchKernelGGL((le_step_edges_kernel<S, ORDER, GZ, GMU>), dim3(n2), block, 0, st, g, fe, prm, le, a,
		     bz, by, phi, phi_out, u_prev, u_next, rho, g_in, g_out, force, csum);
}

}  // namespace

void launch_le_step(const Geom & g, const FeParams & fe, const LbParams & prm, const LeGeom & le,
		    const double * phib, const double * ub, const double * fcor,
		    const double * chface, double ra, int jw, double frw, int je, double fre,
		    const double duy[2], double * sbuf,
		    const double * phi, double * phi_out, const double * u_prev, double * u_next,
		    double * rho, const double * g_in, double * g_out, double * force, double * csum,
		    hipStream_t st, hipEvent_t sums_done) {
  LeStepArgs a;
  a.phib = phib; a.ub = ub; a.fcor = fcor; a.chface = chface;
  a.ra = ra; a.jw = jw; a.je = je; a.frw = frw; a.fre = fre;
  a.sbuf = sbuf; a.duy0 = duy[0]; a.duy1 = duy[1];
  using S = Shape<64, 8>;
  const bool gz = (prm.relax == 0);
  // the external-gradient term is compiled only into the kernels that see
  // one: without it the interior kernel keeps 230 VGPRs, so two of its
  // waves and a plane-sum wave (launched just before it on the second
  // stream) still fit in a SIMD's 512 registers together
#define LB_(O, G, M) launch_both<S, O, G, M>(g, fe, prm, le, a, phi, phi_out, u_prev, u_next, rho, g_in, g_out, force, csum, st, sums_done)
#define LB_G(O, M) do { if (gz) LB_(O, true, M); else LB_(O, false, M); } while (0)
#define LB_O(M) do { if (fe.order == 1) LB_G(1, M); else if (fe.order == 2) LB_G(2, M); \
		     else if (fe.order == 3) LB_G(3, M); else LB_G(4, M); } while (0)
  if (fe.gmu_on) LB_O(true);
  else LB_O(false);
#undef LB_O
#undef LB_G
#undef LB_
}

}  // namespace ldw
