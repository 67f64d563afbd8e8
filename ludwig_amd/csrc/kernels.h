// kernels.h -- launch interface of the HIP kernels (internal to
// libludwig_hip.so; the public C ABI is include/ludwig_hip.h).

#ifndef LDW_KERNELS_H
#define LDW_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ldw {

// Lattice geometry: a periodic nx*ny*nz box, site = (x*ny + y)*nz + z.
struct Geom {
  int nx, ny, nz;
  int64_t nsite;            // nx*ny*nz
};

// Collision parameters, precomputed on the host exactly as the reference
// does (collision.c:1163-1246).
struct LbParams {
  double rtau_shear;
  double rtau_bulk;
  double rtau_ghost[19];    // entries 10..18 used
  double fglobal[3];        // uniform body force
  double rho0;              // shift for the FP32 storage
  int relax;                // 0 M10, 1 BGK, 2 TRT
};

// Symmetric free energy and Cahn-Hilliard parameters
struct FeParams {
  double a, b, kappa;
  double mobility;
  int order;                // advection order 1..5
  int force_method;         // 0 stress divergence, 1 phi grad mu
  int stencil;              // gradient: 7 (3d_7pt_fluid) or 27 (3d_27pt_fluid)
  int gmu_on;               // non-zero: the external gradient gmu acts
  double gmu[3];            // external chemical potential gradient (physics grad_mu)
};

// Launch tuning of the tiled kernels, read once from the environment:
// LDW_XCD_REMAP (1: XCD-contiguous tile ranges, 0: linear order, default) and
// LDW_MIN_BLOCKS (x-chunks are halved until the grid has this many blocks).
int tune_remap(void);
// the tile order of the single-lattice step for ncol column tiles and
// ntile tiles in all (LDW_XCD_REMAP overrides)
int tune_remap_step(int ncol, int ntile);
int tune_min_blocks(void);
// tile of the single-pass step kernel: 0 = 64 x 4, 1 = 64 x 8, 2 = 128 x 4
// (z x y); LDW_STEP_TILE forces one, else chosen by precision and size
int tune_step_tile(int precision, int64_t nsite);

// K1: fused D3Q19 MRT collision + pull propagation.
//   g_in/g_out: 19 planes of nsite (post-collision distributions; when
//   `pull` is 1 the pre-collision value at x is g_in(x - c_p), else g_in(x)).
//   force: 3 planes or nullptr (then only fglobal acts).
//   rho, u: outputs (u 3 planes); may be nullptr.
void launch_collide_f64(const Geom & g, const LbParams & p, const double * g_in,
			double * g_out, const double * force, double * rho,
			double * u, int pull, hipStream_t s);

// Same with FP32 storage of f - w rho0 and FP32 arithmetic.
void launch_collide_f32(const Geom & g, const LbParams & p, const float * g_in,
			float * g_out, const float * force, float * rho,
			float * u, int pull, hipStream_t s);

// Plain pull propagation (materialises one streaming step).
void launch_propagate_f64(const Geom & g, const double * in, double * out,
			  hipStream_t s);
void launch_propagate_f32(const Geom & g, const float * in, float * out,
			  hipStream_t s);

// K2: 7-point gradient + mu + thermodynamic force in one pass.
//   force_mode: 0 -> F = fzero + force (F not read);
//               1 -> F = F + force (read-modify-write).
void launch_phi_force_mu_f64(const Geom & g, const FeParams & fe,
			     const double * phi, double * force, double * mu,
			     int force_mode, const double fzero[3],
			     hipStream_t s);
void launch_phi_force_mu_f32(const Geom & g, const FeParams & fe,
			     const float * phi, float * force, float * mu,
			     int force_mode, const double fzero[3],
			     hipStream_t s);

// mu only (when no force is computed before the Cahn-Hilliard update)
void launch_mu_f64(const Geom & g, const FeParams & fe, const double * phi,
		   double * mu, hipStream_t s);
void launch_mu_f32(const Geom & g, const FeParams & fe, const float * phi,
		   float * mu, hipStream_t s);

// Gradient and Laplacian materialised (for field_grad reads), 7- or
// 27-point stencil
void launch_grad_f64(const Geom & g, int stencil, const double * phi, double * grad,
		     double * delsq, hipStream_t s);

// K3: Cahn-Hilliard advective + diffusive fluxes and the phi update fused.
//   u may be nullptr (no hydrodynamics: zero advective flux).  csum
//   (nullptr, or nsite values kept between steps) selects the compensated
//   update of cahn_hilliard_options_conserve 1.
void launch_cahn_hilliard_f64(const Geom & g, const FeParams & fe,
			      const double * phi, const double * mu,
			      const double * u, double * phi_out, double * csum,
			      hipStream_t s);
void launch_cahn_hilliard_f32(const Geom & g, const FeParams & fe,
			      const float * phi, const float * mu,
			      const float * u, float * phi_out, float * csum,
			      hipStream_t s);

// K23: force (as K2) and the Cahn-Hilliard update (as K3) in one pass,
// 2.5D-tiled through LDS.  Does not write mu.  csum as for K3.
void launch_force_ch_f64(const Geom & g, const FeParams & fe, const double * phi,
			 const double * u, double * force, double * phi_out, double * csum,
			 int force_mode, const double fzero[3], hipStream_t s);
void launch_force_ch_f32(const Geom & g, const FeParams & fe, const float * phi,
			 const float * u, float * force, float * phi_out, float * csum,
			 int force_mode, const double fzero[3], hipStream_t s);

// K123: the whole binary-fluid step (force, Cahn-Hilliard, collision with
// pull propagation) in one pass.  phi -> phi_out, u_prev -> u_next (the
// Cahn-Hilliard part reads u_prev), g_in -> g_out with pull indexing.
void launch_binary_step_f64(const Geom & g, const FeParams & fe, const LbParams & prm,
			    const double * phi, double * phi_out, const double * u_prev,
			    double * u_next, double * rho, const double * g_in,
			    double * g_out, const double fzero[3], double * csum, hipStream_t s);
void launch_binary_step_f32(const Geom & g, const FeParams & fe, const LbParams & prm,
			    const float * phi, float * phi_out, const float * u_prev,
			    float * u_next, float * rho, const float * g_in,
			    float * g_out, const double fzero[3], float * csum, hipStream_t s);

// Replica ensemble (ensemble.hip).  Each replica owns a slice of shared
// pools: its 19 distribution planes start at element off_f of the f pools,
// its phi and rho planes at off_s, its 3 velocity planes at 3*off_s.
struct EnsRep {
  Geom g;
  int64_t off_f;
  int64_t off_s;
  FeParams fe;
  LbParams prm;
};

// One work item: tile [y0, y0+TY) x [z0, z0+TZ) of planes [xb, xe) of replica rep
struct EnsItem {
  int rep, y0, z0, xb, xe, pad[3];
};

// tile shapes: 0 -> 64 x 4, 1 -> 32 x 8, 2 -> 16 x 16, 3 -> 64 x 8,
// 4 -> 128 x 4 (TZ x TY)
constexpr int ENS_NSHAPE = 5;
int ens_shape_tz(int shape);
int ens_shape_ty(int shape);
// resident blocks per CU of the ensemble kernel (0 on error)
int ens_blocks_per_cu(int shape, int precision, int order, int method);

// One step of every replica: phi -> phi_out, u_prev -> u_next, g_in -> g_out
// (pull), rho written.  ctr: two queue counters, zero before the first
// launch; each launch leaves them zero.  M10 collision (ghost modes relax
// to zero).  Returns the launch status.
hipError_t launch_ensemble_step_f64(int shape, int order, int method, int nblocks,
				    const EnsRep * reps, const EnsItem * items, int nitems,
				    unsigned int * ctr, const double * phi, double * phi_out,
				    const double * u_prev, double * u_next, double * rho,
				    const double * g_in, double * g_out, hipStream_t s);
hipError_t launch_ensemble_step_f32(int shape, int order, int method, int nblocks,
				    const EnsRep * reps, const EnsItem * items, int nitems,
				    unsigned int * ctr, const float * phi, float * phi_out,
				    const float * u_prev, float * u_next, float * rho,
				    const float * g_in, float * g_out, hipStream_t s);

// Replace one replica's phi and u by src (4 planes of n floats: phi, ux,
// uy, uz) and its distributions by the equilibrium at (rho, u); FP32
// storage subtracts w rho0
void launch_ensemble_set_state(const float * src, int64_t n, int precision, void * phi, void * u,
			       void * f, const void * rho, double rho0, hipStream_t s);

// b[i] = (float) a[i] for n values of precision 64 (double) or 32 (float)
void launch_to_f32(const void * a, int precision, float * b, int64_t n, hipStream_t s);

// 3x3x3 periodic convolutions, BF16 operands (raw 16-bit), FP32 results
// (conv3d.hip).  Channels-last activations; channel counts multiples of 16,
// at most 160 for the weight gradient.  Weights W[27][cout][cin].
// y (may be null) = bias + conv (+ y when accumulate); act (may be null) =
// bf16(SiLU(that value)) -- the activation fused into the epilogue
hipError_t launch_conv3_fwd(const uint16_t * x, const uint16_t * w, const float * bias, float * y,
			    uint16_t * act, int batch, int nx, int ny, int nz, int cin, int cout,
			    int accumulate, hipStream_t s);
// floats of workspace for launch_conv3_wgrad
int64_t conv3_wgrad_workspace(int batch, int nx, int ny, int nz, int cin, int cout, int * nchunk,
			      int64_t * chunk);
hipError_t launch_conv3_wgrad(const uint16_t * x, const uint16_t * dy, float * dw, float * work,
			      int batch, int nx, int ny, int nz, int cin, int cout, int accumulate,
			      hipStream_t s);
// wf = bf16(w) (same layout), wd[t][ci][co] = bf16(w[26-t][co][ci]) for the
// backward-data pass; either output may be null
hipError_t launch_conv3_weights_prep(const float * w, uint16_t * wf, uint16_t * wd, int cin, int cout,
				     hipStream_t s);

// Surrogate network elementwise kernels (nn_ops.hip)
hipError_t nn_silu_fwd(const float * z, uint16_t * a, int64_t n, hipStream_t s);
hipError_t nn_silu_bwd(const float * z, const float * da, float * dz, uint16_t * dzb, int64_t n,
		       hipStream_t s);
int64_t nn_bias_grad_workspace(int64_t npos, int c);
hipError_t nn_bias_grad(const float * dz, float * db, float * work, int64_t npos, int c,
			int accumulate, hipStream_t s);
hipError_t nn_pool2_fwd(const uint16_t * in, uint16_t * out, int batch, int nx, int ny, int nz, int c,
			hipStream_t s);
hipError_t nn_pool2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c,
			int accumulate, hipStream_t s);
hipError_t nn_up2_fwd(const uint16_t * in, uint16_t * out, int batch, int nx, int ny, int nz, int c,
		      hipStream_t s);
hipError_t nn_up2_bwd(const float * dout, float * din, int batch, int nx, int ny, int nz, int c,
		      hipStream_t s);
hipError_t nn_concat(const uint16_t * a, int ca, const uint16_t * b, int cb, uint16_t * out,
		     int64_t npos, hipStream_t s);
// out (fine fx x fy x fz) = concat(up2(coarse), skip); cu, cs multiples of 8
hipError_t nn_up2_concat(const uint16_t * coarse, int cu, const uint16_t * skip, int cs, uint16_t * out,
			int batch, int fx, int fy, int fz, hipStream_t s);
hipError_t nn_split(const float * d, int ca, int cb, float * da, float * db, int64_t npos,
		    int acc_a, int acc_b, hipStream_t s);
hipError_t nn_to_bf16(const float * a, uint16_t * b, int64_t n, hipStream_t s);
// network input scaling: phi -> (phi - phi_shift) phi_scale, u -> u u_scale,
// prm: the three scalar channels (already scaled)
struct EncodeNorm {
  float phi_shift, phi_scale, u_scale;
  float prm[3];
};
hipError_t nn_encode(const float * state, int64_t nsite, const EncodeNorm & e, uint16_t * out, int c,
		     hipStream_t s);
hipError_t nn_target(const float * now, const float * next, int64_t nsite, float * tgt,
		     float dphi_scale, float du_scale, hipStream_t s);
hipError_t nn_decode(const float * now, const float * y, int cy, int64_t nsite, float * out,
		     float dphi_scale, float du_scale, float phi_max, float u_max, hipStream_t s);
struct ChannelWeights { float w[4]; };
hipError_t nn_mse(const float * y, int cy, const float * tgt, int64_t npos, float * err, uint16_t * dy,
		  float * dyf, const ChannelWeights & w, hipStream_t s);
hipError_t nn_adam(float * w, const float * g, float * m, float * v, int64_t n, float lr_t,
		   float b1, float b2, float eps, float wd, float gscale, hipStream_t s);
hipError_t nn_scale(float * a, int64_t n, float f, hipStream_t s);
// sum of a[i]^2 in FP64 into work[0]; work holds 1 + 512 doubles
hipError_t nn_sumsq(const float * a, int64_t n, double * work, hipStream_t s);
hipError_t nn_sum(const float * a, int64_t n, double * work, hipStream_t s);
hipError_t nn_shift(float * a, int64_t n, float v, hipStream_t s);
// out = the state `in` (4 planes of n^3 floats: phi, ux, uy, uz) under the
// axis permutation perm and the reflections flip (cubic group)
hipError_t nn_cubic(const float * in, float * out, int n, const int perm[3], const int flip[3],
		    hipStream_t s);
// out plane p += (sum now plane p - sum out plane p)/ns for nplane planes of
// ns floats; work holds nn_restore_sums_scratch(nplane) doubles
int nn_restore_sums_scratch(int nplane);
hipError_t nn_restore_sums(const float * now, float * out, int64_t ns, int nplane, double * work,
			   hipStream_t s);

// Plain device copy of nbytes (multiple of 32) for bandwidth calibration
void launch_copy(const void * src, void * dst, int64_t nbytes, hipStream_t s);

// Fill n doubles/floats with a value
void launch_fill_f64(double * a, int64_t n, double v, hipStream_t s);
void launch_fill_f32(float * a, int64_t n, float v, hipStream_t s);

// Compensated (Neumaier/Klein) FP64 sum of n values: partial results in
// work (>= 3*nblocks doubles), final in out[0..2] (sum, cs, ccs).
int  sum_blocks(int64_t n);
void launch_klein_sum_f64(const double * a, int64_t n, double * work,
			  double * out, hipStream_t s);
void launch_klein_sum_f32(const float * a, int64_t n, double * work,
			  double * out, hipStream_t s);
// sum of the 19 planes of g (distribution total), plus rho0*nsite shift
void launch_klein_sum_planes_f32(const float * g, int64_t nsite, int nplane,
				 double * work, double * out, hipStream_t s);

// cahn_hilliard_options_conserve 2: out[0] = the plain FP64 sum of a in
// site order, ((a0 + a1) + a2) + ..., as the reference forms it on one
// thread (one block; its cost grows with n, not with the GPU), then
// a -= (out[0] - target)/n at every site.
void launch_site_order_sum_f64(const double * a, int64_t n, double * out, hipStream_t s);
void launch_site_order_sum_f32(const float * a, int64_t n, double * out, hipStream_t s);
void launch_subtract_drift_f64(double * a, int64_t n, const double * sum, double target,
			       hipStream_t s);
void launch_subtract_drift_f32(float * a, int64_t n, const double * sum, double target,
			       hipStream_t s);

/* ------------------------------------------------------------------------ */
/* Lees-Edwards sliding planes (lees_edwards.hip), FP64                      */
/* ------------------------------------------------------------------------ */

constexpr int LE_MAX_PLANES = 16;
constexpr int LE_MAX_NH = 4;
constexpr int LE_MAX_BUF = 2*LE_MAX_NH*LE_MAX_PLANES;

// Plane geometry of one rank (leesedwards.c:430-455): nxbuf = 2 nh nplane
// buffer planes follow the nx real planes in the "virtual" x index; buffer
// plane b holds real plane loc[b/(2 nh)] - (nh - 1) + b%(2 nh) displaced
// in y, downwards for b%(2 nh) < nh and upwards otherwise.
struct LeGeom {
  int nplane;
  int nh;                       // coordinate halo depth nhalo
  int nxbuf;
  int sep;                      // nx/nplane (integer division)
  int loc[LE_MAX_PLANES];       // 1-based x of the site just below plane p
};

// A y displacement, split as the reference splits it: integer part jdy
// and fraction fr, per buffer plane (or per plane side)
struct LeShift {
  int jdy[LE_MAX_BUF];
  double fr[LE_MAX_BUF];
};

// phi of the buffer planes, four-point Lagrange in y (field.c:459-500)
void launch_le_phi_buffer(const Geom & g, const LeGeom & le, const LeShift & sh,
			  const double * phi, double * phib, hipStream_t s);
// gradient and Laplacian at the real sites (x neighbours across a plane
// from the buffer) and at the buffer sites the reference fills
// (gradient_3d_7pt_fluid.c:231-310, 330-430; gradient_3d_27pt_fluid.c)
void launch_le_grad(const Geom & g, const LeGeom & le, int stencil, const double * phi,
		    const double * phib, double * gr, double * dr, double * gb, double * db,
		    hipStream_t s);
// velocity of the buffer planes, linear in y plus the plane jump
// (hydro.c:350-437)
void launch_le_u_buffer(const Geom & g, const LeGeom & le, const LeShift & sh, double uy,
			const double * u, double * ub, hipStream_t s);
// the flux-form thermodynamic force (phi_force.c:289-520): raw x-face
// fluxes at the planes, per-plane correction sums, then the force
void launch_le_force(const Geom & g, const LeGeom & le, const FeParams & fe,
		     const double * phi, const double * phib, const double * gr,
		     const double * dr, const double * gb, const double * db,
		     double * xface, double * fcor, double ra, double * force,
		     int force_mode, const double fzero[3], hipStream_t s);
// Cahn-Hilliard fluxes with the plane flux averaging
// (phi_cahn_hilliard.c:213-290, 350-413, 613-730) and the update;
// shw/she: the +dy and -dy splits of lees_edw_plane_dy
void launch_le_cahn_hilliard(const Geom & g, const LeGeom & le, const FeParams & fe,
			     const double * phi, const double * phib, const double * dr,
			     const double * db, const double * u, const double * ub,
			     double * xface, int jdy_w, double fr_w, int jdy_e, double fr_e,
			     double * phi_out, double * csum, hipStream_t s);
// the gradient at the real planes either side of each plane and at buffer
// planes nh - 1 and nh only (what the plane pre-pass of the step reads)
void launch_le_grad_planes(const Geom & g, const LeGeom & le, int stencil, const double * phi,
			   const double * phib, double * gr, double * dr, double * gb, double * db,
			   hipStream_t s);
// the raw x-faces at the planes: force fluxes [plane][side][3][nyz] and
// Cahn-Hilliard fluxes [plane][side][nyz] (side 0 fluxe at x = L, side 1
// fluxw at x = L + 1)
void launch_le_force_faces(const Geom & g, const LeGeom & le, const FeParams & fe,
			   const double * phi, const double * phib, const double * gr,
			   const double * dr, const double * gb, const double * db,
			   double * xface, hipStream_t s);
void launch_le_ch_faces(const Geom & g, const LeGeom & le, const FeParams & fe,
			const double * phi, const double * phib, const double * dr,
			const double * db, const double * u, const double * ub,
			double * xface, hipStream_t s);
// the per-plane force correction sums from the force x-faces
void launch_le_plane_sums(const Geom & g, const LeGeom & le, const double * xface, double * fcor,
			  hipStream_t s);
// the whole step in two launches of the tiled single-pass kernel
// (le_step.hip): slab interiors, then (after sums_done) the slab edges with
// the sums fcor; force stored; csum for conserve 1 (or null)
// The edge launch also forms the LB plane reprojection of its sites into
// sbuf (duy: the velocity jump per side), so that launch_le_lb_interp
// completes the plane condition.
void launch_le_step(const Geom & g, const FeParams & fe, const LbParams & prm, const LeGeom & le,
		    const double * phib, const double * ub, const double * fcor,
		    const double * chface, double ra, int jw, double frw, int je, double fre,
		    const double duy[2], double * sbuf,
		    const double * phi, double * phi_out, const double * u_prev, double * u_next,
		    double * rho, const double * g_in, double * g_out, double * force, double * csum,
		    hipStream_t st, hipEvent_t sums_done);
// lb_data_apply_le_boundary_conditions (model_le.c:78-560) on the
// post-collision distributions g: side 0 reprojects cx = +1 at plane
// sites, side 1 cx = -1 above; dj/fr per side, duy the velocity jump
void launch_le_lb_bc(const Geom & g, const LeGeom & le, const int dj[2], const double fr[2],
		     const double duy[2], double * gdist, double * sbuf, hipStream_t s);
// the second half of it alone: the reprojected values in sbuf displaced
// and interpolated into the plane sites of g
void launch_le_lb_interp(const Geom & g, const LeGeom & le, const int dj[2], const double fr[2],
			 double * gdist, const double * sbuf, hipStream_t s);

}  // namespace ldw

#endif
