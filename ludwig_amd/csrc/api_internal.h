// api_internal.h -- helpers of ludwig_api.cpp shared with the other host
// sources of libludwig_hip.so (not part of the public C ABI).

#ifndef LDW_API_INTERNAL_H
#define LDW_API_INTERNAL_H

#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ldw {
namespace api {

// record an error (fatal by default, as the reference's pe_fatal); returns -1
int fail(const char * file, int line, const char * fmt, ...);
// select the device and create the library's stream (once)
int device_init(void);
hipStream_t stream(void);
// device address of the fixed scratch of ldw_nn_restore_sums_noalloc on
// the current device (surrogate_api.cpp); once per device at set-up
int noalloc_scratch_init(void);

// Collision parameters for fixed shear and bulk viscosities
// (collision.c:1163-1246); fglobal may be null (zero body force)
LbParams lb_params(double rho0, double eta_shear, double eta_bulk, int relax,
		   const double fglobal[3]);

// f_p of the equilibrium at (rho, u) as lb_1st_moment_equilib_set()
double equilibrium_f(int p, double rho, const double u[3]);

// phi0 + amp (U - 1/2) at global site (ic, jc, kc), 1-based, with the
// per-site generator of field_phi_init_spinodal()
double spinodal_value(int seed, int ic, int jc, int kc, double phi0, double amp);

}  // namespace api
}  // namespace ldw

#define LDW_FAIL(...) return ldw::api::fail(__FILE__, __LINE__, __VA_ARGS__)

#define LDW_HIP_CHECK(call)						\
  do {									\
    hipError_t err_ = (call);						\
    if (err_ != hipSuccess) {						\
      return ldw::api::fail(__FILE__, __LINE__, "%s: %s", #call,		\
			    hipGetErrorString(err_));			\
    }									\
  } while (0)

#endif
