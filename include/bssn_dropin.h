/* C++ drop-in for the reference per-block entry point.
 *
 * Declares exactly the reference signature of bssnrhs (bssn/include/rhs.h:33-36),
 * so a Dendro build links this library instead of bssn/src/rhs.cpp.  The
 * implementation (libsympygr_amd.so) reads the physics parameters from the
 * same bssn:: globals the reference reads (bssn/src/parameters.cpp:99-109);
 * the library carries weak default definitions that the application's own
 * parameters.cpp overrides at link time.  Arrays are HOST memory, as in the
 * reference; each call stages the block to the GPU (one copy from pinned
 * memory) and copies the interior RHS back.  Calls from several host threads
 * (an OpenMP loop over blocks, as the reference's drivers run) proceed
 * concurrently, each thread with its own staging context and stream.  The
 * reference has no error channel and nothing here throws: a failed call
 * leaves NaN in the block's interior RHS, prints the first failure to stderr
 * and keeps it for sgr_dropin_status().
 */
#ifndef SYMPYGR_BSSN_DROPIN_H
#define SYMPYGR_BSSN_DROPIN_H

#include <stddef.h>
#include <stdint.h>

void bssnrhs(double **uzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *ptmin,
             const double *ptmax, const unsigned int *sz, const unsigned int &bflag);

/* The staged form, bssn/include/rhs.h:39-42 (bssn/src/rhs.cpp:392-825): the
 * same right-hand side computed one variable at a time; here the same
 * kernels serve it.  Like the reference it reads bssn::DISSIPATION_TYPE
 * (0 Kreiss-Oliger, 1 TVB3, 2 TVB5; bssnrhs always uses Kreiss-Oliger). */
void bssnrhs_sep(double **uzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *ptmin,
                 const double *ptmax, const unsigned int *sz, const unsigned int &bflag);

extern "C" {
/* Method switches the reference selects with compile-time macros
 * (BSSN_USE_ADVECTIVE_DERIVS / deriv42adv_*, bssn/include/derivs.h:68-103):
 * adv_mode SGR_ADV_CENTRED or SGR_ADV_UPWIND, precision SGR_PREC_FP64 or
 * SGR_PREC_MIXED, and the HIP device used by bssnrhs.  Default: centred
 * (the shipped 6th-order build), FP64, device 0. */
void sgr_dropin_configure(int32_t adv_mode, int32_t precision, int device);

/* Gauge switches the reference selects with compile-time macros
 * (bssn/src/rhs.cpp:224-243): gauge SGR_GAUGE_STANDARD or SGR_GAUGE_ROCHESTER
 * (USE_ROCHESTER_GAUGE), eta_mode SGR_ETA_CONST or SGR_ETA_FUNC
 * (USE_ETA_FUNC).  Their parameters come from the bssn:: globals BSSN_XI,
 * BSSN_ETA_R0, BSSN_ETA_POWER (bssn/src/parameters.cpp:50-51,123).
 * Default: standard gauge, constant eta. */
void sgr_dropin_configure_gauge(int32_t gauge, int32_t eta_mode);

/* The first failure of bssnrhs / bssnrhs_sep since the last call (an SGR_E*
 * code, SGR_OK if none) and its message; clears it.  No reference
 * counterpart (the reference's bssnrhs cannot fail). */
int sgr_dropin_status(char *buf, size_t len);

/* The parameters the next bssnrhs call will use, as read from the bssn::
 * globals and the two configure calls above (sgr_params of sympygr_amd.h;
 * passed as void * so this header needs no HIP header).  No reference
 * counterpart. */
void sgr_dropin_params(void *params);
}

#endif
