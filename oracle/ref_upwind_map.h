/* Force-included when building the "upwinded" reference variant: 6th-order
 * centred first/second derivatives, the shift-sign-selected 4th-order
 * advective stencils (bssn/src/derivs.cpp:495-906) and the KO stencils, with
 * BSSN_USE_ADVECTIVE_DERIVS switching the agrad_* arrays to their own sweeps
 * (bssn/scripts/bssnrhs_memalloc_adv.h). No order macro is defined, so
 * bssn/include/derivs.h:68-103 leaves these names to this map. */
#define deriv_x deriv64_x
#define deriv_y deriv64_y
#define deriv_z deriv64_z
#define deriv_xx deriv64_xx
#define deriv_yy deriv64_yy
#define deriv_zz deriv64_zz
#define adv_deriv_x deriv42adv_x
#define adv_deriv_y deriv42adv_y
#define adv_deriv_z deriv42adv_z
#define ko_deriv_x ko_deriv42_x
#define ko_deriv_y ko_deriv42_y
#define ko_deriv_z ko_deriv42_z
