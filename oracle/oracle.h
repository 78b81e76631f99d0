/* Test infrastructure: CPU restatement of the reference BSSN RHS
 * (bssn/src/rhs.cpp:80-388 with bssn/src/derivs.cpp stencils).
 *
 * Only the tests, the smoke check and the CPU-baseline leg of the benchmark
 * load this library, and only as the checker / the timed CPU path.  The
 * product (libsympygr_amd.so) never links it.
 */
#ifndef SYMPYGR_ORACLE_H
#define SYMPYGR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physics parameters read by the reference bssnrhs from bssn:: globals
 * (bssn/src/parameters.cpp:99-109, rhs.cpp:148-151,215-219,330). */
typedef struct {
    uint32_t lambda[4];      /* BSSN_LAMBDA (unsigned in the reference) */
    double lambda_f[2];      /* BSSN_LAMBDA_F */
    double eta_const;        /* ETA_CONST */
    double eta_r0;           /* ETA_R0 */
    double eta_damping_exp;  /* ETA_DAMPING_EXP */
    double ko_sigma;         /* KO_DISS_SIGMA */
    int32_t upwind;          /* 0: advective derivs = centred deriv64 (as shipped)
                                1: 4th-order shift-upwinded deriv42adv */
    int32_t rochester;       /* USE_ROCHESTER_GAUGE (rhs.cpp:224-243, GR/bssn.py:194-209) */
    int32_t eta_func;        /* USE_ETA_FUNC (GR/bssn.py:52) */
    uint32_t xi[3];          /* BSSN_XI */
    double eta_func_r0;      /* BSSN_ETA_R0 */
    double eta_power[2];     /* BSSN_ETA_POWER */
    int32_t ccz4;            /* 1: the CCZ4 system (sympygr_amd/codegen/ccz4_eqns.py):
                                25 fields, Theta last; the arrays hold 25 pointers */
    double kappa[3];         /* CCZ4 damping kappa1..3 */
    int32_t dissipation;     /* DISSIPATION_TYPE (rhs.cpp:729-757): 0 Kreiss-Oliger,
                                1 TVB3, 2 TVB5 (BSSN only) */
} oracle_params;

/* One block, same argument meaning as the reference bssnrhs (rhs.h:33-36). */
int oracle_bssnrhs(double **out, const double **in, uint64_t offset,
                   const double *pmin, const double *pmax, const uint32_t *sz,
                   uint32_t bflag, const oracle_params *p);

/* Many blocks; block b uses offsets[b], pmin/pmax[3b..3b+2], sz[3b..3b+2],
 * bflag[b].  nthreads > 0 runs an OpenMP parallel-for over blocks. */
int oracle_bssnrhs_blocks(double **out, const double **in, uint32_t nblocks,
                          const uint64_t *offsets, const double *pmin,
                          const double *pmax, const uint32_t *sz,
                          const uint32_t *bflag, const oracle_params *p,
                          int nthreads);

/* Hamiltonian constraint (out[0]) and momentum constraints (out[1..3]) at
 * the interior points of one block, GR/constraints.py:161-173. */
int oracle_constraints(double **out, const double **in, uint64_t offset, const double *pmin, const double *pmax,
                       const uint32_t *sz, uint32_t bflag);

/* Same, and the real (out[4]) and imaginary (out[5]) parts of Psi4 at the
 * coordinates pmin + (i, j, k) h, GR/constraints.py:46-155. */
int oracle_constraints_psi4(double **out, const double **in, uint64_t offset, const double *pmin,
                            const double *pmax, const uint32_t *sz, uint32_t bflag);

/* Ghost refresh of a block set in place (oracle/ghost_oracle.cpp): every
 * ghost node owned by another block's interior box takes that block's value,
 * coinciding node or 6-point Lagrange interpolant per axis.  -1 when the
 * block coordinates share no common lattice. */
int oracle_ghost_fill(double **u, uint32_t nblocks, const uint64_t *offsets, const uint32_t *sz,
                      const double *pmin, const double *pmax, const uint32_t *bflag);

/* One stencil sweep: which = 0..2 first x/y/z, 3..5 second xx/yy/zz,
 * 6..8 advective x/y/z (needs beta_d), 9..11 Kreiss-Oliger x/y/z,
 * 12..14 TVB3 x/y/z and 15..17 TVB5 x/y/z (beta holds the speeds along the
 * axis).  upwind selects the advective stencil as in oracle_params. */
int oracle_deriv(int which, double *out, const double *u, double h,
                 const uint32_t *sz, const double *beta, uint32_t bflag,
                 int upwind);

/* Characteristic speeds of the TVB dissipation (rhs.cpp:984-1028) at every
 * node of one block, ghost nodes included: lam[d][offset + p]. */
int oracle_spacetime_speeds(double **lam, const double **in, uint64_t offset, const uint32_t *sz);

#ifdef __cplusplus
}
#endif
#endif
