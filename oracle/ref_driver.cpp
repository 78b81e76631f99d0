// Test infrastructure only: a C entry point into the compiled reference
// `bssnrhs` (bssn/src/rhs.cpp:80-388) so the fixture generator can call it
// through ctypes.  This file is ours; it is linked with the reference's own
// sources by oracle/Makefile and is never part of the product library.
#include "rhs.h"
#include "derivs.h"
#include "parameters.h"

extern "C" {

// Parameters are the reference's mutable globals (bssn/src/parameters.cpp:99-109);
// they are set here before every call, exactly as a Dendro driver would.
int ref_bssnrhs(double **out, const double **in, unsigned int offset,
                const double *pmin, const double *pmax, const unsigned int *sz,
                unsigned int bflag, const unsigned int *lambda,
                const double *lambda_f, double eta_const, double eta_r0,
                double eta_damping_exp, double ko_sigma)
{
    for (int i = 0; i < 4; i++) bssn::BSSN_LAMBDA[i] = lambda[i];
    bssn::BSSN_LAMBDA_F[0] = lambda_f[0];
    bssn::BSSN_LAMBDA_F[1] = lambda_f[1];
    bssn::ETA_CONST = eta_const;
    bssn::ETA_R0 = eta_r0;
    bssn::ETA_DAMPING_EXP = eta_damping_exp;
    bssn::KO_DISS_SIGMA = ko_sigma;
    const unsigned int off = offset;
    const unsigned int bf = bflag;
    bssnrhs(out, in, off, pmin, pmax, sz, bf);
    return 0;
}

// The staged form of the same right-hand side (bssn/src/rhs.cpp:392-825):
// one generated file per variable (a_rhs.cpp, b_rhs.cpp, ...), same
// parameters, same boundary and dissipation steps.
int ref_bssnrhs_sep(double **out, const double **in, unsigned int offset,
                    const double *pmin, const double *pmax, const unsigned int *sz,
                    unsigned int bflag, const unsigned int *lambda,
                    const double *lambda_f, double eta_const, double eta_r0,
                    double eta_damping_exp, double ko_sigma)
{
    for (int i = 0; i < 4; i++) bssn::BSSN_LAMBDA[i] = lambda[i];
    bssn::BSSN_LAMBDA_F[0] = lambda_f[0];
    bssn::BSSN_LAMBDA_F[1] = lambda_f[1];
    bssn::ETA_CONST = eta_const;
    bssn::ETA_R0 = eta_r0;
    bssn::ETA_DAMPING_EXP = eta_damping_exp;
    bssn::KO_DISS_SIGMA = ko_sigma;
    const unsigned int off = offset;
    const unsigned int bf = bflag;
    bssnrhs_sep(out, in, off, pmin, pmax, sz, bf);
    return 0;
}

// Many blocks on `nthreads` host threads, the way the reference's own GPU
// example driver runs the CPU path (bssn/cuda_gr/src/gpuBSSNExample.cpp:
// 267-300: an OpenMP parallel for over blocks, one bssnrhs call each, the
// parameter globals set once beforehand).  Block b reads in + offsets[b] and
// writes out + offsets[b]; pmin/pmax/sz hold 3 values per block.
int ref_bssnrhs_blocks(double **out, const double **in, unsigned int nblocks, const unsigned long long *offsets,
                       const double *pmin, const double *pmax, const unsigned int *sz, const unsigned int *bflag,
                       const unsigned int *lambda, const double *lambda_f, double eta_const, double eta_r0,
                       double eta_damping_exp, double ko_sigma, int nthreads)
{
    for (int i = 0; i < 4; i++) bssn::BSSN_LAMBDA[i] = lambda[i];
    bssn::BSSN_LAMBDA_F[0] = lambda_f[0];
    bssn::BSSN_LAMBDA_F[1] = lambda_f[1];
    bssn::ETA_CONST = eta_const;
    bssn::ETA_R0 = eta_r0;
    bssn::ETA_DAMPING_EXP = eta_damping_exp;
    bssn::KO_DISS_SIGMA = ko_sigma;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads)
    for (unsigned int b = 0; b < nblocks; b++) {
        double *o[24];
        const double *u[24];
        for (int f = 0; f < 24; f++) {
            o[f] = out[f] + offsets[b];
            u[f] = in[f] + offsets[b];
        }
        const unsigned int off = 0;
        const unsigned int bf = bflag[b];
        bssnrhs(o, u, off, pmin + 3 * b, pmax + 3 * b, sz + 3 * b, bf);
    }
    return 0;
}

// Parameters of the gauge variants (bssn/src/parameters.cpp:50-51,123):
// the eta function R0 * sqrt(igt^ij d_i chi d_j chi) / (1 - chi^p0)^p1
// (USE_ETA_FUNC) and the Rochester shift weights BSSN_XI (USE_ROCHESTER_GAUGE).
void ref_set_gauge_params(double eta_func_r0, const double *eta_power, const unsigned int *xi)
{
    bssn::BSSN_ETA_R0 = eta_func_r0;
    bssn::BSSN_ETA_POWER[0] = eta_power[0];
    bssn::BSSN_ETA_POWER[1] = eta_power[1];
    for (int i = 0; i < 3; i++) bssn::BSSN_XI[i] = xi[i];
}

// DISSIPATION_TYPE (bssn/src/parameters.cpp:109), read by bssnrhs_sep
// (bssn/src/rhs.cpp:729-757): 0 Kreiss-Oliger, 1 TVB3, 2 TVB5.
void ref_set_dissipation(unsigned int type) { bssn::DISSIPATION_TYPE = type; }

// One TVB sweep (bssn/src/derivs.cpp:1348-2054): which 0..2 disstvb3_{x,y,z},
// 3..5 disstvb5_{x,y,z}; lam is the speed along the sweep axis.
int ref_tvb(int which, double *out, const double *u, const double *lam, double h, const unsigned int *sz,
            unsigned int bflag)
{
    switch (which) {
    case 0: disstvb3_x(out, u, lam, h, sz, bflag); break;
    case 1: disstvb3_y(out, u, lam, h, sz, bflag); break;
    case 2: disstvb3_z(out, u, lam, h, sz, bflag); break;
    case 3: disstvb5_x(out, u, lam, h, sz, bflag); break;
    case 4: disstvb5_y(out, u, lam, h, sz, bflag); break;
    case 5: disstvb5_z(out, u, lam, h, sz, bflag); break;
    default: return -1;
    }
    return 0;
}

// The TVB characteristic speeds of one block (bssn/src/rhs.cpp:984-1028).
void ref_spacetime_speeds(double **lam, const double **in, const unsigned int *sz)
{
    max_spacetime_speeds(lam[0], lam[1], lam[2], in[0], in[6], in[7], in[8], in[12], in[13], in[14], in[15], in[16],
                         in[17], in[1], sz);
}

// Single stencil sweeps (bssn/src/derivs.cpp), selected by a small integer:
// 0..2 deriv_{x,y,z}, 3..5 deriv_{xx,yy,zz}, 6..8 adv_deriv_{x,y,z},
// 9..11 ko_deriv_{x,y,z}.  `beta` is only read by the advective sweeps.
int ref_deriv(int which, double *out, const double *u, double h,
              const unsigned int *sz, const double *beta, unsigned int bflag)
{
    switch (which) {
    case 0: deriv_x(out, u, h, sz, bflag); break;
    case 1: deriv_y(out, u, h, sz, bflag); break;
    case 2: deriv_z(out, u, h, sz, bflag); break;
    case 3: deriv_xx(out, u, h, sz, bflag); break;
    case 4: deriv_yy(out, u, h, sz, bflag); break;
    case 5: deriv_zz(out, u, h, sz, bflag); break;
    case 6: adv_deriv_x(out, u, h, sz, beta, bflag); break;
    case 7: adv_deriv_y(out, u, h, sz, beta, bflag); break;
    case 8: adv_deriv_z(out, u, h, sz, beta, bflag); break;
    case 9: ko_deriv_x(out, u, h, sz, bflag); break;
    case 10: ko_deriv_y(out, u, h, sz, bflag); break;
    case 11: ko_deriv_z(out, u, h, sz, bflag); break;
    default: return -1;
    }
    return 0;
}

void ref_fake_initial_data(double x, double y, double z, double *u)
{
    bssn::fake_initial_data(x, y, z, u);
}

}
