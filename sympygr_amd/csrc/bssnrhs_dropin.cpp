// C++ drop-in bssnrhs(...) over the C ABI (see include/bssn_dropin.h).
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/bssn_dropin.h"
#include "../../include/sympygr_amd.h"

// Parameters the reference reads from bssn:: globals at call time
// (bssn/include/parameters.h, defaults bssn/src/parameters.cpp:99-109).  Weak:
// an application that defines them (Dendro's parameters.cpp) wins.
namespace bssn {
__attribute__((weak)) unsigned int BSSN_LAMBDA[4] = {1, 1, 1, 1};
__attribute__((weak)) double BSSN_LAMBDA_F[2] = {1.0, 0.0};
__attribute__((weak)) double ETA_CONST = 2.0;
__attribute__((weak)) double ETA_R0 = 50.0;
__attribute__((weak)) double ETA_DAMPING_EXP = 1.0;
__attribute__((weak)) double KO_DISS_SIGMA = 0.01;
// gauge-variant parameters (bssn/src/parameters.cpp:50-51,123)
__attribute__((weak)) double BSSN_ETA_R0 = 1.31;
__attribute__((weak)) double BSSN_ETA_POWER[2] = {2.0, 2.0};
__attribute__((weak)) unsigned int BSSN_XI[3] = {0, 0, 0};
// dissipation of bssnrhs_sep (bssn/src/parameters.cpp:109, rhs.cpp:729-757)
__attribute__((weak)) unsigned int DISSIPATION_TYPE = 0;
}  // namespace bssn

namespace {
std::atomic<int32_t> g_adv{SGR_ADV_CENTRED};
std::atomic<int32_t> g_prec{SGR_PREC_FP64};
std::atomic<int> g_device{0};
std::atomic<int32_t> g_gauge{SGR_GAUGE_STANDARD};
std::atomic<int32_t> g_eta{SGR_ETA_CONST};

// first failure since the last sgr_dropin_status call
std::mutex g_status_m;
int g_status = SGR_OK;
std::string g_status_msg;
std::atomic<bool> g_reported{false};

void record_error(int rc, const std::string &msg)
{
    std::lock_guard<std::mutex> lock(g_status_m);
    if (g_status == SGR_OK) {
        g_status = rc;
        g_status_msg = msg;
    }
    if (!g_reported.exchange(true)) std::fprintf(stderr, "sympygr_amd: %s\n", msg.c_str());
}

// The parameters a bssnrhs call uses: the bssn:: globals and the switches.
sgr_params current_params()
{
    sgr_params p;
    for (int i = 0; i < 4; i++) p.lambda[i] = bssn::BSSN_LAMBDA[i];
    p.lambda_f[0] = bssn::BSSN_LAMBDA_F[0];
    p.lambda_f[1] = bssn::BSSN_LAMBDA_F[1];
    p.eta_const = bssn::ETA_CONST;
    p.eta_r0 = bssn::ETA_R0;
    p.eta_damping_exp = bssn::ETA_DAMPING_EXP;
    p.ko_sigma = bssn::KO_DISS_SIGMA;
    p.adv_mode = g_adv;
    p.precision = g_prec;
    p.gauge = g_gauge;
    p.eta_mode = g_eta;
    for (int i = 0; i < 3; i++) p.xi[i] = bssn::BSSN_XI[i];
    p.eta_func_r0 = bssn::BSSN_ETA_R0;
    p.eta_power[0] = bssn::BSSN_ETA_POWER[0];
    p.eta_power[1] = bssn::BSSN_ETA_POWER[1];
    p.formulation = SGR_FORM_BSSN;  // the reference's bssnrhs evolves BSSN
    p.kappa[0] = p.kappa[1] = p.kappa[2] = 0.0;
    p.dissipation = SGR_DISS_KO;    // bssnrhs; bssnrhs_sep reads DISSIPATION_TYPE
    return p;
}

void run(double **uzipVarsRHS, const double **uZipVars, unsigned int offset, const double *ptmin,
         const double *ptmax, const unsigned int *sz, unsigned int bflag, const char *who, bool sep)
{
    sgr_params p = current_params();
    // a DISSIPATION_TYPE other than 0, 1, 2 fails the call (SGR_E_INVALID);
    // the reference prints "Unknown DISSIPATION_TYPE" and adds no dissipation
    if (sep) p.dissipation = bssn::DISSIPATION_TYPE > 2 ? -1 : (int32_t)bssn::DISSIPATION_TYPE;
    const uint32_t s[3] = {sz[0], sz[1], sz[2]};
    const int rc = sgr_bssnrhs_host(uzipVarsRHS, uZipVars, offset, ptmin, ptmax, s, bflag, &p, g_device);
    if (rc != SGR_OK) {
        // bssnrhs has no error channel and may run inside a parallel loop, so
        // nothing is thrown: the block's interior RHS becomes NaN (which the
        // caller's own checks see, as they would a diverging evolution) and
        // the first error is kept for sgr_dropin_status
        char buf[512];
        sgr_last_error(buf, sizeof buf);
        record_error(rc, std::string(who) + ": " + buf);
        if (uzipVarsRHS && sz && sz[0] > 6 && sz[1] > 6 && sz[2] > 6) {
            for (int f = 0; f < SGR_NUM_VARS; f++) {
                if (!uzipVarsRHS[f]) continue;
                for (unsigned int k = 3; k < sz[2] - 3; k++)
                    for (unsigned int j = 3; j < sz[1] - 3; j++)
                        for (unsigned int i = 3; i < sz[0] - 3; i++)
                            uzipVarsRHS[f][offset + i + sz[0] * (j + sz[1] * k)] = NAN;
            }
        }
    }
}
}  // namespace

extern "C" void sgr_dropin_params(void *p)
{
    if (p) *(sgr_params *)p = current_params();
}

extern "C" int sgr_dropin_status(char *buf, size_t len)
{
    std::lock_guard<std::mutex> lock(g_status_m);
    const int rc = g_status;
    if (buf && len) {
        const size_t m = g_status_msg.size() < len - 1 ? g_status_msg.size() : len - 1;
        std::memcpy(buf, g_status_msg.data(), m);
        buf[m] = 0;
    }
    g_status = SGR_OK;
    g_status_msg.clear();
    return rc;
}

extern "C" void sgr_dropin_configure(int32_t adv_mode, int32_t precision, int device)
{
    g_adv = adv_mode;
    g_prec = precision;
    g_device = device;
}

extern "C" void sgr_dropin_configure_gauge(int32_t gauge, int32_t eta_mode)
{
    g_gauge = gauge;
    g_eta = eta_mode;
}

void bssnrhs(double **uzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *ptmin,
             const double *ptmax, const unsigned int *sz, const unsigned int &bflag)
{
    run(uzipVarsRHS, uZipVars, offset, ptmin, ptmax, sz, bflag, "bssnrhs", false);
}

// The staged form computes the same right-hand side (bssn/src/rhs.cpp:392-825)
// with the dissipation DISSIPATION_TYPE selects (rhs.cpp:729-757).
void bssnrhs_sep(double **uzipVarsRHS, const double **uZipVars, const unsigned int &offset, const double *ptmin,
                 const double *ptmax, const unsigned int *sz, const unsigned int &bflag)
{
    run(uzipVarsRHS, uZipVars, offset, ptmin, ptmax, sz, bflag, "bssnrhs_sep", true);
}
