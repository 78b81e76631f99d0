// Test program: an application that defines the bssn:: parameter globals
// itself (as Dendro's bssn/src/parameters.cpp does) and links
// libsympygr_amd.so.  Its definitions must win over the library's weak
// defaults: prints the parameters the drop-in bssnrhs would use.
#include <cstdio>

#include "bssn_dropin.h"
#include "sympygr_amd.h"

namespace bssn {
unsigned int BSSN_LAMBDA[4] = {1, 0, 1, 0};
double BSSN_LAMBDA_F[2] = {0.75, 0.25};
double ETA_CONST = 1.25;
double ETA_R0 = 30.0;
double ETA_DAMPING_EXP = 2.0;
double KO_DISS_SIGMA = 0.37;
double BSSN_ETA_R0 = 1.5;
double BSSN_ETA_POWER[2] = {1.0, 3.0};
unsigned int BSSN_XI[3] = {2, 1, 3};
}  // namespace bssn

int main()
{
    sgr_params p;
    sgr_dropin_params(&p);
    std::printf("%u %u %u %u %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %u %u %u\n", p.lambda[0],
                p.lambda[1], p.lambda[2], p.lambda[3], p.lambda_f[0], p.lambda_f[1], p.eta_const, p.eta_r0,
                p.eta_damping_exp, p.ko_sigma, p.eta_func_r0, p.eta_power[0], p.eta_power[1], p.xi[0], p.xi[1],
                p.xi[2]);
    return 0;
}
