/* Test infrastructure: CPU statement of the ghost (unzip) refresh of a block
 * set whose neighbours may have other spacings and sizes.
 *
 * The reference repository has no ghost fill of its own (the unzip lives in
 * Dendro, outside it), so this is the project's design, written here as
 * plainly as possible and independently of the GPU library's patch tables:
 *
 *  - Every block owns the half-open box of its interior nodes widened by half
 *    a cell: [x_3 - h/2, x_{n-4} + h/2) per axis.  On the octree and uniform
 *    tilings these boxes partition the domain.
 *  - Coordinates are exact integers on the coarsest lattice of unit
 *    u = hmin / (2K) (smallest K = 1..64) on which every block origin and
 *    every half spacing is an integer, so ownership has no rounding ties.
 *  - A ghost node takes the value of the block that owns its position:
 *    per axis either the coinciding node (weight 1) or the 6-point Lagrange
 *    interpolant over source interior nodes p0..p0+5 with
 *    p0 = clamp(q - 2, 3, m - 9) around the node q at or below the position.
 *    Tensor product, summed as sum_c wz_c (sum_b wy_b (sum_a wx_a u)).
 *  - Ghost nodes beyond a flagged (domain boundary) face, or owned by no
 *    block, keep their values.
 *
 * Only interior nodes are read, so the refresh can run in place.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Lat {
    int64_t P[3], H[3];  // node-0 position and spacing in lattice units
    int64_t m[3];        // padded sizes
};

bool make_lattice(uint32_t nb, const uint32_t *sz, const double *pmin, const double *pmax, std::vector<Lat> &L)
{
    double O[3] = {1e300, 1e300, 1e300}, hmin = 1e300;
    for (uint32_t b = 0; b < nb; b++)
        for (int d = 0; d < 3; d++) {
            O[d] = std::fmin(O[d], pmin[3 * b + d]);
            hmin = std::fmin(hmin, (pmax[3 * b + d] - pmin[3 * b + d]) / (sz[3 * b + d] - 1));
        }
    for (int K = 1; K <= 64; K++) {
        const double u = hmin / (2.0 * K);
        bool ok = true;
        L.assign(nb, Lat{});
        for (uint32_t b = 0; b < nb && ok; b++)
            for (int d = 0; d < 3 && ok; d++) {
                const double h = (pmax[3 * b + d] - pmin[3 * b + d]) / (sz[3 * b + d] - 1);
                const double a = (pmin[3 * b + d] - O[d]) / u, c = h / (2.0 * u);
                if (std::fabs(a - std::nearbyint(a)) > 1e-6 || std::fabs(c - std::nearbyint(c)) > 1e-6) ok = false;
                L[b].P[d] = (int64_t)std::nearbyint(a);
                L[b].H[d] = 2 * (int64_t)std::nearbyint(c);
                L[b].m[d] = sz[3 * b + d];
            }
        if (ok) return true;
    }
    return false;
}

// 6-point Lagrange weights at t (nodes 0..5) or the coinciding node.
int axis_weights(int64_t X, const Lat &s, int d, double w[6])
{
    const int64_t lat = X - s.P[d];
    const int64_t q = lat / s.H[d], rem = lat - q * s.H[d];
    if (rem == 0) {
        w[0] = 1.0;
        return (int)q;
    }
    int64_t p0 = q - 2;
    if (p0 < 3) p0 = 3;
    if (p0 > s.m[d] - 9) p0 = s.m[d] - 9;
    const double t = (double)(q - p0) + (double)rem * (1.0 / (double)s.H[d]);
    static const double C[6] = {1.0 / -120.0, 1.0 / 24.0, 1.0 / -12.0, 1.0 / 12.0, 1.0 / -24.0, 1.0 / 120.0};
    for (int j = 0; j < 6; j++) {
        double N = 1.0;
        for (int k = 0; k < 6; k++)
            if (k != j) N = N * (t - (double)k);
        w[j] = N * C[j];
    }
    return (int)p0;
}

}  // namespace

extern "C" int oracle_ghost_fill(double **u, uint32_t nb, const uint64_t *offsets, const uint32_t *sz,
                                 const double *pmin, const double *pmax, const uint32_t *bflag)
{
    std::vector<Lat> L;
    if (!make_lattice(nb, sz, pmin, pmax, L)) return -1;
    for (uint32_t b = 0; b < nb; b++) {
        const Lat &T = L[b];
        int64_t lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
            lo[d] = (bflag[b] >> (2 * d)) & 1 ? 3 : 0;
            hi[d] = (bflag[b] >> (2 * d + 1)) & 1 ? T.m[d] - 3 : T.m[d];
        }
        for (int64_t k = lo[2]; k < hi[2]; k++)
            for (int64_t j = lo[1]; j < hi[1]; j++)
                for (int64_t i = lo[0]; i < hi[0]; i++) {
                    const int64_t idx[3] = {i, j, k};
                    if (i >= 3 && i < T.m[0] - 3 && j >= 3 && j < T.m[1] - 3 && k >= 3 && k < T.m[2] - 3) continue;
                    int64_t X[3];
                    for (int d = 0; d < 3; d++) X[d] = T.P[d] + idx[d] * T.H[d];
                    int owner = -1;
                    for (uint32_t s = 0; s < nb && owner < 0; s++) {
                        if (s == b) continue;
                        bool in = true;
                        for (int d = 0; d < 3 && in; d++) {
                            const int64_t rlo = L[s].P[d] + 5 * (L[s].H[d] / 2);
                            const int64_t rhi = L[s].P[d] + (L[s].m[d] - 4) * L[s].H[d] + L[s].H[d] / 2;
                            in = X[d] >= rlo && X[d] < rhi;
                        }
                        if (in) owner = (int)s;
                    }
                    if (owner < 0) continue;
                    const Lat &S = L[owner];
                    double w[3][6];
                    int p0[3], nt[3];
                    for (int d = 0; d < 3; d++) {
                        const int64_t lat = X[d] - S.P[d];
                        nt[d] = lat % S.H[d] == 0 ? 1 : 6;
                        p0[d] = axis_weights(X[d], S, d, w[d]);
                    }
                    const uint64_t so = offsets[owner];
                    const uint64_t dst = offsets[b] + i + T.m[0] * (j + T.m[1] * k);
                    for (int f = 0; f < 24; f++) {
                        double out = 0.0;
                        for (int c = 0; c < nt[2]; c++) {
                            double plane = 0.0;
                            for (int bb = 0; bb < nt[1]; bb++) {
                                double row = 0.0;
                                const uint64_t r = so + p0[0] + S.m[0] * ((uint64_t)(p0[1] + bb) + S.m[1] * (uint64_t)(p0[2] + c));
                                for (int a = 0; a < nt[0]; a++) row = row + w[0][a] * u[f][r + a];
                                plane = plane + w[1][bb] * row;
                            }
                            out = out + w[2][c] * plane;
                        }
                        u[f][dst] = out;
                    }
                }
    }
    return 0;
}
