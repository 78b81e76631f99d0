// Test driver: builds the ghost refresh table of a block set read from a file
// with the library's own builder (sympygr_amd/csrc/ghost_table.cpp), compiled
// by the host compiler under the address and undefined-behaviour sanitizers.
//
//   ghost_table_sanitize <blocks.bin>
//
// blocks.bin: uint32 count, then `count` sgr_block records.  Prints
// "complete patches nodes interpolated boxes" and, when incomplete, the
// reason on a second line.  Every box is checked against its patch (inside
// it, within the ghost kernels' limits) so the sanitizers see each access the
// kernels' index arithmetic would make on the host.
#include <cstdio>
#include <vector>

#include "ghost_table.h"

using namespace sgr;

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    uint32_t n = 0;
    if (std::fread(&n, sizeof n, 1, f) != 1) return 2;
    std::vector<sgr_block> b(n);
    if (n && std::fread(b.data(), sizeof(sgr_block), n, f) != n) return 2;
    std::fclose(f);
    const GhostTable T = build_ghost_table(b.data(), n);
    uint64_t checked = 0;
    for (int c = 0; c < GHOST_CLASSES && T.complete; c++)
        for (uint32_t k = T.bstart[c]; k < T.bstart[c + 1]; k++) {
            const GhostBox &x = T.boxes[k];
            const GhostPatch &g = T.patches.at((size_t)x.patch);
            const int nt[3] = {c & 1 ? 1 : 6, c & 2 ? 1 : 6, c & 4 ? 1 : 6};
            for (int d = 0; d < 3; d++) {
                // first and last tap of the box's nodes stay on source interior nodes
                const int p0 = ghost_p0(x.lo[d], g.rel[d], g.ht[d], g.hs[d], g.sm[d], nt[d]);
                const int p1 = ghost_p0(x.lo[d] + x.e[d] - 1, g.rel[d], g.ht[d], g.hs[d], g.sm[d], nt[d]) + nt[d] - 1;
                if (p0 < 3 || p1 > g.sm[d] - 4 || x.lo[d] < g.lo[d] || x.lo[d] + x.e[d] > g.lo[d] + g.e[d]) {
                    std::printf("bad box %u axis %d\n", k, d);
                    return 1;
                }
            }
            checked++;
        }
    std::printf("%d %zu %llu %llu %llu\n", T.complete ? 1 : 0, T.patches.size(), (unsigned long long)T.nodes,
                (unsigned long long)T.interp_nodes, (unsigned long long)checked);
    if (!T.complete) std::printf("%s\n", T.msg.c_str());
    return 0;
}
