// Host check of the panel kernel's weighted matvec tile split (skh_device.hpp): for matrix sizes, panel starts
// and grids, every rank's range is non-empty and contiguous, split_owner inverts split_lo on every tile,
// and tall_slots (the row-partial slots a consumer polls) matches a direct count.  Exit 0 when all hold.
#include "skh_device.hpp"
#include <cstdio>
using namespace skh;
int main()
{
    long bad = 0, cases = 0;
    for (int64_t n : {3, 64, 130, 511, 512, 600, 1024, 2048, 3000, 4096, 8192, 12000, 16384}) {
        for (int64_t c0 = 0; c0 <= n - 3; c0 += (n > 4096 ? 96 : 32)) {
            const TallGrid tg = tall_grid(n, c0 + 1);
            for (int64_t G : {8, 32, 100, 256, 512}) {
                const int64_t Ge = G < tg.T ? G : tg.T;
                const TileSplit ts = tile_split(tg, Ge);
                cases++;
                if (split_lo(ts, 0) != 0 || split_lo(ts, Ge) != tg.T) { bad++; continue; }
                for (int64_t g = 0; g < Ge; g++) {
                    const int64_t lo = split_lo(ts, g), hi = split_lo(ts, g + 1);
                    if (hi <= lo) { bad++; printf("empty n=%ld c0=%ld G=%ld g=%ld\n", n, c0, G, g); break; }
                    for (int64_t t = lo; t < hi; t++)
                        if (split_owner(ts, t) != g) { bad++; printf("owner n=%ld c0=%ld G=%ld g=%ld t=%ld -> %ld\n", n, c0, G, g, t, split_owner(ts, t)); goto next; }
                }
                // tall_slots against a direct count of the ranges that meet the strip's tiles
                for (int64_t I = tg.I0; I < tg.nts; I++) {
                    const int64_t f = tall_first(tg, I), cnt = MPS * I + MPS - tg.Jlo;
                    int64_t meet = 0;
                    for (int64_t g = 0; g < Ge; g++)
                        if (split_lo(ts, g) < f + cnt && split_lo(ts, g + 1) > f) meet++;
                    if (tall_slots(tg, I, G) != meet) { bad++; printf("slots n=%ld c0=%ld G=%ld I=%ld\n", n, c0, G, I); break; }
                }
            next:;
            }
        }
    }
    printf("cases %ld bad %ld\n", cases, bad);
    return bad != 0;
}
