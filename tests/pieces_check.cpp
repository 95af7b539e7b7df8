/* Host check of p2m_pieces.h: run_pieces cuts every ring run into pieces
 * that lie in one tile each, follow the run from its first pixel in ring
 * order (wrapping past the ring's end once), end only at a tile boundary,
 * the ring's end or the run's end, and cover the run exactly; and
 * first_two_pieces gives its first two pieces. */
#include "p2m_pieces.h"
#include <cstdio>
int main() {
    long bad = 0, n = 0;
    const int tb = 1000;
    for (int log2w = 6; log2w <= 11; ++log2w)
    for (int npr : {4, 8, 60, 64, 100, 128, 1000, 1024, 4000, 4096, 16384, 16383}) {
        const int W = 1 << log2w;
        for (int lo = 0; lo < npr; lo += (npr > 200 ? 7 : 1))
        for (int np = 1; np <= npr; np += (npr > 200 ? 13 : 1)) {
            ++n;
            long next = lo, covered = 0;
            int count = 0;
            bool ok = true;
            int t2[2], a2[2], l2[2];
            const int k2 = first_two_pieces(lo, np, npr, tb, log2w, t2, a2, l2);
            run_pieces(lo, np, npr, tb, log2w, [&](int tile, int a, int len) {
                if (count < 2 && (count >= k2 || t2[count] != tile || a2[count] != a || l2[count] != len)) ok = false;
                const int col = tile - tb;
                const long start = (long)col * W + a, stop = start + len;
                if (len < 1 || a < 0 || a + len > W || col < 0 || stop > npr) ok = false;
                if (start != next % npr) ok = false;
                covered += len;
                next += len;
                const bool last = covered == np;
                if (!last && a + len != W && stop != npr) ok = false;
                ++count;
            });
            if (covered != np || count < 1 || k2 != (count < 2 ? count : 2)) ok = false;
            if (!ok) { ++bad; if (bad < 5) printf("lo=%d np=%d npr=%d w=%d\n", lo, np, npr, W); }
        }
    }
    printf("checked %ld runs, %ld mismatches\n", n, bad);
    return bad != 0;
}
