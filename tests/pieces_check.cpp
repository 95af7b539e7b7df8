/* Host check of p2m_pieces.h: first_two_pieces against the first two
 * pieces of run_pieces, over runs of every start and length. */
#include "p2m_pieces.h"
#include <cstdio>
#include <vector>
int main() {
    long bad = 0, n = 0;
    for (int log2w = 6; log2w <= 12; ++log2w)
    for (int npr : {4, 8, 60, 64, 100, 128, 1000, 1024, 4000, 4096, 16384, 16383}) {
        for (int lo = 0; lo < npr; lo += (npr > 200 ? 7 : 1))
        for (int np = 1; np <= npr; np += (npr > 200 ? 13 : 1)) {
            std::vector<int> T, L, N;
            run_pieces(lo, np, npr, 1000, log2w, [&](int t, int a, int l) { T.push_back(t); L.push_back(a); N.push_back(l); });
            int t[2], a[2], l[2];
            int k = first_two_pieces(lo, np, npr, 1000, log2w, t, a, l);
            int want = T.size() < 2 ? (int)T.size() : 2;
            ++n;
            if (k != want) { ++bad; if (bad < 5) printf("count lo=%d np=%d npr=%d w=%d: %d vs %d\n", lo, np, npr, 1 << log2w, k, want); continue; }
            for (int j = 0; j < k; ++j)
                if (t[j] != T[j] || a[j] != L[j] || l[j] != N[j]) { ++bad; if (bad < 5) printf("piece %d lo=%d np=%d npr=%d w=%d\n", j, lo, np, npr, 1 << log2w); }
        }
    }
    printf("checked %ld runs, %ld mismatches\n", n, bad);
    return bad != 0;
}
