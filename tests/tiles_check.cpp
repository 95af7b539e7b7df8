/* Host check of the ring tiles (ring_tiles, tile_span in p2m_dev.h) that the
 * library numbers with its host loop and describes with k_tile_desc: for
 * every nside 1..8192 and tile width 64..2048, the rings' first pixels
 * follow one another from 0 to 12 nside^2, each ring's tiles start at
 * multiples of W, hold 1..W pixels, stay inside their ring and cover it
 * exactly, so no tile's pixels first + ip0 .. first + ip0 + width - 1 leave
 * the map. */
#include "p2m_dev.h"
#include <cstdio>
int main() {
    long bad = 0, tiles = 0;
    for (int32_t nside = 1; nside <= 8192; nside *= 2) {
        const int64_t npix = 12LL * nside * nside;
        for (int log2w = 6; log2w <= 11; ++log2w) {
            const int32_t W = 1 << log2w;
            int64_t next_first = 0, tile_base = 0;
            for (int32_t ring = 1; ring <= 4 * nside - 1; ++ring) {
                int32_t npr, kshift;
                int64_t first;
                ring_info(nside, ring, npr, kshift, first);
                /* pixels per ring: 4 min(ring, nside, 4 nside - ring) (Gorski et al. 2005, sec. 4) */
                int32_t m = ring < nside ? ring : nside;
                if (4 * nside - ring < m) m = 4 * nside - ring;
                if (npr != 4 * m || first != next_first) ++bad;
                next_first = first + npr;
                const int32_t nt = ring_tiles(nside, ring, log2w);
                int64_t covered = 0;
                for (int32_t col = 0; col < nt; ++col) {
                    int64_t f;
                    int32_t ip0, width;
                    tile_span(nside, ring, col, log2w, f, ip0, width);
                    ++tiles;
                    if (f != first || ip0 != col * W || width < 1 || width > W || ip0 + width > npr ||
                        f + ip0 + width > npix || ip0 != covered) {
                        if (++bad < 5) printf("nside %d w %d ring %d col %d\n", nside, W, ring, col);
                    }
                    covered += width;
                }
                if (covered != npr) ++bad;
                tile_base += nt;
            }
            if (next_first != npix) ++bad;
        }
    }
    printf("checked %ld tiles, %ld mismatches\n", tiles, bad);
    return bad != 0;
}
