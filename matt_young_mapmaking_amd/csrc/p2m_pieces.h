/* Tile pieces of a ring run: the run [lo, lo+np) of a ring of npr pixels
 * (wrapping past npr to 0) cut at the tile boundaries of width 2^log2w.
 * Host and device code share these, so the enumeration order used to
 * assign segment slots is the one used to write the segments. */
#ifndef P2M_PIECES_H
#define P2M_PIECES_H

#include <hip/hip_runtime.h>
#include <stdint.h>

/* f(tile, lo_in_tile, len) for each piece, in order: the columns of the
 * part before the wrap, then those of the part after it */
template <typename F>
__host__ __device__ static inline void run_pieces(int32_t lo, int32_t np, int32_t npr, int32_t tb, int log2w, F f)
{
    const int32_t a0 = lo, b0 = lo + np;
    for (int piece = 0; piece < 2; ++piece) {
        int32_t a, b;
        if (piece == 0) { a = a0; b = b0 < npr ? b0 : npr; }
        else { if (b0 <= npr) break; a = 0; b = b0 - npr; }
        const int32_t c0 = a >> log2w, c1 = (b - 1) >> log2w;
        for (int32_t col = c0; col <= c1; ++col) {
            const int32_t s0 = col << log2w;
            const int32_t pa = a > s0 ? a : s0;
            const int32_t pb = b < s0 + (1 << log2w) ? b : s0 + (1 << log2w);
            f(tb + col, pa - s0, pb - pa);
        }
    }
}

/* The first two pieces of run_pieces' order (np > 0): their tiles, first
 * in-tile pixels and lengths.  Returns 1 or 2 (the number of pieces, capped
 * at 2). */
__host__ __device__ static inline int first_two_pieces(int32_t lo, int32_t np, int32_t npr, int32_t tb, int log2w,
                                                       int32_t tile[2], int32_t plo[2], int32_t len[2])
{
    const int32_t W = 1 << log2w;
    const int32_t b0 = lo + np;
    const int32_t b = b0 < npr ? b0 : npr;        /* end of the part before the wrap */
    const int32_t c0 = lo >> log2w;
    const int32_t s0 = c0 << log2w;
    const int32_t e0 = b < s0 + W ? b : s0 + W;
    tile[0] = tb + c0; plo[0] = lo - s0; len[0] = e0 - lo;
    if (b > s0 + W) {                             /* a second column before the wrap */
        const int32_t s1 = s0 + W;
        tile[1] = tb + c0 + 1; plo[1] = 0; len[1] = (b < s1 + W ? b : s1 + W) - s1;
        return 2;
    }
    if (b0 > npr) {                               /* the first column after the wrap */
        const int32_t r = b0 - npr;
        tile[1] = tb; plo[1] = 0; len[1] = r < W ? r : W;
        return 2;
    }
    return 1;
}

#endif
