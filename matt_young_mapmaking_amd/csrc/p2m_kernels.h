/*
 * p2m_kernels.h -- launch interface between the host runtime (p2m_api.hip)
 * and the gfx950 kernels (p2m_kernels.hip).
 */
#ifndef P2M_KERNELS_H
#define P2M_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "p2m_dev.h"

/* Statistics counters are sharded: shard k of counter c lives at
 * c + k * P2M_CTR_STRIDE, and workgroups add into shard (block % P2M_SHARDS),
 * so no address takes more than 1/64 of the atomics (same-address device
 * atomics serialise beyond the L2).  The host sums the shards.  Append
 * counters (whose value kernels read back) use shard 0 only. */
#define P2M_SHARDS 64
#define P2M_CTR_STRIDE 16

struct DecodeArgs {
    const float *halos;          /* 7 float32 per halo */
    int64_t n, index_base;
    CosmoDev c;
    TabGeom g;
    double cut_low, cut_high, chihview, mmin, dmodel;
    int32_t precut;
    int32_t profile, nside, npix_fs;
    double fov, dpix;
    uint32_t *cand;              /* chunk indices that passed both cuts, all chunks */
    unsigned long long *n_cand;  /* append counter over all chunks of a call */
    unsigned long long *n_cand0; /* its value before this chunk's cull */
    HaloRec *out;                /* record k of candidate k (all chunks) */
    int64_t cap;
    uint32_t *flat_idx;          /* slots of halos the flat-sky painter covers */
    unsigned long long *n_flat_out;
    uint32_t *keys, *vals;       /* spatial sort key and slot of each record (tiled painters) */
    int32_t nphi;                /* azimuth bins of the key */
    uint32_t key_max;            /* the largest key, also given to inert records */
    unsigned long long *counters;/* sharded: [0] passed radial cut, [2] radius errors, [3] (halo, ring) items */
};

struct PaintArgs {
    const HaloRec *halos;
    int64_t n_halo;
    const uint32_t *flat_idx;     /* flat-sky halo list (tiled painters) */
    const unsigned long long *n_flat_dev;
    int64_t n_flat_host;          /* its length, known on the host */
    uint32_t *fs_cnt;             /* box pixels per listed halo (n_flat_host + 1) */
    unsigned long long *fs_off;   /* their exclusive scan */
    int32_t nside, npix_fs, n_maps, profile;
    double fov, dpix;
    TabGeom g;
    const float *tab_flat[P2M_MAXMAPS];
    int64_t n_flat;
    double *map64[P2M_MAXMAPS];
    double *fs64[P2M_MAXMAPS];
    uint32_t *count;
    unsigned long long *hash;
    uint32_t *fs_count;
    unsigned long long *fs_hash;
    unsigned long long *counters; /* [0] listed, [1] hp updates, [2] fs updates, [3] ring gaps */
};

struct BinArgs {
    const BinHalo *halos;         /* in spatial order */
    int64_t n_halo;
    int32_t nside, log2w;
    uint32_t *nr;                 /* rings per halo (n_halo + 1 entries, last 0) */
    const unsigned long long *ring_off;  /* exclusive scan of nr */
    unsigned long long n_items;   /* total (halo, ring) items */
    uint32_t *group_halo;         /* halo of item 64 g (n_items / 64 + 1 entries) */
    const double *ring_z;         /* ring2z of every ring */
    uint32_t *runs;               /* per item: first in-ring pixel | pixel count << 16 */
    uint32_t *pos;                /* per item: slot of its first tile piece within its tile */
    const int32_t *tile_base;     /* first tile of ring r+1 */
    uint32_t *tile_cnt;           /* segments per tile */
    const unsigned long long *tile_off;  /* exclusive scan of tile_cnt */
    uint2 *segs;                  /* (halo, lo<<16 | len) */
    uint4 *ovf;                   /* pieces after an item's first: (tile, slot, halo, lo<<16 | len) */
    unsigned long long *n_ovf;    /* entries of ovf (one counter, not sharded) */
    int64_t ovf_cap;
    int64_t n_tiles;
    unsigned long long *counters; /* [0] listed, [3] ring gaps, [4] segments */
    unsigned long long *err;      /* sharded: pieces whose slot lies outside their tile's range */
};

/* Per-tile geometry, computed once per tile shape: the pixel-centre z and
 * sin(theta) of the tile's ring, its phi parameters (phi = (ip + off)
 * pi/2 / den, kphi = pi/2 / den), the ring's first map pixel, the ring,
 * the tile's first in-ring pixel and its width. */
struct __attribute__((aligned(16))) TileDesc {
    double zr, sth, off, den, kphi;
    int64_t first;
    int32_t ring, ip0, width, pad;
};

#define P2M_TQ_STRIDE 32           /* tile-queue counters one 128-byte line apart */

struct TileArgs {
    const HaloRec *halos;
    const TileHalo *hot;          /* the same halos, packed for the FP32 painter */
    int32_t nside, log2w, n_rings, n_maps, profile, accumulate;
    int64_t n_tiles;
    const int32_t *tile_base;
    const TileDesc *desc;         /* geometry of each tile */
    const unsigned long long *tile_off;
    const uint2 *segs;
    TabGeom g;
    const float *tab_flat[P2M_MAXMAPS];
    const float *tab_q;           /* corner-quad table (k_build_quad): a pixel's nodes ir-1, ir of its
                                     cell, 4 corners of every map, are 8 n_maps contiguous floats */
    int64_t n_quad;               /* floats in tab_q */
    int64_t n_flat;
    const double *inv_node2;      /* exp(-2 (lr0 + k drt)), k = 0..nrt */
    const float *inv_node2f;      /* the same in float */
    float u_k1, u_k0, u_max, band, fr_clamp;
    int32_t ir_clamp;             /* cell and fraction of r = rmint + 1e-5 (maptable.f90:190, 196, 200) */
    float band_c;                 /* relative chord^2 band decided in FP64 */
    double *map64[P2M_MAXMAPS];
    float *map32[P2M_MAXMAPS];
    uint32_t *count;
    unsigned long long *hash;
    uint4 *fb;                    /* deferred FP64 pixels: (ring, ip, halo, 0), P2M_SHARDS sub-lists */
    unsigned long long *fb_count; /* sharded: sub-list k's length at fb_count[k * P2M_CTR_STRIDE] */
    int64_t fb_cap;               /* total; each sub-list holds fb_cap / P2M_SHARDS */
    uint8_t *lost;                /* per tile: set when its deferred pixels did not all fit in fb;
                                     such a tile writes nothing (it is painted again) */
    int32_t redo;                 /* paint only the tiles marked lost, clearing the mark of each
                                     that now fits */
    int64_t n_halo;               /* records in hot / halos */
    unsigned int *tile_next;      /* the FP32 painter's tile queues: one counter per XCD, P2M_TQ_STRIDE
                                     apart (zeroed by p2m_launch_tile_paint) */
    unsigned long long *counters; /* [1] hp updates, [5] FP64 fallbacks */
    unsigned long long *err;      /* sharded: segments whose halo or pixel range is out of bounds */
    unsigned long long *n_lost;   /* sharded: tiles marked lost */
};

struct FallbackArgs {
    const HaloRec *halos;         /* records in slot order */
    const uint32_t *order;        /* spatial position -> slot */
    const uint4 *fb;
    const unsigned long long *fb_count;
    int64_t fb_cap;
    int32_t nside, n_maps, profile;
    TabGeom g;
    const float *tab_flat[P2M_MAXMAPS];
    int64_t n_flat;
    float *map32[P2M_MAXMAPS];
    uint32_t *count;
    unsigned long long *hash;
    const uint8_t *lost;          /* per tile: entries of a lost tile are skipped (it is painted again) */
    const int32_t *tile_base;     /* first tile of ring r+1 */
    int32_t log2w;
    int64_t n_halo;
    unsigned long long *counters; /* [1] hp updates */
    unsigned long long *err;      /* sharded: entries whose halo is out of range */
};

void p2m_launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s);
void p2m_launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t s);
void p2m_launch_fill_f64(double *p, int64_t n, hipStream_t s);
void p2m_launch_fill_f32(float *p, int64_t n, hipStream_t s);
void p2m_launch_clear_ctr(unsigned long long *ctr, int first, int count, hipStream_t s);
void p2m_launch_decode(const DecodeArgs &a, hipStream_t s);
void p2m_launch_paint_plain_hp(const PaintArgs &a, hipStream_t s);
void p2m_launch_paint_fs(const PaintArgs &a, int plain, hipStream_t s);
void p2m_launch_fs_counts(const PaintArgs &a, hipStream_t s);
void p2m_launch_fs_paint(const PaintArgs &a, hipStream_t s);
void p2m_launch_ring_counts(const BinArgs &b, hipStream_t s);
void p2m_launch_ring_z(int32_t nside, double *z, hipStream_t s);
void p2m_launch_tile_desc(int32_t nside, int32_t log2w, int64_t n_tiles, const int32_t *tile_ring,
                          const int32_t *tile_base, TileDesc *desc, hipStream_t s);
void p2m_launch_seg_count(const BinArgs &b, hipStream_t s);
void p2m_launch_seg_emit(const BinArgs &b, hipStream_t s);
uint32_t p2m_key_max(int32_t nside, int32_t nphi);
int p2m_sort_halos(const HaloRec *in, int64_t n, uint32_t key_max, BinHalo *bin, TileHalo *hot, HaloRec *full,
                   int32_t nrt, int32_t nmht, uint32_t *keys, uint32_t *vals, void *tmp, size_t *tmp_bytes,
                   uint32_t *nr, hipStream_t s);
int p2m_scan_u32(const uint32_t *cnt, unsigned long long *off, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t s);
int p2m_tile_max_log2w(void);
void p2m_launch_build_quad(const float *slice, int64_t n_slice, int32_t nrt, int32_t nmht, int32_t nchit,
                           int32_t n_maps, int32_t slot, float *quad, hipStream_t s);
size_t p2m_tile_lds_bytes(int mode, int n_maps, int log2w, int membership, int nrt);
int p2m_tile_occupancy(int mode, int n_maps, int wg, size_t lds);
int p2m_launch_tile_paint(const TileArgs &a, int mode, int n_maps, int grid, int wg, hipStream_t s);
int p2m_launch_fallback(const FallbackArgs &a, int64_t n_max, hipStream_t s);
void p2m_launch_moments(const float *m32, const double *m64, int64_t n, double mean, double *out,
                        unsigned long long *hist, int nb, float dy, hipStream_t s);
void p2m_launch_dpdy_cumulative(const unsigned long long *hist, int nb, float npix, float *dpdy, hipStream_t s);
void p2m_launch_f64_to_f32(const double *in, float *out, int64_t n, hipStream_t s);
void p2m_launch_f32_to_f64(const float *in, double *out, int64_t n, hipStream_t s);

int p2m_launch_selftest_math(int op, const double *x, const double *y, double *out, int64_t n, hipStream_t s);
int p2m_launch_copy_calib(const void *src, void *dst, int64_t n4, hipStream_t s);

/* make_maptable kernel (p2m_maptable.hip) */
int p2m_launch_make_maptable(int32_t model, double nu, double alpha_cmb, int32_t nchit, int32_t nmht, int32_t nrt,
                             const double *d_rofz, const double *d_zofr, float *d_table, hipStream_t s);

#endif
