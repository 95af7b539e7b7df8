/*
 * pks2map_gpu.h -- C ABI of the MI355X halo->map painter (libp2m.so).
 *
 * The reference has no plugin API for this path: its boundary is the per-halo
 * Fortran calls of the halo loop in src/pks2map_src/pks2map.f90:282-303,
 *     call projecthalo_flatsky(xh,yh,zh,rh,chih,mh,redshifth,vh,profile)   (:300)
 *     call projecthalo_healpix(xh,yh,zh,rh,chih,mh,redshifth,vh,profile)   (:301)
 * (haloproject.f90:38-156) over module-global maps (flatskyvars.f90:8,
 * healpixvars.f90:11), the table of loadmaptable (maptable.f90:150-172) and
 * the distance tables of set_rofztable/set_zofrtable (cosmology.f90:51-105).
 * This ABI replaces that loop at chunk granularity: the Fortran host keeps the
 * CLI, the catalogue and table reads, the products and the writers, and hands
 * whole catalogue chunks to the GPU.
 *
 * Conventions: every function returns 0 on success or a negative P2M_E_*
 * code (p2m_strerror explains it); nothing aborts or prints.  The caller owns
 * every input buffer; the library copies what it needs before returning.
 * Maps stay on the device until p2m_fetch_* copies them out.  One context per
 * host thread; calls on a context are serialised by the caller.
 */
#ifndef PKS2MAP_GPU_H
#define PKS2MAP_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define P2M_OK            0
#define P2M_E_INVALID    -1   /* bad argument (nside not a power of 2, n_maps, ...) */
#define P2M_E_SIZE       -2   /* buffer / header size mismatch */
#define P2M_E_HIP        -3   /* HIP runtime error */
#define P2M_E_NOMEM      -4   /* device or host allocation failed */
#define P2M_E_RADIUS     -5   /* a disc radius outside [0, pi] (pixel_routines.F90:1962) */
#define P2M_E_STATE      -6   /* call out of order (no cosmology / table loaded) */
#define P2M_E_IO         -7   /* file could not be written */
#define P2M_E_NODEVICE   -8   /* no GPU visible */
#define P2M_E_INTERNAL   -9   /* a device-side consistency check failed (segment or tile out of range);
                                 the maps of that call are not valid */

/* precision / algorithm of the painter */
#define P2M_PREC_PLAIN_FP64  0  /* one work item per halo, FP64 everything, FP64 global atomics (baseline) */
#define P2M_PREC_FP64        1  /* tiled painter, FP64 geometry + FP64 profile, FP64 LDS tiles */
#define P2M_PREC_FP32_COMP   2  /* tiled painter, FP64 geometry, FP32 profile with FP64 fallback band;
                                   the FP32 values are summed in FP64 LDS tiles (error-compensated
                                   accumulation) and each map pixel is rounded to FP32 once */

/* flags */
#define P2M_FLAG_MEMBERSHIP  1  /* keep per-pixel contribution count and halo-hash sum */

typedef struct p2m_ctx p2m_ctx;

typedef struct {
    int32_t nside;        /* HEALPix nside (power of 2, <= 8192); 0 = no HEALPix map */
    int32_t npix_fs;      /* flat-sky pixels per side; 0 = no flat-sky map */
    double  fov_deg;      /* flat-sky field of view in degrees (pks2map.f90:78,92) */
    int32_t model;        /* 1..3: bbps_delta(model) (bbps_profile.f90:21) */
    int32_t profile;      /* 1 tSZ (v = 1), 2 kSZ (v = -v_r/c) (pksc.f90:287-295) */
    int32_t n_maps;       /* tables painted in one catalogue pass (1..6) */
    int32_t precision;    /* P2M_PREC_* */
    int64_t chunk_halos;  /* halos per streamed chunk (0: 1<<26; at most 1<<30) */
    int32_t device;       /* HIP device ordinal */
    int32_t flags;        /* P2M_FLAG_* */
} p2m_config;

/* run flags */
#define P2M_RUN_PRECUT       1  /* halos were already radially cut by the caller (pks2map.f90:144-160),
                                   e.g. before the scramble/centre branches; decode skips the cut */

typedef struct {
    double omegam, omegal, h;   /* pks2map.f90:118-121 */
    double zmin, zmax;          /* radial cut (pks2map.f90:73,79,134-135) */
    double chihview;            /* observer shift along painter x (pks2map.f90:80,149,292) */
    double mmin;                /* mass cut (pks2map.f90:84) */
    int32_t flags;              /* P2M_RUN_* */
} p2m_run;

typedef struct {
    int64_t n_in, n_cut, n_painted;
    int64_t n_listed_hp, n_upd_hp, n_upd_fs;
    int64_t n_fallback;     /* FP32 path pixels recomputed in FP64 near cell/edge boundaries */
    int64_t n_segments;     /* (halo, ring, tile) work items */
    int64_t n_ring_gaps;    /* empty rings between non-empty rings (quirk Q15); expect 0 */
    int64_t n_radius_err;   /* halos with disc radius outside [0, pi] (skipped) */
    double  cut_low, cut_high;
    /* device time of the last p2m_paint* call: total, decode, binning (spatial
     * sort to segment lists), healpix painting, flat-sky painting (the tiled
     * modes run it on a second stream beside the binning) */
    double  ms_total, ms_decode, ms_bin, ms_paint, ms_finish;
    int32_t tile_w, tile_wg;  /* tile shape used (autotuned or set) */
    double  ms_tile;          /* the tile painter kernel alone, within ms_paint (0 for plain FP64) */
    /* catalogue handed over in host memory: bytes copied to the device, the
     * copy stream's time from the call's start to its last copy, and whether
     * every block was pinned (page-locked) memory */
    int64_t h2d_bytes;
    double  ms_h2d;
    int32_t h2d_pinned;
    int32_t n_redo_tiles;     /* tiles repainted after the deferred-pixel list grew (0 normally) */
} p2m_stats;

/* Context on one GPU; allocates the maps (zeroed).  Replaces the allocations
 * of pks2map.f90:110-115. */
int  p2m_create(p2m_ctx **ctx, const p2m_config *cfg);
void p2m_destroy(p2m_ctx *ctx);
const char *p2m_strerror(int code);

/* Cosmology, distance tables and cuts.  Replaces pks2map.f90:118-135 and
 * set_rofztable/set_zofrtable (cosmology.f90:51-105). */
int p2m_set_run(p2m_ctx *ctx, const p2m_run *run);
/* rofz / zofr of the tables built by p2m_set_run (cosmology.f90:137-181). */
int p2m_rofz(p2m_ctx *ctx, double z, double *chi);
int p2m_zofr(p2m_ctx *ctx, double chi, double *z);

/* Profile table of map slot `slot` (0..n_maps-1): the 36-byte header and
 * table(2,nrt,nmht,nchit) as read by loadmaptable (maptable.f90:150-172). */
int p2m_load_table(p2m_ctx *ctx, int slot, const int32_t hdr_i[3], const float hdr_f[6],
                   const float *table, int64_t n_table);

/* Paint n halos (7 float32 each, the .pksc record layout of pksc.f90:64-66)
 * from host memory into the maps; accumulates across calls.  Replaces the
 * halo loop pks2map.f90:276-303 for one rank's halos.  `index_base` is the
 * catalogue index of halos[0] (used for the membership hash). */
int p2m_paint(p2m_ctx *ctx, const float *halos, int64_t n, int64_t index_base);
/* Same, with the halos already in device memory (HBM-resident input). */
int p2m_paint_device(p2m_ctx *ctx, const float *d_halos, int64_t n, int64_t index_base);

/* Page-locked host memory for catalogue blocks (hipHostMalloc on the
 * context's device).  Host catalogues in such memory reach the device by
 * DMA overlapped with the decode, at the PCIe rate; pageable memory is
 * staged by the runtime at a lower rate.  Replaces the allocate() of the
 * halo arrays in loadpksc (pksc.f90:54-58).  p2m_host_free releases it. */
int p2m_host_alloc(p2m_ctx *ctx, int64_t bytes, void **ptr);
int p2m_host_free(p2m_ctx *ctx, void *ptr);

/* Streamed catalogue pass: the block read of loadpksc (pksc.f90:51-66) fed
 * to the halo loop (pks2map.f90:276-303) block by block.
 *   p2m_stream_begin(ctx, n_total)   -- n_total: halos of all blocks together
 *   p2m_stream_push(ctx, halos, n, index_base)   -- one block of 7-float records
 *       (host memory, pinned or pageable); copied and decoded asynchronously.
 *       The caller may overwrite a block's buffer once the NEXT push (or the
 *       finish) has returned, so two buffers used in turn keep the copy
 *       engine busy while the host reads the next block.
 *   p2m_stream_finish(ctx)           -- bins and paints every pushed halo
 * p2m_paint is begin + one push + finish.  Pushing more than n_total halos
 * in total returns P2M_E_SIZE. */
int p2m_stream_begin(p2m_ctx *ctx, int64_t n_total);
int p2m_stream_push(p2m_ctx *ctx, const float *halos, int64_t n, int64_t index_base);
int p2m_stream_finish(p2m_ctx *ctx);

/* Zero the maps (and membership arrays). */
int p2m_clear(p2m_ctx *ctx);

/* Copy map slot `slot` out; any pointer may be NULL.  hp* hold 12*nside^2
 * values (RING order), fs* npix_fs^2 values (index j + npix*(k-1) of
 * haloproject.f90:148, 0-based). */
int p2m_fetch_maps(p2m_ctx *ctx, int slot, float *hp32, double *hp64, float *fs32, double *fs64);
/* Membership fingerprints (P2M_FLAG_MEMBERSHIP): per pixel the number of
 * accepted halos and the wrapping sum of splitmix64(catalogue index). */
int p2m_fetch_membership(p2m_ctx *ctx, uint32_t *hp_count, uint64_t *hp_hash,
                         uint32_t *fs_count, uint64_t *fs_hash);
int p2m_get_stats(p2m_ctx *ctx, p2m_stats *st);

/* Tile shape of the tiled painter: width in pixels along a ring (power of 2,
 * 64..2048) and workgroup size (a multiple of 64, 64..512; 64..256 for
 * P2M_PREC_FP32_COMP).
 * 0,0 (and a context that never calls this) = the library picks the shape on
 * the next paint call: with at least 2e5 halos past the cuts and cleared maps
 * it times each candidate shape on those halos (painting, then clearing the
 * maps) and keeps the fastest; otherwise, or with P2M_NO_AUTOTUNE set in the
 * environment, 1024,256.  The shape then holds until the next call here. */
int p2m_set_tile(p2m_ctx *ctx, int32_t tile_w, int32_t tile_wg);

/* Products of pks2map.f90:321-357 computed on the device from map slot
 * `slot`: moments (fs1,fs2,fs3,hp1,hp2,hp3), RMS (fs,hp) and the cumulative
 * dP/dy histogram (10000 bins of 1e-8, quirk Q13 binning). */
int p2m_products(p2m_ctx *ctx, int slot, double moments[6], double rms[2], float dpdy[10000]);

/* p2m_products with no allocation at all (no hipMalloc, hipFree or host
 * heap): it has no counterpart in the reference.  It is for callers that
 * cannot allocate -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation.  Its partial sums, dP/dy histogram and
 * cumulative dP/dy live in a fixed 128 KiB device array per GPU, held until
 * the process ends; no input changes that size (the histogram has 10000 bins
 * and the moments 8 partial sums for any map).  Calls that overlap on one
 * GPU share the array and race: serialise them. */
int p2m_products_noalloc(p2m_ctx *ctx, int slot, double moments[6], double rms[2], float dpdy[10000]);

/* Writers of pks2map.f90:359-381: <base>_fs.map, <base>_hp.fits, <base>_dp.bin
 * for map slot `slot` (file layouts of the reference; FITS written without
 * cfitsio).  Any of the three may be skipped by passing write_* = 0. */
int p2m_write_outputs(p2m_ctx *ctx, int slot, const char *base, int write_fs, int write_hp, int write_dp);

/* make_maptable on the GPU (maptable.f90:25-146, make_maptable.f90:1-46):
 * fills hdr_i/hdr_f and table (2*nrt*nmht*nchit floats, host memory). */
int p2m_make_maptable(int32_t device, int32_t model, double nu, double alpha_cmb,
                      int32_t nchit, int32_t nmht, int32_t nrt,
                      int32_t hdr_i[3], float hdr_f[6], float *table, int64_t n_table);

/* Diagnostic (no counterpart in the reference): evaluate the library's FP64
 * elementary function `op` on the device for n inputs, so callers can check
 * that device and host produce the same bits.  op: 0 sqrt, 1 sin, 2 cos,
 * 3 tan, 4 asin, 5 acos, 6 atan2(x, y), 7 log, 8 exp, 9 pow(x, y). */
int p2m_selftest_math(int32_t device, int32_t op, const double *x, const double *y, double *out, int64_t n);

/* Measurement helper (no counterpart in the reference): copies n_bytes
 * (a multiple of 16, at most 1<<34) between two device buffers `reps` times
 * with a plain streaming kernel and returns the mean kernel time in *ms.
 * Its known traffic (n_bytes read, n_bytes written per launch) calibrates
 * the HBM byte counters of a profiler. */
int p2m_selftest_copy(int32_t device, int64_t n_bytes, int32_t reps, double *ms);

/* Library build identification. */
const char *p2m_version(void);

#ifdef __cplusplus
}
#endif
#endif
