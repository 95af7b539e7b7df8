/*
 * p2m_oracle.h -- CPU restatement of Peak Patch pks2map's halo->map path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library or executables
 * links this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, as the checker.
 *
 * Precision modes (P2MO_MODE_*):
 *   FP64  : every Fortran `real` of the reference evaluated in double from the
 *           float32 file/CLI inputs; named constants keep the reference's
 *           values; elementary functions from p2m_math.h (portable, so the
 *           GPU can reproduce pixel membership bit for bit).  This is the
 *           grading oracle.
 *   ref32 : the reference's own float32 arithmetic (p2m_oracle_ref32.c), to
 *           state how far maps in the reference's precision lie from the
 *           FP64 oracle, and as the CPU baseline.
 * The library is built twice: libp2mo.so (portable math) and libp2mo_libm.so
 * (same source, glibc libm), to show the portable math does not change any
 * membership decision on the test catalogues.
 */
#ifndef P2M_ORACLE_H
#define P2M_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cosmology + distance tables (cosmology.f90:51-214). */
typedef struct {
    double omegam, omegal, h;      /* pks2map.f90:118-121 */
    double rho_0, rhocrit_0;       /* pks2map.f90:127-128 */
    int    nrofz, nzofr;           /* 10000 each */
    double rofz_min, rofz_max, drofz;
    double zofr_min, zofr_max, dzofr;
    double *rofz, *zofr;           /* malloc'd by p2mo_cosmo_init */
} p2mo_cosmo;

/* Profile table (maptable.f90:150-172) with zero padding for the
 * upper-edge index aliasing of interpolate_table (quirk Q6). */
typedef struct {
    int32_t nchit, nmht, nrt;
    double chimint, chimaxt, mhmint, mhmaxt, rmint, rmaxt; /* float32 header values */
    double drt, dchit, dmht;
    const float *t;       /* 2*nrt*nmht*nchit, profile fastest */
    int64_t nt;           /* elements in t */
} p2mo_table;

typedef struct {
    int32_t nside;        /* HEALPix nside (0: no HEALPix map) */
    int32_t npix;         /* flat-sky pixels per side (0: no flat map) */
    double  fov_deg;      /* flat-sky field of view in degrees (CLI value) */
    double  zmin, zmax;   /* redshift cut (CLI values) */
    double  chihview;     /* observer shift along painter x (CLI value) */
    int32_t model;        /* 1..3 */
    int32_t profile;      /* 1 tSZ, 2 kSZ */
    double  mmin;         /* pks2map.f90:84 */
    int32_t nthreads;
} p2mo_run;

typedef struct {
    int64_t n_in, n_cut, n_painted;
    int64_t n_listed_hp;  /* pixels returned by query_disc */
    int64_t n_upd_hp, n_upd_fs;
    int64_t n_ring_gaps;  /* quirk Q15 detector: must be 0 */
    double  cut_low, cut_high;
} p2mo_stats;

/* cosmology */
int    p2mo_cosmo_init(p2mo_cosmo *c, double omegam, double omegal, double h);
void   p2mo_cosmo_free(p2mo_cosmo *c);
double p2mo_zofr(const p2mo_cosmo *c, double r);
double p2mo_rofz(const p2mo_cosmo *c, double z);
double p2mo_rhocrit(const p2mo_cosmo *c, double z);
double p2mo_deltacrit(const p2mo_cosmo *c, double z);

/* table */
int    p2mo_table_init(p2mo_table *tb, const int32_t hdr_i[3], const float hdr_f[6], const float *padded, int64_t n_padded);
double p2mo_interpolate_table(const p2mo_table *tb, const p2mo_cosmo *c, double theta, double z, double *mh, int profile);

/* HEALPix primitives (vendored HEALPix 3.10 semantics) */
int64_t p2mo_nside2npix(int32_t nside);
int32_t p2mo_ring_num(int32_t nside, double z);
double  p2mo_ring2z(int32_t nside, int32_t ir);
void    p2mo_pix2vec_ring(int32_t nside, int64_t ipix, double v[3]);
double  p2mo_angdist(const double v1[3], const double v2[3]);
/* returns nlist (or -1 on radius error); list must hold 12*nside^2 entries */
int64_t p2mo_query_disc(int32_t nside, const double v0[3], double radius, int32_t *list, int64_t *n_gaps);

/* the painter: maps are accumulated into (must be zeroed by caller) */
int p2mo_paint(const p2mo_run *run, const p2mo_cosmo *c, const p2mo_table *tb,
               const float *halos, int64_t n,
               double *hpmap, double *fsmap,
               uint32_t *hp_count, uint64_t *hp_hash,
               uint32_t *fs_count, uint64_t *fs_hash,
               p2mo_stats *st);

/* per-halo setup of the halo loop (pks2map.f90:144-160, 264, 284-298,
 * haloproject.f90:47): out[6*i] = (z, mh, rh, chi, dtheta, 1) for a painted
 * halo, zeros for a halo cut by radius or mass */
int p2mo_halo_setup(const p2mo_run *run, const p2mo_cosmo *c, const float *halos, int64_t n, double *out);

/* per-halo membership fingerprint term */
uint64_t p2mo_halo_hash(int64_t i);

/* make_maptable restatement (maptable.f90:25-146); table has 2*nrt*nmht*nchit
 * floats.  Nodes [node_lo, node_hi) of the flattened (i_chi, j_m) grid are
 * computed (node = (i-1)*nmht + (j-1)); pass 0 and nchit*nmht for all. */
int p2mo_make_maptable(int32_t model, double nu, double alpha_cmb,
                       int32_t nchit, int32_t nmht, int32_t nrt,
                       int64_t node_lo, int64_t node_hi,
                       int32_t hdr_i[3], float hdr_f[6], float *table, int32_t nthreads);

/* elementary functions as the oracle evaluates them (op codes: 0 sqrt, 1 sin,
 * 2 cos, 3 tan, 4 asin, 5 acos, 6 atan2(x,y), 7 log, 8 exp, 9 pow(x,y)) */
int p2mo_math(int op, const double *x, const double *y, double *out, int64_t n);

/* 48-bit LCG of RandomField/random.f90:40-57 */
double p2mo_ranf(int32_t seed[4]);

/* Radial cut with compaction (pks2map.f90:144-160), scramble (:162-187) and
 * centring on the largest halo (:189-262), in float arithmetic with the C
 * library's float functions, on 7-float records in place.  *n is updated to
 * the number kept.  Velocities are carried with their halo (the reference
 * compacts only positions and Rth; quirk Q18). */
void p2mo_host_branches(float *h, int64_t *n, float cut_low, float cut_high, float chihview,
                        int scramble, int centre);

/* ref32 mode (p2m_oracle_ref32.c): the reference's float32 arithmetic with
 * its mixed-mode promotions, on contiguous halo blocks per thread with
 * thread-private float maps summed in thread order (hpmap/fsmap are
 * overwritten; counts/hashes accumulate).  setup (6 doubles per input halo,
 * may be NULL) receives (z, mh, rh, chi, dtheta, 1) of each painted halo. */
int p2mo_paint_ref32(const p2mo_run *run, const int32_t hdr_i[3], const float hdr_f[6],
                     const float *padded, int64_t n_padded, const float *halos, int64_t n,
                     float *hpmap, float *fsmap, uint32_t *hp_count, uint64_t *hp_hash,
                     uint32_t *fs_count, uint64_t *fs_hash, double *setup, p2mo_stats *st);
/* the float32 distance tables of set_rofztable / set_zofrtable (10000 each) */
int p2mo_ref32_tables(double omegam, double omegal, double h, float *rofz, float *zofr);
/* interpolate_table in float32 (mh clamped in place) */
float p2mo_ref32_interpolate(const int32_t hdr_i[3], const float hdr_f[6], const float *padded, int64_t n_padded,
                             double omegam, double omegal, double h, float theta, float z, float *mh, int profile);

#ifdef __cplusplus
}
#endif
#endif
