/*
 * p2m_oracle.c -- CPU restatement of Peak Patch pks2map (TEST INFRASTRUCTURE).
 *
 * This file is the checker the HIP library is graded against.  It follows the
 * reference Fortran statement by statement (citations are file:line in the
 * reference tree).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * FP64 semantics: each Fortran `real` variable or default-real literal in an
 * executable statement is evaluated in double from its decimal text or from
 * the float32 input; named constants (Fortran `parameter`) keep the value the
 * reference gives them, including float rounding of real(4) literals.
 * Evaluation order follows Fortran left-to-right association; x**3 is (x*x)*x.
 *
 * Pinning: tests/test_oracle_*.py check every function here against closed
 * forms, brute force and invariants (see the docstrings of those tests).
 */
#include "p2m_oracle.h"
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "p2m_math.h"   /* p2m_floor/ceil/fabs/nint in both flavours */

#if P2MO_LIBM
#define M_SQRT_(x)   sqrt(x)
#define M_SIN_(x)    sin(x)
#define M_COS_(x)    cos(x)
#define M_TAN_(x)    tan(x)
#define M_ASIN_(x)   asin(x)
#define M_ACOS_(x)   acos(x)
#define M_ATAN_(x)   atan(x)
#define M_ATAN2_(y,x) atan2(y,x)
#define M_LOG_(x)    log(x)
#define M_EXP_(x)    exp(x)
#define M_POW_(x,y)  pow(x,y)
#else
#define M_SQRT_(x)   p2m_sqrt(x)
#define M_SIN_(x)    p2m_sin(x)
#define M_COS_(x)    p2m_cos(x)
#define M_TAN_(x)    p2m_tan(x)
#define M_ASIN_(x)   p2m_asin(x)
#define M_ACOS_(x)   p2m_acos(x)
#define M_ATAN_(x)   p2m_atan(x)
#define M_ATAN2_(y,x) p2m_atan2(y,x)
#define M_LOG_(x)    p2m_log(x)
#define M_EXP_(x)    p2m_exp(x)
#define M_POW_(x,y)  p2m_pow(x,y)
#endif

/* healpix_types.F90:65-73 */
static const double HPX_PI     = 3.141592653589793238462643383279502884197;
static const double HPX_HALFPI = 1.570796326794896619231321691639751442099;
static const double HPX_TWOPI  = 6.283185307179586476925286766559005768394;
static const double HPX_TWOTHIRD = 0.6666666666666666666666666666666666666666;

/* cosmology.f90:8-9: double parameter initialised from a real(4) literal (Q8) */
#define COSMO_PI ((double)3.14159265359f)
/* cosmology.f90:14,23: c and ckms = c/1e5 */
#define COSMO_C    2.99792458e10
#define COSMO_CKMS (COSMO_C / 1e5)

static inline double cube(double x) { return (x * x) * x; }
static inline int64_t fmodulo(int64_t a, int64_t n) { int64_t r = a % n; return r < 0 ? r + n : r; }

/* ===================================================================== */
/* cosmology.f90                                                          */
/* ===================================================================== */

/* dzdr, cosmology.f90:123-133: h*100/ckms*sqrt(omegam*(1.+z)**3+1.-omegam) */
static double dzdr(const p2mo_cosmo *c, double z)
{
    return ((c->h * 100.0) / COSMO_CKMS) * M_SQRT_((c->omegam * cube(1.0 + z) + 1.0) - c->omegam);
}

/* drdz, cosmology.f90:109-119 */
static double drdz(const p2mo_cosmo *c, double z) { return 1.0 / dzdr(c, z); }

int p2mo_cosmo_init(p2mo_cosmo *c, double omegam, double omegal, double h)
{
    c->omegam = omegam; c->omegal = omegal; c->h = h;
    /* pks2map.f90:127-128 */
    c->rho_0 = (2.775e11 * omegam) * (h * h);
    c->rhocrit_0 = 2.775e11 * (h * h);
    /* set_rofztable, cosmology.f90:51-76 */
    c->rofz_min = 0.0; c->rofz_max = 6.0; c->nrofz = 10000;
    c->drofz = (c->rofz_max - c->rofz_min) / (double)(c->nrofz - 1);
    c->rofz = (double *)malloc(sizeof(double) * c->nrofz);
    c->rofz[0] = 0.0;
    for (int i = 2; i <= c->nrofz; ++i) {
        double z = (double)(i - 1) * c->drofz + c->rofz_min;
        c->rofz[i - 1] = c->rofz[i - 2] + drdz(c, z) * c->drofz;
    }
    /* set_zofrtable, cosmology.f90:80-105 */
    c->zofr_min = 0.0; c->zofr_max = 1.3e4; c->nzofr = 10000;
    c->dzofr = (c->zofr_max - c->zofr_min) / (double)(c->nzofr - 1);
    c->zofr = (double *)malloc(sizeof(double) * c->nzofr);
    c->zofr[0] = 0.0;
    for (int i = 2; i <= c->nzofr; ++i) {
        double z = c->zofr[i - 2];
        c->zofr[i - 1] = c->zofr[i - 2] + dzdr(c, z) * c->dzofr;
    }
    return 0;
}

void p2mo_cosmo_free(p2mo_cosmo *c)
{
    free(c->rofz); free(c->zofr); c->rofz = c->zofr = NULL;
}

/* table lookup shared by zofr (cosmology.f90:137-157) and rofz (:161-181) */
static double tab_lookup(const double *t, int n, double tmin, double dt, double x)
{
    if (x != x) return x;
    double u = (x - tmin) / dt + 1.0;          /* itab = ((r-min)/d)+1, truncated */
    if (u < 1.0) return t[0];                   /* itab < 1 */
    if (u >= (double)n) return t[n - 1];        /* itab >= n */
    int itab = (int)u;
    double fint = ((x - (double)(itab - 1) * dt) - tmin) / dt;
    return (1.0 - fint) * t[itab - 1] + fint * t[itab];
}

double p2mo_zofr(const p2mo_cosmo *c, double r) { return tab_lookup(c->zofr, c->nzofr, c->zofr_min, c->dzofr, r); }
double p2mo_rofz(const p2mo_cosmo *c, double z) { return tab_lookup(c->rofz, c->nrofz, c->rofz_min, c->drofz, z); }

/* rhocrit, cosmology.f90:185-195 */
double p2mo_rhocrit(const p2mo_cosmo *c, double z)
{
    double z3 = cube(1.0 + z);
    return (c->rhocrit_0 * (c->omegam * z3 + c->omegal)) / z3;
}

/* deltacrit, cosmology.f90:199-214 */
double p2mo_deltacrit(const p2mo_cosmo *c, double z)
{
    double z3 = cube(1.0 + z);
    double omegaz = (c->omegam * z3) / ((c->omegam * z3 + 1.0) - c->omegam);
    double x = omegaz - 1.0;
    return (18.0 * (COSMO_PI * COSMO_PI) + 82.0 * x) - 39.0 * (x * x);
}

/* bbps_profile.f90:21 */
static double bbps_delta(int model) { return model == 2 ? 500.0 : 200.0; }

/* ===================================================================== */
/* maptable.f90                                                           */
/* ===================================================================== */

int p2mo_table_init(p2mo_table *tb, const int32_t hdr_i[3], const float hdr_f[6], const float *padded, int64_t n_padded)
{
    /* maptable.f90:158 header order: nchit,nmht,nrt,chimint,chimaxt,mhmint,mhmaxt,rmint,rmaxt */
    tb->nchit = hdr_i[0]; tb->nmht = hdr_i[1]; tb->nrt = hdr_i[2];
    tb->chimint = hdr_f[0]; tb->chimaxt = hdr_f[1];
    tb->mhmint = hdr_f[2];  tb->mhmaxt = hdr_f[3];
    tb->rmint = hdr_f[4];   tb->rmaxt = hdr_f[5];
    /* maptable.f90:166-168 */
    tb->dchit = (M_LOG_(tb->chimaxt) - M_LOG_(tb->chimint)) / (double)(tb->nchit - 1);
    tb->drt   = (M_LOG_(tb->rmaxt) - M_LOG_(tb->rmint)) / (double)(tb->nrt - 1);
    tb->dmht  = (M_LOG_(tb->mhmaxt) - M_LOG_(tb->mhmint)) / (double)(tb->nmht - 1);
    tb->t = padded; tb->nt = n_padded;
    return 0;
}

/* Flat column-major index of table(p, ir, imh, ichi) (1-based), used beyond
 * the declared bounds exactly as Fortran without bounds checks would (Q6);
 * elements past the array end read as zero. */
static inline double tval(const p2mo_table *tb, int p, int ir, int imh, int ichi)
{
    int64_t idx = (int64_t)(p - 1) + 2 * ((int64_t)(ir - 1) + (int64_t)tb->nrt * ((int64_t)(imh - 1) + (int64_t)tb->nmht * (int64_t)(ichi - 1)));
    if (idx < 0 || idx >= tb->nt) return 0.0;
    return (double)tb->t[idx];
}

/* interpolate_table, maptable.f90:176-216.  mh is clamped in place (Q3). */
double p2mo_interpolate_table(const p2mo_table *tb, const p2mo_cosmo *c, double theta, double z, double *mh, int profile)
{
    double chi = p2mo_rofz(c, z);
    double r = chi * M_TAN_(theta);
    if (r > tb->rmaxt || chi > tb->chimaxt) return 0.0;
    if (r < tb->rmint) r = tb->rmint + 1e-5;
    if (chi < tb->chimint) chi = tb->chimint + 1e-5;
    if (*mh < tb->mhmint) *mh = tb->mhmint + 1e-5;
    if (*mh > tb->mhmaxt) *mh = tb->mhmaxt - 1e-5;
    double lr = M_LOG_(r), lc = M_LOG_(chi), lm = M_LOG_(*mh);
    double lr0 = M_LOG_(tb->rmint), lc0 = M_LOG_(tb->chimint), lm0 = M_LOG_(tb->mhmint);
    int ir   = (int)((lr - lr0) / tb->drt) + 1;
    int ichi = (int)((lc - lc0) / tb->dchit) + 1;
    int imh  = (int)((lm - lm0) / tb->dmht) + 1;
    double fr = lr - (lr0 + (double)(ir - 1) * tb->drt);
    double fc = lc - (lc0 + (double)(ichi - 1) * tb->dchit);
    double fm = lm - (lm0 + (double)(imh - 1) * tb->dmht);
    /* maptable.f90:204-212: imh+1 corners carry fc, ichi+1 corners carry fm (Q4) */
    double v = tval(tb, profile, ir,     imh,     ichi)     * (1.0 - fr) * (1.0 - fc) * (1.0 - fm);
    v = v + tval(tb, profile, ir + 1, imh,     ichi)     * (fr)       * (1.0 - fc) * (1.0 - fm);
    v = v + tval(tb, profile, ir,     imh + 1, ichi)     * (1.0 - fr) * (fc)       * (1.0 - fm);
    v = v + tval(tb, profile, ir + 1, imh + 1, ichi)     * (fr)       * (fc)       * (1.0 - fm);
    v = v + tval(tb, profile, ir,     imh,     ichi + 1) * (1.0 - fr) * (1.0 - fc) * (fm);
    v = v + tval(tb, profile, ir + 1, imh,     ichi + 1) * (fr)       * (1.0 - fc) * (fm);
    v = v + tval(tb, profile, ir,     imh + 1, ichi + 1) * (1.0 - fr) * (fc)       * (fm);
    v = v + tval(tb, profile, ir + 1, imh + 1, ichi + 1) * (fr)       * (fc)       * (fm);
    return v;
}

/* ===================================================================== */
/* HEALPix 3.10 (vendored): pix_tools.F90, pixel_routines.F90             */
/* ===================================================================== */

/* nside2npix, pix_tools.F90:1946-1969 */
int64_t p2mo_nside2npix(int32_t nside)
{
    if (nside < 1 || nside > 536870912 || ((nside - 1) & nside) != 0) return -1;
    return (12LL * nside) * nside;
}

/* ring_num, pix_tools.F90:490-533 (shift = 0) */
int32_t p2mo_ring_num(int32_t nside, double z)
{
    int32_t iring = (int32_t)p2m_nint((double)nside * (2.0 - 1.5 * z));
    if (z > HPX_TWOTHIRD) {
        iring = (int32_t)p2m_nint((double)nside * M_SQRT_(3.0 * (1.0 - z)));
        if (iring == 0) iring = 1;
    }
    if (z < -HPX_TWOTHIRD) {
        iring = (int32_t)p2m_nint((double)nside * M_SQRT_(3.0 * (1.0 + z)));
        if (iring == 0) iring = 1;
        iring = 4 * nside - iring;
    }
    return iring;
}

/* ring2z, pix_tools.F90:535-560 */
double p2mo_ring2z(int32_t nside, int32_t ir)
{
    double fn = (double)nside;
    if (ir < nside) {
        double tmp = (double)ir;
        return 1.0 - (tmp * tmp) / ((3.0 * fn) * fn);
    } else if (ir < 3 * nside) {
        return ((double)(2 * nside - ir) * 2.0) / (3.0 * fn);
    } else {
        double tmp = (double)(4 * nside - ir);
        return -1.0 + (tmp * tmp) / ((3.0 * fn) * fn);
    }
}

/* pixels_per_ring, pix_tools.F90:802-836 */
static void pixels_per_ring(int32_t nside, int32_t ring, int32_t *npr, int32_t *kshift, int64_t *npnorth)
{
    int32_t m = nside < ring ? nside : ring;
    if (4 * nside - ring < m) m = 4 * nside - ring;
    *npr = 4 * m;
    *kshift = (ring + 1) % 2;
    if (nside == 1) *kshift = 1 - *kshift;
    if (*npr < 4 * nside) *kshift = 1;
    if (npnorth) {
        if (ring <= nside) {
            *npnorth = (int64_t)ring * (ring + 1LL) * 2LL;
        } else if (ring <= 3 * nside) {
            int64_t ncap = (int64_t)nside * (nside + 1LL) * 2LL;
            int64_t ir = ring - nside;
            *npnorth = ncap + 4LL * nside * ir;
        } else {
            int64_t npix = p2mo_nside2npix(nside);
            int64_t ir = 4LL * nside - ring - 1;
            *npnorth = npix - ir * (ir + 1LL) * 2LL;
        }
    }
}

/* pix2vec_ring, pixel_routines.F90:167-359 (I4B variant; no vertices) */
void p2mo_pix2vec_ring(int32_t nside, int64_t ipix, double v[3])
{
    int64_t npix = p2mo_nside2npix(nside);
    int64_t nl2 = 2LL * nside;
    int64_t ncap = nl2 * (nside - 1LL);
    const double half = 0.5;
    double z, sth, phi;
    if (ipix < ncap) {
        int64_t iring = ((int64_t)p2m_floor(M_SQRT_((double)(2 * ipix + 2))) + 1) / 2;   /* cheap_isqrt_4 */
        int64_t iphi = ipix - 2 * iring * (iring - 1);
        double fact2 = (3.0 * nside) * nside;
        z = 1.0 - ((double)iring / fact2) * (double)iring;
        sth = (double)iring * M_SQRT_((1.0 + z) / fact2);
        phi = (((double)iphi + half) * HPX_HALFPI) / (double)iring;
    } else if (ipix < npix - ncap) {
        int64_t nl4 = 4LL * nside;
        int64_t ip = ipix - ncap;
        int64_t iring = ip / nl4 + nside;
        int64_t iphi = ip & (nl4 - 1);
        double fact1 = 1.5 * nside;
        double fodd = half * (double)((iring + nside + 1) & 1);
        z = (double)(nl2 - iring) / fact1;
        sth = M_SQRT_((1.0 - z) * (1.0 + z));
        phi = (((double)iphi + fodd) * HPX_HALFPI) / (double)nside;
    } else {
        int64_t ip = npix - ipix;
        int64_t iring = ((int64_t)p2m_floor(M_SQRT_((double)(2 * ip))) + 1) / 2;
        int64_t iphi = 2 * iring * (iring + 1) - ip;
        double fact2 = (3.0 * nside) * nside;
        z = -1.0 + ((double)iring / fact2) * (double)iring;
        sth = (double)iring * M_SQRT_((1.0 - z) / fact2);
        phi = (((double)iphi + half) * HPX_HALFPI) / (double)iring;
    }
    double cphi = M_COS_(phi), sphi = M_SIN_(phi);
    v[0] = sth * cphi;
    v[1] = sth * sphi;
    v[2] = z;
}

/* angdist + vect_prod, pix_tools.F90:2015-2055 */
double p2mo_angdist(const double v1[3], const double v2[3])
{
    double sprod = (v1[0] * v2[0] + v1[1] * v2[1]) + v1[2] * v2[2];
    double v3[3];
    v3[0] = v1[1] * v2[2] - v1[2] * v2[1];
    v3[1] = v1[2] * v2[0] - v1[0] * v2[2];
    v3[2] = v1[0] * v2[1] - v1[1] * v2[0];
    double vprod = M_SQRT_((v3[0] * v3[0] + v3[1] * v3[1]) + v3[2] * v3[2]);
    return M_ATAN2_(vprod, sprod);
}

/* query_disc (RING, non-inclusive), pixel_routines.F90:1924-2046, with
 * discphirange_at_z (pix_tools.F90:712-760), pixels_on_edge (:762-800) and
 * discedge2fulldisc (pixel_routines.F90:1823-1879).  Rings with no pixel
 * centre inside the disc between non-empty rings (Q15) are counted in n_gaps;
 * each ring then uses its own range (the reference would misalign rows). */
int64_t p2mo_query_disc(int32_t nside, const double v0[3], double radius, int32_t *list, int64_t *n_gaps)
{
    if (!(radius >= 0.0 && radius <= HPX_PI)) return -1;
    double norm = M_SQRT_((v0[0] * v0[0] + v0[1] * v0[1]) + v0[2] * v0[2]);
    double z0 = v0[2] / norm;
    double rlat0 = M_ASIN_(z0);
    double rlat1 = rlat0 + radius;
    double rlat2 = rlat0 - radius;
    double zmax = (rlat1 >= HPX_HALFPI) ? 1.0 : M_SIN_(rlat1);
    int32_t irmin = p2mo_ring_num(nside, zmax);
    irmin = irmin - 1 > 1 ? irmin - 1 : 1;
    double zmin = (rlat2 <= -HPX_HALFPI) ? -1.0 : M_SIN_(rlat2);
    int32_t irmax = p2mo_ring_num(nside, zmin);
    irmax = irmax + 1 < 4 * nside - 1 ? irmax + 1 : 4 * nside - 1;

    /* discphirange_at_z */
    double cosang = M_COS_(radius);
    double x0 = v0[0] / norm, y0 = v0[1] / norm, zz0 = v0[2] / norm;
    double phi0 = 0.0;
    if (x0 != 0.0 || y0 != 0.0) phi0 = M_ATAN2_(y0, x0);
    double a = x0 * x0 + y0 * y0;

    int64_t nlist = 0, gaps = 0;
    int32_t last_nonempty = -1;
    int pending_gap = 0;
    for (int32_t iz = irmin; iz <= irmax; ++iz) {
        double z = p2mo_ring2z(nside, iz);
        double dphi = -1000.0;
        double b = cosang - z * zz0;
        if (a == 0.0) {
            if (b <= 0.0) dphi = HPX_PI;
        } else {
            double cc = 1.0 - z * z;
            if (cc < 1.0e-12) cc = 1.0e-12;
            double cosdphi = b / M_SQRT_(a * cc);
            if (cosdphi < -1.0) dphi = HPX_PI;
            if (p2m_fabs(cosdphi) <= 1.0) dphi = M_ACOS_(cosdphi);
        }
        /* pixels_on_edge */
        int32_t npr, kshift; int64_t npc;
        pixels_per_ring(nside, iz, &npr, &kshift, &npc);
        int64_t lo, hi; int has = 0;
        if (dphi >= HPX_PI) {
            lo = 0; hi = npr - 1; has = 1;
        } else if (dphi >= 0.0) {
            double shift = (double)kshift * 0.5;
            int64_t il = (int64_t)p2m_ceil(((double)npr * (phi0 - dphi)) / HPX_TWOPI - shift);
            int64_t ih = (int64_t)p2m_floor(((double)npr * (phi0 + dphi)) / HPX_TWOPI - shift);
            if (ih >= il) { lo = fmodulo(il, npr); hi = fmodulo(ih, npr); has = 1; }
        }
        if (!has) { if (last_nonempty >= 0) pending_gap = 1; continue; }
        if (pending_gap) { gaps++; pending_gap = 0; }
        last_nonempty = iz;
        /* discedge2fulldisc */
        int64_t np = hi - lo;
        np = fmodulo(np, npr) + 1;
        if (np > npr) np = npr;
        if (list) {
            for (int64_t i = 0; i < np; ++i) {
                int64_t ip = fmodulo(lo + i, npr);
                list[nlist + i] = (int32_t)(npc - npr + ip);
            }
        }
        nlist += np;
    }
    if (n_gaps) *n_gaps += gaps;
    return nlist;
}

/* ===================================================================== */
/* painter: pks2map.f90:144-160, 264, 282-303; haloproject.f90:38-156     */
/* ===================================================================== */

uint64_t p2mo_halo_hash(int64_t i)
{
    /* splitmix64 finaliser of the catalogue index */
    uint64_t z = (uint64_t)i + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline void add_d(double *p, double v)
{
#pragma omp atomic
    *p += v;
}
static inline void add_u32(uint32_t *p, uint32_t v)
{
#pragma omp atomic
    *p += v;
}
static inline void add_u64(uint64_t *p, uint64_t v)
{
#pragma omp atomic
    *p += v;
}

/* Per-halo part of the halo loop (pks2map.f90:144-160, 264, 284-298 and
 * haloproject.f90:47,97-99): returns 0 if the halo fails the radial cut, 1 if
 * it passes the cut but not the mass cut, 2 if it is painted (then the
 * outputs are set). */
typedef struct {
    double redshift, mh, rh, chih, dtheta, vh;
    double ux, uy, uz;
} halo_setup;

static int setup_one(const p2mo_run *run, const p2mo_cosmo *c, const float *h, double cut_low, double cut_high,
                     double mypi, halo_setup *o)
{
    const double dmodel = bbps_delta(run->model);
    const double delta_pksc = 200.0;                                /* pksc.f90:8-9 */
    const double rmax = 4.0;                                        /* haloproject.f90:18 */
    /* radial cut, pks2map.f90:144-160 (painter axes: x=pos3, y=pos2, z=pos1) */
    double xh = (double)h[2], yh = (double)h[1], zh = (double)h[0];
    xh = xh - run->chihview;
    double chih = M_SQRT_((xh * xh + yh * yh) + zh * zh);
    if (chih > cut_high || chih < cut_low) return 0;
    /* mass, pks2map.f90:264 */
    double rth = (double)h[6];
    double mass = (((4.0 / 3.0) * mypi) * c->rho_0) * cube(rth);
    /* velocity factor: profile 1 -> 1 (pksc.f90:293-294); profile 2 -> -v_r/c (intended kSZ, Q18) */
    double vh = 1.0;
    if (run->profile == 2) {
        double px = h[0], py = h[1], pz = h[2];
        double vr = (px * (double)h[3] + py * (double)h[4]) + pz * (double)h[5];
        vr = vr / M_SQRT_((px * px + py * py) + pz * pz);
        vh = -vr / 3.0e5;
    }
    /* halo loop body, pks2map.f90:284-298 */
    xh = (double)h[2]; yh = (double)h[1]; zh = (double)h[0];
    chih = M_SQRT_((xh * xh + yh * yh) + zh * zh);
    double redshift = p2mo_zofr(c, chih);
    xh = xh - run->chihview;
    chih = M_SQRT_((xh * xh + yh * yh) + zh * zh);
    double mh = M_SQRT_(p2mo_deltacrit(c, redshift) / dmodel) * mass;
    if (mass < run->mmin) return 1;
    double rh = M_POW_(((((3.0 * mh) / 4.0) / mypi) / delta_pksc) / p2mo_rhocrit(c, redshift), 1.0 / 3.0);
    /* both painters: dtheta_halo = rmax*asin(r/chi) (haloproject.f90:47,93) */
    o->dtheta = rmax * M_ASIN_(rh / chih);
    o->ux = xh / chih; o->uy = yh / chih; o->uz = zh / chih;          /* haloproject.f90:97-99 */
    o->redshift = redshift; o->mh = mh; o->rh = rh; o->chih = chih; o->vh = vh;
    return 2;
}

int p2mo_halo_setup(const p2mo_run *run, const p2mo_cosmo *c, const float *halos, int64_t n, double *out)
{
    const double mypi = 4.0 * M_ATAN_(1.0);
    const double cut_low = p2mo_rofz(c, run->zmin), cut_high = p2mo_rofz(c, run->zmax);
    for (int64_t i = 0; i < n; ++i) {
        halo_setup hs;
        const int k = setup_one(run, c, halos + 7 * i, cut_low, cut_high, mypi, &hs);
        double *o = out + 6 * i;
        if (k == 2) { o[0] = hs.redshift; o[1] = hs.mh; o[2] = hs.rh; o[3] = hs.chih; o[4] = hs.dtheta; o[5] = 1.0; }
        else { o[0] = o[1] = o[2] = o[3] = o[4] = 0.0; o[5] = 0.0; }
    }
    return 0;
}

int p2mo_paint(const p2mo_run *run, const p2mo_cosmo *c, const p2mo_table *tb,
               const float *halos, int64_t n,
               double *hpmap, double *fsmap,
               uint32_t *hp_count, uint64_t *hp_hash,
               uint32_t *fs_count, uint64_t *fs_hash,
               p2mo_stats *st)
{
    const double mypi = 4.0 * M_ATAN_(1.0);                         /* pks2map.f90:88 */
    const double fov = ((run->fov_deg / 360.0) * 2.0) * mypi;       /* pks2map.f90:92 */
    const double dpix = run->npix > 0 ? fov / (double)run->npix : 1.0; /* :93 */
    const double cut_low = p2mo_rofz(c, run->zmin);                 /* :134 */
    const double cut_high = p2mo_rofz(c, run->zmax);                /* :135 */
    const int32_t nside = run->nside;
    const int32_t npix = run->npix;

    int64_t n_cut = 0, n_painted = 0, n_list = 0, n_upd_hp = 0, n_upd_fs = 0, n_gaps = 0;
    int nth = run->nthreads > 0 ? run->nthreads : 1;

#pragma omp parallel num_threads(nth) reduction(+:n_cut,n_painted,n_list,n_upd_hp,n_upd_fs,n_gaps)
    {
        int32_t *list = NULL;
        int64_t list_cap = 0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < n; ++i) {
            halo_setup hs;
            const int kind = setup_one(run, c, halos + 7 * i, cut_low, cut_high, mypi, &hs);
            if (kind == 0) continue;
            n_cut++;
            if (kind == 1) continue;
            n_painted++;
            const double dtheta = hs.dtheta, redshift = hs.redshift, vh = hs.vh;
            double mh = hs.mh;
            const double ux = hs.ux, uy = hs.uy, uz = hs.uz;
            uint64_t hh = p2mo_halo_hash(i);

            /* ---- projecthalo_flatsky, haloproject.f90:83-156 ---- */
            if (npix > 0 && uz >= 0.0) {
                double th = M_ASIN_(ux), ph = M_ASIN_(uy);
                double thetamin = th - 1.1 * dtheta, thetamax = th + 1.1 * dtheta;
                double phimin = ph - 1.1 * dtheta, phimax = ph + 1.1 * dtheta;
                double hf = fov / 2.0;
                if (!(thetamax < -hf || thetamin > hf) && !(phimax < -hf || phimin > hf)) {
                    if (thetamin <= -hf) thetamin = -hf + 1e-10;
                    if (thetamax >= hf) thetamax = hf - 1e-10;
                    if (phimin <= -hf) phimin = -hf + 1e-10;
                    if (phimax >= hf) phimax = hf - 1e-10;
                    int jmin = (int)((thetamin + hf) / dpix) + 1; if (jmin < 1) jmin = 1;
                    int jmax = (int)((thetamax + hf) / dpix) + 1; if (jmax > npix) jmax = npix;
                    int kmin = (int)((phimin + hf) / dpix) + 1; if (kmin < 1) kmin = 1;
                    int kmax = (int)((phimax + hf) / dpix) + 1; if (kmax > npix) kmax = npix;
                    for (int k = kmin; k <= kmax; ++k) {
                        double phip = -hf + ((double)k - 0.5) * dpix;
                        for (int j = jmin; j <= jmax; ++j) {
                            double thetap = -hf + ((double)j - 0.5) * dpix;
                            double dphi = phip - ph;
                            double dth = thetap - th;
                            double angle = M_SQRT_(dphi * dphi + dth * dth);
                            if (angle < dtheta) {
                                double val = vh * p2mo_interpolate_table(tb, c, angle, redshift, &mh, run->profile);
                                int64_t pix = (int64_t)(j - 1) + (int64_t)npix * (k - 1);
                                if (fsmap) add_d(&fsmap[pix], val);
                                if (fs_count) add_u32(&fs_count[pix], 1u);
                                if (fs_hash) add_u64(&fs_hash[pix], hh);
                                n_upd_fs++;
                            }
                        }
                    }
                }
            }

            /* ---- projecthalo_healpix, haloproject.f90:38-79 ---- */
            if (nside > 0) {
                double v1[3] = { uz, uy, ux };
                /* count first, then fill a list of that size (the reference
                 * allocates a full-sky list, pks2map.f90:115, quirk Q19) */
                int64_t nl = p2mo_query_disc(nside, v1, dtheta, NULL, NULL);
                if (nl < 0) continue;
                if (nl > list_cap) {
                    free(list);
                    list_cap = nl + 1024;
                    list = (int32_t *)malloc(sizeof(int32_t) * (size_t)list_cap);
                }
                nl = p2mo_query_disc(nside, v1, dtheta, list, &n_gaps);
                n_list += nl;
                for (int64_t q = 0; q < nl; ++q) {
                    int32_t jp = list[q];
                    double v2[3];
                    p2mo_pix2vec_ring(nside, jp, v2);
                    double angle = p2mo_angdist(v1, v2);
                    if (angle < dtheta) {
                        double val = vh * p2mo_interpolate_table(tb, c, angle, redshift, &mh, run->profile);
                        if (hpmap) add_d(&hpmap[jp], val);
                        if (hp_count) add_u32(&hp_count[jp], 1u);
                        if (hp_hash) add_u64(&hp_hash[jp], hh);
                        n_upd_hp++;
                    }
                }
            }
        }
        free(list);
    }
    if (st) {
        st->n_in = n; st->n_cut = n_cut; st->n_painted = n_painted;
        st->n_listed_hp = n_list; st->n_upd_hp = n_upd_hp; st->n_upd_fs = n_upd_fs;
        st->n_ring_gaps = n_gaps; st->cut_low = cut_low; st->cut_high = cut_high;
    }
    return 0;
}

/* ===================================================================== */
/* make_maptable: maptable.f90:25-146, integrate_profiles.f90:12-95,     */
/* bbps_profile.f90:9-180, make_maptable.f90:33-46                        */
/* ===================================================================== */

/* bbps_profile.f90:9-40: real(4) parameter arrays (float-rounded values) */
static const float P0_Am[3] = {18.100f, 7.490f, 20.7f};
static const float P0_alpham[3] = {0.154f, 0.226f, -0.074f};
static const float P0_alphaz[3] = {-0.758f, -0.957f, -0.743f};
static const float xc_Am[3] = {0.497f, 0.710f, 0.438f};
static const float xc_alpham[3] = {-0.00865f, -0.0833f, 0.011f};
static const float xc_alphaz[3] = {0.731f, 0.853f, 1.01f};
static const float beta_Am[3] = {4.35f, 4.19f, 3.82f};
static const float beta_alpham[3] = {0.0393f, 0.048f, 0.0375f};
static const float beta_alphaz[3] = {0.415f, 0.615f, 0.535f};
static const float P0_Am_rho[3] = {1.9e-5f, 9.1e-05f, 1.5e-4f};
static const float P0_alpham_rho[3] = {0.29f, 0.10f, 0.14f};
static const float P0_alphaz_rho[3] = {-0.67f, -0.92f, -1.32f};
static const float alpha_Am_rho[3] = {0.88f, 0.69f, 0.68f};
static const float alpha_alpham_rho[3] = {-0.03f, -0.02f, -0.02f};
static const float alpha_alphaz_rho[3] = {0.20f, 0.27f, 0.29f};
static const float beta_Am_rho[3] = {3.83f, 4.43f, 6.4f};
static const float beta_alpham_rho[3] = {0.044f, 0.0035f, 0.028f};
static const float beta_alphaz_rho[3] = {-0.025f, 0.037f, -0.25f};
#define BBPS_GAMMA     ((double)-0.3f)
#define BBPS_XC_RHO    0.5
#define BBPS_GAMMA_RHO ((double)0.2f)
#define IP_DS          ((double)1e-2f)   /* integrate_profiles.f90:8 */

/* cosmology.f90:11-38 physical constants (double parameters) */
#define K_SIGMAT 6.65246e-25
#define K_ME     9.109383e-28
#define K_MUE    ((double)1.136f)
#define K_MP     ((double)1.67e-24f)
#define K_MSUN   1.9891e33
#define K_MPC2KM 3.08567758e19
#define K_MPC    3.08567758e24
#define K_KPC    3.08567758e21
#define K_H100   (100.0 / K_MPC2KM)

static double powk(double x, double y) { return y == 1.0 ? x : M_POW_(x, y); }

/* A0*(mh/1e14)**alpham*(1+z)**alphaz, bbps_profile.f90:91 */
static double bbps_fit(double A0, double alpham, double alphaz, double mh, double z)
{
    return (A0 * M_POW_(mh / 1e14, alpham)) * M_POW_(1.0 + z, alphaz);
}

int p2mo_make_maptable(int32_t model, double nu, double alpha_cmb,
                       int32_t nchit, int32_t nmht, int32_t nrt,
                       int64_t node_lo, int64_t node_hi,
                       int32_t hdr_i[3], float hdr_f[6], float *table, int32_t nthreads)
{
    /* make_maptable.f90:33-43 cosmology */
    const double omegam = 0.25, omegab = 0.043, omegal = 0.75, h = 0.7;
    const double fb = omegab / omegam;
    p2mo_cosmo c;
    p2mo_cosmo_init(&c, omegam, omegal, h);
    c.rho_0 = (2.775e11 * omegam) * (0.7 * 0.7);
    c.rhocrit_0 = 2.775e11 * (0.7 * 0.7);
    const int mi = model - 1;
    /* maptable.f90:56-64 */
    const double Delta = bbps_delta(model);
    const double m2rfac = M_POW_(((3.0 / 4.0) / COSMO_PI) / Delta, 1.0 / 3.0);
    const double ypermass = ((((((3.0 * K_SIGMAT) * (K_H100 * K_H100)) / 8.0) / COSMO_PI) / K_ME) / (COSMO_C * COSMO_C)) * K_MSUN / (double)1.932f;
    const double taupersigma = ((((K_SIGMAT / K_MUE) / K_MP) * K_MSUN) / (K_KPC * K_KPC * K_KPC)) * K_MPC * 1e10;
    const double y0 = ((fb * Delta) * (h * h)) * ypermass;
    const double tau0 = (taupersigma * (h * h)) * m2rfac;
    const double x0const = 0.017612 * nu;
    /* grid, maptable.f90:69-87; bounds are the float32 values written to the header */
    const double chimint = (double)1e2f, chimaxt = (double)7e3f;
    const double mhmint = (double)1e13f, mhmaxt = (double)1e16f;
    const double rmint = (double)1e-2f, rmaxt = (double)10.0f;
    const double dchit = (M_LOG_(chimaxt) - M_LOG_(chimint)) / (double)(nchit - 1);
    const double dmht = (M_LOG_(mhmaxt) - M_LOG_(mhmint)) / (double)(nmht - 1);
    const double drt = (M_LOG_(rmaxt) - M_LOG_(rmint)) / (double)(nrt - 1);
    hdr_i[0] = nchit; hdr_i[1] = nmht; hdr_i[2] = nrt;
    hdr_f[0] = (float)chimint; hdr_f[1] = (float)chimaxt; hdr_f[2] = (float)mhmint;
    hdr_f[3] = (float)mhmaxt; hdr_f[4] = (float)rmint; hdr_f[5] = (float)rmaxt;
    const double rmaxp = 4.0;    /* integrate_profiles.f90:20 */
    int nth = nthreads > 0 ? nthreads : 1;

#pragma omp parallel for num_threads(nth) schedule(dynamic, 1)
    for (int64_t node = node_lo; node < node_hi; ++node) {
        int i = (int)(node / nmht) + 1, j = (int)(node % nmht) + 1;
        double chit = M_EXP_(M_LOG_(chimint) + (double)(i - 1) * dchit);
        double z = p2mo_zofr(&c, chit);
        double zfacy = omegam * cube(1.0 + z) + omegal;
        double zfact = (1.0 / (1.0 + z)) / M_POW_(p2mo_rhocrit(&c, z), 1.0 / 3.0);
        double Gx = 0.0;
        if (nu > 0.0) {
            double x0new = x0const * M_POW_(1.0 + z, alpha_cmb);
            double ex = M_EXP_(x0new);
            Gx = ((x0new * (ex + 1.0)) / (ex - 1.0)) - 4.0;
        }
        double mht = M_EXP_(M_LOG_(mhmint) + (double)(j - 1) * dmht);
        double mht3 = M_POW_(mht, 1.0 / 3.0);
        /* per-(mh,z) constants of bbps_ptilde / bbps_rhotilde (pure functions of mh, z) */
        double P0 = bbps_fit(P0_Am[mi], P0_alpham[mi], P0_alphaz[mi], mht, z);
        double xc = bbps_fit(xc_Am[mi], xc_alpham[mi], xc_alphaz[mi], mht, z);
        double beta = bbps_fit(beta_Am[mi], beta_alpham[mi], beta_alphaz[mi], mht, z);
        double P0r = bbps_fit(P0_Am_rho[mi], P0_alpham_rho[mi], P0_alphaz_rho[mi], mht, z);
        double alr = bbps_fit(alpha_Am_rho[mi], alpha_alpham_rho[mi], alpha_alphaz_rho[mi], mht, z);
        double betar = bbps_fit(beta_Am_rho[mi], beta_alpham_rho[mi], beta_alphaz_rho[mi], mht, z);
        /* integrate_*: chi = rofz(z); rvir = (3*mh/4/pi/Delta/rhocrit(z))**(1./3.) */
        double chi = p2mo_rofz(&c, z);
        double rvir = M_POW_(((((3.0 * mht) / 4.0) / COSMO_PI) / Delta) / p2mo_rhocrit(&c, z), 1.0 / 3.0);
        for (int k = 1; k <= nrt; ++k) {
            double rt = M_EXP_(M_LOG_(rmint) + (double)(k - 1) * drt);
            double theta = M_ATAN_(rt / chit);
            double b = (chi * M_SIN_(theta)) / rvir;
            /* integrate_ptilde (cylindrical cut, s0 = rmax) */
            double yint = 0.0;
            if (!(b > rmaxp)) {
                int nn = (int)(rmaxp / IP_DS) + 1;
                double y = 0.0;
                for (int q = 1; q <= nn; ++q) {
                    double s = ((double)q - 0.5) * IP_DS;
                    double x = M_SQRT_(s * s + b * b);
                    double xx = x / xc;
                    /* bbps_ptilde: P0*(x/xc)**gamma*(1+(x/xc)**alpha)**-beta, alpha = 1 */
                    y = y + (P0 * M_POW_(xx, BBPS_GAMMA)) * M_POW_(1.0 + powk(xx, 1.0), -beta);
                }
                yint = y * IP_DS;
            }
            /* integrate_rhotilde (spherical cut, s0 = sqrt(rmax**2-b**2)) */
            double rint = 0.0;
            if (!(b > rmaxp)) {
                double s0 = M_SQRT_(rmaxp * rmaxp - b * b);
                int nn = (int)(s0 / IP_DS) + 1;
                double rho = 0.0;
                for (int q = 1; q <= nn; ++q) {
                    double s = ((double)q - 0.5) * IP_DS;
                    double x = M_SQRT_(s * s + b * b);
                    double xx = x / BBPS_XC_RHO;
                    rho = rho + (P0r * M_POW_(xx, -1.0 * BBPS_GAMMA_RHO)) *
                                M_POW_(1.0 + M_POW_(xx, alr), (-1.0 * (betar - BBPS_GAMMA_RHO)) / alr);
                }
                rint = 2.0 * (rho * IP_DS);
            }
            double y = ((y0 * mht) * zfacy) * yint;
            if (nu > 0.0) y = y * Gx;
            double tau = ((tau0 * mht3) * zfact) * rint;
            int64_t idx = 2 * ((int64_t)(k - 1) + (int64_t)nrt * ((int64_t)(j - 1) + (int64_t)nmht * (int64_t)(i - 1)));
            table[idx] = (float)y;
            table[idx + 1] = (float)tau;
        }
    }
    p2mo_cosmo_free(&c);
    return 0;
}

/* ranf + ranfmodmult, RandomField/random.f90:40-57, 264-286 */
double p2mo_ranf(int32_t seed[4])
{
    static const int32_t mult[4] = {373, 3707, 1442, 647};
    /* float(Seed(k)) is real(4); the divisors are exact powers of two */
    double r = (double)(float)seed[3] / 4096.0 + (double)(float)seed[2] / 16777216.0
             + (double)(float)seed[1] / 68719476736.0 + (double)(float)seed[0] / 281474976710656.0;
    int32_t j1 = mult[0] * seed[0];
    int32_t j2 = mult[0] * seed[1] + mult[1] * seed[0];
    int32_t j3 = mult[0] * seed[2] + mult[1] * seed[1] + mult[2] * seed[0];
    int32_t j4 = mult[0] * seed[3] + mult[1] * seed[2] + mult[2] * seed[1] + mult[3] * seed[0];
    int32_t k1 = j1, k2 = j2 + k1 / 4096, k3 = j3 + k2 / 4096, k4 = j4 + k3 / 4096;
    seed[0] = k1 % 4096; seed[1] = k2 % 4096; seed[2] = k3 % 4096; seed[3] = k4 % 4096;
    return r;
}

/* host evaluation of the elementary functions the oracle uses (same op codes
 * as the library's p2m_selftest_math) */
int p2mo_math(int op, const double *x, const double *y, double *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) {
        double a = x[i], b = y ? y[i] : 0.0, r = 0.0;
        switch (op) {
        case 0: r = M_SQRT_(a); break;
        case 1: r = M_SIN_(a); break;
        case 2: r = M_COS_(a); break;
        case 3: r = M_TAN_(a); break;
        case 4: r = M_ASIN_(a); break;
        case 5: r = M_ACOS_(a); break;
        case 6: r = M_ATAN2_(a, b); break;
        case 7: r = M_LOG_(a); break;
        case 8: r = M_EXP_(a); break;
        case 9: r = M_POW_(a, b); break;
        default: return -1;
        }
        out[i] = r;
    }
    return 0;
}


/* pks2map.f90:144-262 on the host, default real (float).  Fortran's
 * `2 * randnum - 1` and `2 * 3.14159 * randnum` mix real(4) and real(8):
 * the float factor is formed first, the product is taken in double and
 * rounded to float on assignment. */
void p2mo_host_branches(float *h, int64_t *n, float cut_low, float cut_high, float chihview,
                        int scramble, int centre)
{
    int64_t m = 0;
    for (int64_t i = 0; i < *n; ++i) {                       /* :144-160 */
        float *p = h + 7 * i;
        float xh = p[2] - chihview, yh = p[1], zh = p[0];
        float chih = sqrtf((xh * xh + yh * yh) + zh * zh);
        if (chih > cut_high || chih < cut_low) continue;
        if (m != i) memmove(h + 7 * m, p, 7 * sizeof(float));
        ++m;
    }
    *n = m;
    if (scramble == 1) {                                      /* :162-187 */
        int32_t seed[4] = {13579, 13579, 13579, 13579};       /* seed = 13579*(myid+1), myid 0 */
        const float twopi = 2.0f * 3.14159f;
        for (int64_t i = 0; i < m; ++i) {
            float *p = h + 7 * i;
            float xh = p[0], yh = p[1], zh = p[2];
            float chih = sqrtf((xh * xh + yh * yh) + zh * zh);
            double r = p2mo_ranf(seed);
            float muran = (float)(2.0 * r - 1.0);
            r = p2mo_ranf(seed);
            float phiran = (float)((double)twopi * r);
            float rh = sqrtf(1.0f - muran * muran) * chih;
            p[0] = rh * cosf(phiran);
            p[1] = rh * sinf(phiran);
            p[2] = muran * chih;
        }
    }
    if (centre == 0 || m == 0) return;                        /* :189 */
    float rthmax = h[6];
    for (int64_t i = 1; i < m; ++i) if (h[7 * i + 6] > rthmax) rthmax = h[7 * i + 6];
    float xm = -1e5f, ym = -1e5f, zm = -1e5f;
    for (int64_t i = 0; i < m; ++i)                           /* :200-213: the last match wins */
        if (h[7 * i + 6] == rthmax) { xm = h[7 * i + 2]; ym = h[7 * i + 1]; zm = h[7 * i + 0]; }
    for (int64_t i = 0; i < m; ++i) {                         /* :220-222 */
        if (xm < 0.0f) h[7 * i + 2] = -h[7 * i + 2];
        if (ym < 0.0f) h[7 * i + 1] = -h[7 * i + 1];
        if (zm < 0.0f) h[7 * i + 0] = -h[7 * i + 0];
    }
    xm = fabsf(xm); ym = fabsf(ym); zm = fabsf(zm);
    const float rt = asinf(zm / sqrtf(xm * xm + zm * zm));   /* :227-229 */
    const float xml = xm / fabsf(xm) * sqrtf(xm * xm + zm * zm);
    const float rp = asinf(ym / sqrtf(ym * ym + xml * xml));
    const float ct = cosf(rt), st = sinf(rt), cp = cosf(rp), sp = sinf(rp);
    for (int64_t i = 0; i < m; ++i) {                         /* :232-244 */
        float *p = h + 7 * i;
        float xh = p[2] * ct + p[0] * st;
        float zh = -p[2] * st + p[0] * ct;
        float yh = -xh * sp + p[1] * cp;
        xh = xh * cp + p[1] * sp;
        p[2] = zh;
        p[1] = yh;
        p[0] = xh;
    }
}
