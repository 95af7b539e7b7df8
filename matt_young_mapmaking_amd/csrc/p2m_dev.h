/*
 * p2m_dev.h -- device-side data layout and FP64 routines of the painter.
 *
 * Every FP64 routine here reproduces the arithmetic of the reference path
 * operation for operation (same operands, same association, same elementary
 * functions from p2m_math.h), so pixel membership, profile-table cells and
 * FP64 contributions agree bit for bit with a faithful CPU evaluation.
 * Reference: src/pks2map_src/{pks2map,haloproject,maptable,cosmology}.f90 and
 * the vendored HEALPix 3.10 (src/modules/External/pix_tools.F90,
 * pixel_routines.F90).
 */
#ifndef P2M_DEV_H
#define P2M_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include "p2m_math.h"

#define P2M_MAXMAPS 6
#define P2M_DEV __device__ static inline

/* healpix_types.F90:65-73 */
#define HPX_PI      3.141592653589793238462643383279502884197
#define HPX_HALFPI  1.570796326794896619231321691639751442099
#define HPX_TWOPI   6.283185307179586476925286766559005768394
#define HPX_TWOTHIRD 0.6666666666666666666666666666666666666666
/* cosmology.f90:8-9: double parameter from a real(4) literal */
#define COSMO_PI ((double)3.14159265359f)

/* Distance tables (cosmology.f90:40-43), FP64, built on the host. */
struct CosmoDev {
    const double *rofz;
    const double *zofr;
    double drofz, dzofr, rofz_min, zofr_min;
    int32_t nrofz, nzofr;
    double omegam, omegal, rho_0, rhocrit_0, mypi;
};

/* Profile-table grid shared by all map slots (maptable.f90:9-21, 166-168). */
struct TabGeom {
    int32_t nchit, nmht, nrt;
    double chimint, chimaxt, mhmint, mhmaxt, rmint, rmaxt;
    double drt, dchit, dmht;
    double lr0, lc0, lm0;        /* log of the lower bounds */
};

/* One painted halo after decode: the per-halo constants of both painters
 * and of interpolate_table.  Struct of doubles first for alignment. */
struct HaloRec {
    double hx, hy, hz;       /* hpvec1 = (z,y,x)/chi (haloproject.f90:51-53) */
    double dtheta;           /* rmax*asin(rh/chi) (haloproject.f90:47) */
    double zz0, phi0, a;     /* discphirange_at_z: z0, atan2(y0,x0), x0^2+y0^2 */
    double cosang;           /* cos(radius) */
    double s0;               /* sqrt(a): sin of the centre colatitude */
    double thr;              /* (2 sin(dtheta/2))^2: chord^2 of the disc edge */
    double chi_r;            /* rofz(z) of interpolate_table (maptable.f90:184) */
    double fc, fm;           /* per-halo table fractions (maptable.f90:201-202) */
    double vh;               /* velocity factor (1 for tSZ) */
    double th, ph;           /* flat sky: asin(x/chi), asin(y/chi) (haloproject.f90:105-106) */
    int64_t cat;             /* catalogue index */
    int32_t ichi, imh;       /* per-halo table cells (maptable.f90:197-198) */
    int32_t irmin, irmax;    /* ring range of query_disc */
    int32_t chi_out;         /* chi > chimaxt: every value is 0 */
    int32_t flat;            /* 1 if the flat-sky painter covers this halo */
    int32_t jmin, jmax, kmin, kmax;  /* flat-sky index box (haloproject.f90:124-128) */
};

/* The fields of a painted halo that the binning reads (query_disc per
 * ring), in spatial order. */
struct BinHalo {
    double cosang, zz0, a, phi0;
    int32_t irmin, irmax;
};

/* The fields of a painted halo that the FP32 tile painter stages per
 * segment, packed into one 64-B line (written in spatial order). */
struct __attribute__((aligned(64))) TileHalo {
    double zz0, s0, phi0, thr;   /* as in HaloRec */
    float chi2;                  /* (float)(chi_r * chi_r) */
    float vh;                    /* velocity factor; 0 when chi > chimaxt */
    float fc, fm;                /* table fractions, rounded to float */
    int32_t colbase;             /* nrt * ((imh - 1) + nmht * (ichi - 1)); 0 when chi > chimaxt */
    uint32_t pad;
    unsigned long long hash;     /* membership hash of the catalogue index */
};

/* ------------------------------------------------------------ cosmology */
P2M_DEV double tab_lookup(const double *t, int n, double tmin, double dt, double x)
{
    if (x != x) return x;
    double u = (x - tmin) / dt + 1.0;
    if (u < 1.0) return t[0];
    if (u >= (double)n) return t[n - 1];
    int itab = (int)u;
    double fint = ((x - (double)(itab - 1) * dt) - tmin) / dt;
    return (1.0 - fint) * t[itab - 1] + fint * t[itab];
}

P2M_DEV double cube(double x) { return (x * x) * x; }

/* cosmology.f90:185-195 */
P2M_DEV double rhocrit_z(const CosmoDev &c, double z)
{
    double z3 = cube(1.0 + z);
    return (c.rhocrit_0 * (c.omegam * z3 + c.omegal)) / z3;
}

/* cosmology.f90:199-214 */
P2M_DEV double deltacrit_z(const CosmoDev &c, double z)
{
    double z3 = cube(1.0 + z);
    double omegaz = (c.omegam * z3) / ((c.omegam * z3 + 1.0) - c.omegam);
    double x = omegaz - 1.0;
    return (18.0 * (COSMO_PI * COSMO_PI) + 82.0 * x) - 39.0 * (x * x);
}

/* ------------------------------------------------------------ HEALPix */
/* ring_num, pix_tools.F90:490-533 */
P2M_DEV int32_t ring_num(int32_t nside, double z)
{
    int32_t iring = (int32_t)p2m_nint((double)nside * (2.0 - 1.5 * z));
    if (z > HPX_TWOTHIRD) {
        iring = (int32_t)p2m_nint((double)nside * p2m_sqrt(3.0 * (1.0 - z)));
        if (iring == 0) iring = 1;
    }
    if (z < -HPX_TWOTHIRD) {
        iring = (int32_t)p2m_nint((double)nside * p2m_sqrt(3.0 * (1.0 + z)));
        if (iring == 0) iring = 1;
        iring = 4 * nside - iring;
    }
    return iring;
}

/* ring2z, pix_tools.F90:535-560 */
P2M_DEV double ring2z(int32_t nside, int32_t ir)
{
    double fn = (double)nside;
    if (ir < nside) {
        double tmp = (double)ir;
        return 1.0 - (tmp * tmp) / ((3.0 * fn) * fn);
    } else if (ir < 3 * nside) {
        return ((double)(2 * nside - ir) * 2.0) / (3.0 * fn);
    }
    double tmp = (double)(4 * nside - ir);
    return -1.0 + (tmp * tmp) / ((3.0 * fn) * fn);
}

/* pixels_per_ring, pix_tools.F90:802-836: npr, kshift, first pixel of ring */
P2M_HD void ring_info(int32_t nside, int32_t ring, int32_t &npr, int32_t &kshift, int64_t &first)
{
    int32_t m = nside < ring ? nside : ring;
    if (4 * nside - ring < m) m = 4 * nside - ring;
    npr = 4 * m;
    kshift = (ring + 1) & 1;
    if (nside == 1) kshift = 1 - kshift;
    if (npr < 4 * nside) kshift = 1;
    int64_t npnorth;
    if (ring <= nside) {
        npnorth = (int64_t)ring * (ring + 1LL) * 2LL;
    } else if (ring <= 3 * nside) {
        npnorth = (int64_t)nside * (nside + 1LL) * 2LL + 4LL * nside * (int64_t)(ring - nside);
    } else {
        int64_t irs = 4LL * nside - ring - 1;
        npnorth = 12LL * nside * nside - irs * (irs + 1LL) * 2LL;
    }
    first = npnorth - npr;
}

/* Ring tiles: ring `ring` is cut into ceil(npr / W) tiles of W = 2^log2w
 * consecutive RING pixels, the last one narrower.  Tile `col` starts at
 * in-ring pixel ip0 = col W, holds `width` pixels, and its first map pixel is
 * first + ip0 (first: the ring's first pixel).  The host numbers the tiles
 * ring by ring with ring_tiles; k_tile_desc fills each descriptor with
 * tile_span. */
P2M_HD int32_t ring_tiles(int32_t nside, int32_t ring, int32_t log2w)
{
    int32_t npr, kshift;
    int64_t first;
    ring_info(nside, ring, npr, kshift, first);
    return (npr + (1 << log2w) - 1) >> log2w;
}

P2M_HD void tile_span(int32_t nside, int32_t ring, int32_t col, int32_t log2w, int64_t &first, int32_t &ip0,
                      int32_t &width)
{
    int32_t npr, kshift;
    ring_info(nside, ring, npr, kshift, first);
    ip0 = col << log2w;
    width = npr - ip0 < (1 << log2w) ? npr - ip0 : (1 << log2w);
}

/* Pixel-centre z and sin(theta) of ring `ring`, exactly as pix2vec_ring
 * computes them for any pixel of that ring (pixel_routines.F90:229-311). */
P2M_DEV void ring_zsth(int32_t nside, int32_t ring, double &z, double &sth)
{
    if (ring < nside) {
        double fact2 = (3.0 * nside) * nside;
        double ir = (double)ring;
        z = 1.0 - (ir / fact2) * ir;
        sth = ir * p2m_sqrt((1.0 + z) / fact2);
    } else if (ring <= 3 * nside) {
        double fact1 = 1.5 * nside;
        z = (double)(2 * nside - ring) / fact1;
        sth = p2m_sqrt((1.0 - z) * (1.0 + z));
    } else {
        double fact2 = (3.0 * nside) * nside;
        double ir = (double)(4 * nside - ring);
        z = -1.0 + (ir / fact2) * ir;
        sth = ir * p2m_sqrt((1.0 - z) / fact2);
    }
}

/* phi of pixel ip of ring `ring`: phi = (iphi + off) * HALFPI / den with the
 * association of pix2vec_ring. */
P2M_DEV void ring_phi_param(int32_t nside, int32_t ring, double &off, double &den)
{
    if (ring < nside) { off = 0.5; den = (double)ring; }
    else if (ring <= 3 * nside) { off = 0.5 * (double)((ring + nside + 1) & 1); den = (double)nside; }
    else { off = 0.5; den = (double)(4 * nside - ring); }
}

P2M_DEV double ring_phi(double ip, double off, double den) { return ((ip + off) * HPX_HALFPI) / den; }

/* pix2vec_ring from the pixel index (pixel_routines.F90:167-359, I4B path). */
P2M_DEV void pix2vec_ring(int32_t nside, int64_t ipix, double v[3])
{
    int64_t npix = 12LL * nside * nside;
    int64_t nl2 = 2LL * nside;
    int64_t ncap = nl2 * (nside - 1LL);
    double z, sth, phi;
    if (ipix < ncap) {
        int64_t iring = ((int64_t)p2m_floor(p2m_sqrt((double)(2 * ipix + 2))) + 1) / 2;
        int64_t iphi = ipix - 2 * iring * (iring - 1);
        double fact2 = (3.0 * nside) * nside;
        z = 1.0 - ((double)iring / fact2) * (double)iring;
        sth = (double)iring * p2m_sqrt((1.0 + z) / fact2);
        phi = (((double)iphi + 0.5) * HPX_HALFPI) / (double)iring;
    } else if (ipix < npix - ncap) {
        int64_t nl4 = 4LL * nside;
        int64_t ip = ipix - ncap;
        int64_t iring = ip / nl4 + nside;
        int64_t iphi = ip & (nl4 - 1);
        double fact1 = 1.5 * nside;
        double fodd = 0.5 * (double)((iring + nside + 1) & 1);
        z = (double)(nl2 - iring) / fact1;
        sth = p2m_sqrt((1.0 - z) * (1.0 + z));
        phi = (((double)iphi + fodd) * HPX_HALFPI) / (double)nside;
    } else {
        int64_t ip = npix - ipix;
        int64_t iring = ((int64_t)p2m_floor(p2m_sqrt((double)(2 * ip))) + 1) / 2;
        int64_t iphi = 2 * iring * (iring + 1) - ip;
        double fact2 = (3.0 * nside) * nside;
        z = -1.0 + ((double)iring / fact2) * (double)iring;
        sth = (double)iring * p2m_sqrt((1.0 - z) / fact2);
        phi = (((double)iphi + 0.5) * HPX_HALFPI) / (double)iring;
    }
    double s, c;
    p2m_sincos(phi, &s, &c);
    v[0] = sth * c;
    v[1] = sth * s;
    v[2] = z;
}

/* angdist + vect_prod (pix_tools.F90:2015-2055) */
P2M_DEV double angdist(double ax, double ay, double az, double bx, double by, double bz)
{
    double sprod = (ax * bx + ay * by) + az * bz;
    double v0 = ay * bz - az * by;
    double v1 = az * bx - ax * bz;
    double v2 = ax * by - ay * bx;
    double vprod = p2m_sqrt((v0 * v0 + v1 * v1) + v2 * v2);
    return p2m_atan2(vprod, sprod);
}

/* phi half-width of the disc on a ring at height z (discphirange_at_z,
 * pix_tools.F90:712-760); -1000 = no intersection. */
template <typename H>
P2M_HD double disc_dphi(const H &h, double z)
{
    double dphi = -1000.0;
    double b = h.cosang - z * h.zz0;
    if (h.a == 0.0) {
        if (b <= 0.0) dphi = HPX_PI;
    } else {
        double cc = 1.0 - z * z;
        if (cc < 1.0e-12) cc = 1.0e-12;
        double cosdphi = b / p2m_sqrt(h.a * cc);
        if (cosdphi < -1.0) dphi = HPX_PI;
        if (p2m_fabs(cosdphi) <= 1.0) dphi = p2m_acos(cosdphi);
    }
    return dphi;
}

/* x mod n in [0, n) (modulo of pixels_on_edge); the disc edges lie within a
 * ring's length or two of zero, so the 64-bit division is rarely needed */
P2M_HD int64_t mod_ring(int64_t x, int64_t n)
{
    if (x >= 0 && x < n) return x;
    if (x < 0 && x >= -n) return x + n;
    if (x >= n && x < 2 * n) return x - n;
    const int64_t r = x % n;
    return r < 0 ? r + n : r;
}

/* Pixel range of the disc on ring `ring` (pixels_on_edge, pix_tools.F90:762-800,
 * and discedge2fulldisc, pixel_routines.F90:1823-1879).  Returns the number
 * of pixels np (0 if none) and the first in-ring index lo (pixels lo..lo+np-1
 * modulo npr). */
template <typename H>
P2M_HD int32_t disc_ring_run_z(const H &h, double z, int32_t npr, int32_t kshift, int32_t &lo);

template <typename H>
P2M_DEV int32_t disc_ring_run(const H &h, int32_t nside, int32_t ring, int32_t npr, int32_t kshift, int32_t &lo)
{
    return disc_ring_run_z(h, ring2z(nside, ring), npr, kshift, lo);
}

/* the same, with the ring's z (ring2z) given */
template <typename H>
P2M_HD int32_t disc_ring_run_z(const H &h, double z, int32_t npr, int32_t kshift, int32_t &lo)
{
    double dphi = disc_dphi(h, z);
    if (dphi >= HPX_PI) { lo = 0; return npr; }
    if (!(dphi >= 0.0)) return 0;
    double shift = (double)kshift * 0.5;
    int64_t il = (int64_t)p2m_ceil(((double)npr * (h.phi0 - dphi)) / HPX_TWOPI - shift);
    int64_t ih = (int64_t)p2m_floor(((double)npr * (h.phi0 + dphi)) / HPX_TWOPI - shift);
    if (ih < il) return 0;
    const int64_t l = mod_ring(il, npr);
    const int64_t hh = mod_ring(ih, npr);
    int64_t np = mod_ring(hh - l, npr);
    np += 1;
    if (np > npr) np = npr;
    lo = (int32_t)l;
    return (int32_t)np;
}

/* The run of disc_ring_run_z with the arc half-width dphi = acos(cosdphi)
 * taken as 2 asin(sqrt((1 - cosdphi)/2)) in FP32 (cosdphi itself, and
 * 1 - cosdphi, in FP64 as the reference forms them).  The ends
 * npr (phi0 -+ dphi) / 2pi - shift are formed in FP64 with the FP32 dphi;
 * when either lies within the bound of that dphi's error (1e-6 relative,
 * against a few ulp of the FP32 asin and sqrt) of an integer, where ceil or
 * floor could differ from the reference's, or the arc nearly closes the ring,
 * the exact FP64 evaluation decides.  So the returned run is the exact one
 * for every disc and ring. */
template <typename H>
P2M_HD int32_t disc_ring_run_fast(const H &h, double z, int32_t npr, int32_t kshift, int32_t &lo)
{
    if (h.a == 0.0) return disc_ring_run_z(h, z, npr, kshift, lo);      /* no acos on this branch */
    const double b = h.cosang - z * h.zz0;
    double cc = 1.0 - z * z;
    if (cc < 1.0e-12) cc = 1.0e-12;
    const double cosdphi = b / p2m_sqrt(h.a * cc);
    if (cosdphi < -1.0) { lo = 0; return npr; }
    if (!(p2m_fabs(cosdphi) <= 1.0)) return 0;
    if (1.0 + cosdphi < 1e-5) return disc_ring_run_z(h, z, npr, kshift, lo);   /* dphi near pi */
    const float sh = sqrtf((float)(0.5 * (1.0 - cosdphi)));
    const double dphi = (double)(2.0f * asinf(sh));
    const double shift = (double)kshift * 0.5;
    const double k = (double)npr * (1.0 / HPX_TWOPI);
    const double tl = (h.phi0 - dphi) * k - shift;
    const double th = (h.phi0 + dphi) * k - shift;
    const double eps = k * (1e-6 * dphi + 1e-12) + 1e-9;
    const double fl = tl - p2m_floor(tl), fh = th - p2m_floor(th);
    if (fl < eps || fl > 1.0 - eps || fh < eps || fh > 1.0 - eps) return disc_ring_run_z(h, z, npr, kshift, lo);
    const int64_t il = (int64_t)p2m_ceil(tl);
    const int64_t ih = (int64_t)p2m_floor(th);
    if (ih < il) return 0;
    const int64_t l = mod_ring(il, npr);
    const int64_t hh = mod_ring(ih, npr);
    int64_t np = mod_ring(hh - l, npr);
    np += 1;
    if (np > npr) np = npr;
    lo = (int32_t)l;
    return (int32_t)np;
}

/* ------------------------------------------------------------ profile */
/* Table element table(p, ir, imh, ichi) through the flat column-major index,
 * addressed past the declared bounds as the reference does (quirk Q6). */
P2M_DEV double tab_flat(const float *t, int64_t nt, const TabGeom &g, int p, int ir, int imh, int ichi)
{
    int64_t idx = (int64_t)(p - 1) + 2 * ((int64_t)(ir - 1) + (int64_t)g.nrt * ((int64_t)(imh - 1) + (int64_t)g.nmht * (int64_t)(ichi - 1)));
    if (idx < 0 || idx >= nt) return 0.0;
    return (double)t[idx];
}

/* interpolate_table (maptable.f90:176-216) for a halo whose per-halo part
 * (chi, clamps, ichi/imh, fc/fm) is precomputed in h. */
P2M_DEV double interp_exact(const HaloRec &h, const TabGeom &g, const float *t, int64_t nt, int profile, double theta)
{
    if (h.chi_out) return 0.0;
    double r = h.chi_r * p2m_tan(theta);
    if (r > g.rmaxt) return 0.0;
    if (r < g.rmint) r = g.rmint + 1e-5;
    double lr = p2m_log(r);
    int ir = (int)((lr - g.lr0) / g.drt) + 1;
    double fr = lr - (g.lr0 + (double)(ir - 1) * g.drt);
    double fc = h.fc, fm = h.fm;
    int imh = h.imh, ichi = h.ichi;
    double v = tab_flat(t, nt, g, profile, ir, imh, ichi) * (1.0 - fr) * (1.0 - fc) * (1.0 - fm);
    v = v + tab_flat(t, nt, g, profile, ir + 1, imh, ichi) * (fr) * (1.0 - fc) * (1.0 - fm);
    v = v + tab_flat(t, nt, g, profile, ir, imh + 1, ichi) * (1.0 - fr) * (fc) * (1.0 - fm);
    v = v + tab_flat(t, nt, g, profile, ir + 1, imh + 1, ichi) * (fr) * (fc) * (1.0 - fm);
    v = v + tab_flat(t, nt, g, profile, ir, imh, ichi + 1) * (1.0 - fr) * (1.0 - fc) * (fm);
    v = v + tab_flat(t, nt, g, profile, ir + 1, imh, ichi + 1) * (fr) * (1.0 - fc) * (fm);
    v = v + tab_flat(t, nt, g, profile, ir, imh + 1, ichi + 1) * (1.0 - fr) * (fc) * (fm);
    v = v + tab_flat(t, nt, g, profile, ir + 1, imh + 1, ichi + 1) * (fr) * (fc) * (fm);
    return v;
}

P2M_DEV uint64_t halo_hash(int64_t i)
{
    uint64_t z = (uint64_t)i + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

#endif
