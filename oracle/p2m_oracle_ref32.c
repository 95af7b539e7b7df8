/*
 * p2m_oracle_ref32.c -- the reference's own arithmetic (TEST INFRASTRUCTURE).
 *
 * The "ref32" mode of the oracle: pks2map's halo loop evaluated in the
 * reference's default-real (float32) arithmetic, statement by statement, with
 * Fortran's mixed-mode promotion rules: an expression of default reals is
 * evaluated in float; a double operand (a `double precision` variable or
 * parameter) promotes the rest of that operation to double; assignment to a
 * `real` rounds to float.  The HEALPix routines (query_disc, pix2vec_ring,
 * angdist) are double precision in the reference too and are the FP64
 * oracle's.  Elementary functions of float arguments are the C library's
 * float functions (the reference's Intel libimf is absent: those last bits
 * are parity unpinned).
 *
 * Work layout mirrors the reference's MPI run: contiguous halo blocks of
 * nl = (n-1)/T + 1 halos per thread (pksc.f90:51-52), thread-private float
 * maps zeroed per thread (pks2map.f90:278-279), then a float sum of the
 * private maps in thread order (the MPI_Allreduce of :313-314).  It is the
 * CPU baseline that does the reference's per-pixel work.
 *
 * Only tests/ and bench.py's cpu_baseline leg load it.
 */
#include "p2m_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* cosmology.f90:8-9 (Q8) and :13-23 */
#define R32_PI   ((double)3.14159265359f)
#define R32_CKMS (2.99792458e10 / 1e5)
#define R32_NTAB 10000

typedef struct {
    float omegam, omegal, h, rho_0, rhocrit_0, mypi;
    float rofz_min, rofz_max, drofz, zofr_min, zofr_max, dzofr;
    float rofz[R32_NTAB], zofr[R32_NTAB];
} r32_cosmo;

typedef struct {
    int32_t nchit, nmht, nrt;
    float chimint, chimaxt, mhmint, mhmaxt, rmint, rmaxt;
    float drt, dchit, dmht;
    const float *t;
    int64_t nt;
} r32_table;

static inline float cube_f(float x) { return (x * x) * x; }

/* dzdr, cosmology.f90:123-133: h*100/ckms*sqrt(omegam*(1.+z)**3+1.-omegam);
 * h*100 is real, /ckms promotes to double, the sqrt argument is real */
static float r32_dzdr(const r32_cosmo *c, float z)
{
    const float s = sqrtf((c->omegam * cube_f(1.0f + z) + 1.0f) - c->omegam);
    return (float)((((double)(c->h * 100.0f)) / R32_CKMS) * (double)s);
}

/* set_rofztable / set_zofrtable, cosmology.f90:51-105, in real */
static void r32_cosmo_init(r32_cosmo *c, float omegam, float omegal, float h)
{
    c->omegam = omegam; c->omegal = omegal; c->h = h;
    c->rho_0 = (2.775e11f * omegam) * (h * h);              /* pks2map.f90:127 */
    c->rhocrit_0 = 2.775e11f * (h * h);                      /* :128 */
    c->mypi = 4.0f * atanf(1.0f);                            /* :88 */
    c->rofz_min = 0.0f; c->rofz_max = 6.0f;
    c->drofz = (c->rofz_max - c->rofz_min) / (float)(R32_NTAB - 1);
    c->rofz[0] = 0.0f;
    for (int i = 2; i <= R32_NTAB; ++i) {
        const float z = (float)(i - 1) * c->drofz + c->rofz_min;
        const float drdz = 1.0f / r32_dzdr(c, z);            /* :109-119 */
        c->rofz[i - 1] = c->rofz[i - 2] + drdz * c->drofz;
    }
    c->zofr_min = 0.0f; c->zofr_max = 1.3e4f;
    c->dzofr = (c->zofr_max - c->zofr_min) / (float)(R32_NTAB - 1);
    c->zofr[0] = 0.0f;
    for (int i = 2; i <= R32_NTAB; ++i) {
        const float z = c->zofr[i - 2];
        c->zofr[i - 1] = c->zofr[i - 2] + r32_dzdr(c, z) * c->dzofr;
    }
}

/* zofr / rofz, cosmology.f90:137-181: itab = ((r-min)/d)+1 truncated */
static float r32_lookup(const float *t, float tmin, float dt, float x)
{
    const float u = (x - tmin) / dt + 1.0f;
    if (u != u) return u;
    if (u < 1.0f) return t[0];                               /* int(u) < 1 */
    if (u >= (float)R32_NTAB) return t[R32_NTAB - 1];        /* int(u) >= n */
    const int itab = (int)u;
    const float fint = ((x - (float)(itab - 1) * dt) - tmin) / dt;
    return (1.0f - fint) * t[itab - 1] + fint * t[itab];
}
static float r32_zofr(const r32_cosmo *c, float r) { return r32_lookup(c->zofr, c->zofr_min, c->dzofr, r); }
static float r32_rofz(const r32_cosmo *c, float z) { return r32_lookup(c->rofz, c->rofz_min, c->drofz, z); }

/* rhocrit, cosmology.f90:185-195 */
static float r32_rhocrit(const r32_cosmo *c, float z)
{
    const float z3 = cube_f(1.0f + z);
    return (c->rhocrit_0 * (c->omegam * z3 + c->omegal)) / z3;
}

/* deltacrit, cosmology.f90:199-214: 18*pi**2 is double (pi is a double
 * parameter), so the sum is formed in double and rounded on return */
static float r32_deltacrit(const r32_cosmo *c, float z)
{
    const float z3 = cube_f(1.0f + z);
    const float omegaz = (c->omegam * z3) / ((c->omegam * z3 + 1.0f) - c->omegam);
    const float x = omegaz - 1.0f;
    return (float)((18.0 * (R32_PI * R32_PI) + (double)(82.0f * x)) - (double)(39.0f * (x * x)));
}

/* loadmaptable steps, maptable.f90:166-168, in real */
static void r32_table_init(r32_table *tb, const int32_t hdr_i[3], const float hdr_f[6], const float *t, int64_t nt)
{
    tb->nchit = hdr_i[0]; tb->nmht = hdr_i[1]; tb->nrt = hdr_i[2];
    tb->chimint = hdr_f[0]; tb->chimaxt = hdr_f[1]; tb->mhmint = hdr_f[2];
    tb->mhmaxt = hdr_f[3]; tb->rmint = hdr_f[4]; tb->rmaxt = hdr_f[5];
    tb->dchit = (logf(tb->chimaxt) - logf(tb->chimint)) / (float)(tb->nchit - 1);
    tb->drt = (logf(tb->rmaxt) - logf(tb->rmint)) / (float)(tb->nrt - 1);
    tb->dmht = (logf(tb->mhmaxt) - logf(tb->mhmint)) / (float)(tb->nmht - 1);
    tb->t = t; tb->nt = nt;
}

/* table(p, ir, imh, ichi) by the column-major flat index, zero past the end (Q6) */
static inline float r32_tval(const r32_table *tb, int p, int ir, int imh, int ichi)
{
    const int64_t idx = (int64_t)(p - 1) + 2 * ((int64_t)(ir - 1) + (int64_t)tb->nrt *
                        ((int64_t)(imh - 1) + (int64_t)tb->nmht * (int64_t)(ichi - 1)));
    if (idx < 0 || idx >= tb->nt) return 0.0f;
    return tb->t[idx];
}

/* interpolate_table, maptable.f90:176-216, in real; mh is clamped in place (Q3) */
static float r32_interpolate(const r32_table *tb, const r32_cosmo *c, float theta, float z, float *mh, int profile)
{
    float chi = r32_rofz(c, z);
    float r = chi * tanf(theta);
    if (r > tb->rmaxt || chi > tb->chimaxt) return 0.0f;
    if (r < tb->rmint) r = tb->rmint + 1e-5f;
    if (chi < tb->chimint) chi = tb->chimint + 1e-5f;
    if (*mh < tb->mhmint) *mh = tb->mhmint + 1e-5f;
    if (*mh > tb->mhmaxt) *mh = tb->mhmaxt - 1e-5f;
    const int ir = (int)((logf(r) - logf(tb->rmint)) / tb->drt) + 1;
    const int ichi = (int)((logf(chi) - logf(tb->chimint)) / tb->dchit) + 1;
    const int imh = (int)((logf(*mh) - logf(tb->mhmint)) / tb->dmht) + 1;
    const float fr = logf(r) - (logf(tb->rmint) + (float)(ir - 1) * tb->drt);
    const float fc = logf(chi) - (logf(tb->chimint) + (float)(ichi - 1) * tb->dchit);
    const float fm = logf(*mh) - (logf(tb->mhmint) + (float)(imh - 1) * tb->dmht);
    /* maptable.f90:204-212 (Q4: imh+1 carries fc, ichi+1 carries fm) */
    float v = r32_tval(tb, profile, ir, imh, ichi) * (1.0f - fr) * (1.0f - fc) * (1.0f - fm);
    v = v + r32_tval(tb, profile, ir + 1, imh, ichi) * (fr) * (1.0f - fc) * (1.0f - fm);
    v = v + r32_tval(tb, profile, ir, imh + 1, ichi) * (1.0f - fr) * (fc) * (1.0f - fm);
    v = v + r32_tval(tb, profile, ir + 1, imh + 1, ichi) * (fr) * (fc) * (1.0f - fm);
    v = v + r32_tval(tb, profile, ir, imh, ichi + 1) * (1.0f - fr) * (1.0f - fc) * (fm);
    v = v + r32_tval(tb, profile, ir + 1, imh, ichi + 1) * (fr) * (1.0f - fc) * (fm);
    v = v + r32_tval(tb, profile, ir, imh + 1, ichi + 1) * (1.0f - fr) * (fc) * (fm);
    v = v + r32_tval(tb, profile, ir + 1, imh + 1, ichi + 1) * (fr) * (fc) * (fm);
    return v;
}

int p2mo_ref32_tables(double omegam, double omegal, double h, float *rofz, float *zofr)
{
    r32_cosmo *c = (r32_cosmo *)malloc(sizeof(r32_cosmo));
    if (!c) return -1;
    r32_cosmo_init(c, (float)omegam, (float)omegal, (float)h);
    memcpy(rofz, c->rofz, sizeof(c->rofz));
    memcpy(zofr, c->zofr, sizeof(c->zofr));
    free(c);
    return 0;
}

float p2mo_ref32_interpolate(const int32_t hdr_i[3], const float hdr_f[6], const float *padded, int64_t n_padded,
                             double omegam, double omegal, double h, float theta, float z, float *mh, int profile)
{
    r32_cosmo *c = (r32_cosmo *)malloc(sizeof(r32_cosmo));
    if (!c) return NAN;
    r32_cosmo_init(c, (float)omegam, (float)omegal, (float)h);
    r32_table tb;
    r32_table_init(&tb, hdr_i, hdr_f, padded, n_padded);
    const float v = r32_interpolate(&tb, c, theta, z, mh, profile);
    free(c);
    return v;
}

/* pks2map.f90:144-160 (cut), :264 (mass), :282-303 (halo loop) and
 * haloproject.f90:38-156 in real, on contiguous blocks of halos per thread;
 * the private maps are summed into hpmap / fsmap (overwritten).  setup[6*i]
 * receives, for each painted halo i of the input, (z, mh, rh, chi, dtheta, 1)
 * when setup is not NULL (entries of halos not painted stay untouched). */
int p2mo_paint_ref32(const p2mo_run *run, const int32_t hdr_i[3], const float hdr_f[6],
                     const float *padded, int64_t n_padded, const float *halos, int64_t n,
                     float *hpmap, float *fsmap, uint32_t *hp_count, uint64_t *hp_hash,
                     uint32_t *fs_count, uint64_t *fs_hash, double *setup, p2mo_stats *st)
{
    r32_cosmo *c = (r32_cosmo *)malloc(sizeof(r32_cosmo));
    if (!c) return -1;
    r32_cosmo_init(c, 0.25f, 0.75f, 0.7f);                   /* pks2map.f90:118-121 */
    r32_table tb;
    r32_table_init(&tb, hdr_i, hdr_f, padded, n_padded);
    const int32_t nside = run->nside, npix = run->npix;
    const int64_t hpn = nside > 0 ? 12LL * nside * nside : 0;
    const int64_t fsn = (int64_t)npix * npix;
    /* pks2map.f90:73-92: CLI reals, fov in radians, dpix */
    const float zmin = (float)run->zmin, zmax = (float)run->zmax, chihview = (float)run->chihview;
    const float mmin = (float)run->mmin;
    const float fov = (((float)run->fov_deg / 360.0f) * 2.0f) * c->mypi;
    const float dpix = npix > 0 ? fov / (float)npix : 1.0f;
    const float cut_low = r32_rofz(c, zmin), cut_high = r32_rofz(c, zmax);   /* :134-135 */
    const float dmodel = run->model == 2 ? 500.0f : 200.0f;                  /* bbps_profile.f90:21 */
    const float delta_pksc = 200.0f;                                         /* pksc.f90:8-9 */
    const float rmax = 4.0f;                                                 /* haloproject.f90:18 */
    int T = run->nthreads > 0 ? run->nthreads : 1;
    float **hpl = (float **)calloc((size_t)T, sizeof(float *));
    float **fsl = (float **)calloc((size_t)T, sizeof(float *));
    int fail = 0;
    int64_t n_cut = 0, n_painted = 0, n_list = 0, n_upd_hp = 0, n_upd_fs = 0, n_gaps = 0;

#pragma omp parallel num_threads(T) reduction(+:n_cut,n_painted,n_list,n_upd_hp,n_upd_fs,n_gaps)
    {
        int t = 0, nt = 1;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nt = omp_get_num_threads();
#endif
        /* the rank's block (pksc.f90:51-52) and its zeroed maps (pks2map.f90:278-279) */
        const int64_t nl = n > 0 ? (n - 1) / nt + 1 : 0;
        const int64_t i0 = (int64_t)t * nl, i1 = i0 + nl < n ? i0 + nl : n;
        float *hpm = hpn && hpmap ? (float *)calloc((size_t)hpn, sizeof(float)) : NULL;
        float *fsm = fsn && fsmap ? (float *)calloc((size_t)fsn, sizeof(float)) : NULL;
        if ((hpn && hpmap && !hpm) || (fsn && fsmap && !fsm)) {
#pragma omp atomic write
            fail = 1;
        }
        hpl[t] = hpm; fsl[t] = fsm;
        int32_t *list = NULL;
        int64_t list_cap = 0;
        for (int64_t i = i0; i < i1; ++i) {
            const float *p = halos + 7 * i;
            /* radial cut, pks2map.f90:144-152 */
            float xh = p[2], yh = p[1], zh = p[0];
            xh = xh - chihview;
            float chih = sqrtf((xh * xh + yh * yh) + zh * zh);
            if (chih > cut_high || chih < cut_low) continue;
            n_cut++;
            /* mass, :264 */
            const float rth = p[6];
            const float mass = (((4.0f / 3.0f) * c->mypi) * c->rho_0) * ((rth * rth) * rth);
            /* vrad: profile 1 -> 1 (pksc.f90:76-77); profile 2 -> the intended -v_r/c (Q18) */
            float vh = 1.0f;
            if (run->profile == 2) {
                const float vr = ((p[0] * p[3] + p[1] * p[4]) + p[2] * p[5]) /
                                 powf((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2], 1.0f / 2.0f);
                vh = -vr / 3.e5f;
            }
            /* halo loop body, :284-298 */
            xh = p[2]; yh = p[1]; zh = p[0];
            chih = sqrtf((xh * xh + yh * yh) + zh * zh);
            const float redshift = r32_zofr(c, chih);
            xh = xh - chihview;
            chih = sqrtf((xh * xh + yh * yh) + zh * zh);
            float mh = sqrtf(r32_deltacrit(c, redshift) / dmodel) * mass;
            if (mass < mmin) continue;
            n_painted++;
            const float rh = powf(((((3.0f * mh) / 4.0f) / c->mypi) / delta_pksc) / r32_rhocrit(c, redshift),
                                  1.0f / 3.0f);
            /* ---- projecthalo_flatsky, haloproject.f90:83-156 ---- */
            double dtheta = (double)(rmax * asinf(rh / chih));           /* :93, real -> double variable */
            if (setup) {
                double *s = setup + 6 * i;
                s[0] = redshift; s[1] = mh; s[2] = rh; s[3] = chih; s[4] = dtheta; s[5] = 1.0;
            }
            xh = xh / chih; yh = yh / chih; zh = zh / chih;              /* :97-99, mutates the caller's (Q2) */
            if (!(zh < 0.0f) && npix > 0) {
                const double th = (double)asinf(xh), ph = (double)asinf(yh);     /* :105-106 */
                float thetamin = (float)(th - (double)1.1f * dtheta);
                float thetamax = (float)(th + (double)1.1f * dtheta);
                float phimin = (float)(ph - (double)1.1f * dtheta);
                float phimax = (float)(ph + (double)1.1f * dtheta);
                const float hf = fov / 2.0f;
                if (!(thetamax < -hf || thetamin > hf) && !(phimax < -hf || phimin > hf)) {
                    if (thetamin <= -fov / 2.0f) thetamin = -fov / 2.0f + 1e-10f;
                    if (thetamax >= fov / 2.0f) thetamax = fov / 2.0f - 1e-10f;
                    if (phimin <= -fov / 2.0f) phimin = -fov / 2.0f + 1e-10f;
                    if (phimax >= fov / 2.0f) phimax = fov / 2.0f - 1e-10f;
                    int jmin = (int)((thetamin + fov / 2.0f) / dpix) + 1; if (jmin < 1) jmin = 1;
                    int jmax = (int)((thetamax + fov / 2.0f) / dpix) + 1; if (jmax > npix) jmax = npix;
                    int kmin = (int)((phimin + fov / 2.0f) / dpix) + 1; if (kmin < 1) kmin = 1;
                    int kmax = (int)((phimax + fov / 2.0f) / dpix) + 1; if (kmax > npix) kmax = npix;
                    for (int k = kmin; k <= kmax; ++k) {
                        const float phip = -fov / 2.0f + ((float)k - 0.5f) * dpix;
                        for (int j = jmin; j <= jmax; ++j) {
                            const float thetap = -fov / 2.0f + ((float)j - 0.5f) * dpix;
                            const float dphi = (float)((double)phip - ph);            /* flatskyvars.f90:15 */
                            const float dth = (float)((double)thetap - th);
                            const float angle = sqrtf(dphi * dphi + dth * dth);
                            if ((double)angle < dtheta) {
                                const float val = vh * r32_interpolate(&tb, c, angle, redshift, &mh, run->profile);
                                const int64_t pix = (int64_t)(j - 1) + (int64_t)npix * (k - 1);
                                if (fsm) fsm[pix] = fsm[pix] + val;
                                if (fs_count) {
#pragma omp atomic
                                    fs_count[pix] += 1u;
                                }
                                if (fs_hash) {
                                    const uint64_t hh = p2mo_halo_hash(i);
#pragma omp atomic
                                    fs_hash[pix] += hh;
                                }
                                n_upd_fs++;
                            }
                        }
                    }
                }
            }
            /* ---- projecthalo_healpix, haloproject.f90:38-79 ---- */
            if (nside > 0) {
                dtheta = (double)(rmax * asinf(rh / chih));                 /* :47 */
                double v1[3];
                v1[0] = (double)(zh / chih); v1[1] = (double)(yh / chih); v1[2] = (double)(xh / chih);   /* :51-53 */
                int64_t nlst = p2mo_query_disc(nside, v1, dtheta, NULL, NULL);
                if (nlst < 0) continue;
                if (nlst > list_cap) {
                    free(list);
                    list_cap = nlst + 1024;
                    list = (int32_t *)malloc(sizeof(int32_t) * (size_t)list_cap);
                }
                nlst = p2mo_query_disc(nside, v1, dtheta, list, &n_gaps);
                n_list += nlst;
                for (int64_t q = 0; q < nlst; ++q) {
                    const int32_t jp = list[q];
                    double v2[3];
                    p2mo_pix2vec_ring(nside, jp, v2);
                    const float angle = (float)p2mo_angdist(v1, v2);        /* :64-66 */
                    if ((double)angle < dtheta) {
                        const float val = vh * r32_interpolate(&tb, c, angle, redshift, &mh, run->profile);
                        if (hpm) hpm[jp] = hpm[jp] + val;
                        if (hp_count) {
#pragma omp atomic
                            hp_count[jp] += 1u;
                        }
                        if (hp_hash) {
                            const uint64_t hh = p2mo_halo_hash(i);
#pragma omp atomic
                            hp_hash[jp] += hh;
                        }
                        n_upd_hp++;
                    }
                }
            }
        }
        free(list);
    }
    /* the all-reduce: float sum of the private maps in thread order */
    if (!fail) {
        if (hpmap && hpn) {
#pragma omp parallel for num_threads(T) schedule(static)
            for (int64_t q = 0; q < hpn; ++q) {
                float s = 0.0f;
                for (int t = 0; t < T; ++t) if (hpl[t]) s = s + hpl[t][q];
                hpmap[q] = s;
            }
        }
        if (fsmap && fsn) {
            for (int64_t q = 0; q < fsn; ++q) {
                float s = 0.0f;
                for (int t = 0; t < T; ++t) if (fsl[t]) s = s + fsl[t][q];
                fsmap[q] = s;
            }
        }
    }
    for (int t = 0; t < T; ++t) { free(hpl[t]); free(fsl[t]); }
    free(hpl); free(fsl);
    if (st) {
        st->n_in = n; st->n_cut = n_cut; st->n_painted = n_painted;
        st->n_listed_hp = n_list; st->n_upd_hp = n_upd_hp; st->n_upd_fs = n_upd_fs;
        st->n_ring_gaps = n_gaps; st->cut_low = cut_low; st->cut_high = cut_high;
    }
    free(c);
    return fail ? -1 : 0;
}
