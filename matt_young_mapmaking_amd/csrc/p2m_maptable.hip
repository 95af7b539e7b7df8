/*
 * p2m_maptable.hip -- make_maptable on the GPU: the projected Battaglia et
 * al. (2012) tSZ y (times G(x) at frequency nu) and tau profiles on the
 * (chi, M, r) grid of maptable.f90:25-146, one work item per grid node.
 *
 * FP64 arithmetic with the portable elementary functions of p2m_math.h, in
 * the operation order of maptable.f90, integrate_profiles.f90:12-95 and
 * bbps_profile.f90:60-180; named constants keep the reference's values
 * (real(4) literals rounded to float where the reference declares them so).
 */
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pks2map_gpu.h"
#include "p2m_kernels.h"

namespace {

/* bbps_profile.f90:9-40 (real(4) parameter arrays) */
struct BBPS {
    float P0_Am[3], P0_alpham[3], P0_alphaz[3];
    float xc_Am[3], xc_alpham[3], xc_alphaz[3];
    float beta_Am[3], beta_alpham[3], beta_alphaz[3];
    float P0_Am_rho[3], P0_alpham_rho[3], P0_alphaz_rho[3];
    float alpha_Am_rho[3], alpha_alpham_rho[3], alpha_alphaz_rho[3];
    float beta_Am_rho[3], beta_alpham_rho[3], beta_alphaz_rho[3];
};

const BBPS kBBPS = {
    {18.100f, 7.490f, 20.7f}, {0.154f, 0.226f, -0.074f}, {-0.758f, -0.957f, -0.743f},
    {0.497f, 0.710f, 0.438f}, {-0.00865f, -0.0833f, 0.011f}, {0.731f, 0.853f, 1.01f},
    {4.35f, 4.19f, 3.82f}, {0.0393f, 0.048f, 0.0375f}, {0.415f, 0.615f, 0.535f},
    {1.9e-5f, 9.1e-05f, 1.5e-4f}, {0.29f, 0.10f, 0.14f}, {-0.67f, -0.92f, -1.32f},
    {0.88f, 0.69f, 0.68f}, {-0.03f, -0.02f, -0.02f}, {0.20f, 0.27f, 0.29f},
    {3.83f, 4.43f, 6.4f}, {0.044f, 0.0035f, 0.028f}, {-0.025f, 0.037f, -0.25f}};

struct MTArgs {
    BBPS b;
    int32_t mi, nchit, nmht, nrt;
    double nu, alpha_cmb, Delta, y0, tau0, x0const;
    double omegam, omegal, rhocrit_0;
    double chimint, mhmint, rmint, dchit, dmht, drt;
    const double *rofz, *zofr;
    float *table;
};

}  // namespace

#define MT_DS ((double)1e-2f)            /* integrate_profiles.f90:8 */
#define MT_GAMMA ((double)-0.3f)         /* bbps_profile.f90:25 */
#define MT_GAMMA_RHO ((double)0.2f)      /* bbps_profile.f90:40 */

__device__ static inline double mt_lookup(const double *t, double dt, double x)
{
    double u = (x - 0.0) / dt + 1.0;
    if (u < 1.0) return t[0];
    if (u >= 10000.0) return t[9999];
    int itab = (int)u;
    double fint = ((x - (double)(itab - 1) * dt) - 0.0) / dt;
    return (1.0 - fint) * t[itab - 1] + fint * t[itab];
}

__device__ static inline double mt_fit(float A0, float am, float az, double mh, double z)
{
    return ((double)A0 * p2m_pow(mh / 1e14, (double)am)) * p2m_pow(1.0 + z, (double)az);
}

__global__ void __launch_bounds__(256) k_make_maptable(MTArgs A)
{
    int64_t node = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t total = (int64_t)A.nchit * A.nmht * A.nrt;
    if (node >= total) return;
    int k = (int)(node % A.nrt) + 1;
    int64_t cm = node / A.nrt;
    int j = (int)(cm % A.nmht) + 1;
    int i = (int)(cm / A.nmht) + 1;
    const double dz_r = 6.0 / 9999.0, dr_z = 1.3e4 / 9999.0;
    double chit = p2m_exp(p2m_log(A.chimint) + (double)(i - 1) * A.dchit);
    double z = mt_lookup(A.zofr, dr_z, chit);
    double z1 = 1.0 + z;
    double zc = (z1 * z1) * z1;
    double zfacy = A.omegam * zc + A.omegal;
    double rhoc = (A.rhocrit_0 * (A.omegam * zc + A.omegal)) / zc;
    double zfact = (1.0 / (1.0 + z)) / p2m_pow(rhoc, 1.0 / 3.0);
    double Gx = 0.0;
    if (A.nu > 0.0) {
        double x0new = A.x0const * p2m_pow(1.0 + z, A.alpha_cmb);
        double ex = p2m_exp(x0new);
        Gx = ((x0new * (ex + 1.0)) / (ex - 1.0)) - 4.0;
    }
    double mht = p2m_exp(p2m_log(A.mhmint) + (double)(j - 1) * A.dmht);
    double mht3 = p2m_pow(mht, 1.0 / 3.0);
    const int mi = A.mi;
    double P0 = mt_fit(A.b.P0_Am[mi], A.b.P0_alpham[mi], A.b.P0_alphaz[mi], mht, z);
    double xc = mt_fit(A.b.xc_Am[mi], A.b.xc_alpham[mi], A.b.xc_alphaz[mi], mht, z);
    double beta = mt_fit(A.b.beta_Am[mi], A.b.beta_alpham[mi], A.b.beta_alphaz[mi], mht, z);
    double P0r = mt_fit(A.b.P0_Am_rho[mi], A.b.P0_alpham_rho[mi], A.b.P0_alphaz_rho[mi], mht, z);
    double alr = mt_fit(A.b.alpha_Am_rho[mi], A.b.alpha_alpham_rho[mi], A.b.alpha_alphaz_rho[mi], mht, z);
    double betar = mt_fit(A.b.beta_Am_rho[mi], A.b.beta_alpham_rho[mi], A.b.beta_alphaz_rho[mi], mht, z);
    double chi = mt_lookup(A.rofz, dz_r, z);
    double rvir = p2m_pow(((((3.0 * mht) / 4.0) / COSMO_PI) / A.Delta) / rhoc, 1.0 / 3.0);
    double rt = p2m_exp(p2m_log(A.rmint) + (double)(k - 1) * A.drt);
    double theta = p2m_atan(rt / chit);
    double b = (chi * p2m_sin(theta)) / rvir;
    const double rmaxp = 4.0;
    double yint = 0.0, rint = 0.0;
    if (!(b > rmaxp)) {
        int nn = (int)(rmaxp / MT_DS) + 1;
        double y = 0.0;
        for (int q = 1; q <= nn; ++q) {
            double s = ((double)q - 0.5) * MT_DS;
            double x = p2m_sqrt(s * s + b * b);
            double xx = x / xc;
            y = y + (P0 * p2m_pow(xx, MT_GAMMA)) * p2m_pow(1.0 + xx, -beta);
        }
        yint = y * MT_DS;
        double s0 = p2m_sqrt(rmaxp * rmaxp - b * b);
        nn = (int)(s0 / MT_DS) + 1;
        double rho = 0.0;
        for (int q = 1; q <= nn; ++q) {
            double s = ((double)q - 0.5) * MT_DS;
            double x = p2m_sqrt(s * s + b * b);
            double xx = x / 0.5;
            rho = rho + (P0r * p2m_pow(xx, -1.0 * MT_GAMMA_RHO)) *
                        p2m_pow(1.0 + p2m_pow(xx, alr), (-1.0 * (betar - MT_GAMMA_RHO)) / alr);
        }
        rint = 2.0 * (rho * MT_DS);
    }
    double y = ((A.y0 * mht) * zfacy) * yint;
    if (A.nu > 0.0) y = y * Gx;
    double tau = ((A.tau0 * mht3) * zfact) * rint;
    int64_t idx = 2 * ((int64_t)(k - 1) + (int64_t)A.nrt * ((int64_t)(j - 1) + (int64_t)A.nmht * (int64_t)(i - 1)));
    A.table[idx] = (float)y;
    A.table[idx + 1] = (float)tau;
}

static double mt_dzdr(double h, double om, double z)
{
    double z1 = 1.0 + z;
    return ((h * 100.0) / (2.99792458e10 / 1e5)) * p2m_sqrt((om * ((z1 * z1) * z1) + 1.0) - om);
}

extern "C" int p2m_make_maptable(int32_t device, int32_t model, double nu, double alpha_cmb,
                                 int32_t nchit, int32_t nmht, int32_t nrt,
                                 int32_t hdr_i[3], float hdr_f[6], float *table, int64_t n_table)
{
    if (!hdr_i || !hdr_f || !table) return P2M_E_INVALID;
    if (model < 1 || model > 3 || nchit < 2 || nmht < 2 || nrt < 2) return P2M_E_INVALID;
    if (n_table != 2LL * nchit * nmht * nrt) return P2M_E_SIZE;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { (void)hipGetLastError(); return P2M_E_NODEVICE; }
    if (device < 0 || device >= ndev) return P2M_E_INVALID;
    if (hipSetDevice(device) != hipSuccess) return P2M_E_HIP;
    /* make_maptable.f90:33-43 */
    const double omegam = 0.25, omegab = 0.043, omegal = 0.75, h = 0.7;
    const double fb = omegab / omegam;
    const int n = 10000;
    std::vector<double> rofz(n, 0.0), zofr(n, 0.0);
    double dz = 6.0 / (double)(n - 1), dr = 1.3e4 / (double)(n - 1);
    for (int i = 2; i <= n; ++i) rofz[i - 1] = rofz[i - 2] + (1.0 / mt_dzdr(h, omegam, (double)(i - 1) * dz + 0.0)) * dz;
    for (int i = 2; i <= n; ++i) zofr[i - 1] = zofr[i - 2] + mt_dzdr(h, omegam, zofr[i - 2]) * dr;
    /* maptable.f90:56-64 and the cosmology.f90:13-33 parameters */
    const double sigmat = 6.65246e-25, me = 9.109383e-28, mue = (double)1.136f, mp = (double)1.67e-24f;
    const double msun = 1.9891e33, mpc2km = 3.08567758e19, mpc = 3.08567758e24, kpc = 3.08567758e21;
    const double cc = 2.99792458e10, h100 = 100.0 / mpc2km;
    MTArgs A;
    A.b = kBBPS;
    A.mi = model - 1;
    A.nchit = nchit; A.nmht = nmht; A.nrt = nrt;
    A.nu = nu; A.alpha_cmb = alpha_cmb;
    A.Delta = model == 2 ? 500.0 : 200.0;
    const double m2rfac = p2m_pow(((3.0 / 4.0) / COSMO_PI) / A.Delta, 1.0 / 3.0);
    const double ypermass = ((((((3.0 * sigmat) * (h100 * h100)) / 8.0) / COSMO_PI) / me) / (cc * cc)) * msun / (double)1.932f;
    const double taupersigma = ((((sigmat / mue) / mp) * msun) / (kpc * kpc * kpc)) * mpc * 1e10;
    A.y0 = ((fb * A.Delta) * (h * h)) * ypermass;
    A.tau0 = (taupersigma * (h * h)) * m2rfac;
    A.x0const = 0.017612 * nu;
    A.omegam = omegam; A.omegal = omegal; A.rhocrit_0 = 2.775e11 * (0.7 * 0.7);
    /* maptable.f90:69-87; bounds as the float32 values of the header */
    const double chimint = (double)1e2f, chimaxt = (double)7e3f, mhmint = (double)1e13f, mhmaxt = (double)1e16f;
    const double rmint = (double)1e-2f, rmaxt = (double)10.0f;
    A.chimint = chimint; A.mhmint = mhmint; A.rmint = rmint;
    A.dchit = (p2m_log(chimaxt) - p2m_log(chimint)) / (double)(nchit - 1);
    A.dmht = (p2m_log(mhmaxt) - p2m_log(mhmint)) / (double)(nmht - 1);
    A.drt = (p2m_log(rmaxt) - p2m_log(rmint)) / (double)(nrt - 1);
    hdr_i[0] = nchit; hdr_i[1] = nmht; hdr_i[2] = nrt;
    hdr_f[0] = (float)chimint; hdr_f[1] = (float)chimaxt; hdr_f[2] = (float)mhmint;
    hdr_f[3] = (float)mhmaxt; hdr_f[4] = (float)rmint; hdr_f[5] = (float)rmaxt;
    double *d_r = nullptr, *d_z = nullptr;
    float *d_t = nullptr;
    int rc = 0;
    if (hipMalloc(&d_r, sizeof(double) * n) != hipSuccess || hipMalloc(&d_z, sizeof(double) * n) != hipSuccess ||
        hipMalloc(&d_t, sizeof(float) * n_table) != hipSuccess) {
        rc = P2M_E_NOMEM;
    }
    if (!rc && (hipMemcpy(d_r, rofz.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_z, zofr.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess))
        rc = P2M_E_HIP;
    if (!rc) {
        A.rofz = d_r; A.zofr = d_z; A.table = d_t;
        int64_t total = (int64_t)nchit * nmht * nrt;
        hipLaunchKernelGGL(k_make_maptable, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, 0, A);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = P2M_E_HIP;
    }
    if (!rc && hipMemcpy(table, d_t, sizeof(float) * n_table, hipMemcpyDeviceToHost) != hipSuccess) rc = P2M_E_HIP;
    if (d_r) (void)hipFree(d_r);
    if (d_z) (void)hipFree(d_z);
    if (d_t) (void)hipFree(d_t);
    (void)hipGetLastError();
    return rc;
}
