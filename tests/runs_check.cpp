/* Host check of disc_ring_run_fast (p2m_dev.h): the FP32-arc ring run with
 * its FP64 re-decide band returns exactly the run of the FP64 evaluation
 * disc_ring_run_z (pixels_on_edge + discedge2fulldisc) for random discs of
 * radius 1e-6..2 rad, centres anywhere (poles and equator included), every
 * ring of the disc's latitude band at nside 1..8192.  Also counts how often
 * the band sends a ring to the FP64 path. */
#include "p2m_dev.h"
#include <cstdio>
#include <cmath>
#include <random>

struct H { double cosang, zz0, a, phi0; };

static double ring_z_host(int32_t nside, int32_t ir)
{
    const double fn = nside;
    if (ir < nside) return 1.0 - ((double)ir * ir) / ((3.0 * fn) * fn);
    if (ir < 3 * nside) return ((double)(2 * nside - ir) * 2.0) / (3.0 * fn);
    const double t = 4 * nside - ir;
    return -1.0 + (t * t) / ((3.0 * fn) * fn);
}

int main()
{
    std::mt19937_64 rng(2024);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long n = 0, bad = 0, slow = 0;
    for (int d = 0; d < 20000; ++d) {
        const int32_t nside = 1 << (int)(U(rng) * 14);
        double z0 = 2.0 * U(rng) - 1.0;
        if (d % 50 == 0) z0 = (d % 100 == 0) ? 1.0 : -1.0;          /* polar centres */
        const double ph = 2.0 * M_PI * U(rng) - M_PI;
        const double st = std::sqrt(std::max(0.0, 1.0 - z0 * z0));
        const double rad = std::exp(std::log(1e-6) + U(rng) * (std::log(2.0) - std::log(1e-6)));
        H h;
        h.cosang = p2m_cos(rad);
        h.zz0 = z0;
        const double x0 = st * std::cos(ph), y0 = st * std::sin(ph);
        h.a = x0 * x0 + y0 * y0;
        h.phi0 = (x0 != 0.0 || y0 != 0.0) ? p2m_atan2(y0, x0) : 0.0;
        const double rl0 = p2m_asin(z0);
        const double zmax = rl0 + rad >= HPX_HALFPI ? 1.0 : p2m_sin(rl0 + rad);
        const double zmin = rl0 - rad <= -HPX_HALFPI ? -1.0 : p2m_sin(rl0 - rad);
        int32_t r0 = (int32_t)std::max(1.0, std::floor(nside * std::sqrt(std::max(0.0, 3.0 * (1.0 - zmax)))) - 2);
        int32_t r1 = 4 * nside - 1;
        const double zz = zmin;
        int32_t rb = zz < -2.0 / 3.0 ? 4 * nside - (int32_t)std::floor(nside * std::sqrt(3.0 * (1.0 + zz))) + 2
                                     : (int32_t)(nside * (2.0 - 1.5 * zz)) + 2;
        if (rb < r1) r1 = rb;
        if (zmax <= 2.0 / 3.0) r0 = std::max(1, (int32_t)(nside * (2.0 - 1.5 * zmax)) - 2);
        for (int32_t ring = r0; ring <= r1; ring += 1 + (r1 - r0) / 200) {
            const int32_t m = std::min(std::min(nside, ring), 4 * nside - ring);
            const int32_t npr = 4 * m;
            int32_t kshift = (ring + 1) % 2;
            if (nside == 1) kshift = 1 - kshift;
            if (npr < 4 * nside) kshift = 1;
            const double z = ring_z_host(nside, ring);
            int32_t lo_e = -1, lo_f = -1;
            const int32_t np_e = disc_ring_run_z(h, z, npr, kshift, lo_e);
            const int32_t np_f = disc_ring_run_fast(h, z, npr, kshift, lo_f);
            ++n;
            if (np_e != np_f || (np_e > 0 && lo_e != lo_f)) {
                if (++bad < 10) printf("nside %d ring %d rad %.3e: exact (%d,%d) fast (%d,%d)\n", nside, ring, rad, lo_e, np_e, lo_f, np_f);
            }
            /* how often the band decides in FP64 */
            const double b = h.cosang - z * h.zz0;
            double cc = 1.0 - z * z; if (cc < 1e-12) cc = 1e-12;
            if (h.a != 0.0) {
                const double cd = b / p2m_sqrt(h.a * cc);
                if (p2m_fabs(cd) <= 1.0) {
                    const double dp = (double)(2.0f * asinf(sqrtf((float)(0.5 * (1.0 - cd)))));
                    const double k = npr / HPX_TWOPI, sh = kshift * 0.5;
                    const double tl = (h.phi0 - dp) * k - sh, th = (h.phi0 + dp) * k - sh;
                    const double eps = k * (1e-6 * dp + 1e-12) + 1e-9;
                    const double fl = tl - std::floor(tl), fh = th - std::floor(th);
                    if (1.0 + cd < 1e-5 || fl < eps || fl > 1 - eps || fh < eps || fh > 1 - eps) ++slow;
                }
            }
        }
    }
    printf("checked %ld rings, %ld mismatches, %ld decided in FP64\n", n, bad, slow);
    return bad != 0;
}
