/*
 * p2m_kernels.hip -- gfx950 kernels of the halo->map painter.
 *
 *  k_cull, k_setup   catalogue decode: radial and mass cuts over every record,
 *                    then the FP64 per-halo setup of the survivors
 *                    (pks2map.f90:144-160, 264, 284-298; haloproject.f90:47-55,
 *                    93-128; maptable.f90:184-202; query_disc prologue)
 *  k_gather_halos    the survivors in spatial order (band of rings x azimuth),
 *                    packed for the binning and the FP32 painter, with each
 *                    halo's (halo, ring) item count
 *  k_ring_counts (a pass with no painted halo), k_group_heads, k_seg_count,
 *  k_seg_emit
 *                    bin the (halo, ring) pixel runs of query_disc into ring
 *                    tiles of W pixels (pix_tools.F90:712-800,
 *                    pixel_routines.F90:1823-1879)
 *  k_tile_fast       FP32 painter: workgroups take ring tiles from a queue per
 *                    XCD; per pixel the FP32 chord^2 test and profile value,
 *                    summed in an FP64 LDS tile and written once as FP32;
 *                    pixels near a disc edge or a table-cell edge go to
 *  k_fallback        the exact FP64 decision and value (haloproject.f90:60-75,
 *                    maptable.f90:176-216)
 *  k_tile_exact      FP64 painter over the same tiles, the reference's
 *                    arithmetic per pixel, FP64 LDS tile
 *  k_paint_plain_hp  baseline: one work item per halo, naive query_disc loop,
 *                    pix2vec_ring/angdist/interpolate_table per pixel, FP64
 *                    global atomics (haloproject.f90:55-75)
 *  k_fs_counts, k_fs_paint, k_paint_fs
 *                    flat-sky painter over the flattened (halo, box pixel)
 *                    space; one work item per halo for the baseline
 *                    (haloproject.f90:136-152)
 *  k_moments, k_dpdy_cumulative
 *                    map moments and the dP/dy histogram (pks2map.f90:321-357)
 */
#include "p2m_kernels.h"
#include "p2m_pieces.h"

#include <hipcub/hipcub.hpp>
#include <type_traits>

/* ===================================================================== */
/* helpers                                                                */
/* ===================================================================== */

__device__ static inline double dbl_of(uint32_t lo, uint32_t hi)
{
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* Loaded words that the code below needs on every path: an empty asm that
 * uses them, so that the compiler issues their loads together here instead
 * of moving each behind the branch that first uses it (a memory round trip
 * per branch). */
__device__ static inline void pin4(uint4 v)
{
    asm volatile("" ::"v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
}
__device__ static inline void pin2(uint2 v)
{
    asm volatile("" ::"v"(v.x), "v"(v.y));
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0ULL;
}
__global__ void k_fill_f64(double *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0;
}
__global__ void k_fill_f32(float *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0.0f;
}

static inline int fill_grid(int64_t n) { int64_t g = (n + 255) / 256; return (int)(g < 16384 ? (g < 1 ? 1 : g) : 16384); }

void p2m_launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s) { if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(256), 0, s, p, n, v); }
void p2m_launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t s) { if (n > 0) hipLaunchKernelGGL(k_fill_u64, dim3(fill_grid(n)), dim3(256), 0, s, p, n); }
void p2m_launch_fill_f64(double *p, int64_t n, hipStream_t s) { if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(fill_grid(n)), dim3(256), 0, s, p, n); }
__global__ void k_clear_ctr(unsigned long long *ctr, int first, int count)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < P2M_SHARDS * count) ctr[(t / count) * P2M_CTR_STRIDE + first + t % count] = 0ULL;
}

void p2m_launch_clear_ctr(unsigned long long *ctr, int first, int count, hipStream_t s)
{
    if (count > 0) hipLaunchKernelGGL(k_clear_ctr, dim3((P2M_SHARDS * count + 255) / 256), dim3(256), 0, s, ctr, first, count);
}

void p2m_launch_fill_f32(float *p, int64_t n, hipStream_t s) { if (n > 0) hipLaunchKernelGGL(k_fill_f32, dim3(fill_grid(n)), dim3(256), 0, s, p, n); }

/* wave-aggregated append: returns the slot of this lane (or garbage if !take) */
__device__ static inline unsigned long long wave_append(unsigned long long *counter, bool take)
{
    unsigned long long mask = __ballot(take);
    int lane = __lane_id();
    int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (mask == 0) return 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    unsigned long long below = lane == 0 ? 0ULL : (mask & ((~0ULL) >> (64 - lane)));
    return base + (unsigned long long)__popcll(below);
}

__device__ static inline void wave_count(unsigned long long *counter, unsigned long long v)
{
    /* sum v over the wave, one atomic into this wave's shard of a statistics counter */
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const unsigned shard = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) % P2M_SHARDS;
    if (__lane_id() == 0 && v) atomicAdd(counter + shard * P2M_CTR_STRIDE, v);
}

/* sum v over the workgroup, one atomic per workgroup into its shard of a
 * statistics counter; every thread calls */
__device__ static inline void block_count(unsigned long long *counter, unsigned long long v, unsigned long long *lds16)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    int wv = threadIdx.x >> 6;
    __syncthreads();
    if (__lane_id() == 0) lds16[wv] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += lds16[k];
        if (t) atomicAdd(counter + (blockIdx.x % P2M_SHARDS) * P2M_CTR_STRIDE, t);
    }
}

/* Inclusive prefix sum over the 64 lanes of a wave in registers: DPP row
 * shifts by 1, 2, 4 and 8 scan each row of 16 lanes, then row broadcasts of
 * lane 15 (into rows 1 and 3) and lane 31 (into rows 2 and 3) add the
 * preceding rows' totals.  Lanes whose DPP source lies outside the row, and
 * the rows a broadcast masks off, add 0.  Every lane of the wave must be
 * active. */
__device__ static inline uint32_t wave_incl_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   /* row_shr:1 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   /* row_shr:2 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   /* row_shr:4 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   /* row_shr:8 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   /* row_bcast:15 */
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   /* row_bcast:31 */
    return v;
}

/* ===================================================================== */
/* decode                                                                 */
/* ===================================================================== */

#define CULL_SPAN 8192
#define CULL_U 4                   /* records per thread in flight in k_cull */

/* Cull pass over every record: the radial cut (pks2map.f90:144-160) and the
 * mass cut on the unconverted mass (:264, :296).  A workgroup takes spans of
 * CULL_SPAN consecutive records, collects the survivors' chunk indices in
 * LDS (wave-aggregated LDS atomics), then reserves their range with one
 * global atomic per span and copies them out coalesced.  The FP64 setup
 * then runs on full waves of survivors. */
__global__ void __launch_bounds__(256) k_cull(DecodeArgs A)
{
    const CosmoDev &c = A.c;
    __shared__ uint32_t buf[CULL_SPAN];
    __shared__ unsigned int nbuf;
    __shared__ unsigned long long base_s;
    __shared__ unsigned long long lds16[16];
    const double mconst = ((4.0 / 3.0) * c.mypi) * c.rho_0;
    unsigned long long tot_cut = 0;
    const int64_t nspan = (A.n + CULL_SPAN - 1) / CULL_SPAN;
    const int lane = __lane_id();
    const unsigned long long below = lane == 0 ? 0ULL : ((~0ULL) >> (64 - lane));
    for (int64_t sp = blockIdx.x; sp < nspan; sp += gridDim.x) {
        if (threadIdx.x == 0) nbuf = 0;
        __syncthreads();
        const int64_t i0 = sp * CULL_SPAN;
        /* CULL_U records per thread per step, all four fields of each loaded
         * before any is used: the loads of a step are in flight together */
        for (int kb = 0; kb < CULL_SPAN; kb += CULL_U * (int)blockDim.x) {
            float r0[CULL_U], r1[CULL_U], r2[CULL_U], r6[CULL_U];
#pragma unroll
            for (int u = 0; u < CULL_U; ++u) {
                const int64_t i = i0 + kb + u * (int)blockDim.x + (int)threadIdx.x;
                const float *p = A.halos + 7 * (i < A.n ? i : 0);
                r0[u] = p[0]; r1[u] = p[1]; r2[u] = p[2]; r6[u] = p[6];
            }
#pragma unroll
            for (int u = 0; u < CULL_U; ++u) {
                const int64_t i = i0 + kb + u * (int)blockDim.x + (int)threadIdx.x;
                bool take = false;
                if (i < A.n && kb + u * (int)blockDim.x < CULL_SPAN) {
                    const double xh = (double)r2[u], yh = (double)r1[u], zh = (double)r0[u];
                    const double xs = xh - A.chihview;
                    const double chih = p2m_sqrt((xs * xs + yh * yh) + zh * zh);
                    if (A.precut || !(chih > A.cut_high || chih < A.cut_low)) {
                        tot_cut++;
                        const double mass = mconst * cube((double)r6[u]);
                        take = !(mass < A.mmin);
                    }
                }
                const unsigned long long mask = __ballot(take);
                if (mask) {
                    unsigned int off = 0;
                    if (lane == 0) off = atomicAdd(&nbuf, (unsigned int)__popcll(mask));
                    off = __shfl(off, 0);
                    if (take) buf[off + (unsigned int)__popcll(mask & below)] = (uint32_t)i;
                }
            }
        }
        __syncthreads();
        const unsigned int cnt = nbuf;
        if (threadIdx.x == 0 && cnt) base_s = atomicAdd(A.n_cand, (unsigned long long)cnt);
        __syncthreads();
        for (unsigned int k = threadIdx.x; k < cnt; k += blockDim.x) A.cand[base_s + k] = buf[k];
        __syncthreads();
    }
    block_count(&A.counters[0], tot_cut, lds16);
}

/* FP64 per-halo setup of the survivors: redshift, mass conversion, disc
 * radius (pks2map.f90:284-298, haloproject.f90:47), the query_disc and
 * discphirange_at_z prologues (pixel_routines.F90:1981-2010,
 * pix_tools.F90:729-738), the per-halo part of interpolate_table
 * (maptable.f90:184-202) and the flat-sky box (haloproject.f90:101-128). */
__global__ void __launch_bounds__(256) k_setup(DecodeArgs A)
{
    const CosmoDev &c = A.c;
    const TabGeom &g = A.g;
    __shared__ unsigned long long lds16[16];
    unsigned long long tot_rad = 0, tot_rings = 0;
    const unsigned long long kbeg = *A.n_cand0, kend = *A.n_cand;
    for (unsigned long long k0 = kbeg + blockIdx.x * (unsigned long long)blockDim.x; k0 < kend;
         k0 += (unsigned long long)gridDim.x * blockDim.x) {
        const unsigned long long k = k0 + threadIdx.x;
        bool painted = false;
        HaloRec h;
        if (k < kend) {
            const int64_t i = A.cand[k];
            const float *p = A.halos + 7 * i;
            const float p0 = p[0], p1 = p[1], p2 = p[2], p6 = p[6];
            const double xh = (double)p2, yh = (double)p1, zh = (double)p0;
            const double xs = xh - A.chihview;
            const double chih = p2m_sqrt((xs * xs + yh * yh) + zh * zh);
            const double mass = (((4.0 / 3.0) * c.mypi) * c.rho_0) * cube((double)p6);   /* pks2map.f90:264 */
            double vh = 1.0;
            if (A.profile == 2) {
                double px = p0, py = p1, pz = p2;
                double vr = (px * (double)p[3] + py * (double)p[4]) + pz * (double)p[5];
                vr = vr / p2m_sqrt((px * px + py * py) + pz * pz);
                vh = -vr / 3.0e5;
            }
            const double chi_pre = p2m_sqrt((xh * xh + yh * yh) + zh * zh);          /* :289 */
            const double redshift = tab_lookup(c.zofr, c.nzofr, c.zofr_min, c.dzofr, chi_pre);  /* :290 */
            const double mh = p2m_sqrt(deltacrit_z(c, redshift) / A.dmodel) * mass;    /* :295 */
            const double rh = p2m_pow(((((3.0 * mh) / 4.0) / c.mypi) / 200.0) / rhocrit_z(c, redshift), 1.0 / 3.0); /* :298 */
            const double dtheta = 4.0 * p2m_asin(rh / chih);                      /* haloproject.f90:47 */
            if (!(dtheta >= 0.0 && dtheta <= HPX_PI)) {
                tot_rad++;   /* pixel_routines.F90:1962 would abort */
                /* an inert record: no rings, no flat-sky box */
                h = HaloRec();
                h.dtheta = -1.0; h.irmin = 0; h.irmax = -1; h.flat = 0;
                h.cat = A.index_base + i;
            } else {
                painted = true;
                const double ux = xs / chih, uy = yh / chih, uz = zh / chih;
                h.hx = uz; h.hy = uy; h.hz = ux;
                h.dtheta = dtheta;
                h.vh = vh;
                h.cat = A.index_base + i;
                const double norm = p2m_sqrt((h.hx * h.hx + h.hy * h.hy) + h.hz * h.hz);
                const double z0 = h.hz / norm;
                h.irmin = 0; h.irmax = -1;
                if (A.nside > 0) {
                    const double rlat0 = p2m_asin(z0);
                    const double rlat1 = rlat0 + dtheta, rlat2 = rlat0 - dtheta;
                    const double zmax = (rlat1 >= HPX_HALFPI) ? 1.0 : p2m_sin(rlat1);
                    const int32_t irmin = ring_num(A.nside, zmax);
                    h.irmin = irmin - 1 > 1 ? irmin - 1 : 1;
                    const double zmin = (rlat2 <= -HPX_HALFPI) ? -1.0 : p2m_sin(rlat2);
                    const int32_t irmax = ring_num(A.nside, zmin);
                    h.irmax = irmax + 1 < 4 * A.nside - 1 ? irmax + 1 : 4 * A.nside - 1;
                }
                h.cosang = p2m_cos(dtheta);
                const double x0 = h.hx / norm, y0 = h.hy / norm;
                h.zz0 = h.hz / norm;
                h.phi0 = 0.0;
                if (x0 != 0.0 || y0 != 0.0) h.phi0 = p2m_atan2(y0, x0);
                h.a = x0 * x0 + y0 * y0;
                h.s0 = p2m_sqrt(h.a);
                const double sh = p2m_sin(0.5 * dtheta);
                h.thr = 4.0 * (sh * sh);
                double chi = tab_lookup(c.rofz, c.nrofz, c.rofz_min, c.drofz, redshift);
                h.chi_r = chi;
                h.chi_out = chi > g.chimaxt ? 1 : 0;
                if (chi < g.chimint) chi = g.chimint + 1e-5;
                double mc = mh;
                if (mc < g.mhmint) mc = g.mhmint + 1e-5;
                if (mc > g.mhmaxt) mc = g.mhmaxt - 1e-5;
                const double lc = p2m_log(chi), lm = p2m_log(mc);
                h.ichi = (int)((lc - g.lc0) / g.dchit) + 1;
                h.imh = (int)((lm - g.lm0) / g.dmht) + 1;
                h.fc = lc - (g.lc0 + (double)(h.ichi - 1) * g.dchit);
                h.fm = lm - (g.lm0 + (double)(h.imh - 1) * g.dmht);
                h.flat = 0; h.th = 0.0; h.ph = 0.0;
                h.jmin = h.jmax = h.kmin = h.kmax = 0;
                if (A.npix_fs > 0 && !(uz < 0.0)) {
                    const double th = p2m_asin(ux), ph = p2m_asin(uy);
                    h.th = th; h.ph = ph;
                    double thetamin = th - 1.1 * dtheta, thetamax = th + 1.1 * dtheta;
                    double phimin = ph - 1.1 * dtheta, phimax = ph + 1.1 * dtheta;
                    const double hf = A.fov / 2.0;
                    if (!(thetamax < -hf || thetamin > hf) && !(phimax < -hf || phimin > hf)) {
                        if (thetamin <= -hf) thetamin = -hf + 1e-10;
                        if (thetamax >= hf) thetamax = hf - 1e-10;
                        if (phimin <= -hf) phimin = -hf + 1e-10;
                        if (phimax >= hf) phimax = hf - 1e-10;
                        int jmin = (int)((thetamin + hf) / A.dpix) + 1; if (jmin < 1) jmin = 1;
                        int jmax = (int)((thetamax + hf) / A.dpix) + 1; if (jmax > A.npix_fs) jmax = A.npix_fs;
                        int kmin = (int)((phimin + hf) / A.dpix) + 1; if (kmin < 1) kmin = 1;
                        int kmax = (int)((phimax + hf) / A.dpix) + 1; if (kmax > A.npix_fs) kmax = A.npix_fs;
                        if (jmax >= jmin && kmax >= kmin) {
                            h.flat = 1; h.jmin = jmin; h.jmax = jmax; h.kmin = kmin; h.kmax = kmax;
                        }
                    }
                }
            }
        }
        const bool flat = k < kend && painted && h.flat;
        /* (halo, ring) items of this record, as the binning counts them */
        if (k < kend && k < (unsigned long long)A.cap) tot_rings += h.irmax >= h.irmin ? (unsigned long long)(h.irmax - h.irmin + 1) : 1ULL;
        if (k < kend && k < (unsigned long long)A.cap) {
            A.out[k] = h;
            if (A.keys) {
                /* spatial key: band of 8 rings of the disc centre, then its azimuth bin */
                uint32_t key = A.key_max;
                if (painted && A.nside > 0) {
                    const int32_t rc = ring_num(A.nside, h.zz0);
                    double ph = h.phi0;
                    if (ph < 0.0) ph += HPX_TWOPI;
                    int32_t pb = (int32_t)(ph * (double)A.nphi / HPX_TWOPI);
                    pb = pb < 0 ? 0 : (pb >= A.nphi ? A.nphi - 1 : pb);
                    key = (uint32_t)((rc - 1) >> 3) * (uint32_t)A.nphi + (uint32_t)pb;
                }
                A.keys[k] = key;
                A.vals[k] = (uint32_t)k;
            }
        }
        const unsigned long long fs = wave_append(A.n_flat_out, flat);
        if (flat) A.flat_idx[fs] = (uint32_t)k;
    }
    block_count(&A.counters[2], tot_rad, lds16);
    block_count(&A.counters[3], tot_rings, lds16);
}

/* n_cand0 <- n_cand: the first candidate slot of the next chunk */
__global__ void k_mark_cand(DecodeArgs A)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *A.n_cand0 = *A.n_cand;
}

void p2m_launch_decode(const DecodeArgs &a, hipStream_t s)
{
    if (a.n <= 0) return;
    int64_t gc = (a.n + CULL_SPAN - 1) / CULL_SPAN;
    if (gc > 2048) gc = 2048;
    int64_t g = (a.n + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_mark_cand, dim3(1), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_cull, dim3((unsigned)gc), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_setup, dim3((unsigned)g), dim3(256), 0, s, a);
}

/* ===================================================================== */
/* baseline: plain FP64 painter, one work item per halo                   */
/* ===================================================================== */

__global__ void __launch_bounds__(64) k_paint_plain_hp(PaintArgs A)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    unsigned long long n_list = 0, n_upd = 0, n_gaps = 0;
    if (i < A.n_halo) {
        HaloRec h = A.halos[i];
        const int32_t nside = A.nside;
        int last = 0, pend = 0;
        for (int32_t ring = h.irmin; ring <= h.irmax; ++ring) {
            int32_t npr, kshift; int64_t first;
            ring_info(nside, ring, npr, kshift, first);
            int32_t lo = 0;
            int32_t np = disc_ring_run(h, nside, ring, npr, kshift, lo);
            if (np == 0) { if (last) pend = 1; continue; }
            if (pend) { n_gaps++; pend = 0; }
            last = 1;
            n_list += np;
            for (int32_t k = 0; k < np; ++k) {
                int32_t ip = lo + k; if (ip >= npr) ip -= npr;
                int64_t pix = first + ip;
                double v2[3];
                pix2vec_ring(nside, pix, v2);
                double angle = angdist(h.hx, h.hy, h.hz, v2[0], v2[1], v2[2]);
                if (angle < h.dtheta) {
                    n_upd++;
                    for (int m = 0; m < A.n_maps; ++m) {
                        double val = h.vh * interp_exact(h, A.g, A.tab_flat[m], A.n_flat, A.profile, angle);
                        atomicAdd(&A.map64[m][pix], val);
                    }
                    if (A.count) {
                        atomicAdd(&A.count[pix], 1u);
                        atomicAdd(&A.hash[pix], (unsigned long long)halo_hash(h.cat));
                    }
                }
            }
        }
    }
    wave_count(&A.counters[0], n_list);
    wave_count(&A.counters[1], n_upd);
    wave_count(&A.counters[3], n_gaps);
}

void p2m_launch_paint_plain_hp(const PaintArgs &a, hipStream_t s)
{
    if (a.n_halo <= 0) return;
    hipLaunchKernelGGL(k_paint_plain_hp, dim3((unsigned)((a.n_halo + 63) / 64)), dim3(64), 0, s, a);
}

/* ===================================================================== */
/* flat-sky painter: one wave per halo, FP64, global atomics              */
/* ===================================================================== */

__global__ void __launch_bounds__(256) k_paint_fs(PaintArgs A, int plain)
{
    unsigned long long n_upd = 0;
    if (plain) {
        /* baseline: one work item per halo */
        const int64_t hidx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
        if (hidx < A.n_halo) {
            const HaloRec &hr = A.halos[hidx];
            if (hr.flat) {
                HaloRec h = hr;
                double hf = A.fov / 2.0;
                for (int k = h.kmin; k <= h.kmax; ++k) {
                    double phip = -hf + ((double)k - 0.5) * A.dpix;
                    for (int j = h.jmin; j <= h.jmax; ++j) {
                        double thetap = -hf + ((double)j - 0.5) * A.dpix;
                        double dphi = phip - h.ph, dth = thetap - h.th;
                        double angle = p2m_sqrt(dphi * dphi + dth * dth);
                        if (angle < h.dtheta) {
                            int64_t pix = (int64_t)(j - 1) + (int64_t)A.npix_fs * (k - 1);
                            n_upd++;
                            for (int m = 0; m < A.n_maps; ++m)
                                atomicAdd(&A.fs64[m][pix], h.vh * interp_exact(h, A.g, A.tab_flat[m], A.n_flat, A.profile, angle));
                            if (A.fs_count) {
                                atomicAdd(&A.fs_count[pix], 1u);
                                atomicAdd(&A.fs_hash[pix], (unsigned long long)halo_hash(h.cat));
                            }
                        }
                    }
                }
            }
        }
        wave_count(&A.counters[2], n_upd);
        return;
    }
}

/* box pixels of each listed flat-sky halo (haloproject.f90:124-135) */
__global__ void __launch_bounds__(256) k_fs_counts(PaintArgs A)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= A.n_flat_host; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t n = 0;
        if (i < A.n_flat_host) {
            const HaloRec &h = A.halos[A.flat_idx[i]];
            n = (uint32_t)((h.jmax - h.jmin + 1) * (h.kmax - h.kmin + 1));
        }
        A.fs_cnt[i] = n;
    }
}

/* Flat-sky painter over the flattened (halo, box pixel) space: every wave
 * takes 64 consecutive box pixels, so a few large boxes do not serialise
 * on one wave.  FP64 angle and value (haloproject.f90:130-150), global
 * FP64 atomics into the map. */
__global__ void __launch_bounds__(256) k_fs_paint(PaintArgs A)
{
    const int lane = __lane_id();
    const int64_t nfl = A.n_flat_host;
    const unsigned long long total = A.fs_off[nfl];
    const unsigned long long ngroup = (total + 63) >> 6;
    const unsigned long long nw = ((unsigned long long)gridDim.x * blockDim.x) >> 6;
    const double hf = A.fov / 2.0;
    unsigned long long n_upd = 0;
    for (unsigned long long g = (blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x) >> 6; g < ngroup; g += nw) {
        const unsigned long long c = g << 6;
        int64_t lo = 0, hi = nfl - 1;           /* halo holding pixel c (wave-uniform) */
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if (A.fs_off[mid] <= c) lo = mid; else hi = mid - 1;
        }
        const unsigned long long item = c + lane;
        const int64_t hl = lo + lane;
        const unsigned long long my_off = hl <= nfl ? A.fs_off[hl] : ~0ULL;
        int jj = 0;
        for (int step = 32; step > 0; step >>= 1) {
            const unsigned long long o = __shfl(my_off, jj + step);
            if (o <= item) jj += step;
        }
        const unsigned long long hoff = __shfl(my_off, jj);
        if (item >= total) continue;
        const HaloRec &h = A.halos[A.flat_idx[lo + jj]];
        const int nj = h.jmax - h.jmin + 1;
        const int q = (int)(item - hoff);
        const int k = h.kmin + q / nj, j = h.jmin + q % nj;
        const double phip = -hf + ((double)k - 0.5) * A.dpix;
        const double thetap = -hf + ((double)j - 0.5) * A.dpix;
        const double dphi = phip - h.ph, dth = thetap - h.th;
        const double angle = p2m_sqrt(dphi * dphi + dth * dth);
        if (angle < h.dtheta) {
            const int64_t pix = (int64_t)(j - 1) + (int64_t)A.npix_fs * (k - 1);
            n_upd++;
            for (int m = 0; m < A.n_maps; ++m)
                atomicAdd(&A.fs64[m][pix], h.vh * interp_exact(h, A.g, A.tab_flat[m], A.n_flat, A.profile, angle));
            if (A.fs_count) {
                atomicAdd(&A.fs_count[pix], 1u);
                atomicAdd(&A.fs_hash[pix], (unsigned long long)halo_hash(h.cat));
            }
        }
    }
    wave_count(&A.counters[2], n_upd);
}

void p2m_launch_paint_fs(const PaintArgs &a, int plain, hipStream_t s)
{
    if (a.n_halo <= 0 || a.npix_fs <= 0) return;
    if (plain) hipLaunchKernelGGL(k_paint_fs, dim3((unsigned)((a.n_halo + 255) / 256)), dim3(256), 0, s, a, 1);
}

void p2m_launch_fs_counts(const PaintArgs &a, hipStream_t s)
{
    if (a.n_flat_host <= 0 || a.npix_fs <= 0) return;
    int64_t g = (a.n_flat_host + 1 + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_fs_counts, dim3((unsigned)g), dim3(256), 0, s, a);
}

void p2m_launch_fs_paint(const PaintArgs &a, hipStream_t s)
{
    if (a.n_flat_host <= 0 || a.npix_fs <= 0) return;
    hipLaunchKernelGGL(k_fs_paint, dim3(2048), dim3(256), 0, s, a);
}

/* ===================================================================== */
/* binning of (halo, ring) runs into ring tiles                           */
/* ===================================================================== */

__global__ void __launch_bounds__(256) k_ring_counts(BinArgs B)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i <= B.n_halo; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t nr = 0;
        if (i < B.n_halo) {
            const BinHalo &h = B.halos[i];
            nr = h.irmax >= h.irmin ? (uint32_t)(h.irmax - h.irmin + 1) : 1u;
        }
        B.nr[i] = nr;
    }
}

/* Segment slots are assigned per workgroup through an LDS hash table keyed
 * by tile: a chunk of BIN_CH consecutive (halo, ring) items touches few
 * distinct tiles once the halos are in spatial order, so one global atomic
 * per (chunk, tile) reserves the slots of all its pieces.  A piece whose
 * tile finds no hash slot within BIN_PROBE probes takes one global atomic
 * of its own. */
static_assert(offsetof(BinHalo, cosang) == 0 && offsetof(BinHalo, zz0) == 8 && offsetof(BinHalo, a) == 16 &&
              offsetof(BinHalo, phi0) == 24 && offsetof(BinHalo, irmin) == 32 && offsetof(BinHalo, irmax) == 36 &&
              sizeof(BinHalo) == 40, "BinHalo layout read by k_seg_count");

#define BIN_CH 1024
#define BIN_HASH 1024
#define BIN_HASH_BITS 10
#define BIN_PROBE 32
#define BIN_EMPTY 0xffffffffu
#define BIN_DIRECT 0x80000000u

struct BinHash {
    uint32_t key[BIN_HASH];
    uint32_t cnt[BIN_HASH];
};

struct CountLDS : BinHash {
    uint32_t prec[BIN_CH];       /* per item of the chunk: its first piece's slot code */
};

__device__ static inline int bin_slot(BinHash *L, uint32_t tile)
{
    uint32_t h = (tile * 2654435761u) >> (32 - BIN_HASH_BITS);
    for (int k = 0; k < BIN_PROBE; ++k) {
        const uint32_t prev = atomicCAS(&L->key[h], BIN_EMPTY, tile);
        if (prev == BIN_EMPTY || prev == tile) return (int)h;
        h = (h + 1) & (BIN_HASH - 1);
    }
    return -1;
}

__device__ static inline void bin_reset(BinHash *L)
{
    for (int k = threadIdx.x; k < BIN_HASH; k += blockDim.x) { L->key[k] = BIN_EMPTY; L->cnt[k] = 0; }
}

/* the items [start, end) of one wave (start a multiple of 64), with their
 * halo and ring offset: f(item, halo, k) for each item of this lane.  The
 * halo holding item 64g is group_halo[g]; the 64 items of a group span at
 * most 64 halos (every halo has at least one ring), so each lane finds its
 * halo by a 6-step shuffle search over off[h .. h+63]. */
template <typename F>
__device__ static inline void wave_range(const unsigned long long *off, const uint32_t *group_halo, int64_t n_halo,
                                         unsigned long long start, unsigned long long end, F f)
{
    const int lane = __lane_id();
    for (unsigned long long c = start; c < end; c += 64) {
        const int64_t h = group_halo[c >> 6];
        const unsigned long long item = c + lane;
        const int64_t hl = h + lane;
        const unsigned long long my_off = hl <= n_halo ? off[hl] : ~0ULL;
        int j = 0;
        for (int step = 32; step > 0; step >>= 1) {
            unsigned long long o = __shfl(my_off, j + step);
            if (o <= item) j += step;
        }
        const unsigned long long hoff = __shfl(my_off, j);
        if (item < end) f(item, h + j, item - hoff);
    }
}

/* z of every ring centre (ring2z) */
__global__ void k_ring_z(int32_t nside, double *z)
{
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (r <= 4 * nside - 1) z[r - 1] = ring2z(nside, r);
}

void p2m_launch_ring_z(int32_t nside, double *z, hipStream_t s)
{
    const int n = 4 * nside - 1;
    hipLaunchKernelGGL(k_ring_z, dim3((n + 255) / 256), dim3(256), 0, s, nside, z);
}

/* group_halo[g] = the halo holding item 64 g */
__global__ void __launch_bounds__(256) k_group_heads(BinArgs B)
{
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < B.n_halo; h += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long g0 = (B.ring_off[h] + 63) >> 6, g1 = (B.ring_off[h + 1] + 63) >> 6;
        for (unsigned long long g = g0; g < g1; ++g) B.group_halo[g] = (uint32_t)h;
    }
}

/* Per (halo, ring) item: the query_disc pixel run, its tile pieces and the
 * slot of each piece within its tile's segment list.  The first piece's slot
 * code is (hash slot << 16 | rank within the chunk's pieces of that tile)
 * until the chunk reserves its range in the tile, BIN_DIRECT | slot when
 * reserved on its own, BIN_EMPTY for none.  Later pieces (runs crossing a
 * tile boundary, about 8% of the items) take their slot on their own and are
 * listed in ovf with it. */
__device__ static inline uint32_t slot_of(const BinHash &L, uint32_t code)
{
    if (code == BIN_EMPTY) return 0u;
    if (code & BIN_DIRECT) return code & ~BIN_DIRECT;
    return L.cnt[code >> 16] + (code & 0xffffu);
}

/* XCD-aware split of n work units over the grid: workgroup b runs on XCD
 * b % 8, and the workgroups of one XCD take one contiguous eighth of the
 * units, so the units in flight on an XCD are spatial neighbours whose tile
 * lists meet in its L2 (whole lines written before they leave it). */
__device__ static inline void xcd_range(unsigned long long n, unsigned long long &beg, unsigned long long &end,
                                        unsigned long long &step, unsigned long long lane_units, unsigned long long unit)
{
    const bool xcd = (gridDim.x % 8) == 0 && n >= 8;
    const unsigned long long per_wg = lane_units;     /* units a workgroup takes per step */
    if (xcd) {
        const unsigned long long x = blockIdx.x % 8, l = blockIdx.x / 8;
        beg = x * n / 8 + l * per_wg + unit;
        end = (x + 1) * n / 8;
        step = (unsigned long long)(gridDim.x / 8) * per_wg;
    } else {
        beg = (unsigned long long)blockIdx.x * per_wg + unit;
        end = n;
        step = (unsigned long long)gridDim.x * per_wg;
    }
}

/* eight waves per SIMD (64 VGPRs, a few spilled): the pass waits on loads
 * and LDS atomics far more than it computes, so occupancy pays */
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) k_seg_count(BinArgs B)
{
    __shared__ CountLDS L;
    __shared__ unsigned long long lds16[16];
    unsigned long long n_list = 0, n_seg = 0;
    const unsigned long long nchunk = (B.n_items + BIN_CH - 1) / BIN_CH;
    bin_reset(&L);
    __syncthreads();
    const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const unsigned long long per = BIN_CH / nwv;
    unsigned long long cbeg, cend, cstep;
    xcd_range(nchunk, cbeg, cend, cstep, 1, 0);
    for (unsigned long long ch = cbeg; ch < cend; ch += cstep) {
        const unsigned long long c0 = ch * BIN_CH;
        unsigned long long a = c0 + wv * per, e = a + per;
        if (e > B.n_items) e = B.n_items;
        wave_range(B.ring_off, B.group_halo, B.n_halo, a, e, [&](unsigned long long item, int64_t hi, unsigned long long k) {
            const uint2 *hs = reinterpret_cast<const uint2 *>(B.halos + hi);
            const uint2 q0 = hs[0], q1 = hs[1], q2 = hs[2], q3 = hs[3], q4 = hs[4];
            pin2(q0); pin2(q1); pin2(q2); pin2(q3); pin2(q4);
            BinHalo h;
            h.cosang = dbl_of(q0.x, q0.y); h.zz0 = dbl_of(q1.x, q1.y);
            h.a = dbl_of(q2.x, q2.y); h.phi0 = dbl_of(q3.x, q3.y);
            h.irmin = (int32_t)q4.x; h.irmax = (int32_t)q4.y;
            const int32_t ring = h.irmin + (int32_t)k;
            int32_t lo = 0, np = 0;
            int32_t npr, kshift; int64_t first;
            if (ring <= h.irmax) {
                ring_info(B.nside, ring, npr, kshift, first);
                np = disc_ring_run_fast(h, B.ring_z[ring - 1], npr, kshift, lo);
            }
            /* first pixel | count << 16: lo < npr and np <= npr <= 32768 for nside <= 8192 */
            B.runs[item] = (uint32_t)lo | ((uint32_t)np << 16);
            uint32_t code = BIN_EMPTY;
            if (np > 0) {
                n_list += (unsigned long long)np;
                int j = 0;
                run_pieces(lo, np, npr, B.tile_base[ring - 1], B.log2w, [&](int32_t tile, int32_t plo, int32_t len) {
                    const int sl = j == 0 ? bin_slot(&L, (uint32_t)tile) : -1;
                    if (sl >= 0) {
                        code = ((uint32_t)sl << 16) | atomicAdd(&L.cnt[sl], 1u);
                    } else {
                        const uint32_t slot = atomicAdd(&B.tile_cnt[tile], 1u);
                        if (j == 0) {
                            code = BIN_DIRECT | slot;
                        } else {
                            const unsigned long long o = atomicAdd(B.n_ovf, 1ULL);
                            if (o < (unsigned long long)B.ovf_cap)
                                B.ovf[o] = make_uint4((uint32_t)tile, slot, (uint32_t)hi, ((uint32_t)plo << 16) | (uint32_t)len);
                        }
                    }
                    ++j;
                    n_seg++;
                });
            }
            L.prec[item - c0] = code;
        });
        __syncthreads();
        /* reserve each (chunk, tile) range; cnt[] then holds its first slot */
        for (int k = threadIdx.x; k < BIN_HASH; k += blockDim.x)
            if (L.key[k] != BIN_EMPTY) L.cnt[k] = atomicAdd(&B.tile_cnt[L.key[k]], L.cnt[k]);
        __syncthreads();
        const unsigned long long nin = B.n_items - c0 < BIN_CH ? B.n_items - c0 : BIN_CH;
        for (unsigned long long k = threadIdx.x; k < nin; k += blockDim.x) B.pos[c0 + k] = slot_of(L, L.prec[k]);
        __syncthreads();
        bin_reset(&L);
        __syncthreads();
    }
    block_count(&B.counters[0], n_list, lds16);
    block_count(&B.counters[4], n_seg, lds16);
}

/* Scatter of every piece to its slot: segs[tile_off[tile] + slot].  A wave
 * takes 128 consecutive items, two per lane, with the loads of both issued
 * before either is used (the pass is bound by load latency); the pieces
 * after an item's first come from ovf. */
#define EMIT_U 2
__global__ void __launch_bounds__(256) k_seg_emit(BinArgs B)
{
    __shared__ unsigned long long lds16[16];
    unsigned long long n_gaps = 0, n_err = 0;
    const int lane = __lane_id();
    const unsigned long long nunit = (B.n_items + 64 * EMIT_U - 1) / (64 * EMIT_U);
    unsigned long long ubeg, uend, ustep;
    xcd_range(nunit, ubeg, uend, ustep, blockDim.x >> 6, threadIdx.x >> 6);
    for (unsigned long long un = ubeg; un < uend; un += ustep) {
        const unsigned long long base = un * (64 * EMIT_U);
        unsigned long long item[EMIT_U];
        uint32_t run[EMIT_U], ps[EMIT_U];
        int64_t hgrp[EMIT_U];
        unsigned long long off[EMIT_U];
#pragma unroll
        for (int u = 0; u < EMIT_U; ++u) {
            item[u] = base + 64 * u + lane;
            const unsigned long long it = item[u] < B.n_items ? item[u] : B.n_items - 1;
            run[u] = B.runs[it];
            ps[u] = B.pos[it];
            unsigned long long c = base + 64 * u;
            if (c >= B.n_items) c = (B.n_items - 1) & ~63ULL;
            hgrp[u] = B.group_halo[c >> 6];
        }
#pragma unroll
        for (int u = 0; u < EMIT_U; ++u) {
            const int64_t hl = hgrp[u] + lane;
            off[u] = hl <= B.n_halo ? B.ring_off[hl] : ~0ULL;
        }
        int64_t hi[EMIT_U];
        unsigned long long kk[EMIT_U];
#pragma unroll
        for (int u = 0; u < EMIT_U; ++u) {
            int j = 0;
            for (int step = 32; step > 0; step >>= 1) {
                const unsigned long long o = __shfl(off[u], j + step);
                if (o <= item[u]) j += step;
            }
            hi[u] = hgrp[u] + j;
            kk[u] = item[u] - __shfl(off[u], j);
        }
        int32_t tb[EMIT_U], npr[EMIT_U];
#pragma unroll
        for (int u = 0; u < EMIT_U; ++u) {
            const int64_t h = hi[u] < B.n_halo ? hi[u] : B.n_halo - 1;
            int32_t ring = B.halos[h].irmin + (int32_t)kk[u];
            ring = ring < 1 ? 1 : (ring > 4 * B.nside - 1 ? 4 * B.nside - 1 : ring);
            int32_t kshift; int64_t first;
            ring_info(B.nside, ring, npr[u], kshift, first);
            tb[u] = B.tile_base[ring - 1];
        }
        int32_t tile[EMIT_U][2], plo[EMIT_U][2], len[EMIT_U][2];
        bool live[EMIT_U];
        unsigned long long to[EMIT_U], tn[EMIT_U];
#pragma unroll
        for (int u = 0; u < EMIT_U; ++u) {
            const int32_t np = (int32_t)(run[u] >> 16);
            live[u] = item[u] < B.n_items && np != 0;
            if (live[u]) first_two_pieces((int32_t)(run[u] & 0xffffu), np, npr[u], tb[u], B.log2w, tile[u], plo[u], len[u]);
            const int64_t tl = live[u] && tile[u][0] >= 0 && tile[u][0] < B.n_tiles ? tile[u][0] : 0;
            to[u] = live[u] ? B.tile_off[tl] : 0;
            tn[u] = live[u] ? B.tile_off[tl + 1] : 0;
        }
        /* a slot outside its tile's range is an error, never a store */
#pragma unroll
        for (int u = 0; u < EMIT_U; ++u) {
            if (!live[u]) continue;
            if (tile[u][0] < 0 || tile[u][0] >= B.n_tiles || to[u] + ps[u] >= tn[u]) { n_err++; continue; }
            B.segs[to[u] + ps[u]] = make_uint2((uint32_t)hi[u], ((uint32_t)plo[u][0] << 16) | (uint32_t)len[u][0]);
        }
        /* an empty ring between non-empty rings of one disc, counted once per
         * gap (the row misalignment of discedge2fulldisc, quirk Q15) */
#pragma unroll
        for (int u = 0; u < EMIT_U; ++u) {
            if (item[u] < B.n_items && (run[u] >> 16) == 0 && kk[u] > 0 && (B.runs[item[u] - 1] >> 16) != 0) {
                const unsigned long long endi = B.ring_off[hi[u] + 1];
                for (unsigned long long j = item[u] + 1; j < endi; ++j)
                    if ((B.runs[j] >> 16) != 0) { n_gaps++; break; }
            }
        }
    }
    const unsigned long long nov = *B.n_ovf < (unsigned long long)B.ovf_cap ? *B.n_ovf : (unsigned long long)B.ovf_cap;
    for (unsigned long long o = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; o < nov;
         o += (unsigned long long)gridDim.x * blockDim.x) {
        const uint4 v = B.ovf[o];
        if (v.x >= (uint64_t)B.n_tiles || B.tile_off[v.x] + v.y >= B.tile_off[v.x + 1]) { n_err++; continue; }
        B.segs[B.tile_off[v.x] + v.y] = make_uint2(v.z, v.w);
    }
    block_count(&B.counters[3], n_gaps, lds16);
    block_count(B.err, n_err, lds16);
}

/* ------------------------------------------------------------------ */
/* spatial order of the painted halos                                  */
/* ------------------------------------------------------------------ */

/* The fields read here lie in bytes [32, 112) and [128, 160) of the
 * 176-B record.  They are fetched with seven 16-B loads issued together and
 * one wait (field by field, the compiler put the loads of the FP32 fields
 * behind branches on chi_out, a memory round trip each). */
static_assert(offsetof(HaloRec, hx) == 0 && offsetof(HaloRec, hy) == 8 && offsetof(HaloRec, hz) == 16 &&
              offsetof(HaloRec, dtheta) == 24, "HaloRec layout read by k_fallback");
static_assert(offsetof(HaloRec, zz0) == 32 && offsetof(HaloRec, phi0) == 40 && offsetof(HaloRec, a) == 48 &&
              offsetof(HaloRec, cosang) == 56 && offsetof(HaloRec, s0) == 64 && offsetof(HaloRec, thr) == 72 &&
              offsetof(HaloRec, chi_r) == 80 && offsetof(HaloRec, fc) == 88 && offsetof(HaloRec, fm) == 96 &&
              offsetof(HaloRec, vh) == 104 && offsetof(HaloRec, cat) == 128 && offsetof(HaloRec, ichi) == 136 &&
              offsetof(HaloRec, imh) == 140 && offsetof(HaloRec, irmin) == 144 && offsetof(HaloRec, irmax) == 148 &&
              offsetof(HaloRec, chi_out) == 152 && sizeof(HaloRec) == 176, "HaloRec layout read by k_gather_halos");

__global__ void __launch_bounds__(256) k_gather_halos(const HaloRec *in, const uint32_t *order, int64_t n,
                                                      BinHalo *bin, TileHalo *hot, HaloRec *full, int32_t nrt, int32_t nmht,
                                                      uint32_t *nr)
{
    /* nr: the (halo, ring) items of every halo in spatial order (its ring
     * range, or one item for an inert record), and 0 after the last, as
     * the scan into ring offsets needs */
    if (blockIdx.x == 0 && threadIdx.x == 0) nr[n] = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t o = order[i];
        const uint4 *src = reinterpret_cast<const uint4 *>(in + o);
        const uint4 c2 = src[2], c3 = src[3], c4 = src[4], c5 = src[5], c6 = src[6], c8 = src[8], c9 = src[9];
        const double zz0 = dbl_of(c2.x, c2.y), phi0 = dbl_of(c2.z, c2.w);
        const double a = dbl_of(c3.x, c3.y), cosang = dbl_of(c3.z, c3.w);
        const double s0 = dbl_of(c4.x, c4.y), thr = dbl_of(c4.z, c4.w);
        const double chi_r = dbl_of(c5.x, c5.y), fc = dbl_of(c5.z, c5.w);
        const double fm = dbl_of(c6.x, c6.y), vh = dbl_of(c6.z, c6.w);
        const int64_t cat = (int64_t)(((unsigned long long)c8.y << 32) | c8.x);
        const int32_t ichi = (int32_t)c8.z, imh = (int32_t)c8.w;
        const int32_t irmin = (int32_t)c9.x, irmax = (int32_t)c9.y, chi_out = (int32_t)c9.z;
        BinHalo b;
        b.cosang = cosang; b.zz0 = zz0; b.a = a; b.phi0 = phi0;
        b.irmin = irmin; b.irmax = irmax;
        bin[i] = b;
        nr[i] = irmax >= irmin ? (uint32_t)(irmax - irmin + 1) : 1u;
        if (hot) {
            TileHalo t;
            t.zz0 = zz0; t.s0 = s0; t.phi0 = phi0; t.thr = thr;
            t.chi2 = (float)(chi_r * chi_r);
            t.vh = chi_out ? 0.0f : (float)vh;
            t.fc = (float)fc; t.fm = (float)fm;
            /* a halo beyond the table's chi range paints zeros (vh = 0); its
             * cell indices may lie outside the table, so point it at column 0 */
            t.colbase = chi_out ? 0 : nrt * ((imh - 1) + nmht * (ichi - 1));
            t.pad = 0;
            t.hash = halo_hash(cat);
            hot[i] = t;
        }
        if (full) full[i] = in[o];
    }
}

/* the largest key: the last azimuth bin of the last band, which the inert
 * records share (their place in the order does not matter), so that the
 * keys of nside 4096 fit 16 bits (two radix passes) */
uint32_t p2m_key_max(int32_t nside, int32_t nphi)
{
    const uint32_t nband = (uint32_t)((4 * nside - 2) >> 3) + 1;
    return nband * (uint32_t)nphi - 1;
}

/* Sorts the records by the key k_setup wrote (stable radix sort of
 * (key, slot)), then writes the binning, FP32 and (if asked) full records
 * in that order.  keys and vals hold 2n entries: input then output. */
int p2m_sort_halos(const HaloRec *in, int64_t n, uint32_t key_max, BinHalo *bin, TileHalo *hot, HaloRec *full,
                   int32_t nrt, int32_t nmht, uint32_t *keys, uint32_t *vals, void *tmp, size_t *tmp_bytes,
                   uint32_t *nr, hipStream_t s)
{
    int bits = 1;
    while ((1u << bits) <= key_max && bits < 32) ++bits;
    if (!tmp) {
        hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, *tmp_bytes, keys, keys + n, vals, vals + n,
                                                          (int)n, 0, bits, s);
        return e == hipSuccess ? 0 : -1;
    }
    if (n <= 0) return 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, *tmp_bytes, keys, keys + n, vals, vals + n,
                                                      (int)n, 0, bits, s);
    if (e != hipSuccess) return -1;
    unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_gather_halos, dim3(g), dim3(256), 0, s, in, vals + n, n, bin, hot, full, nrt, nmht, nr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

static unsigned bin_grid(int64_t n) { int64_t g = (n + 255) / 256; return (unsigned)(g > 8192 ? 8192 : (g < 1 ? 1 : g)); }
void p2m_launch_ring_counts(const BinArgs &b, hipStream_t s)
{
    hipLaunchKernelGGL(k_ring_counts, dim3(bin_grid(b.n_halo + 1)), dim3(256), 0, s, b);
}
static unsigned chunk_grid(unsigned long long n_items)
{
    unsigned long long g = (n_items + BIN_CH - 1) / BIN_CH;
    if (g >= 8) g &= ~7ULL;      /* a multiple of 8 for the XCD-aware split */
    return (unsigned)(g > 2048 ? 2048 : (g < 1 ? 1 : g));
}
void p2m_launch_seg_count(const BinArgs &b, hipStream_t s)
{
    if (b.n_items == 0) return;
    hipLaunchKernelGGL(k_group_heads, dim3(bin_grid(b.n_halo)), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_seg_count, dim3(chunk_grid(b.n_items)), dim3(256), 0, s, b);
}
void p2m_launch_seg_emit(const BinArgs &b, hipStream_t s)
{
    if (b.n_items == 0) return;
    hipLaunchKernelGGL(k_seg_emit, dim3(chunk_grid(b.n_items)), dim3(256), 0, s, b);
}

/* exclusive scan of uint32 counts into uint64 offsets with a 64-bit accumulator */
int p2m_scan_u32(const uint32_t *cnt, unsigned long long *off, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t s)
{
    hipError_t e = hipcub::DeviceScan::ExclusiveScan(tmp, *tmp_bytes, cnt, off, hipcub::Sum(), 0ULL, (int)n, s);
    return e == hipSuccess ? 0 : -1;
}


/* ===================================================================== */
/* tiled painter                                                          */
/* ===================================================================== */

/* ------------------------------------------------------------------ */
/* tile painters: common parts                                          */
/* ------------------------------------------------------------------ */

/* Workgroup barrier for LDS traffic only: waits for this wave's LDS
 * operations, not for its outstanding global loads and stores, then
 * s_barrier.  (__syncthreads also drains global memory, which exposes the
 * latency of the map stores at every tile boundary.) */
__device__ static inline void lds_barrier()
{
    /* LDS-only release and acquire fences around the barrier: the compiler
     * keeps LDS accesses on their side of it, and the hardware waits only
     * for LDS traffic (lgkmcnt), not for outstanding global memory */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_waitcnt(0xc07f);      /* lgkmcnt(0); vmcnt, expcnt unconstrained */
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

#define FSEG 256                   /* segments staged per batch */
#define FB_LDS 32                  /* deferred pixels of a tile buffered in LDS (flushed per tile) */
#define PMAX 3072                  /* pixels per batch (>= the widest tile, 2048) */

/* the part of a batch both painters share: segment starts and the pixel
 * -> segment map */
struct BatchLDS {
    uint32_t pre[FSEG + 1];        /* exclusive scan of the staged lengths */
    uint32_t seg_of[PMAX / 4];     /* batch pixel -> segment (bytes) */
    uint32_t wsum[16];
    uint32_t nb;
};

__global__ void __launch_bounds__(256) k_tile_desc(int32_t nside, int32_t log2w, int64_t n_tiles, const int32_t *tile_ring,
                                                   const int32_t *tile_base, TileDesc *desc)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n_tiles; t += (int64_t)gridDim.x * blockDim.x) {
        TileDesc d;
        d.ring = tile_ring[t];
        tile_span(nside, d.ring, (int32_t)(t - tile_base[d.ring - 1]), log2w, d.first, d.ip0, d.width);
        ring_zsth(nside, d.ring, d.zr, d.sth);
        ring_phi_param(nside, d.ring, d.off, d.den);
        d.kphi = HPX_HALFPI / d.den;
        d.pad = 0;
        desc[t] = d;
    }
}

void p2m_launch_tile_desc(int32_t nside, int32_t log2w, int64_t n_tiles, const int32_t *tile_ring,
                          const int32_t *tile_base, TileDesc *desc, hipStream_t s)
{
    if (n_tiles <= 0) return;
    int64_t g = (n_tiles + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_tile_desc, dim3((unsigned)g), dim3(256), 0, s, nside, log2w, n_tiles, tile_ring, tile_base, desc);
}

/* XCD-aware tile order: workgroup b runs on XCD b % 8; the workgroups of
 * one XCD sweep one contiguous eighth of the tiles (ring-major), so tiles in
 * flight on an XCD are neighbours that share halos in its L2. */
__device__ static inline void tile_range(const TileArgs &A, int64_t &tbeg, int64_t &tend, int64_t &tstep)
{
    const bool xcd = (gridDim.x % 8) == 0 && A.n_tiles >= 8;
    tstep = xcd ? gridDim.x / 8 : gridDim.x;
    tbeg = xcd ? (int64_t)(blockIdx.x % 8) * A.n_tiles / 8 + blockIdx.x / 8 : blockIdx.x;
    tend = xcd ? (int64_t)(blockIdx.x % 8 + 1) * A.n_tiles / 8 : A.n_tiles;
}

/* Given this thread's staged length (0 if none; nstage segments staged),
 * forms pre[], takes the leading segments whose pixels fit in PMAX (a
 * segment is at most W <= PMAX pixels, so at least one) and builds the
 * pixel -> segment map.  Returns the number of segments in the batch; P
 * receives its pixel count.  Every thread calls; ends with a barrier. */
__device__ static inline int batch_prepare(BatchLDS *bl, uint32_t len, int nstage, uint32_t &P)
{
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    uint32_t v = wave_incl_scan(len);
    if (lane == 63) bl->wsum[wv] = v;
    lds_barrier();
    uint32_t add = 0;
    for (int k = 0; k < wv; ++k) add += bl->wsum[k];
    v += add;
    if (tid < nstage) bl->pre[tid + 1] = v;
    if (tid == 0) bl->pre[0] = 0;
    lds_barrier();
    if (tid < nstage) {
        const uint32_t e = bl->pre[tid + 1];
        if (e <= PMAX && (tid + 1 == nstage || bl->pre[tid + 2] > PMAX)) bl->nb = (uint32_t)(tid + 1);
    }
    lds_barrier();
    const int nb = (int)bl->nb;
    P = bl->pre[nb];
    /* each thread takes a run of whole words, finds the segment of its first
     * pixel by binary search and walks forward */
    const uint32_t nw = (P + 3) >> 2;
    const uint32_t wpt = (nw + blockDim.x - 1) / blockDim.x;
    const uint32_t w0 = tid * wpt;
    const uint32_t w1 = w0 + wpt < nw ? w0 + wpt : nw;
    if (w0 < w1) {
        const uint32_t p0 = w0 * 4;
        int lo_s = 0, hi_s = nb - 1;
        while (lo_s < hi_s) {
            const int mid = (lo_s + hi_s + 1) >> 1;
            if (bl->pre[mid] <= p0) lo_s = mid; else hi_s = mid - 1;
        }
        uint32_t sg = (uint32_t)lo_s, nxt = bl->pre[sg + 1];
        for (uint32_t w = w0; w < w1; ++w) {
            uint32_t o = 0;
            for (int bb = 0; bb < 4; ++bb) {
                const uint32_t p = w * 4 + bb;
                while (p >= nxt && (int)sg + 1 < nb) { ++sg; nxt = bl->pre[sg + 1]; }
                o |= sg << (8 * bb);
            }
            bl->seg_of[w] = o;
        }
    }
    lds_barrier();
    return nb;
}

/* ------------------------------------------------------------------ */
/* FP64 tiled painter                                                   */
/*                                                                      */
/* The exact arithmetic of the reference per pixel: pix2vec_ring from    */
/* the ring's z and sin(theta) and the pixel's phi, angdist against the  */
/* disc radius, interpolate_table in FP64, accumulated in an FP64 LDS    */
/* tile (ds_add_f64) and reduced to the FP64 map.                        */
/* ------------------------------------------------------------------ */

/* the per-halo values the exact per-pixel code reads, staged per segment */
struct ExactSeg {
    BatchLDS b;
    double hx[FSEG], hy[FSEG], hz[FSEG], dth[FSEG], vh[FSEG], chir[FSEG], fc[FSEG], fm[FSEG];
    unsigned long long hash[FSEG];
    uint32_t lo[FSEG];
    int32_t imh[FSEG], ichi[FSEG], cout[FSEG];
};

/* interpolate_table (maptable.f90:176-216) on staged per-halo values, the
 * same operations in the same order as interp_exact; the table through a
 * raw buffer over its n_flat floats, whose out-of-range reads return 0 as
 * tab_flat's bounds test does (the flat index is never negative: r >= rmint
 * and the cells are clamped to >= 1) */
__device__ static inline double tab_buf(__amdgpu_buffer_rsrc_t tr, const TabGeom &g, int p, int ir, int imh, int ichi)
{
    const int64_t idx = (int64_t)(p - 1) + 2 * ((int64_t)(ir - 1) + (int64_t)g.nrt * ((int64_t)(imh - 1) + (int64_t)g.nmht * (int64_t)(ichi - 1)));
    return (double)__uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(tr, (int)(idx * 4), 0, 0));
}

__device__ static inline double interp_staged(int chi_out, double chi_r, double fc, double fm, int imh, int ichi,
                                              const TabGeom &g, __amdgpu_buffer_rsrc_t tr, int profile, double theta)
{
    if (chi_out) return 0.0;
    double r = chi_r * p2m_tan(theta);
    if (r > g.rmaxt) return 0.0;
    if (r < g.rmint) r = g.rmint + 1e-5;
    double lr = p2m_log(r);
    int ir = (int)((lr - g.lr0) / g.drt) + 1;
    double fr = lr - (g.lr0 + (double)(ir - 1) * g.drt);
    double v = tab_buf(tr, g, profile, ir, imh, ichi) * (1.0 - fr) * (1.0 - fc) * (1.0 - fm);
    v = v + tab_buf(tr, g, profile, ir + 1, imh, ichi) * (fr) * (1.0 - fc) * (1.0 - fm);
    v = v + tab_buf(tr, g, profile, ir, imh + 1, ichi) * (1.0 - fr) * (fc) * (1.0 - fm);
    v = v + tab_buf(tr, g, profile, ir + 1, imh + 1, ichi) * (fr) * (fc) * (1.0 - fm);
    v = v + tab_buf(tr, g, profile, ir, imh, ichi + 1) * (1.0 - fr) * (1.0 - fc) * (fm);
    v = v + tab_buf(tr, g, profile, ir + 1, imh, ichi + 1) * (fr) * (1.0 - fc) * (fm);
    v = v + tab_buf(tr, g, profile, ir, imh + 1, ichi + 1) * (1.0 - fr) * (fc) * (fm);
    v = v + tab_buf(tr, g, profile, ir + 1, imh + 1, ichi + 1) * (fr) * (fc) * (fm);
    return v;
}

template <int NM>
__global__ void __launch_bounds__(512) k_tile_exact(TileArgs A)
{
    extern __shared__ unsigned char smem_raw[];
    const int W = 1 << A.log2w;
    /* LDS: acc (NM*W doubles) | hash (W u64) count (W u32) if membership | ExactSeg */
    double *acc = (double *)smem_raw;
    const size_t acc_bytes = (size_t)NM * W * 8;
    unsigned long long *lhash = (unsigned long long *)(smem_raw + acc_bytes);
    uint32_t *lcount = (uint32_t *)(smem_raw + acc_bytes + (A.count ? (size_t)W * 8 : 0));
    const size_t mem_bytes = A.count ? (size_t)W * 12 : 0;
    ExactSeg *sl = (ExactSeg *)(smem_raw + ((acc_bytes + mem_bytes + 15) & ~(size_t)15));
    __shared__ unsigned long long lds16[16];
    const int tid = threadIdx.x;
    unsigned long long n_upd = 0, n_err = 0;
    __amdgpu_buffer_rsrc_t trs[NM];
    for (int m = 0; m < NM; ++m)
        trs[m] = __builtin_amdgcn_make_buffer_rsrc((void *)A.tab_flat[m], (short)0, (int)(A.n_flat * 4), 0x00020000);
    int64_t tbeg, tend, tstep;
    tile_range(A, tbeg, tend, tstep);
    for (int64_t t = tbeg; t < tend; t += tstep) {
        const TileDesc G = A.desc[t];
        const uint64_t sbeg = A.tile_off[t], send = A.tile_off[t + 1];
        for (int q = tid; q < G.width; q += blockDim.x) {
            for (int m = 0; m < NM; ++m) acc[m * W + q] = 0.0;
            if (A.count) { lcount[q] = 0; lhash[q] = 0ULL; }
        }
        for (uint64_t b0 = sbeg; b0 < send;) {
            lds_barrier();
            int nstage = (int)(send - b0 < FSEG ? send - b0 : FSEG);
            if (nstage > (int)blockDim.x) nstage = (int)blockDim.x;
            uint32_t len = 0;
            if (tid < nstage) {
                const uint2 sg = A.segs[b0 + tid];
                /* a segment must name a record and lie inside the tile */
                const uint32_t slo = sg.y >> 16, slen = sg.y & 0xffffu;
                const bool ok = sg.x < (uint64_t)A.n_halo && slo + slen <= (uint32_t)G.width;
                if (!ok) n_err++;
                const HaloRec &h = A.halos[ok ? sg.x : 0u];
                sl->lo[tid] = slo;
                len = ok ? slen : 0u;
                sl->hx[tid] = h.hx; sl->hy[tid] = h.hy; sl->hz[tid] = h.hz; sl->dth[tid] = h.dtheta;
                sl->vh[tid] = h.vh; sl->chir[tid] = h.chi_r; sl->fc[tid] = h.fc; sl->fm[tid] = h.fm;
                sl->imh[tid] = h.imh; sl->ichi[tid] = h.ichi; sl->cout[tid] = h.chi_out;
                if (A.count) sl->hash[tid] = halo_hash(h.cat);
            }
            uint32_t P;
            const int nb = batch_prepare(&sl->b, len, nstage, P);
            b0 += (uint64_t)nb;
            const unsigned char *segb = (const unsigned char *)sl->b.seg_of;
            for (uint32_t p = tid; p < P; p += blockDim.x) {
                const int s = segb[p];
                const int q = (int)(sl->lo[s] + (p - sl->b.pre[s]));
                const double phi = (((double)(G.ip0 + q) + G.off) * HPX_HALFPI) / G.den;
                double sn, cs;
                p2m_sincos(phi, &sn, &cs);
                const double angle = angdist(sl->hx[s], sl->hy[s], sl->hz[s], G.sth * cs, G.sth * sn, G.zr);
                if (!(angle < sl->dth[s])) continue;
                n_upd++;
                for (int m = 0; m < NM; ++m)
                    atomicAdd(&acc[m * W + q], sl->vh[s] * interp_staged(sl->cout[s], sl->chir[s], sl->fc[s], sl->fm[s],
                                                                         sl->imh[s], sl->ichi[s], A.g, trs[m],
                                                                         A.profile, angle));
                if (A.count) {
                    atomicAdd(&lcount[q], 1u);
                    atomicAdd(&lhash[q], sl->hash[s]);
                }
            }
        }
        lds_barrier();
        const int64_t pix0 = G.first + G.ip0;
        /* branch on accumulate outside the loops (no speculative map loads) */
        if (A.accumulate) {
            for (int q = tid; q < G.width; q += blockDim.x) {
                for (int m = 0; m < NM; ++m) A.map64[m][pix0 + q] += acc[m * W + q];
                if (A.count) {
                    A.count[pix0 + q] += lcount[q];
                    A.hash[pix0 + q] += lhash[q];
                }
            }
        } else {
            for (int q = tid; q < G.width; q += blockDim.x) {
                for (int m = 0; m < NM; ++m) A.map64[m][pix0 + q] = acc[m * W + q];
                if (A.count) {
                    A.count[pix0 + q] = lcount[q];
                    A.hash[pix0 + q] = lhash[q];
                }
            }
        }
        lds_barrier();
    }
    block_count(&A.counters[1], n_upd, lds16);
    block_count(A.err, n_err, lds16);
}

/* --------------------------------------------------------------------- */
/* FP32 tiled painter                                                      */
/*                                                                         */
/* Per pixel: one FP64 FMA for the azimuth offset from the halo centre,    */
/* then FP32: chord^2 = A + B sin^2(dphi/2) with the per-(halo,ring)       */
/* constants A = (z-z0)^2 + (sth-s0)^2, B = 4 sth s0 staged in LDS;        */
/* tan^2(theta) = 4h(1-h)/(1-2h)^2 with h = chord^2/4; the cell coordinate */
/* u = (ln r - ln rmin)/drt from log2(r^2); fr = ln(r/r_node) from         */
/* log1p(r^2/r_node^2 - 1); a lerp of the 4 bracketing radial columns.    */
/* Values are summed in an FP64 LDS tile.  Pixels whose                    */
/* chord^2 lies within band_c of the disc edge, or whose u lies within     */
/* `band` of a cell boundary, the r cut-off or the r clamp, are deferred   */
/* to k_fallback, which decides them with the exact FP64 arithmetic of the */
/* reference.                                                              */
/* --------------------------------------------------------------------- */

/* One array per constant: a wave's pixels read a few consecutive segments,
 * and 4-byte reads of one array (paired by the compiler into read2st64)
 * broadcast in one pass, where 16-byte reads of a struct keep the LDS unit
 * several times busier. */
struct FastSeg {
    uint32_t nb_s, P_s;            /* the batch's segment and pixel counts (set between its barriers) */
    uint32_t seg_of[PMAX / 4];     /* batch pixel -> segment (bytes) */
    uint32_t wsum[16];             /* per wave: staged lengths [0, 8), segments that fit [8, 16) */
    /* the per-tile control words come in two slots, used by alternate
     * tiles, so that a tile's end needs no barrier before the next begins */
    uint32_t nfb[2];               /* deferred pixels of the tile */
    uint32_t lost[2];              /* a deferred pixel of the tile found the list full */
    int32_t next_tile[2];          /* the tile this workgroup paints next */
    unsigned long long fb_base;    /* list slot of the tile's deferred pixels */
    uint4 fb[FB_LDS];
    uint32_t halo[FSEG];
    int16_t qoff[FSEG];            /* in-tile pixel minus batch pixel: lo - pre (-3072 .. 2047) */
    float A[FSEG], B[FSEG], thr_hi[FSEG], chi2[FSEG];   /* chord^2 = A + B sin^2(x); its band top; chi^2 */
    float w00[FSEG], w10[FSEG], w01[FSEG], w11[FSEG];   /* corner weights times the velocity factor */
    float xc[FSEG];                /* azimuth origin x_c */
    int32_t cidx[FSEG];            /* batch pixel of the azimuth origin: p - cidx = ip - ip_c */
    int32_t colbase[FSEG];
};

/* sin^2(x) for |x| <= 2: the odd Taylor polynomial of sin to x^15 (the
 * next term is below 3e-10 there, 6e-12 at pi/2), squared; one branch-free
 * sequence */
__device__ static inline float sin2f(float x)
{
    const float z = x * x;
    float p = __builtin_fmaf(z, -7.6471637e-13f, 1.6059044e-10f);
    p = __builtin_fmaf(z, p, -2.5052108e-8f);
    p = __builtin_fmaf(z, p, 2.7557319e-6f);
    p = __builtin_fmaf(z, p, -1.9841270e-4f);
    p = __builtin_fmaf(z, p, 8.3333333e-3f);
    p = __builtin_fmaf(z, p, -1.6666667e-1f);
    const float sn = __builtin_fmaf(x * z, p, x);
    return sn * sn;
}

/* NF consecutive floats of a raw buffer at byte offset vo + so, in loads of
 * 16 and 8 bytes (dword alignment is enough for buffer loads) */
template <int NF>
__device__ static inline void load_floats(__amdgpu_buffer_rsrc_t r, int vo, int so, float *f)
{
    int k = 0;
#pragma unroll
    for (; k + 4 <= NF; k += 4) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, vo + 4 * k, so, 0);
        f[k] = __uint_as_float((uint32_t)v[0]); f[k + 1] = __uint_as_float((uint32_t)v[1]);
        f[k + 2] = __uint_as_float((uint32_t)v[2]); f[k + 3] = __uint_as_float((uint32_t)v[3]);
    }
#pragma unroll
    for (; k + 2 <= NF; k += 2) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, vo + 4 * k, so, 0);
        f[k] = __uint_as_float((uint32_t)v[0]); f[k + 1] = __uint_as_float((uint32_t)v[1]);
    }
#pragma unroll
    for (; k < NF; ++k) f[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, vo + 4 * k, so, 0));
}

/* Scan of the staged lengths: start receives this thread's first batch
 * pixel and fit whether its segment belongs to the batch, i.e. lies within
 * the first PMAX pixels (a segment is at most W <= PMAX pixels, so at least
 * one does; the prefix is non-decreasing, so the batch is a leading run of
 * the staged segments).  The batch's segment and pixel counts reach
 * fs->nb_s and fs->P_s through LDS atomics, readable after the caller's next
 * barrier.  Every thread calls; one barrier. */
__device__ static inline void batch_prefix(FastSeg *fs, uint32_t len, int nstage, uint32_t &start, bool &fit)
{
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    uint32_t v = wave_incl_scan(len);
    if (lane == 63) fs->wsum[wv] = v;
    if (tid == 0) { fs->nb_s = 0; fs->P_s = 0; }
    lds_barrier();
    uint32_t add = 0;
    for (int k = 0; k < wv; ++k) add += fs->wsum[k];
    v += add;
    start = v - len;
    fit = tid < nstage && v <= PMAX;
    const unsigned long long fm = __ballot(fit);
    if (fm) {
        /* the wave's fitting lanes are lanes 0 .. last */
        const int last = 63 - __clzll((long long)fm);
        if (lane == 0) atomicAdd(&fs->nb_s, (uint32_t)__popcll(fm));
        if (lane == last) atomicMax(&fs->P_s, v);
    }
}

/* The batch's pixel -> segment byte map, written by the segments
 * themselves: thread s (in the batch) owns pixels [a, a + len).  A segment of at
 * most 64 pixels is written by its own thread (bytes up to a word boundary,
 * whole words, the tail bytes); a longer one by its whole wave, 64 words a
 * round, its head and tail bytes by six lanes.  Every byte belongs to one
 * segment and whole-word writes only cover words inside one, so the writes
 * never overlap; there are no reads.  Every lane of the wave must be
 * active.  The caller's barrier publishes the map. */
__device__ static inline void build_seg_map_direct(FastSeg *fs, bool in_batch, uint32_t a, uint32_t len)
{
    unsigned char *mb = reinterpret_cast<unsigned char *>(fs->seg_of);
    uint32_t *mw = fs->seg_of;
    const int tid = threadIdx.x, lane = __lane_id();
    const bool mine = in_batch && len > 0;
    const bool lng = mine && len > 64;
    if (mine && !lng) {
        const uint32_t e = a + len, fill = (uint32_t)tid * 0x01010101u;
        uint32_t p = a;
        for (; p < e && (p & 3u); ++p) mb[p] = (unsigned char)tid;
        for (; p + 4 <= e; p += 4) mw[p >> 2] = fill;
        for (; p < e; ++p) mb[p] = (unsigned char)tid;
    }
    unsigned long long lm = __ballot(lng);
    while (lm) {
        const int L = __ffsll((long long)lm) - 1;
        lm &= lm - 1;
        const uint32_t aL = (uint32_t)__builtin_amdgcn_readlane((int)a, L);
        const uint32_t eL = aL + (uint32_t)__builtin_amdgcn_readlane((int)len, L);
        const uint32_t sL = (uint32_t)((tid & ~63) + L), fL = sL * 0x01010101u;
        const uint32_t w0 = (aL + 3u) >> 2, w1 = eL >> 2;     /* whole words [w0, w1): len > 64, so w0 < w1 */
        for (uint32_t w = w0 + (uint32_t)lane; w < w1; w += 64) mw[w] = fL;
        if (lane < 3) {
            const uint32_t p = aL + (uint32_t)lane;             /* head bytes below 4 w0 */
            if (p < 4u * w0) mb[p] = (unsigned char)sL;
        } else if (lane < 6) {
            const uint32_t p = 4u * w1 + (uint32_t)(lane - 3);   /* tail bytes from 4 w1 */
            if (p < eL) mb[p] = (unsigned char)sL;
        }
    }
}


template <int NM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NM == 1 ? 6 : 1))) k_tile_fast(TileArgs A, const TileDesc *__restrict__ tdesc,
                                                   const unsigned long long *__restrict__ toff)
{
    extern __shared__ unsigned char smem_raw[];
    const int W = 1 << A.log2w;
    /* LDS: acc (NM*W doubles) | hash (W u64) count (W u32) if membership | FastSeg | hashes (FSEG u64) if
     * membership | 1/r_node^2 (nrt+2 floats) */
    double *acc = (double *)smem_raw;
    const size_t acc_bytes = (size_t)NM * W * 8;
    unsigned long long *lhash = (unsigned long long *)(smem_raw + acc_bytes);
    uint32_t *lcount = (uint32_t *)(smem_raw + acc_bytes + (A.count ? (size_t)W * 8 : 0));
    const size_t mem_bytes = A.count ? (size_t)W * 12 : 0;
    FastSeg *sl = (FastSeg *)(smem_raw + ((acc_bytes + mem_bytes + 15) & ~(size_t)15));
    unsigned long long *hh = (unsigned long long *)(sl + 1);   /* membership hashes, if counted */
    float *invn = (float *)(hh + (A.count ? FSEG : 0));
    /* inner edge of the deferral band, relative to the outer one */
    const float lo_band = (float)((1.0 - (double)A.band_c) / (1.0 + (double)A.band_c));
    const float u_zero = A.u_max + A.band;
    /* constants of the pixel loop that an instruction needs in a vector
     * register (a second scalar operand, a select), pinned there once */
    float v_k0 = A.u_k0, v_frc = A.fr_clamp;
    int v_irc = A.ir_clamp;
    asm volatile("" : "+v"(v_k0), "+v"(v_frc), "+v"(v_irc));

    const int nrt = A.g.nrt;
    /* the corner-quad table as a raw buffer: a 32-bit byte offset per pixel,
     * reads past its end return 0 */
    const __amdgpu_buffer_rsrc_t trs =
        __builtin_amdgcn_make_buffer_rsrc((void *)A.tab_q, (short)0, (int)(A.n_quad * 4), 0x00020000);
    const int tid = threadIdx.x;
    unsigned long long n_upd = 0, n_fb = 0, n_err = 0;
    const int64_t fbs = A.fb_cap / P2M_SHARDS;
    unsigned long long *const fbc = A.fb_count + (blockIdx.x % P2M_SHARDS) * P2M_CTR_STRIDE;
    uint4 *const fbr = A.fb + (blockIdx.x % P2M_SHARDS) * fbs;
    if (tid == 0) { sl->nfb[0] = sl->nfb[1] = 0; sl->lost[0] = sl->lost[1] = 0; }
    for (int k = tid; k < nrt + 2; k += blockDim.x) invn[k] = A.inv_node2f[k];
    lds_barrier();

    /* Tiles come from one queue per XCD: workgroup b runs on XCD b % 8 and
     * claims the next tile of that XCD's contiguous eighth of the tiles
     * (ring-major, so the tiles in flight on an XCD share halos in its L2)
     * from the XCD's counter.  A workgroup that becomes resident late, or
     * whose tiles run long, leaves the remaining tiles to the others.
     * Thread 0 claims the following tile with the tile's first segment
     * loads and writes it to the other slot, which the tile's next barrier
     * publishes.
     *
     * Tile k uses slot par = k & 1.  Its end needs no barrier: the next
     * tile's first writes to shared words come after that tile's own
     * barriers, its clear of the accumulator touches only the pixels each
     * thread itself wrote back, and slot par ^ 1 is cleared by thread 0
     * after tile k's post-loop barrier, when tile k-1's readers of it are
     * done. */
    const bool xq = (gridDim.x % 8) == 0 && A.n_tiles >= 8;
    const int xcd = xq ? (int)(blockIdx.x % 8) : 0;
    const int64_t xb = xq ? (int64_t)xcd * A.n_tiles / 8 : 0;
    const int64_t xe = xq ? (int64_t)(xcd + 1) * A.n_tiles / 8 : A.n_tiles;
    unsigned int *const qctr = A.tile_next + xcd * P2M_TQ_STRIDE;
    if (tid == 0) sl->next_tile[0] = (int32_t)(xb + (int64_t)atomicAdd(qctr, 1u));
    lds_barrier();
    for (int par = 0;; par ^= 1) {
        /* every thread reads the same tile: taken as uniform, its descriptor
         * and segment range are scalar loads */
        const int64_t t = (int64_t)__builtin_amdgcn_readfirstlane(sl->next_tile[par]);
        if (t >= xe) break;
        bool claimed = false;
        const TileDesc G = tdesc[t];
        const uint64_t sbeg = toff[t], send = toff[t + 1];
        if (A.redo && !A.lost[t]) {      /* repaint pass: only the tiles that lost pixels */
            lds_barrier();
            if (tid == 0) {
                sl->nfb[par ^ 1] = 0; sl->lost[par ^ 1] = 0;
                sl->next_tile[par ^ 1] = (int32_t)(xb + (int64_t)atomicAdd(qctr, 1u));
            }
            lds_barrier();
            continue;
        }
        const double kphi = G.kphi;
        const double inv_kphi = G.den * (2.0 / HPX_PI);   /* 1 / kphi */
        const float hphi = (float)(0.5 * kphi);          /* x step per pixel along the ring */
        const int32_t mr = G.ring < A.nside ? G.ring : (G.ring < 3 * A.nside ? A.nside : 4 * A.nside - G.ring);
        const int32_t npr = 4 * mr, half_npr = npr / 2;
        uint32_t tu = 0, tf = 0;             /* this tile's updates and deferrals */
        for (int q = tid; q < G.width; q += blockDim.x) {
            for (int m = 0; m < NM; ++m) acc[m * W + q] = 0.0;
            if (A.count) { lcount[q] = 0; lhash[q] = 0ULL; }
        }
        for (uint64_t b0 = sbeg; b0 < send;) {
            /* a later batch restages what the previous one's pixels read
             * (the first batch is ordered after the previous tile's reads by
             * that tile's post-loop barrier) */
            if (b0 != sbeg) lds_barrier();
            int nstage = (int)(send - b0 < FSEG ? send - b0 : FSEG);
            if (nstage > (int)blockDim.x) nstage = (int)blockDim.x;
            uint32_t len = 0;
            /* unconditional (the last staged entry for the idle lanes), so
             * that its wait can move down to the first use */
            /* a streaming load: the segment list is read once */
            const unsigned long long sgw = __builtin_nontemporal_load(
                reinterpret_cast<const unsigned long long *>(A.segs) + b0 + (uint64_t)(tid < nstage ? tid : nstage - 1));
            const uint2 sg = make_uint2((uint32_t)sgw, (uint32_t)(sgw >> 32));
            /* the claim of the following tile is made with the first
             * batch's segment loads, so that one wait covers both */
            if (!claimed) {
                if (tid == 0) sl->next_tile[par ^ 1] = (int32_t)(xb + (int64_t)atomicAdd(qctr, 1u));
                claimed = true;
            }
            uint32_t slo = 0;
            if (tid < nstage) {
                /* a segment must name a record and lie inside the tile; one
                 * that does not is counted as an error and painted empty */
                slo = sg.y >> 16;
                const uint32_t slen = sg.y & 0xffffu;
                const bool ok = sg.x < (uint64_t)A.n_halo && slo + slen <= (uint32_t)G.width;
                if (!ok) n_err++;
                len = ok ? slen : 0u;
            }
            uint32_t start;
            bool fit;
            batch_prefix(sl, len, nstage, start, fit);
            if (fit) {
                const TileHalo &h = A.hot[sg.x < (uint64_t)A.n_halo ? sg.x : 0u];
                /* the record's fields, all requested before the first is used
                 * (one memory round trip; the empty asm keeps them together) */
                const double zz0 = h.zz0, s0 = h.s0, phi0 = h.phi0, thr = h.thr;
                const float chi2 = h.chi2, fc = h.fc, fm = h.fm, vh = h.vh;
                const int32_t colbase = h.colbase;
                asm volatile("" :: "v"(zz0), "v"(s0), "v"(phi0), "v"(thr), "v"(chi2), "v"(fc), "v"(fm), "v"(vh), "v"(colbase));
                sl->halo[tid] = sg.x;
                sl->qoff[tid] = (int16_t)((int32_t)slo - (int32_t)start);
                const double dz = G.zr - zz0, ds = G.sth - s0;
                sl->A[tid] = (float)(dz * dz + ds * ds);
                sl->B[tid] = (float)((4.0 * G.sth) * s0);
                sl->thr_hi[tid] = (float)(thr * (1.0 + (double)A.band_c));
                sl->chi2[tid] = chi2;
                sl->w00[tid] = vh * ((1.0f - fc) * (1.0f - fm));
                sl->w10[tid] = vh * (fc * (1.0f - fm));
                sl->w01[tid] = vh * ((1.0f - fc) * fm);
                sl->w11[tid] = vh * (fc * fm);
                sl->colbase[tid] = colbase;
                /* azimuth origin: the ring pixel ip_c nearest the halo centre's
                 * azimuth, with x_c = ((ip_c + off) kphi - phi0) / 2 folded by
                 * pi (|x_c| <= kphi / 4) in FP64; a pixel's x is then x_c plus
                 * n = ip - ip_c steps of kphi / 2 in FP32.  n may move by npr
                 * (x by pi: sin^2 has period pi): the segment's middle step is
                 * moved into [-npr/2, npr/2] here, so on rings of at least 4 W
                 * pixels every n of the segment is within npr/2 + W/2 and
                 * |x| <= pi/2 + pi/8 + kphi/4; on shorter rings each pixel's n
                 * is folded into [-npr/2, npr/2] as well.  Either way the FP32
                 * error of x stays relative to x. */
                double ipc = __builtin_rint(phi0 * inv_kphi - G.off);
                if (ipc < 0.0) ipc += (double)npr;
                if (ipc >= (double)npr) ipc -= (double)npr;
                double xc = 0.5 * ((ipc + G.off) * kphi - phi0);
                xc -= HPX_PI * __builtin_rint(xc * (1.0 / HPX_PI));
                sl->xc[tid] = (float)xc;
                const int32_t nmid = G.ip0 + (int32_t)slo + (int32_t)(len >> 1) - (int32_t)ipc;
                sl->cidx[tid] = (int32_t)ipc - G.ip0 - ((int32_t)slo - (int32_t)start) +
                                (nmid > half_npr ? npr : (nmid < -half_npr ? -npr : 0));
                if (A.count) hh[tid] = h.hash;
            }
            build_seg_map_direct(sl, fit, start, len);
            lds_barrier();
            const int nb = (int)sl->nb_s;
            const uint32_t P = sl->P_s;
            b0 += (uint64_t)nb;
            const unsigned char *segb = (const unsigned char *)sl->seg_of;
            /* the pixel loop, with or without the per-pixel fold of n (a
             * tile-uniform choice: two copies of the loop) */
            auto paint_pixels = [&](auto fold) {
            for (uint32_t p = tid; p < P; p += blockDim.x) {
                const int s = segb[p];
                /* every per-segment constant the accepted path needs is read
                 * here, in one round of LDS reads: the empty asm below keeps
                 * the compiler from moving the later ones behind the disc test,
                 * where each would cost a round trip of its own */
                const int cix = sl->cidx[s];
                const float xcs = sl->xc[s], As = sl->A[s], Bs = sl->B[s], thr_hi = sl->thr_hi[s];
                const float chi2 = sl->chi2[s];
                const int qo = sl->qoff[s], cb = sl->colbase[s];
                int n = (int)p - cix;
                if constexpr (decltype(fold)::value) n = n > half_npr ? n - npr : (n < -half_npr ? n + npr : n);
                const float x = __builtin_fmaf((float)n, hphi, xcs);
                const float s2 = sin2f(x);
                const float c2 = __builtin_fmaf(Bs, s2, As);
                asm volatile("" :: "v"(chi2), "v"(qo), "v"(cb));
                if (!(c2 < thr_hi)) continue;                     /* outside the disc */
                const int q = (int)p + qo;
                /* tan^2(theta) = 4h(1-h)/(1-2h)^2 with h = chord^2/4 */
                const float dd = __builtin_fmaf(c2, -0.5f, 1.0f);
                const float rd = __builtin_amdgcn_rcpf(dd);
                const float t2 = __builtin_fmaf(-0.25f * c2, c2, c2) * (rd * rd);
                const float r2 = chi2 * t2;
                const float u = __builtin_fmaf(__builtin_amdgcn_logf(r2), A.u_k1, -v_k0);
                const bool zero = u > u_zero;                     /* r > rmaxt: value 0 */
                /* deferred: near the disc edge, a degenerate chord, or u within
                 * the band of a cell boundary 0..u_max (u clamped to that range,
                 * then its nearest boundary) */
                const float ub = __builtin_amdgcn_fmed3f(u, 0.0f, A.u_max);
                const bool defer = (c2 > thr_hi * lo_band) | !(dd > 1e-3f) | (__builtin_fabsf(u - __builtin_rintf(ub)) <= A.band);
                if (defer) {
                    const uint32_t slot = atomicAdd(&sl->nfb[par], 1u);
                    const uint4 e = make_uint4((uint32_t)G.ring, (uint32_t)(G.ip0 + q), sl->halo[s], 0u);
                    if (slot < FB_LDS) sl->fb[slot] = e;
                    else {
                        const unsigned long long gs = atomicAdd(fbc, 1ULL);
                        if (gs < (unsigned long long)fbs) fbr[gs] = e;
                        else sl->lost[par] = 1u;
                    }
                    tf++;
                    continue;
                }
                const bool clampr = u < 0.0f;                     /* r < rmint: r = rmint + 1e-5 */
                /* floor(u) + 1 clamped to [1, nrt-1] (u >= 0 here unless clamped:
                 * truncation is floor); a clamped r (r = rmint + 1e-5) has the cell
                 * and fraction the host computed for it */
                const int ir_u = min(max((int)u + 1, 1), nrt - 1);
                const int ir = clampr ? v_irc : ir_u;
                /* nodes ir-1 and ir of the cell: 4 corners of every map each,
                 * 32 NM contiguous bytes, requested before the radial fraction
                 * so that their latency covers its LDS read and log */
                const int vo = (cb + (ir - 1)) * (16 * NM);
                float tq[8 * NM];
                load_floats<8 * NM>(trs, vo, 0, tq);
                const float w00 = sl->w00[s], w10 = sl->w10[s], w01 = sl->w01[s], w11 = sl->w11[s];
                /* fr = ln(r / r_node) = (ln 2 / 2) log2(r^2 / r_node^2): the ratio is
                 * in [1, 1.08), where the hardware log2 is within one ulp
                 * (7e-9); both sides evaluated, then selected */
                float fr_in = 0.34657359f * __builtin_amdgcn_logf(r2 * invn[ir - 1]);
                asm volatile("" : "+v"(fr_in));   /* computed on every lane: a select, not a branch */
                const float fr = clampr ? v_frc : fr_in;
                tu++;
#pragma unroll
                for (int m = 0; m < NM; ++m) {
                    const float *c0v = tq + 4 * m, *c1v = tq + 4 * (NM + m);
                    /* the corners combined at each node, then one lerp in r */
                    float v0 = w00 * c0v[0];
                    v0 = __builtin_fmaf(w10, c0v[1], v0);
                    v0 = __builtin_fmaf(w01, c0v[2], v0);
                    v0 = __builtin_fmaf(w11, c0v[3], v0);
                    float v1 = w00 * c1v[0];
                    v1 = __builtin_fmaf(w10, c1v[1], v1);
                    v1 = __builtin_fmaf(w01, c1v[2], v1);
                    v1 = __builtin_fmaf(w11, c1v[3], v1);
                    const float val = __builtin_fmaf(fr, v1 - v0, v0);
                    atomicAdd(&acc[m * W + q], (double)(zero ? 0.0f : val));
                }
                if (A.count) {
                    atomicAdd(&lcount[q], 1u);
                    atomicAdd(&lhash[q], hh[s]);
                }
            }
            };
            if (npr < 4 * W) paint_pixels(std::true_type{});
            else paint_pixels(std::false_type{});
        }
        if (!claimed && tid == 0)        /* a tile without segments */
            sl->next_tile[par ^ 1] = (int32_t)(xb + (int64_t)atomicAdd(qctr, 1u));
        lds_barrier();
        /* the tile's deferred pixels still in LDS join the list before the
         * tile is written: a tile whose deferred pixels do not all fit in the
         * list writes nothing and is marked for the repaint pass */
        const uint32_t nfb = sl->nfb[par];
        const uint32_t nfl = nfb < FB_LDS ? nfb : FB_LDS;
        if (nfl) {
            if (tid == 0) sl->fb_base = atomicAdd(fbc, (unsigned long long)nfl);
            lds_barrier();
            for (uint32_t k = tid; k < nfl; k += blockDim.x)
                if (sl->fb_base + k < (unsigned long long)fbs) fbr[sl->fb_base + k] = sl->fb[k];
        }
        const bool lost = sl->lost[par] != 0u || (nfl && sl->fb_base + nfl > (unsigned long long)fbs);
        if (tid == 0) {
            if (lost) { A.lost[t] = 1; atomicAdd(A.n_lost + (blockIdx.x % P2M_SHARDS) * P2M_CTR_STRIDE, 1ULL); }
            else if (A.redo) A.lost[t] = 0;
        }
        if (!lost) { n_upd += (unsigned long long)tu; n_fb += (unsigned long long)tf; }
        const int64_t pix0 = G.first + G.ip0;
        /* the branch on accumulate is taken outside the loops: a select
         * inside lets the compiler load the map speculatively (a map read and
         * a wait on it per tile) even when the tile only stores */
        if (lost) {
        } else if (A.accumulate) {
            for (int q = tid; q < G.width; q += blockDim.x) {
                for (int m = 0; m < NM; ++m) {
                    float *dst = A.map32[m] + pix0 + q;
                    *dst = (float)((double)*dst + acc[m * W + q]);
                }
                if (A.count) {
                    A.count[pix0 + q] += lcount[q];
                    A.hash[pix0 + q] += lhash[q];
                }
            }
        } else {
            for (int q = tid; q < G.width; q += blockDim.x) {
                /* streaming stores: the map is not read again in the pass, so
                 * it should not displace halo records and table rows in L2 */
                for (int m = 0; m < NM; ++m) __builtin_nontemporal_store((float)acc[m * W + q], A.map32[m] + pix0 + q);
                if (A.count) {
                    A.count[pix0 + q] = lcount[q];
                    A.hash[pix0 + q] = lhash[q];
                }
            }
        }
        if (tid == 0) { sl->nfb[par ^ 1] = 0; sl->lost[par ^ 1] = 0; }   /* the next tile's slot */
    }
    /* the staged constants are dead: their space takes the block sums (the
     * kernel has no static LDS, so its occupancy is what its launch asks) */
    unsigned long long *red = (unsigned long long *)sl->fb;
    block_count(&A.counters[1], n_upd, red);
    block_count(&A.counters[5], n_fb, red);
    block_count(A.err, n_err, red);
}

/* Deferred pixels of k_tile_fast, decided and valued with the exact FP64
 * arithmetic of the reference (pix2vec_ring, angdist, interpolate_table),
 * added to the FP32 map with one atomic each. */
__global__ void __launch_bounds__(256) k_fallback(FallbackArgs A)
{
    __shared__ unsigned long long lds16[16];
    /* blockIdx.y: shard of the deferred list */
    const unsigned long long fbs = (unsigned long long)(A.fb_cap / P2M_SHARDS);
    unsigned long long n = A.fb_count[blockIdx.y * P2M_CTR_STRIDE];
    if (n > fbs) n = fbs;
    const uint4 *fbr = A.fb + blockIdx.y * fbs;
    unsigned long long n_upd = 0, n_err = 0;
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const uint4 e = fbr[i];
        const int32_t ring = (int32_t)e.x, ip = (int32_t)e.y;
        if (!(e.z < (uint64_t)A.n_halo && ring >= 1 && ring <= 4 * A.nside - 1 && ip >= 0)) { n_err++; continue; }
        /* the entries of a tile that lost some of its deferred pixels: the
         * repaint pass decides that tile again */
        if (A.lost[A.tile_base[ring - 1] + (ip >> A.log2w)]) continue;
        /* bytes [0, 32), [80, 112) and [128, 160) of the record: the fields
         * of the exact test and value, requested before the ring geometry */
        const uint4 *hs = reinterpret_cast<const uint4 *>(A.halos + A.order[e.z]);
        const uint4 c0 = hs[0], c1 = hs[1], c5 = hs[5], c6 = hs[6], c8 = hs[8], c9 = hs[9];
        int32_t npr, kshift; int64_t first;
        ring_info(A.nside, ring, npr, kshift, first);
        double zr, sth, off, den;
        ring_zsth(A.nside, ring, zr, sth);
        ring_phi_param(A.nside, ring, off, den);
        const double phi = (((double)ip + off) * HPX_HALFPI) / den;
        double sn, cs;
        p2m_sincos(phi, &sn, &cs);
        pin4(c0); pin4(c1); pin4(c5); pin4(c6); pin4(c8); pin4(c9);
        const double hx = dbl_of(c0.x, c0.y), hy = dbl_of(c0.z, c0.w), hz = dbl_of(c1.x, c1.y);
        const double dtheta = dbl_of(c1.z, c1.w);
        const double chi_r = dbl_of(c5.x, c5.y), fc = dbl_of(c5.z, c5.w), fm = dbl_of(c6.x, c6.y);
        const double vh = dbl_of(c6.z, c6.w);
        const int64_t cat = (int64_t)(((unsigned long long)c8.y << 32) | c8.x);
        const int32_t ichi = (int32_t)c8.z, imh = (int32_t)c8.w, chi_out = (int32_t)c9.z;
        const double angle = angdist(hx, hy, hz, sth * cs, sth * sn, zr);
        if (!(angle < dtheta)) continue;
        n_upd++;
        const int64_t pix = first + ip;
        for (int m = 0; m < A.n_maps; ++m) {
            /* built from the kernel argument inside the loop, so that the
             * resource stays in scalar registers */
            const __amdgpu_buffer_rsrc_t tr =
                __builtin_amdgcn_make_buffer_rsrc((void *)A.tab_flat[m], (short)0, (int)(A.n_flat * 4), 0x00020000);
            double v = vh * interp_staged(chi_out, chi_r, fc, fm, imh, ichi, A.g, tr, A.profile, angle);
            atomicAdd(A.map32[m] + pix, (float)v);
        }
        if (A.count) {
            atomicAdd(&A.count[pix], 1u);
            atomicAdd(&A.hash[pix], (unsigned long long)halo_hash(cat));
        }
    }
    block_count(&A.counters[1], n_upd, lds16);
    block_count(A.err, n_err, lds16);
}

int p2m_launch_fallback(const FallbackArgs &a, int64_t n_max, hipStream_t s)
{
    int64_t g = (n_max / P2M_SHARDS + 255) / 256;
    if (g < 1) g = 1;
    if (g > 64) g = 64;
    hipLaunchKernelGGL(k_fallback, dim3((unsigned)g, P2M_SHARDS), dim3(256), 0, s, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* The FP32 painter's table: for every (mass, chi) cell and radial node k,
 * the values of the four corners the pixel weights combine -- (imh, ichi),
 * (imh+1, ichi), (imh, ichi+1), (imh+1, ichi+1) -- for every map, so that a
 * pixel's two nodes k = ir-1, ir are 32 NM contiguous bytes:
 *   quad[((cell * nrt + k) * n_maps + m) * 4 + corner],
 *   cell = (imh-1) + nmht (ichi-1).
 * Corner values come from the profile slice by the reference's flat index
 * (maptable.f90:204-212), so the aliasing past the upper mass or chi edge
 * (Q6) reads the next column or the zero padding as the flat table does. */
__global__ void k_build_quad(const float *slice, int64_t n_slice, int32_t nrt, int32_t nmht, int32_t nchit,
                             int32_t n_maps, int32_t slot, float *quad)
{
    const int64_t n = (int64_t)nrt * nmht * nchit;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int32_t k = (int32_t)(e % nrt);
        const int64_t cell = e / nrt;
        const int32_t imh0 = (int32_t)(cell % nmht), ichi0 = (int32_t)(cell / nmht);   /* 0-based */
        float v[4];
        for (int cr = 0; cr < 4; ++cr) {
            const int64_t idx = (int64_t)k + (int64_t)nrt * ((int64_t)(imh0 + (cr & 1)) + (int64_t)nmht * (ichi0 + (cr >> 1)));
            v[cr] = idx < n_slice ? slice[idx] : 0.0f;
        }
        float *q = quad + (e * n_maps + slot) * 4;
        q[0] = v[0]; q[1] = v[1]; q[2] = v[2]; q[3] = v[3];
    }
}

void p2m_launch_build_quad(const float *slice, int64_t n_slice, int32_t nrt, int32_t nmht, int32_t nchit,
                           int32_t n_maps, int32_t slot, float *quad, hipStream_t s)
{
    const int64_t n = (int64_t)nrt * nmht * nchit;
    if (n <= 0) return;
    int64_t g = (n + 255) / 256;
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL(k_build_quad, dim3((unsigned)g), dim3(256), 0, s, slice, n_slice, nrt, nmht, nchit, n_maps,
                       slot, quad);
}

/* widest tile the painters take: a segment (at most one tile wide) must fit
 * in one batch of PMAX pixels */
int p2m_tile_max_log2w(void)
{
    int l = 6;
    while ((2 << l) <= PMAX) ++l;
    return l;
}

size_t p2m_tile_lds_bytes(int mode, int n_maps, int log2w, int membership, int nrt)
{
    size_t W = (size_t)1 << log2w;
    size_t b = (size_t)n_maps * W * 8;
    if (membership) b += W * 12;
    b = (b + 15) & ~(size_t)15;
    if (mode == 2) b += sizeof(FastSeg) + (membership ? 8 * FSEG : 0) + sizeof(float) * (size_t)(nrt + 2);
    else b += sizeof(ExactSeg);
    return b;
}

/* workgroups of the tile painter resident per CU at this shape (0 on error) */
int p2m_tile_occupancy(int mode, int n_maps, int wg, size_t lds)
{
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
#define P2M_OC(K) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, K, wg, lds)
    switch (mode * 8 + n_maps) {
    case 9: P2M_OC(k_tile_exact<1>); break;   case 10: P2M_OC(k_tile_exact<2>); break;
    case 11: P2M_OC(k_tile_exact<3>); break;  case 12: P2M_OC(k_tile_exact<4>); break;
    case 13: P2M_OC(k_tile_exact<5>); break;  case 14: P2M_OC(k_tile_exact<6>); break;
    case 17: P2M_OC(k_tile_fast<1>); break;   case 18: P2M_OC(k_tile_fast<2>); break;
    case 19: P2M_OC(k_tile_fast<3>); break;   case 20: P2M_OC(k_tile_fast<4>); break;
    case 21: P2M_OC(k_tile_fast<5>); break;   case 22: P2M_OC(k_tile_fast<6>); break;
    default: break;
    }
#undef P2M_OC
    return e == hipSuccess ? nb : 0;
}

int p2m_launch_tile_paint(const TileArgs &a, int mode, int n_maps, int grid, int wg, hipStream_t s)
{
    size_t lds = p2m_tile_lds_bytes(mode, n_maps, a.log2w, a.count != nullptr, a.g.nrt);
    if (lds > 160 * 1024 || wg > 512 || (mode == 2 && wg > 256)) return -1;
    if (mode == 2) {
        if (!a.tile_next) return -1;
        p2m_launch_fill_u32(a.tile_next, 8 * P2M_TQ_STRIDE, 0u, s);   /* empty tile queues */
    }
#define P2M_TE(NMv) case NMv: hipLaunchKernelGGL((k_tile_exact<NMv>), dim3(grid), dim3(wg), lds, s, a); break;
#define P2M_TF(NMv) case NMv: hipLaunchKernelGGL((k_tile_fast<NMv>), dim3(grid), dim3(wg), lds, s, a, a.desc, a.tile_off); break;
    if (mode == 1) {
        switch (n_maps) { P2M_TE(1) P2M_TE(2) P2M_TE(3) P2M_TE(4) P2M_TE(5) P2M_TE(6) default: return -1; }
    } else {
        switch (n_maps) { P2M_TF(1) P2M_TF(2) P2M_TF(3) P2M_TF(4) P2M_TF(5) P2M_TF(6) default: return -1; }
    }
#undef P2M_TE
#undef P2M_TF
    return 0;
}

/* ===================================================================== */
/* products: moments, RMS, dP/dy (pks2map.f90:321-357)                    */
/* ===================================================================== */

__global__ void __launch_bounds__(256) k_moments(const float *m32, const double *m64, int64_t n, double mean,
                                                 double *out /* [0..2]=sum x^k, [3]=sum (x-mean)^2 */,
                                                 unsigned long long *hist, int nb, float dy)
{
    double s1 = 0, s2 = 0, s3 = 0, sd = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float xf = m32 ? m32[i] : (float)m64[i];
        double x = (double)xf;
        s1 += x; s2 += x * x; s3 += (x * x) * x;
        double d = x - mean; sd += d * d;
        if (hist) {
            /* m = int(hpmap(i)/dpdy_dy)+1 in FP32 (pks2map.f90:344) */
            float u = xf / dy;
            if (u > -1.0f && u < (float)nb) {
                int mb = (int)u + 1;
                if (mb > 0 && mb <= nb) atomicAdd(&hist[mb - 1], 1ULL);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o); s2 += __shfl_xor(s2, o); s3 += __shfl_xor(s3, o); sd += __shfl_xor(sd, o);
    }
    if (__lane_id() == 0) {
        atomicAdd(&out[0], s1); atomicAdd(&out[1], s2); atomicAdd(&out[2], s3); atomicAdd(&out[3], sd);
    }
}

void p2m_launch_moments(const float *m32, const double *m64, int64_t n, double mean, double *out,
                        unsigned long long *hist, int nb, float dy, hipStream_t s)
{
    hipLaunchKernelGGL(k_moments, dim3(fill_grid(n) < 2048 ? fill_grid(n) : 2048), dim3(256), 0, s, m32, m64, n, mean, out, hist, nb, dy);
}

/* dpdy[i] = hist[i] / npix, then the cumulative sum, in float and in order
 * (pks2map.f90:347, 355-357), by one thread */
__global__ void k_dpdy_cumulative(const unsigned long long *hist, int nb, float npix, float *dpdy)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    float run = 0.0f;
    for (int i = 0; i < nb; ++i) {
        const float v = (float)hist[i] / npix;
        run = i == 0 ? v : run + v;
        dpdy[i] = run;
    }
}

void p2m_launch_dpdy_cumulative(const unsigned long long *hist, int nb, float npix, float *dpdy, hipStream_t s)
{
    hipLaunchKernelGGL(k_dpdy_cumulative, dim3(1), dim3(64), 0, s, hist, nb, npix, dpdy);
}

__global__ void k_f64_to_f32(const double *in, float *out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) out[i] = (float)in[i];
}
__global__ void k_f32_to_f64(const float *in, double *out, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) out[i] = (double)in[i];
}
void p2m_launch_f64_to_f32(const double *in, float *out, int64_t n, hipStream_t s) { if (n > 0) hipLaunchKernelGGL(k_f64_to_f32, dim3(fill_grid(n)), dim3(256), 0, s, in, out, n); }
void p2m_launch_f32_to_f64(const float *in, double *out, int64_t n, hipStream_t s) { if (n > 0) hipLaunchKernelGGL(k_f32_to_f64, dim3(fill_grid(n)), dim3(256), 0, s, in, out, n); }

/* ===================================================================== */
/* diagnostic: portable math on the device                                */
/* ===================================================================== */

__global__ void k_selftest_math(int op, const double *x, const double *y, double *out, int64_t n)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = x[i], b = y ? y[i] : 0.0, r = 0.0;
    switch (op) {
    case 0: r = p2m_sqrt(a); break;
    case 1: r = p2m_sin(a); break;
    case 2: r = p2m_cos(a); break;
    case 3: r = p2m_tan(a); break;
    case 4: r = p2m_asin(a); break;
    case 5: r = p2m_acos(a); break;
    case 6: r = p2m_atan2(a, b); break;
    case 7: r = p2m_log(a); break;
    case 8: r = p2m_exp(a); break;
    case 9: r = p2m_pow(a, b); break;
    default: r = 0.0; break;
    }
    out[i] = r;
}

int p2m_launch_selftest_math(int op, const double *x, const double *y, double *out, int64_t n, hipStream_t s)
{
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_selftest_math, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, op, x, y, out, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

/* plain streaming copy, 16 B per lane per iteration: the known-traffic
 * kernel that calibrates HBM byte counters and measures copy bandwidth */
__global__ void __launch_bounds__(256) k_copy_calib(const float4 *__restrict__ src, float4 *__restrict__ dst, int64_t n4)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int p2m_launch_copy_calib(const void *src, void *dst, int64_t n4, hipStream_t s)
{
    if (n4 <= 0) return 0;
    int64_t g = (n4 + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(k_copy_calib, dim3((unsigned)g), dim3(256), 0, s, (const float4 *)src, (float4 *)dst, n4);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
