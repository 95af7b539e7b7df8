/*
 * p2m_api.hip -- host runtime behind the C ABI of include/pks2map_gpu.h.
 *
 * Owns the device state of one pks2map run: distance tables, profile tables,
 * maps, per-halo records, tile bins.  A paint call decodes the catalogue in
 * chunks (appending painted halos to one device array), bins every
 * (halo, ring) pixel run into ring tiles, and paints all tiles in one launch,
 * so each map pixel is written once per call.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pks2map_gpu.h"
#include "p2m_kernels.h"

/* Scratch of p2m_products_noalloc: one array per GPU, kept for the life of
 * the process.  Layout: 8 doubles of partial sums, the 10000-bin dP/dy
 * histogram (u64), the cumulative dP/dy (float).  Its device address comes
 * from hipGetSymbolAddress once per device (the array's name is a host-side
 * stand-in in host code). */
#define P2M_NOALLOC_BYTES (128 * 1024)
#define P2M_NOALLOC_HIST_OFF 64
#define P2M_NOALLOC_DPDY_OFF (64 + 8 * 10000)
static_assert(P2M_NOALLOC_DPDY_OFF + 4 * 10000 <= P2M_NOALLOC_BYTES, "no-alloc scratch layout");
static __device__ __align__(256) unsigned char scratch[P2M_NOALLOC_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

/* the device address of this device's scratch (the device must be current) */
static int scratch_for_device(int device, void **out)
{
    if (device < 0 || device >= 64) return P2M_E_INVALID;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return P2M_E_HIP;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return 0;
}

/* last HIP failure of this thread, reported by p2m_strerror(P2M_E_HIP) */
static thread_local char g_hip_detail[256] = "";

static int hip_fail(hipError_t e, const char *what, int line)
{
    snprintf(g_hip_detail, sizeof(g_hip_detail), "HIP runtime error: %s (%s) at p2m_api.hip:%d: %s",
             hipGetErrorName(e), hipGetErrorString(e), line, what);
    (void)hipGetLastError();
    return P2M_E_HIP;
}

#define HIPCHK(x)                                                  \
    do {                                                           \
        hipError_t e_ = (x);                                       \
        if (e_ != hipSuccess) return hip_fail(e_, #x, __LINE__);   \
    } while (0)

namespace {

const double CKMS = 2.99792458e10 / 1e5;   /* cosmology.f90:23 */

struct Ctr {                 /* layout of one shard row of the device counter block; OVF is
                                the paint counter (PAINT + OVF, shard 0 only) of binning
                                overflow-list entries; ERR counts failed device-side range
                                checks, LOST the tiles marked for the repaint pass */
    enum { ERR = 0, DEC = 1, LOST = 2, ITEMS = 4, N_FLAT = 5, N_CAND = 6, N_CAND0 = 7, PAINT = 8, OVF = 7,
           TOTAL = P2M_CTR_STRIDE, ALL = P2M_CTR_STRIDE * P2M_SHARDS };
};

}  // namespace

struct p2m_ctx {
    p2m_config cfg;
    hipStream_t stream = nullptr;
    hipEvent_t ev[10] = {};
    hipStream_t fs_stream = nullptr;            /* the flat-sky painter, beside the binning */
    hipEvent_t ev_dec = nullptr, ev_fs = nullptr;
    unsigned long long *h_pin = nullptr;   /* pinned host words for mid-call readbacks */
    bool have_run = false;
    bool have_table[P2M_MAXMAPS] = {false};
    p2m_run run;
    double cut_low = 0, cut_high = 0, mypi = 0;
    std::vector<double> rofz, zofr;
    double *d_rofz = nullptr, *d_zofr = nullptr;
    CosmoDev cdev;
    int32_t hdr_i[3] = {0, 0, 0};
    float hdr_f[6] = {0, 0, 0, 0, 0, 0};
    TabGeom g;
    float *d_flat[P2M_MAXMAPS] = {nullptr};
    float *d_prof[P2M_MAXMAPS] = {nullptr};    /* the profile's slice, zero-padded (FP32 painter) */
    float *d_quad = nullptr;                    /* corner-quad table of all maps (FP32 painter) */
    int64_t n_flat = 0, n_prof = 0, n_quad = 0;
    double *d_inv_node2 = nullptr;
    float *d_inv_node2f = nullptr;
    uint4 *d_fb = nullptr;
    int64_t fb_cap = 0;
    float fr_clamp = 0.f, u_k1 = 0.f, u_k0 = 0.f, u_max = 0.f;
    int32_t ir_clamp = 1;
    float u_band = 2.5e-4f;     /* half-width of the deferral band around cell boundaries, in cells */
    int64_t hpnpix = 0, fsnpix = 0;
    double *d_hp64[P2M_MAXMAPS] = {nullptr};
    float *d_hp32[P2M_MAXMAPS] = {nullptr};
    double *d_fs64[P2M_MAXMAPS] = {nullptr};
    uint32_t *d_count = nullptr, *d_fs_count = nullptr;
    unsigned long long *d_hash = nullptr, *d_fs_hash = nullptr;
    bool hp_zero = true;        /* next paint may store instead of accumulate */
    bool hp_valid = true;       /* device HEALPix arrays hold the map (false after a lazy clear) */
    HaloRec *d_halo = nullptr;
    uint32_t *d_flat_idx = nullptr;
    uint32_t *d_nr = nullptr;
    unsigned long long *d_ring_off = nullptr;
    uint32_t *d_runs = nullptr, *d_pos = nullptr;   /* per (halo, ring) item: run, first piece's slot */
    int64_t runs_cap = 0;
    uint4 *d_ovf = nullptr;                       /* pieces beyond an item's second */
    int64_t ovf_cap = 0;
    int64_t halo_cap = 0;
    unsigned long long *d_ctr = nullptr;
    int log2w = 0, wg = 0;
    bool tuned = false;
    std::vector<int32_t> tile_base;
    int32_t *d_tile_base = nullptr;
    int32_t *d_tile_ring = nullptr;
    int n_cu = 0;                   /* compute units of the device */
    TileDesc *d_tile_desc = nullptr;
    double *d_ring_z = nullptr;     /* ring2z of every ring */
    uint32_t *d_group = nullptr;    /* halo of every 64th (halo, ring) item */
    int64_t group_cap = 0;
    uint32_t *d_fs_cnt = nullptr;   /* flat-sky box pixels per listed halo, and their scan */
    unsigned long long *d_fs_off = nullptr;
    int64_t fs_cap = 0;
    int64_t n_tiles = 0;
    int n_rings = 0;
    uint32_t *d_tile_cnt = nullptr;
    unsigned long long *d_tile_off = nullptr;
    void *d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    void *d_scan_tmp_fs = nullptr;              /* the flat-sky painter's own scan storage */
    size_t scan_tmp_fs_bytes = 0;
    uint2 *d_segs = nullptr;
    int64_t seg_cap = 0;
    /* host catalogues: two device slots that pieces of a block are copied
     * into on the copy stream, decoded from on the painter's stream */
    hipStream_t h2d_stream = nullptr;
    hipEvent_t ev_copy[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    hipEvent_t ev_push = nullptr, ev_h2d_end = nullptr;
    float *d_slot[2] = {nullptr, nullptr};
    int64_t slot_cap = 0;                       /* halos per slot */
    /* the streamed pass in progress (p2m_stream_*; p2m_paint* run one) */
    bool in_stream = false;
    int64_t stream_total = 0, stream_pushed = 0, stream_h2d_bytes = 0;
    int n_host_push = 0, n_pieces = 0;
    bool stream_pinned = true;
    /* per tile: 1 if the tile's deferred pixels did not all fit in the list
     * (the tile then wrote nothing and is painted again) */
    uint8_t *d_lost = nullptr;
    unsigned int *d_tile_next = nullptr;   /* tile-queue counters of the FP32 painter */
    int n_redo = 0;
    /* the FP32 painter's last launch, kept for the repaint pass */
    TileArgs last_t;
    FallbackArgs last_f;
    int last_grid = 0;
    int64_t last_nlisted = 0;
    bool last_fast = false;
    uint32_t *d_cand = nullptr;
    int64_t cand_cap = 0;
    /* painted halos in spatial order (tiled painters) */
    BinHalo *d_bin = nullptr;      /* fields the binning reads */
    TileHalo *d_hot = nullptr;     /* fields the FP32 painter stages (FP32 mode) */
    HaloRec *d_halo_s = nullptr;   /* full records (FP64 tiled mode) */
    uint32_t *d_keys = nullptr, *d_vals = nullptr;
    void *d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    p2m_stats st;
};

/* ------------------------------------------------------------------ helpers */
static int dmalloc(void **p, size_t bytes)
{
    if (bytes == 0) { *p = nullptr; return 0; }
    if (hipMalloc(p, bytes) != hipSuccess) { *p = nullptr; (void)hipGetLastError(); return P2M_E_NOMEM; }
    return 0;
}
static void dfree(void *p) { if (p) (void)hipFree(p); }

/* cosmology.f90:123-133 (dzdr) and 109-119 (drdz), FP64 */
static double dzdr(double h, double om, double z)
{
    double z1 = 1.0 + z;
    return ((h * 100.0) / CKMS) * p2m_sqrt((om * ((z1 * z1) * z1) + 1.0) - om);
}

static double host_lookup(const std::vector<double> &t, double tmin, double dt, double x)
{
    if (x != x) return x;
    int n = (int)t.size();
    double u = (x - tmin) / dt + 1.0;
    if (u < 1.0) return t[0];
    if (u >= (double)n) return t[n - 1];
    int itab = (int)u;
    double fint = ((x - (double)(itab - 1) * dt) - tmin) / dt;
    return (1.0 - fint) * t[itab - 1] + fint * t[itab];
}

static bool is_pow2(int32_t x) { return x > 0 && (x & (x - 1)) == 0; }

static int build_tiles(p2m_ctx *c, int log2w)
{
    int32_t ns = c->cfg.nside;
    if (ns <= 0) return 0;
    if (c->log2w == log2w && c->d_tile_base) return 0;
    c->n_rings = 4 * ns - 1;
    c->tile_base.assign(c->n_rings, 0);
    int64_t acc = 0;
    for (int r = 1; r <= c->n_rings; ++r) {
        c->tile_base[r - 1] = (int32_t)acc;
        acc += ring_tiles(ns, r, log2w);
    }
    c->n_tiles = acc;
    std::vector<int32_t> tile_ring((size_t)acc);
    for (int r = 1; r <= c->n_rings; ++r) {
        const int64_t e = r < c->n_rings ? c->tile_base[r] : acc;
        for (int64_t t = c->tile_base[r - 1]; t < e; ++t) tile_ring[(size_t)t] = r;
    }
    dfree(c->d_tile_base); dfree(c->d_tile_cnt); dfree(c->d_tile_off); dfree(c->d_tile_ring);
    dfree(c->d_tile_desc); dfree(c->d_lost);
    c->d_tile_base = nullptr; c->d_tile_cnt = nullptr; c->d_tile_off = nullptr;
    c->d_tile_ring = nullptr; c->d_tile_desc = nullptr; c->d_lost = nullptr;
    int rc = 0;
    rc |= dmalloc((void **)&c->d_lost, (size_t)((acc + 3) / 4) * 4);
    if (!c->d_tile_next) rc |= dmalloc((void **)&c->d_tile_next, sizeof(unsigned int) * 8 * P2M_TQ_STRIDE);
    rc |= dmalloc((void **)&c->d_tile_ring, sizeof(int32_t) * acc);
    rc |= dmalloc((void **)&c->d_tile_desc, sizeof(TileDesc) * acc);
    rc |= dmalloc((void **)&c->d_tile_base, sizeof(int32_t) * c->n_rings);
    rc |= dmalloc((void **)&c->d_tile_cnt, sizeof(uint32_t) * (c->n_tiles + 1));
    rc |= dmalloc((void **)&c->d_tile_off, sizeof(unsigned long long) * (c->n_tiles + 1));
    if (rc) return P2M_E_NOMEM;
    HIPCHK(hipMemcpy(c->d_tile_base, c->tile_base.data(), sizeof(int32_t) * c->n_rings, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_tile_ring, tile_ring.data(), sizeof(int32_t) * acc, hipMemcpyHostToDevice));
    p2m_launch_tile_desc(ns, log2w, acc, c->d_tile_ring, c->d_tile_base, c->d_tile_desc, c->stream);
    p2m_launch_fill_u32((uint32_t *)c->d_lost, (acc + 3) / 4, 0u, c->stream);
    HIPCHK(hipGetLastError());
    if (!c->d_ring_z) {
        if (dmalloc((void **)&c->d_ring_z, sizeof(double) * c->n_rings)) return P2M_E_NOMEM;
        p2m_launch_ring_z(ns, c->d_ring_z, c->stream);
        HIPCHK(hipGetLastError());
    }
    c->scan_tmp_bytes = 0;
    c->log2w = log2w;
    return 0;
}

/* ------------------------------------------------------------------ ABI */
extern "C" {

const char *p2m_version(void) { return "p2m 1.0 (gfx950)"; }

const char *p2m_strerror(int code)
{
    switch (code) {
    case P2M_OK: return "ok";
    case P2M_E_INVALID: return "invalid argument";
    case P2M_E_SIZE: return "buffer or header size mismatch";
    case P2M_E_HIP: return g_hip_detail[0] ? g_hip_detail : "HIP runtime error";
    case P2M_E_NOMEM: return "allocation failed";
    case P2M_E_RADIUS: return "disc radius outside [0, pi]";
    case P2M_E_STATE: return "call out of order (run parameters or table missing)";
    case P2M_E_IO: return "file could not be written";
    case P2M_E_NODEVICE: return "no GPU visible";
    case P2M_E_INTERNAL: return g_hip_detail[0] ? g_hip_detail : "device-side consistency check failed";
    default: return "unknown error";
    }
}

int p2m_create(p2m_ctx **out, const p2m_config *cfg)
{
    if (!out || !cfg) return P2M_E_INVALID;
    *out = nullptr;
    if (cfg->nside < 0 || (cfg->nside > 0 && (!is_pow2(cfg->nside) || cfg->nside > 8192))) return P2M_E_INVALID;
    if (cfg->npix_fs < 0 || cfg->npix_fs > 65536) return P2M_E_INVALID;
    if (cfg->n_maps < 1 || cfg->n_maps > P2M_MAXMAPS) return P2M_E_INVALID;
    if (cfg->model < 1 || cfg->model > 3) return P2M_E_INVALID;
    if (cfg->profile < 1 || cfg->profile > 2) return P2M_E_INVALID;
    if (cfg->precision < P2M_PREC_PLAIN_FP64 || cfg->precision > P2M_PREC_FP32_COMP) return P2M_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { (void)hipGetLastError(); return P2M_E_NODEVICE; }
    if (cfg->device < 0 || cfg->device >= ndev) return P2M_E_INVALID;
    HIPCHK(hipSetDevice(cfg->device));
    p2m_ctx *c = new p2m_ctx();
    c->cfg = *cfg;
    {
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, cfg->device) == hipSuccess) c->n_cu = ncu;
        else (void)hipGetLastError();
    }
    if (c->cfg.chunk_halos <= 0) c->cfg.chunk_halos = 1 << 26;
    if (c->cfg.chunk_halos > (1LL << 30)) c->cfg.chunk_halos = 1LL << 30;   /* chunk indices are 32-bit */
    memset(&c->st, 0, sizeof(c->st));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return P2M_E_HIP; }
    if (hipStreamCreateWithFlags(&c->fs_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->h2d_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_dec, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fs, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_copy[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_copy[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_free[0], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_free[1], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_push, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&c->ev_h2d_end) != hipSuccess) {
        (void)hipGetLastError();
        p2m_destroy(c);
        return P2M_E_HIP;
    }
    for (int i = 0; i < 10; ++i) (void)hipEventCreate(&c->ev[i]);
    {
        /* this device's scratch of p2m_products_noalloc, resolved once */
        void *sp = nullptr;
        int e = scratch_for_device(cfg->device, &sp);
        if (e) { p2m_destroy(c); return e; }
    }
    if (hipHostMalloc((void **)&c->h_pin, sizeof(unsigned long long) * (Ctr::ALL + 8), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->h_pin = nullptr;
    }
    c->hpnpix = cfg->nside > 0 ? 12LL * cfg->nside * cfg->nside : 0;
    c->fsnpix = (int64_t)cfg->npix_fs * cfg->npix_fs;
    int rc = 0;
    for (int m = 0; m < cfg->n_maps && !rc; ++m) {
        if (c->hpnpix) {
            if (cfg->precision == P2M_PREC_FP32_COMP) rc |= dmalloc((void **)&c->d_hp32[m], sizeof(float) * c->hpnpix);
            else rc |= dmalloc((void **)&c->d_hp64[m], sizeof(double) * c->hpnpix);
        }
        if (c->fsnpix) rc |= dmalloc((void **)&c->d_fs64[m], sizeof(double) * c->fsnpix);
    }
    if (!rc && (cfg->flags & P2M_FLAG_MEMBERSHIP)) {
        if (c->hpnpix) {
            rc |= dmalloc((void **)&c->d_count, sizeof(uint32_t) * c->hpnpix);
            rc |= dmalloc((void **)&c->d_hash, sizeof(unsigned long long) * c->hpnpix);
        }
        if (c->fsnpix) {
            rc |= dmalloc((void **)&c->d_fs_count, sizeof(uint32_t) * c->fsnpix);
            rc |= dmalloc((void **)&c->d_fs_hash, sizeof(unsigned long long) * c->fsnpix);
        }
    }
    if (!rc) rc |= dmalloc((void **)&c->d_ctr, sizeof(unsigned long long) * Ctr::ALL);
    if (rc) { p2m_destroy(c); return P2M_E_NOMEM; }
    *out = c;
    int e = p2m_clear(c);
    if (e) { p2m_destroy(c); *out = nullptr; return e; }
    return 0;
}

void p2m_destroy(p2m_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->fs_stream) (void)hipStreamSynchronize(c->fs_stream);
    if (c->h2d_stream) (void)hipStreamSynchronize(c->h2d_stream);
    dfree(c->d_slot[0]); dfree(c->d_slot[1]); dfree(c->d_lost); dfree(c->d_tile_next);
    for (int i = 0; i < 2; ++i) {
        if (c->ev_copy[i]) (void)hipEventDestroy(c->ev_copy[i]);
        if (c->ev_free[i]) (void)hipEventDestroy(c->ev_free[i]);
    }
    if (c->ev_push) (void)hipEventDestroy(c->ev_push);
    if (c->ev_h2d_end) (void)hipEventDestroy(c->ev_h2d_end);
    if (c->h2d_stream) (void)hipStreamDestroy(c->h2d_stream);
    for (int m = 0; m < P2M_MAXMAPS; ++m) {
        dfree(c->d_hp64[m]); dfree(c->d_hp32[m]); dfree(c->d_fs64[m]); dfree(c->d_flat[m]); dfree(c->d_prof[m]);
    }
    dfree(c->d_quad);
    dfree(c->d_count); dfree(c->d_hash); dfree(c->d_fs_count); dfree(c->d_fs_hash);
    dfree(c->d_rofz); dfree(c->d_zofr); dfree(c->d_inv_node2); dfree(c->d_inv_node2f); dfree(c->d_fb);
    dfree(c->d_halo); dfree(c->d_flat_idx); dfree(c->d_nr); dfree(c->d_ring_off); dfree(c->d_runs); dfree(c->d_pos); dfree(c->d_ovf); dfree(c->d_ctr);
    dfree(c->d_tile_base); dfree(c->d_tile_cnt); dfree(c->d_tile_off); dfree(c->d_scan_tmp); dfree(c->d_scan_tmp_fs);
    dfree(c->d_tile_ring); dfree(c->d_tile_desc); dfree(c->d_group); dfree(c->d_ring_z); dfree(c->d_fs_cnt); dfree(c->d_fs_off);
    dfree(c->d_segs); dfree(c->d_cand);
    dfree(c->d_halo_s); dfree(c->d_hot); dfree(c->d_bin); dfree(c->d_keys); dfree(c->d_vals); dfree(c->d_sort_tmp);
    for (int i = 0; i < 10; ++i) if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    if (c->ev_dec) (void)hipEventDestroy(c->ev_dec);
    if (c->ev_fs) (void)hipEventDestroy(c->ev_fs);
    if (c->fs_stream) (void)hipStreamDestroy(c->fs_stream);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int p2m_clear(p2m_ctx *c)
{
    if (!c) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    HIPCHK(hipSetDevice(c->cfg.device));
    /* the tiled painters store every HEALPix pixel on the first paint after a
     * clear, so only the plain painter (atomics) needs a zeroed map */
    const bool tiled = c->cfg.precision != P2M_PREC_PLAIN_FP64;
    for (int m = 0; m < c->cfg.n_maps; ++m) {
        if (c->d_hp64[m] && !tiled) p2m_launch_fill_f64(c->d_hp64[m], c->hpnpix, c->stream);
        if (c->d_fs64[m]) p2m_launch_fill_f64(c->d_fs64[m], c->fsnpix, c->stream);
    }
    if (c->d_count && !tiled) p2m_launch_fill_u32(c->d_count, c->hpnpix, 0u, c->stream);
    if (c->d_hash && !tiled) p2m_launch_fill_u64(c->d_hash, c->hpnpix, c->stream);
    if (c->d_fs_count) p2m_launch_fill_u32(c->d_fs_count, c->fsnpix, 0u, c->stream);
    if (c->d_fs_hash) p2m_launch_fill_u64(c->d_fs_hash, c->fsnpix, c->stream);
    HIPCHK(hipGetLastError());
    /* no wait: later work on the painter's stream is ordered after the
     * fills, and every readback waits for the stream first */
    c->hp_zero = true;
    c->hp_valid = !tiled;
    return 0;
}

/* Zero-fill the HEALPix arrays if a lazy clear left them unwritten. */
static int ensure_hp_valid(p2m_ctx *c)
{
    if (c->hp_valid) { HIPCHK(hipStreamSynchronize(c->stream)); return 0; }
    for (int m = 0; m < c->cfg.n_maps; ++m) {
        if (c->d_hp64[m]) p2m_launch_fill_f64(c->d_hp64[m], c->hpnpix, c->stream);
        if (c->d_hp32[m]) p2m_launch_fill_f32(c->d_hp32[m], c->hpnpix, c->stream);
    }
    if (c->d_count) p2m_launch_fill_u32(c->d_count, c->hpnpix, 0u, c->stream);
    if (c->d_hash) p2m_launch_fill_u64(c->d_hash, c->hpnpix, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->hp_valid = true;
    return 0;
}

int p2m_set_run(p2m_ctx *c, const p2m_run *r)
{
    if (!c || !r) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (!(r->h > 0) || !(r->omegam > 0)) return P2M_E_INVALID;
    HIPCHK(hipSetDevice(c->cfg.device));
    c->run = *r;
    /* set_rofztable, cosmology.f90:51-76 */
    const int n = 10000;
    double dz = (6.0 - 0.0) / (double)(n - 1);
    c->rofz.assign(n, 0.0);
    for (int i = 2; i <= n; ++i) {
        double z = (double)(i - 1) * dz + 0.0;
        c->rofz[i - 1] = c->rofz[i - 2] + (1.0 / dzdr(r->h, r->omegam, z)) * dz;
    }
    /* set_zofrtable, cosmology.f90:80-105 */
    double dr = (1.3e4 - 0.0) / (double)(n - 1);
    c->zofr.assign(n, 0.0);
    for (int i = 2; i <= n; ++i) {
        double z = c->zofr[i - 2];
        c->zofr[i - 1] = c->zofr[i - 2] + dzdr(r->h, r->omegam, z) * dr;
    }
    if (!c->d_rofz && dmalloc((void **)&c->d_rofz, sizeof(double) * n)) return P2M_E_NOMEM;
    if (!c->d_zofr && dmalloc((void **)&c->d_zofr, sizeof(double) * n)) return P2M_E_NOMEM;
    HIPCHK(hipMemcpy(c->d_rofz, c->rofz.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_zofr, c->zofr.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    c->mypi = 4.0 * p2m_atan(1.0);                         /* pks2map.f90:88 */
    c->cdev.rofz = c->d_rofz; c->cdev.zofr = c->d_zofr;
    c->cdev.drofz = dz; c->cdev.dzofr = dr; c->cdev.rofz_min = 0.0; c->cdev.zofr_min = 0.0;
    c->cdev.nrofz = n; c->cdev.nzofr = n;
    c->cdev.omegam = r->omegam; c->cdev.omegal = r->omegal;
    c->cdev.rho_0 = (2.775e11 * r->omegam) * (r->h * r->h);     /* pks2map.f90:127 */
    c->cdev.rhocrit_0 = 2.775e11 * (r->h * r->h);               /* :128 */
    c->cdev.mypi = c->mypi;
    c->cut_low = host_lookup(c->rofz, 0.0, dz, r->zmin);        /* :134 */
    c->cut_high = host_lookup(c->rofz, 0.0, dz, r->zmax);       /* :135 */
    c->have_run = true;
    return 0;
}

int p2m_rofz(p2m_ctx *c, double z, double *chi)
{
    if (!c || !chi) return P2M_E_INVALID;
    if (!c->have_run) return P2M_E_STATE;
    *chi = host_lookup(c->rofz, 0.0, 6.0 / 9999.0, z);
    return 0;
}

int p2m_zofr(p2m_ctx *c, double chi, double *z)
{
    if (!c || !z) return P2M_E_INVALID;
    if (!c->have_run) return P2M_E_STATE;
    *z = host_lookup(c->zofr, 0.0, 1.3e4 / 9999.0, chi);
    return 0;
}

int p2m_load_table(p2m_ctx *c, int slot, const int32_t hdr_i[3], const float hdr_f[6], const float *table, int64_t n_table)
{
    if (!c || !hdr_i || !hdr_f || !table) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (slot < 0 || slot >= c->cfg.n_maps) return P2M_E_INVALID;
    if (hdr_i[0] < 2 || hdr_i[1] < 2 || hdr_i[2] < 2) return P2M_E_SIZE;
    int64_t nt = 2LL * hdr_i[0] * hdr_i[1] * hdr_i[2];
    if (n_table != nt) return P2M_E_SIZE;
    if (!(hdr_f[0] > 0 && hdr_f[1] > hdr_f[0] && hdr_f[2] > 0 && hdr_f[3] > hdr_f[2] && hdr_f[4] > 0 && hdr_f[5] > hdr_f[4]))
        return P2M_E_INVALID;
    bool any = false;
    for (int m = 0; m < c->cfg.n_maps; ++m) any |= c->have_table[m];
    if (any && (memcmp(hdr_i, c->hdr_i, sizeof(c->hdr_i)) || memcmp(hdr_f, c->hdr_f, sizeof(c->hdr_f))))
        return P2M_E_SIZE;   /* all slots share one grid */
    HIPCHK(hipSetDevice(c->cfg.device));
    memcpy(c->hdr_i, hdr_i, sizeof(c->hdr_i));
    memcpy(c->hdr_f, hdr_f, sizeof(c->hdr_f));
    if (c->cfg.precision != P2M_PREC_PLAIN_FP64 && c->cfg.nside > 0) {
        /* the tiled painters stage the radial nodes in LDS: the default tile
         * shape must fit with this table; a shape set earlier that no longer
         * fits goes back to the library's choice */
        const int ml = c->cfg.precision == P2M_PREC_FP64 ? 1 : 2;
        const int mem = c->d_count != nullptr;
        if (p2m_tile_lds_bytes(ml, c->cfg.n_maps, 10, mem, hdr_i[2]) > 160 * 1024) return P2M_E_SIZE;
        if (c->tuned && c->log2w && p2m_tile_lds_bytes(ml, c->cfg.n_maps, c->log2w, mem, hdr_i[2]) > 160 * 1024)
            c->tuned = false;
    }
    TabGeom &g = c->g;
    g.nchit = hdr_i[0]; g.nmht = hdr_i[1]; g.nrt = hdr_i[2];
    g.chimint = hdr_f[0]; g.chimaxt = hdr_f[1]; g.mhmint = hdr_f[2]; g.mhmaxt = hdr_f[3];
    g.rmint = hdr_f[4]; g.rmaxt = hdr_f[5];
    /* maptable.f90:166-168 */
    g.dchit = (p2m_log(g.chimaxt) - p2m_log(g.chimint)) / (double)(g.nchit - 1);
    g.drt = (p2m_log(g.rmaxt) - p2m_log(g.rmint)) / (double)(g.nrt - 1);
    g.dmht = (p2m_log(g.mhmaxt) - p2m_log(g.mhmint)) / (double)(g.nmht - 1);
    g.lr0 = p2m_log(g.rmint); g.lc0 = p2m_log(g.chimint); g.lm0 = p2m_log(g.mhmint);
    /* padding covers every index interpolate_table can form past the end (Q6) */
    int64_t pad = 2LL * g.nrt * (g.nmht + 2) + 4LL * g.nrt + 64;
    /* the painters read the tables through buffer resources with 32-bit byte
     * offsets: one padded table, and the corner-quad table of all slots (16
     * bytes per map and node of each cell), must each stay below 2 GiB */
    if ((nt + pad) * 4 > INT32_MAX || (nt / 2) * 16 * c->cfg.n_maps > INT32_MAX) return P2M_E_SIZE;
    c->n_flat = nt + pad;
    if (!c->d_flat[slot] && dmalloc((void **)&c->d_flat[slot], sizeof(float) * c->n_flat)) return P2M_E_NOMEM;
    std::vector<float> flat(c->n_flat, 0.f);
    memcpy(flat.data(), table, sizeof(float) * nt);
    HIPCHK(hipMemcpy(c->d_flat[slot], flat.data(), sizeof(float) * c->n_flat, hipMemcpyHostToDevice));
    if (c->cfg.precision == P2M_PREC_FP32_COMP) {
        /* the profile's slice (column-major ir, imh, ichi), with the same
         * zero padding past the end for the upper-edge aliasing (Q6) */
        c->n_prof = nt / 2 + pad / 2;
        std::vector<float> prof((size_t)c->n_prof, 0.f);
        const int p = c->cfg.profile;
        for (int64_t i = 0; i < nt / 2; ++i) prof[(size_t)i] = table[(p - 1) + 2 * i];
        if (!c->d_prof[slot] && dmalloc((void **)&c->d_prof[slot], sizeof(float) * c->n_prof)) return P2M_E_NOMEM;
        HIPCHK(hipMemcpy(c->d_prof[slot], prof.data(), sizeof(float) * c->n_prof, hipMemcpyHostToDevice));
        c->n_quad = (int64_t)g.nchit * g.nmht * g.nrt * 4 * c->cfg.n_maps;
        if (!c->d_quad && dmalloc((void **)&c->d_quad, sizeof(float) * c->n_quad)) return P2M_E_NOMEM;
        p2m_launch_build_quad(c->d_prof[slot], c->n_prof, g.nrt, g.nmht, g.nchit, c->cfg.n_maps, slot, c->d_quad,
                              c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    /* FP32 path constants: node radii, ln r -> cell coordinate */
    std::vector<double> inv(g.nrt + 2);
    for (int k = 0; k < g.nrt + 2; ++k) inv[k] = p2m_exp(-2.0 * (g.lr0 + (double)k * g.drt));
    std::vector<float> invf(g.nrt + 2);
    for (int k = 0; k < g.nrt + 2; ++k) invf[k] = (float)inv[k];
    if (!c->d_inv_node2 && dmalloc((void **)&c->d_inv_node2, sizeof(double) * (g.nrt + 2))) return P2M_E_NOMEM;
    if (!c->d_inv_node2f && dmalloc((void **)&c->d_inv_node2f, sizeof(float) * (g.nrt + 2))) return P2M_E_NOMEM;
    HIPCHK(hipMemcpy(c->d_inv_node2, inv.data(), sizeof(double) * (g.nrt + 2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_inv_node2f, invf.data(), sizeof(float) * (g.nrt + 2), hipMemcpyHostToDevice));
    c->u_k1 = (float)(0.5 * 0.6931471805599453 / g.drt);   /* log2(r^2) -> (ln r)/drt */
    c->u_k0 = (float)(g.lr0 / g.drt);
    c->u_max = (float)(g.nrt - 1);
    /* the FP32 cell coordinate carries log2's error times u_k1 (about 10 on
     * the standard grid): a finer radial grid widens the band in proportion,
     * and a band of half a cell or more sends every interior pixel to FP64 */
    c->u_band = std::max(2.5e-4f, 2.5e-5f * c->u_k1);
    {
        /* r < rmint is clamped to rmint + 1e-5 (maptable.f90:190); its cell
         * and fraction as interpolate_table forms them (:196, :200), in FP64 */
        const double lrc = p2m_log(g.rmint + 1e-5);
        c->ir_clamp = (int)((lrc - g.lr0) / g.drt) + 1;
        c->fr_clamp = (float)(lrc - (g.lr0 + (double)(c->ir_clamp - 1) * g.drt));
    }
    c->have_table[slot] = true;
    return 0;
}

int p2m_set_tile(p2m_ctx *c, int32_t tile_w, int32_t tile_wg)
{
    if (!c) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (tile_w == 0 && tile_wg == 0) { c->tuned = false; return 0; }
    if (!is_pow2(tile_w) || tile_w < 64 || tile_w > (1 << p2m_tile_max_log2w())) return P2M_E_INVALID;
    if (tile_wg < 64 || tile_wg > 512 || tile_wg % 64) return P2M_E_INVALID;
    if (c->cfg.precision == P2M_PREC_FP32_COMP && tile_wg > 256) return P2M_E_INVALID;
    int l = 0;
    while ((1 << l) < tile_w) ++l;
    if (p2m_tile_lds_bytes(c->cfg.precision == P2M_PREC_FP64 ? 1 : 2, c->cfg.n_maps, l, c->d_count != nullptr,
                           c->have_table[0] ? c->g.nrt : 200) > 160 * 1024)
        return P2M_E_INVALID;
    HIPCHK(hipSetDevice(c->cfg.device));
    int rc = build_tiles(c, l);
    if (rc) return rc;
    c->wg = tile_wg;
    c->tuned = true;
    return 0;
}

}  /* extern "C" */

/* ------------------------------------------------------------------ paint */
/* exclusive scan of n uint32 counts into uint64 offsets, growing the temp storage */
static int scan_counts_on(const uint32_t *cnt, unsigned long long *off, int64_t n, void **tmp, size_t *tmp_bytes,
                          hipStream_t s)
{
    size_t need = 0;
    if (p2m_scan_u32(cnt, off, n, nullptr, &need, s)) return P2M_E_HIP;
    if (need > *tmp_bytes || !*tmp) {
        dfree(*tmp); *tmp = nullptr; *tmp_bytes = 0;
        if (dmalloc(tmp, need + 256)) return P2M_E_NOMEM;
        *tmp_bytes = need + 256;
    }
    size_t bytes = *tmp_bytes;
    if (p2m_scan_u32(cnt, off, n, *tmp, &bytes, s)) return P2M_E_HIP;
    return 0;
}

static int scan_counts(p2m_ctx *c, const uint32_t *cnt, unsigned long long *off, int64_t n)
{
    return scan_counts_on(cnt, off, n, &c->d_scan_tmp, &c->scan_tmp_bytes, c->stream);
}

/* The per-halo buffers exist after the first paint call even for an empty
 * catalogue: the binning writes the terminating ring count nr[n_halo]. */
/* counter block summed over its shards (blocks until the stream is idle) */
static int read_counters(p2m_ctx *c, unsigned long long out[Ctr::TOTAL], unsigned long long *raw = nullptr)
{
    unsigned long long hbuf[Ctr::ALL];
    unsigned long long *h = raw ? raw : hbuf;
    HIPCHK(hipMemcpyAsync(h, c->d_ctr, sizeof(hbuf), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < Ctr::TOTAL; ++k) {
        out[k] = 0;
        for (int sh = 0; sh < P2M_SHARDS; ++sh) out[k] += h[sh * Ctr::TOTAL + k];
    }
    return 0;
}

static int ensure_halo_cap(p2m_ctx *c, int64_t n)
{
    if (n <= c->halo_cap && c->d_halo) return 0;
    dfree(c->d_halo); c->d_halo = nullptr; c->halo_cap = 0;
    dfree(c->d_flat_idx); c->d_flat_idx = nullptr;
    dfree(c->d_nr); c->d_nr = nullptr;
    dfree(c->d_ring_off); c->d_ring_off = nullptr;
    dfree(c->d_halo_s); c->d_halo_s = nullptr;
    dfree(c->d_hot); c->d_hot = nullptr;
    dfree(c->d_bin); c->d_bin = nullptr;
    dfree(c->d_keys); c->d_keys = nullptr;
    dfree(c->d_vals); c->d_vals = nullptr;
    int64_t cap = std::max<int64_t>(n, 1024);
    if (dmalloc((void **)&c->d_halo, sizeof(HaloRec) * cap)) return P2M_E_NOMEM;
    if (dmalloc((void **)&c->d_flat_idx, sizeof(uint32_t) * cap)) return P2M_E_NOMEM;
    if (dmalloc((void **)&c->d_nr, sizeof(uint32_t) * (cap + 1))) return P2M_E_NOMEM;
    if (dmalloc((void **)&c->d_ring_off, sizeof(unsigned long long) * (cap + 1))) return P2M_E_NOMEM;
    if (c->cfg.precision != P2M_PREC_PLAIN_FP64 && c->cfg.nside > 0) {
        if (dmalloc((void **)&c->d_bin, sizeof(BinHalo) * cap)) return P2M_E_NOMEM;
        if (c->cfg.precision == P2M_PREC_FP32_COMP) {
            if (dmalloc((void **)&c->d_hot, sizeof(TileHalo) * cap)) return P2M_E_NOMEM;
        } else {
            if (dmalloc((void **)&c->d_halo_s, sizeof(HaloRec) * cap)) return P2M_E_NOMEM;
        }
        if (dmalloc((void **)&c->d_keys, sizeof(uint32_t) * 2 * cap)) return P2M_E_NOMEM;
        if (dmalloc((void **)&c->d_vals, sizeof(uint32_t) * 2 * cap)) return P2M_E_NOMEM;
    }
    c->halo_cap = cap;
    return 0;
}

static int decode_chunk(p2m_ctx *c, const float *d_src, int64_t n, int64_t base)
{
    DecodeArgs a;
    a.halos = d_src; a.n = n; a.index_base = base;
    a.c = c->cdev; a.g = c->g;
    a.cut_low = c->cut_low; a.cut_high = c->cut_high; a.chihview = c->run.chihview; a.mmin = c->run.mmin;
    a.precut = (c->run.flags & P2M_RUN_PRECUT) ? 1 : 0;
    a.dmodel = c->cfg.model == 2 ? 500.0 : 200.0;
    a.profile = c->cfg.profile; a.nside = c->cfg.nside; a.npix_fs = c->cfg.npix_fs;
    a.fov = ((c->cfg.fov_deg / 360.0) * 2.0) * c->mypi;        /* pks2map.f90:92 */
    a.dpix = c->cfg.npix_fs > 0 ? a.fov / (double)c->cfg.npix_fs : 1.0;
    a.out = c->d_halo; a.cap = c->halo_cap;
    a.flat_idx = c->d_flat_idx;
    a.cand = c->d_cand;
    a.n_cand = c->d_ctr + Ctr::N_CAND;
    a.n_cand0 = c->d_ctr + Ctr::N_CAND0;
    a.n_flat_out = c->d_ctr + Ctr::N_FLAT;
    a.counters = c->d_ctr + Ctr::DEC;
    a.keys = nullptr; a.vals = nullptr; a.nphi = 1; a.key_max = 0;
    if (c->d_bin && c->cfg.nside > 0) {
        a.keys = c->d_keys; a.vals = c->d_vals;
        a.nphi = std::max(1, c->cfg.nside / 128);
        a.key_max = p2m_key_max(c->cfg.nside, a.nphi);
    }
    p2m_launch_decode(a, c->stream);
    return hipGetLastError() == hipSuccess ? 0 : P2M_E_HIP;
}

static int autotune(p2m_ctx *c, int64_t n_halo, int64_t n_flat, unsigned long long n_items);

static int launch_flat(p2m_ctx *c, PaintArgs &pa, int64_t n_flat, hipStream_t s);

/* the main stream waits for whatever the flat-sky stream has enqueued */
static void fs_join(p2m_ctx *c)
{
    if (hipEventRecord(c->ev_fs, c->fs_stream) == hipSuccess) (void)hipStreamWaitEvent(c->stream, c->ev_fs, 0);
    (void)hipGetLastError();
}

static int paint_halos(p2m_ctx *c, int64_t n_halo, int64_t n_flat, unsigned long long n_items, bool allow_tune)
{
    const int mode = c->cfg.precision;
    unsigned long long *pc = c->d_ctr + Ctr::PAINT;
    c->last_fast = false;
    PaintArgs pa;
    memset(&pa, 0, sizeof(pa));
    pa.halos = c->d_halo; pa.n_halo = n_halo; pa.nside = c->cfg.nside; pa.npix_fs = c->cfg.npix_fs;
    pa.flat_idx = c->d_flat_idx; pa.n_flat_dev = c->d_ctr + Ctr::N_FLAT; pa.n_flat_host = n_flat;
    pa.n_maps = c->cfg.n_maps; pa.profile = c->cfg.profile;
    pa.fov = ((c->cfg.fov_deg / 360.0) * 2.0) * c->mypi;
    pa.dpix = c->cfg.npix_fs > 0 ? pa.fov / (double)c->cfg.npix_fs : 1.0;
    pa.g = c->g; pa.n_flat = c->n_flat;
    for (int m = 0; m < c->cfg.n_maps; ++m) { pa.tab_flat[m] = c->d_flat[m]; pa.map64[m] = c->d_hp64[m]; pa.fs64[m] = c->d_fs64[m]; }
    pa.count = c->d_count; pa.hash = c->d_hash; pa.fs_count = c->d_fs_count; pa.fs_hash = c->d_fs_hash;
    pa.counters = pc;

    if (mode != P2M_PREC_PLAIN_FP64 && c->cfg.nside > 0 && !c->tuned) {
        /* tile shape first: the tuner's trials paint and clear both maps, so
         * this call's own painting (the flat-sky map included) starts after
         * the last trial */
        if (allow_tune) { int rc = autotune(c, n_halo, n_flat, n_items); if (rc) return rc; }
        else { int rc = build_tiles(c, 10); if (rc) return rc; c->wg = 256; c->tuned = true; }
    }
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    if (mode == P2M_PREC_PLAIN_FP64) {
        if (c->cfg.nside > 0) p2m_launch_paint_plain_hp(pa, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[3], c->stream));
        if (c->cfg.npix_fs > 0) p2m_launch_paint_fs(pa, 1, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        HIPCHK(hipEventRecord(c->ev[5], c->stream));
        c->hp_zero = false;
        return 0;
    }
    if (c->cfg.nside > 0) {
        /* the flat-sky painter needs only the decoded records: it runs on
         * its own stream beside the binning and the tile painter */
        HIPCHK(hipEventRecord(c->ev_dec, c->stream));
        HIPCHK(hipStreamWaitEvent(c->fs_stream, c->ev_dec, 0));
        { int rc = launch_flat(c, pa, n_flat, c->fs_stream); if (rc) { fs_join(c); return rc; } }
        HIPCHK(hipEventRecord(c->ev_fs, c->fs_stream));
        /* every return from here on, error or not, first orders the work
         * that follows on the main stream after the flat-sky painter */
        struct FsJoin {
            p2m_ctx *c;
            bool armed;
            ~FsJoin() { if (armed) (void)hipStreamWaitEvent(c->stream, c->ev_fs, 0); }
        } fs_guard{c, true};
        /* painted halos in spatial order (keys written by decode):
         * neighbouring halos share tiles */
        {
            const uint32_t key_max = p2m_key_max(c->cfg.nside, std::max(1, c->cfg.nside / 128));
            size_t need = 0;
            if (p2m_sort_halos(c->d_halo, n_halo, key_max, c->d_bin, c->d_hot, c->d_halo_s, c->g.nrt, c->g.nmht,
                               c->d_keys, c->d_vals, nullptr, &need, c->d_nr, c->stream)) return P2M_E_HIP;
            if (need > c->sort_tmp_bytes || !c->d_sort_tmp) {
                dfree(c->d_sort_tmp); c->d_sort_tmp = nullptr; c->sort_tmp_bytes = 0;
                if (dmalloc(&c->d_sort_tmp, need + 256)) return P2M_E_NOMEM;
                c->sort_tmp_bytes = need + 256;
            }
            size_t bytes = c->sort_tmp_bytes;
            if (p2m_sort_halos(c->d_halo, n_halo, key_max, c->d_bin, c->d_hot, c->d_halo_s, c->g.nrt, c->g.nmht,
                               c->d_keys, c->d_vals, c->d_sort_tmp, &bytes, c->d_nr, c->stream)) return P2M_E_HIP;
        }
        const uint32_t *order = c->d_vals + n_halo;   /* spatial position -> slot */
        BinArgs b;
        memset(&b, 0, sizeof(b));
        b.halos = c->d_bin; b.n_halo = n_halo; b.nside = c->cfg.nside; b.log2w = c->log2w;
        b.nr = c->d_nr; b.ring_off = c->d_ring_off;
        b.tile_base = c->d_tile_base; b.tile_cnt = c->d_tile_cnt;
        b.tile_off = c->d_tile_off; b.segs = c->d_segs; b.counters = pc; b.ring_z = c->d_ring_z;
        /* (halo, ring) item space: the gather wrote the item count of every
         * halo (with no halo it did not run) */
        if (n_halo <= 0) p2m_launch_ring_counts(b, c->stream);
        { int rc = scan_counts(c, c->d_nr, c->d_ring_off, n_halo + 1); if (rc) return rc; }
        /* n_items (= ring_off[n_halo]) was counted by the decode */
        if ((int64_t)n_items > c->runs_cap || !c->d_runs || !c->d_pos) {
            dfree(c->d_runs); dfree(c->d_pos); c->d_runs = nullptr; c->d_pos = nullptr; c->runs_cap = 0;
            int64_t cap = (int64_t)(n_items * 1.25) + 1024;
            if (dmalloc((void **)&c->d_runs, sizeof(uint32_t) * cap)) return P2M_E_NOMEM;
            if (dmalloc((void **)&c->d_pos, sizeof(uint32_t) * cap)) return P2M_E_NOMEM;
            c->runs_cap = cap;
        }
        if (!c->d_ovf || c->ovf_cap < (int64_t)(n_items / 8) + (1 << 16)) {
            /* about 8% of the items cross a tile boundary at the default shape */
            dfree(c->d_ovf); c->d_ovf = nullptr; c->ovf_cap = 0;
            const int64_t cap = (int64_t)(n_items / 8) + (1 << 16);
            if (dmalloc((void **)&c->d_ovf, sizeof(uint4) * cap)) return P2M_E_NOMEM;
            c->ovf_cap = cap;
        }
        if ((int64_t)(n_items / 64 + 2) > c->group_cap || !c->d_group) {
            dfree(c->d_group); c->d_group = nullptr; c->group_cap = 0;
            const int64_t cap = (int64_t)((n_items / 64 + 2) * 1.25) + 1024;
            if (dmalloc((void **)&c->d_group, sizeof(uint32_t) * cap)) return P2M_E_NOMEM;
            c->group_cap = cap;
        }
        b.n_items = n_items; b.runs = c->d_runs; b.pos = c->d_pos; b.group_halo = c->d_group;
        b.ovf = c->d_ovf; b.ovf_cap = c->ovf_cap; b.n_ovf = pc + Ctr::OVF;
        b.n_tiles = c->n_tiles; b.err = c->d_ctr + Ctr::ERR;
        p2m_launch_fill_u32(c->d_tile_cnt, c->n_tiles + 1, 0u, c->stream);
        p2m_launch_seg_count(b, c->stream);
        { int rc = scan_counts(c, c->d_tile_cnt, c->d_tile_off, c->n_tiles + 1); if (rc) return rc; }
        /* segment and listed-pixel counts back to the host (pinned memory) */
        unsigned long long nseg = 0, nlisted = 0;
        {
            unsigned long long hbuf[Ctr::ALL + 1];
            unsigned long long *hb = c->h_pin ? c->h_pin : hbuf;
            HIPCHK(hipMemcpyAsync(hb, c->d_ctr, sizeof(unsigned long long) * Ctr::ALL, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipMemcpyAsync(hb + Ctr::ALL, c->d_tile_off + c->n_tiles, sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipEventRecord(c->ev[6], c->stream));
            HIPCHK(hipEventSynchronize(c->ev[6]));
            const unsigned long long n_ovf = hb[Ctr::PAINT + Ctr::OVF];
            if ((int64_t)n_ovf > c->ovf_cap) {
                /* pieces beyond the overflow list (runs spanning three or
                 * more tiles): grow it and count again */
                dfree(c->d_ovf); c->d_ovf = nullptr; c->ovf_cap = 0;
                const int64_t cap = (int64_t)(n_ovf * 1.25) + 1024;
                if (dmalloc((void **)&c->d_ovf, sizeof(uint4) * cap)) return P2M_E_NOMEM;
                c->ovf_cap = cap;
                b.ovf = c->d_ovf; b.ovf_cap = cap;
                p2m_launch_clear_ctr(pc, 0, 1, c->stream);
                p2m_launch_clear_ctr(pc, 4, 1, c->stream);
                p2m_launch_clear_ctr(pc, Ctr::OVF, 1, c->stream);
                p2m_launch_fill_u32(c->d_tile_cnt, c->n_tiles + 1, 0u, c->stream);
                p2m_launch_seg_count(b, c->stream);
                { int rc = scan_counts(c, c->d_tile_cnt, c->d_tile_off, c->n_tiles + 1); if (rc) return rc; }
                HIPCHK(hipMemcpyAsync(hb, c->d_ctr, sizeof(unsigned long long) * Ctr::ALL, hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipMemcpyAsync(hb + Ctr::ALL, c->d_tile_off + c->n_tiles, sizeof(unsigned long long),
                                      hipMemcpyDeviceToHost, c->stream));
                HIPCHK(hipStreamSynchronize(c->stream));
            }
            nseg = hb[Ctr::ALL];
            for (int sh = 0; sh < P2M_SHARDS; ++sh) nlisted += hb[sh * Ctr::TOTAL + Ctr::PAINT + 0];
        }
        if (mode == P2M_PREC_FP32_COMP) {
            /* deferred FP64 pixels: ~5e-4 of the listed pixels in practice;
             * a list that fills up costs a repaint of the tiles that lost
             * pixels, after the list has grown (P2M_FB_CAP caps it, so that
             * the repaint can be exercised) */
            int64_t need = (int64_t)(nlisted / 256) + (1 << 18);
            if (const char *e = getenv("P2M_FB_CAP")) {
                const long long cap = atoll(e);
                if (cap >= P2M_SHARDS && cap < need) need = cap;
            }
            if (need != c->fb_cap && (need > c->fb_cap || getenv("P2M_FB_CAP"))) {
                dfree(c->d_fb); c->d_fb = nullptr; c->fb_cap = 0;
                if (dmalloc((void **)&c->d_fb, sizeof(uint4) * need)) return P2M_E_NOMEM;
                c->fb_cap = need;
            }
        }
        if ((int64_t)nseg > c->seg_cap || !c->d_segs) {
            dfree(c->d_segs); c->d_segs = nullptr; c->seg_cap = 0;
            int64_t cap = (int64_t)(nseg * 1.25) + 1024;
            if (dmalloc((void **)&c->d_segs, sizeof(uint2) * cap)) return P2M_E_NOMEM;
            c->seg_cap = cap;
        }
        b.segs = c->d_segs;
        p2m_launch_seg_emit(b, c->stream);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[3], c->stream));
        TileArgs t;
        memset(&t, 0, sizeof(t));
        t.halos = c->d_halo_s; t.hot = c->d_hot; t.nside = c->cfg.nside; t.log2w = c->log2w; t.n_rings = c->n_rings;
        t.n_maps = c->cfg.n_maps; t.profile = c->cfg.profile; t.accumulate = c->hp_zero ? 0 : 1;
        t.n_tiles = c->n_tiles; t.tile_base = c->d_tile_base; t.desc = c->d_tile_desc; t.tile_off = c->d_tile_off; t.segs = c->d_segs;
        t.g = c->g; t.n_flat = c->n_flat; t.inv_node2 = c->d_inv_node2;
        t.u_k1 = c->u_k1; t.u_k0 = c->u_k0; t.u_max = c->u_max; t.band = c->u_band; t.fr_clamp = c->fr_clamp; t.ir_clamp = c->ir_clamp;
        t.band_c = 4e-6f;
        t.inv_node2f = c->d_inv_node2f;
        t.n_quad = c->n_quad; t.tab_q = c->d_quad;
        t.fb = c->d_fb; t.fb_count = pc + 6; t.fb_cap = c->fb_cap;
        t.lost = c->d_lost; t.tile_next = c->d_tile_next; t.redo = 0; t.n_halo = n_halo; t.err = c->d_ctr + Ctr::ERR; t.n_lost = c->d_ctr + Ctr::LOST;
        for (int m = 0; m < c->cfg.n_maps; ++m) {
            t.tab_flat[m] = c->d_flat[m]; t.map64[m] = c->d_hp64[m]; t.map32[m] = c->d_hp32[m];
        }
        t.count = c->d_count; t.hash = c->d_hash; t.counters = pc;
        /* persistent grid of exactly the resident workgroups (a multiple
         * of 8 for the XCD-aware tile order): no partial last round */
        int grid = (int)std::min<int64_t>(c->n_tiles, 8192);
        {
            const size_t lds = p2m_tile_lds_bytes(mode, c->cfg.n_maps, c->log2w, c->d_count != nullptr, c->g.nrt);
            const int occ = p2m_tile_occupancy(mode, c->cfg.n_maps, c->wg, lds);
            if (occ > 0 && c->n_cu > 0) {
                const int64_t res = ((int64_t)occ * c->n_cu) & ~(int64_t)7;
                if (res >= 8) grid = (int)std::min<int64_t>(c->n_tiles, res);
            }
        }
        if (p2m_launch_tile_paint(t, mode, c->cfg.n_maps, grid, c->wg, c->stream)) return P2M_E_INVALID;
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(c->ev[7], c->stream));
        if (mode == P2M_PREC_FP32_COMP) {
            FallbackArgs f;
            memset(&f, 0, sizeof(f));
            f.halos = c->d_halo; f.order = order; f.fb = c->d_fb; f.fb_count = pc + 6; f.fb_cap = c->fb_cap;
            f.nside = c->cfg.nside; f.n_maps = c->cfg.n_maps; f.profile = c->cfg.profile;
            f.g = c->g; f.n_flat = c->n_flat;
            for (int m = 0; m < c->cfg.n_maps; ++m) { f.tab_flat[m] = c->d_flat[m]; f.map32[m] = c->d_hp32[m]; }
            f.count = c->d_count; f.hash = c->d_hash; f.counters = pc;
            f.lost = c->d_lost; f.tile_base = c->d_tile_base; f.log2w = c->log2w; f.n_halo = n_halo;
            f.err = c->d_ctr + Ctr::ERR;
            if (p2m_launch_fallback(f, (int64_t)(nlisted / 256) + 1024, c->stream)) return P2M_E_HIP;
            c->last_t = t; c->last_f = f; c->last_grid = grid; c->last_nlisted = (int64_t)nlisted;
        }
        c->last_fast = mode == P2M_PREC_FP32_COMP;
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        fs_guard.armed = false;
        HIPCHK(hipStreamWaitEvent(c->stream, c->ev_fs, 0));
    } else {
        HIPCHK(hipEventRecord(c->ev[3], c->stream));
        HIPCHK(hipEventRecord(c->ev[4], c->stream));
        { int rc = launch_flat(c, pa, n_flat, c->stream); if (rc) return rc; }
    }
    HIPCHK(hipEventRecord(c->ev[5], c->stream));
    c->hp_zero = false;
    return 0;
}

/* flat-sky painter over the flattened box pixels on stream s, between
 * events 8 and 9 */
static int launch_flat(p2m_ctx *c, PaintArgs &pa, int64_t n_flat, hipStream_t s)
{
    HIPCHK(hipEventRecord(c->ev[8], s));
    if (c->cfg.npix_fs > 0 && n_flat > 0) {
        if (n_flat + 1 > c->fs_cap || !c->d_fs_cnt) {
            dfree(c->d_fs_cnt); dfree(c->d_fs_off); c->d_fs_cnt = nullptr; c->d_fs_off = nullptr; c->fs_cap = 0;
            const int64_t cap = n_flat + 1 + 1024;
            if (dmalloc((void **)&c->d_fs_cnt, sizeof(uint32_t) * cap)) return P2M_E_NOMEM;
            if (dmalloc((void **)&c->d_fs_off, sizeof(unsigned long long) * cap)) return P2M_E_NOMEM;
            c->fs_cap = cap;
        }
        pa.fs_cnt = c->d_fs_cnt; pa.fs_off = c->d_fs_off;
        p2m_launch_fs_counts(pa, s);
        {
            /* its own scan storage when beside the binning */
            const bool own = s != c->stream;
            int rc = scan_counts_on(c->d_fs_cnt, c->d_fs_off, n_flat + 1, own ? &c->d_scan_tmp_fs : &c->d_scan_tmp,
                                    own ? &c->scan_tmp_fs_bytes : &c->scan_tmp_bytes, s);
            if (rc) return rc;
        }
        p2m_launch_fs_paint(pa, s);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[9], s));
    return 0;
}

/* The FP32 painter's deferred-pixel list filled up in some sub-list: grow
 * it so that every sub-list holds the whole demand, and paint the tiles
 * marked lost again (they wrote nothing; entries they did store are skipped
 * by the fallback), then decide their deferred pixels.  Repeats until no
 * tile is lost; `raw` holds the counters on entry and on return. */
static int repaint_lost(p2m_ctx *c, unsigned long long h[Ctr::TOTAL], unsigned long long raw[Ctr::ALL])
{
    unsigned long long *pc = c->d_ctr + Ctr::PAINT;
    c->n_redo = 0;
    for (int iter = 0;; ++iter) {
        unsigned long long demand = 0;
        bool over = false;
        for (int sh = 0; sh < P2M_SHARDS; ++sh) {
            const unsigned long long v = raw[sh * Ctr::TOTAL + Ctr::PAINT + 6];
            demand += v;
            if ((int64_t)v > c->fb_cap / P2M_SHARDS) over = true;
        }
        if (!over) return 0;
        c->n_redo += (int)h[Ctr::LOST];
        if (iter >= 3) {
            /* cannot happen: the grown list holds every pass's whole demand */
            p2m_launch_fill_u32((uint32_t *)c->d_lost, (c->n_tiles + 3) / 4, 0u, c->stream);
            (void)hipStreamSynchronize(c->stream);
            snprintf(g_hip_detail, sizeof(g_hip_detail), "deferred-pixel list still full after %d repaints", iter);
            return P2M_E_NOMEM;
        }
        const int64_t cap = (int64_t)P2M_SHARDS * (int64_t)(demand + 64);
        dfree(c->d_fb); c->d_fb = nullptr; c->fb_cap = 0;
        if (dmalloc((void **)&c->d_fb, sizeof(uint4) * cap)) return P2M_E_NOMEM;
        c->fb_cap = cap;
        p2m_launch_clear_ctr(pc, 6, 1, c->stream);
        p2m_launch_clear_ctr(c->d_ctr, Ctr::LOST, 1, c->stream);
        TileArgs t = c->last_t;
        t.fb = c->d_fb; t.fb_cap = cap; t.redo = 1;
        if (p2m_launch_tile_paint(t, P2M_PREC_FP32_COMP, c->cfg.n_maps, c->last_grid, c->wg, c->stream))
            return P2M_E_INVALID;
        HIPCHK(hipGetLastError());
        FallbackArgs f = c->last_f;
        f.fb = c->d_fb; f.fb_cap = cap;
        if (p2m_launch_fallback(f, (int64_t)demand + 1024, c->stream)) return P2M_E_HIP;
        int rc = read_counters(c, h, raw);
        if (rc) return rc;
    }
}

static int finish_stats(p2m_ctx *c)
{
    unsigned long long h[Ctr::TOTAL], raw[Ctr::ALL];
    int rc = read_counters(c, h, raw);
    if (rc) return rc;
    c->n_redo = 0;
    if (c->last_fast) {
        rc = repaint_lost(c, h, raw);
        if (rc) return rc;
    }
    if (h[Ctr::ERR]) {
        snprintf(g_hip_detail, sizeof(g_hip_detail),
                 "%llu device-side range checks failed (segment, tile slot or deferred entry out of range)",
                 h[Ctr::ERR]);
        return P2M_E_INTERNAL;
    }
    p2m_stats &s = c->st;
    s.n_cut = (int64_t)h[Ctr::DEC + 0];
    s.n_radius_err = (int64_t)h[Ctr::DEC + 2];
    s.n_painted = (int64_t)h[Ctr::N_CAND] - s.n_radius_err;
    s.n_listed_hp = (int64_t)h[Ctr::PAINT + 0];
    s.n_upd_hp = (int64_t)h[Ctr::PAINT + 1];
    s.n_upd_fs = (int64_t)h[Ctr::PAINT + 2];
    s.n_ring_gaps = (int64_t)h[Ctr::PAINT + 3];
    s.n_segments = (int64_t)h[Ctr::PAINT + 4];
    s.n_fallback = (int64_t)h[Ctr::PAINT + 5];
    s.n_redo_tiles = c->n_redo;
    s.cut_low = c->cut_low; s.cut_high = c->cut_high;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[5]); s.ms_total = ms;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]); s.ms_decode = ms;
    float ms_fs = 0;
    if (c->cfg.precision != P2M_PREC_PLAIN_FP64) (void)hipEventElapsedTime(&ms_fs, c->ev[8], c->ev[9]);
    /* the flat-sky painter of the tiled modes runs on its own stream beside
     * the binning: its time is reported as the finish phase */
    (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
    s.ms_bin = ms;
    (void)hipEventElapsedTime(&ms, c->ev[3], c->ev[4]); s.ms_paint = ms;
    if (c->cfg.precision == P2M_PREC_PLAIN_FP64) { (void)hipEventElapsedTime(&ms, c->ev[4], c->ev[5]); s.ms_finish = ms; }
    else s.ms_finish = ms_fs;
    s.ms_tile = 0.0;
    if (c->cfg.precision != P2M_PREC_PLAIN_FP64 && c->cfg.nside > 0) {
        (void)hipEventElapsedTime(&ms, c->ev[3], c->ev[7]); s.ms_tile = ms;
    }
    s.tile_w = c->log2w ? (1 << c->log2w) : 0;
    s.tile_wg = c->wg;
    return 0;
}

/* page-locked host memory (hipHostMalloc or otherwise pinned)? */
static bool host_is_pinned(const void *p)
{
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) { (void)hipGetLastError(); return false; }
    return at.type == hipMemoryTypeHost;
}

/* A catalogue pass is begin -> push (one or more blocks) -> finish.  begin
 * sizes the per-halo buffers for n_total halos and starts the call's clock;
 * push decodes a block, copying host blocks to the device in pieces on the
 * copy stream (two device slots used in turn, each piece decoded on the
 * painter's stream as soon as it has arrived); finish bins and paints. */
static int pass_begin(p2m_ctx *c, int64_t n_total)
{
    if (!c || n_total < 0) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (!c->have_run) return P2M_E_STATE;
    for (int m = 0; m < c->cfg.n_maps; ++m) if (!c->have_table[m]) return P2M_E_STATE;
    HIPCHK(hipSetDevice(c->cfg.device));
    int rc = ensure_halo_cap(c, n_total);
    if (rc) return rc;
    if (n_total > 0 && n_total > c->cand_cap) {
        dfree(c->d_cand); c->d_cand = nullptr; c->cand_cap = 0;
        if (dmalloc((void **)&c->d_cand, sizeof(uint32_t) * n_total)) return P2M_E_NOMEM;
        c->cand_cap = n_total;
    }
    p2m_launch_fill_u64(c->d_ctr, Ctr::ALL, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    c->in_stream = true;
    c->stream_total = n_total;
    c->stream_pushed = 0;
    c->stream_h2d_bytes = 0;
    c->n_host_push = 0;
    c->n_pieces = 0;
    c->stream_pinned = true;
    return 0;
}

static int pass_push_impl(p2m_ctx *c, const float *src, bool on_device, int64_t n, int64_t index_base)
{
    if (n == 0) return 0;
    if (!on_device) {
        /* the previous host block's copies have finished: its buffer is the
         * caller's again (the double-buffer contract of p2m_stream_push) */
        if (c->n_host_push > 0) HIPCHK(hipEventSynchronize(c->ev_push));
        const bool pinned = host_is_pinned(src);
        c->stream_pinned = c->stream_pinned && pinned;
        /* pieces of 2^21 records (58.7 MB): long enough for the DMA engine's
         * full rate, short enough that the decode trails the copy closely */
        const int64_t piece = std::min<int64_t>(std::min<int64_t>(c->cfg.chunk_halos, 1 << 21), n);
        if (piece > c->slot_cap || !c->d_slot[0] || !c->d_slot[1]) {
            /* slots in use by an earlier piece are released only after the
             * painter's stream has passed them */
            HIPCHK(hipStreamSynchronize(c->stream));
            dfree(c->d_slot[0]); dfree(c->d_slot[1]);
            c->d_slot[0] = c->d_slot[1] = nullptr; c->slot_cap = 0;
            const int64_t cap = std::max<int64_t>(piece, 1 << 16);
            if (dmalloc((void **)&c->d_slot[0], sizeof(float) * 7 * cap)) return P2M_E_NOMEM;
            if (dmalloc((void **)&c->d_slot[1], sizeof(float) * 7 * cap)) return P2M_E_NOMEM;
            c->slot_cap = cap;
            c->n_pieces = 0;
        }
        /* copies start after the call's start mark, i.e. after every earlier
         * use of the slots on the painter's stream */
        if (c->n_host_push == 0) HIPCHK(hipStreamWaitEvent(c->h2d_stream, c->ev[0], 0));
        for (int64_t c0 = 0; c0 < n; c0 += piece) {
            const int64_t nc = std::min(piece, n - c0);
            const int slot = c->n_pieces & 1;
            if (c->n_pieces >= 2) HIPCHK(hipStreamWaitEvent(c->h2d_stream, c->ev_free[slot], 0));
            HIPCHK(hipMemcpyAsync(c->d_slot[slot], src + 7 * c0, sizeof(float) * 7 * nc, hipMemcpyHostToDevice,
                                  c->h2d_stream));
            HIPCHK(hipEventRecord(c->ev_copy[slot], c->h2d_stream));
            HIPCHK(hipStreamWaitEvent(c->stream, c->ev_copy[slot], 0));
            int rc = decode_chunk(c, c->d_slot[slot], nc, index_base + c0);
            if (rc) return rc;
            HIPCHK(hipEventRecord(c->ev_free[slot], c->stream));
            c->n_pieces++;
        }
        HIPCHK(hipEventRecord(c->ev_push, c->h2d_stream));
        c->stream_h2d_bytes += (int64_t)sizeof(float) * 7 * n;
        c->n_host_push++;
        return 0;
    }
    const int64_t chunk = c->cfg.chunk_halos;
    for (int64_t c0 = 0; c0 < n; c0 += chunk) {
        const int64_t nc = std::min(chunk, n - c0);
        int rc = decode_chunk(c, src + 7 * c0, nc, index_base + c0);
        if (rc) return rc;
    }
    return 0;
}

static int pass_push(p2m_ctx *c, const float *src, bool on_device, int64_t n, int64_t index_base)
{
    if (!c || n < 0 || (!src && n > 0) || index_base < 0) return P2M_E_INVALID;
    if (!c->in_stream) return P2M_E_STATE;
    if (n > c->stream_total - c->stream_pushed) return P2M_E_SIZE;
    HIPCHK(hipSetDevice(c->cfg.device));
    int rc = pass_push_impl(c, src, on_device, n, index_base);
    if (rc) {
        /* the pass is abandoned: wait for what was queued */
        (void)hipStreamSynchronize(c->h2d_stream);
        (void)hipStreamSynchronize(c->stream);
        c->in_stream = false;
        return rc;
    }
    c->stream_pushed += n;
    return 0;
}

static int pass_finish(p2m_ctx *c, bool allow_tune)
{
    if (!c) return P2M_E_INVALID;
    if (!c->in_stream) return P2M_E_STATE;
    c->in_stream = false;
    HIPCHK(hipSetDevice(c->cfg.device));
    if (c->n_host_push > 0) HIPCHK(hipEventRecord(c->ev_h2d_end, c->h2d_stream));
    HIPCHK(hipEventRecord(c->ev[1], c->stream));
    unsigned long long hc[Ctr::TOTAL];
    int rc = read_counters(c, hc);
    if (rc) return rc;
    int64_t n_halo = (int64_t)std::min<unsigned long long>(hc[Ctr::N_CAND], (unsigned long long)c->halo_cap);
    int64_t n_flat = (int64_t)std::min<unsigned long long>(hc[Ctr::N_FLAT], (unsigned long long)c->halo_cap);
    rc = paint_halos(c, n_halo, n_flat, hc[Ctr::ITEMS], allow_tune);
    if (rc) {
        /* nothing queued may outlive the call */
        (void)hipStreamSynchronize(c->fs_stream);
        (void)hipStreamSynchronize(c->stream);
        return rc;
    }
    c->hp_valid = true;
    c->st.n_in = c->stream_pushed;
    rc = finish_stats(c);
    if (rc) return rc;
    c->st.h2d_bytes = c->stream_h2d_bytes;
    c->st.h2d_pinned = c->n_host_push > 0 && c->stream_pinned ? 1 : 0;
    c->st.ms_h2d = 0.0;
    if (c->n_host_push > 0) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev_h2d_end);
        c->st.ms_h2d = ms;
    }
    return 0;
}

static int paint_common(p2m_ctx *c, const float *src, bool on_device, int64_t n, int64_t index_base, bool allow_tune)
{
    if (!c || (!src && n > 0) || n < 0) return P2M_E_INVALID;
    int rc = pass_begin(c, n);
    if (rc) return rc;
    rc = pass_push(c, src, on_device, n, index_base);
    if (rc) return rc;
    return pass_finish(c, allow_tune);
}

extern "C" {

int p2m_paint(p2m_ctx *c, const float *halos, int64_t n, int64_t index_base)
{
    return paint_common(c, halos, false, n, index_base, true);
}

int p2m_paint_device(p2m_ctx *c, const float *d_halos, int64_t n, int64_t index_base)
{
    return paint_common(c, d_halos, true, n, index_base, true);
}

int p2m_stream_begin(p2m_ctx *c, int64_t n_total) { return pass_begin(c, n_total); }

int p2m_stream_push(p2m_ctx *c, const float *halos, int64_t n, int64_t index_base)
{
    return pass_push(c, halos, false, n, index_base);
}

int p2m_stream_finish(p2m_ctx *c) { return pass_finish(c, true); }

int p2m_host_alloc(p2m_ctx *c, int64_t bytes, void **ptr)
{
    if (!c || !ptr || bytes <= 0) return P2M_E_INVALID;
    *ptr = nullptr;
    HIPCHK(hipSetDevice(c->cfg.device));
    if (hipHostMalloc(ptr, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return P2M_E_NOMEM;
    }
    return 0;
}

int p2m_host_free(p2m_ctx *c, void *ptr)
{
    if (!c) return P2M_E_INVALID;
    if (!ptr) return 0;
    HIPCHK(hipSetDevice(c->cfg.device));
    /* a block may still be read by a queued copy */
    HIPCHK(hipStreamSynchronize(c->h2d_stream));
    HIPCHK(hipHostFree(ptr));
    return 0;
}

int p2m_get_stats(p2m_ctx *c, p2m_stats *st)
{
    if (!c || !st) return P2M_E_INVALID;
    *st = c->st;
    return 0;
}

int p2m_fetch_maps(p2m_ctx *c, int slot, float *hp32, double *hp64, float *fs32, double *fs64)
{
    if (!c) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (slot < 0 || slot >= c->cfg.n_maps) return P2M_E_INVALID;
    HIPCHK(hipSetDevice(c->cfg.device));
    { int rc = ensure_hp_valid(c); if (rc) return rc; }
    if (c->hpnpix > 0) {
        if (c->d_hp64[slot]) {
            if (hp64) HIPCHK(hipMemcpy(hp64, c->d_hp64[slot], sizeof(double) * c->hpnpix, hipMemcpyDeviceToHost));
            if (hp32) {
                std::vector<double> tmp(c->hpnpix);
                HIPCHK(hipMemcpy(tmp.data(), c->d_hp64[slot], sizeof(double) * c->hpnpix, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < c->hpnpix; ++i) hp32[i] = (float)tmp[i];
            }
        } else {
            if (hp32) HIPCHK(hipMemcpy(hp32, c->d_hp32[slot], sizeof(float) * c->hpnpix, hipMemcpyDeviceToHost));
            if (hp64) {
                std::vector<float> tmp(c->hpnpix);
                HIPCHK(hipMemcpy(tmp.data(), c->d_hp32[slot], sizeof(float) * c->hpnpix, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < c->hpnpix; ++i) hp64[i] = (double)tmp[i];
            }
        }
    }
    if (c->fsnpix > 0 && (fs32 || fs64)) {
        std::vector<double> tmp(c->fsnpix);
        HIPCHK(hipMemcpy(tmp.data(), c->d_fs64[slot], sizeof(double) * c->fsnpix, hipMemcpyDeviceToHost));
        if (fs64) memcpy(fs64, tmp.data(), sizeof(double) * c->fsnpix);
        if (fs32) for (int64_t i = 0; i < c->fsnpix; ++i) fs32[i] = (float)tmp[i];
    }
    return 0;
}

int p2m_fetch_membership(p2m_ctx *c, uint32_t *hc, uint64_t *hh, uint32_t *fc, uint64_t *fh)
{
    if (!c) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (!(c->cfg.flags & P2M_FLAG_MEMBERSHIP)) return P2M_E_STATE;
    HIPCHK(hipSetDevice(c->cfg.device));
    { int rc = ensure_hp_valid(c); if (rc) return rc; }
    if (c->hpnpix > 0) {
        if (hc) HIPCHK(hipMemcpy(hc, c->d_count, sizeof(uint32_t) * c->hpnpix, hipMemcpyDeviceToHost));
        if (hh) HIPCHK(hipMemcpy(hh, c->d_hash, sizeof(uint64_t) * c->hpnpix, hipMemcpyDeviceToHost));
    }
    if (c->fsnpix > 0) {
        if (fc) HIPCHK(hipMemcpy(fc, c->d_fs_count, sizeof(uint32_t) * c->fsnpix, hipMemcpyDeviceToHost));
        if (fh) HIPCHK(hipMemcpy(fh, c->d_fs_hash, sizeof(uint64_t) * c->fsnpix, hipMemcpyDeviceToHost));
    }
    return 0;
}

}  /* extern "C" */

/* ------------------------------------------------------------------ autotuner */
/* Times the binning + tile paint for each candidate tile shape on the
 * halos already decoded, keeps the fastest, and restores the maps. */
static int autotune(p2m_ctx *c, int64_t n_halo, int64_t n_flat, unsigned long long n_items)
{
    static const int cand[][2] = {{9, 256}, {10, 256}, {11, 256}, {11, 128}, {12, 256}};
    int mode_ld = c->cfg.precision == P2M_PREC_FP64 ? 1 : 2;
    int best_l = 10, best_wg = 256;
    /* small problems: the default shape */
    if (n_halo < 200000 || getenv("P2M_NO_AUTOTUNE")) {
        int rc = build_tiles(c, best_l); if (rc) return rc;
        c->wg = best_wg; c->tuned = true;
        return 0;
    }
    /* tune on a copy of the maps' state: paint into the maps, then clear
     * (the call that triggers tuning starts from the maps' current state,
     * so remember whether they were zero and re-zero after each trial) */
    bool was_zero = c->hp_zero;
    if (!was_zero) {
        int rc = build_tiles(c, best_l); if (rc) return rc;
        c->wg = best_wg; c->tuned = true;
        return 0;
    }
    float best_ms = 1e30f;
    for (auto &cd : cand) {
        if (cd[0] > p2m_tile_max_log2w()) continue;
        if (p2m_tile_lds_bytes(mode_ld, c->cfg.n_maps, cd[0], c->d_count != nullptr, c->g.nrt) > 160 * 1024) continue;
        int rc = build_tiles(c, cd[0]); if (rc) return rc;
        c->wg = cd[1]; c->tuned = true;
        c->hp_zero = true;
        /* each trial starts from zero paint counters, as a call does: the
         * deferred-pixel and overflow lists are appended at their counters */
        p2m_launch_clear_ctr(c->d_ctr, Ctr::PAINT, Ctr::TOTAL - Ctr::PAINT, c->stream);
        rc = paint_halos(c, n_halo, n_flat, n_items, false);
        if (rc == P2M_E_INVALID) {
            /* a shape the painter rejects at launch: the trial's flat-sky
             * paint is still on its stream, so the clear waits for it */
            HIPCHK(hipStreamWaitEvent(c->stream, c->ev_fs, 0));
            (void)p2m_clear(c);
            continue;
        }
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(c->stream));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[4]);
        if (ms < best_ms) { best_ms = ms; best_l = cd[0]; best_wg = cd[1]; }
        rc = p2m_clear(c); if (rc) return rc;
        /* a trial's tiles that lost deferred pixels are not repainted: the
         * maps are cleared anyway, and so are the marks */
        p2m_launch_fill_u32((uint32_t *)c->d_lost, (c->n_tiles + 3) / 4, 0u, c->stream);
    }
    /* counters were polluted by the trials: reset the paint counters */
    p2m_launch_clear_ctr(c->d_ctr, Ctr::PAINT, Ctr::TOTAL - Ctr::PAINT, c->stream);
    p2m_launch_clear_ctr(c->d_ctr, Ctr::LOST, 1, c->stream);
    int rc = build_tiles(c, best_l); if (rc) return rc;
    c->wg = best_wg; c->tuned = true; c->hp_zero = true;
    return 0;
}

/* ------------------------------------------------------------------ products + writers */
extern "C" {

int p2m_products(p2m_ctx *c, int slot, double moments[6], double rms[2], float dpdy[10000])
{
    if (!c) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (slot < 0 || slot >= c->cfg.n_maps) return P2M_E_INVALID;
    HIPCHK(hipSetDevice(c->cfg.device));
    { int rc = ensure_hp_valid(c); if (rc) return rc; }
    double *d_out = nullptr;
    unsigned long long *d_hist = nullptr;
    if (dmalloc((void **)&d_out, sizeof(double) * 8)) return P2M_E_NOMEM;
    if (dmalloc((void **)&d_hist, sizeof(unsigned long long) * 10000)) { dfree(d_out); return P2M_E_NOMEM; }
    double res[2][4] = {{0}};
    const int ndpdy = 10000;
    const float dy = 1e-4f / (float)ndpdy;                 /* pks2map.f90:39-40 */
    int rc = 0;
    for (int which = 0; which < 2 && !rc; ++which) {
        int64_t n = which == 0 ? c->fsnpix : c->hpnpix;
        if (n <= 0) continue;
        const float *m32 = which == 1 ? c->d_hp32[slot] : nullptr;
        const double *m64 = which == 1 ? c->d_hp64[slot] : c->d_fs64[slot];
        if (m32) m64 = nullptr;
        p2m_launch_fill_f64(d_out, 8, c->stream);
        if (which == 1) p2m_launch_fill_u64(d_hist, ndpdy, c->stream);
        p2m_launch_moments(m32, m64, n, 0.0, d_out, which == 1 ? d_hist : nullptr, ndpdy, dy, c->stream);
        double h1[8];
        if (hipMemcpyAsync(h1, d_out, sizeof(h1), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { rc = P2M_E_HIP; break; }
        double mean = h1[0] / (double)n;
        p2m_launch_fill_f64(d_out, 8, c->stream);
        p2m_launch_moments(m32, m64, n, mean, d_out, nullptr, ndpdy, dy, c->stream);
        double h2[8];
        if (hipMemcpyAsync(h2, d_out, sizeof(h2), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { rc = P2M_E_HIP; break; }
        res[which][0] = h1[0] / (double)n; res[which][1] = h1[1] / (double)n; res[which][2] = h1[2] / (double)n;
        res[which][3] = p2m_sqrt(h2[3] / (double)n);
        if (which == 1 && dpdy) {
            std::vector<unsigned long long> hist(ndpdy);
            if (hipMemcpy(hist.data(), d_hist, sizeof(unsigned long long) * ndpdy, hipMemcpyDeviceToHost) != hipSuccess) {
                rc = P2M_E_HIP; break;
            }
            /* dpdy = dpdy / real(hpnpix); cumulative (pks2map.f90:347, 355-357) */
            for (int i = 0; i < ndpdy; ++i) dpdy[i] = (float)hist[i] / (float)n;
            for (int i = 1; i < ndpdy; ++i) dpdy[i] = dpdy[i - 1] + dpdy[i];
        }
    }
    dfree(d_out); dfree(d_hist);
    if (rc) return rc;
    if (moments) { for (int k = 0; k < 3; ++k) { moments[k] = res[0][k]; moments[3 + k] = res[1][k]; } }
    if (rms) { rms[0] = res[0][3]; rms[1] = res[1][3]; }
    return 0;
}

/* p2m_products without any allocation: the same kernels with the partial
 * sums, the histogram and the cumulative dP/dy in this device's scratch array.
 * Calls that overlap on one GPU share that array and race. */
int p2m_products_noalloc(p2m_ctx *c, int slot, double moments[6], double rms[2], float dpdy[10000])
{
    if (!c) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (slot < 0 || slot >= c->cfg.n_maps) return P2M_E_INVALID;
    HIPCHK(hipSetDevice(c->cfg.device));
    void *base = nullptr;
    { int rc = scratch_for_device(c->cfg.device, &base); if (rc) return rc; }
    double *d_out = (double *)base;
    unsigned long long *d_hist = (unsigned long long *)((char *)base + P2M_NOALLOC_HIST_OFF);
    float *d_dp = (float *)((char *)base + P2M_NOALLOC_DPDY_OFF);
    { int rc = ensure_hp_valid(c); if (rc) return rc; }
    double res[2][4] = {{0}};
    const int ndpdy = 10000;
    const float dy = 1e-4f / (float)ndpdy;                 /* pks2map.f90:39-40 */
    for (int which = 0; which < 2; ++which) {
        const int64_t n = which == 0 ? c->fsnpix : c->hpnpix;
        if (n <= 0) continue;
        const float *m32 = which == 1 ? c->d_hp32[slot] : nullptr;
        const double *m64 = which == 1 ? c->d_hp64[slot] : c->d_fs64[slot];
        if (m32) m64 = nullptr;
        /* the scratch keeps the last call's contents: cleared by fill kernels */
        p2m_launch_fill_f64(d_out, 8, c->stream);
        if (which == 1) p2m_launch_fill_u64(d_hist, ndpdy, c->stream);
        p2m_launch_moments(m32, m64, n, 0.0, d_out, which == 1 ? d_hist : nullptr, ndpdy, dy, c->stream);
        double h1[8];
        HIPCHK(hipMemcpyAsync(h1, d_out, sizeof(h1), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const double mean = h1[0] / (double)n;
        p2m_launch_fill_f64(d_out, 8, c->stream);
        p2m_launch_moments(m32, m64, n, mean, d_out, nullptr, ndpdy, dy, c->stream);
        double h2[8];
        HIPCHK(hipMemcpyAsync(h2, d_out, sizeof(h2), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        res[which][0] = h1[0] / (double)n; res[which][1] = h1[1] / (double)n; res[which][2] = h1[2] / (double)n;
        res[which][3] = p2m_sqrt(h2[3] / (double)n);
        if (which == 1 && dpdy) {
            p2m_launch_dpdy_cumulative(d_hist, ndpdy, (float)n, d_dp, c->stream);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(dpdy, d_dp, sizeof(float) * ndpdy, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    }
    if (moments) { for (int k = 0; k < 3; ++k) { moments[k] = res[0][k]; moments[3 + k] = res[1][k]; } }
    if (rms) { rms[0] = res[0][3]; rms[1] = res[1][3]; }
    return 0;
}

}  /* extern "C" */

/* FITS card helpers (FITS standard 4.0, sec. 4.2) */
static std::string card(const char *key, const std::string &val, const char *comment)
{
    char b[81];
    std::string s;
    if (val.empty()) {
        snprintf(b, sizeof(b), "%-80s", key);
        return std::string(b, 80);
    }
    snprintf(b, sizeof(b), "%-8s= %20s / %s", key, val.c_str(), comment ? comment : "");
    s = b;
    s.resize(80, ' ');
    return s;
}
static std::string sval(const char *v)
{
    std::string s = "'";
    s += v;
    while (s.size() < 9) s += ' ';
    s += "'";
    while (s.size() < 20) s += ' ';
    return s;
}
static std::string ival(long long v) { char b[32]; snprintf(b, sizeof(b), "%lld", v); return b; }

static bool write_block(FILE *f, const std::vector<std::string> &cards)
{
    std::string txt;
    for (auto &c : cards) txt += c;
    std::string end = "END";
    end.resize(80, ' ');
    txt += end;
    txt.resize(((txt.size() + 2879) / 2880) * 2880, ' ');
    return fwrite(txt.data(), 1, txt.size(), f) == txt.size();
}

/* HEALPix map as FITS BINTABLE (fits_s_inc.f90:787-960 via cfitsio in the
 * reference; the same HDU structure and keywords written directly).  The map
 * streams from the device in pieces of 2^24 pixels through two page-locked
 * buffers: the copy of piece k+1 runs while piece k is byte-swapped to
 * big-endian and written, as the reference writes its rows piece by piece
 * (fits_s_inc.f90:921-933).  An FP64 map is rounded to float on the device
 * first (the reference's map is real). */
static int write_fits_dev(p2m_ctx *c, int slot, const char *path)
{
    const int64_t npix = c->hpnpix;
    const int32_t nside = c->cfg.nside;
    int64_t repeat = 1024;
    if (npix < 1024) repeat = 1;
    if (nside > 1024 * 256) repeat = nside / 256;
    int64_t nrows = npix / repeat;
    char date[32];
    time_t t = time(nullptr);
    struct tm tmv;
    gmtime_r(&t, &tmv);
    strftime(date, sizeof(date), "%Y-%m-%d", &tmv);
    std::vector<std::string> prim = {
        card("SIMPLE", "T", "file does conform to FITS standard"),
        card("BITPIX", ival(32), "number of bits per data pixel"),
        card("NAXIS", ival(0), "number of data axes"),
        card("EXTEND", "T", "FITS dataset may contain extensions"),
        card("DATE", sval(date), "file creation date (YYYY-MM-DD)")};
    char tform[32];
    snprintf(tform, sizeof(tform), "%lldE", (long long)repeat);
    std::vector<std::string> ext = {
        card("XTENSION", sval("BINTABLE"), "binary table extension"),
        card("BITPIX", ival(8), "8-bit bytes"),
        card("NAXIS", ival(2), "2-dimensional binary table"),
        card("NAXIS1", ival(4 * repeat), "width of table in bytes"),
        card("NAXIS2", ival(nrows), "number of rows in table"),
        card("PCOUNT", ival(0), "size of special data area"),
        card("GCOUNT", ival(1), "one data group (required keyword)"),
        card("TFIELDS", ival(1), "number of fields in each row"),
        card("TTYPE1", sval("simulation"), "label for field   1"),
        card("TFORM1", sval(tform), "data format of field: 4-byte REAL"),
        card("NSIDE", ival(nside), "the nside of the map"),
        card("ORDERING", sval("RING"), "the nside of the map")};
    const int64_t piece = std::min<int64_t>(npix, 1 << 24);
    float *hb[2] = {nullptr, nullptr};
    float *d32 = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    FILE *f = nullptr;
    int rc = 0;
    if (hipHostMalloc((void **)&hb[0], sizeof(float) * piece, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&hb[1], sizeof(float) * piece, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        rc = P2M_E_NOMEM;
    }
    if (!rc && c->d_hp64[slot] && dmalloc((void **)&d32, sizeof(float) * 2 * piece)) rc = P2M_E_NOMEM;
    if (!rc && (hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess)) {
        (void)hipGetLastError();
        rc = P2M_E_HIP;
    }
    if (!rc) {
        f = fopen(path, "wb");
        if (!f) rc = P2M_E_IO;
    }
    bool ok = !rc && write_block(f, prim) && write_block(f, ext);
    if (!rc && !ok) rc = P2M_E_IO;
    const int64_t npieces = (npix + piece - 1) / piece;
    auto enqueue = [&](int64_t k) -> int {
        const int64_t i0 = k * piece, nk = std::min(piece, npix - i0);
        const int bslot = (int)(k & 1);
        const float *src = c->d_hp32[slot] + (c->d_hp32[slot] ? i0 : 0);
        if (c->d_hp64[slot]) {
            p2m_launch_f64_to_f32(c->d_hp64[slot] + i0, d32 + bslot * piece, nk, c->stream);
            src = d32 + bslot * piece;
        }
        if (hipMemcpyAsync(hb[bslot], src, sizeof(float) * nk, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipEventRecord(ev[bslot], c->stream) != hipSuccess) {
            (void)hipGetLastError();
            return P2M_E_HIP;
        }
        return 0;
    };
    if (!rc && npieces > 0) rc = enqueue(0);
    for (int64_t k = 0; !rc && k < npieces; ++k) {
        if (k + 1 < npieces) rc = enqueue(k + 1);
        if (rc) break;
        const int bslot = (int)(k & 1);
        if (hipEventSynchronize(ev[bslot]) != hipSuccess) { (void)hipGetLastError(); rc = P2M_E_HIP; break; }
        const int64_t nk = std::min(piece, npix - k * piece);
        uint32_t *u = reinterpret_cast<uint32_t *>(hb[bslot]);
        for (int64_t j = 0; j < nk; ++j) u[j] = __builtin_bswap32(u[j]);
        if (fwrite(u, 4, (size_t)nk, f) != (size_t)nk) rc = P2M_E_IO;
        /* the buffer is refilled by the copy enqueued two pieces later, after
         * this write has returned */
    }
    if (!rc) {
        int64_t bytes = npix * 4, pad = (2880 - bytes % 2880) % 2880;
        std::vector<char> z(pad, 0);
        if (pad && fwrite(z.data(), 1, pad, f) != (size_t)pad) rc = P2M_E_IO;
    }
    (void)hipStreamSynchronize(c->stream);
    if (f && fclose(f) != 0 && !rc) rc = P2M_E_IO;
    if (rc && f) remove(path);
    for (int i = 0; i < 2; ++i) {
        if (hb[i]) (void)hipHostFree(hb[i]);
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    dfree(d32);
    return rc;
}

extern "C" int p2m_write_outputs(p2m_ctx *c, int slot, const char *base, int write_fs, int write_hp, int write_dp)
{
    if (!c || !base) return P2M_E_INVALID;
    if (c->in_stream) return P2M_E_STATE;
    if (slot < 0 || slot >= c->cfg.n_maps) return P2M_E_INVALID;
    std::string b(base);
    int rc = 0;
    if (write_dp && c->hpnpix > 0) {
        std::vector<float> dpdy(10000);
        rc = p2m_products(c, slot, nullptr, nullptr, dpdy.data());
        if (rc) return rc;
        std::string p = b + "_dp.bin";
        FILE *f = fopen(p.c_str(), "wb");
        if (!f) return P2M_E_IO;
        int32_t nd = 10000;
        float mx = 1e-4f;
        bool ok = fwrite(&nd, 4, 1, f) == 1 && fwrite(&mx, 4, 1, f) == 1 && fwrite(dpdy.data(), 4, 10000, f) == 10000;
        if (fclose(f) != 0) ok = false;
        if (!ok) { remove(p.c_str()); return P2M_E_IO; }
    }
    if (write_fs && c->fsnpix > 0) {
        std::vector<float> m(c->fsnpix);
        rc = p2m_fetch_maps(c, slot, nullptr, nullptr, m.data(), nullptr);
        if (rc) return rc;
        std::string p = b + "_fs.map";
        FILE *f = fopen(p.c_str(), "wb");
        if (!f) return P2M_E_IO;
        int32_t np = c->cfg.npix_fs;
        float fov = (float)(((c->cfg.fov_deg / 360.0) * 2.0) * c->mypi);
        bool ok = fwrite(&np, 4, 1, f) == 1 && fwrite(&np, 4, 1, f) == 1 && fwrite(&fov, 4, 1, f) == 1 &&
                  fwrite(&fov, 4, 1, f) == 1 && fwrite(m.data(), 4, (size_t)c->fsnpix, f) == (size_t)c->fsnpix;
        if (fclose(f) != 0) ok = false;
        if (!ok) { remove(p.c_str()); return P2M_E_IO; }
    }
    if (write_hp && c->hpnpix > 0) {
        HIPCHK(hipSetDevice(c->cfg.device));
        rc = ensure_hp_valid(c);
        if (rc) return rc;
        std::string p = b + "_hp.fits";
        rc = write_fits_dev(c, slot, p.c_str());
        if (rc) return rc;
    }
    return 0;
}

extern "C" int p2m_selftest_math(int32_t device, int32_t op, const double *x, const double *y, double *out, int64_t n)
{
    if (!x || !out || n < 0 || op < 0 || op > 9) return P2M_E_INVALID;
    if ((op == 6 || op == 9) && !y) return P2M_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { (void)hipGetLastError(); return P2M_E_NODEVICE; }
    if (device < 0 || device >= ndev) return P2M_E_INVALID;
    HIPCHK(hipSetDevice(device));
    if (n == 0) return 0;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    size_t b = sizeof(double) * (size_t)n;
    int rc = dmalloc((void **)&dx, b);
    if (!rc && y) rc = dmalloc((void **)&dy, b);
    if (!rc) rc = dmalloc((void **)&dout, b);
    if (!rc && hipMemcpy(dx, x, b, hipMemcpyHostToDevice) != hipSuccess) rc = P2M_E_HIP;
    if (!rc && y && hipMemcpy(dy, y, b, hipMemcpyHostToDevice) != hipSuccess) rc = P2M_E_HIP;
    if (!rc && p2m_launch_selftest_math(op, dx, dy, dout, n, 0)) rc = P2M_E_HIP;
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = P2M_E_HIP;
    if (!rc && hipMemcpy(out, dout, b, hipMemcpyDeviceToHost) != hipSuccess) rc = P2M_E_HIP;
    dfree(dx); dfree(dy); dfree(dout);
    (void)hipGetLastError();
    return rc;
}

extern "C" int p2m_selftest_copy(int32_t device, int64_t n_bytes, int32_t reps, double *ms)
{
    if (n_bytes <= 0 || n_bytes % 16 || n_bytes > (1LL << 34) || reps <= 0 || !ms) return P2M_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { (void)hipGetLastError(); return P2M_E_NODEVICE; }
    if (device < 0 || device >= ndev) return P2M_E_INVALID;
    HIPCHK(hipSetDevice(device));
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = dmalloc(&a, (size_t)n_bytes);
    if (!rc) rc = dmalloc(&b, (size_t)n_bytes);
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = P2M_E_HIP;
    if (!rc) {
        p2m_launch_fill_f32((float *)a, n_bytes / 4, 0);
        if (p2m_launch_copy_calib(a, b, n_bytes / 16, 0)) rc = P2M_E_HIP;      /* warm-up */
    }
    if (!rc && hipEventRecord(e0, 0) != hipSuccess) rc = P2M_E_HIP;
    for (int r = 0; r < reps && !rc; ++r)
        if (p2m_launch_copy_calib(a, b, n_bytes / 16, 0)) rc = P2M_E_HIP;
    if (!rc && hipEventRecord(e1, 0) != hipSuccess) rc = P2M_E_HIP;
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = P2M_E_HIP;
    float t = 0.f;
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = P2M_E_HIP;
    if (!rc) *ms = (double)t / reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dfree(a); dfree(b);
    (void)hipGetLastError();
    return rc;
}
