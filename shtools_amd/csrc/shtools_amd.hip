// C ABI of libshtools_amd.so (declared in include/shtools_amd.h): argument
// checks with the reference's exit codes, cached device tables and grid
// geometry, and the kernel launch sequences of the plain transform path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <functional>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/shtools_amd.h"
#include "fftplan.hpp"
#include "checks.hpp"
#include "kernels.hpp"
#include "tables.hpp"

namespace shtg {

// Elements of a workspace FFT's LDS chunk of the largest LDS footprint.
constexpr int kChunkCap = (int)(kMaxLds / sizeof(double2));
static_assert(sizeof(Item4) == sizeof(int4) && sizeof(Item2) == sizeof(int2), "work items as int4 / int2");

// Locking: g_mu guards the table caches (recursion, seeds, geometries,
// gravity rows, Gauss-Legendre nodes) while a call looks its tables up or
// builds them; the launches run outside it, on copies of the table
// descriptors (a table that grows is replaced, and the outgrown arrays stay
// allocated).  g_ws_mu guards the per-device intermediates of the plain
// launchers, g_prof_mu the profiling records, and each batched plan has its
// own mutex for its bindings and workspace.  Host threads calling on
// different streams therefore issue their launches concurrently.
static std::mutex g_mu;
static std::mutex g_ws_mu;
static std::mutex g_prof_mu;

#define HIP_OK(x) (hipSuccess == (x))

// The ring-FFT kernel instantiation for a plan variant (kVarPlain,
// kVarSplit, kVarPrime, kVarPow2).
template <class K>
static K by_variant(int v, K plain, K split, K prime, K pow2)
{
    return v == kVarSplit ? split : v == kVarPrime ? prime : v == kVarPow2 ? pow2 : plain;
}
#define SHTG_FFT_KERNEL(k, v) by_variant((v), k<kVarPlain>, k<kVarSplit>, k<kVarPrime>, k<kVarPow2>)


template <class T>
static T *dev_upload(const std::vector<T> &v)
{
    T *p = nullptr;
    if (v.empty()) return nullptr;
    if (!HIP_OK(hipMalloc(&p, v.size() * sizeof(T)))) return nullptr;
    if (!HIP_OK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice))) {
        (void)hipFree(p);
        return nullptr;
    }
    return p;
}

// ------------------------------------------------------------ caches

struct DevRecurrence {
    int L = -1;
    double *alpha = nullptr, *s = nullptr, *post = nullptr, *post_c = nullptr;
};
struct DevSeeds {
    int L = -1;
    double *seed = nullptr;
};

// Grid geometry: ring-pair slots and the ring FFT plan.
struct Geometry {
    int P = 0;          // ring-pair slots
    int nlat = 0;       // rows of the (non-extended) grid
    int nlong = 0;      // FFT length
    int s0 = 0;         // first slot with analysis weight (DH: 1, the pole carries none)
    double *z = nullptr, *u = nullptr;
    double2 *w = nullptr;  // analysis weights (north, south)
    double2 *ones = nullptr;  // unit weights (adjoint of synthesis)
    int2 *rows = nullptr;
    int *kmax = nullptr;   // per slot: the orders with non-negligible functions (polar skip)
    int *smin = nullptr;   // per order: the first such slot
    FftDev fft{};
    bool ok = false;
};

static std::map<std::tuple<int, int>, DevRecurrence> g_rec;            // (dev, norm)
static std::map<std::tuple<int, int, int>, DevSeeds> g_seeds;          // (dev, norm, csphase)
static std::map<std::tuple<int, int, int>, std::unique_ptr<Geometry>> g_dh;      // (dev, L, sampling)
static std::map<std::tuple<int, int, uint64_t>, std::unique_ptr<Geometry>> g_glq;  // (dev, L, hash)

static bool kernels_configured = false;

static void configure_kernels()
{
    if (kernels_configured) return;
    for (const void *k :
         {(const void *)k_synth_fft<kVarPlain>, (const void *)k_synth_fft<kVarSplit>,
          (const void *)k_synth_fft<kVarPrime>, (const void *)k_anal_fft<kVarPlain>,
          (const void *)k_anal_fft<kVarSplit>, (const void *)k_anal_fft<kVarPrime>, (const void *)k_synth_fft_ws,
          (const void *)k_anal_fft_ws, (const void *)kb_synth_fft<kVarPlain>, (const void *)kb_synth_fft<kVarSplit>,
          (const void *)kb_synth_fft<kVarPrime>, (const void *)kb_anal_fft<kVarPlain>,
          (const void *)kb_anal_fft<kVarSplit>, (const void *)kb_anal_fft<kVarPrime>, (const void *)kb_synth_fft_ws,
          (const void *)kb_anal_fft_ws, (const void *)k_synth_fft_c<false, kVarPlain>,
          (const void *)k_synth_fft_c<false, kVarSplit>, (const void *)k_synth_fft_c<false, kVarPrime>,
          (const void *)k_synth_fft_c<true, kVarPlain>, (const void *)k_anal_fft_c<false, kVarPlain>,
          (const void *)k_anal_fft_c<false, kVarSplit>, (const void *)k_anal_fft_c<false, kVarPrime>,
          (const void *)k_anal_fft_c<true, kVarPlain>, (const void *)k_synth_fft<kVarPow2>,
          (const void *)k_anal_fft<kVarPow2>, (const void *)kb_synth_fft<kVarPow2>, (const void *)kb_anal_fft<kVarPow2>,
          (const void *)k_synth_fft_c<false, kVarPow2>, (const void *)k_anal_fft_c<false, kVarPow2>})
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
    kernels_configured = true;
}

// Tables outgrown by a larger degree stay allocated until the process ends:
// work already enqueued and batch plans may still read them.  Degrees grow
// geometrically, which bounds what is kept.
static std::vector<void *> g_retired;

static int grown_degree(int L, int old)
{
    int Lb = std::max(L, 64);
    if (old >= 0) Lb = std::max(Lb, old + old / 2);
    return Lb;
}

static const DevRecurrence *get_recurrence(int dev, int norm, int L)
{
    auto &r = g_rec[{dev, norm}];
    if (r.L >= L) return &r;
    const int Lb = grown_degree(L, r.L);
    if (r.L >= 0) {
        g_retired.push_back(r.alpha);
        g_retired.push_back(r.s);
        g_retired.push_back(r.post);
        g_retired.push_back(r.post_c);
        r = DevRecurrence{};
    }
    RecurrenceTables t = build_recurrence(norm, Lb);
    r.alpha = dev_upload(t.alpha);
    r.s = dev_upload(t.s);
    r.post = dev_upload(t.post);
    r.post_c = dev_upload(t.post_c);
    if (!r.alpha || !r.s || !r.post || !r.post_c) {
        // nothing has been enqueued on these yet; a later call retries
        for (void *q : {(void *)r.alpha, (void *)r.s, (void *)r.post, (void *)r.post_c})
            if (q) (void)hipFree(q);
        r = DevRecurrence{};
        return nullptr;
    }
    r.L = Lb;
    return &r;
}

static const DevSeeds *get_seeds(int dev, int norm, int csphase, int L, bool cplx = false)
{
    auto &r = g_seeds[{dev, norm + (cplx ? 8 : 0), csphase}];
    if (r.L >= L) return &r;
    const int Lb = grown_degree(L, r.L);
    if (r.L >= 0) { g_retired.push_back(r.seed); r = DevSeeds{}; }
    r.seed = dev_upload(build_seeds(norm, csphase, Lb, cplx));
    if (!r.seed) { r = DevSeeds{}; return nullptr; }
    r.L = Lb;
    return &r;
}

static bool upload_fft_arrays(const FftPlanHost &h, FftDev &d);

// Device arrays of an FFT plan that no work has been enqueued on.
static void free_fft(FftDev &d)
{
    for (const void *q : {(const void *)d.tw, (const void *)d.twst, (const void *)d.perm, (const void *)d.chirp,
                          (const void *)d.bhat, (const void *)d.bhat3, (const void *)d.bhat4})
        if (q) (void)hipFree(const_cast<void *>(q));
    if (d.bhat2 && d.bhat2 != d.bhat) (void)hipFree(const_cast<double2 *>(d.bhat2));
    std::memset(&d, 0, sizeof(d));
}

// Plan upload; on failure nothing stays allocated and d is zeroed.
static bool upload_fft(const FftPlanHost &h, FftDev &d)
{
    if (upload_fft_arrays(h, d)) return true;
    free_fft(d);
    return false;
}

static bool upload_fft_arrays(const FftPlanHost &h, FftDev &d)
{
    std::memset(&d, 0, sizeof(d));
    d.n = h.n;
    d.m = h.m;
    d.nst = h.nst;
    d.blue = h.blue;
    d.lc = h.lc;
    d.half = h.half;
    for (int i = 0; i < h.nst; ++i) {
        d.radix[i] = h.radix[i];
        d.stride[i] = h.stride[i];
        d.twoff[i] = h.twoff[i];
    }
    d.tw = reinterpret_cast<const double2 *>(dev_upload(h.tw));
    d.twst = reinterpret_cast<const double2 *>(dev_upload(h.twst));
    if (!d.twst) return false;
    d.perm = dev_upload(h.perm);
    if (!d.tw || !d.perm) return false;
    if (h.blue) {
        d.chirp = reinterpret_cast<const double2 *>(dev_upload(h.chirp));
        d.bhat = reinterpret_cast<const double2 *>(dev_upload(h.bhat));
        d.bhat2 = h.bhat2.empty() ? d.bhat : reinterpret_cast<const double2 *>(dev_upload(h.bhat2));
        if (!d.chirp || !d.bhat || !d.bhat2) return false;
        if (h.half) {
            d.bhat3 = reinterpret_cast<const double2 *>(dev_upload(h.bhat3));
            d.bhat4 = reinterpret_cast<const double2 *>(dev_upload(h.bhat4));
            if (!d.bhat3 || !d.bhat4) return false;
        }
    }
    return true;
}

// Device arrays of a geometry whose construction failed (nothing has been
// enqueued on them yet).
static void free_geometry(Geometry &g)
{
    for (const void *q : {(const void *)g.z, (const void *)g.u, (const void *)g.w, (const void *)g.rows,
                          (const void *)g.ones, (const void *)g.kmax, (const void *)g.smin})
        if (q) (void)hipFree(const_cast<void *>(q));
    free_fft(g.fft);  // zeroed if upload_fft failed or never ran
    g = Geometry{};
}

// lc: the FFT plan's pruning degree (-1: none); L: the grid's degree.
static bool finish_geometry(Geometry &g, const std::vector<double> &z, const std::vector<double> &u,
                            const std::vector<double> &w, const std::vector<int> &rows, int lc, int L)
{
    g.z = dev_upload(z);
    g.u = dev_upload(u);
    g.w = reinterpret_cast<double2 *>(dev_upload(w));
    g.rows = reinterpret_cast<int2 *>(dev_upload(rows));
    g.ones = reinterpret_cast<double2 *>(dev_upload(std::vector<double>(w.size(), 1.0)));
    const std::vector<int> km = polar_kmax(z, u, L);
    g.kmax = dev_upload(km);
    g.smin = dev_upload(polar_smin(km, L));
    if (!g.z || !g.u || !g.w || !g.rows || !g.ones || !g.kmax || !g.smin) return false;
    FftPlanHost fp = lc >= 0 ? ring_plan(g.nlong, lc) : make_fft_plan(g.nlong, lc);
    if (!upload_fft(fp, g.fft)) return false;
    g.ok = true;
    return true;
}

// DH geometry (src/MakeGridDH.f95:504-510, src/SHExpandDH.f95:287-294,463-479):
// slot s = 0 .. n/2 pairs row s (theta = pi s / n) with row n - s; slot 0 is
// the north pole with the south pole (row n, present only with extend),
// slot n/2 is the equator (unpaired).
static Geometry *get_dh(int dev, int L, int sampling)
{
    auto &slot = g_dh[{dev, L, sampling}];
    if (slot) return slot.get();
    slot.reset(new Geometry());
    Geometry &g = *slot;
    const int n = 2 * L + 2;
    const double pi = std::acos(-1.0);
    g.nlat = n;
    g.nlong = sampling * n;
    g.P = n / 2 + 1;
    g.s0 = 1;
    std::vector<double> z(g.P), u(g.P), w(2 * g.P);
    std::vector<int> rows(2 * g.P);
    std::vector<double> aj;
    dh_weights(n, aj);
    const double c = std::sqrt(2 * pi) * std::sqrt(4.0 * pi) / (double)g.nlong;
    for (int s = 0; s < g.P; ++s) {
        double zz;
        if (s == n / 2) zz = 0.0;
        else zz = std::cos(pi * (double)s / (double)n);
        z[s] = zz;
        u[s] = (s == n / 2) ? 1.0 : std::sqrt((1.0 - zz) * (1.0 + zz));
        rows[2 * s] = s;
        rows[2 * s + 1] = (s == n / 2) ? -1 : n - s;
        w[2 * s] = c * aj[s];
        w[2 * s + 1] = (s == 0 || s == n / 2) ? 0.0 : c * aj[n - s];
    }
    if (!finish_geometry(g, z, u, w, rows, L, L)) {
        free_geometry(g);
        slot.reset();  // a later call retries
        return nullptr;
    }
    return &g;
}

static uint64_t fnv1a(const void *p, size_t nbytes, uint64_t h = 1469598103934665603ULL)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < nbytes; ++i) { h ^= b[i]; h *= 1099511628211ULL; }
    return h;
}

// GLQ geometry (src/MakeGridGLQ.f95:467-480,546-553): slot s pairs row s
// (z = zero(s)) with row nlat-1-s; for odd nlat the middle row is an exact
// equator (z = 0).  Weights 2 pi w / nlong (src/SHExpandGLQ.f95:535-547).
static uint64_t glq_key(int L, const double *zero, const double *wq)
{
    const int nlat = L + 1;
    const uint64_t h = fnv1a(zero, sizeof(double) * nlat);
    return wq ? fnv1a(wq, sizeof(double) * nlat, h) : fnv1a("nw", 2, h);
}

static Geometry *get_glq(int dev, int L, const double *zero, const double *wq)
{
    const int nlat = L + 1;
    auto &slot = g_glq[{dev, L, glq_key(L, zero, wq)}];
    if (slot) return slot.get();
    slot.reset(new Geometry());
    Geometry &g = *slot;
    const double pi = std::acos(-1.0);
    g.nlat = nlat;
    g.nlong = 2 * L + 1;
    g.P = (nlat + 1) / 2;
    std::vector<double> z(g.P), u(g.P), w(2 * g.P);
    std::vector<int> rows(2 * g.P);
    for (int s = 0; s < g.P; ++s) {
        const bool eq = (nlat % 2 != 0) && (s == g.P - 1);
        const double zz = eq ? 0.0 : zero[s];
        z[s] = zz;
        u[s] = eq ? 1.0 : std::sqrt((1.0 - zz) * (1.0 + zz));
        rows[2 * s] = s;
        rows[2 * s + 1] = eq ? -1 : nlat - 1 - s;
        w[2 * s] = wq ? 2.0 * pi * wq[s] / (double)g.nlong : 0.0;
        w[2 * s + 1] = (wq && !eq) ? 2.0 * pi * wq[nlat - 1 - s] / (double)g.nlong : 0.0;
    }
    if (!finish_geometry(g, z, u, w, rows, -1, L)) {
        free_geometry(g);
        slot.reset();  // a later call retries
        return nullptr;
    }
    return &g;
}

// The plain transforms take their intermediates from the device's default
// stream-ordered pool (hipMallocAsync).  Its release threshold is raised
// once per device so that freed blocks stay in the pool across calls
// instead of being unmapped at every synchronisation and mapped again by the
// next call; the pool keeps at most kPoolKeep bytes that way.
constexpr uint64_t kPoolKeep = 8ull << 30;

// Intermediates of one call from the pool, in power-of-two sizes: calls of
// other degrees then reuse the blocks earlier calls freed instead of adding
// blocks of new sizes to the pool.
static hipError_t pool_alloc(void **p, size_t bytes, hipStream_t st)
{
    size_t b = 4096;
    while (b < bytes) b <<= 1;
    return hipMallocAsync(p, b, st);
}

static bool capturing(hipStream_t st);
static int current_device();

// Intermediates of the plain launchers: one buffer per device, reused by
// consecutive calls; a call that finds it held by another host thread's
// call takes its own block from the pool.  Every call first makes its
// stream wait, on the device, for
// the previous user's work (an event recorded at release; on the same
// stream the wait is already met, and a stream handle reused after a
// destroy is covered too).  A larger request replaces the buffer by one of the
// next power-of-two size, the old one freed in stream order after its last
// use.  Under graph capture the call takes its own block from the pool
// instead, since a later growth would free memory the graph still uses.
// (Allocating per call from the pool cost 0.1-0.5 ms of host time per call in
// back-to-back calls.)
struct WsCache {
    void *buf = nullptr;
    size_t size = 0;
    hipEvent_t ev = nullptr;
    bool used = false;    // ev marks the end of the last user's work
    bool leased = false;  // held by a call between ws_acquire and ws_release
};
static std::map<int, WsCache> g_ws;  // under g_ws_mu

struct WsLease {
    void *p = nullptr;
    bool pooled = false;
    int dev = -1;
};

static int ws_acquire(WsLease &l, size_t bytes, hipStream_t st)
{
    if (capturing(st)) {
        l.pooled = true;
        return HIP_OK(pool_alloc(&l.p, bytes, st)) ? ST_OK : ST_ALLOC;
    }
    const int dev = current_device();
    if (dev < 0) return ST_OTHER;
    std::lock_guard<std::mutex> lk(g_ws_mu);
    WsCache &c = g_ws[dev];
    if (c.leased) {
        // another host thread's call holds the buffer: this one takes a
        // block of its own from the pool
        l.pooled = true;
        return HIP_OK(pool_alloc(&l.p, bytes, st)) ? ST_OK : ST_ALLOC;
    }
    if (!c.ev && !HIP_OK(hipEventCreateWithFlags(&c.ev, hipEventDisableTiming))) return ST_OTHER;
    if (c.used && !HIP_OK(hipStreamWaitEvent(st, c.ev, 0))) return ST_OTHER;
    if (bytes > c.size) {
        size_t b = 1 << 20;
        while (b < bytes) b <<= 1;
        void *nb = nullptr;
        if (!HIP_OK(hipMallocAsync(&nb, b, st))) return ST_ALLOC;
        if (c.buf) (void)hipFreeAsync(c.buf, st);
        c.buf = nb;
        c.size = b;
    }
    c.leased = true;
    l.p = c.buf;
    l.pooled = false;
    l.dev = dev;
    return ST_OK;
}

static void ws_release(const WsLease &l, hipStream_t st)
{
    if (l.pooled) {
        (void)hipFreeAsync(l.p, st);
        return;
    }
    std::lock_guard<std::mutex> lk(g_ws_mu);
    WsCache &c = g_ws[l.dev];
    (void)hipEventRecord(c.ev, st);
    c.used = true;
    c.leased = false;
}

static int current_device()
{
    static bool pool_done[64] = {};
    int d = 0;
    if (!HIP_OK(hipGetDevice(&d))) return -1;
    if (d >= 0 && d < 64 && !pool_done[d]) {
        pool_done[d] = true;
        hipMemPool_t pool = nullptr;
        uint64_t keep = kPoolKeep;
        if (HIP_OK(hipDeviceGetDefaultMemPool(&pool, d)))
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    return d;
}

// Waves a plain transform's Legendre launch aims for (8 per SIMD), and the
// fewest ring pairs per lane the plain kernels go down to for it.
constexpr int kLegWaves = 8192;
constexpr int kPlainSynthKMin = 1;
constexpr int kPlainAnalKMin = 4;

// Ring pairs per lane of a plain synthesis: kSynthKMax, or fewer when the
// orders times the blocks of ring pairs would give the launch fewer than
// kLegWaves waves; but 2 rather than 1 while 2 still gives kLegWaves / 4
// waves, since the staged tile (kb < kSynthKMax) runs 2 ring pairs per lane
// faster than twice the waves of 1 (GLQ lmax 1023: 54 vs 74 us, DH2 lmax
// 511: 22.5 vs 25.4 us).
static int synth_lanes(int Lc, int P)
{
    auto waves = [&](int kb) { return (long long)(Lc + 1) * ((P + kb * kWave - 1) / (kb * kWave)); };
    for (int kb = kSynthKMax; kb > kPlainSynthKMin; --kb)
        if (waves(kb) >= kLegWaves) return kb;
    return waves(2) >= kLegWaves / 4 ? 2 : kPlainSynthKMin;
}

// Largest ring pairs per lane of a plain analysis over nslots slots: fewer
// (down to kPlainAnalKMin) when the orders alone give fewer than kLegWaves
// waves, so that the slots split into more blocks.
static int plain_anal_kmax(int Lc, int nslots)
{
    const int want = (kLegWaves + Lc) / (Lc + 1);  // blocks per order wanted
    const int k = (nslots + want * kWave - 1) / (want * kWave);
    return std::max(kPlainAnalKMin, std::min(kAnalKMax, k));
}

// Slot blocks per order of a plain analysis whose waves take `passes`
// passes of K*64 ring pairs: one (a wave per order) for a single pass or
// when the orders alone give the GPU enough waves, else up to one block per
// pass, as many as bring the launch to about kLegWaves waves.
static int anal_blocks(int Lc, int passes)
{
    if (passes <= 1 || Lc + 1 >= kLegWaves) return 1;
    return std::max(1, std::min(passes, (kLegWaves + Lc) / (Lc + 1)));
}

// workgroups of a plain Legendre launch over orders 0..Lc in nb blocks each
// (order_block's mapping)
static unsigned leg_grid(int Lc, int nb) { return 8u * (unsigned)nb * (unsigned)((Lc + 8) / 8); }

// The plain synthesis kernel for kb ring pairs per lane: fewer than the most
// mean a launch of few waves, bounded by its longest waves' reads of their
// rows, which then come through LDS (DH2 lmax 511 synthesis Legendre 44 ->
// 27 us); with the most, the launch is bound by arithmetic and the LDS path
// costs registers and occupancy (DH1 lmax 2047 513 -> 564 us).
typedef decltype(&k_synth_leg<false>) SynthLegKernel;
static SynthLegKernel synth_leg_kernel(int kb) { return kb < kSynthKMax ? k_synth_leg<true> : k_synth_leg<false>; }

// the analysis kernels by ring pairs per lane
typedef decltype(&k_anal_leg<1>) AnalLegKernel;
typedef decltype(&kb_anal_leg<1>) BatchAnalLegKernel;
static AnalLegKernel anal_leg_kernel(int K)
{
    static const AnalLegKernel k[kAnalKMax] = {k_anal_leg<1>, k_anal_leg<2>, k_anal_leg<3>, k_anal_leg<4>,
                                               k_anal_leg<5>, k_anal_leg<6>, k_anal_leg<7>, k_anal_leg<8>};
    return k[std::min(std::max(K, 1), kAnalKMax) - 1];
}
static BatchAnalLegKernel batch_anal_leg_kernel(int K)
{
    static const BatchAnalLegKernel k[kBatchAnalKMax] = {
        kb_anal_leg<1>, kb_anal_leg<2>, kb_anal_leg<3>, kb_anal_leg<4>,  kb_anal_leg<5>,  kb_anal_leg<6>,
        kb_anal_leg<7>, kb_anal_leg<8>, kb_anal_leg<9>, kb_anal_leg<10>, kb_anal_leg<11>, kb_anal_leg<12>};
    return k[std::min(std::max(K, 1), kBatchAnalKMax) - 1];
}

static int grid_blocks(long long work)
{
    long long b = (work + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    if (b > 65535) b = 65535;
    return (int)b;
}

// Zero what lies outside the leading r0 x c0 block of a rows x cols array.
static void zero_outside(double *a, int rows, int cols, long long ld, int r0, int c0, hipStream_t st)
{
    if (r0 >= rows && c0 >= cols) return;
    hipLaunchKernelGGL(k_zero_outside, dim3(grid_blocks((long long)rows * cols)), dim3(kThreads), 0, st, a, rows,
                       cols, ld, r0, c0);
}

// ------------------------------------------------------------ profiling

// When enabled, every kernel launch is bracketed by HIP events recorded on
// the launch stream; shtg_profile_collect() reduces them per kernel id.
static std::atomic<bool> g_prof{false};
struct ProfRec {
    int kid;
    hipEvent_t a, b;
};
static std::vector<ProfRec> g_prof_recs;

struct ProfScope {
    int kid;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(int k, hipStream_t s) : kid(k), st(s)
    {
        if (!g_prof) return;
        if (!HIP_OK(hipEventCreate(&a)) || !HIP_OK(hipEventCreate(&b))) { a = b = nullptr; return; }
        (void)hipEventRecord(a, st);
    }
    void end()
    {
        if (!g_prof || !a) return;
        (void)hipEventRecord(b, st);
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_prof_recs.push_back({kid, a, b});
        a = b = nullptr;
    }
    ~ProfScope() { end(); }
};

// ------------------------------------------------------------ launchers

// Synthesis of rows [0, nrows_out) (+ the repeated column with ext_col).
// ctab scales the coefficients before the Legendre sums (rt.s; rt.post for
// the adjoint of analysis) and roww the output rows (nullptr: 1).
// ring: per-slot (z, u, rho^2) of a radially continued synthesis, z and u
// multiplied by rho (nullptr: the geometry's rings, no continuation).
struct RhoRings {
    const double *z = nullptr, *u = nullptr, *r2 = nullptr;
};

// Intermediates of a synthesis supplied by the caller (the no-allocation
// entry point's scratch) instead of the per-device buffer.
struct ExtWorkspace {
    void *p = nullptr;
    size_t cap = 0;
};

// Bytes of a synthesis' intermediates: packed coefficients, per-order ring
// sums and, for rings beyond the LDS, the FFT workspace.
static size_t synthesis_bytes(const Geometry &g, int Lc)
{
    const size_t npack = (size_t)tri_size(Lc);
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const size_t nws = fft_in_lds(g.fft.m) ? 0 : (size_t)std::min(g.P, kWsGroups) * ws_stride(g.fft.m);
    return npack * sizeof(double2) + nf * sizeof(double4) + nws * sizeof(double2);
}

static int launch_synthesis(const Geometry &g, const DevRecurrence &rt, const DevSeeds &sd, const double *cilm,
                            int d1, int d2, int Lc, int nrows_out, int ext_col, double *grid, long long ld,
                            hipStream_t st, const double *ctab = nullptr, const double2 *roww = nullptr,
                            const RhoRings &ring = RhoRings(), const ExtWorkspace &ext = ExtWorkspace())
{
    if (!ctab) ctab = rt.s;
    const double *rz = ring.r2 ? ring.z : g.z, *ru = ring.r2 ? ring.u : g.u;
    // the polar skip holds for the unit-radius functions only (rho^l may
    // grow what it drops)
    const int *kmax = ring.r2 ? nullptr : g.kmax, *smin = ring.r2 ? nullptr : g.smin;
    const size_t npack = (size_t)tri_size(Lc);
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    WsLease lease;
    const size_t bytes = synthesis_bytes(g, Lc);
    if (ext.p) {
        if (bytes > ext.cap) return ST_SCRATCH;
        lease.p = ext.p;
    } else if (int rc = ws_acquire(lease, bytes, st)) {
        return rc;
    }
    void *ws = lease.p;
    double2 *cpack = (double2 *)ws;
    double4 *fbuf = (double4 *)((char *)ws + npack * sizeof(double2));
    double2 *fws = (double2 *)((char *)fbuf + nf * sizeof(double4));

    {
        ProfScope ps(SHTG_K_PACK, st);
        const int ntm = (Lc + kTM) / kTM, ntl = (Lc + kTL) / kTL;
        hipLaunchKernelGGL(k_pack_synth, dim3(ntm * ntl), dim3(kThreads), 0, st, cilm, d1, d2, ctab, rt.L, cpack, Lc,
                           ntm);
    }
    {
        ProfScope ps(SHTG_K_SYNTH_LEG, st);
        const int kb = synth_lanes(Lc, g.P);
        const int nsb = (g.P + kb * kWave - 1) / (kb * kWave);
        hipLaunchKernelGGL(synth_leg_kernel(kb), dim3(leg_grid(Lc, nsb)), dim3(kWave), 0, st, rt.alpha, rt.L,
                           sd.seed, rz, cpack, Lc, g.P, nsb, fbuf, ring.r2, smin, kb);
    }
    {
        ProfScope ps(SHTG_K_SYNTH_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL(SHTG_FFT_KERNEL(k_synth_fft, plan_variant(g.fft)), dim3(g.P),
                               dim3(plan_threads(g.fft)), lds, st, g.fft, fbuf, g.P, Lc, ru,
                               g.rows, nrows_out, grid, ld, ext_col, roww, kmax);
        else
            hipLaunchKernelGGL(k_synth_fft_ws, dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft, fbuf, g.P, Lc, ru,
                               g.rows, nrows_out, grid, ld, ext_col, roww, fws, ws_stride(g.fft.m), kmax, kChunkCap);
    }
    const bool ok = HIP_OK(hipGetLastError());
    if (!ext.p) ws_release(lease, st);
    if (!ok) return ST_OTHER;
    return ST_OK;
}

// Analysis of rows [0, g.nlat).  The adjoint of synthesis passes unit
// weights w, all slots (s0 = 0), the recursion's s as ptab (no
// normalisation epilogue), the rows it wrote (nrows) and ext_col.
static int launch_analysis(const Geometry &g, const DevRecurrence &rt, const DevSeeds &sd, const double *grid,
                           long long ld, int Lc, double *cilm, int d1, int d2, hipStream_t st,
                           const double2 *w = nullptr, int s0 = -1, const double *ptab = nullptr, int nrows = -1,
                           int ext_col = 0)
{
    if (!w) w = g.w;
    if (s0 < 0) s0 = g.s0;
    if (!ptab) ptab = rt.post;
    if (nrows < 0) nrows = g.nlat;
    const size_t npack = (size_t)tri_size(Lc);
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    const size_t nws = lds_fft ? 0 : (size_t)nwg * ws_stride(g.fft.m);
    // the Legendre sums of each order in blocks of slots (one wave each) when
    // one wave per order would leave the GPU short of waves
    const int K = anal_lanes_per_slot(g.P - s0, plain_anal_kmax(Lc, g.P - s0));
    const int passes = (g.P - s0 + K * kWave - 1) / (K * kWave);
    const int nblk = anal_blocks(Lc, passes);
    const int bslots = (passes + nblk - 1) / nblk * K * kWave;
    WsLease lease;
    const size_t bytes = nblk * npack * sizeof(double2) + nf * sizeof(double4) + nws * sizeof(double2);
    if (int rc = ws_acquire(lease, bytes, st)) return rc;
    void *ws = lease.p;
    double2 *cbuf = (double2 *)ws;
    double4 *gbuf = (double4 *)((char *)ws + nblk * npack * sizeof(double2));
    double2 *fws = (double2 *)((char *)gbuf + nf * sizeof(double4));

    {
        ProfScope ps(SHTG_K_ANAL_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL(SHTG_FFT_KERNEL(k_anal_fft, plan_variant(g.fft)), dim3(g.P),
                               dim3(plan_threads(g.fft)), lds, st, g.fft, grid, ld, g.u, g.rows,
                               nrows, w, g.P, Lc, gbuf, ext_col, (const int *)g.kmax);
        else
            hipLaunchKernelGGL(k_anal_fft_ws, dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft, grid, ld, g.u, g.rows,
                               nrows, w, g.P, Lc, gbuf, ext_col, fws, ws_stride(g.fft.m), (const int *)g.kmax,
                               kChunkCap);
    }
    {
        ProfScope ps(SHTG_K_ANAL_LEG, st);
        hipLaunchKernelGGL(anal_leg_kernel(K), dim3(leg_grid(Lc, nblk)), dim3(kWave), 0, st, rt.alpha, rt.L, sd.seed,
                           g.z, gbuf, Lc, s0, g.P, cbuf, (const int *)g.smin, nblk, bslots, (long long)npack);
    }
    ProfScope ps_un(SHTG_K_UNPACK, st);
    const int ntm = (d2 + kTM - 1) / kTM, ntl = (d1 + kTL - 1) / kTL;
    hipLaunchKernelGGL(k_unpack_anal, dim3(ntm * ntl), dim3(kThreads), 0, st, cbuf, ptab, rt.L, Lc, cilm, d1, d2,
                       ntm, nblk, (long long)npack);
    const bool ok = HIP_OK(hipGetLastError());
    ws_release(lease, st);
    if (!ok) return ST_OTHER;
    return ST_OK;
}

// Complex synthesis (MakeGridDHC / MakeGridGLQC): pack both coefficient
// planes, the real Legendre kernel once per sign of m, complex ring FFTs.
static int launch_synthesis_c(const Geometry &g, const DevRecurrence &rt, const DevSeeds &sd, const double2 *cilm,
                              int d1, int d2, int Lc, int nrows_out, int ext_col, double2 *grid, long long ld,
                              hipStream_t st)
{
    const size_t npack = (size_t)tri_size(Lc);
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    const size_t nws = lds_fft ? 0 : (size_t)nwg * ws_stride(g.fft.m);
    WsLease lease;
    const size_t bytes = 2 * npack * sizeof(double2) + 2 * nf * sizeof(double4) + nws * sizeof(double2);
    if (int rc = ws_acquire(lease, bytes, st)) return rc;
    void *ws = lease.p;
    double2 *cp_p = (double2 *)ws, *cp_m = cp_p + npack;
    double4 *fb_p = (double4 *)(cp_m + npack), *fb_m = fb_p + nf;
    double2 *fws = (double2 *)(fb_m + nf);
    {
        ProfScope ps(SHTG_K_PACK, st);
        const int ntm = (Lc + kTM) / kTM, ntl = (Lc + kTL) / kTL;
        hipLaunchKernelGGL(k_pack_synth_c, dim3(ntm * ntl), dim3(kThreads), 0, st, cilm, d1, d2, rt.s, rt.L, cp_p,
                           cp_m, Lc, ntm);
    }
    {
        ProfScope ps(SHTG_K_SYNTH_LEG, st);
        const int kb = synth_lanes(Lc, g.P);
        const int nsb = (g.P + kb * kWave - 1) / (kb * kWave);
        hipLaunchKernelGGL(synth_leg_kernel(kb), dim3(leg_grid(Lc, nsb)), dim3(kWave), 0, st, rt.alpha, rt.L,
                           sd.seed, g.z, cp_p, Lc, g.P, nsb, fb_p, (const double *)nullptr, (const int *)nullptr, kb);
        hipLaunchKernelGGL(synth_leg_kernel(kb), dim3(leg_grid(Lc, nsb)), dim3(kWave), 0, st, rt.alpha, rt.L,
                           sd.seed, g.z, cp_m, Lc, g.P, nsb, fb_m, (const double *)nullptr, (const int *)nullptr, kb);
    }
    {
        ProfScope ps(SHTG_K_SYNTH_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL((by_variant(plan_variant(g.fft), k_synth_fft_c<false, kVarPlain>, k_synth_fft_c<false, kVarSplit>, k_synth_fft_c<false, kVarPrime>, k_synth_fft_c<false, kVarPow2>)), dim3(g.P),
                               dim3(plan_threads(g.fft)), lds, st, g.fft, fb_p, fb_m,
                               g.P, Lc, g.u, g.rows, nrows_out, grid, ld, ext_col, (double2 *)nullptr, 0LL, 0);
        else
            hipLaunchKernelGGL((k_synth_fft_c<true, kVarPlain>), dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft, fb_p, fb_m,
                               g.P, Lc, g.u, g.rows, nrows_out, grid, ld, ext_col, fws, ws_stride(g.fft.m),
                               kChunkCap);
    }
    const bool ok = HIP_OK(hipGetLastError());
    ws_release(lease, st);
    if (!ok) return ST_OTHER;
    return ST_OK;
}

// Complex analysis (SHExpandDHC / SHExpandGLQC).
static int launch_analysis_c(const Geometry &g, const DevRecurrence &rt, const DevSeeds &sd, const double2 *grid,
                             long long ld, int Lc, double2 *cilm, int d1, int d2, hipStream_t st)
{
    const size_t npack = (size_t)tri_size(Lc);
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    const size_t nws = lds_fft ? 0 : (size_t)nwg * ws_stride(g.fft.m);
    WsLease lease;
    const size_t bytes = 2 * npack * sizeof(double2) + 2 * nf * sizeof(double4) + nws * sizeof(double2);
    if (int rc = ws_acquire(lease, bytes, st)) return rc;
    void *ws = lease.p;
    double2 *cb_p = (double2 *)ws, *cb_m = cb_p + npack;
    double4 *gb_p = (double4 *)(cb_m + npack), *gb_m = gb_p + nf;
    double2 *fws = (double2 *)(gb_m + nf);
    {
        ProfScope ps(SHTG_K_ANAL_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL((by_variant(plan_variant(g.fft), k_anal_fft_c<false, kVarPlain>, k_anal_fft_c<false, kVarSplit>, k_anal_fft_c<false, kVarPrime>, k_anal_fft_c<false, kVarPow2>)), dim3(g.P),
                               dim3(plan_threads(g.fft)), lds, st, g.fft, grid, ld,
                               g.u, g.rows, g.nlat, g.w, g.P, Lc, gb_p, gb_m, (double2 *)nullptr, 0LL, 0);
        else
            hipLaunchKernelGGL((k_anal_fft_c<true, kVarPlain>), dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft, grid, ld,
                               g.u, g.rows, g.nlat, g.w, g.P, Lc, gb_p, gb_m, fws, ws_stride(g.fft.m), kChunkCap);
    }
    {
        ProfScope ps(SHTG_K_ANAL_LEG, st);
        const AnalLegKernel kern = anal_leg_kernel(anal_lanes_per_slot(g.P - g.s0));
        hipLaunchKernelGGL(kern, dim3(leg_grid(Lc, 1)), dim3(kWave), 0, st, rt.alpha, rt.L, sd.seed, g.z, gb_p, Lc, g.s0, g.P,
                           cb_p, (const int *)nullptr, 1, 0, 0LL);
        hipLaunchKernelGGL(kern, dim3(leg_grid(Lc, 1)), dim3(kWave), 0, st, rt.alpha, rt.L, sd.seed, g.z, gb_m, Lc, g.s0, g.P,
                           cb_m, (const int *)nullptr, 1, 0, 0LL);
    }
    {
        ProfScope ps(SHTG_K_UNPACK, st);
        const int ntm = (d2 + kTM - 1) / kTM, ntl = (d1 + kTL - 1) / kTL;
        hipLaunchKernelGGL(k_unpack_anal_c, dim3(ntm * ntl), dim3(kThreads), 0, st, cb_p, cb_m, rt.post_c, rt.L,
                           Lc, cilm, d1, d2, ntm);
    }
    const bool ok = HIP_OK(hipGetLastError());
    ws_release(lease, st);
    if (!ok) return ST_OTHER;
    return ST_OK;
}

// ------------------------------------------------------------ batched plans

// A batch of DH fields of uneven degree, transformed by one launch per stage
// (pack, Legendre, FFT grouped by LDS size, unpack) over work items sorted
// longest first.  Tables, ring geometry, FFT plans, item lists and
// workspace are built once per plan.
struct FftGroup {
    int2 *items = nullptr;
    int nitems = 0;
    size_t lds = 0;
    int threads = 256;       // synthesis workgroup size (heuristic or tuned)
    int threads_anal = 256;  // analysis workgroup size
    int min_threads = 1;     // the least its split plans allow (plan_min_threads)
    int max_threads = kFftMaxThreads;  // the largest its kernel instantiation allows
    int variant = kVarPlain;           // its plans' kind: the kernels' instantiation
};

struct BatchPlan {
    std::mutex mu;  // calls on the plan (bindings, workspace, tuning)
    int dev = -1, nf = 0, norm = 1, csphase = 1, Lmax = 0;
    std::vector<int> L;
    BField *fields = nullptr;
    int4 *items_pack = nullptr;
    int n_pack = 0;
    int4 *items_sleg = nullptr;
    int n_sleg = 0;
    // synthesis of same-degree buckets on MFMA (kb_synth_leg_mfma): buckets of
    // at least mfma_min fields (0: off) leave the VALU items
    int mfma_min = 0;
    int synth_kmax = kSynthKMax;  // ring pairs per lane of the synthesis Legendre waves (1 .. kSynthKMax)
    int anal_kmax = kBatchAnalKMax;  // ring pairs per lane and pass of the analysis waves (1 .. kBatchAnalKMax)
    int4 *items_smfma = nullptr;
    int n_smfma = 0;
    int *bfields = nullptr, *bnf = nullptr;  // per bucket: field indices, count
    std::vector<BField> hfields;             // host copies of fields and polar-skip rows
    std::vector<int> hsmin;
    std::vector<void *> retired;             // replaced item lists (freed with the plan)
    int2 *items_aleg[kBatchAnalKMax] = {};  // by ring pairs per lane K = index + 1
    int n_aleg[kBatchAnalKMax] = {};
    int4 *items_unpack = nullptr;
    int n_unpack = 0;
    std::vector<FftGroup> fft;
    FftGroup fft_ws;              // fields whose ring FFTs run on the global workspace
    std::vector<int> fft_group_of;  // per field: its index in fft (fft.size(): the workspace group)
    double2 *fws = nullptr;       // that workspace: kWsGroups slices of fws_stride
    long long fws_stride = 0;
    int fft_slots = 2;
    double *z = nullptr, *u = nullptr;
    double2 *w = nullptr, *ones = nullptr;
    int2 *rows = nullptr;
    int *kmax = nullptr, *smin = nullptr;  // polar skip tables (per slot; per distinct degree and order)
    FftDev *plans = nullptr;
    std::vector<FftDev> hplans;
    double2 *cbuf = nullptr;  // packed coefficients (synthesis) / sums (analysis)
    double4 *fbuf = nullptr;  // per-order ring sums
    // Field pointers bound at slots 0..3 (synthesis in, synthesis out,
    // analysis in, analysis out).  Each slot keeps a ring of kPtrRing device
    // tables; a new pointer list goes to the next table of the ring, copied
    // in stream order (hipMemcpyAsync on the call's stream from the table's
    // pinned host copy) after the host has waited for that table's previous
    // use (an event recorded after the launches that read it), so work still
    // queued on the stream keeps reading the table it was launched with.  A
    // table a graph capture used is pinned: it is never written again, and
    // the ring replaces it by a new table when it comes round.
    static constexpr int kPtrRing = 4;
    struct PtrTable {
        const void **dev = nullptr;
        const void **host = nullptr;  // pinned staging of the upload
        hipEvent_t done = nullptr;    // after the last launches that read it
        bool recorded = false;
        bool pinned = false;
        std::vector<const void *> ptrs;
    };
    struct PtrSlot {
        PtrTable ring[kPtrRing];
        int cur = -1;  // the bound table
        const void **dev() const { return cur >= 0 ? ring[cur].dev : nullptr; }
    } slot[4];
    std::vector<PtrTable> old_tables;  // pinned tables replaced in their ring
    BTables T{};
};

template <class T>
static T *dev_upload_n(const T *v, size_t n)
{
    std::vector<T> tmp(v, v + n);
    return dev_upload(tmp);
}

static void free_ptr_table(BatchPlan::PtrTable &t)
{
    if (t.done) (void)hipEventSynchronize(t.done);
    (void)hipFree((void *)t.dev);
    if (t.host) (void)hipHostFree((void *)t.host);
    if (t.done) (void)hipEventDestroy(t.done);
    t = BatchPlan::PtrTable{};
}

static void batch_free(BatchPlan *b)
{
    if (!b) return;
    for (auto &s : b->slot)
        for (auto &t : s.ring) free_ptr_table(t);
    for (auto &t : b->old_tables) free_ptr_table(t);
    (void)hipFree(b->fields);
    (void)hipFree(b->items_pack);
    (void)hipFree(b->items_sleg);
    (void)hipFree(b->items_smfma);
    (void)hipFree(b->bfields);
    (void)hipFree(b->bnf);
    for (void *p : b->retired) (void)hipFree(p);
    for (auto p : b->items_aleg) (void)hipFree(p);
    (void)hipFree(b->items_unpack);
    for (auto &g : b->fft) (void)hipFree(g.items);
    (void)hipFree(b->fft_ws.items);
    (void)hipFree(b->fws);
    (void)hipFree(b->z);
    (void)hipFree(b->u);
    (void)hipFree(b->w);
    (void)hipFree(b->ones);
    (void)hipFree(b->rows);
    (void)hipFree(b->kmax);
    (void)hipFree(b->smin);
    (void)hipFree(b->plans);
    for (auto &p : b->hplans) free_fft(p);
    (void)hipFree(b->cbuf);
    (void)hipFree(b->fbuf);
    delete b;
}

// Synthesis Legendre work items of a plan (build_synth_items): same-degree
// buckets of at least mfma_min fields (0: none) to the MFMA kernel, the rest
// to the VALU kernel with b->synth_kmax ring pairs per lane.
static int batch_synth_items(BatchPlan *b, int mfma_min)
{
    const SynthItems it = build_synth_items(b->hfields, b->hsmin, b->synth_kmax, mfma_min);
    // the replaced lists may still be read by queued work or captured graphs:
    // they are kept until the plan is destroyed
    for (void *p : {(void *)b->items_sleg, (void *)b->items_smfma, (void *)b->bfields, (void *)b->bnf})
        if (p) b->retired.push_back(p);
    b->items_sleg = nullptr;
    b->items_smfma = nullptr;
    b->bfields = b->bnf = nullptr;
    b->n_sleg = (int)it.valu.size();
    b->n_smfma = (int)it.mfma.size();
    if (b->n_sleg && !(b->items_sleg = reinterpret_cast<int4 *>(dev_upload(it.valu)))) return ST_ALLOC;
    if (b->n_smfma) {
        b->items_smfma = reinterpret_cast<int4 *>(dev_upload(it.mfma));
        b->bfields = dev_upload(it.bfields);
        b->bnf = dev_upload(it.bnf);
        if (!b->items_smfma || !b->bfields || !b->bnf) return ST_ALLOC;
    }
    b->mfma_min = mfma_min;
    return ST_OK;
}

// Analysis Legendre work items (build_anal_items) with at most b->anal_kmax
// ring pairs per lane.  Replaced lists are kept until the plan is destroyed.
static int batch_anal_items(BatchPlan *b)
{
    std::vector<Item2> ia[kBatchAnalKMax];
    build_anal_items(b->hfields, b->hsmin, b->anal_kmax, ia);
    for (int k = 0; k < kBatchAnalKMax; ++k) {
        if (b->items_aleg[k]) b->retired.push_back(b->items_aleg[k]);
        b->items_aleg[k] = nullptr;
        b->n_aleg[k] = (int)ia[k].size();
        if (b->n_aleg[k] && !(b->items_aleg[k] = reinterpret_cast<int2 *>(dev_upload(ia[k])))) return ST_ALLOC;
    }
    return ST_OK;
}

// sampling: per field (1 or 2).
static int batch_create(int nfields, const int *lmax, const int *grid, int norm, int csphase, BatchPlan **out)
{
    if (nfields <= 0 || !lmax || !grid || !out) return ST_DIMS;
    if (norm < 1 || norm > 4 || (csphase != 1 && csphase != -1)) return ST_BOUNDS;
    int Lmax = 0;
    for (int i = 0; i < nfields; ++i) {
        if (lmax[i] < 0 || lmax[i] > kMaxArgDegree) return ST_BOUNDS;
        if (grid[i] != kGridDH1 && grid[i] != kGridDH2 && grid[i] != kGridGLQ) return ST_BOUNDS;
        Lmax = std::max(Lmax, lmax[i]);
    }
    const int dev = current_device();
    if (dev < 0) return ST_OTHER;
    configure_kernels();
    std::unique_ptr<BatchPlan, void (*)(BatchPlan *)> b(new BatchPlan(), batch_free);
    b->dev = dev;
    b->nf = nfields;
    b->norm = norm;
    b->csphase = csphase;
    b->Lmax = Lmax;
    b->L.assign(lmax, lmax + nfields);
    const DevRecurrence *rt = get_recurrence(dev, norm, Lmax);
    const DevSeeds *sd = get_seeds(dev, norm, csphase, Lmax);
    if (!rt || !sd) return ST_ALLOC;

    // ring geometry, FFT plans, field offsets and work items (layout.hpp)
    BatchLayout lay;
    if (!build_batch_layout(nfields, lmax, grid, lay)) return ST_OTHER;
    for (const FftPlanHost &fp : lay.plans) {
        FftDev fd;
        if (!upload_fft(fp, fd)) return ST_OTHER;
        b->hplans.push_back(fd);
    }
    b->z = dev_upload(lay.z);
    b->u = dev_upload(lay.u);
    b->w = reinterpret_cast<double2 *>(dev_upload(lay.w));
    b->ones = reinterpret_cast<double2 *>(dev_upload(std::vector<double>(lay.w.size(), 1.0)));
    b->rows = reinterpret_cast<int2 *>(dev_upload(lay.rows));
    b->kmax = dev_upload(lay.kmax);
    b->smin = dev_upload(lay.smin);
    b->plans = dev_upload(b->hplans);
    if (!b->z || !b->u || !b->w || !b->ones || !b->rows || !b->kmax || !b->smin || !b->plans) return ST_ALLOC;
    b->fields = dev_upload(lay.fields);
    if (!b->fields) return ST_ALLOC;
    if (!HIP_OK(hipMalloc(&b->cbuf, (size_t)lay.ncpack * sizeof(double2)))) return ST_ALLOC;
    if (!HIP_OK(hipMalloc(&b->fbuf, (size_t)lay.nfbuf * sizeof(double4)))) return ST_ALLOC;
    b->hfields = lay.fields;
    b->hsmin = lay.smin;
    b->items_pack = reinterpret_cast<int4 *>(dev_upload(lay.pack));
    b->items_unpack = reinterpret_cast<int4 *>(dev_upload(lay.unpack));
    b->n_pack = (int)lay.pack.size();
    b->n_unpack = (int)lay.unpack.size();
    if (!b->items_pack || !b->items_unpack) return ST_ALLOC;
    if (int rc = batch_synth_items(b.get(), 0)) return rc;
    if (int rc = batch_anal_items(b.get())) return rc;

    b->fft_slots = fft_slots_per_wg();
    const FftLayout fl = build_fft_groups(lay.fields, lay.plans, b->fft_slots);
    b->fft_group_of = fl.group_of;
    for (const FftGroupLayout &gl : fl.groups) {
        FftGroup g;
        g.items = reinterpret_cast<int2 *>(dev_upload(gl.items));
        g.nitems = (int)gl.items.size();
        g.lds = gl.lds;
        g.threads = g.threads_anal = gl.threads;
        g.min_threads = gl.min_threads;
        g.variant = gl.variant;
        g.max_threads = gl.max_threads;
        if (!g.items) return ST_ALLOC;
        b->fft.push_back(g);
    }
    if (!fl.ws_items.empty()) {
        b->fws_stride = fl.ws_stride;
        b->fft_ws.items = reinterpret_cast<int2 *>(dev_upload(fl.ws_items));
        b->fft_ws.nitems = (int)fl.ws_items.size();
        b->fft_ws.threads = kFftMaxThreads;
        if (!b->fft_ws.items) return ST_ALLOC;
        if (!HIP_OK(hipMalloc(&b->fws, (size_t)kWsGroups * b->fws_stride * sizeof(double2)))) return ST_ALLOC;
    }
    b->T.alpha = rt->alpha;
    b->T.s = rt->s;
    b->T.post = rt->post;
    b->T.seed = sd->seed;
    b->T.Lt = rt->L;
    b->T.z = b->z;
    b->T.u = b->u;
    b->T.w = b->w;
    b->T.roww = nullptr;
    b->T.s0 = 1;
    b->T.rows = b->rows;
    b->T.plans = b->plans;
    b->T.kmax = b->kmax;
    b->T.smin = b->smin;
    *out = b.release();
    return ST_OK;
}

static bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return HIP_OK(hipStreamIsCapturing(st, &cs)) && cs != hipStreamCaptureStatusNone;
}

// Bind field pointers p[0..nf) to slot k (0: synthesis in, 1: synthesis
// out, 2: analysis in, 3: analysis out) for launches on stream st.  An
// unchanged list keeps its table.  A new list goes to the next table of the
// slot's ring (see BatchPlan::PtrTable), uploaded in stream order; it cannot
// be bound under graph capture (ST_CAPTURE: capture after an eager call with
// the same arrays), and a list bound under capture pins its table.
static int batch_bind(BatchPlan *b, int k, const void *const *p, hipStream_t st)
{
    for (int i = 0; i < b->nf; ++i)
        if (!p[i]) return ST_DIMS;
    BatchPlan::PtrSlot &s = b->slot[k];
    const bool cap = capturing(st);
    if (s.cur >= 0 && std::equal(p, p + b->nf, s.ring[s.cur].ptrs.begin())) {
        if (cap) s.ring[s.cur].pinned = true;
        return ST_OK;
    }
    if (cap) return ST_CAPTURE;
    const int next = (s.cur + 1) % BatchPlan::kPtrRing;
    BatchPlan::PtrTable &t = s.ring[next];
    if (t.pinned) {
        // a captured graph reads this table: keep it until the plan goes
        b->old_tables.push_back(t);
        t = BatchPlan::PtrTable{};
    }
    const size_t bytes = sizeof(void *) * (size_t)b->nf;
    if (!t.dev) {
        if (!HIP_OK(hipMalloc((void **)&t.dev, bytes))) { t.dev = nullptr; return ST_ALLOC; }
        if (!HIP_OK(hipHostMalloc((void **)&t.host, bytes, hipHostMallocDefault))) { t.host = nullptr; return ST_ALLOC; }
        if (!HIP_OK(hipEventCreateWithFlags(&t.done, hipEventDisableTiming))) { t.done = nullptr; return ST_OTHER; }
    }
    // the table's previous launches (kPtrRing binds ago) must be done with it
    // and with its staging copy
    if (t.recorded && !HIP_OK(hipEventSynchronize(t.done))) return ST_OTHER;
    std::copy(p, p + b->nf, t.host);
    if (!HIP_OK(hipMemcpyAsync((void *)t.dev, t.host, bytes, hipMemcpyHostToDevice, st))) return ST_OTHER;
    t.ptrs.assign(p, p + b->nf);
    t.recorded = false;
    s.cur = next;
    return ST_OK;
}

// After the launches of a call that read slots k0 and k1: mark their bound
// tables' last use (not under capture: those tables are pinned instead).
static void batch_mark_used(BatchPlan *b, int k0, int k1, hipStream_t st)
{
    if (capturing(st)) return;
    for (int k : {k0, k1}) {
        BatchPlan::PtrSlot &s = b->slot[k];
        if (s.cur < 0) continue;
        BatchPlan::PtrTable &t = s.ring[s.cur];
        t.recorded = HIP_OK(hipEventRecord(t.done, st));
    }
}

// workgroups covering n items under xcd_item's mapping (multiple of 32)
static int xcd_blocks(int n) { return ((n + 31) / 32) * 32; }

// Synthesis of every field with tables T (b->T; the adjoint of analysis
// scales the coefficients by post and the rows by the analysis weights).
static int batch_synthesis(BatchPlan *b, hipStream_t st, const BTables &T)
{
    const double *const *cin = (const double *const *)b->slot[0].dev();
    double *const *gout = (double *const *)b->slot[1].dev();
    {
        ProfScope ps(SHTG_K_PACK, st);
        hipLaunchKernelGGL(kb_pack, dim3(b->n_pack), dim3(kThreads), 0, st, b->fields, b->items_pack, T, cin,
                           b->cbuf);
    }
    {
        ProfScope ps(SHTG_K_SYNTH_LEG, st);
        if (b->n_smfma)
            hipLaunchKernelGGL(kb_synth_leg_mfma, dim3(b->n_smfma), dim3(kWave), 0, st, b->fields, b->items_smfma,
                               b->bfields, b->bnf, T, b->cbuf, b->fbuf);
        if (b->n_sleg)
            hipLaunchKernelGGL(kb_synth_leg, dim3(b->n_sleg), dim3(kWave), 0, st, b->fields, b->items_sleg, T,
                               b->cbuf, b->fbuf);
    }
    {
        ProfScope ps(SHTG_K_SYNTH_FFT, st);
        for (auto &g : b->fft)
            hipLaunchKernelGGL(SHTG_FFT_KERNEL(kb_synth_fft, g.variant), dim3(xcd_blocks(g.nitems)),
                               dim3(g.threads), g.lds, st, b->fields,
                               g.items, g.nitems, b->fft_slots, T, b->fbuf, gout);
        if (b->fft_ws.nitems)
            hipLaunchKernelGGL(kb_synth_fft_ws, dim3(std::min(b->fft_ws.nitems, kWsGroups)), dim3(kFftMaxThreads),
                               kMaxLds, st, b->fields, b->fft_ws.items, b->fft_ws.nitems, b->fft_slots, T, b->fbuf,
                               gout, b->fws, b->fws_stride, kChunkCap);
    }
    batch_mark_used(b, 0, 1, st);
    return HIP_OK(hipGetLastError()) ? ST_OK : ST_OTHER;
}

// Analysis of every field with tables T (b->T; the adjoint of synthesis
// uses unit weights, every slot and the recursion's s as epilogue).
static int batch_analysis(BatchPlan *b, hipStream_t st, const BTables &T)
{
    const double *const *gin = (const double *const *)b->slot[2].dev();
    double *const *cout = (double *const *)b->slot[3].dev();
    {
        ProfScope ps(SHTG_K_ANAL_FFT, st);
        for (auto &g : b->fft)
            hipLaunchKernelGGL(SHTG_FFT_KERNEL(kb_anal_fft, g.variant), dim3(xcd_blocks(g.nitems)),
                               dim3(g.threads_anal), g.lds, st, b->fields,
                               g.items, g.nitems, b->fft_slots, T, gin, b->fbuf);
        if (b->fft_ws.nitems)
            hipLaunchKernelGGL(kb_anal_fft_ws, dim3(std::min(b->fft_ws.nitems, kWsGroups)), dim3(kFftMaxThreads),
                               kMaxLds, st, b->fields, b->fft_ws.items, b->fft_ws.nitems, b->fft_slots, T, gin,
                               b->fbuf, b->fws, b->fws_stride, kChunkCap);
    }
    {
        ProfScope ps(SHTG_K_ANAL_LEG, st);
        for (int k = 0; k < kBatchAnalKMax; ++k)
            if (b->n_aleg[k])
                hipLaunchKernelGGL(batch_anal_leg_kernel(k + 1), dim3(b->n_aleg[k]), dim3(kWave), 0, st, b->fields,
                                   b->items_aleg[k], T, b->fbuf, b->cbuf);
    }
    {
        ProfScope ps(SHTG_K_UNPACK, st);
        hipLaunchKernelGGL(kb_unpack, dim3(b->n_unpack), dim3(kThreads), 0, st, b->fields, b->items_unpack, T,
                           b->cbuf, cout);
    }
    batch_mark_used(b, 2, 3, st);
    return HIP_OK(hipGetLastError()) ? ST_OK : ST_OTHER;
}

// ------------------------------------------------------------ autotuning

// The Legendre launches of one direction on the plan's current items and
// workspace (synthesis: packed coefficients in cbuf -> fbuf; analysis: fbuf
// -> cbuf); best of reps, in ms (negative on a launch error).
static float time_legendre(BatchPlan *b, hipStream_t st, int reps, bool synth, hipEvent_t e0, hipEvent_t e1)
{
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        (void)hipEventRecord(e0, st);
        if (synth) {
            if (b->n_smfma)
                hipLaunchKernelGGL(kb_synth_leg_mfma, dim3(b->n_smfma), dim3(kWave), 0, st, b->fields, b->items_smfma,
                                   b->bfields, b->bnf, b->T, b->cbuf, b->fbuf);
            if (b->n_sleg)
                hipLaunchKernelGGL(kb_synth_leg, dim3(b->n_sleg), dim3(kWave), 0, st, b->fields, b->items_sleg, b->T,
                                   b->cbuf, b->fbuf);
        } else {
            for (int k = 0; k < kBatchAnalKMax; ++k)
                if (b->n_aleg[k])
                    hipLaunchKernelGGL(batch_anal_leg_kernel(k + 1), dim3(b->n_aleg[k]), dim3(kWave), 0, st,
                                       b->fields, b->items_aleg[k], b->T, b->fbuf, b->cbuf);
        }
        (void)hipEventRecord(e1, st);
        float ms = 0.0f;
        if (!HIP_OK(hipGetLastError()) || !HIP_OK(hipEventSynchronize(e1)) || !HIP_OK(hipEventElapsedTime(&ms, e0, e1)))
            return -1.0f;
        best = std::min(best, ms);
    }
    return best;
}

// Legendre part of the autotuner: ring pairs per lane of the synthesis
// waves (2 .. kSynthKMax), then, when the plan has fields of equal degree
// and sampling, the smallest bucket that goes to the MFMA kernel (off, 2, 4,
// 8, 16 -- only thresholds that change which buckets qualify); each time the
// fastest launch is kept.  These choices leave every result bitwise as it
// was.  With tune_analysis, also ring pairs per lane of the analysis waves
// (4, 6, 8), which regroups the sums over ring pairs (results change by
// rounding).
static int batch_tune_legendre(BatchPlan *b, hipStream_t st, int reps, bool tune_analysis)
{
    hipEvent_t e0, e1;
    if (!HIP_OK(hipEventCreate(&e0)) || !HIP_OK(hipEventCreate(&e1))) return ST_OTHER;
    int rc = ST_OK;
    auto pick = [&](std::vector<int> cands, bool synth, std::function<int(int)> apply, int &keep) {
        float best = 1e30f;
        int best_v = keep;
        for (int v : cands) {
            if ((rc = apply(v))) return;
            const float ms = time_legendre(b, st, reps, synth, e0, e1);
            if (ms < 0.0f) { rc = ST_OTHER; return; }
            if (ms < best) { best = ms; best_v = v; }
        }
        keep = best_v;
        rc = apply(best_v);
    };
    int sk = b->synth_kmax, mm = b->mfma_min, ak = b->anal_kmax;
    pick({2, 3, kSynthKMax}, true, [&](int v) { b->synth_kmax = v; return batch_synth_items(b, mm); }, sk);
    // largest same-degree group of the plan
    std::map<std::pair<int, int>, int> cnt;
    int biggest = 0;
    for (const BField &f : b->hfields) biggest = std::max(biggest, ++cnt[{f.L, f.nlong}]);
    if (!rc && biggest >= 2) {
        std::vector<int> thr = {0};
        for (int t : {2, 4, 8, 16})
            if (t <= biggest) thr.push_back(t);
        pick(thr, true, [&](int v) { return batch_synth_items(b, v); }, mm);
    }
    if (!rc && tune_analysis)
        pick({4, 6, 8, kBatchAnalKMax}, false, [&](int v) { b->anal_kmax = v; return batch_anal_items(b); }, ak);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Launch-shape autotuner of a batch plan: for every ring-FFT group, times the
// synthesis and the analysis kernels at 256, 512 and 1024 threads per
// workgroup (the group's items and LDS are unchanged) on the plan's bound
// arrays, and keeps the fastest of each.  The runs overwrite the bound grids
// and the plan's workspace; the caller reruns its transforms afterwards.
// Workgroup sizes the ring-FFT autotuner tries.  Besides powers of two, 384
// and 768 threads: a stage of B butterflies takes ceil(B / threads)
// iterations between barriers, and for the Bluestein lengths of the largest
// rings (m = 5760 .. 6144, one workgroup per CU) 768 threads leave fewer
// lanes idle than 1024.
constexpr int kFftShapes[] = {256, 384, 512, 768, 1024};

static bool valid_fft_shape(int t, int min_threads, int max_threads)
{
    return t % kWave == 0 && t >= kFftShapes[0] && t <= max_threads && t >= min_threads;
}

static int batch_tune(BatchPlan *b, hipStream_t st, int reps, int flags)
{
    const double *const *gin = (const double *const *)b->slot[2].dev();
    double *const *gout = (double *const *)b->slot[1].dev();
    if (!gin || !gout) return ST_OTHER;  // needs a prior synthesis and analysis
    hipEvent_t e0, e1;
    if (!HIP_OK(hipEventCreate(&e0)) || !HIP_OK(hipEventCreate(&e1))) return ST_OTHER;
    int st_out = ST_OK;
    for (auto &g : b->fft) {
        for (int dir = 0; dir < 2; ++dir) {
            float best = 1e30f;
            int best_t = dir ? g.threads_anal : g.threads;
            for (int t : kFftShapes) {
                if (t < g.min_threads || t > g.max_threads) continue;
                float tmin = 1e30f;
                for (int r = 0; r < reps; ++r) {
                    (void)hipEventRecord(e0, st);
                    if (dir == 0)
                        hipLaunchKernelGGL(SHTG_FFT_KERNEL(kb_synth_fft, g.variant),
                                           dim3(xcd_blocks(g.nitems)), dim3(t), g.lds, st, b->fields,
                                           g.items, g.nitems, b->fft_slots, b->T, b->fbuf, gout);
                    else
                        hipLaunchKernelGGL(SHTG_FFT_KERNEL(kb_anal_fft, g.variant),
                                           dim3(xcd_blocks(g.nitems)), dim3(t), g.lds, st, b->fields,
                                           g.items, g.nitems, b->fft_slots, b->T, gin, b->fbuf);
                    (void)hipEventRecord(e1, st);
                    float ms = 0.0f;
                    if (!HIP_OK(hipGetLastError()) || !HIP_OK(hipEventSynchronize(e1)) ||
                        !HIP_OK(hipEventElapsedTime(&ms, e0, e1))) {
                        st_out = ST_OTHER;
                        break;
                    }
                    tmin = std::min(tmin, ms);
                }
                if (st_out) break;
                if (tmin < best) {
                    best = tmin;
                    best_t = t;
                }
            }
            if (st_out) break;
            (dir ? g.threads_anal : g.threads) = best_t;
        }
        if (st_out) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st_out) return st_out;
    return batch_tune_legendre(b, st, reps, (flags & 1) != 0);
}

// Device tables of one call: the geometry (never changed once built) and
// copies of the recursion and seed descriptors, taken under g_mu.
struct CallTables {
    const Geometry *g = nullptr;
    DevRecurrence rt;
    DevSeeds sd;
};

// Whether a call's tables exist already (nothing to build or upload).
static bool tables_ready(int dev, bool dh, int lmax, int sampling, const double *zero, const double *wq, int norm,
                         int csphase, int Lc, bool cplx)
{
    if (!kernels_configured) return false;
    auto r = g_rec.find({dev, norm});
    if (r == g_rec.end() || r->second.L < Lc) return false;
    auto s = g_seeds.find({dev, norm + (cplx ? 8 : 0), csphase});
    if (s == g_seeds.end() || s->second.L < Lc) return false;
    if (dh) {
        auto g = g_dh.find({dev, lmax, sampling});
        return g != g_dh.end() && g->second && g->second->ok;
    }
    auto g = g_glq.find({dev, lmax, glq_key(lmax, zero, wq)});
    return g != g_glq.end() && g->second && g->second->ok;
}

// Under graph capture a call may not build tables (device allocations and
// synchronous uploads are not allowed there): it needs an eager call of the
// same degree, grid and normalisation before the capture, else ST_CAPTURE.
static int get_tables(bool dh, int lmax, int sampling, const double *zero, const double *wq, int norm,
                      int csphase, int Lc, CallTables &t, hipStream_t st, bool cplx = false)
{
    const int dev = current_device();
    if (dev < 0) return ST_OTHER;
    std::lock_guard<std::mutex> lk(g_mu);
    if (capturing(st) && !tables_ready(dev, dh, lmax, sampling, zero, wq, norm, csphase, Lc, cplx))
        return ST_CAPTURE;
    configure_kernels();
    t.g = dh ? get_dh(dev, lmax, sampling) : get_glq(dev, lmax, zero, wq);
    const DevRecurrence *rt = get_recurrence(dev, norm, Lc);
    const DevSeeds *sd = get_seeds(dev, norm, csphase, Lc, cplx);
    if (!t.g || !rt || !sd) return ST_ALLOC;
    t.rt = *rt;
    t.sd = *sd;
    return ST_OK;
}

// Adjoint of synthesis: cilm_bar = S^T grid_bar over the rows synthesis
// writes (with extend the south pole row and the repeated column), all
// slots with unit weight and the raw Legendre functions (no epilogue).
static int adjoint_of_synthesis(const CallTables &t, const Shape &sh, const double *grid_bar, long long ld,
                                double *cilm_bar, int d1, int d2, hipStream_t st)
{
    return launch_analysis(*t.g, t.rt, t.sd, grid_bar, ld, sh.Lc, cilm_bar, d1, d2, st, t.g->ones, 0, t.rt.s,
                           sh.nrows, sh.ext);
}

// Adjoint of analysis: grid_bar = A^T cilm_bar = W S (post * cilm_bar) on
// the rows analysis reads; the rest of the array is zeroed.
static int adjoint_of_analysis(const CallTables &t, const Shape &sh, const double *cilm_bar, int d1, int d2,
                               double *grid_bar, int gd0, int gd1, hipStream_t st)
{
    const Geometry &g = *t.g;
    zero_outside(grid_bar, gd0, gd1, gd1, g.nlat, g.nlong, st);
    return launch_synthesis(g, t.rt, t.sd, cilm_bar, d1, d2, sh.Lc, g.nlat, 0, grid_bar, gd1, st, t.rt.post,
                            g.w);
}

}  // namespace shtg

using namespace shtg;

extern "C" {

int shtg_makegrid_dh(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int norm, int sampling, int csphase,
                     int lmax_calc, int extend, double *grid, int grid_d0, int grid_d1, void *stream)
{
    Shape sh;
    int st = check_makegrid_dh(cilm_d1, cilm_d2, lmax, norm, sampling, csphase, lmax_calc, extend, grid_d0, grid_d1,
                               sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, lmax, sampling, nullptr, nullptr, norm, csphase, sh.Lc, t,
                         (hipStream_t)stream))) return st;
    return launch_synthesis(*t.g, t.rt, t.sd, cilm, cilm_d1, cilm_d2, sh.Lc, sh.nrows, extend, grid, grid_d1,
                            (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------ no-allocation synthesis
//
// shtg_makegrid_dh_noalloc keeps a synthesis' intermediates in `scratch`, a
// device array of kScratchBytes that every GPU the library runs on holds
// until the process ends (each GPU its own copy at its own address).  In
// host code the array's name stands for no device address: its address on
// a device comes from hipGetSymbolAddress, taken once per device with that
// device current (on the entry point's first call there) and kept in
// g_scratch, and only that address is passed to kernels.  Every call on a
// GPU shares that GPU's array: calls that overlap in time on one GPU (other
// streams, other host threads) race on it.  The synthesis writes every
// scratch element it later reads, so nothing needs clearing between calls.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void *g_scratch[kMaxDevices] = {};

static void *scratch_address(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (!HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) || !p) return nullptr;
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

extern "C" {

int shtg_makegrid_dh_noalloc(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int norm, int sampling,
                             int csphase, int lmax_calc, int extend, double *grid, int grid_d0, int grid_d1,
                             void *stream)
{
    Shape sh;
    int st = check_makegrid_dh(cilm_d1, cilm_d2, lmax, norm, sampling, csphase, lmax_calc, extend, grid_d0, grid_d1,
                               sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    int dev = 0;
    if (!HIP_OK(hipGetDevice(&dev))) return ST_OTHER;
    const Geometry *gp;
    DevRecurrence rt;
    DevSeeds sd;
    void *ws;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        // only tables an ordinary call has built already (building them would
        // allocate)
        if (!tables_ready(dev, true, lmax, sampling, nullptr, nullptr, norm, csphase, sh.Lc, false))
            return ST_NOTABLES;
        ws = scratch_address(dev);
        if (!ws) return ST_OTHER;
        gp = g_dh.find({dev, lmax, sampling})->second.get();
        rt = g_rec.find({dev, norm})->second;
        sd = g_seeds.find({dev, norm, csphase})->second;
    }
    const Geometry &g = *gp;
    if (synthesis_bytes(g, sh.Lc) > kScratchBytes) return ST_SCRATCH;
    ExtWorkspace ext;
    ext.p = ws;
    ext.cap = kScratchBytes;
    return launch_synthesis(g, rt, sd, cilm, cilm_d1, cilm_d2, sh.Lc, sh.nrows, extend, grid, grid_d1,
                            (hipStream_t)stream, nullptr, nullptr, RhoRings(), ext);
}

int shtg_shexpand_dh(const double *grid, int grid_d0, int grid_d1, int norm, int sampling, int csphase,
                     int lmax_calc, double *cilm, int cilm_d1, int cilm_d2, void *stream)
{
    Shape sh;
    int st = check_shexpand_dh(grid_d0, grid_d1, norm, sampling, csphase, lmax_calc, cilm_d1, cilm_d2, sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, sh.lmax, sampling, nullptr, nullptr, norm, csphase, sh.Lc, t,
                         (hipStream_t)stream))) return st;
    return launch_analysis(*t.g, t.rt, t.sd, grid, grid_d1, sh.Lc, cilm, cilm_d1, cilm_d2, (hipStream_t)stream);
}

int shtg_makegrid_glq(const double *cilm, int cilm_d1, int cilm_d2, int lmax, const double *zero, int nzero,
                      int norm, int csphase, int lmax_calc, int extend, double *grid, int grid_d0, int grid_d1,
                      void *stream)
{
    Shape sh;
    int st = check_makegrid_glq(cilm_d1, cilm_d2, lmax, nzero, zero, norm, csphase, lmax_calc, extend, grid_d0,
                                grid_d1, sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(false, lmax, 0, zero, nullptr, norm, csphase, sh.Lc, t, (hipStream_t)stream))) return st;
    return launch_synthesis(*t.g, t.rt, t.sd, cilm, cilm_d1, cilm_d2, sh.Lc, sh.nrows, extend, grid, grid_d1,
                            (hipStream_t)stream);
}

int shtg_shexpand_glq(const double *grid, int grid_d0, int grid_d1, const double *w, int nw, const double *zero,
                      int nzero, int norm, int csphase, int lmax_calc, double *cilm, int cilm_d1, int cilm_d2,
                      void *stream)
{
    Shape sh;
    int st = check_shexpand_glq(grid_d0, grid_d1, w, nw, zero, nzero, norm, csphase, lmax_calc, cilm_d1, cilm_d2,
                                sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(false, sh.lmax, 0, zero, w, norm, csphase, sh.Lc, t, (hipStream_t)stream))) return st;
    return launch_analysis(*t.g, t.rt, t.sd, grid, grid_d1, sh.Lc, cilm, cilm_d1, cilm_d2, (hipStream_t)stream);
}

int shtg_makegrid_dh_adjoint(const double *grid_bar, int grid_d0, int grid_d1, int lmax, int norm, int sampling,
                             int csphase, int lmax_calc, int extend, double *cilm_bar, int cilm_d1, int cilm_d2,
                             void *stream)
{
    Shape sh;
    int st = check_makegrid_dh(cilm_d1, cilm_d2, lmax, norm, sampling, csphase, lmax_calc, extend, grid_d0, grid_d1,
                               sh);
    if (st) return st;
    if (!cilm_bar || !grid_bar) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, lmax, sampling, nullptr, nullptr, norm, csphase, sh.Lc, t,
                         (hipStream_t)stream))) return st;
    return adjoint_of_synthesis(t, sh, grid_bar, grid_d1, cilm_bar, cilm_d1, cilm_d2, (hipStream_t)stream);
}

int shtg_shexpand_dh_adjoint(const double *cilm_bar, int cilm_d1, int cilm_d2, int norm, int sampling, int csphase,
                             int lmax_calc, double *grid_bar, int grid_d0, int grid_d1, void *stream)
{
    Shape sh;
    int st = check_shexpand_dh(grid_d0, grid_d1, norm, sampling, csphase, lmax_calc, cilm_d1, cilm_d2, sh);
    if (st) return st;
    if (!cilm_bar || !grid_bar) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, sh.lmax, sampling, nullptr, nullptr, norm, csphase, sh.Lc, t,
                         (hipStream_t)stream))) return st;
    return adjoint_of_analysis(t, sh, cilm_bar, cilm_d1, cilm_d2, grid_bar, grid_d0, grid_d1, (hipStream_t)stream);
}

int shtg_makegrid_glq_adjoint(const double *grid_bar, int grid_d0, int grid_d1, int lmax, const double *zero,
                              int nzero, int norm, int csphase, int lmax_calc, int extend, double *cilm_bar,
                              int cilm_d1, int cilm_d2, void *stream)
{
    Shape sh;
    int st = check_makegrid_glq(cilm_d1, cilm_d2, lmax, nzero, zero, norm, csphase, lmax_calc, extend, grid_d0,
                                grid_d1, sh);
    if (st) return st;
    if (!cilm_bar || !grid_bar) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(false, lmax, 0, zero, nullptr, norm, csphase, sh.Lc, t, (hipStream_t)stream))) return st;
    return adjoint_of_synthesis(t, sh, grid_bar, grid_d1, cilm_bar, cilm_d1, cilm_d2, (hipStream_t)stream);
}

int shtg_shexpand_glq_adjoint(const double *cilm_bar, int cilm_d1, int cilm_d2, const double *w, int nw,
                              const double *zero, int nzero, int norm, int csphase, int lmax_calc, double *grid_bar,
                              int grid_d0, int grid_d1, void *stream)
{
    Shape sh;
    int st = check_shexpand_glq(grid_d0, grid_d1, w, nw, zero, nzero, norm, csphase, lmax_calc, cilm_d1, cilm_d2,
                                sh);
    if (st) return st;
    if (!cilm_bar || !grid_bar) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(false, sh.lmax, 0, zero, w, norm, csphase, sh.Lc, t, (hipStream_t)stream))) return st;
    return adjoint_of_analysis(t, sh, cilm_bar, cilm_d1, cilm_d2, grid_bar, grid_d0, grid_d1, (hipStream_t)stream);
}

int shtg_makegrid_dhc(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int norm, int sampling, int csphase,
                      int lmax_calc, int extend, double *grid, int grid_d0, int grid_d1, void *stream)
{
    Shape sh;
    int st = check_makegrid_dh(cilm_d1, cilm_d2, lmax, norm, sampling, csphase, lmax_calc, extend, grid_d0, grid_d1,
                               sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, lmax, sampling, nullptr, nullptr, norm, csphase, sh.Lc, t,
                         (hipStream_t)stream, true))) return st;
    return launch_synthesis_c(*t.g, t.rt, t.sd, (const double2 *)cilm, cilm_d1, cilm_d2, sh.Lc, sh.nrows, extend,
                              (double2 *)grid, grid_d1, (hipStream_t)stream);
}

int shtg_shexpand_dhc(const double *grid, int grid_d0, int grid_d1, int norm, int sampling, int csphase,
                      int lmax_calc, double *cilm, int cilm_d1, int cilm_d2, void *stream)
{
    Shape sh;
    int st = check_shexpand_dh(grid_d0, grid_d1, norm, sampling, csphase, lmax_calc, cilm_d1, cilm_d2, sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, sh.lmax, sampling, nullptr, nullptr, norm, csphase, sh.Lc, t,
                         (hipStream_t)stream, true))) return st;
    return launch_analysis_c(*t.g, t.rt, t.sd, (const double2 *)grid, grid_d1, sh.Lc, (double2 *)cilm, cilm_d1,
                             cilm_d2, (hipStream_t)stream);
}

int shtg_makegrid_glqc(const double *cilm, int cilm_d1, int cilm_d2, int lmax, const double *zero, int nzero,
                       int norm, int csphase, int lmax_calc, int extend, double *grid, int grid_d0, int grid_d1,
                       void *stream)
{
    Shape sh;
    int st = check_makegrid_glq(cilm_d1, cilm_d2, lmax, nzero, zero, norm, csphase, lmax_calc, extend, grid_d0,
                                grid_d1, sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(false, lmax, 0, zero, nullptr, norm, csphase, sh.Lc, t, (hipStream_t)stream, true))) return st;
    return launch_synthesis_c(*t.g, t.rt, t.sd, (const double2 *)cilm, cilm_d1, cilm_d2, sh.Lc, sh.nrows, extend,
                              (double2 *)grid, grid_d1, (hipStream_t)stream);
}

int shtg_shexpand_glqc(const double *grid, int grid_d0, int grid_d1, const double *w, int nw, const double *zero,
                       int nzero, int norm, int csphase, int lmax_calc, double *cilm, int cilm_d1, int cilm_d2,
                       void *stream)
{
    Shape sh;
    int st = check_shexpand_glq(grid_d0, grid_d1, w, nw, zero, nzero, norm, csphase, lmax_calc, cilm_d1, cilm_d2,
                                sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(false, sh.lmax, 0, zero, w, norm, csphase, sh.Lc, t, (hipStream_t)stream, true))) return st;
    return launch_analysis_c(*t.g, t.rt, t.sd, (const double2 *)grid, grid_d1, sh.Lc, (double2 *)cilm, cilm_d1,
                             cilm_d2, (hipStream_t)stream);
}

int shtg_makegradient_dh(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int sampling, int lmax_calc,
                         int extend, double radius, double *theta, double *phi, int grid_d0, int grid_d1,
                         void *stream)
{
    // argument checks in the order of src/MakeGradientDH.f95:104-226
    Shape sh;
    int st = check_makegrid_dh(cilm_d1, cilm_d2, lmax, 1, sampling, 1, lmax_calc, extend, grid_d0, grid_d1, sh);
    if (st) return st;
    if (!cilm || !theta || !phi) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, lmax, sampling, nullptr, nullptr, 1, 1, sh.Lc, t, (hipStream_t)stream))) return st;
    hipStream_t hs = (hipStream_t)stream;
    const int L1 = sh.Lc + 1, n = 2 * lmax + 2;
    const size_t plane = (size_t)L1 * L1;
    void *ws = nullptr;
    const size_t ngrid = (size_t)grid_d0 * grid_d1;
    if (!HIP_OK(pool_alloc(&ws, (6 * plane + ngrid) * sizeof(double), hs))) return ST_ALLOC;
    double *coefs = (double *)ws, *gb = coefs + 6 * plane;
    hipLaunchKernelGGL(k_gradient_coefs, dim3(grid_blocks((long long)plane)), dim3(kThreads), 0, hs, cilm, cilm_d1,
                       cilm_d2, sh.Lc, coefs);
    st = launch_synthesis(*t.g, t.rt, t.sd, coefs, L1, L1, sh.Lc, sh.nrows, extend, theta, grid_d1, hs);
    if (!st) st = launch_synthesis(*t.g, t.rt, t.sd, coefs + 2 * plane, L1, L1, sh.Lc, sh.nrows, extend, gb,
                                   grid_d1, hs);
    if (!st) st = launch_synthesis(*t.g, t.rt, t.sd, coefs + 4 * plane, L1, L1, sh.Lc, sh.nrows, extend, phi,
                                   grid_d1, hs);
    if (!st) {
        const int ncols = sampling * n + extend;
        hipLaunchKernelGGL(k_gradient_combine, dim3(grid_blocks((long long)sh.nrows * ncols)), dim3(kThreads), 0, hs,
                           theta, phi, gb, sh.nrows, ncols, (long long)grid_d1, n, radius);
        if (!HIP_OK(hipGetLastError())) st = ST_OTHER;
    }
    (void)hipFreeAsync(ws, hs);
    return st;
}

// ------------------------------------------------------------ gravity

// NormalGravity of src/NormalGravity.f95:1-139 (Somigliana's closed form).
static double normal_gravity(double geocentric_lat, double gm, double omega, double a, double b)
{
    const double pi = std::acos(-1.0);
    if (a == b && omega == 0.0) return gm / (a * a);
    if (a == b) {
        const double cl = std::cos(geocentric_lat * pi / 180.0);
        return std::sqrt(gm * gm / (a * a * a * a) + a * omega * omega * cl * cl * (a * omega * omega - 2.0 * gm / (a * a)));
    }
    const double m = omega * omega * a * a * b / gm;
    const double bigE = std::sqrt(a * a - b * b), ep = bigE / b;
    const double q0 = 0.5 * ((1.0 + 3.0 * (b / bigE) * (b / bigE)) * std::atan(bigE / b) - 3.0 * b / bigE);
    const double q0p = 3.0 * (1.0 + (b / bigE) * (b / bigE)) * (1.0 - b / bigE * std::atan(bigE / b)) - 1.0;
    const double ga = gm / (a * b) * (1.0 - m - m * ep * q0p / 6.0 / q0);
    const double gb = gm / (a * a) * (1.0 + m * ep * q0p / 3.0 / q0);
    const double gl = std::atan((a / b) * (a / b) * std::tan(geocentric_lat * pi / 180.0));
    const double c = std::cos(gl), sn = std::sin(gl);
    return (a * ga * c * c + b * gb * sn * sn) / std::sqrt(a * a * c * c + b * b * sn * sn);
}

// Ring and row tables of one gravity grid: per slot rho z, rho u and rho^2
// with rho = r0 / r of the slot's ellipsoid radius, and the GravRow of each
// output row.  Cached by every parameter they depend on.
struct GravTables {
    double *z = nullptr, *u = nullptr, *r2 = nullptr;
    GravRow *rows = nullptr;
    bool ok = false;
};
static std::map<std::vector<uint64_t>, std::unique_ptr<GravTables>> g_grav;

static uint64_t dbits(double x)
{
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    return b;
}

static const GravTables *get_grav(int dev, int lmax, int sampling, int extend, double r0, double a, double f,
                                  double omega, int normal_grav, double gm_ng)
{
    const std::vector<uint64_t> key = {(uint64_t)dev, (uint64_t)lmax, (uint64_t)sampling, (uint64_t)extend,
                                       dbits(r0), dbits(a), dbits(f), dbits(omega), (uint64_t)normal_grav,
                                       dbits(gm_ng)};
    auto &slot = g_grav[key];
    if (slot) return slot->ok ? slot.get() : nullptr;
    slot.reset(new GravTables());
    GravTables &t = *slot;
    const int n = 2 * lmax + 2, P = n / 2 + 1, nrows = n + extend;
    const double pi = std::acos(-1.0);
    // the ellipsoid radius of row pair s (src/MakeGravGridDH.f95:435,622-631)
    auto r_of_slot = [&](int s) {
        if (s == n / 2) return a;
        if (s == 0) return a * (1.0 - f);
        const double lat = pi / 2.0 - pi * (double)s / (double)n;
        const double q = std::cos(lat) * std::cos(lat) + std::sin(lat) * std::sin(lat) / ((1.0 - f) * (1.0 - f));
        return a * std::sqrt(1.0 / q);
    };
    std::vector<double> z(P), u(P), r2(P);
    for (int s = 0; s < P; ++s) {
        const double zz = (s == n / 2) ? 0.0 : std::cos(pi * (double)s / (double)n);
        const double uu = (s == n / 2) ? 1.0 : std::sqrt((1.0 - zz) * (1.0 + zz));
        const double rho = r0 / r_of_slot(s);
        z[s] = rho * zz;
        u[s] = rho * uu;
        r2[s] = rho * rho;
    }
    std::vector<GravRow> rows(nrows);
    const double b = a * (1.0 - f);
    for (int i = 0; i < nrows; ++i) {
        const int s = (2 * i <= n) ? i : n - i;  // the row pair's slot (northern row index)
        const double tn = (s == n / 2) ? pi / 2.0 : pi * (double)s / (double)n;
        GravRow &R = rows[i];
        R.z = (s == n / 2) ? 0.0 : (i == s ? std::cos(tn) : -std::cos(tn));
        R.st = std::sin(tn);
        const double zn = (s == n / 2) ? 0.0 : std::cos(tn);
        R.u2 = (s == n / 2) ? 1.0 : (1.0 - zn) * (1.0 + zn);
        R.rex = r_of_slot(s);
        R.rho = r0 / R.rex;
        R.pole = (s == 0) ? 1.0 : 0.0;
        // centrifugal terms and normal gravity at the row's own colatitude
        // (src/MakeGravGridDH.f95:1004-1050)
        const double th = pi * (double)i / (double)n, lat = pi / 2.0 - th, tl = std::tan(lat);
        const double rr = a * std::sqrt((1.0 + tl * tl) / (1.0 + tl * tl / ((1.0 - f) * (1.0 - f))));
        R.crad = R.cth = R.cpot = 0.0;
        if (omega != 0.0) {
            R.crad = rr * (std::sin(th) * omega) * (std::sin(th) * omega);
            R.cth = std::sin(th) * std::cos(th) * rr * omega * omega;
            R.cpot = 0.5 * (rr * std::sin(th) * omega) * (rr * std::sin(th) * omega);
        }
        R.ng = normal_grav ? normal_gravity(lat * 180.0 / pi, gm_ng, omega, a, b) : 0.0;
    }
    t.z = dev_upload(z);
    t.u = dev_upload(u);
    t.r2 = dev_upload(r2);
    t.rows = dev_upload(rows);
    t.ok = t.z && t.u && t.r2 && t.rows;
    return t.ok ? &t : nullptr;
}

int shtg_makegravgrid_dh(const double *cilm, int cilm_d1, int cilm_d2, int lmax, double gm, double r0, double a,
                         double f, int sampling, int lmax_calc, double omega, int normal_gravity_flag,
                         double normal_gravity_gm, int extend, double *rad, double *theta, double *phi,
                         double *total, double *pot, int grid_d0, int grid_d1, void *stream)
{
    // argument checks in the order of src/MakeGravGridDH.f95:140-300
    Shape sh;
    int st = check_makegrid_dh(cilm_d1, cilm_d2, lmax, 1, sampling, 1, lmax_calc, extend, grid_d0, grid_d1, sh);
    if (st) return st;
    if (normal_gravity_flag != 0 && normal_gravity_flag != 1) return ST_BOUNDS;
    if (!cilm || !rad || !theta || !phi || !total || !pot) return ST_DIMS;
    CallTables t;
    if ((st = get_tables(true, lmax, sampling, nullptr, nullptr, 1, 1, sh.Lc, t, (hipStream_t)stream))) return st;
    const GravTables *gt;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        gt = get_grav(current_device(), lmax, sampling, extend, r0, a, f, omega, normal_gravity_flag,
                      normal_gravity_gm);
    }
    if (!gt) return ST_ALLOC;
    hipStream_t hs = (hipStream_t)stream;
    const int L1 = sh.Lc + 1, n = 2 * lmax + 2;
    const size_t plane = (size_t)L1 * L1;
    void *ws = nullptr;
    if (!HIP_OK(pool_alloc(&ws, 6 * plane * sizeof(double), hs))) return ST_ALLOC;
    double *coefs = (double *)ws;
    hipLaunchKernelGGL(k_gradient_coefs, dim3(grid_blocks((long long)plane)), dim3(kThreads), 0, hs, cilm, cilm_d1,
                       cilm_d2, sh.Lc, coefs);
    RhoRings ring;
    ring.z = gt->z;
    ring.u = gt->u;
    ring.r2 = gt->r2;
    const Geometry &g = *t.g;
    st = launch_synthesis(g, t.rt, t.sd, cilm, cilm_d1, cilm_d2, sh.Lc, sh.nrows, extend, pot, grid_d1, hs,
                          nullptr, nullptr, ring);
    if (!st) st = launch_synthesis(g, t.rt, t.sd, coefs + 2 * plane, L1, L1, sh.Lc, sh.nrows, extend, rad, grid_d1,
                                   hs, nullptr, nullptr, ring);
    if (!st) st = launch_synthesis(g, t.rt, t.sd, coefs, L1, L1, sh.Lc, sh.nrows, extend, theta, grid_d1, hs,
                                   nullptr, nullptr, ring);
    if (!st) st = launch_synthesis(g, t.rt, t.sd, coefs + 4 * plane, L1, L1, sh.Lc, sh.nrows, extend, phi, grid_d1,
                                   hs, nullptr, nullptr, ring);
    if (!st) {
        const int ncols = sampling * n + extend;
        hipLaunchKernelGGL(k_grav_combine, dim3(grid_blocks((long long)sh.nrows * ncols)), dim3(kThreads), 0, hs, rad,
                           theta, phi, total, pot, gt->rows, sh.nrows, ncols, (long long)grid_d1, gm);
        if (!HIP_OK(hipGetLastError())) st = ST_OTHER;
    }
    (void)hipFreeAsync(ws, hs);
    return st;
}

int shtg_normal_gravity(double geocentric_lat, double gm, double omega, double a, double b, double *out)
{
    if (!out) return ST_DIMS;
    *out = normal_gravity(geocentric_lat, gm, omega, a, b);
    return ST_OK;
}

// ------------------------------------------------------------ products

// Gauss-Legendre nodes and weights of nlat points, cached (host).
static const std::pair<std::vector<double>, std::vector<double>> *glq_cached(int nlat)
{
    static std::map<int, std::pair<std::vector<double>, std::vector<double>>> cache;
    auto it = cache.find(nlat);
    if (it != cache.end()) return &it->second;
    std::vector<double> z(nlat), w(nlat);
    if (glq_nodes(nlat, z.data(), w.data())) return nullptr;
    return &(cache[nlat] = {std::move(z), std::move(w)});
}

// ------------------------------------------------------------ plx tables
//
// The GLQ routines' optional table of Legendre functions at the nodes (the
// reference's plx argument): SHGLQ builds it, MakeGridGLQ and SHExpandGLQ
// read it in place of the recursion.  The table is device memory of
// (lmax+1) (lmax+1)(lmax+2)/2 doubles, plx[k (lmax+1) + i] (the Fortran
// array plx(i, k) in C order); a transform streams the part of degrees up
// to lmax_calc once.

}  // extern "C"

namespace shtg {

// Geometry of a plx call (ring FFT plan and, for analysis, the caller's
// weights wq) on the library's own Gauss-Legendre nodes.
static int plx_tables(int lmax, const double *wq, int Lc, CallTables &t, hipStream_t st)
{
    const std::pair<std::vector<double>, std::vector<double>> *zw;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        zw = glq_cached(lmax + 1);  // entries are never changed once made
    }
    if (!zw) return ST_OTHER;
    return get_tables(false, lmax, 0, zw->first.data(), wq, 1, 1, Lc, t, st);
}

static int launch_plx_synthesis(const Geometry &g, const double *cilm, int d1, int d2, int Lc, const double *plx,
                                int ext_col, double *grid, long long ld, hipStream_t st)
{
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    const size_t nws = lds_fft ? 0 : (size_t)nwg * ws_stride(g.fft.m);
    WsLease lease;
    if (int rc = ws_acquire(lease, nf * sizeof(double4) + nws * sizeof(double2), st)) return rc;
    double4 *fbuf = (double4 *)lease.p;
    double2 *fws = (double2 *)((char *)fbuf + nf * sizeof(double4));
    {
        ProfScope ps(SHTG_K_SYNTH_LEG, st);
        const int nblk = (g.P + kWave - 1) / kWave;
        hipLaunchKernelGGL(k_plx_synth_leg, dim3((Lc + 1) * nblk), dim3(kWave), 0, st, cilm, d1, d2, Lc, plx, g.nlat,
                           g.rows, g.P, nblk, fbuf);
    }
    {
        // no ring u: the sums are function values (no order rescale)
        ProfScope ps(SHTG_K_SYNTH_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL(SHTG_FFT_KERNEL(k_synth_fft, plan_variant(g.fft)), dim3(g.P),
                               dim3(plan_threads(g.fft)), lds, st, g.fft, fbuf, g.P, Lc, (const double *)nullptr,
                               g.rows, g.nlat, grid, ld, ext_col, (const double2 *)nullptr, (const int *)nullptr);
        else
            hipLaunchKernelGGL(k_synth_fft_ws, dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft, fbuf, g.P, Lc,
                               (const double *)nullptr, g.rows, g.nlat, grid, ld, ext_col, (const double2 *)nullptr,
                               fws, ws_stride(g.fft.m), (const int *)nullptr, kChunkCap);
    }
    const bool ok = HIP_OK(hipGetLastError());
    ws_release(lease, st);
    return ok ? ST_OK : ST_OTHER;
}

static int launch_plx_analysis(const Geometry &g, const double *grid, long long ld, int Lc, const double *plx,
                               int norm, double *cilm, int d1, int d2, hipStream_t st)
{
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    const size_t nws = lds_fft ? 0 : (size_t)nwg * ws_stride(g.fft.m);
    WsLease lease;
    if (int rc = ws_acquire(lease, nf * sizeof(double4) + nws * sizeof(double2), st)) return rc;
    double4 *gbuf = (double4 *)lease.p;
    double2 *fws = (double2 *)((char *)gbuf + nf * sizeof(double4));
    zero_outside(cilm, 2 * d1, d2, d2, 0, 0, st);
    {
        // no ring u: the rows' coefficients unscaled and apart
        ProfScope ps(SHTG_K_ANAL_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL(SHTG_FFT_KERNEL(k_anal_fft, plan_variant(g.fft)), dim3(g.P),
                               dim3(plan_threads(g.fft)), lds, st, g.fft, grid, ld, (const double *)nullptr, g.rows,
                               g.nlat, g.w, g.P, Lc, gbuf, 0, (const int *)nullptr);
        else
            hipLaunchKernelGGL(k_anal_fft_ws, dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft, grid, ld,
                               (const double *)nullptr, g.rows, g.nlat, g.w, g.P, Lc, gbuf, 0, fws,
                               ws_stride(g.fft.m), (const int *)nullptr, kChunkCap);
    }
    {
        ProfScope ps(SHTG_K_ANAL_LEG, st);
        const int nblk = (Lc + kWave) / kWave;
        hipLaunchKernelGGL(k_plx_anal_leg, dim3((Lc + 1) * nblk), dim3(kWave), 0, st, gbuf, g.P, g.nlat, Lc, plx,
                           g.nlat, norm, nblk, cilm, d1, d2);
    }
    const bool ok = HIP_OK(hipGetLastError());
    ws_release(lease, st);
    return ok ? ST_OK : ST_OTHER;
}

// Complex counterparts (MakeGridGLQC / SHExpandGLQC with plx).
static int launch_plx_synthesis_c(const Geometry &g, const double2 *cilm, int d1, int d2, int Lc, const double *plx,
                                  int ext_col, double2 *grid, long long ld, hipStream_t st)
{
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    const size_t nws = lds_fft ? 0 : (size_t)nwg * ws_stride(g.fft.m);
    WsLease lease;
    if (int rc = ws_acquire(lease, 2 * nf * sizeof(double4) + nws * sizeof(double2), st)) return rc;
    double4 *fb_p = (double4 *)lease.p, *fb_m = fb_p + nf;
    double2 *fws = (double2 *)(fb_m + nf);
    {
        ProfScope ps(SHTG_K_SYNTH_LEG, st);
        const int nblk = (g.P + kWave - 1) / kWave;
        hipLaunchKernelGGL(k_plx_synth_leg_c, dim3((Lc + 1) * nblk), dim3(kWave), 0, st, cilm, d1, d2, Lc, plx,
                           g.nlat, g.rows, g.P, nblk, fb_p, fb_m);
    }
    {
        ProfScope ps(SHTG_K_SYNTH_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL((by_variant(plan_variant(g.fft), k_synth_fft_c<false, kVarPlain>,
                                           k_synth_fft_c<false, kVarSplit>, k_synth_fft_c<false, kVarPrime>, k_synth_fft_c<false, kVarPow2>)),
                               dim3(g.P), dim3(plan_threads(g.fft)), lds, st, g.fft, fb_p, fb_m, g.P, Lc,
                               (const double *)nullptr, g.rows, g.nlat, grid, ld, ext_col, (double2 *)nullptr, 0LL, 0);
        else
            hipLaunchKernelGGL((k_synth_fft_c<true, kVarPlain>), dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft,
                               fb_p, fb_m, g.P, Lc, (const double *)nullptr, g.rows, g.nlat, grid, ld, ext_col, fws,
                               ws_stride(g.fft.m), kChunkCap);
    }
    const bool ok = HIP_OK(hipGetLastError());
    ws_release(lease, st);
    return ok ? ST_OK : ST_OTHER;
}

static int launch_plx_analysis_c(const Geometry &g, const double2 *grid, long long ld, int Lc, const double *plx,
                                 int norm, double2 *cilm, int d1, int d2, hipStream_t st)
{
    const size_t nf = (size_t)(Lc + 1) * g.P;
    const bool lds_fft = fft_in_lds(g.fft.m);
    const int nwg = std::min(g.P, kWsGroups);
    const size_t nws = lds_fft ? 0 : (size_t)nwg * ws_stride(g.fft.m);
    WsLease lease;
    if (int rc = ws_acquire(lease, 2 * nf * sizeof(double4) + nws * sizeof(double2), st)) return rc;
    double4 *gb_p = (double4 *)lease.p, *gb_m = gb_p + nf;
    double2 *fws = (double2 *)(gb_m + nf);
    zero_outside((double *)cilm, 2 * d1, 2 * d2, 2 * d2, 0, 0, st);
    {
        ProfScope ps(SHTG_K_ANAL_FFT, st);
        const size_t lds = (size_t)fft_lds_slots(g.fft.m) * sizeof(double2);
        if (lds_fft)
            hipLaunchKernelGGL((by_variant(plan_variant(g.fft), k_anal_fft_c<false, kVarPlain>,
                                           k_anal_fft_c<false, kVarSplit>, k_anal_fft_c<false, kVarPrime>, k_anal_fft_c<false, kVarPow2>)),
                               dim3(g.P), dim3(plan_threads(g.fft)), lds, st, g.fft, grid, ld, (const double *)nullptr,
                               g.rows, g.nlat, g.w, g.P, Lc, gb_p, gb_m, (double2 *)nullptr, 0LL, 0);
        else
            hipLaunchKernelGGL((k_anal_fft_c<true, kVarPlain>), dim3(nwg), dim3(kFftMaxThreads), kMaxLds, st, g.fft,
                               grid, ld, (const double *)nullptr, g.rows, g.nlat, g.w, g.P, Lc, gb_p, gb_m, fws,
                               ws_stride(g.fft.m), kChunkCap);
    }
    {
        ProfScope ps(SHTG_K_ANAL_LEG, st);
        const int nblk = (Lc + kWave) / kWave;
        hipLaunchKernelGGL(k_plx_anal_leg_c, dim3((Lc + 1) * nblk), dim3(kWave), 0, st, gb_p, gb_m, g.P, g.nlat, Lc,
                           plx, g.nlat, norm, nblk, cilm, d1, d2);
    }
    const bool ok = HIP_OK(hipGetLastError());
    ws_release(lease, st);
    return ok ? ST_OK : ST_OTHER;
}

// SHGLQ's table for degree lmax at the nodes zero (host, lmax + 1 of them)
// into plx, from the recursion tables of its normalisation.
static int launch_plx_table(const DevRecurrence &rt, const DevSeeds &sd, const std::vector<double> &zero, int lmax,
                            double *plx, hipStream_t st)
{
    const int n = lmax + 1;
    // the nodes on the device, from the caller's host copy (which outlives
    // the copy), in a block of the per-device intermediates
    WsLease lease;
    if (int rc = ws_acquire(lease, (size_t)n * sizeof(double), st)) return rc;
    bool ok = HIP_OK(hipMemcpyAsync(lease.p, zero.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
    if (ok) {
        const int nblk = ((n + 1) / 2 + kWave - 1) / kWave;
        hipLaunchKernelGGL(k_plx_table, dim3((lmax + 1) * nblk), dim3(kWave), 0, st, rt.alpha, rt.s, rt.L, sd.seed,
                           (const double *)lease.p, n, lmax, nblk, plx);
        ok = HIP_OK(hipGetLastError());
    }
    ws_release(lease, st);
    return ok ? ST_OK : ST_OTHER;
}

// A degree-0 field on the Gauss-Legendre grid: every node cilm(1,1,1) P_00
// (src/MakeGridGLQ.f95:408-420, which precedes the plx branch).
static int fill_degree0(const double *cilm, int norm, int nlat, int nlong_out, double *grid, long long ld,
                        hipStream_t st)
{
    const double pm2 = (norm == 4) ? 1.0 / std::sqrt(4.0 * std::acos(-1.0)) : 1.0;
    hipLaunchKernelGGL(k_fill, dim3(grid_blocks((long long)nlat * nlong_out)), dim3(kThreads), 0, st, grid, nlat,
                       nlong_out, ld, cilm, pm2);
    return HIP_OK(hipGetLastError()) ? ST_OK : ST_OTHER;
}

}  // namespace shtg

extern "C" {

int shtg_shglq_plx(int lmax, double *zero, double *w, double *plx, int norm, int csphase, int cnorm,
                   void *stream)
{
    // argument checks in the order of src/SHGLQ.f95:64-150
    if (lmax < 0 || lmax > kMaxArgDegree) return ST_BOUNDS;
    if (!zero || !w || !plx) return ST_DIMS;
    if (norm < 1 || norm > 4) return ST_BOUNDS;
    if (csphase != 1 && csphase != -1) return ST_BOUNDS;
    const int n = lmax + 1;
    hipStream_t hs = (hipStream_t)stream;
    const int dev = current_device();
    if (dev < 0) return ST_OTHER;
    const std::pair<std::vector<double>, std::vector<double>> *zw;
    DevRecurrence rt;
    DevSeeds sd;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        zw = glq_cached(n);
        if (!zw) return ST_OTHER;
        if (capturing(hs)) return ST_CAPTURE;  // table uploads and a host-to-device copy follow
        const DevRecurrence *r = get_recurrence(dev, norm, lmax);
        // PLegendreA (norm 3) takes no cnorm (src/SHGLQ.f95:225-227)
        const DevSeeds *s = get_seeds(dev, norm, csphase, lmax, cnorm == 1 && norm != 3);
        if (!r || !s) return ST_ALLOC;
        rt = *r;
        sd = *s;
    }
    std::copy(zw->first.begin(), zw->first.end(), zero);
    std::copy(zw->second.begin(), zw->second.end(), w);
    // the cache's host copy of the nodes is kept for the process
    return launch_plx_table(rt, sd, zw->first, lmax, plx, hs);
}

int shtg_makegrid_glq_plx(const double *cilm, int cilm_d1, int cilm_d2, int lmax, const double *plx, int norm,
                          int csphase, int lmax_calc, int extend, double *grid, int grid_d0, int grid_d1,
                          void *stream)
{
    Shape sh;
    int st = check_makegrid_glq_plx(cilm_d1, cilm_d2, lmax, plx, norm, csphase, lmax_calc, extend, grid_d0, grid_d1,
                                    sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    hipStream_t hs = (hipStream_t)stream;
    if (sh.Lc == 0) return fill_degree0(cilm, norm, lmax + 1, 2 * lmax + 1 + extend, grid, grid_d1, hs);
    CallTables t;
    if ((st = plx_tables(lmax, nullptr, sh.Lc, t, hs))) return st;
    return launch_plx_synthesis(*t.g, cilm, cilm_d1, cilm_d2, sh.Lc, plx, extend, grid, grid_d1, hs);
}

int shtg_shexpand_glq_plx(const double *grid, int grid_d0, int grid_d1, const double *w, int nw, const double *plx,
                          int norm, int csphase, int lmax_calc, double *cilm, int cilm_d1, int cilm_d2, void *stream)
{
    Shape sh;
    int st = check_shexpand_glq_plx(grid_d0, grid_d1, w, nw, plx, norm, csphase, lmax_calc, cilm_d1, cilm_d2, sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    hipStream_t hs = (hipStream_t)stream;
    CallTables t;
    if ((st = plx_tables(sh.lmax, w, sh.Lc, t, hs))) return st;
    return launch_plx_analysis(*t.g, grid, grid_d1, sh.Lc, plx, norm, cilm, cilm_d1, cilm_d2, hs);
}

int shtg_makegrid_glqc_plx(const double *cilm, int cilm_d1, int cilm_d2, int lmax, const double *plx, int norm,
                           int csphase, int lmax_calc, int extend, double *grid, int grid_d0, int grid_d1,
                           void *stream)
{
    Shape sh;
    int st = check_makegrid_glq_plx(cilm_d1, cilm_d2, lmax, plx, norm, csphase, lmax_calc, extend, grid_d0, grid_d1,
                                    sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    hipStream_t hs = (hipStream_t)stream;
    if (sh.Lc == 0) {
        // a degree-0 field: every node cilm(1,1,1) P_00 (src/MakeGridGLQC.f95:411-421)
        const double pm2 = (norm == 4) ? 1.0 / std::sqrt(4.0 * std::acos(-1.0)) : 1.0;
        const int nlong_out = 2 * lmax + 1 + extend;
        hipLaunchKernelGGL(k_fill_c, dim3(grid_blocks((long long)(lmax + 1) * nlong_out)), dim3(kThreads), 0, hs,
                           (double2 *)grid, lmax + 1, nlong_out, (long long)grid_d1, (const double2 *)cilm, pm2);
        return HIP_OK(hipGetLastError()) ? ST_OK : ST_OTHER;
    }
    CallTables t;
    if ((st = plx_tables(lmax, nullptr, sh.Lc, t, hs))) return st;
    return launch_plx_synthesis_c(*t.g, (const double2 *)cilm, cilm_d1, cilm_d2, sh.Lc, plx, extend, (double2 *)grid,
                                  grid_d1, hs);
}

int shtg_shexpand_glqc_plx(const double *grid, int grid_d0, int grid_d1, const double *w, int nw, const double *plx,
                           int norm, int csphase, int lmax_calc, double *cilm, int cilm_d1, int cilm_d2, void *stream)
{
    Shape sh;
    int st = check_shexpand_glq_plx(grid_d0, grid_d1, w, nw, plx, norm, csphase, lmax_calc, cilm_d1, cilm_d2, sh);
    if (st) return st;
    if (!cilm || !grid) return ST_DIMS;
    hipStream_t hs = (hipStream_t)stream;
    CallTables t;
    if ((st = plx_tables(sh.lmax, w, sh.Lc, t, hs))) return st;
    return launch_plx_analysis_c(*t.g, (const double2 *)grid, grid_d1, sh.Lc, plx, norm, (double2 *)cilm, cilm_d1,
                                 cilm_d2, hs);
}

int shtg_shmultiply(double *cilmout, int out_d1, int out_d2, const double *cilm1, int c1_d1, int c1_d2, int lmax1,
                    const double *cilm2, int c2_d1, int c2_d2, int lmax2, int precomp, int norm, int csphase,
                    void *stream)
{
    // argument checks in the order of src/SHMultiply.f95:61-180
    if (lmax1 < 0 || lmax2 < 0 || lmax1 > kMaxArgDegree || lmax2 > kMaxArgDegree - lmax1) return ST_BOUNDS;
    if (c1_d1 < lmax1 + 1 || c1_d2 < lmax1 + 1) return ST_DIMS;
    if (c2_d1 < lmax2 + 1 || c2_d2 < lmax2 + 1) return ST_DIMS;
    const int lmaxout = lmax1 + lmax2;
    if (out_d1 < lmaxout + 1 || out_d2 < lmaxout + 1) return ST_DIMS;
    if (csphase != 1 && csphase != -1) return ST_BOUNDS;
    if (precomp != 0 && precomp != 1) return ST_BOUNDS;
    if (norm < 1 || norm > 4) return ST_BOUNDS;
    if (!cilmout || !cilm1 || !cilm2) return ST_DIMS;
    const int nlat = lmaxout + 1, nlong = 2 * lmaxout + 1;
    const std::pair<std::vector<double>, std::vector<double>> *zw;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        zw = glq_cached(nlat);  // entries are never changed once made
    }
    if (!zw) return ST_OTHER;
    CallTables t;
    int st = get_tables(false, lmaxout, 0, zw->first.data(), zw->second.data(), norm, csphase, lmaxout, t,
                        (hipStream_t)stream);
    if (st) return st;
    hipStream_t hs = (hipStream_t)stream;
    if (precomp == 1 && capturing(hs)) return ST_CAPTURE;  // the table's nodes are copied from the host
    const size_t ng = (size_t)nlat * nlong;
    // precomp = 1: SHGLQ's table of the functions at the nodes, nlat values
    // per (l, m), read by the transforms' table branches
    // (src/SHMultiply.f95:281-325)
    const size_t nplx = (precomp == 1) ? (size_t)nlat * ((size_t)nlat * (nlat + 1) / 2) : 0;
    void *ws = nullptr;
    if (!HIP_OK(pool_alloc(&ws, (2 * ng + nplx) * sizeof(double), hs))) return ST_ALLOC;
    double *g1 = (double *)ws, *g2 = g1 + ng, *plx = (precomp == 1) ? g2 + ng : nullptr;
    if (plx) {
        st = launch_plx_table(t.rt, t.sd, zw->first, lmaxout, plx, hs);
        // each factor synthesised on the product's grid from its leading
        // (lmax+1)^2 block, the table read up to the factor's degree
        auto factor = [&](const double *c, int d1, int d2, int L, double *g) {
            return L == 0 ? fill_degree0(c, norm, nlat, nlong, g, nlong, hs)
                          : launch_plx_synthesis(*t.g, c, d1, d2, L, plx, 0, g, nlong, hs);
        };
        if (!st) st = factor(cilm1, c1_d1, c1_d2, lmax1, g1);
        if (!st) st = factor(cilm2, c2_d1, c2_d2, lmax2, g2);
    } else {
        // each factor synthesised on the product's grid from its leading
        // (lmax+1)^2 block (src/SHMultiply.f95:330-350)
        st = launch_synthesis(*t.g, t.rt, t.sd, cilm1, c1_d1, c1_d2, lmax1, nlat, 0, g1, nlong, hs);
        if (!st) st = launch_synthesis(*t.g, t.rt, t.sd, cilm2, c2_d1, c2_d2, lmax2, nlat, 0, g2, nlong, hs);
    }
    if (!st) {
        hipLaunchKernelGGL(k_mul_inplace, dim3(grid_blocks((long long)ng)), dim3(kThreads), 0, hs, g1, g2,
                           (long long)ng);
        if (!HIP_OK(hipGetLastError())) st = ST_OTHER;
    }
    if (!st)
        st = plx ? launch_plx_analysis(*t.g, g1, nlong, lmaxout, plx, norm, cilmout, out_d1, out_d2, hs)
                 : launch_analysis(*t.g, t.rt, t.sd, g1, nlong, lmaxout, cilmout, out_d1, out_d2, hs);
    (void)hipFreeAsync(ws, hs);
    return st;
}

int shtg_dh_polar_orders(int lmax, int *kmax)
{
    if (lmax < 0 || lmax > kMaxArgDegree) return ST_BOUNDS;
    if (!kmax) return ST_DIMS;
    const int n = 2 * lmax + 2, P = n / 2 + 1;
    const double pi = std::acos(-1.0);
    std::vector<double> z(P), u(P);
    for (int s = 0; s < P; ++s) {
        z[s] = (s == n / 2) ? 0.0 : std::cos(pi * (double)s / (double)n);
        u[s] = (s == n / 2) ? 1.0 : std::sqrt((1.0 - z[s]) * (1.0 + z[s]));
    }
    const std::vector<int> km = polar_kmax(z, u, lmax);
    for (int s = 0; s < P; ++s) kmax[s] = km[s];
    return ST_OK;
}

// src/PreGLQ.f95:153-233: Gauss-Legendre points that integrate exactly a
// polynomial of the given degree, a field of degree `degree`, and a field of
// that degree raised to the power n.  The reference stops the program for a
// negative degree; these return -1 instead.
int shtg_nglq(int degree) { return degree < 0 ? -1 : (degree + 2) / 2; }
int shtg_nglqsh(int degree) { return degree < 0 ? -1 : degree + 1; }
int shtg_nglqshn(int degree, int n)
{
    if (degree < 0) return -1;
    const long long d = ((long long)n + 1) * degree + 1;  // ceiling(d / 2), d of either sign
    return (int)(d >= 0 ? (d + 1) / 2 : -((-d) / 2));
}

int shtg_shglq(int lmax, double *zero, double *w)
{
    if (lmax < 0 || lmax > kMaxArgDegree) return ST_BOUNDS;
    if (!zero || !w) return ST_DIMS;
    return glq_nodes(lmax + 1, zero, w);
}

int shtg_dhaj(int n, int extend, double *aj)
{
    if (extend != 0 && extend != 1) return ST_BOUNDS;
    if (n % 2 != 0 || n < 0) return ST_BOUNDS;
    if (!aj) return ST_DIMS;
    std::vector<double> v;
    dh_weights(n, v);
    for (int j = 0; j < n; ++j) aj[j] = v[j];
    if (extend) aj[n] = 0.0;
    return ST_OK;
}

int shtg_glqgridcoord(int lmax, int extend, double *lat, double *lon)
{
    if (lmax < 0 || lmax > kMaxArgDegree) return ST_BOUNDS;
    if (extend != 0 && extend != 1) return ST_BOUNDS;
    if (!lat || !lon) return ST_DIMS;
    const int nlat = lmax + 1, nlong = 2 * lmax + 1;
    std::vector<double> zero(nlat), w(nlat);
    int st = glq_nodes(nlat, zero.data(), w.data());
    if (st) return st;
    const double pi = std::acos(-1.0);
    for (int i = 0; i < nlong + extend; ++i) lon[i] = 360.0 * (double)i / (double)nlong;
    for (int i = 0; i < nlat; ++i) lat[i] = std::asin(zero[i]) * 180.0 / pi;
    return ST_OK;
}

// P_lm(z) for all l, m <= lmax into p[l*(lmax+1)+m], evaluated on the host
// with exactly the tables, seeds, rescaled recursion and extended-range
// powers that the GPU kernels use (the reference's PlmBar/PlmSchmidt/
// PLegendreA/PlmON family for the transforms' normalisations).
int shtg_plm(int lmax, int norm, int csphase, double z, double *p)
{
    if (lmax < 0 || lmax > kMaxArgDegree || norm < 1 || norm > 4 || (csphase != 1 && csphase != -1))
        return ST_BOUNDS;
    if (!p) return ST_DIMS;
    RecurrenceTables t = build_recurrence(norm, lmax);
    std::vector<double> seed = build_seeds(norm, csphase, lmax);
    const double u = std::sqrt((1.0 - z) * (1.0 + z));
    for (size_t i = 0; i < (size_t)(lmax + 1) * (lmax + 1); ++i) p[i] = 0.0;
    for (int m = 0; m <= lmax; ++m) {
        double r = 1.0;
        if (m > 0) {
            if (u == 0.0) r = 0.0;
            else {
                int e;
                double f = std::frexp(u, &e), rr = 1.0, b = f;
                int re = 0, be = 0, k = m;
                while (k) {
                    if (k & 1) { int q; rr = std::frexp(rr * b, &q); re += be + q; }
                    k >>= 1;
                    if (k) { int q; b = std::frexp(b * b, &q); be = 2 * be + q; }
                }
                r = std::ldexp(rr * 1.0e280, re + e * m);
            }
        }
        const double zz = (r == 0.0) ? 0.0 : z;
        const int64_t o = tri_off(m, lmax);
        double q1 = 0.0, q2 = -seed[m];
        for (int l = m; l <= lmax; ++l) {
            const double q = (t.alpha[o + l - m] * zz) * q1 - q2;
            q2 = q1;
            q1 = q;
            p[(size_t)l * (lmax + 1) + m] = (r == 0.0) ? 0.0 : t.s[o + l - m] * q * r;
        }
    }
    return ST_OK;
}

int shtg_device_count(void)
{
    int n = 0;
    if (!HIP_OK(hipGetDeviceCount(&n))) return 0;
    return n;
}

const char *shtg_version(void) { return "shtools_amd 0.1 (gfx950)"; }

int shtg_batch_create_dh(int nfields, const int *lmax, int sampling, int norm, int csphase,
                         shtg_batch_plan **plan)
{
    if (nfields <= 0) {
        if (plan) *plan = nullptr;
        return ST_DIMS;
    }
    const std::vector<int> samp((size_t)nfields, sampling);
    return shtg_batch_create_dh_fields(nfields, lmax, samp.data(), norm, csphase, plan);
}

int shtg_batch_create_dh_fields(int nfields, const int *lmax, const int *sampling, int norm, int csphase,
                                shtg_batch_plan **plan)
{
    if (plan) *plan = nullptr;
    if (nfields <= 0 || !sampling) return ST_DIMS;
    for (int i = 0; i < nfields; ++i)
        if (sampling[i] != 1 && sampling[i] != 2) return ST_BOUNDS;
    // sampling 1 and 2 are the grid codes kGridDH1 and kGridDH2
    return shtg_batch_create_fields(nfields, lmax, sampling, norm, csphase, plan);
}

int shtg_batch_create_glq(int nfields, const int *lmax, int norm, int csphase, shtg_batch_plan **plan)
{
    if (nfields <= 0) {
        if (plan) *plan = nullptr;
        return ST_DIMS;
    }
    const std::vector<int> grid((size_t)nfields, (int)kGridGLQ);
    return shtg_batch_create_fields(nfields, lmax, grid.data(), norm, csphase, plan);
}

int shtg_batch_create_fields(int nfields, const int *lmax, const int *grid, int norm, int csphase,
                             shtg_batch_plan **plan)
{
    std::lock_guard<std::mutex> lk(g_mu);
    BatchPlan *b = nullptr;
    int st = batch_create(nfields, lmax, grid, norm, csphase, &b);
    if (plan) *plan = reinterpret_cast<shtg_batch_plan *>(b);
    return st;
}

int shtg_batch_destroy(shtg_batch_plan *plan)
{
    // no call on the plan may be in progress (the caller's responsibility)
    batch_free(reinterpret_cast<BatchPlan *>(plan));
    return ST_OK;
}

int shtg_batch_makegrid_dh(shtg_batch_plan *plan, const double *const *cilm, double *const *grid, void *stream)
{
    if (!plan || !cilm || !grid) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    int st = batch_bind(b, 0, (const void *const *)cilm, (hipStream_t)stream);
    if (!st) st = batch_bind(b, 1, (const void *const *)grid, (hipStream_t)stream);
    if (st) return st;
    return batch_synthesis(b, (hipStream_t)stream, b->T);
}

int shtg_batch_shexpand_dh(shtg_batch_plan *plan, const double *const *grid, double *const *cilm, void *stream)
{
    if (!plan || !cilm || !grid) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    int st = batch_bind(b, 2, (const void *const *)grid, (hipStream_t)stream);
    if (!st) st = batch_bind(b, 3, (const void *const *)cilm, (hipStream_t)stream);
    if (st) return st;
    return batch_analysis(b, (hipStream_t)stream, b->T);
}

int shtg_batch_makegrid_dh_adjoint(shtg_batch_plan *plan, const double *const *grid_bar,
                                   double *const *cilm_bar, void *stream)
{
    if (!plan || !cilm_bar || !grid_bar) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    int st = batch_bind(b, 2, (const void *const *)grid_bar, (hipStream_t)stream);
    if (!st) st = batch_bind(b, 3, (const void *const *)cilm_bar, (hipStream_t)stream);
    if (st) return st;
    BTables T = b->T;
    T.w = b->ones;
    T.s0 = 0;
    T.post = T.s;
    return batch_analysis(b, (hipStream_t)stream, T);
}

int shtg_batch_shexpand_dh_adjoint(shtg_batch_plan *plan, const double *const *cilm_bar,
                                   double *const *grid_bar, void *stream)
{
    if (!plan || !cilm_bar || !grid_bar) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    int st = batch_bind(b, 0, (const void *const *)cilm_bar, (hipStream_t)stream);
    if (!st) st = batch_bind(b, 1, (const void *const *)grid_bar, (hipStream_t)stream);
    if (st) return st;
    BTables T = b->T;
    T.s = T.post;
    T.roww = b->w;
    return batch_synthesis(b, (hipStream_t)stream, T);
}

int shtg_batch_tune(shtg_batch_plan *plan, int reps, int flags, void *stream)
{
    if (!plan || reps < 1 || (flags & ~1)) return ST_BOUNDS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    return batch_tune(b, (hipStream_t)stream, reps, flags);
}

int shtg_batch_get_shapes(shtg_batch_plan *plan, int *threads_synth, int *threads_anal, int n)
{
    if (!plan) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    const int ng = (int)b->fft.size();
    for (int i = 0; i < n && i < ng; ++i) {
        if (threads_synth) threads_synth[i] = b->fft[i].threads;
        if (threads_anal) threads_anal[i] = b->fft[i].threads_anal;
    }
    return ng;
}

int shtg_batch_set_mfma(shtg_batch_plan *plan, int min_bucket, void *stream)
{
    if (!plan || min_bucket < 0 || min_bucket == 1) return ST_BOUNDS;
    (void)stream;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    return batch_synth_items(b, min_bucket);
}

int shtg_batch_get_tuning(shtg_batch_plan *plan, int *synth_kmax, int *anal_kmax, int *mfma_min)
{
    if (!plan) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    if (synth_kmax) *synth_kmax = b->synth_kmax;
    if (anal_kmax) *anal_kmax = b->anal_kmax;
    if (mfma_min) *mfma_min = b->mfma_min;
    return ST_OK;
}

int shtg_batch_set_tuning(shtg_batch_plan *plan, int synth_kmax, int anal_kmax, int mfma_min)
{
    if (!plan) return ST_DIMS;
    if (synth_kmax < 1 || synth_kmax > kSynthKMax || anal_kmax < 1 || anal_kmax > kBatchAnalKMax || mfma_min < 0 ||
        mfma_min == 1)
        return ST_BOUNDS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    b->synth_kmax = synth_kmax;
    b->anal_kmax = anal_kmax;
    int st = batch_synth_items(b, mfma_min);
    if (!st) st = batch_anal_items(b);
    return st;
}

int shtg_batch_set_shapes(shtg_batch_plan *plan, const int *threads_synth, const int *threads_anal, int n)
{
    if (!plan || !threads_synth || !threads_anal) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    if (n != (int)b->fft.size()) return ST_DIMS;
    for (int i = 0; i < n; ++i)
        for (int t : {threads_synth[i], threads_anal[i]})
            if (!valid_fft_shape(t, b->fft[i].min_threads, b->fft[i].max_threads)) return ST_BOUNDS;
    for (int i = 0; i < n; ++i) {
        b->fft[i].threads = threads_synth[i];
        b->fft[i].threads_anal = threads_anal[i];
    }
    return ST_OK;
}

int shtg_batch_field_groups(shtg_batch_plan *plan, int *group, int n)
{
    if (!plan || !group) return ST_DIMS;
    BatchPlan *b = reinterpret_cast<BatchPlan *>(plan);
    std::lock_guard<std::mutex> lk(b->mu);
    if (n < b->nf) return ST_DIMS;
    for (int i = 0; i < b->nf; ++i) group[i] = b->fft_group_of[i];
    return (int)b->fft.size() + (b->fft_ws.nitems ? 1 : 0);
}

int shtg_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof = on != 0;
    return ST_OK;
}

int shtg_profile_collect(double *ms, long long *count, int nk)
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    for (int k = 0; k < nk; ++k) { ms[k] = 0.0; count[k] = 0; }
    int st = ST_OK;
    for (auto &r : g_prof_recs) {
        float t = 0.0f;
        if (!HIP_OK(hipEventSynchronize(r.b)) || !HIP_OK(hipEventElapsedTime(&t, r.a, r.b))) st = ST_OTHER;
        if (r.kid >= 0 && r.kid < nk) { ms[r.kid] += t; count[r.kid] += 1; }
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_recs.clear();
    return st;
}

}  // extern "C"
