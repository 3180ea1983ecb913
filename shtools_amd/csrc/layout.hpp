// Host-side layout of the transforms' launches: launch constants, the LDS
// classes of ring FFTs, and the batched plans' geometry tables, field
// offsets and work-item lists.  Plain C++ (no HIP runtime calls), so that it
// is also built and checked on the host alone, under AddressSanitizer.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

#include "fftplan.hpp"
#include "tables.hpp"

namespace shtg {

// ------------------------------------------------------------ launch constants

constexpr int kThreads = 256;  // generic workgroups
// Ring-FFT workgroups have 256, 512 or 1024 threads.  The 1024-thread bound
// caps the kernels at 128 VGPRs, i.e. up to 4 waves per SIMD: the ring FFTs
// are latency-bound, so the stage radices stay small enough (up to 8) to
// keep that occupancy.
constexpr int kFftMaxThreads = 1024;
constexpr int kWave = 64;      // Legendre workgroups are single waves
constexpr int kSynthKMax = 4;  // ring pairs per lane in synthesis (1 .. 4)
constexpr int kAnalKMax = 8;   // ring pairs per lane and pass in analysis (1 .. 8)
// Batched analysis waves go up to 12 ring pairs per lane (246 VGPRs, still
// 2 waves per SIMD like 8): one pass over up to 768 slots instead of two,
// one cross-lane reduction per chunk of degrees instead of two.
constexpr int kBatchAnalKMax = 12;
constexpr int kAnalLC = 8;     // degrees per reduction chunk in analysis
// Coefficient tiles of kTL degrees x kTM orders (pack / unpack).
constexpr int kTM = 64, kTL = 32;
constexpr int kBucketMaxCols = 64;  // up to 32 fields per MFMA bucket wave (4 column tiles)
constexpr int kMfmaSlots = 16;      // slots per MFMA synthesis wave

// Largest LDS footprint a ring-FFT workgroup may use (gfx950: 160 KiB); ring
// FFTs whose work array is larger run on a global workspace of kWsGroups
// slices.
constexpr size_t kMaxLds = 160 * 1024;
constexpr int kWsGroups = 512;
constexpr size_t kComplexBytes = 16;  // one double2 element of a work array

inline bool fft_in_lds(int m) { return (size_t)fft_lds_slots(m) * kComplexBytes <= kMaxLds; }
inline long long ws_stride(int m) { return ((long long)fft_lds_slots(m) + 15) / 16 * 16; }

// Workgroups of a ring FFT with work length m that share a CU's LDS (0: the
// array does not fit, the ring runs on the global workspace).
inline int lds_per_cu(int m)
{
    if (!fft_in_lds(m)) return 0;
    return std::min(4, (int)(kMaxLds / ((size_t)fft_lds_slots(m) * kComplexBytes)));
}

inline int env_int(const char *name, int dflt)
{
    const char *v = std::getenv(name);
    if (!v || !*v) return dflt;
    int x = std::atoi(v);
    return x > 0 ? x : dflt;
}

// Whether a ring of length n with band limit lc runs the split Bluestein
// plan (make_fft_plan): when its whole plan fits fewer than two workgroups
// per CU (one workgroup per CU, or the global workspace beyond the LDS) and
// the split plan fits the LDS, for at most 15% more transform work (3
// transforms of the split length against 2 of the whole one).  Plain and
// batched transforms decide alike, so they stay bitwise equal.
//
// Splitting the one-workgroup-per-CU class (rings of DH2 fields above degree
// ~ 850) gives two workgroups per CU: measured on the 512-field benchmark,
// tuned, synthesis ring FFTs 13.3 -> 12.9 ms and analysis 12.4 -> 12.0 ms
// (whole step 48.9-49.4 -> 47.8 ms).  Rings that fit two or more workgroups
// per CU keep the whole plan: the split array is about 2/3 as long but has as
// many stages, and its second inverse pass adds more LDS traffic than the
// extra workgroups per CU recover (also splitting the two-workgroup class:
// 13.6 and 12.8 ms; splitting every in-LDS ring: synthesis 13.9 -> 17.1 ms).
// SHTG_SPLIT_BELOW = k splits every ring whose whole plan fits fewer than k
// workgroups per CU (default 2; 1: only rings beyond the LDS); it changes
// results by rounding.
inline bool ring_split(int n, int lc)
{
    const int mh = split_length(n, lc);
    if (mh <= 0 || !fft_in_lds(mh)) return false;
    const int mf = bluestein_length(n + 2 * lc);
    if (3.0 * mh > 1.15 * 2.0 * mf) return false;
    static const int below = env_int("SHTG_SPLIT_BELOW", 2);
    return lds_per_cu(mf) < below;
}

inline FftPlanHost ring_plan(int n, int lc) { return make_fft_plan(n, lc, lc >= 0 && ring_split(n, lc)); }

// Tuning of the ring-FFT launches: threads per workgroup by work length and
// ring pairs per batched workgroup.  SHTG_FFT_MAXTHREADS (256/512/1024) and
// SHTG_FFT_SLOTS override the defaults.

inline int fft_threads(int m)
{
    static const int cap = std::min(kFftMaxThreads, env_int("SHTG_FFT_MAXTHREADS", kFftMaxThreads));
    int t = 256;
    if (m > 1024) t = 512;
    if (m > 4096) t = 1024;
    return std::min(t, cap);
}

// Threads per workgroup of a plan's ring FFTs (P: FftPlanHost or the device
// plan): fft_threads(m), and for a split plan at least as many as give every
// thread at most kTurnStash / R turn butterflies (R the last radix), and 512
// for the largest arrays (two such workgroups share a CU, each with up to
// 128 registers per thread).
template <class P>
inline int plan_min_threads(const P &F)
{
    if (!F.half) return 1;
    const int R = F.radix[F.nst - 1], nb = kTurnStash / R;
    const int need = (F.m / R + nb - 1) / nb;
    return (need + kWave - 1) / kWave * kWave;
}

// Largest workgroup of the kernel instantiation a plan runs in (512 for
// prime-radix plans, whose kernels are built for up to 256 registers per
// thread).
template <class P>
inline int plan_max_threads(const P &F)
{
    return plan_variant(F) == kVarPrime ? kFftMaxThreads / 2 : kFftMaxThreads;
}

template <class P>
inline int plan_threads(const P &F)
{
    int t = fft_threads(F.m);
    if (F.half && F.m > 4096) t = std::min(t, kFftMaxThreads / 2);
    t = std::min(t, plan_max_threads(F));
    return std::max(t, plan_min_threads(F));
}

inline int fft_slots_per_wg()
{
    static const int v = std::max(1, std::min(16, env_int("SHTG_FFT_SLOTS", 2)));
    return v;
}

// Ring pairs per lane of an analysis wave for P slots (1 .. kmax): the
// fewest passes of at most kmax * 64 slots, with lanes packed as tightly as
// those passes allow.
inline int anal_lanes_per_slot(int P, int kmax = kAnalKMax)
{
    const int per = kmax * kWave;
    const int np = std::max(1, (P + per - 1) / per);
    return std::max(1, (P + np * kWave - 1) / (np * kWave));
}

// ------------------------------------------------------------ batched plans

// Grids of batched fields: DH with sampling 1 or 2 (src/MakeGridDH.f95) or
// Gauss-Legendre (src/MakeGridGLQ.f95, nodes of src/PreGLQ.f95).
enum { kGridDH1 = 1, kGridDH2 = 2, kGridGLQ = 3 };

// Per-field description of a batch (all fields: lmax_calc = lmax), as the
// batched kernels read it.
struct BField {
    int L;           // degree (grid and expansion)
    int P;           // ring-pair slots
    int nlat;        // grid rows
    int nlong;       // grid columns (FFT length)
    int geo;         // offset of the field's slots in the concatenated ring tables
    int fft;         // index of the field's FFT plan
    long long coff;  // offset of the field in cpack/cbuf (double2 units)
    long long foff;  // offset of the field in fbuf/gbuf (double4 units)
    int sm;          // offset of the field's per-order first slots in the smin table
    int s0;          // first slot with an analysis weight (DH: 1, the pole carries none; GLQ: 0)
};

// Work items as the kernels read them (int4 / int2 on the device).
struct Item4 {
    int x, y, z, w;
};
struct Item2 {
    int x, y;
};

// Ring geometry of every distinct (degree, grid) of a plan, the fields'
// offsets into the plan's workspaces, and the coefficient-tile items.
struct BatchLayout {
    std::vector<double> z, u, w;  // per slot: cos and sin of the northern ring, weights (north, south)
    std::vector<int> rows;        // per slot: (north row, south row or -1)
    std::vector<int> kmax;        // per slot: largest order with non-negligible functions
    std::vector<int> smin;        // per geometry and order: first slot it needs
    std::vector<FftPlanHost> plans;
    std::vector<BField> fields;
    long long ncpack = 0;         // double2 elements of cpack/cbuf
    long long nfbuf = 0;          // double4 elements of fbuf/gbuf
    std::vector<Item4> pack;      // (field, m0, l0) tiles on or below the diagonal, longest first
    std::vector<Item4> unpack;    // every (field, m0, l0) tile
};

struct WorkItem {
    double w;
    int a, b, c;
};
inline bool by_work(const WorkItem &x, const WorkItem &y) { return x.w > y.w; }

// Rows, slots and ring length of a field of degree L on grid `grid`.
struct GridShape {
    int nlat, nlong, P;
};
inline GridShape grid_shape(int L, int grid)
{
    if (grid == kGridGLQ) return {L + 1, 2 * L + 1, (L + 2) / 2};
    const int n = 2 * L + 2;
    return {n, (grid == kGridDH1 ? 1 : 2) * n, n / 2 + 1};
}

// Appends the ring slots of one geometry (z, u, rows, weights) to out.
//
// DH (src/MakeGridDH.f95:504-510, src/SHExpandDH.f95:287-294,463-479): slot
// s = 0 .. n/2 of n = 2L+2 rows pairs row s (theta = pi s / n) with row
// n - s; slot 0 is the north pole with the south pole (row n, present only
// with extend, so its weight is 0), slot n/2 the equator (unpaired).
// Weights sqrt(2 pi) sqrt(4 pi) aj / nlong.
//
// GLQ (src/MakeGridGLQ.f95:467-480, src/SHExpandGLQ.f95:535-547): slot s
// pairs row s (z = zero(s), the nodes of glq_nodes in descending order) with
// row nlat-1-s; for odd nlat the middle row is an exact equator.  Weights
// 2 pi w / nlong.  Returns false when the nodes do not converge.
inline bool append_geometry(int L, int grid, BatchLayout &out)
{
    const double pi = std::acos(-1.0);
    const GridShape g = grid_shape(L, grid);
    if (grid == kGridGLQ) {
        std::vector<double> zero(g.nlat), wq(g.nlat);
        if (glq_nodes(g.nlat, zero.data(), wq.data()) != 0) return false;
        for (int s = 0; s < g.P; ++s) {
            const bool eq = (g.nlat % 2 != 0) && (s == g.P - 1);
            const double zz = eq ? 0.0 : zero[s];
            out.z.push_back(zz);
            out.u.push_back(eq ? 1.0 : std::sqrt((1.0 - zz) * (1.0 + zz)));
            out.rows.push_back(s);
            out.rows.push_back(eq ? -1 : g.nlat - 1 - s);
            out.w.push_back(2.0 * pi * wq[s] / (double)g.nlong);
            out.w.push_back(eq ? 0.0 : 2.0 * pi * wq[g.nlat - 1 - s] / (double)g.nlong);
        }
        return true;
    }
    const int n = g.nlat;
    std::vector<double> aj;
    dh_weights(n, aj);
    const double c = std::sqrt(2 * pi) * std::sqrt(4.0 * pi) / (double)g.nlong;
    for (int s = 0; s < g.P; ++s) {
        const double zz = (s == n / 2) ? 0.0 : std::cos(pi * (double)s / (double)n);
        out.z.push_back(zz);
        out.u.push_back((s == n / 2) ? 1.0 : std::sqrt((1.0 - zz) * (1.0 + zz)));
        out.rows.push_back(s);
        out.rows.push_back((s == n / 2) ? -1 : n - s);
        out.w.push_back(c * aj[s]);
        out.w.push_back((s == 0 || s == n / 2) ? 0.0 : c * aj[n - s]);
    }
    return true;
}

// The layout of a plan of nfields fields of degrees lmax[i] on grids
// grid[i] (kGridDH1, kGridDH2 or kGridGLQ; both already checked).  Every
// distinct (degree, grid) gets one geometry and one ring-FFT plan, built as
// the plain transforms build theirs (pruned for DH rings, full for GLQ
// rings), so that batched and plain results agree.  Returns false when a
// GLQ geometry cannot be built.
inline bool build_batch_layout(int nfields, const int *lmax, const int *grid, BatchLayout &out)
{
    typedef std::pair<int, int> Key;
    std::map<Key, std::pair<int, int>> geo_of;  // -> (geo offset, plan index)
    std::map<Key, int> sm_of;                   // -> offset of its smin row
    for (int i = 0; i < nfields; ++i) {
        const int L = lmax[i];
        const Key key(L, grid[i]);
        if (geo_of.count(key)) continue;
        const int off = (int)out.z.size();
        if (!append_geometry(L, grid[i], out)) return false;
        const std::vector<double> zl(out.z.begin() + off, out.z.end()), ul(out.u.begin() + off, out.u.end());
        const std::vector<int> km = polar_kmax(zl, ul, L);
        out.kmax.insert(out.kmax.end(), km.begin(), km.end());
        const std::vector<int> sm = polar_smin(km, L);
        sm_of[key] = (int)out.smin.size();
        out.smin.insert(out.smin.end(), sm.begin(), sm.end());
        const GridShape g = grid_shape(L, grid[i]);
        out.plans.push_back(grid[i] == kGridGLQ ? make_fft_plan(g.nlong, -1) : ring_plan(g.nlong, L));
        geo_of[key] = {off, (int)out.plans.size() - 1};
    }
    out.fields.assign(nfields, BField{});
    long long coff = 0, foff = 0;
    for (int i = 0; i < nfields; ++i) {
        const int L = lmax[i];
        const Key key(L, grid[i]);
        const GridShape g = grid_shape(L, grid[i]);
        BField &f = out.fields[i];
        f.L = L;
        f.P = g.P;
        f.nlat = g.nlat;
        f.nlong = g.nlong;
        f.geo = geo_of[key].first;
        f.fft = geo_of[key].second;
        f.coff = coff;
        f.foff = foff;
        f.sm = sm_of[key];
        f.s0 = grid[i] == kGridGLQ ? 0 : 1;
        coff += tri_size(L);
        foff += (long long)(L + 1) * f.P;
    }
    out.ncpack = coff;
    out.nfbuf = foff;
    std::vector<WorkItem> wp, wu;
    for (int i = 0; i < nfields; ++i) {
        const BField &f = out.fields[i];
        for (int m0 = 0; m0 <= f.L; m0 += kTM)
            for (int l0 = 0; l0 <= f.L; l0 += kTL) {
                wu.push_back({(double)f.L, i, m0, l0});
                if (l0 + kTL - 1 >= m0) wp.push_back({(double)f.L, i, m0, l0});
            }
    }
    std::stable_sort(wp.begin(), wp.end(), by_work);
    out.pack.clear();
    out.unpack.clear();
    for (auto &x : wp) out.pack.push_back({x.a, x.b, x.c, 0});
    for (auto &x : wu) out.unpack.push_back({x.a, x.b, x.c, 0});
    return true;
}

// Synthesis Legendre work items, longest first: buckets of at least
// mfma_min fields of one degree and sampling (mfma_min >= 2; 0: none), in
// groups of up to kBucketMaxCols / 2, as (bucket, m, kMfmaSlots-slot tile)
// items of the MFMA kernel with their field lists (kBucketMaxCols / 2 entries
// per bucket, padded with the first field) and counts; every other field as
// (field, m, first slot, ring pairs per lane) items of the VALU kernel, in
// blocks of synth_kmax * 64 slots from the order's first slot with
// non-negligible functions (smin), the last block packing its lanes tightly.
struct SynthItems {
    std::vector<Item4> valu, mfma;
    std::vector<int> bfields, bnf;
};

inline SynthItems build_synth_items(const std::vector<BField> &fields, const std::vector<int> &smin, int synth_kmax,
                                    int mfma_min)
{
    const int nfields = (int)fields.size();
    std::map<std::pair<int, int>, std::vector<int>> groups;  // (L, nlong) -> fields
    for (int i = 0; i < nfields; ++i) groups[{fields[i].L, fields[i].nlong}].push_back(i);
    std::vector<int> in_bucket(nfields, -1);
    SynthItems out;
    if (mfma_min >= 2)
        for (auto &kv : groups) {
            const std::vector<int> &mem = kv.second;
            for (size_t k0 = 0; k0 + (size_t)mfma_min <= mem.size(); k0 += kBucketMaxCols / 2) {
                const int cnt = (int)std::min(mem.size() - k0, (size_t)(kBucketMaxCols / 2));
                if (cnt < mfma_min) break;
                const int id = (int)out.bnf.size();
                out.bnf.push_back(cnt);
                for (int k = 0; k < kBucketMaxCols / 2; ++k) out.bfields.push_back(k < cnt ? mem[k0 + k] : mem[k0]);
                for (int k = 0; k < cnt; ++k) in_bucket[mem[k0 + k]] = id;
            }
        }
    std::vector<WorkItem> ws, wm;
    for (int i = 0; i < nfields; ++i) {
        const BField &f = fields[i];
        const int *sm = smin.data() + f.sm;
        if (in_bucket[i] >= 0) {
            const int id = in_bucket[i];
            if (out.bfields[(size_t)id * (kBucketMaxCols / 2)] != i) continue;  // the bucket's first field
            for (int m = 0; m <= f.L; ++m)
                for (int s0 = sm[m] / 16 * 16; s0 < f.P; s0 += kMfmaSlots)
                    wm.push_back({(double)(f.L - m + 1) * out.bnf[id], id, m, s0});
            continue;
        }
        for (int m = 0; m <= f.L; ++m)
            for (int s0 = sm[m]; s0 < f.P; s0 += synth_kmax * kWave) {
                const int kk = std::min(synth_kmax, (f.P - s0 + kWave - 1) / kWave);
                ws.push_back({(double)(f.L - m + 1) * std::min(kk * kWave, f.P - s0), i, m, s0 + (kk << 24)});
            }
    }
    std::stable_sort(ws.begin(), ws.end(), by_work);
    std::stable_sort(wm.begin(), wm.end(), by_work);
    for (auto &x : ws) out.valu.push_back({x.a, x.b, x.c & 0xFFFFFF, x.c >> 24});
    for (auto &x : wm) out.mfma.push_back({x.a, x.b, x.c, 0});
    return out;
}

// Analysis Legendre work items, longest first: one single-wave item (field,
// m) per order, grouped by ring pairs per lane (at most anal_kmax) into one
// launch each (out[K - 1]).  The sums start at slot max(s0, smin): the DH
// pole slot carries no analysis weight.
inline void build_anal_items(const std::vector<BField> &fields, const std::vector<int> &smin, int anal_kmax,
                             std::vector<Item2> (&out)[kBatchAnalKMax])
{
    std::vector<WorkItem> wa;
    for (int i = 0; i < (int)fields.size(); ++i) {
        const BField &f = fields[i];
        const int *sm = smin.data() + f.sm;
        for (int m = 0; m <= f.L; ++m) {
            const int sa = std::max(f.s0, sm[m]);
            wa.push_back({(double)(f.L - m + 1) * (f.P - sa), i, m, anal_lanes_per_slot(f.P - sa, anal_kmax)});
        }
    }
    std::stable_sort(wa.begin(), wa.end(), by_work);
    for (auto &v : out) v.clear();
    for (auto &x : wa) out[x.c - 1].push_back({x.a, x.b});
}

// Ring-FFT work items: (field, first slot) of fft_slots consecutive ring
// pairs, grouped by how many workgroups fit the LDS of a CU (the group's LDS
// is its largest need), then by threads per workgroup and plan variant
// (kVarPlain, kVarSplit, kVarPrime: one kernel instantiation each); fields
// whose rings exceed the LDS form the workspace group.
struct FftGroupLayout {
    std::vector<Item2> items;
    size_t lds = 0;
    int threads = 256;
    int min_threads = 1;
    int max_threads = 1024;  // the instantiation's largest workgroup
    int variant = kVarPlain;
};
struct FftLayout {
    std::vector<FftGroupLayout> groups;
    std::vector<Item2> ws_items;
    long long ws_stride = 0;
    std::vector<int> group_of;  // per field: its group (groups.size(): the workspace group)
};

inline FftLayout build_fft_groups(const std::vector<BField> &fields, const std::vector<FftPlanHost> &plans,
                                  int fft_slots)
{
    typedef std::tuple<int, int, int> GKey;  // (-(WGs per CU), -threads, plan variant)
    std::map<GKey, std::vector<WorkItem>> groups;
    std::map<GKey, size_t> group_lds;
    std::map<GKey, int> group_min;
    std::vector<WorkItem> wsi;
    FftLayout out;
    for (int i = 0; i < (int)fields.size(); ++i) {
        const BField &f = fields[i];
        const FftPlanHost &p = plans[f.fft];
        const size_t need = (size_t)fft_lds_slots(p.m) * kComplexBytes;
        if (!fft_in_lds(p.m)) {
            for (int s = 0; s < f.P; s += fft_slots) wsi.push_back({(double)need, i, s, 0});
            out.ws_stride = std::max(out.ws_stride, ws_stride(p.m));
            continue;
        }
        const int per_cu = std::max(1, (int)(kMaxLds / need));
        const GKey key(-std::min(per_cu, 4), -plan_threads(p), plan_variant(p));
        group_lds[key] = std::max(group_lds[key], need);
        group_min[key] = std::max(group_min[key], plan_min_threads(p));
        for (int s = 0; s < f.P; s += fft_slots) groups[key].push_back({(double)need, i, s, 0});
    }
    out.group_of.assign(fields.size(), (int)groups.size());
    for (auto &kv : groups) {
        for (auto &x : kv.second) out.group_of[x.a] = (int)out.groups.size();
        std::stable_sort(kv.second.begin(), kv.second.end(), by_work);
        FftGroupLayout g;
        for (auto &x : kv.second) g.items.push_back({x.a, x.b});
        g.lds = group_lds[kv.first];
        g.threads = -std::get<1>(kv.first);
        g.min_threads = group_min[kv.first];
        g.variant = std::get<2>(kv.first);
        g.max_threads = g.variant == kVarPrime ? kFftMaxThreads / 2 : kFftMaxThreads;
        out.groups.push_back(std::move(g));
    }
    std::stable_sort(wsi.begin(), wsi.end(), by_work);
    for (auto &x : wsi) out.ws_items.push_back({x.a, x.b});
    return out;
}

}  // namespace shtg
