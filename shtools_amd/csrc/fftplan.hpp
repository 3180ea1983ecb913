// Host planning of the per-ring FFTs run in LDS by the GPU kernels.
//
// Every ring (pair) transform is a complex DFT of length n, forward
// X_k = sum_j x_j e^{-2 pi i jk/n}; inverse transforms are taken as
// conj(F(conj(.))).  When n has no prime factor above 7, apart from at most
// one of the prime radices 11 .. 23 (kPrimeRadices), the kernel runs
// in-place decimation-in-frequency stages (radices up to 8 and that prime,
// see direct_stages) on the length-n array and reads results through a
// digit-reversal table.  Otherwise the DFT is computed by Bluestein's chirp-z
// method on a 5-smooth work length m >= 2n-1 (bluestein_length):
// the chirp-modulated input is transformed by DIF, multiplied by the
// precomputed spectrum of the conjugate chirp (stored in DIF output order,
// scaled by 1/m) and transformed back by the transposed (DIT) stages, which
// consume digit-reversed input and leave natural order.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <functional>
#include <iterator>
#include <map>
#include <vector>

namespace shtg {

constexpr int kMaxStages = 24;

// The radices the device stages unroll in registers; radix 7 is the one
// rolled radix.  Larger register radices cut LDS passes but raise the
// kernels' register count above 128, which halves the waves per SIMD of
// these latency-bound kernels and measured slower.  X(R) is applied to each.
#define SHTG_FFT_RADICES(X) X(2) X(3) X(4) X(5) X(8)
constexpr int kRegRadices[] = {2, 3, 4, 5, 8};

// LDS (or workspace) elements of a ring FFT's work array of length m: the
// array is unpadded (see fpad in sht_device.hpp).
constexpr int fft_lds_slots(int m) { return m; }

// Values a thread of a split plan keeps between its two turns (see
// sht_device.hpp): up to kTurnStash / R turn butterflies of R values.
constexpr int kTurnStash = 10;

struct FftPlanHost {
    int n = 0, m = 0, nst = 0, blue = 0;
    int lc = -1;                // pruning: orders |k| <= lc only (-1: none)
    int half = 0;               // split plan: the ring's columns in two halves of n / 2 (see make_fft_plan)
    int radix[kMaxStages] = {0};
    int stride[kMaxStages] = {0};
    int twoff[kMaxStages] = {0};  // offset of each stage's twiddles in twst
    std::vector<double> tw;     // 2*m: e^{-2 pi i k/m}
    std::vector<double> twst;   // per stage: e^{-2 pi i i / (R S)}, i < S
    std::vector<int> perm;      // m: DIF output position of natural index k
    std::vector<double> chirp;  // 2*n (Bluestein)
    std::vector<double> bhat;   // 2*m (Bluestein), DIF output order, /m
    std::vector<double> bhat2;  // pruned plans: analysis spectrum (bhat is synthesis)
    std::vector<double> bhat3;  // split plans: synthesis spectrum of the second half
    std::vector<double> bhat4;  // split plans: analysis spectrum of the second half
};

// Prime radices a direct plan may end with (one of them, once): a ring
// whose length has such a prime factor runs a direct DFT of that length in
// its last stage (stride 1, no twiddles) instead of Bluestein's three
// transforms of about 1.5 times its length.  Only rings whose array fits
// the LDS (n <= kPrimeMaxLen) take such plans: they run in kernel
// instantiations of their own (kVarPrime), which the workspace kernels for
// longer rings do not have.  X(R) is applied to each.
#define SHTG_PRIME_RADICES(X) X(11) X(13) X(17) X(19) X(23)
#define SHTG_LIST_ITEM(R) R,
constexpr int kPrimeRadices[] = {SHTG_PRIME_RADICES(SHTG_LIST_ITEM)};
#undef SHTG_LIST_ITEM
constexpr int kPrimeMaxLen = 160 * 1024 / 16;

// Kernel instantiations by plan kind: plain plans, split Bluestein plans
// (FftPlanHost::half), direct plans ending with a prime radix, and
// Bluestein plans of a power-of-two length (their turn runs radix 8, 4 or 2).
enum { kVarPlain = 0, kVarSplit = 1, kVarPrime = 2, kVarPow2 = 3 };

inline bool is_prime_radix(int p)
{
    for (int q : kPrimeRadices)
        if (p == q) return true;
    return false;
}

// Fewest-stage factorization of m over the register radices (and radix 7
// when allow7).  Even radices run first, at the large strides; odd radices
// run last, so that the stride-1 stage has an odd radix whenever m has an odd
// factor (consecutive lanes then touch the LDS at odd multiples of 16 bytes,
// which spread over the banks).  Among fewest-stage factorizations one with
// an odd last radix is preferred.  Returns false when m has another prime.
inline bool factor_stages(int m, bool allow7, std::vector<int> &rad)
{
    rad.clear();
    if (m == 1) return true;
    std::vector<int> radices(std::begin(kRegRadices), std::end(kRegRadices));
    if (allow7) radices.push_back(7);
    // best[d]: fewest stages for the cofactor d of m (memo over divisors)
    std::map<int, int> best;
    std::function<int(int)> stages = [&](int r) -> int {
        if (r == 1) return 0;
        auto it = best.find(r);
        if (it != best.end()) return it->second;
        int b = 1 << 20;
        for (int R : radices)
            if (r % R == 0) b = std::min(b, 1 + stages(r / R));
        best[r] = b;
        return b;
    };
    const int k = stages(m);
    if (k >= (1 << 20)) return false;
    // walk one optimal factorization, taking an odd radix last if one fits
    std::vector<int> chosen;
    std::function<bool(int, int, bool)> walk = [&](int r, int left, bool want_odd_last) -> bool {
        if (r == 1) return left == 0;
        for (int R : radices) {
            if (r % R != 0 || stages(r / R) != left - 1) continue;
            if (want_odd_last && left == 1 && R % 2 == 0) continue;
            chosen.push_back(R);
            if (walk(r / R, left - 1, want_odd_last)) return true;
            chosen.pop_back();
        }
        return false;
    };
    if (!walk(m, k, m % 2 == 0 && (m & (m - 1)) != 0)) {
        chosen.clear();
        walk(m, k, false);
    }
    std::vector<int> even, odd;
    for (int R : chosen) (R % 2 == 0 ? even : odd).push_back(R);
    std::sort(even.rbegin(), even.rend());
    std::sort(odd.rbegin(), odd.rend());
    rad = even;
    rad.insert(rad.end(), odd.begin(), odd.end());
    return (int)rad.size() <= kMaxStages;
}

// Stages of a direct plan of length n: the register radices and radix 7,
// then (n <= kPrimeMaxLen) at most one prime radix of kPrimeRadices as the
// last stage.
inline bool direct_stages(int n, std::vector<int> &rad)
{
    if (factor_stages(n, true, rad)) return true;
    if (n > kPrimeMaxLen) return false;
    for (int p : kPrimeRadices) {
        if (n % p != 0 || (n / p) % p == 0) continue;
        if (!factor_stages(n / p, true, rad)) continue;
        rad.push_back(p);
        return (int)rad.size() <= kMaxStages;
    }
    rad.clear();
    return false;
}

inline bool is_5smooth(int m)
{
    for (int p : {2, 3, 5})
        while (m % p == 0) m /= p;
    return m == 1;
}

// Bluestein work length: the 5-smooth m >= lo (up to 1.35 lo) with the
// fewest LDS passes, m (2 stages - 1) (the last DIF and first DIT stage run
// together in registers).  The last radix is odd (3 or 5: m has an odd
// factor, which factor_stages orders last), except for powers of two when
// allow_pow2 (unpruned plans whose array fits the LDS: they run in kernel
// instantiations of their own, kVarPow2, whose turn is radix 8, 4 or 2; GLQ
// lmax 1023 rings, n = 2047: m = 4096 in 4 stages instead of 4608 in 5).
// Other even last radices are skipped (their stride-1 accesses collide on
// the LDS banks).
constexpr int kPow2MaxLen = 160 * 1024 / 16;
inline int bluestein_length(int lo, bool allow_pow2 = false)
{
    int best_m = 0;
    double best_cost = 0.0;
    std::vector<int> rad;
    for (int m = lo; m <= lo + lo / 3 + 8; ++m) {
        if (!is_5smooth(m) || !factor_stages(m, false, rad)) continue;
        const bool pow2 = (m & (m - 1)) == 0;
        if (pow2 ? !(allow_pow2 && m <= kPow2MaxLen) : rad.back() % 2 == 0) continue;
        const double cost = (double)m * (2.0 * rad.size() - 1.0);
        if (best_m == 0 || cost < best_cost) {
            best_m = m;
            best_cost = cost;
        }
    }
    return best_m;
}

// e^{-2 pi i k / m} evaluated in long double.
inline std::complex<long double> root_ld(long long k, long long m)
{
    k %= m;
    if (k < 0) k += m;
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)m;
    return {std::cos(a), std::sin(a)};
}

// Reference host DIF (same stage structure as the device) in long double;
// output left in DIF order.  Used to build the Bluestein spectrum.  The
// roots of unity come from one table of the m-th roots.
inline void host_dif(std::vector<std::complex<long double>> &x, const FftPlanHost &p)
{
    const int m = p.m;
    std::vector<std::complex<long double>> root(m);
    for (int k = 0; k < m; ++k) root[k] = root_ld(k, m);
    for (int st = 0; st < p.nst; ++st) {
        const int R = p.radix[st], S = p.stride[st], Lsub = R * S, step = m / Lsub, mr = m / R;
        std::vector<std::complex<long double>> v(R), y(R);
        for (int b = 0; b < m / R; ++b) {
            const int grp = b / S, i = b - grp * S, base = grp * Lsub + i;
            for (int r = 0; r < R; ++r) v[r] = x[base + r * S];
            for (int k = 0; k < R; ++k) {
                std::complex<long double> acc = 0;
                for (int j = 0; j < R; ++j) acc += v[j] * root[(size_t)((j * k) % R) * mr];
                y[k] = acc * root[(size_t)i * k * step % m];
            }
            for (int r = 0; r < R; ++r) x[base + r * S] = y[r];
        }
    }
}

inline void set_stages(FftPlanHost &p, const std::vector<int> &rad)
{
    p.nst = (int)rad.size();
    int Lsub = p.m;
    for (int s = 0; s < p.nst; ++s) {
        p.radix[s] = rad[s];
        p.stride[s] = Lsub / rad[s];
        Lsub = p.stride[s];
    }
    // digit reversal: natural k -> position
    p.perm.assign(p.m, 0);
    for (int k = 0; k < p.m; ++k) {
        int kk = k, pos = 0;
        for (int s = 0; s < p.nst; ++s) {
            int d = kk % p.radix[s];
            kk /= p.radix[s];
            pos += d * p.stride[s];
        }
        p.perm[k] = pos;
    }
    p.tw.assign(2 * (size_t)p.m, 0.0);
    for (int k = 0; k < p.m; ++k) {
        auto w = root_ld(k, p.m);
        p.tw[2 * k] = (double)w.real();
        p.tw[2 * k + 1] = (double)w.imag();
    }
    p.twst.clear();
    for (int s = 0; s < p.nst; ++s) {
        p.twoff[s] = (int)(p.twst.size() / 2);
        const int Lsub = p.radix[s] * p.stride[s];
        for (int i = 0; i < p.stride[s]; ++i) {
            auto w = root_ld(i, Lsub);
            p.twst.push_back((double)w.real());
            p.twst.push_back((double)w.imag());
        }
    }
    if (p.twst.empty()) p.twst.assign(2, 0.0);
}

// Spectrum (DIF order, scaled by 1/m) of the conjugate chirp of lag
// t + shift placed at the positions p of signed lags t: p = t for t in
// [0, tpos], p = m + t for t in [-tneg, -1].
inline std::vector<double> chirp_spectrum(const FftPlanHost &p, int tpos, int tneg, int shift = 0)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    const int n = p.n;
    auto wconj = [&](long long t) {
        t += shift;
        long long q = ((t % (2LL * n)) + 2LL * n) % (2LL * n);
        q = (q * q) % (2LL * n);
        long double a = pi * (long double)q / (long double)n;
        return std::complex<long double>(std::cos(a), std::sin(a));
    };
    std::vector<std::complex<long double>> b(p.m, 0);
    for (int t = 0; t <= tpos; ++t) b[t] = wconj(t);
    for (int t = 1; t <= tneg; ++t) b[p.m - t] = wconj(-t);
    host_dif(b, p);
    std::vector<double> out(2 * (size_t)p.m);
    for (int q = 0; q < p.m; ++q) {
        out[2 * q] = (double)(b[q].real() / p.m);
        out[2 * q + 1] = (double)(b[q].imag() / p.m);
    }
    return out;
}

// Work length of the split plan of a pruned Bluestein ring (0 if there is
// none): m >= n / 2 + 2 lc.
inline int split_length(int n, int lc)
{
    if (lc < 0 || n % 2 != 0) return 0;
    std::vector<int> rad;
    if (direct_stages(n, rad)) return 0;  // a direct plan needs no Bluestein
    return bluestein_length(n / 2 + 2 * lc);
}

// Plan for length n.  lc >= 0 (even n only) declares that the synthesis
// input and the analysis output occupy only the orders |k| <= lc; Bluestein
// then needs m >= n + 2 lc instead of 2n - 1 (signed indices: negative
// orders sit at the top of the work array).
//
// split (Bluestein rings with lc >= 0 only): the ring's n columns are taken
// as two halves [0, n/2) and [n/2, n), each needing only a convolution of
// n/2 + 2 lc lags, so m >= n/2 + 2 lc (about 2/3 of the above for
// n = 4 (lc + 1)).  Synthesis transforms the chirped orders once and runs
// the turn and the inverse stages once per half of the output columns, with
// the spectrum kept in registers in between; analysis transforms each half
// of the input columns, sums the two products with their chirp spectra in
// the turn and runs the inverse stages once.  Three transforms of the
// shorter length for two of the longer: the same work on 2/3 of the LDS.
// bhat / bhat3 are the synthesis spectra of the two output halves (lags
// shifted by 0 and n/2), bhat2 / bhat4 the analysis spectra of the two input
// halves (shifted by 0 and -n/2).
inline FftPlanHost make_fft_plan(int n, int lc = -1, bool split = false)
{
    FftPlanHost p;
    p.n = n;
    std::vector<int> rad;
    if (n == 1) {
        p.m = 1;
        set_stages(p, rad);
        return p;
    }
    if (direct_stages(n, rad)) {
        p.m = n;
        set_stages(p, rad);
        return p;
    }
    p.blue = 1;
    const bool prune = lc >= 0 && (n % 2 == 0) && n + 2 * lc < 2 * n - 1;
    p.lc = prune ? lc : -1;
    p.half = (split && prune) ? 1 : 0;
    // powers of two for unpruned plans only (the GLQ rings): pruned DH plans
    // measured slower with them, split ones more so
    p.m = bluestein_length(p.half ? n / 2 + 2 * lc : prune ? n + 2 * lc : 2 * n - 1, !p.half && !prune);
    factor_stages(p.m, false, rad);
    set_stages(p, rad);
    const long double pi = 3.14159265358979323846264338327950288L;
    p.chirp.assign(2 * (size_t)n, 0.0);
    for (int j = 0; j < n; ++j) {
        long long q = ((long long)j * j) % (2LL * n);
        long double a = -pi * (long double)q / (long double)n;
        p.chirp[2 * j] = (double)std::cos(a);
        p.chirp[2 * j + 1] = (double)std::sin(a);
    }
    if (p.half) {
        const int h = n / 2;
        p.bhat = chirp_spectrum(p, h - 1 + lc, lc, 0);    // synthesis, outputs [0, h) at 0 ..
        p.bhat3 = chirp_spectrum(p, h - 1 + lc, lc, h);   // .. and [h, n) at 0
        p.bhat2 = chirp_spectrum(p, lc, h - 1 + lc, 0);   // analysis, inputs [0, h) at 0 ..
        p.bhat4 = chirp_spectrum(p, lc, h - 1 + lc, -h);  // .. and [h, n) at 0
    } else if (prune) {
        p.bhat = chirp_spectrum(p, n - 1 + lc, lc);   // synthesis: inputs |j| <= lc, outputs [0, n)
        p.bhat2 = chirp_spectrum(p, lc, n - 1 + lc);  // analysis: inputs [0, n), outputs |k| <= lc
    } else {
        p.bhat = chirp_spectrum(p, n - 1, n - 1);
    }
    return p;
}

// The kernel instantiation a plan needs (P: FftPlanHost or the device plan).
template <class P>
inline int plan_variant(const P &F)
{
    if (F.half) return kVarSplit;
    if (F.blue && (F.m & (F.m - 1)) == 0) return kVarPow2;
    return (F.nst > 0 && F.radix[F.nst - 1] > 8) ? kVarPrime : kVarPlain;
}

}  // namespace shtg
