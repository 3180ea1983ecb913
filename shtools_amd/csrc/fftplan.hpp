// Host planning of the per-ring FFTs run in LDS by the GPU kernels.
//
// Every ring (pair) transform is a complex DFT of length n, forward
// X_k = sum_j x_j e^{-2 pi i jk/n}; inverse transforms are taken as
// conj(F(conj(.))).  When n factors into radices {8, 4, 2, 3, 5, 7} the
// kernel runs in-place decimation-in-frequency stages on the length-n array
// and reads results through a digit-reversal table.  Otherwise the DFT is
// computed by Bluestein's chirp-z method on a 5-smooth work length m >= 2n-1:
// the chirp-modulated input is transformed by DIF, multiplied by the
// precomputed spectrum of the conjugate chirp (stored in DIF output order,
// scaled by 1/m) and transformed back by the transposed (DIT) stages, which
// consume digit-reversed input and leave natural order.
#pragma once

#include <cmath>
#include <complex>
#include <vector>

namespace shtg {

constexpr int kMaxStages = 24;

struct FftPlanHost {
    int n = 0, m = 0, nst = 0, blue = 0;
    int lc = -1;                // pruning: orders |k| <= lc only (-1: none)
    int radix[kMaxStages] = {0};
    int stride[kMaxStages] = {0};
    int twoff[kMaxStages] = {0};  // offset of each stage's twiddles in twst
    std::vector<double> tw;     // 2*m: e^{-2 pi i k/m}
    std::vector<double> twst;   // per stage: e^{-2 pi i i / (R S)}, i < S
    std::vector<int> perm;      // m: DIF output position of natural index k
    std::vector<double> chirp;  // 2*n (Bluestein)
    std::vector<double> bhat;   // 2*m (Bluestein), DIF output order, /m
    std::vector<double> bhat2;  // pruned plans: analysis spectrum (bhat is synthesis)
};

inline bool factor_smooth(int m, std::vector<int> &rad)
{
    rad.clear();
    int r = m;
    while (r % 8 == 0) { rad.push_back(8); r /= 8; }
    if (r % 4 == 0) { rad.push_back(4); r /= 4; }
    if (r % 2 == 0) { rad.push_back(2); r /= 2; }
    static const int primes[] = {3, 5, 7};
    for (int p : primes)
        while (r % p == 0) { rad.push_back(p); r /= p; }
    return r == 1;
}

inline int next_5smooth(int v)
{
    for (int m = v;; ++m) {
        int r = m;
        while (r % 2 == 0) r /= 2;
        while (r % 3 == 0) r /= 3;
        while (r % 5 == 0) r /= 5;
        if (r == 1) return m;
    }
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
// output left in DIF order.  Used to build the Bluestein spectrum.
inline void host_dif(std::vector<std::complex<long double>> &x, const FftPlanHost &p)
{
    int m = p.m;
    for (int st = 0; st < p.nst; ++st) {
        int R = p.radix[st], S = p.stride[st], Lsub = R * S, step = m / Lsub;
        std::vector<std::complex<long double>> v(R), y(R);
        for (int b = 0; b < m / R; ++b) {
            int grp = b / S, i = b - grp * S, base = grp * Lsub + i;
            for (int r = 0; r < R; ++r) v[r] = x[base + r * S];
            for (int k = 0; k < R; ++k) {
                std::complex<long double> acc = 0;
                for (int j = 0; j < R; ++j) acc += v[j] * root_ld((long long)j * k, R);
                y[k] = acc * root_ld((long long)i * k * step, m);
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

// Spectrum (DIF order, scaled by 1/m) of the conjugate chirp placed at the
// positions p of signed lags t: p = t for t in [0, tpos], p = m + t for
// t in [-tneg, -1].
inline std::vector<double> chirp_spectrum(const FftPlanHost &p, int tpos, int tneg)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    const int n = p.n;
    auto wconj = [&](long long t) {
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

// Plan for length n.  lc >= 0 (even n only) declares that the synthesis
// input and the analysis output occupy only the orders |k| <= lc; Bluestein
// then needs m >= n + 2 lc instead of 2n - 1 (signed indices: negative
// orders sit at the top of the work array).
inline FftPlanHost make_fft_plan(int n, int lc = -1)
{
    FftPlanHost p;
    p.n = n;
    std::vector<int> rad;
    if (n == 1) {
        p.m = 1;
        set_stages(p, rad);
        return p;
    }
    if (factor_smooth(n, rad) && (int)rad.size() <= kMaxStages) {
        p.m = n;
        set_stages(p, rad);
        return p;
    }
    p.blue = 1;
    const bool prune = lc >= 0 && (n % 2 == 0) && n + 2 * lc < 2 * n - 1;
    p.lc = prune ? lc : -1;
    p.m = next_5smooth(prune ? n + 2 * lc : 2 * n - 1);
    factor_smooth(p.m, rad);
    set_stages(p, rad);
    const long double pi = 3.14159265358979323846264338327950288L;
    p.chirp.assign(2 * (size_t)n, 0.0);
    for (int j = 0; j < n; ++j) {
        long long q = ((long long)j * j) % (2LL * n);
        long double a = -pi * (long double)q / (long double)n;
        p.chirp[2 * j] = (double)std::cos(a);
        p.chirp[2 * j + 1] = (double)std::sin(a);
    }
    if (prune) {
        p.bhat = chirp_spectrum(p, n - 1 + lc, lc);   // synthesis: inputs |j| <= lc, outputs [0, n)
        p.bhat2 = chirp_spectrum(p, lc, n - 1 + lc);  // analysis: inputs [0, n), outputs |k| <= lc
    } else {
        p.bhat = chirp_spectrum(p, n - 1, n - 1);
    }
    return p;
}

}  // namespace shtg
