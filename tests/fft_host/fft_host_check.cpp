// CPU run of the library's device FFT code (sht_device.hpp compiled for the
// host with hip/hip_runtime.h from this folder) against direct DFT sums.
//
// For every ring length given on the command line it builds the plan the
// library builds (make_fft_plan, pruned for synthesis and analysis when
// even), fills the work array the way the kernels do (fft_prep_in at
// fft_pos), runs fft_forward and reads with fft_get.  Each output checked is
// compared with the direct sum X_k = sum_j x_j e^{-2 pi i jk/n} evaluated in
// long double; the line printed per length is
//   n m blue nstages radices... max_rel_err
// with the error relative to max |X| (1 when the chirp table breaks the
// mirror identity chirp_{n-j} = (-1)^n chirp_j that the kernels rely on, 2
// when fft_forward_chunked, the variant for work arrays outside the LDS,
// differs from fft_forward in any bit).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../shtools_amd/csrc/sht_device.hpp"

using namespace shtg;

static FftDev to_dev(const FftPlanHost &h)
{
    FftDev d{};
    d.n = h.n;
    d.m = h.m;
    d.nst = h.nst;
    d.blue = h.blue;
    d.lc = h.lc;
    for (int s = 0; s < kMaxStages; ++s) {
        d.radix[s] = h.radix[s];
        d.stride[s] = h.stride[s];
        d.twoff[s] = h.twoff[s];
    }
    d.tw = reinterpret_cast<const double2 *>(h.tw.data());
    d.twst = reinterpret_cast<const double2 *>(h.twst.data());
    d.perm = h.perm.data();
    d.chirp = h.blue ? reinterpret_cast<const double2 *>(h.chirp.data()) : nullptr;
    d.bhat = h.blue ? reinterpret_cast<const double2 *>(h.bhat.data()) : nullptr;
    d.bhat2 = h.blue ? reinterpret_cast<const double2 *>((h.bhat2.empty() ? h.bhat : h.bhat2).data()) : nullptr;
    d.half = h.half;
    d.bhat3 = h.half ? reinterpret_cast<const double2 *>(h.bhat3.data()) : nullptr;
    d.bhat4 = h.half ? reinterpret_cast<const double2 *>(h.bhat4.data()) : nullptr;
    return d;
}

// Direct DFT of the sparse input (index, value) at output k; root holds
// the n-th roots of unity.
static std::complex<long double> direct(const std::vector<std::pair<int, double2>> &in, int n, int k,
                                        const std::vector<std::complex<long double>> &root)
{
    std::complex<long double> acc = 0;
    for (const auto &e : in) {
        const long long jk = ((long long)e.first * k) % n;
        acc += std::complex<long double>(e.second.x, e.second.y) * root[jk];
    }
    return acc;
}

// The turn of a split plan run by T virtual threads (each its own stash),
// T the fewest that give every thread at most kTurnStash / R butterflies,
// as the launches do.
template <int MODE>
static void turn_threads(double2 *x, const FftDev &F, const double2 *bhat, std::vector<TurnStash> &keep)
{
    const int R = F.radix[F.nst - 1], nb = kTurnStash / R;
    const int T = (F.m / R + nb - 1) / nb;
    keep.resize(T);
    blockDim.x = T;
    for (int t = 0; t < T; ++t) {
        threadIdx.x = t;
        fft_run_turn_split<MODE>(x, F, bhat, keep[t]);
    }
    blockDim.x = 1;
    threadIdx.x = 0;
}

// Split plan (make_fft_plan(n, lc, true)): synthesis (orders |k| <= lc in,
// every column out, by halves) and analysis (every column in, by halves;
// orders |k| <= lc out) as the kernels run them, against direct sums.
static double run_split(const FftPlanHost &h, std::mt19937_64 &rng, int nsample)
{
    const FftDev F = to_dev(h);
    const int n = h.n, hh = n / 2;
    std::normal_distribution<double> g;
    std::vector<std::complex<long double>> root(n);
    for (int k = 0; k < n; ++k) root[k] = root_ld(k, n);
    double maxerr = 0.0, maxref = 0.0;
    auto check = [&](const std::vector<std::pair<int, double2>> &in, int k, double2 y) {
        const auto r = direct(in, n, k, root);
        maxerr = std::max(maxerr, (double)std::abs(std::complex<long double>(y.x, y.y) - r));
        maxref = std::max(maxref, (double)std::abs(r));
    };
    std::uniform_int_distribution<int> u(0, n - 1);
    // synthesis
    {
        std::vector<double2> x((size_t)h.m, make_double2(0.0, 0.0));
        std::vector<std::pair<int, double2>> in;
        for (int k = -h.lc; k <= h.lc; ++k) {
            const int j = (k + n) % n;
            const double2 v = make_double2(g(rng), g(rng));
            in.push_back({j, v});
            x[fft_pos(F, j)] = fft_prep_in(F, j, v);
        }
        std::vector<TurnStash> keep;
        fft_dif_front(x.data(), F);
        for (int half = 0; half < 2; ++half) {
            if (half == 0)
                turn_threads<kSynth1>(x.data(), F, F.bhat, keep);
            else
                turn_threads<kSynth2>(x.data(), F, F.bhat3, keep);
            fft_dit_back(x.data(), F);
            std::vector<int> ps = {0, hh - 1};
            for (int t = 0; t < nsample && (int)ps.size() < hh; ++t) ps.push_back(u(rng) % hh);
            if (hh <= nsample) {
                ps.clear();
                for (int p = 0; p < hh; ++p) ps.push_back(p);
            }
            for (int p : ps) check(in, half * hh + p, fft_get(x.data(), F, half * hh + p, p));
        }
    }
    // analysis
    {
        std::vector<double2> x((size_t)h.m), val(n);
        std::vector<std::pair<int, double2>> in;
        for (int j = 0; j < n; ++j) {
            val[j] = make_double2(g(rng), g(rng));
            in.push_back({j, val[j]});
        }
        std::vector<TurnStash> keep;
        for (int half = 0; half < 2; ++half) {
            for (int q = 0; q < h.m; ++q)
                x[q] = q < hh ? fft_prep_in(F, half * hh + q, val[half * hh + q]) : make_double2(0.0, 0.0);
            fft_dif_front(x.data(), F);
            if (half == 0)
                turn_threads<kAnal1>(x.data(), F, F.bhat2, keep);
            else
                turn_threads<kAnal2>(x.data(), F, F.bhat4, keep);
        }
        fft_dit_back(x.data(), F);
        for (int k = -h.lc; k <= h.lc; ++k) {
            if (2 * h.lc + 1 > nsample && k != -h.lc && k != h.lc && k != 0 && u(rng) % (2 * h.lc + 1) >= nsample)
                continue;
            const int kk = (k + n) % n;
            check(in, kk, fft_get(x.data(), F, kk, fft_pos(F, kk)));
        }
    }
    return maxerr / maxref;
}

// mode 0: full input and output; 1: synthesis of a pruned plan (orders
// |k| <= lc in, all n out); 2: analysis of a pruned plan (all n in, orders
// |k| <= lc out).
static double run_case(const FftPlanHost &h, int mode, std::mt19937_64 &rng, int nsample)
{
    const FftDev F = to_dev(h);
    const int n = h.n;
    std::normal_distribution<double> g;
    std::vector<double2> x((size_t)h.m, make_double2(0.0, 0.0));
    std::vector<std::pair<int, double2>> in;
    if (mode == 1) {
        for (int k = -h.lc; k <= h.lc; ++k) {
            const int j = (k + n) % n;
            const double2 v = make_double2(g(rng), g(rng));
            in.push_back({j, v});
            x[fft_pos(F, j)] = fft_prep_in(F, j, v);
        }
    } else {
        for (int j = 0; j < n; ++j) {
            const double2 v = make_double2(g(rng), g(rng));
            in.push_back({j, v});
            x[j] = fft_prep_in(F, j, v);
        }
    }
    // the chunked variant (work array outside the LDS, inner stages on
    // chunks of one and of three groups) must give the same bits
    std::vector<std::vector<double2>> xc;
    // (plans ending with a prime radix and power-of-two Bluestein plans run
    // in the LDS only, never chunked)
    if (h.nst >= 2 && plan_variant(h) != kVarPrime && plan_variant(h) != kVarPow2) {
        const int gsz = h.m / h.radix[0];
        for (int cap : {gsz, 3 * gsz}) {
            std::vector<double2> y = x, lds((size_t)cap);
            fft_forward_chunked(y.data(), F, mode == 2 ? F.bhat2 : F.bhat, lds.data(), cap);
            xc.push_back(y);
        }
    }
    if (plan_variant(h) == kVarPow2)
        fft_forward<true, true>(x.data(), F, mode == 2 ? F.bhat2 : F.bhat);
    else
        fft_forward<true>(x.data(), F, mode == 2 ? F.bhat2 : F.bhat);
    for (const auto &y : xc)
        for (int e = 0; e < h.m; ++e)
            if (y[e].x != x[e].x || y[e].y != x[e].y) return 2.0;
    std::vector<int> ks;
    if (mode == 2) {
        for (int k = -h.lc; k <= h.lc; ++k) ks.push_back((k + n) % n);
    } else {
        for (int k = 0; k < n; ++k) ks.push_back(k);
    }
    // sample the outputs (all of them for short rings)
    if ((int)ks.size() > nsample) {
        std::vector<int> pick;
        std::uniform_int_distribution<int> u(0, (int)ks.size() - 1);
        pick.push_back(ks.front());
        pick.push_back(ks.back());
        for (int t = 2; t < nsample; ++t) pick.push_back(ks[u(rng)]);
        ks = pick;
    }
    // the mirror identity the kernels use: chirp_{n-j} = (-1)^n chirp_j
    if (h.blue) {
        for (int j = 1; j < n; ++j) {
            const double2 a = fft_chirp(F, n - j), b = fft_chirp_mirror(F, fft_chirp(F, j));
            if (std::fabs(a.x - b.x) + std::fabs(a.y - b.y) > 4e-16) return 1.0;
        }
    }
    std::vector<std::complex<long double>> root(n);
    for (int k = 0; k < n; ++k) root[k] = root_ld(k, n);
    double maxerr = 0.0, maxref = 0.0;
    for (int k : ks) {
        const int pos = (mode == 2) ? fft_pos(F, k) : k;
        const double2 y = fft_get(x.data(), F, k, pos);
        const auto r = direct(in, n, k, root);
        maxerr = std::max(maxerr, (double)std::abs(std::complex<long double>(y.x, y.y) - r));
        maxref = std::max(maxref, (double)std::abs(r));
    }
    return maxerr / maxref;
}

int main(int argc, char **argv)
{
    std::mt19937_64 rng(7);
    const int nsample = 48;
    for (int a = 1; a < argc; ++a) {
        const int n = std::atoi(argv[a]);
        const int lc = (n % 2 == 0) ? n / 4 - 1 : -1;  // the DH ring: lmax = n/4 - 1
        FftPlanHost full = make_fft_plan(n);
        double err = run_case(full, 0, rng, nsample);
        int nsplit = 0;
        if (n % 2 == 0) {
            FftPlanHost pr = make_fft_plan(n, lc);
            if (pr.lc >= 0) {
                err = std::max(err, run_case(pr, 1, rng, nsample));
                err = std::max(err, run_case(pr, 2, rng, nsample));
            }
            // split plans of the DH2 (lc = n/4 - 1) and DH1 (lc = n/2 - 1) rings
            for (int l : {lc, n / 2 - 1}) {
                FftPlanHost sp = make_fft_plan(n, l, true);
                if (sp.half) {
                    err = std::max(err, run_split(sp, rng, nsample));
                    ++nsplit;
                }
            }
        }
        if (std::getenv("SPLIT_COUNT")) std::printf("split %d ", nsplit);
        std::printf("%d %d %d %d", n, full.m, full.blue, full.nst);
        for (int s = 0; s < full.nst; ++s) std::printf(" %d", full.radix[s]);
        std::printf(" %.3e\n", err);
    }
    return 0;
}
