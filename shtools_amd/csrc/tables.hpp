// Host-side tables for the MI355X spherical-harmonic transforms.
//
// Recursion constants follow the reference's table code (identical in all
// four transforms, e.g. src/MakeGridDH.f95:291-381), re-expressed for the
// GPU recursion
//     q_l = (alpha_l z) q_{l-1} - q_{l-2},   P_lm = s_lm q_l,
// which divides the reference's three-term recursion
//     P_l = z ff1(l,m) P_{l-1} - ff2(l,m) P_{l-2}
// by s_l = ff2(l,m) s_{l-2} (s_m = s_{m+1} = 1), so that alpha_l =
// ff1(l,m) s_{l-1} / s_l.  Each step then costs one multiply and one FMA.
// The s factors are folded into the coefficients before synthesis and into
// the normalisation epilogue after analysis.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace shtg {

// Offset of order m in a triangular [m][l-m] table with maximum degree L.
inline int64_t tri_off(int m, int L) { return (int64_t)m * (L + 1) - (int64_t)m * (m - 1) / 2; }
inline int64_t tri_size(int L) { return (int64_t)(L + 1) * (L + 2) / 2; }

struct RecurrenceTables {
    int L = -1;
    std::vector<double> alpha;  // tri: alpha_{m+j}; entry j = 0 is 0
    std::vector<double> s;      // tri: s_{m+j}
    std::vector<double> post;   // tri: s_{m+j} * (analysis epilogue factor)
    std::vector<double> post_c; // tri: the same for the complex analysis
};

// Recursion coefficients of one order m, l = m+1..L (index l), in extended
// precision: P_l = ff1 z P_{l-1} - ff2 P_{l-2} with the reference's
// coefficients (src/MakeGridDH.f95:321-381: norm 1,4 / 2 / 3 branches,
// the m = l-1 entry with ff2 = 0), each from exact integer ratios with
// one long-double square root.
inline void ff_column(int norm, int m, int L, std::vector<long double> &ff1, std::vector<long double> &ff2)
{
    typedef long double R;
    ff1.assign(L + 2, 0.0L);
    ff2.assign(L + 2, 0.0L);
    for (int l = m + 1; l <= L; ++l) {
        const R lp = (R)(l + m), lm = (R)(l - m), t1 = (R)(2 * l - 1), t3 = (R)(2 * l - 3), t2 = (R)(2 * l + 1);
        R f1 = 0.0L, f2 = 0.0L;
        if (norm == 1 || norm == 4) {
            if (m == 0) {
                if (l == 1) f1 = std::sqrt(3.0L);
                else {
                    f1 = std::sqrt(t1 * t2) / (R)l;
                    f2 = (R)(l - 1) * std::sqrt(t2 / t3) / (R)l;
                }
            } else {
                f1 = std::sqrt(t2 * t1 / (lp * lm));
                if (m <= l - 2) f2 = std::sqrt(t2 * (lm - 1) * (lp - 1) / (t3 * lp * lm));
            }
        } else if (norm == 2) {
            if (m == 0) {
                if (l == 1) f1 = 1.0L;
                else { f1 = t1 / (R)l; f2 = (R)(l - 1) / (R)l; }
            } else {
                f1 = t1 / std::sqrt(lp * lm);
                if (m <= l - 2) f2 = std::sqrt((lm - 1) * (lp - 1) / (lp * lm));
            }
        } else {  // norm 3
            if (m == 0) { f1 = t1 / (R)l; f2 = (R)(l - 1) / (R)l; }
            else { f1 = t1 / lm; f2 = (lp - 1) / lm; }
        }
        ff1[l] = f1;
        ff2[l] = f2;
    }
}

// Analysis epilogue factor per (l, m): src/SHExpandDH.f95:674-702.  The
// complex analysis (src/SHExpandDHC.f95:681-723) omits the halving of the
// unnormalised (norm 3) product for m >= 1.
inline void epilogue_factors(int norm, int L, std::vector<double> &epi /* [l*(L+1)+m] */, bool cplx = false)
{
    const double pi = std::acos(-1.0);
    epi.assign((size_t)(L + 1) * (L + 1), 0.0);
    for (int l = 0; l <= L; ++l) {
        if (norm == 1) {
            for (int m = 0; m <= l; ++m) epi[(size_t)l * (L + 1) + m] = 1.0 / (4 * pi);
        } else if (norm == 2) {
            for (int m = 0; m <= l; ++m) epi[(size_t)l * (L + 1) + m] = (double)(2 * l + 1) / (4 * pi);
        } else if (norm == 3) {
            double prod = 4 * pi / (double)(2 * l + 1);
            epi[(size_t)l * (L + 1)] = 1.0 / prod;
            if (!cplx) prod = prod / 2.0;
            for (int m = 1; m <= l - 1; ++m) {
                prod = prod * (l + m) * (l - m + 1);
                epi[(size_t)l * (L + 1) + m] = 1.0 / prod;
            }
            if (l != 0) epi[(size_t)l * (L + 1) + l] = 1.0 / (prod * 2 * l);
        } else {
            for (int m = 0; m <= l; ++m) epi[(size_t)l * (L + 1) + m] = 1.0;
        }
    }
}

inline RecurrenceTables build_recurrence(int norm, int L)
{
    RecurrenceTables t;
    t.L = L;
    size_t n = (size_t)tri_size(L);
    t.alpha.assign(n, 0.0);
    t.s.assign(n, 0.0);
    t.post.assign(n, 0.0);
    t.post_c.assign(n, 0.0);
    std::vector<double> epi, epi_c;
    epilogue_factors(norm, L, epi);
    epilogue_factors(norm, L, epi_c, true);
    // The running products s and the ratios alpha are formed in long double
    // and rounded once: alpha and s then carry at most half an ulp each.
    // Rounded along the way in double (about 30 ulps after a few thousand
    // degrees), their perturbations add up near the poles, where the
    // recursion's solutions hardly separate: at z = cos(pi/5600), degree
    // 2800, order 0, the functions were 1.5e-11 off against 2.3e-12 now (the
    // reference's own double recursion: 1.2e-12).
    std::vector<long double> ff1, ff2;
    for (int m = 0; m <= L; ++m) {
        ff_column(norm, m, L, ff1, ff2);
        int64_t o = tri_off(m, L);
        long double s_lm2 = 1.0L, s_lm1 = 1.0L;  // s_{l-2}, s_{l-1}
        for (int l = m; l <= L; ++l) {
            long double s, a;
            if (l == m) { s = 1.0L; a = 0.0L; }
            else if (l == m + 1) { s = 1.0L; a = ff1[l]; }
            else { s = ff2[l] * s_lm2; a = ff1[l] * s_lm1 / s; }
            t.alpha[o + (l - m)] = (double)a;
            t.s[o + (l - m)] = (double)s;
            t.post[o + (l - m)] = (double)(s * (long double)epi[(size_t)l * (L + 1) + m]);
            t.post_c[o + (l - m)] = (double)(s * (long double)epi_c[(size_t)l * (L + 1) + m]);
            s_lm2 = s_lm1;
            s_lm1 = s;
        }
    }
    return t;
}

// Seeds q_m of each order: P_00 for m = 0 (MakeGridDH.f95:516-519) and the
// 1e-280-scaled sectoral chain for m >= 1 (MakeGridDH.f95:542-564).  The
// complex transforms' functions carry no sqrt(2) for m >= 1: their chain
// starts at 1e-280 (norms 1, 2, 3) or 1e-280/sqrt(4 pi) (norm 4)
// (MakeGridDHC.f95:485-489).
inline std::vector<double> build_seeds(int norm, int csphase, int L, bool cplx = false)
{
    const double pi = std::acos(-1.0), scalef = 1.0e-280;
    std::vector<double> seed(L + 1, 0.0);
    std::vector<double> sqr(2 * L + 4, 0.0);
    for (int k = 1; k < (int)sqr.size(); ++k) sqr[k] = std::sqrt((double)k);
    seed[0] = (norm == 4) ? 1.0 / std::sqrt(4.0 * pi) : 1.0;
    double pmm;
    switch (norm) {
    case 1: case 2: pmm = cplx ? scalef : sqr[2] * scalef; break;
    case 3: pmm = scalef; break;
    default: pmm = (cplx ? scalef : sqr[2] * scalef) / std::sqrt(4.0 * pi); break;
    }
    for (int m = 1; m <= L; ++m) {
        if (norm == 1 || norm == 4) {
            pmm = csphase * pmm * sqr[2 * m + 1] / sqr[2 * m];
            seed[m] = pmm;
        } else if (norm == 2) {
            pmm = csphase * pmm * sqr[2 * m + 1] / sqr[2 * m];
            seed[m] = pmm / sqr[2 * m + 1];
        } else {
            pmm = csphase * pmm * (double)(2 * m - 1);
            seed[m] = pmm;
        }
    }
    return seed;
}

// src/DHaj.f95:91-101
inline void dh_weights(int n, std::vector<double> &aj)
{
    const double pi = std::acos(-1.0);
    aj.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        double sum1 = 0.0;
        for (int l = 0; l <= n / 2 - 1; ++l)
            sum1 = sum1 + std::sin((double)(2 * l + 1) * pi * (double)j / (double)n) / (double)(2 * l + 1);
        aj[j] = sum1 * std::sin(pi * (double)j / (double)n) * std::sqrt(8.0) / (double)n;
    }
}

// Gauss-Legendre nodes and weights on [-1, 1]: src/PreGLQ.f95:93-149.
// Returns 0, or 5 when Newton's method does not converge.
inline int glq_nodes(int n, double *zero, double *w)
{
    const int itermax = 1000;
    const double eps = 1.0e-15, pi = std::acos(-1.0);
    for (int i = 0; i < n; ++i) { zero[i] = 0.0; w[i] = 0.0; }
    int half = (n + 1) / 2;
    for (int i = 1; i <= half; ++i) {
        int iter = 0;
        double z = std::cos(pi * (i - 0.25) / (n + 0.5)), z1, p1, p2, p3, pp;
        for (;;) {
            ++iter;
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                p3 = p2;
                p2 = p1;
                p1 = ((double)(2 * j - 1) * z * p2 - (double)(j - 1) * p3) / (double)j;
            }
            pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) <= eps) break;
            if (iter > itermax) return 5;
        }
        zero[i - 1] = z;
        zero[n - i] = -z;
        w[i - 1] = 2.0 / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
    return 0;
}

// ------------------------------------------------------------ polar skip
//
// Near the poles the associated Legendre functions of high order are
// exponentially small at every degree: for a ring with sin(theta) = u, the
// 4pi-normalised P_lm(cos theta), m <= l <= L, stay below kPolarEps when m is
// well above L u.  Such (order, ring) pairs contribute nothing representable
// to a field or to its coefficients (below 1e-32 of the coefficients'
// magnitudes, times at most the number of degrees), so the transforms skip
// them: the Legendre kernels do not run their recursion, the ring FFTs treat
// their Fourier coefficients as zero.  The reference sums these terms; the
// results differ from it far below any rounding the transforms incur.
constexpr double kPolarEps = 1e-32;

// kmax[s]: the largest order m <= L whose functions reach kPolarEps at ring
// slot s for some degree l <= L (0 at the poles); non-decreasing in s for
// slots ordered from the pole to the equator (|z| non-increasing), which
// the caller guarantees.  A staircase walk over (s, m) with one recursion per
// step: O((P + L) L) operations.  Without that ordering every slot keeps all
// orders.
inline std::vector<int> polar_kmax(const std::vector<double> &z, const std::vector<double> &u, int L)
{
    const int P = (int)z.size();
    std::vector<int> km(P, L);
    for (int s = 1; s < P; ++s)
        if (std::fabs(z[s]) > std::fabs(z[s - 1])) return km;
    // log of the sectoral value without u^m: P_mm = c_m u^m, c_1 = sqrt(3),
    // c_m = c_m-1 sqrt((2m+1)/(2m))
    std::vector<double> lc(L + 1, 0.0);
    if (L >= 1) lc[1] = 0.5 * std::log(3.0);
    for (int m = 2; m <= L; ++m) lc[m] = lc[m - 1] + 0.5 * std::log((2.0 * m + 1.0) / (2.0 * m));
    const double leps = std::log(kPolarEps);
    // does some P_lm(z_s), m <= l <= L, reach kPolarEps?
    auto reaches = [&](int s, int m) {
        if (m == 0) return true;
        const double zz = z[s], uu = u[s];
        if (uu <= 0.0) return false;
        double off = lc[m] + m * std::log(uu);  // P = p exp(off)
        if (off >= leps) return true;
        // |p| >= thr reaches; recomputed from off (not rescaled) so that a
        // threshold beyond the double range does not stick at infinity
        auto threshold = [&]() { return leps - off > 700.0 ? HUGE_VAL : std::exp(leps - off); };
        double thr = threshold();
        double p2 = 0.0, p1 = 1.0;
        for (int l = m + 1; l <= L; ++l) {
            const double a = std::sqrt((2.0 * l + 1.0) * (2.0 * l - 1.0) / ((double)(l - m) * (double)(l + m)));
            const double b = (l > m + 1) ? std::sqrt((2.0 * l + 1.0) * (double)(l - m - 1) * (double)(l + m - 1) /
                                                     ((2.0 * l - 3.0) * (double)(l - m) * (double)(l + m)))
                                         : 0.0;
            const double p = a * zz * p1 - b * p2;
            p2 = p1;
            p1 = p;
            if (std::fabs(p1) >= thr) return true;
            if (std::fabs(p1) > 1e150) {
                p1 *= 1e-150;
                p2 *= 1e-150;
                off += std::log(1e150);
                thr = threshold();
            }
        }
        return false;
    };
    int m = 0;
    for (int s = 0; s < P; ++s) {
        while (m < L && reaches(s, m + 1)) ++m;
        km[s] = m;
    }
    return km;
}

// smin[m]: the first slot whose kmax reaches m (kmax non-decreasing), at most
// P - 1 so that every order keeps a slot.
inline std::vector<int> polar_smin(const std::vector<int> &kmax, int L)
{
    const int P = (int)kmax.size();
    std::vector<int> sm(L + 1, P - 1);
    int s = 0;
    for (int m = 0; m <= L; ++m) {
        while (s < P && kmax[s] < m) ++s;
        sm[m] = std::min(s, P - 1);
    }
    return sm;
}

}  // namespace shtg
