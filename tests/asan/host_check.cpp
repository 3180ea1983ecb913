// Host-only check of the library's plan construction, built with
// AddressSanitizer and UndefinedBehaviorSanitizer (and libstdc++'s
// container assertions) by tests/test_asan_host.py.
//
// It runs the host code the C interface runs before any launch -- the
// recursion, seed, quadrature and polar-skip tables (tables.hpp), the ring
// FFT plans (fftplan.hpp), the batched plans' geometry, field offsets and
// work-item lists (layout.hpp) and the argument checks of the plain
// transforms (checks.hpp) -- over the degrees and mixes the transforms
// take, and checks the invariants the kernels rely on: every offset inside
// its array, every unit of work covered exactly once, every LDS class large
// enough for its members.  A sanitizer report or a failed invariant ends the
// program with a non-zero status and names the case.
//
// usage: host_check tables | plans | layout | checks
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../shtools_amd/csrc/checks.hpp"
#include "../../shtools_amd/csrc/layout.hpp"

using namespace shtg;

static int g_fail = 0;

#define CHECK(cond, ...)                                              \
    do {                                                              \
        if (!(cond)) {                                                \
            std::fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond); \
            std::fprintf(stderr, __VA_ARGS__);                        \
            std::fprintf(stderr, "\n");                               \
            if (++g_fail > 20) std::exit(1);                          \
        }                                                             \
    } while (0)

// Small deterministic generator (the cases must not depend on the libc).
struct Lcg {
    uint64_t s;
    explicit Lcg(uint64_t seed) : s(seed * 6364136223846793005ULL + 1442695040888963407ULL) {}
    uint32_t next()
    {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (uint32_t)(s >> 33);
    }
    int uniform(int lo, int hi) { return lo + (int)(next() % (uint32_t)(hi - lo + 1)); }
};

// ------------------------------------------------------------ tables

static void check_tables()
{
    int ncases = 0;
    for (int norm = 1; norm <= 4; ++norm)
        for (int L : {0, 1, 2, 3, 17, 64, 255, 700, 1300}) {
            const RecurrenceTables t = build_recurrence(norm, L);
            const size_t n = (size_t)tri_size(L);
            CHECK(t.alpha.size() == n && t.s.size() == n && t.post.size() == n && t.post_c.size() == n,
                  "norm %d L %d", norm, L);
            for (int m = 0; m <= L; ++m) {
                const int64_t o = tri_off(m, L);
                CHECK(o >= 0 && o + (L - m) < (int64_t)n, "tri_off m %d L %d", m, L);
                CHECK(m == L || tri_off(m + 1, L) == o + (L - m + 1), "tri rows contiguous m %d L %d", m, L);
                CHECK(t.alpha[o] == 0.0 && t.s[o] == 1.0, "order start norm %d L %d m %d", norm, L, m);
                for (int j = 0; j <= L - m; ++j)
                    CHECK(std::isfinite(t.alpha[o + j]) && std::isfinite(t.s[o + j]) && t.s[o + j] > 0.0 &&
                              std::isfinite(t.post[o + j]) && std::isfinite(t.post_c[o + j]),
                          "norm %d L %d m %d j %d", norm, L, m, j);
            }
            for (int cs : {1, -1})
                for (bool cplx : {false, true}) {
                    const std::vector<double> sd = build_seeds(norm, cs, L, cplx);
                    CHECK((int)sd.size() == L + 1, "seeds size");
                    // unnormalised seeds (2m-1)!! 1e-280 leave the double range
                    // from order 259, as the reference's do (MakeGridDH.f95:542-564)
                    const int mfin = norm == 3 ? std::min(L, 258) : L;
                    for (int m = 0; m <= mfin; ++m) {
                        CHECK(std::isfinite(sd[m]) && sd[m] != 0.0, "seed norm %d L %d m %d", norm, L, m);
                        // Condon-Shortley phase (-1)^m with csphase -1
                        CHECK(m == 0 || (cs == 1 ? sd[m] > 0.0 : (sd[m] > 0.0) == (m % 2 == 0)),
                              "seed sign norm %d cs %d m %d", norm, cs, m);
                    }
                }
            ++ncases;
        }
    // DH weights integrate sin(theta) d theta over [0, pi] (= 2) exactly for
    // band-limited integrands; GLQ weights sum to 2, nodes descend in (-1, 1)
    for (int L : {0, 1, 5, 63, 400, 1023}) {
        const int n = 2 * L + 2;
        std::vector<double> aj;
        dh_weights(n, aj);
        // the weights are those of the integral of f(theta) sin(theta) over
        // [0, pi] divided by sqrt(2) (src/DHaj.f95): sum sqrt(2) for f = 1,
        // sqrt(2) / 3 for f = cos^2, sqrt(2) / 5 for f = cos^4
        const double pi = std::acos(-1.0);
        double s0 = 0.0, s2 = 0.0, s4 = 0.0;
        for (int j = 0; j < (int)aj.size(); ++j) {
            const double c = std::cos(pi * j / n);
            s0 += aj[j];
            s2 += aj[j] * c * c;
            s4 += aj[j] * c * c * c * c;
        }
        const double r2 = std::sqrt(2.0);
        CHECK((int)aj.size() == n && std::fabs(s0 - r2) < 1e-12, "DH weights L %d: sum %.17g", L, s0);
        CHECK(L < 1 || std::fabs(s2 - r2 / 3.0) < 1e-12, "DH weights L %d: cos^2 %.17g", L, s2);
        CHECK(L < 2 || std::fabs(s4 - r2 / 5.0) < 1e-12, "DH weights L %d: cos^4 %.17g", L, s4);
        std::vector<double> zero(L + 1), w(L + 1);
        CHECK(glq_nodes(L + 1, zero.data(), w.data()) == 0, "glq_nodes L %d", L);
        double ws = 0.0;
        for (int i = 0; i <= L; ++i) {
            ws += w[i];
            CHECK(zero[i] > -1.0 && zero[i] < 1.0 && w[i] > 0.0, "GLQ node L %d i %d", L, i);
            CHECK(i == 0 || zero[i] < zero[i - 1], "GLQ order L %d i %d", L, i);
        }
        CHECK(std::fabs(ws - 2.0) < 1e-13, "GLQ weights L %d: sum %.17g", L, ws);
        ++ncases;
    }
    // polar skip tables: kmax non-decreasing from the pole to the equator,
    // every order keeps the equator, smin the first slot of each order
    for (int L : {0, 1, 2, 40, 333, 1023, 2047}) {
        const int n = 2 * L + 2, P = n / 2 + 1;
        const double pi = std::acos(-1.0);
        std::vector<double> z(P), u(P);
        for (int s = 0; s < P; ++s) {
            z[s] = (s == n / 2) ? 0.0 : std::cos(pi * s / n);
            u[s] = (s == n / 2) ? 1.0 : std::sqrt((1.0 - z[s]) * (1.0 + z[s]));
        }
        const std::vector<int> km = polar_kmax(z, u, L);
        const std::vector<int> sm = polar_smin(km, L);
        CHECK((int)km.size() == P && (int)sm.size() == L + 1, "polar sizes L %d", L);
        CHECK(km[P - 1] == L, "equator keeps every order L %d", L);
        for (int s = 0; s < P; ++s)
            CHECK(km[s] >= 0 && km[s] <= L && (s == 0 || km[s] >= km[s - 1]), "kmax L %d s %d", L, s);
        for (int m = 0; m <= L; ++m) {
            CHECK(sm[m] >= 0 && sm[m] < P && km[sm[m]] >= m && (sm[m] == 0 || km[sm[m] - 1] < m),
                  "smin L %d m %d", L, m);
            CHECK(m == 0 || sm[m] >= sm[m - 1], "smin order L %d m %d", L, m);
        }
        ++ncases;
    }
    std::printf("ok tables %d\n", ncases);
}

// ------------------------------------------------------------ FFT plans

static void check_plan(const FftPlanHost &p, int n, int lc, bool split)
{
    CHECK(p.n == n && p.m >= 1 && p.nst >= 0 && p.nst <= kMaxStages, "plan n %d lc %d", n, lc);
    long long prod = 1;
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.radix[s];
        // register radices and 7; a prime radix only as a direct plan's last stage
        CHECK((R >= 2 && R <= 8 && R != 6) || (is_prime_radix(R) && !p.blue && s == p.nst - 1 && n <= kPrimeMaxLen),
              "radix %d n %d stage %d", R, n, s);
        prod *= R;
        CHECK(p.stride[s] * (long long)prod == p.m, "stride n %d stage %d", n, s);
        CHECK(p.twoff[s] >= 0 && 2 * ((size_t)p.twoff[s] + p.stride[s]) <= p.twst.size(), "twiddles n %d stage %d",
              n, s);
    }
    CHECK(prod == p.m, "radices multiply to m: n %d m %d", n, p.m);
    CHECK((int)p.perm.size() == p.m && p.tw.size() == 2 * (size_t)p.m, "plan arrays n %d", n);
    std::vector<char> seen(p.m, 0);
    for (int k = 0; k < p.m; ++k) {
        CHECK(p.perm[k] >= 0 && p.perm[k] < p.m && !seen[p.perm[k]], "perm n %d k %d", n, k);
        if (p.perm[k] >= 0 && p.perm[k] < p.m) seen[p.perm[k]] = 1;
    }
    if (!p.blue) {
        CHECK(p.m == n && p.half == 0 && p.lc == -1, "direct plan n %d", n);
        return;
    }
    // Bluestein: the work array holds every lag of the convolution
    const int h = n / 2;
    const long long need = p.half ? h + 2LL * lc : p.lc >= 0 ? n + 2LL * lc : 2LL * n - 1;
    CHECK(p.m >= need && p.m <= need + need / 3 + 8, "Bluestein length n %d lc %d m %d", n, lc, p.m);
    CHECK(p.radix[p.nst - 1] % 2 == 1 || (!p.half && (p.m & (p.m - 1)) == 0 && p.m <= kPow2MaxLen &&
                                         plan_variant(p) == kVarPow2),
          "odd last radix, or a power-of-two length in its own instantiation n %d", n);
    CHECK(p.chirp.size() == 2 * (size_t)n && p.bhat.size() == 2 * (size_t)p.m, "chirp arrays n %d", n);
    CHECK(p.lc < 0 || p.bhat2.size() == 2 * (size_t)p.m, "analysis spectrum n %d", n);
    CHECK(!p.half || (split && p.bhat3.size() == 2 * (size_t)p.m && p.bhat4.size() == 2 * (size_t)p.m),
          "split spectra n %d", n);
    if (p.half) CHECK(plan_min_threads(p) <= plan_threads(p) && plan_threads(p) <= kFftMaxThreads,
                      "split plan threads n %d", n);
    for (double v : p.bhat) CHECK(std::isfinite(v), "spectrum n %d", n);
}

static void check_plans()
{
    int ncases = 0;
    for (int lo = 1; lo <= 14000; ++lo) {
        const int m = bluestein_length(lo);
        if (!(m >= lo && is_5smooth(m))) {
            CHECK(false, "bluestein_length(%d) = %d", lo, m);
            break;
        }
    }
    std::vector<int> Ls;
    for (int L = 0; L <= 1300; L += (L < 100 ? 1 : 11)) Ls.push_back(L);
    for (int L : {1023, 1024, 1290, 2047}) Ls.push_back(L);
    for (int L : Ls) {
        // DH rings: sampling 2 and 1 (pruned, split where it pays), GLQ rings
        for (int nlong : {4 * (L + 1), 2 * (L + 1)}) {
            const FftPlanHost p = ring_plan(nlong, L);
            check_plan(p, nlong, L, ring_split(nlong, L));
            CHECK(!p.half || (lds_per_cu(bluestein_length(nlong + 2 * L)) < 2 && fft_in_lds(p.m)),
                  "split only for rings of fewer than two workgroups per CU: n %d", nlong);
            ++ncases;
        }
        const FftPlanHost g = make_fft_plan(2 * L + 1);
        check_plan(g, 2 * L + 1, -1, false);
        ++ncases;
    }
    // explicit split plans of every pruned Bluestein length the DH rings have
    for (int L = 1; L <= 1300; L += 29) {
        const int n = 4 * (L + 1);
        if (split_length(n, L) <= 0) continue;
        check_plan(make_fft_plan(n, L, true), n, L, true);
        ++ncases;
    }
    std::printf("ok plans %d\n", ncases);
}

// ------------------------------------------------------------ batched layouts

// samp: grid codes per field (kGridDH1, kGridDH2, kGridGLQ)
static void check_layout_case(const std::vector<int> &lmax, const std::vector<int> &samp, const char *name)
{
    const int nf = (int)lmax.size();
    BatchLayout lay;
    CHECK(build_batch_layout(nf, lmax.data(), samp.data(), lay), "%s: layout", name);
    const size_t nslots = lay.z.size();
    CHECK(lay.u.size() == nslots && lay.w.size() == 2 * nslots && lay.rows.size() == 2 * nslots &&
              lay.kmax.size() == nslots,
          "%s: slot tables", name);
    CHECK((int)lay.fields.size() == nf, "%s: fields", name);
    long long coff = 0, foff = 0;
    for (int i = 0; i < nf; ++i) {
        const BField &f = lay.fields[i];
        const bool glq = samp[i] == kGridGLQ;
        const int n = glq ? lmax[i] + 1 : 2 * lmax[i] + 2;  // grid rows
        CHECK(f.L == lmax[i] && f.P == (glq ? (n + 1) / 2 : n / 2 + 1) && f.nlat == n &&
                  f.nlong == (glq ? 2 * lmax[i] + 1 : samp[i] * n) && f.s0 == (glq ? 0 : 1),
              "%s: field %d", name, i);
        CHECK(f.geo >= 0 && (size_t)f.geo + f.P <= nslots, "%s: field %d geo", name, i);
        CHECK(f.sm >= 0 && (size_t)f.sm + f.L + 1 <= lay.smin.size(), "%s: field %d smin", name, i);
        CHECK(f.fft >= 0 && f.fft < (int)lay.plans.size() && lay.plans[f.fft].n == f.nlong, "%s: field %d plan",
              name, i);
        CHECK(f.coff == coff && f.foff == foff, "%s: field %d offsets", name, i);
        coff += tri_size(f.L);
        foff += (long long)(f.L + 1) * f.P;
        for (int s = 0; s < f.P; ++s) {
            const int north = lay.rows[2 * (size_t)(f.geo + s)], south = lay.rows[2 * (size_t)(f.geo + s) + 1];
            // DH: row s with n - s, the equator n/2 alone; GLQ: row s with
            // nlat-1-s, the middle row of odd nlat alone
            const bool alone = glq ? (n % 2 == 1 && s == f.P - 1) : (s == n / 2);
            CHECK(north == s && (alone ? south == -1 : south == (glq ? n - 1 - s : n - s)),
                  "%s: field %d rows s %d", name, i, s);
            CHECK(!glq || (lay.w[2 * (size_t)(f.geo + s)] > 0.0 && lay.z[f.geo + s] >= 0.0 &&
                           (s == 0 || lay.z[f.geo + s] < lay.z[f.geo + s - 1])),
                  "%s: GLQ slot %d", name, s);
            CHECK(std::isfinite(lay.w[2 * (size_t)(f.geo + s)]) && std::isfinite(lay.w[2 * (size_t)(f.geo + s) + 1]),
                  "%s: weights", name);
        }
        for (int m = 0; m <= f.L; ++m) {
            const int s0 = lay.smin[(size_t)f.sm + m];
            CHECK(s0 >= 0 && s0 < f.P, "%s: field %d smin m %d", name, i, m);
        }
    }
    CHECK(lay.ncpack == coff && lay.nfbuf == foff, "%s: workspace sizes", name);

    // coefficient tiles: pack covers every tile touching the triangle, unpack every tile
    std::vector<long long> tiles_seen;
    {
        std::vector<std::vector<char>> pk(nf), up(nf);
        for (int i = 0; i < nf; ++i) {
            const int nt = (lmax[i] / kTM + 1) * (lmax[i] / kTL + 1);
            pk[i].assign(nt, 0);
            up[i].assign(nt, 0);
        }
        auto tile = [&](const Item4 &t, std::vector<std::vector<char>> &cov, const char *what) {
            CHECK(t.x >= 0 && t.x < nf, "%s: %s field %d", name, what, t.x);
            if (t.x < 0 || t.x >= nf) return;
            const int L = lmax[t.x];
            CHECK(t.y % kTM == 0 && t.z % kTL == 0 && t.y <= L && t.z <= L, "%s: %s tile", name, what);
            const int k = (t.y / kTM) * (L / kTL + 1) + t.z / kTL;
            CHECK(k >= 0 && k < (int)cov[t.x].size() && !cov[t.x][k]++, "%s: %s tile twice", name, what);
        };
        double prev = 1e300;
        for (const Item4 &t : lay.pack) {
            tile(t, pk, "pack");
            CHECK(t.z + kTL - 1 >= t.y, "%s: pack tile above the triangle", name);
            CHECK(lmax[t.x] <= prev, "%s: pack not longest first", name);
            prev = lmax[t.x];
        }
        for (const Item4 &t : lay.unpack) tile(t, up, "unpack");
        for (int i = 0; i < nf; ++i) {
            const int L = lmax[i];
            for (int m0 = 0; m0 <= L; m0 += kTM)
                for (int l0 = 0; l0 <= L; l0 += kTL) {
                    const int k = (m0 / kTM) * (L / kTL + 1) + l0 / kTL;
                    CHECK(up[i][k] == 1, "%s: unpack misses field %d tile %d", name, i, k);
                    CHECK(pk[i][k] == (l0 + kTL - 1 >= m0 ? 1 : 0), "%s: pack field %d tile %d", name, i, k);
                }
        }
    }

    // synthesis Legendre items: every (field, order, slot >= smin) once
    for (int sk = 1; sk <= kSynthKMax; ++sk)
        for (int mm : {0, sk == 1 ? 2 : 16}) {
            const SynthItems it = build_synth_items(lay.fields, lay.smin, sk, mm);
            const int nb = (int)it.bnf.size();
            CHECK(it.bfields.size() == (size_t)nb * (kBucketMaxCols / 2), "%s: bucket fields", name);
            std::vector<int> bucket_of(nf, -1);
            for (int b = 0; b < nb; ++b) {
                CHECK(it.bnf[b] >= std::max(2, mm) && it.bnf[b] <= kBucketMaxCols / 2, "%s: bucket %d size", name, b);
                for (int k = 0; k < kBucketMaxCols / 2; ++k) {
                    const int f = it.bfields[(size_t)b * (kBucketMaxCols / 2) + k];
                    CHECK(f >= 0 && f < nf, "%s: bucket %d field", name, b);
                    if (f < 0 || f >= nf) continue;
                    const int f0 = it.bfields[(size_t)b * (kBucketMaxCols / 2)];
                    CHECK(lmax[f] == lmax[f0] && samp[f] == samp[f0], "%s: bucket %d mixes degrees", name, b);
                    if (k < it.bnf[b]) {
                        CHECK(bucket_of[f] == -1, "%s: field %d in two buckets", name, f);
                        bucket_of[f] = b;
                    }
                }
            }
            CHECK(mm >= 2 || nb == 0, "%s: buckets without MFMA", name);
            std::vector<std::vector<int>> by_field(nf), by_bucket(nb);
            for (size_t k = 0; k < it.valu.size(); ++k) {
                const Item4 &t = it.valu[k];
                CHECK(t.x >= 0 && t.x < nf && bucket_of[t.x] < 0, "%s: VALU item field %d", name, t.x);
                if (t.x >= 0 && t.x < nf) by_field[t.x].push_back((int)k);
            }
            for (size_t k = 0; k < it.mfma.size(); ++k) {
                const Item4 &t = it.mfma[k];
                CHECK(t.x >= 0 && t.x < nb, "%s: MFMA item bucket %d", name, t.x);
                if (t.x >= 0 && t.x < nb) by_bucket[t.x].push_back((int)k);
            }
            for (int i = 0; i < nf; ++i) {
                if (bucket_of[i] >= 0) continue;
                const BField &f = lay.fields[i];
                std::vector<uint8_t> cov((size_t)(f.L + 1) * f.P, 0);
                for (int k : by_field[i]) {
                    const Item4 &t = it.valu[k];
                    CHECK(t.y >= 0 && t.y <= f.L && t.w >= 1 && t.w <= sk, "%s: VALU item", name);
                    if (t.y < 0 || t.y > f.L) continue;
                    // lanes pack kk ring pairs each: slots z .. z + 64 kk - 1, clipped at P
                    for (int s = t.z; s < std::min(f.P, t.z + kWave * t.w); ++s)
                        CHECK(s >= 0 && !cov[(size_t)t.y * f.P + s]++, "%s: slot twice f %d m %d s %d", name, i,
                              t.y, s);
                    CHECK(t.z + kWave * (t.w - 1) < f.P, "%s: idle lane rows f %d m %d", name, i, t.y);
                }
                for (int m = 0; m <= f.L; ++m) {
                    const int s0 = lay.smin[(size_t)f.sm + m];
                    for (int s = 0; s < f.P; ++s)
                        CHECK(cov[(size_t)m * f.P + s] == (s >= s0), "%s: synthesis cover f %d m %d s %d", name, i, m,
                              s);
                }
            }
            for (int b = 0; b < nb; ++b) {
                const BField &f = lay.fields[it.bfields[(size_t)b * (kBucketMaxCols / 2)]];
                const int ntile = (f.P + kMfmaSlots - 1) / kMfmaSlots + 1;
                std::vector<uint8_t> cov((size_t)(f.L + 1) * ntile, 0);
                for (int k : by_bucket[b]) {
                    const Item4 &t = it.mfma[k];
                    CHECK(t.y >= 0 && t.y <= f.L && t.z >= 0 && t.z < f.P && t.z % 16 == 0, "%s: MFMA tile", name);
                    if (t.y < 0 || t.y > f.L || t.z < 0 || t.z >= f.P) continue;
                    CHECK(!cov[(size_t)t.y * ntile + t.z / kMfmaSlots]++, "%s: MFMA tile twice", name);
                }
                for (int m = 0; m <= f.L; ++m) {
                    const int s0 = lay.smin[(size_t)f.sm + m] / 16 * 16;
                    for (int s = 0; s < f.P; s += kMfmaSlots)
                        CHECK(cov[(size_t)m * ntile + s / kMfmaSlots] == (s >= s0 ? 1 : 0),
                              "%s: MFMA cover bucket %d m %d s %d", name, b, m, s);
                }
            }
        }

    // analysis Legendre items: every (field, order) once, lanes within the bound
    for (int ak : {1, 3, kAnalKMax, kBatchAnalKMax}) {
        std::vector<Item2> ia[kBatchAnalKMax];
        build_anal_items(lay.fields, lay.smin, ak, ia);
        std::vector<std::vector<char>> cov(nf);
        for (int i = 0; i < nf; ++i) cov[i].assign(lay.fields[i].L + 1, 0);
        for (int k = 0; k < kBatchAnalKMax; ++k) {
            CHECK(k < ak || ia[k].empty(), "%s: analysis list %d beyond kmax %d", name, k + 1, ak);
            for (const Item2 &t : ia[k]) {
                CHECK(t.x >= 0 && t.x < nf, "%s: analysis item field", name);
                if (t.x < 0 || t.x >= nf) continue;
                const BField &f = lay.fields[t.x];
                CHECK(t.y >= 0 && t.y <= f.L, "%s: analysis order", name);
                if (t.y < 0 || t.y > f.L) continue;
                CHECK(!cov[t.x][t.y]++, "%s: analysis (%d, %d) twice", name, t.x, t.y);
                const int sa = std::max(1, lay.smin[(size_t)f.sm + t.y]);
                // one pass of 64 lanes x (k+1) ring pairs covers the slots per pass
                const int passes = (f.P - sa + kWave * (k + 1) - 1) / (kWave * (k + 1));
                CHECK(passes * kWave * (k + 1) >= f.P - sa, "%s: analysis lanes", name);
            }
        }
        for (int i = 0; i < nf; ++i)
            for (int m = 0; m <= lay.fields[i].L; ++m) CHECK(cov[i][m] == 1, "%s: analysis misses (%d, %d)", name, i, m);
    }

    // ring-FFT groups: every (field, slot block) once, LDS and threads per class
    for (int slots : {1, 2, 4}) {
        const FftLayout fl = build_fft_groups(lay.fields, lay.plans, slots);
        CHECK((int)fl.group_of.size() == nf, "%s: group_of", name);
        std::vector<std::vector<char>> cov(nf);
        for (int i = 0; i < nf; ++i) cov[i].assign((lay.fields[i].P + slots - 1) / slots, 0);
        auto take = [&](const Item2 &t, int group) {
            CHECK(t.x >= 0 && t.x < nf, "%s: FFT item field", name);
            if (t.x < 0 || t.x >= nf) return;
            const BField &f = lay.fields[t.x];
            CHECK(t.y >= 0 && t.y < f.P && t.y % slots == 0, "%s: FFT item slot", name);
            if (t.y < 0 || t.y >= f.P) return;
            CHECK(!cov[t.x][t.y / slots]++, "%s: FFT block twice", name);
            CHECK(fl.group_of[t.x] == group, "%s: field %d group %d vs %d", name, t.x, fl.group_of[t.x], group);
            const FftPlanHost &p = lay.plans[f.fft];
            if (group < (int)fl.groups.size()) {
                const FftGroupLayout &g = fl.groups[group];
                CHECK(fft_in_lds(p.m) && g.lds >= (size_t)fft_lds_slots(p.m) * kComplexBytes && g.lds <= kMaxLds,
                      "%s: group %d LDS", name, group);
                CHECK(g.threads >= plan_min_threads(p) && g.min_threads >= plan_min_threads(p) &&
                          g.threads <= g.max_threads && g.max_threads == plan_max_threads(p) &&
                          g.threads % kWave == 0,
                      "%s: group %d threads", name, group);
                CHECK(g.variant == plan_variant(p), "%s: group %d plan variant", name, group);
                // prime-radix plans run in the LDS only
                CHECK(plan_variant(p) != kVarPrime || p.n <= kPrimeMaxLen, "%s: prime plan length", name);
            } else {
                CHECK(!fft_in_lds(p.m) && fl.ws_stride >= ws_stride(p.m), "%s: workspace stride", name);
            }
        };
        for (int g = 0; g < (int)fl.groups.size(); ++g)
            for (const Item2 &t : fl.groups[g].items) take(t, g);
        for (const Item2 &t : fl.ws_items) take(t, (int)fl.groups.size());
        for (int i = 0; i < nf; ++i)
            for (size_t b = 0; b < cov[i].size(); ++b) CHECK(cov[i][b] == 1, "%s: FFT misses (%d, %zu)", name, i, b);
    }
}

static void check_layouts()
{
    int ncases = 0;
    auto run = [&](const std::vector<int> &L, const std::vector<int> &S, const std::string &name) {
        check_layout_case(L, S, name.c_str());
        ++ncases;
    };
    run({0}, {2}, "single L0");
    run({1}, {1}, "single L1 sampling 1");
    run({0, 1, 2, 3, 4, 5}, {2, 1, 2, 1, 2, 1}, "smallest degrees");
    run(std::vector<int>(40, 100), std::vector<int>(40, 2), "one degree, 40 fields (buckets)");
    run({1023, 1290, 1100}, {2, 2, 1}, "largest rings (workspace and split)");
    run({0, 1, 2, 3, 4, 5, 6}, std::vector<int>(7, kGridGLQ), "smallest GLQ degrees (odd and even nlat)");
    run({1023, 2047, 100, 100, 101}, {kGridGLQ, kGridGLQ, kGridGLQ, kGridDH2, kGridGLQ}, "GLQ with DH");
    // mixed workloads like the batched benchmark: degrees across the LDS classes
    for (int seed = 1; seed <= 4; ++seed) {
        Lcg r(seed);
        const int nf = r.uniform(1, 80);
        std::vector<int> L(nf), S(nf);
        for (int i = 0; i < nf; ++i) {
            L[i] = (r.next() % 4 == 0) ? r.uniform(0, 40) : r.uniform(41, 1100);
            if (r.next() % 5 == 0 && i > 0) L[i] = L[i - 1];  // repeated degrees form buckets
            const uint32_t k = r.next() % 4;
            S[i] = k == 0 ? kGridDH1 : k == 1 ? kGridGLQ : kGridDH2;
        }
        run(L, S, "random mix seed " + std::to_string(seed));
    }
    std::printf("ok layout %d\n", ncases);
}

// ------------------------------------------------------------ argument checks

static void check_arguments()
{
    int ncases = 0;
    Shape sh;
    // the reference's error classes: bad options -> bounds, short arrays -> dims
    struct DhCase {
        int d1, d2, lmax, norm, samp, cs, lc, ext, g0, g1, want, Lc;
    };
    const DhCase mk[] = {
        {11, 11, 10, 1, 2, 1, 10, 0, 22, 44, ST_OK, 10},
        {11, 11, 10, 1, 1, 1, 10, 0, 22, 22, ST_OK, 10},
        {11, 11, 10, 1, 2, 1, 10, 1, 23, 45, ST_OK, 10},
        {11, 11, 10, 1, 2, 1, 10, 1, 22, 45, ST_DIMS, 0},
        {11, 11, 10, 1, 2, 1, 10, 0, 22, 43, ST_DIMS, 0},
        {6, 6, 10, 1, 2, 1, 10, 0, 22, 44, ST_OK, 5},    // coefficients shorter than lmax
        {11, 11, 10, 1, 2, 1, 3, 0, 22, 44, ST_OK, 3},   // lmax_calc
        {11, 11, 10, 1, 2, 1, 11, 0, 22, 44, ST_BOUNDS, 0},
        {11, 11, 10, 0, 2, 1, 10, 0, 22, 44, ST_BOUNDS, 0},
        {11, 11, 10, 5, 2, 1, 10, 0, 22, 44, ST_BOUNDS, 0},
        {11, 11, 10, 1, 3, 1, 10, 0, 22, 44, ST_BOUNDS, 0},
        {11, 11, 10, 1, 2, 0, 10, 0, 22, 44, ST_BOUNDS, 0},
        {11, 11, 10, 1, 2, 1, 10, 2, 22, 44, ST_BOUNDS, 0},
        {11, 11, -1, 1, 2, 1, 10, 0, 22, 44, ST_BOUNDS, 0},
        {0, 11, 10, 1, 2, 1, 10, 0, 22, 44, ST_BOUNDS, 0},
        {11, 11, 10, 1, 2, 1, -1, 0, 22, 44, ST_BOUNDS, 0},
    };
    for (const DhCase &c : mk) {
        const int st = check_makegrid_dh(c.d1, c.d2, c.lmax, c.norm, c.samp, c.cs, c.lc, c.ext, c.g0, c.g1, sh);
        CHECK(st == c.want && (st != ST_OK || sh.Lc == c.Lc), "MakeGridDH case lmax %d: status %d Lc %d", c.lmax, st,
              sh.Lc);
        ++ncases;
    }
    CHECK(check_shexpand_dh(22, 44, 1, 2, 1, 10, 11, 11, sh) == ST_OK && sh.lmax == 10 && sh.Lc == 10, "SHExpandDH");
    CHECK(check_shexpand_dh(22, 22, 1, 1, 1, 4, 5, 5, sh) == ST_OK && sh.Lc == 4, "SHExpandDH lmax_calc");
    CHECK(check_shexpand_dh(21, 42, 1, 2, 1, 9, 11, 11, sh) == ST_BOUNDS, "SHExpandDH odd n");
    CHECK(check_shexpand_dh(22, 43, 1, 2, 1, 10, 11, 11, sh) == ST_DIMS, "SHExpandDH short rows");
    CHECK(check_shexpand_dh(22, 44, 1, 2, 1, 10, 10, 11, sh) == ST_DIMS, "SHExpandDH short cilm");
    CHECK(check_shexpand_dh(0, 0, 1, 2, 1, 0, 1, 1, sh) != ST_OK, "SHExpandDH empty grid");
    const double zero[12] = {0}, w[12] = {0};
    CHECK(check_makegrid_glq(11, 11, 10, 11, zero, 1, 1, 10, 0, 11, 21, sh) == ST_OK && sh.nrows == 11, "MakeGridGLQ");
    CHECK(check_makegrid_glq(11, 11, 10, 10, zero, 1, 1, 10, 0, 11, 21, sh) == ST_DIMS, "MakeGridGLQ short nodes");
    CHECK(check_makegrid_glq(11, 11, 10, 11, nullptr, 1, 1, 10, 0, 11, 21, sh) == ST_OTHER, "MakeGridGLQ no nodes");
    CHECK(check_makegrid_glq(11, 11, 10, 11, zero, 1, 1, 10, 1, 11, 21, sh) == ST_DIMS, "MakeGridGLQ extend columns");
    CHECK(check_shexpand_glq(11, 21, w, 11, zero, 11, 1, 1, 10, 11, 11, sh) == ST_OK && sh.Lc == 10, "SHExpandGLQ");
    CHECK(check_shexpand_glq(11, 20, w, 11, zero, 11, 1, 1, 10, 11, 11, sh) == ST_DIMS, "SHExpandGLQ short rows");
    CHECK(check_shexpand_glq(11, 21, nullptr, 11, zero, 11, 1, 1, 10, 11, 11, sh) == ST_DIMS, "SHExpandGLQ no weights");
    CHECK(check_shexpand_glq(0, 21, w, 11, zero, 11, 1, 1, 10, 11, 11, sh) == ST_DIMS, "SHExpandGLQ empty grid");
    // the plx branch: the table in place of the nodes
    const double plx[4] = {0};
    CHECK(check_makegrid_glq_plx(11, 11, 10, plx, 1, 1, 10, 0, 11, 21, sh) == ST_OK && sh.Lc == 10,
          "MakeGridGLQ plx");
    CHECK(check_makegrid_glq_plx(11, 11, 10, nullptr, 1, 1, 10, 0, 11, 21, sh) == ST_OTHER, "MakeGridGLQ no table");
    CHECK(check_makegrid_glq_plx(11, 11, -1, plx, 1, 1, 10, 0, 11, 21, sh) == ST_BOUNDS, "MakeGridGLQ plx lmax");
    CHECK(check_makegrid_glq_plx(11, 11, 10, plx, 5, 1, 10, 0, 11, 21, sh) == ST_BOUNDS, "MakeGridGLQ plx norm");
    CHECK(check_shexpand_glq_plx(11, 21, w, 11, plx, 1, 1, 4, 5, 5, sh) == ST_OK && sh.Lc == 4, "SHExpandGLQ plx");
    CHECK(check_shexpand_glq_plx(11, 21, w, 11, nullptr, 1, 1, 10, 11, 11, sh) == ST_OTHER, "SHExpandGLQ no table");
    CHECK(check_shexpand_glq_plx(11, 21, w, 10, plx, 1, 1, 10, 11, 11, sh) == ST_DIMS, "SHExpandGLQ plx weights");
    CHECK(check_shexpand_glq_plx(11, 21, w, 11, plx, 1, 0, 10, 11, 11, sh) == ST_BOUNDS, "SHExpandGLQ plx phase");
    ncases += 22;
    // extreme integers: every check returns a status, and a grid it accepts
    // has dimensions that exist (no overflow in the shape arithmetic)
    const int ext[] = {INT_MIN, INT_MIN + 1, -2, -1, 0, 1, 2, 3, 1000, (INT_MAX - 4) / 4, (INT_MAX - 4) / 4 + 1,
                       INT_MAX / 2, INT_MAX - 1, INT_MAX};
    for (int a : ext)
        for (int b : ext) {
            int st = check_makegrid_dh(b, b, a, 1, 2, 1, a, 1, b, b, sh);
            CHECK(st != ST_OK || (sh.nrows == 2 * a + 3 && sh.Lc <= a), "MakeGridDH lmax %d dims %d", a, b);
            st = check_makegrid_dh(a, a, a, 1, 2, 1, b, 0, INT_MAX, INT_MAX, sh);
            CHECK(st != ST_OK || (sh.Lc >= 0 && sh.Lc <= a), "MakeGridDH lmax %d lmax_calc %d", a, b);
            st = check_shexpand_dh(a, b, 1, 2, 1, INT_MAX, b, b, sh);
            st = check_shexpand_dh(a, b, 1, 2, 1, a / 2 - 1 > 0 ? a / 2 - 1 : 0, INT_MAX, INT_MAX, sh);
            CHECK(st != ST_OK || (a % 2 == 0 && a >= 2 && (long long)b >= 2LL * a), "SHExpandDH n %d cols %d", a, b);
            st = check_makegrid_glq(b, b, a, INT_MAX, zero, 1, 1, a, 1, b, b, sh);
            CHECK(st != ST_OK || (long long)b >= 2LL * a + 2, "MakeGridGLQ lmax %d dims %d", a, b);
            st = check_shexpand_glq(a, b, w, INT_MAX, zero, INT_MAX, 1, 1, INT_MAX, INT_MAX, INT_MAX, sh);
            CHECK(st != ST_OK || (a >= 1 && (long long)b >= 2LL * a - 1), "SHExpandGLQ rows %d cols %d", a, b);
            st = check_makegrid_glq_plx(b, b, a, plx, 1, 1, a, 1, b, b, sh);
            CHECK(st != ST_OK || (long long)b >= 2LL * a + 2, "MakeGridGLQ plx lmax %d dims %d", a, b);
            st = check_shexpand_glq_plx(a, b, w, INT_MAX, plx, 1, 1, INT_MAX, INT_MAX, INT_MAX, sh);
            CHECK(st != ST_OK || (a >= 1 && (long long)b >= 2LL * a - 1), "SHExpandGLQ plx rows %d cols %d", a, b);
            ncases += 8;
        }
    std::printf("ok checks %d\n", ncases);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_check tables|plans|layout|checks\n");
        return 2;
    }
    const std::string what = argv[1];
    if (what == "tables") check_tables();
    else if (what == "plans") check_plans();
    else if (what == "layout") check_layouts();
    else if (what == "checks") check_arguments();
    else return 2;
    return g_fail ? 1 : 0;
}
