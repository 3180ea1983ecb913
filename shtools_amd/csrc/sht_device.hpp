// Device building blocks shared by the plain and batched kernels:
// extended-range powers, the LDS ring FFT, and the Legendre tiles.
#pragma once

#include <hip/hip_runtime.h>

#include "fftplan.hpp"
#include "prime_roots.hpp"

namespace shtg {

// --------------------------------------------------------------- helpers

// Complex product with the fused multiply-adds written out, so that its
// rounding does not depend on which product the compiler would contract in
// a given kernel (the plain and the batched kernels then agree bit for bit).
// Loads through the global address space.  Pointers the batched kernels
// read from tables in memory (plans, field arrays) are generic to the
// compiler, and its flat loads count against the LDS wait counter as well:
// each wait for LDS data would then also wait for a twiddle or spectrum
// load from the caches.  T must live in device global memory.
// (Vector types are loaded by component: their copy constructors take
// generic references, which would turn the load back into a flat one.)
typedef const __attribute__((address_space(1))) double gdouble;
typedef const __attribute__((address_space(1))) int gint;
__device__ __forceinline__ double gld(const double *p) { return *(gdouble *)p; }
__device__ __forceinline__ int gld(const int *p) { return *(gint *)p; }
__device__ __forceinline__ double2 gld(const double2 *p)
{
    gdouble *q = (gdouble *)p;
    return make_double2(q[0], q[1]);
}
__device__ __forceinline__ double4 gld(const double4 *p)
{
    gdouble *q = (gdouble *)p;
    return make_double4(q[0], q[1], q[2], q[3]);
}
// Stores through the global address space (see gld).
__device__ __forceinline__ void gst(double *p, double v) { *(__attribute__((address_space(1))) double *)p = v; }

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

__host__ __device__ __forceinline__ long long tri_offd(int m, int L)
{
    return (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
}

// 1e280 * u^m for m >= 1 and 0 <= u <= 1, i.e. the reference's rescalem
// after m iterations of rescalem = rescalem * u starting from 1/1e-280
// (src/MakeGridDH.f95:548,552).  Binary powering on a (mantissa, exponent)
// pair keeps the full range; the final ldexp rounds into the subnormal range
// or to zero exactly where the iterated product does.
__device__ __forceinline__ double scaled_upow(double u, int m)
{
    if (u == 0.0) return 0.0;
    int e;
    double f = frexp(u, &e);
    double r = 1.0, b = f;
    int re = 0, be = 0;
    int k = m;
    while (k) {
        if (k & 1) {
            int t;
            r = frexp(r * b, &t);
            re += be + t;
        }
        k >>= 1;
        if (k) {
            int t;
            b = frexp(b * b, &t);
            be = 2 * be + t;
        }
    }
    return ldexp(r * 1.0e280, re + e * m);
}

// --------------------------------------------------------------- FFT

struct FftDev {
    int n, m, nst, blue;
    int lc;                // pruned Bluestein: orders |k| <= lc (-1: unpruned)
    int half;              // split Bluestein plan (the ring's columns in two halves; see make_fft_plan)
    int radix[kMaxStages];
    int stride[kMaxStages];
    int twoff[kMaxStages];
    const double2 *tw;     // [m]
    const double2 *twst;   // per-stage twiddles (w_{RS}^i, i < S)
    const int *perm;       // [m]
    const double2 *chirp;  // [n]
    const double2 *bhat;   // [m] (pruned plans: synthesis spectrum)
    const double2 *bhat2;  // [m] pruned plans: analysis spectrum (else = bhat)
    const double2 *bhat3;  // [m] split plans: synthesis spectrum of the second output half
    const double2 *bhat4;  // [m] split plans: analysis spectrum of the second input half
};

// Work-array position of signed order k (|k| <= lc) of a pruned plan, or of
// natural index k otherwise.
__device__ __forceinline__ int fft_pos(const FftDev &P, int k)
{
    return (P.lc >= 0 && k > P.lc) ? k - P.n + P.m : k;
}

// LDS slot of FFT element e (the one place that fixes the layout).  The
// array is unpadded: the stages run the radix-8/4/2 factors first and the odd
// radices last, so the late, small-stride stages access the LDS at odd
// strides, which spread over the banks; padding (one slot per 16 elements)
// broke those patterns and measured 6% slower.
__device__ __forceinline__ int fpad(int e) { return e; }

// Fully unrolled forward DFTs, y_k = sum_j v_j e^{-2 pi i jk/R}, in place.
__device__ __forceinline__ void dft2(double2 &a, double2 &b)
{
    const double2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

__device__ __forceinline__ void dft4(double2 &v0, double2 &v1, double2 &v2, double2 &v3)
{
    const double2 a = cadd(v0, v2), b = csub(v0, v2);
    const double2 c = cadd(v1, v3), d = csub(v1, v3);
    v0 = cadd(a, c);
    v2 = csub(a, c);
    v1 = make_double2(b.x + d.y, b.y - d.x);  // b - i d
    v3 = make_double2(b.x - d.y, b.y + d.x);  // b + i d
}

template <int R>
__device__ __forceinline__ void dft_unrolled(double2 (&v)[R])
{
    if constexpr (R == 2) {
        dft2(v[0], v[1]);
    } else if constexpr (R == 3) {
        constexpr double c = -0.5, s = -0.86602540378443864676;  // e^{-2 pi i/3}
        const double2 t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const double2 m = make_double2(v[0].x + c * t.x, v[0].y + c * t.y);
        v[0] = cadd(v[0], t);
        v[1] = make_double2(m.x - s * d.y, m.y + s * d.x);
        v[2] = make_double2(m.x + s * d.y, m.y - s * d.x);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 5) {
        constexpr double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
        constexpr double s1 = -0.95105651629515357212, s2 = -0.58778525229247312917;
        const double2 t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        const double2 t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        const double2 m1 = make_double2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        const double2 m2 = make_double2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        // i * (s1 d1 + s2 d2) and i * (s2 d1 - s1 d2)
        const double2 n1 = make_double2(-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x);
        const double2 n2 = make_double2(-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x);
        v[0] = cadd(v[0], cadd(t1, t2));
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    } else if constexpr (R == 8) {
        constexpr double h = 0.70710678118654752440;
        double2 b0 = v[0], b1 = v[1], b2 = v[2], b3 = v[3];
        double2 c0 = v[4], c1 = v[5], c2 = v[6], c3 = v[7];
        dft2(b0, c0);
        dft2(b1, c1);
        dft2(b2, c2);
        dft2(b3, c3);
        // c_j *= e^{-i pi j/4}
        c1 = make_double2(h * (c1.x + c1.y), h * (c1.y - c1.x));
        c2 = make_double2(c2.y, -c2.x);
        c3 = make_double2(h * (c3.y - c3.x), -h * (c3.x + c3.y));
        dft4(b0, b1, b2, b3);
        dft4(c0, c1, c2, c3);
        v[0] = b0; v[2] = b1; v[4] = b2; v[6] = b3;
        v[1] = c0; v[3] = c1; v[5] = c2; v[7] = c3;
    }
}

// DFT of prime length R in registers (the last stage of direct plans whose
// length has a prime factor among kPrimeRadices): with t_j = v_j + v_{R-j}
// and d_j = v_j - v_{R-j}, j = 1 .. (R-1)/2,
//     X_k     = v_0 + sum_j t_j cos(2 pi jk/R) - i sum_j d_j sin(2 pi jk/R),
//     X_{R-k} = v_0 + sum_j t_j cos(2 pi jk/R) + i sum_j d_j sin(2 pi jk/R),
// (R-1)^2 fused multiply-adds per butterfly in place of 4 (R-1)^2 for the
// plain sums.  The roots come from the constant table kPrimeRoots
// (prime_roots.hpp) at indices that are the same for every lane, so they
// arrive as scalar operands; the loop over k stays rolled to keep the
// registers for t and d.  Outputs go to x[base + k*S] as they are formed.
template <int R>
__device__ __forceinline__ void dft_prime_store(double2 (&v)[R], double2 *x, int base, int S)
{
    constexpr int H = (R - 1) / 2;
    constexpr int off = prime_roots_offset(R);
    static_assert(off >= 0, "a prime radix with a root table");
    double2 t[H], d[H];
    const double2 x0 = v[0];
#pragma unroll
    for (int j = 1; j <= H; ++j) {
        t[j - 1] = cadd(v[j], v[R - j]);
        d[j - 1] = csub(v[j], v[R - j]);
    }
    double2 sum = x0;
#pragma unroll
    for (int j = 0; j < H; ++j) sum = cadd(sum, t[j]);
    x[fpad(base)] = sum;
#pragma unroll 1
    for (int k = 1; k <= H; ++k) {
        double2 a = x0, b = make_double2(0.0, 0.0);
        int q = 0;
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            q += k;
            if (q >= R) q -= R;  // q = jk mod R
            const double c = kPrimeRoots[off + 2 * q], sn = kPrimeRoots[off + 2 * q + 1];
            a.x = fma(t[j - 1].x, c, a.x);
            a.y = fma(t[j - 1].y, c, a.y);
            b.x = fma(d[j - 1].x, sn, b.x);
            b.y = fma(d[j - 1].y, sn, b.y);
        }
        x[fpad(base + k * S)] = make_double2(a.x + b.y, a.y - b.x);
        x[fpad(base + (R - k) * S)] = make_double2(a.x - b.y, a.y + b.x);
    }
}

// v[r] *= w^r for r = 1..R-1 (R <= 8), w = e^{-2 pi i i/(R S)} of the
// butterfly: the powers come from a product tree of depth 3.
template <int R>
__device__ __forceinline__ void apply_twiddles(double2 (&v)[R], double2 w1)
{
    double2 w[R];
    w[0] = make_double2(1.0, 0.0);
    if constexpr (R > 1) w[1] = w1;
    if constexpr (R > 2) w[2] = cmul(w[1], w[1]);
    if constexpr (R > 3) w[3] = cmul(w[2], w[1]);
    if constexpr (R > 4) w[4] = cmul(w[2], w[2]);
    if constexpr (R > 5) w[5] = cmul(w[4], w[1]);
    if constexpr (R > 6) w[6] = cmul(w[4], w[2]);
    if constexpr (R > 7) w[7] = cmul(w[4], w[3]);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], w[r]);
}

// One in-place stage of radix R over x[0..m): butterflies over elements
// base + r*S; DIF applies twiddles after the butterfly, DIT (the transposed
// stage) before it.  Radices 2, 3, 4, 5 and 8 are unrolled in registers;
// radix 7 (and a prime radix at a stride other than 1) uses a rolled loop
// over outputs that stores each output as it is formed; the prime radices
// at stride 1 use dft_prime_store.
//
// x holds mloc elements (mloc = m, or a run of whole groups of R*S elements
// of the m-element transform: the butterflies and twiddles are the same in
// every group).
template <int R, bool DIT>
__device__ __forceinline__ void fft_stage(double2 *x, int m, int S, const double2 *__restrict__ tw,
                                          const double2 *__restrict__ tws, int mloc)
{
    constexpr bool kUnrolled = (R == 2 || R == 3 || R == 4 || R == 5 || R == 8);
    constexpr bool kPrime = R >= 11 && !DIT;
    const int Lsub = R * S, step = m / Lsub, nb = mloc / R, nt = blockDim.x;
    // butterfly b = grp * S + i starts at element base = grp * Lsub + i; the
    // thread's first (grp, i) come from one division, later ones by stepping
    // b by nt = dg * S + di
    int i, base;
    {
        const int b = threadIdx.x;
        int grp = (int)((float)b * (1.0f / (float)S));
        if (grp * S > b) --grp;
        if ((grp + 1) * S <= b) ++grp;
        i = b - grp * S;
        base = grp * Lsub + i;
    }
    const int dg = nt / S, di = nt - dg * S, dbase = dg * Lsub + di, wrap = Lsub - S;
    for (int b = threadIdx.x; b < nb; b += nt) {
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = x[fpad(base + r * S)];
        if constexpr (kUnrolled) {
            // twiddles w^r, w = e^{-2 pi i i/Lsub}
            const double2 w1 = gld(tws + i);
            if constexpr (DIT) apply_twiddles<R>(v, w1);
            dft_unrolled<R>(v);
            if constexpr (!DIT) apply_twiddles<R>(v, w1);
#pragma unroll
            for (int r = 0; r < R; ++r) x[fpad(base + r * S)] = v[r];
        } else if (kPrime && S == 1) {
            // the last DIF stage of a direct plan: no twiddles
            if constexpr (kPrime) dft_prime_store<R>(v, x, base, S);
        } else {
            if constexpr (DIT) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[r] = cmul(v[r], gld(tw + i * r * step));
            }
            const int st = m / R;
            for (int k = 0; k < R; ++k) {
                double2 acc = v[0];
                int jk = 0;
#pragma unroll
                for (int j = 1; j < R; ++j) {
                    jk += k;
                    if (jk >= R) jk -= R;
                    const double2 c = gld(tw + jk * st);
                    acc.x += v[j].x * c.x - v[j].y * c.y;
                    acc.y += v[j].x * c.y + v[j].y * c.x;
                }
                if constexpr (!DIT) {
                    if (k > 0) acc = cmul(acc, gld(tw + i * k * step));
                }
                x[fpad(base + k * S)] = acc;
            }
        }
        i += di;
        base += dbase;
        if (i >= S) {
            i -= S;
            base += wrap;
        }
    }
}

// kPrimes: the instantiation for direct plans ending with a prime radix
// (kVarPrime), which alone carries the prime butterflies and their
// registers.  Such plans have no DIT stages.
template <bool DIT, bool kPrimes = false>
__device__ __forceinline__ void fft_run_stage(double2 *x, const FftDev &P, int st, int mloc)
{
    const int S = P.stride[st];
    const double2 *tws = P.twst + P.twoff[st];
    switch (P.radix[st]) {
#define SHTG_STAGE_CASE(R) \
    case R: fft_stage<R, DIT>(x, P.m, S, P.tw, tws, mloc); break;
        SHTG_FFT_RADICES(SHTG_STAGE_CASE)
#undef SHTG_STAGE_CASE
    case 7: fft_stage<7, DIT>(x, P.m, S, P.tw, tws, mloc); break;
    default:
        if constexpr (kPrimes && !DIT) {
            switch (P.radix[st]) {
#define SHTG_PRIME_CASE(R) \
    case R: fft_stage<R, false>(x, P.m, S, P.tw, tws, mloc); break;
                SHTG_PRIME_RADICES(SHTG_PRIME_CASE)
#undef SHTG_PRIME_CASE
            default: break;
            }
        }
        break;
    }
}

// Value that the caller stores at natural index j < n (LDS slot fpad(j))
// before fft_forward.
__device__ __forceinline__ double2 fft_prep_in(const FftDev &P, int j, double2 v)
{
    return P.blue ? cmul(v, gld(P.chirp + j)) : v;
}

// The chirp e^{-i pi j^2/n} of natural index j (Bluestein plans).  Indices j
// and n - j share it up to a sign, chirp_{n-j} = (-1)^n chirp_j, so callers
// that handle both read the table once.
__device__ __forceinline__ double2 fft_chirp(const FftDev &P, int j)
{
    return P.blue ? gld(P.chirp + j) : make_double2(1.0, 0.0);
}
__device__ __forceinline__ double2 fft_chirp_mirror(const FftDev &P, double2 c)
{
    return (P.n & 1) ? make_double2(-c.x, -c.y) : c;
}
// fft_prep_in with the chirp of j given.
__device__ __forceinline__ double2 fft_prep_in_c(const FftDev &P, double2 v, double2 c)
{
    return P.blue ? cmul(v, c) : v;
}

// The turn of a Bluestein transform: the last DIF stage, the product with
// the chirp spectrum (conjugated, so that the DIT stages compute the inverse)
// and the first DIT stage.  The last stage has stride 1, so its twiddles are
// all 1 and DIF and DIT touch the same R contiguous elements of one
// butterfly: the three steps run in registers, with one LDS read and one
// write per element.
//
// x holds mloc elements; bhat is the spectrum at x's first element.
template <int R>
__device__ __forceinline__ void fft_turn(double2 *x, int mloc, const double2 *__restrict__ bhat)
{
    for (int b = threadIdx.x; b < mloc / R; b += blockDim.x) {
        const int base = b * R;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = x[fpad(base + r)];
        dft_unrolled<R>(v);
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = conjd(cmul(v[r], gld(bhat + base + r)));
        dft_unrolled<R>(v);
#pragma unroll
        for (int r = 0; r < R; ++r) x[fpad(base + r)] = v[r];
    }
}

// The last radix of a Bluestein plan is 3 or 5, or, for the power-of-two
// lengths of kVarPow2 plans, 8, 4 or 2 (bluestein_length).
template <bool kPow2 = false>
__device__ __forceinline__ void fft_run_turn(double2 *x, const FftDev &P, const double2 *__restrict__ bhat, int mloc)
{
    if constexpr (kPow2) {
        if (P.radix[P.nst - 1] == 8)
            fft_turn<8>(x, mloc, bhat);
        else if (P.radix[P.nst - 1] == 4)
            fft_turn<4>(x, mloc, bhat);
        else
            fft_turn<2>(x, mloc, bhat);
    } else {
        if (P.radix[P.nst - 1] == 3)
            fft_turn<3>(x, mloc, bhat);
        else
            fft_turn<5>(x, mloc, bhat);
    }
}

// In-place forward DFT of the values stored in x (x has P.m slots; for
// Bluestein the caller has zeroed the unused slots).  bhat selects the chirp
// spectrum (P.bhat for synthesis, P.bhat2 for analysis).  Ends with a barrier.
//
// One call site per stage direction (the last DIF stage of a plain plan runs
// in the same loop as the others), which keeps the kernels' code and
// register use down.
template <bool kPrimes = false, bool kPow2 = false>
__device__ __forceinline__ void fft_forward(double2 *x, const FftDev &P, const double2 *__restrict__ bhat)
{
    const int last = P.nst - 1;
    const int ndif = P.blue ? last : last + 1;  // Bluestein: the last stage is part of the turn
    for (int st = 0; st < ndif; ++st) {
        fft_run_stage<false, kPrimes>(x, P, st, P.m);
        __syncthreads();
    }
    if (!P.blue) return;
    fft_run_turn<kPow2>(x, P, bhat, P.m);
    __syncthreads();
    for (int st = last - 1; st >= 0; --st) {
        fft_run_stage<true>(x, P, st, P.m);
        __syncthreads();
    }
}

// fft_forward for a work array x in global memory (rings whose array exceeds
// the LDS).  After the first DIF stage the transform splits into radix[0]
// independent groups of m / radix[0] consecutive elements, and every later
// stage, the Bluestein turn and every DIT stage but the last stay inside a
// group.  So only the first stage and its transpose run on x; the rest runs
// in LDS on chunks of whole groups (lds holds lds_cap elements, at least one
// group): load, all inner stages, store.  x then crosses the memory system
// four times per transform instead of twice per stage.  Ends with a barrier.
__device__ __forceinline__ void fft_forward_chunked(double2 *x, const FftDev &P, const double2 *__restrict__ bhat,
                                                    double2 *lds, int lds_cap)
{
    const int last = P.nst - 1;
    const int gsz = P.m / P.radix[0];
    const int per = max(1, lds_cap / gsz) * gsz;  // elements per chunk
    const int ndif = P.blue ? last : last + 1;
    fft_run_stage<false>(x, P, 0, P.m);
    __syncthreads();
    for (int c0 = 0; c0 < P.m; c0 += per) {
        const int len = min(per, P.m - c0);
        for (int e = threadIdx.x; e < len; e += blockDim.x) lds[fpad(e)] = x[c0 + e];
        __syncthreads();
        for (int st = 1; st < ndif; ++st) {
            fft_run_stage<false>(lds, P, st, len);
            __syncthreads();
        }
        if (P.blue) {
            fft_run_turn(lds, P, bhat + c0, len);
            __syncthreads();
            for (int st = last - 1; st >= 1; --st) {
                fft_run_stage<true>(lds, P, st, len);
                __syncthreads();
            }
        }
        for (int e = threadIdx.x; e < len; e += blockDim.x) x[c0 + e] = lds[fpad(e)];
        __syncthreads();
    }
    if (P.blue && last >= 1) {
        fft_run_stage<true>(x, P, 0, P.m);
        __syncthreads();
    }
}

// ---------------------------------------------------- split Bluestein plans
//
// A split plan (FftDev::half) runs the pieces of fft_forward separately: the
// DIF stages before the turn, the turn once per half of the ring's columns
// (with the spectrum, or the first half's product, kept in registers in
// between) and the DIT stages after it.

// DIF stages of a Bluestein plan before the turn; ends with a barrier.
__device__ __forceinline__ void fft_dif_front(double2 *x, const FftDev &P)
{
    for (int st = 0; st < P.nst - 1; ++st) {
        fft_run_stage<false>(x, P, st, P.m);
        __syncthreads();
    }
}

// DIT stages of a Bluestein plan after the turn; ends with a barrier.
__device__ __forceinline__ void fft_dit_back(double2 *x, const FftDev &P)
{
    for (int st = P.nst - 2; st >= 0; --st) {
        fft_run_stage<true>(x, P, st, P.m);
        __syncthreads();
    }
}

// Values a thread keeps between the two turns of a split plan: its turn
// butterflies' R values each, at most kTurnStash of them (the launch gives
// every thread at most kTurnStash / R butterflies; see plan_min_threads).
struct TurnStash {
    double2 v[kTurnStash];
};

// The turns of a split plan (R the last radix, one butterfly of R contiguous
// elements per step; the spectrum S is the last DIF stage's output):
//   kSynth1: S from x, kept; x <- DIT(conj(S bhat))     (first output half)
//   kSynth2: S from the stash; x <- DIT(conj(S bhat))   (second output half)
//   kAnal1:  S from x; keeps S bhat; x is not written   (first input half)
//   kAnal2:  S from x; x <- DIT(conj(S bhat + kept))    (second input half)
enum { kSynth1 = 0, kSynth2 = 1, kAnal1 = 2, kAnal2 = 3 };
template <int R, int MODE>
__device__ __forceinline__ void fft_turn_split(double2 *x, int m, const double2 *__restrict__ bhat, TurnStash &keep)
{
    // butterfly k of this thread keeps keep.v[k R .. k R + R) (compile-time
    // indices: the values stay in registers)
    constexpr int NB = kTurnStash / R;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        const int b = threadIdx.x + k * blockDim.x;
        if (b >= m / R) break;
        const int base = b * R, i = k * R;
        double2 v[R];
        if constexpr (MODE == kSynth2) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = keep.v[i + r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = x[fpad(base + r)];
            dft_unrolled<R>(v);
        }
        if constexpr (MODE == kSynth1) {
#pragma unroll
            for (int r = 0; r < R; ++r) keep.v[i + r] = v[r];
        }
        if constexpr (MODE == kAnal1) {
#pragma unroll
            for (int r = 0; r < R; ++r) keep.v[i + r] = cmul(v[r], gld(bhat + base + r));
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double2 t = cmul(v[r], gld(bhat + base + r));
                if constexpr (MODE == kAnal2) t = cadd(t, keep.v[i + r]);
                v[r] = conjd(t);
            }
            dft_unrolled<R>(v);
#pragma unroll
            for (int r = 0; r < R; ++r) x[fpad(base + r)] = v[r];
        }
    }
}

template <int MODE>
__device__ __forceinline__ void fft_run_turn_split(double2 *x, const FftDev &P, const double2 *__restrict__ bhat,
                                                   TurnStash &keep)
{
    if (P.radix[P.nst - 1] == 3)
        fft_turn_split<3, MODE>(x, P.m, bhat, keep);
    else
        fft_turn_split<5, MODE>(x, P.m, bhat, keep);
}

// X_k after fft_forward; for Bluestein, pos is the work-array position
// holding order k (fft_pos for signed orders of pruned plans).
__device__ __forceinline__ double2 fft_get(const double2 *x, const FftDev &P, int k, int pos)
{
    if (P.blue) return cmul(conjd(x[fpad(pos)]), gld(P.chirp + k));
    return x[fpad(gld(P.perm + k))];
}

// fft_get with the chirp c of k given (fft_chirp / fft_chirp_mirror).
__device__ __forceinline__ double2 fft_get_c(const double2 *x, const FftDev &P, int k, int pos, double2 c)
{
    if (P.blue) return cmul(conjd(x[fpad(pos)]), c);
    return x[fpad(gld(P.perm + k))];
}

}  // namespace shtg
