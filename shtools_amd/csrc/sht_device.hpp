// Device building blocks shared by the plain and batched kernels:
// extended-range powers, the LDS ring FFT, and the Legendre tiles.
#pragma once

#include <hip/hip_runtime.h>

#include "fftplan.hpp"

namespace shtg {

// --------------------------------------------------------------- helpers

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
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
    int radix[kMaxStages];
    int stride[kMaxStages];
    const double2 *tw;     // [m]
    const int *perm;       // [m]
    const double2 *chirp;  // [n]
    const double2 *bhat;   // [m]
};

// y_k = sum_j v_j e^{-2 pi i jk/R}
template <int R>
__device__ __forceinline__ void dft_small(double2 (&v)[R], const double2 *__restrict__ tw, int m)
{
    if constexpr (R == 2) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        double2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // d * (-i) = (d.y, -d.x)
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = make_double2(b.x + d.y, b.y - d.x);
        v[3] = make_double2(b.x - d.y, b.y + d.x);
    } else {
        double2 w[R];
        const int st = m / R;
#pragma unroll
        for (int q = 0; q < R; ++q) w[q] = tw[q * st];
        double2 y[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            double2 acc = v[0];
#pragma unroll
            for (int j = 1; j < R; ++j) {
                double2 c = w[(j * k) % R];
                acc.x += v[j].x * c.x - v[j].y * c.y;
                acc.y += v[j].x * c.y + v[j].y * c.x;
            }
            y[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = y[k];
    }
}

template <int R, bool DIT>
__device__ void fft_stage(double2 *x, int m, int S, const double2 *__restrict__ tw)
{
    const int Lsub = R * S, step = m / Lsub, nb = m / R;
    const float invS = 1.0f / (float)S;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        // grp = b / S without an integer division (exact after correction)
        int grp = (int)((float)b * invS);
        if (grp * S > b) --grp;
        if ((grp + 1) * S <= b) ++grp;
        const int i = b - grp * S, base = grp * Lsub + i;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = x[base + r * S];
        if constexpr (DIT) {
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[i * r * step]);
            dft_small<R>(v, tw, m);
        } else {
            dft_small<R>(v, tw, m);
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[i * r * step]);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) x[base + r * S] = v[r];
    }
}

template <bool DIT>
__device__ void fft_run_stage(double2 *x, const FftDev &P, int st)
{
    const int S = P.stride[st];
    switch (P.radix[st]) {
    case 2: fft_stage<2, DIT>(x, P.m, S, P.tw); break;
    case 3: fft_stage<3, DIT>(x, P.m, S, P.tw); break;
    case 4: fft_stage<4, DIT>(x, P.m, S, P.tw); break;
    case 5: fft_stage<5, DIT>(x, P.m, S, P.tw); break;
    case 7: fft_stage<7, DIT>(x, P.m, S, P.tw); break;
    case 11: fft_stage<11, DIT>(x, P.m, S, P.tw); break;
    case 13: fft_stage<13, DIT>(x, P.m, S, P.tw); break;
    default: break;
    }
}

// Value that the caller stores at natural index j < n before fft_forward.
__device__ __forceinline__ double2 fft_prep_in(const FftDev &P, int j, double2 v)
{
    return P.blue ? cmul(v, P.chirp[j]) : v;
}

// In-place forward DFT of the n values in x[0..n) (x has P.m slots; for
// Bluestein the caller has zeroed x[n..m)).  Ends with a barrier.
__device__ void fft_forward(double2 *x, const FftDev &P)
{
    for (int st = 0; st < P.nst; ++st) {
        fft_run_stage<false>(x, P, st);
        __syncthreads();
    }
    if (P.blue) {
        for (int q = threadIdx.x; q < P.m; q += blockDim.x) x[q] = conjd(cmul(x[q], P.bhat[q]));
        __syncthreads();
        for (int st = P.nst - 1; st >= 0; --st) {
            fft_run_stage<true>(x, P, st);
            __syncthreads();
        }
    }
}

// X_k after fft_forward.
__device__ __forceinline__ double2 fft_get(const double2 *x, const FftDev &P, int k)
{
    if (P.blue) return cmul(conjd(x[k]), P.chirp[k]);
    return x[P.perm[k]];
}

}  // namespace shtg
