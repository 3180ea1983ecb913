// MI355X kernels of the transform path: per-workgroup tiles, the plain
// (one field per call) kernels and the batched (many fields per launch)
// kernels built from the same tiles.
//
// Synthesis (MakeGridDH / MakeGridGLQ):
//   pack        cilm[2][d1][d2] -> cpack[m][l-m] = (C s_lm, S s_lm) (LDS-tiled
//               transpose; S_l0 is ignored as in src/MakeGridDH.f95:523-538)
//   synth_leg   one workgroup per (order m, block of ring pairs); lanes are
//               ring pairs, the recursion over l runs in registers with
//               wave-uniform (scalar) table and coefficient loads; north and
//               south sums come from even/odd degree parity
//               -> fbuf[m][slot] = (C_N, S_N, C_S, S_S) * rescale
//   synth_fft   one workgroup per ring pair: Hermitian fill of both rings
//               into one complex spectrum, inverse DFT in LDS, rows out
// Analysis (SHExpandDH / SHExpandGLQ):
//   anal_fft    one workgroup per ring pair: both rings as one complex
//               signal, forward DFT in LDS, split, quadrature weights
//               -> gbuf[m][slot] = (E_C, E_S, O_C, O_S)
//   anal_leg    one workgroup per order m over all ring pairs; each lane
//               holds 4 ring pairs per pass, the recursion runs in registers
//               and the per-degree sums over rings are reduced through LDS
//               in chunks of 8 degrees -> cbuf[m][l-m]
//   unpack      cbuf * (s_lm * normalisation) -> cilm[2][d1][d2]
//
// Geometry of a ring-pair slot s: z = cos(theta) of the northern ring, the
// southern ring is its mirror (z -> -z).  Reference loops:
// src/MakeGridDH.f95:504-639, src/SHExpandDH.f95:463-664,
// src/MakeGridGLQ.f95:467-669, src/SHExpandGLQ.f95:445-649.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.hpp"
#include "sht_device.hpp"

namespace shtg {

// Largest workgroup of a ring-FFT kernel instantiation: 512 threads for the
// prime-radix plans (up to 256 registers per thread for the prime
// butterflies of up to 23 values), 1024 otherwise.
constexpr int fft_max_threads(int var) { return var == kVarPrime ? kFftMaxThreads / 2 : kFftMaxThreads; }

// ====================================================================== tiles

// Coefficient tiles of kTL degrees x kTM orders, transposed through LDS so
// that both the [l][m] arrays and the per-order packed arrays are accessed
// contiguously.
typedef double2 CoefTile[kTL][kTM + 1];

// cpack[tri(m) + l - m] = (C s, S s) for the tile's m <= l <= L.
__device__ __forceinline__ void pack_tile(const double *__restrict__ cilm, int d1, int d2, int L,
                                          const double *__restrict__ s_tab, int Lt, double2 *__restrict__ cpack,
                                          int m0, int l0, CoefTile &t)
{
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int r = e / kTM, c = e - r * kTM, l = l0 + r, m = m0 + c;
        double2 v = make_double2(0.0, 0.0);
        if (l <= L && m <= l) {
            v.x = cilm[(size_t)l * d2 + m];
            v.y = (m == 0) ? 0.0 : cilm[((size_t)d1 + l) * d2 + m];
        }
        t[r][c] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int c = e / kTL, r = e - c * kTL, l = l0 + r, m = m0 + c;
        if (l <= L && m <= l) {
            const double s = s_tab[tri_offd(m, Lt) + (l - m)];
            const double2 v = t[r][c];
            cpack[tri_offd(m, L) + (l - m)] = make_double2(v.x * s, v.y * s);
        }
    }
}

// cilm rows [l0, l0+kTL) x columns [m0, m0+kTM) (within d1 x d2): packed
// sums times (s_lm * normalisation) for m <= l <= L, zero elsewhere and S_l0 = 0.
// The packed sums may come in nblk partial blocks cstride apart, added in
// order.
__device__ __forceinline__ void unpack_tile(const double2 *__restrict__ cbuf, const double *__restrict__ post,
                                            int Lt, int L, double *__restrict__ cilm, int d1, int d2, int m0, int l0,
                                            CoefTile &t, int nblk = 1, long long cstride = 0)
{
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int c = e / kTL, r = e - c * kTL, l = l0 + r, m = m0 + c;
        double2 v = make_double2(0.0, 0.0);
        if (l <= L && m <= l) {
            const double p = post[tri_offd(m, Lt) + (l - m)];
            double2 x = cbuf[tri_offd(m, L) + (l - m)];
            for (int b = 1; b < nblk; ++b) {
                const double2 y = cbuf[b * cstride + tri_offd(m, L) + (l - m)];
                x = make_double2(x.x + y.x, x.y + y.y);
            }
            v = make_double2(x.x * p, m == 0 ? 0.0 : x.y * p);
        }
        t[r][c] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int r = e / kTM, c = e - r * kTM, l = l0 + r, m = m0 + c;
        if (l < d1 && m < d2) {
            const double2 v = t[r][c];
            cilm[(size_t)l * d2 + m] = v.x;
            cilm[((size_t)d1 + l) * d2 + m] = v.y;
        }
    }
}

// The reference's rescalem for order m at a ring with sin(theta) = u
// (1 for m = 0, where the recursion is unscaled: src/MakeGridDH.f95:516-540).
__device__ __forceinline__ double order_rescale(double u, int m) { return m == 0 ? 1.0 : scaled_upow(u, m); }

// The rescale a ring FFT applies to order k of a slot with sin(theta) = u:
// order_rescale, or 1 where u < 0 marks sums of actual function values
// (tables of the functions, no extended-range scaling).
__device__ __forceinline__ double ring_rescale(double u, int k) { return u < 0.0 ? 1.0 : order_rescale(u, k); }

// Synthesis sums of order m for the ring pairs slot0 + k*64 + lane (one
// wave).  al: recursion row of order m (al[j] = alpha_{m+j}); cp: packed
// coefficients of order m; frow: fbuf row of order m.  The sums are stored
// without the order's rescale factor, which the ring FFT applies.
//
// kRho: sums of rho^(l-m) P_lm per ring (rho = ringr2[s]^(1/2), the radial
// continuation factor r0/r of the ring): the rescaled recursion
// q_l = alpha_l z q_{l-1} - q_{l-2} becomes, for rho^(l-m) q_l,
// q_l = alpha_l (rho z) q_{l-1} - rho^2 q_{l-2}, with ringz holding rho z.
//
// kStage: the recursion and coefficient rows reach the wave in blocks of 64
// degrees through LDS instead of as wave-uniform scalar loads: one vector
// load per lane fetches a block while the previous one is in use (vector
// loads complete in order, so waiting for one block does not wait for the
// next; scalar loads are waited for all together, which leaves their
// latency on every few degrees of a wave whose rows are not cached), and
// the degree loop reads the block back as wave-wide broadcasts (sal, scp:
// the kernel's LDS blocks of 64 entries).  For the launches of one field,
// whose longest waves bound them; workgroups are one wave.
template <int K, bool kRho, bool kStage>
__device__ __forceinline__ void synth_leg_tile(const double *__restrict__ al, double sd,
                                               const double *__restrict__ ringz, const double *__restrict__ ringr2,
                                               const double2 *__restrict__ cp, int jmax, int P, int slot0,
                                               double4 *__restrict__ frow, double *sal = nullptr,
                                               double2 *scp = nullptr)
{
    const int lane = threadIdx.x & (kWave - 1);
    double z[K], q1[K], q2[K], eC[K], eS[K], oC[K], oS[K], r2[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int s = slot0 + k * kWave + lane;
        z[k] = (s < P) ? ringz[s] : 0.0;
        r2[k] = (kRho && s < P) ? ringr2[s] : 1.0;
        q1[k] = 0.0;
        q2[k] = kRho ? -sd / r2[k] : -sd;  // the first step then yields q_m = seed
        eC[k] = eS[k] = oC[k] = oS[k] = 0.0;
    }
    // explicit fused multiply-adds, so that every kernel built from this
    // tile rounds alike whatever the compiler would contract
    auto step = [&](double a, double z_, double qa, double qb, double r2_) {
        return kRho ? fma(a * z_, qa, -(r2_ * qb)) : fma(a * z_, qa, -qb);
    };
    // the rows of degrees [jb, je] (jb even) from a, c at offset jb, then
    // the last degree if it is even; j runs on across calls
    int j = 0;
    auto degrees = [&](const double *__restrict__ a, const double2 *__restrict__ c, int jb, int je) {
#pragma unroll 4
        for (; j + 1 <= je; j += 2) {
            const double a0 = a[j - jb], a1 = a[j - jb + 1];
            const double2 c0 = c[j - jb], c1 = c[j - jb + 1];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double q = step(a0, z[k], q1[k], q2[k], r2[k]);
                eC[k] = fma(q, c0.x, eC[k]);
                eS[k] = fma(q, c0.y, eS[k]);
                const double q3 = step(a1, z[k], q, q1[k], r2[k]);
                oC[k] = fma(q3, c1.x, oC[k]);
                oS[k] = fma(q3, c1.y, oS[k]);
                q2[k] = q;
                q1[k] = q3;
            }
        }
        if (je == jmax && j == jmax) {
            const double a0 = a[j - jb];
            const double2 c0 = c[j - jb];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const double q = step(a0, z[k], q1[k], q2[k], r2[k]);
                eC[k] = fma(q, c0.x, eC[k]);
                eS[k] = fma(q, c0.y, eS[k]);
            }
        }
    };
    if constexpr (kStage) {
        double pa = 0.0;
        double2 pc = make_double2(0.0, 0.0);
        if (lane <= jmax) {
            pa = al[lane];
            pc = cp[lane];
        }
        for (int jb = 0; jb <= jmax; jb += kWave) {
            __syncthreads();  // the previous block's reads are done
            sal[lane] = pa;
            scp[lane] = pc;
            __syncthreads();
            const int jn = jb + kWave + lane;
            if (jn <= jmax) {
                pa = al[jn];
                pc = cp[jn];
            }
            degrees(sal, scp, jb, min(jmax, jb + kWave - 1));
        }
    } else {
        degrees(al, cp, 0, jmax);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int s = slot0 + k * kWave + lane;
        if (s < P) frow[s] = make_double4(eC[k] + oC[k], eS[k] + oS[k], eC[k] - oC[k], eS[k] - oS[k]);
    }
}

template <int K, bool kRho = false, bool kStage = false>
__device__ __forceinline__ void synth_leg_dispatch(int kk, const double *__restrict__ al, double sd,
                                                   const double *__restrict__ ringz,
                                                   const double2 *__restrict__ cp, int jmax, int P, int slot0,
                                                   double4 *__restrict__ frow,
                                                   const double *__restrict__ ringr2 = nullptr,
                                                   double *sal = nullptr, double2 *scp = nullptr)
{
    static_assert(K == 4, "synthesis tiles are instantiated for 1 .. 4 ring pairs per lane");
    if (kk <= 1) synth_leg_tile<1, kRho, kStage>(al, sd, ringz, ringr2, cp, jmax, P, slot0, frow, sal, scp);
    else if (kk == 2) synth_leg_tile<2, kRho, kStage>(al, sd, ringz, ringr2, cp, jmax, P, slot0, frow, sal, scp);
    else if (kk == 3) synth_leg_tile<3, kRho, kStage>(al, sd, ringz, ringr2, cp, jmax, P, slot0, frow, sal, scp);
    else synth_leg_tile<4, kRho, kStage>(al, sd, ringz, ringr2, cp, jmax, P, slot0, frow, sal, scp);
}

// ------------------------------------------------- cross-lane reductions
//
// 64-bit lane moves and swaps from 32-bit ones (gfx950 v_permlane32_swap,
// v_permlane16_swap and DPP row moves applied to both halves).
__device__ __forceinline__ unsigned lo32(double v) { return (unsigned)__double_as_longlong(v); }
__device__ __forceinline__ unsigned hi32(double v) { return (unsigned)(__double_as_longlong(v) >> 32); }
__device__ __forceinline__ double from32(unsigned lo, unsigned hi)
{
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// a = [a lanes 0-31, b lanes 0-31], b = [a lanes 32-63, b lanes 32-63]
__device__ __forceinline__ void swap32_f64(double &a, double &b)
{
    const auto l = __builtin_amdgcn_permlane32_swap(lo32(a), lo32(b), false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(hi32(a), hi32(b), false, false);
    a = from32(l[0], h[0]);
    b = from32(l[1], h[1]);
}
// per 32 lanes: a = [a lanes 0-15, b lanes 0-15], b = [a lanes 16-31, b lanes 16-31]
__device__ __forceinline__ void swap16_f64(double &a, double &b)
{
    const auto l = __builtin_amdgcn_permlane16_swap(lo32(a), lo32(b), false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(hi32(a), hi32(b), false, false);
    a = from32(l[0], h[0]);
    b = from32(l[1], h[1]);
}
// DPP row moves: 0x140 row_mirror (lane i of a row of 16 reads 15 - i),
// 0x141 row_half_mirror (7 - i within 8), 0x4E / 0xB1 quad_perm xor 2 / 1
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int l = __builtin_amdgcn_update_dpp(0, (int)lo32(v), kCtrl, 0xF, 0xF, false);
    const int h = __builtin_amdgcn_update_dpp(0, (int)hi32(v), kCtrl, 0xF, 0xF, false);
    return from32((unsigned)l, (unsigned)h);
}

// Sums over the 64 lanes of a wave of 16 values per lane v[0..15], by a
// transpose-reduce in registers: each step pairs a lane with one that differs
// in one lane bit (5, 4: permlane swaps; 3, 2: row mirrors, which flip that
// bit; then quad moves for bits 1, 0), keeps half of its values and adds the
// partner's copies of them.  Returns the sum of value vi = (lane bit 2) +
// 2 (bit 3) + 4 (bit 4) + 8 (bit 5), which the 4 lanes of each quad hold.
__device__ __forceinline__ double wave_transpose_sum16(double (&v)[16], int lane)
{
    double a[8], b[4], c[2];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        swap32_f64(v[i], v[i + 8]);
        a[i] = v[i] + v[i + 8];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        swap16_f64(a[i], a[i + 4]);
        b[i] = a[i] + a[i + 4];
    }
    const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const double send = b3 ? b[i] : b[i + 2], keep = b3 ? b[i + 2] : b[i];
        c[i] = keep + dpp_f64<0x140>(send);
    }
    const double send = b2 ? c[0] : c[1], keep = b2 ? c[1] : c[0];
    double d = keep + dpp_f64<0x141>(send);
    d += dpp_f64<0x4E>(d);
    d += dpp_f64<0xB1>(d);
    return d;
}

// Analysis sums of order m over ring pairs [s0, P) (one wave):
// out[j] = sum over slots of q_{m+j}(z) * (E or O by parity of j), as (C, S)
// pairs.  grow holds the order's (E_C, E_S, O_C, O_S) already multiplied by
// the order's rescale factor.  Passes of K * 64 slots; the first pass
// stores, later passes accumulate.  Each chunk of kAnalLC degrees is reduced
// over the wave in registers (wave_transpose_sum16).
template <int K, bool kStage>
__device__ __forceinline__ void anal_leg_tile(const double *__restrict__ al, double sd,
                                              const double *__restrict__ ringz, const double4 *__restrict__ grow,
                                              int jmax, int s0, int P, double2 *__restrict__ out)
{
    constexpr int LC = kAnalLC;
    static_assert(kWave % LC == 0, "a chunk of degrees lies in one staged block");
    const int lane = threadIdx.x & (kWave - 1);
    // kStage: the recursion row in blocks of 64 degrees through LDS, as in
    // synth_leg_tile (one-wave workgroups)
    __shared__ double sal[kWave];
    double pre = 0.0;
    // the recursion coefficients of the chunk of degrees starting at j0
    auto chunk_row = [&](int j0) -> const double * {
        if constexpr (kStage) {
            if ((j0 & (kWave - 1)) == 0) {
                __syncthreads();  // the previous block's reads are done
                sal[lane] = pre;
                __syncthreads();
                const int jn = j0 + kWave + lane;
                pre = (jn <= jmax) ? al[jn] : 0.0;
            }
            return sal + (j0 & (kWave - 1));
        } else {
            return al + j0;
        }
    };
    for (int p0 = s0; p0 < P; p0 += K * kWave) {
        if constexpr (kStage) pre = (lane <= jmax) ? al[lane] : 0.0;
        double z[K], q1[K], q2[K], eC[K], eS[K], oC[K], oS[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int s = p0 + k * kWave + lane;
            double4 g = make_double4(0.0, 0.0, 0.0, 0.0);
            double zz = 0.0;
            if (s < P) {
                g = grow[s];
                zz = ringz[s];
            }
            // a ring whose data vanish (rescale underflow, zero weight) keeps
            // a bounded recursion so that 0 * q stays 0
            if (g.x == 0.0 && g.y == 0.0 && g.z == 0.0 && g.w == 0.0) zz = 0.0;
            z[k] = zz;
            eC[k] = g.x;
            eS[k] = g.y;
            oC[k] = g.z;
            oS[k] = g.w;
            q1[k] = 0.0;
            q2[k] = -sd;
        }
        // Each chunk of LC degrees: K recursion steps per lane and degree, the
        // first ring's product initialising the accumulators.  Full chunks
        // run without guards so the recursion state rotates in registers.
        static_assert(2 * LC == 16, "the register reduction takes 16 values per lane");
        auto reduce_chunk = [&](double (&acc)[LC][2], int j0) {
            double v[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) v[i] = acc[i >> 1][i & 1];
            const double sum = wave_transpose_sum16(v, lane);
            const int vi = ((lane >> 2) & 1) | ((lane >> 2) & 2) | ((lane >> 2) & 4) | ((lane >> 2) & 8);
            const int jj = vi >> 1;
            if ((lane & 3) == 0 && j0 + jj <= jmax) {
                double *o = reinterpret_cast<double *>(out + j0 + jj) + (vi & 1);
                *o = (p0 == s0) ? sum : *o + sum;
            }
        };
        int j0 = 0;
        for (; j0 + LC - 1 <= jmax; j0 += LC) {
            const double *ab = chunk_row(j0);
            double acc[LC][2];
#pragma unroll
            for (int jj = 0; jj < LC; ++jj) {
                const double a = ab[jj];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    // explicit fused multiply-adds, as in the synthesis tile
                    const double q = fma(a * z[k], q1[k], -q2[k]);
                    q2[k] = q1[k];
                    q1[k] = q;
                    const double ec = (jj % 2 == 0) ? eC[k] : oC[k];
                    const double es = (jj % 2 == 0) ? eS[k] : oS[k];
                    if (k == 0) {
                        acc[jj][0] = q * ec;
                        acc[jj][1] = q * es;
                    } else {
                        acc[jj][0] = fma(q, ec, acc[jj][0]);
                        acc[jj][1] = fma(q, es, acc[jj][1]);
                    }
                }
            }
            reduce_chunk(acc, j0);
        }
        if (j0 <= jmax) {
            const double *ab = chunk_row(j0);
            double acc[LC][2];
#pragma unroll
            for (int jj = 0; jj < LC; ++jj) acc[jj][0] = acc[jj][1] = 0.0;
#pragma unroll
            for (int jj = 0; jj < LC; ++jj) {
                if (j0 + jj <= jmax) {
                    const double a = ab[jj];
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const double q = fma(a * z[k], q1[k], -q2[k]);
                        q2[k] = q1[k];
                        q1[k] = q;
                        acc[jj][0] = fma(q, (jj % 2 == 0) ? eC[k] : oC[k], acc[jj][0]);
                        acc[jj][1] = fma(q, (jj % 2 == 0) ? eS[k] : oS[k], acc[jj][1]);
                    }
                }
            }
            reduce_chunk(acc, j0);
        }
    }
}

// ---------------------------------------------------------------- ring FFTs
//
// A ring-pair FFT is split into load (global inputs into registers), fill
// (all m work slots of the LDS array written once: data and zeros, no
// barrier inside), and run (the DFT and the global outputs), so that a
// workgroup transforming several slots can load the next slot's inputs
// while the current one is in its FFT stages.

// The orders a ring slot carries: Lc, or fewer near the poles (kmax: per slot
// the largest order whose Legendre functions are not negligible there;
// nullptr: all).
__device__ __forceinline__ int slot_orders(const int *__restrict__ kmax, int s, int Lc)
{
    return kmax ? min(Lc, kmax[s]) : Lc;
}

// Synthesis inputs of one thread: ring sums of orders threadIdx.x + j *
// blockDim.x, j < kSynthPre (later orders are read in the fill).  Orders
// interleave over the threads so that the fill's LDS writes of a wave touch
// consecutive elements.
constexpr int kSynthPre = 2;
struct SynthRing {
    double4 f[kSynthPre];
};

__device__ __forceinline__ void synth_fft_load(const double4 *__restrict__ fbuf, int P, int Lc, int s,
                                               SynthRing &in)
{
#pragma unroll
    for (int j = 0; j < kSynthPre; ++j) {
        const int k = threadIdx.x + j * blockDim.x;
        in.f[j] = (k <= Lc) ? fbuf[(size_t)k * P + s] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
}

// Hermitian fill of slot s (fbuf rows of order m at stride P hold unscaled
// sums; u = sin(theta) of the slot gives the orders' rescale).
__device__ __forceinline__ void synth_fft_fill(const FftDev &F, const SynthRing &in,
                                               const double4 *__restrict__ fbuf, int P, int Lc, int s, double u,
                                               double2 *lds)
{
    const int n = F.n, m = F.m;
    // data occupy work positions [0, Lc] and [top - Lc, top): negative orders
    // sit at the top of the array in pruned plans, at n - k otherwise
    const int top = (F.lc >= 0) ? m : n;
    for (int q = threadIdx.x; q < m; q += blockDim.x)
        if ((q > Lc && q < top - Lc) || q >= top) lds[fpad(q)] = make_double2(0.0, 0.0);
    // order k carries the extended-range power 1e280 u^k (the reference's
    // iterated rescalem)
    auto put = [&](int k, double r, double4 f) {
        // Z_k = X_k + i Y_k with X (north) and Y (south) Hermitian:
        // X_0 = C_N, X_k = (C_N - i S_N)/2 and X_{n-k} = conj(X_k); stored
        // conjugated (inverse DFT via conj(F(conj Z))).
        f.x *= r;
        f.y *= r;
        f.z *= r;
        f.w *= r;
        if (k == 0) {
            lds[0] = make_double2(f.x, -f.z);  // chirp_0 = 1
        } else {
            const double2 c = fft_chirp(F, k);
            lds[fpad(k)] = fft_prep_in_c(F, make_double2(0.5 * (f.x + f.w), -0.5 * (f.z - f.y)), c);
            lds[fpad(fft_pos(F, n - k))] =
                fft_prep_in_c(F, make_double2(0.5 * (f.x - f.w), -0.5 * (f.y + f.z)), fft_chirp_mirror(F, c));
        }
    };
#pragma unroll
    for (int j = 0; j < kSynthPre; ++j) {
        const int k = threadIdx.x + j * blockDim.x;
        if (k <= Lc) put(k, ring_rescale(u, k), in.f[j]);
    }
    for (int k = threadIdx.x + kSynthPre * blockDim.x; k <= Lc; k += blockDim.x)
        put(k, ring_rescale(u, k), fbuf[(size_t)k * P + s]);
}

// DFT of the filled array and the rows out; the north and south rows are
// scaled by rwt (1 for synthesis, the quadrature weights for the adjoint of
// analysis).
//
// chunk != nullptr: lds is a work array in global memory and the DFT runs
// its inner stages in the LDS chunk (chunk_cap elements) when a group of the
// plan fits there.
// Split plans (FftDev::half): the inverse stages run once per half of the
// output columns, the forward spectrum kept in registers in between.
__device__ __forceinline__ void synth_fft_run_split(const FftDev &F, int2 rw, int nrows, double *__restrict__ grid,
                                                    long long ld, int ext_col, double2 rwt, double2 *lds)
{
    const int n = F.n, h = n / 2;
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
    TurnStash keep;
    fft_dif_front(lds, F);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        if (half == 0)
            fft_run_turn_split<kSynth1>(lds, F, F.bhat, keep);
        else
            fft_run_turn_split<kSynth2>(lds, F, F.bhat3, keep);
        __syncthreads();
        fft_dit_back(lds, F);
        // columns j = half h + p at work positions p
        for (int p = threadIdx.x; p < h; p += blockDim.x) {
            const int j = half * h + p;
            const double2 v = fft_get_c(lds, F, j, p, gld(F.chirp + j));
            const double gn = v.x * rwt.x, gs = -v.y * rwt.y;
            if (north) gst(grid + rw.x * ld + j, gn);
            if (south) gst(grid + rw.y * ld + j, gs);
            if (ext_col && j == 0) {
                if (north) gst(grid + rw.x * ld + n, gn);
                if (south) gst(grid + rw.y * ld + n, gs);
            }
        }
        if (half == 0) __syncthreads();  // the second turn overwrites the array
    }
}

//
// kVar: the kernel instantiation by plan kind (kVarPlain, kVarSplit for
// split plans, kVarPrime for direct plans ending with a prime radix); each
// kind's code and registers stay out of the other instantiations.
template <int kVar = kVarPlain>
__device__ __forceinline__ void synth_fft_run(const FftDev &F, int2 rw, int nrows, double *__restrict__ grid,
                                              long long ld, int ext_col, double2 rwt, double2 *lds,
                                              double2 *chunk = nullptr, int chunk_cap = 0)
{
    if constexpr (kVar == kVarSplit) {
        synth_fft_run_split(F, rw, nrows, grid, ld, ext_col, rwt, lds);
        return;
    }
    const int n = F.n;
    if (chunk && F.nst >= 2 && F.m / F.radix[0] <= chunk_cap)
        fft_forward_chunked(lds, F, F.bhat, chunk, chunk_cap);
    else
        fft_forward<kVar == kVarPrime, kVar == kVarPow2>(lds, F, F.bhat);
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
    // columns j and n - j (1 <= j < n - j) together: one chirp read
    for (int j = threadIdx.x; 2 * j <= n; j += blockDim.x) {
        const double2 c = fft_chirp(F, j);
        const double2 v = fft_get_c(lds, F, j, j, c);
        const double gn = v.x * rwt.x, gs = -v.y * rwt.y;
        if (north) gst(grid + rw.x * ld + j, gn);
        if (south) gst(grid + rw.y * ld + j, gs);
        if (ext_col && j == 0) {
            if (north) gst(grid + rw.x * ld + n, gn);
            if (south) gst(grid + rw.y * ld + n, gs);
        }
        const int j2 = n - j;
        if (j > 0 && j2 > j) {
            const double2 v2 = fft_get_c(lds, F, j2, j2, fft_chirp_mirror(F, c));
            if (north) gst(grid + rw.x * ld + j2, v2.x * rwt.x);
            if (south) gst(grid + rw.y * ld + j2, -v2.y * rwt.y);
        }
    }
}

template <int kVar = kVarPlain>
__device__ __forceinline__ void synth_fft_tile(const FftDev &F, const double4 *__restrict__ fbuf, int P, int Lc,
                                               int s, double u, int2 rw, int nrows, double *__restrict__ grid,
                                               long long ld, int ext_col, double2 rwt, double2 *lds,
                                               double2 *chunk = nullptr, int chunk_cap = 0)
{
    SynthRing in;
    synth_fft_load(fbuf, P, Lc, s, in);
    synth_fft_fill(F, in, fbuf, P, Lc, s, u, lds);
    __syncthreads();
    synth_fft_run<kVar>(F, rw, nrows, grid, ld, ext_col, rwt, lds, chunk, chunk_cap);
}

// Analysis inputs of one thread: ring values at the column pairs q and
// n - q, q = threadIdx.x + j * blockDim.x, j < kAnalPre / 2 (later columns
// are read in the fill).
constexpr int kAnalPre = 4;
struct AnalRing {
    double2 v[kAnalPre];
};

// (north, south) values of column k; ext_col adds column n to column 0 (the
// adjoint of a grid's repeated first column).
__device__ __forceinline__ double2 anal_column(const double *__restrict__ grid, long long ld, int2 rw, bool north,
                                               bool south, int n, int ext_col, int k)
{
    double2 v;
    v.x = north ? gld(grid + rw.x * ld + k) : 0.0;
    v.y = south ? gld(grid + rw.y * ld + k) : 0.0;
    if (ext_col && k == 0) {
        if (north) v.x += gld(grid + rw.x * ld + n);
        if (south) v.y += gld(grid + rw.y * ld + n);
    }
    return v;
}

__device__ __forceinline__ void anal_fft_load(const FftDev &F, const double *__restrict__ grid, long long ld, int2 rw,
                                              int nrows, int ext_col, AnalRing &in)
{
    const int n = F.n;
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
#pragma unroll
    for (int j = 0; j < kAnalPre / 2; ++j) {
        const int q = threadIdx.x + j * blockDim.x, q2 = n - q;
        const bool lo = 2 * q <= n, hi = q > 0 && q2 > q;
        in.v[2 * j] = lo ? anal_column(grid, ld, rw, north, south, n, ext_col, q) : make_double2(0.0, 0.0);
        in.v[2 * j + 1] = hi ? anal_column(grid, ld, rw, north, south, n, ext_col, q2) : make_double2(0.0, 0.0);
    }
}

__device__ __forceinline__ void anal_fft_fill(const FftDev &F, const AnalRing &in, const double *__restrict__ grid,
                                              long long ld, int2 rw, int nrows, int ext_col, double2 *lds)
{
    const int n = F.n, m = F.m, nt = blockDim.x;
    // columns q and n - q (1 <= q < n - q) together: one chirp read
    auto put = [&](int q, double2 a, double2 b) {
        const double2 c = fft_chirp(F, q);
        lds[fpad(q)] = fft_prep_in_c(F, a, c);
        const int q2 = n - q;
        if (q > 0 && q2 > q) lds[fpad(q2)] = fft_prep_in_c(F, b, fft_chirp_mirror(F, c));
    };
#pragma unroll
    for (int j = 0; j < kAnalPre / 2; ++j) {
        const int q = threadIdx.x + j * nt;
        if (2 * q <= n) put(q, in.v[2 * j], in.v[2 * j + 1]);
    }
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
    for (int q = threadIdx.x + (kAnalPre / 2) * nt; 2 * q <= n; q += nt) {
        const int q2 = n - q;
        put(q, anal_column(grid, ld, rw, north, south, n, ext_col, q),
            q2 > q ? anal_column(grid, ld, rw, north, south, n, ext_col, q2) : make_double2(0.0, 0.0));
    }
    for (int q = n + threadIdx.x; q < m; q += nt) lds[fpad(q)] = make_double2(0.0, 0.0);
}

// The orders 0 .. Lc of the transformed ring pair: split into the two rings,
// weighted, times the order's rescale factor, into gbuf as the even and odd
// combinations of the two rings (u < 0: no rescale, the two rings apart).
__device__ __forceinline__ void anal_fft_extract(const FftDev &F, int2 rw, int nrows, double2 w, double u, int P,
                                                 int Lc, int s, double4 *__restrict__ gbuf, const double2 *lds)
{
    const int n = F.n;
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
    const double wn = north ? w.x : 0.0, ws = south ? w.y : 0.0;
    // orders interleave over the threads (a wave reads consecutive elements)
    for (int mm = threadIdx.x; mm <= Lc; mm += blockDim.x) {
        const double r = ring_rescale(u, mm);
        const double2 c = fft_chirp(F, mm);
        const double2 zm = fft_get_c(lds, F, mm, mm, c);
        const int nm = mm == 0 ? 0 : n - mm;
        const double2 zn = mm == 0 ? zm : fft_get_c(lds, F, nm, fft_pos(F, nm), fft_chirp_mirror(F, c));
        const double xr = 0.5 * (zm.x + zn.x), xi = 0.5 * (zm.y - zn.y);
        const double yr = 0.5 * (zm.y + zn.y), yi = 0.5 * (zn.x - zm.x);
        const double f1c = wn * xr, f1s = -wn * xi, f2c = ws * yr, f2s = -ws * yi;
        if (u < 0.0)
            gbuf[(size_t)mm * P + s] = make_double4(f1c, f1s, f2c, f2s);
        else
            gbuf[(size_t)mm * P + s] =
                make_double4((f1c + f2c) * r, (f1s + f2s) * r, (f1c - f2c) * r, (f1s - f2s) * r);
    }
}

// DFT of the filled array; w = (w_north, w_south) include the quadrature
// weight and the reference's constant factors; the outputs carry the
// order's rescale factor for the ring (sin(theta) = u).
template <int kVar = kVarPlain>
__device__ __forceinline__ void anal_fft_run(const FftDev &F, int2 rw, int nrows, double2 w, double u, int P, int Lc,
                                             int s, double4 *__restrict__ gbuf, double2 *lds,
                                             double2 *chunk = nullptr, int chunk_cap = 0)
{
    if (chunk && F.nst >= 2 && F.m / F.radix[0] <= chunk_cap)
        fft_forward_chunked(lds, F, F.bhat2, chunk, chunk_cap);
    else
        fft_forward<kVar == kVarPrime, kVar == kVarPow2>(lds, F, F.bhat2);
    anal_fft_extract(F, rw, nrows, w, u, P, Lc, s, gbuf, lds);
}

// Split plans (FftDev::half): each half of the ring's columns is filled and
// transformed in turn; the first half's product with its chirp spectrum is
// kept in registers and added in the second turn, then the inverse stages
// run once and the orders are extracted as for the whole ring.
__device__ __forceinline__ void anal_fft_tile_split(const FftDev &F, const double *__restrict__ grid, long long ld,
                                                    int2 rw, int nrows, double2 w, double u, int P, int Lc, int s,
                                                    double4 *__restrict__ gbuf, int ext_col, double2 *lds)
{
    const int n = F.n, h = n / 2, m = F.m;
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
    TurnStash keep;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        for (int q = threadIdx.x; q < m; q += blockDim.x) {
            double2 v = make_double2(0.0, 0.0);
            if (q < h) {
                const int col = half * h + q;
                v = fft_prep_in_c(F, anal_column(grid, ld, rw, north, south, n, ext_col, col), gld(F.chirp + col));
            }
            lds[fpad(q)] = v;
        }
        __syncthreads();
        fft_dif_front(lds, F);
        if (half == 0) {
            fft_run_turn_split<kAnal1>(lds, F, F.bhat2, keep);
            __syncthreads();  // the second half's fill overwrites the array
        } else {
            fft_run_turn_split<kAnal2>(lds, F, F.bhat4, keep);
            __syncthreads();
        }
    }
    fft_dit_back(lds, F);
    anal_fft_extract(F, rw, nrows, w, u, P, Lc, s, gbuf, lds);
}

template <int kVar = kVarPlain>
__device__ __forceinline__ void anal_fft_tile(const FftDev &F, const double *__restrict__ grid, long long ld,
                                              int2 rw, int nrows, double2 w, double u, int P, int Lc, int s,
                                              double4 *__restrict__ gbuf, int ext_col, double2 *lds,
                                              double2 *chunk = nullptr, int chunk_cap = 0)
{
    if constexpr (kVar == kVarSplit) {
        anal_fft_tile_split(F, grid, ld, rw, nrows, w, u, P, Lc, s, gbuf, ext_col, lds);
        return;
    }
    AnalRing in;
    anal_fft_load(F, grid, ld, rw, nrows, ext_col, in);
    anal_fft_fill(F, in, grid, ld, rw, nrows, ext_col, lds);
    __syncthreads();
    anal_fft_run<kVar>(F, rw, nrows, w, u, P, Lc, s, gbuf, lds, chunk, chunk_cap);
}

// =============================================================== plain kernels

// grid: ntm x ntl tiles (blockIdx.x = tile_l * ntm + tile_m)
__global__ __launch_bounds__(kThreads) void k_pack_synth(const double *__restrict__ cilm, int d1, int d2,
                                                         const double *__restrict__ s_tab, int Lt,
                                                         double2 *__restrict__ cpack, int Lc, int ntm)
{
    __shared__ CoefTile t;
    const int tl = blockIdx.x / ntm, tm = blockIdx.x - tl * ntm;
    const int m0 = tm * kTM, l0 = tl * kTL;
    if (l0 + kTL - 1 < m0) return;  // above the diagonal: nothing to pack
    pack_tile(cilm, d1, d2, Lc, s_tab, Lt, cpack, m0, l0, t);
}

// grid: (Lc+1) * nsb single-wave workgroups, order m = blockIdx / nsb, slot
// block sb of kSynthKMax*64 ring pairs; the last block takes as many ring
// pairs per lane as its slots need.
// ringr2: per-slot rho^2 of a radially continued synthesis (nullptr: none);
// smin: per order the first slot with non-negligible functions (nullptr: 0).
// Workgroup b of a plain Legendre launch: order m and slot block sb, with
// the nb blocks of an order on one XCD (workgroups go to the 8 XCDs in turn,
// b mod 8), so that their reads of the order's recursion and coefficient
// rows share that XCD's L2; orders are dealt to the XCDs in turn, lowest
// (longest) first.  Grid: 8 nb ceil((Lc+1)/8) workgroups; false for the
// padding ones (m > Lc).
__device__ __forceinline__ bool order_block(int b, int nb, int Lc, int &m, int &sb)
{
    const int t = b >> 3;
    sb = t % nb;
    m = 8 * (t / nb) + (b & 7);
    return m <= Lc;
}

// kStage: the rows through LDS (synth_leg_tile), for launches that bring
// few ring pairs per lane.
template <bool kStage>
__global__ __launch_bounds__(kWave) void k_synth_leg(const double *__restrict__ alpha, int Lt,
                                                     const double *__restrict__ seed,
                                                     const double *__restrict__ ringz,
                                                     const double2 *__restrict__ cpack, int Lc, int P, int nsb,
                                                     double4 *__restrict__ fbuf,
                                                     const double *__restrict__ ringr2,
                                                     const int *__restrict__ smin, int kb)
{
    int m, sb;
    if (!order_block(blockIdx.x, nsb, Lc, m, sb)) return;
    const int slot0 = (smin ? smin[m] : 0) + sb * kb * kWave;
    if (slot0 >= P) return;
    const int kk = min(kb, (P - slot0 + kWave - 1) / kWave);
    const double *al = alpha + tri_offd(m, Lt);
    const double2 *cp = cpack + tri_offd(m, Lc);
    // one pair of staging blocks for every tile instantiation of the kernel
    __shared__ double sal[kStage ? kWave : 1];
    __shared__ double2 scp[kStage ? kWave : 1];
    if (ringr2)
        synth_leg_dispatch<kSynthKMax, true, kStage>(kk, al, seed[m], ringz, cp, Lc - m, P, slot0,
                                                     fbuf + (size_t)m * P, ringr2, sal, scp);
    else
        synth_leg_dispatch<kSynthKMax, false, kStage>(kk, al, seed[m], ringz, cp, Lc - m, P, slot0,
                                                      fbuf + (size_t)m * P, nullptr, sal, scp);
}

// rows[s] = (north row, south row or -1); rows >= nrows are skipped;
// roww[s]: row weights (nullptr: 1).
template <int kVar>
__global__ __launch_bounds__(fft_max_threads(kVar)) void k_synth_fft(FftDev F, const double4 *__restrict__ fbuf, int P, int Lc,
                                                        const double *__restrict__ ringu,
                                                        const int2 *__restrict__ rows, int nrows,
                                                        double *__restrict__ grid, long long ld, int ext_col,
                                                        const double2 *__restrict__ roww,
                                                        const int *__restrict__ kmax)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x;
    synth_fft_tile<kVar>(F, fbuf, P, slot_orders(kmax, s, Lc), s, ringu ? ringu[s] : -1.0, rows[s], nrows, grid, ld,
                         ext_col, roww ? roww[s] : make_double2(1.0, 1.0), lds);
}

template <int kVar>
__global__ __launch_bounds__(fft_max_threads(kVar)) void k_anal_fft(FftDev F, const double *__restrict__ grid, long long ld,
                                                       const double *__restrict__ ringu,
                                                       const int2 *__restrict__ rows, int nrows,
                                                       const double2 *__restrict__ wts, int P, int Lc,
                                                       double4 *__restrict__ gbuf, int ext_col,
                                                       const int *__restrict__ kmax)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x;
    anal_fft_tile<kVar>(F, grid, ld, rows[s], nrows, wts[s], ringu ? ringu[s] : -1.0, P, slot_orders(kmax, s, Lc), s,
                        gbuf, ext_col, lds);
}

// Rings whose FFT work array does not fit the LDS of a CU: the same tiles
// on a per-workgroup slice (ws_stride double2) of a global workspace, each
// workgroup taking slots blockIdx.x, blockIdx.x + gridDim.x, ...
__global__ __launch_bounds__(kFftMaxThreads) void k_synth_fft_ws(FftDev F, const double4 *__restrict__ fbuf, int P,
                                                           int Lc, const double *__restrict__ ringu,
                                                           const int2 *__restrict__ rows, int nrows,
                                                           double *__restrict__ grid, long long ld, int ext_col,
                                                           const double2 *__restrict__ roww, double2 *ws,
                                                           long long ws_stride, const int *__restrict__ kmax,
                                                           int chunk_cap)
{
    extern __shared__ double2 chunk[];
    double2 *x = ws + blockIdx.x * ws_stride;
    for (int s = blockIdx.x; s < P; s += gridDim.x) {
        synth_fft_tile(F, fbuf, P, slot_orders(kmax, s, Lc), s, ringu ? ringu[s] : -1.0, rows[s], nrows, grid, ld,
                       ext_col, roww ? roww[s] : make_double2(1.0, 1.0), x, chunk, chunk_cap);
        __syncthreads();
    }
}

__global__ __launch_bounds__(kFftMaxThreads) void k_anal_fft_ws(FftDev F, const double *__restrict__ grid,
                                                          long long ld, const double *__restrict__ ringu,
                                                          const int2 *__restrict__ rows, int nrows,
                                                          const double2 *__restrict__ wts, int P, int Lc,
                                                          double4 *__restrict__ gbuf, int ext_col, double2 *ws,
                                                          long long ws_stride, const int *__restrict__ kmax,
                                                          int chunk_cap)
{
    extern __shared__ double2 chunk[];
    double2 *x = ws + blockIdx.x * ws_stride;
    for (int s = blockIdx.x; s < P; s += gridDim.x) {
        anal_fft_tile(F, grid, ld, rows[s], nrows, wts[s], ringu ? ringu[s] : -1.0, P, slot_orders(kmax, s, Lc), s,
                      gbuf, ext_col, x, chunk, chunk_cap);
        __syncthreads();
    }
}

// one single-wave workgroup per order m over ring pairs [s0, P), from the
// order's first slot with non-negligible functions (smin; nullptr: s0)
//
// nblk > 1 splits each order's ring pairs into blocks of bslots slots (a
// multiple of K*64) from its first slot: workgroup m * nblk + b sums block b
// into cbuf + b * cstride (zeros when the order has no slots there), and the
// unpack adds the blocks in order.  Degrees far above the GPU's wave count
// then no longer wait on one wave per order.
template <int K>
__global__ __launch_bounds__(kWave) void k_anal_leg(const double *__restrict__ alpha, int Lt,
                                                    const double *__restrict__ seed,
                                                    const double *__restrict__ ringz,
                                                    const double4 *__restrict__ gbuf, int Lc, int s0, int P,
                                                    double2 *__restrict__ cbuf, const int *__restrict__ smin,
                                                    int nblk, int bslots, long long cstride)
{
    int m, b;
    if (!order_block(blockIdx.x, nblk, Lc, m, b)) return;
    const int sa = (smin ? max(s0, smin[m]) : s0) + b * bslots;
    const int se = nblk > 1 ? min(P, sa + bslots) : P;
    double2 *out = cbuf + b * cstride + tri_offd(m, Lc);
    if (sa >= se) {
        for (int j = threadIdx.x; j <= Lc - m; j += kWave) out[j] = make_double2(0.0, 0.0);
        return;
    }
    anal_leg_tile<K, true>(alpha + tri_offd(m, Lt), seed[m], ringz, gbuf + (size_t)m * P, Lc - m, sa, se, out);
}

// grid: ntm x ntl tiles over the whole d1 x d2 array
__global__ __launch_bounds__(kThreads) void k_unpack_anal(const double2 *__restrict__ cbuf,
                                                          const double *__restrict__ post, int Lt, int Lc,
                                                          double *__restrict__ cilm, int d1, int d2, int ntm,
                                                          int nblk, long long cstride)
{
    __shared__ CoefTile t;
    const int tl = blockIdx.x / ntm, tm = blockIdx.x - tl * ntm;
    unpack_tile(cbuf, post, Lt, Lc, cilm, d1, d2, tm * kTM, tl * kTL, t, nblk, cstride);
}

// Zero the part of a rows x cols array (leading dimension ld) outside its
// leading r0 x c0 block.
__global__ __launch_bounds__(kThreads) void k_zero_outside(double *__restrict__ a, int rows, int cols, long long ld,
                                                           int r0, int c0)
{
    const long long total = (long long)rows * cols;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(e / cols), c = (int)(e - (long long)r * cols);
        if (r >= r0 || c >= c0) a[r * ld + c] = 0.0;
    }
}

// Set every element of a rows x cols array (leading dimension ld) to
// src[0] * v (src == nullptr: v).
__global__ __launch_bounds__(kThreads) void k_fill(double *__restrict__ a, int rows, int cols, long long ld,
                                                   const double *__restrict__ src, double v)
{
    if (src) v = src[0] * v;
    const long long total = (long long)rows * cols;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(e / cols), c = (int)(e - (long long)r * cols);
        a[r * ld + c] = v;
    }
}

// ===================================================== precomputed tables
//
// The plx branch of the GLQ routines: a table of the Legendre functions at
// the Gauss-Legendre nodes, the Fortran array plx(n, (L+1)(L+2)/2) of
// src/SHGLQ.f95, here plx[k * ld + i] for node i and k = l(l+1)/2 + m
// (ld = n = L + 1), read by the synthesis and analysis in place of the
// on-the-fly recursion.  The ring FFTs are the transforms' own, told by a
// negative ring u that the sums are actual function values (no order
// rescale) and, in analysis, to keep the two rings of a slot apart.

// The table (src/SHGLQ.f95:202-257): one wave per (m, block of 64 northern
// nodes i < (n+1)/2), each lane the recursion at its node; node n-1-i takes
// the values times (-1)^(l-m).  nblk blocks per order.
__global__ __launch_bounds__(kWave) void k_plx_table(const double *__restrict__ alpha,
                                                     const double *__restrict__ stab, int Lt,
                                                     const double *__restrict__ seed,
                                                     const double *__restrict__ zero, int n, int L, int nblk,
                                                     double *__restrict__ plx)
{
    const int m = blockIdx.x / nblk, b = blockIdx.x - m * nblk;
    const int i = b * kWave + (int)threadIdx.x;
    if (i >= (n + 1) / 2) return;
    const double z = zero[i];
    const double u = sqrt((1.0 - z) * (1.0 + z));
    const double r = order_rescale(u, m);
    const double zz = (r == 0.0) ? 0.0 : z;
    const double *__restrict__ al = alpha + tri_offd(m, Lt);
    const double *__restrict__ sl = stab + tri_offd(m, Lt);
    const int is = n - 1 - i;
    double q1 = 0.0, q2 = -seed[m];  // the first step yields the seed
    for (int l = m; l <= L; ++l) {
        const double q = (al[l - m] * zz) * q1 - q2;
        q2 = q1;
        q1 = q;
        const double p = (r == 0.0) ? 0.0 : sl[l - m] * q * r;
        const size_t k = (size_t)l * (l + 1) / 2 + m;
        plx[k * n + i] = p;
        if (is != i) plx[k * n + is] = ((l - m) & 1) ? -p : p;
    }
}

// Synthesis sums from the table (src/MakeGridGLQ.f95:441-465): one wave per
// (m, block of 64 slots), lanes = slots; fbuf[m][s] = (sum C p, sum S p) of
// the slot's north row and of its south row, degrees from Lc down to m.
__global__ __launch_bounds__(kWave) void k_plx_synth_leg(const double *__restrict__ cilm, int d1, int d2, int Lc,
                                                         const double *__restrict__ plx, int ld,
                                                         const int2 *__restrict__ rows, int P, int nblk,
                                                         double4 *__restrict__ fbuf)
{
    const int m = blockIdx.x / nblk, b = blockIdx.x - m * nblk;
    const int s = b * kWave + (int)threadIdx.x;
    if (s >= P) return;
    const int2 rw = rows[s];
    const bool south = rw.y >= 0;
    const double *__restrict__ cc = cilm + m;
    const double *__restrict__ cs = cilm + (size_t)d1 * d2 + m;
    double cn = 0.0, sn = 0.0, csum = 0.0, ssum = 0.0;
    for (int l = Lc; l >= m; --l) {
        const size_t k = (size_t)l * (l + 1) / 2 + m;
        const double C = cc[(size_t)l * d2], S = (m > 0) ? cs[(size_t)l * d2] : 0.0;
        const double pn = plx[k * ld + rw.x];
        cn = fma(C, pn, cn);
        sn = fma(S, pn, sn);
        if (south) {
            const double ps = plx[k * ld + rw.y];
            csum = fma(C, ps, csum);
            ssum = fma(S, ps, ssum);
        }
    }
    fbuf[(size_t)m * P + s] = make_double4(cn, sn, csum, ssum);
}

// The analysis epilogue of the table's coefficient (l, m): divided by the
// integral of the squared function (src/SHExpandGLQ.f95:659-689; the
// complex transforms, src/SHExpandGLQC.f95, omit the halving of the
// unnormalised product for m >= 1), in the reference's operations.
__device__ __forceinline__ double plx_normalise(double c, int norm, int l, int m, bool cplx)
{
    const double pi = 3.14159265358979323846;
    switch (norm) {
    case 1:
        return c / (4 * pi);
    case 2:
        return (2 * l + 1) * c / (4 * pi);
    case 3: {
        double prod = 4 * pi / (double)(2 * l + 1);
        if (m == 0) return c / prod;
        if (!cplx) prod = prod / 2.0;
        for (int j = 1; j <= m && j <= l - 1; ++j) prod = prod * (l + j) * (l - j + 1);
        return c / ((m == l) ? prod * 2 * l : prod);
    }
    default:
        return c;
    }
}

// Analysis from the table (src/SHExpandGLQ.f95:419-441 and the epilogue
// :659-689): one wave per (m, block of 64 degrees), lanes = degrees l;
// gbuf[m][s] holds the weighted Fourier coefficients of the slot's north
// and south rows apart, summed over the rows in order as the reference does,
// then divided by the normalisation's integral.  nblk blocks per order;
// cilm is zero outside what this writes.
__global__ __launch_bounds__(kWave) void k_plx_anal_leg(const double4 *__restrict__ gbuf, int P, int nlat, int Lc,
                                                        const double *__restrict__ plx, int ld, int norm,
                                                        int nblk, double *__restrict__ cilm, int d1, int d2)
{
    const int m = blockIdx.x / nblk, b = blockIdx.x - m * nblk;
    const int l = m + b * kWave + (int)threadIdx.x;
    if (l > Lc) return;
    const double *__restrict__ row = plx + ((size_t)l * (l + 1) / 2 + m) * ld;
    const double4 *__restrict__ g = gbuf + (size_t)m * P;
    double c = 0.0, sv = 0.0;
    for (int r = 0; r < nlat; ++r) {
        // row r is the north row of slot r or the south row of slot nlat-1-r
        const bool north = r < P;
        const double4 f = g[north ? r : nlat - 1 - r];
        const double p = row[r];
        c = fma(p, north ? f.x : f.z, c);
        sv = fma(p, north ? f.y : f.w, sv);
    }
    cilm[(size_t)l * d2 + m] = plx_normalise(c, norm, l, m, false);
    // S_l0 stays as summed: the reference divides only C_l0 for norm 3
    cilm[((size_t)d1 + l) * d2 + m] = (m == 0 && norm == 3) ? sv : plx_normalise(sv, norm, l, m, false);
}

// MakeGridGLQC with the table (src/MakeGridGLQC.f95:443-468): as
// k_plx_synth_leg for complex coefficients cilm[2][d1][d2]; fb_p[m][s] the
// sums of plane 0 (orders +m), fb_m[m][s] of plane 1 (orders -m, m >= 1;
// the ring FFT applies their (-1)^m), each (Re, Im) of the north row and of
// the south row.
__global__ __launch_bounds__(kWave) void k_plx_synth_leg_c(const double2 *__restrict__ cilm, int d1, int d2, int Lc,
                                                           const double *__restrict__ plx, int ld,
                                                           const int2 *__restrict__ rows, int P, int nblk,
                                                           double4 *__restrict__ fb_p, double4 *__restrict__ fb_m)
{
    const int m = blockIdx.x / nblk, b = blockIdx.x - m * nblk;
    const int s = b * kWave + (int)threadIdx.x;
    if (s >= P) return;
    const int2 rw = rows[s];
    const bool south = rw.y >= 0;
    const double2 *__restrict__ cp = cilm + m;
    const double2 *__restrict__ cm = cilm + (size_t)d1 * d2 + m;
    double4 ap = make_double4(0.0, 0.0, 0.0, 0.0), am = ap;
    for (int l = Lc; l >= m; --l) {
        const size_t k = (size_t)l * (l + 1) / 2 + m;
        const double2 A = cp[(size_t)l * d2];
        const double2 Bm = (m > 0) ? cm[(size_t)l * d2] : make_double2(0.0, 0.0);
        const double pn = plx[k * ld + rw.x];
        ap.x = fma(A.x, pn, ap.x);
        ap.y = fma(A.y, pn, ap.y);
        am.x = fma(Bm.x, pn, am.x);
        am.y = fma(Bm.y, pn, am.y);
        if (south) {
            const double ps = plx[k * ld + rw.y];
            ap.z = fma(A.x, ps, ap.z);
            ap.w = fma(A.y, ps, ap.w);
            am.z = fma(Bm.x, ps, am.z);
            am.w = fma(Bm.y, ps, am.w);
        }
    }
    fb_p[(size_t)m * P + s] = ap;
    fb_m[(size_t)m * P + s] = am;
}

// SHExpandGLQC with the table (src/SHExpandGLQC.f95 plx branch and the
// epilogue): as k_plx_anal_leg for the complex DFTs of the rows, gb_p /
// gb_m holding orders +m and -m (the latter times (-1)^m) of the slot's
// north and south rows apart; complex coefficients cilm[2][d1][d2].
__global__ __launch_bounds__(kWave) void k_plx_anal_leg_c(const double4 *__restrict__ gb_p,
                                                          const double4 *__restrict__ gb_m, int P, int nlat, int Lc,
                                                          const double *__restrict__ plx, int ld, int norm,
                                                          int nblk, double2 *__restrict__ cilm, int d1, int d2)
{
    const int m = blockIdx.x / nblk, b = blockIdx.x - m * nblk;
    const int l = m + b * kWave + (int)threadIdx.x;
    if (l > Lc) return;
    const double *__restrict__ row = plx + ((size_t)l * (l + 1) / 2 + m) * ld;
    const double4 *__restrict__ gp = gb_p + (size_t)m * P;
    const double4 *__restrict__ gm = gb_m + (size_t)m * P;
    double2 cp = make_double2(0.0, 0.0), cm = cp;
    for (int r = 0; r < nlat; ++r) {
        const bool north = r < P;
        const int s = north ? r : nlat - 1 - r;
        const double p = row[r];
        const double4 fp = gp[s];
        cp.x = fma(p, north ? fp.x : fp.z, cp.x);
        cp.y = fma(p, north ? fp.y : fp.w, cp.y);
        if (m > 0) {
            const double4 fm = gm[s];
            cm.x = fma(p, north ? fm.x : fm.z, cm.x);
            cm.y = fma(p, north ? fm.y : fm.w, cm.y);
        }
    }
    cilm[(size_t)l * d2 + m] = make_double2(plx_normalise(cp.x, norm, l, m, true), plx_normalise(cp.y, norm, l, m, true));
    if (m > 0)
        cilm[((size_t)d1 + l) * d2 + m] =
            make_double2(plx_normalise(cm.x, norm, l, m, true), plx_normalise(cm.y, norm, l, m, true));
}

// A degree-0 complex field: every element of a rows x cols array (leading
// dimension ld) src[0] * v.
__global__ __launch_bounds__(kThreads) void k_fill_c(double2 *__restrict__ a, int rows, int cols, long long ld,
                                                     const double2 *__restrict__ src, double v)
{
    const double2 x = make_double2(src[0].x * v, src[0].y * v);
    const long long total = (long long)rows * cols;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(e / cols), c = (int)(e - (long long)r * cols);
        a[r * ld + c] = x;
    }
}

// ============================================================ complex kernels
//
// MakeGridDHC / SHExpandDHC / MakeGridGLQC / SHExpandGLQC
// (src/MakeGridDHC.f95:531-609, src/SHExpandDHC.f95:552-723 and the GLQ
// counterparts): coefficients cilm[2][d1][d2] complex ([0] orders +m, [1]
// orders -m, the latter entering with (-1)^m), rings complex.  The Legendre
// sums reuse the real kernels once per sign of m with (Re, Im) in place of
// (C, S); the ring FFTs are full complex DFTs, one ring at a time.

// cpack[tri(m) + l - m] = plane[l][m] * s_lm for the tile's m <= l <= L (zero
// at m = 0 when zero_m0: the negative-order plane has no m = 0 entry).
__device__ __forceinline__ void pack_tile_c(const double2 *__restrict__ plane, int d2, int L,
                                            const double *__restrict__ s_tab, int Lt, double2 *__restrict__ cpack,
                                            int m0, int l0, bool zero_m0, CoefTile &t)
{
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int r = e / kTM, c = e - r * kTM, l = l0 + r, m = m0 + c;
        double2 v = make_double2(0.0, 0.0);
        if (l <= L && m <= l && !(zero_m0 && m == 0)) v = plane[(size_t)l * d2 + m];
        t[r][c] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int c = e / kTL, r = e - c * kTL, l = l0 + r, m = m0 + c;
        if (l <= L && m <= l) {
            const double sv = s_tab[tri_offd(m, Lt) + (l - m)];
            const double2 v = t[r][c];
            cpack[tri_offd(m, L) + (l - m)] = make_double2(v.x * sv, v.y * sv);
        }
    }
    __syncthreads();
}

// plane[l][m] (within d1 x d2) = cbuf * post for m <= l <= L, zero elsewhere
// (and at m = 0 when zero_m0).
__device__ __forceinline__ void unpack_tile_c(const double2 *__restrict__ cbuf, const double *__restrict__ post,
                                              int Lt, int L, double2 *__restrict__ plane, int d1, int d2, int m0,
                                              int l0, bool zero_m0, CoefTile &t)
{
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int c = e / kTL, r = e - c * kTL, l = l0 + r, m = m0 + c;
        double2 v = make_double2(0.0, 0.0);
        if (l <= L && m <= l && !(zero_m0 && m == 0)) {
            const double pv = post[tri_offd(m, Lt) + (l - m)];
            const double2 x = cbuf[tri_offd(m, L) + (l - m)];
            v = make_double2(x.x * pv, x.y * pv);
        }
        t[r][c] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTL * kTM; e += blockDim.x) {
        const int r = e / kTM, c = e - r * kTM, l = l0 + r, m = m0 + c;
        if (l < d1 && m < d2) plane[(size_t)l * d2 + m] = t[r][c];
    }
    __syncthreads();
}

// grid: ntm x ntl tiles; both planes
__global__ __launch_bounds__(kThreads) void k_pack_synth_c(const double2 *__restrict__ cilm, int d1, int d2,
                                                           const double *__restrict__ s_tab, int Lt,
                                                           double2 *__restrict__ cpack_p,
                                                           double2 *__restrict__ cpack_m, int Lc, int ntm)
{
    __shared__ CoefTile t;
    const int tl = blockIdx.x / ntm, tm = blockIdx.x - tl * ntm;
    const int m0 = tm * kTM, l0 = tl * kTL;
    if (l0 + kTL - 1 < m0) return;
    pack_tile_c(cilm, d2, Lc, s_tab, Lt, cpack_p, m0, l0, false, t);
    pack_tile_c(cilm + (size_t)d1 * d2, d2, Lc, s_tab, Lt, cpack_m, m0, l0, true, t);
}

__global__ __launch_bounds__(kThreads) void k_unpack_anal_c(const double2 *__restrict__ cbuf_p,
                                                            const double2 *__restrict__ cbuf_m,
                                                            const double *__restrict__ post, int Lt, int Lc,
                                                            double2 *__restrict__ cilm, int d1, int d2, int ntm)
{
    __shared__ CoefTile t;
    const int tl = blockIdx.x / ntm, tm = blockIdx.x - tl * ntm;
    unpack_tile_c(cbuf_p, post, Lt, Lc, cilm, d1, d2, tm * kTM, tl * kTL, false, t);
    unpack_tile_c(cbuf_m, post, Lt, Lc, cilm + (size_t)d1 * d2, d1, d2, tm * kTM, tl * kTL, true, t);
}

__device__ __forceinline__ double msign(int m) { return (m & 1) ? -1.0 : 1.0; }

// Complex synthesis of slot s: for the north (ring 0) and south (ring 1)
// row, X_k = F_k and X_{n-k} = (-1)^k G_k (F, G the sums of orders +k, -k,
// times the order's rescale), inverse DFT via conj(DFT(conj X)).
template <int kVar>
__device__ __forceinline__ void synth_fft_c_tile(const FftDev &F, const double4 *__restrict__ fb_p,
                                                 const double4 *__restrict__ fb_m, int P, int Lc, int s, double u,
                                                 int2 rw, int nrows, double2 *__restrict__ grid, long long ld,
                                                 int ext_col, double2 *x, double2 *chunk = nullptr,
                                                 int chunk_cap = 0)
{
    const bool chunked = chunk && F.nst >= 2 && F.m / F.radix[0] <= chunk_cap;
    const int n = F.n, m = F.m;
    const int top = (F.lc >= 0) ? m : n;
    for (int ring = 0; ring < 2; ++ring) {
        const int row = ring ? rw.y : rw.x;
        if (row < 0 || row >= nrows) continue;
        for (int q = threadIdx.x; q < m; q += blockDim.x)
            if ((q > Lc && q < top - Lc) || q >= top) x[fpad(q)] = make_double2(0.0, 0.0);
        for (int k = threadIdx.x; k <= Lc; k += blockDim.x) {
            const double r = ring_rescale(u, k);
            const double4 fp = fb_p[(size_t)k * P + s];
            const double2 a = ring ? make_double2(fp.z * r, fp.w * r) : make_double2(fp.x * r, fp.y * r);
            x[fpad(k)] = fft_prep_in(F, k, conjd(a));
            if (k > 0) {
                const double4 fm = fb_m[(size_t)k * P + s];
                const double g = r * msign(k);
                const double2 b = ring ? make_double2(fm.z * g, fm.w * g) : make_double2(fm.x * g, fm.y * g);
                x[fpad(fft_pos(F, n - k))] = fft_prep_in(F, n - k, conjd(b));
            }
        }
        __syncthreads();
        if constexpr (kVar == kVarSplit) {
            // split plan: the inverse stages once per half of the columns
            TurnStash keep;
            const int h = n / 2;
            fft_dif_front(x, F);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                if (half == 0)
                    fft_run_turn_split<kSynth1>(x, F, F.bhat, keep);
                else
                    fft_run_turn_split<kSynth2>(x, F, F.bhat3, keep);
                __syncthreads();
                fft_dit_back(x, F);
                for (int p = threadIdx.x; p < h; p += blockDim.x) {
                    const int j = half * h + p;
                    const double2 v = conjd(fft_get(x, F, j, p));
                    grid[row * ld + j] = v;
                    if (ext_col && j == 0) grid[row * ld + n] = v;
                }
                __syncthreads();
            }
            continue;
        }
        if (chunked)
            fft_forward_chunked(x, F, F.bhat, chunk, chunk_cap);
        else
            fft_forward<kVar == kVarPrime, kVar == kVarPow2>(x, F, F.bhat);
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const double2 v = conjd(fft_get(x, F, j, j));
            grid[row * ld + j] = v;
            if (ext_col && j == 0) grid[row * ld + n] = v;
        }
        __syncthreads();
    }
}

// Complex analysis of slot s: DFT of each ring, weighted, orders +k and -k
// (the latter with (-1)^k), times the order's rescale; the north pass
// stores (f, f), the south pass turns them into (f_N + f_S, f_N - f_S).
template <int kVar>
__device__ __forceinline__ void anal_fft_c_tile(const FftDev &F, const double2 *__restrict__ grid, long long ld,
                                                int2 rw, int nrows, double2 w, double u, int P, int Lc, int s,
                                                double4 *__restrict__ gb_p, double4 *__restrict__ gb_m, double2 *x,
                                                double2 *chunk = nullptr, int chunk_cap = 0)
{
    const bool chunked = chunk && F.nst >= 2 && F.m / F.radix[0] <= chunk_cap;
    const int n = F.n, m = F.m;
    for (int ring = 0; ring < 2; ++ring) {
        const int row = ring ? rw.y : rw.x;
        if (row < 0 || row >= nrows) continue;
        const double wr = ring ? w.y : w.x;
        if constexpr (kVar == kVarSplit) {
            // split plan: each half of the columns filled and transformed in
            // turn, the products with their spectra summed in the turn
            TurnStash keep;
            const int h = n / 2;
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                for (int q = threadIdx.x; q < m; q += blockDim.x) {
                    const int col = half * h + q;
                    x[fpad(q)] = (q < h) ? fft_prep_in(F, col, grid[row * ld + col]) : make_double2(0.0, 0.0);
                }
                __syncthreads();
                fft_dif_front(x, F);
                if (half == 0)
                    fft_run_turn_split<kAnal1>(x, F, F.bhat2, keep);
                else
                    fft_run_turn_split<kAnal2>(x, F, F.bhat4, keep);
                __syncthreads();
            }
            fft_dit_back(x, F);
        } else {
            for (int q = threadIdx.x; q < m; q += blockDim.x)
                x[fpad(q)] = (q < n) ? fft_prep_in(F, q, grid[row * ld + q]) : make_double2(0.0, 0.0);
            __syncthreads();
            if (chunked)
                fft_forward_chunked(x, F, F.bhat2, chunk, chunk_cap);
            else
                fft_forward<kVar == kVarPrime, kVar == kVarPow2>(x, F, F.bhat2);
        }
        for (int k = threadIdx.x; k <= Lc; k += blockDim.x) {
            const double r = ring_rescale(u, k) * wr;
            const double2 zp = fft_get(x, F, k, k);
            const double2 fp = make_double2(zp.x * r, zp.y * r);
            double2 fm = make_double2(0.0, 0.0);
            if (k > 0) {
                const double2 zm = fft_get(x, F, n - k, fft_pos(F, n - k));
                const double g = r * msign(k);
                fm = make_double2(zm.x * g, zm.y * g);
            }
            double4 &op = gb_p[(size_t)k * P + s];
            double4 &om = gb_m[(size_t)k * P + s];
            if (ring == 0) {
                op = make_double4(fp.x, fp.y, fp.x, fp.y);
                om = make_double4(fm.x, fm.y, fm.x, fm.y);
            } else if (u < 0.0) {
                // unscaled sums: the two rings kept apart
                op.z = fp.x;
                op.w = fp.y;
                om.z = fm.x;
                om.w = fm.y;
            } else {
                const double4 a = op, b = om;
                op = make_double4(a.x + fp.x, a.y + fp.y, a.x - fp.x, a.y - fp.y);
                om = make_double4(b.x + fm.x, b.y + fm.y, b.x - fm.x, b.y - fm.y);
            }
        }
        __syncthreads();
    }
}

// One workgroup per slot (LDS) or per slice of a global workspace (WS).
template <bool WS, int kVar>
__global__ __launch_bounds__(fft_max_threads(kVar)) void k_synth_fft_c(FftDev F, const double4 *__restrict__ fb_p,
                                                          const double4 *__restrict__ fb_m, int P, int Lc,
                                                          const double *__restrict__ ringu,
                                                          const int2 *__restrict__ rows, int nrows,
                                                          double2 *__restrict__ grid, long long ld, int ext_col,
                                                          double2 *ws, long long ws_stride, int chunk_cap)
{
    extern __shared__ double2 lds[];
    double2 *x;
    if constexpr (WS) x = ws + blockIdx.x * ws_stride;
    else x = lds;
    for (int s = blockIdx.x; s < P; s += gridDim.x)
        synth_fft_c_tile<kVar>(F, fb_p, fb_m, P, Lc, s, ringu ? ringu[s] : -1.0, rows[s], nrows, grid, ld, ext_col, x,
                               WS ? lds : nullptr, chunk_cap);
}

template <bool WS, int kVar>
__global__ __launch_bounds__(fft_max_threads(kVar)) void k_anal_fft_c(FftDev F, const double2 *__restrict__ grid,
                                                         long long ld, const double *__restrict__ ringu,
                                                         const int2 *__restrict__ rows, int nrows,
                                                         const double2 *__restrict__ wts, int P, int Lc,
                                                         double4 *__restrict__ gb_p, double4 *__restrict__ gb_m,
                                                         double2 *ws, long long ws_stride, int chunk_cap)
{
    extern __shared__ double2 lds[];
    double2 *x;
    if constexpr (WS) x = ws + blockIdx.x * ws_stride;
    else x = lds;
    for (int s = blockIdx.x; s < P; s += gridDim.x)
        anal_fft_c_tile<kVar>(F, grid, ld, rows[s], nrows, wts[s], ringu ? ringu[s] : -1.0, P, Lc, s, gb_p, gb_m, x,
                              WS ? lds : nullptr, chunk_cap);
}

// ================================================================== gradient
//
// MakeGradientDH (src/MakeGradientDH.f95): with dP_lm/dz = (a_lm P_l-1,m -
// l z P_lm) / u^2, a_lm = sqrt((2l+1)(l^2-m^2)/(2l-1)), the theta component
// (1/r) dF/dtheta = -(sin theta / (r u^2)) (G_A - z G_B) and the phi component
// (1/(r sin theta)) dF/dphi = G_Phi / (r sin theta), where G_X is the 4pi
// synthesis of A_lm = a_l+1,m c_l+1,m, B_lm = l c_lm and
// Phi_lm = (m S_lm, -m C_lm).

// out: [3][2][Lc+1][Lc+1] (A, B, Phi)
__global__ __launch_bounds__(kThreads) void k_gradient_coefs(const double *__restrict__ cilm, int d1, int d2, int Lc,
                                                             double *__restrict__ out)
{
    const int L1 = Lc + 1;
    const size_t plane = (size_t)L1 * L1;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L1 * L1; e += gridDim.x * blockDim.x) {
        const int l = e / L1, m = e - l * L1;
        double c = 0.0, sv = 0.0, cn = 0.0, sn = 0.0;
        if (m <= l) {
            c = cilm[(size_t)l * d2 + m];
            sv = (m == 0) ? 0.0 : cilm[((size_t)d1 + l) * d2 + m];
            if (l + 1 <= Lc) {
                cn = cilm[(size_t)(l + 1) * d2 + m];
                sn = (m == 0) ? 0.0 : cilm[((size_t)d1 + l + 1) * d2 + m];
            }
        }
        const int lp = l + 1;
        const double a = (m <= l && lp <= Lc)
                             ? sqrt((double)(2 * lp + 1)) * sqrt((double)(lp - m)) * sqrt((double)(lp + m)) /
                                   sqrt((double)(2 * lp - 1))
                             : 0.0;
        out[0 * plane + e] = a * cn;
        out[1 * plane + e] = a * sn;
        out[2 * plane + e] = l * c;
        out[3 * plane + e] = l * sv;
        out[4 * plane + e] = m * sv;
        out[5 * plane + e] = -m * c;
    }
}

// a[e] *= b[e] for e < n (the pointwise product of two grids).
__global__ __launch_bounds__(kThreads) void k_mul_inplace(double *__restrict__ a, const double *__restrict__ b,
                                                          long long n)
{
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x)
        a[e] *= b[e];
}

// Per output row of a gravity grid on an ellipsoid: z of the row, sin of the
// pair's northern colatitude, u^2, rho = r0/r and r of the row, the
// centrifugal terms added to rad, theta and pot, the normal gravity taken
// from the magnitude, and whether the row is a pole (theta and phi 0).
struct GravRow {
    double z, st, u2, rho, rex, crad, cth, cpot, ng, pole;
};

// In place: pot holds S(c), rad S(l c), theta S(A), phi S(Phi), the radially
// continued syntheses of the potential coefficients (A_lm = a_l+1,m
// c_l+1,m, Phi_lm = (m S_lm, -m C_lm)).  Then rad = -GM/r^2 (S(l c) + S(c)),
// pot = GM/r S(c), theta = -sin t GM/r^2 (rho S(A) - z S(l c)) / u^2,
// phi = GM/r^2 S(Phi) / sin t, plus the row's centrifugal terms, and
// total = |(rad, theta, phi)| - normal gravity (src/MakeGravGridDH.f95:
// 838-1052 term by term).
__global__ __launch_bounds__(kThreads) void k_grav_combine(double *__restrict__ rad, double *__restrict__ theta,
                                                           double *__restrict__ phi, double *__restrict__ total,
                                                           double *__restrict__ pot,
                                                           const GravRow *__restrict__ rowt, int nrows, int ncols,
                                                           long long ld, double gm)
{
    const long long count = (long long)nrows * ncols;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e / ncols), k = (int)(e - (long long)i * ncols);
        const long long o = (long long)i * ld + k;
        const GravRow R = rowt[i];
        const double gu = pot[o], gb = rad[o], ga = theta[o], gp = phi[o];
        const double r2 = R.rex * R.rex;
        const double vr = -(gb + gu) * gm / r2 + R.crad;
        const double vu = gu * gm / R.rex + R.cpot;
        double vt = 0.0, vp = 0.0;
        if (R.pole == 0.0) {
            vt = -R.st * ((R.rho * ga - R.z * gb) / R.u2) * gm / r2;
            vp = gp * (gm / r2) / R.st;
        }
        vt += R.cth;
        rad[o] = vr;
        theta[o] = vt;
        phi[o] = vp;
        pot[o] = vu;
        total[o] = sqrt(vr * vr + vp * vp + vt * vt) - R.ng;
    }
}

// theta = -(sin t / (r u^2)) (theta - z gb) and phi = phi / (r sin t) row by
// row (t = pi i / n, z = cos t, u^2 = (1 - z)(1 + z)); the pole rows are zero.
__global__ __launch_bounds__(kThreads) void k_gradient_combine(double *__restrict__ theta, double *__restrict__ phi,
                                                               const double *__restrict__ gb, int nrows, int ncols,
                                                               long long ld, int n, double radius)
{
    const double pi = 3.14159265358979323846;
    const long long total = (long long)nrows * ncols;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e / ncols), k = (int)(e - (long long)i * ncols);
        const long long o = (long long)i * ld + k;
        if (i == 0 || i == n) {
            theta[o] = 0.0;
            phi[o] = 0.0;
            continue;
        }
        double t, z, st;
        if (2 * i == n) {
            t = pi / 2.0;
            z = 0.0;
        } else {
            const int in = (2 * i < n) ? i : n - i;  // the northern ring of the pair
            t = pi * (double)in / (double)n;
            z = cos(t);
            if (in != i) z = -z;
        }
        st = sin(t);
        const double u2 = (1.0 - z) * (1.0 + z);
        theta[o] = -st * ((theta[o] - z * gb[(long long)i * ld + k]) / u2) / radius;
        phi[o] = phi[o] / radius / st;
    }
}

// ============================================================= batched kernels

struct BTables {
    const double *alpha, *s, *post, *seed;  // s: pack scale, post: unpack scale
    int Lt;
    const double *z, *u;
    const double2 *w;     // analysis ring weights per slot
    const double2 *roww;  // synthesis row weights per slot (nullptr: 1)
    int s0;               // 1: analysis sums start at each field's s0; 0: at slot 0 (adjoint of synthesis)
    const int2 *rows;
    const FftDev *plans;
    const int *kmax;      // per slot (indexed like z): orders with non-negligible functions
    const int *smin;      // per field (BField::sm) and order: first such slot
};

// items: (field, first order, first degree) of a coefficient tile
__global__ __launch_bounds__(kThreads) void kb_pack(const BField *__restrict__ fields, const int4 *__restrict__ items,
                                                    BTables T, const double *const *__restrict__ cilm,
                                                    double2 *__restrict__ cpack)
{
    __shared__ CoefTile t;
    const int4 it = items[blockIdx.x];
    const BField f = fields[it.x];
    pack_tile(cilm[it.x], f.L + 1, f.L + 1, f.L, T.s, T.Lt, cpack + f.coff, it.y, it.z, t);
}

// items: (field, m, first slot, ring pairs per lane)
__global__ __launch_bounds__(kWave) void kb_synth_leg(const BField *__restrict__ fields,
                                                      const int4 *__restrict__ items, BTables T,
                                                      const double2 *__restrict__ cpack, double4 *__restrict__ fbuf)
{
    const int4 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const int m = it.y;
    synth_leg_dispatch<kSynthKMax>(it.w, T.alpha + tri_offd(m, T.Lt), T.seed[m], T.z + f.geo,
                                   cpack + f.coff + tri_offd(m, f.L), f.L - m, f.P, it.z,
                                   fbuf + f.foff + (size_t)m * f.P);
}

// ---------------------------------------------- same-degree buckets on MFMA
//
// Fields of one degree and sampling (a bucket) share their rings and their
// Legendre functions, so for an order m the synthesis sums of all of them
// are one product
//     F[s][col] = sum_j Q[s][j] W[j][col],  Q[s][j] = q_{m+j}(z_s),
//     W[j][col] = packed coefficient (C s or S s) of field col / 2,
// for the ring-pair slots s, split by the parity of j into the even and odd
// sums whose sum and difference are the north and south rows.  One wave
// takes 16 slots and every column of the bucket, 4 degrees of one parity at
// a time on v_mfma_f64_16x16x4f64 (operand layouts: A[i][k] in lane
// i + 16k, B[k][j] in lane j + 16k, D[i][j] in register i / 4 of lane
// j + 16 (i % 4)).  Lane i computes the recursion of slot s0 + i % 16 shifted
// by 2 (i / 16) degrees, so that at block t it holds q_{8t + 2g} and
// q_{8t + 2g + 1} (g = i / 16): the A operands of the even and odd products
// of that block.  The recursion is thus computed 4 times per slot (2 FP64
// VALU operations per degree and lane), and the products run on the matrix
// pipe: 2 x ceil(2B / 16) MFMAs per block of 8 degrees and 16 slots.
//
// The same sums as the VALU tile (synth_leg_tile), and bitwise the same
// results: the recursion runs the same operations per slot, and the FP64
// MFMA accumulates its 4 products per output as a chain of fused
// multiply-adds in k order, which is the order of the VALU tile's even (odd)
// degree sums.
typedef double v4d __attribute__((ext_vector_type(4)));

// items: (bucket, m, first slot, 0); bf: the bucket's nbf fields.
//
// Every lane runs the recursion of its slot over the same degrees (the
// alpha row is wave-uniform), 8 per block, and lane group g takes q_{8t+2g}
// and q_{8t+2g+1} of the block as its A operands.  The coefficient operands
// of the next block are loaded while the current block's products run.
// (Two 16-slot tiles per wave, sharing each coefficient operand, measured
// slower: more registers, fewer waves.)
template <int NT>
__device__ __forceinline__ void synth_leg_mfma_tile(const BField *__restrict__ fields, const int *__restrict__ bf,
                                                    int nbf, const BField &f0, int m, int s0, const BTables &T,
                                                    const double2 *__restrict__ cpack, double4 *__restrict__ fbuf)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int r = lane & 15, g = lane >> 4;
    const bool g1 = (g & 1) != 0, g2 = (g & 2) != 0;
    const int jmax = f0.L - m;
    const double *__restrict__ al = T.alpha + tri_offd(m, T.Lt);
    const int s = s0 + r;
    const double z = (s < f0.P) ? T.z[f0.geo + s] : 0.0;
    // this lane's B-operand column in each column tile: a field's packed C or
    // S row of order m (nullptr past the bucket's fields)
    const double *wp[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) {
        const int col = 16 * c + r, k = col >> 1;
        wp[c] = (k < nbf) ? reinterpret_cast<const double *>(cpack + fields[bf[k]].coff + tri_offd(m, f0.L)) + (col & 1)
                          : nullptr;
    }
    auto load_w = [&](int t, double (&we)[NT], double (&wo)[NT]) {
        const int je = 8 * t + 2 * g;
#pragma unroll
        for (int c = 0; c < NT; ++c) {
            we[c] = (wp[c] && je <= jmax) ? wp[c][2 * je] : 0.0;
            wo[c] = (wp[c] && je + 1 <= jmax) ? wp[c][2 * je + 2] : 0.0;
        }
    };
    double q1 = 0.0, q2 = -T.seed[m];  // the first step yields the seed
    v4d accE[NT], accO[NT];
#pragma unroll
    for (int c = 0; c < NT; ++c) {
        accE[c] = v4d{0.0, 0.0, 0.0, 0.0};
        accO[c] = v4d{0.0, 0.0, 0.0, 0.0};
    }
    const int nblk = (jmax + 8) / 8, nfull = (jmax + 1) / 8;
    double we[NT], wo[NT];
    load_w(0, we, wo);
    for (int t = 0; t < nblk; ++t) {
        const int j0 = 8 * t;
        double q[8];
        if (t < nfull) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double qa = k == 0 ? q1 : q[k - 1], qb = k == 0 ? q2 : (k == 1 ? q1 : q[k - 2]);
                q[k] = fma(al[j0 + k] * z, qa, -qb);
            }
        } else {
            // the last, partial block: degrees past jmax continue bounded
            // (alpha 0) and meet zero coefficients
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const double qa = k == 0 ? q1 : q[k - 1], qb = k == 0 ? q2 : (k == 1 ? q1 : q[k - 2]);
                q[k] = fma((j0 + k <= jmax ? al[j0 + k] : 0.0) * z, qa, -qb);
            }
        }
        q1 = q[7];
        q2 = q[6];
        const double qe = g2 ? (g1 ? q[6] : q[4]) : (g1 ? q[2] : q[0]);
        const double qo = g2 ? (g1 ? q[7] : q[5]) : (g1 ? q[3] : q[1]);
        double ne[NT], no[NT];
        if (t + 1 < nblk) load_w(t + 1, ne, no);
#pragma unroll
        for (int c = 0; c < NT; ++c) {
            accE[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(qe, we[c], accE[c], 0, 0, 0);
            accO[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(qo, wo[c], accO[c], 0, 0, 0);
        }
        if (t + 1 < nblk) {
#pragma unroll
            for (int c = 0; c < NT; ++c) {
                we[c] = ne[c];
                wo[c] = no[c];
            }
        }
    }
    // D[i][j]: slot s0 + 4 * reg + g, column 16 c + r
#pragma unroll
    for (int c = 0; c < NT; ++c) {
        const int col = 16 * c + r, k = col >> 1;
        if (k >= nbf) continue;
        const BField &f = fields[bf[k]];
        double *base = reinterpret_cast<double *>(fbuf + f.foff + (size_t)m * f.P) + (col & 1);
#pragma unroll
        for (int reg = 0; reg < 4; ++reg) {
            const int ss = s0 + 4 * reg + g;
            if (ss < f.P) {
                const double e = accE[c][reg], o = accO[c][reg];
                base[4 * (size_t)ss] = e + o;
                base[4 * (size_t)ss + 2] = e - o;
            }
        }
    }
}

// one wave per item (bucket, m, first slot, -); bfields: per bucket
// kBucketMaxCols / 2 field indices, nbf: fields per bucket
__global__ __launch_bounds__(kWave) void kb_synth_leg_mfma(const BField *__restrict__ fields,
                                                           const int4 *__restrict__ items,
                                                           const int *__restrict__ bfields,
                                                           const int *__restrict__ nbf, BTables T,
                                                           const double2 *__restrict__ cpack,
                                                           double4 *__restrict__ fbuf)
{
    const int4 it = items[blockIdx.x];
    const int *bf = bfields + it.x * (kBucketMaxCols / 2);
    const int n = nbf[it.x];
    const BField f0 = fields[bf[0]];
    if (n <= 8)
        synth_leg_mfma_tile<1>(fields, bf, n, f0, it.y, it.z, T, cpack, fbuf);
    else if (n <= 16)
        synth_leg_mfma_tile<2>(fields, bf, n, f0, it.y, it.z, T, cpack, fbuf);
    else if (n <= 24)
        synth_leg_mfma_tile<3>(fields, bf, n, f0, it.y, it.z, T, cpack, fbuf);
    else
        synth_leg_mfma_tile<4>(fields, bf, n, f0, it.y, it.z, T, cpack, fbuf);
}

// Ring FFT items (field, first slot): each workgroup transforms nslots
// consecutive ring pairs in turn.  Workgroups are mapped so that groups of 4
// consecutive items run on the same XCD (shared L2 lines of fbuf/gbuf).

__device__ __forceinline__ int xcd_item(int b, int n)
{
    const int x = b & 7, t = b >> 3;
    const int it = (t >> 2) * 32 + x * 4 + (t & 3);
    return it < n ? it : -1;
}

template <int kVar>
__global__ __launch_bounds__(fft_max_threads(kVar)) void kb_synth_fft(const BField *__restrict__ fields,
                                                         const int2 *__restrict__ items, int nitems, int nslots,
                                                         BTables T, const double4 *__restrict__ fbuf,
                                                         double *const *grid)
{
    extern __shared__ double2 lds[];
    const int id = xcd_item(blockIdx.x, nitems);
    if (id < 0) return;
    const int2 it = items[id];
    const BField f = fields[it.x];
    const FftDev &F = T.plans[f.fft];
    const double4 *fb = fbuf + f.foff;
    const int send = min(it.y + nslots, f.P);
    // the next slot's ring sums are loaded while the current slot is in its
    // FFT stages (not for split plans, whose turns keep values in registers)
    SynthRing in;
    if constexpr (kVar != kVarSplit) synth_fft_load(fb, f.P, slot_orders(T.kmax, f.geo + it.y, f.L), it.y, in);
    for (int s = it.y; s < send; ++s) {
        if constexpr (kVar == kVarSplit) synth_fft_load(fb, f.P, slot_orders(T.kmax, f.geo + s, f.L), s, in);
        synth_fft_fill(F, in, fb, f.P, slot_orders(T.kmax, f.geo + s, f.L), s, T.u[f.geo + s], lds);
        __syncthreads();
        if constexpr (kVar != kVarSplit)
            if (s + 1 < send) synth_fft_load(fb, f.P, slot_orders(T.kmax, f.geo + s + 1, f.L), s + 1, in);
        synth_fft_run<kVar>(F, T.rows[f.geo + s], f.nlat, grid[it.x], f.nlong, 0,
                      T.roww ? T.roww[f.geo + s] : make_double2(1.0, 1.0), lds);
        __syncthreads();
    }
}

template <int kVar>
__global__ __launch_bounds__(fft_max_threads(kVar)) void kb_anal_fft(const BField *__restrict__ fields,
                                                        const int2 *__restrict__ items, int nitems, int nslots,
                                                        BTables T, const double *const *grid,
                                                        double4 *__restrict__ gbuf)
{
    extern __shared__ double2 lds[];
    const int id = xcd_item(blockIdx.x, nitems);
    if (id < 0) return;
    const int2 it = items[id];
    const BField f = fields[it.x];
    const FftDev &F = T.plans[f.fft];
    const double *g = grid[it.x];
    const int send = min(it.y + nslots, f.P);
    if constexpr (kVar == kVarSplit) {
        for (int s = it.y; s < send; ++s) {
            anal_fft_tile_split(F, g, f.nlong, T.rows[f.geo + s], f.nlat, T.w[f.geo + s], T.u[f.geo + s], f.P,
                                slot_orders(T.kmax, f.geo + s, f.L), s, gbuf + f.foff, 0, lds);
            __syncthreads();
        }
        return;
    }
    // the next slot's rows are loaded while the current slot is in its FFT
    // stages
    AnalRing in;
    anal_fft_load(F, g, f.nlong, T.rows[f.geo + it.y], f.nlat, 0, in);
    for (int s = it.y; s < send; ++s) {
        const int2 rw = T.rows[f.geo + s];
        anal_fft_fill(F, in, g, f.nlong, rw, f.nlat, 0, lds);
        __syncthreads();
        if (s + 1 < send) anal_fft_load(F, g, f.nlong, T.rows[f.geo + s + 1], f.nlat, 0, in);
        anal_fft_run<kVar>(F, rw, f.nlat, T.w[f.geo + s], T.u[f.geo + s], f.P, slot_orders(T.kmax, f.geo + s, f.L), s,
                     gbuf + f.foff, lds);
        __syncthreads();
    }
}

// Batched ring FFTs on a global workspace (fields whose work array exceeds
// the LDS): workgroup b takes items b, b + gridDim.x, ...
__global__ __launch_bounds__(kFftMaxThreads) void kb_synth_fft_ws(const BField *__restrict__ fields,
                                                            const int2 *__restrict__ items, int nitems, int nslots,
                                                            BTables T, const double4 *__restrict__ fbuf,
                                                            double *const *grid, double2 *ws, long long ws_stride,
                                                            int chunk_cap)
{
    extern __shared__ double2 chunk[];
    double2 *x = ws + blockIdx.x * ws_stride;
    for (int id = blockIdx.x; id < nitems; id += gridDim.x) {
        const int2 it = items[id];
        const BField f = fields[it.x];
        const FftDev &F = T.plans[f.fft];
        for (int s = it.y; s < it.y + nslots && s < f.P; ++s) {
            synth_fft_tile(F, fbuf + f.foff, f.P, slot_orders(T.kmax, f.geo + s, f.L), s, T.u[f.geo + s],
                           T.rows[f.geo + s], f.nlat, grid[it.x], f.nlong, 0,
                           T.roww ? T.roww[f.geo + s] : make_double2(1.0, 1.0), x, chunk, chunk_cap);
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kFftMaxThreads) void kb_anal_fft_ws(const BField *__restrict__ fields,
                                                           const int2 *__restrict__ items, int nitems, int nslots,
                                                           BTables T, const double *const *grid,
                                                           double4 *__restrict__ gbuf, double2 *ws,
                                                           long long ws_stride, int chunk_cap)
{
    extern __shared__ double2 chunk[];
    double2 *x = ws + blockIdx.x * ws_stride;
    for (int id = blockIdx.x; id < nitems; id += gridDim.x) {
        const int2 it = items[id];
        const BField f = fields[it.x];
        const FftDev &F = T.plans[f.fft];
        for (int s = it.y; s < it.y + nslots && s < f.P; ++s) {
            anal_fft_tile(F, grid[it.x], f.nlong, T.rows[f.geo + s], f.nlat, T.w[f.geo + s], T.u[f.geo + s], f.P,
                          slot_orders(T.kmax, f.geo + s, f.L), s, gbuf + f.foff, 0, x, chunk, chunk_cap);
            __syncthreads();
        }
    }
}

// items: (field, m); single-wave workgroups, one launch per ring pairs per
// lane K.  Slots below the field's s0 are skipped (the DH pole slot carries
// no analysis weight) unless T.s0 is 0 (the adjoint of synthesis reads every
// slot), and so are the slots below the order's first slot with
// non-negligible functions.
template <int K>
__global__ __launch_bounds__(kWave) void kb_anal_leg(const BField *__restrict__ fields,
                                                     const int2 *__restrict__ items, BTables T,
                                                     const double4 *__restrict__ gbuf, double2 *__restrict__ cbuf)
{
    const int2 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const int m = it.y;
    const int s0 = min(T.s0, f.s0);
    anal_leg_tile<K, false>(T.alpha + tri_offd(m, T.Lt), T.seed[m], T.z + f.geo, gbuf + f.foff + (size_t)m * f.P,
                     f.L - m, T.smin ? max(s0, T.smin[f.sm + m]) : s0, f.P, cbuf + f.coff + tri_offd(m, f.L));
}

// items: (field, first order, first degree) of a tile of the field's
// [2][L+1][L+1] output
__global__ __launch_bounds__(kThreads) void kb_unpack(const BField *__restrict__ fields,
                                                      const int4 *__restrict__ items, BTables T,
                                                      const double2 *__restrict__ cbuf, double *const *cilm)
{
    __shared__ CoefTile t;
    const int4 it = items[blockIdx.x];
    const BField f = fields[it.x];
    unpack_tile(cbuf + f.coff, T.post, T.Lt, f.L, cilm[it.x], f.L + 1, f.L + 1, it.y, it.z, t);
}

}  // namespace shtg
