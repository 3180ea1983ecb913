// MI355X kernels of the plain (one field per call) transform path.
//
// Synthesis (MakeGridDH / MakeGridGLQ):
//   k_pack_synth   cilm[2][d1][d2] -> cpack[m][l-m] = (C s_lm, S s_lm)
//   k_synth_leg    one workgroup per (order m, block of ring pairs); lanes
//                  are ring pairs, the recursion over l runs in registers with
//                  wave-uniform (scalar) table and coefficient loads; north
//                  and south sums come from even/odd degree parity
//                  -> fbuf[m][slot] = (C_N, S_N, C_S, S_S) * rescale
//   k_synth_fft    one workgroup per ring pair: Hermitian fill of both rings
//                  into one complex spectrum, inverse DFT in LDS, rows out
// Analysis (SHExpandDH / SHExpandGLQ):
//   k_anal_fft     one workgroup per ring pair: both rings as one complex
//                  signal, forward DFT in LDS, split, quadrature weights
//                  -> gbuf[m][slot] = (E_C, E_S, O_C, O_S)
//   k_anal_leg     one workgroup per order m over all ring pairs; lanes hold
//                  K ring pairs each, the recursion runs in registers and the
//                  per-degree sums over rings are reduced through LDS in
//                  chunks of 8 degrees -> cbuf[m][l-m]
//   k_unpack_anal  cbuf * (s_lm * normalisation) -> cilm[2][d1][d2]
//
// Geometry of a ring-pair slot s: z = cos(theta) of the northern ring, the
// southern ring is its mirror (z -> -z).  Reference loops:
// src/MakeGridDH.f95:504-639, src/SHExpandDH.f95:463-664,
// src/MakeGridGLQ.f95:467-669, src/SHExpandGLQ.f95:445-649.
#pragma once

#include <hip/hip_runtime.h>

#include "sht_device.hpp"

namespace shtg {

constexpr int kThreads = 256;

// ------------------------------------------------------------ packing

// cpack[tri(m) + j] = (C_{m+j,m} s, S_{m+j,m} s); S_l0 ignored as in
// src/MakeGridDH.f95:523-538 (only cilm(1,l,1) enters the m = 0 sum).
__global__ __launch_bounds__(kThreads) void k_pack_synth(const double *__restrict__ cilm, int d1, int d2,
                                                         const double *__restrict__ s_tab, int Lt,
                                                         double2 *__restrict__ cpack, int Lc)
{
    const long long total = tri_offd(Lc + 1, Lc);
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        // invert tri_off: find m with tri(m) <= idx < tri(m+1)
        double A = 2.0 * Lc + 3.0;
        int m = (int)floor((A - sqrt(A * A - 8.0 * (double)idx)) * 0.5);
        if (m < 0) m = 0;
        while (m > 0 && tri_offd(m, Lc) > idx) --m;
        while (m < Lc && tri_offd(m + 1, Lc) <= idx) ++m;
        const int j = (int)(idx - tri_offd(m, Lc));
        const int l = m + j;
        const double s = s_tab[tri_offd(m, Lt) + j];
        const double c = cilm[(size_t)l * d2 + m];
        const double sn = (m == 0) ? 0.0 : cilm[((size_t)d1 + l) * d2 + m];
        cpack[idx] = make_double2(c * s, sn * s);
    }
}

// ------------------------------------------------------------ synthesis

template <int K>
__global__ __launch_bounds__(kThreads) void k_synth_leg(const double *__restrict__ alpha, int Lt,
                                                        const double *__restrict__ seed,
                                                        const double *__restrict__ ringz,
                                                        const double *__restrict__ ringu,
                                                        const double2 *__restrict__ cpack, int Lc, int P,
                                                        int nrb, double4 *__restrict__ fbuf)
{
    const int m = blockIdx.x / nrb;
    const int rb = blockIdx.x - m * nrb;
    const double *__restrict__ al = alpha + tri_offd(m, Lt);
    const double2 *__restrict__ cp = cpack + tri_offd(m, Lc);
    const double sd = seed[m];

    double z[K], q1[K], q2[K], r[K], eC[K], eS[K], oC[K], oS[K];
    int slot[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        slot[k] = (rb * K + k) * kThreads + threadIdx.x;
        const bool ok = slot[k] < P;
        double zz = ok ? ringz[slot[k]] : 0.0;
        double uu = ok ? ringu[slot[k]] : 1.0;
        r[k] = (m == 0) ? 1.0 : scaled_upow(uu, m);
        if (r[k] == 0.0) zz = 0.0;  // keeps the unused recursion bounded
        z[k] = zz;
        q1[k] = 0.0;
        q2[k] = -sd;  // so that the first step yields q_m = seed
        eC[k] = eS[k] = oC[k] = oS[k] = 0.0;
    }
    const int jmax = Lc - m;
    int j = 0;
    for (; j + 1 <= jmax; j += 2) {
        const double a0 = al[j], a1 = al[j + 1];
        const double2 c0 = cp[j], c1 = cp[j + 1];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double q = (a0 * z[k]) * q1[k] - q2[k];
            eC[k] += q * c0.x;
            eS[k] += q * c0.y;
            double q3 = (a1 * z[k]) * q - q1[k];
            oC[k] += q3 * c1.x;
            oS[k] += q3 * c1.y;
            q2[k] = q;
            q1[k] = q3;
        }
    }
    if (j == jmax) {
        const double a0 = al[j];
        const double2 c0 = cp[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double q = (a0 * z[k]) * q1[k] - q2[k];
            eC[k] += q * c0.x;
            eS[k] += q * c0.y;
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (slot[k] < P) {
            double4 o;
            if (r[k] == 0.0) {
                o = make_double4(0.0, 0.0, 0.0, 0.0);
            } else {
                o = make_double4((eC[k] + oC[k]) * r[k], (eS[k] + oS[k]) * r[k], (eC[k] - oC[k]) * r[k],
                                 (eS[k] - oS[k]) * r[k]);
            }
            fbuf[(size_t)m * P + slot[k]] = o;
        }
    }
}

// rows[s] = (north row, south row or -1); rows >= nrows are skipped.
__global__ __launch_bounds__(kThreads) void k_synth_fft(FftDev F, const double4 *__restrict__ fbuf, int P, int Lc,
                                                        const int2 *__restrict__ rows, int nrows,
                                                        double *__restrict__ grid, long long ld, int ext_col)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x;
    const int n = F.n;
    const int2 rw = rows[s];
    for (int k = threadIdx.x; k < F.m; k += blockDim.x) {
        double2 zc = make_double2(0.0, 0.0);
        if (k < n) {
            // Z_k = X_k + i Y_k with X (north) and Y (south) Hermitian:
            // X_0 = C_N, X_k = (C_N - i S_N)/2 and X_{n-k} = conj(X_k).
            if (k == 0) {
                double4 f = fbuf[s];
                zc = make_double2(f.x, -f.z);
            } else if (k <= Lc) {
                double4 f = fbuf[(size_t)k * P + s];
                zc = make_double2(0.5 * (f.x + f.w), -0.5 * (f.z - f.y));
            } else if (k >= n - Lc) {
                double4 f = fbuf[(size_t)(n - k) * P + s];
                zc = make_double2(0.5 * (f.x - f.w), -0.5 * (f.y + f.z));
            }
            zc = fft_prep_in(F, k, zc);  // conj(Z_k): inverse via conj(F(conj Z))
        }
        lds[k] = zc;
    }
    __syncthreads();
    fft_forward(lds, F);
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double2 v = fft_get(lds, F, j);
        if (north) grid[rw.x * ld + j] = v.x;
        if (south) grid[rw.y * ld + j] = -v.y;
        if (ext_col && j == 0) {
            if (north) grid[rw.x * ld + n] = v.x;
            if (south) grid[rw.y * ld + n] = -v.y;
        }
    }
}

// ------------------------------------------------------------ analysis

// wts[s] = (w_north, w_south) include the quadrature weight and the
// constant factors of the reference (2 pi w / nlong for GLQ,
// sqrt(2 pi) sqrt(4 pi) aj / nlong for DH).
__global__ __launch_bounds__(kThreads) void k_anal_fft(FftDev F, const double *__restrict__ grid, long long ld,
                                                       const int2 *__restrict__ rows, int nrows,
                                                       const double2 *__restrict__ wts, int P, int Lc,
                                                       double4 *__restrict__ gbuf)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x;
    const int n = F.n;
    const int2 rw = rows[s];
    const bool north = rw.x >= 0 && rw.x < nrows;
    const bool south = rw.y >= 0 && rw.y < nrows;
    for (int k = threadIdx.x; k < F.m; k += blockDim.x) {
        double2 v = make_double2(0.0, 0.0);
        if (k < n) {
            v.x = north ? grid[rw.x * ld + k] : 0.0;
            v.y = south ? grid[rw.y * ld + k] : 0.0;
            v = fft_prep_in(F, k, v);
        }
        lds[k] = v;
    }
    __syncthreads();
    fft_forward(lds, F);
    const double2 w = wts[s];
    const double wn = north ? w.x : 0.0, ws = south ? w.y : 0.0;
    for (int mm = threadIdx.x; mm <= Lc; mm += blockDim.x) {
        const double2 zm = fft_get(lds, F, mm);
        const double2 zn = fft_get(lds, F, mm == 0 ? 0 : n - mm);
        const double xr = 0.5 * (zm.x + zn.x), xi = 0.5 * (zm.y - zn.y);
        const double yr = 0.5 * (zm.y + zn.y), yi = 0.5 * (zn.x - zm.x);
        const double f1c = wn * xr, f1s = -wn * xi, f2c = ws * yr, f2s = -ws * yi;
        gbuf[(size_t)mm * P + s] = make_double4(f1c + f2c, f1s + f2s, f1c - f2c, f1s - f2s);
    }
}

template <int K>
__global__ __launch_bounds__(kThreads) void k_anal_leg(const double *__restrict__ alpha, int Lt,
                                                       const double *__restrict__ seed,
                                                       const double *__restrict__ ringz,
                                                       const double *__restrict__ ringu,
                                                       const double4 *__restrict__ gbuf, int Lc, int P,
                                                       double2 *__restrict__ cbuf)
{
    constexpr int LC = 8;  // degrees per reduction chunk
    __shared__ double red[2 * LC][kThreads];
    const int m = blockIdx.x;
    const double *__restrict__ al = alpha + tri_offd(m, Lt);
    const double sd = seed[m];
    double z[K], q1[K], q2[K], eC[K], eS[K], oC[K], oS[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int s = k * kThreads + threadIdx.x;
        const bool ok = s < P;
        double zz = ok ? ringz[s] : 0.0;
        const double uu = ok ? ringu[s] : 1.0;
        const double r = (m == 0) ? 1.0 : scaled_upow(uu, m);
        double4 g = ok ? gbuf[(size_t)m * P + s] : make_double4(0.0, 0.0, 0.0, 0.0);
        if (r == 0.0 || !ok) {
            zz = 0.0;
            g = make_double4(0.0, 0.0, 0.0, 0.0);
        }
        z[k] = zz;
        eC[k] = g.x * r;
        eS[k] = g.y * r;
        oC[k] = g.z * r;
        oS[k] = g.w * r;
        q1[k] = 0.0;
        q2[k] = -sd;
    }
    const int jmax = Lc - m;
    double2 *__restrict__ out = cbuf + tri_offd(m, Lc);
    for (int j0 = 0; j0 <= jmax; j0 += LC) {
        double acc[LC][2];
#pragma unroll
        for (int jj = 0; jj < LC; ++jj) acc[jj][0] = acc[jj][1] = 0.0;
#pragma unroll
        for (int jj = 0; jj < LC; ++jj) {
            const int j = j0 + jj;
            if (j <= jmax) {
                const double a = al[j];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const double q = (a * z[k]) * q1[k] - q2[k];
                    q2[k] = q1[k];
                    q1[k] = q;
                    if (jj % 2 == 0) {
                        acc[jj][0] += q * eC[k];
                        acc[jj][1] += q * eS[k];
                    } else {
                        acc[jj][0] += q * oC[k];
                        acc[jj][1] += q * oS[k];
                    }
                }
            }
        }
#pragma unroll
        for (int v = 0; v < 2 * LC; ++v) red[v][threadIdx.x] = acc[v >> 1][v & 1];
        __syncthreads();
        {
            const int v = threadIdx.x >> 4, part = threadIdx.x & 15;
            double sum = 0.0;
#pragma unroll
            for (int i = 0; i < 16; ++i) sum += red[v][part * 16 + i];
            sum += __shfl_xor(sum, 8, 16);
            sum += __shfl_xor(sum, 4, 16);
            sum += __shfl_xor(sum, 2, 16);
            sum += __shfl_xor(sum, 1, 16);
            const int jj = v >> 1;
            if (part == 0 && j0 + jj <= jmax) {
                double *o = reinterpret_cast<double *>(out + j0 + jj);
                o[v & 1] = sum;
            }
        }
        __syncthreads();
    }
}

// cilm[i][l][m] = cbuf[tri(m) + l - m].i * post(l, m) for m <= l <= Lc,
// zero elsewhere (including S_l0), over the whole [2][d1][d2] array.
__global__ __launch_bounds__(kThreads) void k_unpack_anal(const double2 *__restrict__ cbuf,
                                                          const double *__restrict__ post, int Lt, int Lc,
                                                          double *__restrict__ cilm, int d1, int d2)
{
    const long long total = 2LL * d1 * d2;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int m = (int)(idx % d2);
        const long long t = idx / d2;
        const int l = (int)(t % d1);
        const int i = (int)(t / d1);
        double v = 0.0;
        if (l <= Lc && m <= l && !(i == 1 && m == 0)) {
            const double2 c = cbuf[tri_offd(m, Lc) + (l - m)];
            v = (i == 0 ? c.x : c.y) * post[tri_offd(m, Lt) + (l - m)];
        }
        cilm[idx] = v;
    }
}

}  // namespace shtg
