// MI355X kernels of the transform path: per-workgroup tiles, the plain
// (one field per call) kernels and the batched (many fields per launch)
// kernels built from the same tiles.
//
// Synthesis (MakeGridDH / MakeGridGLQ):
//   pack        cilm[2][d1][d2] -> cpack[m][l-m] = (C s_lm, S s_lm)
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

#include "sht_device.hpp"

namespace shtg {

constexpr int kThreads = 256;
constexpr int kSynthK = 2;   // ring pairs per lane in synthesis
constexpr int kAnalK = 4;    // ring pairs per lane and pass in analysis
constexpr int kAnalLC = 8;   // degrees per reduction chunk in analysis

// ====================================================================== tiles

// One packed coefficient: C_lm s_lm, S_lm s_lm (S_l0 ignored as in
// src/MakeGridDH.f95:523-538: only cilm(1,l,1) enters the m = 0 sum).
__device__ __forceinline__ double2 pack_one(const double *__restrict__ cilm, int d1, int d2, int l, int m, double s)
{
    const double c = cilm[(size_t)l * d2 + m];
    const double sn = (m == 0) ? 0.0 : cilm[((size_t)d1 + l) * d2 + m];
    return make_double2(c * s, sn * s);
}

// Synthesis sums of order m for the ring pairs slot0 + k*kThreads + tid.
// al: recursion row of order m (al[j] = alpha_{m+j}); cp: packed
// coefficients of order m; frow: fbuf row of order m.
template <int K>
__device__ __forceinline__ void synth_leg_tile(const double *__restrict__ al, double sd,
                                               const double *__restrict__ ringz, const double *__restrict__ ringu,
                                               const double2 *__restrict__ cp, int jmax, int m, int P, int slot0,
                                               double4 *__restrict__ frow)
{
    double z[K], q1[K], q2[K], r[K], eC[K], eS[K], oC[K], oS[K];
    int slot[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        slot[k] = slot0 + k * kThreads + threadIdx.x;
        const bool ok = slot[k] < P;
        double zz = ok ? ringz[slot[k]] : 0.0;
        const double uu = ok ? ringu[slot[k]] : 1.0;
        r[k] = (m == 0) ? 1.0 : scaled_upow(uu, m);
        if (r[k] == 0.0) zz = 0.0;  // keeps the unused recursion bounded
        z[k] = zz;
        q1[k] = 0.0;
        q2[k] = -sd;  // the first step then yields q_m = seed
        eC[k] = eS[k] = oC[k] = oS[k] = 0.0;
    }
    int j = 0;
    for (; j + 1 <= jmax; j += 2) {
        const double a0 = al[j], a1 = al[j + 1];
        const double2 c0 = cp[j], c1 = cp[j + 1];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double q = (a0 * z[k]) * q1[k] - q2[k];
            eC[k] += q * c0.x;
            eS[k] += q * c0.y;
            const double q3 = (a1 * z[k]) * q - q1[k];
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
            const double q = (a0 * z[k]) * q1[k] - q2[k];
            eC[k] += q * c0.x;
            eS[k] += q * c0.y;
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (slot[k] < P) {
            double4 o = make_double4(0.0, 0.0, 0.0, 0.0);
            if (r[k] != 0.0)
                o = make_double4((eC[k] + oC[k]) * r[k], (eS[k] + oS[k]) * r[k], (eC[k] - oC[k]) * r[k],
                                 (eS[k] - oS[k]) * r[k]);
            frow[slot[k]] = o;
        }
    }
}

// Analysis sums of order m over all P ring pairs: out[j] = sum over slots of
// q_{m+j}(z) * (E or O by parity of j), written as (C, S) pairs.  Passes of
// kAnalK * kThreads slots; the first pass stores, later passes accumulate.
// red: kThreads * 2 * kAnalLC doubles of LDS.
__device__ void anal_leg_tile(const double *__restrict__ al, double sd, const double *__restrict__ ringz,
                              const double *__restrict__ ringu, const double4 *__restrict__ grow, int jmax, int m,
                              int P, double2 *__restrict__ out, double *red)
{
    constexpr int K = kAnalK, LC = kAnalLC;
    for (int p0 = 0; p0 < P; p0 += K * kThreads) {
        double z[K], q1[K], q2[K], eC[K], eS[K], oC[K], oS[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int s = p0 + k * kThreads + threadIdx.x;
            const bool ok = s < P;
            double zz = ok ? ringz[s] : 0.0;
            const double uu = ok ? ringu[s] : 1.0;
            const double r = (m == 0) ? 1.0 : scaled_upow(uu, m);
            double4 g = ok ? grow[s] : make_double4(0.0, 0.0, 0.0, 0.0);
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
            for (int v = 0; v < 2 * LC; ++v) red[v * kThreads + threadIdx.x] = acc[v >> 1][v & 1];
            __syncthreads();
            {
                const int v = threadIdx.x >> 4, part = threadIdx.x & 15;
                double sum = 0.0;
#pragma unroll
                for (int i = 0; i < 16; ++i) sum += red[v * kThreads + part * 16 + i];
                sum += __shfl_xor(sum, 8, 16);
                sum += __shfl_xor(sum, 4, 16);
                sum += __shfl_xor(sum, 2, 16);
                sum += __shfl_xor(sum, 1, 16);
                const int jj = v >> 1;
                if (part == 0 && j0 + jj <= jmax) {
                    double *o = reinterpret_cast<double *>(out + j0 + jj) + (v & 1);
                    *o = (p0 == 0) ? sum : *o + sum;
                }
            }
            __syncthreads();
        }
    }
}

// Ring-pair synthesis FFT for slot s: fbuf holds rows of order m at stride P.
__device__ void synth_fft_tile(const FftDev &F, const double4 *__restrict__ fbuf, int P, int Lc, int s, int2 rw,
                               int nrows, double *__restrict__ grid, long long ld, int ext_col, double2 *lds)
{
    const int n = F.n;
    for (int k = threadIdx.x; k < F.m; k += blockDim.x) {
        double2 zc = make_double2(0.0, 0.0);
        if (k < n) {
            // Z_k = X_k + i Y_k with X (north) and Y (south) Hermitian:
            // X_0 = C_N, X_k = (C_N - i S_N)/2 and X_{n-k} = conj(X_k).
            if (k == 0) {
                const double4 f = fbuf[s];
                zc = make_double2(f.x, -f.z);
            } else if (k <= Lc) {
                const double4 f = fbuf[(size_t)k * P + s];
                zc = make_double2(0.5 * (f.x + f.w), -0.5 * (f.z - f.y));
            } else if (k >= n - Lc) {
                const double4 f = fbuf[(size_t)(n - k) * P + s];
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
        const double2 v = fft_get(lds, F, j);
        if (north) grid[rw.x * ld + j] = v.x;
        if (south) grid[rw.y * ld + j] = -v.y;
        if (ext_col && j == 0) {
            if (north) grid[rw.x * ld + n] = v.x;
            if (south) grid[rw.y * ld + n] = -v.y;
        }
    }
}

// Ring-pair analysis FFT for slot s; w = (w_north, w_south) include the
// quadrature weight and the reference's constant factors.
__device__ void anal_fft_tile(const FftDev &F, const double *__restrict__ grid, long long ld, int2 rw, int nrows,
                              double2 w, int P, int Lc, int s, double4 *__restrict__ gbuf, double2 *lds)
{
    const int n = F.n;
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

// =============================================================== plain kernels

__global__ __launch_bounds__(kThreads) void k_pack_synth(const double *__restrict__ cilm, int d1, int d2,
                                                         const double *__restrict__ s_tab, int Lt,
                                                         double2 *__restrict__ cpack, int Lc)
{
    const int m = blockIdx.x;
    const double *__restrict__ s = s_tab + tri_offd(m, Lt);
    double2 *__restrict__ o = cpack + tri_offd(m, Lc);
    for (int j = threadIdx.x; j <= Lc - m; j += blockDim.x) o[j] = pack_one(cilm, d1, d2, m + j, m, s[j]);
}

__global__ __launch_bounds__(kThreads) void k_synth_leg(const double *__restrict__ alpha, int Lt,
                                                        const double *__restrict__ seed,
                                                        const double *__restrict__ ringz,
                                                        const double *__restrict__ ringu,
                                                        const double2 *__restrict__ cpack, int Lc, int P,
                                                        int nrb, double4 *__restrict__ fbuf)
{
    const int m = blockIdx.x / nrb;
    const int rb = blockIdx.x - m * nrb;
    synth_leg_tile<kSynthK>(alpha + tri_offd(m, Lt), seed[m], ringz, ringu, cpack + tri_offd(m, Lc), Lc - m, m, P,
                            rb * kSynthK * kThreads, fbuf + (size_t)m * P);
}

// rows[s] = (north row, south row or -1); rows >= nrows are skipped.
__global__ __launch_bounds__(kThreads) void k_synth_fft(FftDev F, const double4 *__restrict__ fbuf, int P, int Lc,
                                                        const int2 *__restrict__ rows, int nrows,
                                                        double *__restrict__ grid, long long ld, int ext_col)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x;
    synth_fft_tile(F, fbuf, P, Lc, s, rows[s], nrows, grid, ld, ext_col, lds);
}

__global__ __launch_bounds__(kThreads) void k_anal_fft(FftDev F, const double *__restrict__ grid, long long ld,
                                                       const int2 *__restrict__ rows, int nrows,
                                                       const double2 *__restrict__ wts, int P, int Lc,
                                                       double4 *__restrict__ gbuf)
{
    extern __shared__ double2 lds[];
    const int s = blockIdx.x;
    anal_fft_tile(F, grid, ld, rows[s], nrows, wts[s], P, Lc, s, gbuf, lds);
}

__global__ __launch_bounds__(kThreads) void k_anal_leg(const double *__restrict__ alpha, int Lt,
                                                       const double *__restrict__ seed,
                                                       const double *__restrict__ ringz,
                                                       const double *__restrict__ ringu,
                                                       const double4 *__restrict__ gbuf, int Lc, int P,
                                                       double2 *__restrict__ cbuf)
{
    __shared__ double red[2 * kAnalLC * kThreads];
    const int m = blockIdx.x;
    anal_leg_tile(alpha + tri_offd(m, Lt), seed[m], ringz, ringu, gbuf + (size_t)m * P, Lc - m, m, P,
                  cbuf + tri_offd(m, Lc), red);
}

// cilm[i][l][m] = cbuf[tri(m) + l - m].i * post(l, m) for m <= l <= Lc,
// zero elsewhere (including S_l0), over the whole [2][d1][d2] array.
__device__ __forceinline__ double unpack_one(const double2 *__restrict__ cbuf, const double *__restrict__ post,
                                             int Lt, int Lc, int i, int l, int m)
{
    if (l > Lc || m > l || (i == 1 && m == 0)) return 0.0;
    const double2 c = cbuf[tri_offd(m, Lc) + (l - m)];
    return (i == 0 ? c.x : c.y) * post[tri_offd(m, Lt) + (l - m)];
}

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
        cilm[idx] = unpack_one(cbuf, post, Lt, Lc, i, l, m);
    }
}

// ============================================================= batched kernels

// Per-field description of a batch (all fields: lmax_calc = lmax).
struct BField {
    int L;           // degree (grid and expansion)
    int P;           // ring-pair slots
    int nlat;        // grid rows
    int nlong;       // grid columns (FFT length)
    int geo;         // offset of the field's slots in the concatenated ring tables
    int fft;         // index of the field's FFT plan
    long long coff;  // offset of the field in cpack/cbuf (double2 units)
    long long foff;  // offset of the field in fbuf/gbuf (double4 units)
};

struct BTables {
    const double *alpha, *s, *post, *seed;
    int Lt;
    const double *z, *u;
    const double2 *w;
    const int2 *rows;
    const FftDev *plans;
};

// items: (field, m)
__global__ __launch_bounds__(kThreads) void kb_pack(const BField *__restrict__ fields, const int2 *__restrict__ items,
                                                    BTables T, const double *const *__restrict__ cilm,
                                                    double2 *__restrict__ cpack)
{
    const int2 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const int m = it.y, L = f.L;
    const double *__restrict__ s = T.s + tri_offd(m, T.Lt);
    const double *__restrict__ c = cilm[it.x];
    double2 *__restrict__ o = cpack + f.coff + tri_offd(m, L);
    for (int j = threadIdx.x; j <= L - m; j += blockDim.x) o[j] = pack_one(c, L + 1, L + 1, m + j, m, s[j]);
}

// items: (field, m, first slot)
__global__ __launch_bounds__(kThreads) void kb_synth_leg(const BField *__restrict__ fields,
                                                         const int4 *__restrict__ items, BTables T,
                                                         const double2 *__restrict__ cpack,
                                                         double4 *__restrict__ fbuf)
{
    const int4 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const int m = it.y;
    synth_leg_tile<kSynthK>(T.alpha + tri_offd(m, T.Lt), T.seed[m], T.z + f.geo, T.u + f.geo,
                            cpack + f.coff + tri_offd(m, f.L), f.L - m, m, f.P, it.z,
                            fbuf + f.foff + (size_t)m * f.P);
}

// items: (field, slot); all fields of a launch share its LDS size.
__global__ __launch_bounds__(kThreads) void kb_synth_fft(const BField *__restrict__ fields,
                                                         const int2 *__restrict__ items, BTables T,
                                                         const double4 *__restrict__ fbuf, double *const *grid)
{
    extern __shared__ double2 lds[];
    const int2 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const FftDev F = T.plans[f.fft];
    synth_fft_tile(F, fbuf + f.foff, f.P, f.L, it.y, T.rows[f.geo + it.y], f.nlat, grid[it.x], f.nlong, 0, lds);
}

__global__ __launch_bounds__(kThreads) void kb_anal_fft(const BField *__restrict__ fields,
                                                        const int2 *__restrict__ items, BTables T,
                                                        const double *const *grid, double4 *__restrict__ gbuf)
{
    extern __shared__ double2 lds[];
    const int2 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const FftDev F = T.plans[f.fft];
    anal_fft_tile(F, grid[it.x], f.nlong, T.rows[f.geo + it.y], f.nlat, T.w[f.geo + it.y], f.P, f.L, it.y,
                  gbuf + f.foff, lds);
}

// items: (field, m)
__global__ __launch_bounds__(kThreads) void kb_anal_leg(const BField *__restrict__ fields,
                                                        const int2 *__restrict__ items, BTables T,
                                                        const double4 *__restrict__ gbuf,
                                                        double2 *__restrict__ cbuf)
{
    __shared__ double red[2 * kAnalLC * kThreads];
    const int2 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const int m = it.y;
    anal_leg_tile(T.alpha + tri_offd(m, T.Lt), T.seed[m], T.z + f.geo, T.u + f.geo,
                  gbuf + f.foff + (size_t)m * f.P, f.L - m, m, f.P, cbuf + f.coff + tri_offd(m, f.L), red);
}

// items: (field, l): writes row l of C and S of the field's [2][L+1][L+1].
__global__ __launch_bounds__(kThreads) void kb_unpack(const BField *__restrict__ fields,
                                                      const int2 *__restrict__ items, BTables T,
                                                      const double2 *__restrict__ cbuf, double *const *cilm)
{
    const int2 it = items[blockIdx.x];
    const BField f = fields[it.x];
    const int l = it.y, L = f.L;
    double *__restrict__ c = cilm[it.x];
    for (int k = threadIdx.x; k < 2 * (L + 1); k += blockDim.x) {
        const int i = k / (L + 1), m = k - i * (L + 1);
        c[((size_t)i * (L + 1) + l) * (L + 1) + m] = unpack_one(cbuf + f.coff, T.post, T.Lt, L, i, l, m);
    }
}

}  // namespace shtg
