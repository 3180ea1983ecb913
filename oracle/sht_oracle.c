/*
 * ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the SHTOOLS Fortran 95 spherical-harmonic
 * transforms (real, scalar, FP64) that the GPU library replaces.  It is the
 * checker for parity tests, for `__graft_entry__.smoke()` and for the CPU
 * baseline leg of `bench.py`; the shipped library never links, loads or calls
 * it.  Every routine follows the reference's loop structure and arithmetic
 * order so that it reproduces the reference's rounding closely:
 *
 *   or_dhaj            src/DHaj.f95:91-105
 *   or_preglq          src/PreGLQ.f95:93-149 (x1=-1, x2=1 as src/SHGLQ.f95:182-192)
 *   or_glqgridcoord    src/GLQGridCoord.f95:100-119
 *   build_tables       src/MakeGridDH.f95:258-386 (identical in SHExpandDH,
 *                      MakeGridGLQ, SHExpandGLQ)
 *   or_plm             the transforms' own recursion for one z (tables above,
 *                      seeds MakeGridDH.f95:542-603, scaling :228,:552)
 *   or_makegriddh      src/MakeGridDH.f95:107-648
 *   or_shexpanddh      src/SHExpandDH.f95:118-704
 *   or_makegridglq     src/MakeGridGLQ.f95:192-678 (on-the-fly branch)
 *   or_shexpandglq     src/SHExpandGLQ.f95:100-691 (on-the-fly branch)
 *
 * The longitude stage of the reference calls FFTW 3.3.x (not part of the
 * reference tree): r2c X_m = sum_k x_k e^{-2 pi i m k/N}, c2r
 * x_k = sum_m X_m e^{+2 pi i m k/N} over the Hermitian extension, both
 * unnormalised (src/FFTW3.f95:212-218,265-271).  Here that DFT is computed by
 * radix-2 Cooley-Tukey for powers of two and by Bluestein's chirp-z
 * convolution otherwise.
 *
 * Arrays use the Python (C-order) layouts of the pyshtools API:
 * cilm[2][d1][d2] with cilm[i][l][m], grid[nlat][nlon] with rings contiguous.
 * Return value: the reference's exitstatus codes (0 ok, 1 dimensions,
 * 2 bounds, 3 allocation, 5 other).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* ------------------------------------------------------------------ FFT */

typedef struct {
    int n;       /* transform length */
    int m;       /* power-of-two work length (== n when n is a power of 2) */
    cplx *tw;    /* exp(-2 pi i k / m), k < m/2 */
    cplx *chirp; /* Bluestein: exp(-i pi (j^2 mod 2n) / n), j < n */
    cplx *bhat;  /* Bluestein: FFT_m of the conjugate chirp, scaled by 1/m */
    cplx *work;  /* m scratch */
} fftplan;

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void fft_pow2(cplx *a, int n, const cplx *tw, int inverse)
{
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cplx t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int k = 0; k < half; k++) {
            cplx w = tw[k * step];
            if (inverse) w = conj(w);
            for (int i = k; i < n; i += len) {
                cplx u = a[i], v = a[i + half] * w;
                a[i] = u + v;
                a[i + half] = u - v;
            }
        }
    }
}

static fftplan *plan_create(int n)
{
    fftplan *p = (fftplan *)calloc(1, sizeof(fftplan));
    if (!p) return NULL;
    p->n = n;
    if (is_pow2(n)) p->m = n;
    else { p->m = 1; while (p->m < 2 * n - 1) p->m <<= 1; }
    int m = p->m;
    p->tw = (cplx *)malloc(sizeof(cplx) * (m / 2 + 1));
    p->work = (cplx *)malloc(sizeof(cplx) * m);
    for (int k = 0; k < m / 2 + 1; k++) {
        double ang = -2.0 * M_PI * (double)k / (double)m;
        p->tw[k] = cos(ang) + I * sin(ang);
    }
    if (m != n) {
        p->chirp = (cplx *)malloc(sizeof(cplx) * n);
        p->bhat = (cplx *)malloc(sizeof(cplx) * m);
        for (int j = 0; j < n; j++) {
            long long q = ((long long)j * j) % (2LL * n);
            double ang = -M_PI * (double)q / (double)n;
            p->chirp[j] = cos(ang) + I * sin(ang);
        }
        for (int j = 0; j < m; j++) p->bhat[j] = 0;
        p->bhat[0] = conj(p->chirp[0]);
        for (int j = 1; j < n; j++) {
            p->bhat[j] = conj(p->chirp[j]);
            p->bhat[m - j] = conj(p->chirp[j]);
        }
        fft_pow2(p->bhat, m, p->tw, 0);
        for (int j = 0; j < m; j++) p->bhat[j] /= (double)m;
    }
    return p;
}

static void plan_destroy(fftplan *p)
{
    if (!p) return;
    free(p->tw); free(p->work); free(p->chirp); free(p->bhat);
    free(p);
}

/* In-place unnormalised DFT, e^{-} (forward) or e^{+} (inverse). */
static void plan_exec(fftplan *p, cplx *x, int inverse)
{
    int n = p->n, m = p->m;
    if (m == n) { fft_pow2(x, n, p->tw, inverse); return; }
    cplx *a = p->work;
    for (int j = 0; j < n; j++) a[j] = (inverse ? conj(x[j]) : x[j]) * p->chirp[j];
    for (int j = n; j < m; j++) a[j] = 0;
    fft_pow2(a, m, p->tw, 0);
    for (int j = 0; j < m; j++) a[j] *= p->bhat[j];
    fft_pow2(a, m, p->tw, 1);
    for (int k = 0; k < n; k++) {
        cplx v = a[k] * p->chirp[k];
        x[k] = inverse ? conj(v) : v;
    }
}

/* FFTW-style r2c: out[0..n/2] from real in[0..n-1]. */
static void r2c(fftplan *p, const double *in, cplx *buf, cplx *out)
{
    int n = p->n;
    for (int k = 0; k < n; k++) buf[k] = in[k];
    plan_exec(p, buf, 0);
    for (int k = 0; k <= n / 2; k++) out[k] = buf[k];
}

/* FFTW-style c2r: real out[0..n-1] from the Hermitian half in[0..n/2]. */
static void c2r(fftplan *p, const cplx *in, cplx *buf, double *out)
{
    int n = p->n;
    buf[0] = creal(in[0]);
    for (int k = 1; k < (n + 1) / 2; k++) {
        buf[k] = in[k];
        buf[n - k] = conj(in[k]);
    }
    if (n % 2 == 0) buf[n / 2] = creal(in[n / 2]);
    plan_exec(p, buf, 1);
    for (int k = 0; k < n; k++) out[k] = creal(buf[k]);
}

/* Exported for tests of the longitude stage against numpy.fft. */
int or_r2c(int n, const double *in, double *out_ri)
{
    fftplan *p = plan_create(n);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * n), *o = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
    r2c(p, in, buf, o);
    for (int k = 0; k <= n / 2; k++) { out_ri[2 * k] = creal(o[k]); out_ri[2 * k + 1] = cimag(o[k]); }
    free(buf); free(o); plan_destroy(p);
    return 0;
}

int or_c2r(int n, const double *in_ri, double *out)
{
    fftplan *p = plan_create(n);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * n), *i2 = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
    for (int k = 0; k <= n / 2; k++) i2[k] = in_ri[2 * k] + I * in_ri[2 * k + 1];
    c2r(p, i2, buf, out);
    free(buf); free(i2); plan_destroy(p);
    return 0;
}

/* ------------------------------------------------------ quadrature */

/* src/DHaj.f95:91-105 */
int or_dhaj(int n, int extend, double *aj)
{
    if (n % 2 != 0) return 2;
    if (extend != 0 && extend != 1) return 2;
    double pi = acos(-1.0);
    for (int j = 0; j <= n - 1; j++) {
        double sum1 = 0.0;
        for (int l = 0; l <= n / 2 - 1; l++)
            sum1 = sum1 + sin((double)(2 * l + 1) * pi * (double)j / (double)n) / (double)(2 * l + 1);
        aj[j] = sum1 * sin(pi * (double)j / (double)n) * sqrt(8.0) / (double)n;
    }
    if (extend == 1) aj[n] = 0.0;
    return 0;
}

/* src/PreGLQ.f95:93-149 with x1 = -1, x2 = 1. */
int or_preglq(int n, double *zero, double *w)
{
    const int itermax = 1000;
    const double eps = 1.0e-15;
    double pi = acos(-1.0);
    double xm = 0.0, xu = 1.0;   /* (x2+x1)/2, (x2-x1)/2 */
    for (int i = 0; i < n; i++) { zero[i] = 0.0; w[i] = 0.0; }
    int mh = (n + 1) / 2;
    for (int i = 1; i <= mh; i++) {
        int iter = 0;
        double z = cos(pi * (i - 0.25) / (n + 0.5)), z1, p1, p2, p3, pp;
        for (;;) {
            iter++;
            p1 = 1.0; p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                p3 = p2; p2 = p1;
                p1 = ((double)(2 * j - 1) * z * p2 - (double)(j - 1) * p3) / (double)j;
            }
            pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) <= eps) break;
            if (iter > itermax) return 5;
        }
        zero[i - 1] = xm + xu * z;
        zero[n - i] = xm - xu * z;
        w[i - 1] = 2.0 * xu / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
    return 0;
}

/* src/GLQGridCoord.f95:100-119 */
int or_glqgridcoord(int lmax, int extend, double *lat, double *lon)
{
    if (extend != 0 && extend != 1) return 2;
    int nlat = lmax + 1, nlong = 2 * lmax + 1, nlong_out = nlong + extend;
    double *zero = (double *)malloc(sizeof(double) * nlat), *w = (double *)malloc(sizeof(double) * nlat);
    int st = or_preglq(nlat, zero, w);
    if (st == 0) {
        double pi = acos(-1.0);
        for (int i = 0; i < nlong_out; i++) lon[i] = 360.0 * (double)i / (double)nlong;
        for (int i = 0; i < nlat; i++) lat[i] = asin(zero[i]) * 180.0 / pi;
    }
    free(zero); free(w);
    return st;
}

/* ------------------------------------------------- recursion tables */

typedef struct {
    int lc;        /* lmax_comp */
    double *sqr;   /* sqr[k] = sqrt(k), k = 0..2lc+1 (index 0 unused) */
    double *ff1;   /* ff1[l*(lc+1)+m] */
    double *ff2;
} tables;

/* column-major like the reference's ff1(l+1,m+1): consecutive l contiguous */
#define FF(t, a, l, m) ((t)->a[(size_t)(m) * ((t)->lc + 1) + (l)])

/* src/MakeGridDH.f95:291-381 */
static int build_tables(tables *t, int lc, int norm)
{
    t->lc = lc;
    t->sqr = (double *)calloc((size_t)(2 * lc + 2), sizeof(double));
    t->ff1 = (double *)calloc((size_t)(lc + 1) * (lc + 1), sizeof(double));
    t->ff2 = (double *)calloc((size_t)(lc + 1) * (lc + 1), sizeof(double));
    if (!t->sqr || !t->ff1 || !t->ff2) return 3;
    double *sqr = t->sqr;
    for (int l = 1; l <= 2 * lc + 1; l++) sqr[l] = sqrt((double)l);
    int l, m;
    switch (norm) {
    case 1: case 4:
        if (lc != 0) { FF(t, ff1, 1, 0) = sqr[3]; FF(t, ff2, 1, 0) = 0.0; }
        for (l = 2; l <= lc; l++) {
            FF(t, ff1, l, 0) = sqr[2 * l - 1] * sqr[2 * l + 1] / (double)l;
            FF(t, ff2, l, 0) = (double)(l - 1) * sqr[2 * l + 1] / sqr[2 * l - 3] / (double)l;
            for (m = 1; m <= l - 2; m++) {
                FF(t, ff1, l, m) = sqr[2 * l + 1] * sqr[2 * l - 1] / sqr[l + m] / sqr[l - m];
                FF(t, ff2, l, m) = sqr[2 * l + 1] * sqr[l - m - 1] * sqr[l + m - 1] / sqr[2 * l - 3] / sqr[l + m] / sqr[l - m];
            }
            m = l - 1;   /* value of the DO variable after the loop */
            FF(t, ff1, l, l - 1) = sqr[2 * l + 1] * sqr[2 * l - 1] / sqr[l + m] / sqr[l - m];
            FF(t, ff2, l, l - 1) = 0.0;
        }
        break;
    case 2:
        if (lc != 0) { FF(t, ff1, 1, 0) = 1.0; FF(t, ff2, 1, 0) = 0.0; }
        for (l = 2; l <= lc; l++) {
            FF(t, ff1, l, 0) = (double)(2 * l - 1) / (double)l;
            FF(t, ff2, l, 0) = (double)(l - 1) / (double)l;
            for (m = 1; m <= l - 2; m++) {
                FF(t, ff1, l, m) = (double)(2 * l - 1) / sqr[l + m] / sqr[l - m];
                FF(t, ff2, l, m) = sqr[l - m - 1] * sqr[l + m - 1] / sqr[l + m] / sqr[l - m];
            }
            m = l - 1;
            FF(t, ff1, l, l - 1) = (double)(2 * l - 1) / sqr[l + m] / sqr[l - m];
            FF(t, ff2, l, l - 1) = 0.0;
        }
        break;
    case 3:
        for (l = 1; l <= lc; l++) {
            FF(t, ff1, l, 0) = (double)(2 * l - 1) / (double)l;
            FF(t, ff2, l, 0) = (double)(l - 1) / (double)l;
            for (m = 1; m <= l - 1; m++) {
                FF(t, ff1, l, m) = (double)(2 * l - 1) / (double)(l - m);
                FF(t, ff2, l, m) = (double)(l + m - 1) / (double)(l - m);
            }
        }
        break;
    }
    return 0;
}

static void free_tables(tables *t) { free(t->sqr); free(t->ff1); free(t->ff2); }

static inline int fsym(int l, int m) { return ((l - m) % 2 == 0) ? 1 : -1; }

/* Seed of order 0: MakeGridDH.f95:516-519 */
static double pm00(int norm) { return norm == 4 ? 1.0 / sqrt(4.0 * acos(-1.0)) : 1.0; }

/* Scaled seed chain start: MakeGridDH.f95:542-546 */
static double pmm_start(int norm, const double *sqr)
{
    double scalef = 1.0e-280, pi = acos(-1.0);
    switch (norm) {
    case 1: case 2: return sqr[2] * scalef;
    case 3: return scalef;
    default: return sqr[2] * scalef / sqrt(4.0 * pi);
    }
}

/* One step of the seed chain for 1 <= m < lc: returns pm2 (the P_mm used in
 * the sums) and updates pmm.  MakeGridDH.f95:554-564 */
static double pmm_step(int norm, int phase, double *pmm, int m, const double *sqr)
{
    switch (norm) {
    case 1: case 4:
        *pmm = phase * *pmm * sqr[2 * m + 1] / sqr[2 * m];
        return *pmm;
    case 2:
        *pmm = phase * *pmm * sqr[2 * m + 1] / sqr[2 * m];
        return *pmm / sqr[2 * m + 1];
    default:
        *pmm = phase * *pmm * (double)(2 * m - 1);
        return *pmm;
    }
}

/* Scaled seed chain start of the complex transforms, whose functions carry
 * no sqrt(2) for m > 0: MakeGridDHC.f95:485-489, SHExpandDHC.f95:495-498. */
static double pmm_start_c(int norm)
{
    double scalef = 1.0e-280, pi = acos(-1.0);
    return norm == 4 ? scalef / sqrt(4.0 * pi) : scalef;
}

/* Last sectoral term m = lc (MakeGridDH.f95:595-603), times rescalem. */
static double pmm_last(int norm, int phase, double pmm, int lc, const double *sqr, double rescalem)
{
    switch (norm) {
    case 1: case 4: return phase * pmm * sqr[2 * lc + 1] / sqr[2 * lc] * rescalem;
    case 2: return phase * pmm / sqr[2 * lc] * rescalem;
    default: return phase * pmm * (double)(2 * lc - 1) * rescalem;
    }
}

/* All P_lm(z), l,m <= lmax, with the transforms' recursion, seeds and
 * 1e-280 scaling; p[l*(lmax+1)+m].  cplx: the complex normalisation, whose
 * sectoral chain starts without sqrt(2) (src/PlmBar.f95:242-253; norm 3 has
 * none either way). */
static int plm_eval(int lmax, int norm, int csphase, int cplx, double z, double *p)
{
    if (norm < 1 || norm > 4 || (csphase != 1 && csphase != -1) || lmax < 0) return 2;
    tables t;
    int st = build_tables(&t, lmax, norm);
    if (st) { free_tables(&t); return st; }
    const double *sqr = t.sqr;
    int L1 = lmax + 1;
    for (int i = 0; i < L1 * L1; i++) p[i] = 0.0;
    double u = sqrt((1.0 - z) * (1.0 + z));
    double pm2 = pm00(norm), pm1, pv;
    p[0] = pm2;
    if (lmax == 0) { free_tables(&t); return 0; }
    pm1 = FF(&t, ff1, 1, 0) * z * pm2;
    p[1 * L1] = pm1;
    for (int l = 2; l <= lmax; l++) {
        pv = FF(&t, ff1, l, 0) * z * pm1 - FF(&t, ff2, l, 0) * pm2;
        p[l * L1] = pv; pm2 = pm1; pm1 = pv;
    }
    double pmm = cplx ? pmm_start_c(norm) : pmm_start(norm, sqr), rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lmax - 1; m++) {
        rescalem = rescalem * u;
        pm2 = pmm_step(norm, csphase, &pmm, m, sqr);
        p[m * L1 + m] = pm2 * rescalem;
        pm1 = z * FF(&t, ff1, m + 1, m) * pm2;
        p[(m + 1) * L1 + m] = pm1 * rescalem;
        for (int l = m + 2; l <= lmax; l++) {
            pv = z * FF(&t, ff1, l, m) * pm1 - FF(&t, ff2, l, m) * pm2;
            pm2 = pm1; pm1 = pv;
            p[l * L1 + m] = pv * rescalem;
        }
    }
    rescalem = rescalem * u;
    p[lmax * L1 + lmax] = pmm_last(norm, csphase, pmm, lmax, sqr, rescalem);
    free_tables(&t);
    return 0;
}

/* Used to pin the recursion against closed forms. */
int or_plm(int lmax, int norm, int csphase, double z, double *p)
{
    return plm_eval(lmax, norm, csphase, 0, z, p);
}

/* The table of SHGLQ's optional argument plx (src/SHGLQ.f95:202-257): the
 * Legendre functions at the Gauss-Legendre nodes zero[0..lmax], stored as
 * the Fortran array plx(lmax+1, (lmax+1)(lmax+2)/2) in C order,
 * plx[k * (lmax+1) + i] with k = l(l+1)/2 + m.  Nodes i < (n+1)/2 are
 * evaluated (the reference calls PlmBar / PlmSchmidt / PLegendreA / PlmON,
 * cnorm passed to all but PLegendreA); node n-1-i takes them times
 * (-1)^(l-m).  The functions here are those of the transforms' recursion,
 * which the tests compare with the reference's own table. */
int or_plx_table(int lmax, const double *zero, int nzero, int norm, int csphase, int cnorm, double *plx)
{
    if (lmax < 0) return 2;
    if (nzero < lmax + 1) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != 1 && csphase != -1) return 2;
    const int n = lmax + 1, L1 = lmax + 1;
    double *p = (double *)malloc(sizeof(double) * (size_t)L1 * L1);
    if (!p) return 3;
    for (int i = 1; i <= (n + 1) / 2; i++) {
        int st = plm_eval(lmax, norm, csphase, norm != 3 && cnorm == 1, zero[i - 1], p);
        if (st) { free(p); return st; }
        const int mirror = !(i == (n + 1) / 2 && n % 2 != 0);
        const int i_s = n + 1 - i;
        for (int l = 0; l <= lmax; l++)
            for (int m = 0; m <= l; m++) {
                const size_t k = (size_t)l * (l + 1) / 2 + m;
                plx[k * n + (i - 1)] = p[(size_t)l * L1 + m];
                if (mirror) plx[k * n + (i_s - 1)] = plx[k * n + (i - 1)] * (double)fsym(l, m);
            }
    }
    free(p);
    return 0;
}

/* ------------------------------------------------------- synthesis */

/* Coefficients in the reference's column-major order cf[m][l][i] (what f2py
 * hands the Fortran: cilm(i,l,m) with i fastest), l, m <= lc. */
static double *to_fortran_order(const double *cilm, int d1, int d2, int lc)
{
    size_t L1 = (size_t)lc + 1;
    double *cf = (double *)calloc(2 * L1 * L1, sizeof(double));
    if (!cf) return NULL;
    for (int m = 0; m <= lc; m++)
        for (int l = 0; l <= lc; l++)
            for (int i = 0; i < 2; i++)
                cf[(m * L1 + l) * 2 + i] = cilm[((size_t)i * d1 + l) * d2 + m];
    return cf;
}

#define CF(i, l, m) cf[((size_t)(m) * (lc + 1) + (l)) * 2 + (i)]

/* Per-ring Legendre sums for one ring (north: cr/ci) and its mirror
 * (south: csr/csi); entry m>=1 is scaled by rescalem but NOT yet halved.
 * (C - iS) * p is evaluated component-wise.  MakeGridDH.f95:510-609 /
 * MakeGridGLQ.f95:548-642. */
static void synth_ring(const tables *t, int norm, int phase, const double *cf,
                       double z, double u, double *cr, double *ci, double *csr, double *csi,
                       double *c0, double *c0s)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
    double coef0 = 0.0, coef0s = 0.0, tempr, pm2, pm1, p, pmm, rescalem, tr, ti;
    pm2 = pm00(norm);
    tempr = CF(0, 0, 0) * pm2;
    coef0 += tempr; coef0s += tempr;
    pm1 = FF(t, ff1, 1, 0) * z * pm2;
    tempr = CF(0, 1, 0) * pm1;
    coef0 += tempr; coef0s -= tempr;
    for (int l = 2; l <= lc; l++) {
        p = FF(t, ff1, l, 0) * z * pm1 - FF(t, ff2, l, 0) * pm2;
        tempr = CF(0, l, 0) * p;
        coef0 += tempr;
        coef0s += tempr * fsym(l, 0);
        pm2 = pm1; pm1 = p;
    }
    pmm = pmm_start(norm, sqr);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        double ar = 0.0, ai = 0.0, sr = 0.0, si = 0.0;
        rescalem = rescalem * u;
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        tr = CF(0, m, m) * pm2; ti = -CF(1, m, m) * pm2;
        ar += tr; ai += ti; sr += tr; si += ti;
        pm1 = z * FF(t, ff1, m + 1, m) * pm2;
        tr = CF(0, m + 1, m) * pm1; ti = -CF(1, m + 1, m) * pm1;
        ar += tr; ai += ti; sr -= tr; si -= ti;
        for (int l = m + 2; l <= lc; l++) {
            p = z * FF(t, ff1, l, m) * pm1 - FF(t, ff2, l, m) * pm2;
            pm2 = pm1; pm1 = p;
            tr = CF(0, l, m) * p; ti = -CF(1, l, m) * p;
            ar += tr; ai += ti;
            if ((l - m) % 2 == 0) { sr += tr; si += ti; } else { sr -= tr; si -= ti; }
        }
        cr[m] = ar * rescalem; ci[m] = ai * rescalem;
        csr[m] = sr * rescalem; csi[m] = si * rescalem;
    }
    rescalem = rescalem * u;
    pmm = pmm_last(norm, phase, pmm, lc, sqr, rescalem);
    tr = CF(0, lc, lc) * pmm; ti = -CF(1, lc, lc) * pmm;
    cr[lc] += tr; ci[lc] += ti; csr[lc] += tr; csi[lc] += ti;
    *c0 = coef0; *c0s = coef0s;
}

/* Equator ring (z = 0, u = 1): MakeGridDH.f95:421-496; entry m>=1 is
 * returned not yet scaled by rescalem/2. */
static void synth_equator(const tables *t, int norm, int phase, const double *cf,
                          double *cr, double *ci, double *c0, double *rescale_out)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
    double coef0 = 0.0, pm2, p, pmm, rescalem;
    pm2 = pm00(norm);
    coef0 += CF(0, 0, 0) * pm2;
    for (int l = 2; l <= lc; l += 2) {
        p = -FF(t, ff2, l, 0) * pm2;
        pm2 = p;
        coef0 += CF(0, l, 0) * p;
    }
    pmm = pmm_start(norm, sqr);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        cr[m] += CF(0, m, m) * pm2; ci[m] += -CF(1, m, m) * pm2;
        for (int l = m + 2; l <= lc; l += 2) {
            p = -FF(t, ff2, l, m) * pm2;
            cr[m] += CF(0, l, m) * p; ci[m] += -CF(1, l, m) * p;
            pm2 = p;
        }
    }
    pmm = pmm_last(norm, phase, pmm, lc, sqr, 1.0);
    cr[lc] += CF(0, lc, lc) * pmm; ci[lc] += -CF(1, lc, lc) * pmm;
    *c0 = coef0;
    *rescale_out = rescalem;
}

static void pack_half(cplx *coef, const double *cr, const double *ci, int nmax, int ncl)
{
    for (int m = 0; m < ncl; m++) coef[m] = (m < nmax) ? (cr[m] + I * ci[m]) : 0.0;
}

int or_makegriddh(const double *cilm, int d1, int d2, int lmax, int norm, int sampling,
                  int csphase, int lmax_calc, int extend, double *grid, int g0, int g1)
{
    int n = 2 * lmax + 2, nlong, nlat_out, nlong_out;
    if (sampling == 1) nlong = n; else if (sampling == 2) nlong = 2 * n; else return 2;
    if (extend == 0) { nlat_out = n; nlong_out = nlong; }
    else if (extend == 1) { nlat_out = n + 1; nlong_out = nlong + 1; }
    else return 2;
    if (g0 < nlat_out || g1 < nlong_out) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    if (lmax_calc > lmax) return 2;
    int lc = lmax;
    if (d2 - 1 < lc) lc = d2 - 1;
    if (d1 - 1 < lc) lc = d1 - 1;
    if (lmax_calc < lc) lc = lmax_calc;
    double pi = acos(-1.0);
#define G(i, k) grid[(size_t)(i) * g1 + (k)]
    if (lc == 0) {
        double pm2 = (norm == 4) ? 1.0 / sqrt(4.0 * pi) : 1.0;
        for (int i = 0; i < nlat_out; i++)
            for (int k = 0; k < nlong_out; k++) G(i, k) = cilm[0] * pm2;
        return 0;
    }
    tables t;
    int st = build_tables(&t, lc, norm);
    double *cf = to_fortran_order(cilm, d1, d2, lc);
    if (st || !cf) { free_tables(&t); free(cf); return 3; }
    fftplan *plan = plan_create(nlong);
    int ncl = nlong / 2 + 1;
    cplx *coef = (cplx *)malloc(sizeof(cplx) * ncl);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    double *row = (double *)malloc(sizeof(double) * nlong);
    double *cr = (double *)malloc(sizeof(double) * 4 * (lmax + 2));
    double *ci = cr + (lmax + 2), *csr = ci + (lmax + 2), *csi = csr + (lmax + 2);
    int i_eq = n / 2 + 1;   /* 1-based, as in the reference */
    double c0, c0s, rescalem;

    /* equator: MakeGridDH.f95:421-502 */
    for (int k = 0; k < 4 * (lmax + 2); k++) cr[k] = 0.0;
    synth_equator(&t, norm, csphase, cf, cr, ci, &c0, &rescalem);
    cr[0] = c0; ci[0] = 0.0;
    for (int m = 1; m <= lmax; m++) { cr[m] = cr[m] * rescalem / 2.0; ci[m] = ci[m] * rescalem / 2.0; }
    pack_half(coef, cr, ci, lmax + 1, ncl);
    c2r(plan, coef, buf, row);
    for (int k = 0; k < nlong; k++) G(i_eq - 1, k) = row[k];

    /* ring pairs: MakeGridDH.f95:504-639.  Rings are independent, so they
     * are spread over OpenMP threads (each with its own FFT plan and
     * buffers); the arithmetic of every ring is unchanged. */
#pragma omp parallel
    {
        fftplan *tplan = plan_create(nlong);
        cplx *tcoef = (cplx *)malloc(sizeof(cplx) * ncl);
        cplx *tbuf = (cplx *)malloc(sizeof(cplx) * nlong);
        double *trow = (double *)malloc(sizeof(double) * nlong);
        double *tcr = (double *)malloc(sizeof(double) * 4 * (lmax + 2));
        double *tci = tcr + (lmax + 2), *tcsr = tci + (lmax + 2), *tcsi = tcsr + (lmax + 2);
#pragma omp for schedule(dynamic, 1)
        for (int i = 1; i <= i_eq - 1; i++) {
            int i_s = 2 * i_eq - i;
            double theta = pi * (double)(i - 1) / (double)n;
            double z = cos(theta);
            double u = sqrt((1.0 - z) * (1.0 + z));
            double d0, d0s;
            for (int k = 0; k < 4 * (lmax + 2); k++) tcr[k] = 0.0;
            synth_ring(&t, norm, csphase, cf, z, u, tcr, tci, tcsr, tcsi, &d0, &d0s);
            tcr[0] = d0; tci[0] = 0.0;
            for (int m = 1; m <= lmax; m++) { tcr[m] = tcr[m] / 2.0; tci[m] = tci[m] / 2.0; }
            pack_half(tcoef, tcr, tci, lmax + 1, ncl);
            c2r(tplan, tcoef, tbuf, trow);
            for (int k = 0; k < nlong; k++) G(i - 1, k) = trow[k];
            if (!(i == 1 && extend == 0)) {
                tcsr[0] = d0s; tcsi[0] = 0.0;
                for (int m = 1; m <= lmax; m++) { tcsr[m] = tcsr[m] / 2.0; tcsi[m] = tcsi[m] / 2.0; }
                pack_half(tcoef, tcsr, tcsi, lmax + 1, ncl);
                c2r(tplan, tcoef, tbuf, trow);
                for (int k = 0; k < nlong; k++) G(i_s - 1, k) = trow[k];
            }
        }
        free(tcoef); free(tbuf); free(trow); free(tcr);
        plan_destroy(tplan);
    }
    (void)c0s;
    if (extend == 1)
        for (int i = 0; i < nlat_out; i++) G(i, nlong_out - 1) = G(i, 0);
#undef G
    free(coef); free(buf); free(row); free(cr); free(cf);
    plan_destroy(plan);
    free_tables(&t);
    return 0;
}

int or_makegridglq(const double *cilm, int d1, int d2, int lmax, const double *zero, int nzero,
                   int norm, int csphase, int lmax_calc, int extend, double *grid, int g0, int g1)
{
    int nlong = 2 * lmax + 1, nlat = lmax + 1, nlong_out;
    if (extend == 0) nlong_out = nlong; else if (extend == 1) nlong_out = nlong + 1; else return 2;
    if (g1 < nlong_out || g0 < nlat) return 1;
    if (!zero) return 5;
    if (nzero < lmax + 1) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    if (lmax_calc > lmax) return 2;
    int lc = lmax;
    if (d2 - 1 < lc) lc = d2 - 1;
    if (d1 - 1 < lc) lc = d1 - 1;
    if (lmax_calc < lc) lc = lmax_calc;
    double pi = acos(-1.0);
#define G(i, k) grid[(size_t)(i) * g1 + (k)]
    if (lc == 0) {
        double pm2 = (norm == 4) ? 1.0 / sqrt(4.0 * pi) : 1.0;
        for (int i = 0; i < nlat; i++)
            for (int k = 0; k < nlong_out; k++) G(i, k) = cilm[0] * pm2;
        return 0;
    }
    tables t;
    int st = build_tables(&t, lc, norm);
    double *cf = to_fortran_order(cilm, d1, d2, lc);
    if (st || !cf) { free_tables(&t); free(cf); return 3; }
    fftplan *plan = plan_create(nlong);
    int ncl = nlong / 2 + 1;   /* == lmax + 1 */
    cplx *coef = (cplx *)malloc(sizeof(cplx) * ncl);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    double *row = (double *)malloc(sizeof(double) * nlong);
    double *cr = (double *)malloc(sizeof(double) * 4 * (lmax + 2));
    double *ci = cr + (lmax + 2), *csr = ci + (lmax + 2), *csi = csr + (lmax + 2);
    free(coef); free(buf); free(row); free(cr);
    plan_destroy(plan);
    /* rings spread over OpenMP threads as in or_makegriddh */
#pragma omp parallel
    {
    fftplan *plan = plan_create(nlong);
    cplx *coef = (cplx *)malloc(sizeof(cplx) * ncl);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    double *row = (double *)malloc(sizeof(double) * nlong);
    double *cr = (double *)malloc(sizeof(double) * 4 * (lmax + 2));
    double *ci = cr + (lmax + 2), *csr = ci + (lmax + 2), *csi = csr + (lmax + 2);
#pragma omp for schedule(dynamic, 1)
    for (int i = 1; i <= (nlat + 1) / 2; i++) {
        double c0, c0s, rescalem;
        for (int k = 0; k < 4 * (lmax + 2); k++) cr[k] = 0.0;
        if (i == (nlat + 1) / 2 && nlat % 2 != 0) {
            /* exact equator: MakeGridGLQ.f95:472-545 */
            synth_equator(&t, norm, csphase, cf, cr, ci, &c0, &rescalem);
            cr[0] = c0; ci[0] = 0.0;
            for (int m = 1; m <= lc; m++) { cr[m] = cr[m] * rescalem / 2.0; ci[m] = ci[m] * rescalem / 2.0; }
            pack_half(coef, cr, ci, lc + 1, ncl);
            c2r(plan, coef, buf, row);
            for (int k = 0; k < nlong; k++) G(i - 1, k) = row[k];
        } else {
            double z = zero[i - 1];
            double u = sqrt((1.0 - z) * (1.0 + z));
            int i_s = nlat + 1 - i;
            synth_ring(&t, norm, csphase, cf, z, u, cr, ci, csr, csi, &c0, &c0s);
            cr[0] = c0; ci[0] = 0.0;
            for (int m = 1; m <= lc; m++) { cr[m] = cr[m] / 2.0; ci[m] = ci[m] / 2.0; }
            pack_half(coef, cr, ci, lc + 1, ncl);
            c2r(plan, coef, buf, row);
            for (int k = 0; k < nlong; k++) G(i - 1, k) = row[k];
            csr[0] = c0s; csi[0] = 0.0;
            for (int m = 1; m <= lc; m++) { csr[m] = csr[m] / 2.0; csi[m] = csi[m] / 2.0; }
            pack_half(coef, csr, csi, lc + 1, ncl);
            c2r(plan, coef, buf, row);
            for (int k = 0; k < nlong; k++) G(i_s - 1, k) = row[k];
        }
    }
    free(coef); free(buf); free(row); free(cr);
    plan_destroy(plan);
    }
    if (extend == 1)
        for (int i = 0; i < nlat; i++) G(i, nlong_out - 1) = G(i, 0);
#undef G
    free(cf);
    free_tables(&t);
    return 0;
}

/* -------------------------------------------------------- analysis */

/* Accumulate one ring pair into cf (column-major [m][l][i]); f1/f2 are the
 * weighted Fourier coefficients [i][m] of the north/south rings.
 * SHExpandDH.f95:481-580 / SHExpandGLQ.f95:555-645. */
static void anal_ring(const tables *t, int norm, int phase, double z, double u,
                      double *f1, double *f2, int nf, double *cf)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
#define F1(i, m) f1[(size_t)(i) * nf + (m)]
#define F2(i, m) f2[(size_t)(i) * nf + (m)]
    double pm2, pm1, p, pmm, rescalem, ffc[2][2];
    pm2 = pm00(norm);
    CF(0, 0, 0) += pm2 * (F1(0, 0) + F2(0, 0));
    if (lc == 0) return;
    pm1 = FF(t, ff1, 1, 0) * z * pm2;
    CF(0, 1, 0) += pm1 * (F1(0, 0) - F2(0, 0));
    ffc[0][0] = F1(0, 0) - F2(0, 0);   /* [.][0] : fsymsign = -1 */
    ffc[0][1] = F1(0, 0) + F2(0, 0);   /* [.][1] : fsymsign = +1 */
    for (int l = 2; l <= lc; l++) {
        p = FF(t, ff1, l, 0) * z * pm1 - FF(t, ff2, l, 0) * pm2;
        pm2 = pm1; pm1 = p;
        CF(0, l, 0) += p * ffc[0][fsym(l, 0) > 0];
    }
    pmm = pmm_start(norm, sqr);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        rescalem = rescalem * u;
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        for (int i = 0; i < 2; i++) { F1(i, m) = F1(i, m) * rescalem; F2(i, m) = F2(i, m) * rescalem; }
        for (int i = 0; i < 2; i++) CF(i, m, m) += pm2 * (F1(i, m) + F2(i, m));
        pm1 = z * FF(t, ff1, m + 1, m) * pm2;
        for (int i = 0; i < 2; i++) CF(i, m + 1, m) += pm1 * (F1(i, m) - F2(i, m));
        for (int i = 0; i < 2; i++) {
            ffc[i][0] = F1(i, m) - F2(i, m);
            ffc[i][1] = F1(i, m) + F2(i, m);
        }
        for (int l = m + 2; l <= lc; l++) {
            p = z * FF(t, ff1, l, m) * pm1 - FF(t, ff2, l, m) * pm2;
            pm2 = pm1; pm1 = p;
            int s = ((l - m) % 2 == 0);
            CF(0, l, m) += p * ffc[0][s];
            CF(1, l, m) += p * ffc[1][s];
        }
    }
    rescalem = rescalem * u;
    pmm = pmm_last(norm, phase, pmm, lc, sqr, rescalem);
    for (int i = 0; i < 2; i++) CF(i, lc, lc) += pmm * (F1(i, lc) + F2(i, lc));
#undef F1
#undef F2
}

/* Equator: SHExpandDH.f95:584-664 / SHExpandGLQ.f95:470-553. */
static void anal_equator(const tables *t, int norm, int phase, double *f1, int nf, double *cf)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
#define F1(i, m) f1[(size_t)(i) * nf + (m)]
    double pm2, p, pmm, rescalem;
    pm2 = pm00(norm);
    CF(0, 0, 0) += pm2 * F1(0, 0);
    if (lc == 0) return;
    for (int l = 2; l <= lc; l += 2) {
        p = -FF(t, ff2, l, 0) * pm2;
        pm2 = p;
        CF(0, l, 0) += p * F1(0, 0);
    }
    pmm = pmm_start(norm, sqr);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        for (int i = 0; i < 2; i++) F1(i, m) = F1(i, m) * rescalem;
        for (int i = 0; i < 2; i++) CF(i, m, m) += pm2 * F1(i, m);
        for (int l = m + 2; l <= lc; l += 2) {
            p = -FF(t, ff2, l, m) * pm2;
            pm2 = p;
            for (int i = 0; i < 2; i++) CF(i, l, m) += p * F1(i, m);
        }
    }
    pmm = pmm_last(norm, phase, pmm, lc, sqr, rescalem);
    for (int i = 0; i < 2; i++) CF(i, lc, lc) += pmm * F1(i, lc);
#undef F1
}

/* Copy the column-major accumulator back to cilm[2][d1][d2] and apply the
 * normalisation epilogue: SHExpandDH.f95:674-702. */
static void anal_epilogue(int norm, int lc, const double *cf, double *cilm, int d1, int d2)
{
#define CI(i, l, m) cilm[((size_t)(i) * d1 + (l)) * d2 + (m)]
    for (int l = 0; l <= lc; l++)
        for (int m = 0; m <= l; m++)
            for (int i = 0; i < 2; i++) CI(i, l, m) = CF(i, l, m);
    double pi = acos(-1.0), prod;
    switch (norm) {
    case 1:
        for (int l = 0; l <= lc; l++)
            for (int m = 0; m <= l; m++)
                for (int i = 0; i < 2; i++) CI(i, l, m) = CI(i, l, m) / (4 * pi);
        break;
    case 2:
        for (int l = 0; l <= lc; l++)
            for (int m = 0; m <= l; m++)
                for (int i = 0; i < 2; i++) CI(i, l, m) = CI(i, l, m) * (2 * l + 1) / (4 * pi);
        break;
    case 3:
        for (int l = 0; l <= lc; l++) {
            prod = 4 * pi / (double)(2 * l + 1);
            CI(0, l, 0) = CI(0, l, 0) / prod;
            prod = prod / 2.0;
            for (int m = 1; m <= l - 1; m++) {
                prod = prod * (l + m) * (l - m + 1);
                for (int i = 0; i < 2; i++) CI(i, l, m) = CI(i, l, m) / prod;
            }
            if (l != 0)
                for (int i = 0; i < 2; i++) CI(i, l, l) = CI(i, l, l) / (prod * 2 * l);
        }
        break;
    default:
        break;
    }
#undef CI
}

int or_shexpanddh(const double *grid, int n, int g1, int norm, int sampling, int csphase,
                  int lmax_calc, double *cilm, int d1, int d2)
{
    int lmax = n / 2 - 1, lc, nlong;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (sampling != 1 && sampling != 2) return 2;
    if (n % 2 != 0) return 2;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    nlong = (sampling == 1) ? n : 2 * n;
    if (g1 < nlong) return 1;
    if (csphase != -1 && csphase != 1) return 2;
    if (norm < 1 || norm > 4) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0.0;
    double *aj = (double *)malloc(sizeof(double) * (n + 1));
    or_dhaj(n, 0, aj);
    for (int j = 0; j < n; j++) aj[j] = aj[j] * sqrt(4.0 * pi);
    tables t;
    int st = build_tables(&t, lc, norm);
    double *cf = (double *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(double));
    if (st || !cf) { free_tables(&t); free(aj); free(cf); return 3; }
    fftplan *plan = plan_create(nlong);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    cplx *cc = (cplx *)malloc(sizeof(cplx) * (nlong / 2 + 1));
    int nf = n / 2 + 1;
    double *f1 = (double *)calloc((size_t)2 * nf, sizeof(double));
    double *f2 = (double *)calloc((size_t)2 * nf, sizeof(double));
    int i_eq = n / 2 + 1;
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
    for (int i = 2; i <= i_eq - 1; i++) {
        double theta = (i - 1) * pi / (double)n;
        double z = cos(theta);
        double u = sqrt((1.0 - z) * (1.0 + z));
        r2c(plan, &GR(i - 1, 0), buf, cc);
        for (int m = 0; m < n / 2; m++) {
            f1[m] = sqrt(2 * pi) * aj[i - 1] * creal(cc[m]) / (double)nlong;
            f1[nf + m] = -sqrt(2 * pi) * aj[i - 1] * cimag(cc[m]) / (double)nlong;
        }
        int i_s = 2 * i_eq - i;
        r2c(plan, &GR(i_s - 1, 0), buf, cc);
        for (int m = 0; m < n / 2; m++) {
            f2[m] = sqrt(2 * pi) * aj[i_s - 1] * creal(cc[m]) / (double)nlong;
            f2[nf + m] = -sqrt(2 * pi) * aj[i_s - 1] * cimag(cc[m]) / (double)nlong;
        }
        anal_ring(&t, norm, csphase, z, u, f1, f2, nf, cf);
    }
    {
        int i = i_eq;
        r2c(plan, &GR(i - 1, 0), buf, cc);
        for (int m = 0; m < n / 2; m++) {
            f1[m] = sqrt(2 * pi) * aj[i - 1] * creal(cc[m]) / (double)nlong;
            f1[nf + m] = -sqrt(2 * pi) * aj[i - 1] * cimag(cc[m]) / (double)nlong;
        }
        anal_equator(&t, norm, csphase, f1, nf, cf);
    }
#undef GR
    anal_epilogue(norm, lc, cf, cilm, d1, d2);
    free(buf); free(cc); free(f1); free(f2); free(aj); free(cf);
    plan_destroy(plan);
    free_tables(&t);
    return 0;
}

/* ------------------------------------- order-parallel analysis (same sums)
 *
 * The sequential analysis above visits the ring pairs in order and, per
 * ring, every order m; a coefficient (l, m) therefore accumulates its ring
 * contributions in ring order, and nothing of order m depends on another
 * order except two chains that do not depend on the ring data: the seeds
 * pmm (pmm_step) and, per ring, the iterated rescalem = u^m / 1e-280.  So the
 * orders can be summed in parallel, each over the rings in the same order,
 * from precomputed Fourier coefficients, seeds and rescale factors: every
 * coefficient comes out bitwise as in the sequential loop (checked by the
 * tests).  Rings r = 0 .. nr-1 are the ring pairs in processing order
 * (z[r], u[r], north/south coefficients f1/f2 + r * 2 nf), and fe, if not
 * NULL, the equator processed after them. */
static void anal_orders(const tables *t, int norm, int phase, int nr, const double *z, const double *u,
                        const double *f1, const double *f2, const double *fe, int nf, double *cf)
{
    const double *sqr = t->sqr;
    const int lc = t->lc;
    double *seed = (double *)malloc(sizeof(double) * (size_t)(lc + 2));
    double *resc = (double *)malloc(sizeof(double) * (size_t)(nr > 0 ? nr : 1) * (lc + 1));
    double pmm = pmm_start(norm, sqr), pmm_fin;
    for (int m = 1; m <= lc - 1; m++) seed[m] = pmm_step(norm, phase, &pmm, m, sqr);
    pmm_fin = pmm;
    for (int r = 0; r < nr; r++) {
        double rescalem = 1.0 / 1.0e-280;
        for (int m = 1; m <= lc; m++) {
            rescalem = rescalem * u[r];
            resc[(size_t)r * (lc + 1) + m] = rescalem;
        }
    }
#define F1(r, i, m) f1[((size_t)(r) * 2 + (i)) * nf + (m)]
#define F2(r, i, m) f2[((size_t)(r) * 2 + (i)) * nf + (m)]
#define FE(i, m) fe[(size_t)(i) * nf + (m)]
#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m <= lc; m++) {
        double pm2, pm1, p, ffc[2][2];
        if (m == 0) {
            for (int r = 0; r < nr; r++) {
                pm2 = pm00(norm);
                CF(0, 0, 0) += pm2 * (F1(r, 0, 0) + F2(r, 0, 0));
                if (lc == 0) continue;
                pm1 = FF(t, ff1, 1, 0) * z[r] * pm2;
                CF(0, 1, 0) += pm1 * (F1(r, 0, 0) - F2(r, 0, 0));
                ffc[0][0] = F1(r, 0, 0) - F2(r, 0, 0);
                ffc[0][1] = F1(r, 0, 0) + F2(r, 0, 0);
                for (int l = 2; l <= lc; l++) {
                    p = FF(t, ff1, l, 0) * z[r] * pm1 - FF(t, ff2, l, 0) * pm2;
                    pm2 = pm1; pm1 = p;
                    CF(0, l, 0) += p * ffc[0][fsym(l, 0) > 0];
                }
            }
            if (fe) {
                pm2 = pm00(norm);
                CF(0, 0, 0) += pm2 * FE(0, 0);
                for (int l = 2; l <= lc; l += 2) {
                    p = -FF(t, ff2, l, 0) * pm2;
                    pm2 = p;
                    CF(0, l, 0) += p * FE(0, 0);
                }
            }
        } else if (m <= lc - 1) {
            for (int r = 0; r < nr; r++) {
                const double rescalem = resc[(size_t)r * (lc + 1) + m];
                double a1[2], a2[2];
                for (int i = 0; i < 2; i++) { a1[i] = F1(r, i, m) * rescalem; a2[i] = F2(r, i, m) * rescalem; }
                pm2 = seed[m];
                for (int i = 0; i < 2; i++) CF(i, m, m) += pm2 * (a1[i] + a2[i]);
                pm1 = z[r] * FF(t, ff1, m + 1, m) * pm2;
                for (int i = 0; i < 2; i++) CF(i, m + 1, m) += pm1 * (a1[i] - a2[i]);
                for (int i = 0; i < 2; i++) {
                    ffc[i][0] = a1[i] - a2[i];
                    ffc[i][1] = a1[i] + a2[i];
                }
                for (int l = m + 2; l <= lc; l++) {
                    p = z[r] * FF(t, ff1, l, m) * pm1 - FF(t, ff2, l, m) * pm2;
                    pm2 = pm1; pm1 = p;
                    int s = ((l - m) % 2 == 0);
                    CF(0, l, m) += p * ffc[0][s];
                    CF(1, l, m) += p * ffc[1][s];
                }
            }
            if (fe) {
                const double rescalem = 1.0 / 1.0e-280;
                double a[2];
                for (int i = 0; i < 2; i++) a[i] = FE(i, m) * rescalem;
                pm2 = seed[m];
                for (int i = 0; i < 2; i++) CF(i, m, m) += pm2 * a[i];
                for (int l = m + 2; l <= lc; l += 2) {
                    p = -FF(t, ff2, l, m) * pm2;
                    pm2 = p;
                    for (int i = 0; i < 2; i++) CF(i, l, m) += p * a[i];
                }
            }
        } else {  /* m == lc >= 1: the last sectoral term */
            for (int r = 0; r < nr; r++) {
                const double pl = pmm_last(norm, phase, pmm_fin, lc, sqr, resc[(size_t)r * (lc + 1) + lc]);
                for (int i = 0; i < 2; i++) CF(i, lc, lc) += pl * (F1(r, i, lc) + F2(r, i, lc));
            }
            if (fe) {
                const double pl = pmm_last(norm, phase, pmm_fin, lc, sqr, 1.0 / 1.0e-280);
                for (int i = 0; i < 2; i++) CF(i, lc, lc) += pl * FE(i, lc);
            }
        }
    }
#undef F1
#undef F2
#undef FE
    free(seed);
    free(resc);
}

/* SHExpandDH with the orders summed in parallel (anal_orders): the same
 * coefficients, bitwise, as or_shexpanddh, in a fraction of the time on
 * several cores.  Ring Fourier coefficients are computed in parallel too,
 * each thread with its own FFT plan. */
int or_shexpanddh_par(const double *grid, int n, int g1, int norm, int sampling, int csphase,
                      int lmax_calc, double *cilm, int d1, int d2)
{
    int lmax = n / 2 - 1, lc, nlong;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (sampling != 1 && sampling != 2) return 2;
    if (n % 2 != 0) return 2;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    nlong = (sampling == 1) ? n : 2 * n;
    if (g1 < nlong) return 1;
    if (csphase != -1 && csphase != 1) return 2;
    if (norm < 1 || norm > 4) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0.0;
    double *aj = (double *)malloc(sizeof(double) * (n + 1));
    or_dhaj(n, 0, aj);
    for (int j = 0; j < n; j++) aj[j] = aj[j] * sqrt(4.0 * pi);
    tables t;
    int st = build_tables(&t, lc, norm);
    const int nf = n / 2 + 1, i_eq = n / 2 + 1, nr = i_eq - 2;  /* pairs i = 2 .. i_eq - 1 */
    double *cf = (double *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(double));
    double *z = (double *)malloc(sizeof(double) * (size_t)(nr + 1));
    double *u = (double *)malloc(sizeof(double) * (size_t)(nr + 1));
    double *f1 = (double *)calloc((size_t)2 * nf * (nr + 1), sizeof(double));
    double *f2 = (double *)calloc((size_t)2 * nf * (nr + 1), sizeof(double));
    double *fe = (double *)calloc((size_t)2 * nf, sizeof(double));
    if (st || !cf || !z || !u || !f1 || !f2 || !fe) {
        free_tables(&t); free(aj); free(cf); free(z); free(u); free(f1); free(f2); free(fe);
        return 3;
    }
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
#pragma omp parallel
    {
        fftplan *plan = plan_create(nlong);
        cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
        cplx *cc = (cplx *)malloc(sizeof(cplx) * (nlong / 2 + 1));
#pragma omp for schedule(dynamic, 4)
        for (int i = 2; i <= i_eq; i++) {
            const int r = i - 2;
            double *a = (i == i_eq) ? fe : f1 + (size_t)r * 2 * nf;
            r2c(plan, &GR(i - 1, 0), buf, cc);
            for (int m = 0; m < n / 2; m++) {
                a[m] = sqrt(2 * pi) * aj[i - 1] * creal(cc[m]) / (double)nlong;
                a[nf + m] = -sqrt(2 * pi) * aj[i - 1] * cimag(cc[m]) / (double)nlong;
            }
            if (i == i_eq) continue;
            double theta = (i - 1) * pi / (double)n;
            z[r] = cos(theta);
            u[r] = sqrt((1.0 - z[r]) * (1.0 + z[r]));
            int i_s = 2 * i_eq - i;
            double *b = f2 + (size_t)r * 2 * nf;
            r2c(plan, &GR(i_s - 1, 0), buf, cc);
            for (int m = 0; m < n / 2; m++) {
                b[m] = sqrt(2 * pi) * aj[i_s - 1] * creal(cc[m]) / (double)nlong;
                b[nf + m] = -sqrt(2 * pi) * aj[i_s - 1] * cimag(cc[m]) / (double)nlong;
            }
        }
        free(buf); free(cc);
        plan_destroy(plan);
    }
#undef GR
    anal_orders(&t, norm, csphase, nr, z, u, f1, f2, fe, nf, cf);
    anal_epilogue(norm, lc, cf, cilm, d1, d2);
    free(aj); free(cf); free(z); free(u); free(f1); free(f2); free(fe);
    free_tables(&t);
    return 0;
}

/* SHExpandGLQ with the orders summed in parallel (as or_shexpanddh_par). */
int or_shexpandglq_par(const double *grid, int g0, int g1, int lmax, const double *w, int nw,
                       const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                       double *cilm, int d1, int d2)
{
    int lc, nlong = 2 * lmax + 1, nlat = lmax + 1;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    if (g1 < 2 * lmax + 1 || g0 < lmax + 1) return 1;
    if (nw < lmax + 1) return 1;
    if (!zero) return 5;
    if (nzero < lmax + 1) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0.0;
    tables t;
    int st = build_tables(&t, lc, norm);
    const int nf = lmax + 1, half = (nlat + 1) / 2, has_eq = nlat % 2 != 0;
    const int nr = has_eq ? half - 1 : half;
    double *cf = (double *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(double));
    double *z = (double *)malloc(sizeof(double) * (size_t)(nr + 1));
    double *u = (double *)malloc(sizeof(double) * (size_t)(nr + 1));
    double *f1 = (double *)calloc((size_t)2 * nf * (nr + 1), sizeof(double));
    double *f2 = (double *)calloc((size_t)2 * nf * (nr + 1), sizeof(double));
    double *fe = (double *)calloc((size_t)2 * nf, sizeof(double));
    if (st || !cf || !z || !u || !f1 || !f2 || !fe) {
        free_tables(&t); free(cf); free(z); free(u); free(f1); free(f2); free(fe);
        return 3;
    }
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
#pragma omp parallel
    {
        fftplan *plan = plan_create(nlong);
        cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
        cplx *cc = (cplx *)malloc(sizeof(cplx) * (nlong / 2 + 1));
#pragma omp for schedule(dynamic, 4)
        for (int i = 1; i <= half; i++) {
            const int r = i - 1, eq = has_eq && i == half;
            double *a = eq ? fe : f1 + (size_t)r * 2 * nf;
            r2c(plan, &GR(i - 1, 0), buf, cc);
            for (int m = 0; m <= lmax; m++) {
                a[m] = 2 * w[i - 1] * pi * creal(cc[m]) / (double)nlong;
                a[nf + m] = -2 * w[i - 1] * pi * cimag(cc[m]) / (double)nlong;
            }
            if (eq) continue;
            z[r] = zero[i - 1];
            u[r] = sqrt((1.0 - z[r]) * (1.0 + z[r]));
            int i_s = nlat + 1 - i;
            double *b = f2 + (size_t)r * 2 * nf;
            r2c(plan, &GR(i_s - 1, 0), buf, cc);
            for (int m = 0; m <= lmax; m++) {
                b[m] = 2 * w[i_s - 1] * pi * creal(cc[m]) / (double)nlong;
                b[nf + m] = -2 * w[i_s - 1] * pi * cimag(cc[m]) / (double)nlong;
            }
        }
        free(buf); free(cc);
        plan_destroy(plan);
    }
#undef GR
    anal_orders(&t, norm, csphase, nr, z, u, f1, f2, has_eq ? fe : NULL, nf, cf);
    anal_epilogue(norm, lc, cf, cilm, d1, d2);
    free(cf); free(z); free(u); free(f1); free(f2); free(fe);
    free_tables(&t);
    return 0;
}

int or_shexpandglq(const double *grid, int g0, int g1, int lmax, const double *w, int nw,
                   const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                   double *cilm, int d1, int d2)
{
    int lc, nlong = 2 * lmax + 1, nlat = lmax + 1;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    if (g1 < 2 * lmax + 1 || g0 < lmax + 1) return 1;
    if (nw < lmax + 1) return 1;
    if (!zero) return 5;
    if (nzero < lmax + 1) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0.0;
    tables t;
    int st = build_tables(&t, lc, norm);
    double *cf = (double *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(double));
    if (st || !cf) { free_tables(&t); free(cf); return 3; }
    fftplan *plan = plan_create(nlong);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    cplx *cc = (cplx *)malloc(sizeof(cplx) * (nlong / 2 + 1));
    int nf = lmax + 1;
    double *f1 = (double *)calloc((size_t)2 * nf, sizeof(double));
    double *f2 = (double *)calloc((size_t)2 * nf, sizeof(double));
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
    for (int i = 1; i <= (nlat + 1) / 2; i++) {
        r2c(plan, &GR(i - 1, 0), buf, cc);
        for (int m = 0; m <= lmax; m++) {
            f1[m] = 2 * w[i - 1] * pi * creal(cc[m]) / (double)nlong;
            f1[nf + m] = -2 * w[i - 1] * pi * cimag(cc[m]) / (double)nlong;
        }
        if (i == (nlat + 1) / 2 && nlat % 2 != 0) {
            anal_equator(&t, norm, csphase, f1, nf, cf);
        } else {
            double z = zero[i - 1];
            double u = sqrt((1.0 - z) * (1.0 + z));
            int i_s = nlat + 1 - i;
            r2c(plan, &GR(i_s - 1, 0), buf, cc);
            for (int m = 0; m <= lmax; m++) {
                f2[m] = 2 * w[i_s - 1] * pi * creal(cc[m]) / (double)nlong;
                f2[nf + m] = -2 * w[i_s - 1] * pi * cimag(cc[m]) / (double)nlong;
            }
            anal_ring(&t, norm, csphase, z, u, f1, f2, nf, cf);
        }
    }
#undef GR
    anal_epilogue(norm, lc, cf, cilm, d1, d2);
    free(buf); free(cc); free(f1); free(f2); free(cf);
    plan_destroy(plan);
    free_tables(&t);
    return 0;
}

/* ------------------------------------------- GLQ with a plx table */

/* MakeGridGLQ with plx (src/MakeGridGLQ.f95:107-270 checks, :408-420 for
 * lmax_comp = 0, :441-465): every row from the table, degrees summed from
 * lmax_comp down to 0.  plx as or_plx_table lays it out (first dimension
 * lmax + 1). */
int or_makegridglq_plx(const double *cilm, int d1, int d2, int lmax, const double *plx, int norm, int csphase,
                       int lmax_calc, int extend, double *grid, int g0, int g1)
{
    int nlong = 2 * lmax + 1, nlat = lmax + 1, nlong_out;
    if (extend == 0) nlong_out = nlong; else if (extend == 1) nlong_out = nlong + 1; else return 2;
    if (g1 < nlong_out || g0 < nlat) return 1;
    if (!plx) return 5;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    if (lmax_calc > lmax) return 2;
    int lc = lmax;
    if (d2 - 1 < lc) lc = d2 - 1;
    if (d1 - 1 < lc) lc = d1 - 1;
    if (lmax_calc < lc) lc = lmax_calc;
    double pi = acos(-1.0);
#define G(i, k) grid[(size_t)(i) * g1 + (k)]
    if (lc == 0) {
        double pm2 = (norm == 4) ? 1.0 / sqrt(4.0 * pi) : 1.0;
        for (int i = 0; i < nlat; i++)
            for (int k = 0; k < nlong_out; k++) G(i, k) = cilm[0] * pm2;
        return 0;
    }
    double *cf = to_fortran_order(cilm, d1, d2, lc);
    if (!cf) return 3;
    int ncl = nlong / 2 + 1;   /* == lmax + 1 */
#pragma omp parallel
    {
    fftplan *plan = plan_create(nlong);
    cplx *coef = (cplx *)malloc(sizeof(cplx) * ncl);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    double *row = (double *)malloc(sizeof(double) * nlong);
    double *cr = (double *)malloc(sizeof(double) * 2 * (lmax + 2));
    double *ci = cr + (lmax + 2);
#pragma omp for schedule(dynamic, 1)
    for (int i = 1; i <= nlat; i++) {
        double coef0 = 0.0;
        for (int k = 0; k < 2 * (lmax + 2); k++) cr[k] = 0.0;
        for (int l = lc; l >= 0; l--) {
            size_t k = (size_t)l * (l + 1) / 2;   /* m = 0 */
            coef0 = coef0 + CF(0, l, 0) * plx[k * nlat + (i - 1)];
            for (int m = 1; m <= l; m++) {
                k = (size_t)l * (l + 1) / 2 + m;
                const double p = plx[k * nlat + (i - 1)];
                cr[m] = cr[m] + CF(0, l, m) * p / 2.0;
                ci[m] = ci[m] + (-CF(1, l, m)) * p / 2.0;
            }
        }
        cr[0] = coef0; ci[0] = 0.0;
        pack_half(coef, cr, ci, lmax + 1, ncl);
        c2r(plan, coef, buf, row);
        for (int k = 0; k < nlong; k++) G(i - 1, k) = row[k];
    }
    free(coef); free(buf); free(row); free(cr);
    plan_destroy(plan);
    }
    if (extend == 1)
        for (int i = 0; i < nlat; i++) G(i, nlong_out - 1) = G(i, 0);
#undef G
    free(cf);
    return 0;
}

/* SHExpandGLQ with plx (src/SHExpandGLQ.f95:100-250 checks, :419-441 and the
 * epilogue :674-702): every ring's weighted Fourier coefficients times the
 * table, rings in order, then the normalisation. */
int or_shexpandglq_plx(const double *grid, int g0, int g1, int lmax, const double *w, int nw,
                       const double *plx, int norm, int csphase, int lmax_calc, double *cilm, int d1, int d2)
{
    int lc, nlong = 2 * lmax + 1, nlat = lmax + 1;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    if (g1 < 2 * lmax + 1 || g0 < lmax + 1) return 1;
    if (nw < lmax + 1) return 1;
    if (!plx) return 5;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0.0;
    double *cf = (double *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(double));
    if (!cf) return 3;
    fftplan *plan = plan_create(nlong);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    cplx *cc = (cplx *)malloc(sizeof(cplx) * (nlong / 2 + 1));
    int nf = lmax + 1;
    double *f1 = (double *)calloc((size_t)2 * nf, sizeof(double));
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
    for (int i = 1; i <= nlat; i++) {
        r2c(plan, &GR(i - 1, 0), buf, cc);
        for (int m = 0; m <= lmax; m++) {
            f1[m] = 2 * w[i - 1] * pi * creal(cc[m]) / (double)nlong;
            f1[nf + m] = -2 * w[i - 1] * pi * cimag(cc[m]) / (double)nlong;
        }
        size_t k = 0;
        for (int l = 0; l <= lc; l++)
            for (int m = 0; m <= l; m++, k++) {
                const double p = plx[k * nlat + (i - 1)];
                CF(0, l, m) = CF(0, l, m) + p * f1[m];
                CF(1, l, m) = CF(1, l, m) + p * f1[nf + m];
            }
    }
#undef GR
    anal_epilogue(norm, lc, cf, cilm, d1, d2);
    free(buf); free(cc); free(f1); free(cf);
    plan_destroy(plan);
    return 0;
}

/* ------------------------------------------------ complex transforms */

/* Complex coefficients cilm[2][d1][d2] (complex, [0] positive orders C_lm,
 * [1] negative orders C_l,-m) in the reference's column-major order
 * cc[m][l][i], l, m <= lc. */
static cplx *to_fortran_order_c(const cplx *cilm, int d1, int d2, int lc)
{
    size_t L1 = (size_t)lc + 1;
    cplx *cc = (cplx *)calloc(2 * L1 * L1, sizeof(cplx));
    if (!cc) return NULL;
    for (int m = 0; m <= lc; m++)
        for (int l = 0; l <= lc; l++)
            for (int i = 0; i < 2; i++)
                cc[(m * L1 + l) * 2 + i] = cilm[((size_t)i * d1 + l) * d2 + m];
    return cc;
}

#define CCF(i, l, m) cc[((size_t)(m) * (lc + 1) + (l)) * 2 + (i)]
static inline double msign(int m) { return (m % 2 == 0) ? 1.0 : -1.0; }

/* One ring (coef) and its mirror (coefs) of the complex synthesis:
 * coef[m] = sum_l C_lm P_lm, coef[nlong-m] = (-1)^m sum_l C_l,-m P_lm.
 * MakeGridDHC.f95:531-609 / MakeGridGLQC.f95:556-668. */
static void synth_ring_c(const tables *t, int norm, int phase, const cplx *cc, double z, double u,
                         cplx *coef, cplx *coefs, int nlong)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
    double pm2, pm1, p, pmm, rescalem;
    cplx tempc;
    for (int k = 0; k < nlong; k++) { coef[k] = 0; coefs[k] = 0; }
    pm2 = pm00(norm);
    tempc = CCF(0, 0, 0) * pm2;
    coef[0] += tempc; coefs[0] += tempc;
    pm1 = FF(t, ff1, 1, 0) * z * pm2;
    tempc = CCF(0, 1, 0) * pm1;
    coef[0] += tempc; coefs[0] -= tempc;
    for (int l = 2; l <= lc; l++) {
        p = FF(t, ff1, l, 0) * z * pm1 - FF(t, ff2, l, 0) * pm2;
        tempc = CCF(0, l, 0) * p;
        coef[0] += tempc; coefs[0] += tempc * fsym(l, 0);
        pm2 = pm1; pm1 = p;
    }
    pmm = pmm_start_c(norm);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        rescalem = rescalem * u;
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        tempc = CCF(0, m, m) * pm2; coef[m] += tempc; coefs[m] += tempc;
        tempc = CCF(1, m, m) * pm2; coef[nlong - m] += tempc; coefs[nlong - m] += tempc;
        pm1 = z * FF(t, ff1, m + 1, m) * pm2;
        tempc = CCF(0, m + 1, m) * pm1; coef[m] += tempc; coefs[m] -= tempc;
        tempc = CCF(1, m + 1, m) * pm1; coef[nlong - m] += tempc; coefs[nlong - m] -= tempc;
        for (int l = m + 2; l <= lc; l++) {
            p = z * FF(t, ff1, l, m) * pm1 - FF(t, ff2, l, m) * pm2;
            pm2 = pm1; pm1 = p;
            tempc = CCF(0, l, m) * p; coef[m] += tempc; coefs[m] += tempc * fsym(l, m);
            tempc = CCF(1, l, m) * p; coef[nlong - m] += tempc; coefs[nlong - m] += tempc * fsym(l, m);
        }
        coef[m] *= rescalem; coefs[m] *= rescalem;
        coef[nlong - m] *= rescalem * msign(m); coefs[nlong - m] *= rescalem * msign(m);
    }
    rescalem = rescalem * u;
    pmm = pmm_last(norm, phase, pmm, lc, sqr, rescalem);
    tempc = CCF(0, lc, lc) * pmm; coef[lc] += tempc; coefs[lc] += tempc;
    tempc = CCF(1, lc, lc) * pmm * msign(lc); coef[nlong - lc] += tempc; coefs[nlong - lc] += tempc;
}

/* Equator (z = 0, u = 1): MakeGridDHC.f95:452-529 / MakeGridGLQC.f95:473-553. */
static void synth_equator_c(const tables *t, int norm, int phase, const cplx *cc, cplx *coef, int nlong)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
    double pm2, p, pmm, rescalem;
    for (int k = 0; k < nlong; k++) coef[k] = 0;
    pm2 = pm00(norm);
    coef[0] += CCF(0, 0, 0) * pm2;
    for (int l = 2; l <= lc; l += 2) {
        p = -FF(t, ff2, l, 0) * pm2;
        pm2 = p;
        coef[0] += CCF(0, l, 0) * p;
    }
    pmm = pmm_start_c(norm);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        coef[m] += CCF(0, m, m) * pm2;
        coef[nlong - m] += CCF(1, m, m) * pm2;
        for (int l = m + 2; l <= lc; l += 2) {
            p = -FF(t, ff2, l, m) * pm2;
            coef[m] += CCF(0, l, m) * p;
            coef[nlong - m] += CCF(1, l, m) * p;
            pm2 = p;
        }
        coef[m] *= rescalem;
        coef[nlong - m] *= rescalem * msign(m);
    }
    pmm = pmm_last(norm, phase, pmm, lc, sqr, rescalem);
    coef[lc] += CCF(0, lc, lc) * pmm;
    coef[nlong - lc] += CCF(1, lc, lc) * pmm * msign(lc);
}

static int lc_of_makegrid(int d1, int d2, int lmax, int lmax_calc)
{
    int lc = lmax;
    if (d2 - 1 < lc) lc = d2 - 1;
    if (d1 - 1 < lc) lc = d1 - 1;
    if (lmax_calc < lc) lc = lmax_calc;
    return lc;
}

/* src/MakeGridDHC.f95:1-643: grid[g0][g1] complex (interleaved doubles). */
int or_makegriddhc(const double *cilm_ri, int d1, int d2, int lmax, int norm, int sampling,
                   int csphase, int lmax_calc, int extend, double *grid_ri, int g0, int g1)
{
    const cplx *cilm = (const cplx *)cilm_ri;
    cplx *grid = (cplx *)grid_ri;
    int n = 2 * lmax + 2, nlong, nlat_out, nlong_out;
    if (sampling == 1) nlong = n; else if (sampling == 2) nlong = 2 * n; else return 2;
    if (extend == 0) { nlat_out = n; nlong_out = nlong; }
    else if (extend == 1) { nlat_out = n + 1; nlong_out = nlong + 1; }
    else return 2;
    if (g0 < nlat_out || g1 < nlong_out) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    if (lmax_calc > lmax) return 2;
    int lc = lc_of_makegrid(d1, d2, lmax, lmax_calc);
#define G(i, k) grid[(size_t)(i) * g1 + (k)]
    if (lc == 0) {
        double pm2 = pm00(norm);
        for (int i = 0; i < nlat_out; i++)
            for (int k = 0; k < nlong_out; k++) G(i, k) = cilm[0] * pm2;
        return 0;
    }
    tables t;
    int st = build_tables(&t, lc, norm);
    cplx *cc = to_fortran_order_c(cilm, d1, d2, lc);
    if (st || !cc) { free_tables(&t); free(cc); return 3; }
    int i_eq = n / 2 + 1;
    {
        fftplan *plan = plan_create(nlong);
        cplx *coef = (cplx *)malloc(sizeof(cplx) * nlong);
        synth_equator_c(&t, norm, csphase, cc, coef, nlong);
        plan_exec(plan, coef, 1);
        for (int k = 0; k < nlong; k++) G(i_eq - 1, k) = coef[k];
        free(coef);
        plan_destroy(plan);
    }
    const double pi = acos(-1.0);
#pragma omp parallel
    {
        fftplan *plan = plan_create(nlong);
        cplx *coef = (cplx *)malloc(sizeof(cplx) * nlong);
        cplx *coefs = (cplx *)malloc(sizeof(cplx) * nlong);
#pragma omp for schedule(dynamic, 1)
        for (int i = 1; i <= i_eq - 1; i++) {
            int i_s = 2 * i_eq - i;
            double theta = pi * (double)(i - 1) / (double)n;
            double z = cos(theta);
            double u = sqrt((1.0 - z) * (1.0 + z));
            synth_ring_c(&t, norm, csphase, cc, z, u, coef, coefs, nlong);
            plan_exec(plan, coef, 1);
            for (int k = 0; k < nlong; k++) G(i - 1, k) = coef[k];
            if (!(i == 1 && extend == 0)) {
                plan_exec(plan, coefs, 1);
                for (int k = 0; k < nlong; k++) G(i_s - 1, k) = coefs[k];
            }
        }
        free(coef); free(coefs);
        plan_destroy(plan);
    }
    if (extend == 1)
        for (int i = 0; i < nlat_out; i++) G(i, nlong_out - 1) = G(i, 0);
#undef G
    free(cc);
    free_tables(&t);
    return 0;
}

/* src/MakeGridGLQC.f95:1-688, on-the-fly branch (:471-678). */
int or_makegridglqc(const double *cilm_ri, int d1, int d2, int lmax, const double *zero, int nzero,
                    int norm, int csphase, int lmax_calc, int extend, double *grid_ri, int g0, int g1)
{
    const cplx *cilm = (const cplx *)cilm_ri;
    cplx *grid = (cplx *)grid_ri;
    int nlong = 2 * lmax + 1, nlat = lmax + 1, nlong_out;
    if (extend == 0) nlong_out = nlong; else if (extend == 1) nlong_out = nlong + 1; else return 2;
    if (g1 < nlong_out || g0 < nlat) return 1;
    if (!zero) return 5;
    if (nzero < lmax + 1) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    if (lmax_calc > lmax) return 2;
    int lc = lc_of_makegrid(d1, d2, lmax, lmax_calc);
#define G(i, k) grid[(size_t)(i) * g1 + (k)]
    if (lc == 0) {
        double pm2 = pm00(norm);
        for (int i = 0; i < nlat; i++)
            for (int k = 0; k < nlong_out; k++) G(i, k) = cilm[0] * pm2;
        return 0;
    }
    tables t;
    int st = build_tables(&t, lc, norm);
    cplx *cc = to_fortran_order_c(cilm, d1, d2, lc);
    if (st || !cc) { free_tables(&t); free(cc); return 3; }
#pragma omp parallel
    {
        fftplan *plan = plan_create(nlong);
        cplx *coef = (cplx *)malloc(sizeof(cplx) * nlong);
        cplx *coefs = (cplx *)malloc(sizeof(cplx) * nlong);
#pragma omp for schedule(dynamic, 1)
        for (int i = 1; i <= (nlat + 1) / 2; i++) {
            if (i == (nlat + 1) / 2 && nlat % 2 != 0) {
                synth_equator_c(&t, norm, csphase, cc, coef, nlong);
                plan_exec(plan, coef, 1);
                for (int k = 0; k < nlong; k++) G(i - 1, k) = coef[k];
            } else {
                double z = zero[i - 1];
                double u = sqrt((1.0 - z) * (1.0 + z));
                int i_s = nlat + 1 - i;
                synth_ring_c(&t, norm, csphase, cc, z, u, coef, coefs, nlong);
                plan_exec(plan, coef, 1);
                for (int k = 0; k < nlong; k++) G(i - 1, k) = coef[k];
                plan_exec(plan, coefs, 1);
                for (int k = 0; k < nlong; k++) G(i_s - 1, k) = coefs[k];
            }
        }
        free(coef); free(coefs);
        plan_destroy(plan);
    }
    if (extend == 1)
        for (int i = 0; i < nlat; i++) G(i, nlong_out - 1) = G(i, 0);
#undef G
    free(cc);
    free_tables(&t);
    return 0;
}

/* Accumulate one ring pair of the complex analysis into cc ([m][l][i]);
 * f1/f2 are the weighted DFTs of the north/south rings (length nlong,
 * negative orders at nlong - m).  SHExpandDHC.f95:585-664 /
 * SHExpandGLQC.f95:556-656. */
static void anal_ring_c(const tables *t, int norm, int phase, double z, double u, cplx *f1, cplx *f2,
                        int nlong, cplx *cc)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
    double pm2, pm1, p, pmm, rescalem;
    cplx ffc1[2], ffc2[2];
    pm2 = pm00(norm);
    CCF(0, 0, 0) += pm2 * (f1[0] + f2[0]);
    if (lc == 0) return;
    pm1 = FF(t, ff1, 1, 0) * z * pm2;
    CCF(0, 1, 0) += pm1 * (f1[0] - f2[0]);
    ffc1[0] = f1[0] - f2[0];
    ffc1[1] = f1[0] + f2[0];
    for (int l = 2; l <= lc; l++) {
        p = FF(t, ff1, l, 0) * z * pm1 - FF(t, ff2, l, 0) * pm2;
        pm2 = pm1; pm1 = p;
        CCF(0, l, 0) += p * ffc1[fsym(l, 0) > 0];
    }
    pmm = pmm_start_c(norm);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        rescalem = rescalem * u;
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        f1[m] *= rescalem;
        f1[nlong - m] *= rescalem * msign(m);
        f2[m] *= rescalem;
        f2[nlong - m] *= rescalem * msign(m);
        CCF(0, m, m) += pm2 * (f1[m] + f2[m]);
        CCF(1, m, m) += pm2 * (f1[nlong - m] + f2[nlong - m]);
        pm1 = z * FF(t, ff1, m + 1, m) * pm2;
        CCF(0, m + 1, m) += pm1 * (f1[m] - f2[m]);
        CCF(1, m + 1, m) += pm1 * (f1[nlong - m] - f2[nlong - m]);
        ffc1[0] = f1[m] - f2[m];
        ffc1[1] = f1[m] + f2[m];
        ffc2[0] = f1[nlong - m] - f2[nlong - m];
        ffc2[1] = f1[nlong - m] + f2[nlong - m];
        for (int l = m + 2; l <= lc; l++) {
            p = z * FF(t, ff1, l, m) * pm1 - FF(t, ff2, l, m) * pm2;
            pm2 = pm1; pm1 = p;
            int s = fsym(l, m) > 0;
            CCF(0, l, m) += p * ffc1[s];
            CCF(1, l, m) += p * ffc2[s];
        }
    }
    rescalem = rescalem * u;
    pmm = pmm_last(norm, phase, pmm, lc, sqr, rescalem);
    CCF(0, lc, lc) += pmm * (f1[lc] + f2[lc]);
    CCF(1, lc, lc) += pmm * (f1[nlong - lc] + f2[nlong - lc]) * msign(lc);
}

/* Equator of the complex analysis: SHExpandDHC.f95:469-550. */
static void anal_equator_c(const tables *t, int norm, int phase, cplx *f1, int nlong, cplx *cc)
{
    const double *sqr = t->sqr;
    int lc = t->lc;
    double pm2, p, pmm, rescalem;
    pm2 = pm00(norm);
    CCF(0, 0, 0) += pm2 * f1[0];
    if (lc == 0) return;
    for (int l = 2; l <= lc; l += 2) {
        p = -FF(t, ff2, l, 0) * pm2;
        pm2 = p;
        CCF(0, l, 0) += p * f1[0];
    }
    pmm = pmm_start_c(norm);
    rescalem = 1.0 / 1.0e-280;
    for (int m = 1; m <= lc - 1; m++) {
        pm2 = pmm_step(norm, phase, &pmm, m, sqr);
        f1[m] *= rescalem;
        f1[nlong - m] *= rescalem * msign(m);
        CCF(0, m, m) += pm2 * f1[m];
        CCF(1, m, m) += pm2 * f1[nlong - m];
        for (int l = m + 2; l <= lc; l += 2) {
            p = -FF(t, ff2, l, m) * pm2;
            pm2 = p;
            CCF(0, l, m) += p * f1[m];
            CCF(1, l, m) += p * f1[nlong - m];
        }
    }
    pmm = pmm_last(norm, phase, pmm, lc, sqr, rescalem);
    CCF(0, lc, lc) += pmm * f1[lc];
    CCF(1, lc, lc) += msign(lc) * pmm * f1[nlong - lc];
}

/* Normalisation epilogue of the complex analysis: SHExpandDHC.f95:681-723. */
static void anal_epilogue_c(int norm, int lc, const cplx *cc, cplx *cilm, int d1, int d2)
{
#define CI(i, l, m) cilm[((size_t)(i) * d1 + (l)) * d2 + (m)]
    for (int l = 0; l <= lc; l++)
        for (int m = 0; m <= l; m++)
            for (int i = 0; i < 2; i++) CI(i, l, m) = CCF(i, l, m);
    double pi = acos(-1.0), prod;
    switch (norm) {
    case 1:
        for (int l = 0; l <= lc; l++)
            for (int m = 0; m <= l; m++)
                for (int i = 0; i < 2; i++) CI(i, l, m) = CI(i, l, m) / (4 * pi);
        break;
    case 2:
        for (int l = 0; l <= lc; l++)
            for (int m = 0; m <= l; m++)
                for (int i = 0; i < 2; i++) CI(i, l, m) = CI(i, l, m) * (2 * l + 1) / (4 * pi);
        break;
    case 3:
        for (int l = 0; l <= lc; l++) {
            prod = 4 * pi / (double)(2 * l + 1);
            CI(0, l, 0) = CI(0, l, 0) / prod;
            for (int m = 1; m <= l - 1; m++) {
                prod = prod * (l + m) * (l - m + 1);
                for (int i = 0; i < 2; i++) CI(i, l, m) = CI(i, l, m) / prod;
            }
            if (l != 0)
                for (int i = 0; i < 2; i++) CI(i, l, l) = CI(i, l, l) / (prod * 2 * l);
        }
        break;
    default: break;
    }
#undef CI
}

/* src/SHExpandDHC.f95:1-725: grid[n][g1] complex; cilm complex [2][d1][d2]. */
int or_shexpanddhc(const double *grid_ri, int n, int g1, int norm, int sampling, int csphase,
                   int lmax_calc, double *cilm_ri, int d1, int d2)
{
    const cplx *grid = (const cplx *)grid_ri;
    cplx *cilm = (cplx *)cilm_ri;
    int lmax = n / 2 - 1, lc, nlong;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (sampling != 1 && sampling != 2) return 2;
    if (n % 2 != 0) return 2;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    nlong = (sampling == 1) ? n : 2 * n;
    if (g1 < nlong) return 1;
    if (csphase != -1 && csphase != 1) return 2;
    if (norm < 1 || norm > 4) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0;
    double *aj = (double *)malloc(sizeof(double) * (n + 1));
    or_dhaj(n, 0, aj);
    for (int j = 0; j < n; j++) aj[j] = aj[j] * sqrt(4.0 * pi);
    tables t;
    int st = build_tables(&t, lc, norm);
    cplx *cc = (cplx *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(cplx));
    if (st || !cc) { free_tables(&t); free(aj); free(cc); return 3; }
    fftplan *plan = plan_create(nlong);
    cplx *f1 = (cplx *)malloc(sizeof(cplx) * nlong);
    cplx *f2 = (cplx *)malloc(sizeof(cplx) * nlong);
    int i_eq = n / 2 + 1;
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
    {   /* equator first (SHExpandDHC.f95:469-550) */
        int i = i_eq;
        for (int k = 0; k < nlong; k++) f1[k] = GR(i - 1, k);
        plan_exec(plan, f1, 0);
        for (int k = 0; k < nlong; k++) f1[k] = f1[k] * sqrt(2 * pi) * aj[i - 1] / (double)nlong;
        anal_equator_c(&t, norm, csphase, f1, nlong, cc);
    }
    for (int i = 2; i <= i_eq - 1; i++) {
        double theta = (i - 1) * pi / (double)n;
        double z = cos(theta);
        double u = sqrt((1.0 - z) * (1.0 + z));
        int i_s = 2 * i_eq - i;
        for (int k = 0; k < nlong; k++) f1[k] = GR(i - 1, k);
        plan_exec(plan, f1, 0);
        for (int k = 0; k < nlong; k++) f1[k] = f1[k] * sqrt(2 * pi) * aj[i - 1] / (double)nlong;
        for (int k = 0; k < nlong; k++) f2[k] = GR(i_s - 1, k);
        plan_exec(plan, f2, 0);
        for (int k = 0; k < nlong; k++) f2[k] = f2[k] * sqrt(2 * pi) * aj[i_s - 1] / (double)nlong;
        anal_ring_c(&t, norm, csphase, z, u, f1, f2, nlong, cc);
    }
#undef GR
    anal_epilogue_c(norm, lc, cc, cilm, d1, d2);
    free(f1); free(f2); free(aj); free(cc);
    plan_destroy(plan);
    free_tables(&t);
    return 0;
}

/* src/SHExpandGLQC.f95:1-708, on-the-fly branch (:450-660). */
int or_shexpandglqc(const double *grid_ri, int g0, int g1, int lmax, const double *w, int nw,
                    const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                    double *cilm_ri, int d1, int d2)
{
    const cplx *grid = (const cplx *)grid_ri;
    cplx *cilm = (cplx *)cilm_ri;
    int lc, nlong = 2 * lmax + 1, nlat = lmax + 1;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    if (g1 < 2 * lmax + 1 || g0 < lmax + 1) return 1;
    if (nw < lmax + 1) return 1;
    if (!zero) return 5;
    if (nzero < lmax + 1) return 1;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0;
    tables t;
    int st = build_tables(&t, lc, norm);
    cplx *cc = (cplx *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(cplx));
    if (st || !cc) { free_tables(&t); free(cc); return 3; }
    fftplan *plan = plan_create(nlong);
    cplx *f1 = (cplx *)malloc(sizeof(cplx) * nlong);
    cplx *f2 = (cplx *)malloc(sizeof(cplx) * nlong);
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
    for (int i = 1; i <= (nlat + 1) / 2; i++) {
        for (int k = 0; k < nlong; k++) f1[k] = GR(i - 1, k);
        plan_exec(plan, f1, 0);
        for (int k = 0; k < nlong; k++) f1[k] = 2 * w[i - 1] * pi * f1[k] / (double)nlong;
        if (i == (nlat + 1) / 2 && nlat % 2 != 0) {
            anal_equator_c(&t, norm, csphase, f1, nlong, cc);
        } else {
            double z = zero[i - 1];
            double u = sqrt((1.0 - z) * (1.0 + z));
            int i_s = nlat + 1 - i;
            for (int k = 0; k < nlong; k++) f2[k] = GR(i_s - 1, k);
            plan_exec(plan, f2, 0);
            for (int k = 0; k < nlong; k++) f2[k] = 2 * w[i_s - 1] * pi * f2[k] / (double)nlong;
            anal_ring_c(&t, norm, csphase, z, u, f1, f2, nlong, cc);
        }
    }
#undef GR
    anal_epilogue_c(norm, lc, cc, cilm, d1, d2);
    free(f1); free(f2); free(cc);
    plan_destroy(plan);
    free_tables(&t);
    return 0;
}

/* MakeGridGLQC with plx (src/MakeGridGLQC.f95:411-421, :443-468): every row
 * from the table (of complex normalisation, SHGLQ's cnorm = 1), degrees
 * summed from lmax_comp down; the order -m takes cilm(2,l,m) (-1)^m. */
int or_makegridglqc_plx(const double *cilm_ri, int d1, int d2, int lmax, const double *plx, int norm, int csphase,
                        int lmax_calc, int extend, double *grid_ri, int g0, int g1)
{
    const cplx *cilm = (const cplx *)cilm_ri;
    cplx *grid = (cplx *)grid_ri;
    int nlong = 2 * lmax + 1, nlat = lmax + 1, nlong_out;
    if (extend == 0) nlong_out = nlong; else if (extend == 1) nlong_out = nlong + 1; else return 2;
    if (g1 < nlong_out || g0 < nlat) return 1;
    if (!plx) return 5;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    if (lmax_calc > lmax) return 2;
    int lc = lc_of_makegrid(d1, d2, lmax, lmax_calc);
#define G(i, k) grid[(size_t)(i) * g1 + (k)]
    if (lc == 0) {
        double pm2 = pm00(norm);
        for (int i = 0; i < nlat; i++)
            for (int k = 0; k < nlong_out; k++) G(i, k) = cilm[0] * pm2;
        return 0;
    }
    cplx *cc = to_fortran_order_c(cilm, d1, d2, lc);
    if (!cc) return 3;
#pragma omp parallel
    {
        fftplan *plan = plan_create(nlong);
        cplx *coef = (cplx *)malloc(sizeof(cplx) * nlong);
#pragma omp for schedule(dynamic, 1)
        for (int i = 1; i <= nlat; i++) {
            for (int k = 0; k < nlong; k++) coef[k] = 0;
            for (int l = lc; l >= 0; l--) {
                size_t k = (size_t)l * (l + 1) / 2;   /* m = 0 */
                coef[0] = coef[0] + CCF(0, l, 0) * plx[k * nlat + (i - 1)];
                for (int m = 1; m <= l; m++) {
                    k = (size_t)l * (l + 1) / 2 + m;
                    const double p = plx[k * nlat + (i - 1)];
                    coef[m] = coef[m] + CCF(0, l, m) * p;
                    coef[nlong - m] = coef[nlong - m] + CCF(1, l, m) * p * msign(m);
                }
            }
            plan_exec(plan, coef, 1);
            for (int k = 0; k < nlong; k++) G(i - 1, k) = coef[k];
        }
        free(coef);
        plan_destroy(plan);
    }
    if (extend == 1)
        for (int i = 0; i < nlat; i++) G(i, nlong_out - 1) = G(i, 0);
#undef G
    free(cc);
    return 0;
}

/* SHExpandGLQC with plx (src/SHExpandGLQC.f95 plx branch and epilogue):
 * every ring's weighted DFT times the table, rings in order. */
int or_shexpandglqc_plx(const double *grid_ri, int g0, int g1, int lmax, const double *w, int nw,
                        const double *plx, int norm, int csphase, int lmax_calc, double *cilm_ri, int d1, int d2)
{
    const cplx *grid = (const cplx *)grid_ri;
    cplx *cilm = (cplx *)cilm_ri;
    int lc, nlong = 2 * lmax + 1, nlat = lmax + 1;
    if (lmax_calc > lmax) return 2;
    lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (d1 < lc + 1 || d2 < lc + 1) return 1;
    if (g1 < 2 * lmax + 1 || g0 < lmax + 1) return 1;
    if (nw < lmax + 1) return 1;
    if (!plx) return 5;
    if (norm < 1 || norm > 4) return 2;
    if (csphase != -1 && csphase != 1) return 2;
    double pi = acos(-1.0);
    for (size_t k = 0; k < (size_t)2 * d1 * d2; k++) cilm[k] = 0;
    cplx *cc = (cplx *)calloc((size_t)2 * (lc + 1) * (lc + 1), sizeof(cplx));
    if (!cc) return 3;
    fftplan *plan = plan_create(nlong);
    cplx *f1 = (cplx *)malloc(sizeof(cplx) * nlong);
#define GR(i, k) grid[(size_t)(i) * g1 + (k)]
    for (int i = 1; i <= nlat; i++) {
        for (int k = 0; k < nlong; k++) f1[k] = GR(i - 1, k);
        plan_exec(plan, f1, 0);
        for (int k = 0; k < nlong; k++) f1[k] = 2 * f1[k] * w[i - 1] * pi / (double)nlong;
        size_t k = 0;
        for (int l = 0; l <= lc; l++) {
            CCF(0, l, 0) = CCF(0, l, 0) + plx[k * nlat + (i - 1)] * f1[0];
            k++;
            for (int m = 1; m <= l; m++, k++) {
                const double p = plx[k * nlat + (i - 1)];
                CCF(0, l, m) = CCF(0, l, m) + p * f1[m];
                CCF(1, l, m) = CCF(1, l, m) + msign(m) * p * f1[nlong - m];
            }
        }
    }
#undef GR
    anal_epilogue_c(norm, lc, cc, cilm, d1, d2);
    free(f1); free(cc);
    plan_destroy(plan);
    return 0;
}

/* ------------------------------------------------------ gradient */

/* Horizontal gradient of a 4pi-normalised real field (no Condon-Shortley
 * phase) on a DH grid: theta = (1/r) dF/dtheta, phi = (1/(r sin theta))
 * dF/dphi, with dP_lm/dz evaluated term by term as
 * (sqrt((2l+1)(l^2-m^2)/(2l-1)) P_l-1,m - l z P_lm) / u^2.
 * src/MakeGradientDH.f95:1-658. */
int or_makegradientdh(const double *cilm, int d1, int d2, int lmax, int sampling, int lmax_calc,
                      int extend, double radius, double *theta_g, double *phi_g, int g0, int g1)
{
    int n = 2 * lmax + 2, nlong, nlat_out, nlong_out;
    if (sampling == 1) nlong = n; else if (sampling == 2) nlong = 2 * n; else return 2;
    if (extend == 0) { nlat_out = n; nlong_out = nlong; }
    else if (extend == 1) { nlat_out = n + 1; nlong_out = nlong + 1; }
    else return 2;
    if (g0 < nlat_out || g1 < nlong_out) return 1;
    if (lmax_calc > lmax) return 2;
    int lc = lc_of_makegrid(d1, d2, lmax, lmax_calc);
    double pi = acos(-1.0), scalef = 1.0e-280, r = radius;
#define TG(i, k) theta_g[(size_t)(i) * g1 + (k)]
#define PG(i, k) phi_g[(size_t)(i) * g1 + (k)]
    tables t;
    int st = build_tables(&t, lc, 1);
    double *cf = to_fortran_order(cilm, d1, d2, lc);
    if (st || !cf) { free_tables(&t); free(cf); return 3; }
    const double *sqr = t.sqr;
    int ncl = nlong / 2 + 1, i_eq = n / 2 + 1;
    fftplan *plan = plan_create(nlong);
    cplx *coef = (cplx *)malloc(sizeof(cplx) * ncl);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    double *row = (double *)malloc(sizeof(double) * nlong);
    cplx *ct = (cplx *)malloc(sizeof(cplx) * (lmax + 2)), *cts = (cplx *)malloc(sizeof(cplx) * (lmax + 2));
    cplx *cp = (cplx *)malloc(sizeof(cplx) * (lmax + 2)), *cps = (cplx *)malloc(sizeof(cplx) * (lmax + 2));
    /* c2r of (c0, c[1..lmax]/2) with zeros above lmax */
#define EMIT(c0v, carr, dest, i, fac, divsin)                                            \
    do {                                                                                 \
        for (int k = 0; k < ncl; k++) coef[k] = 0;                                       \
        coef[0] = (c0v);                                                                 \
        for (int k = 1; k <= lmax && k < ncl; k++) coef[k] = (carr)[k] / 2.0;            \
        c2r(plan, coef, buf, row);                                                       \
        for (int k = 0; k < nlong; k++)                                                  \
            dest(i, k) = (divsin) ? row[k] / r / (fac) : -(fac) * row[k] / r;            \
    } while (0)
    double pm2, pm1, p, dpl, pmm, rescalem, tempr, ct0, cts0, cp0, cps0;
    cplx tempc;
    /* equator: MakeGradientDH.f95:322-430 */
    {
        double theta = pi / 2.0, z = 0.0;
        for (int k = 0; k < lmax + 2; k++) { ct[k] = 0; cp[k] = 0; }
        ct0 = 0.0; cp0 = 0.0;
        pm2 = 1.0;
        pm1 = 0.0;
        if (lc != 0) {
            pm1 = 0.0;
            dpl = FF(&t, ff1, 1, 0);
            ct0 = ct0 + CF(0, 1, 0) * dpl;
        }
        for (int l = 2; l <= lc; l++) {
            p = -FF(&t, ff2, l, 0) * pm2;
            dpl = l * (sqr[2 * l + 1] / sqr[2 * l - 1] * pm1);
            ct0 = ct0 + CF(0, l, 0) * dpl;
            pm2 = pm1;
            pm1 = p;
        }
        pmm = sqr[2] * scalef;
        rescalem = 1.0 / scalef;
        for (int m = 1; m <= lc - 1; m++) {
            pmm = pmm * sqr[2 * m + 1] / sqr[2 * m];
            pm2 = pmm;
            cp[m] += (CF(1, m, m) + I * CF(0, m, m)) * pm2 * m;
            pm1 = 0.0;
            dpl = sqr[2 * m + 3] * pmm;
            ct[m] += (CF(0, m + 1, m) - I * CF(1, m + 1, m)) * dpl;
            for (int l = m + 2; l <= lc; l++) {
                p = -FF(&t, ff2, l, m) * pm2;
                cp[m] += (CF(1, l, m) + I * CF(0, l, m)) * p * m;
                dpl = sqr[2 * l + 1] * sqr[l - m] * sqr[l + m] / sqr[2 * l - 1] * pm1;
                ct[m] += (CF(0, l, m) - I * CF(1, l, m)) * dpl;
                pm2 = pm1;
                pm1 = p;
            }
            cp[m] *= rescalem;
            ct[m] *= rescalem;
        }
        if (lc != 0) {
            pmm = pmm * sqr[2 * lc + 1] / sqr[2 * lc] * rescalem;
            cp[lc] += (CF(1, lc, lc) + I * CF(0, lc, lc)) * pmm * lc;
            dpl = -lc * z * pmm / 1.0;
            ct[lc] += (CF(0, lc, lc) - I * CF(1, lc, lc)) * dpl;
        }
        EMIT(ct0, ct, TG, i_eq - 1, sin(theta), 0);
        EMIT(cp0, cp, PG, i_eq - 1, sin(theta), 1);
    }
    /* ring pairs: MakeGradientDH.f95:432-640 */
    for (int i = 1; i <= i_eq - 1; i++) {
        int i_s = 2 * i_eq - i;
        double theta = pi * (double)(i - 1) / (double)n;
        double z = cos(theta);
        double u = sqrt((1.0 - z) * (1.0 + z));
        double u2 = u * u;
        for (int k = 0; k < lmax + 2; k++) { ct[k] = 0; cts[k] = 0; cp[k] = 0; cps[k] = 0; }
        ct0 = cts0 = cp0 = cps0 = 0.0;
        pm2 = 1.0;
        pm1 = 0.0;
        if (lc != 0) {
            pm1 = FF(&t, ff1, 1, 0) * z;
            dpl = FF(&t, ff1, 1, 0);
            tempr = CF(0, 1, 0) * dpl;
            ct0 = ct0 + tempr;
            cts0 = cts0 + tempr;
        }
        for (int l = 2; l <= lc; l++) {
            p = FF(&t, ff1, l, 0) * z * pm1 - FF(&t, ff2, l, 0) * pm2;
            dpl = l * (sqr[2 * l + 1] / sqr[2 * l - 1] * pm1 - z * p) / u2;
            tempr = CF(0, l, 0) * dpl;
            ct0 = ct0 + tempr;
            cts0 = cts0 - tempr * fsym(l, 0);
            pm2 = pm1;
            pm1 = p;
        }
        pmm = sqr[2] * scalef;
        rescalem = 1.0 / scalef;
        for (int m = 1; m <= lc - 1; m++) {
            rescalem = rescalem * u;
            pmm = pmm * sqr[2 * m + 1] / sqr[2 * m];
            pm2 = pmm;
            tempc = (CF(1, m, m) + I * CF(0, m, m)) * pm2 * m;
            cp[m] += tempc; cps[m] += tempc;
            dpl = -m * z * pm2 / u2;
            tempc = (CF(0, m, m) - I * CF(1, m, m)) * dpl;
            ct[m] += tempc; cts[m] -= tempc;
            pm1 = z * FF(&t, ff1, m + 1, m) * pm2;
            tempc = (CF(1, m + 1, m) + I * CF(0, m + 1, m)) * pm1 * m;
            cp[m] += tempc; cps[m] -= tempc;
            dpl = (sqr[2 * m + 3] * pmm - z * (m + 1) * pm1) / u2;
            tempc = (CF(0, m + 1, m) - I * CF(1, m + 1, m)) * dpl;
            ct[m] += tempc; cts[m] += tempc;
            for (int l = m + 2; l <= lc; l++) {
                p = z * FF(&t, ff1, l, m) * pm1 - FF(&t, ff2, l, m) * pm2;
                tempc = (CF(1, l, m) + I * CF(0, l, m)) * p * m;
                cp[m] += tempc; cps[m] += tempc * fsym(l, m);
                dpl = (sqr[2 * l + 1] * sqr[l - m] * sqr[l + m] / sqr[2 * l - 1] * pm1 - z * l * p) / u2;
                tempc = (CF(0, l, m) - I * CF(1, l, m)) * dpl;
                ct[m] += tempc; cts[m] -= tempc * fsym(l, m);
                pm2 = pm1;
                pm1 = p;
            }
            ct[m] *= rescalem; cts[m] *= rescalem;
            cp[m] *= rescalem; cps[m] *= rescalem;
        }
        if (lc != 0) {
            rescalem = rescalem * u;
            pmm = pmm * sqr[2 * lc + 1] / sqr[2 * lc] * rescalem;
            tempc = (CF(1, lc, lc) + I * CF(0, lc, lc)) * pmm * lc;
            cp[lc] += tempc; cps[lc] += tempc;
            dpl = -lc * z * pmm / u2;
            tempc = (CF(0, lc, lc) - I * CF(1, lc, lc)) * dpl;
            ct[lc] += tempc; cts[lc] -= tempc;
        }
        if (i == 1) {
            for (int k = 0; k < nlong; k++) { TG(0, k) = 0.0; PG(0, k) = 0.0; }
        } else {
            EMIT(ct0, ct, TG, i - 1, sin(theta), 0);
            EMIT(cp0, cp, PG, i - 1, sin(theta), 1);
        }
        if (!(i == 1 && extend == 0)) {
            if (i == 1) {
                for (int k = 0; k < nlong; k++) { TG(i_s - 1, k) = 0.0; PG(i_s - 1, k) = 0.0; }
            } else {
                EMIT(cts0, cts, TG, i_s - 1, sin(theta), 0);
                EMIT(cps0, cps, PG, i_s - 1, sin(theta), 1);
            }
        }
    }
#undef EMIT
    if (extend == 1)
        for (int i = 0; i < nlat_out; i++) {
            TG(i, nlong_out - 1) = TG(i, 0);
            PG(i, nlong_out - 1) = PG(i, 0);
        }
#undef TG
#undef PG
    free(coef); free(buf); free(row); free(ct); free(cts); free(cp); free(cps); free(cf);
    plan_destroy(plan);
    free_tables(&t);
    return 0;
}

/* ------------------------------------------- gravity on an ellipsoid */

/* Normal gravity of a rotating ellipsoid (Somigliana's closed form) at
 * geocentric latitude lat (degrees).  src/NormalGravity.f95:1-139. */
double or_normalgravity(double geocentric_lat, double gm, double omega, double a, double b)
{
    double pi = acos(-1.0);
    if (a == b && omega == 0.0) return gm / (a * a);
    if (a == b) {
        double cl = cos(geocentric_lat * pi / 180.0);
        double g2 = gm * gm / (a * a * a * a) + a * omega * omega * cl * cl * (a * omega * omega - 2.0 * gm / (a * a));
        return sqrt(g2);
    }
    double m = omega * omega * a * a * b / gm;
    double bigE = sqrt(a * a - b * b), ep = bigE / b;
    double q0 = 0.5 * ((1.0 + 3.0 * (b / bigE) * (b / bigE)) * atan(bigE / b) - 3.0 * b / bigE);
    double q0p = 3.0 * (1.0 + (b / bigE) * (b / bigE)) * (1.0 - b / bigE * atan(bigE / b)) - 1.0;
    double ga = gm / (a * b) * (1.0 - m - m * ep * q0p / 6.0 / q0);
    double gb = gm / (a * a) * (1.0 + m * ep * q0p / 3.0 / q0);
    double gl = atan((a / b) * (a / b) * tan(geocentric_lat * pi / 180.0));
    double c = cos(gl), sn = sin(gl);
    double g = a * ga * c * c + b * gb * sn * sn;
    return g / sqrt(a * a * c * c + b * b * sn * sn);
}

/* Gravity vector (radial, theta, phi), its magnitude and the potential on a
 * flattened ellipsoid (semi-major axis a, flattening f) from 4pi-normalised
 * potential coefficients without the Condon-Shortley phase:
 * V = GM/r sum (r0/r)^l C_lm Y_lm, B = grad V, plus the centrifugal terms
 * when omega != 0 and minus the normal gravity from the magnitude when
 * normal_gravity = 1.  Restates src/MakeGravGridDH.f95:1-1069 ring by ring
 * (equator :430-605, ring pairs :607-1000, centrifugal :1002-1024, magnitude
 * and normal gravity :1026-1052, extend :1056-1066). */
int or_makegravgriddh(const double *cilm, int d1, int d2, int lmax, double gm, double r0, double a, double f,
                      int sampling, int lmax_calc, double omega, int normal_gravity, double normal_gravity_gm,
                      int extend, double *rad_g, double *theta_g, double *phi_g, double *total_g, double *pot_g,
                      int g0, int g1)
{
    int n = 2 * lmax + 2, nlong, nlat_out, nlong_out;
    if (sampling == 1) nlong = n; else if (sampling == 2) nlong = 2 * n; else return 2;
    if (extend == 0) { nlat_out = n; nlong_out = nlong; }
    else if (extend == 1) { nlat_out = n + 1; nlong_out = nlong + 1; }
    else return 2;
    if (g0 < nlat_out || g1 < nlong_out) return 1;
    if (normal_gravity != 0 && normal_gravity != 1) return 2;
    if (lmax_calc > lmax) return 2;
    int lc = lc_of_makegrid(d1, d2, lmax, lmax_calc);
    double pi = acos(-1.0), scalef = 1.0e-280;
#define RG(i, k) rad_g[(size_t)(i) * g1 + (k)]
#define TG(i, k) theta_g[(size_t)(i) * g1 + (k)]
#define PG(i, k) phi_g[(size_t)(i) * g1 + (k)]
#define AG(i, k) total_g[(size_t)(i) * g1 + (k)]
#define UG(i, k) pot_g[(size_t)(i) * g1 + (k)]
    tables t;
    int st = build_tables(&t, lc, 1);
    double *cf = to_fortran_order(cilm, d1, d2, lc);
    if (st || !cf) { free_tables(&t); free(cf); return 3; }
    const double *sqr = t.sqr;
    int ncl = nlong / 2 + 1, i_eq = n / 2 + 1;
    fftplan *plan = plan_create(nlong);
    cplx *coef = (cplx *)malloc(sizeof(cplx) * ncl);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * nlong);
    double *row = (double *)malloc(sizeof(double) * nlong);
    double *pf = (double *)malloc(sizeof(double) * (lmax + 2));
    /* per ring: R, U, T, P sums of the north ring, then of the south ring */
    cplx *cs[8];
    for (int j = 0; j < 8; j++) cs[j] = (cplx *)malloc(sizeof(cplx) * (lmax + 2));
    cplx *cr = cs[0], *cu = cs[1], *ct = cs[2], *cp = cs[3], *crs = cs[4], *cus = cs[5], *cts = cs[6], *cps = cs[7];
    double c0[8];
    /* c2r of (c0, c[1..lmax]/2), zeros above lmax, into row */
#define SYN(c0v, carr)                                                                   \
    do {                                                                                 \
        for (int k = 0; k < ncl; k++) coef[k] = 0;                                       \
        coef[0] = (c0v);                                                                 \
        for (int k = 1; k <= lmax && k < ncl; k++) coef[k] = (carr)[k] / 2.0;            \
        c2r(plan, coef, buf, row);                                                       \
    } while (0)
    double pm2, pm1, p, dpl, pmm, rescalem, tempr, r_ex;
    cplx tempc, cmi, spc;
    /* equator: MakeGravGridDH.f95:430-605 */
    {
        double theta = pi / 2.0, z = 0.0, u = 1.0;
        r_ex = a;
        for (int j = 0; j < 8; j++) { c0[j] = 0.0; for (int k = 0; k < lmax + 2; k++) cs[j][k] = 0; }
        pm2 = 1.0;
        c0[0] -= CF(0, 0, 0) * pm2;
        c0[1] += CF(0, 0, 0) * pm2;
        pm1 = 0.0;
        if (lc != 0) {
            pf[1] = r0 / r_ex;
            for (int l = 2; l <= lc; l++) pf[l] = pf[l - 1] * r0 / r_ex;
            c0[1] += CF(0, 1, 0) * pm1 * pf[1];
            dpl = FF(&t, ff1, 1, 0);
            c0[2] += CF(0, 1, 0) * dpl * pf[1];
        }
        for (int l = 2; l <= lc; l++) {
            p = -FF(&t, ff2, l, 0) * pm2;
            c0[0] += CF(0, l, 0) * p * (-(l + 1)) * pf[l];
            c0[1] += CF(0, l, 0) * p * pf[l];
            dpl = l * (sqr[2 * l + 1] / sqr[2 * l - 1] * pm1);
            c0[2] += CF(0, l, 0) * dpl * pf[l];
            pm2 = pm1;
            pm1 = p;
        }
        pmm = sqr[2] * scalef;
        rescalem = 1.0 / scalef;
        for (int m = 1; m <= lc - 1; m++) {
            pmm = pmm * sqr[2 * m + 1] / sqr[2 * m];
            pm2 = pmm;
            cmi = CF(0, m, m) - I * CF(1, m, m);
            spc = CF(1, m, m) + I * CF(0, m, m);
            cr[m] += cmi * pm2 * (-m - 1) * pf[m];
            cu[m] += cmi * pm2 * pf[m];
            cp[m] += spc * pm2 * pf[m] * m;
            pm1 = 0.0;
            dpl = sqr[2 * m + 3] * pmm;
            ct[m] += (CF(0, m + 1, m) - I * CF(1, m + 1, m)) * dpl * pf[m + 1];
            for (int l = m + 2; l <= lc; l++) {
                p = -FF(&t, ff2, l, m) * pm2;
                cmi = CF(0, l, m) - I * CF(1, l, m);
                cr[m] += cmi * p * (-(l + 1)) * pf[l];
                cu[m] += cmi * p * pf[l];
                cp[m] += (CF(1, l, m) + I * CF(0, l, m)) * p * pf[l] * m;
                dpl = sqr[2 * l + 1] * sqr[l - m] * sqr[l + m] / sqr[2 * l - 1] * pm1;
                ct[m] += cmi * dpl * pf[l];
                pm2 = pm1;
                pm1 = p;
            }
            cr[m] *= rescalem; cu[m] *= rescalem; cp[m] *= rescalem; ct[m] *= rescalem;
        }
        if (lc != 0) {
            pmm = pmm * sqr[2 * lc + 1] / sqr[2 * lc] * rescalem;
            cmi = CF(0, lc, lc) - I * CF(1, lc, lc);
            cr[lc] += cmi * pmm * (-lc - 1) * pf[lc];
            cu[lc] += cmi * pmm * pf[lc];
            cp[lc] += (CF(1, lc, lc) + I * CF(0, lc, lc)) * pmm * pf[lc] * lc;
            dpl = -lc * z * pmm / (u * u);
            ct[lc] += cmi * dpl * pf[lc];
        }
        int i = i_eq - 1;
        SYN(c0[0], cr);
        for (int k = 0; k < nlong; k++) RG(i, k) = row[k] * gm / (r_ex * r_ex);
        SYN(c0[1], cu);
        for (int k = 0; k < nlong; k++) UG(i, k) = row[k] * gm / r_ex;
        SYN(c0[2], ct);
        for (int k = 0; k < nlong; k++) TG(i, k) = -sin(theta) * row[k] * gm / (r_ex * r_ex);
        SYN(c0[3], cp);
        for (int k = 0; k < nlong; k++) PG(i, k) = row[k] * (gm / (r_ex * r_ex)) / sin(theta);
    }
    /* ring pairs: MakeGravGridDH.f95:607-1000 */
    for (int i = 1; i <= i_eq - 1; i++) {
        int i_s = 2 * i_eq - i;
        double theta = pi * (double)(i - 1) / (double)n;
        double z = cos(theta);
        double u = sqrt((1.0 - z) * (1.0 + z));
        double u2 = u * u;
        double lat = pi / 2.0 - theta;
        if (i == 1) {
            r_ex = a * (1.0 - f);
        } else {
            r_ex = cos(lat) * cos(lat) + sin(lat) * sin(lat) / ((1.0 - f) * (1.0 - f));
            r_ex = a * sqrt(1.0 / r_ex);
        }
        for (int j = 0; j < 8; j++) { c0[j] = 0.0; for (int k = 0; k < lmax + 2; k++) cs[j][k] = 0; }
        pm2 = 1.0;
        tempr = -CF(0, 0, 0) * pm2;
        c0[0] += tempr; c0[4] += tempr;
        tempr = CF(0, 0, 0) * pm2;
        c0[1] += tempr; c0[5] += tempr;
        pm1 = 0.0;
        if (lc != 0) {
            pf[1] = r0 / r_ex;
            for (int l = 2; l <= lc; l++) pf[l] = pf[l - 1] * r0 / r_ex;
            pm1 = FF(&t, ff1, 1, 0) * z;
            tempr = CF(0, 1, 0) * pm1 * (-2) * pf[1];
            c0[0] += tempr; c0[4] -= tempr;
            tempr = CF(0, 1, 0) * pm1 * pf[1];
            c0[1] += tempr; c0[5] -= tempr;
            dpl = FF(&t, ff1, 1, 0);
            tempr = CF(0, 1, 0) * dpl * pf[1];
            c0[2] += tempr; c0[6] += tempr;
        }
        for (int l = 2; l <= lc; l++) {
            p = FF(&t, ff1, l, 0) * z * pm1 - FF(&t, ff2, l, 0) * pm2;
            tempr = CF(0, l, 0) * p * (-(l + 1)) * pf[l];
            c0[0] += tempr; c0[4] += tempr * fsym(l, 0);
            tempr = CF(0, l, 0) * p * pf[l];
            c0[1] += tempr; c0[5] += tempr * fsym(l, 0);
            dpl = l * (sqr[2 * l + 1] / sqr[2 * l - 1] * pm1 - z * p) / u2;
            tempr = CF(0, l, 0) * dpl * pf[l];
            c0[2] += tempr; c0[6] -= tempr * fsym(l, 0);
            pm2 = pm1;
            pm1 = p;
        }
        pmm = sqr[2] * scalef;
        rescalem = 1.0 / scalef;
        for (int m = 1; m <= lc - 1; m++) {
            rescalem = rescalem * u;
            pmm = pmm * sqr[2 * m + 1] / sqr[2 * m];
            pm2 = pmm;
            cmi = CF(0, m, m) - I * CF(1, m, m);
            spc = CF(1, m, m) + I * CF(0, m, m);
            tempc = cmi * pm2 * (-m - 1) * pf[m];
            cr[m] += tempc; crs[m] += tempc;
            tempc = cmi * pm2 * pf[m];
            cu[m] += tempc; cus[m] += tempc;
            tempc = spc * pm2 * pf[m] * m;
            cp[m] += tempc; cps[m] += tempc;
            dpl = -m * z * pm2 / u2;
            tempc = cmi * dpl * pf[m];
            ct[m] += tempc; cts[m] -= tempc;
            pm1 = z * FF(&t, ff1, m + 1, m) * pm2;
            cmi = CF(0, m + 1, m) - I * CF(1, m + 1, m);
            spc = CF(1, m + 1, m) + I * CF(0, m + 1, m);
            tempc = cmi * pm1 * (-m - 2) * pf[m + 1];
            cr[m] += tempc; crs[m] -= tempc;
            tempc = cmi * pm1 * pf[m + 1];
            cu[m] += tempc; cus[m] -= tempc;
            tempc = spc * pm1 * pf[m + 1] * m;
            cp[m] += tempc; cps[m] -= tempc;
            dpl = (sqr[2 * m + 3] * pmm - z * (m + 1) * pm1) / u2;
            tempc = cmi * dpl * pf[m + 1];
            ct[m] += tempc; cts[m] += tempc;
            for (int l = m + 2; l <= lc; l++) {
                p = z * FF(&t, ff1, l, m) * pm1 - FF(&t, ff2, l, m) * pm2;
                cmi = CF(0, l, m) - I * CF(1, l, m);
                tempc = cmi * p * (-(l + 1)) * pf[l];
                cr[m] += tempc; crs[m] += tempc * fsym(l, m);
                tempc = cmi * p * pf[l];
                cu[m] += tempc; cus[m] += tempc * fsym(l, m);
                tempc = (CF(1, l, m) + I * CF(0, l, m)) * p * pf[l] * m;
                cp[m] += tempc; cps[m] += tempc * fsym(l, m);
                dpl = (sqr[2 * l + 1] * sqr[l - m] * sqr[l + m] / sqr[2 * l - 1] * pm1 - z * l * p) / u2;
                tempc = cmi * dpl * pf[l];
                ct[m] += tempc; cts[m] -= tempc * fsym(l, m);
                pm2 = pm1;
                pm1 = p;
            }
            for (int j = 0; j < 8; j++) cs[j][m] *= rescalem;
        }
        if (lc != 0) {
            rescalem = rescalem * u;
            pmm = pmm * sqr[2 * lc + 1] / sqr[2 * lc] * rescalem;
            cmi = CF(0, lc, lc) - I * CF(1, lc, lc);
            tempc = cmi * pmm * (-lc - 1) * pf[lc];
            cr[lc] += tempc; crs[lc] += tempc;
            tempc = cmi * pmm * pf[lc];
            cu[lc] += tempc; cus[lc] += tempc;
            tempc = (CF(1, lc, lc) + I * CF(0, lc, lc)) * pmm * pf[lc] * lc;
            cp[lc] += tempc; cps[lc] += tempc;
            dpl = -lc * z * pmm / u2;
            tempc = cmi * dpl * pf[lc];
            ct[lc] += tempc; cts[lc] -= tempc;
        }
        double g2 = gm / (r_ex * r_ex);
        for (int side = 0; side < 2; side++) {
            if (side == 1 && i == 1 && extend == 0) break;
            int row_i = side ? i_s - 1 : i - 1, o = side ? 4 : 0;
            SYN(c0[o + 0], cs[o + 0]);
            for (int k = 0; k < nlong; k++) RG(row_i, k) = row[k] * g2;
            SYN(c0[o + 1], cs[o + 1]);
            for (int k = 0; k < nlong; k++) UG(row_i, k) = row[k] * gm / r_ex;
            if (i == 1) {
                for (int k = 0; k < nlong; k++) { TG(row_i, k) = 0.0; PG(row_i, k) = 0.0; }
            } else {
                SYN(c0[o + 2], cs[o + 2]);
                for (int k = 0; k < nlong; k++) TG(row_i, k) = -sin(theta) * row[k] * g2;
                SYN(c0[o + 3], cs[o + 3]);
                for (int k = 0; k < nlong; k++) PG(row_i, k) = row[k] * g2 / sin(theta);
            }
        }
    }
#undef SYN
    /* centrifugal terms: MakeGravGridDH.f95:1002-1024 */
    if (omega != 0.0) {
        for (int i = 0; i < nlat_out; i++) {
            double theta = pi * (double)i / (double)n, lat = pi / 2.0 - theta, tl = tan(lat);
            double rr = a * sqrt((1.0 + tl * tl) / (1.0 + tl * tl / ((1.0 - f) * (1.0 - f))));
            for (int k = 0; k < nlong; k++) {
                RG(i, k) += rr * (sin(theta) * omega) * (sin(theta) * omega);
                TG(i, k) += sin(theta) * cos(theta) * rr * omega * omega;
                UG(i, k) += 0.5 * (rr * sin(theta) * omega) * (rr * sin(theta) * omega);
            }
        }
    }
    for (int i = 0; i < nlat_out; i++)
        for (int k = 0; k < nlong; k++)
            AG(i, k) = sqrt(RG(i, k) * RG(i, k) + PG(i, k) * PG(i, k) + TG(i, k) * TG(i, k));
    if (normal_gravity == 1) {
        double b = a * (1.0 - f);
        for (int i = 0; i < nlat_out; i++) {
            double theta = pi * (double)i / (double)n, lat = (pi / 2.0 - theta) * 180.0 / pi;
            double ng = or_normalgravity(lat, normal_gravity_gm, omega, a, b);
            for (int k = 0; k < nlong; k++) AG(i, k) -= ng;
        }
    }
    if (extend == 1)
        for (int i = 0; i < nlat_out; i++) {
            RG(i, nlong_out - 1) = RG(i, 0);
            TG(i, nlong_out - 1) = TG(i, 0);
            PG(i, nlong_out - 1) = PG(i, 0);
            AG(i, nlong_out - 1) = AG(i, 0);
            UG(i, nlong_out - 1) = UG(i, 0);
        }
#undef RG
#undef TG
#undef PG
#undef AG
#undef UG
    for (int j = 0; j < 8; j++) free(cs[j]);
    free(coef); free(buf); free(row); free(pf); free(cf);
    plan_destroy(plan);
    free_tables(&t);
    return 0;
}

/* ------------------------------------------- products of fields */

/* Spherical harmonic coefficients (degree lmax1 + lmax2) of the pointwise
 * product of two real fields: both synthesised on the Gauss-Legendre grid of
 * degree lmax1 + lmax2, multiplied and expanded.  Restates
 * src/SHMultiply.f95:1-368: argument checks :61-180 (precomp :137-151),
 * nodes :182-231, the leading (2, lmax+1, lmax+1) blocks of the inputs; with
 * precomp = 0 (the only mode of the f2py interface, src/pyshtools.pyf:596-620)
 * the transforms recur the functions (:256-280, :330-350), with precomp = 1
 * they read SHGLQ's table of them at the nodes (:281-325). */
int or_shmultiply(double *out, int od1, int od2, const double *c1, int c1d1, int c1d2, int lmax1,
                  const double *c2, int c2d1, int c2d2, int lmax2, int precomp, int norm, int csphase)
{
    if (lmax1 < 0 || lmax2 < 0) return 2;
    if (c1d1 < lmax1 + 1 || c1d2 < lmax1 + 1) return 1;
    if (c2d1 < lmax2 + 1 || c2d2 < lmax2 + 1) return 1;
    int lmaxout = lmax1 + lmax2;
    if (od1 < lmaxout + 1 || od2 < lmaxout + 1) return 1;
    if (csphase != 1 && csphase != -1) return 2;
    if (precomp != 0 && precomp != 1) return 2;
    if (norm < 1 || norm > 4) return 2;
    int nlat = lmaxout + 1, nlong = 2 * lmaxout + 1;
    double *zero = (double *)malloc(sizeof(double) * nlat), *w = (double *)malloc(sizeof(double) * nlat);
    double *plx = NULL;
    double *s1 = (double *)calloc((size_t)2 * (lmax1 + 1) * (lmax1 + 1), sizeof(double));
    double *s2 = (double *)calloc((size_t)2 * (lmax2 + 1) * (lmax2 + 1), sizeof(double));
    double *g1 = (double *)malloc(sizeof(double) * (size_t)nlat * nlong);
    double *g2 = (double *)malloc(sizeof(double) * (size_t)nlat * nlong);
    int st = (zero && w && s1 && s2 && g1 && g2) ? 0 : 3;
    if (!st) st = or_preglq(nlat, zero, w);
    if (!st && precomp == 1) {
        plx = (double *)malloc(sizeof(double) * (size_t)nlat * ((size_t)nlat * (nlat + 1) / 2));
        st = plx ? or_plx_table(lmaxout, zero, nlat, norm, csphase, 0, plx) : 3;
    }
    if (!st) {
        for (int i = 0; i < 2; i++)
            for (int l = 0; l <= lmax1; l++)
                for (int m = 0; m <= lmax1; m++)
                    s1[((size_t)i * (lmax1 + 1) + l) * (lmax1 + 1) + m] = c1[((size_t)i * c1d1 + l) * c1d2 + m];
        for (int i = 0; i < 2; i++)
            for (int l = 0; l <= lmax2; l++)
                for (int m = 0; m <= lmax2; m++)
                    s2[((size_t)i * (lmax2 + 1) + l) * (lmax2 + 1) + m] = c2[((size_t)i * c2d1 + l) * c2d2 + m];
        st = plx ? or_makegridglq_plx(s1, lmax1 + 1, lmax1 + 1, lmaxout, plx, norm, csphase, lmaxout, 0, g1, nlat,
                                      nlong)
                 : or_makegridglq(s1, lmax1 + 1, lmax1 + 1, lmaxout, zero, nlat, norm, csphase, lmaxout, 0, g1, nlat,
                                  nlong);
    }
    if (!st)
        st = plx ? or_makegridglq_plx(s2, lmax2 + 1, lmax2 + 1, lmaxout, plx, norm, csphase, lmaxout, 0, g2, nlat,
                                      nlong)
                 : or_makegridglq(s2, lmax2 + 1, lmax2 + 1, lmaxout, zero, nlat, norm, csphase, lmaxout, 0, g2, nlat,
                                  nlong);
    if (!st) {
        for (size_t e = 0; e < (size_t)nlat * nlong; e++) g1[e] *= g2[e];
        st = plx ? or_shexpandglq_plx(g1, nlat, nlong, lmaxout, w, nlat, plx, norm, csphase, lmaxout, out, od1, od2)
                 : or_shexpandglq(g1, nlat, nlong, lmaxout, w, nlat, zero, nlat, norm, csphase, lmaxout, out, od1,
                                  od2);
    }
    free(zero); free(w); free(s1); free(s2); free(g1); free(g2); free(plx);
    return st;
}

/* ------------------------------------------- batched CPU round trips */

/* One MakeGridDH + SHExpandDH round trip per field, fields spread over the
 * OpenMP threads with each reference-style call single-threaded.  Inputs are
 * packed back to back: field f has cilm at cilm + coff[f] (shape
 * [2][L+1][L+1]), grid scratch at grid + goff[f] and output at out + coff[f].
 * Returns the first nonzero status. */
int or_roundtrip_dh_batch(int nfields, const int *lmax, const long long *coff, const long long *goff,
                          int norm, int sampling, int csphase, const double *cilm, double *grid,
                          double *out, int nthreads)
{
    int status = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(| : status)
#endif
    for (int f = 0; f < nfields; f++) {
        int L = lmax[f], n = 2 * L + 2, nlong = sampling * n;
        int s = or_makegriddh(cilm + coff[f], L + 1, L + 1, L, norm, sampling, csphase, L, 0,
                              grid + goff[f], n, nlong);
        if (!s) s = or_shexpanddh(grid + goff[f], n, nlong, norm, sampling, csphase, L,
                                  out + coff[f], L + 1, L + 1);
        status |= s;
    }
    (void)nthreads;
    return status;
}
