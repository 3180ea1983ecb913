/*
 * shtools_amd -- MI355X-native real spherical-harmonic transforms (FP64).
 *
 * C ABI of libshtools_amd.so.  Each transform entry point replaces one
 * SHTOOLS routine behind the pyshtools backend plugin API
 * (pyshtools/backends/backends.py:26-31 registers backend modules whose
 * MakeGridDH/SHExpandDH/MakeGridGLQ/SHExpandGLQ are injected into
 * pyshtools.expand by pyshtools/expand/__init__.py:91-110).
 *
 * Conventions (all entry points):
 *   - arrays are C-order (row-major) device pointers: cilm[2][d1][d2] with
 *     cilm[i][l][m] (i = 0: C_lm, i = 1: S_lm), grid[nlat][nlon_alloc] with
 *     latitude rings contiguous;
 *   - work is enqueued on `stream` (a hipStream_t, NULL = default stream) and
 *     the call returns before it completes;
 *   - the return value is the reference's exitstatus: 0 ok, 1 improper array
 *     dimensions, 2 improper bounds of an input variable, 3 allocation
 *     failure, 5 other (e.g. neither zero nor plx given, no GPU); and one
 *     of this library's own: 6, the call would have to build device tables
 *     or bind new arrays while its stream is captured into a HIP graph
 *     (make one eager call of the same shape on the same arrays first);
 *     7 and 8 come from shtg_makegrid_dh_noalloc only (see there);
 *   - norm 1 = 4pi (geodesy), 2 = Schmidt, 3 = unnormalised, 4 = orthonormal;
 *     csphase 1 omits and -1 applies the Condon-Shortley phase (-1)^m;
 *   - calls are thread-safe; host-side tables are cached per (norm, csphase,
 *     degree) and per grid.  A lock is held only while a call looks its
 *     tables up (or builds them); host threads calling on different streams
 *     enqueue their work concurrently.  Calls on one batched plan take the
 *     plan's own lock.
 */
#ifndef SHTOOLS_AMD_H
#define SHTOOLS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MakeGridDH: src/MakeGridDH.f95:1-2 (C ABI include/shtools.h:104-116;
 * f2py src/pyshtools.pyf:256-276).  Output grid has
 * 2(lmax+1)+extend rows and sampling*2(lmax+1)+extend columns; grid_d0/d1
 * are the allocated dimensions (must be at least that). */
int shtg_makegrid_dh(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int norm,
                     int sampling, int csphase, int lmax_calc, int extend,
                     double *grid, int grid_d0, int grid_d1, void *stream);

/* MakeGridDH without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Arguments
 * and results are those of shtg_makegrid_dh.  The intermediates live in a
 * 64 MiB device array of the library that each GPU holds until the process
 * ends (enough for lmax up to about 1290); a call that needs more returns 7.
 * The call builds no tables either: the degree, sampling and normalisation
 * must have been used by an ordinary call on this device first, else it
 * returns 8.  All calls on one GPU share that GPU's array: calls that
 * overlap in time on one GPU (several streams or host threads) race. */
int shtg_makegrid_dh_noalloc(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int norm,
                             int sampling, int csphase, int lmax_calc, int extend,
                             double *grid, int grid_d0, int grid_d1, void *stream);

/* SHExpandDH: src/SHExpandDH.f95:1-2 (include/shtools.h:92-102;
 * src/pyshtools.pyf:235-254).  n = grid_d0 must be even; the grid has at
 * least n (sampling 1) or 2n (sampling 2) columns.  cilm receives
 * (2, lmax_calc+1, lmax_calc+1) coefficients in a [2][cilm_d1][cilm_d2]
 * array that is fully overwritten (zero where m > l). */
int shtg_shexpand_dh(const double *grid, int grid_d0, int grid_d1, int norm, int sampling,
                     int csphase, int lmax_calc, double *cilm, int cilm_d1, int cilm_d2,
                     void *stream);

/* MakeGridGLQ, on-the-fly branch: src/MakeGridGLQ.f95:1-2,467-669
 * (include/shtools.h:161-172; src/pyshtools.pyf:357-377).  `zero` is a HOST
 * array of nzero >= lmax+1 Gauss-Legendre nodes (from shtg_shglq). */
int shtg_makegrid_glq(const double *cilm, int cilm_d1, int cilm_d2, int lmax,
                      const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                      int extend, double *grid, int grid_d0, int grid_d1, void *stream);

/* MakeGridGLQ with a table of Legendre functions (the plx branch):
 * src/MakeGridGLQ.f95:1-2,408-420,441-465 (include/shtools.h:161-172 with
 * plx present).  plx is a DEVICE array of (lmax+1) * (lmax+1)(lmax+2)/2
 * doubles laid out as the reference's Fortran array plx(lmax+1, k):
 * plx[k*(lmax+1) + i] = P_lm at node i, k = l(l+1)/2 + m, as
 * shtg_shglq_plx writes it.  The table decides normalisation and phase;
 * norm and csphase are checked as the reference checks them, and norm sets
 * the value of a degree-0 field.  plx == NULL is exit status 5 (neither
 * table nor nodes). */
int shtg_makegrid_glq_plx(const double *cilm, int cilm_d1, int cilm_d2, int lmax,
                          const double *plx, int norm, int csphase, int lmax_calc, int extend,
                          double *grid, int grid_d0, int grid_d1, void *stream);

/* SHExpandGLQ with a table of Legendre functions (the plx branch):
 * src/SHExpandGLQ.f95:1-2,419-441,659-689 (include/shtools.h:150-159 with
 * plx present).  lmax = grid_d0 - 1; `w` is a HOST array of >= lmax+1
 * weights; plx a DEVICE table as for shtg_makegrid_glq_plx; norm selects
 * the normalisation epilogue. */
int shtg_shexpand_glq_plx(const double *grid, int grid_d0, int grid_d1, const double *w, int nw,
                          const double *plx, int norm, int csphase, int lmax_calc, double *cilm,
                          int cilm_d1, int cilm_d2, void *stream);

/* MakeGridGLQC and SHExpandGLQC with a table of Legendre functions (their
 * plx branches: src/MakeGridGLQC.f95:411-421,443-468 and src/SHExpandGLQC.f95,
 * include/shtools.h:174-195 with plx present): as shtg_makegrid_glq_plx and
 * shtg_shexpand_glq_plx for complex coefficients and grids (interleaved
 * (re, im) doubles, the layouts of shtg_makegrid_glqc / shtg_shexpand_glqc),
 * with a table of the complex normalisation (shtg_shglq_plx, cnorm = 1). */
int shtg_makegrid_glqc_plx(const double *cilm, int cilm_d1, int cilm_d2, int lmax,
                           const double *plx, int norm, int csphase, int lmax_calc, int extend,
                           double *grid, int grid_d0, int grid_d1, void *stream);
int shtg_shexpand_glqc_plx(const double *grid, int grid_d0, int grid_d1, const double *w, int nw,
                           const double *plx, int norm, int csphase, int lmax_calc, double *cilm,
                           int cilm_d1, int cilm_d2, void *stream);

/* SHExpandGLQ, on-the-fly branch: src/SHExpandGLQ.f95:1-2,445-649
 * (include/shtools.h:150-159; src/pyshtools.pyf:334-355).  lmax =
 * grid_d0 - 1; `w` and `zero` are HOST arrays of >= lmax+1 weights and
 * nodes. */
int shtg_shexpand_glq(const double *grid, int grid_d0, int grid_d1, const double *w, int nw,
                      const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                      double *cilm, int cilm_d1, int cilm_d2, void *stream);

/* Complex transforms: MakeGridDHC src/MakeGridDHC.f95:1-2 (f2py
 * src/pyshtools.pyf:299-319), SHExpandDHC src/SHExpandDHC.f95:1-2
 * (src/pyshtools.pyf:278-297), MakeGridGLQC src/MakeGridGLQC.f95:1-2
 * (src/pyshtools.pyf:402-422), SHExpandGLQC src/SHExpandGLQC.f95:1-2
 * (src/pyshtools.pyf:379-400), on-the-fly branches.  Arguments as for the
 * real transforms; cilm is a complex [2][d1][d2] array ([0] orders +m, [1]
 * orders -m) and the grid a complex [d0][d1] array, both as interleaved
 * (re, im) doubles (the layout of numpy/torch complex128). */
int shtg_makegrid_dhc(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int norm,
                      int sampling, int csphase, int lmax_calc, int extend,
                      double *grid, int grid_d0, int grid_d1, void *stream);
int shtg_shexpand_dhc(const double *grid, int grid_d0, int grid_d1, int norm, int sampling,
                      int csphase, int lmax_calc, double *cilm, int cilm_d1, int cilm_d2,
                      void *stream);
int shtg_makegrid_glqc(const double *cilm, int cilm_d1, int cilm_d2, int lmax,
                       const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                       int extend, double *grid, int grid_d0, int grid_d1, void *stream);
int shtg_shexpand_glqc(const double *grid, int grid_d0, int grid_d1, const double *w, int nw,
                       const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                       double *cilm, int cilm_d1, int cilm_d2, void *stream);

/* MakeGradientDH: src/MakeGradientDH.f95:1-658 (f2py src/pyshtools.pyf:
 * 2021-2043).  Horizontal gradient of a 4pi-normalised real field without
 * the Condon-Shortley phase: theta = (1/r) dF/dtheta and phi =
 * (1/(r sin theta)) dF/dphi on DH grids of 2(lmax+1)+extend rows and
 * sampling*2(lmax+1)+extend columns (both [grid_d0][grid_d1]); the pole rows
 * are zero. */
int shtg_makegradient_dh(const double *cilm, int cilm_d1, int cilm_d2, int lmax, int sampling,
                         int lmax_calc, int extend, double radius, double *theta, double *phi,
                         int grid_d0, int grid_d1, void *stream);

/* MakeGravGridDH: src/MakeGravGridDH.f95:1-1069 (f2py src/pyshtools.pyf:
 * 1440-1477; Python wrapper src/PythonWrapper.f95:1690-1735, which always
 * passes omega and pot).  Gravity vector (rad positive upward, theta, phi),
 * its magnitude and the potential, in SI units, on a flattened ellipsoid
 * (semi-major axis a, flattening f) from 4pi-normalised potential
 * coefficients without the Condon-Shortley phase referenced to r0:
 * V = GM/r sum (r0/r)^l C_lm Y_lm.  omega adds the centrifugal terms;
 * normal_gravity = 1 removes the normal gravity (NormalGravity.f95) computed
 * with normal_gravity_gm from the magnitude.  All five grids are
 * [grid_d0][grid_d1] with 2(lmax+1)+extend rows and
 * sampling*2(lmax+1)+extend columns. */
int shtg_makegravgrid_dh(const double *cilm, int cilm_d1, int cilm_d2, int lmax, double gm, double r0,
                         double a, double f, int sampling, int lmax_calc, double omega,
                         int normal_gravity, double normal_gravity_gm, int extend, double *rad,
                         double *theta, double *phi, double *total, double *pot, int grid_d0,
                         int grid_d1, void *stream);

/* NormalGravity: src/NormalGravity.f95:1-139 (host).  Normal gravity of a
 * rotating ellipsoid (semi-axes a >= b, GM, omega) at geocentric latitude
 * geocentric_lat (degrees), by Somigliana's closed form; into *out. */
int shtg_normal_gravity(double geocentric_lat, double gm, double omega, double a, double b, double *out);

/* SHMultiply: src/SHMultiply.f95:1-368, include/shtools.h:254-265 (f2py
 * src/pyshtools.pyf:596-620 fixes precomp = 0).  Coefficients
 * [2][out_d1][out_d2] (at least lmax1+lmax2+1 square) of the pointwise
 * product of the fields of the leading (lmax1+1)^2 and (lmax2+1)^2 blocks of
 * cilm1 and cilm2: both synthesised on the Gauss-Legendre grid of degree
 * lmax1+lmax2, multiplied, and expanded.  precomp = 0 recurs the Legendre
 * functions in every pass; precomp = 1 first builds SHGLQ's table of them at
 * the nodes (as shtg_shglq_plx, (lmax1+lmax2+1)^2 (lmax1+lmax2+2)/2 doubles
 * of stream-ordered device memory for the call) and reads it in every pass
 * (src/SHMultiply.f95:281-325); other values are exit status 2, and
 * precomp = 1 under stream capture is status 6. */
int shtg_shmultiply(double *cilmout, int out_d1, int out_d2, const double *cilm1, int c1_d1, int c1_d2,
                    int lmax1, const double *cilm2, int c2_d1, int c2_d2, int lmax2, int precomp, int norm,
                    int csphase, void *stream);

/* The polar skip of the DH transforms (host; no counterpart in the
 * reference, which sums every term): kmax[s] for the ring-pair slots
 * s = 0 .. lmax+1 (slot s holds rows s and 2 lmax + 2 - s, colatitude
 * pi s / (2 lmax + 2)) is the largest order whose 4pi-normalised functions
 * reach 1e-32 at some degree <= lmax; orders above it are not summed there. */
int shtg_dh_polar_orders(int lmax, int *kmax);

/* Adjoints (transposes) of the four transforms, for gradients through them
 * (no counterpart in the reference; the differentiable path of the
 * transform API).  Each takes the argument list of its forward call with
 * the roles of input and output exchanged, and applies the exact transpose
 * of the linear map that call computes:
 *   makegrid adjoint:  cilm_bar = S^T grid_bar, S the synthesis
 *                      (MakeGridDH.f95:504-643 / MakeGridGLQ.f95:467-669),
 *                      including the extended south-pole row and repeated
 *                      column; entries of cilm_bar the synthesis does not
 *                      read (l > lmax_calc, m > l, S_l0) are zero;
 *   shexpand adjoint:  grid_bar = A^T cilm_bar, A the analysis
 *                      (SHExpandDH.f95:463-702 / SHExpandGLQ.f95:445-689);
 *                      grid entries the analysis does not read are zero.
 * The kernels recompute the Legendre functions as the forward ones do. */
int shtg_makegrid_dh_adjoint(const double *grid_bar, int grid_d0, int grid_d1, int lmax, int norm,
                             int sampling, int csphase, int lmax_calc, int extend,
                             double *cilm_bar, int cilm_d1, int cilm_d2, void *stream);
int shtg_shexpand_dh_adjoint(const double *cilm_bar, int cilm_d1, int cilm_d2, int norm, int sampling,
                             int csphase, int lmax_calc, double *grid_bar, int grid_d0, int grid_d1,
                             void *stream);
int shtg_makegrid_glq_adjoint(const double *grid_bar, int grid_d0, int grid_d1, int lmax,
                              const double *zero, int nzero, int norm, int csphase, int lmax_calc,
                              int extend, double *cilm_bar, int cilm_d1, int cilm_d2, void *stream);
int shtg_shexpand_glq_adjoint(const double *cilm_bar, int cilm_d1, int cilm_d2, const double *w,
                              int nw, const double *zero, int nzero, int norm, int csphase,
                              int lmax_calc, double *grid_bar, int grid_d0, int grid_d1,
                              void *stream);

/* Batched DH transforms of many fields of uneven degree (no counterpart in
 * the reference, which transforms one field per call; semantics per field
 * are those of MakeGridDH/SHExpandDH with lmax_calc = lmax, extend = 0).
 * A plan fixes the degrees, sampling, norm and csphase; each call enqueues a
 * handful of launches covering every field, with work items ordered longest
 * first.  cilm[i] points to a device [2][lmax[i]+1][lmax[i]+1] array and
 * grid[i] to a device [2lmax[i]+2][sampling(2lmax[i]+2)] array; the arrays of
 * pointers are host arrays (uploaded when they change).  Calls on one plan
 * are serialised on the host by the plan's lock, but their work must not
 * overlap in time on different streams (they share its workspace); plans
 * are independent of each other.  shtg_batch_destroy must not run while
 * another call on the plan does. */
typedef struct shtg_batch_plan shtg_batch_plan;
int shtg_batch_create_dh(int nfields, const int *lmax, int sampling, int norm, int csphase,
                         shtg_batch_plan **plan);
/* The same with a sampling (1 or 2) per field: grid[i] is then
 * [2lmax[i]+2][sampling[i](2lmax[i]+2)]. */
int shtg_batch_create_dh_fields(int nfields, const int *lmax, const int *sampling, int norm, int csphase,
                                shtg_batch_plan **plan);
/* Batched GLQ transforms: per field MakeGridGLQ / SHExpandGLQ
 * (src/MakeGridGLQ.f95:108-480, src/SHExpandGLQ.f95:100-649; f2py
 * src/pyshtools.pyf:334-377) with lmax_calc = lmax, extend = 0,
 * on the nodes and weights of shtg_shglq (SHGLQ), which the plan computes
 * itself: grid[i] is a device [lmax[i]+1][2lmax[i]+1] array.  The batched
 * calls, adjoints, tuning and field groups below apply unchanged. */
int shtg_batch_create_glq(int nfields, const int *lmax, int norm, int csphase, shtg_batch_plan **plan);
/* A plan mixing grids: grid[i] = 1 or 2 (DH with that sampling) or 3 (GLQ). */
int shtg_batch_create_fields(int nfields, const int *lmax, const int *grid, int norm, int csphase,
                             shtg_batch_plan **plan);
int shtg_batch_makegrid_dh(shtg_batch_plan *plan, const double *const *cilm, double *const *grid,
                           void *stream);
int shtg_batch_shexpand_dh(shtg_batch_plan *plan, const double *const *grid, double *const *cilm,
                           void *stream);
/* Adjoints of the batched transforms (per field as shtg_makegrid_dh_adjoint
 * and shtg_shexpand_dh_adjoint). */
int shtg_batch_makegrid_dh_adjoint(shtg_batch_plan *plan, const double *const *grid_bar,
                                   double *const *cilm_bar, void *stream);
int shtg_batch_shexpand_dh_adjoint(shtg_batch_plan *plan, const double *const *cilm_bar,
                                   double *const *grid_bar, void *stream);
int shtg_batch_destroy(shtg_batch_plan *plan);
/* Same-degree buckets on the matrix cores (no counterpart in the
 * reference): fields of one degree and sampling, in groups of up to 32, whose
 * group has at least min_bucket fields have their synthesis Legendre sums
 * computed as FP64 MFMA products (one recursion per ring pair and order
 * shared by the group); min_bucket = 0 turns this off (the default).  The
 * grids are bitwise those of the per-field path (the FP64 MFMA accumulates
 * its products as an in-order chain of fused multiply-adds, the order of
 * the per-field sums).  `stream` is unused (kept for symmetry). */
int shtg_batch_set_mfma(shtg_batch_plan *plan, int min_bucket, void *stream);
/* Launch-shape autotuner (no counterpart in the reference): times each
 * ring-FFT group of the plan at 256, 384, 512, 768 and 1024 threads per workgroup
 * (best of `reps` runs, synthesis and analysis separately), then the
 * Legendre launches: ring pairs per lane of the synthesis waves (2, 3, 4)
 * and the same-degree bucket size from which the synthesis goes to the MFMA
 * kernel (off, 2, 4, 8, 16, where the plan has such buckets), on the arrays
 * bound by the plan's last makegrid and shexpand calls, keeping the fastest
 * of each; none of these changes any result.  flags bit 0 also tunes ring
 * pairs per lane of the analysis waves (4, 6, 8, 12), which regroups the
 * analysis sums (coefficients then change by rounding).  Synchronises
 * `stream`; overwrites those grids and the plan's workspace, so rerun the
 * transforms afterwards.  Returns 5 if the plan has not been called in both
 * directions yet. */
int shtg_batch_tune(shtg_batch_plan *plan, int reps, int flags, void *stream);
/* The Legendre launch choices of a plan (from shtg_batch_tune, or set here
 * to reuse an earlier tuning): ring pairs per lane of the synthesis waves
 * (1 .. 4) and of the analysis waves (1 .. 12), and the MFMA bucket size
 * (0: off, else >= 2; see shtg_batch_set_mfma).  Results do not depend on
 * them beyond rounding (analysis sums are associated by pass). */
int shtg_batch_get_tuning(shtg_batch_plan *plan, int *synth_kmax, int *anal_kmax, int *mfma_min);
int shtg_batch_set_tuning(shtg_batch_plan *plan, int synth_kmax, int anal_kmax, int mfma_min);
/* Set the ring-FFT workgroup sizes (a multiple of 64 from 256 to 1024 per
 * group, at least what the group's plans need; n = the number of groups
 * shtg_batch_get_shapes returns). */
int shtg_batch_set_shapes(shtg_batch_plan *plan, const int *threads_synth, const int *threads_anal, int n);
/* Workgroup sizes of the plan's ring-FFT groups (up to n of them); returns
 * the number of groups. */
int shtg_batch_get_shapes(shtg_batch_plan *plan, int *threads_synth, int *threads_anal, int n);
/* The ring-FFT group of each field (group[0 .. nfields), n >= nfields):
 * fields of one group share a launch and an LDS size class; the last group,
 * numbered after the shapes of shtg_batch_get_shapes, holds the fields whose
 * rings run on the global workspace.  Returns the number of groups. */
int shtg_batch_field_groups(shtg_batch_plan *plan, int *group, int n);

/* Host helpers. */
/* SHGLQ nodes and weights (include/shtools.h:142-148; src/SHGLQ.f95:187-192 ->
 * src/PreGLQ.f95:93-149):
 * zero[lmax+1] descending from ~+1, w[lmax+1]. */
int shtg_shglq(int lmax, double *zero, double *w);
/* SHGLQ with its table plx (include/shtools.h:142-148 with plx present;
 * src/SHGLQ.f95:202-257): zero and w as shtg_shglq (HOST arrays), plx a
 * DEVICE array of (lmax+1) * (lmax+1)(lmax+2)/2 doubles receiving the
 * Legendre functions of normalisation norm and phase csphase at the nodes
 * (plx[k*(lmax+1) + i], k = l(l+1)/2 + m; cnorm = 1: the complex
 * normalisation, which omits the sqrt(2) of orders m >= 1; norm 3 has
 * none), computed on the stream.  Not under graph capture (status 6). */
int shtg_shglq_plx(int lmax, double *zero, double *w, double *plx, int norm, int csphase, int cnorm,
                   void *stream);
/* NGLQ, NGLQSH, NGLQSHN (include/shtools.h:1031-1033; src/PreGLQ.f95:153-233):
 * Gauss-Legendre points needed to integrate exactly a polynomial of degree
 * `degree` (ceiling((degree + 1) / 2)), a field of degree `degree`
 * (degree + 1), and that field raised to the power n (ceiling(((n + 1)
 * degree + 1) / 2)).  -1 for a negative degree (the reference stops). */
int shtg_nglq(int degree);
int shtg_nglqsh(int degree);
int shtg_nglqshn(int degree, int n);
/* DHaj (include/shtools.h:1034; src/DHaj.f95:91-105): aj[n+extend]
 * Driscoll-Healy weights. */
int shtg_dhaj(int n, int extend, double *aj);
/* GLQGridCoord (include/shtools.h:198; src/GLQGridCoord.f95:100-119):
 * lat[lmax+1], lon[2lmax+1+extend]. */
int shtg_glqgridcoord(int lmax, int extend, double *lat, double *lon);

/* Associated Legendre functions P_lm(z), l, m <= lmax, into the host array
 * p[(lmax+1)*(lmax+1)] (row l, column m), computed on the host with the same
 * tables, seeds, extended-range scaling and recursion as the GPU kernels;
 * norm/csphase as above (the reference's PlmBar, PlmSchmidt, PLegendreA and
 * PlmON: src/PlmBar.f95:153-288, include/shtools.h:20-90). */
int shtg_plm(int lmax, int norm, int csphase, double z, double *p);

/* Per-kernel timing with HIP events recorded on each launch stream (no
 * counterpart in the reference; used by the benchmark).  While enabled,
 * every launch is bracketed by two events; collect() waits for them and
 * returns per kernel id the summed milliseconds and launch counts, then
 * forgets them. */
enum {
    SHTG_K_PACK = 0,
    SHTG_K_SYNTH_LEG = 1,
    SHTG_K_SYNTH_FFT = 2,
    SHTG_K_ANAL_FFT = 3,
    SHTG_K_ANAL_LEG = 4,
    SHTG_K_UNPACK = 5,
    SHTG_K_COUNT = 16
};
int shtg_profile_enable(int on);
int shtg_profile_collect(double *ms, long long *count, int nk);

/* Library status. */
int shtg_device_count(void);
const char *shtg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SHTOOLS_AMD_H */
