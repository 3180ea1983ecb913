// Argument checks of the plain transforms: the reference's checks of
// dimensions and options, returning the library's status codes, and the
// shapes they fix.  Plain C++ (no HIP runtime calls), so that it is also
// built and checked on the host alone, under AddressSanitizer.
#pragma once

#include <climits>

namespace shtg {

// Status codes of the C interface (include/shtools_amd.h).
enum { ST_OK = 0, ST_DIMS = 1, ST_BOUNDS = 2, ST_ALLOC = 3, ST_OTHER = 5, ST_CAPTURE = 6, ST_SCRATCH = 7,
       ST_NOTABLES = 8 };

// Largest degree whose grids have int dimensions: a DH grid of degree L
// has up to 4L + 5 columns (sampling 2, extend).
constexpr int kMaxArgDegree = (INT_MAX - 8) / 4;

// Shapes of one plain transform after the reference's argument checks.
struct Shape {
    int lmax = 0;      // grid degree
    int Lc = 0;        // degrees computed
    int nrows = 0;     // grid rows written by synthesis (with extend)
    int ext = 0;       // extend
};

// src/MakeGridDH.f95:107-251
inline int check_makegrid_dh(int cilm_d1, int cilm_d2, int lmax, int norm, int sampling, int csphase,
                             int lmax_calc, int extend, int grid_d0, int grid_d1, Shape &sh)
{
    if (lmax < 0 || lmax > kMaxArgDegree) return ST_BOUNDS;
    const int n = 2 * lmax + 2;
    int nlong;
    if (sampling == 1) nlong = n;
    else if (sampling == 2) nlong = 2 * n;
    else return ST_BOUNDS;
    if (extend != 0 && extend != 1) return ST_BOUNDS;
    const int nlat_out = n + extend, nlong_out = nlong + extend;
    if (grid_d0 < nlat_out || grid_d1 < nlong_out) return ST_DIMS;
    if (norm < 1 || norm > 4) return ST_BOUNDS;
    if (csphase != 1 && csphase != -1) return ST_BOUNDS;
    if (lmax_calc > lmax) return ST_BOUNDS;
    if (lmax_calc < 0 || cilm_d1 < 1 || cilm_d2 < 1) return ST_BOUNDS;
    int Lc = lmax;
    if (cilm_d2 - 1 < Lc) Lc = cilm_d2 - 1;
    if (cilm_d1 - 1 < Lc) Lc = cilm_d1 - 1;
    if (lmax_calc < Lc) Lc = lmax_calc;
    sh.lmax = lmax;
    sh.Lc = Lc;
    sh.nrows = nlat_out;
    sh.ext = extend;
    return ST_OK;
}

// src/SHExpandDH.f95:118-270
inline int check_shexpand_dh(int grid_d0, int grid_d1, int norm, int sampling, int csphase, int lmax_calc,
                             int cilm_d1, int cilm_d2, Shape &sh)
{
    const int n = grid_d0;
    if (n > 2 * kMaxArgDegree + 2) return ST_BOUNDS;
    const int lmax = n / 2 - 1;
    if (lmax_calc > lmax) return ST_BOUNDS;
    if (lmax_calc < 0) return ST_BOUNDS;
    const int Lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (sampling != 1 && sampling != 2) return ST_BOUNDS;
    if (n % 2 != 0 || n < 2) return ST_BOUNDS;
    if (cilm_d1 < Lc + 1 || cilm_d2 < Lc + 1) return ST_DIMS;
    const int nlong = sampling * n;
    if (grid_d1 < nlong) return ST_DIMS;
    if (csphase != 1 && csphase != -1) return ST_BOUNDS;
    if (norm < 1 || norm > 4) return ST_BOUNDS;
    sh.lmax = lmax;
    sh.Lc = Lc;
    sh.nrows = n;
    return ST_OK;
}

// src/MakeGridGLQ.f95:108-270
inline int check_makegrid_glq(int cilm_d1, int cilm_d2, int lmax, int nzero, const double *zero, int norm,
                              int csphase, int lmax_calc, int extend, int grid_d0, int grid_d1, Shape &sh)
{
    if (lmax < 0 || lmax > kMaxArgDegree) return ST_BOUNDS;
    const int nlong = 2 * lmax + 1, nlat = lmax + 1;
    if (extend != 0 && extend != 1) return ST_BOUNDS;
    const int nlong_out = nlong + extend;
    if (grid_d1 < nlong_out || grid_d0 < nlat) return ST_DIMS;
    if (!zero) return ST_OTHER;
    if (nzero < lmax + 1) return ST_DIMS;
    if (norm < 1 || norm > 4) return ST_BOUNDS;
    if (csphase != 1 && csphase != -1) return ST_BOUNDS;
    if (lmax_calc > lmax) return ST_BOUNDS;
    if (lmax_calc < 0 || cilm_d1 < 1 || cilm_d2 < 1) return ST_BOUNDS;
    int Lc = lmax;
    if (cilm_d2 - 1 < Lc) Lc = cilm_d2 - 1;
    if (cilm_d1 - 1 < Lc) Lc = cilm_d1 - 1;
    if (lmax_calc < Lc) Lc = lmax_calc;
    sh.lmax = lmax;
    sh.Lc = Lc;
    sh.nrows = nlat;
    sh.ext = extend;
    return ST_OK;
}

// src/SHExpandGLQ.f95:100-250
inline int check_shexpand_glq(int grid_d0, int grid_d1, const double *w, int nw, const double *zero, int nzero,
                              int norm, int csphase, int lmax_calc, int cilm_d1, int cilm_d2, Shape &sh)
{
    if (grid_d0 < 1) return ST_DIMS;
    if (grid_d0 - 1 > kMaxArgDegree) return ST_BOUNDS;
    const int lmax = grid_d0 - 1;
    if (lmax_calc > lmax) return ST_BOUNDS;
    if (lmax_calc < 0) return ST_BOUNDS;
    const int Lc = lmax_calc < lmax ? lmax_calc : lmax;
    if (cilm_d1 < Lc + 1 || cilm_d2 < Lc + 1) return ST_DIMS;
    if (grid_d1 < 2 * lmax + 1 || grid_d0 < lmax + 1) return ST_DIMS;
    if (!w || nw < lmax + 1) return ST_DIMS;
    if (!zero) return ST_OTHER;
    if (nzero < lmax + 1) return ST_DIMS;
    if (norm < 1 || norm > 4) return ST_BOUNDS;
    if (csphase != 1 && csphase != -1) return ST_BOUNDS;
    sh.lmax = lmax;
    sh.Lc = Lc;
    sh.nrows = lmax + 1;
    return ST_OK;
}

// The same checks with the table plx in place of the nodes (the
// reference's plx branch): the table's size follows from lmax, and no table
// at all is exit status 5 as for missing nodes (src/MakeGridGLQ.f95:162-190,
// src/SHExpandGLQ.f95:166-193).
inline int check_makegrid_glq_plx(int cilm_d1, int cilm_d2, int lmax, const double *plx, int norm, int csphase,
                                  int lmax_calc, int extend, int grid_d0, int grid_d1, Shape &sh)
{
    if (lmax < 0 || lmax > kMaxArgDegree) return ST_BOUNDS;
    return check_makegrid_glq(cilm_d1, cilm_d2, lmax, lmax + 1, plx, norm, csphase, lmax_calc, extend, grid_d0,
                              grid_d1, sh);
}

inline int check_shexpand_glq_plx(int grid_d0, int grid_d1, const double *w, int nw, const double *plx, int norm,
                                  int csphase, int lmax_calc, int cilm_d1, int cilm_d2, Shape &sh)
{
    return check_shexpand_glq(grid_d0, grid_d1, w, nw, plx, grid_d0, norm, csphase, lmax_calc, cilm_d1, cilm_d2,
                              sh);
}

}  // namespace shtg
