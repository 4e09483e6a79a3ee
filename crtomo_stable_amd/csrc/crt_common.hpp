// crt_common.hpp -- shared types of the CRTomo HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

// Error numbers follow the reference's errmod codes (src/error.txt):
//   19 band width exceeded, 26 element geometry, 27 zero diagonal,
//   28 non-positive pivot, 35 beta geometry, 73 duplicate electrodes,
//   82 zero voltage, 93 vanishing geometric factor, 94/97 allocation.  Library-only
//   codes start at 900.
enum {
    CRT_OK = 0,
    CRT_ERR_BAND = 19,
    CRT_ERR_GEOM = 26,
    CRT_ERR_ZERODIAG = 27,
    CRT_ERR_PIVOT = 28,
    CRT_ERR_BETA = 35,
    CRT_ERR_ELEC = 73,
    CRT_ERR_ZEROVOLT = 82,
    CRT_ERR_KFAK = 93,
    CRT_ERR_BOUNDARY_NODE = 30,
    CRT_ERR_BOUNDARY_MIX = 101,
    CRT_ERR_ALLOC = 97,
    CRT_ERR_TYPE = 18,
    CRT_ERR_HIP = 900,
    CRT_ERR_STATE = 901,
    CRT_ERR_ARG = 902,
    CRT_ERR_UNSUPPORTED = 903,
    CRT_ERR_SCRATCH = 904,
};

#define CRT_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "crtomo_hip: %s failed at %s:%d: %s\n", #call, __FILE__, __LINE__, \
                    hipGetErrorString(e_));                                                \
            return CRT_ERR_HIP;                                                            \
        }                                                                                  \
    } while (0)

// ---------------------------------------------------------------- complex FP64
struct zd {
    double re, im;
};

__host__ __device__ __forceinline__ zd zmake(double r, double i) { return zd{r, i}; }
__host__ __device__ __forceinline__ zd zadd(zd a, zd b) { return zd{a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ zd zsub(zd a, zd b) { return zd{a.re - b.re, a.im - b.im}; }
__host__ __device__ __forceinline__ zd zmul(zd a, zd b) {
    return zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ __forceinline__ zd zscale(zd a, double s) { return zd{a.re * s, a.im * s}; }
// a - b*c
__host__ __device__ __forceinline__ zd zfms(zd a, zd b, zd c) {
    return zd{a.re - b.re * c.re + b.im * c.im, a.im - b.re * c.im - b.im * c.re};
}
// a + b*c
__host__ __device__ __forceinline__ zd zfma(zd a, zd b, zd c) {
    return zd{a.re + b.re * c.re - b.im * c.im, a.im + b.re * c.im + b.im * c.re};
}
__host__ __device__ __forceinline__ double zabs(zd a) { return hypot(a.re, a.im); }
// 1/z
__host__ __device__ __forceinline__ zd zinv(zd z) {
    // Smith's algorithm for 1/(c+id)
    if (fabs(z.re) >= fabs(z.im)) {
        double r = z.im / z.re, den = z.re + z.im * r;
        return zd{1.0 / den, -r / den};
    } else {
        double r = z.re / z.im, den = z.re * r + z.im;
        return zd{r / den, -1.0 / den};
    }
}
__host__ __device__ __forceinline__ zd zdiv(zd a, zd b) {
    // Smith's algorithm for a/b
    if (fabs(b.re) >= fabs(b.im)) {
        double r = b.im / b.re, den = b.re + b.im * r;
        return zd{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
    } else {
        double r = b.re / b.im, den = b.re * r + b.im;
        return zd{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
    }
}
// principal square root (branch cut on the negative real axis)
__host__ __device__ __forceinline__ zd zsqrt(zd z) {
    double r = hypot(z.re, z.im);
    if (r == 0.0) return zd{0.0, z.im};
    if (z.re >= 0.0) {
        double t = sqrt(0.5 * (r + z.re));
        return zd{t, z.im / (2.0 * t)};
    }
    double t = sqrt(0.5 * (r - z.re));
    return zd{fabs(z.im) / (2.0 * t), copysign(t, z.im)};
}

// 1/sqrt(m) for finite m > 0: hardware estimate y0 and one third-order step
// y = y0 (1 + e/2 + 3e^2/8), e = 1 - m y0^2, which cubes the estimate's error
__device__ __forceinline__ double rsqrt_f64(double m) {
    double y0 = __builtin_amdgcn_rsq(m);
    double e = fma(-m * y0, y0, 1.0);
    return fma(y0 * e, fma(e, 0.375, 0.5), y0);
}

// Pivot of column c: L = principal sqrt(s) (branch cut on the negative real
// axis, as Fortran's CDSQRT) and 1/L = conj(L)/|s|, from two reciprocal square
// roots.  bad is set for a zero (or non-finite) pivot: src/chol.f90 error 28.
__device__ __forceinline__ void pivot_fast(zd s, zd &L, zd &inv, bool &bad) {
    double m2 = fma(s.re, s.re, s.im * s.im);
    bad = !(m2 > 0.0) || !(m2 < 1.0e300);
    double rm = rsqrt_f64(m2);  // 1/|s|
    double mag = m2 * rm;       // |s|
    double h = 0.5 * (mag + fabs(s.re));
    double rh = rsqrt_f64(h);   // 1/t, t = sqrt((|s| + |Re s|)/2)
    double t = h * rh;
    double q = 0.5 * fabs(s.im) * rh;  // |Im s|/(2t)
    bool pos = s.re >= 0.0;
    L = zd{pos ? t : q, copysign(pos ? q : t, s.im)};
    inv = zd{L.re * rm, -L.im * rm};
}

// complex FP32 (panels of the mixed-precision factorisation)
struct alignas(8) zf {
    float re, im;
};

// ---------------------------------------------------------------- geometry
// Band tile used by the FP64 factorisation: rows are grouped into blocks of
// CRT_NB rows; the factor is stored as block-column panels
//   panel[J][a][r][c] = L(J*NB + a*NB + r, J*NB + c),  a = 0..D,
// where D = ceil(mb / NB) is the number of sub-diagonal blocks.
#ifndef CRT_NB
#define CRT_NB 8
#endif

struct DevGrid {
    int sanz, mb, elanz, relanz, eanz, maxn, npk, kwnanz, nanz;
    int nblk, D;              // row blocks and sub-diagonal block count
    int nnz;                  // lower-triangle nonzeros
    int fstride;              // per-system stride of fak / dga / y arrays (rows padded to a block)
    int vstride;              // per-system stride of the CSR values: nnz entries, then 0 and 1
    // CSR of the lower triangle of A in node order (0-based), diagonal last
    const int *row_ptr;       // sanz+1
    const int *col;           // nnz
    const int *crow;          // nnz: row of each entry
    const int *cptr;          // nnz+1 contribution ranges
    const int *contrib;       // >=0: e*npk+imn (area); <0: -(rel*npk+imn)-1 (mixed boundary)
    const int *rnr;           // relanz (0-based area element)
    const int *enr;           // eanz (0-based node)
    const double *elbg;       // [k][imn][e]
    const double *relbg;      // [imn][rel]
    const double *kg;         // [k][l][rel]
    const zd *sigma;          // elanz
    // 16x16 tile map of the lower band: for block row I and block diagonal d
    // (0..D16), tmap[((I*(D16+1) + d)*16 + r)*16 + c] is the index into a system's
    // values of entry (16I + r, 16(I-d) + c): its CSR index, nnz if the entry is zero,
    // nnz+1 (the value 1) for the unit diagonal of padding rows past sanz, so every
    // tile element is a plain load
    const int *tmap;
    int tmapD;                // D16 the map was built for (0 if none)
    // systems of different structure in one launch (the condensed forward's interior
    // pieces): system s reads the map at tmap + sys[s].x * tmap_sys; 0: one map for all
    long tmap_sys = 0;
    // Every contribution list holds its area elements before its boundary elements
    // (checked on the host), so the area part of an entry is the same sum for every
    // source: area[k*nnz + q] holds it per wavenumber, csplit[q] is where the boundary
    // contributions of entry q start.  area null: every entry is summed in full.
    const int *csplit;
    const zd *area;
    // 1 for the DC inversion (kompadc / scaldc / choldc on DBLE(sigma)): a diagonal entry
    // or pivot that is not real and positive is an error there (src/scaldc.f90:54-58,
    // src/choldc.f90:66-70), where the complex routines only stop on zero
    int dc;
    // Dirichlet nodes (type-13 elements, src/randb.f90; homogeneous boundary values,
    // src/randb2.f90): per CSR entry 0 as assembled, 1 forced to zero (a row or column
    // of such a node), 2 forced to one (its diagonal); null when the grid has none
    const unsigned char *dmask;
};

// Gauss-Newton CG state shared by crt_cg.hip and crt_api.hip
struct CGScalars {
    double dr, dr0, dr1, alpha, beta;
    double rz, rz1;  // r^H M r of this and the previous iteration (preconditioned CG)
    int k, done, ncg, pad;
};

// factorised sparse approximate inverse M = G^H G of the CG operator: row j of the
// lower-triangular G has CRT_FSAI_NP entries on the columns fsai_pat[j*NP + s]
// (-1 unused, the last slot is j itself)
#define CRT_FSAI_NP 8

struct CGArgs {
    int nanz, manz, ltri, nx, maxn, ncgmax;
    double lam, eps;
    const zd *J;            // (nanz, manz) column-major
    const float2 *J32;      // the same in complex FP32 for the iterations, or null (FP64)
    const double *Jr;       // real J of the DC / FPI branches (sensdc or DBLE(sens)), or null
    const float *Jr32;      // the real J in FP32 for the iterations of those branches, or null
    int part;               // CG right-hand side: 0 bvec, 1 its real part (DC), 2 its imaginary part (FPI)
    const double *w;        // nanz: wmatd * wdfak
    const double *smatm;    // (manz, ncol) column-major
    const int *nachbar;     // (manz, maxn+1) column-major, 1-based (ltri 1/2)
    const double *cgfac;    // manz
    zd *r, *p, *q, *bv, *dpar;  // manz
    zd *ap;                 // nanz (W-weighted after reduction: wap)
    zd *part_ap;            // [ncb][nanz]
    double *part_dr1;       // per block of the J^H kernel
    double *part_rr;        // per block of the update kernel
    int n_dr1, n_rr, ncb, cb;
    const zd *fg;           // FSAI: G entries [manz][CRT_FSAI_NP], or null (Jacobi only)
    const int *fpat;        // FSAI: column of each G entry, -1 unused
    const int *ftp;         // FSAI: transposed pattern, entry index i*NP+s of G(i, j), -1 unused
    zd *t, *z;              // FSAI: G r and z = G^H G r
    double *part_rz;        // FSAI: per block Re(r^H z)
    const CGScalars *sc;    // state after the previous iteration (read only)
    CGScalars *scn;         // state of this iteration (written by block 0)
    double *cgres;          // ncgmax+1
};
