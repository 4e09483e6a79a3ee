// crt_precal.hip -- element matrices per wavenumber and mixed-boundary factors.
//
// elbg(e, imn, k) = S1 + k^2 S2 (types 3, 5) or the condensed four-triangle
// quadrilateral (type 8); relbg(rel, imn) = boundary mass matrix; kg(rel,l,k) =
// beta(l,k) with the modified Bessel K0/K1 (src/precal.f90:79-215,
// src/elem3.f90:68-94, src/elem5.f90:42-67, src/elem8.f90:102-194,
// src/elem1.f90:50-64, src/beta.f90:69-121, src/bessk0.f90 / bessk1.f90 /
// bessi0.f90 / bessi1.f90).  All FP64; one thread per (element, wavenumber)
// and per (boundary element, electrode, wavenumber).
#include "crt_common.hpp"

namespace {

// Polynomial approximations of Abramowitz & Stegun 9.8.1-9.8.8 (the same
// coefficient tables the reference evaluates, kept alike so that kg agrees).
__device__ double bessi0(double x) {
    double ax = fabs(x), y;
    if (ax < 3.75) {
        y = (x / 3.75) * (x / 3.75);
        return 1.0 + y * (3.5156229 + y * (3.0899424 + y * (1.2067492 + y * (0.2659732 +
               y * (0.360768e-1 + y * 0.45813e-2)))));
    }
    y = 3.75 / ax;
    return (exp(ax) / sqrt(ax)) * (0.39894228 + y * (0.1328592e-1 + y * (0.225319e-2 +
           y * (-0.157565e-2 + y * (0.916281e-2 + y * (-0.2057706e-1 + y * (0.2635537e-1 +
           y * (-0.1647633e-1 + y * 0.392377e-2))))))));
}

__device__ double bessi1(double x) {
    double ax = fabs(x), y;
    if (ax < 3.75) {
        y = (x / 3.75) * (x / 3.75);
        return x * (0.5 + y * (0.87890594 + y * (0.51498869 + y * (0.15084934 + y * (0.2658733e-1 +
               y * (0.301532e-2 + y * 0.32411e-3))))));
    }
    y = 3.75 / ax;
    return (exp(ax) / sqrt(ax)) * (0.39894228 + y * (-0.3988024e-1 + y * (-0.362018e-2 +
           y * (0.163801e-2 + y * (-0.1031555e-1 + y * (0.2282967e-1 + y * (-0.2895312e-1 +
           y * (0.1787654e-1 + y * -0.420059e-2))))))));
}

__device__ double bessk0(double x) {
    double y;
    if (x <= 2.0) {
        y = x * x / 4.0;
        return (-log(x / 2.0) * bessi0(x)) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 +
               y * (0.3488590e-1 + y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5))))));
    } else if (x > 5e2) {
        return 0.0;
    }
    y = 2.0 / x;
    return (exp(-x) / sqrt(x)) * (1.25331414 + y * (-0.7832358e-1 + y * (0.2189568e-1 +
           y * (-0.1062446e-1 + y * (0.587872e-2 + y * (-0.251540e-2 + y * 0.53208e-3))))));
}

__device__ double bessk1(double x) {
    double y;
    if (x <= 2.0) {
        y = x * x / 4.0;
        return (log(x / 2.0) * bessi1(x)) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 +
               y * (-0.18156897 + y * (-0.1919402e-1 + y * (-0.110404e-2 + y * -0.4686e-4))))));
    } else if (x > 5e2) {
        return 0.0;
    }
    y = 2.0 / x;
    return (exp(-x) / sqrt(x)) * (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 +
           y * (0.1504268e-1 + y * (-0.780353e-2 + y * (0.325614e-2 + y * -0.68245e-3))))));
}

// base matrices of the linear triangle (column-major data of src/elem3.f90:55-58)
__constant__ int c_s1[9] = {1, -1, 0, -1, 1, 0, 0, 0, 0};
__constant__ int c_s2[9] = {2, -1, -1, -1, 0, 1, -1, 1, 0};
__constant__ int c_s3[9] = {1, 0, -1, 0, 0, 0, -1, 0, 1};
__constant__ int c_s4[9] = {2, 1, 1, 1, 2, 1, 1, 1, 2};
// bilinear parallelogram (src/elem5.f90:36-39)
__constant__ int c_p1[16] = {2, -2, -1, 1, -2, 2, 1, -1, -1, 1, 2, -2, 1, -1, -2, 2};
__constant__ int c_p2[16] = {1, 0, -1, 0, 0, -1, 0, 1, -1, 0, 1, 0, 0, 1, 0, -1};
__constant__ int c_p3[16] = {2, 1, -1, -2, 1, 2, -2, -1, -1, -2, 2, 1, -2, -1, 1, 2};
__constant__ int c_p4[16] = {4, 2, 1, 2, 2, 4, 2, 1, 1, 2, 4, 2, 2, 1, 2, 4};

// Element matrices written straight to elbg(e, imn, k) (packed lower triangle,
// m outer, n <= m).  All indices are compile-time after unrolling, so nothing
// lives in private (scratch) memory.
__device__ __forceinline__ int elbg_tri(const double *xk, const double *yk, double kw, double *dst, size_t ld) {
    double x21 = xk[1] - xk[0], x31 = xk[2] - xk[0], y21 = yk[1] - yk[0], y31 = yk[2] - yk[0];
    double det = x21 * y31 - x31 * y21;
    if (det <= 0.0) return CRT_ERR_GEOM;
    double a = (x31 * x31 + y31 * y31) / det, b = -(x31 * x21 + y31 * y21) / det,
           c = (x21 * x21 + y21 * y21) / det;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = 0; n <= m; ++n) {
            const int t = n * 3 + m, imn = m * (m + 1) / 2 + n;
            double elmas = (a * c_s1[t] + b * c_s2[t] + c * c_s3[t]) / 2.0;
            double elmam = det * c_s4[t] / 2.4e1;
            dst[imn * ld] = elmas + elmam * kw * kw;
        }
    return 0;
}

__device__ __forceinline__ int elbg_par(const double *xk, const double *yk, double kw, double *dst, size_t ld) {
    double x21 = xk[1] - xk[0], x31 = xk[3] - xk[0], y21 = yk[1] - yk[0], y31 = yk[3] - yk[0];
    double det = x21 * y31 - x31 * y21;
    if (det <= 0.0) return CRT_ERR_GEOM;
    double a = (x31 * x31 + y31 * y31) / det, b = -(x31 * x21 + y31 * y21) / det,
           c = (x21 * x21 + y21 * y21) / det;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n <= m; ++n) {
            const int t = n * 4 + m, imn = m * (m + 1) / 2 + n;
            double elmas = (a * c_p1[t] + 3.0 * b * c_p2[t] + c * c_p3[t]) / 6.0;
            double elmam = det * c_p4[t] / 3.6e1;
            dst[imn * ld] = elmas + elmam * kw * kw;
        }
    return 0;
}

// sub-triangle k of the four-triangle quadrilateral: local node ids
__device__ __forceinline__ void q8_tri(int k, int &i0, int &i1, int &i2) {
    i0 = (k == 0 || k == 3) ? 0 : (k == 1 ? 1 : 4);
    i1 = (k == 0) ? 1 : (k == 3 ? 4 : 2);
    i2 = (k == 0 || k == 1) ? 4 : 3;
}

__device__ __forceinline__ int elbg_quad8(const double *xk, const double *yk, double kw, double *dst, size_t ld) {
    double ms[5][5], mm[5][5], xd[5], yd[5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) { ms[i][j] = 0.0; mm[i][j] = 0.0; }
    double a1 = xk[2] - xk[0], a2 = xk[3] - xk[1];
    if (fabs(a1) <= 1e-12 || fabs(a2) <= 1e-12) return CRT_ERR_GEOM;
    a1 = (yk[2] - yk[0]) / a1;
    a2 = (yk[3] - yk[1]) / a2;
    if (fabs(a1 - a2) <= 1e-12) return CRT_ERR_GEOM;
    xd[4] = (a1 * xk[0] - yk[0] - a2 * xk[1] + yk[1]) / (a1 - a2);
    yd[4] = yk[0] + a1 * (xd[4] - xk[0]);
#pragma unroll
    for (int i = 0; i < 4; ++i) { xd[i] = xk[i]; yd[i] = yk[i]; }
    int bad = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int ik[3];
        q8_tri(k, ik[0], ik[1], ik[2]);
        double x21 = xd[ik[1]] - xd[ik[0]], x31 = xd[ik[2]] - xd[ik[0]];
        double y21 = yd[ik[1]] - yd[ik[0]], y31 = yd[ik[2]] - yd[ik[0]];
        double det = x21 * y31 - x31 * y21;
        if (det <= 1e-12) bad = 1;
        double a = (x31 * x31 + y31 * y31) / det, b = -(x31 * x21 + y31 * y21) / det,
               c = (x21 * x21 + y21 * y21) / det;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int t = j * 3 + i;
                ms[ik[i]][ik[j]] += (a * c_s1[t] + b * c_s2[t] + c * c_s3[t]) / 2.0;
                mm[ik[i]][ik[j]] += det * c_s4[t] / 2.4e1;
            }
    }
    if (bad) return CRT_ERR_GEOM;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) ms[i][j] = ms[i][j] + kw * kw * mm[i][j];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n <= m; ++n) {
            const int imn = m * (m + 1) / 2 + n;
            dst[imn * ld] = ms[m][n] - ms[m][4] * ms[4][n] / ms[4][4];
        }
    return 0;
}

}  // namespace

// one thread per (area element, wavenumber); errors are flagged with a plain
// store (any non-zero code stops the caller)
__global__ void __launch_bounds__(256) crt_precal_area_kernel(
    int elanz, int maxn, int npk, int kwnanz, const int *__restrict__ etyp,
    const int *__restrict__ nrel, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ kwn, double *__restrict__ elbg, int *__restrict__ status) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= elanz * kwnanz) return;
    int e = t % elanz, k = t / elanz;
    int typ = etyp[e];
    double xk[4], yk[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        int nd = (m < 3 || typ != 3) ? nrel[(size_t)e * maxn + m] - 1 : 0;
        xk[m] = x[nd];
        yk[m] = y[nd];
    }
    double *dst = elbg + (size_t)k * npk * elanz + e;
    const size_t ld = (size_t)elanz;
    int err;
    if (typ == 3) err = elbg_tri(xk, yk, kwn[k], dst, ld);
    else if (typ == 5) err = elbg_par(xk, yk, kwn[k], dst, ld);
    else if (typ == 8) err = elbg_quad8(xk, yk, kwn[k], dst, ld);
    else err = CRT_ERR_TYPE;
    if (err) *status = err;
}

// one thread per (boundary element, electrode, wavenumber); type-11 only
__global__ void __launch_bounds__(256) crt_precal_bnd_kernel(
    int elanz, int relanz, int maxn, int npk, int eanz, int kwnanz, const int *__restrict__ btyp,
    const int *__restrict__ nrel, const double *__restrict__ x, const double *__restrict__ y,
    const int *__restrict__ enr, const double *__restrict__ kwn, double sytop,
    double *__restrict__ relbg, double *__restrict__ kg, int *__restrict__ status) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)relanz * eanz * kwnanz) return;
    int rel = (int)(t % relanz);
    int l = (int)((t / relanz) % eanz);
    int k = (int)(t / ((long)relanz * eanz));
    if (btyp[rel] != 11) return;
    const int *nd = nrel + (size_t)(elanz + rel) * maxn;
    double xk0 = x[nd[0] - 1], yk0 = y[nd[0] - 1], xk1 = x[nd[1] - 1], yk1 = y[nd[1] - 1];
    if (l == 0 && k == 0) {
        double x21 = xk1 - xk0, y21 = yk1 - yk0;
        double len = sqrt(x21 * x21 + y21 * y21);
        relbg[(size_t)0 * relanz + rel] = len * 2.0 / 6.0;
        relbg[(size_t)1 * relanz + rel] = len * 1.0 / 6.0;
        relbg[(size_t)2 * relanz + rel] = len * 2.0 / 6.0;
    }
    int en = enr[l];
    double xs = x[en], ys = y[en];
    double xr = (xk0 + xk1) / 2.0, yr = (yk0 + yk1) / 2.0;
    double x3 = xr - xs;
    double y3m = (yr - sytop) - (ys - sytop);
    double y3p = (yr - sytop) + (ys - sytop);
    double r1m = sqrt(x3 * x3 + y3m * y3m), r1p = sqrt(x3 * x3 + y3p * y3p);
    double x4 = yk0 - yk1, y4 = xk1 - xk0;
    double r2 = sqrt(x4 * x4 + y4 * y4);
    if (r1m < 1e-12 || r1p < 1e-12 || r2 < 1e-12) { *status = CRT_ERR_BETA; return; }
    double cm = (x3 * x4 + y3m * y4) / r1m / r2;
    double cp = (x3 * x4 + y3p * y4) / r1p / r2;
    double kw = kwn[k];
    double bk0m = bessk0(r1m * kw), bk0p = bessk0(r1p * kw);
    double bk1m = bessk1(r1m * kw), bk1p = bessk1(r1p * kw);
    double b;
    if (bk0m + bk0p == 0.0) b = (cm + cp) / 2.0;
    else b = (cm * bk1m + cp * bk1p) / (bk0m + bk0p);
    kg[((size_t)k * eanz + l) * relanz + rel] = kw * b;
}

int crt_launch_precal(hipStream_t s, int elanz, int relanz, int maxn, int npk, int eanz, int kwnanz,
                      const int *etyp, const int *btyp, const int *nrel, const double *x,
                      const double *y, const int *enr, const double *kwn, double sytop, double *elbg,
                      double *relbg, double *kg, int *status) {
    if (elanz > 0) {
        int n = elanz * kwnanz;
        hipLaunchKernelGGL(crt_precal_area_kernel, dim3((n + 255) / 256), dim3(256), 0, s, elanz, maxn,
                           npk, kwnanz, etyp, nrel, x, y, kwn, elbg, status);
    }
    if (relanz > 0 && eanz > 0) {
        long n = (long)relanz * eanz * kwnanz;
        hipLaunchKernelGGL(crt_precal_bnd_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           elanz, relanz, maxn, npk, eanz, kwnanz, btyp, nrel, x, y, enr, kwn, sytop,
                           relbg, kg, status);
    }
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// ------------------------------------------------ singularity removal, analytic solution
// Reference conductivity (src/refsig.f90:33-70): sigma0 = exp(sum_e A_e log sigma_e /
// sum_e A_e) over the area elements, A_e by Gauss' area formula (triangles from nodes
// 1-3, quadrilaterals as the triangles 1-2-3 and 1-3-4).  One workgroup: each thread sums
// a stride of the elements, the partial sums meet in a fixed-order tree.
__global__ void __launch_bounds__(256) crt_refsig_kernel(int elanz, int maxn, const int *__restrict__ etyp,
                                                         const int *__restrict__ nrel, const double *__restrict__ x,
                                                         const double *__restrict__ y, const zd *__restrict__ sigma,
                                                         zd *__restrict__ sigma0) {
    __shared__ double red[3][256];
    double sr = 0.0, si = 0.0, ar = 0.0;
    for (int e = threadIdx.x; e < elanz; e += 256) {
        const int *nd = nrel + (size_t)e * maxn;
        const double xk = x[nd[0] - 1], yk = y[nd[0] - 1];
        const double ax = x[nd[1] - 1] - xk, ay = y[nd[1] - 1] - yk;
        const double bx = x[nd[2] - 1] - xk, by = y[nd[2] - 1] - yk;
        double dum = 0.0;
        if (etyp[e] == 3) {
            dum = fabs(ax * by - ay * bx) / 2.0;
        } else {
            const double cx = x[nd[3] - 1] - xk, cy = y[nd[3] - 1] - yk;
            dum = (fabs(ax * by - ay * bx) + fabs(bx * cy - by * cx)) / 2.0;
        }
        const zd s = sigma[e];
        sr += dum * log(hypot(s.re, s.im));
        si += dum * atan2(s.im, s.re);
        ar += dum;
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    red[2][threadIdx.x] = ar;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int v = 0; v < 3; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = exp(red[0][0] / red[2][0]), ph = red[1][0] / red[2][0];
        *sigma0 = zd{m * cos(ph), m * sin(ph)};
    }
}

// Analytic potential of a unit current in a homogeneous half-space of conductivity
// sigma0 (src/potana.f90:51-97), one workgroup per (electrode, wavenumber) system r:
// pota[r][j] = (K0(k rm) + K0(k rp)) / (2 pi sigma0), rm and rp the distances of node j
// to the electrode and to its mirror image in y = 0; the electrode's own node gets five
// times the largest value of the other nodes.
__global__ void __launch_bounds__(256) crt_potana_kernel(int sanz, int nrhs, const int4 *__restrict__ rhs,
                                                         const int *__restrict__ enr, const double *__restrict__ x,
                                                         const double *__restrict__ y, const double *__restrict__ kwn,
                                                         const zd *__restrict__ sigma0, zd *__restrict__ pota,
                                                         size_t pstride) {
    __shared__ double red[256];
    const int r = blockIdx.x;
    if (r >= nrhs) return;
    const int4 q = rhs[r];
    const int idum = enr[q.x];
    const double kk = kwn[q.y], xk1 = x[idum], yk1 = y[idum];
    zd *p = pota + (size_t)r * pstride;
    double pmax = 0.0;
    for (int j = threadIdx.x; j < sanz; j += 256) {
        if (j == idum) continue;
        const double x21 = x[j] - xk1, y21m = y[j] - yk1, y21p = y[j] + yk1;
        const double rm = sqrt(x21 * x21 + y21m * y21m), rp = sqrt(x21 * x21 + y21p * y21p);
        const double dum = bessk0(rm * kk) + bessk0(rp * kk);
        pmax = fmax(pmax, dum);
        p[j] = zd{dum, 0.0};
    }
    red[threadIdx.x] = pmax;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    const double pi = acos(-1.0);
    const zd s0 = *sigma0;
    const double den = s0.re * s0.re + s0.im * s0.im;
    const zd sc = zd{0.5 / pi * s0.re / den, -0.5 / pi * s0.im / den};
    for (int j = threadIdx.x; j < sanz; j += 256) {
        const double v = j == idum ? 5.0 * red[0] : p[j].re;
        p[j] = zd{v * sc.re, v * sc.im};
    }
}

int crt_launch_refsig(hipStream_t s, int elanz, int maxn, const int *etyp, const int *nrel, const double *x,
                      const double *y, const zd *sigma, zd *sigma0) {
    hipLaunchKernelGGL(crt_refsig_kernel, dim3(1), dim3(256), 0, s, elanz, maxn, etyp, nrel, x, y, sigma, sigma0);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_potana(hipStream_t s, int sanz, int nrhs, const int4 *rhs, const int *enr, const double *x,
                      const double *y, const double *kwn, const zd *sigma0, zd *pota, size_t pstride) {
    if (nrhs < 1) return 0;
    hipLaunchKernelGGL(crt_potana_kernel, dim3(nrhs), dim3(256), 0, s, sanz, nrhs, rhs, enr, x, y, kwn, sigma0, pota,
                       pstride);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
