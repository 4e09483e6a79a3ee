/*
 * crt_oracle.c -- plain-C restatement of the CRTomo forward / Jacobian / CG path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 * Every function names the reference file:line it restates.  Built with
 * gcc -O2 and no fast-math (the parity build of the reference uses -O2,
 * .github/workflows/build_crtomo.yml:22).
 *
 * Pinning of each function (what checks it besides itself) is listed in
 * oracle/README.md.
 */
#include "crt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NPK(maxn) ((maxn) * ((maxn) + 1) / 2)

static inline int imax_(int a, int b) { return a > b ? a : b; }
static inline int imin_(int a, int b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ Bessel */
/* Polynomial approximations of Abramowitz & Stegun 9.8.1-9.8.8, evaluated in
 * Horner form with the same coefficient tables as src/bessi0.f90:11-24,
 * src/bessi1.f90:11-24, src/bessk0.f90:9-26, src/bessk1.f90:8-25. */
double orc_bessi0(double x) {
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

double orc_bessi1(double x) {
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

double orc_bessk0(double x) {
    double y;
    if (x <= 2.0) {
        y = x * x / 4.0;
        return (-log(x / 2.0) * orc_bessi0(x)) + (-0.57721566 + y * (0.42278420 + y * (0.23069756 +
               y * (0.3488590e-1 + y * (0.262698e-2 + y * (0.10750e-3 + y * 0.74e-5))))));
    } else if (x > 5e2) {
        return 0.0;
    }
    y = 2.0 / x;
    return (exp(-x) / sqrt(x)) * (1.25331414 + y * (-0.7832358e-1 + y * (0.2189568e-1 +
           y * (-0.1062446e-1 + y * (0.587872e-2 + y * (-0.251540e-2 + y * 0.53208e-3))))));
}

double orc_bessk1(double x) {
    double y;
    if (x <= 2.0) {
        y = x * x / 4.0;
        return (log(x / 2.0) * orc_bessi1(x)) + (1.0 / x) * (1.0 + y * (0.15443144 + y * (-0.67278579 +
               y * (-0.18156897 + y * (-0.1919402e-1 + y * (-0.110404e-2 + y * -0.4686e-4))))));
    } else if (x > 5e2) {
        return 0.0;
    }
    y = 2.0 / x;
    return (exp(-x) / sqrt(x)) * (1.25331414 + y * (0.23498619 + y * (-0.3655620e-1 +
           y * (0.1504268e-1 + y * (-0.780353e-2 + y * (0.325614e-2 + y * -0.68245e-3))))));
}

/* -------------------------------------------------------------- quadrature */
/* Gauss-Legendre nodes/weights by Newton iteration on P_n (the textbook method
 * used by src/gauleg.f90:9-40): nodes ascending in [x1,x2]. */
void orc_gauleg(double x1, double x2, double *x, double *w, int n) {
    const double eps = 3e-14;
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; ++i) {
        double z = cos(3.141592653589793 * ((double)i - 0.25) / ((double)n + 0.5)), z1, pp;
        do {
            double p1 = 1.0, p2 = 0.0, p3;
            for (int j = 1; j <= n; ++j) {
                p3 = p2; p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > eps);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* Lanczos log-gamma (src/gammln.f90:6-25, coefficient table of the published
 * Lanczos g=5, n=6 approximation). */
double orc_gammln(double xx) {
    static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                  -1.231739572450155, .1208650973866179e-2, -.5395239384953e-5};
    const double stp = 2.5066282746310005;
    double x = xx, y = x, tmp = x + 5.5, ser = 1.000000000190015;
    tmp = (x + 0.5) * log(tmp) - tmp;
    for (int j = 0; j < 6; ++j) { y += 1.0; ser += cof[j] / y; }
    return tmp + log(stp * ser / x);
}

/* Gauss-Laguerre nodes/weights by Newton iteration with the standard initial
 * guesses (src/gaulag.f90:6-44). */
void orc_gaulag(double *x, double *w, int n, double alf) {
    const double eps = 3e-14;
    double z = 0.0, z1, pp = 0.0, p1, p2, p3;
    for (int i = 1; i <= n; ++i) {
        if (i == 1) {
            z = (1.0 + alf) * (3.0 + 0.92 * alf) / (1.0 + 2.4 * n + 1.8 * alf);
        } else if (i == 2) {
            z = z + (15.0 + 6.25 * alf) / (1.0 + 0.9 * alf + 2.5 * n);
        } else {
            double ai = (double)(i - 2);
            z = z + ((1.0 + 2.55 * ai) / (1.9 * ai) + 1.26 * ai * alf / (1.0 + 3.5 * ai)) *
                        (z - x[i - 3]) / (1.0 + 0.3 * alf);
        }
        p2 = 0.0;
        for (int its = 1; its <= 10; ++its) {
            p1 = 1.0; p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                p3 = p2; p2 = p1;
                p1 = ((double)(2 * j - 1) + alf - z) * p2 - ((double)(j - 1) + alf) * p3;
                p1 /= (double)j;
            }
            pp = ((double)n * p1 - ((double)n + alf) * p2) / z;
            z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) <= eps) break;
        }
        x[i - 1] = z;
        w[i - 1] = -exp(orc_gammln(alf + n) - orc_gammln((double)n)) / (pp * n * p2);
    }
}

/* Wavenumber set (src/rwaven.f90:52-151): amin/amax over current-potential
 * electrode distances, ganz = int(real(6*log10(amax/amin))) clamped to [2,26],
 * kwnanz = ganz+4, k0 = 1/(2*amin).  Returns errnr (73 on duplicate electrodes). */
int orc_rwaven(const orc_grid *g, int nanz, const int *strnr, const int *vnr,
               int *kwnanz, double *kwn, double *kwnwi, double *amin_out, double *amax_out) {
    double amin = 1e9, amax = 0.0;
    for (int i = 0; i < nanz; ++i) {
        int e[4];
        double xe[4] = {0, 0, 0, 0}, ye[4] = {0, 0, 0, 0};
        e[0] = strnr[i] % 10000; e[1] = (strnr[i] - e[0]) / 10000;
        e[2] = vnr[i] % 10000;   e[3] = (vnr[i] - e[2]) / 10000;
        for (int j = 0; j < 4; ++j)
            if (e[j] > 0) { int nd = g->enr[e[j] - 1] - 1; xe[j] = g->x[nd]; ye[j] = g->y[nd]; }
        for (int m = 2; m < 4; ++m) {
            if (e[m] <= 0) continue;
            for (int j = 0; j < 2; ++j) {
                if (e[j] <= 0) continue;
                double x21 = xe[m] - xe[j], y21 = ye[m] - ye[j];
                double d = sqrt(x21 * x21 + y21 * y21);
                amin = fmin(d, amin);
                amax = fmax(d, amax);
            }
        }
    }
    if (amin_out) *amin_out = amin;
    if (amax_out) *amax_out = amax;
    if (amin < 1e-12 || amax < 1e-12) return 73;
    const int lanz = 4;
    int ganz = (int)(float)(6.0 * log10(amax / amin));  /* INT(REAL(...)): REAL*4 conversion */
    ganz = imax_(ganz, 2);
    ganz = imin_(ganz, 30 - lanz);
    *kwnanz = ganz + lanz;
    double k0 = 1.0 / (2.0 * amin);
    orc_gauleg(0.0, 1.0, kwn, kwnwi, ganz);
    for (int i = 0; i < ganz; ++i) {
        kwnwi[i] = 2.0 * k0 * kwnwi[i] * kwn[i];
        kwn[i] = k0 * kwn[i] * kwn[i];
    }
    orc_gaulag(kwn + ganz, kwnwi + ganz, lanz, 0.0);
    for (int i = ganz; i < ganz + lanz; ++i) {
        kwnwi[i] = k0 * kwnwi[i] * exp(kwn[i]);
        kwn[i] = k0 * (kwn[i] + 1.0);
    }
    return 0;
}

/* --------------------------------------------------------- element matrices */
/* linear triangle (src/elem3.f90:68-94): stiffness elmas, mass elmam (3x3, col-major n) */
static int elem3(const double *xk, const double *yk, double *elmas, double *elmam, int ld) {
    static const int s1[9] = {1, -1, 0, -1, 1, 0, 0, 0, 0};
    static const int s2[9] = {2, -1, -1, -1, 0, 1, -1, 1, 0};
    static const int s3[9] = {1, 0, -1, 0, 0, 0, -1, 0, 1};
    static const int s4[9] = {2, 1, 1, 1, 2, 1, 1, 1, 2};
    double x21 = xk[1] - xk[0], x31 = xk[2] - xk[0];
    double y21 = yk[1] - yk[0], y31 = yk[2] - yk[0];
    double det = x21 * y31 - x31 * y21;
    if (det <= 0.0) return 26;
    double a = (x31 * x31 + y31 * y31) / det;
    double b = -(x31 * x21 + y31 * y21) / det;
    double c = (x21 * x21 + y21 * y21) / det;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int t = j * 3 + i; /* data statements fill column-major */
            elmas[j * ld + i] = (a * s1[t] + b * s2[t] + c * s3[t]) / 2.0;
            elmam[j * ld + i] = det * s4[t] / 2.4e1;
        }
    return 0;
}

/* parallelogram bilinear (src/elem5.f90:42-67) */
static int elem5(const double *xk, const double *yk, double *elmas, double *elmam, int ld) {
    static const int s1[16] = {2, -2, -1, 1, -2, 2, 1, -1, -1, 1, 2, -2, 1, -1, -2, 2};
    static const int s2[16] = {1, 0, -1, 0, 0, -1, 0, 1, -1, 0, 1, 0, 0, 1, 0, -1};
    static const int s3[16] = {2, 1, -1, -2, 1, 2, -2, -1, -1, -2, 2, 1, -2, -1, 1, 2};
    static const int s4[16] = {4, 2, 1, 2, 2, 4, 2, 1, 1, 2, 4, 2, 2, 1, 2, 4};
    double x21 = xk[1] - xk[0], x31 = xk[3] - xk[0];
    double y21 = yk[1] - yk[0], y31 = yk[3] - yk[0];
    double det = x21 * y31 - x31 * y21;
    if (det <= 0.0) return 26;
    double a = (x31 * x31 + y31 * y31) / det;
    double b = -(x31 * x21 + y31 * y21) / det;
    double c = (x21 * x21 + y21 * y21) / det;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            int t = j * 4 + i;
            elmas[j * ld + i] = (a * s1[t] + 3.0 * b * s2[t] + c * s3[t]) / 6.0;
            elmam[j * ld + i] = det * s4[t] / 3.6e1;
        }
    return 0;
}

/* quadrilateral of four linear triangles around the diagonal intersection, with
 * static condensation of the centre node (src/elem8.f90:102-194).  kel (4x4). */
static int elem8(const double *xk, const double *yk, double kwert, double *kel, int ld) {
    static const int s1[9] = {1, -1, 0, -1, 1, 0, 0, 0, 0};
    static const int s2[9] = {2, -1, -1, -1, 0, 1, -1, 1, 0};
    static const int s3[9] = {1, 0, -1, 0, 0, 0, -1, 0, 1};
    static const int s4[9] = {2, 1, 1, 1, 2, 1, 1, 1, 2};
    static const int tri[4][3] = {{1, 2, 5}, {2, 3, 5}, {5, 3, 4}, {1, 5, 4}};
    double elmas[25] = {0}, elmam[25] = {0}, xd[5], yd[5];
    double a1 = xk[2] - xk[0], a2 = xk[3] - xk[1];
    if (fabs(a1) <= 1e-12 || fabs(a2) <= 1e-12) return 26;
    a1 = (yk[2] - yk[0]) / a1;
    a2 = (yk[3] - yk[1]) / a2;
    if (fabs(a1 - a2) <= 1e-12) return 26;
    xd[4] = (a1 * xk[0] - yk[0] - a2 * xk[1] + yk[1]) / (a1 - a2);
    yd[4] = yk[0] + a1 * (xd[4] - xk[0]);
    for (int i = 0; i < 4; ++i) { xd[i] = xk[i]; yd[i] = yk[i]; }
    for (int k = 0; k < 4; ++k) {
        int ik[3] = {tri[k][0] - 1, tri[k][1] - 1, tri[k][2] - 1};
        double x21 = xd[ik[1]] - xd[ik[0]], x31 = xd[ik[2]] - xd[ik[0]];
        double y21 = yd[ik[1]] - yd[ik[0]], y31 = yd[ik[2]] - yd[ik[0]];
        double det = x21 * y31 - x31 * y21;
        if (det <= 1e-12) return 26;
        double a = (x31 * x31 + y31 * y31) / det;
        double b = -(x31 * x21 + y31 * y21) / det;
        double c = (x21 * x21 + y21 * y21) / det;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                int t = j * 3 + i;
                elmas[ik[j] * 5 + ik[i]] += (a * s1[t] + b * s2[t] + c * s3[t]) / 2.0;
                elmam[ik[j] * 5 + ik[i]] += det * s4[t] / 2.4e1;
            }
    }
    for (int i = 0; i < 25; ++i) elmas[i] = elmas[i] + kwert * kwert * elmam[i];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            kel[j * ld + i] = elmas[j * 5 + i] - elmas[4 * 5 + i] * elmas[j * 5 + 4] / elmas[4 * 5 + 4];
    return 0;
}

/* linear boundary segment mass matrix (src/elem1.f90:50-64) */
static void elem1(const double *xk, const double *yk, double *elmam, int ld) {
    double x21 = xk[1] - xk[0], y21 = yk[1] - yk[0];
    double l = sqrt(x21 * x21 + y21 * y21);
    static const int s4[4] = {2, 1, 1, 2};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) elmam[j * ld + i] = l * s4[j * 2 + i] / 6.0;
}

/* mixed-boundary factor beta(l,k) (src/beta.f90:69-121) */
static double beta_fn(const double *xk, const double *yk, double xs, double ys, double sytop,
                      double kw, int *err) {
    double xr = (xk[0] + xk[1]) / 2.0, yr = (yk[0] + yk[1]) / 2.0;
    double x3 = xr - xs;
    double y3m = (yr - sytop) - (ys - sytop);
    double y3p = (yr - sytop) + (ys - sytop);
    double r1m = sqrt(x3 * x3 + y3m * y3m), r1p = sqrt(x3 * x3 + y3p * y3p);
    double x4 = yk[0] - yk[1], y4 = xk[1] - xk[0];
    double r2 = sqrt(x4 * x4 + y4 * y4);
    *err = 0;
    if (r1m < 1e-12 || r1p < 1e-12 || r2 < 1e-12) { *err = 35; return 0.0; }
    double cm = (x3 * x4 + y3m * y4) / r1m / r2;
    double cp = (x3 * x4 + y3p * y4) / r1p / r2;
    double bk0m = orc_bessk0(r1m * kw), bk0p = orc_bessk0(r1p * kw);
    double bk1m = orc_bessk1(r1m * kw), bk1p = orc_bessk1(r1p * kw);
    double b;
    if (bk0m + bk0p == 0.0) b = (cm + cp) / 2.0;
    else b = (cm * bk1m + cp * bk1p) / (bk0m + bk0p);
    return kw * b;
}

/* mean height of the no-flow boundary (src/bsytop.f90:23-66) */
double orc_bsytop(const orc_grid *g) {
    double sytop = 0.0;
    int iel = 0;
    for (int t = 0; t < g->typanz; ++t) {
        if (g->typ[t] == 12) {
            int nkel = g->selanz[t];
            sytop = 0.0;
            for (int j = 1; j <= g->nelanz[t]; ++j, ++iel) {
                double sp = 0.0;
                for (int ik = 0; ik < nkel; ++ik) sp += g->y[g->nrel[iel * g->maxn + ik] - 1];
                sp = sp / (double)nkel;
                if (j == 1) sytop = sp;
                else sytop = (sytop * (double)(j - 1) + sp) / (double)j;
            }
        } else {
            iel += g->nelanz[t];
        }
    }
    return sytop;
}

/* src/precal.f90:79-215 */
int orc_precal(const orc_grid *g, int kwnanz, const double *kwn, int lsytop,
               double *elbg, double *relbg, double *kg, int *lbeta, double *sytop_out) {
    const int maxn = g->maxn, npk = NPK(maxn), elanz = g->elanz, relanz = g->relanz;
    double xk[8], yk[8], elmas[64], elmam[64];
    double sytop = lsytop ? orc_bsytop(g) : 0.0;
    if (sytop_out) *sytop_out = sytop;
    int lb = 0, lrandb = 0, iel = 0;
    for (int t = 0; t < g->typanz; ++t) {
        int ntyp = g->typ[t], nkel = g->selanz[t];
        for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
            for (int m = 0; m < nkel; ++m) {
                int nd = g->nrel[iel * maxn + m] - 1;
                xk[m] = g->x[nd]; yk[m] = g->y[nd];
            }
            if (ntyp == 11) {
                if (lrandb) return 101;
                lb = 1;
                elem1(xk, yk, elmam, 8);
                int rel = iel - elanz;
                if (rel < 0) return 36;
                int imn = 0;
                for (int m = 0; m < nkel; ++m)
                    for (int n = 0; n <= m; ++n, ++imn) relbg[(size_t)imn * relanz + rel] = elmam[n * 8 + m];
                for (int l = 0; l < g->eanz; ++l) {
                    int nd = g->enr[l] - 1;
                    for (int k = 0; k < kwnanz; ++k) {
                        int err;
                        double bt = beta_fn(xk, yk, g->x[nd], g->y[nd], sytop, kwn[k], &err);
                        if (err) return err;
                        kg[((size_t)k * g->eanz + l) * relanz + rel] = bt;
                    }
                }
            } else if (ntyp == 12) {
                continue;
            } else if (ntyp == 13) {
                if (lb) return 101;
                lrandb = 1;
                continue;
            } else if (ntyp == 8) {
                for (int k = 0; k < kwnanz; ++k) {
                    int err = elem8(xk, yk, kwn[k], elmas, 8);
                    if (err) return err;
                    int imn = 0;
                    for (int m = 0; m < nkel; ++m)
                        for (int n = 0; n <= m; ++n, ++imn)
                            elbg[((size_t)k * npk + imn) * elanz + iel] = elmas[n * 8 + m];
                }
            } else {
                int err;
                if (ntyp == 3) err = elem3(xk, yk, elmas, elmam, 8);
                else if (ntyp == 5) err = elem5(xk, yk, elmas, elmam, 8);
                else return 18;
                if (err) return err;
                for (int k = 0; k < kwnanz; ++k) {
                    int imn = 0;
                    for (int m = 0; m < nkel; ++m)
                        for (int n = 0; n <= m; ++n, ++imn)
                            elbg[((size_t)k * npk + imn) * elanz + iel] =
                                elmas[n * 8 + m] + elmam[n * 8 + m] * kwn[k] * kwn[k];
                }
            }
        }
    }
    *lbeta = lb;  /* lrandb: the forward loops apply randb / randdc (has_type13) */
    return 0;
}

/* ---------------------------------------------------------- band assembly */
/* src/kompab.f90:80-138,162 (lsr off, no electrode capacitances) */
int orc_kompab(const orc_grid *g, const zdbl *sigma, const double *elbg, const double *relbg,
               const double *kg, int kwnanz, int l, int k, zdbl *a, zdbl *b) {
    const int maxn = g->maxn, npk = NPK(maxn), mb = g->mb, sanz = g->sanz;
    (void)kwnanz;
    memset(a, 0, sizeof(zdbl) * (size_t)(mb + 1) * sanz);
    memset(b, 0, sizeof(zdbl) * (size_t)sanz);
    int iel = 0;
    for (int t = 0; t < g->typanz; ++t) {
        int ntyp = g->typ[t], nkel = g->selanz[t];
        for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
            if (ntyp > 11) continue;
            int ikl = 0;
            for (int kk = 0; kk < nkel; ++kk) {
                int nzp = g->nrel[iel * maxn + kk];
                for (int ll = 0; ll <= kk; ++ll) {
                    int nnp = g->nrel[iel * maxn + ll];
                    if (abs(nzp - nnp) > mb) return 19;
                    int im = imax_(nzp, nnp) * mb + imin_(nzp, nnp);
                    double dum;
                    zdbl s;
                    if (ntyp == 11) {
                        int rel = iel - g->elanz;
                        dum = relbg[(size_t)ikl * g->relanz + rel] *
                              kg[((size_t)(k - 1) * g->eanz + (l - 1)) * g->relanz + rel];
                        s = sigma[g->rnr[rel] - 1];
                    } else {
                        dum = elbg[((size_t)(k - 1) * npk + ikl) * g->elanz + iel];
                        s = sigma[iel];
                    }
                    a[im - 1] += CMPLX(dum * creal(s), dum * cimag(s));
                    ++ikl;
                }
            }
        }
    }
    b[g->enr[l - 1] - 1] = CMPLX(-1.0, 0.0);
    return 0;
}

/* src/scalab.f90:52-76 */
int orc_scalab(int sanz, int mb, zdbl *a, zdbl *b, double *fak) {
    for (int i = 1; i <= sanz; ++i) {
        int idi = i * (mb + 1);
        double dum = cabs(a[idi - 1]);
        if (dum <= 0.0) return 27;
        a[idi - 1] = CMPLX(creal(a[idi - 1]) / dum, cimag(a[idi - 1]) / dum);
        fak[i - 1] = 1.0 / sqrt(dum);
        b[i - 1] = b[i - 1] * fak[i - 1];
        if (i == 1) continue;
        int i0 = i * mb, ja = imax_(1, i - mb);
        for (int j = ja; j <= i - 1; ++j) a[i0 + j - 1] = a[i0 + j - 1] * (fak[i - 1] * fak[j - 1]);
    }
    return 0;
}

/* complex-symmetric band Cholesky, row oriented, no conjugation (src/chol.f90:46-78) */
int orc_chol(int sanz, int mb, zdbl *a) {
    const int m1 = mb + 1;
    for (int i = 1; i <= sanz; ++i) {
        int idi = i * m1, fi = imax_(1, i - mb), i0 = idi - i;
        for (int j = fi; j <= i; ++j) {
            int ij = i0 + j, j0 = j * mb;
            zdbl s = a[ij - 1];
            for (int k = fi; k <= j - 1; ++k) s -= a[i0 + k - 1] * a[j0 + k - 1];
            if (j < i) {
                a[ij - 1] = s / a[j * m1 - 1];
            } else {
                if (cabs(s) <= 0.0) return 28;
                a[idi - 1] = csqrt(s);
            }
        }
    }
    return 0;
}

/* forward/back substitution, pot = -A^{-1} b (src/vre.f90:45-69) */
void orc_vre(int sanz, int mb, const zdbl *a, const zdbl *b, zdbl *pot) {
    const int m1 = mb + 1;
    int idi = 0;
    for (int i = 1; i <= sanz; ++i) {
        idi = i * m1;
        zdbl s = b[i - 1];
        int i0 = idi - i, jlow = imax_(1, i - mb);
        for (int j = jlow; j <= i - 1; ++j) s -= a[i0 + j - 1] * pot[j - 1];
        pot[i - 1] = s / a[idi - 1];
    }
    for (int i = sanz; i >= 1; --i) {
        pot[i - 1] = -pot[i - 1] / a[idi - 1];
        int jlow = imax_(1, i - mb), i0 = idi - i;
        for (int j = jlow; j <= i - 1; ++j) pot[j - 1] += a[i0 + j - 1] * pot[i - 1];
        idi -= m1;
    }
}

/* src/kompb.f90:39-66 */
void orc_kompb(int sanz, int enode, zdbl *b, const double *fak) {
    for (int i = 0; i < sanz; ++i) b[i] = 0.0;
    b[enode - 1] = CMPLX(-fak[enode - 1], 0.0);
}

/* dense complex-symmetric Cholesky: strict lower L in a(j,i) j>i (column-major
 * a(n,n), i.e. a[(i-1)*n + (j-1)]), diagonal in p (src/cholz.f90:43-76). */
int orc_cholz(zdbl *a, zdbl *p, int n) {
    for (int i = 1; i <= n; ++i)
        for (int j = i; j <= n; ++j) {
            zdbl s = a[(size_t)(j - 1) * n + (i - 1)]; /* a(i,j) */
            for (int k = i - 1; k >= 1; --k) s -= a[(size_t)(k - 1) * n + (i - 1)] * a[(size_t)(k - 1) * n + (j - 1)];
            if (i == j) {
                if (cabs(s) <= 0.0) return 28;
                p[i - 1] = csqrt(s);
            } else {
                a[(size_t)(i - 1) * n + (j - 1)] = s / p[i - 1]; /* a(j,i) */
            }
        }
    return 0;
}

/* forward loop over wavenumbers and electrodes (src/fem.f90:336-397) */
/* ------------------------------------------------------- Dirichlet boundaries */
/* Prescribed boundary values of boundary.dat (src/rrandb.f90: rwdanz, rwdnr, rwd; the
 * DC twins use the real parts, rwddc), applied by randb2 / randbdc2 when n > 0. */
static int g_rwdanz = 0;
static const int *g_rwdnr = NULL;
static const zdbl *g_rwd = NULL;

void orc_set_boundary(int n, const int *nodes, const zdbl *values) {
    g_rwdanz = n > 0 ? n : 0;
    g_rwdnr = nodes;
    g_rwd = values;
}

static int has_type13(const orc_grid *g) {
    for (int t = 0; t < g->typanz; ++t)
        if (g->typ[t] == 13) return 1;
    return 0;
}

/* src/randb.f90:40-84: every node k of a type-13 element gets b(k) = 0, a(k,k) = 1 and
 * the rest of its row and column zero.  T is zdbl (randb) or double (randdc). */
#define RANDB_BODY(T)                                                                      \
    const int sanz = g->sanz, mb = g->mb, m1 = mb + 1;                                     \
    int iel = 0;                                                                           \
    for (int t = 0; t < g->typanz; ++t) {                                                  \
        const int ntyp = g->typ[t], nkel = g->selanz[t];                                   \
        for (int j2 = 0; j2 < g->nelanz[t]; ++j2, ++iel) {                                 \
            if (ntyp != 13) continue;                                                      \
            for (int ir = 0; ir < nkel; ++ir) {                                            \
                const int k = g->nrel[(size_t)iel * g->maxn + ir];                         \
                b[k - 1] = (T)0.0;                                                         \
                const int idk = k * m1;                                                    \
                a[idk - 1] = (T)1.0;                                                       \
                if (k != 1) {                                                              \
                    const int ia = imax_(1, mb + 2 - k);                                   \
                    for (int i = ia; i <= mb; ++i) a[idk + i - m1 - 1] = (T)0.0;           \
                }                                                                          \
                if (k == sanz) continue;                                                   \
                const int ia = imax_(1, k - sanz + m1);                                    \
                for (int i = ia; i <= mb; ++i) {                                           \
                    const int jj = k - i + m1;                                             \
                    a[(jj - 1) * m1 + i - 1] = (T)0.0;                                     \
                }                                                                          \
            }                                                                              \
        }                                                                                  \
    }

void orc_randb(const orc_grid *g, zdbl *a, zdbl *b) { RANDB_BODY(zdbl) }
void orc_randdc(const orc_grid *g, double *a, double *b) { RANDB_BODY(double) }

/* src/randb2.f90:29-79 (randbdc2.f90 with the real values): node k of the list takes
 * the value w: b(k) = -w, a(k,k) = 1, and every other entry of row and column k moves
 * w * a(j,k) into b(j) before it is zeroed. */
#define RANDB2_BODY(T, VAL)                                                                \
    const int sanz = g->sanz, mb = g->mb, m1 = mb + 1;                                     \
    for (int ir = 0; ir < g_rwdanz; ++ir) {                                                \
        const int k = g_rwdnr[ir];                                                         \
        const T w = VAL(g_rwd[ir]);                                                        \
        b[k - 1] = -w;                                                                     \
        const int idk = k * m1;                                                            \
        a[idk - 1] = (T)1.0;                                                               \
        if (k != 1) {                                                                      \
            const int ia = imax_(1, mb + 2 - k);                                           \
            for (int i = ia; i <= mb; ++i) {                                               \
                const int jj = k + i - m1, ki = idk + i - m1;                              \
                b[jj - 1] = b[jj - 1] + w * a[ki - 1];                                     \
                a[ki - 1] = (T)0.0;                                                        \
            }                                                                              \
        }                                                                                  \
        if (k == sanz) continue;                                                           \
        const int ia = imax_(1, k - sanz + m1);                                            \
        for (int i = ia; i <= mb; ++i) {                                                   \
            const int jj = k - i + m1, ji = (jj - 1) * m1 + i;                             \
            b[jj - 1] = b[jj - 1] + w * a[ji - 1];                                         \
            a[ji - 1] = (T)0.0;                                                            \
        }                                                                                  \
    }

#define ZVAL(z) (z)
#define RVAL(z) creal(z)
void orc_randb2(const orc_grid *g, zdbl *a, zdbl *b) { RANDB2_BODY(zdbl, ZVAL) }
void orc_randbdc2(const orc_grid *g, double *a, double *b) { RANDB2_BODY(double, RVAL) }

/* ------------------------------------------------ singularity removal, sink */
/* src/refsig.f90:33-70: sigma0 = exp(sum_e area_e log sigma_e / sum_e area_e) over the
 * area elements (types 3, 5, 8), areas by Gauss' formula */
void orc_refsig(const orc_grid *g, const zdbl *sigma, zdbl *sigma0) {
    const int maxn = g->maxn;
    double area = 0.0;
    zdbl s0 = 0.0;
    int iel = 0;
    for (int t = 0; t < g->typanz; ++t) {
        for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
            const int ty = g->typ[t];
            if (ty > 10) continue;
            const int *nd = g->nrel + (size_t)iel * maxn;
            const double xk = g->x[nd[0] - 1], yk = g->y[nd[0] - 1];
            const double ax = g->x[nd[1] - 1] - xk, ay = g->y[nd[1] - 1] - yk;
            const double bx = g->x[nd[2] - 1] - xk, by = g->y[nd[2] - 1] - yk;
            double dum = 0.0;
            if (ty == 3) {
                dum = fabs(ax * by - ay * bx) / 2.0;
            } else if (ty == 5 || ty == 8) {
                const double cx = g->x[nd[3] - 1] - xk, cy = g->y[nd[3] - 1] - yk;
                dum = (fabs(ax * by - ay * bx) + fabs(bx * cy - by * cx)) / 2.0;
            }
            s0 += dum * clog(sigma[iel]);
            area += dum;
        }
    }
    *sigma0 = cexp(s0 / area);
}

/* src/potana.f90:51-97: the homogeneous half-space potential of a unit current at
 * electrode l for wavenumber k, K0 of the distances to the source and its mirror image
 * in y = 0, the source node set to 5 times the largest value, scaled by (1/2pi)/sigma0 */
void orc_potana(const orc_grid *g, const double *kwn, int l, int k, const zdbl *sigma0p, zdbl *pota) {
    const double pi = acos(-1.0), kk = kwn[k - 1];
    const zdbl sigma0 = *sigma0p;
    const int idum = g->enr[l - 1];
    const double xk1 = g->x[idum - 1], yk1 = g->y[idum - 1];
    double potmax = 0.0;
    for (int j = 1; j <= g->sanz; ++j) {
        if (j == idum) continue;
        const double x21 = g->x[j - 1] - xk1, y21m = g->y[j - 1] - yk1, y21p = g->y[j - 1] + yk1;
        const double rm = sqrt(x21 * x21 + y21m * y21m), rp = sqrt(x21 * x21 + y21p * y21p);
        const double dum = orc_bessk0(rm * kk) + orc_bessk0(rp * kk);
        potmax = fmax(potmax, dum);
        pota[j - 1] = dum;
    }
    pota[idum - 1] = 5.0 * potmax;
    const zdbl dum2 = (0.5 / pi) / sigma0;
    for (int j = 0; j < g->sanz; ++j) pota[j] *= dum2;
}

/* src/kompab.f90:80-183 with the singularity removal (lsr: the right-hand side
 * sum_e elbg (sigma_e - sigma0) pota, no unit source, :129-133,162) and the fictitious
 * sink (lsink: b(nsink) = 1, :183) */
static int kompab_sw(const orc_grid *g, const zdbl *sigma, const double *elbg, const double *relbg,
                     const double *kg, int l, int k, int lsr, zdbl sigma0, const zdbl *pota, int lsink,
                     int nsink, zdbl *a, zdbl *b) {
    const int maxn = g->maxn, npk = NPK(maxn), mb = g->mb, sanz = g->sanz;
    memset(a, 0, sizeof(zdbl) * (size_t)(mb + 1) * sanz);
    memset(b, 0, sizeof(zdbl) * (size_t)sanz);
    int iel = 0;
    for (int t = 0; t < g->typanz; ++t) {
        int ntyp = g->typ[t], nkel = g->selanz[t];
        for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
            if (ntyp > 11) continue;
            int ikl = 0;
            for (int kk = 0; kk < nkel; ++kk) {
                int nzp = g->nrel[iel * maxn + kk];
                for (int ll = 0; ll <= kk; ++ll) {
                    int nnp = g->nrel[iel * maxn + ll];
                    if (abs(nzp - nnp) > mb) return 19;
                    int im = imax_(nzp, nnp) * mb + imin_(nzp, nnp);
                    double dum;
                    zdbl s;
                    if (ntyp == 11) {
                        int rel = iel - g->elanz;
                        dum = relbg[(size_t)ikl * g->relanz + rel] *
                              kg[((size_t)(k - 1) * g->eanz + (l - 1)) * g->relanz + rel];
                        s = sigma[g->rnr[rel] - 1];
                    } else {
                        dum = elbg[((size_t)(k - 1) * npk + ikl) * g->elanz + iel];
                        s = sigma[iel];
                    }
                    a[im - 1] += CMPLX(dum * creal(s), dum * cimag(s));
                    if (lsr) {
                        const zdbl d2 = dum * (s - sigma0);
                        b[nzp - 1] += d2 * pota[nnp - 1];
                        if (nnp != nzp) b[nnp - 1] += d2 * pota[nzp - 1];
                    }
                    ++ikl;
                }
            }
        }
    }
    if (!lsr) b[g->enr[l - 1] - 1] = CMPLX(-1.0, 0.0);
    if (lsink) b[nsink - 1] = CMPLX(1.0, 0.0);
    return 0;
}

/* The forward loop of src/fem.f90:302-397 with its switches: refsig, lsr forced off
 * without mixed boundaries (:311); per (l, k) potana when lsr or lana (:350), the
 * assembly (kompab, randb, randb2, scalab, chol) when lbeta, l = 1 or lsr and not lana
 * (:355-370), else kompb with the sink (src/kompb.f90:49-58); kpot = pota (lana), or
 * pot fak (+ pota with lsr) (:383-391).  Errors as orc_forward. */
int orc_forward_sw(const orc_grid *g, const zdbl *sigma, int kwnanz, const double *kwn, const double *elbg,
                   const double *relbg, const double *kg, int lbeta, int lsr, int lana, int lsink, int nsink,
                   zdbl *kpot) {
    const int sanz = g->sanz, mb = g->mb, eanz = g->eanz;
    zdbl *a = (zdbl *)malloc(sizeof(zdbl) * (size_t)(mb + 1) * sanz);
    zdbl *b = (zdbl *)malloc(sizeof(zdbl) * (size_t)sanz);
    zdbl *pot = (zdbl *)malloc(sizeof(zdbl) * (size_t)sanz);
    zdbl *pota = (zdbl *)malloc(sizeof(zdbl) * (size_t)sanz);
    double *fak = (double *)malloc(sizeof(double) * (size_t)sanz);
    zdbl sigma0;
    orc_refsig(g, sigma, &sigma0);
    if (!lbeta) lsr = 0;
    if (lsr) lana = 0; /* src/fem.f90:198-203 */
    int err = 0;
    for (int k = 1; k <= kwnanz && !err; ++k) {
        for (int l = 1; l <= eanz && !err; ++l) {
            if (lsr || lana) orc_potana(g, kwn, l, k, &sigma0, pota);
            if ((lbeta || l == 1 || lsr) && !lana) {
                err = kompab_sw(g, sigma, elbg, relbg, kg, l, k, lsr, sigma0, pota, lsink, nsink, a, b);
                if (!err && has_type13(g)) orc_randb(g, a, b);
                if (!err && g_rwdanz) orc_randb2(g, a, b);
                if (!err) err = orc_scalab(sanz, mb, a, b, fak);
                if (!err) err = orc_chol(sanz, mb, a);
            } else if (!lana) {
                orc_kompb(sanz, g->enr[l - 1], b, fak);
                if (lsink) b[nsink - 1] = fak[nsink - 1];
            }
            if (err) break;
            if (!lana) orc_vre(sanz, mb, a, b, pot);
            zdbl *kp = kpot + ((size_t)(k - 1) * eanz + (l - 1)) * sanz;
            for (int j = 0; j < sanz; ++j) {
                if (lana) kp[j] = pota[j];
                else kp[j] = pot[j] * fak[j] + (lsr ? pota[j] : 0.0);
            }
        }
    }
    free(a); free(b); free(pot); free(pota); free(fak);
    return err;
}

/* src/bkfak.f90:48-136: the half-space geometric factor of each measurement from the
 * electrode positions and their mirror images in y = 0, kfak = 4 pi / ((r1 - r2) - (r3 -
 * r4)); errnr 93 when the denominator vanishes */
int orc_bkfak(const orc_grid *g, int nanz, const int *strnr, const int *vnr, double *kfak) {
    const double pi = acos(-1.0);
    double xk[4] = {0, 0, 0, 0}, yk[4] = {0, 0, 0, 0};
    for (int i = 0; i < nanz; ++i) {
        const int e[4] = {strnr[i] % 10000, strnr[i] / 10000, vnr[i] % 10000, vnr[i] / 10000};
        for (int m = 0; m < 4; ++m)
            if (e[m] > 0) {
                xk[m] = g->x[g->enr[e[m] - 1] - 1];
                yk[m] = g->y[g->enr[e[m] - 1] - 1];
            }
        double r[4];
        /* r1: electrodes (4, 2), r2: (4, 1), r3: (3, 2), r4: (3, 1), 0-based below */
        const int pu[4] = {3, 3, 2, 2}, pv[4] = {1, 0, 1, 0};
        for (int m = 0; m < 4; ++m) {
            const int u = pu[m], v = pv[m];
            if (e[u] > 0 && e[v] > 0) {
                const double dx = xk[u] - xk[v], dym = yk[u] - yk[v], dyp = yk[u] + yk[v];
                r[m] = 1.0 / sqrt(dx * dx + dym * dym) + 1.0 / sqrt(dx * dx + dyp * dyp);
            } else {
                r[m] = 0.0;
            }
        }
        const double dum = (r[0] - r[1]) - (r[2] - r[3]);
        if (fabs(dum) == 0.0) return 93;
        kfak[i] = 4.0 * pi / dum;
    }
    return 0;
}

int orc_forward(const orc_grid *g, const zdbl *sigma, int kwnanz, const double *elbg,
                const double *relbg, const double *kg, int lbeta, zdbl *kpot,
                int nsys, const int *sysl, const int *sysk) {
    const int sanz = g->sanz, mb = g->mb, eanz = g->eanz;
    zdbl *a = (zdbl *)malloc(sizeof(zdbl) * (size_t)(mb + 1) * sanz);
    zdbl *b = (zdbl *)malloc(sizeof(zdbl) * (size_t)sanz);
    zdbl *pot = (zdbl *)malloc(sizeof(zdbl) * (size_t)sanz);
    double *fak = (double *)malloc(sizeof(double) * (size_t)sanz);
    int err = 0;
    int ntot = nsys > 0 ? nsys : kwnanz * eanz;
    int lastk = -1;
    for (int s = 0; s < ntot && !err; ++s) {
        int k, l;
        if (nsys > 0) { k = sysk[s]; l = sysl[s]; }
        else { k = s / eanz + 1; l = s % eanz + 1; }
        if (lbeta || l == 1 || k != lastk) {
            err = orc_kompab(g, sigma, elbg, relbg, kg, kwnanz, l, k, a, b);
            /* src/fem.f90:362-363 */
            if (!err && has_type13(g)) orc_randb(g, a, b);
            if (!err && g_rwdanz) orc_randb2(g, a, b);
            if (!err) err = orc_scalab(sanz, mb, a, b, fak);
            if (!err) err = orc_chol(sanz, mb, a);
            if (!err && !lbeta && l != 1) orc_kompb(sanz, g->enr[l - 1], b, fak);
        } else {
            orc_kompb(sanz, g->enr[l - 1], b, fak);
        }
        lastk = k;
        if (err) break;
        orc_vre(sanz, mb, a, b, pot);
        zdbl *kp = kpot + ((size_t)(k - 1) * eanz + (l - 1)) * sanz;
        for (int j = 0; j < sanz; ++j) kp[j] = pot[j] * fak[j];
    }
    free(a); free(b); free(pot); free(fak);
    return err;
}

/* src/rtrafo.f90:73-88 (swrtr = 1) */
void orc_rtrafo(int sanz, int eanz, int kwnanz, const double *kwnwi, const zdbl *kpot, zdbl *hpot) {
    const double pi = acos(-1.0);
    for (int l = 0; l < eanz; ++l)
        for (int j = 0; j < sanz; ++j) {
            zdbl s = 0.0;
            for (int k = 0; k < kwnanz; ++k) s += kpot[((size_t)k * eanz + l) * sanz + j] * kwnwi[k];
            hpot[(size_t)l * sanz + j] = s / pi;
        }
}

static zdbl volt_of(int strnr, int vnr, const int *enr, int sanz, const zdbl *hpot) {
    int e1 = strnr % 10000, e2 = (strnr - e1) / 10000;
    int e3 = vnr % 10000, e4 = (vnr - e3) / 10000;
#define HP(e, c) (((e) > 0 && (c) > 0) ? hpot[(size_t)((c)-1) * sanz + enr[(e)-1] - 1] : (zdbl)0.0)
    zdbl d1 = HP(e4, e2), d2 = HP(e4, e1), d3 = HP(e3, e2), d4 = HP(e3, e1);
#undef HP
    return (d1 - d2) - (d3 - d4);
}

/* src/bvolt.f90:41-79 */
void orc_bvolt(int nanz, const int *strnr, const int *vnr, const int *enr, int sanz,
               const zdbl *hpot, const double *kfak, zdbl *volt, zdbl *sigmaa) {
    for (int i = 0; i < nanz; ++i) {
        volt[i] = volt_of(strnr[i], vnr[i], enr, sanz, hpot);
        sigmaa[i] = (kfak ? kfak[i] : 1.0) * volt[i];
    }
}

/* src/bvolti.f90:43-89: sigmaa = -log(V), errnr 82 on V == 0 */
int orc_bvolti(int nanz, const int *strnr, const int *vnr, const int *enr, int sanz,
               const zdbl *hpot, zdbl *volt, zdbl *sigmaa) {
    for (int i = 0; i < nanz; ++i) {
        volt[i] = volt_of(strnr[i], vnr[i], enr, sanz, hpot);
        if (cabs(volt[i]) == 0.0) return 82;
        sigmaa[i] = -clog(volt[i]);
    }
    return 0;
}

/* per-element reciprocity kernel shared by bsens/bsensi (src/bsensi.f90:134-176) */
static zdbl sens_elem(const orc_grid *g, int iel, int nkel, int e1, int e2, int e3, int e4,
                      int kwnanz, const double *kwnwi, const zdbl *kpot, const double *elbg) {
    const int sanz = g->sanz, eanz = g->eanz, npk = NPK(g->maxn);
    zdbl dum = 0.0;
    for (int k = 0; k < kwnanz; ++k) {
        zdbl h = 0.0, sup[4] = {0, 0, 0, 0};
        const zdbl *kp = kpot + (size_t)k * eanz * sanz;
        for (int mi = 0; mi < nkel; ++mi) {
            int nzp = g->nrel[iel * g->maxn + mi];
            for (int mj = 0; mj < nkel; ++mj) {
                int nnp = g->nrel[iel * g->maxn + mj];
                int ia = imax_(mi, mj) + 1, ib = imin_(mi, mj) + 1;
                int imn = ia * (ia - 1) / 2 + ib;
                if (e1 > 0) sup[0] = kp[(size_t)(e1 - 1) * sanz + nnp - 1];
                if (e2 > 0) sup[1] = kp[(size_t)(e2 - 1) * sanz + nnp - 1];
                if (e3 > 0) sup[2] = kp[(size_t)(e3 - 1) * sanz + nzp - 1];
                if (e4 > 0) sup[3] = kp[(size_t)(e4 - 1) * sanz + nzp - 1];
                zdbl d = (sup[1] - sup[0]) * (sup[3] - sup[2]);
                h += elbg[((size_t)k * npk + imn - 1) * g->elanz + iel] * d;
            }
        }
        dum += h * kwnwi[k];
    }
    return dum;
}

/* src/bsensi.f90:98-187 (swrtr = 1, mnr = identity) */
void orc_bsensi(const orc_grid *g, int nanz, const int *strnr, const int *vnr,
                int kwnanz, const double *kwnwi, const zdbl *kpot, const double *elbg,
                const zdbl *sigma, const zdbl *volt, zdbl *sens) {
    const double pi = acos(-1.0);
    for (int i = 0; i < nanz; ++i) {
        int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
        int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
        int iel = 0;
        for (int t = 0; t < g->typanz; ++t) {
            if (g->typ[t] > 10) continue;
            for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
                zdbl dum = sens_elem(g, iel, g->selanz[t], e1, e2, e3, e4, kwnanz, kwnwi, kpot, elbg);
                dum = dum / pi;
                sens[(size_t)iel * nanz + i] = (dum * sigma[iel]) / volt[i];
            }
        }
    }
}

/* src/bsens.f90:103-191 (swrtr = 1, mnr = identity): sens = -dV/dsigma */
void orc_bsens(const orc_grid *g, int nanz, const int *strnr, const int *vnr, const double *strom,
               int kwnanz, const double *kwnwi, const zdbl *kpot, const double *elbg, zdbl *sens) {
    const double pi = acos(-1.0);
    for (int i = 0; i < nanz; ++i) {
        int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
        int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
        int iel = 0;
        for (int t = 0; t < g->typanz; ++t) {
            if (g->typ[t] > 10) continue;
            for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
                zdbl dum = sens_elem(g, iel, g->selanz[t], e1, e2, e3, e4, kwnanz, kwnwi, kpot, elbg);
                dum = dum * (strom[i] / pi);
                sens[(size_t)iel * nanz + i] = 0.0 - dum;
            }
        }
    }
}

/* ------------------------------------------------------------ regularisation */
static double centroid(const orc_grid *g, int iel, int nkel, int axis) {
    double s = 0.0;
    for (int l = 0; l < nkel; ++l) {
        int nd = g->nrel[iel * g->maxn + l] - 1;
        s += axis ? g->y[nd] : g->x[nd];
    }
    return s;
}

/* smooth regularisation on a regular nx x nz quad grid (src/bsmatm_mod.f90:190-362,
 * no discontinuities: ndis_x = ndis_z = 0).  smatm (manz,3) column-major. */
void orc_bsmatmreg(const orc_grid *g, int nx, int nz, double alfx, double alfz, double *smatm) {
    const int manz = g->elanz;
    memset(smatm, 0, sizeof(double) * 3 * (size_t)manz);
#define KK(i, j) (((i)-1) * nx + (j))
#define NY(e, l) g->y[g->nrel[((e)-1) * g->maxn + (l)-1] - 1]
#define NXC(e, l) g->x[g->nrel[((e)-1) * g->maxn + (l)-1] - 1]
    for (int i = 1; i <= nz; ++i) {
        for (int j = 1; j <= nx; ++j) {
            int kk = KK(i, j);
            double dzleft = fabs(NY(kk, 4) - NY(kk, 1));
            double dzright = fabs(NY(kk, 3) - NY(kk, 2));
            double xmean = centroid(g, kk - 1, 4, 0) / 4.0;
            if (j > 1) {
                double xleft = centroid(g, KK(i, j - 1) - 1, 4, 0) / 4.0;
                smatm[kk - 1] = 1.0 * alfx * dzleft / fabs(xmean - xleft);
            }
            if (j < nx) {
                double xright = centroid(g, KK(i, j + 1) - 1, 4, 0) / 4.0;
                double dum = 1.0 * alfx * dzright / fabs(xright - xmean);
                smatm[kk - 1] += dum;
                smatm[manz + kk - 1] = -dum;
            }
            double dxup = fabs(NXC(kk, 3) - NXC(kk, 4));
            double dxdown = fabs(NXC(kk, 2) - NXC(kk, 1));
            double zmean = centroid(g, kk - 1, 4, 1) / 4.0;
            if (i > 1) {
                double zup = centroid(g, KK(i - 1, j) - 1, 4, 1) / 4.0;
                double dum = 1.0 * alfz * dxup / fabs(zup - zmean);
                smatm[kk - 1] += dum;
            }
            if (i < nz) {
                double zdown = centroid(g, KK(i + 1, j) - 1, 4, 1) / 4.0;
                double dum = 1.0 * alfz * dxdown / fabs(zmean - zdown);
                smatm[kk - 1] += dum;
                smatm[2 * (size_t)manz + kk - 1] = -dum;
            }
        }
    }
#undef KK
#undef NY
#undef NXC
}

/* edge neighbours of area elements (src/bnachbar.f90:44-90).  nachbar (manz, maxn+1)
 * column-major, last column = neighbour count. */
void orc_bnachbar(const orc_grid *g, int *nachbar) {
    const int elanz = g->elanz, maxn = g->maxn;
    memset(nachbar, 0, sizeof(int) * (size_t)elanz * (maxn + 1));
    for (int i = 0; i < elanz; ++i)
        for (int ik = 0; ik < maxn; ++ik) {
            int ik1 = g->nrel[i * maxn + ik], ik2 = g->nrel[i * maxn + (ik + 1) % maxn];
            for (int j = 0; j < elanz; ++j) {
                if (j == i) continue;
                for (int jk = 0; jk < maxn; ++jk) {
                    int jk1 = g->nrel[j * maxn + jk], jk2 = g->nrel[j * maxn + (jk + 1) % maxn];
                    if ((ik1 == jk1 && ik2 == jk2) || (ik1 == jk2 && ik2 == jk1)) {
                        nachbar[(size_t)ik * elanz + i] = j + 1;
                        nachbar[(size_t)maxn * elanz + i] += 1;
                    }
                }
            }
        }
}

/* triangular/unstructured smoothness (src/bsmatm_mod.f90:364-460, no decoupling).
 * Element centroids espx/espy as in src/relem.f90:128-138. smatm (manz, maxn+1). */
void orc_bsmatmtri(const orc_grid *g, const int *nachbar, double alfx, double alfz, double *smatm) {
    const int elanz = g->elanz, maxn = g->maxn;
    double *espx = (double *)malloc(sizeof(double) * elanz), *espy = (double *)malloc(sizeof(double) * elanz);
    int iel = 0;
    for (int t = 0; t < g->typanz; ++t) {
        if (g->typ[t] >= 10) continue;
        for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
            double sx_ = 0.0, sy_ = 0.0;
            for (int k = 0; k < g->selanz[t]; ++k) {
                int nd = g->nrel[iel * maxn + k] - 1;
                sx_ += g->x[nd]; sy_ += g->y[nd];
            }
            espx[iel] = sx_ / g->selanz[t];
            espy[iel] = sy_ / g->selanz[t];
        }
    }
    memset(smatm, 0, sizeof(double) * (size_t)elanz * (maxn + 1));
    for (int i = 0; i < elanz; ++i) {
        for (int k = 0; k < maxn; ++k) {
            int n1 = g->nrel[i * maxn + k] - 1, n2 = g->nrel[i * maxn + (k + 1) % maxn] - 1;
            double dx = g->x[n1] - g->x[n2], dy = g->y[n1] - g->y[n2];
            double edglen = sqrt(pow(dx, 2.0) + pow(dy, 2.0));
            int nb = nachbar[(size_t)k * elanz + i];
            if (nb > 0) {
                double ex = espx[i] - espx[nb - 1], ey = espy[i] - espy[nb - 1];
                double dist = sqrt(pow(ex, 2.0) + pow(ey, 2.0));
                double ang = atan2(ey, ex);
                double alfgeo = sqrt(pow(alfx * cos(ang), 2.0) + pow(alfz * sin(ang), 2.0));
                double rs = edglen / dist * alfgeo;
                smatm[(size_t)k * elanz + i] = -rs;
                smatm[(size_t)maxn * elanz + i] += rs;
            }
        }
    }
    free(espx); free(espy);
}

/* R p for the model stencil, with optional per-cell scaling c (NULL = 1):
 * ltri 0 -> src/update.f90:65-75 / src/cg_mod.f90:564-576;
 * ltri 1/2 -> src/update.f90:107-124 / src/cg_mod.f90:600-611 */
static void rmatvec(const orc_cg_problem *p, const double *c, const zdbl *v, zdbl *out) {
    const int manz = p->manz, nx = p->nx;
    const double *s = p->smatm;
    for (int i = 1; i <= manz; ++i) {
        zdbl cd = 0.0;
        if (p->ltri == 0) {
            if (c) {
                if (i > 1) cd = v[i - 2] * (s[manz + i - 2] * c[i - 2]);
                if (i < manz) cd += v[i] * (s[manz + i - 1] * c[i]);
                if (i > nx) cd += v[i - nx - 1] * (s[2 * manz + i - nx - 1] * c[i - nx - 1]);
                if (i < manz - nx + 1) cd += v[i + nx - 1] * (s[2 * manz + i - 1] * c[i + nx - 1]);
                out[i - 1] = cd + v[i - 1] * (s[i - 1] * c[i - 1]);
            } else {
                if (i > 1) cd = s[manz + i - 2] * v[i - 2];
                if (i < manz) cd += s[manz + i - 1] * v[i];
                if (i > nx) cd += s[2 * manz + i - nx - 1] * v[i - nx - 1];
                if (i < manz - nx + 1) cd += s[2 * manz + i - 1] * v[i + nx - 1];
                out[i - 1] = cd + s[i - 1] * v[i - 1];
            }
        } else {
            for (int ij = 0; ij < p->maxn; ++ij) {
                int in = p->nachbar[(size_t)ij * manz + i - 1];
                if (in != 0) {
                    if (c) cd += v[in - 1] * s[(size_t)ij * manz + i - 1] * c[in - 1];
                    else cd += s[(size_t)ij * manz + i - 1] * v[in - 1];
                }
            }
            double dg = s[(size_t)p->maxn * manz + i - 1];
            if (c) out[i - 1] = cd + v[i - 1] * dg * c[i - 1];
            else out[i - 1] = cd + dg * v[i - 1];
        }
    }
}

static double rdiag(const orc_cg_problem *p, int j) {
    return p->ltri == 0 ? p->smatm[j] : p->smatm[(size_t)p->maxn * p->manz + j];
}

/* RHS and Jacobi scaling (src/update.f90:63-136,179-231,236-293), complex case,
 * no prior, no reference model, no difference inversion. */
/* src/blam0.f90:101-168: jtj(j) = |sum_i sum_k conj(J(i,j)) J(i,k) w_i| (lfpi: real
 * parts, DABS), lammax = sum(jtj)/manz * 2/(alfx+alfz) * 5.  Same loop order as the
 * reference (j outer, i, k inner). */
double orc_blam0(const orc_cg_problem *p, int fpi, double alfx, double alfz) {
    const int nanz = p->nanz, manz = p->manz;
    const zdbl *J = p->sens;
    double *jtj = (double *)malloc(sizeof(double) * (size_t)manz);
    for (int j = 0; j < manz; ++j) {
        if (fpi) {
            double dum = 0.0;
            for (int i = 0; i < nanz; ++i) {
                double wi = p->wmatd[i] * (double)p->wdfak[i];
                double aij = creal(J[(size_t)j * nanz + i]);
                for (int k = 0; k < manz; ++k) dum = dum + aij * creal(J[(size_t)k * nanz + i]) * wi;
            }
            jtj[j] = fabs(dum);
        } else {
            zdbl cdum = 0.0;
            for (int i = 0; i < nanz; ++i) {
                zdbl wi = p->wmatd[i] * (double)p->wdfak[i];
                zdbl aij = conj(J[(size_t)j * nanz + i]);
                for (int k = 0; k < manz; ++k) cdum = cdum + aij * J[(size_t)k * nanz + i] * wi;
            }
            jtj[j] = cabs(cdum);
        }
    }
    double lammax = 0.0;
    for (int j = 0; j < manz; ++j) lammax += jtj[j];
    free(jtj);
    lammax = lammax / (double)manz;
    lammax = lammax * 2.0 / (alfx + alfz);
    lammax = lammax * 5.0;
    return lammax;
}

void orc_update_rhs(const orc_cg_problem *p, double *cgfac, zdbl *bvec) {
    const int nanz = p->nanz, manz = p->manz;
    rmatvec(p, NULL, p->par, bvec);
    for (int j = 0; j < manz; ++j) {
        const zdbl *col = p->sens + (size_t)j * nanz;
        double dum = 0.0;
        for (int i = 0; i < nanz; ++i) dum += creal(conj(col[i]) * col[i]) * p->wmatd[i] * (double)p->wdfak[i];
        dum = dum + p->lam * rdiag(p, j);
        cgfac[j] = 1.0 / sqrt(dum);
    }
    for (int j = 0; j < manz; ++j) {
        const zdbl *col = p->sens + (size_t)j * nanz;
        zdbl cd = 0.0;
        for (int i = 0; i < nanz; ++i) cd += conj(col[i]) * (p->wmatd[i] * (double)p->wdfak[i]) * (p->dat[i] - p->sigmaa[i]);
        bvec[j] = cd - p->lam * bvec[j];
        bvec[j] = bvec[j] * cgfac[j];
    }
}

/* ap = J (C p) (src/cg_mod.f90:537-543); out = C (J^H W ap + lam R C p)
 * (src/cg_mod.f90:564-576 / 600-611, 715-725) */
void orc_cg_matvec(const orc_cg_problem *p, const double *cgfac, const zdbl *pvec, zdbl *ap, zdbl *out) {
    const int nanz = p->nanz, manz = p->manz;
    for (int i = 0; i < nanz; ++i) ap[i] = 0.0;
    for (int i = 0; i < nanz; ++i) {
        zdbl s = 0.0;
        for (int j = 0; j < manz; ++j) s += pvec[j] * p->sens[(size_t)j * nanz + i] * cgfac[j];
        ap[i] = s;
    }
    rmatvec(p, cgfac, pvec, out);
    for (int j = 0; j < manz; ++j) {
        const zdbl *col = p->sens + (size_t)j * nanz;
        zdbl cd = 0.0;
        for (int i = 0; i < nanz; ++i) cd += conj(col[i]) * (p->wmatd[i] * (double)p->wdfak[i]) * ap[i];
        out[j] = cd + p->lam * out[j];
        out[j] = out[j] * cgfac[j];
    }
}

/* Fletcher-Reeves CG on the scaled normal equations (src/cg_mod.f90:426-518).
 * cgres[0] = ncg, cgres[k] = REAL*4(eps*dr/dr0) history. */
void orc_cjggra(const orc_cg_problem *p, const double *cgfac, const zdbl *bvec0,
                zdbl *dpar, double *cgres, int *ncg_out) {
    const int manz = p->manz, nanz = p->nanz;
    zdbl *r = (zdbl *)malloc(sizeof(zdbl) * manz), *pv = (zdbl *)calloc(manz, sizeof(zdbl));
    zdbl *bv = (zdbl *)malloc(sizeof(zdbl) * manz), *ap = (zdbl *)malloc(sizeof(zdbl) * nanz);
    double dr = 0, dr0 = 0, dr1 = 0, beta, alpha;
    int ncg = p->ncgmax;
    for (int j = 0; j < manz; ++j) { dpar[j] = 0.0; r[j] = bvec0[j]; }
    for (int k = 1; k <= p->ncgmax; ++k) {
        dr = 0.0;
        for (int j = 0; j < manz; ++j) dr += creal(conj(r[j]) * r[j]);
        if (k == 1) {
            dr0 = dr * p->eps;
            beta = 0.0;
        } else {
            if (dr <= dr0) { ncg = k - 1; break; }
            beta = dr / dr1;
        }
        for (int j = 0; j < manz; ++j) pv[j] = r[j] + beta * pv[j];
        orc_cg_matvec(p, cgfac, pv, ap, bv);
        dr1 = 0.0;
        for (int j = 0; j < manz; ++j) dr1 += creal(conj(pv[j]) * bv[j]);
        alpha = dr / dr1;
        for (int j = 0; j < manz; ++j) { dpar[j] += alpha * pv[j]; r[j] -= alpha * bv[j]; }
        dr1 = dr;
        cgres[k] = (double)(float)(p->eps * dr / dr0);
    }
    cgres[0] = (double)(float)ncg;
    *ncg_out = ncg;
    free(r); free(pv); free(bv); free(ap);
}

/* ------------------------------------------------------------- DC twins */
/* The real-valued inversion (ldc) and the final phase improvement (lfpi):
 * kompadc/scaldc/choldc/kompbdc/vredc in the forward loop of src/inv.f90:355-398,
 * rtrafo's DC branch, bvolti's DC branch, bsendc, and update/cjggdc. */

/* src/kompadc.f90:70-125,128 (lsr off, no sink): real band with DBLE(sigma) */
int orc_kompadc(const orc_grid *g, const zdbl *sigma, const double *elbg, const double *relbg,
                const double *kg, int l, int k, double *a, double *b) {
    const int maxn = g->maxn, npk = NPK(maxn), mb = g->mb, sanz = g->sanz;
    memset(a, 0, sizeof(double) * (size_t)(mb + 1) * sanz);
    memset(b, 0, sizeof(double) * (size_t)sanz);
    int iel = 0;
    for (int t = 0; t < g->typanz; ++t) {
        int ntyp = g->typ[t], nkel = g->selanz[t];
        for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
            if (ntyp > 11) continue;
            int ikl = 0;
            for (int kk = 0; kk < nkel; ++kk) {
                int nzp = g->nrel[iel * maxn + kk];
                for (int ll = 0; ll <= kk; ++ll) {
                    int nnp = g->nrel[iel * maxn + ll];
                    if (abs(nzp - nnp) > mb) return 19;
                    int im = imax_(nzp, nnp) * mb + imin_(nzp, nnp);
                    double dum, dum2;
                    if (ntyp == 11) {
                        int rel = iel - g->elanz;
                        dum = relbg[(size_t)ikl * g->relanz + rel] *
                              kg[((size_t)(k - 1) * g->eanz + (l - 1)) * g->relanz + rel];
                        dum2 = creal(sigma[g->rnr[rel] - 1]);
                    } else {
                        dum = elbg[((size_t)(k - 1) * npk + ikl) * g->elanz + iel];
                        dum2 = creal(sigma[iel]);
                    }
                    a[im - 1] = a[im - 1] + dum * dum2;
                    ++ikl;
                }
            }
        }
    }
    b[g->enr[l - 1] - 1] = -1.0;
    return 0;
}

/* src/scaldc.f90:49-75: unit diagonal, errnr 27 on a non-positive diagonal */
int orc_scaldc(int sanz, int mb, double *a, double *b, double *fak) {
    for (int i = 1; i <= sanz; ++i) {
        int idi = i * (mb + 1);
        double dum = a[idi - 1];
        if (dum <= 0.0) return 27;
        a[idi - 1] = 1.0;
        fak[i - 1] = 1.0 / sqrt(dum);
        b[i - 1] = b[i - 1] * fak[i - 1];
        if (i == 1) continue;
        int i0 = i * mb, ja = imax_(1, i - mb);
        for (int j = ja; j <= i - 1; ++j) a[i0 + j - 1] = a[i0 + j - 1] * fak[i - 1] * fak[j - 1];
    }
    return 0;
}

/* real band Cholesky, row oriented, errnr 28 on s <= 0 (src/choldc.f90:44-76) */
int orc_choldc(int sanz, int mb, double *a) {
    const int m1 = mb + 1;
    for (int i = 1; i <= sanz; ++i) {
        int idi = i * m1, fi = imax_(1, i - mb), i0 = idi - i;
        for (int j = fi; j <= i; ++j) {
            int ij = i0 + j, j0 = j * mb;
            double s = a[ij - 1];
            for (int k = fi; k <= j - 1; ++k) s = s - a[i0 + k - 1] * a[j0 + k - 1];
            if (j < i) {
                a[ij - 1] = s / a[j * m1 - 1];
            } else {
                if (s <= 0.0) return 28;
                a[idi - 1] = sqrt(s);
            }
        }
    }
    return 0;
}

/* pot = -A^{-1} b in real arithmetic (src/vredc.f90:45-76) */
void orc_vredc(int sanz, int mb, const double *a, const double *b, double *pot) {
    const int m1 = mb + 1;
    int idi = 0;
    for (int i = 1; i <= sanz; ++i) {
        idi = i * m1;
        double s = b[i - 1];
        int i0 = idi - i, jlow = imax_(1, i - mb);
        for (int j = jlow; j <= i - 1; ++j) s = s - a[i0 + j - 1] * pot[j - 1];
        pot[i - 1] = s / a[idi - 1];
    }
    for (int i = sanz; i >= 1; --i) {
        pot[i - 1] = -pot[i - 1] / a[idi - 1];
        int jlow = imax_(1, i - mb), i0 = idi - i;
        for (int j = jlow; j <= i - 1; ++j) pot[j - 1] = pot[j - 1] + a[i0 + j - 1] * pot[i - 1];
        idi -= m1;
    }
}

/* DC forward loop (src/inv.f90:355-398; kompbdc src/kompbdc.f90:38-43):
 * kpotdc(j,l,k) = DBLE(pot(j)) * fak(j), layout [k][l][j] as orc_forward */
int orc_forward_dc(const orc_grid *g, const zdbl *sigma, int kwnanz, const double *elbg,
                   const double *relbg, const double *kg, int lbeta, double *kpotdc) {
    const int sanz = g->sanz, mb = g->mb, eanz = g->eanz;
    double *a = (double *)malloc(sizeof(double) * (size_t)(mb + 1) * sanz);
    double *b = (double *)malloc(sizeof(double) * (size_t)sanz);
    double *pot = (double *)malloc(sizeof(double) * (size_t)sanz);
    double *fak = (double *)malloc(sizeof(double) * (size_t)sanz);
    int err = 0;
    for (int k = 1; k <= kwnanz && !err; ++k)
        for (int l = 1; l <= eanz && !err; ++l) {
            if (lbeta || l == 1) {
                err = orc_kompadc(g, sigma, elbg, relbg, kg, l, k, a, b);
                /* src/inv.f90:379-380 */
                if (!err && has_type13(g)) orc_randdc(g, a, b);
                if (!err && g_rwdanz) orc_randbdc2(g, a, b);
                if (!err) err = orc_scaldc(sanz, mb, a, b, fak);
                if (!err) err = orc_choldc(sanz, mb, a);
            } else {
                for (int i = 0; i < sanz; ++i) b[i] = 0.0;
                b[g->enr[l - 1] - 1] = -fak[g->enr[l - 1] - 1];
            }
            if (err) break;
            orc_vredc(sanz, mb, a, b, pot);
            double *kp = kpotdc + ((size_t)(k - 1) * eanz + (l - 1)) * sanz;
            for (int j = 0; j < sanz; ++j) kp[j] = pot[j] * fak[j];
        }
    free(a); free(b); free(pot); free(fak);
    return err;
}

/* src/rtrafo.f90:48-66 (ldc) */
void orc_rtrafo_dc(int sanz, int eanz, int kwnanz, const double *kwnwi, const double *kpotdc, double *hpotdc) {
    const double pi = acos(-1.0);
    for (int l = 0; l < eanz; ++l)
        for (int j = 0; j < sanz; ++j) {
            double s = 0.0;
            for (int k = 0; k < kwnanz; ++k) s = s + kpotdc[((size_t)k * eanz + l) * sanz + j] * kwnwi[k];
            hpotdc[(size_t)l * sanz + j] = s / pi;
        }
}

/* src/bvolti.f90:55-66,80-89 (ldc): real superposition, volt = dcmplx(V), sigmaa = -log(volt) */
int orc_bvolti_dc(int nanz, const int *strnr, const int *vnr, const int *enr, int sanz,
                  const double *hpotdc, zdbl *volt, zdbl *sigmaa) {
    for (int i = 0; i < nanz; ++i) {
        int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
        int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
#define HD(e, c) ((double)(imin_(e, 1) * imin_(c, 1)) * \
                  hpotdc[(size_t)(imax_(c, 1) - 1) * sanz + enr[imax_(e, 1) - 1] - 1])
        double d1 = HD(e4, e2), d2 = HD(e4, e1), d3 = HD(e3, e2), d4 = HD(e3, e1);
#undef HD
        volt[i] = CMPLX((d1 - d2) - (d3 - d4), 0.0);
        if (cabs(volt[i]) == 0.0) return 82;
        sigmaa[i] = -clog(volt[i]);
    }
    return 0;
}

/* src/bsendc.f90:93-175 (swrtr = 1, mnr = identity): sensdc nanz x elanz column-major */
void orc_bsendc(const orc_grid *g, int nanz, const int *strnr, const int *vnr, int kwnanz,
                const double *kwnwi, const double *kpotdc, const double *elbg, const zdbl *sigma,
                const zdbl *volt, double *sensdc) {
    const double pi = acos(-1.0);
    const int sanz = g->sanz, eanz = g->eanz, npk = NPK(g->maxn);
    memset(sensdc, 0, sizeof(double) * (size_t)nanz * g->elanz);
    for (int i = 0; i < nanz; ++i) {
        int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
        int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
        double sup[4] = {0.0, 0.0, 0.0, 0.0};
        int iel = 0;
        for (int t = 0; t < g->typanz; ++t) {
            const int nkel = g->selanz[t];
            if (g->typ[t] > 10) continue;
            for (int j = 0; j < g->nelanz[t]; ++j, ++iel) {
                double dum = 0.0;
                for (int k = 0; k < kwnanz; ++k) {
                    const double *kp = kpotdc + (size_t)k * eanz * sanz;
                    double h = 0.0;
                    for (int mi = 0; mi < nkel; ++mi) {
                        int nzp = g->nrel[iel * g->maxn + mi];
                        for (int mj = 0; mj < nkel; ++mj) {
                            int nnp = g->nrel[iel * g->maxn + mj];
                            int ia = imax_(mi, mj) + 1, ib = imin_(mi, mj) + 1;
                            int imn = ia * (ia - 1) / 2 + ib;
                            if (e1 > 0) sup[0] = kp[(size_t)(e1 - 1) * sanz + nnp - 1];
                            if (e2 > 0) sup[1] = kp[(size_t)(e2 - 1) * sanz + nnp - 1];
                            if (e3 > 0) sup[2] = kp[(size_t)(e3 - 1) * sanz + nzp - 1];
                            if (e4 > 0) sup[3] = kp[(size_t)(e4 - 1) * sanz + nzp - 1];
                            double d = (sup[1] - sup[0]) * (sup[3] - sup[2]);
                            h = h + elbg[((size_t)k * npk + imn - 1) * g->elanz + iel] * d;
                        }
                    }
                    dum = dum + h * kwnwi[k];
                }
                dum = dum / pi;
                sensdc[(size_t)iel * nanz + i] = sensdc[(size_t)iel * nanz + i] +
                                                 dum * creal(sigma[iel] / volt[i]);
            }
        }
    }
}

/* real Jacobian entry of the DC (kind 1: sensdc) or FPI (kind 2: DBLE(sens)) branches */
static double jre(const orc_cg_problem *p, int kind, const double *sensdc, int i, int j) {
    const size_t o = (size_t)j * p->nanz + i;
    return kind == 1 ? sensdc[o] : creal(p->sens[o]);
}

/* update RHS and scaling, DC/FPI branches (src/update.f90:182-191,243-251,261-270,276-290) */
void orc_update_rhs_dc(const orc_cg_problem *p, int kind, const double *sensdc, double *cgfac, zdbl *bvec) {
    const int nanz = p->nanz, manz = p->manz;
    rmatvec(p, NULL, p->par, bvec);
    for (int j = 0; j < manz; ++j) {
        double dum = 0.0;
        for (int i = 0; i < nanz; ++i) {
            double a = jre(p, kind, sensdc, i, j);
            dum = dum + a * a * p->wmatd[i] * (double)p->wdfak[i];
        }
        dum = dum + p->lam * rdiag(p, j);
        cgfac[j] = 1.0 / sqrt(dum);
    }
    for (int j = 0; j < manz; ++j) {
        zdbl cd = 0.0;
        for (int i = 0; i < nanz; ++i)
            cd = cd + (jre(p, kind, sensdc, i, j) * p->wmatd[i] * (double)p->wdfak[i]) * (p->dat[i] - p->sigmaa[i]);
        bvec[j] = cd - p->lam * bvec[j];
        bvec[j] = bvec[j] * cgfac[j];
    }
}

/* R p of bpdc (src/cg_mod.f90:247-264) / bpdctri (:288-296), real, scaled by cgfac */
static void rmatvec_dc(const orc_cg_problem *p, const double *c, const double *v, double *out) {
    const int manz = p->manz, nx = p->nx;
    const double *s = p->smatm;
    for (int i = 1; i <= manz; ++i) {
        double dum = 0.0;
        if (p->ltri == 0) {
            if (i > 1) dum = v[i - 2] * s[manz + i - 2] * c[i - 2];
            if (i < manz) dum = dum + v[i] * s[manz + i - 1] * c[i];
            if (i > nx) dum = dum + v[i - nx - 1] * s[2 * manz + i - nx - 1] * c[i - nx - 1];
            if (i < manz - nx + 1) dum = dum + v[i + nx - 1] * s[2 * manz + i - 1] * c[i + nx - 1];
            out[i - 1] = dum + v[i - 1] * s[i - 1] * c[i - 1];
        } else {
            for (int ij = 0; ij < p->maxn; ++ij) {
                int in = p->nachbar[(size_t)ij * manz + i - 1];
                if (in != 0) dum = dum + v[in - 1] * s[(size_t)ij * manz + i - 1] * c[in - 1];
            }
            out[i - 1] = dum + v[i - 1] * s[(size_t)p->maxn * manz + i - 1] * c[i - 1];
        }
    }
}

/* cjggdc (src/cg_mod.f90:106-196) with bapdc (:198-226), bpdc/bpdctri, bbdc (:385-419).
 * bvec0 is the complex scaled RHS; the CG runs on its real (kind 1) or imaginary
 * (kind 2) part.  dpar = dcmplx(real CG solution), unscaled. */
void orc_cjggdc(const orc_cg_problem *p, int kind, const double *sensdc, const double *cgfac,
                const zdbl *bvec0, zdbl *dpar, double *cgres, int *ncg_out) {
    const int manz = p->manz, nanz = p->nanz;
    double *r = (double *)malloc(sizeof(double) * manz), *pv = (double *)calloc(manz, sizeof(double));
    double *bv = (double *)malloc(sizeof(double) * manz), *ap = (double *)malloc(sizeof(double) * nanz);
    double *dp = (double *)calloc(manz, sizeof(double));
    double dr = 0, dr0 = 0, dr1 = 0, beta, alpha;
    int ncg = p->ncgmax;
    for (int j = 0; j < manz; ++j) r[j] = kind == 2 ? cimag(bvec0[j]) : creal(bvec0[j]);
    for (int k = 1; k <= p->ncgmax; ++k) {
        dr = 0.0;
        for (int j = 0; j < manz; ++j) dr = dr + r[j] * r[j];
        if (k == 1) {
            dr0 = dr * p->eps;
            beta = 0.0;
        } else {
            beta = dr / dr1;
        }
        if (dr <= dr0) { ncg = k - 1; break; }
        for (int j = 0; j < manz; ++j) pv[j] = r[j] + beta * pv[j];
        for (int i = 0; i < nanz; ++i) {
            double s = 0.0;
            for (int j = 0; j < manz; ++j) s = s + pv[j] * jre(p, kind, sensdc, i, j) * cgfac[j];
            ap[i] = s;
        }
        rmatvec_dc(p, cgfac, pv, bv);
        for (int j = 0; j < manz; ++j) {
            double dum = 0.0;
            for (int i = 0; i < nanz; ++i)
                dum = dum + jre(p, kind, sensdc, i, j) * p->wmatd[i] * (double)p->wdfak[i] * ap[i];
            bv[j] = dum + p->lam * bv[j];
            bv[j] = bv[j] * cgfac[j];
        }
        dr1 = 0.0;
        for (int j = 0; j < manz; ++j) dr1 = dr1 + pv[j] * bv[j];
        alpha = dr / dr1;
        for (int j = 0; j < manz; ++j) { dp[j] = dp[j] + alpha * pv[j]; r[j] = r[j] - alpha * bv[j]; }
        dr1 = dr;
        cgres[k] = (double)(float)(p->eps * dr / dr0);
    }
    cgres[0] = (double)(float)ncg;
    *ncg_out = ncg;
    for (int j = 0; j < manz; ++j) dpar[j] = CMPLX(dp[j], 0.0);
    free(r); free(pv); free(bv); free(ap); free(dp);
}

/* -------------------------------------------------------------- CutMcK */
/* Cuthill-McKee with start nodes of minimal degree, best band width kept
 * (cutmckee/cutmck.f:126-300).  Returns mb, or -1 on a disconnected graph. */
/* ---------------------------------------------------------------- buncert */

/* column-major 1-based element (i, j) of an n x n matrix */
#define M2(a, n, i, j) (a)[(size_t)((j) - 1) * (size_t)(n) + (size_t)((i) - 1)]

/* src/bmcm_mod.f90:226-244: upper triangle (k, j), j >= k, mirrored */
static void orc_bata(const orc_cg_problem *p, double *ata) {
    const int n = p->manz, nanz = p->nanz;
    for (int k = 1; k <= n; ++k)
        for (int j = k; j <= n; ++j) {
            zdbl cdum = 0.0;
            for (int i = 1; i <= nanz; ++i) {
                const zdbl a = p->sens[(size_t)(k - 1) * nanz + (i - 1)];
                const zdbl b = p->sens[(size_t)(j - 1) * nanz + (i - 1)];
                cdum = cdum + conj(a) * b * p->wmatd[i - 1] * (double)p->wdfak[i - 1];
            }
            M2(ata, n, k, j) = creal(cdum);
            M2(ata, n, j, k) = creal(cdum);
        }
}

/* src/bmcm_mod.f90:317-383, loops and write order as there */
static int orc_bata_reg(const orc_cg_problem *p, const double *ata, double *reg) {
    const int n = p->manz, nx = p->nx, maxn = p->maxn;
    const double lam = p->lam;
    const double *sm = p->smatm;
    memcpy(reg, ata, sizeof(double) * (size_t)n * n);
    if (p->ltri == 0) {
#define S0(i, c) sm[(size_t)((c) - 1) * n + (size_t)((i) - 1)]
        for (int j = 1; j <= n; ++j)
            for (int i = 1; i <= n; ++i) {
                M2(reg, n, j, j) = M2(ata, n, j, j) + lam * S0(i, 1);
                if (i + 1 < n) {
                    M2(reg, n, i + 1, j) = M2(ata, n, i + 1, j) + lam * S0(i + 1, 2);
                    M2(reg, n, j, i + 1) = M2(reg, n, i + 1, j);
                }
                if (i + nx < n) {
                    M2(reg, n, i + nx, j) = M2(ata, n, i + nx, j) + lam * S0(i + nx, 3);
                    M2(reg, n, j, i + nx) = M2(reg, n, i + nx, j);
                }
            }
#undef S0
    } else if (p->ltri == 3 || p->ltri == 4) {
        for (int i = 1; i <= n; ++i) M2(reg, n, i, i) = M2(ata, n, i, i) + lam * sm[i - 1];
    } else if (p->ltri == 1 || p->ltri == 2 || (p->ltri > 4 && p->ltri < 15)) {
#define SN(j, k) sm[(size_t)((k) - 1) * n + (size_t)((j) - 1)]
#define NB(j, k) p->nachbar[(size_t)((k) - 1) * n + (size_t)((j) - 1)]
        for (int j = 1; j <= n; ++j) {
            M2(reg, n, j, j) = M2(ata, n, j, j) + lam * SN(j, maxn + 1);
            for (int k = 1; k <= NB(j, maxn + 1); ++k) {
                const int ik = NB(j, k);
                if (ik == 0) continue;
                M2(reg, n, ik, j) = M2(ata, n, ik, j) + lam * SN(j, k);
                M2(reg, n, j, ik) = M2(reg, n, ik, j);
            }
        }
#undef SN
#undef NB
    } else {
        return -1;
    }
    return 0;
}

/* src/chold.f90 (stops on a non-positive pivot: here 108) */
static int orc_chold(double *a, double *pd, int n) {
    for (int j = 1; j <= n; ++j) {
        for (int k = 1; k <= j - 1; ++k)
            for (int i = j; i <= n; ++i) M2(a, n, i, j) = M2(a, n, i, j) - M2(a, n, i, k) * M2(a, n, j, k);
        if (M2(a, n, j, j) <= 0) return 108;
        pd[j - 1] = sqrt(M2(a, n, j, j));
        M2(a, n, j, j) = pd[j - 1];
        for (int i = j + 1; i <= n; ++i) M2(a, n, i, j) = M2(a, n, i, j) / pd[j - 1];
    }
    return 0;
}

/* src/linvd.f90: L^-1 in the lower triangle, then (L^-1)^T L^-1 in the upper */
static void orc_linvd(double *a, const double *pd, int n) {
    for (int i = 1; i <= n; ++i) {
        M2(a, n, i, i) = 1. / pd[i - 1];
        for (int j = i + 1; j <= n; ++j) {
            double s = 0.;
            for (int k = i; k <= j - 1; ++k) s = s - M2(a, n, j, k) * M2(a, n, k, i);
            M2(a, n, j, i) = s / pd[j - 1];
            M2(a, n, i, j) = 0.;
        }
    }
    for (int i = 1; i <= n; ++i) {
        M2(a, n, i, i) = M2(a, n, i, i) * M2(a, n, i, i);
        for (int k = i + 1; k <= n; ++k) M2(a, n, i, i) = M2(a, n, i, i) + M2(a, n, k, i) * M2(a, n, k, i);
        for (int j = i + 1; j <= n; ++j)
            for (int k = j; k <= n; ++k) M2(a, n, i, j) = M2(a, n, i, j) + M2(a, n, k, i) * M2(a, n, k, j);
    }
}

/* src/gauss_dble.f90: Gauss-Jordan in place, no pivoting */
static int orc_gauss_dble(double *a, int n) {
    double *temp = (double *)malloc(sizeof(double) * (size_t)n);
    const double eps = 2.220446049250313e-16;
    for (int i = 1; i <= n; ++i) {
        if (fabs(M2(a, n, i, i)) < eps) { free(temp); return 108; }
        const double d = 1.0 / M2(a, n, i, i);
        for (int r = 1; r <= n; ++r) temp[r - 1] = M2(a, n, r, i);
        for (int j = 1; j <= n; ++j) {
            const double c = M2(a, n, i, j) * d;
            for (int r = 1; r <= n; ++r) M2(a, n, r, j) = M2(a, n, r, j) - temp[r - 1] * c;
            M2(a, n, i, j) = c;
        }
        for (int r = 1; r <= n; ++r) M2(a, n, r, i) = temp[r - 1] * (-d);
        M2(a, n, i, i) = d;
    }
    free(temp);
    return 0;
}

/* MATMUL(a, b), k ascending */
static void orc_matmul(const double *a, const double *b, double *c, int n) {
    for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= n; ++i) {
            double s = 0.0;
            for (int k = 1; k <= n; ++k) s += M2(a, n, i, k) * M2(b, n, k, j);
            M2(c, n, i, j) = s;
        }
}

int orc_buncert(const orc_cg_problem *p, int lres, int lcov2, int gauss, double *ata, double *ata_reg,
                double *cov, double *res, double *cov2) {
    const int n = p->manz;
    orc_bata(p, ata);
    if (orc_bata_reg(p, ata, ata_reg)) return -1;
    if (!cov) return 0;
    memcpy(cov, ata_reg, sizeof(double) * (size_t)n * n);
    if (!gauss) {
        double *pd = (double *)malloc(sizeof(double) * (size_t)n);
        int e = orc_chold(cov, pd, n);
        if (!e) orc_linvd(cov, pd, n);
        free(pd);
        if (e) return e;
        for (int i = 1; i <= n; ++i)
            for (int j = 1; j <= i - 1; ++j) M2(cov, n, i, j) = M2(cov, n, j, i);
    } else {
        int e = orc_gauss_dble(cov, n);
        if (e) return e;
    }
    if (lres && res) {
        orc_matmul(cov, ata, res, n);
        if (lcov2 && cov2) orc_matmul(res, cov, cov2, n);
    }
    return 0;
}
#undef M2

int orc_cutmck(int sanz, int typanz, const int *nelanz, const int *selanz, int maxn,
               const int *nrel, int elanz, int *perm, int *start_used) {
    enum { GRMAX = 500, SPMAX = 100 };
    int *graph = (int *)calloc((size_t)GRMAX * sanz, sizeof(int));
    int *grad = (int *)calloc(sanz, sizeof(int));
    int *neu = (int *)calloc(sanz, sizeof(int)), *neuin = (int *)calloc(sanz, sizeof(int));
    int *level = (int *)calloc(sanz + 1, sizeof(int)), *start = (int *)calloc(sanz + 1, sizeof(int));
    char *num = (char *)calloc(sanz, 1);
#define G(k, n) graph[(size_t)((n)-1) * GRMAX + (k)-1]
    int idum = 0;
    for (int l = 0; l < typanz; ++l) {
        for (int m = 1; m <= nelanz[l]; ++m) {
            const int *el = nrel + (size_t)(idum + m - 1) * maxn;
            for (int i = 1; i <= selanz[l] - 1; ++i) {
                int nzp = el[i - 1];
                for (int j = i + 1; j <= selanz[l]; ++j) {
                    int nnp = el[j - 1], dup = 0;
                    for (int k = 1; k <= GRMAX; ++k) {
                        if (G(k, nzp) == nnp) { dup = 1; break; }
                        if (G(k, nzp) > 0) continue;
                        G(k, nzp) = nnp;
                        grad[nzp - 1]++;
                        break;
                    }
                    if (dup) continue;
                    grad[nnp - 1]++;
                    if (grad[nnp - 1] <= GRMAX) G(grad[nnp - 1], nnp) = nzp;
                }
            }
        }
        idum += nelanz[l];
    }
    int mingd = grad[0];
    for (int i = 1; i < sanz; ++i) mingd = imin_(mingd, grad[i]);
    int spanz = imax_(elanz / 10, 1000);
    spanz = imin_(spanz, SPMAX);
    int k = 0;
    for (;;) {
        int done = 0;
        for (int i = 1; i <= sanz; ++i) {
            if (grad[i - 1] == mingd) {
                start[++k] = i;
                if (k >= spanz) { done = 1; break; }
            }
        }
        if (done) break;
        if (mingd > elanz / 10) break;
        mingd++;
    }
    spanz = k - 1;
    int mmin = sanz, kbdm = 0, bad = 0;
    for (int is = 1; is <= spanz && !bad; ++is) {
        int nstart = start[is];
        neu[0] = nstart;
        neuin[nstart - 1] = 1;
        memset(num, 0, sanz);
        num[nstart - 1] = 1;
        level[1] = 1;
        int levs = 1, leve = 1, nlev = 1, l = 1;
        for (;;) {
            for (int j = levs; j <= leve; ++j) {
                int nzp = neu[j - 1], gradzp = grad[nzp - 1];
                for (;;) {
                    int mingr = GRMAX, kk = 0;
                    for (int i = 1; i <= gradzp; ++i) {
                        int nnp = G(i, nzp);
                        if (num[nnp - 1] || grad[nnp - 1] > mingr) continue;
                        mingr = grad[nnp - 1];
                        kk = nnp;
                    }
                    if (kk == 0) break;
                    l++;
                    neu[l - 1] = kk;
                    neuin[kk - 1] = l;
                    num[kk - 1] = 1;
                }
            }
            levs = levs + level[nlev];
            nlev++;
            level[nlev] = l - levs + 1;
            leve = leve + level[nlev];
            if (level[nlev] == 0 && leve < sanz) { bad = 1; break; }
            if (!(leve < sanz)) break;
        }
        if (bad) break;
        int mb = 0;
        for (int i = 1; i <= sanz; ++i) {
            int nzp = neuin[i - 1];
            for (int j = 1; j <= grad[i - 1]; ++j) {
                int kk = neuin[G(j, i) - 1];
                mb = imax_(mb, abs(kk - nzp));
            }
        }
        if (mb < mmin) {
            mmin = mb;
            kbdm = is;
            memcpy(perm, neuin, sizeof(int) * sanz);
        }
    }
#undef G
    if (start_used) *start_used = kbdm > 0 ? start[kbdm] : 0;
    free(graph); free(grad); free(neu); free(neuin); free(level); free(start); free(num);
    return bad ? -1 : mmin;
}
