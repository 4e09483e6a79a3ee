// crt_cg.hip -- Gauss-Newton normal-equation step on the device.
//
// RHS and Jacobi scaling (src/update.f90:63-136,179-231,236-293):
//   cgfac_j = (sum_i |J_ij|^2 w_i + lam R_jj)^(-1/2),   w_i = wmatd_i * wdfak_i
//   bvec_j  = cgfac_j (sum_i conj(J_ij) w_i (d_i - f_i) - lam (R m)_j)
// Fletcher-Reeves CG on B = C (J^H W J + lam R) C (src/cg_mod.f90:426-518):
//   ap = J (C p)                         bap  (src/cg_mod.f90:537-543)
//   t  = R (C p)                         bp / bptri (src/cg_mod.f90:564-576, 600-611)
//   Bp = C (J^H (W ap) + lam t)          bb   (src/cg_mod.f90:715-725)
// with real alpha = dr / Re(p^H Bp), beta = dr / dr_prev and the stop test
// dr <= eps * dr_0 on the squared residual.  Scalars live on the device; every
// reduction has a fixed order, so the iteration is deterministic.  J is
// column-major (nanz, manz) complex FP64, as written by the sensitivity kernel.
// The DC inversion and the final phase improvement (cjggdc, src/cg_mod.f90:106-196)
// run the same kernels on a real J (imaginary parts zero throughout) from the real
// or the imaginary part of bvec; the real copy halves the bytes per iteration.
#include "crt_common.hpp"

#include <algorithm>

#define CG_BLK 256
#define CG_JHW_COLS 4  // columns per wave of the J^H pass

namespace {

// Jacobian entry as complex FP64 from the FP64, the FP32 or the real copy
__device__ __forceinline__ zd ldj(const zd *p) { return *p; }
__device__ __forceinline__ zd ldj(const double *p) { return zd{*p, 0.0}; }
__device__ __forceinline__ zd ldj(const float *p) { return zd{(double)*p, 0.0}; }
__device__ __forceinline__ zd ldj(const float2 *p) {
    const float2 v = *p;
    return zd{(double)v.x, (double)v.y};
}

__device__ __forceinline__ double block_sum(double v, double *red) {
    // fixed-order tree reduction over a 256-thread block
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

// sum of a partial array, identical in every block
__device__ __forceinline__ double reduce_partials(const double *part, int n, double *red) {
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    return block_sum(v, red);
}

// (R x)_i with optional per-cell scaling c on the neighbour values
__device__ __forceinline__ zd rstencil(const CGArgs &A, const zd *x, int i) {
    const int manz = A.manz, nx = A.nx;
    const double *s = A.smatm;
    zd cd = zd{0.0, 0.0};
    if (A.ltri == 0) {
        // 5-point stencil of the regular grid (src/cg_mod.f90:564-576)
        if (i > 0) cd = zfma(cd, x[i - 1], zd{s[manz + i - 1], 0.0});
        if (i < manz - 1) cd = zfma(cd, x[i + 1], zd{s[manz + i], 0.0});
        if (i >= nx) cd = zfma(cd, x[i - nx], zd{s[2 * manz + i - nx], 0.0});
        if (i < manz - nx) cd = zfma(cd, x[i + nx], zd{s[2 * manz + i], 0.0});
        return zfma(cd, x[i], zd{s[i], 0.0});
    }
    // neighbour stencil of unstructured grids (src/cg_mod.f90:600-611)
    for (int ij = 0; ij < A.maxn; ++ij) {
        int in = A.nachbar[(size_t)ij * manz + i];
        if (in != 0) cd = zfma(cd, x[in - 1], zd{s[(size_t)ij * manz + i], 0.0});
    }
    return zfma(cd, x[i], zd{s[(size_t)A.maxn * manz + i], 0.0});
}

}  // namespace

// ----------------------------------------------------------- update: RHS pass
// One wave per column j: cgfac_j and bvec_j in a single pass over J.
template <class JT>
struct RhsArgs {
    int nanz, manz;
    double lam;
    const JT *J;
    const double *w;
    const zd *res;       // d - f (nanz)
    const zd *rm;        // R m (manz)
    const double *rdiag; // R_jj (manz)
    double *cgfac;
    zd *bvec;
};

template <class JT>
__global__ void __launch_bounds__(256) crt_cg_rhs_kernel(RhsArgs<JT> A) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x >> 6) + wv;
    if (j >= A.manz) return;
    const JT *col = A.J + (size_t)j * A.nanz;
    double nrm = 0.0;
    zd s = zd{0.0, 0.0};
    for (int i = lane; i < A.nanz; i += 64) {
        zd v = ldj(col + i);
        double wi = A.w[i];
        nrm += (v.re * v.re + v.im * v.im) * wi;
        zd d = A.res[i];
        // conj(v) * w * d
        s.re += wi * (v.re * d.re + v.im * d.im);
        s.im += wi * (v.re * d.im - v.im * d.re);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        nrm += __shfl_xor(nrm, o, 64);
        s.re += __shfl_xor(s.re, o, 64);
        s.im += __shfl_xor(s.im, o, 64);
    }
    if (lane == 0) {
        double c = 1.0 / sqrt(nrm + A.lam * A.rdiag[j]);
        A.cgfac[j] = c;
        zd b = zd{s.re - A.lam * A.rm[j].re, s.im - A.lam * A.rm[j].im};
        A.bvec[j] = zd{b.re * c, b.im * c};
    }
}

// R m (unscaled) for the RHS
__global__ void __launch_bounds__(256) crt_cg_rmul_kernel(CGArgs A, const zd *x, zd *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.manz) return;
    out[i] = rstencil(A, x, i);
}

// ------------------------------------------------------------- CG iteration
// init: r = bvec, p = 0, dpar = 0, |r|^2 partials
__global__ void __launch_bounds__(CG_BLK) crt_cg_init_kernel(CGArgs A, const zd *bvec) {
    __shared__ double red[8];
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (j < A.manz) {
        zd b = bvec[j];
        if (A.part) b = zd{A.part == 1 ? b.re : b.im, 0.0};  // bvecdc (src/cg_mod.f90:123-127)
        A.r[j] = b;
        A.p[j] = zd{0.0, 0.0};
        A.dpar[j] = zd{0.0, 0.0};
        v = b.re * b.re + b.im * b.im;
    }
    double s = block_sum(v, red);
    if (threadIdx.x == 0) A.part_rr[blockIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        CGScalars z{};
        z.k = 0;
        z.done = 0;
        z.ncg = A.ncgmax;
        *A.scn = z;
    }
}

// K_c: dr, stop test, beta; p = r + beta p; q = C p
// Reads the previous state A.sc, block 0 writes the new state A.scn.
__global__ void __launch_bounds__(CG_BLK) crt_cg_dir_kernel(CGArgs A) {
    __shared__ double red[8];
    const CGScalars cur = *A.sc;
    if (cur.done) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *A.scn = cur;
        return;
    }
    const int k = cur.k + 1;  // 1-based iteration number
    double dr = reduce_partials(A.part_rr, A.n_rr, red);
    // preconditioned CG: beta from r^H M r; the stop test stays on |r|^2
    const double rz = A.fg ? reduce_partials(A.part_rz, A.n_rr, red) : dr;
    double dr0, beta;
    bool stop = false;
    if (k == 1) {
        dr0 = dr * A.eps;
        beta = 0.0;
        // cjggdc tests the stop before its first step as well (src/cg_mod.f90:143-150):
        // a zero right-hand side ends the DC / FPI solve with ncg = 0 and dpar = 0;
        // cjggra tests from the second iteration on (src/cg_mod.f90:455-461)
        stop = A.part != 0 && dr <= dr0;
    } else {
        dr0 = cur.dr0;
        stop = dr <= dr0;
        beta = A.fg ? rz / cur.rz1 : dr / cur.dr1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        CGScalars nx = cur;
        if (stop) {
            nx.done = 1;
            nx.ncg = k - 1;
        } else {
            nx.dr = dr;
            nx.rz = rz;
            nx.dr0 = dr0;
            nx.beta = beta;
            nx.k = k;
            if (k == A.ncgmax) nx.ncg = A.ncgmax;
            A.cgres[k] = (double)(float)(A.eps * dr / dr0);
        }
        *A.scn = nx;
    }
    if (stop) return;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < A.manz) {
        zd r = A.fg ? A.z[j] : A.r[j], p = A.p[j];
        zd pn = zd{r.re + beta * p.re, r.im + beta * p.im};
        A.p[j] = pn;
        double c = A.cgfac[j];
        A.q[j] = zd{pn.re * c, pn.im * c};
    }
}

// K_d: partial ap over column chunks: part_ap[cb][i] = sum_{j in chunk} J(i,j) q_j
template <class JT>
__global__ void __launch_bounds__(CG_BLK) crt_cg_jq_kernel(CGArgs A, const JT *J) {
    __shared__ zd qs[1024];
    if (A.scn->done) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int cb = blockIdx.y;
    const int j0 = cb * A.cb, j1 = min(A.manz, j0 + A.cb);
    for (int t = threadIdx.x; t < j1 - j0; t += blockDim.x) qs[t] = A.q[j0 + t];
    __syncthreads();
    if (i >= A.nanz) return;
    zd acc0 = zd{0.0, 0.0}, acc1 = zd{0.0, 0.0};
    const JT *col = J + (size_t)j0 * A.nanz + i;
    int j = j0;
    // 16-byte entries: four column loads in flight; even columns into acc0, odd ones
    // into acc1, in column order (the same sums as two at a time).  The 8-byte copies
    // measured slower this way (C3: 0.229 -> 0.25 ms per pass) and keep two.
    if constexpr (sizeof(JT) >= 16) {
        for (; j + 3 < j1; j += 4) {
            const size_t n = A.nanz;
            zd v0 = ldj(col), v1 = ldj(col + n), v2 = ldj(col + 2 * n), v3 = ldj(col + 3 * n);
            acc0 = zfma(acc0, v0, qs[j - j0]);
            acc1 = zfma(acc1, v1, qs[j + 1 - j0]);
            acc0 = zfma(acc0, v2, qs[j + 2 - j0]);
            acc1 = zfma(acc1, v3, qs[j + 3 - j0]);
            col += 4 * n;
        }
    }
    for (; j + 1 < j1; j += 2) {
        zd v0 = ldj(col), v1 = ldj(col + A.nanz);
        acc0 = zfma(acc0, v0, qs[j - j0]);
        acc1 = zfma(acc1, v1, qs[j + 1 - j0]);
        col += 2 * (size_t)A.nanz;
    }
    if (j < j1) acc0 = zfma(acc0, ldj(col), qs[j - j0]);
    A.part_ap[(size_t)cb * A.nanz + i] = zadd(acc0, acc1);
}

// K_e: ap = sum over chunks, then wap = w * ap.  64 rows per block: wave v of 16 sums
// the chunks cb = v (mod 16) in order, wave 0 adds the 16 sums in order
constexpr int CG_RED_WAVES = 16;
__global__ void __launch_bounds__(64 * CG_RED_WAVES) crt_cg_apred_kernel(CGArgs A) {
    __shared__ zd red[CG_RED_WAVES][64];
    if (A.scn->done) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    zd s = zd{0.0, 0.0};
    if (i < A.nanz)
        for (int cb = wv; cb < A.ncb; cb += CG_RED_WAVES) s = zadd(s, A.part_ap[(size_t)cb * A.nanz + i]);
    red[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && i < A.nanz) {
        zd t = red[0][lane];
        for (int v = 1; v < CG_RED_WAVES; ++v) t = zadd(t, red[v][lane]);
        const double wi = A.w[i];
        A.ap[i] = zd{t.re * wi, t.im * wi};
    }
}

// K_f: Bp_j = C_j (J^H wap + lam R q)_j, CG_JHW_COLS columns per wave, dr1 partials.
// The column groups run from the last to the first: K_d streams J from its first column to
// its last, so this pass starts on the columns K_d read last, which the memory-side cache
// may still hold, and the next K_d on the ones this pass read last (C3: 0.978 -> 0.959 ms
// per FP64 iteration, the FP32 copy unchanged).  Each group keeps its own partial slot, so
// the sums are the same in either order.
template <class JT>
__global__ void __launch_bounds__(CG_BLK) crt_cg_jhw_kernel(CGArgs A, const JT *J) {
    __shared__ double red[8];
    if (A.scn->done) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int bx = gridDim.x - 1 - blockIdx.x;
    const int jbase = (bx * (blockDim.x >> 6) + wv) * CG_JHW_COLS;
    double dpart = 0.0;
    if (jbase < A.manz) {
        const int nc = min(CG_JHW_COLS, A.manz - jbase);
        // the wave's columns share every load of wap (columns past manz re-read the
        // last one and are dropped); per column one accumulation chain in the order
        // i = lane, lane+64, ... (the sums do not depend on the unroll)
        const JT *col[CG_JHW_COLS];
#pragma unroll
        for (int c = 0; c < CG_JHW_COLS; ++c) col[c] = J + (size_t)(jbase + min(c, nc - 1)) * A.nanz;
        zd s[CG_JHW_COLS];
#pragma unroll
        for (int c = 0; c < CG_JHW_COLS; ++c) s[c] = zd{0.0, 0.0};
        constexpr int U = sizeof(JT) >= 16 ? 2 : 4;  // rows per lane in flight
        int i = lane;
        for (; i + 64 * (U - 1) < A.nanz; i += 64 * U) {
            zd a[U], v[U][CG_JHW_COLS];
#pragma unroll
            for (int u = 0; u < U; ++u) a[u] = A.ap[i + 64 * u];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int c = 0; c < CG_JHW_COLS; ++c) v[u][c] = ldj(col[c] + i + 64 * u);
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int c = 0; c < CG_JHW_COLS; ++c) {
                    s[c].re += v[u][c].re * a[u].re + v[u][c].im * a[u].im;
                    s[c].im += v[u][c].re * a[u].im - v[u][c].im * a[u].re;
                }
        }
        for (; i < A.nanz; i += 64) {
            const zd a = A.ap[i];
#pragma unroll
            for (int c = 0; c < CG_JHW_COLS; ++c) {
                const zd v = ldj(col[c] + i);
                s[c].re += v.re * a.re + v.im * a.im;
                s[c].im += v.re * a.im - v.im * a.re;
            }
        }
#pragma unroll
        for (int c = 0; c < CG_JHW_COLS; ++c)
            for (int o = 32; o >= 1; o >>= 1) {
                s[c].re += __shfl_xor(s[c].re, o, 64);
                s[c].im += __shfl_xor(s[c].im, o, 64);
            }
        if (lane == 0) {
            for (int c = 0; c < nc; ++c) {
                const int j = jbase + c;
                zd t = rstencil(A, A.q, j);
                double cf = A.cgfac[j];
                zd b = zd{(s[c].re + A.lam * t.re) * cf, (s[c].im + A.lam * t.im) * cf};
                A.bv[j] = b;
                zd p = A.p[j];
                dpart += p.re * b.re + p.im * b.im;
            }
        }
    }
    double s = block_sum(dpart, red);
    if (threadIdx.x == 0) A.part_dr1[bx] = s;
}

// K_h: alpha; dpar += alpha p; r -= alpha Bp; |r|^2 partials
__global__ void __launch_bounds__(CG_BLK) crt_cg_upd_kernel(CGArgs A) {
    __shared__ double red[8];
    if (A.scn->done) return;
    double pbp = reduce_partials(A.part_dr1, A.n_dr1, red);
    double dr = A.scn->dr;
    double alpha = (A.fg ? A.scn->rz : dr) / pbp;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (j < A.manz) {
        zd p = A.p[j], b = A.bv[j], r = A.r[j], d = A.dpar[j];
        A.dpar[j] = zd{d.re + alpha * p.re, d.im + alpha * p.im};
        zd rn = zd{r.re - alpha * b.re, r.im - alpha * b.im};
        A.r[j] = rn;
        v = rn.re * rn.re + rn.im * rn.im;
    }
    double s = block_sum(v, red);
    if (threadIdx.x == 0) A.part_rr[blockIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.scn->alpha = alpha;
        A.scn->dr1 = dr;
        A.scn->rz1 = A.scn->rz;
    }
}

// final scaling dpar <- C dpar (src/update.f90:302-307) and par += step dpar
// (FPI: dpar = i * dpar first, src/update.f90:302-304)
__global__ void __launch_bounds__(256) crt_cg_final_kernel(int manz, double step, const double *cgfac, zd *dpar,
                                                           zd *par, int part) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= manz) return;
    zd d = dpar[j];
    if (part == 2) d = zd{0.0, d.re};
    double c = cgfac[j];
    d = zd{d.re * c, d.im * c};
    dpar[j] = d;
    if (par) par[j] = zd{par[j].re + step * d.re, par[j].im + step * d.im};
}

// ------------------------------------------------- FSAI preconditioner M = G^H G
// t = G r
__global__ void __launch_bounds__(CG_BLK) crt_cg_pg_kernel(CGArgs A) {
    if (A.scn->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.manz) return;
    zd t = zd{0.0, 0.0};
#pragma unroll
    for (int s = 0; s < CRT_FSAI_NP; ++s) {
        const int c = A.fpat[(size_t)j * CRT_FSAI_NP + s];
        if (c >= 0) t = zfma(t, A.fg[(size_t)j * CRT_FSAI_NP + s], A.r[c]);
    }
    A.t[j] = t;
}

// z = G^H t and the partial sums of Re(r^H z)
__global__ void __launch_bounds__(CG_BLK) crt_cg_pgh_kernel(CGArgs A) {
    __shared__ double red[8];
    if (A.scn->done) return;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (j < A.manz) {
        zd z = zd{0.0, 0.0};
#pragma unroll
        for (int s = 0; s < CRT_FSAI_NP; ++s) {
            const int e = A.ftp[(size_t)j * CRT_FSAI_NP + s];
            if (e >= 0) {
                const zd g = A.fg[e], t = A.t[e / CRT_FSAI_NP];
                // conj(g) * t
                z.re += g.re * t.re + g.im * t.im;
                z.im += g.re * t.im - g.im * t.re;
            }
        }
        A.z[j] = z;
        const zd r = A.r[j];
        v = r.re * z.re + r.im * z.im;
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) A.part_rz[blockIdx.x] = s;
}

// entry R(a, b) of the regularisation stencil
__device__ __forceinline__ double rentry(const CGArgs &A, int a, int b) {
    const int manz = A.manz;
    const double *s = A.smatm;
    if (A.ltri == 0) {
        if (a == b) return s[a];
        if (b == a + 1) return s[manz + a];
        if (a == b + 1) return s[manz + b];
        if (b == a + A.nx) return s[2 * manz + a];
        if (a == b + A.nx) return s[2 * manz + b];
        return 0.0;
    }
    if (a == b) return s[(size_t)A.maxn * manz + a];
    for (int ij = 0; ij < A.maxn; ++ij)
        if (A.nachbar[(size_t)ij * manz + a] == b + 1) return s[(size_t)ij * manz + a];
    return 0.0;
}

// Row j of G: the Hermitian block B(S,S) of B = C (J^H W J + lam R) C on the
// pattern S (last entry j), y = B(S,S)^-1 e_last, g = y / sqrt(y_last).  One wave
// per cell: lanes reduce the weighted column products, lane 0 solves.
__global__ void __launch_bounds__(256) crt_cg_fsai_kernel(CGArgs A, zd *g) {
    constexpr int NP = CRT_FSAI_NP;
    constexpr int NG = NP * (NP + 1) / 2;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x >> 6) + wv;
    if (j >= A.manz) return;
    int pc[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) pc[s] = A.fpat[(size_t)j * NP + s];
    double gr[NG], gi[NG];
#pragma unroll
    for (int q = 0; q < NG; ++q) gr[q] = gi[q] = 0.0;
    for (int i = lane; i < A.nanz; i += 64) {
        const double wi = A.w[i];
        zd v[NP];
#pragma unroll
        for (int s = 0; s < NP; ++s)
            v[s] = pc[s] < 0 ? zd{0.0, 0.0}
                             : (A.Jr ? zd{A.Jr[(size_t)pc[s] * A.nanz + i], 0.0} : A.J[(size_t)pc[s] * A.nanz + i]);
        int q = 0;
#pragma unroll
        for (int a = 0; a < NP; ++a)
#pragma unroll
            for (int b = a; b < NP; ++b, ++q) {
                // w conj(J_a) J_b
                gr[q] += wi * (v[a].re * v[b].re + v[a].im * v[b].im);
                gi[q] += wi * (v[a].re * v[b].im - v[a].im * v[b].re);
            }
    }
#pragma unroll
    for (int q = 0; q < NG; ++q)
        for (int o = 32; o >= 1; o >>= 1) {
            gr[q] += __shfl_xor(gr[q], o, 64);
            gi[q] += __shfl_xor(gi[q], o, 64);
        }
    if (lane != 0) return;
    // Hermitian block (upper triangle stored row-wise), unused slots -> identity
    double c[NP];
#pragma unroll
    for (int s = 0; s < NP; ++s) c[s] = pc[s] >= 0 ? A.cgfac[pc[s]] : 0.0;
    zd m[NP][NP + 1];
    int q = 0;
#pragma unroll
    for (int a = 0; a < NP; ++a)
#pragma unroll
        for (int b = a; b < NP; ++b, ++q) {
            zd e;
            if (pc[a] < 0 || pc[b] < 0) {
                e = zd{a == b ? 1.0 : 0.0, 0.0};
            } else {
                const double rr = A.lam * rentry(A, pc[a], pc[b]);
                e = zd{c[a] * c[b] * (gr[q] + rr), c[a] * c[b] * gi[q]};
            }
            m[a][b] = e;
            m[b][a] = zd{e.re, -e.im};
        }
#pragma unroll
    for (int a = 0; a < NP; ++a) m[a][NP] = zd{a == NP - 1 ? 1.0 : 0.0, 0.0};
    // Gaussian elimination (Hermitian positive definite: no pivoting)
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const zd inv = zdiv(zd{1.0, 0.0}, m[k][k]);
#pragma unroll
        for (int b = k; b <= NP; ++b) m[k][b] = zmul(m[k][b], inv);
#pragma unroll
        for (int a = k + 1; a < NP; ++a) {
            const zd f = m[a][k];
#pragma unroll
            for (int b = k; b <= NP; ++b) m[a][b] = zsub(m[a][b], zmul(f, m[k][b]));
        }
    }
    zd y[NP];
#pragma unroll
    for (int a = NP - 1; a >= 0; --a) {
        zd acc = m[a][NP];
#pragma unroll
        for (int b = a + 1; b < NP; ++b) acc = zsub(acc, zmul(m[a][b], y[b]));
        y[a] = acc;
    }
    const double yl = y[NP - 1].re;
    const double sc = yl > 0.0 ? 1.0 / sqrt(yl) : 0.0;
#pragma unroll
    for (int s = 0; s < NP; ++s)
        g[(size_t)j * NP + s] = pc[s] >= 0 ? zd{y[s].re * sc, y[s].im * sc} : zd{0.0, 0.0};
}

// ---------------------------------------------------------------- launchers
int crt_cg_geometry(int nanz, int manz, int *ncb, int *cb, int *n_dr1, int *n_rr) {
    int rb = (nanz + CG_BLK - 1) / CG_BLK;
    // aim for >= 2048 blocks in the J q pass, at most 1024 columns per chunk
    int want = std::max(1, 2048 / std::max(rb, 1));
    int c = std::max(16, std::min(1024, (manz + want - 1) / want));
    *cb = c;
    *ncb = (manz + c - 1) / c;
    *n_dr1 = (manz + 4 * CG_JHW_COLS - 1) / (4 * CG_JHW_COLS);
    *n_rr = (manz + CG_BLK - 1) / CG_BLK;
    return 0;
}

int crt_launch_cg_rhs(hipStream_t s, int nanz, int manz, double lam, const zd *J, const double *Jr,
                      const double *w, const zd *res, const zd *rm, const double *rdiag, double *cgfac, zd *bvec) {
    if (Jr) {
        RhsArgs<double> a{nanz, manz, lam, Jr, w, res, rm, rdiag, cgfac, bvec};
        hipLaunchKernelGGL(crt_cg_rhs_kernel<double>, dim3((manz + 3) / 4), dim3(256), 0, s, a);
    } else {
        RhsArgs<zd> a{nanz, manz, lam, J, w, res, rm, rdiag, cgfac, bvec};
        hipLaunchKernelGGL(crt_cg_rhs_kernel<zd>, dim3((manz + 3) / 4), dim3(256), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cg_rmul(hipStream_t s, const CGArgs &a, const zd *x, zd *out) {
    hipLaunchKernelGGL(crt_cg_rmul_kernel, dim3((a.manz + 255) / 256), dim3(256), 0, s, a, x, out);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cg_init(hipStream_t s, const CGArgs &a, const zd *bvec) {
    hipLaunchKernelGGL(crt_cg_init_kernel, dim3(a.n_rr), dim3(CG_BLK), 0, s, a, bvec);
    if (a.fg) {
        hipLaunchKernelGGL(crt_cg_pg_kernel, dim3(a.n_rr), dim3(CG_BLK), 0, s, a);
        hipLaunchKernelGGL(crt_cg_pgh_kernel, dim3(a.n_rr), dim3(CG_BLK), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cg_iter(hipStream_t s, const CGArgs &a) {
    hipLaunchKernelGGL(crt_cg_dir_kernel, dim3(a.n_rr), dim3(CG_BLK), 0, s, a);
    const dim3 gq((a.nanz + CG_BLK - 1) / CG_BLK, a.ncb);
    if (a.Jr32)
        hipLaunchKernelGGL(crt_cg_jq_kernel<float>, gq, dim3(CG_BLK), 0, s, a, a.Jr32);
    else if (a.Jr)
        hipLaunchKernelGGL(crt_cg_jq_kernel<double>, gq, dim3(CG_BLK), 0, s, a, a.Jr);
    else if (a.J32)
        hipLaunchKernelGGL(crt_cg_jq_kernel<float2>, gq, dim3(CG_BLK), 0, s, a, a.J32);
    else
        hipLaunchKernelGGL(crt_cg_jq_kernel<zd>, gq, dim3(CG_BLK), 0, s, a, a.J);
    hipLaunchKernelGGL(crt_cg_apred_kernel, dim3((a.nanz + 63) / 64), dim3(64 * CG_RED_WAVES), 0, s, a);
    if (a.Jr32)
        hipLaunchKernelGGL(crt_cg_jhw_kernel<float>, dim3(a.n_dr1), dim3(CG_BLK), 0, s, a, a.Jr32);
    else if (a.Jr)
        hipLaunchKernelGGL(crt_cg_jhw_kernel<double>, dim3(a.n_dr1), dim3(CG_BLK), 0, s, a, a.Jr);
    else if (a.J32)
        hipLaunchKernelGGL(crt_cg_jhw_kernel<float2>, dim3(a.n_dr1), dim3(CG_BLK), 0, s, a, a.J32);
    else
        hipLaunchKernelGGL(crt_cg_jhw_kernel<zd>, dim3(a.n_dr1), dim3(CG_BLK), 0, s, a, a.J);
    hipLaunchKernelGGL(crt_cg_upd_kernel, dim3(a.n_rr), dim3(CG_BLK), 0, s, a);
    if (a.fg) {
        hipLaunchKernelGGL(crt_cg_pg_kernel, dim3(a.n_rr), dim3(CG_BLK), 0, s, a);
        hipLaunchKernelGGL(crt_cg_pgh_kernel, dim3(a.n_rr), dim3(CG_BLK), 0, s, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cg_final(hipStream_t s, int manz, double step, const double *cgfac, zd *dpar, zd *par, int part) {
    hipLaunchKernelGGL(crt_cg_final_kernel, dim3((manz + 255) / 256), dim3(256), 0, s, manz, step, cgfac, dpar, par,
                       part);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// real parts of a complex array (the real Jacobian of the DC / FPI branches, and
// the DC getters)
__global__ void __launch_bounds__(256) crt_real_part_kernel(size_t n, const zd *x, double *out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = x[t].re;
}

int crt_launch_real_part(hipStream_t s, size_t n, const zd *x, double *out) {
    hipLaunchKernelGGL(crt_real_part_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, out);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// complex FP32 copy of J for the CG iterations (round to nearest)
__global__ void __launch_bounds__(256) crt_cg_j32_kernel(size_t n, const zd *J, float2 *J32) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const zd v = J[t];
    J32[t] = make_float2((float)v.re, (float)v.im);
}

__global__ void __launch_bounds__(256) crt_cg_jr32_kernel(size_t n, const double *Jr, float *Jr32) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) Jr32[t] = (float)Jr[t];
}

int crt_launch_cg_jr32(hipStream_t s, size_t n, const double *Jr, float *Jr32) {
    hipLaunchKernelGGL(crt_cg_jr32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, Jr, Jr32);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cg_j32(hipStream_t s, size_t n, const zd *J, float2 *J32) {
    hipLaunchKernelGGL(crt_cg_j32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, J, J32);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cg_fsai(hipStream_t s, const CGArgs &a, zd *g) {
    hipLaunchKernelGGL(crt_cg_fsai_kernel, dim3((a.manz + 3) / 4), dim3(256), 0, s, a, g);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
