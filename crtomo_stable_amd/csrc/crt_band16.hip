// crt_band16.hip -- FP64 complex-symmetric band factorisation on the matrix cores.
//
// Same mathematics as crt_band.hip (src/kompab.f90, src/scalab.f90,
// src/chol.f90, src/vre.f90, src/fem.f90:383-394) with 16x16 tiles:
//   * the active window of the band (tiles (J+a, J+b), 0 <= b <= a <= D,
//     D = ceil(mb/16)) lives in registers as v_mfma_f64_16x16x4 accumulators
//     (real and imaginary part, 8 VGPRs each);
//   * tiles on diagonal d = a-b slide one position per step; wave 0 owns the
//     main diagonal, wave g >= 1 owns diagonals g and D+1-g, so every wave
//     holds D+1 tiles at most and the rank-16 updates are balanced.  The code
//     of each wave is instantiated separately, so every slot index is a
//     compile-time constant;
//   * per block column J: wave 0 factors the diagonal tile and runs the
//     forward substitution of the factor's right-hand side (one LDS exchange
//     per column), waves 1.. solve the panel rows X = A L^-T column by column
//     and stage the entering row block from the CSR values, then every wave
//     applies its rank-16 updates on the matrix cores;
//   * the panel column and the entering row block are staged in LDS (about
//     60 KB at D = 6), so two workgroups share a CU and one system's pivot
//     chain overlaps the other's matrix-core updates.
// Complex products use four real MFMAs: Tr -= Ar Br - Ai Bi, Ti -= Ar Bi + Ai Br.
// Operand/result lanes of __builtin_amdgcn_mfma_f64_16x16x4f64 (checked on the
// device): A[i][k] at lane i+16k, B[k][j] at lane j+16k, D[i][j] at lane
// j+16(i%4), element i/4.
#include "crt_common.hpp"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int T16 = 16;
constexpr int PS = 17;  // padded row stride (complex) of LDS panel tiles: conflict-free column reads

__device__ __forceinline__ zd shfl_z16(zd v, int src) { return zd{__shfl(v.re, src, 64), __shfl(v.im, src, 64)}; }
__device__ __forceinline__ zd shfl_xor_z16(zd v, int m) {
    return zd{__shfl_xor(v.re, m, 64), __shfl_xor(v.im, m, 64)};
}

__device__ __forceinline__ double readlane_d(double x, int src) {
    long long b = __double_as_longlong(x);
    int lo = __builtin_amdgcn_readlane((int)b, src);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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

__device__ __forceinline__ zd readlane_z(zd v, int src) { return zd{readlane_d(v.re, src), readlane_d(v.im, src)}; }
__device__ __forceinline__ zd zsel(bool c, zd a, zd b) { return zd{c ? a.re : b.re, c ? a.im : b.im}; }

template <int D>
struct Cfg {
    static constexpr int W = 1 + D / 2 + (D & 1);  // waves
    static constexpr int NT = 64 * W;               // threads
};

// diagonals held by wave WV: segment A (diagonal dA, LA positions) and, for
// 1 <= WV <= D/2, segment B (diagonal dB = D+1-WV, LB = WV positions)
template <int D, int WV>
struct WaveCfg {
    static constexpr bool pair = WV >= 1 && WV <= D / 2;
    static constexpr int dA = WV == 0 ? 0 : (pair ? WV : (D + 1) / 2);
    static constexpr int LA = WV == 0 ? D + 1 : (pair ? D + 1 - WV : (D + 1) / 2);
    static constexpr int dB = pair ? D + 1 - WV : 0;
    static constexpr int LB = pair ? WV : 0;
};

}  // namespace

struct Factor16Args {
    DevGrid g;
    int nsys;
    const int2 *sys;
    const double *fak;  // [sys][nblk*16]
    const zd *vals;     // [sys][nnz]
    zd *panels;         // [sys][nblk][D+1][16][16]
    zd *y;              // [sys][nblk*16]
    zd *dinv;           // [sys][nblk*16] reciprocal pivots 1/L(i,i)
    int *status;
};

namespace {

// LDS of one workgroup; a static layout keeps every address an immediate
template <int D>
struct F16Lds {
    zd P[(D + 1) * T16 * PS];  // panel tiles of block column J, 16 x PS each
    zd S[(D + 1) * 256];       // staging tiles of the entering row block, 16 x 16 each
    zd bring[(D + 2) * T16];   // right-hand-side blocks J .. J+D+1
    zd col[32];                // current column of the diagonal factor, [16] = y of the column
    zd pinv[T16];              // reciprocal pivots of the diagonal tile
    zd yj[T16];                // forward-substituted right-hand side of block J
};

// CSR rows of block Ie (contiguous, values scaled by crt_assemble_kernel) into
// the staging tiles; padding rows past sanz get a unit diagonal
template <int D>
__device__ __forceinline__ void stage_block(const DevGrid &g, int Ie, int t0, int nthr, const zd *vals,
                                            const double *fak, int enode, F16Lds<D> &m) {
    const int qb = g.row_ptr[Ie * T16], qe = g.row_ptr[min(Ie * T16 + T16, g.sanz)];
    for (int q = qb + t0; q < qe; q += nthr) {
        int i = g.crow[q], j = g.col[q];
        m.S[(Ie - j / T16) * 256 + (i - Ie * T16) * T16 + (j % T16)] = vals[q];
    }
    if (t0 < T16) {
        int i = Ie * T16 + t0;
        if (i >= g.sanz) m.S[t0 * T16 + t0] = zd{1.0, 0.0};
        zd bv = zd{0.0, 0.0};
        if (i < g.sanz && i == enode) bv = zd{-fak[i], 0.0};
        m.bring[(Ie % (D + 2)) * T16 + t0] = bv;
    }
}

__device__ __forceinline__ d4 tile_re(const zd *t, int jc, int rq) {
    return d4{t[rq * T16 + jc].re, t[(4 + rq) * T16 + jc].re, t[(8 + rq) * T16 + jc].re, t[(12 + rq) * T16 + jc].re};
}
__device__ __forceinline__ d4 tile_im(const zd *t, int jc, int rq) {
    return d4{t[rq * T16 + jc].im, t[(4 + rq) * T16 + jc].im, t[(8 + rq) * T16 + jc].im, t[(12 + rq) * T16 + jc].im};
}

// T -= Pa Pb^T for 16x16 complex tiles in LDS (row stride PS), T in result layout
__device__ __forceinline__ void mfma_update(d4 &tr, d4 &ti, const zd *Pa, const zd *Pb, int jc, int rq) {
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        zd va = Pa[jc * PS + 4 * kc + rq];
        zd vb = Pb[jc * PS + 4 * kc + rq];
        tr = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.re, vb.re, tr, 0, 0, 0);
        tr = __builtin_amdgcn_mfma_f64_16x16x4f64(va.im, vb.im, tr, 0, 0, 0);
        ti = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.re, vb.im, ti, 0, 0, 0);
        ti = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.im, vb.re, ti, 0, 0, 0);
    }
}

template <int D, int WV>
__device__ __forceinline__ void factor16_wave(const Factor16Args &A, F16Lds<D> &m) {
    using WC = WaveCfg<D, WV>;
    constexpr int NT = Cfg<D>::NT;
    constexpr int LA = WC::LA, LB = WC::LB, dA = WC::dA, dB = WC::dB;
    constexpr int LBX = LB > 0 ? LB : 1;
    const DevGrid &g = A.g;
    const int nblk = g.nblk;
    const int s = blockIdx.x;
    const int lane = threadIdx.x & 63, tid = WV * 64 + lane;
    const int enode = g.enr[A.sys[s].x];
    const double *fak = A.fak + (size_t)s * nblk * T16;
    const zd *vals = A.vals + (size_t)s * g.nnz;
    zd *panel = A.panels + (size_t)s * nblk * (D + 1) * 256;
    zd *yout = A.y + (size_t)s * nblk * T16;
    const int jc = lane & 15, rq = lane >> 4;

    d4 Ar[LA], Ai[LA], Br[LBX], Bi[LBX];
#pragma unroll
    for (int q = 0; q < LA; ++q) Ar[q] = Ai[q] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < LBX; ++q) Br[q] = Bi[q] = d4{0.0, 0.0, 0.0, 0.0};
    int err = 0;

    for (int J = -(D + 1); J < nblk; ++J) {
        const int Ie = J + D + 1;  // entering row block
        // ------------------------------------------------ A: panel tiles -> LDS
        if (J >= 0) {
            zd *pa = m.P + dA * T16 * PS;
#pragma unroll
            for (int v = 0; v < 4; ++v) pa[(4 * v + rq) * PS + jc] = zd{Ar[0][v], Ai[0][v]};
            if constexpr (LB > 0) {
                zd *pb = m.P + dB * T16 * PS;
#pragma unroll
                for (int v = 0; v < 4; ++v) pb[(4 * v + rq) * PS + jc] = zd{Br[0][v], Bi[0][v]};
            }
        }
        __syncthreads();
        // ---------------- B: wave 0 factors the diagonal tile | others stage Ie
        if (WV == 0 && J >= 0) {
            // right-looking factorisation of the diagonal tile in the result
            // layout; column c travels through LDS (col), the next pivot is formed
            // from two lane reads so the pivot chain does not wait for the exchange
            zd a[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) a[v] = m.P[(4 * v + rq) * PS + jc];
            // lanes of column group jc carry the right-hand side entry jc
            zd yv = m.bring[(J % (D + 2)) * T16 + jc];
            zd pnext = readlane_z(a[0], 0);
#pragma unroll
            for (int c = 0; c < T16; ++c) {
                zd Lc, iv;
                bool bad;
                pivot_fast(pnext, Lc, iv, bad);
                if (bad) err = CRT_ERR_PIVOT;
                zd cv[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) cv[v] = zsel(4 * v + rq == c, Lc, zmul(a[v], iv));
                const zd yc = zmul(yv, iv);
                if (c + 1 < T16) {
                    const int n = c + 1;
                    zd dn = readlane_z(a[n / 4], n + 16 * (n % 4));  // a(n, n)
                    zd ln = readlane_z(cv[n / 4], c + 16 * (n % 4)); // L(n, c)
                    pnext = zfms(dn, ln, ln);
                }
                if (jc == c) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) m.col[4 * v + rq] = cv[v];
                    if (rq == 0) {
                        m.col[16] = yc;
                        m.pinv[c] = iv;
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) a[v] = zsel(jc == c, cv[v], a[v]);
                yv = zsel(jc == c, yc, yv);
                __builtin_amdgcn_wave_barrier();
                const zd ljc = m.col[jc], ycol = m.col[16];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = 4 * v + rq;
                    a[v] = zsel(jc > c && i >= jc, zfms(a[v], m.col[i], ljc), a[v]);
                }
                yv = zsel(jc > c, zfms(yv, ljc, ycol), yv);
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int i = 4 * v + rq;
                m.P[i * PS + jc] = (jc <= i) ? a[v] : zd{0.0, 0.0};
            }
            if (rq == 0) {
                m.yj[jc] = yv;
                m.bring[(J % (D + 2)) * T16 + jc] = yv;
                yout[(size_t)J * T16 + jc] = yv;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane < T16) A.dinv[(size_t)s * nblk * T16 + (size_t)J * T16 + lane] = m.pinv[lane];
        } else if (Ie < nblk) {
            if (J < 0) stage_block<D>(g, Ie, tid, NT, vals, fak, enode, m);
            else stage_block<D>(g, Ie, tid - 64, NT - 64, vals, fak, enode, m);
        }
        __syncthreads();
        // ------------------ C: panel rows X = A L^-T (waves 1..), RHS update
        if (WV >= 1 && J >= 0) {
            const int t = tid - 64;
            const int a = 1 + t / T16, r = t % T16;
            if (t < T16 * D && J + a < nblk) {
                zd *row = m.P + (a * T16 + r) * PS;
                zd x[T16];
#pragma unroll
                for (int c = 0; c < T16; ++c) x[c] = row[c];
                zd bacc = m.bring[((J + a) % (D + 2)) * T16 + r];
#pragma unroll
                for (int c = 0; c < T16; ++c) {
                    x[c] = zmul(x[c], m.pinv[c]);
                    bacc = zfms(bacc, x[c], m.yj[c]);
#pragma unroll
                    for (int i = c + 1; i < T16; ++i) x[i] = zfms(x[i], x[c], m.P[i * PS + c]);
                }
#pragma unroll
                for (int c = 0; c < T16; ++c) row[c] = x[c];
                m.bring[((J + a) % (D + 2)) * T16 + r] = bacc;
            }
        }
        __syncthreads();
        // ------------------- D: store panel, rank-16 updates on the matrix cores, slide
        if (J >= 0) {
            zd *pj = panel + (size_t)J * (D + 1) * 256;
            for (int t = tid; t < (D + 1) * 256; t += NT) {
                int a = t >> 8, rc = t & 255;
                pj[t] = (J + a < nblk) ? m.P[(a * T16 + (rc >> 4)) * PS + (rc & 15)] : zd{0.0, 0.0};
            }
#pragma unroll
            for (int p = 1; p < LA; ++p)
                if (J + p + dA < nblk)
                    mfma_update(Ar[p], Ai[p], m.P + (p + dA) * T16 * PS, m.P + p * T16 * PS, jc, rq);
            if constexpr (LB > 0) {
#pragma unroll
                for (int p = 1; p < LB; ++p)
                    if (J + p + dB < nblk)
                        mfma_update(Br[p], Bi[p], m.P + (p + dB) * T16 * PS, m.P + p * T16 * PS, jc, rq);
            }
        }
        // slide: position p -> p-1; the entering tiles take the last positions
#pragma unroll
        for (int p = 0; p + 1 < LA; ++p) { Ar[p] = Ar[p + 1]; Ai[p] = Ai[p + 1]; }
        {
            zd *sa = m.S + dA * 256;
            Ar[LA - 1] = tile_re(sa, jc, rq);
            Ai[LA - 1] = tile_im(sa, jc, rq);
#pragma unroll
            for (int v = 0; v < 4; ++v) sa[(4 * v + rq) * T16 + jc] = zd{0.0, 0.0};
        }
        if constexpr (LB > 0) {
#pragma unroll
            for (int p = 0; p + 1 < LB; ++p) { Br[p] = Br[p + 1]; Bi[p] = Bi[p + 1]; }
            zd *sb = m.S + dB * 256;
            Br[LB - 1] = tile_re(sb, jc, rq);
            Bi[LB - 1] = tile_im(sb, jc, rq);
#pragma unroll
            for (int v = 0; v < 4; ++v) sb[(4 * v + rq) * T16 + jc] = zd{0.0, 0.0};
        }
        __syncthreads();
    }
    if (WV == 0 && lane == 0 && err && A.status[s] == 0) A.status[s] = err;
}

template <int D, int WV>
__device__ __forceinline__ void factor16_dispatch(int wave, const Factor16Args &A, F16Lds<D> &m) {
    if constexpr (WV < Cfg<D>::W) {
        if (wave == WV) factor16_wave<D, WV>(A, m);
        else factor16_dispatch<D, WV + 1>(wave, A, m);
    }
}

}  // namespace

template <int D>
__global__ void __launch_bounds__(Cfg<D>::NT, 2) crt_factor16_kernel(Factor16Args A) {
    __shared__ F16Lds<D> m;
    if ((int)blockIdx.x >= A.nsys) return;
    for (int t = threadIdx.x; t < (D + 1) * 256; t += Cfg<D>::NT) m.S[t] = zd{0.0, 0.0};
    __syncthreads();
    factor16_dispatch<D, 0>(threadIdx.x >> 6, A, m);
}

size_t crt_factor16_lds_bytes(int D) {
    return ((size_t)(D + 1) * T16 * PS + (size_t)(D + 1) * 256 + (size_t)(D + 2) * T16 + 64) * sizeof(zd);
}
static_assert(sizeof(F16Lds<6>) == 62208, "LDS layout of the 16x16 factorisation");

#define CRT_F16_MAXD 10

template <int D>
static int launch16(hipStream_t st, const Factor16Args &a) {
    hipLaunchKernelGGL(crt_factor16_kernel<D>, dim3(a.nsys), dim3(Cfg<D>::NT), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_factor16_supported(int mb) {
    int D = (mb + 15) / 16;
    return D >= 1 && D <= CRT_F16_MAXD;
}

int crt_launch_factor16(hipStream_t st, const DevGrid &g16, int nsys, const int2 *sys, const double *fak,
                        const zd *vals, zd *panels, zd *y, zd *dinv, int *status) {
    Factor16Args a{g16, nsys, sys, fak, vals, panels, y, dinv, status};
    switch (g16.D) {
        case 1: return launch16<1>(st, a);
        case 2: return launch16<2>(st, a);
        case 3: return launch16<3>(st, a);
        case 4: return launch16<4>(st, a);
        case 5: return launch16<5>(st, a);
        case 6: return launch16<6>(st, a);
        case 7: return launch16<7>(st, a);
        case 8: return launch16<8>(st, a);
        case 9: return launch16<9>(st, a);
        case 10: return launch16<10>(st, a);
        default: return CRT_ERR_UNSUPPORTED;
    }
}

// ------------------------------------------------------------------- solve
// One workgroup (4 waves) per right-hand side.  Per block the off-diagonal
// tiles L(J, J-a) (forward) or L(J+a, J)^T (backward), a = 1..D, are split over
// waves 1..3, each prefetching its tiles for the next block while the current
// one is reduced; wave 0 substitutes through the diagonal tile with the
// reciprocal pivots of the factorisation (no divisions).  The forward pass
// starts at the electrode's block (the right-hand side is zero above it) and is
// skipped when the factorisation already produced y.  Lanes hold tile entries
// (r, c) = (l/16 + 4v, l%16), v = 0..3.
struct Solve16Args {
    int sanz, nblk, kwnanz, eanz, nrhs;
    const int4 *rhs;     // (l, k, factor system, have_y)
    const int *enr;
    const double *fak;
    const zd *panels;
    const zd *dinv;
    const zd *y;
    zd *yscr;
    zd *kpotN;
};

namespace {
constexpr int SW = 4;  // waves per right-hand side

template <int D>
__device__ __forceinline__ void solve16_fetch_diag(const zd *panel, const zd *dinv, int nblk, int J, int c,
                                                   zd (&drow)[T16], zd &pnext) {
    const int Jc = J < 0 ? 0 : (J < nblk ? J : nblk - 1);
    const zd *Ld = panel + (size_t)Jc * (D + 1) * 256 + c * T16;
#pragma unroll
    for (int cc = 0; cc < T16; ++cc) drow[cc] = zd{Ld[cc].re, Ld[cc].im};
    const zd *pv = dinv + (size_t)Jc * T16 + c;
    pnext = zd{pv->re, pv->im};
}

__device__ __forceinline__ void solve16_store_diag(zd *dtile, int lane, const zd (&drow)[T16]) {
    if (lane < T16)
#pragma unroll
        for (int cc = 0; cc < T16; ++cc) {
            dtile[lane * PS + cc].re = drow[cc].re;
            dtile[lane * PS + cc].im = drow[cc].im;
        }
}

template <int D>
__global__ void __launch_bounds__(64 * SW, D <= 6 ? 4 : 2) crt_solve16_kernel(Solve16Args A) {
    constexpr int NA = (D + SW - 2) / (SW - 1);  // off-diagonal tiles per worker wave
    __shared__ zd ring[(D + 1) * T16];
    __shared__ zd part[SW][T16];
    __shared__ zd dtile[T16 * PS];  // diagonal tile of the current block
    const int rid = blockIdx.x;
    if (rid >= A.nrhs) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = lane & 15, rq = lane >> 4;
    const int nblk = A.nblk;
    const int4 q4 = A.rhs[rid];
    const int l = q4.x, k = q4.y, fs = q4.z;
    const double *fak = A.fak + (size_t)fs * nblk * T16;
    const zd *panel = A.panels + (size_t)fs * nblk * (D + 1) * 256;
    const zd *dinv = A.dinv + (size_t)fs * nblk * T16;
    const int enode = A.enr[l];
    zd *yscr = A.yscr + (size_t)rid * nblk * T16;
    const int J0 = q4.w ? 0 : enode / T16;  // y is zero above the electrode's block

    // The two wave roles run separate loops with the same barrier sequence, so
    // their registers are allocated independently.
    if (wave == 0) {
        // ---- wave 0: substitution through the diagonal tiles
        zd pcur, drow[T16], pnext;
        if (!q4.w) {
            solve16_fetch_diag<D>(panel, dinv, nblk, J0, c, drow, pnext);
            solve16_store_diag(dtile, lane, drow);
            pcur = pnext;
            for (int J = J0; J < nblk; ++J) {
                solve16_fetch_diag<D>(panel, dinv, nblk, J + 1, c, drow, pnext);
                const int r = c;  // lane r = lane & 15 carries row r of the block
                const int i = J * T16 + r;
                zd yv = (i < A.sanz && i == enode) ? zd{-fak[i], 0.0} : zd{0.0, 0.0};
                __syncthreads();
#pragma unroll
                for (int w = 1; w < SW; ++w) yv = zsub(yv, part[w][r]);
#pragma unroll
                for (int cc = 0; cc < T16; ++cc) {
                    const zd ycc = zmul(readlane_z(yv, cc), readlane_z(pcur, cc));
                    if (r == cc) yv = ycc;
                    if (r > cc) yv = zfms(yv, dtile[r * PS + cc], ycc);
                }
                if (lane < T16) {
                    ring[(J % (D + 1)) * T16 + lane] = yv;
                    yscr[(size_t)J * T16 + lane] = yv;
                }
                __builtin_amdgcn_wave_barrier();
                solve16_store_diag(dtile, lane, drow);
                pcur = pnext;
                __syncthreads();
            }
        }
        const zd *yg = q4.w ? A.y + (size_t)fs * nblk * T16 : yscr;
        __syncthreads();  // the forward pass has finished with dtile
        solve16_fetch_diag<D>(panel, dinv, nblk, nblk - 1, c, drow, pnext);
        solve16_store_diag(dtile, lane, drow);
        pcur = pnext;
        zd ycur = nblk - 1 >= J0 ? yg[(size_t)(nblk - 1) * T16 + c] : zd{0.0, 0.0};
        for (int J = nblk - 1; J >= 0; --J) {
            solve16_fetch_diag<D>(panel, dinv, nblk, J - 1, c, drow, pnext);
            const zd ynxt = J - 1 >= J0 ? yg[(size_t)(J - 1) * T16 + c] : zd{0.0, 0.0};
            __syncthreads();
            zd xv = zd{-ycur.re, -ycur.im};  // lane c carries column c of the block
#pragma unroll
            for (int w = 1; w < SW; ++w) xv = zsub(xv, part[w][c]);
#pragma unroll
            for (int cc = T16 - 1; cc >= 0; --cc) {
                const zd xcc = zmul(readlane_z(xv, cc), readlane_z(pcur, cc));
                if (c == cc) xv = xcc;
                if (c < cc) xv = zfms(xv, dtile[cc * PS + c], xcc);
            }
            if (lane < T16) {
                ring[(J % (D + 1)) * T16 + lane] = xv;
                const int i = J * T16 + lane;
                if (i < A.sanz) {
                    const double f = fak[i];
                    A.kpotN[((size_t)i * A.kwnanz + k) * A.eanz + l] = zd{xv.re * f, xv.im * f};
                }
            }
            __builtin_amdgcn_wave_barrier();
            solve16_store_diag(dtile, lane, drow);
            pcur = pnext;
            ycur = ynxt;
            __syncthreads();
        }
        return;
    }

    // ---- waves 1..3: products with the off-diagonal tiles a = wave + 3u
    zd cur[NA][4];
    if (!q4.w) {
        auto fetch_f = [&](int J, zd (&t)[NA][4]) {
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                const int a = wave + (SW - 1) * u;
                const bool live = a <= D && J - a >= J0 && J < nblk;
                const zd *Lt = panel + ((size_t)(live ? J - a : 0) * (D + 1) + (live ? a : 0)) * 256;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const zd *p = Lt + (rq + 4 * v) * T16 + c;
                    t[u][v] = live ? zd{p->re, p->im} : zd{0.0, 0.0};
                }
            }
        };
        fetch_f(J0, cur);
        for (int J = J0; J < nblk; ++J) {
            zd nxt[NA][4];
            fetch_f(J + 1, nxt);
            // acc(r) = sum_a sum_c L(J, J-a)(r, c) y_{J-a}(c)
            zd acc[4] = {zd{0, 0}, zd{0, 0}, zd{0, 0}, zd{0, 0}};
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                const int a = wave + (SW - 1) * u;
                if (a <= D && J - a >= J0) {
                    const zd yv = ring[((J - a) % (D + 1)) * T16 + c];
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[v] = zfma(acc[v], cur[u][v], yv);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 1));
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 2));
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 4));
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 8));
                if (c == 0) part[wave][rq + 4 * v] = acc[v];
            }
            __syncthreads();
            __syncthreads();
#pragma unroll
            for (int u = 0; u < NA; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) cur[u][v] = nxt[u][v];
        }
    }
    auto fetch_b = [&](int J, zd (&t)[NA][4]) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = wave + (SW - 1) * u;
            const bool live = a <= D && J >= 0 && J + a < nblk;
            const zd *Lt = panel + ((size_t)(live ? J : 0) * (D + 1) + (live ? a : 0)) * 256;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const zd *p = Lt + (rq + 4 * v) * T16 + c;
                t[u][v] = live ? zd{p->re, p->im} : zd{0.0, 0.0};
            }
        }
    };
    fetch_b(nblk - 1, cur);
    __syncthreads();
    for (int J = nblk - 1; J >= 0; --J) {
        zd nxt[NA][4];
        fetch_b(J - 1, nxt);
        // acc(c) = sum_a sum_r L(J+a, J)(r, c) x_{J+a}(r)
        zd acc = zd{0.0, 0.0};
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = wave + (SW - 1) * u;
            if (a <= D && J + a < nblk) {
                const zd *xr = ring + ((J + a) % (D + 1)) * T16;
#pragma unroll
                for (int v = 0; v < 4; ++v) acc = zfma(acc, cur[u][v], xr[rq + 4 * v]);
            }
        }
        acc = zadd(acc, shfl_xor_z16(acc, 16));
        acc = zadd(acc, shfl_xor_z16(acc, 32));
        if (rq == 0) part[wave][c] = acc;
        __syncthreads();
        __syncthreads();
#pragma unroll
        for (int u = 0; u < NA; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) cur[u][v] = nxt[u][v];
    }
}

template <int D>
int launch_solve16(hipStream_t st, const Solve16Args &a) {
    hipLaunchKernelGGL(crt_solve16_kernel<D>, dim3(a.nrhs), dim3(64 * SW), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
}  // namespace

int crt_launch_solve16(hipStream_t st, int sanz, int nblk, int D, int kwnanz, int eanz, int nrhs, const int4 *rhs,
                       const int *enr, const double *fak, const zd *panels, const zd *dinv, const zd *y, zd *yscr,
                       zd *kpotN) {
    Solve16Args a{sanz, nblk, kwnanz, eanz, nrhs, rhs, enr, fak, panels, dinv, y, yscr, kpotN};
    switch (D) {
        case 1: return launch_solve16<1>(st, a);
        case 2: return launch_solve16<2>(st, a);
        case 3: return launch_solve16<3>(st, a);
        case 4: return launch_solve16<4>(st, a);
        case 5: return launch_solve16<5>(st, a);
        case 6: return launch_solve16<6>(st, a);
        case 7: return launch_solve16<7>(st, a);
        case 8: return launch_solve16<8>(st, a);
        case 9: return launch_solve16<9>(st, a);
        case 10: return launch_solve16<10>(st, a);
        default: return CRT_ERR_UNSUPPORTED;
    }
}
