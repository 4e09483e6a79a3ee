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

#include <type_traits>

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int T16 = 16;
constexpr int PS = 17;  // padded row stride (complex) of LDS panel tiles: conflict-free column reads

// Workgroup barrier that orders LDS only: global loads stay in flight across it and
// global stores drain in the background.  For phases whose hand-offs between waves
// go through LDS alone (global data written in the kernel is not read back by
// another wave before a full __syncthreads()).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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

// r = 1/p = conj(p)/|p|^2 with the hardware reciprocal and two Newton steps;
// bad as in pivot_fast
__device__ __forceinline__ zd crecip(zd p, bool &bad) {
    const double m2 = fma(p.re, p.re, p.im * p.im);
    bad = !(m2 > 0.0) || !(m2 < 1.0e300);
    double y = __builtin_amdgcn_rcp(m2);
    double e = fma(-m2, y, 1.0);
    y = fma(y, e, y);
    e = fma(-m2, y, 1.0);
    y = fma(y, e, y);
    return zd{p.re * y, -p.im * y};
}

__device__ __forceinline__ zd readlane_z(zd v, int src) { return zd{readlane_d(v.re, src), readlane_d(v.im, src)}; }
// 16-byte loads and stores of a complex value (the address must be 16-byte aligned)
__device__ __forceinline__ zd ldz(const zd *p) {
    const double2 v = *reinterpret_cast<const double2 *>(p);
    return zd{v.x, v.y};
}
__device__ __forceinline__ void stz(zd *p, zd z) { *reinterpret_cast<double2 *>(p) = make_double2(z.re, z.im); }

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
struct alignas(16) F16Lds {
    zd P[(D + 1) * T16 * PS];         // panel tiles of block column J, 16 x PS each
    zd U[T16 * PS];                   // unit lower factor u(i,c) = a(i,c)/p_c of the diagonal tile
    zd S[(D + 1) * 256];              // staging tiles of the entering row block, 16 x 16 each
    zd bring[(D + 2) * T16];          // right-hand-side blocks J .. J+D+1
    zd xcol[Cfg<D>::W][2][T16];       // per-wave column exchange of the panel solve
    zd col[32];                       // current column of the diagonal tile, [16] = its RHS entry
    zd pinv[T16];                     // 1/L(c,c)
    zd sdiag[T16];                    // L(c,c)
    zd yj[T16];                       // forward-substituted right-hand side of block J
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

// (dr, di) = T - Pa Pb^T for 16x16 complex tiles in LDS (row stride PS), T and
// the result in the accumulator layout; the result may go to another slot, so
// the window slides without register moves
__device__ __forceinline__ void mfma_update(d4 &dr, d4 &di, const d4 &tr, const d4 &ti, const zd *Pa, const zd *Pb,
                                            int jc, int rq) {
    zd va[4], vb[4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        va[kc] = ldz(Pa + jc * PS + 4 * kc + rq);
        vb[kc] = ldz(Pb + jc * PS + 4 * kc + rq);
    }
    d4 xr = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[0].re, vb[0].re, tr, 0, 0, 0);
    d4 xi = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[0].re, vb[0].im, ti, 0, 0, 0);
    xr = __builtin_amdgcn_mfma_f64_16x16x4f64(va[0].im, vb[0].im, xr, 0, 0, 0);
    xi = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[0].im, vb[0].re, xi, 0, 0, 0);
#pragma unroll
    for (int kc = 1; kc < 4; ++kc) {
        xr = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[kc].re, vb[kc].re, xr, 0, 0, 0);
        xi = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[kc].re, vb[kc].im, xi, 0, 0, 0);
        xr = __builtin_amdgcn_mfma_f64_16x16x4f64(va[kc].im, vb[kc].im, xr, 0, 0, 0);
        xi = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[kc].im, vb[kc].re, xi, 0, 0, 0);
    }
    dr = xr;
    di = xi;
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
    const zd *vals = A.vals + (size_t)s * g.vstride;
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
        // ---------------- B: wave 0 factors the diagonal tile | others stage Ie
        if (WV == 0 && J >= 0) {
            // Right-looking elimination of the 16x16 diagonal tile in the result
            // layout with the square roots deferred: the update uses the pivots p_c
            // as a(i,j) -= a(i,c) a(j,c)/p_c, so only a complex reciprocal sits on
            // the column-to-column chain; the next pivot is formed from two lane
            // reads before the column exchange.  L(i,c) = a(i,c)/sqrt(p_c) at the end.
            zd a[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) a[v] = zd{Ar[0][v], Ai[0][v]};
            zd yv = m.bring[(J % (D + 2)) * T16 + jc];  // lanes of column group jc carry b(jc)
            zd p = readlane_z(a[0], 0);
            bool bad;
            zd r = crecip(p, bad);
            if (bad) err = CRT_ERR_PIVOT;
            zd pk = zsel(lane == 0, p, zd{0.0, 0.0});  // lane c keeps p_c
#pragma unroll
            for (int c = 0; c < T16; ++c) {
                if (jc == c) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) m.col[4 * v + rq] = a[v];
                    if (rq == 0) m.col[16] = yv;
                }
                zd dn, l2;
                if (c + 1 < T16) {
                    const int n = c + 1;
                    dn = readlane_z(a[n / 4], n + 16 * (n % 4));       // a(n, n)
                    const zd ln = readlane_z(a[n / 4], c + 16 * (n % 4));  // a(n, c)
                    l2 = zmul(ln, ln);
                }
                __builtin_amdgcn_wave_barrier();
                zd ci[4];
                const zd cj = m.col[jc], wc = m.col[16];
#pragma unroll
                for (int v = 0; v < 4; ++v) ci[v] = m.col[4 * v + rq];
                __builtin_amdgcn_sched_barrier(0);  // one exchange: issue every read before the first use
                const zd t = zsel(jc > c, zmul(cj, r), zd{0.0, 0.0});  // u(jc, c)
#pragma unroll
                for (int v = 0; v < 4; ++v) a[v] = zfms(a[v], ci[v], t);
                yv = zfms(yv, wc, t);
                if (rq == 0) m.U[jc * PS + c] = t;
                if (c + 1 < T16) {
                    p = zfms(dn, l2, r);
                    r = crecip(p, bad);
                    if (bad) err = CRT_ERR_PIVOT;
                    pk = zsel(lane == c + 1, p, pk);
                }
                __builtin_amdgcn_wave_barrier();
            }
            // square roots of the pivots (lane c), then scale the columns
            zd sc, ivc;
            bool badc;
            pivot_fast(pk, sc, ivc, badc);
            if (lane < T16) {
                m.pinv[lane] = ivc;
                m.sdiag[lane] = sc;
                A.dinv[(size_t)s * nblk * T16 + (size_t)J * T16 + lane] = ivc;
            }
            __builtin_amdgcn_wave_barrier();
            const zd ivj = m.pinv[jc], sj = m.sdiag[jc];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 4 * v + rq;
                m.P[i * PS + jc] = i > jc ? zmul(a[v], ivj) : (i == jc ? sj : zd{0.0, 0.0});
            }
            if (rq == 0) {
                const zd y = zmul(yv, ivj);
                m.yj[jc] = y;
                yout[(size_t)J * T16 + jc] = y;
            }
        } else if (Ie < nblk) {
            if (J < 0) stage_block<D>(g, Ie, tid, NT, vals, fak, enode, m);
            else stage_block<D>(g, Ie, tid - 64, NT - 64, vals, fak, enode, m);
        }
        __syncthreads();
        // ---- C: each wave solves its panel tiles X = A L^-T in registers, RHS update
        if (WV >= 1 && J >= 0) {
            // Y = A U^-T column by column (unit diagonal), X = Y diag(1/L(c,c))
            zd xa[4], xb[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                xa[v] = zd{Ar[0][v], Ai[0][v]};
                xb[v] = zd{Br[0][v], Bi[0][v]};
            }
            zd *xca = m.xcol[WV][0], *xcb = m.xcol[WV][1];
#pragma unroll
            for (int c = 0; c < T16; ++c) {
                if (jc == c) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        xca[4 * v + rq] = xa[v];
                        if constexpr (LB > 0) xcb[4 * v + rq] = xb[v];
                    }
                }
                __builtin_amdgcn_wave_barrier();
                const zd u = m.U[jc * PS + c];
                zd ya[4], yb[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    ya[v] = xca[4 * v + rq];
                    if constexpr (LB > 0) yb[v] = xcb[4 * v + rq];
                }
                __builtin_amdgcn_sched_barrier(0);  // one exchange: issue every read before the first use
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    xa[v] = zfms(xa[v], ya[v], u);
                    if constexpr (LB > 0) xb[v] = zfms(xb[v], yb[v], u);
                }
                __builtin_amdgcn_wave_barrier();
            }
            const zd ivj = m.pinv[jc];
            zd *pa = m.P + dA * T16 * PS;
#pragma unroll
            for (int v = 0; v < 4; ++v) pa[(4 * v + rq) * PS + jc] = zmul(xa[v], ivj);
            if constexpr (LB > 0) {
                zd *pb = m.P + dB * T16 * PS;
#pragma unroll
                for (int v = 0; v < 4; ++v) pb[(4 * v + rq) * PS + jc] = zmul(xb[v], ivj);
            }
            __builtin_amdgcn_wave_barrier();
            // b(J+d) -= X(J+d, J) y_J: lanes 0..15 for the A tile, 16..31 for the B tile
            const int half = lane >> 4, rr = lane & 15;
            const int d = half == 0 ? dA : dB;
            if ((half == 0 || (LB > 0 && half == 1)) && J + d < nblk) {
                const zd *row = m.P + (d * T16 + rr) * PS;
                zd *bp = m.bring + ((J + d) % (D + 2)) * T16 + rr;
                zd acc = *bp;
#pragma unroll
                for (int c = 0; c < T16; ++c) acc = zfms(acc, row[c], m.yj[c]);
                *bp = acc;
            }
        }
        __syncthreads();
        // ------------------- D: store panel, rank-16 updates on the matrix cores, slide
        if (J >= 0) {
            zd *pj = panel + (size_t)J * (D + 1) * 256;
            constexpr int NPS = ((D + 1) * 256 + NT - 1) / NT;
            zd pv[NPS];
#pragma unroll
            for (int u = 0; u < NPS; ++u) {
                const int t = tid + u * NT;
                const int a = t >> 8, rc = t & 255;
                pv[u] = (t < (D + 1) * 256 && J + a < nblk) ? m.P[(a * T16 + (rc >> 4)) * PS + (rc & 15)]
                                                             : zd{0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < NPS; ++u)
                if (tid + u * NT < (D + 1) * 256) pj[tid + u * NT] = pv[u];
        }
        // position p -> p-1: each rank-16 update writes its result into the slot
        // of the next position; the entering tiles take the last positions
#pragma unroll
        for (int p = 1; p < LA; ++p) {
            if (J >= 0 && J + p + dA < nblk)
                mfma_update(Ar[p - 1], Ai[p - 1], Ar[p], Ai[p], m.P + (p + dA) * T16 * PS, m.P + p * T16 * PS, jc,
                            rq);
            else {
                Ar[p - 1] = Ar[p];
                Ai[p - 1] = Ai[p];
            }
        }
        {
            zd *sa = m.S + dA * 256;
            Ar[LA - 1] = tile_re(sa, jc, rq);
            Ai[LA - 1] = tile_im(sa, jc, rq);
#pragma unroll
            for (int v = 0; v < 4; ++v) sa[(4 * v + rq) * T16 + jc] = zd{0.0, 0.0};
        }
        if constexpr (LB > 0) {
#pragma unroll
            for (int p = 1; p < LB; ++p) {
                if (J >= 0 && J + p + dB < nblk)
                    mfma_update(Br[p - 1], Bi[p - 1], Br[p], Bi[p], m.P + (p + dB) * T16 * PS, m.P + p * T16 * PS,
                                jc, rq);
                else {
                    Br[p - 1] = Br[p];
                    Bi[p - 1] = Bi[p];
                }
            }
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

static_assert(sizeof(F16Lds<6>) <= 80 * 1024, "two workgroups per CU at D = 6");

// ---------------------------------------------------------------------------
// Look-ahead variant (D >= 2), 1 + D/2 + (D odd) wavefronts, two workgroups per CU.
//   wave 0       pivot wave: takes the diagonal tile (J+1,J+1) handed over by
//                wave 1, applies its last rank-16 update X(J+1,J) X(J+1,J)^T and
//                factors it (deferred square roots; L^-1 from the same column
//                operations applied to an identity block) while the other waves
//                update the window for panel J;
//   waves 1..    diagonals g and D+1-g (pairs) or the middle diagonal, all
//                positions, plus a contiguous share of diagonal 0, positions
//                2..D; the lowest diagonal-0 tile of each share moves to the next
//                lower wave (wave 1: to the pivot wave) through an LDS hand-off.
// Per block column J: S1 -- the owners of the panel tiles compute X = A (L^-1)^T
// on the matrix cores and update the right-hand side, the receivers take their
// hand-off tiles; S2 -- rank-16 updates and the slide on waves 1.., the
// diagonal factorisation of block J+1 on wave 0.  Entering tiles are gathered
// from the CSR values through the tile map, so the window needs no staging LDS.
namespace {

template <int D>
struct La4 {
    static constexpr int NPAIR = D / 2;
    static constexpr int MID = D & 1;
    static constexpr int NU = NPAIR + MID;  // update waves 1..NU
    static constexpr int W = NU + 1;
    static constexpr int NT = 64 * W;
    static constexpr int NZ = D - 1;        // diagonal-0 positions 2..D
    static constexpr int zcount(int u) { return NZ / NU + ((u - 1) < NZ % NU ? 1 : 0); }
    static constexpr int zlo(int u) { return u == 1 ? 2 : zlo(u - 1) + zcount(u - 1); }
};

template <int D>
struct alignas(16) La4Lds {
    zd P[D * T16 * PS];       // X(J+d, J), d = 1..D
    zd A1[T16 * PS];          // A(J+1, J) from wave 1, then X(J+1, J) on the pivot wave
    zd Li[T16 * PS];          // L(J,J)^-1, lower, row-major
    zd HO[La4<D>::NU][256];   // hand-off tiles: HO[u-1] from wave u to wave u-1 (u = 1: to wave 0)
    zd bring[(D + 2) * T16];
    zd col[48];
    zd pinv[T16], sdiag[T16];
    zd yj[T16];
};

__device__ __forceinline__ void mfma_mul(d4 &xr, d4 &xi, const zd *Pa, const zd *Pb, int jc, int rq) {
    zd va[4], vb[4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        va[kc] = ldz(Pa + jc * PS + 4 * kc + rq);
        vb[kc] = ldz(Pb + jc * PS + 4 * kc + rq);
    }
    const d4 z = d4{0.0, 0.0, 0.0, 0.0};
    xr = __builtin_amdgcn_mfma_f64_16x16x4f64(va[0].re, vb[0].re, z, 0, 0, 0);
    xi = __builtin_amdgcn_mfma_f64_16x16x4f64(va[0].re, vb[0].im, z, 0, 0, 0);
    xr = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[0].im, vb[0].im, xr, 0, 0, 0);
    xi = __builtin_amdgcn_mfma_f64_16x16x4f64(va[0].im, vb[0].re, xi, 0, 0, 0);
#pragma unroll
    for (int kc = 1; kc < 4; ++kc) {
        xr = __builtin_amdgcn_mfma_f64_16x16x4f64(va[kc].re, vb[kc].re, xr, 0, 0, 0);
        xi = __builtin_amdgcn_mfma_f64_16x16x4f64(va[kc].re, vb[kc].im, xi, 0, 0, 0);
        xr = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[kc].im, vb[kc].im, xr, 0, 0, 0);
        xi = __builtin_amdgcn_mfma_f64_16x16x4f64(va[kc].im, vb[kc].re, xi, 0, 0, 0);
    }
}

template <int D>
__device__ __forceinline__ zd *la4_tile(La4Lds<D> &m, int d) { return m.P + (d - 1) * T16 * PS; }

// map entries (CSR indices) of this lane's four elements of tile (I, I-d)
template <int D>
__device__ __forceinline__ void la4_map(const DevGrid &g, const int *tm, int I, int d, int jc, int rq, int (&mi)[4]) {
    const int *mp = tm + ((size_t)(I * (D + 1) + d) * T16) * T16 + jc;
#pragma unroll
    for (int v = 0; v < 4; ++v) mi[v] = I < g.nblk ? mp[(4 * v + rq) * T16] : g.nnz;
}

// tile values in the accumulator layout from the map entries
__device__ __forceinline__ void la4_vals(const zd *vals, const int (&mi)[4], d4 &tr, d4 &ti) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        // plain loads (zero and unit entries are values too): nothing waits on them here
        const zd e = ldz(vals + mi[v]);
        tr[v] = e.re;
        ti[v] = e.im;
    }
}

// diagonal factorisation of block Jn on the pivot wave, in three parts so the
// column loop can be split across a barrier: init, columns [C0, C1), finish
struct DiagChain {
    zd a[4], z[4], yv, p, r, pk;
};

__device__ __forceinline__ void diag_init(DiagChain &q, int lane, const d4 &tr, const d4 &ti, zd yv, int &err) {
    const int jc = lane & 15, rq = lane >> 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        q.a[v] = zd{tr[v], ti[v]};
        q.z[v] = zd{4 * v + rq == jc ? 1.0 : 0.0, 0.0};
    }
    q.yv = yv;
    q.p = readlane_z(q.a[0], 0);
    bool bad;
    q.r = crecip(q.p, bad);
    if (bad) err = CRT_ERR_PIVOT;
    q.pk = zsel(lane == 0, q.p, zd{0.0, 0.0});
}

template <int C0, int C1, class LDS>
__device__ __forceinline__ void diag_cols(DiagChain &q, LDS &m, int lane, int &err) {
    const int jc = lane & 15, rq = lane >> 4;
#pragma unroll
    for (int c = C0; c < C1; ++c) {
        if (jc == c) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                stz(m.col + 4 * v + rq, q.a[v]);
                stz(m.col + 16 + 4 * v + rq, q.z[v]);
            }
            if (rq == 0) stz(m.col + 32, q.yv);
        }
        zd dn, l2;
        if (c + 1 < T16) {
            const int n = c + 1;
            dn = readlane_z(q.a[n / 4], n + 16 * (n % 4));
            const zd ln = readlane_z(q.a[n / 4], c + 16 * (n % 4));
            l2 = zmul(ln, ln);
        }
        __builtin_amdgcn_wave_barrier();
        // Column c changes a(i, jc) and z(i, jc) for jc > c only. Rows i <= c of a are
        // never read again (the factor keeps i > jc), and z(i, c) is zero for i > c, so
        // element v of a is skipped once 4v+3 <= c and element v of z while 4v > c.
        zd ci[4], zi[4];
        const zd cj = ldz(m.col + jc), wc = ldz(m.col + 32);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (4 * v + 3 > c) ci[v] = ldz(m.col + 4 * v + rq);
            if (4 * v <= c) zi[v] = ldz(m.col + 16 + 4 * v + rq);
        }
        __builtin_amdgcn_sched_barrier(0);
        const zd t = zsel(jc > c, zmul(cj, q.r), zd{0.0, 0.0});  // u(jc, c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (4 * v + 3 > c) q.a[v] = zfms(q.a[v], ci[v], t);
            if (4 * v <= c) q.z[v] = zfms(q.z[v], zi[v], t);
        }
        q.yv = zfms(q.yv, wc, t);
        if (c + 1 < T16) {
            bool bad;
            q.p = zfms(dn, l2, q.r);
            q.r = crecip(q.p, bad);
            if (bad) err = CRT_ERR_PIVOT;
            q.pk = zsel(lane == c + 1, q.p, q.pk);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <int D, class LDS>
__device__ __forceinline__ void diag_finish(const Factor16Args &A, LDS &m, int s, int Jn, int lane, DiagChain &q) {
    const int nblk = A.g.nblk;
    const int jc = lane & 15, rq = lane >> 4;
    zd sc, ivc;
    bool badc;
    pivot_fast(q.pk, sc, ivc, badc);
    if (lane < T16) {
        m.pinv[lane] = ivc;
        m.sdiag[lane] = sc;
        A.dinv[(size_t)s * nblk * T16 + (size_t)Jn * T16 + lane] = ivc;
    }
    __builtin_amdgcn_wave_barrier();
    const zd ivj = m.pinv[jc], sj = m.sdiag[jc];
    zd *pg = A.panels + ((size_t)s * nblk + Jn) * (D + 1) * 256;
    // the diagonal tile keeps L on and below the diagonal and, above it, the strict
    // lower triangle of L^-1 transposed (tile(i, jc) = L^-1(jc, i), i < jc): the
    // substitution applies L^-T as a product instead of a serial chain
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = 4 * v + rq;
        const zd li = zmul(q.z[v], ivj);  // L^-1(jc, i) = (L^-T)(i, jc)
        stz(pg + i * T16 + jc, i > jc ? zmul(q.a[v], ivj) : (i == jc ? sj : li));
        stz(m.Li + jc * PS + i, li);
    }
    if (rq == 0) {
        const zd y = zmul(q.yv, ivj);
        m.yj[jc] = y;
        A.y[(size_t)s * nblk * T16 + (size_t)Jn * T16 + jc] = y;
    }
}

template <int D, class LDS>
__device__ __forceinline__ void la4_factor_diag(const Factor16Args &A, LDS &m, int s, int Jn, int lane,
                                                const d4 &tr, const d4 &ti, zd yv, int &err) {
    DiagChain q;
    diag_init(q, lane, tr, ti, yv, err);
    diag_cols<0, T16>(q, m, lane, err);
    diag_finish<D>(A, m, s, Jn, lane, q);
}

// columns of the next diagonal block the pivot wave factorises before the first
// barrier of a step (the rest after it)
constexpr int LA4_SPLIT = 4;

template <int D, int WV>
__device__ __forceinline__ void la4_wave(const Factor16Args &A, La4Lds<D> &m) {
    using C = La4<D>;
    constexpr int NT = C::NT;
    constexpr bool isP = WV == 0;
    constexpr bool isPair = WV >= 1 && WV <= C::NPAIR;
    constexpr bool isMid = C::MID && WV == C::NPAIR + 1;
    constexpr int dA = isPair ? WV : (isMid ? (D + 1) / 2 : 1);
    constexpr int LA = isPair ? D + 1 - WV : (isMid ? (D + 1) / 2 : 1);
    constexpr int dB = isPair ? D + 1 - WV : 0;
    constexpr int LB = isPair ? WV : 0;
    constexpr int LBX = LB > 0 ? LB : 1;
    constexpr int NZW = isP ? 0 : C::zcount(WV);     // diagonal-0 slots of this wave
    constexpr int ZLO = isP ? 0 : C::zlo(WV);        // position of slot 0
    constexpr int NZX = NZW > 0 ? NZW : 1;
    constexpr bool topZ = !isP && WV == C::NU;       // holds position D of diagonal 0
    const DevGrid &g = A.g;
    const int nblk = g.nblk;
    const int s = blockIdx.x;
    const int lane = threadIdx.x & 63, tid = WV * 64 + lane;
    const int enode = g.enr[A.sys[s].x];
    const int *tm = g.tmap + (size_t)A.sys[s].x * g.tmap_sys;
    const double *fak = A.fak + (size_t)s * nblk * T16;
    const zd *vals = A.vals + (size_t)s * g.vstride;
    zd *panel = A.panels + (size_t)s * nblk * (D + 1) * 256;
    const int jc = lane & 15, rq = lane >> 4;

    d4 Ar[LA], Ai[LA], Br[LBX], Bi[LBX], Zr[NZX], Zi[NZX];
#pragma unroll
    for (int q = 0; q < LA; ++q) Ar[q] = Ai[q] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < LBX; ++q) Br[q] = Bi[q] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < NZX; ++q) Zr[q] = Zi[q] = d4{0.0, 0.0, 0.0, 0.0};
    d4 Tr = d4{0.0, 0.0, 0.0, 0.0}, Ti = Tr;  // pivot wave: tile (J+1, J+1)
    DiagChain chain;                           // pivot wave: its factorisation across S1 and S2
    int err = 0;
    // map entries of the entering tiles, loaded one block row ahead
    int mA[4], mB[4], mZ[4];
    if constexpr (!isP) {
        la4_map<D>(g, tm, 0, dA, jc, rq, mA);
        if constexpr (LB > 0) la4_map<D>(g, tm, 0, dB, jc, rq, mB);
        if constexpr (topZ) la4_map<D>(g, tm, 0, 0, jc, rq, mZ);
    }

    for (int J = -(D + 1); J < nblk; ++J) {
        const int Ie = J + D + 1;
        // ================================================================ S1
        if constexpr (isP) {
            if (lane < T16) {
                const int i = Ie * T16 + lane;
                m.bring[(Ie % (D + 2)) * T16 + lane] =
                    (i < g.sanz && i == enode) ? zd{-fak[i], 0.0} : zd{0.0, 0.0};
            }
            const zd *h = m.HO[0];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const zd e = ldz(h + (4 * v + rq) * T16 + jc);
                Tr[v] = e.re;
                Ti[v] = e.im;
            }
            const int Jn = J + 1;
            if (Jn >= 0 && Jn < nblk) {
                zd *brow = m.bring + (Jn % (D + 2)) * T16;
                if (J >= 0) {
                    // X(J+1, J) = A(J+1, J) (L^-1)^T from the tile wave 1 left in A1 (the same
                    // product wave 1 forms for the panel), then b(J+1) -= X y_J and the
                    // rank-16 update of the diagonal tile
                    d4 xr, xi;
                    mfma_mul(xr, xi, m.A1, m.Li, jc, rq);
                    __builtin_amdgcn_wave_barrier();
#pragma unroll
                    for (int v = 0; v < 4; ++v) stz(m.A1 + (4 * v + rq) * PS + jc, zd{xr[v], xi[v]});
                    __builtin_amdgcn_wave_barrier();
                    if (lane < T16) {
                        const zd *row = m.A1 + lane * PS;
                        zd acc = ldz(brow + lane);
#pragma unroll
                        for (int c = 0; c < T16; ++c) acc = zfms(acc, ldz(row + c), ldz(m.yj + c));
                        stz(brow + lane, acc);
                    }
                    mfma_update(Tr, Ti, Tr, Ti, m.A1, m.A1, jc, rq);
                    __builtin_amdgcn_wave_barrier();
                }
                diag_init(chain, lane, Tr, Ti, ldz(brow + jc), err);
                diag_cols<0, LA4_SPLIT>(chain, m, lane, err);
            }
        } else {
            if (J >= 0) {
                // X = A (L^-1)^T for the panel tiles at position 0, then b(J+d) -= X y_J
                const bool liveA = J + dA < nblk, liveB = LB > 0 && J + dB < nblk;
                zd *pa = la4_tile<D>(m, dA);
                zd *pb = la4_tile<D>(m, LB > 0 ? dB : dA);
                if (liveA)
#pragma unroll
                    for (int v = 0; v < 4; ++v) stz(pa + (4 * v + rq) * PS + jc, zd{Ar[0][v], Ai[0][v]});
                if constexpr (LB > 0)
                    if (liveB)
#pragma unroll
                        for (int v = 0; v < 4; ++v) stz(pb + (4 * v + rq) * PS + jc, zd{Br[0][v], Bi[0][v]});
                __builtin_amdgcn_wave_barrier();
                d4 xar, xai, xbr, xbi;
                if (liveA) mfma_mul(xar, xai, pa, m.Li, jc, rq);
                if constexpr (LB > 0)
                    if (liveB) mfma_mul(xbr, xbi, pb, m.Li, jc, rq);
                __builtin_amdgcn_wave_barrier();
                // panel tiles X(J+d, J) to LDS for the updates and straight to the panel in
                // global memory (rows 4v..4v+3 of a tile are 64 consecutive entries)
                zd *pj = panel + (size_t)J * (D + 1) * 256 + rq * T16 + jc;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const zd xa = liveA ? zd{xar[v], xai[v]} : zd{0.0, 0.0};
                    stz(pa + (4 * v + rq) * PS + jc, xa);
                    stz(pj + dA * 256 + 64 * v, xa);
                    if constexpr (LB > 0) {
                        const zd xb = liveB ? zd{xbr[v], xbi[v]} : zd{0.0, 0.0};
                        stz(pb + (4 * v + rq) * PS + jc, xb);
                        stz(pj + dB * 256 + 64 * v, xb);
                    }
                }
                __builtin_amdgcn_wave_barrier();
                const int half = lane >> 4, rr = lane & 15;
                const int d = half == 0 ? dA : dB;
                // b(J+1) is left to the pivot wave, which needs it first
                if ((half == 0 && liveA && dA != 1) || (LB > 0 && half == 1 && liveB)) {
                    const zd *row = la4_tile<D>(m, d) + rr * PS;
                    zd *bp = m.bring + ((J + d) % (D + 2)) * T16 + rr;
                    zd acc = ldz(bp);
#pragma unroll
                    for (int c = 0; c < T16; ++c) acc = zfms(acc, ldz(row + c), ldz(m.yj + c));
                    stz(bp, acc);
                }
            }
            if constexpr (!topZ) {
                // the top diagonal-0 slot takes the tile handed down by wave WV+1
                const zd *h = m.HO[WV];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const zd e = ldz(h + (4 * v + rq) * T16 + jc);
                    Zr[NZX - 1][v] = e.re;
                    Zi[NZX - 1][v] = e.im;
                }
            }
        }
        lds_barrier();
        // ================================================================ S2
        if constexpr (isP) {
            const int Jn = J + 1;
            if (Jn >= 0 && Jn < nblk) {
                diag_cols<LA4_SPLIT, T16>(chain, m, lane, err);
                diag_finish<D>(A, m, s, Jn, lane, chain);
            }
        } else {
            // entering tiles of block row Ie (positions D-d at the next step)
            d4 eAr, eAi, eBr, eBi, eZr, eZi;
            la4_vals(vals, mA, eAr, eAi);
            if constexpr (LB > 0) la4_vals(vals, mB, eBr, eBi);
            if constexpr (topZ) la4_vals(vals, mZ, eZr, eZi);
            la4_map<D>(g, tm, Ie + 1, dA, jc, rq, mA);
            if constexpr (LB > 0) la4_map<D>(g, tm, Ie + 1, dB, jc, rq, mB);
            if constexpr (topZ) la4_map<D>(g, tm, Ie + 1, 0, jc, rq, mZ);
            // diagonal 0: slot q holds position ZLO+q; position ZLO goes down to wave WV-1
            {
                zd *h = m.HO[WV - 1];
                d4 hr = Zr[0], hi = Zi[0];
                if (J >= 0 && J + ZLO < nblk) {
                    const zd *xp = la4_tile<D>(m, ZLO);
                    mfma_update(hr, hi, Zr[0], Zi[0], xp, xp, jc, rq);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) stz(h + (4 * v + rq) * T16 + jc, zd{hr[v], hi[v]});
            }
#pragma unroll
            for (int q = 1; q < NZW; ++q) {
                const int pp = ZLO + q;
                if (J >= 0 && J + pp < nblk) {
                    const zd *xp = la4_tile<D>(m, pp);
                    mfma_update(Zr[q - 1], Zi[q - 1], Zr[q], Zi[q], xp, xp, jc, rq);
                } else {
                    Zr[q - 1] = Zr[q];
                    Zi[q - 1] = Zi[q];
                }
            }
            if constexpr (topZ) {
                Zr[NZX - 1] = eZr;
                Zi[NZX - 1] = eZi;
            }
            // own diagonals: position p -> p-1, the entering tile takes the last position
#pragma unroll
            for (int p = 1; p < LA; ++p) {
                if (J >= 0 && J + p + dA < nblk)
                    mfma_update(Ar[p - 1], Ai[p - 1], Ar[p], Ai[p], la4_tile<D>(m, p + dA), la4_tile<D>(m, p), jc,
                                rq);
                else {
                    Ar[p - 1] = Ar[p];
                    Ai[p - 1] = Ai[p];
                }
            }
            Ar[LA - 1] = eAr;
            Ai[LA - 1] = eAi;
            if constexpr (LB > 0) {
#pragma unroll
                for (int p = 1; p < LB; ++p) {
                    if (J >= 0 && J + p + dB < nblk)
                        mfma_update(Br[p - 1], Bi[p - 1], Br[p], Bi[p], la4_tile<D>(m, p + dB), la4_tile<D>(m, p),
                                    jc, rq);
                    else {
                        Br[p - 1] = Br[p];
                        Bi[p - 1] = Bi[p];
                    }
                }
                Br[LB - 1] = eBr;
                Bi[LB - 1] = eBi;
            }
            if constexpr (dA == 1) {
                // A(J+2, J+1) is final now: hand it to the pivot wave for the next step
#pragma unroll
                for (int v = 0; v < 4; ++v) stz(m.A1 + (4 * v + rq) * PS + jc, zd{Ar[0][v], Ai[0][v]});
            }
        }
        lds_barrier();
    }
    if (isP && lane == 0 && err && A.status[s] == 0) A.status[s] = err;
}

template <int D, int WV>
__device__ __forceinline__ void la4_dispatch(int wave, const Factor16Args &A, La4Lds<D> &m) {
    if constexpr (WV < La4<D>::W) {
        if (wave == WV) la4_wave<D, WV>(A, m);
        else la4_dispatch<D, WV + 1>(wave, A, m);
    }
}

}  // namespace

template <int D>
__global__ void __launch_bounds__(La4<D>::NT, 2) crt_factor16la_kernel(Factor16Args A) {
    __shared__ La4Lds<D> m;
    if ((int)blockIdx.x >= A.nsys) return;
    for (int t = threadIdx.x; t < La4<D>::NU * 256; t += La4<D>::NT) m.HO[t >> 8][t & 255] = zd{0.0, 0.0};
    __syncthreads();
    la4_dispatch<D, 0>(threadIdx.x >> 6, A, m);
}

// ---------------------------------------------------------------------------
// Streaming left-looking variant for wide bands (D = 11..16), where the window
// of D(D+1)/2 tiles no longer fits on one CU.  Per block column J:
//   stage   L(J, J-m), m = 1..D (row block J of the previous panels) into LDS;
//   update  C(a) = A(J+a, J) - sum_m L(J+a, J-m) L(J, J-m)^T on the matrix cores:
//           the pivot wave takes C(0) and the factor's right-hand side, three update
//           waves the off-diagonal tiles, whose operands L(J+a, J-m) stream from the
//           panels with the next operand's load in flight during each product;
//   factor  C(0) on the pivot wave (as in the look-ahead variant: L, L^-1, y), as
//           soon as its own products are done, while the update waves still run;
//   solve   X(a) = C(a) (L^-1)^T on the matrix cores, stored as panel column J.
// Four waves and at most 256 VGPRs per lane keep two systems on each CU, so one
// system's serial diagonal factorisation overlaps the other's products.
namespace {
constexpr int SWV = 4;        // wavefronts of the streaming variant
constexpr int SNU = SWV - 1;  // update waves

template <int D>
struct StCfg {
    static constexpr int NA = (D + SNU - 1) / SNU;  // off-diagonal tiles per update wave
    // tile a (1..D, larger = void) of update wave w (1..SNU) in round u, in snake order:
    // tile a takes D-a products, so each wave gets long and short chains alike
    static constexpr int tile(int w, int u) { return SNU * u + 1 + ((u & 1) ? (SNU - w) : (w - 1)); }
};

template <int D>
struct alignas(16) StLds {
    static_assert(D >= SNU, "solve staging fits in Lrow");
    zd Lrow[D * T16 * PS];  // L(J, J-m), m = 1..D; the first SNU tiles stage C(a) in the solve
    zd Li[T16 * PS];
    zd col[48];
    zd pinv[T16], sdiag[T16];
    zd yj[T16];
    zd yr[(D + 1) * T16];   // y of the last D+1 block rows (ring), for the right-hand side
};

// row block J of panel columns J-m into LDS (all SWV*64 threads, one element each per tile)
template <int D>
__device__ __forceinline__ void st_stage(const zd *panel, StLds<D> &m, int J, int tid) {
    zd v[D];
#pragma unroll
    for (int mm = 1; mm <= D; ++mm)
        v[mm - 1] = J - mm >= 0 ? ldz(panel + ((size_t)(J - mm) * (D + 1) + mm) * 256 + tid) : zd{0.0, 0.0};
#pragma unroll
    for (int mm = 1; mm <= D; ++mm) stz(m.Lrow + (mm - 1) * T16 * PS + (tid >> 4) * PS + (tid & 15), v[mm - 1]);
}

// this lane's A-operand entries of L(J+a, J-mm) (row-major 16x16 panel tile)
template <int D>
__device__ __forceinline__ void st_operand(const zd *panel, int J, int a, int mm, int jc, int rq, zd (&o)[4]) {
    const zd *p = panel + ((size_t)(J - mm) * (D + 1) + a + mm) * 256 + jc * T16 + rq;
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) o[kc] = ldz(p + 4 * kc);
}

// T -= Pa Pb^T with Pa's A-operand entries in registers and Pb in LDS (stride PS)
__device__ __forceinline__ void mfma_update_rv(d4 &tr, d4 &ti, const zd (&va)[4], const zd *Pb, int jc, int rq) {
    zd vb[4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) vb[kc] = ldz(Pb + jc * PS + 4 * kc + rq);
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        tr = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[kc].re, vb[kc].re, tr, 0, 0, 0);
        ti = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[kc].re, vb[kc].im, ti, 0, 0, 0);
        tr = __builtin_amdgcn_mfma_f64_16x16x4f64(va[kc].im, vb[kc].im, tr, 0, 0, 0);
        ti = __builtin_amdgcn_mfma_f64_16x16x4f64(-va[kc].im, vb[kc].re, ti, 0, 0, 0);
    }
}

template <int D>
__device__ __forceinline__ void st_pivot_wave(const Factor16Args &A, StLds<D> &m) {
    const DevGrid &g = A.g;
    const int nblk = g.nblk, s = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int enode = g.enr[A.sys[s].x];
    const int *tm = g.tmap + (size_t)A.sys[s].x * g.tmap_sys;
    const double *fak = A.fak + (size_t)s * nblk * T16;
    const zd *vals = A.vals + (size_t)s * g.vstride;
    const zd *panel = A.panels + (size_t)s * nblk * (D + 1) * 256;
    int err = 0;
    int mi[4];
    d4 dr, di;
    la4_map<D>(g, tm, 0, 0, jc, rq, mi);
    la4_vals(vals, mi, dr, di);
    la4_map<D>(g, tm, 1, 0, jc, rq, mi);
    for (int J = 0; J < nblk; ++J) {
        st_stage<D>(panel, m, J, tid);
        lds_barrier();
        const int mx = J < D ? J : D;
        // right-hand side of block J: b_J - sum_m L(J, J-m) y_{J-m}
        const int i = J * T16 + jc;
        zd yv = (i < g.sanz && i == enode) ? zd{-fak[i], 0.0} : zd{0.0, 0.0};
        for (int mm = 1; mm <= mx; ++mm) {
            const zd *lr = m.Lrow + (mm - 1) * T16 * PS + jc * PS;
            const zd *yp = m.yr + ((J - mm) % (D + 1)) * T16;
#pragma unroll
            for (int c = 0; c < T16; ++c) yv = zfms(yv, ldz(lr + c), ldz(yp + c));
        }
        for (int mm = 1; mm <= mx; ++mm) {
            const zd *lr = m.Lrow + (mm - 1) * T16 * PS;
            mfma_update(dr, di, dr, di, lr, lr, jc, rq);
        }
        // the diagonal needs C(0) only: it overlaps the update waves' products
        la4_factor_diag<D>(A, m, s, J, lane, dr, di, yv, err);
        if (lane < T16) stz(m.yr + (J % (D + 1)) * T16 + lane, ldz(m.yj + lane));
        lds_barrier();
        // diagonal tile of column J+1 (map loaded one column ahead)
        la4_vals(vals, mi, dr, di);
        la4_map<D>(g, tm, J + 2, 0, jc, rq, mi);
        __syncthreads();  // panel column J is read by the next stage
    }
    if (lane == 0 && err && A.status[s] == 0) A.status[s] = err;
}

// The products of update wave W in one block step away from the band's ends
// (every tile live, tile a taking D-a products), in issue order: tile by tile, m upwards.
template <int D, int W>
struct StOps {
    struct Op {
        int u, a, mm;
    };
    static constexpr int count() {
        int n = 0;
        for (int u = 0; u < StCfg<D>::NA; ++u) {
            const int a = StCfg<D>::tile(W, u);
            if (a <= D) n += D - a;
        }
        return n;
    }
    static constexpr int N = count();
    static constexpr Op op(int i) {
        for (int u = 0; u < StCfg<D>::NA; ++u) {
            const int a = StCfg<D>::tile(W, u);
            if (a > D) continue;
            for (int mm = 1; mm <= D - a; ++mm)
                if (i-- == 0) return Op{u, a, mm};
        }
        return Op{0, 0, 0};
    }
};
constexpr int ST_PD = 3;  // streamed operands in flight per update wave

// product I of the steady-state sequence: consume operand slot I % ST_PD, refill it
// with operand I + ST_PD, then T -= L(J+a, J-m) L(J, J-m)^T on the matrix cores
template <int D, int W, int I, int NA>
__device__ __forceinline__ void st_pipe(const zd *panel, const StLds<D> &m, int J, int jc, int rq, zd (&buf)[ST_PD][4],
                                        d4 (&cr)[NA], d4 (&ci)[NA]) {
    if constexpr (I < StOps<D, W>::N) {
        constexpr auto o = StOps<D, W>::op(I);
        zd cur[4];
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            cur[kc].re = buf[I % ST_PD][kc].re;
            cur[kc].im = buf[I % ST_PD][kc].im;
        }
        if constexpr (I + ST_PD < StOps<D, W>::N) {
            constexpr auto on = StOps<D, W>::op(I + ST_PD);
            st_operand<D>(panel, J, on.a, on.mm, jc, rq, buf[I % ST_PD]);
        }
        mfma_update_rv(cr[o.u], ci[o.u], cur, m.Lrow + (o.mm - 1) * T16 * PS, jc, rq);
        // keep each product's loads and MFMAs together (bounds the live registers)
        __builtin_amdgcn_sched_barrier(0);
        st_pipe<D, W, I + 1, NA>(panel, m, J, jc, rq, buf, cr, ci);
    }
}

template <int D, int W, int I>
__device__ __forceinline__ void st_prime(const zd *panel, int J, int jc, int rq, zd (&buf)[ST_PD][4]) {
    if constexpr (I < ST_PD && I < StOps<D, W>::N) {
        constexpr auto o = StOps<D, W>::op(I);
        st_operand<D>(panel, J, o.a, o.mm, jc, rq, buf[I]);
        st_prime<D, W, I + 1>(panel, J, jc, rq, buf);
    }
}

template <int D, int W>
__device__ __forceinline__ void st_update_wave(const Factor16Args &A, StLds<D> &m) {
    using C = StCfg<D>;
    constexpr int NA = C::NA;
    const DevGrid &g = A.g;
    const int nblk = g.nblk, s = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int *tm = g.tmap + (size_t)A.sys[s].x * g.tmap_sys;
    const zd *vals = A.vals + (size_t)s * g.vstride;
    zd *panel = A.panels + (size_t)s * nblk * (D + 1) * 256;
    d4 cr[NA], ci[NA];
    int mi[NA][4];
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = C::tile(W, u);
        if (a <= D) {
            la4_map<D>(g, tm, a, a, jc, rq, mi[u]);
            la4_vals(vals, mi[u], cr[u], ci[u]);
            la4_map<D>(g, tm, 1 + a, a, jc, rq, mi[u]);
        }
    }
    for (int J = 0; J < nblk; ++J) {
        st_stage<D>(panel, m, J, tid);
        lds_barrier();
        // C(a) -= sum_m L(J+a, J-m) L(J, J-m)^T
        if (J >= D && J + D < nblk) {
            // away from the ends: a fixed sequence of products, ST_PD operands in flight
            zd buf[ST_PD][4];
            st_prime<D, W, 0>(panel, J, jc, rq, buf);
            st_pipe<D, W, 0, NA>(panel, m, J, jc, rq, buf, cr, ci);
        } else {
        // near the ends: the next operand in flight
        zd nx[4];
        bool have = false;  // nx holds the first operand of round u
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = C::tile(W, u);
            if (a <= D) {
                const int mx = J + a < nblk ? (J < D - a ? J : D - a) : 0;
                if (mx >= 1 && !have) st_operand<D>(panel, J, a, 1, jc, rq, nx);
                have = false;
                for (int mm = 1; mm <= mx; ++mm) {
                    zd cur[4];
#pragma unroll
                    for (int kc = 0; kc < 4; ++kc) {
                        cur[kc].re = nx[kc].re;
                        cur[kc].im = nx[kc].im;
                    }
                    if (mm < mx) {
                        st_operand<D>(panel, J, a, mm + 1, jc, rq, nx);
                    } else if (u + 1 < NA) {
                        const int a1 = C::tile(W, u + 1);
                        const int mx1 = J + a1 < nblk ? (J < D - a1 ? J : D - a1) : 0;
                        if (a1 <= D && mx1 >= 1) {
                            st_operand<D>(panel, J, a1, 1, jc, rq, nx);
                            have = true;
                        }
                    }
                    mfma_update_rv(cr[u], ci[u], cur, m.Lrow + (mm - 1) * T16 * PS, jc, rq);
                }
            }
        }
        }
        lds_barrier();  // L^-1 of block J is in LDS
        // X(a) = C(a) (L^-1)^T, stored as panel column J
        zd *stg = m.Lrow + (W - 1) * T16 * PS;
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = C::tile(W, u);
            if (a <= D) {
                zd *pg = panel + ((size_t)J * (D + 1) + a) * 256;
                if (J + a < nblk) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) stz(stg + (4 * v + rq) * PS + jc, zd{cr[u][v], ci[u][v]});
                    __builtin_amdgcn_wave_barrier();
                    d4 xr, xi;
                    mfma_mul(xr, xi, stg, m.Li, jc, rq);
#pragma unroll
                    for (int v = 0; v < 4; ++v) stz(pg + (4 * v + rq) * T16 + jc, zd{xr[v], xi[v]});
                    __builtin_amdgcn_wave_barrier();
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) stz(pg + (4 * v + rq) * T16 + jc, zd{0.0, 0.0});
                }
                // entering tile of column J+1 (map loaded one column ahead), map of column J+2
                la4_vals(vals, mi[u], cr[u], ci[u]);
                la4_map<D>(g, tm, J + 2 + a, a, jc, rq, mi[u]);
            }
        }
        __syncthreads();  // panel column J is read by the next stage
    }
}

template <int D, int W>
__device__ __forceinline__ void st_dispatch(int wave, const Factor16Args &A, StLds<D> &m) {
    if constexpr (W == 0) {
        if (wave == 0) st_pivot_wave<D>(A, m);
        else st_dispatch<D, 1>(wave, A, m);
    } else if constexpr (W < SWV) {
        if (wave == W) st_update_wave<D, W>(A, m);
        else st_dispatch<D, W + 1>(wave, A, m);
    }
}
}  // namespace

template <int D>
__global__ void __launch_bounds__(SWV * 64, 2) crt_factor16st_kernel(Factor16Args A) {
    __shared__ StLds<D> m;
    if ((int)blockIdx.x >= A.nsys) return;
    st_dispatch<D, 0>(threadIdx.x >> 6, A, m);
}

#define CRT_F16_MAXD 16

// kernel variants: 1 look-ahead (D = 2..10), 2 streaming (D >= 3), 3 plain
// right-looking (D <= 10); 0 picks 1 up to D = 10 (3 at D = 1) and 2 above
// (CRT_FACTOR_ST=1 and CRT_FACTOR_V2=1 change that default)
int crt_factor16_resolve(int D, int variant) {
    if (variant != 0) return variant;
    static const bool stream = getenv("CRT_FACTOR_ST") != nullptr;
    static const bool v2 = getenv("CRT_FACTOR_V2") != nullptr;
    return (D >= SNU && (D > 10 || stream)) ? 2 : ((D >= 2 && D <= 10 && !v2) ? 1 : 3);
}

template <int D>
static int launch16(hipStream_t st, const Factor16Args &a, int variant) {
    variant = crt_factor16_resolve(D, variant);
    if (variant == 2) {
        if constexpr (D >= SNU) {
            if (a.g.tmapD != D) return CRT_ERR_UNSUPPORTED;
            hipLaunchKernelGGL(crt_factor16st_kernel<D>, dim3(a.nsys), dim3(SWV * 64), 0, st, a);
            return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
        }
        return CRT_ERR_UNSUPPORTED;
    }
    if (variant == 1) {
        if constexpr (D >= 2 && D <= 10) {
            if (a.g.tmapD != D) return CRT_ERR_UNSUPPORTED;
            hipLaunchKernelGGL(crt_factor16la_kernel<D>, dim3(a.nsys), dim3(La4<D>::NT), 0, st, a);
            return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
        }
        return CRT_ERR_UNSUPPORTED;
    }
    if (variant != 3) return CRT_ERR_ARG;
    if constexpr (D <= 10) {
        hipLaunchKernelGGL(crt_factor16_kernel<D>, dim3(a.nsys), dim3(Cfg<D>::NT), 0, st, a);
        return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
    }
    return CRT_ERR_UNSUPPORTED;
}

// workgroups of the factor kernel that launch16 picks for (D, variant) resident on the
// whole device at once, 0 when unknown
template <int D>
static int resident16(int variant) {
    variant = crt_factor16_resolve(D, variant);
    int per_cu = 0, dev = 0, cus = 0;
    hipError_t e = hipErrorInvalidValue;
    if (variant == 2) {
        if constexpr (D >= SNU)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, crt_factor16st_kernel<D>, SWV * 64, 0);
    } else if (variant == 1) {
        if constexpr (D >= 2 && D <= 10)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, crt_factor16la_kernel<D>, La4<D>::NT, 0);
    } else if (variant == 3) {
        if constexpr (D <= 10)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, crt_factor16_kernel<D>, Cfg<D>::NT, 0);
    }
    if (e != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}

int crt_factor16_resident(int D, int variant) {
    switch (D) {
        case 1: return resident16<1>(variant);
        case 2: return resident16<2>(variant);
        case 3: return resident16<3>(variant);
        case 4: return resident16<4>(variant);
        case 5: return resident16<5>(variant);
        case 6: return resident16<6>(variant);
        case 7: return resident16<7>(variant);
        case 8: return resident16<8>(variant);
        case 9: return resident16<9>(variant);
        case 10: return resident16<10>(variant);
        case 11: return resident16<11>(variant);
        case 12: return resident16<12>(variant);
        case 13: return resident16<13>(variant);
        case 14: return resident16<14>(variant);
        case 15: return resident16<15>(variant);
        case 16: return resident16<16>(variant);
        default: return 0;
    }
}

int crt_factor16_supported(int mb) {
    int D = (mb + 15) / 16;
    return D >= 1 && D <= CRT_F16_MAXD;
}

int crt_launch_factor16(hipStream_t st, const DevGrid &g16, int nsys, const int2 *sys, const double *fak,
                        const zd *vals, zd *panels, zd *y, zd *dinv, int *status, int variant) {
    Factor16Args a{g16, nsys, sys, fak, vals, panels, y, dinv, status};
    switch (g16.D) {
        case 1: return launch16<1>(st, a, variant);
        case 2: return launch16<2>(st, a, variant);
        case 3: return launch16<3>(st, a, variant);
        case 4: return launch16<4>(st, a, variant);
        case 5: return launch16<5>(st, a, variant);
        case 6: return launch16<6>(st, a, variant);
        case 7: return launch16<7>(st, a, variant);
        case 8: return launch16<8>(st, a, variant);
        case 9: return launch16<9>(st, a, variant);
        case 10: return launch16<10>(st, a, variant);
        case 11: return launch16<11>(st, a, variant);
        case 12: return launch16<12>(st, a, variant);
        case 13: return launch16<13>(st, a, variant);
        case 14: return launch16<14>(st, a, variant);
        case 15: return launch16<15>(st, a, variant);
        case 16: return launch16<16>(st, a, variant);
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
// PT is the panel element type: zd, or zf for the FP32 factorisation (read into
// FP64 arithmetic).  MODE 0: x scaled by fak into kpotN.  MODE 3: as MODE 0 for the
// scaled right-hand side in yscr (overwritten by y) instead of the unit source.  MODE 1 (first solve of
// the refinement): also x into xacc.  MODE 2 (refinement sweep): the right-hand
// side is the negated residual in yscr (overwritten by y), x is added to xacc,
// kpotN gets fak * xacc, and stats[2 rid + 0/1] the max-norms of fak * x and of
// fak * xacc.
template <class PT>
struct SolveArgs {
    int sanz, nblk, kwnanz, eanz, nrhs;
    const int4 *rhs;     // (l, k, factor system, have_y)
    const int *enr;
    const double *fak;
    const PT *panels;
    const PT *dinv;
    const PT *y;
    zd *yscr;
    zd *kpotN;
    zd *xacc;            // MODE 1/2: [rhs][nblk*16]
    double *stats;       // MODE 2: [rhs][2]
    const int *skip;     // non-zero *skip: the launch does nothing (a converged refinement)
};
using Solve16Args = SolveArgs<zd>;

namespace {
constexpr int SW = 4;  // waves per right-hand side

template <int D, class PT>
__device__ __forceinline__ void solve16_fetch_diag(const PT *panel, const PT *dinv, int nblk, int J, int c,
                                                   zd (&drow)[T16], zd &pnext) {
    const int Jc = J < 0 ? 0 : (J < nblk ? J : nblk - 1);
    const PT *Ld = panel + (size_t)Jc * (D + 1) * 256 + c * T16;
#pragma unroll
    for (int cc = 0; cc < T16; ++cc) drow[cc] = zd{Ld[cc].re, Ld[cc].im};
    const PT *pv = dinv + (size_t)Jc * T16 + c;
    pnext = zd{pv->re, pv->im};
}

__device__ __forceinline__ double cmax_abs(zd z) { return fmax(fabs(z.re), fabs(z.im)); }

// What the substitution wave needs of block J in the backward pass, per lane (row
// c = lane % 16, column quarter rq = lane / 16): four entries L(J,J)(c, 4rq..4rq+3)
// of the diagonal tile, 1/L(c,c), y(c) and fak(c).  Fetched three blocks ahead, so
// the loads of a block have two block steps to arrive.
struct BackStage {
    zd d[4], p, y;
    double f;
};

__device__ __forceinline__ void solve16_store_diag(zd *dtile, int lane, const zd (&drow)[T16]) {
    if (lane < T16)
#pragma unroll
        for (int cc = 0; cc < T16; ++cc) {
            dtile[lane * PS + cc].re = drow[cc].re;
            dtile[lane * PS + cc].im = drow[cc].im;
        }
}

}  // namespace

// INV: the diagonal tiles hold L^-1 above the diagonal (the look-ahead, streaming and
// FP32 factorisations write it), so the backward pass applies L(J,J)^-T as a product
template <int D, class PT = zd, int MODE = 0, bool INV = false>
__global__ void __launch_bounds__(64 * SW, D <= 6 ? 4 : 2) crt_solve16_kernel(SolveArgs<PT> A) {
    constexpr int NA = (D + SW - 2) / (SW - 1);  // off-diagonal tiles per worker wave
    __shared__ zd ring[(D + 1) * T16];
    __shared__ zd part[SW][T16];
    __shared__ zd dtile[T16 * PS];  // diagonal tile of the current block
    const int rid = blockIdx.x;
    if (rid >= A.nrhs) return;
    if (A.skip && *A.skip) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int c = lane & 15, rq = lane >> 4;
    const int nblk = A.nblk;
    const int4 q4 = A.rhs[rid];
    const int l = q4.x, k = q4.y, fs = q4.z;
    const double *fak = A.fak + (size_t)fs * nblk * T16;
    const PT *panel = A.panels + (size_t)fs * nblk * (D + 1) * 256;
    const PT *dinv = A.dinv + (size_t)fs * nblk * T16;
    const int enode = A.enr[l];
    zd *yscr = A.yscr + (size_t)rid * nblk * T16;
    // a refinement sweep (MODE 2) and a general right-hand side (MODE 3) run the forward
    // pass over every block; otherwise y is zero above the electrode's block
    const bool have_y = MODE >= 2 ? false : q4.w != 0;
    const int J0 = (MODE >= 2 || have_y) ? 0 : enode / T16;

    // The two wave roles run separate loops with the same barrier sequence, so
    // their registers are allocated independently.
    if (wave == 0) {
        // ---- wave 0: substitution through the diagonal tiles
        zd pcur, drow[T16], pnext;
        if (!have_y) {
            solve16_fetch_diag<D>(panel, dinv, nblk, J0, c, drow, pnext);
            solve16_store_diag(dtile, lane, drow);
            pcur = pnext;
            for (int J = J0; J < nblk; ++J) {
                solve16_fetch_diag<D>(panel, dinv, nblk, J + 1, c, drow, pnext);
                const int r = c;  // lane r = lane & 15 carries row r of the block
                const int i = J * T16 + r;
                zd yv;
                if constexpr (MODE >= 2) yv = yscr[(size_t)J * T16 + r];
                else yv = (i < A.sanz && i == enode) ? zd{-fak[i], 0.0} : zd{0.0, 0.0};
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
        const PT *yf = A.y + (size_t)fs * nblk * T16;
        const zd *yg = nullptr;
        if constexpr (std::is_same<PT, zd>::value) yg = have_y ? yf : yscr;
        // y of row i: the factor's own (have_y) or this right-hand side's forward pass
        auto yat = [&](int i) -> zd {
            if constexpr (std::is_same<PT, zd>::value) {
                return yg[i];
            } else {
                if (have_y) {
                    const PT e = yf[i];
                    return zd{e.re, e.im};
                }
                return yscr[i];
            }
        };
        auto fetch = [&](int J, BackStage &b) {
            const int Jc = J < 0 ? 0 : (J < nblk ? J : nblk - 1);
            const PT *Ld = panel + (size_t)Jc * (D + 1) * 256 + c * T16 + 4 * rq;
#pragma unroll
            for (int j = 0; j < 4; ++j) b.d[j] = zd{Ld[j].re, Ld[j].im};
            const PT pv = dinv[(size_t)Jc * T16 + c];
            b.p = zd{pv.re, pv.im};
            const int i = J * T16 + c;
            b.y = J >= J0 && J >= 0 ? yat(i) : zd{0.0, 0.0};
            b.f = J >= 0 && i < A.sanz ? fak[i] : 0.0;
        };
        auto store = [&](const BackStage &b) {
#pragma unroll
            for (int j = 0; j < 4; ++j) dtile[c * PS + 4 * rq + j] = b.d[j];
        };
        auto move = [&](BackStage &to, const BackStage &from) {
#pragma unroll
            for (int j = 0; j < 4; ++j) to.d[j] = zd{from.d[j].re, from.d[j].im};
            to.p = zd{from.p.re, from.p.im};
            to.y = zd{from.y.re, from.y.im};
            to.f = from.f;
        };
        __syncthreads();  // the forward pass has finished with dtile
        BackStage s1, s2;
        zd ycur;      // y(c) of the current block
        double fcur;  // fak of the current block's row c
        {
            BackStage s0;
            fetch(nblk - 1, s0);
            fetch(nblk - 2, s1);
            fetch(nblk - 3, s2);
            store(s0);
            pcur = s0.p;
            ycur = s0.y;
            fcur = s0.f;
        }
        double dmax = 0.0, xmax = 0.0;
        for (int J = nblk - 1; J >= 0; --J) {
            BackStage sn;
            fetch(J - 3, sn);
            __syncthreads();
            zd xv = zd{-ycur.re, -ycur.im};  // lane c carries column c of the block
#pragma unroll
            for (int w = 1; w < SW; ++w) xv = zsub(xv, part[w][c]);
            if constexpr (INV) {
                // x(c) = rv(c) / L(c,c) + sum_{r > c} L^-1(r, c) rv(r); lane (c, rq) takes
                // r = 4rq..4rq+3, L^-1(r, c) sits at tile(c, r)
                zd sacc = zd{0.0, 0.0};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = 4 * rq + j;
                    const zd rv = shfl_z16(xv, r);
                    if (r > c) sacc = zfma(sacc, dtile[c * PS + r], rv);
                }
                sacc = zadd(sacc, shfl_xor_z16(sacc, 16));
                sacc = zadd(sacc, shfl_xor_z16(sacc, 32));
                xv = zadd(zmul(xv, pcur), sacc);
            } else {
#pragma unroll
                for (int cc = T16 - 1; cc >= 0; --cc) {
                    const zd xcc = zmul(readlane_z(xv, cc), readlane_z(pcur, cc));
                    if (c == cc) xv = xcc;
                    if (c < cc) xv = zfms(xv, dtile[cc * PS + c], xcc);
                }
            }
            if (lane < T16) {
                ring[(J % (D + 1)) * T16 + lane] = xv;
                const int i = J * T16 + lane;
                if (i < A.sanz) {
                    const double f = fcur;
                    zd xo = xv;
                    if constexpr (MODE == 1 || MODE == 2) {
                        zd *xa = A.xacc + (size_t)rid * nblk * T16 + i;
                        if constexpr (MODE == 2) xo = zadd(*xa, xv);
                        *xa = xo;
                        dmax = fmax(dmax, f * cmax_abs(xv));
                        xmax = fmax(xmax, f * cmax_abs(xo));
                    }
                    A.kpotN[((size_t)i * A.kwnanz + k) * A.eanz + l] = zd{xo.re * f, xo.im * f};
                }
            }
            __builtin_amdgcn_wave_barrier();
            store(s1);
            pcur = s1.p;
            ycur = s1.y;
            fcur = s1.f;
            move(s1, s2);
            move(s2, sn);
            __syncthreads();
        }
        if constexpr (MODE == 1 || MODE == 2) {
#pragma unroll
            for (int m = 1; m < T16; m <<= 1) {
                dmax = fmax(dmax, __shfl_xor(dmax, m, 64));
                xmax = fmax(xmax, __shfl_xor(xmax, m, 64));
            }
            if (lane == 0) {
                A.stats[2 * (size_t)rid] = dmax;
                A.stats[2 * (size_t)rid + 1] = xmax;
            }
        }
        return;
    }

    // ---- waves 1..3: products with the off-diagonal tiles a = wave + 3u; the tiles of
    // PD blocks beyond the current one are in flight (the loop waits on HBM latency,
    // not bandwidth: with FP32 panels, half the bytes, it takes as long)
    constexpr int PD = (NA <= 3 && D > 6) ? 2 : 1;  // D <= 6: four workgroups per CU, 128 VGPRs
    zd cur[NA][4], nx1[NA][4];
    auto rotate = [&](zd (&nxt)[NA][4]) {
#pragma unroll
        for (int u = 0; u < NA; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if constexpr (PD == 2) {
                    cur[u][v] = nx1[u][v];
                    nx1[u][v] = nxt[u][v];
                } else {
                    cur[u][v] = nxt[u][v];
                }
            }
    };
    if (!have_y) {
        auto fetch_f = [&](int J, zd (&t)[NA][4]) {
#pragma unroll
            for (int u = 0; u < NA; ++u) {
                const int a = wave + (SW - 1) * u;
                const bool live = a <= D && J - a >= J0 && J < nblk;
                const PT *Lt = panel + ((size_t)(live ? J - a : 0) * (D + 1) + (live ? a : 0)) * 256;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const PT *p = Lt + (rq + 4 * v) * T16 + c;
                    t[u][v] = live ? zd{p->re, p->im} : zd{0.0, 0.0};
                }
            }
        };
        fetch_f(J0, cur);
        if constexpr (PD == 2) fetch_f(J0 + 1, nx1);
        for (int J = J0; J < nblk; ++J) {
            zd nxt[NA][4];
            fetch_f(J + PD, nxt);
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
            rotate(nxt);
        }
    }
    auto fetch_b = [&](int J, zd (&t)[NA][4]) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = wave + (SW - 1) * u;
            const bool live = a <= D && J >= 0 && J + a < nblk;
            const PT *Lt = panel + ((size_t)(live ? J : 0) * (D + 1) + (live ? a : 0)) * 256;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const PT *p = Lt + (rq + 4 * v) * T16 + c;
                t[u][v] = live ? zd{p->re, p->im} : zd{0.0, 0.0};
            }
        }
    };
    fetch_b(nblk - 1, cur);
    if constexpr (PD == 2) fetch_b(nblk - 2, nx1);
    __syncthreads();
    for (int J = nblk - 1; J >= 0; --J) {
        zd nxt[NA][4];
        fetch_b(J - PD, nxt);
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
        rotate(nxt);
    }
}

namespace {

template <int D, class PT, int MODE, bool INV>
int launch_solve16(hipStream_t st, const SolveArgs<PT> &a) {
    hipLaunchKernelGGL((crt_solve16_kernel<D, PT, MODE, INV>), dim3(a.nrhs), dim3(64 * SW), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

template <class PT, int MODE, bool INV>
int solve16_dispatch(hipStream_t st, int D, const SolveArgs<PT> &a) {
    switch (D) {
        case 1: return launch_solve16<1, PT, MODE, INV>(st, a);
        case 2: return launch_solve16<2, PT, MODE, INV>(st, a);
        case 3: return launch_solve16<3, PT, MODE, INV>(st, a);
        case 4: return launch_solve16<4, PT, MODE, INV>(st, a);
        case 5: return launch_solve16<5, PT, MODE, INV>(st, a);
        case 6: return launch_solve16<6, PT, MODE, INV>(st, a);
        case 7: return launch_solve16<7, PT, MODE, INV>(st, a);
        case 8: return launch_solve16<8, PT, MODE, INV>(st, a);
        case 9: return launch_solve16<9, PT, MODE, INV>(st, a);
        case 10: return launch_solve16<10, PT, MODE, INV>(st, a);
        case 11: return launch_solve16<11, PT, MODE, INV>(st, a);
        case 12: return launch_solve16<12, PT, MODE, INV>(st, a);
        case 13: return launch_solve16<13, PT, MODE, INV>(st, a);
        case 14: return launch_solve16<14, PT, MODE, INV>(st, a);
        case 15: return launch_solve16<15, PT, MODE, INV>(st, a);
        case 16: return launch_solve16<16, PT, MODE, INV>(st, a);
        default: return CRT_ERR_UNSUPPORTED;
    }
}
}  // namespace

// inv: the factor variant stored L^-1 above the diagonal of its diagonal tiles
int crt_launch_solve16(hipStream_t st, int sanz, int nblk, int D, int kwnanz, int eanz, int nrhs, const int4 *rhs,
                       const int *enr, const double *fak, const zd *panels, const zd *dinv, const zd *y, zd *yscr,
                       zd *kpotN, int inv) {
    Solve16Args a{sanz, nblk, kwnanz, eanz, nrhs, rhs, enr, fak, panels, dinv, y, yscr, kpotN, nullptr, nullptr};
    return inv ? solve16_dispatch<zd, 0, true>(st, D, a) : solve16_dispatch<zd, 0, false>(st, D, a);
}

// general right-hand sides (yscr[rhs][nblk*16], scaled; overwritten by y)
int crt_launch_solve16_rhs(hipStream_t st, int sanz, int nblk, int D, int kwnanz, int eanz, int nrhs, const int4 *rhs,
                           const int *enr, const double *fak, const zd *panels, const zd *dinv, zd *yscr, zd *kpotN,
                           int inv) {
    Solve16Args a{sanz, nblk, kwnanz, eanz, nrhs, rhs, enr, fak, panels, dinv, nullptr, yscr, kpotN, nullptr, nullptr};
    return inv ? solve16_dispatch<zd, 3, true>(st, D, a) : solve16_dispatch<zd, 3, false>(st, D, a);
}

// substitution through the FP32 panels for the refinement: mode 1 is the first
// solve (x into xacc), mode 2 a sweep (negated residual in yscr, x added to xacc,
// the max-norms of the correction and of the sum into stats)
int crt_launch_solve16_f32(hipStream_t st, int sanz, int nblk, int D, int kwnanz, int eanz, int nrhs,
                           const int4 *rhs, const int *enr, const double *fak, const zf *panels, const zf *dinv,
                           const zf *y, zd *yscr, zd *kpotN, zd *xacc, double *stats, int mode, const int *skip) {
    SolveArgs<zf> a{sanz, nblk, kwnanz, eanz, nrhs, rhs, enr, fak, panels, dinv, y, yscr, kpotN, xacc, stats, skip};
    if (mode == 1) return solve16_dispatch<zf, 1, true>(st, D, a);
    if (mode == 2) return solve16_dispatch<zf, 2, true>(st, D, a);
    return CRT_ERR_ARG;
}
