// crt_band16.hip -- FP64 complex-symmetric band factorisation on the matrix cores.
//
// Same mathematics as crt_band.hip (src/kompab.f90, src/scalab.f90,
// src/chol.f90, src/vre.f90, src/fem.f90:383-394) with 16x16 tiles:
//   * the active window of the band (tiles (J+a, J+b), 0 <= b <= a <= D,
//     D = ceil(mb/16)) lives in registers as v_mfma_f64_16x16x4 accumulators
//     (real and imaginary part, 8 VGPRs each);
//   * tiles on diagonal d = a-b slide one position per step; wave 0 owns the
//     main diagonal, wave g >= 1 owns diagonals g and D+1-g, so every wave
//     holds D+1 tiles at most and the rank-16 updates are balanced;
//   * the factored panel column J and the entering row block J+D+1 are staged
//     in LDS (about 8 KB per tile column), so two workgroups share a CU and one
//     system's pivot chain overlaps the other's matrix-core updates.
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

// complex pivot: L = sqrt(s), inv = 1/L = conj(L)/|s|
__device__ __forceinline__ void pivot_inv(zd s, zd &L, zd &inv, double &mag) {
    L = zsqrt(s);
    mag = zabs(s);
    double rm = 1.0 / mag;
    inv = zd{L.re * rm, -L.im * rm};
}

template <int D>
struct Cfg {
    static constexpr int W = 1 + D / 2 + (D & 1);  // waves
    static constexpr int NT = 64 * W;               // threads
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
    int *status;
};

template <int D>
__global__ void __launch_bounds__(Cfg<D>::NT) crt_factor16_kernel(Factor16Args A) {
    constexpr int W = Cfg<D>::W;
    constexpr int NT = Cfg<D>::NT;
    extern __shared__ double smem16[];
    const DevGrid &g = A.g;
    const int nblk = g.nblk;  // blocks of 16 rows
    const int s = blockIdx.x;
    if (s >= A.nsys) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int l = A.sys[s].x;
    const int enode = g.enr[l];
    const double *fak = A.fak + (size_t)s * nblk * T16;
    const zd *vals = A.vals + (size_t)s * g.nnz;
    zd *panel = A.panels + (size_t)s * nblk * (D + 1) * 256;
    zd *yout = A.y + (size_t)s * nblk * T16;

    zd *P = reinterpret_cast<zd *>(smem16);        // (D+1) tiles, 16 x PS
    zd *S = P + (D + 1) * T16 * PS;                // (D+1) staging tiles, 16 x 16
    zd *bring = S + (D + 1) * 256;                 // (D+2) x 16
    zd *col = bring + (D + 2) * T16;               // 16
    zd *pinv = col + T16;                          // 16
    zd *yj = pinv + T16;                           // 16

    // diagonals of this wave: A segment (positions 0..LA-1 in acc[0..LA-1]),
    // B segment reversed (position p in acc[D-p], p = 0..LB-1)
    int dA, LA, dB, LB;
    if (wave == 0) { dA = 0; LA = D + 1; dB = 0; LB = 0; }
    else if (wave <= D / 2) { dA = wave; LA = D + 1 - wave; dB = D + 1 - wave; LB = wave; }
    else { dA = (D + 1) / 2; LA = (D + 1) / 2; dB = 0; LB = 0; }

    d4 ar[D + 1], ai[D + 1];
#pragma unroll
    for (int q = 0; q <= D; ++q) {
        ar[q] = d4{0.0, 0.0, 0.0, 0.0};
        ai[q] = d4{0.0, 0.0, 0.0, 0.0};
    }
    for (int t = tid; t < (D + 1) * 256; t += NT) S[t] = zd{0.0, 0.0};
    __syncthreads();

    // result-layout coordinates of this lane: column jc, rows 4v + rq
    const int jc = lane & 15, rq = lane >> 4;
    int err = 0;

    for (int J = -(D + 1); J < nblk; ++J) {
        const int Ie = J + D + 1;  // entering row block
        // ------------------------------------------------ A: panel tiles -> LDS
        if (J >= 0) {
            zd *pa = P + dA * T16 * PS;
#pragma unroll
            for (int v = 0; v < 4; ++v) pa[(4 * v + rq) * PS + jc] = zd{ar[0][v], ai[0][v]};
            if (LB > 0) {
                zd *pb = P + dB * T16 * PS;
#pragma unroll
                for (int v = 0; v < 4; ++v) pb[(4 * v + rq) * PS + jc] = zd{ar[D][v], ai[D][v]};
            }
        }
        __syncthreads();
        // ------------------- B: wave 0 factors the diagonal tile | others stage Ie
        if (wave == 0 && J >= 0) {
            zd a[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) a[v] = P[(4 * v + rq) * PS + jc];
            zd yv = bring[(J % (D + 2)) * T16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < T16; ++c) {
                zd sp = shfl_z16(a[c / 4], c + 16 * (c % 4));
                zd Lc, iv;
                double mag;
                pivot_inv(sp, Lc, iv, mag);
                if (!(mag > 0.0)) err = CRT_ERR_PIVOT;
                if (jc == c) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        int i = 4 * v + rq;
                        if (i == c) a[v] = Lc;
                        else if (i > c) a[v] = zmul(a[v], iv);
                        if (i >= c) col[i] = a[v];
                    }
                }
                __builtin_amdgcn_wave_barrier();
                zd ljc = col[jc];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    int i = 4 * v + rq;
                    if (jc > c && i >= jc) a[v] = zfms(a[v], col[i], ljc);
                }
                // forward substitution of the right-hand side (lanes 0..15 hold y(lane))
                zd yc = zmul(shfl_z16(yv, c), iv);
                if ((lane & 15) == c) yv = yc;
                if ((lane & 15) > c) yv = zfms(yv, col[lane & 15], yc);
                if (lane == 0) pinv[c] = iv;
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int i = 4 * v + rq;
                P[i * PS + jc] = (jc <= i) ? a[v] : zd{0.0, 0.0};
            }
            if (lane < T16) {
                yj[lane] = yv;
                bring[(J % (D + 2)) * T16 + lane] = yv;
                yout[(size_t)J * T16 + lane] = yv;
            }
        } else if (Ie < nblk && (wave != 0 || J < 0)) {
            // CSR entries of row block Ie are contiguous; scaled values from crt_assemble_kernel
            const int t0 = (J < 0) ? tid : tid - 64, nthr = (J < 0) ? NT : NT - 64;
            const int qb = g.row_ptr[Ie * T16], qe = g.row_ptr[min(Ie * T16 + T16, g.sanz)];
            for (int q = qb + t0; q < qe; q += nthr) {
                int i = g.crow[q], j = g.col[q];
                S[(Ie - j / T16) * 256 + (i - Ie * T16) * T16 + (j % T16)] = vals[q];
            }
            if (t0 < T16) {
                int i = Ie * T16 + t0;
                if (i >= g.sanz) S[t0 * T16 + t0] = zd{1.0, 0.0};  // padding rows: unit diagonal
                zd bv = zd{0.0, 0.0};
                if (i < g.sanz && i == enode) bv = zd{-fak[i], 0.0};
                bring[(Ie % (D + 2)) * T16 + t0] = bv;
            }
        }
        __syncthreads();
        // ------------------------------- C: TRSM X = A L^-T of the panel, RHS update
        if (J >= 0 && tid < T16 * D) {
            const int a = 1 + tid / T16, r = tid % T16;
            if (J + a < nblk) {
                zd *row = P + (a * T16 + r) * PS;
                zd x[T16];
#pragma unroll
                for (int c = 0; c < T16; ++c) x[c] = row[c];
                zd bacc = bring[((J + a) % (D + 2)) * T16 + r];
#pragma unroll
                for (int c = 0; c < T16; ++c) {
                    zd sacc = x[c];
#pragma unroll
                    for (int cp = 0; cp < c; ++cp) sacc = zfms(sacc, x[cp], P[c * PS + cp]);
                    x[c] = zmul(sacc, pinv[c]);
                    bacc = zfms(bacc, x[c], yj[c]);
                }
#pragma unroll
                for (int c = 0; c < T16; ++c) row[c] = x[c];
                bring[((J + a) % (D + 2)) * T16 + r] = bacc;
            }
        }
        __syncthreads();
        // ------------------- D: store panel, rank-16 updates on the matrix cores, slide
        if (J >= 0) {
            zd *pj = panel + (size_t)J * (D + 1) * 256;
            for (int t = tid; t < (D + 1) * 256; t += NT) {
                int a = t >> 8, rc = t & 255;
                pj[t] = (J + a < nblk) ? P[(a * T16 + (rc >> 4)) * PS + (rc & 15)] : zd{0.0, 0.0};
            }
#pragma unroll
            for (int q = 0; q <= D; ++q) {
                int d, p;
                bool live;
                if (q < LA) { d = dA; p = q; live = true; }
                else { d = dB; p = D - q; live = (LB > 0); }
                live = live && p >= 1 && (J + p + d < nblk);
                if (live) {
                    const zd *Pa = P + (p + d) * T16 * PS;  // L(J+p+d, J)
                    const zd *Pb = P + p * T16 * PS;        // L(J+p, J)
#pragma unroll
                    for (int kc = 0; kc < 4; ++kc) {
                        zd va = Pa[jc * PS + 4 * kc + rq];
                        zd vb = Pb[jc * PS + 4 * kc + rq];
                        ar[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.re, vb.re, ar[q], 0, 0, 0);
                        ar[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.im, vb.im, ar[q], 0, 0, 0);
                        ai[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.re, vb.im, ai[q], 0, 0, 0);
                        ai[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.im, vb.re, ai[q], 0, 0, 0);
                    }
                }
            }
        }
        // slide: position p -> p-1 on both diagonals; the entering tiles take the last slots
        {
            zd na[4], nb[4];
            const zd *sa = S + dA * 256;
            const zd *sb = S + dB * 256;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                na[v] = sa[(4 * v + rq) * T16 + jc];
                nb[v] = (LB > 0) ? sb[(4 * v + rq) * T16 + jc] : zd{0.0, 0.0};
            }
            d4 orr[D + 1], oii[D + 1];
#pragma unroll
            for (int q = 0; q <= D; ++q) { orr[q] = ar[q]; oii[q] = ai[q]; }
#pragma unroll
            for (int q = 0; q <= D; ++q) {
                d4 nr, ni;
                if (q < LA - 1) { nr = orr[q + 1 <= D ? q + 1 : D]; ni = oii[q + 1 <= D ? q + 1 : D]; }
                else if (q == LA - 1) { nr = d4{na[0].re, na[1].re, na[2].re, na[3].re}; ni = d4{na[0].im, na[1].im, na[2].im, na[3].im}; }
                else if (q == LA) { nr = d4{nb[0].re, nb[1].re, nb[2].re, nb[3].re}; ni = d4{nb[0].im, nb[1].im, nb[2].im, nb[3].im}; }
                else { nr = orr[q >= 1 ? q - 1 : 0]; ni = oii[q >= 1 ? q - 1 : 0]; }
                ar[q] = nr;
                ai[q] = ni;
            }
            // the staging tiles of this wave are consumed: clear them for the next block
            zd *wa = S + dA * 256;
#pragma unroll
            for (int v = 0; v < 4; ++v) wa[(4 * v + rq) * T16 + jc] = zd{0.0, 0.0};
            if (LB > 0) {
                zd *wb = S + dB * 256;
#pragma unroll
                for (int v = 0; v < 4; ++v) wb[(4 * v + rq) * T16 + jc] = zd{0.0, 0.0};
            }
        }
        __syncthreads();
    }
    if (wave == 0 && lane == 0 && err && A.status[s] == 0) A.status[s] = err;
}

size_t crt_factor16_lds_bytes(int D) {
    return ((size_t)(D + 1) * T16 * PS + (size_t)(D + 1) * 256 + (size_t)(D + 2) * T16 + 3 * T16) * sizeof(zd);
}

#define CRT_F16_MAXD 10

template <int D>
static int launch16(hipStream_t st, const Factor16Args &a) {
    size_t lds = crt_factor16_lds_bytes(D);
    (void)hipFuncSetAttribute((const void *)crt_factor16_kernel<D>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(crt_factor16_kernel<D>, dim3(a.nsys), dim3(Cfg<D>::NT), lds, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_factor16_supported(int mb) {
    int D = (mb + 15) / 16;
    return D >= 1 && D <= CRT_F16_MAXD;
}

int crt_launch_factor16(hipStream_t st, const DevGrid &g16, int nsys, const int2 *sys, const double *fak,
                        const zd *vals, zd *panels, zd *y, int *status) {
    Factor16Args a{g16, nsys, sys, fak, vals, panels, y, status};
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
// One wavefront per right-hand side on 16x16 panels.  Lane l holds entries
// (r, c) = (l/16 + 4v, l%16), v = 0..3, of a tile.
struct Solve16Args {
    int sanz, nblk, D, kwnanz, eanz, nrhs;
    const int4 *rhs;     // (l, k, factor system, have_y)
    const int *enr;
    const double *fak;
    const zd *panels;
    const zd *y;
    zd *yscr;
    zd *kpotN;
};

__global__ void __launch_bounds__(256) crt_solve16_kernel(Solve16Args A) {
    extern __shared__ double smem_s16[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rid = blockIdx.x * (blockDim.x >> 6) + wave;
    if (rid >= A.nrhs) return;
    const int D = A.D, nblk = A.nblk;
    const int4 rq4 = A.rhs[rid];
    const int l = rq4.x, k = rq4.y, fs = rq4.z;
    const double *fak = A.fak + (size_t)fs * nblk * T16;
    const zd *panel = A.panels + (size_t)fs * nblk * (D + 1) * 256;
    zd *ring = reinterpret_cast<zd *>(smem_s16) + (size_t)wave * ((D + 1) * T16 + 256);
    zd *tile = ring + (D + 1) * T16;
    const int c = lane & 15, rq = lane >> 4;
    const int enode = A.enr[l];
    zd *yscr = A.yscr + (size_t)rid * nblk * T16;

    if (!rq4.w) {
        for (int J = 0; J < nblk; ++J) {
            // acc(r) = sum_a sum_c L(J, J-a)(r, c) y_{J-a}(c)
            zd acc[4] = {zd{0, 0}, zd{0, 0}, zd{0, 0}, zd{0, 0}};
            for (int a = 1; a <= D && a <= J; ++a) {
                const zd *Lt = panel + ((size_t)(J - a) * (D + 1) + a) * 256;
                zd yv = ring[((J - a) % (D + 1)) * T16 + c];
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[v] = zfma(acc[v], Lt[(rq + 4 * v) * T16 + c], yv);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 1));
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 2));
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 4));
                acc[v] = zadd(acc[v], shfl_xor_z16(acc[v], 8));
            }
            // diagonal tile into LDS, rhs of row r into lane r
            const zd *Ld = panel + (size_t)J * (D + 1) * 256;
#pragma unroll
            for (int v = 0; v < 4; ++v) tile[(rq + 4 * v) * T16 + c] = Ld[(rq + 4 * v) * T16 + c];
            // row r = rq + 4v is held (after the reduction) by every lane with lane>>4 == rq
            zd rv = zd{0.0, 0.0};
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                zd t = shfl_z16(acc[v], 16 * ((lane & 15) % 4));
                if ((lane & 15) / 4 == v) rv = t;
            }
            int i = J * T16 + (lane & 15);
            zd bv = (i < A.sanz && i == enode) ? zd{-fak[i], 0.0} : zd{0.0, 0.0};
            zd yv = zsub(bv, rv);
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int cc = 0; cc < T16; ++cc) {
                zd ycc = zdiv(shfl_z16(yv, cc), tile[cc * T16 + cc]);
                if ((lane & 15) == cc) yv = ycc;
                if ((lane & 15) > cc) yv = zfms(yv, tile[(lane & 15) * T16 + cc], ycc);
            }
            if (lane < T16) {
                ring[(J % (D + 1)) * T16 + lane] = yv;
                yscr[(size_t)J * T16 + lane] = yv;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    const zd *yg = rq4.w ? A.y + (size_t)fs * nblk * T16 : yscr;
    for (int J = nblk - 1; J >= 0; --J) {
        // acc(c) = sum_a sum_r L(J+a, J)(r, c) x_{J+a}(r)
        zd acc = zd{0.0, 0.0};
        for (int a = 1; a <= D && J + a < nblk; ++a) {
            const zd *Lt = panel + ((size_t)J * (D + 1) + a) * 256;
            const zd *xr = ring + ((J + a) % (D + 1)) * T16;
#pragma unroll
            for (int v = 0; v < 4; ++v) acc = zfma(acc, Lt[(rq + 4 * v) * T16 + c], xr[rq + 4 * v]);
        }
        acc = zadd(acc, shfl_xor_z16(acc, 16));
        acc = zadd(acc, shfl_xor_z16(acc, 32));
        const zd *Ld = panel + (size_t)J * (D + 1) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v) tile[(rq + 4 * v) * T16 + c] = Ld[(rq + 4 * v) * T16 + c];
        zd yJ = yg[(size_t)J * T16 + c];
        zd xv = zd{-yJ.re - acc.re, -yJ.im - acc.im};  // lanes hold the rhs of column lane&15
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int cc = T16 - 1; cc >= 0; --cc) {
            zd xcc = zdiv(shfl_z16(xv, cc), tile[cc * T16 + cc]);
            if (c == cc) xv = xcc;
            if (c < cc) xv = zfms(xv, tile[cc * T16 + c], xcc);
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < T16) {
            ring[(J % (D + 1)) * T16 + lane] = xv;
            int i = J * T16 + lane;
            if (i < A.sanz) {
                double f = fak[i];
                A.kpotN[((size_t)i * A.kwnanz + k) * A.eanz + l] = zd{xv.re * f, xv.im * f};
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

int crt_launch_solve16(hipStream_t st, int sanz, int nblk, int D, int kwnanz, int eanz, int nrhs, const int4 *rhs,
                       const int *enr, const double *fak, const zd *panels, const zd *y, zd *yscr, zd *kpotN) {
    Solve16Args a{sanz, nblk, D, kwnanz, eanz, nrhs, rhs, enr, fak, panels, y, yscr, kpotN};
    const int waves = 4;
    size_t lds = (size_t)waves * ((D + 1) * T16 + 256) * sizeof(zd);
    (void)hipFuncSetAttribute((const void *)crt_solve16_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(crt_solve16_kernel, dim3((nrhs + waves - 1) / waves), dim3(64 * waves), lds, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
