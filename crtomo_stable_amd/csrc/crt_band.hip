// crt_band.hip -- batched complex-symmetric band factorisation and substitution.
//
// One workgroup factorises one (electrode l, wavenumber k) system
//     A_lk = sum_e sigma_e elbg(e,.,k) + sum_rel sigma_rnr(rel) relbg(rel,.) kg(rel,l,k)
// (src/kompab.f90:80-138), scaled to a unit-modulus diagonal
// (src/scalab.f90:52-76), factorised A = L L^T without conjugation
// (src/chol.f90:46-78) and forward-substituted y = L^{-1} b with the unit
// source b = -fak(enr(l)) e_enr(l) (src/kompab.f90:162, src/kompb.f90:58).
// The back substitution L^T x = -y and kpot = x * fak (src/vre.f90:45-69,
// src/fem.f90:383-394) run in crt_solve_kernel, one wavefront per right-hand
// side.
//
// Storage.  Rows are grouped in blocks of NB = 8.  The factor is written as
// block-column panels panel[J][a][r][c] = L(8(J+a)+r, 8J+c), a = 0..D with
// D = ceil(mb/8), so a panel is one contiguous (D+1) KB record.  During the
// factorisation the active window (block rows J..J+D+1, lower triangle) sits
// in LDS, one 8x8 tile per slot; block (I,K) with d = I-K lives in slot
// base[d] + K mod (D+2-d), so the window slides by one block per step with no
// data movement.  The matrix is never materialised in HBM: each entering row
// block is assembled from the CSR contribution lists straight into LDS.
#include "crt_common.hpp"

#define NB CRT_NB
#define NB2 (NB * NB)

namespace {

__device__ __forceinline__ zd shfl_z(zd v, int src) {
    return zd{__shfl(v.re, src, 64), __shfl(v.im, src, 64)};
}
__device__ __forceinline__ zd shfl_xor_z(zd v, int m) {
    return zd{__shfl_xor(v.re, m, 64), __shfl_xor(v.im, m, 64)};
}

// Sum of the contributions to A(i,j) in the reference's element order.
// Contraction is off so the sum is formed exactly as the Fortran loop does.
__device__ __forceinline__ zd assemble_entry(const DevGrid &g, int e0, int e1, int l, int k) {
#pragma clang fp contract(off)
    zd s = zd{0.0, 0.0};
    for (int q = e0; q < e1; ++q) {
        int c = g.contrib[q];
        double dum;
        zd sg;
        if (c >= 0) {
            int e = c / g.npk, imn = c - e * g.npk;
            dum = g.elbg[((size_t)k * g.npk + imn) * g.elanz + e];
            sg = g.sigma[e];
        } else {
            int cc = -c - 1;
            int rel = cc / g.npk, imn = cc - rel * g.npk;
            dum = g.relbg[(size_t)imn * g.relanz + rel] * g.kg[((size_t)k * g.eanz + l) * g.relanz + rel];
            sg = g.sigma[g.rnr[rel]];
        }
        s.re = s.re + dum * sg.re;
        s.im = s.im + dum * sg.im;
    }
    return s;
}

// A(i,j) of system (l, k) for CSR entry q: the wavenumber's area sum (when the grid has
// one) continued with the entry's boundary contributions, the same operations in the
// same order as assemble_entry, so the value is identical
__device__ __forceinline__ zd assemble_q(const DevGrid &g, int q, int l, int k) {
#pragma clang fp contract(off)
    // randb / randb2 with zero values (src/randb.f90:40-84): identity row and column
    if (g.dmask) {
        const unsigned char m = g.dmask[q];
        if (m) return zd{m == 2 ? 1.0 : 0.0, 0.0};
    }
    if (!g.area) return assemble_entry(g, g.cptr[q], g.cptr[q + 1], l, k);
    zd s = g.area[(size_t)k * g.nnz + q];
    const int e1 = g.cptr[q + 1];
    for (int p = g.csplit[q]; p < e1; ++p) {
        const int cc = -g.contrib[p] - 1;
        const int rel = cc / g.npk, imn = cc - rel * g.npk;
        const double dum = g.relbg[(size_t)imn * g.relanz + rel] * g.kg[((size_t)k * g.eanz + l) * g.relanz + rel];
        const zd sg = g.sigma[g.rnr[rel]];
        s.re = s.re + dum * sg.re;
        s.im = s.im + dum * sg.im;
    }
    return s;
}

}  // namespace

// -------------------------------------------------------------- area sums
// area[k*nnz + q]: the area-element contributions of entry q at wavenumber k
// (src/kompab.f90:113-121 for the elements of types 3, 5, 8), one thread per (k, q)
__global__ void __launch_bounds__(256) crt_area_kernel(DevGrid g, int K, zd *__restrict__ area) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)K * g.nnz) return;
    const int k = (int)(t / g.nnz), q = (int)(t - (long)k * g.nnz);
    area[t] = assemble_entry(g, g.cptr[q], g.csplit[q], 0, k);
}

int crt_launch_area(hipStream_t st, const DevGrid &g, int K, zd *area) {
    const long n = (long)K * g.nnz;
    hipLaunchKernelGGL(crt_area_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, K, area);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// ------------------------------------------------------------------ diagonal
// fak_i = 1/sqrt|a_ii|, dga_i = |a_ii| per system (src/scalab.f90:54-66)
__global__ void __launch_bounds__(256) crt_diag_kernel(DevGrid g, int nsys, const int2 *__restrict__ sys,
                                                       double *__restrict__ fak, double *__restrict__ dga,
                                                       int *__restrict__ status) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nsys * g.sanz) return;
    int s = (int)(t / g.sanz), i = (int)(t - (long)s * g.sanz);
    int2 lk = sys[s];
    int q = g.row_ptr[i + 1] - 1;  // diagonal is the last CSR entry of the row
    zd a = assemble_q(g, q, lk.x, lk.y);
    double dum = zabs(a);
    size_t o = (size_t)s * g.fstride + i;
    if (!(dum > 0.0) || (g.dc && !(a.re > 0.0))) {
        status[s] = CRT_ERR_ZERODIAG;  // plain store: any non-zero code fails the call
        dum = 1.0;
    }
    fak[o] = 1.0 / sqrt(dum);
    dga[o] = dum;
}

// choldc stops on a pivot s <= 0 (src/choldc.f90:66-70).  The DC forward runs the
// complex factorisation on real matrices: while every s > 0 the pivots and their
// reciprocals stay real and positive, and the first s < 0 makes its square root, and
// every reciprocal pivot after it, non-real (s = 0 is caught by the factor kernels).
// One thread per (system, row) over the reciprocal pivots dinv.
template <class Z>
__global__ void __launch_bounds__(256) crt_dc_pivot_kernel(int nsys, int sanz, int fstride, const Z *__restrict__ dinv,
                                                           int *__restrict__ status) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nsys * sanz) return;
    const int s = (int)(t / sanz), i = (int)(t - (long)s * sanz);
    const Z r = dinv[(size_t)s * fstride + i];
    // an earlier error of this system (scaldc's 27) stays: the routines report in order
    if ((r.im != 0 || !(r.re > 0)) && status[s] == 0) status[s] = CRT_ERR_PIVOT;
}

// ----------------------------------------------------------------- assemble
// Scaled lower-triangle nonzeros of every system, one thread per (system, entry):
// a_ii/|a_ii| on the diagonal, a_ij fak_i fak_j elsewhere (src/scalab.f90:52-76).
__global__ void __launch_bounds__(256) crt_assemble_kernel(DevGrid g, int nsys, const int2 *__restrict__ sys,
                                                           const double *__restrict__ fak,
                                                           const double *__restrict__ dga, zd *__restrict__ vals) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nsys * g.vstride) return;
    int s = (int)(t / g.vstride), q = (int)(t - (long)s * g.vstride);
    if (q >= g.nnz) {
        // the two constants after the nonzeros (zero entries and padding diagonals)
        vals[t] = zd{q == g.nnz ? 0.0 : 1.0, 0.0};
        return;
    }
    int2 lk = sys[s];
    int i = g.crow[q], j = g.col[q];
    zd a = assemble_q(g, q, lk.x, lk.y);
    const double *fs = fak + (size_t)s * g.fstride;
    zd v;
    if (j == i) {
        double dum = dga[(size_t)s * g.fstride + i];
        v = zd{a.re / dum, a.im / dum};
    } else {
        double f = fs[i] * fs[j];
        v = zd{a.re * f, a.im * f};
    }
    vals[t] = v;
}

// --------------------------------------------------------------- factorise
struct FactorArgs {
    DevGrid g;
    int nsys;
    const int2 *sys;    // (l, k) per system, 0-based
    const double *fak;  // [sys][nblk*NB]
    const double *dga;
    const zd *vals;     // [sys][nnz] scaled nonzeros (crt_assemble_kernel)
    zd *panels;         // [sys][nblk][D+1][NB][NB]
    zd *y;              // [sys][nblk*NB]
    int *status;
    const zd *band;     // optional: factor a given scaled band matrix (reference layout) instead
};

__global__ void __launch_bounds__(256) crt_factor_kernel(FactorArgs A) {
    extern __shared__ double smem_d[];
    const DevGrid &g = A.g;
    const int D = g.D, nblk = g.nblk, tid = threadIdx.x;
    const int s = blockIdx.x;
    if (s >= A.nsys) return;
    const int l = A.sys[s].x, k = A.sys[s].y;
    const int enode = A.band ? -1 : g.enr[l];
    const double *fak = A.band ? nullptr : A.fak + (size_t)s * nblk * NB;
    const zd *vals = A.band ? nullptr : A.vals + (size_t)s * g.vstride;
    zd *panel = A.panels + (size_t)s * nblk * (D + 1) * NB2;
    zd *yout = A.y + (size_t)s * nblk * NB;

    // LDS carve-up
    const int nslot = (D + 1) * (D + 4) / 2;
    zd *slots = reinterpret_cast<zd *>(smem_d);               // nslot * 64
    zd *bring = slots + (size_t)nslot * NB2;                  // (D+2) * 8
    zd *ljj = bring + (D + 2) * NB;                           // 64 (factored diag block)
    zd *linv = ljj + NB2;                                     // 8 inverse pivots
    zd *yj = linv + NB;                                       // 8
    int *base = reinterpret_cast<int *>(yj + NB);             // D+1
    int *jm = base + (D + 1);                                 // D+1 : J mod (D+2-d)
    unsigned char *pairs = reinterpret_cast<unsigned char *>(jm + (D + 1));  // 2*npairs
    const int npairs = D * (D + 1) / 2;

    if (tid == 0) {
        int b = 0;
        for (int d = 0; d <= D; ++d) { base[d] = b; b += D + 2 - d; jm[d] = 0; }
        int p = 0;
        for (int a = 1; a <= D; ++a)
            for (int bb = 1; bb <= a; ++bb) { pairs[2 * p] = (unsigned char)a; pairs[2 * p + 1] = (unsigned char)bb; ++p; }
    }
    __syncthreads();

    auto slot_of = [&](int I, int K) -> zd * {
        int d = I - K, m = D + 2 - d;
        return slots + (size_t)(base[d] + K % m) * NB2;
    };

    // Entries of row block I: the CSR entries [row_ptr[8I], row_ptr[8I+8]) are
    // contiguous; a given band matrix contributes (mb+1) entries per row.
    auto block_count = [&](int I) -> int {
        if (A.band) return NB * (g.mb + 1);
        return g.row_ptr[min(I * NB + NB, g.sanz)] - g.row_ptr[I * NB];
    };
    auto load_entry = [&](int I, int t, int &i, int &j, zd &v) -> bool {
        if (A.band) {
            int r = t / (g.mb + 1), o = t - r * (g.mb + 1);
            i = I * NB + r;
            j = i - g.mb + o;
            if (i >= g.sanz || j < 0) return false;
            v = A.band[(size_t)i * (g.mb + 1) + o];
            return true;
        }
        int q = g.row_ptr[I * NB] + t;
        i = g.crow[q];
        j = g.col[q];
        v = vals[q];
        return true;
    };
    auto put_entry = [&](int I, int i, int j, zd v) {
        slot_of(I, j / NB)[(i - I * NB) * NB + (j % NB)] = v;
    };
    // unit diagonal of padding rows and the unit source of the RHS block
    auto block_extras = [&](int I, int t0) {
        if (t0 < NB) {
            int i = I * NB + t0;
            if (i >= g.sanz) slot_of(I, I)[t0 * NB + t0] = zd{1.0, 0.0};
            zd bv = zd{0.0, 0.0};
            if (i < g.sanz && i == enode) bv = zd{-fak[i], 0.0};
            bring[(I % (D + 2)) * NB + t0] = bv;
        }
    };
    // direct scatter of a whole block (set-up)
    auto scatter_rows = [&](int I, int t0, int nthr) {
        const int cnt = block_count(I);
        for (int t = t0; t < cnt; t += nthr) {
            int i, j;
            zd v;
            if (load_entry(I, t, i, j, v)) put_entry(I, i, j, v);
        }
        block_extras(I, t0);
    };

    // ---- initial window: row blocks 0..min(D, nblk-1)
    for (int q = tid; q < nslot * NB2; q += blockDim.x) slots[q] = zd{0.0, 0.0};
    __syncthreads();
    for (int I = 0; I <= D && I < nblk; ++I) scatter_rows(I, tid, blockDim.x);
    __syncthreads();

    const int wave = tid >> 6, lane = tid & 63;
    const int ntrsm = ((NB * D + 63) / 64) * 64;  // threads of the TRSM phase
    const int nspare = (int)blockDim.x - ntrsm;    // threads that assemble the entering block
    // one prefetched CSR entry per spare thread, loaded a step ahead
    int pf_i = -1, pf_j = 0;
    zd pf_v = zd{0.0, 0.0};
    auto prefetch = [&](int I) {
        pf_i = -1;
        if (tid < ntrsm || I >= nblk) return;
        int t = tid - ntrsm;
        if (t < block_count(I) && !load_entry(I, t, pf_i, pf_j, pf_v)) pf_i = -1;
    };
    prefetch(D + 1);
    int err = 0;
    for (int J = 0; J < nblk; ++J) {
        const int Ie = J + D + 1;  // entering row block
        // ---------------- phase 1: diag block (wave 0) | zero entering slots (others)
        if (wave == 0) {
            const int r = lane >> 3, c = lane & 7;
            zd *dj = slot_of(J, J);
            zd v = dj[lane];
            zd yv = bring[(J % (D + 2)) * NB + (lane & 7)];
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) {
                zd pv = shfl_z(v, cc * NB + cc);
                zd piv = zsqrt(pv);
                if (zabs(pv) <= 0.0 || (A.g.dc && !(pv.re > 0.0 && pv.im == 0.0))) err = CRT_ERR_PIVOT;
                zd pinv = zinv(piv);
                if (r == cc && c == cc) v = piv;
                if (r > cc && c == cc) v = zmul(v, pinv);
                zd lr = shfl_z(v, r * NB + cc);   // L(r, cc)
                zd lc = shfl_z(v, c * NB + cc);   // L(c, cc)
                if (r > cc && c > cc && c <= r) v = zfms(v, lr, lc);
                // forward substitution on the RHS block (lanes 0..7 hold y(c))
                zd ycc = shfl_z(yv, cc);
                ycc = zmul(ycc, pinv);
                if (lane == cc) yv = ycc;
                zd lcol = shfl_z(v, (lane & 7) * NB + cc);  // L(lane, cc) after division
                if ((lane & 7) > cc) yv = zfms(yv, lcol, ycc);
                if (lane == 0) linv[cc] = pinv;
            }
            if (c <= r) dj[lane] = v;
            ljj[lane] = (c <= r) ? v : zd{0.0, 0.0};
            if (lane < NB) {
                yj[lane] = yv;
                bring[(J % (D + 2)) * NB + lane] = yv;
                yout[(size_t)J * NB + lane] = yv;
            }
        } else if (Ie < nblk) {
            // zero the slots of the entering row block (they held column J-1)
            int nthr = blockDim.x - 64, t0 = tid - 64;
            for (int q = t0; q < (D + 1) * NB2; q += nthr) {
                int K = Ie - D + q / NB2;  // K = J+1 .. J+D+1
                slot_of(Ie, K)[q % NB2] = zd{0.0, 0.0};
            }
        }
        __syncthreads();
        // ---------------- phase 2: TRSM + RHS update (rows 8..8D+7) | assemble entering block
        {
            int nrow = NB * D;
            if (tid < nrow) {
                int a = 1 + tid / NB, r = tid % NB;
                if (J + a < nblk) {
                    zd *blk = slot_of(J + a, J);
                    zd xr[NB];
#pragma unroll
                    for (int c = 0; c < NB; ++c) xr[c] = blk[r * NB + c];
                    zd bacc = bring[((J + a) % (D + 2)) * NB + r];
#pragma unroll
                    for (int c = 0; c < NB; ++c) {
                        zd sacc = xr[c];
#pragma unroll
                        for (int cp = 0; cp < c; ++cp) sacc = zfms(sacc, xr[cp], ljj[c * NB + cp]);
                        xr[c] = zmul(sacc, linv[c]);
                        bacc = zfms(bacc, xr[c], yj[c]);
                    }
#pragma unroll
                    for (int c = 0; c < NB; ++c) blk[r * NB + c] = xr[c];
                    bring[((J + a) % (D + 2)) * NB + r] = bacc;
                }
            } else if (tid >= ntrsm && Ie < nblk) {
                const int t0 = tid - ntrsm;
                if (pf_i >= 0) put_entry(Ie, pf_i, pf_j, pf_v);
                // entries beyond one per spare thread are loaded directly
                const int cnt = block_count(Ie);
                for (int t = nspare + t0; t < cnt; t += nspare) {
                    int i, j;
                    zd v;
                    if (load_entry(Ie, t, i, j, v)) put_entry(Ie, i, j, v);
                }
                block_extras(Ie, t0);
                prefetch(Ie + 1);
            }
            if (nspare <= 0 && Ie < nblk) {
                __syncthreads();
                scatter_rows(Ie, tid, blockDim.x);
            }
        }
        __syncthreads();
        // ---------------- phase 3: store panel J | rank-8 update of the trailing window
        {
            zd *pj = panel + (size_t)J * (D + 1) * NB2;
            for (int q = tid; q < (D + 1) * NB2; q += blockDim.x) {
                int a = q / NB2;
                pj[q] = (J + a < nblk) ? slot_of(J + a, J)[q % NB2] : zd{0.0, 0.0};
            }
            int nitem = npairs * 16;
            for (int it = tid; it < nitem; it += blockDim.x) {
                int p = it >> 4, tile = it & 15;
                int a = pairs[2 * p], b = pairs[2 * p + 1];
                if (J + a >= nblk) continue;
                int r0 = (tile >> 2) * 2, c0 = (tile & 3) * 2;
                if (a == b && c0 > r0 + 1) continue;  // strictly upper tile of a diagonal block
                const zd *La = slot_of(J + a, J);
                const zd *Lb = slot_of(J + b, J);
                zd *T = slot_of(J + a, J + b);
                zd t00 = T[r0 * NB + c0], t01 = T[r0 * NB + c0 + 1];
                zd t10 = T[(r0 + 1) * NB + c0], t11 = T[(r0 + 1) * NB + c0 + 1];
#pragma unroll
                for (int q = 0; q < NB; ++q) {
                    zd a0 = La[r0 * NB + q], a1 = La[(r0 + 1) * NB + q];
                    zd b0 = Lb[c0 * NB + q], b1 = Lb[(c0 + 1) * NB + q];
                    t00 = zfms(t00, a0, b0); t01 = zfms(t01, a0, b1);
                    t10 = zfms(t10, a1, b0); t11 = zfms(t11, a1, b1);
                }
                T[r0 * NB + c0] = t00; T[r0 * NB + c0 + 1] = t01;
                T[(r0 + 1) * NB + c0] = t10; T[(r0 + 1) * NB + c0 + 1] = t11;
            }
        }
        __syncthreads();
    }
    if (wave == 0 && lane == 0 && err && A.status[s] == 0) A.status[s] = err;
}

size_t crt_factor_lds_bytes(int D) {
    size_t nslot = (size_t)(D + 1) * (D + 4) / 2;
    size_t b = nslot * NB2 * sizeof(zd) + (size_t)(D + 2) * NB * sizeof(zd) + NB2 * sizeof(zd) +
               2 * NB * sizeof(zd) + 2 * (D + 1) * sizeof(int) + (size_t)D * (D + 1) + 16;
    return (b + 15) & ~(size_t)15;
}

// ------------------------------------------------------------------- solve
// One wavefront per right-hand side.  If have_y, y is the forward result of the
// factorisation; otherwise the forward sweep is run from b = -fak(enode) e_enode.
// Back substitution L^T x = -y, then kpot = x * fak written node-major:
// kpotN[(node*K + k)*eanz + l] (src/vre.f90:59-69, src/fem.f90:383-394).
struct SolveArgs {
    int sanz, nblk, D, kwnanz, eanz;
    int nrhs;
    const int4 *rhs;     // (l, k, factor system index, have_y)
    const int *enr;
    const double *fak;   // [fsys][nblk*NB]
    const zd *panels;    // [fsys][nblk][D+1][64]
    const zd *y;         // [fsys][nblk*NB]
    zd *yscr;            // [nrhs][nblk*NB] forward result of right-hand sides without have_y
    zd *kpotN;
    const zd *bdense;    // optional dense right-hand side (drop-in vre); fak = 1
    zd *pot;             // optional dense solution output (drop-in vre)
};

__global__ void __launch_bounds__(256) crt_solve_kernel(SolveArgs A) {
    extern __shared__ double smem_s[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wpb = blockDim.x >> 6;
    const int rid = blockIdx.x * wpb + wave;
    if (rid >= A.nrhs) return;
    const int D = A.D, nblk = A.nblk;
    const int4 rq = A.rhs[rid];
    const int l = rq.x, k = rq.y, fs = rq.z;
    const double *fak = A.fak ? A.fak + (size_t)fs * nblk * NB : nullptr;
    const zd *panel = A.panels + (size_t)fs * nblk * (D + 1) * NB2;
    zd *ring = reinterpret_cast<zd *>(smem_s) + (size_t)wave * (D + 1) * NB;
    const int r = lane >> 3, c = lane & 7;
    const int enode = A.bdense ? -1 : A.enr[l];
    zd *yscr = A.yscr + (size_t)rid * nblk * NB;

    if (!rq.w) {
        // forward sweep: y_J = L_JJ^{-1} (b_J - sum_a L(J, J-a) y_{J-a})
        for (int J = 0; J < nblk; ++J) {
            zd acc = zd{0.0, 0.0};
            for (int a = 1; a <= D && a <= J; ++a) {
                zd Lv = panel[((size_t)(J - a) * (D + 1) + a) * NB2 + lane];  // L(J*8+r, (J-a)*8+c)
                zd yv = ring[((J - a) % (D + 1)) * NB + c];
                acc = zfma(acc, Lv, yv);
            }
            // reduce over c (lanes r*8 .. r*8+7)
            acc = zadd(acc, shfl_xor_z(acc, 1));
            acc = zadd(acc, shfl_xor_z(acc, 2));
            acc = zadd(acc, shfl_xor_z(acc, 4));
            int i = J * NB + r;
            zd bv = zd{0.0, 0.0};
            if (A.bdense) {
                if (i < A.sanz) bv = A.bdense[i];
            } else if (i < A.sanz && i == enode) {
                bv = zd{-fak[i], 0.0};
            }
            zd rhs = zsub(bv, acc);  // valid on lanes with c == 0 for row r
            zd Ld = panel[((size_t)J * (D + 1)) * NB2 + lane];
            zd yv = shfl_z(rhs, (lane & 7) * NB);  // lane holds rhs of row (lane & 7)
#pragma unroll
            for (int cc = 0; cc < NB; ++cc) {
                zd piv = shfl_z(Ld, cc * NB + cc);
                zd ycc = zdiv(shfl_z(yv, cc), piv);
                if ((lane & 7) == cc) yv = ycc;
                zd lcol = shfl_z(Ld, (lane & 7) * NB + cc);
                if ((lane & 7) > cc) yv = zfms(yv, lcol, ycc);
            }
            if (lane < NB) {
                ring[(J % (D + 1)) * NB + lane] = yv;
                yscr[(size_t)J * NB + lane] = yv;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // backward sweep: L^T x = -y ; the ring holds x of the D blocks below J
    const zd *yglob = rq.w ? A.y + (size_t)fs * nblk * NB : yscr;
    for (int J = nblk - 1; J >= 0; --J) {
        zd acc = zd{0.0, 0.0};
        for (int a = 1; a <= D && J + a < nblk; ++a) {
            zd Lv = panel[((size_t)J * (D + 1) + a) * NB2 + lane];  // L((J+a)*8+r, J*8+c)
            zd xv = ring[((J + a) % (D + 1)) * NB + r];
            acc = zfma(acc, Lv, xv);
        }
        // reduce over r (lanes c, c+8, ..., c+56)
        acc = zadd(acc, shfl_xor_z(acc, 8));
        acc = zadd(acc, shfl_xor_z(acc, 16));
        acc = zadd(acc, shfl_xor_z(acc, 32));
        zd yJ = yglob[(size_t)J * NB + c];
        zd rhs = zsub(zd{-yJ.re, -yJ.im}, acc);  // lanes 0..7 hold rhs(c)
        zd Ld = panel[((size_t)J * (D + 1)) * NB2 + lane];
        zd xv = rhs;
#pragma unroll
        for (int cc = NB - 1; cc >= 0; --cc) {
            zd piv = shfl_z(Ld, cc * NB + cc);
            zd xcc = zdiv(shfl_z(xv, cc), piv);
            if ((lane & 7) == cc) xv = xcc;
            zd lrow = shfl_z(Ld, cc * NB + (lane & 7));  // L(cc, lane&7): row cc, column < cc
            if ((lane & 7) < cc) xv = zfms(xv, lrow, xcc);
        }
        __builtin_amdgcn_wave_barrier();
        if (lane < NB) {
            ring[(J % (D + 1)) * NB + lane] = xv;
            int i = J * NB + lane;
            if (i < A.sanz) {
                if (A.pot) {
                    A.pot[i] = xv;
                } else {
                    double f = fak[i];
                    A.kpotN[((size_t)i * A.kwnanz + k) * A.eanz + l] = zd{xv.re * f, xv.im * f};
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

size_t crt_solve_lds_bytes(int D, int waves) { return (size_t)waves * (D + 1) * NB * sizeof(zd); }

int crt_launch_diag(hipStream_t st, const DevGrid &g, int nsys, const int2 *sys, double *fak, double *dga,
                    int *status) {
    long n = (long)nsys * g.sanz;
    hipLaunchKernelGGL(crt_diag_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nsys, sys,
                       fak, dga, status);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

template <class Z>
static int launch_dc_pivot(hipStream_t st, int nsys, int sanz, int fstride, const Z *dinv, int *status) {
    const long n = (long)nsys * sanz;
    hipLaunchKernelGGL(crt_dc_pivot_kernel<Z>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, nsys, sanz,
                       fstride, dinv, status);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
int crt_launch_dc_pivot(hipStream_t st, int nsys, int sanz, int fstride, const zd *dinv, int *status) {
    return launch_dc_pivot(st, nsys, sanz, fstride, dinv, status);
}
int crt_launch_dc_pivot32(hipStream_t st, int nsys, int sanz, int fstride, const zf *dinv, int *status) {
    return launch_dc_pivot(st, nsys, sanz, fstride, dinv, status);
}

// kpotN[(node K + k) E + l] = 0 for every node of the listed (l, k) systems
__global__ void __launch_bounds__(256) crt_zero_sources_kernel(int sanz, int K, int E, int np,
                                                               const int2 *__restrict__ pairs, zd *__restrict__ kpotN) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)np * sanz) return;
    const int p = (int)(t / sanz), j = (int)(t - (long)p * sanz);
    const int2 lk = pairs[p];
    kpotN[((size_t)j * K + lk.y) * E + lk.x] = zd{0.0, 0.0};
}

int crt_launch_zero_sources(hipStream_t st, int sanz, int K, int E, int np, const int2 *pairs, zd *kpotN) {
    const long n = (long)np * sanz;
    hipLaunchKernelGGL(crt_zero_sources_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sanz, K, E, np,
                       pairs, kpotN);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_assemble(hipStream_t st, const DevGrid &g, int nsys, const int2 *sys, const double *fak,
                        const double *dga, zd *vals) {
    long n = (long)nsys * g.vstride;
    hipLaunchKernelGGL(crt_assemble_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nsys, sys, fak,
                       dga, vals);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_factor(hipStream_t st, const DevGrid &g, int nsys, const int2 *sys, const double *fak,
                      const double *dga, const zd *vals, zd *panels, zd *y, int *status) {
    FactorArgs a{g, nsys, sys, fak, dga, vals, panels, y, status, nullptr};
    size_t lds = crt_factor_lds_bytes(g.D);
    if (lds > 160 * 1024) return CRT_ERR_UNSUPPORTED;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)crt_factor_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                            160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL(crt_factor_kernel, dim3(nsys), dim3(256), lds, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// panel[J][a][r][c] <-> reference band layout a[i*(mb+1) + j-i+mb] (lower factor)
__global__ void __launch_bounds__(256) crt_panel_band_kernel(int sanz, int mb, int nblk, int D, int to_band,
                                                             zd *__restrict__ panel, zd *__restrict__ band) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)sanz * (mb + 1)) return;
    int i = (int)(t / (mb + 1)), o = (int)(t - (long)i * (mb + 1));
    int j = i - mb + o;
    if (j < 0) {
        if (to_band) band[t] = zd{0.0, 0.0};
        return;
    }
    int J = j / NB, a = i / NB - J;
    size_t pi = (((size_t)J * (D + 1) + a) * NB + (i % NB)) * NB + (j % NB);
    if (to_band) band[t] = panel[pi];
    else panel[pi] = band[t];
}

int crt_launch_panel_band(hipStream_t st, int sanz, int mb, int nblk, int D, int to_band, zd *panel, zd *band) {
    long n = (long)sanz * (mb + 1);
    hipLaunchKernelGGL(crt_panel_band_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sanz, mb, nblk, D,
                       to_band, panel, band);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// factor one scaled band matrix given in the reference layout (drop-in chol)
int crt_launch_factor_band(hipStream_t st, const DevGrid &g, const int2 *sys, const zd *band, zd *panels, zd *y,
                           int *status) {
    FactorArgs a{g, 1, sys, nullptr, nullptr, nullptr, panels, y, status, band};
    size_t lds = crt_factor_lds_bytes(g.D);
    if (lds > 160 * 1024) return CRT_ERR_UNSUPPORTED;
    (void)hipFuncSetAttribute((const void *)crt_factor_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    hipLaunchKernelGGL(crt_factor_kernel, dim3(1), dim3(256), lds, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_solve(hipStream_t st, int sanz, int nblk, int D, int kwnanz, int eanz, int nrhs,
                     const int4 *rhs, const int *enr, const double *fak, const zd *panels, const zd *y,
                     zd *yscr, zd *kpotN, const zd *bdense, zd *pot) {
    SolveArgs a{sanz, nblk, D, kwnanz, eanz, nrhs, rhs, enr, fak, panels, y, yscr, kpotN, bdense, pot};
    const int waves = 4;
    size_t lds = crt_solve_lds_bytes(D, waves);
    hipLaunchKernelGGL(crt_solve_kernel, dim3((nrhs + waves - 1) / waves), dim3(64 * waves), lds, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// ------------------------------------------------ general right-hand sides
// The right-hand side of system r = (l, k, factor fs) when it is not a unit source,
// scaled as src/scalab.f90:64 scales it (yscr[r][i] = fak[fs][i] b_i, zero on the
// padding rows):
//   lsr:   b_i = sum over the entries q of row and column i of the system's matrix of
//          sum_p dum_p (sigma_p - sigma0) pota_r(other node of q), dum_p and sigma_p of
//          the entry's contributions (the singularity removal of src/kompab.f90:129-133:
//          the matrix of sigma - sigma0 times the analytic potential);
//   else:  b = -e_enr(l) (src/kompab.f90:162, src/kompb.f90:58);
//   lsink: b_nsink = 1 (src/kompab.f90:183, src/kompb.f90:59), set last.
// tptr/tq: per node the strictly lower entries whose column it is.
__global__ void __launch_bounds__(256) crt_gen_rhs_kernel(DevGrid g, int nrhs, const int4 *__restrict__ rhs, int lsr,
                                                          const zd *__restrict__ sigma0, const zd *__restrict__ pota,
                                                          size_t pstride, const int *__restrict__ tptr,
                                                          const int *__restrict__ tq, int nsink0,
                                                          const double *__restrict__ fak, int ystride,
                                                          zd *__restrict__ yscr) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nrhs * ystride) return;
    const int r = (int)(t / ystride), i = (int)(t - (long)r * ystride);
    zd *yo = yscr + (size_t)r * ystride;
    if (i >= g.sanz) {
        yo[i] = zd{0.0, 0.0};
        return;
    }
    const int4 q4 = rhs[r];
    const int l = q4.x, k = q4.y;
    zd b = zd{0.0, 0.0};
    if (lsr) {
        const zd s0 = *sigma0;
        const zd *p = pota + (size_t)r * pstride;
        auto dq = [&](int q) {
            zd d = zd{0.0, 0.0};
            for (int pp = g.cptr[q]; pp < g.cptr[q + 1]; ++pp) {
                const int c = g.contrib[pp];
                double dum;
                zd sg;
                if (c >= 0) {
                    const int e = c / g.npk, imn = c - e * g.npk;
                    dum = g.elbg[((size_t)k * g.npk + imn) * g.elanz + e];
                    sg = g.sigma[e];
                } else {
                    const int cc = -c - 1;
                    const int rel = cc / g.npk, imn = cc - rel * g.npk;
                    dum = g.relbg[(size_t)imn * g.relanz + rel] * g.kg[((size_t)k * g.eanz + l) * g.relanz + rel];
                    sg = g.sigma[g.rnr[rel]];
                }
                d.re += dum * (sg.re - s0.re);
                d.im += dum * (sg.im - s0.im);
            }
            return d;
        };
        for (int q = g.row_ptr[i]; q < g.row_ptr[i + 1]; ++q) b = zfma(b, dq(q), p[g.col[q]]);
        for (int u = tptr[i]; u < tptr[i + 1]; ++u) {
            const int q = tq[u];
            b = zfma(b, dq(q), p[g.crow[q]]);
        }
    } else if (i == g.enr[l]) {
        b = zd{-1.0, 0.0};
    }
    if (i == nsink0) b = zd{1.0, 0.0};
    const double f = fak[(size_t)q4.z * g.fstride + i];
    yo[i] = zd{b.re * f, b.im * f};
}

// kpot of the systems r: + pota (the singularity removal's analytic part, src/fem.f90:389)
// or = pota (the analytic solution alone, src/fem.f90:384-386)
__global__ void __launch_bounds__(256) crt_pota_kpot_kernel(int sanz, int K, int E, int nrhs,
                                                            const int4 *__restrict__ rhs, const zd *__restrict__ pota,
                                                            size_t pstride, int add, zd *__restrict__ kpotN) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nrhs * sanz) return;
    const int r = (int)(t / sanz), i = (int)(t - (long)r * sanz);
    const int4 q4 = rhs[r];
    zd *o = kpotN + ((size_t)i * K + q4.y) * E + q4.x;
    const zd p = pota[(size_t)r * pstride + i];
    *o = add ? zd{o->re + p.re, o->im + p.im} : p;
}

int crt_launch_gen_rhs(hipStream_t st, const DevGrid &g, int nrhs, const int4 *rhs, int lsr, const zd *sigma0,
                       const zd *pota, size_t pstride, const int *tptr, const int *tq, int nsink0, const double *fak,
                       int ystride, zd *yscr) {
    const long n = (long)nrhs * ystride;
    if (n == 0) return 0;
    hipLaunchKernelGGL(crt_gen_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nrhs, rhs, lsr,
                       sigma0, pota, pstride, tptr, tq, nsink0, fak, ystride, yscr);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_pota_kpot(hipStream_t st, int sanz, int K, int E, int nrhs, const int4 *rhs, const zd *pota,
                         size_t pstride, int add, zd *kpotN) {
    const long n = (long)nrhs * sanz;
    if (n == 0) return 0;
    hipLaunchKernelGGL(crt_pota_kpot_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sanz, K, E, nrhs,
                       rhs, pota, pstride, add, kpotN);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
