// crt_condense.hip -- the forward by static condensation onto the mixed-boundary nodes.
//
// The system of source l at wavenumber k (src/kompab.f90:80-138)
//     A(l,k) = sum_e sigma_e elbg(e,.,k) + sum_rel sigma_rnr(rel) relbg(rel,.) kg(rel,l,k)
// depends on the source only through the type-11 boundary terms, and those couple
// only nodes of type-11 elements.  With the nodes split into G (every node of a
// type-11 element) and I (the rest), the interior blocks A_II and A_IG are the same
// for every source, and A x = b is solved exactly (b = e_enr(l): the unit source of
// src/kompb.f90:58 with the sign of the back substitution, src/vre.f90:45-69):
//     A_II = L L^T                          one band factorisation per wavenumber
//     [Z | U] = A_II^-1 [A_IG | b_I]        multi-right-hand-side band substitution
//     S(l) = A_GG(l) - A_GI Z               dense, complex symmetric
//     S(l) x_G = b_G - A_GI U(:,l)          dense factorisation per (l, k)
//     x_I = U(:,l) - Z x_G                  matrix product
// This is block Gaussian elimination of the reference's system with the boundary
// nodes numbered last: the potentials equal those of the band Cholesky of the whole
// system (src/chol.f90, src/vre.f90) up to rounding.  Rows are scaled symmetrically
// by f = 1/sqrt|a_ii| (src/scalab.f90:52-76: on the interior rows the same numbers;
// on the boundary rows the area part of the diagonal, which is shared by the
// sources), so the kernels work on F A F and kpot = f x_hat.
//
// Tiles are 16 x 16 complex FP64 on v_mfma_f64_16x16x4f64; a complex product is
// four real ones.  Lanes: A[i][k] at lane i + 16(k%4) of step k/4, B[k][j] at lane
// j + 16(k%4), D[i][j] at lane j + 16(i%4), element i/4.
#include "crt_common.hpp"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int T16 = 16;

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// (cr, ci) += a b for one k-step of complex operands
__device__ __forceinline__ void cmma(d4 &cr, d4 &ci, zd a, zd b) {
    cr = mfma4(a.re, b.re, cr);
    cr = mfma4(-a.im, b.im, cr);
    ci = mfma4(a.re, b.im, ci);
    ci = mfma4(a.im, b.re, ci);
}

__device__ __forceinline__ zd ldz(const zd *p) {
    const double2 v = *reinterpret_cast<const double2 *>(p);
    return zd{v.x, v.y};
}
__device__ __forceinline__ void stz(zd *p, zd z) { *reinterpret_cast<double2 *>(p) = make_double2(z.re, z.im); }

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// boundary part of entry q of the full matrix for source l at wavenumber k: the
// type-11 contributions of its list (after the area ones), src/kompab.f90:111-116
__device__ __forceinline__ zd boundary_part(const DevGrid &g, int q, int l, int k) {
    zd s = zd{0.0, 0.0};
    const int e1 = g.cptr[q + 1];
    for (int p = g.csplit[q]; p < e1; ++p) {
        const int cc = -g.contrib[p] - 1;
        const int rel = cc / g.npk, imn = cc - rel * g.npk;
        const double dum = g.relbg[(size_t)imn * g.relanz + rel] * g.kg[((size_t)k * g.eanz + l) * g.relanz + rel];
        const zd sg = g.sigma[g.rnr[rel]];
        s.re += dum * sg.re;
        s.im += dum * sg.im;
    }
    return s;
}

}  // namespace

// ------------------------------------------------------------ small kernels
// area sums of the interior entries: areaI[kk][q'] = area[k0 + kk][qmap[q']]
__global__ void __launch_bounds__(256) crt_cond_gather_kernel(int kc, int k0, int nnzF, int nnzI,
                                                              const int *__restrict__ qmap,
                                                              const zd *__restrict__ area, zd *__restrict__ areaI) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)kc * nnzI) return;
    const int kk = (int)(t / nnzI), q = (int)(t - (long)kk * nnzI);
    areaI[t] = area[(size_t)(k0 + kk) * nnzF + qmap[q]];
}

// fG[kk][cg] = 1/sqrt|area part of a_gg| (1 where it vanishes and for padding)
__global__ void __launch_bounds__(256) crt_cond_fg_kernel(int kc, int k0, int nG, int nGp, int nnzF,
                                                          const int *__restrict__ qdiag,
                                                          const zd *__restrict__ area, double *__restrict__ fG) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kc * nGp) return;
    const int kk = t / nGp, cg = t - kk * nGp;
    double f = 1.0;
    if (cg < nG) {
        const double m = zabs(area[(size_t)(k0 + kk) * nnzF + qdiag[cg]]);
        if (m > 0.0) f = 1.0 / sqrt(m);
    }
    fG[t] = f;
}

// values of the right-hand-side entries of [A_IG | b_I] (scaled): an A_IG entry
// area(q) fI(i') fG(c), a source entry fI(i')
__global__ void __launch_bounds__(256) crt_cond_rhs_kernel(int kc, int k0, int nent, int nnzF, int nGp, int fstrideI,
                                                           const int *__restrict__ erow, const int *__restrict__ ecolg,
                                                           const int *__restrict__ eq, const zd *__restrict__ area,
                                                           const double *__restrict__ fI, const double *__restrict__ fG,
                                                           zd *__restrict__ rv) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)kc * nent) return;
    const int kk = (int)(t / nent), e = (int)(t - (long)kk * nent);
    const double fi = fI[(size_t)kk * fstrideI + erow[e]];
    const int q = eq[e];
    zd v;
    if (q >= 0) {
        const zd a = area[(size_t)(k0 + kk) * nnzF + q];
        const double f = fi * fG[(size_t)kk * nGp + ecolg[e]];
        v = zd{a.re * f, a.im * f};
    } else {
        v = zd{fi, 0.0};
    }
    rv[t] = v;
}

// ------------------------------------------------- multi-RHS band substitution
// X <- L^-1 X (forward) or X <- L^-T X (backward) for the interior factor of each
// wavenumber of the chunk, 16 columns per workgroup.  Factor layout of the band
// kernels: panel[J][a] holds the row-major tile L(J+a, J); the diagonal tile holds
// L on and below its diagonal and L^-1 above it (tile(c, r) = L^-1(r, c), r > c);
// dinv the reciprocal pivots.  Per block J the products with the D off-diagonal
// tiles are split over the waves (wave 0 takes a = 1, waves 1..3 the rest), the
// partial sums meet in LDS, and wave 0 applies the inverse diagonal tile on the
// matrix cores; the last D result tiles stay in an LDS ring (slot J mod D is rewritten
// after the barrier of block J, when its old tile J - D has been read).  The forward
// pass takes its right-hand side from the sparse entry lists (sorted by row; each
// entry is handed to the lane that holds its element) and starts at the tile's first
// nonzero block j0; X above it is zero and is not written (the backward pass reads
// zeros there).  At D <= 7 the 52 KB of LDS let three workgroups share a CU.
struct TrsmArgs {
    int nblk, ldx, ntile, kc, nvirt;
    const zd *panels;   // [kk][nblk][D+1][256]
    const zd *dinv;     // [kk][nblk*16]
    zd *X;              // [kk][ntile*16][ldx]
    const int *tptr;    // ntile+1: entries of each column tile
    const int *erow;    // row (interior index) of each entry
    const int *ecol;    // column within the tile
    const zd *rv;       // [kk][nent]
    int nent;
    const int *j0;      // first nonzero block of each column tile
};

template <int D, bool BWD>
__global__ void __launch_bounds__(256, D <= 7 ? 3 : (D <= 14 ? 2 : 1)) crt_cond_trsm_kernel(TrsmArgs A) {
    constexpr int R = D;                              // ring slots
    constexpr int NA = D >= 2 ? (D - 1 + 2) / 3 : 1;  // tiles per worker wave (a = 2 + w-1 + 3u)
    __shared__ zd ring[R][256];
    __shared__ zd part[2][3][256];
    // XCD-aware order: consecutive virtual ids (tiles of one wavenumber) on one XCD
    const int b = blockIdx.x;
    const int per = A.nvirt / 8;
    const int v = (b % 8) * per + b / 8;
    const int kk = v / A.ntile, t = v - kk * A.ntile;
    if (kk >= A.kc) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int nblk = A.nblk;
    const zd *panel = A.panels + (size_t)kk * nblk * (D + 1) * 256;
    const zd *dinv = A.dinv + (size_t)kk * nblk * T16;
    zd *X = A.X + ((size_t)kk * A.ntile + t) * T16 * A.ldx;
    const int J0 = A.j0[t];
    for (int i = threadIdx.x; i < R * 256; i += 256) (&ring[0][0])[i] = zd{};
    __syncthreads();
    if (J0 >= nblk && !BWD) return;

    // operand loaders.  Forward: A = L(J, J-a) = panel[J-a][a], A[i][k] = tile(i, k).
    // Backward: A = L(J+a, J)^T = panel[J][a]^T, A[i][k] = tile(k, i).
    auto load_off = [&](int J, int a, zd (&o)[4]) {
        const int Jt = BWD ? J : J - a;
        const bool live = BWD ? (J >= 0 && J + a < nblk) : (J < nblk && J - a >= J0);
        const zd *tp = panel + ((size_t)(live ? Jt : 0) * (D + 1) + a) * 256;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            o[ks] = live ? (BWD ? ldz(tp + 64 * ks + lane) : ldz(tp + jc * T16 + 4 * ks + rq)) : zd{};
    };
    // inverse diagonal tile: forward Linv(i, k), backward Linv(k, i)
    auto load_diag = [&](int J, zd (&o)[4]) {
        const bool live = J >= 0 && J < nblk;
        const zd *tp = panel + (size_t)(live ? J : 0) * (D + 1) * 256;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int k = 4 * ks + rq;
            zd e = zd{};
            if (live) {
                if (jc == k) e = ldz(dinv + (size_t)J * T16 + jc);
                else if (BWD ? (k > jc) : (jc > k)) e = BWD ? ldz(tp + jc * T16 + k) : ldz(tp + 64 * ks + lane);
            }
            o[ks] = e;
        }
    };

    if (wave == 0) {
        zd a1[4], a1n[4], dg[4], dgn[4];
        zd xin[4] = {zd{}, zd{}, zd{}, zd{}}, xinn[4];
        int wb = BWD ? 0 : A.tptr[t], wc = 0;
        const int cend = BWD ? 0 : A.tptr[t + 1];
        const zd *rv = BWD ? nullptr : A.rv + (size_t)kk * A.nent;
        int wrow = INT32_MAX, wcol = 0;
        zd wv = zd{};
        auto load_window = [&]() {
            const int p = wb + lane;
            wrow = p < cend ? A.erow[p] : INT32_MAX;
            wcol = p < cend ? A.ecol[p] : 0;
            wv = p < cend ? rv[p] : zd{};
        };
        if (!BWD) load_window();
        auto load_x = [&](int J, zd (&o)[4]) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                o[u] = (J >= J0 && J >= 0 && J < nblk) ? ldz(X + (size_t)jc * A.ldx + (size_t)J * T16 + 4 * u + rq)
                                                       : zd{};
        };
        const int Js = BWD ? nblk - 1 : J0;
        const int step = BWD ? -1 : 1;
        load_off(Js, 1, a1);
        load_diag(Js, dg);
        if (BWD) load_x(Js, xin);
        for (int J = Js; BWD ? J >= 0 : J < nblk; J += step) {
            load_off(J + step, 1, a1n);
            load_diag(J + step, dgn);
            if (BWD) load_x(J + step, xinn);
            d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = cr;
            {
                const zd *B = ring[((J - step) % R + R) % R];
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) cmma(cr, ci, a1[ks], B[64 * ks + lane]);
            }
            // right-hand-side entries of block J (forward): a window of 64 entries sits in
            // registers (lane i holds entry wb + i, wc of them used), each entry goes to
            // the lane that holds element (r, c) = (4u + rq, jc)
            zd T[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) T[u] = BWD ? xin[u] : zd{};
            if (!BWD) {
                const int rend = (J + 1) * T16;
                for (;;) {
                    const bool in = lane >= wc && wrow < rend;
                    const int n = __popcll(__ballot(in));
                    for (int i = wc; i < wc + n; ++i) {
                        const int ri = __shfl(wrow, i, 64) - J * T16, ci_ = __shfl(wcol, i, 64);
                        const double vr = __shfl(wv.re, i, 64), vi = __shfl(wv.im, i, 64);
                        const bool mine = jc == ci_ && rq == (ri & 3);
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (mine && (ri >> 2) == u) {
                                T[u].re += vr;
                                T[u].im += vi;
                            }
                    }
                    wc += n;
                    if (wc < 64) break;
                    wb += 64;
                    wc = 0;
                    load_window();
                }
            }
            lds_barrier();  // the worker waves' partial sums of block J
            const int pb = J & 1;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int o = 64 * u + lane;
                T[u].re -= cr[u] + part[pb][0][o].re + part[pb][1][o].re + part[pb][2][o].re;
                T[u].im -= ci[u] + part[pb][0][o].im + part[pb][1][o].im + part[pb][2][o].im;
            }
            // X_J = Linv T (forward) or Linv^T T (backward); T in the result layout is
            // the B operand of step ks = element ks
            d4 xr = {0.0, 0.0, 0.0, 0.0}, xi = xr;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) cmma(xr, xi, dg[ks], T[ks]);
            zd *slot = ring[(J % R + R) % R];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const zd xv = zd{xr[u], xi[u]};
                slot[64 * u + lane] = xv;
                stz(X + (size_t)jc * A.ldx + (size_t)J * T16 + 4 * u + rq, xv);
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                a1[ks] = a1n[ks];
                dg[ks] = dgn[ks];
                if (BWD) xin[ks] = xinn[ks];
            }
        }
        return;
    }

    // worker waves 1..3: a = wave + 1 + 3u
    zd cur[NA][4], nxt[NA][4];
    const int Js = BWD ? nblk - 1 : J0;
    const int step = BWD ? -1 : 1;
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = wave + 1 + 3 * u;
        if (a <= D) load_off(Js, a, cur[u]);
        else
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) cur[u][ks] = zd{};
    }
    for (int J = Js; BWD ? J >= 0 : J < nblk; J += step) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = wave + 1 + 3 * u;
            if (a <= D) load_off(J + step, a, nxt[u]);
        }
        d4 cr = {0.0, 0.0, 0.0, 0.0}, ci = cr;
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = wave + 1 + 3 * u;
            if (a <= D) {
                const zd *B = ring[((J - step * a) % R + R) % R];
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) cmma(cr, ci, cur[u][ks], B[64 * ks + lane]);
            }
        }
        zd *pw = part[J & 1][wave - 1];
#pragma unroll
        for (int u = 0; u < 4; ++u) pw[64 * u + lane] = zd{cr[u], ci[u]};
        lds_barrier();
#pragma unroll
        for (int u = 0; u < NA; ++u)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) cur[u][ks] = nxt[u][ks];
    }
}

// ------------------------------------------------------ Schur complement
// S0[kk](cg, c) = A_GG^area(cg, c) - sum_i A_GI(cg, i) Z(i, c)   (lower tiles)
// G[kk](cg, l)  = b_G(cg, l)       - sum_i A_GI(cg, i) U(i, l)   (column nGp + l of X)
// scaled; padding rows and columns of S0 are the identity.  S0 is written in the
// dense kernel's layout: column-major 16 x 16 tiles packed by tile rows, tile (I, J),
// I >= J, at I(I+1)/2 + J.  One thread per entry.
struct SchurArgs {
    int kc, k0, nG, nGp, E, ldx, ntile, nnzF, fstrideI;
    const int *gi_ptr, *gi_row, *gi_q;  // A_GI by boundary row
    const int *gmap;                   // [nGp][nGp] full-grid entry of (cg, c), -1 none
    const int *eG;                     // [E] boundary index of the electrode node, -1 interior
    const zd *area;                    // [K][nnzF]
    const double *fI, *fG;
    const zd *X;
    zd *S0, *Gr;
};

__global__ void __launch_bounds__(256) crt_cond_schur_kernel(SchurArgs A) {
    const int nt = A.nGp / T16;
    const long ns = (long)nt * (nt + 1) / 2 * 256, ng = (long)A.E * A.nGp, per = ns + ng;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)A.kc * per) return;
    const int kk = (int)(t / per);
    const long o = t - (long)kk * per;
    const int k = A.k0 + kk;
    const bool zpart = o < ns;
    int cg, col;
    if (zpart) {
        const int tile = (int)(o / 256), e = (int)(o % 256);
        int I = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
        while ((I + 1) * (I + 2) / 2 <= tile) ++I;
        while (I * (I + 1) / 2 > tile) --I;
        const int J = tile - I * (I + 1) / 2;
        cg = 16 * I + e % 16;
        col = 16 * J + e / 16;
    } else {
        const long og = o - ns;
        cg = (int)(og % A.nGp);
        col = A.nGp + (int)(og / A.nGp);  // X column nGp + l
    }
    zd s = zd{0.0, 0.0};
    if (cg < A.nG) {
        const double fg = A.fG[(size_t)kk * A.nGp + cg];
        const zd *Xc = A.X + ((size_t)kk * A.ntile * T16 + col) * A.ldx;
        for (int p = A.gi_ptr[cg]; p < A.gi_ptr[cg + 1]; ++p) {
            const int i = A.gi_row[p];
            const zd a = A.area[(size_t)k * A.nnzF + A.gi_q[p]];
            const double f = fg * A.fI[(size_t)kk * A.fstrideI + i];
            const zd x = Xc[i];
            s.re -= f * (a.re * x.re - a.im * x.im);
            s.im -= f * (a.re * x.im + a.im * x.re);
        }
        if (zpart) {
            const int q = col < A.nG ? A.gmap[(size_t)col * A.nGp + cg] : -1;
            if (q >= 0) {
                const zd a = A.area[(size_t)k * A.nnzF + q];
                const double f = fg * A.fG[(size_t)kk * A.nGp + col];
                s.re += a.re * f;
                s.im += a.im * f;
            }
        } else if (A.eG[col - A.nGp] == cg) {
            s.re += fg;
        }
    } else if (zpart && col == cg) {
        s = zd{1.0, 0.0};
    }
    if (zpart) A.S0[(size_t)kk * ns + o] = s;
    else A.Gr[((size_t)kk * A.E + (col - A.nGp)) * A.nGp + cg] = s;
}

// scaled boundary terms of every system: bval[s][p] = (type-11 part of entry bq[p] for
// source l at wavenumber k) fG(g1[p]) fG(g2[p]), s = kk E + l
__global__ void __launch_bounds__(256) crt_cond_bval_kernel(DevGrid g, int nsys, int E, int k0, int nGp, int nb,
                                                            const int *__restrict__ bq, const int2 *__restrict__ bg,
                                                            const double *__restrict__ fG, zd *__restrict__ bval) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nsys * nb) return;
    const int s = (int)(t / nb), p = (int)(t - (long)s * nb);
    const int l = s % E, kk = s / E;
    const zd bp = boundary_part(g, bq[p], l, k0 + kk);
    const int2 gg = bg[p];
    const double f = fG[(size_t)kk * nGp + gg.x] * fG[(size_t)kk * nGp + gg.y];
    bval[t] = zd{bp.re * f, bp.im * f};
}

// --------------------------------------------------- dense factor and solve
// One workgroup (eight waves) per system (l, kk): S(l) = S0 + the boundary terms of source l
// (scaled: S0's tile plus the tile's list of entries that hold boundary terms),
// factorised S = L L^T without conjugation by tile columns (left-looking:
// the tiles of column J are accumulated in registers over the row J tiles staged in
// LDS, wave 0 factors the diagonal tile, each wave divides its tiles by L_JJ^T), then
// L L^T x = g by substitution.  Tiles are column-major 16 x 16, packed by tile rows:
// tile (I, J), I >= J, at I(I+1)/2 + J.  A zero diagonal of S(l) before scaling is
// error 27 (src/scalab.f90:54-58), a zero pivot 28 (src/chol.f90:66-70); with dc a
// diagonal or pivot that is not real and positive is (src/scaldc.f90, src/choldc.f90).
struct DenseArgs {
    DevGrid g;  // full grid: boundary contributions, area sums
    int nG, nGp, nt, E, K, k0;
    const zd *S0, *Gr;
    const int *bptr, *bent, *qdiag, *nodeG;  // boundary entries of each tile: element c*16 + r
    const zd *bval;                          // [s][nb] their values for each system
    int nb;
    const double *fG;
    zd *Lt, *Li, *XG, *kpotN;
    int *status;  // [kk][E]
};

namespace {
constexpr int DCH = 8;  // row-J tiles staged per round
constexpr int DNW = 8;  // waves per system

struct alignas(16) DenseLds {
    zd rowJ[DCH][256];
    zd dt[256], li[256];
    zd ws[DNW][256];
    zd pt[DNW][T16];
    zd pinv[T16];
    int err;
};
}  // namespace

template <int NI>
__global__ void __launch_bounds__(64 * DNW, 1) crt_cond_dense_kernel(DenseArgs A) {
    constexpr int DQ = NI >= 3 ? 4 : 8;  // tiles per batch of loads
    __shared__ DenseLds m;
    extern __shared__ zd yv[];  // nGp
    const int s = blockIdx.x;
    const int l = s % A.E, kk = s / A.E;
    const int k = A.k0 + kk;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int nt = A.nt, nGp = A.nGp, nG = A.nG;
    const size_t ntri = (size_t)nt * (nt + 1) / 2;
    zd *Ls = A.Lt + (size_t)s * ntri * 256;
    zd *Lis = A.Li + (size_t)s * nt * 256;
    const zd *S0 = A.S0 + (size_t)kk * ntri * 256;
    const zd *bval = A.bval + (size_t)s * A.nb;
    const double *fG = A.fG + (size_t)kk * nGp;
    const DevGrid &g = A.g;
    if (threadIdx.x == 0) m.err = 0;
    __syncthreads();
    auto tri = [](int I, int J) { return (size_t)I * (I + 1) / 2 + J; };
    // scalab's check of every diagonal before the factorisation
    for (int cg = threadIdx.x; cg < nG; cg += 64 * DNW) {
        const int q = A.qdiag[cg];
        const zd bp = boundary_part(g, q, l, k);
        const zd a = g.area[(size_t)k * g.nnz + q];
        const zd full = zd{a.re + bp.re, a.im + bp.im};
        if (!(zabs(full) > 0.0) || (g.dc && !(full.re > 0.0))) m.err = CRT_ERR_ZERODIAG;
    }
    __syncthreads();
    if (m.err) {
        if (threadIdx.x == 0) A.status[s] = m.err;
        return;
    }

    for (int J = 0; J < nt; ++J) {
        d4 ar[NI], ai[NI];
        // S(l) entries of the owned tiles I = J + wave + DNW u, column J
#pragma unroll
        for (int u = 0; u < NI; ++u) {
            const int I = J + wave + DNW * u;
            const zd *st = S0 + tri(I < nt ? I : J, J) * 256;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const zd e = I < nt ? ldz(st + jc * T16 + 4 * v + rq) : zd{};
                ar[u][v] = e.re;
                ai[u][v] = e.im;
            }
            if (I < nt) {
                // the tile's boundary terms of source l: lane i reads entry p0 + i, then
                // each goes to the lane that holds its element (r, c) = (4v + rq, jc)
                const int ti = (int)tri(I, J);
                const int pe = A.bptr[ti + 1];
                for (int p0 = A.bptr[ti]; p0 < pe; p0 += 64) {
                    const int p = p0 + lane, n = min(64, pe - p0);
                    const int ev = p < pe ? A.bent[p] : 0;
                    const zd bv = p < pe ? bval[p] : zd{};
                    for (int i = 0; i < n; ++i) {
                        const int e = __shfl(ev, i, 64), r = e % T16, c = e / T16;
                        const double vr = __shfl(bv.re, i, 64), vi = __shfl(bv.im, i, 64);
                        const bool mine = lane == c + 16 * (r % 4);
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (mine && v == r / 4) {
                                ar[u][v] += vr;
                                ai[u][v] += vi;
                            }
                    }
                }
            }
        }
        for (int p0 = 0; p0 < J; p0 += DCH) {
            const int np = min(DCH, J - p0);
            __syncthreads();
            const zd *src = Ls + tri(J, p0) * 256;
            for (int i = threadIdx.x; i < np * 256; i += 64 * DNW) (&m.rowJ[0][0])[i] = ldz(src + i);
            __syncthreads();
            // per owned tile: the chunk's tiles L(I, p0..) (contiguous) loaded DQ at a time,
            // all in flight together, then the products
#pragma unroll
            for (int u = 0; u < NI; ++u) {
                const int I = J + wave + DNW * u;
                if (I < nt) {
                    const zd *tp = Ls + tri(I, p0) * 256;
                    for (int pq = 0; pq < np; pq += DQ) {
                        zd ao[DQ][4];
#pragma unroll
                        for (int j = 0; j < DQ; ++j)
#pragma unroll
                            for (int ks = 0; ks < 4; ++ks)
                                ao[j][ks] = pq + j < np ? ldz(tp + (size_t)(pq + j) * 256 + 64 * ks + lane) : zd{};
#pragma unroll
                        for (int j = 0; j < DQ; ++j)
                            if (pq + j < np)
#pragma unroll
                                for (int ks = 0; ks < 4; ++ks)
                                    cmma(ar[u], ai[u], zd{-ao[j][ks].re, -ao[j][ks].im}, m.rowJ[pq + j][64 * ks + lane]);
                    }
                }
            }
        }
        // diagonal tile: wave 0 factors it in LDS
        if (wave == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) m.dt[jc * T16 + 4 * v + rq] = zd{ar[0][v], ai[0][v]};
            __builtin_amdgcn_wave_barrier();
            int err = 0;
#pragma unroll
            for (int c = 0; c < T16; ++c) {
                const zd sv = m.dt[c * T16 + c];
                const zd L = zsqrt(sv);
                if (!(zabs(sv) > 0.0) || (g.dc && !(sv.re > 0.0))) err = CRT_ERR_PIVOT;
                const zd inv = zinv(L);
                // lane (r, cq) owns entries (r, cq + 4w), w = 0..3
                const int r = jc;
                const zd lr = zmul(m.dt[c * T16 + r], inv);
                zd lc[4];
#pragma unroll
                for (int w = 0; w < 4; ++w) lc[w] = zmul(m.dt[c * T16 + rq + 4 * w], inv);
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const int cc = rq + 4 * w;
                    zd *e = &m.dt[cc * T16 + r];
                    if (cc > c && r >= cc) *e = zfms(*e, lr, lc[w]);
                    else if (cc == c && r > c) *e = lr;
                    else if (cc == c && r == c) *e = L;
                }
                if (lane == 0) m.pinv[c] = inv;
                __builtin_amdgcn_wave_barrier();
            }
            if (err && lane == 0) m.err = m.err ? m.err : err;
            // inverse of the lower triangle by rows, X(r, j) = (e_r - sum_{c<r} L(r,c) X(c, j)) / L(r,r):
            // lane (j, q) sums the terms c = q (mod 4)
            for (int r = 0; r < T16; ++r) {
                zd sacc = zd{};
                for (int c = rq; c < r; c += 4) sacc = zfma(sacc, m.dt[c * T16 + r], m.li[jc * T16 + c]);
                sacc.re += __shfl_xor(sacc.re, 16, 64);
                sacc.im += __shfl_xor(sacc.im, 16, 64);
                sacc.re += __shfl_xor(sacc.re, 32, 64);
                sacc.im += __shfl_xor(sacc.im, 32, 64);
                const zd xv = zmul(zd{(r == jc ? 1.0 : 0.0) - sacc.re, -sacc.im}, m.pinv[r]);
                if (rq == 0) m.li[jc * T16 + r] = xv;
                __builtin_amdgcn_wave_barrier();
            }
            zd *dst = Ls + tri(J, J) * 256;
            zd *ldst = Lis + (size_t)J * 256;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = 64 * u + lane, c = e / T16, r = e % T16;
                stz(dst + e, r >= c ? m.dt[e] : zd{});
                stz(ldst + e, m.li[e]);
            }
        }
        __syncthreads();
        // panel tiles L_IJ = S_IJ L_JJ^-T of the owned I > J
#pragma unroll
        for (int u = 0; u < NI; ++u) {
            const int I = J + wave + DNW * u;
            if (I > J && I < nt) {
                zd *w = m.ws[wave];
#pragma unroll
                for (int v = 0; v < 4; ++v) w[jc * T16 + 4 * v + rq] = zd{ar[u][v], ai[u][v]};
                __builtin_amdgcn_wave_barrier();
                d4 xr = {0.0, 0.0, 0.0, 0.0}, xi = xr;
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) cmma(xr, xi, w[64 * ks + lane], m.li[64 * ks + lane]);
                __builtin_amdgcn_wave_barrier();
                zd *dst = Ls + tri(I, J) * 256;
#pragma unroll
                for (int v = 0; v < 4; ++v) stz(dst + jc * T16 + 4 * v + rq, zd{xr[v], xi[v]});
            }
        }
    }
    __syncthreads();
    if (m.err) {
        if (threadIdx.x == 0) A.status[s] = m.err;
        return;
    }
    // ---- L L^T x = g: forward over tile rows, backward over tile columns
    const zd *gr = A.Gr + ((size_t)kk * A.E + l) * nGp;
    for (int i = threadIdx.x; i < nGp; i += 64 * DNW) yv[i] = gr[i];
    __syncthreads();
    for (int J = 0; J < nt; ++J) {
        // lane (r, q) of step w holds L(r, 4w + q)
        zd acc = zd{};
        for (int p = wave; p < J; p += DNW) {
            const zd *tp = Ls + tri(J, p) * 256;
#pragma unroll
            for (int w = 0; w < 4; ++w) acc = zfma(acc, ldz(tp + 64 * w + lane), yv[16 * p + 4 * w + rq]);
        }
        acc.re += __shfl_xor(acc.re, 16, 64);
        acc.im += __shfl_xor(acc.im, 16, 64);
        acc.re += __shfl_xor(acc.re, 32, 64);
        acc.im += __shfl_xor(acc.im, 32, 64);
        if (lane < T16) m.pt[wave][lane] = acc;
        __syncthreads();
        if (wave == 0) {
            if (lane < T16) {
                const int r = lane;
                zd tv = yv[16 * J + r];
#pragma unroll
                for (int w = 0; w < DNW; ++w) tv = zsub(tv, m.pt[w][r]);
                m.ws[0][r] = tv;
            }
            __builtin_amdgcn_wave_barrier();
            const zd *lt = Lis + (size_t)J * 256;
            zd y = zd{};
#pragma unroll
            for (int w = 0; w < 4; ++w) y = zfma(y, ldz(lt + 64 * w + lane), m.ws[0][4 * w + rq]);
            y.re += __shfl_xor(y.re, 16, 64);
            y.im += __shfl_xor(y.im, 16, 64);
            y.re += __shfl_xor(y.re, 32, 64);
            y.im += __shfl_xor(y.im, 32, 64);
            if (lane < T16) yv[16 * J + lane] = y;
        }
        __syncthreads();
    }
    for (int J = nt - 1; J >= 0; --J) {
        // out(c) = sum_r L(I,J)(r, c) x(16I + r): lane (r, q) of step w holds L(r, 4w + q)
        zd acc[4] = {zd{}, zd{}, zd{}, zd{}};
        for (int I = J + 1 + wave; I < nt; I += DNW) {
            const zd *tp = Ls + tri(I, J) * 256;
            const zd xr = yv[16 * I + jc];
#pragma unroll
            for (int w = 0; w < 4; ++w) acc[w] = zfma(acc[w], ldz(tp + 64 * w + lane), xr);
        }
#pragma unroll
        for (int w = 0; w < 4; ++w)
#pragma unroll
            for (int o = 1; o < T16; o <<= 1) {
                acc[w].re += __shfl_xor(acc[w].re, o, 64);
                acc[w].im += __shfl_xor(acc[w].im, o, 64);
            }
        if (jc == 0)
#pragma unroll
            for (int w = 0; w < 4; ++w) m.pt[wave][4 * w + rq] = acc[w];
        __syncthreads();
        if (wave == 0) {
            if (lane < T16) {
                zd tv = yv[16 * J + lane];
#pragma unroll
                for (int w = 0; w < DNW; ++w) tv = zsub(tv, m.pt[w][lane]);
                m.ws[0][lane] = tv;
            }
            __builtin_amdgcn_wave_barrier();
            // x_J(c) = sum_r Linv(r, c) t(r)
            const zd *lt = Lis + (size_t)J * 256;
            zd xo[4];
            const zd tr = m.ws[0][jc];
#pragma unroll
            for (int w = 0; w < 4; ++w) xo[w] = zmul(ldz(lt + 64 * w + lane), tr);
#pragma unroll
            for (int w = 0; w < 4; ++w)
#pragma unroll
                for (int o = 1; o < T16; o <<= 1) {
                    xo[w].re += __shfl_xor(xo[w].re, o, 64);
                    xo[w].im += __shfl_xor(xo[w].im, o, 64);
                }
            __builtin_amdgcn_wave_barrier();
            if (jc == 0)
#pragma unroll
                for (int w = 0; w < 4; ++w) yv[16 * J + 4 * w + rq] = xo[w];
        }
        __syncthreads();
    }
    zd *xg = A.XG + ((size_t)kk * A.E + l) * nGp;
    for (int i = threadIdx.x; i < nGp; i += 64 * DNW) {
        const zd x = yv[i];
        xg[i] = x;
        if (i < nG) {
            const double f = fG[i];
            A.kpotN[((size_t)A.nodeG[i] * A.K + k) * A.E + l] = zd{x.re * f, x.im * f};
        }
    }
}

// ------------------------------------------------ interior potentials
// kpot(node(i), k, l) = fI(i) (U(i, l) - sum_c Z(i, c) x_G(c, l)): each wave one
// 16-row tile times four 16-column tiles of x_G, K-loop over the boundary tiles.
struct GemmArgs {
    int kc, k0, nI, ldx, ntile, nGp, E, K, fstrideI, rgroups, cgroups;
    const zd *X, *XG;
    const double *fI;
    const int *nodeI;
    zd *kpotN;
};

__global__ void __launch_bounds__(256) crt_cond_gemm_kernel(GemmArgs A) {
    const int b = blockIdx.x;
    const int rg = b % A.rgroups, cgp = (b / A.rgroups) % A.cgroups, kk = b / (A.rgroups * A.cgroups);
    if (kk >= A.kc) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int rt = 4 * rg + wave;  // row tile
    if (rt * T16 >= A.ldx) return;
    const int k = A.k0 + kk;
    const zd *X = A.X + (size_t)kk * A.ntile * T16 * A.ldx;
    const zd *XG = A.XG + (size_t)kk * A.E * A.nGp;
    d4 cr[4], ci[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) cr[c] = ci[c] = d4{0.0, 0.0, 0.0, 0.0};
    const int nkt = A.nGp / T16;
    for (int kt = 0; kt < nkt; ++kt) {
        zd ao[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) ao[ks] = ldz(X + (size_t)(16 * kt + 4 * ks + rq) * A.ldx + 16 * rt + jc);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int lcol = 16 * (4 * cgp + c) + jc;
            if (16 * (4 * cgp + c) < A.E) {
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const zd bo = lcol < A.E ? ldz(XG + (size_t)lcol * A.nGp + 16 * kt + 4 * ks + rq) : zd{};
                    cmma(cr[c], ci[c], ao[ks], bo);
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int lcol = 16 * (4 * cgp + c) + jc;
        if (lcol >= A.E) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * rt + 4 * v + rq;
            if (i >= A.nI) continue;
            const zd u = X[(size_t)(A.nGp + lcol) * A.ldx + i];
            const double f = A.fI[(size_t)kk * A.fstrideI + i];
            A.kpotN[((size_t)A.nodeI[i] * A.K + k) * A.E + lcol] = zd{(u.re - cr[c][v]) * f, (u.im - ci[c][v]) * f};
        }
    }
}

// ------------------------------------------------------------------ launchers
int crt_launch_cond_gather(hipStream_t st, int kc, int k0, int nnzF, int nnzI, const int *qmap, const zd *area,
                           zd *areaI) {
    const long n = (long)kc * nnzI;
    hipLaunchKernelGGL(crt_cond_gather_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, kc, k0, nnzF, nnzI,
                       qmap, area, areaI);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cond_fg(hipStream_t st, int kc, int k0, int nG, int nGp, int nnzF, const int *qdiag, const zd *area,
                       double *fG) {
    const int n = kc * nGp;
    hipLaunchKernelGGL(crt_cond_fg_kernel, dim3((n + 255) / 256), dim3(256), 0, st, kc, k0, nG, nGp, nnzF, qdiag, area,
                       fG);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cond_rhs(hipStream_t st, int kc, int k0, int nent, int nnzF, int nGp, int fstrideI, const int *erow,
                        const int *ecolg, const int *eq, const zd *area, const double *fI, const double *fG,
                        zd *rv) {
    const long n = (long)kc * nent;
    if (n == 0) return 0;
    hipLaunchKernelGGL(crt_cond_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, kc, k0, nent, nnzF,
                       nGp, fstrideI, erow, ecolg, eq, area, fI, fG, rv);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

template <int D>
static int launch_trsm(hipStream_t st, TrsmArgs a, bool bwd) {
    a.nvirt = (a.ntile * a.kc + 7) / 8 * 8;
    if (bwd) hipLaunchKernelGGL((crt_cond_trsm_kernel<D, true>), dim3(a.nvirt), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((crt_cond_trsm_kernel<D, false>), dim3(a.nvirt), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cond_trsm(hipStream_t st, int D, int bwd, int nblk, int ldx, int ntile, int kc, const zd *panels,
                         const zd *dinv, zd *X, const int *tptr, const int *erow, const int *ecol, const zd *rv,
                         int nent, const int *j0) {
    TrsmArgs a{nblk, ldx, ntile, kc, 0, panels, dinv, X, tptr, erow, ecol, rv, nent, j0};
    switch (D) {
        case 2: return launch_trsm<2>(st, a, bwd);
        case 3: return launch_trsm<3>(st, a, bwd);
        case 4: return launch_trsm<4>(st, a, bwd);
        case 5: return launch_trsm<5>(st, a, bwd);
        case 6: return launch_trsm<6>(st, a, bwd);
        case 7: return launch_trsm<7>(st, a, bwd);
        case 8: return launch_trsm<8>(st, a, bwd);
        case 9: return launch_trsm<9>(st, a, bwd);
        case 10: return launch_trsm<10>(st, a, bwd);
        case 11: return launch_trsm<11>(st, a, bwd);
        case 12: return launch_trsm<12>(st, a, bwd);
        case 13: return launch_trsm<13>(st, a, bwd);
        case 14: return launch_trsm<14>(st, a, bwd);
        case 15: return launch_trsm<15>(st, a, bwd);
        case 16: return launch_trsm<16>(st, a, bwd);
        default: return CRT_ERR_UNSUPPORTED;
    }
}

int crt_launch_cond_schur(hipStream_t st, int kc, int k0, int nG, int nGp, int E, int ldx, int ntile, int nnzF,
                          int fstrideI, const int *gi_ptr, const int *gi_row, const int *gi_q, const int *gmap,
                          const int *eG, const zd *area, const double *fI, const double *fG, const zd *X, zd *S0,
                          zd *Gr) {
    SchurArgs a{kc, k0, nG, nGp, E, ldx, ntile, nnzF, fstrideI, gi_ptr, gi_row, gi_q, gmap, eG, area, fI, fG, X, S0, Gr};
    const long nt = nGp / T16;
    const long n = (long)kc * (nt * (nt + 1) / 2 * 256 + (long)E * nGp);
    hipLaunchKernelGGL(crt_cond_schur_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// largest boundary tile count the dense kernel takes (five owned tiles per wave)
int crt_cond_dense_max_tiles() { return 40; }

template <int NI>
static int launch_dense(hipStream_t st, const DenseArgs &a, int nsys) {
    const size_t lds = (size_t)a.nGp * sizeof(zd);
    hipLaunchKernelGGL(crt_cond_dense_kernel<NI>, dim3(nsys), dim3(64 * DNW), lds, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cond_dense(hipStream_t st, const DevGrid &g, int nG, int nGp, int E, int K, int k0, int nsys,
                          const zd *S0, const zd *Gr, const int *bptr, const int *bent, const zd *bval, int nb,
                          const int *qdiag, const int *nodeG, const double *fG, zd *Lt, zd *Li, zd *XG, zd *kpotN,
                          int *status) {
    const int nt = nGp / T16;
    if (nt < 1 || nt > crt_cond_dense_max_tiles() || nGp % T16) return CRT_ERR_ARG;
    DenseArgs a{g,  nG,    nGp,  nt, E,     K,  k0, S0, Gr, bptr, bent, qdiag, nodeG, bval,
                nb, fG,    Lt,   Li, XG,    kpotN, status};
    switch ((nt + DNW - 1) / DNW) {
        case 1: return launch_dense<1>(st, a, nsys);
        case 2: return launch_dense<2>(st, a, nsys);
        case 3: return launch_dense<3>(st, a, nsys);
        case 4: return launch_dense<4>(st, a, nsys);
        case 5: return launch_dense<5>(st, a, nsys);
        default: return CRT_ERR_UNSUPPORTED;
    }
}

int crt_launch_cond_bval(hipStream_t st, const DevGrid &g, int nsys, int E, int k0, int nGp, int nb, const int *bq,
                         const int2 *bg, const double *fG, zd *bval) {
    const long n = (long)nsys * nb;
    if (n == 0) return 0;
    hipLaunchKernelGGL(crt_cond_bval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nsys, E, k0, nGp,
                       nb, bq, bg, fG, bval);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cond_gemm(hipStream_t st, int kc, int k0, int nI, int ldx, int ntile, int nGp, int E, int K,
                         int fstrideI, const zd *X, const zd *XG, const double *fI, const int *nodeI, zd *kpotN) {
    const int rgroups = (ldx / T16 + 3) / 4, cgroups = ((E + 15) / 16 + 3) / 4;
    GemmArgs a{kc, k0, nI, ldx, ntile, nGp, E, K, fstrideI, rgroups, cgroups, X, XG, fI, nodeI, kpotN};
    const long n = (long)rgroups * cgroups * kc;
    hipLaunchKernelGGL(crt_cond_gemm_kernel, dim3((unsigned)n), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
