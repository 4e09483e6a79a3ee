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
// The interior band is cut into P pieces by windows of one band width, the
// separators S (no entry reaches across a window): A_II is block diagonal over the
// pieces, factorised side by side, and S joins the dense block B = [S, G], whose S
// part is eliminated once per wavenumber (the same for every source) before the
// per-source factorisation of the G part.
// This is block Gaussian elimination of the reference's system with the boundary
// nodes numbered last: the potentials equal those of the band Cholesky of the whole
// system (src/chol.f90, src/vre.f90) up to rounding.  Rows are scaled symmetrically
// by f = 1/sqrt|a_ii| (src/scalab.f90:52-76: on the interior rows the same numbers;
// on the boundary rows the area part of the diagonal, which is shared by the
// sources), so the kernels work on F A F and kpot = f x_hat.
//
// Tiles are 16 x 16 complex FP64 on v_mfma_f64_16x16x4f64; a complex product is
// three real ones (Gauss) in the substitution and the interior product, four in the
// dense factorisation.  Lanes: A[i][k] at lane i + 16(k%4) of step k/4, B[k][j] at lane
// j + 16(k%4), D[i][j] at lane j + 16(i%4), element i/4.
#include "crt_common.hpp"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int T16 = 16;

__device__ __forceinline__ d4 mfma4(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// RE: the real arithmetic of the DC inversion (src/choldc.f90, src/vredc.f90) on the
// same complex storage.  The DC model and every matrix built from it have zero
// imaginary parts, so the kernels read only the real parts, run one real MFMA chain
// per product and store +0 imaginary parts.

// (cr, ci) += a b for one k-step of complex operands
template <bool RE>
__device__ __forceinline__ void cmma(d4 &cr, d4 &ci, zd a, zd b) {
    cr = mfma4(a.re, b.re, cr);
    if constexpr (!RE) {
        cr = mfma4(-a.im, b.im, cr);
        ci = mfma4(a.re, b.im, ci);
        ci = mfma4(a.im, b.re, ci);
    }
}

// complex products with three real ones (Gauss): p1 = sum ar br, p2 = sum ai bi,
// p3 = sum (ar + ai)(br + bi); the product is (p1 - p2, p3 - p1 - p2).  The
// imaginary part loses |re|/|im| ulps to the subtraction (1e-13 relative for the
// milliradian phases of these potentials).
struct C3 {
    d4 p1, p2, p3;
};
__device__ __forceinline__ void c3zero(C3 &c) { c.p1 = c.p2 = c.p3 = d4{0.0, 0.0, 0.0, 0.0}; }
template <bool RE>
__device__ __forceinline__ void c3ma(C3 &c, zd a, zd b) {
    c.p1 = mfma4(a.re, b.re, c.p1);
    if constexpr (!RE) {
        c.p2 = mfma4(a.im, b.im, c.p2);
        c.p3 = mfma4(a.re + a.im, b.re + b.im, c.p3);
    }
}
template <bool RE>
__device__ __forceinline__ zd c3get(const C3 &c, int v) {
    if constexpr (RE) return zd{c.p1[v], 0.0};
    else return zd{c.p1[v] - c.p2[v], c.p3[v] - c.p1[v] - c.p2[v]};
}

__device__ __forceinline__ zd ldz(const zd *p) {
    const double2 v = *reinterpret_cast<const double2 *>(p);
    return zd{v.x, v.y};
}
// the real part alone under RE
template <bool RE>
__device__ __forceinline__ zd ldzr(const zd *p) {
    if constexpr (RE) return zd{p->re, 0.0};
    else return ldz(p);
}
template <bool RE>
__device__ __forceinline__ zd zmul_r(zd a, zd b) {
    if constexpr (RE) return zd{a.re * b.re, 0.0};
    else return zmul(a, b);
}
// a + b c and a - b c
template <bool RE>
__device__ __forceinline__ zd zfma_r(zd a, zd b, zd c) {
    if constexpr (RE) return zd{a.re + b.re * c.re, 0.0};
    else return zfma(a, b, c);
}
template <bool RE>
__device__ __forceinline__ zd zfms_r(zd a, zd b, zd c) {
    if constexpr (RE) return zd{a.re - b.re * c.re, 0.0};
    else return zfms(a, b, c);
}
// one butterfly step of a wave sum: a += the value of lane (lane ^ m)
template <bool RE>
__device__ __forceinline__ void zxor(zd &a, int m) {
    a.re += __shfl_xor(a.re, m, 64);
    if constexpr (!RE) a.im += __shfl_xor(a.im, m, 64);
}
// pivot of a DC column: L = sqrt(s), 1/L; bad for a pivot that is not positive and
// finite (src/choldc.f90 error 28)
__device__ __forceinline__ void pivot_real(double s, zd &L, zd &inv, bool &bad) {
    bad = !(s > 0.0) || !(s < 1.0e300);
    const double rh = rsqrt_f64(bad ? 1.0 : s);
    L = zd{s * rh, 0.0};
    inv = zd{rh, 0.0};
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

// fB[kk][b] = 1/sqrt|area part of a_bb| for the nodes of B (1 where it vanishes and
// for padding).  A separator node's diagonal is the whole diagonal: zero (dc: not
// real and positive) is error 27 of the wavenumber's first system (src/scalab.f90:54-58).
__global__ void __launch_bounds__(256) crt_cond_fb_kernel(int kc, int k0, int nS, int nBp, int nnzF, int dc, int nP,
                                                          const int *__restrict__ qdiag,
                                                          const zd *__restrict__ area, double *__restrict__ fB,
                                                          int *__restrict__ status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kc * nBp) return;
    const int kk = t / nBp, b = t - kk * nBp;
    double f = 1.0;
    const int q = qdiag[b];
    if (q >= 0) {
        const zd a = area[(size_t)(k0 + kk) * nnzF + q];
        const double m = zabs(a);
        if (m > 0.0) f = 1.0 / sqrt(m);
        if (b < nS && (!(m > 0.0) || (dc && !(a.re > 0.0)))) status[(size_t)kk * nP] = CRT_ERR_ZERODIAG;
    }
    fB[t] = f;
}

// values of the right-hand-side entries of each piece's [A_pB | b_p] (scaled): an A_pB
// entry area(q) fI(row) fB(b), a source entry fI(row).  fI of interior row g at
// wavenumber kk: fak of the chunk's system kk nP (every piece of a wavenumber scales
// the interior rows alike).
__global__ void __launch_bounds__(256) crt_cond_rhs_kernel(int kc, int k0, int nent, int nnzF, int nBp, int fstrideC,
                                                           int nP, const int *__restrict__ erowg,
                                                           const int *__restrict__ ebcol, const int *__restrict__ eq,
                                                           const zd *__restrict__ area, const double *__restrict__ fI,
                                                           const double *__restrict__ fB, zd *__restrict__ rv) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)kc * nent) return;
    const int kk = (int)(t / nent), e = (int)(t - (long)kk * nent);
    const double fi = fI[(size_t)kk * nP * fstrideC + erowg[e]];
    const int q = eq[e];
    zd v;
    if (q >= 0) {
        const zd a = area[(size_t)(k0 + kk) * nnzF + q];
        const double f = fi * fB[(size_t)kk * nBp + ebcol[e]];
        v = zd{a.re * f, a.im * f};
    } else {
        v = zd{fi, 0.0};
    }
    rv[t] = v;
}

// ------------------------------------------------- multi-RHS band substitution
// X <- L^-1 X (forward) or X <- L^-T X (backward) for the factor of each interior
// piece of each wavenumber of the chunk (system kk nP + p), 16 columns per workgroup.  Factor layout of the band
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
    int nblk, ldx, ntile, nsys, nP, nvirt;
    const zd *panels;   // [s][nblk][D+1][256]
    const zd *dinv;     // [s][nblk*16]
    zd *X;              // [s][ntile*16][ldx]
    const int *tptr;    // nP*ntile+1: entries of each (piece, column tile)
    const int *erow;    // row (within the piece) of each entry
    const int *ecol;    // column within the tile
    const zd *rv;       // [kk][nent]
    int nent;
    const int *j0;      // first nonzero block of each (piece, column tile)
};

template <int D, bool BWD, bool RE>
__global__ void __launch_bounds__(256, D <= 7 ? 3 : (D <= 14 ? 2 : 1)) crt_cond_trsm_kernel(TrsmArgs A) {
    constexpr int R = D;                              // ring slots
    constexpr int NA = D >= 2 ? (D - 1 + 2) / 3 : 1;  // tiles per worker wave (a = 2 + w-1 + 3u)
    __shared__ zd ring[R][256];
    __shared__ zd part[2][3][256];
    // XCD-aware order: consecutive virtual ids (tiles of one wavenumber) on one XCD
    const int b = blockIdx.x;
    const int per = A.nvirt / 8;
    const int v = (b % 8) * per + b / 8;
    const int sy = v / A.ntile, t = v - sy * A.ntile;
    if (sy >= A.nsys) return;
    const int kk = sy / A.nP, pt = (sy - kk * A.nP) * A.ntile + t;  // (piece, tile) of the entry lists
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int nblk = A.nblk;
    const zd *panel = A.panels + (size_t)sy * nblk * (D + 1) * 256;
    const zd *dinv = A.dinv + (size_t)sy * nblk * T16;
    zd *X = A.X + ((size_t)sy * A.ntile + t) * T16 * A.ldx;
    const int J0 = A.j0[pt];
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
            o[ks] = live ? (BWD ? ldzr<RE>(tp + 64 * ks + lane) : ldzr<RE>(tp + jc * T16 + 4 * ks + rq)) : zd{};
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
                if (jc == k) e = ldzr<RE>(dinv + (size_t)J * T16 + jc);
                else if (BWD ? (k > jc) : (jc > k)) e = BWD ? ldzr<RE>(tp + jc * T16 + k) : ldzr<RE>(tp + 64 * ks + lane);
            }
            o[ks] = e;
        }
    };

    if (wave == 0) {
        zd a1[4], a1n[4], dg[4], dgn[4];
        zd xin[4] = {zd{}, zd{}, zd{}, zd{}}, xinn[4];
        int wb = BWD ? 0 : A.tptr[pt], wc = 0;
        const int cend = BWD ? 0 : A.tptr[pt + 1];
        const zd *rv = BWD ? nullptr : A.rv + (size_t)kk * A.nent;
        int wrow = INT32_MAX, wcol = 0;
        zd wv = zd{};
        auto load_window = [&]() {
            const int p = wb + lane;
            wrow = p < cend ? A.erow[p] : INT32_MAX;
            wcol = p < cend ? A.ecol[p] : 0;
            wv = p < cend ? ldzr<RE>(rv + p) : zd{};
        };
        if (!BWD) load_window();
        auto load_x = [&](int J, zd (&o)[4]) {
#pragma unroll
            for (int u = 0; u < 4; ++u)
                o[u] = (J >= J0 && J >= 0 && J < nblk) ? ldzr<RE>(X + (size_t)jc * A.ldx + (size_t)J * T16 + 4 * u + rq)
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
            C3 acc;
            c3zero(acc);
            {
                const zd *B = ring[((J - step) % R + R) % R];
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) c3ma<RE>(acc, a1[ks], B[64 * ks + lane]);
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
                        const double vr = __shfl(wv.re, i, 64), vi = RE ? 0.0 : __shfl(wv.im, i, 64);
                        const bool mine = jc == ci_ && rq == (ri & 3);
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (mine && (ri >> 2) == u) {
                                T[u].re += vr;
                                if constexpr (!RE) T[u].im += vi;
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
                const zd a0 = c3get<RE>(acc, u);
                T[u].re -= a0.re + part[pb][0][o].re + part[pb][1][o].re + part[pb][2][o].re;
                if constexpr (!RE) T[u].im -= a0.im + part[pb][0][o].im + part[pb][1][o].im + part[pb][2][o].im;
            }
            // X_J = Linv T (forward) or Linv^T T (backward); T in the result layout is
            // the B operand of step ks = element ks
            C3 xa;
            c3zero(xa);
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) c3ma<RE>(xa, dg[ks], T[ks]);
            zd *slot = ring[(J % R + R) % R];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const zd xv = c3get<RE>(xa, u);
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
        C3 acc;
        c3zero(acc);
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = wave + 1 + 3 * u;
            if (a <= D) {
                const zd *B = ring[((J - step * a) % R + R) % R];
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) c3ma<RE>(acc, cur[u][ks], B[64 * ks + lane]);
            }
        }
        zd *pw = part[J & 1][wave - 1];
#pragma unroll
        for (int u = 0; u < 4; ++u) pw[64 * u + lane] = c3get<RE>(acc, u);
        lds_barrier();
#pragma unroll
        for (int u = 0; u < NA; ++u)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) cur[u][ks] = nxt[u][ks];
    }
}

// ------------------------------------------------------ Schur complement
// T[kk](b, c) = A_BB^area(b, c) - sum_i A_Bi(b, i) Z_p(i, c)   (lower tiles over B)
// G[kk](b, l) = b_B(b, l)      - sum_i A_Bi(b, i) U_p(i, l)
// scaled, the sums over the interior neighbours i of b in every piece p (Z_p(., c) is
// zero where c is not coupled to piece p, U_p(., l) where electrode l is not in it);
// padding rows and columns of T are the identity.  T is written in the dense
// kernels' layout: column-major 16 x 16 tiles packed by tile rows, tile (I, J),
// I >= J, at I(I+1)/2 + J.  One thread per entry.
struct SchurArgs {
    int kc, k0, nBp, E, ldx, ntile, nnzF, fstrideC, nP;
    const int *gi_ptr, *gi_row, *gi_piece, *gi_q;  // A_Bi by B row: combined interior row, piece, entry
    const int *roff;                              // first combined row of each piece
    const int *bcol, *ucol;                       // [p][nBp], [p][E]: column of b / of electrode l in piece p
    const int *gmap;                              // [nBp][nBp] full-grid entry of (b, c), -1 none
    const int *eB;                                // [E] B index of the electrode node, -1 interior
    const zd *area;                               // [K][nnzF]
    const double *fI, *fB;
    const zd *X;                                  // [kk nP + p][ntile*16][ldx]
    zd *T, *Gr;
};

__global__ void __launch_bounds__(256) crt_cond_schur_kernel(SchurArgs A) {
    const int ntb = A.nBp / T16;
    const long ns = (long)ntb * (ntb + 1) / 2 * 256, ng = (long)A.E * A.nBp, per = ns + ng;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)A.kc * per) return;
    const int kk = (int)(t / per);
    const long o = t - (long)kk * per;
    const int k = A.k0 + kk;
    const bool tpart = o < ns;
    int b, col, l = 0;
    if (tpart) {
        const int tile = (int)(o / 256), e = (int)(o % 256);
        int I = (int)((sqrt(8.0 * tile + 1.0) - 1.0) * 0.5);
        while ((I + 1) * (I + 2) / 2 <= tile) ++I;
        while (I * (I + 1) / 2 > tile) --I;
        const int J = tile - I * (I + 1) / 2;
        b = 16 * I + e % 16;
        col = 16 * J + e / 16;
    } else {
        const long og = o - ns;
        b = (int)(og % A.nBp);
        l = (int)(og / A.nBp);
        col = -1;
    }
    zd s = zd{0.0, 0.0};
    const double fb = A.fB[(size_t)kk * A.nBp + b];
    const int q0 = A.gmap[(size_t)b * A.nBp + b];  // >= 0 for a node of B, -1 for padding
    if (q0 >= 0) {
        for (int p = A.gi_ptr[b]; p < A.gi_ptr[b + 1]; ++p) {
            const int g = A.gi_row[p], pc = A.gi_piece[p];
            const int xc = tpart ? A.bcol[(size_t)pc * A.nBp + col] : A.ucol[(size_t)pc * A.E + l];
            if (xc < 0) continue;
            const zd a = A.area[(size_t)k * A.nnzF + A.gi_q[p]];
            const double f = fb * A.fI[(size_t)kk * A.nP * A.fstrideC + g];
            const zd x = A.X[(((size_t)kk * A.nP + pc) * A.ntile * T16 + xc) * A.ldx + (g - A.roff[pc])];
            s.re -= f * (a.re * x.re - a.im * x.im);
            s.im -= f * (a.re * x.im + a.im * x.re);
        }
        if (tpart) {
            const int q = A.gmap[(size_t)col * A.nBp + b];
            if (q >= 0) {
                const zd a = A.area[(size_t)k * A.nnzF + q];
                const double f = fb * A.fB[(size_t)kk * A.nBp + col];
                s.re += a.re * f;
                s.im += a.im * f;
            }
        } else if (A.eB[l] == b) {
            s.re += fb;
        }
    } else if (tpart && col == b) {
        s = zd{1.0, 0.0};
    }
    if (tpart) A.T[(size_t)kk * ns + o] = s;
    else A.Gr[((size_t)kk * A.E + l) * A.nBp + b] = s;
}

// scaled boundary terms of every system: bval[s][p] = (type-11 part of entry bq[p] for
// source l at wavenumber k) fG(g1[p]) fG(g2[p]), s = kk E + l, fG of wavenumber kk at
// fG + kk nGp
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
// One workgroup (eight waves) per system.  Source systems (l, kk): S(l) = S0 + the
// boundary terms of source l over the G block (scaled: S0's tile plus the tile's list
// of entries that hold boundary terms), factorised S = L L^T without conjugation by
// tile columns (left-looking: the tiles of column J are accumulated in registers over
// the row J tiles staged in LDS, wave 0 factors the diagonal tile, each wave divides
// its tiles by L_JJ^T); then L_B L_B^T x = g over all of B by substitution, the
// separator columns of L_B from the wavenumber's prefix.  Prefix systems (PRE, one per
// wavenumber): the same factorisation of T over B for the JS separator tile columns
// (L_SS, L_GS), and the G block's Schur complement S0 = T_GG - L_GS L_GS^T for the
// source systems.  Tiles are column-major 16 x 16, packed by tile rows: tile (I, J),
// I >= J, at I(I+1)/2 + J.  A zero diagonal of S(l) before scaling is error 27
// (src/scalab.f90:54-58), a zero pivot 28 (src/chol.f90:66-70); with dc a diagonal or
// pivot that is not real and positive is (src/scaldc.f90, src/choldc.f90).
struct DenseArgs {
    DevGrid g;  // full grid: boundary contributions, area sums
    int nG, nt, nBp, JS, E, K, k0;  // nt: G tiles, JS: separator tiles, nBp = 16 (JS + nt)
    const zd *S0, *T, *Gr;          // S0 [kk][G tiles], T [kk][B tiles], Gr [kk][E][nBp]
    const int *bptr, *bent, *qdiag, *nodeB;  // boundary entries per G tile (element c*16 + r); G diagonals
    const zd *bval;                          // [s][nb] values of the boundary entries
    int nb;
    const double *fB;                        // [kk][nBp]
    zd *Lt, *Li;                             // [s][G tiles], [s][nt]
    zd *Lpre, *Lipre;                        // [kk][B tiles] (separator columns), [kk][JS]
    zd *S0w;                                 // prefix output: S0 [kk][G tiles]
    zd *XB, *kpotN;                          // XB [kk][E][nBp]
    int *status;
    int sstride;                             // status slots per system
};

namespace {
constexpr int DCH = 8;  // row-J tiles staged per round
constexpr int DNW = 8;  // waves per system (prefix, and source systems of more than 20 tiles)

// rowJ (the staged row tiles of the column updates) and ws (each wave's panel tile, and
// wave 0's scratch in the substitution) are live in phases that barriers separate
template <int NW>
struct alignas(16) DenseLds {
    union {
        zd rowJ[DCH][256];
        zd ws[NW][256];
    };
    zd dt[256], li[256];
    zd lip[T16][T16 + 1];  // the inverse while it is formed, columns padded against bank conflicts
    zd pt[NW][T16];
    zd pinv[T16];
    int err;
};
}  // namespace

// NW waves per system.  With four (source systems of up to 20 tiles) two systems share
// a CU, so one system's diagonal chain and loads overlap the other's products.
template <int NI, bool PRE, int NW, bool RE>
__global__ void __launch_bounds__(64 * NW, NW == 4 ? 2 : 1) crt_cond_dense_kernel(DenseArgs A) {
    __shared__ DenseLds<NW> m;
    extern __shared__ zd yv[];  // nBp
    const int s = blockIdx.x;
    const int l = PRE ? 0 : s % A.E, kk = PRE ? s : s / A.E;
    const int k = A.k0 + kk;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int JS = A.JS, ntB = A.nBp / T16, nG = A.nG;
    // tile rows and columns of the matrix factorised here: the prefix factorises the
    // separator block T_SS alone (L_GS and S0 follow in crt_cond_lgs_kernel and
    // crt_cond_s0_kernel, row by row and tile by tile)
    const int nt = PRE ? JS : A.nt;
    const size_t ntri = (size_t)A.nt * (A.nt + 1) / 2, ntriB = (size_t)ntB * (ntB + 1) / 2;
    zd *Ls = PRE ? A.Lpre + (size_t)kk * ntriB * 256 : A.Lt + (size_t)s * ntri * 256;
    zd *Lis = PRE ? A.Lipre + (size_t)kk * JS * 256 : A.Li + (size_t)s * A.nt * 256;
    const zd *M = PRE ? A.T + (size_t)kk * ntriB * 256 : A.S0 + (size_t)kk * ntri * 256;
    const zd *bval = A.bval + (size_t)(PRE ? 0 : s) * A.nb;
    const double *fB = A.fB + (size_t)kk * A.nBp;
    const DevGrid &g = A.g;
    if (threadIdx.x == 0) m.err = 0;
    __syncthreads();
    auto tri = [](int I, int J) { return (size_t)I * (I + 1) / 2 + J; };
    if (!PRE) {
        // scalab's check of every boundary diagonal before the factorisation
        for (int cg = threadIdx.x; cg < nG; cg += 64 * NW) {
            const int q = A.qdiag[cg];
            const zd bp = boundary_part(g, q, l, k);
            const zd a = g.area[(size_t)k * g.nnz + q];
            const zd full = zd{a.re + bp.re, a.im + bp.im};
            if (!(zabs(full) > 0.0) || (g.dc && !(full.re > 0.0))) m.err = CRT_ERR_ZERODIAG;
        }
        __syncthreads();
        if (m.err) {
            if (threadIdx.x == 0) A.status[(size_t)s * A.sstride] = m.err;
            return;
        }
    }

    // the prefix factorises the separator columns only; S0 = T_GG - L_GS L_GS^T is formed
    // by crt_cond_s0_kernel over every tile at once
    for (int J = 0; J < (PRE ? JS : nt); ++J) {
        const int pend = J;  // columns of L in this column's update
        d4 ar[NI], ai[NI];
        // S(l) entries of the owned tiles I = J + wave + NW u, column J
#pragma unroll
        for (int u = 0; u < NI; ++u) {
            const int I = J + wave + NW * u;
            const zd *st = M + tri(I < nt ? I : J, J) * 256;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const zd e = I < nt ? ldzr<RE>(st + jc * T16 + 4 * v + rq) : zd{};
                ar[u][v] = e.re;
                ai[u][v] = e.im;
            }
            if (!PRE && I < nt) {
                // the tile's boundary terms of source l: lane i reads entry p0 + i, then
                // each goes to the lane that holds its element (r, c) = (4v + rq, jc)
                const int ti = (int)tri(I, J);
                const int pe = A.bptr[ti + 1];
                for (int p0 = A.bptr[ti]; p0 < pe; p0 += 64) {
                    const int p = p0 + lane, n = min(64, pe - p0);
                    const int ev = p < pe ? A.bent[p] : 0;
                    const zd bv = p < pe ? ldzr<RE>(bval + p) : zd{};
                    for (int i = 0; i < n; ++i) {
                        const int e = __shfl(ev, i, 64), r = e % T16, c = e / T16;
                        const double vr = __shfl(bv.re, i, 64), vi = RE ? 0.0 : __shfl(bv.im, i, 64);
                        const bool mine = lane == c + 16 * (r % 4);
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            if (mine && v == r / 4) {
                                ar[u][v] += vr;
                                if constexpr (!RE) ai[u][v] += vi;
                            }
                    }
                }
            }
        }
        for (int p0 = 0; p0 < pend; p0 += DCH) {
            const int np = min(DCH, pend - p0);
            __syncthreads();
            const zd *src = Ls + tri(J, p0) * 256;
            for (int i = threadIdx.x; i < np * 256; i += 64 * NW) {
                if constexpr (RE) (&m.rowJ[0][0])[i].re = src[i].re;
                else (&m.rowJ[0][0])[i] = ldz(src + i);
            }
            __syncthreads();
            // per owned tile: the chunk's tiles L(I, p0..) (contiguous), PW at a time, the
            // next group in flight while the current one is multiplied (the pipeline runs
            // on across the owned tiles)
            constexpr int PW = 2;
            const int nv = min(NI, max(0, (nt - (J + wave) + NW - 1) / NW));
            zd bc[PW][4], bn[PW][4];
            auto ld2 = [&](int u, int pq, zd (&o)[PW][4]) {
                const zd *tp = Ls + tri(J + wave + NW * u, p0) * 256;
                const bool live = u < nv;
#pragma unroll
                for (int j = 0; j < PW; ++j)
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks)
                        o[j][ks] = live && pq + j < np ? ldzr<RE>(tp + (size_t)(pq + j) * 256 + 64 * ks + lane) : zd{};
            };
            ld2(0, 0, bc);
#pragma unroll
            for (int u = 0; u < NI; ++u) {
                if (u < nv) {
                    for (int pq = 0; pq < np; pq += PW) {
                        const bool last = pq + PW >= np;
                        ld2(last ? u + 1 : u, last ? 0 : pq + PW, bn);
#pragma unroll
                        for (int j = 0; j < PW; ++j)
                            if (pq + j < np)
#pragma unroll
                                for (int ks = 0; ks < 4; ++ks)
                                    cmma<RE>(ar[u], ai[u], zd{-bc[j][ks].re, -bc[j][ks].im}, m.rowJ[pq + j][64 * ks + lane]);
#pragma unroll
                        for (int j = 0; j < PW; ++j)
                            for (int ks = 0; ks < 4; ++ks) bc[j][ks] = bn[j][ks];
                    }
                }
            }
        }
        // diagonal tile: wave 0 factors it in LDS
        if (wave == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) m.dt[jc * T16 + 4 * v + rq] = zd{ar[0][v], RE ? 0.0 : ai[0][v]};
            __builtin_amdgcn_wave_barrier();
            int err = 0;
#pragma unroll
            for (int c = 0; c < T16; ++c) {
                const zd sv = m.dt[c * T16 + c];
                zd L, inv;
                bool bad;
                if constexpr (RE) pivot_real(sv.re, L, inv, bad);
                else pivot_fast(sv, L, inv, bad);
                if (bad || (g.dc && !(sv.re > 0.0))) err = CRT_ERR_PIVOT;
                // lane (r, cq) owns entries (r, cq + 4w), w = 0..3
                const int r = jc;
                const zd lr = zmul_r<RE>(m.dt[c * T16 + r], inv);
                zd lc[4];
#pragma unroll
                for (int w = 0; w < 4; ++w) lc[w] = zmul_r<RE>(m.dt[c * T16 + rq + 4 * w], inv);
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const int cc = rq + 4 * w;
                    zd *e = &m.dt[cc * T16 + r];
                    if (cc > c && r >= cc) *e = zfms_r<RE>(*e, lr, lc[w]);
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
                for (int c = rq; c < r; c += 4) sacc = zfma_r<RE>(sacc, m.dt[c * T16 + r], m.lip[jc][c]);
                zxor<RE>(sacc, 16);
                zxor<RE>(sacc, 32);
                const zd xv = zmul_r<RE>(zd{(r == jc ? 1.0 : 0.0) - sacc.re, -sacc.im}, m.pinv[r]);
                if (rq == 0) m.lip[jc][r] = xv;
                __builtin_amdgcn_wave_barrier();
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = 64 * u + lane;
                m.li[e] = m.lip[e / T16][e % T16];
            }
            __builtin_amdgcn_wave_barrier();
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
            const int I = J + wave + NW * u;
            if (I > J && I < nt) {
                zd *w = m.ws[wave];
#pragma unroll
                for (int v = 0; v < 4; ++v) w[jc * T16 + 4 * v + rq] = zd{ar[u][v], RE ? 0.0 : ai[u][v]};
                __builtin_amdgcn_wave_barrier();
                d4 xr = {0.0, 0.0, 0.0, 0.0}, xi = xr;
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) cmma<RE>(xr, xi, w[64 * ks + lane], m.li[64 * ks + lane]);
                __builtin_amdgcn_wave_barrier();
                zd *dst = Ls + tri(I, J) * 256;
#pragma unroll
                for (int v = 0; v < 4; ++v) stz(dst + jc * T16 + 4 * v + rq, zd{xr[v], RE ? 0.0 : xi[v]});
            }
        }
    }
    __syncthreads();
    if (m.err) {
        if (threadIdx.x == 0) A.status[(size_t)s * A.sstride] = m.err;
        return;
    }
    if constexpr (PRE) return;
    // ---- L_B L_B^T x = g over B: forward over tile rows, backward over tile columns
    const zd *Lp = A.Lpre + (size_t)kk * ntriB * 256, *Lip = A.Lipre + (size_t)kk * JS * 256;
    auto tileL = [&](int I, int J) -> const zd * {
        return J < JS ? Lp + tri(I, J) * 256 : Ls + tri(I - JS, J - JS) * 256;
    };
    auto tileLi = [&](int J) -> const zd * { return J < JS ? Lip + (size_t)J * 256 : Lis + (size_t)(J - JS) * 256; };
    const int nBp = A.nBp;
    const zd *gr = A.Gr + ((size_t)kk * A.E + l) * nBp;
    for (int i = threadIdx.x; i < nBp; i += 64 * NW) yv[i] = gr[i];
    __syncthreads();
    for (int J = 0; J < ntB; ++J) {
        // lane (r, q) of step w holds L(r, 4w + q)
        zd acc = zd{};
        for (int p = wave; p < J; p += NW) {
            const zd *tp = tileL(J, p);
#pragma unroll
            for (int w = 0; w < 4; ++w) acc = zfma_r<RE>(acc, ldzr<RE>(tp + 64 * w + lane), yv[16 * p + 4 * w + rq]);
        }
        zxor<RE>(acc, 16);
        zxor<RE>(acc, 32);
        if (lane < T16) m.pt[wave][lane] = acc;
        __syncthreads();
        if (wave == 0) {
            if (lane < T16) {
                const int r = lane;
                zd tv = yv[16 * J + r];
#pragma unroll
                for (int w = 0; w < NW; ++w) tv = zsub(tv, m.pt[w][r]);
                m.ws[0][r] = tv;
            }
            __builtin_amdgcn_wave_barrier();
            const zd *lt = tileLi(J);
            zd y = zd{};
#pragma unroll
            for (int w = 0; w < 4; ++w) y = zfma_r<RE>(y, ldzr<RE>(lt + 64 * w + lane), m.ws[0][4 * w + rq]);
            zxor<RE>(y, 16);
            zxor<RE>(y, 32);
            if (lane < T16) yv[16 * J + lane] = y;
        }
        __syncthreads();
    }
    for (int J = ntB - 1; J >= 0; --J) {
        // out(c) = sum_r L(I,J)(r, c) x(16I + r): lane (r, q) of step w holds L(r, 4w + q)
        zd acc[4] = {zd{}, zd{}, zd{}, zd{}};
        for (int I = J + 1 + wave; I < ntB; I += NW) {
            const zd *tp = tileL(I, J);
            const zd xr = yv[16 * I + jc];
#pragma unroll
            for (int w = 0; w < 4; ++w) acc[w] = zfma_r<RE>(acc[w], ldzr<RE>(tp + 64 * w + lane), xr);
        }
#pragma unroll
        for (int w = 0; w < 4; ++w)
#pragma unroll
            for (int o = 1; o < T16; o <<= 1) zxor<RE>(acc[w], o);
        if (jc == 0)
#pragma unroll
            for (int w = 0; w < 4; ++w) m.pt[wave][4 * w + rq] = acc[w];
        __syncthreads();
        if (wave == 0) {
            if (lane < T16) {
                zd tv = yv[16 * J + lane];
#pragma unroll
                for (int w = 0; w < NW; ++w) tv = zsub(tv, m.pt[w][lane]);
                m.ws[0][lane] = tv;
            }
            __builtin_amdgcn_wave_barrier();
            // x_J(c) = sum_r Linv(r, c) t(r)
            const zd *lt = tileLi(J);
            zd xo[4];
            const zd tr = m.ws[0][jc];
#pragma unroll
            for (int w = 0; w < 4; ++w) xo[w] = zmul_r<RE>(ldzr<RE>(lt + 64 * w + lane), tr);
#pragma unroll
            for (int w = 0; w < 4; ++w)
#pragma unroll
                for (int o = 1; o < T16; o <<= 1) zxor<RE>(xo[w], o);
            __builtin_amdgcn_wave_barrier();
            if (jc == 0)
#pragma unroll
                for (int w = 0; w < 4; ++w) yv[16 * J + 4 * w + rq] = xo[w];
        }
        __syncthreads();
    }
    zd *xb = A.XB + ((size_t)kk * A.E + l) * nBp;
    for (int i = threadIdx.x; i < nBp; i += 64 * NW) {
        const zd x = yv[i];
        xb[i] = x;
        const int nd = A.nodeB[i];
        if (nd >= 0) {
            const double f = fB[i];
            A.kpotN[((size_t)nd * A.K + k) * A.E + l] = zd{x.re * f, x.im * f};
        }
    }
}

// ------------------------------------------------ L_GS of the prefix
// L(I, p) = (T(I, p) - sum_{q < p} L(I, q) L(p, q)^T) L(p, p)^-T for the tile rows I of
// the G block (JS <= I < ntB) and the separator columns p < JS, one wave per (wavenumber,
// row): the separator block's factor L(p, q) and the inverses of its diagonal tiles come
// from the prefix kernel; the row's own earlier tiles are read back from global memory
// after the wave's stores are complete.
struct LgsArgs {
    int kc, ntB, JS;
    const zd *T;          // [kk][B tiles]
    zd *Lpre;             // [kk][B tiles]
    const zd *Lipre;      // [kk][JS]
};

template <bool RE>
__global__ void __launch_bounds__(256) crt_cond_lgs_kernel(LgsArgs A) {
    __shared__ zd ws[4][256];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int nr = A.ntB - A.JS;
    const long w = (long)blockIdx.x * 4 + wave;
    if (w >= (long)A.kc * nr) return;
    const int kk = (int)(w / nr), I = A.JS + (int)(w - (long)kk * nr);
    const long ntriB = (long)A.ntB * (A.ntB + 1) / 2;
    auto tri = [](long a, long b) { return a * (a + 1) / 2 + b; };
    const zd *Tk = A.T + kk * ntriB * 256, *Lik = A.Lipre + (long)kk * A.JS * 256;
    zd *Lk = A.Lpre + kk * ntriB * 256;
    zd *wsm = ws[wave];
    for (int p = 0; p < A.JS; ++p) {
        d4 ar, ai;
        const zd *src = Tk + tri(I, p) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const zd e = ldzr<RE>(src + jc * T16 + 4 * v + rq);
            ar[v] = e.re;
            ai[v] = e.im;
        }
        for (int q = 0; q < p; ++q) {
            const zd *la = Lk + tri(I, q) * 256, *lb = Lk + tri(p, q) * 256;
            zd a4[4], b4[4];
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                a4[ks] = ldzr<RE>(la + 64 * ks + lane);
                b4[ks] = ldzr<RE>(lb + 64 * ks + lane);
            }
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) cmma<RE>(ar, ai, zd{-a4[ks].re, -a4[ks].im}, b4[ks]);
        }
        // X = C L(p, p)^-T with C staged column-major in the wave's LDS tile
#pragma unroll
        for (int v = 0; v < 4; ++v) wsm[jc * T16 + 4 * v + rq] = zd{ar[v], RE ? 0.0 : ai[v]};
        __builtin_amdgcn_wave_barrier();
        d4 xr = {0.0, 0.0, 0.0, 0.0}, xi = xr;
        const zd *li = Lik + (long)p * 256;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) cmma<RE>(xr, xi, wsm[64 * ks + lane], ldzr<RE>(li + 64 * ks + lane));
        __builtin_amdgcn_wave_barrier();
        zd *dst = Lk + tri(I, p) * 256;
#pragma unroll
        for (int v = 0; v < 4; ++v) stz(dst + jc * T16 + 4 * v + rq, zd{xr[v], RE ? 0.0 : xi[v]});
        __threadfence_block();  // the row's later columns read this tile back
    }
}

int crt_launch_cond_lgs(hipStream_t st, int kc, int ntB, int JS, const zd *T, zd *Lpre, const zd *Lipre, int re) {
    LgsArgs a{kc, ntB, JS, T, Lpre, Lipre};
    const long n = (long)kc * (ntB - JS);
    if (n <= 0) return 0;
    if (re) hipLaunchKernelGGL(crt_cond_lgs_kernel<true>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(crt_cond_lgs_kernel<false>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// ------------------------------------------------ S0 of the prefix
// S0 = T_GG - L_GS L_GS^T per wavenumber, one wave per tile (I, J), I >= J, of the G
// block: the T tile minus its JS products L(JS + I, p) L(JS + J, p)^T over the separator
// columns the prefix factorised (tiles in the packed column-major layout of the dense
// kernels, operands in the MFMA lanes as there).
struct S0Args {
    int kc, nt, ntB, JS;
    const zd *T, *Lpre;  // [kk][B tiles]
    zd *S0w;             // [kk][G tiles]
};

template <bool RE>
__global__ void __launch_bounds__(256) crt_cond_s0_kernel(S0Args A) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const long ntri = (long)A.nt * (A.nt + 1) / 2, ntriB = (long)A.ntB * (A.ntB + 1) / 2;
    const long w = (long)blockIdx.x * 4 + wave;
    if (w >= A.kc * ntri) return;
    const int kk = (int)(w / ntri);
    int t = (int)(w - kk * ntri), I = 0;
    while (t > I) {  // tile t = I (I + 1) / 2 + J
        t -= I + 1;
        ++I;
    }
    const int J = t;
    auto tri = [](long a, long b) { return a * (a + 1) / 2 + b; };
    const zd *Tk = A.T + kk * ntriB * 256, *Lk = A.Lpre + kk * ntriB * 256;
    const zd *src = Tk + tri(I + A.JS, J + A.JS) * 256;
    d4 ar, ai;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const zd e = ldzr<RE>(src + jc * T16 + 4 * v + rq);
        ar[v] = e.re;
        ai[v] = e.im;
    }
    const zd *la = Lk + tri(I + A.JS, 0) * 256, *lb = Lk + tri(J + A.JS, 0) * 256;
    for (int p = 0; p < A.JS; ++p) {
        zd a4[4], b4[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            a4[ks] = ldzr<RE>(la + (size_t)p * 256 + 64 * ks + lane);
            b4[ks] = ldzr<RE>(lb + (size_t)p * 256 + 64 * ks + lane);
        }
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) cmma<RE>(ar, ai, zd{-a4[ks].re, -a4[ks].im}, b4[ks]);
    }
    zd *o = A.S0w + kk * ntri * 256 + tri(I, J) * 256;
#pragma unroll
    for (int v = 0; v < 4; ++v) stz(o + jc * T16 + 4 * v + rq, zd{ar[v], RE ? 0.0 : ai[v]});
}

int crt_launch_cond_s0(hipStream_t st, int kc, int nt, int ntB, int JS, const zd *T, const zd *Lpre, zd *S0w,
                       int re) {
    S0Args a{kc, nt, ntB, JS, T, Lpre, S0w};
    const long n = (long)kc * nt * (nt + 1) / 2;
    if (re) hipLaunchKernelGGL(crt_cond_s0_kernel<true>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(crt_cond_s0_kernel<false>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// ------------------------------------------------ interior potentials
// kpot(node(i), k, l) = fI(i) (U_p(i, l) - sum_c Z_p(i, c) x_B(b(c), l)) for the rows i of
// piece p: each wave one 16-row tile times four 16-column tiles of x_B (gathered
// through the piece's column map), K-loop over the piece's coupled columns.
struct GemmArgs {
    int nsys, nP, k0, ldx, ntile, ncolP, nBp, E, K, fstrideC, rgroups, cgroups;
    const zd *X, *XB;
    const double *fI;
    const int *nodeI, *roff, *plen, *nbc, *colmap, *ucol;
    zd *kpotN;
};

template <bool RE>
__global__ void __launch_bounds__(256) crt_cond_gemm_kernel(GemmArgs A) {
    const int b = blockIdx.x;
    const int rg = b % A.rgroups, cgp = (b / A.rgroups) % A.cgroups, sy = b / (A.rgroups * A.cgroups);
    if (sy >= A.nsys) return;
    const int kk = sy / A.nP, p = sy - kk * A.nP;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int rt = 4 * rg + wave;  // row tile
    const int plen = A.plen[p];
    if (rt * T16 >= plen) return;
    const int k = A.k0 + kk;
    const zd *X = A.X + (size_t)sy * A.ntile * T16 * A.ldx;
    const zd *XB = A.XB + (size_t)kk * A.E * A.nBp;
    const int *cmap = A.colmap + (size_t)p * A.ncolP;
    C3 acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) c3zero(acc[c]);
    const int nkt = (A.nbc[p] + 15) / T16;
    for (int kt = 0; kt < nkt; ++kt) {
        zd ao[4];
        int bc[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            ao[ks] = ldzr<RE>(X + (size_t)(16 * kt + 4 * ks + rq) * A.ldx + 16 * rt + jc);
            bc[ks] = cmap[16 * kt + 4 * ks + rq];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int lcol = 16 * (4 * cgp + c) + jc;
            if (16 * (4 * cgp + c) < A.E) {
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) {
                    const zd bo = (lcol < A.E && bc[ks] >= 0) ? ldzr<RE>(XB + (size_t)lcol * A.nBp + bc[ks]) : zd{};
                    c3ma<RE>(acc[c], ao[ks], bo);
                }
            }
        }
    }
    const int r0 = A.roff[p];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int lcol = 16 * (4 * cgp + c) + jc;
        if (lcol >= A.E) continue;
        const int uc = A.ucol[(size_t)p * A.E + lcol];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = 16 * rt + 4 * v + rq;
            if (i >= plen) continue;
            const zd u = uc >= 0 ? ldzr<RE>(X + (size_t)uc * A.ldx + i) : zd{};
            const double f = A.fI[(size_t)kk * A.nP * A.fstrideC + r0 + i];
            const zd zx = c3get<RE>(acc[c], v);
            A.kpotN[((size_t)A.nodeI[r0 + i] * A.K + k) * A.E + lcol] = zd{(u.re - zx.re) * f, (u.im - zx.im) * f};
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

int crt_launch_cond_fb(hipStream_t st, int kc, int k0, int nS, int nBp, int nnzF, int dc, int nP, const int *qdiag,
                       const zd *area, double *fB, int *status) {
    const int n = kc * nBp;
    hipLaunchKernelGGL(crt_cond_fb_kernel, dim3((n + 255) / 256), dim3(256), 0, st, kc, k0, nS, nBp, nnzF, dc, nP, qdiag,
                       area, fB, status);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cond_rhs(hipStream_t st, int kc, int k0, int nent, int nnzF, int nBp, int fstrideC, int nP,
                        const int *erowg, const int *ebcol, const int *eq, const zd *area, const double *fI,
                        const double *fB, zd *rv) {
    const long n = (long)kc * nent;
    if (n == 0) return 0;
    hipLaunchKernelGGL(crt_cond_rhs_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, kc, k0, nent, nnzF,
                       nBp, fstrideC, nP, erowg, ebcol, eq, area, fI, fB, rv);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

template <int D>
static int launch_trsm(hipStream_t st, TrsmArgs a, bool bwd, bool re) {
    a.nvirt = (a.ntile * a.nsys + 7) / 8 * 8;
    const dim3 gr(a.nvirt), bl(256);
    if (re) {
        if (bwd) hipLaunchKernelGGL((crt_cond_trsm_kernel<D, true, true>), gr, bl, 0, st, a);
        else hipLaunchKernelGGL((crt_cond_trsm_kernel<D, false, true>), gr, bl, 0, st, a);
    } else {
        if (bwd) hipLaunchKernelGGL((crt_cond_trsm_kernel<D, true, false>), gr, bl, 0, st, a);
        else hipLaunchKernelGGL((crt_cond_trsm_kernel<D, false, false>), gr, bl, 0, st, a);
    }
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_cond_trsm(hipStream_t st, int D, int bwd, int nblk, int ldx, int ntile, int nsys, int nP,
                         const zd *panels, const zd *dinv, zd *X, const int *tptr, const int *erow, const int *ecol,
                         const zd *rv, int nent, const int *j0, int re) {
    TrsmArgs a{nblk, ldx, ntile, nsys, nP, 0, panels, dinv, X, tptr, erow, ecol, rv, nent, j0};
    switch (D) {
        case 2: return launch_trsm<2>(st, a, bwd, re);
        case 3: return launch_trsm<3>(st, a, bwd, re);
        case 4: return launch_trsm<4>(st, a, bwd, re);
        case 5: return launch_trsm<5>(st, a, bwd, re);
        case 6: return launch_trsm<6>(st, a, bwd, re);
        case 7: return launch_trsm<7>(st, a, bwd, re);
        case 8: return launch_trsm<8>(st, a, bwd, re);
        case 9: return launch_trsm<9>(st, a, bwd, re);
        case 10: return launch_trsm<10>(st, a, bwd, re);
        case 11: return launch_trsm<11>(st, a, bwd, re);
        case 12: return launch_trsm<12>(st, a, bwd, re);
        case 13: return launch_trsm<13>(st, a, bwd, re);
        case 14: return launch_trsm<14>(st, a, bwd, re);
        case 15: return launch_trsm<15>(st, a, bwd, re);
        case 16: return launch_trsm<16>(st, a, bwd, re);
        default: return CRT_ERR_UNSUPPORTED;
    }
}

int crt_launch_cond_schur(hipStream_t st, int kc, int k0, int nBp, int E, int ldx, int ntile, int nnzF, int fstrideC,
                          int nP, const int *gi_ptr, const int *gi_row, const int *gi_piece, const int *gi_q,
                          const int *roff, const int *bcol, const int *ucol, const int *gmap, const int *eB,
                          const zd *area, const double *fI, const double *fB, const zd *X, zd *T, zd *Gr) {
    SchurArgs a{kc,   k0,   nBp,  E,     ldx,  ntile,  nnzF, fstrideC, nP, gi_ptr, gi_row, gi_piece, gi_q,
                roff, bcol, ucol, gmap,  eB,   area,   fI,   fB,       X,  T,      Gr};
    const long nt = nBp / T16;
    const long n = (long)kc * (nt * (nt + 1) / 2 * 256 + (long)E * nBp);
    hipLaunchKernelGGL(crt_cond_schur_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// largest boundary tile count of the source systems and separator tile count of the
// prefix (five and six owned tiles per wave)
int crt_cond_dense_max_tiles() { return 40; }
int crt_cond_prefix_max_tiles() { return 48; }

int crt_launch_cond_bval(hipStream_t st, const DevGrid &g, int nsys, int E, int k0, int fstride, int nb, const int *bq,
                         const int2 *bg, const double *fG, zd *bval) {
    const long n = (long)nsys * nb;
    if (n == 0) return 0;
    hipLaunchKernelGGL(crt_cond_bval_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nsys, E, k0,
                       fstride, nb, bq, bg, fG, bval);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

template <int NI, bool PRE, int NW = DNW>
static int launch_dense(hipStream_t st, const DenseArgs &a, int nsys) {
    const size_t lds = (size_t)a.nBp * sizeof(zd);
    if (a.g.dc) hipLaunchKernelGGL((crt_cond_dense_kernel<NI, PRE, NW, true>), dim3(nsys), dim3(64 * NW), lds, st, a);
    else hipLaunchKernelGGL((crt_cond_dense_kernel<NI, PRE, NW, false>), dim3(nsys), dim3(64 * NW), lds, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// pre = 0: the source systems (nsys = kc E); pre = 1: the prefixes (nsys = kc)
int crt_launch_cond_dense(hipStream_t st, const DevGrid &g, int pre, int nG, int nt, int nBp, int JS, int E, int K,
                          int k0, int nsys, const zd *S0, const zd *T, const zd *Gr, const int *bptr, const int *bent,
                          const int *qdiag, const int *nodeB, const zd *bval, int nb, const double *fB, zd *Lt,
                          zd *Li, zd *Lpre, zd *Lipre, zd *S0w, zd *XB, zd *kpotN, int *status, int sstride) {
    const int ntB = nBp / T16;
    if (nt < 1 || nt > crt_cond_dense_max_tiles() || nBp % T16 || ntB != nt + JS) return CRT_ERR_ARG;
    if (pre && (JS < 1 || JS > crt_cond_prefix_max_tiles())) return CRT_ERR_ARG;
    DenseArgs a{g,     nG,     nt,   nBp,  JS,   E,     K,     k0,   S0,  T,       Gr,      bptr,  bent,
                qdiag, nodeB, bval, nb,   fB,   Lt,    Li,    Lpre, Lipre, S0w,   XB,      kpotN, status, sstride};
    if (pre) {
        switch ((JS + DNW - 1) / DNW) {
            case 1: return launch_dense<1, true>(st, a, nsys);
            case 2: return launch_dense<2, true>(st, a, nsys);
            case 3: return launch_dense<3, true>(st, a, nsys);
            case 4: return launch_dense<4, true>(st, a, nsys);
            case 5: return launch_dense<5, true>(st, a, nsys);
            case 6: return launch_dense<6, true>(st, a, nsys);
            default: return CRT_ERR_UNSUPPORTED;
        }
    }
    if (nt <= 20) switch ((nt + 3) / 4) {
            case 1: return launch_dense<1, false, 4>(st, a, nsys);
            case 2: return launch_dense<2, false, 4>(st, a, nsys);
            case 3: return launch_dense<3, false, 4>(st, a, nsys);
            case 4: return launch_dense<4, false, 4>(st, a, nsys);
            case 5: return launch_dense<5, false, 4>(st, a, nsys);
            default: return CRT_ERR_UNSUPPORTED;
        }
    switch ((nt + DNW - 1) / DNW) {
        case 1: return launch_dense<1, false>(st, a, nsys);
        case 2: return launch_dense<2, false>(st, a, nsys);
        case 3: return launch_dense<3, false>(st, a, nsys);
        case 4: return launch_dense<4, false>(st, a, nsys);
        case 5: return launch_dense<5, false>(st, a, nsys);
        default: return CRT_ERR_UNSUPPORTED;
    }
}

int crt_launch_cond_gemm(hipStream_t st, int nsys, int nP, int k0, int ldx, int ntile, int ncolP, int nBp, int E,
                         int K, int fstrideC, const zd *X, const zd *XB, const double *fI, const int *nodeI,
                         const int *roff, const int *plen, const int *nbc, const int *colmap, const int *ucol,
                         zd *kpotN, int re) {
    const int rgroups = (ldx / T16 + 3) / 4, cgroups = ((E + 15) / 16 + 3) / 4;
    GemmArgs a{nsys, nP,   k0,  ldx,  ntile, ncolP, nBp,    E,    K,  fstrideC, rgroups, cgroups,
               X,    XB,   fI,  nodeI, roff, plen,  nbc,    colmap, ucol, kpotN};
    const long n = (long)rgroups * cgroups * nsys;
    if (re) hipLaunchKernelGGL(crt_cond_gemm_kernel<true>, dim3((unsigned)n), dim3(256), 0, st, a);
    else hipLaunchKernelGGL(crt_cond_gemm_kernel<false>, dim3((unsigned)n), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
