// crt_band32.hip -- mixed-precision forward: the complex-symmetric band
// factorisation in FP32 on the matrix cores (FP32 products and accumulation),
// won back by FP64 iterative refinement.
//
// The factorisation follows the streaming left-looking scheme of the FP64
// kernel crt_factor16st_kernel (crt_band16.hip): 16x16 tiles, a pivot wave that
// updates and factors the diagonal tile of block column J (deferred square
// roots, L^-1 from the same column operations) and runs the factor's own
// forward substitution, three update waves that form
//     C(a) = A(J+a, J) - sum_m L(J+a, J-m) L(J, J-m)^T,   a = 1..D,
// and X(a) = C(a) L(J,J)^-T, stored as panel column J.  Panels, reciprocal
// pivots and y are complex FP32: a v_mfma_f32_16x16x4f32 takes half the cycles
// of the FP64 one and the panels half the bytes.
//
// Refinement of every (electrode, wavenumber) system on the scaled matrix
// A = F K F of src/scalab.f90 (F = diag(fak)), whose potential solves A x = g,
// g = fak(enode) e_enode (src/kompab.f90:162, src/vre.f90:45-69):
//     x0 = A32^-1 g;   n_s = A x_s - g (FP64, the CSR values);   x_{s+1} = x_s - A32^-1 n_s
// with the substitution kernel (crt_band16.hip) reading the FP32 panels in FP64
// arithmetic.  The host stops when max|F d| <= tol * max|F x| for every system.
//
// Accumulator layout: the FP64 kernels hold tile entry (row 4v+rq, column jc) in
// element v of lane jc + 16 rq.  v_mfma_f32_16x16x4f32 writes row 4 rq + v there,
// so every A operand is read with its rows permuted, row arow(i) = 4(i%4) + i/4 at
// lane i: the product then comes out in the FP64 kernels' layout, and the tile
// map, the diagonal chain and the stores are the same as there.
#include "crt_common.hpp"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int T16 = 16;
constexpr int PSF = 17;  // padded row stride (complex) of LDS tiles
constexpr int SWV = 4;   // wavefronts: pivot wave + three update waves
constexpr int SNU = SWV - 1;

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ zf ldc(const zf *p) {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    return zf{v.x, v.y};
}
__device__ __forceinline__ void stc(zf *p, zf z) { *reinterpret_cast<float2 *>(p) = make_float2(z.re, z.im); }
__device__ __forceinline__ zf cmul(zf a, zf b) { return zf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
// a - b*c
__device__ __forceinline__ zf cfms(zf a, zf b, zf c) {
    return zf{a.re - b.re * c.re + b.im * c.im, a.im - b.re * c.im - b.im * c.re};
}
__device__ __forceinline__ zf csel(bool c, zf a, zf b) { return zf{c ? a.re : b.re, c ? a.im : b.im}; }
__device__ __forceinline__ float readlane_f(float x, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), src));
}
__device__ __forceinline__ zf readlane_c(zf v, int src) { return zf{readlane_f(v.re, src), readlane_f(v.im, src)}; }

// 1/sqrt(m): hardware estimate and one Newton step
__device__ __forceinline__ float rsqrt_f32(float m) {
    const float y = __builtin_amdgcn_rsqf(m);
    return y * fmaf(-0.5f * m * y, y, 1.5f);
}

// L = principal sqrt(s) and 1/L = conj(L)/|s| (as pivot_fast in crt_band16.hip)
__device__ __forceinline__ void pivot_f32(zf s, zf &L, zf &inv) {
    const float m2 = fmaf(s.re, s.re, s.im * s.im);
    const float rm = rsqrt_f32(m2);
    const float mag = m2 * rm;
    const float h = 0.5f * (mag + fabsf(s.re));
    const float rh = rsqrt_f32(h);
    const float t = h * rh;
    const float q = 0.5f * fabsf(s.im) * rh;
    const bool pos = s.re >= 0.0f;
    L = zf{pos ? t : q, copysignf(pos ? q : t, s.im)};
    inv = zf{L.re * rm, -L.im * rm};
}

// 1/p = conj(p)/|p|^2, hardware reciprocal and one Newton step; bad for a zero or
// non-finite pivot (src/chol.f90 error 28)
__device__ __forceinline__ zf crecip_f32(zf p, bool &bad) {
    const float m2 = fmaf(p.re, p.re, p.im * p.im);
    bad = !(m2 > 0.0f) || !(m2 < 1.0e36f);
    float y = __builtin_amdgcn_rcpf(m2);
    y = fmaf(y, fmaf(-m2, y, 1.0f), y);
    return zf{p.re * y, -p.im * y};
}

__device__ __forceinline__ int arow(int i) { return 4 * (i & 3) + (i >> 2); }

__device__ __forceinline__ f4 mfma32(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// this lane's A-operand entries of an LDS tile (row stride PSF), rows permuted
__device__ __forceinline__ void lds_aop(const zf *Pa, int jc, int rq, zf (&va)[4]) {
    const zf *p = Pa + arow(jc) * PSF + rq;
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) va[kc] = ldc(p + 4 * kc);
}

// T -= Pa Pb^T: Pa's A-operand entries in registers, Pb in LDS
__device__ __forceinline__ void mfma32_update(f4 &tr, f4 &ti, const zf (&va)[4], const zf *Pb, int jc, int rq) {
    zf vb[4];
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) vb[kc] = ldc(Pb + jc * PSF + 4 * kc + rq);
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) {
        tr = mfma32(-va[kc].re, vb[kc].re, tr);
        ti = mfma32(-va[kc].re, vb[kc].im, ti);
        tr = mfma32(va[kc].im, vb[kc].im, tr);
        ti = mfma32(-va[kc].im, vb[kc].re, ti);
    }
}

// X = Pa Pb^T, both in LDS
__device__ __forceinline__ void mfma32_mul(f4 &xr, f4 &xi, const zf *Pa, const zf *Pb, int jc, int rq) {
    zf va[4], vb[4];
    lds_aop(Pa, jc, rq, va);
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) vb[kc] = ldc(Pb + jc * PSF + 4 * kc + rq);
    const f4 z = f4{0.0f, 0.0f, 0.0f, 0.0f};
    xr = mfma32(va[0].re, vb[0].re, z);
    xi = mfma32(va[0].re, vb[0].im, z);
    xr = mfma32(-va[0].im, vb[0].im, xr);
    xi = mfma32(va[0].im, vb[0].re, xi);
#pragma unroll
    for (int kc = 1; kc < 4; ++kc) {
        xr = mfma32(va[kc].re, vb[kc].re, xr);
        xi = mfma32(va[kc].re, vb[kc].im, xi);
        xr = mfma32(-va[kc].im, vb[kc].im, xr);
        xi = mfma32(va[kc].im, vb[kc].re, xi);
    }
}

// map entries (CSR indices) of this lane's four elements of tile (I, I-d)
template <int D>
__device__ __forceinline__ void tile_map(const DevGrid &g, int I, int d, int jc, int rq, int (&mi)[4]) {
    const int *mp = g.tmap + ((size_t)(I * (D + 1) + d) * T16) * T16 + jc;
#pragma unroll
    for (int v = 0; v < 4; ++v) mi[v] = I < g.nblk ? mp[(4 * v + rq) * T16] : g.nnz;
}

// FP64 CSR values of the tile, rounded to FP32, in the accumulator layout
__device__ __forceinline__ void tile_vals(const zd *vals, const int (&mi)[4], f4 &tr, f4 &ti) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const double2 e = *reinterpret_cast<const double2 *>(vals + mi[v]);
        tr[v] = (float)e.x;
        ti[v] = (float)e.y;
    }
}

}  // namespace

struct Factor32Args {
    DevGrid g;
    int nsys;
    const int2 *sys;
    const double *fak;  // [sys][nblk*16]
    const zd *vals;     // [sys][vstride] scaled FP64 CSR values
    zf *panels;         // [sys][nblk][D+1][16][16]
    zf *y;              // [sys][nblk*16] the factor's forward-substituted right-hand side
    zf *dinv;           // [sys][nblk*16] reciprocal pivots 1/L(i,i)
    int *status;
};

namespace {

template <int D>
struct alignas(16) St32Lds {
    static_assert(D >= SNU, "the solve staging tiles fit in Lrow");
    zf Lrow[D * T16 * PSF];  // L(J, J-m), m = 1..D; tiles 0..2 stage C(a) in the solve
    zf Li[T16 * PSF];        // L(J,J)^-1 by rows
    zf col[48];
    zf pinv[T16], sdiag[T16];
    zf yj[T16];
    zf yr[(D + 1) * T16];    // y of the last D+1 block rows
};

template <int D>
__device__ __forceinline__ void st32_stage(const zf *panel, St32Lds<D> &m, int J, int tid) {
    zf v[D];
#pragma unroll
    for (int mm = 1; mm <= D; ++mm)
        v[mm - 1] = J - mm >= 0 ? ldc(panel + ((size_t)(J - mm) * (D + 1) + mm) * 256 + tid) : zf{0.0f, 0.0f};
#pragma unroll
    for (int mm = 1; mm <= D; ++mm) stc(m.Lrow + (mm - 1) * T16 * PSF + (tid >> 4) * PSF + (tid & 15), v[mm - 1]);
}

// this lane's A-operand entries of L(J+a, J-mm), rows permuted
template <int D>
__device__ __forceinline__ void st32_operand(const zf *panel, int J, int a, int mm, int jc, int rq, zf (&o)[4]) {
    const zf *p = panel + ((size_t)(J - mm) * (D + 1) + a + mm) * 256 + arow(jc) * T16 + rq;
#pragma unroll
    for (int kc = 0; kc < 4; ++kc) o[kc] = ldc(p + 4 * kc);
}

// diagonal factorisation of one 16x16 tile on the pivot wave: column operations
// with 1/p (deferred square roots) on the tile, on an identity block (-> L^-1)
// and on the block's right-hand side
struct Chain32 {
    zf a[4], z[4], yv, p, r, pk;
};

template <int D>
__device__ __forceinline__ void chain32(const Factor32Args &A, St32Lds<D> &m, int s, int Jn, int lane, const f4 &tr,
                                        const f4 &ti, zf yv0, int &err) {
    const int jc = lane & 15, rq = lane >> 4;
    Chain32 q;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        q.a[v] = zf{tr[v], ti[v]};
        q.z[v] = zf{4 * v + rq == jc ? 1.0f : 0.0f, 0.0f};
    }
    q.yv = yv0;
    q.p = readlane_c(q.a[0], 0);
    bool bad;
    q.r = crecip_f32(q.p, bad);
    if (bad) err = CRT_ERR_PIVOT;
    q.pk = csel(lane == 0, q.p, zf{0.0f, 0.0f});
#pragma unroll
    for (int c = 0; c < T16; ++c) {
        if (jc == c) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                stc(m.col + 4 * v + rq, q.a[v]);
                stc(m.col + 16 + 4 * v + rq, q.z[v]);
            }
            if (rq == 0) stc(m.col + 32, q.yv);
        }
        zf dn, l2;
        if (c + 1 < T16) {
            const int n = c + 1;
            dn = readlane_c(q.a[n / 4], n + 16 * (n % 4));
            const zf ln = readlane_c(q.a[n / 4], c + 16 * (n % 4));
            l2 = cmul(ln, ln);
        }
        __builtin_amdgcn_wave_barrier();
        zf ci[4], zi[4];
        const zf cj = ldc(m.col + jc), wc = ldc(m.col + 32);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (4 * v + 3 > c) ci[v] = ldc(m.col + 4 * v + rq);
            if (4 * v <= c) zi[v] = ldc(m.col + 16 + 4 * v + rq);
        }
        __builtin_amdgcn_sched_barrier(0);
        const zf t = csel(jc > c, cmul(cj, q.r), zf{0.0f, 0.0f});  // u(jc, c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if (4 * v + 3 > c) q.a[v] = cfms(q.a[v], ci[v], t);
            if (4 * v <= c) q.z[v] = cfms(q.z[v], zi[v], t);
        }
        q.yv = cfms(q.yv, wc, t);
        if (c + 1 < T16) {
            q.p = cfms(dn, l2, q.r);
            q.r = crecip_f32(q.p, bad);
            if (bad) err = CRT_ERR_PIVOT;
            q.pk = csel(lane == c + 1, q.p, q.pk);
        }
        __builtin_amdgcn_wave_barrier();
    }
    // L(i,c) = u(i,c) sqrt(p_c), L^-1 rows scaled by 1/sqrt(p), y scaled likewise
    const int nblk = A.g.nblk;
    zf sc, ivc;
    pivot_f32(q.pk, sc, ivc);
    if (lane < T16) {
        m.pinv[lane] = ivc;
        m.sdiag[lane] = sc;
        A.dinv[(size_t)s * nblk * T16 + (size_t)Jn * T16 + lane] = ivc;
    }
    __builtin_amdgcn_wave_barrier();
    const zf ivj = m.pinv[jc], sj = m.sdiag[jc];
    zf *pg = A.panels + ((size_t)s * nblk + Jn) * (D + 1) * 256;
    // L^-1 above the diagonal of the diagonal tile (tile(i, jc) = L^-1(jc, i), i < jc),
    // as the FP64 kernels store it for the substitution
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int i = 4 * v + rq;
        const zf li = cmul(q.z[v], ivj);  // L^-1(jc, i)
        stc(pg + i * T16 + jc, i > jc ? cmul(q.a[v], ivj) : (i == jc ? sj : li));
        stc(m.Li + jc * PSF + i, li);
    }
    if (rq == 0) {
        const zf y = cmul(q.yv, ivj);
        m.yj[jc] = y;
        A.y[(size_t)s * nblk * T16 + (size_t)Jn * T16 + jc] = y;
    }
}

template <int D>
__device__ __forceinline__ void st32_pivot_wave(const Factor32Args &A, St32Lds<D> &m) {
    const DevGrid &g = A.g;
    const int nblk = g.nblk, s = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int enode = g.enr[A.sys[s].x];
    const double *fak = A.fak + (size_t)s * nblk * T16;
    const zd *vals = A.vals + (size_t)s * g.vstride;
    const zf *panel = A.panels + (size_t)s * nblk * (D + 1) * 256;
    int err = 0;
    int mi[4];
    f4 dr, di;
    tile_map<D>(g, 0, 0, jc, rq, mi);
    tile_vals(vals, mi, dr, di);
    tile_map<D>(g, 1, 0, jc, rq, mi);
    for (int J = 0; J < nblk; ++J) {
        st32_stage<D>(panel, m, J, tid);
        lds_barrier();
        const int mx = J < D ? J : D;
        // right-hand side of block J: b_J - sum_m L(J, J-m) y_{J-m}
        const int i = J * T16 + jc;
        zf yv = (i < g.sanz && i == enode) ? zf{(float)-fak[i], 0.0f} : zf{0.0f, 0.0f};
        for (int mm = 1; mm <= mx; ++mm) {
            const zf *lr = m.Lrow + (mm - 1) * T16 * PSF + jc * PSF;
            const zf *yp = m.yr + ((J - mm) % (D + 1)) * T16;
#pragma unroll
            for (int c = 0; c < T16; ++c) yv = cfms(yv, ldc(lr + c), ldc(yp + c));
        }
        for (int mm = 1; mm <= mx; ++mm) {
            const zf *lr = m.Lrow + (mm - 1) * T16 * PSF;
            zf va[4];
            lds_aop(lr, jc, rq, va);
            mfma32_update(dr, di, va, lr, jc, rq);
        }
        chain32<D>(A, m, s, J, lane, dr, di, yv, err);
        if (lane < T16) stc(m.yr + (J % (D + 1)) * T16 + lane, ldc(m.yj + lane));
        lds_barrier();  // L^-1 of block J is in LDS
        tile_vals(vals, mi, dr, di);
        tile_map<D>(g, J + 2, 0, jc, rq, mi);
        __syncthreads();  // panel column J is read by the next stage
    }
    if (lane == 0 && err && A.status[s] == 0) A.status[s] = err;
}

template <int D>
struct St32Cfg {
    static constexpr int NA = (D + SNU - 1) / SNU;  // off-diagonal tiles per update wave
    // tile a (1..D, larger = void) of update wave w in round u, in snake order
    static constexpr int tile(int w, int u) { return SNU * u + 1 + ((u & 1) ? (SNU - w) : (w - 1)); }
};

template <int D, int W>
__device__ __forceinline__ void st32_update_wave(const Factor32Args &A, St32Lds<D> &m) {
    using C = St32Cfg<D>;
    constexpr int NA = C::NA;
    const DevGrid &g = A.g;
    const int nblk = g.nblk, s = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const zd *vals = A.vals + (size_t)s * g.vstride;
    zf *panel = A.panels + (size_t)s * nblk * (D + 1) * 256;
    f4 cr[NA], ci[NA];
    int mi[NA][4];
#pragma unroll
    for (int u = 0; u < NA; ++u) {
        const int a = C::tile(W, u);
        if (a <= D) {
            tile_map<D>(g, a, a, jc, rq, mi[u]);
            tile_vals(vals, mi[u], cr[u], ci[u]);
            tile_map<D>(g, 1 + a, a, jc, rq, mi[u]);
        }
    }
    for (int J = 0; J < nblk; ++J) {
        st32_stage<D>(panel, m, J, tid);
        lds_barrier();
        // C(a) -= sum_m L(J+a, J-m) L(J, J-m)^T, the next operand in flight
        zf nx[4];
        bool have = false;  // nx holds the first operand of round u
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = C::tile(W, u);
            if (a <= D) {
                const int mx = J + a < nblk ? (J < D - a ? J : D - a) : 0;
                if (mx >= 1 && !have) st32_operand<D>(panel, J, a, 1, jc, rq, nx);
                have = false;
                for (int mm = 1; mm <= mx; ++mm) {
                    zf cur[4];
#pragma unroll
                    for (int kc = 0; kc < 4; ++kc) {
                        cur[kc].re = nx[kc].re;
                        cur[kc].im = nx[kc].im;
                    }
                    if (mm < mx) {
                        st32_operand<D>(panel, J, a, mm + 1, jc, rq, nx);
                    } else if (u + 1 < NA) {
                        const int a1 = C::tile(W, u + 1);
                        const int mx1 = J + a1 < nblk ? (J < D - a1 ? J : D - a1) : 0;
                        if (a1 <= D && mx1 >= 1) {
                            st32_operand<D>(panel, J, a1, 1, jc, rq, nx);
                            have = true;
                        }
                    }
                    mfma32_update(cr[u], ci[u], cur, m.Lrow + (mm - 1) * T16 * PSF, jc, rq);
                }
            }
        }
        lds_barrier();  // L^-1 of block J is in LDS
        // X(a) = C(a) (L^-1)^T, stored as panel column J
        zf *stg = m.Lrow + (W - 1) * T16 * PSF;
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int a = C::tile(W, u);
            if (a <= D) {
                zf *pg = panel + ((size_t)J * (D + 1) + a) * 256;
                if (J + a < nblk) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) stc(stg + (4 * v + rq) * PSF + jc, zf{cr[u][v], ci[u][v]});
                    __builtin_amdgcn_wave_barrier();
                    f4 xr, xi;
                    mfma32_mul(xr, xi, stg, m.Li, jc, rq);
#pragma unroll
                    for (int v = 0; v < 4; ++v) stc(pg + (4 * v + rq) * T16 + jc, zf{xr[v], xi[v]});
                    __builtin_amdgcn_wave_barrier();
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) stc(pg + (4 * v + rq) * T16 + jc, zf{0.0f, 0.0f});
                }
                // entering tile of column J+1 (map loaded one column ahead), map of column J+2
                tile_vals(vals, mi[u], cr[u], ci[u]);
                tile_map<D>(g, J + 2 + a, a, jc, rq, mi[u]);
            }
        }
        __syncthreads();  // panel column J is read by the next stage
    }
}

}  // namespace

template <int D>
__global__ void __launch_bounds__(SWV * 64, 2) crt_factor32st_kernel(Factor32Args A) {
    __shared__ St32Lds<D> m;
    if ((int)blockIdx.x >= A.nsys) return;
    const int wave = threadIdx.x >> 6;
    if (wave == 0) st32_pivot_wave<D>(A, m);
    else if (wave == 1) st32_update_wave<D, 1>(A, m);
    else if (wave == 2) st32_update_wave<D, 2>(A, m);
    else st32_update_wave<D, 3>(A, m);
}

// Negated residual n = A x - g of every right-hand side of a batch, in the form the
// substitution kernel takes as its right-hand side; rows past sanz get 0.  A is
// applied from the lower-triangle CSR values through the full-row map
// (frow: row pointers, fent: (column, CSR index) per entry of the full row).
__global__ void crt_refine_resid_kernel(int sanz, int fstride, int vstride, int nrhs, const int4 *rhs, const int *enr,
                                        const int *frow, const int2 *fent, const double *fak, const zd *vals,
                                        const zd *x, zd *nres, const int *skip) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int rid = blockIdx.y;
    if (i >= fstride || rid >= nrhs) return;
    if (skip && *skip) return;
    const int4 q = rhs[rid];
    zd acc = zd{0.0, 0.0};
    if (i < sanz) {
        const zd *v = vals + (size_t)q.z * vstride;
        const zd *xr = x + (size_t)rid * fstride;
        for (int p = frow[i]; p < frow[i + 1]; ++p) {
            const int2 e = fent[p];
            acc = zfma(acc, v[e.y], xr[e.x]);
        }
        if (i == enr[q.x]) acc.re -= fak[(size_t)q.z * fstride + i];
    }
    nres[(size_t)rid * fstride + i] = acc;
}

namespace {
template <int D>
int launch32(hipStream_t st, const Factor32Args &a) {
    if constexpr (D >= SNU) {
        if (a.g.tmapD != D) return CRT_ERR_UNSUPPORTED;
        hipLaunchKernelGGL(crt_factor32st_kernel<D>, dim3(a.nsys), dim3(SWV * 64), 0, st, a);
        return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
    }
    return CRT_ERR_UNSUPPORTED;
}
}  // namespace

// FP32 factorisation needs the 16x16 tile map and at least three sub-diagonal tiles
int crt_factor32_supported(int mb) {
    const int D = (mb + 15) / 16;
    return D >= SNU && D <= 16;
}

int crt_launch_factor32(hipStream_t st, const DevGrid &g16, int nsys, const int2 *sys, const double *fak,
                        const zd *vals, zf *panels, zf *y, zf *dinv, int *status) {
    Factor32Args a{g16, nsys, sys, fak, vals, panels, y, dinv, status};
    switch (g16.D) {
        case 3: return launch32<3>(st, a);
        case 4: return launch32<4>(st, a);
        case 5: return launch32<5>(st, a);
        case 6: return launch32<6>(st, a);
        case 7: return launch32<7>(st, a);
        case 8: return launch32<8>(st, a);
        case 9: return launch32<9>(st, a);
        case 10: return launch32<10>(st, a);
        case 11: return launch32<11>(st, a);
        case 12: return launch32<12>(st, a);
        case 13: return launch32<13>(st, a);
        case 14: return launch32<14>(st, a);
        case 15: return launch32<15>(st, a);
        case 16: return launch32<16>(st, a);
        default: return CRT_ERR_UNSUPPORTED;
    }
}

int crt_launch_refine_resid(hipStream_t st, int sanz, int fstride, int vstride, int nrhs, const int4 *rhs,
                            const int *enr, const int *frow, const int2 *fent, const double *fak, const zd *vals,
                            const zd *x, zd *nres, const int *skip) {
    if (nrhs <= 0) return 0;
    if (nrhs > 65535) return CRT_ERR_UNSUPPORTED;
    dim3 grid((fstride + 255) / 256, nrhs);
    hipLaunchKernelGGL(crt_refine_resid_kernel, grid, dim3(256), 0, st, sanz, fstride, vstride, nrhs, rhs, enr, frow,
                       fent, fak, vals, x, nres, skip);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// Convergence of a refinement on the device, for sweeps captured in a graph: state[0]
// 0 running, 1 converged (max over the right-hand sides of |fak d| / |fak x| at most
// tol), 2 stalled at most accept (the FP64 rounding floor), 3 failed (stalled above
// accept, or not finite); state[1] counts the sweeps that ran; rat[0] the previous
// ratio, rat[1] the last.  The same rule as the host loop of the per-context forward.
__global__ void __launch_bounds__(64) crt_refine_check_kernel(int nrhs, const double *stats, double tol,
                                                              double accept, int *state, double *rat) {
    if (state[0] != 0) return;
    const int lane = threadIdx.x;
    double r = 0.0;
    for (int q = lane; q < nrhs; q += 64) {
        const double d = stats[2 * q], x = stats[2 * q + 1];
        const double rq = x > 0.0 ? d / x : (d > 0.0 ? HUGE_VAL : 0.0);
        r = (rq <= r) ? r : rq;  // NaN propagates as not converged
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double t = __shfl_xor(r, o, 64);
        r = (t <= r) ? r : t;
    }
    if (lane == 0) {
        state[1] += 1;
        rat[1] = r;
        if (r <= tol) state[0] = 1;
        else if (!(r < 0.5 * rat[0])) state[0] = r <= accept ? 2 : 3;
        else rat[0] = r;
    }
}

__global__ void crt_refine_init_kernel(int *state, double *rat) {
    state[0] = 0;
    state[1] = 0;
    rat[0] = HUGE_VAL;
    rat[1] = 0.0;
}

int crt_launch_refine_init(hipStream_t st, int *state, double *rat) {
    hipLaunchKernelGGL(crt_refine_init_kernel, dim3(1), dim3(1), 0, st, state, rat);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_refine_check(hipStream_t st, int nrhs, const double *stats, double tol, double accept, int *state,
                            double *rat) {
    hipLaunchKernelGGL(crt_refine_check_kernel, dim3(1), dim3(64), 0, st, nrhs, stats, tol, accept, state, rat);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
