// crt_volt_sens.hip -- inverse-Fourier back-transform, voltages and sensitivities.
//
// kpot is kept node-major on the device, kpotN[(node*K + k)*eanz + l], so the
// values an element or an electrode needs are contiguous.
//
//  * hpot(j,l) = sum_k kpot(j,l,k) w_k / pi          (src/rtrafo.f90:73-88)
//  * V = (hpot(e4,e2)-hpot(e4,e1)) - (hpot(e3,e2)-hpot(e3,e1)), pole
//    electrodes (0) masked                          (src/bvolt.f90:44-63)
//  * bvolt: sigmaa = kfak V; bvolti: sigmaa = -log V (src/bvolt.f90:77, src/bvolti.f90:80-88)
//  * J(i,e) = [sum_k w_k/pi sum_{m,n} elbg(e,mn,k) (phiB-phiA)_n (phiN-phiM)_m] sigma_e / V_i
//    (bsensi, src/bsensi.f90:98-187) or -(...) strom_i (bsens, src/bsens.f90:103-191)
#include "crt_common.hpp"

#include <algorithm>
#include <cstdlib>

// hpotE[e][l]: back-transformed potential at electrode node e for source l
__global__ void __launch_bounds__(256) crt_hpot_elec_kernel(int eanz, int K, const int *__restrict__ enr,
                                                            const double *__restrict__ kwnwi,
                                                            const zd *__restrict__ kpotN,
                                                            zd *__restrict__ hpotE) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= eanz * eanz) return;
    int l = t % eanz, e = t / eanz;
    const double pi = acos(-1.0);
    const zd *p = kpotN + (size_t)enr[e] * K * eanz + l;
    zd s = zd{0.0, 0.0};
    for (int k = 0; k < K; ++k) {
        zd v = p[(size_t)k * eanz];
        s.re = s.re + v.re * kwnwi[k];
        s.im = s.im + v.im * kwnwi[k];
    }
    hpotE[t] = zd{s.re / pi, s.im / pi};
}

// full hpot in the reference layout hpot(sanz, eanz): out[l*sanz + j]
__global__ void __launch_bounds__(256) crt_hpot_full_kernel(int sanz, int eanz, int K,
                                                            const double *__restrict__ kwnwi,
                                                            const zd *__restrict__ kpotN,
                                                            zd *__restrict__ hpot) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)sanz * eanz) return;
    int l = (int)(t % eanz), j = (int)(t / eanz);
    const double pi = acos(-1.0);
    const zd *p = kpotN + (size_t)j * K * eanz + l;
    zd s = zd{0.0, 0.0};
    for (int k = 0; k < K; ++k) {
        zd v = p[(size_t)k * eanz];
        s.re = s.re + v.re * kwnwi[k];
        s.im = s.im + v.im * kwnwi[k];
    }
    hpot[(size_t)l * sanz + j] = zd{s.re / pi, s.im / pi};
}

// meas[i] = (e1, e2, e3, e4) 1-based electrode numbers (0 = pole / absent)
__global__ void __launch_bounds__(256) crt_volt_kernel(int nanz, int eanz, int mode, const int4 *__restrict__ meas,
                                                       const zd *__restrict__ hpotE,
                                                       const double *__restrict__ kfak, zd *__restrict__ volt,
                                                       zd *__restrict__ sigmaa, int *__restrict__ status) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nanz) return;
    int4 m = meas[i];
    auto hp = [&](int pot, int src) -> zd {
        if (pot <= 0 || src <= 0) return zd{0.0, 0.0};
        return hpotE[(size_t)(pot - 1) * eanz + (src - 1)];
    };
    zd d1 = hp(m.w, m.y), d2 = hp(m.w, m.x), d3 = hp(m.z, m.y), d4 = hp(m.z, m.x);
    zd v = zsub(zsub(d1, d2), zsub(d3, d4));
    // mode 2, the DC branch of bvolti: volt = dcmplx(V) with a +0 imaginary part, so
    // that -log(V) of a negative V has the imaginary part -pi (src/bvolti.f90:65,89)
    if (mode == 2) v.im = 0.0;
    volt[i] = v;
    if (mode >= 1) {
        double a = zabs(v);
        if (a == 0.0) {
            *status = CRT_ERR_ZEROVOLT;
            sigmaa[i] = zd{0.0, 0.0};
        } else {
            sigmaa[i] = zd{-log(a), -atan2(v.im, v.re)};
        }
    } else {
        double f = kfak ? kfak[i] : 1.0;
        sigmaa[i] = zd{f * v.re, f * v.im};
    }
}

// One workgroup per area element; threads stride over measurements.
// LDS: potentials of the element's nodes for all (k, l), and its elbg(., k).
struct SensArgs {
    int elanz, nanz, eanz, K, maxn, npk, mode;  // mode 1: bsensi, 0: bsens
    const int *nrel;       // (neltot, maxn) 1-based
    const int *nkel;       // nodes per area element
    const int4 *meas;      // 1-based electrodes
    const double *kwnwi;
    const zd *kpotN;
    const double *elbg;    // [k][imn][e]
    const zd *sigma;
    const zd *volt;
    const double *strom;
    zd *sens;              // (nanz, manz) column-major
};

// Per (measurement, wavenumber) the 4 x nk potentials of the element's nodes
// are read once: u = phiB - phiA, v = phiN - phiM, and the bilinear form
// h = sum_m v_m sum_n elbg(m,n) u_n runs on registers (elbg from LDS broadcast).
template <int NK>
__device__ __forceinline__ void sens_element(const SensArgs &A, int e, const zd *pot, const double *eb) {
    constexpr int NP = NK * (NK + 1) / 2;
    const int K = A.K, E = A.eanz;
    const double pi = acos(-1.0);
    for (int i = threadIdx.x; i < A.nanz; i += blockDim.x) {
        const int4 m = A.meas[i];  // x=e1 (A), y=e2 (B), z=e3 (M), w=e4 (N)
        const int ia = max(m.x, 1) - 1, ib = max(m.y, 1) - 1, im = max(m.z, 1) - 1, in = max(m.w, 1) - 1;
        zd dum = zd{0.0, 0.0};
        for (int k = 0; k < K; ++k) {
            zd u[NK], v[NK];
#pragma unroll
            for (int n = 0; n < NK; ++n) {
                const zd *pk = pot + ((size_t)n * K + k) * E;
                zd pa = pk[ia], pb = pk[ib], pm = pk[im], pn = pk[in];
                if (m.x <= 0) pa = zd{0.0, 0.0};
                if (m.y <= 0) pb = zd{0.0, 0.0};
                if (m.z <= 0) pm = zd{0.0, 0.0};
                if (m.w <= 0) pn = zd{0.0, 0.0};
                u[n] = zsub(pb, pa);
                v[n] = zsub(pn, pm);
            }
            const double *w = eb + k * NP;
            zd h = zd{0.0, 0.0};
#pragma unroll
            for (int mi = 0; mi < NK; ++mi) {
                zd t = zd{0.0, 0.0};
#pragma unroll
                for (int mj = 0; mj < NK; ++mj) {
                    const int hi = mi > mj ? mi : mj, lo = mi > mj ? mj : mi;
                    const double wm = w[hi * (hi + 1) / 2 + lo];
                    t.re = fma(wm, u[mj].re, t.re);
                    t.im = fma(wm, u[mj].im, t.im);
                }
                h = zfma(h, v[mi], t);
            }
            dum.re = fma(h.re, A.kwnwi[k], dum.re);
            dum.im = fma(h.im, A.kwnwi[k], dum.im);
        }
        zd out;
        if (A.mode == 1) {
            zd t = zd{dum.re / pi, dum.im / pi};
            out = zdiv(zmul(t, A.sigma[e]), A.volt[i]);
        } else {
            double f = A.strom ? A.strom[i] / pi : 1.0 / pi;
            out = zd{-(dum.re * f), -(dum.im * f)};
        }
        A.sens[(size_t)e * A.nanz + i] = out;
    }
}

__global__ void __launch_bounds__(256) crt_sens_kernel(SensArgs A) {
    extern __shared__ double smem_v[];
    const int e = blockIdx.x;
    if (e >= A.elanz) return;
    const int K = A.K, E = A.eanz, nk = A.nkel[e];
    zd *pot = reinterpret_cast<zd *>(smem_v);                           // [m][k][l]
    double *eb = reinterpret_cast<double *>(pot + (size_t)nk * K * E);  // [k][imn]
    const int per_node = K * E;
    for (int q = threadIdx.x; q < nk * per_node; q += blockDim.x) {
        int m = q / per_node, r = q - m * per_node;
        int node = A.nrel[(size_t)e * A.maxn + m] - 1;
        pot[q] = A.kpotN[(size_t)node * per_node + r];
    }
    const int np = nk * (nk + 1) / 2;
    for (int q = threadIdx.x; q < K * np; q += blockDim.x) {
        int k = q / np, imn = q - k * np;
        eb[q] = A.elbg[((size_t)k * A.npk + imn) * A.elanz + e];
    }
    __syncthreads();
    if (nk == 3) sens_element<3>(A, e, pot, eb);
    else sens_element<4>(A, e, pot, eb);
}

// Electrode-pair form of the same sum.  Per element e and source pair (p, q):
//   G(p,q) = sum_k w_k sum_{a,b} elbg(e,ab,k) phi_{k,p}(n_a) phi_{k,q}(n_b)     (symmetric)
// so that sum_k w_k sum_{a,b} elbg (phiN-phiM)(n_a) (phiB-phiA)(n_b)
//        = G(N,B) - G(N,A) - G(M,B) + G(M,A)   (pole electrodes: G = 0).
// G is one (E x K*nk) x (K*nk x E) product per element, computed in 4x4
// register blocks from LDS (Phi and T = w S Phi), then every measurement needs
// four lookups instead of 4*nk potentials per wavenumber.
template <int NTH, int R>
__global__ void __launch_bounds__(NTH) crt_sens_g_kernel(SensArgs A, int Ep) {
    extern __shared__ double smem_g[];
    const int e = blockIdx.x;
    if (e >= A.elanz) return;
    const int tid = threadIdx.x;
    const int K = A.K, E = A.eanz, nk = A.nkel[e], KA = K * nk;
    zd *Phi = reinterpret_cast<zd *>(smem_g);  // [KA][Ep]
    zd *T = Phi + (size_t)KA * Ep;             // [KA][Ep]
    zd *G = Phi;                               // [Ep][Ep], after Phi and T are consumed
    for (int t = tid; t < KA * Ep; t += NTH) {
        const int ka = t / Ep, p = t - ka * Ep;
        const int k = ka / nk, a = ka - k * nk;
        const int node = A.nrel[(size_t)e * A.maxn + a] - 1;
        Phi[t] = p < E ? A.kpotN[((size_t)node * K + k) * E + p] : zd{0.0, 0.0};
    }
    __syncthreads();
    for (int t = tid; t < KA * Ep; t += NTH) {
        const int ka = t / Ep, q = t - ka * Ep;
        const int k = ka / nk, a = ka - k * nk;
        const double w = A.kwnwi[k];
        zd acc = zd{0.0, 0.0};
        for (int b = 0; b < nk; ++b) {
            const int hi = a > b ? a : b, lo = a > b ? b : a;
            const double sab = A.elbg[((size_t)k * A.npk + hi * (hi + 1) / 2 + lo) * A.elanz + e];
            const zd ph = Phi[(k * nk + b) * Ep + q];
            acc.re = fma(sab, ph.re, acc.re);
            acc.im = fma(sab, ph.im, acc.im);
        }
        T[t] = zd{w * acc.re, w * acc.im};
    }
    __syncthreads();
    const int nb4 = Ep / 4;
    zd acc[R][4][4];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[r][i][j] = zd{0.0, 0.0};
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int blk = tid + r * NTH;
        if (blk < nb4 * nb4) {
            const int pb = blk / nb4, qb = blk - pb * nb4;
            for (int ka = 0; ka < KA; ++ka) {
                zd ph[4], tq[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    ph[i] = Phi[ka * Ep + 4 * pb + i];
                    tq[i] = T[ka * Ep + 4 * qb + i];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[r][i][j] = zfma(acc[r][i][j], ph[i], tq[j]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int blk = tid + r * NTH;
        if (blk < nb4 * nb4) {
            const int pb = blk / nb4, qb = blk - pb * nb4;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) G[(4 * pb + i) * Ep + 4 * qb + j] = acc[r][i][j];
        }
    }
    __syncthreads();
    const double pi = acos(-1.0);
    for (int i = tid; i < A.nanz; i += NTH) {
        const int4 m = A.meas[i];  // x=e1 (A), y=e2 (B), z=e3 (M), w=e4 (N)
        auto gv = [&](int p, int q) -> zd { return (p > 0 && q > 0) ? G[(p - 1) * Ep + (q - 1)] : zd{0.0, 0.0}; };
        const zd dum = zadd(zsub(gv(m.w, m.y), gv(m.w, m.x)), zsub(gv(m.z, m.x), gv(m.z, m.y)));
        zd out;
        if (A.mode == 1) {
            zd t = zd{dum.re / pi, dum.im / pi};
            out = zdiv(zmul(t, A.sigma[e]), A.volt[i]);
        } else {
            double f = A.strom ? A.strom[i] / pi : 1.0 / pi;
            out = zd{-(dum.re * f), -(dum.im * f)};
        }
        A.sens[(size_t)e * A.nanz + i] = out;
    }
}

// Electrode-pair form on the matrix cores.  G is symmetric (S_k is), so only its
// upper 16x16 tiles are formed: G[P,Q] = Phi[:,P]^T T[:,Q] with v_mfma_f64_16x16x4f64
// (four real products per complex one), accumulated in registers over chunks of KC
// wavenumbers whose Phi and T = w S Phi rows are staged in LDS.  The tiles then go to
// LDS (over the Phi/T staging) for the measurement lookups.
typedef double sd4 __attribute__((ext_vector_type(4)));
constexpr int SGW = 8;  // wavefronts per element

__device__ __forceinline__ int sens_upper_tile(int P, int Q, int nt) { return P * nt - P * (P - 1) / 2 + (Q - P); }

template <int NT>
__global__ void __launch_bounds__(SGW * 64) crt_sens_gm_kernel(SensArgs A, int KC) {
    constexpr int EP = 16 * NT;
    constexpr int NUT = NT * (NT + 1) / 2;
    constexpr int TPW = (NUT + SGW - 1) / SGW;
    extern __shared__ double smem_gm[];
    const int e = blockIdx.x;
    if (e >= A.elanz) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int K = A.K, E = A.eanz, nk = A.nkel[e];
    const int KAc = (KC * nk + 3) & ~3;
    zd *Phi = reinterpret_cast<zd *>(smem_gm);  // [KAc][EP]
    zd *T = Phi + (size_t)KAc * EP;             // [KAc][EP]
    zd *G = Phi;                                // [NUT][256], after the last chunk
    sd4 gr[TPW], gi[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) gr[u] = gi[u] = sd4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += KC) {
        const int kc = min(KC, K - k0);
        for (int t = tid; t < KAc * EP; t += SGW * 64) {
            const int ka = t / EP, p = t - ka * EP;
            const int kl = ka / nk, a = ka - kl * nk;
            double2 v = make_double2(0.0, 0.0);
            if (kl < kc && p < E) {
                const int node = A.nrel[(size_t)e * A.maxn + a] - 1;
                v = *reinterpret_cast<const double2 *>(A.kpotN + ((size_t)node * K + k0 + kl) * E + p);
            }
            *reinterpret_cast<double2 *>(Phi + t) = v;
        }
        __syncthreads();
        for (int t = tid; t < KAc * EP; t += SGW * 64) {
            const int ka = t / EP, q = t - ka * EP;
            const int kl = ka / nk, a = ka - kl * nk;
            zd acc = zd{0.0, 0.0};
            if (kl < kc) {
                const int k = k0 + kl;
                const double w = A.kwnwi[k];
                for (int b = 0; b < nk; ++b) {
                    const int hi = a > b ? a : b, lo = a > b ? b : a;
                    const double sab = A.elbg[((size_t)k * A.npk + hi * (hi + 1) / 2 + lo) * A.elanz + e];
                    const zd ph = Phi[(kl * nk + b) * EP + q];
                    acc.re = fma(sab, ph.re, acc.re);
                    acc.im = fma(sab, ph.im, acc.im);
                }
                acc = zd{w * acc.re, w * acc.im};
            }
            *reinterpret_cast<double2 *>(T + t) = make_double2(acc.re, acc.im);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            if (wave + SGW * u < NUT) {
                // tile (P, Q), P <= Q, numbered as in sens_upper_tile
                int t = wave + SGW * u, P = 0;
                while (t >= NT - P) {
                    t -= NT - P;
                    ++P;
                }
                const zd *pa = Phi + rq * EP + 16 * P + jc;
                const zd *pb = T + rq * EP + 16 * (P + t) + jc;
                for (int ks = 0; ks < KAc; ks += 4) {
                    const double2 a2 = *reinterpret_cast<const double2 *>(pa + ks * EP);
                    const double2 b2 = *reinterpret_cast<const double2 *>(pb + ks * EP);
                    const zd a = zd{a2.x, a2.y}, b = zd{b2.x, b2.y};
                    gr[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, gr[u], 0, 0, 0);
                    gi[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, gi[u], 0, 0, 0);
                    gr[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b.im, gr[u], 0, 0, 0);
                    gi[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, gi[u], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int t = wave + SGW * u;
        if (t < NUT)
#pragma unroll
            for (int v = 0; v < 4; ++v) G[t * 256 + (4 * v + rq) * 16 + jc] = zd{gr[u][v], gi[u][v]};
    }
    __syncthreads();
    const double pi = acos(-1.0);
    auto gv = [&](int p, int q) -> zd {
        if (p <= 0 || q <= 0) return zd{0.0, 0.0};
        const int a = min(p, q) - 1, b = max(p, q) - 1;
        const double2 v =
            *reinterpret_cast<const double2 *>(G + sens_upper_tile(a >> 4, b >> 4, NT) * 256 + (a & 15) * 16 + (b & 15));
        return zd{v.x, v.y};
    };
    for (int i = tid; i < A.nanz; i += SGW * 64) {
        const int4 m = A.meas[i];  // x=e1 (A), y=e2 (B), z=e3 (M), w=e4 (N)
        const zd dum = zadd(zsub(gv(m.w, m.y), gv(m.w, m.x)), zsub(gv(m.z, m.x), gv(m.z, m.y)));
        zd out;
        if (A.mode == 1) {
            zd t = zd{dum.re / pi, dum.im / pi};
            out = zdiv(zmul(t, A.sigma[e]), A.volt[i]);
        } else {
            double f = A.strom ? A.strom[i] / pi : 1.0 / pi;
            out = zd{-(dum.re * f), -(dum.im * f)};
        }
        A.sens[(size_t)e * A.nanz + i] = out;
    }
}

namespace {
__device__ __forceinline__ zd ldz(const zd *p) {
    const double2 v = *reinterpret_cast<const double2 *>(p);
    return zd{v.x, v.y};
}
__device__ __forceinline__ void stz(zd *p, zd z) { *reinterpret_cast<double2 *>(p) = make_double2(z.re, z.im); }
}  // namespace

// potentials per thread staged in registers for the next stage (fewer where two
// workgroups share a CU's registers)
constexpr int sens_pf(int nt) { return nt <= 4 ? 4 : 6; }

// The same G tiles, but each workgroup runs through a contiguous range of elements
// (neighbouring elements share nodes: their potential rows stay in the XCD's L2) and
// the potentials and element matrix of the next (element, wavenumber chunk) are loaded
// into registers while the current one is multiplied and looked up, so the global
// loads do not stand between the stages.  The per-measurement factor is precomputed
// (vfac: sigma-free part of dum sigma / (pi V) for bsensi, strom / pi for bsens), so
// the lookups multiply instead of dividing.
template <int NT>
__global__ void __launch_bounds__(SGW * 64, NT <= 4 ? 4 : 2) crt_sens_gp_kernel(SensArgs A, int KC,
                                                                   const zd *__restrict__ vfac) {
    constexpr int EP = 16 * NT;
    constexpr int NUT = NT * (NT + 1) / 2;
    constexpr int TPW = (NUT + SGW - 1) / SGW;
    extern __shared__ double smem_gm[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int jc = lane & 15, rq = lane >> 4;
    const int K = A.K, E = A.eanz, maxn = A.maxn, npk = A.npk;
    const int KAm = (KC * maxn + 3) & ~3;        // LDS rows of the largest element
    zd *Phi = reinterpret_cast<zd *>(smem_gm);    // [KAc][EP]
    zd *T = Phi + (size_t)KAm * EP;               // [KAc][EP]
    zd *G = Phi;                                  // [NUT][256], after the last chunk
    const size_t gz = (size_t)NUT * 256 > (size_t)2 * KAm * EP ? (size_t)NUT * 256 : (size_t)2 * KAm * EP;
    double *S = reinterpret_cast<double *>(Phi + gz);  // [KC][npk]
    const int nch = (K + KC - 1) / KC;
    const int e0 = (int)((long)A.elanz * blockIdx.x / gridDim.x);
    const int e1 = (int)((long)A.elanz * (blockIdx.x + 1) / gridDim.x);
    const int nst = (e1 - e0) * nch;
    constexpr int SENS_PF = sens_pf(NT);
    const int nslot = (KAm * EP + SGW * 64 - 1) / (SGW * 64);  // <= SENS_PF (checked at launch)
    zd pre[SENS_PF];
    double spre = 0.0;
    auto fetch = [&](int st) {
        const int e = e0 + st / nch, k0 = (st % nch) * KC;
        const int kc = min(KC, K - k0), nk = A.nkel[e];
        const int KAc = (kc * nk + 3) & ~3;
#pragma unroll
        for (int j = 0; j < SENS_PF; ++j) {
            const int t = tid + j * SGW * 64;
            zd v = zd{0.0, 0.0};
            if (j < nslot && t < KAc * EP) {
                const int ka = t / EP, p = t - ka * EP;
                const int kl = ka / nk, a = ka - kl * nk;
                if (kl < kc && p < E) {
                    const int node = A.nrel[(size_t)e * maxn + a] - 1;
                    v = ldz(A.kpotN + ((size_t)node * K + k0 + kl) * E + p);
                }
            }
            pre[j] = v;
        }
        spre = tid < kc * npk ? A.elbg[((size_t)(k0 + tid / npk) * npk + tid % npk) * A.elanz + e] : 0.0;
    };
    sd4 gr[TPW], gi[TPW];
    if (nst > 0) fetch(0);
    for (int st = 0; st < nst; ++st) {
        const int e = e0 + st / nch, ch = st % nch, k0 = ch * KC;
        const int kc = min(KC, K - k0), nk = A.nkel[e];
        const int KAc = (kc * nk + 3) & ~3;
        __syncthreads();  // the previous stage is done with Phi, T, G and S
#pragma unroll
        for (int j = 0; j < SENS_PF; ++j) {
            const int t = tid + j * SGW * 64;
            if (j < nslot && t < KAc * EP) stz(Phi + t, pre[j]);
        }
        if (tid < kc * npk) S[tid] = spre;
        if (ch == 0)
#pragma unroll
            for (int u = 0; u < TPW; ++u) gr[u] = gi[u] = sd4{0.0, 0.0, 0.0, 0.0};
        __syncthreads();
        if (st + 1 < nst) fetch(st + 1);  // in flight during this stage
        // T = w S Phi, rows (k, a)
        for (int t = tid; t < KAc * EP; t += SGW * 64) {
            const int ka = t / EP, q = t - ka * EP;
            const int kl = ka / nk, a = ka - kl * nk;
            zd acc = zd{0.0, 0.0};
            if (kl < kc) {
                const double w = A.kwnwi[k0 + kl];
                for (int b = 0; b < nk; ++b) {
                    const int hi = a > b ? a : b, lo = a > b ? b : a;
                    const double sab = S[kl * npk + hi * (hi + 1) / 2 + lo];
                    const zd ph = Phi[(kl * nk + b) * EP + q];
                    acc.re = fma(sab, ph.re, acc.re);
                    acc.im = fma(sab, ph.im, acc.im);
                }
                acc = zd{w * acc.re, w * acc.im};
            }
            stz(T + t, acc);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            if (wave + SGW * u < NUT) {
                int t = wave + SGW * u, P = 0;
                while (t >= NT - P) {
                    t -= NT - P;
                    ++P;
                }
                const zd *pa = Phi + rq * EP + 16 * P + jc;
                const zd *pb = T + rq * EP + 16 * (P + t) + jc;
                for (int ks = 0; ks < KAc; ks += 4) {
                    const zd a = ldz(pa + ks * EP), b = ldz(pb + ks * EP);
                    gr[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.re, gr[u], 0, 0, 0);
                    gi[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.re, b.im, gi[u], 0, 0, 0);
                    gr[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a.im, b.im, gr[u], 0, 0, 0);
                    gi[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a.im, b.re, gi[u], 0, 0, 0);
                }
            }
        }
        if (ch != nch - 1) continue;
        __syncthreads();
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            const int t = wave + SGW * u;
            if (t < NUT)
#pragma unroll
                for (int v = 0; v < 4; ++v) G[t * 256 + (4 * v + rq) * 16 + jc] = zd{gr[u][v], gi[u][v]};
        }
        __syncthreads();
        auto gv = [&](int p, int q) -> zd {
            if (p <= 0 || q <= 0) return zd{0.0, 0.0};
            const int a = min(p, q) - 1, b = max(p, q) - 1;
            return ldz(G + sens_upper_tile(a >> 4, b >> 4, NT) * 256 + (a & 15) * 16 + (b & 15));
        };
        const zd sg = A.sigma[e];
        zd *out = A.sens + (size_t)e * A.nanz;
        // four measurements per thread and round, their loads issued together
        constexpr int LB = 4;
        for (int i0 = tid; i0 < A.nanz; i0 += LB * SGW * 64) {
            int4 m[LB];
            zd f[LB];
#pragma unroll
            for (int j = 0; j < LB; ++j) {
                const int i = i0 + j * SGW * 64;
                m[j] = i < A.nanz ? A.meas[i] : int4{0, 0, 0, 0};  // x=e1 (A), y=e2 (B), z=e3 (M), w=e4 (N)
                f[j] = i < A.nanz ? ldz(vfac + i) : zd{0.0, 0.0};
            }
#pragma unroll
            for (int j = 0; j < LB; ++j) {
                const int i = i0 + j * SGW * 64;
                const zd dum = zadd(zsub(gv(m[j].w, m[j].y), gv(m[j].w, m[j].x)), zsub(gv(m[j].z, m[j].x), gv(m[j].z, m[j].y)));
                if (i < A.nanz)
                    stz(out + i, A.mode == 1 ? zmul(zmul(dum, sg), f[j]) : zd{-(dum.re * f[j].re), -(dum.im * f[j].re)});
            }
        }
    }
}

// per-measurement factor of the pipelined form: 1 / (pi V) (bsensi), strom / pi (bsens)
__global__ void __launch_bounds__(256) crt_sens_vfac_kernel(int nanz, int mode, const zd *__restrict__ volt,
                                                            const double *__restrict__ strom, zd *__restrict__ vfac) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nanz) return;
    const double pi = acos(-1.0);
    if (mode == 1) {
        vfac[i] = zdiv(zd{1.0, 0.0}, zd{pi * volt[i].re, pi * volt[i].im});
    } else {
        vfac[i] = zd{(strom ? strom[i] : 1.0) / pi, 0.0};
    }
}

template <int NT>
static int launch_sens_gp(hipStream_t s, const SensArgs &a, int KC, size_t lds, zd *vfac) {
    (void)hipFuncSetAttribute((const void *)crt_sens_gp_kernel<NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    hipLaunchKernelGGL(crt_sens_vfac_kernel, dim3((a.nanz + 255) / 256), dim3(256), 0, s, a.nanz, a.mode, a.volt,
                       a.strom, vfac);
    // resident workgroups: two per CU when the LDS allows, each a contiguous element range
    const int per_cu = lds <= 80 * 1024 ? 2 : 1;  // registers hold two (NT <= 4) when the LDS does
    static const int mult = getenv("CRT_SENS_WG_MULT") ? std::max(1, atoi(getenv("CRT_SENS_WG_MULT"))) : 4;
    const int nwg = std::max(1, std::min(a.elanz, 256 * per_cu * mult));
    hipLaunchKernelGGL(crt_sens_gp_kernel<NT>, dim3(nwg), dim3(SGW * 64), lds, s, a, KC, vfac);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

template <int NT>
static int launch_sens_gm(hipStream_t s, const SensArgs &a, int KC, size_t lds) {
    (void)hipFuncSetAttribute((const void *)crt_sens_gm_kernel<NT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    hipLaunchKernelGGL(crt_sens_gm_kernel<NT>, dim3(a.elanz), dim3(SGW * 64), lds, s, a, KC);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// kpot in the reference layout kpot(sanz, eanz, kwnanz): out[(k*eanz + l)*sanz + j]
__global__ void __launch_bounds__(256) crt_kpot_ref_kernel(int sanz, int eanz, int K, const zd *__restrict__ kpotN,
                                                           zd *__restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)sanz * eanz * K) return;
    int j = (int)(t % sanz);
    long r = t / sanz;
    int l = (int)(r % eanz), k = (int)(r / eanz);
    out[t] = kpotN[((size_t)j * K + k) * eanz + l];
}

int crt_launch_hpot_elec(hipStream_t s, int eanz, int K, const int *enr, const double *kwnwi, const zd *kpotN,
                         zd *hpotE) {
    int n = eanz * eanz;
    hipLaunchKernelGGL(crt_hpot_elec_kernel, dim3((n + 255) / 256), dim3(256), 0, s, eanz, K, enr, kwnwi, kpotN, hpotE);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_hpot_full(hipStream_t s, int sanz, int eanz, int K, const double *kwnwi, const zd *kpotN, zd *hpot) {
    long n = (long)sanz * eanz;
    hipLaunchKernelGGL(crt_hpot_full_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sanz, eanz, K, kwnwi,
                       kpotN, hpot);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_volt(hipStream_t s, int nanz, int eanz, int mode, const int4 *meas, const zd *hpotE,
                    const double *kfak, zd *volt, zd *sigmaa, int *status) {
    hipLaunchKernelGGL(crt_volt_kernel, dim3((nanz + 255) / 256), dim3(256), 0, s, nanz, eanz, mode, meas, hpotE,
                       kfak, volt, sigmaa, status);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

size_t crt_sens_lds_bytes(int maxn, int K, int eanz) {
    return (size_t)maxn * K * eanz * sizeof(zd) + (size_t)K * 10 * sizeof(double);
}

int crt_launch_sens(hipStream_t s, int elanz, int nanz, int eanz, int K, int maxn, int npk, int mode,
                    const int *nrel, const int *nkel, const int4 *meas, const double *kwnwi, const zd *kpotN,
                    const double *elbg, const zd *sigma, const zd *volt, const double *strom, zd *sens,
                    zd *vfac) {
    SensArgs a{elanz, nanz, eanz, K, maxn, npk, mode, nrel, nkel, meas, kwnwi, kpotN, elbg, sigma, volt, strom, sens};
    static const int form = getenv("CRT_SENS_FORM") ? atoi(getenv("CRT_SENS_FORM")) : 3;
    // electrode-pair form on the matrix cores: up to 128 electrodes; 3 (default) the
    // pipelined workgroups over element ranges, 2 one workgroup per element
    if (form >= 2 && eanz <= 128) {
        const int nt = (eanz + 15) / 16, nut = nt * (nt + 1) / 2;
        const size_t gbytes = (size_t)nut * 256 * sizeof(zd);
        // wavenumber chunk: Phi and T rows within max(G, 80 KB) of LDS (two workgroups per CU)
        const size_t budget = std::max(gbytes, (size_t)80 * 1024);
        int KC = K;
        while (KC > 1 && (size_t)2 * ((KC * maxn + 3) & ~3) * 16 * nt * sizeof(zd) > budget) --KC;
        // CRT_SENS_KC: a smaller chunk (more chunks per element) for tests of the
        // cross-chunk accumulation; never more than fits
        if (const char *ev = getenv("CRT_SENS_KC")) {
            const int f = atoi(ev);
            if (f >= 1 && f < KC) KC = f;
        }
        const size_t lds = std::max(gbytes, (size_t)2 * ((KC * maxn + 3) & ~3) * 16 * nt * sizeof(zd));
        // the pipelined form: chunks of equal size whose Phi and T rows and element
        // matrices fit the same budget (C1: two chunks of 7 and 6 wavenumbers in 57 KB,
        // so two workgroups share a CU)
        int KP = K;
        for (int nch = 1; nch <= K; ++nch) {
            KP = (K + nch - 1) / nch;
            if ((size_t)2 * ((KP * maxn + 3) & ~3) * 16 * nt * sizeof(zd) + (size_t)KP * npk * sizeof(double) <=
                budget)
                break;
        }
        if (const char *ev = getenv("CRT_SENS_KC")) {
            const int f = atoi(ev);
            if (f >= 1 && f < KP) KP = f;
        }
        const size_t ldsp = std::max(gbytes, (size_t)2 * ((KP * maxn + 3) & ~3) * 16 * nt * sizeof(zd)) +
                            (size_t)KP * npk * sizeof(double);
        const int nslot = (((KP * maxn + 3) & ~3) * 16 * nt + SGW * 64 - 1) / (SGW * 64);
        if (form == 3 && vfac && ldsp <= 160 * 1024 && nslot <= sens_pf(nt)) {
            switch (nt) {
                case 1: return launch_sens_gp<1>(s, a, KP, ldsp, vfac);
                case 2: return launch_sens_gp<2>(s, a, KP, ldsp, vfac);
                case 3: return launch_sens_gp<3>(s, a, KP, ldsp, vfac);
                case 4: return launch_sens_gp<4>(s, a, KP, ldsp, vfac);
                case 5: return launch_sens_gp<5>(s, a, KP, ldsp, vfac);
                case 6: return launch_sens_gp<6>(s, a, KP, ldsp, vfac);
                case 7: return launch_sens_gp<7>(s, a, KP, ldsp, vfac);
                case 8: return launch_sens_gp<8>(s, a, KP, ldsp, vfac);
                default: break;
            }
        }
        if (lds <= 160 * 1024) {
            switch (nt) {
                case 1: return launch_sens_gm<1>(s, a, KC, lds);
                case 2: return launch_sens_gm<2>(s, a, KC, lds);
                case 3: return launch_sens_gm<3>(s, a, KC, lds);
                case 4: return launch_sens_gm<4>(s, a, KC, lds);
                case 5: return launch_sens_gm<5>(s, a, KC, lds);
                case 6: return launch_sens_gm<6>(s, a, KC, lds);
                case 7: return launch_sens_gm<7>(s, a, KC, lds);
                case 8: return launch_sens_gm<8>(s, a, KC, lds);
                default: break;
            }
        }
    }
    // electrode-pair form when its LDS fits
    {
        const int Ep = (eanz + 3) / 4 * 4, KA = K * maxn, nb = (Ep / 4) * (Ep / 4);
        const size_t lg = std::max((size_t)2 * KA * Ep, (size_t)Ep * Ep) * sizeof(zd);
        if (form >= 1 && lg <= 160 * 1024 && nb <= 1024) {
            if (nb <= 256) {
                (void)hipFuncSetAttribute((const void *)crt_sens_g_kernel<256, 1>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
                hipLaunchKernelGGL((crt_sens_g_kernel<256, 1>), dim3(elanz), dim3(256), lg, s, a, Ep);
            } else {
                (void)hipFuncSetAttribute((const void *)crt_sens_g_kernel<512, 2>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
                hipLaunchKernelGGL((crt_sens_g_kernel<512, 2>), dim3(elanz), dim3(512), lg, s, a, Ep);
            }
            return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
        }
    }
    size_t lds = crt_sens_lds_bytes(maxn, K, eanz);
    if (lds > 160 * 1024) return CRT_ERR_UNSUPPORTED;
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)crt_sens_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL(crt_sens_kernel, dim3(elanz), dim3(256), lds, s, a);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_kpot_ref(hipStream_t s, int sanz, int eanz, int K, const zd *kpotN, zd *out) {
    long n = (long)sanz * eanz * K;
    hipLaunchKernelGGL(crt_kpot_ref_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sanz, eanz, K, kpotN,
                       out);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// Geometric factors of a homogeneous half-space (src/bkfak.f90:48-136), one thread per
// measurement: with the electrodes A, B (current) and M, N (potential) and their
// mirror images in y = 0, r(P, Q) = 1/|P - Q| + 1/|P - Q'|, kfak = 4 pi / ((r(N,B) -
// r(N,A)) - (r(M,B) - r(M,A))); a vanishing denominator is errnr 93.
__global__ void __launch_bounds__(256) crt_bkfak_kernel(int nanz, const int4 *__restrict__ meas,
                                                        const int *__restrict__ enr, const double *__restrict__ x,
                                                        const double *__restrict__ y, double *__restrict__ kfak,
                                                        int *__restrict__ status) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nanz) return;
    const int4 m = meas[i];
    const int e[4] = {m.x, m.y, m.z, m.w};
    double xk[4], yk[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        xk[j] = e[j] > 0 ? x[enr[e[j] - 1]] : 0.0;
        yk[j] = e[j] > 0 ? y[enr[e[j] - 1]] : 0.0;
    }
    auto rr = [&](int u, int v) {
        if (e[u] <= 0 || e[v] <= 0) return 0.0;
        const double dx = xk[u] - xk[v], dym = yk[u] - yk[v], dyp = yk[u] + yk[v];
        return 1.0 / sqrt(dx * dx + dym * dym) + 1.0 / sqrt(dx * dx + dyp * dyp);
    };
    const double dum = (rr(3, 1) - rr(3, 0)) - (rr(2, 1) - rr(2, 0));
    if (fabs(dum) == 0.0) {
        *status = CRT_ERR_KFAK;
        kfak[i] = 0.0;
        return;
    }
    kfak[i] = 4.0 * acos(-1.0) / dum;
}

int crt_launch_bkfak(hipStream_t st, int nanz, const int4 *meas, const int *enr, const double *x, const double *y,
                     double *kfak, int *status) {
    if (nanz < 1) return 0;
    hipLaunchKernelGGL(crt_bkfak_kernel, dim3((nanz + 255) / 256), dim3(256), 0, st, nanz, meas, enr, x, y, kfak,
                       status);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
