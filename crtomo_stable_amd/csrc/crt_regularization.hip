// crt_regularization.hip -- model-side set-up of the Gauss-Newton step.
//
//   bnachbar   edge neighbours of the area elements (src/bnachbar.f90:44-90).  The
//              reference compares every element pair (O(elanz^2 maxn^2)); here the
//              edges are sorted by their node pair, which yields the same table:
//              for edge k of element i, the count of edges (j, jk), j != i, with the
//              same unordered node pair, and the largest such j (the reference
//              overwrites while it scans j upwards).
//   bsmatmreg  5-point smoothness on a regular nx x nz grid (src/bsmatm_mod.f90:190-362,
//              no discontinuities: ndis_x = ndis_z = 0 at :218-224).
//   bsmatmtri  neighbour smoothness on unstructured grids (src/bsmatm_mod.f90:364-460)
//              with the element centroids of src/relem.f90:128-138.
//   blam0      starting regularisation parameter (src/blam0.f90:120-175):
//              lam_0 = 5 * 2/(alfx+alfz) * mean_j |sum_i conj(J_ij) w_i sum_k J_ik|;
//              the reference's triple loop over (j, i, k) factorises into J^H W (J 1),
//              two passes over J.
#include "crt_common.hpp"
#include "../../include/crtomo_hip.h"

#include <algorithm>
#include <cmath>
#include <vector>

// ------------------------------------------------------------------ bnachbar
int crt_bnachbar(int elanz, int maxn, const int *nrel, int *nachbar) {
    if (elanz <= 0 || maxn <= 0 || !nrel || !nachbar) return CRT_ERR_ARG;
    struct Edge {
        long long key;
        int el;
    };
    const size_t ne = (size_t)elanz * maxn;
    std::vector<Edge> ed(ne);
    for (int i = 0; i < elanz; ++i)
        for (int k = 0; k < maxn; ++k) {
            long long a = nrel[(size_t)i * maxn + k], b = nrel[(size_t)i * maxn + (k + 1) % maxn];
            if (a > b) std::swap(a, b);
            ed[(size_t)i * maxn + k] = Edge{(a << 32) | (b & 0xffffffffLL), i};
        }
    std::vector<Edge> srt(ed);
    std::sort(srt.begin(), srt.end(), [](const Edge &p, const Edge &q) {
        return p.key != q.key ? p.key < q.key : p.el < q.el;
    });
    std::fill(nachbar, nachbar + ne + elanz, 0);
    for (int i = 0; i < elanz; ++i)
        for (int k = 0; k < maxn; ++k) {
            const Edge me = ed[(size_t)i * maxn + k];
            auto lo = std::lower_bound(srt.begin(), srt.end(), me.key,
                                       [](const Edge &p, long long key) { return p.key < key; });
            int cnt = 0, last = 0;
            for (auto it = lo; it != srt.end() && it->key == me.key; ++it)
                if (it->el != i) {
                    ++cnt;
                    last = it->el + 1;  // sorted by element: the last one is the largest
                }
            if (cnt) nachbar[(size_t)k * elanz + i] = last;
            nachbar[(size_t)maxn * elanz + i] += cnt;
        }
    return 0;
}

// ---------------------------------------------------------------- bsmatmreg
namespace {

__device__ __forceinline__ double centroid_sum(const int *nrel, int maxn, const double *c, int e, int nk) {
    double s = 0.0;
    for (int l = 0; l < nk; ++l) s += c[nrel[(size_t)e * maxn + l] - 1];
    return s;
}

__global__ void __launch_bounds__(256) crt_bsmatmreg_kernel(int nx, int nz, int manz, int maxn, const int *nrel,
                                                            const double *x, const double *y, double alfx,
                                                            double alfz, double *smatm) {
    // same operation order as the reference; no contraction into fused multiply-adds
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx * nz) return;
    const int i = t / nx + 1, j = t % nx + 1;
    const int kk = (i - 1) * nx + j - 1;  // 0-based element
    auto nd = [&](int e, int l) { return nrel[(size_t)e * maxn + l - 1] - 1; };
    const double dzleft = fabs(y[nd(kk, 4)] - y[nd(kk, 1)]);
    const double dzright = fabs(y[nd(kk, 3)] - y[nd(kk, 2)]);
    const double xmean = centroid_sum(nrel, maxn, x, kk, 4) / 4.0;
    double diag = 0.0, cx = 0.0, cz = 0.0;
    if (j > 1) {
        const double xleft = centroid_sum(nrel, maxn, x, kk - 1, 4) / 4.0;
        diag = 1.0 * alfx * dzleft / fabs(xmean - xleft);
    }
    if (j < nx) {
        const double xright = centroid_sum(nrel, maxn, x, kk + 1, 4) / 4.0;
        const double dum = 1.0 * alfx * dzright / fabs(xright - xmean);
        diag += dum;
        cx = -dum;
    }
    const double dxup = fabs(x[nd(kk, 3)] - x[nd(kk, 4)]);
    const double dxdown = fabs(x[nd(kk, 2)] - x[nd(kk, 1)]);
    const double zmean = centroid_sum(nrel, maxn, y, kk, 4) / 4.0;
    if (i > 1) {
        const double zup = centroid_sum(nrel, maxn, y, kk - nx, 4) / 4.0;
        diag += 1.0 * alfz * dxup / fabs(zup - zmean);
    }
    if (i < nz) {
        const double zdown = centroid_sum(nrel, maxn, y, kk + nx, 4) / 4.0;
        const double dum = 1.0 * alfz * dxdown / fabs(zmean - zdown);
        diag += dum;
        cz = -dum;
    }
    smatm[kk] = diag;
    smatm[manz + kk] = cx;
    smatm[2 * (size_t)manz + kk] = cz;
}

// --------------------------------------------------------------- bsmatmtri
__global__ void __launch_bounds__(256) crt_bsmatmtri_kernel(int elanz, int maxn, const int *nrel, const int *nkel,
                                                            const double *x, const double *y, const int *nachbar,
                                                            double alfx, double alfz, double *smatm) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= elanz) return;
    auto espx = [&](int e) { return centroid_sum(nrel, maxn, x, e, nkel[e]) / nkel[e]; };
    auto espy = [&](int e) { return centroid_sum(nrel, maxn, y, e, nkel[e]) / nkel[e]; };
    const double exi = espx(i), eyi = espy(i);
    double diag = 0.0;
    for (int k = 0; k < maxn; ++k) {
        const int n1 = nrel[(size_t)i * maxn + k] - 1, n2 = nrel[(size_t)i * maxn + (k + 1) % maxn] - 1;
        const double dx = x[n1] - x[n2], dy = y[n1] - y[n2];
        const double edglen = sqrt(dx * dx + dy * dy);
        const int nb = nachbar[(size_t)k * elanz + i];
        double off = 0.0;
        if (nb > 0) {
            const double ex = exi - espx(nb - 1), ey = eyi - espy(nb - 1);
            const double dist = sqrt(ex * ex + ey * ey);
            const double ang = atan2(ey, ex);
            const double ax = alfx * cos(ang), az = alfz * sin(ang);
            const double alfgeo = sqrt(ax * ax + az * az);
            const double rs = edglen / dist * alfgeo;
            off = -rs;
            diag += rs;
        }
        smatm[(size_t)k * elanz + i] = off;
    }
    smatm[(size_t)maxn * elanz + i] = diag;
}

// ------------------------------------------------------------------- blam0
// s_i = w_i sum_k J_ik over a block of columns (partial per column block)
__global__ void __launch_bounds__(256) crt_blam0_rows_kernel(int nanz, int manz, int cb, int fpi, const zd *J,
                                                             zd *part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nanz) return;
    const int k0 = blockIdx.y * cb, k1 = min(manz, k0 + cb);
    zd s = zd{0.0, 0.0};
    for (int k = k0; k < k1; ++k) {
        const zd v = J[(size_t)k * nanz + i];
        s.re += v.re;
        if (!fpi) s.im += v.im;
    }
    part[(size_t)blockIdx.y * nanz + i] = s;
}

__global__ void __launch_bounds__(256) crt_blam0_wsum_kernel(int nanz, int ncb, const double *w, const zd *part,
                                                             zd *s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nanz) return;
    zd a = zd{0.0, 0.0};
    for (int b = 0; b < ncb; ++b) a = zadd(a, part[(size_t)b * nanz + i]);
    s[i] = zd{a.re * w[i], a.im * w[i]};
}

// |sum_i conj(J_ij) s_i| per column, one wave per column
__global__ void __launch_bounds__(256) crt_blam0_cols_kernel(int nanz, int manz, int fpi, const zd *J, const zd *s,
                                                             double *jtj) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x >> 6) + wv;
    if (j >= manz) return;
    const zd *col = J + (size_t)j * nanz;
    double re = 0.0, im = 0.0;
    for (int i = lane; i < nanz; i += 64) {
        const zd v = col[i], d = s[i];
        if (fpi) {
            re += v.re * d.re;
        } else {
            re += v.re * d.re + v.im * d.im;
            im += v.re * d.im - v.im * d.re;
        }
    }
    for (int o = 32; o >= 1; o >>= 1) {
        re += __shfl_xor(re, o, 64);
        im += __shfl_xor(im, o, 64);
    }
    if (lane == 0) jtj[j] = fpi ? fabs(re) : hypot(re, im);
}

// fixed-order sum of jtj in one block
__global__ void __launch_bounds__(256) crt_blam0_sum_kernel(int manz, const double *jtj, double a2, double *out) {
    __shared__ double red[256];
    double v = 0.0;
    for (int j = threadIdx.x; j < manz; j += 256) v += jtj[j];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (double)manz * a2 * 5.0;
}

}  // namespace

int crt_launch_bsmatmreg(hipStream_t s, int nx, int nz, int manz, int maxn, const int *nrel, const double *x,
                         const double *y, double alfx, double alfz, double *smatm) {
    if (nx <= 0 || nz <= 0 || nx * nz != manz || maxn < 4) return CRT_ERR_ARG;
    hipLaunchKernelGGL(crt_bsmatmreg_kernel, dim3((manz + 255) / 256), dim3(256), 0, s, nx, nz, manz, maxn, nrel, x,
                       y, alfx, alfz, smatm);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_bsmatmtri(hipStream_t s, int elanz, int maxn, const int *nrel, const int *nkel, const double *x,
                         const double *y, const int *nachbar, double alfx, double alfz, double *smatm) {
    hipLaunchKernelGGL(crt_bsmatmtri_kernel, dim3((elanz + 255) / 256), dim3(256), 0, s, elanz, maxn, nrel, nkel, x,
                       y, nachbar, alfx, alfz, smatm);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

size_t crt_blam0_scratch(int nanz, int manz, int *ncb, int *cb) {
    // enough column blocks to fill the device; partial sums are (ncb, nanz) complex
    int rows = (nanz + 255) / 256;
    int want = std::max(1, 2048 / std::max(rows, 1));
    int c = std::max(64, (manz + want - 1) / want);
    *cb = c;
    *ncb = (manz + c - 1) / c;
    return (size_t)(*ncb) * nanz;
}

int crt_launch_blam0(hipStream_t s, int nanz, int manz, int fpi, const zd *J, const double *w, zd *part, int ncb,
                     int cb, zd *svec, double *jtj, double a2, double *out) {
    hipLaunchKernelGGL(crt_blam0_rows_kernel, dim3((nanz + 255) / 256, ncb), dim3(256), 0, s, nanz, manz, cb, fpi, J,
                       part);
    hipLaunchKernelGGL(crt_blam0_wsum_kernel, dim3((nanz + 255) / 256), dim3(256), 0, s, nanz, ncb, w, part, svec);
    hipLaunchKernelGGL(crt_blam0_cols_kernel, dim3((manz + 3) / 4), dim3(256), 0, s, nanz, manz, fpi, J, svec, jtj);
    hipLaunchKernelGGL(crt_blam0_sum_kernel, dim3(1), dim3(256), 0, s, manz, jtj, a2, out);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

// ------------------------------------------------------------- grid statistics
// besp_elem (src/besp_elem.f90:34-63): over every ordered pair of area-element
// centroids the smallest and largest distance and the extremes of |dx| and |dy|.  The
// square root is monotone and correctly rounded, so the extremes of the squared distance
// fl(fl(dx dx) + fl(dy dy)) give those of the distance; every reduction is a min or a max
// and therefore exact in any order.  Per block: partial [min d2, max d2, min |dx|,
// max |dx|, min |dy|, max |dy|].
__global__ void __launch_bounds__(256) crt_besp_pairs_kernel(int n, const double *__restrict__ cx,
                                                             const double *__restrict__ cy, double *__restrict__ part) {
#pragma clang fp contract(off)
    __shared__ double sx[256], sy[256];
    __shared__ double red[6][256];
    double v[6] = {HUGE_VAL, 0.0, HUGE_VAL, 0.0, HUGE_VAL, 0.0};
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double xi = i < n ? cx[i] : 0.0, yi = i < n ? cy[i] : 0.0;
    for (int j0 = 0; j0 < n; j0 += 256) {
        const int j = j0 + threadIdx.x;
        sx[threadIdx.x] = j < n ? cx[j] : 0.0;
        sy[threadIdx.x] = j < n ? cy[j] : 0.0;
        __syncthreads();
        if (i < n) {
            const int jn = min(256, n - j0);
            for (int t = 0; t < jn; ++t) {
                if (j0 + t == i) continue;
                const double ax = fabs(xi - sx[t]), ay = fabs(yi - sy[t]);
                const double d2 = ax * ax + ay * ay;
                v[0] = fmin(v[0], d2); v[1] = fmax(v[1], d2);
                v[2] = fmin(v[2], ax); v[3] = fmax(v[3], ax);
                v[4] = fmin(v[4], ay); v[5] = fmax(v[5], ay);
            }
        }
        __syncthreads();
    }
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int q = 0; q < 6; ++q)
                red[q][threadIdx.x] = (q & 1) ? fmax(red[q][threadIdx.x], red[q][threadIdx.x + w])
                                              : fmin(red[q][threadIdx.x], red[q][threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

int crt_launch_besp_pairs(hipStream_t s, int n, const double *cx, const double *cy, double *part, int *nblocks) {
    *nblocks = (n + 255) / 256;
    hipLaunchKernelGGL(crt_besp_pairs_kernel, dim3(*nblocks), dim3(256), 0, s, n, cx, cy, part);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
