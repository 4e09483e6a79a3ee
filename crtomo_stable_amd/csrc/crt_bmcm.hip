// crt_bmcm.hip -- linearised model uncertainty on the device (src/bmcm_mod.f90:49-175,
// buncert): dense manz x manz FP64 matrices
//   ata     = Re(J^H W J), J^T W J for the DC inversion    (bata, :176-289)
//   ata_reg = ata + lam R, in the reference's write order  (bata_reg, :291-413)
//   cov     = ata_reg^-1  (chold + linvd, :415-615, or the Gauss_dble branch)
//   res     = cov ata                                      (bres, :617-688)
//   cov2    = res cov                                      (bmcm2, :690-765)
// The products and the factorisations are plain dense linear algebra: rocBLAS
// (dsyrk, dgemm) and rocSOLVER (dpotrf + dpotri, or dgetrf + dgetri); the
// kernels here scale J, mirror triangles and apply the regularisation fill.
// Matrices are column-major with leading dimension manz, as the reference's.
#include "crt_common.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace {

// Z = sqrt(w) J in the real view of J: column k holds (re, im) of every
// measurement, so Re(J^H W J) = Z^T Z.  real_only (the DC inversion's bata, which
// reads the real sensdc, src/bmcm_mod.f90:197-206): the imaginary parts count as zero.
__global__ void crt_bmcm_scale_kernel(size_t n, int nanz, const double *J, const double *w, int real_only,
                                      double *Z) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int i = (int)((idx % (2 * (size_t)nanz)) / 2);
    Z[idx] = (real_only && (idx & 1)) ? 0.0 : J[idx] * sqrt(w[i]);
}

// a(r, c) = a(c, r) for r < c: the upper triangle from the lower one
__global__ void crt_bmcm_mirror_kernel(int n, double *a) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (r >= n || r >= c) return;
    a[(size_t)c * n + r] = a[(size_t)r * n + c];
}

// bata_reg for ltri 0 (src/bmcm_mod.f90:317-340).  The reference loops
//   do j = 1, manz; do i = 1, manz
//     (1) reg(j,j) = ata(j,j) + lam smatm(i,1)
//     (2) if (i+1 < manz)  reg(i+1,j) = ata(i+1,j) + lam smatm(i+1,2); reg(j,i+1) = reg(i+1,j)
//     (3) if (i+nx < manz) reg(i+nx,j) = ata(i+nx,j) + lam smatm(i+nx,3); reg(j,i+nx) = reg(i+nx,j)
// so every entry keeps the value of the last write that reached it.  For entry (a, b)
// (1-based) the candidate writes, keyed by their place (j, i, statement) in that loop,
// are (1) at (a, manz, 1) when a = b; (2) at (b, a-1, 2) and (a, b-1, 2); (3) at
// (b, a-nx, 3) and (a, b-nx, 3), each where its condition holds; the latest key gives
// the value, no write leaves ata (ata is symmetric, so a mirrored write equals a
// direct one).
__global__ void crt_bmcm_reg0_kernel(int n, int nx, double lam, const double *sm, const double *ata, double *reg) {
    const int a0 = blockIdx.x * blockDim.x + threadIdx.x, b0 = blockIdx.y;
    if (a0 >= n) return;
    const int a = a0 + 1, b = b0 + 1;
    long long best = -1;  // key j * 4 (n + 1) + i * 4 + statement
    double s = 0.0;
    auto cand = [&](int j, int i, int st, double v) {
        const long long key = ((long long)j * (n + 1) + i) * 4 + st;
        if (key > best) { best = key; s = v; }
    };
    if (a == b) cand(a, n, 1, sm[n - 1]);
    if (a >= 2 && a <= n - 1) cand(b, a - 1, 2, sm[(size_t)n + a - 1]);
    if (b >= 2 && b <= n - 1) cand(a, b - 1, 2, sm[(size_t)n + b - 1]);
    if (a - nx >= 1 && a <= n - 1) cand(b, a - nx, 3, sm[2 * (size_t)n + a - 1]);
    if (b - nx >= 1 && b <= n - 1) cand(a, b - nx, 3, sm[2 * (size_t)n + b - 1]);
    const size_t q = (size_t)b0 * n + a0;
    reg[q] = best < 0 ? ata[q] : ata[q] + lam * s;
}

// bata_reg for ltri 1/2 (src/bmcm_mod.f90:360-383):
//   do j: reg(j,j) = ata(j,j) + lam smatm(j,maxn+1)
//         do k = 1, nachbar(j,maxn+1): ik = nachbar(j,k); if (ik /= 0)
//             reg(ik,j) = ata(ik,j) + lam smatm(j,k); reg(j,ik) = reg(ik,j)
// One thread per (j, k) writes entry (ik, j) and its mirror with the value of the
// latest write to that entry: keys (j, 0) for the diagonal, (j, k) for the direct
// write, (a, k') for the mirrored write of column a = ik with nachbar(a, k') = j.
__global__ void crt_bmcm_regn_kernel(int n, int maxn, double lam, const double *sm, const int *nb, const double *ata,
                                     double *reg) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * (maxn + 1)) return;
    const int j = t / (maxn + 1) + 1, k = t % (maxn + 1);  // k = 0: the diagonal
    auto NB = [&](int jj, int kk) { return nb[(size_t)(kk - 1) * n + jj - 1]; };
    auto SM = [&](int jj, int kk) { return sm[(size_t)(kk - 1) * n + jj - 1]; };
    // value of entry (a, b) after the whole loop
    auto final_value = [&](int a, int b, double &v) {
        long long best = -1;
        double s = 0.0;
        if (a == b) { best = (long long)a * (maxn + 2); s = SM(a, maxn + 1); }
        const int nbb = min(NB(b, maxn + 1), maxn), nba = min(NB(a, maxn + 1), maxn);
        for (int kk = 1; kk <= nbb; ++kk)
            if (NB(b, kk) == a) {
                const long long key = (long long)b * (maxn + 2) + kk;
                if (key > best) { best = key; s = SM(b, kk); }
            }
        for (int kk = 1; kk <= nba; ++kk)
            if (NB(a, kk) == b) {
                const long long key = (long long)a * (maxn + 2) + kk;
                if (key > best) { best = key; s = SM(a, kk); }
            }
        v = ata[(size_t)(b - 1) * n + a - 1] + lam * s;
    };
    if (k == 0) {
        double v;
        final_value(j, j, v);
        reg[(size_t)(j - 1) * n + j - 1] = v;
        return;
    }
    if (k > min(NB(j, maxn + 1), maxn)) return;
    const int ik = NB(j, k);
    if (ik <= 0 || ik > n) return;
    double v;
    final_value(ik, j, v);
    reg[(size_t)(j - 1) * n + ik - 1] = v;
    final_value(j, ik, v);
    reg[(size_t)(ik - 1) * n + j - 1] = v;
}

}  // namespace

#define CRT_ROCBLAS(call)                                                                \
    do {                                                                                 \
        rocblas_status s_ = (call);                                                      \
        if (s_ != rocblas_status_success) {                                              \
            fprintf(stderr, "crtomo_hip: %s failed: %s\n", #call, rocblas_status_to_string(s_)); \
            rocblas_destroy_handle(h);                                                   \
            return CRT_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

// The whole buncert on device buffers (each manz*manz, z 2*nanz*manz, info 1 int,
// ipiv manz ints).  res/cov2 may be null when not asked.  Returns 0, 108 (the
// reference's errnr) for a matrix the factorisation rejects, or a library code.
int crt_bmcm_run(hipStream_t st, int nanz, int manz, const zd *J, const double *w, int real_only, int ltri, int nx,
                 int maxn, const double *smatm, const int *nachbar, double lam, int gauss, double *ata, double *reg,
                 double *cov, double *res, double *cov2, double *z, int *info, int *ipiv) {
    if (nanz <= 0 || manz <= 0) return CRT_ERR_ARG;
    if ((size_t)manz * manz > (size_t)INT32_MAX || 2 * (size_t)nanz * manz > (size_t)INT32_MAX)
        return CRT_ERR_UNSUPPORTED;
    const int n = manz;
    const size_t nz = 2 * (size_t)nanz * manz;
    hipLaunchKernelGGL(crt_bmcm_scale_kernel, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, nz, nanz,
                       reinterpret_cast<const double *>(J), w, real_only, z);
    CRT_HIP(hipGetLastError());
    rocblas_handle h;
    if (rocblas_create_handle(&h) != rocblas_status_success) return CRT_ERR_HIP;
    CRT_ROCBLAS(rocblas_set_stream(h, st));
    const double one = 1.0, zero = 0.0;
    CRT_ROCBLAS(rocblas_dsyrk(h, rocblas_fill_lower, rocblas_operation_transpose, n, 2 * nanz, &one, z, 2 * nanz,
                              &zero, ata, n));
    const dim3 gsq((n + 255) / 256, n);
    hipLaunchKernelGGL(crt_bmcm_mirror_kernel, gsq, dim3(256), 0, st, n, ata);
    if (hipMemcpyAsync(reg, ata, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) {
        rocblas_destroy_handle(h);
        return CRT_ERR_HIP;
    }
    if (ltri == 0) {
        hipLaunchKernelGGL(crt_bmcm_reg0_kernel, gsq, dim3(256), 0, st, n, nx, lam, smatm, ata, reg);
    } else {
        const int nt = n * (maxn + 1);
        hipLaunchKernelGGL(crt_bmcm_regn_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, n, maxn, lam, smatm,
                           nachbar, ata, reg);
    }
    if (hipMemcpyAsync(cov, reg, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) {
        rocblas_destroy_handle(h);
        return CRT_ERR_HIP;
    }
    int hinfo = 0;
    if (!gauss) {
        CRT_ROCBLAS(rocsolver_dpotrf(h, rocblas_fill_lower, n, cov, n, info));
        if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rocblas_destroy_handle(h);
            return CRT_ERR_HIP;
        }
        if (hinfo == 0) {
            CRT_ROCBLAS(rocsolver_dpotri(h, rocblas_fill_lower, n, cov, n, info));
            hipLaunchKernelGGL(crt_bmcm_mirror_kernel, gsq, dim3(256), 0, st, n, cov);
        }
    } else {
        CRT_ROCBLAS(rocsolver_dgetrf(h, n, n, cov, n, ipiv, info));
        if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rocblas_destroy_handle(h);
            return CRT_ERR_HIP;
        }
        if (hinfo == 0) CRT_ROCBLAS(rocsolver_dgetri(h, n, cov, n, ipiv, info));
    }
    if (hinfo == 0) {
        if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            rocblas_destroy_handle(h);
            return CRT_ERR_HIP;
        }
    }
    if (hinfo != 0) {
        rocblas_destroy_handle(h);
        return 108;
    }
    if (res) {
        CRT_ROCBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, cov, n, ata, n,
                                  &zero, res, n));
        if (cov2)
            CRT_ROCBLAS(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, res, n, cov,
                                      n, &zero, cov2, n));
    }
    const bool ok = hipStreamSynchronize(st) == hipSuccess;
    rocblas_destroy_handle(h);
    return ok ? 0 : CRT_ERR_HIP;
}
