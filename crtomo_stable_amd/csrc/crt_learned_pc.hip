// crt_learned_pc.hip -- learned sparse approximate inverse for the Gauss-Newton CG.
//
// A per-cell network 8 -> 64 -> 64 -> 64 -> 5 (tanh) reads local features of the
// scaled normal matrix B = C (J^H W J + lam R) C of a regular nx x nz grid
//   f0      log10 coverage diag(J^H W J) of the cell
//   f1, f2  B(W, j) (real, imaginary): coupling to the left neighbour (data + lam R)
//   f3, f4  B(N, j): coupling to the upper neighbour
//   f5, f6  B(N, W): coupling between those two neighbours
//   f7      log10(lam R_jj / coverage)
// (standardised with the training set's mean and spread) and emits the row of the
// lower-triangular factor G of M = G^H G on the regularisation stencil pattern
// {N, W, j} as a correction y to an analytic estimate b (pc_baseline):
// (b0+y0) + i (b1+y1) = g_N, (b2+y2) + i (b3+y3) = g_W, exp(b4+y4) = g_jj > 0, so M
// is Hermitian positive definite for any weights.  The targets are the exact FSAI rows
// of systems the inversion produces (crt_cg.hip); training is full-batch Adam on
// the mean squared error with the forward, backward and update kernels below.
// Inference uses the weights rounded to BF16 with FP32 accumulation.
#include "crt_common.hpp"

#include <cmath>

#define PC_NI 8
#define PC_NH 64
#define PC_NO 5
// parameter offsets in one flat FP32 vector
#define PC_W1 0
#define PC_B1 (PC_W1 + PC_NH * PC_NI)
#define PC_W2 (PC_B1 + PC_NH)
#define PC_B2 (PC_W2 + PC_NH * PC_NH)
#define PC_W3 (PC_B2 + PC_NH)
#define PC_B3 (PC_W3 + PC_NH * PC_NH)
#define PC_W4 (PC_B3 + PC_NH)
#define PC_B4 (PC_W4 + PC_NO * PC_NH)
#define PC_NP (PC_B4 + PC_NO)  // 9221

int crt_pc_nparams() { return PC_NP; }

namespace {

__device__ __forceinline__ float bf16_round(float x) {
    // round to nearest even on the upper 16 bits (finite inputs)
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// one dense layer out = W in + b (W row-major [no][ni]) from LDS, optional tanh
template <int NO, int NIN, bool ACT>
__device__ __forceinline__ void pc_layer(const float *W, const float *b, const float (&in)[NIN], float (&out)[NO]) {
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        float a = b[o];
#pragma unroll
        for (int k = 0; k < NIN; ++k) a = fmaf(W[o * NIN + k], in[k], a);
        out[o] = ACT ? tanhf(a) : a;
    }
}

// Outputs of the FSAI row on {N, W, j} when the N-W coupling is neglected: with
// bn = B(N, j), bw = B(W, j) and d = 1 - |bn|^2 - |bw|^2 (kept >= 0.05),
// g_N = -bn / sqrt(d), g_W = -bw / sqrt(d), log g_jj = -log(d) / 2.  The network
// learns the correction to these values.
__device__ __forceinline__ void pc_baseline(const float *f, float (&b)[PC_NO]) {
    const float bwr = f[1], bwi = f[2], bnr = f[3], bni = f[4];
    const float d = fmaxf(1.f - (bnr * bnr + bni * bni) - (bwr * bwr + bwi * bwi), 0.05f);
    const float sc = 1.f / sqrtf(d);
    b[0] = -bnr * sc;
    b[1] = -bni * sc;
    b[2] = -bwr * sc;
    b[3] = -bwi * sc;
    b[4] = -0.5f * logf(d);
}

}  // namespace

// ---------------------------------------------------------------- features}  // namespace

// ---------------------------------------------------------------- features
// per cell: raw features, the pattern slots of N and W (or -1), and the teacher row
struct PcCell {
    int manz, nx, nz, nanz;
    double lam;
    const double *cgfac, *smatm, *cy;  // cy: cell centroid depth coordinate
    double ymin, ymax;
    const zd *J;                       // (nanz, manz) column-major
    const double *w;                   // nanz
};

// B_jW and B_jN per cell: one wave per cell, two weighted column products
__global__ void __launch_bounds__(256) crt_pc_coupling_kernel(PcCell A, zd *cpl) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x >> 6) + wv;
    if (j >= A.manz) return;
    const int manz = A.manz, nx = A.nx, ix = j % nx, iz = j / nx;
    const int cw = ix > 0 ? j - 1 : -1, cn = iz > 0 ? j - nx : -1;
    const zd *cj = A.J + (size_t)j * A.nanz;
    double wr = 0.0, wi = 0.0, nr = 0.0, ni = 0.0, xr = 0.0, xi = 0.0;
    for (int i = lane; i < A.nanz; i += 64) {
        const zd vj = cj[i];
        const double wt = A.w[i];
        const zd vw = cw >= 0 ? A.J[(size_t)cw * A.nanz + i] : zd{0.0, 0.0};
        const zd vn = cn >= 0 ? A.J[(size_t)cn * A.nanz + i] : zd{0.0, 0.0};
        // w conj(J_a) J_b for (W, j), (N, j), (N, W)
        wr += wt * (vw.re * vj.re + vw.im * vj.im);
        wi += wt * (vw.re * vj.im - vw.im * vj.re);
        nr += wt * (vn.re * vj.re + vn.im * vj.im);
        ni += wt * (vn.re * vj.im - vn.im * vj.re);
        xr += wt * (vn.re * vw.re + vn.im * vw.im);
        xi += wt * (vn.re * vw.im - vn.im * vw.re);
    }
    for (int o = 32; o >= 1; o >>= 1) {
        wr += __shfl_xor(wr, o, 64);
        wi += __shfl_xor(wi, o, 64);
        nr += __shfl_xor(nr, o, 64);
        ni += __shfl_xor(ni, o, 64);
        xr += __shfl_xor(xr, o, 64);
        xi += __shfl_xor(xi, o, 64);
    }
    if (lane != 0) return;
    const double c = A.cgfac[j];
    zd bw = zd{0.0, 0.0}, bn = zd{0.0, 0.0}, bx = zd{0.0, 0.0};
    if (cw >= 0) {
        const double cc = c * A.cgfac[cw];
        bw = zd{cc * (wr + A.lam * A.smatm[manz + cw]), cc * wi};
    }
    if (cn >= 0) {
        const double cc = c * A.cgfac[cn];
        bn = zd{cc * (nr + A.lam * A.smatm[2 * (size_t)manz + cn]), cc * ni};
    }
    if (cw >= 0 && cn >= 0) {
        // N and W are diagonal neighbours: no stencil coupling
        const double cc = A.cgfac[cn] * A.cgfac[cw];
        bx = zd{cc * xr, cc * xi};
    }
    cpl[3 * (size_t)j] = bw;
    cpl[3 * (size_t)j + 1] = bn;
    cpl[3 * (size_t)j + 2] = bx;
}

__global__ void __launch_bounds__(256) crt_pc_features_kernel(PcCell A, const zd *cpl, float *feat) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.manz) return;
    const double cf = A.cgfac[j];
    const double lr = A.lam * A.smatm[j];
    const double cov = fmax(1.0 / (cf * cf) - lr, 1e-300);
    const zd bw = cpl[3 * (size_t)j], bn = cpl[3 * (size_t)j + 1], bx = cpl[3 * (size_t)j + 2];
    float f[PC_NI];
    f[0] = (float)log10(cov);
    f[1] = (float)bw.re;
    f[2] = (float)bw.im;
    f[3] = (float)bn.re;
    f[4] = (float)bn.im;
    f[5] = (float)bx.re;
    f[6] = (float)bx.im;
    f[7] = (float)log10(fmax(lr, 1e-300) / cov);
#pragma unroll
    for (int k = 0; k < PC_NI; ++k) feat[(size_t)j * PC_NI + k] = f[k];
}

// teacher targets from the FSAI rows g (pattern: [.., N, W, j] right-aligned) and
// the output mask (1 where the neighbour exists)
__global__ void __launch_bounds__(256) crt_pc_targets_kernel(int manz, int nx, const int *pat, const zd *g,
                                                             const float *feat, float *tgt, float *mask) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= manz) return;
    constexpr int NP = CRT_FSAI_NP;
    zd gn = zd{0.0, 0.0}, gw = zd{0.0, 0.0}, gs = zd{1.0, 0.0};
    bool hn = false, hw = false;
    for (int s = 0; s < NP; ++s) {
        const int c = pat[(size_t)j * NP + s];
        const zd v = g[(size_t)j * NP + s];
        if (c == j) gs = v;
        else if (c == j - nx && c >= 0) { gn = v; hn = true; }
        else if (c == j - 1 && c >= 0) { gw = v; hw = true; }
    }
    float *t = tgt + (size_t)j * PC_NO, *m = mask + (size_t)j * PC_NO;
    float b[PC_NO];
    pc_baseline(feat + (size_t)j * PC_NI, b);
    t[0] = (float)gn.re - b[0]; t[1] = (float)gn.im - b[1]; t[2] = (float)gw.re - b[2]; t[3] = (float)gw.im - b[3];
    t[4] = (float)log(fmax(gs.re, 1e-30)) - b[4];
    m[0] = m[1] = hn ? 1.f : 0.f;
    m[2] = m[3] = hw ? 1.f : 0.f;
    m[4] = 1.f;
}

// feature standardisation: mean and spread per feature (one block, fixed order)
__global__ void __launch_bounds__(256) crt_pc_norm_kernel(int manz, const float *feat, float *norm) {
    __shared__ double red[2][256];
    for (int k = 0; k < PC_NI; ++k) {
        double s = 0.0, q = 0.0;
        for (int j = threadIdx.x; j < manz; j += 256) {
            const double v = feat[(size_t)j * PC_NI + k];
            s += v;
            q += v * v;
        }
        red[0][threadIdx.x] = s;
        red[1][threadIdx.x] = q;
        __syncthreads();
        for (int o = 128; o >= 1; o >>= 1) {
            if ((int)threadIdx.x < o) {
                red[0][threadIdx.x] += red[0][threadIdx.x + o];
                red[1][threadIdx.x] += red[1][threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double mu = red[0][0] / manz, var = red[1][0] / manz - mu * mu;
            norm[k] = (float)mu;
            // features nearly constant on the first system (lambda, say) are scaled by at most 10
            norm[PC_NI + k] = (float)(1.0 / sqrt(fmax(var, 0.0) + 1e-2));
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- training
// forward + backward of one cell per thread; stores the per-cell products needed for
// the weight gradients (inputs, activations, deltas) and the loss per block
__global__ void __launch_bounds__(64) crt_pc_fwdbwd_kernel(int manz, const float *params, const float *norm,
                                                            const float *feat, const float *tgt, const float *mask,
                                                            float *act, float *delta, double *part_loss) {
    __shared__ float P[PC_NP];
    __shared__ double red[64];
    for (int t = threadIdx.x; t < PC_NP; t += blockDim.x) P[t] = params[t];
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double loss = 0.0;
    if (j < manz) {
        // per cell: [x | h1 | h2 | h3] and [d1 | d2 | d3 | d4]
        float *ac = act + (size_t)j * (PC_NI + 3 * PC_NH);
        float *de = delta + (size_t)j * (3 * PC_NH + PC_NO);
        float x[PC_NI], ha[PC_NH], hb[PC_NH], y[PC_NO];
#pragma unroll
        for (int k = 0; k < PC_NI; ++k) x[k] = (feat[(size_t)j * PC_NI + k] - norm[k]) * norm[PC_NI + k];
        pc_layer<PC_NH, PC_NI, true>(P + PC_W1, P + PC_B1, x, ha);
#pragma unroll
        for (int k = 0; k < PC_NI; ++k) ac[k] = x[k];
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) ac[PC_NI + k] = ha[k];
        pc_layer<PC_NH, PC_NH, true>(P + PC_W2, P + PC_B2, ha, hb);
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) ac[PC_NI + PC_NH + k] = hb[k];
        pc_layer<PC_NH, PC_NH, true>(P + PC_W3, P + PC_B3, hb, ha);
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) ac[PC_NI + 2 * PC_NH + k] = ha[k];
        pc_layer<PC_NO, PC_NH, false>(P + PC_W4, P + PC_B4, ha, y);
        // loss = sum over masked outputs of (y - t)^2 / manz; dL/dy
        float d4[PC_NO];
#pragma unroll
        for (int o = 0; o < PC_NO; ++o) {
            const float e = (y[o] - tgt[(size_t)j * PC_NO + o]) * mask[(size_t)j * PC_NO + o];
            loss += (double)e * e;
            d4[o] = 2.f * e / (float)manz;
        }
        // backward: d3 = (W4^T d4)(1 - h3^2), d2 = (W3^T d3)(1 - h2^2), d1 = (W2^T d2)(1 - h1^2);
        // ha holds h3, the other activations come back from the cell's row
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) {
            float a = 0.f;
#pragma unroll
            for (int o = 0; o < PC_NO; ++o) a = fmaf(P[PC_W4 + o * PC_NH + k], d4[o], a);
            hb[k] = a * (1.f - ha[k] * ha[k]);
        }
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) de[2 * PC_NH + k] = hb[k];
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) {
            float a = 0.f;
#pragma unroll
            for (int o = 0; o < PC_NH; ++o) a = fmaf(P[PC_W3 + o * PC_NH + k], hb[o], a);
            const float h = ac[PC_NI + PC_NH + k];
            ha[k] = a * (1.f - h * h);
        }
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) de[PC_NH + k] = ha[k];
#pragma unroll
        for (int k = 0; k < PC_NH; ++k) {
            float a = 0.f;
#pragma unroll
            for (int o = 0; o < PC_NH; ++o) a = fmaf(P[PC_W2 + o * PC_NH + k], ha[o], a);
            const float h = ac[PC_NI + k];
            de[k] = a * (1.f - h * h);
        }
#pragma unroll
        for (int o = 0; o < PC_NO; ++o) de[3 * PC_NH + o] = d4[o];
    }
    red[threadIdx.x] = loss / manz;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int t = 0; t < (int)blockDim.x; ++t) s += red[t];
        part_loss[blockIdx.x] = s;
    }
}

// gradient partials: one parameter per thread over one chunk of cells, in order
#define PC_NCH 64
__global__ void __launch_bounds__(256) crt_pc_wgrad_kernel(int manz, const float *act, const float *delta,
                                                           double *gpart) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= PC_NP) return;
    const int ch = blockIdx.y, per = (manz + PC_NCH - 1) / PC_NCH;
    const int j0 = ch * per, j1 = min(manz, j0 + per);
    const int NA = PC_NI + 3 * PC_NH, ND = 3 * PC_NH + PC_NO;
    int di, ai;  // delta index and activation index (-1: bias)
    if (p < PC_B1) { di = (p - PC_W1) / PC_NI; ai = (p - PC_W1) % PC_NI; }
    else if (p < PC_W2) { di = p - PC_B1; ai = -1; }
    else if (p < PC_B2) { di = PC_NH + (p - PC_W2) / PC_NH; ai = PC_NI + (p - PC_W2) % PC_NH; }
    else if (p < PC_W3) { di = PC_NH + (p - PC_B2); ai = -1; }
    else if (p < PC_B3) { di = 2 * PC_NH + (p - PC_W3) / PC_NH; ai = PC_NI + PC_NH + (p - PC_W3) % PC_NH; }
    else if (p < PC_W4) { di = 2 * PC_NH + (p - PC_B3); ai = -1; }
    else if (p < PC_B4) { di = 3 * PC_NH + (p - PC_W4) / PC_NH; ai = PC_NI + 2 * PC_NH + (p - PC_W4) % PC_NH; }
    else { di = 3 * PC_NH + (p - PC_B4); ai = -1; }
    double g = 0.0;
    for (int j = j0; j < j1; ++j) {
        const float d = delta[(size_t)j * ND + di];
        g += ai < 0 ? (double)d : (double)d * act[(size_t)j * NA + ai];
    }
    gpart[(size_t)ch * PC_NP + p] = g;
}

__global__ void __launch_bounds__(256) crt_pc_gsum_kernel(const double *gpart, float *grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= PC_NP) return;
    double g = 0.0;
    for (int ch = 0; ch < PC_NCH; ++ch) g += gpart[(size_t)ch * PC_NP + p];
    grad[p] = (float)g;
}

__global__ void __launch_bounds__(256) crt_pc_adam_kernel(float *params, float *m, float *v, const float *grad,
                                                          float lr, float b1, float b2, float bc1, float bc2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= PC_NP) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * g, vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
}

// ---------------------------------------------------------------- inference
// G rows from the network (BF16 weights, FP32 accumulation), in the FSAI layout;
// prior_only: the analytic estimate alone (the network's output taken as zero)
__global__ void __launch_bounds__(64) crt_pc_infer_kernel(int manz, int nx, const float *params, const float *norm,
                                                           const float *feat, const int *pat, zd *g, int prior_only) {
    __shared__ float P[PC_NP];
    if (!prior_only)
        for (int t = threadIdx.x; t < PC_NP; t += blockDim.x) P[t] = bf16_round(params[t]);
    __syncthreads();
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= manz) return;
    float x[PC_NI], h1[PC_NH], h2[PC_NH], y[PC_NO] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if (!prior_only) {
#pragma unroll
        for (int k = 0; k < PC_NI; ++k) x[k] = (feat[(size_t)j * PC_NI + k] - norm[k]) * norm[PC_NI + k];
        pc_layer<PC_NH, PC_NI, true>(P + PC_W1, P + PC_B1, x, h1);
        pc_layer<PC_NH, PC_NH, true>(P + PC_W2, P + PC_B2, h1, h2);
        pc_layer<PC_NH, PC_NH, true>(P + PC_W3, P + PC_B3, h2, h1);
        pc_layer<PC_NO, PC_NH, false>(P + PC_W4, P + PC_B4, h1, y);
    }
    float b[PC_NO];
    pc_baseline(feat + (size_t)j * PC_NI, b);
#pragma unroll
    for (int o = 0; o < PC_NO; ++o) y[o] += b[o];
    constexpr int NP = CRT_FSAI_NP;
    for (int s = 0; s < NP; ++s) {
        const int c = pat[(size_t)j * NP + s];
        zd v = zd{0.0, 0.0};
        if (c == j) v = zd{(double)expf(y[4]), 0.0};
        else if (c >= 0 && c == j - nx) v = zd{(double)y[0], (double)y[1]};
        else if (c >= 0 && c == j - 1) v = zd{(double)y[2], (double)y[3]};
        g[(size_t)j * NP + s] = v;
    }
}

// ------------------------------------------------------------------ launchers
int crt_launch_pc_features(hipStream_t s, const PcCell &a, zd *cpl, float *feat) {
    hipLaunchKernelGGL(crt_pc_coupling_kernel, dim3((a.manz + 3) / 4), dim3(256), 0, s, a, cpl);
    hipLaunchKernelGGL(crt_pc_features_kernel, dim3((a.manz + 255) / 256), dim3(256), 0, s, a, cpl, feat);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_pc_targets(hipStream_t s, int manz, int nx, const int *pat, const zd *g, const float *feat, float *tgt,
                          float *mask) {
    hipLaunchKernelGGL(crt_pc_targets_kernel, dim3((manz + 255) / 256), dim3(256), 0, s, manz, nx, pat, g, feat, tgt,
                       mask);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_pc_norm(hipStream_t s, int manz, const float *feat, float *norm) {
    hipLaunchKernelGGL(crt_pc_norm_kernel, dim3(1), dim3(256), 0, s, manz, feat, norm);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_pc_blocks(int manz) { return (manz + 63) / 64; }

size_t crt_pc_gpart_size() { return (size_t)PC_NCH * PC_NP; }

int crt_launch_pc_fwdbwd(hipStream_t s, int manz, const float *params, const float *norm, const float *feat,
                         const float *tgt, const float *mask, float *act, float *delta, double *part_loss,
                         double *gpart, float *grad) {
    hipLaunchKernelGGL(crt_pc_fwdbwd_kernel, dim3(crt_pc_blocks(manz)), dim3(64), 0, s, manz, params, norm, feat, tgt,
                       mask, act, delta, part_loss);
    hipLaunchKernelGGL(crt_pc_wgrad_kernel, dim3((PC_NP + 255) / 256, PC_NCH), dim3(256), 0, s, manz, act, delta,
                       gpart);
    hipLaunchKernelGGL(crt_pc_gsum_kernel, dim3((PC_NP + 255) / 256), dim3(256), 0, s, gpart, grad);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_pc_adam(hipStream_t s, float *params, float *m, float *v, const float *grad, float lr, int t) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - powf(b1, (float)t), bc2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(crt_pc_adam_kernel, dim3((PC_NP + 255) / 256), dim3(256), 0, s, params, m, v, grad, lr, b1, b2,
                       bc1, bc2);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}

int crt_launch_pc_infer(hipStream_t s, int manz, int nx, const float *params, const float *norm, const float *feat,
                        const int *pat, zd *g, int prior_only) {
    hipLaunchKernelGGL(crt_pc_infer_kernel, dim3((manz + 63) / 64), dim3(64), 0, s, manz, nx, params, norm, feat, pat,
                       g, prior_only);
    return hipGetLastError() == hipSuccess ? 0 : CRT_ERR_HIP;
}
