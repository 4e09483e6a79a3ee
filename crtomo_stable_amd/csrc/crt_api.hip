// crt_api.hip -- host side of the CRTomo HIP library: context, device-resident
// state, assembly map, wavenumber bookkeeping, CutMcK and the C-ABI.
#include "crt_common.hpp"
#include "../../include/crtomo_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

// kernel launchers (crt_precal.hip, crt_band.hip, crt_volt_sens.hip)
int crt_launch_precal(hipStream_t, int, int, int, int, int, int, const int *, const int *, const int *,
                      const double *, const double *, const int *, const double *, double, double *, double *,
                      double *, int *);
int crt_launch_diag(hipStream_t, const DevGrid &, int, const int2 *, double *, double *, int *);
int crt_launch_dc_pivot(hipStream_t, int, int, int, const zd *, int *);
int crt_launch_zero_sources(hipStream_t, int, int, int, int, const int2 *, zd *);
int crt_launch_dc_pivot32(hipStream_t, int, int, int, const zf *, int *);
int crt_launch_assemble(hipStream_t, const DevGrid &, int, const int2 *, const double *, const double *, zd *);
int crt_launch_area(hipStream_t, const DevGrid &, int, zd *);
int crt_launch_factor(hipStream_t, const DevGrid &, int, const int2 *, const double *, const double *, const zd *,
                      zd *, zd *, int *);
size_t crt_factor_lds_bytes(int D);
int crt_launch_solve(hipStream_t, int, int, int, int, int, int, const int4 *, const int *, const double *,
                     const zd *, const zd *, zd *, zd *, const zd *, zd *);
int crt_launch_panel_band(hipStream_t, int, int, int, int, int, zd *, zd *);
int crt_launch_factor_band(hipStream_t, const DevGrid &, const int2 *, const zd *, zd *, zd *, int *);
int crt_factor16_supported(int mb);
int crt_factor16_resident(int D, int variant);
int crt_launch_factor16(hipStream_t, const DevGrid &, int, const int2 *, const double *, const zd *, zd *, zd *,
                        zd *, int *, int);
int crt_launch_solve16(hipStream_t, int, int, int, int, int, int, const int4 *, const int *, const double *,
                       const zd *, const zd *, const zd *, zd *, zd *, int);
int crt_factor16_resolve(int D, int variant);
// mixed precision (crt_band32.hip, FP32-panel solve modes in crt_band16.hip)
int crt_factor32_supported(int mb);
int crt_launch_factor32(hipStream_t, const DevGrid &, int, const int2 *, const double *, const zd *, zf *, zf *, zf *,
                        int *);
int crt_launch_refine_resid(hipStream_t, int, int, int, int, const int4 *, const int *, const int *, const int2 *,
                            const double *, const zd *, const zd *, zd *, const int *skip = nullptr);
int crt_launch_solve16_f32(hipStream_t, int, int, int, int, int, int, const int4 *, const int *, const double *,
                           const zf *, const zf *, const zf *, zd *, zd *, zd *, double *, int,
                           const int *skip = nullptr);
int crt_launch_refine_init(hipStream_t, int *, double *);
int crt_launch_refine_check(hipStream_t, int, const double *, double, double, int *, double *);
// general right-hand sides: singularity removal, analytic solution, fictitious sink
int crt_launch_refsig(hipStream_t, int, int, const int *, const int *, const double *, const double *, const zd *,
                      zd *);
int crt_launch_potana(hipStream_t, int, int, const int4 *, const int *, const double *, const double *,
                      const double *, const zd *, zd *, size_t);
int crt_launch_gen_rhs(hipStream_t, const DevGrid &, int, const int4 *, int, const zd *, const zd *, size_t,
                       const int *, const int *, int, const double *, int, zd *);
int crt_launch_pota_kpot(hipStream_t, int, int, int, int, const int4 *, const zd *, size_t, int, zd *);
int crt_launch_solve16_rhs(hipStream_t, int, int, int, int, int, int, const int4 *, const int *, const double *,
                           const zd *, const zd *, zd *, zd *, int);
int crt_launch_bkfak(hipStream_t, int, const int4 *, const int *, const double *, const double *, double *, int *);
// forward by static condensation onto the mixed-boundary nodes (crt_condense.hip)
int crt_launch_cond_gather(hipStream_t, int, int, int, int, const int *, const zd *, zd *);
int crt_launch_cond_fb(hipStream_t, int, int, int, int, int, int, int, const int *, const zd *, double *, int *);
int crt_launch_cond_rhs(hipStream_t, int, int, int, int, int, int, int, const int *, const int *, const int *,
                        const zd *, const double *, const double *, zd *);
int crt_launch_cond_trsm(hipStream_t, int, int, int, int, int, int, int, const zd *, const zd *, zd *, const int *,
                         const int *, const int *, const zd *, int, const int *, int);
int crt_launch_cond_schur(hipStream_t, int, int, int, int, int, int, int, int, int, const int *, const int *,
                          const int *, const int *, const int *, const int *, const int *, const int *, const int *,
                          const zd *, const double *, const double *, const zd *, zd *, zd *);
int crt_cond_dense_max_tiles();
int crt_launch_cond_s0(hipStream_t, int, int, int, int, const zd *, const zd *, zd *, int);
int crt_launch_cond_lgs(hipStream_t, int, int, int, const zd *, zd *, const zd *, int);
int crt_cond_prefix_max_tiles();
int crt_launch_cond_bval(hipStream_t, const DevGrid &, int, int, int, int, int, const int *, const int2 *,
                         const double *, zd *);
int crt_launch_cond_dense(hipStream_t, const DevGrid &, int, int, int, int, int, int, int, int, int, const zd *,
                          const zd *, const zd *, const int *, const int *, const int *, const int *, const zd *, int,
                          const double *, zd *, zd *, zd *, zd *, zd *, zd *, zd *, int *, int);
int crt_launch_cond_gemm(hipStream_t, int, int, int, int, int, int, int, int, int, int, const zd *, const zd *,
                         const double *, const int *, const int *, const int *, const int *, const int *, const int *,
                         zd *, int);
// model uncertainty (crt_bmcm.hip)
int crt_bmcm_run(hipStream_t, int, int, const zd *, const double *, int, int, int, int, const double *, const int *, double,
                 int, double *, double *, double *, double *, double *, double *, int *, int *);
int crt_launch_hpot_elec(hipStream_t, int, int, const int *, const double *, const zd *, zd *);
int crt_launch_hpot_full(hipStream_t, int, int, int, const double *, const zd *, zd *);
int crt_launch_volt(hipStream_t, int, int, int, const int4 *, const zd *, const double *, zd *, zd *, int *);
int crt_launch_sens(hipStream_t, int, int, int, int, int, int, int, const int *, const int *, const int4 *,
                    const double *, const zd *, const double *, const zd *, const zd *, const double *, zd *, zd *);
int crt_launch_kpot_ref(hipStream_t, int, int, int, const zd *, zd *);

// CG (crt_cg.hip)
int crt_cg_geometry(int, int, int *, int *, int *, int *);
int crt_launch_cg_j32(hipStream_t, size_t, const zd *, float2 *);
int crt_launch_cg_jr32(hipStream_t, size_t, const double *, float *);
int crt_launch_cg_fsai(hipStream_t, const CGArgs &, zd *);
// learned preconditioner (crt_learned_pc.hip)
struct PcCell {
    int manz, nx, nz, nanz;
    double lam;
    const double *cgfac, *smatm, *cy;
    double ymin, ymax;
    const zd *J;
    const double *w;
};
int crt_pc_nparams();
int crt_pc_blocks(int);
size_t crt_pc_gpart_size();
int crt_launch_pc_features(hipStream_t, const PcCell &, zd *, float *);
int crt_launch_pc_targets(hipStream_t, int, int, const int *, const zd *, const float *, float *, float *);
int crt_launch_pc_norm(hipStream_t, int, const float *, float *);
int crt_launch_pc_fwdbwd(hipStream_t, int, const float *, const float *, const float *, const float *, const float *,
                         float *, float *, double *, double *, float *);
int crt_launch_pc_adam(hipStream_t, float *, float *, float *, const float *, float, int);
int crt_launch_pc_infer(hipStream_t, int, int, const float *, const float *, const float *, const int *, zd *, int);
int crt_launch_real_part(hipStream_t, size_t, const zd *, double *);
int crt_launch_cg_rhs(hipStream_t, int, int, double, const zd *, const double *, const double *, const zd *,
                      const zd *, const double *, double *, zd *);
int crt_launch_cg_rmul(hipStream_t, const CGArgs &, const zd *, zd *);
int crt_launch_cg_init(hipStream_t, const CGArgs &, const zd *);
int crt_launch_cg_iter(hipStream_t, const CGArgs &);
int crt_launch_cg_final(hipStream_t, int, double, const double *, zd *, zd *, int);
int crt_launch_bsmatmreg(hipStream_t, int, int, int, int, const int *, const double *, const double *, double, double,
                         double *);
int crt_launch_bsmatmtri(hipStream_t, int, int, const int *, const int *, const double *, const double *, const int *,
                         double, double, double *);
size_t crt_blam0_scratch(int, int, int *, int *);
int crt_launch_besp_pairs(hipStream_t, int, const double *, const double *, double *, int *);
int crt_launch_blam0(hipStream_t, int, int, int, const zd *, const double *, zd *, int, int, zd *, double *, double,
                     double *);

// ------------------------------------------------------- static CG scratch
// crt_cjg_noalloc keeps every buffer of its solve (weights, cgfac, right-hand side, the
// CG vectors, partial sums, scalar states, residual history) in this array instead of
// context buffers, so that the call allocates nothing.  Each GPU holds its own copy,
// from module load until the process ends; calls on one GPU share it.
constexpr size_t CRT_SCRATCH_BYTES = (size_t)64 << 20;
constexpr int CRT_SCRATCH_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[CRT_SCRATCH_BYTES];
static std::mutex scratch_mu;
static unsigned char *scratch_dev[CRT_SCRATCH_DEVICES];  // device address of `scratch` per GPU

// Called with `device` current (crt_create): where that GPU holds `scratch`.  The
// array's name is a host-side stand-in, never a device address: kernels and copies
// take the pointer looked up here.
static int scratch_lookup(int device) {
    if (device < 0 || device >= CRT_SCRATCH_DEVICES) return CRT_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_dev[device]) return 0;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CRT_ERR_HIP;
    scratch_dev[device] = static_cast<unsigned char *>(p);
    return 0;
}

static unsigned char *scratch_base(int device) {
    std::lock_guard<std::mutex> lk(scratch_mu);
    return device >= 0 && device < CRT_SCRATCH_DEVICES ? scratch_dev[device] : nullptr;
}

// w = wmatd * wdfak on the device (the weights of the normal equations)
__global__ void __launch_bounds__(256) crt_noalloc_weights_kernel(int n, const int *__restrict__ wdfak,
                                                                  double *__restrict__ w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = w[i] * (double)wdfak[i];
}

namespace {

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int ensure(size_t count) {
        if (count <= n && p) return 0;
        release();
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) { p = nullptr; return CRT_ERR_ALLOC; }
        n = count;
        return 0;
    }
    int upload(const T *h, size_t count, hipStream_t s) {
        int e = ensure(count);
        if (e) return e;
        if (count && hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) return CRT_ERR_HIP;
        return 0;
    }
};

enum { T_DIAG = 0, T_FACTOR, T_SOLVE, T_VOLT, T_SENS, T_PRECAL, T_CG, T_CGMV, T_REFINE, T_SCHUR, T_PRODUCT, T_PREFIX,
       T_N };

}  // namespace

struct crt_ctx {
    int device = 0;
    // the single-system drop-ins (crt_chol, crt_vre and the twins built on them) share
    // the workspace below and the stream; the reference calls chol/vre from its OpenMP
    // loop over wavenumbers (src/fem.f90:328-380), so these calls take this lock and
    // may come from several host threads at once
    std::recursive_mutex dropin_mu;
    hipStream_t stream = nullptr;
    hipEvent_t ev0[T_N], ev1[T_N];
    double times[T_N] = {0};
    // grid
    bool have_grid = false, have_survey = false, have_kwn = false, have_precal = false, have_sigma = false,
         have_kpot = false, have_volt = false, have_sens = false;
    int sanz = 0, mb = 0, typanz = 0, maxn = 0, elanz = 0, relanz = 0, eanz = 0, npk = 0, nblk = 0, D = 0;
    std::vector<int> typ, nelanz, selanz, nrel, rnr, enr0, etyp, btyp, nkel;
    std::vector<double> x, y;
    int nanz = 0;
    std::vector<int> strnr, vnr;
    int kwnanz = 0;
    std::vector<double> kwn, kwnwi;
    int lbeta = 0;
    // Dirichlet boundaries: nodes of type-13 elements (lrandb, src/randb.f90) and of the
    // boundary values of boundary.dat (lrandb2, src/randb2.f90; the device runs zero
    // values only); dmask holds the per-entry identity rows and columns they make
    bool has13 = false;
    std::vector<unsigned char> dnode13;
    std::vector<int> bnd_nodes;
    bool bnd_zero = true;
    std::vector<int> h_col, h_crow;  // CSR column and row of every entry (host copy)
    DBuf<unsigned char> d_dmask;
    bool dmask_ok = false, dmask_any = false;
    std::vector<unsigned char> dnode;  // every Dirichlet node of the last dmask build
    // device state
    DBuf<int> d_nrel, d_etyp, d_btyp, d_rnr, d_enr, d_row_ptr, d_col, d_crow, d_cptr, d_contrib, d_nkel, d_status;
    DBuf<int> d_tmap;
    int tmapD = 0;
    // full rows of the symmetric matrix over the lower-triangle CSR: row pointers and
    // (column, CSR index) per entry, for the refinement residual
    DBuf<int> d_frow;
    DBuf<int2> d_fent;
    // area contributions before boundary ones in every entry's list: the area part is
    // summed once per wavenumber (d_area) and shared by every source
    bool asplit = false;
    DBuf<int> d_csplit;
    DBuf<zd> d_area;
    // forward precision: 64 (FP64 factorisation) or 32 (FP32 factorisation + FP64
    // iterative refinement until max|fak d| <= refine_tol max|fak x| per system)
    int fwd_bits = 64;
    double refine_tol = 1e-14;     // stop once the correction is this small
    double refine_accept = 1e-12;  // ... or once it stalls at most this large
    int refine_max = 12;
    int refine_sweeps = 0;      // most sweeps any batch of the last forward needed
    double refine_ratio = 0.0;  // largest final correction ratio of the last forward
    int refine_fallbacks = 0;   // batches of the last forward redone in FP64
    DBuf<zd> d_xacc;
    DBuf<double> d_rstat;
    int nnz = 0;
    DBuf<int4> d_meas, d_rhs;
    DBuf<int2> d_sys;
    DBuf<double> d_x, d_y, d_kwn, d_kwnwi, d_elbg, d_relbg, d_kg, d_fak, d_dga, d_kfak, d_strom;
    DBuf<zd> d_sigma, d_kpotN, d_panels, d_yf, d_dinv, d_yscr, d_hpotE, d_volt, d_sigmaa, d_sens, d_tmp, d_vals;
    // normal-equation state
    int ltri = -1, rnx = 0, rncol = 0, sens_nanz = 0, sens_manz = 0;
    int reg_manz = 0;  // cells of the installed regularisation (smatm, rdiag, nachbar rows)
    DBuf<double> d_smatm, d_rdiag, d_w, d_cgfac, d_part_dr1, d_part_rr, d_cgres;
    DBuf<int> d_nachbar;
    DBuf<double> d_jtj;
    DBuf<float2> d_sens32;  // complex FP32 copy of J for the CG iterations
    int cg_jbits = 64;      // storage precision of J in the CG iterations
    bool sens32_ok = false;
    // 0: complex inversion; 1: DC inversion (ldc); 2: final phase improvement (lfpi)
    int inv_kind = 0;
    DBuf<double> d_sensr;   // real J of the DC / FPI CG (sensdc, or DBLE(sens))
    bool sensr_ok = false;
    DBuf<float> d_sensr32;  // its FP32 copy for the iterations
    bool sensr32_ok = false;
    DBuf<double> d_re;      // real parts for the DC getters
    int cg_prec = 0;                      // 0 Jacobi scaling only, 1/2 FSAI patterns
    std::vector<int> h_nachbar;           // host copy of the neighbour table (ltri 1/2)
    DBuf<int> d_fpat, d_ftp;
    DBuf<zd> d_fg, d_t, d_z;
    DBuf<double> d_part_rz;
    // learned preconditioner: parameters, Adam moments, feature standardisation
    DBuf<float> d_pc_par, d_pc_m, d_pc_v, d_pc_norm, d_pc_feat, d_pc_tgt, d_pc_mask, d_pc_act, d_pc_delta, d_pc_grad;
    DBuf<double> d_pc_loss, d_pc_gpart, d_cy;
    DBuf<zd> d_pc_cpl;
    DBuf<float> d_pc_tfeat, d_pc_ttgt, d_pc_tmask;  // training set: every system trained on so far
    int pc_tn = 0;
    int pc_steps = 0, pc_manz = 0, pc_fallbacks = 0;
    bool pc_ready = false, pc_norm_set = false;
    double cy_min = 0.0, cy_max = 0.0;
    // the last CG system (for training on it)
    double last_lam = 0.0;
    bool have_cgsys = false;
    DBuf<zd> d_res, d_rm, d_bvec, d_r, d_p, d_q, d_bv, d_dpar, d_ap, d_part_ap, d_par;
    DBuf<CGScalars> d_sc;
    int cg_check_every = 16;
    size_t panel_budget = (size_t)96 << 30;  // bytes of factor panels kept in flight
    // forward method: -1 condensed where it applies, else the band factorisation of
    // every system; 0 always the band factorisation (the reference's algorithm);
    // 1 condensed (903 where it does not apply)
    int fwd_method = -1;
    // crmod.cfg's forward switches (src/fem.f90:177-203): singularity removal, the
    // analytic solution alone, the fictitious sink (0-based node, -1 none)
    int sw_lsr = 0, sw_lana = 0, sw_nsink = -1;
    DBuf<zd> d_pota, d_sigma0;
    DBuf<zd> d_vfac;          // per-measurement factor of the Jacobian's lookups
    DBuf<int> d_tptr, d_tq;   // per node the strictly lower CSR entries of its column
    bool tcol_built = false;
    int last_method = 0;
    // static condensation plan of the grid (boundary nodes G, interior nodes I)
    struct Cond {
        bool built = false, ok = false;
        bool tuned = false;                           // pieces chosen by timing (first forward)
        int force_np = 0;                             // pieces the plan starts from (0: 8)
        int P = 1;                                    // interior band pieces
        int nI = 0, nG = 0, nS = 0, nSp = 0, nGp = 0, nBp = 0, JS = 0, nt = 0, ntB = 0;
        int mbI = 0, DI = 0, nblkP = 0, nnzI = 0, E = 0, ncolP = 0, ntileP = 0, nent = 0, order = 0, nb = 0;
        std::vector<int> nodeI, nodeB, roff, plen, nbc;
        DBuf<int> d_row_ptr, d_col, d_crow, d_zero, d_tmap, d_qmap, d_enr0;
        DBuf<int> d_tptr, d_erow, d_erowg, d_ecol, d_ebcol, d_eq, d_j0;
        DBuf<int> d_gi_ptr, d_gi_row, d_gi_piece, d_gi_q, d_gmap, d_qdiag, d_qdiagG, d_nodeI, d_nodeB, d_eB;
        DBuf<int> d_bcol, d_ucol, d_colmap, d_roff, d_plen, d_nbc, d_bptr, d_bent, d_bq;
        DBuf<int2> d_bg;
        DBuf<zd> areaI, vals, panels, yI, dinv, X, T, S0, Gr, XB, Lt, Li, Lpre, Lipre, rv, bval;
        DBuf<double> fak, dga, fB;
        DBuf<int2> sysI;
        DBuf<int> stI, stD;
    } cond;

    DevGrid devgrid() const {
        DevGrid g;
        g.sanz = sanz; g.mb = mb; g.elanz = elanz; g.relanz = relanz; g.eanz = eanz; g.maxn = maxn;
        g.npk = npk; g.kwnanz = kwnanz; g.nanz = nanz; g.nblk = nblk; g.D = D; g.nnz = nnz; g.vstride = nnz + 2;
        g.fstride = nblk * CRT_NB;
        g.row_ptr = d_row_ptr.p; g.col = d_col.p; g.crow = d_crow.p; g.cptr = d_cptr.p; g.contrib = d_contrib.p;
        g.rnr = d_rnr.p; g.enr = d_enr.p; g.elbg = d_elbg.p; g.relbg = d_relbg.p; g.kg = d_kg.p;
        g.sigma = d_sigma.p;
        g.tmap = d_tmap.p;
        g.tmapD = tmapD;
        g.csplit = asplit ? d_csplit.p : nullptr;
        g.area = nullptr;  // set by the callers that computed the area sums
        g.dc = inv_kind == 1 ? 1 : 0;
        g.dmask = dmask_any ? d_dmask.p : nullptr;
        return g;
    }
    void tic(int i) { (void)hipEventRecord(ev0[i], stream); }
    void toc(int i) { (void)hipEventRecord(ev1[i], stream); }
    void collect(int i) {
        float ms = 0.f;
        if (hipEventSynchronize(ev1[i]) == hipSuccess && hipEventElapsedTime(&ms, ev0[i], ev1[i]) == hipSuccess)
            times[i] = ms;
    }
    // device -> host copy ordered after the work queued on the context stream (which
    // does not synchronise with the null stream)
    hipError_t copy_out(void *dst, const void *src, size_t bytes) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream);
        return e == hipSuccess ? hipStreamSynchronize(stream) : e;
    }
    // n complex values to the host; the DC inversion hands out their real parts
    // (kpotdc, hpotdc, sensdc are REAL*8 in src/alloci.f90:24-28)
    int copy_out_kind(double *dst, const zd *src, size_t n) {
        if (inv_kind != 1) return copy_out(dst, src, n * sizeof(zd)) == hipSuccess ? 0 : CRT_ERR_HIP;
        int e = d_re.ensure(n);
        e = e ? e : crt_launch_real_part(stream, n, src, d_re.p);
        if (e) return e;
        return copy_out(dst, d_re.p, n * sizeof(double)) == hipSuccess ? 0 : CRT_ERR_HIP;
    }
    int status_check() {
        int st = 0;
        if (copy_out(&st, d_status.p, sizeof(int)) != hipSuccess) return CRT_ERR_HIP;
        return st;
    }
    int status_reset() {
        int e = d_status.ensure(1);
        if (e) return e;
        return hipMemsetAsync(d_status.p, 0, sizeof(int), stream) == hipSuccess ? 0 : CRT_ERR_HIP;
    }
};

// ------------------------------------------------------------------ helpers
namespace {

// host restatement of the wavenumber set (src/rwaven.f90:52-151)
void gauleg_h(double x1, double x2, double *xo, double *w, int n) {
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; ++i) {
        double z = std::cos(3.141592653589793 * (i - 0.25) / (n + 0.5)), z1, pp;
        do {
            double p1 = 1.0, p2 = 0.0, p3;
            for (int j = 1; j <= n; ++j) {
                p3 = p2; p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (std::fabs(z - z1) > 3e-14);
        xo[i - 1] = xm - xl * z;
        xo[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}
double gammln_h(double xx) {
    static const double cof[6] = {76.18009172947146, -86.50532032941677, 24.01409824083091,
                                  -1.231739572450155, .1208650973866179e-2, -.5395239384953e-5};
    double x = xx, yv = x, tmp = x + 5.5, ser = 1.000000000190015;
    tmp = (x + 0.5) * std::log(tmp) - tmp;
    for (int j = 0; j < 6; ++j) { yv += 1.0; ser += cof[j] / yv; }
    return tmp + std::log(2.5066282746310005 * ser / x);
}
void gaulag_h(double *xo, double *w, int n, double alf) {
    double z = 0, z1, pp = 0, p1, p2 = 0, p3;
    for (int i = 1; i <= n; ++i) {
        if (i == 1) z = (1.0 + alf) * (3.0 + 0.92 * alf) / (1.0 + 2.4 * n + 1.8 * alf);
        else if (i == 2) z += (15.0 + 6.25 * alf) / (1.0 + 0.9 * alf + 2.5 * n);
        else {
            double ai = i - 2;
            z += ((1.0 + 2.55 * ai) / (1.9 * ai) + 1.26 * ai * alf / (1.0 + 3.5 * ai)) * (z - xo[i - 3]) /
                 (1.0 + 0.3 * alf);
        }
        for (int its = 1; its <= 10; ++its) {
            p1 = 1.0; p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                p3 = p2; p2 = p1;
                p1 = (((double)(2 * j - 1) + alf - z) * p2 - ((double)(j - 1) + alf) * p3) / (double)j;
            }
            pp = (n * p1 - (n + alf) * p2) / z;
            z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs(z - z1) <= 3e-14) break;
        }
        xo[i - 1] = z;
        w[i - 1] = -std::exp(gammln_h(alf + n) - gammln_h((double)n)) / (pp * n * p2);
    }
}

// mean height of the no-flow boundary (src/bsytop.f90:23-66)
double bsytop_h(const crt_ctx *c) {
    double sytop = 0.0;
    int iel = 0;
    for (int t = 0; t < c->typanz; ++t) {
        if (c->typ[t] == 12) {
            int nk = c->selanz[t];
            sytop = 0.0;
            for (int j = 1; j <= c->nelanz[t]; ++j, ++iel) {
                double sp = 0.0;
                for (int ik = 0; ik < nk; ++ik) sp += c->y[c->nrel[(size_t)iel * c->maxn + ik] - 1];
                sp = sp / (double)nk;
                sytop = (j == 1) ? sp : (sytop * (double)(j - 1) + sp) / (double)j;
            }
        } else {
            iel += c->nelanz[t];
        }
    }
    return sytop;
}

}  // namespace

// ------------------------------------------------------------------ batches
// Forward of many models (contexts) as one graph: every model's systems are
// assembled, factorised and solved in chunks through one shared workspace, all
// launches captured once into a hipGraph and replayed by crt_batch_forward.
struct crt_batch {
    struct Chunk {
        size_t s0;  // first system
        int nb;     // systems
        size_t r0;  // first right-hand side in the concatenated per-chunk lists
        int nr;     // right-hand sides
    };
    struct Model {
        crt_ctx *c = nullptr;
        bool use16 = true;
        // mixed precision (flags bit 1): FP32 factorisation on the matrix cores, refine_max
        // FP64 refinement sweeps per chunk in the graph, each skipped once the chunk's
        // device check has accepted it; cstate/crat per chunk (crt_refine_check_kernel)
        bool mixed = false;
        DBuf<int> cstate;
        DBuf<double> crat;
        size_t max_nr = 0;
        int fallback = 0;  // the last launch redid this model in FP64
        int nblk = 0, D = 0, NBt = 16, variant = 0;
        size_t fstride = 0, panel_elems = 0, nsys = 0, nrhs = 0, batch = 0, nyscr = 0;
        std::vector<Chunk> chunks;
        DBuf<int2> sys;
        DBuf<int4> rhs;
        DBuf<int> status;  // 1 + nsys
        int slot = 0;      // workspace slot and stream
    };
    // One workspace slot per stream: model i runs on slot i % nslot, so the partly
    // filled last round of one model's launches overlaps the next model's on another
    // stream.  Slot 0 uses the batch's own stream.
    struct Slot {
        hipStream_t st = nullptr;
        hipEvent_t join = nullptr;
        zd *panels = nullptr, *yf = nullptr, *dinv = nullptr, *vals = nullptr, *yscr = nullptr;
        double *fak = nullptr, *dga = nullptr;
        zd *xacc = nullptr;        // mixed: FP64 potentials per right-hand side of a chunk
        double *rstat = nullptr;   // mixed: |fak d|, |fak x| per right-hand side
    };
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, fork = nullptr;
    std::vector<std::unique_ptr<Model>> m;
    std::vector<Slot> slots;
    DBuf<zd> panels, yf, dinv, vals, yscr, xacc;
    DBuf<double> fak, dga, rstat;
    int mixed_sweeps = 0, mixed_fallbacks = 0;  // last launch: most sweeps of a chunk, models redone
    double mixed_ratio = 0.0;                   // last launch: largest accepted ratio
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float ms = 0.f;
    int variant[17] = {0};  // factor kernel variant per band width in tiles (0: default)
    float tune_ms[17][4] = {};
    size_t nsys_total = 0;
    ~crt_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (fork) (void)hipEventDestroy(fork);
        for (size_t i = 0; i < slots.size(); ++i) {
            if (slots[i].join) (void)hipEventDestroy(slots[i].join);
            if (i > 0 && slots[i].st) (void)hipStreamDestroy(slots[i].st);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// systems and right-hand sides of one model as crt_forward orders them, split into
// chunks of at most `cap` systems
int batch_plan(crt_batch::Model &M, size_t ws_bytes) {
    crt_ctx *c = M.c;
    M.use16 = crt_factor16_supported(c->mb) && !getenv("CRT_FACTOR8");
    M.NBt = M.use16 ? 16 : CRT_NB;
    const int K = c->kwnanz, E = c->eanz;
    M.nblk = (c->sanz + M.NBt - 1) / M.NBt;
    M.D = std::max(1, (c->mb + M.NBt - 1) / M.NBt);
    if (!M.use16 && crt_factor_lds_bytes(M.D) > 160 * 1024) return CRT_ERR_UNSUPPORTED;
    std::vector<int2> sys;
    std::vector<int4> rhs;
    for (int k = 0; k < K; ++k) {
        if (c->lbeta) {
            for (int l = 0; l < E; ++l) {
                rhs.push_back(int4{l, k, (int)sys.size(), 1});
                sys.push_back(int2{l, k});
            }
        } else {
            int fs = (int)sys.size();
            sys.push_back(int2{0, k});
            for (int l = 0; l < E; ++l) rhs.push_back(int4{l, k, fs, l == 0 ? 1 : 0});
        }
    }
    M.nsys = sys.size();
    M.nrhs = rhs.size();
    M.fstride = (size_t)M.nblk * M.NBt;
    M.panel_elems = (size_t)M.nblk * (M.D + 1) * M.NBt * M.NBt;
    const size_t per_sys = M.panel_elems * sizeof(zd) + M.fstride * (2 * sizeof(zd) + 2 * sizeof(double)) +
                           ((size_t)c->nnz + 2) * sizeof(zd);
    M.batch = std::max<size_t>(1, std::min(M.nsys, ws_bytes / per_sys));
    if (M.use16 && M.batch < M.nsys) {
        // whole multiples of the resident workgroup count (see crt_forward)
        const size_t res = (size_t)crt_factor16_resident(M.D, 0);
        if (res > 0 && M.batch >= res) M.batch = M.batch / res * res;
    }
    M.nyscr = c->lbeta ? 1 : M.nrhs;
    if (M.mixed) {
        if (!M.use16 || !crt_factor32_supported(c->mb) || c->tmapD != M.D) return CRT_ERR_UNSUPPORTED;
    }
    std::vector<int4> rall;
    M.chunks.clear();
    for (size_t s0 = 0; s0 < M.nsys; s0 += M.batch) {
        const int nb = (int)std::min(M.batch, M.nsys - s0);
        crt_batch::Chunk ch{s0, nb, rall.size(), 0};
        for (auto q : rhs)
            if ((size_t)q.z >= s0 && (size_t)q.z < s0 + nb) rall.push_back(int4{q.x, q.y, q.z - (int)s0, q.w});
        ch.nr = (int)(rall.size() - ch.r0);
        M.max_nr = std::max(M.max_nr, (size_t)ch.nr);
        M.chunks.push_back(ch);
    }
    if (M.mixed) {
        M.nyscr = std::max(M.nyscr, M.max_nr);
        int e = M.cstate.ensure(2 * M.chunks.size());
        e = e ? e : M.crat.ensure(2 * M.chunks.size());
        if (e) return e;
    }
    int e = M.sys.upload(sys.data(), sys.size(), c->stream);
    e = e ? e : M.rhs.upload(rall.data(), rall.size(), c->stream);
    e = e ? e : M.status.ensure(M.nsys + 1);
    e = e ? e : c->d_kpotN.ensure((size_t)c->sanz * K * E);
    e = e ? e : (c->asplit ? c->d_area.ensure((size_t)K * c->nnz) : 0);
    if (e) return e;
    return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : CRT_ERR_HIP;
}

DevGrid batch_grid(const crt_batch::Model &M) {
    DevGrid g = M.c->devgrid();
    g.nblk = M.nblk;
    g.D = M.D;
    g.fstride = (int)M.fstride;
    if (M.c->asplit) g.area = M.c->d_area.p;
    return g;
}

// the model's per-wavenumber area sums (before its chunks: sigma may change between calls)
int batch_area(crt_batch *b, crt_batch::Model &M) {
    if (!M.c->asplit) return 0;
    DevGrid g = batch_grid(M);
    g.area = nullptr;
    return crt_launch_area(b->slots[M.slot].st, g, M.c->kwnanz, M.c->d_area.p);
}

// diagonal scaling and assembly of one chunk into the model's workspace slot
int batch_assemble(crt_batch *b, crt_batch::Model &M, const crt_batch::Chunk &ch) {
    const DevGrid g = batch_grid(M);
    const crt_batch::Slot &w = b->slots[M.slot];
    int e = crt_launch_diag(w.st, g, ch.nb, M.sys.p + ch.s0, w.fak, w.dga, M.status.p + 1 + ch.s0);
    return e ? e : crt_launch_assemble(w.st, g, ch.nb, M.sys.p + ch.s0, w.fak, w.dga, w.vals);
}

int batch_factor(crt_batch *b, crt_batch::Model &M, const crt_batch::Chunk &ch, int variant) {
    const DevGrid g = batch_grid(M);
    const crt_batch::Slot &w = b->slots[M.slot];
    if (M.mixed) {
        int e = crt_launch_factor32(w.st, g, ch.nb, M.sys.p + ch.s0, w.fak, w.vals, reinterpret_cast<zf *>(w.panels),
                                    reinterpret_cast<zf *>(w.yf), reinterpret_cast<zf *>(w.dinv),
                                    M.status.p + 1 + ch.s0);
        if (!e && g.dc)
            e = crt_launch_dc_pivot32(w.st, ch.nb, M.c->sanz, g.fstride, reinterpret_cast<zf *>(w.dinv),
                                      M.status.p + 1 + ch.s0);
        return e;
    }
    if (M.use16) {
        int e = crt_launch_factor16(w.st, g, ch.nb, M.sys.p + ch.s0, w.fak, w.vals, w.panels, w.yf, w.dinv,
                                    M.status.p + 1 + ch.s0, variant);
        if (!e && g.dc) e = crt_launch_dc_pivot(w.st, ch.nb, M.c->sanz, g.fstride, w.dinv, M.status.p + 1 + ch.s0);
        return e;
    }
    return crt_launch_factor(w.st, g, ch.nb, M.sys.p + ch.s0, w.fak, w.dga, w.vals, w.panels, w.yf,
                             M.status.p + 1 + ch.s0);
}

int batch_solve(crt_batch *b, crt_batch::Model &M, const crt_batch::Chunk &ch) {
    crt_ctx *c = M.c;
    const int K = c->kwnanz, E = c->eanz;
    const crt_batch::Slot &w = b->slots[M.slot];
    if (M.mixed) {
        // FP32 substitution, then refine_max sweeps (FP64 residual, FP32 correction), each
        // skipped once the chunk's check has accepted it
        const int ci = (int)(&ch - M.chunks.data());
        int *state = M.cstate.p + 2 * ci;
        double *rat = M.crat.p + 2 * ci;
        const zf *pan = reinterpret_cast<const zf *>(w.panels), *di = reinterpret_cast<const zf *>(w.dinv),
                 *y32 = reinterpret_cast<const zf *>(w.yf);
        const DevGrid g = batch_grid(M);
        int e = crt_launch_refine_init(w.st, state, rat);
        e = e ? e : crt_launch_solve16_f32(w.st, c->sanz, M.nblk, M.D, K, E, ch.nr, M.rhs.p + ch.r0, c->d_enr.p, w.fak,
                                           pan, di, y32, w.yscr, c->d_kpotN.p, w.xacc, w.rstat, 1);
        for (int sw = 0; sw < c->refine_max && !e; ++sw) {
            e = crt_launch_refine_resid(w.st, c->sanz, g.fstride, g.vstride, ch.nr, M.rhs.p + ch.r0, c->d_enr.p,
                                        c->d_frow.p, c->d_fent.p, w.fak, w.vals, w.xacc, w.yscr, state);
            e = e ? e : crt_launch_solve16_f32(w.st, c->sanz, M.nblk, M.D, K, E, ch.nr, M.rhs.p + ch.r0, c->d_enr.p,
                                               w.fak, pan, di, y32, w.yscr, c->d_kpotN.p, w.xacc, w.rstat, 2, state);
            e = e ? e : crt_launch_refine_check(w.st, ch.nr, w.rstat, c->refine_tol, c->refine_accept, state, rat);
        }
        return e;
    }
    if (M.use16)
        return crt_launch_solve16(w.st, c->sanz, M.nblk, M.D, K, E, ch.nr, M.rhs.p + ch.r0, c->d_enr.p, w.fak,
                                  w.panels, w.dinv, w.yf, w.yscr, c->d_kpotN.p,
                                  crt_factor16_resolve(M.D, M.variant) != 3);
    return crt_launch_solve(w.st, c->sanz, M.nblk, M.D, K, E, ch.nr, M.rhs.p + ch.r0, c->d_enr.p, w.fak, w.panels,
                            w.yf, w.yscr, c->d_kpotN.p, nullptr, nullptr);
}

// time each factor kernel variant that exists for band width D on the first chunk
// of the first model of that width; keep the fastest
int batch_tune(crt_batch *b) {
    for (int D = 1; D <= 16; ++D) {
        crt_batch::Model *M = nullptr;
        for (auto &q : b->m)
            if (q->use16 && q->D == D) { M = q.get(); break; }
        if (!M) continue;
        int cand[3], nc = 0;
        if (D >= 2 && D <= 10) cand[nc++] = 1;
        if (D >= 3) cand[nc++] = 2;
        if (D <= 10) cand[nc++] = 3;
        if (nc < 2) continue;
        const crt_batch::Chunk &ch = M->chunks[0];
        float best = 0.f;
        for (int i = 0; i < nc; ++i) {
            float t = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                int e = batch_assemble(b, *M, ch);
                if (e) return e;
                CRT_HIP(hipEventRecord(b->e0, b->slots[M->slot].st));
                e = batch_factor(b, *M, ch, cand[i]);
                if (e) return e;
                CRT_HIP(hipEventRecord(b->e1, b->slots[M->slot].st));
                CRT_HIP(hipEventSynchronize(b->e1));
                float ms = 0.f;
                CRT_HIP(hipEventElapsedTime(&ms, b->e0, b->e1));
                t = rep == 0 ? ms : std::min(t, ms);
            }
            b->tune_ms[D][cand[i]] = t;
            if (i == 0 || t < best) {
                best = t;
                b->variant[D] = cand[i];
            }
        }
    }
    return 0;
}

}  // namespace

// ------------------------------------------------------------------ C-ABI
extern "C" {

const char *crt_version(void) { return "crtomo_hip 0.1 (gfx950, FP64 band path)"; }

int crt_create(crt_ctx **out, int device) {
    if (!out) return CRT_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return CRT_ERR_HIP;
    CRT_HIP(hipSetDevice(device));
    crt_ctx *c = new crt_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return CRT_ERR_HIP; }
    for (int i = 0; i < T_N; ++i) {
        (void)hipEventCreate(&c->ev0[i]);
        (void)hipEventCreate(&c->ev1[i]);
    }
    if (const char *b = getenv("CRT_PANEL_BUDGET_GB")) c->panel_budget = (size_t)(atof(b) * (double)(1ull << 30));
    if (c->d_status.ensure(1)) { crt_destroy(c); return CRT_ERR_ALLOC; }
    if (int e = scratch_lookup(device)) { crt_destroy(c); return e; }
    *out = c;
    return 0;
}

void crt_destroy(crt_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < T_N; ++i) {
        (void)hipEventDestroy(c->ev0[i]);
        (void)hipEventDestroy(c->ev1[i]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int crt_set_grid(crt_ctx *c, int sanz, int mb, int typanz, const int *typ, const int *nelanz, const int *selanz,
                 int maxn, const int *nrel, const double *x, const double *y, const int *rnr, int eanz,
                 const int *enr) {
    if (!c || sanz <= 0 || mb < 0 || typanz <= 0 || maxn <= 0 || maxn > 4) return CRT_ERR_ARG;
    CRT_HIP(hipSetDevice(c->device));
    c->have_grid = c->have_precal = c->have_kpot = c->have_volt = c->have_sens = false;
    c->cond.built = c->cond.ok = c->cond.tuned = false;
    c->cond.force_np = 0;
    c->tcol_built = false;
    c->sw_lsr = c->sw_lana = 0;
    c->sw_nsink = -1;
    c->sanz = sanz; c->mb = mb; c->typanz = typanz; c->maxn = maxn; c->eanz = eanz;
    c->npk = maxn * (maxn + 1) / 2;
    c->typ.assign(typ, typ + typanz);
    c->nelanz.assign(nelanz, nelanz + typanz);
    c->selanz.assign(selanz, selanz + typanz);
    c->elanz = c->relanz = 0;
    bool boundary_seen = false;
    for (int t = 0; t < typanz; ++t) {
        if (typ[t] > 10) { c->relanz += nelanz[t]; boundary_seen = true; }
        else {
            if (boundary_seen) return CRT_ERR_ARG;  // area groups precede boundary groups
            if (typ[t] != 3 && typ[t] != 5 && typ[t] != 8) return CRT_ERR_TYPE;
            c->elanz += nelanz[t];
        }
    }
    const int neltot = c->elanz + c->relanz;
    c->nrel.assign(nrel, nrel + (size_t)neltot * maxn);
    c->x.assign(x, x + sanz);
    c->y.assign(y, y + sanz);
    c->rnr.assign(rnr, rnr + c->relanz);
    for (int i = 0; i < c->relanz; ++i)
        if (rnr[i] < 1 || rnr[i] > c->elanz) return CRT_ERR_ARG;
    c->enr0.resize(eanz);
    for (int i = 0; i < eanz; ++i) {
        if (enr[i] < 1 || enr[i] > sanz) return CRT_ERR_ARG;
        c->enr0[i] = enr[i] - 1;
    }
    c->etyp.resize(c->elanz);
    c->nkel.resize(c->elanz);
    c->btyp.resize(std::max(c->relanz, 1));
    {
        int iel = 0;
        for (int t = 0; t < typanz; ++t)
            for (int j = 0; j < nelanz[t]; ++j, ++iel) {
                if (typ[t] < 10) { c->etyp[iel] = typ[t]; c->nkel[iel] = selanz[t]; }
                else c->btyp[iel - c->elanz] = typ[t];
            }
    }
    // nodes of the type-13 (Dirichlet) elements
    c->has13 = false;
    c->dnode13.assign(sanz, 0);
    {
        int iel = 0;
        for (int t = 0; t < typanz; ++t)
            for (int j = 0; j < nelanz[t]; ++j, ++iel) {
                if (typ[t] != 13) continue;
                c->has13 = true;
                for (int m = 0; m < selanz[t]; ++m) {
                    const int nd = nrel[(size_t)iel * maxn + m];
                    if (nd < 1 || nd > sanz) return CRT_ERR_ARG;
                    c->dnode13[nd - 1] = 1;
                }
            }
    }
    c->bnd_nodes.clear();
    c->bnd_zero = true;
    c->dmask_ok = false;
    c->dmask_any = false;
    // assembly map: contributions to each lower-triangle entry in the element order
    // of src/kompab.f90:80-138 (area elements, then type-11 boundary elements)
    struct Ent { int i, j, code; };
    std::vector<Ent> ents;
    ents.reserve((size_t)neltot * 10 + sanz);
    {
        int iel = 0;
        for (int t = 0; t < typanz; ++t) {
            const int nk = selanz[t];
            for (int j = 0; j < nelanz[t]; ++j, ++iel) {
                if (typ[t] > 11) continue;
                int ikl = 0;
                for (int kk = 0; kk < nk; ++kk) {
                    int nzp = nrel[(size_t)iel * maxn + kk];
                    for (int ll = 0; ll <= kk; ++ll, ++ikl) {
                        int nnp = nrel[(size_t)iel * maxn + ll];
                        if (nzp < 1 || nnp < 1 || nzp > sanz || nnp > sanz) return CRT_ERR_ARG;
                        if (std::abs(nzp - nnp) > mb) return CRT_ERR_BAND;
                        int i = std::max(nzp, nnp) - 1, jj = std::min(nzp, nnp) - 1;
                        int code = (typ[t] == 11) ? -((iel - c->elanz) * c->npk + ikl) - 1 : iel * c->npk + ikl;
                        ents.push_back({i, jj, code});
                    }
                }
            }
        }
        for (int i = 0; i < sanz; ++i) ents.push_back({i, i, INT32_MIN});  // guarantees a diagonal entry
    }
    std::stable_sort(ents.begin(), ents.end(), [](const Ent &a, const Ent &b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    std::vector<int> row_ptr(sanz + 1, 0), col, crow, cptr, contrib;
    col.reserve(ents.size()); crow.reserve(ents.size()); cptr.reserve(ents.size() + 1); contrib.reserve(ents.size());
    for (size_t q = 0; q < ents.size();) {
        size_t q1 = q;
        while (q1 < ents.size() && ents[q1].i == ents[q].i && ents[q1].j == ents[q].j) ++q1;
        col.push_back(ents[q].j);
        crow.push_back(ents[q].i);
        cptr.push_back((int)contrib.size());
        for (size_t r = q; r < q1; ++r)
            if (ents[r].code != INT32_MIN) contrib.push_back(ents[r].code);
        row_ptr[ents[q].i + 1]++;
        q = q1;
    }
    cptr.push_back((int)contrib.size());
    for (int i = 0; i < sanz; ++i) row_ptr[i + 1] += row_ptr[i];
    // split point of each list: area contributions first, then boundary ones
    std::vector<int> csplit(col.size());
    c->asplit = true;
    for (size_t q = 0; q < col.size(); ++q) {
        int p = cptr[q];
        while (p < cptr[q + 1] && contrib[p] >= 0) ++p;
        csplit[q] = p;
        for (int r = p; r < cptr[q + 1]; ++r)
            if (contrib[r] >= 0) c->asplit = false;
    }
    if (csplit.empty()) csplit.push_back(0);
    if (contrib.empty()) contrib.push_back(0);
    c->h_col = col;
    c->h_crow = crow;

    c->nblk = (sanz + CRT_NB - 1) / CRT_NB;
    c->D = std::max(1, (mb + CRT_NB - 1) / CRT_NB);
    hipStream_t s = c->stream;
    std::vector<int> rnr0(std::max(c->relanz, 1), 0);
    for (int i = 0; i < c->relanz; ++i) rnr0[i] = rnr[i] - 1;
    int e = 0;
    e = e ? e : c->d_nrel.upload(c->nrel.data(), c->nrel.size(), s);
    e = e ? e : c->d_etyp.upload(c->etyp.data(), std::max((size_t)1, c->etyp.size()), s);
    e = e ? e : c->d_nkel.upload(c->nkel.data(), std::max((size_t)1, c->nkel.size()), s);
    e = e ? e : c->d_btyp.upload(c->btyp.data(), c->btyp.size(), s);
    e = e ? e : c->d_rnr.upload(rnr0.data(), rnr0.size(), s);
    e = e ? e : c->d_enr.upload(c->enr0.data(), std::max((size_t)1, c->enr0.size()), s);
    e = e ? e : c->d_x.upload(c->x.data(), sanz, s);
    e = e ? e : c->d_y.upload(c->y.data(), sanz, s);
    e = e ? e : c->d_row_ptr.upload(row_ptr.data(), row_ptr.size(), s);
    e = e ? e : c->d_col.upload(col.data(), col.size(), s);
    e = e ? e : c->d_crow.upload(crow.data(), crow.size(), s);
    c->nnz = (int)col.size();
    e = e ? e : c->d_cptr.upload(cptr.data(), cptr.size(), s);
    e = e ? e : c->d_contrib.upload(contrib.data(), contrib.size(), s);
    e = e ? e : c->d_csplit.upload(csplit.data(), csplit.size(), s);
    if (e) return e;
    // tile map of the 16x16 block band (used by the factorisation for band widths up to 160)
    c->tmapD = 0;
    {
        const int D16 = std::max(1, (mb + 15) / 16), nb16 = (sanz + 15) / 16;
        if (D16 <= 16) {
            const int nnz = (int)col.size();
            std::vector<int> tmap((size_t)nb16 * (D16 + 1) * 256, nnz);
            for (int i = 0; i < sanz; ++i)
                for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
                    const int j = col[q], I = i / 16, d = I - j / 16;
                    tmap[((size_t)(I * (D16 + 1) + d) * 16 + (i % 16)) * 16 + (j % 16)] = q;
                }
            for (int i = sanz; i < nb16 * 16; ++i)
                tmap[((size_t)((i / 16) * (D16 + 1)) * 16 + (i % 16)) * 16 + (i % 16)] = nnz + 1;
            if ((e = c->d_tmap.upload(tmap.data(), tmap.size(), s))) return e;
            c->tmapD = D16;
        }
    }
    {
        std::vector<int> frow(sanz + 1, 0);
        for (int i = 0; i < sanz; ++i)
            for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
                frow[i + 1]++;
                if (col[q] != i) frow[col[q] + 1]++;
            }
        for (int i = 0; i < sanz; ++i) frow[i + 1] += frow[i];
        std::vector<int2> fent(std::max(frow[sanz], 1));
        std::vector<int> fill(frow.begin(), frow.end() - 1);
        for (int i = 0; i < sanz; ++i)
            for (int q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
                const int j = col[q];
                fent[fill[i]++] = int2{j, q};
                if (j != i) fent[fill[j]++] = int2{i, q};
            }
        e = c->d_frow.upload(frow.data(), frow.size(), s);
        e = e ? e : c->d_fent.upload(fent.data(), fent.size(), s);
        if (e) return e;
    }
    CRT_HIP(hipStreamSynchronize(s));
    c->have_grid = true;
    return 0;
}

int crt_set_survey(crt_ctx *c, int nanz, const int *strnr, const int *vnr) {
    if (!c || nanz <= 0) return CRT_ERR_ARG;
    if (!c->have_grid) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    c->nanz = nanz;
    c->strnr.assign(strnr, strnr + nanz);
    c->vnr.assign(vnr, vnr + nanz);
    std::vector<int4> meas(nanz);
    for (int i = 0; i < nanz; ++i) {
        int e1 = strnr[i] % 10000, e2 = (strnr[i] - e1) / 10000;
        int e3 = vnr[i] % 10000, e4 = (vnr[i] - e3) / 10000;
        int ee[4] = {e1, e2, e3, e4};
        for (int q = 0; q < 4; ++q)
            if (ee[q] < 0 || ee[q] > c->eanz) return CRT_ERR_ARG;
        meas[i] = int4{e1, e2, e3, e4};
    }
    int e = c->d_meas.upload(meas.data(), nanz, c->stream);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->have_survey = true;
    c->have_volt = c->have_sens = false;
    return 0;
}

int crt_rwaven(crt_ctx *c, int *kwnanz, double *kwn, double *kwnwi, double *amin_o, double *amax_o) {
    if (!c) return CRT_ERR_ARG;
    if (!c->have_survey) return CRT_ERR_STATE;
    double amin = 1e9, amax = 0.0;
    for (int i = 0; i < c->nanz; ++i) {
        int e[4];
        double xe[4] = {0, 0, 0, 0}, ye[4] = {0, 0, 0, 0};
        e[0] = c->strnr[i] % 10000; e[1] = (c->strnr[i] - e[0]) / 10000;
        e[2] = c->vnr[i] % 10000;   e[3] = (c->vnr[i] - e[2]) / 10000;
        for (int j = 0; j < 4; ++j)
            if (e[j] > 0) { int nd = c->enr0[e[j] - 1]; xe[j] = c->x[nd]; ye[j] = c->y[nd]; }
        for (int m = 2; m < 4; ++m) {
            if (e[m] <= 0) continue;
            for (int j = 0; j < 2; ++j) {
                if (e[j] <= 0) continue;
                double x21 = xe[m] - xe[j], y21 = ye[m] - ye[j];
                double d = std::sqrt(x21 * x21 + y21 * y21);
                amin = std::fmin(d, amin);
                amax = std::fmax(d, amax);
            }
        }
    }
    if (amin_o) *amin_o = amin;
    if (amax_o) *amax_o = amax;
    if (amin < 1e-12 || amax < 1e-12) return CRT_ERR_ELEC;
    const int lanz = 4;
    int ganz = (int)(float)(6.0 * std::log10(amax / amin));  // INT(REAL(...)) of src/rwaven.f90:105
    ganz = std::min(std::max(ganz, 2), 30 - lanz);
    int n = ganz + lanz;
    std::vector<double> k(n), w(n);
    double k0 = 1.0 / (2.0 * amin);
    gauleg_h(0.0, 1.0, k.data(), w.data(), ganz);
    for (int i = 0; i < ganz; ++i) {
        w[i] = 2.0 * k0 * w[i] * k[i];
        k[i] = k0 * k[i] * k[i];
    }
    gaulag_h(k.data() + ganz, w.data() + ganz, lanz, 0.0);
    for (int i = ganz; i < n; ++i) {
        w[i] = k0 * w[i] * std::exp(k[i]);
        k[i] = k0 * (k[i] + 1.0);
    }
    if (kwnanz) *kwnanz = n;
    if (kwn) std::copy(k.begin(), k.end(), kwn);
    if (kwnwi) std::copy(w.begin(), w.end(), kwnwi);
    return crt_set_wavenumbers(c, n, k.data(), w.data());
}

int crt_set_wavenumbers(crt_ctx *c, int kwnanz, const double *kwn, const double *kwnwi) {
    if (!c || kwnanz <= 0 || kwnanz > 30) return CRT_ERR_ARG;
    CRT_HIP(hipSetDevice(c->device));
    c->kwnanz = kwnanz;
    c->kwn.assign(kwn, kwn + kwnanz);
    c->kwnwi.assign(kwnwi, kwnwi + kwnanz);
    int e = c->d_kwn.upload(c->kwn.data(), kwnanz, c->stream);
    e = e ? e : c->d_kwnwi.upload(c->kwnwi.data(), kwnanz, c->stream);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->have_kwn = true;
    c->have_precal = c->have_kpot = false;
    return 0;
}

int crt_set_boundary_values(crt_ctx *c, int n, const int *nodes, const double *values) {
    if (!c || n < 0 || (n > 0 && (!nodes || !values))) return CRT_ERR_ARG;
    if (!c->have_grid) return CRT_ERR_STATE;
    for (int i = 0; i < n; ++i)
        if (nodes[i] < 1 || nodes[i] > c->sanz) return CRT_ERR_BOUNDARY_NODE;  // src/rrandb.f90:62-66
    c->bnd_nodes.assign(nodes, nodes + n);
    c->bnd_zero = true;
    for (int i = 0; i < 2 * n; ++i) c->bnd_zero = c->bnd_zero && values[i] == 0.0;
    c->dmask_ok = false;
    c->have_kpot = c->have_volt = c->have_sens = false;
    return 0;
}

int crt_precal(crt_ctx *c, int lsytop, int *lbeta) {
    if (!c) return CRT_ERR_ARG;
    if (!c->have_grid || !c->have_kwn) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    const int K = c->kwnanz;
    int e = c->d_elbg.ensure((size_t)c->elanz * c->npk * K);
    e = e ? e : c->d_relbg.ensure((size_t)std::max(c->relanz, 1) * c->npk);
    e = e ? e : c->d_kg.ensure((size_t)std::max(c->relanz, 1) * std::max(c->eanz, 1) * K);
    if (e) return e;
    CRT_HIP(hipMemsetAsync(c->d_elbg.p, 0, c->d_elbg.n * sizeof(double), c->stream));
    CRT_HIP(hipMemsetAsync(c->d_relbg.p, 0, c->d_relbg.n * sizeof(double), c->stream));
    CRT_HIP(hipMemsetAsync(c->d_kg.p, 0, c->d_kg.n * sizeof(double), c->stream));
    if ((e = c->status_reset())) return e;
    double sytop = lsytop ? bsytop_h(c) : 0.0;
    c->tic(T_PRECAL);
    e = crt_launch_precal(c->stream, c->elanz, c->relanz, c->maxn, c->npk, c->eanz, K, c->d_etyp.p, c->d_btyp.p,
                          c->d_nrel.p, c->d_x.p, c->d_y.p, c->d_enr.p, c->d_kwn.p, sytop, c->d_elbg.p, c->d_relbg.p,
                          c->d_kg.p, c->d_status.p);
    c->toc(T_PRECAL);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->collect(T_PRECAL);
    int st = c->status_check();
    if (st) return st;
    c->lbeta = 0;
    for (int t = 0; t < c->typanz; ++t)
        if (c->typ[t] == 11) c->lbeta = 1;
    // mixed and Dirichlet boundaries exclude each other (src/precal.f90:102-106,130-135)
    if (c->lbeta && c->has13) return CRT_ERR_BOUNDARY_MIX;
    if (lbeta) *lbeta = c->lbeta;
    c->have_precal = true;
    return 0;
}

int crt_get_precal(crt_ctx *c, double *elbg, double *relbg, double *kg) {
    if (!c) return CRT_ERR_ARG;
    if (!c->have_precal) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    const int K = c->kwnanz;
    if (elbg) CRT_HIP(c->copy_out(elbg, c->d_elbg.p, (size_t)c->elanz * c->npk * K * sizeof(double)));
    if (relbg && c->relanz)
        CRT_HIP(c->copy_out(relbg, c->d_relbg.p, (size_t)c->relanz * c->npk * sizeof(double)));
    if (kg && c->relanz)
        CRT_HIP(c->copy_out(kg, c->d_kg.p, (size_t)c->relanz * c->eanz * K * sizeof(double)));
    return 0;
}

int crt_set_sigma(crt_ctx *c, const double *sigma) {
    if (!c || !sigma) return CRT_ERR_ARG;
    if (!c->have_grid) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    const zd *zs = reinterpret_cast<const zd *>(sigma);
    std::vector<zd> sre;
    if (c->inv_kind == 1) {
        // the DC forward assembles DBLE(sigma) (src/kompadc.f90:112-118) and bsendc
        // takes DBLE(sigma/volt) with a real volt (src/bsendc.f90:166-167)
        sre.resize(c->elanz);
        for (int i = 0; i < c->elanz; ++i) sre[i] = zd{zs[i].re, 0.0};
        zs = sre.data();
    }
    int e = c->d_sigma.upload(zs, c->elanz, c->stream);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->have_sigma = true;
    c->have_kpot = c->have_volt = false;
    return 0;
}

int crt_set_inversion_kind(crt_ctx *c, int kind) {
    if (!c || kind < 0 || kind > 2) return CRT_ERR_ARG;
    if (kind != c->inv_kind) {
        // the conductivity and everything derived from it belong to the previous kind
        c->have_sigma = c->have_kpot = c->have_volt = c->have_sens = false;
        c->sens32_ok = c->sensr_ok = c->sensr32_ok = false;
    }
    c->inv_kind = kind;
    return 0;
}

// Identity rows and columns of the Dirichlet nodes (randb for type-13 elements,
// randb2 for zero boundary values): per CSR entry 1 when it touches such a node, 2 on
// its diagonal.  Refuses prescribed non-zero values, whose lifting changes the
// right-hand side of every system (903); with `no_dirichlet_sources` also a source
// electrode on a Dirichlet node (the batched path does not clear such sources).
static int build_dmask(crt_ctx *c, bool no_dirichlet_sources = false) {
    if (!c->bnd_zero) return CRT_ERR_UNSUPPORTED;
    if (!c->dmask_ok) {
        c->dnode = c->dnode13;
        for (int nd : c->bnd_nodes) c->dnode[nd - 1] = 1;
        c->dmask_any = false;
        for (unsigned char v : c->dnode) c->dmask_any = c->dmask_any || v;
        if (c->dmask_any) {
            std::vector<unsigned char> m(c->h_col.size());
            for (size_t q = 0; q < m.size(); ++q) {
                const int i = c->h_crow[q], j = c->h_col[q];
                m[q] = (c->dnode[i] || c->dnode[j]) ? (i == j ? 2 : 1) : 0;
            }
            if (int e = c->d_dmask.upload(m.data(), m.size(), c->stream)) return e;
        }
        c->dmask_ok = true;
    }
    if (c->dmask_any && no_dirichlet_sources)
        for (int nd : c->enr0)
            if (c->dnode[nd]) return CRT_ERR_UNSUPPORTED;
    return 0;
}

// The systems whose right-hand side randb / randb2 clear: the source electrode sits on a
// Dirichlet node and the system is assembled by kompab (every (l, k) with mixed
// boundaries, l = 1 per k otherwise; the other sources of a shared factor get kompb's
// fresh right-hand side, src/kompb.f90:55-60, which keeps the source).  Their
// potentials are zero, as the reference's solve of a zero right-hand side gives.
static int zero_cleared_sources(crt_ctx *c) {
    if (!c->dmask_any) return 0;
    std::vector<int2> pairs;
    for (int l = 0; l < c->eanz; ++l)
        if (c->dnode[c->enr0[l]] && (c->lbeta || l == 0))
            for (int k = 0; k < c->kwnanz; ++k) pairs.push_back(int2{l, k});
    if (pairs.empty()) return 0;
    DBuf<int2> d_pairs;
    int e = d_pairs.upload(pairs.data(), pairs.size(), c->stream);
    e = e ? e : crt_launch_zero_sources(c->stream, c->sanz, c->kwnanz, c->eanz, (int)pairs.size(), d_pairs.p,
                                        c->d_kpotN.p);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// FP32 factorisation of one batch of systems, first solve and FP64 refinement
// sweeps (crt_band32.hip).  Returns 0 when every right-hand side met the
// tolerance, -1 when the batch has to be redone in FP64 (no convergence within
// refine_max sweeps, a growing correction, or an FP32 pivot failure), else an
// error code.  Device times (ms) of the factorisation, the first solve and the
// sweeps go to tf, ts, tr.
static int forward_batch_f32(crt_ctx *c, const DevGrid &g, int nb, size_t s0, int nrhs, int D, int &sweeps,
                             double &ratio, double &tf, double &ts, double &tr) {
    const int K = c->kwnanz, E = c->eanz, nblk = g.nblk, fstride = g.fstride;
    zf *panels = reinterpret_cast<zf *>(c->d_panels.p);
    zf *y32 = reinterpret_cast<zf *>(c->d_yf.p);
    zf *dinv32 = reinterpret_cast<zf *>(c->d_dinv.p);
    c->tic(T_FACTOR);
    int e = crt_launch_factor32(c->stream, g, nb, c->d_sys.p + s0, c->d_fak.p, c->d_vals.p, panels, y32, dinv32,
                                c->d_status.p + 1 + s0);
    if (e) return e;
    // DC: a non-positive pivot sends the batch to the FP64 kernels, which report it
    if (g.dc && (e = crt_launch_dc_pivot32(c->stream, nb, c->sanz, fstride, dinv32, c->d_status.p + 1 + s0))) return e;
    c->toc(T_FACTOR);
    c->tic(T_SOLVE);
    e = crt_launch_solve16_f32(c->stream, c->sanz, nblk, D, K, E, nrhs, c->d_rhs.p, c->d_enr.p, c->d_fak.p, panels,
                               dinv32, y32, c->d_yscr.p, c->d_kpotN.p, c->d_xacc.p, c->d_rstat.p, 1);
    if (e) return e;
    c->toc(T_SOLVE);
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->collect(T_FACTOR); tf = c->times[T_FACTOR];
    c->collect(T_SOLVE); ts = c->times[T_SOLVE];
    std::vector<int> st(nb);
    CRT_HIP(c->copy_out(st.data(), c->d_status.p + 1 + s0, nb * sizeof(int)));
    for (int v : st)
        if (v) return -1;
    std::vector<double> h(2 * (size_t)nrhs);
    double prev = HUGE_VAL;
    tr = 0.0;
    sweeps = 0;
    for (int sw = 1; sw <= c->refine_max; ++sw) {
        c->tic(T_REFINE);
        e = crt_launch_refine_resid(c->stream, c->sanz, fstride, g.vstride, nrhs, c->d_rhs.p, c->d_enr.p,
                                    c->d_frow.p, c->d_fent.p, c->d_fak.p, c->d_vals.p, c->d_xacc.p, c->d_yscr.p);
        e = e ? e : crt_launch_solve16_f32(c->stream, c->sanz, nblk, D, K, E, nrhs, c->d_rhs.p, c->d_enr.p,
                                           c->d_fak.p, panels, dinv32, y32, c->d_yscr.p, c->d_kpotN.p, c->d_xacc.p,
                                           c->d_rstat.p, 2);
        if (e) return e;
        c->toc(T_REFINE);
        CRT_HIP(c->copy_out(h.data(), c->d_rstat.p, h.size() * sizeof(double)));
        c->collect(T_REFINE);
        tr += c->times[T_REFINE];
        sweeps = sw;
        ratio = 0.0;
        for (int q = 0; q < nrhs; ++q) {
            const double d = h[2 * q], x = h[2 * q + 1];
            const double rq = x > 0.0 ? d / x : (d > 0.0 ? HUGE_VAL : 0.0);
            if (!(rq <= ratio)) ratio = rq;  // NaN propagates as "not converged"
        }
        if (getenv("CRT_DEBUG")) fprintf(stderr, "crt_forward fp32: batch at %zu sweep %d ratio %.3e\n", s0, sw, ratio);
        if (ratio <= c->refine_tol) return 0;
        // no longer halving: at the FP64 rounding floor of the residual (accept when
        // that floor is below refine_accept) or not converging (redo in FP64)
        if (!(ratio < 0.5 * prev)) return ratio <= c->refine_accept ? 0 : -1;
        prev = ratio;
    }
    return -1;
}

// ------------------------------------------------ condensed forward: plan
// The boundary nodes G are the nodes of the type-11 elements (their terms hold the
// source through kg, src/kompab.f90:111-116); the other nodes form the interior band,
// numbered in the order of smallest band width among the reference's (CutMcK)
// numbering and the nodes sorted by x or by y.  The band is cut into P pieces by
// windows of one band width (no entry reaches across a window): the windows are the
// separators S, the pieces independent band systems, factorised side by side.  The
// dense block is B = [S, G]: S is eliminated once per wavenumber, G per source.  G
// is numbered by the first interior neighbour, and so are each piece's coupled B
// columns, so the right-hand sides of a 16-column tile start close together.  Not
// applicable (ok false): no boundary nodes, an interior band of one tile or more than
// 16, more than 640 boundary nodes or 768 in B.
static int cond_plan(crt_ctx *c) {
    auto &P = c->cond;
    if (P.built) return 0;
    P.built = true;
    P.ok = false;
    const int sanz = c->sanz, nnz = c->nnz;
    const std::vector<int> &col = c->h_col, &crow = c->h_crow;
    std::vector<char> inG(sanz, 0);
    {
        int iel = 0;
        for (int t = 0; t < c->typanz; ++t)
            for (int j = 0; j < c->nelanz[t]; ++j, ++iel) {
                if (c->typ[t] != 11) continue;
                for (int m = 0; m < c->selanz[t]; ++m) inG[c->nrel[(size_t)iel * c->maxn + m] - 1] = 1;
            }
    }
    std::vector<int> I0, G0;
    for (int i = 0; i < sanz; ++i) (inG[i] ? G0 : I0).push_back(i);
    const int nI0 = (int)I0.size();
    P.nG = (int)G0.size();
    if (P.nG == 0 || nI0 == 0) return 0;
    P.nGp = (P.nG + 15) / 16 * 16;
    P.nt = P.nGp / 16;
    if (P.nt > crt_cond_dense_max_tiles()) return 0;
    // interior numbering of smallest band width
    std::vector<int> best, pos(sanz, -1);
    int bestw = INT32_MAX;
    for (int order = 0; order < 3; ++order) {
        std::vector<int> ord = I0;
        if (order == 1)
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
                return c->x[a] != c->x[b] ? c->x[a] < c->x[b] : c->y[a] < c->y[b];
            });
        if (order == 2)
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
                return c->y[a] != c->y[b] ? c->y[a] < c->y[b] : c->x[a] < c->x[b];
            });
        std::vector<int> ps(sanz, -1);
        for (int i = 0; i < nI0; ++i) ps[ord[i]] = i;
        int w = 0;
        for (int q = 0; q < nnz; ++q) {
            const int a = ps[crow[q]], b = ps[col[q]];
            if (a >= 0 && b >= 0) w = std::max(w, std::abs(a - b));
        }
        if (w < bestw) {
            bestw = w;
            best = ord;
            pos = ps;
            P.order = order;
        }
    }
    P.mbI = bestw;
    P.DI = std::max(1, (bestw + 15) / 16);
    if (P.DI < 2 || P.DI > 16) return 0;
    // pieces: [start, start + len) of the interior positions, separators of width mbI between.
    // The most pieces (up to force_np, else 8) of at least 32 band widths each whose
    // separators fit the prefix kernel: the pieces' factorisations run side by side, each a
    // chain of block steps as long as the piece, while the separators' elimination grows
    // with their count; the first condensed forward of a grid times the candidates
    // (crt_forward) and keeps the fastest.
    int np = P.force_np > 0 ? P.force_np : 8, minlen = 32;
    if (const char *ev = getenv("CRT_COND_PIECES")) np = std::max(1, atoi(ev));
    if (const char *ev = getenv("CRT_COND_MINLEN")) minlen = std::max(1, atoi(ev));
    const int w = bestw;
    for (; np > 1; --np) {
        const int L = (nI0 - (np - 1) * w) / np;
        const int nS = (np - 1) * w;
        if (L >= minlen * w && (nS + 15) / 16 <= crt_cond_prefix_max_tiles()) break;
    }
    std::vector<int> pst, plen;
    {
        const int L = (nI0 - (np - 1) * w) / np;
        int a = 0;
        for (int p = 0; p < np; ++p) {
            pst.push_back(a);
            const int len = p == np - 1 ? nI0 - a : L;
            plen.push_back(len);
            a += len + w;
        }
    }
    P.P = np;
    std::vector<int> piece(sanz, -1), prow(sanz, -1);  // piece and combined row of each interior node
    std::vector<int> S;                                 // separator nodes in position order
    P.roff.assign(np, 0);
    P.plen = plen;
    P.nodeI.clear();
    for (int p = 0; p < np; ++p) {
        P.roff[p] = (int)P.nodeI.size();
        for (int r = 0; r < plen[p]; ++r) {
            const int nd = best[pst[p] + r];
            piece[nd] = p;
            prow[nd] = (int)P.nodeI.size();
            P.nodeI.push_back(nd);
        }
    }
    for (int r = 0; r < nI0; ++r)
        if (piece[best[r]] < 0) S.push_back(best[r]);
    P.nI = (int)P.nodeI.size();
    P.nS = (int)S.size();
    P.nSp = (P.nS + 15) / 16 * 16;
    P.JS = P.nSp / 16;
    P.nBp = P.nSp + P.nGp;
    P.ntB = P.nBp / 16;
    P.nblkP = 0;
    for (int p = 0; p < np; ++p) P.nblkP = std::max(P.nblkP, (plen[p] + 15) / 16);
    // combined interior CSR over the pieces (no entry joins two pieces)
    struct E3 { int r, cc, q; };
    std::vector<E3> ie;
    for (int q = 0; q < nnz; ++q) {
        const int i = crow[q], j = col[q];
        if (piece[i] < 0 || piece[j] < 0) continue;
        if (piece[i] != piece[j]) return CRT_ERR_STATE;
        ie.push_back({std::max(prow[i], prow[j]), std::min(prow[i], prow[j]), q});
    }
    std::sort(ie.begin(), ie.end(), [](const E3 &u, const E3 &v) { return u.r != v.r ? u.r < v.r : u.cc < v.cc; });
    P.nnzI = (int)ie.size();
    std::vector<int> rp(P.nI + 1, 0), ci(P.nnzI), cr(P.nnzI), qm(P.nnzI);
    for (int e = 0; e < P.nnzI; ++e) {
        rp[ie[e].r + 1]++;
        ci[e] = ie[e].cc;
        cr[e] = ie[e].r;
        qm[e] = ie[e].q;
    }
    for (int i = 0; i < P.nI; ++i) rp[i + 1] += rp[i];
    for (int i = 0; i < P.nI; ++i)
        if (rp[i + 1] == rp[i] || ci[rp[i + 1] - 1] != i) return CRT_ERR_STATE;  // every row ends on its diagonal
    // one tile map per piece over its local rows (padding rows: unit diagonal)
    const int D = P.DI;
    const size_t tstride = (size_t)P.nblkP * (D + 1) * 256;
    std::vector<int> tmap(tstride * np, P.nnzI);
    for (int p = 0; p < np; ++p) {
        int *tm = tmap.data() + tstride * p;
        for (int r = 0; r < plen[p]; ++r) {
            const int g = P.roff[p] + r;
            for (int q = rp[g]; q < rp[g + 1]; ++q) {
                const int cl = ci[q] - P.roff[p], I = r / 16, d = I - cl / 16;
                if (d < 0 || d > D) return CRT_ERR_STATE;
                tm[((size_t)(I * (D + 1) + d) * 16 + (r % 16)) * 16 + (cl % 16)] = q;
            }
        }
        for (int r = plen[p]; r < P.nblkP * 16; ++r)
            tm[((size_t)((r / 16) * (D + 1)) * 16 + (r % 16)) * 16 + (r % 16)] = P.nnzI + 1;
    }
    // B numbering: S in position order, padding, then G by its first interior neighbour
    std::vector<int> key(sanz, INT32_MAX);
    for (int q = 0; q < nnz; ++q) {
        const int i = crow[q], j = col[q];
        if (inG[i] && pos[j] >= 0) key[i] = std::min(key[i], pos[j]);
        if (inG[j] && pos[i] >= 0) key[j] = std::min(key[j], pos[i]);
    }
    std::stable_sort(G0.begin(), G0.end(), [&](int a, int b) { return key[a] < key[b]; });
    const int nBp = P.nBp, nSp = P.nSp;
    std::vector<int> bidx(sanz, -1);
    P.nodeB.assign(nBp, -1);
    for (int i = 0; i < P.nS; ++i) {
        bidx[S[i]] = i;
        P.nodeB[i] = S[i];
    }
    for (int g = 0; g < P.nG; ++g) {
        bidx[G0[g]] = nSp + g;
        P.nodeB[nSp + g] = G0[g];
    }
    // B-B entries, B diagonals, couplings of each B node to the pieces
    std::vector<int> gmap((size_t)nBp * nBp, -1), qdiagB(nBp, -1);
    struct Cp { int b, p, r, q; };  // B node b coupled to local row r of piece p by entry q
    std::vector<Cp> cpl;
    for (int q = 0; q < nnz; ++q) {
        const int i = crow[q], j = col[q];
        const int bi = bidx[i], bj = bidx[j];
        if (bi >= 0 && bj >= 0) {
            gmap[(size_t)bj * nBp + bi] = q;
            gmap[(size_t)bi * nBp + bj] = q;
            if (i == j) qdiagB[bi] = q;
        } else if (bi >= 0 || bj >= 0) {
            const int b = bi >= 0 ? bi : bj, x = bi >= 0 ? j : i;
            cpl.push_back({b, piece[x], prow[x] - P.roff[piece[x]], q});
        }
    }
    for (int b = 0; b < nBp; ++b)
        if ((P.nodeB[b] >= 0) != (qdiagB[b] >= 0)) return CRT_ERR_STATE;
    std::vector<int> qdiagG(P.nG);
    for (int g = 0; g < P.nG; ++g) qdiagG[g] = qdiagB[nSp + g];
    // entries that hold boundary terms (the node pairs of the type-11 elements), per
    // lower tile of the G block: (element c*16 + r of the tile, entry, G indices)
    const size_t ntri = (size_t)P.nt * (P.nt + 1) / 2;
    struct BE { int e, q, g1, g2; };
    std::vector<std::vector<BE>> bt(ntri);
    {
        std::vector<char> seen((size_t)P.nGp * P.nGp, 0);
        int iel = 0;
        for (int t = 0; t < c->typanz; ++t)
            for (int j = 0; j < c->nelanz[t]; ++j, ++iel) {
                if (c->typ[t] != 11) continue;
                const int ns = c->selanz[t];
                for (int a = 0; a < ns; ++a)
                    for (int b = 0; b <= a; ++b) {
                        int g1 = bidx[c->nrel[(size_t)iel * c->maxn + a] - 1] - nSp;
                        int g2 = bidx[c->nrel[(size_t)iel * c->maxn + b] - 1] - nSp;
                        if (g1 < g2) std::swap(g1, g2);
                        if (seen[(size_t)g2 * P.nGp + g1]) continue;
                        seen[(size_t)g2 * P.nGp + g1] = 1;
                        const int q = gmap[(size_t)(nSp + g2) * nBp + nSp + g1];
                        if (q < 0) return CRT_ERR_STATE;
                        const size_t ti = (size_t)(g1 / 16) * (g1 / 16 + 1) / 2 + g2 / 16;
                        bt[ti].push_back({(g2 % 16) * 16 + g1 % 16, q, g1, g2});
                    }
            }
    }
    std::vector<int> bptr(ntri + 1, 0), bent, bq;
    std::vector<int2> bg;
    for (size_t ti = 0; ti < ntri; ++ti) {
        for (auto &b : bt[ti]) {
            bent.push_back(b.e);
            bq.push_back(b.q);
            bg.push_back(int2{b.g1, b.g2});
        }
        bptr[ti + 1] = (int)bent.size();
    }
    P.nb = (int)bent.size();
    if (bent.empty()) { bent.push_back(0); bq.push_back(0); bg.push_back(int2{0, 0}); }
    // columns of each piece: its coupled B nodes by first coupled row, then its electrodes
    P.E = c->eanz;
    const int E = P.E;
    std::vector<int> first((size_t)np * nBp, INT32_MAX);
    for (auto &cp : cpl) first[(size_t)cp.p * nBp + cp.b] = std::min(first[(size_t)cp.p * nBp + cp.b], cp.r);
    std::vector<int> bcol((size_t)np * nBp, -1), ucol((size_t)np * std::max(E, 1), -1);
    std::vector<std::vector<int>> cols(np);  // B index of each coupled column
    P.nbc.assign(np, 0);
    int ncmax = 0;
    std::vector<int> ne(np, 0);
    for (int p = 0; p < np; ++p) {
        std::vector<int> bs;
        for (int b = 0; b < nBp; ++b)
            if (first[(size_t)p * nBp + b] != INT32_MAX) bs.push_back(b);
        std::stable_sort(bs.begin(), bs.end(), [&](int a, int b) {
            return first[(size_t)p * nBp + a] < first[(size_t)p * nBp + b];
        });
        for (size_t cix = 0; cix < bs.size(); ++cix) bcol[(size_t)p * nBp + bs[cix]] = (int)cix;
        cols[p] = bs;
        P.nbc[p] = (int)bs.size();
        for (int l = 0; l < E; ++l)
            if (piece[c->enr0[l]] == p) ucol[(size_t)p * E + l] = P.nbc[p] + ne[p]++;
        ncmax = std::max(ncmax, P.nbc[p] + ne[p]);
    }
    P.ncolP = std::max(16, (ncmax + 15) / 16 * 16);
    P.ntileP = P.ncolP / 16;
    std::vector<int> colmap((size_t)np * P.ncolP, -1);
    for (int p = 0; p < np; ++p)
        for (int cix = 0; cix < P.nbc[p]; ++cix) colmap[(size_t)p * P.ncolP + cix] = cols[p][cix];
    // right-hand-side entries per (piece, column tile), by row
    struct RE { int p, row, colx, rowg, b, q; };
    std::vector<RE> re;
    for (auto &cp : cpl) re.push_back({cp.p, cp.r, bcol[(size_t)cp.p * nBp + cp.b], P.roff[cp.p] + cp.r, cp.b, cp.q});
    std::vector<int> eB(std::max(E, 1), -1);
    for (int l = 0; l < E; ++l) {
        const int nd = c->enr0[l];
        if (piece[nd] >= 0) re.push_back({piece[nd], prow[nd] - P.roff[piece[nd]], ucol[(size_t)piece[nd] * E + l],
                                          prow[nd], -1, -1});
        else eB[l] = bidx[nd];
    }
    std::sort(re.begin(), re.end(), [](const RE &u, const RE &v) {
        if (u.p != v.p) return u.p < v.p;
        if (u.colx / 16 != v.colx / 16) return u.colx / 16 < v.colx / 16;
        return u.row != v.row ? u.row < v.row : u.colx < v.colx;
    });
    P.nent = (int)re.size();
    const int npt = np * P.ntileP;
    std::vector<int> tptr(npt + 1, 0), erow(std::max(P.nent, 1)), erowg(std::max(P.nent, 1)), ecol(std::max(P.nent, 1)),
        ebcol(std::max(P.nent, 1)), eq(std::max(P.nent, 1)), j0(npt, P.nblkP);
    for (int e = 0; e < P.nent; ++e) {
        tptr[re[e].p * P.ntileP + re[e].colx / 16 + 1]++;
        erow[e] = re[e].row;
        erowg[e] = re[e].rowg;
        ecol[e] = re[e].colx % 16;
        ebcol[e] = re[e].b;
        eq[e] = re[e].q;
    }
    for (int t = 0; t < npt; ++t) tptr[t + 1] += tptr[t];
    for (int t = 0; t < npt; ++t)
        if (tptr[t] < tptr[t + 1]) j0[t] = erow[tptr[t]] / 16;
    // A_Bi lists by B row: (combined interior row, piece, entry)
    std::vector<std::vector<Cp>> gl(nBp);
    for (auto &cp : cpl) gl[cp.b].push_back(cp);
    std::vector<int> gptr(nBp + 1, 0), grow_, gpc, gq;
    for (int b = 0; b < nBp; ++b) {
        std::sort(gl[b].begin(), gl[b].end(), [](const Cp &u, const Cp &v) { return u.p != v.p ? u.p < v.p : u.r < v.r; });
        for (auto &cp : gl[b]) {
            grow_.push_back(P.roff[cp.p] + cp.r);
            gpc.push_back(cp.p);
            gq.push_back(cp.q);
        }
        gptr[b + 1] = (int)grow_.size();
    }
    if (grow_.empty()) { grow_.push_back(0); gpc.push_back(0); gq.push_back(0); }
    std::vector<int> zero(P.nnzI + 1, 0), enr0(std::max(np, 1), 0);
    hipStream_t st = c->stream;
    int e = 0;
    e = e ? e : P.d_row_ptr.upload(rp.data(), rp.size(), st);
    e = e ? e : P.d_col.upload(ci.data(), ci.size(), st);
    e = e ? e : P.d_crow.upload(cr.data(), cr.size(), st);
    e = e ? e : P.d_qmap.upload(qm.data(), qm.size(), st);
    e = e ? e : P.d_zero.upload(zero.data(), zero.size(), st);
    e = e ? e : P.d_enr0.upload(enr0.data(), enr0.size(), st);
    e = e ? e : P.d_tmap.upload(tmap.data(), tmap.size(), st);
    e = e ? e : P.d_tptr.upload(tptr.data(), tptr.size(), st);
    e = e ? e : P.d_erow.upload(erow.data(), erow.size(), st);
    e = e ? e : P.d_erowg.upload(erowg.data(), erowg.size(), st);
    e = e ? e : P.d_ecol.upload(ecol.data(), ecol.size(), st);
    e = e ? e : P.d_ebcol.upload(ebcol.data(), ebcol.size(), st);
    e = e ? e : P.d_eq.upload(eq.data(), eq.size(), st);
    e = e ? e : P.d_j0.upload(j0.data(), j0.size(), st);
    e = e ? e : P.d_gi_ptr.upload(gptr.data(), gptr.size(), st);
    e = e ? e : P.d_gi_row.upload(grow_.data(), grow_.size(), st);
    e = e ? e : P.d_gi_piece.upload(gpc.data(), gpc.size(), st);
    e = e ? e : P.d_gi_q.upload(gq.data(), gq.size(), st);
    e = e ? e : P.d_gmap.upload(gmap.data(), gmap.size(), st);
    e = e ? e : P.d_qdiag.upload(qdiagB.data(), qdiagB.size(), st);
    e = e ? e : P.d_qdiagG.upload(qdiagG.data(), qdiagG.size(), st);
    e = e ? e : P.d_nodeI.upload(P.nodeI.data(), P.nodeI.size(), st);
    e = e ? e : P.d_nodeB.upload(P.nodeB.data(), P.nodeB.size(), st);
    e = e ? e : P.d_eB.upload(eB.data(), eB.size(), st);
    e = e ? e : P.d_bcol.upload(bcol.data(), bcol.size(), st);
    e = e ? e : P.d_ucol.upload(ucol.data(), ucol.size(), st);
    e = e ? e : P.d_colmap.upload(colmap.data(), colmap.size(), st);
    e = e ? e : P.d_roff.upload(P.roff.data(), P.roff.size(), st);
    e = e ? e : P.d_plen.upload(P.plen.data(), P.plen.size(), st);
    e = e ? e : P.d_nbc.upload(P.nbc.data(), P.nbc.size(), st);
    e = e ? e : P.d_bptr.upload(bptr.data(), bptr.size(), st);
    e = e ? e : P.d_bent.upload(bent.data(), bent.size(), st);
    e = e ? e : P.d_bq.upload(bq.data(), bq.size(), st);
    e = e ? e : P.d_bg.upload(bg.data(), bg.size(), st);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(st));
    P.ok = true;
    return 0;
}

// ------------------------------------------------ condensed forward: solve
// Every wavenumber chunk: the interior area sums, scaling, assembly and band
// factorisation of every piece (the band kernels on the pieces' grids), the
// substitutions for each piece's [Z | U], the Schur complement T over B and the
// right-hand sides, the separators' elimination (prefix), the dense factorisation and
// solve of every (l, k) over G, and the interior potentials.  Errors: a zero interior
// or separator diagonal or pivot names the wavenumber's first source (those rows are
// the same for every source), a boundary one its own source; the first failing
// system in the reference's loop order (wavenumber, then electrode) is returned.
static int forward_condensed(crt_ctx *c) {
    auto &P = c->cond;
    const int K = c->kwnanz, E = c->eanz, nP = P.P;
    hipStream_t st = c->stream;
    DevGrid g = c->devgrid();
    int e = c->d_area.ensure((size_t)K * c->nnz);
    e = e ? e : crt_launch_area(st, g, K, c->d_area.p);
    if (e) return e;
    g.area = c->d_area.p;
    const int nGp = P.nGp, nBp = P.nBp, nt = P.nt, ntB = P.ntB, JS = P.JS, ldx = P.nblkP * 16, D = P.DI;
    const int fstrideC = (P.nI + 15) / 16 * 16;
    const size_t ntri = (size_t)nt * (nt + 1) / 2, ntriB = (size_t)ntB * (ntB + 1) / 2;
    const size_t per_k = sizeof(zd) * ((size_t)nP * P.ncolP * ldx + (size_t)nP * P.nblkP * (D + 1) * 256 +
                                       2 * (size_t)nP * ldx + (size_t)nP * (P.nnzI + 2) + P.nnzI + ntriB * 256 * 2 +
                                       ntri * 256 + 2 * (size_t)E * nBp + (size_t)E * (ntri + nt) * 256 +
                                       (size_t)JS * 256 + P.nent + (size_t)E * P.nb) +
                         sizeof(double) * (2 * (size_t)nP * fstrideC + nBp);
    const int kc = (int)std::max<size_t>(1, std::min<size_t>(K, c->panel_budget / per_k));
    const int ns = kc * nP;
    e = P.areaI.ensure((size_t)kc * P.nnzI);
    e = e ? e : P.vals.ensure((size_t)ns * (P.nnzI + 2));
    e = e ? e : P.panels.ensure((size_t)ns * P.nblkP * (D + 1) * 256);
    e = e ? e : P.yI.ensure((size_t)ns * std::max<size_t>(ldx, fstrideC));
    e = e ? e : P.dinv.ensure((size_t)ns * ldx);
    e = e ? e : P.fak.ensure((size_t)ns * fstrideC);
    e = e ? e : P.dga.ensure((size_t)ns * fstrideC);
    e = e ? e : P.fB.ensure((size_t)kc * nBp);
    e = e ? e : P.rv.ensure((size_t)kc * std::max(P.nent, 1));
    e = e ? e : P.X.ensure((size_t)ns * P.ncolP * ldx);
    e = e ? e : P.T.ensure((size_t)kc * ntriB * 256);
    e = e ? e : P.S0.ensure((size_t)kc * ntri * 256);
    e = e ? e : P.Gr.ensure((size_t)kc * E * nBp);
    e = e ? e : P.XB.ensure((size_t)kc * E * nBp);
    e = e ? e : P.Lt.ensure((size_t)kc * E * ntri * 256);
    e = e ? e : P.Li.ensure((size_t)kc * E * nt * 256);
    e = e ? e : P.Lpre.ensure((size_t)kc * (JS ? ntriB : 1) * 256);
    e = e ? e : P.Lipre.ensure((size_t)kc * std::max(JS, 1) * 256);
    e = e ? e : P.bval.ensure((size_t)kc * E * std::max(P.nb, 1));
    e = e ? e : P.sysI.ensure(ns);
    e = e ? e : P.stI.ensure((size_t)K * nP);
    e = e ? e : P.stD.ensure((size_t)K * E);
    e = e ? e : c->d_kpotN.ensure((size_t)c->sanz * K * E);
    if (e) return e;
    std::vector<int2> sysI(ns);
    for (int kk = 0; kk < kc; ++kk)
        for (int p = 0; p < nP; ++p) sysI[(size_t)kk * nP + p] = int2{p, kk};
    CRT_HIP(hipMemcpyAsync(P.sysI.p, sysI.data(), ns * sizeof(int2), hipMemcpyHostToDevice, st));
    CRT_HIP(hipMemsetAsync(P.stI.p, 0, (size_t)K * nP * sizeof(int), st));
    CRT_HIP(hipMemsetAsync(P.stD.p, 0, (size_t)K * E * sizeof(int), st));
    // the pieces' grids: rows of all pieces in one CSR for the scaling and assembly
    // (gI), each piece's own tile map over its padded local rows for the factorisation (gP)
    DevGrid gI = g;
    gI.sanz = P.nI; gI.mb = P.mbI; gI.nnz = P.nnzI; gI.fstride = fstrideC; gI.vstride = P.nnzI + 2;
    gI.row_ptr = P.d_row_ptr.p; gI.col = P.d_col.p; gI.crow = P.d_crow.p;
    gI.cptr = P.d_zero.p; gI.csplit = P.d_zero.p; gI.contrib = P.d_zero.p;
    gI.enr = P.d_enr0.p;
    gI.area = P.areaI.p;
    gI.dmask = nullptr;
    gI.tmap = nullptr; gI.tmapD = 0; gI.tmap_sys = 0;
    DevGrid gP = gI;
    gP.sanz = ldx; gP.nblk = P.nblkP; gP.D = D;
    gP.tmap = P.d_tmap.p; gP.tmapD = D; gP.tmap_sys = (long)P.nblkP * (D + 1) * 256;
    int variant = D <= 10 ? 1 : 2;
    if (const char *ev = getenv("CRT_COND_VARIANT")) {
        const int v = atoi(ev);
        if ((v == 1 && D <= 10) || (v == 2 && D >= 3)) variant = v;
    }
    double tdiag = 0, tfac = 0, tsol = 0, tsch = 0, tprod = 0, tpre = 0;
    for (int k0 = 0; k0 < K; k0 += kc) {
        const int n = std::min(kc, K - k0), nsn = n * nP;
        int *stI = P.stI.p + (size_t)k0 * nP;
        c->tic(T_DIAG);
        e = crt_launch_cond_gather(st, n, k0, c->nnz, P.nnzI, P.d_qmap.p, c->d_area.p, P.areaI.p);
        // the systems (piece, kk) read the area sums of wavenumber kk: sys.y
        e = e ? e : crt_launch_diag(st, gI, nsn, P.sysI.p, P.fak.p, P.dga.p, stI);
        e = e ? e : crt_launch_assemble(st, gI, nsn, P.sysI.p, P.fak.p, P.dga.p, P.vals.p);
        e = e ? e : crt_launch_cond_fb(st, n, k0, P.nS, nBp, c->nnz, g.dc, nP, P.d_qdiag.p, c->d_area.p, P.fB.p, stI);
        e = e ? e : crt_launch_cond_rhs(st, n, k0, P.nent, c->nnz, nBp, fstrideC, nP, P.d_erowg.p, P.d_ebcol.p,
                                        P.d_eq.p, c->d_area.p, P.fak.p, P.fB.p, P.rv.p);
        if (e) return e;
        c->toc(T_DIAG);
        c->tic(T_FACTOR);
        e = crt_launch_factor16(st, gP, nsn, P.sysI.p, P.fak.p, P.vals.p, P.panels.p, P.yI.p, P.dinv.p, stI, variant);
        if (!e && g.dc) e = crt_launch_dc_pivot(st, nsn, ldx, ldx, P.dinv.p, stI);
        if (e) return e;
        c->toc(T_FACTOR);
        c->tic(T_SOLVE);
        e = crt_launch_cond_trsm(st, D, 0, P.nblkP, ldx, P.ntileP, nsn, nP, P.panels.p, P.dinv.p, P.X.p, P.d_tptr.p,
                                 P.d_erow.p, P.d_ecol.p, P.rv.p, P.nent, P.d_j0.p, g.dc);
        e = e ? e : crt_launch_cond_trsm(st, D, 1, P.nblkP, ldx, P.ntileP, nsn, nP, P.panels.p, P.dinv.p, P.X.p,
                                         P.d_tptr.p, P.d_erow.p, P.d_ecol.p, P.rv.p, P.nent, P.d_j0.p, g.dc);
        e = e ? e : crt_launch_cond_schur(st, n, k0, nBp, E, ldx, P.ntileP, c->nnz, fstrideC, nP, P.d_gi_ptr.p,
                                          P.d_gi_row.p, P.d_gi_piece.p, P.d_gi_q.p, P.d_roff.p, P.d_bcol.p,
                                          P.d_ucol.p, P.d_gmap.p, P.d_eB.p, c->d_area.p, P.fak.p, P.fB.p, P.X.p,
                                          JS ? P.T.p : P.S0.p, P.Gr.p);
        if (e) return e;
        c->toc(T_SOLVE);
        c->tic(T_PREFIX);
        if (JS)  // the separators, once per wavenumber: L_SS, L_GS and S0 = T_GG - L_GS L_GS^T
            e = crt_launch_cond_dense(st, g, 1, P.nG, nt, nBp, JS, E, K, k0, n, nullptr, P.T.p, P.Gr.p, P.d_bptr.p,
                                      P.d_bent.p, P.d_qdiagG.p, P.d_nodeB.p, P.bval.p, P.nb, P.fB.p, P.Lt.p, P.Li.p,
                                      P.Lpre.p, P.Lipre.p, P.S0.p, P.XB.p, c->d_kpotN.p, stI, nP);
        if (JS) e = e ? e : crt_launch_cond_lgs(st, n, ntB, JS, P.T.p, P.Lpre.p, P.Lipre.p, g.dc);
        if (JS) e = e ? e : crt_launch_cond_s0(st, n, nt, ntB, JS, P.T.p, P.Lpre.p, P.S0.p, g.dc);
        e = e ? e : crt_launch_cond_bval(st, g, n * E, E, k0, nBp, P.nb, P.d_bq.p, P.d_bg.p, P.fB.p + P.nSp, P.bval.p);
        if (e) return e;
        c->toc(T_PREFIX);
        c->tic(T_SCHUR);
        e = crt_launch_cond_dense(st, g, 0, P.nG, nt, nBp, JS, E, K, k0, n * E, P.S0.p, P.T.p, P.Gr.p,
                                          P.d_bptr.p, P.d_bent.p, P.d_qdiagG.p, P.d_nodeB.p, P.bval.p, P.nb, P.fB.p,
                                          P.Lt.p, P.Li.p, P.Lpre.p, P.Lipre.p, P.S0.p, P.XB.p, c->d_kpotN.p,
                                          P.stD.p + (size_t)k0 * E, 1);
        if (e) return e;
        c->toc(T_SCHUR);
        c->tic(T_PRODUCT);
        e = crt_launch_cond_gemm(st, nsn, nP, k0, ldx, P.ntileP, P.ncolP, nBp, E, K, fstrideC, P.X.p, P.XB.p, P.fak.p,
                                 P.d_nodeI.p, P.d_roff.p, P.d_plen.p, P.d_nbc.p, P.d_colmap.p, P.d_ucol.p,
                                 c->d_kpotN.p, g.dc);
        if (e) return e;
        c->toc(T_PRODUCT);
        CRT_HIP(hipStreamSynchronize(st));
        c->collect(T_DIAG); tdiag += c->times[T_DIAG];
        c->collect(T_FACTOR); tfac += c->times[T_FACTOR];
        c->collect(T_SOLVE); tsol += c->times[T_SOLVE];
        c->collect(T_PREFIX); tpre += c->times[T_PREFIX];
        c->collect(T_SCHUR); tsch += c->times[T_SCHUR];
        c->collect(T_PRODUCT); tprod += c->times[T_PRODUCT];
    }
    c->times[T_DIAG] = tdiag; c->times[T_FACTOR] = tfac; c->times[T_SOLVE] = tsol; c->times[T_SCHUR] = tsch;
    c->times[T_PRODUCT] = tprod;
    c->times[T_PREFIX] = tpre;
    c->times[T_REFINE] = 0.0;
    std::vector<int> sI((size_t)K * nP), sD((size_t)K * E);
    CRT_HIP(c->copy_out(sI.data(), P.stI.p, sI.size() * sizeof(int)));
    CRT_HIP(c->copy_out(sD.data(), P.stD.p, sD.size() * sizeof(int)));
    for (int k = 0; k < K; ++k) {
        for (int p = 0; p < nP; ++p)
            if (sI[(size_t)k * nP + p]) return sI[(size_t)k * nP + p];
        for (int l = 0; l < E; ++l)
            if (sD[(size_t)k * E + l]) return sD[(size_t)k * E + l];
    }
    return 0;
}

// The interior pieces of the condensed forward, chosen by timing one forward of each
// candidate count on the grid's first condensed forward (the latency of the pieces'
// factorisations against the separators' elimination and the longer boundary
// substitutions: C1 3 or 4, C3 6 of 2..8).  CRT_COND_PIECES fixes the count instead,
// CRT_COND_TUNE=0 keeps the plan's own choice.
static int tune_pieces(crt_ctx *c) {
    auto &P = c->cond;
    if (P.tuned) return 0;
    P.tuned = true;
    const char *tv = getenv("CRT_COND_TUNE");
    if (getenv("CRT_COND_PIECES") || (tv && atoi(tv) == 0)) return 0;
    int best = P.P;
    double best_ms = 0.0;
    std::vector<int> seen;
    for (int cand : {2, 3, 4, 5, 6, 8}) {
        P.built = false;
        P.force_np = cand;
        if (int e = cond_plan(c)) return e;
        if (!P.ok) return 0;
        if (std::find(seen.begin(), seen.end(), P.P) != seen.end()) continue;
        seen.push_back(P.P);
        if (int e = forward_condensed(c)) return e;
        double ms = 0.0;
        for (int t : {T_DIAG, T_FACTOR, T_SOLVE, T_PREFIX, T_SCHUR, T_PRODUCT}) ms += c->times[t];
        if (best_ms == 0.0 || ms < best_ms) {
            best_ms = ms;
            best = P.P;
        }
    }
    P.built = false;
    P.force_np = best;
    return cond_plan(c);
}

int crt_set_forward_method(crt_ctx *c, int method) {
    if (!c || method < -1 || method > 1) return CRT_ERR_ARG;
    c->fwd_method = method;
    return 0;
}

int crt_get_forward_method(crt_ctx *c, int *method) {
    if (!c || !method) return CRT_ERR_ARG;
    *method = c->last_method;
    return 0;
}

int crt_get_condense_info(crt_ctx *c, int *info) {
    if (!c || !info) return CRT_ERR_ARG;
    const auto &P = c->cond;
    const int v[8] = {P.ok ? 1 : 0, P.nI, P.nG, P.mbI, P.DI, P.order, P.P, P.nS};
    for (int i = 0; i < 8; ++i) info[i] = v[i];
    return 0;
}

// per node the strictly lower CSR entries of its column (the transpose of the row
// lists), for the right-hand sides that sum over every entry of a row and column
static int build_tcol(crt_ctx *c) {
    if (c->tcol_built) return 0;
    const int sanz = c->sanz, nnz = c->nnz;
    std::vector<int> cnt(sanz + 1, 0), tq;
    for (int q = 0; q < nnz; ++q)
        if (c->h_crow[q] != c->h_col[q]) ++cnt[c->h_col[q] + 1];
    for (int i = 0; i < sanz; ++i) cnt[i + 1] += cnt[i];
    tq.assign(std::max(cnt[sanz], 1), 0);
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int q = 0; q < nnz; ++q)
        if (c->h_crow[q] != c->h_col[q]) tq[pos[c->h_col[q]]++] = q;
    int e = c->d_tptr.upload(cnt.data(), cnt.size(), c->stream);
    e = e ? e : c->d_tq.upload(tq.data(), tq.size(), c->stream);
    if (e) return e;
    c->tcol_built = true;
    return 0;
}

// lana (src/fem.f90:350, 384-386): kpot is the analytic half-space potential of each
// source, no system is assembled or solved
static int forward_analytic(crt_ctx *c) {
    const int K = c->kwnanz, E = c->eanz, sanz = c->sanz;
    const int chunk = std::max(1, std::min(K * E, 1024));
    int e = c->d_sigma0.ensure(1);
    e = e ? e : c->d_pota.ensure((size_t)chunk * sanz);
    e = e ? e : c->d_rhs.ensure(chunk);
    e = e ? e : c->d_kpotN.ensure((size_t)sanz * K * E);
    if (e) return e;
    c->have_kpot = false;
    c->last_method = 0;
    if ((e = crt_launch_refsig(c->stream, c->elanz, c->maxn, c->d_etyp.p, c->d_nrel.p, c->d_x.p, c->d_y.p,
                               c->d_sigma.p, c->d_sigma0.p)))
        return e;
    std::vector<int4> r;
    for (int s0 = 0; s0 < K * E; s0 += chunk) {
        const int n = std::min(chunk, K * E - s0);
        r.clear();
        for (int s = s0; s < s0 + n; ++s) r.push_back(int4{s % E, s / E, 0, 0});
        CRT_HIP(hipMemcpyAsync(c->d_rhs.p, r.data(), n * sizeof(int4), hipMemcpyHostToDevice, c->stream));
        e = crt_launch_potana(c->stream, sanz, n, c->d_rhs.p, c->d_enr.p, c->d_x.p, c->d_y.p, c->d_kwn.p,
                              c->d_sigma0.p, c->d_pota.p, sanz);
        e = e ? e : crt_launch_pota_kpot(c->stream, sanz, K, E, n, c->d_rhs.p, c->d_pota.p, sanz, 0, c->d_kpotN.p);
        if (e) return e;
        CRT_HIP(hipStreamSynchronize(c->stream));
    }
    for (int t : {T_DIAG, T_FACTOR, T_SOLVE, T_SCHUR, T_PRODUCT, T_PREFIX, T_REFINE}) c->times[t] = 0.0;
    c->have_kpot = true;
    c->have_volt = false;
    return 0;
}

int crt_bkfak(crt_ctx *c, double *kfak) {
    if (!c || !kfak) return CRT_ERR_ARG;
    if (!c->have_grid || !c->have_survey) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    int e = c->d_kfak.ensure(c->nanz);
    e = e ? e : c->d_status.ensure(1);
    if (e) return e;
    CRT_HIP(hipMemsetAsync(c->d_status.p, 0, sizeof(int), c->stream));
    if ((e = crt_launch_bkfak(c->stream, c->nanz, c->d_meas.p, c->d_enr.p, c->d_x.p, c->d_y.p, c->d_kfak.p,
                              c->d_status.p)))
        return e;
    int st = 0;
    CRT_HIP(c->copy_out(&st, c->d_status.p, sizeof(int)));
    if (st) return st;
    CRT_HIP(c->copy_out(kfak, c->d_kfak.p, (size_t)c->nanz * sizeof(double)));
    return 0;
}

int crt_set_fem_switches(crt_ctx *c, int lsr, int lana, int nsink) {
    if (!c) return CRT_ERR_ARG;
    if (!c->have_grid) return CRT_ERR_STATE;
    if (nsink != 0 && (nsink < 1 || nsink > c->sanz)) return CRT_ERR_ARG;
    c->sw_lsr = lsr ? 1 : 0;
    c->sw_lana = lana ? 1 : 0;
    c->sw_nsink = nsink - 1;
    c->have_kpot = c->have_volt = false;
    return 0;
}

int crt_forward(crt_ctx *c) {
    if (!c) return CRT_ERR_ARG;
    if (!c->have_precal || !c->have_sigma) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    if (int e0 = build_dmask(c)) return e0;
    // crmod.cfg's switches (src/fem.f90:191-203, 311): lsr only with mixed boundaries,
    // lsr switches lana off; both belong to CRMod's complex forward
    const bool lsr = c->sw_lsr && c->lbeta, lana = c->sw_lana && !c->sw_lsr;
    const bool gen = lsr || c->sw_nsink >= 0;
    if ((c->sw_lsr || c->sw_lana) && c->inv_kind != 0) return CRT_ERR_UNSUPPORTED;
    if (lana) return forward_analytic(c);
    // the condensed solve where the grid has mixed boundaries and no Dirichlet rows
    const bool cond_ok = c->fwd_method != 0 && c->lbeta && !c->dmask_any && c->asplit && c->fwd_bits == 64 && !gen;
    if (cond_ok) {
        if (int e0 = cond_plan(c)) return e0;
    }
    const bool cond = cond_ok && c->cond.ok;
    if (c->fwd_method == 1 && !cond) return CRT_ERR_UNSUPPORTED;
    if (cond) {
        c->refine_sweeps = 0;
        c->refine_ratio = 0.0;
        c->refine_fallbacks = 0;
        c->last_method = 1;
        c->have_kpot = false;
        if (int e0 = tune_pieces(c)) return e0;
        if (int e0 = forward_condensed(c)) return e0;
        c->have_kpot = true;
        c->have_volt = false;
        return 0;
    }
    c->last_method = 0;
    c->times[T_SCHUR] = c->times[T_PRODUCT] = c->times[T_PREFIX] = 0.0;
    // 16x16 matrix-core tiles when the band fits (mb <= 256), else 8x8 LDS tiles
    const bool use16 = crt_factor16_supported(c->mb) && !getenv("CRT_FACTOR8");
    const int NBt = use16 ? 16 : CRT_NB;
    const int K = c->kwnanz, E = c->eanz;
    const int nblk = (c->sanz + NBt - 1) / NBt, D = std::max(1, (c->mb + NBt - 1) / NBt);
    if (!use16 && crt_factor_lds_bytes(D) > 160 * 1024) return CRT_ERR_UNSUPPORTED;
    const bool fp32 = c->fwd_bits == 32;
    if (fp32 && !(use16 && crt_factor32_supported(c->mb) && c->tmapD == D)) return CRT_ERR_UNSUPPORTED;
    // general right-hand sides run through the 16-row FP64 substitution, without
    // Dirichlet rows (the reference applies randb / randb2 to them as well)
    if (gen && (!use16 || fp32 || c->dmask_any)) return CRT_ERR_UNSUPPORTED;
    // systems to factorise: every (l,k) with mixed boundaries, else l = 1 per k
    // (src/fem.f90:355: lbeta.or.l.eq.1)
    std::vector<int2> sys;
    std::vector<int4> rhs;
    for (int k = 0; k < K; ++k) {
        if (c->lbeta) {
            for (int l = 0; l < E; ++l) {
                rhs.push_back(int4{l, k, (int)sys.size(), 1});
                sys.push_back(int2{l, k});
            }
        } else {
            int fs = (int)sys.size();
            sys.push_back(int2{0, k});
            for (int l = 0; l < E; ++l) rhs.push_back(int4{l, k, fs, l == 0 ? 1 : 0});
        }
    }
    const size_t fstride = (size_t)nblk * NBt;
    const size_t panel_elems = (size_t)nblk * (D + 1) * NBt * NBt;
    size_t per_sys = panel_elems * sizeof(zd) + fstride * (2 * sizeof(zd) + 2 * sizeof(double)) +
                     (size_t)c->nnz * sizeof(zd);
    size_t batch = std::max<size_t>(1, std::min(sys.size(), c->panel_budget / per_sys));
    if (use16 && !fp32 && batch < sys.size()) {
        // one workgroup per system: a launch takes ceil(nb / resident) rounds, so batches
        // that are whole multiples of the resident count leave no partly filled round
        // but the last (C3: 1344 systems as 512 + 512 + 320 instead of 672 + 672)
        const size_t res = (size_t)crt_factor16_resident(D, 0);
        if (res > 0 && batch >= res) batch = batch / res * res;
    }
    int e = c->d_panels.ensure(batch * panel_elems);
    e = e ? e : c->d_yf.ensure(batch * fstride);
    e = e ? e : c->d_dinv.ensure(batch * fstride);
    e = e ? e : c->d_fak.ensure(batch * fstride);
    e = e ? e : c->d_dga.ensure(batch * fstride);
    e = e ? e : c->d_vals.ensure(batch * ((size_t)c->nnz + 2));
    e = e ? e : c->d_kpotN.ensure((size_t)c->sanz * K * E);
    e = e ? e : c->d_sys.ensure(sys.size());
    e = e ? e : c->d_rhs.ensure(rhs.size());
    e = e ? e : c->d_status.ensure(sys.size() + 1);
    if (e) return e;
    size_t nyscr = c->lbeta ? 1 : rhs.size();
    // refinement: per right-hand side of a batch the potentials x and the residual
    const size_t rhs_batch = c->lbeta ? batch : std::min(rhs.size(), batch * (size_t)E);
    if (gen) {
        nyscr = std::max(nyscr, rhs_batch);
        if ((e = build_tcol(c))) return e;
        if ((e = c->d_sigma0.ensure(1))) return e;
        if (lsr && (e = c->d_pota.ensure(rhs_batch * (size_t)c->sanz))) return e;
        if (lsr && (e = crt_launch_refsig(c->stream, c->elanz, c->maxn, c->d_etyp.p, c->d_nrel.p, c->d_x.p, c->d_y.p,
                                          c->d_sigma.p, c->d_sigma0.p)))
            return e;
    }
    if (fp32) {
        nyscr = std::max(nyscr, rhs_batch);
        if ((e = c->d_xacc.ensure(rhs_batch * fstride))) return e;
        if ((e = c->d_rstat.ensure(2 * rhs_batch))) return e;
    }
    if ((e = c->d_yscr.ensure(nyscr * fstride))) return e;
    CRT_HIP(hipMemcpyAsync(c->d_sys.p, sys.data(), sys.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream));
    CRT_HIP(hipMemsetAsync(c->d_status.p, 0, (sys.size() + 1) * sizeof(int), c->stream));
    DevGrid g = c->devgrid();
    g.nblk = nblk;
    g.D = D;
    g.fstride = (int)fstride;
    if (c->asplit) {
        if ((e = c->d_area.ensure((size_t)K * c->nnz))) return e;
        if ((e = crt_launch_area(c->stream, g, K, c->d_area.p))) return e;
        g.area = c->d_area.p;
    }
    double tdiag = 0, tfac = 0, tsol = 0, tref = 0;
    c->refine_sweeps = 0;
    c->refine_ratio = 0.0;
    c->refine_fallbacks = 0;
    for (size_t s0 = 0; s0 < sys.size(); s0 += batch) {
        int nb = (int)std::min(batch, sys.size() - s0);
        // right-hand sides of this batch, factor index relative to the batch
        std::vector<int4> r;
        for (auto q : rhs)
            if ((size_t)q.z >= s0 && (size_t)q.z < s0 + nb) r.push_back(int4{q.x, q.y, q.z - (int)s0, q.w});
        CRT_HIP(hipMemcpyAsync(c->d_rhs.p, r.data(), r.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream));
        c->tic(T_DIAG);
        if ((e = crt_launch_diag(c->stream, g, nb, c->d_sys.p + s0, c->d_fak.p, c->d_dga.p, c->d_status.p + 1 + s0)))
            return e;
        if ((e = crt_launch_assemble(c->stream, g, nb, c->d_sys.p + s0, c->d_fak.p, c->d_dga.p, c->d_vals.p)))
            return e;
        c->toc(T_DIAG);
        bool done = false;
        if (fp32) {
            int sweeps = 0;
            double ratio = 0.0, tf = 0.0, ts = 0.0, tr = 0.0;
            e = forward_batch_f32(c, g, nb, s0, (int)r.size(), D, sweeps, ratio, tf, ts, tr);
            if (e > 0) return e;
            tfac += tf; tsol += ts; tref += tr;
            c->refine_sweeps = std::max(c->refine_sweeps, sweeps);
            if (e == 0) {
                c->refine_ratio = std::max(c->refine_ratio, ratio);
                done = true;
            } else {
                // no convergence (or an FP32 pivot failure): this batch again in FP64
                c->refine_fallbacks++;
                CRT_HIP(hipMemsetAsync(c->d_status.p + 1 + s0, 0, nb * sizeof(int), c->stream));
            }
        }
        if (!done) {
            c->tic(T_FACTOR);
            if (use16) {
                e = crt_launch_factor16(c->stream, g, nb, c->d_sys.p + s0, c->d_fak.p, c->d_vals.p, c->d_panels.p,
                                        c->d_yf.p, c->d_dinv.p, c->d_status.p + 1 + s0, 0);
                if (!e && g.dc)
                    e = crt_launch_dc_pivot(c->stream, nb, c->sanz, (int)fstride, c->d_dinv.p, c->d_status.p + 1 + s0);
            }
            else
                e = crt_launch_factor(c->stream, g, nb, c->d_sys.p + s0, c->d_fak.p, c->d_dga.p, c->d_vals.p,
                                      c->d_panels.p, c->d_yf.p, c->d_status.p + 1 + s0);
            if (e) return e;
            c->toc(T_FACTOR);
            c->tic(T_SOLVE);
            if (gen) {
                // the singularity removal's or the sink's right-hand side, the substitution
                // from it, and the analytic part added (src/fem.f90:350-389)
                const int nr = (int)r.size();
                if (lsr)
                    e = crt_launch_potana(c->stream, c->sanz, nr, c->d_rhs.p, c->d_enr.p, c->d_x.p, c->d_y.p,
                                          c->d_kwn.p, c->d_sigma0.p, c->d_pota.p, c->sanz);
                e = e ? e : crt_launch_gen_rhs(c->stream, g, nr, c->d_rhs.p, lsr ? 1 : 0, c->d_sigma0.p, c->d_pota.p,
                                               c->sanz, c->d_tptr.p, c->d_tq.p, c->sw_nsink, c->d_fak.p,
                                               (int)fstride, c->d_yscr.p);
                e = e ? e : crt_launch_solve16_rhs(c->stream, c->sanz, nblk, D, K, E, nr, c->d_rhs.p, c->d_enr.p,
                                                   c->d_fak.p, c->d_panels.p, c->d_dinv.p, c->d_yscr.p,
                                                   c->d_kpotN.p, crt_factor16_resolve(D, 0) != 3);
                if (!e && lsr)
                    e = crt_launch_pota_kpot(c->stream, c->sanz, K, E, nr, c->d_rhs.p, c->d_pota.p, c->sanz, 1,
                                             c->d_kpotN.p);
            } else if (use16)
                e = crt_launch_solve16(c->stream, c->sanz, nblk, D, K, E, (int)r.size(), c->d_rhs.p, c->d_enr.p,
                                       c->d_fak.p, c->d_panels.p, c->d_dinv.p, c->d_yf.p, c->d_yscr.p, c->d_kpotN.p,
                                       crt_factor16_resolve(D, 0) != 3);
            else
                e = crt_launch_solve(c->stream, c->sanz, nblk, D, K, E, (int)r.size(), c->d_rhs.p, c->d_enr.p,
                                     c->d_fak.p, c->d_panels.p, c->d_yf.p, c->d_yscr.p, c->d_kpotN.p, nullptr, nullptr);
            if (e) return e;
            c->toc(T_SOLVE);
        }
        CRT_HIP(hipStreamSynchronize(c->stream));
        c->collect(T_DIAG); tdiag += c->times[T_DIAG];
        if (!done) {
            c->collect(T_FACTOR); tfac += c->times[T_FACTOR];
            c->collect(T_SOLVE); tsol += c->times[T_SOLVE];
        }
    }
    c->times[T_REFINE] = tref;
    c->times[T_DIAG] = tdiag; c->times[T_FACTOR] = tfac; c->times[T_SOLVE] = tsol;
    std::vector<int> st(sys.size() + 1);
    CRT_HIP(c->copy_out(st.data(), c->d_status.p, st.size() * sizeof(int)));
    // the first failing system in the reference's loop order (wavenumber, then
    // electrode: src/fem.f90:328-333) names the error
    int worst = st[0];
    for (size_t q = 1; q < st.size() && !worst; ++q) worst = st[q];
    if (worst && getenv("CRT_DEBUG")) {
        for (size_t q = 1; q < st.size(); ++q)
            if (st[q]) {
                fprintf(stderr, "crt_forward: system %zu (l=%d k=%d) status %d\n", q - 1, sys[q - 1].x, sys[q - 1].y,
                        st[q]);
                if (q - 1 < batch) {
                    std::vector<double> dg(c->sanz);
                    (void)c->copy_out(dg.data(), c->d_dga.p + (q - 1) * fstride, c->sanz * sizeof(double));
                    for (int i = 0; i < c->sanz; ++i)
                        if (!(dg[i] > 0.0) || dg[i] == 1.0) {
                            fprintf(stderr, "  row %d |a_ii| = %g\n", i, dg[i]);
                            break;
                        }
                }
                break;
            }
    }
    if (worst) return worst;
    if ((e = zero_cleared_sources(c))) return e;
    c->have_kpot = true;
    c->have_volt = false;
    return 0;
}

int crt_get_kpot(crt_ctx *c, double *kpot) {
    if (!c || !kpot) return CRT_ERR_ARG;
    if (!c->have_kpot) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    size_t n = (size_t)c->sanz * c->eanz * c->kwnanz;
    int e = c->d_tmp.ensure(n);
    if (e) return e;
    if ((e = crt_launch_kpot_ref(c->stream, c->sanz, c->eanz, c->kwnanz, c->d_kpotN.p, c->d_tmp.p))) return e;
    return c->copy_out_kind(kpot, c->d_tmp.p, n);
}

int crt_voltages(crt_ctx *c, int mode, const double *kfak, double *volt, double *sigmaa) {
    if (!c || (mode != 0 && mode != 1)) return CRT_ERR_ARG;
    if (!c->have_kpot || !c->have_survey) return CRT_ERR_STATE;
    if (c->inv_kind == 1 && mode == 0) return CRT_ERR_UNSUPPORTED;  // bvolt has no DC branch
    CRT_HIP(hipSetDevice(c->device));
    int e = c->d_hpotE.ensure((size_t)c->eanz * c->eanz);
    e = e ? e : c->d_volt.ensure(c->nanz);
    e = e ? e : c->d_sigmaa.ensure(c->nanz);
    if (e) return e;
    const double *dk = nullptr;
    if (kfak && mode == 0) {
        if ((e = c->d_kfak.upload(kfak, c->nanz, c->stream))) return e;
        dk = c->d_kfak.p;
    }
    if ((e = c->status_reset())) return e;
    c->tic(T_VOLT);
    e = crt_launch_hpot_elec(c->stream, c->eanz, c->kwnanz, c->d_enr.p, c->d_kwnwi.p, c->d_kpotN.p, c->d_hpotE.p);
    // mode 2 of the kernel: bvolti's DC branch, volt = dcmplx(real superposition)
    e = e ? e : crt_launch_volt(c->stream, c->nanz, c->eanz, c->inv_kind == 1 ? 2 : mode, c->d_meas.p, c->d_hpotE.p,
                                dk, c->d_volt.p,
                                c->d_sigmaa.p, c->d_status.p);
    c->toc(T_VOLT);
    if (e) return e;
    if (volt) CRT_HIP(hipMemcpyAsync(volt, c->d_volt.p, c->nanz * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
    if (sigmaa) CRT_HIP(hipMemcpyAsync(sigmaa, c->d_sigmaa.p, c->nanz * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->collect(T_VOLT);
    int st = c->status_check();
    c->have_volt = true;
    return st;
}

int crt_get_hpot(crt_ctx *c, double *hpot) {
    if (!c || !hpot) return CRT_ERR_ARG;
    if (!c->have_kpot) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    size_t n = (size_t)c->sanz * c->eanz;
    int e = c->d_tmp.ensure(n);
    if (e) return e;
    if ((e = crt_launch_hpot_full(c->stream, c->sanz, c->eanz, c->kwnanz, c->d_kwnwi.p, c->d_kpotN.p, c->d_tmp.p)))
        return e;
    return c->copy_out_kind(hpot, c->d_tmp.p, n);
}

int crt_sens(crt_ctx *c, int mode, const double *strom) {
    if (!c || (mode != 0 && mode != 1)) return CRT_ERR_ARG;
    if (!c->have_kpot || !c->have_survey) return CRT_ERR_STATE;
    if (mode == 1 && !c->have_volt) return CRT_ERR_STATE;
    if (c->inv_kind == 1 && mode == 0) return CRT_ERR_UNSUPPORTED;  // the DC twin is bsendc (of bsensi)
    CRT_HIP(hipSetDevice(c->device));
    int e = c->d_sens.ensure((size_t)c->nanz * c->elanz);
    e = e ? e : c->d_vfac.ensure(c->nanz);
    if (e) return e;
    const double *ds = nullptr;
    if (strom && mode == 0) {
        if ((e = c->d_strom.upload(strom, c->nanz, c->stream))) return e;
        ds = c->d_strom.p;
    }
    c->tic(T_SENS);
    e = crt_launch_sens(c->stream, c->elanz, c->nanz, c->eanz, c->kwnanz, c->maxn, c->npk, mode, c->d_nrel.p,
                        c->d_nkel.p, c->d_meas.p, c->d_kwnwi.p, c->d_kpotN.p, c->d_elbg.p, c->d_sigma.p,
                        c->d_volt.p ? c->d_volt.p : nullptr, ds, c->d_sens.p, c->d_vfac.p);
    c->toc(T_SENS);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->collect(T_SENS);
    c->have_sens = true;
    c->sens32_ok = false;
    c->sensr_ok = false;
    c->sensr32_ok = false;
    c->sens_nanz = c->nanz;
    c->sens_manz = c->elanz;
    return 0;
}

int crt_get_sens(crt_ctx *c, double *sens) {
    if (!c || !sens) return CRT_ERR_ARG;
    if (!c->have_sens) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)(c->sens_nanz ? c->sens_nanz : c->nanz) * (c->sens_manz ? c->sens_manz : c->elanz);
    return c->copy_out_kind(sens, c->d_sens.p, n);
}

// ------------------------------------------------------------- normal equations
int crt_set_sens(crt_ctx *c, int nanz, int manz, const double *sens) {
    if (!c || !sens || nanz <= 0 || manz <= 0) return CRT_ERR_ARG;
    CRT_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)nanz * manz;
    const zd *zs = reinterpret_cast<const zd *>(sens);
    std::vector<zd> ex;
    if (c->inv_kind == 1) {
        // DC: a real sensdc(nanz, manz)
        ex.resize(n);
        for (size_t t = 0; t < n; ++t) ex[t] = zd{sens[t], 0.0};
        zs = ex.data();
    }
    int e = c->d_sens.upload(zs, n, c->stream);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->sens_nanz = nanz;
    c->sens_manz = manz;
    c->have_sens = true;
    c->sens32_ok = false;
    c->sensr_ok = false;
    c->sensr32_ok = false;
    return 0;
}

int crt_set_regularization(crt_ctx *c, int ltri, int nx, int manz, int maxn, const double *smatm,
                           const int *nachbar) {
    if (!c || !smatm || manz <= 0) return CRT_ERR_ARG;
    if (ltri != 0 && ltri != 1 && ltri != 2) return CRT_ERR_UNSUPPORTED;
    if ((ltri == 1 || ltri == 2) && (!nachbar || maxn <= 0)) return CRT_ERR_ARG;
    if (ltri == 0 && nx <= 0) return CRT_ERR_ARG;
    CRT_HIP(hipSetDevice(c->device));
    int ncol = ltri == 0 ? 3 : maxn + 1;
    std::vector<double> rd(manz);
    for (int j = 0; j < manz; ++j) rd[j] = smatm[(size_t)(ltri == 0 ? 0 : maxn) * manz + j];
    int e = c->d_smatm.upload(smatm, (size_t)manz * ncol, c->stream);
    e = e ? e : c->d_rdiag.upload(rd.data(), manz, c->stream);
    if (!e && ltri != 0) e = c->d_nachbar.upload(nachbar, (size_t)manz * (maxn + 1), c->stream);
    if (ltri != 0) c->h_nachbar.assign(nachbar, nachbar + (size_t)manz * (maxn + 1));
    if (!e && ltri == 0) e = c->d_nachbar.ensure(1);
    if (e) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->ltri = ltri;
    c->rnx = nx;
    c->rncol = ncol;
    c->reg_manz = manz;
    return 0;
}

int crt_bsmatm(crt_ctx *c, int ltri, int nx, int nz, double alfx, double alfz, double *smatm) {
    if (!c) return CRT_ERR_ARG;
    if (!c->have_grid) return CRT_ERR_STATE;
    if (ltri != 0 && ltri != 1 && ltri != 2) return CRT_ERR_UNSUPPORTED;
    CRT_HIP(hipSetDevice(c->device));
    const int manz = c->elanz, maxn = c->maxn;
    int e = 0, ncol;
    if (ltri == 0) {
        // regular grid: quads numbered row by row from the surface
        if (nx <= 0 || nz <= 0 || (long)nx * nz != manz || maxn < 4) return CRT_ERR_ARG;
        for (int i = 0; i < manz; ++i)
            if (c->nkel[i] != 4) return CRT_ERR_UNSUPPORTED;
        ncol = 3;
        e = c->d_smatm.ensure((size_t)manz * ncol);
        e = e ? e : c->d_nachbar.ensure(1);
        e = e ? e : crt_launch_bsmatmreg(c->stream, nx, nz, manz, maxn, c->d_nrel.p, c->d_x.p, c->d_y.p, alfx, alfz,
                                         c->d_smatm.p);
    } else {
        // every edge of every element is a node pair of its own: one element type
        for (int i = 0; i < manz; ++i)
            if (c->nkel[i] != maxn) return CRT_ERR_UNSUPPORTED;
        ncol = maxn + 1;
        std::vector<int> nb((size_t)manz * (maxn + 1));
        e = crt_bnachbar(manz, maxn, c->nrel.data(), nb.data());
        e = e ? e : c->d_nachbar.upload(nb.data(), nb.size(), c->stream);
        c->h_nachbar = nb;
        e = e ? e : c->d_smatm.ensure((size_t)manz * ncol);
        e = e ? e : crt_launch_bsmatmtri(c->stream, manz, maxn, c->d_nrel.p, c->d_nkel.p, c->d_x.p, c->d_y.p,
                                         c->d_nachbar.p, alfx, alfz, c->d_smatm.p);
    }
    e = e ? e : c->d_rdiag.ensure(manz);
    if (e) return e;
    CRT_HIP(hipMemcpyAsync(c->d_rdiag.p, c->d_smatm.p + (size_t)(ltri == 0 ? 0 : maxn) * manz, manz * sizeof(double),
                           hipMemcpyDeviceToDevice, c->stream));
    if (smatm)
        CRT_HIP(hipMemcpyAsync(smatm, c->d_smatm.p, (size_t)manz * ncol * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->ltri = ltri;
    c->rnx = nx;
    c->rncol = ncol;
    c->reg_manz = manz;
    return 0;
}

// Sizes of the resident Jacobian, and the check that every buffer the normal-equation
// kernels index by (measurement, cell) covers them: J holds nanz*manz entries and, when
// the regularisation is used, smatm / rdiag / nachbar were installed for the same manz
// cells.  A mismatch (J and R set for different grids) is refused before any launch.
static int jsizes(const crt_ctx *c, bool need_reg, int *nanz, int *manz) {
    const int n = c->sens_nanz ? c->sens_nanz : c->nanz, m = c->sens_manz ? c->sens_manz : c->elanz;
    if (n <= 0 || m <= 0 || (size_t)n * m > c->d_sens.n) return CRT_ERR_STATE;
    if (need_reg) {
        if (c->ltri < 0 || c->reg_manz != m || (size_t)m * c->rncol > c->d_smatm.n || (size_t)m > c->d_rdiag.n)
            return CRT_ERR_STATE;
        if (c->ltri != 0 && (size_t)m * c->rncol > c->d_nachbar.n) return CRT_ERR_STATE;
    }
    *nanz = n;
    *manz = m;
    return 0;
}

static int cg_prepare(crt_ctx *c, int nanz, int manz, int ncgmax, double lam, double eps, CGArgs &a) {
    int ncb, cb, n_dr1, n_rr;
    crt_cg_geometry(nanz, manz, &ncb, &cb, &n_dr1, &n_rr);
    int e = 0;
    e = e ? e : c->d_r.ensure(manz);
    e = e ? e : c->d_p.ensure(manz);
    e = e ? e : c->d_q.ensure(manz);
    e = e ? e : c->d_bv.ensure(manz);
    e = e ? e : c->d_dpar.ensure(manz);
    e = e ? e : c->d_ap.ensure(nanz);
    e = e ? e : c->d_part_ap.ensure((size_t)ncb * nanz);
    e = e ? e : c->d_part_dr1.ensure(n_dr1);
    e = e ? e : c->d_part_rr.ensure(n_rr);
    e = e ? e : c->d_cgres.ensure((size_t)ncgmax + 1);
    e = e ? e : c->d_sc.ensure(2);
    if (e) return e;
    CRT_HIP(hipMemsetAsync(c->d_cgres.p, 0, ((size_t)ncgmax + 1) * sizeof(double), c->stream));
    a = CGArgs{};
    a.nanz = nanz; a.manz = manz; a.ltri = c->ltri; a.nx = c->rnx; a.maxn = c->rncol - 1; a.ncgmax = ncgmax;
    a.lam = lam; a.eps = eps;
    a.J = c->d_sens.p; a.w = c->d_w.p; a.smatm = c->d_smatm.p; a.nachbar = c->d_nachbar.p;
    a.J32 = nullptr;
    a.Jr = nullptr;
    a.Jr32 = nullptr;
    a.part = c->inv_kind;
    if (c->inv_kind != 0) {
        // DC / FPI: the reference's Jacobi-scaled CG on a real J (src/cg_mod.f90:106-196);
        // the FP32 copy and the FSAI patterns as for the complex inversion (the learned
        // preconditioner is trained on complex systems only)
        if (c->cg_prec >= 3) return CRT_ERR_UNSUPPORTED;
        const size_t n = (size_t)nanz * manz;
        if (!c->sensr_ok) {
            if ((e = c->d_sensr.ensure(n))) return e;
            if ((e = crt_launch_real_part(c->stream, n, c->d_sens.p, c->d_sensr.p))) return e;
            c->sensr_ok = true;
            c->sensr32_ok = false;
        }
        a.Jr = c->d_sensr.p;
        if (c->cg_jbits == 32) {
            if (!c->sensr32_ok) {
                if ((e = c->d_sensr32.ensure(n))) return e;
                if ((e = crt_launch_cg_jr32(c->stream, n, c->d_sensr.p, c->d_sensr32.p))) return e;
                c->sensr32_ok = true;
            }
            a.Jr32 = c->d_sensr32.p;
        }
    } else if (c->cg_jbits == 32) {
        const size_t n = (size_t)nanz * manz;
        if (!c->sens32_ok) {
            if ((e = c->d_sens32.ensure(n))) return e;
            if ((e = crt_launch_cg_j32(c->stream, n, c->d_sens.p, c->d_sens32.p))) return e;
            c->sens32_ok = true;
        }
        a.J32 = c->d_sens32.p;
    }
    a.fg = nullptr; a.fpat = nullptr; a.ftp = nullptr; a.t = nullptr; a.z = nullptr; a.part_rz = nullptr;
    a.cgfac = c->d_cgfac.p;
    a.r = c->d_r.p; a.p = c->d_p.p; a.q = c->d_q.p; a.bv = c->d_bv.p; a.dpar = c->d_dpar.p; a.ap = c->d_ap.p;
    a.part_ap = c->d_part_ap.p; a.part_dr1 = c->d_part_dr1.p; a.part_rr = c->d_part_rr.p;
    a.n_dr1 = n_dr1; a.n_rr = n_rr; a.ncb = ncb; a.cb = cb;
    a.sc = c->d_sc.p; a.scn = c->d_sc.p; a.cgres = c->d_cgres.p;
    return 0;
}

// Fletcher-Reeves iterations on the device; the host looks at the stop flag
// every cg_check_every iterations.
// FSAI pattern (row j: lower-triangular columns, j itself last) and its transpose
static void fsai_pattern(const crt_ctx *c, int manz, int kind, std::vector<int> &pat, std::vector<int> &tp) {
    constexpr int NP = CRT_FSAI_NP;
    pat.assign((size_t)manz * NP, -1);
    tp.assign((size_t)manz * NP, -1);
    std::vector<int> nin(manz, 0);
    for (int j = 0; j < manz; ++j) {
        std::vector<int> cols;
        if (c->ltri == 0) {
            const int nx = c->rnx, ix = j % nx, iz = j / nx;
            auto add = [&](int dz, int dx) {
                const int x = ix + dx, z = iz + dz;
                if (x >= 0 && x < nx && z >= 0) cols.push_back(z * nx + x);
            };
            if (kind == 2) {
                add(-2, 0); add(-1, -1); add(-1, 0); add(-1, 1); add(0, -2); add(0, -1);
            } else {
                add(-1, 0); add(0, -1);
            }
        } else {
            const int mx = (int)(c->h_nachbar.size() / manz) - 1;
            for (int ij = 0; ij < mx; ++ij) {
                const int nb = c->h_nachbar[(size_t)ij * manz + j] - 1;
                if (nb >= 0 && nb < j && std::find(cols.begin(), cols.end(), nb) == cols.end()) cols.push_back(nb);
            }
            std::sort(cols.begin(), cols.end());
        }
        cols.push_back(j);
        const int off = NP - (int)cols.size();
        for (size_t s = 0; s < cols.size(); ++s) pat[(size_t)j * NP + off + s] = cols[s];
    }
    for (int j = 0; j < manz; ++j)
        for (int s = 0; s < NP; ++s) {
            const int col = pat[(size_t)j * NP + s];
            if (col >= 0 && nin[col] < NP) tp[(size_t)col * NP + nin[col]++] = j * NP + s;
        }
}

// build G of M = G^H G for the current cgfac and lambda (after the RHS pass)
static int pc_features(crt_ctx *c, const CGArgs &a);

static int cg_fsai(crt_ctx *c, CGArgs &a) {
    c->last_lam = a.lam;
    c->have_cgsys = true;
    if (c->cg_prec == 0) return 0;
    const int manz = a.manz;
    const bool learned = c->cg_prec >= 3, prior_only = c->cg_prec == 4;
    if (learned && ((!prior_only && !c->pc_ready) || c->ltri != 0)) return CRT_ERR_STATE;
    std::vector<int> pat, tp;
    fsai_pattern(c, manz, learned ? 1 : c->cg_prec, pat, tp);
    int e = c->d_fpat.upload(pat.data(), pat.size(), c->stream);
    e = e ? e : c->d_ftp.upload(tp.data(), tp.size(), c->stream);
    e = e ? e : c->d_fg.ensure((size_t)manz * CRT_FSAI_NP);
    e = e ? e : c->d_t.ensure(manz);
    e = e ? e : c->d_z.ensure(manz);
    e = e ? e : c->d_part_rz.ensure(a.n_rr);
    if (e) return e;
    a.fpat = c->d_fpat.p;
    a.ftp = c->d_ftp.p;
    if (learned) {
        if ((e = pc_features(c, a))) return e;
        e = crt_launch_pc_infer(c->stream, manz, c->rnx, c->d_pc_par.p, c->d_pc_norm.p, c->d_pc_feat.p, c->d_fpat.p,
                                c->d_fg.p, prior_only ? 1 : 0);
    } else {
        e = crt_launch_cg_fsai(c->stream, a, c->d_fg.p);
    }
    if (e) return e;
    a.fg = c->d_fg.p;
    a.t = c->d_t.p;
    a.z = c->d_z.p;
    a.part_rz = c->d_part_rz.p;
    return 0;
}

// features of the cells of the current system (cgfac, lambda, R, depth)
static int pc_features(crt_ctx *c, const CGArgs &a) {
    const int manz = a.manz;
    int e = 0;
    {
        // cell centroid depth coordinate (mean of the element's node y)
        std::vector<double> cy(manz);
        for (int i = 0; i < manz; ++i) {
            double sy = 0.0;
            for (int l = 0; l < c->nkel[i]; ++l) sy += c->y[c->nrel[(size_t)i * c->maxn + l] - 1];
            cy[i] = sy / c->nkel[i];
        }
        c->cy_min = *std::min_element(cy.begin(), cy.end());
        c->cy_max = *std::max_element(cy.begin(), cy.end());
        if ((e = c->d_cy.upload(cy.data(), manz, c->stream))) return e;
    }
    if ((e = c->d_pc_feat.ensure((size_t)manz * 8))) return e;
    if ((e = c->d_pc_cpl.ensure((size_t)manz * 3))) return e;
    PcCell pc{manz, c->rnx, manz / std::max(c->rnx, 1), a.nanz, a.lam, a.cgfac, a.smatm, c->d_cy.p, c->cy_min,
              c->cy_max, a.J, a.w};
    return crt_launch_pc_features(c->stream, pc, c->d_pc_cpl.p, c->d_pc_feat.p);
}

static void pc_init_params(std::vector<float> &par) {
    // Xavier-uniform weights from a fixed linear congruential sequence, zero biases
    const int NI = 8, NH = 64, NO = 5;
    par.assign(crt_pc_nparams(), 0.f);
    unsigned long long st = 0x2545F4914F6CDD1DULL;
    auto uni = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    size_t o = 0;
    const int shp[4][2] = {{NH, NI}, {NH, NH}, {NH, NH}, {NO, NH}};
    for (int l = 0; l < 4; ++l) {
        const int no = shp[l][0], ni = shp[l][1];
        const double a = std::sqrt(6.0 / (ni + no)) * (l == 3 ? 0.1 : 1.0);
        for (int q = 0; q < no * ni; ++q) par[o + q] = (float)(a * uni());
        o += (size_t)no * ni + no;
    }
}

// The two scalar states alternate between a.scn[0] and a.scn[1]; the residual history
// goes to a.cgres (both set by the caller: context buffers or the static scratch).
static int cg_iterate(crt_ctx *c, CGArgs a, const zd *bvec, int *ncg) {
    int e;
    CGScalars *S = a.scn;
    CGArgs a0 = a, a1 = a;
    a0.sc = S; a0.scn = S + 1;
    a1.sc = S + 1; a1.scn = S;
    CGArgs ai = a;
    ai.scn = S;  // init writes state 0
    c->tic(T_CG);
    if ((e = crt_launch_cg_init(c->stream, ai, bvec))) return e;
    CGScalars sc{};
    int it = 0;
    double tmv = 0.0;
    while (it < a.ncgmax) {
        int n = std::min(c->cg_check_every, a.ncgmax - it);
        c->tic(T_CGMV);
        for (int q = 0; q < n; ++q, ++it)
            if ((e = crt_launch_cg_iter(c->stream, (it & 1) ? a1 : a0))) return e;
        c->toc(T_CGMV);
        CRT_HIP(hipMemcpyAsync(&sc, S + (it & 1), sizeof(CGScalars), hipMemcpyDeviceToHost, c->stream));
        CRT_HIP(hipStreamSynchronize(c->stream));
        c->collect(T_CGMV);
        tmv += c->times[T_CGMV];
        if (sc.done) break;
    }
    c->toc(T_CG);
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->collect(T_CG);
    int done_its = sc.done ? sc.ncg : a.ncgmax;
    *ncg = done_its;
    // per-iteration device time of the batches (the last batch may include no-op launches)
    c->times[T_CGMV] = it > 0 ? tmv / it : 0.0;
    double ncgd = (double)(float)(*ncg);
    CRT_HIP(hipMemcpyAsync(a.cgres, &ncgd, sizeof(double), hipMemcpyHostToDevice, c->stream));
    CRT_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

// the learned preconditioner falls back to the reference's Jacobi-scaled CG when it
// does not reach the stop test within ncgmax iterations
static int cg_fallback(crt_ctx *c, const CGArgs &a, int ncgmax, int *n) {
    if (c->cg_prec < 3 || !a.fg || *n < ncgmax) return 0;
    CGArgs aj = a;
    aj.fg = nullptr;
    ++c->pc_fallbacks;
    return cg_iterate(c, aj, c->d_bvec.p, n);
}

static int cg_weights(crt_ctx *c, int nanz, const double *wmatd, const int *wdfak) {
    std::vector<double> w(nanz);
    for (int i = 0; i < nanz; ++i) w[i] = wmatd[i] * (double)wdfak[i];
    return c->d_w.upload(w.data(), nanz, c->stream);
}

int crt_set_cg_preconditioner(crt_ctx *c, int kind) {
    if (!c || kind < 0 || kind > 4) return CRT_ERR_ARG;
    c->cg_prec = kind;
    return 0;
}

int crt_pc_size(void) { return 16 + crt_pc_nparams(); }

int crt_pc_weights(crt_ctx *c, float *get, const float *set) {
    if (!c) return CRT_ERR_ARG;
    CRT_HIP(hipSetDevice(c->device));
    const int np = crt_pc_nparams();
    if (set) {
        int e = c->d_pc_norm.upload(set, 16, c->stream);
        e = e ? e : c->d_pc_par.upload(set + 16, np, c->stream);
        std::vector<float> zero(np, 0.f);
        e = e ? e : c->d_pc_m.upload(zero.data(), np, c->stream);
        e = e ? e : c->d_pc_v.upload(zero.data(), np, c->stream);
        if (e) return e;
        CRT_HIP(hipStreamSynchronize(c->stream));
        c->pc_steps = 0;
        c->pc_ready = c->pc_norm_set = true;
    }
    if (get) {
        if (!c->pc_ready) return CRT_ERR_STATE;
        CRT_HIP(c->copy_out(get, c->d_pc_norm.p, 16 * sizeof(float)));
        CRT_HIP(c->copy_out(get + 16, c->d_pc_par.p, np * sizeof(float)));
    }
    return 0;
}

int crt_pc_train(crt_ctx *c, int steps, double lr, double *loss_first, double *loss_last) {
    if (!c || steps < 1 || !(lr > 0.0)) return CRT_ERR_ARG;
    if (!c->have_cgsys || !c->have_sens || c->ltri != 0) return CRT_ERR_STATE;
    if (c->inv_kind != 0) return CRT_ERR_UNSUPPORTED;  // trained on complex systems only
    CRT_HIP(hipSetDevice(c->device));
    int nanz = 0, manz = 0;
    if (int e0 = jsizes(c, true, &nanz, &manz)) return e0;
    const int np = crt_pc_nparams();
    int e = 0;
    if (!c->pc_ready) {
        std::vector<float> par, zero(np, 0.f);
        pc_init_params(par);
        e = c->d_pc_par.upload(par.data(), np, c->stream);
        e = e ? e : c->d_pc_m.upload(zero.data(), np, c->stream);
        e = e ? e : c->d_pc_v.upload(zero.data(), np, c->stream);
        e = e ? e : c->d_pc_norm.ensure(16);
        if (e) return e;
        c->pc_steps = 0;
    }
    // teacher: the FSAI rows on the stencil pattern of the last system
    CGArgs a;
    if ((e = cg_prepare(c, nanz, manz, 1, c->last_lam, 1e-30, a))) return e;
    std::vector<int> pat, tp;
    fsai_pattern(c, manz, 1, pat, tp);
    e = c->d_fpat.upload(pat.data(), pat.size(), c->stream);
    e = e ? e : c->d_ftp.upload(tp.data(), tp.size(), c->stream);
    e = e ? e : c->d_fg.ensure((size_t)manz * CRT_FSAI_NP);
    if (e) return e;
    a.fpat = c->d_fpat.p;
    if ((e = crt_launch_cg_fsai(c->stream, a, c->d_fg.p))) return e;
    if ((e = pc_features(c, a))) return e;
    e = c->d_pc_tgt.ensure((size_t)manz * 5);
    e = e ? e : c->d_pc_mask.ensure((size_t)manz * 5);
    e = e ? e : c->d_pc_act.ensure((size_t)manz * (8 + 3 * 64));
    e = e ? e : c->d_pc_delta.ensure((size_t)manz * (3 * 64 + 5));
    e = e ? e : c->d_pc_grad.ensure(np);
    e = e ? e : c->d_pc_gpart.ensure(crt_pc_gpart_size());
    e = e ? e : c->d_pc_loss.ensure(crt_pc_blocks(manz));
    e = e ? e : crt_launch_pc_targets(c->stream, manz, c->rnx, c->d_fpat.p, c->d_fg.p, c->d_pc_feat.p, c->d_pc_tgt.p,
                                      c->d_pc_mask.p);
    if (e) return e;
    // append this system to the training set (at most 2^22 cells, oldest dropped)
    {
        const int cap = 1 << 22;
        int keep = std::min(c->pc_tn, cap - manz);
        if (keep < 0) keep = 0;
        const int drop = c->pc_tn - keep, tot = keep + manz;
        DBuf<float> nf, nt, nm;
        e = nf.ensure((size_t)tot * 8);
        e = e ? e : nt.ensure((size_t)tot * 5);
        e = e ? e : nm.ensure((size_t)tot * 5);
        if (e) return e;
        if (keep) {
            CRT_HIP(hipMemcpyAsync(nf.p, c->d_pc_tfeat.p + (size_t)drop * 8, (size_t)keep * 8 * sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
            CRT_HIP(hipMemcpyAsync(nt.p, c->d_pc_ttgt.p + (size_t)drop * 5, (size_t)keep * 5 * sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
            CRT_HIP(hipMemcpyAsync(nm.p, c->d_pc_tmask.p + (size_t)drop * 5, (size_t)keep * 5 * sizeof(float),
                                   hipMemcpyDeviceToDevice, c->stream));
        }
        CRT_HIP(hipMemcpyAsync(nf.p + (size_t)keep * 8, c->d_pc_feat.p, (size_t)manz * 8 * sizeof(float),
                               hipMemcpyDeviceToDevice, c->stream));
        CRT_HIP(hipMemcpyAsync(nt.p + (size_t)keep * 5, c->d_pc_tgt.p, (size_t)manz * 5 * sizeof(float),
                               hipMemcpyDeviceToDevice, c->stream));
        CRT_HIP(hipMemcpyAsync(nm.p + (size_t)keep * 5, c->d_pc_mask.p, (size_t)manz * 5 * sizeof(float),
                               hipMemcpyDeviceToDevice, c->stream));
        CRT_HIP(hipStreamSynchronize(c->stream));
        std::swap(c->d_pc_tfeat.p, nf.p); std::swap(c->d_pc_tfeat.n, nf.n);
        std::swap(c->d_pc_ttgt.p, nt.p); std::swap(c->d_pc_ttgt.n, nt.n);
        std::swap(c->d_pc_tmask.p, nm.p); std::swap(c->d_pc_tmask.n, nm.n);
        c->pc_tn = tot;
    }
    const int ntr = c->pc_tn;
    e = c->d_pc_act.ensure((size_t)ntr * (8 + 3 * 64));
    e = e ? e : c->d_pc_delta.ensure((size_t)ntr * (3 * 64 + 5));
    e = e ? e : c->d_pc_loss.ensure(crt_pc_blocks(ntr));
    if (e) return e;
    c->pc_manz = ntr;
    if (!c->pc_norm_set) {
        if ((e = crt_launch_pc_norm(c->stream, manz, c->d_pc_feat.p, c->d_pc_norm.p))) return e;
        c->pc_norm_set = true;
    }
    const int nb = crt_pc_blocks(ntr);
    std::vector<double> pl(nb);
    auto loss = [&](double *out) -> int {
        if (!out) return 0;
        if (c->copy_out(pl.data(), c->d_pc_loss.p, nb * sizeof(double)) != hipSuccess) return CRT_ERR_HIP;
        double t = 0.0;
        for (double v : pl) t += v;
        *out = t;
        return 0;
    };
    for (int it = 0; it < steps; ++it) {
        e = crt_launch_pc_fwdbwd(c->stream, ntr, c->d_pc_par.p, c->d_pc_norm.p, c->d_pc_tfeat.p, c->d_pc_ttgt.p,
                                 c->d_pc_tmask.p, c->d_pc_act.p, c->d_pc_delta.p, c->d_pc_loss.p, c->d_pc_gpart.p,
                                 c->d_pc_grad.p);
        if (e) return e;
        if (it == 0 && (e = loss(loss_first))) return e;
        if ((e = crt_launch_pc_adam(c->stream, c->d_pc_par.p, c->d_pc_m.p, c->d_pc_v.p, c->d_pc_grad.p, (float)lr,
                                    ++c->pc_steps)))
            return e;
    }
    // loss after the last update
    e = crt_launch_pc_fwdbwd(c->stream, ntr, c->d_pc_par.p, c->d_pc_norm.p, c->d_pc_tfeat.p, c->d_pc_ttgt.p,
                             c->d_pc_tmask.p, c->d_pc_act.p, c->d_pc_delta.p, c->d_pc_loss.p, c->d_pc_gpart.p,
                             c->d_pc_grad.p);
    if (e || (e = loss(loss_last))) return e;
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->pc_ready = true;
    return 0;
}

int crt_pc_reset(crt_ctx *c) {
    if (!c) return CRT_ERR_ARG;
    c->pc_tn = c->pc_manz = c->pc_steps = 0;
    c->pc_ready = c->pc_norm_set = false;
    return 0;
}

int crt_pc_grad(crt_ctx *c, double *loss, float *grad) {
    // loss and gradient of the current parameters on the training system of the last
    // crt_pc_train call (no update)
    if (!c || !grad) return CRT_ERR_ARG;
    if (!c->pc_ready || c->pc_manz == 0) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    const int manz = c->pc_manz;
    const int nb = crt_pc_blocks(manz);
    int e = crt_launch_pc_fwdbwd(c->stream, manz, c->d_pc_par.p, c->d_pc_norm.p, c->d_pc_tfeat.p, c->d_pc_ttgt.p,
                                 c->d_pc_tmask.p, c->d_pc_act.p, c->d_pc_delta.p, c->d_pc_loss.p, c->d_pc_gpart.p,
                                 c->d_pc_grad.p);
    if (e) return e;
    std::vector<double> pl(nb);
    CRT_HIP(c->copy_out(pl.data(), c->d_pc_loss.p, nb * sizeof(double)));
    CRT_HIP(c->copy_out(grad, c->d_pc_grad.p, crt_pc_nparams() * sizeof(float)));
    if (loss) {
        double t = 0.0;
        for (double v : pl) t += v;
        *loss = t;
    }
    return 0;
}

int crt_pc_data(crt_ctx *c, int *n, float *feat, float *tgt, float *mask) {
    // the network inputs' raw features, targets and mask of the training set; *n: the
    // caller's capacity in cells on entry, the set's size on return
    if (!c || !n) return CRT_ERR_ARG;
    if (c->pc_manz == 0) return CRT_ERR_STATE;
    const size_t manz = c->pc_manz;
    const int cap = *n;
    *n = c->pc_manz;
    if ((feat || tgt || mask) && cap < c->pc_manz) return CRT_ERR_ARG;
    if (feat) CRT_HIP(c->copy_out(feat, c->d_pc_tfeat.p, manz * 8 * sizeof(float)));
    if (tgt) CRT_HIP(c->copy_out(tgt, c->d_pc_ttgt.p, manz * 5 * sizeof(float)));
    if (mask) CRT_HIP(c->copy_out(mask, c->d_pc_tmask.p, manz * 5 * sizeof(float)));
    return 0;
}

int crt_set_cg_precision(crt_ctx *c, int jbits) {
    if (!c || (jbits != 64 && jbits != 32)) return CRT_ERR_ARG;
    c->cg_jbits = jbits;
    if (jbits == 64) {
        c->d_sens32.release();
        c->d_sensr32.release();
    }
    c->sens32_ok = false;
    c->sensr_ok = false;
    c->sensr32_ok = false;
    return 0;
}

int crt_set_forward_precision(crt_ctx *c, int bits, double tol, int max_sweeps) {
    if (!c || (bits != 64 && bits != 32) || !(tol > 0.0) || max_sweeps < 1) return CRT_ERR_ARG;
    if (bits == 32 && c->have_grid && !(crt_factor32_supported(c->mb) && c->tmapD == (c->mb + 15) / 16))
        return CRT_ERR_UNSUPPORTED;
    c->fwd_bits = bits;
    c->refine_tol = tol;
    c->refine_accept = std::max(tol, 1e-12);
    c->refine_max = max_sweeps;
    return 0;
}

int crt_get_refine_info(crt_ctx *c, int *sweeps, double *ratio, int *fallbacks) {
    if (!c) return CRT_ERR_ARG;
    if (sweeps) *sweeps = c->refine_sweeps;
    if (ratio) *ratio = c->refine_ratio;
    if (fallbacks) *fallbacks = c->refine_fallbacks;
    return 0;
}

int crt_cjg(crt_ctx *c, double lam, double eps, int ncgmax, const double *wmatd, const int *wdfak,
            const double *cgfac, const double *bvec, double *dpar, double *cgres, int *ncg) {
    if (!c || !wmatd || !wdfak || !cgfac || !bvec || ncgmax < 0) return CRT_ERR_ARG;
    if (!c->have_sens || c->ltri < 0) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    int nanz = 0, manz = 0;
    if (int e0 = jsizes(c, true, &nanz, &manz)) return e0;
    int e = cg_weights(c, nanz, wmatd, wdfak);
    e = e ? e : c->d_cgfac.upload(cgfac, manz, c->stream);
    e = e ? e : c->d_bvec.upload(reinterpret_cast<const zd *>(bvec), manz, c->stream);
    if (e) return e;
    CGArgs a;
    if ((e = cg_prepare(c, nanz, manz, ncgmax, lam, eps, a))) return e;
    if ((e = cg_fsai(c, a))) return e;
    int n = 0;
    if ((e = cg_iterate(c, a, c->d_bvec.p, &n))) return e;
    if ((e = cg_fallback(c, a, ncgmax, &n))) return e;
    if (dpar) CRT_HIP(c->copy_out(dpar, c->d_dpar.p, manz * sizeof(zd)));
    if (cgres) CRT_HIP(c->copy_out(cgres, c->d_cgres.p, ((size_t)ncgmax + 1) * sizeof(double)));
    if (ncg) *ncg = n;
    return 0;
}

// crt_cjg without an allocation: the same kernels and order as crt_cjg (so the same
// results), every buffer of the solve taken from the device's static scratch array.
// Complex inversion, FP64 J, Jacobi scaling only; a solve whose buffers exceed the
// array returns CRT_ERR_SCRATCH.  Calls that overlap on one GPU race on the array.
int crt_cjg_noalloc(crt_ctx *c, double lam, double eps, int ncgmax, const double *wmatd, const int *wdfak,
                    const double *cgfac, const double *bvec, double *dpar, double *cgres, int *ncg) {
    if (!c || !wmatd || !wdfak || !cgfac || !bvec || ncgmax < 0) return CRT_ERR_ARG;
    if (!c->have_sens) return CRT_ERR_STATE;
    if (c->inv_kind != 0 || c->cg_jbits != 64 || c->cg_prec != 0) return CRT_ERR_UNSUPPORTED;
    int nanz = 0, manz = 0;
    if (int e0 = jsizes(c, true, &nanz, &manz)) return e0;
    CRT_HIP(hipSetDevice(c->device));
    unsigned char *base = scratch_base(c->device);
    if (!base) return CRT_ERR_STATE;
    int ncb, cb, n_dr1, n_rr;
    crt_cg_geometry(nanz, manz, &ncb, &cb, &n_dr1, &n_rr);
    // layout: 256-byte aligned pieces, in this order
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off = (off + bytes + 255) & ~(size_t)255;
        return base + at;
    };
    double *w = reinterpret_cast<double *>(take(nanz * sizeof(double)));
    int *wd = reinterpret_cast<int *>(take(nanz * sizeof(int)));
    double *cf = reinterpret_cast<double *>(take(manz * sizeof(double)));
    zd *bv0 = reinterpret_cast<zd *>(take(manz * sizeof(zd)));
    zd *r = reinterpret_cast<zd *>(take(manz * sizeof(zd)));
    zd *p = reinterpret_cast<zd *>(take(manz * sizeof(zd)));
    zd *q = reinterpret_cast<zd *>(take(manz * sizeof(zd)));
    zd *bv = reinterpret_cast<zd *>(take(manz * sizeof(zd)));
    zd *dp = reinterpret_cast<zd *>(take(manz * sizeof(zd)));
    zd *ap = reinterpret_cast<zd *>(take(nanz * sizeof(zd)));
    zd *part_ap = reinterpret_cast<zd *>(take((size_t)ncb * nanz * sizeof(zd)));
    double *part_dr1 = reinterpret_cast<double *>(take(n_dr1 * sizeof(double)));
    double *part_rr = reinterpret_cast<double *>(take(n_rr * sizeof(double)));
    double *res = reinterpret_cast<double *>(take(((size_t)ncgmax + 1) * sizeof(double)));
    CGScalars *sc = reinterpret_cast<CGScalars *>(take(2 * sizeof(CGScalars)));
    if (off > CRT_SCRATCH_BYTES) return CRT_ERR_SCRATCH;
    CRT_HIP(hipMemcpyAsync(w, wmatd, nanz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CRT_HIP(hipMemcpyAsync(wd, wdfak, nanz * sizeof(int), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(crt_noalloc_weights_kernel, dim3((nanz + 255) / 256), dim3(256), 0, c->stream, nanz, wd, w);
    CRT_HIP(hipGetLastError());
    CRT_HIP(hipMemcpyAsync(cf, cgfac, manz * sizeof(double), hipMemcpyHostToDevice, c->stream));
    CRT_HIP(hipMemcpyAsync(bv0, bvec, manz * sizeof(zd), hipMemcpyHostToDevice, c->stream));
    // the residual history starts at zero (the array keeps the previous call's values)
    CRT_HIP(hipMemsetAsync(res, 0, ((size_t)ncgmax + 1) * sizeof(double), c->stream));
    CGArgs a{};
    a.nanz = nanz; a.manz = manz; a.ltri = c->ltri; a.nx = c->rnx; a.maxn = c->rncol - 1; a.ncgmax = ncgmax;
    a.lam = lam; a.eps = eps;
    a.J = c->d_sens.p; a.J32 = nullptr; a.Jr = nullptr; a.part = 0;
    a.w = w; a.smatm = c->d_smatm.p; a.nachbar = c->d_nachbar.p; a.cgfac = cf;
    a.r = r; a.p = p; a.q = q; a.bv = bv; a.dpar = dp; a.ap = ap;
    a.part_ap = part_ap; a.part_dr1 = part_dr1; a.part_rr = part_rr;
    a.n_dr1 = n_dr1; a.n_rr = n_rr; a.ncb = ncb; a.cb = cb;
    a.fg = nullptr; a.fpat = nullptr; a.ftp = nullptr; a.t = nullptr; a.z = nullptr; a.part_rz = nullptr;
    a.sc = sc; a.scn = sc; a.cgres = res;
    int n = 0;
    if (int e = cg_iterate(c, a, bv0, &n)) return e;
    if (dpar) CRT_HIP(c->copy_out(dpar, dp, manz * sizeof(zd)));
    if (cgres) CRT_HIP(c->copy_out(cgres, res, ((size_t)ncgmax + 1) * sizeof(double)));
    if (ncg) *ncg = n;
    return 0;
}

int crt_buncert(crt_ctx *c, double lam, const double *wmatd, const int *wdfak, int lres, int lcov2, int gauss,
                double *ata, double *ata_reg, double *cov, double *res, double *cov2, double *diag) {
    if (!c || !wmatd || !wdfak) return CRT_ERR_ARG;
    if (!c->have_sens || c->ltri < 0) return CRT_ERR_STATE;
    // the DC inversion's bata reads sensdc (src/bmcm_mod.f90:197-206); the final phase
    // improvement runs the complex branches with the caller's lam_cri and wmatdr
    // (src/bmcm_mod.f90:56-83)
    CRT_HIP(hipSetDevice(c->device));
    int nanz = 0, manz = 0;
    if (int e0 = jsizes(c, true, &nanz, &manz)) return e0;
    int e = cg_weights(c, nanz, wmatd, wdfak);
    if (e) return e;
    const size_t n2 = (size_t)manz * manz;
    const bool want_res = lres != 0, want_cov2 = lres != 0 && lcov2 != 0;
    // dense work buffers live for this call only (five manz^2 matrices at most)
    DBuf<double> d_ata, d_reg, d_cov, d_res, d_cov2, d_z;
    DBuf<int> d_info, d_ipiv;
    e = d_ata.ensure(n2);
    e = e ? e : d_reg.ensure(n2);
    e = e ? e : d_cov.ensure(n2);
    e = e ? e : (want_res ? d_res.ensure(n2) : 0);
    e = e ? e : (want_cov2 ? d_cov2.ensure(n2) : 0);
    e = e ? e : d_z.ensure(2 * (size_t)nanz * manz);
    e = e ? e : d_info.ensure(1);
    e = e ? e : d_ipiv.ensure(manz);
    if (e) return e;
    e = crt_bmcm_run(c->stream, nanz, manz, c->d_sens.p, c->d_w.p, c->inv_kind == 1, c->ltri, c->rnx, c->rncol - 1,
                     c->d_smatm.p,
                     c->d_nachbar.p, lam, gauss ? 1 : 0, d_ata.p, d_reg.p, d_cov.p, want_res ? d_res.p : nullptr,
                     want_cov2 ? d_cov2.p : nullptr, d_z.p, d_info.p, d_ipiv.p);
    const bool have_cov = e == 0;
    if (e && e != 108) return e;
    const double *mats[5] = {d_ata.p, d_reg.p, have_cov ? d_cov.p : nullptr, have_cov && want_res ? d_res.p : nullptr,
                             have_cov && want_cov2 ? d_cov2.p : nullptr};
    double *outs[5] = {ata, ata_reg, cov, res, cov2};
    for (int q = 0; q < 5; ++q) {
        if (!mats[q]) continue;
        if (outs[q]) CRT_HIP(c->copy_out(outs[q], mats[q], n2 * sizeof(double)));
        if (diag) {
            CRT_HIP(hipMemcpy2DAsync(diag + (size_t)q * manz, sizeof(double), mats[q], (manz + 1) * sizeof(double),
                                     sizeof(double), manz, hipMemcpyDeviceToHost, c->stream));
            CRT_HIP(hipStreamSynchronize(c->stream));
        }
    }
    return e;
}

// besp_elem (src/besp_elem.f90:1-97): grid statistics of the area-element centroids
// (espx, espy as src/relem.f90:128-143 forms them: the sum of the node coordinates in
// node order over the node count).  The O(elanz^2) pair scan runs on the device; the
// neighbour distances abst (to the last edge neighbour of src/bnachbar.f90, as the
// reference's loop leaves it), their extremes, mean, mean absolute deviation and median
// (MDIAN1: sorted, mean of the middle pair for an even count) on the host.
int crt_besp_elem(crt_ctx *c, double *stats, const char *gstat_path) {
    if (!c || !stats) return CRT_ERR_ARG;
    if (!c->have_grid) return CRT_ERR_STATE;
    const int n = c->elanz, maxn = c->maxn;
    if (n < 2) return CRT_ERR_ARG;
    CRT_HIP(hipSetDevice(c->device));
    std::vector<double> cx(n, 0.0), cy(n, 0.0);
    for (int e = 0; e < n; ++e) {
        const int nk = c->nkel[e];
        for (int m = 0; m < nk; ++m) {
            const int nd = c->nrel[(size_t)e * maxn + m] - 1;
            cx[e] = cx[e] + c->x[nd];
            cy[e] = cy[e] + c->y[nd];
        }
        cx[e] = cx[e] / nk;
        cy[e] = cy[e] / nk;
    }
    DBuf<double> d_cx, d_cy, d_part;
    int nb = 0;
    int e = d_cx.upload(cx.data(), n, c->stream);
    e = e ? e : d_cy.upload(cy.data(), n, c->stream);
    e = e ? e : d_part.ensure((size_t)6 * ((n + 255) / 256));
    e = e ? e : crt_launch_besp_pairs(c->stream, n, d_cx.p, d_cy.p, d_part.p, &nb);
    if (e) return e;
    std::vector<double> part((size_t)6 * nb);
    CRT_HIP(c->copy_out(part.data(), d_part.p, part.size() * sizeof(double)));
    double v[6] = {HUGE_VAL, 0.0, HUGE_VAL, 0.0, HUGE_VAL, 0.0};
    for (int b = 0; b < nb; ++b)
        for (int q = 0; q < 6; ++q) v[q] = (q & 1) ? std::max(v[q], part[6 * b + q]) : std::min(v[q], part[6 * b + q]);
    // the reference's running values start at 1e5 (minima) and 0 (maxima)
    const double grid_min = std::min(1e5, std::sqrt(v[0])), grid_max = std::sqrt(v[1]);
    // grid_minx = MAX(MIN(grid_minx, ax), grid_min) per pair ends at the larger of the
    // smallest |dx| and the final grid_min (grid_min never grows); the same for y
    const double grid_minx = std::max(std::min(1e5, v[2]), grid_min), grid_maxx = v[3];
    const double grid_miny = std::max(std::min(1e5, v[4]), grid_min), grid_maxy = v[5];
    // neighbour distances
    std::vector<int> nb_((size_t)n * (maxn + 1));
    if ((e = crt_bnachbar(n, maxn, c->nrel.data(), nb_.data()))) return e;
    std::vector<double> abst(n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < maxn; ++j) {
            const int k = nb_[(size_t)j * n + i];
            if (k == 0) continue;
            const double dx = cx[i] - cx[k - 1], dy = cy[i] - cy[k - 1];
            abst[i] = std::sqrt(dx * dx + dy * dy);
        }
    double esp_max = abst[0], esp_min = abst[0], sum = 0.0;
    for (double a : abst) { esp_max = std::max(esp_max, a); esp_min = std::min(esp_min, a); sum += a; }
    const double esp_mit = sum / n;
    double sd = 0.0;
    for (double a : abst) sd += std::sqrt((a - esp_mit) * (a - esp_mit));
    const double esp_std = sd / std::max(1, n - 1);
    std::vector<double> srt(abst);
    std::sort(srt.begin(), srt.end());
    const int n2 = n / 2;
    const double esp_med = (2 * n2 == n) ? 0.5 * (srt[n2 - 1] + srt[n2]) : srt[n2];
    const double out[11] = {esp_min, esp_max, esp_mit, esp_med, esp_std, grid_min, grid_max,
                            grid_minx, grid_maxx, grid_miny, grid_maxy};
    std::memcpy(stats, out, sizeof(out));
    if (gstat_path) {
        // inv.gstat as src/besp_elem.f90:74-93 writes it
        FILE *f = fopen(gstat_path, "w");
        if (!f) return 1;
        fprintf(f, "Grid statistics:\n\n");
        fprintf(f, "%20sGridcells:\t%10d\n", "", n);
        fprintf(f, "%20sESP Min/Max:\t%10.4f%10.4f\n", "", esp_min, esp_max);
        fprintf(f, "%20sGRID Min/Max:\t%10.4f%10.4f\n", "", grid_min, grid_max);
        fprintf(f, "%20sGRID-x Min/Max:\t%10.4f%10.4f\n", "", grid_minx, grid_maxx);
        fprintf(f, "%20sGRID-y Min/Max:\t%10.4f%10.4f\n", "", grid_miny, grid_maxy);
        fprintf(f, "%20sMean/Median/Var:\t%10.4f%10.4f%10.4f\n", "", esp_mit, esp_med, esp_std);
        if (fclose(f) != 0) return 4;
    }
    return 0;
}

int crt_blam0(crt_ctx *c, int fpi, const double *wmatd, const int *wdfak, double alfx, double alfz,
              double *lammax) {
    if (!c || !wmatd || !wdfak || !lammax || !(alfx + alfz != 0.0)) return CRT_ERR_ARG;
    if (!c->have_sens) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    int nanz = 0, manz = 0;
    if (int e0 = jsizes(c, false, &nanz, &manz)) return e0;
    int ncb, cb;
    const size_t np = crt_blam0_scratch(nanz, manz, &ncb, &cb);
    int e = cg_weights(c, nanz, wmatd, wdfak);
    e = e ? e : c->d_part_ap.ensure(np);
    e = e ? e : c->d_ap.ensure(nanz);
    e = e ? e : c->d_jtj.ensure((size_t)manz + 1);
    // the DC branch over sensdc is the FPI form on a J with zero imaginary parts
    const int rpart = (fpi || c->inv_kind == 1) ? 1 : 0;
    e = e ? e : crt_launch_blam0(c->stream, nanz, manz, rpart, c->d_sens.p, c->d_w.p, c->d_part_ap.p, ncb, cb,
                                 c->d_ap.p, c->d_jtj.p, 2.0 / (alfx + alfz), c->d_jtj.p + manz);
    if (e) return e;
    // the context stream does not synchronise with the null stream: copy on it
    CRT_HIP(hipMemcpyAsync(lammax, c->d_jtj.p + manz, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CRT_HIP(hipStreamSynchronize(c->stream));
    return 0;
}

int crt_update(crt_ctx *c, double lam, double eps, int ncgmax, double step, const double *wmatd, const int *wdfak,
               const double *dat, const double *sigmaa, double *par, double *dpar, double *cgres, int *ncg,
               double *cgfac, double *bvec0) {
    if (!c || !wmatd || !wdfak || !dat || !par || ncgmax < 0) return CRT_ERR_ARG;
    if (!c->have_sens || c->ltri < 0) return CRT_ERR_STATE;
    if (!sigmaa && !c->have_volt) return CRT_ERR_STATE;
    CRT_HIP(hipSetDevice(c->device));
    int nanz = 0, manz = 0;
    if (int e0 = jsizes(c, true, &nanz, &manz)) return e0;
    int e = cg_weights(c, nanz, wmatd, wdfak);
    // residual d - f(m)
    std::vector<zd> res(nanz), sa(nanz);
    if (sigmaa) std::memcpy(sa.data(), sigmaa, nanz * sizeof(zd));
    else CRT_HIP(c->copy_out(sa.data(), c->d_sigmaa.p, nanz * sizeof(zd)));
    const zd *dd = reinterpret_cast<const zd *>(dat);
    for (int i = 0; i < nanz; ++i) res[i] = zd{dd[i].re - sa[i].re, dd[i].im - sa[i].im};
    e = e ? e : c->d_res.upload(res.data(), nanz, c->stream);
    e = e ? e : c->d_par.upload(reinterpret_cast<const zd *>(par), manz, c->stream);
    e = e ? e : c->d_rm.ensure(manz);
    e = e ? e : c->d_cgfac.ensure(manz);
    e = e ? e : c->d_bvec.ensure(manz);
    if (e) return e;
    CGArgs a;
    if ((e = cg_prepare(c, nanz, manz, ncgmax, lam, eps, a))) return e;
    if ((e = crt_launch_cg_rmul(c->stream, a, c->d_par.p, c->d_rm.p))) return e;
    if ((e = crt_launch_cg_rhs(c->stream, nanz, manz, lam, c->d_sens.p, a.Jr, c->d_w.p, c->d_res.p, c->d_rm.p,
                               c->d_rdiag.p, c->d_cgfac.p, c->d_bvec.p)))
        return e;
    if (cgfac) CRT_HIP(hipMemcpyAsync(cgfac, c->d_cgfac.p, manz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (bvec0) CRT_HIP(hipMemcpyAsync(bvec0, c->d_bvec.p, manz * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
    if ((e = cg_fsai(c, a))) return e;
    int n = 0;
    if ((e = cg_iterate(c, a, c->d_bvec.p, &n))) return e;
    if ((e = cg_fallback(c, a, ncgmax, &n))) return e;
    if ((e = crt_launch_cg_final(c->stream, manz, step, c->d_cgfac.p, c->d_dpar.p, c->d_par.p, a.part))) return e;
    if (dpar) CRT_HIP(hipMemcpyAsync(dpar, c->d_dpar.p, manz * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
    CRT_HIP(hipMemcpyAsync(par, c->d_par.p, manz * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
    if (cgres) CRT_HIP(hipMemcpyAsync(cgres, c->d_cgres.p, ((size_t)ncgmax + 1) * sizeof(double),
                                      hipMemcpyDeviceToHost, c->stream));
    CRT_HIP(hipStreamSynchronize(c->stream));
    if (ncg) *ncg = n;
    return 0;
}

int crt_get_times(crt_ctx *c, double *ms, int n) {
    if (!c || !ms) return CRT_ERR_ARG;
    for (int i = 0; i < n && i < T_N; ++i) ms[i] = c->times[i];
    return 0;
}

// Cuthill-McKee renumbering (cutmckee/cutmck.f:126-300), restated:
//   * node graph of every element's node pairs, neighbours kept in the order the pairs
//     first appear (at most 500 per node, as the reference's graph array);
//   * start candidates: the nodes of the smallest degree in node order, the degree
//     raised until 100 candidates are found or it passes elanz/10; the last candidate
//     found is not tried (the reference counts one past its list);
//   * from each start a breadth-first numbering in which every node hands out its
//     unnumbered neighbours by rising degree, equal degrees latest-first;
//   * the first start with the smallest band width max |new(i) - new(j)| wins.
// perm[old - 1] = new number, 1-based.
namespace {

struct NodeGraph {
    std::vector<std::vector<int>> nb;  // 1-based neighbours in first-appearance order
    int add_pair(int a, int b) {
        auto &la = nb[a - 1];
        if (std::find(la.begin(), la.end(), b) != la.end()) return 0;
        if (la.size() >= 500 || nb[b - 1].size() >= 500) return CRT_ERR_UNSUPPORTED;
        la.push_back(b);
        nb[b - 1].push_back(a);
        return 0;
    }
    int degree(int v) const { return (int)nb[v - 1].size(); }
};

// breadth-first numbering from `start`; false when the graph is not connected
bool cm_number(const NodeGraph &g, int start, std::vector<int> &order, std::vector<int> &newnum) {
    const int n = (int)g.nb.size();
    order.clear();
    std::fill(newnum.begin(), newnum.end(), 0);
    order.push_back(start);
    newnum[start - 1] = 1;
    std::vector<std::pair<int, int>> cand;  // (degree, -position)
    for (size_t head = 0; head < order.size(); ++head) {
        const auto &lst = g.nb[order[head] - 1];
        cand.clear();
        for (int pos = 0; pos < (int)lst.size(); ++pos)
            if (!newnum[lst[pos] - 1]) cand.emplace_back(g.degree(lst[pos]), -pos);
        std::sort(cand.begin(), cand.end());
        for (const auto &cp : cand) {
            const int v = lst[-cp.second];
            order.push_back(v);
            newnum[v - 1] = (int)order.size();
        }
    }
    return (int)order.size() == n;
}

}  // namespace

int crt_cutmck(int sanz, int typanz, const int *nelanz, const int *selanz, int maxn, const int *nrel, int elanz,
               int *perm, int *mb_out) {
    if (sanz <= 0 || !perm) return CRT_ERR_ARG;
    NodeGraph g;
    g.nb.resize(sanz);
    size_t first = 0;  // first element of the current type group
    for (int t = 0; t < typanz; ++t) {
        for (int m = 0; m < nelanz[t]; ++m) {
            const int *el = nrel + (first + m) * maxn;
            for (int i = 0; i < selanz[t]; ++i)
                for (int j = i + 1; j < selanz[t]; ++j) {
                    if (el[i] < 1 || el[i] > sanz || el[j] < 1 || el[j] > sanz) return CRT_ERR_ARG;
                    if (int e = g.add_pair(el[i], el[j])) return e;
                }
        }
        first += nelanz[t];
    }
    int dmin = sanz;
    for (int v = 1; v <= sanz; ++v) dmin = std::min(dmin, g.degree(v));
    if (dmin == 0) return CRT_ERR_ARG;
    const int want = 100;
    std::vector<int> starts;
    for (int d = dmin;; ++d) {
        for (int v = 1; v <= sanz && (int)starts.size() < want; ++v)
            if (g.degree(v) == d) starts.push_back(v);
        if ((int)starts.size() >= want || d > elanz / 10) break;
    }
    if (!starts.empty()) starts.pop_back();
    std::vector<int> order, newnum(sanz);
    int best = sanz + 1;
    for (int st : starts) {
        if (!cm_number(g, st, order, newnum)) return CRT_ERR_ARG;  // disconnected grid
        int bw = 0;
        for (int v = 1; v <= sanz; ++v)
            for (int w : g.nb[v - 1]) bw = std::max(bw, std::abs(newnum[v - 1] - newnum[w - 1]));
        if (bw < best) {
            best = bw;
            std::copy(newnum.begin(), newnum.end(), perm);
        }
    }
    if (best > sanz) return CRT_ERR_ARG;
    if (mb_out) *mb_out = best;
    return 0;
}

// ------------------------------------------------------ single-system drop-ins
static DevGrid band_grid(int sanz, int mb) {
    DevGrid g{};
    g.sanz = sanz;
    g.mb = mb;
    g.nblk = (sanz + CRT_NB - 1) / CRT_NB;
    g.D = std::max(1, (mb + CRT_NB - 1) / CRT_NB);
    g.fstride = g.nblk * CRT_NB;
    return g;
}

// chol(a): in-place complex-symmetric band Cholesky of a scaled band matrix in
// the reference layout (src/chol.f90:46-78); errnr 28 on a vanishing pivot.
int crt_chol(crt_ctx *c, int sanz, int mb, double *a) {
    if (!c || !a || sanz <= 0 || mb < 0) return CRT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(c->dropin_mu);
    CRT_HIP(hipSetDevice(c->device));
    DevGrid g = band_grid(sanz, mb);
    if (crt_factor_lds_bytes(g.D) > 160 * 1024) return CRT_ERR_UNSUPPORTED;
    size_t nband = (size_t)sanz * (mb + 1), npanel = (size_t)g.nblk * (g.D + 1) * CRT_NB * CRT_NB;
    int e = c->d_tmp.upload(reinterpret_cast<const zd *>(a), nband, c->stream);
    e = e ? e : c->d_panels.ensure(npanel);
    e = e ? e : c->d_yf.ensure((size_t)g.nblk * CRT_NB);
    e = e ? e : c->d_sys.ensure(1);
    e = e ? e : c->d_status.ensure(2);
    if (e) return e;
    int2 s0{0, 0};
    CRT_HIP(hipMemcpyAsync(c->d_sys.p, &s0, sizeof(int2), hipMemcpyHostToDevice, c->stream));
    CRT_HIP(hipMemsetAsync(c->d_status.p, 0, 2 * sizeof(int), c->stream));
    c->tic(T_FACTOR);
    if ((e = crt_launch_factor_band(c->stream, g, c->d_sys.p, c->d_tmp.p, c->d_panels.p, c->d_yf.p,
                                    c->d_status.p + 1)))
        return e;
    c->toc(T_FACTOR);
    if ((e = crt_launch_panel_band(c->stream, sanz, mb, g.nblk, g.D, 1, c->d_panels.p, c->d_tmp.p))) return e;
    CRT_HIP(hipMemcpyAsync(a, c->d_tmp.p, nband * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
    int st[2] = {0, 0};
    CRT_HIP(hipMemcpyAsync(st, c->d_status.p, 2 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->collect(T_FACTOR);
    c->have_kpot = false;  // the forward workspace was reused
    return std::max(st[0], st[1]);
}

// choldc(a) (src/choldc.f90:7): the real band LL^T of the DC path, run by the complex
// kernel on the same numbers with zero imaginary parts (which stay exactly zero while
// every pivot is positive).  Where the reference stops on s <= 0 (errnr 28) the complex
// factorisation would take an imaginary square root instead: a diagonal entry of L
// that is not real and positive returns 28, and a is then left as it was.
int crt_choldc(crt_ctx *c, int sanz, int mb, double *a) {
    if (!c || !a || sanz <= 0 || mb < 0) return CRT_ERR_ARG;
    const size_t n = (size_t)sanz * (mb + 1);
    std::vector<double> z(2 * n);
    for (size_t i = 0; i < n; ++i) {
        z[2 * i] = a[i];
        z[2 * i + 1] = 0.0;
    }
    const int e = crt_chol(c, sanz, mb, z.data());
    if (e) return e;
    for (int i = 0; i < sanz; ++i) {
        const size_t d = (size_t)i * (mb + 1) + mb;
        if (z[2 * d + 1] != 0.0 || !(z[2 * d] > 0.0)) return CRT_ERR_PIVOT;
    }
    for (size_t i = 0; i < n; ++i) a[i] = z[2 * i];
    return 0;
}

// vredc(a, b, pot) (src/vredc.f90:1): pot = -A^{-1} b for the real factor and real
// right-hand side of the DC path, complex pot out as the reference's dcmplx(potdc)
int crt_vredc(crt_ctx *c, int sanz, int mb, const double *a, const double *b, double *pot) {
    if (!c || !a || !b || !pot || sanz <= 0 || mb < 0) return CRT_ERR_ARG;
    const size_t n = (size_t)sanz * (mb + 1);
    std::vector<double> za(2 * n, 0.0), zb(2 * (size_t)sanz, 0.0);
    for (size_t i = 0; i < n; ++i) za[2 * i] = a[i];
    for (int i = 0; i < sanz; ++i) zb[2 * (size_t)i] = b[i];
    return crt_vre(c, sanz, mb, za.data(), zb.data(), pot);
}

// vre(a, b, pot): pot = -A^{-1} b from the band factor (src/vre.f90:45-69).
int crt_vre(crt_ctx *c, int sanz, int mb, const double *a, const double *b, double *pot) {
    if (!c || !a || !b || !pot || sanz <= 0 || mb < 0) return CRT_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lk(c->dropin_mu);
    CRT_HIP(hipSetDevice(c->device));
    DevGrid g = band_grid(sanz, mb);
    size_t nband = (size_t)sanz * (mb + 1), npanel = (size_t)g.nblk * (g.D + 1) * CRT_NB * CRT_NB;
    int e = c->d_tmp.upload(reinterpret_cast<const zd *>(a), nband, c->stream);
    e = e ? e : c->d_panels.ensure(npanel);
    e = e ? e : c->d_yscr.ensure((size_t)g.nblk * CRT_NB);
    e = e ? e : c->d_rhs.ensure(1);
    e = e ? e : c->d_res.upload(reinterpret_cast<const zd *>(b), sanz, c->stream);
    e = e ? e : c->d_rm.ensure(sanz);
    if (e) return e;
    if ((e = crt_launch_panel_band(c->stream, sanz, mb, g.nblk, g.D, 0, c->d_panels.p, c->d_tmp.p))) return e;
    int4 r0{0, 0, 0, 0};
    CRT_HIP(hipMemcpyAsync(c->d_rhs.p, &r0, sizeof(int4), hipMemcpyHostToDevice, c->stream));
    c->tic(T_SOLVE);
    if ((e = crt_launch_solve(c->stream, sanz, g.nblk, g.D, 1, 1, 1, c->d_rhs.p, nullptr, nullptr, c->d_panels.p,
                              nullptr, c->d_yscr.p, nullptr, c->d_res.p, c->d_rm.p)))
        return e;
    c->toc(T_SOLVE);
    CRT_HIP(hipMemcpyAsync(pot, c->d_rm.p, sanz * sizeof(zd), hipMemcpyDeviceToHost, c->stream));
    CRT_HIP(hipStreamSynchronize(c->stream));
    c->collect(T_SOLVE);
    c->have_kpot = false;
    return 0;
}

// cholz(a, p, n): dense complex-symmetric Cholesky, strict lower factor into
// a(j,i) j>i and the diagonal into p (src/cholz.f90:43-76).  Runs as a band
// factorisation with mb = n-1.
int crt_cholz(crt_ctx *c, int n, double *a, double *p) {
    if (!c || !a || !p || n <= 0) return CRT_ERR_ARG;
    const int mb = std::max(n - 1, 0);
    std::vector<zd> band((size_t)n * (mb + 1), zd{0.0, 0.0});
    const zd *A = reinterpret_cast<const zd *>(a);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) band[(size_t)i * (mb + 1) + (j - i + mb)] = A[(size_t)i * n + j];  // a(j,i) = a(i,j)
    int e = crt_chol(c, n, mb, reinterpret_cast<double *>(band.data()));
    if (e) return e;
    zd *Aw = reinterpret_cast<zd *>(a);
    zd *P = reinterpret_cast<zd *>(p);
    for (int i = 0; i < n; ++i) {
        P[i] = band[(size_t)i * (mb + 1) + mb];
        for (int j = i + 1; j < n; ++j) Aw[(size_t)i * n + j] = band[(size_t)j * (mb + 1) + (i - j + mb)];  // a(j,i) = L(j,i)
    }
    return 0;
}

// ------------------------------------------------------------------ batches
int crt_batch_create(crt_ctx *const *ctxs, int n, size_t workspace_bytes, int flags, crt_batch **out) {
    for (int i = 0; ctxs && i < n; ++i)  // unit sources only: no forward switches in a batch
        if (ctxs[i] && (ctxs[i]->sw_lsr || ctxs[i]->sw_lana || ctxs[i]->sw_nsink >= 0)) return CRT_ERR_UNSUPPORTED;
    if (!ctxs || n <= 0 || !out || (flags & ~3)) return CRT_ERR_ARG;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!ctxs[i]) return CRT_ERR_ARG;
        if (!ctxs[i]->have_precal || !ctxs[i]->have_sigma) return CRT_ERR_STATE;
        if (ctxs[i]->device != ctxs[0]->device) return CRT_ERR_ARG;
        CRT_HIP(hipSetDevice(ctxs[i]->device));
        if (int e0 = build_dmask(ctxs[i], true)) return e0;
    }
    std::unique_ptr<crt_batch> b(new crt_batch());
    b->device = ctxs[0]->device;
    CRT_HIP(hipSetDevice(b->device));
    CRT_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    CRT_HIP(hipEventCreate(&b->e0));
    CRT_HIP(hipEventCreate(&b->e1));
    CRT_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    if (workspace_bytes == 0) workspace_bytes = (size_t)64 << 30;
    int nslot = 2;
    if (const char *v = getenv("CRT_BATCH_STREAMS")) nslot = std::max(1, std::min(8, atoi(v)));
    nslot = std::min(nslot, n);
    b->slots.resize(nslot);
    for (int q = 0; q < nslot; ++q) {
        if (q == 0) b->slots[q].st = b->stream;
        else CRT_HIP(hipStreamCreateWithFlags(&b->slots[q].st, hipStreamNonBlocking));
        CRT_HIP(hipEventCreateWithFlags(&b->slots[q].join, hipEventDisableTiming));
    }
    size_t np = 0, nf = 0, nv = 0, ny = 0, nx = 0, nr = 0;
    for (int i = 0; i < n; ++i) {
        auto M = std::make_unique<crt_batch::Model>();
        M->c = ctxs[i];
        M->slot = i % nslot;
        M->mixed = (flags & 2) != 0;
        int e = batch_plan(*M, workspace_bytes / nslot);
        if (e) return e;
        np = std::max(np, M->batch * M->panel_elems);
        nf = std::max(nf, M->batch * M->fstride);
        nv = std::max(nv, M->batch * ((size_t)M->c->nnz + 2));
        ny = std::max(ny, M->nyscr * M->fstride);
        if (M->mixed) {
            nx = std::max(nx, M->max_nr * M->fstride);
            nr = std::max(nr, 2 * M->max_nr);
        }
        b->nsys_total += M->nsys;
        b->m.push_back(std::move(M));
    }
    int e = b->panels.ensure(np * nslot);
    e = e ? e : b->yf.ensure(nf * nslot);
    e = e ? e : b->dinv.ensure(nf * nslot);
    e = e ? e : b->fak.ensure(nf * nslot);
    e = e ? e : b->dga.ensure(nf * nslot);
    e = e ? e : b->vals.ensure(nv * nslot);
    e = e ? e : b->yscr.ensure(ny * nslot);
    e = e ? e : (nx ? b->xacc.ensure(nx * nslot) : 0);
    e = e ? e : (nr ? b->rstat.ensure(nr * nslot) : 0);
    if (e) return e;
    for (int q = 0; q < nslot; ++q) {
        crt_batch::Slot &w = b->slots[q];
        w.panels = b->panels.p + q * np;
        w.yf = b->yf.p + q * nf;
        w.dinv = b->dinv.p + q * nf;
        w.fak = b->fak.p + q * nf;
        w.dga = b->dga.p + q * nf;
        w.vals = b->vals.p + q * nv;
        w.yscr = b->yscr.p + q * ny;
        w.xacc = nx ? b->xacc.p + q * nx : nullptr;
        w.rstat = nr ? b->rstat.p + q * nr : nullptr;
    }
    if (flags & 1) {
        for (auto &M : b->m) CRT_HIP(hipMemsetAsync(M->status.p, 0, (M->nsys + 1) * sizeof(int), b->stream));
        for (auto &M : b->m)
            if ((e = batch_area(b.get(), *M))) return e;
        for (auto &w : b->slots) CRT_HIP(hipStreamSynchronize(w.st));  // the tuning runs on the slot streams
        if ((e = batch_tune(b.get()))) return e;
    }
    for (auto &M : b->m) M->variant = M->use16 ? b->variant[M->D] : 0;
    *out = b.release();
    return 0;
}

int crt_batch_forward(crt_batch *b) {
    if (!b) return CRT_ERR_ARG;
    CRT_HIP(hipSetDevice(b->device));
    for (auto &M : b->m) {
        if (!M->c->have_sigma) return CRT_ERR_STATE;
        // cleared outside the graph: a memset node is not replayed reliably
        CRT_HIP(hipMemsetAsync(M->status.p, 0, (M->nsys + 1) * sizeof(int), b->stream));
    }
    if (!b->exec) {
        CRT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        // the other slot streams join the capture through an event of the origin stream
        CRT_HIP(hipEventRecord(b->fork, b->stream));
        for (size_t q = 1; q < b->slots.size(); ++q) CRT_HIP(hipStreamWaitEvent(b->slots[q].st, b->fork, 0));
        int e = 0;
        for (auto &M : b->m) {
            if ((e = batch_area(b, *M))) break;
            for (const auto &ch : M->chunks) {
                e = batch_assemble(b, *M, ch);
                e = e ? e : batch_factor(b, *M, ch, M->variant);
                e = e ? e : batch_solve(b, *M, ch);
                if (e) break;
            }
            if (e) break;
        }
        for (size_t q = 1; q < b->slots.size(); ++q) {
            if (hipEventRecord(b->slots[q].join, b->slots[q].st) != hipSuccess ||
                hipStreamWaitEvent(b->stream, b->slots[q].join, 0) != hipSuccess)
                e = e ? e : CRT_ERR_HIP;
        }
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->stream, &g);
        if (e || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return e ? e : CRT_ERR_HIP;
        }
        b->graph = g;
        if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            b->exec = nullptr;
            return CRT_ERR_HIP;
        }
    }
    CRT_HIP(hipEventRecord(b->e0, b->stream));
    CRT_HIP(hipGraphLaunch(b->exec, b->stream));
    CRT_HIP(hipEventRecord(b->e1, b->stream));
    CRT_HIP(hipEventSynchronize(b->e1));
    CRT_HIP(hipEventElapsedTime(&b->ms, b->e0, b->e1));
    int worst = 0;
    b->mixed_sweeps = b->mixed_fallbacks = 0;
    b->mixed_ratio = 0.0;
    for (auto &M : b->m) {
        std::vector<int> st(M->nsys + 1);
        CRT_HIP(hipMemcpyAsync(st.data(), M->status.p, st.size() * sizeof(int), hipMemcpyDeviceToHost, b->stream));
        CRT_HIP(hipStreamSynchronize(b->stream));
        // as crt_forward: the first failing system names the model's error, the first
        // failing model the batch's
        int w = st[0];
        for (size_t q = 1; q < st.size() && !w; ++q) w = st[q];
        M->fallback = 0;
        if (M->mixed) {
            // a chunk whose refinement did not converge, or an FP32 pivot failure: the
            // model again in FP64 through its own forward, which also names a genuine error
            const size_t nc = M->chunks.size();
            std::vector<int> cs(2 * nc);
            std::vector<double> cr(2 * nc);
            CRT_HIP(hipMemcpy(cs.data(), M->cstate.p, cs.size() * sizeof(int), hipMemcpyDeviceToHost));
            CRT_HIP(hipMemcpy(cr.data(), M->crat.p, cr.size() * sizeof(double), hipMemcpyDeviceToHost));
            bool ok = w == 0;
            for (size_t q = 0; q < nc; ++q) {
                b->mixed_sweeps = std::max(b->mixed_sweeps, cs[2 * q + 1]);
                ok = ok && (cs[2 * q] == 1 || cs[2 * q] == 2);
                if (cs[2 * q] == 1 || cs[2 * q] == 2) b->mixed_ratio = std::max(b->mixed_ratio, cr[2 * q + 1]);
            }
            if (!ok) {
                // the model's chunks again with the FP64 kernels, directly on the batch's
                // workspace (no allocation); their status names a genuine error
                M->fallback = 1;
                b->mixed_fallbacks++;
                M->mixed = false;
                const hipStream_t sst = b->slots[M->slot].st;
                int e = hipMemsetAsync(M->status.p, 0, (M->nsys + 1) * sizeof(int), sst) == hipSuccess ? 0
                                                                                          : CRT_ERR_HIP;
                e = e ? e : batch_area(b, *M);
                for (const auto &ch : M->chunks) {
                    e = e ? e : batch_assemble(b, *M, ch);
                    e = e ? e : batch_factor(b, *M, ch, M->variant);
                    e = e ? e : batch_solve(b, *M, ch);
                }
                M->mixed = true;
                if (e) return e;
                CRT_HIP(hipStreamSynchronize(sst));
                CRT_HIP(hipMemcpy(st.data(), M->status.p, st.size() * sizeof(int), hipMemcpyDeviceToHost));
                w = st[0];
                for (size_t q = 1; q < st.size() && !w; ++q) w = st[q];
            }
        }
        if (!worst) worst = w;
        M->c->have_kpot = w == 0;
        M->c->have_volt = false;
    }
    return worst;
}

int crt_batch_refine_info(crt_batch *b, int *sweeps, double *ratio, int *fallbacks) {
    if (!b) return CRT_ERR_ARG;
    if (sweeps) *sweeps = b->mixed_sweeps;
    if (ratio) *ratio = b->mixed_ratio;
    if (fallbacks) *fallbacks = b->mixed_fallbacks;
    return 0;
}

int crt_batch_info(crt_batch *b, double *ms, long long *nsys, int *variant, float *tune_ms) {
    if (!b) return CRT_ERR_ARG;
    if (ms) *ms = b->ms;
    if (nsys) *nsys = (long long)b->nsys_total;
    if (variant)
        for (int D = 0; D <= 16; ++D) variant[D] = b->variant[D];
    if (tune_ms)
        for (int D = 0; D <= 16; ++D)
            for (int v = 0; v < 4; ++v) tune_ms[D * 4 + v] = b->tune_ms[D][v];
    return 0;
}

int crt_batch_destroy(crt_batch *b) {
    if (!b) return 0;
    (void)hipSetDevice(b->device);
    for (auto &w : b->slots) (void)hipStreamSynchronize(w.st);
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return 0;
}

}  // extern "C"
