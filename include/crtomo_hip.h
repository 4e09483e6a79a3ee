/*
 * crtomo_hip.h -- C-ABI of the MI355X (gfx950) CRTomo forward / Jacobian / CG library.
 *
 * Conventions (identical to the reference's Fortran module data, so a Fortran
 * shim passes its arrays unchanged):
 *   - node, element, electrode and model numbers are 1-based;
 *   - nrel is (neltot x maxn) element-major: nrel[iel*maxn + m] (the transpose
 *     of the Fortran nrel(neltot, maxn) array, which a shim passes as
 *     TRANSPOSE(nrel) or builds element-major);
 *   - x[j], y[j] are the coordinates of node j (sx(snr(j)), sy(snr(j)));
 *   - complex arrays are interleaved (re, im) doubles, the memory layout of
 *     Fortran COMPLEX*16;
 *   - multi-dimensional outputs use the Fortran column-major layouts of the
 *     reference modules: elbg(elanz,npk,kwnanz), kg(relanz,eanz,kwnanz),
 *     kpot(sanz,eanz,kwnanz), hpot(sanz,eanz), sens(nanz,manz).
 * Every function returns 0 or the reference's errnr code (src/error.txt):
 *   19 band width exceeded, 26 element geometry, 27 zero diagonal, 30 boundary
 *   node beyond sanz, 101 mixed and Dirichlet boundaries together,
 *   28 non-positive pivot, 35 boundary geometry, 73 duplicate electrodes,
 *   82 zero voltage, 97 allocation; 900+ are library codes (900 HIP error,
 *   901 call order, 902 argument, 903 unsupported configuration, 904 a solve larger
 *   than the static scratch of crt_cjg_noalloc).
 * A context is bound to one device and one HIP stream; calls on one context
 * must not overlap, except the single-system drop-ins crt_chol, crt_vre,
 * crt_choldc, crt_vredc and crt_cholz, which serialise on the context and may be
 * called from several host threads at once (the reference calls chol/vre inside
 * its OpenMP loop over wavenumbers).
 */
#ifndef CRTOMO_HIP_H
#define CRTOMO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct crt_ctx crt_ctx;

/* Create / destroy a context on HIP device `device`. */
int crt_create(crt_ctx **ctx, int device);
void crt_destroy(crt_ctx *ctx);
const char *crt_version(void);

/* Grid (replaces the elemmod/electrmod state read by src/relem.f90:60-151 and
 * src/relectr.f90).  Builds the assembly map used by the forward solve. */
int crt_set_grid(crt_ctx *ctx, int sanz, int mb, int typanz, const int *typ, const int *nelanz,
                 const int *selanz, int maxn, const int *nrel, const double *x, const double *y,
                 const int *rnr, int eanz, const int *enr);

/* Measurements (datmod strnr/vnr, src/rdatm.f90). */
int crt_set_survey(crt_ctx *ctx, int nanz, const int *strnr, const int *vnr);

/* Wavenumbers and weights (src/rwaven.f90:22).  kwn/kwnwi need room for 30. */
int crt_rwaven(crt_ctx *ctx, int *kwnanz, double *kwn, double *kwnwi, double *amin, double *amax);
/* Or set them explicitly (2D: kwnanz = 1, kwn = 0, swrtr = 0 in the reference). */
int crt_set_wavenumbers(crt_ctx *ctx, int kwnanz, const double *kwn, const double *kwnwi);

/* Prescribed boundary values (boundary.dat, src/rrandb.f90; randb2 in the forward,
 * src/randb2.f90:29-79): n nodes (1-based; errnr 30 beyond sanz) with complex values
 * (interleaved; the DC inversion takes the real parts as randbdc2 does).  Zero values
 * make the nodes Dirichlet nodes as the type-13 elements of the grid do (randb,
 * src/randb.f90:40-84, which crt_set_grid reads from the element types): identity rows
 * and columns, zero potential.  A source electrode on a Dirichlet node has the zero
 * potential of its cleared right-hand side where kompab assembles its system and
 * kompb's unit source elsewhere, as in the reference.  crt_forward returns 903 for
 * non-zero values (crt_batch_create also for a source on a Dirichlet node).  n = 0
 * removes them. */
int crt_set_boundary_values(crt_ctx *ctx, int n, const int *nodes, const double *values);

/* crmod.cfg's forward switches (src/fem.f90:177-203; reset by crt_set_grid):
 *   lsr:   singularity removal (mswitch bit 2): per (electrode, wavenumber) the analytic
 *          half-space potential pota of a unit current in the reference conductivity
 *          sigma0 (src/refsig.f90, src/potana.f90) is removed: the right-hand side is
 *          (A(sigma) - A(sigma0)) pota (src/kompab.f90:129-133) and kpot = the solution +
 *          pota (src/fem.f90:389); only with mixed boundaries (otherwise off, :311);
 *   lana:  the analytic solution alone (mswitch bit 0): kpot = pota, no system is
 *          solved (src/fem.f90:384-386); off when lsr is on (:198-203);
 *   nsink: the fictitious sink node (lsink, 1-based; 0 none): b(nsink) = 1 in every
 *          system (src/kompab.f90:183, src/kompb.f90:59, the DC twins alike).
 * lsr and lana are CRMod's complex forward only (903 with crt_set_inversion_kind 1/2 at
 * crt_forward); general right-hand sides run on the band method with 16-row tiles (903
 * for bands wider than 256, the FP32 forward or Dirichlet rows) and not in crt_batch_*. */
int crt_set_fem_switches(crt_ctx *ctx, int lsr, int lana, int nsink);

/* Geometric factors of the survey's measurements for a homogeneous half-space
 * (bkfak, src/bkfak.f90:48-136; wkfak, mswitch bit 1, with mixed boundaries): kfak
 * (nanz) for crt_voltages mode 0.  Returns 93 when a denominator vanishes. */
int crt_bkfak(crt_ctx *ctx, double *kfak);

/* Element matrices and mixed-boundary factors on the device (src/precal.f90:13);
 * errnr 101 for a grid with both mixed (type 11) and Dirichlet (type 13) boundaries. */
int crt_precal(crt_ctx *ctx, int lsytop, int *lbeta);
int crt_get_precal(crt_ctx *ctx, double *elbg, double *relbg, double *kg);

/* Complex conductivity per area element (sigmamod sigma). */
int crt_set_sigma(crt_ctx *ctx, const double *sigma);

/* Which of CRTomo's inversions the context serves (femmod ldc, invmod lfpi):
 * 0 (default) the complex inversion; 1 the DC inversion (ldc): the forward assembles
 * DBLE(sigma) (kompadc/scaldc/choldc/vredc, src/inv.f90:355-398), crt_voltages mode 1
 * runs bvolti's DC branch (src/bvolti.f90:55-66), crt_sens mode 1 runs bsendc
 * (src/bsendc.f90:1), crt_get_kpot/crt_get_hpot/crt_get_sens hand out REAL*8 kpotdc,
 * hpotdc, sensdc and crt_set_sens takes a REAL*8 sensdc(nanz, manz); 2 the final phase
 * improvement (lfpi).  For 1 and 2, crt_update/crt_cjg run cjggdc (src/cg_mod.f90:106)
 * on the real Jacobian from DBLE(bvec) (DC) or DIMAG(bvec) (FPI), and crt_update stores
 * dpar = i*dpar for FPI (src/update.f90:302-304); crt_blam0 takes the sensdc branch for 1.
 * crt_buncert runs bata on the real Jacobian for 1 (src/bmcm_mod.f90:197-206) and the
 * complex branches for 2 (the caller passes lam_cri and wmatdr, :56-83).
 * The condensed forward runs its kernels in real arithmetic for 1.  The FP32 copy of
 * J and the FSAI preconditioners work on the real J for 1 and 2.  Not for 1/2:
 * bvolt/bsens (mode 0, DC) and the learned preconditioner (903).  A change of kind
 * discards sigma and its results: call crt_set_sigma after it. */
int crt_set_inversion_kind(crt_ctx *ctx, int kind);

/* The forward loop of src/fem.f90:328-397 / src/inv.f90:415-484 for every
 * (electrode, wavenumber): kompab + scalab + chol + vre (or kompb + vre).
 * kpot stays on the device. */
int crt_forward(crt_ctx *ctx);
int crt_get_kpot(crt_ctx *ctx, double *kpot);

/* rtrafo + bvolt (mode 0, kfak may be NULL = 1) or rtrafo + bvolti (mode 1)
 * (src/rtrafo.f90:13, src/bvolt.f90:7, src/bvolti.f90:8). */
int crt_voltages(crt_ctx *ctx, int mode, const double *kfak, double *volt, double *sigmaa);
int crt_get_hpot(crt_ctx *ctx, double *hpot);

/* Sensitivities: bsensi (mode 1, src/bsensi.f90:23) or bsens (mode 0,
 * src/bsens.f90:23, strom may be NULL = 1).  J stays on the device until the next
 * crt_sens / crt_set_sens or a new grid, survey, boundary values, forward switches or
 * inversion kind; forwards of other models keep it, as the reference's lambda search
 * updates with the Jacobian of the iteration's start model (src/inv.f90:704-736). */
int crt_sens(crt_ctx *ctx, int mode, const double *strom);
int crt_get_sens(crt_ctx *ctx, double *sens);

/* Upload a Jacobian sens(nanz, manz) (column-major complex) instead of crt_sens. */
int crt_set_sens(crt_ctx *ctx, int nanz, int manz, const double *sens);

/* Regularisation matrix for the CG (src/bsmatm_mod.f90:190-460):
 * ltri 0: smatm(manz,3) 5-point stencil of an nx-wide regular grid;
 * ltri 1/2: smatm(manz,maxn+1) with nachbar(manz,maxn+1) (1-based neighbours). */
int crt_set_regularization(crt_ctx *ctx, int ltri, int nx, int manz, int maxn, const double *smatm,
                           const int *nachbar);

/* cjg (src/cg_mod.f90:84, complex case -> cjggra :426): Fletcher-Reeves CG on
 * C (J^H W J + lam R) C y = bvec with W = wmatd*wdfak; writes the raw CG
 * solution dpar (before the cgfac scaling of src/update.f90:302), the residual
 * history cgres(ncgmax+1) (cgres(1) = ncg) and ncg. */
int crt_cjg(crt_ctx *ctx, double lam, double eps, int ncgmax, const double *wmatd, const int *wdfak,
            const double *cgfac, const double *bvec, double *dpar, double *cgres, int *ncg);

/* crt_cjg without any allocation.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as crt_cjg;
 * every buffer of the solve (weights, cgfac, right-hand side, CG vectors, partial
 * sums, scalar states, residual history) lives in one static 64 MiB device array per
 * GPU, held from library load until the process ends (about 12 MB for 40000 cells,
 * 4500 data and ncgmax 4000).  It needs the Jacobian and regularisation already on
 * the context (crt_sens / crt_set_sens, crt_set_regularization / crt_bsmatm), the
 * complex inversion, FP64 J and the Jacobi scaling (903 otherwise); a solve whose
 * buffers exceed the array returns 904.  Calls on one GPU share the array: calls that
 * overlap on one GPU (from any contexts or threads) race. */
int crt_cjg_noalloc(crt_ctx *ctx, double lam, double eps, int ncgmax, const double *wmatd, const int *wdfak,
                    const double *cgfac, const double *bvec, double *dpar, double *cgres, int *ncg);

/* update (src/update.f90:27, complex case, no prior / reference model): RHS and
 * cgfac from J on the device, cjg, dpar <- cgfac*dpar, par += step*dpar.
 * sigmaa = NULL uses the forward values of the last crt_voltages call.
 * cgfac and bvec0 (the scaled RHS) may be NULL. */
int crt_update(crt_ctx *ctx, double lam, double eps, int ncgmax, double step, const double *wmatd,
               const int *wdfak, const double *dat, const double *sigmaa, double *par, double *dpar,
               double *cgres, int *ncg, double *cgfac, double *bvec0);

/* Storage precision of J in the CG iterations of crt_cjg / crt_update: 64 (default,
 * the reference's arithmetic) or 32 (complex FP32 copy of J, made once per Jacobian;
 * products accumulate and all inner products run in FP64; the RHS and cgfac pass
 * keeps FP64 J).  No counterpart in the reference: the mixed-precision form halves
 * the bytes of each iteration.  The DC and final-phase-improvement kinds keep a real
 * FP32 copy of their real J (4 bytes per entry). */
int crt_set_cg_precision(crt_ctx *ctx, int jbits);

/* Preconditioner of the CG iterations of crt_cjg / crt_update.  0 (default): the
 * reference's Jacobi scaling C alone.  1: in addition a factorised sparse
 * approximate inverse M = G^H G of B = C (J^H W J + lam R) C, G lower triangular
 * on the regularisation stencil (row j: its lower neighbours and j), built on the
 * device after cgfac from the Hermitian blocks B(S_j, S_j).  2: the same on a wider
 * lower pattern (regular grids: two cells left, the three cells above, two above).
 * 3: the learned preconditioner below; 4: its analytic prior alone (the FSAI row
 * on the stencil with the coupling between the two neighbours neglected, no network,
 * no training).  The stop test stays on |r|^2 of the scaled
 * system, as in the reference; only the iteration count changes.  No counterpart
 * in the reference.  The DC and final-phase-improvement kinds take kinds 0-2 (their FSAI
 * is built on the real J); 3 and 4 return 903 there. */
int crt_set_cg_preconditioner(crt_ctx *ctx, int kind);

/* Learned preconditioner (kind 3 of crt_set_cg_preconditioner, regular grids): a
 * per-cell network 8 -> 64 -> 64 -> 64 -> 5 emits the row of G on the stencil
 * pattern {N, W, j} (as a correction to the estimate that neglects the N-W
 * coupling) from the cell's coverage, its couplings B(W,j), B(N,j), B(N,W) in the
 * scaled normal matrix and lam R_jj / coverage; inference with BF16 weights and
 * FP32 accumulation.  If it does not reach the stop test
 * within ncgmax iterations, the CG reruns with the Jacobi scaling alone.  No
 * counterpart in the reference.
 *   crt_pc_train: adds the system of the last crt_update / crt_cjg call to the
 *     training set (targets: that system's FSAI rows, kind 1) and runs `steps`
 *     full-batch Adam steps (learning rate lr) on every system added so far.
 *     loss_first / loss_last (may be NULL) receive the mean squared error before
 *     the first and after the last step.
 *   crt_pc_weights: get and/or set the model, crt_pc_size() floats (16 feature
 *     standardisation values, then the parameters).
 *   crt_pc_grad: loss and parameter gradient on the training set (no update).
 *   crt_pc_data: the training set's raw features (n x 8), targets (the corrections
 *     to the analytic estimate) and mask (n x 5); *n holds the buffers' capacity in
 *     cells on entry (902 when it is too small; NULL buffers only query) and the
 *     number of cells added so far on return.
 *   crt_pc_reset: drops the training set, the model and its standardisation (the
 *     next crt_pc_train starts a new network). */
int crt_pc_train(crt_ctx *ctx, int steps, double lr, double *loss_first, double *loss_last);
int crt_pc_size(void);
int crt_pc_weights(crt_ctx *ctx, float *get, const float *set);
int crt_pc_grad(crt_ctx *ctx, double *loss, float *grad);
int crt_pc_data(crt_ctx *ctx, int *n, float *feat, float *tgt, float *mask);
int crt_pc_reset(crt_ctx *ctx);

/* Smoothness regularisation built on the device from the grid of crt_set_grid and
 * installed for crt_cjg / crt_update (replaces crt_set_regularization):
 * ltri 0 -> bsmatmreg (src/bsmatm_mod.f90:190-362) on the regular nx x nz quad grid,
 * smatm (manz, 3); ltri 1/2 -> bnachbar + bsmatmtri (src/bnachbar.f90:44-90,
 * src/bsmatm_mod.f90:364-460), smatm (manz, maxn+1).  smatm (column-major, may be
 * NULL) receives a copy. */
int crt_bsmatm(crt_ctx *ctx, int ltri, int nx, int nz, double alfx, double alfz, double *smatm);

/* Edge neighbours of the area elements (src/bnachbar.f90:44-90), host code:
 * nrel (elanz, maxn) element-major 1-based; nachbar (elanz, maxn+1) column-major,
 * last column = neighbour count. */
int crt_bnachbar(int elanz, int maxn, const int *nrel, int *nachbar);

/* Grid statistics of besp_elem (src/besp_elem.f90:1-97) for the grid of crt_set_grid:
 * stats[11] = esp_min, esp_max, esp_mit, esp_med, esp_std (distances of each cell's
 * centroid to its last edge neighbour), grid_min, grid_max, grid_minx, grid_maxx,
 * grid_miny, grid_maxy (over all centroid pairs, the O(elanz^2) scan on the device).
 * gstat_path (may be NULL) receives inv.gstat in the reference's format. */
int crt_besp_elem(crt_ctx *ctx, double *stats, const char *gstat_path);

/* Starting regularisation parameter (src/blam0.f90:120-175) from the Jacobian on
 * the device: lammax = 5 * 2/(alfx+alfz) * mean_j |(J^H W J 1)_j|, W = wmatd*wdfak;
 * fpi != 0 uses the real parts only (the lfpi branch).  The fixed-lambda branches
 * (llamf, lamnull) stay with the caller. */
int crt_blam0(crt_ctx *ctx, int fpi, const double *wmatd, const int *wdfak, double alfx, double alfz,
              double *lammax);

/* CRMod tomodir without the Fortran host: crmod.cfg (the ***FILES*** block: elem,
 * elec, rho and config paths, then the potentials / measurements / sensitivities
 * switches and output paths; relative paths from the directory of crmod.cfg),
 * elem.dat (src/relem.f90:60-151 with the border check of :197-229), elec.dat
 * (src/relectr.f90), rho.dat (src/rsigma.f90:168), config.dat (src/rdatm.f90);
 * runs the forward and writes volt.dat when the measurements switch is T, the
 * potential files of src/bpot.f90 + src/wpot.f90 when the potentials switch is T
 * (errnr 15 for a base name too long for the numbering) and the sensitivity files of
 * src/bsens.f90 + src/wsens.f90 when the sensitivities switch is T (file n of a series:
 * the name with n zero-filled to int(log10(nanz))+1 digits before its extension). */
int crt_tomodir_crmod(crt_ctx *ctx, const char *cfg_path);
/* Sizes of the tomodir's grid, electrodes and configurations (host only). */
int crt_tomodir_sizes(const char *cfg_path, int *sanz, int *mb, int *elanz, int *eanz, int *nanz);
/* volt.dat as src/wdatm.f90:58-76 writes it: nanz, then strnr, vnr, REAL*4 |sigmaa|
 * and 1000 * phase; a phase near +-pi is shifted by pi and swaps the potential
 * electrodes of vnr.  Host only. */
int crt_write_volt(const char *path, int nanz, const int *strnr, const int *vnr, const double *sigmaa);

/* crtomo.cfg as src/rall.f90:141-593 reads it: defaults, '#' comment lines
 * (src/get_unit.f90 read_comments), paths cut at their first blank (clear_string), the
 * optional lines with their fallbacks (mswitch, seed/variance of the prior, ltri read
 * as I2 with the fixed-lambda bits 32/64 and its value, the MGS beta, the seed of coupled
 * noise), the switches derived from mswitch and the validation errors (61 itmax, 89
 * regular grid, 91 background, 104 error model, 105 phase error model).  Logical
 * switches are 0/1; paths hold at most 80 characters. */
typedef struct crt_crtomo_cfg {
    int mswitch;
    char elem[81], elec[81], volt[81], ramd[81];
    int ldiff;
    char dd0[81], dm0[81], dfm0[81], dstart[81];
    int iseedpri;
    double modl_stdn;
    int nx, nz;
    double alfx, alfz;
    int itmax, ldc, lrobust, lfphai;
    double stabw0, stabm0, stabpA1, stabpB, stabpA2, stabp0;
    int lrho0;
    double bet0, pha0;
    int lagain, swrtr, lsink, nsink, lrandb2;
    char drandb[81];
    int ltri, llamf;
    double lamfix, betamgs;
    int iseed;
    int lffhom, lnse, lnse2, lnsepri, lstart, lprior;
    int lsens, lcov1, lres, lcov2, lgauss, lelerr, lphi0, lsytop, lvario, lverb;
} crt_crtomo_cfg;
int crt_read_crtomo_cfg(const char *cfg_path, crt_crtomo_cfg *cfg);

/* CRTomo tomodir without the Fortran host (the inversion of src/inv.f90:136-1100):
 * crtomo.cfg (crt_read_crtomo_cfg; relative paths from its directory, where crt.lamnull
 * and inv.lastmod are read and written as the reference does in its working directory),
 * elem.dat, elec.dat and the measurements with their error model (src/rdati.f90); the
 * starting model (src/bsigm0.f90); then the Gauss-Newton loop with the reference's
 * polarity check, misfit, convergence tests, lambda search, step-length control,
 * parabolic step fit and final phase improvement, every forward, Jacobian,
 * regularisation, blam0 and update through this library; writes inv.ctr, run.ctr,
 * eps.ctr, cjg.ctr, rhoNN.mag/.modl/.pha/.sig and voltNN.dat per iteration
 * (src/wout.f90), coverage.mag (src/bbsens.f90) and, with mswitch bits 1-3, the .diag
 * files of buncert.  Refused with 903: another dataset (lagain), data noise (a negative
 * error level or crt.noisemod), difference and prior inversions, robust inversion,
 * variograms, regularisations other than ltri 0, 1, 2, decouplings.dat and
 * electrode_capacitances.dat.  iterations / final_rms (may be NULL) receive the last
 * iteration index and its data RMS. */
int crt_tomodir_crtomo(crt_ctx *ctx, const char *cfg_path, int *iterations, double *final_rms);

/* Cuthill-McKee renumbering (cutmckee/cutmck.f:126-337): perm[old-1] = new
 * number; returns the band width in *mb. */
int crt_cutmck(int sanz, int typanz, const int *nelanz, const int *selanz, int maxn, const int *nrel,
               int elanz, int *perm, int *mb);

/* Linearised model uncertainty of the complex inversion, buncert
 * (src/bmcm_mod.f90:49-175): ata = Re(J^H W J) (bata), ata_reg = ata + lam R in
 * the reference's own fill order (bata_reg; ltri 0, 1, 2), cov = ata_reg^-1 by
 * Cholesky (chold + linvd) or, with gauss, by LU (the Gauss_dble branch), res =
 * cov ata (bres, when lres) and cov2 = res cov (bmcm2, when lres and lcov2).
 * Uses J on the device (crt_sens / crt_set_sens) and the regularisation of
 * crt_set_regularization / crt_bsmatm; W = wmatd * wdfak.  Each output is a
 * manz x manz column-major host array or NULL; diag (NULL or 5*manz) gets the
 * five main diagonals in that order (what the reference writes to ata.diag,
 * ata_reg.diag, cov1_m.diag, res_m.diag, cov2_m.diag).  Returns 108 when the
 * factorisation rejects ata_reg (the reference's errnr; ata and ata_reg are
 * still returned). */
int crt_buncert(crt_ctx *ctx, double lam, const double *wmatd, const int *wdfak, int lres, int lcov2, int gauss,
                double *ata, double *ata_reg, double *cov, double *res, double *cov2, double *diag);

/* Single-system drop-ins for the named Fortran routines, host arrays in/out:
 * chol(a) (src/chol.f90:7), vre(a,b,pot) (src/vre.f90:5), cholz(a,p,n,...)
 * (src/cholz.f90:1).  Band layout: entry (i,j), j<=i, 1-based, at
 * a[i*mb + j - 1], length (mb+1)*sanz. */
int crt_chol(crt_ctx *ctx, int sanz, int mb, double *a);
int crt_vre(crt_ctx *ctx, int sanz, int mb, const double *a, const double *b, double *pot);
int crt_cholz(crt_ctx *ctx, int n, double *a, double *p);
/* DC twins: choldc(a) (src/choldc.f90:7), real band in/out, errnr 28 where the
 * reference finds s <= 0; vredc(a,b,pot) (src/vredc.f90:1), real band factor and
 * right-hand side, complex pot (sanz COMPLEX*16) out. */
int crt_choldc(crt_ctx *ctx, int sanz, int mb, double *a);
int crt_vredc(crt_ctx *ctx, int sanz, int mb, const double *a, const double *b, double *pot);

/* Forward precision (no counterpart in the reference, whose chol/vre run in
 * COMPLEX*16, src/chol.f90:7, src/vre.f90:5).  bits 64 (default): FP64 band
 * factorisation.  bits 32: the factorisation runs in complex FP32 on the matrix
 * cores and every (electrode, wavenumber) system is refined in FP64 (residual
 * A x - g from the FP64 matrix, correction through the FP32 factor) until the
 * correction's max-norm is below tol times the potentials' (or stalls below
 * max(tol, 1e-12)), at most max_sweeps sweeps; a batch that does not converge
 * is redone in FP64.  Needs a band of 33..256 nodes (3..16 tiles of 16): 903 otherwise.
 * crt_get_refine_info: most sweeps of any batch, largest final correction
 * ratio, batches redone in FP64, all of the last crt_forward. */
int crt_set_forward_precision(crt_ctx *ctx, int bits, double tol, int max_sweeps);
int crt_get_refine_info(crt_ctx *ctx, int *sweeps, double *ratio, int *fallbacks);

/* How crt_forward solves the systems of a grid with mixed boundaries (no counterpart
 * in the reference, which factorises every (electrode, wavenumber) system,
 * src/fem.f90:328-397; the potentials agree to rounding):
 *   0  the band factorisation of every system (kompab + scalab + chol + vre);
 *   1  static condensation onto the mixed-boundary nodes: the interior block, the same
 *      for every source, factorised once per wavenumber (cut into independent band
 *      pieces by separators that join the dense block), a dense complex-symmetric
 *      factorisation of the boundary Schur complement per system; 903 where it does
 *      not apply (no type-11 elements, Dirichlet rows, the FP32 forward, an interior
 *      band of one or more than 16 tiles of 16, more than 640 boundary nodes);
 *  -1  (default) 1 where it applies, else 0.
 * crt_get_forward_method: the method of the last crt_forward. */
int crt_set_forward_method(crt_ctx *ctx, int method);
int crt_get_forward_method(crt_ctx *ctx, int *method);
/* The condensation plan of the grid (made by the first condensed crt_forward):
 * info[8] = applicable (0/1), interior nodes in the pieces, boundary nodes, interior
 * band width, interior band tiles of 16, interior numbering (0 the grid's, 1 by x,
 * 2 by y), interior pieces, separator nodes. */
int crt_get_condense_info(crt_ctx *ctx, int *info);

/* Device time of the last call of each phase, milliseconds:
 * [0] diag, [1] factor, [2] solve, [3] voltages, [4] sensitivities,
 * [5] precal, [6] cg total, [7] cg matvec (mean per iteration),
 * [8] refinement sweeps of the FP32 forward (total), [9] the dense boundary
 * factorisations and solves of the condensed forward (total), [10] its interior
 * potentials (the product with the boundary solution), [11] its elimination of the
 * separators (once per wavenumber) with the boundary terms of every source.  The
 * condensed forward counts its interior band factorisation as [1] and its
 * multi-right-hand-side substitutions and Schur complements as [2]. */
int crt_get_times(crt_ctx *ctx, double *ms, int n);

/* Batched forward of many models of uneven size (crt_forward of every context,
 * as one graph launch).  No counterpart in the reference, whose host runs one
 * model per process (src/fem.f90:328-397); a caller with many surveys or models
 * (model ensembles, training data for the learned preconditioner) replaces N
 * crt_forward calls by one crt_batch_forward.
 *   crt_batch_create: plans each context's systems (every context needs precal
 *     and sigma; all on one device) in chunks through one shared workspace of
 *     at most `workspace_bytes` (0: 64 GiB); flags bit 0 times each factor
 *     kernel variant per band width on the first chunk and keeps the fastest.
 *   crt_batch_forward: first call captures every launch into a hipGraph, later
 *     calls replay it; the contexts' sigma may change between calls, their grid,
 *     survey and wavenumbers may not.  Leaves kpot on each context, as
 *     crt_forward does (crt_voltages, crt_sens follow per context).
 *   crt_batch_info: device time of the last graph launch (ms), systems per
 *     launch, variant[17] per band width (0 default), tune_ms[17*4] timings.
 *   flags bit 1: the mixed-precision variant: each chunk's systems factorised in
 *     complex FP32 on the matrix cores, solved in FP32 and refined in FP64 (residual
 *     from the FP64 matrix, correction through the FP32 factor) for up to each
 *     context's refine_max sweeps inside the graph, a sweep skipped once the device
 *     check has accepted the chunk (crt_set_forward_precision's tolerance and rule);
 *     a model with a chunk that did not converge, or an FP32 pivot failure, is redone
 *     with the FP64 kernels on the batch's workspace after the graph.  Band widths of 3..16 tiles (903
 *     otherwise).  crt_batch_refine_info: the most sweeps a chunk ran, the largest
 *     accepted correction ratio and the number of models redone, of the last launch.
 * A batch owns its stream; calls on a batch and on its contexts must not overlap. */
typedef struct crt_batch crt_batch;
int crt_batch_create(crt_ctx *const *ctxs, int n, size_t workspace_bytes, int flags, crt_batch **out);
int crt_batch_forward(crt_batch *b);
int crt_batch_info(crt_batch *b, double *ms, long long *nsys, int *variant, float *tune_ms);
int crt_batch_refine_info(crt_batch *b, int *sweeps, double *ratio, int *fallbacks);
int crt_batch_destroy(crt_batch *b);

#ifdef __cplusplus
}
#endif
#endif
