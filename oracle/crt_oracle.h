/*
 * crt_oracle.h -- CPU restatement of the CRTomo forward / Jacobian / CG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (crtomo_stable_amd) never links, loads or calls it.
 *
 * Conventions follow the reference's Fortran data exactly:
 *   - node, electrode, element and model numbers are 1-based;
 *   - nrel is (neltot x maxn) row-major (element-major): nrel[iel*maxn + m];
 *   - node coordinates are given per node (x[j] = sx(snr(j)), y[j] = sy(snr(j)));
 *   - band matrices use the reference layout: entry (i,j), j<=i, i-j<=mb,
 *     1-based, lives at a[i*mb + j - 1]  (a((mb+1)*sanz), src/kompab.f90:108-110);
 *   - elbg is (elanz, npk, kwnanz) column-major, kg is (relanz, eanz, kwnanz)
 *     column-major, kpot is (sanz, eanz, kwnanz) column-major, hpot (sanz, eanz),
 *     sens is (nanz, manz) column-major -- the layouts of the reference modules
 *     (src/femmod.f90, src/alloci.f90).
 * Complex numbers are C99 double _Complex (layout = Fortran COMPLEX*16).
 */
#ifndef CRT_ORACLE_H
#define CRT_ORACLE_H
#include <complex.h>

typedef double _Complex zdbl;

typedef struct {
    int sanz;          /* number of nodes */
    int mb;            /* band width */
    int typanz;        /* number of element types */
    const int *typ;    /* element type per type group (3,5,8,11,12,13) */
    const int *nelanz; /* elements per type group */
    const int *selanz; /* nodes per element per type group */
    int maxn;          /* max_nr_element_nodes */
    const int *nrel;   /* (neltot x maxn) row-major, 1-based node numbers */
    const double *x;   /* node x (sanz) */
    const double *y;   /* node y (sanz) */
    const int *rnr;    /* (relanz) area element adjacent to each boundary element, 1-based */
    int elanz;         /* area elements */
    int relanz;        /* boundary elements */
    int eanz;          /* electrodes */
    const int *enr;    /* (eanz) electrode node numbers, 1-based */
} orc_grid;

/* Bessel approximations (src/bessk0.f90, bessk1.f90, bessi0.f90, bessi1.f90). */
double orc_bessk0(double x);
double orc_bessk1(double x);
double orc_bessi0(double x);
double orc_bessi1(double x);

/* Quadrature (src/gauleg.f90, gaulag.f90, gammln.f90, rwaven.f90). */
void orc_gauleg(double x1, double x2, double *x, double *w, int n);
void orc_gaulag(double *x, double *w, int n, double alf);
double orc_gammln(double xx);
int orc_rwaven(const orc_grid *g, int nanz, const int *strnr, const int *vnr,
               int *kwnanz, double *kwn, double *kwnwi, double *amin, double *amax);

/* Element matrices (src/precal.f90). elbg: elanz*npk*kwnanz, relbg: relanz*npk,
 * kg: relanz*eanz*kwnanz.  sytop: mean height of the no-flow boundary or 0. */
int orc_precal(const orc_grid *g, int kwnanz, const double *kwn, int lsytop,
               double *elbg, double *relbg, double *kg, int *lbeta, double *sytop);
double orc_bsytop(const orc_grid *g);

/* Single-system steps (src/kompab.f90, scalab.f90, chol.f90, vre.f90, kompb.f90). */
int orc_kompab(const orc_grid *g, const zdbl *sigma, const double *elbg, const double *relbg,
               const double *kg, int kwnanz, int l, int k, zdbl *a, zdbl *b);
int orc_scalab(int sanz, int mb, zdbl *a, zdbl *b, double *fak);
int orc_chol(int sanz, int mb, zdbl *a);
void orc_vre(int sanz, int mb, const zdbl *a, const zdbl *b, zdbl *pot);
void orc_kompb(int sanz, int enode, zdbl *b, const double *fak);
/* dense complex-symmetric Cholesky (src/cholz.f90) */
int orc_cholz(zdbl *a, zdbl *p, int n);

/* Whole forward loop over (k, l) (src/fem.f90:328-397, src/inv.f90:415-484).
 * Systems with index list: if nsys>0 only (l,k) pairs sysl[i], sysk[i] (1-based) are
 * solved (used to time a bounded sample); kpot columns of other systems untouched. */
/* Dirichlet boundaries: type-13 elements (src/randb.f90, src/randdc.f90) and the
 * prescribed values of boundary.dat (src/randb2.f90, src/randbdc2.f90); the forward
 * loops apply them after kompab / kompadc as src/fem.f90:362-363 and
 * src/inv.f90:379-380 do.  orc_set_boundary keeps the caller's arrays (n = 0: none). */
void orc_set_boundary(int n, const int *nodes, const zdbl *values);
void orc_randb(const orc_grid *g, zdbl *a, zdbl *b);
void orc_randdc(const orc_grid *g, double *a, double *b);
void orc_randb2(const orc_grid *g, zdbl *a, zdbl *b);
void orc_randbdc2(const orc_grid *g, double *a, double *b);
/* reference conductivity (src/refsig.f90), analytic half-space potential (src/potana.f90) */
void orc_refsig(const orc_grid *g, const zdbl *sigma, zdbl *sigma0);
void orc_potana(const orc_grid *g, const double *kwn, int l, int k, const zdbl *sigma0p, zdbl *pota);
/* the forward loop of src/fem.f90:302-397 with singularity removal (lsr), the analytic
 * solution alone (lana) and the fictitious sink (lsink, nsink 1-based) */
int orc_forward_sw(const orc_grid *g, const zdbl *sigma, int kwnanz, const double *kwn, const double *elbg,
                   const double *relbg, const double *kg, int lbeta, int lsr, int lana, int lsink, int nsink,
                   zdbl *kpot);
/* half-space geometric factors (src/bkfak.f90); 93 on a vanishing denominator */
int orc_bkfak(const orc_grid *g, int nanz, const int *strnr, const int *vnr, double *kfak);
int orc_forward(const orc_grid *g, const zdbl *sigma, int kwnanz, const double *elbg,
                const double *relbg, const double *kg, int lbeta, zdbl *kpot,
                int nsys, const int *sysl, const int *sysk);

/* Back-transform and voltages (src/rtrafo.f90, bvolt.f90, bvolti.f90). */
void orc_rtrafo(int sanz, int eanz, int kwnanz, const double *kwnwi, const zdbl *kpot, zdbl *hpot);
void orc_bvolt(int nanz, const int *strnr, const int *vnr, const int *enr, int sanz,
               const zdbl *hpot, const double *kfak, zdbl *volt, zdbl *sigmaa);
int orc_bvolti(int nanz, const int *strnr, const int *vnr, const int *enr, int sanz,
               const zdbl *hpot, zdbl *volt, zdbl *sigmaa);

/* Sensitivities (src/bsensi.f90, src/bsens.f90). sens: nanz x elanz column-major. */
void orc_bsensi(const orc_grid *g, int nanz, const int *strnr, const int *vnr,
                int kwnanz, const double *kwnwi, const zdbl *kpot, const double *elbg,
                const zdbl *sigma, const zdbl *volt, zdbl *sens);
void orc_bsens(const orc_grid *g, int nanz, const int *strnr, const int *vnr, const double *strom,
               int kwnanz, const double *kwnwi, const zdbl *kpot, const double *elbg, zdbl *sens);

/* Regularisation (src/bsmatm_mod.f90:190-460, src/bnachbar.f90, src/relem.f90 espx/espy). */
void orc_bsmatmreg(const orc_grid *g, int nx, int nz, double alfx, double alfz, double *smatm);
void orc_bnachbar(const orc_grid *g, int *nachbar);
void orc_bsmatmtri(const orc_grid *g, const int *nachbar, double alfx, double alfz, double *smatm);

/* Normal-equation step (src/update.f90:63-136,179-231,236-293,302-351, src/cg_mod.f90:426-726).
 * ltri 0: smatm (manz,3), R stencil with nx; ltri 1/2: smatm (manz,maxn+1) with nachbar.
 * Outputs: cgfac, bvec0 (scaled RHS), dpar (scaled update), cgres (ncgmax+1), ncg. */
typedef struct {
    int nanz, manz, ltri, nx, maxn, ncgmax;
    double lam, eps;
    const zdbl *sens;       /* nanz x manz column-major */
    const double *wmatd;    /* nanz */
    const int *wdfak;       /* nanz */
    const zdbl *dat;        /* nanz */
    const zdbl *sigmaa;     /* nanz */
    const zdbl *par;        /* manz */
    const double *smatm;    /* manz x ncol column-major */
    const int *nachbar;     /* manz x maxn column-major (ltri 1/2) */
} orc_cg_problem;
void orc_update_rhs(const orc_cg_problem *p, double *cgfac, zdbl *bvec);
/* starting regularisation parameter (src/blam0.f90:66-168, computed branch):
 * the reference's triple loop over (j, i, k); fpi: real parts only (lfpi). */
double orc_blam0(const orc_cg_problem *p, int fpi, double alfx, double alfz);
void orc_cjggra(const orc_cg_problem *p, const double *cgfac, const zdbl *bvec,
                zdbl *dpar, double *cgres, int *ncg);
/* one CG matvec  out = C (J^H W J + lam R) C p  (src/cg_mod.f90:520-545,547-577,699-726) */
void orc_cg_matvec(const orc_cg_problem *p, const double *cgfac, const zdbl *pvec, zdbl *ap, zdbl *out);

/* DC twins (real-valued inversion ldc, final phase improvement lfpi):
 * src/kompadc.f90, scaldc.f90, choldc.f90, vredc.f90, kompbdc.f90 and the DC forward loop
 * src/inv.f90:355-398 (kpotdc [k][l][j]); rtrafo's and bvolti's DC branches; bsendc;
 * update's DC/FPI branches (kind 1: sensdc, kind 2: DBLE(sens)) and cjggdc. */
int orc_kompadc(const orc_grid *g, const zdbl *sigma, const double *elbg, const double *relbg,
                const double *kg, int l, int k, double *a, double *b);
int orc_scaldc(int sanz, int mb, double *a, double *b, double *fak);
int orc_choldc(int sanz, int mb, double *a);
void orc_vredc(int sanz, int mb, const double *a, const double *b, double *pot);
int orc_forward_dc(const orc_grid *g, const zdbl *sigma, int kwnanz, const double *elbg,
                   const double *relbg, const double *kg, int lbeta, double *kpotdc);
void orc_rtrafo_dc(int sanz, int eanz, int kwnanz, const double *kwnwi, const double *kpotdc, double *hpotdc);
int orc_bvolti_dc(int nanz, const int *strnr, const int *vnr, const int *enr, int sanz,
                  const double *hpotdc, zdbl *volt, zdbl *sigmaa);
void orc_bsendc(const orc_grid *g, int nanz, const int *strnr, const int *vnr, int kwnanz,
                const double *kwnwi, const double *kpotdc, const double *elbg, const zdbl *sigma,
                const zdbl *volt, double *sensdc);
void orc_update_rhs_dc(const orc_cg_problem *p, int kind, const double *sensdc, double *cgfac, zdbl *bvec);
void orc_cjggdc(const orc_cg_problem *p, int kind, const double *sensdc, const double *cgfac,
                const zdbl *bvec0, zdbl *dpar, double *cgres, int *ncg);

/* Linearised model uncertainty (src/bmcm_mod.f90:49-175 buncert, complex case, lam = p->lam):
 * bata (:176-289), bata_reg (:291-413, the reference's own write order per ltri 0, 1/2/5..14,
 * 3/4), bmcm (:415-615: chold + linvd + lower fill, or Gauss_dble when gauss), bres
 * (:617-688), bmcm2 (:690-765).  Matrices are manz x manz column-major; ata and ata_reg
 * are required, cov/res/cov2 may be NULL (res needs lres, cov2 needs lres and lcov2).
 * Returns 0, or 108 where the reference stops on a non-positive pivot (chold) or sets
 * errnr 108 after Gauss_dble. */
int orc_buncert(const orc_cg_problem *p, int lres, int lcov2, int gauss, double *ata, double *ata_reg,
                double *cov, double *res, double *cov2);

/* Cuthill-McKee renumbering (cutmckee/cutmck.f:126-337). Writes perm (sanz, 1-based new
 * number of each old node) and returns the band width; start_used gets the chosen start. */
int orc_cutmck(int sanz, int typanz, const int *nelanz, const int *selanz, int maxn,
               const int *nrel, int elanz, int *perm, int *start_used);

#endif
