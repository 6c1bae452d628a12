/*
 * epsr_oracle.h -- CPU restatement of the EPSR refinement hot loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it to compute
 * the reference answer on the host.  The product (libepsr_mi355x.so) never
 * links, loads or calls it.
 *
 * Every function follows the Python translation of EPSR26 statement by
 * statement in IEEE float32 (compiled -ffp-contract=off, no fast-math), and
 * cites the reference file:line it restates (paths relative to the
 * reference repository root).  Two libm flavours are built from the same
 * source:
 *   liboracle_glibc.so  -- expf/cosf/sinf/logf/... from the host glibc,
 *                          exactly as the reference calls them;
 *   liboracle_cr.so     -- the same functions evaluated in double and
 *                          rounded once (correctly rounded), the arithmetic
 *                          the GPU kernels use.
 *
 * Index conventions (shared with include/epsr_mi355x.h):
 *   atoms, molecules, atom types, pair types, datasets: 0-based;
 *   radial-grid and Q-grid indices keep the reference's 1-based numbering,
 *   so every grid array has a leading unused element 0 and table[ic*ld + iz]
 *   is the reference's table(iz, ic+1).
 */
#ifndef EPSR_ORACLE_H
#define EPSR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- arithmetic contract (src/epsr26/fortran/fmath.py) ---- */
int   ora_nint(float x);               /* fmath.py:222-239 */
int   ora_ifix(float x);               /* fmath.py:252-260 */
float ora_powi(float x, int m);        /* fmath.py:186-200 */
float ora_sqrtf(float x);              /* fmath.py:152-157 */
float ora_expf(float x);
float ora_logf(float x);
float ora_sinf(float x);
float ora_cosf(float x);
float ora_atanf(float x);
float ora_sinhf(float x);
float ora_powf(float x, float y);
int   ora_libm_is_cr(void);

/* ---- ran1 (src/epsr26/library/auxiliary_routines.py:324-385) ---- */
typedef struct { int32_t mcall, idum, iy, iv[32]; } ora_ran1_t;
float ora_ran1(ora_ran1_t *s);
void  ora_ran1_stream(ora_ran1_t *s, int n, float *out);

/* ---- box description ---- */
typedef struct {
    int nmol, natom, ntype, npairs, nmoltype;
    const float *xa, *ya, *za;        /* [natom] atom offsets from centre (box fractions) */
    const int *itype;                 /* [natom] */
    const int *npairref;              /* [ntype*ntype] */
    const int *n1extr, *n2extr;       /* [npairs] */
    const int *naf, *nal;             /* [nmol] first/last atom (inclusive) */
    const int *mtype;                 /* [nmol] */
    const float *cfrac;               /* [ntype] */
    float natomr, sizefactor, dmax, cellst;
} ora_box_t;

/* molecule centres are mutable state and passed separately */
typedef struct { float *xadd, *yadd, *zadd; } ora_centres_t;

/* ---- minimum image (src/epsr26/library/minimagec.py:48-62) ---- */
void ora_minimage(float *x, float *y, float *z);

/* ---- atogen (src/epsr26/library/ato_manipulation_rev4.py:3640-3899) ----
 * wav[(k*npairs + ic)*ldw + iz], k=0 intra, k=1 inter, iz = 1..npt1.
 * irefpcof[ic] indexes rminex.  Returns the (possibly adjusted) ngrsamples.
 * csampled (ntype) and origin flags (nmol, may be NULL) are diagnostics.  */
int ora_atogen(const ora_box_t *b, const ora_centres_t *c, int npt1, int ldw,
               const float *ycoh, const int *irefpcof, const float *rminex,
               int ngrsamples, ora_ran1_t *rng, float *wav, float *potex,
               float *potexmax, float *csampled, int *origin_flags);
/* raw integer counts from the same loops (before the float normalisation):
 * cnt[(k*npairs+ic)*ldw+iz] and potex counts; used by the exact-count tests. */
int ora_atogen_counts(const ora_box_t *b, const ora_centres_t *c, int npt1, int ldw,
                      const int *irefpcof, const float *rminex, int ngrsamples,
                      ora_ran1_t *rng, int64_t *cnt, int64_t *potexcnt);

/* ---- gensq (src/epsr26/programs/epsr_standalone_rev3.py:4730-4802) ----
 * foqols[i*ldf + j] = reference foqols(j, i) (i = Q index 1..nq, j = r index)
 * sofq[(k*npairs + ic)*ldq + i]. natomsphere[type] (NULL = all 1). */
void ora_gensq(int nq, int npairs, int npt1, int ldw, const float *wav,
               const float *ycoh, const float *foqols, int ldf,
               const int *n2extr, const float *natomsphere, float *sofq, int ldq);

/* ---- stog + lorch (src/epsr26/library/ato_manipulation_rev4.py:4066-4124) ----
 * q[1..], r[1..]; sofq[in*ldq + i]; gofr[in*ldg + j]; nqmin/nqmax per column (1-based). */
float ora_lorch(float x, float pi);
void ora_stog(float rho, const float *q, const float *sofq, int ldq, const float *r,
              float *gofr, int ldg, const int *nqmin, const int *nqmax, int nr,
              int npairs, const float *xwt);

/* ---- potential tables and pair energies (src/epsr26/library/molfit.py) ---- */
typedef struct {
    int ld, npt2;
    const float *xstar;                /* [ld] */
    const float *xavs, *xavs1;         /* [npairs*ld] empirical potential + gradient */
    const float *ecoul, *ecoul1;       /* [npairs*ld] reference potential + gradient */
} ora_tables_t;

float ora_en_atom_inter_scale(float x1, float y1, float z1, float x2, float y2, float z2,
                              float stepcl, const ora_tables_t *t, int icref, float scalefac);
float ora_en_mole_pair(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t, int im, int jm);

/* ---- cell lists (src/epsr26/library/cellneighbourlist.py:23-330) ---- */
typedef struct {
    int ncelldiva, ncelldivb, ncelldivc, ncell, ncelldivsq, maxcellno, ncellneighb;
    int *nmolsincell;      /* [ncell] */
    int *molsincell;       /* [ncell*maxcellno] */
    int *molcellno;        /* [nmol] (0-based cell) */
    int *iacellneighb, *ibcellneighb, *iccellneighb;   /* [ncell] */
} ora_cells_t;
int  ora_cells_init(ora_cells_t *cl, int nmolcellin, const float *xadd, const float *yadd,
                    const float *zadd, int nmol, float rmax);
void ora_cells_free(ora_cells_t *cl);
int  ora_make_neighbour_list_from_cells(const ora_cells_t *cl, int ic, int *nbref);
int  ora_get_nnbrefmax(const ora_cells_t *cl);

/* ---- pair-energy ledger (src/epsr26/library/ergpairreference.py:63-213) ---- */
typedef struct {
    int nmol, nnbrefmax, nergpair, nergpairmax;
    float *ergpair;        /* [nergpairmax] (0-based entry numbers) */
    int *molergpair;       /* [nmol*nnbrefmax] partner (0-based) */
    int *ergpairref;       /* [nmol*nnbrefmax] entry in ergpair (0-based) */
    int *nergpairm;        /* [nmol] */
} ora_ledger_t;
void ora_ledger_free(ora_ledger_t *l);

/* ---- fmolec1 (src/epsr26/library/ato_manipulation_rev4.py:2374-2528) ----
 * eintra[nmol] is the intramolecular energy of each molecule (en_mole_intra).
 * Builds the ledger; returns ftest (fp32 as the reference), ftestb, and the
 * FP64 total of the same stored pair energies. */
int ora_fmolec1(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t,
                const ora_cells_t *cl, const float *eintra, ora_ledger_t *l,
                float *ftest, float *ftestb, double *ftest_f64);

/* ---- press (src/epsr26/library/ato_manipulation_rev4.py:3901-4063) ----
 * Whole-box intermolecular energy (kJ/mol per molecule) and virial pressure (kbar)
 * over the ledger's partner lists; slote/slotw [nmol*nnbrefmax] (may be NULL)
 * receive each ledger slot's energy and virial (0 where the partner is not later). */
void ora_press(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t, const ora_ledger_t *l,
               int npt1, float simt, float volume, float *energy, float *pressc, double *presst,
               float *slote, float *slotw);

/* ---- sumsq / sumgr (src/epsr26/programs/epsr_standalone_rev3.py:3642-3685, 3578-3637) ----
 * running averages and error bars, intra (k = 0) and inter (k = 1) rows
 * [(k*npairs + ic)*ld + j]; fc1, fc2 the accumulation factors of common /sumvalues/. */
void ora_sumsq(int npairs, int nq, int ldq, const float *qwind, double fc1, double fc2, float *soq, float *sq2);
void ora_sumgr(int npairs, int nz, int ldw, const float *wav, const float *ycoh, double fc1, double fc2,
               float *gor, float *gr2, float *entsig);

/* ---- invert_cwt (src/epsr26/programs/epsr_standalone_rev3.py:5423-5672) ----
 * cwt [nfile*npairs], cwtt [npairs*(npairs+nfile)], carray [npairs+nfile],
 * carrayold [(npairs+nfile)^2], 0-based; returns the number of step-size rounds. */
int ora_invert_cwt(double *cwt, double *cwtt, double *carray, double *carrayold, float feedback, const float *efacm,
                   int nfile, int npairs, ora_ran1_t *rng);

/* ---- en_change (src/epsr26/library/molfit.py:805-919) ----
 * Uses the ledger list of im (make_neighbour_list_from_mols) and the current
 * coordinates; ergnew[nnbref] out (slot 0 = ergold[0]). Returns fdiff. */
float ora_en_change(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t,
                    const ora_ledger_t *l, int im, float *ergnew, int *nnbref_out);

/* ---- molfit, whole-molecule branch (src/epsr26/library/molfit.py:51-541,671-759) ---- */
typedef struct {
    float *wholerot, *wholetrans;      /* [nmoltype] step sizes (mutated) */
    int *nwhole, *nmovem;              /* [2*nmoltype]: [0]=attempts,[1]=rejections */
    int *nstepreject;                  /* [nmol] */
    float simt, stepr, stepm, rejrate;
    int ireset;
} ora_mcstate_t;
/* xa/ya/za are mutated by rotation: pass writable copies through b (cast). */
int ora_molfit(const ora_box_t *b, float *xa, float *ya, float *za, const ora_centres_t *c,
               const ora_tables_t *t, const ora_cells_t *cl, ora_ledger_t *l,
               ora_mcstate_t *mc, ora_ran1_t *rng, float *fadds, int *naccept_out);

/* ---- delayed acceptance (no counterpart in the reference) ----
 * ora_molfit with a two-stage Metropolis test: des[nmol] is the surrogate's dEs of each
 * molecule's move; stage 1 passes when ora_da_uniform(seed, im) < etest(dEs/simt), stage 2
 * tests ran1 < etest((fdiff - dEs)/simt); the ran1 number is drawn for every move.  des NULL
 * is plain molfit.  target_out[nmol] (nullable, by molecule): the exact energy change of
 * each move over its partners later in the chain, FP64 sum of en_change's terms;
 * fdiff_out[nmol] (nullable, by molecule): en_change's fdiff of every move. */
float ora_da_uniform(unsigned long long seed, int m);
int ora_molfit_da(const ora_box_t *b, float *xa, float *ya, float *za, const ora_centres_t *c,
                  const ora_tables_t *t, const ora_cells_t *cl, ora_ledger_t *l,
                  ora_mcstate_t *mc, ora_ran1_t *rng, float *fadds, int *naccept_out,
                  const float *des, unsigned long long seed, int *nrej_out, double *target_out,
                  float *fdiff_out);

/* ---- intramolecular energies and moves (ora_intra.c; molfit.py:545-655, 922-1524;
 *      ato_manipulation_rev4.py:442-475) ----
 * Atoms, molecules and rotational groups 0-based.  Bonds: atom ia has nbond[ia] entries
 * iabond/tbond[ia*ldb + j] (partner atom, length in molecule box units; a negative length
 * is an angle constraint).  imageatom[ia] is -1 when ia has none.  Rotational groups of
 * molecule im: nrf[im]..nrl[im] (empty when nrl < nrf); group g rotates labrot[rotf[g]..rotl[g]]
 * about the axis labaxi[2g] -> labaxi[2g+1]; isdih[g] marks a DIH group with ndihedralang[g]
 * dihedrals (atom pairs labrotdih[(g*ldd + kd)*2 + 0/1], angles dihedralang[g*ldd + kd],
 * radians).  xav/xav1 [npairs*ld]: the reference potential the intramolecular non-bonded
 * pairs use (potential_par.xav/xav1).  tempang/tempdih are set by ora_molfit_full from the
 * weights in ora_mcfull_t. */
typedef struct {
    const int *qtype, *imageatom, *mref, *nbond, *iabond;
    const float *tbond;
    int ldb;
    const float *amassr, *amassa;
    const int *ngrot, *nrf, *nrl, *labaxi, *rotf, *rotl, *labrot, *isdih, *ndihedralang, *labrotdih;
    const float *dihedralang;
    int ldd;
    const float *xav, *xav1;
    float dmaxmol, tempb, simt, tempang, tempdih, refine_intra_fac;
    int refine_intra;
} ora_intra_t;

float ora_en_atom_intra(const ora_box_t *b, const ora_tables_t *t, const ora_intra_t *in, int ia, int ja,
                        float *sumevib, float *sumeang);
float ora_en_dihedral(const float *co1, const float *co2, const float *cd, float dihedralang);
float ora_en_mole_dihedral(const ora_box_t *b, const ora_intra_t *in, int im);
float ora_en_mole_intra(const ora_box_t *b, const ora_tables_t *t, const ora_intra_t *in, int im, float *sumevib,
                        float *sumeang, float *sumedih);

/* the complete molfit cycle (molfit.py:51-541): the shake branch when *nisum >= intra,
 * else whole-molecule moves over the cells with internal rotations every rotfreq cycles;
 * all four step sizes adapted.  Counters [2*nmoltype]: [2*mt] attempts, [2*mt+1]
 * rejections.  Returns the number of Metropolis tests. */
typedef struct {
    float *wholerot, *wholetrans, *internrot, *interntrans;   /* [nmoltype] */
    int *nwhole, *nmovem, *ninter, *nmovei;                    /* [2*nmoltype] */
    int *nstepreject;                                          /* [nmol] */
    float simt, stepr, stepm, stepri, stepmi, rejrate, tempangwt, tempdihwt;
    int ireset;
} ora_mcfull_t;
int ora_molfit_full(const ora_box_t *b, float *xa, float *ya, float *za, const ora_centres_t *c,
                    const ora_tables_t *t, ora_intra_t *in, const ora_cells_t *cl, ora_ledger_t *l,
                    ora_mcfull_t *mc, ora_ran1_t *rng, float *fadds, float *faddbs, int *nisum, int intra,
                    int rotfreq);

/* ---- the surrogate's descriptors and network (ora_surrogate.c) ---- */
void ora_descriptor(int na, const float *xm, const float *ym, const float *zm, const int *tm, int nb,
                    const float *xn, const float *yn, const float *zn, const int *tn, const int *npairref,
                    int ntype, int npairs, float dmax, int ng, float rc, float r0, float dmu, float width,
                    float xscale, double *out);
void ora_mlp_forward(int d, int h, const float *w, int bf16, int nrows, const float *x, int ldx, double *y);
double ora_mlp_grad(int d, int h, const float *w, int bf16, int npair, const float *x, int ldx,
                    const float *target, double *grad);

/* ---- grids (src/epsr26/programs/epsr_standalone_rev3.py:2947-3071, 3319-3399, 5391-5419) ---- */
typedef struct { int npt, npt1, npt2, nptmin, nptmax; } ora_grid_t;
void  ora_inita_grid(float cellst, float rmaxgr, float rminpt, float rmaxpt, int mpz,
                     float *x, ora_grid_t *g);
void  ora_uniform(float *ycoh, float xst, float rho, float rmaxgr, int ndim);
double ora_j1qr(float q, float r);
float ora_pofr(float delr);
void  ora_foqols(int nq, const float *q, int npt1, const float *x, const float *ycoh,
                 float rho, float fwhm, float fwhmq, float lorchqwidth, float lorchqwidthq,
                 float rmaxgr, float *foqols, int ldf);

/* ---- truncation and reference potential (ato_manipulation_rev4.py:589-730;
 *      epsr_standalone_rev3.py:3732-3786, 5675-5872) ---- */
float ora_truncate(float xx, float rminpt, float rmaxpt);
float ora_truncatec(float xx, float rmaxpt);
float ora_chsph(float sig1, float sig2, float r);
void  ora_generate_rp(int npt, int npt2, const float *x, const float *xstar, const float *wfac,
                      const float *wfacc, int npairs, const float *ecor, const float *dcor,
                      const float *qcor, const float *qradius_i, const float *qradius_j,
                      float power, float refpotfac, const float *pcoefex, const float *rminex,
                      float expecf, float expecfpwr,
                      float *epot, float *ecoul, float *epot1, float *ecoul1, int ld);

/* ---- EP bases and tables (epsr_standalone_rev3.py:4924-4939, 5100-5153,
 *      5229-5326, 5875-6023; lngamma_mod.py:15-30) ---- */
float ora_expand(int nterms, int j, const float *expr, const float *coeff, float summax);
float ora_dnext(float d, float dfac, float dstep, float *rootd);
int   ora_calc_pwrexpq(int lpt, int nterms, float dmin, float dstep, float dmax, float sigmab,
                       float rho, const float *xin, float *expr, float rmin);
int   ora_calc_pwrexpr_grad(int lpt, int nterms, float dmin, float dstep, float dmax,
                            float sigmab, float rho, const float *xin, float *expr,
                            float rmin, float fbroad, float rbroad);
/* pcoefp[ip*ldp + ic], ic 1-based coefficient index (1..nterms); returns pressfac */
float ora_generate_ep(int npt, int npt2, const float *x, const float *xstar, const float *wfac,
                      int npairs, const int *irefpcof, int nterms, float *pcoefp, int ldp,
                      float pdmin, float pdstep, float pdmax, float psigma2, float rho,
                      float rbroad, float dmax, float absolute_energy, float erequnit,
                      float ereqstep, const float *ecoul, float *xavs, float *xavs1,
                      float *xavtot, float *xavtot1, int ld, float *energabs_out);

/* ---- fit_expansion + smoop (epsr_standalone_rev3.py:4805-4920; auxiliary_routines.py:744-764) ----
 * yin/yout 1-based [npt+1]; wt/wtold 1-based [nterms+1]; expr flat 1-based. */
void ora_smoop(int nqmin, int nqmax, int nsmoo, float *diff);
void ora_fit_expansion(const float *yin, const float *expr, float *yout, float *wt, float *wtold,
                       float *stepr, float *acceptr, float *rejrat, float *chisqnew, int nmin,
                       int npt, int ncfirst, int nterms, int nits, int nfreq, ora_ran1_t *rng,
                       int *ntries_out);

/* ptgnsq pieces (neutron weights); layouts as the library's epsr_ptgnsq_diff */
void  ora_ptgnsq_diff(int nfile, int nq, int npairs, const float *q, const float *data, int ldq,
                      const float *qwind, const float *cwt, const float *szeros, const int *nqmin,
                      const int *nqmax, float qminptemp, float *qtemp);
float ora_ptgnsq_rfactor(int i0, int i1, const float *fit, const float *qt);

/* the data-minus-simulation block of the epsr program (epsr_standalone_rev3.py:1703-1830)
 * and ftback (:5330-5389).  Datasets 0..nfile-1; Q index 1..nq in rows of ldq; partial ip
 * of soqintra/soqinter at [ip*ldq + iq] (the running averages); ciwt/cwt [nfile*npairs]
 * (neutron datasets); weightx [nx*ldq*npairs] double, [((ix-1)*ldq + iq)*npairs + ip];
 * xfile[id]: 0 neutron, else the 1-based X-ray index; x [npt1+1] the r grid. */
typedef struct {
    int nfile, nq, ldq, npairs;
    const float *q, *datai, *soqintra, *soqinter, *ciwt, *cwt;
    const double *weightx;
    const int *xfile, *nqmind, *nqmaxd, *nqmin, *nqmax;
    const float *szeros, *porodamp, *rshmin;
    float porodq, porodpwr;
    int nsmoop;
    const float *x;
    int npt1;
    float rho;
} ora_dms_t;
void ora_ftback(int nqmin, int nqmax, const float *q, const float *x, int npt1, float rho, float rmin, float *sofq,
                float *grt);
void ora_data_minus_sim(const ora_dms_t *d, float *datat, float *szerostemp, float *dts, float *grt);
void ora_ptgnsq_diff_x(int nfile, int nq, int npairs, const float *q, const float *data, int ldq, const float *qwind,
                       const float *cwt, const double *weightx, const int *xfile, const float *szeros,
                       const int *nqmin, const int *nqmax, float qminptemp, float *qtemp);

#ifdef __cplusplus
}
#endif
#endif
