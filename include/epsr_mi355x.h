/*
 * epsr_mi355x.h -- C ABI of libepsr_mi355x.so, the MI355X (gfx950) implementation of
 * the EPSR refinement hot loop.
 *
 * Conventions
 *  - Every array argument documented "dev" is a device pointer (HBM), e.g. the
 *    data_ptr() of a torch tensor on the GPU; "host" arrays are ordinary memory.
 *    The library borrows caller buffers for the duration of a call.
 *  - Atoms, molecules, atom types, pair types and datasets are 0-based.
 *    Radial-grid and Q-grid indices keep the reference's 1-based numbering: grid
 *    arrays have an unused element 0 and table[ic*ld + iz] is the reference's
 *    table(iz, ic+1).
 *  - The ran1 state is 35 int32 in the reference's order (mcall, idum, iy, iv(1:32)),
 *    host memory, in/out: after a call it equals the state the reference leaves.
 *  - Every entry point returns 0 on success and a negative code on error;
 *    epsr_last_error() then describes it.  Calls are synchronous with respect to
 *    the host unless documented otherwise (they run on the context's stream).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *  see the comment above each entry point.
 */
#ifndef EPSR_MI355X_H
#define EPSR_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EPSR_OK 0
#define EPSR_ERR_ARG (-1)
#define EPSR_ERR_HIP (-2)
#define EPSR_ERR_UNSUPPORTED (-3)
#define EPSR_ERR_CAPACITY (-4)

typedef struct epsr_ctx epsr_ctx;
typedef struct epsr_box epsr_box;
typedef struct epsr_ledger epsr_ledger;

/* ---- context ---- */
int epsr_ctx_create(int device, epsr_ctx **out);
void epsr_ctx_destroy(epsr_ctx *ctx);
const char *epsr_last_error(const epsr_ctx *ctx);
/* run subsequent calls on this hipStream_t (NULL = the context's own stream) */
int epsr_ctx_set_stream(epsr_ctx *ctx, void *hip_stream);
int epsr_ctx_synchronize(epsr_ctx *ctx);
/* GPU time (ms) of the last launch of the named kernel family, measured with HIP
 * events on the launch stream; families: "atogen_inter", "gensq", "stog",
 * "mc_pair", "mc_resolve", "fmolec1_pair".  Returns -1 if never launched. */
double epsr_ctx_last_kernel_ms(const epsr_ctx *ctx, const char *family);
/* the summed GPU time (ms) and the number (*count) of the named family's launches since
 * the previous call of this function (those read by epsr_ctx_last_kernel_ms included);
 * -1 with *count = 0 when there were none.  Several launches of a family in one call
 * (epsr_molfit_cycles) are each counted. */
double epsr_ctx_kernel_ms_sum(epsr_ctx *ctx, const char *family, long long *count);
int epsr_ctx_enable_timing(epsr_ctx *ctx, int on);

/* ---- host ran1 stream: n deviates of auxiliary_routines.ran1
 * (src/epsr26/library/auxiliary_routines.py:324-385) from the 35-int32 state (in/out).
 * Host-only; used by the origin selection and the MC streams. */
int epsr_ran1_stream(int32_t *ran1_state, int n, float *out);

/* ---- box topology (static per run): copied to HBM once ----
 * Mirrors the molecule/atom COMMON data the path reads:
 * src/epsr26/common/molecom.py:54,138-151 and clustercom.py:30-51. */
typedef struct {
    int nmol, natom, ntype, npairs, nmoltype;
    const int *itype;        /* host [natom]            molecom.itype */
    const int *naf, *nal;    /* host [nmol]             clustercom.naf/nal (inclusive) */
    const int *mtype;        /* host [nmol]             clustercom.mtype */
    const int *npairref;     /* host [ntype*ntype]      molecom.npairref */
    const int *n1extr, *n2extr; /* host [npairs]        molecom.n1extr/n2extr */
    const float *cfrac;      /* host [ntype]            molecom.cfrac */
    float natomr, sizefactor, dmax, cellst;
} epsr_box_desc_t;

int epsr_box_create(epsr_ctx *ctx, const epsr_box_desc_t *desc, epsr_box **out);

/* Launch shapes of a box's kernels (no counterpart in the reference: the reference has
 * no launches), set by the autotuner per box size: EPSR_LAUNCH_ATOGEN_WORKGROUPS, the
 * resident workgroups of atogen's persistent pair kernel (each flushes one histogram;
 * default 768), and EPSR_LAUNCH_PAIR_BLOCK, the workgroup size of mc_pairs and fmolec1's pair
 * kernel (256, 512 or 1024; default 1024).  Results do not depend on them.
 * EPSR_LAUNCH_RESOLVE_PASSES caps the passes of molfit's one-workgroup resolution (default:
 * no cap, nmol + 1 passes, which always suffice; a cap also leaves the grid-wide passes at
 * one, so that the finisher has work); a cycle that ends with moves undecided returns
 * EPSR_ERR_HIP.  The cap (>= 0) exists to test that path: results do not depend on it when
 * it is large enough. */
#define EPSR_LAUNCH_ATOGEN_WORKGROUPS 1
#define EPSR_LAUNCH_PAIR_BLOCK 2
#define EPSR_LAUNCH_RESOLVE_PASSES 3
int epsr_box_set_launch(epsr_box *box, int knob, int value);

/* HIP graphs of a context's stream (no counterpart in the reference): work issued on
 * the stream between epsr_graph_begin and epsr_graph_end is captured, not run, and
 * epsr_graph_launch replays it.  Entry points that read results back to the host (MC
 * cycles, atogen's origin selection, fmolec1) cannot be captured; gensq and stog can,
 * after one eager call has sized the context's scratch. */
typedef struct epsr_graph epsr_graph;
int epsr_graph_begin(epsr_ctx *ctx);
int epsr_graph_end(epsr_ctx *ctx, epsr_graph **out);
int epsr_graph_launch(epsr_ctx *ctx, const epsr_graph *graph);
void epsr_graph_destroy(epsr_graph *graph);
void epsr_box_destroy(epsr_box *box);

/* Mutable coordinates, dev arrays owned by the caller: atom offsets from the
 * molecule centre (molecom.xa/ya/za [natom]) and centres (clustercom.xadd/yadd/zadd
 * [nmol]), box-fraction units. */
typedef struct {
    float *xa, *ya, *za;
    float *xadd, *yadd, *zadd;
} epsr_coords_t;

/* ---- atogen: replaces ato_manipulation_rev4.atogen(wav, ngrsamples)
 * (src/epsr26/library/ato_manipulation_rev4.py:3640-3899; Fortran
 * fortran_src/ato_manipulation_rev4.f:3699).
 * wav dev [2*npairs*ldw]: wav[(k*npairs+ic)*ldw+iz] = reference wav(k+1, iz, ic+1).
 * ycoh dev [npt1+1]; irefpcof dev [npairs]; rminex dev [*]; potex/potexmax dev [npairs].
 * counts (optional, dev uint32 [2*npairs*ldw]) receives the raw histogram counts.
 * ngrsamples host in/out.  */
int epsr_atogen(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, int npt1,
                int ldw, const float *ycoh, const int *irefpcof, const float *rminex,
                int *ngrsamples, int32_t *ran1_state, float *wav, float *potex, float *potexmax,
                uint32_t *counts);

/* ---- atogen's binning, checked.  No counterpart in the reference: a check, for callers
 * and tests, that the pair kernel's bin and close-contact decisions (hardware square root
 * with exact per-launch thresholds near bin edges) equal the reference's correctly rounded
 * NINT(sqrtf(dsq) / stepcl) and sqrtf(dsq) * dmax < rabs for every float32 squared distance
 * whose bit pattern lies in [lo_bits, hi_bits) (hi_bits at most that of +inf).
 * counts (host [3]) receives the disagreements, the values checked and those near a bin
 * edge (the threshold branch).  Synchronises the stream. */
int epsr_atogen_bin_check(epsr_ctx *ctx, float stepcl, int npt1, float dmax, float rabs,
                          uint32_t lo_bits, uint32_t hi_bits, unsigned long long *counts);

/* ---- gensq: replaces epsr_standalone_rev3.gensq(nq, q, sofq, wav)
 * (src/epsr26/programs/epsr_standalone_rev3.py:4730-4802; Fortran :4737).
 * foqols dev [(nq+1)*ldf]: foqols[i*ldf+j] = reference foqols(j, i).
 * natomsphere dev [npairs] (value for each pair's second type; NULL = 1).
 * sofq dev [2*npairs*ldq]: sofq[(k*npairs+ic)*ldq+i] = reference sofq(k+1, i, ic+1).
 * mode 0: each output is the reference's sequential float32 sum (bit-identical);
 * mode 1: FP32 MFMA GEMM [nq x npt1] . [npt1 x 2*npairs] (within 1e-5 of sum |terms|). */
int epsr_gensq(epsr_ctx *ctx, int nq, int npairs, int npt1, int ldw, const float *wav,
               const float *ycoh, const float *foqols, int ldf, const float *natomsphere,
               float *sofq, int ldq, int mode);

/* ---- gensq without allocation.  No counterpart in the reference: epsr_gensq mode 1 (the
 * FP32 MFMA GEMM, same kernels, same results) for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation.  Its scratch (column sums and 16 partial tiles, 4*npairs rounded up to 256
 * bytes + 64 * ceil(2*npairs/16) * 16 * nq bytes) lives in a fixed 8 MiB device array of the
 * library, one per GPU, held until the process ends; larger nq x npairs are refused with
 * EPSR_ERR_ARG.  Calls that overlap on one GPU share that array and race. */
int epsr_gensq_noalloc(epsr_ctx *ctx, int nq, int npairs, int npt1, int ldw, const float *wav, const float *ycoh,
                       const float *foqols, int ldf, const float *natomsphere, float *sofq, int ldq);

/* ---- gensq of many boxes on one grid (replicas of one system).  No counterpart in the
 * reference: epsr_gensq mode 1 of every box as one GEMM, the boxes' columns side by side
 * (the transform matrix foqols is read once); each box's sofq equals its epsr_gensq mode 1
 * result bit for bit.  nq, npt1, ycoh, foqols, ldf are shared; boxes is a host array. */
typedef struct {
    int npairs, ldw, ldq;
    const float *wav;          /* dev [2*npairs*ldw] */
    const float *natomsphere;  /* dev [npairs] or NULL */
    float *sofq;               /* dev [2*npairs*ldq] */
} epsr_gensq_box_t;
int epsr_gensq_batch(epsr_ctx *ctx, int nq, int npt1, const float *ycoh, const float *foqols, int ldf, int nbox,
                     const epsr_gensq_box_t *boxes);

/* ---- stog: replaces ato_manipulation_rev4.stog(rho, q, sofq, r, gofr, nqmin, nqmax,
 * nr, npairs, nqdim, nrdim, npdim, xwt) (ato_manipulation_rev4.py:4066-4111; Fortran
 * :4154) including lorch (:4114-4124).
 * q dev [nqdim+1], r dev [nr+1], sofq dev [ncols*ldq] (sofq[in*ldq+i]), gofr dev
 * [ncols*ldg] (gofr[in*ldg+j]); nqmin/nqmax/xwt dev [ncols].
 * mode 0: per-output sequential sum in the reference's order with correctly
 * rounded sinf (matches a correctly rounded libm bit for bit);
 * mode 1: FP32 MFMA GEMM [nr x nq] . [nq x ncols] with the kernel q sin(qr)
 * generated on the fly and the Lorch window folded into the right operand. */
int epsr_stog(epsr_ctx *ctx, float rho, const float *q, const float *sofq, int ldq,
              const float *r, float *gofr, int ldg, const int *nqmin, const int *nqmax,
              int nr, int ncols, const float *xwt, int mode);

/* ---- stog of many boxes on one grid (replicas of one system).  No counterpart in the
 * reference: epsr_stog mode 1 of every box as one GEMM (q sin(qr) generated once for all
 * columns); each box's gofr equals its epsr_stog mode 1 result bit for bit.  rho, q, ldq,
 * r and nr are shared; boxes is a host array. */
typedef struct {
    const float *sofq;            /* dev [ncols*ldq] */
    float *gofr;                  /* dev [ncols*ldg] */
    const int *nqmin, *nqmax;     /* dev [ncols] */
    const float *xwt;             /* dev [ncols] */
    int ncols, ldg;
} epsr_stog_box_t;
int epsr_stog_batch(epsr_ctx *ctx, float rho, const float *q, int ldq, const float *r, int nr, int nbox,
                    const epsr_stog_box_t *boxes);

/* ---- potential tables, dev arrays [npairs*ld] (table[ic*ld+iz] = reference
 * table(iz, ic+1)) and xstar dev [ld]: potential_par.xavs/xavs1 (empirical potential
 * and gradient), potential_par.ecoul/ecoul1 (reference potential and gradient),
 * epsr16.xstar; npt2 is the last usable index (src/epsr26/common/potential_par.py:22-86). */
typedef struct {
    int ld, npt2;
    const float *xstar;
    const float *xavs, *xavs1, *ecoul, *ecoul1;
} epsr_tables_t;

/* ---- cell lists + pair-energy ledger, device resident ----
 * epsr_ledger_create allocates the object; epsr_cells_init replaces
 * cellneighbourlist.initialise_cellneighbourlist(nmolcellin, xadd, yadd, zadd, nmol, rmax)
 * (src/epsr26/library/cellneighbourlist.py:23-96, makecellneighbourlist :102-194) for a
 * cubic box, reading the centres from coords (dev).  rmax in box fractions. */
int epsr_ledger_create(epsr_ctx *ctx, const epsr_box *box, epsr_ledger **out);
void epsr_ledger_destroy(epsr_ledger *l);
int epsr_cells_init(epsr_ctx *ctx, const epsr_box *box, epsr_ledger *l, const epsr_coords_t *coords,
                    int nmolcellin, float rmax);

/* ---- fmolec1: replaces ato_manipulation_rev4.fmolec1(ierr, num_threads, ftest, ftestb,
 * ftethered, fcoordconstraints) (src/epsr26/library/ato_manipulation_rev4.py:2374-2528;
 * Fortran :2344) with the ledger routines initialise_ergpair/set_ergpair
 * (src/epsr26/library/ergpairreference.py:63-126): builds the ledger from the cells of
 * the last epsr_cells_init.  eintra dev [nmol]: intramolecular energy of each
 * molecule (molfit.en_mole_intra).  Outputs (host): ftest and ftestb per molecule
 * accumulated in FP64 (ftest64, ftestb64) and rounded to float (ftest, ftestb);
 * nnbrefmax and the number of stored entries (nergpair). */
int epsr_fmolec1(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                 epsr_ledger *l, const float *eintra, double *ftest64, double *ftestb64, float *ftest,
                 float *ftestb, int *nnbrefmax, int *nergpair);

/* copy the ledger to host in the reference's layout: nergpairm [nmol],
 * molergpair/ergpairref [nmol*ldl] (row im, slot k at im*ldl+k), ergpair [nergpair];
 * also the chain order of the MC cycle (molecules by cell) in order [nmol]. */
int epsr_ledger_export(epsr_ctx *ctx, const epsr_ledger *l, int ldl, int *nergpairm, int *molergpair,
                       int *ergpairref, float *ergpair, int *order);

/* ---- en_change: replaces molfit.en_change(im) (src/epsr26/library/molfit.py:805-919)
 * after make_neighbour_list_from_mols(im) (ergpairreference.py:165-188): energy change of
 * molecule im at its current coordinates against the ledger.  ergnew (dev, optional)
 * receives the ledger-slot energies (slot 0 = stored self energy). */
int epsr_en_change(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                   const epsr_ledger *l, int im, float *ergnew, float *fdiff);

/* ---- en_mole_pair: replaces molfit.en_mole_pair(im, jm) (molfit.py:1527-1578); npair
 * pairs (dev im[npair], jm[npair]) evaluated at once into energy (dev [npair]). */
int epsr_en_mole_pair(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                      int npair, const int *im, const int *jm, float *energy);

/* ---- sumsq: replaces epsr_standalone_rev3.sumsq(nq, qwind)
 * (src/epsr26/programs/epsr_standalone_rev3.py:3642-3685; Fortran epsr_standalone_rev3.f:sumsq):
 * the running average and error bar of every intra (k = 0) and inter (k = 1) partial S(Q),
 * Q points 1..nq.  qwind dev [2*npairs*ldq] (gensq's sofq, [(k*npairs + ic)*ldq + j]);
 * soq/sq2 dev, same layout (epsr16.soqintra/soqinter and sq2intra/sq2inter), updated in
 * place; fc1, fc2: common /sumvalues/ (epsr_sum_factors in pipeline.py). */
int epsr_sumsq(epsr_ctx *ctx, int npairs, int nq, int ldq, const float *qwind, double fc1, double fc2,
               float *soq, float *sq2);

/* ---- sumgr: replaces epsr_standalone_rev3.sumgr(nz, wav) (:3578-3637): the same for the
 * partial g(r), r points 2..nz, with add1 = wav / ycoh (0 where ycoh <= 0); wav dev
 * [2*npairs*ldw], ycoh dev [ldw]; gor/gr2 dev [2*npairs*ldw] (gorintra/gorinter,
 * gr2intra/gr2inter); entsig dev [npairs*ldw] receives each pair's zero limit
 * (epsr16.entsig). */
int epsr_sumgr(epsr_ctx *ctx, int npairs, int nz, int ldw, const float *wav, const float *ycoh, double fc1,
               double fc2, float *gor, float *gr2, float *entsig);

/* ---- invert_cwt: replaces epsr_standalone_rev3.invert_cwt(cwt, cwtt, carray, carrayold,
 * feedback, efacm, nfile, npairs, mfile, mpairs) (src/epsr26/programs/epsr_standalone_rev3.py:
 * 5423-5672; Fortran epsr_standalone_rev3.f:invert_cwt): the Monte Carlo transpose of the
 * weights matrix, host arrays (real*8) in and out, 0-based and row-major: cwt [nfile rows,
 * ldc] (file i, pair j; scaled in place by feedback * efacm(i) as the reference does),
 * cwtt [npairs rows, ldt >= npairs + nfile] (pair j, column i), carray [npairs + nfile],
 * carrayold [(npairs + nfile) rows, ldo] (both optional); ran1_state in/out.  Host work
 * (a serial walk over at most 51 x 51 elements); itimes_out: step-size rounds taken. */
int epsr_invert_cwt(epsr_ctx *ctx, double *cwt, int ldc, double *cwtt, int ldt, double *carray, double *carrayold,
                    int ldo, float feedback, const float *efacm, int nfile, int npairs, int32_t *ran1_state,
                    int *itimes_out);

/* ---- press: replaces ato_manipulation_rev4.press(energy, pressc, energf, pressf, energabs,
 * energthreeb, prthreebout) (src/epsr26/library/ato_manipulation_rev4.py:3901-4063; Fortran
 * ato_manipulation_rev4.f:press) without three-body terms or coordination constraints:
 * the whole-box intermolecular energy and virial pressure over the ledger's partner lists
 * (make_neighbour_list_from_mols, ergpairreference.py:165-188) at the current coordinates.
 * Outputs (host): energy = the reference's float32 sum / nmol (kJ/mol), pressc (kbar) and
 * presst (the double virial sum).  npt1: the last bin used (2 <= npt1 < ld); simt =
 * 0.008314 T; volume in A^3.  slote/slotw (host, optional, [nmol*ldl], ldl >= nnbrefmax):
 * each ledger slot's energy and virial in the reference's layout, 0 where the partner is
 * not later in the numbering.  abs_energy_ep and the per-dataset energf/pressf
 * normalisation are host arithmetic (press.py). */
int epsr_press(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
               const epsr_ledger *l, int npt1, float simt, float volume, float *energy, float *pressc,
               double *presst, int ldl, float *slote, float *slotw);

/* ---- molfit: replaces molfit.molfit(fadds, faddbs, nisum, intra, rotfreq, ireset)
 * (src/epsr26/library/molfit.py:51-541) for the whole-molecule branch (nisum < intra,
 * no internal rotational groups, no tethered molecules): one MC cycle over the cells,
 * rotatm + movem (:671-759) trial moves, en_change, Metropolis, ledger write-back and the
 * step-size adaptation.  Host arrays: wholerot/wholetrans [nmoltype] (in/out),
 * nwhole/nmovem [2*nmoltype] (out: attempts, rejections), nstepreject dev [nmol]
 * (in/out).  fadds in/out.  Trial moves are evaluated in parallel for every molecule
 * and both states of each earlier neighbour; the serial Metropolis chain then resolves
 * them in the reference's order (bit-identical to a correctly rounded libm). */
typedef struct {
    float simt, stepr, stepm, rejrate;
    int ireset;
} epsr_mc_params_t;

/* per-cycle statistics: moves attempted and accepted, moves decided by forming the
 * exact float32 sum, movem retries, push-row entries the resolver read, the resolver's
 * algorithmic bytes (thresholds, in-block deltas, row offsets, pushed rows, acceptance
 * bits) and the GPU time of the pair and resolver kernels (with timing enabled) */
typedef struct {
    int moves, accepted, exact_resolutions, retries;
    double ms_pairs, ms_resolve;
    long long pushed_entries;
    double resolver_bytes;
    int stage1_rejected;   /* delayed acceptance: moves the surrogate stage rejected */
    int resolver_passes;   /* passes of the parallel resolution */
    int cycles;                /* cycles the call ran (the fields above are the last one's) */
    long long accepted_total;  /* moves accepted over those cycles */
} epsr_mc_stats_t;

int epsr_molfit(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                epsr_ledger *l, const epsr_mc_params_t *p, float *wholerot, float *wholetrans, int *nwhole,
                int *nmovem, int *nstepreject, int32_t *ran1_state, float *fadds, epsr_mc_stats_t *stats);

/* ncycles rigid-molecule molfit cycles in a row: the program's MC loop over itimes
 * (src/epsr26/programs/epsr_standalone_rev3.py:1313-1319) for boxes without intramolecular
 * moves.  Every output is what ncycles calls of epsr_molfit leave (coordinates, ledger,
 * step sizes, the last cycle's counters and stats, nstepreject, ran1 state, fadds).  The
 * cycles run back to back on the GPU: the step sizes adapt on the device between cycles, a
 * host thread draws each cycle's ran1 values while earlier cycles run, and the host waits
 * once.  A cycle in which movem must double a step (or that finds the tables changed)
 * commits nothing, nor do the cycles queued after it; the call then runs it alone and goes
 * on.  Up to 64 cycles are queued at a time. */
int epsr_molfit_cycles(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                       epsr_ledger *l, const epsr_mc_params_t *p, int ncycles, float *wholerot, float *wholetrans,
                       int *nwhole, int *nmovem, int *nstepreject, int32_t *ran1_state, float *fadds,
                       epsr_mc_stats_t *stats);

/* ==== intramolecular energies and moves ========================================
 * The molecule-internal data of molecom/clustercom/molecom_mod that en_mole_intra,
 * movei, rotati and centreofmassi read (src/epsr26/common/molecom.py, molecom_mod.py,
 * clustercom.py: nbond, iabond, tbond, imageatom, qtype, amassr, amassa, nrf/nrl, labaxi,
 * rotf/rotl, labrot, rotgrp, ndihedralang, labrotdih, dihedralang), host arrays, 0-based:
 * atom ia has nbond[ia] bonds iabond/tbond[ia*ldb + j] (partner atom; length in the
 * molecule's box units, negative for an angle constraint); imageatom[ia] = -1 for none;
 * molecule im owns rotational groups nrf[im]..nrl[im] (none when nrl < nrf); group g turns
 * labrot[rotf[g]..rotl[g]] about the axis labaxi[2g] -> labaxi[2g+1]; isdih[g] marks a DIH
 * group with ndihedralang[g] dihedrals: atom pairs labrotdih[(g*ldd + k)*2 + 0/1], angles
 * dihedralang[g*ldd + k] (radians).  dmaxmol = dmax * sizefactor; tempb the bond weight;
 * refine_intra / refine_intra_fac the EP weighting of bonded pairs. */
typedef struct {
    int ldb;
    const int *qtype, *imageatom, *nbond, *iabond;
    const float *tbond, *amassr, *amassa;
    int ngroup, nlabrot, ldd;
    const int *nrf, *nrl, *labaxi, *rotf, *rotl, *labrot, *isdih, *ndihedralang, *labrotdih;
    const float *dihedralang;
    float dmaxmol, tempb, refine_intra_fac;
    int refine_intra;
} epsr_intra_desc_t;
/* no counterpart call in the reference: uploads the topology above (once per box) */
int epsr_box_set_intra(epsr_box *box, const epsr_intra_desc_t *desc);

/* ---- en_mole_intra: replaces molfit.en_mole_intra(im, ...) (src/epsr26/library/molfit.py:
 * 1327-1349, en_atom_intra :922-1047, en_mole_dihedral :1353-1407, en_dihedral :1495-1524)
 * for every molecule at once: eintra dev [nmol] at the current coordinates.  xav/xav1 dev
 * [npairs*tables->ld] (potential_par.xav/xav1: the intramolecular reference potential,
 * zero unless coul_intra; NULL = zero); tempang and tempdih are molecom.tempang and
 * molecom.tempdih (fmolec1 uses tempangreq and tempdihreq * tempb). */
int epsr_en_mole_intra(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                       const float *xav, const float *xav1, float simt, float tempang, float tempdih, float *eintra);

/* ---- molfit, every branch: replaces molfit.molfit(fadds, faddbs, nisum, intra, rotfreq,
 * ireset) (src/epsr26/library/molfit.py:51-541, movei :545-565, rotati :589-655,
 * en_change_intra :763-784, centreofmassi ato_manipulation_rev4.py:442-475) for boxes
 * without tethered molecules, three-body terms or coordination constraints.
 * When *nisum >= intra: the shake cycle (movei of every molecule judged on its
 * intramolecular energy, then internal rotations judged on intra + inter energy);
 * otherwise whole-molecule moves over the cells, preceded, when stepri > 0 and nisum is a
 * multiple of rotfreq, by internal rotations of every molecule with rotational groups.
 * Cycles without intramolecular moves take epsr_molfit's parallel resolution; the others
 * walk the chain in order in one workgroup.  Step sizes host [nmoltype] in/out, counters
 * host [2*nmoltype] out ([2*mt] attempts, [2*mt+1] rejections), nstepreject dev [nmol],
 * fadds/faddbs/nisum host in/out, ran1 state host in/out.  Needs epsr_box_set_intra
 * when the cycle moves atoms inside molecules. */
typedef struct {
    float simt, stepr, stepm, stepri, stepmi, rejrate, tempangwt, tempdihwt;
    int ireset, intra, rotfreq;
} epsr_mc_full_params_t;
typedef struct {
    float *wholerot, *wholetrans, *internrot, *interntrans;
    int *nwhole, *nmovem, *ninter, *nmovei;
} epsr_mc_steps_t;
int epsr_molfit_full(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                     const float *xav, const float *xav1, epsr_ledger *l, const epsr_mc_full_params_t *p,
                     epsr_mc_steps_t *steps, int *nstepreject, int32_t *ran1_state, float *fadds, float *faddbs,
                     int *nisum, epsr_mc_stats_t *stats);

/* ==== learned Delta-E surrogate and delayed-acceptance MC =====================
 * No counterpart in the reference: these entry points add a two-stage Metropolis step
 * beside molfit.  The surrogate is a site-energy network f over radial descriptors of a
 * molecule: for the move of the molecule at chain position k, dEs = f(d_k(trial)) -
 * f(d_k(current)), where d_k sums, over the molecule's ledger partners that come LATER in
 * the chain (still at their positions at the start of the cycle when k moves), Gaussians
 * of every atom-pair distance below rc into the channel of the pair's types.  Since dEs
 * depends only on the current state and the trial and changes sign when the two are
 * swapped, the two-stage rule
 *   stage 1: accept with min(1, exp(-dEs/simt))      (uniform u1 = splitmix64(seed, molecule))
 *   stage 2: accept with min(1, exp(-(dE - dEs)/simt))  (the ran1 number of plain molfit)
 * keeps detailed balance with respect to the exact energy (Christen and Fox, 2005); dE is
 * en_change's fdiff, formed exactly as in plain molfit. */

/* descriptor definition: Gaussian g of channel c = npairref(type_a, type_b) adds
 * exp(-((r - (r0 + g*dmu)) / width)^2) * 0.5*(cos(pi r / rc) + 1) for r < rc (Angstrom),
 * summed in 2^-32 fixed point (order-independent, deterministic), times xscale */
typedef struct {
    int ng;             /* Gaussians per channel, 1..32 */
    float rc, r0, dmu, width, xscale;
} epsr_desc_params_t;

/* descriptors of every chain position k for the current coordinates (row 2k) and the
 * trial coordinates of its molecule (row 2k+1), over its later ledger partners at their
 * current coordinates.  x dev bf16 [2*nmol*ldx], ldx >= npairs*ng, multiple of 32 (zero
 * padded); xf dev fp32 [2*nmol*ldx] or NULL. */
int epsr_descriptors(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_coords_t *trial,
                     const epsr_ledger *l, const epsr_desc_params_t *dp, int ldx, uint16_t *x, float *xf);

/* the network: y = w3 . silu(W2 silu(W1 x + b1) + b2) + b3, W1 [h][d], W2 [h][h];
 * parameters in one fp32 vector w = [W1 | b1 | W2 | b2 | w3 | b3] (master copy), wb the
 * bf16 copy of W1 then W2 the forward pass reads (epsr_mlp_sync refreshes it), m and v the
 * Adam moments (NULL when not training).  d a multiple of 32, h = 128. */
typedef struct {
    int d, h;
    float *w;
    uint16_t *wb;
    float *m, *v;
    int step;
} epsr_mlp_t;
long long epsr_mlp_nparams(int d, int h);
int epsr_mlp_sync(epsr_ctx *ctx, const epsr_mlp_t *mlp);
/* y dev [nrows] = f(x row), x dev bf16 [nrows*ldx]; BF16 MFMA, FP32 accumulation */
int epsr_mlp_forward(epsr_ctx *ctx, const epsr_mlp_t *mlp, int nrows, const uint16_t *x, int ldx, float *y);
/* loss = mean_i (f(x[2i+1]) - f(x[2i]) - t_i)^2 over npair pairs, t_i = target[i] clamped
 * to [-tclip, tclip] when tclip > 0, and its gradient (dev [nparams], the layout of w) with
 * respect to w through the bf16 forward pass; loss host out (nullable) */
int epsr_mlp_grad(epsr_ctx *ctx, const epsr_mlp_t *mlp, int npair, const uint16_t *x, int ldx, const float *target,
                  float tclip, float *grad, float *loss);
/* one Adam step on w (m, v, step updated) and the bf16 copy refreshed */
int epsr_mlp_adam(epsr_ctx *ctx, epsr_mlp_t *mlp, const float *grad, float lr, float beta1, float beta2, float eps);

/* delayed acceptance: the network, the descriptor definition, the stage-1 seed, and
 * des (dev [nmol] by molecule, or NULL) that replaces the network's dEs when given */
typedef struct {
    const epsr_mlp_t *mlp;
    epsr_desc_params_t desc;
    unsigned long long seed;
    const float *des;
} epsr_da_t;

/* what a cycle records by chain position (each pointer nullable): the descriptors (x,
 * rows 2k and 2k+1, as epsr_descriptors), the exact energy change over the later
 * partners (target, the quantity dEs models), dEs and the stage-1 outcome (pass) */
typedef struct {
    const epsr_desc_params_t *desc;   /* used when the cycle has no delayed acceptance */
    int ldx;
    uint16_t *x;
    float *target;
    float *des;
    int *pass;
} epsr_mc_record_t;

/* molfit with delayed acceptance (da non-NULL) and/or recording (rec non-NULL); with
 * both NULL it is epsr_molfit. */
int epsr_molfit_da(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *coords, const epsr_tables_t *tables,
                   epsr_ledger *l, const epsr_mc_params_t *p, float *wholerot, float *wholetrans, int *nwhole,
                   int *nmovem, int *nstepreject, int32_t *ran1_state, float *fadds, const epsr_da_t *da,
                   const epsr_mc_record_t *rec, epsr_mc_stats_t *stats);

/* ---- truncation functions of initpt (src/epsr26/programs/epsr_standalone_rev3.py:3755-3767):
 * wfac = truncate(x, rminpt, rmaxpt), wfacc = truncatec(x, rmaxpt)
 * (src/epsr26/library/ato_manipulation_rev4.py:693-727); x, wfac, wfacc dev [npt+1]. */
int epsr_truncation(epsr_ctx *ctx, int npt, const float *x, float rminpt, float rmaxpt, float *wfac, float *wfacc);

/* ---- generate_rp: replaces epsr_standalone_rev3.generate_rp() (:5675-5872) for n-6
 * Lennard-Jones + Coulomb between charge spheres (chsph, ato_manipulation_rev4.py:589-653)
 * + the exponential repulsive core; per-pair parameters dev [npairs]
 * (molecom.ecor/dcor/qcor, qradius of the two types, potential_par.pcoefex/rminex). */
typedef struct {
    int npairs;
    const float *ecor, *dcor, *qcor;
    const float *qradius_i, *qradius_j;   /* NULL: point charges */
    const float *pcoefex, *rminex;        /* NULL: no repulsive core */
    float power, refpotfac, expecf, expecfpwr;
} epsr_rp_params_t;
int epsr_generate_rp(epsr_ctx *ctx, int npt, int npt2, const float *x, const float *xstar, const float *wfac,
                     const float *wfacc, const epsr_rp_params_t *p, float *epot, float *ecoul, float *epot1,
                     float *ecoul1, int ld);

/* ---- generate_ep: replaces epsr_standalone_rev3.generate_ep(changepotfac=False)
 * (:5875-6023) with calc_pwrexpr_grad (:5229-5310), expand (:4924-4939), abs_energy_ep
 * (ato_manipulation_rev4.py:4127-4157) and multiply_pot_coeffs (:6493-6510) for the
 * Poisson basis.  pcoefp host [nref*ldp] (row irefp, coefficient m at irefp*ldp+m,
 * m = 1..nterms), multiplied in place by the returned pressfac; irefpcof host [npairs];
 * x, xstar, wfac, ecoul dev; xavs/xavs1/xavtot/xavtot1 dev [npairs*ld] in/out. */
typedef struct {
    float pdmin, pdstep, pdmax, psigma2, rho, rbroad, dmax, absolute_energy, erequnit, ereqstep;
} epsr_ep_params_t;
int epsr_generate_ep(epsr_ctx *ctx, int npt, int npt2, const float *x, const float *xstar, const float *wfac,
                     int npairs, const int *irefpcof, int nterms, float *pcoefp, int ldp, const epsr_ep_params_t *p,
                     const float *ecoul, float *xavs, float *xavs1, float *xavtot, float *xavtot1, int ld,
                     float *pressfac, float *energabs);

/* ---- calc_pwrexpq: the Q-space Poisson basis of the EP fit
 * (epsr_standalone_rev3.py:5100-5153); q dev [lpt+1], expr dev [lpt*nterms+1] with
 * expr[(i-1)*nterms+m] = basis m at Q point i; ncfirst host out. */
int epsr_calc_pwrexpq(epsr_ctx *ctx, int lpt, int nterms, float dmin, float dstep, float dmax, float sigmab,
                      float rho, const float *q, float *expr, float rmin, int *ncfirst);

/* ---- foqols: the r -> Q transform matrix built by inita
 * (src/epsr26/programs/epsr_standalone_rev3.py:3025-3071, j1qr :3319-3329, pofr :3389-3399).
 * q dev [nq+1], x dev [npt1+1], r2 dev [npt1+1] (bin edges of the serial r1/r2
 * recurrence), ycoh dev [npt1+1]; foqols dev [(nq+1)*ldf].  Double-precision j1qr
 * differences and correctly rounded expf/sinf/cosf. */
int epsr_foqols(epsr_ctx *ctx, int nq, const float *q, int npt1, const float *x, const float *r2,
                const float *ycoh, float rho, float fwhm, float fwhmq, float lorchqwidth,
                float lorchqwidthq, float rmaxgr, float *foqols, int ldf);

/* ---- expand: replaces epsr_standalone_rev3.expand(nterms, j, expr, coeff, summax)
 * (:4924-4939) for every j = 1..npt at once; expr dev flat 1-based (row j at
 * (j-1)*nterms), coeff dev [nterms+1], out dev [npt+1]. */
int epsr_expand_all(epsr_ctx *ctx, int nterms, int npt, const float *expr, const float *coeff, float summax,
                    float *out);

/* ---- smoop: replaces auxiliary_routines.smoop(nqmin, nqmax, nsmoo, diff) (:744-764);
 * diff dev, smoothed in place. */
int epsr_smoop(epsr_ctx *ctx, int nqmin, int nqmax, int nsmoo, float *diff);

/* ---- fit_expansion: replaces epsr_standalone_rev3.fit_expansion(yin, expr, yout, wt, wtold,
 * stepr, acceptr, rejrat, chisqnew, nmin, npt, ncfirst, nterms, nits, nfreq) (:4806-4920)
 * for nchain independent chains (datasets or replicas).  Chain c uses yin + c*ldy,
 * expr + c*lde (lde = 0: one basis for all), yout + c*ldo, wt/wtold + c*ldw (all dev,
 * 1-based), the host scalars stepr/acceptr/rejrat/chisqnew [nchain] (in/out, as the
 * reference's return tuple) and the host ran1 state ran1_state + 35*c (in/out).
 * EPSR_FIT_MC replays the reference's Metropolis fit exactly (correctly rounded expf);
 * EPSR_FIT_NORMAL solves the FP64 normal equations for coefficients ncfirst..nterms
 * (no ran1 use; acceptr = 0.05 and rejrat unchanged).  ntries (host, nullable) returns
 * the tries run per chain. */
#define EPSR_FIT_MC 0
#define EPSR_FIT_NORMAL 1
int epsr_fit_expansion(epsr_ctx *ctx, int nchain, const float *yin, int ldy, const float *expr, long long lde,
                       float *yout, int ldo, float *wt, float *wtold, int ldw, float *stepr, float *acceptr,
                       float *rejrat, float *chisqnew, int nmin, int npt, int ncfirst, int nterms, int nits,
                       int nfreq, int32_t *ran1_state, int mode, int *ntries);

/* ---- ptgnsq (epsr_standalone_rev3.py:3789-4021):
 * epsr_ptgnsq_diff forms qtemp(i, id) for every dataset (:3827-3867): q dev [nq+1],
 * data dev [nfile*ldq] (data[id*ldq+i]), qwind dev [npairs*ldq] (intermolecular
 * partials, qwind[ic*ldq+i]), cwt host [nfile*npairs], szeros host [nfile], nqmin/nqmax
 * host [nfile] (1-based), qtemp dev out [nfile*ldq].  X-ray datasets (xfile host
 * [nfile], 0 for neutron data, else the 1-based X-ray file, at most nxfile; NULL: none)
 * take the Q-dependent weights weightx dev [nxfile*ldq*npairs] (double,
 * weightx[((ix-1)*ldq + i)*npairs + ic], the reference's x_rayf0.weightx).  Each term of
 * a weighted sum is added in double and the sum rounded to float, as the reference's
 * float64 weight arrays make it. */
int epsr_ptgnsq_diff(epsr_ctx *ctx, int nfile, int nq, int npairs, const float *q, const float *data, int ldq,
                     const float *qwind, const float *cwt, const float *szeros, const int *nqmin, const int *nqmax,
                     float qminptemp, float *qtemp, const double *weightx, const int *xfile, int nxfile);
/* ---- the data-minus-simulation block of the epsr program between the running averages
 * and ptgnsq (epsr_standalone_rev3.py:1703-1830, with ftback :5330-5389), which the
 * reference keeps inline in runepsr17.  Per dataset: the intramolecular partials (ciwt, or
 * weightx for X-ray data) out of the data datai on its Q range nqmind..nqmaxd and out of
 * the zero limit szeros; then data - intermolecular simulation (cwt or weightx) on the
 * fitted range nqmin..nqmax with the Porod term (porodamp > 0) and the extrapolation to
 * Q = 0 (szeros non-zero); with nsmoop > 0 the part of that difference from r < rshmin
 * (ftback) is subtracted from the data.  Outputs datat dev [nfile*ldq] and szerostemp dev
 * [nfile] (what the program hands ptgnsq as data and szeros); dts dev [ldq] is work
 * space; grt dev [npt1+1] is the program's r-indexed work array, which keeps what earlier
 * calls left in the entries a call does not write (zero it once).  sinf and powf are
 * correctly rounded.  Synchronises the stream. */
typedef struct {
    int nfile, nq, ldq, npairs, nxfile;
    const float *q;                  /* dev [ldq] */
    const float *datai;              /* dev [nfile*ldq] */
    const float *soqintra, *soqinter;/* dev [npairs*ldq], the running averages */
    const double *weightx;           /* dev [nxfile*ldq*npairs] or NULL */
    const float *x;                  /* dev [npt1+1], the r grid */
    const float *ciwt, *cwt;         /* host [nfile*npairs] */
    const int *xfile;                /* host [nfile] */
    const int *nqmind, *nqmaxd, *nqmin, *nqmax;   /* host [nfile], 1-based */
    const float *szeros, *porodamp, *rshmin;       /* host [nfile] */
    float porodq, porodpwr;
    int nsmoop, npt1;
    float rho;
} epsr_dms_desc_t;
int epsr_data_minus_sim(epsr_ctx *ctx, const epsr_dms_desc_t *desc, float *datat, float *szerostemp, float *dts,
                        float *grt);

/* rfactorep of one dataset over i0..i1 (:3977-3991); fit, qtemp dev; rfactor host out. */
int epsr_ptgnsq_rfactor(epsr_ctx *ctx, int i0, int i1, const float *fit, const float *qtemp, float *rfactor);
/* the coefficient update of dataset id (:3962-3975), host arrays: pcoefd[id*ldp+ic],
 * pcoefp[irefp*ldp+ic], cwtpot[ip*ldc+id], pwtold [nterms+1]. */
int epsr_ptgnsq_update(epsr_ctx *ctx, int id, int nterms, int npairs, const float *pwtold, const int *irefpcof,
                       const float *efacp, const float *cwtpot, int ldc, float pfact, float *pcoefd, float *pcoefp,
                       int ldp);

#ifdef __cplusplus
}
#endif
#endif
