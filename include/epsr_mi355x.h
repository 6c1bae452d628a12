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

/* ---- foqols: the r -> Q transform matrix built by inita
 * (src/epsr26/programs/epsr_standalone_rev3.py:3025-3071, j1qr :3319-3329, pofr :3389-3399).
 * q dev [nq+1], x dev [npt1+1], r2 dev [npt1+1] (bin edges of the serial r1/r2
 * recurrence), ycoh dev [npt1+1]; foqols dev [(nq+1)*ldf].  Double-precision j1qr
 * differences and correctly rounded expf/sinf/cosf. */
int epsr_foqols(epsr_ctx *ctx, int nq, const float *q, int npt1, const float *x, const float *r2,
                const float *ycoh, float rho, float fwhm, float fwhmq, float lorchqwidth,
                float lorchqwidthq, float rmaxgr, float *foqols, int ldf);

#ifdef __cplusplus
}
#endif
#endif
