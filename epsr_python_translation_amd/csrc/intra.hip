// Intramolecular energies and moves on the GPU: en_mole_intra for every molecule, and the
// MC cycles that move atoms inside molecules (the shake every `intra` cycles and the
// internal rotations every `rotfreq` cycles), which walk the chain in order.
//
// Reference: src/epsr26/library/molfit.py:51-541 (molfit), :545-565 (movei/movea),
// :589-655 (rotati), :671-759 (rotatm, movem), :763-784 (en_change_intra), :805-919
// (en_change), :922-1047 (en_atom_intra), :1327-1349 (en_mole_intra), :1353-1407
// (en_mole_dihedral), :1495-1524 (en_dihedral); src/epsr26/library/ato_manipulation_rev4.py:
// 442-475 (centreofmassi); src/epsr26/library/ergpairreference.py:165-210 (the ledger rows
// make_neighbour_list_from_mols / save_energies / save_erg_pairs work on).
//
// A cycle with intramolecular moves gives every molecule two dependent Metropolis tests
// (internal rotation, then the whole-molecule move, or movei then the internal rotation),
// and the second trial depends on the first outcome; the parallel resolution of the
// whole-molecule cycle does not cover that, so these cycles run the reference's loop in
// order in one workgroup: per move, the lanes evaluate the ledger row's pair energies and
// the molecule's intramolecular pair terms in parallel, and one lane forms the reference's
// sequential float32 sums and the Metropolis test.  Cycles without intramolecular moves go
// to epsr_molfit.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "pairs.h"

void intra_free(IntraTopo *t)
{
    if (!t) return;
    hipFree(t->d_qtype); hipFree(t->d_imageatom); hipFree(t->d_nbond); hipFree(t->d_iabond);
    hipFree(t->d_tbond); hipFree(t->d_amassr); hipFree(t->d_amassa);
    hipFree(t->d_nrf); hipFree(t->d_nrl); hipFree(t->d_labaxi); hipFree(t->d_rotf); hipFree(t->d_rotl);
    hipFree(t->d_labrot); hipFree(t->d_isdih); hipFree(t->d_ndih); hipFree(t->d_labrotdih);
    hipFree(t->d_dihedralang);
    delete t;
}

template <class T>
static int put(epsr_ctx *ctx, T **dst, const T *src, size_t n)
{
    HIP_CHECK(ctx, hipMalloc((void **)dst, sizeof(T) * (n > 0 ? n : 1)));
    if (n && src) HIP_CHECK(ctx, hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    if (n && !src) HIP_CHECK(ctx, hipMemset(*dst, 0, sizeof(T) * n));
    return EPSR_OK;
}

extern "C" int epsr_box_set_intra(epsr_box *box, const epsr_intra_desc_t *d)
{
    if (!box) return EPSR_ERR_ARG;
    epsr_ctx *ctx = box->ctx;
    ARG_CHECK(ctx, d && d->qtype && d->imageatom && d->nbond && d->amassr && d->amassa && d->nrf && d->nrl,
              "epsr_box_set_intra: null argument");
    ARG_CHECK(ctx, d->ldb >= 0 && (d->ldb == 0 || (d->iabond && d->tbond)), "epsr_box_set_intra: bonds");
    ARG_CHECK(ctx, d->ngroup >= 0 && d->nlabrot >= 0 && d->ldd >= 1, "epsr_box_set_intra: group sizes");
    ARG_CHECK(ctx, d->ngroup == 0 || (d->labaxi && d->rotf && d->rotl && d->labrot && d->isdih && d->ndihedralang &&
                                      d->labrotdih && d->dihedralang),
              "epsr_box_set_intra: rotational groups");
    ARG_CHECK(ctx, d->dmaxmol > 0.0f, "epsr_box_set_intra: dmaxmol %g", (double)d->dmaxmol);
    ARG_CHECK(ctx, box->max_atoms_per_mol <= 128, "epsr_box_set_intra: molecules of %d atoms (at most 128)",
              box->max_atoms_per_mol);
    const int na = box->natom, nm = box->nmol;
    // every index must stay inside its array
    for (int a = 0; a < na; ++a) {
        ARG_CHECK(ctx, d->nbond[a] >= 0 && d->nbond[a] <= d->ldb, "epsr_box_set_intra: atom %d has %d bonds", a,
                  d->nbond[a]);
        ARG_CHECK(ctx, d->imageatom[a] >= -1 && d->imageatom[a] < na, "epsr_box_set_intra: image atom of %d", a);
        for (int j = 0; j < d->nbond[a]; ++j)
            ARG_CHECK(ctx, d->iabond[(size_t)a * d->ldb + j] >= 0 && d->iabond[(size_t)a * d->ldb + j] < na,
                      "epsr_box_set_intra: bond %d of atom %d", j, a);
    }
    std::vector<int> ngrot(nm, 0);
    int maxg = 0;
    for (int m = 0; m < nm; ++m) {
        if (d->nrl[m] >= d->nrf[m]) {
            ARG_CHECK(ctx, d->nrf[m] >= 0 && d->nrl[m] < d->ngroup, "epsr_box_set_intra: groups of molecule %d", m);
            ngrot[m] = d->nrl[m] - d->nrf[m] + 1;
        }
        maxg = std::max(maxg, ngrot[m]);
    }
    for (int g = 0; g < d->ngroup; ++g) {
        ARG_CHECK(ctx, d->labaxi[2 * g] >= 0 && d->labaxi[2 * g] < na && d->labaxi[2 * g + 1] >= 0 &&
                           d->labaxi[2 * g + 1] < na,
                  "epsr_box_set_intra: axis of group %d", g);
        ARG_CHECK(ctx, d->rotf[g] >= 0 && d->rotl[g] < d->nlabrot && d->rotl[g] >= d->rotf[g] - 1,
                  "epsr_box_set_intra: atoms of group %d", g);
        ARG_CHECK(ctx, d->ndihedralang[g] >= 0 && d->ndihedralang[g] <= d->ldd, "epsr_box_set_intra: dihedrals %d",
                  g);
        for (int k = 0; k < 2 * d->ndihedralang[g]; ++k)
            ARG_CHECK(ctx, d->labrotdih[(size_t)g * d->ldd * 2 + k] >= 0 && d->labrotdih[(size_t)g * d->ldd * 2 + k] < na,
                      "epsr_box_set_intra: dihedral atoms of group %d", g);
    }
    for (int k = 0; k < d->nlabrot; ++k)
        ARG_CHECK(ctx, d->labrot[k] >= 0 && d->labrot[k] < na, "epsr_box_set_intra: labrot[%d]", k);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    IntraTopo *t = new IntraTopo();
    t->ldb = d->ldb;
    t->ngroup = d->ngroup;
    t->nlabrot = d->nlabrot;
    t->ldd = d->ldd;
    t->refine_intra = d->refine_intra;
    t->refine_intra_fac = d->refine_intra_fac;
    t->dmaxmol = d->dmaxmol;
    t->tempb = d->tempb;
    t->h_ngrot = ngrot;
    t->max_groups = maxg;
    const size_t nb = (size_t)na * d->ldb, ng = (size_t)d->ngroup;
    int rc;
    if ((rc = put(ctx, &t->d_qtype, d->qtype, na)) || (rc = put(ctx, &t->d_imageatom, d->imageatom, na)) ||
        (rc = put(ctx, &t->d_nbond, d->nbond, na)) || (rc = put(ctx, &t->d_iabond, d->iabond, nb)) ||
        (rc = put(ctx, &t->d_tbond, d->tbond, nb)) || (rc = put(ctx, &t->d_amassr, d->amassr, box->npairs)) ||
        (rc = put(ctx, &t->d_amassa, d->amassa, box->ntype)) || (rc = put(ctx, &t->d_nrf, d->nrf, nm)) ||
        (rc = put(ctx, &t->d_nrl, d->nrl, nm)) || (rc = put(ctx, &t->d_labaxi, d->labaxi, 2 * ng)) ||
        (rc = put(ctx, &t->d_rotf, d->rotf, ng)) || (rc = put(ctx, &t->d_rotl, d->rotl, ng)) ||
        (rc = put(ctx, &t->d_labrot, d->labrot, (size_t)d->nlabrot)) || (rc = put(ctx, &t->d_isdih, d->isdih, ng)) ||
        (rc = put(ctx, &t->d_ndih, d->ndihedralang, ng)) ||
        (rc = put(ctx, &t->d_labrotdih, d->labrotdih, ng * d->ldd * 2)) ||
        (rc = put(ctx, &t->d_dihedralang, d->dihedralang, ng * d->ldd))) {
        intra_free(t);
        return rc;
    }
    intra_free(box->intra);
    const_cast<epsr_box *>(box)->intra = t;
    return EPSR_OK;
}

// ------------------------------------------------------------------ device side

// what the intramolecular energy reads, by value
struct IntraDev {
    const int *qtype, *imageatom, *nbond, *iabond;
    const float *tbond, *amassr, *amassa;
    const int *nrf, *nrl, *labaxi, *rotf, *rotl, *labrot, *isdih, *ndih, *labrotdih;
    const float *dang;
    const int *itype, *npairref;
    const float *xstar, *xav, *xav1, *xavs, *xavs1;   // raw tables [npairs*ld]; xav NULL = 0
    int ldb, ldd, ld, npt2, ntype, refine_intra;
    float refine_intra_fac, stepcl, sizefactor, tempb, simt, tempang, tempdih;
};

__device__ __forceinline__ float cr_cosf(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float cr_sinf(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cr_expf(float x) { return (float)exp((double)x); }

__device__ __forceinline__ int ifix_bin(float rat)
{
    return rat < 2.0e9f ? (int)truncf(rat) + 2 : 0x7fffffff;
}

// en_atom_intra (molfit.py:922-1047) of atoms ia < ja of molecule im
__device__ float atom_intra(const IntraDev &d, const float *__restrict__ xa, const float *__restrict__ ya,
                            const float *__restrict__ za, int im, int ia, int ja)
{
    if (ja == d.imageatom[ia]) return 0.0f;
    float sume = 0.0f;
    const int icref = d.npairref[d.itype[ia] * d.ntype + d.itype[ja]];
    const size_t o = (size_t)icref * d.ld;
    const bool notq = d.qtype[ia] == 0 && d.qtype[ja] == 0;
    const float xdifs = xa[ja] - xa[ia], ydifs = ya[ja] - ya[ia], zdifs = za[ja] - za[ia];
    int jref = 0, iflag = 0, image = ia;
    while (jref < d.nbond[image] && iflag == 0) {
        jref = jref + 1;
        if (ja == d.iabond[(size_t)image * d.ldb + jref - 1]) iflag = jref;
    }
    if (iflag == 0) {
        image = -1;
        if (d.imageatom[ia] >= 0) {
            image = d.imageatom[ia];
            jref = 0;
            while (jref < d.nbond[image] && iflag == 0) {
                jref = jref + 1;
                if (ja == d.iabond[(size_t)image * d.ldb + jref - 1]) iflag = jref;
            }
        }
    }
    if (iflag > 0 && image >= 0) {
        const float xdif = dev_mic(xdifs), ydif = dev_mic(ydifs), zdif = dev_mic(zdifs);
        const float dsq = xdif * xdif + ydif * ydif + zdif * zdif;
        const float dtest = sqrtf(dsq);
        const int iz = ifix_bin(dtest / d.stepcl);
        const float wtms = d.amassr[icref];
        const float tbo = d.tbond[(size_t)image * d.ldb + jref - 1];
        const float tb = fabsf(tbo) > d.stepcl ? fabsf(tbo) : d.stepcl;
        float wtm = wtms / tb;
        if (tbo < 0.0f) wtm = wtm * d.tempang;
        const float ddif = tb - dtest;
        sume = wtm * ddif * ddif * d.tempb * d.simt;
        if (d.refine_intra && iz <= d.npt2 && notq) {
            const float dtdif = dtest - d.xstar[iz];
            sume = sume + d.refine_intra_fac * (d.xavs[o + iz] + d.xavs1[o + iz] * dtdif);
        }
        return sume;
    }
    // atom pairs of a dihedral of the molecule's groups are left to en_mole_dihedral
    bool found = false;
    for (int ir = d.nrf[im]; ir <= d.nrl[im] && !found; ++ir)
        for (int kd = 0; kd < d.ndih[ir] && !found; ++kd) {
            const int io1 = d.labrotdih[((size_t)ir * d.ldd + kd) * 2];
            const int io2 = d.labrotdih[((size_t)ir * d.ldd + kd) * 2 + 1];
            found = (io1 == ia && io2 == ja) || (io1 == ja && io2 == ia);
        }
    if (found) return 0.0f;
    const float xdif = dev_mic(xdifs * d.sizefactor), ydif = dev_mic(ydifs * d.sizefactor),
                zdif = dev_mic(zdifs * d.sizefactor);
    const float dsq = xdif * xdif + ydif * ydif + zdif * zdif;
    const float dtest = sqrtf(dsq);
    const int iz = ifix_bin(dtest / d.stepcl);
    if (iz <= d.npt2) {
        const float dtdif = dtest - d.xstar[iz];
        const float v0 = d.xav ? d.xav[o + iz] : 0.0f, v1 = d.xav1 ? d.xav1[o + iz] : 0.0f;
        sume = v0 + v1 * dtdif;
        if (d.refine_intra && notq) sume = sume + d.refine_intra_fac * (d.xavs[o + iz] + d.xavs1[o + iz] * dtdif);
    }
    return sume;
}

__device__ __forceinline__ float scapro(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void vecpro(const float *a, const float *b, float *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// en_dihedral (molfit.py:1495-1524)
__device__ float dihedral(const float *co1, const float *co2, const float *cd, float ang)
{
    float cv1[3], cv2[3], cpar[3];
    vecpro(co1, cd, cv1);
    vecpro(co2, cd, cv2);
    float dmag = sqrtf(scapro(cv1, cv1));
    cv1[0] = cv1[0] / dmag; cv1[1] = cv1[1] / dmag; cv1[2] = cv1[2] / dmag;
    dmag = sqrtf(scapro(cv2, cv2));
    cv2[0] = cv2[0] / dmag; cv2[1] = cv2[1] / dmag; cv2[2] = cv2[2] / dmag;
    float cosphi = scapro(cv1, cv2);
    if (cosphi > 1.0f) cosphi = 1.0f;
    if (cosphi < -1.0f) cosphi = -1.0f;
    vecpro(cv1, cv2, cpar);
    const float sinphi = scapro(cpar, cd);
    const float difc = cr_cosf(ang) - cosphi, difs = cr_sinf(ang) - sinphi;
    return difc * difc + difs * difs;
}

// en_mole_dihedral (molfit.py:1353-1407)
__device__ float mole_dihedral(const IntraDev &d, const float *__restrict__ xa, const float *__restrict__ ya,
                               const float *__restrict__ za, int im)
{
    float e = 0.0f;
    for (int ir = d.nrf[im]; ir <= d.nrl[im]; ++ir) {
        if (!d.isdih[ir]) continue;
        const int a1 = d.labaxi[2 * ir], a2 = d.labaxi[2 * ir + 1];
        const float x1 = xa[a1], y1 = ya[a1], z1 = za[a1], x2 = xa[a2], y2 = ya[a2], z2 = za[a2];
        float cd[3] = {x2 - x1, y2 - y1, z2 - z1};
        const float dmag = sqrtf(scapro(cd, cd));
        cd[0] = cd[0] / dmag; cd[1] = cd[1] / dmag; cd[2] = cd[2] / dmag;
        for (int kd = 0; kd < d.ndih[ir]; ++kd) {
            const int io1 = d.labrotdih[((size_t)ir * d.ldd + kd) * 2];
            const int io2 = d.labrotdih[((size_t)ir * d.ldd + kd) * 2 + 1];
            const float co1[3] = {xa[io1] - x1, ya[io1] - y1, za[io1] - z1};
            const float co2[3] = {xa[io2] - x2, ya[io2] - y2, za[io2] - z2};
            e = e + dihedral(co1, co2, cd, d.dang[(size_t)ir * d.ldd + kd]);
        }
    }
    return e;
}

// en_mole_intra (molfit.py:1327-1349) of one molecule, one lane
__device__ float mole_intra_serial(const IntraDev &d, const float *__restrict__ xa, const float *__restrict__ ya,
                                   const float *__restrict__ za, int im, int a0, int a1)
{
    float e = 0.0f;
    for (int ia = a0; ia < a1; ++ia)
        for (int ja = ia + 1; ja <= a1; ++ja) e = e + atom_intra(d, xa, ya, za, im, ia, ja);
    if (d.nrl[im] >= d.nrf[im] && d.tempdih > 0.0f) e = e + d.tempdih * d.simt * mole_dihedral(d, xa, ya, za, im);
    return e;
}

__global__ void intra_all(int nmol, const int *__restrict__ naf, const int *__restrict__ nal, const float *xa,
                          const float *ya, const float *za, IntraDev d, float *__restrict__ out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmol) return;
    out[m] = mole_intra_serial(d, xa, ya, za, m, naf[m], nal[m]);
}

static int intra_dev(epsr_ctx *ctx, const epsr_box *box, const epsr_tables_t *tb, const float *xav, const float *xav1,
                     float simt, float tempang, float tempdih, IntraDev &d)
{
    const IntraTopo *t = box->intra;
    ARG_CHECK(ctx, t, "no intramolecular topology (call epsr_box_set_intra)");
    ARG_CHECK(ctx, tb && tb->xstar && tb->ld > 0 && tb->npt2 >= 1, "intramolecular energies need the tables' grid");
    ARG_CHECK(ctx, !t->refine_intra || (tb->xavs && tb->xavs1), "refine_intra needs xavs/xavs1");
    d = IntraDev{t->d_qtype, t->d_imageatom, t->d_nbond, t->d_iabond, t->d_tbond, t->d_amassr, t->d_amassa,
                 t->d_nrf, t->d_nrl, t->d_labaxi, t->d_rotf, t->d_rotl, t->d_labrot, t->d_isdih, t->d_ndih,
                 t->d_labrotdih, t->d_dihedralang, box->d_itype, box->d_npairref, tb->xstar, xav, xav1, tb->xavs,
                 tb->xavs1, t->ldb, t->ldd, tb->ld, std::min(tb->npt2, tb->ld - 1), box->ntype, t->refine_intra,
                 t->refine_intra_fac, box->cellst / t->dmaxmol, box->sizefactor, t->tempb, simt, tempang, tempdih};
    return EPSR_OK;
}

extern "C" int epsr_en_mole_intra(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                                  const float *xav, const float *xav1, float simt, float tempang, float tempdih,
                                  float *eintra)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && eintra, "epsr_en_mole_intra: null argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    IntraDev d;
    int rc = intra_dev(ctx, box, tb, xav, xav1, simt, tempang, tempdih, d);
    if (rc) return rc;
    intra_all<<<ceil_div(box->nmol, 64), 64, 0, ctx->stream>>>(box->nmol, box->d_naf, box->d_nal, c->xa, c->ya, c->za,
                                                               d, eintra);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, ctx_sync(ctx));
    return EPSR_OK;
}

// ------------------------------------------------------------------ the in-order cycle

struct SerialArgs {
    int nmol, nmt, shake, rot_cycle, stream_len, natom;
    float temp, trstepmin, rotstepmin, stepr, stepm, stepri, stepmi;
    const int *order, *naf, *nal, *mtype;
    float *xa, *ya, *za, *xadd, *yadd, *zadd;
    float4 *P;
    const int *rowstart, *cnt, *partner, *entry;
    float *ergpair;
    const float *stream;
    float *steps;   // [4*nmt]: wholerot, wholetrans, internrot, interntrans
    int *counts;    // [8*nmt]: nwhole[2], nmovem[2], ninter[2], nmovei[2] per type
    int *nstepreject;
    float *fout;    // fadd, faddb
    int *iout;      // ran1 values used, error, Metropolis tests, acceptances
};

#define SER_THREADS 512
#define SER_MAXAT 128

__device__ __forceinline__ float etest_of(float ediff)
{
    if (ediff <= 0.0f) return 10.0f;
    if (ediff < 20.0f) return cr_expf(-ediff);
    return 0.0f;
}

// en_mole_intra of molecule im by the whole workgroup: pair terms by the lanes in the
// reference's (ia, ja) order, summed in that order by lane 0, then the dihedral term
__device__ float mole_intra_block(const IntraDev &d, const float *xa, const float *ya, const float *za, int im, int a0,
                                  int n, float *terms, float *bcast)
{
    const int np = n * (n - 1) / 2;
    for (int t = threadIdx.x; t < np; t += blockDim.x) {
        int i = 0, rem = t;
        while (rem >= n - 1 - i) {
            rem -= n - 1 - i;
            ++i;
        }
        terms[t] = atom_intra(d, xa, ya, za, im, a0 + i, a0 + i + 1 + rem);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float e = 0.0f;
        for (int t = 0; t < np; ++t) e = e + terms[t];
        if (d.nrl[im] >= d.nrf[im] && d.tempdih > 0.0f) e = e + d.tempdih * d.simt * mole_dihedral(d, xa, ya, za, im);
        *bcast = e;
    }
    __syncthreads();
    return *bcast;
}

// en_change (molfit.py:805-919) against the rows eo/en of the ledger: en[1..nn) by the
// lanes (molecule's atoms at P[a0..a0+n)), fdiff by lane 0 in slot order
__device__ float en_change_block(const SerialArgs &a, const PairTab &t, int row, int nn, int a0, int n, const float *eo,
                                 float *en, float *bcast)
{
    for (int s = 1 + threadIdx.x; s < nn; s += blockDim.x) {
        const int jm = a.partner[row + s];
        const int b0 = a.naf[jm];
        en[s] = 0.0f + mol_pair(a.P + a0, n, a.P + b0, a.nal[jm] - b0 + 1, t);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float fdiff = 0.0f;
        for (int s = 1; s < nn; ++s) fdiff = fdiff + en[s] - eo[s];
        *bcast = fdiff;
    }
    __syncthreads();
    return *bcast;
}

// one MC cycle in chain order (molfit.py:51-463), one workgroup
__global__ void __launch_bounds__(SER_THREADS)
mc_serial_cycle(SerialArgs a, IntraDev d, PairTab t)
{
    extern __shared__ float smem[];
    const int nrowmax = a.iout[4];
    float *eo = smem, *en = eo + nrowmax, *terms = en + nrowmax;
    float *xr = terms + SER_MAXAT * (SER_MAXAT - 1) / 2;   // saved coordinates [3*SER_MAXAT]
    __shared__ float bc[8];
    __shared__ int ib[8];
    __shared__ float steps[4 * 16];
    __shared__ int counts[8 * 16];
    const int tid = threadIdx.x;
    for (int i = tid; i < 4 * a.nmt; i += blockDim.x) steps[i] = a.steps[i];
    for (int i = tid; i < 8 * a.nmt; i += blockDim.x) counts[i] = 0;
    for (int i = tid; i < 2 * nrowmax; i += blockDim.x) eo[i] = 0.0f;
    if (tid == 0) {
        ib[0] = 0;   // ran1 cursor
        ib[1] = 0;   // error
        ib[2] = 0;   // Metropolis tests
        ib[3] = 0;   // acceptances
        bc[4] = 0.0f;   // fadd
        bc[5] = 0.0f;   // faddb
    }
    __syncthreads();
    const float pi = 3.14159265358979323846f;
    // ran1 values are read at ib[0] by lane 0 only; an exhausted stream sets the error and
    // ends the cycle at the next molecule
    auto draw = [&](int k) -> float {
        const int i = ib[0] + k;
        return i < a.stream_len ? a.stream[i] : 2.0f;
    };
    for (int pos = 0; pos < a.nmol; ++pos) {
        if (ib[1]) break;
        const int im = a.shake ? pos : a.order[pos];
        const int a0 = a.naf[im], n = a.nal[im] - a0 + 1, mt = a.mtype[im];
        const int row = a.rowstart[im], nn = a.cnt[im];
        const int ngrot = d.nrl[im] >= d.nrf[im] ? d.nrl[im] - d.nrf[im] + 1 : 0;
        // make_neighbour_list_from_mols: ergold = ergnew = the ledger row
        for (int s = tid; s < nn; s += blockDim.x) {
            const float v = a.ergpair[a.entry[row + s]];
            eo[s] = v;
            en[s] = v;
        }
        // the worst case of this molecule's draws must fit the stream (movem's doubling
        // checks its own)
        if (tid == 0 && ib[0] + 3 * n + ngrot + 8 > a.stream_len) ib[1] = 1;
        __syncthreads();
        if (ib[1]) break;
        if (a.shake && n > 1 && a.stepmi > 0.0f) {
            // movei + en_change_intra (molfit.py:141-190)
            if (tid == 0) {
                float s = steps[3 * a.nmt + mt];   // interntrans of the type
                if (a.stepmi > 0.0f && s == 0.0f) s = a.trstepmin;
                steps[3 * a.nmt + mt] = s;
                counts[8 * mt + 6] += 1;
                bc[0] = s;
            }
            __syncthreads();
            const float step = bc[0];
            const int cur = ib[0];
            if (tid < n) {
                const int ia = a0 + tid;
                xr[3 * tid] = a.xa[ia];
                xr[3 * tid + 1] = a.ya[ia];
                xr[3 * tid + 2] = a.za[ia];
                const float r0 = a.stream[cur + 3 * tid], r1 = a.stream[cur + 3 * tid + 1],
                            r2 = a.stream[cur + 3 * tid + 2];
                a.za[ia] = a.za[ia] + step * (2.0f * r0 - 1.0f);
                a.ya[ia] = a.ya[ia] + step * (2.0f * r1 - 1.0f);
                a.xa[ia] = a.xa[ia] + step * (2.0f * r2 - 1.0f);
            }
            __syncthreads();
            if (tid == 0) ib[0] = cur + 3 * n;
            const float e = mole_intra_block(d, a.xa, a.ya, a.za, im, a0, n, terms, bc + 1);
            if (tid == 0) {
                en[0] = e;
                const float fdiffb = en[0] - eo[0];
                const float ediff = fdiffb * a.temp;
                const float r = draw(0);
                ib[0] += 1;
                ib[2] += 1;
                const int acc = r < etest_of(ediff);
                ib[5] = acc;
                if (acc) {
                    bc[5] = bc[5] + fdiffb;
                    ib[3] += 1;
                } else {
                    counts[8 * mt + 7] += 1;
                }
            }
            __syncthreads();
            if (ib[5]) {
                for (int s = tid; s < nn; s += blockDim.x) eo[s] = en[s];
                // centreofmassi (ato_manipulation_rev4.py:442-475)
                if (tid == 0) {
                    float x1 = 0.0f, y1 = 0.0f, z1 = 0.0f, summass = 0.0f;
                    for (int ia = a0; ia < a0 + n; ++ia) {
                        if (d.qtype[ia] != 0) continue;
                        const float amass = d.amassa[d.itype[ia]];
                        x1 = x1 + a.xa[ia] * amass;
                        y1 = y1 + a.ya[ia] * amass;
                        z1 = z1 + a.za[ia] * amass;
                        summass = summass + amass;
                    }
                    if (summass > 0.0f) {
                        x1 = x1 / summass;
                        y1 = y1 / summass;
                        z1 = z1 / summass;
                    }
                    bc[1] = x1; bc[2] = y1; bc[3] = z1;
                }
                __syncthreads();
                if (tid < n) {
                    const int ia = a0 + tid;
                    a.xa[ia] = a.xa[ia] - bc[1];
                    a.ya[ia] = a.ya[ia] - bc[2];
                    a.za[ia] = a.za[ia] - bc[3];
                }
            } else if (tid < n) {
                const int ia = a0 + tid;
                a.xa[ia] = xr[3 * tid];
                a.ya[ia] = xr[3 * tid + 1];
                a.za[ia] = xr[3 * tid + 2];
            }
            __syncthreads();
            // save_erg_pairs
            for (int s = tid; s < nn; s += blockDim.x) a.ergpair[a.entry[row + s]] = eo[s];
            __syncthreads();
        }
        const bool rot_now = a.stepri > 0.0f && ngrot > 0 && (a.shake || a.rot_cycle);
        if (rot_now) {
            // internal rotations + en_change_intra + en_change (molfit.py:191-265, 294-361)
            if (tid < n) {
                const int ia = a0 + tid;
                xr[3 * tid] = a.xa[ia];
                xr[3 * tid + 1] = a.ya[ia];
                xr[3 * tid + 2] = a.za[ia];
            }
            __syncthreads();
            if (tid == 0) {
                if (steps[2 * a.nmt + mt] == 0.0f) steps[2 * a.nmt + mt] = a.rotstepmin;
                counts[8 * mt + 4] += 1;
                const float step = steps[2 * a.nmt + mt];
                for (int ig = d.nrf[im]; ig <= d.nrl[im]; ++ig) {
                    // rotati (molfit.py:589-655)
                    const float phi = pi * step * (2.0f * draw(0) - 1.0f);
                    ib[0] += 1;
                    const float cphi = cr_cosf(phi);
                    const float sphi = copysignf(fabsf(sqrtf(1.0f - cphi * cphi)), phi);
                    const int x1 = d.labaxi[2 * ig], x2 = d.labaxi[2 * ig + 1];
                    float p[3] = {a.xa[x2] - a.xa[x1], a.ya[x2] - a.ya[x1], a.za[x2] - a.za[x1]};
                    const float rab2 = scapro(p, p);
                    if (rab2 > 0.0f) {
                        const float rab = sqrtf(rab2);
                        p[0] = p[0] / rab; p[1] = p[1] / rab; p[2] = p[2] / rab;
                        for (int k = d.rotf[ig]; k <= d.rotl[ig]; ++k) {
                            const int nat = d.labrot[k];
                            const float xv[3] = {a.xa[nat] - a.xa[x1], a.ya[nat] - a.ya[x1], a.za[nat] - a.za[x1]};
                            const float frac = scapro(p, xv);
                            float xp[3], xrv[3], xrp[3];
                            for (int i = 0; i < 3; ++i) xp[i] = frac * p[i];
                            for (int i = 0; i < 3; ++i) xrv[i] = xv[i] - xp[i];
                            vecpro(p, xrv, xrp);
                            a.xa[nat] = a.xa[x1] + xp[0];
                            a.ya[nat] = a.ya[x1] + xp[1];
                            a.za[nat] = a.za[x1] + xp[2];
                            a.xa[nat] = a.xa[nat] + xrv[0] * cphi + xrp[0] * sphi;
                            a.ya[nat] = a.ya[nat] + xrv[1] * cphi + xrp[1] * sphi;
                            a.za[nat] = a.za[nat] + xrv[2] * cphi + xrp[2] * sphi;
                        }
                    }
                }
            }
            __syncthreads();
            const float e = mole_intra_block(d, a.xa, a.ya, a.za, im, a0, n, terms, bc + 1);
            if (tid < n) {
                const int ia = a0 + tid;
                const float4 q = a.P[ia];
                a.P[ia] = make_float4(a.xa[ia] + a.xadd[im], a.ya[ia] + a.yadd[im], a.za[ia] + a.zadd[im], q.w);
            }
            if (tid == 0) en[0] = e;
            __syncthreads();
            const float fdiff = en_change_block(a, t, row, nn, a0, n, eo, en, bc + 2);
            if (tid == 0) {
                const float fdiffb = en[0] - eo[0];
                const float ediff = (fdiff + fdiffb) * a.temp;
                const float r = draw(0);
                ib[0] += 1;
                ib[2] += 1;
                const int acc = r < etest_of(ediff);
                ib[5] = acc;
                if (acc) {
                    a.nstepreject[im] = 0;
                    bc[4] = bc[4] + fdiff;
                    bc[5] = bc[5] + fdiffb;
                    ib[3] += 1;
                } else {
                    a.nstepreject[im] += 1;
                    counts[8 * mt + 5] += 1;
                }
            }
            __syncthreads();
            if (ib[5]) {
                for (int s = tid; s < nn; s += blockDim.x) eo[s] = en[s];
            } else if (tid < n) {
                const int ia = a0 + tid;
                a.xa[ia] = xr[3 * tid];
                a.ya[ia] = xr[3 * tid + 1];
                a.za[ia] = xr[3 * tid + 2];
                const float4 q = a.P[ia];
                a.P[ia] = make_float4(a.xa[ia] + a.xadd[im], a.ya[ia] + a.yadd[im], a.za[ia] + a.zadd[im], q.w);
            }
            __syncthreads();
            for (int s = tid; s < nn; s += blockDim.x) a.ergpair[a.entry[row + s]] = eo[s];
            __syncthreads();
        } else if (ngrot == 0 && tid == 0) {
            steps[2 * a.nmt + mt] = 0.0f;
        }
        if (a.shake) continue;
        // whole-molecule rotation and translation (molfit.py:362-463)
        if (tid < n) {
            const int ia = a0 + tid;
            xr[3 * tid] = a.xa[ia];
            xr[3 * tid + 1] = a.ya[ia];
            xr[3 * tid + 2] = a.za[ia];
        }
        if (tid == 0) {
            bc[1] = a.xadd[im];
            bc[2] = a.yadd[im];
            bc[3] = a.zadd[im];
            int rot = 0;
            if (a.stepr > 0.0f && n > 1) {
                if (steps[mt] <= 0.0f) steps[mt] = a.rotstepmin;
                counts[8 * mt + 0] += 1;
                rot = 1;
            } else {
                steps[mt] = 0.0f;
            }
            ib[6] = rot ? ib[0] : -1;   // the rotation's three values
            if (rot) ib[0] += 3;
        }
        __syncthreads();
        if (ib[6] >= 0 && tid < n) {
            // rotatm (molfit.py:671-711): z, then y, then x
            const int ia = a0 + tid;
            const float st = steps[mt];
            float x = a.xa[ia], y = a.ya[ia], z = a.za[ia];
            for (int axis = 0; axis < 3; ++axis) {
                const float phi = pi * st * (2.0f * a.stream[ib[6] + axis] - 1.0f);
                const float cphi = cr_cosf(phi);
                const float sphi = copysignf(fabsf(sqrtf(1.0f - cphi * cphi)), phi);
                if (axis == 0) {
                    const float s = x * cphi - y * sphi;
                    y = x * sphi + y * cphi;
                    x = s;
                } else if (axis == 1) {
                    const float s = z * cphi - x * sphi;
                    x = z * sphi + x * cphi;
                    z = s;
                } else {
                    const float s = y * cphi - z * sphi;
                    z = y * sphi + z * cphi;
                    y = s;
                }
            }
            a.xa[ia] = x;
            a.ya[ia] = y;
            a.za[ia] = z;
        }
        if (tid == 0) {
            if (a.stepm > 0.0f) {
                // movem (molfit.py:732-759): the step doubles until the centre moves
                if (steps[a.nmt + mt] <= 0.0f) steps[a.nmt + mt] = a.trstepmin;
                counts[8 * mt + 2] += 1;
                float sz = steps[a.nmt + mt];
                const float xs = a.xadd[im], ys = a.yadd[im], zs = a.zadd[im];
                a.zadd[im] = a.zadd[im] + sz * (2.0f * draw(0) - 1.0f);
                a.yadd[im] = a.yadd[im] + sz * (2.0f * draw(1) - 1.0f);
                a.xadd[im] = a.xadd[im] + sz * (2.0f * draw(2) - 1.0f);
                ib[0] += 3;
                bool nochange = xs == a.xadd[im] && ys == a.yadd[im] && zs == a.zadd[im];
                while (nochange) {
                    if (ib[0] + 4 > a.stream_len) {
                        ib[1] = 1;
                        break;
                    }
                    sz = sz * 2.0f;
                    a.zadd[im] = a.zadd[im] + sz * (2.0f * draw(0) - 1.0f);
                    a.yadd[im] = a.yadd[im] + sz * (2.0f * draw(1) - 1.0f);
                    a.xadd[im] = a.xadd[im] + sz * (2.0f * draw(2) - 1.0f);
                    ib[0] += 3;
                    nochange = xs == a.xadd[im] && ys == a.yadd[im] && zs == a.zadd[im];
                }
                steps[a.nmt + mt] = sz;
            } else {
                steps[a.nmt + mt] = 0.0f;
            }
        }
        __syncthreads();
        if (ib[1]) break;
        if (tid < n) {
            const int ia = a0 + tid;
            const float4 q = a.P[ia];
            a.P[ia] = make_float4(a.xa[ia] + a.xadd[im], a.ya[ia] + a.yadd[im], a.za[ia] + a.zadd[im], q.w);
        }
        __syncthreads();
        const float fdiff = en_change_block(a, t, row, nn, a0, n, eo, en, bc + 0);
        if (tid == 0) {
            const float ediff = (fdiff + 0.0f) * a.temp;
            const float r = draw(0);
            ib[0] += 1;
            ib[2] += 1;
            const int acc = r < etest_of(ediff);
            ib[5] = acc;
            if (acc) {
                a.nstepreject[im] = 0;
                bc[4] = bc[4] + fdiff;
                bc[5] = bc[5] + 0.0f;
                ib[3] += 1;
            } else {
                a.nstepreject[im] += 1;
                a.xadd[im] = bc[1];
                a.yadd[im] = bc[2];
                a.zadd[im] = bc[3];
                if (a.stepr > 0.0f) counts[8 * mt + 1] += 1;
                if (a.stepm > 0.0f) counts[8 * mt + 3] += 1;
            }
        }
        __syncthreads();
        if (ib[5]) {
            for (int s = tid; s < nn; s += blockDim.x) eo[s] = en[s];
        } else if (tid < n) {
            const int ia = a0 + tid;
            a.xa[ia] = xr[3 * tid];
            a.ya[ia] = xr[3 * tid + 1];
            a.za[ia] = xr[3 * tid + 2];
            const float4 q = a.P[ia];
            a.P[ia] = make_float4(a.xa[ia] + a.xadd[im], a.ya[ia] + a.yadd[im], a.za[ia] + a.zadd[im], q.w);
        }
        __syncthreads();
        for (int s = tid; s < nn; s += blockDim.x) a.ergpair[a.entry[row + s]] = eo[s];
        __syncthreads();
    }
    __syncthreads();
    for (int i = tid; i < 4 * a.nmt; i += blockDim.x) a.steps[i] = steps[i];
    for (int i = tid; i < 8 * a.nmt; i += blockDim.x) a.counts[i] = counts[i];
    if (tid == 0) {
        a.fout[0] = bc[4];
        a.fout[1] = bc[5];
        a.iout[0] = ib[0];
        a.iout[1] = ib[1];
        a.iout[2] = ib[2];
        a.iout[3] = ib[3];
    }
}

__global__ void abs_pos_all(int natom, const float *__restrict__ xa, const float *__restrict__ ya,
                            const float *__restrict__ za, const float *__restrict__ xadd,
                            const float *__restrict__ yadd, const float *__restrict__ zadd, const int *__restrict__ amol,
                            const int *__restrict__ itype, float4 *__restrict__ P)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    const int m = amol[a];
    P[a] = make_float4(xa[a] + xadd[m], ya[a] + yadd[m], za[a] + zadd[m], __int_as_float(itype[a]));
}

static float adapt(float rejrate, int att, int rej)
{
    float rat;
    if (rej > 0) {
        rat = rejrate * (float)att / (float)rej;
        if (rat > 1.1f) rat = 1.1f;
        if (rat < 0.9f) rat = 0.9f;
    } else {
        rat = att > 0 ? 1.1f : 1.0f;
    }
    return rat;
}

static void clamp_step(float &s, float mx, float mn)
{
    if (s > mx) s = mx;
    if (s < mn && s != 0.0f) s = mn;
}

extern "C" int epsr_molfit_full(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                                const float *xav, const float *xav1, epsr_ledger *l, const epsr_mc_full_params_t *p,
                                epsr_mc_steps_t *S, int *nstepreject, int32_t *ran1_state, float *fadds,
                                float *faddbs, int *nisum, epsr_mc_stats_t *stats)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && tb && l && p && S && nstepreject && ran1_state && fadds && faddbs && nisum,
              "epsr_molfit_full: null argument");
    ARG_CHECK(ctx, S->wholerot && S->wholetrans && S->internrot && S->interntrans && S->nwhole && S->nmovem &&
                       S->ninter && S->nmovei,
              "epsr_molfit_full: null step array");
    ARG_CHECK(ctx, l->ledger_ready, "epsr_molfit_full: no ledger (call epsr_fmolec1)");
    ARG_CHECK(ctx, p->simt > 0.0f, "epsr_molfit_full: simt %g", (double)p->simt);
    const int nmol = box->nmol, nmt = box->nmoltype;
    const bool shake = *nisum >= p->intra;
    bool any_groups = false;
    if (box->intra)
        for (int g : box->intra->h_ngrot) any_groups = any_groups || g > 0;
    bool rot_cycle = false;
    if (!shake && p->stepri > 0.0f && any_groups) {
        // nisum == rotfreq * (nisum / rotfreq) (molfit.py:295), Fortran integer division
        ARG_CHECK(ctx, p->rotfreq != 0, "epsr_molfit_full: rotfreq = 0 with internal rotations");
        rot_cycle = *nisum == p->rotfreq * (*nisum / p->rotfreq);
    }
    const bool shake_moves = shake && (p->stepmi > 0.0f || (p->stepri > 0.0f && any_groups));
    float trstepmax = powf(1.0f / box->natomr, 1.0f / 3.0f);
    float trstepmin = 0.0001f * trstepmax;
    if (trstepmin < 1.0e-6f) trstepmin = 1.0e-6f;
    const float rotstepmax = 1.0f, rotstepmin = 1.0e-5f;
    std::vector<float> wr(S->wholerot, S->wholerot + nmt), wt(S->wholetrans, S->wholetrans + nmt),
        ir(S->internrot, S->internrot + nmt), it(S->interntrans, S->interntrans + nmt);
    std::vector<int> cnt(8 * (size_t)nmt, 0);
    float fadd = 0.0f, faddb = 0.0f;
    int ntests = 0, nacc = 0;
    HIP_CHECK(ctx, hipSetDevice(ctx->device));

    if (!shake && !rot_cycle) {
        // whole-molecule moves only: the parallel resolution
        epsr_mc_params_t q{p->simt, p->stepr, p->stepm, p->rejrate, 1};   // adaptation below
        std::vector<int> nwh(2 * nmt), nmv(2 * nmt);
        epsr_mc_stats_t st{};
        int rc = epsr_molfit(ctx, box, c, tb, l, &q, wr.data(), wt.data(), nwh.data(), nmv.data(), nstepreject,
                             ran1_state, fadds, &st);
        if (rc) return rc;
        for (int mt = 0; mt < nmt; ++mt) {
            cnt[8 * mt + 0] = nwh[2 * mt];
            cnt[8 * mt + 1] = nwh[2 * mt + 1];
            cnt[8 * mt + 2] = nmv[2 * mt];
            cnt[8 * mt + 3] = nmv[2 * mt + 1];
        }
        // molecules without rotational groups zero their type's internal rotation step
        // (molfit.py:353-355), whatever the cycle
        for (int m = 0; m < nmol; ++m)
            if (!box->intra || box->intra->h_ngrot[m] == 0) ir[box->h_mtype[m]] = 0.0f;
        ntests = st.moves;
        nacc = st.accepted;
        if (stats) *stats = st;
        *faddbs = *faddbs + faddb / (float)nmol;
    } else if (shake && !shake_moves) {
        // a shake cycle that moves nothing: only the internal rotation step of types
        // without rotational groups is zeroed (molfit.py:266-268)
        for (int m = 0; m < nmol; ++m)
            if (!box->intra || box->intra->h_ngrot[m] == 0) ir[box->h_mtype[m]] = 0.0f;
        *fadds = *fadds + fadd / (float)nmol;
        *faddbs = *faddbs + faddb / (float)nmol;
        if (stats) *stats = epsr_mc_stats_t{};
    } else {
        ARG_CHECK(ctx, !shake_moves || box->intra, "epsr_molfit_full: intramolecular moves need epsr_box_set_intra");
        ARG_CHECK(ctx, nmt <= 16, "epsr_molfit_full: %d molecule types (at most 16 in the in-order cycle)", nmt);
        IntraDev d;
        int rc = intra_dev(ctx, box, tb, xav, xav1, p->simt, p->tempangwt, p->tempdihwt * box->intra->tempb, d);
        if (rc) return rc;
        hipStream_t st = ctx->stream;
        // the ran1 values the cycle can use: every draw of every branch plus room for movem's
        // doubling
        long long need = 0;
        for (int m = 0; m < nmol; ++m) {
            const int n = box->h_nal[m] - box->h_naf[m] + 1;
            const int g = box->intra ? box->intra->h_ngrot[m] : 0;
            need += 3LL * n + g + 1 + 3 + 3 + 1 + 1;
        }
        need += 3 * 4096;
        ARG_CHECK(ctx, need < (1LL << 30), "epsr_molfit_full: stream of %lld values", need);
        const int nrowmax = std::max(1, l->maxcnt);
        const size_t lds = sizeof(float) * (2 * (size_t)nrowmax + SER_MAXAT * (SER_MAXAT - 1) / 2 + 3 * SER_MAXAT);
        ARG_CHECK(ctx, lds <= 150 * 1024, "epsr_molfit_full: ledger rows of %d slots exceed LDS", nrowmax);
        Ran1 rng;
        ran1_load(rng, ran1_state);
        const Ran1 rng0 = rng;
        std::vector<float> hstream((size_t)need);
        ran1_fill(rng, hstream.data(), need);
        StreamTemp T(ctx);
        float *dstream = (float *)T.alloc(sizeof(float) * (size_t)need);
        float4 *dP = (float4 *)T.alloc(sizeof(float4) * (size_t)box->natom);
        float4 *dT = (float4 *)T.alloc(sizeof(float4) * (size_t)box->npairs *
                                       std::max(2, std::min(tb->ld, tb->npt2 + 1)));
        int *dtmp = (int *)T.alloc(sizeof(int) * 4);
        float *dsteps = (float *)T.alloc(sizeof(float) * 4 * nmt);
        int *dcounts = (int *)T.alloc(sizeof(int) * 8 * nmt);
        float *dfout = (float *)T.alloc(sizeof(float) * 2);
        int *diout = (int *)T.alloc(sizeof(int) * 8);
        if (!dstream || !dP || !dT || !dtmp || !dsteps || !dcounts || !dfout || !diout) return EPSR_ERR_HIP;
        PreparedTables pt;
        if ((rc = mc_prepare_tables(ctx, box, tb, dT, dtmp, pt))) return rc;
        std::vector<float> hsteps(4 * (size_t)nmt);
        for (int mt = 0; mt < nmt; ++mt) {
            hsteps[mt] = wr[mt];
            hsteps[nmt + mt] = wt[mt];
            hsteps[2 * nmt + mt] = ir[mt];
            hsteps[3 * nmt + mt] = it[mt];
        }
        int hiin[8] = {0, 0, 0, 0, nrowmax, 0, 0, 0};
        HIP_CHECK(ctx, hipMemcpyAsync(dstream, hstream.data(), sizeof(float) * need, hipMemcpyHostToDevice, st));
        HIP_CHECK(ctx, hipMemcpyAsync(dsteps, hsteps.data(), sizeof(float) * 4 * nmt, hipMemcpyHostToDevice, st));
        HIP_CHECK(ctx, hipMemcpyAsync(diout, hiin, sizeof(int) * 8, hipMemcpyHostToDevice, st));
        abs_pos_all<<<ceil_div(box->natom, 256), 256, 0, st>>>(box->natom, c->xa, c->ya, c->za, c->xadd, c->yadd,
                                                               c->zadd, box->d_amol, box->d_itype, dP);
        SerialArgs a{nmol, nmt, shake ? 1 : 0, rot_cycle ? 1 : 0, (int)need, box->natom, 1.0f / p->simt, trstepmin,
                     rotstepmin, p->stepr, p->stepm, p->stepri, p->stepmi, l->d_order, box->d_naf, box->d_nal,
                     box->d_mtype, c->xa, c->ya, c->za, c->xadd, c->yadd, c->zadd, dP, l->d_rowstart, l->d_cnt,
                     l->d_partner, l->d_entry, l->d_ergpair, dstream, dsteps, dcounts, nstepreject, dfout, diout};
        timer_begin(ctx, "mc_serial");
        mc_serial_cycle<<<1, SER_THREADS, lds, st>>>(a, d, pt.tab);
        timer_end(ctx, "mc_serial");
        HIP_CHECK(ctx, hipGetLastError());
        float hf[2];
        int hio[8];
        HIP_CHECK(ctx, hipMemcpyAsync(hsteps.data(), dsteps, sizeof(float) * 4 * nmt, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(cnt.data(), dcounts, sizeof(int) * 8 * nmt, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(hf, dfout, sizeof(float) * 2, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(hio, diout, sizeof(int) * 8, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
        if (hio[1])
            return set_error(ctx, EPSR_ERR_CAPACITY,
                             "epsr_molfit_full: the cycle used up its %lld ran1 values (movem doubled its step too "
                             "often); coordinates and ledger are left mid-cycle",
                             need);
        for (int mt = 0; mt < nmt; ++mt) {
            wr[mt] = hsteps[mt];
            wt[mt] = hsteps[nmt + mt];
            ir[mt] = hsteps[2 * nmt + mt];
            it[mt] = hsteps[3 * nmt + mt];
        }
        // the caller's ran1 state after exactly the values the cycle read
        rng = rng0;
        for (int k = 0; k < hio[0]; ++k) (void)ran1_next(rng);
        ran1_store(rng, ran1_state);
        fadd = hf[0];
        faddb = hf[1];
        ntests = hio[2];
        nacc = hio[3];
        *fadds = *fadds + fadd / (float)nmol;
        *faddbs = *faddbs + faddb / (float)nmol;
        if (stats) {
            *stats = epsr_mc_stats_t{};
            stats->moves = ntests;
            stats->accepted = nacc;
            stats->cycles = 1;
            stats->accepted_total = nacc;
            stats->ms_resolve = ctx->timing ? epsr_ctx_last_kernel_ms(ctx, "mc_serial") : -1.0;
        }
    }
    *nisum = shake ? 1 : *nisum + 1;
    // step-size adaptation (molfit.py:465-536)
    if (p->ireset != 1) {
        for (int mt = 0; mt < nmt; ++mt) {
            it[mt] = it[mt] * adapt(p->rejrate, cnt[8 * mt + 6], cnt[8 * mt + 7]);
            clamp_step(it[mt], trstepmax, trstepmin);
            ir[mt] = ir[mt] * adapt(p->rejrate, cnt[8 * mt + 4], cnt[8 * mt + 5]);
            clamp_step(ir[mt], rotstepmax, rotstepmin);
            wr[mt] = wr[mt] * adapt(p->rejrate, cnt[8 * mt + 0], cnt[8 * mt + 1]);
            clamp_step(wr[mt], rotstepmax, rotstepmin);
            wt[mt] = wt[mt] * adapt(p->rejrate, cnt[8 * mt + 2], cnt[8 * mt + 3]);
            clamp_step(wt[mt], trstepmax, trstepmin);
        }
    }
    for (int mt = 0; mt < nmt; ++mt) {
        S->wholerot[mt] = wr[mt];
        S->wholetrans[mt] = wt[mt];
        S->internrot[mt] = ir[mt];
        S->interntrans[mt] = it[mt];
        S->nwhole[2 * mt] = cnt[8 * mt + 0];
        S->nwhole[2 * mt + 1] = cnt[8 * mt + 1];
        S->nmovem[2 * mt] = cnt[8 * mt + 2];
        S->nmovem[2 * mt + 1] = cnt[8 * mt + 3];
        S->ninter[2 * mt] = cnt[8 * mt + 4];
        S->ninter[2 * mt + 1] = cnt[8 * mt + 5];
        S->nmovei[2 * mt] = cnt[8 * mt + 6];
        S->nmovei[2 * mt + 1] = cnt[8 * mt + 7];
    }
    return EPSR_OK;
}
