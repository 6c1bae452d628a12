// Cell lists, the pair-energy ledger (fmolec1), en_change, en_mole_pair and the MC
// cycle (molfit) on the GPU.
//
// Reference: src/epsr26/library/cellneighbourlist.py:23-330 (cells),
// src/epsr26/library/ergpairreference.py:63-213 (ledger),
// src/epsr26/library/ato_manipulation_rev4.py:2374-2528 (fmolec1),
// src/epsr26/library/molfit.py:51-541, 671-759, 805-919, 1250-1279, 1527-1578.
//
// The MC cycle visits molecules in a fixed chain order (cells, then molecules in a
// cell) and every acceptance depends on the one before.  The GPU version splits it:
//   A  trial moves of every molecule at once (their ran1 offsets are known up front:
//      3 for the rotation, 3 for the translation, 1 for the Metropolis test);
//   B  for every ledger slot (molecule m, partner n) the pair energies the chain can
//      need: E(m_trial, n_current) and, when n moves earlier in the chain, also
//      E(m_trial, n_trial) -- massively parallel, one lane per slot;
//   C  the decisions: from rigorous bounds on the reference's float32 sums, every move
//      whose outcome is already fixed by the decided part of its neighbourhood is
//      decided at once, in passes over the undecided moves; a move whose earlier
//      partners are all decided is decided exactly (its sequential float32 sum when the
//      bounds leave it open), so the earliest undecided move always decides;
//   D  the accepted trial positions and the final ledger values are written back.
// Because every pair energy and every sum is the reference's float32 expression, the
// result is the reference's (with a correctly rounded cosf/expf).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <numeric>

#include "common.h"
#include "pairs.h"
#include "surrogate.h"

template <class T>
static int dev_alloc(epsr_ctx *ctx, T **p, size_t n)
{
    if (*p) hipFree(*p);
    *p = nullptr;
    HIP_CHECK(ctx, hipMalloc((void **)p, sizeof(T) * (n > 0 ? n : 1)));
    return EPSR_OK;
}

extern "C" int epsr_ledger_create(epsr_ctx *ctx, const epsr_box *box, epsr_ledger **out)
{
    if (!ctx || !box || !out) return set_error(ctx, EPSR_ERR_ARG, "epsr_ledger_create: null argument");
    epsr_ledger *l = new epsr_ledger();
    l->ctx = ctx;
    l->box = box;
    l->nmol = box->nmol;
    int rc;
    if ((rc = dev_alloc(ctx, &l->d_order, box->nmol)) || (rc = dev_alloc(ctx, &l->d_pos, box->nmol)) ||
        (rc = dev_alloc(ctx, &l->d_rowstart, box->nmol + 1)) || (rc = dev_alloc(ctx, &l->d_cnt, box->nmol)) ||
        (rc = dev_alloc(ctx, &l->d_rowk, box->nmol + 1))) {
        epsr_ledger_destroy(l);
        return rc;
    }
    *out = l;
    return EPSR_OK;
}

extern "C" void epsr_ledger_destroy(epsr_ledger *l)
{
    if (!l) return;
    hipFree(l->d_rowstart); hipFree(l->d_cnt); hipFree(l->d_order); hipFree(l->d_pos); hipFree(l->d_rowk);
    hipFree(l->d_partner); hipFree(l->d_entry); hipFree(l->d_mirror); hipFree(l->d_slotmol);
    hipFree(l->d_ergpair);
    hipFree(l->d_roff); hipFree(l->d_toff); hipFree(l->d_moff);
    gen_join(l);
    // no upload may still read the pinned windows
    if (l->up_stream) hipStreamSynchronize(l->up_stream);
    if (l->ev_up) hipEventSynchronize(l->ev_up);
    for (float *q : l->pin)
        if (q) hipHostFree(q);
    if (l->ev_up) hipEventDestroy(l->ev_up);
    if (l->hio) hipHostFree(l->hio);
    for (float *q : l->d_stream) hipFree(q);
    if (l->hstage) hipHostFree(l->hstage);
    if (l->hsc) hipHostFree(l->hsc);
    hipFree(l->d_cells);
    hipFree(l->d_bvslot);
    hipFree(l->d_pn);
    hipFree(l->d_mw);
    hipFree(l->d_srec);
    hipFree(l->d_tab);
    if (l->ev_stage) hipEventDestroy(l->ev_stage);
    for (void *w : l->wbuf) hipFree(w);
    hipFree(l->d_bres);
    if (l->up_stream) hipStreamDestroy(l->up_stream);
    delete l;
}

// ------------------------------------------------------------------ host arithmetic
static int h_nint(float xf)
{
    double x = (double)xf;
    if (!(x >= -9.2233720368547758e18 && x < 9.2233720368547758e18)) return 0;
    double t = std::trunc(x), f = x - t;
    int64_t ti = (int64_t)t;
    if (f >= 0.5) ti += 1;
    else if (f <= -0.5) ti -= 1;
    return (int)(int32_t)(uint32_t)(uint64_t)ti;
}

static int h_ifix(float xf)
{
    double x = (double)xf;
    if (!(x > -2147483649.0 && x < 2147483648.0)) return (int)0x80000000u;
    return (int)std::trunc(x);
}

__host__ __device__ static inline int fold(int ix, int n)
{
    while (ix < 0) ix += n;
    while (ix >= n) ix -= n;
    return ix;
}

// grow-only pinned staging buffer of the ledger; a stream synchronisation first, since
// copies from the old buffer may still be in flight
static int ledger_stage(epsr_ctx *ctx, epsr_ledger *l, size_t bytes)
{
    if (bytes <= l->hstage_cap) return EPSR_OK;
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (l->hstage) HIP_CHECK(ctx, hipHostFree(l->hstage));
    l->hstage = nullptr;
    l->hstage_cap = 0;
    const size_t want = bytes + bytes / 8 + 4096;
    HIP_CHECK(ctx, hipHostMalloc(&l->hstage, want));
    l->hstage_cap = want;
    return EPSR_OK;
}

// grow-only device work buffer `i` of the ledger (kept between calls)
static int ledger_work(epsr_ctx *ctx, epsr_ledger *l, int i, size_t bytes, void **out)
{
    if (bytes > l->wcap[i]) {
        if (l->wbuf[i]) {
            HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
            HIP_CHECK(ctx, hipFree(l->wbuf[i]));
        }
        l->wbuf[i] = nullptr;
        l->wcap[i] = 0;
        const size_t want = bytes + bytes / 8 + 4096;
        HIP_CHECK(ctx, hipMalloc(&l->wbuf[i], want));
        l->wcap[i] = want;
    }
    *out = l->wbuf[i];
    return EPSR_OK;
}

enum { LW_CELL = 0, LW_ROW = 1, LW_TMP = 2, LW_REC = 3, LW_SUM = 4 };

// exclusive sums with the ledger's temporary storage
template <class T>
static int ledger_scan(epsr_ctx *ctx, epsr_ledger *l, const T *in, T *out, int n)
{
    size_t tb = 0;
    HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, n, ctx->stream));
    void *tmp = nullptr;
    int rc = ledger_work(ctx, l, LW_TMP, tb + 16, &tmp);
    if (rc) return rc;
    HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(tmp, tb, in, out, n, ctx->stream));
    return EPSR_OK;
}

// device half of the cell functions: the cell of every molecule (atomcellno,
// cellneighbourlist.py:317-330 with the cubic lattice: fc = z, fb = y - fc*0,
// fa = x - fb*0 - fc*0 in box fractions; each coordinate x - NINT(x), then
// IFIX((x1 + 0.5) n) folded into [0, n)), and the identity as the sort's values
__device__ static int d_cellcoord(float x, int n)
{
    // NINT from the double value with the host's 64-bit wrap (h_nint), IFIX (h_ifix)
    const double xd = (double)x;
    long long ti = 0;
    if (xd >= -9.2233720368547758e18 && xd < 9.2233720368547758e18) {
        const double t = trunc(xd), f = xd - t;
        ti = (long long)t;
        if (f >= 0.5) ti += 1;
        else if (f <= -0.5) ti -= 1;
    }
    const float x1 = x - (float)(int)(unsigned)(unsigned long long)ti;
    const double v = (double)((x1 + 0.5f) * (float)n);
    const int ix = (v > -2147483649.0 && v < 2147483648.0) ? (int)trunc(v) : (int)0x80000000u;
    return fold(ix, n);
}

__global__ void cell_assign(int nmol, const float *__restrict__ xadd, const float *__restrict__ yadd,
                            const float *__restrict__ zadd, int na, int nb, int nc, int *__restrict__ mcell,
                            int *__restrict__ ident)
{
    const int im = blockIdx.x * blockDim.x + threadIdx.x;
    if (im >= nmol) return;
    const float fc = zadd[im] / 1.0f;
    const float fb = (yadd[im] - fc * 0.0f) / 1.0f;
    const float fa = (xadd[im] - fb * 0.0f - fc * 0.0f) / 1.0f;
    mcell[im] = d_cellcoord(fa, na) + na * (d_cellcoord(fb, nb) + nb * d_cellcoord(fc, nc));
    ident[im] = im;
}

// from the molecules sorted by cell (stable, so each cell keeps increasing molecule
// numbers, as molsincell lists them): each cell's start in the chain (cstart, ncell + 1)
// and count (nmc), and every molecule's chain position
__global__ void cell_bounds(const int *__restrict__ skey, int nmol, int ncell, int *__restrict__ cstart)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nmol) return;
    const int lo = i == 0 ? 0 : skey[i - 1] + 1;
    const int hi = i == nmol ? ncell : skey[i];
    for (int cc = lo; cc <= hi; ++cc) cstart[cc] = i;
}

__global__ void cell_counts(const int *__restrict__ cstart, int ncell, const int *__restrict__ order, int nmol,
                            int *__restrict__ nmc, int *__restrict__ pos)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ncell) nmc[t] = cstart[t + 1] - cstart[t];
    if (t < nmol) pos[order[t]] = t;
}

extern "C" int epsr_cells_init(epsr_ctx *ctx, const epsr_box *box, epsr_ledger *l, const epsr_coords_t *c,
                               int nmolcellin, float rmax)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && l && c && nmolcellin > 0, "epsr_cells_init: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int nmol = box->nmol;
    hipStream_t st = ctx->stream;
    int rc;
    // initialise_cellneighbourlist (:23-96): the lattice depends on nmol alone
    int nmolcell = std::min(nmolcellin, nmol);
    int ncelldiv = h_nint(powf((float)nmol / (float)nmolcell, 1.0f / 3.0f));
    if (ncelldiv < 1) ncelldiv = 1;
    float cellsize = 2.0f * 1.0f / (float)ncelldiv;
    l->na = l->nb = l->nc = 2 * h_ifix(1.0f / cellsize) + 1;
    l->ncell = l->na * l->nb * l->nc;
    // makecellneighbourlist (:102-194); the offsets depend on the lattice and rmax only
    if (l->nbr_na != l->na || l->nbr_rmax != rmax) {
        float cella = 1.0f / (float)l->na, cellb = 1.0f / (float)l->nb, cellc = 1.0f / (float)l->nc;
        float cella2 = 0.5f * cella, cellb2 = 0.5f * cellb, cellc2 = 0.5f * cellc;
        float xc[8], yc[8], zc[8];
        int nf = 0;
        for (int ic = -1; ic <= 1; ic += 2)
            for (int ib = -1; ib <= 1; ib += 2)
                for (int ia = -1; ia <= 1; ia += 2) {
                    float cx = (float)ic * cellc2 * 0.0f, cy = (float)ic * cellc2 * 0.0f, cz = (float)ic * cellc2 * 1.0f;
                    float bx = (float)ib * cellb2 * 0.0f, by = (float)ib * cellb2 * 1.0f, bz = (float)ib * cellb2 * 0.0f;
                    float ax = (float)ia * cella2 * 1.0f, ay = (float)ia * cella2 * 0.0f, az = (float)ia * cella2 * 0.0f;
                    xc[nf] = ax + bx + cx; yc[nf] = ay + by + cy; zc[nf] = az + bz + cz;
                    nf++;
                }
        int nra = (l->na - 1) / 2, nrb = (l->nb - 1) / 2, nrc = (l->nc - 1) / 2;
        float rmaxsq = rmax * rmax;
        l->iaoff.clear(); l->iboff.clear(); l->icoff.clear();
        for (int ic = -nrc; ic <= nrc; ++ic) {
            float cx = (float)ic * cellc * 0.0f, cy = (float)ic * cellc * 0.0f, cz = (float)ic * cellc * 1.0f;
            for (int ib = -nrb; ib <= nrb; ++ib) {
                float bx = (float)ib * cellb * 0.0f, by = (float)ib * cellb * 1.0f, bz = (float)ib * cellb * 0.0f;
                for (int ia = -nra; ia <= nra; ++ia) {
                    float ax = (float)ia * cella * 1.0f, ay = (float)ia * cella * 0.0f, az = (float)ia * cella * 0.0f;
                    float xdc = ax + bx + cx, ydc = ay + by + cy, zdc = az + bz + cz;
                    bool found = false;
                    for (int i1 = 0; i1 < 8 && !found; ++i1)
                        for (int i2 = 0; i2 < 8 && !found; ++i2) {
                            float xd = xdc + xc[i1] - xc[i2], yd = ydc + yc[i1] - yc[i2], zd = zdc + zc[i1] - zc[i2];
                            float rsq = xd * xd + yd * yd + zd * zd;
                            found = (rsq < rmaxsq) || (ia == 0 && ib == 0 && ic == 0);
                        }
                    if (found) {
                        l->iaoff.push_back(ia); l->iboff.push_back(ib); l->icoff.push_back(ic);
                    }
                }
            }
        }
        l->ncellneighb = (int)l->iaoff.size();
        l->nbr_na = l->na;
        l->nbr_rmax = rmax;
        l->off_uploaded = false;
    }
    // the cells of the molecules, the chain order (cells in order, molecules in cell
    // order) and the chain positions, on the device; the neighbour offsets go up when
    // they change (through the pinned staging buffer, whose previous upload this
    // stream has ordered before)
    const int nbr = l->ncellneighb, ncell = l->ncell;
    const size_t ncw = 3 * (size_t)nbr + 2 * (size_t)ncell + 2 + (size_t)nmol;
    if (l->cap_cells < ncw) {
        const size_t cap = ncw + ncw / 8 + 64;
        if ((rc = dev_alloc(ctx, &l->d_cells, cap))) return rc;
        l->cap_cells = cap;
        l->off_uploaded = false;
    }
    if (!l->off_uploaded) {
        if ((rc = ledger_stage(ctx, l, sizeof(int) * 3 * (size_t)nbr))) return rc;
        if (l->ev_stage) HIP_CHECK(ctx, hipEventSynchronize(l->ev_stage));
        int *hu = (int *)l->hstage;
        std::copy(l->iaoff.begin(), l->iaoff.end(), hu);
        std::copy(l->iboff.begin(), l->iboff.end(), hu + nbr);
        std::copy(l->icoff.begin(), l->icoff.end(), hu + 2 * nbr);
        HIP_CHECK(ctx, hipMemcpyAsync(l->d_cells, hu, sizeof(int) * 3 * nbr, hipMemcpyHostToDevice, st));
        if (!l->ev_stage) HIP_CHECK(ctx, hipEventCreateWithFlags(&l->ev_stage, hipEventDisableTiming));
        HIP_CHECK(ctx, hipEventRecord(l->ev_stage, st));
        l->off_uploaded = true;
    }
    int *dnmc = l->d_cells + 3 * nbr, *dcst = dnmc + ncell, *dmcell = dcst + ncell + 2;
    int *kv = nullptr;
    if ((rc = ledger_work(ctx, l, LW_REC, sizeof(int) * 2 * ((size_t)nmol + 1), (void **)&kv))) return rc;
    int *ident = kv, *skey = kv + nmol + 1;
    cell_assign<<<ceil_div(nmol, 256), 256, 0, st>>>(nmol, c->xadd, c->yadd, c->zadd, l->na, l->nb, l->nc, dmcell,
                                                     ident);
    {
        int bits = 1;
        while ((1 << bits) < ncell) ++bits;
        size_t tbytes = 0;
        HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, dmcell, skey, ident, l->d_order, nmol, 0,
                                                         bits, st));
        void *tmpb = nullptr;
        if ((rc = ledger_work(ctx, l, LW_TMP, tbytes + 16, &tmpb))) return rc;
        HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(tmpb, tbytes, dmcell, skey, ident, l->d_order, nmol, 0, bits,
                                                         st));
    }
    cell_bounds<<<ceil_div(nmol + 1, 256), 256, 0, st>>>(skey, nmol, ncell, dcst);
    cell_counts<<<ceil_div(std::max(ncell, nmol), 256), 256, 0, st>>>(dcst, ncell, l->d_order, nmol, dnmc, l->d_pos);
    HIP_CHECK(ctx, hipGetLastError());
    l->nnbrefmax = 0;   // get_nnbrefmax: formed on the GPU by epsr_fmolec1
    l->cells_ready = true;
    l->ledger_ready = false;
    return EPSR_OK;
}

// ------------------------------------------------------------------ pair energy
// T (near bins) then F (far bins) in one buffer
__global__ void pack_tables(const float *__restrict__ xavs, const float *__restrict__ xavs1,
                            const float *__restrict__ ecoul, const float *__restrict__ ecoul1, int ld, int npairs,
                            int nte, int nnear, float4 *__restrict__ T)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nfar = nte - nnear;
    if (t < npairs * nnear) {
        int ic = t / nnear, iz = t - ic * nnear;
        size_t o = (size_t)ic * ld + iz;
        T[t] = make_float4(xavs[o], xavs1[o], ecoul[o], ecoul1[o]);
        return;
    }
    t -= npairs * nnear;
    if (t >= npairs * nfar) return;
    int ic = t / nfar, iz = nnear + (t - ic * nfar);
    size_t o = (size_t)ic * ld + iz;
    ((float2 *)(T + (size_t)npairs * nnear))[t] = make_float2(ecoul[o], ecoul1[o]);
}

// out[0]: the last bin (up to lim) where any of the four tables has a non-zero value;
// out[1]: the same for xavs and xavs1
__global__ void last_nonzero(const float *__restrict__ a, const float *__restrict__ b, const float *__restrict__ c,
                             const float *__restrict__ d, int ld, int npairs, int lim, int *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * ld) return;
    int iz = t % ld;
    if (iz > lim) return;
    // by value: a bin whose four entries are all +0 or -0 adds exactly nothing to a pair sum
    // (the lookup gives +0 or -0 there, and a sum that starts at +0 is never -0, so adding
    // either leaves it unchanged), the same as a bin beyond the packed tables; likewise
    // a -0 of xavs or xavs1 read as the +0 of the far part only changes the sign of a
    // zero lookup
    const bool ab = a[t] != 0.0f || b[t] != 0.0f;
    if (ab) atomicMax(out + 1, iz);
    if (ab || c[t] != 0.0f || d[t] != 0.0f) atomicMax(out, iz);
}

// absolute atom positions (x = xa + xadd, one fp32 add) with the type in .w
__global__ void abs_positions(const float *__restrict__ xa, const float *__restrict__ ya, const float *__restrict__ za,
                              const float *__restrict__ xadd, const float *__restrict__ yadd,
                              const float *__restrict__ zadd, const int *__restrict__ amol,
                              const int *__restrict__ itype, float4 *__restrict__ P, int natom)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    int m = amol[a];
    P[a] = make_float4(xa[a] + xadd[m], ya[a] + yadd[m], za[a] + zadd[m], __int_as_float(itype[a]));
}

// abs_positions of the current (Pold) and trial (Pnew) coordinates in one launch
__global__ void __launch_bounds__(256)
abs_positions2(int natom, const float *__restrict__ xa, const float *__restrict__ ya, const float *__restrict__ za,
               const float *__restrict__ xadd, const float *__restrict__ yadd, const float *__restrict__ zadd,
               const float *__restrict__ txa, const float *__restrict__ tya, const float *__restrict__ tza,
               const float *__restrict__ txadd, const float *__restrict__ tyadd, const float *__restrict__ tzadd,
               const int *__restrict__ amol, const int *__restrict__ itype, float4 *__restrict__ Pold,
               float4 *__restrict__ Pnew)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 2 * natom) {
        const bool trial = t >= natom;
        const int a = trial ? t - natom : t;
        const int m = amol[a];
        const float4 p = trial ? make_float4(txa[a] + txadd[m], tya[a] + tyadd[m], tza[a] + tzadd[m],
                                             __int_as_float(itype[a]))
                               : make_float4(xa[a] + xadd[m], ya[a] + yadd[m], za[a] + zadd[m],
                                             __int_as_float(itype[a]));
        (trial ? Pnew : Pold)[a] = p;
    }
}


// the last bin (up to npt2) where any of the four tables is non-zero, into dlast[0], and
// where xavs or xavs1 is, into dlast[1]
static int tables_extent(epsr_ctx *ctx, const epsr_box *box, const epsr_tables_t *tb, int *dlast)
{
    hipStream_t st = ctx->stream;
    const int np = box->npairs, ld = tb->ld;
    const int lim = std::min(tb->npt2, ld - 1);
    HIP_CHECK(ctx, hipMemsetAsync(dlast, 0, 2 * sizeof(int), st));
    last_nonzero<<<ceil_div((long)np * ld, 256), 256, 0, st>>>(tb->xavs, tb->xavs1, tb->ecoul, tb->ecoul1, ld, np, lim,
                                                               dlast);
    return EPSR_OK;
}

// the description of tables packed in dT up to bin `last`, four-wide up to bin `lastnear`
static void tables_describe(const epsr_box *box, const epsr_tables_t *tb, int last, int lastnear, float4 *dT,
                            PreparedTables &pt)
{
    const int np = box->npairs;
    const int nte = std::max(2, last + 1);
    const int nnear = std::min(nte, std::max(2, lastnear + 1));
    pt.tab.T = dT;
    pt.tab.F = (const float2 *)(dT + (size_t)np * nnear);
    pt.tab.nnear = nnear;
    pt.tab.xs = tb->xstar;
    pt.tab.npr = box->d_npairref;
    pt.tab.nte = nte;
    pt.tab.npt2 = tb->npt2;
    pt.tab.ntype = box->ntype;
    pt.tab.stepcl = box->cellst / box->dmax;
    pt.tab.inv_stepcl = 1.0f / pt.tab.stepcl;
    pt.lds_bytes = sizeof(float4) * (size_t)np * nnear + sizeof(float2) * (size_t)np * (nte - nnear) +
                   sizeof(float) * nte + sizeof(int) * box->ntype * box->ntype + 64;
    pt.in_lds = pt.lds_bytes <= 150 * 1024;
}

// pack the tables up to bin `last` into dT, four-wide up to bin `lastnear`
static int tables_pack(epsr_ctx *ctx, const epsr_box *box, const epsr_tables_t *tb, int last, int lastnear, float4 *dT,
                       PreparedTables &pt)
{
    tables_describe(box, tb, last, lastnear, dT, pt);
    const int np = box->npairs;
    const long items = (long)np * pt.tab.nte;
    pack_tables<<<ceil_div(items, 256), 256, 0, ctx->stream>>>(tb->xavs, tb->xavs1, tb->ecoul, tb->ecoul1, tb->ld, np,
                                                              pt.tab.nte, pt.tab.nnear, dT);
    return EPSR_OK;
}

// dscratch_int: two ints of device memory; the extents go to extents_out if given
static int prepare_tables(epsr_ctx *ctx, const epsr_box *box, const epsr_tables_t *tb, float4 *dT, int *dscratch_int,
                          PreparedTables &pt, int *extents_out = nullptr)
{
    int rc = tables_extent(ctx, box, tb, dscratch_int);
    if (rc) return rc;
    int last[2] = {0, 0};
    HIP_CHECK(ctx, hipMemcpyAsync(last, dscratch_int, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (extents_out) {
        extents_out[0] = last[0];
        extents_out[1] = last[1];
    }
    return tables_pack(ctx, box, tb, last[0], last[1], dT, pt);
}

int mc_prepare_tables(epsr_ctx *ctx, const epsr_box *box, const epsr_tables_t *tb, float4 *dT, int *dscratch_int,
                      PreparedTables &pt)
{
    return prepare_tables(ctx, box, tb, dT, dscratch_int, pt);
}

// the table extents measured for the arrays of tb (by fmolec1 or a cycle), kept in the
// ledger so that MC cycles on the same arrays need not wait for them
static void tab_remember(epsr_ledger *l, const epsr_tables_t *tb, int last, int lastnear)
{
    l->tab_ptr[0] = tb->xstar;
    l->tab_ptr[1] = tb->xavs;
    l->tab_ptr[2] = tb->xavs1;
    l->tab_ptr[3] = tb->ecoul;
    l->tab_ptr[4] = tb->ecoul1;
    l->tab_ld = tb->ld;
    l->tab_npt2 = tb->npt2;
    l->tab_last = last;
    l->tab_lastnear = lastnear;
    l->tab_ok = true;
    l->tab_packed = false;
}

// the ledger's buffer for packed tables, large enough for tb
static int ledger_tables(epsr_ctx *ctx, const epsr_box *box, const epsr_tables_t *tb, epsr_ledger *l)
{
    const size_t need = (size_t)box->npairs * std::max(2, std::min(tb->ld, tb->npt2 + 1));
    l->tab_packed = false;
    if (l->cap_tab < need) {
        HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        HIP_CHECK(ctx, hipFree(l->d_tab));
        l->d_tab = nullptr;
        l->cap_tab = 0;
        HIP_CHECK(ctx, hipMalloc((void **)&l->d_tab, sizeof(float4) * need));
        l->cap_tab = need;
    }
    return EPSR_OK;
}

static bool tab_known(const epsr_ledger *l, const epsr_tables_t *tb)
{
    return l->tab_ok && l->tab_ptr[0] == tb->xstar && l->tab_ptr[1] == tb->xavs && l->tab_ptr[2] == tb->xavs1 &&
           l->tab_ptr[3] == tb->ecoul && l->tab_ptr[4] == tb->ecoul1 && l->tab_ld == tb->ld &&
           l->tab_npt2 == tb->npt2;
}

// bit 30 of the cycle's no-change count (a count of moves, far below 2^30) marks table
// extents measured on the device beyond the ones the cycle packed
#define TABLES_STALE (1 << 30)

// a packing of the tables to verify against the arrays it came from
struct TabCheck {
    const float *a, *b, *c, *d;   // xavs, xavs1, ecoul, ecoul1, [np][ld]
    const float4 *T;              // the packing (null: nothing to verify)
    int ld, np, lim, nte, nnear;
};

// any entry of the arrays (bins up to lim) that the packing does not reproduce sets
// TABLES_STALE in *flag: a changed bit pattern within the packed extents, or a non-zero
// value where the packing keeps none
__device__ void tables_verify(const TabCheck &tc, int *flag)
{
    if (!tc.T) return;
    const float2 *F = (const float2 *)(tc.T + (size_t)tc.np * tc.nnear);
    const int nfar = tc.nte - tc.nnear;
    bool bad = false;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tc.np * tc.ld; t += gridDim.x * blockDim.x) {
        const int ic = t / tc.ld, iz = t - ic * tc.ld;
        if (iz > tc.lim) continue;
        const int a = __float_as_int(tc.a[t]), b = __float_as_int(tc.b[t]);
        const int c = __float_as_int(tc.c[t]), d = __float_as_int(tc.d[t]);
        if (iz < tc.nnear) {
            const float4 v = tc.T[ic * tc.nnear + iz];
            bad |= a != __float_as_int(v.x) || b != __float_as_int(v.y) || c != __float_as_int(v.z) ||
                   d != __float_as_int(v.w);
        } else if (iz < tc.nte) {
            // (zeros of either sign stand for the +0 of the far part, as the packing measures)
            const float2 f = F[ic * nfar + (iz - tc.nnear)];
            bad |= tc.a[t] != 0.0f || tc.b[t] != 0.0f || c != __float_as_int(f.x) || d != __float_as_int(f.y);
        } else {
            bad |= tc.a[t] != 0.0f || tc.b[t] != 0.0f || tc.c[t] != 0.0f || tc.d[t] != 0.0f;
        }
    }
    if (bad) atomicOr(flag, TABLES_STALE);
}

// copy the tables into LDS; returns a PairTab pointing at LDS
__device__ PairTab tables_to_lds(const PairTab &g, int npairs, char *smem)
{
    float4 *T = (float4 *)smem;
    const int nT = npairs * g.nnear, nF = npairs * (g.nte - g.nnear);
    float2 *F = (float2 *)(T + nT);
    float *xs = (float *)(F + nF);
    int *npr = (int *)(xs + g.nte);
    for (int i = threadIdx.x; i < nT; i += blockDim.x) T[i] = g.T[i];
    for (int i = threadIdx.x; i < nF; i += blockDim.x) F[i] = g.F[i];
    for (int i = threadIdx.x; i < g.nte; i += blockDim.x) xs[i] = g.xs[i];
    for (int i = threadIdx.x; i < g.ntype * g.ntype; i += blockDim.x) npr[i] = g.npr[i];
    __syncthreads();
    PairTab t = g;
    t.T = T;
    t.F = F;
    t.xs = xs;
    t.npr = npr;
    return t;
}

// ------------------------------------------------------------------ fmolec1 kernels
// F1: for every chain position k (molecule im) and every position inn >= 1 of its cell's
// neighbour list with jm > im: E = en_mole_pair(im, jm) (ato_manipulation_rev4.py:2454-2462).
// One wave per chain row; the non-zero energies of a row and their partners are packed to
// the front of the row's item range (rowoff, int64 prefix over k) in neighbour order, and
// their number goes to rowcnt[k].
// position of the r-th set bit (0-based) of m, for r < popcount(m)
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int r)
{
    int pos = 0, c = __popc((unsigned)m);
    if (r >= c) { r -= c; m >>= 32; pos += 32; }
    c = __popc((unsigned)m & 0xffffu);
    if (r >= c) { r -= c; m >>= 16; pos += 16; }
    c = __popc((unsigned)m & 0xffu);
    if (r >= c) { r -= c; m >>= 8; pos += 8; }
    c = __popc((unsigned)m & 0xfu);
    if (r >= c) { r -= c; m >>= 4; pos += 4; }
    c = __popc((unsigned)m & 0x3u);
    if (r >= c) { r -= c; m >>= 2; pos += 2; }
    if (r >= (int)(m & 1u)) pos += 1;
    return pos;
}

// (8 waves per SIMD: two 1024-lane workgroups per CU when their tables fit LDS twice)
// NQ: candidate queues, one per partner molecule type when the box has at most NQ types
// (NQ = 1: one queue for all)
template <bool LDS, int NQ>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 1024), amdgpu_waves_per_eu(NQ == 1 ? 8 : 4)))
fmolec1_pairs(const long long *__restrict__ rowoff, int nmol, const int *__restrict__ order,
              const int *__restrict__ mcell, const int *__restrict__ cstart, const int *__restrict__ nbflat,
              const float4 *__restrict__ P, const int *__restrict__ naf, const int *__restrict__ nal,
              const int *__restrict__ mtype, PairTab gt, int npairs, float *__restrict__ rowE,
              int *__restrict__ rowJ, int *__restrict__ rowcnt)
{
    extern __shared__ char smem[];
    PairTab t = LDS ? tables_to_lds(gt, npairs, smem) : gt;
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * (blockDim.x >> 6);
    const unsigned long long below = (1ull << lane) - 1ull;   // lanes before this one
    for (int k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < nmol; k += nw) {
        const long long r0 = rowoff[k];
        const int len = (int)(rowoff[k + 1] - r0);
        const int im = order[k];
        const int *nb = nbflat + cstart[mcell[im]];
        const float4 *P1 = P + naf[im];
        const int n1 = nal[im] - naf[im] + 1;
        // Only candidates later in the numbering are evaluated (half of them).  Each gets its
        // index ci among the evaluated candidates, in candidate order, and joins the queue of
        // its molecule type (one partner per lane); a full queue is evaluated 64 at a time,
        // so the lanes of a batch meet partners of one size and none idles on an earlier
        // partner.  The energies are staged at their index ci, then the non-zero ones are
        // compacted in candidate order.
        int qj[NQ], qi[NQ], qn[NQ];   // lane's queued partner and its index; queue lengths
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            qj[c] = 0;
            qi[c] = 0;
            qn[c] = 0;
        }
        int nev = 0;    // evaluated candidates so far (wave-uniform)
        int base = 0;   // stored entries (one queue: written in order at once)
        auto flush = [&](int c, int nvalid) {
            float e = 0.0f;
            const int j = qj[c];
            if (lane < nvalid) e = mol_pair_any(P1, n1, P + naf[j], nal[j] - naf[j] + 1, t);
            if constexpr (NQ == 1) {
                const bool keep = e != 0.0f;
                const unsigned long long mask = __ballot(keep);
                if (keep) {
                    const int at = base + __popcll(mask & below);
                    rowE[r0 + at] = e;
                    rowJ[r0 + at] = j;
                }
                base += __popcll(mask);
            } else if (lane < nvalid) {
                rowE[r0 + qi[c]] = e;
                rowJ[r0 + qi[c]] = j;
            }
        };
        for (int i0 = 0; i0 < len; i0 += 64) {   // wave-uniform trip count
            const int inn = i0 + lane;
            int jm = 0;
            bool work = false;
            if (inn < len) {
                jm = nb[inn];
                work = inn >= 1 && jm > im;
            }
            const unsigned long long wm = __ballot(work);
            const int ci = NQ == 1 ? 0 : nev + __popcll(wm & below);
            nev += __popcll(wm);
            const int cls = (NQ == 1 || !work) ? 0 : mtype[jm];
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                const unsigned long long cm = NQ == 1 ? wm : __ballot(work && cls == c);
                const int cnt = __popcll(cm);
                if (cnt == 0) continue;
                // queue lanes qn .. qn + cnt - 1 (those below 64) take the class's candidates in order
                const int r = lane - qn[c];
                const bool take = r >= 0 && r < cnt;
                const int src = take ? nth_set_bit(cm, r) : lane;
                const int vj = __shfl(jm, src), vi = NQ == 1 ? 0 : __shfl(ci, src);
                if (take) {
                    qj[c] = vj;
                    qi[c] = vi;
                }
                if (qn[c] + cnt >= 64) {
                    flush(c, 64);
                    const int r2 = lane + 64 - qn[c];   // the rest: ranks 64 - qn .. cnt - 1
                    const bool rest = r2 < cnt;
                    const int src2 = rest ? nth_set_bit(cm, r2) : lane;
                    const int vj2 = __shfl(jm, src2), vi2 = NQ == 1 ? 0 : __shfl(ci, src2);
                    qj[c] = rest ? vj2 : 0;
                    qi[c] = rest ? vi2 : 0;
                    qn[c] = qn[c] + cnt - 64;
                } else {
                    qn[c] += cnt;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c)
            if (qn[c] > 0) flush(c, qn[c]);
        if constexpr (NQ > 1) {
        // the staged entries written by other lanes of the wave, then compacted in order
        // (a chunk is read whole before it is written to entries at or before it)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (int i0 = 0; i0 < nev; i0 += 64) {
            const int idx = i0 + lane;
            float e = 0.0f;
            int j = 0;
            if (idx < nev) {
                e = rowE[r0 + idx];
                j = rowJ[r0 + idx];
            }
            const bool keep = idx < nev && e != 0.0f;
            const unsigned long long mask = __ballot(keep);
            if (keep) {
                const int at = base + __popcll(mask & below);
                rowE[r0 + at] = e;
                rowJ[r0 + at] = j;
            }
            base += __popcll(mask);
        }
        }
        if (lane == 0) rowcnt[k] = base;
    }
}

// cat[m] = (eintra[m] != 0) for the self entries, cat[2 nmol] = 0 (rows in between come
// from fmolec1_pairs); its exclusive scan numbers the ledger entries: self entries in
// molecule order, then the stored pairs in (chain position, neighbour position) order
__global__ void fmolec1_self_flags(const float *__restrict__ eintra, int nmol, int *__restrict__ cat)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nmol) cat[t] = eintra[t] != 0.0f;
    if (t == nmol) cat[2 * nmol] = 0;
}

// records of the per-molecule lists and the stored energies.  ent is the scan above:
// entry e < nself = ent[nmol] is a self entry, e = nself + p the pair p.  Record r < nself
// is the self entry r (key: its molecule's chain position); pair p gives records
// nself + 2p (key: the row's chain position) and nself + 2p + 1 (key: the partner's);
// val = r.  Keyed by chain position, the lists are laid out in chain order, so the lists
// of neighbouring molecules (and the entries they share) lie close together in memory.
__global__ void fmolec1_entries(int nmol, const int *__restrict__ order, const int *__restrict__ pos,
                                const long long *__restrict__ rowoff, const float *__restrict__ rowE,
                                const int *__restrict__ rowJ, const int *__restrict__ ent,
                                const float *__restrict__ eintra, int *__restrict__ key, int *__restrict__ val,
                                float *__restrict__ ergpair)
{
    const int nself = ent[nmol];
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < nmol; m += gridDim.x * blockDim.x) {
        const float e = eintra[m];
        if (e != 0.0f) {
            const int r = ent[m];
            key[r] = pos[m];
            val[r] = r;
            ergpair[r] = e;
        }
    }
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * (blockDim.x >> 6);
    for (int k = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); k < nmol; k += nw) {
        const int e0 = ent[nmol + k], c = ent[nmol + k + 1] - e0;
        const long long r0 = rowoff[k];
        for (int j = lane; j < c; j += 64) {
            const int e = e0 + j, r = nself + 2 * (e - nself);
            key[r] = k;
            val[r] = r;
            key[r + 1] = pos[rowJ[r0 + j]];
            val[r + 1] = r + 1;
            ergpair[e] = rowE[r0 + j];
        }
    }
}

// list starts by chain position from the records sorted by chain position: rowk[k] = first
// record of position k (the next position's for an empty list), rowk[nmol] = nrec
__global__ void fmolec1_rowstart(const int *__restrict__ skey, int nrec, int nmol, int *__restrict__ rowk)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nrec) return;
    const int lo = i == 0 ? 0 : skey[i - 1] + 1;
    const int hi = i == nrec ? nmol : skey[i];
    for (int k = lo; k <= hi; ++k) rowk[k] = i;
}

// the lists by molecule (rowstart[m], cnt[m]; rowstart[nmol] = nrec) and the longest list
__global__ void fmolec1_counts(const int *__restrict__ rowk, const int *__restrict__ pos, int nmol,
                               int *__restrict__ rowstart, int *__restrict__ cnt, int *__restrict__ maxcnt)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int v = 0;
    if (t < nmol) {
        const int k = pos[t];
        v = rowk[k + 1] - rowk[k];
        rowstart[t] = rowk[k];
        cnt[t] = v;
    }
    if (t == nmol) rowstart[nmol] = rowk[nmol];
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_down(v, off));
    if ((threadIdx.x & 63) == 0) atomicMax(maxcnt, v);
}

__global__ void fmolec1_rows(const int *__restrict__ skey, const int *__restrict__ sval, int nrec, int nself,
                             const int *__restrict__ key, const int *__restrict__ order, int *__restrict__ partner,
                             int *__restrict__ entry, int *__restrict__ slotmol, int *__restrict__ sortedpos)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    const int m = order[skey[i]], r = sval[i];
    slotmol[i] = m;
    sortedpos[r] = i;
    if (r < nself) {
        partner[i] = m;
        entry[i] = r;
    } else {
        const int q = r - nself;
        partner[i] = order[key[nself + (q ^ 1)]];   // the other record of the pair holds the partner
        entry[i] = nself + (q >> 1);
    }
}

__global__ void fmolec1_mirror(const int *__restrict__ sval, int nrec, int nself, const int *__restrict__ sortedpos,
                               int *__restrict__ mirror)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    int r = sval[i];
    mirror[i] = r < nself ? -1 : sortedpos[nself + ((r - nself) ^ 1)];
}

// slots (past the first of their list) whose partner comes earlier in the chain: the MC
// cycle evaluates their pair at both new positions
__global__ void fmolec1_bv_flags(int nrec, const int *__restrict__ slotmol, const int *__restrict__ partner,
                                 const int *__restrict__ rowstart, const int *__restrict__ pos, int *__restrict__ flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nrec) return;
    if (i == nrec) {
        flag[i] = 0;
        return;
    }
    const int m = slotmol[i];
    flag[i] = i != rowstart[m] && pos[partner[i]] < pos[m];
}

__global__ void fmolec1_bv_list(int nrec, const int *__restrict__ flag, const int *__restrict__ idx,
                                int *__restrict__ bvslot, int *__restrict__ nbv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrec && flag[i]) bvslot[idx[i]] = i;
    if (i == nrec) *nbv = idx[nrec];
}

// per slot, what the MC cycle's kernels would otherwise gather: the partner's chain
// position (pn), the mirror slot unless it is the partner's first slot (mw, else -1), and
// srec = (molecule | first-slot bit 31, partner, first atom | atom count << 24 of both)
__global__ void fmolec1_slot_aux(int nrec, const int *__restrict__ slotmol, const int *__restrict__ partner,
                                 const int *__restrict__ mirror, const int *__restrict__ rowstart,
                                 const int *__restrict__ pos, const int *__restrict__ naf, const int *__restrict__ nal,
                                 int *__restrict__ pn, int *__restrict__ mw, int4 *__restrict__ srec)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    const int m = slotmol[i], n = partner[i], mi = mirror[i];
    pn[i] = pos[n];
    mw[i] = (mi >= 0 && mi != rowstart[n]) ? mi : -1;
    srec[i] = make_int4(m | (i == rowstart[m] ? (int)0x80000000u : 0), n, naf[m] | ((nal[m] - naf[m] + 1) << 24),
                        naf[n] | ((nal[n] - naf[n] + 1) << 24));
}

__global__ void max_int(const int *__restrict__ a, int n, int *__restrict__ out)
{
    int v = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v = max(v, a[i]);
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_down(v, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

// FP64 sums of two float arrays in a fixed order: per-block partials (blockIdx.y picks
// the array), then one block per array adds the partials
__global__ void __launch_bounds__(256)
sum_f64_parts(const float *__restrict__ a, int na, const float *__restrict__ b, int nb, double *__restrict__ part)
{
    __shared__ double s[256];
    const float *x = blockIdx.y ? b : a;
    const int n = blockIdx.y ? nb : na;
    double v = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v += (double)x[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.y * gridDim.x + blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) sum_f64_final(const double *__restrict__ part, int nparts, double *__restrict__ out)
{
    __shared__ double s[256];
    double v = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 256) v += part[blockIdx.x * nparts + i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// grid size for persistent loops
// a grid that fills every CU: as many workgroups of `block` threads per CU as the LDS
// (160 KB) and 32 waves allow, and no more than the items need
static int persistent_blocks(size_t lds_bytes, long long items, int block)
{
    int per_cu = 32 / std::max(1, block / 64);
    if (lds_bytes > 0) per_cu = std::min<int>(per_cu, (int)((160 * 1024) / lds_bytes));
    per_cu = std::max(1, per_cu);
    long long want = (items + block - 1) / block;
    return (int)std::max(1LL, std::min<long long>(want, 256LL * per_cu));
}

struct Scratch {
    char *base = nullptr;
    size_t off = 0;
    template <class T>
    T *take(size_t n)
    {
        size_t o = off;
        off += (sizeof(T) * (n > 0 ? n : 1) + 255) & ~(size_t)255;
        return (T *)(base + o);
    }
};

// neighbour cells of every cell (cellneighbourlist.py:199-219 order) and their counts
__global__ void cell_neighbours(int ncell, int nbr, int na, int nb, int nc, const int *__restrict__ iaoff,
                                const int *__restrict__ iboff, const int *__restrict__ icoff,
                                const int *__restrict__ nmolsincell, int *__restrict__ nbcell,
                                int *__restrict__ nbcnt)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ncell * nbr) return;
    int cc = (int)(t / nbr), kk = (int)(t - (long long)cc * nbr);
    int icc = cc / (na * nb), ib = (cc - icc * na * nb) / na, ia = cc - na * (ib + nb * icc);
    int cref = fold(ia + iaoff[kk], na) + na * (fold(ib + iboff[kk], nb) + nb * fold(icc + icoff[kk], nc));
    nbcell[t] = cref;
    nbcnt[t] = nmolsincell[cref];
}

// per cell: exclusive prefix of the neighbour-cell counts (nbpre, nbr + 1 per cell), the
// length of the cell's neighbour molecule list (rowlen) and of its flattened copy
// (flatlen, 0 for an empty cell); one block per cell
__global__ void __launch_bounds__(256)
cell_prefix(int nbr, const int *__restrict__ nbcnt, const int *__restrict__ nmolsincell, int *__restrict__ nbpre,
            int *__restrict__ rowlen, int *__restrict__ flatlen)
{
    __shared__ int sw[4];
    const int cc = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int *cnt = nbcnt + (size_t)cc * nbr;
    int *pre = nbpre + (size_t)cc * (nbr + 1);
    int carry = 0;
    for (int k0 = 0; k0 < nbr; k0 += 256) {
        const int k = k0 + threadIdx.x;
        int v = k < nbr ? cnt[k] : 0, x = v;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) sw[wv] = x;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wv; ++w) before += sw[w];
        if (k < nbr) pre[k] = before + x - v;
        carry += sw[0] + sw[1] + sw[2] + sw[3];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pre[nbr] = carry;
        rowlen[cc] = carry;
        flatlen[cc] = nmolsincell[cc] > 0 ? carry : 0;
    }
}

// flattened neighbour molecule lists of the non-empty cells
// (the molecules of a cell are the chain order's run from its start, cellstart)
__global__ void cell_flatten(int ncell, int nbr, const int *__restrict__ nbcell, const int *__restrict__ nbcnt,
                             const int *__restrict__ nbpre, const int *__restrict__ cstart,
                             const int *__restrict__ nmolsincell, const int *__restrict__ order,
                             const int *__restrict__ cellstart, int *__restrict__ nbflat)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ncell * nbr) return;
    int cc = (int)(t / nbr), kk = (int)(t - (long long)cc * nbr);
    if (nmolsincell[cc] == 0) return;
    int *dst = nbflat + cstart[cc] + nbpre[(size_t)cc * (nbr + 1) + kk];
    const int *src = order + cellstart[nbcell[t]];
    for (int j = 0, n = nbcnt[t]; j < n; ++j) dst[j] = src[j];
}

// row length of every chain position (its cell's list), as int64 for the scan
__global__ void chain_rowlen(int nmol, const int *__restrict__ order, const int *__restrict__ mcell,
                             const int *__restrict__ rowlen, long long *__restrict__ out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nmol) out[k] = rowlen[mcell[order[k]]];
    if (k == nmol) out[k] = 0;
}

// the host-side scalars of the list set-up: candidate items (clamped to INT_MAX), flat
// list length, longest list, last non-zero table bins (all four tables, xavs/xavs1)
__global__ void fmolec1_report(const long long *__restrict__ total, const int *__restrict__ nflat,
                               const int *__restrict__ tail, int *__restrict__ rep)
{
    const long long t = *total;
    rep[0] = t >= 0x7fffffffLL ? 0x7fffffff : (int)t;
    rep[1] = *nflat;
    rep[2] = tail[0];
    rep[3] = tail[1];
    rep[4] = tail[2];
}

__global__ void copy_two(const int *__restrict__ a, const int *__restrict__ b, int *__restrict__ out)
{
    out[0] = *a;
    out[1] = *b;
}

extern "C" int epsr_fmolec1(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                            epsr_ledger *l, const float *eintra, double *ftest64, double *ftestb64, float *ftest,
                            float *ftestb, int *nnbrefmax_out, int *nergpair_out)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && tb && l && eintra, "epsr_fmolec1: null argument");
    ARG_CHECK(ctx, l->cells_ready, "epsr_fmolec1: call epsr_cells_init first");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int nmol = box->nmol;
    int rc;
    // the slot records pack a first atom in 24 bits and an atom count in 8
    ARG_CHECK(ctx, box->natom < (1 << 24) && box->max_atoms_per_mol < 256,
              "epsr_fmolec1: %d atoms (limit 2^24) or %d atoms in a molecule (limit 255)", box->natom,
              box->max_atoms_per_mol);
    if (!l->hsc) HIP_CHECK(ctx, hipHostMalloc((void **)&l->hsc, 64));
    int *hsc = l->hsc;
    // neighbour lists on the GPU from the cell arrays cells_init uploaded: per cell the
    // neighbour cells and their prefix counts, the flattened molecule list of every
    // non-empty cell and the row offsets of the (chain position, neighbour position) items
    const int ncell = l->ncell, nbr = l->ncellneighb;
    const long long ncn = (long long)ncell * nbr;
    const size_t cw_ints = 2 * (size_t)ncn + (size_t)ncell * (nbr + 1) + 3 * ((size_t)ncell + 1) + 16;
    int *cw = nullptr;
    long long *drow = nullptr;
    if ((rc = ledger_work(ctx, l, LW_CELL, sizeof(int) * cw_ints, (void **)&cw))) return rc;
    if ((rc = ledger_work(ctx, l, LW_ROW, sizeof(long long) * 2 * ((size_t)nmol + 1), (void **)&drow))) return rc;
    long long *dlen = drow + nmol + 1;
    const int *doff = l->d_cells, *dnmc = doff + 3 * nbr, *dcst = dnmc + ncell, *dmcell = dcst + ncell + 2;
    int *dnbcell = cw, *dnbcnt = dnbcell + ncn, *dnbpre = dnbcnt + ncn;
    int *drowlen = dnbpre + (size_t)ncell * (nbr + 1), *dflatlen = drowlen + ncell + 1, *dcstart = dflatlen + ncell + 1;
    // [0] longest list, [1..2] table extents, [4..8] report, [10..11] entry counts
    int *dtail = dcstart + ncell + 1;
    cell_neighbours<<<(unsigned)ceil_div(ncn, 256), 256, 0, st>>>(ncell, nbr, l->na, l->nb, l->nc, doff, doff + nbr,
                                                                  doff + 2 * nbr, dnmc, dnbcell, dnbcnt);
    cell_prefix<<<ncell, 256, 0, st>>>(nbr, dnbcnt, dnmc, dnbpre, drowlen, dflatlen);
    HIP_CHECK(ctx, hipMemsetAsync(dflatlen + ncell, 0, sizeof(int), st));
    HIP_CHECK(ctx, hipMemsetAsync(dtail, 0, sizeof(int), st));
    if ((rc = ledger_scan<int>(ctx, l, dflatlen, dcstart, ncell + 1))) return rc;
    max_int<<<std::max(1, std::min(64, ceil_div(ncell, 256))), 256, 0, st>>>(drowlen, ncell, dtail);
    chain_rowlen<<<ceil_div(nmol + 1, 256), 256, 0, st>>>(nmol, l->d_order, dmcell, drowlen, dlen);
    if ((rc = ledger_scan<long long>(ctx, l, dlen, drow, nmol + 1))) return rc;
    if ((rc = tables_extent(ctx, box, tb, dtail + 1))) return rc;
    fmolec1_report<<<1, 1, 0, st>>>(drow + nmol, dcstart + ncell, dtail, dtail + 4);
    HIP_CHECK(ctx, hipMemcpyAsync(hsc, dtail + 4, 5 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    const long long total = hsc[0];
    const int nflat = hsc[1], last = hsc[3], lastnear = hsc[4];
    l->nnbrefmax = hsc[2];
    tab_remember(l, tb, last, lastnear);
    ARG_CHECK(ctx, total < 0x7fffffffLL, "epsr_fmolec1: candidate pairs exceed the 2^31 item limit");

    Scratch S;
    auto layout = [&](Scratch &x) {
        x.take<float4>(box->natom);
        x.take<int>((size_t)nflat + 1);
        x.take<float>(total); x.take<int>(total);
        x.take<int>(2 * (size_t)nmol + 1); x.take<int>(2 * (size_t)nmol + 1);
    };
    {
        Scratch q;
        layout(q);
        S.base = (char *)ctx_scratch(ctx, q.off);
        if (!S.base) return EPSR_ERR_HIP;
    }
    float4 *dP = S.take<float4>(box->natom);
    int *dnb = S.take<int>((size_t)nflat + 1);
    float *rowE = S.take<float>(total);
    int *rowJ = S.take<int>(total);
    int *cat = S.take<int>(2 * (size_t)nmol + 1);
    int *ent = S.take<int>(2 * (size_t)nmol + 1);
    cell_flatten<<<(unsigned)ceil_div(ncn, 256), 256, 0, st>>>(ncell, nbr, dnbcell, dnbcnt, dnbpre, dcstart, dnmc,
                                                               l->d_order, dcst, dnb);
    PreparedTables pt;
    if ((rc = ledger_tables(ctx, box, tb, l))) return rc;
    if ((rc = tables_pack(ctx, box, tb, last, lastnear, l->d_tab, pt))) return rc;
    l->tab_packed = true;
    abs_positions<<<ceil_div(box->natom, 256), 256, 0, st>>>(c->xa, c->ya, c->za, c->xadd, c->yadd, c->zadd,
                                                             box->d_amol, box->d_itype, dP, box->natom);
    {
        const int B = box->pair_block;
        timer_begin(ctx, "fmolec1_pair");
        // one candidate queue per molecule type (up to 4 types; more share one queue)
        const int nq = box->nmoltype <= 1 ? 1 : box->nmoltype <= 2 ? 2 : box->nmoltype <= 4 ? 4 : 1;
        const size_t lds = pt.in_lds ? pt.lds_bytes : 0;
        const int g = persistent_blocks(lds, (long long)nmol * 64, B);
#define FM1_LAUNCH(L, Q)                                                                                               \
    fmolec1_pairs<L, Q><<<g, B, lds, st>>>(drow, nmol, l->d_order, dmcell, dcstart, dnb, dP, box->d_naf, box->d_nal,  \
                                           box->d_mtype, pt.tab, box->npairs, rowE, rowJ, cat + nmol)
        if (pt.in_lds) {
            if (nq == 1) FM1_LAUNCH(true, 1);
            else if (nq == 2) FM1_LAUNCH(true, 2);
            else FM1_LAUNCH(true, 4);
        } else {
            if (nq == 1) FM1_LAUNCH(false, 1);
            else if (nq == 2) FM1_LAUNCH(false, 2);
            else FM1_LAUNCH(false, 4);
        }
#undef FM1_LAUNCH
        timer_end(ctx, "fmolec1_pair");
    }
    fmolec1_self_flags<<<ceil_div(nmol + 1, 256), 256, 0, st>>>(eintra, nmol, cat);
    if ((rc = ledger_scan<int>(ctx, l, cat, ent, 2 * nmol + 1))) return rc;
    copy_two<<<1, 1, 0, st>>>(ent + nmol, ent + 2 * nmol, dtail + 10);
    HIP_CHECK(ctx, hipMemcpyAsync(hsc, dtail + 10, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    const int nself = hsc[0];
    const int nent = hsc[1];
    const int npair = nent - nself;
    const int nrec = nself + 2 * npair;

    // the per-molecule lists: records sorted by their molecule's chain position (stable, so
    // a list keeps the self entry first and its pairs in entry order)
    if (l->cap_slots < (size_t)nrec + 1) {
        const size_t cap = (size_t)nrec + nrec / 8 + 64;
        if ((rc = dev_alloc(ctx, &l->d_partner, cap)) || (rc = dev_alloc(ctx, &l->d_entry, cap)) ||
            (rc = dev_alloc(ctx, &l->d_mirror, cap)) || (rc = dev_alloc(ctx, &l->d_slotmol, cap)) ||
            (rc = dev_alloc(ctx, &l->d_pn, cap)) || (rc = dev_alloc(ctx, &l->d_mw, cap)) ||
            (rc = dev_alloc(ctx, &l->d_srec, cap)))
            return rc;
        l->cap_slots = cap;
    }
    if (l->cap_entries < (size_t)nent + 1) {
        const size_t cap = (size_t)nent + nent / 8 + 64;
        if ((rc = dev_alloc(ctx, &l->d_ergpair, cap))) return rc;
        l->cap_entries = cap;
    }
    int *key = nullptr;
    const size_t nr1 = (size_t)nrec + 1;
    if ((rc = ledger_work(ctx, l, LW_REC, sizeof(int) * 5 * nr1, (void **)&key))) return rc;
    int *val = key + nr1, *skey = val + nr1, *sval = skey + nr1, *spos = sval + nr1, *rowk = l->d_rowk;
    {
        int g = persistent_blocks(0, (long long)nmol * 64, 256);
        fmolec1_entries<<<g, 256, 0, st>>>(nmol, l->d_order, l->d_pos, drow, rowE, rowJ, ent, eintra, key, val,
                                           l->d_ergpair);
    }
    if (nrec > 0) {
        int bits = 1;
        while ((1 << bits) < nmol) ++bits;
        size_t tbytes = 0;
        HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, key, skey, val, sval, nrec, 0, bits, st));
        void *tmpb = nullptr;
        if ((rc = ledger_work(ctx, l, LW_TMP, tbytes + 16, &tmpb))) return rc;
        HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(tmpb, tbytes, key, skey, val, sval, nrec, 0, bits, st));
    }
    const int gsum = 256;
    double *dpart = nullptr;
    if ((rc = ledger_work(ctx, l, LW_SUM, sizeof(double) * (2 * gsum + 4), (void **)&dpart))) return rc;
    double *dout = dpart + 2 * gsum;   // [0] pairs, [1] self energies, then (int) the longest list, nbv
    int *dmaxcnt = (int *)(dout + 2), *dnbv = dmaxcnt + 1;
    HIP_CHECK(ctx, hipMemsetAsync(dmaxcnt, 0, 2 * sizeof(int), st));
    fmolec1_rowstart<<<ceil_div(nrec + 1, 256), 256, 0, st>>>(skey, nrec, nmol, rowk);
    fmolec1_counts<<<ceil_div(nmol + 1, 256), 256, 0, st>>>(rowk, l->d_pos, nmol, l->d_rowstart, l->d_cnt, dmaxcnt);
    if (nrec > 0) {
        fmolec1_rows<<<ceil_div(nrec, 256), 256, 0, st>>>(skey, sval, nrec, nself, key, l->d_order, l->d_partner,
                                                          l->d_entry, l->d_slotmol, spos);
        fmolec1_mirror<<<ceil_div(nrec, 256), 256, 0, st>>>(sval, nrec, nself, spos, l->d_mirror);
        if (l->cap_bv < (size_t)nrec) {
            const size_t cap = (size_t)nrec + nrec / 8 + 64;
            if ((rc = dev_alloc(ctx, &l->d_bvslot, cap))) return rc;
            l->cap_bv = cap;
        }
        // key and val are free again: the flags and their scan
        fmolec1_bv_flags<<<ceil_div(nrec + 1, 256), 256, 0, st>>>(nrec, l->d_slotmol, l->d_partner, l->d_rowstart,
                                                                  l->d_pos, key);
        if ((rc = ledger_scan<int>(ctx, l, key, val, nrec + 1))) return rc;
        fmolec1_bv_list<<<ceil_div(nrec + 1, 256), 256, 0, st>>>(nrec, key, val, l->d_bvslot, dnbv);
        fmolec1_slot_aux<<<ceil_div(nrec, 256), 256, 0, st>>>(nrec, l->d_slotmol, l->d_partner, l->d_mirror,
                                                              l->d_rowstart, l->d_pos, box->d_naf, box->d_nal,
                                                              l->d_pn, l->d_mw, l->d_srec);
    }
    // FP64 totals of the stored pair energies and the molecules' own energies
    sum_f64_parts<<<dim3(gsum, 2), 256, 0, st>>>(l->d_ergpair + nself, npair, eintra, nmol, dpart);
    sum_f64_final<<<2, 256, 0, st>>>(dpart, gsum, dout);
    double *hs = (double *)hsc;
    HIP_CHECK(ctx, hipMemcpyAsync(hs, dout, 2 * sizeof(double) + 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    const int maxcnt = ((const int *)(hs + 2))[0];
    l->nbv = ((const int *)(hs + 2))[1];
    // set_ergpair caps every list at nnbrefmax; the cells bound a list by the molecules of
    // the neighbour cells, which is nnbrefmax, so the cap can never bind
    ARG_CHECK(ctx, maxcnt <= l->nnbrefmax, "epsr_fmolec1: a ledger list of %d exceeds nnbrefmax %d", maxcnt,
              l->nnbrefmax);
    l->maxcnt = maxcnt;
    l->nself = nself;
    l->plan_ready = false;
    l->npairent = npair;
    l->nslots = nrec;
    l->nentries = nent;
    l->ledger_ready = true;
    if (ftest64) *ftest64 = hs[0] / (double)nmol;
    if (ftestb64) *ftestb64 = hs[1] / (double)nmol;
    if (ftest) *ftest = (float)(hs[0] / (double)nmol);
    if (ftestb) *ftestb = (float)(hs[1] / (double)nmol);
    if (nnbrefmax_out) *nnbrefmax_out = l->nnbrefmax;
    if (nergpair_out) *nergpair_out = nent;
    return EPSR_OK;
}

extern "C" int epsr_ledger_export(epsr_ctx *ctx, const epsr_ledger *l, int ldl, int *nergpairm, int *molergpair,
                                  int *ergpairref, float *ergpair, int *order)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, l && l->ledger_ready, "epsr_ledger_export: no ledger built");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int nmol = l->nmol;
    hipStream_t st = ctx->stream;
    std::vector<int> partner(l->nslots), entry(l->nslots), rowstart(nmol + 1), cnt(nmol);
    HIP_CHECK(ctx, hipMemcpyAsync(partner.data(), l->d_partner, sizeof(int) * l->nslots, hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipMemcpyAsync(entry.data(), l->d_entry, sizeof(int) * l->nslots, hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipMemcpyAsync(rowstart.data(), l->d_rowstart, sizeof(int) * (nmol + 1), hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipMemcpyAsync(cnt.data(), l->d_cnt, sizeof(int) * nmol, hipMemcpyDeviceToHost, st));
    if (ergpair)
        HIP_CHECK(ctx, hipMemcpyAsync(ergpair, l->d_ergpair, sizeof(float) * l->nentries, hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    for (int m = 0; m < nmol; ++m) {
        int n = cnt[m];
        ARG_CHECK(ctx, n <= ldl, "epsr_ledger_export: ldl %d < list length %d", ldl, n);
        if (nergpairm) nergpairm[m] = n;
        for (int s = 0; s < n; ++s) {
            if (molergpair) molergpair[(size_t)m * ldl + s] = partner[rowstart[m] + s];
            if (ergpairref) ergpairref[(size_t)m * ldl + s] = entry[rowstart[m] + s];
        }
    }
    if (order) {
        HIP_CHECK(ctx, hipMemcpyAsync(order, l->d_order, sizeof(int) * nmol, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
    }
    return EPSR_OK;
}

// ------------------------------------------------------------------ en_change / en_mole_pair
// en_change(im) at the current coordinates: one workgroup; lane s computes the ledger
// slot s >= 1, lane 0 then forms the sequential float32 difference (molfit.py:846-890).
__global__ void en_change_kernel(int im, const int *__restrict__ rowstart, const int *__restrict__ cnt,
                                 const int *__restrict__ partner, const int *__restrict__ entry,
                                 const float *__restrict__ ergpair, const float4 *__restrict__ P,
                                 const int *__restrict__ naf, const int *__restrict__ nal, PairTab t,
                                 float *__restrict__ ergnew_out, float *__restrict__ fdiff_out, float *__restrict__ work)
{
    const int row = rowstart[im], c = cnt[im];
    const float4 *P1 = P + naf[im];
    const int n1 = nal[im] - naf[im] + 1;
    for (int s = threadIdx.x; s < c; s += blockDim.x) {
        float old = ergpair[entry[row + s]];
        float v = old;
        if (s >= 1) {
            int n = partner[row + s];
            v = 0.0f + mol_pair(P1, n1, P + naf[n], nal[n] - naf[n] + 1, t);
        }
        work[2 * s] = v;
        work[2 * s + 1] = old;
        if (ergnew_out) ergnew_out[s] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float fdiff = 0.0f;
        for (int s = 1; s < c; ++s) fdiff = fdiff + work[2 * s] - work[2 * s + 1];
        *fdiff_out = fdiff;
    }
}

extern "C" int epsr_en_change(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                              const epsr_ledger *l, int im, float *ergnew, float *fdiff)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && tb && l && fdiff, "epsr_en_change: null argument");
    ARG_CHECK(ctx, l->ledger_ready, "epsr_en_change: no ledger (call epsr_fmolec1)");
    ARG_CHECK(ctx, im >= 0 && im < box->nmol, "epsr_en_change: molecule %d out of range", im);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    Scratch S;
    Scratch q;
    q.take<float4>(box->natom); q.take<float4>((size_t)box->npairs * std::max(2, std::min(tb->ld, tb->npt2 + 1)));
    q.take<int>(4); q.take<float>(2 * (size_t)l->nnbrefmax + 2); q.take<float>(1);
    S.base = (char *)ctx_scratch(ctx, q.off);
    if (!S.base) return EPSR_ERR_HIP;
    float4 *dP = S.take<float4>(box->natom);
    float4 *dT = S.take<float4>((size_t)box->npairs * std::max(2, std::min(tb->ld, tb->npt2 + 1)));
    int *dtmp = S.take<int>(4);
    float *work = S.take<float>(2 * (size_t)l->nnbrefmax + 2);
    float *dfd = S.take<float>(1);
    PreparedTables pt;
    int rc = prepare_tables(ctx, box, tb, dT, dtmp, pt);
    if (rc) return rc;
    abs_positions<<<ceil_div(box->natom, 256), 256, 0, st>>>(c->xa, c->ya, c->za, c->xadd, c->yadd, c->zadd,
                                                             box->d_amol, box->d_itype, dP, box->natom);
    en_change_kernel<<<1, 256, 0, st>>>(im, l->d_rowstart, l->d_cnt, l->d_partner, l->d_entry, l->d_ergpair, dP,
                                       box->d_naf, box->d_nal, pt.tab, ergnew, dfd, work);
    HIP_CHECK(ctx, hipMemcpyAsync(fdiff, dfd, sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    return EPSR_OK;
}

__global__ void en_mole_pair_kernel(int npair, const int *__restrict__ im, const int *__restrict__ jm,
                                    const float4 *__restrict__ P, const int *__restrict__ naf,
                                    const int *__restrict__ nal, PairTab t, float *__restrict__ out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npair) return;
    int a = im[k], b = jm[k];
    out[k] = mol_pair(P + naf[a], nal[a] - naf[a] + 1, P + naf[b], nal[b] - naf[b] + 1, t);
}

extern "C" int epsr_en_mole_pair(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                                 int npair, const int *im, const int *jm, float *energy)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && tb && im && jm && energy && npair >= 0, "epsr_en_mole_pair: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (npair == 0) return EPSR_OK;
    hipStream_t st = ctx->stream;
    Scratch S, q;
    q.take<float4>(box->natom); q.take<float4>((size_t)box->npairs * std::max(2, std::min(tb->ld, tb->npt2 + 1)));
    q.take<int>(4);
    S.base = (char *)ctx_scratch(ctx, q.off);
    if (!S.base) return EPSR_ERR_HIP;
    float4 *dP = S.take<float4>(box->natom);
    float4 *dT = S.take<float4>((size_t)box->npairs * std::max(2, std::min(tb->ld, tb->npt2 + 1)));
    int *dtmp = S.take<int>(4);
    PreparedTables pt;
    int rc = prepare_tables(ctx, box, tb, dT, dtmp, pt);
    if (rc) return rc;
    abs_positions<<<ceil_div(box->natom, 256), 256, 0, st>>>(c->xa, c->ya, c->za, c->xadd, c->yadd, c->zadd,
                                                             box->d_amol, box->d_itype, dP, box->natom);
    en_mole_pair_kernel<<<ceil_div(npair, 128), 128, 0, st>>>(npair, im, jm, dP, box->d_naf, box->d_nal, pt.tab, energy);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    return EPSR_OK;
}

// ------------------------------------------------------------------ molfit
__device__ __forceinline__ float cr_cos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float cr_exp(float x) { return (float)exp((double)x); }

// A: trial moves.  rotatm (molfit.py:671-711) about z, y, x with correctly rounded
// cosf; movem (:732-759) with the translation triplet at toff (the host has resolved
// any step doubling).  nochange is reported so the host can resolve it.
// the step sizes (rotation, translation) of up to 16 molecule types, passed by value
#define STEP_TYPES 16
struct StepArgs {
    float w[2 * STEP_TYPES];
};

// w[j] of a by-value StepArgs without dynamic indexing of the argument (which would
// copy it to private memory)
__device__ __forceinline__ float step_of(const StepArgs &sa, int j)
{
    float v = 0.0f;
#pragma unroll
    for (int t = 0; t < 2 * STEP_TYPES; ++t)
        if (t == j) v = sa.w[t];
    return v;
}

__global__ void mc_trial(int nmol, const int *__restrict__ order, const int *__restrict__ naf,
                         const int *__restrict__ nal, const float *__restrict__ xa, const float *__restrict__ ya,
                         const float *__restrict__ za, const float *__restrict__ xadd, const float *__restrict__ yadd,
                         const float *__restrict__ zadd, const float *__restrict__ stream, const int *__restrict__ roff,
                         const int *__restrict__ toff, const float *__restrict__ rstep, const float *__restrict__ tstep,
                         const int *__restrict__ mtype, const float *__restrict__ wtype, StepArgs sa,
                         float *__restrict__ txa,
                         float *__restrict__ tya, float *__restrict__ tza, float *__restrict__ txadd,
                         float *__restrict__ tyadd, float *__restrict__ tzadd, int *__restrict__ nochange,
                         TabCheck tc, int *__restrict__ zero, int nzero, int *__restrict__ zero2, int nzero2)
{
    // the cycle's set-up rides along: verify the packed tables, clear the counters in zero
    // and zero2
    tables_verify(tc, nochange);
    if (blockIdx.x == 0) {
        for (int i = threadIdx.x; i < nzero; i += blockDim.x) zero[i] = 0;
        for (int i = threadIdx.x; i < nzero2; i += blockDim.x) zero2[i] = 0;
    }
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmol) return;
    const float pi = 3.14159265358979323846f;
    int m = order[k];
    int a0 = naf[m], a1 = nal[m];
    for (int a = a0; a <= a1; ++a) {
        txa[a] = xa[a];
        tya[a] = ya[a];
        tza[a] = za[a];
    }
    if (roff[k] >= 0 && a1 > a0) {
        float st = rstep ? rstep[k] : (wtype ? wtype[2 * mtype[m]] : step_of(sa, 2 * mtype[m]));
        for (int axis = 0; axis < 3; ++axis) {
            float phi = pi * st * (2.0f * stream[roff[k] + axis] - 1.0f);
            float cphi = cr_cos(phi);
            float sphi = copysignf(fabsf(sqrtf(1.0f - cphi * cphi)), phi);
            for (int a = a0; a <= a1; ++a) {
                float x = txa[a], y = tya[a], z = tza[a];
                if (axis == 0) {
                    txa[a] = x * cphi - y * sphi;
                    tya[a] = x * sphi + y * cphi;
                } else if (axis == 1) {
                    tza[a] = z * cphi - x * sphi;
                    txa[a] = z * sphi + x * cphi;
                } else {
                    tya[a] = y * cphi - z * sphi;
                    tza[a] = y * sphi + z * cphi;
                }
            }
        }
    }
    float x = xadd[m], y = yadd[m], z = zadd[m];
    float nx = x, ny = y, nz = z;
    if (toff[k] >= 0) {
        float st = tstep ? tstep[k] : (wtype ? wtype[2 * mtype[m] + 1] : step_of(sa, 2 * mtype[m] + 1));
        nz = z + st * (2.0f * stream[toff[k]] - 1.0f);
        ny = y + st * (2.0f * stream[toff[k] + 1] - 1.0f);
        nx = x + st * (2.0f * stream[toff[k] + 2] - 1.0f);
        if (nx == x && ny == y && nz == z) atomicAdd(nochange, 1);
    }
    txadd[m] = nx;
    tyadd[m] = ny;
    tzadd[m] = nz;
}

// B: the pair energies a slot can need.  For slot i (molecule m, partner n, slot >= 1):
// A = E(m_trial, n_current), Bv = E(m_trial, n_trial) when n moves before m, and the
// ledger value at the start of the cycle.
// (8 waves per SIMD: two 1024-lane workgroups per CU when their tables fit LDS twice)
// COOP (boxes of molecules of at most three atoms): the wave loads its 64 partners' atoms together,
// consecutive lanes fetching consecutive atoms (about a third of the cache lines of three
// separate loads per lane), through a wave-private LDS area of 192 float4 at coop_off;
// the arithmetic is unchanged
template <bool LDS, bool COOP>
__global__ void __attribute__((amdgpu_flat_work_group_size(1, 1024), amdgpu_waves_per_eu(8)))
mc_pairs(int nslots, int nbv, const int *__restrict__ bvslot, const int4 *__restrict__ srec,
         const int *__restrict__ entry, const int *__restrict__ pos, const float *__restrict__ ergpair,
         const float4 *__restrict__ Pold, const float4 *__restrict__ Pnew, PairTab gt, int npairs,
         const float *__restrict__ shift, float *__restrict__ A, float *__restrict__ Bv, float *__restrict__ L,
         int coop_off)
{
    extern __shared__ char smem[];
    PairTab t = LDS ? tables_to_lds(gt, npairs, smem) : gt;
    const int lane = threadIdx.x & 63;
    float4 *stage = COOP ? (float4 *)(smem + coop_off) + (threadIdx.x >> 6) * 192 : nullptr;
    // items [0, nslots): A (and L) of every slot; items [nslots, nslots + nbv): Bv of the
    // slots whose partner moves earlier in the chain (bvslot, from the ledger), so every
    // lane of a wave evaluates one molecule pair.  srec[i] holds the slot's molecule (bit
    // 31: first slot of its list), partner and both molecules' first atom and atom count.
    // The slots lie in chain order; every XCD takes one contiguous eighth of each range
    // (gridDim.x a multiple of 8), so the ledger entries a slot reads (its own, or the
    // earlier partner's nearby in the chain) come through the L2 that read their neighbours.
    for (int part = 0; part < 2; ++part) {
        const bool bv = part == 1;
        const XcdRange xr = xcd_range(bv ? nbv : nslots);
        const float4 *P2a = bv ? Pnew : Pold;
        // the wave's items are w0 .. w0 + 63 (w0 is the same in every lane)
        for (long long w0 = xr.first - lane; w0 < xr.hi; w0 += xr.stride) {
            const long long it = w0 + lane;
            int i = 0;
            int4 r = make_int4(0, 0, 0, 0);
            bool doit = false;
            if (it < xr.hi) {
                i = bv ? bvslot[it] : (int)it;
                r = srec[i];
                if (!bv) L[i] = ergpair[entry[i]];
                doit = bv || r.x >= 0;
                // rejected by the surrogate stage: no energy
                if (doit && shift && isnan(shift[pos[r.x & 0x7fffffff]])) doit = false;
            }
            const float4 *P1 = Pnew + (r.z & 0xffffff);
            const int n1 = (int)((unsigned)r.z >> 24), n2 = (int)((unsigned)r.w >> 24);
            float e = 0.0f;
            if constexpr (COOP) {
                // every molecule has at most three atoms: lane f of load k fetches atom f % 3
                // of item f / 3's partner (when the partner has it)
                const int b2 = r.w & 0xffffff;
                const int nd = doit ? n2 : 0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int f = 64 * k + lane, src = f / 3, atom = f - 3 * src;
                    const int bb = __shfl(b2, src), na = __shfl(nd, src);
                    stage[f] = atom < na ? P2a[bb + atom] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const float4 q[3] = {stage[3 * lane], stage[3 * lane + 1], stage[3 * lane + 2]};
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (doit) e = mol_pair3_q(P1, n1, q, n2, t);
            } else if (doit) {
                e = mol_pair_any(P1, n1, P2a + (r.w & 0xffffff), n2, t);
            }
            if (doit) {
                if (bv) Bv[i] = e;
                else A[i] = e;
            }
        }
    }
}

// C: the Metropolis chain (molfit.py:397-414), one workgroup.  acc bits by chain
// position live in LDS; lane 0 sums each move in the reference's order.
__global__ void __launch_bounds__(256)
mc_resolve(int nmol, const int *__restrict__ order, const int *__restrict__ pos, const int *__restrict__ rowstart,
           const int *__restrict__ cnt, const int *__restrict__ partner, const int *__restrict__ mirror,
           const float *__restrict__ A, const float *__restrict__ Bv, const float *__restrict__ L,
           const float *__restrict__ stream, const int *__restrict__ moff, float temp, int maxc,
           unsigned *__restrict__ acc_out, float *__restrict__ fdiff_out, float *__restrict__ res)
{
    extern __shared__ unsigned sm[];
    const int nw = (nmol + 31) >> 5;
    unsigned *bits = sm;
    float *term = (float *)(bits + nw);   // [2*maxc]
    for (int i = threadIdx.x; i < nw; i += blockDim.x) bits[i] = 0u;
    float fadd = 0.0f;
    int nacc = 0;
    __syncthreads();
    for (int k = 0; k < nmol; ++k) {
        const int m = order[k];
        const int row = rowstart[m], c = cnt[m];
        for (int s = 1 + threadIdx.x; s < c; s += blockDim.x) {
            int i = row + s;
            int n = partner[i];
            int pn = pos[n];
            bool a = pn < k && ((bits[pn >> 5] >> (pn & 31)) & 1u);
            int mi = mirror[i];
            bool wrote = a && mi != rowstart[n];      // n's move rewrote this entry
            term[2 * s] = a ? Bv[i] : A[i];
            term[2 * s + 1] = wrote ? A[mi] : L[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            float fdiff = 0.0f;
            for (int s = 1; s < c; ++s) fdiff = fdiff + term[2 * s] - term[2 * s + 1];
            float ediff = (fdiff + 0.0f) * temp;
            float etest;
            if (ediff <= 0.0f) etest = 10.0f;
            else if (ediff < 20.0f) etest = cr_exp(-ediff);
            else etest = 0.0f;
            bool accept = stream[moff[k]] < etest;
            if (accept) {
                bits[k >> 5] |= 1u << (k & 31);
                fadd = fadd + fdiff;
                nacc++;
            }
            fdiff_out[k] = fdiff;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nw; i += blockDim.x) acc_out[i] = bits[i];
    if (threadIdx.x == 0) {
        res[0] = fadd;
        res[1] = (float)nacc;
    }
}

// D: write back positions, counters and the ledger (nothing when redo is set: movem's
// step doubling makes the host redo the cycle)
__global__ void mc_commit_mols(int nmol, const int *__restrict__ order, const int *__restrict__ naf,
                               const int *__restrict__ nal, const int *__restrict__ mtype,
                               const unsigned *__restrict__ accb, const int *__restrict__ roff,
                               const int *__restrict__ toff, const float *__restrict__ txa,
                               const float *__restrict__ tya, const float *__restrict__ tza,
                               const float *__restrict__ txadd, const float *__restrict__ tyadd,
                               const float *__restrict__ tzadd, float *__restrict__ xa, float *__restrict__ ya,
                               float *__restrict__ za, float *__restrict__ xadd, float *__restrict__ yadd,
                               float *__restrict__ zadd, int *__restrict__ nstepreject, int count_rot,
                               const int *__restrict__ redo, int *__restrict__ rej)
{
    if (*redo) return;
    const int kk = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = kk < nmol;
    const int k = valid ? kk : nmol - 1;
    int m = order[k];
    bool acc = valid && ((accb[k >> 5] >> (k & 31)) & 1u);
    if (acc) {
        for (int a = naf[m]; a <= nal[m]; ++a) {
            xa[a] = txa[a];
            ya[a] = tya[a];
            za[a] = tza[a];
        }
        xadd[m] = txadd[m];
        yadd[m] = tyadd[m];
        zadd[m] = tzadd[m];
        nstepreject[m] = 0;
    } else if (valid) {
        nstepreject[m] += 1;
    }
    // rejection counters per molecule type, one atomic per type present in the wave
    const int mt = mtype[m];
    bool rr = valid && !acc && count_rot, rt = valid && !acc && toff[k] >= 0;
    unsigned long long live = __ballot(rr || rt);
    while (live) {
        const int t = __shfl(mt, __ffsll((long long)live) - 1);
        const unsigned long long same = __ballot(mt == t);
        const int nr = __popcll(__ballot(rr && mt == t)), nt = __popcll(__ballot(rt && mt == t));
        if ((int)(threadIdx.x & 63) == __ffsll((long long)(same & live)) - 1) {
            if (nr) atomicAdd(&rej[2 * t], nr);
            if (nt) atomicAdd(&rej[2 * t + 1], nt);
        }
        live &= ~same;
    }
}

// ledger write-back (molfit.py:417-447): an accepted move rewrites the entries of its list's
// slots s >= 1 with their new energies, in chain order, so an entry ends with the value of
// the later of its two molecules' writes.  One wave per accepted chain position k writes
// the entries whose last write is its own: for a partner earlier in the chain, always
// (the energy against the partner's final position: Bv when the partner moved, else A);
// for a later partner, unless the partner was accepted and the pair is not in slot 0 of
// the partner's list (mw[i] >= 0), when the partner's move writes last.  Exactly one move
// writes each rewritten entry.
__global__ void __launch_bounds__(256)
mc_commit_ledger(int nmol, const int *__restrict__ rowk, const int *__restrict__ pnv, const int *__restrict__ mw,
                 const int *__restrict__ entry, const unsigned *__restrict__ accb, const float *__restrict__ A,
                 const float *__restrict__ Bv, const int *__restrict__ redo, float *__restrict__ ergpair)
{
    const int lane = threadIdx.x & 63;
    const int k = xcd_block() * 4 + (threadIdx.x >> 6);   // chain order, each XCD a contiguous part
    if (k >= nmol || *redo) return;
    if (!((accb[k >> 5] >> (k & 31)) & 1u)) return;
    const int row = rowk[k], c = rowk[k + 1] - row;
    for (int s = 1 + lane; s < c; s += 64) {
        const int i = row + s, pn = pnv[i];
        const bool an = (accb[pn >> 5] >> (pn & 31)) & 1u;
        if (pn < k) ergpair[entry[i]] = an ? Bv[i] : A[i];
        else if (!an || mw[i] < 0) ergpair[entry[i]] = A[i];
    }
}

// ------------------------------------------------------------------ fast resolution
// The reference decides move k from fdiff_k, the float32 sequential sum over its slots
// s >= 1 of (+ergnew_s, -ergold_s).  Every slot has two variants: partner not yet moved
// (or rejected) and partner accepted earlier in the chain.  Let e0_s = A - L (first
// variant) and delta_s = (Bv - L') - e0_s with L' the ledger value the partner's move
// wrote.  Then, in exact arithmetic, fdiff_k = base_k + sum of delta_s over earlier
// accepted partners, with base_k = sum of e0_s.  A move is decided from this value
// when the decision is the same for every float32 result within the a-priori rounding
// bound of the reference's sequential sum; otherwise the exact sequential float32 sum
// is formed and used (same result as the reference in every case).
//   bound_k = u * (1 + 1e-3) * W_k        (recursive summation: W_k bounds the sum of
//                                           the moduli of all partial sums)
//           + u * D_k * (npb_k + 2)       (float32 rounding and accumulation of deltas)
// with u = 2^-24, D_k = sum |delta_s| and npb_k the number of earlier partners.
//
// The Metropolis test r < RN(exp(-ediff)), ediff = RN(fdiff * temp), is decided in log
// space with nlr = -log(r) (double, relative error < 1e-15):
//   accept for every ediff <= hi  if  hi <= 0  or  (hi < 20 and hi <= nlr - 2^-22)
//   reject for every ediff >= lo  if  lo >= 20 or (lo > 0 and lo >= nlr + 1e-12)
// (exp(-x) >= r (1 + 2^-22) gives RN(exp(-x)) > r; exp(-x) < r gives RN(exp(-x)) <= r),
// where [lo, hi] = temp * [fdiff - bound, fdiff + bound] widened by 2^-22 relative for
// the float32 rounding of fdiff and of the product.  Both conditions are monotone in
// fdiff, so they become two float32 thresholds on the pushed sum accde_k:
//   accept when accde_k <= fa_k, reject when accde_k >= fr_k, else sum exactly.

__device__ __forceinline__ float etest_of(float ediff)
{
    if (ediff <= 0.0f) return 10.0f;
    if (ediff < 20.0f) return (float)exp(-(double)ediff);
    return 0.0f;
}

// float32 rounding towards -inf / +inf of a double
__device__ __forceinline__ float f_down(double x)
{
    float f = (float)x;
    return (double)f > x ? nextafterf(f, -INFINITY) : f;
}
__device__ __forceinline__ float f_up(double x)
{
    float f = (float)x;
    return (double)f < x ? nextafterf(f, INFINITY) : f;
}

// thresholds on accde for a move with base sum b, rounding bound bnd and nlr = -log(r)
__device__ void mc_thresholds(double b, double bnd, double tempd, double nlr, float &fa, float &fr)
{
    const double rel = 2.384185791015625e-07;   // 2^-22
    const double tiny = 1e-30;
    fa = -INFINITY;
    fr = INFINITY;
    if (!(tempd > 0.0) || !isfinite(tempd) || !isfinite(b) || !isfinite(bnd) || !isfinite(nlr)) return;
    // accept iff hi' <= ta, hi' = h + |h| rel + tiny, h = (fdiff + bnd) temp
    const double ta = nlr - rel < 20.0 ? fmax(0.0, nlr - rel) : 19.999999999;
    const double ha = (ta - tiny) / (ta - tiny >= 0.0 ? 1.0 + rel : 1.0 - rel);
    // reject iff lo' >= tr (> 0), lo' = l - |l| rel - tiny, l = (fdiff - bnd) temp
    const double tr = fmin(20.0, nlr + 1e-12);
    const double lr = (tr + tiny) / (1.0 - rel);
    const double va = ha / tempd, vr = lr / tempd;
    // cover the double roundings of these few operations
    const double slack = 1e-14 * (fabs(va) + fabs(vr) + bnd + fabs(b)) + 1e-300;
    fa = f_down(va - bnd - b - slack);
    fr = f_up(vr + bnd - b + slack);
}

// one wave per molecule: the deltas of its slots (delta, the change of its energy change
// when that earlier partner is accepted) and the decision thresholds of its move (thr, by
// chain position)
#define PREP_CH 4
__global__ void __launch_bounds__(256)
mc_prepare(int nmol, const int *__restrict__ rowk, const int *__restrict__ pnv, const int *__restrict__ mw,
           const float *__restrict__ A, const float *__restrict__ Bv, const float *__restrict__ L,
           const float *__restrict__ stream, const int *__restrict__ moff, float temp,
           const float *__restrict__ shift, float *__restrict__ delta, float2 *__restrict__ thr,
           unsigned char *__restrict__ state, int *__restrict__ wout, int *__restrict__ nout)
{
    // Each rounding of the sequential sum errs by at most u |partial sum|.  With
    // P_s = |sum_{s'<=s} e0| + sum_{s'<=s, earlier partner} |delta| the partial sums obey
    // |S_2s| <= P_s and |S_2s-1| <= P_s-1 + max|ergnew_s| whatever the earlier decisions,
    // so W = sum_s (P_s-1 + max|ergnew_s| + P_s) bounds sum_j |S_j|.  One wave per
    // molecule, taken in chain order (each XCD a contiguous part of the chain, whose lists
    // and their partners' lists lie together); the slots are scanned in order, 64 at a
    // time, with the loads of PREP_CH chunks issued together.
    //
    // The first pass of the resolution rides along: with every earlier partner still
    // undecided, a move whose pushed sum lies on one side of its thresholds for every
    // outcome (the sum of its negative deltas at or above fr: rejected; of its positive ones
    // at or below fa: accepted) is decided here; the others go to the work list wout.
    const int pm = xcd_block() * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (pm >= nmol) return;
    const int row = rowk[pm], c = rowk[pm + 1] - row;
    const float sh = shift ? shift[pm] : 0.0f;
    if (isnan(sh)) {
        // rejected by the surrogate stage: never accepted, receives no pushes
        if (lane == 0) {
            thr[pm] = make_float2(-INFINITY, -INFINITY);
            state[pm] = 2;
        }
        return;
    }
    double carry_e = 0.0, carry_d = 0.0, pprev = 0.0, w = 0.0, dneg = 0.0, dpos = 0.0;
    int npb = 0;
    for (int g0 = 1; g0 < c; g0 += 64 * PREP_CH) {
        double e0[PREP_CH], ad[PREP_CH];
        float maxnew[PREP_CH];
#pragma unroll
        for (int j = 0; j < PREP_CH; ++j) {
            const int s = g0 + 64 * j + lane;
            e0[j] = 0.0;
            ad[j] = 0.0;
            maxnew[j] = 0.0f;
            if (s < c) {
                int i = row + s;
                bool before = pnv[i] < pm;
                float a = A[i], l = L[i];
                e0[j] = (double)a - (double)l;
                maxnew[j] = fabsf(a);
                float dl = 0.0f;
                if (before) {
                    int mi = mw[i];
                    float lw = mi >= 0 ? A[mi] : l;
                    float bv = Bv[i];
                    dl = (float)(((double)bv - (double)lw) - e0[j]);
                    maxnew[j] = fmaxf(maxnew[j], fabsf(bv));
                    ad[j] = fabs((double)dl);
                    if (dl < 0.0f) dneg += (double)dl;
                    else dpos += (double)dl;
                    npb++;
                }
                delta[i] = dl;
            }
        }
#pragma unroll
        for (int j = 0; j < PREP_CH; ++j) {
            if (g0 + 64 * j >= c) break;   // wave-uniform
            const int s = g0 + 64 * j + lane;
            double se = e0[j], sd = ad[j];
            for (int off = 1; off < 64; off <<= 1) {
                double te = __shfl_up(se, off), td = __shfl_up(sd, off);
                if (lane >= off) {
                    se += te;
                    sd += td;
                }
            }
            double pe = carry_e + se, pd = carry_d + sd;
            double ps = fabs(pe) + pd;
            double pm1 = __shfl_up(ps, 1);
            if (lane == 0) pm1 = pprev;
            if (s < c) w += pm1 + (double)maxnew[j] + ps;
            carry_e = __shfl(pe, 63);
            carry_d = __shfl(pd, 63);
            pprev = __shfl(ps, 63);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        w += __shfl_down(w, off);
        npb += __shfl_down(npb, off);
        dneg += __shfl_down(dneg, off);
        dpos += __shfl_down(dpos, off);
    }
    if (lane == 0) {
        const double u = 5.9604644775390625e-08;   // 2^-24
        const double bnd = u * (1.001 * w + carry_d * (double)(npb + 2)) + 1e-12 * fabs(carry_e) + 1e-30;
        const double nlr = -log((double)stream[moff[pm]]);
        float fa, fr;
        mc_thresholds(carry_e - (double)sh, bnd, (double)temp, nlr, fa, fr);
        thr[pm] = make_float2(fa, fr);
        const int d = dpos <= (double)fa ? 1 : (dneg >= (double)fr ? 2 : 0);
        state[pm] = (unsigned char)d;
        if (!d) wout[atomicAdd(nout, 1)] = pm;
    }
}

// fdiff of molfit.py:397-405: the float32 sum fdiff = fdiff + new_s - old_s over the slots
// s in [1, c) in order, from (new, old) pairs in LDS read eight at a time ahead of the adds
__device__ __forceinline__ float serial_fdiff(const float2 *term, int c)
{
    float fdiff = 0.0f;
    int s = 1;
    for (; s + 8 <= c; s += 8) {
        float2 v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = term[s + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) fdiff = fdiff + v[j].x - v[j].y;
    }
    for (; s < c; ++s) fdiff = fdiff + term[s].x - term[s].y;
    return fdiff;
}

// exact fdiff of every accepted move (for fadd): one wave per chain position gathers
// the selected variant of every slot into LDS, lane 0 sums them in the reference's order
__global__ void __launch_bounds__(256)
mc_fdiff_accepted(int nmol, const int *__restrict__ rowk, const int *__restrict__ pnv, const int *__restrict__ mw,
                  const float *__restrict__ A, const float *__restrict__ Bv,
                  const float *__restrict__ L, const unsigned *__restrict__ bits, int stride,
                  float *__restrict__ fdiff_out, unsigned *__restrict__ bits_out, int nwords)
{
    extern __shared__ float2 fterm[];   // [4 waves][stride >= longest list]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int k = xcd_block() * 4 + wv;   // chain order, each XCD a contiguous part
    // a copy of the acceptance bits beside the fdiffs, when asked for (nwords <= nmol)
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (bits_out && g < nwords) bits_out[g] = bits[g];
    if (k >= nmol) return;
    if (!((bits[k >> 5] >> (k & 31)) & 1u)) return;
    const int row = rowk[k], c = rowk[k + 1] - row;
    float2 *t = fterm + (size_t)wv * stride;
    // the terms, four chunks of 64 slots with their loads in flight together
    for (int g0 = 1; g0 < c; g0 += 256) {
        float2 v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = g0 + 64 * j + lane;
            if (s < c) {
                const int i = row + s;
                const int pn = pnv[i];
                const bool a = pn < k && ((bits[pn >> 5] >> (pn & 31)) & 1u);
                const int mi = mw[i];
                v[j] = make_float2(a ? Bv[i] : A[i], (a && mi >= 0) ? A[mi] : L[i]);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (g0 + 64 * j + lane < c) t[g0 + 64 * j + lane] = v[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (lane == 0) fdiff_out[k] = serial_fdiff(t, c);
}

// ------------------------------------------------------------------ parallel resolution
// The decisions of a cycle, resolved by every CU at once.  A move's pushed sum is the sum
// of delta over its earlier partners that were accepted; while some of them are
// undecided it lies between sure + (negative undecided deltas) and sure + (positive
// ones).  A move whose whole range falls on one side of its thresholds is decided without
// waiting for those partners (every outcome gives the same decision); a move with no
// undecided earlier partner is decided exactly as the chain would (thresholds, or the
// reference's sequential float32 sum when they are ambiguous).  Passes repeat over the
// undecided moves; the earliest undecided move in the chain always decides, so every
// pass makes progress and the result is the chain's.  Sums are FP64: within 1e-16 of the
// exact sum, far inside the float32 accumulation bound the thresholds allow for.
// state[k]: 0 undecided, 1 accepted, 2 rejected.

// one wave decides move k if it can; returns the decision (0: still undecided)
// pnv[i]: chain position of slot i's partner; mw[i]: the mirror slot, -1 when it is the
// partner's first slot (a slot that no move rewrites)
__device__ int mc_pass_move(int k, const int *__restrict__ rowk, const int *__restrict__ pnv,
                            const int *__restrict__ mw,
                            const float *__restrict__ A, const float *__restrict__ Bv, const float *__restrict__ L,
                            const float *__restrict__ delta, const float2 *__restrict__ thr,
                            const float *__restrict__ stream, const int *__restrict__ moff, float temp,
                            const float *__restrict__ shift, volatile unsigned char *state, float2 *term,
                            int *nexact)
{
    const int lane = threadIdx.x & 63;
    const int row = rowk[k], c = rowk[k + 1] - row;
    double sure = 0.0, lo = 0.0, hi = 0.0;
    int und = 0;
    for (int s = 1 + lane; s < c; s += 64) {
        const int i = row + s;
        const int pn = pnv[i];
        if (pn < k) {
            const int sp = state[pn];
            const double d = (double)delta[i];
            if (sp == 1) {
                sure += d;
            } else if (sp == 0) {
                und = 1;
                if (d < 0.0) lo += d;
                else hi += d;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sure += __shfl_xor(sure, off);
        lo += __shfl_xor(lo, off);
        hi += __shfl_xor(hi, off);
        und |= __shfl_xor(und, off);
    }
    const float2 t = thr[k];
    if (sure + hi <= (double)t.x) return 1;
    if (sure + lo >= (double)t.y) return 2;
    if (und) return 0;
    // every earlier partner decided, thresholds ambiguous: the reference's sum
    for (int s = 1 + lane; s < c; s += 64) {
        const int i = row + s;
        const int pn = pnv[i];
        const bool acc = pn < k && state[pn] == 1;
        const int mi = mw[i];
        term[s] = make_float2(acc ? Bv[i] : A[i], (acc && mi >= 0) ? A[mi] : L[i]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int dd = 0;
    if (lane == 0) {
        const float fdiff = serial_fdiff(term, c);
        const float ediff = (shift ? fdiff - shift[k] : fdiff + 0.0f) * temp;
        dd = stream[moff[k]] < etest_of(ediff) ? 1 : 2;
        atomicAdd(nexact, 1);
    }
    __builtin_amdgcn_wave_barrier();
    return __shfl(dd, 0);
}

// one pass over a work list (the whole chain when first): decided moves write state,
// the others go to the next list
__global__ void __launch_bounds__(256)
mc_resolve_pass(int nmol, const int *__restrict__ rowk, const int *__restrict__ pnv, const int *__restrict__ mw,
                const float *__restrict__ A, const float *__restrict__ Bv,
                const float *__restrict__ L, const float *__restrict__ delta, const float2 *__restrict__ thr,
                const float *__restrict__ stream, const int *__restrict__ moff, float temp,
                const float *__restrict__ shift, int maxc, int first, const int *__restrict__ win,
                const int *__restrict__ nin, int *__restrict__ wout, int *__restrict__ nout,
                unsigned char *state, int *nexact)
{
    extern __shared__ float2 pterm[];   // [4 waves][maxc]
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float2 *term = pterm + (size_t)wv * maxc;
    const int n = first ? nmol : *nin;
    for (int w = xcd_block() * 4 + wv; w < n; w += gridDim.x * 4) {
        const int k = first ? w : win[w];
        const int d = mc_pass_move(k, rowk, pnv, mw, A, Bv, L, delta, thr, stream, moff, temp, shift, state, term,
                                   nexact);
        if (lane == 0) {
            if (d) state[k] = (unsigned char)d;
            else wout[atomicAdd(nout, 1)] = k;
        }
    }
}

// the remaining passes in one workgroup (up to 16 waves, as many as the per-wave term
// rows of maxc slots fit in LDS), with a barrier between passes, until no move is
// undecided; lists ping-pong between wa and wb.  Then the acceptance bits by chain
// position from the decisions.
__global__ void __launch_bounds__(1024)
mc_resolve_finish(const int *__restrict__ rowk, const int *__restrict__ pnv, const int *__restrict__ mw,
                  const float *__restrict__ A, const float *__restrict__ Bv, const float *__restrict__ L,
                  const float *__restrict__ delta, const float2 *__restrict__ thr, const float *__restrict__ stream,
                  const int *__restrict__ moff, float temp, const float *__restrict__ shift, int maxc,
                  int *__restrict__ wa, const int *__restrict__ na, int *__restrict__ wb, unsigned char *state,
                  int *nexact, int *npass, int nmol, unsigned *__restrict__ bits, int maxpass,
                  int *__restrict__ undecided)
{
    extern __shared__ float2 fterm2[];   // [waves][maxc]
    __shared__ int ncur, nnext, nund;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (int)(blockDim.x >> 6);
    float2 *term = fterm2 + (size_t)wv * maxc;
    if (threadIdx.x == 0) {
        ncur = *na;
        nnext = 0;
    }
    __syncthreads();
    int *cur = wa, *nxt = wb;
    // every pass decides at least the earliest undecided move of the chain, so nmol passes
    // always suffice; the cap only guards against a fault in that argument, and what is left
    // undecided is counted for the host
    int passes = 0;
    while (ncur > 0 && passes < maxpass) {
        for (int w = wv; w < ncur; w += nwv) {
            const int k = cur[w];
            const int d = mc_pass_move(k, rowk, pnv, mw, A, Bv, L, delta, thr, stream, moff, temp, shift, state,
                                       term, nexact);
            if (lane == 0) {
                if (d) state[k] = (unsigned char)d;
                else nxt[atomicAdd(&nnext, 1)] = k;
            }
        }
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0) {
            ncur = nnext;
            nnext = 0;
        }
        passes++;
        int *t = cur;
        cur = nxt;
        nxt = t;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *npass = passes;
        nund = 0;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < (nmol + 31) / 32; w += blockDim.x) {
        unsigned b = 0u;
        int und = 0;
        for (int j = 0; j < 32; ++j) {
            const int k = w * 32 + j;
            if (k < nmol && state[k] == 1) b |= 1u << j;
            if (k < nmol && state[k] == 0) ++und;
        }
        bits[w] = b;
        if (und) atomicAdd(&nund, und);
    }
    __syncthreads();
    if (threadIdx.x == 0) *undecided = nund;
}


// Delayed acceptance, stage 1, by chain position k: dEs from the network output of
// rows 2k (current) and 2k+1 (trial), or from des by molecule; the move passes with
// probability min(1, exp(-dEs/simt)) tested in the float32 form of the Metropolis test,
// against the uniform of its molecule.  shift[k] = dEs for a pass, NaN for a rejection.
__global__ void da_stage1(int nmol, const int *__restrict__ order, const float *__restrict__ y,
                          const float *__restrict__ des, float temp, unsigned long long seed,
                          float *__restrict__ shift, float *__restrict__ rec_des, int *__restrict__ rec_pass,
                          int *__restrict__ nrej)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmol) return;
    const int m = order[k];
    const float d = des ? des[m] : y[2 * k + 1] - y[2 * k];
    const float e1 = d * temp;
    const bool pass = da_uniform(seed, m) < etest_of(e1);
    shift[k] = pass ? d : __int_as_float(0x7fc00000);
    if (rec_des) rec_des[k] = d;
    if (rec_pass) rec_pass[k] = pass ? 1 : 0;
    if (!pass) atomicAdd(nrej, 1);
}

// the exact energy change of the move at chain position k over its later partners
// (still at their cycle-start positions): sum of A - L over those slots, FP64, one wave
// per move; NaN where the surrogate stage rejected the move (its A was not formed)
__global__ void __launch_bounds__(64)
mc_record_target(int nmol, const int *__restrict__ order, const int *__restrict__ pos,
                 const int *__restrict__ rowstart, const int *__restrict__ cnt, const int *__restrict__ partner,
                 const float *__restrict__ A, const float *__restrict__ L, const float *__restrict__ shift,
                 float *__restrict__ target)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    const int m = order[k];
    const int row = rowstart[m], c = cnt[m];
    double s = 0.0;
    for (int j = 1 + lane; j < c; j += 64)
        if (pos[partner[row + j]] > k) s += (double)A[row + j] - (double)L[row + j];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    if (lane == 0) target[k] = (shift && isnan(shift[k])) ? __int_as_float(0x7fc00000) : (float)s;
}

__host__ __device__ static float adapt_ratio(float rejrate, int att, int rej)
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

// ---- ran1 windows.  A window holds len values of the stream from a state, with the
// state at every 256th value; a cycle whose ran1 state lies in a window reads its stream
// there, at that offset (other consumers, such as atogen's origin draws, may have moved
// the state on since the window was drawn).
static void fill_window(Ran1 from, float *out, int len, std::vector<int32_t> &ck)
{
    ck.assign(35 * ((size_t)len / 256 + 1), 0);
    for (int o = 0;; o += 256) {
        ran1_store(from, &ck[35 * ((size_t)o / 256)]);
        if (o >= len) break;
        const int n = std::min(256, len - o);
        ran1_fill(from, out + o, n);
        if (n < 256) break;
    }
}

// the state at offset o of window b
static void window_state(const epsr_ledger *l, int b, int o, Ran1 &out)
{
    ran1_load(out, &l->win_ck[b][35 * ((size_t)o / 256)]);
    for (int k = 0; k < o % 256; ++k) (void)ran1_next(out);
}

// the offset of state in window b with room for need values after it, or -1
static int window_offset(const epsr_ledger *l, int b, const int32_t *state, int need)
{
    if (b < 0 || l->win_len[b] < need) return -1;
    const uint32_t o = (uint32_t)state[0] - (uint32_t)l->win_ck[b][0];   // calls made since its start
    if (o > (uint32_t)(l->win_len[b] - need)) return -1;
    Ran1 r;
    window_state(l, b, (int)o, r);
    int32_t a[35];
    ran1_store(r, a);
    return std::memcmp(a, state, sizeof(a)) == 0 ? (int)o : -1;
}

// a host thread draws window buf (len values from `from`); the window is usable once
// gen_finish has joined it.  With upload, the thread also copies it to d_stream[buf] on
// the ledger's upload stream and records ev_up there (both exist before the call).
static void gen_start(epsr_ledger *l, int buf, const Ran1 &from, int len, bool upload = false)
{
    l->gen_buf = buf;
    l->gen_len = len;
    l->gen_err = hipSuccess;
    l->win_len[buf] = 0;
    float *out = l->pin[buf];
    std::vector<int32_t> *ck = &l->win_ck[buf];
    float *dst = upload ? l->d_stream[buf] : nullptr;
    hipStream_t us = l->up_stream;
    hipEvent_t ev = l->ev_up;
    hipError_t *err = &l->gen_err;
    const int dev = l->ctx->device;
    l->gen = std::thread([from, out, len, ck, dst, us, ev, err, dev]() {
        fill_window(from, out, len, *ck);
        if (!dst) return;
        hipError_t e = hipSetDevice(dev);
        if (e == hipSuccess) e = hipMemcpyAsync(dst, out, sizeof(float) * (size_t)len, hipMemcpyHostToDevice, us);
        if (e == hipSuccess) e = hipEventRecord(ev, us);
        *err = e;
    });
}

static void gen_finish(epsr_ledger *l)
{
    gen_join(l);
    if (l->gen_buf >= 0) {
        l->win_len[l->gen_buf] = l->gen_len;
        if (l->gen_err != hipSuccess) {
            // its upload failed: the window is not used
            l->win_len[l->gen_buf] = 0;
            if (l->pin_ahead == l->gen_buf) l->pin_ahead = -1;
        }
    }
    l->gen_err = hipSuccess;
    l->gen_buf = -1;
}

// Step-size adaptation after a cycle on the device (molfit.py:465-536, the host loop of a
// single cycle restated), so that cycles follow each other without the host.  blk: the
// cycle's result block (16 + 2 nmt ints; [9] the no-change flag, [16..] rejections per
// type); att: attempts per type; w: the step sizes the cycle used, replaced by the next
// cycle's (a zero step of a type that moves set to the minimum, molfit.py:375-392).  The
// block and the adapted steps go to slot (16 + 2 nmt ints, then 2 nmt floats).  A cycle
// whose trial moves did not all change committed nothing, and neither do the ones after
// it: their steps are left alone.
__global__ void mc_adapt(int nmt, const int *__restrict__ att, const int *__restrict__ type_nat,
                         const int *__restrict__ blk, float *__restrict__ w, int *__restrict__ slot, float rejrate,
                         int ireset, int stepr_on, int stepm_on, float rotstepmax, float rotstepmin,
                         float trstepmax, float trstepmin)
{
    const int nblk = 16 + 2 * nmt;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) slot[i] = blk[i];
    if (blk[9] != 0) return;
    float *after = (float *)(slot + nblk);
    const int *rej = blk + 16;
    for (int mt = threadIdx.x; mt < nmt; mt += blockDim.x) {
        float wr = w[2 * mt], wt = w[2 * mt + 1];
        if (ireset != 1) {
            float rat = adapt_ratio(rejrate, att[2 * mt], rej[2 * mt]);
            wr = wr * rat;
            if (wr > rotstepmax) wr = rotstepmax;
            if (wr < rotstepmin && wr != 0.0f) wr = rotstepmin;
            rat = adapt_ratio(rejrate, att[2 * mt + 1], rej[2 * mt + 1]);
            wt = wt * rat;
            if (wt > trstepmax) wt = trstepmax;
            if (wt < trstepmin && wt != 0.0f) wt = trstepmin;
        }
        after[2 * mt] = wr;
        after[2 * mt + 1] = wt;
        if (type_nat[mt] != 0) {
            if (stepr_on && type_nat[mt] > 1) {
                if (wr <= 0.0f) wr = rotstepmin;
            } else {
                wr = 0.0f;
            }
            if (stepm_on) {
                if (wt <= 0.0f) wt = trstepmin;
            } else {
                wt = 0.0f;
            }
        }
        w[2 * mt] = wr;
        w[2 * mt + 1] = wt;
    }
}

// the plan of the chain (molfit.py:367-403) on the device.  flags: 1 stepr > 0, 2 stepm > 0.
// Values per move: 3 for a rotation (molecules of more than one atom), 3 for a translation,
// 1 for the Metropolis test
__global__ void mc_plan_counts(int nmol, const int *__restrict__ order, const int *__restrict__ naf,
                               const int *__restrict__ nal, int flags, int *__restrict__ cnt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmol) return;
    const int m = order[k];
    const bool rot = (flags & 1) && nal[m] > naf[m];
    cnt[k] = (rot ? 3 : 0) + ((flags & 2) ? 3 : 0) + 1;
}

// offsets from the exclusive sums in base (overwritten by moff): roff -1 without
// rotation, -3 for a single atom when stepr > 0; toff -1 without translation
__global__ void mc_plan_offsets(int nmol, const int *__restrict__ order, const int *__restrict__ naf,
                                const int *__restrict__ nal, int flags, int *__restrict__ roff,
                                int *__restrict__ toff, int *__restrict__ base_moff)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmol) return;
    const int m = order[k];
    const bool rot = (flags & 1) && nal[m] > naf[m];
    const int b = base_moff[k];
    roff[k] = rot ? b : ((flags & 1) ? -3 : -1);
    const int t = b + (rot ? 3 : 0);
    toff[k] = (flags & 2) ? t : -1;
    base_moff[k] = t + ((flags & 2) ? 3 : 0);
}

// the most cycles one pass of epsr_molfit_cycles keeps on the device (more are run in turns)
static const int MC_BATCH_MAX = 64;

static int molfit_impl(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                       epsr_ledger *l, const epsr_mc_params_t *p, int ncycles, float *wholerot, float *wholetrans,
                       int *nwhole, int *nmovem, int *nstepreject, int32_t *ran1_state, float *fadds,
                       const epsr_da_t *da, const epsr_mc_record_t *rec, epsr_mc_stats_t *stats)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && tb && l && p && wholerot && wholetrans && nwhole && nmovem && nstepreject &&
                       ran1_state && fadds,
              "epsr_molfit: null argument");
    ARG_CHECK(ctx, l->ledger_ready, "epsr_molfit: no ledger (call epsr_fmolec1)");
    ARG_CHECK(ctx, ncycles >= 1, "epsr_molfit_cycles: ncycles=%d", ncycles);
    if (ncycles > 1) {
        // cycles back to back need the parallel resolution and no surrogate stage or
        // recording; otherwise, and beyond MC_BATCH_MAX cycles, they are run in turns
        const bool par = sizeof(float2) * 4 * (size_t)std::max(1, l->maxcnt) <= 160 * 1024;
        const int first = (da || rec || !par) ? 1 : std::min(ncycles, MC_BATCH_MAX);
        if (first < ncycles) {
            epsr_mc_stats_t s0{};
            int rc0 = molfit_impl(ctx, box, c, tb, l, p, first, wholerot, wholetrans, nwhole, nmovem, nstepreject,
                                  ran1_state, fadds, da, rec, &s0);
            if (rc0) return rc0;
            rc0 = molfit_impl(ctx, box, c, tb, l, p, ncycles - first, wholerot, wholetrans, nwhole, nmovem,
                              nstepreject, ran1_state, fadds, da, rec, stats);
            if (!rc0 && stats) {
                stats->cycles += s0.cycles;
                stats->accepted_total += s0.accepted_total;
            }
            return rc0;
        }
    }
    const bool batch = ncycles > 1;
    // the window the previous call left (drawn and uploaded by a host thread) is ready once
    // that thread is joined
    if (l->gen_buf >= 0 && l->gen_buf == l->pin_ahead) gen_finish(l);
    // delayed acceptance and recording: descriptor rows of width ldx
    const bool use_net = da && !da->des;
    ARG_CHECK(ctx, !da || da->des || da->mlp, "epsr_molfit_da: neither a network nor des");
    const bool want_x = use_net || (rec && rec->x);
    const epsr_desc_params_t *dpar = da ? &da->desc : (rec ? rec->desc : nullptr);
    int ldx = 0;
    if (use_net) ldx = da->mlp->d;
    if (rec && rec->x) {
        ARG_CHECK(ctx, !use_net || rec->ldx == ldx, "epsr_molfit_da: rec->ldx=%d differs from the network's d=%d",
                  rec ? rec->ldx : 0, ldx);
        ldx = rec->ldx;
    }
    ARG_CHECK(ctx, !want_x || dpar, "epsr_molfit_da: recording descriptors needs rec->desc");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int nmol = box->nmol, natom = box->natom, nmt = box->nmoltype;
    const float temp = 1.0f / p->simt;
    float trstepmax = 1.0f / box->natomr;
    trstepmax = powf(trstepmax, 1.0f / 3.0f);
    float trstepmin = 0.0001f * trstepmax;
    if (trstepmin < 1.0e-6f) trstepmin = 1.0e-6f;
    const float rotstepmax = 1.0f, rotstepmin = 1.0e-5f;

    // the plan of the chain (molfit.py:367-403): per move, the ran1 offsets of its
    // rotation (-1 none, -3 a single atom with stepr > 0), translation and Metropolis
    // values; it depends on the chain order and the stepr/stepm flags only
    const int flags = (p->stepr > 0.0f ? 1 : 0) | (p->stepm > 0.0f ? 2 : 0);
    if (!l->plan_ready || l->plan_flags != flags) {
        // attempts per type and the stream length do not depend on the chain order
        l->att.assign(2 * nmt, 0);
        l->type_nat.assign(nmt, 0);
        int cur = 0;
        for (int m = 0; m < nmol; ++m) {
            const int mt = box->h_mtype[m];
            const int nat = box->h_nal[m] - box->h_naf[m] + 1;
            l->type_nat[mt] = nat;
            if (p->stepr > 0.0f && nat > 1) {
                l->att[2 * mt] += 1;
                cur += 3;
            }
            if (p->stepm > 0.0f) {
                l->att[2 * mt + 1] += 1;
                cur += 3;
            }
            cur += 1;
        }
        l->plan_total = cur;
        // the offsets do: values per move, their exclusive sum over the chain, the offsets
        int rc0;
        if (!l->d_roff && ((rc0 = dev_alloc(ctx, &l->d_roff, nmol)) || (rc0 = dev_alloc(ctx, &l->d_toff, nmol)) ||
                           (rc0 = dev_alloc(ctx, &l->d_moff, nmol))))
            return rc0;
        mc_plan_counts<<<ceil_div(nmol, 256), 256, 0, st>>>(nmol, l->d_order, box->d_naf, box->d_nal, flags,
                                                            l->d_toff);
        if ((rc0 = ledger_scan<int>(ctx, l, l->d_toff, l->d_moff, nmol))) return rc0;
        mc_plan_offsets<<<ceil_div(nmol, 256), 256, 0, st>>>(nmol, l->d_order, box->d_naf, box->d_nal, flags,
                                                             l->d_roff, l->d_toff, l->d_moff);
        l->plan_flags = flags;
        l->plan_ready = true;
    }
    // step sizes of every type for this cycle: the first move of a type fixes a zero
    // step (molfit.py:375-392), and every move of the type then uses it
    // pinned host block of the cycle's small transfers: the device result block
    // (16 + 2 nmt ints), the no-change flag, the step sizes per type (2 nmt floats).  The
    // previous call ended with a synchronisation, so nothing reads it now.
    const int nwords = (nmol + 31) / 32;
    // back-to-back cycles: per cycle its result block and adapted steps (hw words), fdiffs
    // and acceptance bits; before them the constants and first steps (5 nmt words)
    const int hw = 16 + 4 * nmt;
    const size_t bres_words = batch ? 5 * (size_t)nmt + (size_t)ncycles * ((size_t)hw + nmol + nwords) : 0;
    const size_t hio_need = sizeof(int) * (20 + 4 * (size_t)nmt) + sizeof(float) * (size_t)nmol + sizeof(int) * nwords +
                            sizeof(int) * bres_words;
    if (l->hio_cap < hio_need) {
        if (l->hio) HIP_CHECK(ctx, hipHostFree(l->hio));
        l->hio = nullptr;
        l->hio_cap = 0;
        HIP_CHECK(ctx, hipHostMalloc((void **)&l->hio, hio_need));
        l->hio_cap = hio_need;
    }
    int *hblk = l->hio;
    std::fill(hblk, hblk + 20 + 4 * nmt, 0);
    int *hpc = hblk, *rej = hblk + 16, *hflag = hblk + 9;
    const float *hres = (const float *)(hblk + 12);
    float *wtype = (float *)(hblk + 20 + 2 * nmt);
    // the exact fdiff of every move and the acceptance bits, for fadd on the host
    float *hfd = (float *)(hblk + 20 + 4 * nmt);
    const unsigned *hbits = (const unsigned *)(hfd + nmol);
    int *hbres = (int *)(hbits + nwords);   // the batch block (bres_words)
    std::vector<int> att(l->att);
    std::vector<float> wr(wholerot, wholerot + nmt), wt(wholetrans, wholetrans + nmt);
    for (int mt = 0; mt < nmt; ++mt) {
        if (l->type_nat[mt] == 0) continue;   // no molecule of this type
        if (p->stepr > 0.0f && l->type_nat[mt] > 1) {
            if (wr[mt] <= 0.0f) wr[mt] = rotstepmin;
        } else {
            wr[mt] = 0.0f;
        }
        if (p->stepm > 0.0f) {
            if (wt[mt] <= 0.0f) wt[mt] = trstepmin;
        } else {
            wt[mt] = 0.0f;
        }
        wtype[2 * mt] = wr[mt];
        wtype[2 * mt + 1] = wt[mt];
    }
    int total = l->plan_total;
    // the ran1 stream of the cycle: drawn ahead by the previous call when it starts from
    // this state, else drawn now; both in pinned memory
    // windows of the call's values plus room for atogen's origin draws (one per molecule)
    // before the next call
    const int wlen = ncycles * total + nmol + 256;
    if (l->pin_cap < (size_t)wlen + 2 * nmt + 8) {
        gen_join(l);
        l->gen_buf = -1;
        if (l->ev_up) HIP_CHECK(ctx, hipEventSynchronize(l->ev_up));
        for (int i = 0; i < 3; ++i) {
            if (l->pin[i]) HIP_CHECK(ctx, hipHostFree(l->pin[i]));
            l->pin[i] = nullptr;
            l->win_len[i] = 0;
        }
        l->pin_cap = 0;
        l->pin_ahead = -1;
        for (int i = 0; i < 3; ++i)
            HIP_CHECK(ctx, hipHostMalloc((void **)&l->pin[i], sizeof(float) * ((size_t)wlen + 2 * nmt + 8)));
        l->pin_cap = (size_t)wlen + 2 * nmt + 8;
    }
    Ran1 rng;
    ran1_load(rng, ran1_state);
    const Ran1 rng0 = rng;
    // the cycle's stream: in the window the previous call uploaded when ran1_state lies in
    // it, else in a window drawn now
    int use = -1, off = 0;
    bool uploaded = false;   // the window is in l->d_stream[use], behind the event ev_up
    {
        const int o = window_offset(l, l->pin_ahead, ran1_state, ncycles * total);
        if (o >= 0) {
            use = l->pin_ahead;
            off = o;
            uploaded = l->ahead_uploaded;
        }
    }
    if (use < 0) {
        gen_finish(l);
        use = l->pin_ahead == 0 ? 1 : 0;
        if (l->ev_up) HIP_CHECK(ctx, hipEventSynchronize(l->ev_up));   // no upload still reads the pinned buffers
        fill_window(rng, l->pin[use], wlen, l->win_ck[use]);
        l->win_len[use] = wlen;
    }
    l->pin_ahead = -1;
    l->ahead_uploaded = false;
    window_state(l, use, off + ncycles * total, rng);   // the state after the call's stream
    // the device copies of the windows
    if (l->dstream_cap < (size_t)wlen + 8) {
        gen_finish(l);
        HIP_CHECK(ctx, hipStreamSynchronize(st));
        if (l->ev_up) HIP_CHECK(ctx, hipEventSynchronize(l->ev_up));
        const size_t cap = (size_t)wlen + wlen / 8 + 64;
        for (int i = 0; i < 3; ++i) {
            int rc0;
            if ((rc0 = dev_alloc(ctx, &l->d_stream[i], cap))) return rc0;
        }
        l->dstream_cap = cap;
        uploaded = false;
    }
    // back-to-back cycles: every cycle's values are in window `use`; while they run, a host
    // thread draws the window of the next call (nb) from this call's end state
    int nb = -1;
    int *dbres = nullptr;
    if (batch) {
        gen_finish(l);
        if (l->ev_up) HIP_CHECK(ctx, hipEventSynchronize(l->ev_up));
        if (l->d_bres_cap < bres_words) {
            l->d_bres_cap = 0;
            int rc0;
            if ((rc0 = dev_alloc(ctx, &l->d_bres, bres_words))) return rc0;
            l->d_bres_cap = bres_words;
        }
        dbres = l->d_bres;
        if (!l->up_stream) HIP_CHECK(ctx, hipStreamCreateWithFlags(&l->up_stream, hipStreamNonBlocking));
        if (!l->ev_up) HIP_CHECK(ctx, hipEventCreateWithFlags(&l->ev_up, hipEventDisableTiming));
        nb = (use + 1) % 3;
        gen_start(l, nb, rng, wlen, true);
    }

    // device buffers.  The parallel resolution keeps the terms of an exact sum (the
    // longest ledger list) in LDS for each of its waves.
    const int nslots = l->nslots;
    const int maxc = std::max(1, l->maxcnt);
    const size_t lds_pass = sizeof(float2) * 4 * (size_t)maxc;
    const bool fast = lds_pass <= 160 * 1024;
    Scratch S, q;
    auto layout = [&](Scratch &x) {
        x.take<float4>(natom); x.take<float4>(natom);
        x.take<int>(8);
        x.take<float>(2 * nmt);
        x.take<float>(natom); x.take<float>(natom); x.take<float>(natom);
        x.take<float>(nmol); x.take<float>(nmol); x.take<float>(nmol);
        x.take<float>(nslots); x.take<float>(nslots); x.take<float>(nslots);
        x.take<unsigned>((nmol + 31) / 32); x.take<float>(nmol);
        x.take<float>(nslots);
        x.take<float2>(fast ? nmol : 0);
        x.take<float>(da ? 2 * (size_t)nmol : 0); x.take<float>(da ? nmol : 0);
        x.take<uint16_t>(want_x && !(rec && rec->x) ? 2 * (size_t)nmol * ldx : 0);
        x.take<unsigned char>(nmol); x.take<int>(nmol); x.take<int>(nmol); x.take<int>(16 + 2 * nmt);
    };
    layout(q);
    S.base = (char *)ctx_scratch(ctx, q.off);
    if (!S.base) return EPSR_ERR_HIP;
    float4 *Pold = S.take<float4>(natom), *Pnew = S.take<float4>(natom);
    // [2], [3] flags of a redone cycle, [4..5] table extents
    int *dtmp = S.take<int>(8);
    // the ran1 stream of the cycle (in the ledger) and, for more than STEP_TYPES molecule
    // types, the step sizes of every type (else they go by value to mc_trial)
    float *dstream = l->d_stream[use] + off;
    float *dwtype = S.take<float>(2 * nmt);
    if (batch) dwtype = (float *)(dbres + 3 * nmt);   // the steps the device adapts between cycles
    float *txa = S.take<float>(natom), *tya = S.take<float>(natom), *tza = S.take<float>(natom);
    float *txadd = S.take<float>(nmol), *tyadd = S.take<float>(nmol), *tzadd = S.take<float>(nmol);
    float *dA = S.take<float>(nslots), *dB = S.take<float>(nslots), *dL = S.take<float>(nslots);
    unsigned *dacc = S.take<unsigned>((nmol + 31) / 32);
    float *dfdiff = S.take<float>(nmol);
    float *ddelta = S.take<float>(nslots);
    float2 *dthr = S.take<float2>(fast ? nmol : 0);
    float *dy = S.take<float>(da ? 2 * (size_t)nmol : 0);
    float *dshift = da ? S.take<float>(nmol) : (S.take<float>(0), nullptr);
    uint16_t *dX = S.take<uint16_t>(want_x && !(rec && rec->x) ? 2 * (size_t)nmol * ldx : 0);
    unsigned char *dstate = S.take<unsigned char>(nmol);
    int *dW0 = S.take<int>(nmol), *dW1 = S.take<int>(nmol);
    // the cycle's counters and results, copied to the host in one transfer: [0..7] list
    // counts of passes 0..2, exact sums, passes of the finisher; [8] stage-1 rejections;
    // [9] moves that did not change (| TABLES_STALE); [12..15] fadd and the acceptances of
    // the one-workgroup resolution;
    // [16..] rejections per type
    int *dblk = S.take<int>(16 + 2 * nmt);
    int *dpc = dblk, *dnrej = dblk + 8, *dnochange = dblk + 9, *drej = dblk + 16;
    float *dres = (float *)(dblk + 12);
    ARG_CHECK(ctx, !da || fast, "epsr_molfit_da: delayed acceptance needs the parallel resolution (LDS too small)");

    // the packed tables: the ledger keeps the packing of the arrays it last measured, and
    // mc_trial verifies it against them (a cycle that finds them changed in place commits
    // nothing and runs again with a fresh packing); else they are measured and packed now
    PreparedTables pt;
    int rc;
    TabCheck tc{tb->xavs, tb->xavs1, tb->ecoul, tb->ecoul1, nullptr, tb->ld, box->npairs,
                std::min(tb->npt2, tb->ld - 1), 0, 0};
    if (tab_known(l, tb) && l->tab_packed) {
        tables_describe(box, tb, l->tab_last, l->tab_lastnear, l->d_tab, pt);
        tc.T = l->d_tab;
        tc.nte = pt.tab.nte;
        tc.nnear = pt.tab.nnear;
    } else {
        if ((rc = ledger_tables(ctx, box, tb, l))) return rc;
        int ext[2];
        if ((rc = prepare_tables(ctx, box, tb, l->d_tab, dtmp + 4, pt, ext))) return rc;
        tab_remember(l, tb, ext[0], ext[1]);
        l->tab_packed = true;
    }
    const bool steps_by_value = !batch && nmt <= STEP_TYPES;
    StepArgs sa{};
    if (steps_by_value)
        for (int j = 0; j < 2 * nmt; ++j) sa.w[j] = wtype[j];
    if (uploaded) {
        HIP_CHECK(ctx, hipStreamWaitEvent(st, l->ev_up, 0));
    } else {
        HIP_CHECK(ctx, hipMemcpyAsync(dstream, l->pin[use] + off, sizeof(float) * (size_t)ncycles * total,
                                      hipMemcpyHostToDevice, st));
    }
    if (batch) {
        // attempts and molecules per type, and the first cycle's steps (a type without
        // molecules keeps its given steps, which the adaptation still clamps)
        float *w0 = (float *)(hbres + 3 * nmt);
        for (int mt = 0; mt < nmt; ++mt) {
            hbres[2 * mt] = att[2 * mt];
            hbres[2 * mt + 1] = att[2 * mt + 1];
            hbres[2 * nmt + mt] = l->type_nat[mt];
            w0[2 * mt] = wr[mt];
            w0[2 * mt + 1] = wt[mt];
        }
        HIP_CHECK(ctx, hipMemcpyAsync(dbres, hbres, sizeof(int) * 5 * nmt, hipMemcpyHostToDevice, st));
    } else if (!steps_by_value) {
        HIP_CHECK(ctx, hipMemcpyAsync(dwtype, wtype, sizeof(float) * 2 * nmt, hipMemcpyHostToDevice, st));
    }
    HIP_CHECK(ctx, hipMemsetAsync(dblk + 8, 0, 4 * sizeof(int), st));
    mc_trial<<<ceil_div(nmol, 128), 128, 0, st>>>(nmol, l->d_order, box->d_naf, box->d_nal, c->xa, c->ya, c->za, c->xadd,
                                                  c->yadd, c->zadd, dstream, l->d_roff, l->d_toff, nullptr, nullptr,
                                                  box->d_mtype, steps_by_value ? nullptr : dwtype, sa, txa, tya, tza,
                                                  txadd, tyadd, tzadd, dnochange, tc, drej, 2 * nmt, dpc, 8);

    int nexact = nmol, npasses = 0;
    // B..D of the cycle from trial positions in txa..tzadd and the stream at stream; the
    // commits are skipped when *redo is set
    // slot >= 0: cycle `slot` of a batch, whose fdiffs and acceptance bits stay on the device
    auto run = [&](const float *stream, const int *moff, const int *roff, const int *toff, const int *redo,
                   int slot) -> int {
        int rc1;
        bool joined = false;
        abs_positions2<<<ceil_div(2 * (long)natom, 256), 256, 0, st>>>(
            natom, c->xa, c->ya, c->za, c->xadd, c->yadd, c->zadd, txa, tya, tza, txadd, tyadd, tzadd, box->d_amol,
            box->d_itype, Pold, Pnew);
        if (want_x) {
            uint16_t *X = (rec && rec->x) ? rec->x : dX;
            if ((rc1 = surr_descriptors_launch(ctx, box, l, Pold, Pnew, dpar, ldx, X, nullptr))) return rc1;
            if (use_net && (rc1 = surr_forward_launch(ctx, da->mlp, 2 * nmol, X, ldx, dy))) return rc1;
        }
        if (da)
            da_stage1<<<ceil_div(nmol, 256), 256, 0, st>>>(nmol, l->d_order, dy, da->des, temp, da->seed, dshift,
                                                          rec ? rec->des : nullptr, rec ? rec->pass : nullptr, dnrej);
        timer_begin(ctx, "mc_pair");
        // three-atom molecules: the partners' atoms loaded by the wave together, staged in
        // 192 float4 of LDS per wave after the tables
        const size_t coop_off = (pt.lds_bytes + 15) & ~(size_t)15;
        const size_t coop_lds = coop_off + (size_t)(box->pair_block / 64) * 192 * sizeof(float4);
        if (pt.in_lds && box->max_atoms_per_mol <= 3 && coop_lds <= 160 * 1024) {
            int g = xcd_grid(persistent_blocks(coop_lds, (long long)nslots + l->nbv, box->pair_block));
            mc_pairs<true, true><<<g, box->pair_block, coop_lds, st>>>(
                nslots, l->nbv, l->d_bvslot, l->d_srec, l->d_entry, l->d_pos, l->d_ergpair, Pold, Pnew, pt.tab,
                box->npairs, dshift, dA, dB, dL, (int)coop_off);
        } else if (pt.in_lds) {
            int g = xcd_grid(persistent_blocks(pt.lds_bytes, (long long)nslots + l->nbv, box->pair_block));
            mc_pairs<true, false><<<g, box->pair_block, pt.lds_bytes, st>>>(
                nslots, l->nbv, l->d_bvslot, l->d_srec, l->d_entry, l->d_pos, l->d_ergpair, Pold, Pnew, pt.tab,
                box->npairs, dshift, dA, dB, dL, 0);
        } else {
            int g = xcd_grid(persistent_blocks(0, (long long)nslots + l->nbv, box->pair_block));
            mc_pairs<false, false><<<g, box->pair_block, 0, st>>>(
                nslots, l->nbv, l->d_bvslot, l->d_srec, l->d_entry, l->d_pos, l->d_ergpair, Pold, Pnew, pt.tab,
                box->npairs, dshift, dA, dB, dL, 0);
        }
        timer_end(ctx, "mc_pair");
        if (rec && rec->target)
            mc_record_target<<<nmol, 64, 0, st>>>(nmol, l->d_order, l->d_pos, l->d_rowstart, l->d_cnt,
                                                  l->d_partner, dA, dL, dshift, rec->target);
        if (fast) {
            // the deltas, the thresholds and the first decisions (every earlier partner
            // undecided); the moves left go to dW0 (count dpc[0])
            mc_prepare<<<xcd_grid(ceil_div(nmol, 4)), 256, 0, st>>>(nmol, l->d_rowk, l->d_pn, l->d_mw, dA, dB, dL,
                                                                    stream, moff, temp, dshift, ddelta, dthr, dstate,
                                                                    dW0, dpc + 0);
            const int nwf = std::max(1, std::min(16, (int)((150 * 1024) / (sizeof(float2) * (size_t)maxc))));
            const int gpass = xcd_grid(std::min(ceil_div(nmol, 4), 2048));
            timer_begin(ctx, "mc_resolve");
            // grid-wide passes over the lists (dW0 -> dW1 -> dW0), then the finisher; a capped
            // finisher (the test hook) follows the first decisions alone
            const bool capped = box->resolve_passes >= 0;
            if (!capped) {
                mc_resolve_pass<<<gpass, 256, lds_pass, st>>>(nmol, l->d_rowk, l->d_pn, l->d_mw, dA, dB, dL, ddelta,
                                                              dthr, stream, moff, temp, dshift, maxc, 0, dW0, dpc + 0,
                                                              dW1, dpc + 1, dstate, dpc + 3);
                mc_resolve_pass<<<gpass, 256, lds_pass, st>>>(nmol, l->d_rowk, l->d_pn, l->d_mw, dA, dB, dL, ddelta,
                                                              dthr, stream, moff, temp, dshift, maxc, 0, dW1, dpc + 1,
                                                              dW0, dpc + 2, dstate, dpc + 3);
            }
            mc_resolve_finish<<<1, 64 * nwf, sizeof(float2) * nwf * (size_t)maxc, st>>>(
                l->d_rowk, l->d_pn, l->d_mw, dA, dB, dL, ddelta, dthr, stream, moff, temp, dshift, maxc, dW0,
                dpc + (capped ? 0 : 2), dW1, dstate, dpc + 3, dpc + 4, nmol, dacc,
                capped ? box->resolve_passes : nmol + 1, dpc + 5);
            timer_end(ctx, "mc_resolve");
            // fadd only reports: the exact fdiff of the accepted moves is formed on the side
            // stream beside the commits, and summed on the host (a sequential float sum is
            // several times faster on a CPU core than on one GPU lane)
            hipStream_t side = nullptr;
            if ((rc1 = ctx_side(ctx, &side))) return rc1;
            HIP_CHECK(ctx, hipEventRecord(ctx->ev_fork, st));
            HIP_CHECK(ctx, hipStreamWaitEvent(side, ctx->ev_fork, 0));
            float *fdo = dfdiff;
            unsigned *bo = nullptr;
            if (slot >= 0) {
                fdo = (float *)(dbres + 5 * nmt + (size_t)ncycles * hw) + (size_t)slot * nmol;
                bo = (unsigned *)(dbres + 5 * nmt + (size_t)ncycles * (hw + nmol)) + (size_t)slot * nwords;
            }
            mc_fdiff_accepted<<<xcd_grid(ceil_div(nmol, 4)), 256, sizeof(float2) * 4 * (size_t)maxc, side>>>(
                nmol, l->d_rowk, l->d_pn, l->d_mw, dA, dB, dL, dacc, maxc, fdo, bo, nwords);
            if (slot < 0) {
                HIP_CHECK(ctx, hipMemcpyAsync(hfd, dfdiff, sizeof(float) * nmol, hipMemcpyDeviceToHost, side));
                HIP_CHECK(ctx,
                          hipMemcpyAsync((void *)hbits, dacc, sizeof(unsigned) * nwords, hipMemcpyDeviceToHost, side));
            }
            HIP_CHECK(ctx, hipEventRecord(ctx->ev_join, side));
            joined = true;
        } else {
            size_t lds = sizeof(unsigned) * ((nmol + 31) / 32) + sizeof(float) * 2 * (size_t)maxc + 64;
            ARG_CHECK(ctx, lds <= 160 * 1024, "epsr_molfit: %d molecules x %d ledger slots exceed LDS", nmol, maxc);
            timer_begin(ctx, "mc_resolve");
            mc_resolve<<<1, 256, lds, st>>>(nmol, l->d_order, l->d_pos, l->d_rowstart, l->d_cnt, l->d_partner,
                                            l->d_mirror, dA, dB, dL, stream, moff, temp, maxc, dacc, dfdiff, dres);
            timer_end(ctx, "mc_resolve");
        }
        mc_commit_ledger<<<xcd_grid(ceil_div(nmol, 4)), 256, 0, st>>>(nmol, l->d_rowk, l->d_pn, l->d_mw, l->d_entry,
                                                                   dacc, dA, dB, redo, l->d_ergpair);
        mc_commit_mols<<<ceil_div(nmol, 128), 128, 0, st>>>(nmol, l->d_order, box->d_naf, box->d_nal, box->d_mtype,
                                                            dacc, roff, toff, txa, tya, tza, txadd, tyadd, tzadd, c->xa,
                                                            c->ya, c->za, c->xadd, c->yadd, c->zadd, nstepreject,
                                                            p->stepr > 0.0f ? 1 : 0, redo, drej);
        if (joined) HIP_CHECK(ctx, hipStreamWaitEvent(st, ctx->ev_join, 0));
        if (slot < 0)
            HIP_CHECK(ctx, hipMemcpyAsync(hblk, dblk, sizeof(int) * (16 + 2 * (size_t)nmt), hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipGetLastError());
        return EPSR_OK;
    };
    if ((rc = run(dstream, l->d_moff, l->d_roff, l->d_toff, dnochange, batch ? 0 : -1))) return rc;
    if (batch) {
        // cycles 1..n-1, the steps adapted on the device after every cycle; one wait at the end
        auto adapt = [&](int k) {
            mc_adapt<<<1, 64, 0, st>>>(nmt, dbres, dbres + 2 * nmt, dblk, dwtype, dbres + 5 * nmt + (size_t)k * hw,
                                       p->rejrate, p->ireset, p->stepr > 0.0f ? 1 : 0, p->stepm > 0.0f ? 1 : 0,
                                       rotstepmax, rotstepmin, trstepmax, trstepmin);
        };
        adapt(0);
        for (int k = 1; k < ncycles; ++k) {
            const float *sk = dstream + (size_t)k * total;
            mc_trial<<<ceil_div(nmol, 128), 128, 0, st>>>(nmol, l->d_order, box->d_naf, box->d_nal, c->xa, c->ya,
                                                          c->za, c->xadd, c->yadd, c->zadd, sk, l->d_roff, l->d_toff,
                                                          nullptr, nullptr, box->d_mtype, dwtype, sa, txa, tya, tza,
                                                          txadd, tyadd, tzadd, dnochange, tc, drej, 2 * nmt, dpc, 8);
            if ((rc = run(sk, l->d_moff, l->d_roff, l->d_toff, dnochange, k))) return rc;
            adapt(k);
        }
        HIP_CHECK(ctx, hipMemcpyAsync(hbres + 5 * nmt, dbres + 5 * nmt, sizeof(int) * (bres_words - 5 * (size_t)nmt),
                                      hipMemcpyDeviceToHost, st));
        if ((rc = ctx_wait(ctx))) return rc;
        // the next call's window: its thread goes on drawing and uploading after the return
        // (the next call joins it)
        l->pin_ahead = nb;
        l->ahead_uploaded = true;
        // the cycles in order.  The first whose trial moves did not all change (or that
        // found the tables changed in place) committed nothing, and neither did any after it
        const int *slots = hbres + 5 * nmt;
        const float *fds = (const float *)(slots + (size_t)ncycles * hw);
        const unsigned *bts = (const unsigned *)(fds + (size_t)ncycles * nmol);
        int done = 0, nacc = 0;
        long long nacc_all = 0;
        for (; done < ncycles; ++done) {
            const int *b = slots + (size_t)done * hw;
            if (b[9] != 0) break;
            if (b[5] != 0)
                return set_error(ctx, EPSR_ERR_HIP,
                                 "epsr_molfit_cycles: cycle %d left %d moves undecided after %d passes; coordinates "
                                 "and ledger are not the chain's",
                                 done, b[5], 3 + b[4]);
            const float *fd = fds + (size_t)done * nmol;
            const unsigned *bw = bts + (size_t)done * nwords;
            float fadd = 0.0f;
            nacc = 0;
            for (int w = 0; w < nwords; ++w)
                for (unsigned bb = bw[w]; bb; bb &= bb - 1u) {   // the accepted moves of the word, in order
                    fadd = fadd + fd[32 * w + __builtin_ctz(bb)];
                    ++nacc;
                }
            *fadds = *fadds + fadd / (float)nmol;
            nacc_all += nacc;
        }
        if (done > 0) {
            const int *b = slots + (size_t)(done - 1) * hw;
            const float *after = (const float *)(b + 16 + 2 * nmt);
            for (int mt = 0; mt < nmt; ++mt) {
                wholerot[mt] = after[2 * mt];
                wholetrans[mt] = after[2 * mt + 1];
                nwhole[2 * mt] = att[2 * mt];
                nwhole[2 * mt + 1] = b[16 + 2 * mt];
                nmovem[2 * mt] = att[2 * mt + 1];
                nmovem[2 * mt + 1] = b[16 + 2 * mt + 1];
            }
            Ran1 r;
            window_state(l, use, off + done * total, r);   // the state at the start of cycle `done`
            ran1_store(r, ran1_state);
            if (stats) {
                stats->moves = nmol;
                stats->accepted = nacc;
                stats->exact_resolutions = b[3];
                stats->pushed_entries = 0;
                stats->resolver_passes = (box->resolve_passes >= 0 ? 1 : 3) + b[4];   // (the first in mc_prepare)
                stats->resolver_bytes = (double)nslots * 13.0 + (double)nmol * (sizeof(float2) + 1);
                stats->retries = 0;
                stats->ms_pairs = ctx->timing ? epsr_ctx_last_kernel_ms(ctx, "mc_pair") : -1.0;
                stats->ms_resolve = ctx->timing ? epsr_ctx_last_kernel_ms(ctx, "mc_resolve") : -1.0;
                stats->stage1_rejected = b[8];
                stats->cycles = done;
                stats->accepted_total = nacc_all;
            }
        }
        if (done < ncycles) {
            // that cycle on its own (it doubles movem's step, or measures the tables again),
            // then the rest
            epsr_mc_stats_t s1{};
            if ((rc = molfit_impl(ctx, box, c, tb, l, p, 1, wholerot, wholetrans, nwhole, nmovem, nstepreject,
                                  ran1_state, fadds, da, rec, &s1)))
                return rc;
            if (done + 1 < ncycles) {
                if ((rc = molfit_impl(ctx, box, c, tb, l, p, ncycles - done - 1, wholerot, wholetrans, nwhole, nmovem,
                                      nstepreject, ran1_state, fadds, da, rec, stats)))
                    return rc;
                if (stats) {
                    stats->cycles += done + 1;
                    stats->accepted_total += nacc_all + s1.accepted_total;
                }
            } else if (stats) {
                *stats = s1;
                stats->cycles = done + 1;
                stats->accepted_total = nacc_all + s1.accepted_total;
            }
        }
        return EPSR_OK;
    }
    // while the GPU runs the cycle: the next cycle's stream, from this one's end state
    // (the window the thread of the previous call drew, when this cycle's end state lies
    // in it), and a thread for the window after, from the next cycle's expected end
    const Ran1 rend = rng;
    {
        int32_t rend_a[35];
        ran1_store(rend, rend_a);
        const int gb = l->gen_buf;
        gen_finish(l);
        int nxt = -1, o2 = 0;
        if (gb >= 0 && gb != use) {
            const int o = window_offset(l, gb, rend_a, total);
            if (o >= 0) {
                nxt = gb;
                o2 = o;
            }
        }
        if (nxt < 0) {
            nxt = (use + 1) % 3;
            if (l->ev_up) HIP_CHECK(ctx, hipEventSynchronize(l->ev_up));
            fill_window(rend, l->pin[nxt], wlen, l->win_ck[nxt]);
            l->win_len[nxt] = wlen;
        }
        // its upload, beside this cycle's kernels (the buffer is not this cycle's)
        hipStream_t side = nullptr;
        if ((rc = ctx_side(ctx, &side))) return rc;
        if (!l->ev_up) HIP_CHECK(ctx, hipEventCreateWithFlags(&l->ev_up, hipEventDisableTiming));
        HIP_CHECK(ctx, hipMemcpyAsync(l->d_stream[nxt], l->pin[nxt], sizeof(float) * l->win_len[nxt],
                                      hipMemcpyHostToDevice, side));
        HIP_CHECK(ctx, hipEventRecord(l->ev_up, side));
        l->pin_ahead = nxt;
        // the third buffer (its last upload finished before this cycle's kernels started)
        Ran1 r3;
        window_state(l, nxt, o2 + total, r3);
        gen_start(l, 3 - use - nxt, r3, wlen);
    }
    if ((rc = ctx_wait(ctx))) return rc;
    l->ahead_uploaded = true;
    if (*hflag & TABLES_STALE) {
        // the tables changed in place since their extents were measured; nothing was
        // committed and no state was updated: measure them and run the cycle again
        l->tab_ok = false;
        return molfit_impl(ctx, box, c, tb, l, p, 1, wholerot, wholetrans, nwhole, nmovem, nstepreject, ran1_state,
                           fadds, da, rec, stats);
    }
    int retries = 0;
    if (*hflag) {
        // movem doubles its step until the centre moves (molfit.py:747-758); the doubled
        // step persists for the molecule type.  Nothing was committed: resolve the chain on
        // the host from the current centres, then run the cycle again with those offsets.
        std::vector<float> cx(nmol), cy(nmol), cz(nmol);
        std::vector<int> horder(nmol);
        HIP_CHECK(ctx, hipMemcpyAsync(horder.data(), l->d_order, sizeof(int) * nmol, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(cx.data(), c->xadd, sizeof(float) * nmol, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(cy.data(), c->yadd, sizeof(float) * nmol, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(cz.data(), c->zadd, sizeof(float) * nmol, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
        rng = rng0;
        std::vector<float> hstream;
        auto draw = [&](int idx) {
            while ((int)hstream.size() <= idx) hstream.push_back(ran1_next(rng));
            return hstream[idx];
        };
        std::vector<int> roff(nmol), toff(nmol), moff(nmol);
        std::vector<float> rstep(nmol), tstep(nmol);
        wr.assign(wholerot, wholerot + nmt);
        wt.assign(wholetrans, wholetrans + nmt);
        int cursor = 0;
        for (int k = 0; k < nmol; ++k) {
            int m = horder[k], mt = box->h_mtype[m];
            int nat = box->h_nal[m] - box->h_naf[m] + 1;
            if (p->stepr > 0.0f && nat > 1) {
                if (wr[mt] <= 0.0f) wr[mt] = rotstepmin;
                roff[k] = cursor;
                rstep[k] = wr[mt];
                cursor += 3;
            } else {
                wr[mt] = 0.0f;
                roff[k] = p->stepr > 0.0f ? -3 : -1;
            }
            if (p->stepm > 0.0f) {
                if (wt[mt] <= 0.0f) wt[mt] = trstepmin;
                float sz = wt[mt];
                while (true) {
                    float nz = cz[m] + sz * (2.0f * draw(cursor) - 1.0f);
                    float ny = cy[m] + sz * (2.0f * draw(cursor + 1) - 1.0f);
                    float nx = cx[m] + sz * (2.0f * draw(cursor + 2) - 1.0f);
                    if (!(nx == cx[m] && ny == cy[m] && nz == cz[m])) break;
                    sz = sz * 2.0f;
                    cursor += 3;
                    retries++;
                }
                toff[k] = cursor;
                tstep[k] = sz;
                wt[mt] = sz;
                cursor += 3;
            } else {
                wt[mt] = 0.0f;
                toff[k] = -1;
            }
            moff[k] = cursor;
            draw(cursor);
            cursor += 1;
        }
        total = cursor;
        hstream.resize(total);
        StreamTemp T(ctx);
        float *dstream2 = (float *)T.alloc(sizeof(float) * (total + 8));
        int *droff = (int *)T.alloc(sizeof(int) * nmol), *dtoff = (int *)T.alloc(sizeof(int) * nmol);
        int *dmoff = (int *)T.alloc(sizeof(int) * nmol);
        float *drstep = (float *)T.alloc(sizeof(float) * nmol), *dtstep = (float *)T.alloc(sizeof(float) * nmol);
        if (!dstream2 || !droff || !dtoff || !dmoff || !drstep || !dtstep) return EPSR_ERR_HIP;
        HIP_CHECK(ctx, hipMemcpyAsync(dstream2, hstream.data(), sizeof(float) * total, hipMemcpyHostToDevice, st));
        HIP_CHECK(ctx, hipMemcpyAsync(droff, roff.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, st));
        HIP_CHECK(ctx, hipMemcpyAsync(dtoff, toff.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, st));
        HIP_CHECK(ctx, hipMemcpyAsync(dmoff, moff.data(), sizeof(int) * nmol, hipMemcpyHostToDevice, st));
        HIP_CHECK(ctx, hipMemcpyAsync(drstep, rstep.data(), sizeof(float) * nmol, hipMemcpyHostToDevice, st));
        HIP_CHECK(ctx, hipMemcpyAsync(dtstep, tstep.data(), sizeof(float) * nmol, hipMemcpyHostToDevice, st));
        // counters of the redone cycle start at zero; its trial moves all change
        HIP_CHECK(ctx, hipMemsetAsync(dblk + 8, 0, 4 * sizeof(int), st));
        HIP_CHECK(ctx, hipMemsetAsync(dtmp + 2, 0, 2 * sizeof(int), st));
        TabCheck none = tc;
        none.T = nullptr;
        mc_trial<<<ceil_div(nmol, 128), 128, 0, st>>>(nmol, l->d_order, box->d_naf, box->d_nal, c->xa, c->ya, c->za,
                                                      c->xadd, c->yadd, c->zadd, dstream2, droff, dtoff, drstep,
                                                      dtstep, box->d_mtype, steps_by_value ? nullptr : dwtype, sa, txa,
                                                      tya, tza, txadd, tyadd, tzadd, dtmp + 2, none, drej, 2 * nmt, dpc, 8);
        if ((rc = run(dstream2, dmoff, droff, dtoff, dtmp + 3, -1))) return rc;
        HIP_CHECK(ctx, hipStreamSynchronize(st));
    }
    if (fast) {
        nexact = hpc[3];
        npasses = (box->resolve_passes >= 0 ? 1 : 3) + hpc[4];
        if (hpc[5] != 0)
            return set_error(ctx, EPSR_ERR_HIP,
                             "epsr_molfit: the resolution left %d moves undecided after %d passes; coordinates and "
                             "ledger are not the chain's",
                             hpc[5], npasses);
    }
    // counters and the step-size adaptation (molfit.py:465-536)
    for (int mt = 0; mt < nmt; ++mt) {
        nwhole[2 * mt] = att[2 * mt];
        nwhole[2 * mt + 1] = rej[2 * mt];
        nmovem[2 * mt] = att[2 * mt + 1];
        nmovem[2 * mt + 1] = rej[2 * mt + 1];
    }
    if (p->ireset != 1) {
        for (int mt = 0; mt < nmt; ++mt) {
            float rat = adapt_ratio(p->rejrate, nwhole[2 * mt], nwhole[2 * mt + 1]);
            wr[mt] = wr[mt] * rat;
            if (wr[mt] > rotstepmax) wr[mt] = rotstepmax;
            if (wr[mt] < rotstepmin && wr[mt] != 0.0f) wr[mt] = rotstepmin;
            rat = adapt_ratio(p->rejrate, nmovem[2 * mt], nmovem[2 * mt + 1]);
            wt[mt] = wt[mt] * rat;
            if (wt[mt] > trstepmax) wt[mt] = trstepmax;
            if (wt[mt] < trstepmin && wt[mt] != 0.0f) wt[mt] = trstepmin;
        }
    }
    std::copy(wr.begin(), wr.end(), wholerot);
    std::copy(wt.begin(), wt.end(), wholetrans);
    // fadd = fadd + fdiff over the accepted moves in chain order (molfit.py:405)
    float fadd = hres[0];
    int nacc = (int)hres[1];
    if (fast) {
        fadd = 0.0f;
        nacc = 0;
        for (int w = 0; w < nwords; ++w)
            for (unsigned b = hbits[w]; b; b &= b - 1u) {   // the accepted moves of the word, in order
                fadd = fadd + hfd[32 * w + __builtin_ctz(b)];
                ++nacc;
            }
    }
    *fadds = *fadds + fadd / (float)nmol;
    ran1_store(rng, ran1_state);
    if (stats) {
        stats->moves = nmol;
        stats->accepted = nacc;
        stats->exact_resolutions = nexact;
        stats->pushed_entries = 0;
        stats->resolver_passes = npasses;
        // the first pass reads every slot once (partner, its chain position, delta, its
        // state) and each move's thresholds, and writes the decisions
        stats->resolver_bytes = fast ? (double)nslots * 13.0 + (double)nmol * (sizeof(float2) + 1) : 0.0;
        stats->retries = retries;
        stats->ms_pairs = ctx->timing ? epsr_ctx_last_kernel_ms(ctx, "mc_pair") : -1.0;
        stats->ms_resolve = ctx->timing ? epsr_ctx_last_kernel_ms(ctx, "mc_resolve") : -1.0;
        stats->stage1_rejected = hblk[8];
        stats->cycles = 1;
        stats->accepted_total = nacc;
    }
    return EPSR_OK;
}

extern "C" int epsr_molfit(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                           epsr_ledger *l, const epsr_mc_params_t *p, float *wholerot, float *wholetrans, int *nwhole,
                           int *nmovem, int *nstepreject, int32_t *ran1_state, float *fadds, epsr_mc_stats_t *stats)
{
    return molfit_impl(ctx, box, c, tb, l, p, 1, wholerot, wholetrans, nwhole, nmovem, nstepreject, ran1_state, fadds,
                       nullptr, nullptr, stats);
}

extern "C" int epsr_molfit_cycles(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c,
                                  const epsr_tables_t *tb, epsr_ledger *l, const epsr_mc_params_t *p, int ncycles,
                                  float *wholerot, float *wholetrans, int *nwhole, int *nmovem, int *nstepreject,
                                  int32_t *ran1_state, float *fadds, epsr_mc_stats_t *stats)
{
    return molfit_impl(ctx, box, c, tb, l, p, ncycles, wholerot, wholetrans, nwhole, nmovem, nstepreject, ran1_state,
                       fadds, nullptr, nullptr, stats);
}

extern "C" int epsr_molfit_da(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                              epsr_ledger *l, const epsr_mc_params_t *p, float *wholerot, float *wholetrans,
                              int *nwhole, int *nmovem, int *nstepreject, int32_t *ran1_state, float *fadds,
                              const epsr_da_t *da, const epsr_mc_record_t *rec, epsr_mc_stats_t *stats)
{
    return molfit_impl(ctx, box, c, tb, l, p, 1, wholerot, wholetrans, nwhole, nmovem, nstepreject, ran1_state, fadds,
                       da, rec, stats);
}


