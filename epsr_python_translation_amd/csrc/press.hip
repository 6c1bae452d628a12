// press: the whole-box intermolecular energy and virial pressure over the ledger's
// partner lists (src/epsr26/library/ato_manipulation_rev4.py:3901-4063).
//
// Work split: one workgroup per molecule m, one lane per ledger slot of m.  A lane whose
// partner n comes later in the numbering forms the pair's float32 energy and virial over
// every atom pair (atoms of m outer, atoms of n inner) exactly as the reference orders
// them; other lanes leave 0.  The two per-slot values go to host memory and the host
// adds them in the reference's order (molecule, then ledger slot): the energy as a
// float32 running sum, the virial as a double one.  Those two sums are sequential by
// definition, and a host loop over a few million values is cheaper than the same chain
// of dependent adds on one GPU lane.
#include "common.h"

namespace {

__global__ void __launch_bounds__(128)
press_slots(int nmol, const int *__restrict__ rowstart, const int *__restrict__ cnt, const int *__restrict__ partner,
            const int *__restrict__ naf, const int *__restrict__ nal, const int *__restrict__ itype,
            const int *__restrict__ npairref, int ntype, const float *__restrict__ xa, const float *__restrict__ ya,
            const float *__restrict__ za, const float *__restrict__ xadd, const float *__restrict__ yadd,
            const float *__restrict__ zadd, const float *__restrict__ xstar, const float *__restrict__ xavs,
            const float *__restrict__ xavs1, const float *__restrict__ ecoul, const float *__restrict__ ecoul1,
            int ld, int npt1, float stepcl, float *__restrict__ slote, float *__restrict__ slotw)
{
    const int im = blockIdx.x;
    const int row = rowstart[im], c = cnt[im];
    const float xc1 = xadd[im], yc1 = yadd[im], zc1 = zadd[im];
    const int a0 = naf[im], a1 = nal[im];
    const float zmax = (float)(npt1 - 1);   // IFIX(rat) + 2 <= npt1  <=>  rat < npt1 - 1 (rat >= 0)
    for (int s = threadIdx.x; s < c; s += blockDim.x) {
        const int jm = partner[row + s];
        float ergold = 0.0f, ergnew = 0.0f;
        if (jm > im) {
            const float xc2 = xadd[jm], yc2 = yadd[jm], zc2 = zadd[jm];
            const float xcdif = dev_mic(xc2 - xc1), ycdif = dev_mic(yc2 - yc1), zcdif = dev_mic(zc2 - zc1);
            const int b0 = naf[jm], b1 = nal[jm];
            for (int ia = a0; ia <= a1; ++ia) {
                const int it = itype[ia];
                const float x1 = xa[ia] + xc1, y1 = ya[ia] + yc1, z1 = za[ia] + zc1;
                for (int ja = b0; ja <= b1; ++ja) {
                    const int icref = npairref[it * ntype + itype[ja]];
                    const float x2 = xa[ja] + xc2, y2 = ya[ja] + yc2, z2 = za[ja] + zc2;
                    const float xdif = dev_mic(x2 - x1), ydif = dev_mic(y2 - y1), zdif = dev_mic(z2 - z1);
                    const float dsq = xdif * xdif + ydif * ydif + zdif * zdif;
                    const float dtest = sqrtf(dsq);
                    const float rat = dtest / stepcl;
                    if (!(rat < zmax)) continue;
                    const int iz = (int)truncf(rat) + 2;
                    const size_t o = (size_t)icref * ld;
                    const float dtdif = dtest - xstar[iz];
                    const float prfac = dtest > 0.0f ? (xcdif * xdif + ycdif * ydif + zcdif * zdif) / dtest : 0.0f;
                    const float enpot = ecoul[o + iz] + ecoul1[o + iz] * dtdif;
                    const float prpot = 0.5f * (ecoul1[o + iz] + ecoul1[o + iz - 1]) * prfac;
                    const float enwt = xavs[o + iz] + xavs1[o + iz] * dtdif;
                    const float prwt = 0.5f * (xavs1[o + iz] + xavs1[o + iz - 1]) * prfac;
                    ergold = ergold + enwt + enpot;
                    // (ergnew + prwt) in float32, + prpot in double, stored to float32
                    const float sw = ergnew + prwt;
                    ergnew = (float)((double)sw + (double)prpot);
                }
            }
        }
        slote[row + s] = ergold;
        slotw[row + s] = ergnew;
    }
}

}  // namespace

extern "C" int epsr_press(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_tables_t *tb,
                          const epsr_ledger *l, int npt1, float simt, float volume, float *energy, float *pressc,
                          double *presst, int ldl, float *slote_h, float *slotw_h)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && tb && l && energy && pressc && presst, "epsr_press: null argument");
    ARG_CHECK(ctx, l->ledger_ready, "epsr_press: no ledger (call epsr_fmolec1)");
    ARG_CHECK(ctx, npt1 >= 2 && npt1 < tb->ld, "epsr_press: npt1 %d outside [2, ld %d)", npt1, tb->ld);
    ARG_CHECK(ctx, !(slote_h || slotw_h) || ldl >= l->nnbrefmax, "epsr_press: ldl %d < nnbrefmax %d", ldl,
              l->nnbrefmax);
    ARG_CHECK(ctx, box->dmax > 0.0f && box->cellst > 0.0f, "epsr_press: box without dmax or cellst");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int nmol = box->nmol, ns = l->nslots;
    float sumc = 0.0f;
    double pt = 0.0;
    if (nmol >= 2 && ns > 0) {
        float *dE = (float *)ctx_scratch(ctx, sizeof(float) * 2 * (size_t)ns);
        if (!dE) return EPSR_ERR_HIP;
        float *dW = dE + ns;
        const float stepcl = box->cellst / box->dmax;
        press_slots<<<nmol - 1, 128, 0, st>>>(nmol, l->d_rowstart, l->d_cnt, l->d_partner, box->d_naf, box->d_nal,
                                               box->d_itype, box->d_npairref, box->ntype, c->xa, c->ya, c->za,
                                               c->xadd, c->yadd, c->zadd, tb->xstar, tb->xavs, tb->xavs1, tb->ecoul,
                                               tb->ecoul1, tb->ld, npt1, stepcl, dE, dW);
        HIP_CHECK(ctx, hipGetLastError());
        const size_t hb = sizeof(float) * 2 * (size_t)ns + sizeof(int) * (2 * (size_t)nmol + 1);
        char *h = (char *)ctx_pinned(ctx, hb);
        if (!h) return EPSR_ERR_HIP;
        float *hE = (float *)h, *hW = hE + ns;
        int *hrow = (int *)(hW + ns), *hcnt = hrow + nmol + 1;
        HIP_CHECK(ctx, hipMemcpyAsync(hE, dE, sizeof(float) * 2 * (size_t)ns, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(hrow, l->d_rowstart, sizeof(int) * (nmol + 1), hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(hcnt, l->d_cnt, sizeof(int) * nmol, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
        for (int m = 0; m < nmol - 1; ++m) {
            const float *e = hE + hrow[m], *w = hW + hrow[m];
            for (int s = 0; s < hcnt[m]; ++s) {
                sumc = sumc + e[s];
                pt = pt + (double)w[s];
            }
        }
        if (slote_h || slotw_h) {
            for (int m = 0; m < nmol; ++m)
                for (int s = 0; s < hcnt[m]; ++s) {
                    const bool last = m == nmol - 1;
                    if (slote_h) slote_h[(size_t)m * ldl + s] = last ? 0.0f : hE[hrow[m] + s];
                    if (slotw_h) slotw_h[(size_t)m * ldl + s] = last ? 0.0f : hW[hrow[m] + s];
                }
        }
    }
    *energy = sumc / (float)nmol;
    *pressc = (float)((((double)((float)nmol * simt)) - pt / 3.0) * 100.0 / (double)(6.0222f * volume));
    *presst = pt;
    return EPSR_OK;
}
