// press: the whole-box intermolecular energy and virial pressure over the ledger's
// partner lists (src/epsr26/library/ato_manipulation_rev4.py:3901-4063).
//
// Work split: one workgroup per molecule m, one lane per ledger slot of m.  A lane whose
// partner n comes later in the numbering forms the pair's float32 energy and virial over
// every atom pair (atoms of m outer, atoms of n inner) exactly as the reference orders
// them; other lanes leave 0.  The host adds the per-slot values in the reference's order
// (molecule, then ledger slot): the energy as a float32 running sum, the virial as a
// double one.  Those two sums are sequential by definition, and a host loop is cheaper
// than the same chain of dependent adds on one GPU lane.  Two things shorten it:
//  - a slot whose energy and virial are both zero is left out.  Adding a zero changes a
//    running sum only when the sum is -0, and neither sum can be: each starts at +0, and
//    a round-to-nearest sum that is exactly zero is +0 unless both terms are -0.  About
//    half the slots (the partners numbered lower) are such zeros;
//  - the remaining (energy, virial) pairs are gathered on the GPU into molecule order, one
//    contiguous array, and copied to the host in chunks, each summed as soon as it lands
//    while the next is in flight.
#include <algorithm>

#include "common.h"

namespace {

__global__ void __launch_bounds__(128)
press_slots(int nmol, const int *__restrict__ rowstart, const int *__restrict__ cnt, const int *__restrict__ partner,
            const int *__restrict__ naf, const int *__restrict__ nal, const int *__restrict__ itype,
            const int *__restrict__ npairref, int ntype, const float *__restrict__ xa, const float *__restrict__ ya,
            const float *__restrict__ za, const float *__restrict__ xadd, const float *__restrict__ yadd,
            const float *__restrict__ zadd, const float *__restrict__ xstar, const float *__restrict__ xavs,
            const float *__restrict__ xavs1, const float *__restrict__ ecoul, const float *__restrict__ ecoul1,
            int ld, int npt1, float stepcl, float *__restrict__ slote, float *__restrict__ slotw,
            int *__restrict__ nnz)
{
    __shared__ int nzs;
    if (threadIdx.x == 0) nzs = 0;
    __syncthreads();
    int mine = 0;
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
        mine += (ergold != 0.0f || ergnew != 0.0f) ? 1 : 0;
    }
    if (mine) atomicAdd(&nzs, mine);
    __syncthreads();
    if (threadIdx.x == 0) nnz[im] = nzs;
}

// off[m] = nnz[0] + ... + nnz[m-1] for m <= n (one workgroup of 1024 lanes, each over a
// contiguous run of molecules)
__global__ void __launch_bounds__(1024) press_scan(int n, const int *__restrict__ nnz, int *__restrict__ off)
{
    __shared__ int part[1024];
    const int per = (n + 1023) / 1024, lo = min(n, (int)threadIdx.x * per), hi = min(n, lo + per);
    int t = 0;
    for (int m = lo; m < hi; ++m) t += nnz[m];
    part[threadIdx.x] = t;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // inclusive scan of the run totals
        const int v = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int base = part[threadIdx.x] - t;
    for (int m = lo; m < hi; ++m) {
        off[m] = base;
        base += nnz[m];
    }
    if (threadIdx.x == 1023) off[n] = part[1023];
}

// one wave per molecule m < n: its non-zero slots, in slot order, to out[off[m]...]
__global__ void __launch_bounds__(256)
press_gather(int n, const int *__restrict__ rowstart, const int *__restrict__ cnt, const int *__restrict__ off,
             const float *__restrict__ slote, const float *__restrict__ slotw, float2 *__restrict__ out)
{
    const int lane = threadIdx.x & 63, m = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (m >= n) return;
    const int row = rowstart[m], c = cnt[m];
    const unsigned long long below = (1ull << lane) - 1ull;
    int base = off[m];
    for (int s0 = 0; s0 < c; s0 += 64) {
        const int s = s0 + lane;
        float e = 0.0f, w = 0.0f;
        if (s < c) {
            e = slote[row + s];
            w = slotw[row + s];
        }
        const bool keep = e != 0.0f || w != 0.0f;
        const unsigned long long mask = __ballot(keep);
        if (keep) out[base + __popcll(mask & below)] = make_float2(e, w);
        base += __popcll(mask);
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
        const int nm = nmol - 1;
        // scratch: slot energies and virials, non-zero counts and offsets, the gathered pairs
        const size_t oE = 0, oW = oE + sizeof(float) * (size_t)ns, oN = oW + sizeof(float) * (size_t)ns,
                     oO = oN + sizeof(int) * (size_t)nm, oV = (oO + sizeof(int) * ((size_t)nm + 1) + 15) & ~(size_t)15,
                     sb = oV + sizeof(float2) * (size_t)ns;
        char *scr = (char *)ctx_scratch(ctx, sb);
        if (!scr) return EPSR_ERR_HIP;
        float *dE = (float *)(scr + oE), *dW = (float *)(scr + oW);
        int *dN = (int *)(scr + oN), *dO = (int *)(scr + oO);
        float2 *dV = (float2 *)(scr + oV);
        const float stepcl = box->cellst / box->dmax;
        timer_begin(ctx, "press_slots");
        press_slots<<<nm, 128, 0, st>>>(nmol, l->d_rowstart, l->d_cnt, l->d_partner, box->d_naf, box->d_nal,
                                        box->d_itype, box->d_npairref, box->ntype, c->xa, c->ya, c->za, c->xadd,
                                        c->yadd, c->zadd, tb->xstar, tb->xavs, tb->xavs1, tb->ecoul, tb->ecoul1,
                                        tb->ld, npt1, stepcl, dE, dW, dN);
        HIP_CHECK(ctx, hipGetLastError());
        timer_end(ctx, "press_slots");
        timer_begin(ctx, "press_gather");
        press_scan<<<1, 1024, 0, st>>>(nm, dN, dO);
        HIP_CHECK(ctx, hipGetLastError());
        press_gather<<<(nm + 3) / 4, 256, 0, st>>>(nm, l->d_rowstart, l->d_cnt, dO, dE, dW, dV);
        HIP_CHECK(ctx, hipGetLastError());
        timer_end(ctx, "press_gather");
        const bool want = slote_h || slotw_h;
        const size_t hs = want ? sizeof(float) * 2 * (size_t)ns + sizeof(int) * (2 * (size_t)nmol + 1) : 0;
        const size_t hb = 16 + sizeof(float2) * (size_t)ns + hs;
        char *h = (char *)ctx_pinned(ctx, hb);
        if (!h) return EPSR_ERR_HIP;
        int *htot = (int *)h;
        float2 *hV = (float2 *)(h + 16);
        HIP_CHECK(ctx, hipMemcpyAsync(htot, dO + nm, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
        const long long tot = *htot;
        ARG_CHECK(ctx, tot >= 0 && tot <= ns, "epsr_press: %lld non-zero slots of %d", tot, ns);
        // chunks of at least 4096 pairs, at most 8 of them
        const long long chunk = std::max(4096LL, (tot + 7) / 8);
        int nch = 0;
        for (long long lo = 0; lo < tot; lo += chunk, ++nch) {
            if (!ctx->ev_chunk[nch]) HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->ev_chunk[nch], hipEventDisableTiming));
            const long long n = std::min(chunk, tot - lo);
            HIP_CHECK(ctx, hipMemcpyAsync(hV + lo, dV + lo, sizeof(float2) * n, hipMemcpyDeviceToHost, st));
            HIP_CHECK(ctx, hipEventRecord(ctx->ev_chunk[nch], st));
        }
        for (int k = 0; k < nch; ++k) {
            HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_chunk[k]));
            const long long lo = k * chunk, hi = std::min(tot, lo + chunk);
            for (long long i = lo; i < hi; ++i) {
                sumc = sumc + hV[i].x;
                pt = pt + (double)hV[i].y;
            }
        }
        if (want) {
            float *hE = (float *)(h + 16 + sizeof(float2) * (size_t)ns), *hW = hE + ns;
            int *hrow = (int *)(hW + ns), *hcnt = hrow + nmol + 1;
            HIP_CHECK(ctx, hipMemcpyAsync(hE, dE, sizeof(float) * 2 * (size_t)ns, hipMemcpyDeviceToHost, st));
            HIP_CHECK(ctx, hipMemcpyAsync(hrow, l->d_rowstart, sizeof(int) * (nmol + 1), hipMemcpyDeviceToHost, st));
            HIP_CHECK(ctx, hipMemcpyAsync(hcnt, l->d_cnt, sizeof(int) * nmol, hipMemcpyDeviceToHost, st));
            HIP_CHECK(ctx, hipStreamSynchronize(st));
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
