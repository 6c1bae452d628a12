// atogen: pair-distance histograms of one box (the reference's
// src/epsr26/library/ato_manipulation_rev4.py:3640-3899).
//
// Work split
//  * host: the origin selection, which is a serial ran1 walk over the molecules
//    (:3735-3779); O(nmol) integer work, no pair touched.
//  * GPU: absolute atom positions, then the O(N_origin * N) inter-molecular pair loop
//    as a tiled all-pairs kernel: one origin atom per lane, partner atoms staged
//    through LDS 256 at a time (one float4 per atom), per-workgroup uint32 histograms
//    in LDS, merged into the global counts with atomics.  Intra-molecular pairs run
//    in a small kernel of their own.  A last kernel turns counts into the normalised
//    float32 wav of the reference.
//  * counts are exact integers; the reference's float32 accumulation of equal
//    power-of-two increments equals increment * min(count, 2^24), which the
//    normalisation kernel reproduces.
#include "common.h"

#define ATOGEN_BLOCK 512       // origin atoms (lanes) per workgroup of the pair kernel
#define ATOGEN_MIN_SLICE 256   // fewest partner atoms per work item

// absolute positions x2 = xa + xadd (one fp32 add, as :3794-3796), type in .w
__global__ void atogen_positions(const float *__restrict__ xa, const float *__restrict__ ya,
                                 const float *__restrict__ za, const float *__restrict__ xadd,
                                 const float *__restrict__ yadd, const float *__restrict__ zadd,
                                 const int *__restrict__ amol, const int *__restrict__ itype,
                                 float4 *__restrict__ P, int natom)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    int m = amol[a];
    P[a] = make_float4(xa[a] + xadd[m], ya[a] + yadd[m], za[a] + zadd[m], __int_as_float(itype[a]));
}

// intramolecular pairs (:3698-3724): one thread per molecule.  Rigid molecules put
// every copy of a bond in the same bin, so the counts go through a per-workgroup LDS
// histogram (or straight to global memory when it does not fit).
__global__ void atogen_intra(const float *__restrict__ xa, const float *__restrict__ ya,
                             const float *__restrict__ za, const int *__restrict__ naf,
                             const int *__restrict__ nal, const int *__restrict__ itype,
                             const int *__restrict__ npairref, int ntype, int npairs, int nmol, float sizefactor,
                             float stepcl, int npt1, int ldw, int use_lds, uint32_t *__restrict__ counts)
{
    extern __shared__ uint32_t hist[];
    const int ldh = npt1 + 1;
    if (use_lds) {
        for (int i = threadIdx.x; i < npairs * ldh; i += blockDim.x) hist[i] = 0u;
        __syncthreads();
    }
    int im = blockIdx.x * blockDim.x + threadIdx.x;
    if (im < nmol) {
        int a0 = naf[im], a1 = nal[im];
        for (int ia = a0; ia < a1; ++ia) {
            float x1 = xa[ia], y1 = ya[ia], z1 = za[ia];
            int it = itype[ia];
            for (int ja = ia + 1; ja <= a1; ++ja) {
                int icref = npairref[it * ntype + itype[ja]];
                float xx = dev_mic((x1 - xa[ja]) * sizefactor);
                float yy = dev_mic((y1 - ya[ja]) * sizefactor);
                float zz = dev_mic((z1 - za[ja]) * sizefactor);
                float dtest = sqrtf(xx * xx + yy * yy + zz * zz);
                int iz1 = dev_nint(dtest / stepcl) + 2;
                if (iz1 > 1 && iz1 <= npt1) {
                    if (use_lds) atomicAdd(&hist[icref * ldh + iz1], 1u);
                    else atomicAdd(&counts[(size_t)icref * ldw + iz1], 1u);
                }
            }
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < npairs * ldh; i += blockDim.x) {
            uint32_t v = hist[i];
            if (v) {
                int ic = i / ldh, iz = i - ic * ldh;
                atomicAdd(&counts[(size_t)ic * ldw + iz], v);
            }
        }
    }
}

// intermolecular pairs (:3780-3837).
// origin atoms: oatoms[0..norigin); for origin atom ia the partners are the atoms
// ja >= lo (lo = nal(mol(ia))+1 when every molecule is an origin, else 0) whose
// molecule differs from mol(ia).
//  * the minimum-image components enter only the squared distance (mic_sq).
//  * potex (read on the device): close-contact counting only when some rminex is non-zero.
//  * oatoms null: the origins are atoms 0..norigin-1 (every molecule but the last).
// |rminex(irefpcof(ic))| per pair, and *flag = 1 when one is non-zero
__global__ void atogen_rminex(int npairs, const int *__restrict__ irefpcof, const float *__restrict__ rminex,
                              float *__restrict__ rabs, unsigned *__restrict__ flag)
{
    const int ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= npairs) return;
    const float v = 1.0f * fabsf(rminex[irefpcof[ic]]);
    rabs[ic] = v;
    if (v > 0.0f) atomicOr(flag, 1u);
}

// The reference's bin of a squared distance: NINT(fl(sqrtf(dsq)) / stepcl) (:3800-3806),
// correctly rounded square root and division.
__device__ __forceinline__ int atogen_exact_bin(float dsq, float stepcl) { return dev_nint(sqrtf(dsq) / stepcl); }

// Both atogen_exact_bin(d) and the close-contact test fl(sqrtf(d) * dmax) < r are monotone
// in d (correctly rounded operations are), so each is a threshold on the squared distance.
//   thr[b], 1 <= b < npt1: the smallest float d whose bin is at least b (thr[0] = 0);
//   pth[ic]: the smallest float d with fl(sqrtf(d) * dmax) >= rabs[ic] -- a pair of column
//            ic is a close contact exactly when dsq < pth[ic] (0 when rabs is 0).
// Bisection over the bit patterns of non-negative floats, which are ordered like their values.
__global__ void atogen_thresholds(float stepcl, int npt1, float *__restrict__ thr, int npairs,
                                  const float *__restrict__ rabs, float dmax, float *__restrict__ pth)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npt1) {
        if (i == 0) {
            thr[0] = 0.0f;
            return;
        }
        // the bin of 0 is 0 < i; that of (2 (i + 2) stepcl)^2 is about 2 i + 4 >= i
        float hv = 2.0f * (float)(i + 2) * stepcl;
        hv = hv * hv;
        uint32_t lo = 0u, hi = __float_as_uint(hv);
        while (hi - lo > 1u) {
            const uint32_t mid = lo + (hi - lo) / 2u;
            if (atogen_exact_bin(__uint_as_float(mid), stepcl) >= i) hi = mid;
            else lo = mid;
        }
        thr[i] = __uint_as_float(hi);
        return;
    }
    const int ic = i - npt1;
    if (ic >= npairs) return;
    const float r = rabs[ic];
    if (!(r > 0.0f)) {
        pth[ic] = 0.0f;
        return;
    }
    uint32_t lo = 0u, hi = 0x7f800000u;   // f(0) = 0 < r <= f(+inf) = +inf
    while (hi - lo > 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (sqrtf(__uint_as_float(mid)) * dmax >= r) hi = mid;
        else lo = mid;
    }
    pth[ic] = __uint_as_float(hi);
}

// The bin atogen_exact_bin(dsq, stepcl) from the hardware square root.  qa = v_sqrt(dsq) x
// fl(1/stepcl) lies within 8 units of roundoff (relative; square root within 2 ulp, two
// roundings) of the correctly rounded quotient q, itself two roundings off the exact one.
// Away from half-integers -- |qa - rint(qa)| < bnd = 0.5 - 1.2e-6 (npt1 + 2), a margin of
// more than 20 units of roundoff of q for every q up to npt1 + 2 -- NINT(q) = rint(qa).  Near
// the half-integer b - 1/2 the bin is b - 1 or b, and it is b exactly when dsq >= thr[b].
// A bin past lastb = npt1 - 1 is never counted (the counted bins end at npt1 - 2), so there
// b needs no threshold; for q past npt1 + 2 both paths give bins past lastb as well.
__device__ __forceinline__ int atogen_bin(float dsq, float inv_stepcl, float bnd, int lastb, const float *thr)
{
    const float qa = __builtin_amdgcn_sqrtf(dsq) * inv_stepcl;
    const float rq = rintf(qa);
    const float dq = qa - rq;
    int n = (int)rq;
    if (!(fabsf(dq) < bnd)) {
        const int b = n + (dq > 0.0f ? 1 : 0);   // the nearest half-integer is b - 1/2, b >= 1
        n = b <= lastb ? b - (dsq < thr[b] ? 1 : 0) : b;
    }
    return n;
}

__device__ __forceinline__ float atogen_bnd(int npt1) { return 0.5f - 1.2e-6f * (float)(npt1 + 2); }

// Every float squared distance lo + i (i < count, as bit patterns) through atogen_bin and
// the close-contact threshold against the correctly rounded computation.  out[0]: the
// disagreements (a counted bin that differs, counted on one side only, or a close-contact
// decision that differs); out[1]: the values checked; out[2]: those that took the
// threshold branch of atogen_bin (near a bin edge).
__global__ void atogen_bin_sweep(float stepcl, float inv_stepcl, int npt1, const float *__restrict__ thr,
                                 float dmax, float rabs, const float *__restrict__ pth, uint32_t lo, uint32_t count,
                                 unsigned long long *__restrict__ out)
{
    const float bnd = atogen_bnd(npt1);
    const float p = pth[0];
    unsigned long long nbad = 0, nseen = 0, nedge = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
        const float dsq = __uint_as_float(lo + (uint32_t)i);
        const int n = atogen_bin(dsq, inv_stepcl, bnd, npt1 - 1, thr);
        const int e = atogen_exact_bin(dsq, stepcl);
        const bool in = n <= npt1 - 2, ine = e <= npt1 - 2;
        if (in != ine || (in && n != e)) ++nbad;
        if ((dsq < p) != (sqrtf(dsq) * dmax < rabs)) ++nbad;
        const float qa = __builtin_amdgcn_sqrtf(dsq) * inv_stepcl;
        if (!(fabsf(qa - rintf(qa)) < bnd)) ++nedge;
        ++nseen;
    }
    if (nbad) atomicAdd(&out[0], nbad);
    atomicAdd(&out[1], nseen);
    if (nedge) atomicAdd(&out[2], nedge);
}

// partner atoms p0..p1e-1 against this lane's origin atom (p1), into the workgroup's
// histograms.  Every lane of a wave takes the same partner at the same step, so the
// partners' positions are wave-uniform (scalar) loads, four at a time.  A lane counts the
// pair when the partner is outside its own molecule [mlo, mhi] (and, with FULL -- every
// molecule an origin, partners after the origin's molecule -- at or past lo = mhi + 1);
// with FULL, from the wave's largest lo (kfree) on every lane counts every partner.
//  * bin: atogen_bin, exact; close contacts: dsq < pth[g], exact (atogen_thresholds).
//  * PACK (one group of columns, at most 4 atom types, no close-contact counting): the
//    byte offset in LDS of bin 2 of each partner type's column for this lane's origin
//    type sits in two registers, 16 bits each (pk0: types 0, 1; pk1: types 2, 3); otherwise the pair
//    column comes from the LDS table snp and the column test against [ic0, ic0 + ngrp).
template <bool POTEX, bool FULL, bool PACK>
__device__ __forceinline__ void atogen_scan(const float4 *__restrict__ P, int p0, int p1e, int kfree, int lo,
                                            int mlo, int mhi, float4 p1, int rowoff, uint32_t pk0, uint32_t pk1,
                                            int ldh, int npt1, float inv_stepcl, const float *thr,
                                            const int *snp, const float *pth, uint32_t *hist, uint32_t *spot, int ic0,
                                            int ngrp)
{
    const float bnd = atogen_bnd(npt1);
    auto pair = [&](const float4 &p2) {
        const float xx = mic_sq(p1.x - p2.x), yy = mic_sq(p1.y - p2.y), zz = mic_sq(p1.z - p2.z);
        const float dsq = xx * xx + yy * yy + zz * zz;
        const int n = atogen_bin(dsq, inv_stepcl, bnd, npt1 - 1, thr);
        const bool in = n <= npt1 - 2;   // iz1 = n + 2 in 2..npt1
        const int t2 = __float_as_int(p2.w);
        if (PACK) {
            const uint32_t w = t2 < 2 ? pk0 : pk1;
            const uint32_t base = (w >> ((t2 & 1) * 16)) & 0xffffu;   // byte offset of bin 2
            if (in) atomicAdd((uint32_t *)((char *)hist + base + 4 * n), 1u);
        } else {
            const int g = snp[rowoff + t2] - ic0;   // column within this launch's group
            if ((unsigned)g < (unsigned)ngrp) {
                if (in) atomicAdd(&hist[g * ldh + n + 2], 1u);
                if (POTEX && dsq < pth[g]) atomicAdd(&spot[g], 1u);
            }
        }
    };
    auto take = [&](int k, const float4 &p2) {
        if (FULL ? k >= lo : (k < mlo || k > mhi)) pair(p2);
    };
    const int k1 = FULL ? min(p1e, max(p0, kfree)) : p1e;
    int k = p0;
    for (; k + 4 <= k1; k += 4) {
        const float4 a0 = P[k], a1 = P[k + 1], a2 = P[k + 2], a3 = P[k + 3];
        take(k, a0);
        take(k + 1, a1);
        take(k + 2, a2);
        take(k + 3, a3);
    }
    for (; k < k1; ++k) take(k, P[k]);
    if (FULL) {
        for (; k + 4 <= p1e; k += 4) {
            const float4 a0 = P[k], a1 = P[k + 1], a2 = P[k + 2], a3 = P[k + 3];
            pair(a0);
            pair(a1);
            pair(a2);
            pair(a3);
        }
        for (; k < p1e; ++k) pair(P[k]);
    }
}

// A persistent grid: gridDim.x workgroups take the work items (origin block bx of
// ATOGEN_BLOCK origin atoms x partner slice by, item = bx * gy + by) from a queue (one
// counter per column group, zero at launch) until it runs dry, accumulating into one LDS
// histogram that is flushed to the global counts once, at the end -- the flush costs
// gridDim.x histograms, not one per item.
// blockIdx.y selects the group of pair columns [ic0, ic0 + ngrp) this workgroup counts,
// so histograms larger than LDS split into groups (each group repeats the distances).
__global__ void __launch_bounds__(ATOGEN_BLOCK)
atogen_inter_fast(const float4 *__restrict__ P, const int *__restrict__ amol, const int *__restrict__ naf,
                  const int *__restrict__ nal,
                  const int *__restrict__ oatoms, int norigin, int full, int natom,
                  const int *__restrict__ npairref_g, int ntype, int npairs, int npt1, int ldw, float inv_stepcl,
                  const float *__restrict__ thr_g, const float *__restrict__ pth_g,
                  const unsigned *__restrict__ potex_flag, uint32_t *__restrict__ counts,
                  uint32_t *__restrict__ potex_counts, unsigned *__restrict__ queue, int gx, int gy, int ngrp)
{
    const bool potex = *potex_flag != 0u;   // some |rminex| is non-zero
    extern __shared__ uint32_t smem[];
    const int ldh = npt1 + 1;
    const int ic0 = blockIdx.y * ngrp;
    const int ng = min(ngrp, npairs - ic0);
    uint32_t *hist = smem;
    uint32_t *spot = hist + (size_t)ngrp * ldh;
    int *snp = (int *)(spot + ngrp);
    float *pth = (float *)(snp + ntype * ntype);
    float *thr = pth + ngrp;

    const int tid = threadIdx.x;
    for (int i = tid; i < ng * ldh; i += ATOGEN_BLOCK) hist[i] = 0u;
    for (int i = tid; i < ng; i += ATOGEN_BLOCK) {
        spot[i] = 0u;
        pth[i] = pth_g[ic0 + i];
    }
    for (int i = tid; i < ntype * ntype; i += ATOGEN_BLOCK) snp[i] = npairref_g[i];
    for (int i = tid; i < npt1; i += ATOGEN_BLOCK) thr[i] = thr_g[i];
    __syncthreads();

    // packed offsets: one group of columns, at most 4 types, byte offsets below 2^16
    const bool pack = !potex && ng == npairs && ntype <= 4 && 4 * (npairs * ldh + 2) < 65536;
    const int nitems = gx * gy;
    __shared__ int s_item;
    for (;;) {
        // the next work item of this group, from its queue: in full mode the items come
        // heaviest first (origin block 0 has the most partners)
        if (tid == 0) s_item = (int)atomicAdd(&queue[blockIdx.y], 1u);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= nitems) break;
        const int bx = item / gy, by = item - bx * gy;
        const int oi = bx * ATOGEN_BLOCK + tid;
        const bool active = oi < norigin;
        float4 p1 = make_float4(0.f, 0.f, 0.f, 0.f);
        int lo = natom, mlo = 0, mhi = -1, rowoff = 0;
        uint32_t pk0 = 0u, pk1 = 0u;
        if (active) {
            int ia = oatoms ? oatoms[oi] : oi;
            p1 = P[ia];
            const int mymol = amol[ia];
            mlo = naf[mymol];
            mhi = nal[mymol];
            lo = full ? mhi + 1 : 0;
            rowoff = __float_as_int(p1.w) * ntype;
            if (pack)
                for (int t = 0; t < ntype; ++t) {
                    const uint32_t o = (uint32_t)(4 * (snp[rowoff + t] * ldh + 2)) << ((t & 1) * 16);
                    if (t < 2) pk0 |= o;
                    else pk1 |= o;
                }
        }
        const int first = bx * ATOGEN_BLOCK;
        const int blo = full ? nal[amol[oatoms ? oatoms[first] : first]] + 1 : 0;
        const int span = natom - blo;
        const int per = (span + gy - 1) / gy;
        const int p0 = blo + by * per;
        const int p1e = min(natom, p0 + per);
        // the largest lo of this lane's wave (full mode: the origins are atoms 0..norigin-1,
        // so it is the lo of the wave's last active origin)
        int kfree = natom;
        if (full) {
            const int wlast = min(norigin - 1, (oi | 63));
            kfree = wlast >= 0 ? nal[amol[wlast]] + 1 : natom;
            kfree = __builtin_amdgcn_readfirstlane(kfree);   // the same in every lane of the wave
        }
        if (active) {
            // the pair loop, specialised on the launch-wide flags
            if (potex) {
                if (full) atogen_scan<true, true, false>(P, p0, p1e, kfree, lo, mlo, mhi, p1, rowoff, 0u, 0u, ldh, npt1, inv_stepcl, thr, snp, pth, hist, spot, ic0, ng);
                else atogen_scan<true, false, false>(P, p0, p1e, kfree, lo, mlo, mhi, p1, rowoff, 0u, 0u, ldh, npt1, inv_stepcl, thr, snp, pth, hist, spot, ic0, ng);
            } else if (pack) {
                if (full) atogen_scan<false, true, true>(P, p0, p1e, kfree, lo, mlo, mhi, p1, rowoff, pk0, pk1, ldh, npt1, inv_stepcl, thr, snp, pth, hist, spot, ic0, ng);
                else atogen_scan<false, false, true>(P, p0, p1e, kfree, lo, mlo, mhi, p1, rowoff, pk0, pk1, ldh, npt1, inv_stepcl, thr, snp, pth, hist, spot, ic0, ng);
            } else {
                if (full) atogen_scan<false, true, false>(P, p0, p1e, kfree, lo, mlo, mhi, p1, rowoff, 0u, 0u, ldh, npt1, inv_stepcl, thr, snp, pth, hist, spot, ic0, ng);
                else atogen_scan<false, false, false>(P, p0, p1e, kfree, lo, mlo, mhi, p1, rowoff, 0u, 0u, ldh, npt1, inv_stepcl, thr, snp, pth, hist, spot, ic0, ng);
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < ng * ldh; i += ATOGEN_BLOCK) {
        uint32_t v = hist[i];
        if (v) {
            int g = i / ldh, iz = i - g * ldh;
            atomicAdd(&counts[((size_t)npairs + ic0 + g) * ldw + iz], v);
        }
    }
    if (potex)
        for (int i = tid; i < ng; i += ATOGEN_BLOCK)
            if (spot[i]) atomicAdd(&potex_counts[ic0 + i], spot[i]);
}

// normalisation (:3859-3897).  norm[ic] = cprod, norm[npairs+ic] = cprod2 (host
// computed in float32 exactly as the reference); potex follows cprod2.
__global__ void atogen_normalise(const uint32_t *__restrict__ counts, const uint32_t *__restrict__ pcounts,
                                 const float *__restrict__ norm, const float *__restrict__ cfrac_jt,
                                 const float *__restrict__ ycoh, int npairs, int npt1, int ldw,
                                 float addwav, float *__restrict__ wav, float *__restrict__ potex,
                                 float *__restrict__ potexmax)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 2 * npairs * npt1;
    if (i < npairs) {
        uint32_t pc = pcounts[i];
        float v = (float)(pc > 16777216u ? 16777216u : pc);
        float c2 = norm[npairs + i];
        if (c2 > 0.0f) v = v * cfrac_jt[i] / c2;
        potex[i] = v;
        potexmax[i] = 0.0f;
    }
    if (i >= total) return;
    int k = i / (npairs * npt1);
    int r = i - k * npairs * npt1;
    int ic = r / npt1;
    int iz = r - ic * npt1 + 1;
    size_t o = ((size_t)k * npairs + ic) * ldw + iz;
    uint32_t c = counts[o];
    float inc = k == 0 ? 2.0f : addwav;
    float v = inc * (float)(c > 16777216u ? 16777216u : c);
    float d = norm[k * npairs + ic];
    if (d > 0.0f) v = v / d;
    else v = (k == 0) ? 0.0f : ycoh[iz];
    wav[o] = v;
}

extern "C" int epsr_atogen(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, int npt1, int ldw,
                           const float *ycoh, const int *irefpcof, const float *rminex, int *ngrsamples,
                           int32_t *ran1_state, float *wav, float *potex, float *potexmax, uint32_t *counts_out)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && ycoh && irefpcof && rminex && ngrsamples && ran1_state && wav && potex && potexmax,
              "epsr_atogen: null argument");
    ARG_CHECK(ctx, npt1 >= 2 && ldw >= npt1 + 1, "epsr_atogen: need npt1 >= 2 and ldw >= npt1+1 (npt1=%d ldw=%d)", npt1, ldw);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int nmol = box->nmol, natom = box->natom, npairs = box->npairs, ntype = box->ntype;
    hipStream_t st = ctx->stream;

    // ---- origin selection on the host (:3735-3779), ran1 drawn for every candidate
    int ngr = *ngrsamples;
    if (ngr <= 0) ngr = 1000;
    if (ngr > nmol) ngr = nmol;
    float probaccept = (float)ngr / (float)nmol;
    std::vector<int> molsampled(box->nmoltype, 0);
    std::vector<float> molfraction(box->nmoltype, 0.0f);
    std::vector<float> csampled(ntype, 0.0f);
    for (int im = 0; im < nmol; ++im) molfraction[box->h_mtype[im]] = molfraction[box->h_mtype[im]] + 1.0f;
    for (int t = 0; t < box->nmoltype; ++t) molfraction[t] = molfraction[t] / (float)nmol;
    int im2 = nmol - 1;
    bool full = false;
    float addwav = 1.0f;
    if (probaccept >= 1.0f) {
        im2 = nmol - 2;
        full = true;
        addwav = 2.0f;
        for (int t = 0; t < ntype; ++t)
            csampled[t] = csampled[t] + (float)box->h_type_count[(size_t)(nmol - 1) * ntype + t];
    }
    Ran1 rng;
    ran1_load(rng, ran1_state);
    std::vector<int> oatoms;
    int norigin = 0;
    // one ran1 value per candidate molecule, drawn as a block
    std::vector<float> rv((size_t)im2 + 1);
    ran1_fill(rng, rv.data(), (long long)im2 + 1);
    if (full) {
        // every candidate is taken (r < 1 <= probaccept): the origins are the atoms of
        // molecules 0..im2, a contiguous range the kernel enumerates itself
        for (int im = 0; im <= im2; ++im) {
            molsampled[box->h_mtype[im]] += 1;
            for (int t = 0; t < ntype; ++t)
                csampled[t] = csampled[t] + (float)box->h_type_count[(size_t)im * ntype + t];
        }
        norigin = im2 >= 0 ? box->h_nal[im2] + 1 : 0;
    } else {
        oatoms.reserve((size_t)ngr * box->max_atoms_per_mol * 2);
        for (int im = 0; im <= im2; ++im) {
            int mt = box->h_mtype[im];
            float r = rv[im];
            if (r < probaccept || molsampled[mt] < 1 || molfraction[mt] < probaccept) {
                molsampled[mt] += 1;
                for (int t = 0; t < ntype; ++t)
                    csampled[t] = csampled[t] + (float)box->h_type_count[(size_t)im * ntype + t];
                for (int a = box->h_naf[im]; a <= box->h_nal[im]; ++a) oatoms.push_back(a);
            }
        }
        norigin = (int)oatoms.size();
    }
    ran1_store(rng, ran1_state);
    *ngrsamples = ngr;

    // ---- normalisation constants (:3862-3869), float32 as the reference
    std::vector<float> hnorm(2 * npairs), hcfj(npairs);
    for (int ic = 0; ic < npairs; ++ic) {
        int it = box->h_n1extr[ic], jt = box->h_n2extr[ic];
        float cprod = box->h_cfrac[it] * box->h_cfrac[jt] * box->natomr;
        float cprod2 = csampled[it] * box->h_cfrac[jt];
        if (it != jt) {
            cprod = cprod * 2.0f;
            cprod2 = cprod2 + csampled[jt] * box->h_cfrac[it];
        }
        hnorm[ic] = cprod;
        hnorm[npairs + ic] = cprod2;
        hcfj[ic] = box->h_cfrac[jt];
    }

    // ---- device scratch layout
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    size_t oP = take(sizeof(float4) * natom);
    size_t oO = take(sizeof(int) * (oatoms.size() > 0 ? oatoms.size() : 1));
    size_t oC = counts_out ? 0 : take(sizeof(uint32_t) * 2 * (size_t)npairs * ldw);
    size_t oPC = take(sizeof(uint32_t) * npairs);
    size_t oN = take(sizeof(float) * 2 * npairs);
    size_t oF = take(sizeof(float) * npairs);
    size_t oR = take(sizeof(float) * npairs);
    size_t oPT = take(sizeof(float) * npairs);
    size_t oT = take(sizeof(float) * npt1);
    size_t oMX = take(sizeof(unsigned));
    size_t oQ = take(sizeof(unsigned) * npairs);
    char *scr = (char *)ctx_scratch(ctx, off);
    if (!scr) return EPSR_ERR_HIP;
    float4 *dP = (float4 *)(scr + oP);
    int *dO = (int *)(scr + oO);
    uint32_t *dC = counts_out ? counts_out : (uint32_t *)(scr + oC);
    uint32_t *dPC = (uint32_t *)(scr + oPC);
    float *dN = (float *)(scr + oN), *dF = (float *)(scr + oF), *dR = (float *)(scr + oR);
    float *dPT = (float *)(scr + oPT), *dT = (float *)(scr + oT);
    unsigned *dMX = (unsigned *)(scr + oMX);
    unsigned *dQ = (unsigned *)(scr + oQ);

    // staging of the small host arrays through pinned memory
    const int nlist = (int)oatoms.size();
    size_t hb = sizeof(int) * (size_t)nlist + sizeof(float) * 4 * npairs;
    char *pin = (char *)ctx_pinned(ctx, hb);
    if (!pin) return EPSR_ERR_HIP;
    HIP_CHECK(ctx, hipStreamSynchronize(st));   // the pinned buffer may still feed a previous copy
    memcpy(pin, oatoms.data(), sizeof(int) * (size_t)nlist);
    float *pf = (float *)(pin + sizeof(int) * (size_t)nlist);
    memcpy(pf, hnorm.data(), sizeof(float) * 2 * npairs);
    memcpy(pf + 2 * npairs, hcfj.data(), sizeof(float) * npairs);
    if (nlist) HIP_CHECK(ctx, hipMemcpyAsync(dO, pin, sizeof(int) * (size_t)nlist, hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dN, pf, sizeof(float) * 2 * npairs, hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dF, pf + 2 * npairs, sizeof(float) * npairs, hipMemcpyHostToDevice, st));

    // |rminex(irefpcof(ic))| per pair, gathered on the device; dMX says whether any is
    // non-zero (the pair kernel then counts close contacts)
    HIP_CHECK(ctx, hipMemsetAsync(dMX, 0, sizeof(unsigned), st));
    atogen_rminex<<<ceil_div(npairs, 64), 64, 0, st>>>(npairs, irefpcof, rminex, dR, dMX);

    HIP_CHECK(ctx, hipMemsetAsync(dC, 0, sizeof(uint32_t) * 2 * (size_t)npairs * ldw, st));
    HIP_CHECK(ctx, hipMemsetAsync(dPC, 0, sizeof(uint32_t) * npairs, st));
    HIP_CHECK(ctx, hipMemsetAsync(dQ, 0, sizeof(unsigned) * npairs, st));
    const float stepcl = box->cellst / box->dmax;

    atogen_positions<<<ceil_div(natom, 256), 256, 0, st>>>(c->xa, c->ya, c->za, c->xadd, c->yadd, c->zadd,
                                                            box->d_amol, box->d_itype, dP, natom);
    {
        size_t hbytes = sizeof(uint32_t) * (size_t)npairs * (npt1 + 1);
        int use_lds = hbytes <= 96 * 1024;
        const int blk = 256;
        atogen_intra<<<ceil_div(nmol, blk), blk, use_lds ? hbytes : 0, st>>>(
            c->xa, c->ya, c->za, box->d_naf, box->d_nal, box->d_itype, box->d_npairref, ntype, npairs, nmol,
            box->sizefactor, stepcl, npt1, ldw, use_lds, dC);
    }
    if (norigin > 0) {
        // pair columns per launch group: the histogram of a group and the tiles within a
        // workgroup's share of LDS (at least two resident workgroups per CU)
        const size_t fixed = sizeof(int) * ntype * ntype + sizeof(float) * npt1 + 16;
        const size_t percol = sizeof(uint32_t) * ((size_t)npt1 + 1) + sizeof(uint32_t) + sizeof(float);
        const size_t budget = 64 * 1024;
        ARG_CHECK(ctx, fixed + percol <= budget, "epsr_atogen: %d bins do not fit in LDS", npt1);
        const int ngrp = (int)std::min<size_t>((size_t)npairs, (budget - fixed) / percol);
        const int ngroups = ceil_div(npairs, ngrp);
        const size_t lds = fixed + percol * ngrp + 16;
        const int gx = ceil_div(norigin, ATOGEN_BLOCK);
        // resident workgroups per group, and partner slices so that each walks several items
        const int wres = std::max(1, std::min(box->atogen_workgroups / ngroups, 1 << 16));
        int gy = std::max(1, std::min(256, ceil_div(8L * wres, gx)));
        gy = std::min(gy, std::max(1, natom / ATOGEN_MIN_SLICE));
        const int nwg = std::min(wres, gx * gy);
        const float inv_stepcl = 1.0f / stepcl;
        atogen_thresholds<<<ceil_div(npt1 + npairs, 64), 64, 0, st>>>(stepcl, npt1, dT, npairs, dR, box->dmax, dPT);
        timer_begin(ctx, "atogen_inter");
        hipLaunchKernelGGL(atogen_inter_fast, dim3(nwg, ngroups), dim3(ATOGEN_BLOCK), lds, st, dP, box->d_amol,
                           box->d_naf, box->d_nal, nlist ? dO : nullptr, norigin, full ? 1 : 0, natom, box->d_npairref, ntype,
                           npairs, npt1, ldw, inv_stepcl, dT, dPT, dMX, dC, dPC, dQ, gx, gy, ngrp);
        timer_end(ctx, "atogen_inter");
    }
    int total = 2 * npairs * npt1;
    atogen_normalise<<<ceil_div(std::max(total, npairs), 256), 256, 0, st>>>(dC, dPC, dN, dF, ycoh, npairs, npt1, ldw,
                                                                              addwav, wav, potex, potexmax);
    // no synchronisation: the outputs are device arrays, ordered on the stream
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}

extern "C" int epsr_atogen_bin_check(epsr_ctx *ctx, float stepcl, int npt1, float dmax, float rabs, uint32_t lo_bits,
                                     uint32_t hi_bits, unsigned long long *counts)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, counts, "epsr_atogen_bin_check: null argument");
    ARG_CHECK(ctx, npt1 >= 2 && stepcl > 0.0f && dmax > 0.0f && rabs >= 0.0f,
              "epsr_atogen_bin_check: need npt1 >= 2, stepcl > 0, dmax > 0, rabs >= 0");
    ARG_CHECK(ctx, lo_bits <= hi_bits && hi_bits <= 0x7f800000u,
              "epsr_atogen_bin_check: need lo_bits <= hi_bits <= the bits of +inf");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    const size_t oT = take(sizeof(float) * npt1), oR = take(sizeof(float)), oP = take(sizeof(float)),
                 oB = take(3 * sizeof(unsigned long long));
    char *scr = (char *)ctx_scratch(ctx, off);
    if (!scr) return EPSR_ERR_HIP;
    float *dT = (float *)(scr + oT), *dR = (float *)(scr + oR), *dP = (float *)(scr + oP);
    unsigned long long *dB = (unsigned long long *)(scr + oB);
    HIP_CHECK(ctx, hipMemcpyAsync(dR, &rabs, sizeof(float), hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemsetAsync(dB, 0, 3 * sizeof(unsigned long long), st));
    atogen_thresholds<<<ceil_div(npt1 + 1, 64), 64, 0, st>>>(stepcl, npt1, dT, 1, dR, dmax, dP);
    const uint32_t count = hi_bits - lo_bits;
    if (count > 0)
        atogen_bin_sweep<<<4096, 256, 0, st>>>(stepcl, 1.0f / stepcl, npt1, dT, dmax, rabs, dP, lo_bits, count, dB);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipMemcpyAsync(counts, dB, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    return EPSR_OK;
}
