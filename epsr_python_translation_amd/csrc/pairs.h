// Pair energies by table lookup, shared by the MC kernels (mc.hip) and the
// intramolecular cycle (intra.hip): packed tables, the correctly rounded square root,
// en_atom_inter_scale and en_mole_pair.
//
// Reference: src/epsr26/library/molfit.py:1250-1279 (en_atom_inter_scale),
// :1527-1578 (en_mole_pair).
#pragma once
#include "common.h"

// Packed tables: bins [0, nnear) as T[ic*nnear + iz] = (xavs, xavs1, ecoul, ecoul1)(iz, ic);
// bins [nnear, nte), where xavs and xavs1 are +0 in every column, as
// F[ic*(nte - nnear) + iz - nnear] = (ecoul, ecoul1)(iz, ic); beyond nte every table entry
// is +0, where the reference's lookup also gives +0.  The split keeps the tables of a
// long Coulomb tail small enough for two workgroups' copies per CU in LDS.
struct PairTab {
    const float4 *T;
    const float2 *F;
    const float *xs;
    const int *npr;
    int nte, nnear, npt2, ntype;
    float stepcl, inv_stepcl;   // inv_stepcl = 1/stepcl rounded to float
};

// IEEE float square root.  For 2^-96 <= x < inf: the hardware estimate s (within 1 ulp)
// or a neighbour, chosen by the signs of the exact residuals x - s_dn*s and x - s_up*s
// (the same selection the compiler's correctly rounded sqrtf makes, without its scaling
// of tiny arguments and its special-value fix-up, which that range does not need);
// other arguments take sqrtf itself.
__device__ __forceinline__ float sqrt_rn(float x)
{
    if (!(x >= 0x1p-96f && x < INFINITY)) return sqrtf(x);
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sdn = __int_as_float(__float_as_int(s) - 1), sup = __int_as_float(__float_as_int(s) + 1);
    const float rdn = fmaf(-sdn, s, x), rup = fmaf(-sup, s, x);
    float r = rdn <= 0.0f ? sdn : s;
    return rup > 0.0f ? sup : r;
}

// en_atom_inter_scale (molfit.py:1250-1279) with scalefac = 1; the minimum-image
// components enter only the squared distance (mic_sq).  The bin is
// trunc(dtest / stepcl) + 2 with the correctly rounded quotient Q: q = dtest * inv_stepcl
// lies within 1.2e-7 q of dtest / stepcl (two roundings of 2^-24) and Q within 6e-8 Q of
// it, so when q's fraction keeps clear of 0 and 1 by 1.2e-6 q, trunc(q) = trunc(Q);
// otherwise (about 2.4e-6 q of the pairs) the division is made.
__device__ __forceinline__ float atom_pair(float x1, float y1, float z1, float x2, float y2, float z2, int ic,
                                           const PairTab &t)
{
    const float xd = mic_sq(x2 - x1), yd = mic_sq(y2 - y1), zd = mic_sq(z2 - z1);
    float dsq = xd * xd + yd * yd + zd * zd;
    float dtest = sqrt_rn(dsq);
    const float q = dtest * t.inv_stepcl;
    const float tq = truncf(q), fq = q - tq, mq = 1.2e-6f * q;
    int iz;
    if (fq > mq && fq < 1.0f - mq && q < 4194304.0f) {
        iz = (int)tq + 2;
    } else {
        const float rat = dtest / t.stepcl;
        if (!(rat < 2.0e9f)) return 0.0f;
        iz = (int)truncf(rat) + 2;
    }
    float sume = 0.0f;
    if (iz <= t.npt2 && iz < t.nte) {
        float4 v;
        if (iz < t.nnear) {
            v = t.T[ic * t.nnear + iz];
        } else {
            const float2 f = t.F[ic * (t.nte - t.nnear) + (iz - t.nnear)];
            v = make_float4(0.0f, 0.0f, f.x, f.y);   // the stored +0 of xavs, xavs1
        }
        float dtdif = dtest - t.xs[iz];
        sume = v.x + v.y * dtdif;
        sume = sume + v.z + v.w * dtdif;
    }
    return sume;
}

// The same arithmetic as atom_pair, split in three steps so that a group of pairs has
// all its table reads in flight at once: the bins (no branch but a rare fix-up), the
// table reads, then the energies.
// pair_bin_fast: dtest (sqrt_rn's correction, branch-free) and the bin from the rounded
// reciprocal; *ok is false when the reciprocal cannot settle the bin or dsq lies outside
// sqrt_rn's fast range -- those pairs take pair_bin_exact.
__device__ __forceinline__ int pair_bin_fast(float dsq, const PairTab &t, float &dtest, bool &ok)
{
    const float s = __builtin_amdgcn_sqrtf(dsq);
    const float sdn = __int_as_float(__float_as_int(s) - 1), sup = __int_as_float(__float_as_int(s) + 1);
    const float rdn = fmaf(-sdn, s, dsq), rup = fmaf(-sup, s, dsq);
    const float r = rdn <= 0.0f ? sdn : s;
    dtest = rup > 0.0f ? sup : r;
    const float q = dtest * t.inv_stepcl;
    const float tq = truncf(q), fq = q - tq, mq = 1.2e-6f * q;
    ok = dsq >= 0x1p-96f && dsq < INFINITY && fq > mq && fq < 1.0f - mq && q < 4194304.0f;
    return (int)tq + 2;
}

// atom_pair's division path: dtest and the bin, INT_MAX (no table entry) where atom_pair
// gives 0 outright
struct PairBin {
    float dtest;
    int iz;
};
__device__ __forceinline__ PairBin pair_bin_exact(float dsq, const PairTab &t)
{
    const float d = sqrt_rn(dsq);
    const float rat = d / t.stepcl;
    return {d, rat < 2.0e9f ? (int)truncf(rat) + 2 : 0x7fffffff};
}

// G atom pairs, member g pairing p[g / NQ] with q[g % NQ] (valid: bit g), added to sume
// in the order of g
template <int G, int NQ>
__device__ __forceinline__ float pair_block(float sume, const float4 *p, const float4 *q, unsigned valid,
                                            const PairTab &t)
{
    float dsq[G], dt[G];
    int iz[G], ic[G];
    bool fix = false;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        dsq[g] = 0.0f;
        dt[g] = 0.0f;
        iz[g] = 0x7fffffff;
        ic[g] = 0;
        if ((valid >> g) & 1u) {
            const float4 &pa = p[g / NQ], &qa = q[g % NQ];
            ic[g] = t.npr[__float_as_int(pa.w) * t.ntype + __float_as_int(qa.w)];
            const float xd = mic_sq(qa.x - pa.x), yd = mic_sq(qa.y - pa.y), zd = mic_sq(qa.z - pa.z);
            dsq[g] = xd * xd + yd * yd + zd * zd;
            bool ok;
            iz[g] = pair_bin_fast(dsq[g], t, dt[g], ok);
            if (!ok) {
                iz[g] = -1;   // marks the pair for the exact bin below
                fix = true;
            }
        }
    }
    if (fix) {
#pragma unroll
        for (int g = 0; g < G; ++g)
            if (iz[g] < 0) {
                const PairBin b = pair_bin_exact(dsq[g], t);
                dt[g] = b.dtest;
                iz[g] = b.iz;
            }
    }
    float4 v[G];
    float xs[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const bool in = iz[g] <= t.npt2 && iz[g] < t.nte;
        const bool near = iz[g] < t.nnear;
        v[g] = t.T[in && near ? ic[g] * t.nnear + iz[g] : 0];
        xs[g] = t.xs[in ? iz[g] : 0];
        if (in && !near) {
            const float2 f = t.F[ic[g] * (t.nte - t.nnear) + (iz[g] - t.nnear)];
            v[g] = make_float4(0.0f, 0.0f, f.x, f.y);   // the stored +0 of xavs, xavs1
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if ((valid >> g) & 1u) {
            float e = 0.0f;
            if (iz[g] <= t.npt2 && iz[g] < t.nte) {
                const float dtdif = dt[g] - xs[g];
                e = v[g].x + v[g].y * dtdif;
                e = e + v[g].z + v[g].w * dtdif;
            }
            sume = sume + e;
        }
    }
    return sume;
}

// en_mole_pair (molfit.py:1527-1578): sequential sum, molecule 1 outer, molecule 2 inner;
// each atom of molecule 1 against molecule 2's atoms four at a time (pair_block)
__device__ __forceinline__ float mol_pair(const float4 *__restrict__ P1, int na1, const float4 *__restrict__ P2,
                                          int nb2, const PairTab &t)
{
    float sume = 0.0f;
    for (int ia = 0; ia < na1; ++ia) {
        const float4 p = P1[ia];
        for (int j0 = 0; j0 < nb2; j0 += 4) {
            float4 q[4];
            const int ng = min(4, nb2 - j0);
#pragma unroll
            for (int g = 0; g < 4; ++g) q[g] = g < ng ? P2[j0 + g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            sume = pair_block<4, 4>(sume, &p, q, (1u << ng) - 1u, t);
        }
    }
    return sume;
}

// mol_pair for molecules of at most three atoms: the six positions are loaded before any
// pair is evaluated, and each atom of molecule 1 meets the three of molecule 2 as a group
__device__ __forceinline__ float mol_pair3(const float4 *__restrict__ P1, int n1, const float4 *__restrict__ P2,
                                           int n2, const PairTab &t)
{
    float4 p[3], q[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        p[j] = j < n1 ? P1[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        q[j] = j < n2 ? P2[j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const unsigned rowm = (1u << n2) - 1u;
    float sume = 0.0f;
#pragma unroll
    for (int ia = 0; ia < 3; ++ia) {
        if (ia >= n1) break;
        sume = pair_block<3, 3>(sume, &p[ia], q, rowm, t);
    }
    return sume;
}

// mol_pair3 with molecule 2's three atoms given (loaded by the caller)
__device__ __forceinline__ float mol_pair3_q(const float4 *__restrict__ P1, int n1, const float4 (&q)[3], int n2,
                                             const PairTab &t)
{
    const unsigned rowm = (1u << n2) - 1u;
    float sume = 0.0f;
#pragma unroll
    for (int ia = 0; ia < 3; ++ia) {
        if (ia >= n1) break;
        const float4 p = P1[ia];   // (one atom at a time: fewer registers live)
        sume = pair_block<3, 3>(sume, &p, q, rowm, t);
    }
    return sume;
}

// the pair energy of two molecules, through mol_pair3 when neither has more than three atoms
__device__ __forceinline__ float mol_pair_any(const float4 *__restrict__ P1, int na1, const float4 *__restrict__ P2,
                                              int nb2, const PairTab &t)
{
    return (na1 <= 3 && nb2 <= 3) ? mol_pair3(P1, na1, P2, nb2, t) : mol_pair(P1, na1, P2, nb2, t);
}

// Tables prepared for a call: packed float4 rows up to the last non-zero bin.
struct PreparedTables {
    PairTab tab;
    size_t lds_bytes;
    bool in_lds;
};

// measure and pack the tables of tb into dT (dscratch_int: 2 ints of device scratch)
int mc_prepare_tables(epsr_ctx *ctx, const epsr_box *box, const epsr_tables_t *tb, float4 *dT, int *dscratch_int,
                      PreparedTables &pt);
