/*
 * ora_core.c -- oracle: arithmetic contract, ran1, minimum image, atogen,
 * gensq, stog.  TEST INFRASTRUCTURE ONLY (see epsr_oracle.h).
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ fmath */

/* src/epsr26/fortran/fmath.py:222-239 -- NINT: round half away from zero on the
 * double value; outside int64 -> 0, outside int32 wraps. */
int ora_nint(float xf)
{
    double x = (double)xf;
    if (!(x >= -9.2233720368547758e18 && x < 9.2233720368547758e18)) return 0;
    double t = trunc(x);
    double frac = x - t;
    int64_t ti = (int64_t)t;
    if (frac >= 0.5) ti += 1;
    else if (frac <= -0.5) ti -= 1;
    return (int)(int32_t)(uint32_t)(uint64_t)ti;
}

/* fmath.py:252-260 -- IFIX: truncation, out of range -> INT_MIN */
int ora_ifix(float xf)
{
    double x = (double)xf;
    if (!(x > -2147483649.0 && x < 2147483648.0)) return (int)0x80000000u;
    return (int)trunc(x);
}

/* fmath.py:186-200 -- libgcc __powisf2 */
float ora_powi(float x, int m)
{
    unsigned n = m < 0 ? (unsigned)(-(long)m) : (unsigned)m;
    float y = (n % 2u) ? x : 1.0f;
    while (1) {
        n >>= 1;
        if (n == 0) break;
        x = x * x;
        if (n % 2u) y = y * x;
    }
    return m < 0 ? 1.0f / y : y;
}

/* fmath.py:152-157 -- correctly rounded sqrt; negative -> NaN */
float ora_sqrtf(float x) { return x < 0.0f ? NAN : sqrtf(x); }

#ifdef ORACLE_CR_MATH
/* correctly rounded flavour: evaluate in double, round once to float */
float ora_expf(float x)  { return (float)exp((double)x); }
float ora_logf(float x)  { return (float)log((double)x); }
float ora_sinf(float x)  { return (float)sin((double)x); }
float ora_cosf(float x)  { return (float)cos((double)x); }
float ora_atanf(float x) { return (float)atan((double)x); }
float ora_sinhf(float x) { return (float)sinh((double)x); }
float ora_powf(float x, float y) { return (float)pow((double)x, (double)y); }
int ora_libm_is_cr(void) { return 1; }
#else
/* glibc flavour: the calls the reference makes through ctypes (fmath.py:46-149) */
float ora_expf(float x)  { return expf(x); }
float ora_logf(float x)  { return logf(x); }
float ora_sinf(float x)  { return sinf(x); }
float ora_cosf(float x)  { return cosf(x); }
float ora_atanf(float x) { return atanf(x); }
float ora_sinhf(float x) { return sinhf(x); }
float ora_powf(float x, float y) { return powf(x, y); }
int ora_libm_is_cr(void) { return 0; }
#endif

/* ------------------------------------------------------------------ ran1 */
/* src/epsr26/common/ran1com.py:12-20 constants; auxiliary_routines.py:324-385 */
#define R_IA 16807
#define R_IM 2147483647
#define R_IQ 127773
#define R_IR 2836
#define R_NTAB 32
#define R_NDIV 67108864

float ora_ran1(ora_ran1_t *s)
{
    const float AM = 4.656612873077393e-10f;
    const float RNMX = 0.9999998807907104f;
    int32_t idum = s->idum, iy = s->iy;
    if (idum <= 0 || iy == 0) {
        int64_t neg = -(int64_t)idum;
        int32_t w = (int32_t)(uint32_t)(uint64_t)neg;     /* wrap32(-idum) */
        idum = w > 1 ? w : 1;
        for (int j = R_NTAB + 8; j >= 1; --j) {
            int32_t k = idum / R_IQ;
            idum = R_IA * (idum - k * R_IQ) - R_IR * k;
            if (idum < 0) idum += R_IM;
            if (j <= R_NTAB) s->iv[j - 1] = idum;
        }
        iy = s->iv[0];
    }
    int32_t k = idum / R_IQ;
    idum = R_IA * (idum - k * R_IQ) - R_IR * k;
    if (idum < 0) idum += R_IM;
    int j = 1 + iy / R_NDIV;
    iy = s->iv[j - 1];
    s->iv[j - 1] = idum;
    float r = AM * (float)iy;
    if (r > RNMX) r = RNMX;
    s->idum = idum;
    s->iy = iy;
    s->mcall = (int32_t)((uint32_t)s->mcall + 1u);
    return r;
}

void ora_ran1_stream(ora_ran1_t *s, int n, float *out)
{
    for (int i = 0; i < n; ++i) out[i] = ora_ran1(s);
}

/* ------------------------------------------------------------------ minimage */
/* src/epsr26/library/minimagec.py:48-62 (cubic branch of minimagec, :22-25) */
void ora_minimage(float *x, float *y, float *z)
{
    *z = *z - (float)ora_nint(*z);
    *y = *y - (float)ora_nint(*y);
    *x = *x - (float)ora_nint(*x);
}

/* ------------------------------------------------------------------ atogen */
/* Origin selection of atogen, ato_manipulation_rev4.py:3735-3779.  Fills
 * origin[im] (1 = sampled), csampled[type], returns probaccept>=1 flag via *full. */
static int atogen_select(const ora_box_t *b, int *ngrsamples, ora_ran1_t *rng,
                         int *origin, float *csampled, int *full)
{
    int nmol = b->nmol;
    if (*ngrsamples <= 0) *ngrsamples = 1000;
    if (*ngrsamples > nmol) *ngrsamples = nmol;
    float probaccept = (float)(*ngrsamples) / (float)nmol;
    int *molsampled = calloc((size_t)b->nmoltype, sizeof(int));
    float *molfraction = calloc((size_t)b->nmoltype, sizeof(float));
    for (int it = 0; it < b->ntype; ++it) csampled[it] = 0.0f;
    for (int im = 0; im < nmol; ++im) molfraction[b->mtype[im]] = molfraction[b->mtype[im]] + 1.0f;
    for (int it = 0; it < b->nmoltype; ++it) molfraction[it] = molfraction[it] / (float)nmol;
    int im1 = 0, im2 = nmol - 1;
    *full = 0;
    if (probaccept >= 1.0f) {
        im2 = nmol - 2;
        *full = 1;
        for (int ia = b->naf[nmol - 1]; ia <= b->nal[nmol - 1]; ++ia)
            csampled[b->itype[ia]] = csampled[b->itype[ia]] + 1.0f;
    }
    for (int im = 0; im < nmol; ++im) origin[im] = 0;
    for (int im = im1; im <= im2; ++im) {
        int mt = b->mtype[im];
        float r = ora_ran1(rng);            /* always drawn: first operand of the or */
        if (r < probaccept || molsampled[mt] < 1 || molfraction[mt] < probaccept) {
            molsampled[mt] += 1;
            origin[im] = 1;
            for (int ia = b->naf[im]; ia <= b->nal[im]; ++ia)
                csampled[b->itype[ia]] = csampled[b->itype[ia]] + 1.0f;
        }
    }
    free(molsampled);
    free(molfraction);
    return 0;
}

/* inter loop (:3780-3837) of origin molecule im: its atoms against every atom of the
 * molecules after it (full) or of every other molecule; cnt: the [npairs][ldw] inter
 * histograms */
static void atogen_origin(const ora_box_t *b, const ora_centres_t *c, int im, int full, float stepcl, int npt1,
                          int ldw, const int *irefpcof, const float *rminex, int64_t *cnt, int64_t *potexcnt)
{
    float xc1 = c->xadd[im], yc1 = c->yadd[im], zc1 = c->zadd[im];
    int jm1 = full ? im + 1 : 0;
    for (int ia = b->naf[im]; ia <= b->nal[im]; ++ia) {
        float x1 = b->xa[ia] + xc1, y1 = b->ya[ia] + yc1, z1 = b->za[ia] + zc1;
        int it = b->itype[ia];
        for (int jm = jm1; jm < b->nmol; ++jm) {
            if (jm == im) continue;
            float xc2 = c->xadd[jm], yc2 = c->yadd[jm], zc2 = c->zadd[jm];
            for (int ja = b->naf[jm]; ja <= b->nal[jm]; ++ja) {
                int icref = b->npairref[it * b->ntype + b->itype[ja]];
                float x2 = b->xa[ja] + xc2, y2 = b->ya[ja] + yc2, z2 = b->za[ja] + zc2;
                float xx = x1 - x2, yy = y1 - y2, zz = z1 - z2;
                ora_minimage(&xx, &yy, &zz);
                float dtest = ora_sqrtf(xx * xx + yy * yy + zz * zz);
                int iz1 = ora_nint(dtest / stepcl) + 2;
                if (iz1 > 1 && iz1 <= npt1) cnt[(size_t)icref * ldw + iz1] += 1;
                float dtesttest = dtest * b->dmax;
                if (dtesttest < 1.0f * fabsf(rminex[irefpcof[icref]])) potexcnt[icref] += 1;
            }
        }
    }
}

/* intra (:3698-3724) and inter (:3780-3837) pair loops, integer counts */
static void atogen_loops(const ora_box_t *b, const ora_centres_t *c, int npt1, int ldw,
                         const int *irefpcof, const float *rminex, const int *origin, int full,
                         int64_t *cnt, int64_t *potexcnt)
{
    const int np = b->npairs;
    const float stepcl = b->cellst / b->dmax;
    const float sf = b->sizefactor;
    for (int im = 0; im < b->nmol; ++im) {
        if (b->nal[im] - b->naf[im] + 1 > 1) {
            for (int ia = b->naf[im]; ia < b->nal[im]; ++ia) {
                float x1 = b->xa[ia], y1 = b->ya[ia], z1 = b->za[ia];
                for (int ja = ia + 1; ja <= b->nal[im]; ++ja) {
                    int icref = b->npairref[b->itype[ia] * b->ntype + b->itype[ja]];
                    float xx = (x1 - b->xa[ja]) * sf;
                    float yy = (y1 - b->ya[ja]) * sf;
                    float zz = (z1 - b->za[ja]) * sf;
                    ora_minimage(&xx, &yy, &zz);
                    float dtest = ora_sqrtf(xx * xx + yy * yy + zz * zz);
                    int iz1 = ora_nint(dtest / stepcl) + 2;
                    if (iz1 > 1 && iz1 <= npt1) cnt[(size_t)icref * ldw + iz1] += 1;
                }
            }
        }
    }
#ifdef ORACLE_OMP
    /* the origin molecules across threads, each counting into its own integer
     * histograms; integer sums do not depend on the order */
#pragma omp parallel
    {
        int64_t *lc = calloc((size_t)np * ldw, sizeof(int64_t));
        int64_t *lp = calloc((size_t)np, sizeof(int64_t));
#pragma omp for schedule(dynamic, 8)
        for (int im = 0; im < b->nmol; ++im)
            if (origin[im]) atogen_origin(b, c, im, full, stepcl, npt1, ldw, irefpcof, rminex, lc, lp);
#pragma omp critical
        {
            for (size_t k = 0; k < (size_t)np * ldw; ++k) cnt[(size_t)np * ldw + k] += lc[k];
            for (int k = 0; k < np; ++k) potexcnt[k] += lp[k];
        }
        free(lc);
        free(lp);
    }
#else
    for (int im = 0; im < b->nmol; ++im)
        if (origin[im]) atogen_origin(b, c, im, full, stepcl, npt1, ldw, irefpcof, rminex, cnt + (size_t)np * ldw,
                                      potexcnt);
#endif
}

/* The reference accumulates +addwav in float32.  Equal power-of-two increments
 * sum exactly up to 2^24 increments and then stall (round-to-even), so the
 * float value is addwav * min(k, 2^24). */
static float fsum_of_equal_increments(int64_t k, float a)
{
    if (k > 16777216) k = 16777216;
    return a * (float)k;
}

int ora_atogen_counts(const ora_box_t *b, const ora_centres_t *c, int npt1, int ldw,
                      const int *irefpcof, const float *rminex, int ngrsamples,
                      ora_ran1_t *rng, int64_t *cnt, int64_t *potexcnt)
{
    int *origin = malloc(sizeof(int) * (size_t)b->nmol);
    float *cs = malloc(sizeof(float) * (size_t)b->ntype);
    int full = 0;
    atogen_select(b, &ngrsamples, rng, origin, cs, &full);
    memset(cnt, 0, sizeof(int64_t) * (size_t)2 * b->npairs * ldw);
    memset(potexcnt, 0, sizeof(int64_t) * (size_t)b->npairs);
    atogen_loops(b, c, npt1, ldw, irefpcof, rminex, origin, full, cnt, potexcnt);
    free(origin);
    free(cs);
    return ngrsamples;
}

/* ato_manipulation_rev4.py:3640-3899 */
int ora_atogen(const ora_box_t *b, const ora_centres_t *c, int npt1, int ldw,
               const float *ycoh, const int *irefpcof, const float *rminex,
               int ngrsamples, ora_ran1_t *rng, float *wav, float *potex,
               float *potexmax, float *csampled, int *origin_flags)
{
    const int np = b->npairs;
    int *origin = malloc(sizeof(int) * (size_t)b->nmol);
    float *cs = malloc(sizeof(float) * (size_t)b->ntype);
    int64_t *cnt = calloc((size_t)2 * np * ldw, sizeof(int64_t));
    int64_t *pcnt = calloc((size_t)np, sizeof(int64_t));
    int full = 0;
    atogen_select(b, &ngrsamples, rng, origin, cs, &full);
    atogen_loops(b, c, npt1, ldw, irefpcof, rminex, origin, full, cnt, pcnt);
    float addwav = full ? 2.0f : 1.0f;
    for (int ic = 0; ic < np; ++ic) {
        potexmax[ic] = 0.0f;
        potex[ic] = fsum_of_equal_increments(pcnt[ic], 1.0f);
        for (int iz = 1; iz <= npt1; ++iz) {
            wav[(size_t)ic * ldw + iz] = fsum_of_equal_increments(cnt[(size_t)ic * ldw + iz], 2.0f);
            wav[((size_t)np + ic) * ldw + iz] =
                fsum_of_equal_increments(cnt[((size_t)np + ic) * ldw + iz], addwav);
        }
    }
    /* normalise sums (:3859-3897) */
    for (int ic = 0; ic < np; ++ic) {
        int it = b->n1extr[ic], jt = b->n2extr[ic];
        float cprod = b->cfrac[it] * b->cfrac[jt] * b->natomr;
        float cprod2 = cs[it] * b->cfrac[jt];
        if (it != jt) {
            cprod = cprod * 2.0f;
            cprod2 = cprod2 + cs[jt] * b->cfrac[it];
        }
        float *w1 = wav + (size_t)ic * ldw, *w2 = wav + ((size_t)np + ic) * ldw;
        if (cprod > 0.0f) {
            for (int iz = 1; iz <= npt1; ++iz) w1[iz] = w1[iz] / cprod;
        } else {
            for (int iz = 1; iz <= npt1; ++iz) w1[iz] = 0.0f;
        }
        if (cprod2 > 0.0f) {
            for (int iz = 1; iz <= npt1; ++iz) w2[iz] = w2[iz] / cprod2;
            potex[ic] = potex[ic] * b->cfrac[jt] / cprod2;
        } else {
            for (int iz = 1; iz <= npt1; ++iz) w2[iz] = ycoh[iz];
        }
    }
    if (csampled) memcpy(csampled, cs, sizeof(float) * (size_t)b->ntype);
    if (origin_flags) memcpy(origin_flags, origin, sizeof(int) * (size_t)b->nmol);
    free(origin); free(cs); free(cnt); free(pcnt);
    return ngrsamples;
}

/* ------------------------------------------------------------------ gensq */
/* epsr_standalone_rev3.py:4730-4802 (coarse graining off) */
void ora_gensq(int nq, int npairs, int npt1, int ldw, const float *wav,
               const float *ycoh, const float *foqols, int ldf,
               const int *n2extr, const float *natomsphere, float *sofq, int ldq)
{
    for (int ic = 0; ic < npairs; ++ic) {
        const float *wav1 = wav + (size_t)ic * ldw;
        const float *wav2 = wav + ((size_t)npairs + ic) * ldw;
        float nas = natomsphere ? natomsphere[n2extr[ic]] : 1.0f;
        for (int i = 1; i <= nq; ++i) {
            const float *fq = foqols + (size_t)i * ldf;
            float sum1 = 0.0f, sum2 = 0.0f, sum3 = 0.0f;
            for (int j = 2; j <= npt1; ++j) {
                float add1 = wav1[j];
                if (fabsf(add1) > 0.0f) sum1 = sum1 + add1 * fq[j];
            }
            for (int j = 2; j <= npt1; ++j) {
                float add2 = (wav2[j] - ycoh[j]) * nas;
                sum2 = sum2 + add2 * fq[j];
                sum3 = sum3 + wav2[j];
            }
            if (sum3 <= 0.0f) sum2 = 0.0f;
            sofq[(size_t)ic * ldq + i] = sum1;
            sofq[((size_t)npairs + ic) * ldq + i] = sum2;
        }
    }
}

/* ------------------------------------------------------------------ stog */
/* ato_manipulation_rev4.py:4114-4124 */
float ora_lorch(float x, float pi)
{
    if (x == 0.0f) return 1.0f;
    if (x < pi) return ora_sinf(x) / x;
    return 0.0f;
}

/* ato_manipulation_rev4.py:4066-4111 */
void ora_stog(float rho, const float *q, const float *sofq, int ldq, const float *r,
              float *gofr, int ldg, const int *nqmin, const int *nqmax, int nr,
              int npairs, const float *xwt)
{
    const float pi = 3.14159265358979323846f;
    float qstep = q[4] - q[3];
    float sfac = 0.5f * qstep / pi / pi / rho;
    for (int j = 1; j <= nr; ++j) {
        float rr = r[j];
        for (int in = 0; in < npairs; ++in) {
            float qmax = pi / q[nqmax[in]];
            const float *sq = sofq + (size_t)in * ldq;
            float sumi = 0.0f;
            for (int i = nqmin[in]; i <= nqmax[in]; ++i) {
                float qq = q[i];
                float prod;
                if (rr > 0.0f) {
                    float qr = qq * rr;
                    prod = qq * ora_sinf(qr) * ora_lorch(qmax * qq, pi);
                } else {
                    prod = qq * qq * ora_lorch(qmax * qq, pi);
                }
                sumi = sumi + sq[i] * prod;
            }
            if (rr > 0.0f) gofr[(size_t)in * ldg + j] = xwt[in] + sfac * sumi / rr;
            else gofr[(size_t)in * ldg + j] = xwt[in] + sfac * sumi;
        }
    }
}
