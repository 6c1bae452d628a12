/*
 * ora_ep.c -- oracle: radial/Q grids, foqols, truncation functions, the
 * reference potential tables (generate_rp), the Poisson EP bases,
 * generate_ep, expand, smoop and the Metropolis fit fit_expansion.
 * TEST INFRASTRUCTURE ONLY (see epsr_oracle.h).
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI4ATAN 3.14159265358979323846f

/* epsr_standalone_rev3.py:2947-2999 (rlog branch inactive, as in the reference) */
void ora_inita_grid(float cellst, float rmaxgr, float rminpt, float rmaxpt, int mpz,
                    float *x, ora_grid_t *g)
{
    float celstt = cellst;
    float x1dold = -0.5f * celstt;
    float cellhf = 0.25f * celstt;
    x[1] = 0.0f;
    int j = 2;
    g->npt1 = 2; g->npt2 = 2; g->nptmin = 2; g->nptmax = 2;
    float x1d = x1dold + celstt;
    while (j < mpz && x1d < rmaxgr) {
        x[j] = x1d - cellhf;
        x1dold = x1d;
        cellhf = 0.5f * celstt;
        if (x1d < rminpt) g->nptmin = j + 1;
        if (x1d < rmaxpt) g->nptmax = j + 1;
        if (x1d < rmaxgr) g->npt1 = j;
        j = j + 1;
        x1d = x1dold + celstt;
    }
    x[j] = x1d - cellhf;
    g->npt = j;
    g->npt2 = j - 2;
    if (g->npt1 > g->npt2) g->npt1 = g->npt2;
}

/* epsr_standalone_rev3.py:5391-5419 */
void ora_uniform(float *ycoh, float xst, float rho, float rmaxgr, int ndim)
{
    const float pi = PI4ATAN;
    float rhofac = 4.0f * pi * rho / 3.0f;
    float x1 = 0.0f, x13 = 0.0f;
    float xstep = xst * 0.5f;
    float x2 = x1 + xstep;
    ycoh[1] = 0.0f;
    int j = 1;
    while (x2 < rmaxgr && j < ndim) {
        j = j + 1;
        x2 = x1 + xstep;
        float x23 = x2 * x2 * x2;
        ycoh[j] = rhofac * (x23 - x13);
        xstep = xst;
        x1 = x2;
        x13 = x23;
    }
}

/* epsr_standalone_rev3.py:3319-3329 (real*8 result; double sin/cos from libm) */
double ora_j1qr(float q, float r)
{
    if (q > 0.0f) {
        double qr = (double)(q * r);
        return (sin(qr) - qr * cos(qr)) / (double)(q * q * q);
    }
    return (double)(r * r * r / 3.0f);
}

/* epsr_standalone_rev3.py:3389-3399 */
float ora_pofr(float delr)
{
    if (delr <= 0.0f) return 1.0f;
    float delr3 = 3.0f / (delr * delr * delr);
    return delr3 * (ora_sinf(delr) - delr * ora_cosf(delr));
}

/* epsr_standalone_rev3.py:3025-3071: foqols[i*ldf + j] = foqols(j, i) */
void ora_foqols(int nq, const float *q, int npt1, const float *x, const float *ycoh,
                float rho, float fwhm, float fwhmq, float lorchqwidth, float lorchqwidthq,
                float rmaxgr, float *foqols, int ldf)
{
    const float pi = PI4ATAN;
    float sqrt2ln2 = ora_sqrtf(2.0f * ora_logf(2.0f));
    float sigres = 0.5f * fwhm / sqrt2ln2;
    float sigrsq = 0.5f * fwhmq / sqrt2ln2;
    float trfac = 4.0f * pi * rho;
    float lorchlim = 4.4934f / rmaxgr;
    double trfac_d = (double)trfac;
    for (int i = 1; i <= nq; ++i) {
        float qq = q[i];
        float r1 = 0.0f;
        double j1qr1 = ora_j1qr(qq, r1);
        float *fo = foqols + (size_t)i * ldf;
        for (int j = 1; j <= npt1; ++j) {
            float rr = x[j];
            float rstep = rr - r1;
            float r2 = rr + rstep;
            double j1qr2 = ora_j1qr(qq, r2);
            if (ycoh[j] > 0.0f) fo[j] = (float)(trfac_d * (j1qr2 - j1qr1) / (double)ycoh[j]);
            else fo[j] = 0.0f;
            r1 = r2;
            j1qr1 = j1qr2;
            float qsig = sigres + qq * sigrsq;
            float qrs = qsig * rr;
            float qr2 = 0.5f * qrs * qrs;
            float sums = 0.0f;
            if (qr2 < 30.0f) {
                sums = ora_expf(-qr2);
                if (lorchqwidth > 0.0f || lorchqwidthq > 0.0f) {
                    float bwidth = rr * lorchlim * (lorchqwidth + lorchqwidthq * qq);
                    sums = sums * ora_pofr(bwidth);
                }
            }
            fo[j] = fo[j] * sums;
        }
    }
}

/* ato_manipulation_rev4.py:693-712 */
float ora_truncate(float xx, float rminpt, float rmaxpt)
{
    const float pi = PI4ATAN;
    float decay = rmaxpt - rminpt;
    if (xx >= rmaxpt) return 0.0f;
    if (xx <= rminpt) return 1.0f;
    float rat = (xx - rminpt) * pi / decay;
    return 0.5f * (1.0f + ora_cosf(rat));
}

/* ato_manipulation_rev4.py:715-727 */
float ora_truncatec(float xx, float rmaxpt)
{
    if (xx >= rmaxpt) return 0.0f;
    float rat = xx / rmaxpt;
    float a = 1.0f - rat;
    float a2 = a * a;
    float t = a2 * a2;
    return t * (1.0f + rat * (1.6f + rat * 0.4f));
}

/* ato_manipulation_rev4.py:614-653 */
static float sum1(float r, float a, float b, float aup, float alo)
{
    float cof[7];
    float a2 = a * a, b2 = b * b, r2 = r * r;
    cof[6] = 1.0f / 6.0f;
    cof[5] = 3.0f * r / 5.0f;
    cof[4] = -(3.0f * (a2 - r2) + b2) / 4.0f;
    cof[3] = -r * (3.0f * (a2 + b2) - r2) / 3.0f;
    cof[2] = 3.0f * b2 * (a2 - r2) / 2.0f;
    cof[1] = r * b2 * (3.0f * a2 - r2);
    float fac = 3.0f / (8.0f * a2 * a * b2 * b * r);
    float sumu = 0.0f, suml = 0.0f;
    for (int i = 1; i <= 6; ++i) {
        int ir = 7 - i;
        sumu = aup * (sumu + cof[ir]);
        suml = alo * (suml + cof[ir]);
    }
    return fac * (sumu - suml);
}

static float sum2(float r, float a, float b, float aup, float alo)
{
    float cof[4];
    (void)a;
    float b2 = b * b;
    cof[3] = -1.0f / 3.0f;
    cof[2] = 0.0f;
    cof[1] = b2;
    float fac = 3.0f / (4.0f * r * b2 * b);
    float sumu = 0.0f, suml = 0.0f;
    for (int i = 1; i <= 3; ++i) {
        int ir = 4 - i;
        sumu = aup * (sumu + cof[ir]);
        suml = alo * (suml + cof[ir]);
    }
    return fac * (sumu - suml);
}

/* ato_manipulation_rev4.py:589-611 */
float ora_chsph(float sig1, float sig2, float r)
{
    float ra = sig1, rb = sig2;
    if (sig2 > sig1) { ra = sig2; rb = sig1; }
    if (r >= (ra + rb)) return 1.0f / r;
    if (r <= (ra - rb)) {
        if (ra > 0.0f) return (3.0f * (ra * ra - 0.2f * rb * rb) - r * r) / (2.0f * ra * ra * ra);
        return 1.0e20f;
    }
    return sum1(r, ra, rb, ra - r, -rb) + sum2(r, ra, rb, rb, ra - r);
}

/* epsr_standalone_rev3.py:5675-5872: n-6 LJ + Coulomb (+ exponential repulsive
 * core); no extra Gaussians, no intramolecular copy, no container forces. */
void ora_generate_rp(int npt, int npt2, const float *x, const float *xstar, const float *wfac,
                     const float *wfacc, int npairs, const float *ecor, const float *dcor,
                     const float *qcor, const float *qradius_i, const float *qradius_j,
                     float power, float refpotfac, const float *pcoefex, const float *rminex,
                     float expecf, float expecfpwr,
                     float *epot, float *ecoul, float *epot1, float *ecoul1, int ld)
{
    float gradlimit = expecfpwr, loggradlimit;
    if (gradlimit > 0.0f) loggradlimit = ora_logf(gradlimit);
    else { gradlimit = 0.0f; loggradlimit = 0.0f; }
    float rat = 6.0f / power;
    float rat1 = ora_sqrtf(rat);
    float rat2 = 1.0f / rat1;
    float power1 = (power + 6.0f) / (power - 6.0f);
    float rat3 = ora_powf(rat1, power1);
    float prefac = (rat2 - rat1) * rat3;
    prefac = 1.0f / prefac;
    for (int ic = 0; ic < npairs; ++ic) {
        float *ep = epot + (size_t)ic * ld, *ec = ecoul + (size_t)ic * ld;
        float *ep1 = epot1 + (size_t)ic * ld, *ec1 = ecoul1 + (size_t)ic * ld;
        float ra = qradius_i[ic], rb = qradius_j[ic];
        float ect = ecor[ic], qct = qcor[ic] * 1.0f, dc = dcor[ic];
        for (int j = 2; j <= npt; ++j) {
            float x2 = x[j];
            float ddif = dc / x2;
            float sumlj;
            if (ddif > 0.0f && ect > 0.0f) {
                float fac = prefac * ect;
                float xx = ora_logf(ddif);
                float pwr1 = 0.5f * (power + 6.0f) * xx;
                if (fabsf(pwr1) > 20.0f) pwr1 = copysignf(20.0f, pwr1);
                float pwr2 = 0.5f * (power - 6.0f) * xx;
                if (fabsf(pwr2) > 40.0f) pwr2 = copysignf(40.0f, pwr2);
                sumlj = 2.0f * fac * ora_sinhf(pwr2) * ora_expf(pwr1);
            } else {
                sumlj = 0.0f;
            }
            float sumc = qct * ora_chsph(ra, rb, x2);
            sumlj = refpotfac * sumlj * wfac[j];
            sumc = refpotfac * sumc * wfacc[j];
            ep[j] = sumlj + sumc;
            ec[j] = sumlj + sumc;
        }
        if (pcoefex && rminex && fabsf(pcoefex[ic]) > 0.0f && fabsf(rminex[ic]) > 0.0f) {
            float expecfuse = expecf * ora_sqrtf(fabsf(rminex[ic]));
            float rminextest = fabsf(rminex[ic]);
            for (int j = 1; j <= npt2; ++j) {
                float ddiff = (rminextest - x[j]) / expecfuse;
                float add;
                if (gradlimit > 0.0f && ddiff >= loggradlimit)
                    add = pcoefex[ic] * gradlimit * (1.0f + ddiff - loggradlimit);
                else
                    add = pcoefex[ic] * ora_expf(ddiff);
                ec[j] = ec[j] + add * wfac[j];
            }
        }
        ep[1] = ep[2];
        ec[1] = ec[2];
        float s1 = ep[1], c1 = ec[1];
        float x1 = xstar[1];
        for (int j = 2; j <= npt; ++j) {
            float x2 = xstar[j];
            float xdif = x2 - x1;
            float s2 = ep[j], c2 = ec[j];
            ep1[j - 1] = (s2 - s1) / xdif;
            ec1[j - 1] = (c2 - c1) / xdif;
            x1 = x2; s1 = s2; c1 = c2;
        }
        ep1[npt] = 0.0f;
        ec1[npt] = 0.0f;
    }
}

/* epsr_standalone_rev3.py:4924-4939 */
float ora_expand(int nterms, int j, const float *expr, const float *coeff, float summax)
{
    float sum = 0.0f;
    size_t jref0 = (size_t)(j - 1) * nterms;
    for (int i = 1; i <= nterms; ++i) {
        sum = sum + expr[jref0 + i] * coeff[i];
        if (fabsf(sum) > summax) sum = copysignf(summax, sum);
    }
    return sum;
}

/* epsr_standalone_rev3.py:5313-5326 */
float ora_dnext(float d, float dfac, float dstep, float *rootd)
{
    float dn = d + dstep;
    float dtest = d * dfac;
    if (dtest > dn) {
        dn = dtest;
        *rootd = *rootd * ora_sqrtf(dfac);
    }
    return dn;
}

/* epsr_standalone_rev3.py:5100-5153 (expr is 1-based flat: expr[(i-1)*nterms + m]) */
int ora_calc_pwrexpq(int lpt, int nterms, float dmin, float dstep, float dmax, float sigmab,
                     float rho, const float *xin, float *expr, float rmin)
{
    (void)dmax; (void)rho;
    int ncfirst = 0;
    float d = dstep, rootd = 1.0f;
    for (int m = 1; m <= nterms; ++m) {
        if (d < rmin) ncfirst = m + 1;
        float sigma = rootd * sigmab;
        int n = (sigma <= d && sigma > 0.0f && m > 1) ? ora_ifix(d / sigma - 1.0f) : 0;
        float an2 = (float)(n + 2);
        float an2fac = 1.0f / an2;
        for (int i = 1; i <= lpt; ++i) {
            float q = xin[i];
            float qsigma = q * sigma;
            float qs2 = qsigma * qsigma;
            float onemqs2 = 1.0f - qs2;
            float onepqs2 = 1.0f + qs2;
            float rtonepqs2 = ora_sqrtf(onepqs2);
            float alpha = ora_atanf(qsigma);
            float alphan = (float)n * alpha;
            float term;
            if (qsigma > 0.0f) term = 2.0f * ora_cosf(alphan) + onemqs2 * ora_sinf(alphan) / qsigma;
            else term = an2;
            float pn = term * an2fac / ora_powi(rtonepqs2, 4 + n);
            expr[(size_t)(i - 1) * nterms + m] = pn;
        }
        d = ora_dnext(d, dmin, dstep, &rootd);
    }
    return ncfirst;
}

/* lngamma_mod.py:15-30: gamln[i] = log((i-1)!) for i = 1..nterms+1 */
static double *fctrln(int nterms)
{
    int ncalc = nterms + 1;
    double *g = calloc((size_t)ncalc + 2, sizeof(double));
    double a = 0.0, v = 0.0;
    for (int i = 1; i <= ncalc; ++i) {
        g[i] = v;
        a = a + 1.0;
        v = v + log(a);
    }
    return g;
}

/* epsr_standalone_rev3.py:5229-5310 */
int ora_calc_pwrexpr_grad(int lpt, int nterms, float dmin, float dstep, float dmax,
                          float sigmab, float rho, const float *xin, float *expr,
                          float rmin, float fbroad, float rbroad)
{
    const float pi = 3.141592653f;
    const float expmax = 25.0f;
    int ntmax = ora_ifix(dmax / sigmab + 3.0f);
    double *gamln = fctrln(ntmax);
    int ncfirst = 0;
    float d = dstep, rootd = 1.0f;
    for (int m = 1; m <= nterms; ++m) {
        if (d < rmin) ncfirst = m + 1;
        float sigma = rootd * fbroad * sigmab;
        int n = (sigma <= d && sigma > 0.0f && m > 1) ? ora_ifix(d / sigma - 1.0f) : 0;
        int nref = n + 3;
        if (nref > ntmax + 1) {   /* gamln(1..ntmax+1): the reference raises IndexError */
            free(gamln);
            return -1;
        }
        float gfac = 4.0f * pi * rho * sigma * sigma * sigma;
        float gfacln = ora_logf(gfac);
        float sigln = (float)(gamln[nref] + (double)gfacln);
        if (m == 1) sigln = sigln - rmin / sigma;
        for (int i = 1; i <= lpt; ++i) {
            size_t jref = (size_t)(i - 1) * nterms + m;
            float r = xin[i];
            float fact = rbroad * r;
            float rsigma = r / sigma;
            float expon;
            if (r > 0.0f) {
                float rsigmaln = ora_logf(rsigma);
                expon = (float)n * rsigmaln - rsigma - sigln - fact;
                fact = (float)n / r - 1.0f / sigma - rbroad;
            } else if (n == 0) {
                expon = -rsigma - sigln - fact;
                fact = -(1.0f / sigma + rbroad);
            } else if (n == 1) {
                expon = -rsigma - sigln - fact;
                fact = 1.0f / sigma;
            } else {
                expon = -2.0f * expmax;
                fact = 0.0f;
            }
            expr[jref] = expon > -expmax ? fact * ora_expf(expon) : 0.0f;
        }
        d = ora_dnext(d, dmin, dstep, &rootd);
    }
    free(gamln);
    return ncfirst;
}

/* epsr_standalone_rev3.py:5875-6023 with abs_energy_ep (ato_manipulation_rev4.py:4127-4157)
 * and multiply_pot_coeffs (:6493-6510).  Poisson basis, npgauss = 0. */
float ora_generate_ep(int npt, int npt2, const float *x, const float *xstar, const float *wfac,
                      int npairs, const int *irefpcof, int nterms, float *pcoefp, int ldp,
                      float pdmin, float pdstep, float pdmax, float psigma2, float rho,
                      float rbroad, float dmax, float absolute_energy, float erequnit,
                      float ereqstep, const float *ecoul, float *xavs, float *xavs1,
                      float *xavtot, float *xavtot1, int ld, float *energabs_out)
{
    const float summax = (float)exp(25.0);
    float *expgr = calloc((size_t)npt2 * nterms + 1, sizeof(float));
    float *pwtold = calloc((size_t)nterms + 1, sizeof(float));
    if (ora_calc_pwrexpr_grad(npt2, nterms, pdmin, pdstep, pdmax, psigma2, rho, x, expgr, 0.0f, 1.0f,
                              rbroad) < 0) {
        free(expgr);
        free(pwtold);
        return NAN;   /* a basis position needs ln((n+2)!) beyond the tabulated range */
    }
    for (int ip = 0; ip < npairs; ++ip) {
        const float *pc = pcoefp + (size_t)irefpcof[ip] * ldp;
        for (int ic = 2; ic <= nterms; ++ic) pwtold[ic] = pc[ic];
        pwtold[1] = 0.0f;
        for (int j = 1; j <= npt2; ++j) {
            float term = ora_expand(nterms, j, expgr, pwtold, summax);
            xavs1[(size_t)ip * ld + j] = term * pc[1];
        }
    }
    for (int ip = 0; ip < npairs; ++ip) {
        float *xs = xavs + (size_t)ip * ld, *xs1 = xavs1 + (size_t)ip * ld;
        for (int j = 1; j <= npt2; ++j) xs1[j] = xs1[j] * dmax;
        xs[npt] = 0.0f;
        xs1[npt] = 0.0f;
        float x1 = xstar[npt], xts1 = xs1[npt];
        for (int j = 1; j <= npt2; ++j) {
            int jref = npt - j;
            float x2 = xstar[jref], xts2 = xs1[jref];
            xs[jref] = xs[jref + 1] - 0.5f * (x1 - x2) * (xts1 + xts2);
            x1 = x2;
            xts1 = xts2;
        }
        for (int j = 1; j <= npt2; ++j) xs[j] = xs[j] * wfac[j];
    }
    /* abs_energy_ep over the coefficient sets referenced by the pairs */
    int maxref = 0;
    for (int ip = 0; ip < npairs; ++ip) if (irefpcof[ip] + 1 > maxref) maxref = irefpcof[ip] + 1;
    float energabs = 0.0f;
    for (int ip = 0; ip < maxref; ++ip) {
        float mn = 0.0f, mx = 0.0f;
        for (int i = 2; i <= nterms; ++i) {
            float term = pcoefp[(size_t)ip * ldp + i];
            mn = fminf(mn, term);
            mx = fmaxf(mx, term);
        }
        energabs = fmaxf(energabs, mx - mn);
    }
    float pressfac = 1.0f;
    if (fabsf(absolute_energy) == 0.0f) {
        pressfac = 0.0f;
        energabs = 0.0f;
    } else if (fabsf(energabs) > 0.0f) {
        if (fabsf(energabs) < fabsf(erequnit)) pressfac = fabsf(erequnit) / fabsf(energabs);
        else {
            pressfac = fabsf(absolute_energy) / fabsf(energabs);
            if (pressfac > 1.0f && ereqstep == 0.0f) pressfac = 1.0f;
        }
    }
    for (int ip = 0; ip < npairs; ++ip) {
        float *pc = pcoefp + (size_t)irefpcof[ip] * ldp;
        for (int ic = 2; ic <= nterms; ++ic) pc[ic] = pc[ic] * pressfac;
    }
    for (int ip = 0; ip < npairs; ++ip) {
        float *xs = xavs + (size_t)ip * ld, *xs1 = xavs1 + (size_t)ip * ld;
        float *xt = xavtot + (size_t)ip * ld, *xt1 = xavtot1 + (size_t)ip * ld;
        const float *ec = ecoul + (size_t)ip * ld;
        for (int j = 1; j <= npt2; ++j) {
            xs[j] = xs[j] * pressfac;
            xt[j] = ec[j] + xs[j];
        }
        float s1 = xs[1], t1 = xt[1];
        float x1 = xstar[1];
        for (int j = 2; j <= npt; ++j) {
            float x2 = xstar[j];
            float xdif = x2 - x1;
            float s2 = xs[j], t2 = xt[j];
            xs1[j - 1] = (s2 - s1) / xdif;
            xt1[j - 1] = (t2 - t1) / xdif;
            x1 = x2; s1 = s2; t1 = t2;
        }
        xs1[npt] = 0.0f;
        xt1[npt] = 0.0f;
    }
    free(expgr);
    free(pwtold);
    if (energabs_out) *energabs_out = energabs;
    return pressfac;
}

/* auxiliary_routines.py:744-764 */
void ora_smoop(int nqmin, int nqmax, int nsmoo, float *diff)
{
    int n1 = nqmin, n2 = nqmax;
    if (n2 - 2 < n1 || nsmoo == 0) return;
    for (int it = 1; it <= nsmoo; ++it) {
        float sm = 0.0f, sn = diff[n1];
        float s1 = 0.25f * (sm + sn);
        for (int i = n1; i <= n2; ++i) {
            sm = sn;
            int j = i + 1;
            sn = j <= n2 ? diff[j] : 0.0f;
            float s2 = 0.25f * (sm + sn);
            diff[i] = s1 + s2;
            s1 = s2;
        }
    }
}

static float chisq(const float *yin, const float *expr, const float *wt, int nmin, int npt,
                   int nterms, float summax)
{
    float c = 0.0f;
    for (int i = nmin; i <= npt; ++i) {
        float term = yin[i] - ora_expand(nterms, i, expr, wt, summax);
        c = c + term * term;
    }
    return c;
}

/* epsr_standalone_rev3.py:4805-4920 */
void ora_fit_expansion(const float *yin, const float *expr, float *yout, float *wt, float *wtold,
                       float *stepr, float *acceptr, float *rejrat, float *chisqnew, int nmin,
                       int npt, int ncfirst, int nterms, int nits, int nfreq, ora_ran1_t *rng,
                       int *ntries_out)
{
    const float summax = (float)exp(25.0);
    const int nstep = 100, ntry = 150;
    *acceptr = 0.05f;
    int npoints = npt - nmin + 1;
    float wtmax = 0.0f;
    for (int it = 1; it <= nterms; ++it) {
        wt[it] = wtold[it];
        wtmax = fmaxf(fabsf(wt[it]), wtmax);
    }
    if (*stepr < 1.0e-6f * wtmax) *stepr = 2.0e-6f * wtmax;
    int itry = 0;
    if (nits > 0) {
        int icount = 0;
        float cnew = chisq(yin, expr, wt, nmin, npt, nterms, summax);
        float chisqold = cnew;
        float chisqtest = -1.0f;
        while (itry < 10 || (itry < ntry && chisqtest < 1.0e-4f)) {
            itry = itry + 1;
            float accept = *acceptr * chisqold / (float)npoints;
            float rr = 0.0f;
            float chisqstart = chisqold / (float)npoints;
            for (int it = 1; it <= nstep; ++it) {
                for (int ic = ncfirst; ic <= nterms; ++ic) {
                    if (fabsf(wtold[ic]) < 1.0f) wt[ic] = wtold[ic] + *stepr * (2.0f * ora_ran1(rng) - 1.0f);
                    else wt[ic] = wtold[ic] * (1.0f + *stepr * (2.0f * ora_ran1(rng) - 1.0f));
                }
                cnew = chisq(yin, expr, wt, nmin, npt, nterms, summax);
                float exponent = (cnew - chisqold) / accept;
                float test;
                if (exponent > 0.0f && exponent <= 20.0f) test = ora_expf(-exponent) - ora_ran1(rng);
                else if (exponent > 20.0f) test = -1.0f;
                else test = 1.0f;
                if (test >= 0.0f) {
                    icount = icount + 1;
                    if (icount == nfreq) {
                        icount = 0;
                        ora_smoop(ncfirst, nterms, 1, wt);
                        cnew = chisq(yin, expr, wt, nmin, npt, nterms, summax);
                    }
                    chisqold = cnew;
                    for (int ic = 1; ic <= nterms; ++ic) wtold[ic] = wt[ic];
                } else {
                    rr = rr + 1.0f;
                }
            }
            rr = rr / (float)nstep;
            *rejrat = rr;
            float chisqfinish = chisqold / (float)npoints;
            if (chisqstart > 0.0f) chisqtest = (chisqfinish - chisqstart) / chisqstart;
            else chisqtest = chisqfinish - chisqstart;
            float rejfac = rr > 0.0f ? 0.75f / rr : 2.0f;
            *stepr = *stepr * rejfac;
        }
        for (int ic = 1; ic <= nterms; ++ic) wt[ic] = wtold[ic];
    }
    float c = 0.0f;
    for (int i = 1; i <= npt; ++i) {
        float term = ora_expand(nterms, i, expr, wt, summax);
        yout[i] = term;
        if (i >= nmin) {
            float d = yin[i] - term;
            c = c + d * d;
        }
    }
    *chisqnew = c;
    if (ntries_out) *ntries_out = itry;
}

/* epsr_standalone_rev3.py:3827-3867 with neutron weights (ix = 0): the difference
 * between each dataset and its cwt-weighted intermolecular partials */
void ora_ptgnsq_diff(int nfile, int nq, int npairs, const float *q, const float *data, int ldq,
                     const float *qwind, const float *cwt, const float *szeros, const int *nqmin,
                     const int *nqmax, float qminptemp, float *qtemp)
{
    ora_ptgnsq_diff_x(nfile, nq, npairs, q, data, ldq, qwind, cwt, NULL, NULL, szeros, nqmin, nqmax, qminptemp,
                      qtemp);
}

float ora_ptgnsq_rfactor(int i0, int i1, const float *fit, const float *qt)
{
    float r = 0.0f;
    for (int i = i0; i <= i1; ++i) {
        if (fabsf(qt[i]) > 0.0f) r = r + fabsf(fit[i] + qt[i]) / fabsf(qt[i]);
        else r = r + fabsf(fit[i] + qt[i]);
    }
    return r / (float)(i1 - i0 + 1);
}
