// Potential tables on the GPU: truncation functions (initpt), the reference potential
// (generate_rp), the Poisson-basis empirical potential (calc_pwrexpr_grad, expand,
// generate_ep) and the Q-space basis used by the EP fit (calc_pwrexpq).
//
// Reference: src/epsr26/programs/epsr_standalone_rev3.py:3732-3786 (initpt),
// 5675-5872 (generate_rp), 5875-6023 (generate_ep), 5100-5153 (calc_pwrexpq),
// 5229-5310 (calc_pwrexpr_grad), 5313-5326 (dnext), 4924-4939 (expand);
// src/epsr26/library/ato_manipulation_rev4.py:589-730 (chsph, truncate, truncatec),
// 4127-4157 (abs_energy_ep); src/epsr26/library/lngamma_mod.py:15-30 (fctrln).
//
// Element-wise work (one lane per table entry or basis value) runs on the GPU with
// correctly rounded expf/logf/sinhf/cosf/sinf/atanf; the O(nterms) serial parameter
// recurrences (dnext, gamln) and the O(npairs*nterms) coefficient bookkeeping
// (abs_energy_ep, multiply_pot_coeffs) run on the host.
#include <cmath>

#include "common.h"

__device__ __forceinline__ float cr_expf(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float cr_logf(float x) { return (float)log((double)x); }
__device__ __forceinline__ float cr_cosf(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float cr_sinf(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cr_sinhf(float x) { return (float)sinh((double)x); }
__device__ __forceinline__ float cr_atanf(float x) { return (float)atan((double)x); }

// libgcc __powisf2 (src/epsr26/fortran/fmath.py:186-200)
__device__ __forceinline__ float dev_powi(float x, int m)
{
    unsigned n = m < 0 ? (unsigned)(-m) : (unsigned)m;
    float y = (n & 1u) ? x : 1.0f;
    while (true) {
        n >>= 1;
        if (n == 0) break;
        x = x * x;
        if (n & 1u) y = y * x;
    }
    return m < 0 ? 1.0f / y : y;
}

// ------------------------------------------------------------------ truncation
// truncate (:693-712) and truncatec (:715-727); wfac(npt) = wfacc(npt) = 0 (initpt)
__global__ void truncation_kernel(int npt, const float *__restrict__ x, float rminpt, float rmaxpt,
                                  float *__restrict__ wfac, float *__restrict__ wfacc)
{
    const float pi = 3.14159265358979323846f;
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > npt) return;
    float xx = x[j];
    float w;
    float decay = rmaxpt - rminpt;
    if (xx >= rmaxpt) w = 0.0f;
    else if (xx <= rminpt) w = 1.0f;
    else {
        float rat = (xx - rminpt) * pi / decay;
        w = 0.5f * (1.0f + cr_cosf(rat));
    }
    float wc;
    if (xx >= rmaxpt) wc = 0.0f;
    else {
        float rat = xx / rmaxpt;
        float a = 1.0f - rat;
        float a2 = a * a;
        wc = a2 * a2;
        wc = wc * (1.0f + rat * (1.6f + rat * 0.4f));
    }
    if (j == npt) {
        w = 0.0f;
        wc = 0.0f;
    }
    wfac[j] = w;
    wfacc[j] = wc;
}

extern "C" int epsr_truncation(epsr_ctx *ctx, int npt, const float *x, float rminpt, float rmaxpt, float *wfac,
                               float *wfacc)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, x && wfac && wfacc && npt >= 2, "epsr_truncation: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    truncation_kernel<<<ceil_div(npt, 256), 256, 0, ctx->stream>>>(npt, x, rminpt, rmaxpt, wfac, wfacc);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return EPSR_OK;
}

// ------------------------------------------------------------------ generate_rp
// chsph (:589-611) with sum1/sum2 (:614-653)
__device__ float dev_sum1(float r, float a, float b, float aup, float alo)
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
    float su = 0.0f, sl = 0.0f;
    for (int ir = 6; ir >= 1; --ir) {
        su = aup * (su + cof[ir]);
        sl = alo * (sl + cof[ir]);
    }
    return fac * (su - sl);
}

__device__ float dev_sum2(float r, float b, float aup, float alo)
{
    float b2 = b * b;
    float cof1 = b2, cof2 = 0.0f, cof3 = -1.0f / 3.0f;
    float fac = 3.0f / (4.0f * r * b2 * b);
    float su = 0.0f, sl = 0.0f;
    su = aup * (su + cof3); sl = alo * (sl + cof3);
    su = aup * (su + cof2); sl = alo * (sl + cof2);
    su = aup * (su + cof1); sl = alo * (sl + cof1);
    return fac * (su - sl);
}

__device__ float dev_chsph(float sig1, float sig2, float r)
{
    float ra = sig1, rb = sig2;
    if (sig2 > sig1) {
        ra = sig2;
        rb = sig1;
    }
    if (r >= (ra + rb)) return 1.0f / r;
    if (r <= (ra - rb)) {
        if (ra > 0.0f) return (3.0f * (ra * ra - 0.2f * rb * rb) - r * r) / (2.0f * ra * ra * ra);
        return 1.0e20f;
    }
    return dev_sum1(r, ra, rb, ra - r, -rb) + dev_sum2(r, rb, rb, ra - r);
}

// values: one lane per (pair, j); n-6 LJ (:5744-5760) + Coulomb (:5761-5767) and the
// exponential repulsive core (:5806-5822)
__global__ void generate_rp_values(int npt, int npt2, const float *__restrict__ x, const float *__restrict__ wfac,
                                   const float *__restrict__ wfacc, int npairs, const float *__restrict__ ecor,
                                   const float *__restrict__ dcor, const float *__restrict__ qcor,
                                   const float *__restrict__ qri, const float *__restrict__ qrj,
                                   const float *__restrict__ pcoefex, const float *__restrict__ rminex, float power,
                                   float prefac, float refpotfac, float expecf, float gradlimit, float loggradlimit,
                                   float *__restrict__ epot, float *__restrict__ ecoul, int ld)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * npt) return;
    int ic = t / npt;
    int j = t - ic * npt + 1;
    if (j < 2) return;
    float ra = qri ? qri[ic] : 0.0f, rb = qrj ? qrj[ic] : 0.0f;
    float ect = ecor[ic], qct = qcor[ic] * 1.0f, dc = dcor[ic];
    float x2 = x[j];
    float ddif = dc / x2;
    float sumlj = 0.0f;
    if (ddif > 0.0f && ect > 0.0f) {
        float fac = prefac * ect;
        float xx = cr_logf(ddif);
        float pwr1 = 0.5f * (power + 6.0f) * xx;
        if (fabsf(pwr1) > 20.0f) pwr1 = copysignf(20.0f, pwr1);
        float pwr2 = 0.5f * (power - 6.0f) * xx;
        if (fabsf(pwr2) > 40.0f) pwr2 = copysignf(40.0f, pwr2);
        sumlj = 2.0f * fac * cr_sinhf(pwr2) * cr_expf(pwr1);
    }
    float sumc = qct * dev_chsph(ra, rb, x2);
    sumlj = refpotfac * sumlj * wfac[j];
    sumc = refpotfac * sumc * wfacc[j];
    float v = sumlj + sumc;
    size_t o = (size_t)ic * ld + j;
    epot[o] = v;
    float ec = v;
    if (pcoefex && rminex && fabsf(pcoefex[ic]) > 0.0f && fabsf(rminex[ic]) > 0.0f && j <= npt2) {
        float expecfuse = expecf * sqrtf(fabsf(rminex[ic]));
        float rminextest = fabsf(rminex[ic]);
        float dd = (rminextest - x[j]) / expecfuse;
        float add;
        if (gradlimit > 0.0f && dd >= loggradlimit) add = pcoefex[ic] * gradlimit * (1.0f + dd - loggradlimit);
        else add = pcoefex[ic] * cr_expf(dd);
        ec = ec + add * wfac[j];
    }
    ecoul[o] = ec;
}

// j = 1 gets the exponential core term too (the core loop runs from j = 1), then
// epot(1) = epot(2), ecoul(1) = ecoul(2) (:5824-5826) -- note ecoul(1) is overwritten
__global__ void generate_rp_ends(int npt, int npairs, float *__restrict__ epot, float *__restrict__ ecoul, int ld)
{
    int ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= npairs) return;
    epot[(size_t)ic * ld + 1] = epot[(size_t)ic * ld + 2];
    ecoul[(size_t)ic * ld + 1] = ecoul[(size_t)ic * ld + 2];
}

// forward-difference gradients (:5830-5850): g(j-1) = (v(j) - v(j-1)) / (xstar(j) - xstar(j-1)), g(npt) = 0
__global__ void gradients_kernel(int npt, int npairs, const float *__restrict__ xstar, const float *__restrict__ v1,
                                 float *__restrict__ g1, const float *__restrict__ v2, float *__restrict__ g2, int ld)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * npt) return;
    int ic = t / npt;
    int j = t - ic * npt + 1;
    size_t o = (size_t)ic * ld;
    if (j == npt) {
        g1[o + npt] = 0.0f;
        if (g2) g2[o + npt] = 0.0f;
        return;
    }
    // entry j is the difference of j+1 and j
    float xdif = xstar[j + 1] - xstar[j];
    g1[o + j] = (v1[o + j + 1] - v1[o + j]) / xdif;
    if (g2) g2[o + j] = (v2[o + j + 1] - v2[o + j]) / xdif;
}

extern "C" int epsr_generate_rp(epsr_ctx *ctx, int npt, int npt2, const float *x, const float *xstar,
                                const float *wfac, const float *wfacc, const epsr_rp_params_t *p, float *epot,
                                float *ecoul, float *epot1, float *ecoul1, int ld)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, x && xstar && wfac && wfacc && p && epot && ecoul && epot1 && ecoul1, "epsr_generate_rp: null");
    ARG_CHECK(ctx, p->ecor && p->dcor && p->qcor && npt >= 3 && ld > npt, "epsr_generate_rp: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // scalar constants (:5719-5728), float32 with correctly rounded powf
    float power = p->power;
    float rat = 6.0f / power;
    float rat1 = sqrtf(rat);
    float rat2 = 1.0f / rat1;
    float power1 = (power + 6.0f) / (power - 6.0f);
    float rat3 = (float)pow((double)rat1, (double)power1);
    float prefac = (rat2 - rat1) * rat3;
    prefac = 1.0f / prefac;
    float gradlimit = p->expecfpwr, loggradlimit = 0.0f;
    if (gradlimit > 0.0f) loggradlimit = (float)log((double)gradlimit);
    else gradlimit = 0.0f;
    int n = p->npairs * npt;
    generate_rp_values<<<ceil_div(n, 256), 256, 0, st>>>(npt, npt2, x, wfac, wfacc, p->npairs, p->ecor, p->dcor, p->qcor,
                                                         p->qradius_i, p->qradius_j, p->pcoefex, p->rminex, power, prefac,
                                                         p->refpotfac, p->expecf, gradlimit, loggradlimit, epot, ecoul,
                                                         ld);
    generate_rp_ends<<<ceil_div(p->npairs, 64), 64, 0, st>>>(npt, p->npairs, epot, ecoul, ld);
    gradients_kernel<<<ceil_div(n, 256), 256, 0, st>>>(npt, p->npairs, xstar, epot, epot1, ecoul, ecoul1, ld);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    return EPSR_OK;
}

// ------------------------------------------------------------------ EP bases
struct BasisTerm {
    float sigma, sigln;
    int n;
};

// the serial per-term parameters of calc_pwrexpr_grad / calc_pwrexpq: positions by
// dnext (:5313-5326), powers n, and the log normalisation with gamln = log((n+2)!)
static bool basis_terms(int nterms, float dmin, float dstep, float dmax, float sigmab, float rho, float rmin,
                        float fbroad, bool rspace, std::vector<BasisTerm> &out, int &ncfirst)
{
    const float pi = 3.141592653f;
    out.resize(nterms + 1);
    int ntmax = (int)std::trunc((double)(dmax / sigmab + 3.0f));
    std::vector<double> gamln(ntmax + 3, 0.0);
    {
        double g = 0.0, a = 0.0;
        for (int i = 1; i <= ntmax + 1; ++i) {
            gamln[i] = g;
            a = a + 1.0;
            g = g + std::log(a);
        }
    }
    ncfirst = 0;
    float d = dstep, rootd = 1.0f;
    for (int m = 1; m <= nterms; ++m) {
        if (d < rmin) ncfirst = m + 1;
        BasisTerm bt;
        if (rspace) {
            float sigma = rootd * fbroad * sigmab;
            int n = (sigma <= d && sigma > 0.0f && m > 1) ? (int)std::trunc((double)(d / sigma - 1.0f)) : 0;
            float gfac = 4.0f * pi * rho * sigma * sigma * sigma;
            float gfacln = (float)std::log((double)gfac);
            int nref = n + 3;
            if (nref > ntmax + 1) return false;   // beyond gamln(1..ntmax+1)
            float sigln = (float)(gamln[nref] + (double)gfacln);
            if (m == 1) sigln = sigln - rmin / sigma;
            bt.sigma = sigma;
            bt.sigln = sigln;
            bt.n = n;
        } else {
            float sigma = rootd * sigmab;
            int n = (sigma <= d && sigma > 0.0f && m > 1) ? (int)std::trunc((double)(d / sigma - 1.0f)) : 0;
            bt.sigma = sigma;
            bt.sigln = 0.0f;
            bt.n = n;
        }
        out[m] = bt;
        float dn = d + dstep, dtest = d * dmin;
        if (dtest > dn) {
            dn = dtest;
            rootd = rootd * sqrtf(dmin);
        }
        d = dn;
    }
    return true;
}

// calc_pwrexpr_grad element (:5266-5306); expr flat 1-based [(i-1)*nterms + m]
__global__ void pwrexpr_grad_kernel(int lpt, int nterms, const BasisTerm *__restrict__ bt, const float *__restrict__ xin,
                                    float rbroad, float *__restrict__ expr)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lpt * nterms) return;
    int i = t / nterms + 1;
    int m = t - (i - 1) * nterms + 1;
    const float expmax = 25.0f;
    BasisTerm b = bt[m];
    float r = xin[i];
    float fact = rbroad * r;
    float rsigma = r / b.sigma;
    float expon;
    if (r > 0.0f) {
        float rsigmaln = cr_logf(rsigma);
        expon = (float)b.n * rsigmaln - rsigma - b.sigln - fact;
        fact = (float)b.n / r - 1.0f / b.sigma - rbroad;
    } else if (b.n == 0) {
        expon = -rsigma - b.sigln - fact;
        fact = -(1.0f / b.sigma + rbroad);
    } else if (b.n == 1) {
        expon = -rsigma - b.sigln - fact;
        fact = 1.0f / b.sigma;
    } else {
        expon = -2.0f * expmax;
        fact = 0.0f;
    }
    expr[(size_t)(i - 1) * nterms + m] = expon > -expmax ? fact * cr_expf(expon) : 0.0f;
}

// calc_pwrexpq element (:5127-5150)
__global__ void pwrexpq_kernel(int lpt, int nterms, const BasisTerm *__restrict__ bt, const float *__restrict__ xin,
                               float *__restrict__ expr)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lpt * nterms) return;
    int i = t / nterms + 1;
    int m = t - (i - 1) * nterms + 1;
    BasisTerm b = bt[m];
    int n = b.n;
    float an2 = (float)(n + 2);
    float an2fac = 1.0f / an2;
    float q = xin[i];
    float qsigma = q * b.sigma;
    float qs2 = qsigma * qsigma;
    float onemqs2 = 1.0f - qs2;
    float onepqs2 = 1.0f + qs2;
    float rt = sqrtf(onepqs2);
    float alpha = cr_atanf(qsigma);
    float alphan = (float)n * alpha;
    float term = qsigma > 0.0f ? 2.0f * cr_cosf(alphan) + onemqs2 * cr_sinf(alphan) / qsigma : an2;
    expr[(size_t)(i - 1) * nterms + m] = term * an2fac / dev_powi(rt, 4 + n);
}

extern "C" int epsr_calc_pwrexpq(epsr_ctx *ctx, int lpt, int nterms, float dmin, float dstep, float dmax, float sigmab,
                                 float rho, const float *q, float *expr, float rmin, int *ncfirst)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, q && expr && lpt >= 1 && nterms >= 1, "epsr_calc_pwrexpq: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    std::vector<BasisTerm> bt;
    int ncf = 0;
    (void)basis_terms(nterms, dmin, dstep, dmax, sigmab, rho, rmin, 1.0f, false, bt, ncf);
    BasisTerm *dbt = (BasisTerm *)ctx_scratch(ctx, sizeof(BasisTerm) * (nterms + 1));
    if (!dbt) return EPSR_ERR_HIP;
    HIP_CHECK(ctx, hipMemcpyAsync(dbt, bt.data(), sizeof(BasisTerm) * (nterms + 1), hipMemcpyHostToDevice, ctx->stream));
    pwrexpq_kernel<<<ceil_div((long)lpt * nterms, 256), 256, 0, ctx->stream>>>(lpt, nterms, dbt, q, expr);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ncfirst) *ncfirst = ncf;
    return EPSR_OK;
}

// expand (:4924-4939): sequential dot product clamped at +-summax after every term
__device__ __forceinline__ float dev_expand(int nterms, const float *__restrict__ row, const float *__restrict__ coeff,
                                            float summax)
{
    float sum = 0.0f;
    for (int i = 1; i <= nterms; ++i) {
        sum = sum + row[i] * coeff[i];
        if (fabsf(sum) > summax) sum = copysignf(summax, sum);
    }
    return sum;
}

// xavs1(j, ip) = expand(nterms, j, expgr, pwtold) * pcoefp(1, irefp) (:5920-5930);
// coef rows [ip*(nterms+1) + m] hold pwtold (m >= 2) with pwtold(1) = 0, scale[ip] = pcoefp(1, irefp)
__global__ void ep_gradient_values(int npt2, int nterms, int npairs, const float *__restrict__ expgr,
                                   const float *__restrict__ coef, const float *__restrict__ scale, float summax,
                                   float *__restrict__ xavs1, int ld)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * npt2) return;
    int ip = t / npt2;
    int j = t - ip * npt2 + 1;
    const float *row = expgr + (size_t)(j - 1) * nterms;   // row[m] = expgr((j-1)*nterms + m)
    float term = dev_expand(nterms, row, coef + (size_t)ip * (nterms + 1), summax);
    xavs1[(size_t)ip * ld + j] = term * scale[ip];
}

// per pair, serial: x dmax, backward trapezoid integration from the last point, x wfac
// (:5932-5958)
__global__ void ep_integrate(int npt, int npt2, int npairs, const float *__restrict__ xstar,
                             const float *__restrict__ wfac, float dmax, float *__restrict__ xavs,
                             float *__restrict__ xavs1, int ld)
{
    int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= npairs) return;
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

// xavs *= pressfac, xavtot = ecoul + xavs for j <= npt2 (:5985-5991)
__global__ void ep_scale_total(int npt2, int npairs, float pressfac, const float *__restrict__ ecoul,
                               float *__restrict__ xavs, float *__restrict__ xavtot, int ld)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs * npt2) return;
    int ip = t / npt2;
    int j = t - ip * npt2 + 1;
    size_t o = (size_t)ip * ld + j;
    xavs[o] = xavs[o] * pressfac;
    xavtot[o] = ecoul[o] + xavs[o];
}

extern "C" int epsr_generate_ep(epsr_ctx *ctx, int npt, int npt2, const float *x, const float *xstar,
                                const float *wfac, int npairs, const int *irefpcof, int nterms, float *pcoefp, int ldp,
                                const epsr_ep_params_t *p, const float *ecoul, float *xavs, float *xavs1,
                                float *xavtot, float *xavtot1, int ld, float *pressfac_out, float *energabs_out)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, x && xstar && wfac && irefpcof && pcoefp && p && ecoul && xavs && xavs1 && xavtot && xavtot1,
              "epsr_generate_ep: null argument");
    ARG_CHECK(ctx, npt >= 3 && npt2 >= 1 && npt2 <= npt - 2 && ld > npt && nterms >= 1 && ldp > nterms,
              "epsr_generate_ep: bad sizes");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const float summax = (float)std::exp(25.0);
    std::vector<BasisTerm> bt;
    int ncf = 0;
    // calc_pwrexpr_grad(npt2, nterms, pdmin, pdstep, pdmax, psigma2, rho, x, expgr, 0, ncfirst, 1, rbroad)
    ARG_CHECK(ctx, basis_terms(nterms, p->pdmin, p->pdstep, p->pdmax, p->psigma2, p->rho, 0.0f, 1.0f, true, bt, ncf),
              "epsr_generate_ep: a basis function position needs ln((n+2)!) beyond the table of pdmax/psigma2+3 "
              "terms (pdstep x ncoeffp exceeds pdmax)");
    int maxref = 0;
    for (int ip = 0; ip < npairs; ++ip) maxref = std::max(maxref, irefpcof[ip] + 1);
    std::vector<float> coef((size_t)npairs * (nterms + 1), 0.0f), scale(npairs);
    for (int ip = 0; ip < npairs; ++ip) {
        const float *pc = pcoefp + (size_t)irefpcof[ip] * ldp;
        for (int m = 2; m <= nterms; ++m) coef[(size_t)ip * (nterms + 1) + m] = pc[m];
        scale[ip] = pc[1];
    }
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    size_t oB = take(sizeof(BasisTerm) * (nterms + 1)), oE = take(sizeof(float) * ((size_t)npt2 * nterms + 1));
    size_t oC = take(sizeof(float) * coef.size()), oS = take(sizeof(float) * npairs);
    char *scr = (char *)ctx_scratch(ctx, off);
    if (!scr) return EPSR_ERR_HIP;
    BasisTerm *dbt = (BasisTerm *)(scr + oB);
    float *dexp = (float *)(scr + oE), *dcoef = (float *)(scr + oC), *dscale = (float *)(scr + oS);
    HIP_CHECK(ctx, hipMemcpyAsync(dbt, bt.data(), sizeof(BasisTerm) * (nterms + 1), hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dcoef, coef.data(), sizeof(float) * coef.size(), hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dscale, scale.data(), sizeof(float) * npairs, hipMemcpyHostToDevice, st));
    pwrexpr_grad_kernel<<<ceil_div((long)npt2 * nterms, 256), 256, 0, st>>>(npt2, nterms, dbt, x, p->rbroad, dexp);
    ep_gradient_values<<<ceil_div((long)npairs * npt2, 256), 256, 0, st>>>(npt2, nterms, npairs, dexp, dcoef, dscale,
                                                                          summax, xavs1, ld);
    ep_integrate<<<ceil_div(npairs, 64), 64, 0, st>>>(npt, npt2, npairs, xstar, wfac, p->dmax, xavs, xavs1, ld);
    // abs_energy_ep (ato_manipulation_rev4.py:4127-4157) and the pressure factor (:5966-5980)
    float energabs = 0.0f;
    for (int ip = 0; ip < maxref; ++ip) {
        float mn = 0.0f, mx = 0.0f;
        for (int i = 2; i <= nterms; ++i) {
            float term = pcoefp[(size_t)ip * ldp + i];
            mn = std::min(mn, term);
            mx = std::max(mx, term);
        }
        energabs = std::max(energabs, mx - mn);
    }
    float pressfac = 1.0f;
    if (fabsf(p->absolute_energy) == 0.0f) {
        pressfac = 0.0f;
        energabs = 0.0f;
    } else if (fabsf(energabs) > 0.0f) {
        if (fabsf(energabs) < fabsf(p->erequnit)) pressfac = fabsf(p->erequnit) / fabsf(energabs);
        else {
            pressfac = fabsf(p->absolute_energy) / fabsf(energabs);
            if (pressfac > 1.0f && p->ereqstep == 0.0f) pressfac = 1.0f;
        }
    }
    // multiply_pot_coeffs (epsr_standalone_rev3.py:6493-6510)
    for (int ip = 0; ip < npairs; ++ip) {
        float *pc = pcoefp + (size_t)irefpcof[ip] * ldp;
        for (int m = 2; m <= nterms; ++m) pc[m] = pc[m] * pressfac;
    }
    ep_scale_total<<<ceil_div((long)npairs * npt2, 256), 256, 0, st>>>(npt2, npairs, pressfac, ecoul, xavs, xavtot, ld);
    gradients_kernel<<<ceil_div((long)npairs * npt, 256), 256, 0, st>>>(npt, npairs, xstar, xavs, xavs1, xavtot,
                                                                        xavtot1, ld);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    if (pressfac_out) *pressfac_out = pressfac;
    if (energabs_out) *energabs_out = energabs;
    return EPSR_OK;
}
