// The empirical-potential fit: expand, smoop, fit_expansion (exact Metropolis replay and
// an FP64 normal-equation mode) and the data-minus-fit differences of ptgnsq.
//
// Reference: src/epsr26/programs/epsr_standalone_rev3.py:4805-4939 (fit_expansion,
// expand), :3789-4021 (ptgnsq); src/epsr26/library/auxiliary_routines.py:744-764 (smoop).
//
// Exact replay: one workgroup per chain (a dataset, or a replica of one).  The ran1
// stream is drawn on the host in chunks of ten tries and consumed through a cursor:
// each step takes nterms - ncfirst + 1 values for the trial coefficients and one more
// when the Metropolis test needs it.  The chain state lives in device memory between
// launches, so a launch runs whole tries while the stream lasts and the host refills
// it until every chain has met the reference's stopping rule.  Every float32 operation
// is the reference's, in its order: the trial coefficients, expand (a sequential dot
// product clamped after every term), term * term, and chi^2 as a sequential sum over
// the Q points formed on one lane after the squares are computed in parallel.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include <cstring>
#include <vector>

#define FIT_NTMAX 128
#define FIT_NSTEP 100
#define FIT_NTRY 150

struct FitChain {
    float stepr, acceptr, rejrat, chisqold, chisqtest;
    int icount, itry, cursor, done, started;
};

// expand (epsr_standalone_rev3.py:4924-4939): expr flat, 1-based, row j at (j-1)*nterms
__device__ __forceinline__ float dev_expand_row(int nterms, int j, const float *__restrict__ expr, const float *coeff,
                                                float summax)
{
    float sum = 0.0f;
    const float *e = expr + (size_t)(j - 1) * nterms;
    for (int i = 1; i <= nterms; ++i) {
        sum = sum + e[i] * coeff[i];
        if (fabsf(sum) > summax) sum = copysignf(summax, sum);
    }
    return sum;
}

// smoop (auxiliary_routines.py:744-764) on one lane
__device__ void dev_smoop(int n1, int n2, int nsmoo, float *diff)
{
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

// chi^2 = sum_{i=nmin..npt} (yin_i - expand_i(w))^2: squares in parallel, the sum in
// order on thread 0; every thread returns the sum
__device__ float dev_chisq(const float *__restrict__ yin, const float *__restrict__ expr, const float *w, int nmin,
                           int npt, int nterms, float summax, float *sq, float *bcast)
{
    for (int i = nmin + (int)threadIdx.x; i <= npt; i += blockDim.x) {
        float term = yin[i] - dev_expand_row(nterms, i, expr, w, summax);
        sq[i] = term * term;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float c = 0.0f;
        for (int i = nmin; i <= npt; ++i) c = c + sq[i];
        *bcast = c;
    }
    __syncthreads();
    float c = *bcast;
    __syncthreads();
    return c;
}

__device__ __forceinline__ float cr_expf(float x) { return (float)exp((double)x); }

__global__ void __launch_bounds__(256)
fit_mc_kernel(const float *__restrict__ yin_all, int ldy, const float *__restrict__ expr_all, long long lde,
              float *__restrict__ wt_all, float *__restrict__ wtold_all, int ldw, FitChain *__restrict__ chains,
              const float *__restrict__ stream_all, long long lds_stream, const int *__restrict__ navail, int nmin,
              int npt, int ncfirst, int nterms, int nfreq, float summax)
{
    extern __shared__ float fsm[];
    float *w = fsm;                       // [FIT_NTMAX + 1] trial coefficients
    float *wo = w + FIT_NTMAX + 1;        // [FIT_NTMAX + 1] accepted coefficients
    float *bc = wo + FIT_NTMAX + 1;       // [8] broadcast slots
    float *sq = bc + 8;                   // [npt + 1] squared differences
    const int c = blockIdx.x;
    const float *yin = yin_all + (size_t)c * ldy;
    const float *expr = expr_all + (size_t)c * lde;
    float *wt = wt_all + (size_t)c * ldw;
    float *wtold = wtold_all + (size_t)c * ldw;
    const float *rs = stream_all + (size_t)c * lds_stream;
    FitChain s = chains[c];
    if (s.done) return;
    const int tid = threadIdx.x;
    for (int i = tid; i <= nterms; i += blockDim.x) {
        wo[i] = wtold[i];
        w[i] = s.started ? wt[i] : wtold[i];
    }
    __syncthreads();
    const int npoints = npt - nmin + 1;
    const int nd = nterms - ncfirst + 1;
    const int avail = navail[c];
    if (!s.started) {
        // chi^2 of the starting coefficients (epsr_standalone_rev3.py:4838-4841)
        s.chisqold = dev_chisq(yin, expr, w, nmin, npt, nterms, summax, sq, bc);
        s.chisqtest = -1.0f;
        s.icount = 0;
        s.itry = 0;
        s.started = 1;
    }
    while (true) {
        if (!(s.itry < 10 || (s.itry < FIT_NTRY && s.chisqtest < 1.0e-4f))) {
            s.done = 1;
            break;
        }
        if (s.cursor + FIT_NSTEP * (nd + 1) > avail) break;   // the host refills the stream
        s.itry = s.itry + 1;
        const float accept = s.acceptr * s.chisqold / (float)npoints;
        float rr = 0.0f;
        const float chisqstart = s.chisqold / (float)npoints;
        for (int it = 1; it <= FIT_NSTEP; ++it) {
            for (int ic = ncfirst + tid; ic <= nterms; ic += blockDim.x) {
                const float r = rs[s.cursor + (ic - ncfirst)];
                if (fabsf(wo[ic]) < 1.0f) w[ic] = wo[ic] + s.stepr * (2.0f * r - 1.0f);
                else w[ic] = wo[ic] * (1.0f + s.stepr * (2.0f * r - 1.0f));
            }
            s.cursor += nd;
            __syncthreads();
            float cnew = dev_chisq(yin, expr, w, nmin, npt, nterms, summax, sq, bc);
            const float exponent = (cnew - s.chisqold) / accept;
            float test;
            if (exponent > 0.0f && exponent <= 20.0f) {
                test = cr_expf(-exponent) - rs[s.cursor];
                s.cursor += 1;
            } else if (exponent > 20.0f) {
                test = -1.0f;
            } else {
                test = 1.0f;
            }
            if (test >= 0.0f) {
                s.icount = s.icount + 1;
                if (s.icount == nfreq) {
                    s.icount = 0;
                    if (tid == 0) dev_smoop(ncfirst, nterms, 1, w);
                    __syncthreads();
                    cnew = dev_chisq(yin, expr, w, nmin, npt, nterms, summax, sq, bc);
                }
                s.chisqold = cnew;
                for (int ic = 1 + tid; ic <= nterms; ic += blockDim.x) wo[ic] = w[ic];
            } else {
                rr = rr + 1.0f;
            }
            __syncthreads();
        }
        rr = rr / (float)FIT_NSTEP;
        s.rejrat = rr;
        const float chisqfinish = s.chisqold / (float)npoints;
        if (chisqstart > 0.0f) s.chisqtest = (chisqfinish - chisqstart) / chisqstart;
        else s.chisqtest = chisqfinish - chisqstart;
        const float rejfac = rr > 0.0f ? 0.75f / rr : 2.0f;
        s.stepr = s.stepr * rejfac;
    }
    for (int i = tid; i <= nterms; i += blockDim.x) {
        wtold[i] = wo[i];
        wt[i] = s.done ? wo[i] : w[i];
    }
    if (tid == 0) chains[c] = s;
}

// start of a fit: wt = wtold, acceptr = 0.05, the minimum step (:4824-4833)
__global__ void fit_start_kernel(int nchain, const float *__restrict__ wtold_all, float *__restrict__ wt_all, int ldw,
                                 int nterms, FitChain *__restrict__ chains)
{
    const int c = blockIdx.x;
    if (c >= nchain) return;
    const float *wtold = wtold_all + (size_t)c * ldw;
    float *wt = wt_all + (size_t)c * ldw;
    for (int i = threadIdx.x; i <= nterms; i += blockDim.x) wt[i] = wtold[i];
    if (threadIdx.x == 0) {
        FitChain s = chains[c];
        s.acceptr = 0.05f;
        float wtmax = 0.0f;
        for (int it = 1; it <= nterms; ++it) wtmax = fmaxf(fabsf(wtold[it]), wtmax);
        if (s.stepr < 1.0e-6f * wtmax) s.stepr = 2.0e-6f * wtmax;
        chains[c] = s;
    }
}

// end of a fit (:4911-4919): wt = wtold, yout = expand(wt) for i = 1..npt, chi^2 over
// nmin..npt in order
__global__ void __launch_bounds__(256)
fit_finish_kernel(const float *__restrict__ yin_all, int ldy, const float *__restrict__ expr_all, long long lde,
                  float *__restrict__ yout_all, int ldo, float *__restrict__ wt_all, const float *__restrict__ wtold_all,
                  int ldw, int nmin, int npt, int nterms, float summax, int copy_old, float *__restrict__ chisq_out)
{
    extern __shared__ float fsm[];
    float *w = fsm;
    float *sq = w + FIT_NTMAX + 1;
    const int c = blockIdx.x;
    const float *yin = yin_all + (size_t)c * ldy;
    const float *expr = expr_all + (size_t)c * lde;
    float *yout = yout_all + (size_t)c * ldo;
    float *wt = wt_all + (size_t)c * ldw;
    const float *wtold = wtold_all + (size_t)c * ldw;
    for (int i = threadIdx.x; i <= nterms; i += blockDim.x) {
        float v = copy_old ? wtold[i] : wt[i];
        w[i] = v;
        wt[i] = v;
    }
    __syncthreads();
    for (int i = 1 + (int)threadIdx.x; i <= npt; i += blockDim.x) {
        float term = dev_expand_row(nterms, i, expr, w, summax);
        yout[i] = term;
        float d = yin[i] - term;
        sq[i] = d * d;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float cs = 0.0f;
        for (int i = nmin; i <= npt; ++i) cs = cs + sq[i];
        chisq_out[c] = cs;
    }
}

// FP64 least squares: the coefficients ncfirst..nterms minimising
// sum_{i=nmin..npt} (yin_i - sum_m expr_im w_m)^2 with w_1..w_{ncfirst-1} = wtold held,
// through the normal equations G w = h (Cholesky; a relative 1e-13 ridge on the
// diagonal keeps a rank-deficient basis solvable).  The expand clamp is not applied.
__global__ void __launch_bounds__(256)
fit_ne_kernel(const float *__restrict__ yin_all, int ldy, const float *__restrict__ expr_all, long long lde,
              float *__restrict__ wt_all, float *__restrict__ wtold_all, int ldw, int nmin, int npt, int ncfirst,
              int nterms, int *__restrict__ status)
{
    extern __shared__ double dsm[];
    const int nd = nterms - ncfirst + 1;
    double *G = dsm;                       // [nd * nd]
    double *h = G + (size_t)nd * nd;       // [nd]
    double *ws = h + nd;                   // [nterms + 1]
    const int c = blockIdx.x;
    const float *yin = yin_all + (size_t)c * ldy;
    const float *expr = expr_all + (size_t)c * lde;
    float *wt = wt_all + (size_t)c * ldw;
    float *wtold = wtold_all + (size_t)c * ldw;
    const int tid = threadIdx.x;
    for (int i = tid; i <= nterms; i += blockDim.x) ws[i] = (double)wtold[i];
    __syncthreads();
    // G and h: one entry per thread pass, sums over the points in FP64
    const int nent = nd * (nd + 1) / 2 + nd;
    for (int e = tid; e < nent; e += blockDim.x) {
        double acc = 0.0;
        if (e < nd * (nd + 1) / 2) {
            int a = 0, rem = e;
            while (rem >= nd - a) {
                rem -= nd - a;
                ++a;
            }
            const int b = a + rem;
            for (int i = nmin; i <= npt; ++i) {
                const float *row = expr + (size_t)(i - 1) * nterms;
                acc += (double)row[ncfirst + a] * (double)row[ncfirst + b];
            }
            G[a * nd + b] = acc;
            G[b * nd + a] = acc;
        } else {
            const int a = e - nd * (nd + 1) / 2;
            for (int i = nmin; i <= npt; ++i) {
                const float *row = expr + (size_t)(i - 1) * nterms;
                double y = (double)yin[i];
                for (int m = 1; m < ncfirst; ++m) y -= (double)row[m] * ws[m];
                acc += (double)row[ncfirst + a] * y;
            }
            h[a] = acc;
        }
    }
    __syncthreads();
    if (tid == 0) {
        double tr = 0.0;
        for (int a = 0; a < nd; ++a) tr += G[a * nd + a];
        const double ridge = 1e-13 * (tr > 0.0 ? tr / nd : 1.0);
        int ok = 1;
        // Cholesky G = L L^T in place (lower triangle)
        for (int j = 0; j < nd && ok; ++j) {
            double d = G[j * nd + j] + ridge;
            for (int k = 0; k < j; ++k) d -= G[j * nd + k] * G[j * nd + k];
            if (!(d > 0.0)) {
                ok = 0;
                break;
            }
            d = sqrt(d);
            G[j * nd + j] = d;
            for (int i = j + 1; i < nd; ++i) {
                double v = G[i * nd + j];
                for (int k = 0; k < j; ++k) v -= G[i * nd + k] * G[j * nd + k];
                G[i * nd + j] = v / d;
            }
        }
        if (ok) {
            for (int i = 0; i < nd; ++i) {   // L z = h
                double v = h[i];
                for (int k = 0; k < i; ++k) v -= G[i * nd + k] * h[k];
                h[i] = v / G[i * nd + i];
            }
            for (int i = nd - 1; i >= 0; --i) {   // L^T w = z
                double v = h[i];
                for (int k = i + 1; k < nd; ++k) v -= G[k * nd + i] * h[k];
                h[i] = v / G[i * nd + i];
            }
            for (int a = 0; a < nd; ++a) ws[ncfirst + a] = h[a];
        }
        status[c] = ok;
    }
    __syncthreads();
    for (int i = 1 + tid; i <= nterms; i += blockDim.x) {
        float v = (float)ws[i];
        wt[i] = v;
        wtold[i] = v;
    }
}

// expand(nterms, j, expr, coeff, summax) for j = 1..npt
__global__ void expand_all_kernel(int nterms, int npt, const float *__restrict__ expr, const float *__restrict__ coeff,
                                  float summax, float *__restrict__ out)
{
    __shared__ float cf[FIT_NTMAX + 1];
    for (int i = threadIdx.x; i <= nterms; i += blockDim.x) cf[i] = coeff[i];
    __syncthreads();
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j <= npt) out[j] = dev_expand_row(nterms, j, expr, cf, summax);
}

__global__ void smoop_kernel(int n1, int n2, int nsmoo, float *diff)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) dev_smoop(n1, n2, nsmoo, diff);
}

static int fit_check(epsr_ctx *ctx, int nmin, int npt, int ncfirst, int nterms)
{
    ARG_CHECK(ctx, nterms >= 1 && nterms <= FIT_NTMAX, "fit: nterms %d outside 1..%d", nterms, FIT_NTMAX);
    ARG_CHECK(ctx, ncfirst >= 1 && ncfirst <= nterms, "fit: ncfirst %d outside 1..nterms", ncfirst);
    ARG_CHECK(ctx, nmin >= 1 && npt >= nmin, "fit: bad point range %d..%d", nmin, npt);
    return EPSR_OK;
}

extern "C" int epsr_expand_all(epsr_ctx *ctx, int nterms, int npt, const float *expr, const float *coeff, float summax,
                               float *out)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, expr && coeff && out && nterms >= 1 && nterms <= FIT_NTMAX && npt >= 0,
              "epsr_expand_all: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (npt > 0) expand_all_kernel<<<ceil_div(npt, 256), 256, 0, ctx->stream>>>(nterms, npt, expr, coeff, summax, out);
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}

extern "C" int epsr_smoop(epsr_ctx *ctx, int nqmin, int nqmax, int nsmoo, float *diff)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, diff && nqmin >= 0, "epsr_smoop: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    smoop_kernel<<<1, 64, 0, ctx->stream>>>(nqmin, nqmax, nsmoo, diff);
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}

extern "C" int epsr_fit_expansion(epsr_ctx *ctx, int nchain, const float *yin, int ldy, const float *expr, long long lde,
                                  float *yout, int ldo, float *wt, float *wtold, int ldw, float *stepr, float *acceptr,
                                  float *rejrat, float *chisqnew, int nmin, int npt, int ncfirst, int nterms, int nits,
                                  int nfreq, int32_t *ran1_state, int mode, int *ntries)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, nchain >= 1 && yin && expr && yout && wt && wtold && stepr && acceptr && rejrat && chisqnew,
              "epsr_fit_expansion: null argument");
    int rc = fit_check(ctx, nmin, npt, ncfirst, nterms);
    if (rc) return rc;
    ARG_CHECK(ctx, ldy >= npt + 1 && ldo >= npt + 1 && ldw >= nterms + 1 && (lde == 0 || lde >= (long long)npt * nterms + 1),
              "epsr_fit_expansion: leading dimension too small");
    ARG_CHECK(ctx, mode == EPSR_FIT_MC || mode == EPSR_FIT_NORMAL, "epsr_fit_expansion: unknown mode %d", mode);
    ARG_CHECK(ctx, mode != EPSR_FIT_MC || ran1_state, "epsr_fit_expansion: the replay needs the ran1 state");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const float summax = (float)std::exp(25.0);   // expf(25), correctly rounded (:4722, :4820)
    const int nd = nterms - ncfirst + 1;
    StreamTemp tmp(ctx);
    FitChain *dch = (FitChain *)tmp.alloc(sizeof(FitChain) * nchain);
    float *dchi = (float *)tmp.alloc(sizeof(float) * nchain);
    int *dstat = (int *)tmp.alloc(sizeof(int) * nchain);
    if (!dch || !dchi || !dstat) return EPSR_ERR_HIP;
    std::vector<FitChain> hch(nchain);
    for (int c = 0; c < nchain; ++c) {
        FitChain s{};
        s.stepr = stepr[c];
        s.acceptr = acceptr[c];
        s.rejrat = rejrat[c];
        s.chisqtest = -1.0f;
        hch[c] = s;
    }
    HIP_CHECK(ctx, hipMemcpyAsync(dch, hch.data(), sizeof(FitChain) * nchain, hipMemcpyHostToDevice, st));
    fit_start_kernel<<<nchain, 64, 0, st>>>(nchain, wtold, wt, ldw, nterms, dch);
    std::vector<int> tries(nchain, 0);
    if (nits > 0 && mode == EPSR_FIT_MC) {
        // the ran1 stream, drawn in chunks of ten tries per chain
        const long long per_try = (long long)FIT_NSTEP * (nd + 1);
        const long long chunk = 10 * per_try;
        const long long cap = FIT_NTRY * per_try + 8;
        float *dstream = (float *)tmp.alloc(sizeof(float) * (size_t)(cap * nchain));
        int *davail = (int *)tmp.alloc(sizeof(int) * nchain);
        if (!dstream || !davail) return EPSR_ERR_HIP;
        std::vector<Ran1> gen(nchain);
        std::vector<std::vector<std::pair<long long, Ran1>>> snap(nchain);   // (offset, state) per chunk
        std::vector<long long> have(nchain, 0);
        std::vector<float> buf((size_t)chunk);
        for (int c = 0; c < nchain; ++c) ran1_load(gen[c], ran1_state + 35 * c);
        const size_t lds = sizeof(float) * (2 * (FIT_NTMAX + 1) + 8 + (size_t)npt + 1);
        ARG_CHECK(ctx, lds <= 160 * 1024, "epsr_fit_expansion: %d points exceed LDS", npt);
        std::vector<int> avail(nchain);
        while (true) {
            bool any = false;
            for (int c = 0; c < nchain; ++c) {
                if (hch[c].done) continue;
                if (have[c] - hch[c].cursor >= per_try || have[c] >= cap - 8) continue;
                long long n = std::min(chunk, cap - 8 - have[c]);
                snap[c].push_back({have[c], gen[c]});
                ran1_fill(gen[c], buf.data(), n);
                HIP_CHECK(ctx, hipMemcpyAsync(dstream + (size_t)c * cap + have[c], buf.data(), sizeof(float) * n,
                                              hipMemcpyHostToDevice, st));
                HIP_CHECK(ctx, hipStreamSynchronize(st));   // buf is reused
                have[c] += n;
                any = true;
            }
            (void)any;
            int progress_before = 0;
            for (int c = 0; c < nchain; ++c) progress_before += hch[c].itry + hch[c].done;
            for (int c = 0; c < nchain; ++c) avail[c] = (int)have[c];
            HIP_CHECK(ctx, hipMemcpyAsync(davail, avail.data(), sizeof(int) * nchain, hipMemcpyHostToDevice, st));
            fit_mc_kernel<<<nchain, 256, lds, st>>>(yin, ldy, expr, lde, wt, wtold, ldw, dch, dstream, cap, davail, nmin,
                                                    npt, ncfirst, nterms, nfreq, summax);
            HIP_CHECK(ctx, hipMemcpyAsync(hch.data(), dch, sizeof(FitChain) * nchain, hipMemcpyDeviceToHost, st));
            HIP_CHECK(ctx, hipStreamSynchronize(st));
            bool all = true;
            int progress_after = 0;
            for (int c = 0; c < nchain; ++c) {
                all = all && hch[c].done;
                progress_after += hch[c].itry + hch[c].done;
            }
            if (all) break;
            ARG_CHECK(ctx, progress_after > progress_before, "epsr_fit_expansion: no progress with a full stream");
        }
        // the ran1 state after the values each chain consumed
        for (int c = 0; c < nchain; ++c) {
            const long long cur = hch[c].cursor;
            size_t k = 0;
            while (k + 1 < snap[c].size() && snap[c][k + 1].first <= cur) ++k;
            Ran1 r = snap[c].empty() ? gen[c] : snap[c][k].second;
            long long off = snap[c].empty() ? cur : snap[c][k].first;
            for (long long i = off; i < cur; ++i) (void)ran1_next(r);
            ran1_store(r, ran1_state + 35 * c);
            tries[c] = hch[c].itry;
        }
    } else if (nits > 0) {
        ARG_CHECK(ctx, nd <= 96, "epsr_fit_expansion: the normal-equation mode takes at most 96 free coefficients");
        const size_t lds = sizeof(double) * ((size_t)nd * nd + nd + nterms + 1);
        fit_ne_kernel<<<nchain, 256, lds, st>>>(yin, ldy, expr, lde, wt, wtold, ldw, nmin, npt, ncfirst, nterms, dstat);
        std::vector<int> ok(nchain);
        HIP_CHECK(ctx, hipMemcpyAsync(ok.data(), dstat, sizeof(int) * nchain, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipMemcpyAsync(hch.data(), dch, sizeof(FitChain) * nchain, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
        for (int c = 0; c < nchain; ++c)
            ARG_CHECK(ctx, ok[c], "epsr_fit_expansion: chain %d: the normal matrix is not positive definite", c);
    } else {
        HIP_CHECK(ctx, hipMemcpyAsync(hch.data(), dch, sizeof(FitChain) * nchain, hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
    }
    const size_t lds_fin = sizeof(float) * (FIT_NTMAX + 2 + (size_t)npt + 1);
    ARG_CHECK(ctx, lds_fin <= 160 * 1024, "epsr_fit_expansion: %d points exceed LDS", npt);
    fit_finish_kernel<<<nchain, 256, lds_fin, st>>>(yin, ldy, expr, lde, yout, ldo, wt, wtold, ldw, nmin, npt, nterms,
                                                    summax, nits > 0 ? 1 : 0, dchi);
    std::vector<float> chi(nchain);
    HIP_CHECK(ctx, hipMemcpyAsync(chi.data(), dchi, sizeof(float) * nchain, hipMemcpyDeviceToHost, st));
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));
    for (int c = 0; c < nchain; ++c) {
        stepr[c] = hch[c].stepr;
        acceptr[c] = hch[c].acceptr;
        rejrat[c] = mode == EPSR_FIT_MC ? hch[c].rejrat : rejrat[c];
        chisqnew[c] = chi[c];
        if (ntries) ntries[c] = tries[c];
    }
    return EPSR_OK;
}

// ------------------------------------------------------------------ ptgnsq
// qtemp(i, id): data minus the cwt-weighted sum of the intermolecular partials, with the
// extrapolation to Q = 0 below qmin (epsr_standalone_rev3.py:3827-3867); one thread per
// (Q point, dataset)
// sum = f32(sum + w * s): the reference's weights are float64 arrays (weights.cwt, and
// x_rayf0.weightx for X-ray data), so each term of a weighted sum is added in double and
// the sum rounded to float
__device__ __forceinline__ float wsum_step(float sum, double w, float s) { return (float)((double)sum + w * (double)s); }

__global__ void ptgnsq_diff_kernel(int nfile, int nq, int npairs, const float *__restrict__ q,
                                   const float *__restrict__ data, int ldq, const float *__restrict__ qwind,
                                   const float *__restrict__ cwt, const double *__restrict__ weightx,
                                   const int *__restrict__ xfile, const float *__restrict__ szeros,
                                   const int *__restrict__ nqmin, const int *__restrict__ nqmax, float qminptemp,
                                   float *__restrict__ qtemp)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfile * nq) return;
    const int id = t / nq, i = t - id * nq + 1;
    const float *dat = data + (size_t)id * ldq;
    const int nqmn = nqmin[id];
    const int ix = xfile ? xfile[id] : 0;
    float grads0 = 0.0f, conss0 = 0.0f;
    if (szeros[id] != 0.0f && nqmn > 1) {
        grads0 = (dat[nqmn] - szeros[id]) / (q[nqmn] - q[1]);
        conss0 = szeros[id];
    }
    float sum = 0.0f;
    if (ix > 0) {
        const double *w = weightx + ((size_t)(ix - 1) * ldq + i) * npairs;
        for (int ic = 0; ic < npairs; ++ic) sum = wsum_step(sum, w[ic], qwind[(size_t)ic * ldq + i]);
    } else {
        for (int ic = 0; ic < npairs; ++ic) sum = wsum_step(sum, (double)cwt[id * npairs + ic], qwind[(size_t)ic * ldq + i]);
    }
    float v;
    if (q[i] < fmaxf(qminptemp, q[nqmn])) {
        v = szeros[id] != 0.0f ? conss0 + grads0 * (q[i] - q[1]) - sum : 0.0f;
    } else if (i >= nqmn && i <= nqmax[id]) {
        v = dat[i] - sum;
    } else {
        v = 0.0f;
    }
    qtemp[(size_t)id * ldq + i] = v;
}

extern "C" int epsr_ptgnsq_diff(epsr_ctx *ctx, int nfile, int nq, int npairs, const float *q, const float *data,
                                int ldq, const float *qwind, const float *cwt, const float *szeros, const int *nqmin,
                                const int *nqmax, float qminptemp, float *qtemp, const double *weightx,
                                const int *xfile, int nxfile)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, nfile >= 1 && nq >= 1 && npairs >= 1 && ldq >= nq + 1 && q && data && qwind && cwt && szeros &&
                       nqmin && nqmax && qtemp,
              "epsr_ptgnsq_diff: bad argument");
    for (int id = 0; id < nfile; ++id) {
        ARG_CHECK(ctx, nqmin[id] >= 1 && nqmin[id] <= nq && nqmax[id] <= nq,
                  "epsr_ptgnsq_diff: dataset %d: Q range %d..%d outside 1..%d", id, nqmin[id], nqmax[id], nq);
        ARG_CHECK(ctx, !xfile || (xfile[id] >= 0 && xfile[id] <= nxfile && (xfile[id] == 0 || weightx)),
                  "epsr_ptgnsq_diff: dataset %d: X-ray file %d outside 0..%d (or no weightx)", id,
                  xfile ? xfile[id] : 0, nxfile);
    }
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    StreamTemp tmp(ctx);
    float *dcwt = (float *)tmp.alloc(sizeof(float) * nfile * npairs);
    float *dsz = (float *)tmp.alloc(sizeof(float) * nfile);
    int *dmn = (int *)tmp.alloc(sizeof(int) * nfile), *dmx = (int *)tmp.alloc(sizeof(int) * nfile);
    int *dxf = xfile ? (int *)tmp.alloc(sizeof(int) * nfile) : nullptr;
    if (!dcwt || !dsz || !dmn || !dmx || (xfile && !dxf)) return EPSR_ERR_HIP;
    HIP_CHECK(ctx, hipMemcpyAsync(dcwt, cwt, sizeof(float) * nfile * npairs, hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dsz, szeros, sizeof(float) * nfile, hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dmn, nqmin, sizeof(int) * nfile, hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dmx, nqmax, sizeof(int) * nfile, hipMemcpyHostToDevice, st));
    if (dxf) HIP_CHECK(ctx, hipMemcpyAsync(dxf, xfile, sizeof(int) * nfile, hipMemcpyHostToDevice, st));
    ptgnsq_diff_kernel<<<ceil_div(nfile * nq, 256), 256, 0, st>>>(nfile, nq, npairs, q, data, ldq, qwind, dcwt, weightx,
                                                                   dxf, dsz, dmn, dmx, qminptemp, qtemp);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));   // the host arguments are released on return
    return EPSR_OK;
}

// ------------------------------------------------------------------ data minus simulation
// The block of the epsr program between the running averages and ptgnsq
// (epsr_standalone_rev3.py:1703-1830), one workgroup, the datasets in order:
//  1. datat = datai - sum_ip w soqintra (w: ciwt, or weightx for X-ray data) on the data's
//     Q range, szerostemp = szeros - the same sum at Q index 1 (when szeros is non-zero);
//  2. per dataset dts = datat - sum_ic w soqinter on the fitted range (plus the Porod term
//     porodamp (porodq / q)^porodpwr when porodamp > 0), extrapolated below it from
//     szerostemp when szeros is non-zero, zero elsewhere;
//  3. with nsmoop > 0, ftback (:5330-5389): the part of dts from r < rshmin (a sine
//     transform to g(r) on r < rshmin and back), subtracted from datat on the fitted range.
// grt is the program's one r-indexed work array: entries a call does not write keep what
// earlier calls left.  sinf and powf are the correctly rounded values (float of the
// double functions).
__device__ __forceinline__ float cr_sinf(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cr_powf(float x, float y) { return (float)pow((double)x, (double)y); }

__device__ __forceinline__ float dms_pair_sum(const epsr_dms_desc_t &d, const float *__restrict__ ciw,
                                              const float *__restrict__ cw, const int *__restrict__ xf, int id,
                                              int iq, const float *__restrict__ soq, bool intra)
{
    const int ix = xf[id];
    float sum = 0.0f;
    if (ix > 0) {
        const double *w = d.weightx + ((size_t)(ix - 1) * d.ldq + iq) * d.npairs;
        for (int ip = 0; ip < d.npairs; ++ip) sum = wsum_step(sum, w[ip], soq[(size_t)ip * d.ldq + iq]);
    } else {
        const float *w = (intra ? ciw : cw) + (size_t)id * d.npairs;
        for (int ip = 0; ip < d.npairs; ++ip) sum = wsum_step(sum, (double)w[ip], soq[(size_t)ip * d.ldq + iq]);
    }
    return sum;
}

// per-dataset arrays staged on the device: [0] xfile, [1..4] nqmind, nqmaxd, nqmin, nqmax
// (int), then szeros, porodamp, rshmin (float), then ciwt, cwt [nfile*npairs]
__global__ void __launch_bounds__(1024)
data_minus_sim_kernel(epsr_dms_desc_t d, const int *__restrict__ iv, const float *__restrict__ fv,
                      float *__restrict__ datat, float *__restrict__ szerostemp, float *__restrict__ dts,
                      float *__restrict__ grt)
{
    const int nf = d.nfile, nq = d.nq, ldq = d.ldq;
    const int *xf = iv, *nqmind = iv + nf, *nqmaxd = iv + 2 * nf, *nqmin = iv + 3 * nf, *nqmax = iv + 4 * nf;
    const float *szeros = fv, *porodamp = fv + nf, *rshmin = fv + 2 * nf;
    const float *ciw = fv + 3 * nf, *cw = ciw + (size_t)nf * d.npairs;
    const float *q = d.q;
    const int tid = threadIdx.x, nt = blockDim.x;
    __shared__ int s_jfirst;
    // 1. the intramolecular terms out of the data
    for (int t = tid; t < nf * nq; t += nt) {
        const int id = t / nq, iq = t - id * nq + 1;
        float v = 0.0f;
        if (iq >= nqmind[id] && iq <= nqmaxd[id])
            v = d.datai[(size_t)id * ldq + iq] - dms_pair_sum(d, ciw, cw, xf, id, iq, d.soqintra, true);
        datat[(size_t)id * ldq + iq] = v;
    }
    for (int id = tid; id < nf; id += nt)
        szerostemp[id] = szeros[id] != 0.0f ? szeros[id] - dms_pair_sum(d, ciw, cw, xf, id, 1, d.soqintra, true) : 0.0f;
    __syncthreads();
    const float pi = 3.141592653f;
    for (int id = 0; id < nf; ++id) {
        float *dt = datat + (size_t)id * ldq;
        const int nmn = nqmin[id], nmx = nqmax[id];
        float grads0, conss0;
        if (szeros[id] != 0.0f && nmn > 1) {
            grads0 = (dt[nmn] - szerostemp[id]) / (q[nmn] - q[1]);
            conss0 = szerostemp[id];
        } else {
            grads0 = 0.0f;
            conss0 = dt[nmn];
        }
        // 2. data - simulation
        for (int i = 1 + tid; i <= nq; i += nt) {
            const float sum = dms_pair_sum(d, ciw, cw, xf, id, i, d.soqinter, false);
            float v;
            if (i < nmn) {
                v = szeros[id] != 0.0f ? conss0 + grads0 * (q[i] - q[1]) - sum : 0.0f;
            } else if (i <= nmx) {
                v = dt[i] - sum;
                if (porodamp[id] > 0.0f) v = v + porodamp[id] * cr_powf(d.porodq / q[i], d.porodpwr);
            } else {
                v = 0.0f;
            }
            // ftback starts from sofq * q on the fitted range
            if (d.nsmoop > 0 && i >= nmn && i <= nmx) v = v * q[i];
            dts[i] = v;
        }
        if (d.nsmoop <= 0) {
            __syncthreads();
            continue;
        }
        // 3. ftback: g(r) on r < rshmin from dts q, then the S(Q) of that truncated g(r)
        if (tid == 0) {
            int j = 2;
            while (j <= d.npt1 && d.x[j] < rshmin[id]) ++j;
            s_jfirst = j;
        }
        __syncthreads();
        const int jfirst = s_jfirst;
        const int jsave = jfirst > 2 ? jfirst - 1 : 2;
        const float qstep = q[4] - q[3];
        const float sfac = 0.5f * qstep / pi / pi / d.rho;
        for (int j = 2 + tid; j < jfirst; j += nt) {
            const float rr = d.x[j];
            float sum = 0.0f;
            for (int i = nmn; i <= nmx; ++i) sum = sum + dts[i] * cr_sinf(q[i] * rr);
            grt[j] = sfac * sum;
        }
        __syncthreads();
        const float gfac = 4.0f * pi * d.rho;
        for (int i = nmn + tid; i <= nmx; i += nt) {
            const float qq = q[i];
            float r1 = 0.0f, sum = 0.0f;
            if (qq > 0.0f) {
                for (int j = 2; j <= jsave; ++j) {
                    const float rr = d.x[j];
                    const float r2 = 2.0f * rr - r1;
                    const float rstep = r2 - r1;
                    sum = sum + grt[j] * cr_sinf(qq * rr) * rstep;
                    r1 = r2;
                }
                dt[i] = dt[i] - gfac * sum / qq;
            } else {
                for (int j = 2; j <= jsave; ++j) {
                    const float rr = d.x[j];
                    const float r2 = 2.0f * rr - r1;
                    const float rstep = r2 - r1;
                    sum = sum + grt[j] * rr * rstep;
                    r1 = r2;
                }
                dt[i] = dt[i] - gfac * sum;
            }
        }
        __syncthreads();
    }
}

extern "C" int epsr_data_minus_sim(epsr_ctx *ctx, const epsr_dms_desc_t *d, float *datat, float *szerostemp,
                                   float *dts, float *grt)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, d && datat && szerostemp && dts && grt, "epsr_data_minus_sim: null argument");
    ARG_CHECK(ctx, d->nfile >= 1 && d->nq >= 4 && d->npairs >= 1 && d->ldq >= d->nq + 1 && d->q && d->datai &&
                       d->soqintra && d->soqinter && d->ciwt && d->cwt && d->xfile && d->nqmind && d->nqmaxd &&
                       d->nqmin && d->nqmax && d->szeros && d->porodamp && d->rshmin && d->x && d->npt1 >= 2,
              "epsr_data_minus_sim: bad descriptor");
    const int nf = d->nfile;
    for (int id = 0; id < nf; ++id) {
        ARG_CHECK(ctx, d->nqmin[id] >= 1 && d->nqmin[id] <= d->nqmax[id] && d->nqmax[id] <= d->nq,
                  "epsr_data_minus_sim: dataset %d: fitted Q range %d..%d outside 1..%d", id, d->nqmin[id],
                  d->nqmax[id], d->nq);
        ARG_CHECK(ctx, d->xfile[id] >= 0 && d->xfile[id] <= d->nxfile && (d->xfile[id] == 0 || d->weightx),
                  "epsr_data_minus_sim: dataset %d: X-ray file %d outside 0..%d (or no weightx)", id, d->xfile[id],
                  d->nxfile);
    }
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    StreamTemp tmp(ctx);
    const size_t ni = 5 * (size_t)nf, nfl = 3 * (size_t)nf + 2 * (size_t)nf * d->npairs;
    int *div = (int *)tmp.alloc(sizeof(int) * ni);
    float *dfv = (float *)tmp.alloc(sizeof(float) * nfl);
    if (!div || !dfv) return EPSR_ERR_HIP;
    std::vector<int> hi(ni);
    std::vector<float> hf(nfl);
    for (int id = 0; id < nf; ++id) {
        hi[id] = d->xfile[id];
        hi[nf + id] = d->nqmind[id];
        hi[2 * nf + id] = d->nqmaxd[id];
        hi[3 * nf + id] = d->nqmin[id];
        hi[4 * nf + id] = d->nqmax[id];
        hf[id] = d->szeros[id];
        hf[nf + id] = d->porodamp[id];
        hf[2 * nf + id] = d->rshmin[id];
    }
    std::memcpy(hf.data() + 3 * nf, d->ciwt, sizeof(float) * nf * d->npairs);
    std::memcpy(hf.data() + 3 * nf + (size_t)nf * d->npairs, d->cwt, sizeof(float) * nf * d->npairs);
    HIP_CHECK(ctx, hipMemcpyAsync(div, hi.data(), sizeof(int) * ni, hipMemcpyHostToDevice, st));
    HIP_CHECK(ctx, hipMemcpyAsync(dfv, hf.data(), sizeof(float) * nfl, hipMemcpyHostToDevice, st));
    data_minus_sim_kernel<<<1, 1024, 0, st>>>(*d, div, dfv, datat, szerostemp, dts, grt);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(st));   // the staged host arrays are released on return
    return EPSR_OK;
}

// rfactorep of one dataset (epsr_standalone_rev3.py:3977-3991): sequential on one lane
__global__ void ptgnsq_rfactor_kernel(int i0, int i1, const float *__restrict__ fit, const float *__restrict__ qt,
                                      float *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    float r = 0.0f;
    for (int i = i0; i <= i1; ++i) {
        if (fabsf(qt[i]) > 0.0f) r = r + fabsf(fit[i] + qt[i]) / fabsf(qt[i]);
        else r = r + fabsf(fit[i] + qt[i]);
    }
    *out = r / (float)(i1 - i0 + 1);
}

extern "C" int epsr_ptgnsq_rfactor(epsr_ctx *ctx, int i0, int i1, const float *fit, const float *qtemp,
                                   float *rfactor)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, fit && qtemp && rfactor && i0 >= 1 && i1 >= i0, "epsr_ptgnsq_rfactor: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    StreamTemp tmp(ctx);
    float *d = (float *)tmp.alloc(sizeof(float));
    if (!d) return EPSR_ERR_HIP;
    ptgnsq_rfactor_kernel<<<1, 64, 0, ctx->stream>>>(i0, i1, fit, qtemp, d);
    HIP_CHECK(ctx, hipMemcpyAsync(rfactor, d, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return EPSR_OK;
}

// the coefficient update of ptgnsq (epsr_standalone_rev3.py:3962-3975), host arrays:
// pcoefd[id*ldp + ic] = pwtold[ic] and
// pcoefp[irefp*ldp + ic] += ((pfact * efacp[irefp]) * cwtpot[ip*ldc + id]) * pwtold[ic]
// for ic = 2..nterms, every pair ip (irefp = irefpcof[ip]), all in float32
extern "C" int epsr_ptgnsq_update(epsr_ctx *ctx, int id, int nterms, int npairs, const float *pwtold,
                                  const int *irefpcof, const float *efacp, const float *cwtpot, int ldc, float pfact,
                                  float *pcoefd, float *pcoefp, int ldp)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, pwtold && irefpcof && efacp && cwtpot && pcoefd && pcoefp && nterms >= 1 && ldp >= nterms + 1,
              "epsr_ptgnsq_update: bad argument");
    for (int ic = 2; ic <= nterms; ++ic) {
        pcoefd[(size_t)id * ldp + ic] = pwtold[ic];
        for (int ip = 0; ip < npairs; ++ip) {
            const int irefp = irefpcof[ip];
            const float a = pfact * efacp[irefp];
            const float b = a * cwtpot[(size_t)ip * ldc + id];
            const float cc = b * pwtold[ic];
            pcoefp[(size_t)irefp * ldp + ic] = pcoefp[(size_t)irefp * ldp + ic] + cc;
        }
    }
    return EPSR_OK;
}
