// invert_cwt: the Monte Carlo "smart transpose" of the weights matrix
// (src/epsr26/programs/epsr_standalone_rev3.py:5423-5672).  A serial Metropolis walk over
// a matrix of at most (npairs + nfile)^2 <= 51^2 elements, driven by ran1: host code, run
// where the caller's ran1 state lives; there is no parallel work in it for a GPU.
#include "common.h"

#include <cmath>
#include <vector>

namespace {

// the cross array carrayold (nnew x nnew) from cwt and cwtt and its distance from the unit
// matrix (epsr_standalone_rev3.py:5484-5501 and :5617-5634)
double cross_total(int nfile, int npairs, const std::vector<double> &cwt, const std::vector<double> &cwtt,
                   std::vector<double> &co, double feedback1d)
{
    const int nnew = npairs + nfile;
    double total = 0.0;
    for (int i1 = 0; i1 < nnew; ++i1)
        for (int i2 = 0; i2 < nnew; ++i2) {
            double sum;
            if (i1 < nfile) {
                sum = 0.0;
                for (int j = 0; j < npairs; ++j) sum = sum + cwt[(size_t)i1 * npairs + j] * cwtt[(size_t)j * nnew + i2];
            } else {
                sum = feedback1d * cwtt[(size_t)(i1 - nfile) * nnew + i2];
            }
            co[(size_t)i1 * nnew + i2] = sum;
            if (i1 == i2) {
                const double t = sum - 1.0;
                total = total + t * t;
            } else {
                total = total + sum * sum;
            }
        }
    return total;
}

}  // namespace

extern "C" int epsr_invert_cwt(epsr_ctx *ctx, double *cwt_io, int ldc, double *cwtt_io, int ldt, double *carray_out,
                               double *carrayold_out, int ldo, float feedback, const float *efacm, int nfile,
                               int npairs, int32_t *ran1_state, int *itimes_out)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, cwt_io && cwtt_io && efacm && ran1_state && nfile > 0 && npairs > 0, "epsr_invert_cwt: bad argument");
    const int nnew = npairs + nfile;
    ARG_CHECK(ctx, ldc >= npairs && ldt >= nnew && (!carrayold_out || ldo >= nnew),
              "epsr_invert_cwt: leading dimensions too small");
    Ran1 rng;
    ran1_load(rng, ran1_state);
    std::vector<double> cwt((size_t)nfile * npairs), cwtt((size_t)npairs * nnew), co((size_t)nnew * nnew),
        carray(nnew, 0.0);
    for (int i = 0; i < nfile; ++i)
        for (int j = 0; j < npairs; ++j) cwt[(size_t)i * npairs + j] = cwt_io[(size_t)i * ldc + j];
    for (int j = 0; j < npairs; ++j)
        for (int i = 0; i < nnew; ++i) cwtt[(size_t)j * nnew + i] = cwtt_io[(size_t)j * ldt + i];
    const int niter = 50000;
    const float feedback1 = 1.0f - feedback;
    const double feedback1d = (double)feedback1;
    for (int i = 0; i < nfile; ++i)
        for (int j = 0; j < npairs; ++j)
            cwt[(size_t)i * npairs + j] = cwt[(size_t)i * npairs + j] * (double)feedback * (double)efacm[i];
    const float totaltest = 1.0e-15f;
    double total = cross_total(nfile, npairs, cwt, cwtt, co, feedback1d);
    double totalold = total;
    float amp = (float)((double)0.01f * total);
    double temp = (totalold - (double)nfile) / (double)(nnew * nnew);
    int itimes = 1, iter = 0, nreject = 0;
    while ((totalold - (double)nfile) >= (double)totaltest && itimes <= 1000) {
        for (int i2 = 0; i2 < nnew; ++i2) {
            for (int j = 0; j < npairs; ++j) {
                iter = iter + 1;
                const double add = (double)((2.0f * ran1_next(rng) - 1.0f) * amp);
                for (int i1 = 0; i1 < nfile; ++i1)
                    carray[i1] = co[(size_t)i1 * nnew + i2] + cwt[(size_t)i1 * npairs + j] * add;
                const int ip = j + nfile;
                carray[ip] = co[(size_t)ip * nnew + i2] + feedback1d * add;
                double tdif = 0.0;
                auto term = [&](int i1) {
                    const double sum = carray[i1], sumold = co[(size_t)i1 * nnew + i2];
                    if (i1 == i2) {
                        const double t = sum - 1.0, told = sumold - 1.0;
                        tdif = tdif + t * t - told * told;
                    } else {
                        tdif = tdif + sum * sum - sumold * sumold;
                    }
                };
                for (int i1 = 0; i1 < nfile; ++i1) term(i1);
                term(ip);
                tdif = tdif / temp;
                bool accept;
                if (tdif <= 0.0) accept = true;
                else if (tdif > 20.0) accept = false;
                else accept = (double)ran1_next(rng) <= std::exp(-tdif);
                if (accept) {
                    total = total + tdif * temp;
                    totalold = total;
                    cwtt[(size_t)j * nnew + i2] = cwtt[(size_t)j * nnew + i2] + add;
                    for (int i1 = 0; i1 < nfile; ++i1) co[(size_t)i1 * nnew + i2] = carray[i1];
                    co[(size_t)ip * nnew + i2] = carray[ip];
                } else {
                    nreject = nreject + 1;
                }
            }
        }
        if (iter >= niter) {
            float fac = nreject > 0 ? 0.5f * (float)(iter - nreject) / (float)nreject : 4.0f;
            if (fac < 0.95f) fac = 0.95f;
            if (fac > 1.05f) fac = 1.05f;
            amp = amp * fac;
            itimes = itimes + 1;
            iter = 0;
            nreject = 0;
            total = cross_total(nfile, npairs, cwt, cwtt, co, feedback1d);
            totalold = total;
            temp = (totalold - (double)nfile) / (double)(nnew * nnew);
        }
    }
    // the cross array of the result, then the transpose scaled by the feedback factors
    // (epsr_standalone_rev3.py:5643-5664)
    for (int j1 = 0; j1 < npairs; ++j1) {
        for (int j2 = 0; j2 < npairs; ++j2) {
            double sum = 0.0;
            for (int i = 0; i < nfile; ++i) sum = sum + cwtt[(size_t)j1 * nnew + i] * cwt[(size_t)i * npairs + j2];
            carray[j2] = sum + cwtt[(size_t)j1 * nnew + j2 + nfile] * feedback1d;
        }
        for (int i = 0; i < nfile; ++i)
            cwtt[(size_t)j1 * nnew + i] = cwtt[(size_t)j1 * nnew + i] * (double)feedback * (double)efacm[i];
        for (int i = nfile; i < nnew; ++i) cwtt[(size_t)j1 * nnew + i] = cwtt[(size_t)j1 * nnew + i] * feedback1d;
    }
    for (int i = 0; i < nfile; ++i)
        for (int j = 0; j < npairs; ++j) cwt_io[(size_t)i * ldc + j] = cwt[(size_t)i * npairs + j];
    for (int j = 0; j < npairs; ++j)
        for (int i = 0; i < nnew; ++i) cwtt_io[(size_t)j * ldt + i] = cwtt[(size_t)j * nnew + i];
    if (carray_out)
        for (int i = 0; i < nnew; ++i) carray_out[i] = carray[i];
    if (carrayold_out)
        for (int i1 = 0; i1 < nnew; ++i1)
            for (int i2 = 0; i2 < nnew; ++i2) carrayold_out[(size_t)i1 * ldo + i2] = co[(size_t)i1 * nnew + i2];
    ran1_store(rng, ran1_state);
    if (itimes_out) *itimes_out = itimes;
    return EPSR_OK;
}
