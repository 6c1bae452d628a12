/*
 * ora_weights.c -- oracle: invert_cwt, the Monte Carlo "smart transpose" of the
 * weights matrix.  TEST INFRASTRUCTURE ONLY (see epsr_oracle.h).
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

/* epsr_standalone_rev3.py:5423-5672, 0-based: cwt[i*npairs + j] (file i, pair j),
 * cwtt[j*nnew + i] (pair j, column i < nnew = npairs + nfile), carrayold[i1*nnew + i2];
 * cwt and cwtt are updated in place as the reference updates them. */
static double ora_cross(int nfile, int npairs, const double *cwt, const double *cwtt, double *co, double fb1)
{
    int nnew = npairs + nfile;
    double total = 0.0;
    for (int i1 = 0; i1 < nnew; ++i1)
        for (int i2 = 0; i2 < nnew; ++i2) {
            double sum = 0.0;
            if (i1 < nfile)
                for (int j = 0; j < npairs; ++j) sum = sum + cwt[i1 * npairs + j] * cwtt[j * nnew + i2];
            else
                sum = fb1 * cwtt[(i1 - nfile) * nnew + i2];
            co[i1 * nnew + i2] = sum;
            total = i1 == i2 ? total + (sum - 1.0) * (sum - 1.0) : total + sum * sum;
        }
    return total;
}

int ora_invert_cwt(double *cwt, double *cwtt, double *carray, double *carrayold, float feedback, const float *efacm,
                   int nfile, int npairs, ora_ran1_t *rng)
{
    const int nnew = npairs + nfile, niter = 50000;
    const double fb1 = (double)(1.0f - feedback);
    for (int i = 0; i < nfile; ++i)
        for (int j = 0; j < npairs; ++j) cwt[i * npairs + j] = cwt[i * npairs + j] * (double)feedback * (double)efacm[i];
    double total = ora_cross(nfile, npairs, cwt, cwtt, carrayold, fb1), totalold = total;
    float amp = (float)((double)0.01f * total);
    double temp = (totalold - (double)nfile) / (double)(nnew * nnew);
    int itimes = 1, iter = 0, nreject = 0;
    while ((totalold - (double)nfile) >= (double)1.0e-15f && itimes <= 1000) {
        for (int i2 = 0; i2 < nnew; ++i2)
            for (int j = 0; j < npairs; ++j) {
                ++iter;
                double add = (double)((2.0f * ora_ran1(rng) - 1.0f) * amp);
                for (int i1 = 0; i1 < nfile; ++i1) carray[i1] = carrayold[i1 * nnew + i2] + cwt[i1 * npairs + j] * add;
                int ip = j + nfile;
                carray[ip] = carrayold[ip * nnew + i2] + fb1 * add;
                double tdif = 0.0;
                for (int k = 0; k <= nfile; ++k) {
                    int i1 = k < nfile ? k : ip;
                    double sum = carray[i1], sumold = carrayold[i1 * nnew + i2];
                    if (i1 == i2) tdif = tdif + (sum - 1.0) * (sum - 1.0) - (sumold - 1.0) * (sumold - 1.0);
                    else tdif = tdif + sum * sum - sumold * sumold;
                }
                tdif = tdif / temp;
                int acc = tdif <= 0.0 ? 1 : (tdif > 20.0 ? 0 : ((double)ora_ran1(rng) <= exp(-tdif)));
                if (acc) {
                    total = total + tdif * temp;
                    totalold = total;
                    cwtt[j * nnew + i2] = cwtt[j * nnew + i2] + add;
                    for (int i1 = 0; i1 < nfile; ++i1) carrayold[i1 * nnew + i2] = carray[i1];
                    carrayold[ip * nnew + i2] = carray[ip];
                } else {
                    ++nreject;
                }
            }
        if (iter >= niter) {
            float fac = nreject > 0 ? 0.5f * (float)(iter - nreject) / (float)nreject : 4.0f;
            fac = fac < 0.95f ? 0.95f : (fac > 1.05f ? 1.05f : fac);
            amp = amp * fac;
            ++itimes;
            iter = nreject = 0;
            total = ora_cross(nfile, npairs, cwt, cwtt, carrayold, fb1);
            totalold = total;
            temp = (totalold - (double)nfile) / (double)(nnew * nnew);
        }
    }
    for (int j1 = 0; j1 < npairs; ++j1) {
        for (int j2 = 0; j2 < npairs; ++j2) {
            double sum = 0.0;
            for (int i = 0; i < nfile; ++i) sum = sum + cwtt[j1 * nnew + i] * cwt[i * npairs + j2];
            carray[j2] = sum + cwtt[j1 * nnew + j2 + nfile] * fb1;
        }
        for (int i = 0; i < nfile; ++i) cwtt[j1 * nnew + i] = cwtt[j1 * nnew + i] * (double)feedback * (double)efacm[i];
        for (int i = nfile; i < nnew; ++i) cwtt[j1 * nnew + i] = cwtt[j1 * nnew + i] * fb1;
    }
    return itimes;
}
