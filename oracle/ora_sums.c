/*
 * ora_sums.c -- oracle: the running averages of S(Q) and g(r) with their error bars
 * (sumsq, sumgr).  TEST INFRASTRUCTURE ONLY (see epsr_oracle.h).
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stddef.h>

/* one element of the accumulation (epsr_standalone_rev3.py:3610-3630, 3663-3680); the
 * work variables of common /sumvalues/ are real*8 */
static void accumulate(double add1, double fc1, double fc2, float *avg, float *err)
{
    double sumtot1 = (double)*avg, sumtot2 = (double)*err;
    double add2 = add1 * add1;
    sumtot2 = sumtot2 * sumtot2 + sumtot1 * sumtot1;
    sumtot1 = fc2 * sumtot1 + fc1 * add1;
    sumtot2 = fc2 * sumtot2 + fc1 * add2;
    double errort = sumtot2 - sumtot1 * sumtot1;
    errort = errort > 0.0 ? sqrt(errort) : 0.0;
    *avg = (float)sumtot1;
    *err = (float)errort;
}

/* epsr_standalone_rev3.py:3642-3685: for every pair and Q point 1..nq, intra (k = 0) and
 * inter (k = 1); arrays [(k*npairs + ic)*ldq + j] */
void ora_sumsq(int npairs, int nq, int ldq, const float *qwind, double fc1, double fc2, float *soq, float *sq2)
{
    for (int id = 0; id < npairs; ++id)
        for (int j = 1; j <= nq; ++j)
            for (int k = 0; k < 2; ++k) {
                size_t o = ((size_t)k * npairs + id) * ldq + j;
                accumulate((double)qwind[o], fc1, fc2, &soq[o], &sq2[o]);
            }
}

/* epsr_standalone_rev3.py:3578-3637: r points 2..nz, add1 = wav / ycoh (float32, 0 where
 * ycoh <= 0); then the zero limit of each pair, a double running sum of
 * ycoh (g_inter - 1) in float32 terms, stored in entsig [ic*ldw + j] (entsig(1) = 0) */
void ora_sumgr(int npairs, int nz, int ldw, const float *wav, const float *ycoh, double fc1, double fc2,
               float *gor, float *gr2, float *entsig)
{
    for (int id = 0; id < npairs; ++id) {
        double zerolimit = 0.0;
        entsig[(size_t)id * ldw + 1] = 0.0f;
        for (int j = 2; j <= nz; ++j) {
            for (int k = 0; k < 2; ++k) {
                size_t o = ((size_t)k * npairs + id) * ldw + j;
                double add1 = ycoh[j] > 0.0f ? (double)(wav[o] / ycoh[j]) : 0.0;
                accumulate(add1, fc1, fc2, &gor[o], &gr2[o]);
            }
            const float g = gor[((size_t)npairs + id) * ldw + j];
            zerolimit = zerolimit + (double)(ycoh[j] * (g - 1.0f));
            entsig[(size_t)id * ldw + j] = (float)zerolimit;
        }
    }
}
