/*
 * ora_data.c -- oracle: the data-minus-simulation block of the epsr program and ftback,
 * and ptgnsq's differences with X-ray weights.  TEST INFRASTRUCTURE ONLY (see
 * epsr_oracle.h).
 *
 * Arithmetic follows the reference's float32 statements.  Its weights are float64 arrays
 * (weights.ciwt/cwt hold float32 products, x_rayf0.weightx full doubles), so each step
 * of a weighted sum, `sum = f32(sum + w * s)`, is formed in double and rounded to float.
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stddef.h>

/* epsr_standalone_rev3.py:5330-5389 (ftback): the part of sofq[nqmin..nqmax] that comes
 * from r < rmin, in place.  grt[j] (r index j) keeps what earlier calls left beyond the
 * entries this call writes, as the program's one work array does. */
void ora_ftback(int nqmin, int nqmax, const float *q, const float *x, int npt1, float rho, float rmin, float *sofq,
                float *grt)
{
    for (int i = nqmin; i <= nqmax; ++i) sofq[i] = sofq[i] * q[i];
    const float pi = 3.141592653f;
    const float qstep = q[4] - q[3];
    const float sfac = 0.5f * qstep / pi / pi / rho;
    int jsave = 2;
    for (int j = 2; j <= npt1; ++j) {
        const float rr = x[j];
        if (rr >= rmin) break;
        jsave = j;
        float sum = 0.0f;
        for (int i = nqmin; i <= nqmax; ++i) {
            const float qr = q[i] * rr;
            sum = sum + sofq[i] * ora_sinf(qr);
        }
        grt[j] = sfac * sum;
    }
    const float gfac = 4.0f * pi * rho;
    for (int i = nqmin; i <= nqmax; ++i) {
        const float qq = q[i];
        float r1 = 0.0f, sum = 0.0f;
        if (qq > 0.0f) {
            for (int j = 2; j <= jsave; ++j) {
                const float rr = x[j];
                const float r2 = 2.0f * rr - r1;
                const float rstep = r2 - r1;
                const float qr = qq * rr;
                sum = sum + grt[j] * ora_sinf(qr) * rstep;
                r1 = r2;
            }
            sofq[i] = gfac * sum / qq;
        } else {
            for (int j = 2; j <= jsave; ++j) {
                const float rr = x[j];
                const float r2 = 2.0f * rr - r1;
                const float rstep = r2 - r1;
                sum = sum + grt[j] * rr * rstep;
                r1 = r2;
            }
            sofq[i] = gfac * sum;
        }
    }
}

/* the weighted sum over pairs of soq at Q index iq for dataset id (X-ray: weightx of
 * file ix >= 1; neutron: ciwt for the intramolecular partials, cwt for the others), each
 * term added in double and the sum rounded to float */
static float pair_sum(const ora_dms_t *d, int id, int iq, const float *soq)
{
    const int ix = d->xfile[id];
    float sum = 0.0f;
    if (ix > 0) {
        const double *w = d->weightx + ((size_t)(ix - 1) * d->ldq + iq) * d->npairs;
        for (int ip = 0; ip < d->npairs; ++ip) sum = (float)((double)sum + w[ip] * (double)soq[(size_t)ip * d->ldq + iq]);
    } else {
        const float *w = (soq == d->soqintra ? d->ciwt : d->cwt) + (size_t)id * d->npairs;
        for (int ip = 0; ip < d->npairs; ++ip)
            sum = (float)((double)sum + (double)w[ip] * (double)soq[(size_t)ip * d->ldq + iq]);
    }
    return sum;
}

/* epsr_standalone_rev3.py:1703-1830: the intramolecular terms out of the data (datat,
 * szerostemp), then per dataset the difference data - simulation with the extrapolation
 * below qmin and the Porod term (dts), whose part from r < rshmin (ftback) is subtracted
 * from datat when nsmoop > 0.  dts [ldq] and grt [npt1 + 1] are work arrays (grt persists
 * across calls, as the program's). */
void ora_data_minus_sim(const ora_dms_t *d, float *datat, float *szerostemp, float *dts, float *grt)
{
    const int nq = d->nq, ldq = d->ldq;
    for (int id = 0; id < d->nfile; ++id) {
        float *dt = datat + (size_t)id * ldq;
        for (int iq = 1; iq <= nq; ++iq) {
            if (iq >= d->nqmind[id] && iq <= d->nqmaxd[id]) dt[iq] = d->datai[(size_t)id * ldq + iq] - pair_sum(d, id, iq, d->soqintra);
            else dt[iq] = 0.0f;
        }
        szerostemp[id] = 0.0f;
        if (d->szeros[id] != 0.0f) szerostemp[id] = d->szeros[id] - pair_sum(d, id, 1, d->soqintra);
    }
    for (int id = 0; id < d->nfile; ++id) {
        float *dt = datat + (size_t)id * ldq;
        const int nmn = d->nqmin[id], nmx = d->nqmax[id];
        float grads0, conss0;
        if (d->szeros[id] != 0.0f && nmn > 1) {
            grads0 = (dt[nmn] - szerostemp[id]) / (d->q[nmn] - d->q[1]);
            conss0 = szerostemp[id];
        } else {
            grads0 = 0.0f;
            conss0 = dt[nmn];
        }
        for (int i = 1; i <= nq; ++i) {
            const float sum = pair_sum(d, id, i, d->soqinter);
            if (i < nmn) {
                dts[i] = d->szeros[id] != 0.0f ? conss0 + grads0 * (d->q[i] - d->q[1]) - sum : 0.0f;
            } else if (i <= nmx) {
                dts[i] = dt[i] - sum;
                if (d->porodamp[id] > 0.0f) dts[i] = dts[i] + d->porodamp[id] * ora_powf(d->porodq / d->q[i], d->porodpwr);
            } else {
                dts[i] = 0.0f;
            }
        }
        if (d->nsmoop > 0) {
            ora_ftback(nmn, nmx, d->q, d->x, d->npt1, d->rho, d->rshmin[id], dts, grt);
            for (int i = nmn; i <= nmx; ++i) dt[i] = dt[i] - dts[i];
        }
    }
}

/* ptgnsq's differences (epsr_standalone_rev3.py:3827-3867): per dataset and Q the
 * weighted sum of the intermolecular partials (cwt, or for a dataset with xfile[id] = ix
 * >= 1 the Q-dependent weightx[((ix-1)*ldq + i)*npairs + ic]), each term added in double
 * and rounded to float, subtracted from the data in the fitted range, extrapolated below */
void ora_ptgnsq_diff_x(int nfile, int nq, int npairs, const float *q, const float *data, int ldq, const float *qwind,
                       const float *cwt, const double *weightx, const int *xfile, const float *szeros,
                       const int *nqmin, const int *nqmax, float qminptemp, float *qtemp)
{
    for (int id = 0; id < nfile; ++id) {
        const float *dat = data + (size_t)id * ldq;
        const int nqmn = nqmin[id];
        const int ix = xfile ? xfile[id] : 0;
        float grads0 = 0.0f, conss0 = 0.0f;
        if (szeros[id] != 0.0f && nqmn > 1) {
            grads0 = (dat[nqmn] - szeros[id]) / (q[nqmn] - q[1]);
            conss0 = szeros[id];
        }
        for (int i = 1; i <= nq; ++i) {
            float sum = 0.0f;
            if (ix > 0) {
                const double *w = weightx + ((size_t)(ix - 1) * ldq + i) * npairs;
                for (int ic = 0; ic < npairs; ++ic) sum = (float)((double)sum + w[ic] * (double)qwind[(size_t)ic * ldq + i]);
            } else {
                for (int ic = 0; ic < npairs; ++ic)
                    sum = (float)((double)sum + (double)cwt[id * npairs + ic] * (double)qwind[(size_t)ic * ldq + i]);
            }
            const float qlim = qminptemp > q[nqmn] ? qminptemp : q[nqmn];
            float v;
            if (q[i] < qlim) v = szeros[id] != 0.0f ? conss0 + grads0 * (q[i] - q[1]) - sum : 0.0f;
            else if (i >= nqmn && i <= nqmax[id]) v = dat[i] - sum;
            else v = 0.0f;
            qtemp[(size_t)id * ldq + i] = v;
        }
    }
}
