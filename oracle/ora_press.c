/*
 * ora_press.c -- oracle: press, the whole-box intermolecular energy and virial
 * pressure.  TEST INFRASTRUCTURE ONLY (see epsr_oracle.h).
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stddef.h>

/* src/epsr26/library/ato_manipulation_rev4.py:3901-4063 (no three-body terms, no
 * coordination constraints; energf/pressf and energabs are host bookkeeping).
 * For molecule im (all but the last), over its ledger partners jm > im in ledger
 * order, one float32 energy and one virial per partner from every atom pair
 * (ia outer, ja inner) whose bin iz = IFIX(r/stepcl) + 2 is <= npt1:
 *   ergold += enwt; ergold += enpot               (float32, in that order)
 *   ergnew  = float32(double(float32(ergnew + prwt)) + double(prpot))
 * with prfac = (Rc . r)/|r| (centre separation on the atom separation), the slopes
 * averaged over bins iz and iz-1.  sumc (float32) and presst (double) add the
 * partners' values in ledger order.  Outputs: energy = sumc/nmol (kJ/mol),
 * pressc (kbar), presst (the virial sum), and per ledger slot the pair values
 * (slote/slotw, [nmol*nnbrefmax], 0 where jm <= im) when not NULL. */
void ora_press(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t, const ora_ledger_t *l,
               int npt1, float simt, float volume, float *energy, float *pressc, double *presst_out,
               float *slote, float *slotw)
{
    const float stepcl = b->cellst / b->dmax;
    const float *xadd = c->xadd, *yadd = c->yadd, *zadd = c->zadd;
    float sumc = 0.0f;
    double presst = 0.0;
    for (int im = 0; im < b->nmol - 1; ++im) {
        const float xc1 = xadd[im], yc1 = yadd[im], zc1 = zadd[im];
        const int nn = l->nergpairm[im];
        for (int jref = 0; jref < nn; ++jref) {
            const int jm = l->molergpair[(size_t)im * l->nnbrefmax + jref];
            float ergold = 0.0f, ergnew = 0.0f;
            if (jm > im) {
                for (int ia = b->naf[im]; ia <= b->nal[im]; ++ia) {
                    const int it = b->itype[ia];
                    const float x1 = b->xa[ia] + xc1, y1 = b->ya[ia] + yc1, z1 = b->za[ia] + zc1;
                    for (int ja = b->naf[jm]; ja <= b->nal[jm]; ++ja) {
                        const float xc2 = xadd[jm], yc2 = yadd[jm], zc2 = zadd[jm];
                        const int icref = b->npairref[it * b->ntype + b->itype[ja]];
                        const float x2 = b->xa[ja] + xc2, y2 = b->ya[ja] + yc2, z2 = b->za[ja] + zc2;
                        float xdif = x2 - x1, ydif = y2 - y1, zdif = z2 - z1;
                        ora_minimage(&xdif, &ydif, &zdif);
                        float xcdif = xc2 - xc1, ycdif = yc2 - yc1, zcdif = zc2 - zc1;
                        ora_minimage(&xcdif, &ycdif, &zcdif);
                        const float dsq = xdif * xdif + ydif * ydif + zdif * zdif;
                        const float dtest = ora_sqrtf(dsq);
                        const float rat = dtest / stepcl;
                        const int iz = ora_ifix(rat) + 2;
                        if (iz <= npt1) {
                            const size_t o = (size_t)icref * t->ld;
                            const float dtdif = dtest - t->xstar[iz];
                            const float prfac = dtest > 0.0f ? (xcdif * xdif + ycdif * ydif + zcdif * zdif) / dtest : 0.0f;
                            const float enpot = t->ecoul[o + iz] + t->ecoul1[o + iz] * dtdif;
                            const float prpotf = 0.5f * (t->ecoul1[o + iz] + t->ecoul1[o + iz - 1]) * prfac;
                            const float enwt = t->xavs[o + iz] + t->xavs1[o + iz] * dtdif;
                            const float prwt = 0.5f * (t->xavs1[o + iz] + t->xavs1[o + iz - 1]) * prfac;
                            ergold = ergold + enwt + enpot;
                            const float s = ergnew + prwt;
                            ergnew = (float)((double)s + (double)prpotf);
                        }
                    }
                }
            }
            if (slote) slote[(size_t)im * l->nnbrefmax + jref] = ergold;
            if (slotw) slotw[(size_t)im * l->nnbrefmax + jref] = ergnew;
            sumc = sumc + ergold;
            presst = presst + (double)ergnew;
        }
    }
    *energy = sumc / (float)b->nmol;
    *pressc = (float)((((double)((float)b->nmol * simt)) - presst / 3.0) * 100.0 / (double)(6.0222f * volume));
    *presst_out = presst;
}
