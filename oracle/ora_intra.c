/*
 * ora_intra.c -- oracle: intramolecular energies (en_mole_intra, en_atom_intra,
 * en_mole_dihedral, en_dihedral), the intramolecular moves (movei/movea, rotati,
 * centreofmassi) and the complete molfit cycle: the shake branch, internal rotations
 * every rotfreq cycles and the whole-molecule moves, with the step adaptation of all four
 * step sizes.
 * TEST INFRASTRUCTURE ONLY (see epsr_oracle.h).
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI4ATAN 3.14159265358979323846f

/* auxiliary_routines.py:231-244 */
static float scapro(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void vecpro(const float *a, const float *b, float *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* molfit.py:922-1047 (no Q-type refinement beyond the reference's notqtype test) */
float ora_en_atom_intra(const ora_box_t *b, const ora_tables_t *t, const ora_intra_t *in, int ia, int ja,
                        float *sumevib, float *sumeang)
{
    if (ja == in->imageatom[ia]) return 0.0f;
    const float *xa = b->xa, *ya = b->ya, *za = b->za;
    float sume = 0.0f;
    const float stepcl = b->cellst / in->dmaxmol;
    int icref = b->npairref[b->itype[ia] * b->ntype + b->itype[ja]];
    const size_t o = (size_t)icref * t->ld;
    int notqtype = in->qtype[ia] == 0 && in->qtype[ja] == 0;
    float xdifs = xa[ja] - xa[ia], ydifs = ya[ja] - ya[ia], zdifs = za[ja] - za[ia];
    float xdif = xdifs, ydif = ydifs, zdif = zdifs;
    ora_minimage(&xdif, &ydif, &zdif);
    /* bond search in the atom, then in its image atom */
    int jref = 0, iflag = 0, image = ia;
    while (jref < in->nbond[image] && iflag == 0) {
        jref = jref + 1;
        if (ja == in->iabond[(size_t)image * in->ldb + jref - 1]) iflag = jref;
    }
    if (iflag == 0) {
        image = -1;
        if (in->imageatom[ia] >= 0) {
            image = in->imageatom[ia];
            jref = 0;
            while (jref < in->nbond[image] && iflag == 0) {
                jref = jref + 1;
                if (ja == in->iabond[(size_t)image * in->ldb + jref - 1]) iflag = jref;
            }
        }
    }
    if (iflag > 0 && image >= 0) {
        float dsq = xdif * xdif + ydif * ydif + zdif * zdif;
        float dtest = ora_sqrtf(dsq);
        float rat = dtest / stepcl;
        int iz = ora_ifix(rat) + 2;
        float wtms = in->amassr[icref];
        float tbo = in->tbond[(size_t)image * in->ldb + jref - 1];
        float tb = fabsf(tbo) > stepcl ? fabsf(tbo) : stepcl;
        float wtm = wtms / tb;
        if (tbo < 0.0f) wtm = wtm * in->tempang;
        float ddif = tb - dtest;
        sume = wtm * ddif * ddif * in->tempb * in->simt;
        if (tbo > 0.0f && sumevib) *sumevib = *sumevib + sume;
        if (tbo < 0.0f && sumeang) *sumeang = *sumeang + sume;
        if (in->refine_intra && iz <= t->npt2 && notqtype) {
            float dtdif = dtest - t->xstar[iz];
            sume = sume + in->refine_intra_fac * (t->xavs[o + iz] + t->xavs1[o + iz] * dtdif);
        }
    } else {
        /* atom pairs of a dihedral are left to en_mole_dihedral */
        int im = in->mref[ia];
        int found = 0;
        for (int ir = in->nrf[im]; ir <= in->nrl[im] && !found; ++ir)
            for (int kd = 0; kd < in->ndihedralang[ir] && !found; ++kd) {
                int io1 = in->labrotdih[((size_t)ir * in->ldd + kd) * 2];
                int io2 = in->labrotdih[((size_t)ir * in->ldd + kd) * 2 + 1];
                found = (io1 == ia && io2 == ja) || (io1 == ja && io2 == ia);
            }
        if (!found) {
            xdif = xdifs * b->sizefactor;
            ydif = ydifs * b->sizefactor;
            zdif = zdifs * b->sizefactor;
            ora_minimage(&xdif, &ydif, &zdif);
            float dsq = xdif * xdif + ydif * ydif + zdif * zdif;
            float dtest = ora_sqrtf(dsq);
            float rat = dtest / stepcl;
            int iz = ora_ifix(rat) + 2;
            if (iz <= t->npt2) {
                float dtdif = dtest - t->xstar[iz];
                sume = in->xav[o + iz] + in->xav1[o + iz] * dtdif;
                if (in->refine_intra && notqtype)
                    sume = sume + in->refine_intra_fac * (t->xavs[o + iz] + t->xavs1[o + iz] * dtdif);
            }
        } else {
            sume = 0.0f;
        }
    }
    return sume;
}

/* molfit.py:1495-1524 */
float ora_en_dihedral(const float *co1, const float *co2, const float *cd, float dihedralang)
{
    float cv1[3], cv2[3], cpar[3];
    vecpro(co1, cd, cv1);
    vecpro(co2, cd, cv2);
    float dmag = ora_sqrtf(scapro(cv1, cv1));
    cv1[0] = cv1[0] / dmag; cv1[1] = cv1[1] / dmag; cv1[2] = cv1[2] / dmag;
    dmag = ora_sqrtf(scapro(cv2, cv2));
    cv2[0] = cv2[0] / dmag; cv2[1] = cv2[1] / dmag; cv2[2] = cv2[2] / dmag;
    float cosphi = scapro(cv1, cv2);
    if (cosphi > 1.0f) cosphi = 1.0f;
    if (cosphi < -1.0f) cosphi = -1.0f;
    vecpro(cv1, cv2, cpar);
    float sinphi = scapro(cpar, cd);
    float difc = ora_cosf(dihedralang) - cosphi;
    float difs = ora_sinf(dihedralang) - sinphi;
    return difc * difc + difs * difs;
}

/* molfit.py:1353-1407 */
float ora_en_mole_dihedral(const ora_box_t *b, const ora_intra_t *in, int im)
{
    const float *xa = b->xa, *ya = b->ya, *za = b->za;
    float e = 0.0f;
    for (int ir = in->nrf[im]; ir <= in->nrl[im]; ++ir) {
        if (!in->isdih[ir]) continue;
        int ia1 = in->labaxi[2 * ir], ia2 = in->labaxi[2 * ir + 1];
        float x1 = xa[ia1], y1 = ya[ia1], z1 = za[ia1];
        float x2 = xa[ia2], y2 = ya[ia2], z2 = za[ia2];
        float cd[3] = {x2 - x1, y2 - y1, z2 - z1};
        float dmag = ora_sqrtf(scapro(cd, cd));
        cd[0] = cd[0] / dmag; cd[1] = cd[1] / dmag; cd[2] = cd[2] / dmag;
        for (int kd = 0; kd < in->ndihedralang[ir]; ++kd) {
            int io1 = in->labrotdih[((size_t)ir * in->ldd + kd) * 2];
            int io2 = in->labrotdih[((size_t)ir * in->ldd + kd) * 2 + 1];
            float co1[3] = {xa[io1] - x1, ya[io1] - y1, za[io1] - z1};
            float co2[3] = {xa[io2] - x2, ya[io2] - y2, za[io2] - z2};
            e = e + ora_en_dihedral(co1, co2, cd, in->dihedralang[(size_t)ir * in->ldd + kd]);
        }
    }
    return e;
}

/* molfit.py:1327-1349 */
float ora_en_mole_intra(const ora_box_t *b, const ora_tables_t *t, const ora_intra_t *in, int im, float *sumevib,
                        float *sumeang, float *sumedih)
{
    float e = 0.0f;
    for (int ia = b->naf[im]; ia < b->nal[im]; ++ia)
        for (int ja = ia + 1; ja <= b->nal[im]; ++ja) e = e + ora_en_atom_intra(b, t, in, ia, ja, sumevib, sumeang);
    if (in->ngrot[im] > 0 && in->tempdih > 0.0f) {
        float sume = in->tempdih * in->simt * ora_en_mole_dihedral(b, in, im);
        e = e + sume;
        if (sumedih) *sumedih = *sumedih + sume;
    }
    return e;
}

/* ato_manipulation_rev4.py:442-475 */
static void centreofmassi(const ora_box_t *b, const ora_intra_t *in, float *xa, float *ya, float *za, int im)
{
    float x1 = 0.0f, y1 = 0.0f, z1 = 0.0f, summass = 0.0f;
    for (int ia = b->naf[im]; ia <= b->nal[im]; ++ia) {
        if (in->qtype[ia] != 0) continue;
        float amass = in->amassa[b->itype[ia]];
        x1 = x1 + xa[ia] * amass;
        y1 = y1 + ya[ia] * amass;
        z1 = z1 + za[ia] * amass;
        summass = summass + amass;
    }
    if (summass > 0.0f) {
        x1 = x1 / summass;
        y1 = y1 / summass;
        z1 = z1 / summass;
    }
    for (int ia = b->naf[im]; ia <= b->nal[im]; ++ia) {
        xa[ia] = xa[ia] - x1;
        ya[ia] = ya[ia] - y1;
        za[ia] = za[ia] - z1;
    }
}

/* molfit.py:545-565: every atom of the molecule, z then y then x */
static void movei(const ora_box_t *b, float *xa, float *ya, float *za, int im, float step, ora_ran1_t *rng)
{
    for (int ia = b->naf[im]; ia <= b->nal[im]; ++ia) {
        za[ia] = za[ia] + step * (2.0f * ora_ran1(rng) - 1.0f);
        ya[ia] = ya[ia] + step * (2.0f * ora_ran1(rng) - 1.0f);
        xa[ia] = xa[ia] + step * (2.0f * ora_ran1(rng) - 1.0f);
    }
}

/* molfit.py:589-655: rotate the atoms of group ig about its axis */
static void rotati(const ora_intra_t *in, float *xa, float *ya, float *za, int ig, float step, ora_ran1_t *rng)
{
    const float pi = PI4ATAN;
    float phi = pi * step * (2.0f * ora_ran1(rng) - 1.0f);
    float cphi = ora_cosf(phi);
    float sphi = copysignf(fabsf(ora_sqrtf(1.0f - cphi * cphi)), phi);
    int a1 = in->labaxi[2 * ig], a2 = in->labaxi[2 * ig + 1];
    float p[3] = {xa[a2] - xa[a1], ya[a2] - ya[a1], za[a2] - za[a1]};
    float rab2 = scapro(p, p);
    if (!(rab2 > 0.0f)) return;
    float rab = ora_sqrtf(rab2);
    for (int i = 0; i < 3; ++i) p[i] = p[i] / rab;
    for (int k = in->rotf[ig]; k <= in->rotl[ig]; ++k) {
        int nat = in->labrot[k];
        int nax = a1;
        float xv[3] = {xa[nat] - xa[nax], ya[nat] - ya[nax], za[nat] - za[nax]};
        float frac = scapro(p, xv);
        float xp[3], xr[3], xrp[3];
        for (int i = 0; i < 3; ++i) xp[i] = frac * p[i];
        for (int i = 0; i < 3; ++i) xr[i] = xv[i] - xp[i];
        vecpro(p, xr, xrp);
        xa[nat] = xa[nax] + xp[0];
        ya[nat] = ya[nax] + xp[1];
        za[nat] = za[nax] + xp[2];
        xa[nat] = xa[nat] + xr[0] * cphi + xrp[0] * sphi;
        ya[nat] = ya[nat] + xr[1] * cphi + xrp[1] * sphi;
        za[nat] = za[nat] + xr[2] * cphi + xrp[2] * sphi;
    }
}

/* molfit.py:671-711 */
static void rotatm(const ora_box_t *b, float *xa, float *ya, float *za, int im, float stepsz, ora_ran1_t *rng)
{
    const float pi = PI4ATAN;
    int i1 = b->naf[im], i2 = b->nal[im];
    if (i2 <= i1) return;
    float phi = pi * stepsz * (2.0f * ora_ran1(rng) - 1.0f);
    float cphi = ora_cosf(phi);
    float sphi = copysignf(fabsf(ora_sqrtf(1.0f - cphi * cphi)), phi);
    for (int ia = i1; ia <= i2; ++ia) {
        float s = xa[ia] * cphi - ya[ia] * sphi;
        ya[ia] = xa[ia] * sphi + ya[ia] * cphi;
        xa[ia] = s;
    }
    phi = pi * stepsz * (2.0f * ora_ran1(rng) - 1.0f);
    cphi = ora_cosf(phi);
    sphi = copysignf(fabsf(ora_sqrtf(1.0f - cphi * cphi)), phi);
    for (int ia = i1; ia <= i2; ++ia) {
        float s = za[ia] * cphi - xa[ia] * sphi;
        xa[ia] = za[ia] * sphi + xa[ia] * cphi;
        za[ia] = s;
    }
    phi = pi * stepsz * (2.0f * ora_ran1(rng) - 1.0f);
    cphi = ora_cosf(phi);
    sphi = copysignf(fabsf(ora_sqrtf(1.0f - cphi * cphi)), phi);
    for (int ia = i1; ia <= i2; ++ia) {
        float s = ya[ia] * cphi - za[ia] * sphi;
        za[ia] = ya[ia] * sphi + za[ia] * cphi;
        ya[ia] = s;
    }
}

/* molfit.py:732-759 */
static float movem(const ora_centres_t *c, int im, float stepsz, ora_ran1_t *rng)
{
    float xs = c->xadd[im], ys = c->yadd[im], zs = c->zadd[im];
    c->zadd[im] = c->zadd[im] + stepsz * (2.0f * ora_ran1(rng) - 1.0f);
    c->yadd[im] = c->yadd[im] + stepsz * (2.0f * ora_ran1(rng) - 1.0f);
    c->xadd[im] = c->xadd[im] + stepsz * (2.0f * ora_ran1(rng) - 1.0f);
    int nochange = xs == c->xadd[im] && ys == c->yadd[im] && zs == c->zadd[im];
    while (nochange) {
        stepsz = stepsz * 2.0f;
        c->zadd[im] = c->zadd[im] + stepsz * (2.0f * ora_ran1(rng) - 1.0f);
        c->yadd[im] = c->yadd[im] + stepsz * (2.0f * ora_ran1(rng) - 1.0f);
        c->xadd[im] = c->xadd[im] + stepsz * (2.0f * ora_ran1(rng) - 1.0f);
        nochange = xs == c->xadd[im] && ys == c->yadd[im] && zs == c->zadd[im];
    }
    return stepsz;
}

/* molfit.py:805-919 over explicit ergold/ergnew rows (make_neighbour_list_from_mols,
 * ergpairreference.py:165-188): ergnew[1..] recomputed, ergnew[0] untouched */
static float en_change_rows(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t, const ora_ledger_t *l,
                            int im, const float *ergold, float *ergnew)
{
    const int M = l->nnbrefmax;
    const float stepcl = b->cellst / b->dmax;
    int nnbref = l->nergpairm[im];
    for (int inn = 1; inn < nnbref; ++inn) ergnew[inn] = 0.0f;
    for (int inn = 1; inn < nnbref; ++inn) {
        int jm = l->molergpair[(size_t)im * M + inn];
        float sume = 0.0f;
        for (int ia = b->naf[im]; ia <= b->nal[im]; ++ia) {
            int it = b->itype[ia];
            float x1 = b->xa[ia] + c->xadd[im], y1 = b->ya[ia] + c->yadd[im], z1 = b->za[ia] + c->zadd[im];
            for (int ja = b->naf[jm]; ja <= b->nal[jm]; ++ja) {
                int icref = b->npairref[it * b->ntype + b->itype[ja]];
                float x2 = b->xa[ja] + c->xadd[jm], y2 = b->ya[ja] + c->yadd[jm], z2 = b->za[ja] + c->zadd[jm];
                sume = sume + ora_en_atom_inter_scale(x1, y1, z1, x2, y2, z2, stepcl, t, icref, 1.0f);
            }
        }
        ergnew[inn] = ergnew[inn] + sume;
    }
    float fdiff = 0.0f;
    for (int inn = 1; inn < nnbref; ++inn) fdiff = fdiff + ergnew[inn] - ergold[inn];
    return fdiff;
}

static float etest_of(float ediff)
{
    if (ediff <= 0.0f) return 10.0f;
    if (ediff < 20.0f) return ora_expf(-ediff);
    return 0.0f;
}

static float adapt(float rejrate, int att, int rej)
{
    float rat;
    if (rej > 0) {
        rat = rejrate * (float)att / (float)rej;
        if (rat > 1.1f) rat = 1.1f;
        if (rat < 0.9f) rat = 0.9f;
    } else {
        rat = att > 0 ? 1.1f : 1.0f;
    }
    return rat;
}

static void clamp_step(float *s, float mx, float mn)
{
    if (*s > mx) *s = mx;
    if (*s < mn && *s != 0.0f) *s = mn;
}

/* molfit.py:51-541: the complete cycle (no tethered molecules, three-body terms or
 * coordination constraints) */
int ora_molfit_full(const ora_box_t *b0, float *xa, float *ya, float *za, const ora_centres_t *c,
                    const ora_tables_t *t, ora_intra_t *in, const ora_cells_t *cl, ora_ledger_t *l,
                    ora_mcfull_t *mc, ora_ran1_t *rng, float *fadds, float *faddbs, int *nisum, int intra,
                    int rotfreq)
{
    ora_box_t bb = *b0;
    bb.xa = xa; bb.ya = ya; bb.za = za;
    const ora_box_t *b = &bb;
    const int M = l->nnbrefmax;
    const int nt = b->nmoltype;
    float temp = 1.0f / mc->simt;
    for (int k = 0; k < 2 * nt; ++k) mc->nmovei[k] = mc->ninter[k] = mc->nwhole[k] = mc->nmovem[k] = 0;
    float trstepmax = ora_powf(1.0f / b->natomr, 1.0f / 3.0f);
    float trstepmin = 0.0001f * trstepmax;
    if (trstepmin < 1.0e-6f) trstepmin = 1.0e-6f;
    const float rotstepmax = 1.0f, rotstepmin = 1.0e-5f;
    float fadd = 0.0f, faddb = 0.0f;
    in->tempang = mc->tempangwt;
    in->tempdih = mc->tempdihwt * in->tempb;
    int maxat = 1;
    for (int im = 0; im < b->nmol; ++im)
        if (b->nal[im] - b->naf[im] + 1 > maxat) maxat = b->nal[im] - b->naf[im] + 1;
    float *xr = malloc(sizeof(float) * (size_t)maxat * 3);
    float *ergold = malloc(sizeof(float) * (size_t)(M > 0 ? M : 1));
    float *ergnew = malloc(sizeof(float) * (size_t)(M > 0 ? M : 1));
    int nmoves = 0;
    const int shake = *nisum >= intra;
    int nchain = 0;
    for (int cc = 0; cc < cl->ncell; ++cc) nchain += cl->nmolsincell[cc];
    for (int pos = 0, cc = 0, iref = 0; pos < (shake ? b->nmol : nchain); ++pos) {
        int im;
        if (shake) {
            im = pos;
        } else {
            while (iref >= cl->nmolsincell[cc]) { cc++; iref = 0; }
            im = cl->molsincell[(size_t)cc * cl->maxcellno + iref];
            iref++;
        }
        int n = b->nal[im] - b->naf[im] + 1;
        int mt = b->mtype[im];
        /* make_neighbour_list_from_mols */
        int nnbref = l->nergpairm[im];
        for (int k = 0; k < nnbref; ++k) {
            ergold[k] = l->ergpair[l->ergpairref[(size_t)im * M + k]];
            ergnew[k] = ergold[k];
        }
#define SAVE_XR() for (int k = 0; k < n; ++k) { xr[3*k] = xa[b->naf[im]+k]; xr[3*k+1] = ya[b->naf[im]+k]; xr[3*k+2] = za[b->naf[im]+k]; }
#define RESTORE_XR() for (int k = 0; k < n; ++k) { xa[b->naf[im]+k] = xr[3*k]; ya[b->naf[im]+k] = xr[3*k+1]; za[b->naf[im]+k] = xr[3*k+2]; }
#define SAVE_ENERGIES() for (int k = 0; k < nnbref; ++k) ergold[k] = ergnew[k]
#define SAVE_ERG_PAIRS() for (int k = 0; k < nnbref; ++k) l->ergpair[l->ergpairref[(size_t)im * M + k]] = ergold[k]
        if (shake) {
            if (n > 1 && mc->stepmi > 0.0f) {
                float stepmithis = mc->interntrans[mt];
                if (mc->stepmi > 0.0f && stepmithis == 0.0f) stepmithis = trstepmin;
                mc->interntrans[mt] = stepmithis;
                SAVE_XR();
                mc->nmovei[2 * mt] += 1;
                movei(b, xa, ya, za, im, stepmithis, rng);
                float e = ora_en_mole_intra(b, t, in, im, NULL, NULL, NULL);
                ergnew[0] = e;
                float fdiffb = ergnew[0] - ergold[0];
                float ediff = fdiffb * temp;
                nmoves++;
                if (ora_ran1(rng) < etest_of(ediff)) {
                    faddb = faddb + fdiffb;
                    SAVE_ENERGIES();
                    centreofmassi(b, in, xa, ya, za, im);
                } else {
                    mc->nmovei[2 * mt + 1] += 1;
                    RESTORE_XR();
                }
                SAVE_ERG_PAIRS();
            }
        }
        int rot_now = mc->stepri > 0.0f && in->ngrot[im] > 0 &&
                      (shake || *nisum == rotfreq * (*nisum / rotfreq));
        if (rot_now) {
            if (mc->internrot[mt] == 0.0f) mc->internrot[mt] = rotstepmin;
            mc->ninter[2 * mt] += 1;
            SAVE_XR();
            for (int ig = in->nrf[im]; ig <= in->nrl[im]; ++ig) rotati(in, xa, ya, za, ig, mc->internrot[mt], rng);
            float e = ora_en_mole_intra(b, t, in, im, NULL, NULL, NULL);
            ergnew[0] = e;
            float fdiffb = ergnew[0] - ergold[0];
            float fdiff = en_change_rows(b, c, t, l, im, ergold, ergnew);
            float ediff = (fdiff + fdiffb) * temp;
            nmoves++;
            if (ora_ran1(rng) < etest_of(ediff)) {
                mc->nstepreject[im] = 0;
                fadd = fadd + fdiff;
                faddb = faddb + fdiffb;
                SAVE_ENERGIES();
            } else {
                mc->nstepreject[im] += 1;
                RESTORE_XR();
                mc->ninter[2 * mt + 1] += 1;
            }
            SAVE_ERG_PAIRS();
        } else if (in->ngrot[im] == 0) {
            mc->internrot[mt] = 0.0f;
        }
        if (shake) continue;
        /* whole-molecule rotation and translation */
        SAVE_XR();
        float xs = c->xadd[im], ys = c->yadd[im], zs = c->zadd[im];
        if (mc->stepr > 0.0f && n > 1) {
            if (mc->wholerot[mt] <= 0.0f) mc->wholerot[mt] = rotstepmin;
            mc->nwhole[2 * mt] += 1;
            rotatm(b, xa, ya, za, im, mc->wholerot[mt], rng);
        } else {
            mc->wholerot[mt] = 0.0f;
        }
        if (mc->stepm > 0.0f) {
            if (mc->wholetrans[mt] <= 0.0f) mc->wholetrans[mt] = trstepmin;
            mc->nmovem[2 * mt] += 1;
            mc->wholetrans[mt] = movem(c, im, mc->wholetrans[mt], rng);
        } else {
            mc->wholetrans[mt] = 0.0f;
        }
        float fdiff = en_change_rows(b, c, t, l, im, ergold, ergnew);
        float ediff = (fdiff + 0.0f) * temp;
        nmoves++;
        if (ora_ran1(rng) < etest_of(ediff)) {
            mc->nstepreject[im] = 0;
            fadd = fadd + fdiff;
            faddb = faddb + 0.0f;
            SAVE_ENERGIES();
        } else {
            mc->nstepreject[im] += 1;
            RESTORE_XR();
            c->xadd[im] = xs; c->yadd[im] = ys; c->zadd[im] = zs;
            if (mc->stepr > 0.0f) mc->nwhole[2 * mt + 1] += 1;
            if (mc->stepm > 0.0f) mc->nmovem[2 * mt + 1] += 1;
        }
        SAVE_ERG_PAIRS();
#undef SAVE_XR
#undef RESTORE_XR
#undef SAVE_ENERGIES
#undef SAVE_ERG_PAIRS
    }
    free(xr);
    free(ergold);
    free(ergnew);
    *nisum = shake ? 1 : *nisum + 1;
    if (mc->ireset != 1) {
        for (int mt = 0; mt < nt; ++mt) {
            mc->interntrans[mt] = mc->interntrans[mt] * adapt(mc->rejrate, mc->nmovei[2 * mt], mc->nmovei[2 * mt + 1]);
            clamp_step(&mc->interntrans[mt], trstepmax, trstepmin);
            mc->internrot[mt] = mc->internrot[mt] * adapt(mc->rejrate, mc->ninter[2 * mt], mc->ninter[2 * mt + 1]);
            clamp_step(&mc->internrot[mt], rotstepmax, rotstepmin);
            mc->wholerot[mt] = mc->wholerot[mt] * adapt(mc->rejrate, mc->nwhole[2 * mt], mc->nwhole[2 * mt + 1]);
            clamp_step(&mc->wholerot[mt], rotstepmax, rotstepmin);
            mc->wholetrans[mt] = mc->wholetrans[mt] * adapt(mc->rejrate, mc->nmovem[2 * mt], mc->nmovem[2 * mt + 1]);
            clamp_step(&mc->wholetrans[mt], trstepmax, trstepmin);
        }
    }
    *fadds = *fadds + fadd / (float)b->nmol;
    *faddbs = *faddbs + faddb / (float)b->nmol;
    return nmoves;
}
