/*
 * ora_energy.c -- oracle: pair energies by table lookup, cell lists, the
 * pair-energy ledger, fmolec1, en_change and the whole-molecule MC cycle.
 * TEST INFRASTRUCTURE ONLY (see epsr_oracle.h).
 */
#include "epsr_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI4ATAN 3.14159265358979323846f

/* src/epsr26/library/molfit.py:1250-1279 */
float ora_en_atom_inter_scale(float x1, float y1, float z1, float x2, float y2, float z2,
                              float stepcl, const ora_tables_t *t, int icref, float scalefac)
{
    float sume = 0.0f;
    float xdif = x2 - x1, ydif = y2 - y1, zdif = z2 - z1;
    ora_minimage(&xdif, &ydif, &zdif);
    float dsq = xdif * xdif + ydif * ydif + zdif * zdif;
    float dtest = ora_sqrtf(dsq);
    float rat = dtest / stepcl;
    int iz = ora_ifix(rat) + 2;
    const size_t o = (size_t)icref * t->ld;
    if (iz <= t->npt2) {
        float dtdif = dtest - t->xstar[iz];
        sume = t->xavs[o + iz] + t->xavs1[o + iz] * dtdif;
    }
    dtest = dtest * scalefac;
    rat = dtest / stepcl;
    iz = ora_ifix(rat) + 2;
    if (iz <= t->npt2) {
        float dtdif = dtest - t->xstar[iz];
        sume = sume + t->ecoul[o + iz] + t->ecoul1[o + iz] * dtdif;
    }
    return sume;
}

/* molfit.py:1527-1578 (no tethering, no coordination constraints, scalefac 1) */
float ora_en_mole_pair(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t, int im, int jm)
{
    const float stepcl = b->cellst / b->dmax;
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
    return sume;
}

/* ------------------------------------------------------------------ cell lists */
/* cellneighbourlist.py:309-330 (cubic lattice in box fractions: a=(1,0,0), b=(0,1,0), c=(0,0,1)) */
static int cellcoord(float x, int ncelldiv)
{
    float x1 = x - (float)ora_nint(x);
    int ix = ora_ifix((x1 + 0.5f) * (float)ncelldiv);
    while (ix < 0) ix += ncelldiv;
    while (ix >= ncelldiv) ix -= ncelldiv;
    return ix;
}

static int atomcellno(const ora_cells_t *cl, float x, float y, float z)
{
    float fc = z / 1.0f;
    float fb = (y - fc * 0.0f) / 1.0f;
    float fa = (x - fb * 0.0f - fc * 0.0f) / 1.0f;
    int ia = cellcoord(fa, cl->ncelldiva);
    int ib = cellcoord(fb, cl->ncelldivb);
    int ic = cellcoord(fc, cl->ncelldivc);
    return ia + cl->ncelldiva * (ib + cl->ncelldivb * ic);      /* 0-based cell */
}

/* cellneighbourlist.py:102-194 */
static void makecellneighbourlist(ora_cells_t *cl, float rmax)
{
    const float latc[3][3] = {{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}};
    float xc[8], yc[8], zc[8];
    float cella = 1.0f / (float)cl->ncelldiva, cellb = 1.0f / (float)cl->ncelldivb,
          cellc = 1.0f / (float)cl->ncelldivc;
    float cella2 = 0.5f * cella, cellb2 = 0.5f * cellb, cellc2 = 0.5f * cellc;
    int nf = 0;
    for (int ic = -1; ic <= 1; ic += 2) {
        float cx = (float)ic * cellc2 * latc[2][0], cy = (float)ic * cellc2 * latc[2][1],
              cz = (float)ic * cellc2 * latc[2][2];
        for (int ib = -1; ib <= 1; ib += 2) {
            float bx = (float)ib * cellb2 * latc[1][0], by = (float)ib * cellb2 * latc[1][1],
                  bz = (float)ib * cellb2 * latc[1][2];
            for (int ia = -1; ia <= 1; ia += 2) {
                float ax = (float)ia * cella2 * latc[0][0], ay = (float)ia * cella2 * latc[0][1],
                      az = (float)ia * cella2 * latc[0][2];
                xc[nf] = ax + bx + cx; yc[nf] = ay + by + cy; zc[nf] = az + bz + cz;
                nf++;
            }
        }
    }
    int nra = (cl->ncelldiva - 1) / 2, nrb = (cl->ncelldivb - 1) / 2, nrc = (cl->ncelldivc - 1) / 2;
    float rmaxsq = rmax * rmax;
    int found_n = 0;
    for (int ic = -nrc; ic <= nrc; ++ic) {
        float cx = (float)ic * cellc * latc[2][0], cy = (float)ic * cellc * latc[2][1],
              cz = (float)ic * cellc * latc[2][2];
        for (int ib = -nrb; ib <= nrb; ++ib) {
            float bx = (float)ib * cellb * latc[1][0], by = (float)ib * cellb * latc[1][1],
                  bz = (float)ib * cellb * latc[1][2];
            for (int ia = -nra; ia <= nra; ++ia) {
                float ax = (float)ia * cella * latc[0][0], ay = (float)ia * cella * latc[0][1],
                      az = (float)ia * cella * latc[0][2];
                float xdc = ax + bx + cx, ydc = ay + by + cy, zdc = az + bz + cz;
                int found = 0;
                for (int i1 = 0; i1 < 8 && !found; ++i1)
                    for (int i2 = 0; i2 < 8 && !found; ++i2) {
                        float xd = xdc + xc[i1] - xc[i2], yd = ydc + yc[i1] - yc[i2], zd = zdc + zc[i1] - zc[i2];
                        float rsq = xd * xd + yd * yd + zd * zd;
                        found = (rsq < rmaxsq) || (ia == 0 && ib == 0 && ic == 0);
                    }
                if (found) {
                    cl->iacellneighb[found_n] = ia; cl->ibcellneighb[found_n] = ib; cl->iccellneighb[found_n] = ic;
                    found_n++;
                }
            }
        }
    }
    cl->ncellneighb = found_n;
}

/* cellneighbourlist.py:23-96 */
int ora_cells_init(ora_cells_t *cl, int nmolcellin, const float *xadd, const float *yadd,
                   const float *zadd, int nmol, float rmax)
{
    memset(cl, 0, sizeof(*cl));
    int nmolcell = nmolcellin > nmol ? nmol : nmolcellin;
    int ncelldiv = ora_nint(ora_powf((float)nmol / (float)nmolcell, 1.0f / 3.0f));
    if (ncelldiv < 1) ncelldiv = 1;
    float cellsize = 2.0f * 1.0f / (float)ncelldiv;
    cl->ncelldiva = 2 * ora_ifix(1.0f / cellsize) + 1;
    cl->ncelldivb = 2 * ora_ifix(1.0f / cellsize) + 1;
    cl->ncelldivc = 2 * ora_ifix(1.0f / cellsize) + 1;
    cl->ncell = cl->ncelldiva * cl->ncelldivb * cl->ncelldivc;
    cl->ncelldivsq = cl->ncelldiva * cl->ncelldivb;
    cl->nmolsincell = calloc((size_t)cl->ncell, sizeof(int));
    cl->molcellno = calloc((size_t)nmol, sizeof(int));
    cl->iacellneighb = calloc((size_t)cl->ncell, sizeof(int));
    cl->ibcellneighb = calloc((size_t)cl->ncell, sizeof(int));
    cl->iccellneighb = calloc((size_t)cl->ncell, sizeof(int));
    cl->maxcellno = 0;
    for (int im = 0; im < nmol; ++im) {
        int c = atomcellno(cl, xadd[im], yadd[im], zadd[im]);
        cl->molcellno[im] = c;
        cl->nmolsincell[c] += 1;
        if (cl->nmolsincell[c] > cl->maxcellno) cl->maxcellno = cl->nmolsincell[c];
    }
    cl->molsincell = calloc((size_t)cl->ncell * (cl->maxcellno > 0 ? cl->maxcellno : 1), sizeof(int));
    for (int c = 0; c < cl->ncell; ++c) cl->nmolsincell[c] = 0;
    for (int im = 0; im < nmol; ++im) {
        int c = cl->molcellno[im];
        cl->molsincell[(size_t)c * cl->maxcellno + cl->nmolsincell[c]] = im;
        cl->nmolsincell[c] += 1;
    }
    makecellneighbourlist(cl, rmax);
    return 0;
}

void ora_cells_free(ora_cells_t *cl)
{
    free(cl->nmolsincell); free(cl->molcellno); free(cl->iacellneighb);
    free(cl->ibcellneighb); free(cl->iccellneighb); free(cl->molsincell);
    memset(cl, 0, sizeof(*cl));
}

static int fold(int ix, int n) { while (ix < 0) ix += n; while (ix >= n) ix -= n; return ix; }

/* cellneighbourlist.py:199-219 (findiaibic :252-263, getcellreference :266-279) */
int ora_make_neighbour_list_from_cells(const ora_cells_t *cl, int c, int *nbref)
{
    int icc = c / cl->ncelldivsq;
    int ib = (c - icc * cl->ncelldivsq) / cl->ncelldiva;
    int ia = c - cl->ncelldiva * (ib + cl->ncelldivb * icc);
    int n = 0;
    for (int k = 0; k < cl->ncellneighb; ++k) {
        int a = fold(ia + cl->iacellneighb[k], cl->ncelldiva);
        int bb = fold(ib + cl->ibcellneighb[k], cl->ncelldivb);
        int cc = fold(icc + cl->iccellneighb[k], cl->ncelldivc);
        int cref = a + cl->ncelldiva * (bb + cl->ncelldivb * cc);
        for (int i = 0; i < cl->nmolsincell[cref]; ++i) nbref[n++] = cl->molsincell[(size_t)cref * cl->maxcellno + i];
    }
    return n;
}

/* cellneighbourlist.py:222-249 */
int ora_get_nnbrefmax(const ora_cells_t *cl)
{
    int best = 0;
    for (int c = 0; c < cl->ncell; ++c) {
        int icc = c / cl->ncelldivsq;
        int ib = (c - icc * cl->ncelldivsq) / cl->ncelldiva;
        int ia = c - cl->ncelldiva * (ib + cl->ncelldivb * icc);
        int n = 0;
        for (int k = 0; k < cl->ncellneighb; ++k) {
            int a = fold(ia + cl->iacellneighb[k], cl->ncelldiva);
            int bb = fold(ib + cl->ibcellneighb[k], cl->ncelldivb);
            int cc = fold(icc + cl->iccellneighb[k], cl->ncelldivc);
            n += cl->nmolsincell[a + cl->ncelldiva * (bb + cl->ncelldivb * cc)];
        }
        if (n > best) best = n;
    }
    return best;
}

/* ------------------------------------------------------------------ ledger */
void ora_ledger_free(ora_ledger_t *l)
{
    free(l->ergpair); free(l->molergpair); free(l->ergpairref); free(l->nergpairm);
    memset(l, 0, sizeof(*l));
}

/* ergpairreference.py:63-99 */
static void initialise_ergpair(ora_ledger_t *l, const ora_cells_t *cl, int nmol)
{
    l->nmol = nmol;
    l->nnbrefmax = ora_get_nnbrefmax(cl);
    l->nergpair = 0;
    /* the reference starts at 100*nmol and grows by 20% (reallocate_ergpair);
     * growth does not change any value, so the bound is allocated at once */
    l->nergpairmax = nmol + nmol * (l->nnbrefmax > 0 ? l->nnbrefmax : 1) / 2 + 16;
    l->ergpair = calloc((size_t)l->nergpairmax, sizeof(float));
    l->molergpair = calloc((size_t)nmol * (l->nnbrefmax > 0 ? l->nnbrefmax : 1), sizeof(int));
    l->ergpairref = calloc((size_t)nmol * (l->nnbrefmax > 0 ? l->nnbrefmax : 1), sizeof(int));
    l->nergpairm = calloc((size_t)nmol, sizeof(int));
}

/* ergpairreference.py:102-126 */
static void set_ergpair(ora_ledger_t *l, int im, int jm, float energy)
{
    const int M = l->nnbrefmax;
    if (l->nergpairm[im] < M && l->nergpairm[jm] < M && energy != 0.0f) {
        int e = l->nergpair++;
        l->ergpair[e] = energy;
        if (jm == im) {
            l->nergpairm[im] = 1;
            l->molergpair[(size_t)im * M + 0] = im;
            l->ergpairref[(size_t)im * M + 0] = e;
        } else {
            int pi = l->nergpairm[im]++, pj = l->nergpairm[jm]++;
            l->molergpair[(size_t)im * M + pi] = jm;
            l->molergpair[(size_t)jm * M + pj] = im;
            l->ergpairref[(size_t)im * M + pi] = e;
            l->ergpairref[(size_t)jm * M + pj] = e;
        }
    }
}

/* ato_manipulation_rev4.py:2374-2528 */
int ora_fmolec1(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t,
                const ora_cells_t *cl, const float *eintra, ora_ledger_t *l,
                float *ftest_out, float *ftestb_out, double *ftest_f64)
{
    initialise_ergpair(l, cl, b->nmol);
    float ftest = 0.0f, ftestb = 0.0f;
    double f64 = 0.0;
    for (int im = 0; im < b->nmol; ++im) {
        float energy = eintra ? eintra[im] : 0.0f;
        ftestb = ftestb + energy;
        set_ergpair(l, im, im, energy);
    }
    int *nbref = malloc(sizeof(int) * (size_t)(l->nnbrefmax > 0 ? l->nnbrefmax : 1));
    float *ergold = malloc(sizeof(float) * (size_t)(l->nnbrefmax > 0 ? l->nnbrefmax : 1));
    for (int cc = 0; cc < cl->ncell; ++cc) {
        if (cl->nmolsincell[cc] <= 0) continue;
        int nnbref = ora_make_neighbour_list_from_cells(cl, cc, nbref);
        for (int iref = 0; iref < cl->nmolsincell[cc]; ++iref) {
            int im = cl->molsincell[(size_t)cc * cl->maxcellno + iref];
            for (int inn = 0; inn < nnbref; ++inn) {
                int jm = nbref[inn];
                ergold[inn] = 0.0f;
                if (jm > im) ergold[inn] = ora_en_mole_pair(b, c, t, im, jm);
            }
            /* the store loop starts at the second neighbour (reference :2464-2475) */
            for (int inn = 1; inn < nnbref; ++inn) {
                int jm = nbref[inn];
                if (ergold[inn] != 0.0f && jm > im) {
                    set_ergpair(l, im, jm, ergold[inn]);
                    ftest = ftest + ergold[inn];
                    f64 += (double)ergold[inn];
                }
            }
        }
    }
    free(nbref);
    free(ergold);
    if (ftest_f64) *ftest_f64 = f64 / (double)b->nmol;
    *ftest_out = ftest / (float)b->nmol;
    *ftestb_out = ftestb / (float)b->nmol;
    return 0;
}

/* molfit.py:805-919 (no three-body, container or coordination terms) */
float ora_en_change(const ora_box_t *b, const ora_centres_t *c, const ora_tables_t *t,
                    const ora_ledger_t *l, int im, float *ergnew, int *nnbref_out)
{
    const int M = l->nnbrefmax;
    const float stepcl = b->cellst / b->dmax;
    int nnbref = l->nergpairm[im];
    float *ergold = malloc(sizeof(float) * (size_t)(nnbref > 0 ? nnbref : 1));
    for (int k = 0; k < nnbref; ++k) {
        ergold[k] = l->ergpair[l->ergpairref[(size_t)im * M + k]];
        ergnew[k] = ergold[k];
    }
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
    free(ergold);
    if (nnbref_out) *nnbref_out = nnbref;
    return fdiff;
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

/* molfit.py:732-759: the step is doubled until the move changes the position */
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

static float adapt_ratio(float rejrate, int att, int rej)
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

/* molfit.py:51-541: whole-molecule rotation + translation branch (nisum < intra,
 * no internal rotational groups, no tethered molecules). */
/* the splitmix64 uniform of delayed acceptance's stage 1 (the GPU's da_uniform) */
float ora_da_uniform(unsigned long long seed, int m)
{
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(m + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((float)(z >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

static float etest_of(float ediff)
{
    if (ediff <= 0.0f) return 10.0f;
    if (ediff < 20.0f) return ora_expf(-ediff);
    return 0.0f;
}

static int molfit_impl(const ora_box_t *b0, float *xa, float *ya, float *za, const ora_centres_t *c,
                       const ora_tables_t *t, const ora_cells_t *cl, ora_ledger_t *l,
                       ora_mcstate_t *mc, ora_ran1_t *rng, float *fadds, int *naccept_out,
                       const float *des, unsigned long long seed, int *nrej_out, double *target_out,
                       float *fdiff_out)
{
    ora_box_t bb = *b0;
    bb.xa = xa; bb.ya = ya; bb.za = za;
    const ora_box_t *b = &bb;
    const int M = l->nnbrefmax;
    float temp = 1.0f / mc->simt;
    for (int mt = 0; mt < b->nmoltype; ++mt) {
        mc->nwhole[2 * mt] = mc->nwhole[2 * mt + 1] = 0;
        mc->nmovem[2 * mt] = mc->nmovem[2 * mt + 1] = 0;
    }
    float trstepmax = 1.0f / b->natomr;
    trstepmax = ora_powf(trstepmax, 1.0f / 3.0f);
    float trstepmin = 0.0001f * trstepmax;
    if (trstepmin < 1.0e-6f) trstepmin = 1.0e-6f;
    const float rotstepmax = 1.0f, rotstepmin = 1.0e-5f;
    float fadd = 0.0f;
    int nacc = 0, nrej1 = 0;
    int maxat = 0;
    for (int im = 0; im < b->nmol; ++im)
        if (b->nal[im] - b->naf[im] + 1 > maxat) maxat = b->nal[im] - b->naf[im] + 1;
    float *xr = malloc(sizeof(float) * (size_t)maxat * 3);
    float *ergnew = malloc(sizeof(float) * (size_t)(M > 0 ? M : 1));
    /* chain position of every molecule (cells in order, molecules in a cell) */
    int *cpos = malloc(sizeof(int) * (size_t)b->nmol);
    for (int cc = 0, k = 0; cc < cl->ncell; ++cc)
        for (int iref = 0; iref < cl->nmolsincell[cc]; ++iref) cpos[cl->molsincell[(size_t)cc * cl->maxcellno + iref]] = k++;
    for (int cc = 0; cc < cl->ncell; ++cc) {
        for (int iref = 0; iref < cl->nmolsincell[cc]; ++iref) {
            int im = cl->molsincell[(size_t)cc * cl->maxcellno + iref];
            int n = b->nal[im] - b->naf[im] + 1;
            int mt = b->mtype[im];
            for (int k = 0; k < n; ++k) {
                xr[3 * k] = xa[b->naf[im] + k];
                xr[3 * k + 1] = ya[b->naf[im] + k];
                xr[3 * k + 2] = za[b->naf[im] + k];
            }
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
            int nnbref = 0;
            int accept;
            float fdiff = 0.0f;
            if (target_out) {
                /* exact energy change over the partners later in the chain */
                int nn2 = 0;
                (void)ora_en_change(b, c, t, l, im, ergnew, &nn2);
                double s = 0.0;
                for (int k = 1; k < nn2; ++k) {
                    int jm = l->molergpair[(size_t)im * M + k];
                    if (cpos[jm] > cpos[im])
                        s += (double)ergnew[k] - (double)l->ergpair[l->ergpairref[(size_t)im * M + k]];
                }
                target_out[im] = s;
            }
            if (fdiff_out) {
                int nn3 = 0;
                fdiff_out[im] = ora_en_change(b, c, t, l, im, ergnew, &nn3);
            }
            if (!des) {
                fdiff = ora_en_change(b, c, t, l, im, ergnew, &nnbref);
                float ediff = (fdiff + 0.0f) * temp;
                accept = ora_ran1(rng) < etest_of(ediff);
            } else {
                /* delayed acceptance: stage 1 on the surrogate's dEs, stage 2 on dE - dEs;
                 * the ran1 number of the Metropolis test is drawn either way */
                float d = des[im];
                int pass = ora_da_uniform(seed, im) < etest_of(d * temp);
                if (pass) {
                    fdiff = ora_en_change(b, c, t, l, im, ergnew, &nnbref);
                    accept = ora_ran1(rng) < etest_of((fdiff - d) * temp);
                } else {
                    (void)ora_ran1(rng);
                    accept = 0;
                    nrej1++;
                }
            }
            if (accept) {
                mc->nstepreject[im] = 0;
                fadd = fadd + fdiff;
                nacc++;
                /* save_energies + save_erg_pairs (ergpairreference.py:191-210) */
                for (int k = 0; k < nnbref; ++k) l->ergpair[l->ergpairref[(size_t)im * M + k]] = ergnew[k];
            } else {
                mc->nstepreject[im] += 1;
                for (int k = 0; k < n; ++k) {
                    xa[b->naf[im] + k] = xr[3 * k];
                    ya[b->naf[im] + k] = xr[3 * k + 1];
                    za[b->naf[im] + k] = xr[3 * k + 2];
                }
                c->xadd[im] = xs; c->yadd[im] = ys; c->zadd[im] = zs;
                if (mc->stepr > 0.0f) mc->nwhole[2 * mt + 1] += 1;
                if (mc->stepm > 0.0f) mc->nmovem[2 * mt + 1] += 1;
            }
        }
    }
    free(xr);
    free(ergnew);
    free(cpos);
    if (mc->ireset != 1) {
        for (int mt = 0; mt < b->nmoltype; ++mt) {
            float rat = adapt_ratio(mc->rejrate, mc->nwhole[2 * mt], mc->nwhole[2 * mt + 1]);
            mc->wholerot[mt] = mc->wholerot[mt] * rat;
            if (mc->wholerot[mt] > rotstepmax) mc->wholerot[mt] = rotstepmax;
            if (mc->wholerot[mt] < rotstepmin && mc->wholerot[mt] != 0.0f) mc->wholerot[mt] = rotstepmin;
            rat = adapt_ratio(mc->rejrate, mc->nmovem[2 * mt], mc->nmovem[2 * mt + 1]);
            mc->wholetrans[mt] = mc->wholetrans[mt] * rat;
            if (mc->wholetrans[mt] > trstepmax) mc->wholetrans[mt] = trstepmax;
            if (mc->wholetrans[mt] < trstepmin && mc->wholetrans[mt] != 0.0f) mc->wholetrans[mt] = trstepmin;
        }
    }
    *fadds = *fadds + fadd / (float)b->nmol;
    if (naccept_out) *naccept_out = nacc;
    if (nrej_out) *nrej_out = nrej1;
    return 0;
}

int ora_molfit(const ora_box_t *b, float *xa, float *ya, float *za, const ora_centres_t *c,
               const ora_tables_t *t, const ora_cells_t *cl, ora_ledger_t *l,
               ora_mcstate_t *mc, ora_ran1_t *rng, float *fadds, int *naccept_out)
{
    return molfit_impl(b, xa, ya, za, c, t, cl, l, mc, rng, fadds, naccept_out, NULL, 0ull, NULL, NULL, NULL);
}

int ora_molfit_da(const ora_box_t *b, float *xa, float *ya, float *za, const ora_centres_t *c,
                  const ora_tables_t *t, const ora_cells_t *cl, ora_ledger_t *l,
                  ora_mcstate_t *mc, ora_ran1_t *rng, float *fadds, int *naccept_out,
                  const float *des, unsigned long long seed, int *nrej_out, double *target_out,
                  float *fdiff_out)
{
    return molfit_impl(b, xa, ya, za, c, t, cl, l, mc, rng, fadds, naccept_out, des, seed, nrej_out, target_out,
                       fdiff_out);
}
