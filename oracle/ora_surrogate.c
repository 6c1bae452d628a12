/* CPU oracle of the learned Delta-E surrogate (TEST INFRASTRUCTURE ONLY).
 *
 * The surrogate has no counterpart in the reference; this file restates, in FP64, the
 * definitions the GPU implements (include/epsr_mi355x.h, "learned Delta-E surrogate"):
 * the radial descriptors of a molecule over its later ledger partners, and the
 * two-hidden-layer SiLU network with its loss and gradient.  Its checks: the descriptor
 * against an independent numpy evaluation and closed forms (tests/test_oracle_surrogate.py),
 * the gradient against central finite differences of the loss. */
#include <math.h>
#include <string.h>

#include "epsr_oracle.h"

static const double PI_D = 3.14159265358979323846;

/* minimum image of one component, as minimagec (cubic): x - nint(x) in float32 */
static float mic1(float x) { return x - (float)ora_nint(x); }

void ora_descriptor(int na, const float *xm, const float *ym, const float *zm, const int *tm, int nb,
                    const float *xn, const float *yn, const float *zn, const int *tn, const int *npairref,
                    int ntype, int npairs, float dmax, int ng, float rc, float r0, float dmu, float width,
                    float xscale, double *out)
{
    memset(out, 0, sizeof(double) * (size_t)npairs * ng);
    for (int a = 0; a < na; ++a)
        for (int b = 0; b < nb; ++b) {
            float dx = mic1(xn[b] - xm[a]), dy = mic1(yn[b] - ym[a]), dz = mic1(zn[b] - zm[a]);
            float r = ora_sqrtf(dx * dx + dy * dy + dz * dz) * dmax;   /* float32, as the GPU */
            if (!(r < rc)) continue;
            int ch = npairref[tm[a] * ntype + tn[b]];
            double fc = 0.5 * (cos(PI_D * (double)r / (double)rc) + 1.0);
            for (int g = 0; g < ng; ++g) {
                double t = ((double)r - ((double)r0 + g * (double)dmu)) / (double)width;
                out[ch * ng + g] += exp(-t * t) * fc;
            }
        }
    for (int i = 0; i < npairs * ng; ++i) out[i] *= (double)xscale;
}

/* float32 -> bfloat16 -> float32, round to nearest even */
static double bf(double v)
{
    float f = (float)v;
    unsigned u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return v;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double silu(double z) { return z / (1.0 + exp(-z)); }
static double dsilu(double z)
{
    double s = 1.0 / (1.0 + exp(-z));
    return s * (1.0 + z * (1.0 - s));
}

/* parameter layout [W1 (h x d) | b1 | W2 (h x h) | b2 | w3 | b3] */
typedef struct {
    const float *W1, *b1, *W2, *b2, *w3, *b3;
} net_t;

static net_t unpack(int d, int h, const float *w)
{
    net_t n;
    n.W1 = w;
    n.b1 = w + (size_t)h * d;
    n.W2 = n.b1 + h;
    n.b2 = n.W2 + (size_t)h * h;
    n.w3 = n.b2 + h;
    n.b3 = n.w3 + h;
    return n;
}

/* one row: z1, h1 (bf16 when asked), z2, h2, y */
static double row_forward(int d, int h, const net_t *n, int bfl, const float *x, double *z1, double *h1, double *z2,
                          double *h2)
{
    for (int j = 0; j < h; ++j) {
        double s = 0.0;
        for (int k = 0; k < d; ++k) {
            double wv = bfl ? bf(n->W1[(size_t)j * d + k]) : (double)n->W1[(size_t)j * d + k];
            double xv = bfl ? bf(x[k]) : (double)x[k];
            s += wv * xv;
        }
        z1[j] = s + n->b1[j];
        h1[j] = bfl ? bf(silu(z1[j])) : silu(z1[j]);
    }
    double y = 0.0;
    for (int j = 0; j < h; ++j) {
        double s = 0.0;
        for (int i = 0; i < h; ++i) {
            double wv = bfl ? bf(n->W2[(size_t)j * h + i]) : (double)n->W2[(size_t)j * h + i];
            s += wv * h1[i];
        }
        z2[j] = s + n->b2[j];
        h2[j] = silu(z2[j]);
        y += h2[j] * n->w3[j];
    }
    return y + n->b3[0];
}

void ora_mlp_forward(int d, int h, const float *w, int bfl, int nrows, const float *x, int ldx, double *y)
{
    net_t n = unpack(d, h, w);
    double z1[512], h1[512], z2[512], h2[512];
    if (h > 512) return;
    for (int r = 0; r < nrows; ++r) y[r] = row_forward(d, h, &n, bfl, x + (size_t)r * ldx, z1, h1, z2, h2);
}

double ora_mlp_grad(int d, int h, const float *w, int bfl, int npair, const float *x, int ldx, const float *target,
                    double *grad)
{
    net_t n = unpack(d, h, w);
    const size_t np = (size_t)h * d + h + (size_t)h * h + h + h + 1;
    memset(grad, 0, sizeof(double) * np);
    double *gW1 = grad, *gb1 = grad + (size_t)h * d, *gW2 = gb1 + h, *gb2 = gW2 + (size_t)h * h, *gw3 = gb2 + h;
    double z1[2][512], h1[2][512], z2[2][512], h2[2][512], y[2];
    double dz2[512], dh1[512];
    if (h > 512) return NAN;
    double loss = 0.0;
    for (int p = 0; p < npair; ++p) {
        for (int s = 0; s < 2; ++s)
            y[s] = row_forward(d, h, &n, bfl, x + (size_t)(2 * p + s) * ldx, z1[s], h1[s], z2[s], h2[s]);
        double e = (y[1] - y[0]) - (double)target[p];
        loss += e * e;
        double g = 2.0 * e / npair;
        for (int s = 0; s < 2; ++s) {
            double dy = s ? g : -g;
            const float *xr = x + (size_t)(2 * p + s) * ldx;
            for (int j = 0; j < h; ++j) {
                gw3[j] += dy * h2[s][j];
                dz2[j] = dy * n.w3[j] * dsilu(z2[s][j]);
                gb2[j] += dz2[j];
            }
            for (int i = 0; i < h; ++i) {
                double acc = 0.0;
                for (int j = 0; j < h; ++j) {
                    double wv = bfl ? bf(n.W2[(size_t)j * h + i]) : (double)n.W2[(size_t)j * h + i];
                    gW2[(size_t)j * h + i] += dz2[j] * h1[s][i];
                    acc += dz2[j] * wv;
                }
                dh1[i] = acc * dsilu(z1[s][i]);
                gb1[i] += dh1[i];
            }
            for (int j = 0; j < h; ++j)
                for (int k = 0; k < d; ++k) gW1[(size_t)j * d + k] += dh1[j] * (bfl ? bf(xr[k]) : (double)xr[k]);
        }
    }
    return loss / npair;
}
