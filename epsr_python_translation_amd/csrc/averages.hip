// The running averages of S(Q) and g(r) with their error bars, sumsq and sumgr
// (src/epsr26/programs/epsr_standalone_rev3.py:3642-3685, 3578-3637).
//
// Every (row, point) element is independent: one lane reads its input and the two
// float32 accumulators, updates them in double (the reference's real*8 work variables,
// no contraction) and writes them back -- 12 bytes in, 8 out, HBM-bound.  sumgr's zero
// limit is a double running sum along r for each pair, which the reference evaluates in
// order; one lane per pair walks it after the update.
#include "common.h"

namespace {

// correctly rounded double square root: the hardware result corrected by one ulp where
// the exact residual says a neighbour is closer
__device__ __forceinline__ double sqrt_rn_f64(double x)
{
    double r = __builtin_sqrt(x);
    if (!(x > 0.0) || !(x < __builtin_inf())) return r;
    // r > 0: its neighbours are one step of the bit pattern either way
    const long long br = __double_as_longlong(r);
    double lo = __longlong_as_double(br - 1), hi = __longlong_as_double(br + 1);
    // |c^2 - x| compared through fma residuals (exact for these magnitudes)
    double er = __builtin_fabs(__builtin_fma(r, r, -x));
    double el = __builtin_fabs(__builtin_fma(lo, lo, -x));
    double eh = __builtin_fabs(__builtin_fma(hi, hi, -x));
    if (el < er) return lo;
    if (eh < er) return hi;
    return r;
}

__device__ __forceinline__ void accumulate(double add1, double fc1, double fc2, float *avg, float *err)
{
    double sumtot1 = (double)*avg, sumtot2 = (double)*err;
    const double add2 = add1 * add1;
    sumtot2 = sumtot2 * sumtot2 + sumtot1 * sumtot1;
    sumtot1 = fc2 * sumtot1 + fc1 * add1;
    sumtot2 = fc2 * sumtot2 + fc1 * add2;
    double errort = sumtot2 - sumtot1 * sumtot1;
    errort = errort > 0.0 ? sqrt_rn_f64(errort) : 0.0;
    *avg = (float)sumtot1;
    *err = (float)errort;
}

// rows = 2*npairs, points jlo..jhi of each row (stride ld); ycoh null: add1 = x
__global__ void running_average(int rows, int jlo, int npts, int ld, const float *__restrict__ x,
                                const float *__restrict__ ycoh, double fc1, double fc2, float *__restrict__ avg,
                                float *__restrict__ err)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)rows * npts) return;
    const int row = (int)(t / npts), j = jlo + (int)(t - (long)row * npts);
    const size_t o = (size_t)row * ld + j;
    double add1;
    if (ycoh) {
        const float y = ycoh[j];
        add1 = y > 0.0f ? (double)(x[o] / y) : 0.0;
    } else {
        add1 = (double)x[o];
    }
    accumulate(add1, fc1, fc2, avg + o, err + o);
}

// sumgr's zero limit of pair ic: entsig(j) = sum over j' = 2..j of ycoh (g_inter - 1).
// One workgroup per pair: the float32 terms are formed in parallel into LDS, one lane
// then runs the double running sum in order over LDS, and the results go out in parallel.
constexpr int ZL_CHUNK = 4096;
__global__ void __launch_bounds__(256)
zero_limit(int npairs, int nz, int ld, const float *__restrict__ ycoh, const float *__restrict__ gor,
           float *__restrict__ entsig)
{
    __shared__ float term[ZL_CHUNK];
    __shared__ double carry;
    const int ic = blockIdx.x;
    const float *g = gor + ((size_t)npairs + ic) * ld;
    float *e = entsig + (size_t)ic * ld;
    if (threadIdx.x == 0) {
        e[1] = 0.0f;
        carry = 0.0;
    }
    for (int j0 = 2; j0 <= nz; j0 += ZL_CHUNK) {
        const int n = min(ZL_CHUNK, nz - j0 + 1);
        for (int k = threadIdx.x; k < n; k += blockDim.x) term[k] = ycoh[j0 + k] * (g[j0 + k] - 1.0f);
        __syncthreads();
        if (threadIdx.x == 0) {
            double z = carry;
            for (int k = 0; k < n; ++k) {
                z = z + (double)term[k];
                term[k] = (float)z;
            }
            carry = z;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += blockDim.x) e[j0 + k] = term[k];
        __syncthreads();
    }
}

}  // namespace

extern "C" int epsr_sumsq(epsr_ctx *ctx, int npairs, int nq, int ldq, const float *qwind, double fc1, double fc2,
                          float *soq, float *sq2)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, npairs > 0 && nq >= 1 && ldq > nq && qwind && soq && sq2, "epsr_sumsq: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const long n = 2L * npairs * nq;
    running_average<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(2 * npairs, 1, nq, ldq, qwind, nullptr,
                                                                          fc1, fc2, soq, sq2);
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}

extern "C" int epsr_sumgr(epsr_ctx *ctx, int npairs, int nz, int ldw, const float *wav, const float *ycoh,
                          double fc1, double fc2, float *gor, float *gr2, float *entsig)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, npairs > 0 && nz >= 2 && ldw > nz && wav && ycoh && gor && gr2 && entsig,
              "epsr_sumgr: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const long n = 2L * npairs * (nz - 1);
    running_average<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(2 * npairs, 2, nz - 1, ldw, wav, ycoh, fc1,
                                                                          fc2, gor, gr2);
    zero_limit<<<npairs, 256, 0, ctx->stream>>>(npairs, nz, ldw, ycoh, gor, entsig);
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}
