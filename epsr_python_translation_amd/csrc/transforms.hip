// r <-> Q transforms: gensq (histograms -> partial S(Q)) and stog (S(Q) -> g(r)).
//
// Two modes each:
//  mode 0  one lane per output element, summing in the reference's order in float32
//          (gensq is then bit-identical to the reference; stog is bit-identical to
//          the reference evaluated with a correctly rounded sinf);
//  mode 1  the same contraction as an FP32 MFMA GEMM (v_mfma_f32_16x16x4_f32) with
//          FP32 accumulation; the right-hand operand is assembled on the fly.
#include <cstring>
#include <mutex>
#include <vector>

#include "common.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ gensq, mode 0
// src/epsr26/programs/epsr_standalone_rev3.py:4730-4802.  One thread per (Q point i,
// column): columns 0..npairs-1 are the intramolecular sums, npairs..2npairs-1 the
// intermolecular ones.
__global__ void gensq_exact(int nq, int npairs, int npt1, int ldw, const float *__restrict__ wav,
                            const float *__restrict__ ycoh, const float *__restrict__ foqols, int ldf,
                            const float *__restrict__ natomsphere, float *__restrict__ sofq, int ldq)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int ncol = 2 * npairs;
    if (t >= nq * ncol) return;
    int i = t / ncol + 1;
    int col = t - (i - 1) * ncol;
    const float *fq = foqols + (size_t)i * ldf;
    if (col < npairs) {
        const float *w1 = wav + (size_t)col * ldw;
        float sum1 = 0.0f;
        for (int j = 2; j <= npt1; ++j) {
            float a = w1[j];
            if (fabsf(a) > 0.0f) sum1 = sum1 + a * fq[j];
        }
        sofq[(size_t)col * ldq + i] = sum1;
    } else {
        int ic = col - npairs;
        const float *w2 = wav + ((size_t)npairs + ic) * ldw;
        float nas = natomsphere ? natomsphere[ic] : 1.0f;
        float sum2 = 0.0f, sum3 = 0.0f;
        for (int j = 2; j <= npt1; ++j) {
            float w = w2[j];
            float add2 = (w - ycoh[j]) * nas;
            sum2 = sum2 + add2 * fq[j];
            sum3 = sum3 + w;
        }
        if (sum3 <= 0.0f) sum2 = 0.0f;
        sofq[((size_t)npairs + ic) * ldq + i] = sum2;
    }
}

// ------------------------------------------------------------------ MFMA GEMM core
// One wave computes a 16 x 16 tile of C = A . B over k in [k0, k1), A row-major
// (element (m,k) at A[m*lda + k]) and B given by a functor b(k, n).  Lane l holds
// A(m0 + l%16, k + l/16) and B(k + l/16, n0 + l%16); the result element
// (m0 + 4*(l/16) + v, n0 + l%16) is in acc[v].
template <class BF>
__device__ __forceinline__ floatx4 mfma_tile_16x16(const float *__restrict__ A, int lda, int M, int m0, int n0,
                                                   int k0, int k1, BF b)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = lane >> 4;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    const int m = m0 + r;
    const bool mrow = m < M;
    for (int k = k0; k < k1; k += 4) {
        int kk = k + kq;
        float a = (mrow && kk < k1) ? A[(size_t)m * lda + kk] : 0.0f;
        float bv = (kk < k1) ? b(kk, n0 + r) : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv, acc, 0, 0, 0);
    }
    return acc;
}

// gensq, mode 1.  B(j, col): intramolecular wav for col < npairs, (wav2 - ycoh) *
// natomsphere for the intermolecular columns; rows j = 2..npt1.  grid.x: 16-row tiles of Q
// points, grid.y: 16-column tiles, grid.z: GENSQ_KSPLIT slices of the r rows (a handful of
// columns gives too few tiles to fill the GPU otherwise); one wave per workgroup writes its
// slice's partial tile, and gensq_reduce adds the slices in a fixed order.
constexpr int GENSQ_KSPLIT = 16;
__global__ void __launch_bounds__(64) gensq_mfma(int nq, int npairs, int npt1, int ldw, const float *__restrict__ wav,
                                                 const float *__restrict__ ycoh, const float *__restrict__ foqols,
                                                 int ldf, const float *__restrict__ natomsphere,
                                                 float *__restrict__ part)
{
    const int ncol = 2 * npairs, ncolp = (int)gridDim.y * 16;
    const int m0 = blockIdx.x * 16;   // Q rows, 0-based: Q index = m + 1
    const int n0 = blockIdx.y * 16;
    const int kz = blockIdx.z;
    auto bfun = [&](int j, int col) -> float {
        if (col >= ncol) return 0.0f;
        if (col < npairs) return wav[(size_t)col * ldw + j];
        int ic = col - npairs;
        float nas = natomsphere ? natomsphere[ic] : 1.0f;
        return (wav[((size_t)npairs + ic) * ldw + j] - ycoh[j]) * nas;
    };
    // A rows start at foqols + ldf (Q index 1); columns j are direct indices 2..npt1
    const int K0 = 2, K1 = npt1 + 1;
    const int span = (((K1 - K0) + (int)gridDim.z - 1) / (int)gridDim.z + 3) & ~3;
    const int kb = K0 + kz * span, ke = min(K1, kb + span);
    floatx4 acc = mfma_tile_16x16(foqols + ldf, ldf, nq, m0, n0, kb, ke, bfun);
    const int lane = threadIdx.x & 63;
    const int col = n0 + (lane & 15);
    for (int v = 0; v < 4; ++v) {
        int m = m0 + 4 * (lane >> 4) + v;
        if (m < nq) part[((size_t)kz * ncolp + col) * nq + m] = acc[v];
    }
}

__global__ void gensq_reduce(int nq, int npairs, int ncolp, int ksplit, const float *__restrict__ part,
                             const float *__restrict__ colsum, float *__restrict__ sofq, int ldq)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2L * npairs * nq) return;
    const int col = (int)(t / nq), m = (int)(t - (long)col * nq);
    float sum = 0.0f;
    for (int kz = 0; kz < ksplit; ++kz) sum = sum + part[((size_t)kz * ncolp + col) * nq + m];
    const bool zero = col >= npairs && colsum[col - npairs] <= 0.0f;
    sofq[(size_t)col * ldq + m + 1] = zero ? 0.0f : sum;
}

// sum3 of gensq (:4785-4793) is only tested for <= 0.  The intermolecular histogram is
// non-negative, so the sequential float32 sum is <= 0 exactly when every term is zero,
// which a parallel maximum decides in any order.  One workgroup per pair.
__global__ void gensq_colsum(int npairs, int npt1, int ldw, const float *__restrict__ wav, float *__restrict__ out)
{
    __shared__ float s[256];
    int ic = blockIdx.x;
    const float *w2 = wav + ((size_t)npairs + ic) * ldw;
    float m = 0.0f;
    for (int j = 2 + threadIdx.x; j <= npt1; j += blockDim.x) m = fmaxf(m, w2[j]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[ic] = s[0];
}

// scratch of mode 1: column sums (256-byte aligned) then GENSQ_KSPLIT partial tiles
static size_t gensq_scratch_bytes(int nq, int npairs)
{
    const int ncolp = 16 * ceil_div(2 * npairs, 16);
    const size_t csb = (sizeof(float) * npairs + 255) & ~(size_t)255;
    return csb + sizeof(float) * (size_t)GENSQ_KSPLIT * ncolp * nq;
}

// mode 1 on stream st with its scratch at sc (gensq_scratch_bytes)
static void gensq_mfma_launch(hipStream_t st, int nq, int npairs, int npt1, int ldw, const float *wav,
                              const float *ycoh, const float *foqols, int ldf, const float *natomsphere, float *sofq,
                              int ldq, char *sc)
{
    const int ncolp = 16 * ceil_div(2 * npairs, 16);
    const size_t csb = (sizeof(float) * npairs + 255) & ~(size_t)255;
    float *cs = (float *)sc, *part = (float *)(sc + csb);
    gensq_colsum<<<npairs, 256, 0, st>>>(npairs, npt1, ldw, wav, cs);
    dim3 grid(ceil_div(nq, 16), ncolp / 16, GENSQ_KSPLIT);
    gensq_mfma<<<grid, 64, 0, st>>>(nq, npairs, npt1, ldw, wav, ycoh, foqols, ldf, natomsphere, part);
    gensq_reduce<<<ceil_div(2L * npairs * nq, 256), 256, 0, st>>>(nq, npairs, ncolp, GENSQ_KSPLIT, part, cs, sofq,
                                                                  ldq);
}

extern "C" int epsr_gensq(epsr_ctx *ctx, int nq, int npairs, int npt1, int ldw, const float *wav, const float *ycoh,
                          const float *foqols, int ldf, const float *natomsphere, float *sofq, int ldq, int mode)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, wav && ycoh && foqols && sofq, "epsr_gensq: null argument");
    ARG_CHECK(ctx, nq >= 1 && npairs >= 1 && npt1 >= 2 && ldw > npt1 && ldf > npt1 && ldq > nq,
              "epsr_gensq: bad sizes nq=%d npairs=%d npt1=%d ldw=%d ldf=%d ldq=%d", nq, npairs, npt1, ldw, ldf, ldq);
    ARG_CHECK(ctx, mode == 0 || mode == 1, "epsr_gensq: mode must be 0 or 1");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    timer_begin(ctx, "gensq");
    if (mode == 0) {
        int n = nq * 2 * npairs;
        gensq_exact<<<ceil_div(n, 64), 64, 0, st>>>(nq, npairs, npt1, ldw, wav, ycoh, foqols, ldf, natomsphere, sofq, ldq);
    } else {
        char *sc = (char *)ctx_scratch(ctx, gensq_scratch_bytes(nq, npairs));
        if (!sc) return EPSR_ERR_HIP;
        gensq_mfma_launch(st, nq, npairs, npt1, ldw, wav, ycoh, foqols, ldf, natomsphere, sofq, ldq, sc);
    }
    timer_end(ctx, "gensq");
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, ctx_sync(ctx));
    return EPSR_OK;
}

// ------------------------------------------------------------------ gensq without allocation
// The scratch of mode 1 (column sums and the GENSQ_KSPLIT partial tiles) in a file-scope
// device array, so that a call allocates nothing.  Each GPU holds its own copy at its own
// address, found once per device with hipGetSymbolAddress; host code uses only that address.
#define GENSQ_NOALLOC_BYTES ((size_t)8 << 20)
static __device__ __align__(256) char scratch[GENSQ_NOALLOC_BYTES];

static std::mutex scratch_mu;
static void *scratch_dev[256];   // the array's address on each device (nullptr: not looked up yet)

static int scratch_address(epsr_ctx *ctx, void **out)
{
    std::lock_guard<std::mutex> g(scratch_mu);
    ARG_CHECK(ctx, ctx->device >= 0 && ctx->device < 256, "epsr_gensq_noalloc: device %d", ctx->device);
    if (!scratch_dev[ctx->device]) {
        void *p = nullptr;
        HIP_CHECK(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ARG_CHECK(ctx, p != nullptr, "epsr_gensq_noalloc: no address for the scratch array");
        scratch_dev[ctx->device] = p;
    }
    *out = scratch_dev[ctx->device];
    return EPSR_OK;
}

extern "C" int epsr_gensq_noalloc(epsr_ctx *ctx, int nq, int npairs, int npt1, int ldw, const float *wav,
                                  const float *ycoh, const float *foqols, int ldf, const float *natomsphere,
                                  float *sofq, int ldq)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, wav && ycoh && foqols && sofq, "epsr_gensq_noalloc: null argument");
    ARG_CHECK(ctx, nq >= 1 && npairs >= 1 && npt1 >= 2 && ldw > npt1 && ldf > npt1 && ldq > nq,
              "epsr_gensq_noalloc: bad sizes nq=%d npairs=%d npt1=%d ldw=%d ldf=%d ldq=%d", nq, npairs, npt1, ldw,
              ldf, ldq);
    const size_t need = gensq_scratch_bytes(nq, npairs);
    ARG_CHECK(ctx, need <= GENSQ_NOALLOC_BYTES,
              "epsr_gensq_noalloc: nq=%d x npairs=%d needs %zu bytes of scratch, above its %zu", nq, npairs, need,
              GENSQ_NOALLOC_BYTES);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    void *sc = nullptr;
    int rc = scratch_address(ctx, &sc);
    if (rc) return rc;
    // every value the kernels read from the scratch is written first in the same call
    gensq_mfma_launch(ctx->stream, nq, npairs, npt1, ldw, wav, ycoh, foqols, ldf, natomsphere, sofq, ldq, (char *)sc);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, ctx_sync(ctx));
    return EPSR_OK;
}

// ------------------------------------------------------------------ stog
// correctly rounded single-precision sine: double evaluation, one rounding
__device__ __forceinline__ float cr_sinf(float x) { return (float)sin((double)x); }

// ato_manipulation_rev4.py:4114-4124
__device__ __forceinline__ float dev_lorch(float x, float pi)
{
    if (x == 0.0f) return 1.0f;
    if (x < pi) return cr_sinf(x) / x;
    return 0.0f;
}

// sfac = 0.5*qstep/pi/pi/rho with qstep = q(4) - q(3) (:4076-4077), on the device
__device__ __forceinline__ float stog_sfac(const float *__restrict__ q, float rho)
{
    const float pi = 3.14159265358979323846f;
    const float qstep = q[4] - q[3];
    return 0.5f * qstep / pi / pi / rho;
}

// ato_manipulation_rev4.py:4066-4111, one thread per (r point j, column in)
__global__ void stog_exact(float rho, const float *__restrict__ q, const float *__restrict__ sofq, int ldq,
                           const float *__restrict__ r, float *__restrict__ gofr, int ldg,
                           const int *__restrict__ nqmin, const int *__restrict__ nqmax, int nr, int ncols,
                           const float *__restrict__ xwt)
{
    const float pi = 3.14159265358979323846f;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nr * ncols) return;
    int in = t / nr;
    int j = t - in * nr + 1;
    const float sfac = stog_sfac(q, rho);
    float rr = r[j];
    float qmax = pi / q[nqmax[in]];
    const float *sq = sofq + (size_t)in * ldq;
    float sumi = 0.0f;
    for (int i = nqmin[in]; i <= nqmax[in]; ++i) {
        float qq = q[i];
        float prod;
        if (rr > 0.0f) prod = qq * cr_sinf(qq * rr) * dev_lorch(qmax * qq, pi);
        else prod = qq * qq * dev_lorch(qmax * qq, pi);
        sumi = sumi + sq[i] * prod;
    }
    gofr[(size_t)in * ldg + j] = rr > 0.0f ? xwt[in] + sfac * sumi / rr : xwt[in] + sfac * sumi;
}

// B(i, in) = sofq(i, in) * lorch(pi q_i / q_nqmax(in)) on [nqmin, nqmax], else 0
__global__ void stog_prepare_b(const float *__restrict__ q, const float *__restrict__ sofq, int ldq,
                               const int *__restrict__ nqmin, const int *__restrict__ nqmax, int nqpad, int ncols,
                               float *__restrict__ B)
{
    const float pi = 3.14159265358979323846f;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nqpad * ncols) return;
    int in = t / nqpad;
    int i = t - in * nqpad;
    float v = 0.0f;
    if (i >= nqmin[in] && i <= nqmax[in]) {
        float qmax = pi / q[nqmax[in]];
        v = sofq[(size_t)in * ldq + i] * dev_lorch(qmax * q[i], pi);
    }
    B[(size_t)in * nqpad + i] = v;
}

// stog, mode 1: A(j, i) = q_i sin(q_i r_j) (q_i^2 at r = 0) generated in registers.
// grid.z slices of the Q points (STOG_KSPLIT): each wave writes its slice's partial
// tile; stog_reduce adds the slices in a fixed order and applies rho, the 1/r and xwt
constexpr int STOG_KSPLIT = 8;
__global__ void __launch_bounds__(64) stog_mfma(const float *__restrict__ q, int nqpad, const float *__restrict__ B,
                                                const float *__restrict__ r, int nr, int ncols,
                                                float *__restrict__ part)
{
    const int lane = threadIdx.x & 63;
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16, kz = blockIdx.z;
    const int ncolp = (int)gridDim.y * 16;
    const int rl = lane & 15, kq = lane >> 4;
    const int jrow = m0 + rl + 1;   // r index (1-based)
    const float rr = jrow <= nr ? r[jrow] : 0.0f;
    const int ncol = n0 + rl;
    const int span = ((nqpad + (int)gridDim.z - 1) / (int)gridDim.z + 3) & ~3;
    const int kb = kz * span, ke = min(nqpad, kb + span);
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k = kb; k < ke; k += 4) {
        int i = k + kq;
        float qq = i < ke ? q[i] : 0.0f;
        float a = 0.0f;
        if (jrow <= nr && i < ke && i >= 1) a = rr > 0.0f ? qq * sinf(qq * rr) : qq * qq;
        float bv = (ncol < ncols && i < ke) ? B[(size_t)ncol * nqpad + i] : 0.0f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv, acc, 0, 0, 0);
    }
    const int col = n0 + (lane & 15);
    for (int v = 0; v < 4; ++v) {
        int j = m0 + 4 * kq + v;   // 0-based row
        if (j < nr) part[((size_t)kz * ncolp + col) * nr + j] = acc[v];
    }
}

__global__ void stog_reduce(float rho, const float *__restrict__ q, int nr, int ncols, int ncolp, int ksplit,
                            const float *__restrict__ part, const float *__restrict__ r, float *__restrict__ gofr,
                            int ldg, const float *__restrict__ xwt)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncols * nr) return;
    const int col = (int)(t / nr), j0 = (int)(t - (long)col * nr);
    float sum = 0.0f;
    for (int kz = 0; kz < ksplit; ++kz) sum = sum + part[((size_t)kz * ncolp + col) * nr + j0];
    const float sfac = stog_sfac(q, rho);
    const int j = j0 + 1;
    const float rj = r[j];
    gofr[(size_t)col * ldg + j] = rj > 0.0f ? xwt[col] + sfac * sum / rj : xwt[col] + sfac * sum;
}

extern "C" int epsr_stog(epsr_ctx *ctx, float rho, const float *q, const float *sofq, int ldq, const float *r,
                         float *gofr, int ldg, const int *nqmin, const int *nqmax, int nr, int ncols, const float *xwt,
                         int mode)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, q && sofq && r && gofr && nqmin && nqmax && xwt, "epsr_stog: null argument");
    ARG_CHECK(ctx, nr >= 1 && ncols >= 1 && ldg > nr, "epsr_stog: bad sizes nr=%d ncols=%d ldg=%d", nr, ncols, ldg);
    ARG_CHECK(ctx, mode == 0 || mode == 1, "epsr_stog: mode must be 0 or 1");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    timer_begin(ctx, "stog");
    if (mode == 0) {
        int n = nr * ncols;
        stog_exact<<<ceil_div(n, 64), 64, 0, st>>>(rho, q, sofq, ldq, r, gofr, ldg, nqmin, nqmax, nr, ncols, xwt);
    } else {
        // the Q range used is 1..ldq-1 (nqmax <= ldq-1 is the caller's contract)
        int nqpad = ldq;
        const int ncolp = 16 * ceil_div(ncols, 16);
        const size_t bb = (sizeof(float) * (size_t)nqpad * ncols + 255) & ~(size_t)255;
        char *sc = (char *)ctx_scratch(ctx, bb + sizeof(float) * (size_t)STOG_KSPLIT * ncolp * nr);
        if (!sc) return EPSR_ERR_HIP;
        float *B = (float *)sc, *part = (float *)(sc + bb);
        stog_prepare_b<<<ceil_div(nqpad * ncols, 256), 256, 0, st>>>(q, sofq, ldq, nqmin, nqmax, nqpad, ncols, B);
        dim3 grid(ceil_div(nr, 16), ncolp / 16, STOG_KSPLIT);
        stog_mfma<<<grid, 64, 0, st>>>(q, nqpad, B, r, nr, ncols, part);
        stog_reduce<<<ceil_div((long)ncols * nr, 256), 256, 0, st>>>(rho, q, nr, ncols, ncolp, STOG_KSPLIT, part, r,
                                                                     gofr, ldg, xwt);
    }
    timer_end(ctx, "stog");
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, ctx_sync(ctx));
    return EPSR_OK;
}

// ------------------------------------------------------------------ batched transforms
// gensq and stog of many boxes on the same grids (replicas of one system) as one GEMM each:
// the boxes' columns side by side in the right-hand operand, so the transform matrix (A)
// is read once for all of them.  A result element depends only on its row of A, its
// column of B and the k-slices, which are the single-box mode 1's, so every box's output
// equals what epsr_gensq / epsr_stog mode 1 give it alone, bit for bit.

struct GensqBoxDev {
    const float *wav, *natomsphere;
    float *sofq, *colsum;
    int npairs, ldw, ldq, col0;   // col0: the box's first column in the batch
};

__global__ void __launch_bounds__(64) gensq_mfma_batch(int nq, int npt1, const float *__restrict__ ycoh,
                                                       const float *__restrict__ foqols, int ldf,
                                                       const GensqBoxDev *__restrict__ bx,
                                                       const int *__restrict__ colbox, int ncol,
                                                       float *__restrict__ part)
{
    const int ncolp = (int)gridDim.y * 16;
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16, kz = blockIdx.z;
    auto bfun = [&](int j, int col) -> float {
        if (col >= ncol) return 0.0f;
        const GensqBoxDev &b = bx[colbox[col]];
        const int c = col - b.col0;
        if (c < b.npairs) return b.wav[(size_t)c * b.ldw + j];
        const int ic = c - b.npairs;
        const float nas = b.natomsphere ? b.natomsphere[ic] : 1.0f;
        return (b.wav[((size_t)b.npairs + ic) * b.ldw + j] - ycoh[j]) * nas;
    };
    const int K0 = 2, K1 = npt1 + 1;
    const int span = (((K1 - K0) + (int)gridDim.z - 1) / (int)gridDim.z + 3) & ~3;
    const int kb = K0 + kz * span, ke = min(K1, kb + span);
    floatx4 acc = mfma_tile_16x16(foqols + ldf, ldf, nq, m0, n0, kb, ke, bfun);
    const int lane = threadIdx.x & 63;
    const int col = n0 + (lane & 15);
    for (int v = 0; v < 4; ++v) {
        int m = m0 + 4 * (lane >> 4) + v;
        if (m < nq) part[((size_t)kz * ncolp + col) * nq + m] = acc[v];
    }
}

// gensq_colsum over every box of a batch: one workgroup per intermolecular column
__global__ void gensq_colsum_batch(int npt1, const GensqBoxDev *__restrict__ bx, const int *__restrict__ colbox)
{
    __shared__ float s[256];
    const int col = blockIdx.x;
    const GensqBoxDev &b = bx[colbox[col]];
    const int c = col - b.col0;
    if (c < b.npairs) return;   // (uniform over the workgroup)
    const int ic = c - b.npairs;
    const float *w2 = b.wav + ((size_t)b.npairs + ic) * b.ldw;
    float m = 0.0f;
    for (int j = 2 + threadIdx.x; j <= npt1; j += blockDim.x) m = fmaxf(m, w2[j]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmaxf(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) b.colsum[ic] = s[0];
}

__global__ void gensq_reduce_batch(int nq, int ncol, int ncolp, int ksplit, const float *__restrict__ part,
                                   const GensqBoxDev *__restrict__ bx, const int *__restrict__ colbox)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncol * nq) return;
    const int col = (int)(t / nq), m = (int)(t - (long)col * nq);
    float sum = 0.0f;
    for (int kz = 0; kz < ksplit; ++kz) sum = sum + part[((size_t)kz * ncolp + col) * nq + m];
    const GensqBoxDev &b = bx[colbox[col]];
    const int c = col - b.col0;
    const bool zero = c >= b.npairs && b.colsum[c - b.npairs] <= 0.0f;
    b.sofq[(size_t)c * b.ldq + m + 1] = zero ? 0.0f : sum;
}

// the small per-box tables of a batch in the context's scratch: the descriptors, then the
// box of every column (uploaded from pinned memory on the context's stream)
static int batch_tables(epsr_ctx *ctx, const void *desc, size_t desc_bytes, const std::vector<int> &colbox,
                        char *dev)
{
    const size_t cb = sizeof(int) * colbox.size();
    char *h = (char *)ctx_pinned(ctx, desc_bytes + cb);
    if (!h) return EPSR_ERR_HIP;
    std::memcpy(h, desc, desc_bytes);
    std::memcpy(h + desc_bytes, colbox.data(), cb);
    HIP_CHECK(ctx, hipMemcpyAsync(dev, h, desc_bytes + cb, hipMemcpyHostToDevice, ctx->stream));
    return EPSR_OK;
}

extern "C" int epsr_gensq_batch(epsr_ctx *ctx, int nq, int npt1, const float *ycoh, const float *foqols, int ldf,
                                int nbox, const epsr_gensq_box_t *boxes)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, ycoh && foqols && boxes && nbox >= 1, "epsr_gensq_batch: null argument or no box");
    ARG_CHECK(ctx, !ctx_capturing(ctx), "epsr_gensq_batch: not for graph capture (it uploads its box tables from host memory)");
    ARG_CHECK(ctx, nq >= 1 && npt1 >= 2 && ldf > npt1, "epsr_gensq_batch: bad sizes nq=%d npt1=%d ldf=%d", nq, npt1,
              ldf);
    std::vector<GensqBoxDev> bd(nbox);
    std::vector<int> colbox;
    int ncol = 0;
    for (int b = 0; b < nbox; ++b) {
        const epsr_gensq_box_t &x = boxes[b];
        ARG_CHECK(ctx, x.wav && x.sofq, "epsr_gensq_batch: box %d has a null array", b);
        ARG_CHECK(ctx, x.npairs >= 1 && x.ldw > npt1 && x.ldq > nq, "epsr_gensq_batch: box %d: npairs=%d ldw=%d ldq=%d",
                  b, x.npairs, x.ldw, x.ldq);
        bd[b] = GensqBoxDev{x.wav, x.natomsphere, x.sofq, nullptr, x.npairs, x.ldw, x.ldq, ncol};
        for (int c = 0; c < 2 * x.npairs; ++c) colbox.push_back(b);
        ncol += 2 * x.npairs;
    }
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int ncolp = 16 * ceil_div(ncol, 16);
    // scratch: column sums of every box (256-byte aligned each), the tables, the partial tiles
    size_t off = 0;
    std::vector<size_t> csoff(nbox);
    for (int b = 0; b < nbox; ++b) {
        csoff[b] = off;
        off += (sizeof(float) * bd[b].npairs + 255) & ~(size_t)255;
    }
    const size_t tab = off, descb = sizeof(GensqBoxDev) * nbox;
    off += (descb + sizeof(int) * ncol + 255) & ~(size_t)255;
    const size_t partoff = off;
    off += sizeof(float) * (size_t)GENSQ_KSPLIT * ncolp * nq;
    char *sc = (char *)ctx_scratch(ctx, off);
    if (!sc) return EPSR_ERR_HIP;
    for (int b = 0; b < nbox; ++b) bd[b].colsum = (float *)(sc + csoff[b]);
    int rc = batch_tables(ctx, bd.data(), descb, colbox, sc + tab);
    if (rc) return rc;
    const GensqBoxDev *dbx = (const GensqBoxDev *)(sc + tab);
    const int *dcol = (const int *)(sc + tab + descb);
    timer_begin(ctx, "gensq");
    gensq_colsum_batch<<<ncol, 256, 0, st>>>(npt1, dbx, dcol);
    dim3 grid(ceil_div(nq, 16), ncolp / 16, GENSQ_KSPLIT);
    gensq_mfma_batch<<<grid, 64, 0, st>>>(nq, npt1, ycoh, foqols, ldf, dbx, dcol, ncol, (float *)(sc + partoff));
    gensq_reduce_batch<<<ceil_div((long)ncol * nq, 256), 256, 0, st>>>(nq, ncol, ncolp, GENSQ_KSPLIT,
                                                                      (const float *)(sc + partoff), dbx, dcol);
    timer_end(ctx, "gensq");
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, ctx_sync(ctx));
    return EPSR_OK;
}

struct StogBoxDev {
    const float *sofq, *xwt;
    const int *nqmin, *nqmax;
    float *gofr;
    int ldg, ncols, col0;
};

__global__ void stog_prepare_b_batch(const float *__restrict__ q, int nqpad, const StogBoxDev *__restrict__ bx,
                                     const int *__restrict__ colbox, int ncol, float *__restrict__ B)
{
    const float pi = 3.14159265358979323846f;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nqpad * ncol) return;
    const int col = t / nqpad, i = t - col * nqpad;
    const StogBoxDev &b = bx[colbox[col]];
    const int in = col - b.col0;
    float v = 0.0f;
    if (i >= b.nqmin[in] && i <= b.nqmax[in]) {
        const float qmax = pi / q[b.nqmax[in]];
        v = b.sofq[(size_t)in * nqpad + i] * dev_lorch(qmax * q[i], pi);
    }
    B[(size_t)col * nqpad + i] = v;
}

__global__ void stog_reduce_batch(float rho, const float *__restrict__ q, int nr, int ncol, int ncolp, int ksplit,
                                  const float *__restrict__ part, const float *__restrict__ r,
                                  const StogBoxDev *__restrict__ bx, const int *__restrict__ colbox)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncol * nr) return;
    const int col = (int)(t / nr), j0 = (int)(t - (long)col * nr);
    float sum = 0.0f;
    for (int kz = 0; kz < ksplit; ++kz) sum = sum + part[((size_t)kz * ncolp + col) * nr + j0];
    const float sfac = stog_sfac(q, rho);
    const int j = j0 + 1;
    const float rj = r[j];
    const StogBoxDev &b = bx[colbox[col]];
    const int in = col - b.col0;
    b.gofr[(size_t)in * b.ldg + j] = rj > 0.0f ? b.xwt[in] + sfac * sum / rj : b.xwt[in] + sfac * sum;
}

extern "C" int epsr_stog_batch(epsr_ctx *ctx, float rho, const float *q, int ldq, const float *r, int nr, int nbox,
                               const epsr_stog_box_t *boxes)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, q && r && boxes && nbox >= 1, "epsr_stog_batch: null argument or no box");
    ARG_CHECK(ctx, !ctx_capturing(ctx), "epsr_stog_batch: not for graph capture (it uploads its box tables from host memory)");
    ARG_CHECK(ctx, nr >= 1 && ldq >= 2, "epsr_stog_batch: bad sizes nr=%d ldq=%d", nr, ldq);
    std::vector<StogBoxDev> bd(nbox);
    std::vector<int> colbox;
    int ncol = 0;
    for (int b = 0; b < nbox; ++b) {
        const epsr_stog_box_t &x = boxes[b];
        ARG_CHECK(ctx, x.sofq && x.gofr && x.nqmin && x.nqmax && x.xwt, "epsr_stog_batch: box %d has a null array",
                  b);
        ARG_CHECK(ctx, x.ncols >= 1 && x.ldg > nr, "epsr_stog_batch: box %d: ncols=%d ldg=%d", b, x.ncols, x.ldg);
        bd[b] = StogBoxDev{x.sofq, x.xwt, x.nqmin, x.nqmax, x.gofr, x.ldg, x.ncols, ncol};
        for (int c = 0; c < x.ncols; ++c) colbox.push_back(b);
        ncol += x.ncols;
    }
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int nqpad = ldq;   // the Q range used is 1..ldq-1, as epsr_stog mode 1
    const int ncolp = 16 * ceil_div(ncol, 16);
    const size_t descb = sizeof(StogBoxDev) * nbox;
    const size_t tabb = (descb + sizeof(int) * ncol + 255) & ~(size_t)255;
    const size_t bb = (sizeof(float) * (size_t)nqpad * ncol + 255) & ~(size_t)255;
    char *sc = (char *)ctx_scratch(ctx, tabb + bb + sizeof(float) * (size_t)STOG_KSPLIT * ncolp * nr);
    if (!sc) return EPSR_ERR_HIP;
    int rc = batch_tables(ctx, bd.data(), descb, colbox, sc);
    if (rc) return rc;
    const StogBoxDev *dbx = (const StogBoxDev *)sc;
    const int *dcol = (const int *)(sc + descb);
    float *B = (float *)(sc + tabb), *part = (float *)(sc + tabb + bb);
    timer_begin(ctx, "stog");
    stog_prepare_b_batch<<<ceil_div((long)nqpad * ncol, 256), 256, 0, st>>>(q, nqpad, dbx, dcol, ncol, B);
    dim3 grid(ceil_div(nr, 16), ncolp / 16, STOG_KSPLIT);
    stog_mfma<<<grid, 64, 0, st>>>(q, nqpad, B, r, nr, ncol, part);
    stog_reduce_batch<<<ceil_div((long)ncol * nr, 256), 256, 0, st>>>(rho, q, nr, ncol, ncolp, STOG_KSPLIT, part, r,
                                                                      dbx, dcol);
    timer_end(ctx, "stog");
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, ctx_sync(ctx));
    return EPSR_OK;
}

// ------------------------------------------------------------------ foqols
// j1qr (src/epsr26/programs/epsr_standalone_rev3.py:3319-3329): real*8 result
__device__ __forceinline__ double dev_j1qr(float q, float r)
{
    if (q > 0.0f) {
        double qr = (double)(q * r);
        return (sin(qr) - qr * cos(qr)) / (double)(q * q * q);
    }
    return (double)(r * r * r / 3.0f);
}

__device__ __forceinline__ float cr_expf(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float cr_cosf(float x) { return (float)cos((double)x); }

// pofr (:3389-3399)
__device__ __forceinline__ float dev_pofr(float delr)
{
    if (delr <= 0.0f) return 1.0f;
    float delr3 = 3.0f / (delr * delr * delr);
    return delr3 * (cr_sinf(delr) - delr * cr_cosf(delr));
}

// the transform matrix of inita (:3035-3071), one thread per (Q point i, r point j);
// r2[j] are the bin edges of the serial r1/r2 recurrence (r2[0] = 0).
__global__ void foqols_kernel(int nq, const float *__restrict__ q, int npt1, const float *__restrict__ x,
                              const float *__restrict__ r2, const float *__restrict__ ycoh, float trfac, float sigres,
                              float sigrsq, float lorchlim, float lw, float lwq, float *__restrict__ foqols, int ldf)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * npt1) return;
    int i = t / npt1 + 1;
    int j = t - (i - 1) * npt1 + 1;
    float qq = q[i];
    float rr = x[j];
    float r1 = j == 1 ? 0.0f : r2[j - 1];
    double j1 = dev_j1qr(qq, r1), j2 = dev_j1qr(qq, r2[j]);
    float fo = ycoh[j] > 0.0f ? (float)((double)trfac * (j2 - j1) / (double)ycoh[j]) : 0.0f;
    float qsig = sigres + qq * sigrsq;
    float qrs = qsig * rr;
    float qr2 = 0.5f * qrs * qrs;
    float sums = 0.0f;
    if (qr2 < 30.0f) {
        sums = cr_expf(-qr2);
        if (lw > 0.0f || lwq > 0.0f) {
            float bwidth = rr * lorchlim * (lw + lwq * qq);
            sums = sums * dev_pofr(bwidth);
        }
    }
    foqols[(size_t)i * ldf + j] = fo * sums;
}

extern "C" int epsr_foqols(epsr_ctx *ctx, int nq, const float *q, int npt1, const float *x, const float *r2,
                           const float *ycoh, float rho, float fwhm, float fwhmq, float lorchqwidth,
                           float lorchqwidthq, float rmaxgr, float *foqols, int ldf)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, q && x && r2 && ycoh && foqols, "epsr_foqols: null argument");
    ARG_CHECK(ctx, nq >= 1 && npt1 >= 1 && ldf > npt1, "epsr_foqols: bad sizes");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const float pi = 3.14159265358979323846f;
    // scalar constants of inita (:3029-3039), float32 as the reference
    float sqrt2ln2 = sqrtf(2.0f * (float)log(2.0));
    float sigres = 0.5f * fwhm / sqrt2ln2;
    float sigrsq = 0.5f * fwhmq / sqrt2ln2;
    float trfac = 4.0f * pi * rho;
    float lorchlim = 4.4934f / rmaxgr;
    int n = nq * npt1;
    foqols_kernel<<<ceil_div(n, 256), 256, 0, ctx->stream>>>(nq, q, npt1, x, r2, ycoh, trfac, sigres, sigrsq,
                                                              lorchlim, lorchqwidth, lorchqwidthq, foqols, ldf);
    HIP_CHECK(ctx, hipGetLastError());
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return EPSR_OK;
}
