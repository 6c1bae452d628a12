// The learned Delta-E surrogate of delayed-acceptance MC: radial descriptors of a
// molecule's move, a two-hidden-layer SiLU network over them (BF16 weights and
// activations, FP32 accumulation on the gfx950 MFMA units), its gradient (FP32 MFMA)
// and an Adam step.  The reference has no counterpart; the descriptors read the same
// ledger rows as en_change (src/epsr26/library/molfit.py:805-919).
#include <cmath>

#include "common.h"
#include "surrogate.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

#define MLP_H 128
#define DESC_MAXD 2048   // channels * Gaussians

// float -> bf16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// ------------------------------------------------------------------ descriptors
// One wave per row (chain position k, current or trial).  The later partners' atoms are
// gathered into LDS a batch at a time, and every (atom of the molecule, partner atom) pair
// of a batch with r < rc is staged in the wave's LDS list (compacted by ballot) with its
// cutoff factor, formed once per pair.  At a flush the list is sorted by channel (a
// counting sort); then, channel by channel, lane g of each half-wave sums the g-th
// Gaussian of every other pair in a register, in 2^-32 fixed point; the two halves' sums
// are added and go to the row's accumulators in LDS once per channel.  Integer sums do not depend
// on the order of the pairs, so a row is a deterministic function of the coordinates.
#define DESC_WCAP 192    // staged pairs per wave
#define DESC_PCAP 128    // partner atoms gathered per batch
#define DESC_MAXCH 256   // channels (type pairs)
// LDS of the wave, in bytes: D accumulators (8 bytes each), the staged pairs (r, factor)
// and their channels, the sorted pairs, the partner atoms, channel counts and starts
#define DESC_LDS(D) (8 * ((size_t)(D) + 1) + 20 * DESC_WCAP + 16 * DESC_PCAP + 8 * (DESC_MAXCH + 2))
__global__ void __launch_bounds__(64)
surr_descriptors(int nmol, const int *__restrict__ order, const int *__restrict__ pos,
                 const int *__restrict__ rowstart, const int *__restrict__ cnt, const int *__restrict__ partner,
                 const int *__restrict__ naf, const int *__restrict__ nal, const int *__restrict__ npr, int ntype,
                 const float4 *__restrict__ Pold, const float4 *__restrict__ Pnew, float dmax, DescDev dp, int nch,
                 int maxat, int ldx, uint16_t *__restrict__ X, float *__restrict__ Xf)
{
    extern __shared__ unsigned long long dsm[];
    const int lane = threadIdx.x;
    const int row = blockIdx.x;
    const int D = nch * dp.ng;
    unsigned long long *acc = dsm;                              // [D]
    float4 *pq = (float4 *)(acc + D + (D & 1));                 // [DESC_PCAP] partner atoms
    float2 *lrf = (float2 *)(pq + DESC_PCAP);                   // [DESC_WCAP] staged (r, factor)
    float2 *srt = lrf + DESC_WCAP;                              // [DESC_WCAP] sorted by channel
    int *lch = (int *)(srt + DESC_WCAP);                        // [DESC_WCAP] staged channels
    int *chn = lch + DESC_WCAP;                                 // [DESC_MAXCH + 1] counts, cursors
    int *chs = chn + DESC_MAXCH + 1;                            // [DESC_MAXCH + 1] starts
    for (int i = lane; i < D; i += 64) acc[i] = 0ull;
    const int k = row >> 1;
    const int m = order[k];
    const int a0 = naf[m], na = nal[m] - a0 + 1;
    const float4 *Pm = ((row & 1) ? Pnew : Pold) + a0;
    const int rs = rowstart[m], c = cnt[m];
    const int spc = min(64, DESC_PCAP / maxat);   // slots per partner chunk: a chunk always fits pq
    const float rcf = dp.rc;
    const float pi_rc = 3.14159265358979323846f / rcf;
    const int h = lane >> 5, g = lane & 31;
    const float mu = dp.r0 + (float)g * dp.dmu;
    int nl = 0;   // staged pairs (wave-uniform)

    auto wave_sync = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto flush = [&]() {
        for (int i = lane; i <= nch; i += 64) chn[i] = 0;
        wave_sync();
        for (int e = lane; e < nl; e += 64) atomicAdd(&chn[lch[e]], 1);
        wave_sync();
        if (lane == 0) {
            int run = 0;
            for (int ch = 0; ch < nch; ++ch) {
                chs[ch] = run;
                run += chn[ch];
                chn[ch] = chs[ch];
            }
            chs[nch] = run;
        }
        wave_sync();
        for (int e = lane; e < nl; e += 64) srt[atomicAdd(&chn[lch[e]], 1)] = lrf[e];
        wave_sync();
        for (int ch = 0; ch < nch; ++ch) {
            const int e0 = chs[ch], e1 = chs[ch + 1];
            if (e0 == e1) continue;
            unsigned long long sum = 0ull;
#pragma unroll 4
            for (int e = e0 + h; e < e1; e += 2) {
                const float2 v = srt[e];
                const float t = (v.x - mu) * dp.inv_w;
                // the hardware exponential: within a few 1e-7 relative of exp, far inside
                // the descriptors' tolerance against the FP64 definition
                const float G = __expf(-t * t) * v.y;   // < 1, so G * 2^32 fits 32 bits
                sum += (unsigned long long)(unsigned)(G * 4294967296.0f);
            }
            sum += __shfl_xor(sum, 32);   // the two half-waves' sums, then one writer per Gaussian
            if (h == 0 && g < dp.ng) acc[ch * dp.ng + g] += sum;
        }
        wave_sync();
        nl = 0;
    };

    // the later partners' atoms, DESC_PCAP at a time, in slot order, then every (atom of the
    // molecule, partner atom) pair of the batch, 64 at a time
    int s0 = 1;
    while (s0 < c) {
        int np = 0;
        for (; s0 < c; s0 += spc) {
            const int s = s0 + lane;
            int nat = 0, b0 = 0;
            if (lane < spc && s < c) {
                const int n = partner[rs + s];
                if (pos[n] > k) {
                    b0 = naf[n];
                    nat = nal[n] - b0 + 1;
                }
            }
            int pre = nat;   // inclusive prefix of the atom counts over the lanes
            for (int off = 1; off < 64; off <<= 1) {
                const int t = __shfl_up(pre, off);
                if (lane >= off) pre += t;
            }
            const int tot = __shfl(pre, 63);
            if (np + tot > DESC_PCAP) break;   // (then np > 0: a chunk alone fits, spc * maxat <= DESC_PCAP)
            for (int j = 0; j < nat; ++j) pq[np + pre - nat + j] = Pold[b0 + j];
            np += tot;
        }
        wave_sync();
        const int npairs = na * np;
        // lane's pair i = a * np + t, advanced by 64 each step without a division
        int a = 0, t = lane;
        while (np > 0 && t >= np) {
            t -= np;
            ++a;
        }
        for (int base = 0; base < npairs; base += 64) {
            const int i = base + lane;
            bool in = false;
            float r = 0.0f;
            int ch = 0;
            if (i < npairs) {
                const float4 p = Pm[a], q = pq[t];
                // only the distance is used: x - rint(x) and the hardware square root
                // (the descriptors are checked against the FP64 definition, not bit for bit)
                const float dx = mic_sq(q.x - p.x), dy = mic_sq(q.y - p.y), dz = mic_sq(q.z - p.z);
                r = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz) * dmax;
                in = r < rcf;
                ch = npr[__float_as_int(p.w) * ntype + __float_as_int(q.w)];
            }
            t += 64;
            while (t >= np) {
                t -= np;
                ++a;
            }
            const unsigned long long mask = __ballot(in);
            if (in) {
                const int at = nl + __popcll(mask & ((1ull << lane) - 1ull));
                lrf[at] = make_float2(r, 0.5f * (__cosf(pi_rc * r) + 1.0f));
                lch[at] = ch;
            }
            nl += __popcll(mask);
            if (nl > DESC_WCAP - 64) {
                wave_sync();
                flush();
            }
        }
        wave_sync();
    }
    if (nl > 0) {
        wave_sync();
        flush();
    }
    for (int i = lane; i < ldx; i += 64) {
        float v = i < D ? (float)((double)acc[i] * 2.3283064365386963e-10) * dp.xscale : 0.0f;
        X[(size_t)row * ldx + i] = f2bf(v);
        if (Xf) Xf[(size_t)row * ldx + i] = v;
    }
}

int surr_descriptors_launch(epsr_ctx *ctx, const epsr_box *box, const epsr_ledger *l, const float4 *Pold,
                            const float4 *Pnew, const epsr_desc_params_t *dp, int ldx, uint16_t *x, float *xf)
{
    ARG_CHECK(ctx, dp->ng >= 1 && dp->ng <= 32 && dp->rc > 0.0f && dp->width > 0.0f,
              "descriptors: ng must be 1..32, rc and width positive");
    const int D = box->npairs * dp->ng;
    ARG_CHECK(ctx, D <= DESC_MAXD, "descriptors: %d channels x %d Gaussians exceed %d", box->npairs, dp->ng,
              DESC_MAXD);
    ARG_CHECK(ctx, ldx >= D && ldx % 32 == 0, "descriptors: ldx=%d must be >= %d and a multiple of 32", ldx, D);
    ARG_CHECK(ctx, box->max_atoms_per_mol <= DESC_PCAP, "descriptors: molecules of %d atoms are too large",
              box->max_atoms_per_mol);
    ARG_CHECK(ctx, box->npairs <= DESC_MAXCH, "descriptors: %d channels exceed %d", box->npairs, DESC_MAXCH);
    DescDev d{dp->ng, dp->rc, dp->r0, dp->dmu, 1.0f / dp->width, dp->xscale};
    timer_begin(ctx, "surr_descriptors");
    surr_descriptors<<<2 * box->nmol, 64, DESC_LDS(D), ctx->stream>>>(
        box->nmol, l->d_order, l->d_pos, l->d_rowstart, l->d_cnt, l->d_partner, box->d_naf, box->d_nal,
        box->d_npairref, box->ntype, Pold, Pnew, box->dmax, d, box->npairs, std::max(1, box->max_atoms_per_mol), ldx,
        x, xf);
    timer_end(ctx, "surr_descriptors");
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}

// absolute positions (x = xa + xadd) with the type in .w, as the MC cycle forms them
__global__ void surr_positions(const float *__restrict__ xa, const float *__restrict__ ya,
                               const float *__restrict__ za, const float *__restrict__ xadd,
                               const float *__restrict__ yadd, const float *__restrict__ zadd,
                               const int *__restrict__ amol, const int *__restrict__ itype, float4 *__restrict__ P,
                               int natom)
{
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natom) return;
    int m = amol[a];
    P[a] = make_float4(xa[a] + xadd[m], ya[a] + yadd[m], za[a] + zadd[m], __int_as_float(itype[a]));
}

extern "C" int epsr_descriptors(epsr_ctx *ctx, const epsr_box *box, const epsr_coords_t *c, const epsr_coords_t *trial,
                                const epsr_ledger *l, const epsr_desc_params_t *dp, int ldx, uint16_t *x, float *xf)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, box && c && trial && l && dp && x, "epsr_descriptors: null argument");
    ARG_CHECK(ctx, l->ledger_ready, "epsr_descriptors: no ledger (call epsr_fmolec1)");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int natom = box->natom;
    StreamTemp T(ctx);
    float4 *P = (float4 *)T.alloc(sizeof(float4) * 2 * (size_t)natom);
    if (!P) return EPSR_ERR_HIP;
    hipStream_t st = ctx->stream;
    surr_positions<<<ceil_div(natom, 256), 256, 0, st>>>(c->xa, c->ya, c->za, c->xadd, c->yadd, c->zadd, box->d_amol,
                                                        box->d_itype, P, natom);
    surr_positions<<<ceil_div(natom, 256), 256, 0, st>>>(trial->xa, trial->ya, trial->za, trial->xadd, trial->yadd,
                                                        trial->zadd, box->d_amol, box->d_itype, P + natom, natom);
    return surr_descriptors_launch(ctx, box, l, P, P + natom, dp, ldx, x, xf);
}

// ------------------------------------------------------------------ forward pass
// 64 rows per workgroup of four waves.  Layer 1: wave w forms columns 32w..32w+31 of
// the 64 x 128 pre-activation as 4 x 2 tiles of v_mfma_f32_16x16x32_bf16 (lane l holds
// row/column l%16 and k = 8*(l/16)..+7 of each operand, both read as 16-byte vectors);
// the bf16 activations go to LDS and feed layer 2 the same way; layer 3 is a dot
// product per row.  Training passes also keep Z1, bf16 H1 and Z2.
struct MlpDev {
    const uint16_t *W1, *W2;
    const float *b1, *b2, *w3, *b3;
    int d;
};

__device__ __forceinline__ bf16x8 ld_bf16x8(const uint16_t *p) { return __builtin_bit_cast(bf16x8, *(const uint4 *)p); }

__global__ void __launch_bounds__(256)
mlp_forward_k(int nrows, const uint16_t *__restrict__ X, int ldx, MlpDev w, float *__restrict__ y,
              float *__restrict__ Z1, uint16_t *__restrict__ H1, float *__restrict__ Z2)
{
    __shared__ __align__(16) uint16_t Hs[64][MLP_H + 8];
    __shared__ float H2s[64][MLP_H + 4];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, kq = lane >> 4;
    const int r0 = blockIdx.x * 64;
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0u, 0u, 0u, 0u));
    floatx4 acc[4][2];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < w.d; k0 += 32) {
        bf16x8 b[2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) b[ct] = ld_bf16x8(w.W1 + (size_t)(32 * wv + 16 * ct + lr) * w.d + k0 + 8 * kq);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const int r = r0 + 16 * rt + lr;
            const bf16x8 a = r < nrows ? ld_bf16x8(X + (size_t)r * ldx + k0 + 8 * kq) : zero8;
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[ct], acc[rt][ct], 0, 0, 0);
        }
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = 16 * rt + 4 * kq + v, col = 32 * wv + 16 * ct + lr;
                const float z = acc[rt][ct][v] + w.b1[col];
                const uint16_t h = f2bf(silu(z));
                Hs[rr][col] = h;
                if (Z1 && r0 + rr < nrows) {
                    Z1[(size_t)(r0 + rr) * MLP_H + col] = z;
                    H1[(size_t)(r0 + rr) * MLP_H + col] = h;
                }
                acc[rt][ct][v] = 0.0f;
            }
    __syncthreads();
    for (int k0 = 0; k0 < MLP_H; k0 += 32) {
        bf16x8 b[2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) b[ct] = ld_bf16x8(w.W2 + (size_t)(32 * wv + 16 * ct + lr) * MLP_H + k0 + 8 * kq);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
            const bf16x8 a = ld_bf16x8(&Hs[16 * rt + lr][k0 + 8 * kq]);
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[ct], acc[rt][ct], 0, 0, 0);
        }
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rr = 16 * rt + 4 * kq + v, col = 32 * wv + 16 * ct + lr;
                const float z = acc[rt][ct][v] + w.b2[col];
                H2s[rr][col] = silu(z);
                if (Z2 && r0 + rr < nrows) Z2[(size_t)(r0 + rr) * MLP_H + col] = z;
            }
    __syncthreads();
    const int rr = threadIdx.x >> 2, part = threadIdx.x & 3;
    float s = 0.0f;
    for (int j = part * 32; j < part * 32 + 32; ++j) s = s + H2s[rr][j] * w.w3[j];
    s = s + __shfl_xor(s, 1);
    s = s + __shfl_xor(s, 2);
    if (part == 0 && r0 + rr < nrows) y[r0 + rr] = s + w.b3[0];
}

static MlpDev mlp_dev(const epsr_mlp_t *m)
{
    const size_t d = (size_t)m->d, h = MLP_H;
    const float *w = m->w;
    MlpDev o;
    o.W1 = m->wb;
    o.W2 = m->wb + h * d;
    o.b1 = w + h * d;
    o.b2 = w + h * d + h + h * h;
    o.w3 = o.b2 + h;
    o.b3 = o.w3 + h;
    o.d = m->d;
    return o;
}

static int mlp_check(epsr_ctx *ctx, const epsr_mlp_t *m)
{
    ARG_CHECK(ctx, m && m->w && m->wb, "mlp: null parameters");
    ARG_CHECK(ctx, m->h == MLP_H && m->d >= 32 && m->d % 32 == 0, "mlp: h must be %d and d a multiple of 32 (d=%d h=%d)",
              MLP_H, m->d, m->h);
    return EPSR_OK;
}

int surr_forward_launch(epsr_ctx *ctx, const epsr_mlp_t *m, int nrows, const uint16_t *x, int ldx, float *y)
{
    int rc = mlp_check(ctx, m);
    if (rc) return rc;
    ARG_CHECK(ctx, ldx >= m->d && ldx % 8 == 0, "mlp: ldx=%d must be >= d=%d and a multiple of 8", ldx, m->d);
    if (nrows <= 0) return EPSR_OK;
    timer_begin(ctx, "mlp_forward");
    mlp_forward_k<<<ceil_div(nrows, 64), 256, 0, ctx->stream>>>(nrows, x, ldx, mlp_dev(m), y, nullptr, nullptr,
                                                                 nullptr);
    timer_end(ctx, "mlp_forward");
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}

// ------------------------------------------------------------------ backward pass
// C[z][m][n] = sum over k in chunk z of A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn] (fp32 or bf16 storage), on v_mfma_f32_16x16x4_f32: 64 x 64
// tiles per workgroup, 32 x 32 per wave.
__device__ __forceinline__ float ldv(const float *p) { return *p; }
__device__ __forceinline__ float ldv(const uint16_t *p) { return bf2f(*p); }

template <class TA, class TB>
__global__ void __launch_bounds__(256)
gemm_f32(int M, int N, int K, const TA *__restrict__ A, long long sam, long long sak, const TB *__restrict__ B,
         long long sbk, long long sbn, int kchunk, float *__restrict__ C)
{
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, kq = lane >> 4;
    const int m0 = blockIdx.x * 64 + 32 * (wv >> 1), n0 = blockIdx.y * 64 + 32 * (wv & 1);
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k = kb; k < ke; k += 4) {
        const int kk = k + kq;
        float a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int m = m0 + 16 * i + lr, n = n0 + 16 * i + lr;
            a[i] = (m < M && kk < ke) ? ldv(A + (long long)m * sam + (long long)kk * sak) : 0.0f;
            b[i] = (n < N && kk < ke) ? ldv(B + (long long)kk * sbk + (long long)n * sbn) : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    float *Cz = C + (size_t)blockIdx.z * M * N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + 16 * i + 4 * kq + v, n = n0 + 16 * j + lr;
                if (m < M && n < N) Cz[(size_t)m * N + n] = acc[i][j][v];
            }
}

// out[i] = sum over z (in order) of part[z][i]
__global__ void sum_parts(int n, int nz, const float *__restrict__ part, float *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < nz; ++z) s = s + part[(size_t)z * n + i];
    out[i] = s;
}

// output gradient of the pair loss: dy = +-2 e / npair, loss partials per block
__global__ void __launch_bounds__(256)
mlp_out_grad(int npair, const float *__restrict__ y, const float *__restrict__ target, float tclip,
             float *__restrict__ dy, double *__restrict__ lpart)
{
    __shared__ double s[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < npair) {
        const float t = tclip > 0.0f ? fminf(fmaxf(target[i], -tclip), tclip) : target[i];
        const float e = (y[2 * i + 1] - y[2 * i]) - t;
        const float g = 2.0f * e / (float)npair;
        dy[2 * i] = -g;
        dy[2 * i + 1] = g;
        l = (double)e * (double)e;
    }
    s[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = s[0];
}

__global__ void sum_loss(int n, int npair, const double *__restrict__ lpart, float *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += lpart[i];
    out[0] = (float)(s / (double)npair);
}

// dZ2 = dy w3 silu'(Z2) and H2 = silu(Z2) (in place of Z2)
__global__ void mlp_dz2(int nrows, const float *__restrict__ dy, const float *__restrict__ w3, float *__restrict__ Z2,
                        float *__restrict__ dZ2)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nrows * MLP_H) return;
    const int r = (int)(i / MLP_H), j = (int)(i % MLP_H);
    const float z = Z2[i];
    dZ2[i] = dy[r] * w3[j] * dsilu(z);
    Z2[i] = silu(z);
}

// dZ1 = dH1 silu'(Z1), in place
__global__ void mlp_dz1(int nrows, const float *__restrict__ Z1, float *__restrict__ dH)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nrows * MLP_H) return;
    dH[i] = dH[i] * dsilu(Z1[i]);
}

// column sums of M [nrows][128] (times rowscale[r] when given) over chunks of rows
__global__ void __launch_bounds__(128)
col_sums(int nrows, const float *__restrict__ M, const float *__restrict__ rowscale, int chunk, float *__restrict__ part)
{
    const int j = threadIdx.x;
    const int r0 = blockIdx.x * chunk, r1 = min(nrows, r0 + chunk);
    float s = 0.0f;
    for (int r = r0; r < r1; ++r) {
        const float v = M[(size_t)r * MLP_H + j];
        s = s + (rowscale ? v * rowscale[r] : v);
    }
    part[(size_t)blockIdx.x * MLP_H + j] = s;
}

__global__ void mlp_to_bf16(size_t n, const float *__restrict__ w, uint16_t *__restrict__ wb)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

__global__ void adam_k(size_t n, float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

extern "C" long long epsr_mlp_nparams(int d, int h) { return (long long)h * d + h + (long long)h * h + h + h + 1; }

extern "C" int epsr_mlp_sync(epsr_ctx *ctx, const epsr_mlp_t *m)
{
    if (!ctx) return EPSR_ERR_ARG;
    int rc = mlp_check(ctx, m);
    if (rc) return rc;
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const size_t hd = (size_t)MLP_H * m->d, hh = (size_t)MLP_H * MLP_H;
    mlp_to_bf16<<<ceil_div((long)hd, 256), 256, 0, ctx->stream>>>(hd, m->w, m->wb);
    mlp_to_bf16<<<ceil_div((long)hh, 256), 256, 0, ctx->stream>>>(hh, m->w + hd + MLP_H, m->wb + hd);
    HIP_CHECK(ctx, hipGetLastError());
    return EPSR_OK;
}

extern "C" int epsr_mlp_forward(epsr_ctx *ctx, const epsr_mlp_t *m, int nrows, const uint16_t *x, int ldx, float *y)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, x && y && nrows >= 0, "epsr_mlp_forward: bad argument");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return surr_forward_launch(ctx, m, nrows, x, ldx, y);
}

extern "C" int epsr_mlp_grad(epsr_ctx *ctx, const epsr_mlp_t *m, int npair, const uint16_t *x, int ldx,
                             const float *target, float tclip, float *grad, float *loss)
{
    if (!ctx) return EPSR_ERR_ARG;
    int rc = mlp_check(ctx, m);
    if (rc) return rc;
    ARG_CHECK(ctx, x && target && grad && npair >= 1, "epsr_mlp_grad: bad argument");
    ARG_CHECK(ctx, ldx >= m->d && ldx % 8 == 0, "epsr_mlp_grad: ldx=%d must be >= d=%d and a multiple of 8", ldx, m->d);
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int R = 2 * npair, d = m->d, H = MLP_H;
    const int kchunk = 1024, nz = ceil_div(R, kchunk);
    const int lblocks = ceil_div(npair, 256), cchunk = 256, ncb = ceil_div(R, cchunk);
    StreamTemp T(ctx);
    float *y = (float *)T.alloc(sizeof(float) * R);
    float *Z1 = (float *)T.alloc(sizeof(float) * (size_t)R * H);
    uint16_t *H1 = (uint16_t *)T.alloc(sizeof(uint16_t) * (size_t)R * H);
    float *Z2 = (float *)T.alloc(sizeof(float) * (size_t)R * H);
    float *dZ2 = (float *)T.alloc(sizeof(float) * (size_t)R * H);
    float *dH = (float *)T.alloc(sizeof(float) * (size_t)R * H);
    float *dy = (float *)T.alloc(sizeof(float) * R);
    double *lpart = (double *)T.alloc(sizeof(double) * lblocks);
    float *dloss = (float *)T.alloc(sizeof(float));
    float *part = (float *)T.alloc(sizeof(float) * (size_t)nz * H * std::max(d, H));
    float *cpart = (float *)T.alloc(sizeof(float) * (size_t)ncb * H);
    if (!y || !Z1 || !H1 || !Z2 || !dZ2 || !dH || !dy || !lpart || !dloss || !part || !cpart) return EPSR_ERR_HIP;
    const MlpDev w = mlp_dev(m);
    float *gW1 = grad, *gb1 = grad + (size_t)H * d, *gW2 = gb1 + H, *gb2 = gW2 + (size_t)H * H, *gw3 = gb2 + H,
          *gb3 = gw3 + H;
    timer_begin(ctx, "mlp_train");
    mlp_forward_k<<<ceil_div(R, 64), 256, 0, st>>>(R, x, ldx, w, y, Z1, H1, Z2);
    mlp_out_grad<<<lblocks, 256, 0, st>>>(npair, y, target, tclip, dy, lpart);
    sum_loss<<<1, 64, 0, st>>>(lblocks, npair, lpart, dloss);
    const size_t RH = (size_t)R * H;
    mlp_dz2<<<ceil_div((long)RH, 256), 256, 0, st>>>(R, dy, w.w3, Z2, dZ2);   // Z2 now holds H2
    // dw3 = H2^T dy, db2 = colsum dZ2, db3 = 0 (b3 cancels in f(new) - f(old))
    col_sums<<<ncb, H, 0, st>>>(R, Z2, dy, cchunk, cpart);
    sum_parts<<<1, H, 0, st>>>(H, ncb, cpart, gw3);
    col_sums<<<ncb, H, 0, st>>>(R, dZ2, nullptr, cchunk, cpart);
    sum_parts<<<1, H, 0, st>>>(H, ncb, cpart, gb2);
    HIP_CHECK(ctx, hipMemsetAsync(gb3, 0, sizeof(float), st));
    // dW2 = dZ2^T H1
    gemm_f32<float, uint16_t><<<dim3(ceil_div(H, 64), ceil_div(H, 64), nz), 256, 0, st>>>(H, H, R, dZ2, 1, H, H1, H, 1,
                                                                                          kchunk, part);
    sum_parts<<<ceil_div((long)H * H, 256), 256, 0, st>>>(H * H, nz, part, gW2);
    // dH1 = dZ2 W2, then dZ1
    gemm_f32<float, uint16_t><<<dim3(ceil_div(R, 64), ceil_div(H, 64), 1), 256, 0, st>>>(R, H, H, dZ2, H, 1, w.W2, H, 1,
                                                                                        H, dH);
    mlp_dz1<<<ceil_div((long)RH, 256), 256, 0, st>>>(R, Z1, dH);
    col_sums<<<ncb, H, 0, st>>>(R, dH, nullptr, cchunk, cpart);
    sum_parts<<<1, H, 0, st>>>(H, ncb, cpart, gb1);
    // dW1 = dZ1^T X
    gemm_f32<float, uint16_t><<<dim3(ceil_div(H, 64), ceil_div(d, 64), nz), 256, 0, st>>>(H, d, R, dH, 1, H, x, ldx, 1,
                                                                                         kchunk, part);
    sum_parts<<<ceil_div((long)H * d, 256), 256, 0, st>>>(H * d, nz, part, gW1);
    timer_end(ctx, "mlp_train");
    HIP_CHECK(ctx, hipGetLastError());
    if (loss) {
        HIP_CHECK(ctx, hipMemcpyAsync(loss, dloss, sizeof(float), hipMemcpyDeviceToHost, st));
        HIP_CHECK(ctx, hipStreamSynchronize(st));
    }
    return EPSR_OK;
}

extern "C" int epsr_mlp_adam(epsr_ctx *ctx, epsr_mlp_t *m, const float *grad, float lr, float beta1, float beta2,
                             float eps)
{
    if (!ctx) return EPSR_ERR_ARG;
    int rc = mlp_check(ctx, m);
    if (rc) return rc;
    ARG_CHECK(ctx, grad && m->m && m->v, "epsr_mlp_adam: null gradient or moments");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    m->step += 1;
    const float c1 = 1.0f - powf(beta1, (float)m->step), c2 = 1.0f - powf(beta2, (float)m->step);
    const size_t n = (size_t)epsr_mlp_nparams(m->d, m->h);
    adam_k<<<ceil_div((long)n, 256), 256, 0, ctx->stream>>>(n, m->w, m->m, m->v, grad, lr, beta1, beta2, eps, c1, c2);
    HIP_CHECK(ctx, hipGetLastError());
    return epsr_mlp_sync(ctx, m);
}
