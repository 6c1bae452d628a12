// Internal interface between the MC cycle and the surrogate kernels.
#pragma once
#include "common.h"

struct DescDev {
    int ng;
    float rc, r0, dmu, inv_w, xscale;
};

// descriptors of every chain position (rows 2k: Pold, 2k+1: Pnew) into x (bf16) and xf
int surr_descriptors_launch(epsr_ctx *ctx, const epsr_box *box, const epsr_ledger *l, const float4 *Pold,
                            const float4 *Pnew, const epsr_desc_params_t *dp, int ldx, uint16_t *x, float *xf);
// y[r] = f(x row r)
int surr_forward_launch(epsr_ctx *ctx, const epsr_mlp_t *m, int nrows, const uint16_t *x, int ldx, float *y);

// the stage-1 uniform of molecule m: splitmix64 (Steele, Lea and Flood, 2014) of
// seed + (m + 1) * golden gamma, top 24 bits, in (0, 1)
__host__ __device__ inline float da_uniform(unsigned long long seed, int m)
{
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(m + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((float)(z >> 40) + 0.5f) * (1.0f / 16777216.0f);
}
