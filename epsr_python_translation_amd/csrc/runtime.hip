// Context, box topology upload, scratch memory, kernel timing and the host ran1.
#include <cstdarg>

#include <algorithm>

#include "common.h"

int set_error(epsr_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

extern "C" int epsr_ctx_create(int device, epsr_ctx **out)
{
    if (!out) return EPSR_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return EPSR_ERR_HIP;
    epsr_ctx *ctx = new epsr_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return EPSR_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    *out = ctx;
    return EPSR_OK;
}

extern "C" void epsr_ctx_destroy(epsr_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->timers) {
        for (hipEvent_t e : kv.second.start) hipEventDestroy(e);
        for (hipEvent_t e : kv.second.stop) hipEventDestroy(e);
    }
    if (ctx->side) {
        hipStreamSynchronize(ctx->side);
        hipStreamDestroy(ctx->side);
    }
    if (ctx->ev_fork) hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_done) hipEventDestroy(ctx->ev_done);
    if (ctx->ev_join) hipEventDestroy(ctx->ev_join);
    for (hipEvent_t e : ctx->ev_chunk)
        if (e) hipEventDestroy(e);
    if (ctx->scratch) hipFree(ctx->scratch);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

int ctx_wait(epsr_ctx *ctx)
{
    if (!ctx->ev_done) HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->ev_done, hipEventDisableTiming));
    HIP_CHECK(ctx, hipEventRecord(ctx->ev_done, ctx->stream));
    for (;;) {
        const hipError_t e = hipEventQuery(ctx->ev_done);
        if (e == hipSuccess) return EPSR_OK;
        if (e != hipErrorNotReady)
            return set_error(ctx, EPSR_ERR_HIP, "hipEventQuery failed: %s", hipGetErrorString(e));
    }
}

int ctx_side(epsr_ctx *ctx, hipStream_t *out)
{
    if (!ctx->side) {
        HIP_CHECK(ctx, hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking));
        HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
        HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    }
    *out = ctx->side;
    return EPSR_OK;
}

extern "C" const char *epsr_last_error(const epsr_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

extern "C" int epsr_ctx_set_stream(epsr_ctx *ctx, void *s)
{
    if (!ctx) return EPSR_ERR_ARG;
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return EPSR_OK;
}

extern "C" int epsr_ctx_synchronize(epsr_ctx *ctx)
{
    if (!ctx) return EPSR_ERR_ARG;
    HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return EPSR_OK;
}

extern "C" int epsr_ctx_enable_timing(epsr_ctx *ctx, int on)
{
    if (!ctx) return EPSR_ERR_ARG;
    ctx->timing = on != 0;
    return EPSR_OK;
}

bool ctx_capturing(epsr_ctx *ctx) { return ctx->capturing; }

hipError_t ctx_sync(epsr_ctx *ctx) { return ctx_capturing(ctx) ? hipSuccess : hipStreamSynchronize(ctx->stream); }

// read the durations of every recorded launch of t (waits for the last one)
static void timer_flush(KernelTimer &t)
{
    if (t.n == 0) return;
    hipEventSynchronize(t.stop[t.n - 1]);
    for (int i = 0; i < t.n; ++i) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, t.start[i], t.stop[i]);
        t.sum_ms += ms;
        t.count += 1;
        t.last_ms = ms;
    }
    t.n = 0;
}

// at most this many launches of a family are kept unread; the next one reads them first
static const int TIMER_RECORDS = 1024;

void timer_begin(epsr_ctx *ctx, const char *family)
{
    if (!ctx->timing || ctx_capturing(ctx)) return;
    KernelTimer &t = ctx->timers[family];
    if (t.n == TIMER_RECORDS) timer_flush(t);
    if (t.n == (int)t.start.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        hipEventCreate(&a);
        hipEventCreate(&b);
        t.start.push_back(a);
        t.stop.push_back(b);
    }
    hipEventRecord(t.start[t.n], ctx->stream);
}

void timer_end(epsr_ctx *ctx, const char *family)
{
    if (!ctx->timing || ctx_capturing(ctx)) return;
    KernelTimer &t = ctx->timers[family];
    if (t.n == (int)t.start.size()) return;   // no begin
    hipEventRecord(t.stop[t.n], ctx->stream);
    t.n += 1;
}

extern "C" double epsr_ctx_last_kernel_ms(const epsr_ctx *cctx, const char *family)
{
    epsr_ctx *ctx = const_cast<epsr_ctx *>(cctx);
    if (!ctx || !family) return -1.0;
    auto it = ctx->timers.find(family);
    if (it == ctx->timers.end()) return -1.0;
    KernelTimer &t = it->second;
    timer_flush(t);
    return t.last_ms;
}

extern "C" double epsr_ctx_kernel_ms_sum(epsr_ctx *ctx, const char *family, long long *count)
{
    if (count) *count = 0;
    if (!ctx || !family) return -1.0;
    auto it = ctx->timers.find(family);
    if (it == ctx->timers.end()) return -1.0;
    KernelTimer &t = it->second;
    timer_flush(t);
    if (count) *count = t.count;
    const double s = t.count ? t.sum_ms : -1.0;
    t.sum_ms = 0.0;
    t.count = 0;
    return s;
}

void *ctx_scratch(epsr_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->scratch_bytes) return ctx->scratch;
    if (ctx_capturing(ctx)) {
        set_error(ctx, EPSR_ERR_ARG, "scratch of %zu bytes needed while capturing: run the work once before capture",
                  bytes);
        return nullptr;
    }
    if (ctx->scratch) {
        hipStreamSynchronize(ctx->stream);
        hipFree(ctx->scratch);
        ctx->scratch = nullptr;
        ctx->scratch_bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 4096;
    if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
        set_error(ctx, EPSR_ERR_HIP, "hipMalloc of %zu bytes of scratch failed", want);
        ctx->scratch = nullptr;
        return nullptr;
    }
    ctx->scratch_bytes = want;
    return ctx->scratch;
}

void *ctx_pinned(epsr_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->pinned_bytes) return ctx->pinned;
    if (ctx->pinned) {
        hipStreamSynchronize(ctx->stream);
        hipHostFree(ctx->pinned);
        ctx->pinned = nullptr;
        ctx->pinned_bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 4096;
    if (hipHostMalloc(&ctx->pinned, want, hipHostMallocDefault) != hipSuccess) {
        set_error(ctx, EPSR_ERR_HIP, "hipHostMalloc of %zu bytes failed", want);
        ctx->pinned = nullptr;
        return nullptr;
    }
    ctx->pinned_bytes = want;
    return ctx->pinned;
}

// ------------------------------------------------------------------ box

template <class T>
static int upload(epsr_ctx *ctx, T **dst, const T *src, size_t n)
{
    HIP_CHECK(ctx, hipMalloc((void **)dst, sizeof(T) * (n > 0 ? n : 1)));
    if (n) HIP_CHECK(ctx, hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
    return EPSR_OK;
}

extern "C" int epsr_box_create(epsr_ctx *ctx, const epsr_box_desc_t *d, epsr_box **out)
{
    if (!ctx || !d || !out) return set_error(ctx, EPSR_ERR_ARG, "epsr_box_create: null argument");
    ARG_CHECK(ctx, d->nmol > 0 && d->natom > 0 && d->ntype > 0, "epsr_box_create: empty box");
    ARG_CHECK(ctx, d->npairs == d->ntype * (d->ntype + 1) / 2, "epsr_box_create: npairs != ntype(ntype+1)/2");
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    epsr_box *b = new epsr_box();
    b->ctx = ctx;
    b->nmol = d->nmol; b->natom = d->natom; b->ntype = d->ntype; b->npairs = d->npairs;
    b->nmoltype = d->nmoltype;
    b->natomr = d->natomr; b->sizefactor = d->sizefactor; b->dmax = d->dmax; b->cellst = d->cellst;
    b->h_itype.assign(d->itype, d->itype + d->natom);
    b->h_naf.assign(d->naf, d->naf + d->nmol);
    b->h_nal.assign(d->nal, d->nal + d->nmol);
    b->h_mtype.assign(d->mtype, d->mtype + d->nmol);
    b->h_npairref.assign(d->npairref, d->npairref + d->ntype * d->ntype);
    b->h_n1extr.assign(d->n1extr, d->n1extr + d->npairs);
    b->h_n2extr.assign(d->n2extr, d->n2extr + d->npairs);
    b->h_cfrac.assign(d->cfrac, d->cfrac + d->ntype);
    for (int i = 0; i < d->ntype * d->ntype; ++i)
        if (d->npairref[i] < 0 || d->npairref[i] >= d->npairs) {
            delete b;
            return set_error(ctx, EPSR_ERR_ARG, "epsr_box_create: npairref[%d] = %d outside [0,%d)", i, d->npairref[i],
                             d->npairs);
        }
    std::vector<int> amol(d->natom, -1);
    b->h_type_count.assign((size_t)d->nmol * d->ntype, 0);
    int prev_end = -1;
    for (int im = 0; im < d->nmol; ++im) {
        int a0 = d->naf[im], a1 = d->nal[im];
        if (a0 != prev_end + 1 || a1 < a0 || a1 >= d->natom || d->mtype[im] < 0 || d->mtype[im] >= d->nmoltype) {
            delete b;
            return set_error(ctx, EPSR_ERR_ARG,
                             "epsr_box_create: molecule %d atoms [%d,%d] are not contiguous in order", im, a0, a1);
        }
        prev_end = a1;
        b->max_atoms_per_mol = std::max(b->max_atoms_per_mol, a1 - a0 + 1);
        for (int a = a0; a <= a1; ++a) {
            amol[a] = im;
            int t = d->itype[a];
            if (t < 0 || t >= d->ntype) {
                delete b;
                return set_error(ctx, EPSR_ERR_ARG, "epsr_box_create: atom %d has type %d", a, t);
            }
            b->h_type_count[(size_t)im * d->ntype + t] += 1;
        }
    }
    if (prev_end != d->natom - 1) {
        delete b;
        return set_error(ctx, EPSR_ERR_ARG, "epsr_box_create: molecules cover %d of %d atoms", prev_end + 1, d->natom);
    }
    int rc;
    if ((rc = upload(ctx, &b->d_itype, d->itype, d->natom)) || (rc = upload(ctx, &b->d_naf, d->naf, d->nmol)) ||
        (rc = upload(ctx, &b->d_nal, d->nal, d->nmol)) || (rc = upload(ctx, &b->d_mtype, d->mtype, d->nmol)) ||
        (rc = upload(ctx, &b->d_npairref, d->npairref, (size_t)d->ntype * d->ntype)) ||
        (rc = upload(ctx, &b->d_amol, amol.data(), d->natom))) {
        epsr_box_destroy(b);
        return rc;
    }
    *out = b;
    return EPSR_OK;
}

extern "C" void epsr_box_destroy(epsr_box *b)
{
    if (!b) return;
    hipFree(b->d_itype); hipFree(b->d_naf); hipFree(b->d_nal); hipFree(b->d_mtype);
    hipFree(b->d_npairref); hipFree(b->d_amol);
    intra_free(b->intra);
    delete b;
}

extern "C" int epsr_box_set_launch(epsr_box *b, int knob, int value)
{
    if (!b) return EPSR_ERR_ARG;
    epsr_ctx *ctx = b->ctx;
    switch (knob) {
    case EPSR_LAUNCH_ATOGEN_WORKGROUPS:
        ARG_CHECK(ctx, value >= 1 && value <= (1 << 20), "epsr_box_set_launch: atogen workgroups %d", value);
        b->atogen_workgroups = value;
        return EPSR_OK;
    case EPSR_LAUNCH_PAIR_BLOCK:
        ARG_CHECK(ctx, value == 256 || value == 512 || value == 1024, "epsr_box_set_launch: pair block %d", value);
        b->pair_block = value;
        return EPSR_OK;
    case EPSR_LAUNCH_RESOLVE_PASSES:
        ARG_CHECK(ctx, value >= 0, "epsr_box_set_launch: resolver passes %d", value);
        b->resolve_passes = value;
        return EPSR_OK;
    default:
        return set_error(ctx, EPSR_ERR_ARG, "epsr_box_set_launch: unknown knob %d", knob);
    }
}

// ------------------------------------------------------------------ graphs
struct epsr_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int epsr_graph_begin(epsr_ctx *ctx)
{
    if (!ctx) return EPSR_ERR_ARG;
    HIP_CHECK(ctx, hipSetDevice(ctx->device));
    ARG_CHECK(ctx, ctx->stream != nullptr && ctx->stream != hipStreamPerThread,
              "epsr_graph_begin: capture needs a stream of its own (epsr_ctx_set_stream)");
    HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return EPSR_OK;
}

extern "C" int epsr_graph_end(epsr_ctx *ctx, epsr_graph **out)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, out, "epsr_graph_end: null argument");
    epsr_graph *g = new epsr_graph();
    ctx->capturing = false;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
    if (e != hipSuccess) {
        delete g;
        return set_error(ctx, EPSR_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g->graph);
        delete g;
        return set_error(ctx, EPSR_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    *out = g;
    return EPSR_OK;
}

extern "C" int epsr_graph_launch(epsr_ctx *ctx, const epsr_graph *g)
{
    if (!ctx) return EPSR_ERR_ARG;
    ARG_CHECK(ctx, g && g->exec, "epsr_graph_launch: null graph");
    HIP_CHECK(ctx, hipGraphLaunch(g->exec, ctx->stream));
    return EPSR_OK;
}

extern "C" void epsr_graph_destroy(epsr_graph *g)
{
    if (!g) return;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
}

// ------------------------------------------------------------------ ran1
// Park-Miller minimal standard generator with the Bays-Durham shuffle, in the
// reference's integer arithmetic (auxiliary_routines.py:324-385; constants in
// src/epsr26/common/ran1com.py:12-20).
float ran1_next(Ran1 &s)
{
    const int32_t IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836, NTAB = 32, NDIV = 67108864;
    const float AM = 4.656612873077393e-10f, RNMX = 0.9999998807907104f;
    int32_t idum = s.idum, iy = s.iy;
    if (idum <= 0 || iy == 0) {
        int32_t w = (int32_t)(uint32_t)(uint64_t)(-(int64_t)idum);
        idum = w > 1 ? w : 1;
        for (int j = NTAB + 8; j >= 1; --j) {
            int32_t k = idum / IQ;
            idum = IA * (idum - k * IQ) - IR * k;
            if (idum < 0) idum += IM;
            if (j <= NTAB) s.iv[j - 1] = idum;
        }
        iy = s.iv[0];
    }
    int32_t k = idum / IQ;
    idum = IA * (idum - k * IQ) - IR * k;
    if (idum < 0) idum += IM;
    int j = 1 + iy / NDIV;
    iy = s.iv[j - 1];
    s.iv[j - 1] = idum;
    float r = AM * (float)iy;
    if (r > RNMX) r = RNMX;
    s.idum = idum;
    s.iy = iy;
    s.mcall = (int32_t)((uint32_t)s.mcall + 1u);
    return r;
}

// a * x mod (2^31 - 1) for 0 < a, x < 2^31 - 1: the 62-bit product folded at bit 31
// (2^31 = 1 mod 2^31 - 1); the sum is below 2 (2^31 - 1), and never a multiple of it
static inline uint32_t mulmod_m31(uint32_t x, uint32_t a)
{
    const uint64_t p = (uint64_t)x * a;
    uint32_t r = (uint32_t)(p & 0x7fffffffu) + (uint32_t)(p >> 31);
    if (r >= 0x7fffffffu) r -= 0x7fffffffu;
    return r;
}

// n values of the same stream.  IA*idum mod (2^31 - 1) is exact as Schrage's
// factorisation is; the shuffle index iy/NDIV is iy >> 26 (NDIV = 2^26, iy >= 0).  In
// blocks: the generator values first, eight chains apart (each steps by IA^8, so the
// multiplications are independent), then the shuffle (a chain of table reads), then the
// conversion to float.
void ran1_fill(Ran1 &s, float *out, long long n)
{
    const float AM = 4.656612873077393e-10f, RNMX = 0.9999998807907104f;
    const uint32_t IA = 16807u;
    uint32_t ia8 = 1u;
    for (int k = 0; k < 8; ++k) ia8 = mulmod_m31(ia8, IA);
    constexpr int BLOCK = 4096;
    uint32_t v[BLOCK];
    long long i = 0;
    while (i < n) {
        // the (re)seeding call, taken whenever idum <= 0 or iy == 0 as ran1 does; in a
        // seeded state neither is ever reached again (idum stays in [1, 2^31 - 2], iy
        // takes earlier idum values)
        if (s.idum <= 0 || s.iy == 0) {
            out[i++] = ran1_next(s);
            continue;
        }
        const int nb = (int)std::min<long long>(BLOCK, n - i);
        uint32_t x = (uint32_t)s.idum, c[8];
        for (int k = 0; k < 8 && k < nb; ++k) v[k] = c[k] = x = mulmod_m31(x, IA);
        int k8 = 8;
        for (; k8 + 8 <= nb; k8 += 8)   // the eight chains in registers
            for (int q = 0; q < 8; ++q) v[k8 + q] = c[q] = mulmod_m31(c[q], ia8);
        for (; k8 < nb; ++k8) v[k8] = mulmod_m31(v[k8 - 8], ia8);
        const int32_t last = (int32_t)v[nb - 1];
        int32_t iy = s.iy;
        int32_t iv[32];
        for (int j = 0; j < 32; ++j) iv[j] = s.iv[j];
        for (int k = 0; k < nb; ++k) {
            const int j = iy >> 26;
            iy = iv[j];
            iv[j] = (int32_t)v[k];
            v[k] = (uint32_t)iy;
        }
        for (int k = 0; k < nb; ++k) {
            const float f = AM * (float)(int32_t)v[k];
            out[i + k] = f > RNMX ? RNMX : f;
        }
        s.idum = last;
        s.iy = iy;
        for (int j = 0; j < 32; ++j) s.iv[j] = iv[j];
        s.mcall = (int32_t)((uint32_t)s.mcall + (uint32_t)nb);
        i += nb;
    }
}

void ran1_load(Ran1 &s, const int32_t *a)
{
    s.mcall = a[0]; s.idum = a[1]; s.iy = a[2];
    for (int i = 0; i < 32; ++i) s.iv[i] = a[3 + i];
}

void ran1_store(const Ran1 &s, int32_t *a)
{
    a[0] = s.mcall; a[1] = s.idum; a[2] = s.iy;
    for (int i = 0; i < 32; ++i) a[3 + i] = s.iv[i];
}

// exposed for the host-side tests of the generator
extern "C" int epsr_ran1_stream(int32_t *state, int n, float *out)
{
    if (!state || (n > 0 && !out)) return EPSR_ERR_ARG;
    Ran1 s;
    ran1_load(s, state);
    ran1_fill(s, out, n);
    ran1_store(s, state);
    return EPSR_OK;
}
