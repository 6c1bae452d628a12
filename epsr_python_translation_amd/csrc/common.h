// Internal definitions shared by the libepsr_mi355x.so translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/epsr_mi355x.h"

// ------------------------------------------------------------------ context

// event pairs of the launches of one kernel family not yet read (n of them), the
// duration of the last one read, and the sum and count of those read since the last sum
struct KernelTimer {
    std::vector<hipEvent_t> start, stop;
    int n = 0;
    double last_ms = -1.0;
    double sum_ms = 0.0;
    long long count = 0;
};

struct epsr_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    bool timing = false;
    std::map<std::string, KernelTimer> timers;
    // growable device scratch, reused across calls
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // growable pinned host staging buffer
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    // between epsr_graph_begin and epsr_graph_end
    bool capturing = false;
    // a second stream (default priority) for work that can overlap the main stream's, and
    // the events that fork it from and join it back to the main stream
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // marks the end of a call's work for ctx_wait
    hipEvent_t ev_done = nullptr;
    // mark the chunks of a copy that the host consumes as they land (epsr_press)
    hipEvent_t ev_chunk[8] = {};
};

// intramolecular topology of a box (epsr_box_set_intra), device copies
struct IntraTopo {
    int ldb = 0, ngroup = 0, nlabrot = 0, ldd = 1, refine_intra = 0, max_groups = 0;
    float dmaxmol = 1.f, tempb = 0.f, refine_intra_fac = 0.f;
    std::vector<int> h_ngrot;   // groups per molecule
    int *d_qtype = nullptr, *d_imageatom = nullptr, *d_nbond = nullptr, *d_iabond = nullptr;
    float *d_tbond = nullptr, *d_amassr = nullptr, *d_amassa = nullptr;
    int *d_nrf = nullptr, *d_nrl = nullptr, *d_labaxi = nullptr, *d_rotf = nullptr, *d_rotl = nullptr;
    int *d_labrot = nullptr, *d_isdih = nullptr, *d_ndih = nullptr, *d_labrotdih = nullptr;
    float *d_dihedralang = nullptr;
};
void intra_free(IntraTopo *t);

struct epsr_box {
    epsr_ctx *ctx = nullptr;
    IntraTopo *intra = nullptr;
    int nmol = 0, natom = 0, ntype = 0, npairs = 0, nmoltype = 0;
    float natomr = 0.f, sizefactor = 1.f, dmax = 1.f, cellst = 0.f;
    // host copies
    std::vector<int> h_itype, h_naf, h_nal, h_mtype, h_npairref, h_n1extr, h_n2extr;
    std::vector<float> h_cfrac;
    std::vector<int> h_type_count;   // [nmol*ntype] atoms of each type per molecule
    int max_atoms_per_mol = 0;
    // device copies
    int *d_itype = nullptr, *d_naf = nullptr, *d_nal = nullptr, *d_mtype = nullptr;
    int *d_npairref = nullptr, *d_amol = nullptr;
    // launch shapes (epsr_box_set_launch): workgroups atogen's pair kernel aims for, and
    // the workgroup size of the pair-energy kernels
    int atogen_workgroups = 768, pair_block = 1024;
    // the cap on the one-workgroup resolver's passes (-1: nmol + 1); a smaller cap is a test hook
    int resolve_passes = -1;
};

// ------------------------------------------------------------------ ledger

struct epsr_ledger {
    epsr_ctx *ctx = nullptr;
    const epsr_box *box = nullptr;
    int nmol = 0;
    // cells: the lattice and neighbour offsets on the host; the cell of every molecule,
    // the chain order (d_order) and chain positions (d_pos) on the device
    int na = 0, nb = 0, nc = 0, ncell = 0, ncellneighb = 0, nnbrefmax = 0;
    std::vector<int> iaoff, iboff, icoff;
    bool cells_ready = false;
    // ledger (device)
    int nself = 0, npairent = 0, nslots = 0, nentries = 0;
    int *d_rowstart = nullptr, *d_cnt = nullptr, *d_order = nullptr, *d_pos = nullptr;
    // the lists lie in chain order: d_rowk[k] is the first slot of chain position k's list
    // (d_rowk[nmol] = the slot count), so the list of position k is [rowk[k], rowk[k + 1])
    int *d_rowk = nullptr;
    int *d_partner = nullptr, *d_entry = nullptr, *d_mirror = nullptr, *d_slotmol = nullptr;
    // per slot: partner's chain position, mirror unless the partner's first slot (else -1),
    // and the packed record of the MC pair kernel
    int *d_pn = nullptr, *d_mw = nullptr;
    int4 *d_srec = nullptr;
    float *d_ergpair = nullptr;
    size_t cap_slots = 0, cap_entries = 0;
    // the MC plan of the chain order: ran1 offsets of every move's rotation, translation
    // and Metropolis values (for the stepr/stepm flags in plan_flags), attempts per type
    bool plan_ready = false;
    int plan_flags = 0, plan_total = 0;
    std::vector<int> att, type_nat;
    int *d_roff = nullptr, *d_toff = nullptr, *d_moff = nullptr;
    // pinned ran1 windows: pin[b] holds win_len[b] values of the stream (0: none) and
    // win_ck[b] the state at every 256th of them, from the window's start; a cycle reads its
    // stream in the window where its ran1 state lies.  pin_ahead: the window the previous
    // call uploaded to d_stream[pin_ahead] (complete when ahead_uploaded, behind ev_up)
    float *pin[3] = {nullptr, nullptr, nullptr};
    size_t pin_cap = 0;
    int win_len[3] = {0, 0, 0};
    std::vector<int32_t> win_ck[3];
    int pin_ahead = -1;
    float *d_stream[3] = {nullptr, nullptr, nullptr};
    size_t dstream_cap = 0;
    bool ahead_uploaded = false;
    hipEvent_t ev_up = nullptr;
    // a host thread drawing window gen_buf (gen_len values) for the cycle after next
    std::thread gen;
    int gen_buf = -1, gen_len = 0;
    hipError_t gen_err = hipSuccess;   // of the thread's upload, when it uploads
    hipStream_t up_stream = nullptr;   // the stream of such uploads
    // pinned block of a cycle's small transfers (results down, step sizes up)
    int *hio = nullptr;
    size_t hio_cap = 0;
    // cycles run back to back (epsr_molfit_cycles): the device block of constants, step
    // sizes and per-cycle results
    int *d_bres = nullptr;
    size_t d_bres_cap = 0;
    bool ledger_ready = false;
    int maxcnt = 0;   // the longest list
    // table extents measured for the arrays in tab_ptr (xstar, xavs, xavs1, ecoul, ecoul1)
    bool tab_ok = false;
    const void *tab_ptr[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int tab_ld = 0, tab_npt2 = 0, tab_last = 0, tab_lastnear = 0;
    // their packing (valid when tab_packed)
    float4 *d_tab = nullptr;
    size_t cap_tab = 0;
    bool tab_packed = false;
    // slots (not the first of a list) whose partner comes earlier in the chain order
    int *d_bvslot = nullptr;
    int nbv = 0;
    size_t cap_bv = 0;
    // neighbour offsets were formed for this lattice size and rmax
    int nbr_na = 0;
    float nbr_rmax = -1.0f;
    // the cell arrays of cells_init ([offsets a, b, c][molecules per cell][start of each
    // cell in the chain order, ncell + 2][cell of each molecule]), the pinned staging
    // buffer of the offsets' upload (ev_stage: after it) and a small pinned block for
    // fmolec1's scalars
    int *d_cells = nullptr;
    size_t cap_cells = 0;
    bool off_uploaded = false;
    hipEvent_t ev_stage = nullptr;
    void *hstage = nullptr;
    size_t hstage_cap = 0;
    int *hsc = nullptr;
    // fmolec1's device work buffers (grow-only, kept between calls)
    void *wbuf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t wcap[5] = {0, 0, 0, 0, 0};
};

// wait for the ledger's stream-drawing thread, if one runs
inline void gen_join(epsr_ledger *l)
{
    if (l->gen.joinable()) l->gen.join();
}

// ------------------------------------------------------------------ errors

int set_error(epsr_ctx *ctx, int code, const char *fmt, ...);

#define HIP_CHECK(ctx, expr)                                                             \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_error((ctx), EPSR_ERR_HIP, "%s failed: %s (%s:%d)", #expr,         \
                             hipGetErrorString(e_), __FILE__, __LINE__);                  \
    } while (0)

#define ARG_CHECK(ctx, cond, ...)                                                        \
    do {                                                                                 \
        if (!(cond)) return set_error((ctx), EPSR_ERR_ARG, __VA_ARGS__);                  \
    } while (0)

// true while the context's stream is being captured into a graph: no host reads, no
// synchronisation, no allocation may happen then
bool ctx_capturing(epsr_ctx *ctx);
// synchronise the context's stream unless it is being captured
hipError_t ctx_sync(epsr_ctx *ctx);
// the context's side stream and its fork/join events, created on first use
int ctx_side(epsr_ctx *ctx, hipStream_t *out);
// wait for the work issued so far on the context's stream by polling an event (a short
// wait wakes sooner than a stream synchronisation)
int ctx_wait(epsr_ctx *ctx);

// scratch management (device); returns nullptr and sets the error on failure
void *ctx_scratch(epsr_ctx *ctx, size_t bytes);
void *ctx_pinned(epsr_ctx *ctx, size_t bytes);

// stream-ordered temporaries of one call, released (stream-ordered) on every return path
struct StreamTemp {
    epsr_ctx *ctx;
    std::vector<void *> p;
    explicit StreamTemp(epsr_ctx *c) : ctx(c) {}
    StreamTemp(const StreamTemp &) = delete;
    StreamTemp &operator=(const StreamTemp &) = delete;
    void *alloc(size_t bytes)
    {
        void *q = nullptr;
        if (hipMallocAsync(&q, bytes > 0 ? bytes : 1, ctx->stream) != hipSuccess) {
            set_error(ctx, EPSR_ERR_HIP, "hipMallocAsync of %zu bytes failed", bytes);
            return nullptr;
        }
        p.push_back(q);
        return q;
    }
    ~StreamTemp()
    {
        for (void *q : p) (void)hipFreeAsync(q, ctx->stream);
    }
};

// kernel timing with HIP events on the launch stream
void timer_begin(epsr_ctx *ctx, const char *family);
void timer_end(epsr_ctx *ctx, const char *family);

// ------------------------------------------------------------------ host ran1

struct Ran1 {
    int32_t mcall, idum, iy, iv[32];
};
// src/epsr26/library/auxiliary_routines.py:324-385 (Park-Miller + Bays-Durham shuffle)
float ran1_next(Ran1 &s);
void ran1_fill(Ran1 &s, float *out, long long n);   // n values of ran1_next, state in registers
void ran1_load(Ran1 &s, const int32_t *a);
void ran1_store(const Ran1 &s, int32_t *a);

// ------------------------------------------------------------------ device arithmetic
// The reference's float32 contract (src/epsr26/fortran/fmath.py).  All translation
// units are compiled with -ffp-contract=off and correctly rounded division and
// square root, so a*b+c is two roundings and x/y, sqrtf(x) are correctly rounded.

// NINT: round half away from zero (fmath.py:222-239); |x| < 2^31 on this path.
__device__ __forceinline__ int dev_nint(float x)
{
    float t = truncf(x);
    float f = x - t;   // exact
    int ti = (int)t;
    ti += f >= 0.5f ? 1 : 0;
    ti -= f <= -0.5f ? 1 : 0;
    return ti;
}

__device__ __forceinline__ float dev_anint(float x)
{
    // branch-free: t is integral, so (t + up) - dn is exact
    float t = truncf(x);
    float f = x - t;
    float up = f >= 0.5f ? 1.0f : 0.0f;
    float dn = f <= -0.5f ? 1.0f : 0.0f;
    return (t + up) - dn;
}

// minimagec, cubic branch (src/epsr26/library/minimagec.py:48-62)
__device__ __forceinline__ float dev_mic(float x) { return x - dev_anint(x); }

// dev_mic(x) where only its square is used: x - rint(x).  NINT (half away from zero) and
// rint (half to even) differ only at half-integers, where the two differences are +0.5
// and -0.5 (both exact), with the same square; elsewhere the two are the same float.
__device__ __forceinline__ float mic_sq(float x) { return x - rintf(x); }

// IFIX (fmath.py:252-260) for arguments known to lie in [0, 2^31)
__device__ __forceinline__ int dev_ifix_pos(float x) { return (int)truncf(x); }

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// XCD-aware work split.  MI355X dispatches workgroup b of a grid to XCD b % 8, and each XCD
// has an L2 of its own.  For a grid whose size is a multiple of 8 (xcd_grid), xcd_block()
// renumbers the workgroups so that every XCD holds one contiguous range of logical blocks,
// and xcd_range() gives every XCD one contiguous eighth of n items, walked by its
// workgroups in order: work that touches neighbouring data (chain-ordered ledger lists and
// the lists of their partners) then finds it in the L2 of the XCD that touched it last.
inline int xcd_grid(long long blocks) { return (int)(((blocks < 8 ? 8 : blocks) + 7) / 8 * 8); }
__device__ __forceinline__ int xcd_block() { return (int)((blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3)); }
struct XcdRange {
    long long lo, hi;   // this XCD's items [lo, hi)
    long long first, stride;   // this thread's first item and the stride of its loop
};
__device__ __forceinline__ XcdRange xcd_range(long long n)
{
    const long long x = blockIdx.x & 7u, nj = gridDim.x >> 3, j = blockIdx.x >> 3;
    XcdRange r;
    r.lo = n * x / 8;
    r.hi = n * (x + 1) / 8;
    r.first = r.lo + j * blockDim.x + threadIdx.x;
    r.stride = nj * blockDim.x;
    return r;
}
