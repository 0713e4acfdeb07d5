// ps4_lib.cpp -- host side of libps4.so: the C ABI of include/ps4.h.
//
// A context owns one HIP device: the resident factor tables (built once, like
// the table set-up of exhaustive/s4.c:436-447), the work-queue control block,
// the 4-set queue, the extension buffer and one default-priority stream.  A
// search cuts its intervals into superwindows ("items"), groups them into
// batches sized to the 4-set queue, and per batch launches reset -> k_rq ->
// k_ext on the stream, then reads back counters and extensions.  Extensions
// are rare (about 1e-5 per S), so turning them into the reference's result
// lines (M:/5:/6-set, exhaustive/s4.c:245-277) happens here on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "ps4.h"
#include "ps4_internal.h"

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;

extern "C" hipError_t ps4_launch_rq(hipStream_t st, u32 grid, u32 block, u32 tile, const ps4_item *items, u32 n_items,
                                    u32 bin_cap, u32 nb_alloc, ps4_ctl *ctl, u64 *ctr, u32 n_rids, u64 *scratch,
                                    ps4_quad *quads, u64 quad_cap, u32 cap_reps, u32 verify);
extern "C" hipError_t ps4_launch_ext(hipStream_t st, u32 grid, const ps4_quad *quads, u64 quad_cap, ps4_ctl *ctl,
                                     const u32 *oftab, u64 tab_n, const uint16_t *ndiv, const ps4_qrtab *qr,
                                     u64 *ctr, u32 n_rids, ps4_hit *hits, u32 hit_cap, u32 cap_divs, u64 *exq,
                                     int chunk);
extern "C" hipError_t ps4_launch_reset(hipStream_t st, ps4_ctl *ctl, u64 *ctr, u32 n_ctr);
extern "C" hipError_t ps4_launch_triple(hipStream_t st, const u64 *abc, u64 n, u64 d_max, const u64 *ftab,
                                        u64 tab_n, u64 *d_out, u32 cap, u32 *count, u32 *err);
extern "C" hipError_t ps4_launch_fill_u32(hipStream_t st, u32 *p, u32 n, u32 v);
extern "C" hipError_t ps4_noalloc_scratch(int device, void **ptr);
extern "C" unsigned ps4_noalloc_cap(void);
extern "C" hipError_t ps4_launch_elem(hipStream_t st, const ps4_quad *quads, const ps4_ctl *ctl, u64 quad_cap,
                                      u64 n_max, ps4_quad *out, u32 cap, u32 *count, u32 grid);
extern "C" hipError_t ps4_launch_sqtri(hipStream_t st, u64 n_max, unsigned long long *total);
extern "C" int ps4_rq_threads(void);
extern "C" int ps4_rq_tile(void);
extern "C" int ps4_ext_chunk(void);
extern "C" int ps4_ext_threads(void);
extern "C" int ps4_ext_queue_cap(void);
extern "C" int ps4_rq_window(void);
extern "C" int ps4_rq_nb_max(void);
extern "C" void *ps4_k_rq_ptr(u32 block, u32 tile);
extern "C" void *ps4_k_ext_ptr(int chunk);

namespace {

// ---------------------------------------------------------------------------
// exact host arithmetic for the result lines
// ---------------------------------------------------------------------------

u64 isqrt64(u64 n)
{
    u64 r = (u64)std::sqrt((double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) ++r;
    return r;
}

u64 isqrt128(u128 n)
{
    if ((n >> 64) == 0) return isqrt64((u64)n);
    u64 r = (u64)sqrtl((long double)n);
    while ((u128)r * r > n) --r;
    while ((u128)(r + 1) * (r + 1) <= n) ++r;
    return r;
}

bool is_square(u128 n)
{
    const u64 r = isqrt128(n);
    return (u128)r * r == n;
}

std::string dec(u128 v)
{
    char tmp[48];
    int n = 0;
    while (v >> 64) { tmp[n++] = (char)('0' + (int)(v % 10)); v /= 10; }   // 128-bit digits (rare)
    u64 w = (u64)v;
    do { tmp[n++] = (char)('0' + (int)(w % 10)); w /= 10; } while (w);     // 64-bit digits
    std::string s;
    while (n) s.push_back(tmp[--n]);
    return s;
}

// Decimal digits of v at p (no terminator); returns the end.
char *put_dec(char *p, u128 v)
{
    char tmp[48];
    int n = 0;
    while (v >> 64) { tmp[n++] = (char)('0' + (int)(v % 10)); v /= 10; }
    u64 w = (u64)v;
    do { tmp[n++] = (char)('0' + (int)(w % 10)); w /= 10; } while (w);
    while (n) *p++ = tmp[--n];
    return p;
}

// Append "tag [x...] S=..." with members largest first, and the verify note of
// exhaustive/s4.c:209-220, to out.  n <= 6 members.
void append_set(std::string &out, const char *tag, const u128 *xin, int n, u64 S)
{
    u128 x[6];
    for (int i = 0; i < n; ++i) {   // insertion sort, descending
        u128 v = xin[i];
        int j = i;
        while (j > 0 && x[j - 1] < v) { x[j] = x[j - 1]; --j; }
        x[j] = v;
    }
    bool ok = true;
    for (int i = 0; i < n && ok; ++i)
        for (int j = i + 1; j < n; ++j)
            if (x[i] == x[j] || !is_square(x[i] + x[j])) { ok = false; break; }
    char buf[512];
    char *p = buf;
    for (const char *t = tag; *t; ++t) *p++ = *t;
    *p++ = ' ';
    *p++ = '[';
    for (int i = 0; i < n; ++i) {
        if (i) *p++ = ',';
        p = put_dec(p, x[i]);
    }
    *p++ = ']';
    *p++ = ' ';
    *p++ = 'S';
    *p++ = '=';
    p = put_dec(p, S);
    if (!ok) for (const char *t = "  *** VERIFY FAILED ***"; *t; ++t) *p++ = *t;
    *p++ = '\n';
    out.append(buf, (size_t)(p - buf));
}

std::string format_set(const char *tag, std::vector<u128> x, u64 S)
{
    std::string s;
    append_set(s, tag, x.data(), (int)x.size(), S);
    return s;
}

// Prime factors of m ascending (trial division; m < 2^33 here).
void factor_into(u64 m, std::vector<u64> &p, std::vector<int> &e)
{
    for (u64 d = 2; m > 1; d += (d == 2 ? 1 : 2)) {
        if (d * d > m) d = m;
        if (m % d) continue;
        int k = 0;
        while (m % d == 0) { m /= d; ++k; }
        size_t i = 0;
        while (i < p.size() && p[i] != d) ++i;
        if (i == p.size()) { p.push_back(d); e.push_back(k); }
        else e[i] += k;
    }
}

struct Quad6 {
    u64 rab, rac, rad, rbc, rbd, rcd;
    u64 a, b, c, d, S;
};

Quad6 members(const u32 *r)
{
    Quad6 q;
    q.rab = r[0]; q.rac = r[1]; q.rad = r[2]; q.rbc = r[3]; q.rbd = r[4]; q.rcd = r[5];
    q.a = (q.rab * q.rab + q.rac * q.rac - q.rbc * q.rbc) / 2;
    q.b = q.rab * q.rab - q.a;
    q.c = q.rac * q.rac - q.a;
    q.d = q.rad * q.rad - q.a;
    q.S = q.a + q.b + q.c + q.d;
    return q;
}

}  // namespace

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

struct ps4_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    hipStream_t st2 = nullptr;   // triple queries of a serving step, beside the census batch on st
    uint32_t flags = 0;
    int64_t dump4_limit = 0;
    u32 cap_reps = 512, cap_divs = 16384, cap_ext = 64;
    u64 S_hi_max = 0, tab_n = 0;
    u32 n_cu = 0, rq_grid = 0, ext_grid = 0;
    // tuning: the grids above may be lowered to these resident maxima (which
    // sized the scratch); superwindow shift, batch fill, k_ext form threshold
    u32 rq_grid_max = 0, ext_grid_max = 0;
    // launch shapes: k_rq threads per workgroup and points per thread load, k_ext
    // tiles per work claim; rq_occ[i][j]: resident k_rq workgroups per CU of
    // block 128 << i, tile 2 << j
    u32 rq_block = 256, rq_tile = 4, ext_chunk = 4;
    u32 rq_occ[3][3] = {{0}};
    u32 sw_shift = 0;
    double batch_frac = 0.8;
    u32 W = 0, nb_alloc = 0, bin_cap = 0;
    u32 scratch_wgs = 0;       // workgroups the window-bin scratch holds
    u64 quad_cap = 0;
    u32 hit_cap = 0, items_cap = 0, ctr_cap = 0;
    // device
    u64 *d_ftab = nullptr;
    u32 *d_oftab = nullptr;   // odd n -> its prime powers (the extension kernel's factor table)
    uint16_t *d_ndiv = nullptr;
    ps4_qrtab *d_qr = nullptr;
    ps4_ctl *d_ctl = nullptr;
    u64 *d_ctr = nullptr;
    ps4_item *d_items = nullptr;
    u64 *d_scratch = nullptr;
    ps4_quad *d_quads = nullptr;
    u64 *d_exq = nullptr;      // per-workgroup exact-test queues of the extension kernel
    // Two of each, one per batch parity: the next batch is planned and
    // enqueued while the current one runs, so its host items, read-backs,
    // extensions and events must not share a buffer with the current one's.
    // ev[p]: before k_rq, after k_rq, after k_ext, after the read-backs.
    ps4_hit *d_hits[2] = {nullptr, nullptr};
    ps4_ctl *h_ctl[2] = {nullptr, nullptr};        // pinned host
    u64 *h_ctr[2] = {nullptr, nullptr};
    ps4_item *h_items[2] = {nullptr, nullptr};
    hipEvent_t ev[2][4] = {};
    ps4_hit *h_hits = nullptr;
    // state
    ps4_alert alert = nullptr;
    void *alert_user = nullptr;
    bool test6_fired = false;
    std::string last_error;
    // kernel time accumulated over the launches since the last reset (ms)
    double ms_rq = 0.0, ms_ext = 0.0;
    u64 n_batches = 0;
    u32 density_pct = 100;     // PS4_TEST_DENSITY_PCT: scales the planner's 4-set estimate (tests of the redo path)
    u64 n_redo = 0;            // batches redone after an overflow (since open)
    hipGraphExec_t gexec = nullptr;   // PS4_GRAPH: the batch sequence, updated per batch
    // triple queries
    u64 *d_tabc = nullptr, *d_tout = nullptr;
    u32 *d_tcnt = nullptr, *d_terr = nullptr;
    u32 tq_cap = 0, tq_out = 0;
    bool debug_sync = false;
    void *na_scratch = nullptr;   // this device's fixed scratch (ps4_extend_triple_noalloc)
    // element-range mode: largest member allowed (0: off), the device buffer of
    // the 4-sets that qualify in a batch, and those collected so far
    u64 elem_max = 0;
    ps4_quad *d_eout = nullptr;
    u32 *d_ecnt = nullptr;
    unsigned long long *d_etri = nullptr;
    std::vector<ps4_quad> elem_sets;
    // serving stream (ps4_submit_*, ps4_step): requests in arrival order
    struct StreamReq {
        u64 tag = 0;
        bool triple = false;
        u64 lo = 0, cur = 0, hi = 0;   // a range request: [lo, hi), searched up to cur
        u64 abc[3] = {0, 0, 0}, d_max = 0;
        ps4_counters c;                // counters so far
        std::string lines;             // result lines so far
    };
    std::vector<StreamReq> sq;
};

#define HIP_OK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ctx->last_error = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return PS4_E_HIP;                                                            \
        }                                                                                \
    } while (0)

namespace {

// Upper bound on 4-sets per S near S (observed 0.03 at 1e7 .. 0.31 at 1.2e13).
double quad_density(u64 S)
{
    const double l2 = std::log2((double)(S < 2 ? 2 : S));
    double d = 0.02 * l2 - 0.3;
    return d < 0.08 ? 0.08 : (d > 0.7 ? 0.7 : d);
}

int build_tables(ps4_ctx *ctx)
{
    // n < 2*isqrt(S_hi)+64 covers every root sum ra+rb (exhaustive/s4.c:440)
    const u64 n = 2 * isqrt64(ctx->S_hi_max) + 64;
    ctx->tab_n = n;
    std::vector<u32> spf(n, 0);
    for (u64 i = 2; i < n; ++i) {
        if (spf[i]) continue;
        spf[i] = (u32)i;
        if (i * i < n)
            for (u64 j = i * i; j < n; j += i)
                if (!spf[j]) spf[j] = (u32)i;
    }
    std::vector<u64> ftab(n, 0);
    std::vector<uint16_t> ndiv(n, 0);
    std::vector<u32> ex(n, 0);
    if (n > 1) ndiv[1] = 1;
    for (u64 i = 2; i < n; ++i) {
        const u32 p = spf[i];
        const u64 q = i / p;
        u32 e;
        u64 cof;
        if (q > 1 && spf[q] == p) { e = ex[q] + 1; cof = ftab[q] & 0xFFFFFFFFull; }
        else { e = 1; cof = q; }
        ex[i] = e;
        ftab[i] = cof | ((u64)p << 32) | ((u64)e << 59);
        ndiv[i] = (uint16_t)((e + 1) * ndiv[cof]);
    }
    ps4_qrtab qr;
    memset(&qr, 0, sizeof qr);
    auto isq = [](u32 x, u32 m) { for (u32 r = 0; r < m; ++r) if ((r * r) % m == x) return true; return false; };
    for (u32 a = 0; a < 64; ++a) for (u32 r = 0; r < 64; ++r) if (isq((r * r + a) % 64, 64)) qr.t64[a] |= 1ull << r;
    for (u32 a = 0; a < 63; ++a) for (u32 r = 0; r < 63; ++r) if (isq((r * r + a) % 63, 63)) qr.t63[a] |= 1ull << r;
    for (u32 a = 0; a < 65; ++a)
        for (u32 b = 0; b < 64; ++b) { const u32 r = (b + 33) % 65; if (isq((r * r + a) % 65, 65)) qr.t65[a] |= 1ull << b; }
    for (u32 a = 0; a < 11; ++a) for (u32 r = 0; r < 11; ++r) if (isq((r * r + a) % 11, 11)) qr.t11[a] |= 1u << r;
    for (u32 a = 0; a < 17; ++a) for (u32 r = 0; r < 17; ++r) if (isq((r * r + a) % 17, 17)) qr.t17[a] |= 1u << r;
    for (u32 a = 0; a < 19; ++a) for (u32 r = 0; r < 19; ++r) if (isq((r * r + a) % 19, 19)) qr.t19[a] |= 1u << r;
    for (u32 a = 0; a < 23; ++a) for (u32 r = 0; r < 12; ++r) if (isq((r * r + a) % 23, 23)) qr.t23[a] |= 1u << r;
    // the extension kernel's odd-factor table: for odd i < n, entry i / 2 holds
    // the distinct prime powers of i as p << 6 | e, ascending (at most 7 for
    // i < 2^26), then k | d(i) << 4 with k their number (one gather factors an
    // odd part completely)
    const u64 n_odd = (n + 1) / 2;
    std::vector<u32> oftab(8 * n_odd, 0);
    for (u64 i = 1; i < n; i += 2) {
        u32 *w = &oftab[8 * (i >> 1)];
        u32 k = 0;
        for (u64 m = i; m > 1; m = ftab[m] & 0xFFFFFFFFull) {
            if (k == 7) { ctx->last_error = "odd-factor table: more than 7 primes"; return PS4_E_INTERNAL; }
            w[k++] = (u32)(((ftab[m] >> 32) & 0x7FFFFFFull) << 6) | (u32)(ftab[m] >> 59);
        }
        w[7] = k | ((u32)ndiv[i] << 4);
    }
    HIP_OK(hipMalloc(&ctx->d_ftab, n * sizeof(u64)));
    HIP_OK(hipMalloc(&ctx->d_oftab, oftab.size() * sizeof(u32)));
    HIP_OK(hipMalloc(&ctx->d_ndiv, n * sizeof(uint16_t)));
    HIP_OK(hipMalloc(&ctx->d_qr, sizeof(ps4_qrtab)));
    HIP_OK(hipMemcpy(ctx->d_ftab, ftab.data(), n * sizeof(u64), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ctx->d_oftab, oftab.data(), oftab.size() * sizeof(u32), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ctx->d_ndiv, ndiv.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(ctx->d_qr, &qr, sizeof qr, hipMemcpyHostToDevice));
    return 0;
}

u64 env_u64(const char *name, u64 dflt)
{
    const char *s = getenv(name);
    return (s && *s) ? strtoull(s, nullptr, 10) : dflt;
}

// resident k_rq workgroups per CU of a launch shape (block 128/256/512, tile 2/4/8)
u32 rq_occ_of(const ps4_ctx *ctx, u32 block, u32 tile)
{
    const u32 i = block == 128 ? 0 : block == 256 ? 1 : 2, j = tile == 2 ? 0 : tile == 4 ? 1 : 2;
    return ctx->rq_occ[i][j];
}

int ensure_ctr(ps4_ctx *ctx, u32 n)
{
    if (n <= ctx->ctr_cap) return 0;
    u32 cap = std::max<u32>(n, 64);
    if (ctx->d_ctr) { (void)hipFree(ctx->d_ctr); ctx->d_ctr = nullptr; }
    for (u64 *&h : ctx->h_ctr)
        if (h) { (void)hipHostFree(h); h = nullptr; }
    HIP_OK(hipMalloc(&ctx->d_ctr, (size_t)cap * CT_N * sizeof(u64)));
    for (u64 *&h : ctx->h_ctr) HIP_OK(hipHostMalloc(&h, (size_t)cap * CT_N * sizeof(u64)));
    ctx->ctr_cap = cap;
    return 0;
}

// The representation kernel's window bins: one region per workgroup of the
// largest grid the tuning may select (rq_grid_max), so that lowering the grid,
// growing the bins, then raising the grid again never indexes past the end.
int alloc_scratch(ps4_ctx *ctx)
{
    if (ctx->d_scratch) { (void)hipFree(ctx->d_scratch); ctx->d_scratch = nullptr; }
    ctx->scratch_wgs = ctx->rq_grid_max;
    HIP_OK(hipMalloc(&ctx->d_scratch, (size_t)ctx->scratch_wgs * ctx->nb_alloc * ctx->bin_cap * sizeof(u64)));
    return 0;
}

// Superwindow length at S: a power of two near 2*sqrt(S), so walking the
// ~0.7*sqrt(S) lattice columns once costs a few operations per S.
u64 superwindow(const ps4_ctx *ctx, u64 S)
{
    const u64 target = 2 * isqrt64(S) + 1;
    u64 sw = ctx->W;
    const u64 swmax = (u64)ctx->W * ctx->nb_alloc;
    while (sw < target && sw < swmax) sw <<= 1;
    return std::max<u64>(sw >> ctx->sw_shift, ctx->W);
}

struct Interval {
    u64 lo, hi, cur;
    u32 rid;
};

struct Group {
    u32 rid;
    Quad6 q;
    u32 best;
    std::vector<u128> ext;
};

// Order of extensions inside an M: line: the divisor index t of the reference's
// mixed-radix divisor list of D (exhaustive/s4.c:149-156, 185-204).
void order_extensions(Group &g)
{
    if (g.ext.size() < 2) return;
    const Quad6 &q = g.q;
    u64 A, B, ra, rb;
    switch (g.best) {
    case 0: A = q.a; B = q.b; ra = q.rac; rb = q.rbc; break;
    case 1: A = q.c; B = q.d; ra = q.rac; rb = q.rad; break;
    case 2: A = q.a; B = q.c; ra = q.rab; rb = q.rbc; break;
    case 3: A = q.b; B = q.d; ra = q.rab; rb = q.rad; break;
    case 4: A = q.a; B = q.d; ra = q.rab; rb = q.rbd; break;
    default: A = q.b; B = q.c; ra = q.rab; rb = q.rac; break;
    }
    std::vector<u64> p;
    std::vector<int> e;
    factor_into(ra - rb, p, e);
    factor_into(ra + rb, p, e);
    std::vector<std::pair<u64, u128>> keyed;
    for (u128 x : g.ext) {
        const u64 u = isqrt128(x + A), w = isqrt128(x + B);
        u64 fd = u - w, t = 0, stride = 1;
        for (size_t i = 0; i < p.size(); ++i) {
            u64 dig = 0;
            while (fd % p[i] == 0) { fd /= p[i]; ++dig; }
            t += dig * stride;
            stride *= (u64)(e[i] + 1);
        }
        keyed.push_back({t, x});
    }
    std::sort(keyed.begin(), keyed.end(), [](const std::pair<u64, u128> &x, const std::pair<u64, u128> &y) { return x.first < y.first; });
    for (size_t i = 0; i < keyed.size(); ++i) g.ext[i] = keyed[i].second;
}

// What a finished batch leaves for the host: its interval rows, the 4-set
// queue (only when 4: lines are wanted) and the extensions found.
struct Pending {
    u32 rid_lo = 0, rid_hi = 0;
    std::vector<ps4_quad> qs;
    std::vector<ps4_hit> hits;
};

// The device sequence of one batch on the context's stream: the items copy,
// counter reset, k_rq, (k_elem), k_ext and the read-backs of the control block
// and counter rows into pinned memory.  Kernel time is taken by events around
// the two census kernels.  Returns once everything is enqueued.
int enqueue_batch(ps4_ctx *ctx, int bp, u32 n_items, u32 n_rids, u64 s_top, bool noext, bool graph, double quads_expected)
{
    const u32 n_ctr = n_rids * CT_N;
    if (ctx->rq_grid > ctx->scratch_wgs) { ctx->last_error = "k_rq grid above its scratch"; return PS4_E_INTERNAL; }
    HIP_OK(hipMemcpyAsync(ctx->d_items, ctx->h_items[bp], n_items * sizeof(ps4_item), hipMemcpyHostToDevice, ctx->st));
    // PS4_DEBUG_SYNC=1: wait after every launch and name the kernel that failed
    auto dbg = [&](const char *what) -> int {
        if (!ctx->debug_sync) return 0;
        fprintf(stderr, "ps4: %s (items %u)\n", what, n_items);
        fflush(stderr);
        const hipError_t e = hipStreamSynchronize(ctx->st);
        if (e != hipSuccess) {
            ctx->last_error = std::string("after ") + what + ": " + hipGetErrorString(e);
            return PS4_E_HIP;
        }
        return 0;
    };
    HIP_OK(ps4_launch_reset(ctx->st, ctx->d_ctl, ctx->d_ctr, n_ctr));
    if (int rc = dbg("k_reset")) return rc;
    if (!graph) HIP_OK(hipEventRecord(ctx->ev[bp][0], ctx->st));
    // A small batch launches only the workgroups it can keep busy: every
    // workgroup of the persistent kernels claims work by an atomic on one
    // counter, so idle ones only add to the contention (about 0.2 ms each for
    // the two kernels at full grids).  k_rq takes one item per claim; k_ext
    // one chunk of 4 tiles of 64 queued 4-sets (its queue is filled in whole
    // tiles, so the upper bound on 4-sets plus one tile per item bounds the chunks).
    const u32 g_rq = std::max<u32>(1u, std::min<u32>(ctx->rq_grid, n_items));
    const double chunks = (quads_expected + 64.0 * (double)n_items) / (64.0 * ctx->ext_chunk) + 1.0;
    const u32 g_ext = (u32)std::max<double>(1.0, std::min<double>((double)ctx->ext_grid, chunks));
    (void)s_top;
    HIP_OK(ps4_launch_rq(ctx->st, g_rq, ctx->rq_block, ctx->rq_tile, ctx->d_items, n_items, ctx->bin_cap, ctx->nb_alloc, ctx->d_ctl,
                         ctx->d_ctr, n_rids, ctx->d_scratch, ctx->d_quads, ctx->quad_cap, ctx->cap_reps,
                         (ctx->flags & PS4_VERIFY) ? 1u : 0u));
    if (!graph) HIP_OK(hipEventRecord(ctx->ev[bp][1], ctx->st));
    if (int rc = dbg("k_rq")) return rc;
    if (ctx->elem_max) {
        HIP_OK(ps4_launch_fill_u32(ctx->st, ctx->d_ecnt, 1, 0));
        HIP_OK(ps4_launch_elem(ctx->st, ctx->d_quads, ctx->d_ctl, ctx->quad_cap, ctx->elem_max, ctx->d_eout,
                               PS4_ELEM_CAP, ctx->d_ecnt, ctx->n_cu * 8));
    }
    if (!noext)
        HIP_OK(ps4_launch_ext(ctx->st, g_ext, ctx->d_quads, ctx->quad_cap,
                              ctx->d_ctl, ctx->d_oftab, ctx->tab_n, ctx->d_ndiv, ctx->d_qr, ctx->d_ctr, n_rids,
                              ctx->d_hits[bp], ctx->hit_cap, ctx->cap_divs, ctx->d_exq, (int)ctx->ext_chunk));
    if (!graph) HIP_OK(hipEventRecord(ctx->ev[bp][2], ctx->st));
    if (!noext)
        if (int rc = dbg("k_ext")) return rc;
    HIP_OK(hipMemcpyAsync(ctx->h_ctl[bp], ctx->d_ctl, sizeof(ps4_ctl), hipMemcpyDeviceToHost, ctx->st));
    HIP_OK(hipMemcpyAsync(ctx->h_ctr[bp], ctx->d_ctr, (size_t)n_ctr * sizeof(u64), hipMemcpyDeviceToHost, ctx->st));
    return 0;
}

// One batch launched: directly, or (graph) captured and launched as one
// hipGraph.  A failure inside the capture ends the capture and drops the
// partial graph, so the stream is usable again.
int launch_batch(ps4_ctx *ctx, int bp, u32 n_items, u32 n_rids, u64 s_top, bool noext, bool graph, double quads_expected)
{
    if (!graph) {
        if (int rc = enqueue_batch(ctx, bp, n_items, n_rids, s_top, noext, false, quads_expected)) return rc;
        HIP_OK(hipEventRecord(ctx->ev[bp][3], ctx->st));
        return 0;
    }
    HIP_OK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_batch(ctx, bp, n_items, n_rids, s_top, noext, true, quads_expected);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->st, &g);
    if (rc || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        if (!rc) ctx->last_error = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec);
        return rc ? rc : PS4_E_HIP;
    }
    // same node structure from batch to batch (only the arguments and
    // sizes change), so the executable graph is updated in place; it is
    // built anew when the structure changes (an update that fails)
    if (ctx->gexec) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res;
        if (hipGraphExecUpdate(ctx->gexec, g, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(ctx->gexec);
            ctx->gexec = nullptr;
        }
    }
    hipError_t e = hipSuccess;
    if (!ctx->gexec) e = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_OK(e);
    HIP_OK(hipEventRecord(ctx->ev[bp][0], ctx->st));
    HIP_OK(hipGraphLaunch(ctx->gexec, ctx->st));
    HIP_OK(hipEventRecord(ctx->ev[bp][1], ctx->st));
    HIP_OK(hipEventRecord(ctx->ev[bp][2], ctx->st));
    HIP_OK(hipEventRecord(ctx->ev[bp][3], ctx->st));
    return 0;
}

// The result lines of a finished batch (exhaustive/s4.c:237-277): "4:" lines
// in the reference's single-thread order (dump_left: per interval, the 4-set
// lines still to print), and the extensions grouped per 4-set into M:, 5: and
// 6-set lines, with the 5-set histogram, 5-set and multi counts and the
// extension maximum added to out[rid].  text[r - rid_lo] gets interval r's lines.
int make_lines(ps4_ctx *ctx, const Pending &pb, std::vector<int64_t> &dump_left, ps4_counters *out,
               std::vector<std::string> &text)
{
    const u32 rid_lo = pb.rid_lo, rid_hi = pb.rid_hi;
    text.assign(rid_hi - rid_lo + 1, std::string());
    // ---- "4:" lines (exhaustive/s4.c:237-241), in the reference's single-thread order
    if (!pb.qs.empty()) {
        const u64 nq = pb.qs.size();
        const std::vector<ps4_quad> &qs = pb.qs;
        struct K { u64 S, px, py, pz; u32 cls; size_t i; };
        std::vector<K> keys;
        keys.reserve(nq);
        for (u64 i = 0; i < nq; ++i) {
            const ps4_quad &x = qs[i];
            if (x.rid == PS4_RID_NONE) continue;   // padding of a window's block
            if (x.rid < rid_lo || x.rid > rid_hi) { ctx->last_error = "4-set queue entry of another batch"; return PS4_E_INTERNAL; }
            const u64 S = (u64)x.r_ab * x.r_ab + (u64)x.r_cd * x.r_cd;
            const u32 cls = x.r_ad < x.r_bc ? 1u : 0u;
            keys.push_back({S, x.r_cd, x.r_bd, cls ? x.r_ad : x.r_bc, cls, (size_t)i});
        }
        std::sort(keys.begin(), keys.end(), [&](const K &x, const K &y) {
            if (qs[x.i].rid != qs[y.i].rid) return qs[x.i].rid < qs[y.i].rid;
            if (x.S != y.S) return x.S < y.S;
            if (x.px != y.px) return x.px < y.px;
            if (x.py != y.py) return x.py < y.py;
            if (x.pz != y.pz) return x.pz < y.pz;
            return x.cls < y.cls;
        });
        char line[160];
        for (const K &k : keys) {
            const ps4_quad &x = qs[k.i];
            if (dump_left[x.rid] <= 0) continue;
            --dump_left[x.rid];
            const u32 r6[6] = {x.r_ab, x.r_ac, x.r_ad, x.r_bc, x.r_bd, x.r_cd};
            const Quad6 q = members(r6);
            snprintf(line, sizeof line, "4: [%lld,%lld,%lld,%lld] S=%llu\n", (long long)q.a, (long long)q.b,
                     (long long)q.c, (long long)q.d, (unsigned long long)q.S);
            text[x.rid - rid_lo] += line;
        }
    }
    // ---- extensions -> M:, 5: and 6-set lines (exhaustive/s4.c:245-277)
    const u32 nh = (u32)pb.hits.size();
    if (nh) {
        const ps4_hit *hits = pb.hits.data();
        // order: interval, S, then the 4-set's roots (one group per 4-set)
        struct HK { u64 S, k1, k2, k3; u32 rid, i; };   // roots < 2^26: two per key word
        std::vector<HK> hk(nh);
        for (u32 i = 0; i < nh; ++i) {
            const ps4_hit &h = hits[i];
            const u64 S = (u64)h.r[0] * h.r[0] + (u64)h.r[5] * h.r[5];
            hk[i] = {S, ((u64)h.r[0] << 26) | h.r[1], ((u64)h.r[2] << 26) | h.r[3], ((u64)h.r[4] << 26) | h.r[5],
                     h.rid, i};
        }
        std::sort(hk.begin(), hk.end(), [](const HK &x, const HK &y) {
            if (x.rid != y.rid) return x.rid < y.rid;
            if (x.S != y.S) return x.S < y.S;
            if (x.k1 != y.k1) return x.k1 < y.k1;
            if (x.k2 != y.k2) return x.k2 < y.k2;
            return x.k3 < y.k3;
        });
        std::vector<u32> idx(nh);
        for (u32 i = 0; i < nh; ++i) idx[i] = hk[i].i;
        auto key = [&](const ps4_hit &h) { return std::make_tuple(h.rid, h.r[0], h.r[1], h.r[2], h.r[3], h.r[4], h.r[5]); };
        for (u32 i = 0; i < nh;) {
            const ps4_hit &h0 = hits[idx[i]];
            Group g;
            g.rid = h0.rid;
            g.q = members(h0.r);
            g.best = h0.pad;
            u32 j = i;
            while (j < nh && key(hits[idx[j]]) == key(h0)) {
                const ps4_hit &h = hits[idx[j]];
                g.ext.push_back(((u128)h.e_hi << 64) | h.e_lo);
                ++j;
            }
            i = j;
            ps4_counters &o = out[g.rid];
            const u64 ne = g.ext.size();
            if (ne > ctx->cap_ext) { ctx->last_error = "extension capacity exceeded"; return PS4_E_CAPACITY; }
            o.max_extensions = std::max(o.max_extensions, ne);
            order_extensions(g);
            std::string &t = text[g.rid - rid_lo];
            const Quad6 &q = g.q;
            if (ne >= 2) {
                ++o.multi;
                std::string m = "M: [" + std::to_string((unsigned long long)q.a) + "," + std::to_string((unsigned long long)q.b) +
                                "," + std::to_string((unsigned long long)q.c) + "," + std::to_string((unsigned long long)q.d) +
                                "] ext=[";
                for (size_t k = 0; k < ne; ++k) { if (k) m += ","; m += dec(g.ext[k]); }
                m += "] S=" + std::to_string((unsigned long long)q.S) + "\n";
                t += m;
                if ((ctx->flags & PS4_TEST6) && !ctx->test6_fired) {
                    ctx->test6_fired = true;
                    const std::string l = format_set("TEST-6SET (NOT A SOLUTION, --test-sixset drill):",
                                                     {q.a, q.b, q.c, q.d, g.ext[0], g.ext[1]}, q.S);
                    if (ctx->alert) ctx->alert(ctx->alert_user, l.c_str()); else t += l;
                }
            }
            for (size_t k = 0; k < ne; ++k) {
                if (g.ext[k] <= q.a) continue;
                ++o.sets5;
                int lg = 0;
                for (u128 x = g.ext[k]; x > 1; x >>= 1) ++lg;
                ++o.hist5[lg > 99 ? 99 : lg];
                const u128 v5[5] = {q.a, q.b, q.c, q.d, g.ext[k]};
                append_set(t, "5:", v5, 5, q.S);
            }
            for (size_t k = 0; k < ne; ++k)
                for (size_t s = k + 1; s < ne; ++s)
                    if (is_square(g.ext[k] + g.ext[s])) {
                        ++o.sets6;
                        const std::string l = format_set("!!!!!!!! FOUND 6-SET:", {q.a, q.b, q.c, q.d, g.ext[k], g.ext[s]}, q.S);
                        if (ctx->alert) ctx->alert(ctx->alert_user, l.c_str()); else t += l;
                    }
        }
    }
    return 0;
}

int run_job(ps4_ctx *ctx, size_t n, const u64 *lo_in, const u64 *hi_in, ps4_sink sink, void *const *users,
            void *single_user, ps4_counters *out)
{
    std::vector<Interval> iv(n);
    for (size_t i = 0; i < n; ++i) {
        u64 lo = lo_in[i] < 1 ? 1 : lo_in[i];
        u64 hi = hi_in[i];
        if (hi < lo) { ctx->last_error = "S_hi < S_lo"; return PS4_E_ARG; }
        if (hi > ctx->S_hi_max) { ctx->last_error = "S_hi beyond the tables of this context (S_hi_max)"; return PS4_E_ARG; }
        iv[i] = {lo, hi, lo, (u32)i};
        memset(&out[i], 0, sizeof(ps4_counters));
    }
    if (int rc = ensure_ctr(ctx, (u32)n)) return rc;
    std::vector<int64_t> dump_left(n, ctx->dump4_limit);
    const bool noext = (ctx->flags & PS4_NOEXT) != 0;
    // PS4_GRAPH: each batch's sequence is captured and launched as one graph
    const bool graph = (ctx->flags & PS4_GRAPH) && !ctx->debug_sync;
    bool want_dump_any = false;
    for (size_t r = 0; r < n; ++r) want_dump_any |= dump_left[r] > 0;
    // Two batches in flight: the next one is planned and enqueued before the
    // host waits for the current one, so the device does not idle between
    // them.  One at a time where a batch's results are read from a buffer the
    // next batch would overwrite (the 4-set dump, the element mode), and for
    // graph launches and the debugging mode.
    const size_t depth = (want_dump_any || ctx->elem_max || graph || ctx->debug_sync) ? 1 : 2;
    double budget_scale = 1.0;
    size_t next_iv = 0;
    int parity = 0;
    // a batch in flight: its intervals, buffer parity, and the interval state
    // before it was planned (restored if it has to be redone)
    struct Flight {
        u32 rid_lo = 0, rid_hi = 0;
        int bp = 0;
        size_t next_iv = 0;
        std::vector<Interval> save;
    };
    std::deque<Flight> fl;
    // the lines of a finished batch are made on the host while the next batch
    // runs on the device (batches emit in order)
    Pending pending;
    bool have_pending = false;
    auto emit = [&](const Pending &pb) -> int {
        const u32 rid_lo = pb.rid_lo, rid_hi = pb.rid_hi;
        std::vector<std::string> text;
        if (int rc = make_lines(ctx, pb, dump_left, out, text)) return rc;
        if (sink)
            for (u32 r = rid_lo; r <= rid_hi; ++r) {
                const std::string &t = text[r - rid_lo];
                if (t.empty()) continue;
                void *u = users ? users[r] : single_user;
                if (sink(u, t.data(), t.size())) { ctx->last_error = "sink returned nonzero"; return PS4_E_SINK; }
            }
        return 0;
    };
    // an error after a launch: let the device finish before returning
    auto fail = [&](int rc) { (void)hipStreamSynchronize(ctx->st); return rc; };
    for (;;) {
        // ---- plan and enqueue batches of superwindows until depth are in flight
        while (fl.size() < depth) {
            while (next_iv < n && iv[next_iv].cur >= iv[next_iv].hi) ++next_iv;
            if (next_iv >= n) break;
            Flight f;
            f.bp = parity;
            parity ^= 1;
            f.next_iv = next_iv;
            f.save.assign(iv.begin() + next_iv, iv.end());
            ps4_item *items = ctx->h_items[f.bp];
            double quads_expected = 0.0;
            // element mode: a batch's 4-sets must also fit the buffer of qualifying ones
            double quads_budget = ctx->batch_frac *
                                  (double)(ctx->elem_max ? std::min<u64>(ctx->quad_cap, PS4_ELEM_CAP) : ctx->quad_cap) *
                                  budget_scale;
            // the planner's estimate of 4-sets per S (an upper bound; scaled down
            // only by the test hook that drills the redo path)
            const double dscale = (double)ctx->density_pct / 100.0;
            auto density = [&](u64 S) { return quad_density(S) * dscale; };
            // A job's last batch has no device work to hide its host work (the
            // lines), so a large job ends with a small batch: when what is left
            // fits in this batch but is more than twice the tail (a sixth of a
            // full batch, at most 1.4*10^8 4-sets: about 2 ms of host lines at
            // 10^10), this batch leaves the tail for the last.  Only below
            // S = 10^11: higher up a batch has few lines (5-sets are 10x rarer
            // per S at 10^12 than at 10^10), and a small batch there needs short
            // superwindows, whose column walks cost about what long ones do.
            {
                const double tail = std::min(quads_budget / 6.0, 1.4e8);
                double left = 0.0;
                u64 top = 0;
                for (size_t i = next_iv; i < n; ++i) {
                    left += density(iv[i].hi) * (double)(iv[i].hi - iv[i].cur);
                    top = std::max(top, iv[i].hi);
                }
                // (not in the element mode: its lines are made after the last batch)
                if (!ctx->elem_max && top <= 100000000000ull && left <= quads_budget && left > 2.0 * tail)
                    quads_budget = left - tail;
            }
            u32 n_items = 0;
            u64 s_top = 0;   // largest S of the batch
            u32 rid_lo = iv[next_iv].rid, rid_hi = rid_lo;
            // little work left (small queries): shorter superwindows, so that the
            // representation kernel still has about one per workgroup
            u64 remaining = 0;
            for (size_t i = next_iv; i < n; ++i) remaining += iv[i].hi - iv[i].cur;
            u64 item_max = ctx->W;
            // null entries an item may leave in the 4-set queue: each wave of the
            // representation kernel ends an item with up to two partly used tiles
            const double item_pad = 128.0 * (double)(ctx->rq_block / 64);
            while (item_max * ctx->rq_grid < remaining && item_max < ((u64)1 << 40)) item_max <<= 1;
            // at least two superwindows per representation workgroup in a batch: at
            // high S a batch's 4-set budget covers fewer S, and superwindows of
            // ~2 sqrt(S) would leave workgroups idle (at S ~ 9*10^11: 820 per batch
            // for 1024 workgroups)
            auto sw_at = [&](u64 S) {
                u64 sw = superwindow(ctx, S);
                const double span = quads_budget / density(S);
                while (sw > ctx->W && (double)sw * 2.0 * (double)ctx->rq_grid > span) sw >>= 1;
                return sw;
            };
            for (size_t i = next_iv; i < n && n_items < ctx->items_cap && quads_expected < quads_budget; ++i) {
                Interval &v = iv[i];
                while (v.cur < v.hi && n_items < ctx->items_cap && quads_expected < quads_budget) {
                    u64 len = std::min(sw_at(v.cur), item_max);
                    if (len > v.hi - v.cur) len = v.hi - v.cur;
                    const double dq = density(v.cur + len) * (double)len + item_pad;
                    if (n_items > 0 && quads_expected + dq > quads_budget) {
                        // shrink the last item to the remaining budget, in whole windows
                        const double left = quads_budget - quads_expected - item_pad;
                        u64 room = left > 0.0 ? (u64)(left / density(v.cur + len)) : 0;
                        room = room / ctx->W * ctx->W;
                        if (room == 0) { quads_expected = quads_budget; break; }
                        len = room;
                    }
                    items[n_items++] = {v.cur, v.cur + len, v.rid, 0};
                    s_top = std::max(s_top, v.cur + len);
                    quads_expected += density(v.cur + len) * (double)len + item_pad;
                    v.cur += len;
                    rid_hi = v.rid;
                }
            }
            f.rid_lo = rid_lo;
            f.rid_hi = rid_hi;
            if (int rc = launch_batch(ctx, f.bp, n_items, rid_hi + 1, s_top, noext, graph, quads_expected)) return fail(rc);
            fl.push_back(std::move(f));
        }
        if (fl.empty()) break;
        if (have_pending) {
            have_pending = false;
            if (int rc = emit(pending)) return fail(rc);
        }
        // ---- the oldest batch in flight: wait for it and its read-backs
        Flight f = std::move(fl.front());
        fl.pop_front();
        const int bp = f.bp;
        HIP_OK(hipEventSynchronize(ctx->ev[bp][3]));
        float t_rq = 0.f, t_ext = 0.f;
        HIP_OK(hipEventElapsedTime(&t_rq, ctx->ev[bp][0], ctx->ev[bp][1]));
        HIP_OK(hipEventElapsedTime(&t_ext, ctx->ev[bp][1], ctx->ev[bp][2]));
        if (graph) { t_rq += t_ext; t_ext = 0.f; }   // one figure for the whole graph
        ctx->ms_rq += t_rq;
        ctx->ms_ext += t_ext;
        ++ctx->n_batches;
        const ps4_ctl *hc = ctx->h_ctl[bp];
        const u32 err = hc->err;
        if (err & ERR_CAP_REPS) { ctx->last_error = "representation capacity exceeded"; return fail(PS4_E_CAPACITY); }
        if (err & ERR_CAP_DIVS) { ctx->last_error = "divisor capacity exceeded"; return fail(PS4_E_CAPACITY); }
        if (err & ERR_VERIFY) { ctx->last_error = "quadruple verification failed"; return fail(PS4_E_VERIFY); }
        if (err & (ERR_ARENA | ERR_FACTORS | ERR_LDS_OVF | ERR_BADQUAD | ERR_BOUNDS)) {
            char buf[96];
            snprintf(buf, sizeof buf, "kernel capacity flag 0x%x", err);
            ctx->last_error = buf;
            return fail(PS4_E_INTERNAL);
        }
        if (err & (ERR_QUAD_OVF | ERR_HIT_OVF | ERR_BIN_OVF)) {
            // never truncate: redo this batch smaller (or with larger point bins);
            // the batch enqueued after it ran on the intervals as planned after
            // this one, so it is dropped and planned again
            HIP_OK(hipStreamSynchronize(ctx->st));
            fl.clear();
            std::copy(f.save.begin(), f.save.end(), iv.begin() + f.next_iv);
            next_iv = f.next_iv;
            if (err & ERR_BIN_OVF) {
                if (ctx->bin_cap >= 4 * ctx->W) { ctx->last_error = "window point bins overflow"; return PS4_E_INTERNAL; }
                ctx->bin_cap *= 2;
                if (int rc = alloc_scratch(ctx)) return rc;
            }
            if (err & (ERR_QUAD_OVF | ERR_HIT_OVF)) budget_scale *= 0.5;
            if (budget_scale < 1e-6) { ctx->last_error = "batch cannot shrink further"; return PS4_E_INTERNAL; }
            ++ctx->n_redo;
            continue;
        }
        // ---- counters
        for (u32 r = f.rid_lo; r <= f.rid_hi; ++r) {
            const u64 *c = ctx->h_ctr[bp] + (size_t)r * CT_N;
            ps4_counters &o = out[r];
            o.S_done += c[CT_S_DONE];
            o.pairs += c[CT_PAIRS];
            o.S_ge3 += c[CT_S_GE3];
            o.triples += c[CT_TRIPLES];
            o.sets4 += c[CT_SETS4];
            o.ext_cand += c[CT_EXT_CAND];
            o.ext_sqrt += c[CT_EXT_SQRT];
            o.max_reps = std::max(o.max_reps, c[CT_MAX_REPS]);
            o.max_divisors = std::max(o.max_divisors, c[CT_MAX_DIVS]);
        }
        // ---- element mode (one batch in flight): the batch's 4-sets with largest member <= elem_max
        if (ctx->elem_max) {
            u32 ne = 0;
            HIP_OK(hipMemcpy(&ne, ctx->d_ecnt, sizeof ne, hipMemcpyDeviceToHost));
            if (ne > PS4_ELEM_CAP) { ctx->last_error = "element-mode buffer overflow"; return fail(PS4_E_INTERNAL); }
            const size_t old = ctx->elem_sets.size();
            ctx->elem_sets.resize(old + ne);
            if (ne)
                HIP_OK(hipMemcpy(ctx->elem_sets.data() + old, ctx->d_eout, (size_t)ne * sizeof(ps4_quad),
                                 hipMemcpyDeviceToHost));
        }
        // ---- this batch's 4-set dump (one batch in flight) and extensions, to be
        // turned into lines while the next batch runs on the device (the copies
        // read buffers that batch does not write)
        Pending cur;
        cur.rid_lo = f.rid_lo;
        cur.rid_hi = f.rid_hi;
        bool want_dump = false;
        for (u32 r = f.rid_lo; r <= f.rid_hi; ++r) want_dump |= dump_left[r] > 0;
        if (want_dump && hc->quad_count) {
            cur.qs.resize(hc->quad_count);
            HIP_OK(hipMemcpy(cur.qs.data(), ctx->d_quads, cur.qs.size() * sizeof(ps4_quad), hipMemcpyDeviceToHost));
        }
        if (const u32 nh = hc->hit_count) {
            HIP_OK(hipMemcpy(ctx->h_hits, ctx->d_hits[bp], (size_t)nh * sizeof(ps4_hit), hipMemcpyDeviceToHost));
            cur.hits.assign(ctx->h_hits, ctx->h_hits + nh);
        }
        pending = std::move(cur);
        have_pending = true;
        budget_scale = std::min(1.0, budget_scale * 2.0);
    }
    if (have_pending)
        if (int rc = emit(pending)) return rc;
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" const char *ps4_strerror(int rc)
{
    switch (rc) {
    case PS4_OK: return "ok";
    case PS4_E_ARG: return "invalid argument";
    case PS4_E_CAPACITY: return "capacity exceeded; the interval is NOT certified complete";
    case PS4_E_HIP: return "HIP runtime error";
    case PS4_E_OOM: return "out of memory";
    case PS4_E_VERIFY: return "quadruple verification failed";
    case PS4_E_SINK: return "output sink failed";
    case PS4_E_INTERNAL: return "internal error";
    default: return "unknown error";
    }
}

extern "C" const char *ps4_last_error(ps4_ctx *ctx) { return ctx ? ctx->last_error.c_str() : ""; }

extern "C" void ps4_close(ps4_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    if (ctx->st2) (void)hipStreamSynchronize(ctx->st2);
    void *dev[] = {ctx->d_ftab, ctx->d_oftab, ctx->d_ndiv, ctx->d_qr, ctx->d_ctl, ctx->d_ctr, ctx->d_items, ctx->d_scratch,
                   ctx->d_quads, ctx->d_hits[0], ctx->d_hits[1], ctx->d_exq, ctx->d_eout, ctx->d_ecnt, ctx->d_etri,
                   ctx->d_tabc, ctx->d_tout, ctx->d_tcnt, ctx->d_terr};
    for (void *p : dev) if (p) (void)hipFree(p);
    void *host[] = {ctx->h_ctl[0], ctx->h_ctl[1], ctx->h_ctr[0], ctx->h_ctr[1], ctx->h_items[0], ctx->h_items[1], ctx->h_hits};
    for (void *p : host) if (p) (void)hipHostFree(p);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    for (auto &evs : ctx->ev)
        for (hipEvent_t e : evs) if (e) (void)hipEventDestroy(e);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    if (ctx->st2) (void)hipStreamDestroy(ctx->st2);
    delete ctx;
}

static int open_impl(const ps4_config *cfg, ps4_ctx *ctx)
{
    ctx->device = cfg->device;
    ctx->flags = cfg->flags;
    ctx->dump4_limit = cfg->dump4_limit;
    if (cfg->cap_reps) ctx->cap_reps = cfg->cap_reps;
    if (cfg->cap_divs) ctx->cap_divs = cfg->cap_divs;
    // the extension kernel holds the divisor lists of any 4-set with d(D) <= 16384
    // in its LDS arena (at most 748 entries); a larger cap would not fit
    if (ctx->cap_divs > 16384) { ctx->last_error = "cap_divs above 16384 exceeds the divisor-list arena"; return PS4_E_ARG; }
    if (cfg->cap_ext) ctx->cap_ext = cfg->cap_ext;
    ctx->S_hi_max = cfg->S_hi_max;
    // 2^50: keeps every root below 2^26 (the odd-factor table's entries and the
    // kernels' 32-bit root arithmetic) and every divisor product an exact double
    if (ctx->S_hi_max < 2 || ctx->S_hi_max > (1ull << 50)) { ctx->last_error = "S_hi_max out of range"; return PS4_E_ARG; }
    HIP_OK(hipSetDevice(ctx->device));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, ctx->device));
    ctx->n_cu = (u32)prop.multiProcessorCount;
    HIP_OK(hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&ctx->st2, hipStreamNonBlocking));
    HIP_OK(ps4_noalloc_scratch(ctx->device, &ctx->na_scratch));
    for (auto &evs : ctx->ev)
        for (hipEvent_t &e : evs) HIP_OK(hipEventCreate(&e));
    if (int rc = build_tables(ctx)) return rc;
    // resident workgroups per CU of every compiled launch shape; the window-bin
    // scratch holds the largest resident grid of them all
    u32 occ_max = 1;
    for (u32 i = 0; i < 3; ++i)
        for (u32 j = 0; j < 3; ++j) {
            int b = 0;
            HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, ps4_k_rq_ptr(128u << i, 2u << j), (int)(128u << i), 0));
            ctx->rq_occ[i][j] = (u32)std::max(1, b);
            occ_max = std::max(occ_max, ctx->rq_occ[i][j]);
        }
    int b_ext = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b_ext, ps4_k_ext_ptr(ps4_ext_chunk()), ps4_ext_threads(), 0));
    ctx->rq_block = (u32)ps4_rq_threads();
    ctx->rq_tile = (u32)ps4_rq_tile();
    ctx->ext_chunk = (u32)ps4_ext_chunk();
    ctx->rq_grid_max = (u32)env_u64("PS4_RQ_GRID", (u64)ctx->n_cu * occ_max);
    ctx->rq_grid = std::min(ctx->rq_grid_max, ctx->n_cu * rq_occ_of(ctx, ctx->rq_block, ctx->rq_tile));
    ctx->ext_grid = (u32)env_u64("PS4_EXT_GRID", (u64)ctx->n_cu * (u64)std::max(1, b_ext));
    ctx->ext_grid_max = ctx->ext_grid;
    ctx->W = (u32)ps4_rq_window();
    ctx->debug_sync = env_u64("PS4_DEBUG_SYNC", 0) != 0;
    ctx->density_pct = (u32)std::min<u64>(100, std::max<u64>(1, env_u64("PS4_TEST_DENSITY_PCT", 100)));
    // superwindows up to 2*sqrt(S_hi_max), at most 256 windows
    u32 nb = 1;
    while ((u64)ctx->W * nb < 2 * isqrt64(ctx->S_hi_max) + 1 && nb < 512) nb <<= 1;
    ctx->nb_alloc = std::min<u32>(nb, (u32)ps4_rq_nb_max());
    ctx->bin_cap = ctx->W / 2 + 1024;
    // 4-set queue: enough for the whole range when it is small, else 2^31 entries
    // (64 GiB of the 288 GB: [1,10^10) in two batches and a small last one; each
    // batch boundary costs the two kernels' tails), at most 40% of the free
    // device memory
    const double whole = quad_density(ctx->S_hi_max) * (double)ctx->S_hi_max * 1.5 + 65536.0;
    u64 qcap = env_u64("PS4_QUAD_CAP", 1ull << 31);
    if ((double)qcap > whole) qcap = (u64)whole;
    size_t fr = 0, tot = 0;
    HIP_OK(hipMemGetInfo(&fr, &tot));
    while (qcap > (1u << 16) && (double)qcap * sizeof(ps4_quad) > 0.4 * (double)fr) qcap >>= 1;
    ctx->quad_cap = qcap;
    ctx->hit_cap = 1u << 20;
    ctx->items_cap = 1u << 20;
    HIP_OK(hipMalloc(&ctx->d_ctl, sizeof(ps4_ctl)));
    HIP_OK(hipMalloc(&ctx->d_items, (size_t)ctx->items_cap * sizeof(ps4_item)));
    for (int bp = 0; bp < 2; ++bp) {
        HIP_OK(hipHostMalloc(&ctx->h_ctl[bp], sizeof(ps4_ctl)));
        HIP_OK(hipHostMalloc(&ctx->h_items[bp], (size_t)ctx->items_cap * sizeof(ps4_item)));
        HIP_OK(hipMalloc(&ctx->d_hits[bp], (size_t)ctx->hit_cap * sizeof(ps4_hit)));
    }
    HIP_OK(hipMalloc(&ctx->d_quads, (size_t)ctx->quad_cap * sizeof(ps4_quad)));
    HIP_OK(hipMalloc(&ctx->d_exq, (size_t)ctx->ext_grid * (size_t)ps4_ext_queue_cap() * sizeof(u64)));
    HIP_OK(hipHostMalloc(&ctx->h_hits, (size_t)ctx->hit_cap * sizeof(ps4_hit)));
    if (int rc = alloc_scratch(ctx)) return rc;
    if (int rc = ensure_ctr(ctx, 64)) return rc;
    return 0;
}

extern "C" int ps4_open(const ps4_config *cfg, ps4_ctx **ctx_out)
{
    if (!cfg || !ctx_out) return PS4_E_ARG;
    *ctx_out = nullptr;
    ps4_ctx *ctx = new (std::nothrow) ps4_ctx();
    if (!ctx) return PS4_E_OOM;
    const int rc = open_impl(cfg, ctx);
    if (rc) {
        fprintf(stderr, "ps4_open: %s\n", ctx->last_error.c_str());
        ps4_close(ctx);
        return rc;
    }
    *ctx_out = ctx;
    return 0;
}

extern "C" int ps4_get_tuning(ps4_ctx *ctx, ps4_tuning *cur, ps4_tuning *max)
{
    if (!ctx) return PS4_E_ARG;
    if (cur) {
        memset(cur, 0, sizeof *cur);
        cur->rq_grid = ctx->rq_grid;
        cur->ext_grid = ctx->ext_grid;
        cur->rq_block = ctx->rq_block;
        cur->sw_shift = ctx->sw_shift;
        cur->batch_pct = (u32)std::lround(ctx->batch_frac * 100.0);
        cur->n_cu = ctx->n_cu;
        cur->rq_tile = ctx->rq_tile;
        cur->ext_chunk = ctx->ext_chunk;
    }
    if (max) {
        memset(max, 0, sizeof *max);
        max->rq_grid = ctx->rq_grid_max;
        max->ext_grid = ctx->ext_grid_max;
        max->rq_block = 512;
        max->sw_shift = PS4_SW_SHIFT_MAX;
        max->batch_pct = PS4_BATCH_PCT_MAX;
        max->n_cu = ctx->n_cu;
        max->rq_tile = 8;
        max->ext_chunk = 16;
    }
    return 0;
}

extern "C" int ps4_set_tuning(ps4_ctx *ctx, const ps4_tuning *t)
{
    if (!ctx || !t) return PS4_E_ARG;
    const bool shape_ok = (t->rq_block == 0 || t->rq_block == 128 || t->rq_block == 256 || t->rq_block == 512) &&
                          (t->rq_tile == 0 || t->rq_tile == 2 || t->rq_tile == 4 || t->rq_tile == 8) &&
                          (t->ext_chunk == 0 || t->ext_chunk == 1 || t->ext_chunk == 4 || t->ext_chunk == 16);
    if (t->rq_grid > ctx->rq_grid_max || t->ext_grid > ctx->ext_grid_max || t->sw_shift > PS4_SW_SHIFT_MAX ||
        t->batch_pct > PS4_BATCH_PCT_MAX || !shape_ok) {
        ctx->last_error = "tuning value above its maximum or not a compiled launch shape (ps4_get_tuning)";
        return PS4_E_ARG;
    }
    ctx->rq_block = t->rq_block ? t->rq_block : (u32)ps4_rq_threads();
    ctx->rq_tile = t->rq_tile ? t->rq_tile : (u32)ps4_rq_tile();
    ctx->ext_chunk = t->ext_chunk ? t->ext_chunk : (u32)ps4_ext_chunk();
    ctx->rq_grid = t->rq_grid ? t->rq_grid : std::min(ctx->rq_grid_max, ctx->n_cu * rq_occ_of(ctx, ctx->rq_block, ctx->rq_tile));
    ctx->ext_grid = t->ext_grid ? t->ext_grid : ctx->ext_grid_max;
    ctx->sw_shift = t->sw_shift;
    ctx->batch_frac = (t->batch_pct ? t->batch_pct : 80u) / 100.0;
    return 0;
}

extern "C" int ps4_parse_tuning(const char *text, ps4_tuning *t)
{
    if (!text || !t) return PS4_E_ARG;
    memset(t, 0, sizeof *t);
    const char *p = text;
    while (*p) {
        const char *eol = strchr(p, '\n');
        std::string line(p, eol ? (size_t)(eol - p) : strlen(p));
        p = eol ? eol + 1 : p + line.size();
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        auto trim = [](std::string x) {
            const size_t a = x.find_first_not_of(" \t\r"), b = x.find_last_not_of(" \t\r");
            return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
        };
        line = trim(line);
        if (line.empty()) continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos) return PS4_E_ARG;
        const std::string key = trim(line.substr(0, eq)), val = trim(line.substr(eq + 1));
        if (val.empty() || val.find_first_not_of("0123456789") != std::string::npos || val.size() > 10) return PS4_E_ARG;
        const u64 v = strtoull(val.c_str(), nullptr, 10);
        if (v > 0xFFFFFFFFull) return PS4_E_ARG;
        if (key == "rq_grid") t->rq_grid = (u32)v;
        else if (key == "ext_grid") t->ext_grid = (u32)v;
        else if (key == "rq_block") t->rq_block = (u32)v;
        else if (key == "sw_shift") t->sw_shift = (u32)v;
        else if (key == "batch_pct") t->batch_pct = (u32)v;
        else if (key == "rq_tile") t->rq_tile = (u32)v;
        else if (key == "ext_chunk") t->ext_chunk = (u32)v;
        else return PS4_E_ARG;
    }
    return 0;
}

extern "C" int ps4_set_dump4(ps4_ctx *ctx, int64_t limit)
{
    if (!ctx) return PS4_E_ARG;
    ctx->dump4_limit = limit;
    return 0;
}

extern "C" int ps4_set_alert(ps4_ctx *ctx, ps4_alert alert, void *user)
{
    if (!ctx) return PS4_E_ARG;
    ctx->alert = alert;
    ctx->alert_user = user;
    return 0;
}

extern "C" int ps4_search(ps4_ctx *ctx, uint64_t S_lo, uint64_t S_hi, ps4_sink sink, void *user, ps4_counters *out)
{
    if (!ctx || !out) return PS4_E_ARG;
    HIP_OK(hipSetDevice(ctx->device));
    return run_job(ctx, 1, &S_lo, &S_hi, sink, nullptr, user, out);
}

extern "C" int ps4_search_batch(ps4_ctx *ctx, size_t n, const uint64_t *lo, const uint64_t *hi, ps4_sink sink,
                                void *const *users, ps4_counters *out)
{
    if (!ctx || !out || (n && (!lo || !hi))) return PS4_E_ARG;
    if (n == 0) return 0;
    HIP_OK(hipSetDevice(ctx->device));
    return run_job(ctx, n, lo, hi, sink, users, nullptr, out);
}

// Kernel time (ms, HIP events on the context's stream) accumulated since the
// last reset, and the number of batches; reset when reset != 0.
extern "C" uint64_t ps4_redo_count(ps4_ctx *ctx) { return ctx ? ctx->n_redo : 0; }

extern "C" int ps4_kernel_time(ps4_ctx *ctx, double *ms_rq, double *ms_ext, uint64_t *batches, int reset)
{
    if (!ctx) return PS4_E_ARG;
    if (ms_rq) *ms_rq = ctx->ms_rq;
    if (ms_ext) *ms_ext = ctx->ms_ext;
    if (batches) *batches = ctx->n_batches;
    if (reset) { ctx->ms_rq = ctx->ms_ext = 0.0; ctx->n_batches = 0; }
    return 0;
}

// Triple queries, n at once (one workgroup each on the device).  abc holds 3n
// values; query i writes its count to n_out[i] and up to cap values to
// d_out[i*cap ...], ascending.
namespace {

// A triple of distinct positive integers with square pair sums, inside the
// tables of the context?  (PS4_E_ARG with the reason otherwise.)
int check_triple(ps4_ctx *ctx, const uint64_t *abc)
{
    u64 v[3] = {abc[0], abc[1], abc[2]};
    std::sort(v, v + 3);
    if (v[0] == 0 || v[0] == v[1] || v[1] == v[2] || !is_square((u128)v[0] + v[1]) || !is_square((u128)v[0] + v[2]) ||
        !is_square((u128)v[1] + v[2])) {
        ctx->last_error = "not a triple of distinct positive integers with square pair sums";
        return PS4_E_ARG;
    }
    if (2 * isqrt64(v[2] + v[0]) + 2 >= ctx->tab_n) { ctx->last_error = "triple beyond the tables of this context"; return PS4_E_ARG; }
    return 0;
}

const u32 TQ_PER = 1u << 16;   // answers one triple query may stage

// Enqueue n checked triple queries on stream st (their counts and answers land
// in the context's triple buffers, grown here when n exceeds them).
int triple_enqueue(ps4_ctx *ctx, hipStream_t st, size_t n, const uint64_t *abc, uint64_t d_max)
{
    if (n > ctx->tq_cap) {
        HIP_OK(hipStreamSynchronize(ctx->st));
        HIP_OK(hipStreamSynchronize(ctx->st2));
        void *old[] = {ctx->d_tabc, ctx->d_tout, ctx->d_tcnt, ctx->d_terr};
        for (void *p : old) if (p) (void)hipFree(p);
        const u32 c = (u32)std::max<size_t>(n, 64);
        HIP_OK(hipMalloc(&ctx->d_tabc, (size_t)c * 3 * sizeof(u64)));
        HIP_OK(hipMalloc(&ctx->d_tout, (size_t)c * TQ_PER * sizeof(u64)));
        HIP_OK(hipMalloc(&ctx->d_tcnt, (size_t)c * sizeof(u32)));
        HIP_OK(hipMalloc(&ctx->d_terr, sizeof(u32)));
        ctx->tq_cap = c;
    }
    HIP_OK(hipMemcpyAsync(ctx->d_tabc, abc, n * 3 * sizeof(u64), hipMemcpyHostToDevice, st));
    HIP_OK(ps4_launch_fill_u32(st, ctx->d_tcnt, (u32)n, 0u));
    HIP_OK(ps4_launch_fill_u32(st, ctx->d_terr, 1u, 0u));
    HIP_OK(ps4_launch_triple(st, ctx->d_tabc, n, d_max, ctx->d_ftab, ctx->tab_n, ctx->d_tout, TQ_PER, ctx->d_tcnt,
                             ctx->d_terr));
    return 0;
}

// After triple_enqueue on st: wait, then query i's answers ascending into
// ans[i] (all of them) and its count into n_out[i].
int triple_collect(ps4_ctx *ctx, hipStream_t st, size_t n, std::vector<std::vector<u64>> &ans)
{
    std::vector<u32> cnt(n);
    u32 terr = 0;
    HIP_OK(hipMemcpyAsync(cnt.data(), ctx->d_tcnt, n * sizeof(u32), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(&terr, ctx->d_terr, sizeof(u32), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (terr) { ctx->last_error = "divisor lists of a triple exceed the kernel's arena"; return PS4_E_INTERNAL; }
    ans.assign(n, std::vector<u64>());
    for (size_t i = 0; i < n; ++i) {
        if (cnt[i] > TQ_PER) { ctx->last_error = "too many solutions for one triple"; return PS4_E_CAPACITY; }
        if (!cnt[i]) continue;
        ans[i].resize(cnt[i]);
        HIP_OK(hipMemcpy(ans[i].data(), ctx->d_tout + i * TQ_PER, cnt[i] * sizeof(u64), hipMemcpyDeviceToHost));
        std::sort(ans[i].begin(), ans[i].end());
    }
    return 0;
}

}  // namespace

// Triple queries, n at once (one workgroup each on the device).  abc holds 3n
// values; query i writes its count to n_out[i] and up to cap values to
// d_out[i*cap ...], ascending.
extern "C" int ps4_extend_triple_batch(ps4_ctx *ctx, size_t n, const uint64_t *abc, uint64_t d_max,
                                       uint64_t *d_out, size_t cap, size_t *n_out)
{
    if (!ctx || !abc || !n_out || (cap && !d_out)) return PS4_E_ARG;
    if (n == 0) return 0;
    HIP_OK(hipSetDevice(ctx->device));
    for (size_t i = 0; i < n; ++i)
        if (int rc = check_triple(ctx, abc + 3 * i)) return rc;
    if (int rc = triple_enqueue(ctx, ctx->st, n, abc, d_max)) return rc;
    std::vector<std::vector<u64>> ans;
    if (int rc = triple_collect(ctx, ctx->st, n, ans)) return rc;
    for (size_t i = 0; i < n; ++i) {
        n_out[i] = ans[i].size();
        for (size_t k = 0; k < ans[i].size() && k < cap; ++k) d_out[i * cap + k] = ans[i][k];
    }
    return 0;
}

extern "C" int ps4_extend_triple(ps4_ctx *ctx, const uint64_t abc[3], uint64_t d_max, uint64_t *d_out, size_t cap,
                                 size_t *n_out)
{
    return ps4_extend_triple_batch(ctx, 1, abc, d_max, d_out, cap, n_out);
}

// ---------------------------------------------------------------------------
// serving stream: continuous batching over one context
// ---------------------------------------------------------------------------

namespace {

void add_counters(ps4_counters &a, const ps4_counters &b)
{
    a.S_done += b.S_done; a.pairs += b.pairs; a.S_ge3 += b.S_ge3; a.triples += b.triples; a.sets4 += b.sets4;
    a.sets5 += b.sets5; a.sets6 += b.sets6; a.multi += b.multi; a.ext_cand += b.ext_cand; a.ext_sqrt += b.ext_sqrt;
    for (int i = 0; i < 100; ++i) a.hist5[i] += b.hist5[i];
    a.max_reps = std::max(a.max_reps, b.max_reps);
    a.max_divisors = std::max(a.max_divisors, b.max_divisors);
    a.max_extensions = std::max(a.max_extensions, b.max_extensions);
}

int append_sink(void *user, const char *lines, size_t n)
{
    ((std::string *)user)->append(lines, n);
    return 0;
}

}  // namespace

extern "C" int ps4_submit_range(ps4_ctx *ctx, uint64_t tag, uint64_t S_lo, uint64_t S_hi)
{
    if (!ctx) return PS4_E_ARG;
    const u64 lo = S_lo < 1 ? 1 : S_lo;
    if (S_hi < lo) { ctx->last_error = "S_hi < S_lo"; return PS4_E_ARG; }
    if (S_hi > ctx->S_hi_max) { ctx->last_error = "S_hi beyond the tables of this context (S_hi_max)"; return PS4_E_ARG; }
    ps4_ctx::StreamReq r;
    r.tag = tag;
    r.lo = r.cur = lo;
    r.hi = S_hi;
    memset(&r.c, 0, sizeof r.c);
    ctx->sq.push_back(std::move(r));
    return 0;
}

extern "C" int ps4_submit_triple(ps4_ctx *ctx, uint64_t tag, const uint64_t abc[3], uint64_t d_max)
{
    if (!ctx || !abc) return PS4_E_ARG;
    if (int rc = check_triple(ctx, abc)) return rc;
    ps4_ctx::StreamReq r;
    r.tag = tag;
    r.triple = true;
    for (int i = 0; i < 3; ++i) r.abc[i] = abc[i];
    r.d_max = d_max;
    memset(&r.c, 0, sizeof r.c);
    ctx->sq.push_back(std::move(r));
    return 0;
}

// One step of the stream: the range requests in order of their remaining
// work (shortest first) take the next part of their interval as long as the
// step's expected 4-sets fit max_quads and the free part of the 4-set queue
// (a request larger than what is left is cut, so every step is short), all
// run as one batch of shared persistent launches; the pending triple queries
// of one d_max (up to the triple buffers, at most 256) run beside them on a
// second stream.  Requests finished by this step are handed to done.
extern "C" int ps4_step(ps4_ctx *ctx, uint64_t max_quads, ps4_done_fn done, void *user, ps4_stream_state *state)
{
    if (!ctx) return PS4_E_ARG;
    HIP_OK(hipSetDevice(ctx->device));
    std::vector<ps4_ctx::StreamReq> &q = ctx->sq;
    const double free_quads = ctx->batch_frac * (double)ctx->quad_cap;
    // default budget 2^26 expected 4-sets (about 1.3*10^8 S at 10^12): the
    // largest requests of the serving workload (10^8 S) fit in one step.
    // Cutting them into steps of ~1 ms of work (2^23) cost more in per-step
    // work (launches, the kernels' tails, a column walk per superwindow) than
    // it saved the small requests: p99 16.0 -> 7.9 ms at 200 requests/s, p50
    // unchanged (1.05 -> 1.02 ms).
    double budget = max_quads ? std::min((double)max_quads, free_quads) : std::min(free_quads, 67108864.0);
    // ---- admission: triples of the oldest pending triple's d_max
    std::vector<size_t> tri;
    for (size_t i = 0; i < q.size() && tri.size() < 256; ++i)
        if (q[i].triple && (tri.empty() || q[i].d_max == q[tri[0]].d_max)) tri.push_back(i);
    // ---- admission: range parts, shortest remaining first
    std::vector<size_t> order;
    for (size_t i = 0; i < q.size(); ++i)
        if (!q[i].triple) order.push_back(i);
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return q[x].hi - q[x].cur < q[y].hi - q[y].cur; });
    std::vector<size_t> rng;
    std::vector<u64> los, his;
    for (size_t i : order) {
        ps4_ctx::StreamReq &r = q[i];
        u64 len = r.hi - r.cur;
        const double dens = quad_density(r.hi);
        if (len > 0 && dens * (double)len > budget) {
            if (!rng.empty() && budget < dens * (double)ctx->W) break;
            len = std::max<u64>((u64)(budget / dens) / ctx->W * ctx->W, ctx->W);
            len = std::min(len, r.hi - r.cur);
        }
        rng.push_back(i);
        los.push_back(r.cur);
        his.push_back(r.cur + len);
        budget -= dens * (double)len;
        if (budget <= 0.0) break;
    }
    // ---- device work: the triples on the second stream, the range parts as one batched search
    std::vector<std::vector<u64>> tans;
    std::vector<u64> tabc(3 * tri.size());
    for (size_t k = 0; k < tri.size(); ++k)
        for (int j = 0; j < 3; ++j) tabc[3 * k + j] = q[tri[k]].abc[j];
    if (!tri.empty())
        if (int rc = triple_enqueue(ctx, ctx->st2, tri.size(), tabc.data(), q[tri[0]].d_max)) return rc;
    std::vector<ps4_counters> cs(rng.size());
    std::vector<std::string> txt(rng.size());
    std::vector<void *> users(rng.size());
    for (size_t k = 0; k < rng.size(); ++k) users[k] = &txt[k];
    int rc_r = 0;
    if (!rng.empty()) {
        const int64_t dump4 = ctx->dump4_limit;
        ctx->dump4_limit = 0;
        rc_r = run_job(ctx, rng.size(), los.data(), his.data(), append_sink, users.data(), nullptr, cs.data());
        ctx->dump4_limit = dump4;
    }
    int rc_t = tri.empty() ? 0 : triple_collect(ctx, ctx->st2, tri.size(), tans);
    // a failed part is run again alone, so that only the request it belongs to fails
    std::vector<int> rc_each(rng.size(), 0);
    if (rc_r) {
        for (size_t k = 0; k < rng.size(); ++k) {
            txt[k].clear();
            void *u = &txt[k];
            const int64_t dump4 = ctx->dump4_limit;
            ctx->dump4_limit = 0;
            rc_each[k] = run_job(ctx, 1, &los[k], &his[k], append_sink, &u, nullptr, &cs[k]);
            ctx->dump4_limit = dump4;
        }
    }
    std::vector<int> rc_tri(tri.size(), rc_t);
    if (rc_t) {   // one query at a time, likewise
        for (size_t k = 0; k < tri.size(); ++k) {
            std::vector<std::vector<u64>> one;
            rc_tri[k] = triple_enqueue(ctx, ctx->st2, 1, &tabc[3 * k], q[tri[k]].d_max);
            if (!rc_tri[k]) rc_tri[k] = triple_collect(ctx, ctx->st2, 1, one);
            if (!rc_tri[k]) { tans.resize(tri.size()); tans[k] = one[0]; }
        }
    }
    // ---- completions
    std::vector<char> gone(q.size(), 0);
    for (size_t k = 0; k < rng.size(); ++k) {
        ps4_ctx::StreamReq &r = q[rng[k]];
        if (!rc_each[k]) {
            add_counters(r.c, cs[k]);
            r.lines += txt[k];
            r.cur = his[k];
        }
        if (rc_each[k] || r.cur >= r.hi) {
            ps4_done d;
            memset(&d, 0, sizeof d);
            d.tag = r.tag;
            d.rc = rc_each[k];
            d.kind = 0;
            d.counters = r.c;
            d.lines = r.lines.c_str();
            d.nbytes = r.lines.size();
            if (done) done(user, &d);
            gone[rng[k]] = 1;
        }
    }
    for (size_t k = 0; k < tri.size(); ++k) {
        ps4_done d;
        memset(&d, 0, sizeof d);
        d.tag = q[tri[k]].tag;
        d.rc = rc_tri[k];
        d.kind = 1;
        if (!rc_tri[k]) {
            d.d = tans[k].data();
            d.n_d = tans[k].size();
        }
        if (done) done(user, &d);
        gone[tri[k]] = 1;
    }
    size_t w = 0;
    for (size_t i = 0; i < q.size(); ++i)
        if (!gone[i]) {
            if (w != i) q[w] = std::move(q[i]);
            ++w;
        }
    q.resize(w);
    if (state) {
        memset(state, 0, sizeof *state);
        for (const auto &r : q) (r.triple ? state->pending_triples : state->pending_ranges) += 1;
        state->quad_slots = (uint64_t)free_quads;
        state->hit_slots = ctx->hit_cap;
        state->triple_slots = std::max<u64>(ctx->tq_cap, 256);
    }
    return 0;
}

// The triple query without any allocation: the query, its count and its
// answers are staged in the device's fixed scratch (ps4_noalloc_scratch), the
// answers are copied into the caller's d_out and sorted there in place.
extern "C" int ps4_extend_triple_noalloc(ps4_ctx *ctx, const uint64_t abc[3], uint64_t d_max, uint64_t *d_out,
                                         size_t cap, size_t *n_out)
{
    if (!ctx || !abc || !n_out || (cap && !d_out)) return PS4_E_ARG;
    *n_out = 0;
    if (int rc = check_triple(ctx, abc)) return rc;
    if (!ctx->na_scratch) { ctx->last_error = "no fixed scratch on this device"; return PS4_E_INTERNAL; }
    HIP_OK(hipSetDevice(ctx->device));
    u64 *s_abc = (u64 *)ctx->na_scratch;
    u32 *s_ce = (u32 *)(s_abc + 3);          // count, error word
    u64 *s_out = s_abc + 8;
    const u32 na_cap = ps4_noalloc_cap();
    HIP_OK(hipMemcpyAsync(s_abc, abc, 3 * sizeof(u64), hipMemcpyHostToDevice, ctx->st));
    HIP_OK(ps4_launch_fill_u32(ctx->st, s_ce, 2u, 0u));   // the last call's count and error word
    HIP_OK(ps4_launch_triple(ctx->st, s_abc, 1, d_max, ctx->d_ftab, ctx->tab_n, s_out, na_cap, s_ce, s_ce + 1));
    u32 ce[2] = {0u, 0u};
    HIP_OK(hipMemcpyAsync(ce, s_ce, sizeof ce, hipMemcpyDeviceToHost, ctx->st));
    HIP_OK(hipStreamSynchronize(ctx->st));
    if (ce[1]) { ctx->last_error = "divisor lists of a triple exceed the kernel's arena"; return PS4_E_INTERNAL; }
    *n_out = ce[0];
    if (ce[0] > na_cap) { ctx->last_error = "more answers than the fixed scratch stages (65536)"; return PS4_E_CAPACITY; }
    if (!cap || !ce[0]) return 0;
    if (ce[0] > cap) { ctx->last_error = "d_out holds fewer values than the query has answers"; return PS4_E_CAPACITY; }
    HIP_OK(hipMemcpy(d_out, s_out, (size_t)ce[0] * sizeof(u64), hipMemcpyDeviceToHost));
    std::sort(d_out, d_out + ce[0]);
    return 0;
}


extern "C" int ps4_search_elements(ps4_ctx *ctx, uint64_t n_max, ps4_sink sink, void *user, uint64_t *n_sets,
                                   uint64_t *n_triples)
{
    if (!ctx) return PS4_E_ARG;
    if (!n_sets || !n_triples) { ctx->last_error = "null output"; return PS4_E_ARG; }
    *n_sets = 0;
    *n_triples = 0;
    if (n_max < 1) return 0;
    // a + b + c + d <= (n - 3) + (n - 2) + (n - 1) + n: the census of S < 4n - 5 holds them all
    // (no 4-set below n = 4)
    const u64 lo = 1, hi = n_max >= 4 ? 4 * n_max - 5 : 1;
    if (n_max > PS4_ELEM_NMAX || hi > ctx->S_hi_max) {
        ctx->last_error = "n_max beyond 10^10 or beyond the tables (S_hi_max >= 4 n_max - 6 is needed)";
        return PS4_E_ARG;
    }
    HIP_OK(hipSetDevice(ctx->device));
    if (!ctx->d_eout) {
        HIP_OK(hipMalloc(&ctx->d_eout, (size_t)PS4_ELEM_CAP * sizeof(ps4_quad)));
        HIP_OK(hipMalloc(&ctx->d_ecnt, sizeof(u32)));
        HIP_OK(hipMalloc(&ctx->d_etri, sizeof(unsigned long long)));
    }
    // the square triples
    HIP_OK(ps4_launch_fill_u32(ctx->st, (u32 *)ctx->d_etri, 2, 0));
    HIP_OK(ps4_launch_sqtri(ctx->st, n_max, ctx->d_etri));
    unsigned long long ntri = 0;
    HIP_OK(hipMemcpyAsync(&ntri, ctx->d_etri, sizeof ntri, hipMemcpyDeviceToHost, ctx->st));
    HIP_OK(hipStreamSynchronize(ctx->st));
    // the 4-sets: a census without extensions and without lines, filtered on the device
    const uint32_t flags = ctx->flags;
    const int64_t dump4 = ctx->dump4_limit;
    ctx->flags = PS4_NOEXT;
    ctx->dump4_limit = 0;
    ctx->elem_max = n_max;
    ctx->elem_sets.clear();
    ps4_counters c;
    const int rc = hi > lo ? run_job(ctx, 1, &lo, &hi, nullptr, nullptr, nullptr, &c) : 0;
    ctx->flags = flags;
    ctx->dump4_limit = dump4;
    ctx->elem_max = 0;
    if (rc) { ctx->elem_sets.clear(); return rc; }
    // members ascending, the list in ascending order of (a, b, c, d)
    std::vector<std::array<u64, 4>> sets;
    sets.reserve(ctx->elem_sets.size());
    for (const ps4_quad &x : ctx->elem_sets) {
        const u32 r6[6] = {x.r_ab, x.r_ac, x.r_ad, x.r_bc, x.r_bd, x.r_cd};
        const Quad6 q = members(r6);
        sets.push_back({(u64)q.d, (u64)q.c, (u64)q.b, (u64)q.a});
    }
    std::vector<ps4_quad>().swap(ctx->elem_sets);
    std::sort(sets.begin(), sets.end());
    *n_sets = sets.size();
    *n_triples = ntri;
    if (sink) {
        std::string t;
        for (size_t i = 0; i < sets.size(); ++i) {
            t += '[';
            for (int k = 0; k < 4; ++k) {
                char num[48];
                if (k) t += ',';
                t.append(num, (size_t)(put_dec(num, sets[i][k]) - num));
            }
            t += "]\n";
            if (t.size() >= (1u << 20) || i + 1 == sets.size()) {
                if (sink(user, t.data(), t.size())) { ctx->last_error = "sink returned nonzero"; return PS4_E_SINK; }
                t.clear();
            }
        }
    }
    return 0;
}
