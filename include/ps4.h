/*
 * ps4.h -- C ABI of the MI355X census engine (libps4.so).
 *
 * The reference (panjanek/pairwise-square-sums, exhaustive/s4.c) is a process
 * CLI with no FFI; these entry points are the library cut through it: what its
 * main() does per interval (exhaustive/s4.c:414-509) minus argv/stdio, with the
 * hot loops (process_chunk :280-342, handle_4set :225-278, extend4 :160-207)
 * running as HIP kernels on one gfx950 device.  The drop-in CLI `ps4`
 * (pairwise-square-sums_amd/csrc/ps4_cli.c) is the reference's own interface
 * rebuilt on top of this ABI.
 *
 * Plain C types only.  The caller owns every array it passes; text handed to a
 * sink belongs to the library and is valid only during the callback.  A
 * context is used from one host thread at a time.  Functions return 0 on
 * success or a negative PS4_E_* code; nothing is printed.
 */
#ifndef PS4_H
#define PS4_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (ps4_strerror gives the text). */
#define PS4_OK          0
#define PS4_E_ARG      -1   /* bad argument: S_hi <= S_lo, range beyond the tables, NULL pointer   */
#define PS4_E_CAPACITY -2   /* a reference capacity was exceeded: reps > cap_reps, divisors >
                               cap_divs or extensions > cap_ext (s4.c fatal(), exhaustive/s4.c:56-59,
                               176, 201, 321).  The interval is NOT certified complete.          */
#define PS4_E_HIP      -3   /* a HIP runtime call failed                                           */
#define PS4_E_OOM      -4   /* device or host allocation failed                                    */
#define PS4_E_VERIFY   -5   /* PS4_VERIFY found a 4-set whose six roots do not check
                               (exhaustive/s4.c:229-233)                                           */
#define PS4_E_SINK     -6   /* the sink callback returned nonzero                                  */
#define PS4_E_INTERNAL -7   /* an internal invariant failed                                        */

/* Flags of ps4_config.flags: the reference's switches (exhaustive/s4.c:416-423). */
#define PS4_NOEXT   0x1u    /* --noext: count 4-sets, do not extend (:242)                          */
#define PS4_VERIFY  0x2u    /* --verify: re-check the six roots of every 4-set (:229-233)           */
#define PS4_TEST6   0x4u    /* --test-sixset: exercise the 6-set alert on the first M: line (:255-260) */
#define PS4_GRAPH   0x8u    /* launch each batch's sequence (copy, reset, kernels, read-backs) as one
                               captured hipGraph, updated in place from batch to batch; the kernel
                               time is then reported as one figure (ms_rq).  No s4 counterpart    */

/* Opaque context: resident tables, device buffers, streams. */
typedef struct ps4_ctx ps4_ctx;

/* Counter vector: the stats_t fields the reference reports (exhaustive/s4.c:80-91)
 * with the aggregation of :488-506 (sums, and maxima for the three max_*). */
typedef struct {
    uint64_t S_done, pairs, S_ge3, triples, sets4, sets5, sets6, multi, ext_cand, ext_sqrt;
    uint64_t hist5[100];
    uint64_t max_reps, max_divisors, max_extensions;
} ps4_counters;

typedef struct {
    int device;            /* HIP device ordinal                                              */
    uint64_t S_hi_max;     /* largest S_hi any later call may use; sizes the factor tables
                              (exhaustive/s4.c:439-447: n < 2*isqrt(S_hi)+64)                   */
    uint32_t flags;        /* PS4_NOEXT | PS4_VERIFY | PS4_TEST6 | PS4_GRAPH                    */
    int64_t dump4_limit;   /* emit the first dump4_limit 4-sets per search as "4:" lines
                              (--dump4 N, :237-241; 0 = none)                                   */
    uint32_t cap_reps, cap_divs, cap_ext;   /* capacity limits; 0 = the reference's 512 / 16384 / 64
                                               (S4_MAX_REPS/DIVS/EXT, :45-54); cap_divs may not
                                               exceed 16384 (PS4_E_ARG)                           */
} ps4_config;

/* Receives result text in the reference's line grammar ("5: [...] S=..",
 * "M: [...] ext=[...] S=..", "!!!!!!!! FOUND 6-SET: [...]", "4: [...]";
 * exhaustive/s4.c:214-220, 239, 251-253, 274).  Return 0 to continue. */
typedef int (*ps4_sink)(void *user, const char *lines, size_t nbytes);

/* Receives each 6-set line (and the PS4_TEST6 drill line) the moment it is
 * found, before ordinary output: the durable-alert hook of exhaustive/s4.c:129-137. */
typedef void (*ps4_alert)(void *user, const char *line);

/* Open a context on cfg->device and build the resident factor tables for
 * S < cfg->S_hi_max.  Replaces the table set-up of exhaustive/s4.c:436-447. */
int ps4_open(const ps4_config *cfg, ps4_ctx **ctx_out);

/* Census of every S in [S_lo, S_hi): the work of one whole s4 invocation
 * (exhaustive/s4.c:449-506; per chunk process_chunk :280-342).  Result lines
 * go to sink (may be NULL), counters to *out.  S_lo < 1 is taken as 1 (:427). */
int ps4_search(ps4_ctx *ctx, uint64_t S_lo, uint64_t S_hi, ps4_sink sink, void *user, ps4_counters *out);

/* Many independent intervals [lo[i], hi[i]) in one persistent launch with a
 * device work-stealing queue; out[i] and sink(users[i], ...) equal what
 * ps4_search would give for interval i alone.  Intervals may overlap.  No
 * reference counterpart beyond calling s4 once per interval (the block
 * controller, run_exhaustive.py, drives s4 block by block). */
int ps4_search_batch(ps4_ctx *ctx, size_t n, const uint64_t *lo, const uint64_t *hi,
                     ps4_sink sink, void *const *users, ps4_counters *out);

/* Triple query: every d in [1, d_max] \ {a,b,c} with d+a, d+b, d+c square,
 * ascending.  The reference has no such entry point; it is extend4
 * (exhaustive/s4.c:160-207) applied to three known members.  *n_out receives
 * the full count; at most cap values are written. */
int ps4_extend_triple(ps4_ctx *ctx, const uint64_t abc[3], uint64_t d_max,
                      uint64_t *d_out, size_t cap, size_t *n_out);

/* The triple query of ps4_extend_triple without any allocation: the query, its
 * answer count and its answers are staged in a fixed device scratch of the
 * library (one per device, 65536 answers, 512 KiB, held until the process
 * ends), the answers are copied into d_out and sorted there.  No counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops).  *n_out receives the
 * count; a count above 65536 is PS4_E_CAPACITY, and so is a count above cap
 * when cap > 0 (d_out must hold every answer).  Calls that overlap on one
 * device share the scratch and race. */
int ps4_extend_triple_noalloc(ps4_ctx *ctx, const uint64_t abc[3], uint64_t d_max,
                              uint64_t *d_out, size_t cap, size_t *n_out);

/* ---- Serving: continuous batching over one context.  No reference
 * counterpart (s4 answers one interval per process, exhaustive/s4.c:416-426);
 * the census of a range request is that of ps4_search, a triple request that
 * of ps4_extend_triple.  Requests are queued by ps4_submit_*; each ps4_step
 * runs one short batch: the range requests in order of their remaining work
 * (shortest first) each take the next part of their interval while the
 * step's expected 4-sets fit its budget and the free part of the 4-set queue,
 * all in shared persistent launches, and the pending triple queries run
 * beside them on a second stream.  A request's result is the same as when it
 * runs alone. */
typedef struct {
    uint64_t tag;              /* the caller's id of the request                          */
    int rc;                    /* 0, or the PS4_E_* this request failed with              */
    int kind;                  /* 0 range, 1 triple                                       */
    ps4_counters counters;     /* range: as ps4_search gives them                         */
    const char *lines;         /* range: its result lines (valid during the callback)     */
    size_t nbytes;
    const uint64_t *d;         /* triple: every answer, ascending (valid during the call)  */
    size_t n_d;
} ps4_done;
typedef void (*ps4_done_fn)(void *user, const ps4_done *done);

typedef struct {
    uint64_t pending_ranges, pending_triples;   /* requests still queued after the step        */
    uint64_t quad_slots;       /* 4-set queue entries one step may fill                          */
    uint64_t hit_slots;        /* extension entries one step may hold                            */
    uint64_t triple_slots;     /* triple queries one step may hold                              */
} ps4_stream_state;

/* Queue a range request [S_lo, S_hi) (checked as by ps4_search). */
int ps4_submit_range(ps4_ctx *ctx, uint64_t tag, uint64_t S_lo, uint64_t S_hi);
/* Queue a triple request (checked as by ps4_extend_triple). */
int ps4_submit_triple(ps4_ctx *ctx, uint64_t tag, const uint64_t abc[3], uint64_t d_max);
/* Run one step; requests it finishes go to done(user, ...).  max_quads: the
 * step's budget of expected 4-sets (0: 2^26, a whole 10^8-wide request at 10^12);
 * *state (may be NULL) gets what remains queued and the step's capacities. */
int ps4_step(ps4_ctx *ctx, uint64_t max_quads, ps4_done_fn done, void *user, ps4_stream_state *state);

/* Install the durable-alert hook.  With a hook installed, 6-set lines and the
 * PS4_TEST6 drill line go to the hook only; without one they go to the sink. */
int ps4_set_alert(ps4_ctx *ctx, ps4_alert alert, void *user);

/* Change the "4:" line budget of later searches (--dump4 N, exhaustive/s4.c:237;
 * the reference's budget is per thread, a driver that mirrors its threads as
 * sub-intervals sets it per sub-interval). */
int ps4_set_dump4(ps4_ctx *ctx, int64_t limit);

/* Triple queries in one launch: abc holds 3n values, query i writes its count
 * to n_out[i] and its first cap values (ascending) to d_out[i*cap ...]. */
int ps4_extend_triple_batch(ps4_ctx *ctx, size_t n, const uint64_t *abc, uint64_t d_max,
                            uint64_t *d_out, size_t cap, size_t *n_out);

/* Element-range search: every 4-set {a < b < c < d} with all six pair sums
 * square and d <= n_max, written through the sink as "[a,b,c,d]\n" lines in
 * ascending order of (a, b, c, d); *n_sets gets their number and *n_triples
 * the number of square triples {a < b < c} (a+b, a+c, b+c square) with
 * c <= n_max.  The reference ranges over S only (exhaustive/s4.c:416-426);
 * this is the same census (process_chunk, :280-342) over S < 4 n_max - 5,
 * filtered by largest member, without extensions.  Needs S_hi_max >=
 * 4 n_max - 6 and n_max <= 10^10. */
int ps4_search_elements(ps4_ctx *ctx, uint64_t n_max, ps4_sink sink, void *user, uint64_t *n_sets,
                        uint64_t *n_triples);

/* Device time of the census kernels (ms, HIP events on the context's stream)
 * accumulated since the last reset, and the number of batches launched.  No
 * reference counterpart (the reference reports per-thread seconds,
 * exhaustive/s4.c:496). */
int ps4_kernel_time(ps4_ctx *ctx, double *ms_rq, double *ms_ext, uint64_t *batches, int reset);

/* Batches run again since ps4_open because their 4-sets or extensions
 * overflowed the device queues (each is redone smaller: results are never
 * truncated, as s4's FATAL checks never let a count pass silently,
 * exhaustive/s4.c:45-54).  A diagnostic; no reference counterpart. */
uint64_t ps4_redo_count(ps4_ctx *ctx);

/* Launch shape and batch plan of later searches: what the autotuner
 * (pairwise-square-sums_amd/ps4/autotune.py) varies.  Results never depend on
 * these, only the time does.  No reference counterpart: the shape knobs of s4
 * are its thread count and log2 chunk (exhaustive/s4.c:424-426, 451-456). */
typedef struct {
    uint32_t rq_grid;        /* workgroups of the persistent representation kernel (k_rq)          */
    uint32_t ext_grid;       /* single-wave workgroups of the extension kernel (k_ext)              */
    uint32_t rq_block;       /* threads per k_rq workgroup: 128, 256 or 512 (a compiled shape)      */
    uint32_t sw_shift;       /* subrange split: superwindows 2^-sw_shift of the default length (a
                                power of two near 2 sqrt(S)), never shorter than one 4096-S window  */
    uint32_t batch_pct;      /* a batch is planned to fill this percentage of the 4-set queue       */
    uint32_t n_cu;           /* the device's compute units (reported by ps4_get_tuning; ignored by
                                ps4_set_tuning)                                                     */
    uint32_t rq_tile;        /* window points a k_rq thread loads at once in its bucket passes:
                                2, 4 or 8                                                           */
    uint32_t ext_chunk;      /* tiles of 64 queued 4-sets a k_ext workgroup claims at once: 1, 4, 16 */
} ps4_tuning;

/* The values in force (*cur) and the largest each field may take (*max; grids:
 * the workgroups resident at once, which sized the scratch at ps4_open).
 * Either pointer may be NULL. */
int ps4_get_tuning(ps4_ctx *ctx, ps4_tuning *cur, ps4_tuning *max);

/* Set every field; a field of 0 selects its default (the resident maximum of
 * the shape for the grids, 256 / 4 / 4 for rq_block / rq_tile / ext_chunk, 0 for
 * sw_shift, 80 for batch_pct).  A value above its maximum, or a shape that is
 * not compiled, is PS4_E_ARG and changes nothing. */
int ps4_set_tuning(ps4_ctx *ctx, const ps4_tuning *t);

/* Parse a tuning file: "key = value" lines, keys named as the ps4_tuning
 * fields, '#' starts a comment; fields not named are 0.  PS4_E_ARG on an
 * unknown key or a malformed value.  Needs no context and no device. */
int ps4_parse_tuning(const char *text, ps4_tuning *t);

/* Close a context and free its device memory. */
void ps4_close(ps4_ctx *ctx);

/* Text for a PS4_E_* code. */
const char *ps4_strerror(int rc);

/* Detail of the last error on ctx (e.g. which capacity), or "" . */
const char *ps4_last_error(ps4_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* PS4_H */
