/*
 * ps4_cli.c -- the drop-in command line: the reference's s4 interface
 * (exhaustive/s4.c:414-509) with the census running on the GPU through
 * include/ps4.h.
 *
 *   ps4 S_lo S_hi [threads] [log2 chunk] [--out FILE] [--ckpt FILE] [--resume]
 *                 [--noext] [--verify] [--dump4 N] [--test-sixset] [--device N] [--tune FILE]
 *   ps4 --elements N [--out FILE] [--device N]
 *
 * --elements N (no s4 counterpart: the element-range form of the search) prints
 * every 4-set {a<b<c<d} with six square pair sums and d <= N as "[a,b,c,d]"
 * lines in ascending order, and on stderr one line
 * "ELEMENTS: n_max=N sets=... square_triples=..." (square triples {a<b<c}
 * with c <= N).
 *
 * --tune FILE (no s4 counterpart) applies a tuning file written by the
 * autotuner (ps4/autotune.py): launch shape and batch plan only, the output is
 * the same with or without it.
 *
 * Same result lines on stdout (or appended to --out), same stderr reports
 * (header, progress, per-thread lines, TOTAL, 5-set histogram, CAPACITY), same
 * s4ckpt v1 checkpoint, same exit codes (0 ok, 1 usage/checkpoint/IO, 2 FATAL).
 * "threads" keeps its meaning for the outputs: the interval is cut into that
 * many contiguous regions exactly as exhaustive/s4.c:451-456 does, and each
 * region gets its own counters, per-thread line and checkpoint position; the
 * regions are searched one after another on the device.  Default: the online
 * CPU count, as the reference's threads=0 (exhaustive/s4.c:428-431), so the
 * header and per-thread lines read as its would.
 * Checkpoints are written every 5 s (exhaustive/s4.c:470-478) between blocks
 * of about 2^33 S; PS4_CKPT_SECS=x (a drill setting, no s4 counterpart) sets
 * the interval, 0 = after every block.
 * Deviations: S_hi < S_lo is rejected (the reference wraps around, :451), and
 * --dump4 N applies per region in S order (the reference's per-thread order).
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "ps4.h"

typedef unsigned long long ull;

static FILE *out_fp;
static const char *opt_ckpt;

static double now_sec(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void die_fatal(const char *msg)
{
    fprintf(stderr, "FATAL: %s; this interval is NOT certified complete\n", msg);
    fflush(stderr);
    exit(2);
}

static int sink_out(void *user, const char *lines, size_t n)
{
    (void)user;
    if (fwrite(lines, 1, n, out_fp) != n || fflush(out_fp) != 0) die_fatal("result output failed");
    return 0;
}

/* 6-set lines and the --test-sixset drill: at once, fsync'd, to SIXSET_FOUND.txt,
 * the result stream and stderr (exhaustive/s4.c:129-137). */
static void alert_out(void *user, const char *line)
{
    (void)user;
    FILE *f = fopen("SIXSET_FOUND.txt", "a");
    if (f) { fputs(line, f); fflush(f); fsync(fileno(f)); fclose(f); }
    fputs(line, out_fp);
    fflush(out_fp);
    fsync(fileno(out_fp));
    fputs(line, stderr);
    fflush(stderr);
}

typedef struct {
    ull lo, hi, next;      /* region and next unsearched S */
    ps4_counters c;
    double secs;
} region_t;

static void add_counters(ps4_counters *a, const ps4_counters *b)
{
    a->S_done += b->S_done; a->pairs += b->pairs; a->S_ge3 += b->S_ge3; a->triples += b->triples;
    a->sets4 += b->sets4; a->sets5 += b->sets5; a->sets6 += b->sets6; a->multi += b->multi;
    a->ext_cand += b->ext_cand; a->ext_sqrt += b->ext_sqrt;
    for (int i = 0; i < 100; i++) a->hist5[i] += b->hist5[i];
    if (b->max_reps > a->max_reps) a->max_reps = b->max_reps;
    if (b->max_divisors > a->max_divisors) a->max_divisors = b->max_divisors;
    if (b->max_extensions > a->max_extensions) a->max_extensions = b->max_extensions;
}

/* s4ckpt v1 (exhaustive/s4.c:367-385): header, one line per region, DONE. */
static void write_ckpt(ull S_lo, ull S_hi, int T, ull W, const region_t *rg, int done)
{
    if (!opt_ckpt) return;
    fflush(out_fp);
    fsync(fileno(out_fp));
    char tmp[4200];
    snprintf(tmp, sizeof tmp, "%s.tmp", opt_ckpt);
    FILE *f = fopen(tmp, "w");
    if (!f) { perror("checkpoint open"); return; }
    fprintf(f, "s4ckpt v1\n%llu %llu %d %llu\n", S_lo, S_hi, T, W);
    for (int t = 0; t < T; t++) {
        const ps4_counters *c = &rg[t].c;
        fprintf(f, "%d %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu\n", t, rg[t].next, rg[t].hi,
                (ull)c->S_done, (ull)c->pairs, (ull)c->S_ge3, (ull)c->triples, (ull)c->sets4, (ull)c->sets5,
                (ull)c->sets6, (ull)c->multi, (ull)c->ext_cand, (ull)c->ext_sqrt);
    }
    if (done) fprintf(f, "DONE\n");
    fflush(f);
    fsync(fileno(f));
    fclose(f);
    if (rename(tmp, opt_ckpt) != 0) perror("checkpoint rename");
}

static int ckpt_threads(void)
{
    FILE *f = fopen(opt_ckpt, "r");
    if (!f) return -1;
    char hdr[64];
    ull a, b, d;
    int c = -1;
    if (!fgets(hdr, sizeof hdr, f) || fscanf(f, "%llu %llu %d %llu", &a, &b, &c, &d) != 4) c = -1;
    fclose(f);
    return c;
}

/* 0 = no checkpoint, 1 = positions restored, 2 = complete (exhaustive/s4.c:394-412). */
static int read_ckpt(ull S_lo, ull S_hi, int T, ull W, region_t *rg)
{
    FILE *f = fopen(opt_ckpt, "r");
    if (!f) return 0;
    char hdr[64];
    if (!fgets(hdr, sizeof hdr, f) || strncmp(hdr, "s4ckpt v1", 9)) {
        fprintf(stderr, "checkpoint %s: bad header\n", opt_ckpt);
        exit(1);
    }
    ull a, b, d;
    int c;
    if (fscanf(f, "%llu %llu %d %llu\n", &a, &b, &c, &d) != 4 || a != S_lo || b != S_hi || c != T || d != W) {
        fprintf(stderr, "checkpoint %s was written for S_lo=%llu S_hi=%llu threads=%d chunk=%llu; current parameters differ\n",
                opt_ckpt, a, b, c, d);
        exit(1);
    }
    for (int t = 0; t < T; t++) {
        int id;
        ull ns, rh, v[10];
        if (fscanf(f, "%d %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu %llu\n", &id, &ns, &rh, &v[0], &v[1], &v[2],
                   &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9]) != 13 || id != t || rh != rg[t].hi) {
            fprintf(stderr, "checkpoint %s: bad line for thread %d\n", opt_ckpt, t);
            exit(1);
        }
        rg[t].next = ns > rg[t].hi ? rg[t].hi : ns;
        ps4_counters *k = &rg[t].c;
        k->S_done = v[0]; k->pairs = v[1]; k->S_ge3 = v[2]; k->triples = v[3]; k->sets4 = v[4];
        k->sets5 = v[5]; k->sets6 = v[6]; k->multi = v[7]; k->ext_cand = v[8]; k->ext_sqrt = v[9];
    }
    char tail[16];
    const int done = fgets(tail, sizeof tail, f) != NULL && !strncmp(tail, "DONE", 4);
    fclose(f);
    return done ? 2 : 1;
}

/* Apply a tuning file (ps4_parse_tuning grammar); 0 or the ps4 error code. */
static int apply_tuning(ps4_ctx *ctx, const char *path)
{
    if (!path) return 0;
    FILE *f = fopen(path, "r");
    if (!f) { perror(path); return PS4_E_ARG; }
    char buf[4096];
    const size_t n = fread(buf, 1, sizeof buf - 1, f);
    fclose(f);
    buf[n] = 0;
    ps4_tuning t;
    int rc = ps4_parse_tuning(buf, &t);
    if (!rc) rc = ps4_set_tuning(ctx, &t);
    if (rc) fprintf(stderr, "ps4: tuning file %s: %s %s\n", path, ps4_strerror(rc), ps4_last_error(ctx));
    return rc;
}

static int run_elements(ull n_max, int device, const char *out, const char *tune)
{
    if (out) { out_fp = fopen(out, "a"); if (!out_fp) { perror(out); return 1; } }
    else out_fp = stdout;
    ps4_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.device = device;
    cfg.S_hi_max = n_max >= 4 ? 4 * n_max - 5 : 2;
    ps4_ctx *ctx = NULL;
    int rc = ps4_open(&cfg, &ctx);
    if (rc) { fprintf(stderr, "FATAL: %s: %s\n", ps4_strerror(rc), ctx ? ps4_last_error(ctx) : ""); return 2; }
    if (apply_tuning(ctx, tune)) { ps4_close(ctx); return 1; }
    uint64_t ns = 0, nt = 0;
    rc = ps4_search_elements(ctx, n_max, sink_out, NULL, &ns, &nt);
    if (rc) {
        fprintf(stderr, "FATAL: %s: %s\n", ps4_strerror(rc), ps4_last_error(ctx));
        ps4_close(ctx);
        return rc == PS4_E_ARG ? 1 : 2;
    }
    fflush(out_fp);
    fprintf(stderr, "ELEMENTS: n_max=%llu sets=%llu square_triples=%llu\n", n_max, (ull)ns, (ull)nt);
    ps4_close(ctx);
    if (out) fclose(out_fp);
    return 0;
}

int main(int argc, char **argv)
{
    ull S_lo = 1, S_hi = 10000000ull, W = 1ull << 20;
    int T = 0, pos = 0, device = 0, opt_noext = 0, opt_verify = 0, opt_resume = 0, opt_test6 = 0;
    long long opt_dump4 = 0;
    ull opt_elements = 0;
    const char *opt_out = NULL, *opt_tune = NULL;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--noext")) opt_noext = 1;
        else if (!strcmp(argv[i], "--verify")) opt_verify = 1;
        else if (!strcmp(argv[i], "--resume")) opt_resume = 1;
        else if (!strcmp(argv[i], "--test-sixset")) opt_test6 = 1;
        else if (!strcmp(argv[i], "--dump4") && i + 1 < argc) opt_dump4 = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--out") && i + 1 < argc) opt_out = argv[++i];
        else if (!strcmp(argv[i], "--ckpt") && i + 1 < argc) opt_ckpt = argv[++i];
        else if (!strcmp(argv[i], "--device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (!strcmp(argv[i], "--tune") && i + 1 < argc) opt_tune = argv[++i];
        else if (!strcmp(argv[i], "--elements") && i + 1 < argc) opt_elements = strtoull(argv[++i], 0, 10);
        else {
            pos++;
            if (pos == 1) S_lo = strtoull(argv[i], 0, 10);
            else if (pos == 2) S_hi = strtoull(argv[i], 0, 10);
            else if (pos == 3) T = atoi(argv[i]);
            else if (pos == 4) W = 1ull << atoi(argv[i]);
        }
    }
    if (opt_elements) return run_elements(opt_elements, device, opt_out, opt_tune);
    if (S_lo < 1) S_lo = 1;
    if (S_hi < S_lo) { fprintf(stderr, "S_hi < S_lo\n"); return 1; }
    if (T <= 0) {
        const long nc = sysconf(_SC_NPROCESSORS_ONLN);
        T = nc > 0 ? (int)nc : 1;
        if (opt_ckpt && opt_resume) { int c = ckpt_threads(); if (c > 0) T = c; }
    }
    if (opt_ckpt && !opt_resume && access(opt_ckpt, F_OK) == 0) {
        fprintf(stderr, "checkpoint %s exists: add --resume to continue it, or delete it to start over\n", opt_ckpt);
        return 1;
    }
    if (opt_out) { out_fp = fopen(opt_out, "a"); if (!out_fp) { perror(opt_out); return 1; } }
    else out_fp = stdout;

    region_t *rg = calloc((size_t)T, sizeof(region_t));
    const ull total = S_hi - S_lo;
    for (int t = 0; t < T; t++) {
        rg[t].lo = S_lo + (ull)((unsigned __int128)total * (unsigned)t / (unsigned)T);
        rg[t].hi = S_lo + (ull)((unsigned __int128)total * (unsigned)(t + 1) / (unsigned)T);
        rg[t].next = rg[t].lo;
    }
    const int rs = (opt_ckpt && opt_resume) ? read_ckpt(S_lo, S_hi, T, W, rg) : 0;
    if (rs == 2) {
        fprintf(stderr, "checkpoint %s marks S in [%llu, %llu) as complete; nothing to do\n", opt_ckpt, S_lo, S_hi);
        return 0;
    }
    fprintf(stderr, "s4: S in [%llu, %llu), %d threads, chunk %llu, ext=%s%s%s\n", S_lo, S_hi, T, W, opt_noext ? "off" : "on",
            opt_ckpt ? ", checkpointing" : "", rs == 1 ? ", RESUMED from checkpoint" : "");
    if (rs == 1)
        for (int t = 0; t < T; t++)
            fprintf(stderr, "  thread %2d resumes at S=%llu (region end %llu)\n", t, rg[t].next, rg[t].hi);
    fflush(stderr);

    ps4_config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.device = device;
    cfg.S_hi_max = S_hi > 2 ? S_hi : 2;
    cfg.flags = (opt_noext ? PS4_NOEXT : 0u) | (opt_verify ? PS4_VERIFY : 0u) | (opt_test6 ? PS4_TEST6 : 0u);
    cfg.dump4_limit = opt_dump4;
    /* capacity limits for failure drills (the reference rebuilds with -DS4_MAX_*) */
    if (getenv("PS4_CAP_REPS")) cfg.cap_reps = (unsigned)atoi(getenv("PS4_CAP_REPS"));
    if (getenv("PS4_CAP_DIVS")) cfg.cap_divs = (unsigned)atoi(getenv("PS4_CAP_DIVS"));
    if (getenv("PS4_CAP_EXT")) cfg.cap_ext = (unsigned)atoi(getenv("PS4_CAP_EXT"));
    ps4_ctx *ctx = NULL;
    int rc = ps4_open(&cfg, &ctx);
    if (rc) { fprintf(stderr, "ps4: cannot open device %d: %s\n", device, ps4_strerror(rc)); return 1; }
    if (apply_tuning(ctx, opt_tune)) { ps4_close(ctx); return 1; }
    ps4_set_alert(ctx, alert_out, NULL);

    /* blocks: whole chunks, about 2^33 S, so positions stay on chunk starts */
    ull block = W;
    while (block < (1ull << 33)) block <<= 1;
    ull total_chunks = 0, done_chunks = 0;
    for (int t = 0; t < T; t++) total_chunks += (rg[t].hi - rg[t].next + W - 1) / W;
    ps4_counters run;      /* this invocation only: histogram and capacity maxima */
    memset(&run, 0, sizeof run);
    const double t0 = now_sec();
    double last_ck = t0, last_rep = t0;
    const double ck_every = getenv("PS4_CKPT_SECS") ? atof(getenv("PS4_CKPT_SECS")) : 5.0;
    for (int t = 0; t < T; t++) {
        const double ts = now_sec();
        long long left = opt_dump4;
        while (rg[t].next < rg[t].hi) {
            ull lo = rg[t].next, hi = lo + block;
            if (hi > rg[t].hi || hi < lo) hi = rg[t].hi;
            ps4_set_dump4(ctx, left);
            ps4_counters c;
            rc = ps4_search(ctx, lo, hi, sink_out, NULL, &c);
            if (rc == PS4_E_CAPACITY || rc == PS4_E_VERIFY) die_fatal(ps4_last_error(ctx));
            if (rc) {
                fprintf(stderr, "ps4: %s (%s)\n", ps4_strerror(rc), ps4_last_error(ctx));
                die_fatal("device search failed");
            }
            if (left > 0) left = left > (long long)c.sets4 ? left - (long long)c.sets4 : 0;
            add_counters(&rg[t].c, &c);
            add_counters(&run, &c);
            rg[t].next = hi;
            done_chunks += (hi - lo + W - 1) / W;
            const double nw = now_sec();
            if (nw - last_ck >= ck_every) { write_ckpt(S_lo, S_hi, T, W, rg, 0); last_ck = nw; }
            if (nw - last_rep >= 30.0) {
                last_rep = nw;
                ull s4 = 0, s5 = 0, s6 = 0;
                for (int k = 0; k < T; k++) { s4 += rg[k].c.sets4; s5 += rg[k].c.sets5; s6 += rg[k].c.sets6; }
                fprintf(stderr, "[%.0fs] chunks %llu/%llu  4-sets %llu  5-sets %llu  6-sets %llu\n", nw - t0, done_chunks,
                        total_chunks, s4, s5, s6);
                fflush(stderr);
            }
        }
        rg[t].secs = now_sec() - ts;
    }
    const double el = now_sec() - t0;
    write_ckpt(S_lo, S_hi, T, W, rg, 1);
    ps4_counters tot;
    memset(&tot, 0, sizeof tot);
    for (int t = 0; t < T; t++) {
        add_counters(&tot, &rg[t].c);
        fprintf(stderr, "  thread %2d: S [%llu,%llu) 4-sets %llu 5-sets %llu  %.1fs\n", t, rg[t].lo, rg[t].hi,
                (ull)rg[t].c.sets4, (ull)rg[t].c.sets5, rg[t].secs);
    }
    fprintf(stderr,
            "TOTAL: S_done=%llu pairs=%llu S_ge3=%llu triples=%llu 4-sets=%llu 5-sets=%llu 6-sets=%llu multi-ext-4sets=%llu "
            "ext_cand=%llu ext_sqrt=%llu  wall=%.1fs%s\n",
            (ull)tot.S_done, (ull)tot.pairs, (ull)tot.S_ge3, (ull)tot.triples, (ull)tot.sets4, (ull)tot.sets5,
            (ull)tot.sets6, (ull)tot.multi, (ull)tot.ext_cand, (ull)tot.ext_sqrt, el,
            rs == 1 ? " (counters cumulative over resumed runs; wall time of this run only)" : "");
    fprintf(stderr, "5-sets by log2(largest element)%s:", rs == 1 ? " (this run only)" : "");
    for (int i = 0; i < 100; i++)
        if (run.hist5[i]) fprintf(stderr, " %d:%llu", i, (ull)run.hist5[i]);
    fprintf(stderr, "\n");
    fprintf(stderr, "CAPACITY: reps=%llu/%u divisors=%llu/%u extensions=%llu/%u%s\n", (ull)run.max_reps,
            cfg.cap_reps ? cfg.cap_reps : 512u, (ull)run.max_divisors, cfg.cap_divs ? cfg.cap_divs : 16384u,
            (ull)run.max_extensions, cfg.cap_ext ? cfg.cap_ext : 64u,
            rs == 1 ? " (maxima this resumed invocation only)" : "");
    ps4_close(ctx);
    if (out_fp != stdout) fclose(out_fp);
    free(rg);
    return 0;
}
