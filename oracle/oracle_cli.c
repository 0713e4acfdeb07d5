/*
 * oracle_cli.c -- command-line face of the CPU restatement (TEST INFRASTRUCTURE).
 *
 *   ps4_oracle S_lo S_hi [--noext] [--dump4 N]
 *   ps4_oracle --triple a b c d_max
 *
 * Result lines go to stdout; a TOTAL line in the reference's key=value grammar
 * (exhaustive/s4.c:498-500) and the capacity line (:504-506) go to stderr.
 * Single-threaded.  Used by tests and by bench.py's CPU-baseline leg only.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "cpu_ref.h"

static void put_line(void *user, const char *line)
{
    (void)user;
    fputs(line, stdout);
}

int main(int argc, char **argv)
{
    if (argc >= 6 && !strcmp(argv[1], "--triple")) {
        uint64_t abc[3] = {strtoull(argv[2], 0, 10), strtoull(argv[3], 0, 10), strtoull(argv[4], 0, 10)};
        uint64_t dmax = strtoull(argv[5], 0, 10);
        uint64_t out[4096];
        int64_t n = ora_extend_triple(abc, dmax, out, 4096);
        if (n < 0) { fprintf(stderr, "not a square triple\n"); return 1; }
        for (int64_t i = 0; i < n && i < 4096; i++) printf("%llu\n", (unsigned long long)out[i]);
        return 0;
    }
    uint64_t lo = 1, hi = 10000000ull;
    unsigned flags = 0;
    int64_t dump4 = 0;
    int pos = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "--noext")) flags |= ORA_NOEXT;
        else if (!strcmp(argv[i], "--dump4") && i + 1 < argc) dump4 = atoll(argv[++i]);
        else if (pos == 0) { lo = strtoull(argv[i], 0, 10); pos++; }
        else if (pos == 1) { hi = strtoull(argv[i], 0, 10); pos++; }
    }
    ora_counters c;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int rc = ora_census(lo, hi, flags, dump4, put_line, NULL, &c);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    fflush(stdout);
    if (rc) { fprintf(stderr, "FATAL: capacity exceeded; this interval is NOT certified complete\n"); return 2; }
    double el = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    fprintf(stderr,
            "TOTAL: S_done=%llu pairs=%llu S_ge3=%llu triples=%llu 4-sets=%llu 5-sets=%llu 6-sets=%llu "
            "multi-ext-4sets=%llu ext_cand=%llu ext_sqrt=%llu  wall=%.3fs\n",
            (unsigned long long)c.S_done, (unsigned long long)c.pairs, (unsigned long long)c.S_ge3,
            (unsigned long long)c.triples, (unsigned long long)c.sets4, (unsigned long long)c.sets5,
            (unsigned long long)c.sets6, (unsigned long long)c.multi, (unsigned long long)c.ext_cand,
            (unsigned long long)c.ext_sqrt, el);
    fprintf(stderr, "5-sets by log2(largest element):");
    for (int i = 0; i < 100; i++)
        if (c.hist5[i]) fprintf(stderr, " %d:%llu", i, (unsigned long long)c.hist5[i]);
    fprintf(stderr, "\nCAPACITY: reps=%llu/512 divisors=%llu/16384 extensions=%llu/64\n",
            (unsigned long long)c.max_reps, (unsigned long long)c.max_divisors, (unsigned long long)c.max_extensions);
    return 0;
}
