/*
 * cpu_ref.h -- CPU restatement of the reference census (TEST INFRASTRUCTURE).
 *
 * This is the oracle: only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it, and only as the checker.  The product
 * (libps4.so / the ps4 CLI) never links, loads or calls anything here.
 *
 * Every function restates exhaustive/s4.c of panjanek/pairwise-square-sums;
 * the line cited beside each declaration is the reference code it follows.
 */
#ifndef PS4_ORACLE_CPU_REF_H
#define PS4_ORACLE_CPU_REF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Counter vector with the meaning of stats_t (exhaustive/s4.c:80-91, 488-506). */
typedef struct {
    uint64_t S_done, pairs, S_ge3, triples, sets4, sets5, sets6, multi, ext_cand, ext_sqrt;
    uint64_t hist5[100];
    uint64_t max_reps, max_divisors, max_extensions;
} ora_counters;

/* Receives one finished output line (with trailing '\n'). */
typedef void (*ora_line_fn)(void *user, const char *line);

#define ORA_NOEXT 1u   /* --noext  (exhaustive/s4.c:242) */

/* Census of every S in [S_lo, S_hi): exhaustive/s4.c:280-342 (process_chunk),
 * :225-278 (handle_4set), :160-207 (extend4).  dump4 > 0 emits the first dump4
 * 4-sets as "4:" lines (:237-241) in S order.  Returns 0, or -1 on a capacity
 * overflow (reps > 512, divisors > 16384, extensions > 64: :176, :201, :321). */
int ora_census(uint64_t S_lo, uint64_t S_hi, unsigned flags, int64_t dump4,
               ora_line_fn fn, void *user, ora_counters *out);

/* Triple query: every d in [1, d_max] \ {a,b,c} with d+a, d+b, d+c all square,
 * written ascending.  No reference entry point exists; this is the extension
 * step of exhaustive/s4.c:160-207 with three known members instead of four.
 * Returns the number of values found (may exceed cap; only cap are written),
 * or -1 if the input is not a valid square triple. */
int64_t ora_extend_triple(const uint64_t abc[3], uint64_t d_max, uint64_t *d_out, size_t cap);

/* Exact primitives (exhaustive/s4.c:95-108). */
uint64_t ora_isqrt64(uint64_t n);
uint64_t ora_isqrt128(uint64_t lo, uint64_t hi);
int      ora_is_square128(uint64_t lo, uint64_t hi);

/* Divisor count d(n) by trial division (the quantity ndivtab holds, :446-447). */
uint32_t ora_num_divisors(uint64_t n);

/* Mixed-radix divisor list of (m1*m2) with the prime order of fact_add(m1) then
 * fact_add(m2) (exhaustive/s4.c:140-156).  Returns the count, or -1 if > cap. */
int ora_divisors_of_product(uint64_t m1, uint64_t m2, uint64_t *out, int cap);

/* Element-range mode (no s4.c entry point; the plain search by largest
 * element): the number of square triples {a<b<c} (a+b, a+c, b+c square) with
 * c <= n_max, and every 4-set {a<b<c<d} with six square pair sums and
 * d <= n_max as "[a,b,c,d]\n" lines in ascending order (returns their count,
 * -1 when out of memory).  Brute force over the roots of a+b, a+c, b+c. */
uint64_t ora_square_triples(uint64_t n_max);
int64_t ora_elements(uint64_t n_max, ora_line_fn fn, void *user);

#ifdef __cplusplus
}
#endif
#endif
