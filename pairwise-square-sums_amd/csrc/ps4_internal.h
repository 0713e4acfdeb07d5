/*
 * ps4_internal.h -- layouts shared by the HIP kernels and the host library.
 * Not part of the public ABI (include/ps4.h is).
 */
#ifndef PS4_INTERNAL_H
#define PS4_INTERNAL_H
#include <stdint.h>

/* One superwindow of the representation/triple kernel: S in [lo, hi) of
 * interval `rid`.  The host cuts every interval into such items. */
typedef struct {
    uint64_t lo, hi;
    uint32_t rid, pad;
} ps4_item;

/* A 4-set as the triple kernel emits it: its six pair-sum roots and the
 * interval it belongs to.  Members and S follow from the roots:
 *   S = r_ab^2 + r_cd^2,  a = (r_ab^2 + r_ac^2 - r_bc^2)/2,  b = r_ab^2 - a,
 *   c = r_ac^2 - a,  d = r_ad^2 - a,  and a > b > c > d > 0. */
typedef struct {
    uint32_t r_ab, r_ac, r_ad, r_bc, r_bd, r_cd;
    uint32_t rid, pad;
} ps4_quad;

/* rid of a null entry: the queue is filled in whole 64-entry tiles, each
 * with the 4-sets of one superwindow; the slots a superwindow leaves unused
 * carry this rid (and zero roots). */
#define PS4_RID_NONE 0xFFFFFFFFu
// largest tuning values: superwindow shift, batch fill of the 4-set queue (%)
#define PS4_SW_SHIFT_MAX 8u
#define PS4_BATCH_PCT_MAX 95u
// element-range mode: qualifying 4-sets one batch can return (the batch's
// 4-set budget is held below it)
#define PS4_ELEM_CAP (1u << 25)
// largest n_max of element-range mode (the square-triple count takes about
// sqrt(2 n_max) threads of up to sqrt(2 n_max) steps)
#define PS4_ELEM_NMAX 10000000000ull

/* An extension e of a 4-set (e + each member is square). */
typedef struct {
    uint32_t r[6];          /* the 4-set's roots, order of ps4_quad */
    uint32_t rid, pad;
    uint64_t e_lo, e_hi;    /* e as a 128-bit value */
} ps4_hit;

/* Per-interval counter slots on the device. */
enum {
    CT_S_DONE = 0, CT_PAIRS, CT_S_GE3, CT_TRIPLES, CT_SETS4, CT_EXT_CAND, CT_EXT_SQRT,
    CT_MAX_REPS, CT_MAX_DIVS, CT_N
};

/* Error bits raised by the kernels (ps4_ctl.err). */
enum {
    ERR_QUAD_OVF   = 1u << 0,  /* 4-set queue full: rerun with shorter batches      */
    ERR_BIN_OVF    = 1u << 1,  /* point bin full: rerun with a smaller window        */
    ERR_LDS_OVF    = 1u << 2,  /* per-window LDS lists full: smaller window          */
    ERR_HIT_OVF    = 1u << 3,  /* extension buffer full: shorter batches             */
    ERR_CAP_REPS   = 1u << 4,  /* k > cap_reps on a counted S (reference fatal)      */
    ERR_CAP_DIVS   = 1u << 5,  /* d(D) > cap_divs (reference fatal)                  */
    ERR_VERIFY     = 1u << 6,  /* a 4-set failed its root check                      */
    ERR_ARENA      = 1u << 7,  /* one 4-set's divisor lists exceed the LDS arena     */
    ERR_FACTORS    = 1u << 8,  /* more distinct primes than the kernel holds         */
    ERR_BADQUAD    = 1u << 9,  /* a queued 4-set fails its root identities           */
    ERR_BOUNDS     = 1u << 10  /* an index derived from data left its buffer (the
                                  access was skipped)                                */
};

/* Work-queue control block (device memory, reset before every launch pair). */
typedef struct {
    uint32_t item_next;     /* next superwindow to claim                           */
    uint32_t tile_next;     /* next 4-set tile to claim                            */
    uint32_t err;           /* ERR_* bits                                          */
    uint32_t hit_count;     /* extensions written                                  */
    uint64_t quad_count;    /* 4-sets written                                      */
    uint64_t pad;
} ps4_ctl;

/* QR shift tables, one 64-bit word per offset a: bit r of T_m[a] is set iff
 * (r*r + a) mod m is a square mod m.  T_65 is indexed by b = (u + 32) mod 65
 * taken mod 64: bit b is the test of u = b - 32, and b = 0 also stands for
 * u = 32, whose square is that of u = -32. */
typedef struct {
    uint64_t t64[64];
    uint64_t t63[63];
    uint64_t t65[65];
    uint32_t t11[11], t17[17], t19[19], t23[23];   // t23: folded, bit min(r, 23 - r)
} ps4_qrtab;

#endif
