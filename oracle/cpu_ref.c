/*
 * cpu_ref.c -- CPU restatement of the reference census (TEST INFRASTRUCTURE ONLY).
 *
 * This file is the oracle that the HIP path is checked against.  It is loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg,
 * and by nothing that ships.  It is written for clarity, not speed: plain
 * loops, trial division, a fresh square root per column.
 *
 * Algorithm and counter semantics follow exhaustive/s4.c of
 * panjanek/pairwise-square-sums; each function names the lines it restates.
 * Parity pins (tests/test_oracle.py): the reference TOTAL vectors held in
 * exhaustive/logs/ (*.err), the known answers of exhaustive/README.md:157-159 and
 * exhaustive/test_verify.py:13,31-32, fixtures produced by the reference s4.c
 * (tests/golden/), brute force on tiny ranges, and closed forms for the
 * integer square roots and divisor counts.
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;
typedef uint64_t u64;
typedef int64_t i64;

#define ORA_CAP_REPS 512     /* exhaustive/s4.c:46-48 */
#define ORA_CAP_EXT 64       /* exhaustive/s4.c:49-51 */
#define ORA_CAP_DIVS 16384   /* exhaustive/s4.c:52-54 */

/* ------------------------------------------------------------------ */
/* exact square roots (exhaustive/s4.c:95-108)                          */
/* ------------------------------------------------------------------ */

/* floor(sqrt(n)) for n < 2^64.  A long-double estimate corrected until
 * r^2 <= n < (r+1)^2 holds exactly. */
uint64_t ora_isqrt64(uint64_t n)
{
    u64 r = (u64)sqrtl((long double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while ((u128)r * r > n) r--;
    while ((u128)(r + 1) * (r + 1) <= n) r++;
    return r;
}

static u64 isqrt_u128(u128 n)
{
    if ((n >> 64) == 0) return ora_isqrt64((u64)n);
    long double x = (long double)n;
    u64 r = (u64)sqrtl(x);
    while ((u128)r * r > n) r--;
    while ((u128)(r + 1) * (r + 1) <= n) r++;
    return r;
}

uint64_t ora_isqrt128(uint64_t lo, uint64_t hi)
{
    return isqrt_u128(((u128)hi << 64) | lo);
}

static int is_square_u128(u128 n)
{
    u64 r = isqrt_u128(n);
    return (u128)r * r == n;
}

int ora_is_square128(uint64_t lo, uint64_t hi)
{
    return is_square_u128(((u128)hi << 64) | lo);
}

/* ------------------------------------------------------------------ */
/* divisor counts and factorisation (exhaustive/s4.c:75-76, 140-156)    */
/* ------------------------------------------------------------------ */

uint32_t ora_num_divisors(uint64_t n)
{
    if (n == 0) return 0;
    uint32_t d = 1;
    for (u64 p = 2; p * p <= n; p += (p == 2 ? 1 : 2)) {
        uint32_t e = 0;
        while (n % p == 0) { n /= p; e++; }
        d *= e + 1;
    }
    if (n > 1) d *= 2;
    return d;
}

typedef struct { u64 p[40]; int e[40]; int n; } ora_fact;

/* Append the prime factors of m in ascending order, merging exponents of primes
 * already present (the order fact_add produces, exhaustive/s4.c:141-148). */
static void fact_append(ora_fact *f, u64 m)
{
    for (u64 p = 2; m > 1; p += (p == 2 ? 1 : 2)) {
        if (p * p > m) p = m;          /* m itself is prime */
        if (m % p) continue;
        int e = 0;
        while (m % p == 0) { m /= p; e++; }
        int i;
        for (i = 0; i < f->n; i++)
            if (f->p[i] == p) { f->e[i] += e; break; }
        if (i == f->n) { f->p[f->n] = p; f->e[f->n] = e; f->n++; }
    }
}

/* Divisors in mixed-radix order, first prime the fastest digit
 * (exhaustive/s4.c:149-156): index t <-> digits, complement at nd-1-t. */
static int fact_divisors(const ora_fact *f, u64 *out, int cap)
{
    long nd = 1;
    for (int i = 0; i < f->n; i++) nd *= (f->e[i] + 1);
    if (nd > cap) return -1;
    for (long t = 0; t < nd; t++) {
        long rest = t;
        u64 v = 1;
        for (int i = 0; i < f->n; i++) {
            int digit = (int)(rest % (f->e[i] + 1));
            rest /= (f->e[i] + 1);
            for (int k = 0; k < digit; k++) v *= f->p[i];
        }
        out[t] = v;
    }
    return (int)nd;
}

int ora_divisors_of_product(uint64_t m1, uint64_t m2, uint64_t *out, int cap)
{
    ora_fact f;
    f.n = 0;
    fact_append(&f, m1);
    fact_append(&f, m2);
    return fact_divisors(&f, out, cap);
}

/* ------------------------------------------------------------------ */
/* output formatting (exhaustive/s4.c:110-114, 209-220)                 */
/* ------------------------------------------------------------------ */

static void u128_dec(char *buf, u128 v)
{
    char tmp[48];
    int n = 0;
    do { tmp[n++] = (char)('0' + (int)(v % 10)); v /= 10; } while (v);
    for (int i = 0; i < n; i++) buf[i] = tmp[n - 1 - i];
    buf[n] = 0;
}

static int all_pairs_square(const u128 *x, int n)
{
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (x[i] == x[j] || !is_square_u128(x[i] + x[j])) return 0;
    return 1;
}

/* "tag [x1,...,xn] S=..\n" with the members sorted largest first. */
static void emit_set(ora_line_fn fn, void *user, const char *tag, const u128 *in, int n, u64 S)
{
    u128 x[8];
    memcpy(x, in, (size_t)n * sizeof(u128));
    for (int i = 1; i < n; i++)          /* insertion sort, descending */
        for (int j = i; j > 0 && x[j] > x[j - 1]; j--) { u128 t = x[j]; x[j] = x[j - 1]; x[j - 1] = t; }
    char line[1024], num[48];
    size_t len = (size_t)snprintf(line, sizeof line, "%s [", tag);
    for (int i = 0; i < n; i++) {
        u128_dec(num, x[i]);
        len += (size_t)snprintf(line + len, sizeof line - len, "%s%s", i ? "," : "", num);
    }
    snprintf(line + len, sizeof line - len, "] S=%llu%s\n", (unsigned long long)S,
             all_pairs_square(x, n) ? "" : "  *** VERIFY FAILED ***");
    fn(user, line);
}

/* ------------------------------------------------------------------ */
/* census state                                                          */
/* ------------------------------------------------------------------ */

typedef struct {
    unsigned flags;
    int64_t dump4_left;
    ora_line_fn fn;
    void *user;
    ora_counters *c;
    uint16_t *ndiv;     /* d(n) for n < ndiv_n, sieved (the table of exhaustive/s4.c:446-447) */
    u64 ndiv_n;
    u64 *divs;          /* scratch, ORA_CAP_DIVS entries */
    int failed;
} census_t;

static int square_residue(u64 x, u64 m)
{
    for (u64 r = 0; r < m; r++) if ((r * r) % m == x % m) return 1;
    return 0;
}

static int sqres64[64], sqres63[63], sqres65[65];

static void init_residues(void)
{
    static int done = 0;
    if (done) return;
    for (u64 x = 0; x < 64; x++) sqres64[x] = square_residue(x, 64);
    for (u64 x = 0; x < 63; x++) sqres63[x] = square_residue(x, 63);
    for (u64 x = 0; x < 65; x++) sqres65[x] = square_residue(x, 65);
    done = 1;
}

/* (u^2 + delta) mod m is a square residue, delta signed. */
static int qr_ok(u64 u, i64 delta, u64 m, const int *tab)
{
    i64 dm = delta % (i64)m;
    if (dm < 0) dm += (i64)m;
    u64 um = u % m;
    return tab[(um * um + (u64)dm) % m];
}

/* extend4, exhaustive/s4.c:160-207.  v[] are the members in the order the
 * triple step produced them, rt[i][j] = sqrt(v[i]+v[j]).  Returns the number of
 * extensions written to ext[] in divisor-index order, or -1 on overflow. */
static int extend_four(census_t *cs, const u64 v[4], const u64 rt[4][4], u128 *ext)
{
    /* the six pairs (i,j) with the third index k used for the root difference,
     * in the reference's scan order (exhaustive/s4.c:163) */
    static const int order[6][3] = {{0, 1, 2}, {2, 3, 0}, {0, 2, 1}, {1, 3, 0}, {0, 3, 1}, {1, 2, 0}};
    u64 best = ~0ull;
    int bi = 0, bj = 1, bk = 2;
    for (int t = 0; t < 6; t++) {
        int i = order[t][0], j = order[t][1], k = order[t][2];
        if (v[i] < v[j]) { int s = i; i = j; j = s; }
        u64 ra = rt[i][k], rb = rt[j][k];
        u64 est = (u64)cs->ndiv[ra - rb] * cs->ndiv[ra + rb];
        if (est < best) { best = est; bi = i; bj = j; bk = k; }
    }
    u64 A = v[bi];
    u64 m1 = rt[bi][bk] - rt[bj][bk], m2 = rt[bi][bk] + rt[bj][bk];
    int nd = ora_divisors_of_product(m1, m2, cs->divs, ORA_CAP_DIVS);
    if (nd < 0) return -1;
    if ((u64)nd > cs->c->max_divisors) cs->c->max_divisors = (u64)nd;
    int others[2], no = 0;
    for (int t = 0; t < 4; t++) if (t != bi && t != bj) others[no++] = t;
    u64 C1 = v[others[0]], C2 = v[others[1]];
    i64 d1 = (i64)C1 - (i64)A, d2 = (i64)C2 - (i64)A;
    int ne = 0;
    for (int t = 0; t < nd; t++) {
        u64 f = cs->divs[t], g = cs->divs[nd - 1 - t];
        if (f >= g || ((f ^ g) & 1)) continue;
        u64 u = (f + g) / 2;
        cs->c->ext_cand++;
        if (!qr_ok(u, d1, 64, sqres64) || !qr_ok(u, d2, 64, sqres64)) continue;
        if (!qr_ok(u, d1, 63, sqres63) || !qr_ok(u, d2, 63, sqres63)) continue;
        if (!qr_ok(u, d1, 65, sqres65) || !qr_ok(u, d2, 65, sqres65)) continue;
        u128 uu = (u128)u * u;
        if (uu <= A) continue;
        u128 e = uu - A;
        if (e == v[0] || e == v[1] || e == v[2] || e == v[3]) continue;
        cs->c->ext_sqrt++;
        if (!is_square_u128(e + C1) || !is_square_u128(e + C2)) continue;
        if (ne >= ORA_CAP_EXT) return -1;
        ext[ne++] = e;
    }
    if ((u64)ne > cs->c->max_extensions) cs->c->max_extensions = (u64)ne;
    return ne;
}

/* handle_4set, exhaustive/s4.c:225-278 (without the --verify abort and the
 * --test-sixset drill, which are driver behaviour). */
static void found_four(census_t *cs, u64 S, const u64 v[4], const u64 rt[4][4])
{
    char line[256];
    cs->c->sets4++;
    if (cs->dump4_left > 0) {
        cs->dump4_left--;
        snprintf(line, sizeof line, "4: [%lld,%lld,%lld,%lld] S=%llu\n", (long long)v[0], (long long)v[1],
                 (long long)v[2], (long long)v[3], (unsigned long long)S);
        cs->fn(cs->user, line);
    }
    if (cs->flags & ORA_NOEXT) return;
    u128 ext[ORA_CAP_EXT];
    int ne = extend_four(cs, v, rt, ext);
    if (ne < 0) { cs->failed = 1; return; }
    if (ne == 0) return;
    u64 vmax = v[0];
    for (int i = 1; i < 4; i++) if (v[i] > vmax) vmax = v[i];
    if (ne >= 2) {
        cs->c->multi++;
        char buf[4096], num[48];
        size_t len = (size_t)snprintf(buf, sizeof buf, "M: [%llu,%llu,%llu,%llu] ext=[", (unsigned long long)v[0],
                                      (unsigned long long)v[1], (unsigned long long)v[2], (unsigned long long)v[3]);
        for (int t = 0; t < ne; t++) {
            u128_dec(num, ext[t]);
            len += (size_t)snprintf(buf + len, sizeof buf - len, "%s%s", t ? "," : "", num);
        }
        snprintf(buf + len, sizeof buf - len, "] S=%llu\n", (unsigned long long)S);
        cs->fn(cs->user, buf);
    }
    for (int t = 0; t < ne; t++) {
        if (ext[t] <= vmax) continue;
        cs->c->sets5++;
        int lg = 0;
        for (u128 e = ext[t]; e > 1; e >>= 1) lg++;
        cs->c->hist5[lg > 99 ? 99 : lg]++;
        u128 x[5] = {v[0], v[1], v[2], v[3], ext[t]};
        emit_set(cs->fn, cs->user, "5:", x, 5, S);
    }
    for (int t = 0; t < ne; t++)
        for (int s = t + 1; s < ne; s++)
            if (is_square_u128(ext[t] + ext[s])) {
                cs->c->sets6++;
                u128 x[6] = {v[0], v[1], v[2], v[3], ext[t], ext[s]};
                emit_set(cs->fn, cs->user, "!!!!!!!! FOUND 6-SET:", x, 6, S);
            }
}

/* One S with its k representations S = p[t]^2 + q[t]^2, p ascending.
 * Triple and class rules of exhaustive/s4.c:316-341. */
static void visit_sum(census_t *cs, u64 S, const u64 *p, const u64 *q, unsigned k)
{
    cs->c->S_done++;
    if (k > cs->c->max_reps) cs->c->max_reps = k;
    if (k < 3 || (S & 3) == 2) return;
    cs->c->S_ge3++;
    if (k > ORA_CAP_REPS) { cs->failed = 1; return; }
    for (unsigned x = 0; x < k; x++)
        for (unsigned y = x + 1; y < k; y++)
            for (unsigned z = y + 1; z < k; z++) {
                cs->c->triples++;
                unsigned parity = (unsigned)((p[x] + p[y] + p[z]) & 1);
                for (int cls = 0; cls < 2; cls++) {
                    /* class 0: sums (Q_x,Q_y,Q_z), needs even p-sum;
                     * class 1: sums (Q_x,Q_y,P_z), needs p-sum = S mod 2 and p_z != q_z */
                    if (cls == 0 && parity != 0) continue;
                    if (cls == 1 && (parity != (unsigned)(S & 1) || p[z] == q[z])) continue;
                    i128 s1 = (i128)(q[x] * q[x]), s2 = (i128)(q[y] * q[y]);
                    i128 s3 = cls == 0 ? (i128)(q[z] * q[z]) : (i128)(p[z] * p[z]);
                    i128 SS = (i128)S;
                    i128 a = (s1 + s2 + s3 - SS) / 2, b = (s1 - s2 - s3 + SS) / 2;
                    i128 c = (-s1 + s2 - s3 + SS) / 2, d = (-s1 - s2 + s3 + SS) / 2;
                    if (a <= 0 || b <= 0 || c <= 0 || d <= 0) continue;
                    u64 v[4] = {(u64)a, (u64)b, (u64)c, (u64)d};
                    u64 r_ab = q[x], r_ac = q[y], r_cd = p[x], r_bd = p[y];
                    u64 r_ad = cls == 0 ? q[z] : p[z];
                    u64 r_bc = cls == 0 ? p[z] : q[z];
                    u64 rt[4][4] = {{0, r_ab, r_ac, r_ad}, {r_ab, 0, r_bc, r_bd}, {r_ac, r_bc, 0, r_cd}, {r_ad, r_bd, r_cd, 0}};
                    found_four(cs, S, v, rt);
                    if (cs->failed) return;
                }
            }
}

int ora_census(uint64_t S_lo, uint64_t S_hi, unsigned flags, int64_t dump4,
               ora_line_fn fn, void *user, ora_counters *out)
{
    memset(out, 0, sizeof *out);
    if (S_lo < 1) S_lo = 1;
    if (S_hi <= S_lo) return 0;
    init_residues();
    census_t cs;
    memset(&cs, 0, sizeof cs);
    cs.flags = flags; cs.dump4_left = dump4; cs.fn = fn; cs.user = user; cs.c = out;
    /* d(n) sieve over n < 2*isqrt(S_hi)+64: every root sum ra+rb lies below it */
    cs.ndiv_n = 2 * ora_isqrt64(S_hi) + 64;
    cs.ndiv = (uint16_t *)calloc(cs.ndiv_n, sizeof(uint16_t));
    cs.divs = (u64 *)malloc(ORA_CAP_DIVS * sizeof(u64));
    if (!cs.ndiv || !cs.divs) { free(cs.ndiv); free(cs.divs); return -2; }
    for (u64 d = 1; d < cs.ndiv_n; d++)
        for (u64 m = d; m < cs.ndiv_n; m += d) cs.ndiv[m]++;

    const u64 BLK = 1ull << 18;
    u64 pmax_all = ora_isqrt64((S_hi - 1) / 2);
    u64 *qstart = (u64 *)malloc((pmax_all + 2) * sizeof(u64));
    uint32_t *cnt = (uint32_t *)malloc(BLK * sizeof(uint32_t));
    uint32_t *pos = (uint32_t *)malloc((BLK + 1) * sizeof(uint32_t));
    u64 cap_pv = BLK;
    uint32_t *pv = (uint32_t *)malloc(cap_pv * sizeof(uint32_t));
    u64 pl[ORA_CAP_REPS], ql[ORA_CAP_REPS];
    int rc = 0;
    for (u64 B0 = S_lo; B0 < S_hi && !cs.failed; B0 = (B0 + BLK < S_hi ? B0 + BLK : S_hi)) {
        u64 B1 = B0 + BLK < S_hi ? B0 + BLK : S_hi;
        u64 n = B1 - B0;
        u64 pmax = ora_isqrt64((B1 - 1) / 2);
        memset(cnt, 0, n * sizeof(uint32_t));
        /* every lattice point 1 <= p <= q with p^2+q^2 in [B0,B1) */
        for (u64 p = 1; p <= pmax; p++) {
            u64 pp = p * p, q = p;
            if (B0 > pp + pp) { q = ora_isqrt64(B0 - pp); if (q * q + pp < B0) q++; }
            qstart[p] = q;
            for (; pp + q * q < B1; q++) cnt[pp + q * q - B0]++;
        }
        u64 total = 0;
        for (u64 i = 0; i < n; i++) { pos[i] = (uint32_t)total; total += cnt[i]; }
        out->pairs += total;
        if (total > cap_pv) { cap_pv = total; pv = (uint32_t *)realloc(pv, cap_pv * sizeof(uint32_t)); }
        for (u64 p = 1; p <= pmax; p++) {          /* ascending p: each bucket ends up sorted */
            u64 pp = p * p;
            for (u64 q = qstart[p]; pp + q * q < B1; q++) pv[pos[pp + q * q - B0]++] = (uint32_t)p;
        }
        for (u64 i = 0; i < n && !cs.failed; i++) {
            u64 S = B0 + i;
            unsigned k = cnt[i];
            u64 base = pos[i] - k;
            if (k >= 3 && (S & 3) != 2 && k <= ORA_CAP_REPS)
                for (unsigned t = 0; t < k; t++) { pl[t] = pv[base + t]; ql[t] = ora_isqrt64(S - pl[t] * pl[t]); }
            visit_sum(&cs, S, pl, ql, k);
        }
    }
    if (cs.failed) rc = -1;
    free(qstart); free(cnt); free(pos); free(pv); free(cs.ndiv); free(cs.divs);
    return rc;
}

/* ------------------------------------------------------------------ */
/* triple query                                                          */
/* ------------------------------------------------------------------ */

static int cmp_u64(const void *a, const void *b)
{
    u64 x = *(const u64 *)a, y = *(const u64 *)b;
    return x < y ? -1 : x > y;
}

int64_t ora_extend_triple(const uint64_t abc[3], uint64_t d_max, uint64_t *d_out, size_t cap)
{
    u64 x[3] = {abc[0], abc[1], abc[2]};
    qsort(x, 3, sizeof(u64), cmp_u64);
    u64 c = x[0], b = x[1], a = x[2];                      /* a > b > c */
    if (c == 0 || c == b || b == a) return -1;
    if (!is_square_u128((u128)a + b) || !is_square_u128((u128)a + c) || !is_square_u128((u128)b + c)) return -1;
    /* d + a = u^2, d + b = w^2  =>  (u - w)(u + w) = a - b */
    u64 D = a - b;
    ora_fact f;
    f.n = 0;
    fact_append(&f, D);
    long nd = 1;
    for (int i = 0; i < f.n; i++) nd *= f.e[i] + 1;
    u64 *divs = (u64 *)malloc((size_t)nd * sizeof(u64));
    if (!divs) return -1;
    fact_divisors(&f, divs, (int)nd);
    u64 *found = (u64 *)malloc(((size_t)nd + 1) * sizeof(u64));
    int64_t nf = 0;
    for (long t = 0; t < nd; t++) {
        u64 lo = divs[t], hi = divs[nd - 1 - t];
        if (lo >= hi || ((lo ^ hi) & 1)) continue;
        u128 u = (lo + hi) / 2;
        u128 uu = u * u;
        if (uu <= a) continue;
        u128 d = uu - a;
        if (d > d_max || d == a || d == b || d == c) continue;
        if (!is_square_u128(d + c)) continue;
        found[nf++] = (u64)d;
    }
    qsort(found, (size_t)nf, sizeof(u64), cmp_u64);
    for (int64_t i = 0; i < nf && (size_t)i < cap; i++) d_out[i] = found[i];
    free(divs);
    free(found);
    return nf;
}

/* ------------------------------------------------------------------ */
/* element-range mode (no s4.c entry point: the plain search by largest  */
/* element of the task's first configuration), by brute force           */
/* ------------------------------------------------------------------ */

/* Square triples {a < b < c}: a+b = x^2, a+c = y^2, b+c = z^2 with x < y < z,
 * x+y+z even, a = (x^2+y^2-z^2)/2 >= 1, c = (y^2+z^2-x^2)/2 <= n_max; every z
 * tried one by one. */
uint64_t ora_square_triples(uint64_t n_max)
{
    u64 n = 0;
    for (u64 x = 1; x * x <= 2 * n_max; x++)
        for (u64 y = x + 1; y * y <= 2 * n_max + x * x; y++)
            for (u64 z = y + 1; z * z < x * x + y * y; z++) {
                if ((x + y + z) & 1) continue;
                const u64 c2 = y * y + z * z - x * x;
                if (c2 > 2 * n_max) break;
                n++;
            }
    return n;
}

typedef struct { u64 v[4]; } ora_set4;

static int cmp_set4(const void *p, const void *q)
{
    const ora_set4 *a = (const ora_set4 *)p, *b = (const ora_set4 *)q;
    for (int i = 0; i < 4; i++)
        if (a->v[i] != b->v[i]) return a->v[i] < b->v[i] ? -1 : 1;
    return 0;
}

static int is_square_u64(u64 n)
{
    const u64 r = ora_isqrt64(n);
    return r * r == n;
}

/* Every 4-set {a < b < c < d} with six square pair sums and d <= n_max: each
 * square triple {a, b, c} as above, extended by every d in (c, n_max] with
 * d + c = w^2 and d + a, d + b square. */
int64_t ora_elements(uint64_t n_max, ora_line_fn fn, void *user)
{
    size_t cap = 1024, n = 0;
    ora_set4 *sets = (ora_set4 *)malloc(cap * sizeof(ora_set4));
    if (!sets) return -1;
    for (u64 x = 1; x * x <= 2 * n_max; x++)
        for (u64 y = x + 1; y * y <= 2 * n_max + x * x; y++)
            for (u64 z = y + 1; z * z < x * x + y * y; z++) {
                if ((x + y + z) & 1) continue;
                if (y * y + z * z - x * x > 2 * n_max) break;
                const u64 A = (x * x + y * y - z * z) / 2, B = x * x - A, C = y * y - A;   /* A < B < C */
                for (u64 w = ora_isqrt64(2 * C) + 1; w * w - C <= n_max; w++) {
                    const u64 d = w * w - C;
                    if (d <= C) continue;
                    if (!is_square_u64(d + A) || !is_square_u64(d + B)) continue;
                    if (n == cap) {
                        cap *= 2;
                        ora_set4 *t = (ora_set4 *)realloc(sets, cap * sizeof(ora_set4));
                        if (!t) { free(sets); return -1; }
                        sets = t;
                    }
                    sets[n].v[0] = A; sets[n].v[1] = B; sets[n].v[2] = C; sets[n].v[3] = d;
                    n++;
                }
            }
    qsort(sets, n, sizeof(ora_set4), cmp_set4);
    char line[128];
    for (size_t i = 0; i < n && fn; i++) {
        snprintf(line, sizeof line, "[%llu,%llu,%llu,%llu]\n", (unsigned long long)sets[i].v[0],
                 (unsigned long long)sets[i].v[1], (unsigned long long)sets[i].v[2], (unsigned long long)sets[i].v[3]);
        fn(user, line);
    }
    free(sets);
    return (int64_t)n;
}
