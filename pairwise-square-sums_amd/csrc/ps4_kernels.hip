// ps4_kernels.hip -- gfx950 kernels of the census (exhaustive/s4.c:160-342 of
// panjanek/pairwise-square-sums, redesigned for CDNA4).
//
// k_rq   one workgroup per superwindow of S values (persistent, work queue):
//        walks every lattice column p once, bins the points (p,q), p<=q,
//        p^2+q^2 in the superwindow into 4096-wide windows in a per-workgroup
//        global scratch, then per window builds the representation buckets in
//        LDS (count, scan, fill -- process_chunk passes 1-3,
//        exhaustive/s4.c:285-310) and enumerates the representation triples of
//        every S with >= 3 representations, S != 2 mod 4, into 4-sets (pass 4,
//        :311-341).  Triples are spread evenly over the threads; each wave
//        appends its 4-sets to its own 64-entry tiles of a global queue.
// k_ext  single-wave workgroups over the queued 4-sets (persistent, claims of
//        4 tiles of 64): each lane holds one 4-set at a time and sets it up --
//        the pair choice (:163-171), the factorisation of D = (ra-rb)(ra+rb) by
//        one gather of each odd part's entry in a resident odd-factor table, the
//        closed-form candidate count, and the split of its divisors into two
//        lists; a pass then writes the lists of as many held 4-sets as fit an
//        LDS arena and flattens all their divisor pairs over the lanes.  Each
//        candidate is filtered by per-4-set quadratic-residue masks (mod 64, 63,
//        65: the ext_sqrt rule of :192-198), by the exact e>0 / e not in the set
//        conditions, then by residue masks mod 11, 17, 19, 23 and a size bound;
//        the few survivors get the exact 128-bit square tests (:199-200).
//
// All integer work is exact: uint64, doubles only where every value is an
// integer below 2^53, and unsigned __int128 for e = u^2 - A.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>

#include "ps4_internal.h"

typedef uint64_t u64;
typedef uint32_t u32;
typedef int64_t i64;
typedef int32_t i32;
typedef unsigned __int128 u128;

// ----------------------------------------------------------------------------
// small helpers
// ----------------------------------------------------------------------------

// floor(sqrt(n)), n < 2^64: FP64 estimate, exact integer correction
// (the closed form r^2 <= n < (r+1)^2; exhaustive/s4.c:95-100).
__device__ __forceinline__ u64 d_isqrt64(u64 n)
{
    u64 r = (u64)__builtin_sqrt((double)n);
    if (r > 0xFFFFFFFFull) r = 0xFFFFFFFFull;
    while (r * r > n) --r;
    while (r < 0xFFFFFFFFull && (r + 1) * (r + 1) <= n) ++r;
    return r;
}

// floor(sqrt(n)) for 1 <= n < 2^52 from the hardware reciprocal square root
// refined by one Newton step (a few operations, where the correctly rounded
// square root is a long sequence); every candidate root t < 2^26 has t^2 and
// (t+1)^2 exact in doubles, so the corrections are exact whatever the estimate.
__device__ __forceinline__ u32 d_isqrt52r(u64 n)
{
    const double x = (double)n;
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y;
    s = __builtin_fma(s, __builtin_fma(-s, 0.5 * y, 0.5), s);
    double t = __builtin_trunc(s);
    while (t * t > x) t -= 1.0;
    while ((t + 1.0) * (t + 1.0) <= x) t += 1.0;
    return (u32)t;
}

// n is a perfect square, n < 2^106 (exhaustive/s4.c:101-108).
__device__ __noinline__ bool d_is_square128_slow(u128 n)
{
    const double x = (double)(u64)(n >> 64) * 18446744073709551616.0 + (double)(u64)n;
    u64 r = (u64)__builtin_sqrt(x);
    while ((u128)r * r > n) --r;
    while ((u128)(r + 1) * (r + 1) <= n) ++r;
    return (u128)r * r == n;
}

// The same test without loops for n < 2^96: x = n rounded once (the high word
// is exact, the fused multiply-add rounds once), so sqrt(x) is within
// k * 2^-53 + a few ulp(k) < 0.25 of k = sqrt(n) when n = k^2 (k < 2^48), and
// the nearest integer r is k; r^2 == n then decides exactly.
__device__ __forceinline__ bool d_is_square128(u128 n)
{
    const u64 hi = (u64)(n >> 64), lo = (u64)n;
    if (hi >> 32) return d_is_square128_slow(n);
    const double x = __builtin_fma((double)hi, 18446744073709551616.0, (double)lo);
    const u64 r = (u64)__builtin_rint(__builtin_sqrt(x));
    return (u128)r * r == n;
}

// Takes values in registers at this point: an empty asm statement with them as
// inputs, so the compiler issues their loads before it and cannot sink them
// into branches of their own.  (The host emulation of these sources defines it
// as nothing.)
#ifndef PS4_KEEP
#define PS4_KEEP(...) asm volatile("" ::__VA_ARGS__)
#endif
// Makes a register value opaque (an empty asm statement that may have changed
// it), so that every use starts from the same compare.
#ifndef PS4_OPAQUE
#define PS4_OPAQUE(x) asm volatile("" : "+v"(x))
#endif
// (a << 3) + b for a uniform b, in one instruction (the compiler splits it)
#ifndef PS4_LSHL3_ADD
#define PS4_LSHL3_ADD(a, b)                                                               \
    ({                                                                                    \
        u32 r_;                                                                           \
        asm("v_lshl_add_u32 %0, %1, 3, %2" : "=v"(r_) : "v"((u32)(a)), "s"((u32)(b)));    \
        r_;                                                                               \
    })
#endif
// popcount(x) + a in one instruction (the compiler adds separately)
#ifndef PS4_BCNT_ADD
#define PS4_BCNT_ADD(x, a)                                                                  \
    ({                                                                                    \
        u32 r_;                                                                           \
        asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r_) : "v"((u32)(x)), "v"((u32)(a)));       \
        r_;                                                                               \
    })
#endif

// Two 16-bit halves of a word at once: the halfwise maximum of a and b, and
// |a.lo - b.lo| + |a.hi - b.hi| + c.
#ifndef PS4_MAX_U16X2
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
#define PS4_MAX_U16X2(a, b)                                                                                     \
    __builtin_bit_cast(u32, __builtin_elementwise_max(__builtin_bit_cast(u16x2, (u32)(a)), __builtin_bit_cast(u16x2, (u32)(b))))
#define PS4_SAD_U16(a, b, c) __builtin_amdgcn_sad_u16((u32)(a), (u32)(b), (u32)(c))
#endif

__device__ __forceinline__ u32 lane_id() { return threadIdx.x & 63u; }

// A value every thread of the workgroup holds (read from LDS after a barrier),
// made provably uniform so that branches and loop exits around barriers stay
// scalar and the compiler cannot split them across lanes.
__device__ __forceinline__ u32 uni(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }

// set bits of a wave-wide mask (a ballot) in the lanes below this one
__device__ __forceinline__ u32 lanes_below(u64 m)
{
    return __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
}

// Inclusive prefix sum over the 64 lanes of a wave (all lanes active): four
// row_shr steps inside each 16-lane row, then the row totals by row_bcast:15
// (into rows 1 and 3) and row_bcast:31 (into rows 2 and 3).  Each step is one
// VALU add with a DPP source; a lane without a source adds 0.
__device__ __forceinline__ u32 wave_incl_scan(u32 x)
{
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);   // row_shr:1
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);   // row_shr:2
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);   // row_shr:4
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);   // row_shr:8
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);   // row_bcast:15
    x += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);   // row_bcast:31
    return x;
}

// Exclusive block scan; *total gets the block sum.  Contains two barriers.
template <int NT>
__device__ __forceinline__ u32 block_excl_scan(u32 v, u32 *s_w, u32 *total)
{
    const u32 incl = wave_incl_scan(v);
    const u32 wid = threadIdx.x >> 6;
    if (lane_id() == 63) s_w[wid] = incl;
    __syncthreads();
    u32 pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        const u32 t = s_w[w];
        pre += ((u32)w < wid) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return pre + incl - v;
}

// Two exclusive block scans sharing their two barriers; *t1, *t2 get the sums.
template <int NT>
__device__ __forceinline__ void block_excl_scan2(u32 v1, u32 v2, u32 *s_w, u32 *e1, u32 *e2, u32 *t1, u32 *t2)
{
    const u32 i1 = wave_incl_scan(v1), i2 = wave_incl_scan(v2);
    const u32 wid = threadIdx.x >> 6;
    if (lane_id() == 63) { s_w[wid] = i1; s_w[NT / 64 + wid] = i2; }
    __syncthreads();
    u32 p1 = 0, p2 = 0, s1 = 0, s2 = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) {
        const u32 a = s_w[w], b = s_w[NT / 64 + w];
        p1 += ((u32)w < wid) ? a : 0u;
        p2 += ((u32)w < wid) ? b : 0u;
        s1 += a;
        s2 += b;
    }
    __syncthreads();
    *e1 = p1 + i1 - v1;
    *e2 = p2 + i2 - v2;
    *t1 = s1;
    *t2 = s2;
}

// Block-wide OR of a predicate, uniform result (barriers only, no library call).
template <int NT>
__device__ __forceinline__ bool block_any(bool p, u32 *s_w)
{
    u32 tot;
    (void)block_excl_scan<NT>(p ? 1u : 0u, s_w, &tot);
    return uni(tot) != 0;
}

template <int NT>
__device__ __forceinline__ bool block_all(bool p, u32 *s_w)
{
    return !block_any<NT>(!p, s_w);
}

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int m)
{
    const u32 lo = __shfl_xor((u32)v, m, 64), hi = __shfl_xor((u32)(v >> 32), m, 64);
    return ((u64)hi << 32) | lo;
}

template <int NT>
__device__ __forceinline__ u64 block_sum_u64(u64 v, u64 *s_w)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
    const u32 wid = threadIdx.x >> 6;
    if (lane_id() == 0) s_w[wid] = v;
    __syncthreads();
    u64 t = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) t += s_w[w];
    __syncthreads();
    return t;
}

template <int NT>
__device__ __forceinline__ u64 block_max_u64(u64 v, u64 *s_w)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) { const u64 o = shfl_xor_u64(v, m); v = o > v ? o : v; }
    const u32 wid = threadIdx.x >> 6;
    if (lane_id() == 0) s_w[wid] = v;
    __syncthreads();
    u64 t = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) t = s_w[w] > t ? s_w[w] : t;
    __syncthreads();
    return t;
}

// C(k,3) for 2 <= k <= 2000: C(k,2) * (k - 2) stays below 2^32 and is a
// multiple of 3 (x / 3 = mulhi(x, 0xAAAAAAAB) >> 1 for every 32-bit x)
__device__ __forceinline__ u32 tri3(u32 k) { return __umulhi((k * (k - 1) >> 1) * (k - 2), 0xAAAAAAABu) >> 1; }

// ----------------------------------------------------------------------------
// k_rq: representations and 4-sets
// ----------------------------------------------------------------------------

// threads per workgroup (NT) and window points a thread loads at once in the
// bucket passes (UN) are template parameters of k_rq (the launch shape the
// autotuner chooses among); 256 and 4 by default
#define RQ_NT_DEFAULT 256
#define RQ_UN_DEFAULT 4
#define RQ_LOGW 12          // windows of 4096 S: 29 KB of LDS, 5 workgroups (20 waves) per CU
#define RQ_W (1u << RQ_LOGW)
#define RQ_PV_CAP 1536      // representations of the counted S of a window (about 800 at
                            // S ~ 10^15, at most about 900 in samples; one S holds <= 512)
#define RQ_QS_CAP 512       // counted S of a window (about 140, at most about 160 in samples)
#define RQ_NB_MAX 1024
#define RQ_K_MAX 2000u      // largest bucket whose triples a window indexes (C(2000,3) < 2^31)
#define RQ_RUNS 2           // column runs a thread walks in lockstep (4 measured the same, and
                            // spills one register at the 96-register bound)

struct RqArgs {
    const ps4_item *items;
    u32 n_items;
    u32 bin_cap;        // points per window held in scratch
    u32 nb_alloc;       // windows per superwindow the scratch holds (<= RQ_NB_MAX)
    ps4_ctl *ctl;
    u64 *ctr;           // [rid][CT_N]
    u64 *scratch;       // per workgroup: RQ_NB_MAX * bin_cap points
    ps4_quad *quads;
    u64 quad_cap;
    u32 cap_reps;
    u32 verify;
    u32 n_rids;         // counter rows in ctr
};

// A wave's place in the 4-set queue.  The queue is filled a tile (64 entries,
// the extension kernel's unit) at a time: a wave appends its 4-sets at
// consecutive slots of its current tile and takes its next tile from a run of
// tiles it claimed with one atomic, 1, 2, 4, then 8 tiles per claim (the
// queue's counter is one address, so claims stay rare).  Every tile holds the
// 4-sets of one interval: when a wave moves to an item of another interval,
// and at its end, the unused slots of its tiles are written as null entries.
#define RQ_RUN_MAX 8u
struct RqTiles {
    u64 cur;     // first slot of the tile being filled
    u64 run;     // first slot of the next unused tile of the claimed run
    u32 fill;    // entries of the current tile written (64: none left, or no tile yet)
    u32 left;    // tiles left in the run
    u32 next;    // tiles of the next claim
    u32 rid;     // interval of the tiles (PS4_RID_NONE: none)
};

// a uniform 64-bit value: lane 0's
__device__ __forceinline__ u64 uni64(u64 v) { return ((u64)uni((u32)(v >> 32)) << 32) | uni((u32)v); }

__device__ __forceinline__ void rq_tiles_reset(RqTiles &tl, u32 rid)
{
    tl.cur = tl.run = 0;
    tl.fill = 64u;
    tl.left = 0;
    tl.next = 1;
    tl.rid = rid;
}

// Append one 4-set {w0, w1} for every lane with `on` (every lane active).
__device__ __forceinline__ void rq_put(RqTiles &tl, bool on, const uint4 &w0, const uint4 &w1, ps4_quad *out, u64 ocap,
                                      ps4_ctl *ctl, u32 *err)
{
    const u64 m = __ballot(on);
    if (m == 0ull) return;
    const u32 cnt = (u32)__popcll(m);
    const u32 idx = tl.fill + lanes_below(m);
    const bool cross = tl.fill + cnt >= 64u;   // uniform: the current tile fills up
    u64 nx = 0;
    if (cross) {
        if (tl.left == 0) {   // claim the next run of tiles
            u64 v = 0;
            if (lane_id() == 0)
                v = atomicAdd((unsigned long long *)&ctl->quad_count, (unsigned long long)(64u * tl.next));
            tl.run = uni64(v);
            tl.left = tl.next;
            tl.next = tl.next < RQ_RUN_MAX ? 2u * tl.next : RQ_RUN_MAX;
        }
        nx = tl.run;
        tl.run += 64u;
        --tl.left;
    }
    if (on) {
        const u64 pos = (idx < 64u ? tl.cur : nx) + (idx & 63u);
        if (pos < ocap) {
            uint4 *o = (uint4 *)(out + pos);
            o[0] = w0;
            o[1] = w1;
        } else {
            *err |= ERR_QUAD_OVF;
        }
    }
    tl.fill += cnt;
    if (cross) {
        tl.cur = nx;
        tl.fill -= 64u;
    }
}

// Null entries in the unused slots of a wave's tiles: the rest of the current
// tile and the unused tiles of its run (every lane active).
__device__ __forceinline__ void rq_pad(const RqTiles &tl, ps4_quad *out, u64 ocap, u32 *err)
{
    const u32 l = lane_id();
    const uint4 z0 = make_uint4(0u, 0u, 0u, 0u), z1 = make_uint4(0u, 0u, PS4_RID_NONE, 0u);
    for (u32 t = 0; t <= tl.left; ++t) {
        const u64 pos = (t == 0 ? tl.cur : tl.run + 64u * (t - 1u)) + l;
        if (t > 0 || l >= tl.fill) {
            if (pos < ocap) {
                uint4 *o = (uint4 *)(out + pos);
                o[0] = z0;
                o[1] = z1;
            } else {
                *err |= ERR_QUAD_OVF;
            }
        }
    }
}

// Walk the representation triples with flat indices [c0, c1) of one window
// (lexicographic index triples x<y<z inside each S, S in bucket order) and
// append the 4-sets of both classes to the wave's tiles (exhaustive/s4.c:324-339);
// returns their number.  Every lane runs nit >= c1 - c0 iterations (the tile
// appends are wave-wide), those past its range without effect.  The buckets
// are not sorted: the three p values of each triple are put in ascending order
// here, which is the order the classes are defined in (and makes every 4-set
// come out a > b > c > d).  s_qv holds the q of each representation
// (p^2 + q^2 = S), so q is read rather than recomputed; p ascending is q
// descending.  The window must have at least one triple.
__device__ u32 rq_walk(u32 c0, u32 c1, u32 nit, const u32 *s_sk, const u32 *s_st, const u32 *s_tp, const u32 *s_pv,
                       const u32 *s_qv, u32 nq, u64 B0, RqTiles &tl, ps4_quad *out, u64 ocap, ps4_ctl *ctl, u32 rid,
                       u32 verify, u32 *err)
{
    u32 j = 0, x = 0, y = 1, z = 2;   // a lane without triples stays on the first one
    if (c0 < c1) {
        u32 lo = 0, hi = nq;
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (s_tp[mid] <= c0) lo = mid; else hi = mid;
        }
        j = lo;
        u32 tau = c0 - s_tp[j];
        const u32 kk = s_sk[j] >> 16;
        for (;;) {
            const u32 m = kk - 1 - x, c = m * (m - 1) / 2;
            if (tau < c) break;
            tau -= c;
            ++x;
        }
        y = x + 1;
        while (tau >= kk - 1 - y) { tau -= kk - 1 - y; ++y; }
        z = y + 1 + tau;
    }
    u32 sk = s_sk[j], k = sk >> 16, st = s_st[j];
    const double B0d = (double)B0;
    const u32 b0 = (u32)B0;
    double Sd = B0d + (double)(sk & 0xFFFFu);
    u32 Sodd = (b0 + sk) & 1u;
    u32 n = 0;
    u32 vx = s_pv[st + x], vy = s_pv[st + y], vz = s_pv[st + z];
    u32 wx = s_qv[st + x], wy = s_qv[st + y], wz = s_qv[st + z];
    for (u32 i = 0; i < nit; ++i) {
        const u32 c = c0 + i;
        const bool live = c < c1;
        const u32 lo2 = min(vx, vy), hi2 = max(vx, vy);
        const u32 px = min(lo2, vz), pz = max(hi2, vz), py = max(lo2, min(hi2, vz));
        const double Px = (double)px * px, Py = (double)py * py, Pz = (double)pz * pz;
        const double sxy = Px + Py;
        const u32 par = (px + py + pz) & 1u;
        // class A: sums (Q_x,Q_y,Q_z), p-sum even, d = (P_x+P_y-P_z)/2 > 0
        const bool ca = live && (par == 0) && (sxy > Pz);
        // class B: sums (Q_x,Q_y,P_z), p-sum = S mod 2, 2P_z != S, d = (P_x+P_y+P_z-S)/2 > 0
        const bool cb = live && (par == Sodd) && (Pz + Pz != Sd) && (sxy + Pz > Sd);
        n += (u32)ca + (u32)cb;
        const u32 wlo = min(wx, wy), whi = max(wx, wy);
        const u32 qx = max(whi, wz), qz = min(wlo, wz), qy = max(wlo, min(whi, wz));
        const uint4 tail = make_uint4(py, px, rid, 0u);
        rq_put(tl, ca, make_uint4(qx, qy, qz, pz), tail, out, ocap, ctl, err);
        rq_put(tl, cb, make_uint4(qx, qy, pz, qz), tail, out, ocap, ctl, err);
        if (verify && (ca || cb)) {
            // the six roots of both 4-sets (exhaustive/s4.c:229-233)
            for (int cl = 0; cl < 2; ++cl) {
                if (cl == 0 ? !ca : !cb) continue;
                const u64 rab = qx, rac = qy, rad = cl == 0 ? qz : pz, rbc = cl == 0 ? pz : qz, rbd = py, rcd = px;
                const u64 va = (rab * rab + rac * rac - rbc * rbc) / 2;
                const u64 vb = rab * rab - va, vc = rac * rac - va, vd = rad * rad - va;
                const bool ok = va > vb && vb > vc && vc > vd && vd > 0 && vb + vc == rbc * rbc &&
                                vb + vd == rbd * rbd && vc + vd == rcd * rcd && va + vd == rad * rad &&
                                va + vb + vc + vd == B0 + (sk & 0xFFFFu);
                if (!ok) *err |= ERR_VERIFY;
            }
        }
        // the next triple in lexicographic order: z advances, a z at k moves y
        // on, a y at k - 1 moves x on, and an x at k - 2 moves to the next
        // counted S (an exec-masked block, rare per lane).  The three p (and q)
        // of the next triple are read together, so one LDS round trip per
        // triple; every index stays inside the current or next S's bucket, and
        // a lane past its range stays where it is.
        const bool tz = live && z + 1u == k, ty = tz && y + 2u == k, tn = ty && x + 3u == k;
        if (tn && c + 1 < c1) {
            ++j;
            sk = s_sk[j];
            k = sk >> 16;
            st = s_st[j];
            Sd = B0d + (double)(sk & 0xFFFFu);
            Sodd = (b0 + sk) & 1u;
        }
        x = tn ? 0u : x + (ty ? 1u : 0u);
        y = ty ? x + 1u : y + (tz ? 1u : 0u);
        z = tz ? y + 1u : z + (live ? 1u : 0u);
        vx = s_pv[st + x];
        vy = s_pv[st + y];
        vz = s_pv[st + z];
        wx = s_qv[st + x];
        wy = s_qv[st + y];
        wz = s_qv[st + z];
    }
    return n;
}

// The 256-thread shape is held to 5 waves per SIMD (<= 96 VGPRs, no spills), so
// that 5 workgroups (20 waves) share a CU: its LDS allows 5, and with 98 VGPRs
// only 4 fit (k_rq -6.5 % at S <= 10^10, -9.5 % at 9*10^11).  The other shapes
// are held by their LDS (128 threads) or workgroup size (512) anyway.
template <int NT, int UN>
__global__ __launch_bounds__(NT, NT == 256 ? 5 : 1) void k_rq(RqArgs a)
{
    __shared__ u32 s_cnt[RQ_W / 2];        // u16 bucket counts, two per word; later offsets
    __shared__ u32 s_pv[RQ_PV_CAP];        // p of every representation of a counted S
    __shared__ u32 s_qv[RQ_PV_CAP];        // and its q
    __shared__ u32 s_sk[RQ_QS_CAP];        // counted S: window offset | k << 16
    __shared__ u32 s_st[RQ_QS_CAP];        // first slot in s_pv
    __shared__ u32 s_tp[RQ_QS_CAP + 1];    // triple prefix
    __shared__ u32 s_bin[RQ_NB_MAX];       // points per window of the superwindow
    __shared__ u32 s_w[2 * NT / 64];
    __shared__ u64 s_w64[NT / 64];
    __shared__ u32 s_item;

    const u32 tid = threadIdx.x;
    u64 *scr = a.scratch + (u64)blockIdx.x * a.nb_alloc * a.bin_cap;
    u32 err = 0;
    RqTiles tl;   // this wave's tiles of the 4-set queue
    rq_tiles_reset(tl, PS4_RID_NONE);

    for (;;) {
        // The barrier first: it separates this claim from the previous item's
        // counter flush, so the two tid == 0 regions cannot be merged across the
        // loop back-edge (which would leave the other lanes spinning on a stale item).
        __syncthreads();
        if (tid == 0) s_item = atomicAdd(&a.ctl->item_next, 1u);
        __syncthreads();
        const u32 it = uni(s_item);
        if (it >= a.n_items) break;
        ps4_item item = a.items[it];
        item.rid = uni(item.rid);
        if (item.rid != tl.rid) {   // tiles hold one interval's 4-sets
            rq_pad(tl, a.quads, a.quad_cap, &err);
            rq_tiles_reset(tl, item.rid);
        }
        const u64 X0 = item.lo, X1 = item.hi, len = X1 - X0;
        const u32 nb = (u32)((len + RQ_W - 1) >> RQ_LOGW);
        // a malformed item (never produced by the host) is skipped, not walked
        if (X1 <= X0 || nb > a.nb_alloc || nb > RQ_NB_MAX || item.rid >= a.n_rids) {
            if (tid == 0) atomicOr(&a.ctl->err, (u32)ERR_BOUNDS);
            continue;
        }
        for (u32 i = tid; i < nb; i += NT) s_bin[i] = 0;
        __syncthreads();

        // ---- every lattice point 1 <= p <= q, X0 <= p^2+q^2 < X1, binned by window.
        // The columns p = 1..pmax are cut into NT * RQ_RUNS runs of equal length;
        // a thread follows RQ_RUNS runs in lockstep (runs tid, tid + NT, ...), each
        // along the inner circle: q the least admissible root of its column and
        // off = p^2 + q^2 - X0.  All of it fits in 32 bits: q > p gives
        // off < 2q - 1 < 2^27, and q = p (the diagonal) gives off = 2p^2 - X0 <
        // len.  Moving to the next column takes q down by at most one: if q >= p + 2,
        // p^2 + (q-1)^2 < X0 gives (p+1)^2 + (q-2)^2 = p^2 + (q-1)^2 + 2(p - q + 2) < X0;
        // if q = p + 1 the new column's diagonal bounds q; if q = p, q moves onto the
        // new diagonal.  So a step is straight-line selects; the points of a step's
        // columns (rare at high S, where a column meets a narrow window in a point
        // with probability about len / 2q) are emitted only when some lane has one,
        // their bin slots claimed together.  About 20 VALU operations per column:
        // when a short job at high S leaves one wave per SIMD, the walk is bound
        // by that wave's issue rate (a 4096-wide superwindow at 9e11 walks
        // 6.7e5 columns in about 0.3 ms; the 64-bit loop with branches it
        // replaces took 0.8 ms).
        {
            const u32 pmax = (u32)d_isqrt64((X1 - 1) >> 1);
            const u32 runs = (u32)NT * RQ_RUNS;
            const u32 per = (pmax + runs - 1) / runs;   // columns per run (uniform)
            const u32 ulen = (u32)len;
            // a run's columns are p0 .. p0 + per - 1, cut at pmax: as the walk takes
            // per steps, a run is at a column of its own exactly while p <= pmax
            u32 cp[RQ_RUNS], cq[RQ_RUNS], co[RQ_RUNS];
#pragma unroll
            for (int r = 0; r < RQ_RUNS; ++r) {
                const u64 p0 = 1 + (u64)(tid + (u32)r * NT) * per;
                cp[r] = (u32)min(p0, (u64)pmax + 1u);
                const u64 pp = (u64)cp[r] * cp[r];
                u64 q;
                if (X0 <= 2 * pp) {
                    q = cp[r];
                } else {
                    const u64 t = X0 - pp;   // in [1, 2^50)
                    q = d_isqrt52r(t);
                    if (q * q < t) ++q;
                }
                cq[r] = (u32)q;
                co[r] = (u32)(pp + q * q - X0);
            }
            // a point: offset in its window | p << RQ_LOGW | q << 38 (p < 2^25 and
            // q < 2^26 for S < 2^50), in the bin of its window
            auto point = [](u32 pc, u32 qc, u32 o) {
                return ((u64)qc << 38) | ((u64)pc << RQ_LOGW) | (o & (RQ_W - 1));
            };
            for (u32 st = 0; st < per; ++st) {
                u32 hm = 0;   // bit r: run r's column has a point at its least root
#pragma unroll
                for (int r = 0; r < RQ_RUNS; ++r)   // p <= pmax and off < len, as sign bits (|differences| < 2^31)
                    hm |= (((co[r] - ulen) & ~(pmax - cp[r])) >> 31) << r;
                if (__ballot(hm != 0u) != 0ull) {
                    // the slots of the first points, claimed before any is used (a lane
                    // without a point adds 0 to one of 32 counters, one per LDS bank)
                    u32 sl[RQ_RUNS];
#pragma unroll
                    for (int r = 0; r < RQ_RUNS; ++r) {
                        const bool h = (hm >> r) & 1u;   // o < len: the bin is below nb
                        sl[r] = atomicAdd(&s_bin[h ? co[r] >> RQ_LOGW : lane_id() & 31u], h ? 1u : 0u);
                    }
#pragma unroll
                    for (int r = 0; r < RQ_RUNS; ++r)   // (element indices below nb_alloc * bin_cap < 2^23)
                        if (((hm >> r) & 1u) && sl[r] < a.bin_cap)
                            scr[(co[r] >> RQ_LOGW) * a.bin_cap + sl[r]] = point(cp[r], cq[r], co[r]);
                    // the further points of these columns (several per column at low S)
#pragma unroll
                    for (int r = 0; r < RQ_RUNS; ++r) {
                        u32 o = co[r] + 2u * cq[r] + 1u, qq = cq[r] + 1u;
                        bool more = ((hm >> r) & 1u) && o < ulen;
                        while (__ballot(more) != 0ull) {
                            if (more) {
                                const u32 bin = o >> RQ_LOGW;
                                const u32 slot = atomicAdd(&s_bin[bin], 1u);
                                if (slot < a.bin_cap) scr[bin * a.bin_cap + slot] = point(cp[r], qq, o);
                            }
                            o += 2u * qq + 1u;
                            ++qq;
                            more = more && o < ulen;
                        }
                    }
                }
                // next column: p^2 grows by 2p + 1; then one step of q down, or onto
                // the diagonal (q was p - 1); runs past their last column go on
                // harmlessly (their columns never count)
#pragma unroll
                for (int r = 0; r < RQ_RUNS; ++r) {
                    // (the tests as sign bits of differences below 2^31 in size, and
                    // selects: VALU work only, nothing through the scalar unit)
                    const u32 pc = cp[r] + 1u, qc = cq[r];
                    const u32 t = 2u * qc - 1u;
                    u32 o = co[r] + 2u * cp[r] + 1u;
                    const bool dec = (((pc - qc) & ~(o - t)) >> 31) != 0u;   // q > p and o >= 2q - 1
                    const bool diag = ((qc - pc) >> 31) != 0u;               // q < p
                    o = o - (dec ? t : 0u) + (diag ? 2u * pc - 1u : 0u);
                    cq[r] = qc - (dec ? 1u : 0u) + (diag ? 1u : 0u);
                    cp[r] = pc;
                    co[r] = o;
                }
            }
        }
        __syncthreads();

        u64 c_pairs = 0, c_sge3 = 0, c_trip = 0, c_sets4 = 0, c_maxk = 0;
        for (u32 b = 0; b < nb; ++b) {
            const u64 B0 = X0 + ((u64)b << RQ_LOGW);
            u32 npts = uni(s_bin[b]);
            if (npts > a.bin_cap) { err |= ERR_BIN_OVF; npts = a.bin_cap; }
            const u64 *pts = scr + (u64)b * a.bin_cap;
            for (u32 i = tid; i < RQ_W / 8; i += NT) ((uint4 *)s_cnt)[i] = make_uint4(0u, 0u, 0u, 0u);
            __syncthreads();
            // pass 1: count (exhaustive/s4.c:285-298); UN points per thread
            // loaded together (independent loads in flight), then counted
            for (u32 i0 = tid; i0 < npts; i0 += UN * NT) {
                u32 sv[UN];
#pragma unroll
                for (u32 u = 0; u < UN; ++u) {
                    const u32 i = i0 + u * NT;
                    sv[u] = (u32)pts[i < npts ? i : i0] & (RQ_W - 1u);
                }
#pragma unroll
                for (u32 u = 0; u < UN; ++u) PS4_KEEP("v"(sv[u]));   // all loads issued before the branches
#pragma unroll
                for (u32 u = 0; u < UN; ++u)
                    if (i0 + u * NT < npts) atomicAdd(&s_cnt[sv[u] >> 1], 1u << ((sv[u] & 1u) << 4));
            }
            __syncthreads();
            // pass 2: which S count, offsets (exhaustive/s4.c:300-305, 314-321).
            // Thread t owns the SPT = W / 256 S from s0 = SPT t: SPT / 2 count
            // words, read as 16-byte words.  Counted S (k >= 3, S != 2 mod 4) are a few per cent,
            // so one sweep sums and maxes the counts two at a time and marks the
            // counted S in a 32-bit mask; only the marked S are visited after.
            constexpr u32 SPT = RQ_W / NT;
            static_assert(SPT % 8 == 0 && SPT <= 32, "whole 16-byte words, one 32-bit mask of counted S per thread");
            const u32 s0 = tid * SPT;
            const u32 b0 = (u32)B0;   // S mod 4 = (b0 + s) mod 4, and s0 = 0 mod 4
            // bit 15 (low half) / 31 (high half) set where that S may count: word
            // w holds S = s0 + 2w and s0 + 2w + 1, so the pattern alternates with w
            u32 okm[2];
#pragma unroll
            for (u32 par = 0; par < 2; ++par) {
                u32 m = 0x80008000u;
#pragma unroll
                for (u32 h = 0; h < 2; ++h)
                    if (((b0 + 2u * par + h) & 3u) == 2u) m &= ~(0x8000u << (16u * h));
                okm[par] = m;
            }
            uint4 *cw = (uint4 *)s_cnt + tid * (SPT / 8);
            uint4 cv[SPT / 8];
#pragma unroll
            for (u32 i = 0; i < SPT / 8; ++i) cv[i] = cw[i];
            u32 w_pairs = 0, pmax = 0, cm = 0;
#pragma unroll
            for (u32 w = 0; w < SPT / 2; ++w) {
                u32 &word = (w & 3u) == 0 ? cv[w >> 2].x : (w & 3u) == 1 ? cv[w >> 2].y : (w & 3u) == 2 ? cv[w >> 2].z
                                                                                                       : cv[w >> 2].w;
                w_pairs = PS4_SAD_U16(word, 0u, w_pairs);   // + both counts
                pmax = PS4_MAX_U16X2(pmax, word);
                // a count >= 3 sets the top bit of its half (counts stay below 0x8003,
                // so no carry crosses into the high half)
                const u32 g = (word + 0x7FFD7FFDu) & okm[w & 1u];
                cm |= (((g >> 15) & 1u) | ((g >> 30) & 2u)) << (2u * w);
                // the halves of S that do not count become 0xFFFF (the fill below skips
                // them); counted halves keep k until their offset replaces it
                word |= ~((g >> 15) * 0xFFFFu);
            }
            const u32 w_maxk = max(pmax & 0xFFFFu, pmax >> 16);
            const u32 nq = (u32)__builtin_popcount(cm);
            u32 qr = 0, tr = 0;
            bool big = false;
            for (u32 m = cm; m != 0u; m &= m - 1u) {
                const u32 i = (u32)__builtin_ctz(m);
                const u32 k = (s_cnt[(s0 + i) >> 1] >> ((i & 1u) << 4)) & 0xFFFFu;
                qr += k;
                if (k > a.cap_reps) err |= ERR_CAP_REPS;
                if (k > RQ_K_MAX) big = true;
                tr += tri3(k < RQ_K_MAX ? k : RQ_K_MAX);
            }
            c_pairs += w_pairs;
            c_maxk = w_maxk > c_maxk ? w_maxk : c_maxk;
            c_sge3 += nq;
            c_trip += tr;
            // two scans: counted S and their representations packed in one word
            // (each sum stays below 2^16: at most RQ_W S and bin_cap <= 4 RQ_W
            // points per window), and the triples with bit 31 set by a bucket too
            // large for the 32-bit triple index (without one, the sum of C(k,3)
            // over <= 2560 representations stays below 2^31)
            u32 tot_p, tot_t, b_p, b_tr;
            block_excl_scan2<NT>(nq | (qr << 16), big ? 0x80000000u : tr, s_w, &b_p, &b_tr, &tot_p, &tot_t);
            tot_p = uni(tot_p);
            tot_t = uni(tot_t);
            u32 b_nq = b_p & 0xFFFFu, b_qr = b_p >> 16;
            const u32 tot_nq = tot_p & 0xFFFFu, tot_qr = tot_p >> 16, tot_tr = tot_t;
            // window lists beyond the LDS capacity, or a bucket too large for the
            // 32-bit triple index: the window is not processed and the batch fails
            if ((tot_t >> 31) || tot_nq > RQ_QS_CAP || tot_qr > RQ_PV_CAP) {
                err |= ERR_LDS_OVF;
                continue;
            }
#pragma unroll
            for (u32 i = 0; i < SPT / 8; ++i) cw[i] = cv[i];
            // the counted S in order: bucket records, and the bucket offset in
            // place of k (the thread's own words)
            for (u32 m = cm; m != 0u; m &= m - 1u) {
                const u32 i = (u32)__builtin_ctz(m), s = s0 + i, sh = (i & 1u) << 4;
                u32 *pw = s_cnt + (s >> 1);
                const u32 word = *pw, k = (word >> sh) & 0xFFFFu;
                s_sk[b_nq] = s | (k << 16);
                s_st[b_nq] = b_qr;
                s_tp[b_nq] = b_tr;
                *pw = (word & ~(0xFFFFu << sh)) | (b_qr << sh);
                ++b_nq;
                b_qr += k;
                b_tr += tri3(k);
            }
            if (tid == 0) s_tp[tot_nq] = tot_tr;
            __syncthreads();
            // pass 3: fill the buckets of counted S only (exhaustive/s4.c:307-310)
            for (u32 i0 = tid; i0 < npts; i0 += UN * NT) {
                u64 pv[UN];
#pragma unroll
                for (u32 u = 0; u < UN; ++u) {
                    const u32 i = i0 + u * NT;
                    pv[u] = pts[i < npts ? i : i0];
                }
#pragma unroll
                for (u32 u = 0; u < UN; ++u) {
                    const u32 s = (u32)pv[u] & (RQ_W - 1u), sh = (s & 1u) << 4;
                    if (i0 + u * NT < npts && ((s_cnt[s >> 1] >> sh) & 0xFFFFu) != 0xFFFFu) {
                        const u32 old = atomicAdd(&s_cnt[s >> 1], 1u << sh);
                        const u32 slot = (old >> sh) & 0xFFFFu;
                        if (slot < tot_qr) {
                            s_pv[slot] = (u32)(pv[u] >> RQ_LOGW) & 0x1FFFFFFu;
                            s_qv[slot] = (u32)(pv[u] >> 38);
                        }
                    }
                }
            }
            __syncthreads();
            // pass 4: triples -> 4-sets, appended to the wave's tiles of the queue
            const u32 T = tot_tr;   // uniform
            if (T > 0) {
                const u32 c0 = (u32)(((u64)tid * T) / NT), c1 = (u32)(((u64)(tid + 1) * T) / NT);
                c_sets4 += rq_walk(c0, c1, (T + NT - 1) / NT, s_sk, s_st, s_tp, s_pv, s_qv, tot_nq, B0, tl,
                                   a.quads, a.quad_cap, a.ctl, item.rid, a.verify, &err);
            }
            __syncthreads();
        }
        // ---- counters of this superwindow
        const u64 t_pairs = block_sum_u64<NT>(c_pairs, s_w64);
        const u64 t_sge3 = block_sum_u64<NT>(c_sge3, s_w64);
        const u64 t_trip = block_sum_u64<NT>(c_trip, s_w64);
        const u64 t_sets4 = block_sum_u64<NT>(c_sets4, s_w64);
        const u64 t_maxk = block_max_u64<NT>(c_maxk, s_w64);
        if (err) atomicOr(&a.ctl->err, err);
        if (tid == 0) {
            unsigned long long *c = (unsigned long long *)(a.ctr + (u64)item.rid * CT_N);
            atomicAdd(c + CT_S_DONE, (unsigned long long)len);
            atomicAdd(c + CT_PAIRS, (unsigned long long)t_pairs);
            atomicAdd(c + CT_S_GE3, (unsigned long long)t_sge3);
            atomicAdd(c + CT_TRIPLES, (unsigned long long)t_trip);
            atomicAdd(c + CT_SETS4, (unsigned long long)t_sets4);
            atomicMax(c + CT_MAX_REPS, (unsigned long long)t_maxk);
        }
    }
    rq_pad(tl, a.quads, a.quad_cap, &err);
    if (err) atomicOr(&a.ctl->err, err);
}

// ----------------------------------------------------------------------------
// k_ext: extension of queued 4-sets
// ----------------------------------------------------------------------------

#define EX_NT 64          // one wave per workgroup: every barrier below is a wave's own
#define EX_ARENA 861      // doubles of divisor lists per pass, + 3 spare (the workgroup's LDS
                          // stays at 13632 bytes, so 12 workgroups share a CU's 160 KB); one
                          // 4-set with d(D) <= 16384 needs at most 748 (nL + nH over the best
                          // prefix split, reached by the radices 2, 11, 33, 2, 11)
#define EX_CMAX 8192      // candidates per pass; one 4-set with d(D) <= 16384 has <= 8192
#define EX_CHUNK_DEFAULT 4   // tiles of 64 4-sets claimed per work-queue atomic (a template
                             // parameter of k_ext: 1, 4 or 16)
#define EX_Q1 128         // LDS stack of candidates past the residue filters (a round of 64
                          // runs once it holds >= 64; each push adds at most 64)
#define EX_Q2 128         // queue of candidates for the exact test, per workgroup in global
                          // memory, kept across passes and tiles (drained at >= 64 and at
                          // the end; a round of the first queue adds at most 64)
#define EX_OSLOT 7        // distinct primes of an odd n < 2^26 (3*5*7*...*23 > 2^26)
#define EX_NPOS 15        // stage positions of D': the power of two, then EX_OSLOT slots for the
                          // odd part of ra - rb and EX_OSLOT for that of ra + rb

// The odd-factor table: for odd n < tab_n, entry (n - 1) / 2 is two 16-byte
// words {w0..w7}: w0..w6 the distinct prime powers of n as p << 6 | e in
// ascending order of p (0 past the last), w7 = k | d(n) << 4 with k the
// number of distinct primes.  One gather factors an odd part completely.
struct __align__(16) ExOdd {
    uint4 a, b;
};

struct ExArgs {
    const ps4_quad *quads;
    u64 quad_cap;
    ps4_ctl *ctl;
    const ExOdd *oftab;     // odd n -> its factorisation
    u64 tab_n;              // entries in ndiv (n < tab_n); oftab holds the odd n below it
    const uint16_t *ndiv;   // n -> d(n)
    const ps4_qrtab *qr;
    u64 *ctr;
    ps4_hit *hits;
    uint4 *exq;             // EX_Q2 words per workgroup: entries {u (52 bits) | pair choice << 52,
                            // 4-set index}
    u32 hit_cap;
    u32 cap_divs;
    u32 n_rids;             // counter rows in ctr
    double neg_m1;          // -262080 (a kernel argument, so that it sits in scalar registers)
};

// Per-4-set parameters of the candidate loop: one record of three 16-byte
// words per 4-set of a pass (read together, before any branch):
//   q0 = {aL | aL2 << 16,  aH | aH2 << 16,  magic,  (-8 nL mod 2^16) | 8 coff << 16}
//         aL, aL2: byte offsets in the arena of L[0] and L[nL-1]; aH, aH2: of
//         H[0] and H[nH-1]; coff < 8192: the 4-set's first candidate in the
//         pass; magic = ceil(2^23 / nL) gives t / nL = (t << 9) * magic >> 32
//         exactly for t < 8192 and nL < 1024 (checked for every pair on the host)
//   q1 = {mask mod 64 (lo, hi),  mask mod 63 (lo, hi)}
//   q2 = {mask mod 65, bit (u + 32) mod 65 mod 64 (lo, hi),  msm (lo, hi)}: msm
//         holds the masks mod 11 (bit 1), 23 folded (12), 17 (25), 19 (42);
//         bits 61..63 hold the pair choice
// Read only for queued candidates: s_mem[k] = {isqrt(A), F1, F2, F3} (the roots
// of the pair sums containing A), s_aux[k] = isqrt(2A) when 2A is a square,
// else 0, and s_qix[k] = the 4-set's queue index.
#define EX_PAR_WORDS 3


// bit r (< 64) of the 64-bit value hi:lo, as 0 or 1
__device__ __forceinline__ u32 bit64(u32 lo, u32 hi, u32 r) { return (u32)((((u64)hi << 32) | lo) >> r) & 1u; }

// x mod m for an integer-valued double |x| < 2^53; result in [0, m).
__device__ __forceinline__ double dmod(double x, double m, double inv_m)
{
    const double q = __builtin_floor(x * inv_m);
    double r = __builtin_fma(-q, m, x);
    r = r < 0.0 ? r + m : r;
    r = r >= m ? r - m : r;
    return r;
}

// a * b + c as one full-rate 24-bit multiply-add (a, b taken mod 2^24, as
// unsigned or as signed numbers): left to itself the compiler picks a 32-bit or
// a 64-bit multiply for these
#ifndef PS4_MAD_U24
#define PS4_MAD_U24(a, b, c)                                                              \
    ({                                                                                    \
        u32 r_;                                                                           \
        asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r_) : "v"((u32)(a)), "v"((u32)(b)), "v"((u32)(c))); \
        r_;                                                                               \
    })
#define PS4_MAD_I24(a, b, c)                                                              \
    ({                                                                                    \
        u32 r_;                                                                           \
        asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r_) : "v"((u32)(a)), "s"((int)(b)), "v"((u32)(c))); \
        r_;                                                                               \
    })
#endif

// (a mod 2^24) * (b mod 2^24) >> 32: the 24-bit multiply-high
__device__ __forceinline__ u32 mulhi_u24(u32 a, u32 b)
{
    return (u32)(((u64)(a & 0xFFFFFFu) * (u64)(b & 0xFFFFFFu)) >> 32);
}

// s mod 63 and s mod 65 for s < 2^20 (a representative of a class mod
// 262080 = 64*63*65): floor(s / m) = (8 s * ceil(2^29 / m)) >> 32 by the 24-bit
// multiply-high, exact because the rounding error s * 2^-29 stays below 1/m.
// s mod 11, 17, 19 and 23 for s < 2^17 (a representative of a class mod
// 81719 = 11*17*19*23) likewise: floor(s / m) = (128 s * ceil(2^25 / m)) >> 32.
// All constants are checked exhaustively on the host.  The multiply-high is a
// plain 64-bit product of the masked operands, not __umul24: with __umul24 the
// compiler emitted an arithmetic shift of the product, which is wrong once the
// product reaches 2^31 (a residue then came out near 90000 and indexed far past
// the residue tables).
__device__ __forceinline__ u32 mod63(u32 s) { return PS4_MAD_I24(mulhi_u24(s << 3, 8521761u), -63, s); }
__device__ __forceinline__ u32 mod65(u32 s) { return PS4_MAD_I24(mulhi_u24(s << 3, 8259553u), -65, s); }
__device__ __forceinline__ u32 mod11(u32 s) { return PS4_MAD_I24(mulhi_u24(s << 7, 3050403u), -11, s); }
__device__ __forceinline__ u32 mod17(u32 s) { return PS4_MAD_I24(mulhi_u24(s << 7, 1973791u), -17, s); }
__device__ __forceinline__ u32 mod19(u32 s) { return PS4_MAD_I24(mulhi_u24(s << 7, 1766023u), -19, s); }
__device__ __forceinline__ u32 mod23(u32 s) { return PS4_MAD_I24(mulhi_u24(s << 7, 1458889u), -23, s); }

#define M1 262080.0
#define M2 81719.0   // 11 * 17 * 19 * 23

// x mod M2 in [0, M2) for an integer-valued double |x| < 2^52 (the
// representative of rep_m1's kind, then two conditional subtractions as
// unsigned minima).
__device__ __forceinline__ u32 mod_m2(double x)
{
    const double q = __builtin_floor(x * (1.0 / M2));
    u32 r = (u32)(__builtin_fma(-q, M2, x) + M2);   // [0, 3*M2)
    r = r < r - 81719u ? r : r - 81719u;
    r = r < r - 81719u ? r : r - 81719u;
    return r & 0x1FFFFu;   // (no-op: tells the compiler r < 2^17, so 24-bit multiplies suffice)
}

// A representative in [0, 3*M1) of x mod M1 for an integer-valued double
// |x| < 2^52: the floor of the rounded quotient is within one of the true
// quotient, so x - q*M1 (exact by the fused multiply-add) lies in [-M1, 2*M1).
// Residues mod 64, 63 and 65 do not depend on which representative.
__device__ __forceinline__ u32 rep_m1(double x)
{
    const double q = __builtin_floor(x * (1.0 / M1));
    return (u32)(__builtin_fma(-q, M1, x) + M1);
}

// The chosen pair of a 4-set (exhaustive/s4.c:163-171 and the switch that
// follows): A = v_i, the other two members C1, C2 in index order, the three
// roots of pair sums containing A, and the pair roots ra > rb with
// D = v_i - v_j = (ra - rb)(ra + rb).  Members a > b > c > d (index 0..3).
struct ExPair {
    u64 A, C1, C2;
    u32 F1, F2, F3, ra, rb;   // roots (below 2^26)
};

__device__ __forceinline__ ExPair ex_pair(u32 best, u64 va, u64 vb, u64 vc, u64 vd, u32 rab, u32 rac, u32 rad,
                                          u32 rbc, u32 rbd, u32 rcd)
{
    // as selects (the lanes of a wave choose different pairs):
    //   best 0: A=va C1=vc C2=vd F=(rab,rac,rad) ra=rac rb=rbc   (i,j,k) = (0,1,2)
    //        1: A=vc C1=va C2=vb F=(rac,rbc,rcd) ra=rac rb=rad           (2,3,0)
    //        2: A=va C1=vb C2=vd F=(rab,rac,rad) ra=rab rb=rbc           (0,2,1)
    //        3: A=vb C1=va C2=vc F=(rab,rbc,rbd) ra=rab rb=rad           (1,3,0)
    //        4: A=va C1=vb C2=vc F=(rab,rac,rad) ra=rab rb=rbd           (0,3,1)
    //        5: A=vb C1=va C2=vd F=(rab,rbc,rbd) ra=rab rb=rac           (1,2,0)
    const bool b1 = best == 1, odd = (best & 1u) != 0, b35 = best == 3 || best == 5;
    ExPair p;
    p.A = b1 ? vc : (b35 ? vb : va);
    p.C1 = best == 0 ? vc : (odd ? va : vb);
    p.C2 = b1 ? vb : ((best == 3 || best == 4) ? vc : vd);
    p.F1 = b1 ? rac : rab;
    p.F2 = odd ? rbc : rac;
    p.F3 = b1 ? rcd : (b35 ? rbd : rad);
    p.ra = best < 2 ? rac : rab;
    p.rb = (best == 0 || best == 2) ? rbc : ((best == 1 || best == 3) ? rad : (best == 4 ? rbd : rac));
    return p;
}

// One surviving candidate: exact e = u^2 - A and the two square tests of
// exhaustive/s4.c:195-200.  Rare (after all residue filters), so out of line;
// A, C1, C2 are rederived from the queued 4-set.
__device__ __forceinline__ u32 ext_exact(double u, const ps4_quad *quads, u64 qidx, u32 best, ps4_ctl *ctl,
                                      ps4_hit *hits, u32 hit_cap)
{
    const ps4_quad qd = quads[qidx];
    // every field in registers here: no load of this entry stays in flight past
    // the drain (a pending load would make the candidate loop wait for all
    // outstanding memory operations, its queue stores included, every iteration)
    PS4_KEEP("v"(qd.r_ab), "v"(qd.r_ac), "v"(qd.r_ad), "v"(qd.r_bc), "v"(qd.r_bd), "v"(qd.r_cd), "v"(qd.rid));
    const u64 rab = qd.r_ab, rac = qd.r_ac, rad = qd.r_ad, rbc = qd.r_bc;
    const u64 va = (rab * rab + rac * rac - rbc * rbc) >> 1;
    const u64 vb = rab * rab - va, vc = rac * rac - va, vd = rad * rad - va;
    const ExPair P = ex_pair(best, va, vb, vc, vd, qd.r_ab, qd.r_ac, qd.r_ad, qd.r_bc, qd.r_bd, qd.r_cd);
    const u64 uu = (u64)u;
    const u128 e = (u128)uu * uu - P.A;
    if (!d_is_square128(e + P.C1) || !d_is_square128(e + P.C2)) return 0;
    const u32 slot = atomicAdd(&ctl->hit_count, 1u);
    if (slot >= hit_cap) return ERR_HIT_OVF;
    ps4_hit h;
    h.r[0] = qd.r_ab; h.r[1] = qd.r_ac; h.r[2] = qd.r_ad;
    h.r[3] = qd.r_bc; h.r[4] = qd.r_bd; h.r[5] = qd.r_cd;
    h.rid = qd.rid;
    h.pad = best;
    h.e_lo = (u64)e;
    h.e_hi = (u64)(e >> 64);
    hits[slot] = h;
    return 0;
}

// The entry of odd n in the odd-factor table (n < tab_n, checked by the caller).
__device__ __forceinline__ ExOdd ex_odd_at(const ExOdd *oftab, u32 n)
{
    return oftab[n >> 1];
}

// Does p divide the odd o (p, o < 2^26)?  The refined reciprocal has relative
// error far below 2^-40, so when p | o the rounded quotient is o / p exactly,
// and the fused multiply-add then gives 0 exactly; otherwise it cannot.
__device__ __forceinline__ bool ex_divides(u32 p, double o)
{
    const double dp = (double)p;
    double r = __builtin_amdgcn_rcp(dp);
    r = __builtin_fma(__builtin_fma(-dp, r, 1.0), r, r);
    const double q = __builtin_rint(o * r);
    return p > 1u && __builtin_fma(-q, dp, o) == 0.0;
}

// The prime powers of D' in stage order, as words p << 6 | e (0: no stage;
// six exponent bits: the power of two of D < 2^51 can exceed 31):
// position 0 the power of two 2^e2p, positions 1..7 those of the odd part o1
// of ra - rb (entry f1), positions 8..14 those of the odd part of ra + rb
// (entry f2) whose prime does not divide o1.  A prime dividing both (it then
// divides ra and rb) keeps one stage, in o1's position, with the two exponents
// added.  *nd gets d of the odd part of D.
__device__ __forceinline__ void ex_stages(const ExOdd &f1, const ExOdd &f2, u32 o1, u32 e2p, u32 st[EX_NPOS], u32 *nd)
{
    const u32 w1[EX_OSLOT] = {f1.a.x, f1.a.y, f1.a.z, f1.a.w, f1.b.x, f1.b.y, f1.b.z};
    const u32 w2[EX_OSLOT] = {f2.a.x, f2.a.y, f2.a.z, f2.a.w, f2.b.x, f2.b.y, f2.b.z};
    u32 d = (f1.b.w >> 4) * (f2.b.w >> 4);
    const double o = (double)o1;
    const u32 n2 = f2.b.w & 15u;   // slots of o2 in use (the others hold 0)
    // The merges are recorded in three values only (the exponents added to o1's
    // slots, six bits each; the merged o2 slots; d), so that the branches around
    // them, taken by some lane in most waves, do not carry the stage words; the
    // stage words are written once afterwards.
    u64 add1 = 0;
    u32 gone2 = 0;
#pragma unroll
    for (int j = 0; j < EX_OSLOT; ++j) {
        const u32 w = w2[j];
        if ((u32)j < n2 && ex_divides(w >> 6, o)) {   // a prime of both parts
            const u32 p = w >> 6, eb = w & 63u;
#pragma unroll
            for (int i = 0; i < EX_OSLOT; ++i) {
                if ((w1[i] >> 6) == p) {
                    // d: the factors (ea + 1)(eb + 1) of the two parts become ea + eb + 1.
                    // The quotient is exact in single precision: den divides d, and
                    // d <= d(o1) d(o2) <= 2^16 (an odd n < 2^26 + 64 has at most 256
                    // divisors), so the hardware reciprocal (1 ulp) leaves an error
                    // below 2^-6 before rounding
                    const u32 ea = w1[i] & 63u;
                    const float den = (float)((ea + 1u) * (eb + 1u));
                    d = (u32)__builtin_rintf((float)d * __builtin_amdgcn_rcpf(den)) * (ea + eb + 1u);
                    add1 += (u64)eb << (6 * i);
                }
            }
            gone2 |= 1u << j;
        }
    }
    st[0] = e2p ? (2u << 6) | e2p : 0u;
#pragma unroll
    for (int i = 0; i < EX_OSLOT; ++i) st[1 + i] = w1[i] + (u32)((add1 >> (6 * i)) & 63u);
#pragma unroll
    for (int j = 0; j < EX_OSLOT; ++j) st[1 + EX_OSLOT + j] = ((gone2 >> j) & 1u) ? 0u : w2[j];
    *nd = d;
}

// The largest value of v (0..7) over the wave, uniform.
__device__ __forceinline__ u32 wave_max7(u32 v)
{
    u32 m = 0;
    for (u32 t = 7; t > 0; --t)
        if (__ballot(v >= t) != 0ull) { m = t; break; }
    return m;
}

// Stage position s is walked by the loops below when it can hold a stage of
// some lane: the power of two, and the first n1 / n2 slots of the two odd parts
// (n1, n2: the largest numbers of distinct primes of the parts in the wave).
__device__ __forceinline__ bool ex_pos_live(int s, u32 n1, u32 n2)
{
    return s == 0 || (s <= EX_OSLOT ? (u32)(s - 1) < n1 : (u32)(s - 1 - EX_OSLOT) < n2);
}

// One candidate: 4-set record k of the pass, candidate index t of that 4-set.
// Straight-line: the record and the four list entries are loaded before the
// tests (ex_lists), and every test is a select.  ok: the candidate
// passes the residue filters mod 64, 63, 65.  The queued ok candidates then
// get the remaining tests of the ext_sqrt count (e > 0, e not a member) and
// the later filters (ex_drain1).
struct ExCand {
    double us;   // u + EX_UOFF
    bool ok;
};

// The candidate loop carries u + EX_UOFF: 64512 = 16 * 64 * 63 is 0 mod 64 and
// mod 63 and 32 mod 65, so its residue r65 mod 65 is (u + 32) mod 65, whose
// values 64 and 0 (u = 32 and u = -32) have the same square: the mod-65 test
// is one 64-bit shift by r65 mod 64.  The offset rides in the first product's
// fused add, and the stack entries hold u + EX_UOFF.
#define EX_UOFF 64512.0
// 1/262080 lowered by 2^-45 relative (with the two roundings of this
// expression, by 2^-45 +- 2^-52): for an integer 0 <= x < 2^53 the rounded
// product x * EX_INV_M1 lies below x / 262080 and within 2^-9 of it, so its
// floor q is floor(x / 262080) or one less, and x - 262080 q is in [0, 2 * 262080).
#define EX_INV_M1 ((1.0 / 262080.0) * (1.0 - 0x1p-45))

// The four list entries of candidate t: L[i], H[j], L[nL-1-i], H[nH-1-j].
struct ExLists {
    double l0, h0, l1, h1;
};

// The double at byte offset off of the arena.
__device__ __forceinline__ double ar_at(const double *s_ar, u32 off)
{
    return *(const double *)((const char *)s_ar + off);
}

// g8 = 8 g, g the candidate's index in the pass; t = g - coff is its index in
// its 4-set, i = t mod nL, j = t / nL; q0: the first word of the 4-set's record
// (with 8 coff in its top half).
__device__ __forceinline__ ExLists ex_lists(const uint4 &q0, const double *s_ar, u32 g8)
{
    const u32 t8 = g8 - (q0.w >> 16);
    const u32 j = mulhi_u24(t8 << 6, q0.z);
    const u32 i8 = PS4_MAD_U24(j, q0.w, t8) & 0xFFFFu;   // 8 (t - j nL) < 2^16 from the low halves
    const u32 j8 = j << 3;
    return ExLists{ar_at(s_ar, (q0.x & 0xFFFFu) + i8), ar_at(s_ar, (q0.y & 0xFFFFu) + j8),
                   ar_at(s_ar, (q0.x >> 16) - i8), ar_at(s_ar, (q0.y >> 16) - j8)};
}

// q1, q2: the mask words of the 4-set's record.  neg_m1 = -M1, passed in by
// the launch so that it sits in scalar registers.
__device__ __forceinline__ ExCand ex_cand(const uint4 &q1, const uint4 &q2, const ExLists &Ls, double neg_m1)
{
    ExCand c;
    c.us = __builtin_fma(Ls.l1, Ls.h1, __builtin_fma(Ls.l0, Ls.h0, EX_UOFF));   // H holds halves when u = (f+g)/2
    const double q = __builtin_floor(c.us * EX_INV_M1);
    const u32 r1 = (u32)__builtin_fma(q, neg_m1, c.us);   // (u + EX_UOFF) mod 262080, + 0 or 262080: below 2^20
    const u32 r63 = mod63(r1), r65 = mod65(r1);
    u32 qrb = bit64(q1.x, q1.y, r1 & 63u) & bit64(q1.z, q1.w, r63) & bit64(q2.x, q2.y, r65 & 63u);
    PS4_OPAQUE(qrb);   // (else the ballot and the branch on c.ok get a test each)
    c.ok = qrb != 0u;
    return c;
}

// Stack entries: an integer-valued us = u + EX_UOFF in [0, 2^52) as the
// mantissa of us + 2^52 (exact below 2^53), whose exponent field 0x433 gets
// the record index k < 64 added (one add to the high word).
__device__ __forceinline__ u64 ex_qent(double us, u32 k)
{
    const u64 b = __builtin_bit_cast(u64, us + 4503599627370496.0);
    return ((u64)((u32)(b >> 32) + (k << 20)) << 32) | (u32)b;
}
__device__ __forceinline__ u32 ex_qk(u64 e) { return (u32)(e >> 52) - 0x433u; }
// u from the low 52 bits of an entry (us = u + EX_UOFF)
__device__ __forceinline__ double ex_qu(u64 e)
{
    return __builtin_bit_cast(double, (e & ((1ull << 52) - 1ull)) | 0x4330000000000000ull) -
           (4503599627370496.0 + EX_UOFF);
}

// The rest of the ext_sqrt tests for a queued candidate u of 4-set record k:
// e = u^2 - A > 0 (u > isqrt(A)) and e not a member (u is not a root of a pair
// sum containing A, and u^2 != 2A); every root is below 2^26 for S < 2^50.
__device__ __forceinline__ bool ex_counted(double u, const uint4 &m, u32 aux)
{
    const u32 ui = (u32)__builtin_fmin(u, 4294967295.0);
    return (ui > m.x) & (ui != m.y) & (ui != m.z) & (ui != m.w) & (ui != (aux & 0x3FFFFFFu));
}

// The later filters of a counted candidate u of record word q2: the residue
// masks mod 11, 23, 17, 19 (e + C1 and e + C2 both squares mod each; the
// mask mod 23 folded, as u and -u pass alike: bit min(r, 23 - r)).  True: it
// needs the exact test.  (No mask mod 13: it would reject nothing, a square mod
// 65 being a square mod 13, and every counted candidate passed the mod-65 mask.)
__device__ __forceinline__ bool ex_late(double u, const uint4 &q2)
{
    const u32 r2 = mod_m2(u);
    const u32 r23 = mod23(r2);
    const u32 m2b = bit64(q2.z, q2.w, 1u + mod11(r2)) & bit64(q2.z, q2.w, 12u + min(r23, 23u - r23)) &
                    bit64(q2.z, q2.w, 25u + mod17(r2)) & bit64(q2.z, q2.w, 42u + mod19(r2));
    return m2b != 0u;
}

__device__ __forceinline__ u64 wave_sum_u64(u64 v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += shfl_xor_u64(v, m);
    return v;
}
__device__ __forceinline__ u64 wave_max_u64(u64 v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) { const u64 o = shfl_xor_u64(v, m); v = o > v ? o : v; }
    return v;
}
// exclusive prefix sum over the wave; *total gets the sum (uniform)
__device__ __forceinline__ u32 wave_excl_scan(u32 v, u32 *total)
{
    const u32 incl = wave_incl_scan(v);
    *total = (u32)__builtin_amdgcn_readlane((int)incl, 63);
    return incl - v;
}

// Add a wave's accumulated counters of interval `rid` to the device counters.
__device__ __forceinline__ void ex_flush(u64 *ctr, u32 n_rids, u32 rid, u64 cand, u64 esq, u64 md)
{
    const u64 tc = wave_sum_u64(cand), te = wave_sum_u64(esq), tm = wave_max_u64(md);
    if (threadIdx.x == 0 && rid < n_rids) {
        unsigned long long *c = (unsigned long long *)(ctr + (u64)rid * CT_N);
        atomicAdd(c + CT_EXT_CAND, (unsigned long long)tc);
        atomicAdd(c + CT_EXT_SQRT, (unsigned long long)te);
        atomicMax(c + CT_MAX_DIVS, (unsigned long long)tm);
    }
}

// Workgroup = one wave; a tile is 64 queued 4-sets, one per lane for the setup:
//   validation, pair choice, the factorisation of D = 2^e2 * o1 * o2 (o1, o2
//   the odd parts of ra - rb and ra + rb) by one gather of each odd part's
//   entry in the odd-factor table, d(D), the candidate count ncand = d(D')/2
//   (exhaustive/s4.c:172-186), the split of the stage sequence into L and H
//   parts, residue masks, e > 0 and e-member bounds.
// Then passes: the lists L and H (all divisors are L[i]*H[j]) of as many
// 4-sets as the arena holds, and their candidates spread over the lanes with
// stride 64 (lane l takes candidates l, l+64, ...; the 4-set of a candidate
// comes from a bitmap of 4-set starts).  Candidate t in [0, ncand) of a 4-set
// is the divisor pair
//   f = L[i] * H[j],  g = D'/f = L[nL-1-i] * H[nH-1-j],   t = i + nL*j,
// i.e. the index pair (t, d(D')-1-t) of the mixed-radix divisor list: every
// unordered pair f != g exactly once, and f + g needs no ordering.
// Counters are summed per wave over consecutive tiles of one interval and
// added to the device counters when the interval changes and at the end.
template <int EX_CHUNK>
__global__ __launch_bounds__(EX_NT, 3) void k_ext(ExArgs a)   // 3 waves per SIMD: <= 168 VGPRs
{
    __shared__ double s_ar[EX_ARENA + 3];      // L and H lists, exact doubles (+ 3 spare: see the list building)
    __shared__ uint4 s_par[(EX_NT + 1) * EX_PAR_WORDS];   // parameters of the 4-sets of the current pass,
                                                          // and the guard record after them
    __shared__ uint4 s_mem[EX_NT];             // their e > 0 and member bounds
    __shared__ u32 s_aux[EX_NT];               // isqrt(2A) or 0
    __shared__ u32 s_qix[EX_NT];               // their queue indices
    __shared__ uint4 s_bits4[EX_CMAX / 128 + 1];   // bit g: a 4-set's candidates start at g + 1
    __shared__ u64 s_q1[EX_Q1];                // candidates past the residue filters: u | record << 52
    u32 *s_bits = (u32 *)s_bits4;

    const u32 tid = threadIdx.x;
    const u64 qc = a.ctl->quad_count;   // written by the previous kernel on this stream
    u64 nquads = ((u64)uni((u32)(qc >> 32)) << 32) | uni((u32)qc);
    if (nquads > a.quad_cap) nquads = a.quad_cap;
    const u64 ntiles = (nquads + EX_NT - 1) / EX_NT;
    const u64 nchunks = (ntiles + EX_CHUNK - 1) / EX_CHUNK;
    const ps4_qrtab *qr = a.qr;
    uint4 *exq2 = a.exq + (u64)blockIdx.x * EX_Q2;   // candidates for the exact test
    u32 qn2 = 0;                                     // entries in it

    // this wave's counters of interval acc_rid (lane partial sums)
    u32 acc_rid = 0xFFFFFFFFu;
    u64 acc_cand = 0, acc_esq = 0, acc_md = 0;
    u32 err = 0;

    // exact tests of every entry of the second queue
    auto drain2 = [&]() {
        __syncthreads();   // the queue's stores are visible to every lane
        for (u32 q = 0; q < qn2; q += EX_NT) {
            if (q + tid < qn2) {
                const uint4 e = exq2[q + tid];
                const double u = ex_qu((u64)e.x | ((u64)(e.y & 0xFFFFFu) << 32));
                err |= ext_exact(u, a.quads, (u64)e.z | ((u64)e.w << 32), e.y >> 20, a.ctl, a.hits, a.hit_cap);
            }
        }
        __syncthreads();   // read before the next stores
        qn2 = 0;
    };

    // ---- The wave's stream of queued 4-sets: chunks of EX_CHUNK tiles, claimed
    // one ahead.  Each lane holds one set-up 4-set until a pass takes it, and
    // then the next 4-set of the stream, so every pass fills the arena with the
    // 4-sets the lanes hold (a whole tile need not be used up first).
    u32 chunk = 0, nxt = 0;
    if (tid == 0) {
        chunk = atomicAdd(&a.ctl->tile_next, 1u);
        nxt = atomicAdd(&a.ctl->tile_next, 1u);
    }
    chunk = uni(chunk);   // lane 0's values (every lane is active here)
    nxt = uni(nxt);
    const u64 CH = (u64)EX_CHUNK * EX_NT;   // 4-sets per chunk
    u64 s_pos = (u64)chunk * CH;            // next stream index to hand out
    u64 s_end = chunk < nchunks ? min(s_pos + CH, nquads) : s_pos;   // end of the current chunk (s_pos: none left)
    // The next stream indices for the lanes of mask, in lane order (~0: none
    // left); every lane must be active.
    auto take = [&](u64 mask) -> u64 {
        const u32 k = (u32)__popcll(mask), r = lanes_below(mask);
        const u64 avail = s_end - s_pos;
        u64 qi = ~0ull;
        if ((mask >> tid) & 1ull) {
            if (r < avail) {
                qi = s_pos + r;
            } else if (nxt < nchunks) {
                const u64 b = (u64)nxt * CH + (r - avail);
                if (b < nquads) qi = b;
            }
        }
        if ((u64)k < avail) {
            s_pos += k;
        } else if (nxt < nchunks) {   // the current chunk is used up: move to the claimed one
            s_end = min((u64)nxt * CH + CH, nquads);
            s_pos = min((u64)nxt * CH + (k - avail), s_end);
            u32 n = 0;
            if (tid == 0) n = atomicAdd(&a.ctl->tile_next, 1u);
            nxt = uni(n);
        } else {
            s_pos = s_end;
        }
        return qi;
    };

    // the lane's set-up 4-set (has: one waits for a pass) and its parameters
    bool has = false;
    u32 qix = 0, lrid = 0;                  // its queue index and interval
    u32 ncand = 0, lsize = 0, nL = 0;       // candidates, nL + nH and nL of its lists
    u32 lpk = 0;                            // sL | e2p << 4 | half << 10 | pair choice << 11 | d(D) << 14
    u32 o1n = 0, o2n = 0;                   // odd parts of ra - rb, ra + rb | their numbers of primes << 28
    u64 pm64 = 0, pm63 = 0, pm65 = 0, pmsm = 0;
    u32 psqA = 0, pf1 = 0, pf2 = 0, pf3 = 0, pf4 = 0;
    // the entry each free lane sets up next, loaded ahead (index pqi, ~0: none)
    u64 pqi = take(~0ull);
    ps4_quad qpf = {0u, 0u, 0u, 0u, 0u, 0u, PS4_RID_NONE, 0u};
    if (pqi != ~0ull) qpf = a.quads[pqi];
    for (;;) {
        // ---- lanes without a 4-set set up their next one
        const bool refill = !has;
        bool valid = refill && pqi != ~0ull;
        const ps4_quad qd = qpf;
        valid = valid && qd.rid != PS4_RID_NONE;   // padding of a window's block
        u64 my_maxdiv = 0;
        u32 sL = 0, e2p = 0, half = 0, best = 0, nH = 0;   // this round's values (lpk keeps them)
        u32 e2 = 0, ndp = 0, n1 = 0, n2 = 0;
        u32 st[EX_NPOS];                        // stage words of D' (ex_stages)
#pragma unroll
        for (int q = 0; q < EX_NPOS; ++q) st[q] = 0u;
        if (refill) {
            ncand = 0;
            qix = (u32)pqi;
            lrid = qd.rid;
        }
        // The odd-factor entries of D's two odd parts: gathered below for the
        // lanes set up now, and gathered again here for the lanes that hold a
        // 4-set over from an earlier pass (stage words are not kept in registers
        // across the candidate loop), so that one merge into stage words serves
        // both kinds of lane.
        const bool carried = has;
        ExOdd f1 = {}, f2 = {};
        if (carried) {
            f1 = ex_odd_at(a.oftab, o1n & 0x0FFFFFFFu);
            f2 = ex_odd_at(a.oftab, o2n & 0x0FFFFFFFu);
        }
        u32 o1 = 1;
        if (valid) {
            u32 o2 = 1;
                // The divisor counts of the pair choice below, gathered first so
                // that their latency overlaps the checks; each index is clamped
                // into the table (a no-op for a well-formed 4-set, whose indices
                // are below rab + rac < tab_n), so no entry can read outside it.
                // (the roots and the table indices are 32-bit: tab_n < 2^27)
                const u32 ra_[6] = {qd.r_ac, qd.r_ac, qd.r_ab, qd.r_ab, qd.r_ab, qd.r_ab};
                const u32 rb_[6] = {qd.r_bc, qd.r_ad, qd.r_bc, qd.r_ad, qd.r_bd, qd.r_ac};
                const u32 nlast = (u32)a.tab_n - 1u;
                u32 nd_[12];
#pragma unroll
                for (int t = 0; t < 6; ++t) {
                    // (32-bit byte offsets: the loads take the table base from scalar registers)
                    nd_[2 * t] = *(const uint16_t *)((const char *)a.ndiv + 2u * min(ra_[t] - rb_[t], nlast));
                    nd_[2 * t + 1] = *(const uint16_t *)((const char *)a.ndiv + 2u * min(ra_[t] + rb_[t], nlast));
                }
                const u64 rab = qd.r_ab, rac = qd.r_ac, rad = qd.r_ad, rbc = qd.r_bc, rbd = qd.r_bd, rcd = qd.r_cd;
                const u64 va = (rab * rab + rac * rac - rbc * rbc) >> 1;
                const u64 vb = rab * rab - va, vc = rac * rac - va, vd = rad * rad - va;
                // every root difference and sum below indexes the tables: a queued
                // 4-set must satisfy its identities and lie inside them, else it is
                // reported (ERR_BADQUAD) and not extended
                const bool good = qd.rid < a.n_rids && va > vb && vb > vc && vc > vd && vd > 0 &&
                                  vb + vc == rbc * rbc && vb + vd == rbd * rbd && vc + vd == rcd * rcd &&
                                  va + vd == rad * rad && rab + rac < a.tab_n && rac > rbc && rac > rad &&
                                  rab > rbc && rab > rad && rab > rbd && rab > rac;
                if (!good) {
                    err |= ERR_BADQUAD;
                    valid = false;
                }
                if (valid) {
                    // Pair choice: fewest estimated divisors d(ra-rb)*d(ra+rb) of
                    // D = v_i - v_j, first minimum in the scan order of
                    // exhaustive/s4.c:163-171 (members are a > b > c > d, so the
                    // reference's swap never fires).
                    u32 bestv = 0xFFFFFFFFu;
#pragma unroll
                    for (int t = 0; t < 6; ++t) {
                        const u32 e = nd_[2 * t] * nd_[2 * t + 1];
                        if (e < bestv) { bestv = e; best = (u32)t; }
                    }
                    const ExPair P = ex_pair(best, va, vb, vc, vd, qd.r_ab, qd.r_ac, qd.r_ad, qd.r_bc, qd.r_bd, qd.r_cd);
                    // D = m1 * m2 = 2^e2 * o1 * o2: the odd parts are factored by one
                    // gather each, issued ahead of the residue tables
                    const u32 m1 = P.ra - P.rb, m2 = P.ra + P.rb;
                    const u32 z1 = (u32)__builtin_ctz(m1), z2 = (u32)__builtin_ctz(m2);
                    e2 = z1 + z2;
                    o1 = m1 >> z1;
                    o2 = m2 >> z2;
                    o1n = o1;
                    o2n = o2;
                    f1 = ex_odd_at(a.oftab, o1);
                    f2 = ex_odd_at(a.oftab, o2);
                    // quadratic-residue masks of this 4-set (read together with the
                    // odd-factor entries, so that the reads overlap): bit r of M_m set iff
                    // r^2 + (C1-A) and r^2 + (C2-A) are both squares mod m (the
                    // filters of exhaustive/s4.c:192-194)
                    const double d1 = (double)(i64)(P.C1 - P.A), d2 = (double)(i64)(P.C2 - P.A);
                    // representatives mod 262080 (residues mod 64, 63, 65 follow) and
                    // exact residues mod 81719; |C - A| < 2^51
                    // The residues index small tables.  The representatives are
                    // clamped to the domains on which every residue below is exact
                    // and in range (all s < 2^20 for mod 63 and 65, s < 81719 for
                    // mod 11..23; checked exhaustively on the host), so no index
                    // can leave its table, whatever the input; for |C - A| < 2^51
                    // the clamps change nothing.
                    const u32 r1a = rep_m1(d1) & 0xFFFFFu, r1b = rep_m1(d2) & 0xFFFFFu;
                    const u32 r2a = min(mod_m2(d1), 81718u), r2b = min(mod_m2(d2), 81718u);
                    const u32 ix[14] = {r1a & 63u, r1b & 63u, mod63(r1a), mod63(r1b), mod65(r1a), mod65(r1b), mod11(r2a),
                                        mod11(r2b), mod17(r2a), mod17(r2b), mod19(r2a), mod19(r2b), mod23(r2a), mod23(r2b)};
                    // (read here, combined after the walk)
                    const u64 t64a = qr->t64[ix[0]], t64b = qr->t64[ix[1]], t63a = qr->t63[ix[2]],
                              t63b = qr->t63[ix[3]], t65a = qr->t65[ix[4]], t65b = qr->t65[ix[5]];
                    const u32 t11a = qr->t11[ix[6]], t11b = qr->t11[ix[7]], t17a = qr->t17[ix[8]], t17b = qr->t17[ix[9]],
                              t19a = qr->t19[ix[10]], t19b = qr->t19[ix[11]], t23a = qr->t23[ix[12]],
                              t23b = qr->t23[ix[13]];
                    // e = u^2 - A > 0 <=> u > isqrt(A);  e = v_m <=> u^2 = A + v_m,
                    // i.e. u is a root of a pair sum containing A, or u^2 = 2A
                    // (every root is below 2^26 for S < 2^50)
                    const u64 sa = d_isqrt52r(P.A), s2a = d_isqrt52r(2 * P.A);
                    psqA = (u32)sa;
                    pf1 = P.F1;
                    pf2 = P.F2;
                    pf3 = P.F3;
                    pf4 = (s2a * s2a == 2 * P.A) ? (u32)s2a : 0u;   // u >= 2: 0 never matches
                    pm64 = t64a & t64b;
                    pm63 = t63a & t63b;
                    pm65 = t65a & t65b;
                    pmsm = ((u64)(t11a & t11b) << 1) | ((u64)(t23a & t23b) << 12) |
                           ((u64)(t17a & t17b) << 25) | ((u64)(t19a & t19b) << 42);
                    // Divisor pairs f < g of D' = D (D odd, u = (f+g)/2) or D/4
                    // (4 | D, u = f+g); none when D = 2 mod 4.  Their number is the
                    // reference's ext_cand.
                    half = (e2 == 0);
                    e2p = e2 >= 2 ? e2 - 2 : 0u;
                }
            }
        if (valid || carried) {
            // stage words of D' (a lane carried over keeps its o1 and e2p)
            u32 ndodd;   // d of the odd part (< 2^32 for D < 2^51)
            ex_stages(f1, f2, carried ? (o1n & 0x0FFFFFFFu) : o1, carried ? ((lpk >> 4) & 63u) : e2p, st, &ndodd);
            if (valid) {
                n1 = f1.b.w & 15u;
                n2 = f2.b.w & 15u;
                o1n |= n1 << 28;
                o2n |= n2 << 28;
                const u64 ndD = (u64)(e2 + 1) * ndodd;
                my_maxdiv = ndD;
                if (ndD > a.cap_divs) err |= ERR_CAP_DIVS;
                if (e2 != 1 && ndD <= a.cap_divs) {
                    ndp = (e2p + 1) * ndodd;
                    ncand = ndp / 2;
                }
            }
        }
        // Split the stage sequence into a prefix L (positions < sL) and the
        // rest H, minimising |L| + |H| = nL + nH over the prefixes; ndp / nl is
        // exact (nl divides ndp < 2^24) by a rounded single-precision quotient.
        // Positions no lane set up now uses are skipped (uniform).
        {
            const u32 n1max = wave_max7(n1), n2max = wave_max7(n2);
            if (ncand > 0 && refill) {
                const float fndp = (float)ndp;
                u32 nl = 1, bestc = 1 + ndp, bestnl = 1;
                sL = 0;
#pragma unroll
                for (int q = 0; q < EX_NPOS; ++q) {
                    if (!ex_pos_live(q, n1max, n2max)) continue;
                    nl *= (st[q] & 63u) + 1u;
                    const u32 nh = (u32)__builtin_rintf(fndp * __builtin_amdgcn_rcpf((float)nl));
                    const u32 c = nl + nh;
                    if (c < bestc) { bestc = c; sL = (u32)q + 1u; bestnl = nl; }
                }
                nL = bestnl;
                nH = bestc - bestnl;
                lsize = bestc;
                if (nL * nH != ndp || nl != ndp || ncand > EX_CMAX) { err |= ERR_FACTORS; ncand = 0; }   // never
            }
        }
        // A 4-set set up now that yields no candidates is counted at once (into
        // its own interval: rare); the others are counted by the pass that
        // takes them (ext_cand, the divisor maximum and ext_sqrt together).
        if (valid && ncand == 0 && lrid < a.n_rids)
            atomicMax((unsigned long long *)(a.ctr + (u64)lrid * CT_N + CT_MAX_DIVS), (unsigned long long)my_maxdiv);
        if (refill && ncand > 0) lpk = sL | (e2p << 4) | (half << 10) | (best << 11) | ((u32)my_maxdiv << 14);
        sL = lpk & 15u;
        e2p = (lpk >> 4) & 63u;
        half = (lpk >> 10) & 1u;
        best = (lpk >> 11) & 7u;
        const u32 lmd = lpk >> 14;
        if (refill) has = ncand > 0;
        if (__ballot(has) == 0ull) {      // nothing to extend: take more, or stop
            if (s_pos >= s_end) break;
            pqi = take(~0ull);
            if (pqi != ~0ull) qpf = a.quads[pqi];
            continue;
        }

        // ---- one pass: the divisor lists of as many held 4-sets as the arena
        //      holds, then all their candidates spread over the lanes
        {
            // a pass takes 4-sets of one interval: that of the first lane holding
            // one (so its counters go to the wave's counters of that interval)
            const u64 hm = __ballot(has);
            const u32 prid = (u32)__builtin_amdgcn_readlane((int)lrid, (int)__builtin_ctzll(hm));
            const bool pending = has && lrid == prid;
            if (prid != acc_rid) {
                ex_flush(a.ctr, a.n_rids, acc_rid, acc_cand, acc_esq, acc_md);
                acc_cand = acc_esq = acc_md = 0;
                acc_rid = prid;
            }
            u32 ltot, ptot, ctot;
            const u32 loff = wave_excl_scan(pending ? lsize : 0u, &ltot);
            const u32 coff = wave_excl_scan(pending ? ncand : 0u, &ptot);
            bool fits = pending && loff + lsize <= EX_ARENA && coff + ncand <= EX_CMAX;
            bool drop = false;
            if (pending && loff == 0 && !fits) { err |= ERR_ARENA; drop = true; }   // never (bounds above)
            if (fits) {
                acc_cand += ncand;
                acc_md = lmd > acc_md ? (u64)lmd : acc_md;
            }
            const u64 fmask = __ballot(fits);
            const u32 slot = (u32)__popcll(fmask & ((1ull << tid) - 1ull));
            (void)wave_excl_scan(fits ? ncand : 0u, &ctot);
            // the stage positions the fitting 4-sets use (uniform)
            const u32 p1max = wave_max7(fits ? (o1n >> 28) : 0u), p2max = wave_max7(fits ? (o2n >> 28) : 0u);
            u32 w_esq = 0, my_unc = 0;   // ext_sqrt: residue-filter passes (the wave's, from its ballots), minus
                                         // the queued ones that fail the rest (per lane)
            // fitting 4-sets are a prefix of the pending ones: coff is each one's
            // first candidate in the pass
            // whole 128-bit words, and one more: the fetch past the pass's last
            // iteration reads it
            for (u32 w = tid; w < ((ctot + 127) / 128 + 1) * 4; w += EX_NT) s_bits[w] = 0u;
            __syncthreads();
            // The guard record after the pass's records: the candidates of the last
            // iteration past ctot (fewer than 128) fall to it, so the loop needs no
            // range tests.  Its masks are 0 (no candidate passes) and its lists are
            // the first 256 entries of the arena read as L = {x}, H of 256 entries
            // (nL = 1: j = t, i = 0); its first candidate is 128 before ctot (or
            // 0), so t < 256 and every read stays inside the arena.
            if (tid == 0 && ctot > 0u) {
                uint4 *pw = s_par + (u32)__popcll(fmask) * EX_PAR_WORDS;
                const u32 gc = ctot >= 128u ? ctot - 128u : 0u;
                pw[0] = make_uint4(2040u << 16, 2040u << 16, 0x800000u, ((0u - 8u) & 0xFFFFu) | ((8u * gc) << 16));
                pw[1] = make_uint4(0u, 0u, 0u, 0u);
                pw[2] = make_uint4(0u, 0u, 0u, 0u);
                atomicOr(&s_bits[(ctot - 1u) >> 5], 1u << ((ctot - 1u) & 31u));
            }
                if (fits) {
                    uint4 *pw = s_par + slot * EX_PAR_WORDS;
                    const u32 aL = 8u * loff, aH = 8u * (loff + nL);
                    pw[0] = make_uint4(aL | ((aL + 8u * (nL - 1u)) << 16), aH | ((aH + 8u * (lsize - nL - 1u)) << 16),
                                       (0x800000u + nL - 1u) / nL, ((0u - 8u * nL) & 0xFFFFu) | ((8u * coff) << 16));
                    pw[1] = make_uint4((u32)pm64, (u32)(pm64 >> 32), (u32)pm63, (u32)(pm63 >> 32));
                    pw[2] = make_uint4((u32)pm65, (u32)(pm65 >> 32), (u32)pmsm, (u32)(pmsm >> 32) | (best << 29));
                    s_mem[slot] = make_uint4(psqA, pf1, pf2, pf3);
                    s_aux[slot] = pf4;
                    s_qix[slot] = qix;
                    if (coff > 0u) atomicOr(&s_bits[(coff - 1u) >> 5], 1u << ((coff - 1u) & 31u));
                    // mixed-radix lists over the stages: stage s multiplies the list
                    // so far by p_s^1 .. p_s^e (earlier stages are the less
                    // significant digits, so index n and n_max - n are complements);
                    // positions before sL build L, the rest H.  When u = (f + g) / 2
                    // the H list holds halves (exact doubles), so that
                    // u = L[i] * H[j] + L[i'] * H[j'] in both cases.
                    const u32 *sw = st;
                    const double h0 = half ? 0.5 : 1.0;
                    double *lst = s_ar + loff;
                    u32 n = 1;
                    lst[0] = 1.0;
                    bool inH = false;
#pragma unroll
                    for (int q = 0; q < EX_NPOS; ++q) {
                        if (!ex_pos_live(q, p1max, p2max)) continue;
                        if (!inH && (u32)q >= sL) {   // start the H list
                            lst = s_ar + loff + nL;
                            n = 1;
                            lst[0] = h0;
                            inH = true;
                        }
                        const u32 e = sw[q] & 63u;
                        if (e == 0u) continue;
                        // entry (ee, j) = lst[j] * p^ee at ee * cur + j: each entry of the
                        // list so far is read once and its e multiples are written, so
                        // the only read-after-write dependence is between stages
                        const u32 cur = n;
                        const double dp = (double)(sw[q] >> 6);
                        u32 j = 0;
                        for (; j + 4 <= cur; j += 4) {   // four loads in flight
                            double y0 = lst[j], y1 = lst[j + 1], y2 = lst[j + 2], y3 = lst[j + 3];
                            for (u32 ee = 1; ee <= e; ++ee) {
                                y0 *= dp; y1 *= dp; y2 *= dp; y3 *= dp;
                                double *o = lst + ee * cur + j;
                                o[0] = y0; o[1] = y1; o[2] = y2; o[3] = y3;
                            }
                        }
                        if (j < cur) {   // 1 to 3 entries left: one group, its loads issued together
                            // (the two reads past cur stay inside the arena's 3 spare
                            // doubles; the products of entries past cur go to the
                            // spare slot s_ar[EX_ARENA + 2], never to a list)
                            const bool w1 = j + 1u < cur, w2 = j + 2u < cur;
                            double y0 = lst[j], y1 = lst[j + 1], y2 = lst[j + 2];
                            double *spare = s_ar + EX_ARENA + 2;
                            for (u32 ee = 1; ee <= e; ++ee) {
                                y0 *= dp; y1 *= dp; y2 *= dp;
                                double *o = lst + ee * cur + j;
                                o[0] = y0;
                                *(w1 ? o + 1 : spare) = y1;
                                *(w2 ? o + 2 : spare) = y2;
                            }
                        }
                        n = cur * (e + 1);
                    }
                    if (!inH) s_ar[loff + nL] = h0;   // empty H part: the list {1}
                }
                __syncthreads();
            __syncthreads();
            // the entries the lanes free after this pass set up next: loaded now,
            // used after the candidate loop
            pqi = take(__ballot(!has || fits || drop));
            if (pqi != ~0ull) qpf = a.quads[pqi];
                // lane l takes candidates l, l+64, l+128, ... of the pass, two per
                // iteration (g and g + 64: two independent load/compute chains).
                // Pipelined: while this iteration's list entries load, the first
                // record words of the next iteration and the bitmap words of the
                // one after are read, and the next iteration's mask words once this
                // iteration's tests have used theirs (LDS reads complete in order,
                // so each wait covers only older reads).
                u32 kp = 0;     // marks before the next fetched base
                u32 qn1 = 0;    // entries in the first queue
                // the records of candidates base + tid and base + 64 + tid from the
                // bitmap words of base: every 4-set but the first marks the candidate
                // before its first, and so does the guard record, so the record of
                // candidate g is the number of marks below g (a ballot-style prefix
                // count, v_mbcnt), the guard for g >= ctot; the fetch past the last
                // iteration reads a cleared word, so every record index is at most
                // the guard's (the number of fitting 4-sets <= EX_NT)
                u32 ka, kb, ia, ib;       // ia, ib: byte offsets of the records read (of ka, kb or 0)
                uint4 A0, A1, A2, B0, B1, B2;   // their words
                auto fetch = [&](u32 base, const uint4 &wv) {
                    // (prefix and population counts with kp as their addend)
                    ka = __builtin_amdgcn_mbcnt_hi(wv.y, __builtin_amdgcn_mbcnt_lo(wv.x, kp));
                    kp = PS4_BCNT_ADD(wv.y, PS4_BCNT_ADD(wv.x, kp));
                    kb = __builtin_amdgcn_mbcnt_hi(wv.w, __builtin_amdgcn_mbcnt_lo(wv.z, kp));
                    kp = PS4_BCNT_ADD(wv.w, PS4_BCNT_ADD(wv.z, kp));
                    // byte offsets of the records (record indices are at most 64, the guard's)
                    ia = PS4_MAD_U24(ka, EX_PAR_WORDS * 16u, 0u);
                    ib = PS4_MAD_U24(kb, EX_PAR_WORDS * 16u, 0u);
                    A0 = *(const uint4 *)((const char *)s_par + ia);
                    B0 = *(const uint4 *)((const char *)s_par + ib);
                };
                auto fetch_masks = [&]() {
                    A1 = *(const uint4 *)((const char *)s_par + ia + 16u);
                    A2 = *(const uint4 *)((const char *)s_par + ia + 32u);
                    B1 = *(const uint4 *)((const char *)s_par + ib + 16u);
                    B2 = *(const uint4 *)((const char *)s_par + ib + 32u);
                };
                // (read unconditionally: beyond the pass the bitmap holds stale bits,
                // which only choose records of candidates that are discarded)
                fetch(0u, s_bits4[0]);
                fetch_masks();
                uint4 wnext = s_bits4[1];
                for (u32 base = 0; base < ctot; base += 128) {
                    const u32 ga8 = 8u * (base + tid), gb8 = ga8 + 512u;
                    const u32 ka0 = ka, kb0 = kb;
                    const ExLists LA = ex_lists(A0, s_ar, ga8), LB = ex_lists(B0, s_ar, gb8);
                    __builtin_amdgcn_sched_barrier(0);   // both candidates' list reads issued before any use
                    fetch(base + 128u, wnext);
                    wnext = s_bits4[min((base + 256u) >> 7, EX_CMAX / 128u)];
                    PS4_KEEP("v"(LA.l0), "v"(LA.h0), "v"(LA.l1), "v"(LA.h1), "v"(LB.l0), "v"(LB.h0), "v"(LB.l1),
                             "v"(LB.h1));   // both candidates' list entries (the next reads may still be in flight)
                    const ExCand A = ex_cand(A1, A2, LA, a.neg_m1);
                    const ExCand B = ex_cand(B1, B2, LB, a.neg_m1);
                    fetch_masks();
                    const bool oka = A.ok, okb = B.ok;   // (the guard record's candidates never pass)
                    // candidates past the residue filters wait on an LDS stack (their
                    // order does not matter) for the rest of the ext_sqrt tests and the
                    // later filters, which run 64 at a time with every lane busy; the
                    // survivors go to the exact-test queue.  The workgroup is one wave
                    // and its LDS operations complete in order, so a wave barrier (no
                    // instruction: it keeps the compiler from moving memory operations
                    // across it) orders the stack's writes before a round's reads.
                    auto round1 = [&]() {   // the top min(64, qn1) entries
                        __builtin_amdgcn_wave_barrier();
                        const u32 m = qn1 < 64u ? qn1 : 64u, b1 = qn1 - m;
                        const bool on = tid < m;
                        // the entry, then every read of its record at once (one wait
                        // for all of them: the round is bound by these latencies);
                        // lanes without an entry read record 0 and count nothing
                        const u64 ent = s_q1[on ? b1 + tid : 0u];
                        const u32 k = on ? ex_qk(ent) : 0u;
                        const double u = ex_qu(ent);
                        const u32 aux = s_aux[k];
                        const uint4 mem = s_mem[k];
                        const u64 q2h = *(const u64 *)((const char *)s_par + k * (EX_PAR_WORDS * 16u) + 40u);   // msm
                        const u64 qi = s_qix[k];
                        const bool counted = on && ex_counted(u, mem, aux);
                        my_unc += (on && !counted) ? 1u : 0u;
                        const bool go = counted && ex_late(u, make_uint4(0u, 0u, (u32)q2h, (u32)(q2h >> 32)));
                        const u64 bg = __ballot(go);
                        if (go) {
                            // record | pair choice << 6
                            exq2[qn2 + lanes_below(bg)] =
                                make_uint4((u32)ent, ((u32)(ent >> 32) & 0xFFFFFu) | ((u32)(q2h >> 61) << 20), (u32)qi,
                                           (u32)(qi >> 32));
                        }
                        qn2 += (u32)__popcll(bg);
                        qn1 = b1;
                        if (qn2 >= 64u) drain2();
                    };
                    // (the stack holds < 64 entries before each push of at most 64)
                    const u64 boa = __ballot(oka);
                    if (oka) *(u64 *)((char *)s_q1 + PS4_LSHL3_ADD(lanes_below(boa), 8u * qn1)) = ex_qent(A.us, ka0);
                    qn1 += (u32)__popcll(boa);
                    w_esq += (u32)__popcll(boa);
                    if (qn1 >= 64u) round1();
                    const u64 bob = __ballot(okb);
                    if (okb) *(u64 *)((char *)s_q1 + PS4_LSHL3_ADD(lanes_below(bob), 8u * qn1)) = ex_qent(B.us, kb0);
                    qn1 += (u32)__popcll(bob);
                    w_esq += (u32)__popcll(bob);
                    while (qn1 >= 64u || (base + 128u >= ctot && qn1 > 0u))   // (the pass's records
                        round1();                                              //  are about to be replaced)
                }
            __syncthreads();
            if (fits || drop) has = false;
            acc_esq += tid == 0 ? w_esq : 0u;   // (w_esq < 2^14: at most the pass's 8192 candidates)
            acc_esq -= my_unc;   // lane sums may wrap; the wave's sum is exact
        }
    }
    if (qn2 > 0u) drain2();
    ex_flush(a.ctr, a.n_rids, acc_rid, acc_cand, acc_esq, acc_md);
    if (err) atomicOr(&a.ctl->err, err);
}

// ----------------------------------------------------------------------------
// launch wrappers (called by the host library)
// ----------------------------------------------------------------------------

// The compiled launch shapes of k_rq: threads per workgroup x window points a
// thread loads at once (the autotuner's block and tile).
#define RQ_SHAPES(X) X(128, 2) X(128, 4) X(128, 8) X(256, 2) X(256, 4) X(256, 8) X(512, 2) X(512, 4) X(512, 8)

extern "C" hipError_t ps4_launch_rq(hipStream_t st, u32 grid, u32 block, u32 tile, const ps4_item *items, u32 n_items,
                                    u32 bin_cap, u32 nb_alloc, ps4_ctl *ctl, u64 *ctr, u32 n_rids, u64 *scratch,
                                    ps4_quad *quads, u64 quad_cap, u32 cap_reps, u32 verify)
{
    RqArgs a;
    a.n_rids = n_rids;
    a.items = items; a.n_items = n_items; a.bin_cap = bin_cap; a.nb_alloc = nb_alloc; a.ctl = ctl; a.ctr = ctr;
    a.scratch = scratch;
    a.quads = quads; a.quad_cap = quad_cap; a.cap_reps = cap_reps; a.verify = verify;
#define RQ_LAUNCH(nt, un)                                                                 \
    if (block == nt && tile == un) {                                                      \
        hipLaunchKernelGGL((k_rq<nt, un>), dim3(grid), dim3(nt), 0, st, a);               \
        return hipGetLastError();                                                         \
    }
    RQ_SHAPES(RQ_LAUNCH)
#undef RQ_LAUNCH
    return hipErrorInvalidValue;
}

extern "C" hipError_t ps4_launch_ext(hipStream_t st, u32 grid, const ps4_quad *quads, u64 quad_cap, ps4_ctl *ctl,
                                     const uint32_t *oftab, u64 tab_n, const uint16_t *ndiv, const ps4_qrtab *qr,
                                     u64 *ctr, u32 n_rids, ps4_hit *hits, u32 hit_cap, u32 cap_divs, u64 *exq,
                                     int chunk)
{
    ExArgs a;
    a.n_rids = n_rids;
    a.quads = quads; a.quad_cap = quad_cap; a.ctl = ctl; a.oftab = (const ExOdd *)oftab; a.tab_n = tab_n; a.ndiv = ndiv; a.qr = qr;
    a.ctr = ctr;
    a.hits = hits; a.hit_cap = hit_cap; a.cap_divs = cap_divs; a.exq = (uint4 *)exq;
    a.neg_m1 = -M1;
    switch (chunk) {
    case 1: hipLaunchKernelGGL(k_ext<1>, dim3(grid), dim3(EX_NT), 0, st, a); break;
    case 4: hipLaunchKernelGGL(k_ext<4>, dim3(grid), dim3(EX_NT), 0, st, a); break;
    case 16: hipLaunchKernelGGL(k_ext<16>, dim3(grid), dim3(EX_NT), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Reset of the control block and the counters, as a kernel (no memset nodes,
// so the launch sequence stays capturable).
__global__ void k_reset(ps4_ctl *ctl, u64 *ctr, u32 n_ctr)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctl->item_next = 0; ctl->tile_next = 0; ctl->err = 0; ctl->hit_count = 0;
        ctl->quad_count = 0; ctl->pad = 0;
    }
    if (i < n_ctr) ctr[i] = 0;
}

extern "C" hipError_t ps4_launch_reset(hipStream_t st, ps4_ctl *ctl, u64 *ctr, u32 n_ctr)
{
    const u32 n = n_ctr > 0 ? n_ctr : 1;
    hipLaunchKernelGGL(k_reset, dim3((n + 255) / 256), dim3(256), 0, st, ctl, ctr, n_ctr);
    return hipGetLastError();
}

extern "C" int ps4_rq_threads(void) { return RQ_NT_DEFAULT; }
extern "C" int ps4_rq_tile(void) { return RQ_UN_DEFAULT; }
extern "C" int ps4_ext_chunk(void) { return EX_CHUNK_DEFAULT; }
extern "C" int ps4_ext_threads(void) { return EX_NT; }
extern "C" int ps4_ext_queue_cap(void) { return 2 * EX_Q2; }   // u64 words per workgroup
extern "C" int ps4_rq_window(void) { return (int)RQ_W; }
extern "C" int ps4_rq_nb_max(void) { return RQ_NB_MAX; }

// The kernel of a launch shape (for occupancy queries), or NULL.
extern "C" void *ps4_k_rq_ptr(u32 block, u32 tile)
{
#define RQ_PTR(nt, un) \
    if (block == nt && tile == un) return (void *)&k_rq<nt, un>;
    RQ_SHAPES(RQ_PTR)
#undef RQ_PTR
    return nullptr;
}
extern "C" void *ps4_k_ext_ptr(int chunk)
{
    return chunk == 1 ? (void *)&k_ext<1> : chunk == 4 ? (void *)&k_ext<4> : chunk == 16 ? (void *)&k_ext<16> : nullptr;
}

// ----------------------------------------------------------------------------
// k_triple: d with d+a, d+b, d+c square for given square triples {a,b,c}
// ----------------------------------------------------------------------------
// One workgroup per query.  With a > b > c, d + a = u^2 and d + b = w^2 give
// (u-w)(u+w) = a - b = (r_ac - r_bc)(r_ac + r_bc): the divisor step of
// extend4 (exhaustive/s4.c:172-190) with one remaining member to test.

#define TR_NT 256
#define TR_LIST 1024

__global__ __launch_bounds__(TR_NT) void k_triple(const u64 *abc, u64 d_max, const u64 *ftab, u64 tab_n,
                                                  u64 *d_out, u32 cap, u32 *count, u32 *err_out)
{
    __shared__ double s_L[TR_LIST];
    __shared__ u32 s_n[4];     // nL, nH, half, ok
    __shared__ u64 s_v[3];
    const u32 tid = threadIdx.x;
    const u64 qi = blockIdx.x;
    if (tid == 0) {
        u64 x = abc[3 * qi], y = abc[3 * qi + 1], z = abc[3 * qi + 2], t;
        if (x < y) { t = x; x = y; y = t; }
        if (y < z) { t = y; y = z; z = t; }
        if (x < y) { t = x; x = y; y = t; }
        s_v[0] = x; s_v[1] = y; s_v[2] = z;
        const u64 rac = d_isqrt64(x + z), rbc = d_isqrt64(y + z);
        u64 m1 = rac - rbc, m2 = rac + rbc;
        u32 ok = (rac * rac == x + z && rbc * rbc == y + z && m2 < tab_n && x > y && y > z && z > 0) ? 1u : 0u;
        u32 nL = 0, nH = 0, half = 0;
        if (ok) {
            const u32 z1 = (u32)__builtin_ctzll(m1), z2 = (u32)__builtin_ctzll(m2), e2 = z1 + z2;
            m1 >>= z1;
            m2 >>= z2;
            u32 pr[32], ex[32], nf = 0;
            if (e2 >= 2) { pr[0] = 2; ex[0] = e2 - 2; nf = (e2 > 2) ? 1u : 0u; }
            for (int pass = 0; pass < 2 && ok; ++pass) {
                u64 m = pass ? m2 : m1;
                for (int guard = 0; m > 1; ++guard) {
                    if (m >= tab_n || guard >= 64) { ok = 0; break; }
                    const u64 en = ftab[m];
                    const u32 p = (u32)(en >> 32) & 0x7FFFFFFu, e = (u32)(en >> 59);
                    m = en & 0xFFFFFFFFull;
                    u32 i = 0;
                    while (i < nf && pr[i] != p) ++i;
                    if (i < nf) ex[i] += e;
                    else if (nf < 32) { pr[nf] = p; ex[nf] = e; ++nf; }
                    else ok = 0;
                }
            }
            if (e2 == 1) ok = 0;   // a - b = 2 mod 4: no divisor pair of equal parity
            half = (e2 == 0);
            if (ok) {
                u64 nd = 1;
                for (u32 i = 0; i < nf; ++i) nd *= ex[i] + 1;
                // split into L (prefix) and H with nL + nH small
                u64 nl = 1, bestc = 1 + nd, bestnl = 1;
                u32 sL = 0;
                for (u32 s = 1; s <= nf; ++s) {
                    nl *= ex[s - 1] + 1;
                    const u64 c = nl + nd / nl;
                    if (c < bestc) { bestc = c; sL = s; bestnl = nl; }
                }
                nL = (u32)bestnl;
                nH = (u32)(nd / bestnl);
                if (nL + nH > TR_LIST) { ok = 0; atomicOr(err_out, ERR_ARENA); }
                if (ok) {
                    double *lst = s_L;
                    u32 n = 1;
                    lst[0] = 1.0;
                    for (u32 s = 0; s < nf; ++s) {
                        if (s == sL) { lst = s_L + nL; n = 1; lst[0] = 1.0; }
                        const u32 cur = n;
                        double pk = 1.0;
                        for (u32 ee = 0; ee < ex[s]; ++ee) {
                            pk *= (double)pr[s];
                            for (u32 j = 0; j < cur; ++j) lst[n++] = lst[j] * pk;
                        }
                    }
                    if (sL == nf) s_L[nL] = 1.0;
                }
            }
        }
        s_n[0] = nL; s_n[1] = nH; s_n[2] = half; s_n[3] = ok;
    }
    __syncthreads();
    if (!s_n[3]) return;
    const u32 nL = s_n[0], nH = s_n[1];
    const bool half = s_n[2] != 0;
    const u64 A = s_v[0], B = s_v[1], C = s_v[2];
    const u32 total = nL * nH;
    for (u32 t = tid; t < total; t += TR_NT) {
        const u32 j = t / nL, i = t - j * nL;
        const double f = s_L[i] * s_L[nL + j], g = s_L[nL - 1 - i] * s_L[nL + nH - 1 - j];
        if (!(f < g)) continue;
        const u64 u = (u64)(half ? (f + g) * 0.5 : f + g);
        const u128 uu = (u128)u * u;
        if (uu <= A) continue;
        const u128 d = uu - A;
        if (d > d_max || d == A || d == B || d == C) continue;
        if (!d_is_square128(d + C)) continue;
        const u32 slot = atomicAdd(&count[qi], 1u);
        if (slot < cap) d_out[qi * cap + slot] = (u64)d;
    }
}

extern "C" hipError_t ps4_launch_triple(hipStream_t st, const u64 *abc, u64 n, u64 d_max, const u64 *ftab,
                                        u64 tab_n, u64 *d_out, u32 cap, u32 *count, u32 *err)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_triple, dim3((u32)n), dim3(TR_NT), 0, st, abc, d_max, ftab, tab_n, d_out, cap, count, err);
    return hipGetLastError();
}

// Fixed device scratch of the allocation-free triple query
// (ps4_extend_triple_noalloc): words 0..2 the query's members, word 3 its
// answer count (low half) and error word (high half), words 8.. the staging of
// its answers (PS4_NA_CAP of them).  Each device holds its own copy until the
// process ends; host code reaches it only through the address
// ps4_noalloc_scratch() returns (the array's name is not a device address in
// host code).  Calls that overlap on one device share it and race.
#define PS4_NA_CAP 65536u
#define PS4_NA_WORDS (8u + PS4_NA_CAP)
static __device__ __align__(256) u64 scratch[PS4_NA_WORDS];

// The device address of the scratch on `device`, looked up once per device
// (the first call for a device must have it current: the library's ps4_open
// makes this call right after hipSetDevice) and kept for later calls.
#define PS4_NA_DEVICES 64
static void *g_na_ptr[PS4_NA_DEVICES];
static std::mutex g_na_mu;
extern "C" hipError_t ps4_noalloc_scratch(int device, void **ptr)
{
    if (device < 0 || device >= PS4_NA_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_na_ptr[device] = p;
    }
    *ptr = g_na_ptr[device];
    return hipSuccess;
}

extern "C" unsigned ps4_noalloc_cap(void) { return PS4_NA_CAP; }

// Fill n 32-bit words with a value (used instead of memset so launch
// sequences stay free of memset nodes).
// ----------------------------------------------------------------------------
// element-range mode: the 4-sets of a census whose largest member is <= n_max,
// and the square triples {a < b < c} with c <= n_max
// ----------------------------------------------------------------------------

// Copy the queued 4-sets with largest member a = (r_ab^2 + r_ac^2 - r_bc^2) / 2
// <= n_max to out (grid-stride; slots by atomic, *count may exceed cap, which
// the host reports).
__global__ void k_elem(const ps4_quad *quads, const ps4_ctl *ctl, u64 quad_cap, u64 n_max, ps4_quad *out,
                       u32 cap, u32 *count)
{
    u64 nq = ctl->quad_count;
    if (nq > quad_cap) nq = quad_cap;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += (u64)gridDim.x * blockDim.x) {
        const ps4_quad q = quads[i];
        if (q.rid == PS4_RID_NONE) continue;   // padding of a window's block
        const u64 a = ((u64)q.r_ab * q.r_ab + (u64)q.r_ac * q.r_ac - (u64)q.r_bc * q.r_bc) >> 1;
        if (a > n_max) continue;
        const u32 slot = atomicAdd(count, 1u);
        if (slot < cap) out[slot] = q;
    }
}

// Square triples {a < b < c}: a + b = x^2, a + c = y^2, b + c = z^2 with
// x < y < z, x + y + z even, a = (x^2 + y^2 - z^2) / 2 >= 1 and
// c = (y^2 + z^2 - x^2) / 2 <= n_max.  For one x and each y the admissible z
// form the range y < z <= min(isqrt(x^2 + y^2 - 2), isqrt(2 n_max + x^2 - y^2))
// of one parity, counted in closed form; y stops where the range is empty for
// good (both bounds fall with y).  One thread per x (grid-stride), a sum per
// wave added to *total.
__global__ void k_sqtri(u64 n_max, u64 x_max, unsigned long long *total)
{
    u64 acc = 0;
    for (u64 x = 1 + (u64)blockIdx.x * blockDim.x + threadIdx.x; x <= x_max; x += (u64)gridDim.x * blockDim.x) {
        const u64 xx = x * x;
        for (u64 y = x + 1;; ++y) {
            const u64 yy = y * y;
            if (2 * y + 3 > xx) break;                 // a >= 1 needs (y + 1)^2 <= x^2 + y^2 - 2
            if (yy + (y + 1) * (y + 1) > 2 * n_max + xx) break;   // c <= n_max needs z = y + 1 to fit
            const u64 z1 = d_isqrt64(xx + yy - 2), z2 = d_isqrt64(2 * n_max + xx - yy);
            const u64 zmax = z1 < z2 ? z1 : z2;
            const u64 zl = y + 1 + ((y + 1 + x + y) & 1u);   // z = x + y (mod 2)
            if (zmax >= zl) acc += (zmax - zl) / 2 + 1;
        }
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0 && acc) atomicAdd(total, (unsigned long long)acc);
}

extern "C" hipError_t ps4_launch_elem(hipStream_t st, const ps4_quad *quads, const ps4_ctl *ctl, u64 quad_cap,
                                      u64 n_max, ps4_quad *out, u32 cap, u32 *count, u32 grid)
{
    hipLaunchKernelGGL(k_elem, dim3(grid), dim3(256), 0, st, quads, ctl, quad_cap, n_max, out, cap, count);
    return hipGetLastError();
}

extern "C" hipError_t ps4_launch_sqtri(hipStream_t st, u64 n_max, unsigned long long *total)
{
    // x < y and c <= n_max give x^2 + 6x + 5 <= 2 n_max: x <= isqrt(2 n_max)
    u64 x_max = (u64)sqrt((double)(2 * n_max));
    while (x_max * x_max > 2 * n_max) --x_max;
    while ((x_max + 1) * (x_max + 1) <= 2 * n_max) ++x_max;
    const u32 grid = (u32)((x_max + 255) / 256);
    hipLaunchKernelGGL(k_sqtri, dim3(grid), dim3(256), 0, st, n_max, x_max, total);
    return hipGetLastError();
}

__global__ void k_fill_u32(u32 *p, u32 n, u32 v)
{
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

extern "C" hipError_t ps4_launch_fill_u32(hipStream_t st, u32 *p, u32 n, u32 v)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}
