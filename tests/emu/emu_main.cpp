// emu_main.cpp -- runs the census library with its kernels emulated on the CPU
// (hip_host_shim.h) and compares every result with the CPU restatement.
// Test infrastructure only.
//
//   ps4_emu [--noext] [--dump4 N] [--smax S] [--batch] lo hi [lo hi ...]
//   (--batch: all ranges in one ps4_search_batch call, so tiles mix intervals)
//   ps4_emu --triple a b c dmax
//   ps4_emu --elements n_max
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cpu_ref.h"
#include "ps4.h"

static int sink(void *user, const char *lines, size_t n)
{
    ((std::string *)user)->append(lines, n);
    return 0;
}
static void ora_line(void *user, const char *line) { ((std::string *)user)->append(line); }

static std::vector<std::string> split_sorted(const std::string &s)
{
    std::vector<std::string> v;
    size_t i = 0;
    while (i < s.size()) {
        size_t j = s.find('\n', i);
        if (j == std::string::npos) j = s.size();
        v.push_back(s.substr(i, j - i));
        i = j + 1;
    }
    std::sort(v.begin(), v.end());
    return v;
}

int main(int argc, char **argv)
{
    unsigned flags = 0;
    long long dump4 = 0;
    unsigned long long smax = 0;
    std::vector<unsigned long long> nums;
    bool triple = false, batch = false, elements = false, stream = false;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--noext")) flags |= PS4_NOEXT;
        else if (!strcmp(argv[i], "--dump4")) dump4 = atoll(argv[++i]);
        else if (!strcmp(argv[i], "--smax")) smax = strtoull(argv[++i], 0, 10);
        else if (!strcmp(argv[i], "--triple")) triple = true;
        else if (!strcmp(argv[i], "--batch")) batch = true;
        else if (!strcmp(argv[i], "--elements")) elements = true;
        else if (!strcmp(argv[i], "--stream")) stream = true;
        else nums.push_back(strtoull(argv[i], 0, 10));
    }
    ps4_config cfg;
    memset(&cfg, 0, sizeof cfg);
    unsigned long long top = 0;
    for (auto x : nums) top = std::max(top, x);
    cfg.S_hi_max = smax ? smax : std::max((elements ? 4 : 1) * top + 1, 100ull);
    cfg.flags = flags;
    cfg.dump4_limit = dump4;
    ps4_ctx *ctx = nullptr;
    int rc = ps4_open(&cfg, &ctx);
    if (rc) { printf("open failed %d\n", rc); return 2; }
    if (const char *tune = getenv("PS4_EMU_TUNE")) {   // a tuning file's text, e.g. "rq_block = 512"
        ps4_tuning t;
        if (ps4_parse_tuning(tune, &t) || ps4_set_tuning(ctx, &t)) { printf("bad tuning\n"); return 2; }
    }
    int bad = 0;
    if (elements) {
        std::string got, want;
        uint64_t ns = 0, nt = 0;
        rc = ps4_search_elements(ctx, nums[0], sink, &got, &ns, &nt);
        const int64_t m = ora_elements(nums[0], ora_line, &want);
        const uint64_t mt = ora_square_triples(nums[0]);
        const bool ok = rc == 0 && got == want && (int64_t)ns == m && nt == mt;
        printf("elements %llu %s sets=%llu want=%lld triples=%llu want=%llu rc=%d\n", (unsigned long long)nums[0],
               ok ? "OK" : "MISMATCH", (unsigned long long)ns, (long long)m, (unsigned long long)nt,
               (unsigned long long)mt, rc);
        ps4_close(ctx);
        return ok ? 0 : 1;
    }
    if (triple) {
        uint64_t abc[3] = {nums[0], nums[1], nums[2]};
        uint64_t out[4096], want[4096];
        size_t n = 0;
        rc = ps4_extend_triple(ctx, abc, nums[3], out, 4096, &n);
        int64_t m = ora_extend_triple(abc, nums[3], want, 4096);
        bool ok = rc == 0 && (int64_t)n == m && std::equal(out, out + n, want);
        printf("triple %s n=%zu want=%lld rc=%d\n", ok ? "OK" : "MISMATCH", n, (long long)m, rc);
        ps4_close(ctx);
        return ok ? 0 : 1;
    }
    if (stream) {   // every range through the serving stream, small steps, plus one triple query
        struct Got { ps4_counters c; std::string text; int rc; bool seen; std::vector<uint64_t> d; };
        std::vector<Got> got(nums.size() / 2 + 1);
        for (size_t r = 0; r + 1 < nums.size(); r += 2) ps4_submit_range(ctx, r / 2, nums[r], nums[r + 1]);
        const uint64_t abc[3] = {482, 359, 2};
        ps4_submit_triple(ctx, nums.size() / 2, abc, 1000000);
        auto cb = [](void *u, const ps4_done *d) {
            Got &g = (*(std::vector<Got> *)u)[d->tag];
            g.c = d->counters;
            g.text.assign(d->lines ? d->lines : "", d->nbytes);
            g.rc = d->rc;
            g.seen = true;
            g.d.assign(d->d, d->d + d->n_d);
        };
        ps4_stream_state stt;
        int steps = 0;
        do {
            rc = ps4_step(ctx, 20000, cb, &got, &stt);
            if (rc) { printf("step rc=%d %s\n", rc, ps4_last_error(ctx)); return 1; }
            ++steps;
        } while (stt.pending_ranges + stt.pending_triples > 0 && steps < 100000);
        int bad_s = 0;
        for (size_t r = 0; r + 1 < nums.size(); r += 2) {
            Got &g = got[r / 2];
            std::string want;
            ora_counters w;
            ora_census(nums[r], nums[r + 1], 0u, 0, ora_line, &want, &w);
            const bool ok = g.seen && g.rc == 0 && g.c.sets4 == w.sets4 && g.c.ext_cand == w.ext_cand &&
                            g.c.ext_sqrt == w.ext_sqrt && g.c.sets5 == w.sets5 && g.c.pairs == w.pairs &&
                            g.c.max_divisors == w.max_divisors && split_sorted(g.text) == split_sorted(want);
            printf("stream [%llu,%llu) %s 4-sets=%llu\n", (unsigned long long)nums[r], (unsigned long long)nums[r + 1],
                   ok ? "OK" : "MISMATCH", (unsigned long long)g.c.sets4);
            bad_s |= !ok;
        }
        uint64_t want_d[64];
        const int64_t m = ora_extend_triple(abc, 1000000, want_d, 64);
        const Got &gt = got.back();
        const bool tok = gt.seen && gt.rc == 0 && (int64_t)gt.d.size() == m && std::equal(gt.d.begin(), gt.d.end(), want_d);
        printf("stream triple %s n=%zu steps=%d\n", tok ? "OK" : "MISMATCH", gt.d.size(), steps);
        ps4_close(ctx);
        return (bad_s || !tok) ? 1 : 0;
    }
    const size_t nr = nums.size() / 2;
    std::vector<std::string> btext(nr);
    std::vector<ps4_counters> bc(nr);
    if (batch) {
        std::vector<uint64_t> los(nr), his(nr);
        std::vector<void *> users(nr);
        for (size_t r = 0; r < nr; ++r) { los[r] = nums[2 * r]; his[r] = nums[2 * r + 1]; users[r] = &btext[r]; }
        rc = ps4_search_batch(ctx, nr, los.data(), his.data(), sink, users.data(), bc.data());
        if (rc) { printf("batch search rc=%d %s\n", rc, ps4_last_error(ctx)); ps4_close(ctx); return 1; }
    }
    for (size_t i = 0; i + 1 < nums.size(); i += 2) {
        const unsigned long long lo = nums[i], hi = nums[i + 1];
        std::string got_text, want_text;
        ps4_counters c;
        if (batch) {
            got_text = btext[i / 2];
            c = bc[i / 2];
        } else {
            rc = ps4_search(ctx, lo, hi, sink, &got_text, &c);
            if (rc) { printf("[%llu,%llu) search rc=%d %s\n", lo, hi, rc, ps4_last_error(ctx)); bad = 1; continue; }
        }
        ora_counters w;
        int orc = ora_census(lo, hi, (flags & PS4_NOEXT) ? ORA_NOEXT : 0u, dump4, ora_line, &want_text, &w);
        if (orc) { printf("oracle rc=%d\n", orc); bad = 1; continue; }
        const uint64_t g[] = {c.S_done, c.pairs, c.S_ge3, c.triples, c.sets4, c.sets5, c.sets6, c.multi, c.ext_cand,
                              c.ext_sqrt, c.max_reps, c.max_divisors, c.max_extensions};
        const uint64_t e[] = {w.S_done, w.pairs, w.S_ge3, w.triples, w.sets4, w.sets5, w.sets6, w.multi, w.ext_cand,
                              w.ext_sqrt, w.max_reps, w.max_divisors, w.max_extensions};
        const char *names[] = {"S_done", "pairs", "S_ge3", "triples", "4-sets", "5-sets", "6-sets", "multi",
                               "ext_cand", "ext_sqrt", "max_reps", "max_divisors", "max_extensions"};
        bool ok = true;
        for (int k = 0; k < 13; ++k)
            if (g[k] != e[k]) { printf("  %s got %llu want %llu\n", names[k], (unsigned long long)g[k], (unsigned long long)e[k]); ok = false; }
        if (memcmp(c.hist5, w.hist5, sizeof c.hist5)) { printf("  hist5 differs\n"); ok = false; }
        auto gl = split_sorted(got_text), wl = split_sorted(want_text);
        if (gl != wl) {
            printf("  lines differ: got %zu want %zu\n", gl.size(), wl.size());
            std::vector<std::string> d;
            std::set_symmetric_difference(gl.begin(), gl.end(), wl.begin(), wl.end(), std::back_inserter(d));
            for (size_t k = 0; k < d.size() && k < 6; ++k) printf("    %s\n", d[k].c_str());
            ok = false;
        }
        printf("[%llu,%llu) %s 4-sets=%llu ext_cand=%llu lines=%zu\n", lo, hi, ok ? "OK" : "MISMATCH",
               (unsigned long long)c.sets4, (unsigned long long)c.ext_cand, gl.size());
        fflush(stdout);
        if (!ok) bad = 1;
    }
    ps4_close(ctx);
    return bad;
}
