// The persisted tuning table: per transform size, the single-pass kernel variant or the
// multi-pass split and columns per workgroup that measured fastest on this GPU.
//
// File: tab-separated lines "log2n  kind  params  gb_s  batch", '#' comments, where kind is
// "single" (params "nt=0|1,layout=0..2[,tw64=0..3[,xlane=0|1]]"), "multi" (params
// "split=a,b[,c[,d]];cols=x,y[,..][;layout=0|2]", log2 N_1 first; layout: the passes' LDS layout,
// padded when absent) or "ragged" (params "batched_min=b;split=s;tile_run=t[;order=o][;grid=g]": ragged
// plans' choices, one line, log2n the largest item size of the workload it was tuned on, batch its
// item count; order 0 groups first, 1 persistent kernel first, 2 side by side, 3 the groups split
// over two streams, then the persistent kernel, 4 as 3 with the persistent kernel on the stream
// with fewer bytes, 5 as 3 with the persistent kernel first on the plan's stream; grid the
// persistent kernel's workgroups, 0 or absent for as many as are resident at once). The batch the line was measured at sets its working-set regime
// (runtime.hpp: tune_regime); a size may have a line per regime. Written by `python -m onefft_amd.tune`; read once, at the first plan creation,
// from $ONEFFT_TUNE_FILE or tune_gfx950.tsv next to the library. Plan creation stays O(n) with no
// measuring, as with the reference's planner (its tuning is compiled in, src/onefft.c:652-697).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "runtime.hpp"

namespace onefft {
namespace {

TuneEntry g_table[kTuneMaxLog2 + 1][2];
RaggedTune g_ragged;
std::once_flag g_once;

void load_table();

std::string table_path() {
    if (const char *env = std::getenv("ONEFFT_TUNE_FILE")) return env;
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&load_table), &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t slash = p.rfind('/');
        return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/tune_gfx950.tsv";
    }
    return "tune_gfx950.tsv";
}

// "a,b,c" -> up to `max` ints; returns the count, 0 on a malformed list
int parse_list(const char *s, int *out, int max) {
    int count = 0;
    while (*s && count < max) {
        char *end;
        const long v = std::strtol(s, &end, 10);
        if (end == s) return 0;
        out[count++] = (int)v;
        if (*end != ',') break;
        s = end + 1;
    }
    return count;
}

void parse_line(const char *line) {
    int log2n = -1;
    double gbs = 0;
    unsigned long long batch = 0;
    char kind[16] = {0}, params[128] = {0};
    if (std::sscanf(line, "%d %15s %127s %lf %llu", &log2n, kind, params, &gbs, &batch) != 5) return;
    if (log2n < 0 || log2n > kTuneMaxLog2 || batch == 0) return;
    if (!std::strcmp(kind, "ragged")) {
        unsigned b = 0, s = 0, t = 0, o = 0;
        if (std::sscanf(params, "batched_min=%u;split=%u;tile_run=%u;order=%u", &b, &s, &t, &o) < 3) return;
        if (b <= (unsigned)kMaxSinglePassLog2 || b > (unsigned)kRaggedMaxLog2 + 1) return;
        if (s <= (unsigned)kRaggedSingleMaxLog2 || s > (unsigned)kRaggedMaxLog2 + 1) return;
        if (t < 1 || t > 256 || o >= kRaggedOrders) return;
        long g = 0;
        if (const char *gp = std::strstr(params, "grid=")) g = std::atol(gp + 5);
        if (g < 0 || g > (1L << 20)) return;
        g_ragged.set = true;
        g_ragged.batched_min = b;
        g_ragged.split = s;
        g_ragged.tile_run = t;
        g_ragged.order = o;
        g_ragged.grid = (uint32_t)g;
        return;
    }
    const int regime = tune_regime(uint64_t(1) << log2n, batch);
    TuneEntry e;
    if (!std::strcmp(kind, "single")) {
        int nt = 1, layout = 0, tw64 = 0, xlane = 0;  // tables written before FP64 bases existed meant tw64=0
        if (std::sscanf(params, "nt=%d,layout=%d,tw64=%d,xlane=%d", &nt, &layout, &tw64, &xlane) < 2) return;
        e.kind = TuneEntry::kSingle;
        e.variant.nt = nt != 0;
        e.variant.layout = layout < 0 ? 0 : layout > 2 ? 2 : layout;
        e.variant.tw64 = tw64 < 0 ? 0 : tw64 > 3 ? 3 : tw64;
        e.variant.xlane = xlane != 0;
    } else if (!std::strcmp(kind, "multi")) {
        const char *sp = std::strstr(params, "split="), *cp = std::strstr(params, "cols=");
        if (!sp || !cp) return;
        e.passes = parse_list(sp + 6, e.split, kMaxPasses);
        const int nc = parse_list(cp + 5, e.cols, kMaxPasses);
        if (e.passes < 2 || nc != e.passes) return;
        if (const char *lp = std::strstr(params, "layout=")) {
            const int layout = std::atoi(lp + 7);
            if (layout != kPadded && layout != kHalves) return;
            e.pass_layout = layout;
        }
        int sum = 0;
        for (int i = 0; i < e.passes; i++) {
            if (e.split[i] < kPassMinLog2 || e.split[i] > kPassMaxLog2) return;
            if (!pass_cols_ok(e.split[i], e.cols[i], e.pass_layout)) return;
            sum += e.split[i];
        }
        if (sum != log2n) return;
        e.kind = TuneEntry::kMulti;
    } else {
        return;
    }
    g_table[log2n][regime] = e;
}

void load_table() {
    FILE *f = std::fopen(table_path().c_str(), "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        parse_line(line);
    }
    std::fclose(f);
}

}  // namespace

const RaggedTune &ragged_tune() {
    std::call_once(g_once, load_table);
    return g_ragged;
}

const TuneEntry &tune_lookup(unsigned log2n, int regime) {
    std::call_once(g_once, load_table);
    static const TuneEntry none;
    return log2n <= (unsigned)kTuneMaxLog2 && (regime == 0 || regime == 1) ? g_table[log2n][regime] : none;
}

}  // namespace onefft
