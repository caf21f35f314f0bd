// Autotuner table: graph signatures and the launch settings measured best
// for them.  Host code only (no HIP), so the table logic runs and is tested
// without a GPU; the trials themselves live in the engine (cent_tune).
//
// A signature buckets the properties the level kernels' cost depends on:
// vertex count (log2), average degree (half-log2 steps), degree skew
// (log2 of max degree over average degree) and BFS depth (log2 of the
// eccentricity probe + 1).  Lookups take the exact signature if present,
// else the nearest entry within an L1 distance of 2 buckets.
//
// Entries are per call type (CENT_TUNE_BC or CENT_TUNE_CC): one graph can
// want the source-parallel step for betweenness and the 512-source
// bit-parallel step for closeness.
//
// File format, one entry per line, '#' starts a comment:
//   lg_n deg_x2 skew lg_ecc what  hub_deg chunk_edges grid_waves frontier push_pct  ms
// (a line without the `what` field is a betweenness entry; a line without
// push_pct leaves it 0, the engine default).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/cent.h"

namespace {

struct Entry {
    cent_tune_sig sig;
    int32_t what;
    cent_tune_params p;
};

std::mutex g_mu;
std::vector<Entry> g_table;

int ilog2_round(double x) { return x <= 1.0 ? 0 : (int)lround(log2(x)); }

int sig_dist(const cent_tune_sig& a, const cent_tune_sig& b) {
    return abs(a.lg_n - b.lg_n) + abs(a.deg_x2 - b.deg_x2) + abs(a.skew - b.skew) + abs(a.lg_ecc - b.lg_ecc);
}

bool same_sig(const cent_tune_sig& a, const cent_tune_sig& b) { return sig_dist(a, b) == 0; }

}  // namespace

extern "C" int cent_tune_params_check(const cent_tune_params* p) {
    if (!p) return CENT_E_INVAL;
    if (p->hub_deg < 32 || p->hub_deg > (1 << 20)) return CENT_E_INVAL;
    if (p->chunk_edges < 64 || p->chunk_edges > (1 << 22)) return CENT_E_INVAL;
    if (p->grid_waves < 0 || p->grid_waves > (1 << 22)) return CENT_E_INVAL;
    if (p->frontier < CENT_FRONTIER_AUTO || p->frontier > CENT_FRONTIER_SOURCE) return CENT_E_INVAL;
    if (p->push_pct < -1 || p->push_pct > 1000) return CENT_E_INVAL;
    return CENT_OK;
}

extern "C" int cent_tune_signature(const int32_t* row_ptr, int32_t n, int64_t nnz, int32_t ecc,
                                   cent_tune_sig* out) {
    if (!out || n < 0 || nnz < 0 || ecc < 0 || (n > 0 && !row_ptr)) return CENT_E_INVAL;
    int32_t maxdeg = 0;
    for (int32_t v = 0; v < n; ++v) {
        const int32_t d = row_ptr[v + 1] - row_ptr[v];
        if (d > maxdeg) maxdeg = d;
    }
    const double avg = n > 0 ? (double)nnz / (double)n : 0.0;
    out->lg_n = ilog2_round((double)n);
    out->deg_x2 = avg > 1.0 ? (int)lround(2.0 * log2(avg)) : 0;
    out->skew = avg > 0.0 ? ilog2_round((double)maxdeg / avg) : 0;
    out->lg_ecc = ilog2_round((double)ecc + 1.0);
    return CENT_OK;
}

static bool what_ok(int32_t what) { return what == CENT_TUNE_BC || what == CENT_TUNE_CC; }

extern "C" int cent_tune_table_put(const cent_tune_sig* sig, int32_t what, const cent_tune_params* p) {
    if (!sig || !what_ok(what)) return CENT_E_INVAL;
    if (int st = cent_tune_params_check(p)) return st;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& e : g_table)
        if (e.what == what && same_sig(e.sig, *sig)) {
            e.p = *p;
            return CENT_OK;
        }
    g_table.push_back(Entry{*sig, what, *p});
    return CENT_OK;
}

extern "C" int cent_tune_table_get(const cent_tune_sig* sig, int32_t what, cent_tune_params* out) {
    if (!sig || !out || !what_ok(what)) return CENT_E_INVAL;
    std::lock_guard<std::mutex> lk(g_mu);
    int best = -1, bd = 3;
    for (size_t i = 0; i < g_table.size(); ++i) {
        if (g_table[i].what != what) continue;
        const int d = sig_dist(g_table[i].sig, *sig);
        if (d < bd) {
            bd = d;
            best = (int)i;
        }
    }
    if (best < 0) return CENT_E_NOTFOUND;
    *out = g_table[(size_t)best].p;
    return CENT_OK;
}

extern "C" int32_t cent_tune_table_size(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    return (int32_t)g_table.size();
}

extern "C" void cent_tune_table_clear(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_table.clear();
}

extern "C" int cent_tune_table_load(const char* path) {
    if (!path) return CENT_E_INVAL;
    FILE* f = fopen(path, "r");
    if (!f) return CENT_E_IO;
    std::vector<Entry> add;
    char line[512];
    int st = CENT_OK;
    while (fgets(line, sizeof line, f)) {
        char* hash = strchr(line, '#');
        if (hash) *hash = 0;
        char* p = line;
        while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
        if (!*p) continue;
        Entry e{};
        int v[11] = {0};
        double ms = 0.0;
        int nf = 0;
        {   // count the fields: 9 (betweenness entry), 10 (with `what`), 11 (and push_pct)
            char tmp[512];
            strncpy(tmp, p, sizeof tmp - 1);
            tmp[sizeof tmp - 1] = 0;
            char* save = nullptr;
            for (char* t = strtok_r(tmp, " \t\r\n", &save); t; t = strtok_r(nullptr, " \t\r\n", &save)) ++nf;
        }
        int got = -1;
        if (nf == 11) {
            got = sscanf(p, "%d %d %d %d %d %d %d %d %d %d %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                         &v[7], &v[8], &v[9], &ms);
            e.sig = cent_tune_sig{v[0], v[1], v[2], v[3]};
            e.what = v[4];
            e.p = cent_tune_params{v[5], v[6], v[7], v[8], v[9], ms};
            got = got == 11 ? 0 : -1;
        } else if (nf == 10) {
            got = sscanf(p, "%d %d %d %d %d %d %d %d %d %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                         &v[7], &v[8], &ms);
            e.sig = cent_tune_sig{v[0], v[1], v[2], v[3]};
            e.what = v[4];
            e.p = cent_tune_params{v[5], v[6], v[7], v[8], 0, ms};
            got = got == 10 ? 0 : -1;
        } else if (nf == 9) {
            got = sscanf(p, "%d %d %d %d %d %d %d %d %lf", &v[0], &v[1], &v[2], &v[3], &v[5], &v[6], &v[7], &v[8],
                         &ms);
            e.sig = cent_tune_sig{v[0], v[1], v[2], v[3]};
            e.what = CENT_TUNE_BC;
            e.p = cent_tune_params{v[5], v[6], v[7], v[8], 0, ms};
            got = got == 9 ? 0 : -1;
        }
        if (got != 0 || !what_ok(e.what) || cent_tune_params_check(&e.p) != CENT_OK) {
            st = CENT_E_IO;
            break;
        }
        add.push_back(e);
    }
    fclose(f);
    if (st) return st;  // a malformed file changes nothing
    for (const auto& e : add) cent_tune_table_put(&e.sig, e.what, &e.p);
    return CENT_OK;
}

extern "C" int cent_tune_table_save(const char* path) {
    if (!path) return CENT_E_INVAL;
    std::vector<Entry> snap;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        snap = g_table;
    }
    FILE* f = fopen(path, "w");
    if (!f) return CENT_E_IO;
    fprintf(f, "# cent autotuner table: lg_n deg_x2 skew lg_ecc what  hub_deg chunk_edges grid_waves frontier push_pct  ms\n");
    for (const auto& e : snap)
        fprintf(f, "%d %d %d %d %d  %d %d %d %d %d  %.4f\n", e.sig.lg_n, e.sig.deg_x2, e.sig.skew, e.sig.lg_ecc, e.what,
                e.p.hub_deg, e.p.chunk_edges, e.p.grid_waves, e.p.frontier, e.p.push_pct, e.p.ms);
    const bool ok = fflush(f) == 0;
    fclose(f);
    return ok ? CENT_OK : CENT_E_IO;
}
