/*
 * Text edge-list loader for the reference's dataset format (Appendix of the
 * reference's Lib/FileReader/FileReader.c:9-21 and Lib/AdjList/AdjList.c:13-150):
 * first line "n m" (trailing text ignored), then m lines "u v".  Ids are
 * 0-based if any edge endpoint is 0 (AdjList.c:39-86), 1-based otherwise; a
 * 1-based file keeps its ids and vertex 0 becomes an isolated vertex.
 * Lines starting with '%' or '#' are skipped.
 */
#include <stdio.h>
#include <ctype.h>
#include "../common.h"

static int next_line(FILE* f, char* buf, int cap) {
    while (fgets(buf, cap, f)) {
        const char* p = buf;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '%' || *p == '#' || *p == '\n' || *p == '\r' || *p == 0) continue;
        return 1;
    }
    return 0;
}

static int parse2(const char* s, long long* a, long long* b) {
    char* e1; char* e2;
    *a = strtoll(s, &e1, 10);
    if (e1 == s) return 0;
    *b = strtoll(e1, &e2, 10);
    if (e2 == e1) return 0;
    return 1;
}

int cent_load_edgelist(const char* path, struct CSR** out) { return cent_load_edgelist_ex(path, 0, out); }

int cent_load_edgelist_ex(const char* path, int32_t flags, struct CSR** out) {
    if (!path || !out || (flags & ~CENT_CSR_RAW)) return CENT_E_INVAL;
    FILE* f = fopen(path, "r");
    if (!f) return CENT_E_IO;
    char buf[4096];
    long long nn, mm;
    if (!next_line(f, buf, sizeof buf) || !parse2(buf, &nn, &mm) || nn < 0 || mm < 0 ||
        nn > INT32_MAX - 2) {
        fclose(f);
        return CENT_E_IO;
    }
    int32_t* u = (int32_t*)malloc(sizeof(int32_t) * (size_t)(mm > 0 ? mm : 1));
    int32_t* v = (int32_t*)malloc(sizeof(int32_t) * (size_t)(mm > 0 ? mm : 1));
    if (!u || !v) { free(u); free(v); fclose(f); return CENT_E_OOM; }
    int64_t m = 0;
    int zero = 0;
    long long mx = -1;
    while (m < mm && next_line(f, buf, sizeof buf)) {
        long long a, b;
        if (!parse2(buf, &a, &b) || a < 0 || b < 0 || a > INT32_MAX - 2 || b > INT32_MAX - 2) {
            free(u); free(v); fclose(f);
            return CENT_E_IO;
        }
        if (a == 0 || b == 0) zero = 1;
        if (a > mx) mx = a;
        if (b > mx) mx = b;
        u[m] = (int32_t)a; v[m] = (int32_t)b; ++m;
    }
    fclose(f);
    if (m != mm) { free(u); free(v); return CENT_E_IO; }
    int32_t n = zero ? (int32_t)nn : (int32_t)nn + 1;
    if (mx >= n) { free(u); free(v); return CENT_E_IO; }
    struct CSR* g = NULL;
    int st = cent_csr_from_edges(n, u, v, m, flags, &g, NULL);
    free(u); free(v);
    if (st) return st;
    g->startAtZero = zero;
    g->startNodeID = zero ? 0 : 1;
    g->endNodeID = n - 1;
    g->totalNodeNumber = (int32_t)nn;
    *out = g;
    return CENT_OK;
}

/* Writes csr in the reference's edge-list format: "count m" (count =
 * endNodeID - startNodeID + 1), then each undirected edge once as "u v" with
 * u < v, in row order.  The reference's loader takes the id base from the
 * first edge line (0-based iff it names vertex 0, Lib/AdjList/AdjList.c:39),
 * so a 0-based graph whose vertex 0 has no edge cannot be written
 * (CENT_E_UNSUPPORTED). */
int cent_write_edgelist(const char* path, const struct CSR* csr) {
    if (!path || !csr) return CENT_E_INVAL;
    const int32_t n = csr->csrVSize - 1;
    if (csr->startNodeID == 0 && csr->csrESize > 0 && n > 0 && csr->csrV[1] == csr->csrV[0])
        return CENT_E_UNSUPPORTED;
    FILE* f = fopen(path, "w");
    if (!f) return CENT_E_IO;
    int ok = fprintf(f, "%d %lld\n", csr->totalNodeNumber, (long long)(csr->csrESize / 2)) > 0;
    for (int32_t u = 0; ok && u < n; ++u) {
        int self = 0;  /* a self-loop is listed twice in its row: written once */
        for (int32_t j = csr->csrV[u]; ok && j < csr->csrV[u + 1]; ++j)
            if (csr->csrE[j] > u || (csr->csrE[j] == u && (self++ & 1) == 0))
                ok = fprintf(f, "%d %d\n", u, csr->csrE[j]) > 0;
    }
    if (fclose(f) != 0) ok = 0;
    return ok ? CENT_OK : CENT_E_IO;
}
