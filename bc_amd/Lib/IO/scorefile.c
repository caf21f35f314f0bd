/* The reference's on-disk result formats, so results can be written for and
 * checked against its files (Seq_Brandes/Brandes.c:2313-2322 writes scores;
 * Seq_CC/CC.cpp:2621-2684 reads answers).
 *
 * BC score file: a header line "start end" (startNodeID, endNodeID), then one
 * "%f" line per vertex start..end, in id order.
 *
 * Closeness answer file "<dataset>_ans.txt": one "id cc" line per vertex,
 * found next to a ".mtx" dataset by replacing the ".mtx" suffix of the file
 * name with "_ans.txt". */
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/cent.h"

int cent_write_bc_scores(const char* path, const struct CSR* csr, const double* bc) {
    if (!path || !csr || (!bc && csr->endNodeID >= csr->startNodeID)) return CENT_E_INVAL;
    FILE* f = fopen(path, "w");
    if (!f) return CENT_E_IO;
    int ok = fprintf(f, "%d %d\n", csr->startNodeID, csr->endNodeID) > 0;
    for (int32_t v = csr->startNodeID; ok && v <= csr->endNodeID; ++v) ok = fprintf(f, "%f\n", bc[v]) > 0;
    if (fclose(f) != 0) ok = 0;
    return ok ? CENT_OK : CENT_E_IO;
}

int cent_read_bc_scores(const char* path, int32_t* start, int32_t* end, double* out, int32_t cap) {
    if (!path || !start || !end || cap < 0 || (cap > 0 && !out)) return CENT_E_INVAL;
    FILE* f = fopen(path, "r");
    if (!f) return CENT_E_IO;
    int32_t s = 0, e = 0;
    int st = CENT_OK;
    if (fscanf(f, "%d %d", &s, &e) != 2 || s < 0 || e < s - 1) st = CENT_E_IO;
    else if (e >= cap) st = CENT_E_INVAL;  /* out must hold ids 0..end */
    for (int32_t v = s; st == CENT_OK && v <= e; ++v)
        if (fscanf(f, "%lf", &out[v]) != 1) st = CENT_E_IO;
    fclose(f);
    if (st == CENT_OK) {
        *start = s;
        *end = e;
    }
    return st;
}

int cent_ans_path(const char* dataset_path, char* out, size_t cap) {
    if (!dataset_path || !out || cap == 0) return CENT_E_INVAL;
    const char* slash = strrchr(dataset_path, '/');
    const char* name = slash ? slash + 1 : dataset_path;
    const char* ext = strstr(name, ".mtx");
    if (!ext) return CENT_E_INVAL;  /* the reference only derives it from .mtx names */
    const size_t head = (size_t)(ext - dataset_path);
    if (head + strlen("_ans.txt") + 1 > cap) return CENT_E_INVAL;
    memcpy(out, dataset_path, head);
    strcpy(out + head, "_ans.txt");
    return CENT_OK;
}

int cent_read_cc_answers(const char* path, const struct CSR* csr, int64_t* ans) {
    if (!path || !csr || !ans) return CENT_E_INVAL;
    FILE* f = fopen(path, "r");
    if (!f) return CENT_E_IO;
    int st = CENT_OK;
    for (int32_t k = csr->startNodeID; st == CENT_OK && k <= csr->endNodeID; ++k) {
        long long id = 0, cc = 0;
        if (fscanf(f, "%lld %lld", &id, &cc) != 2) st = CENT_E_IO;
        else if (id < 0 || id >= csr->csrVSize - 1) st = CENT_E_IO;
        else ans[id] = (int64_t)cc;
    }
    fclose(f);
    return st;
}

int cent_write_cc_answers(const char* path, const struct CSR* csr, const int64_t* cc) {
    if (!path || !csr || (!cc && csr->endNodeID >= csr->startNodeID)) return CENT_E_INVAL;
    FILE* f = fopen(path, "w");
    if (!f) return CENT_E_IO;
    int ok = 1;
    for (int32_t v = csr->startNodeID; ok && v <= csr->endNodeID; ++v) ok = fprintf(f, "%d %lld\n", v, (long long)cc[v]) > 0;
    if (fclose(f) != 0) ok = 0;
    return ok ? CENT_OK : CENT_E_IO;
}
