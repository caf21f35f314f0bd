/*
 * The reference's Seq_CC main (Seq_CC/CC.cpp:2687-2827) with its hot path
 * bound to libcent.so through include/cent.h: cent_load_edgelist replaces
 * buildGraph + createCSR, and compute_D1_AP_CC_shareBased_DegreeOrder
 * (CC.cpp:1728, called at CC.cpp:2812) assigns int64 farness instead of the
 * int CCs array.  Prints one "id farness" line per vertex
 * startNodeID..endNodeID, then computeCC's value for the same vertices.
 *
 *   cc_main <graph file>
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "cent.h"

int main(int argc, char* argv[]) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <graph>\n", argv[0]);
        return 2;
    }
    struct CSR* csr = NULL;
    int st = cent_load_edgelist(argv[1], &csr);
    if (st) { fprintf(stderr, "load: %s\n", cent_strerror(st)); return 1; }
    int64_t* CCs = (int64_t*)calloc((size_t)csr->csrVSize, sizeof(int64_t));
    int64_t* plain = (int64_t*)calloc((size_t)csr->csrVSize, sizeof(int64_t));
    if (!CCs || !plain) return 1;
    st = compute_D1_AP_CC_shareBased_DegreeOrder(csr, CCs);
    if (st) { fprintf(stderr, "compute_D1_AP_CC_shareBased_DegreeOrder: %s\n", cent_strerror(st)); return 1; }
    st = computeCC(csr, plain);
    if (st) { fprintf(stderr, "computeCC: %s\n", cent_strerror(st)); return 1; }
    for (int32_t v = csr->startNodeID; v <= csr->endNodeID; ++v)
        printf("%d %lld %lld\n", v, (long long)CCs[v], (long long)plain[v]);
    cent_csr_free(csr);
    free(CCs);
    free(plain);
    return 0;
}
