/*
 * The reference's Seq_Brandes main (Seq_Brandes/Brandes.c:2239-2293) with its
 * hot path bound to libcent.so through include/cent.h: buildGraph + createCSR
 * become cent_load_edgelist, the float BC array becomes double, computeBC
 * becomes the engine's reference-formula shim, and the scores are written in
 * the reference's BC_Score format (Brandes.c:2313-2322).
 *
 *   brandes_main <graph file> <score file> [textbook]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cent.h"

int main(int argc, char* argv[]) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s <graph> <score file> [textbook]\n", argv[0]);
        return 2;
    }
    struct CSR* csr = NULL;
    int st = cent_load_edgelist(argv[1], &csr);                  /* was buildGraph + createCSR */
    if (st) { fprintf(stderr, "load: %s\n", cent_strerror(st)); return 1; }
    double* BCs = (double*)calloc((size_t)csr->csrVSize, sizeof(double));  /* was float */
    if (!BCs) return 1;
    const int textbook = argc > 3 && strcmp(argv[3], "textbook") == 0;
    st = textbook ? computeBC(csr, BCs) : computeBC_compat(csr, BCs);
    if (st) { fprintf(stderr, "computeBC: %s\n", cent_strerror(st)); return 1; }
    st = cent_write_bc_scores(argv[2], csr, BCs);
    if (st) { fprintf(stderr, "write: %s\n", cent_strerror(st)); return 1; }
    printf("BC of %d vertices (ids %d..%d) written to %s\n", csr->totalNodeNumber, csr->startNodeID,
           csr->endNodeID, argv[2]);
    cent_csr_free(csr);
    free(BCs);
    return 0;
}
