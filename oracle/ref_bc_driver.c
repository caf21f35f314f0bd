/*
 * Driver for the reference's own computeBC (Seq_Brandes/Brandes.c:224-372),
 * compiled from the reference sources by oracle/Makefile into oracle/_ref/
 * (never committed, never shipped in the product).  Used only by tests/.
 * Usage: ref_bc <graph-file> <out-file>; output one "%.9g" FP32 score per id
 * from startNodeID to endNodeID.  Sources start at startNodeID (the reference
 * starts at the global nextBFS = 0, which reads csrV[-1]-style padding on
 * 1-based graphs, Brandes.c:24,249; the driver sets it first).
 */
#include <stdio.h>
#include <stdlib.h>
#include "../Lib/CSR/CSR.h"

extern int nextBFS;
void computeBC(struct CSR* _csr, float* _BCs);

int main(int argc, char** argv) {
    if (argc != 3) { fprintf(stderr, "usage: ref_bc graph out\n"); return 2; }
    struct Graph* g = buildGraph(argv[1]);
    struct CSR* csr = createCSR(g);
    float* bc = (float*)calloc(sizeof(float), csr->csrVSize);
    nextBFS = csr->startNodeID;
    computeBC(csr, bc);
    FILE* f = fopen(argv[2], "w");
    if (!f) return 3;
    for (int v = csr->startNodeID; v <= csr->endNodeID; ++v) fprintf(f, "%.9g\n", bc[v]);
    fclose(f);
    return 0;
}
