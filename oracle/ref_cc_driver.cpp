/*
 * Driver for the reference's own closeness code, compiled from the reference
 * sources by oracle/Makefile into oracle/_ref/ (never committed, never shipped
 * in the product).  Used only by tests/ to check the CPU oracle against the
 * reference itself.  Usage: ref_cc {cc|d1ap} <graph-file> <out-file>
 * Output: one line "<id> <closeness>" per original id.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "headers.h"

int* computeCC(struct CSR* _csr, int* _CCs);                               /* Seq_CC/CC.cpp:37 */
void compute_D1_AP_CC_shareBased_DegreeOrder(struct CSR* _csr, int* _CCs); /* Seq_CC/CC.cpp:1728 */

int main(int argc, char** argv) {
    if (argc != 4) { fprintf(stderr, "usage: ref_cc {cc|d1ap} graph out\n"); return 2; }
    struct Graph* g = buildGraph(argv[2]);
    struct CSR* csr = createCSR(g);
    int start = csr->startNodeID, end = csr->endNodeID;
    int* out = NULL;
    if (!strcmp(argv[1], "cc")) {
        out = (int*)calloc(sizeof(int), csr->csrVSize);
        computeCC(csr, out);
    } else {
        compute_D1_AP_CC_shareBased_DegreeOrder(csr, csr->CCs);
        out = csr->CCs;
    }
    FILE* f = fopen(argv[3], "w");
    if (!f) return 3;
    for (int v = start; v <= end; ++v) fprintf(f, "%d %d\n", v, out[v]);
    fclose(f);
    return 0;
}
