// Kruskal tensors (reference: src/ktensor/ktensor.c:23-78, src/ktensor/dump.c:23-54).
#include "internal.h"

#include <cstdio>
#include <cstdlib>

extern "C" {

int ptiNewKruskalTensor(ptiKruskalTensor *ktsr, ptiIndex nmodes, const ptiIndex ndims[], ptiIndex rank) {
    ktsr->nmodes = nmodes;
    ktsr->rank = rank;
    ktsr->ndims = static_cast<ptiIndex *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndex)));
    ktsr->lambda = static_cast<ptiValue *>(std::calloc(rank ? rank : 1, sizeof(ptiValue)));
    PTI_CHECK_ALLOC(ktsr->ndims, "KruskalTns New");
    PTI_CHECK_ALLOC(ktsr->lambda, "KruskalTns New");
    for (ptiIndex i = 0; i < nmodes; ++i) ktsr->ndims[i] = ndims[i];
    ktsr->fit = 0.0;
    ktsr->factors = nullptr;
    return 0;
}

void ptiFreeKruskalTensor(ptiKruskalTensor *ktsr) {
    if (ktsr->factors) {
        for (ptiIndex i = 0; i < ktsr->nmodes; ++i) {
            if (ktsr->factors[i]) {
                ptiFreeMatrix(ktsr->factors[i]);
                std::free(ktsr->factors[i]);
            }
        }
        std::free(ktsr->factors);
    }
    std::free(ktsr->ndims);
    std::free(ktsr->lambda);
    ktsr->factors = nullptr;
    ktsr->ndims = nullptr;
    ktsr->lambda = nullptr;
    ktsr->rank = 0;
    ktsr->fit = 0.0;
    ktsr->nmodes = 0;
}

int ptiDumpKruskalTensor(ptiKruskalTensor *ktsr, FILE *fp) {
    if (std::fprintf(fp, "nmodes: %" PRIu32 ", rank: %" PRIu32 "\n", ktsr->nmodes, ktsr->rank) < 0)
        PTI_FAIL(PTIERR_OS_ERROR + errno, "KruskalTns Dump", std::strerror(errno));
    for (ptiIndex m = 0; m < ktsr->nmodes; ++m) std::fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, ktsr->ndims[m]);
    std::fputs("\n", fp);
    std::fprintf(fp, "fit: %lf\n", ktsr->fit);
    std::fprintf(fp, "lambda:\n");
    for (ptiIndex r = 0; r < ktsr->rank; ++r) std::fprintf(fp, "%f ", (double)ktsr->lambda[r]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "Factor matrices:\n");
    for (ptiIndex m = 0; m < ktsr->nmodes; ++m) PTI_CHECK(ptiDumpMatrix(ktsr->factors[m], fp), "KruskalTns Dump");
    return 0;
}

}  // extern "C"
