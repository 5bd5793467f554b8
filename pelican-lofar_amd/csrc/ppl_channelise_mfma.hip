// ppl_channelise_mfma.hip -- split-precision DFT channeliser (BF16x3 / FP16x2
// MFMA GEMMs with FP32 accumulation).
#include "ppl_internal.h"

struct ppl_ppf_plan;

int ppl_mfma_create(ppl_ppf_plan*)
{
    ppl::set_error("split-precision DFT channeliser is not available in this build");
    return PPL_E_CONFIG;
}
void ppl_mfma_destroy(ppl_ppf_plan*) {}
int ppl_mfma_run(ppl_ppf_plan*, const int16_t*, const uint32_t*, unsigned, float*, int, hipStream_t)
{
    ppl::set_error("split-precision DFT channeliser is not available in this build");
    return PPL_E_CONFIG;
}
