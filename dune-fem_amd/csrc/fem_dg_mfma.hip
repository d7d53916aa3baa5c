// BF16-MFMA variant of the SIPG DG apply (FP32 data, FP32 accumulation).
#include "fem_dg.hpp"

namespace femamd {

void dg_apply_bf16(const DGArgs<float> &, int, const float *, float *, const int *, hipStream_t)
{
    throw Error(FEM_ERR_UNSUPPORTED, "BF16 MFMA DG apply not built yet");
}

}  // namespace femamd
