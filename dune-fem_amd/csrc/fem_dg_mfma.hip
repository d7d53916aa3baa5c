// The BF16-MFMA DG apply is the contraction policy Ctr<K, float, true> of the
// sum-factorised kernel in fem_dg.hip (dg_apply_bf16); this translation unit
// is kept empty so that the build lists stay stable.
#include "fem_dg.hpp"
