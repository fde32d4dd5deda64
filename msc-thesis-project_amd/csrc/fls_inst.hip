// One explicit instantiation of a FastLanes kernel family per compilation,
// selected with -DFLS_FAMILY / -DFLS_T / -DFLS_S / -DFLS_NT / -DFLS_PRED by the
// Makefile, so the per-bit-width specialisations of each family and variant
// compile in parallel.
#include "fls_kernels.hpp"

#if FLS_FAMILY == 0
template struct flsgpu_hip::FlsDecompress<FLS_T, FLS_S, FLS_NT>;
#elif FLS_FAMILY == 1
template struct flsgpu_hip::FlsAny<FLS_T, FLS_S>;
#elif FLS_FAMILY == 2
template struct flsgpu_hip::FlsFilterSum<FLS_T, FLS_S, FLS_PRED>;
#elif FLS_FAMILY == 3
template struct flsgpu_hip::FlsCompute<FLS_T>;
#elif FLS_FAMILY == 4
template struct flsgpu_hip::FlsMulti<FLS_T>;
#elif FLS_FAMILY == 5
template struct flsgpu_hip::FlsDecompressStaged<FLS_T, FLS_S>;
#else
#error "unknown FLS_FAMILY"
#endif
