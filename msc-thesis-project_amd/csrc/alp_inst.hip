// One explicit instantiation of an ALP kernel family per compilation,
// selected with -DALP_FAMILY / -DALP_T / -DALP_S / -DALP_NT / -DALP_LAYOUT /
// -DALP_PATCH by the Makefile (family 0 decompress, 1 any, 2 multi, 3
// decompress staged through LDS).
#include "alp_kernels.hpp"

#if ALP_FAMILY == 0
template struct flsgpu_hip::AlpDecompress<ALP_T, ALP_S, ALP_NT, ALP_LAYOUT, ALP_PATCH>;
#elif ALP_FAMILY == 1
template struct flsgpu_hip::AlpAny<ALP_T, ALP_LAYOUT, ALP_PATCH>;
#elif ALP_FAMILY == 2
template struct flsgpu_hip::AlpMulti<ALP_T, ALP_LAYOUT, ALP_PATCH>;
#elif ALP_FAMILY == 3
template struct flsgpu_hip::AlpDecompressStaged<ALP_T, ALP_S, ALP_LAYOUT, ALP_PATCH>;
#else
#error "unknown ALP_FAMILY"
#endif
