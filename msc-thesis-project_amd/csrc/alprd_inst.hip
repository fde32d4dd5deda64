// Explicit instantiation of the ALP-RD decoder, selected with -DRD_T / -DRD_NT
// (RD_NT 2: the staged decoder, S = 2 for float and 4 for double; RD_NT 3:
// staged with S = 8, double only).
#include "alprd_kernels.hpp"

#if RD_NT == 2
template struct flsgpu_hip::AlprdDecompressStaged<RD_T, sizeof(RD_T) == 4 ? 2 : 4>;
#elif RD_NT == 3
template struct flsgpu_hip::AlprdDecompressStaged<RD_T, 8>;
#else
template struct flsgpu_hip::AlprdDecompress<RD_T, RD_NT>;
#endif
