// Explicit instantiation of the ALP-RD decoder, selected with -DRD_T / -DRD_NT.
#include "alprd_kernels.hpp"

template struct flsgpu_hip::AlprdDecompress<RD_T, RD_NT>;
