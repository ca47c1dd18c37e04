// D3Q27 kernel variants (FP32/FP16S/FP16C x SRT/TRT).
#include "lbm_launch.hpp"
namespace fx {
const LbmKernels* kernels_q27(uint32_t storage, bool trt, bool plain) { return select_kernels<27>(storage, trt, plain); }
}
