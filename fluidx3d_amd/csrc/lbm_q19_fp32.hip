// D3Q19 FP32 kernel variants (SRT/TRT x extension subsets).
#include "lbm_launch.hpp"
namespace fx {
template const LbmKernels* kernels_for<19, ST_FP32>(bool, uint32_t);
}
