// D3Q19 FP16C kernel variants (SRT/TRT x extension subsets).
#include "lbm_launch.hpp"
namespace fx {
template const LbmKernels* kernels_for<19, ST_FP16C>(bool, uint32_t);
}
