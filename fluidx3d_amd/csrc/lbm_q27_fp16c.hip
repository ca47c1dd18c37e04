// D3Q27 FP16C kernel variants (SRT/TRT x extension subsets).
#include "lbm_launch.hpp"
namespace fx {
template const LbmKernels* kernels_for<27, ST_FP16C>(bool, uint32_t);
}
