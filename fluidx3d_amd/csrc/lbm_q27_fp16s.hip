// D3Q27 FP16S kernel variants (SRT/TRT x extension subsets).
#include "lbm_launch.hpp"
namespace fx {
template const LbmKernels* kernels_for<27, ST_FP16S>(bool, uint32_t);
}
