// inst_level.hip -- explicit instantiation of the sync-free kernels for PolLevel
// (one translation unit per policy so the library builds in parallel).
#include "syncfree.hpp"

namespace sptrsv {
template hipError_t launch_solve<PolLevel, 1>(int, int, int, int, const SolveArgs&, hipStream_t);
}  // namespace sptrsv
