// inst_f32.hip -- explicit instantiation of the sync-free kernels for PolF32
// (one translation unit per policy so the library builds in parallel).
#include "syncfree.hpp"

namespace sptrsv {
template hipError_t launch_solve<PolF32, 64>(int, int, int, int, const SolveArgs&, hipStream_t);
}  // namespace sptrsv
