// inst_f64.hip -- explicit instantiation of the sync-free kernels for PolF64
// (one translation unit per policy so the library builds in parallel).
#include "syncfree.hpp"

namespace sptrsv {
template hipError_t launch_solve<PolF64, 64>(int, int, int, int, const SolveArgs&, hipStream_t);
}  // namespace sptrsv
