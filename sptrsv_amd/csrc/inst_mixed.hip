// inst_mixed.hip -- explicit instantiation of the sync-free kernels for PolMixed
// (one translation unit per policy so the library builds in parallel).
#include "syncfree.hpp"

namespace sptrsv {
template hipError_t launch_solve<PolMixed, 64>(int, int, int, int, const SolveArgs&, hipStream_t);
}  // namespace sptrsv
