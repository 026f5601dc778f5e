// The gang kernel (sdim_amd/csrc/gang.hip) reads the XCC id and waits on store
// counters with inline GPU assembly, so the CPU build of the engine leaves it
// out: a submission that would need a gang fails with "gang kernel launch
// failed".  CPU-build tests run with gang_min_n = 0.
#include <hip/hip_runtime.h>

#include "engine.h"

namespace sdim {
template <int D>
int prepare_gang() { return 0; }
template <int D>
int launch_gang_d(const KParams&, const GangParams&, int, int, hipStream_t) { return -2; }
template int prepare_gang<2>();
template int prepare_gang<3>();
template int prepare_gang<5>();
template int prepare_gang<7>();
template int prepare_gang<11>();
template int launch_gang_d<2>(const KParams&, const GangParams&, int, int, hipStream_t);
template int launch_gang_d<3>(const KParams&, const GangParams&, int, int, hipStream_t);
template int launch_gang_d<5>(const KParams&, const GangParams&, int, int, hipStream_t);
template int launch_gang_d<7>(const KParams&, const GangParams&, int, int, hipStream_t);
template int launch_gang_d<11>(const KParams&, const GangParams&, int, int, hipStream_t);
}  // namespace sdim
