// Node-level path cuts: the expansion and counting walk for chains of 3 specular vertices.
#include "pathcuts_kernels.hpp"

namespace palz {
namespace cuts {

template void launch_expand<3>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, CutEntry*, uint32_t*,
                                hipStream_t, bool);
template void launch_walk<3, kCount>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);

}  // namespace cuts
}  // namespace palz
