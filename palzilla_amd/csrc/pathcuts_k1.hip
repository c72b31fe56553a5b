// Node-level path cuts: the kernels for chains of 1 specular vertex.
#include "pathcuts_kernels.hpp"

namespace palz {
namespace cuts {

template void launch_expand<1>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, CutEntry*, uint32_t*,
                                hipStream_t, bool);
template void launch_walk<1, kCount>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);
template void launch_walk<1, kWrite>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);
template void launch_walk<1, kEmit>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);

}  // namespace cuts
}  // namespace palz
