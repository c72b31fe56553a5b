// Node-level path cuts: the kernels for chains of 2 specular vertices.
#include "pathcuts_kernels.hpp"

namespace palz {
namespace cuts {

template void launch_expand<2>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, CutEntry*, uint32_t*,
                                hipStream_t, bool);
template void launch_walk<2, kCount>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);
template void launch_walk<2, kWrite>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);
template void launch_walk<2, kEmit>(const CutsArgs&, const CutEntry*, uint32_t, uint32_t*, hipStream_t);

}  // namespace cuts
}  // namespace palz
