// Instantiates the row-pass kernels of the pass shapes PassDualSelfF32, PassDualSelfBf16
// (row_pass.hpp); one translation unit per few shapes keeps the build parallel.
#include "row_pass.hpp"

namespace cbmk {
namespace rp {

template hipError_t dispatch_ve<PassDualSelfF32>(const RowPass &, const Launch &, hipStream_t);
template hipError_t dispatch_ve<PassDualSelfBf16>(const RowPass &, const Launch &, hipStream_t);

}  // namespace rp
}  // namespace cbmk
