// Instantiates the row-pass kernels of the pass shapes PassDualF32, PassDualBf16
// (row_pass.hpp); one translation unit per few shapes keeps the build parallel.
#include "row_pass.hpp"

namespace cbmk {
namespace rp {

template hipError_t dispatch_ve<PassDualF32>(const RowPass &, const Launch &, hipStream_t);
template hipError_t dispatch_ve<PassDualBf16>(const RowPass &, const Launch &, hipStream_t);

}  // namespace rp
}  // namespace cbmk
