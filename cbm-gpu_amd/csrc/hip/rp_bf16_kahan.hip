// Instantiates the row-pass kernels of the pass shapes PassBf16Kahan
// (row_pass.hpp); one translation unit per few shapes keeps the build parallel.
#include "row_pass.hpp"

namespace cbmk {
namespace rp {

template hipError_t dispatch_ve<PassBf16Kahan>(const RowPass &, const Launch &, hipStream_t);

}  // namespace rp
}  // namespace cbmk
