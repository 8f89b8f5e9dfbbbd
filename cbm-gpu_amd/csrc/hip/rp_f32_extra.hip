// Instantiates the row-pass kernels of the pass shapes PassF32Kahan, PassF32Self
// (row_pass.hpp); one translation unit per few shapes keeps the build parallel.
#include "row_pass.hpp"

namespace cbmk {
namespace rp {

template hipError_t dispatch_ve<PassF32Kahan>(const RowPass &, const Launch &, hipStream_t);
template hipError_t dispatch_ve<PassF32Self>(const RowPass &, const Launch &, hipStream_t);

}  // namespace rp
}  // namespace cbmk
