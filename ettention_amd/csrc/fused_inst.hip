// Instantiates the fused plane kernels of one arithmetic mode and storage
// type.  Built once per (mode, type) pair with
//   -DETN_FUSED_FAST=true|false -DETN_FUSED_STORAGE=F32|F16|BF16
//   -DETN_FUSED_ENTRY=launch_fused_<fast|exact>_<F32|F16|BF16>
// so that the kernel variants compile in parallel translation units.
#include "fused_kernel.hpp"

#if !defined(ETN_FUSED_FAST) || !defined(ETN_FUSED_STORAGE) || !defined(ETN_FUSED_ENTRY)
#error "fused_inst.hip needs ETN_FUSED_FAST, ETN_FUSED_STORAGE and ETN_FUSED_ENTRY"
#endif

namespace etn {
namespace dev {

void ETN_FUSED_ENTRY(bool mask, bool over, bool vec, bool cap, const FusedLaunch& l, hipStream_t stream)
{
    using S = ETN_FUSED_STORAGE;
    constexpr bool kFast = ETN_FUSED_FAST;
    if (mask && over) launch_fused_typed<S, kFast, true, true>(vec, cap, l, stream);
    else if (mask) launch_fused_typed<S, kFast, true, false>(vec, cap, l, stream);
    else if (over) launch_fused_typed<S, kFast, false, true>(vec, cap, l, stream);
    else launch_fused_typed<S, kFast, false, false>(vec, cap, l, stream);
}

}  // namespace dev
}  // namespace etn
