// Fast transform kernels instantiated for the Goldilocks 2-limb policy (code in fast_impl.hpp).
#include "fast_impl.hpp"

namespace hmpc_amd
{
    hipError_t launch_fast_ntt_l2(transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                  hipStream_t s)
    {
        return fast::launch_ntt<fast::arith_gold>(T, fast::montgomery_params(T), n, inverse, in, out, s);
    }

    hipError_t launch_fast_polymul_l2(transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                      u32* workspace, hipStream_t s)
    {
        return fast::launch_polymul<fast::arith_gold>(T, fast::montgomery_params(T), n, a, b, out, workspace, s);
    }
}
