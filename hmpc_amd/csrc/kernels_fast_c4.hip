// Fast tile kernels instantiated for the 4-limb CIOS policy (any modulus with 4q < 2^128, code
// in fast_impl.hpp and jobs_impl.hpp): transforms, products, and BGV encryption over tiles with the
// plaintext modulus as a Montgomery constant.
#include "jobs_impl.hpp"

namespace hmpc_amd
{
    hipError_t launch_fast_ntt_c4(transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                  hipStream_t s)
    {
        return fast::launch_ntt<fast::arith_cios<4>>(T, fast::montgomery_params(T), n, inverse, in, out, s);
    }

    hipError_t launch_fast_polymul_c4(transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                      u32* workspace, hipStream_t s)
    {
        return fast::launch_polymul<fast::arith_cios<4>>(T, fast::montgomery_params(T), n, a, b, out, workspace, s);
    }

    hipError_t launch_fast_bgv_passes_c4(transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s)
    {
        return fast::launch_bgv_passes<fast::arith_cios<4>>(T.elems, base, n, ws, fast::montgomery_params(T), s);
    }
}
