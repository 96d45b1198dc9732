// Persistent job kernel instantiated for the Goldilocks 2-limb policy at 16 elements per thread (code
// in jobs_impl.hpp).
#include "jobs_impl.hpp"

namespace hmpc_amd
{
    hipError_t fast_jobs_occupancy_l2_e16(int* blocks_per_cu)
    {
        return fast::jobs_occupancy_e<fast::arith_gold, 16>(blocks_per_cu);
    }

    hipError_t launch_fast_jobs_l2_e16(job_desc const* table, job_desc const& base, u32 njobs, unsigned grid, u32* counter,
                                       u32* scratch, u64 slot_words, fast_params const& P, hipStream_t s)
    {
        return fast::launch_jobs_e<fast::arith_gold, 16>(table, base, njobs, grid, counter, scratch, slot_words, P, s);
    }
}
