// Persistent job kernel instantiated for the Goldilocks 2-limb policy at 8 elements per thread, and
// the dispatch by elements per thread (the 16-element variant is in kernels_jobs_l2_e16.hip; code in
// jobs_impl.hpp).
#include "jobs_impl.hpp"

namespace hmpc_amd
{
    hipError_t fast_jobs_occupancy_l2_e16(int* blocks_per_cu);
    hipError_t launch_fast_jobs_l2_e16(job_desc const* table, job_desc const& base, u32 njobs, unsigned grid, u32* counter,
                                       u32* scratch, u64 slot_words, fast_params const& P, hipStream_t s);

    hipError_t fast_jobs_occupancy_l2(int elems, int* blocks_per_cu)
    {
        if (elems == 8)
            return fast::jobs_occupancy_e<fast::arith_gold, 8>(blocks_per_cu);
        return fast_jobs_occupancy_l2_e16(blocks_per_cu);
    }

    hipError_t launch_fast_jobs_l2(transform_args const& T, job_desc const* table, job_desc const& base, u32 njobs,
                                   unsigned grid, u32* counter, u32* scratch, u64 slot_words, hipStream_t s)
    {
        if (T.elems == 8)
            return fast::launch_jobs_e<fast::arith_gold, 8>(table, base, njobs, grid, counter, scratch, slot_words,
                                                            fast::montgomery_params(T), s);
        return launch_fast_jobs_l2_e16(table, base, njobs, grid, counter, scratch, slot_words, fast::montgomery_params(T), s);
    }
}
