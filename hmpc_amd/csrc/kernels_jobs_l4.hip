// Persistent job kernel instantiated for the radix-2^26 4-limb policy (code in jobs_impl.hpp).
#include "jobs_impl.hpp"

namespace hmpc_amd
{
    hipError_t fast_jobs_occupancy_l4(int elems, int* blocks_per_cu)
    {
        return fast::jobs_occupancy<fast::arith_r26_round_prefetch>(elems, blocks_per_cu);
    }

    hipError_t launch_fast_jobs_l4(transform_args const& T, job_desc const* table, job_desc const& base, u32 njobs,
                                   unsigned grid, u32* counter, u32* scratch, u64 slot_words, hipStream_t s)
    {
        return fast::launch_jobs<fast::arith_r26_round_prefetch>(T.elems, table, base, njobs, grid, counter, scratch, slot_words, T.r26, s);
    }

    unsigned two_pass_tiles(int log_n, bool rows)
    {
        return fast::geom_for(log_n, rows).blocks_per_poly;
    }
}
