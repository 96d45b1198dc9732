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

    hipError_t launch_fast_bgv_passes_l4(transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s)
    {
        return fast::launch_bgv_passes<fast::arith_r26>(T.elems, base, n, ws, T.r26, s);
    }

    hipError_t launch_fast_bgv_table_l4(transform_args const& T, int uniform_log_n, job_desc const* jobs, u32 const* map_cols,
                                         u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32* ws, hipStream_t s)
    {
        return fast::launch_bgv_table<fast::arith_r26>(T.elems, uniform_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, ws,
                                                    T.r26, s);
    }

    unsigned two_pass_tiles(int log_n, bool rows)
    {
        return fast::geom_for(log_n, rows).blocks_per_poly;
    }

    hipError_t launch_fast_polymul_table_l4(transform_args const& T, int uniform_log_n, job_desc const* jobs, u32 const* map_cols,
                                             u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                             u32 icols_blocks, u32* ws, hipStream_t s)
    {
        return fast::launch_polymul_table<fast::arith_r26>(T.elems, uniform_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, map_icols,
                                                        icols_blocks, ws, T.r26, s);
    }
}
