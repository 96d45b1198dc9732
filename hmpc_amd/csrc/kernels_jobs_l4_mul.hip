// Product tile launches over job tables for the radix-2^26 4-limb policy (code in jobs_impl.hpp).
#include "jobs_impl.hpp"

namespace hmpc_amd
{
    hipError_t launch_fast_polymul_table_l4(transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                             u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                             u32 icols_blocks, u32* ws, hipStream_t s)
    {
        return fast::launch_polymul_table<fast::arith_r26>(T.elems, uniform_log_n, max_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, map_icols,
                                                        icols_blocks, ws, T.r26, s);
    }
}
