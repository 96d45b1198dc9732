// Persistent job kernel instantiated for the Goldilocks 2-limb policy (code in jobs_impl.hpp).
#include "jobs_impl.hpp"

namespace hmpc_amd
{
    namespace
    {
        fast_params gold_params(transform_args const& T)
        {
            fast_params P{};
            for (int l = 0; l < max_limbs; ++l)
            {
                P.q[l] = T.P.q[l];
                P.q2[l] = T.q2[l];
                P.r2[l] = T.r2[l];
            }
            P.nprime = T.P.nprime;
            return P;
        }
    }

    hipError_t fast_jobs_occupancy_l2(int elems, int* blocks_per_cu)
    {
        return fast::jobs_occupancy<fast::arith_gold>(elems, blocks_per_cu);
    }

    hipError_t launch_fast_jobs_l2(transform_args const& T, job_desc const* table, job_desc const& base, u32 njobs,
                                   unsigned grid, u32* counter, u32* scratch, u64 slot_words, hipStream_t s)
    {
        return fast::launch_jobs<fast::arith_gold>(T.elems, table, base, njobs, grid, counter, scratch, slot_words,
                                                   gold_params(T), s);
    }

    hipError_t launch_fast_bgv_passes_l2(transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s)
    {
        return fast::launch_bgv_passes<fast::arith_gold>(T.elems, base, n, ws, gold_params(T), s);
    }

    hipError_t launch_fast_bgv_table_l2(transform_args const& T, int uniform_log_n, job_desc const* jobs, u32 const* map_cols,
                                         u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32* ws, hipStream_t s)
    {
        return fast::launch_bgv_table<fast::arith_gold>(T.elems, uniform_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, ws,
                                                    gold_params(T), s);
    }

    hipError_t launch_fast_polymul_table_l2(transform_args const& T, int uniform_log_n, job_desc const* jobs, u32 const* map_cols,
                                             u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                             u32 icols_blocks, u32* ws, hipStream_t s)
    {
        return fast::launch_polymul_table<fast::arith_gold>(T.elems, uniform_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, map_icols,
                                                        icols_blocks, ws, gold_params(T), s);
    }
}
