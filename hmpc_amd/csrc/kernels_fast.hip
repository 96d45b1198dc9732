// Dispatch of the fast transform path by limb count (instantiations in kernels_fast_l2.hip
// and kernels_fast_l4.hip, device code in fast_impl.hpp).
#include "kernels.hpp"
#include "fastmath.hpp"
#include "jobs.hpp"

namespace hmpc_amd
{
    hipError_t launch_fast_ntt_l2(transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                  hipStream_t s);
    hipError_t launch_fast_ntt_l4(transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                  hipStream_t s);
    hipError_t launch_fast_polymul_l2(transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                      u32* workspace, hipStream_t s);
    hipError_t launch_fast_polymul_l4(transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                      u32* workspace, hipStream_t s);
    hipError_t launch_fast_ntt_c4(transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                  hipStream_t s);
    hipError_t launch_fast_ntt_c5(transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                  hipStream_t s);
    hipError_t launch_fast_polymul_c4(transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                      u32* workspace, hipStream_t s);
    hipError_t launch_fast_polymul_c5(transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                      u32* workspace, hipStream_t s);
    hipError_t launch_fast_bgv_passes_c4(transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s);
    hipError_t launch_fast_bgv_passes_c5(transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s);

    hipError_t fast_jobs_occupancy_l2(int elems, int* blocks_per_cu);
    hipError_t fast_jobs_occupancy_l4(int elems, int* blocks_per_cu);
    hipError_t launch_fast_jobs_l2(transform_args const& T, job_desc const* table, job_desc const& base, u32 njobs,
                                   unsigned grid, u32* counter, u32* scratch, u64 slot_words, hipStream_t s);
    hipError_t launch_fast_jobs_l4(transform_args const& T, job_desc const* table, job_desc const& base, u32 njobs,
                                   unsigned grid, u32* counter, u32* scratch, u64 slot_words, hipStream_t s);

    hipError_t launch_fast_bgv_passes_l2(transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s);
    hipError_t launch_fast_bgv_passes_l4(transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s);

    hipError_t launch_fast_bgv_table_l2(transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                        u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32* ws, hipStream_t s);
    hipError_t launch_fast_bgv_table_l4(transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                        u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32* ws, hipStream_t s);

    hipError_t launch_fast_polymul_table_l2(transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                            u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                            u32 icols_blocks, u32* ws, hipStream_t s);
    hipError_t launch_fast_polymul_table_l4(transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                            u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                            u32 icols_blocks, u32* ws, hipStream_t s);

    bool fast_path_supported(int L, int form, int log_n)
    {
        // special-form moduli only; whole-polynomial tiles up to 2^12, two passes of at most
        // 2^12 points up to 2^20
        bool special = (L == 4 && form == form_r26) || (L == 2 && form == form_goldilocks);
        return special && log_n >= 2 && log_n <= 20;
    }

    namespace
    {
        bool cios_form(int L, int form) { return (L == 4 || L == 5) && (form == form_lazy || form == form_sparse); }
    }

    bool fast_transform_supported(int L, int form, int log_n)
    {
        return fast_path_supported(L, form, log_n) || (cios_form(L, form) && log_n >= 2 && log_n <= 20);
    }

    hipError_t launch_fast_ntt(int L, transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                               hipStream_t s)
    {
        if (L == 4 && T.form == form_r26)
            return launch_fast_ntt_l4(T, n, inverse, in, out, s);
        if (L == 2 && T.form == form_goldilocks)
            return launch_fast_ntt_l2(T, n, inverse, in, out, s);
        if (cios_form(L, T.form))
            return L == 4 ? launch_fast_ntt_c4(T, n, inverse, in, out, s) : launch_fast_ntt_c5(T, n, inverse, in, out, s);
        return hipErrorInvalidValue;
    }

    std::size_t fast_polymul_workspace_words(int L, int log_n, std::size_t n)
    {
        return log_n <= 12 ? 0 : 2 * static_cast<std::size_t>(ws_limbs(L)) * (n << log_n);
    }

    hipError_t launch_fast_polymul(int L, transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                   u32* workspace, hipStream_t s)
    {
        if (L == 4 && T.form == form_r26)
            return launch_fast_polymul_l4(T, n, a, b, out, workspace, s);
        if (L == 2 && T.form == form_goldilocks)
            return launch_fast_polymul_l2(T, n, a, b, out, workspace, s);
        if (cios_form(L, T.form))
            return L == 4 ? launch_fast_polymul_c4(T, n, a, b, out, workspace, s)
                          : launch_fast_polymul_c5(T, n, a, b, out, workspace, s);
        return hipErrorInvalidValue;
    }

    hipError_t fast_jobs_occupancy(int L, int elems, int* blocks_per_cu)
    {
        if (L == 4)
            return fast_jobs_occupancy_l4(elems, blocks_per_cu);
        if (L == 2)
            return fast_jobs_occupancy_l2(elems, blocks_per_cu);
        return hipErrorInvalidValue;
    }

    hipError_t launch_fast_jobs(int L, transform_args const& T, job_desc const* table, job_desc const& base, u32 njobs,
                                unsigned grid, u32* counter, u32* scratch, u64 slot_words, hipStream_t s)
    {
        if (L == 4)
            return launch_fast_jobs_l4(T, table, base, njobs, grid, counter, scratch, slot_words, s);
        if (L == 2)
            return launch_fast_jobs_l2(T, table, base, njobs, grid, counter, scratch, slot_words, s);
        return hipErrorInvalidValue;
    }

    hipError_t launch_fast_bgv_passes(int L, transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s)
    {
        if (L == 4 && T.form == form_r26)
            return launch_fast_bgv_passes_l4(T, base, n, ws, s);
        if (L == 2 && T.form == form_goldilocks)
            return launch_fast_bgv_passes_l2(T, base, n, ws, s);
        if (cios_form(L, T.form))
            return L == 4 ? launch_fast_bgv_passes_c4(T, base, n, ws, s) : launch_fast_bgv_passes_c5(T, base, n, ws, s);
        return hipErrorInvalidValue;
    }

    hipError_t launch_fast_bgv_table(int L, transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                     u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32* ws, hipStream_t s)
    {
        if (L == 4)
            return launch_fast_bgv_table_l4(T, uniform_log_n, max_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, ws, s);
        if (L == 2)
            return launch_fast_bgv_table_l2(T, uniform_log_n, max_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, ws, s);
        return hipErrorInvalidValue;
    }

    hipError_t launch_fast_polymul_table(int L, transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                         u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                         u32 icols_blocks, u32* ws, hipStream_t s)
    {
        if (L == 4)
            return launch_fast_polymul_table_l4(T, uniform_log_n, max_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, map_icols, icols_blocks,
                                                ws, s);
        if (L == 2)
            return launch_fast_polymul_table_l2(T, uniform_log_n, max_log_n, jobs, map_cols, cols_blocks, map_rows, rows_blocks, map_icols, icols_blocks,
                                                ws, s);
        return hipErrorInvalidValue;
    }
}
