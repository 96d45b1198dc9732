// Persistent job kernel of the fast path, gfx950.
//
// One launch runs a list of independent jobs (jobs.hpp). The grid is sized to the number of
// workgroups the device holds at once; every workgroup claims the next job with one atomic on a
// counter and runs it to completion, from inputs in HBM to outputs in HBM, then claims the next,
// until the list is exhausted (every workgroup exits once the counter passes the job count). A
// job never waits for another workgroup, so there is no inter-workgroup synchronisation and no
// cross-XCD coherence traffic: the only state shared between the passes of one job is the
// workgroup's own scratch slot, written and read by the same workgroup (visible after its
// barrier).
//
// Per job, for N = 2^log_n:
//   N <= 4096: the whole polynomial is one LDS tile (fast_impl.hpp); forward rounds, product
//              and inverse rounds run back to back.
//   N  > 4096: the two-pass decomposition N = N1 * N2 of fast_impl.hpp, pass by pass over the
//              job's tiles: forward column tiles -> scratch; row tiles (forward, product,
//              inverse) -> scratch; inverse column tiles -> output.
// BGV jobs sample u, p v + m and p w on the capability streams straight into the column tiles
// (three transforms instead of four: NTT(v) p + NTT(m) = NTT(p v + m)), and combine with the
// public key in the row pass: c0 = A o U + X0, c1 = B o U + X1
// (hmpc/expr/crypto/lhe/enc.hpp:19-69, hmpc/expr/crypto/lhe/randomness.hpp:80).
#pragma once

#include "fast_impl.hpp"
#include "jobs.hpp"

namespace hmpc_amd
{
    namespace fast
    {
        template<typename A, int E>
        __device__ __forceinline__ void scatter_last(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td)
        {
            constexpr int WM = shape<E>::wmax;
            int n = round_count<WM>(td.s);
            int a = round_start<WM>(td.s, n - 1, false);
            int w = round_width<WM>(td.s, n - 1, false);
            if (w == WM)
                scatter<A, E, WM>(x, lds, td, a, false);
            else if (WM > 1 && w == 1)
                scatter<A, E, 1>(x, lds, td, a, false);
            else if constexpr (WM > 2)
            {
                if (w == 2)
                    scatter<A, E, 2>(x, lds, td, a, false);
                else if constexpr (WM > 3)
                {
                    if (w == 3)
                        scatter<A, E, 3>(x, lds, td, a, false);
                }
            }
        }

        // ------------------------------------------------------------------ capability sampling
        // Fills the tile with stream `which` (0: u, 1: p v + m, 2: p w) of a BGV job. Butterfly
        // pair k < N/2 draws both of its coefficients (k, k + N/2) from one ChaCha20 block with
        // counter low32(poly N + k) (hmpc/expr/number_theoretic_transform.hpp:326-329,
        // hmpc/comp/queue.hpp:191-201). Column tiles hold rows r and r + N1/2 of each column,
        // so both coefficients of a pair land in the same tile.
        template<typename A, int T, int Which>
        __device__ __forceinline__ void sample_tile(u32* lds, job_desc const& J, bool cols, int log_n1, int log_seqs, u32 col0,
                                                    typename A::params const& P)
        {
            constexpr int which = Which;
            int const log_n = static_cast<int>(J.log_n);
            u32 const half = 1u << (log_n - 1);
            int const pairs = sample_pairs(cols, log_n, log_n1, log_seqs);
            int const bits = J.bits[which];
            u32 const* nonce = J.nonce[which];
            long long const p = static_cast<long long>(J.p_int);
            for (int e = tid_x(); e < pairs; e += T)
            {
                sample_slot sl = sample_slot_of(cols, e, log_n, log_n1, log_seqs, col0);
                u32 const k = sl.k;
                int const elo = sl.elo, ehi = sl.ehi;
                u32 blk[16];
                chacha20_block(blk, J.key, static_cast<u32>((static_cast<u64>(J.poly) << log_n) + k), nonce[0], nonce[1],
                               nonce[2]);
                long long zl = capability_small_words(blk[0], blk[1], bits);
                long long zh = capability_two_words(bits) ? capability_small_words(blk[2], blk[3], bits)
                                                          : centered_binomial_words(blk[1], 0u, bits);
                u32 xl[A::rl], xh[A::rl];
                if constexpr (A::mont_plaintext)
                {
                    // any plaintext modulus: p v + m = mont(v R, p R) + m R (p as a Montgomery constant)
                    A::from_signed(xl, zl, P);
                    A::from_signed(xh, zh, P);
                    if constexpr (which != 0)
                    {
                        A::mul_const(xl, J.p_mont, P);
                        A::mul_const(xh, J.p_mont, P);
                    }
                    if constexpr (which == 1)
                    {
                        A::add_signed(xl, J.msg[J.msg_off + k], P);
                        A::add_signed(xh, J.msg[J.msg_off + k + half], P);
                    }
                }
                else
                {
                    if constexpr (which != 0)
                    {
                        zl *= p;
                        zh *= p;
                    }
                    if constexpr (which == 1)
                    {
                        zl += J.msg[J.msg_off + k];
                        zh += J.msg[J.msg_off + k + half];
                    }
                    A::from_signed(xl, zl, P);
                    A::from_signed(xh, zh, P);
                }
                A::lds_put(lds, elo, xl);
                A::lds_put(lds, ehi, xh);
            }
        }

        // ------------------------------------------------------------------ jobs
        // Only one register tile is live at any time (the other operand streams from scratch),
        // which keeps the kernel within 128 VGPRs.
        template<typename A, int E>
        __device__ __forceinline__ void run_polymul_job(job_desc const& J, u32* lds, u32* scratch, typename A::params const& P)
        {
            constexpr int T = tile / E;
            constexpr int WM = shape<E>::wmax;
            int const log_n = static_cast<int>(J.log_n);
            std::size_t const N = std::size_t{1} << log_n;
            u32 x[E][A::rl];
            if (log_n <= tile_log)
            {
                tile_desc td;
                io_desc ia, io, is;
                job_single_tile(log_n, J.in_count, J.in_off, td, ia);
                job_single_tile(log_n, J.out_count, J.out_off, td, io);
                job_single_tile(log_n, N, 0, td, is);
                // NTT(a) -> scratch
                tile_load<A, T>(lds, J.in0, ia);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, false);
                tile_store<A, T, true>(lds, scratch, is, false, P);
                __syncthreads();
                // NTT(b) in registers, times NTT(a), inverse
                tile_load<A, T>(lds, J.in1, ia);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, true);
                with_operands_last<A, E, 0>(x, td, is.valid, scratch, N, 0, nullptr, 0, 0, P);
                dispatch_round_regs<A, E>(round_width<WM>(td.s, 0, true), x, lds, td, 1u, 0u, J.twi, P, true);
                __syncthreads();
                tile_rounds<A, E, true>(x, lds, td, 1u, 0u, J.twi, P, true, 1, false);
                tile_store<A, T>(lds, J.out0, io, true, P);
                __syncthreads();
                return;
            }
            u32* sa = scratch;
            u32* sb = scratch + A::wsl * N;
            pass_geom const gc = geom_for(log_n, false);
            pass_geom const gr = geom_for(log_n, true);
            // forward column pass of a and b
            for (u32 t = 0; t < 2 * gc.blocks_per_poly; ++t)
            {
                u32 which = t >= gc.blocks_per_poly ? 1u : 0u;
                u32 ct = t - which * gc.blocks_per_poly;
                tile_desc td;
                io_desc ii, is;
                job_cols_tile(gc, ct, J.in_count, J.in_off, td, ii);
                job_cols_tile(gc, ct, N, 0, td, is);
                tile_load<A, T>(lds, which ? J.in1 : J.in0, ii);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, false);
                tile_store<A, T, true>(lds, which ? sb : sa, is, ws_fold_columns(log_n), P);
                __syncthreads();
            }
            // row pass: NTT(a) rows back into scratch; NTT(b) rows in registers, product,
            // inverse rows (scaled later) into scratch
            for (u32 t = 0; t < gr.blocks_per_poly; ++t)
            {
                tile_desc td;
                io_desc is;
                u32 hi0 = job_rows_tile(gr, t, N, 0, td, is);
                u32 base0 = (1u << gr.log_n1) + hi0;
                tile_load<A, T, true>(lds, sa, is);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, base0, 1u, J.twf, P, false, 0, false);
                tile_store<A, T, true>(lds, sa, is, false, P);
                __syncthreads();
                tile_load<A, T, true>(lds, sb, is);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, base0, 1u, J.twf, P, false, 0, true);
                with_operands_last<A, E, 0>(x, td, is.valid, sa, N, is.base, nullptr, 0, 0, P);
                dispatch_round_regs<A, E>(round_width<WM>(td.s, 0, true), x, lds, td, base0, 1u, J.twi, P, false);
                __syncthreads();
                tile_rounds<A, E, true>(x, lds, td, base0, 1u, J.twi, P, false, 1, false);
                tile_store<A, T, true>(lds, sa, is, false, P);
                __syncthreads();
            }
            // inverse column pass with the final scaling, canonical output
            for (u32 t = 0; t < gc.blocks_per_poly; ++t)
            {
                tile_desc td;
                io_desc is, io;
                job_cols_tile(gc, t, N, 0, td, is);
                job_cols_tile(gc, t, J.out_count, J.out_off, td, io);
                tile_load<A, T, true>(lds, sa, is);
                __syncthreads();
                tile_rounds<A, E, true>(x, lds, td, 1u, 0u, J.twi, P, true, 0, false, gc.log_n2);
                tile_store<A, T>(lds, J.out0, io, true, P);
                __syncthreads();
            }
        }

        // c = k o U + X: X in registers after its forward transform, U and k read per element,
        // result written through LDS with coalesced stores (canonical)
        template<typename A, int E, bool WL = false>
        __device__ __forceinline__ void bgv_emit(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, u32 const* U,
                                                 std::size_t u_count, std::size_t u_base, u32 const* key,
                                                 io_desc const& ik, u32* out, io_desc const& io,
                                                 typename A::params const& P)
        {
            constexpr int T = tile / E;
            with_operands_last<A, E, 1>(x, td, ik.valid, U, u_count, u_base, key, ik.count, ik.base, P);
            scatter_last<A, E>(x, lds, td);
            tile_sync<WL>();
            tile_store<A, T, false, WL>(lds, out, io, true, P);
            tile_sync<WL>();
        }

        // the same with NTT(u) in registers (the unit order of x)
        template<typename A, int E, bool WL = false>
        __device__ __forceinline__ void bgv_emit_regs(u32 (&x)[E][A::rl], u32 const (&u)[E][A::rl], u32* lds, tile_desc const& td,
                                                      u32 const* key, io_desc const& ik, u32* out, io_desc const& io,
                                                      typename A::params const& P)
        {
            constexpr int T = tile / E;
            with_key_regs_last<A, E>(x, u, td, ik.valid, key, ik.count, ik.base, P);
            scatter_last<A, E>(x, lds, td);
            tile_sync<WL>();
            tile_store<A, T, false, WL>(lds, out, io, true, P);
            tile_sync<WL>();
        }

        // The one-wave BGV row passes keep NTT(u) in registers while they transform X0 and X1 (40 more
        // VGPRs): no write-back of U's rows and no re-read per ciphertext output (1.0 of 2.5 GB per
        // batch of 1024 at N = 2^14; BGV rows 0.72 -> 0.66 ms). Their launch bounds allow 3 waves per
        // SIMD, so that rows of 2^7 fit without spilling (134 VGPRs; as fast as 128 with 14 spilled).
        // The other row kernels, whose mixed-length variants would spill heavily, read U back from the
        // workspace.
        template<typename A>
        constexpr bool bgv_u_in_registers()
        {
            return A::wave_blocks;
        }

        template<typename A, int E, int Which>
        __device__ __forceinline__ void bgv_columns(job_desc const& J, u32* lds, u32* dst, pass_geom const& gc,
                                                    u32 (&x)[E][A::rl], typename A::params const& P)
        {
            constexpr int T = tile / E;
            std::size_t const N = std::size_t{1} << J.log_n;
            for (u32 ct = 0; ct < gc.blocks_per_poly; ++ct)
            {
                tile_desc td;
                io_desc is;
                job_cols_tile(gc, ct, N, 0, td, is);
                sample_tile<A, T, Which>(lds, J, true, gc.log_n1, gc.log_seqs, ct << gc.log_seqs, P);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, false);
                tile_store<A, T, true>(lds, dst, is, ws_fold_columns(gc.log_n), P);
                __syncthreads();
            }
        }

        template<typename A, int E>
        __device__ __forceinline__ void run_bgv_job(job_desc const& J, u32* lds, u32* scratch, typename A::params const& P)
        {
            constexpr int T = tile / E;
            int const log_n = static_cast<int>(J.log_n);
            std::size_t const N = std::size_t{1} << log_n;
            u32 x[E][A::rl];
            if (log_n <= tile_log)
            {
                // public key and scratch planes hold N elements: in_count == N, in_off == 0
                tile_desc td;
                io_desc ik, io, is;
                job_single_tile(log_n, J.in_count, J.in_off, td, ik);
                job_single_tile(log_n, J.out_count, J.out_off, td, io);
                job_single_tile(log_n, N, 0, td, is);
                sample_tile<A, T, 0>(lds, J, false, 0, 0, 0u, P);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, false);
                tile_store<A, T, true>(lds, scratch, is, false, P);
                __syncthreads();
                sample_tile<A, T, 1>(lds, J, false, 0, 0, 0u, P);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, true);
                bgv_emit<A, E>(x, lds, td, scratch, N, 0, J.in0, ik, J.out0, io, P);
                sample_tile<A, T, 2>(lds, J, false, 0, 0, 0u, P);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, true);
                bgv_emit<A, E>(x, lds, td, scratch, N, 0, J.in1, ik, J.out1, io, P);
                return;
            }
            pass_geom const gc = geom_for(log_n, false);
            pass_geom const gr = geom_for(log_n, true);
            u32* su = scratch;
            // sample and run the column stages of u, p v + m, p w
            bgv_columns<A, E, 0>(J, lds, su, gc, x, P);
            bgv_columns<A, E, 1>(J, lds, scratch + A::wsl * N, gc, x, P);
            bgv_columns<A, E, 2>(J, lds, scratch + 2 * A::wsl * N, gc, x, P);
            // row stages: U back into scratch; X0, X1 in registers combined with the key
            for (u32 t = 0; t < gr.blocks_per_poly; ++t)
            {
                tile_desc td;
                io_desc is, ik, io;
                u32 hi0 = job_rows_tile(gr, t, N, 0, td, is);
                job_rows_tile(gr, t, J.in_count, J.in_off, td, ik);
                job_rows_tile(gr, t, J.out_count, J.out_off, td, io);
                u32 base0 = (1u << gr.log_n1) + hi0;
                tile_load<A, T, true>(lds, su, is);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, base0, 1u, J.twf, P, false, 0, false);
                tile_store<A, T, true>(lds, su, is, false, P);
                __syncthreads();
                for (int which = 1; which < 3; ++which)
                {
                    tile_load<A, T, true>(lds, scratch + which * A::wsl * N, is);
                    __syncthreads();
                    tile_rounds<A, E, false>(x, lds, td, base0, 1u, J.twf, P, false, 0, true);
                    bool first = which == 1;
                    bgv_emit<A, E>(x, lds, td, su, N, is.base, first ? J.in0 : J.in1, ik, first ? J.out0 : J.out1, io, P);
                }
            }
        }

        // ------------------------------------------------------------------ the persistent kernel
        template<typename A>
        __global__ void k_reset_counter(u32* counter)
        {
            if (threadIdx.x == 0)
                counter[0] = 0;
        }

        // Jobs come from `table` or, when it is null, from `base` with per-job strides (job j of
        // a uniform batch: element offsets + j N, capability polynomial index + j).
        template<typename A, int E>
        __global__ void __launch_bounds__(tile / E) k_jobs(job_desc const* __restrict__ table, job_desc base, u32 njobs,
                                                           u32* __restrict__ counter, u32* __restrict__ scratch,
                                                           u64 slot_words, typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            u32* slot = scratch + static_cast<std::size_t>(blockIdx.x) * slot_words;
            for (;;)
            {
                // the tile is free here (every job ends with a barrier after its last LDS read)
                if (threadIdx.x == 0)
                    lds[0] = atomicAdd(counter, 1u);
                __syncthreads();
                u32 j = __builtin_amdgcn_readfirstlane(lds[0]);
                __syncthreads();
                if (j >= njobs)
                    break;
                job_desc J;
                if (table)
                    J = table[j];
                else
                {
                    J = base;
                    u64 step = static_cast<u64>(j) << J.log_n;
                    if (J.type == job_polymul)
                        J.in_off += step;
                    J.out_off += step;
                    J.msg_off += step;
                    J.poly += j;
                }
                if (J.type == job_polymul)
                    run_polymul_job<A, E>(J, lds, slot, P);
                else
                    run_bgv_job<A, E>(J, lds, slot, P);
            }
        }

        // ------------------------------------------------------------------ two-pass BGV (N > 4096)
        // The batched encryption of n ciphertexts as two launches over tiles instead of one
        // workgroup per ciphertext: every ciphertext's tiles run in parallel (short latency for
        // small batches). Workspace: the three column-pass results [3][wsl][n N].
        //   column pass: blockIdx.x = (p * 3 + which) * tiles + t  (sample stream `which` of ciphertext p)
        //   row pass:    blockIdx.x = p * tiles + t
        // column tiles of 2^7-point columns (32 per tile) on the radix-2^26 policy at 8 elements per
        // thread: the forward rounds at stages 0-2 and 3-5 and the inverse rounds at 1-3 and 4-6 have
        // wave-uniform twiddles and take them in expanded form (as the streaming column passes)
        template<typename A, int E, int S>
        constexpr bool cols_expanded()
        {
            return std::is_same<A, arith_r26>::value && E == 8 && S == 7;
        }

        template<typename A, int E, int Which, int S = 0, bool X = false>
        __device__ __forceinline__ void bgv_column_tile(job_desc const& J, u32* lds, u32* dst, pass_geom const& gc, u32 ct,
                                                        std::size_t count, u32 p, typename A::params const& P,
                                                        u32 const* __restrict__ twx = nullptr)
        {
            constexpr int T = tile / E;
            u32 x[E][A::rl];
            tile_desc td;
            io_desc is;
            cols_setup(gc, p * gc.blocks_per_poly + ct, count, td, is);
            int log_n1 = gc.log_n1, log_seqs = gc.log_seqs;
            if constexpr (S > 0)
            {
                // the column length as a compile-time constant
                log_n1 = S;
                log_seqs = tile_log - S;
                td.s = S;
                td.log_seqs = log_seqs;
                is.log_seqs = log_seqs;
            }
            sample_tile<A, T, Which>(lds, J, true, log_n1, log_seqs, ct << log_seqs, P);
            __syncthreads();
            if constexpr (X)
            {
                // radix-2^26 columns of 2^7, 32 per tile: stages 0-2 and 3-5 have wave-uniform
                // twiddles (as in the streaming column passes) and take them in expanded form from
                // twx (a restrict kernel argument, so that the loads are scalar); stage 6 as before
                static_assert(std::is_same<A, arith_r26>::value && E == 8 && S == 7, "expanded column rounds");
                round_expanded<A, E, false>(x, lds, td, 0, 1u, 0u, twx, P);
                __syncthreads();
                round_expanded<A, E, false>(x, lds, td, 3, 1u, 0u, twx, P);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 2, false);
            }
            else
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, false);
            tile_store<A, T, true>(lds, dst, is, ws_fold_columns(gc.log_n), P);
        }

        template<typename A, int E, int S = 0>
        __global__ void __launch_bounds__(tile / E, E == 8 ? 4 : 1) k_bgv_cols(job_desc base, u32* __restrict__ ws,
                                                                              pass_geom gc, u64 n, typename A::params P,
                                                                              u32 const* __restrict__ twx)
        {
            extern __shared__ __align__(16) u32 lds[];
            u32 ct = blockIdx.x % gc.blocks_per_poly;
            u32 r = blockIdx.x / gc.blocks_per_poly;
            u32 which = r % 3;
            u32 p = r / 3;
            job_desc J = base;
            J.poly += p;
            J.msg_off += static_cast<u64>(p) << J.log_n;
            std::size_t count = static_cast<std::size_t>(n) << J.log_n;
            u32* dst = ws + which * A::wsl * count;
            constexpr bool X = std::is_same<A, arith_r26>::value && E == 8 && S == 7;
            if (which == 0)
                bgv_column_tile<A, E, 0, S, X>(J, lds, dst, gc, ct, count, p, P, twx);
            else if (which == 1)
                bgv_column_tile<A, E, 1, S, X>(J, lds, dst, gc, ct, count, p, P, twx);
            else
                bgv_column_tile<A, E, 2, S, X>(J, lds, dst, gc, ct, count, p, P, twx);
        }

        template<typename A, int E, int S = 0>
        __global__ void __launch_bounds__(A::wave_blocks ? 64 : tile / E, A::wave_blocks ? 3 : (E == 8 ? 4 : 1)) k_bgv_rows(job_desc base, u32* __restrict__ ws,
                                                                              pass_geom gr, u64 n, typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            constexpr int T = tile / E;
            std::size_t count = static_cast<std::size_t>(n) << base.log_n;
            // one-wave workgroups: block b serves wave b % (tile / 64 E) of tile b / (tile / 64 E)
            static_assert(!A::wave_blocks || (E == 8 && S > 0 && (1 << S) <= 64 * E),
                          "one-wave workgroups need wave-local rows at 8 elements per thread");
            u32 const tb = A::wave_blocks ? blockIdx.x / static_cast<u32>(tile / E / 64) : blockIdx.x;
            u32 p = tb / gr.blocks_per_poly;
            u32 rt = tb - p * gr.blocks_per_poly;
            tile_desc td;
            io_desc is, ik, io;
            u32 hi0 = rows_setup(gr, tb, count, td, is);  // workspace planes of n N
            job_rows_tile(gr, rt, base.in_count, base.in_off, td, ik);  // public key (one polynomial)
            if constexpr (S > 0)
            {
                td.s = S;  // the row length as a compile-time constant
                td.log_seqs = tile_log - S;
            }
            io = is;
            io.count = base.out_count;
            io.base = base.out_off + is.base;
            u32 base0 = (1u << gr.log_n1) + hi0;
            u32* su = ws;
            u32 x[E][A::rl];
            // with the wave-major unit order rows of at most 64 E elements are wave-local: no
            // workgroup barrier
            constexpr bool WL = A::wave_units && S > 0 && (1 << S) <= 64 * E;
            if constexpr (bgv_u_in_registers<A>())
            {
                // NTT(u) stays in registers (the last round's unit order, shared by X0 and X1)
                u32 u[E][A::rl];
                tile_load<A, T, true, WL>(lds, su, is);
                tile_sync<WL>();
                row_rounds_forward<A, E, S, WL>(u, lds, td, base0, 1u, base.twf, base.twfx, P, true);
                tile_sync<WL>();
                for (int which = 1; which < 3; ++which)
                {
                    tile_load<A, T, true, WL>(lds, ws + which * A::wsl * count, is);
                    tile_sync<WL>();
                    row_rounds_forward<A, E, S, WL>(x, lds, td, base0, 1u, base.twf, base.twfx, P, true);
                    bool first = which == 1;
                    bgv_emit_regs<A, E, WL>(x, u, lds, td, first ? base.in0 : base.in1, ik, first ? base.out0 : base.out1, io, P);
                }
            }
            else
            {
                // NTT(u) rows back into the workspace, read per element below
                tile_load<A, T, true, WL>(lds, su, is);
                tile_sync<WL>();
                tile_rounds<A, E, false, WL>(x, lds, td, base0, 1u, base.twf, P, false, 0, false);
                tile_store<A, T, true, WL>(lds, su, is, false, P);
                tile_sync<WL>();
                for (int which = 1; which < 3; ++which)
                {
                    tile_load<A, T, true, WL>(lds, ws + which * A::wsl * count, is);
                    tile_sync<WL>();
                    tile_rounds<A, E, false, WL>(x, lds, td, base0, 1u, base.twf, P, false, 0, true);
                    bool first = which == 1;
                    bgv_emit<A, E, WL>(x, lds, td, su, count, is.base, first ? base.in0 : base.in1, ik,
                                       first ? base.out0 : base.out1, io, P);
                }
            }
        }

        template<typename A>
        hipError_t launch_bgv_passes(int elems, job_desc const& base, u64 n, u32* ws, typename A::params const& P,
                                     hipStream_t s)
        {
            return with_elems<A>(elems, [&](auto EC) -> hipError_t {
                constexpr int E = decltype(EC)::value;
                constexpr std::size_t bytes = tile_bytes<A>();
                int log_n = static_cast<int>(base.log_n);
                if (log_n <= tile_log)
                    return hipErrorInvalidValue;
                pass_geom gc = geom_for(log_n, false);
                pass_geom gr = geom_for(log_n, true);
                hipError_t e = hipSuccess;
                // column length 2^7 from N = 2^14 on; row lengths 2^7 .. 2^10 at N = 2^14 .. 2^17
                // the radix-2^26 columns of 2^7 take the expanded table (bgv_column_tile)
                bool const cols7 = gc.log_n1 == 7 && (!std::is_same<A, arith_r26>::value || E != 8 || base.twfx);
                auto cols_kernel = cols7 ? k_bgv_cols<A, E, 7> : k_bgv_cols<A, E, 0>;
                using AB = typename bgv_rows_arith<A>::type;
                auto rows_kernel = k_bgv_rows<A, E, 0>;
                switch (gr.log_n2)
                {
                case 7: rows_kernel = k_bgv_rows<AB, E, 7>; break;
                case 8: rows_kernel = k_bgv_rows<AB, E, 8>; break;
                case 9: rows_kernel = k_bgv_rows<AB, E, 9>; break;
                case 10: rows_kernel = k_bgv_rows<AB, E, 10>; break;
                default: break;
                }
                // radix-2^26 rows of up to 64 E elements: one wave per workgroup (arith_r26_bgv_rows1)
                unsigned rows_grid = static_cast<unsigned>(n * gr.blocks_per_poly), rows_threads = tile / E;
                std::size_t rows_bytes = bytes;
                if constexpr (std::is_same<A, arith_r26>::value && E == 8)
                {
                    // rows of 2^9 read the expanded forward table up to entry 8 * 2^log_n1
                    bool const expanded = base.twfx && (std::size_t{8} << gr.log_n1) <= r26_tw_entries;
                    if (gr.log_n2 >= 7 && gr.log_n2 <= 9 && (gr.log_n2 < 9 || expanded))
                    {
                        rows_kernel = gr.log_n2 == 7   ? k_bgv_rows<arith_r26_bgv_rows1, E, 7>
                                      : gr.log_n2 == 8 ? k_bgv_rows<arith_r26_bgv_rows1, E, 8>
                                                       : k_bgv_rows<arith_r26_bgv_rows1, E, 9>;
                        rows_grid *= static_cast<unsigned>(tile / E / 64);
                        rows_threads = 64;
                        rows_bytes = bytes / static_cast<std::size_t>(tile / E / 64);
                    }
                }
                if ((e = allow_lds(cols_kernel, bytes)) != hipSuccess || (e = allow_lds(rows_kernel, rows_bytes)) != hipSuccess)
                    return e;
                {
                    probe_scope ps(s, "fast bgv columns");
                    cols_kernel<<<static_cast<unsigned>(3 * n * gc.blocks_per_poly), tile / E, bytes, s>>>(base, ws, gc, n, P,
                                                                                                        base.twfx);
                }
                {
                    probe_scope ps(s, "fast bgv rows");
                    rows_kernel<<<rows_grid, rows_threads, rows_bytes, s>>>(base, ws, gr, n, P);
                }
                return hipGetLastError();
            });
        }

        // ------------------------------------------------------------------ tiles over a job table
        // The two launches above for a table of BGV jobs of any two-pass degrees, keys and nonces
        // (one ciphertext each): block b of a launch works on job map[b] >> 10, local tile
        // map[b] & 1023 (at most 3 x 256 tiles per job, at N = 2^20). Each job keeps its three
        // column results in its own workspace region [3][wsl][N] at ws + ws_off.
        // S > 0: every job of the table has this column length (the pass length as a compile-time
        // constant, as in the uniform launches); likewise the row length of the row kernels below
        template<typename A, int E, int S = 0>
        __global__ void __launch_bounds__(tile / E, E == 8 ? 4 : 1) k_bgv_cols_table(job_desc const* __restrict__ jobs,
                                                                                    u32 const* __restrict__ map,
                                                                                    u32* __restrict__ ws,
                                                                                    typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            u32 m = map[blockIdx.x];
            job_desc J = jobs[m >> 10];
            u32 local = m & 1023u;
            pass_geom gc = geom_for(static_cast<int>(J.log_n), false);
            u32 which = local / gc.blocks_per_poly;
            u32 ct = local - which * gc.blocks_per_poly;
            std::size_t N = std::size_t{1} << J.log_n;
            u32* dst = ws + J.ws_off + which * A::wsl * N;
            // mixed degrees (S = 0): the column length 2^7 (N >= 2^14) at compile time too
            if (S == 7 || (S == 0 && gc.log_n1 == 7))
            {
                if (which == 0)
                    bgv_column_tile<A, E, 0, 7, cols_expanded<A, E, 7>()>(J, lds, dst, gc, ct, N, 0, P, J.twfx);
                else if (which == 1)
                    bgv_column_tile<A, E, 1, 7, cols_expanded<A, E, 7>()>(J, lds, dst, gc, ct, N, 0, P, J.twfx);
                else
                    bgv_column_tile<A, E, 2, 7, cols_expanded<A, E, 7>()>(J, lds, dst, gc, ct, N, 0, P, J.twfx);
            }
            else
            {
                if (which == 0)
                    bgv_column_tile<A, E, 0, 0>(J, lds, dst, gc, ct, N, 0, P);
                else if (which == 1)
                    bgv_column_tile<A, E, 1, 0>(J, lds, dst, gc, ct, N, 0, P);
                else
                    bgv_column_tile<A, E, 2, 0>(J, lds, dst, gc, ct, N, 0, P);
            }
        }

        template<typename A, int E, int S>
        __device__ __forceinline__ void bgv_rows_table_tile(job_desc const& J, u32* lds, u32* ws, u32 rt,
                                                            typename A::params const& P)
        {
            constexpr int T = tile / E;
            pass_geom gr = geom_for(static_cast<int>(J.log_n), true);
            std::size_t N = std::size_t{1} << J.log_n;
            u32* su = ws + J.ws_off;
            tile_desc td;
            io_desc is, ik, io;
            u32 hi0 = job_rows_tile(gr, rt, N, 0, td, is);
            job_rows_tile(gr, rt, J.in_count, J.in_off, td, ik);
            job_rows_tile(gr, rt, J.out_count, J.out_off, td, io);
            if constexpr (S > 0)
            {
                td.s = S;
                td.log_seqs = tile_log - S;
            }
            u32 base0 = (1u << gr.log_n1) + hi0;
            u32 x[E][A::rl];
            // wave-local rows (the one-wave policy) need no workgroup barrier
            constexpr bool WL = A::wave_units && S > 0 && (1 << S) <= 64 * E;
            if constexpr (bgv_u_in_registers<A>())
            {
                u32 u[E][A::rl];
                tile_load<A, T, true, WL>(lds, su, is);
                tile_sync<WL>();
                row_rounds_forward<A, E, S, WL>(u, lds, td, base0, 1u, J.twf, J.twfx, P, true);
                tile_sync<WL>();
                for (int which = 1; which < 3; ++which)
                {
                    tile_load<A, T, true, WL>(lds, su + which * A::wsl * N, is);
                    tile_sync<WL>();
                    row_rounds_forward<A, E, S, WL>(x, lds, td, base0, 1u, J.twf, J.twfx, P, true);
                    bool first = which == 1;
                    bgv_emit_regs<A, E, WL>(x, u, lds, td, first ? J.in0 : J.in1, ik, first ? J.out0 : J.out1, io, P);
                }
            }
            else
            {
                tile_load<A, T, true, WL>(lds, su, is);
                tile_sync<WL>();
                tile_rounds<A, E, false, WL>(x, lds, td, base0, 1u, J.twf, P, false, 0, false);
                tile_store<A, T, true, WL>(lds, su, is, false, P);
                tile_sync<WL>();
                for (int which = 1; which < 3; ++which)
                {
                    tile_load<A, T, true, WL>(lds, su + which * A::wsl * N, is);
                    tile_sync<WL>();
                    tile_rounds<A, E, false, WL>(x, lds, td, base0, 1u, J.twf, P, false, 0, true);
                    bool first = which == 1;
                    bgv_emit<A, E, WL>(x, lds, td, su, N, is.base, first ? J.in0 : J.in1, ik, first ? J.out0 : J.out1, io, P);
                }
            }
        }

        // the row policies of the table kernels: with W1 (radix-2^26 at 8 elements per thread, rows
        // of at most 2^9 elements) one wave per workgroup, block b serving wave b % 8 of map entry
        // b / 8
        template<typename A, bool W1>
        struct table_rows_policy
        {
            using polymul = typename rows_arith<A>::type;
            using bgv = typename bgv_rows_arith<A>::type;
        };
        template<>
        struct table_rows_policy<arith_r26, true>
        {
            using polymul = arith_r26_rows1;
            using bgv = arith_r26_bgv_rows1;
        };

        // whether the row launch of a table whose largest degree is 2^max_log_n runs one wave per
        // workgroup: every job's rows must be wave-local (two-pass degrees up to 2^16 at E = 8)
        template<typename A, int E>
        bool table_rows_one_wave(int max_log_n)
        {
            return std::is_same<A, arith_r26>::value && E == 8 && max_log_n > tile_log &&
                   geom_for(max_log_n, true).log_n2 <= 9;
        }

        // S > 0: every job has this row length; S = 0: mixed degrees, each block branching (uniformly)
        // to the row length of its own job (2^7..2^10 at compile time, others at run time)
        // W1: one wave per workgroup (table_rows_one_wave holds for the table: every row length is
        // 2^7 .. 2^9, the only cases the W1 instantiation handles)
        template<typename A, int E, int S = 0, bool W1 = false>
        __global__ void __launch_bounds__(W1 ? 64 : tile / E, W1 ? 3 : (E == 8 ? 4 : 1)) k_bgv_rows_table(job_desc const* __restrict__ jobs,
                                                                                    u32 const* __restrict__ map,
                                                                                    u32* __restrict__ ws,
                                                                                    typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            using AB = typename table_rows_policy<A, W1>::bgv;
            u32 m = map[W1 ? blockIdx.x / static_cast<u32>(tile / E / 64) : blockIdx.x];
            job_desc J = jobs[m >> 10];
            u32 rt = m & 1023u;
            if constexpr (S > 0)
                bgv_rows_table_tile<AB, E, S>(J, lds, ws, rt, P);
            else if constexpr (W1)
            {
                switch (geom_for(static_cast<int>(J.log_n), true).log_n2)
                {
                case 7: bgv_rows_table_tile<AB, E, 7>(J, lds, ws, rt, P); break;
                case 8: bgv_rows_table_tile<AB, E, 8>(J, lds, ws, rt, P); break;
                case 9: bgv_rows_table_tile<AB, E, 9>(J, lds, ws, rt, P); break;
                default: break;
                }
            }
            else
            {
                switch (geom_for(static_cast<int>(J.log_n), true).log_n2)
                {
                case 7: bgv_rows_table_tile<AB, E, 7>(J, lds, ws, rt, P); break;
                case 8: bgv_rows_table_tile<AB, E, 8>(J, lds, ws, rt, P); break;
                case 9: bgv_rows_table_tile<AB, E, 9>(J, lds, ws, rt, P); break;
                case 10: bgv_rows_table_tile<AB, E, 10>(J, lds, ws, rt, P); break;
                default: bgv_rows_table_tile<A, E, 0>(J, lds, ws, rt, P); break;
                }
            }
        }

        // log_n > 0: every job of the table has degree 2^log_n (pass lengths at compile time where
        // they are specialized); 0: mixed degrees
        template<typename A>
        hipError_t launch_bgv_table(int elems, int log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                    u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32* ws, typename A::params const& P,
                                    hipStream_t s)
        {
            return with_elems<A>(elems, [&](auto EC) -> hipError_t {
                constexpr int E = decltype(EC)::value;
                constexpr std::size_t bytes = tile_bytes<A>();
                pass_geom gc = geom_for(log_n, false);
                pass_geom gr = geom_for(log_n, true);
                bool const uniform = log_n > tile_log;
                auto cols_kernel = uniform && gc.log_n1 == 7 ? k_bgv_cols_table<A, E, 7> : k_bgv_cols_table<A, E, 0>;
                auto rows_kernel = k_bgv_rows_table<A, E, 0>;
                switch (uniform ? gr.log_n2 : 0)
                {
                case 7: rows_kernel = k_bgv_rows_table<A, E, 7>; break;
                case 8: rows_kernel = k_bgv_rows_table<A, E, 8>; break;
                case 9: rows_kernel = k_bgv_rows_table<A, E, 9>; break;
                case 10: rows_kernel = k_bgv_rows_table<A, E, 10>; break;
                default: break;
                }
                unsigned rows_grid = rows_blocks, rows_threads = tile / E;
                std::size_t rows_bytes = bytes;
                if constexpr (std::is_same<A, arith_r26>::value && E == 8)
                {
                    if (table_rows_one_wave<A, E>(max_log_n))
                    {
                        switch (uniform ? gr.log_n2 : 0)
                        {
                        case 7: rows_kernel = k_bgv_rows_table<A, E, 7, true>; break;
                        case 8: rows_kernel = k_bgv_rows_table<A, E, 8, true>; break;
                        case 9: rows_kernel = k_bgv_rows_table<A, E, 9, true>; break;
                        default: rows_kernel = k_bgv_rows_table<A, E, 0, true>; break;
                        }
                        rows_grid = rows_blocks * static_cast<unsigned>(tile / E / 64);
                        rows_threads = 64;
                        rows_bytes = bytes / static_cast<std::size_t>(tile / E / 64);
                    }
                }
                hipError_t e = hipSuccess;
                if ((e = allow_lds(cols_kernel, bytes)) != hipSuccess || (e = allow_lds(rows_kernel, rows_bytes)) != hipSuccess)
                    return e;
                {
                    probe_scope ps(s, "fast bgv columns");
                    cols_kernel<<<cols_blocks, tile / E, bytes, s>>>(jobs, map_cols, ws, P);
                }
                {
                    probe_scope ps(s, "fast bgv rows");
                    rows_kernel<<<rows_grid, rows_threads, rows_bytes, s>>>(jobs, map_rows, ws, P);
                }
                return hipGetLastError();
            });
        }

        // The two-pass product over a table of jobs of any two-pass degrees (one polynomial pair
        // each), in three launches: forward columns of a and b (2 x column tiles per job), rows
        // (forward, product, inverse), inverse columns with the final scaling. Job j keeps
        // NTT(a), NTT(b) and the row results in its own workspace region [2][wsl][N] at
        // ws + ws_off. Blocks find their job and tile through map as above.
        template<typename A, int E, int S>
        __device__ __forceinline__ void polymul_cols_table_tile(job_desc const& J, u32* lds, u32* ws, u32 local,
                                                                pass_geom const& gc, typename A::params const& P)
        {
            constexpr int T = tile / E;
            u32 which = local >= gc.blocks_per_poly ? 1u : 0u;
            u32 ct = local - which * gc.blocks_per_poly;
            std::size_t N = std::size_t{1} << J.log_n;
            tile_desc td;
            io_desc ii, is;
            job_cols_tile(gc, ct, J.in_count, J.in_off, td, ii);
            job_cols_tile(gc, ct, N, 0, td, is);
            if constexpr (S > 0)
            {
                td.s = S;
                td.log_seqs = tile_log - S;
                ii.log_seqs = tile_log - S;
                is.log_seqs = tile_log - S;
            }
            tile_load<A, T>(lds, which ? J.in1 : J.in0, ii);
            __syncthreads();
            u32 x[E][A::rl];
            if constexpr (cols_expanded<A, E, S>())
            {
                // stages 0-2 and 3-5 on wave-uniform expanded twiddles, stage 6 as before
                round_expanded<A, E, false>(x, lds, td, 0, 1u, 0u, J.twfx, P);
                __syncthreads();
                round_expanded<A, E, false>(x, lds, td, 3, 1u, 0u, J.twfx, P);
                __syncthreads();
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 2, false);
            }
            else
                tile_rounds<A, E, false>(x, lds, td, 1u, 0u, J.twf, P, false, 0, false);
            tile_store<A, T, true>(lds, ws + J.ws_off + which * A::wsl * N, is, ws_fold_columns(static_cast<int>(J.log_n)), P);
        }

        // S > 0: every job has this column length; S = 0: mixed degrees, the column length 2^7
        // (N >= 2^14) at compile time too
        template<typename A, int E, int S = 0>
        __global__ void __launch_bounds__(tile / E) k_polymul_cols_table(job_desc const* __restrict__ jobs,
                                                                         u32 const* __restrict__ map, u32* __restrict__ ws,
                                                                         typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            u32 m = map[blockIdx.x];
            job_desc J = jobs[m >> 10];
            pass_geom gc = geom_for(static_cast<int>(J.log_n), false);
            if (S == 7 || (S == 0 && gc.log_n1 == 7))
                polymul_cols_table_tile<A, E, 7>(J, lds, ws, m & 1023u, gc, P);
            else
                polymul_cols_table_tile<A, E, 0>(J, lds, ws, m & 1023u, gc, P);
        }

        template<typename A, int E, int S>
        __device__ __forceinline__ void polymul_rows_table_tile(job_desc const& J, u32* lds, u32* ws, u32 rt,
                                                                typename A::params const& P)
        {
            constexpr int T = tile / E;
            constexpr int WM = shape<E>::wmax;
            pass_geom gr = geom_for(static_cast<int>(J.log_n), true);
            std::size_t N = std::size_t{1} << J.log_n;
            u32* sa = ws + J.ws_off;
            u32* sb = sa + A::wsl * N;
            tile_desc td;
            io_desc is;
            u32 hi0 = job_rows_tile(gr, rt, N, 0, td, is);
            if constexpr (S > 0)
            {
                td.s = S;
                td.log_seqs = tile_log - S;
            }
            u32 base0 = (1u << gr.log_n1) + hi0;
            // wave-local rows as in the uniform row kernel
            constexpr bool WL = A::wave_units && S > 0 && (1 << S) <= 64 * E;
            u32 x[E][A::rl];
            if constexpr (A::wave_blocks)
            {
                // one-wave rows: NTT(a) stays in registers next to NTT(b), as in the uniform kernel
                u32 xa[E][A::rl];
                tile_load<A, T, true, WL>(lds, sa, is);
                tile_sync<WL>();
                row_rounds_forward<A, E, S, WL>(xa, lds, td, base0, 1u, J.twf, J.twfx, P, true);
                tile_sync<WL>();
                tile_load<A, T, true, WL>(lds, sb, is);
                tile_sync<WL>();
                row_rounds_forward<A, E, S, WL>(x, lds, td, base0, 1u, J.twf, J.twfx, P, true);
#pragma unroll
                for (int k = 0; k < E; ++k)
                {
                    A::point(xa[k], x[k], P);
#pragma unroll
                    for (int l = 0; l < A::rl; ++l)
                        x[k][l] = xa[k][l];
                }
            }
            else
            {
                tile_load<A, T, true, WL>(lds, sa, is);
                tile_sync<WL>();
                tile_rounds<A, E, false, WL>(x, lds, td, base0, 1u, J.twf, P, false, 0, false);
                tile_store<A, T, true, WL>(lds, sa, is, false, P);
                tile_sync<WL>();
                tile_load<A, T, true, WL>(lds, sb, is);
                tile_sync<WL>();
                tile_rounds<A, E, false, WL>(x, lds, td, base0, 1u, J.twf, P, false, 0, true);
                with_operands_last<A, E, 0>(x, td, is.valid, sa, N, is.base, nullptr, 0, 0, P);
            }
            dispatch_round_regs<A, E>(round_width<WM>(td.s, 0, true), x, lds, td, base0, 1u, J.twi, P, false);
            tile_sync<WL>();
            row_rounds_inverse_rest<A, E, S, WL>(x, lds, td, base0, 1u, J.twi, J.twix, P);
            tile_store<A, T, true, WL>(lds, sa, is, false, P);
        }

        // S > 0: every job has this row length (A then the row policy); S = 0: mixed degrees, each
        // block branching (uniformly) to its own job's row length, 2^7..2^10 with the row policy
        // W1: one wave per workgroup, as k_bgv_rows_table
        template<typename A, int E, int S = 0, bool W1 = false>
        __global__ void __launch_bounds__(W1 ? 64 : tile / E, W1 ? 3 : (E == 8 ? 4 : 1)) k_polymul_rows_table(job_desc const* __restrict__ jobs,
                                                                                        u32 const* __restrict__ map,
                                                                                        u32* __restrict__ ws,
                                                                                        typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            using AR = typename table_rows_policy<A, W1>::polymul;
            u32 m = map[W1 ? blockIdx.x / static_cast<u32>(tile / E / 64) : blockIdx.x];
            job_desc J = jobs[m >> 10];
            u32 rt = m & 1023u;
            if constexpr (S > 0 && W1)
                polymul_rows_table_tile<AR, E, S>(J, lds, ws, rt, P);
            else if constexpr (S > 0)
                polymul_rows_table_tile<A, E, S>(J, lds, ws, rt, P);
            else if constexpr (W1)
            {
                switch (geom_for(static_cast<int>(J.log_n), true).log_n2)
                {
                case 7: polymul_rows_table_tile<AR, E, 7>(J, lds, ws, rt, P); break;
                case 8: polymul_rows_table_tile<AR, E, 8>(J, lds, ws, rt, P); break;
                case 9: polymul_rows_table_tile<AR, E, 9>(J, lds, ws, rt, P); break;
                default: break;
                }
            }
            else
            {
                switch (geom_for(static_cast<int>(J.log_n), true).log_n2)
                {
                case 7: polymul_rows_table_tile<AR, E, 7>(J, lds, ws, rt, P); break;
                case 8: polymul_rows_table_tile<AR, E, 8>(J, lds, ws, rt, P); break;
                case 9: polymul_rows_table_tile<AR, E, 9>(J, lds, ws, rt, P); break;
                case 10: polymul_rows_table_tile<AR, E, 10>(J, lds, ws, rt, P); break;
                default: polymul_rows_table_tile<A, E, 0>(J, lds, ws, rt, P); break;
                }
            }
        }

        template<typename A, int E, int S>
        __device__ __forceinline__ void polymul_icols_table_tile(job_desc const& J, u32* lds, u32 const* ws, u32 ct,
                                                                 pass_geom const& gc, typename A::params const& P)
        {
            constexpr int T = tile / E;
            std::size_t N = std::size_t{1} << J.log_n;
            tile_desc td;
            io_desc is, io;
            job_cols_tile(gc, ct, N, 0, td, is);
            job_cols_tile(gc, ct, J.out_count, J.out_off, td, io);
            if constexpr (S > 0)
            {
                td.s = S;
                td.log_seqs = tile_log - S;
                is.log_seqs = tile_log - S;
                io.log_seqs = tile_log - S;
            }
            tile_load<A, T, true>(lds, ws + J.ws_off, is);
            __syncthreads();
            u32 x[E][A::rl];
            if constexpr (cols_expanded<A, E, S>())
            {
                // stage 0 as before, stages 1-3 and 4-6 (with the scaled final stage) on
                // wave-uniform expanded twiddles
                tile_rounds<A, E, true>(x, lds, td, 1u, 0u, J.twi, P, true, 0, false, gc.log_n2, 1);
                round_expanded<A, E, true>(x, lds, td, 1, 1u, 0u, J.twix, P, gc.log_n2);
                __syncthreads();
                round_expanded<A, E, true>(x, lds, td, 4, 1u, 0u, J.twix, P, gc.log_n2, true);
                __syncthreads();
            }
            else
                tile_rounds<A, E, true>(x, lds, td, 1u, 0u, J.twi, P, true, 0, false, gc.log_n2);
            tile_store<A, T>(lds, J.out0, io, true, P);
        }

        // S as for the forward column kernel
        template<typename A, int E, int S = 0>
        __global__ void __launch_bounds__(tile / E, E == 8 ? 4 : 1) k_polymul_icols_table(job_desc const* __restrict__ jobs,
                                                                                          u32 const* __restrict__ map,
                                                                                          u32 const* __restrict__ ws,
                                                                                          typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            u32 m = map[blockIdx.x];
            job_desc J = jobs[m >> 10];
            pass_geom gc = geom_for(static_cast<int>(J.log_n), false);
            if (S == 7 || (S == 0 && gc.log_n1 == 7))
                polymul_icols_table_tile<A, E, 7>(J, lds, ws, m & 1023u, gc, P);
            else
                polymul_icols_table_tile<A, E, 0>(J, lds, ws, m & 1023u, gc, P);
        }

        template<typename A>
        hipError_t launch_polymul_table(int elems, int log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                        u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                        u32 icols_blocks, u32* ws, typename A::params const& P, hipStream_t s)
        {
            return with_elems<A>(elems, [&](auto EC) -> hipError_t {
                constexpr int E = decltype(EC)::value;
                constexpr std::size_t bytes = tile_bytes<A>();
                using AR = typename rows_arith<A>::type;
                pass_geom gc = geom_for(log_n, false);
                pass_geom gr = geom_for(log_n, true);
                bool const uniform = log_n > tile_log;
                bool const c7 = uniform && gc.log_n1 == 7;
                auto cols_kernel = c7 ? k_polymul_cols_table<A, E, 7> : k_polymul_cols_table<A, E, 0>;
                auto icols_kernel = c7 ? k_polymul_icols_table<A, E, 7> : k_polymul_icols_table<A, E, 0>;
                auto rows_kernel = k_polymul_rows_table<A, E, 0>;
                switch (uniform ? gr.log_n2 : 0)
                {
                case 7: rows_kernel = k_polymul_rows_table<AR, E, 7>; break;
                case 8: rows_kernel = k_polymul_rows_table<AR, E, 8>; break;
                case 9: rows_kernel = k_polymul_rows_table<AR, E, 9>; break;
                case 10: rows_kernel = k_polymul_rows_table<AR, E, 10>; break;
                default: break;
                }
                unsigned rows_grid = rows_blocks, rows_threads = tile / E;
                std::size_t rows_bytes = bytes;
                if constexpr (std::is_same<A, arith_r26>::value && E == 8)
                {
                    if (table_rows_one_wave<A, E>(max_log_n))
                    {
                        switch (uniform ? gr.log_n2 : 0)
                        {
                        case 7: rows_kernel = k_polymul_rows_table<A, E, 7, true>; break;
                        case 8: rows_kernel = k_polymul_rows_table<A, E, 8, true>; break;
                        case 9: rows_kernel = k_polymul_rows_table<A, E, 9, true>; break;
                        default: rows_kernel = k_polymul_rows_table<A, E, 0, true>; break;
                        }
                        rows_grid = rows_blocks * static_cast<unsigned>(tile / E / 64);
                        rows_threads = 64;
                        rows_bytes = bytes / static_cast<std::size_t>(tile / E / 64);
                    }
                }
                hipError_t e = hipSuccess;
                if ((e = allow_lds(cols_kernel, bytes)) != hipSuccess || (e = allow_lds(rows_kernel, rows_bytes)) != hipSuccess ||
                    (e = allow_lds(icols_kernel, bytes)) != hipSuccess)
                    return e;
                {
                    probe_scope ps(s, "fast polymul forward columns");
                    cols_kernel<<<cols_blocks, tile / E, bytes, s>>>(jobs, map_cols, ws, P);
                }
                {
                    probe_scope ps(s, "fast polymul rows");
                    rows_kernel<<<rows_grid, rows_threads, rows_bytes, s>>>(jobs, map_rows, ws, P);
                }
                {
                    probe_scope ps(s, "fast polymul inverse columns");
                    icols_kernel<<<icols_blocks, tile / E, bytes, s>>>(jobs, map_icols, ws, P);
                }
                return hipGetLastError();
            });
        }

        // one elements-per-thread variant (the Goldilocks variants compile in separate units)
        template<typename A, int E>
        hipError_t jobs_occupancy_e(int* blocks_per_cu)
        {
            constexpr std::size_t bytes = tile_bytes<A>();
            hipError_t e = allow_lds(k_jobs<A, E>, bytes);
            if (e != hipSuccess)
                return e;
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_jobs<A, E>, tile / E, bytes);
        }

        template<typename A, int E>
        hipError_t launch_jobs_e(job_desc const* table, job_desc const& base, u32 njobs, unsigned grid, u32* counter,
                                 u32* scratch, u64 slot_words, typename A::params const& P, hipStream_t s)
        {
            constexpr std::size_t bytes = tile_bytes<A>();
            hipError_t e = allow_lds(k_jobs<A, E>, bytes);
            if (e != hipSuccess)
                return e;
            k_reset_counter<A><<<1, 64, 0, s>>>(counter);
            {
                probe_scope ps(s, "fast persistent jobs");
                k_jobs<A, E><<<grid, tile / E, bytes, s>>>(table, base, njobs, counter, scratch, slot_words, P);
            }
            return hipGetLastError();
        }

        template<typename A>
        hipError_t jobs_occupancy(int elems, int* blocks_per_cu)
        {
            return with_elems<A>(elems, [&](auto EC) -> hipError_t {
                return jobs_occupancy_e<A, decltype(EC)::value>(blocks_per_cu);
            });
        }

        template<typename A>
        hipError_t launch_jobs(int elems, job_desc const* table, job_desc const& base, u32 njobs, unsigned grid,
                               u32* counter, u32* scratch, u64 slot_words, typename A::params const& P, hipStream_t s)
        {
            return with_elems<A>(elems, [&](auto EC) -> hipError_t {
                return launch_jobs_e<A, decltype(EC)::value>(table, base, njobs, grid, counter, scratch, slot_words, P, s);
            });
        }
    }
}
