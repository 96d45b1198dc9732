// Internal launcher interface between the C ABI (api.hip) and the HIP kernels (kernels.hip).
#pragma once

#include "modarith.hpp"
#include "chacha.hpp"
#include "r26.hpp"
#include "jobs.hpp"

#include <cstddef>

namespace hmpc_amd
{
    // Optional per-launch timing. While a probe is installed for the calling thread, the fast
    // path brackets each kernel launch with begin/end callbacks on the launch stream (the
    // library's timing probe records HIP events there).
    struct launch_probe
    {
        virtual void begin(hipStream_t s, char const* kernel) = 0;
        virtual void end(hipStream_t s, char const* kernel) = 0;
        virtual ~launch_probe() = default;
    };
    launch_probe*& current_probe();

    struct probe_scope
    {
        hipStream_t s;
        char const* kernel;
        probe_scope(hipStream_t stream, char const* name) : s(stream), kernel(name)
        {
            if (launch_probe* p = current_probe())
                p->begin(s, kernel);
        }
        ~probe_scope()
        {
            if (launch_probe* p = current_probe())
                p->end(s, kernel);
        }
    };

    // Everything a kernel needs to know about a modulus at one transform size.
    struct transform_args
    {
        mod_params P;
        u32 one[max_limbs];    // R mod q
        u32 r2[max_limbs];     // R^2 mod q
        int log_n;
        const u32* fwd;        // [L][N] forward table (hmpc's roots, device)
        const u32* inv;        // [L][N] inverse table
        const u32* fwd_em;     // [N][L] element-major copies for vector loads
        const u32* inv_em;
        const u32* fwd_fast;   // tables in the fast path's format (r26: [N][8] words)
        const u32* inv_fast;
        const u32* fwd_tw;     // r26: the first r26_tw_entries entries of fwd_fast / inv_fast expanded
        const u32* inv_tw;     // for r26_mont_tw (r26_tw_stride words each); null for other forms
        u32 q2[max_limbs];     // 2q
        int form;              // arith_form of the fast path
        int elems;             // fast path: elements per thread (4, 8 or 16)
        p26 r26;               // constants of the radix-2^26 form (q = 2^124 + c)
    };

    // Decomposition N = N1 * N2 used by the two-pass fused transforms.
    struct split_plan
    {
        int log_n1;  // column transforms of N1 points (first stages of the forward transform)
        int log_n2;  // row transforms of N2 points
        int cols;    // columns per workgroup in the column pass
        int rows;    // rows per workgroup in the row pass
        bool single; // whole polynomial in one workgroup's LDS
    };

    split_plan plan_for(int limbs, int log_n, int tiles_in_lds);

    hipError_t launch_to_element_major(int L, u32 const* planar, u32* em, std::size_t count, hipStream_t s);
    // [4][N] canonical table of w R -> [N][8] radix-2^26 limbs of w 2^130 mod q
    hipError_t launch_r26_table(u32 const* planar, u32* out, std::size_t count, p26 const& P, hipStream_t s);
    // entries of the expanded twiddle tables (r26_mont_tw): the twiddles of the first 10 forward
    // stages and of the inverse's last 10, which the streaming column passes (stages 0-6) and the
    // one-wave row kernel's wave-uniform rounds (stages 7-9 of N = 2^16) take per wave
    constexpr std::size_t r26_tw_entries = 1024;
    hipError_t launch_r26_tw_table(u32 const* fast, u32* out, std::size_t count, p26 const& P, hipStream_t s);

    // fast path (L = 2, 4): register-resident radix-16 rounds, see kernels_fast.hip
    bool fast_path_supported(int L, int form, int log_n);
    // the fast tile kernels of transforms and products: the special-form moduli above, and any
    // modulus of 4 or 5 limbs with 4q < R (the CIOS policy)
    bool fast_transform_supported(int L, int form, int log_n);
    hipError_t launch_fast_ntt(int L, transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                               hipStream_t s);
    std::size_t fast_polymul_workspace_words(int L, int log_n, std::size_t n);
    hipError_t launch_fast_polymul(int L, transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                                   u32* workspace, hipStream_t s);
    // persistent job kernel (jobs_impl.hpp): workgroups the device holds at once per CU, and
    // the launch (counter reset + kernel) of `njobs` jobs from `table` or the uniform `base`
    hipError_t fast_jobs_occupancy(int L, int elems, int* blocks_per_cu);
    hipError_t launch_fast_jobs(int L, transform_args const& T, job_desc const* table, job_desc const& base, u32 njobs,
                                unsigned grid, u32* counter, u32* scratch, u64 slot_words, hipStream_t s);
    // batched BGV of n ciphertexts, N > 4096, as two launches over tiles (workspace
    // [3][ws_limbs(L)][n N])
    hipError_t launch_fast_bgv_passes(int L, transform_args const& T, job_desc const& base, u64 n, u32* ws, hipStream_t s);
    // the same over a table of BGV jobs of two-pass degrees (one ciphertext each): block b runs
    // job map[b] >> 10, local tile map[b] & 1023; job j's column results live at
    // ws + jobs[j].ws_off ([3][ws_limbs(L)][N]). uniform_log_n > 0: every job has degree
    // 2^uniform_log_n (kernels with the pass lengths at compile time); 0: mixed degrees. max_log_n:
    // the largest degree of the table (up to 2^16 the radix-2^26 row launch runs one wave per
    // workgroup)
    hipError_t launch_fast_bgv_table(int L, transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                     u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32* ws, hipStream_t s);
    // the two-pass product over a table of jobs of two-pass degrees (one pair each): forward
    // columns (2 x column tiles per job), rows, inverse columns; job j's workspace region is
    // ws + jobs[j].ws_off ([2][ws_limbs(L)][N]); maps and uniform_log_n as above
    hipError_t launch_fast_polymul_table(int L, transform_args const& T, int uniform_log_n, int max_log_n, job_desc const* jobs, u32 const* map_cols,
                                         u32 cols_blocks, u32 const* map_rows, u32 rows_blocks, u32 const* map_icols,
                                         u32 icols_blocks, u32* ws, hipStream_t s);
    // tiles per polynomial of the column (rows = false) or row pass at N = 2^log_n > 4096
    unsigned two_pass_tiles(int log_n, bool rows);

    hipError_t launch_make_twiddles(int L, u32* table, int log_n, u32 const* root, u32 const* norm, u32 const* one,
                                    bool inverse, mod_params const& P, hipStream_t s);

    hipError_t launch_ntt_plain(int L, transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                u32* scratch, hipStream_t s);

    hipError_t launch_ntt_fused(int L, transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                hipStream_t s);

    hipError_t launch_polymul(int L, transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                              u32* workspace, hipStream_t s);
    std::size_t polymul_workspace_words(int L, int log_n, std::size_t n);

    hipError_t launch_pointwise(int L, mod_params const& P, int op, std::size_t count, u32 const* a, u32 const* b,
                                std::size_t b_count, u32* out, hipStream_t s);
    // both parts of a ciphertext in one launch (kernels_ext.hip): out_p = a_p op b_p with b
    // broadcast with period b_count, or out_p = a_p op c for a constant c (L words)
    hipError_t launch_ct_binary(int L, mod_params const& P, int op, std::size_t count, u32 const* a0, u32 const* a1,
                                u32 const* b0, u32 const* b1, std::size_t b_count, u32* out0, u32* out1, hipStream_t s);
    hipError_t launch_ct_scalar(int L, mod_params const& P, int op, std::size_t count, u32 const* a0, u32 const* a1,
                                u32 const* c, u32* out0, u32* out1, hipStream_t s);
    // out = a b + c
    hipError_t launch_fma(int L, mod_params const& P, std::size_t count, u32 const* a, u32 const* b, u32 const* c, u32* out,
                          hipStream_t s);
    // full reduction (op 0: sum, 1: product; one = R mod q) of `count` elements to out (L words);
    // partials holds reduce_parts(count) x L words
    unsigned reduce_parts(std::size_t count);
    hipError_t launch_reduce(int L, mod_params const& P, u32 const* one, int op, std::size_t count, u32 const* in,
                             u32* partials, u32* out, hipStream_t s);
    // out[b][r] = sum_c M[b m_stride + r cols + c] v[b v_stride + c] (strides 0 broadcast); r26: the
    // radix-2^26 constants when q = 2^124 + c (4 limbs), else null
    hipError_t launch_matvec(int L, mod_params const& P, p26 const* r26, std::size_t batch, std::size_t rows,
                             std::size_t cols, u32 const* M, std::size_t m_count, std::size_t m_stride, u32 const* v,
                             std::size_t v_count, std::size_t v_stride, u32* out, hipStream_t s);
    // tensor cipher: limb-planar [L][count] <-> element-major words [count][L] XOR the ChaCha20
    // keystream (key, nonce, first block counter0)
    hipError_t launch_cipher(int L, bool encrypt, std::size_t count, u32 const* key8, u32 const* nonce3, u32 counter0,
                             u32 const* in, u32* out, hipStream_t s);
    hipError_t launch_pointwise_scalar(int L, mod_params const& P, int op, std::size_t count, u32 const* a,
                                       u32 const* c, u32* out, hipStream_t s);
    hipError_t launch_from_u32(int L, mod_params const& P, u32 const* r2, std::size_t count, u32 const* words, u32* out,
                               hipStream_t s);
    hipError_t launch_to_unsigned(int L, mod_params const& P, std::size_t count, u32 const* in, u32* out, hipStream_t s);
    hipError_t launch_centered_low32(int L, mod_params const& P, std::size_t count, u32 const* in, u32* out,
                                     hipStream_t s);

    hipError_t launch_sample_uniform(int L, mod_params const& P, int bit_size, chacha_key const& key, u64 offset,
                                     std::size_t count, u32* out, hipStream_t s);
    hipError_t launch_sample_cbd(int L, mod_params const& P, u32 const* r2, chacha_key const& key, u64 offset, int bits,
                                 std::size_t count, u32* out, hipStream_t s);
    hipError_t launch_sample_words(chacha_key const& key, u64 offset, std::size_t count, u32* out, hipStream_t s);

    // Capability-path CBD fused with the forward transform's first stages.
    hipError_t launch_sample_cbd_ntt(int L, transform_args const& T, chacha_key const& key, u32 const* nonce, int bits,
                                     std::size_t n, u32* out, hipStream_t s);

    struct bgv_args
    {
        chacha_key key;
        u32 nonce_u[3], nonce_v[3], nonce_w[3];
        int bits_u, bits_v, bits_w;
        u32 p_mont[max_limbs];
    };
    std::size_t bgv_workspace_words(int L, int log_n, std::size_t n);
    hipError_t launch_bgv_encrypt(int L, transform_args const& T, bgv_args const& B, std::size_t n, u32 const* pk_a,
                                  u32 const* pk_b, u32 const* messages, u32* c0, u32* c1, u32* workspace,
                                  hipStream_t s);
    // per-ciphertext plain path pieces
    hipError_t launch_sample_cbd_capability(int L, mod_params const& P, u32 const* r2, chacha_key const& key,
                                            u32 const* nonce, int bits, int log_n, std::size_t n, u32* out,
                                            std::size_t first_poly, hipStream_t s);

    // ---------------------------------------------------------------- general plaintext path (kernels_lhe.hip)
    // source modulus of a cross-modulus conversion (q, floor(q / 2), N' padded with zero limbs)
    struct src_modulus
    {
        u32 q[max_limbs];
        u32 half[max_limbs];
        u32 nprime;
        int limbs;
    };
    // out (target modulus, L limbs, constants r2 = R^2 mod q and c32 = 2^32 R mod q) = centred lift of
    // in (source modulus S), limb-planar, `count` elements
    hipError_t launch_cast(src_modulus const& S, int L, mod_params const& P, u32 const* r2, u32 const* c32, std::size_t count,
                           u32 const* in, u32* out, hipStream_t s);

    enum cap_distribution : int
    {
        cap_uniform = 0,
        cap_centered_binomial = 1,
        cap_binomial = 2,
        cap_drown_signed = 3,
        cap_drown_unsigned = 4,
    };
    enum cap_layout : int
    {
        cap_layout_element = 0,  // one stream per element (element-wise kernels)
        cap_layout_ntt = 1,      // stream of index b N + t: coefficients t, t + N/2 (forward stage 0)
        cap_layout_intt = 2,     // stream of index b N + 2t: coefficients 2t, 2t + 1 (inverse stage 0)
    };
    constexpr int cap_max_words = 12;  // widest draw: 384 bits
    struct cap_args
    {
        chacha_key key;
        u32 nonce[3];
        int dist;
        int bits;           // centred binomial: 4 x variance (<= 64); binomial: count (<= 256); drowning: width
        int layout;
        int log_n;          // pair layouts: polynomials of 2^log_n coefficients
        u64 total;          // elements
        int uniform_words;  // uniform<mod q>: ceil((bit_size + 80) / 32) words of bit_size + 80 bits
        int uniform_bits;
    };
    // c32 = 2^32 R mod q, kuni = 2^(-32 uniform_words) mod q (plain)
    hipError_t launch_sample_capability(int L, mod_params const& P, u32 const* r2, u32 const* c32, u32 const* kuni,
                                        cap_args const& C, u32* out, hipStream_t s);
    // c0 = A u + p v + M, c1 = B u + p w with A, B broadcast over the polynomials (period N)
    hipError_t launch_bgv_combine(int L, mod_params const& P, std::size_t count, std::size_t N, u32 const* pk_a, u32 const* pk_b,
                                  u32 const* U, u32 const* V, u32 const* W, u32 const* M, u32 const* p_const, u32* c0, u32* c1,
                                  hipStream_t s);
}
