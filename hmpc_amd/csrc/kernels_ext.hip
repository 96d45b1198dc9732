// Element-wise kernels around the hot path, gfx950:
//   - homomorphic operations on BGV ciphertexts (c0 and c1 in one launch);
//   - the ChaCha20 tensor cipher (keystream XOR over the limbs of a tensor);
//   - fused multiply-add, full reductions (sum, product) and matrix-vector products over mod q.
// All are HBM-bound streaming kernels (the cipher: ChaCha-bound at one block per 16 words).
#include "kernels.hpp"
#include "chacha.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hmpc_amd
{
    namespace
    {
        constexpr int ext_threads = 256;

        unsigned ext_grid(std::size_t work)
        {
            return static_cast<unsigned>((work + ext_threads - 1) / ext_threads);
        }

        template<typename F>
        hipError_t ext_limbs(int L, F&& f)
        {
            switch (L)
            {
            case 1: return f(std::integral_constant<int, 1>{});
            case 2: return f(std::integral_constant<int, 2>{});
            case 3: return f(std::integral_constant<int, 3>{});
            case 4: return f(std::integral_constant<int, 4>{});
            case 5: return f(std::integral_constant<int, 5>{});
            case 6: return f(std::integral_constant<int, 6>{});
            case 7: return f(std::integral_constant<int, 7>{});
            case 8: return f(std::integral_constant<int, 8>{});
            case 9: return f(std::integral_constant<int, 9>{});
            default: return hipErrorInvalidValue;
            }
        }

        template<int L>
        __device__ __forceinline__ void ct_apply(int op, u32 (&r)[L], u32 const (&x)[L], u32 const (&y)[L],
                                                 mod_params const& P)
        {
            if (op == 0)
                mod_add<L>(r, x, y, P);
            else if (op == 1)
                mod_sub<L>(r, x, y, P);
            else
                mod_mul<L>(r, x, y, P);
        }

        // blockIdx.y selects the ciphertext part: out_p = a_p op b_p, b broadcast with period b_count
        template<int L>
        __global__ void __launch_bounds__(ext_threads) k_ct_binary(int op, u64 count, u32 const* __restrict__ a0,
                                                                   u32 const* __restrict__ a1, u32 const* __restrict__ b0,
                                                                   u32 const* __restrict__ b1, u64 b_count,
                                                                   u32* __restrict__ out0, u32* __restrict__ out1,
                                                                   mod_params P)
        {
            u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
            if (e >= count)
                return;
            bool second = blockIdx.y != 0;
            u64 eb = b_count == count ? e : e % b_count;
            u32 x[L], y[L], r[L];
            load_planar<L>(x, second ? a1 : a0, count, e);
            load_planar<L>(y, second ? b1 : b0, b_count, eb);
            ct_apply<L>(op, r, x, y, P);
            store_planar<L>(second ? out1 : out0, count, e, r);
        }

        template<int L>
        __global__ void __launch_bounds__(ext_threads) k_ct_scalar(int op, u64 count, u32 const* __restrict__ a0,
                                                                   u32 const* __restrict__ a1, residue<L> c,
                                                                   u32* __restrict__ out0, u32* __restrict__ out1,
                                                                   mod_params P)
        {
            u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
            if (e >= count)
                return;
            bool second = blockIdx.y != 0;
            u32 x[L], r[L];
            load_planar<L>(x, second ? a1 : a0, count, e);
            ct_apply<L>(op, r, x, c.v, P);
            store_planar<L>(second ? out1 : out0, count, e, r);
        }

        // Tensor cipher, one ChaCha20 block per thread: block t covers keystream words
        // [16 t, 16 t + 16), and keystream word w masks limb w % L of element w / L
        // (hmpc/expr/crypto/cipher.hpp:124-160: element i takes words [i L, i L + L) of one
        // stream with counter = block index). Encryption reads the limb-planar tensor and writes
        // element-major words; decryption the reverse.
        template<int L, bool Enc>
        __global__ void __launch_bounds__(ext_threads) k_cipher(u64 count, u32 const* __restrict__ in,
                                                                u32* __restrict__ out, chacha_key key, u32 counter0,
                                                                u32 n0, u32 n1, u32 n2, bool aligned)
        {
            u64 t = u64{blockIdx.x} * blockDim.x + threadIdx.x;
            u64 words = count * L;
            u64 w0 = t * 16;
            if (w0 >= words)
                return;
            u32 ks[16];
            chacha20_block(ks, key, counter0 + static_cast<u32>(t), n0, n1, n2);
            if constexpr (L == 1 || L == 2 || L == 4)
            {
                // a whole block of whole elements (16 / L of them, a multiple of 4): 16-byte loads
                // per limb plane and four 16-byte stores of the element-major words; the planes and
                // the output are 16-byte aligned when the buffers are and count is a multiple of 4
                constexpr int per = 16 / L;
                u64 e0 = t * per;
                if (aligned && e0 + per <= count && count % 4 == 0)
                {
                    u32 v[16];
                    if constexpr (Enc)
                    {
#pragma unroll
                        for (int j = 0; j < L; ++j)
#pragma unroll
                            for (int q = 0; q < per; q += 4)
                            {
                                uint4 x = *reinterpret_cast<uint4 const*>(in + j * count + e0 + q);
                                v[(q + 0) * L + j] = x.x;
                                v[(q + 1) * L + j] = x.y;
                                v[(q + 2) * L + j] = x.z;
                                v[(q + 3) * L + j] = x.w;
                            }
#pragma unroll
                        for (int k = 0; k < 16; k += 4)
                            *reinterpret_cast<uint4*>(out + w0 + k) =
                                make_uint4(v[k] ^ ks[k], v[k + 1] ^ ks[k + 1], v[k + 2] ^ ks[k + 2], v[k + 3] ^ ks[k + 3]);
                    }
                    else
                    {
#pragma unroll
                        for (int k = 0; k < 16; k += 4)
                        {
                            uint4 x = *reinterpret_cast<uint4 const*>(in + w0 + k);
                            v[k] = x.x ^ ks[k];
                            v[k + 1] = x.y ^ ks[k + 1];
                            v[k + 2] = x.z ^ ks[k + 2];
                            v[k + 3] = x.w ^ ks[k + 3];
                        }
#pragma unroll
                        for (int j = 0; j < L; ++j)
#pragma unroll
                            for (int q = 0; q < per; q += 4)
                                *reinterpret_cast<uint4*>(out + j * count + e0 + q) =
                                    make_uint4(v[(q + 0) * L + j], v[(q + 1) * L + j], v[(q + 2) * L + j], v[(q + 3) * L + j]);
                    }
                    return;
                }
            }
#pragma unroll
            for (int k = 0; k < 16; ++k)
            {
                u64 w = w0 + k;
                if (w >= words)
                    break;
                u64 i = w / L;
                u64 j = w - i * L;
                if constexpr (Enc)
                    out[w] = in[j * count + i] ^ ks[k];
                else
                    out[j * count + i] = in[w] ^ ks[k];
            }
        }
    }

    namespace
    {
        // out = a b + c (examples/fma.cpp:56-67: x[i] = a[i] * b[i] + c[i])
        template<int L>
        __global__ void __launch_bounds__(ext_threads) k_fma(u64 count, u32 const* __restrict__ a, u32 const* __restrict__ b,
                                                             u32 const* __restrict__ c, u32* __restrict__ out, mod_params P)
        {
            u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
            if (e >= count)
                return;
            u32 x[L], y[L], z[L], t[L], r[L];
            load_planar<L>(x, a, count, e);
            load_planar<L>(y, b, count, e);
            load_planar<L>(z, c, count, e);
            mod_mul<L>(t, x, y, P);
            mod_add<L>(r, t, z, P);
            store_planar<L>(out, count, e, r);
        }

        template<int L>
        __device__ __forceinline__ void combine(int op, u32 (&acc)[L], u32 const (&v)[L], mod_params const& P)
        {
            u32 r[L];
            if (op == 0)
                mod_add<L>(r, acc, v, P);
            else
                mod_mul<L>(r, acc, v, P);
#pragma unroll
            for (int l = 0; l < L; ++l)
                acc[l] = r[l];
        }

        template<int L>
        __device__ __forceinline__ void identity(int op, u32 (&acc)[L], u32 const (&one)[L])
        {
#pragma unroll
            for (int l = 0; l < L; ++l)
                acc[l] = op == 0 ? 0u : one[l];
        }

        // workgroup tree reduction through LDS (ext_threads x L words); result in thread 0
        template<int L>
        __device__ __forceinline__ void block_reduce(int op, u32 (&acc)[L], u32* lds, mod_params const& P)
        {
            int t = threadIdx.x;
#pragma unroll
            for (int l = 0; l < L; ++l)
                lds[l * ext_threads + t] = acc[l];
            __syncthreads();
            for (int h = ext_threads / 2; h > 0; h >>= 1)
            {
                if (t < h)
                {
                    u32 v[L];
#pragma unroll
                    for (int l = 0; l < L; ++l)
                        v[l] = lds[l * ext_threads + t + h];
                    combine<L>(op, acc, v, P);
#pragma unroll
                    for (int l = 0; l < L; ++l)
                        lds[l * ext_threads + t] = acc[l];
                }
                __syncthreads();
            }
        }

        // pass 1: grid-stride partial reductions, one partial per workgroup into partials[L][grid];
        // pass 2 (one workgroup): the partials to out (L words). Sums and products mod q are
        // associative and commutative, so the result does not depend on the order.
        template<int L>
        __global__ void __launch_bounds__(ext_threads) k_reduce(int op, u64 count, u32 const* __restrict__ in,
                                                                u32* __restrict__ partials, residue<L> one, mod_params P)
        {
            __shared__ u32 lds[L * ext_threads];
            u32 acc[L];
            identity<L>(op, acc, one.v);
            for (u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x; e < count; e += u64{gridDim.x} * blockDim.x)
            {
                u32 v[L];
                load_planar<L>(v, in, count, e);
                combine<L>(op, acc, v, P);
            }
            block_reduce<L>(op, acc, lds, P);
            if (threadIdx.x == 0)
                store_planar<L>(partials, gridDim.x, blockIdx.x, acc);
        }

        template<int L>
        __global__ void __launch_bounds__(ext_threads) k_reduce_final(int op, u32 parts, u32 const* __restrict__ partials,
                                                                      u32* __restrict__ out, residue<L> one, mod_params P)
        {
            __shared__ u32 lds[L * ext_threads];
            u32 acc[L];
            identity<L>(op, acc, one.v);
            for (u32 e = threadIdx.x; e < parts; e += ext_threads)
            {
                u32 v[L];
                load_planar<L>(v, partials, parts, e);
                combine<L>(op, acc, v, P);
            }
            block_reduce<L>(op, acc, lds, P);
            if (threadIdx.x == 0)
                store_planar<L>(out, 1, 0, acc);
        }

        // Arithmetic of the matrix-vector products, as a policy: accumulator words W, a product
        // added into an accumulator, the sum of two accumulators, `settle` (called at least every
        // 16 products into one accumulator and before `merge`) and the canonical result.
        //   mv_mont<L>: hmpc's Montgomery form for any odd q: each product reduced, each sum mod q
        //   mv_r26:     q = 2^124 + c (4 limbs): radix-2^26 products (r26.hpp) summed without
        //               reductions; a fold every 16 products keeps the limbs below 2^31
        template<int L>
        struct mv_mont
        {
            static constexpr int W = L;
            using params = mod_params;
            __device__ __forceinline__ static void product_add(u32 (&acc)[W], u32 const (&x)[L], u32 const (&y)[L],
                                                               params const& P)
            {
                u32 t[L];
                mod_mul<L>(t, x, y, P);
                combine<L>(0, acc, t, P);
            }
            __device__ __forceinline__ static void settle(u32 (&)[W], params const&) {}
            __device__ __forceinline__ static void merge(u32 (&acc)[W], u32 const (&o)[W], params const& P)
            {
                combine<L>(0, acc, o, P);
            }
            __device__ __forceinline__ static void finish(u32 (&out)[L], u32 const (&acc)[W], params const&)
            {
#pragma unroll
                for (int l = 0; l < L; ++l)
                    out[l] = acc[l];
            }
        };

        // Bounds: canonical inputs unpack to limbs < 2^26 (x_4 < 2^21), so each r26_mont product is
        // < 2q (1 + 2^-19) with limbs < 2^26 (limb 4 < 2^22). A folded accumulator (< 2q, limbs
        // < 2^28) plus 16 products keeps limbs < 2^28 + 2^30 < 2^31, the fold's input bound. The sum
        // S of products a R b R 2^-130 is a b R / 4 summed: the result is 4 S, folded and made
        // canonical.
        struct mv_r26
        {
            static constexpr int W = 5;
            using params = p26;
            __device__ __forceinline__ static void product_add(u32 (&acc)[W], u32 const (&x)[4], u32 const (&y)[4],
                                                               params const& P)
            {
                u32 a[5], b[5], t[5];
                r26_unpack(a, x[0], x[1], x[2], x[3]);
                r26_unpack(b, y[0], y[1], y[2], y[3]);
                r26_mont(t, a, b, P);
#pragma unroll
                for (int l = 0; l < 5; ++l)
                    acc[l] += t[l];
            }
            __device__ __forceinline__ static void settle(u32 (&acc)[W], params const& P) { r26_fold(acc, P); }
            // both settled: limbs < 2^29, folded again
            __device__ __forceinline__ static void merge(u32 (&acc)[W], u32 const (&o)[W], params const& P)
            {
#pragma unroll
                for (int l = 0; l < 5; ++l)
                    acc[l] += o[l];
                r26_fold(acc, P);
            }
            __device__ __forceinline__ static void finish(u32 (&out)[4], u32 const (&acc)[W], params const& P)
            {
                u32 x[5];
#pragma unroll
                for (int l = 0; l < 5; ++l)
                    x[l] = acc[l] << 2;  // settled limbs < 2^28
                r26_fold(x, P);
                r26_canonical(out, x, P);
            }
        };

        // workgroup tree of settled accumulators through LDS; result in thread 0
        template<typename V>
        __device__ __forceinline__ void block_merge(u32 (&acc)[V::W], u32* lds, typename V::params const& P)
        {
            constexpr int W = V::W;
            int t = threadIdx.x;
#pragma unroll
            for (int l = 0; l < W; ++l)
                lds[l * ext_threads + t] = acc[l];
            __syncthreads();
            for (int h = ext_threads / 2; h > 0; h >>= 1)
            {
                if (t < h)
                {
                    u32 v[W];
#pragma unroll
                    for (int l = 0; l < W; ++l)
                        v[l] = lds[l * ext_threads + t + h];
                    V::merge(acc, v, P);
#pragma unroll
                    for (int l = 0; l < W; ++l)
                        lds[l * ext_threads + t] = acc[l];
                }
                __syncthreads();
            }
        }

        // out[b][r] = sum_c M[b][r][c] v[b][c] (hmpc/expr/matrix_vector_product.hpp:92-133), one wave
        // per output row: lanes stride over the columns (coalesced per limb plane), then a
        // cross-lane tree. Matrix and vector strides of 0 broadcast them over the batch.
        template<int L, typename V>
        __global__ void __launch_bounds__(ext_threads) k_matvec(u64 batch, u64 rows, u64 cols, u32 const* __restrict__ M,
                                                                u64 m_count, u64 m_stride, u32 const* __restrict__ v,
                                                                u64 v_count, u64 v_stride, u32* __restrict__ out,
                                                                typename V::params P)
        {
            constexpr int waves = ext_threads / 64;
            constexpr int W = V::W;
            u64 row = u64{blockIdx.x} * waves + threadIdx.x / 64;
            int lane = threadIdx.x & 63;
            if (row >= batch * rows)
                return;
            u64 b = row / rows;
            u64 r = row - b * rows;
            u64 m0 = b * m_stride + r * cols;
            u64 v0 = b * v_stride;
            // four columns in flight per lane (independent loads and two accumulators, 2 products
            // each per iteration: settled every 8 iterations)
            u32 acc[W] = {};
            u32 acc2[W] = {};
            u64 c = lane;
            for (int it = 1; c + 192 < cols; c += 256, ++it)
            {
                u32 x[4][L], y[4][L];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                {
                    load_planar<L>(x[u], M, m_count, m0 + c + 64 * u);
                    load_planar<L>(y[u], v, v_count, v0 + c + 64 * u);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    V::product_add((u & 1) ? acc2 : acc, x[u], y[u], P);
                if ((it & 7) == 0)
                {
                    V::settle(acc, P);
                    V::settle(acc2, P);
                }
            }
            V::settle(acc, P);
            V::settle(acc2, P);
            // at most 3 more products per lane before the next settle
            for (; c < cols; c += 64)
            {
                u32 x[L], y[L];
                load_planar<L>(x, M, m_count, m0 + c);
                load_planar<L>(y, v, v_count, v0 + c);
                V::product_add(acc, x, y, P);
            }
            V::settle(acc, P);
            V::merge(acc, acc2, P);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1)
            {
                u32 o[W];
#pragma unroll
                for (int l = 0; l < W; ++l)
                    o[l] = static_cast<u32>(__shfl_xor(static_cast<int>(acc[l]), off, 64));
                V::merge(acc, o, P);
            }
            if (lane == 0)
            {
                u32 res[L];
                V::finish(res, acc, P);
                store_planar<L>(out, batch * rows, row, res);
            }
        }
    }

    namespace
    {
        // long rows: one workgroup per output row, the columns strided over all its threads (more
        // loads in flight per row), then the workgroup tree
        template<int L, typename V>
        __global__ void __launch_bounds__(ext_threads) k_matvec_wide(u64 rows, u64 cols, u32 const* __restrict__ M,
                                                                     u64 m_count, u64 m_stride, u32 const* __restrict__ v,
                                                                     u64 v_count, u64 v_stride, u32* __restrict__ out,
                                                                     u64 out_count, typename V::params P)
        {
            constexpr int W = V::W;
            __shared__ u32 lds[W * ext_threads];
            u64 row = blockIdx.x;
            u64 b = row / rows;
            u64 r = row - b * rows;
            u64 m0 = b * m_stride + r * cols;
            u64 v0 = b * v_stride;
            u32 acc[W] = {};
            u32 acc2[W] = {};
            u64 c = threadIdx.x;
            constexpr u64 T = ext_threads;
            for (int it = 1; c + 3 * T < cols; c += 4 * T, ++it)
            {
                u32 x[4][L], y[4][L];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                {
                    load_planar<L>(x[u], M, m_count, m0 + c + T * u);
                    load_planar<L>(y[u], v, v_count, v0 + c + T * u);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    V::product_add((u & 1) ? acc2 : acc, x[u], y[u], P);
                if ((it & 7) == 0)
                {
                    V::settle(acc, P);
                    V::settle(acc2, P);
                }
            }
            V::settle(acc, P);
            V::settle(acc2, P);
            for (; c < cols; c += T)
            {
                u32 x[L], y[L];
                load_planar<L>(x, M, m_count, m0 + c);
                load_planar<L>(y, v, v_count, v0 + c);
                V::product_add(acc, x, y, P);
            }
            V::settle(acc, P);
            V::merge(acc, acc2, P);
            block_merge<V>(acc, lds, P);
            if (threadIdx.x == 0)
            {
                u32 res[L];
                V::finish(res, acc, P);
                store_planar<L>(out, out_count, row, res);
            }
        }

        template<int L, typename V>
        hipError_t matvec_with(typename V::params const& P, std::size_t batch, std::size_t rows, std::size_t cols,
                               u32 const* M, std::size_t m_count, std::size_t m_stride, u32 const* v, std::size_t v_count,
                               std::size_t v_stride, u32* out, hipStream_t s)
        {
            if (cols >= 1024)
            {
                k_matvec_wide<L, V><<<static_cast<unsigned>(batch * rows), ext_threads, 0, s>>>(
                    rows, cols, M, m_count, m_stride, v, v_count, v_stride, out, batch * rows, P);
                return hipGetLastError();
            }
            std::size_t waves = ext_threads / 64;
            unsigned grid = static_cast<unsigned>((batch * rows + waves - 1) / waves);
            k_matvec<L, V><<<grid, ext_threads, 0, s>>>(batch, rows, cols, M, m_count, m_stride, v, v_count, v_stride, out, P);
            return hipGetLastError();
        }
    }

    hipError_t launch_fma(int L, mod_params const& P, std::size_t count, u32 const* a, u32 const* b, u32 const* c, u32* out,
                          hipStream_t s)
    {
        return ext_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_fma<LL><<<ext_grid(count), ext_threads, 0, s>>>(count, a, b, c, out, P);
            return hipGetLastError();
        });
    }

    unsigned reduce_parts(std::size_t count)
    {
        std::size_t g = (count + ext_threads - 1) / ext_threads;
        return static_cast<unsigned>(g < 1 ? 1 : (g > 1024 ? 1024 : g));
    }

    hipError_t launch_reduce(int L, mod_params const& P, u32 const* one, int op, std::size_t count, u32 const* in,
                             u32* partials, u32* out, hipStream_t s)
    {
        return ext_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            residue<LL> o;
            for (int l = 0; l < LL; ++l)
                o.v[l] = one[l];
            unsigned parts = reduce_parts(count);
            k_reduce<LL><<<parts, ext_threads, 0, s>>>(op, count, in, partials, o, P);
            k_reduce_final<LL><<<1, ext_threads, 0, s>>>(op, parts, partials, out, o, P);
            return hipGetLastError();
        });
    }

    hipError_t launch_matvec(int L, mod_params const& P, p26 const* r26, std::size_t batch, std::size_t rows,
                             std::size_t cols, u32 const* M, std::size_t m_count, std::size_t m_stride, u32 const* v,
                             std::size_t v_count, std::size_t v_stride, u32* out, hipStream_t s)
    {
        if (r26 && L == 4)
            return matvec_with<4, mv_r26>(*r26, batch, rows, cols, M, m_count, m_stride, v, v_count, v_stride, out, s);
        return ext_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            return matvec_with<LL, mv_mont<LL>>(P, batch, rows, cols, M, m_count, m_stride, v, v_count, v_stride, out, s);
        });
    }

    hipError_t launch_ct_binary(int L, mod_params const& P, int op, std::size_t count, u32 const* a0, u32 const* a1,
                                u32 const* b0, u32 const* b1, std::size_t b_count, u32* out0, u32* out1, hipStream_t s)
    {
        return ext_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            dim3 grid(ext_grid(count), 2);
            k_ct_binary<LL><<<grid, ext_threads, 0, s>>>(op, count, a0, a1, b0, b1, b_count, out0, out1, P);
            return hipGetLastError();
        });
    }

    hipError_t launch_ct_scalar(int L, mod_params const& P, int op, std::size_t count, u32 const* a0, u32 const* a1,
                                u32 const* c, u32* out0, u32* out1, hipStream_t s)
    {
        return ext_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            residue<LL> cv;
            for (int l = 0; l < LL; ++l)
                cv.v[l] = c[l];
            dim3 grid(ext_grid(count), 2);
            k_ct_scalar<LL><<<grid, ext_threads, 0, s>>>(op, count, a0, a1, cv, out0, out1, P);
            return hipGetLastError();
        });
    }

    hipError_t launch_cipher(int L, bool encrypt, std::size_t count, u32 const* key8, u32 const* nonce3, u32 counter0,
                             u32 const* in, u32* out, hipStream_t s)
    {
        return ext_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            chacha_key key;
            for (int i = 0; i < 8; ++i)
                key.k[i] = key8[i];
            std::size_t blocks = (count * LL + 15) / 16;
            bool aligned = ((reinterpret_cast<std::uintptr_t>(in) | reinterpret_cast<std::uintptr_t>(out)) & 15u) == 0;
            if (encrypt)
                k_cipher<LL, true><<<ext_grid(blocks), ext_threads, 0, s>>>(count, in, out, key, counter0, nonce3[0],
                                                                            nonce3[1], nonce3[2], aligned);
            else
                k_cipher<LL, false><<<ext_grid(blocks), ext_threads, 0, s>>>(count, in, out, key, counter0, nonce3[0],
                                                                             nonce3[1], nonce3[2], aligned);
            return hipGetLastError();
        });
    }
}
