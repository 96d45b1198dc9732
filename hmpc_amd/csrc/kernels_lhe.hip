// Kernels of hmpc's general plaintext path for linear BGV (gfx950):
//  - cross-modulus conversion with the centred lift (hmpc/ints/mod.hpp:259-307, used by
//    hmpc::expr::cast in hmpc/expr/crypto/lhe/enc.hpp:43-57 and dec.hpp:33-44);
//  - the capability-path samplers in every layout hmpc's kernels consume them in (element-wise,
//    forward-NTT pairs, inverse-NTT pairs) and every distribution of hmpc/expr/random;
//  - the element-wise combination of an encryption's transformed parts,
//    c0 = A u + p v + M and c1 = B u + p w (hmpc/expr/crypto/lhe/enc.hpp:65-68).
// These are element-wise, memory- or ChaCha-bound kernels on the generic 32-bit-limb arithmetic
// (any odd modulus of 1..9 limbs); the transforms between them are the library's NTT kernels.
#include "kernels.hpp"

namespace hmpc_amd
{
    namespace
    {
        constexpr int block_threads = 256;

        template<typename F>
        hipError_t with_limbs(int L, F&& f)
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

        unsigned grid_for(std::size_t work)
        {
            return static_cast<unsigned>((work + block_threads - 1) / block_threads);
        }

        template<int L>
        residue<L> as_residue(u32 const* x)
        {
            residue<L> r;
            for (int l = 0; l < L; ++l)
                r.v[l] = x[l];
            return r;
        }

        // (value mod q) R mod q of an integer held in `words` 32-bit words (w[0] least significant).
        // At most L words: one Montgomery product with R^2 (the integer is below R). Wider: Horner
        // from the top word, r = r 2^32 + w_i, where mont(r, c32) = r 2^32 for c32 = 2^32 R mod q.
        template<int L>
        __device__ void mont_from_words(u32 (&r)[L], u32 const* w, int words, residue<L> const& r2, residue<L> const& c32,
                                        mod_params const& P)
        {
            u32 x[L];
            if (words <= L)
            {
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = l < words ? w[l] : 0u;
                mod_mul<L>(r, x, r2.v, P);
                return;
            }
#pragma unroll
            for (int l = 0; l < L; ++l)
                r[l] = 0u;
            for (int i = words - 1; i >= 0; --i)
            {
                u32 t[L], d[L];
                mod_mul<L>(t, r, c32.v, P);
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = l == 0 ? w[i] : 0u;
                mod_mul<L>(d, x, r2.v, P);
                mod_add<L>(r, t, d, P);
            }
        }

        template<int L>
        __device__ __forceinline__ void mod_negate(u32 (&r)[L], mod_params const& P)
        {
            u32 z[L];
#pragma unroll
            for (int l = 0; l < L; ++l)
                z[l] = 0u;
            mod_sub<L>(r, z, r, P);
        }
    }

    // ------------------------------------------------------------------ cross-modulus conversion
    // Element e: the canonical integer v of the source value (REDC with the source's limb count,
    // a run-time 1..9), the centred value s = v - P for v > floor(P / 2) (else v), and the result
    // (s mod Q) R_Q mod Q. This equals hmpc's construction (offset Q - P when Q > P, offset_to<P>
    // otherwise, then the unsigned constructor): both give the one canonical residue of s.
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_cast(u64 count, u32 const* __restrict__ in, src_modulus S,
                                                            u32* __restrict__ out, residue<L> r2, residue<L> c32,
                                                            mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        int const Ls = S.limbs;
        u32 acc[max_limbs + 1];
#pragma unroll
        for (int l = 0; l < max_limbs; ++l)
            acc[l] = l < Ls ? in[l * count + e] : 0u;
        acc[max_limbs] = 0u;
        // REDC with Ls word rounds: acc = x 2^(-32 Ls) mod P, below 2P
#pragma unroll
        for (int round = 0; round < max_limbs; ++round)
        {
            if (round < Ls)
            {
                u32 m = acc[0] * S.nprime;
                u64 c = (static_cast<u64>(m) * S.q[0] + acc[0]) >> 32;
#pragma unroll
                for (int j = 1; j <= max_limbs; ++j)
                {
                    u64 s = (j < max_limbs ? static_cast<u64>(m) * S.q[j] : u64{0}) + acc[j] + c;
                    acc[j - 1] = static_cast<u32>(s);
                    c = s >> 32;
                }
                acc[max_limbs] = static_cast<u32>(c);
            }
        }
        // canonical v: subtract P once when acc >= P
        u32 d[max_limbs + 1];
        u32 br = 0;
#pragma unroll
        for (int l = 0; l <= max_limbs; ++l)
            d[l] = subb(acc[l], l < max_limbs ? S.q[l] : 0u, br, br);
        bool take = br == 0;
        u32 v[max_limbs];
#pragma unroll
        for (int l = 0; l < max_limbs; ++l)
            v[l] = take ? d[l] : acc[l];
        // negative iff v > floor(P / 2)
        br = 0;
#pragma unroll
        for (int l = 0; l < max_limbs; ++l)
            (void)subb(S.half[l], v[l], br, br);
        bool negative = br != 0;
        if (negative)
        {
            br = 0;
#pragma unroll
            for (int l = 0; l < max_limbs; ++l)
                v[l] = subb(S.q[l], v[l], br, br);
        }
        u32 r[L];
        mont_from_words<L>(r, v, Ls, r2, c32, P);
        if (negative)
            mod_negate<L>(r, P);
        store_planar<L>(out, count, e, r);
    }

    hipError_t launch_cast(src_modulus const& S, int L, mod_params const& P, u32 const* r2, u32 const* c32, std::size_t count,
                           u32 const* in, u32* out, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_cast<LL><<<grid_for(count), block_threads, 0, s>>>(count, in, S, out, as_residue<LL>(r2), as_residue<LL>(c32), P);
            return hipGetLastError();
        });
    }

    // ------------------------------------------------------------------ capability samplers
    namespace
    {
        // hmpc's number_generator on the device: the words of the ChaCha20 stream (key, nonce,
        // counter), block after block (hmpc/random/number_generator.hpp:41-91)
        struct dev_stream
        {
            chacha_key const* key;
            u32 n0, n1, n2, counter;
            u32 blk[16];
            int idx;

            __device__ dev_stream(chacha_key const* k, u32 const* nonce, u32 c)
                : key(k), n0(nonce[0]), n1(nonce[1]), n2(nonce[2]), counter(c), idx(16)
            {
            }

            __device__ u32 next()
            {
                if (idx == 16)
                {
                    chacha20_block(blk, *key, counter, n0, n1, n2);
                    ++counter;
                    idx = 0;
                }
                return blk[idx++];
            }
        };

        // one draw in Montgomery form
        template<int L>
        __device__ void draw(u32 (&r)[L], dev_stream& g, cap_args const& C, residue<L> const& r2, residue<L> const& c32,
                             residue<L> const& kuni, mod_params const& P)
        {
            u32 w[cap_max_words];
            switch (C.dist)
            {
            case cap_uniform:
            {
                // uniform<mod q>: bit_size + 80 bits, REDC with that many words' rounds (hmpc/random/
                // uniform.hpp:61-66, hmpc/ints/mod.hpp:210-215): v 2^(-32 W) = mont(v R, 2^(-32 W))
                for (int i = 0; i < C.uniform_words; ++i)
                    w[i] = g.next();
                if (C.uniform_bits % 32)
                    w[C.uniform_words - 1] &= (1u << (C.uniform_bits % 32)) - 1u;
                u32 t[L];
                mont_from_words<L>(t, w, C.uniform_words, r2, c32, P);
                mod_mul<L>(r, t, kuni.v, P);
                return;
            }
            case cap_centered_binomial:
            {
                u32 w0 = g.next();
                u32 w1 = capability_two_words(C.bits) ? g.next() : 0u;
                mod_from_small_signed<L>(r, capability_small_words(w0, w1, C.bits), r2.v, P);
                return;
            }
            case cap_binomial:
            {
                int words = (C.bits + 31) / 32;
                u32 pop = 0;
                for (int i = 0; i < words; ++i)
                {
                    u32 x = g.next();
                    if (i == words - 1 && (C.bits % 32))
                        x &= (1u << (C.bits % 32)) - 1u;
                    pop += static_cast<u32>(__builtin_popcount(x));
                }
                u32 x[L];
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = l == 0 ? pop : 0u;
                mod_mul<L>(r, x, r2.v, P);
                return;
            }
            default:
            {
                // drown_uniform: a bigint of `bits` bits, the top word masked (unsigned) or
                // sign-extended from bit bits - 1 (signed) (hmpc/random/uniform.hpp:34-47,
                // hmpc/core/bit_span.hpp:180-200), then (value mod q) R mod q
                int words = (C.bits + 31) / 32;
                for (int i = 0; i < words; ++i)
                    w[i] = g.next();
                int top = C.bits % 32;
                if (top)
                    w[words - 1] &= (1u << top) - 1u;
                bool negative = C.dist == cap_drown_signed && ((w[words - 1] >> ((C.bits - 1) % 32)) & 1u);
                if (negative)
                {
                    // magnitude 2^bits - raw: two's complement within `words` words after the mask
                    if (top)
                        w[words - 1] |= ~((1u << top) - 1u);
                    u32 c = 1;
                    for (int i = 0; i < words; ++i)
                        w[i] = addc(~w[i], 0u, c, c);
                }
                mont_from_words<L>(r, w, words, r2, c32, P);
                if (negative)
                    mod_negate<L>(r, P);
                return;
            }
            }
        }
    }

    // one thread per stream: an element (element layout) or a pair (NTT / INTT stage-0 layouts)
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_sample_capability(cap_args C, residue<L> r2, residue<L> c32,
                                                                         residue<L> kuni, u32* __restrict__ out,
                                                                         mod_params P)
    {
        u64 id = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        u64 total = C.total;
        u32 r[L];
        if (C.layout == cap_layout_element)
        {
            if (id >= total)
                return;
            dev_stream g(&C.key, C.nonce, static_cast<u32>(id));
            draw<L>(r, g, C, r2, c32, kuni, P);
            store_planar<L>(out, total, id, r);
            return;
        }
        u64 half = u64{1} << (C.log_n - 1);
        if (id >= total / 2)
            return;
        u64 b = id >> (C.log_n - 1);
        u64 t = id & (half - 1);
        u64 first = C.layout == cap_layout_ntt ? (b << C.log_n) + t : (b << C.log_n) + 2 * t;
        u64 second = C.layout == cap_layout_ntt ? first + half : first + 1;
        dev_stream g(&C.key, C.nonce, static_cast<u32>(first));
        draw<L>(r, g, C, r2, c32, kuni, P);
        store_planar<L>(out, total, first, r);
        draw<L>(r, g, C, r2, c32, kuni, P);
        store_planar<L>(out, total, second, r);
    }

    hipError_t launch_sample_capability(int L, mod_params const& P, u32 const* r2, u32 const* c32, u32 const* kuni,
                                        cap_args const& C, u32* out, hipStream_t s)
    {
        std::size_t units = C.layout == cap_layout_element ? C.total : C.total / 2;
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_sample_capability<LL><<<grid_for(units), block_threads, 0, s>>>(C, as_residue<LL>(r2), as_residue<LL>(c32),
                                                                             as_residue<LL>(kuni), out, P);
            return hipGetLastError();
        });
    }

    // ------------------------------------------------------------------ c0, c1 of an encryption
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_bgv_combine(u64 count, u64 N, u32 const* __restrict__ pk_a,
                                                                   u32 const* __restrict__ pk_b, u32 const* __restrict__ U,
                                                                   u32 const* __restrict__ V, u32 const* __restrict__ W,
                                                                   u32 const* __restrict__ M, residue<L> p_const,
                                                                   u32* __restrict__ c0, u32* __restrict__ c1, mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u64 k = e & (N - 1);
        u32 u[L], x[L], t[L], y[L];
        load_planar<L>(u, U, count, e);
        load_planar<L>(x, pk_a, N, k);
        mod_mul<L>(t, x, u, P);
        load_planar<L>(x, V, count, e);
        mod_mul<L>(y, x, p_const.v, P);
        mod_add<L>(t, t, y, P);
        load_planar<L>(x, M, count, e);
        mod_add<L>(t, t, x, P);
        store_planar<L>(c0, count, e, t);
        load_planar<L>(x, pk_b, N, k);
        mod_mul<L>(t, x, u, P);
        load_planar<L>(x, W, count, e);
        mod_mul<L>(y, x, p_const.v, P);
        mod_add<L>(t, t, y, P);
        store_planar<L>(c1, count, e, t);
    }

    hipError_t launch_bgv_combine(int L, mod_params const& P, std::size_t count, std::size_t N, u32 const* pk_a, u32 const* pk_b,
                                  u32 const* U, u32 const* V, u32 const* W, u32 const* M, u32 const* p_const, u32* c0, u32* c1,
                                  hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_bgv_combine<LL><<<grid_for(count), block_threads, 0, s>>>(count, N, pk_a, pk_b, U, V, W, M,
                                                                       as_residue<LL>(p_const), c0, c1, P);
            return hipGetLastError();
        });
    }
}
