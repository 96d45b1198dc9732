// CPU ORACLE — test infrastructure only.
//
// This file is a plain C++ restatement of hmpc's algorithms for the NTT / BGV /
// ChaCha20 path. It exists to check the HIP library (and to serve as the timed CPU
// baseline); nothing in the product links, loads or calls it. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// Every function names the reference file:line it restates (paths relative to the
// hmpc source tree). Arithmetic is exact 32-bit-limb integer work, limb-planar
// storage `u32[L][count]` (hmpc/comp/accessor.hpp:191-197).
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   chacha20_block            RFC 8439 vectors held in tests/crypto/chacha.cpp:9-196
//   montgomery constants      tests/ints/mod.cpp:18-42 (150-bit p), closed forms a*R mod q
//   root_of_unity             tests/ints/poly_mod.cpp:78-79, tests/ints/num/theory/root_of_unity.cpp
//   ntt_forward / inverse     tests/ints/poly_mod.cpp:92-103 KATs, closed-form DFT, schoolbook
//                             negacyclic convolution theorem, INTT(NTT(x)) == x
//   samplers (explicit key)   closed form from ChaCha words (python big-int recomputation)
//   capability CBD layout     parity unpinned by any reference vector: checked only by a
//                             python recomputation from ChaCha words following
//                             hmpc/comp/queue.hpp:185-253 and
//                             hmpc/expr/number_theoretic_transform.hpp:320-333
//   bgv encrypt / decrypt     decrypt round trip (tests/expr/crypto/lhe/enc.cpp:65-72 shape),
//                             linear-algebra recomputation in python big ints
//   cast (centred lift)       the cross-modulus KATs of tests/ints/mod.cpp:60-78 (q = 99 against the
//                             150-bit p) and a python big-int closed form (signed centred value
//                             mod Q, Montgomery form) over edge values in both directions
//   capability samplers       python recomputation from ChaCha words (textbook ChaCha, tests/bigref.py)
//                             per layout and distribution; signed conversions against the
//                             KATs of tests/ints/mod.cpp:50-56 and closed forms
//   bgv with an Rp plaintext  python big-int recomputation at small N of c0 = A U + p V + M and
//                             c1 = B U + p W from independently computed samples and closed-form
//                             DFTs, and decryption of hmpc's own test case (enc.cpp:15-72)
//   cipher_encrypt / decrypt  RFC 8439 section 2.4.2 vector held in tests/crypto/cipher.cpp:8-36,
//                             the round trip of tests/expr/crypto/cipher.cpp, and the continuous
//                             keystream of chacha20_stream (one stream across hmpc's batches)

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace oracle
{
    using u32 = std::uint32_t;
    using u64 = std::uint64_t;

    constexpr int max_limbs = 16;

    // ---------------------------------------------------------------- limb kernels
    // hmpc/core/multiply.hpp:70-131 (extended_multiply_add: a*b + c + d as (lo, hi))
    inline void mad(u32 a, u32 b, u32 c, u32 d, u32& lo, u32& hi)
    {
        u64 t = static_cast<u64>(a) * b + c + d;
        lo = static_cast<u32>(t);
        hi = static_cast<u32>(t >> 32);
    }

    // hmpc/core/num/multiply.hpp:17-101: schoolbook product, row by row over `right`.
    inline void num_multiply(u32* result, int result_limbs, u32 const* left, int left_limbs, u32 const* right, int right_limbs)
    {
        u32 inter[2 * max_limbs + 2] = {};
        int inter_limbs = std::min(result_limbs, left_limbs + right_limbs);
        for (int i = 0; i < std::min(inter_limbs, right_limbs); ++i)
        {
            int left_overlap = std::min(inter_limbs - i, left_limbs);
            u32 carry = 0;
            for (int j = 0; j < left_overlap; ++j)
            {
                int k = i + j;
                u32 current = (i == 0) ? 0 : inter[k];
                u32 lo, hi;
                mad(left[j], right[i], current, carry, lo, hi);
                inter[k] = lo;
                carry = (k + 1 < result_limbs) ? hi : 0;
            }
            if (inter_limbs > left_limbs + i)
                inter[i + left_limbs] = carry;
        }
        for (int i = 0; i < result_limbs; ++i)
            result[i] = i < inter_limbs ? inter[i] : 0;
    }

    // a >= b over `limbs` limbs (hmpc/core/num/compare.hpp:154-168)
    inline bool greater_equal(u32 const* a, int a_limbs, u32 const* b, int b_limbs)
    {
        int n = std::max(a_limbs, b_limbs);
        for (int i = n - 1; i >= 0; --i)
        {
            u32 x = i < a_limbs ? a[i] : 0;
            u32 y = i < b_limbs ? b[i] : 0;
            if (x != y)
                return x > y;
        }
        return true;
    }

    // result = a - b over `limbs`, returns borrow (hmpc/core/subtract.hpp:21-60)
    inline u32 sub_limbs(u32* result, u32 const* a, u32 const* b, int limbs)
    {
        u32 borrow = 0;
        for (int i = 0; i < limbs; ++i)
        {
            u64 t = static_cast<u64>(a[i]) - b[i] - borrow;
            result[i] = static_cast<u32>(t);
            borrow = static_cast<u32>(t >> 63);
        }
        return borrow;
    }

    // result = a + b over `limbs`, returns carry (hmpc/core/add.hpp:118-192)
    inline u32 add_limbs(u32* result, u32 const* a, u32 const* b, int limbs)
    {
        u32 carry = 0;
        for (int i = 0; i < limbs; ++i)
        {
            u64 t = static_cast<u64>(a[i]) + b[i] + carry;
            result[i] = static_cast<u32>(t);
            carry = static_cast<u32>(t >> 32);
        }
        return carry;
    }

    inline int bit_width(u32 const* x, int limbs)
    {
        for (int i = limbs - 1; i >= 0; --i)
            if (x[i] != 0)
                return 32 * i + (32 - __builtin_clz(x[i]));
        return 0;
    }

    struct modulus
    {
        int limbs = 0;     // L = ceil(bit_size / 32)  (hmpc/ints/mod.hpp:90-93)
        int bit_size = 0;  // bit_width(q - 1)
        u32 q[max_limbs] = {};
        u32 nprime = 0;    // -q^{-1} mod 2^32 (hmpc/ints/mod.hpp:98-103)
        u32 r1[max_limbs] = {};  // R mod q
        u32 r2[max_limbs] = {};  // R^2 mod q
    };

    // hmpc/core/num/modulo.hpp:30-66: SOS Montgomery reduction with r rounds.
    // value has `value_limbs` limbs, requires value < B^r * q. Output canonical, L limbs.
    inline void montgomery_reduce(u32* result, u32 const* value, int value_limbs, modulus const& m, int r)
    {
        int p = m.limbs;
        int t_limbs = r + p + 1;
        u32 T[3 * max_limbs + 2] = {};
        for (int i = 0; i < std::min(value_limbs, t_limbs); ++i)
            T[i] = value[i];
        for (int i = 0; i < r; ++i)
        {
            u32 mi = T[i] * m.nprime;
            u32 carry = 0;
            for (int j = 0; j < p; ++j)
            {
                u32 lo, hi;
                mad(mi, m.q[j], T[i + j], carry, lo, hi);
                T[i + j] = lo;
                carry = hi;
            }
            for (int k = i + p; k < t_limbs; ++k)
            {
                u64 s = static_cast<u64>(T[k]) + carry;
                T[k] = static_cast<u32>(s);
                carry = static_cast<u32>(s >> 32);
            }
        }
        u32 const* S = T + r;  // p + 1 limbs
        u32 diff[max_limbs + 1];
        if (greater_equal(S, p + 1, m.q, p))
        {
            u32 qext[max_limbs + 1] = {};
            std::copy(m.q, m.q + p, qext);
            sub_limbs(diff, S, qext, p + 1);
            std::copy(diff, diff + p, result);
        }
        else
        {
            std::copy(S, S + p, result);
        }
    }

    // hmpc/ints/mod.hpp:17-34 (HMPC_OPERATOR_BODY_ADD)
    inline void mod_add(u32* result, u32 const* a, u32 const* b, modulus const& m)
    {
        int L = m.limbs;
        u32 sum[max_limbs + 1];
        sum[L] = add_limbs(sum, a, b, L);
        u32 qext[max_limbs + 1] = {};
        std::copy(m.q, m.q + L, qext);
        u32 diff[max_limbs + 1];
        u32 underflow = sub_limbs(diff, sum, qext, L + 1);
        std::copy(underflow ? sum : diff, (underflow ? sum : diff) + L, result);
    }

    // hmpc/ints/mod.hpp:35-40 (HMPC_OPERATOR_BODY_SUBTRACT)
    inline void mod_sub(u32* result, u32 const* a, u32 const* b, modulus const& m)
    {
        int L = m.limbs;
        u32 diff[max_limbs];
        u32 underflow = sub_limbs(diff, a, b, L);
        u32 masked[max_limbs];
        for (int i = 0; i < L; ++i)
            masked[i] = underflow ? m.q[i] : 0;
        add_limbs(result, diff, masked, L);
    }

    // hmpc/ints/mod.hpp:41-46 (HMPC_OPERATOR_BODY_MULTIPLY): multiply, then REDC with r = L.
    inline void mod_mul(u32* result, u32 const* a, u32 const* b, modulus const& m)
    {
        int L = m.limbs;
        u32 product[2 * max_limbs];
        num_multiply(product, 2 * L, a, L, b, L);
        montgomery_reduce(result, product, 2 * L, m, L);
    }

    // hmpc/ints/mod.hpp:165-173: from an unsigned value of <= L limbs: REDC(value * R^2 mod q).
    inline void mod_from_unsigned(u32* result, u32 const* value, int value_limbs, modulus const& m)
    {
        int L = m.limbs;
        if (value_limbs <= L)
        {
            u32 product[2 * max_limbs + 1];
            num_multiply(product, value_limbs + L, value, value_limbs, m.r2, L);
            montgomery_reduce(result, product, value_limbs + L, m, L);
        }
        else
        {
            // hmpc/ints/mod.hpp:180-189: two reductions, the second with R' R^2 mod q.
            u32 y[max_limbs];
            montgomery_reduce(y, value, value_limbs, m, value_limbs);
            // factor = R' * R^2 mod q, with R' = 2^(32 value_limbs): compute as REDC chain on host.
            // R' mod q in Montgomery form: mont(R') = R' R mod q. We need R' R^2 mod q = mont(R') * R mod q.
            // Build 2^(32 value_limbs) mod q by repeated doubling from 1.
            u32 x[max_limbs + 1] = {};
            x[0] = 1;
            for (int bit = 0; bit < 32 * value_limbs; ++bit)
            {
                u32 d[max_limbs + 1];
                d[L] = add_limbs(d, x, x, L);
                u32 qext[max_limbs + 1] = {};
                std::copy(m.q, m.q + L, qext);
                u32 s[max_limbs + 1];
                u32 underflow = sub_limbs(s, d, qext, L + 1);
                std::copy(underflow ? d : s, (underflow ? d : s) + L, x);
            }
            // factor = x * R^2 mod q = REDC(x * R^3) ... simpler: factor = mont(x) * R mod q = REDC(mont(x) * R^2)
            u32 mx[max_limbs], factor[max_limbs];
            mod_from_unsigned(mx, x, L, m);
            mod_from_unsigned(factor, mx, L, m);
            u32 product[2 * max_limbs];
            num_multiply(product, 2 * L, y, L, factor, L);
            montgomery_reduce(result, product, 2 * L, m, L);
        }
    }

    // hmpc/ints/mod.hpp:356-361: out-conversion REDC(x) with r = L.
    inline void mod_to_unsigned(u32* result, u32 const* x, modulus const& m)
    {
        montgomery_reduce(result, x, m.limbs, m, m.limbs);
    }

    // hmpc/ints/mod.hpp:219-226: small signed value -> add q when negative, then from_unsigned.
    inline void mod_from_small_signed(u32* result, long long value, modulus const& m)
    {
        int L = m.limbs;
        u32 v[max_limbs] = {};
        if (value >= 0)
        {
            v[0] = static_cast<u32>(value);
            if (L > 1)
                v[1] = static_cast<u32>(static_cast<u64>(value) >> 32);
        }
        else
        {
            // q + value, value small and negative
            u32 neg[max_limbs] = {};
            u64 a = static_cast<u64>(-value);
            neg[0] = static_cast<u32>(a);
            if (L > 1)
                neg[1] = static_cast<u32>(a >> 32);
            sub_limbs(v, m.q, neg, L);
        }
        mod_from_unsigned(result, v, L, m);
    }

    // hmpc/ints/mod.hpp:381-434 (runtime pow): square-and-multiply over exponent bits, LSB first.
    inline void mod_pow(u32* result, u32 const* base, u32 const* exponent, int exponent_limbs, modulus const& m)
    {
        int L = m.limbs;
        u32 r[max_limbs], v[max_limbs];
        std::copy(m.r1, m.r1 + L, r);  // one
        std::copy(base, base + L, v);
        int bits = bit_width(exponent, exponent_limbs);
        for (int i = 0; i < bits; ++i)
        {
            if ((exponent[i / 32] >> (i % 32)) & 1u)
                mod_mul(r, r, v, m);
            mod_mul(v, v, v, m);
        }
        std::copy(r, r + L, result);
    }

    inline bool mod_is_zero(u32 const* x, int L)
    {
        for (int i = 0; i < L; ++i)
            if (x[i])
                return false;
        return true;
    }

    inline bool mod_equal(u32 const* x, u32 const* y, int L)
    {
        for (int i = 0; i < L; ++i)
            if (x[i] != y[i])
                return false;
        return true;
    }

    // Inverse of a Montgomery value via Fermat (q prime): equals the extended-Euclid inverse
    // of hmpc/ints/mod.hpp:343-354 whenever it exists.
    inline void mod_invert(u32* result, u32 const* x, modulus const& m)
    {
        int L = m.limbs;
        u32 e[max_limbs];
        u32 two[max_limbs] = {2};
        sub_limbs(e, m.q, two, L);
        mod_pow(result, x, e, L, m);
    }

    inline bool setup_modulus(modulus& m, u32 const* q, int q_limbs)
    {
        int qbits = bit_width(q, q_limbs);
        if (qbits < 2 || (q[0] & 1u) == 0)
            return false;
        u32 qm1[max_limbs + 1] = {};
        u32 one[max_limbs + 1] = {1};
        std::copy(q, q + q_limbs, qm1);
        sub_limbs(qm1, qm1, one, q_limbs);
        m.bit_size = bit_width(qm1, q_limbs);
        m.limbs = (m.bit_size + 31) / 32;
        if (m.limbs > max_limbs || bit_width(q, q_limbs) > 32 * m.limbs)
            return false;
        std::fill(m.q, m.q + max_limbs, 0u);
        std::copy(q, q + m.limbs, m.q);
        // N' = -q^{-1} mod 2^32 by Newton iteration
        u32 inv = 1;
        for (int i = 0; i < 6; ++i)
            inv *= 2u - m.q[0] * inv;
        m.nprime = 0u - inv;
        // R mod q and R^2 mod q by doubling from 1
        int L = m.limbs;
        u32 x[max_limbs + 1] = {};
        x[0] = 1;
        u32 qext[max_limbs + 1] = {};
        std::copy(m.q, m.q + L, qext);
        for (int bit = 0; bit < 64 * L; ++bit)
        {
            u32 d[max_limbs + 1];
            d[L] = add_limbs(d, x, x, L);
            u32 s[max_limbs + 1];
            u32 underflow = sub_limbs(s, d, qext, L + 1);
            std::copy(underflow ? d : s, (underflow ? d : s) + L, x);
            if (bit == 32 * L - 1)
                std::copy(x, x + L, m.r1);
        }
        std::copy(x, x + L, m.r2);
        return true;
    }

    // ---------------------------------------------------------------- ChaCha20
    // hmpc/crypto/chacha.hpp:13-47 (quarter round, double round)
    inline u32 rotl(u32 x, int r) { return (x << r) | (x >> (32 - r)); }
    inline void quarter_round(u32& a, u32& b, u32& c, u32& d)
    {
        a += b; d ^= a; d = rotl(d, 16);
        c += d; b ^= c; b = rotl(b, 12);
        a += b; d ^= a; d = rotl(d, 8);
        c += d; b ^= c; b = rotl(b, 7);
    }

    // hmpc/crypto/chacha.hpp:100-109,144-166: state [c0..c3, key[8], counter, nonce[3]]
    inline void chacha20_block(u32 const* key, u32 counter, u32 const* nonce, u32* out)
    {
        u32 s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
        for (int i = 0; i < 8; ++i)
            s[4 + i] = key[i];
        s[12] = counter;
        s[13] = nonce[0];
        s[14] = nonce[1];
        s[15] = nonce[2];
        u32 x[16];
        std::copy(s, s + 16, x);
        for (int r = 0; r < 10; ++r)
        {
            quarter_round(x[0], x[4], x[8], x[12]);
            quarter_round(x[1], x[5], x[9], x[13]);
            quarter_round(x[2], x[6], x[10], x[14]);
            quarter_round(x[3], x[7], x[11], x[15]);
            quarter_round(x[0], x[5], x[10], x[15]);
            quarter_round(x[1], x[6], x[11], x[12]);
            quarter_round(x[2], x[7], x[8], x[13]);
            quarter_round(x[3], x[4], x[9], x[14]);
        }
        for (int i = 0; i < 16; ++i)
            out[i] = x[i] + s[i];
    }

    // hmpc/random/number_generator.hpp:41-104: buffered word stream, block generated on
    // construction, counter incremented after each block, refill after 16 words.
    struct number_generator
    {
        u32 key[8];
        u32 nonce[3];
        u32 counter;
        u32 block[16];
        int index;

        number_generator(u32 const* k, u32 const* n, u32 c)
        {
            std::copy(k, k + 8, key);
            std::copy(n, n + 3, nonce);
            counter = c;
            refill();
        }
        void refill()
        {
            chacha20_block(key, counter, nonce, block);
            counter += 1;
            index = 0;
        }
        u32 next()
        {
            u32 w = block[index++];
            if (index == 16)
                refill();
            return w;
        }
        // `uniform(bit_span<bits>)`: ceil(bits/32) words, top word masked to bits % 32 (if non-zero)
        // (hmpc/core/bit_span.hpp:188-200 normal write)
        void uniform_bits(u32* out, int bits)
        {
            int words = (bits + 31) / 32;
            for (int i = 0; i < words; ++i)
                out[i] = next();
            if (bits % 32)
                out[words - 1] &= (1u << (bits % 32)) - 1u;
        }
    };

    // hmpc/random/binomial.hpp:30-123: popcount of `bits` random bits minus bits/2.
    inline long long centered_binomial(number_generator& g, int bits)
    {
        u32 w[8] = {};
        g.uniform_bits(w, bits);
        int words = (bits + 31) / 32;
        long long pop = 0;
        for (int i = 0; i < words; ++i)
            pop += __builtin_popcount(w[i]);
        return pop - bits / 2;
    }

    // A capability draw: the centred binomial above, or at width 0 the uniform ternary draw that
    // hmpc does not have (the C3 workload's "ternary uniform" secret and ephemeral key): two
    // words w0, w1 as w = w1 * 2^32 + w0, then floor(3 w / 2^64) - 1, within 2^-63 of uniform.
    inline long long capability_small(number_generator& g, int bits)
    {
        if (bits != 0)
            return centered_binomial(g, bits);
        unsigned __int128 w = g.next();
        w |= static_cast<unsigned __int128>(g.next()) << 32;
        return static_cast<long long>((w * 3) >> 64) - 1;
    }

    // hmpc/random/uniform.hpp:34-66 + hmpc/ints/mod.hpp:210-215:
    // draw bit_size + 80 bits, then REDC with r = max(L, W) rounds.
    inline void uniform_mod(u32* out, number_generator& g, modulus const& m)
    {
        int bits = m.bit_size + 80;
        int W = (bits + 31) / 32;
        u32 v[max_limbs + 4] = {};
        g.uniform_bits(v, bits);
        montgomery_reduce(out, v, W, m, std::max(m.limbs, W));
    }

    // hmpc/expr/random/number_generator.hpp:52-72: explicit-key stream for element `index`
    // of a tensor with `elements` elements: nonce = index + elements * offset (96-bit), counter 0.
    inline number_generator explicit_generator(u32 const* key, u64 index, u64 elements, u64 offset)
    {
        u64 v = index + elements * offset;
        u32 nonce[3] = {static_cast<u32>(v), static_cast<u32>(v >> 32), 0u};
        return number_generator(key, nonce, 0u);
    }

    // 96-bit nonce += value (hmpc/comp/queue.hpp:242-253)
    inline void nonce_add(u32* nonce, u64 value)
    {
        u64 lo = static_cast<u64>(nonce[0]) + static_cast<u32>(value);
        nonce[0] = static_cast<u32>(lo);
        u64 mid = static_cast<u64>(nonce[1]) + static_cast<u32>(value >> 32) + (lo >> 32);
        nonce[1] = static_cast<u32>(mid);
        nonce[2] += static_cast<u32>(mid >> 32);
    }

    // ---------------------------------------------------------------- roots and tables
    inline u32 bit_reverse(u32 v, int bits)
    {
        u32 r = 0;
        for (int i = 0; i < bits; ++i)
            r |= ((v >> i) & 1u) << (bits - 1 - i);
        return r;
    }

    // hmpc/ints/num/theory/root_of_unity.hpp:19-54: deterministic search with the all-zero
    // ChaCha stream; candidate x (Montgomery representation) -> g = x^((q-1)/degree), accept if
    // g^(degree/2) != 1 and g^degree == 1.
    inline bool root_of_unity(u32* result, modulus const& m, int log_degree)
    {
        int L = m.limbs;
        // exponent = (q - 1) / degree, requires degree | q - 1
        u32 qm1[max_limbs] = {};
        u32 one_int[max_limbs] = {1};
        sub_limbs(qm1, m.q, one_int, L);
        for (int b = 0; b < log_degree; ++b)
            if ((qm1[b / 32] >> (b % 32)) & 1u)
                return false;
        u32 exponent[max_limbs] = {};
        for (int b = 0; b < 32 * L - log_degree; ++b)
        {
            int src = b + log_degree;
            if ((qm1[src / 32] >> (src % 32)) & 1u)
                exponent[b / 32] |= 1u << (b % 32);
        }
        u32 half_degree[max_limbs] = {}, degree[max_limbs] = {};
        // degree < 2^32 for all supported transforms
        degree[0] = 1u << log_degree;
        half_degree[0] = 1u << (log_degree - 1);
        u32 zero_key[8] = {}, zero_nonce[3] = {};
        number_generator g(zero_key, zero_nonce, 0u);
        for (int attempt = 0; attempt < 4096; ++attempt)
        {
            u32 x[max_limbs] = {};
            g.uniform_bits(x, m.bit_size);
            if (!greater_equal(qm1, L, x, L))  // x > q - 1
                sub_limbs(x, x, m.q, L);
            if (mod_is_zero(x, L))
                continue;
            u32 gval[max_limbs], t1[max_limbs], t2[max_limbs];
            mod_pow(gval, x, exponent, L, m);
            mod_pow(t1, gval, half_degree, 1, m);
            mod_pow(t2, gval, degree, 1, m);
            if (!mod_equal(t1, m.r1, L) && mod_equal(t2, m.r1, L))
            {
                std::copy(gval, gval + L, result);
                return true;
            }
        }
        return false;
    }

    // hmpc/expr/number_theoretic_transform.hpp:35-88: roots[0] = 1 (forward) or N^{-1}
    // (inverse); roots[brv(i)] = root^i; inverse roots[1] pre-multiplied by N^{-1}.
    // Output row-major [N][L] (element-major) for internal use.
    inline bool make_roots(std::vector<u32>& roots, modulus const& m, int log_n, bool inverse)
    {
        int L = m.limbs;
        u32 N = 1u << log_n;
        u32 phi[max_limbs];
        if (!root_of_unity(phi, m, log_n + 1))
            return false;
        u32 root[max_limbs];
        if (inverse)
            mod_invert(root, phi, m);
        else
            std::copy(phi, phi + L, root);
        u32 nval[max_limbs] = {N};
        u32 nmont[max_limbs], norm[max_limbs];
        mod_from_unsigned(nmont, nval, 1, m);
        mod_invert(norm, nmont, m);
        roots.assign(static_cast<size_t>(N) * L, 0u);
        if (inverse)
            std::copy(norm, norm + L, &roots[0]);
        else
            std::copy(m.r1, m.r1 + L, &roots[0]);
        std::copy(root, root + L, &roots[static_cast<size_t>(bit_reverse(1, log_n)) * L]);
        u32 power[max_limbs];
        std::copy(root, root + L, power);
        for (u32 i = 2; i < N; ++i)
        {
            mod_mul(power, power, root, m);
            u32 j = bit_reverse(i, log_n);
            if (inverse && j == 1)
                mod_mul(&roots[static_cast<size_t>(j) * L], power, norm, m);
            else
                std::copy(power, power + L, &roots[static_cast<size_t>(j) * L]);
        }
        return true;
    }

    // ---------------------------------------------------------------- threading helper
    inline void parallel_for(size_t count, int threads, std::function<void(size_t, size_t)> const& body)
    {
        if (threads <= 1 || count < 2)
        {
            body(0, count);
            return;
        }
        std::vector<std::thread> pool;
        size_t chunk = (count + threads - 1) / threads;
        for (int t = 0; t < threads; ++t)
        {
            size_t begin = t * chunk;
            size_t end = std::min(count, begin + chunk);
            if (begin >= end)
                break;
            pool.emplace_back([&, begin, end] { body(begin, end); });
        }
        for (auto& th : pool)
            th.join();
    }

    // element accessors for limb-planar storage
    inline void load(u32* x, u32 const* planes, size_t count, size_t index, int L)
    {
        for (int l = 0; l < L; ++l)
            x[l] = planes[static_cast<size_t>(l) * count + index];
    }
    inline void store(u32* planes, size_t count, size_t index, u32 const* x, int L)
    {
        for (int l = 0; l < L; ++l)
            planes[static_cast<size_t>(l) * count + index] = x[l];
    }

    // ---------------------------------------------------------------- cross-modulus cast
    // value mod q (value of value_limbs limbs, q of L limbs) by binary long division
    inline void bigint_mod(u32* result, u32 const* value, int value_limbs, u32 const* q, int L)
    {
        u32 r[max_limbs + 1] = {};
        u32 qext[max_limbs + 1] = {};
        std::copy(q, q + L, qext);
        for (int bit = 32 * value_limbs - 1; bit >= 0; --bit)
        {
            u32 d[max_limbs + 1] = {};
            add_limbs(d, r, r, L + 1);
            d[0] |= (value[bit / 32] >> (bit % 32)) & 1u;
            u32 s[max_limbs + 1];
            u32 underflow = sub_limbs(s, d, qext, L + 1);
            std::copy(underflow ? d : s, (underflow ? d : s) + L + 1, r);
        }
        std::copy(r, r + L, result);
    }

    // hmpc/ints/mod.hpp:146-158 (offset_to<Other>): (modulus - Other mod modulus) mod modulus
    inline void offset_to(u32* result, u32 const* other, int other_limbs, modulus const& m)
    {
        int L = m.limbs;
        u32 r[max_limbs];
        bigint_mod(r, other, other_limbs, m.q, L);
        if (mod_is_zero(r, L))
            std::fill(result, result + L, 0u);
        else
            sub_limbs(result, m.q, r, L);
    }

    // limbs needed for an integer of `bits` bits (at least one)
    inline int limbs_for_bits(int bits) { return bits <= 0 ? 1 : (bits + 31) / 32; }

    // hmpc/ints/mod.hpp:259-307 (mod<Q>(mod<P>)): the canonical integer v of x in `from`; values
    // above floor(P / 2) are shifted by an offset congruent to -P mod Q (Q > P: Q - P; otherwise
    // offset_to<P>), then the integer enters `to` as an unsigned value (one REDC, or two for a value
    // wider than to's R, :165-189).
    inline void mod_cast(u32* result, u32 const* x, modulus const& from, modulus const& to)
    {
        int Lf = from.limbs, Lt = to.limbs;
        int W = std::max(Lf, Lt) + 1;
        u32 v[max_limbs + 1] = {};
        mod_to_unsigned(v, x, from);
        u32 half[max_limbs + 1] = {};
        for (int l = 0; l < Lf; ++l)
            half[l] = (from.q[l] >> 1) | (l + 1 < Lf ? from.q[l + 1] << 31 : 0u);
        bool above_half = !greater_equal(half, Lf, v, Lf);
        u32 offset[max_limbs + 1] = {};
        u32 normalized[max_limbs + 2] = {};
        if (!greater_equal(from.q, Lf, to.q, Lt))  // to.q > from.q
        {
            u32 fq[max_limbs + 1] = {};
            std::copy(from.q, from.q + Lf, fq);
            sub_limbs(offset, to.q, fq, Lt);
            if (!above_half)
                std::fill(offset, offset + Lt, 0u);
            add_limbs(normalized, v, offset, Lt);
            mod_from_unsigned(result, normalized, Lt, to);
        }
        else
        {
            offset_to(offset, from.q, Lf, to);
            // max value P + offset - 1 fixes the width of the normalized integer
            u32 top[max_limbs + 2] = {};
            u32 fq[max_limbs + 2] = {};
            std::copy(from.q, from.q + Lf, fq);
            add_limbs(top, fq, offset, W);
            u32 one[max_limbs + 2] = {1};
            sub_limbs(top, top, one, W);
            int width = limbs_for_bits(bit_width(top, W));
            if (!above_half)
                std::fill(offset, offset + Lt, 0u);
            add_limbs(normalized, v, offset, W);
            mod_from_unsigned(result, normalized, width, to);
        }
    }

    // ---------------------------------------------------------------- capability samplers
    enum distribution
    {
        dist_uniform = 0,         // hmpc/expr/random/uniform.hpp:41-45 -> hmpc/random/uniform.hpp:61-66
        dist_centered_binomial = 1,  // hmpc/expr/random/binomial.hpp (centred), hmpc/random/binomial.hpp:88-128
        dist_binomial = 2,        // hmpc/expr/random/binomial.hpp:42-46 (uncentred), hmpc/random/binomial.hpp:30-58
        dist_drown_signed = 3,    // hmpc/expr/random/uniform.hpp:91-95 with_sign, hmpc/random/uniform.hpp:34-47
        dist_drown_unsigned = 4,  // the same without sign
    };

    // hmpc/ints/mod.hpp:228-253 (mod(sbigint<Bits>) for Bits >= bit_size) and :219-226 (Bits <
    // bit_size): a signed integer of `bits` bits in two's complement over `words` words
    inline void mod_from_signed_bigint(u32* result, u32 const* value, int words, int bits, modulus const& m)
    {
        int L = m.limbs;
        bool negative = (value[(bits - 1) / 32] >> ((bits - 1) % 32)) & 1u;
        if (bits < m.bit_size)
        {
            // q masked by the sign plus the (sign-extended) value, in bit_size bits
            u32 ext[max_limbs] = {};
            for (int l = 0; l < L; ++l)
                ext[l] = l < words ? value[l] : (negative ? 0xffffffffu : 0u);
            u32 qm[max_limbs] = {};
            for (int l = 0; l < L; ++l)
                qm[l] = negative ? m.q[l] : 0u;
            u32 v[max_limbs];
            add_limbs(v, ext, qm, L);
            if (m.bit_size % 32)
                v[L - 1] &= (1u << (m.bit_size % 32)) - 1u;
            mod_from_unsigned(result, v, L, m);
            return;
        }
        // M = 2^(bits - 1) + offset_to<2^(bits - 1)>: the next multiple of q at or above 2^(bits - 1)
        int W = std::max(words, L) + 1;
        u32 pow2[max_limbs + 2] = {};
        pow2[(bits - 1) / 32] = 1u << ((bits - 1) % 32);
        u32 off[max_limbs] = {};
        offset_to(off, pow2, W, m);
        u32 M[max_limbs + 2] = {}, offw[max_limbs + 2] = {};
        std::copy(off, off + L, offw);
        add_limbs(M, pow2, offw, W);
        u32 top[max_limbs + 2] = {}, one[max_limbs + 2] = {1};
        sub_limbs(top, M, one, W);
        int width = limbs_for_bits(bit_width(top, W));
        u32 ext[max_limbs + 2] = {};
        for (int l = 0; l < W; ++l)
            ext[l] = l < words ? value[l] : (negative ? 0xffffffffu : 0u);
        if (!negative)
            std::fill(M, M + W, 0u);
        u32 normalized[max_limbs + 2] = {};
        add_limbs(normalized, ext, M, W);  // value + M for negative values (mod 2^(32 W))
        mod_from_unsigned(result, normalized, width, m);
    }

    // one draw of `dist` from the generator, in Montgomery form
    inline void sample_one(u32* out, number_generator& g, int dist, int bits, modulus const& m)
    {
        switch (dist)
        {
        case dist_uniform:
            uniform_mod(out, g, m);
            return;
        case dist_centered_binomial:
            mod_from_small_signed(out, capability_small(g, bits), m);
            return;
        case dist_binomial:
        {
            // popcount of `bits` random bits as an unsigned integer (hmpc/random/binomial.hpp:30-58)
            u32 w[16] = {};
            g.uniform_bits(w, bits);
            u32 pop = 0;
            for (int i = 0; i < (bits + 31) / 32; ++i)
                pop += static_cast<u32>(__builtin_popcount(w[i]));
            mod_from_unsigned(out, &pop, 1, m);
            return;
        }
        case dist_drown_signed:
        case dist_drown_unsigned:
        {
            // a bigint of `bits` bits: ceil(bits / 32) words, the top one masked (unsigned) or
            // sign-extended from bit bits - 1 (signed), hmpc/core/bit_span.hpp:180-200
            int words = (bits + 31) / 32;
            u32 w[16] = {};
            for (int i = 0; i < words; ++i)
                w[i] = g.next();
            int top = bits % 32;
            if (dist == dist_drown_unsigned)
            {
                if (top)
                    w[words - 1] &= (1u << top) - 1u;
                mod_from_unsigned(out, w, words, m);
            }
            else
            {
                if (top)
                {
                    u32 sign = (w[words - 1] >> (top - 1)) & 1u;
                    u32 mask = (1u << top) - 1u;
                    w[words - 1] = sign ? (w[words - 1] | ~mask) : (w[words - 1] & mask);
                }
                mod_from_signed_bigint(out, w, words, bits, m);
            }
            return;
        }
        default:
            std::fill(out, out + m.limbs, 0u);
        }
    }

    enum capability_layout
    {
        layout_element = 0,  // element-wise kernel: one stream per element (hmpc/comp/queue.hpp:387-412)
        layout_ntt = 1,      // forward NTT stage 0: pair (t, t + N/2) from the stream of index t
        layout_intt = 2,     // inverse NTT stage 0: pair (2t, 2t + 1) from the stream of index 2t
    };

    // Capability-path samples of n polynomials of N = 2^log_n coefficients (coefficient domain, in
    // the order the consuming kernel reads them). The capability of linear index i is the ChaCha20
    // stream (key, kernel nonce, counter = low32(i)) (hmpc/comp/queue.hpp:185-218).
    //   element: coefficient e draws from the stream of e (hmpc/comp/queue.hpp:400-409)
    //   ntt:     t < N/2 of polynomial b: stream b N + t gives coefficient t, then t + N/2
    //            (hmpc/expr/number_theoretic_transform.hpp:320-329)
    //   intt:    stream b N + 2t gives coefficient 2t, then 2t + 1 (:480-489)
    inline void sample_capability(modulus const& m, u32 const* key, u32 const* nonce, int dist, int bits, int layout,
                                  int log_n, size_t n, u32* out)
    {
        int L = m.limbs;
        size_t N = size_t{1} << log_n, total = n * N;
        u32 x[max_limbs];
        if (layout == layout_element)
        {
            for (size_t e = 0; e < total; ++e)
            {
                number_generator g(key, nonce, static_cast<u32>(e));
                sample_one(x, g, dist, bits, m);
                store(out, total, e, x, L);
            }
            return;
        }
        size_t half = N / 2;
        for (size_t b = 0; b < n; ++b)
            for (size_t t = 0; t < half; ++t)
            {
                size_t first = layout == layout_ntt ? b * N + t : b * N + 2 * t;
                size_t second = layout == layout_ntt ? first + half : first + 1;
                number_generator g(key, nonce, static_cast<u32>(first));
                sample_one(x, g, dist, bits, m);
                store(out, total, first, x, L);
                sample_one(x, g, dist, bits, m);
                store(out, total, second, x, L);
            }
    }

    // ---------------------------------------------------------------- NTT
    // Forward transform, following hmpc/expr/number_theoretic_transform.hpp:262-393 stage by
    // stage: stage 0 reads [n][N] and writes the transposed scratch [N][n]; stages
    // 1..log2N-2 update the scratch in place (:144-197); the last stage writes [n][N] back.
    // Scratch is element-major [N*n][L] here (layout inside the oracle is free).
    inline void ntt_forward(modulus const& m, int log_n, size_t n, u32 const* in, u32* out,
                            std::vector<u32> const& roots, int threads)
    {
        int L = m.limbs;
        size_t N = size_t{1} << log_n;
        size_t half = N / 2;
        size_t total = n * N;
        std::vector<u32> scratch(total * L);
        auto S = [&](size_t idx) { return &scratch[idx * L]; };
        auto W = [&](size_t idx) { return &roots[idx * L]; };
        // stage 0 (K2)
        parallel_for(n * half, threads, [&](size_t b, size_t e) {
            for (size_t id = b; id < e; ++id)
            {
                size_t i = id / half, tid = id % half;
                u32 u[max_limbs], v[max_limbs];
                load(u, in, total, tid + i * N, L);
                load(v, in, total, tid + half + i * N, L);
                mod_mul(v, v, W(1), m);
                mod_add(S(i + tid * n), u, v, m);
                mod_sub(S(i + (tid + half) * n), u, v, m);
            }
        });
        // stages 1 .. log2N-2 (K3)
        for (int it = 1; it < log_n - 1; ++it)
        {
            size_t length = size_t{1} << it;
            size_t step = (N / length) / 2;
            parallel_for(half * n, threads, [&](size_t b, size_t e) {
                for (size_t id = b; id < e; ++id)
                {
                    size_t tid = id / n, i = id % n;
                    size_t psi_step = tid / step;
                    size_t target = psi_step * step * 2 + tid % step;
                    u32 const* psi = W(length + psi_step);
                    u32* pu = S(i + target * n);
                    u32* pv = S(i + (target + step) * n);
                    u32 u[max_limbs], v[max_limbs];
                    std::copy(pu, pu + L, u);
                    mod_mul(v, pv, psi, m);
                    mod_add(pu, u, v, m);
                    mod_sub(pv, u, v, m);
                }
            });
        }
        // last stage (K4)
        parallel_for(half * n, threads, [&](size_t b, size_t e) {
            for (size_t id = b; id < e; ++id)
            {
                size_t tid = id / n, i = id % n;
                size_t target = 2 * tid;
                u32 const* psi = W(half + tid);
                u32 u[max_limbs], v[max_limbs], r[max_limbs];
                std::copy(S(i + target * n), S(i + target * n) + L, u);
                mod_mul(v, S(i + (target + 1) * n), psi, m);
                mod_add(r, u, v, m);
                store(out, total, target + i * N, r, L);
                mod_sub(r, u, v, m);
                store(out, total, target + 1 + i * N, r, L);
            }
        });
    }

    // Inverse transform, following hmpc/expr/number_theoretic_transform.hpp:419-549 and the
    // inverse inner_transform (:90-142): GS butterflies (u+v, (u-v) psi), final stage scaled.
    inline void ntt_inverse(modulus const& m, int log_n, size_t n, u32 const* in, u32* out,
                            std::vector<u32> const& roots, int threads)
    {
        int L = m.limbs;
        size_t N = size_t{1} << log_n;
        size_t half = N / 2;
        size_t total = n * N;
        std::vector<u32> scratch(total * L);
        auto S = [&](size_t idx) { return &scratch[idx * L]; };
        auto W = [&](size_t idx) { return &roots[idx * L]; };
        // stage 0 (K5): pairs (2 tid, 2 tid + 1), psi = roots[N/2 + tid]
        parallel_for(n * half, threads, [&](size_t b, size_t e) {
            for (size_t id = b; id < e; ++id)
            {
                size_t i = id / half, tid = id % half;
                size_t target = 2 * tid;
                u32 u[max_limbs], v[max_limbs], d[max_limbs];
                load(u, in, total, target + i * N, L);
                load(v, in, total, target + 1 + i * N, L);
                mod_add(S(i + target * n), u, v, m);
                mod_sub(d, u, v, m);
                mod_mul(S(i + (target + 1) * n), d, W(half + tid), m);
            }
        });
        // stages 1 .. log2N-2 (K6)
        for (int it = 1; it < log_n - 1; ++it)
        {
            size_t length = N >> (it + 1);
            size_t step = (N / length) / 2;
            parallel_for(half * n, threads, [&](size_t b, size_t e) {
                for (size_t id = b; id < e; ++id)
                {
                    size_t tid = id / n, i = id % n;
                    size_t psi_step = tid / step;
                    size_t target = psi_step * step * 2 + tid % step;
                    u32 const* psi = W(length + psi_step);
                    u32* pu = S(i + target * n);
                    u32* pv = S(i + (target + step) * n);
                    u32 u[max_limbs], v[max_limbs], d[max_limbs];
                    std::copy(pu, pu + L, u);
                    std::copy(pv, pv + L, v);
                    mod_add(pu, u, v, m);
                    mod_sub(d, u, v, m);
                    mod_mul(pv, d, psi, m);
                }
            });
        }
        // last stage (K7): (u+v) * roots[0] (= N^{-1}), (u-v) * roots[1] (pre-scaled)
        parallel_for(half * n, threads, [&](size_t b, size_t e) {
            for (size_t id = b; id < e; ++id)
            {
                size_t tid = id / n, i = id % n;
                u32 u[max_limbs], v[max_limbs], t[max_limbs], r[max_limbs];
                std::copy(S(i + tid * n), S(i + tid * n) + L, u);
                std::copy(S(i + (tid + half) * n), S(i + (tid + half) * n) + L, v);
                mod_add(t, u, v, m);
                mod_mul(r, t, W(0), m);
                store(out, total, tid + i * N, r, L);
                mod_sub(t, u, v, m);
                mod_mul(r, t, W(1), m);
                store(out, total, tid + half + i * N, r, L);
            }
        });
    }
}

using namespace oracle;

extern "C"
{
    // ------------------------------------------------------------------ modulus
    void* oracle_modulus_new(u32 const* q, int q_limbs)
    {
        auto* m = new modulus;
        if (!setup_modulus(*m, q, q_limbs))
        {
            delete m;
            return nullptr;
        }
        return m;
    }
    void oracle_modulus_free(void* m) { delete static_cast<modulus*>(m); }
    int oracle_modulus_limbs(void* m) { return static_cast<modulus*>(m)->limbs; }
    int oracle_modulus_bit_size(void* m) { return static_cast<modulus*>(m)->bit_size; }
    // out: nprime (1 word), r1 (L), r2 (L)
    void oracle_modulus_constants(void* mp, u32* nprime, u32* r1, u32* r2)
    {
        auto& m = *static_cast<modulus*>(mp);
        *nprime = m.nprime;
        std::copy(m.r1, m.r1 + m.limbs, r1);
        std::copy(m.r2, m.r2 + m.limbs, r2);
    }

    // ------------------------------------------------------------------ scalar ops (contiguous limbs)
    void oracle_from_unsigned(void* mp, u32 const* value, int value_limbs, u32* out)
    {
        mod_from_unsigned(out, value, value_limbs, *static_cast<modulus*>(mp));
    }
    void oracle_from_small_signed(void* mp, long long value, u32* out)
    {
        mod_from_small_signed(out, value, *static_cast<modulus*>(mp));
    }
    void oracle_to_unsigned(void* mp, u32 const* x, u32* out) { mod_to_unsigned(out, x, *static_cast<modulus*>(mp)); }
    void oracle_mul(void* mp, u32 const* a, u32 const* b, u32* out) { mod_mul(out, a, b, *static_cast<modulus*>(mp)); }
    void oracle_add(void* mp, u32 const* a, u32 const* b, u32* out) { mod_add(out, a, b, *static_cast<modulus*>(mp)); }
    void oracle_sub(void* mp, u32 const* a, u32 const* b, u32* out) { mod_sub(out, a, b, *static_cast<modulus*>(mp)); }
    void oracle_pow(void* mp, u32 const* base, u32 const* e, int e_limbs, u32* out)
    {
        mod_pow(out, base, e, e_limbs, *static_cast<modulus*>(mp));
    }
    void oracle_invert(void* mp, u32 const* x, u32* out) { mod_invert(out, x, *static_cast<modulus*>(mp)); }
    void oracle_montgomery_reduce(void* mp, u32 const* value, int value_limbs, int rounds, u32* out)
    {
        montgomery_reduce(out, value, value_limbs, *static_cast<modulus*>(mp), rounds);
    }

    // root of unity of order 2^log_degree (Montgomery form); inverse selects phi^{-1}
    int oracle_root_of_unity(void* mp, int log_degree, int inverse, u32* out)
    {
        auto& m = *static_cast<modulus*>(mp);
        u32 phi[max_limbs];
        if (!root_of_unity(phi, m, log_degree))
            return 1;
        if (inverse)
            mod_invert(out, phi, m);
        else
            std::copy(phi, phi + m.limbs, out);
        return 0;
    }

    // twiddle table, limb-planar [L][N]
    int oracle_roots(void* mp, int log_n, int inverse, u32* out)
    {
        auto& m = *static_cast<modulus*>(mp);
        std::vector<u32> roots;
        if (!make_roots(roots, m, log_n, inverse != 0))
            return 1;
        size_t N = size_t{1} << log_n;
        for (size_t i = 0; i < N; ++i)
            store(out, N, i, &roots[i * m.limbs], m.limbs);
        return 0;
    }

    // ------------------------------------------------------------------ ChaCha20
    void oracle_chacha20_block(u32 const* key, u32 counter, u32 const* nonce, u32* out)
    {
        chacha20_block(key, counter, nonce, out);
    }
    // first `words` words of the stream (key, nonce, counter)
    void oracle_chacha20_stream(u32 const* key, u32 const* nonce, u32 counter, size_t words, u32* out)
    {
        number_generator g(key, nonce, counter);
        for (size_t i = 0; i < words; ++i)
            out[i] = g.next();
    }

    // Tensor cipher, hmpc/expr/crypto/cipher.hpp:121-161 (enc) and :253-300 (dec) over
    // hmpc/crypto/cipher.hpp:30-45: work item b owns batch_size = lcm(16, L) words, i.e.
    // elements_per_batch = batch_size / L elements, and a generator started at block counter
    // counter0 + b * (batch_size / 16); each element takes the next L words as its mask and is
    // XORed with it. `planar` is u32[L][count], `words` u32[count][L].
    static size_t gcd_size(size_t a, size_t b)
    {
        while (b)
        {
            size_t t = a % b;
            a = b;
            b = t;
        }
        return a;
    }

    static void cipher_batches(int limbs, size_t count, u32 const* key, u32 const* nonce, u32 counter0, bool encrypt,
                               u32 const* in, u32* out)
    {
        size_t L = static_cast<size_t>(limbs);
        size_t batch_size = 16 / gcd_size(16, L) * L;
        size_t per_batch = batch_size / L;
        size_t blocks_per_batch = batch_size / 16;
        size_t batches = (count * L + batch_size - 1) / batch_size;
        for (size_t b = 0; b < batches; ++b)
        {
            number_generator g(key, nonce, static_cast<u32>(counter0 + b * blocks_per_batch));
            for (size_t e = 0; e < per_batch; ++e)
            {
                size_t i = b * per_batch + e;
                if (i >= count)
                    break;
                for (size_t j = 0; j < L; ++j)
                {
                    u32 mask = g.next();
                    if (encrypt)
                        out[i * L + j] = in[j * count + i] ^ mask;
                    else
                        out[j * count + i] = in[i * L + j] ^ mask;
                }
            }
        }
    }

    void oracle_cipher_encrypt(int limbs, size_t count, u32 const* key, u32 const* nonce, u32 counter0, u32 const* planar,
                               u32* words)
    {
        cipher_batches(limbs, count, key, nonce, counter0, true, planar, words);
    }

    void oracle_cipher_decrypt(int limbs, size_t count, u32 const* key, u32 const* nonce, u32 counter0, u32 const* words,
                               u32* planar)
    {
        cipher_batches(limbs, count, key, nonce, counter0, false, words, planar);
    }

    // ------------------------------------------------------------------ NTT
    int oracle_ntt(void* mp, int log_n, size_t n, int inverse, u32 const* in, u32* out, int threads)
    {
        auto& m = *static_cast<modulus*>(mp);
        if (log_n < 2)
            return 2;
        std::vector<u32> roots;
        if (!make_roots(roots, m, log_n, inverse != 0))
            return 1;
        if (inverse)
            ntt_inverse(m, log_n, n, in, out, roots, threads);
        else
            ntt_forward(m, log_n, n, in, out, roots, threads);
        return 0;
    }

    // point-wise ops over `count` elements; b is broadcast with period b_count (b index = i % b_count)
    // op: 0 add, 1 sub, 2 mul  (hmpc/expr/binary_expression.hpp:8-163 executed per element)
    int oracle_pointwise(void* mp, int op, size_t count, u32 const* a, u32 const* b, size_t b_count, u32* out, int threads)
    {
        auto& m = *static_cast<modulus*>(mp);
        int L = m.limbs;
        parallel_for(count, threads, [&](size_t s, size_t e) {
            for (size_t i = s; i < e; ++i)
            {
                u32 x[max_limbs], y[max_limbs], r[max_limbs];
                load(x, a, count, i, L);
                load(y, b, b_count, i % b_count, L);
                if (op == 0)
                    mod_add(r, x, y, m);
                else if (op == 1)
                    mod_sub(r, x, y, m);
                else
                    mod_mul(r, x, y, m);
                store(out, count, i, r, L);
            }
        });
        return 0;
    }

    // ------------------------------------------------------------------ samplers
    // explicit key uniform<mod q> (hmpc/expr/random/uniform_from_number_generator.hpp:11-47)
    int oracle_sample_uniform(void* mp, u32 const* key, u64 offset, size_t count, u32* out, int threads)
    {
        auto& m = *static_cast<modulus*>(mp);
        parallel_for(count, threads, [&](size_t s, size_t e) {
            for (size_t i = s; i < e; ++i)
            {
                auto g = explicit_generator(key, i, count, offset);
                u32 x[max_limbs];
                uniform_mod(x, g, m);
                store(out, count, i, x, m.limbs);
            }
        });
        return 0;
    }

    // explicit key centred binomial with `bits` = 4 * variance
    // (hmpc/expr/random/binomial_from_number_generator.hpp:11-47)
    int oracle_sample_cbd(void* mp, u32 const* key, u64 offset, int bits, size_t count, u32* out, int threads)
    {
        auto& m = *static_cast<modulus*>(mp);
        parallel_for(count, threads, [&](size_t s, size_t e) {
            for (size_t i = s; i < e; ++i)
            {
                auto g = explicit_generator(key, i, count, offset);
                u32 x[max_limbs];
                mod_from_small_signed(x, centered_binomial(g, bits), m);
                store(out, count, i, x, m.limbs);
            }
        });
        return 0;
    }

    // explicit key 32-bit words (first word of each element's stream): BGV messages in [0, 2^32)
    int oracle_sample_words(u32 const* key, u64 offset, size_t count, u32* out)
    {
        for (size_t i = 0; i < count; ++i)
        {
            auto g = explicit_generator(key, i, count, offset);
            out[i] = g.next();
        }
        return 0;
    }

    // Capability-path centred binomial as consumed by the forward NTT's stage 0
    // (hmpc/comp/queue.hpp:185-218,242-253; hmpc/expr/number_theoretic_transform.hpp:320-333):
    // butterfly (i, tid) builds one generator with counter = low32(i*N + tid) and the kernel
    // nonce; coefficient tid draws first, tid + N/2 second. Output: coefficient-domain values.
    int oracle_sample_cbd_capability(void* mp, u32 const* key, u32 const* nonce, int bits, int log_n, size_t n, u32* out)
    {
        auto& m = *static_cast<modulus*>(mp);
        size_t N = size_t{1} << log_n, half = N / 2, total = n * N;
        for (size_t i = 0; i < n; ++i)
            for (size_t tid = 0; tid < half; ++tid)
            {
                number_generator g(key, nonce, static_cast<u32>(i * N + tid));
                u32 x[max_limbs];
                mod_from_small_signed(x, capability_small(g, bits), m);
                store(out, total, i * N + tid, x, m.limbs);
                mod_from_small_signed(x, capability_small(g, bits), m);
                store(out, total, i * N + tid + half, x, m.limbs);
            }
        return 0;
    }

    // ------------------------------------------------------------------ BGV
    // hmpc/expr/crypto/lhe/enc.hpp:19-69 with hmpc/expr/crypto/lhe/randomness.hpp:80-110:
    //   u, v, w = NTT(CBD) on the capability path with nonces nonce, nonce + nN, nonce + 2nN
    //   message = NTT(mod_q(m)) (unsigned lift of a 32-bit word, examples/bgv.cpp:66-69)
    //   c0 = (A*u + v*p) + message ; c1 = B*u + w*p  (A, B, p in NTT/Montgomery form)
    // pk_a, pk_b: [L][N]; messages: u32 [n*N]; c0, c1: [L][n*N]. nonce is advanced by 3nN.
    int oracle_bgv_encrypt(void* mp, int log_n, size_t n, u32 const* pk_a, u32 const* pk_b, u32 const* messages,
                           u32 const* p_mont, u32 const* key, u32* nonce, int bits_u, int bits_v, int bits_w,
                           u32* c0, u32* c1, int threads)
    {
        auto& m = *static_cast<modulus*>(mp);
        int L = m.limbs;
        size_t N = size_t{1} << log_n, total = n * N;
        std::vector<u32> fr, ir;
        if (!make_roots(fr, m, log_n, false))
            return 1;
        std::vector<u32> tmp(total * L), u(total * L), v(total * L), w(total * L), msg(total * L);
        oracle_sample_cbd_capability(mp, key, nonce, bits_u, log_n, n, tmp.data());
        nonce_add(nonce, total);
        ntt_forward(m, log_n, n, tmp.data(), u.data(), fr, threads);
        oracle_sample_cbd_capability(mp, key, nonce, bits_v, log_n, n, tmp.data());
        nonce_add(nonce, total);
        ntt_forward(m, log_n, n, tmp.data(), v.data(), fr, threads);
        parallel_for(total, threads, [&](size_t s, size_t e) {
            for (size_t i = s; i < e; ++i)
            {
                u32 x[max_limbs];
                u32 word = messages[i];
                mod_from_unsigned(x, &word, 1, m);
                store(tmp.data(), total, i, x, L);
            }
        });
        ntt_forward(m, log_n, n, tmp.data(), msg.data(), fr, threads);
        parallel_for(total, threads, [&](size_t s, size_t e) {
            for (size_t i = s; i < e; ++i)
            {
                u32 a[max_limbs], uu[max_limbs], vv[max_limbs], mm[max_limbs], t1[max_limbs], t2[max_limbs];
                load(a, pk_a, N, i % N, L);
                load(uu, u.data(), total, i, L);
                load(vv, v.data(), total, i, L);
                load(mm, msg.data(), total, i, L);
                mod_mul(t1, a, uu, m);
                mod_mul(t2, vv, p_mont, m);
                mod_add(t1, t1, t2, m);
                mod_add(t1, t1, mm, m);
                store(c0, total, i, t1, L);
            }
        });
        oracle_sample_cbd_capability(mp, key, nonce, bits_w, log_n, n, tmp.data());
        nonce_add(nonce, total);
        ntt_forward(m, log_n, n, tmp.data(), w.data(), fr, threads);
        parallel_for(total, threads, [&](size_t s, size_t e) {
            for (size_t i = s; i < e; ++i)
            {
                u32 b[max_limbs], uu[max_limbs], ww[max_limbs], t1[max_limbs], t2[max_limbs];
                load(b, pk_b, N, i % N, L);
                load(uu, u.data(), total, i, L);
                load(ww, w.data(), total, i, L);
                mod_mul(t1, b, uu, m);
                mod_mul(t2, ww, p_mont, m);
                mod_add(t1, t1, t2, m);
                store(c1, total, i, t1, L);
            }
        });
        return 0;
    }

    // hmpc/expr/crypto/lhe/dec.hpp:13-45: x = INTT(c0 - s*c1), then centred lift; here the
    // plaintext modulus is 2^32, so the message is the low 32 bits of the centred value.
    int oracle_bgv_decrypt(void* mp, int log_n, size_t n, u32 const* s_ntt, u32 const* c0, u32 const* c1, u32* messages, int threads)
    {
        auto& m = *static_cast<modulus*>(mp);
        int L = m.limbs;
        size_t N = size_t{1} << log_n, total = n * N;
        std::vector<u32> x(total * L), y(total * L), ir;
        if (!make_roots(ir, m, log_n, true))
            return 1;
        parallel_for(total, threads, [&](size_t st, size_t e) {
            for (size_t i = st; i < e; ++i)
            {
                u32 a[max_limbs], b[max_limbs], s[max_limbs], t[max_limbs];
                load(a, c0, total, i, L);
                load(b, c1, total, i, L);
                load(s, s_ntt, N, i % N, L);
                mod_mul(t, s, b, m);
                mod_sub(t, a, t, m);
                store(x.data(), total, i, t, L);
            }
        });
        ntt_inverse(m, log_n, n, x.data(), y.data(), ir, threads);
        // centred lift: value > (q-1)/2 -> value - q (two's complement low word = value - q mod 2^32)
        u32 half[max_limbs] = {};
        for (int l = 0; l < L; ++l)
        {
            half[l] = m.q[l] >> 1;
            if (l + 1 < L)
                half[l] |= m.q[l + 1] << 31;
        }
        for (size_t i = 0; i < total; ++i)
        {
            u32 t[max_limbs], v[max_limbs];
            load(t, y.data(), total, i, L);
            mod_to_unsigned(v, t, m);
            bool negative = !greater_equal(half, L, v, L);
            messages[i] = negative ? v[0] - m.q[0] : v[0];
        }
        return 0;
    }

    // schoolbook negacyclic product a*b mod (X^N + 1), Montgomery in and out, [L][N] each
    int oracle_schoolbook_negacyclic(void* mp, int log_n, u32 const* a, u32 const* b, u32* out)
    {
        auto& m = *static_cast<modulus*>(mp);
        int L = m.limbs;
        size_t N = size_t{1} << log_n;
        std::vector<u32> acc(N * L, 0u);
        for (size_t i = 0; i < N; ++i)
        {
            u32 x[max_limbs];
            load(x, a, N, i, L);
            for (size_t j = 0; j < N; ++j)
            {
                u32 y[max_limbs], p[max_limbs];
                load(y, b, N, j, L);
                mod_mul(p, x, y, m);
                size_t k = i + j;
                u32* target = &acc[(k % N) * L];
                if (k < N)
                    mod_add(target, target, p, m);
                else
                    mod_sub(target, target, p, m);
            }
        }
        for (size_t k = 0; k < N; ++k)
            store(out, N, k, &acc[k * L], L);
        return 0;
    }

    // ------------------------------------------------------------------ cast, capability samplers
    void oracle_cast(void* from, void* to, size_t count, u32 const* in, u32* out)
    {
        auto& f = *static_cast<modulus*>(from);
        auto& t = *static_cast<modulus*>(to);
        for (size_t e = 0; e < count; ++e)
        {
            u32 x[max_limbs], y[max_limbs];
            load(x, in, count, e, f.limbs);
            mod_cast(y, x, f, t);
            store(out, count, e, y, t.limbs);
        }
    }

    void oracle_from_signed_bigint(void* mp, u32 const* value, int words, int bits, u32* out)
    {
        mod_from_signed_bigint(out, value, words, bits, *static_cast<modulus*>(mp));
    }

    int oracle_sample_capability(void* mp, u32 const* key, u32 const* nonce, int dist, int bits, int layout, int log_n,
                                 size_t n, u32* out)
    {
        if (layout != layout_element && log_n < 1)
            return 1;
        sample_capability(*static_cast<modulus*>(mp), key, nonce, dist, bits, layout, log_n, n, out);
        return 0;
    }

    // ------------------------------------------------------------------ BGV with an Rp plaintext
    // hmpc/expr/crypto/lhe/enc.hpp:19-69 for a plaintext over mod p (mp):
    //   message = NTT_q(cast<mod q>(INTT_p(x)))  (NTT-form plaintext; coefficient form: no INTT_p)
    //   p = the plaintext modulus as a mod q constant (:59-63, constant_cast = mod q from unsigned)
    //   u = NTT(CBD(bits_u)), v = NTT(v_dist(v_bits)), w = NTT(CBD(bits_w)) on the capability path
    //   (layout ntt), nonces nonce, nonce + nN, nonce + 2nN in hmpc's cache order (u, v, w; the
    //   message transforms draw no randomness); v_dist = centred binomial is randomness()
    //   (randomness.hpp:80-110), drown_signed is drowning_randomness() (:112-143)
    //   c0 = A u + v p + message, c1 = B u + w p
    int oracle_bgv_encrypt_plaintext(void* mqp, void* mpp, int log_n, size_t n, u32 const* pk_a, u32 const* pk_b,
                                     u32 const* plaintext, int plaintext_ntt, u32 const* key, u32* nonce, int bits_u,
                                     int v_dist, int v_bits, int bits_w, u32* c0, u32* c1, int threads)
    {
        auto& mq = *static_cast<modulus*>(mqp);
        auto& mp = *static_cast<modulus*>(mpp);
        int L = mq.limbs, Lp = mp.limbs;
        size_t N = size_t{1} << log_n, total = n * N;
        std::vector<u32> fr;
        if (!make_roots(fr, mq, log_n, false))
            return 1;
        std::vector<u32> xp(total * Lp), tmp(total * L), msg(total * L), u(total * L), v(total * L), w(total * L);
        if (plaintext_ntt)
        {
            std::vector<u32> ipr;
            if (!make_roots(ipr, mp, log_n, true))
                return 2;
            ntt_inverse(mp, log_n, n, plaintext, xp.data(), ipr, threads);
        }
        else
            std::copy(plaintext, plaintext + total * Lp, xp.begin());
        oracle_cast(mpp, mqp, total, xp.data(), tmp.data());
        ntt_forward(mq, log_n, n, tmp.data(), msg.data(), fr, threads);
        u32 p_const[max_limbs];
        mod_from_unsigned(p_const, mp.q, Lp, mq);
        sample_capability(mq, key, nonce, dist_centered_binomial, bits_u, layout_ntt, log_n, n, tmp.data());
        nonce_add(nonce, total);
        ntt_forward(mq, log_n, n, tmp.data(), u.data(), fr, threads);
        sample_capability(mq, key, nonce, v_dist, v_bits, layout_ntt, log_n, n, tmp.data());
        nonce_add(nonce, total);
        ntt_forward(mq, log_n, n, tmp.data(), v.data(), fr, threads);
        sample_capability(mq, key, nonce, dist_centered_binomial, bits_w, layout_ntt, log_n, n, tmp.data());
        nonce_add(nonce, total);
        ntt_forward(mq, log_n, n, tmp.data(), w.data(), fr, threads);
        parallel_for(total, threads, [&](size_t st, size_t e) {
            for (size_t i = st; i < e; ++i)
            {
                u32 a[max_limbs], b[max_limbs], uu[max_limbs], x[max_limbs], t1[max_limbs], t2[max_limbs];
                load(a, pk_a, N, i % N, L);
                load(b, pk_b, N, i % N, L);
                load(uu, u.data(), total, i, L);
                mod_mul(t1, a, uu, mq);
                load(x, v.data(), total, i, L);
                mod_mul(t2, x, p_const, mq);
                mod_add(t1, t1, t2, mq);
                load(x, msg.data(), total, i, L);
                mod_add(t1, t1, x, mq);
                store(c0, total, i, t1, L);
                mod_mul(t1, b, uu, mq);
                load(x, w.data(), total, i, L);
                mod_mul(t2, x, p_const, mq);
                mod_add(t1, t1, t2, mq);
                store(c1, total, i, t1, L);
            }
        });
        return 0;
    }

    // hmpc/expr/crypto/lhe/dec.hpp:13-45: x = INTT_q(c0 - s c1), cast to mod p (centred lift), then
    // NTT_p for an NTT-form plaintext
    int oracle_bgv_decrypt_plaintext(void* mqp, void* mpp, int log_n, size_t n, u32 const* s_ntt, u32 const* c0,
                                     u32 const* c1, int plaintext_ntt, u32* out, int threads)
    {
        auto& mq = *static_cast<modulus*>(mqp);
        auto& mp = *static_cast<modulus*>(mpp);
        int L = mq.limbs, Lp = mp.limbs;
        size_t N = size_t{1} << log_n, total = n * N;
        std::vector<u32> ir;
        if (!make_roots(ir, mq, log_n, true))
            return 1;
        std::vector<u32> x(total * L), y(total * L), z(total * Lp);
        parallel_for(total, threads, [&](size_t st, size_t e) {
            for (size_t i = st; i < e; ++i)
            {
                u32 a[max_limbs], b[max_limbs], s[max_limbs], t[max_limbs];
                load(a, c0, total, i, L);
                load(b, c1, total, i, L);
                load(s, s_ntt, N, i % N, L);
                mod_mul(t, s, b, mq);
                mod_sub(t, a, t, mq);
                store(x.data(), total, i, t, L);
            }
        });
        ntt_inverse(mq, log_n, n, x.data(), y.data(), ir, threads);
        if (plaintext_ntt)
        {
            std::vector<u32> fpr;
            if (!make_roots(fpr, mp, log_n, false))
                return 2;
            oracle_cast(mqp, mpp, total, y.data(), z.data());
            ntt_forward(mp, log_n, n, z.data(), out, fpr, threads);
        }
        else
            oracle_cast(mqp, mpp, total, y.data(), out);
        return 0;
    }

    // ------------------------------------------------------------------ CPU baselines
    // Batched polynomial multiplication INTT(NTT(a) o NTT(b)) over n pairs, hmpc's pass
    // structure, `threads` host threads. Returns wall seconds (tables excluded).
    double oracle_polymul(void* mp, int log_n, size_t n, u32 const* a, u32 const* b, u32* out, int threads)
    {
        auto& m = *static_cast<modulus*>(mp);
        int L = m.limbs;
        size_t total = n << log_n;
        std::vector<u32> fr, ir;
        if (!make_roots(fr, m, log_n, false) || !make_roots(ir, m, log_n, true))
            return -1.0;
        std::vector<u32> fa(total * L), fb(total * L), prod(total * L);
        auto start = std::chrono::steady_clock::now();
        ntt_forward(m, log_n, n, a, fa.data(), fr, threads);
        ntt_forward(m, log_n, n, b, fb.data(), fr, threads);
        oracle_pointwise(mp, 2, total, fa.data(), fb.data(), total, prod.data(), threads);
        ntt_inverse(m, log_n, n, prod.data(), out, ir, threads);
        auto stop = std::chrono::steady_clock::now();
        return std::chrono::duration<double>(stop - start).count();
    }
}
