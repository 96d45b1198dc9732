// Per-lane ChaCha20 (RFC 8439) for CDNA4: each lane owns a full 16-word state in VGPRs and
// runs the 20 rounds with 32-bit adds, xors and rotates (v_alignbit / v_lshl_or), no
// cross-lane traffic. Block layout and counter rule follow hmpc/crypto/chacha.hpp:100-166:
// state = [constants, key[8], counter, nonce[3]], counter += 1 after every block.
#pragma once

#include "modarith.hpp"

namespace hmpc_amd
{
    HMPC_HD u32 rotl32(u32 x, int r)
    {
        return (x << r) | (x >> (32 - r));
    }

#define HMPC_QR(a, b, c, d)      \
    a += b; d ^= a; d = rotl32(d, 16); \
    c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  \
    c += d; b ^= c; b = rotl32(b, 7);

    struct chacha_key
    {
        u32 k[8];
    };

    // One 64-byte block: out = rounds(state) + state.
    HMPC_HD void chacha20_block(u32 (&out)[16], chacha_key const& key, u32 counter, u32 n0, u32 n1, u32 n2)
    {
        u32 x0 = 0x61707865u, x1 = 0x3320646eu, x2 = 0x79622d32u, x3 = 0x6b206574u;
        u32 x4 = key.k[0], x5 = key.k[1], x6 = key.k[2], x7 = key.k[3];
        u32 x8 = key.k[4], x9 = key.k[5], x10 = key.k[6], x11 = key.k[7];
        u32 x12 = counter, x13 = n0, x14 = n1, x15 = n2;
#pragma unroll
        for (int r = 0; r < 10; ++r)
        {
            HMPC_QR(x0, x4, x8, x12)
            HMPC_QR(x1, x5, x9, x13)
            HMPC_QR(x2, x6, x10, x14)
            HMPC_QR(x3, x7, x11, x15)
            HMPC_QR(x0, x5, x10, x15)
            HMPC_QR(x1, x6, x11, x12)
            HMPC_QR(x2, x7, x8, x13)
            HMPC_QR(x3, x4, x9, x14)
        }
        out[0] = x0 + 0x61707865u;
        out[1] = x1 + 0x3320646eu;
        out[2] = x2 + 0x79622d32u;
        out[3] = x3 + 0x6b206574u;
        out[4] = x4 + key.k[0];
        out[5] = x5 + key.k[1];
        out[6] = x6 + key.k[2];
        out[7] = x7 + key.k[3];
        out[8] = x8 + key.k[4];
        out[9] = x9 + key.k[5];
        out[10] = x10 + key.k[6];
        out[11] = x11 + key.k[7];
        out[12] = x12 + counter;
        out[13] = x13 + n0;
        out[14] = x14 + n1;
        out[15] = x15 + n2;
    }
#undef HMPC_QR

    // centred binomial value from `bits` (= 4 * variance, 1..64) random bits starting at
    // words[first]: popcount of the masked words minus bits/2 (hmpc/random/binomial.hpp:30-123).
    template<int Bits>
    HMPC_HD int centered_binomial_from(u32 const* words)
    {
        constexpr int W = (Bits + 31) / 32;
        int pop = 0;
#pragma unroll
        for (int i = 0; i < W; ++i)
        {
            u32 w = words[i];
            if (i == W - 1 && (Bits % 32) != 0)
                w &= (1u << (Bits % 32)) - 1u;
            pop += __builtin_popcount(w);
        }
        return pop - Bits / 2;
    }

    // runtime-width variant for bits <= 64 random bits held in two words (the second is
    // ignored when bits <= 32); explicit words keep the block in VGPRs (no dynamic indexing)
    HMPC_HD int centered_binomial_words(u32 w0, u32 w1, int bits)
    {
        u32 m0 = bits >= 32 ? w0 : (w0 & ((1u << bits) - 1u));
        u32 m1 = bits <= 32 ? 0u : (bits == 64 ? w1 : (w1 & ((1u << (bits - 32)) - 1u)));
        return __builtin_popcount(m0) + __builtin_popcount(m1) - bits / 2;
    }

    // Width 0 on a capability stream selects the uniform ternary draw, which hmpc does not have:
    // t = floor(3 w / 2^64) - 1 for the 64-bit word w = w1 : w0. Each of -1, 0, 1 takes
    // floor or ceil of 2^64 / 3 values of w, so the draw is within 2^-63 of uniform.
    constexpr int ternary_bits = 0;

    HMPC_HD int ternary_words(u32 w0, u32 w1)
    {
        u64 w = (static_cast<u64>(w1) << 32) | w0;
        u64 twice = w << 1;
        int hi = static_cast<int>(w >> 63) + ((twice + w) < twice ? 1 : 0);  // high word of 2w + w
        return hi - 1;
    }

    // whether a capability draw of this width takes two words (ternary and widths above 32)
    HMPC_HD bool capability_two_words(int bits) { return bits == ternary_bits || bits > 32; }

    // one capability draw: the centred binomial of `bits` bits, or the ternary draw at width 0
    HMPC_HD int capability_small_words(u32 w0, u32 w1, int bits)
    {
        return bits == ternary_bits ? ternary_words(w0, w1) : centered_binomial_words(w0, w1, bits);
    }

    // Montgomery form of a small signed value v (|v| < 2^31): (v mod q) * R mod q.
    // hmpc/ints/mod.hpp:219-226: add q when negative, then multiply by R^2 and reduce.
    template<int L>
    HMPC_HD void mod_from_small_signed(u32 (&r)[L], int v, u32 const (&r2)[L], mod_params const& P)
    {
        u32 x[L];
        u32 mag = v < 0 ? static_cast<u32>(-static_cast<long long>(v)) : static_cast<u32>(v);
        if (v < 0)
        {
            u32 br = 0;
#pragma unroll
            for (int i = 0; i < L; ++i)
                x[i] = subb(P.q[i], i == 0 ? mag : 0u, br, br);
        }
        else
        {
#pragma unroll
            for (int i = 0; i < L; ++i)
                x[i] = i == 0 ? mag : 0u;
        }
        mod_mul<L>(r, x, r2, P);
    }
}
