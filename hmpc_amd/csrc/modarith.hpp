// Exact multi-limb modular arithmetic on 32-bit limbs for CDNA4 (gfx950) and the host.
//
// Values are canonical Montgomery residues a*R mod q in [0, q), R = 2^(32 L), exactly the
// representation of hmpc::ints::mod<q> (hmpc/ints/mod.hpp:73-379). Any reduction strategy
// is allowed internally as long as results come out canonical; the functions here use
// CIOS Montgomery multiplication (one v_mad_u64_u32 per limb product) and branch-free
// conditional subtraction (v_cndmask), the same constant-time discipline as
// hmpc/core/select.hpp:12-66.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#define HMPC_HD __host__ __device__ __forceinline__

namespace hmpc_amd
{
    using u32 = std::uint32_t;
    using u64 = std::uint64_t;

    constexpr int max_limbs = 9;

    // Kernel-argument view of a registered modulus (passed by value: lives in SGPRs).
    struct mod_params
    {
        u32 q[max_limbs];
        u32 nprime;  // -q^{-1} mod 2^32 (hmpc/ints/mod.hpp:98-103)
        int limbs;
    };

    template<int L>
    struct residue
    {
        u32 v[L];
    };

    // a + b + carry_in -> (sum, carry_out)
    HMPC_HD u32 addc(u32 a, u32 b, u32 carry_in, u32& carry_out)
    {
        u64 s = static_cast<u64>(a) + b + carry_in;
        carry_out = static_cast<u32>(s >> 32);
        return static_cast<u32>(s);
    }

    // a - b - borrow_in -> (diff, borrow_out)
    HMPC_HD u32 subb(u32 a, u32 b, u32 borrow_in, u32& borrow_out)
    {
        u64 d = static_cast<u64>(a) - b - borrow_in;
        borrow_out = static_cast<u32>(d >> 63);
        return static_cast<u32>(d);
    }

    // (a + b) mod q for canonical a, b
    template<int L>
    HMPC_HD void mod_add(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], mod_params const& P)
    {
        u32 s[L], d[L];
        u32 c = 0, br = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            s[i] = addc(a[i], b[i], c, c);
#pragma unroll
        for (int i = 0; i < L; ++i)
            d[i] = subb(s[i], P.q[i], br, br);
        // keep the difference when the sum carried out or did not underflow
        bool take_diff = (c != 0) | (br == 0);
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = take_diff ? d[i] : s[i];
    }

    // (a - b) mod q for canonical a, b
    template<int L>
    HMPC_HD void mod_sub(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], mod_params const& P)
    {
        u32 d[L];
        u32 br = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            d[i] = subb(a[i], b[i], br, br);
        u32 mask = 0u - br;
        u32 c = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = addc(d[i], P.q[i] & mask, c, c);
    }

    // Montgomery product a*b*R^{-1} mod q (CIOS), canonical output for canonical inputs.
    template<int L>
    HMPC_HD void mod_mul(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], mod_params const& P)
    {
        u32 t[L + 2];
#pragma unroll
        for (int i = 0; i < L + 2; ++i)
            t[i] = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
        {
            u32 c = 0;
#pragma unroll
            for (int j = 0; j < L; ++j)
            {
                u64 s = static_cast<u64>(a[j]) * b[i] + t[j] + c;
                t[j] = static_cast<u32>(s);
                c = static_cast<u32>(s >> 32);
            }
            u64 s = static_cast<u64>(t[L]) + c;
            t[L] = static_cast<u32>(s);
            t[L + 1] = static_cast<u32>(s >> 32);

            u32 m = t[0] * P.nprime;
            s = static_cast<u64>(m) * P.q[0] + t[0];
            c = static_cast<u32>(s >> 32);
#pragma unroll
            for (int j = 1; j < L; ++j)
            {
                s = static_cast<u64>(m) * P.q[j] + t[j] + c;
                t[j - 1] = static_cast<u32>(s);
                c = static_cast<u32>(s >> 32);
            }
            s = static_cast<u64>(t[L]) + c;
            t[L - 1] = static_cast<u32>(s);
            t[L] = t[L + 1] + static_cast<u32>(s >> 32);
        }
        // t < 2q: subtract q once if t >= q
        u32 d[L];
        u32 br = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            d[i] = subb(t[i], P.q[i], br, br);
        bool take_diff = (t[L] != 0) | (br == 0);
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = take_diff ? d[i] : t[i];
    }

    // Montgomery reduction of a wide value (W words, value < 2^(32 W)) with W word rounds:
    // result = value * 2^(-32 W) mod q, canonical. Used for the uniform sampler's drowning draw
    // (hmpc/ints/mod.hpp:210-215, where the round count is max(L, W) = W). Shift-register form:
    // each round adds m q (m = acc_0 N' mod 2^32 zeroes the low word) and drops the low word,
    // so the accumulator stays W words wide plus one carry word.
    template<int L, int W>
    HMPC_HD void montgomery_reduce_wide(u32 (&r)[L], u32 const (&value)[W], mod_params const& P)
    {
        static_assert(W > L, "wide reduction expects more words than limbs");
        u32 acc[W];
        u32 top = 0;  // word W of the accumulator
#pragma unroll
        for (int i = 0; i < W; ++i)
            acc[i] = value[i];
#pragma unroll
        for (int round = 0; round < W; ++round)
        {
            u32 m = acc[0] * P.nprime;
            u64 s = static_cast<u64>(m) * P.q[0] + acc[0];
            u64 c = s >> 32;
#pragma unroll
            for (int j = 1; j < W; ++j)
            {
                s = (j < L ? static_cast<u64>(m) * P.q[j] : u64{0}) + acc[j] + c;
                acc[j - 1] = static_cast<u32>(s);
                c = s >> 32;
            }
            s = static_cast<u64>(top) + c;
            acc[W - 1] = static_cast<u32>(s);
            top = static_cast<u32>(s >> 32);
        }
        // acc < 2q (acc fits L + 1 words): subtract q once if acc >= q
        u32 d[L];
        u32 br = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            d[i] = subb(acc[i], P.q[i], br, br);
        u32 high = acc[L] | top;
        bool take_diff = (high != 0) | (br == 0);
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = take_diff ? d[i] : acc[i];
    }

    template<int L>
    HMPC_HD void copy(u32 (&r)[L], u32 const (&a)[L])
    {
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = a[i];
    }

    // limb-planar loads/stores: limb l of element e at base[l * count + e]
    template<int L>
    HMPC_HD void load_planar(u32 (&x)[L], u32 const* __restrict__ base, std::size_t count, std::size_t e)
    {
#pragma unroll
        for (int l = 0; l < L; ++l)
            x[l] = base[l * count + e];
    }

    template<int L>
    HMPC_HD void store_planar(u32* __restrict__ base, std::size_t count, std::size_t e, u32 const (&x)[L])
    {
#pragma unroll
        for (int l = 0; l < L; ++l)
            base[l * count + e] = x[l];
    }
}
