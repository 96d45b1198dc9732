// Radix-2^26 redundant arithmetic for moduli q = 2^124 + c, 0 < c < 2^24 (the 125-bit primes
// of the workloads: 2^124 + 2^20 + 2^18 + 1 and 2^124 + 0x78001).
//
// Why: on gfx950 v_mad_u64_u32 runs at full rate, but 32-bit-limb Montgomery code spends most
// of its instructions on 64-bit register-pair moves and carry adds around each mad. With
// 26-bit limbs every partial product accumulates in place into a 64-bit column sum
// (acc = a * b + acc: one 64-bit mad, no carries), and additions are limb-wise.
//
// Representation: x = sum_j x_j 2^(26 j), j = 0..4, limbs x_j held in 32 bits (redundant). The
// transforms' values are signed digits: two's-complement limbs |x_j| < 2^31, so that their
// butterflies (x + t, x - t) and (x + y, x - y) need no added multiple of q; point-wise products
// fold one operand to non-negative limbs first. Internal Montgomery radix R' = 2^130: mont(a, b) = a b 2^-130 mod q. Data stays in
// hmpc's representation a*R mod q (R = 2^128); twiddles are stored as w * 2^130 mod q so that
// mont(a R, w 2^130) = (a w) R. A product of two data values is corrected by 4 = R' / R.
//
// Bounds (q > 2^124, so 64 q > 2^130). A signed-digit value with |x_j| < 2^31 for j < 4 and
// |x_4| < 2^30 - 2^6 lies within (|x_4| + 2^6) 2^104 < 1024 q of zero, so limb bounds bound values:
//   mont(a, b): a signed-digit (|a_j| < 2^31), b in [0, q) with limbs < 2^27: column sums stay
//               below 2^61 in magnitude; the result (T + M q) / 2^130 has digits r_0..r_3 in
//               [0, 2^26) and a signed top limb, and lies in (-|a| q / 2^130, |a| q / 2^130 +
//               q (1 + 2^-19)): within (-q, 2q (1 + 2^-19)) for |a| < 64 q, in [0, 2q (1 + 2^-19))
//               with r_4 < 2^22 for a >= 0, v(a) v(b) < 2^130 q; for any such a,
//               |r_4| < (|a_4| + 2^6) / 2^6 + 2^20 + 2^2. With b in [0, 2q) and limbs < 2^28
//               (a folded value) and |a_j| < 2^30.2 the sums stay below 2^61 as well and the result
//               lies in (-2|a| q / 2^130, 2|a| q / 2^130 + q (1 + 2^-19))
//   mont_tw(a, V): a times a twiddle w held in its expanded form V_i = 2^(26 (i + 1)) w mod q,
//               i = 0..4 (canonical limbs: < 2^26, top limb <= 2^20; V_4 is the stored w 2^130):
//               z = sum_i a_i V_i = a w 2^26 mod q collects in five columns, and one Montgomery
//               round divides by 2^26. For |a_j| <= A < 2^31 the columns stay below 2^59 in
//               magnitude; the result has digits 0..3 in [0, 2^26) and |r_4| < 5 A / 2^6 + 2^20
//               + 2^8: its value is bounded by a's limbs, not by a's value.
//   forward butterfly (x + t, x - t), t = mont(y, w): |value| grows by < 2q (1 + 2^-18) per stage,
//               limbs 0..3 by < 2^26 and limb 4 by < 2^22 while |y| < 64 q. From inputs in [0, q)
//               with limbs (< 2^27, < 2^26, < 2^26, < 2^26, < 2^22) (canonical data, or
//               from_signed samples) S stages stay below q + 2.0001 q S in magnitude with |limbs| <
//               (S + 2) x those: a mont input needs |limbs| < 2^31, so up to 29 stages (a whole
//               transform of N <= 2^29, both of its passes) run without reductions. With
//               t = mont_tw(y, V) limb 4 grows by < 5 A / 2^6 + 2^20 + 2^8 per stage instead
//               (A: y's largest limb); the product's schedule (the streaming column passes'
//               3-stage rounds and, at rows of 2^9, the row kernel's first forward and last
//               inverse rounds on mont_tw, the rest on mont) is checked stage by stage for
//               N = 2^14..2^17 in tests/test_r26_host.py: every mont and mont_tw input keeps
//               |limbs| < 2^30.2, every folded sum |x_4| < 2^29.6
//   fold(x):    normalize, then x_4 >= 2^20 carries h = x_4 >> 20 copies of 2^124 = q - c:
//               replace by -h c (plus one q in borrowed digits to keep limbs >= 0); result < 2q
//               with limbs < 2^28, for any limbs < 2^31. fold_signed adds 1024 q = 2^134 + 1024 c
//               first (limb 4 += 2^30), so that a signed-digit value (|x_j| < 2^31 - 2^26 for
//               j < 4, |x_4| < 2^30 - 2^6) folds the same way
//   inverse butterfly (x + y, mont(x - y, w)) on signed digits, x + y folded (signed) at stages
//               1, 4, 7, ... of the transform (r26_inverse_fold_stage). Folded values are in [0, 2q)
//               with limbs < 2^27.2 and mont outputs within (-q, 2q (1 + 2^-19)) with digits 0..3
//               < 2^26, so the sums of the two stages after a fold stay within 4.2q and 8.4q with
//               limbs < 2^28.2 and < 2^29.2, and the third stage's sum (< 16.8q, limbs < 2^30.2)
//               meets the signed fold's bounds; from point-wise products (a forward result,
//               |a| < 35.01q with |limbs| < 2^30.2 after at most 17 stages, times a folded one:
//               within (-4.4q, 8.4q), digits 0..3 < 2^28 and |top limb| < 2^24 after the x4
//               correction) stage 0 gives < 16.8q, and stage 1's sums stay within 33.6q with
//               |limbs| < 2^30. Every mont input stays within 33.6q with |limbs| < 2^30. With
//               mont_tw the top limbs of the sums grow faster (by the bound above); the fold
//               stages stay the same and the checked schedule keeps them within the signed fold.
#pragma once

#include "modarith.hpp"

#include <type_traits>

namespace hmpc_amd
{
    constexpr u32 m26 = (1u << 26) - 1u;

    struct p26
    {
        u32 c;         // q = 2^124 + c
        u32 n26;       // -q^-1 mod 2^26
        u32 qb[5];     // q with limbs qb_0 >= 2^26, qb_1 >= 2^10 (bias of the fold)
        u32 b1024[2];  // 1024 c as two radix-2^26 limbs (with 2^30 in limb 4: 1024 q, the signed fold's bias)
        u32 q32[4];    // q as 32-bit limbs (final canonical subtraction)
        u32 rr[5];     // R 2^130 mod q (R = 2^128), normalized: mont(z, rr) = z R mod q
        u32 t20;       // 2^20 (top radix-2^26 limb of q), kept in a register so that m * 2^20 + T
                       // stays one v_mad_u64_u32 instead of a 64-bit shift and a 64-bit add
    };

    // every constant but rr (which needs R^2 mod q) for q = 2^124 + c, 0 < c < 2^24 odd
    inline void r26_constants(p26& R, u32 c)
    {
        R.c = c;
        u32 inv = 1;
        for (int i = 0; i < 6; ++i)
            inv *= 2u - c * inv;  // c^-1 mod 2^32 (q = c mod 2^32)
        R.n26 = (0u - inv) & m26;
        u32 const qb[5] = {c + (1u << 26), (1u << 26) - 1, (1u << 26) - 1, (1u << 26) - 1, (1u << 20) - 1};
        for (int i = 0; i < 5; ++i)
        {
            R.qb[i] = qb[i];
        }
        u64 const c1024 = static_cast<u64>(c) << 10;  // < 2^34
        R.b1024[0] = static_cast<u32>(c1024) & m26;
        R.b1024[1] = static_cast<u32>(c1024 >> 26);
        R.q32[0] = c;
        R.q32[1] = 0;
        R.q32[2] = 0;
        R.q32[3] = 1u << 28;
        R.t20 = 1u << 20;
    }

    HMPC_HD void r26_unpack(u32 (&x)[5], u32 w0, u32 w1, u32 w2, u32 w3)
    {
        x[0] = w0 & m26;
        x[1] = ((w0 >> 26) | (w1 << 6)) & m26;
        x[2] = ((w1 >> 20) | (w2 << 12)) & m26;
        x[3] = ((w2 >> 14) | (w3 << 18)) & m26;
        x[4] = w3 >> 8;
    }

    // normalized limbs (x_j < 2^26 for j < 4) and value < 2^128 -> 4 words
    HMPC_HD void r26_pack(u32 (&w)[4], u32 const (&x)[5])
    {
        w[0] = x[0] | (x[1] << 26);
        w[1] = (x[1] >> 6) | (x[2] << 20);
        w[2] = (x[2] >> 12) | (x[3] << 14);
        w[3] = (x[3] >> 18) | (x[4] << 8);
    }

    // carry-propagate limbs 0..3 into limb 4 (limb 4 keeps the excess), limbs in [0, 2^32)
    HMPC_HD void r26_normalize(u32 (&x)[5])
    {
        x[1] += x[0] >> 26;
        x[0] &= m26;
        x[2] += x[1] >> 26;
        x[1] &= m26;
        x[3] += x[2] >> 26;
        x[2] &= m26;
        x[4] += x[3] >> 26;
        x[3] &= m26;
    }

    // the fold's reduction of a normalized value with x_4 in [0, 2^31 + 2^5): result in [0, 2q),
    // limbs < 2^28
    HMPC_HD void r26_fold_top(u32 (&x)[5], p26 const& P)
    {
        u32 h = x[4] >> 20;
        x[4] &= (1u << 20) - 1u;
        u64 hc = static_cast<u64>(h) * P.c;  // < 2^12 * 2^24
        u32 lo = static_cast<u32>(hc) & m26;
        u32 hi = static_cast<u32>(hc >> 26);
        x[0] = x[0] + P.qb[0] - lo;
        x[1] = x[1] + P.qb[1] - hi;
        x[2] += P.qb[2];
        x[3] += P.qb[3];
        x[4] += P.qb[4];
    }

    // value -> congruent value < 2q, limbs < 2^28, for any limbs < 2^31
    HMPC_HD void r26_fold(u32 (&x)[5], p26 const& P)
    {
        r26_normalize(x);  // x_4 < 2^31 + 2^5
        r26_fold_top(x, P);
    }

    // signed-digit value (|x_j| < 2^31 - 2^26 for j < 4, |x_4| < 2^30 - 2^6) -> congruent value in
    // [0, 2q), limbs < 2^28: 1024 q = 2^134 + 1024 c is added first (limb 4 += 2^30, 1024 c into
    // limbs 0 and 1), which makes the value positive; the signed carry chain then leaves digits
    // 0..3 in [0, 2^26) and limb 4 in [0, 2^31), the range of r26_fold_top
    HMPC_HD void r26_fold_signed(u32 (&x)[5], p26 const& P)
    {
        x[0] += P.b1024[0];
        x[1] += P.b1024[1];
        x[4] += 1u << 30;
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            x[j + 1] += static_cast<u32>(static_cast<int>(x[j]) >> 26);
            x[j] &= m26;
        }
        r26_fold_top(x, P);
    }

    // Montgomery product a b 2^-130 mod q. Signed: a has signed digits (|a_j| < 2^31), else limbs
    // in [0, 2^31); b has limbs in [0, 2^27). Result digits 0..3 in [0, 2^26); the top limb is
    // signed for a signed a, < 2^22 otherwise (bounds in the header).
    //
    // Product scanning: column k (weight 2^(26 k)) is accumulated after column k - 1 is finished,
    // starting from that column's carry, so every carry enters as the addend of a 64-bit mad
    // instead of a separate 64-bit add: 25 product mads, per reduction round one v_mul_lo_u32, the
    // m c mad and one 64-bit shift, and per result limb a mask and a shift (54 instructions against
    // 62 when all columns are summed first and carried afterwards; the results are identical).
    // An empty register constraint on a column sum: the compiler cannot reassociate the sum
    // across it, so each accumulation stays a mad whose addend is the running sum (left free, it
    // moves the carry to the end of the column as a separate 64-bit add). No-op on the host.
    template<typename T>
    HMPC_HD void r26_pin(T& t)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        asm("" : "+v"(t));
#else
        (void)t;
#endif
    }

    template<bool Signed>
    HMPC_HD void r26_mont_t(u32 (&r)[5], u32 const (&a)[5], u32 const (&b)[5], p26 const& P)
    {
        using acc = typename std::conditional<Signed, long long, u64>::type;
        // q in radix 2^26: limb 0 = c, limb 4 = 2^20, others 0. Only the last quotient digit is
        // masked to 26 bits: with m_0..m_3 < 2^32 the quotient M = sum m_i 2^(26 i) stays below
        // 2^130 + 2^111, so the result (T + M q) / 2^130 keeps its bound and each round saves the
        // mask. For a signed a, m q is added as unsigned 64-bit products onto the two's-complement
        // column sums (exact: no sum leaves (-2^62, 2^62)); the low 26 bits of T_k + m_k c are
        // zero, so the (arithmetic) shift carries exactly. Column k >= 4 also takes m_(k-4) 2^20,
        // the top limb of m_(k-4) q.
        u32 m[5];
        acc carry = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k)
        {
            acc t = carry;
#pragma unroll
            for (int i = (k > 4 ? k - 4 : 0); i <= (k < 4 ? k : 4); ++i)
            {
                if constexpr (Signed)
                    t = static_cast<long long>(static_cast<int>(a[i])) * static_cast<long long>(static_cast<int>(b[k - i])) + t;
                else
                    t = static_cast<u64>(a[i]) * b[k - i] + t;
                r26_pin(t);
            }
            if (k >= 4)
                t = static_cast<acc>(static_cast<u64>(m[k - 4]) * P.t20 + static_cast<u64>(t));
            if (k < 5)
            {
                u32 mk = static_cast<u32>(t) * P.n26;
                if (k == 4)
                    mk &= m26;
                m[k] = mk;
                t = static_cast<acc>(static_cast<u64>(mk) * P.c + static_cast<u64>(t));
            }
            else
                r[k - 5] = static_cast<u32>(t) & m26;
            carry = t >> 26;
        }
        r[4] = static_cast<u32>(carry);
    }

    // non-negative operands (inverse transforms, point-wise products of folded values)
    HMPC_HD void r26_mont(u32 (&r)[5], u32 const (&a)[5], u32 const (&b)[5], p26 const& P)
    {
        r26_mont_t<false>(r, a, b, P);
    }

    // a signed-digit first operand (forward-transform values)
    HMPC_HD void r26_mont_signed(u32 (&r)[5], u32 const (&a)[5], u32 const (&b)[5], p26 const& P)
    {
        r26_mont_t<true>(r, a, b, P);
    }

    // forward (Cooley-Tukey) butterfly on signed digits: (x, y) -> (x + t, x - t), t = mont(y, w)
    HMPC_HD void r26_forward(u32 (&x)[5], u32 (&y)[5], u32 const (&w)[5], p26 const& P)
    {
        u32 t[5];
        r26_mont_signed(t, y, w, P);
#pragma unroll
        for (int l = 0; l < 5; ++l)
        {
            u32 xl = x[l];
            x[l] = xl + t[l];
            y[l] = xl - t[l];
        }
    }

    // inverse (Gentleman-Sande) butterfly on signed digits: (x, y) -> (x + y, mont(x - y, w)), the
    // sum folded (to [0, 2q)) when `fold`
    HMPC_HD void r26_inverse(u32 (&x)[5], u32 (&y)[5], u32 const (&w)[5], p26 const& P, bool fold)
    {
        u32 d[5];
#pragma unroll
        for (int l = 0; l < 5; ++l)
        {
            d[l] = x[l] - y[l];
            x[l] = x[l] + y[l];
        }
        if (fold)
            r26_fold_signed(x, P);
        r26_mont_signed(y, d, w, P);
    }

    // the inverse stages whose sums x + y are folded (bounds in the header): 1, 4, 7, ...
    HMPC_HD bool r26_inverse_fold_stage(int i) { return i % 3 == 1; }

    // the inverse's last stage with its scaling: (x, y) -> (mont(x + y, w0), mont(x - y, w1))
    HMPC_HD void r26_inverse_last(u32 (&x)[5], u32 (&y)[5], u32 const (&w0)[5], u32 const (&w1)[5], p26 const& P)
    {
        u32 s[5], d[5];
#pragma unroll
        for (int l = 0; l < 5; ++l)
        {
            d[l] = x[l] - y[l];
            s[l] = x[l] + y[l];
        }
        r26_mont_signed(x, s, w0, P);
        r26_mont_signed(y, d, w1, P);
    }

    // ---- products by a twiddle in expanded form
    // V holds the twiddle w as five canonical values V_i = 2^(26 (i + 1)) w mod q, limb j of V_i at
    // V[5 i + j] (r26_expand_twiddle). a w 2^26 = sum_i a_i V_i (mod q) then collects in columns
    // 0..4 only: 25 mads, and a single Montgomery round (m = z_0 (-q^-1) mod 2^26, z + m q,
    // divided by 2^26) returns a w, 38 instructions against mont's 54. The cost is the twiddle's
    // size (25 words): for wave-uniform twiddles it comes from scalar loads and the mads take
    // its limbs from SGPRs. Bounds in the header.
    constexpr int r26_tw_words = 25;     // words of an expanded twiddle
    constexpr int r26_tw_stride = 32;    // table stride in words

    HMPC_HD void r26_mont_tw(u32 (&r)[5], u32 const (&a)[5], u32 const (&V)[r26_tw_words], p26 const& P)
    {
        long long t = 0;
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            t = static_cast<long long>(static_cast<int>(a[i])) * static_cast<long long>(static_cast<int>(V[5 * i])) + t;
            r26_pin(t);
        }
        // masked to 26 bits, so that m 2^20 adds less than 2^20 to the top limb
        u32 const m = (static_cast<u32>(t) * P.n26) & m26;
        t = static_cast<long long>(static_cast<u64>(m) * P.c + static_cast<u64>(t));
        long long carry = t >> 26;
#pragma unroll
        for (int k = 1; k < 5; ++k)
        {
            t = carry;
#pragma unroll
            for (int i = 0; i < 5; ++i)
            {
                t = static_cast<long long>(static_cast<int>(a[i])) * static_cast<long long>(static_cast<int>(V[5 * i + k])) + t;
                r26_pin(t);
            }
            if (k == 4)
                t = static_cast<long long>(static_cast<u64>(m) * P.t20 + static_cast<u64>(t));
            r[k - 1] = static_cast<u32>(t) & m26;
            carry = t >> 26;
        }
        r[4] = static_cast<u32>(carry);
    }

    // the butterflies with expanded twiddles (the same values as r26_forward and r26_inverse,
    // modulo q)
    HMPC_HD void r26_forward_tw(u32 (&x)[5], u32 (&y)[5], u32 const (&V)[r26_tw_words], p26 const& P)
    {
        u32 t[5];
        r26_mont_tw(t, y, V, P);
#pragma unroll
        for (int l = 0; l < 5; ++l)
        {
            u32 xl = x[l];
            x[l] = xl + t[l];
            y[l] = xl - t[l];
        }
    }

    HMPC_HD void r26_inverse_tw(u32 (&x)[5], u32 (&y)[5], u32 const (&V)[r26_tw_words], p26 const& P, bool fold)
    {
        u32 d[5];
#pragma unroll
        for (int l = 0; l < 5; ++l)
        {
            d[l] = x[l] - y[l];
            x[l] = x[l] + y[l];
        }
        if (fold)
            r26_fold_signed(x, P);
        r26_mont_tw(y, d, V, P);
    }

    // point-wise product of two forward-transform values a R, b R (signed digits) -> a b R: b is
    // folded to [0, 2q), a stays the signed operand, and the Montgomery product is corrected by
    // R' / R = 4. For |a| < 35.01 q with |a_j| < 2^30.2 (forward outputs of up to 17 stages) the
    // result lies in (-4.4 q, 8.4 q) with digits 0..3 in [0, 2^28) and |r_4| < 2^24.
    HMPC_HD void r26_point(u32 (&a)[5], u32 (&b)[5], p26 const& P)
    {
        r26_fold_signed(b, P);
        u32 t[5];
        r26_mont_signed(t, a, b, P);
#pragma unroll
        for (int l = 0; l < 5; ++l)
            a[l] = t[l] << 2;
    }

    // signed integer z -> a value congruent to z R mod q (R = 2^128) with limbs
    // (< 2^27, < 2^26, < 2^26, < 2^26, < 2^20) and value < q, without a Montgomery product:
    // 2^128 = 16 (q - c) = -16 c mod q, so z R = -16 c z, and t = 16 c |z| < 2^92 < q for any
    // |z| < 2^64. z < 0: t itself; z >= 0: q - t, formed limb-wise from q's borrowed digits qb
    // (qb_j >= t_j for every limb: no borrows).
    HMPC_HD void r26_from_signed(u32 (&x)[5], long long z, p26 const& P)
    {
        u64 az = z < 0 ? static_cast<u64>(0) - static_cast<u64>(z) : static_cast<u64>(z);
        u32 c16 = P.c << 4;  // < 2^28
        u64 t0 = static_cast<u64>(static_cast<u32>(az) & m26) * c16;
        u64 t1 = static_cast<u64>(static_cast<u32>(az >> 26) & m26) * c16 + (t0 >> 26);
        u64 t2 = static_cast<u64>(static_cast<u32>(az >> 52)) * c16 + (t1 >> 26);
        u32 t[5] = {static_cast<u32>(t0) & m26, static_cast<u32>(t1) & m26, static_cast<u32>(t2) & m26,
                    static_cast<u32>(t2 >> 26), 0u};
        bool neg = z < 0;
#pragma unroll
        for (int l = 0; l < 5; ++l)
            x[l] = neg ? t[l] : P.qb[l] - t[l];
    }

    // x (value < 2q after a fold, limbs < 2^28) -> canonical 32-bit limbs
    HMPC_HD void r26_canonical(u32 (&w)[4], u32 const (&xin)[5], p26 const& P)
    {
        u32 x[5] = {xin[0], xin[1], xin[2], xin[3], xin[4]};
        r26_normalize(x);
        r26_pack(w, x);
        u32 d[4];
        u32 br = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            d[i] = subb(w[i], P.q32[i], br, br);
        bool take = br == 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            w[i] = take ? d[i] : w[i];
    }

    // a twiddle table entry W = w 2^130 mod q (canonical radix-2^26 limbs) -> its expanded form
    // for r26_mont_tw: V_i = 2^(26 (i + 1)) w mod q is mont(W, 2^(26 (i + 1))) for i < 4 (a unit
    // limb as the second operand) and W itself for i = 4; each made canonical (fold, final
    // subtraction, unpacked again)
    HMPC_HD void r26_expand_twiddle(u32 (&V)[r26_tw_words], u32 const (&W)[5], p26 const& P)
    {
#pragma unroll
        for (int i = 0; i < 5; ++i)
        {
            u32 v[5] = {W[0], W[1], W[2], W[3], W[4]};
            if (i < 4)
            {
                u32 e[5] = {0u, 0u, 0u, 0u, 0u};
                e[i + 1] = 1u;
                r26_mont(v, W, e, P);
            }
            r26_fold(v, P);
            u32 w[4];
            r26_canonical(w, v, P);
            u32 x[5];
            r26_unpack(x, w[0], w[1], w[2], w[3]);
#pragma unroll
            for (int l = 0; l < 5; ++l)
                V[5 * i + l] = x[l];
        }
    }
}
