// Arithmetic "forms" for the fast transform path. Every form returns the same canonical
// residues at the end of a transform; they differ in how much reduction work they skip:
//
//   form_generic     canonical CIOS Montgomery, canonical butterflies (any odd q)
//   form_lazy        CIOS without the final subtraction, Harvey butterflies with values kept
//                    in [0, 4q) (forward) / [0, 2q) (inverse); needs 4q < R
//   form_sparse      form_lazy, and the CIOS reduction skips limbs q_1..q_{L-2} known to be
//                    zero (q = q_0 + q_{L-1} 2^(32(L-1)), e.g. 2^124 + 2^20 + 2^18 + 1)
//   form_goldilocks  L = 2, q = 2^64 - 2^32 + 1: Montgomery reduction by shifts and adds
//                    (q^-1 = 2^32 + 1 mod 2^64), no multiplies in the reduction
//
// Bounds used by the lazy forms (R = 2^(32 L), q < R / 4):
//   mont_lazy(a, b) = (a b + m q) / R < a b / R + q  <  2q   whenever a b < 4 q^2
#pragma once

#include "modarith.hpp"

namespace hmpc_amd
{
    enum arith_form : int
    {
        form_generic = 0,
        form_lazy = 1,
        form_sparse = 2,
        form_goldilocks = 3,
        form_r26 = 4,  // q = 2^124 + c, c < 2^24: radix-2^26 redundant limbs (r26.hpp)
    };

    struct fast_params
    {
        u32 q[max_limbs];
        u32 q2[max_limbs];  // 2q (lazy forms)
        u32 r2[max_limbs];  // R^2 mod q (conversion of integers to Montgomery form)
        u32 nprime;
    };

    template<int L>
    HMPC_HD bool ge_limbs(u32 const (&a)[L], u32 const* b)
    {
        u32 br = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            (void)subb(a[i], b[i], br, br);
        return br == 0;
    }

    // r = a >= b ? a - b : a  (b given as L words)
    template<int L>
    HMPC_HD void cond_sub(u32 (&r)[L], u32 const (&a)[L], u32 const* b)
    {
        u32 d[L];
        u32 br = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            d[i] = subb(a[i], b[i], br, br);
        bool take = br == 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = take ? d[i] : a[i];
    }

    template<int L>
    HMPC_HD void add_plain(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L])
    {
        u32 c = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = addc(a[i], b[i], c, c);
    }

    // r = a - b + k (k = 2q), all < 2^(32 L); the result is positive for a >= 0, b < k
    template<int L>
    HMPC_HD void sub_plus(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], u32 const* k)
    {
        u32 t[L];
        u32 c = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            t[i] = addc(a[i], k[i], c, c);
        u32 br = 0;
#pragma unroll
        for (int i = 0; i < L; ++i)
            r[i] = subb(t[i], b[i], br, br);
    }

    // CIOS Montgomery product; Sparse skips q_1..q_{L-2}; Lazy skips the final subtraction.
    template<int L, bool Sparse, bool Lazy>
    HMPC_HD void mont_cios(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], fast_params const& P)
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
                if (Sparse && j < L - 1)
                    s = static_cast<u64>(t[j]) + c;
                else
                    s = static_cast<u64>(m) * P.q[j] + t[j] + c;
                t[j - 1] = static_cast<u32>(s);
                c = static_cast<u32>(s >> 32);
            }
            s = static_cast<u64>(t[L]) + c;
            t[L - 1] = static_cast<u32>(s);
            t[L] = t[L + 1] + static_cast<u32>(s >> 32);
        }
        if (Lazy)
        {
            // t < 2q < R: t[L] == 0
#pragma unroll
            for (int i = 0; i < L; ++i)
                r[i] = t[i];
        }
        else
        {
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
    }

    // Goldilocks Montgomery product (R = 2^64), canonical for canonical inputs:
    //   x = a b (128 bits); m = x_lo (1 + 2^32) mod 2^64 (= x_lo q^-1);
    //   high word of m q = m - (m >> 32) - [carry of x_lo + (x_lo << 32)];
    //   result = x_hi - high(m q) (mod q), in [0, q) because x_hi < q and high(m q) < q.
    HMPC_HD void mont_goldilocks(u32 (&r)[2], u32 const (&a)[2], u32 const (&b)[2])
    {
        u64 t = static_cast<u64>(a[0]) * b[0];
        u32 x0 = static_cast<u32>(t);
        t = static_cast<u64>(a[1]) * b[0] + (t >> 32);
        u32 p1 = static_cast<u32>(t);
        u32 p2 = static_cast<u32>(t >> 32);
        t = static_cast<u64>(a[0]) * b[1] + p1;
        u32 x1 = static_cast<u32>(t);
        t = static_cast<u64>(a[1]) * b[1] + p2 + (t >> 32);
        u64 xl = (static_cast<u64>(x1) << 32) | x0;
        u64 xh = t;
        u64 s = xl + (xl << 32);
        u64 e = s < xl ? 1u : 0u;
        u64 hb = s - (s >> 32) - e;
        u64 d = xh - hb;
        u64 borrow = xh < hb ? 0xffffffffull : 0ull;
        d -= borrow;
        r[0] = static_cast<u32>(d);
        r[1] = static_cast<u32>(d >> 32);
    }

    template<int L, int Form>
    struct fast_ops
    {
        static constexpr bool lazy = Form == form_lazy || Form == form_sparse;

        HMPC_HD static void mul(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], fast_params const& P)
        {
            if constexpr (Form == form_goldilocks)
                mont_goldilocks(r, a, b);
            else
                mont_cios<L, Form == form_sparse, lazy>(r, a, b, P);
        }

        HMPC_HD static void add(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], fast_params const& P)
        {
            u32 s[L], d[L];
            u32 c = 0, br = 0;
#pragma unroll
            for (int i = 0; i < L; ++i)
                s[i] = addc(a[i], b[i], c, c);
#pragma unroll
            for (int i = 0; i < L; ++i)
                d[i] = subb(s[i], P.q[i], br, br);
            bool take_diff = (c != 0) | (br == 0);
#pragma unroll
            for (int i = 0; i < L; ++i)
                r[i] = take_diff ? d[i] : s[i];
        }

        HMPC_HD static void sub(u32 (&r)[L], u32 const (&a)[L], u32 const (&b)[L], fast_params const& P)
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

        // forward (CT) butterfly: (x, y) -> (x + y w, x - y w)
        // lazy: x, y in [0, 4q) -> outputs in [0, 4q)
        HMPC_HD static void fwd(u32 (&x)[L], u32 (&y)[L], u32 const (&w)[L], fast_params const& P)
        {
            u32 t[L];
            mul(t, y, w, P);
            if constexpr (lazy)
            {
                u32 xr[L];
                cond_sub<L>(xr, x, P.q2);
                add_plain<L>(x, xr, t);
                sub_plus<L>(y, xr, t, P.q2);
            }
            else
            {
                u32 xo[L];
                copy<L>(xo, x);
                add(x, xo, t, P);
                sub(y, xo, t, P);
            }
        }

        // inverse (GS) butterfly: (x, y) -> (x + y, (x - y) w)
        // lazy: x, y in [0, 2q) -> outputs in [0, 2q)
        HMPC_HD static void inv(u32 (&x)[L], u32 (&y)[L], u32 const (&w)[L], fast_params const& P)
        {
            if constexpr (lazy)
            {
                u32 s[L], d[L];
                add_plain<L>(s, x, y);
                sub_plus<L>(d, x, y, P.q2);
                cond_sub<L>(x, s, P.q2);
                mul(y, d, w, P);
            }
            else
            {
                u32 s[L], d[L];
                add(s, x, y, P);
                sub(d, x, y, P);
                copy<L>(x, s);
                mul(y, d, w, P);
            }
        }

        // last inverse stage: (x, y) -> ((x + y) w0, (x - y) w1)
        HMPC_HD static void inv_last(u32 (&x)[L], u32 (&y)[L], u32 const (&w0)[L], u32 const (&w1)[L], fast_params const& P)
        {
            u32 s[L], d[L];
            if constexpr (lazy)
            {
                add_plain<L>(s, x, y);
                sub_plus<L>(d, x, y, P.q2);
            }
            else
            {
                add(s, x, y, P);
                sub(d, x, y, P);
            }
            mul(x, s, w0, P);
            mul(y, d, w1, P);
        }

        // [0, 4q) -> [0, 2q)  (identity for canonical forms)
        HMPC_HD static void half_reduce(u32 (&x)[L], fast_params const& P)
        {
            if constexpr (lazy)
            {
                u32 t[L];
                cond_sub<L>(t, x, P.q2);
                copy<L>(x, t);
            }
        }

        // [0, 4q) -> [0, q)
        HMPC_HD static void canonical(u32 (&x)[L], fast_params const& P)
        {
            if constexpr (lazy)
            {
                u32 t[L];
                cond_sub<L>(t, x, P.q2);
                cond_sub<L>(x, t, P.q);
            }
        }
    };
}
