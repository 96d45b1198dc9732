// Fast transform path for the special-form moduli of the workloads, gfx950:
//   arith_r26   q = 2^124 + c (4 limbs in memory), radix-2^26 redundant arithmetic (r26.hpp)
//   arith_gold  q = 2^64 - 2^32 + 1 (2 limbs), Goldilocks Montgomery reduction (fastmath.hpp)
// Included by kernels_fast_l2.hip and kernels_fast_l4.hip, which instantiate one policy each
// (separate translation units build in parallel).
//
// A workgroup owns a 4096-element tile in LDS (element-major, XOR-swizzled so that strided
// lanes hit distinct banks). Each of its 4096 / E threads keeps E elements (E = 4, 8 or 16, a
// tuning knob) in VGPRs and runs up to log2(E) radix-2 stages on them between two LDS
// exchanges ("rounds"), so an S-point sub-transform costs ceil(log S / log E) exchanges instead
// of log S. At the start of a round every thread issues the loads of all twiddles the round
// needs (element-major tables in the policy's format), so their latencies overlap.
//
// Round geometry (s = log S, round = stages [a, b), w = b - a). A "unit" is the group of
// 2^w elements that interact in the round; a thread handles E / 2^w units.
//   forward (CT, stage i pairs bit s-1-i):  j = F << (s-a) | k << (s-b) | Lo,  F: top a bits
//       twiddle index  (base << i) + (F << (i-a) | k >> (b-i))
//   inverse (GS, stage i pairs bit i):      j = F << b | k << a | Lo,          Lo: low a bits
//       twiddle index  base * (S >> (i+1)) + (F << (b-i-1) | k >> (i-a+1))
// base = 1 for column passes and whole-polynomial tiles, N1 + hi for the row pass of row hi
// (the two-pass decomposition is described in kernels.hip). Forward rounds are full-width
// first and partial last, inverse rounds partial first, so the forward's last round and the
// inverse's first round group the same elements: a product stays in registers between them.
#pragma once

#include "kernels.hpp"
#include "fastmath.hpp"
#include "r26.hpp"
#include "tile_geometry.hpp"

#include <atomic>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

namespace hmpc_amd
{
    namespace fast
    {
        // Thread index as an opaque value per use. Every per-thread LDS / twiddle address below is
        // a function of it; kept opaque, that arithmetic is recomputed where it is used (a few
        // integer ops) instead of being hoisted out of the job kernel's tile loops and held in
        // VGPRs across them, which spilled registers.
        __device__ __forceinline__ int tid_x()
        {
            int t;
            asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"(static_cast<int>(threadIdx.x)));
            return t;
        }

        // The thread's index in its 4096-element tile: the thread index, or for one-wave workgroups
        // (wave_blocks) the lane plus 64 times the wave's place in its tile (blockIdx.x modulo the
        // tile's waves)
        template<typename A, int E>
        __device__ __forceinline__ int tile_thread()
        {
            if constexpr (A::wave_blocks)
                return tid_x() + static_cast<int>((blockIdx.x % static_cast<u32>(tile / E / 64)) << 6);
            else
                return tid_x();
        }

        // Global accesses of the policies: `base` points at element 0 of a tile (or of an operand's
        // elements) in limb plane 0, planes `count` words apart, and `off` is the element's offset
        // from it (below 2^28 for every tile geometry). The plane pointers are uniform, so the
        // accesses address a scalar base plus a 32-bit vector byte offset shared by all limbs.
        __device__ __forceinline__ u32 ld_off(u32 const* __restrict__ p, u32 off)
        {
            return *reinterpret_cast<u32 const*>(reinterpret_cast<char const*>(p) + (off << 2));
        }

        __device__ __forceinline__ void st_off(u32* __restrict__ p, u32 off, u32 v)
        {
            *reinterpret_cast<u32*>(reinterpret_cast<char*>(p) + (off << 2)) = v;
        }

        // ------------------------------------------------------------------ arithmetic policies
        // r26: registers and LDS hold 5 limbs of 26 bits per element; global memory keeps hmpc's
        // 4 x 32-bit limb-planar layout (converted once on load and once on store per kernel).
        struct arith_r26
        {
            static constexpr int rl = 5;    // limbs per element in registers and LDS
            static constexpr int twl = rl;  // words per twiddle in registers
            static constexpr int iol = 4;   // limbs per element in global memory (hmpc layout)
            static constexpr int wsl = 5;   // limbs per element in private workspaces (raw limbs)
            static constexpr int tws = 8;   // twiddle table stride in words (5 used)
            static constexpr bool stage_prefetch = true;
            static constexpr bool wave_units = false;  // unit order (unit_index)
            static constexpr bool wave_blocks = false;  // one wave per workgroup (tile_thread)
            static constexpr bool mont_plaintext = false;  // BGV: p v + m formed as a signed 64-bit integer
            using params = p26;

            __device__ __forceinline__ static void load_global(u32 (&x)[rl], u32 const* __restrict__ src, std::size_t count,
                                                               u32 g)
            {
                r26_unpack(x, ld_off(src, g), ld_off(src + count, g), ld_off(src + 2 * count, g), ld_off(src + 3 * count, g));
            }

            // canonical: final output; otherwise a lazy value < 2q for the next pass
            __device__ __forceinline__ static void store_global(u32* __restrict__ dst, std::size_t count, u32 g,
                                                                u32 const (&xin)[rl], bool canonical, params const& P)
            {
                u32 x[rl] = {xin[0], xin[1], xin[2], xin[3], xin[4]};
                r26_fold_signed(x, P);  // forward results are signed digits
                u32 w[4];
                if (canonical)
                    r26_canonical(w, x, P);
                else
                {
                    r26_normalize(x);
                    r26_pack(w, x);
                }
#pragma unroll
                for (int l = 0; l < 4; ++l)
                    st_off(dst + l * count, g, w[l]);
            }

            // private workspaces between passes hold the raw redundant limbs (no reduction, no
            // repacking): 25% more bytes, about 50 fewer VALU operations per element and pass. A
            // forward column result is folded first only when the whole transform is longer
            // than the lazy forward bound allows (fold = true).
            __device__ __forceinline__ static void load_ws(u32 (&x)[rl], u32 const* __restrict__ src, std::size_t count,
                                                           u32 g)
            {
#pragma unroll
                for (int l = 0; l < rl; ++l)
                    x[l] = ld_off(src + l * count, g);
            }

            __device__ __forceinline__ static void store_ws(u32* __restrict__ dst, std::size_t count, u32 g,
                                                            u32 const (&xin)[rl], bool fold, params const& P)
            {
                u32 x[rl] = {xin[0], xin[1], xin[2], xin[3], xin[4]};
                if (fold)
                    r26_fold_signed(x, P);
#pragma unroll
                for (int l = 0; l < rl; ++l)
                    st_off(dst + l * count, g, x[l]);
            }

            // LDS slot of tile element e (an XOR swizzle: linear over GF(2), so the slot of e | d for
            // disjoint bits is phys(e) ^ phys(d))
            __device__ __forceinline__ static int phys(int e) { return swizzle_16(e); }

            __device__ __forceinline__ static void lds_get_at(u32 (&x)[rl], u32 const* lds, int slot)
            {
                int p = slot * rl;
#pragma unroll
                for (int l = 0; l < rl; ++l)
                    x[l] = lds[p + l];
            }

            __device__ __forceinline__ static void lds_put_at(u32* lds, int slot, u32 const (&x)[rl])
            {
                int p = slot * rl;
#pragma unroll
                for (int l = 0; l < rl; ++l)
                    lds[p + l] = x[l];
            }

            __device__ __forceinline__ static void lds_put(u32* lds, int e, u32 const (&x)[rl]) { lds_put_at(lds, phys(e), x); }

            __device__ __forceinline__ static void tw_get(u32 (&w)[rl], u32 const* __restrict__ tw, u32 idx)
            {
                uint4 v = reinterpret_cast<uint4 const*>(tw)[2 * idx];
                w[0] = v.x;
                w[1] = v.y;
                w[2] = v.z;
                w[3] = v.w;
                w[4] = tw[8 * static_cast<std::size_t>(idx) + 4];
            }

            // the same entry from a 32-bit byte offset off the uniform table base (tables stay far
            // below 4 GiB): one shift per twiddle instead of two 64-bit address computations. The
            // row kernels use it (product rows 2.3% faster); the column kernels ran 3-4% slower.
            __device__ __forceinline__ static void tw_get_offset32(u32 (&w)[rl], u32 const* __restrict__ tw, u32 idx)
            {
                char const* base = reinterpret_cast<char const*>(tw);
                u32 const off = idx << 5;
                uint4 v = *reinterpret_cast<uint4 const*>(base + off);
                w[0] = v.x;
                w[1] = v.y;
                w[2] = v.z;
                w[3] = v.w;
                w[4] = *reinterpret_cast<u32 const*>(base + off + 16u);
            }

            // (x, y) -> (x + t, x - t), t = mont(y, w), on signed digits
            __device__ __forceinline__ static void fwd(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[rl], params const& P)
            {
                r26_forward(x, y, w, P);
            }

            // (x, y) -> (x + y, mont(x - y, w)) on signed digits, x + y folded when `fold` (odd stages)
            __device__ __forceinline__ static void inv(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[rl], params const& P,
                                                       bool fold)
            {
                r26_inverse(x, y, w, P, fold);
            }

            __device__ __forceinline__ static void inv_last(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w0)[rl],
                                                            u32 const (&w1)[rl], params const& P)
            {
                r26_inverse_last(x, y, w0, w1, P);
            }

            // point-wise product of two forward results a R, b R (signed digits) -> a b R, only b
            // folded (bounds with r26_point)
            __device__ __forceinline__ static void point(u32 (&a)[rl], u32 (&b)[rl], params const& P)
            {
                r26_point(a, b, P);
            }

            // x = k R + x for k canonical (limbs < 2^26) and u, x forward-transform outputs (signed
            // digits, |value| < 35 q, |limbs| < 19 2^26): mont(u, k) lies in (-q, 2q) without folding
            // u; 4 mont(u, k) adds digits < 2^28 (a signed top limb < 2^24 in magnitude)
            __device__ __forceinline__ static void point_add(u32 (&x)[rl], u32 const (&u)[rl], u32 const (&k)[rl],
                                                             params const& P)
            {
                u32 t[rl];
                r26_mont_signed(t, u, k, P);
#pragma unroll
                for (int l = 0; l < rl; ++l)
                    x[l] += t[l] << 2;
            }

            // signed integer -> z R mod q as a transform input (< q, limbs < 2^27)
            __device__ __forceinline__ static void from_signed(u32 (&x)[rl], long long z, params const& P)
            {
                r26_from_signed(x, z, P);
            }
        };

        // the same arithmetic with all of a round's twiddles loaded up front: for kernels that run
        // one workgroup per CU anyway (the persistent job kernel), where the loads' latency is not
        // covered by another workgroup
        struct arith_r26_round_prefetch : arith_r26
        {
            static constexpr bool stage_prefetch = false;
        };

        // Rounds whose twiddles are wave-uniform (the streaming column passes' 3-stage rounds, the
        // one-wave row kernel's forward first and inverse last rounds): twiddles in the expanded
        // form of r26_mont_tw (25 words per entry of the expanded tables, through scalar loads, the
        // mads taking them from SGPRs), 38 instead of 54 instructions per product. Elements, LDS
        // and global accesses as the base policy.
        template<typename Base>
        struct arith_tw : Base
        {
            using typename Base::params;
            static_assert(std::is_same<params, p26>::value, "expanded twiddles are a radix-2^26 form");
            static constexpr int rl = Base::rl;
            static constexpr int twl = r26_tw_words;

            // read through the constant address space: the tables are read-only in every kernel, and
            // a uniform index then gives scalar loads also where the table pointer comes from memory
            // (a job table), not only from a restrict kernel argument
            __device__ __forceinline__ static void tw_get(u32 (&w)[twl], u32 const* __restrict__ tw, u32 idx)
            {
                using const_u32 = __attribute__((address_space(4))) u32 const;
                const_u32* p = (const_u32*)(tw) + static_cast<std::size_t>(idx) * r26_tw_stride;  // address-space cast
#pragma unroll
                for (int l = 0; l < twl; ++l)
                    w[l] = p[l];
            }

            __device__ __forceinline__ static void fwd(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[twl], params const& P)
            {
                r26_forward_tw(x, y, w, P);
            }

            __device__ __forceinline__ static void inv(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[twl], params const& P,
                                                       bool fold)
            {
                r26_inverse_tw(x, y, w, P, fold);
            }

            // the scaled last stage keeps mont with the stored twiddles (V_4 = w 2^130 of entries 0
            // and 1; the loads of V_0..V_3 fall away): its two twiddles are the same for every tile,
            // and held across a tile loop they take 10 SGPRs instead of 50, which the other
            // stages' twiddles need (with 50 the inverse column pass spilled SGPRs to VGPR lanes)
            __device__ __forceinline__ static void inv_last(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w0)[twl],
                                                            u32 const (&w1)[twl], params const& P)
            {
                u32 const s0[rl] = {w0[20], w0[21], w0[22], w0[23], w0[24]};
                u32 const s1[rl] = {w1[20], w1[21], w1[22], w1[23], w1[24]};
                r26_inverse_last(x, y, s0, s1, P);
            }
        };

        using arith_r26_tw = arith_tw<arith_r26>;

        // The LDS slot order of the row kernels (rows of 2^7 .. 2^10 elements). With 64 banks of 4
        // bytes, element slots of 5 words and E = 8 every gather and scatter of every round of these
        // rows, and the 64 consecutive elements of a tile load or store, reach 64 distinct banks
        // (e ^ ((e >> 4) & 15) leaves 2- and 4-way conflicts in most of their rounds); at E = 4 and
        // 16 it is at most 2-way.
        struct arith_r26_rows : arith_r26
        {
            static constexpr bool wave_units = true;  // wave-local row tiles (rows of <= 64 E elements)

            __device__ __forceinline__ static void tw_get(u32 (&w)[rl], u32 const* __restrict__ tw, u32 idx)
            {
                tw_get_offset32(w, tw, idx);
            }

            __device__ __forceinline__ static int phys(int e) { return swizzle_rows(e); }

            __device__ __forceinline__ static void lds_put(u32* lds, int e, u32 const (&x)[rl]) { lds_put_at(lds, phys(e), x); }
        };

        // The product's row kernel with one wave per workgroup: its rows are wave-local, so each
        // wave-sized block of a tile (64 E elements, wave b % (tile / 64 E) of tile b / (tile / 64 E))
        // runs as a workgroup of its own with a 64 E-slot LDS block (10 KiB at E = 8). Workgroups of
        // eight waves load, transform and store in step; single waves start and end independently, so
        // the memory phases of some overlap the arithmetic of others. A wave's tile elements share
        // their bits above log2(64 E), which the slot order keeps (its XOR terms reach bit 5 only), so
        // the slot of element e is swizzle_rows(e) modulo 64 E: a one-to-one map of the wave's elements
        // onto its block, still XOR-linear.
        struct arith_r26_rows1 : arith_r26_rows
        {
            static constexpr bool wave_blocks = true;

            __device__ __forceinline__ static int phys(int e) { return swizzle_rows(e) & 511; }

            __device__ __forceinline__ static void lds_put(u32* lds, int e, u32 const (&x)[rl]) { lds_put_at(lds, phys(e), x); }
        };

        using arith_r26_rows1_tw = arith_tw<arith_r26_rows1>;

        // the BGV row kernels: twiddles from 32-bit offsets as in the product's row kernel
        struct arith_r26_bgv_rows : arith_r26
        {
            __device__ __forceinline__ static void tw_get(u32 (&w)[rl], u32 const* __restrict__ tw, u32 idx)
            {
                tw_get_offset32(w, tw, idx);
            }
        };

        // the BGV row kernel with one wave per workgroup (as arith_r26_rows1 for the product)
        struct arith_r26_bgv_rows1 : arith_r26_bgv_rows
        {
            static constexpr bool wave_units = true;
            static constexpr bool wave_blocks = true;
            // whole-round twiddle loads: 146 VGPRs, still 3 waves per SIMD, rows 0.66 -> 0.65 ms
            static constexpr bool stage_prefetch = false;

            __device__ __forceinline__ static int phys(int e) { return swizzle_rows(e) & 511; }

            __device__ __forceinline__ static void lds_put(u32* lds, int e, u32 const (&x)[rl]) { lds_put_at(lds, phys(e), x); }
        };

        // the arithmetic of the product's row kernel: the radix-2^26 rows are wave-local in the row
        // kernels' LDS slot order (twiddles stage by stage: with signed-digit forward butterflies
        // this ran 2.4% faster than whole-round prefetch)
        template<typename A>
        struct rows_arith
        {
            using type = A;
        };
        template<>
        struct rows_arith<arith_r26>
        {
            using type = arith_r26_rows;
        };

        template<typename A>
        struct bgv_rows_arith
        {
            using type = A;
        };
        template<>
        struct bgv_rows_arith<arith_r26>
        {
            using type = arith_r26_bgv_rows;
        };

        struct arith_gold
        {
            static constexpr int rl = 2;
            static constexpr int twl = rl;
            static constexpr int iol = 2;
            static constexpr int wsl = 2;
            static constexpr int tws = 2;
            static constexpr bool stage_prefetch = false;
            static constexpr bool wave_units = false;
            static constexpr bool wave_blocks = false;
            static constexpr bool mont_plaintext = false;
            using params = fast_params;
            using ops = fast_ops<2, form_goldilocks>;

            __device__ __forceinline__ static void load_global(u32 (&x)[rl], u32 const* __restrict__ src, std::size_t count,
                                                               u32 g)
            {
                x[0] = ld_off(src, g);
                x[1] = ld_off(src + count, g);
            }

            __device__ __forceinline__ static void store_global(u32* __restrict__ dst, std::size_t count, u32 g,
                                                                u32 const (&x)[rl], bool, params const&)
            {
                st_off(dst, g, x[0]);
                st_off(dst + count, g, x[1]);
            }

            // values are canonical throughout: workspaces use the global format
            __device__ __forceinline__ static void load_ws(u32 (&x)[rl], u32 const* __restrict__ src, std::size_t count,
                                                           u32 g)
            {
                load_global(x, src, count, g);
            }

            __device__ __forceinline__ static void store_ws(u32* __restrict__ dst, std::size_t count, u32 g,
                                                            u32 const (&x)[rl], bool, params const& P)
            {
                store_global(dst, count, g, x, false, P);
            }

            __device__ __forceinline__ static int phys(int e) { return swizzle_16(e); }

            __device__ __forceinline__ static void lds_get_at(u32 (&x)[rl], u32 const* lds, int slot)
            {
                uint2 v = reinterpret_cast<uint2 const*>(lds)[slot];
                x[0] = v.x;
                x[1] = v.y;
            }

            __device__ __forceinline__ static void lds_put_at(u32* lds, int slot, u32 const (&x)[rl])
            {
                reinterpret_cast<uint2*>(lds)[slot] = make_uint2(x[0], x[1]);
            }

            __device__ __forceinline__ static void lds_put(u32* lds, int e, u32 const (&x)[rl]) { lds_put_at(lds, phys(e), x); }

            __device__ __forceinline__ static void tw_get(u32 (&w)[rl], u32 const* __restrict__ tw, u32 idx)
            {
                uint2 v = reinterpret_cast<uint2 const*>(tw)[idx];
                w[0] = v.x;
                w[1] = v.y;
            }

            __device__ __forceinline__ static void fwd(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[rl], params const& P)
            {
                ops::fwd(x, y, w, P);
            }

            __device__ __forceinline__ static void inv(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[rl], params const& P,
                                                       bool)
            {
                ops::inv(x, y, w, P);
            }

            __device__ __forceinline__ static void inv_last(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w0)[rl],
                                                            u32 const (&w1)[rl], params const& P)
            {
                ops::inv_last(x, y, w0, w1, P);
            }

            __device__ __forceinline__ static void point(u32 (&a)[rl], u32 (&b)[rl], params const& P)
            {
                u32 t[rl];
                ops::mul(t, a, b, P);
                a[0] = t[0];
                a[1] = t[1];
            }

            // x = k u + x (canonical)
            __device__ __forceinline__ static void point_add(u32 (&x)[rl], u32 const (&u)[rl], u32 const (&k)[rl],
                                                             params const& P)
            {
                u32 t[rl], s[rl];
                ops::mul(t, k, u, P);
                ops::add(s, t, x, P);
                x[0] = s[0];
                x[1] = s[1];
            }

            // signed integer |z| < 2^63 -> z R mod q (canonical): |z| < q is canonical, so the
            // Montgomery product with R^2 is |z| R; negatives are q - that (0 stays 0)
            __device__ __forceinline__ static void from_signed(u32 (&x)[rl], long long z, params const& P)
            {
                u64 az = z < 0 ? static_cast<u64>(-z) : static_cast<u64>(z);
                u32 a[rl] = {static_cast<u32>(az), static_cast<u32>(az >> 32)};
                u32 r2[rl] = {P.r2[0], P.r2[1]};
                u32 t[rl], zero[rl] = {0u, 0u}, n[rl];
                ops::mul(t, a, r2, P);
                ops::sub(n, zero, t, P);
                bool neg = z < 0;
                x[0] = neg ? n[0] : t[0];
                x[1] = neg ? n[1] : t[1];
            }
        };

        // Any other modulus with 4q < R (R = 2^(32 L)) at 4 or 5 limbs, e.g. the 125-bit prime of
        // hmpc's examples/ntt.cpp and the 150-bit prime of examples/bgv.cpp: CIOS Montgomery
        // products on 32-bit limbs (fastmath.hpp, the lazy form) with Harvey butterflies, values in
        // [0, 4q) through forward transforms and [0, 2q) through inverse ones. Registers, LDS,
        // workspaces and global memory all hold hmpc's L x 32-bit limbs; twiddles come from the
        // element-major tables ([N][L] words).
        template<int L>
        struct arith_cios
        {
            static constexpr int rl = L;
            static constexpr int twl = rl;
            static constexpr int iol = L;
            static constexpr int wsl = L;
            static constexpr int tws = L;
            static constexpr bool stage_prefetch = true;
            static constexpr bool wave_units = false;
            static constexpr bool wave_blocks = false;
            // BGV sampling with any plaintext modulus (given as a Montgomery constant):
            // from_signed, mul_const and add_signed below
            static constexpr bool mont_plaintext = true;
            using params = fast_params;
            using ops = fast_ops<L, form_lazy>;

            __device__ __forceinline__ static void load_global(u32 (&x)[rl], u32 const* __restrict__ src, std::size_t count,
                                                               u32 g)
            {
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = ld_off(src + l * count, g);
            }

            // canonical: [0, 4q) -> [0, q); otherwise the lazy value itself (the next pass accepts
            // [0, 4q) forward and [0, 2q) inverse, which is what each pass leaves)
            __device__ __forceinline__ static void store_global(u32* __restrict__ dst, std::size_t count, u32 g,
                                                                u32 const (&xin)[rl], bool canonical, params const& P)
            {
                u32 x[rl];
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = xin[l];
                if (canonical)
                    ops::canonical(x, P);
#pragma unroll
                for (int l = 0; l < L; ++l)
                    st_off(dst + l * count, g, x[l]);
            }

            __device__ __forceinline__ static void load_ws(u32 (&x)[rl], u32 const* __restrict__ src, std::size_t count,
                                                           u32 g)
            {
                load_global(x, src, count, g);
            }

            __device__ __forceinline__ static void store_ws(u32* __restrict__ dst, std::size_t count, u32 g,
                                                            u32 const (&x)[rl], bool, params const& P)
            {
                store_global(dst, count, g, x, false, P);
            }

            __device__ __forceinline__ static int phys(int e) { return swizzle_16(e); }

            __device__ __forceinline__ static void lds_get_at(u32 (&x)[rl], u32 const* lds, int slot)
            {
                if constexpr (L == 4)
                {
                    uint4 v = reinterpret_cast<uint4 const*>(lds)[slot];
                    x[0] = v.x;
                    x[1] = v.y;
                    x[2] = v.z;
                    x[3] = v.w;
                }
                else
                {
#pragma unroll
                    for (int l = 0; l < L; ++l)
                        x[l] = lds[slot * L + l];
                }
            }

            __device__ __forceinline__ static void lds_put_at(u32* lds, int slot, u32 const (&x)[rl])
            {
                if constexpr (L == 4)
                    reinterpret_cast<uint4*>(lds)[slot] = make_uint4(x[0], x[1], x[2], x[3]);
                else
                {
#pragma unroll
                    for (int l = 0; l < L; ++l)
                        lds[slot * L + l] = x[l];
                }
            }

            __device__ __forceinline__ static void lds_put(u32* lds, int e, u32 const (&x)[rl]) { lds_put_at(lds, phys(e), x); }

            __device__ __forceinline__ static void tw_get(u32 (&w)[rl], u32 const* __restrict__ tw, u32 idx)
            {
                if constexpr (L == 4)
                {
                    uint4 v = reinterpret_cast<uint4 const*>(tw)[idx];
                    w[0] = v.x;
                    w[1] = v.y;
                    w[2] = v.z;
                    w[3] = v.w;
                }
                else
                {
#pragma unroll
                    for (int l = 0; l < L; ++l)
                        w[l] = tw[static_cast<std::size_t>(idx) * L + l];
                }
            }

            __device__ __forceinline__ static void fwd(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[rl], params const& P)
            {
                ops::fwd(x, y, w, P);
            }

            __device__ __forceinline__ static void inv(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w)[rl], params const& P,
                                                       bool)
            {
                ops::inv(x, y, w, P);
            }

            __device__ __forceinline__ static void inv_last(u32 (&x)[rl], u32 (&y)[rl], u32 const (&w0)[rl],
                                                            u32 const (&w1)[rl], params const& P)
            {
                ops::inv_last(x, y, w0, w1, P);
            }

            // a b R from forward results a R, b R in [0, 4q): both halved to [0, 2q), so that
            // a b < 4 q^2 and the lazy product lies in [0, 2q), the inverse transform's input range
            __device__ __forceinline__ static void point(u32 (&a)[rl], u32 (&b)[rl], params const& P)
            {
                ops::half_reduce(a, P);
                ops::half_reduce(b, P);
                u32 t[rl];
                ops::mul(t, a, b, P);
#pragma unroll
                for (int l = 0; l < L; ++l)
                    a[l] = t[l];
            }

            // x = k u + x for k canonical and u, x forward results in [0, 4q): u and x halved to
            // [0, 2q), k u < 2 q^2 gives a lazy product in [0, 2q), the sum lies in [0, 4q) (the
            // canonical store's input range)
            __device__ __forceinline__ static void point_add(u32 (&x)[rl], u32 const (&u)[rl], u32 const (&k)[rl],
                                                             params const& P)
            {
                u32 uh[rl], t[rl], xh[rl];
#pragma unroll
                for (int l = 0; l < L; ++l)
                {
                    uh[l] = u[l];
                    xh[l] = x[l];
                }
                ops::half_reduce(uh, P);
                ops::half_reduce(xh, P);
                ops::mul(t, k, uh, P);
                add_plain<L>(x, xh, t);
            }

            // signed integer |z| < 2^63 -> z R mod q in [0, 2q]: |z| < q, so the lazy Montgomery
            // product with R^2 is |z| R in [0, 2q); a negative z takes 2q minus it
            __device__ __forceinline__ static void from_signed(u32 (&x)[rl], long long z, params const& P)
            {
                u64 az = z < 0 ? static_cast<u64>(0) - static_cast<u64>(z) : static_cast<u64>(z);
                u32 a[rl] = {}, r2[rl], t[rl];
                a[0] = static_cast<u32>(az);
                a[1] = static_cast<u32>(az >> 32);
#pragma unroll
                for (int l = 0; l < L; ++l)
                    r2[l] = P.r2[l];
                ops::mul(t, a, r2, P);
                u32 n[rl];
                u32 br = 0;
#pragma unroll
                for (int l = 0; l < L; ++l)
                    n[l] = subb(P.q2[l], t[l], br, br);
                bool const neg = z < 0;
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = neg ? n[l] : t[l];
            }

            // x (in [0, 2q]) -> x c R^-1 for a canonical constant c: in [0, 2q)
            __device__ __forceinline__ static void mul_const(u32 (&x)[rl], u32 const* c, params const& P)
            {
                u32 cc[rl], t[rl];
#pragma unroll
                for (int l = 0; l < L; ++l)
                    cc[l] = c[l];
                ops::mul(t, x, cc, P);
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = t[l];
            }

            // x (in [0, 2q)) + m R for an unsigned 32-bit m: in [0, 4q)
            __device__ __forceinline__ static void add_signed(u32 (&x)[rl], u32 m, params const& P)
            {
                u32 y[rl], s[rl];
                from_signed(y, static_cast<long long>(m), P);
                add_plain<L>(s, x, y);
#pragma unroll
                for (int l = 0; l < L; ++l)
                    x[l] = s[l];
            }
        };

        template<int E>
        struct shape
        {
            static constexpr int threads = tile / E;
            static constexpr int wmax = ilog2(E);
        };

        // unit g of this thread in a round of width W, in the policy's unit order (unit_index_at)
        template<typename A, int E, int W>
        __device__ __forceinline__ int unit_index(int g)
        {
            return unit_index_at<E, W, A::wave_units>(tile_thread<A, E>(), g);
        }

        // Wave-local tiles (every exchange of a wave stays within its own tile elements) replace
        // the workgroup barrier by a wave barrier with wavefront-scope fences: a wave's LDS and
        // vector memory operations complete in order, so lanes of one wave see each other's
        // stores after it without waiting for the other waves.
        __device__ __forceinline__ void wave_sync()
        {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }

        template<bool WL>
        __device__ __forceinline__ void tile_sync()
        {
            if constexpr (WL)
                wave_sync();
            else
                __syncthreads();
        }

        // The elements of a unit are elem_of(unit, k = 0) | unit_offset(k) with disjoint bits, so
        // their LDS slots are the unit's slot XOR a slot offset that depends on k and the round only
        // (uniform over the workgroup, a compile-time constant for compile-time row lengths): the
        // swizzle is evaluated once per unit instead of once per element. (Pinning the offset in a
        // scalar register with readfirstlane cut the column kernels' VALU count further but ran
        // their forward pass 3% slower.)
        template<typename A>
        __device__ __forceinline__ int slot_offset(int d)
        {
            return A::phys(d);
        }

        template<typename A, int E, int W>
        __device__ __forceinline__ void gather(u32 (&x)[E][A::rl], u32 const* lds, tile_desc const& td, int a, bool inv)
        {
            constexpr int K = 1 << W;
#pragma unroll
            for (int g = 0; g < E / K; ++g)
            {
                unit_t un = unit_of(td, a, W, inv, unit_index<A, E, W>(g));
                int const slot = A::phys(elem_of(td, un.c, j_of(un, 0, a, W, td.s, inv)));
#pragma unroll
                for (int k = 0; k < K; ++k)
                    A::lds_get_at(x[g * K + k], lds, slot ^ slot_offset<A>(unit_offset(td, a, W, inv, k)));
            }
        }

        template<typename A, int E, int W>
        __device__ __forceinline__ void scatter(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, int a, bool inv)
        {
            constexpr int K = 1 << W;
#pragma unroll
            for (int g = 0; g < E / K; ++g)
            {
                unit_t un = unit_of(td, a, W, inv, unit_index<A, E, W>(g));
                int const slot = A::phys(elem_of(td, un.c, j_of(un, 0, a, W, td.s, inv)));
#pragma unroll
                for (int k = 0; k < K; ++k)
                    A::lds_put_at(lds, slot ^ slot_offset<A>(unit_offset(td, a, W, inv, k)), x[g * K + k]);
            }
        }

        // twiddles of one round: per unit g, stage II occupies entries
        //   forward: [2^II - 1, 2^(II+1) - 1), entry (2^II - 1) + (k >> (W - II))
        //   inverse: [2^W - 2^(W-II), 2^W - 2^(W-II-1)), entry 2^W - 2^(W-II) + (k >> (II + 1))
        template<int W>
        constexpr int tw_slot(bool inv, int ii, int k)
        {
            return inv ? ((1 << W) - (1 << (W - ii)) + (k >> (ii + 1))) : ((1 << ii) - 1 + (k >> (W - ii)));
        }

        // per unit g of the thread: the round's twiddle base (tw_round_base), from which every
        // twiddle index of the round is one shift and a constant
        template<typename A, int E, bool Inv, int W>
        __device__ __forceinline__ void round_bases(u32 (&xr)[E >> W], tile_desc const& td, int a, u32 base0,
                                                    u32 base_step)
        {
#pragma unroll
            for (int g = 0; g < (E >> W); ++g)
            {
                unit_t un = unit_of(td, a, W, Inv, unit_index<A, E, W>(g));
                xr[g] = tw_round_base<Inv, W>(base0 + un.c * base_step, td.s, a, un.F);
            }
        }

        template<typename A, int E, bool Inv, int W, int II>
        __device__ __forceinline__ void prefetch_stage(u32 (&tw_regs)[E >> W][(1 << W) - 1][A::twl], u32 const (&xr)[E >> W],
                                                       u32 const* __restrict__ tw)
        {
            constexpr int count = Inv ? (1 << (W - II - 1)) : (1 << II);
            constexpr int first = Inv ? ((1 << W) - (1 << (W - II))) : ((1 << II) - 1);
#pragma unroll
            for (int g = 0; g < (E >> W); ++g)
#pragma unroll
                for (int m = 0; m < count; ++m)
                    A::tw_get(tw_regs[g][first + m], tw, tw_index_round<Inv, W, II>(xr[g], m));
        }

        template<typename A, int E, bool Inv, int W, int II>
        __device__ __forceinline__ void prefetch_upto(u32 (&tw_regs)[E >> W][(1 << W) - 1][A::twl], u32 const (&xr)[E >> W],
                                                      u32 const* __restrict__ tw)
        {
            if constexpr (II < W - 1)
            {
                prefetch_stage<A, E, Inv, W, II>(tw_regs, xr, tw);
                prefetch_upto<A, E, Inv, W, II + 1>(tw_regs, xr, tw);
            }
        }

        // g0: index of the round's first stage within the whole inverse transform (it selects the
        // stages that fold: r26_inverse_fold_stage)
        template<typename A, int E, bool Inv, int W, int II, bool Last>
        __device__ __forceinline__ void stage(u32 (&x)[E][A::rl], u32 const (&tw_regs)[E >> W][(1 << W) - 1][A::twl],
                                              u32 const (&w0)[A::twl], u32 const (&w1)[A::twl], typename A::params const& P,
                                              int g0)
        {
            constexpr int K = 1 << W;
            constexpr int lb = Inv ? II : (W - 1 - II);
            bool const fold = r26_inverse_fold_stage(g0 + II);
#pragma unroll
            for (int g = 0; g < E / K; ++g)
            {
#pragma unroll
                for (int k = 0; k < K; ++k)
                {
                    if (k & (1 << lb))
                        continue;
                    int const slot = tw_slot<W>(Inv, II, k);
                    if constexpr (!Inv)
                        A::fwd(x[g * K + k], x[g * K + (k | (1 << lb))], tw_regs[g][slot], P);
                    else if constexpr (Last)
                        A::inv_last(x[g * K + k], x[g * K + (k | (1 << lb))], w0, w1, P);
                    else
                        A::inv(x[g * K + k], x[g * K + (k | (1 << lb))], tw_regs[g][slot], P, fold);
                }
            }
        }

        template<typename A, int E, bool Inv, int W, int II>
        __device__ __forceinline__ void stages_upto(u32 (&x)[E][A::rl], u32 const (&tw_regs)[E >> W][(1 << W) - 1][A::twl],
                                                    u32 const (&w0)[A::twl], u32 const (&w1)[A::twl],
                                                    typename A::params const& P, int g0)
        {
            if constexpr (II < W - 1)
            {
                stage<A, E, Inv, W, II, false>(x, tw_regs, w0, w1, P, g0);
                stages_upto<A, E, Inv, W, II + 1>(x, tw_regs, w0, w1, P, g0);
            }
        }

        // stages II .. W-2 of a round, each preceded by the loads of its own twiddles
        template<typename A, int E, bool Inv, int W, int II>
        __device__ __forceinline__ void stages_prefetched(u32 (&x)[E][A::rl], u32 (&tw_regs)[E >> W][(1 << W) - 1][A::twl],
                                                          u32 const (&xr)[E >> W], u32 const* __restrict__ tw,
                                                          typename A::params const& P, int g0)
        {
            if constexpr (II < W - 1)
            {
                u32 const z[A::twl] = {};
                prefetch_stage<A, E, Inv, W, II>(tw_regs, xr, tw);
                stage<A, E, Inv, W, II, false>(x, tw_regs, z, z, P, g0);
                stages_prefetched<A, E, Inv, W, II + 1>(x, tw_regs, xr, tw, P, g0);
            }
        }

        // butterflies of one round: all twiddle loads first, then the stages. The round's last
        // stage is the inverse's scaled final stage when the round ends the transform. off: the
        // inverse stages before this tile's sub-transform (a column pass after a row pass whose
        // results were not folded); stage a + II of the tile is stage off + a + II of the inverse.
        template<typename A, int E, bool Inv, int W>
        __device__ __forceinline__ void compute_round(u32 (&x)[E][A::rl], tile_desc const& td, int a, u32 base0,
                                                      u32 base_step, u32 const* __restrict__ tw,
                                                      typename A::params const& P, bool final_scale, int off)
        {
            u32 tw_regs[E >> W][(1 << W) - 1][A::twl];
            bool const scaled = Inv && final_scale && (a + W == td.s);
            u32 w0[A::twl] = {}, w1[A::twl] = {};
            int const g0 = off + a;
            u32 xr[E >> W];
            round_bases<A, E, Inv, W>(xr, td, a, base0, base_step);
            if constexpr (A::stage_prefetch)
            {
                // twiddles loaded stage by stage: at most one stage's twiddles live (fewer VGPRs,
                // the latency is covered by the other workgroup on the CU)
                stages_prefetched<A, E, Inv, W, 0>(x, tw_regs, xr, tw, P, g0);
                if (scaled)
                {
                    A::tw_get(w0, tw, 0);
                    A::tw_get(w1, tw, 1);
                    stage<A, E, Inv, W, W - 1, true>(x, tw_regs, w0, w1, P, g0);
                }
                else
                {
                    prefetch_stage<A, E, Inv, W, W - 1>(tw_regs, xr, tw);
                    stage<A, E, Inv, W, W - 1, false>(x, tw_regs, w0, w1, P, g0);
                }
                return;
            }
            prefetch_upto<A, E, Inv, W, 0>(tw_regs, xr, tw);
            if (scaled)
            {
                A::tw_get(w0, tw, 0);
                A::tw_get(w1, tw, 1);
            }
            else
                prefetch_stage<A, E, Inv, W, W - 1>(tw_regs, xr, tw);
            stages_upto<A, E, Inv, W, 0>(x, tw_regs, w0, w1, P, g0);
            if (scaled)
                stage<A, E, Inv, W, W - 1, true>(x, tw_regs, w0, w1, P, g0);
            else
                stage<A, E, Inv, W, W - 1, false>(x, tw_regs, w0, w1, P, g0);
        }

        template<typename A, int E, bool Inv, int W>
        __device__ __forceinline__ void one_round(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, int a, u32 base0,
                                                  u32 base_step, u32 const* __restrict__ tw, typename A::params const& P,
                                                  bool final_scale, bool keep, int off)
        {
            gather<A, E, W>(x, lds, td, a, Inv);
            compute_round<A, E, Inv, W>(x, td, a, base0, base_step, tw, P, final_scale, off);
            if (!keep)
                scatter<A, E, W>(x, lds, td, a, Inv);
        }

        template<typename A, int E, bool Inv>
        __device__ __forceinline__ void dispatch_round(int w, u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, int a,
                                                       u32 base0, u32 base_step, u32 const* __restrict__ tw,
                                                       typename A::params const& P, bool final_scale, bool keep, int off)
        {
            constexpr int WM = shape<E>::wmax;
            if (w == WM)
                one_round<A, E, Inv, WM>(x, lds, td, a, base0, base_step, tw, P, final_scale, keep, off);
            else if (WM > 1 && w == 1)
                one_round<A, E, Inv, 1>(x, lds, td, a, base0, base_step, tw, P, final_scale, keep, off);
            else if constexpr (WM > 2)
            {
                if (w == 2)
                    one_round<A, E, Inv, 2>(x, lds, td, a, base0, base_step, tw, P, final_scale, keep, off);
                else if constexpr (WM > 3)
                {
                    if (w == 3)
                        one_round<A, E, Inv, 3>(x, lds, td, a, base0, base_step, tw, P, final_scale, keep, off);
                }
            }
        }

        // All rounds of the tile's sub-transform, LDS -> LDS. With keep_last (forward only) the
        // last round's results stay in x instead of being written back. (Width dispatch by
        // explicit branches: a lambda over x would force the array into scratch memory.)
        // WL: a wave-local tile (rows of at most 64 E elements), synchronized per wave.
        template<typename A, int E, bool Inv, bool WL = false>
        __device__ __forceinline__ void tile_rounds(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, u32 base0,
                                                    u32 base_step, u32 const* __restrict__ tw, typename A::params const& P,
                                                    bool final_scale, int first_round, bool keep_last, int off = 0,
                                                    int end_round = -1)
        {
            constexpr int WM = shape<E>::wmax;
            int n = end_round < 0 ? round_count<WM>(td.s) : end_round;
            for (int r = first_round; r < n; ++r)
            {
                int a = round_start<WM>(td.s, r, Inv);
                int w = round_width<WM>(td.s, r, Inv);
                bool keep = keep_last && r == n - 1;
                dispatch_round<A, E, Inv>(w, x, lds, td, a, base0, base_step, tw, P, final_scale, keep, off);
                if (!keep)
                    tile_sync<WL>();
            }
        }

        // inverse round 0 computed on registers that already hold its elements, then written to LDS
        template<typename A, int E, int W>
        __device__ __forceinline__ void regs_round(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, u32 base0,
                                                   u32 base_step, u32 const* __restrict__ tw, typename A::params const& P,
                                                   bool final_scale)
        {
            compute_round<A, E, true, W>(x, td, 0, base0, base_step, tw, P, final_scale, 0);
            scatter<A, E, W>(x, lds, td, 0, true);
        }

        template<typename A, int E>
        __device__ __forceinline__ void dispatch_round_regs(int w, u32 (&x)[E][A::rl], u32* lds, tile_desc const& td,
                                                            u32 base0, u32 base_step, u32 const* __restrict__ tw,
                                                            typename A::params const& P, bool final_scale)
        {
            constexpr int WM = shape<E>::wmax;
            if (w == WM)
                regs_round<A, E, WM>(x, lds, td, base0, base_step, tw, P, final_scale);
            else if (WM > 1 && w == 1)
                regs_round<A, E, 1>(x, lds, td, base0, base_step, tw, P, final_scale);
            else if constexpr (WM > 2)
            {
                if (w == 2)
                    regs_round<A, E, 2>(x, lds, td, base0, base_step, tw, P, final_scale);
                else if constexpr (WM > 3)
                {
                    if (w == 3)
                        regs_round<A, E, 3>(x, lds, td, base0, base_step, tw, P, final_scale);
                }
            }
        }

        // Ws: src is a private workspace (A::wsl limb planes, raw format) rather than data in the
        // hmpc layout. WL: each wave moves its own elements wave 64 E + 64 i + lane (E = tile / T),
        // the ones its rounds work on in a wave-local tile.
        template<typename A, int T, bool Ws = false, bool WL = false>
        __device__ __forceinline__ void tile_load(u32* lds, u32 const* __restrict__ src, io_desc const& io)
        {
            constexpr int E = tile / T;
            constexpr int step = WL ? 64 : T;
            int const t = tile_thread<A, E>();
            int const e0 = WL ? (t & ~63) * E + (t & 63) : t;
            // e = e0 + d with d a multiple of the step above e0's bits: slot phys(e0) ^ phys(d)
            int const slot0 = A::phys(e0);
            u32 const* const src_t = src + io.base;
            // gidx_rel is additive over elements with disjoint bits, so element e0 + i step is at the
            // thread's offset off0 from a base that moves by a uniform amount per i
            u32 const off0 = gidx_rel(io, e0);
            if (io.valid >= tile)
            {
                // a whole tile (every tile but the last of a batch of small polynomials): no checks
#pragma unroll
                for (int i = 0; i < E; ++i)
                {
                    u32 x[A::rl];
                    if constexpr (Ws)
                        A::load_ws(x, src_t + gidx_rel(io, i * step), io.count, off0);
                    else
                        A::load_global(x, src_t + gidx_rel(io, i * step), io.count, off0);
                    A::lds_put_at(lds, slot0 ^ slot_offset<A>(i * step), x);
                }
                return;
            }
#pragma unroll
            for (int i = 0; i < E; ++i)
            {
                u32 x[A::rl];
                if (e0 + i * step < io.valid)
                {
                    if constexpr (Ws)
                        A::load_ws(x, src_t + gidx_rel(io, i * step), io.count, off0);
                    else
                        A::load_global(x, src_t + gidx_rel(io, i * step), io.count, off0);
                }
                else
                {
#pragma unroll
                    for (int l = 0; l < A::rl; ++l)
                        x[l] = 0;
                }
                A::lds_put_at(lds, slot0 ^ slot_offset<A>(i * step), x);
            }
        }

        // Ws: dst is a private workspace (raw format; canonical is then the fold flag of store_ws)
        template<typename A, int T, bool Ws = false, bool WL = false>
        __device__ __forceinline__ void tile_store(u32 const* lds, u32* __restrict__ dst, io_desc const& io, bool canonical,
                                                   typename A::params const& P)
        {
            constexpr int E = tile / T;
            constexpr int step = WL ? 64 : T;
            int const t = tile_thread<A, E>();
            int const e0 = WL ? (t & ~63) * E + (t & 63) : t;
            int const slot0 = A::phys(e0);
            u32* const dst_t = dst + io.base;
            u32 const off0 = gidx_rel(io, e0);
            if (io.valid >= tile)
            {
#pragma unroll
                for (int i = 0; i < E; ++i)
                {
                    u32 x[A::rl];
                    A::lds_get_at(x, lds, slot0 ^ slot_offset<A>(i * step));
                    if constexpr (Ws)
                        A::store_ws(dst_t + gidx_rel(io, i * step), io.count, off0, x, canonical, P);
                    else
                        A::store_global(dst_t + gidx_rel(io, i * step), io.count, off0, x, canonical, P);
                }
                return;
            }
#pragma unroll
            for (int i = 0; i < E; ++i)
            {
                if (e0 + i * step >= io.valid)
                    continue;
                u32 x[A::rl];
                A::lds_get_at(x, lds, slot0 ^ slot_offset<A>(i * step));
                if constexpr (Ws)
                    A::store_ws(dst_t + gidx_rel(io, i * step), io.count, off0, x, canonical, P);
                else
                    A::store_global(dst_t + gidx_rel(io, i * step), io.count, off0, x, canonical, P);
            }
        }

        template<typename A>
        constexpr std::size_t tile_bytes() { return std::size_t{4} * A::rl * tile; }

        // ------------------------------------------------------------------ second operands
        // After the forward transform's last round (stages [s - W, s)) of a row-layout tile, the
        // thread's unit g holds the K = 2^W consecutive tile elements (unit_index(g) << W) + k, so a
        // second operand of the same elements is read straight from its private workspace (K
        // consecutive words per limb plane, raw format; the workgroup wrote it itself just
        // before, so it is served from L2) instead of occupying a second register tile.
        //   Mode 0: x = a o x            (a = NTT(a) of a product, lazy)
        //   Mode 1: x = k o u + x        (u = NTT(u), k = public-key polynomial, canonical)
        // Only the tile's first `valid` elements are backed by memory (a whole-tile job of degree
        // N < 4096 fills one of the tile's 4096 / N sequences); the others are left untouched.
        template<typename A, int E, int W, int Mode>
        __device__ __forceinline__ void with_operands(u32 (&x)[E][A::rl], int valid, u32 const* __restrict__ src,
                                                      std::size_t src_count, std::size_t src_base, u32 const* __restrict__ key,
                                                      std::size_t key_count, std::size_t key_base, typename A::params const& P)
        {
            constexpr int K = 1 << W;
#pragma unroll
            for (int g = 0; g < E / K; ++g)
            {
                int const e0 = unit_index<A, E, W>(g) << W;
#pragma unroll
                for (int k = 0; k < K; ++k)
                {
                    if (e0 + k >= valid)
                        continue;
                    u32 v[A::rl];
                    A::load_ws(v, src + src_base, src_count, static_cast<u32>(e0 + k));
                    if constexpr (Mode == 0)
                    {
                        A::point(v, x[g * K + k], P);
#pragma unroll
                        for (int l = 0; l < A::rl; ++l)
                            x[g * K + k][l] = v[l];
                    }
                    else
                    {
                        u32 kv[A::rl];
                        A::load_global(kv, key + key_base, key_count, static_cast<u32>(e0 + k));
                        A::point_add(x[g * K + k], v, kv, P);
                    }
                }
            }
        }

        template<typename A, int E, int Mode>
        __device__ __forceinline__ void with_operands_last(u32 (&x)[E][A::rl], tile_desc const& td, int valid, u32 const* src,
                                                           std::size_t src_count, std::size_t src_base, u32 const* key,
                                                           std::size_t key_count, std::size_t key_base,
                                                           typename A::params const& P)
        {
            constexpr int WM = shape<E>::wmax;
            int w = round_width<WM>(td.s, round_count<WM>(td.s) - 1, false);
            if (w == WM)
                with_operands<A, E, WM, Mode>(x, valid, src, src_count, src_base, key, key_count, key_base, P);
            else if (WM > 1 && w == 1)
                with_operands<A, E, 1, Mode>(x, valid, src, src_count, src_base, key, key_count, key_base, P);
            else if constexpr (WM > 2)
            {
                if (w == 2)
                    with_operands<A, E, 2, Mode>(x, valid, src, src_count, src_base, key, key_count, key_base, P);
                else if constexpr (WM > 3)
                {
                    if (w == 3)
                        with_operands<A, E, 3, Mode>(x, valid, src, src_count, src_base, key, key_count, key_base, P);
                }
            }
        }

        // x += key o u for the register operand u (the same unit order as x) and the key read per
        // element (BGV rows with NTT(u) kept in registers)
        template<typename A, int E, int W>
        __device__ __forceinline__ void with_key_regs(u32 (&x)[E][A::rl], u32 const (&u)[E][A::rl], int valid,
                                                      u32 const* __restrict__ key, std::size_t key_count, std::size_t key_base,
                                                      typename A::params const& P)
        {
            constexpr int K = 1 << W;
#pragma unroll
            for (int g = 0; g < E / K; ++g)
            {
                int const e0 = unit_index<A, E, W>(g) << W;
#pragma unroll
                for (int k = 0; k < K; ++k)
                {
                    if (e0 + k >= valid)
                        continue;
                    u32 kv[A::rl];
                    A::load_global(kv, key + key_base, key_count, static_cast<u32>(e0 + k));
                    A::point_add(x[g * K + k], u[g * K + k], kv, P);
                }
            }
        }

        template<typename A, int E>
        __device__ __forceinline__ void with_key_regs_last(u32 (&x)[E][A::rl], u32 const (&u)[E][A::rl], tile_desc const& td,
                                                           int valid, u32 const* key, std::size_t key_count, std::size_t key_base,
                                                           typename A::params const& P)
        {
            constexpr int WM = shape<E>::wmax;
            int w = round_width<WM>(td.s, round_count<WM>(td.s) - 1, false);
            if (w == WM)
                with_key_regs<A, E, WM>(x, u, valid, key, key_count, key_base, P);
            else if (WM > 1 && w == 1)
                with_key_regs<A, E, 1>(x, u, valid, key, key_count, key_base, P);
            else if constexpr (WM > 2)
            {
                if (w == 2)
                    with_key_regs<A, E, 2>(x, u, valid, key, key_count, key_base, P);
                else if constexpr (WM > 3)
                {
                    if (w == 3)
                        with_key_regs<A, E, 3>(x, u, valid, key, key_count, key_base, P);
                }
            }
        }

        // ------------------------------------------------------------------ kernels
        // mode: 0 forward column pass, 1 forward row pass (final), 2 inverse row pass,
        //       3 inverse column pass (final), 4 forward whole-poly, 5 inverse whole-poly
        // WsIn: `in` is a private workspace (the row results of the two-pass product)
        template<typename A, int E, int Mode, bool WsIn = false, int S = 0>
        __global__ void __launch_bounds__(tile / E, E == 8 ? 4 : 1) k_pass(u32 const* __restrict__ in, u32* __restrict__ out,
                                                           u32 const* __restrict__ tw, pass_geom G, u64 n,
                                                           typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            constexpr bool Inv = Mode == 2 || Mode == 3 || Mode == 5;
            constexpr int T = tile / E;
            std::size_t count = static_cast<std::size_t>(n) << G.log_n;
            tile_desc td;
            io_desc io;
            u32 base0 = 1, base_step = 0;
            bool final_scale = false, canonical = false;
            if constexpr (Mode == 0 || Mode == 3)
            {
                cols_setup(G, blockIdx.x, count, td, io);
                if constexpr (S > 0)
                {
                    td.s = S;
                    td.log_seqs = tile_log - S;
                    io.log_seqs = tile_log - S;
                }
                final_scale = Mode == 3;
                canonical = Mode == 3;
            }
            else if constexpr (Mode == 1 || Mode == 2)
            {
                u32 hi0 = rows_setup(G, blockIdx.x, count, td, io);
                base0 = (1u << G.log_n1) + hi0;
                base_step = 1;
                canonical = Mode == 1;
            }
            else
            {
                single_setup(G, blockIdx.x, n, td, io);
                final_scale = Mode == 5;
                canonical = true;
            }
            tile_load<A, T, WsIn>(lds, in, io);
            __syncthreads();
            u32 x[E][A::rl];
            // a workspace input holds the row pass's unfolded results: the column stages continue
            // the inverse's stage numbering after the row pass's log_n2 stages
            tile_rounds<A, E, Inv>(x, lds, td, base0, base_step, tw, P, final_scale, 0, false, WsIn ? G.log_n2 : 0);
            tile_store<A, T>(lds, out, io, canonical, P);
        }

        // ------------------------------------------------------------------ rounds with wave-uniform twiddles
        // (the streaming column passes and the one-wave row kernel's first forward and last
        // inverse rounds)
        // a twiddle whose index is the same on every lane of the wave
        template<typename A>
        __device__ __forceinline__ void tw_get_wave(u32 (&w)[A::twl], u32 const* __restrict__ tw, u32 idx)
        {
            A::tw_get(w, tw, __builtin_amdgcn_readfirstlane(idx));
        }

        // a twiddle whose index takes at most two values on the wave, lane 0's and lane 63's
        // (two scalar loads and a per-lane select)
        template<typename A>
        __device__ __forceinline__ void tw_get_wave2(u32 (&w)[A::twl], u32 const* __restrict__ tw, u32 idx)
        {
            u32 const i0 = __builtin_amdgcn_readfirstlane(idx);
            u32 const i1 = __builtin_amdgcn_readlane(idx, 63);
            u32 w0[A::twl], w1[A::twl];
            A::tw_get(w0, tw, i0);
            A::tw_get(w1, tw, i1);
            bool const first = idx == i0;
#pragma unroll
            for (int l = 0; l < A::twl; ++l)
                w[l] = first ? w0[l] : w1[l];
        }

        // The expanded twiddles of the columns' rounds are the same for every tile, and hoisted out
        // of the tile loop they would hold 25 SGPRs each for the whole loop (150 in the forward
        // pass) and spill to VGPR lanes, one v_readlane per use. Adding an opaque zero (an empty
        // asm) to the table pointer at the top of each iteration keeps each stage's scalar loads in
        // the loop (scalar-cache hits), so its twiddles live only through that stage. The pointer
        // stays an offset of the kernel's restrict argument: an opaque pointer would lose the
        // no-clobber proof that lets uniform loads be scalar loads.
        __device__ __forceinline__ u32 const* opaque_table(u32 const* tw)
        {
            u32 z = 0;
            asm volatile("" : "+s"(z));
            return tw + z;
        }

        // stage II of a full-width round with expanded twiddles (A::twl > A::rl), each loaded just
        // before the first butterfly that uses it. Loaded a stage at a time they held up to four
        // expanded twiddles (100 SGPRs) and spilled SGPRs to VGPR lanes; here an empty asm makes the
        // next load's index depend on the butterfly just done (through an SGPR that stays zero), so
        // one expanded twiddle is live at a time. Butterfly k uses entry m = slot(k) - first of the
        // stage, which does not decrease with k.
        template<typename A, int E, bool Inv, int W, int II>
        __device__ __forceinline__ void stage_each(u32 (&x)[E][A::rl], u32 X, u32 const* __restrict__ tw,
                                                   typename A::params const& P, int g0)
        {
            static_assert((1 << W) == E, "full-width round");
            constexpr int K = 1 << W;
            constexpr int lb = Inv ? II : (W - 1 - II);
            constexpr int first = Inv ? ((1 << W) - (1 << (W - II))) : ((1 << II) - 1);
            bool const fold = r26_inverse_fold_stage(g0 + II);
            u32 dep = 0;
            u32 w[A::twl];
            int loaded = -1;
#pragma unroll
            for (int k = 0; k < K; ++k)
            {
                if (k & (1 << lb))
                    continue;
                int const m = tw_slot<W>(Inv, II, k) - first;
                if (m != loaded)
                {
                    tw_get_wave<A>(w, tw, tw_index_round<Inv, W, II>(X, m) + dep);
                    loaded = m;
                }
                u32 (&lo)[A::rl] = x[k];
                u32 (&hi)[A::rl] = x[k | (1 << lb)];
                if constexpr (!Inv)
                    A::fwd(lo, hi, w, P);
                else
                    A::inv(lo, hi, w, P, fold);
                asm volatile("" : "+s"(dep) : "v"(hi[4]));
            }
        }

        // stages II.. of a full-width round (one unit per thread) with per-wave twiddles; with
        // `scaled` the inverse's last stage is the scaled final stage (twiddles 0 and 1)
        template<typename A, int E, bool Inv, int W, int II>
        __device__ __forceinline__ void stages_wave(u32 (&x)[E][A::rl], u32 (&twr)[1][(1 << W) - 1][A::twl], u32 X,
                                                    u32 const* __restrict__ tw, typename A::params const& P, int g0,
                                                    bool scaled)
        {
            if constexpr (II < W)
            {
                constexpr int count = Inv ? (1 << (W - II - 1)) : (1 << II);
                constexpr int first = Inv ? ((1 << W) - (1 << (W - II))) : ((1 << II) - 1);
                if (Inv && II == W - 1 && scaled)
                {
                    u32 w0[A::twl], w1[A::twl];
                    tw_get_wave<A>(w0, tw, 0u);
                    tw_get_wave<A>(w1, tw, 1u);
                    stage<A, E, Inv, W, II, true>(x, twr, w0, w1, P, g0);
                }
                else if constexpr (A::twl > A::rl)
                    stage_each<A, E, Inv, W, II>(x, X, tw, P, g0);
                else
                {
                    u32 const z[A::twl] = {};
#pragma unroll
                    for (int m = 0; m < count; ++m)
                        tw_get_wave<A>(twr[0][first + m], tw, tw_index_round<Inv, W, II>(X, m));
                    stage<A, E, Inv, W, II, false>(x, twr, z, z, P, g0);
                }
                stages_wave<A, E, Inv, W, II + 1>(x, twr, X, tw, P, g0, scaled);
            }
        }

        template<typename A, int E, bool Inv, int W>
        __device__ __forceinline__ void round_wave(u32 (&x)[E][A::rl], tile_desc const& td, int a, u32 const* __restrict__ tw,
                                                   typename A::params const& P, int g0, bool scaled, u32 base0 = 1u,
                                                   u32 base_step = 0u)
        {
            static_assert((1 << W) == E, "full-width round");
            u32 twr[1][(1 << W) - 1][A::twl];
            u32 xr[1];
            round_bases<A, E, Inv, W>(xr, td, a, base0, base_step);
            stages_wave<A, E, Inv, W, 0>(x, twr, xr[0], tw, P, g0, scaled);
        }

        // polynomial product, whole polynomials per tile (N <= 4096) or the row pass of the
        // two-pass product: forward(A), forward(B), A o B, inverse. The row pass works in place
        // (a == out) and asks for 4 waves per SIMD (2 workgroups per CU with the 80 KiB tile).
        // S > 0: the row length log2 (G.log_n2) as a compile-time constant, so that the round
        // structure and every LDS address are resolved at compile time (and only the round widths
        // this length uses are emitted)
        // One-wave rows of 2^9 (A::wave_blocks, S = 9) hold one row per wave, so the twiddles of the
        // forward transforms' first round and of the inverse's last round (units of stride 2^6,
        // whose index depends on the row only) are wave-uniform: those rounds take the expanded
        // tables (twfx, twix) through round_wave.
        // off: the inverse stages before this tile's sub-transform (as in compute_round); scaled: the
        // round ends the inverse with its scaled final stage
        template<typename A, int E, bool Inv>
        __device__ __forceinline__ void round_expanded(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, int a,
                                                       u32 base0, u32 base_step, u32 const* __restrict__ tw,
                                                       typename A::params const& P, int off = 0, bool scaled = false)
        {
            gather<A, E, 3>(x, lds, td, a, Inv);
            round_wave<arith_tw<A>, E, Inv, 3>(x, td, a, opaque_table(tw), P, off + a, scaled, base0, base_step);
            scatter<A, E, 3>(x, lds, td, a, Inv);
        }

        // one-wave row tiles of 2^9 (one row per wave) take their first forward and last inverse
        // round on expanded twiddles
        template<typename A, int E, int S>
        constexpr bool rows_expanded()
        {
            return std::is_same<typename A::params, p26>::value && A::wave_blocks && S == 9 && E == 8;
        }

        // the forward rounds of a row tile (twx: the expanded forward table, read when
        // rows_expanded)
        template<typename A, int E, int S, bool WL>
        __device__ __forceinline__ void row_rounds_forward(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, u32 base0,
                                                           u32 base_step, u32 const* __restrict__ tw, u32 const* __restrict__ twx,
                                                           typename A::params const& P, bool keep_last)
        {
            constexpr bool XW = rows_expanded<A, E, S>();
            if constexpr (XW)
            {
                round_expanded<A, E, false>(x, lds, td, 0, base0, base_step, twx, P);
                tile_sync<WL>();
            }
            tile_rounds<A, E, false, WL>(x, lds, td, base0, base_step, tw, P, false, XW ? 1 : 0, keep_last);
        }

        // the inverse rounds after the first (which ran on registers, dispatch_round_regs); the last
        // (stages 6-8) on the expanded inverse table twx when rows_expanded
        template<typename A, int E, int S, bool WL>
        __device__ __forceinline__ void row_rounds_inverse_rest(u32 (&x)[E][A::rl], u32* lds, tile_desc const& td, u32 base0,
                                                                u32 base_step, u32 const* __restrict__ tw,
                                                                u32 const* __restrict__ twx, typename A::params const& P)
        {
            if constexpr (rows_expanded<A, E, S>())
            {
                tile_rounds<A, E, true, WL>(x, lds, td, base0, base_step, tw, P, false, 1, false, 0, 2);
                round_expanded<A, E, true>(x, lds, td, 6, base0, base_step, twx, P);
                tile_sync<WL>();
            }
            else
                tile_rounds<A, E, true, WL>(x, lds, td, base0, base_step, tw, P, false, 1, false);
        }

        template<typename A, int E, bool Single, int S = 0>
        __global__ void __launch_bounds__(A::wave_blocks ? 64 : tile / E, A::wave_blocks ? 3 : ((E == 8 && !Single) ? 4 : 1)) k_polymul(u32 const* a, u32 const* __restrict__ b,
                                                              u32* out, u32 const* __restrict__ twf,
                                                              u32 const* __restrict__ twi, u32 const* __restrict__ twfx,
                                                              u32 const* __restrict__ twix, pass_geom G, u64 n,
                                                              typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            constexpr int T = tile / E;
            constexpr int WM = shape<E>::wmax;
            std::size_t count = static_cast<std::size_t>(n) << G.log_n;
            tile_desc td;
            io_desc io;
            u32 base0 = 1, base_step = 0;
            if constexpr (Single)
            {
                single_setup(G, blockIdx.x, n, td, io);
            }
            else
            {
                // one-wave workgroups: block b serves wave b % (tile / 64 E) of tile b / (tile / 64 E)
                static_assert(!A::wave_blocks || (E == 8 && S > 0 && (1 << S) <= 64 * E),
                              "one-wave workgroups need wave-local rows at 8 elements per thread");
                u32 const tb = A::wave_blocks ? blockIdx.x / static_cast<u32>(tile / E / 64) : blockIdx.x;
                u32 hi0 = rows_setup(G, tb, count, td, io);
                if constexpr (S > 0)
                {
                    td.s = S;
                    td.log_seqs = tile_log - S;
                }
                base0 = (1u << G.log_n1) + hi0;
                base_step = 1;
            }
            int w0 = round_width<WM>(td.s, 0, true);
            if constexpr (Single)
            {
                u32 xa[E][A::rl];
                u32 xb[E][A::rl];
                tile_load<A, T>(lds, a, io);
                __syncthreads();
                tile_rounds<A, E, false>(xa, lds, td, base0, base_step, twf, P, false, 0, true);
                __syncthreads();
                tile_load<A, T>(lds, b, io);
                __syncthreads();
                tile_rounds<A, E, false>(xb, lds, td, base0, base_step, twf, P, false, 0, true);
#pragma unroll
                for (int k = 0; k < E; ++k)
                    A::point(xa[k], xb[k], P);
                // inverse round 0 groups the same elements as the forward's last round
                dispatch_round_regs<A, E>(w0, xa, lds, td, base0, base_step, twi, P, true);
                __syncthreads();
                tile_rounds<A, E, true>(xa, lds, td, base0, base_step, twi, P, true, 1, false);
                tile_store<A, T>(lds, out, io, true, P);
            }
            else
            {
                // two-pass row tile: NTT(a) rows go back to the workspace (a == out, in place) and
                // are read per element next to NTT(b) in registers, so one register tile is live.
                // With the wave-major unit order, rows of at most 64 E elements (S <= 9 at E = 8) are
                // wave-local: each wave loads, transforms, multiplies and stores whole rows of its
                // own, with no workgroup barrier.
                constexpr bool WL = A::wave_units && S > 0 && (1 << S) <= 64 * E;
                u32 x[E][A::rl];
                if constexpr (A::wave_blocks)
                {
                    // one-wave rows: NTT(a) stays in registers next to NTT(b)
                    u32 xa[E][A::rl];
                    tile_load<A, T, true, WL>(lds, a, io);
                    tile_sync<WL>();
                    row_rounds_forward<A, E, S, WL>(xa, lds, td, base0, base_step, twf, twfx, P, true);
                    tile_sync<WL>();
                    tile_load<A, T, true, WL>(lds, b, io);
                    tile_sync<WL>();
                    row_rounds_forward<A, E, S, WL>(x, lds, td, base0, base_step, twf, twfx, P, true);
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
                    tile_load<A, T, true, WL>(lds, a, io);
                    tile_sync<WL>();
                    tile_rounds<A, E, false, WL>(x, lds, td, base0, base_step, twf, P, false, 0, false);
                    tile_store<A, T, true, WL>(lds, out, io, false, P);
                    tile_sync<WL>();
                    tile_load<A, T, true, WL>(lds, b, io);
                    tile_sync<WL>();
                    tile_rounds<A, E, false, WL>(x, lds, td, base0, base_step, twf, P, false, 0, true);
                    with_operands_last<A, E, 0>(x, td, io.valid, out, io.count, io.base, nullptr, 0, 0, P);
                }
                dispatch_round_regs<A, E>(w0, x, lds, td, base0, base_step, twi, P, false);
                tile_sync<WL>();
                row_rounds_inverse_rest<A, E, S, WL>(x, lds, td, base0, base_step, twi, twix, P);
                tile_store<A, T, true, WL>(lds, out, io, false, P);
            }
        }

        // forward column pass of both inputs of the two-pass product (blockIdx.x = 2 * tile + which)
        template<typename A, int E, int S = 0>
        __global__ void __launch_bounds__(tile / E) k_polymul_cols(u32 const* __restrict__ a, u32 const* __restrict__ b,
                                                                   u32* __restrict__ wa, u32* __restrict__ wb,
                                                                   u32 const* __restrict__ tw, pass_geom G, u64 n,
                                                                   typename A::params P)
        {
            extern __shared__ __align__(16) u32 lds[];
            constexpr int T = tile / E;
            std::size_t count = static_cast<std::size_t>(n) << G.log_n;
            u32 which = blockIdx.x & 1u;
            tile_desc td;
            io_desc io;
            cols_setup(G, blockIdx.x >> 1, count, td, io);
            if constexpr (S > 0)
            {
                td.s = S;
                td.log_seqs = tile_log - S;
                io.log_seqs = tile_log - S;
            }
            tile_load<A, T>(lds, which ? b : a, io);
            __syncthreads();
            u32 x[E][A::rl];
            tile_rounds<A, E, false>(x, lds, td, 1u, 0u, tw, P, false, 0, false);
            tile_store<A, T, true>(lds, which ? wb : wa, io, ws_fold_columns(G.log_n), P);
        }

        // ------------------------------------------------------------------ streaming column passes
        // The column passes of the radix-2^26 product (columns of 2^7, 32 columns per tile, 8
        // elements per thread) as persistent workgroups that walk the tiles t = blockIdx.x,
        // blockIdx.x + gridDim.x, ... and load tile t + gridDim.x into registers while tile t is
        // transformed. With one tile per workgroup the two workgroups of a CU load, transform and
        // store in step, so memory and VALU time add up (0.52 + 0.53 ms for the forward pass of 512
        // pairs at N = 2^16, 0.85-0.89 ms together); streamed, they overlap (0.67 ms, same results).
        //
        // The prefetch only overlaps if no later vector load waits behind it (vector loads complete
        // in issue order), so these passes take no twiddle from a vector load inside the loop:
        //   - the 3-stage rounds have wave-uniform twiddles in this geometry (forward stages 0-5,
        //     inverse stages 1-6: the unit's top bits F are 0 or the wave index), taken in expanded
        //     form (arith_tw, r26_mont_tw) with scalar loads from lane 0's index;
        //   - forward stage 6 and inverse stage 0 run in registers in the pair layout
        //     (cols_pair_row: one v_permlane32_swap per word and pair from or to the 3-stage
        //     layout instead of an LDS exchange and a barrier); their twiddles (two per wave and
        //     pair, not depending on the tile) are loaded once per workgroup and kept in VGPRs
        //     (forward), or taken with two scalar loads and a per-lane select (inverse, whose 5-word
        //     prefetch leaves no room).
        // tests/cpp/jobs_index_check.cpp checks the per-wave uniformity, the pair layout and that each
        // pass moves every tile element once, inside its polynomial. Round-0 (forward) and pair-layout
        // (inverse) elements are loaded straight into registers and the last round's stored straight
        // from them (each wave access is two 128-byte column segments per limb plane, as in
        // tile_load), so LDS carries a single exchange per tile.

        // tile element of element k of this thread's unit g in round (a, W)
        template<typename A, int E, int W>
        __device__ __forceinline__ int unit_elem(tile_desc const& td, int a, bool inv, int g, int k)
        {
            unit_t un = unit_of(td, a, W, inv, unit_index<A, E, W>(g));
            return elem_of(td, un.c, j_of(un, static_cast<u32>(k), a, W, td.s, inv));
        }

        // pair layout <-> 3-stage-round layout of the streaming column passes (cols_pair_row): for
        // each pair m, the lower half-wave's element 2m + 1 and the upper half's element 2m change
        // places (v_permlane32_swap gives (a_lo | b_lo, a_hi | b_hi) for (a, b)); its own inverse
        template<typename A, int E>
        __device__ __forceinline__ void cols_pair_swap(u32 (&x)[E][A::rl])
        {
#pragma unroll
            for (int m = 0; m < E / 2; ++m)
#pragma unroll
                for (int l = 0; l < A::rl; ++l)
                {
                    auto r = __builtin_amdgcn_permlane32_swap(x[2 * m][l], x[2 * m + 1][l], false, false);
                    x[2 * m][l] = r[0];
                    x[2 * m + 1][l] = r[1];
                }
        }

        // forward column pass of both product inputs, tile t = 2 * column tile + operand (as
        // k_polymul_cols): rounds (stages 0-2) and (3-5) through LDS, stage 6 in registers after
        // the half-wave exchange into the pair layout
        template<typename A, int E, int S>
        __global__ void __launch_bounds__(tile / E, 4) k_polymul_cols_stream(u32 const* __restrict__ a, u32 const* __restrict__ b,
                                                                             u32* __restrict__ wa, u32* __restrict__ wb,
                                                                             u32 const* __restrict__ tw, u32 const* __restrict__ twx,
                                                                             pass_geom G, u64 n, typename A::params P, u32 tiles)
        {
            static_assert(E == 8 && S == 7, "streaming column geometry: 2^7-point columns, 8 elements per thread");
            static_assert(std::is_same<A, arith_r26>::value, "the rounds take expanded twiddles (arith_r26_tw)");
            using AT = arith_r26_tw;
            extern __shared__ __align__(16) u32 lds[];
            std::size_t const count = static_cast<std::size_t>(n) << G.log_n;
            tile_desc const td{S, tile_log - S, false};
            int const t0 = tid_x();
            // stage 6: the pairs' twiddles do not depend on the tile (two per wave and pair), loaded
            // once; the empty asm uses make the loads complete before the tile loop (inside it they
            // would wait behind the prefetch)
            u32 tw6[E / 2][1][A::rl];
#pragma unroll
            for (int m = 0; m < E / 2; ++m)
                A::tw_get(tw6[m][0], tw, cols_pair_twiddle(t0, m));
#pragma unroll
            for (int m = 0; m < E / 2; ++m)
#pragma unroll
                for (int l = 0; l < A::rl; ++l)
                    asm volatile("" ::"v"(tw6[m][0][l]));
            bool const fold = ws_fold_columns(G.log_n);
            io_desc io;
            tile_desc tdd;
            u32 t = blockIdx.x;
            if (t >= tiles)
                return;
            cols_setup(G, t >> 1, count, tdd, io);
            // round 0: one unit per thread, element k at e0 + unit_offset(k) (disjoint bits, so the
            // global offsets add: gidx_rel is additive over them)
            int const e_in = unit_elem<A, E, 3>(td, 0, false, 0, 0);
            u32 const o_in = gidx_rel(io, e_in);
            u32 raw[E][4];
            {
                u32 const* src = ((t & 1u) ? b : a) + io.base;
#pragma unroll
                for (int k = 0; k < E; ++k)
#pragma unroll
                    for (int l = 0; l < 4; ++l)
                        raw[k][l] = ld_off(src + l * count + gidx_rel(io, unit_offset(td, 0, 3, false, k)), o_in);
            }
            // stores in the pair layout: element (m, k) at e_out + the rows' offset
            int const e_out = cols_pair_elem(t0, 0, 0);
            u32 const o_out = gidx_rel(io, e_out);
            for (;;)
            {
                u32* const dst = ((t & 1u) ? wb : wa) + io.base;
                u32 x[E][A::rl];
#pragma unroll
                for (int k = 0; k < E; ++k)
                    r26_unpack(x[k], raw[k][0], raw[k][1], raw[k][2], raw[k][3]);
                u32 const tn = t + gridDim.x;
                io_desc ion;
                if (tn < tiles)
                {
                    cols_setup(G, tn >> 1, count, tdd, ion);
                    u32 const* src = ((tn & 1u) ? b : a) + ion.base;
#pragma unroll
                    for (int k = 0; k < E; ++k)
#pragma unroll
                        for (int l = 0; l < 4; ++l)
                            raw[k][l] = ld_off(src + l * count + gidx_rel(ion, unit_offset(td, 0, 3, false, k)), o_in);
                }
                u32 const* const twt = opaque_table(twx);
                round_wave<AT, E, false, 3>(x, td, 0, twt, P, 0, false);
                __syncthreads();  // the previous tile's gather is done
                scatter<A, E, 3>(x, lds, td, 0, false);
                __syncthreads();
                gather<A, E, 3>(x, lds, td, 3, false);
                round_wave<AT, E, false, 3>(x, td, 3, twt, P, 0, false);
                cols_pair_swap<A, E>(x);
                {
                    u32 const z[A::twl] = {};
                    stage<A, E, false, 1, 0, false>(x, tw6, z, z, P, 0);
                }
#pragma unroll
                for (int m = 0; m < E / 2; ++m)
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        A::store_ws(dst + gidx_rel(io, cols_pair_elem(t0, m, k) - e_out), count, o_out, x[2 * m + k], fold, P);
                if (tn >= tiles)
                    break;
                t = tn;
                io = ion;
            }
        }

        // inverse column pass of the product from the row pass's workspace (raw limbs), with the
        // N^-1 scaling and canonical output (as k_pass<A, E, 3, true, 7>): stage 0 in registers in
        // the pair layout, the half-wave exchange, then rounds (1-3) and (4-6); `off` = log N2, the
        // inverse stages the row pass ran before
        template<typename A, int E, int S>
        __global__ void __launch_bounds__(tile / E, 4) k_icols_stream(u32 const* __restrict__ in, u32* __restrict__ out,
                                                                      u32 const* __restrict__ tw, u32 const* __restrict__ twx,
                                                                      pass_geom G, u64 n, typename A::params P, u32 tiles)
        {
            static_assert(E == 8 && S == 7, "streaming column geometry: 2^7-point columns, 8 elements per thread");
            static_assert(std::is_same<A, arith_r26>::value, "the rounds take expanded twiddles (arith_r26_tw)");
            using AT = arith_r26_tw;
            extern __shared__ __align__(16) u32 lds[];
            std::size_t const count = static_cast<std::size_t>(n) << G.log_n;
            tile_desc const td{S, tile_log - S, false};
            int const off = G.log_n2;
            int const t0 = tid_x();
            io_desc io;
            tile_desc tdd;
            u32 t = blockIdx.x;
            if (t >= tiles)
                return;
            cols_setup(G, t, count, tdd, io);
            // loads in the pair layout
            int const e_in = cols_pair_elem(t0, 0, 0);
            u32 const o_in = gidx_rel(io, e_in);
            u32 raw[E][A::wsl];
#pragma unroll
            for (int m = 0; m < E / 2; ++m)
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    A::load_ws(raw[2 * m + k], in + io.base + gidx_rel(io, cols_pair_elem(t0, m, k) - e_in), count, o_in);
            // last round (stages 4-6): one unit per thread
            int const e_out = unit_elem<A, E, 3>(td, 4, true, 0, 0);
            u32 const o_out = gidx_rel(io, e_out);
            for (;;)
            {
                u32* const dst = out + io.base;
                u32 x[E][A::rl];
#pragma unroll
                for (int k = 0; k < E; ++k)
#pragma unroll
                    for (int l = 0; l < A::rl; ++l)
                        x[k][l] = raw[k][l];
                u32 const tn = t + gridDim.x;
                io_desc ion;
                if (tn < tiles)
                {
                    cols_setup(G, tn, count, tdd, ion);
#pragma unroll
                    for (int m = 0; m < E / 2; ++m)
#pragma unroll
                        for (int k = 0; k < 2; ++k)
                            A::load_ws(raw[2 * m + k], in + ion.base + gidx_rel(ion, cols_pair_elem(t0, m, k) - e_in), count, o_in);
                }
                {
                    // stage 0: two twiddles per wave and pair (lanes 0-31 and 32-63)
                    u32 const z[A::twl] = {};
                    u32 tw1[E / 2][1][A::rl];
#pragma unroll
                    for (int m = 0; m < E / 2; ++m)
                        tw_get_wave2<A>(tw1[m][0], tw, cols_pair_twiddle(t0, m));
                    stage<A, E, true, 1, 0, false>(x, tw1, z, z, P, off);
                }
                cols_pair_swap<A, E>(x);
                u32 const* const twt = opaque_table(twx);
                round_wave<AT, E, true, 3>(x, td, 1, twt, P, off + 1, false);
                __syncthreads();  // the previous tile's gather is done
                scatter<A, E, 3>(x, lds, td, 1, true);
                __syncthreads();
                gather<A, E, 3>(x, lds, td, 4, true);
                round_wave<AT, E, true, 3>(x, td, 4, twt, P, off + 4, true);
#pragma unroll
                for (int k = 0; k < E; ++k)
                {
                    int const d = unit_elem<A, E, 3>(td, 4, true, 0, k) - e_out;
                    A::store_global(dst + gidx_rel(io, d), count, o_out, x[k], true, P);
                }
                if (tn >= tiles)
                    break;
                t = tn;
                io = ion;
            }
        }

        // ------------------------------------------------------------------ host launchers
        // compute units of the current device (cached per device; 0 if the query fails)
        inline int device_cus()
        {
            static std::atomic<int> cache[64];
            int dev = 0;
            if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
                return 0;
            int v = cache[dev].load(std::memory_order_relaxed);
            if (v == 0 && hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
                cache[dev].store(v, std::memory_order_relaxed);
            return v > 0 ? v : 0;
        }

        // kernel-argument constants of the 32-bit-limb Montgomery policies (Goldilocks, CIOS)
        inline fast_params montgomery_params(transform_args const& T)
        {
            fast_params P{};
            for (int l = 0; l < max_limbs; ++l)
            {
                P.q[l] = T.P.q[l];
                P.q2[l] = T.q2[l];
                P.r2[l] = T.r2[l];
            }
            P.nprime = T.P.nprime;
            return P;
        }

        template<typename A, typename F>
        hipError_t with_elems(int elems, F&& f)
        {
            if constexpr (A::rl >= 4)
            {
                if (elems == 4)
                    return f(std::integral_constant<int, 4>{});
                return f(std::integral_constant<int, 8>{});
            }
            else
            {
                if (elems == 8)
                    return f(std::integral_constant<int, 8>{});
                return f(std::integral_constant<int, 16>{});
            }
        }

        // dynamic LDS above the 64 KiB default needs the kernel attribute raised (once per kernel
        // and process; the attribute call is not free)
        template<typename K>
        hipError_t allow_lds(K kernel, std::size_t bytes)
        {
            if (bytes <= 65536)
                return hipSuccess;
            static std::mutex m;
            static std::vector<std::pair<const void*, std::size_t>> done;
            const void* f = reinterpret_cast<const void*>(kernel);
            std::lock_guard<std::mutex> lock(m);
            for (auto const& d : done)
                if (d.first == f && d.second >= bytes)
                    return hipSuccess;
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
            if (e == hipSuccess)
                done.push_back({f, bytes});
            return e;
        }

        // whole-polynomial tiles (N <= 4096)
        template<typename A, int E>
        hipError_t launch_whole_tiles(transform_args const& T, typename A::params const& P, std::size_t n, bool inverse,
                                      u32 const* in, u32* out, hipStream_t s)
        {
            constexpr int TH = tile / E;
            constexpr std::size_t bytes = tile_bytes<A>();
            u32 const* tw = inverse ? T.inv_fast : T.fwd_fast;
            hipError_t e = hipSuccess;
            pass_geom G = geom_for(T.log_n, true);
            unsigned grid = static_cast<unsigned>((n + (std::size_t{1} << G.log_seqs) - 1) >> G.log_seqs);
            if (!inverse)
            {
                if ((e = allow_lds(k_pass<A, E, 4>, bytes)) != hipSuccess)
                    return e;
                probe_scope ps(s, "fast ntt forward (whole tile)");
                k_pass<A, E, 4><<<grid, TH, bytes, s>>>(in, out, tw, G, n, P);
            }
            else
            {
                if ((e = allow_lds(k_pass<A, E, 5>, bytes)) != hipSuccess)
                    return e;
                probe_scope ps(s, "fast ntt inverse (whole tile)");
                k_pass<A, E, 5><<<grid, TH, bytes, s>>>(in, out, tw, G, n, P);
            }
            return hipGetLastError();
        }

        template<typename A>
        hipError_t launch_ntt(transform_args const& T, typename A::params const& P, std::size_t n, bool inverse,
                              u32 const* in, u32* out, hipStream_t s)
        {
            // the latency schedule of small whole-tile Goldilocks calls: 4 elements per thread
            // (1024 threads per tile)
            if constexpr (A::rl == 2)
            {
                if (T.elems == 4 && T.log_n <= tile_log)
                    return launch_whole_tiles<A, 4>(T, P, n, inverse, in, out, s);
            }
            return with_elems<A>(T.elems, [&](auto EC) -> hipError_t {
                constexpr int E = decltype(EC)::value;
                constexpr int TH = tile / E;
                constexpr std::size_t bytes = tile_bytes<A>();
                u32 const* tw = inverse ? T.inv_fast : T.fwd_fast;
                hipError_t e = hipSuccess;
                if (T.log_n <= tile_log)
                    return launch_whole_tiles<A, E>(T, P, n, inverse, in, out, s);
                pass_geom gc = geom_for(T.log_n, false);
                pass_geom gr = geom_for(T.log_n, true);
                unsigned grid_c = static_cast<unsigned>(n * gc.blocks_per_poly);
                unsigned grid_r = static_cast<unsigned>(n * gr.blocks_per_poly);
                if (!inverse)
                {
                    if ((e = allow_lds(k_pass<A, E, 0>, bytes)) != hipSuccess || (e = allow_lds(k_pass<A, E, 1>, bytes)) != hipSuccess)
                        return e;
                    {
                        probe_scope ps(s, "fast ntt forward columns");
                        k_pass<A, E, 0><<<grid_c, TH, bytes, s>>>(in, out, tw, gc, n, P);
                    }
                    {
                        probe_scope ps(s, "fast ntt forward rows");
                        k_pass<A, E, 1><<<grid_r, TH, bytes, s>>>(out, out, tw, gr, n, P);
                    }
                }
                else
                {
                    if ((e = allow_lds(k_pass<A, E, 2>, bytes)) != hipSuccess || (e = allow_lds(k_pass<A, E, 3>, bytes)) != hipSuccess)
                        return e;
                    {
                        probe_scope ps(s, "fast ntt inverse rows");
                        k_pass<A, E, 2><<<grid_r, TH, bytes, s>>>(in, out, tw, gr, n, P);
                    }
                    {
                        probe_scope ps(s, "fast ntt inverse columns");
                        k_pass<A, E, 3><<<grid_c, TH, bytes, s>>>(out, out, tw, gc, n, P);
                    }
                }
                return hipGetLastError();
            });
        }

        template<typename A>
        hipError_t launch_polymul(transform_args const& T, typename A::params const& P, std::size_t n, u32 const* a,
                                  u32 const* b, u32* out, u32* workspace, hipStream_t s)
        {
            return with_elems<A>(T.elems, [&](auto EC) -> hipError_t {
                constexpr int E = decltype(EC)::value;
                constexpr int TH = tile / E;
                constexpr std::size_t bytes = tile_bytes<A>();
                hipError_t e = hipSuccess;
                if (T.log_n <= tile_log)
                {
                    pass_geom G = geom_for(T.log_n, true);
                    unsigned grid = static_cast<unsigned>((n + (std::size_t{1} << G.log_seqs) - 1) >> G.log_seqs);
                    if ((e = allow_lds(k_polymul<A, E, true>, bytes)) != hipSuccess)
                        return e;
                    {
                        probe_scope ps(s, "fast polymul (whole tile)");
                        k_polymul<A, E, true><<<grid, TH, bytes, s>>>(a, b, out, T.fwd_fast, T.inv_fast, T.fwd_tw, T.inv_tw,
                                                                      G, n, P);
                    }
                    return hipGetLastError();
                }
                std::size_t count = n << T.log_n;
                u32* wa = workspace;
                u32* wb = workspace + A::wsl * count;
                pass_geom gc = geom_for(T.log_n, false);
                pass_geom gr = geom_for(T.log_n, true);
                unsigned grid_c = static_cast<unsigned>(n * gc.blocks_per_poly);
                unsigned grid_r = static_cast<unsigned>(n * gr.blocks_per_poly);
                // the column length is 2^7 from N = 2^14 on (2^6 at N = 2^13)
                auto cols_kernel = gc.log_n1 == 7 ? k_polymul_cols<A, E, 7> : k_polymul_cols<A, E, 0>;
                auto icols_kernel = gc.log_n1 == 7 ? k_pass<A, E, 3, true, 7> : k_pass<A, E, 3, true, 0>;
                if ((e = allow_lds(cols_kernel, bytes)) != hipSuccess || (e = allow_lds(icols_kernel, bytes)) != hipSuccess)
                    return e;
                using AR = typename rows_arith<A>::type;
                auto rows_kernel = k_polymul<A, E, false, 0>;
                switch (gr.log_n2)
                {
                case 7: rows_kernel = k_polymul<AR, E, false, 7>; break;
                case 8: rows_kernel = k_polymul<AR, E, false, 8>; break;
                case 9: rows_kernel = k_polymul<AR, E, false, 9>; break;
                case 10: rows_kernel = k_polymul<AR, E, false, 10>; break;
                default: break;
                }
                // radix-2^26 rows of up to 64 E elements: one wave per workgroup (arith_r26_rows1)
                unsigned rows_grid = grid_r, rows_threads = TH;
                std::size_t rows_bytes = bytes;
                if constexpr (std::is_same<A, arith_r26>::value && E == 8)
                {
                    // rows of 2^9 read the expanded tables up to entry 8 * 2^log_n1 (their first and
                    // last rounds' indices)
                    bool const expanded = T.fwd_tw && T.inv_tw && (std::size_t{8} << gr.log_n1) <= r26_tw_entries;
                    if (gr.log_n2 >= 7 && gr.log_n2 <= 9 && (gr.log_n2 < 9 || expanded))
                    {
                        rows_kernel = gr.log_n2 == 7   ? k_polymul<arith_r26_rows1, E, false, 7>
                                      : gr.log_n2 == 8 ? k_polymul<arith_r26_rows1, E, false, 8>
                                                       : k_polymul<arith_r26_rows1, E, false, 9>;
                        rows_grid = grid_r * static_cast<unsigned>(tile / E / 64);
                        rows_threads = 64;
                        rows_bytes = bytes / static_cast<std::size_t>(tile / E / 64);
                    }
                }
                if ((e = allow_lds(rows_kernel, rows_bytes)) != hipSuccess)
                    return e;
                // the radix-2^26 column passes stream their tiles through two resident workgroups
                // per CU (k_polymul_cols_stream, k_icols_stream)
                bool stream_cols = false;
                unsigned grid_s = 0;
                if constexpr (std::is_same<A, arith_r26>::value && E == 8)
                {
                    int const cus = device_cus();
                    if (gc.log_n1 == 7 && cus > 0 && T.fwd_tw && T.inv_tw)
                    {
                        if ((e = allow_lds(k_polymul_cols_stream<A, E, 7>, bytes)) != hipSuccess ||
                            (e = allow_lds(k_icols_stream<A, E, 7>, bytes)) != hipSuccess)
                            return e;
                        stream_cols = true;
                        grid_s = 2u * static_cast<unsigned>(cus);
                    }
                }
                {
                    probe_scope ps(s, "fast polymul forward columns");
                    if constexpr (std::is_same<A, arith_r26>::value && E == 8)
                    {
                        if (stream_cols)
                            k_polymul_cols_stream<A, E, 7><<<grid_s < 2 * grid_c ? grid_s : 2 * grid_c, TH, bytes, s>>>(
                                a, b, wa, wb, T.fwd_fast, T.fwd_tw, gc, n, P, 2 * grid_c);
                        else
                            cols_kernel<<<2 * grid_c, TH, bytes, s>>>(a, b, wa, wb, T.fwd_fast, gc, n, P);
                    }
                    else
                        cols_kernel<<<2 * grid_c, TH, bytes, s>>>(a, b, wa, wb, T.fwd_fast, gc, n, P);
                }
                {
                    probe_scope ps(s, "fast polymul rows");
                    rows_kernel<<<rows_grid, rows_threads, rows_bytes, s>>>(wa, wb, wa, T.fwd_fast, T.inv_fast, T.fwd_tw, T.inv_tw,
                                                                            gr, n, P);
                }
                {
                    probe_scope ps(s, "fast polymul inverse columns");
                    if constexpr (std::is_same<A, arith_r26>::value && E == 8)
                    {
                        if (stream_cols)
                            k_icols_stream<A, E, 7><<<grid_s < grid_c ? grid_s : grid_c, TH, bytes, s>>>(wa, out, T.inv_fast, T.inv_tw,
                                                                                                      gc, n, P, grid_c);
                        else
                            icols_kernel<<<grid_c, TH, bytes, s>>>(wa, out, T.inv_fast, gc, n, P);
                    }
                    else
                        icols_kernel<<<grid_c, TH, bytes, s>>>(wa, out, T.inv_fast, gc, n, P);
                }
                return hipGetLastError();
            });
        }
    }
}
