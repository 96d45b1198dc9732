// HIP kernels for gfx950: twiddle tables, hmpc's per-stage NTT ("plain" path), LDS-staged
// fused transforms, point-wise ops, ChaCha20 samplers and batched BGV encryption.
//
// Transform semantics (hmpc/expr/number_theoretic_transform.hpp):
//   forward stage s (CT, :158-193):  m = 2^s, t = N/(2m), group g: psi = W[m + g],
//                                    (u, v) -> (u + v psi, u - v psi) on pairs (j, j + t)
//   inverse stage s (GS, :104-138):  h = N/2^(s+1), step = 2^s, psi = Winv[h + g],
//                                    (u, v) -> (u + v, (u - v) psi); last stage scales the sum by
//                                    Winv[0] = N^-1 and uses the pre-scaled Winv[1] (:536-546)
// Two-pass decomposition N = N1 * N2, j = hi * N2 + lo:
//   forward stages 0..log N1 - 1 act on columns (fixed lo) as an N1-point transform with
//   table indices 2^s + g; stages log N1.. act on rows (fixed hi) as an N2-point transform
//   with indices (N1 + hi) 2^s' + g'. The inverse runs rows first (indices (N1 + hi) h' + g')
//   and columns last (indices h + g, with the final scaling). The output order is therefore
//   exactly hmpc's bit-reversed order, with no extra permutation pass.
#include "kernels.hpp"

#include <type_traits>

namespace hmpc_amd
{
    launch_probe*& current_probe()
    {
        thread_local launch_probe* probe = nullptr;
        return probe;
    }
}

namespace hmpc_amd
{
    namespace
    {
        constexpr int block_threads = 256;
        constexpr std::size_t lds_budget = 65536;

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

        __device__ __forceinline__ u32 bit_reverse(u32 v, int bits)
        {
            return __builtin_bitreverse32(v) >> (32 - bits);
        }

        template<int L>
        __device__ __forceinline__ void load_w(u32 (&w)[L], u32 const* __restrict__ table, u32 N, u32 idx)
        {
#pragma unroll
            for (int l = 0; l < L; ++l)
                w[l] = table[static_cast<std::size_t>(l) * N + idx];
        }

        template<int L>
        __device__ __forceinline__ void lds_load(u32 (&x)[L], u32 const* tile, int plane, int e)
        {
#pragma unroll
            for (int l = 0; l < L; ++l)
                x[l] = tile[l * plane + e];
        }

        template<int L>
        __device__ __forceinline__ void lds_store(u32* tile, int plane, int e, u32 const (&x)[L])
        {
#pragma unroll
            for (int l = 0; l < L; ++l)
                tile[l * plane + e] = x[l];
        }
    }

    // ------------------------------------------------------------------ twiddle tables
    // roots[0] = 1 (forward) or N^-1 (inverse); roots[brv(i)] = root^i; inverse roots[1]
    // pre-multiplied by N^-1 (hmpc/expr/number_theoretic_transform.hpp:59-83).
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_make_twiddles(u32* table, int log_n, residue<L> root, residue<L> norm,
                                                                     residue<L> one, int inverse, mod_params P)
    {
        u32 N = 1u << log_n;
        u32 i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= N)
            return;
        u32 r[L];
        if (i == 0)
        {
            copy<L>(r, inverse ? norm.v : one.v);
        }
        else
        {
            u32 acc[L], base[L];
            copy<L>(acc, one.v);
            copy<L>(base, root.v);
            for (u32 e = i; e != 0; e >>= 1)
            {
                if (e & 1u)
                    mod_mul<L>(acc, acc, base, P);
                mod_mul<L>(base, base, base, P);
            }
            u32 j = bit_reverse(i, log_n);
            if (inverse && j == 1)
                mod_mul<L>(r, acc, norm.v, P);
            else
                copy<L>(r, acc);
            i = j;
        }
        store_planar<L>(table, N, i, r);
    }

    hipError_t launch_make_twiddles(int L, u32* table, int log_n, u32 const* root, u32 const* norm, u32 const* one,
                                    bool inverse, mod_params const& P, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            residue<LL> r, nm, o;
            for (int l = 0; l < LL; ++l)
            {
                r.v[l] = root[l];
                nm.v[l] = norm[l];
                o.v[l] = one[l];
            }
            k_make_twiddles<LL><<<grid_for(std::size_t{1} << log_n), block_threads, 0, s>>>(table, log_n, r, nm, o,
                                                                                            inverse ? 1 : 0, P);
            return hipGetLastError();
        });
    }

    template<int L>
    __global__ void __launch_bounds__(block_threads) k_to_element_major(u32 const* __restrict__ planar, u32* __restrict__ em,
                                                                        u64 count)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
#pragma unroll
        for (int l = 0; l < L; ++l)
            em[e * L + l] = planar[l * count + e];
    }

    hipError_t launch_to_element_major(int L, u32 const* planar, u32* em, std::size_t count, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_to_element_major<LL><<<grid_for(count), block_threads, 0, s>>>(planar, em, count);
            return hipGetLastError();
        });
    }

    __global__ void __launch_bounds__(block_threads) k_r26_table(u32 const* __restrict__ planar, u32* __restrict__ out, u64 count,
                                                                 p26 P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 w[4];
#pragma unroll
        for (int l = 0; l < 4; ++l)
            w[l] = planar[l * count + e];
        // 4 w R mod q (w R canonical, 4 q < 2^127): shift, then three conditional subtractions
        u32 t[4] = {w[0] << 2, (w[1] << 2) | (w[0] >> 30), (w[2] << 2) | (w[1] >> 30), (w[3] << 2) | (w[2] >> 30)};
        for (int rep = 0; rep < 3; ++rep)
        {
            u32 d[4];
            u32 br = 0;
#pragma unroll
            for (int l = 0; l < 4; ++l)
                d[l] = subb(t[l], P.q32[l], br, br);
#pragma unroll
            for (int l = 0; l < 4; ++l)
                t[l] = br ? t[l] : d[l];
        }
        u32 x[5];
        r26_unpack(x, t[0], t[1], t[2], t[3]);
#pragma unroll
        for (int l = 0; l < 5; ++l)
            out[e * 8 + l] = x[l];
#pragma unroll
        for (int l = 5; l < 8; ++l)
            out[e * 8 + l] = 0;
    }

    hipError_t launch_r26_table(u32 const* planar, u32* out, std::size_t count, p26 const& P, hipStream_t s)
    {
        k_r26_table<<<grid_for(count), block_threads, 0, s>>>(planar, out, count, P);
        return hipGetLastError();
    }

    // the first `count` entries of a radix-2^26 table (stride 8) in the expanded form of
    // r26_mont_tw (25 words per entry, stride r26_tw_stride, the rest zero)
    __global__ void __launch_bounds__(block_threads) k_r26_tw_table(u32 const* __restrict__ fast, u32* __restrict__ out, u32 count,
                                                                    p26 P)
    {
        u32 e = blockIdx.x * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 W[5], V[r26_tw_words];
#pragma unroll
        for (int l = 0; l < 5; ++l)
            W[l] = fast[8 * e + l];
        r26_expand_twiddle(V, W, P);
#pragma unroll
        for (int l = 0; l < r26_tw_stride; ++l)
            out[r26_tw_stride * e + l] = l < r26_tw_words ? V[l] : 0u;
    }

    hipError_t launch_r26_tw_table(u32 const* fast, u32* out, std::size_t count, p26 const& P, hipStream_t s)
    {
        k_r26_tw_table<<<grid_for(count), block_threads, 0, s>>>(fast, out, static_cast<u32>(count), P);
        return hipGetLastError();
    }

    // ------------------------------------------------------------------ plain path (hmpc's launch structure)
    // K2 / K5: first stage, [n][N] -> transposed scratch [N][n]
    template<int L, bool Inv>
    __global__ void __launch_bounds__(block_threads) k_plain_first(u32 const* __restrict__ in, u32* __restrict__ scratch,
                                                                   u32 const* __restrict__ W, int log_n, u64 n, mod_params P)
    {
        u64 half = u64{1} << (log_n - 1);
        u64 N = half << 1;
        u64 count = n * N;
        u64 id = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (id >= n * half)
            return;
        u64 i = id >> (log_n - 1);
        u64 tid = id & (half - 1);
        u32 u[L], v[L], w[L], r[L];
        if (!Inv)
        {
            load_planar<L>(u, in, count, tid + i * N);
            load_planar<L>(v, in, count, tid + half + i * N);
            load_w<L>(w, W, static_cast<u32>(N), 1);
            mod_mul<L>(v, v, w, P);
            mod_add<L>(r, u, v, P);
            store_planar<L>(scratch, count, i + tid * n, r);
            mod_sub<L>(r, u, v, P);
            store_planar<L>(scratch, count, i + (tid + half) * n, r);
        }
        else
        {
            u64 target = 2 * tid;
            load_planar<L>(u, in, count, target + i * N);
            load_planar<L>(v, in, count, target + 1 + i * N);
            load_w<L>(w, W, static_cast<u32>(N), static_cast<u32>(half + tid));
            mod_add<L>(r, u, v, P);
            store_planar<L>(scratch, count, i + target * n, r);
            mod_sub<L>(r, u, v, P);
            mod_mul<L>(r, r, w, P);
            store_planar<L>(scratch, count, i + (target + 1) * n, r);
        }
    }

    // K3 / K6: one middle stage over the scratch
    template<int L, bool Inv>
    __global__ void __launch_bounds__(block_threads) k_plain_mid(u32* __restrict__ scratch, u32 const* __restrict__ W,
                                                                 int log_n, int it, u64 n, mod_params P)
    {
        u64 half = u64{1} << (log_n - 1);
        u64 N = half << 1;
        u64 count = n * N;
        u64 id = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (id >= n * half)
            return;
        u64 tid = id / n;
        u64 i = id - tid * n;
        int log_step = Inv ? it : (log_n - 1 - it);
        u64 step = u64{1} << log_step;
        u64 length = Inv ? (N >> (it + 1)) : (u64{1} << it);
        u64 psi_step = tid >> log_step;
        u64 target = (psi_step << (log_step + 1)) + (tid & (step - 1));
        u32 u[L], v[L], w[L], r[L];
        load_w<L>(w, W, static_cast<u32>(N), static_cast<u32>(length + psi_step));
        load_planar<L>(u, scratch, count, i + target * n);
        load_planar<L>(v, scratch, count, i + (target + step) * n);
        if (!Inv)
        {
            mod_mul<L>(v, v, w, P);
            mod_add<L>(r, u, v, P);
            store_planar<L>(scratch, count, i + target * n, r);
            mod_sub<L>(r, u, v, P);
            store_planar<L>(scratch, count, i + (target + step) * n, r);
        }
        else
        {
            mod_add<L>(r, u, v, P);
            store_planar<L>(scratch, count, i + target * n, r);
            mod_sub<L>(r, u, v, P);
            mod_mul<L>(r, r, w, P);
            store_planar<L>(scratch, count, i + (target + step) * n, r);
        }
    }

    // K4 / K7: last stage, scratch [N][n] -> [n][N]
    template<int L, bool Inv>
    __global__ void __launch_bounds__(block_threads) k_plain_last(u32 const* __restrict__ scratch, u32* __restrict__ out,
                                                                  u32 const* __restrict__ W, int log_n, u64 n, mod_params P)
    {
        u64 half = u64{1} << (log_n - 1);
        u64 N = half << 1;
        u64 count = n * N;
        u64 id = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (id >= n * half)
            return;
        u64 tid = id / n;
        u64 i = id - tid * n;
        u32 u[L], v[L], w[L], r[L];
        if (!Inv)
        {
            u64 target = 2 * tid;
            load_w<L>(w, W, static_cast<u32>(N), static_cast<u32>(half + tid));
            load_planar<L>(u, scratch, count, i + target * n);
            load_planar<L>(v, scratch, count, i + (target + 1) * n);
            mod_mul<L>(v, v, w, P);
            mod_add<L>(r, u, v, P);
            store_planar<L>(out, count, target + i * N, r);
            mod_sub<L>(r, u, v, P);
            store_planar<L>(out, count, target + 1 + i * N, r);
        }
        else
        {
            u32 w0[L];
            load_w<L>(w0, W, static_cast<u32>(N), 0);
            load_w<L>(w, W, static_cast<u32>(N), 1);
            load_planar<L>(u, scratch, count, i + tid * n);
            load_planar<L>(v, scratch, count, i + (tid + half) * n);
            u32 t[L];
            mod_add<L>(t, u, v, P);
            mod_mul<L>(r, t, w0, P);
            store_planar<L>(out, count, tid + i * N, r);
            mod_sub<L>(t, u, v, P);
            mod_mul<L>(r, t, w, P);
            store_planar<L>(out, count, tid + half + i * N, r);
        }
    }

    hipError_t launch_ntt_plain(int L, transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                u32* scratch, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) -> hipError_t {
            constexpr int LL = decltype(LC)::value;
            std::size_t work = n << (T.log_n - 1);
            unsigned grid = grid_for(work);
            u32 const* W = inverse ? T.inv : T.fwd;
            if (!inverse)
                k_plain_first<LL, false><<<grid, block_threads, 0, s>>>(in, scratch, W, T.log_n, n, T.P);
            else
                k_plain_first<LL, true><<<grid, block_threads, 0, s>>>(in, scratch, W, T.log_n, n, T.P);
            for (int it = 1; it < T.log_n - 1; ++it)
            {
                if (!inverse)
                    k_plain_mid<LL, false><<<grid, block_threads, 0, s>>>(scratch, W, T.log_n, it, n, T.P);
                else
                    k_plain_mid<LL, true><<<grid, block_threads, 0, s>>>(scratch, W, T.log_n, it, n, T.P);
            }
            if (!inverse)
                k_plain_last<LL, false><<<grid, block_threads, 0, s>>>(scratch, out, W, T.log_n, n, T.P);
            else
                k_plain_last<LL, true><<<grid, block_threads, 0, s>>>(scratch, out, W, T.log_n, n, T.P);
            return hipGetLastError();
        });
    }

    // ------------------------------------------------------------------ LDS tiles
    // A tile holds `seqs` independent sequences of S = 2^log_s points, limb-planar in LDS
    // (plane size S * seqs words). Column layout: element (k, c) at k * seqs + c (consecutive
    // lanes touch consecutive banks); row layout: element (c, k) at c * S + k.
    struct tile_shape
    {
        int log_s;
        int log_seqs;
        bool rows;
        __device__ int plane() const { return 1 << (log_s + log_seqs); }
        __device__ int elem(int k, int c) const { return rows ? ((c << log_s) + k) : ((k << log_seqs) + c); }
    };

    // All log_s stages of the sub-transform on every sequence of the tile.
    // Twiddle index for sequence c: forward (base_c << s) + g, inverse base_c * h + g, with
    // base_c = base0 + c * base_step. `final_scale` applies the inverse's last-stage scaling.
    template<int L, bool Inv>
    __device__ void tile_transform(u32* tile, tile_shape sh, u32 const* __restrict__ W, u32 N, u32 base0, u32 base_step,
                                   bool final_scale, mod_params const& P)
    {
        int const plane = sh.plane();
        int const log_half = sh.log_s - 1;
        int const total = 1 << (log_half + sh.log_seqs);
        for (int s = 0; s < sh.log_s; ++s)
        {
            int log_step = Inv ? s : (sh.log_s - 1 - s);
            int step = 1 << log_step;
            bool last = Inv && final_scale && (s == sh.log_s - 1);
            for (int b = threadIdx.x; b < total; b += blockDim.x)
            {
                int c, t;
                if (sh.rows)
                {
                    c = b >> log_half;
                    t = b & ((1 << log_half) - 1);
                }
                else
                {
                    c = b & ((1 << sh.log_seqs) - 1);
                    t = b >> sh.log_seqs;
                }
                int g = t >> log_step;
                int j = (g << (log_step + 1)) + (t & (step - 1));
                u32 base = base0 + static_cast<u32>(c) * base_step;
                int e0 = sh.elem(j, c);
                int e1 = sh.elem(j + step, c);
                u32 u[L], v[L], w[L], r[L];
                lds_load<L>(u, tile, plane, e0);
                lds_load<L>(v, tile, plane, e1);
                if (!Inv)
                {
                    load_w<L>(w, W, N, (base << s) + static_cast<u32>(g));
                    mod_mul<L>(v, v, w, P);
                    mod_add<L>(r, u, v, P);
                    lds_store<L>(tile, plane, e0, r);
                    mod_sub<L>(r, u, v, P);
                    lds_store<L>(tile, plane, e1, r);
                }
                else if (!last)
                {
                    u32 h = 1u << (sh.log_s - 1 - s);
                    load_w<L>(w, W, N, base * h + static_cast<u32>(g));
                    mod_add<L>(r, u, v, P);
                    lds_store<L>(tile, plane, e0, r);
                    mod_sub<L>(r, u, v, P);
                    mod_mul<L>(r, r, w, P);
                    lds_store<L>(tile, plane, e1, r);
                }
                else
                {
                    u32 w0[L], t0[L];
                    load_w<L>(w0, W, N, 0);
                    load_w<L>(w, W, N, 1);
                    mod_add<L>(t0, u, v, P);
                    mod_mul<L>(r, t0, w0, P);
                    lds_store<L>(tile, plane, e0, r);
                    mod_sub<L>(t0, u, v, P);
                    mod_mul<L>(r, t0, w, P);
                    lds_store<L>(tile, plane, e1, r);
                }
            }
            __syncthreads();
        }
    }

    // global <-> LDS for a column tile: rows k < S, columns c0 .. c0 + C of poly p
    template<int L>
    __device__ void cols_load(u32* tile, tile_shape sh, u32 const* __restrict__ src, std::size_t count, std::size_t poly_base,
                              int log_n2, int c0)
    {
        int plane = sh.plane();
        int C = 1 << sh.log_seqs;
        for (int e = threadIdx.x; e < plane; e += blockDim.x)
        {
            int k = e >> sh.log_seqs;
            int c = e & (C - 1);
            std::size_t g = poly_base + (static_cast<std::size_t>(k) << log_n2) + c0 + c;
#pragma unroll
            for (int l = 0; l < L; ++l)
                tile[l * plane + e] = src[l * count + g];
        }
    }

    template<int L>
    __device__ void cols_store(u32 const* tile, tile_shape sh, u32* __restrict__ dst, std::size_t count, std::size_t poly_base,
                               int log_n2, int c0)
    {
        int plane = sh.plane();
        int C = 1 << sh.log_seqs;
        for (int e = threadIdx.x; e < plane; e += blockDim.x)
        {
            int k = e >> sh.log_seqs;
            int c = e & (C - 1);
            std::size_t g = poly_base + (static_cast<std::size_t>(k) << log_n2) + c0 + c;
#pragma unroll
            for (int l = 0; l < L; ++l)
                dst[l * count + g] = tile[l * plane + e];
        }
    }

    // rows are contiguous: element e of the tile is global element start + e
    template<int L>
    __device__ void rows_load(u32* tile, int plane, u32 const* __restrict__ src, std::size_t count, std::size_t start)
    {
        for (int e = threadIdx.x; e < plane; e += blockDim.x)
        {
#pragma unroll
            for (int l = 0; l < L; ++l)
                tile[l * plane + e] = src[l * count + start + e];
        }
    }

    template<int L>
    __device__ void rows_store(u32 const* tile, int plane, u32* __restrict__ dst, std::size_t count, std::size_t start)
    {
        for (int e = threadIdx.x; e < plane; e += blockDim.x)
        {
#pragma unroll
            for (int l = 0; l < L; ++l)
                dst[l * count + start + e] = tile[l * plane + e];
        }
    }

    // N = N1 * N2 with N1 = 2^floor(log N / 2); the widest column / row blocks whose tiles fit
    // the LDS budget when `tiles_cols` / `tiles_rows` tiles are resident at once (-1: none fits).
    split_plan make_split(int limbs, int log_n, int tiles_cols, int tiles_rows)
    {
        split_plan p{};
        p.single = false;
        p.log_n1 = log_n / 2;
        p.log_n2 = log_n - p.log_n1;
        std::size_t n1 = std::size_t{1} << p.log_n1, n2 = std::size_t{1} << p.log_n2;
        std::size_t ec = lds_budget / (static_cast<std::size_t>(tiles_cols) * 4 * limbs);
        std::size_t er = lds_budget / (static_cast<std::size_t>(tiles_rows) * 4 * limbs);
        int lc = 0;
        while ((n1 << (lc + 1)) <= ec && (std::size_t{1} << (lc + 1)) <= n2)
            ++lc;
        int lr = 0;
        while ((n2 << (lr + 1)) <= er && (std::size_t{1} << (lr + 1)) <= n1)
            ++lr;
        p.cols = n1 <= ec ? lc : -1;
        p.rows = n2 <= er ? lr : -1;
        return p;
    }

    split_plan plan_for(int limbs, int log_n, int tiles_in_lds)
    {
        std::size_t elems = lds_budget / (static_cast<std::size_t>(tiles_in_lds) * 4 * limbs);
        if ((std::size_t{1} << log_n) <= elems)
        {
            split_plan p{};
            p.single = true;
            p.log_n1 = log_n;
            return p;
        }
        return make_split(limbs, log_n, tiles_in_lds, tiles_in_lds);
    }

    // ------------------------------------------------------------------ fused transforms
    // whole polynomial in LDS (small N)
    template<int L, bool Inv>
    __global__ void __launch_bounds__(block_threads) k_ntt_single(u32 const* __restrict__ in, u32* __restrict__ out,
                                                                  u32 const* __restrict__ W, int log_n, u64 n, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.x;
        u32 N = 1u << log_n;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        tile_shape sh{log_n, 0, true};
        rows_load<L>(lds, N, in, count, p * N);
        __syncthreads();
        tile_transform<L, Inv>(lds, sh, W, N, 1u, 0u, true, P);
        rows_store<L>(lds, N, out, count, p * N);
    }

    // column pass: blockIdx.x = column block, blockIdx.y = polynomial
    template<int L, bool Inv>
    __global__ void __launch_bounds__(block_threads) k_ntt_cols(u32 const* __restrict__ in, u32* __restrict__ out,
                                                                u32 const* __restrict__ W, int log_n, int log_n1, int log_cols,
                                                                u64 n, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.y;
        int log_n2 = log_n - log_n1;
        int c0 = blockIdx.x << log_cols;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        tile_shape sh{log_n1, log_cols, false};
        cols_load<L>(lds, sh, in, count, p << log_n, log_n2, c0);
        __syncthreads();
        tile_transform<L, Inv>(lds, sh, W, 1u << log_n, 1u, 0u, true, P);
        cols_store<L>(lds, sh, out, count, p << log_n, log_n2, c0);
    }

    // row pass: blockIdx.x = row block, blockIdx.y = polynomial
    template<int L, bool Inv>
    __global__ void __launch_bounds__(block_threads) k_ntt_rows(u32 const* __restrict__ in, u32* __restrict__ out,
                                                                u32 const* __restrict__ W, int log_n, int log_n1, int log_rows,
                                                                u64 n, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.y;
        int log_n2 = log_n - log_n1;
        u32 hi0 = blockIdx.x << log_rows;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        tile_shape sh{log_n2, log_rows, true};
        int plane = sh.plane();
        std::size_t start = (p << log_n) + (static_cast<std::size_t>(hi0) << log_n2);
        rows_load<L>(lds, plane, in, count, start);
        __syncthreads();
        tile_transform<L, Inv>(lds, sh, W, 1u << log_n, (1u << log_n1) + hi0, 1u, false, P);
        rows_store<L>(lds, plane, out, count, start);
    }

    hipError_t launch_ntt_fused(int L, transform_args const& T, std::size_t n, bool inverse, u32 const* in, u32* out,
                                hipStream_t s)
    {
        split_plan plan = plan_for(L, T.log_n, 1);
        if (!plan.single && (plan.cols < 0 || plan.rows < 0))
            return hipErrorInvalidValue;
        return with_limbs(L, [&](auto LC) -> hipError_t {
            constexpr int LL = decltype(LC)::value;
            u32 const* W = inverse ? T.inv : T.fwd;
            if (plan.single)
            {
                std::size_t bytes = (std::size_t{4} * LL) << T.log_n;
                if (!inverse)
                    k_ntt_single<LL, false><<<static_cast<unsigned>(n), block_threads, bytes, s>>>(in, out, W, T.log_n, n, T.P);
                else
                    k_ntt_single<LL, true><<<static_cast<unsigned>(n), block_threads, bytes, s>>>(in, out, W, T.log_n, n, T.P);
                return hipGetLastError();
            }
            dim3 gc(1u << (plan.log_n2 - plan.cols), static_cast<unsigned>(n));
            dim3 gr(1u << (plan.log_n1 - plan.rows), static_cast<unsigned>(n));
            std::size_t bc = (std::size_t{4} * LL) << (plan.log_n1 + plan.cols);
            std::size_t br = (std::size_t{4} * LL) << (plan.log_n2 + plan.rows);
            if (!inverse)
            {
                k_ntt_cols<LL, false><<<gc, block_threads, bc, s>>>(in, out, W, T.log_n, plan.log_n1, plan.cols, n, T.P);
                k_ntt_rows<LL, false><<<gr, block_threads, br, s>>>(out, out, W, T.log_n, plan.log_n1, plan.rows, n, T.P);
            }
            else
            {
                k_ntt_rows<LL, true><<<gr, block_threads, br, s>>>(in, out, W, T.log_n, plan.log_n1, plan.rows, n, T.P);
                k_ntt_cols<LL, true><<<gc, block_threads, bc, s>>>(out, out, W, T.log_n, plan.log_n1, plan.cols, n, T.P);
            }
            return hipGetLastError();
        });
    }

    // ------------------------------------------------------------------ fused polynomial product
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_polymul_single(u32 const* __restrict__ a, u32 const* __restrict__ b,
                                                                      u32* __restrict__ out, u32 const* __restrict__ Wf,
                                                                      u32 const* __restrict__ Wi, int log_n, u64 n, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.x;
        u32 N = 1u << log_n;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        u32* ta = lds;
        u32* tb = lds + L * N;
        tile_shape sh{log_n, 0, true};
        rows_load<L>(ta, N, a, count, p * N);
        rows_load<L>(tb, N, b, count, p * N);
        __syncthreads();
        tile_transform<L, false>(ta, sh, Wf, N, 1u, 0u, false, P);
        tile_transform<L, false>(tb, sh, Wf, N, 1u, 0u, false, P);
        for (int e = threadIdx.x; e < static_cast<int>(N); e += blockDim.x)
        {
            u32 x[L], y[L];
            lds_load<L>(x, ta, N, e);
            lds_load<L>(y, tb, N, e);
            mod_mul<L>(x, x, y, P);
            lds_store<L>(ta, N, e, x);
        }
        __syncthreads();
        tile_transform<L, true>(ta, sh, Wi, N, 1u, 0u, true, P);
        rows_store<L>(ta, N, out, count, p * N);
    }

    // column pass over both inputs: blockIdx.z selects a or b
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_polymul_cols_fwd(u32 const* __restrict__ a, u32 const* __restrict__ b,
                                                                        u32* __restrict__ wa, u32* __restrict__ wb,
                                                                        u32 const* __restrict__ W, int log_n, int log_n1,
                                                                        int log_cols, u64 n, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.y;
        u32 const* src = blockIdx.z ? b : a;
        u32* dst = blockIdx.z ? wb : wa;
        int log_n2 = log_n - log_n1;
        int c0 = blockIdx.x << log_cols;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        tile_shape sh{log_n1, log_cols, false};
        cols_load<L>(lds, sh, src, count, p << log_n, log_n2, c0);
        __syncthreads();
        tile_transform<L, false>(lds, sh, W, 1u << log_n, 1u, 0u, false, P);
        cols_store<L>(lds, sh, dst, count, p << log_n, log_n2, c0);
    }

    // row pass: forward rows of a and b, product, inverse rows; result into wa
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_polymul_rows(u32* __restrict__ wa, u32 const* __restrict__ wb,
                                                                    u32 const* __restrict__ Wf, u32 const* __restrict__ Wi,
                                                                    int log_n, int log_n1, int log_rows, u64 n, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.y;
        int log_n2 = log_n - log_n1;
        u32 hi0 = blockIdx.x << log_rows;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        tile_shape sh{log_n2, log_rows, true};
        int plane = sh.plane();
        u32* ta = lds;
        u32* tb = lds + L * plane;
        std::size_t start = (p << log_n) + (static_cast<std::size_t>(hi0) << log_n2);
        rows_load<L>(ta, plane, wa, count, start);
        rows_load<L>(tb, plane, wb, count, start);
        __syncthreads();
        u32 base = (1u << log_n1) + hi0;
        tile_transform<L, false>(ta, sh, Wf, 1u << log_n, base, 1u, false, P);
        tile_transform<L, false>(tb, sh, Wf, 1u << log_n, base, 1u, false, P);
        for (int e = threadIdx.x; e < plane; e += blockDim.x)
        {
            u32 x[L], y[L];
            lds_load<L>(x, ta, plane, e);
            lds_load<L>(y, tb, plane, e);
            mod_mul<L>(x, x, y, P);
            lds_store<L>(ta, plane, e, x);
        }
        __syncthreads();
        tile_transform<L, true>(ta, sh, Wi, 1u << log_n, base, 1u, false, P);
        rows_store<L>(ta, plane, wa, count, start);
    }

    std::size_t polymul_workspace_words(int L, int log_n, std::size_t n)
    {
        split_plan plan = plan_for(L, log_n, 2);
        return plan.single ? 0 : 2 * static_cast<std::size_t>(L) * (n << log_n);
    }

    hipError_t launch_polymul(int L, transform_args const& T, std::size_t n, u32 const* a, u32 const* b, u32* out,
                              u32* workspace, hipStream_t s)
    {
        split_plan plan = plan_for(L, T.log_n, 2);
        if (!plan.single)
            plan = make_split(L, T.log_n, 1, 2);
        if (!plan.single && (plan.cols < 0 || plan.rows < 0))
            return hipErrorInvalidValue;
        return with_limbs(L, [&](auto LC) -> hipError_t {
            constexpr int LL = decltype(LC)::value;
            if (plan.single)
            {
                std::size_t bytes = (std::size_t{8} * LL) << T.log_n;
                k_polymul_single<LL><<<static_cast<unsigned>(n), block_threads, bytes, s>>>(a, b, out, T.fwd, T.inv, T.log_n,
                                                                                          n, T.P);
                return hipGetLastError();
            }
            // the column passes use one tile, the row pass two
            split_plan one = make_split(LL, T.log_n, 1, 2);
            std::size_t count = n << T.log_n;
            u32* wa = workspace;
            u32* wb = workspace + LL * count;
            dim3 gc(1u << (one.log_n2 - one.cols), static_cast<unsigned>(n), 2);
            dim3 gr(1u << (plan.log_n1 - plan.rows), static_cast<unsigned>(n));
            std::size_t bc = (std::size_t{4} * LL) << (one.log_n1 + one.cols);
            std::size_t br = (std::size_t{8} * LL) << (plan.log_n2 + plan.rows);
            k_polymul_cols_fwd<LL><<<gc, block_threads, bc, s>>>(a, b, wa, wb, T.fwd, T.log_n, one.log_n1, one.cols, n, T.P);
            k_polymul_rows<LL><<<gr, block_threads, br, s>>>(wa, wb, T.fwd, T.inv, T.log_n, plan.log_n1, plan.rows, n, T.P);
            dim3 gi(1u << (one.log_n2 - one.cols), static_cast<unsigned>(n));
            k_ntt_cols<LL, true><<<gi, block_threads, bc, s>>>(wa, out, T.inv, T.log_n, one.log_n1, one.cols, n, T.P);
            return hipGetLastError();
        });
    }

    // ------------------------------------------------------------------ point-wise
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_pointwise(int op, u64 count, u32 const* __restrict__ a,
                                                                 u32 const* __restrict__ b, u64 b_count, u32* __restrict__ out,
                                                                 mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u64 eb = b_count == count ? e : e % b_count;
        u32 x[L], y[L], r[L];
        load_planar<L>(x, a, count, e);
        load_planar<L>(y, b, b_count, eb);
        if (op == 0)
            mod_add<L>(r, x, y, P);
        else if (op == 1)
            mod_sub<L>(r, x, y, P);
        else
            mod_mul<L>(r, x, y, P);
        store_planar<L>(out, count, e, r);
    }

    hipError_t launch_pointwise(int L, mod_params const& P, int op, std::size_t count, u32 const* a, u32 const* b,
                                std::size_t b_count, u32* out, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_pointwise<LL><<<grid_for(count), block_threads, 0, s>>>(op, count, a, b, b_count, out, P);
            return hipGetLastError();
        });
    }

    template<int L>
    __global__ void __launch_bounds__(block_threads) k_pointwise_scalar(int op, u64 count, u32 const* __restrict__ a,
                                                                        residue<L> c, u32* __restrict__ out, mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 x[L], r[L];
        load_planar<L>(x, a, count, e);
        if (op == 0)
            mod_add<L>(r, x, c.v, P);
        else if (op == 1)
            mod_sub<L>(r, x, c.v, P);
        else
            mod_mul<L>(r, x, c.v, P);
        store_planar<L>(out, count, e, r);
    }

    hipError_t launch_pointwise_scalar(int L, mod_params const& P, int op, std::size_t count, u32 const* a, u32 const* c,
                                       u32* out, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            residue<LL> cv;
            for (int l = 0; l < LL; ++l)
                cv.v[l] = c[l];
            k_pointwise_scalar<LL><<<grid_for(count), block_threads, 0, s>>>(op, count, a, cv, out, P);
            return hipGetLastError();
        });
    }

    template<int L>
    __device__ __forceinline__ void lift_u32(u32 (&r)[L], u32 word, u32 const (&r2)[L], mod_params const& P)
    {
        u32 x[L];
#pragma unroll
        for (int l = 0; l < L; ++l)
            x[l] = l == 0 ? word : 0u;
        mod_mul<L>(r, x, r2, P);
    }

    template<int L>
    __global__ void __launch_bounds__(block_threads) k_from_u32(u64 count, u32 const* __restrict__ words, residue<L> r2,
                                                                u32* __restrict__ out, mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 r[L];
        lift_u32<L>(r, words[e], r2.v, P);
        store_planar<L>(out, count, e, r);
    }

    hipError_t launch_from_u32(int L, mod_params const& P, u32 const* r2, std::size_t count, u32 const* words, u32* out,
                               hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            residue<LL> rv;
            for (int l = 0; l < LL; ++l)
                rv.v[l] = r2[l];
            k_from_u32<LL><<<grid_for(count), block_threads, 0, s>>>(count, words, rv, out, P);
            return hipGetLastError();
        });
    }

    template<int L>
    __device__ __forceinline__ void to_unsigned(u32 (&r)[L], u32 const (&x)[L], mod_params const& P)
    {
        u32 one[L];
#pragma unroll
        for (int l = 0; l < L; ++l)
            one[l] = l == 0 ? 1u : 0u;
        mod_mul<L>(r, x, one, P);
    }

    template<int L>
    __global__ void __launch_bounds__(block_threads) k_to_unsigned(u64 count, u32 const* __restrict__ in, u32* __restrict__ out,
                                                                   mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 x[L], r[L];
        load_planar<L>(x, in, count, e);
        to_unsigned<L>(r, x, P);
        store_planar<L>(out, count, e, r);
    }

    hipError_t launch_to_unsigned(int L, mod_params const& P, std::size_t count, u32 const* in, u32* out, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_to_unsigned<LL><<<grid_for(count), block_threads, 0, s>>>(count, in, out, P);
            return hipGetLastError();
        });
    }

    // centred lift of REDC(x): values above (q - 1) / 2 map to value - q; keep the low word
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_centered_low32(u64 count, u32 const* __restrict__ in, u32* __restrict__ out,
                                                                      mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 x[L], v[L];
        load_planar<L>(x, in, count, e);
        to_unsigned<L>(v, x, P);
        // negative iff v > (q - 1) / 2  <=>  2 v > q - 1  <=>  2 v >= q
        u32 d[L + 1];
        u32 carry = 0, br = 0;
        u32 twice[L + 1];
#pragma unroll
        for (int l = 0; l < L; ++l)
            twice[l] = addc(v[l], v[l], carry, carry);
        twice[L] = carry;
#pragma unroll
        for (int l = 0; l < L; ++l)
            d[l] = subb(twice[l], P.q[l], br, br);
        d[L] = subb(twice[L], 0u, br, br);
        bool negative = br == 0;
        out[e] = negative ? v[0] - P.q[0] : v[0];
    }

    hipError_t launch_centered_low32(int L, mod_params const& P, std::size_t count, u32 const* in, u32* out, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            k_centered_low32<LL><<<grid_for(count), block_threads, 0, s>>>(count, in, out, P);
            return hipGetLastError();
        });
    }

    // ------------------------------------------------------------------ samplers
    __device__ __forceinline__ void explicit_nonce(u64 index, u64 count, u64 offset, u32& n0, u32& n1)
    {
        u64 v = index + count * offset;
        n0 = static_cast<u32>(v);
        n1 = static_cast<u32>(v >> 32);
    }

    // uniform<mod q> with drowning: W = ceil((bit_size + 80) / 32) words, top word masked,
    // REDC with max(L, W) = W rounds (hmpc/random/uniform.hpp:61-66, hmpc/ints/mod.hpp:210-215)
    template<int L, int W>
    __global__ void __launch_bounds__(block_threads) k_sample_uniform(chacha_key key, u64 offset, u64 count, int bits,
                                                                      u32* __restrict__ out, mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 n0, n1;
        explicit_nonce(e, count, offset, n0, n1);
        u32 blk[16];
        chacha20_block(blk, key, 0u, n0, n1, 0u);
        u32 v[W];
#pragma unroll
        for (int i = 0; i < W; ++i)
            v[i] = blk[i];
        if (bits % 32)
            v[W - 1] &= (1u << (bits % 32)) - 1u;
        u32 r[L];
        montgomery_reduce_wide<L, W>(r, v, P);
        store_planar<L>(out, count, e, r);
    }

    hipError_t launch_sample_uniform(int L, mod_params const& P, int bit_size, chacha_key const& key, u64 offset,
                                     std::size_t count, u32* out, hipStream_t s)
    {
        int bits = bit_size + 80;
        int W = (bits + 31) / 32;
        return with_limbs(L, [&](auto LC) -> hipError_t {
            constexpr int LL = decltype(LC)::value;
            // bit_size in (32(L-1), 32L]  =>  W in {L + 2, L + 3}
            if (W == LL + 2)
                k_sample_uniform<LL, LL + 2><<<grid_for(count), block_threads, 0, s>>>(key, offset, count, bits, out, P);
            else if (W == LL + 3)
                k_sample_uniform<LL, LL + 3><<<grid_for(count), block_threads, 0, s>>>(key, offset, count, bits, out, P);
            else
                return hipErrorInvalidValue;
            return hipGetLastError();
        });
    }

    template<int L>
    __global__ void __launch_bounds__(block_threads) k_sample_cbd(chacha_key key, u64 offset, u64 count, int bits, residue<L> r2,
                                                                  u32* __restrict__ out, mod_params P)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 n0, n1;
        explicit_nonce(e, count, offset, n0, n1);
        u32 blk[16];
        chacha20_block(blk, key, 0u, n0, n1, 0u);
        int val = centered_binomial_words(blk[0], blk[1], bits);
        u32 r[L];
        mod_from_small_signed<L>(r, val, r2.v, P);
        store_planar<L>(out, count, e, r);
    }

    hipError_t launch_sample_cbd(int L, mod_params const& P, u32 const* r2, chacha_key const& key, u64 offset, int bits,
                                 std::size_t count, u32* out, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            residue<LL> rv;
            for (int l = 0; l < LL; ++l)
                rv.v[l] = r2[l];
            k_sample_cbd<LL><<<grid_for(count), block_threads, 0, s>>>(key, offset, count, bits, rv, out, P);
            return hipGetLastError();
        });
    }

    __global__ void __launch_bounds__(block_threads) k_sample_words(chacha_key key, u64 offset, u64 count, u32* __restrict__ out)
    {
        u64 e = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        if (e >= count)
            return;
        u32 n0, n1;
        explicit_nonce(e, count, offset, n0, n1);
        u32 blk[16];
        chacha20_block(blk, key, 0u, n0, n1, 0u);
        out[e] = blk[0];
    }

    hipError_t launch_sample_words(chacha_key const& key, u64 offset, std::size_t count, u32* out, hipStream_t s)
    {
        k_sample_words<<<grid_for(count), block_threads, 0, s>>>(key, offset, count, out);
        return hipGetLastError();
    }

    // Capability stream of butterfly (poly, j), j < N/2: counter = low32(poly * N + j); the
    // first draw is coefficient j, the second j + N/2 (both from one block: bits <= 64).
    struct cap_stream
    {
        chacha_key key;
        u32 nonce[3];
        int bits;
    };

    template<int L>
    __device__ __forceinline__ void cap_pair(u32 (&lo)[L], u32 (&hi)[L], cap_stream const& cs, u64 poly, int log_n, u32 j,
                                             u32 const (&r2)[L], mod_params const& P)
    {
        u32 counter = static_cast<u32>((poly << log_n) + j);
        u32 blk[16];
        chacha20_block(blk, cs.key, counter, cs.nonce[0], cs.nonce[1], cs.nonce[2]);
        int a = capability_small_words(blk[0], blk[1], cs.bits);
        int b = capability_two_words(cs.bits) ? capability_small_words(blk[2], blk[3], cs.bits)
                                              : centered_binomial_words(blk[1], 0u, cs.bits);
        mod_from_small_signed<L>(lo, a, r2, P);
        mod_from_small_signed<L>(hi, b, r2, P);
    }

    // coefficient-domain capability sampling (plain path): out[poly][N] for polys first..first+n
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_sample_cbd_capability(cap_stream cs, int log_n, u64 n, u64 first,
                                                                             residue<L> r2, u32* __restrict__ out, u64 out_count,
                                                                             mod_params P)
    {
        u64 id = u64{blockIdx.x} * blockDim.x + threadIdx.x;
        u64 half = u64{1} << (log_n - 1);
        if (id >= n * half)
            return;
        u64 p = id >> (log_n - 1);
        u32 j = static_cast<u32>(id & (half - 1));
        u32 lo[L], hi[L];
        cap_pair<L>(lo, hi, cs, first + p, log_n, j, r2.v, P);
        store_planar<L>(out, out_count, (p << log_n) + j, lo);
        store_planar<L>(out, out_count, (p << log_n) + j + half, hi);
    }

    hipError_t launch_sample_cbd_capability(int L, mod_params const& P, u32 const* r2, chacha_key const& key,
                                            u32 const* nonce, int bits, int log_n, std::size_t n, u32* out,
                                            std::size_t first_poly, hipStream_t s)
    {
        return with_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            residue<LL> rv;
            for (int l = 0; l < LL; ++l)
                rv.v[l] = r2[l];
            cap_stream cs{key, {nonce[0], nonce[1], nonce[2]}, bits};
            k_sample_cbd_capability<LL><<<grid_for(n << (log_n - 1)), block_threads, 0, s>>>(cs, log_n, n, first_poly, rv, out,
                                                                                           n << log_n, P);
            return hipGetLastError();
        });
    }

    // Fill a column tile with capability samples. mode 0: x; mode 1: p x + lift(m); mode 2: p x.
    template<int L>
    __device__ void cols_sample(u32* tile, tile_shape sh, cap_stream const& cs, u64 poly, int log_n, int log_n2, int c0,
                                int mode, u32 const (&p_mont)[L], u32 const* __restrict__ messages, u32 const (&r2)[L],
                                mod_params const& P)
    {
        int plane = sh.plane();
        int C = 1 << sh.log_seqs;
        int half_rows = 1 << (sh.log_s - 1);
        int pairs = half_rows << sh.log_seqs;
        u32 half = 1u << (log_n - 1);
        for (int e = threadIdx.x; e < pairs; e += blockDim.x)
        {
            int k = e >> sh.log_seqs;
            int c = e & (C - 1);
            u32 j = (static_cast<u32>(k) << log_n2) + c0 + c;
            u32 lo[L], hi[L];
            cap_pair<L>(lo, hi, cs, poly, log_n, j, r2, P);
            if (mode != 0)
            {
                mod_mul<L>(lo, lo, p_mont, P);
                mod_mul<L>(hi, hi, p_mont, P);
            }
            if (mode == 1)
            {
                u32 m[L];
                lift_u32<L>(m, messages[(poly << log_n) + j], r2, P);
                mod_add<L>(lo, lo, m, P);
                lift_u32<L>(m, messages[(poly << log_n) + j + half], r2, P);
                mod_add<L>(hi, hi, m, P);
            }
            lds_store<L>(tile, plane, sh.elem(k, c), lo);
            lds_store<L>(tile, plane, sh.elem(k + half_rows, c), hi);
        }
    }

    // capability CBD -> first (column) stages of the forward transform
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_cap_cols(cap_stream cs, u32* __restrict__ out, u32 const* __restrict__ W,
                                                                int log_n, int log_n1, int log_cols, u64 n, residue<L> r2,
                                                                mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.y;
        int log_n2 = log_n - log_n1;
        int c0 = blockIdx.x << log_cols;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        tile_shape sh{log_n1, log_cols, false};
        u32 pm[L] = {};
        cols_sample<L>(lds, sh, cs, p, log_n, log_n2, c0, 0, pm, nullptr, r2.v, P);
        __syncthreads();
        tile_transform<L, false>(lds, sh, W, 1u << log_n, 1u, 0u, false, P);
        cols_store<L>(lds, sh, out, count, p << log_n, log_n2, c0);
    }

    hipError_t launch_sample_cbd_ntt(int L, transform_args const& T, chacha_key const& key, u32 const* nonce, int bits,
                                     std::size_t n, u32* out, hipStream_t s)
    {
        split_plan plan = make_split(L, T.log_n, 1, 1);
        if (plan.cols < 0 || plan.rows < 0)
            return hipErrorInvalidValue;
        return with_limbs(L, [&](auto LC) -> hipError_t {
            constexpr int LL = decltype(LC)::value;
            residue<LL> rv;
            for (int l = 0; l < LL; ++l)
                rv.v[l] = T.r2[l];
            cap_stream cs{key, {nonce[0], nonce[1], nonce[2]}, bits};
            dim3 gc(1u << (plan.log_n2 - plan.cols), static_cast<unsigned>(n));
            dim3 gr(1u << (plan.log_n1 - plan.rows), static_cast<unsigned>(n));
            std::size_t bc = (std::size_t{4} * LL) << (plan.log_n1 + plan.cols);
            std::size_t br = (std::size_t{4} * LL) << (plan.log_n2 + plan.rows);
            k_cap_cols<LL><<<gc, block_threads, bc, s>>>(cs, out, T.fwd, T.log_n, plan.log_n1, plan.cols, n, rv, T.P);
            k_ntt_rows<LL, false><<<gr, block_threads, br, s>>>(out, out, T.fwd, T.log_n, plan.log_n1, plan.rows, n, T.P);
            return hipGetLastError();
        });
    }

    // ------------------------------------------------------------------ batched BGV encryption
    // Linearity of the transform: NTT(v) p + NTT(m) = NTT(p v + m) and NTT(w) p = NTT(p w), so a
    // ciphertext needs three transforms (u, p v + m, p w) instead of four; results are the
    // same residues (exact arithmetic mod q).
    // Pass 1 (columns): sample u, p v + m, p w on the capability streams and run the column stages.
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_bgv_cols(cap_stream su, cap_stream sv, cap_stream sw,
                                                                u32 const* __restrict__ messages, residue<L> p_mont,
                                                                u32* __restrict__ ws, u32 const* __restrict__ W, int log_n,
                                                                int log_n1, int log_cols, u64 n, residue<L> r2, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.y;
        int log_n2 = log_n - log_n1;
        int c0 = blockIdx.x << log_cols;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        tile_shape sh{log_n1, log_cols, false};
        for (int which = 0; which < 3; ++which)
        {
            cap_stream st = which == 0 ? su : (which == 1 ? sv : sw);
            cols_sample<L>(lds, sh, st, p, log_n, log_n2, c0, which, p_mont.v, messages, r2.v, P);
            __syncthreads();
            tile_transform<L, false>(lds, sh, W, 1u << log_n, 1u, 0u, false, P);
            cols_store<L>(lds, sh, ws + static_cast<std::size_t>(which) * L * count, count, p << log_n, log_n2, c0);
            __syncthreads();
        }
    }

    // Pass 2 (rows): finish the three transforms and combine with the public key
    //   c0 = A o U + X0,  c1 = B o U + X1
    template<int L>
    __global__ void __launch_bounds__(block_threads) k_bgv_rows(u32 const* __restrict__ ws, u32 const* __restrict__ pk_a,
                                                                u32 const* __restrict__ pk_b, u32* __restrict__ c0,
                                                                u32* __restrict__ c1, u32 const* __restrict__ W, int log_n,
                                                                int log_n1, int log_rows, u64 n, mod_params P)
    {
        extern __shared__ u32 lds[];
        u64 p = blockIdx.y;
        int log_n2 = log_n - log_n1;
        u32 hi0 = blockIdx.x << log_rows;
        std::size_t count = static_cast<std::size_t>(n) << log_n;
        u32 N = 1u << log_n;
        tile_shape sh{log_n2, log_rows, true};
        int plane = sh.plane();
        u32* tu = lds;
        u32* tx = lds + L * plane;
        std::size_t row0 = static_cast<std::size_t>(hi0) << log_n2;
        std::size_t start = (p << log_n) + row0;
        u32 base = (1u << log_n1) + hi0;
        rows_load<L>(tu, plane, ws, count, start);
        __syncthreads();
        tile_transform<L, false>(tu, sh, W, N, base, 1u, false, P);
        for (int which = 1; which < 3; ++which)
        {
            rows_load<L>(tx, plane, ws + static_cast<std::size_t>(which) * L * count, count, start);
            __syncthreads();
            tile_transform<L, false>(tx, sh, W, N, base, 1u, false, P);
            u32 const* pk = which == 1 ? pk_a : pk_b;
            u32* dst = which == 1 ? c0 : c1;
            for (int e = threadIdx.x; e < plane; e += blockDim.x)
            {
                u32 uu[L], xx[L], kk[L];
                lds_load<L>(uu, tu, plane, e);
                lds_load<L>(xx, tx, plane, e);
                load_planar<L>(kk, pk, N, row0 + e);
                mod_mul<L>(kk, kk, uu, P);
                mod_add<L>(xx, kk, xx, P);
                store_planar<L>(dst, count, start + e, xx);
            }
            __syncthreads();
        }
    }

    std::size_t bgv_workspace_words(int L, int log_n, std::size_t n)
    {
        return 3 * static_cast<std::size_t>(L) * (n << log_n);
    }

    hipError_t launch_bgv_encrypt(int L, transform_args const& T, bgv_args const& B, std::size_t n, u32 const* pk_a,
                                  u32 const* pk_b, u32 const* messages, u32* c0, u32* c1, u32* workspace, hipStream_t s)
    {
        split_plan pc = make_split(L, T.log_n, 1, 2);
        split_plan& pr = pc;
        if (pc.cols < 0 || pc.rows < 0)
            return hipErrorInvalidValue;
        return with_limbs(L, [&](auto LC) -> hipError_t {
            constexpr int LL = decltype(LC)::value;
            residue<LL> rv, pm;
            for (int l = 0; l < LL; ++l)
            {
                rv.v[l] = T.r2[l];
                pm.v[l] = B.p_mont[l];
            }
            cap_stream su{B.key, {B.nonce_u[0], B.nonce_u[1], B.nonce_u[2]}, B.bits_u};
            cap_stream sv{B.key, {B.nonce_v[0], B.nonce_v[1], B.nonce_v[2]}, B.bits_v};
            cap_stream sw{B.key, {B.nonce_w[0], B.nonce_w[1], B.nonce_w[2]}, B.bits_w};
            dim3 gc(1u << (pc.log_n2 - pc.cols), static_cast<unsigned>(n));
            dim3 gr(1u << (pr.log_n1 - pr.rows), static_cast<unsigned>(n));
            std::size_t bc = (std::size_t{4} * LL) << (pc.log_n1 + pc.cols);
            std::size_t br = (std::size_t{8} * LL) << (pr.log_n2 + pr.rows);
            k_bgv_cols<LL><<<gc, block_threads, bc, s>>>(su, sv, sw, messages, pm, workspace, T.fwd, T.log_n, pc.log_n1,
                                                        pc.cols, n, rv, T.P);
            k_bgv_rows<LL><<<gr, block_threads, br, s>>>(workspace, pk_a, pk_b, c0, c1, T.fwd, T.log_n, pr.log_n1, pr.rows,
                                                        n, T.P);
            return hipGetLastError();
        });
    }
}
