// CPU check of the persistent job kernel's indexing (k_jobs in hmpc_amd/csrc/jobs_impl.hpp): the
// access sequence of run_polymul_job and run_bgv_job is replayed for every thread of the workgroup
// with the kernel's own geometry code (hmpc_amd/csrc/tile_geometry.hpp), for both job types, both
// limb policies (4 limbs: 5 workspace planes; 2 limbs: 2), every elements-per-thread choice and
// N = 2^2 .. 2^17, and every index is checked:
//   operands, messages   within the job's own polynomial [off, off + N) of their planes
//   public key           within its N elements
//   outputs              within the job's polynomial of the output planes
//   scratch              within the job's slot (job_scratch_words), planes of N elements
//   twiddles             below N;  LDS elements below the 4096-element tile
// It also counts the second-operand reads that the `valid` guard of with_operands skips for
// whole-tile jobs of degree N < 4096: without that guard those reads would fall outside the N-element
// key and scratch planes. Prints "ok" and the counts, or the first failure.
#include "../../hmpc_amd/csrc/jobs.hpp"
#include "../../hmpc_amd/csrc/kernels.hpp"
#include "../../hmpc_amd/csrc/tile_geometry.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace hmpc_amd;
using namespace hmpc_amd::fast;

namespace
{
    long long checks = 0, guarded = 0;
    int failures = 0;

    void require(bool ok, std::string const& what)
    {
        ++checks;
        if (!ok && failures++ < 20)
            std::printf("FAILED: %s\n", what.c_str());
    }

    struct job_case
    {
        int log_n;
        int E;
        int wsl;          // workspace planes per element
        int io_limbs;     // limbs in global memory (2 or 4)
        bool wave_units;  // unit order of the policy (false for the job kernel's policies)
        std::size_t N, in_count, in_off, out_count, out_off, msg_off;
        u64 slot_words;
    };

    // a buffer region: plane size `count`, allowed element range [lo, hi), planes at word offset
    // `plane0` (scratch: the slot holds `slot_words`)
    struct region
    {
        char const* name;
        std::size_t count, lo, hi;
        u64 plane0;
        int planes;
        u64 slot_words;  // 0: not a scratch region
    };

    void access(job_case const& J, region const& r, std::size_t idx, char const* where)
    {
        require(idx >= r.lo && idx < r.hi && idx < r.count,
                std::string(where) + ": " + r.name + " index " + std::to_string(idx) + " outside [" + std::to_string(r.lo) +
                    ", " + std::to_string(r.hi) + ") N=2^" + std::to_string(J.log_n) + " E=" + std::to_string(J.E));
        if (r.slot_words)
            require(r.plane0 + static_cast<u64>(r.planes - 1) * r.count + idx < r.slot_words,
                    std::string(where) + ": scratch word outside the slot, N=2^" + std::to_string(J.log_n));
    }

    // tile_load / tile_store of the job kernel (no wave-local order): every thread t moves tile
    // elements t, t + T, ... below io.valid
    void tile_io(job_case const& J, io_desc const& io, region const& r, char const* where)
    {
        int const T = tile / J.E;
        for (int t = 0; t < T; ++t)
            for (int e = t; e < tile; e += T)
            {
                require((t & (e - t)) == 0, "tile_load/tile_store element is not e0 | d");
                require(gidx_rel(io, e) < (1u << 28), "tile element offset beyond 2^28");
                require(gidx_rel(io, e) == gidx_rel(io, t) + gidx_rel(io, e - t), "tile offset not additive");
                if (e < io.valid)
                    access(J, r, gidx(io, e), where);
            }
    }

    template<int E, int W>
    void round_accesses(job_case const& J, tile_desc const& td, int a, bool inv, u32 base0, u32 base_step, bool scaled)
    {
        constexpr int K = 1 << W;
        int const T = tile / E;
        for (int t = 0; t < T; ++t)
            for (int g = 0; g < E / K; ++g)
            {
                int u = J.wave_units ? unit_index_at<E, W, true>(t, g) : unit_index_at<E, W, false>(t, g);
                unit_t un = unit_of(td, a, W, inv, u);
                int const el0 = elem_of(td, un.c, j_of(un, 0u, a, W, td.s, inv));
                for (int k = 0; k < K; ++k)
                {
                    int el = elem_of(td, un.c, j_of(un, static_cast<u32>(k), a, W, td.s, inv));
                    require(el >= 0 && el < tile, "LDS element outside the tile");
                    int d = unit_offset(td, a, W, inv, static_cast<u32>(k));
                    require((el0 & d) == 0 && (el0 | d) == el, "unit element is not element 0 | unit_offset(k)");
                }
                u32 base = base0 + un.c * base_step;
                for (int ii = 0; ii < W; ++ii)
                {
                    if (scaled && ii == W - 1)
                        continue;  // the scaled final stage reads twiddles 0 and 1
                    int i = a + ii;
                    u32 h = 1u << (td.s - i - 1);
                    int count = inv ? (1 << (W - ii - 1)) : (1 << ii);
                    for (int m = 0; m < count; ++m)
                    {
                        u32 idx = 0;
                        switch (ii)
                        {
                        case 0: idx = inv ? tw_index<true, W, 0>(base, h, i, un.F, m) : tw_index<false, W, 0>(base, h, i, un.F, m); break;
                        case 1: if constexpr (W > 1) idx = inv ? tw_index<true, W, 1>(base, h, i, un.F, m) : tw_index<false, W, 1>(base, h, i, un.F, m); break;
                        case 2: if constexpr (W > 2) idx = inv ? tw_index<true, W, 2>(base, h, i, un.F, m) : tw_index<false, W, 2>(base, h, i, un.F, m); break;
                        default: if constexpr (W > 3) idx = inv ? tw_index<true, W, 3>(base, h, i, un.F, m) : tw_index<false, W, 3>(base, h, i, un.F, m); break;
                        }
                        require(idx < J.N, "twiddle index " + std::to_string(idx) + " >= N=2^" + std::to_string(J.log_n));
                        u32 X = inv ? tw_round_base<true, W>(base, td.s, a, un.F) : tw_round_base<false, W>(base, td.s, a, un.F);
                        u32 idx2 = 0;
                        switch (ii)
                        {
                        case 0: idx2 = inv ? tw_index_round<true, W, 0>(X, m) : tw_index_round<false, W, 0>(X, m); break;
                        case 1: if constexpr (W > 1) idx2 = inv ? tw_index_round<true, W, 1>(X, m) : tw_index_round<false, W, 1>(X, m); break;
                        case 2: if constexpr (W > 2) idx2 = inv ? tw_index_round<true, W, 2>(X, m) : tw_index_round<false, W, 2>(X, m); break;
                        default: if constexpr (W > 3) idx2 = inv ? tw_index_round<true, W, 3>(X, m) : tw_index_round<false, W, 3>(X, m); break;
                        }
                        require(idx2 == idx, "round-base twiddle index differs");
                    }
                }
            }
    }

    template<int E>
    void rounds(job_case const& J, tile_desc const& td, bool inv, u32 base0, u32 base_step, bool final_scale, int first_round)
    {
        constexpr int WM = ilog2(E);
        int n = round_count<WM>(td.s);
        for (int r = first_round; r < n; ++r)
        {
            int a = round_start<WM>(td.s, r, inv);
            int w = round_width<WM>(td.s, r, inv);
            bool scaled = inv && final_scale && a + w == td.s;
            switch (w)
            {
            case 1: round_accesses<E, 1>(J, td, a, inv, base0, base_step, scaled); break;
            case 2: if constexpr (WM >= 2) round_accesses<E, 2>(J, td, a, inv, base0, base_step, scaled); break;
            case 3: if constexpr (WM >= 3) round_accesses<E, 3>(J, td, a, inv, base0, base_step, scaled); break;
            case 4: if constexpr (WM >= 4) round_accesses<E, 4>(J, td, a, inv, base0, base_step, scaled); break;
            default: require(false, "round width");
            }
        }
    }

    // with_operands after the forward's last round: src (and key) at base + e0 + k for e0 + k < valid
    template<int E, int W>
    void operands_w(job_case const& J, int valid, region const& src, std::size_t src_base, region const* key, std::size_t key_base)
    {
        constexpr int K = 1 << W;
        int const T = tile / E;
        for (int t = 0; t < T; ++t)
            for (int g = 0; g < E / K; ++g)
            {
                int e0 = (J.wave_units ? unit_index_at<E, W, true>(t, g) : unit_index_at<E, W, false>(t, g)) << W;
                for (int k = 0; k < K; ++k)
                {
                    if (e0 + k >= valid)
                    {
                        // skipped by the guard: this read would lie beyond the plane's N elements
                        if (src_base + static_cast<std::size_t>(e0 + k) >= src.hi)
                            ++guarded;
                        continue;
                    }
                    access(J, src, src_base + static_cast<std::size_t>(e0 + k), "second operand");
                    if (key)
                        access(J, *key, key_base + static_cast<std::size_t>(e0 + k), "public key");
                }
            }
    }

    template<int E>
    void operands(job_case const& J, tile_desc const& td, int valid, region const& src, std::size_t src_base, region const* key,
                  std::size_t key_base)
    {
        constexpr int WM = ilog2(E);
        int w = round_width<WM>(td.s, round_count<WM>(td.s) - 1, false);
        switch (w)
        {
        case 1: operands_w<E, 1>(J, valid, src, src_base, key, key_base); break;
        case 2: if constexpr (WM >= 2) operands_w<E, 2>(J, valid, src, src_base, key, key_base); break;
        case 3: if constexpr (WM >= 3) operands_w<E, 3>(J, valid, src, src_base, key, key_base); break;
        case 4: if constexpr (WM >= 4) operands_w<E, 4>(J, valid, src, src_base, key, key_base); break;
        default: require(false, "operand round width");
        }
    }

    void sample(job_case const& J, bool cols, int log_n1, int log_seqs, u32 col0, region const& msg)
    {
        std::size_t half = J.N / 2;
        int pairs = sample_pairs(cols, J.log_n, log_n1, log_seqs);
        for (int e = 0; e < pairs; ++e)
        {
            sample_slot sl = sample_slot_of(cols, e, J.log_n, log_n1, log_seqs, col0);
            require(sl.k < half, "sampled pair index >= N/2");
            require(sl.elo >= 0 && sl.elo < tile && sl.ehi >= 0 && sl.ehi < tile, "sampled LDS element outside the tile");
            access(J, msg, J.msg_off + sl.k, "message");
            access(J, msg, J.msg_off + sl.k + half, "message");
        }
    }

    region scratch_plane(job_case const& J, int which)
    {
        return {"scratch", J.N, 0, J.N, static_cast<u64>(which) * J.wsl * J.N, J.wsl, J.slot_words};
    }

    template<int E>
    void polymul_job(job_case const& J)
    {
        region in{"operand", J.in_count, J.in_off, J.in_off + J.N, 0, J.io_limbs, 0};
        region out{"output", J.out_count, J.out_off, J.out_off + J.N, 0, J.io_limbs, 0};
        int const log_n = J.log_n;
        if (log_n <= tile_log)
        {
            tile_desc td;
            io_desc ia, io, is;
            job_single_tile(log_n, J.in_count, J.in_off, td, ia);
            job_single_tile(log_n, J.out_count, J.out_off, td, io);
            job_single_tile(log_n, J.N, 0, td, is);
            region sa = scratch_plane(J, 0);
            tile_io(J, ia, in, "load a");
            rounds<E>(J, td, false, 1u, 0u, false, 0);
            tile_io(J, is, sa, "store NTT(a)");
            tile_io(J, ia, in, "load b");
            rounds<E>(J, td, false, 1u, 0u, false, 0);
            operands<E>(J, td, is.valid, sa, 0, nullptr, 0);
            rounds<E>(J, td, true, 1u, 0u, true, 0);
            tile_io(J, io, out, "store product");
            return;
        }
        pass_geom const gc = geom_for(log_n, false);
        pass_geom const gr = geom_for(log_n, true);
        region sa = scratch_plane(J, 0), sb = scratch_plane(J, 1);
        for (u32 t = 0; t < 2 * gc.blocks_per_poly; ++t)
        {
            u32 which = t >= gc.blocks_per_poly ? 1u : 0u;
            u32 ct = t - which * gc.blocks_per_poly;
            tile_desc td;
            io_desc ii, is;
            job_cols_tile(gc, ct, J.in_count, J.in_off, td, ii);
            job_cols_tile(gc, ct, J.N, 0, td, is);
            tile_io(J, ii, in, "column load");
            rounds<E>(J, td, false, 1u, 0u, false, 0);
            tile_io(J, is, which ? sb : sa, "column store");
        }
        for (u32 t = 0; t < gr.blocks_per_poly; ++t)
        {
            tile_desc td;
            io_desc is;
            u32 hi0 = job_rows_tile(gr, t, J.N, 0, td, is);
            u32 base0 = (1u << gr.log_n1) + hi0;
            tile_io(J, is, sa, "row load a");
            rounds<E>(J, td, false, base0, 1u, false, 0);
            tile_io(J, is, sa, "row store a");
            tile_io(J, is, sb, "row load b");
            rounds<E>(J, td, false, base0, 1u, false, 0);
            operands<E>(J, td, is.valid, sa, is.base, nullptr, 0);
            rounds<E>(J, td, true, base0, 1u, false, 0);
            tile_io(J, is, sa, "row store product");
        }
        for (u32 t = 0; t < gc.blocks_per_poly; ++t)
        {
            tile_desc td;
            io_desc is, io;
            job_cols_tile(gc, t, J.N, 0, td, is);
            job_cols_tile(gc, t, J.out_count, J.out_off, td, io);
            tile_io(J, is, sa, "inverse column load");
            rounds<E>(J, td, true, 1u, 0u, true, 0);
            tile_io(J, io, out, "inverse column store");
        }
    }

    template<int E>
    void bgv_job(job_case const& J)
    {
        region key{"public key", J.N, 0, J.N, 0, J.io_limbs, 0};
        region out{"ciphertext", J.out_count, J.out_off, J.out_off + J.N, 0, J.io_limbs, 0};
        region msg{"message", J.out_count, J.msg_off, J.msg_off + J.N, 0, 1, 0};
        int const log_n = J.log_n;
        if (log_n <= tile_log)
        {
            tile_desc td;
            io_desc ik, io, is;
            job_single_tile(log_n, J.N, 0, td, ik);
            job_single_tile(log_n, J.out_count, J.out_off, td, io);
            job_single_tile(log_n, J.N, 0, td, is);
            region su = scratch_plane(J, 0);
            sample(J, false, 0, 0, 0u, msg);
            rounds<E>(J, td, false, 1u, 0u, false, 0);
            tile_io(J, is, su, "store NTT(u)");
            for (int which = 1; which < 3; ++which)
            {
                sample(J, false, 0, 0, 0u, msg);
                rounds<E>(J, td, false, 1u, 0u, false, 0);
                operands<E>(J, td, ik.valid, su, 0, &key, ik.base);
                tile_io(J, io, out, "store c");
            }
            return;
        }
        pass_geom const gc = geom_for(log_n, false);
        pass_geom const gr = geom_for(log_n, true);
        for (int which = 0; which < 3; ++which)
        {
            region dst = scratch_plane(J, which);
            for (u32 ct = 0; ct < gc.blocks_per_poly; ++ct)
            {
                tile_desc td;
                io_desc is;
                job_cols_tile(gc, ct, J.N, 0, td, is);
                sample(J, true, gc.log_n1, gc.log_seqs, ct << gc.log_seqs, msg);
                rounds<E>(J, td, false, 1u, 0u, false, 0);
                tile_io(J, is, dst, "column store");
            }
        }
        region su = scratch_plane(J, 0);
        for (u32 t = 0; t < gr.blocks_per_poly; ++t)
        {
            tile_desc td;
            io_desc is, ik, io;
            u32 hi0 = job_rows_tile(gr, t, J.N, 0, td, is);
            job_rows_tile(gr, t, J.N, 0, td, ik);
            job_rows_tile(gr, t, J.out_count, J.out_off, td, io);
            u32 base0 = (1u << gr.log_n1) + hi0;
            tile_io(J, is, su, "row load u");
            rounds<E>(J, td, false, base0, 1u, false, 0);
            tile_io(J, is, su, "row store u");
            for (int which = 1; which < 3; ++which)
            {
                tile_io(J, is, scratch_plane(J, which), "row load x");
                rounds<E>(J, td, false, base0, 1u, false, 0);
                operands<E>(J, td, ik.valid, su, is.base, &key, ik.base);
                tile_io(J, io, out, "row store c");
            }
        }
    }

    template<int E>
    void run_case(int log_n, int io_limbs)
    {
        std::size_t const N = std::size_t{1} << log_n;
        for (u32 type : {job_polymul, job_bgv})
        {
            job_case J{};
            J.log_n = log_n;
            J.E = E;
            J.io_limbs = io_limbs;
            J.wsl = ws_limbs(io_limbs);
            J.wave_units = false;
            J.N = N;
            // the middle job of three: neighbours on both sides
            J.in_count = type == job_bgv ? N : 3 * N;
            J.in_off = type == job_bgv ? 0 : N;
            J.out_count = 3 * N;
            J.out_off = N;
            J.msg_off = N;
            J.slot_words = job_scratch_words(type, log_n, io_limbs);
            if (type == job_polymul)
                polymul_job<E>(J);
            else
                bgv_job<E>(J);
        }
    }
}

// The kernels address LDS as swizzle(e0) ^ swizzle(d) for elements e0 | d with disjoint bits:
// every swizzle is checked linear on all such pairs of the tile and a permutation of the tile's
// slots, and the unit decomposition (unit element k = element 0 | unit_offset(k)) is checked for
// every tile geometry, round and unit order, and the wave-local tile_load / tile_store order.
template<int W, int II>
void tw_round_stage(bool inv, u32 base, int s, int a, u32 F)
{
    if constexpr (II < W)
    {
        int i = a + II;
        u32 h = 1u << (s - i - 1);
        int count = inv ? (1 << (W - II - 1)) : (1 << II);
        u32 X = inv ? tw_round_base<true, W>(base, s, a, F) : tw_round_base<false, W>(base, s, a, F);
        for (int m = 0; m < count; ++m)
        {
            u32 want = inv ? tw_index<true, W, II>(base, h, i, F, m) : tw_index<false, W, II>(base, h, i, F, m);
            u32 got = inv ? tw_index_round<true, W, II>(X, m) : tw_index_round<false, W, II>(X, m);
            require(got == want, "round-base twiddle index differs");
        }
        tw_round_stage<W, II + 1>(inv, base, s, a, F);
    }
}

template<int W>
void tw_round_check(bool inv, u32 base, int s, int a, u32 F)
{
    tw_round_stage<W, 0>(inv, base, s, a, F);
}

template<int E, int W>
void unit_offsets_all()
{
    for (int s = W; s <= tile_log; ++s)
        for (int rows = 0; rows < 2; ++rows)
            for (int inv = 0; inv < 2; ++inv)
                for (int a = 0; a + W <= s; ++a)
                    for (int wave = 0; wave < 2; ++wave)
                    {
                        tile_desc td{s, tile_log - s, rows != 0};
                        for (int t = 0; t < tile / E; ++t)
                            for (int g = 0; g < E >> W; ++g)
                            {
                                int u = wave ? unit_index_at<E, W, true>(t, g) : unit_index_at<E, W, false>(t, g);
                                unit_t un = unit_of(td, a, W, inv != 0, u);
                                int el0 = elem_of(td, un.c, j_of(un, 0u, a, W, s, inv != 0));
                                for (int k = 0; k < (1 << W); ++k)
                                {
                                    int el = elem_of(td, un.c, j_of(un, static_cast<u32>(k), a, W, s, inv != 0));
                                    int d = unit_offset(td, a, W, inv != 0, static_cast<u32>(k));
                                    require((el0 & d) == 0 && (el0 | d) == el && el < tile,
                                            "unit offset s=" + std::to_string(s) + " a=" + std::to_string(a) +
                                                " W=" + std::to_string(W) + " E=" + std::to_string(E));
                                }
                                // twiddle indices: whole-polynomial / column base 1, row bases N1 + row
                                for (u32 base : {1u, 128u + un.c, 256u + 3u * un.c})
                                    tw_round_check<W>(inv != 0, base, s, a, un.F);
                            }
                    }
}

template<int E>
void wave_local_io()
{
    int const T = tile / E;
    for (int t = 0; t < T; ++t)
    {
        int e0 = (t & ~63) * E + (t & 63);
        for (int e = e0; e < e0 + E * 64; e += 64)
        {
            require((e0 & (e - e0)) == 0 && e < tile, "wave-local tile_load element is not e0 | d");
            // offsets of row tiles and of column tiles of every width
            for (int ls = 0; ls <= 10; ++ls)
            {
                io_desc io{0, 0, 1u << 13, ls, ls == 0, tile};
                require(gidx_rel(io, e) == gidx_rel(io, e0) + gidx_rel(io, e - e0), "wave-local tile offset not additive");
            }
        }
    }
}

void swizzles()
{
    int (*const fns[2])(int) = {swizzle_16, swizzle_rows};
    for (auto f : fns)
    {
        std::vector<char> seen(tile, 0);
        for (int e = 0; e < tile; ++e)
        {
            int p = f(e);
            require(p >= 0 && p < tile && !seen[p], "swizzle is not a permutation of the tile");
            if (p >= 0 && p < tile)
                seen[p] = 1;
            // every d whose bits are disjoint from e
            int const free = (tile - 1) & ~e;
            for (int d = free;; d = (d - 1) & free)
            {
                if (f(e | d) != (f(e) ^ f(d)) && failures++ < 20)
                    std::printf("FAILED: swizzle not linear at e=%d d=%d\n", e, d);
                ++checks;
                if (d == 0)
                    break;
            }
        }
    }
}

// The streaming column passes of the radix-2^26 product (k_polymul_cols_stream, k_icols_stream in
// hmpc_amd/csrc/fast_impl.hpp): 2^7-point columns, 32 columns per tile, 8 elements per thread, 512
// threads. Their twiddles come from scalar loads at lane 0's index (3-stage rounds: forward stages
// 0-5, inverse stages 1-6), which needs the index to be the same on every lane of a wave, and from a
// lane-0 / lane-63 pair (the pair layout of forward stage 6 and inverse stage 0), which needs at most
// those two values per wave; their direct loads and stores address element e0 + d as the offset of
// e0 plus the offset of d. Checked for every thread, unit and element at N = 2^14 .. 2^17 (every
// degree whose column length is 2^7 in the radix-2^26 product), together with: each pass reads and
// writes every tile element exactly once, inside the tile's polynomial.
template<int W>
int unit_elem_at(tile_desc const& td, int a, bool inv, int t, int g, int k)
{
    unit_t un = unit_of(td, a, W, inv, unit_index_at<8, W, false>(t, g));
    return elem_of(td, un.c, j_of(un, static_cast<u32>(k), a, W, td.s, inv));
}

template<int W>
u32 round_x(tile_desc const& td, int a, bool inv, int t, int g)
{
    unit_t un = unit_of(td, a, W, inv, unit_index_at<8, W, false>(t, g));
    return inv ? tw_round_base<true, W>(1u, td.s, a, un.F) : tw_round_base<false, W>(1u, td.s, a, un.F);
}

template<int W>
void elements_once(tile_desc const& td, int a, bool inv, io_desc const& io, std::size_t poly_lo, std::size_t N, char const* what)
{
    std::vector<int> seen(tile, 0);
    for (int t = 0; t < 512; ++t)
    {
        int const e0 = unit_elem_at<W>(td, a, inv, t, 0, 0);
        for (int g = 0; g < (8 >> W); ++g)
            for (int k = 0; k < (1 << W); ++k)
            {
                int const e = unit_elem_at<W>(td, a, inv, t, g, k);
                int const d = e - e0;
                require(e >= 0 && e < tile && d >= 0 && (e0 & d) == 0, std::string(what) + ": element is not e0 | d");
                require(gidx_rel(io, e) == gidx_rel(io, e0) + gidx_rel(io, d), std::string(what) + ": offset not additive");
                std::size_t const gi = io.base + gidx_rel(io, e);
                require(gi >= poly_lo && gi < poly_lo + N, std::string(what) + ": element outside its polynomial");
                if (e >= 0 && e < tile)
                    ++seen[e];
            }
    }
    for (int e = 0; e < tile; ++e)
        require(seen[e] == 1, std::string(what) + ": tile element not moved exactly once");
}

// the pair layout (cols_pair_row): every tile element once, inside its polynomial, offsets additive
void pairs_once(io_desc const& io, std::size_t poly_lo, std::size_t N, char const* what)
{
    std::vector<int> seen(tile, 0);
    for (int t = 0; t < 512; ++t)
    {
        int const e0 = cols_pair_elem(t, 0, 0);
        for (int m = 0; m < 4; ++m)
            for (int k = 0; k < 2; ++k)
            {
                int const e = cols_pair_elem(t, m, k);
                int const d = e - e0;
                require(e >= 0 && e < tile && d >= 0 && (e0 & d) == 0, std::string(what) + ": element is not e0 | d");
                require(gidx_rel(io, e) == gidx_rel(io, e0) + gidx_rel(io, d), std::string(what) + ": offset not additive");
                std::size_t const gi = io.base + gidx_rel(io, e);
                require(gi >= poly_lo && gi < poly_lo + N, std::string(what) + ": element outside its polynomial");
                if (e >= 0 && e < tile)
                    ++seen[e];
            }
    }
    for (int e = 0; e < tile; ++e)
        require(seen[e] == 1, std::string(what) + ": tile element not moved exactly once");
}

// The 3-stage rounds next to the pair layout (forward stages 3-5, inverse stages 1-3) hold rows
// 16 (t >> 6) + 2 k + ((t >> 5) & 1) of column t & 31; the half-wave exchange (element 2m + 1 of the
// lower half-wave <-> element 2m of the upper half, as v_permlane32_swap does) turns that layout into
// the pair layout and back, the pairs are the partners of forward stage 6 / inverse stage 0 (rows j
// and j + 1, j even), and their twiddle is the generic stage index.
void pair_exchange()
{
    tile_desc const td{7, tile_log - 7, false};
    auto round_row = [&](int t, int k, bool inv) {
        unit_t un = unit_of(td, inv ? 1 : 3, 3, inv, unit_index_at<8, 3, false>(t, 0));
        require(static_cast<int>(un.c) == (t & 31), "3-stage round: unit column is not t & 31");
        return static_cast<int>(j_of(un, static_cast<u32>(k), inv ? 1 : 3, 3, 7, inv));
    };
    for (int inv = 0; inv < 2; ++inv)
        for (int t = 0; t < 512; ++t)
        {
            bool const upper = (t & 32) != 0;
            for (int m = 0; m < 4; ++m)
            {
                // what the exchange leaves in this lane's elements 2m, 2m + 1
                int const lo = upper ? t - 32 : t, hi = lo + 32;
                int const after0 = upper ? round_row(lo, 2 * m + 1, inv != 0) : round_row(lo, 2 * m, inv != 0);
                int const after1 = upper ? round_row(hi, 2 * m + 1, inv != 0) : round_row(hi, 2 * m, inv != 0);
                require(after0 == cols_pair_row(t, m, 0) && after1 == cols_pair_row(t, m, 1),
                        "half-wave exchange does not give the pair layout");
                int const j = cols_pair_row(t, m, 0);
                require((j & 1) == 0 && cols_pair_row(t, m, 1) == j + 1, "pair is not (j, j + 1), j even");
                // forward stage 6 (base 1, F = j >> 1) and inverse stage 0 use the same table entry
                u32 const fwd = tw_index<false, 1, 0>(1u, 1u, 6, static_cast<u32>(j >> 1), 0);
                u32 const inv0 = tw_index<true, 1, 0>(1u, 1u << 6, 0, static_cast<u32>(j >> 1), 0);
                require(cols_pair_twiddle(t, m) == fwd && cols_pair_twiddle(t, m) == inv0, "pair twiddle index");
            }
        }
}

void streaming_columns()
{
    pair_exchange();
    for (int log_n = 14; log_n <= 17; ++log_n)
    {
        pass_geom const G = geom_for(log_n, false);
        require(G.log_n1 == 7 && G.log_seqs == 5, "streaming columns need 2^7-point columns, 32 per tile");
        tile_desc const td{7, tile_log - 7, false};
        std::size_t const N = std::size_t{1} << log_n, n = 2, count = n * N;
        for (u32 ct = 0; ct < n * G.blocks_per_poly; ++ct)
        {
            tile_desc tdd;
            io_desc io;
            cols_setup(G, ct, count, tdd, io);
            require(tdd.s == td.s && tdd.log_seqs == td.log_seqs && !tdd.rows, "column tile geometry");
            std::size_t const poly_lo = (ct / G.blocks_per_poly) * N;
            elements_once<3>(td, 0, false, io, poly_lo, N, "forward round-0 loads");
            pairs_once(io, poly_lo, N, "forward stage-6 stores / inverse stage-0 loads (pair layout)");
            elements_once<3>(td, 4, true, io, poly_lo, N, "inverse last-round stores");
        }
        for (int w = 0; w < 8; ++w)
            for (int lane = 0; lane < 64; ++lane)
            {
                int const t = 64 * w + lane;
                // per-wave uniform rounds (the per-stage indices are functions of X)
                require(round_x<3>(td, 0, false, t, 0) == round_x<3>(td, 0, false, 64 * w, 0), "forward stages 0-2 not wave-uniform");
                require(round_x<3>(td, 3, false, t, 0) == round_x<3>(td, 3, false, 64 * w, 0), "forward stages 3-5 not wave-uniform");
                require(round_x<3>(td, 1, true, t, 0) == round_x<3>(td, 1, true, 64 * w, 0), "inverse stages 1-3 not wave-uniform");
                require(round_x<3>(td, 4, true, t, 0) == round_x<3>(td, 4, true, 64 * w, 0), "inverse stages 4-6 not wave-uniform");
                for (int m = 0; m < 4; ++m)
                {
                    // the pair layout's twiddles: lane 0's or lane 63's index, inside the 2^7 entries
                    u32 const x = cols_pair_twiddle(t, m);
                    require(x == cols_pair_twiddle(64 * w, m) || x == cols_pair_twiddle(64 * w + 63, m),
                            "pair twiddles other than lane 0's and lane 63's");
                    require(x < 128u, "pair twiddle index");
                }
                require(tw_index_round<false, 3, 2>(round_x<3>(td, 3, false, t, 0), 3) < 128u, "forward twiddle index");
                require(tw_index_round<true, 3, 0>(round_x<3>(td, 1, true, t, 0), 3) < 128u, "inverse twiddle index");
                static_assert(r26_tw_entries >= 128, "the streaming columns' twiddles are expanded");
            }
    }
}

// The product's row kernel with one wave per workgroup (arith_r26_rows1, 8 elements per thread,
// rows of 2^7 .. 2^9): wave w of a tile may touch only the tile elements [512 w, 512 w + 512) in every
// round of the forward and inverse row transforms and in its loads and stores, and its LDS slot
// swizzle_rows(e) & 511 must map those 512 elements one-to-one onto the 512 slots of its block,
// XOR-linearly (the kernels address slot(e0 | d) as slot(e0) ^ slot(d)).
template<int W>
void rows1_rounds(int s)
{
    tile_desc td{s, tile_log - s, true};
    for (int inv = 0; inv < 2; ++inv)
        for (int a = 0; a + W <= s; ++a)
            for (int t = 0; t < 512; ++t)
                for (int g = 0; g < (8 >> W); ++g)
                {
                    unit_t un = unit_of(td, a, W, inv != 0, unit_index_at<8, W, true>(t, g));
                    for (int k = 0; k < (1 << W); ++k)
                    {
                        int e = elem_of(td, un.c, j_of(un, static_cast<u32>(k), a, W, s, inv != 0));
                        require(e >> 9 == t >> 6, "one-wave rows: a round reaches another wave's elements, s=" + std::to_string(s));
                    }
                }
}

void rows1_blocks()
{
    for (int w = 0; w < 8; ++w)
    {
        std::vector<int> seen(512, 0);
        for (int e = 512 * w; e < 512 * w + 512; ++e)
        {
            int const p = swizzle_rows(e) & 511;
            require(!seen[p], "one-wave rows: slot map is not one-to-one on the wave's block");
            seen[p] = 1;
            int const free = 511 & ~e;
            for (int d = free;; d = (d - 1) & free)
            {
                require((swizzle_rows(e | d) & 511) == ((swizzle_rows(e) & 511) ^ (swizzle_rows(d) & 511)),
                        "one-wave rows: slot map not XOR-linear");
                if (d == 0)
                    break;
            }
        }
    }
    // wave-local tile loads and stores: thread t moves elements e0 + 64 i, e0 = (t & ~63) * 8 + (t & 63)
    for (int t = 0; t < 512; ++t)
        for (int i = 0; i < 8; ++i)
            require(((t & ~63) * 8 + (t & 63) + 64 * i) >> 9 == t >> 6, "one-wave rows: tile load leaves the wave's block");
    for (int s = 7; s <= 9; ++s)
    {
        rows1_rounds<1>(s);
        rows1_rounds<2>(s);
        rows1_rounds<3>(s);
    }
}

// The one-wave row kernel at rows of 2^9 takes the twiddles of the forward transforms' first
// round (stages 0-2) and of the inverse's last round (stages 6-8) per wave, from the expanded
// tables: every lane of a wave must have lane 0's row c and top bits F in those rounds (so the
// round's twiddle bases agree), and for N = 2^16 (columns of 2^7: base = 2^7 + row, rows 0..127)
// every index must lie inside the expanded tables (r26_tw_entries).
void rows1_uniform()
{
    // every degree whose rows have 2^9 elements has columns of at most 2^7: the rows' indices
    // (base 2^log_n1 + row) stay inside the expanded tables
    for (int log_n = 8; log_n <= 20; ++log_n)
    {
        pass_geom const g = geom_for(log_n, true);
        if (g.log_n2 == 9)
            require((std::size_t{8} << g.log_n1) <= r26_tw_entries, "rows of 2^9 beyond the expanded tables");
    }
    tile_desc const td{9, tile_log - 9, true};
    for (int inv = 0; inv < 2; ++inv)
    {
        int const a = inv ? 6 : 0;
        for (int t = 0; t < 512; ++t)
        {
            unit_t const un = unit_of(td, a, 3, inv != 0, unit_index_at<8, 3, true>(t, 0));
            unit_t const u0 = unit_of(td, a, 3, inv != 0, unit_index_at<8, 3, true>(t & ~63, 0));
            require(un.c == u0.c && un.F == u0.F, "one-wave rows of 2^9: a round on expanded twiddles is not wave-uniform");
            for (u32 row = 0; row < 128; ++row)
            {
                u32 const base = 128u + row;
                for (int ii = 0; ii < 3; ++ii)
                {
                    u32 const X = inv ? tw_round_base<true, 3>(base, 9, a, un.F) : tw_round_base<false, 3>(base, 9, a, un.F);
                    int const count = inv ? (1 << (3 - ii - 1)) : (1 << ii);
                    for (int m = 0; m < count; ++m)
                    {
                        u32 const idx = inv ? (ii == 0 ? tw_index_round<true, 3, 0>(X, m)
                                                       : ii == 1 ? tw_index_round<true, 3, 1>(X, m) : tw_index_round<true, 3, 2>(X, m))
                                            : (ii == 0 ? tw_index_round<false, 3, 0>(X, m)
                                                       : ii == 1 ? tw_index_round<false, 3, 1>(X, m) : tw_index_round<false, 3, 2>(X, m));
                        require(idx < r26_tw_entries, "one-wave rows of 2^9: twiddle index outside the expanded tables");
                    }
                }
            }
        }
    }
}

int main()
{
    swizzles();
    streaming_columns();
    rows1_blocks();
    rows1_uniform();
    unit_offsets_all<4, 1>();
    unit_offsets_all<4, 2>();
    unit_offsets_all<8, 1>();
    unit_offsets_all<8, 2>();
    unit_offsets_all<8, 3>();
    unit_offsets_all<16, 1>();
    unit_offsets_all<16, 2>();
    unit_offsets_all<16, 3>();
    unit_offsets_all<16, 4>();
    wave_local_io<4>();
    wave_local_io<8>();
    wave_local_io<16>();
    for (int log_n = 2; log_n <= 17; ++log_n)
    {
        // 4 limbs (radix-2^26 policy): 4 or 8 elements per thread; 2 limbs (Goldilocks): 8 or 16
        run_case<4>(log_n, 4);
        run_case<8>(log_n, 4);
        run_case<8>(log_n, 2);
        run_case<16>(log_n, 2);
    }
    if (guarded == 0)
        require(false, "expected second-operand reads beyond N for whole-tile jobs of N < 4096");
    std::printf("%s checks=%lld guarded=%lld\n", failures ? "FAILED" : "ok", checks, guarded);
    return failures ? 1 : 0;
}
