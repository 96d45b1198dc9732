// Tile geometry of the fast transform kernels (fast_impl.hpp, jobs_impl.hpp): how a 4096-element
// LDS tile is laid out, which elements and twiddles each thread's units touch in a round, and how
// tiles map to global (limb-planar) indices in the column, row and whole-polynomial passes. Pure
// index arithmetic, callable on the host as well, so that the indexing of the kernels is checked on
// the CPU (tests/cpp/jobs_index_check.cpp) with the code the GPU runs.
#pragma once

#include "modarith.hpp"

#include <cstddef>

namespace hmpc_amd
{
    namespace fast
    {
        constexpr int tile_log = 12;
        constexpr int tile = 1 << tile_log;

        constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

        // ------------------------------------------------------------------ tile geometry
        struct tile_desc
        {
            int s;         // log2 of the sub-transform length
            int log_seqs;  // s + log_seqs = 12
            bool rows;     // row layout (c << s) + j, else column layout (j << log_seqs) + c
        };

        struct unit_t
        {
            u32 c, F, Lo;
        };

        HMPC_HD unit_t unit_of(tile_desc const& td, int a, int w, bool inv, int u)
        {
            int fbits = td.s - w;
            u32 f, c;
            if (td.rows)
            {
                f = static_cast<u32>(u) & ((1u << fbits) - 1u);
                c = static_cast<u32>(u) >> fbits;
            }
            else
            {
                c = static_cast<u32>(u) & ((1u << td.log_seqs) - 1u);
                f = static_cast<u32>(u) >> td.log_seqs;
            }
            unit_t r;
            r.c = c;
            if (!inv)
            {
                int lo_bits = td.s - a - w;
                r.F = f >> lo_bits;
                r.Lo = f & ((1u << lo_bits) - 1u);
            }
            else
            {
                r.F = f >> a;
                r.Lo = f & ((1u << a) - 1u);
            }
            return r;
        }

        HMPC_HD u32 j_of(unit_t const& un, u32 k, int a, int w, int s, bool inv)
        {
            return inv ? ((un.F << (a + w)) | (k << a) | un.Lo) : ((un.F << (s - a)) | (k << (s - a - w)) | un.Lo);
        }

        HMPC_HD int elem_of(tile_desc const& td, u32 c, u32 j)
        {
            return td.rows ? static_cast<int>((c << td.s) + j) : static_cast<int>((j << td.log_seqs) + c);
        }

        // LDS slot swizzles of the fast kernels' tiles (XOR of shifted bit fields: linear over GF(2))
        HMPC_HD int swizzle_16(int e) { return e ^ ((e >> 4) & 15); }
        HMPC_HD int swizzle_rows(int e) { return e ^ ((e >> 3) & 63) ^ ((e >> 10) & 15); }

        // element k of a unit relative to its element 0: elem_of(un.c, j_of(un, k)) equals
        // elem_of(un.c, j_of(un, 0)) | unit_offset(k), the two with disjoint bits
        HMPC_HD int unit_offset(tile_desc const& td, int a, int w, bool inv, u32 k)
        {
            return elem_of(td, 0u, k << (inv ? a : td.s - a - w));
        }

        // round schedule: forward full rounds first, inverse partial round first
        template<int WMAX>
        HMPC_HD int round_count(int s) { return (s + WMAX - 1) / WMAX; }
        template<int WMAX>
        HMPC_HD int round_start(int s, int r, bool inv)
        {
            int part = s % WMAX;
            if (!inv || part == 0)
                return WMAX * r;
            return r == 0 ? 0 : part + WMAX * (r - 1);
        }
        template<int WMAX>
        __host__ HMPC_HD int round_width(int s, int r, bool inv)
        {
            int part = s % WMAX;
            int n = (s + WMAX - 1) / WMAX;
            if (part == 0)
                return WMAX;
            if (!inv)
                return r == n - 1 ? part : WMAX;
            return r == 0 ? part : WMAX;
        }

        // global (limb-planar) <-> LDS tile. Element e of the tile maps to global index
        // base + e (rows) or base + (e >> log_seqs) * stride + (e & (C - 1)) (columns).
        struct io_desc
        {
            std::size_t base;   // global element index of tile element 0
            std::size_t count;  // plane size (elements per limb plane)
            u32 stride;         // column layout: distance between rows (N2)
            int log_seqs;
            bool rows;
            int valid;          // tile elements backed by real data (tail tiles)
        };

        // offset of tile element e from the tile's element 0 (io.base): below 2^12 for row tiles,
        // below 2^(12 - log_seqs) stride for column tiles (2^20 at N = 2^20)
        HMPC_HD u32 gidx_rel(io_desc const& io, int e)
        {
            if (io.rows)
                return static_cast<u32>(e);
            return static_cast<u32>(e >> io.log_seqs) * io.stride + static_cast<u32>(e & ((1 << io.log_seqs) - 1));
        }

        HMPC_HD std::size_t gidx(io_desc const& io, int e) { return io.base + gidx_rel(io, e); }

        // The lazy forward transform runs up to 29 stages without reductions (r26.hpp); a
        // forward column result of a transform longer than 18 stages is folded before it goes to
        // the workspace. (No prime 2^124 + c with c < 2^24 has 2^19 | q - 1, so radix-2^26
        // transforms stop at N = 2^17 and the fold is a guard only.)
        __host__ HMPC_HD bool ws_fold_columns(int log_n) { return log_n > 18; }

        // The pair layout of the streaming column passes (2^7-point columns, 32 columns per tile, 8
        // elements per thread, 512 threads): thread t holds the four pairs (j, j + 1) of rows
        // j = 16 (t >> 6) + 4 m + 2 h, m = 0..3, h = (t >> 5) & 1, of column t & 31. In the columns'
        // 3-stage rounds starting at stage 3 (forward) or 1 (inverse) thread t holds rows
        // 16 (t >> 6) + 2 k + h, k = 0..7: the two layouts differ by exchanging the odd elements of
        // the lower half-wave with the even elements of the upper half (one v_permlane32_swap per
        // word and pair), and the pairs are the partners of forward stage 6 and inverse stage 0,
        // whose twiddle is entry 2^6 + (j >> 1) of the column table.
        HMPC_HD int cols_pair_row(int t, int m, int k) { return 16 * (t >> 6) + 4 * m + 2 * ((t >> 5) & 1) + k; }
        HMPC_HD u32 cols_pair_twiddle(int t, int m) { return static_cast<u32>(64 + (cols_pair_row(t, m, 0) >> 1)); }
        // tile element of pair m, element k of thread t
        HMPC_HD int cols_pair_elem(int t, int m, int k) { return (cols_pair_row(t, m, k) << (tile_log - 7)) + (t & 31); }

        struct pass_geom
        {
            int log_n, log_n1, log_n2;
            int log_seqs;  // sequences per tile (columns, rows or polynomials)
            bool single;
            u32 blocks_per_poly;
        };

        HMPC_HD void cols_setup(pass_geom const& G, u32 bid, std::size_t count, tile_desc& td, io_desc& io)
        {
            u32 p = bid / G.blocks_per_poly;
            u32 cb = bid - p * G.blocks_per_poly;
            td = tile_desc{G.log_n1, G.log_seqs, false};
            io.base = (static_cast<std::size_t>(p) << G.log_n) + (static_cast<std::size_t>(cb) << G.log_seqs);
            io.count = count;
            io.stride = 1u << G.log_n2;
            io.log_seqs = G.log_seqs;
            io.rows = false;
            io.valid = tile;
        }

        HMPC_HD u32 rows_setup(pass_geom const& G, u32 bid, std::size_t count, tile_desc& td, io_desc& io)
        {
            u32 p = bid / G.blocks_per_poly;
            u32 rb = bid - p * G.blocks_per_poly;
            u32 hi0 = rb << G.log_seqs;
            td = tile_desc{G.log_n2, G.log_seqs, true};
            io.base = (static_cast<std::size_t>(p) << G.log_n) + (static_cast<std::size_t>(hi0) << G.log_n2);
            io.count = count;
            io.stride = 0;
            io.log_seqs = G.log_seqs;
            io.rows = true;
            io.valid = tile;
            return hi0;
        }

        HMPC_HD void single_setup(pass_geom const& G, u32 bid, std::size_t n, tile_desc& td, io_desc& io)
        {
            td = tile_desc{G.log_n, G.log_seqs, true};
            std::size_t first = static_cast<std::size_t>(bid) << G.log_seqs;
            std::size_t polys = n - first < (std::size_t{1} << G.log_seqs) ? n - first : (std::size_t{1} << G.log_seqs);
            io.base = first << G.log_n;
            io.count = n << G.log_n;
            io.stride = 0;
            io.log_seqs = G.log_seqs;
            io.rows = true;
            io.valid = static_cast<int>(polys << G.log_n);
        }

        HMPC_HD pass_geom geom_for(int log_n, bool rows_pass)
        {
            pass_geom G{};
            G.log_n = log_n;
            if (log_n <= tile_log)
            {
                G.single = true;
                G.log_n1 = log_n;
                G.log_n2 = 0;
                G.log_seqs = tile_log - log_n;
                G.blocks_per_poly = 1;
                return G;
            }
            G.single = false;
            // column tiles of at least 32 columns: each tile row is then 128 contiguous bytes per
            // limb plane (whole cache lines; 16-column tiles fetched every line twice)
            G.log_n1 = log_n / 2 < 7 ? log_n / 2 : 7;
            if (log_n - G.log_n1 > tile_log)
                G.log_n1 = log_n - tile_log;
            G.log_n2 = log_n - G.log_n1;
            if (rows_pass)
            {
                G.log_seqs = tile_log - G.log_n2;
                G.blocks_per_poly = 1u << (G.log_n1 - G.log_seqs);
            }
            else
            {
                G.log_seqs = tile_log - G.log_n1;
                G.blocks_per_poly = 1u << (G.log_n2 - G.log_seqs);
            }
            return G;
        }

        // ------------------------------------------------------------------ tile placement
        HMPC_HD void job_cols_tile(pass_geom const& G, u32 t, std::size_t count, std::size_t off,
                                                      tile_desc& td, io_desc& io)
        {
            cols_setup(G, t, count, td, io);
            io.base += off;
        }

        HMPC_HD u32 job_rows_tile(pass_geom const& G, u32 t, std::size_t count, std::size_t off,
                                                     tile_desc& td, io_desc& io)
        {
            u32 hi0 = rows_setup(G, t, count, td, io);
            io.base += off;
            return hi0;
        }

        HMPC_HD void job_single_tile(int log_n, std::size_t count, std::size_t off, tile_desc& td,
                                                        io_desc& io)
        {
            td = tile_desc{log_n, tile_log - log_n, true};
            io.base = off;
            io.count = count;
            io.stride = 0;
            io.log_seqs = tile_log - log_n;
            io.rows = true;
            io.valid = 1 << log_n;
        }

        // Unit g of thread t in a round of width W (E elements per thread): t + (tile / E) g, or with
        // WaveUnits the wave-major order in which each wave owns the 64 E / 2^W consecutive units
        // [wave 64 E / 2^W, (wave + 1) 64 E / 2^W), lane-consecutive within each g. In a row-layout
        // tile whose rows are at most 64 E elements long every round of a wave then touches only
        // the wave's own rows (tile elements [wave 64 E, (wave + 1) 64 E)). Both orders are the same
        // for full-width rounds (2^W = E, one unit per thread).
        template<int E, int W, bool WaveUnits>
        HMPC_HD int unit_index_at(int t, int g)
        {
            if constexpr ((1 << W) == E)
                return t;
            else if constexpr (WaveUnits)
                return ((((t >> 6) * (E >> W)) + g) << 6) | (t & 63);
            else
                return t + (tile / E) * g;
        }

        // twiddle table index of entry m of stage II of a round of width W for a unit with top
        // bits F (base: 1, or N1 + hi for a row; h = 2^(s - i - 1) with i = a + II)
        template<bool Inv, int W, int II>
        HMPC_HD u32 tw_index(u32 base, u32 h, int i, u32 F, int m)
        {
            return Inv ? base * h + ((F << (W - II - 1)) | static_cast<u32>(m)) : (base << i) + ((F << II) | static_cast<u32>(m));
        }

        // The same indices from one value per unit and round: every twiddle of a round of width W
        // starting at stage a is (X << shift(II)) + m, X = (base << a) + F forward and
        // (base << (s - a - W)) + F inverse, shift(II) = II forward and W - II - 1 inverse.
        template<bool Inv, int W>
        HMPC_HD u32 tw_round_base(u32 base, int s, int a, u32 F)
        {
            return Inv ? (base << (s - a - W)) + F : (base << a) + F;
        }

        template<bool Inv, int W, int II>
        HMPC_HD u32 tw_index_round(u32 X, int m)
        {
            return (X << (Inv ? W - II - 1 : II)) + static_cast<u32>(m);
        }

        // capability sampling into a tile: pair e of the tile's pairs draws coefficients k and
        // k + N/2 of its polynomial (one ChaCha20 block with counter poly N + k) into tile elements
        // elo and ehi. Column tiles (cols) hold rows r and r + N1/2 of each of their 2^log_seqs
        // columns, starting at column col0; whole-polynomial tiles hold the polynomial in order.
        struct sample_slot
        {
            u32 k;
            int elo, ehi;
        };

        HMPC_HD int sample_pairs(bool cols, int log_n, int log_n1, int log_seqs)
        {
            return cols ? ((1 << (log_n1 - 1)) << log_seqs) : (1 << (log_n - 1));
        }

        HMPC_HD sample_slot sample_slot_of(bool cols, int e, int log_n, int log_n1, int log_seqs, u32 col0)
        {
            sample_slot s;
            if (cols)
            {
                int r = e >> log_seqs;
                int c = e & ((1 << log_seqs) - 1);
                s.k = (static_cast<u32>(r) << (log_n - log_n1)) + col0 + static_cast<u32>(c);
                s.elo = e;
                s.ehi = e + ((1 << (log_n1 - 1)) << log_seqs);
            }
            else
            {
                s.k = static_cast<u32>(e);
                s.elo = e;
                s.ehi = e + (1 << (log_n - 1));
            }
            return s;
        }
    }
}
