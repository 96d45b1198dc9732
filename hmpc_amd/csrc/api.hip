// C ABI of libhmpc_hip.so (declared in include/hmpc_hip.h): contexts, registered moduli with
// resident twiddle tables, device memory, and the dispatch of every compute entry point to
// the gfx950 kernels in kernels.hip. Host-side number theory (Montgomery constants, the root
// search of hmpc/ints/num/theory/root_of_unity.hpp, inverses) runs here on the CPU with the
// same limb templates the kernels use.
#include "kernels.hpp"
#include "fastmath.hpp"
#include "server_pool.hpp"

#include "../../include/hmpc_hip.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <type_traits>
#include <deque>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

using namespace hmpc_amd;

namespace
{
    thread_local std::string last_error;

    int fail(int code, std::string const& msg)
    {
        last_error = msg;
        return code;
    }

    int hip_fail(hipError_t e, char const* where)
    {
        return fail(e == hipErrorOutOfMemory ? HMPC_ERR_OUT_OF_MEMORY : HMPC_ERR_DEVICE,
                    std::string(where) + ": " + hipGetErrorString(e));
    }

#define HMPC_HIP(call)                               \
    do                                               \
    {                                                \
        hipError_t e_ = (call);                      \
        if (e_ != hipSuccess)                        \
            return hip_fail(e_, #call);              \
    } while (0)

    template<typename F>
    auto host_limbs(int L, F&& f)
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
        default: return f(std::integral_constant<int, 9>{});
        }
    }

    // ---------------------------------------------------------------- host number theory
    int bit_width(u32 const* x, int words)
    {
        for (int i = words - 1; i >= 0; --i)
            if (x[i])
                return 32 * i + 32 - __builtin_clz(x[i]);
        return 0;
    }

    // x = 2 x mod q over L limbs (x < q)
    void double_mod(u32* x, u32 const* q, int L)
    {
        u32 d[max_limbs + 1], s[max_limbs + 1];
        u32 c = 0;
        for (int i = 0; i < L; ++i)
            d[i] = addc(x[i], x[i], c, c);
        d[L] = c;
        u32 br = 0;
        for (int i = 0; i < L; ++i)
            s[i] = subb(d[i], q[i], br, br);
        s[L] = subb(d[L], 0u, br, br);
        for (int i = 0; i < L; ++i)
            x[i] = br ? d[i] : s[i];
    }

    template<int L>
    void pow_mod(u32 (&r)[L], u32 const (&base)[L], u32 const* e, int e_words, u32 const (&one)[L], mod_params const& P)
    {
        u32 acc[L], b[L];
        copy<L>(acc, one);
        copy<L>(b, base);
        int bits = bit_width(e, e_words);
        for (int i = 0; i < bits; ++i)
        {
            if ((e[i / 32] >> (i % 32)) & 1u)
                mod_mul<L>(acc, acc, b, P);
            mod_mul<L>(b, b, b, P);
        }
        copy<L>(r, acc);
    }

    template<int L>
    bool equal(u32 const (&a)[L], u32 const (&b)[L])
    {
        for (int i = 0; i < L; ++i)
            if (a[i] != b[i])
                return false;
        return true;
    }
}

namespace
{
    // HIP events around each fast-path launch (launch_probe), aggregated per kernel name
    struct event_probe final : launch_probe
    {
        struct record
        {
            char const* kernel;
            hipEvent_t start, stop;
        };
        std::vector<record> open, done;
        std::vector<hipEvent_t> pool;
        std::vector<std::pair<std::string, std::pair<std::size_t, double>>> stats;

        hipEvent_t take()
        {
            if (!pool.empty())
            {
                hipEvent_t e = pool.back();
                pool.pop_back();
                return e;
            }
            hipEvent_t e = nullptr;
            return hipEventCreate(&e) == hipSuccess ? e : nullptr;
        }

        void begin(hipStream_t s, char const* kernel) override
        {
            hipEvent_t e = take();
            if (e && hipEventRecord(e, s) == hipSuccess)
                open.push_back({kernel, e, nullptr});
            else
                open.push_back({kernel, nullptr, nullptr});
        }

        void end(hipStream_t s, char const* kernel) override
        {
            if (open.empty())
                return;
            record r = open.back();
            open.pop_back();
            hipEvent_t e = take();
            if (r.start && e && hipEventRecord(e, s) == hipSuccess)
            {
                r.stop = e;
                done.push_back(r);
            }
            (void)kernel;
        }

        // wait for the recorded launches and fold them into the per-kernel statistics
        void collect()
        {
            for (auto& r : done)
            {
                float ms = 0.0f;
                if (hipEventSynchronize(r.stop) == hipSuccess && hipEventElapsedTime(&ms, r.start, r.stop) == hipSuccess)
                {
                    auto it = std::find_if(stats.begin(), stats.end(), [&](auto const& x) { return x.first == r.kernel; });
                    if (it == stats.end())
                    {
                        stats.push_back({r.kernel, {0, 0.0}});
                        it = stats.end() - 1;
                    }
                    it->second.first += 1;
                    it->second.second += ms;
                }
                pool.push_back(r.start);
                pool.push_back(r.stop);
            }
            done.clear();
        }

        void reset()
        {
            collect();
            stats.clear();
        }

        ~event_probe() override
        {
            for (auto& r : done)
            {
                (void)hipEventDestroy(r.start);
                (void)hipEventDestroy(r.stop);
            }
            for (hipEvent_t e : pool)
                (void)hipEventDestroy(e);
        }
    };
}

// ------------------------------------------------------------------ allocation-free workspace
// Workspace of hmpc_polymul_noalloc: a file-scope device array of a fixed size. Every GPU the library
// runs on holds its own copy from module load until the process ends. In host code `scratch` is only
// a stand-in for the symbol: its device address on each GPU comes from hipGetSymbolAddress, taken once
// per device when the first context on it is created (with that device current), and only that
// address is passed to kernels. Calls that overlap on one GPU share (and race on) the array.
namespace
{
    constexpr std::size_t noalloc_scratch_bytes = std::size_t{64} << 20;
    constexpr int max_devices = 64;
    std::mutex noalloc_mutex;
    void* noalloc_address[max_devices] = {};

    // the array's address on `device` (current), resolved once
    int noalloc_resolve(int device);
}
static __device__ __align__(256) u32 scratch[noalloc_scratch_bytes / 4];

namespace
{
    int noalloc_resolve(int device)
    {
        if (device < 0 || device >= max_devices)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "device ordinal beyond the allocation-free workspace table");
        std::lock_guard<std::mutex> lock(noalloc_mutex);
        if (noalloc_address[device])
            return HMPC_OK;
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
        noalloc_address[device] = ptr;
        return HMPC_OK;
    }
}

struct hmpc_context_s
{
    int device = 0;
    bool timing = false;
    event_probe probe;
    u32 key[8] = {};
    u32 nonce[3] = {};
    hipDeviceProp_t props{};
    std::mutex ws_mutex;
    std::map<hipStream_t, std::pair<void*, std::size_t>> workspaces;

    // scratch of at least `bytes` for work ordered on `stream` (grown on demand)
    int workspace(hipStream_t stream, std::size_t bytes, void** out)
    {
        std::lock_guard<std::mutex> lock(ws_mutex);
        auto& w = workspaces[stream];
        if (w.second < bytes)
        {
            if (w.first)
            {
                HMPC_HIP(hipStreamSynchronize(stream));
                HMPC_HIP(hipFree(w.first));
                w.first = nullptr;
                w.second = 0;
            }
            HMPC_HIP(hipMalloc(&w.first, bytes));
            w.second = bytes;
        }
        *out = w.first;
        return HMPC_OK;
    }

    ~hmpc_context_s()
    {
        for (auto& [s, w] : workspaces)
            if (w.first)
                (void)hipFree(w.first);
    }
};

struct table_pair
{
    u32* fwd = nullptr;
    u32* inv = nullptr;
    u32* fwd_em = nullptr;
    u32* inv_em = nullptr;
    u32* fwd_fast = nullptr;  // r26 tables (owned) or aliases of the em tables
    u32* inv_fast = nullptr;
    u32* fwd_tw = nullptr;    // r26: expanded twiddles of the first entries (owned with the r26 tables)
    u32* inv_tw = nullptr;
    bool fast_owned = false;
    u32 root[max_limbs] = {};
    u32 inverse_root[max_limbs] = {};
};

struct hmpc_modulus_s
{
    hmpc_context ctx = nullptr;
    int device = 0;  // the context's device, kept for destruction in any order with the context
    int limbs = 0;
    int bit_size = 0;
    mod_params P{};
    u32 one[max_limbs] = {};
    u32 r2[max_limbs] = {};
    u32 q2[max_limbs] = {};
    int form = form_generic;  // arithmetic form of the fast path
    int variant = 0;          // 0: fast path where supported, 1: portable LDS radix-2 kernels
    p26 r26{};                // constants of the radix-2^26 form
    std::map<int, table_pair> tables;
    std::map<int, int> elems; // fast path: elements per thread chosen per log_n (tuner)
    std::map<int, int> passes; // fast path: 1 = one launch per pass, 0 = the persistent job kernel (absent: default)
    std::map<int, int> bgv_elems; // fast BGV: elements per thread per log_n (tuner)
    std::map<int, int> bgv_passes; // fast BGV, N > 4096: 1 = two launches over tiles, 0 = persistent jobs (tuner)
    int occupancy[17] = {};   // persistent job kernel: workgroups per CU by elements per thread

    // per-pass kernels (transforms, products): 8 elements per thread at 4 limbs, 16 at 2; the
    // persistent job kernel: 4 at 4 limbs (it holds one workgroup per CU). Measured on MI355X over
    // N = 2^12..2^17 with 2^25 coefficients per call: products at 4 limbs take 2.42-2.95 ms at 8
    // against 2.46-3.34 ms at 4 per pass, and 2.65-3.71 ms on the persistent kernel.
    int default_elems(int) const { return limbs >= 4 ? 8 : 16; }
    int jobs_default_elems() const { return limbs == 4 ? 4 : 16; }
    // elements per thread for one call of n transforms: the tuned schedule, else the default,
    // except that a Goldilocks batch too small to give every CU a tile runs more threads per tile
    // (whole tiles: 4 elements, 1024 threads; two passes: 8). One N = 4096 NTT -> INTT took 41.1 us
    // at 16 and 31.5 us at 8 elements per thread.
    int call_elems(int log_n, std::size_t n) const
    {
        auto e = elems.find(log_n);
        if (e != elems.end())
            return e->second;
        int const lt = 12;  // log2 of the fast tile
        std::size_t tiles = log_n <= lt ? (n + (std::size_t{1} << (lt - log_n)) - 1) >> (lt - log_n) : n << (log_n - lt);
        if (limbs == 2 && tiles < static_cast<std::size_t>(ctx->props.multiProcessorCount))
            return log_n <= lt ? 4 : 8;
        return default_elems(log_n);
    }
    // tuned choice, else the tile launches (at 1024 ciphertexts of N = 2^14 they take 1.78 ms
    // against 2.3 ms for the persistent kernel, and they are the tuner's choice at every batch size
    // measured)
    bool bgv_tiles(int log_n, std::size_t n) const
    {
        (void)n;
        if (log_n <= 12)
            return false;
        auto p = bgv_passes.find(log_n);
        if (p != bgv_passes.end())
            return p->second != 0;
        return true;
    }
    // tuned choice, else 8 elements per thread for the tile launches (N > 4096: the tuner's choice
    // there, and what the one-wave row kernels of the radix-2^26 moduli run at) and 4 (4 limbs) or 8
    // for the persistent kernel (N <= 4096)
    int bgv_elems_for(int log_n) const
    {
        auto e = bgv_elems.find(log_n);
        if (e != bgv_elems.end())
            return e->second;
        return log_n > 12 || limbs != 4 ? 8 : 4;
    }
    // tile launches over job tables (batches, serving): every two-pass BGV job unless the tuner
    // chose the persistent kernel for its degree; 8 elements per thread unless tuned
    bool table_tiles(int log_n) const
    {
        if (log_n <= 12)
            return false;
        auto p = bgv_passes.find(log_n);
        return p == bgv_passes.end() || p->second != 0;
    }
    int table_elems_for(int log_n) const
    {
        auto p = bgv_passes.find(log_n);
        auto e = bgv_elems.find(log_n);
        return p != bgv_passes.end() && p->second != 0 && e != bgv_elems.end() ? e->second : 8;
    }
    // the same for products: the three tile launches unless the tuner chose the persistent kernel
    bool table_polymul(int log_n) const
    {
        if (log_n <= 12)
            return false;
        auto p = passes.find(log_n);
        return p == passes.end() || p->second != 0;
    }
    int polymul_elems_for(int log_n) const
    {
        auto e = elems.find(log_n);
        if (e != elems.end())
            return e->second;
        return table_polymul(log_n) ? default_elems(log_n) : jobs_default_elems();
    }
    // tile launches over a job table for this job, and with which elements per thread
    bool table_job(job_desc const& J, int* e) const
    {
        int log_n = static_cast<int>(J.log_n);
        bool bgv = J.type == job_bgv;
        *e = bgv ? table_elems_for(log_n) : polymul_elems_for(log_n);
        return bgv ? table_tiles(log_n) : table_polymul(log_n);
    }
    // uniform products: one launch per pass unless the persistent job kernel is chosen
    bool persistent(int log_n) const
    {
        auto p = passes.find(log_n);
        return p != passes.end() && p->second == 0;
    }

    ~hmpc_modulus_s()
    {
        for (auto& [k, t] : tables)
        {
            for (u32* p : {t.fwd, t.inv, t.fwd_em, t.inv_em})
                if (p)
                    (void)hipFree(p);
            if (t.fast_owned)
                for (u32* p : {t.fwd_fast, t.inv_fast, t.fwd_tw, t.inv_tw})
                    if (p)
                        (void)hipFree(p);
        }
    }

    // fast(): the special-form moduli, served by every fast kernel (job kernels, BGV, serving);
    // fast_transform(): transforms and products on the fast tile kernels, which also take any
    // 4- or 5-limb modulus with 4q < R
    bool fast(int log_n) const { return variant == 0 && fast_path_supported(limbs, form, log_n); }
    bool fast_transform(int log_n) const { return variant == 0 && fast_transform_supported(limbs, form, log_n); }

    int prepare(int log_n);

    int args(int log_n, transform_args& T)
    {
        int rc = prepare(log_n);
        if (rc)
            return rc;
        auto& t = tables[log_n];
        T.P = P;
        std::copy(one, one + max_limbs, T.one);
        std::copy(r2, r2 + max_limbs, T.r2);
        T.log_n = log_n;
        T.fwd = t.fwd;
        T.inv = t.inv;
        T.fwd_em = t.fwd_em;
        T.inv_em = t.inv_em;
        T.fwd_fast = t.fwd_fast;
        T.inv_fast = t.inv_fast;
        T.fwd_tw = t.fwd_tw;
        T.inv_tw = t.inv_tw;
        T.r26 = r26;
        std::copy(q2, q2 + max_limbs, T.q2);
        T.form = form;
        auto e = elems.find(log_n);
        T.elems = e != elems.end() ? e->second : default_elems(log_n);
        return HMPC_OK;
    }
};

namespace
{
    hipError_t dispatch_ntt(hmpc_modulus_s const& m, transform_args const& T, std::size_t n, bool inverse, u32 const* in,
                            u32* out, hipStream_t s)
    {
        if (m.fast_transform(T.log_n))
            return launch_fast_ntt(m.limbs, T, n, inverse, in, out, s);
        return launch_ntt_fused(m.limbs, T, n, inverse, in, out, s);
    }

    // hmpc/ints/num/theory/root_of_unity.hpp:19-54 with the all-zero ChaCha20 stream
    // (compiletime_number_generator, hmpc/random/number_generator.hpp:106-110).
    template<int L>
    bool root_search(hmpc_modulus_s const& m, int log_degree, u32 (&out)[L])
    {
        mod_params const& P = m.P;
        u32 qm1[L];
        u32 br = 0;
        for (int i = 0; i < L; ++i)
            qm1[i] = subb(P.q[i], i == 0 ? 1u : 0u, br, br);
        for (int b = 0; b < log_degree; ++b)
            if ((qm1[b / 32] >> (b % 32)) & 1u)
                return false;
        u32 exponent[L] = {};
        for (int b = 0; b + log_degree < 32 * L; ++b)
        {
            int src = b + log_degree;
            if ((qm1[src / 32] >> (src % 32)) & 1u)
                exponent[b / 32] |= 1u << (b % 32);
        }
        u32 one[L];
        for (int i = 0; i < L; ++i)
            one[i] = m.one[i];
        u32 half_degree[1] = {1u << (log_degree - 1)};
        u32 degree[2] = {log_degree < 32 ? (1u << log_degree) : 0u, log_degree < 32 ? 0u : 1u};
        chacha_key zero{};
        u32 block[16];
        u32 counter = 0;
        int index = 16;
        auto next = [&]() {
            if (index == 16)
            {
                chacha20_block(block, zero, counter++, 0u, 0u, 0u);
                index = 0;
            }
            return block[index++];
        };
        for (int attempt = 0; attempt < 4096; ++attempt)
        {
            u32 x[L];
            for (int i = 0; i < L; ++i)
                x[i] = next();
            if (m.bit_size % 32)
                x[L - 1] &= (1u << (m.bit_size % 32)) - 1u;
            // x > q - 1  ->  x -= q
            bool greater = false;
            for (int i = L - 1; i >= 0; --i)
                if (x[i] != qm1[i])
                {
                    greater = x[i] > qm1[i];
                    break;
                }
            if (greater)
            {
                u32 b2 = 0;
                for (int i = 0; i < L; ++i)
                    x[i] = subb(x[i], P.q[i], b2, b2);
            }
            bool zero_x = true;
            for (int i = 0; i < L; ++i)
                zero_x &= x[i] == 0;
            if (zero_x)
                continue;
            u32 g[L], t1[L], t2[L];
            pow_mod<L>(g, x, exponent, L, one, P);
            pow_mod<L>(t1, g, half_degree, 1, one, P);
            pow_mod<L>(t2, g, degree, 2, one, P);
            if (!equal<L>(t1, one) && equal<L>(t2, one))
            {
                copy<L>(out, g);
                return true;
            }
        }
        return false;
    }

    template<int L>
    void invert_mont(u32 (&r)[L], u32 const (&x)[L], hmpc_modulus_s const& m)
    {
        u32 e[L];
        u32 br = 0;
        for (int i = 0; i < L; ++i)
            e[i] = subb(m.P.q[i], i == 0 ? 2u : 0u, br, br);
        u32 one[L];
        for (int i = 0; i < L; ++i)
            one[i] = m.one[i];
        pow_mod<L>(r, x, e, L, one, m.P);
    }

    // makes `device` current; a failure is consumed here, so that it is not reported later as the
    // error of an unrelated launch (hipGetLastError after a launch)
    struct device_guard
    {
        explicit device_guard(int device)
        {
            if (hipSetDevice(device) != hipSuccess)
                (void)hipGetLastError();
        }
    };

    // installs the context's timing probe for the duration of one entry point
    struct probe_guard
    {
        launch_probe* previous;
        explicit probe_guard(hmpc_context ctx) : previous(current_probe())
        {
            current_probe() = ctx && ctx->timing ? &ctx->probe : nullptr;
        }
        ~probe_guard() { current_probe() = previous; }
    };

    hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

    bool supported_log_n(int log_n) { return log_n >= 2 && log_n <= 20; }
}

int hmpc_modulus_s::prepare(int log_n)
{
    if (!supported_log_n(log_n))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "log_n must be in [2, 20]");
    if (tables.count(log_n))
        return HMPC_OK;
    device_guard g(ctx->device);
    table_pair t;
    auto release = [&]() {
        for (u32* p : {t.fwd, t.inv, t.fwd_em, t.inv_em})
            if (p)
                (void)hipFree(p);
        if (t.fast_owned)
            for (u32* p : {t.fwd_fast, t.inv_fast, t.fwd_tw, t.inv_tw})
                if (p)
                    (void)hipFree(p);
    };
    // the root search fails for a modulus without a 2N-th root; any later failure is the device's
    char const* where = nullptr;
    hipError_t err = hipSuccess;
    bool root = host_limbs(limbs, [&](auto LC) {
        constexpr int LL = decltype(LC)::value;
        u32 phi[LL], iphi[LL], nval[LL] = {}, nm[LL], ninv[LL], r2v[LL];
        if (!root_search<LL>(*this, log_n + 1, phi))
            return false;
        invert_mont<LL>(iphi, phi, *this);
        nval[0] = 1u << log_n;
        for (int i = 0; i < LL; ++i)
            r2v[i] = r2[i];
        mod_mul<LL>(nm, nval, r2v, P);
        invert_mont<LL>(ninv, nm, *this);
        std::copy(phi, phi + LL, t.root);
        std::copy(iphi, iphi + LL, t.inverse_root);
        auto step = [&](hipError_t e, char const* w) {
            if (err == hipSuccess && e != hipSuccess)
            {
                err = e;
                where = w;
            }
            return err == hipSuccess;
        };
        std::size_t bytes = (std::size_t{4} * LL) << log_n;
        for (u32** p : {&t.fwd, &t.inv, &t.fwd_em, &t.inv_em})
            if (!step(hipMalloc(p, bytes), "hipMalloc (twiddle tables)"))
                return true;
        if (!step(launch_make_twiddles(LL, t.fwd, log_n, phi, ninv, one, false, P, nullptr), "twiddle kernel") ||
            !step(launch_make_twiddles(LL, t.inv, log_n, iphi, ninv, one, true, P, nullptr), "twiddle kernel") ||
            !step(launch_to_element_major(LL, t.fwd, t.fwd_em, std::size_t{1} << log_n, nullptr), "table relayout") ||
            !step(launch_to_element_major(LL, t.inv, t.inv_em, std::size_t{1} << log_n, nullptr), "table relayout"))
            return true;
        if (form == form_r26)
        {
            std::size_t fb = (std::size_t{4} * 8) << log_n;
            if (!step(hipMalloc(&t.fwd_fast, fb), "hipMalloc (radix-2^26 tables)") ||
                !step(hipMalloc(&t.inv_fast, fb), "hipMalloc (radix-2^26 tables)"))
                return true;
            t.fast_owned = true;
            if (!step(launch_r26_table(t.fwd, t.fwd_fast, std::size_t{1} << log_n, r26, nullptr), "radix-2^26 table") ||
                !step(launch_r26_table(t.inv, t.inv_fast, std::size_t{1} << log_n, r26, nullptr), "radix-2^26 table"))
                return true;
            std::size_t const xn = std::min(r26_tw_entries, std::size_t{1} << log_n);
            std::size_t const xb = std::size_t{4} * r26_tw_stride * xn;
            if (!step(hipMalloc(&t.fwd_tw, xb), "hipMalloc (expanded twiddles)") ||
                !step(hipMalloc(&t.inv_tw, xb), "hipMalloc (expanded twiddles)") ||
                !step(launch_r26_tw_table(t.fwd_fast, t.fwd_tw, xn, r26, nullptr), "expanded twiddle table") ||
                !step(launch_r26_tw_table(t.inv_fast, t.inv_tw, xn, r26, nullptr), "expanded twiddle table"))
                return true;
        }
        else
        {
            t.fwd_fast = t.fwd_em;
            t.inv_fast = t.inv_em;
        }
        step(hipStreamSynchronize(nullptr), "hipStreamSynchronize after the table kernels");
        return true;
    });
    if (!root)
        return fail(HMPC_ERR_INVALID_MODULUS, "no primitive 2N-th root of unity for this modulus (2N must divide q - 1)");
    if (err != hipSuccess)
    {
        release();
        return hip_fail(err, where);
    }
    tables[log_n] = t;
    return HMPC_OK;
}

extern "C" {

const char* hmpc_last_error(void) { return last_error.c_str(); }

// ------------------------------------------------------------------ context
int hmpc_context_create(int device, const uint32_t* key, hmpc_context* out)
{
    if (!out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "out is null");
    int count = 0;
    HMPC_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "no such device");
    auto ctx = std::make_unique<hmpc_context_s>();
    ctx->device = device;
    HMPC_HIP(hipSetDevice(device));
    HMPC_HIP(hipGetDeviceProperties(&ctx->props, device));
    int rc = noalloc_resolve(device);
    if (rc)
        return rc;
    if (key)
        std::copy(key, key + 8, ctx->key);
    else
    {
        std::random_device rd;
        for (auto& k : ctx->key)
            k = rd();
    }
    *out = ctx.release();
    return HMPC_OK;
}

int hmpc_context_destroy(hmpc_context ctx)
{
    if (ctx)
    {
        device_guard g(ctx->device);
        (void)hipDeviceSynchronize();
        delete ctx;
    }
    return HMPC_OK;
}

int hmpc_context_get_rng(hmpc_context ctx, uint32_t* key8, uint32_t* nonce3)
{
    if (!ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null context");
    if (key8)
        std::copy(ctx->key, ctx->key + 8, key8);
    if (nonce3)
        std::copy(ctx->nonce, ctx->nonce + 3, nonce3);
    return HMPC_OK;
}

int hmpc_context_set_rng(hmpc_context ctx, const uint32_t* key8, const uint32_t* nonce3)
{
    if (!ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null context");
    if (key8)
        std::copy(key8, key8 + 8, ctx->key);
    if (nonce3)
        std::copy(nonce3, nonce3 + 3, ctx->nonce);
    return HMPC_OK;
}

int hmpc_context_wait(hmpc_context ctx)
{
    if (!ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null context");
    device_guard g(ctx->device);
    HMPC_HIP(hipDeviceSynchronize());
    return HMPC_OK;
}

int hmpc_context_info(hmpc_context ctx, hmpc_device_info* out)
{
    if (!ctx || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof(*out));
    std::snprintf(out->name, sizeof(out->name), "%s", ctx->props.name);
    std::snprintf(out->arch, sizeof(out->arch), "%s", ctx->props.gcnArchName);
    out->compute_units = ctx->props.multiProcessorCount;
    out->wavefront_size = ctx->props.warpSize;
    out->clock_khz = ctx->props.clockRate;
    out->memory_clock_khz = ctx->props.memoryClockRate;
    out->memory_bus_width = ctx->props.memoryBusWidth;
    out->global_memory = ctx->props.totalGlobalMem;
    out->lds_per_workgroup = ctx->props.sharedMemPerBlock;
    return HMPC_OK;
}

// ------------------------------------------------------------------ memory
int hmpc_malloc(hmpc_context ctx, size_t bytes, void** out)
{
    if (!ctx || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    device_guard g(ctx->device);
    HMPC_HIP(hipMalloc(out, std::max<size_t>(bytes, 4)));
    return HMPC_OK;
}

int hmpc_free(hmpc_context ctx, void* ptr)
{
    if (!ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null context");
    device_guard g(ctx->device);
    if (ptr)
        HMPC_HIP(hipFree(ptr));
    return HMPC_OK;
}

int hmpc_memcpy_h2d(hmpc_context ctx, void* dst, const void* src, size_t bytes)
{
    device_guard g(ctx->device);
    HMPC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return HMPC_OK;
}

int hmpc_memcpy_d2h(hmpc_context ctx, void* dst, const void* src, size_t bytes)
{
    device_guard g(ctx->device);
    HMPC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return HMPC_OK;
}

int hmpc_memcpy_d2d(hmpc_context ctx, void* dst, const void* src, size_t bytes, void* stream)
{
    device_guard g(ctx->device);
    HMPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_host_alloc(hmpc_context ctx, size_t bytes, void** out)
{
    if (!ctx || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    device_guard g(ctx->device);
    *out = nullptr;
    if (bytes == 0)
        return HMPC_OK;
    HMPC_HIP(hipHostMalloc(out, bytes, hipHostMallocDefault));
    return HMPC_OK;
}

int hmpc_host_free(hmpc_context ctx, void* ptr)
{
    if (!ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null context");
    device_guard g(ctx->device);
    if (ptr)
        HMPC_HIP(hipHostFree(ptr));
    return HMPC_OK;
}

int hmpc_memcpy_h2d_async(hmpc_context ctx, void* dst, const void* src, size_t bytes, void* stream)
{
    if (!ctx || (bytes && (!dst || !src)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    device_guard g(ctx->device);
    if (bytes)
        HMPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_memcpy_d2h_async(hmpc_context ctx, void* dst, const void* src, size_t bytes, void* stream)
{
    if (!ctx || (bytes && (!dst || !src)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    device_guard g(ctx->device);
    if (bytes)
        HMPC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_stream_create(hmpc_context ctx, void** stream)
{
    device_guard g(ctx->device);
    hipStream_t s;
    HMPC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return HMPC_OK;
}

int hmpc_stream_destroy(hmpc_context ctx, void* stream)
{
    device_guard g(ctx->device);
    if (stream)
    {
        HMPC_HIP(hipStreamSynchronize(as_stream(stream)));
        {
            std::lock_guard<std::mutex> lock(ctx->ws_mutex);
            auto it = ctx->workspaces.find(as_stream(stream));
            if (it != ctx->workspaces.end())
            {
                if (it->second.first)
                    HMPC_HIP(hipFree(it->second.first));
                ctx->workspaces.erase(it);
            }
        }
        HMPC_HIP(hipStreamDestroy(as_stream(stream)));
    }
    return HMPC_OK;
}

int hmpc_stream_synchronize(hmpc_context ctx, void* stream)
{
    device_guard g(ctx->device);
    HMPC_HIP(hipStreamSynchronize(as_stream(stream)));
    return HMPC_OK;
}

// ------------------------------------------------------------------ moduli
int hmpc_modulus_create(hmpc_context ctx, const uint32_t* q, int q_words, hmpc_modulus* out)
{
    if (!ctx || !q || !out || q_words <= 0 || q_words > 64)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad modulus arguments");
    int qbits = bit_width(q, q_words);
    if (qbits < 2 || (q[0] & 1u) == 0)
        return fail(HMPC_ERR_INVALID_MODULUS, "modulus must be odd and > 1");
    std::vector<u32> qm1(q, q + q_words);
    {
        u32 br = 1;
        for (auto& w : qm1)
            w = subb(w, 0u, br, br);
    }
    int bit_size = bit_width(qm1.data(), q_words);
    int L = (bit_size + 31) / 32;
    if (L > max_limbs || qbits > 32 * L)
        return fail(HMPC_ERR_INVALID_MODULUS, "modulus wider than 288 bits is not supported");
    auto m = std::make_unique<hmpc_modulus_s>();
    m->ctx = ctx;
    m->device = ctx->device;
    m->limbs = L;
    m->bit_size = bit_size;
    m->P.limbs = L;
    std::copy(q, q + L, m->P.q);
    u32 inv = 1;
    for (int i = 0; i < 6; ++i)
        inv *= 2u - q[0] * inv;
    m->P.nprime = 0u - inv;
    u32 x[max_limbs] = {1};
    for (int b = 0; b < 64 * L; ++b)
    {
        double_mod(x, m->P.q, L);
        if (b == 32 * L - 1)
            std::copy(x, x + L, m->one);
    }
    std::copy(x, x + L, m->r2);
    {
        u32 c = 0;
        for (int i = 0; i < L; ++i)
            m->q2[i] = addc(m->P.q[i], m->P.q[i], c, c);
    }
    bool lazy_ok = (m->P.q[L - 1] >> 30) == 0;  // 4q < R
    bool middle_zero = L >= 3;
    for (int i = 1; i + 1 < L; ++i)
        middle_zero &= m->P.q[i] == 0;
    if (L == 2 && m->P.q[0] == 1u && m->P.q[1] == 0xffffffffu)
        m->form = form_goldilocks;
    else if (L == 4 && m->P.q[3] == 0x10000000u && m->P.q[2] == 0 && m->P.q[1] == 0 && m->P.q[0] < (1u << 24))
    {
        // q = 2^124 + c: radix-2^26 constants (see r26.hpp for the bounds they satisfy)
        m->form = form_r26;
        p26& R = m->r26;
        r26_constants(R, m->P.q[0]);
        // R 2^130 mod q = 4 R^2 mod q, as normalized radix-2^26 limbs
        u32 rr[4] = {m->r2[0], m->r2[1], m->r2[2], m->r2[3]};
        double_mod(rr, m->P.q, 4);
        double_mod(rr, m->P.q, 4);
        r26_unpack(R.rr, rr[0], rr[1], rr[2], rr[3]);
    }
    else if (lazy_ok && middle_zero)
        m->form = form_sparse;
    else if (lazy_ok)
        m->form = form_lazy;
    else
        m->form = form_generic;
    *out = m.release();
    return HMPC_OK;
}

int hmpc_modulus_destroy(hmpc_modulus mod)
{
    if (mod)
    {
        device_guard g(mod->device);
        (void)hipDeviceSynchronize();
        delete mod;
    }
    return HMPC_OK;
}

int hmpc_modulus_info(hmpc_modulus mod, int* limbs, int* bit_size, uint32_t* nprime, uint32_t* r_mod_q, uint32_t* r2_mod_q)
{
    if (!mod)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null modulus");
    if (limbs)
        *limbs = mod->limbs;
    if (bit_size)
        *bit_size = mod->bit_size;
    if (nprime)
        *nprime = mod->P.nprime;
    if (r_mod_q)
        std::copy(mod->one, mod->one + mod->limbs, r_mod_q);
    if (r2_mod_q)
        std::copy(mod->r2, mod->r2 + mod->limbs, r2_mod_q);
    return HMPC_OK;
}

int hmpc_modulus_set_kernels(hmpc_modulus mod, int variant)
{
    if (!mod || variant < 0 || variant > 1)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "variant must be 0 (fast) or 1 (portable)");
    mod->variant = variant;
    return HMPC_OK;
}

int hmpc_modulus_set_schedule(hmpc_modulus mod, int log_n, int elems_per_thread, int persistent)
{
    if (!mod || log_n < 2 || log_n > 20)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "log_n must be in [2, 20]");
    bool ok = elems_per_thread == 0 ||
              ((mod->limbs == 4 || mod->limbs == 5) && (elems_per_thread == 4 || elems_per_thread == 8)) ||
              (mod->limbs == 2 && (elems_per_thread == 8 || elems_per_thread == 16));
    if (!ok)
        return fail(HMPC_ERR_INVALID_ARGUMENT,
                    "elements per thread: 4 or 8 (4 and 5 limbs), 8 or 16 (2 limbs), 0 = default");
    if (persistent < -2 || persistent > 1)
        return fail(HMPC_ERR_INVALID_ARGUMENT,
                    "persistent: 1 (job kernel), 0 (one launch per pass), -1 (unchanged), -2 (default)");
    if (elems_per_thread == 0)
        mod->elems.erase(log_n);
    else
        mod->elems[log_n] = elems_per_thread;
    if (persistent == -2)
        mod->passes.erase(log_n);
    else if (persistent >= 0)
        mod->passes[log_n] = persistent ? 0 : 1;
    return HMPC_OK;
}

int hmpc_modulus_set_bgv_schedule(hmpc_modulus mod, int log_n, int elems_per_thread, int tiles)
{
    if (!mod || log_n < 2 || log_n > 20)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "log_n must be in [2, 20]");
    bool ok = elems_per_thread == 0 ||
              ((mod->limbs == 4 || mod->limbs == 5) && (elems_per_thread == 4 || elems_per_thread == 8)) ||
              (mod->limbs == 2 && (elems_per_thread == 8 || elems_per_thread == 16));
    if (!ok)
        return fail(HMPC_ERR_INVALID_ARGUMENT,
                    "elements per thread: 4 or 8 (4 and 5 limbs), 8 or 16 (2 limbs), 0 = default");
    if (tiles < -1 || tiles > 1)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "tiles: 1 (two launches over tiles), 0 (persistent jobs), -1 (default)");
    if (elems_per_thread == 0)
        mod->bgv_elems.erase(log_n);
    else
        mod->bgv_elems[log_n] = elems_per_thread;
    if (tiles < 0)
        mod->bgv_passes.erase(log_n);
    else
        mod->bgv_passes[log_n] = tiles;
    return HMPC_OK;
}

int hmpc_modulus_get_schedule(hmpc_modulus mod, int log_n, int* elems_per_thread, int* persistent, int* fast)
{
    if (!mod || log_n < 2 || log_n > 20)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "log_n must be in [2, 20]");
    auto e = mod->elems.find(log_n);
    if (elems_per_thread)
        *elems_per_thread = e != mod->elems.end() ? e->second : mod->default_elems(log_n);
    if (persistent)
        *persistent = mod->persistent(log_n) ? 1 : 0;
    if (fast)
        *fast = mod->fast_transform(log_n) ? 1 : 0;
    return HMPC_OK;
}

int hmpc_modulus_prepare(hmpc_modulus mod, int log_n)
{
    if (!mod)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null modulus");
    return mod->prepare(log_n);
}

int hmpc_modulus_root(hmpc_modulus mod, int log_n, int inverse, uint32_t* out)
{
    if (!mod || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    int rc = mod->prepare(log_n);
    if (rc)
        return rc;
    auto& t = mod->tables[log_n];
    std::copy(inverse ? t.inverse_root : t.root, (inverse ? t.inverse_root : t.root) + mod->limbs, out);
    return HMPC_OK;
}

int hmpc_modulus_twiddles(hmpc_modulus mod, int log_n, int inverse, const uint32_t** out)
{
    if (!mod || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    int rc = mod->prepare(log_n);
    if (rc)
        return rc;
    auto& t = mod->tables[log_n];
    *out = inverse ? t.inv : t.fwd;
    return HMPC_OK;
}

// ------------------------------------------------------------------ NTT
int hmpc_ntt(hmpc_modulus mod, int log_n, size_t n, int inverse, const uint32_t* in, uint32_t* out, void* stream)
{
    if (!mod || (n && (!in || !out)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0)
        return HMPC_OK;
    if (n > 0x7fffffffu)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "batch too large");
    transform_args T;
    int rc = mod->args(log_n, T);
    if (rc)
        return rc;
    T.elems = mod->call_elems(log_n, n);
    device_guard g(mod->ctx->device);
    probe_guard pg(mod->ctx);
    HMPC_HIP(dispatch_ntt(*mod, T, n, inverse != 0, in, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_ntt_plain(hmpc_modulus mod, int log_n, size_t n, int inverse, const uint32_t* in, uint32_t* out, uint32_t* scratch,
                   void* stream)
{
    if (!mod || (n && (!in || !out)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0)
        return HMPC_OK;
    transform_args T;
    int rc = mod->args(log_n, T);
    if (rc)
        return rc;
    device_guard g(mod->ctx->device);
    hipStream_t s = as_stream(stream);
    u32* sc = scratch;
    if (!sc)
    {
        void* w = nullptr;
        rc = mod->ctx->workspace(s, (std::size_t{4} * mod->limbs * n) << log_n, &w);
        if (rc)
            return rc;
        sc = static_cast<u32*>(w);
    }
    HMPC_HIP(launch_ntt_plain(mod->limbs, T, n, inverse != 0, in, out, sc, s));
    return HMPC_OK;
}

// ------------------------------------------------------------------ point-wise
int hmpc_pointwise(hmpc_modulus mod, int op, size_t count, const uint32_t* a, const uint32_t* b, size_t b_count,
                   uint32_t* out, void* stream)
{
    if (!mod || op < 0 || op > 2)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad op");
    if (count == 0)
        return HMPC_OK;
    if (!a || !b || !out || b_count == 0 || b_count > count)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad operands");
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_pointwise(mod->limbs, mod->P, op, count, a, b, b_count, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_pointwise_scalar(hmpc_modulus mod, int op, size_t count, const uint32_t* a, const uint32_t* c_host, uint32_t* out,
                          void* stream)
{
    if (!mod || op < 0 || op > 2 || !c_host)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad op");
    if (count == 0)
        return HMPC_OK;
    if (!a || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad operands");
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_pointwise_scalar(mod->limbs, mod->P, op, count, a, c_host, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_ciphertext_op(hmpc_modulus mod, int op, size_t count, const uint32_t* a0, const uint32_t* a1, const uint32_t* b0,
                       const uint32_t* b1, size_t b_count, uint32_t* out0, uint32_t* out1, void* stream)
{
    if (!mod || op < 0 || op > 2)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad op");
    if (count == 0)
        return HMPC_OK;
    if (!a0 || !a1 || !b0 || !b1 || !out0 || !out1 || b_count == 0 || b_count > count)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad operands");
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_ct_binary(mod->limbs, mod->P, op, count, a0, a1, b0, b1, b_count, out0, out1, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_ciphertext_scalar(hmpc_modulus mod, int op, size_t count, const uint32_t* a0, const uint32_t* a1,
                           const uint32_t* c_host, uint32_t* out0, uint32_t* out1, void* stream)
{
    if (!mod || op < 0 || op > 2 || !c_host)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad op");
    if (count == 0)
        return HMPC_OK;
    if (!a0 || !a1 || !out0 || !out1)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad operands");
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_ct_scalar(mod->limbs, mod->P, op, count, a0, a1, c_host, out0, out1, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_fma(hmpc_modulus mod, size_t count, const uint32_t* a, const uint32_t* b, const uint32_t* c, uint32_t* out,
             void* stream)
{
    if (!mod)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null modulus");
    if (count == 0)
        return HMPC_OK;
    if (!a || !b || !c || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad operands");
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_fma(mod->limbs, mod->P, count, a, b, c, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_reduce(hmpc_modulus mod, int op, size_t count, const uint32_t* in, uint32_t* out, void* stream)
{
    if (!mod || (op != HMPC_ADD && op != HMPC_MUL))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "reduce: HMPC_ADD or HMPC_MUL");
    if (!out || (count && !in))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad operands");
    device_guard g(mod->ctx->device);
    hipStream_t s = as_stream(stream);
    void* w = nullptr;
    int rc = mod->ctx->workspace(s, static_cast<std::size_t>(reduce_parts(count)) * mod->limbs * 4, &w);
    if (rc)
        return rc;
    HMPC_HIP(launch_reduce(mod->limbs, mod->P, mod->one, op == HMPC_ADD ? 0 : 1, count, in, static_cast<u32*>(w), out, s));
    return HMPC_OK;
}

int hmpc_matvec(hmpc_modulus mod, size_t batch, size_t rows, size_t cols, const uint32_t* M, int broadcast_matrix,
                const uint32_t* v, int broadcast_vector, uint32_t* out, void* stream)
{
    if (!mod)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null modulus");
    if (batch == 0 || rows == 0)
        return HMPC_OK;
    if (!out || (cols && (!M || !v)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad operands");
    device_guard g(mod->ctx->device);
    std::size_t m_batches = broadcast_matrix ? 1 : batch;
    std::size_t v_batches = broadcast_vector ? 1 : batch;
    p26 const* r26 = mod->form == form_r26 && mod->variant == 0 ? &mod->r26 : nullptr;
    HMPC_HIP(launch_matvec(mod->limbs, mod->P, r26, batch, rows, cols, M, m_batches * rows * cols,
                           broadcast_matrix ? 0 : rows * cols, v, v_batches * cols, broadcast_vector ? 0 : cols, out,
                           as_stream(stream)));
    return HMPC_OK;
}

namespace
{
    int cipher_call(hmpc_context ctx, bool encrypt, int limbs, size_t count, const uint32_t* key8, const uint32_t* nonce3,
                    uint32_t counter0, const uint32_t* in, uint32_t* out, void* stream)
    {
        if (!ctx || !key8 || !nonce3 || limbs < 1 || limbs > 9)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "cipher: context, key, nonce and 1..9 limbs required");
        if (count == 0)
            return HMPC_OK;
        if (!in || !out || in == out)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "cipher: distinct input and output buffers required");
        device_guard g(ctx->device);
        HMPC_HIP(launch_cipher(limbs, encrypt, count, key8, nonce3, counter0, in, out, as_stream(stream)));
        return HMPC_OK;
    }
}

int hmpc_cipher_encrypt(hmpc_context ctx, int limbs, size_t count, const uint32_t* key8, const uint32_t* nonce3,
                        uint32_t counter0, const uint32_t* in, uint32_t* out, void* stream)
{
    return cipher_call(ctx, true, limbs, count, key8, nonce3, counter0, in, out, stream);
}

int hmpc_cipher_decrypt(hmpc_context ctx, int limbs, size_t count, const uint32_t* key8, const uint32_t* nonce3,
                        uint32_t counter0, const uint32_t* in, uint32_t* out, void* stream)
{
    return cipher_call(ctx, false, limbs, count, key8, nonce3, counter0, in, out, stream);
}

int hmpc_from_u32(hmpc_modulus mod, size_t count, const uint32_t* words, uint32_t* out, void* stream)
{
    if (!mod)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null modulus");
    if (count == 0)
        return HMPC_OK;
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_from_u32(mod->limbs, mod->P, mod->r2, count, words, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_to_unsigned(hmpc_modulus mod, size_t count, const uint32_t* in, uint32_t* out, void* stream)
{
    if (!mod)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null modulus");
    if (count == 0)
        return HMPC_OK;
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_to_unsigned(mod->limbs, mod->P, count, in, out, as_stream(stream)));
    return HMPC_OK;
}

// ------------------------------------------------------------------ sampling
namespace
{
    chacha_key make_key(const uint32_t* k)
    {
        chacha_key key{};
        std::copy(k, k + 8, key.k);
        return key;
    }

    void nonce_add(u32* nonce, u64 value)
    {
        u32 c = 0;
        nonce[0] = addc(nonce[0], static_cast<u32>(value), 0u, c);
        nonce[1] = addc(nonce[1], static_cast<u32>(value >> 32), c, c);
        nonce[2] = addc(nonce[2], 0u, c, c);
    }

    bool valid_bits(int bits) { return bits >= 2 && bits <= 64 && bits % 2 == 0; }
    // capability-stream draws also take width 0, the uniform ternary draw
    bool valid_cap_bits(int bits) { return bits == ternary_bits || valid_bits(bits); }
    constexpr char const* cap_bits_message = "bits = 4 * variance must be even and in [2, 64], or 0 (uniform ternary)";
}

int hmpc_sample_uniform(hmpc_modulus mod, const uint32_t* key8, uint64_t offset, size_t count, uint32_t* out, void* stream)
{
    if (!mod || !key8 || (count && !out))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (count == 0)
        return HMPC_OK;
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_sample_uniform(mod->limbs, mod->P, mod->bit_size, make_key(key8), offset, count, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_sample_cbd(hmpc_modulus mod, const uint32_t* key8, uint64_t offset, int bits, size_t count, uint32_t* out,
                    void* stream)
{
    if (!mod || !key8 || (count && !out))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (!valid_bits(bits))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bits = 4 * variance must be even and in [2, 64]");
    if (count == 0)
        return HMPC_OK;
    device_guard g(mod->ctx->device);
    HMPC_HIP(launch_sample_cbd(mod->limbs, mod->P, mod->r2, make_key(key8), offset, bits, count, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_sample_words(hmpc_context ctx, const uint32_t* key8, uint64_t offset, size_t count, uint32_t* out, void* stream)
{
    if (!ctx || !key8 || (count && !out))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (count == 0)
        return HMPC_OK;
    device_guard g(ctx->device);
    HMPC_HIP(launch_sample_words(make_key(key8), offset, count, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_sample_cbd_ntt(hmpc_context ctx, hmpc_modulus mod, int bits, int log_n, size_t n, uint32_t* out, void* stream)
{
    if (!ctx || !mod || (n && !out))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (!valid_cap_bits(bits))
        return fail(HMPC_ERR_INVALID_ARGUMENT, cap_bits_message);
    transform_args T;
    int rc = mod->args(log_n, T);
    if (rc)
        return rc;
    if (n == 0)
        return HMPC_OK;
    device_guard g(ctx->device);
    HMPC_HIP(launch_sample_cbd_ntt(mod->limbs, T, make_key(ctx->key), ctx->nonce, bits, n, out, as_stream(stream)));
    nonce_add(ctx->nonce, static_cast<u64>(n) << log_n);
    return HMPC_OK;
}

// ------------------------------------------------------------------ persistent job kernel
namespace
{
    // workgroups of the persistent kernel the device holds at once, capped by the job count
    int jobs_grid(hmpc_modulus_s& m, int elems, std::size_t njobs, unsigned* grid)
    {
        int& occ = m.occupancy[elems];
        if (occ == 0)
        {
            device_guard g(m.ctx->device);
            hipError_t e = fast_jobs_occupancy(m.limbs, elems, &occ);
            if (e != hipSuccess)
                return hip_fail(e, "fast_jobs_occupancy");
            occ = std::max(occ, 1);
        }
        std::size_t g = static_cast<std::size_t>(occ) * static_cast<std::size_t>(m.ctx->props.multiProcessorCount);
        *grid = static_cast<unsigned>(std::max<std::size_t>(1, std::min(g, njobs)));
        return HMPC_OK;
    }

    constexpr std::size_t counter_bytes = 256;

    // stream workspace laid out as [job counter, 256 B][grid scratch slots]
    int jobs_workspace(hmpc_context ctx, hipStream_t s, unsigned grid, u64 slot_words, u32** counter, u32** scratch)
    {
        void* w = nullptr;
        int rc = ctx->workspace(s, counter_bytes + static_cast<std::size_t>(grid) * slot_words * 4, &w);
        if (rc)
            return rc;
        *counter = static_cast<u32*>(w);
        *scratch = static_cast<u32*>(w) + counter_bytes / 4;
        return HMPC_OK;
    }

    // plaintext modulus given in Montgomery form -> integer, when it is below 2^40 (the fused
    // samplers form p v + m as a signed 64-bit integer before converting it)
    bool small_plaintext(hmpc_modulus_s const& m, u32 const* p_mont, u64* out)
    {
        return host_limbs(m.limbs, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            u32 a[LL], b[LL] = {}, r[LL];
            for (int i = 0; i < LL; ++i)
                a[i] = p_mont[i];
            b[0] = 1;
            mod_mul<LL>(r, a, b, m.P);
            u64 v = r[0];
            if constexpr (LL >= 2)
                v |= static_cast<u64>(r[1]) << 32;
            for (int i = 2; i < LL; ++i)
                if (r[i])
                    return false;
            *out = v;
            return v < (u64{1} << 40);
        });
    }

    job_desc polymul_job(transform_args const& T, int log_n, u32 const* a, u32 const* b, u32* out, u64 count, u64 off)
    {
        job_desc J{};
        J.type = job_polymul;
        J.log_n = static_cast<u32>(log_n);
        J.twf = T.fwd_fast;
        J.twi = T.inv_fast;
        J.twfx = T.fwd_tw;
        J.twix = T.inv_tw;
        J.in0 = a;
        J.in1 = b;
        J.in_count = count;
        J.in_off = off;
        J.out0 = out;
        J.out_count = count;
        J.out_off = off;
        return J;
    }

    job_desc bgv_job(transform_args const& T, bgv_args const& B, u64 p_int, int log_n, u32 const* pk_a, u32 const* pk_b,
                     u32 const* messages, u32* c0, u32* c1, u64 count, u64 off, u32 poly)
    {
        job_desc J{};
        J.type = job_bgv;
        J.log_n = static_cast<u32>(log_n);
        J.poly = poly;
        J.bits[0] = B.bits_u;
        J.bits[1] = B.bits_v;
        J.bits[2] = B.bits_w;
        J.key = B.key;
        std::copy(B.nonce_u, B.nonce_u + 3, J.nonce[0]);
        std::copy(B.nonce_v, B.nonce_v + 3, J.nonce[1]);
        std::copy(B.nonce_w, B.nonce_w + 3, J.nonce[2]);
        J.p_int = p_int;
        J.twf = T.fwd_fast;
        J.twi = T.inv_fast;
        J.twfx = T.fwd_tw;
        J.twix = T.inv_tw;
        J.in0 = pk_a;
        J.in1 = pk_b;
        J.in_count = u64{1} << log_n;
        J.in_off = 0;
        J.out0 = c0;
        J.out1 = c1;
        J.out_count = count;
        J.out_off = off;
        J.msg = messages;
        J.msg_off = off;
        return J;
    }

    // uniform batch of n jobs (job j at element offset j N) on the persistent kernel
    int run_uniform_jobs(hmpc_modulus_s& m, transform_args const& T, job_desc const& base, std::size_t n, hipStream_t s)
    {
        unsigned grid = 0;
        int rc = jobs_grid(m, T.elems, n, &grid);
        if (rc)
            return rc;
        u64 slot = job_scratch_words(base.type, T.log_n, m.limbs);
        u32 *counter = nullptr, *scratch = nullptr;
        rc = jobs_workspace(m.ctx, s, grid, slot, &counter, &scratch);
        if (rc)
            return rc;
        HMPC_HIP(launch_fast_jobs(m.limbs, T, nullptr, base, static_cast<u32>(n), grid, counter, scratch, slot, s));
        return HMPC_OK;
    }
}

namespace
{
    // Tile launches over tables of jobs of two-pass degrees (BGV: one ciphertext each, two
    // launches; products: one pair each, three launches), grouped by (modulus, elements per
    // thread, kind) and cut into chunks whose workspaces stay within tile_ws_cap_words; the chunks
    // of one stream run in order and share one workspace.
    constexpr u64 tile_ws_cap_words = u64{1} << 27;  // 512 MiB

    // a job of a tile launch with the arithmetic parameters of its modulus (the job itself carries its
    // degree's twiddle tables), captured by the caller
    struct tile_input
    {
        hmpc_modulus mod;
        int elems;
        job_desc job;
        transform_args T;
    };

    struct tile_chunk
    {
        hmpc_modulus mod = nullptr;
        transform_args T{};
        int elems = 0;
        u32 type = 0;
        int log_n = 0;  // a degree of the chunk (the arithmetic parameters do not depend on it)
        bool uniform = false;  // every job of the chunk has degree 2^log_n
        int max_log_n = 0;     // the largest degree of the chunk
        std::size_t job_off = 0, njobs = 0;
        // into the map: forward column pass, row pass, inverse column pass (products)
        std::size_t cols_off = 0, cols_blocks = 0, rows_off = 0, rows_blocks = 0, icols_off = 0, icols_blocks = 0;
        u64 ws_words = 0;
        u64 ws_base = 0;  // the chunk's workspace offset (words): 0 with a shared workspace
        u64 cost = 0;     // sum of N log N over its jobs (scheduling weight)
    };

    struct tile_plan
    {
        std::vector<job_desc> jobs;
        std::vector<u32> map;  // per block: job index << 10 | local tile
        std::vector<tile_chunk> chunks;
        u64 ws_words = 0;      // the largest chunk's workspace

        bool empty() const { return jobs.empty(); }

        // in: (modulus, elements per thread, job); order is kept within a group. by_degree: a group
        // (and so each chunk) holds one degree, and its launches run the kernels with the pass
        // lengths fixed at compile time. That pays where launches on several streams fill each
        // other's tails (the server's slots: peak 7,300 -> 8,200 requests/s); a batch keeps
        // degrees mixed, so that each pass is one launch (its uneven batch: 3.6 against 3.8 ms
        // split by degree). separate: every chunk gets a workspace of its own (ws_words is then the
        // sum, not the largest), so that chunks may run concurrently.
        void build(std::vector<tile_input> const& in, bool by_degree, bool separate = false)
        {
            jobs.clear();
            map.clear();
            chunks.clear();
            ws_words = 0;
            std::vector<std::tuple<hmpc_modulus, int, u32, u32>> groups;
            auto key = [by_degree](tile_input const& t) {
                return std::make_tuple(t.mod, t.elems, t.job.type, by_degree ? t.job.log_n : 0u);
            };
            for (auto const& t : in)
            {
                auto g = key(t);
                if (std::find(groups.begin(), groups.end(), g) == groups.end())
                    groups.push_back(g);
            }
            for (auto const& g : groups)
            {
                hmpc_modulus mod = std::get<0>(g);
                u32 type = std::get<2>(g);
                u64 streams = type == job_bgv ? 3 : 2;
                tile_chunk c;
                auto close = [&]() {
                    if (c.njobs)
                    {
                        if (separate)
                        {
                            c.ws_base = ws_words;
                            ws_words += c.ws_words;
                        }
                        else
                            ws_words = std::max(ws_words, c.ws_words);
                        chunks.push_back(c);
                    }
                    c = tile_chunk{};
                };
                for (auto const& t : in)
                {
                    if (key(t) != g)
                        continue;
                    job_desc J = t.job;
                    u64 words = streams * static_cast<u64>(ws_limbs(mod->limbs)) << J.log_n;
                    if (c.njobs && (c.ws_words + words > tile_ws_cap_words || c.njobs == (1u << 22)))
                        close();
                    if (!c.njobs)
                    {
                        c.mod = mod;
                        c.T = t.T;
                        c.elems = std::get<1>(g);
                        c.type = type;
                        c.log_n = static_cast<int>(J.log_n);
                        c.uniform = by_degree;
                        c.job_off = jobs.size();
                    }
                    J.ws_off = c.ws_words;
                    c.ws_words += words;
                    c.max_log_n = std::max(c.max_log_n, static_cast<int>(J.log_n));
                    c.cost += static_cast<u64>(J.log_n) << J.log_n;
                    c.njobs += 1;
                    jobs.push_back(J);
                }
                close();
            }
            auto section = [&](tile_chunk const& c, bool rows, u32 per_poly, std::size_t& off, std::size_t& blocks) {
                off = map.size();
                for (std::size_t j = 0; j < c.njobs; ++j)
                    for (u32 l = 0, e = per_poly * two_pass_tiles(static_cast<int>(jobs[c.job_off + j].log_n), rows); l < e; ++l)
                        map.push_back(static_cast<u32>(j) << 10 | l);
                blocks = map.size() - off;
            };
            for (auto& c : chunks)
            {
                bool bgv = c.type == job_bgv;
                section(c, false, bgv ? 3 : 2, c.cols_off, c.cols_blocks);
                section(c, true, 1, c.rows_off, c.rows_blocks);
                if (!bgv)
                    section(c, false, 1, c.icols_off, c.icols_blocks);
            }
        }

        // the launches of one chunk; device copies of `jobs` and `map` at d_jobs / d_map; ws holds
        // ws_words
        int launch_chunk(tile_chunk const& c, job_desc const* d_jobs, u32 const* d_map, u32* ws, hipStream_t s) const
        {
            transform_args T = c.T;
            T.elems = c.elems;
            auto n = [](std::size_t v) { return static_cast<u32>(v); };
            u32* w = ws + c.ws_base;
            if (c.type == job_bgv)
                HMPC_HIP(launch_fast_bgv_table(c.mod->limbs, T, c.uniform ? c.log_n : 0, c.max_log_n, d_jobs + c.job_off, d_map + c.cols_off,
                                               n(c.cols_blocks), d_map + c.rows_off, n(c.rows_blocks), w, s));
            else
                HMPC_HIP(launch_fast_polymul_table(c.mod->limbs, T, c.uniform ? c.log_n : 0, c.max_log_n, d_jobs + c.job_off, d_map + c.cols_off,
                                                   n(c.cols_blocks), d_map + c.rows_off, n(c.rows_blocks), d_map + c.icols_off,
                                                   n(c.icols_blocks), w, s));
            return HMPC_OK;
        }

        int launch(job_desc const* d_jobs, u32 const* d_map, u32* ws, hipStream_t s) const
        {
            for (auto const& c : chunks)
            {
                int rc = launch_chunk(c, d_jobs, d_map, ws, s);
                if (rc)
                    return rc;
            }
            return HMPC_OK;
        }
    };
}

// ------------------------------------------------------------------ fused batches
int hmpc_polymul(hmpc_modulus mod, int log_n, size_t n, const uint32_t* a, const uint32_t* b, uint32_t* out, void* stream)
{
    if (!mod || (n && (!a || !b || !out)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    transform_args T;
    int rc = mod->args(log_n, T);
    if (rc)
        return rc;
    if (n == 0)
        return HMPC_OK;
    if (n > 0x7fffffffu)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "batch too large");
    device_guard g(mod->ctx->device);
    probe_guard pg(mod->ctx);
    hipStream_t s = as_stream(stream);
    void* w = nullptr;
    bool fast = mod->fast_transform(log_n);
    if (mod->fast(log_n) && mod->persistent(log_n))
    {
        auto e = mod->elems.find(log_n);
        T.elems = e != mod->elems.end() ? e->second : mod->jobs_default_elems();
        return run_uniform_jobs(*mod, T, polymul_job(T, log_n, a, b, out, static_cast<u64>(n) << log_n, 0), n, s);
    }
    std::size_t words = fast ? fast_polymul_workspace_words(mod->limbs, log_n, n) : polymul_workspace_words(mod->limbs, log_n, n);
    if (words)
    {
        rc = mod->ctx->workspace(s, words * 4, &w);
        if (rc)
            return rc;
    }
    if (fast)
        HMPC_HIP(launch_fast_polymul(mod->limbs, T, n, a, b, out, static_cast<u32*>(w), s));
    else
        HMPC_HIP(launch_polymul(mod->limbs, T, n, a, b, out, static_cast<u32*>(w), s));
    return HMPC_OK;
}

// hmpc_polymul's per-pass kernels with their workspace in the file-scope array (no allocation: the
// tables of N must be prepared, and the workspace must fit the array). The workspace needs no
// clearing: every word the kernels read was written by the same call before.
int hmpc_polymul_noalloc(hmpc_modulus mod, int log_n, size_t n, const uint32_t* a, const uint32_t* b, uint32_t* out,
                         void* stream)
{
    if (!mod || (n && (!a || !b || !out)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (!supported_log_n(log_n))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "log_n must be in [2, 20]");
    if (!mod->tables.count(log_n))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "hmpc_polymul_noalloc: prepare the tables first (hmpc_modulus_prepare)");
    if (n == 0)
        return HMPC_OK;
    if (n > 0x7fffffffu)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "batch too large");
    bool fast = mod->fast_transform(log_n);
    std::size_t words = fast ? fast_polymul_workspace_words(mod->limbs, log_n, n) : polymul_workspace_words(mod->limbs, log_n, n);
    if (words * 4 > noalloc_scratch_bytes)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "hmpc_polymul_noalloc: the workspace of " + std::to_string(words * 4) +
                                                   " bytes exceeds the fixed " + std::to_string(noalloc_scratch_bytes) +
                                                   " bytes (fewer pairs per call)");
    transform_args T;
    int rc = mod->args(log_n, T);  // tables present: no allocation
    if (rc)
        return rc;
    device_guard g(mod->ctx->device);
    u32* w = static_cast<u32*>(noalloc_address[mod->ctx->device]);
    if (!w)
        return fail(HMPC_ERR_DEVICE, "allocation-free workspace not resolved for this device");
    hipStream_t s = as_stream(stream);
    if (fast)
    {
        auto e = mod->elems.find(log_n);
        T.elems = e != mod->elems.end() ? e->second : mod->default_elems(log_n);
        HMPC_HIP(launch_fast_polymul(mod->limbs, T, n, a, b, out, w, s));
    }
    else
        HMPC_HIP(launch_polymul(mod->limbs, T, n, a, b, out, w, s));
    return HMPC_OK;
}

namespace
{
    int bgv_check(hmpc_context ctx, hmpc_modulus mod, size_t n, const uint32_t* pk_a, const uint32_t* pk_b,
                  const uint32_t* messages, const uint32_t* p_mont, int bu, int bv, int bw, uint32_t* c0, uint32_t* c1)
    {
        if (!ctx || !mod || !p_mont || (n && (!pk_a || !pk_b || !messages || !c0 || !c1)))
            return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
        if (!valid_cap_bits(bu) || !valid_cap_bits(bv) || !valid_cap_bits(bw))
            return fail(HMPC_ERR_INVALID_ARGUMENT, cap_bits_message);
        if (n > 0x7fffffffu)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "batch too large");
        return HMPC_OK;
    }

    bgv_args make_bgv_args(hmpc_context ctx, hmpc_modulus mod, int log_n, size_t n, const uint32_t* p_mont, int bu, int bv,
                           int bw)
    {
        bgv_args B{};
        B.key = make_key(ctx->key);
        std::copy(ctx->nonce, ctx->nonce + 3, B.nonce_u);
        std::copy(ctx->nonce, ctx->nonce + 3, B.nonce_v);
        nonce_add(B.nonce_v, static_cast<u64>(n) << log_n);
        std::copy(B.nonce_v, B.nonce_v + 3, B.nonce_w);
        nonce_add(B.nonce_w, static_cast<u64>(n) << log_n);
        B.bits_u = bu;
        B.bits_v = bv;
        B.bits_w = bw;
        std::copy(p_mont, p_mont + mod->limbs, B.p_mont);
        return B;
    }
}

int hmpc_bgv_encrypt(hmpc_context ctx, hmpc_modulus mod, int log_n, size_t n, const uint32_t* pk_a, const uint32_t* pk_b,
                     const uint32_t* messages, const uint32_t* p_mont, int bits_u, int bits_v, int bits_w, uint32_t* c0,
                     uint32_t* c1, void* stream)
{
    int rc = bgv_check(ctx, mod, n, pk_a, pk_b, messages, p_mont, bits_u, bits_v, bits_w, c0, c1);
    if (rc)
        return rc;
    transform_args T;
    rc = mod->args(log_n, T);
    if (rc)
        return rc;
    if (n == 0)
        return HMPC_OK;
    device_guard g(ctx->device);
    probe_guard pg(ctx);
    hipStream_t s = as_stream(stream);
    bgv_args B = make_bgv_args(ctx, mod, log_n, n, p_mont, bits_u, bits_v, bits_w);
    u64 p_int = 0;
    if (mod->fast(log_n) && small_plaintext(*mod, p_mont, &p_int))
    {
        T.elems = mod->bgv_elems_for(log_n);
        job_desc J = bgv_job(T, B, p_int, log_n, pk_a, pk_b, messages, c0, c1, static_cast<u64>(n) << log_n, 0, 0);
        if (mod->bgv_tiles(log_n, n))
        {
            void* w = nullptr;
            rc = ctx->workspace(s, 3 * static_cast<std::size_t>(ws_limbs(mod->limbs)) * (n << log_n) * 4, &w);
            if (rc)
                return rc;
            HMPC_HIP(launch_fast_bgv_passes(mod->limbs, T, J, n, static_cast<u32*>(w), s));
        }
        else
            rc = run_uniform_jobs(*mod, T, J, n, s);
        if (rc)
            return rc;
        nonce_add(ctx->nonce, 3 * (static_cast<u64>(n) << log_n));
        return HMPC_OK;
    }
    if (mod->fast_transform(log_n) && log_n > 12)
    {
        // 4- or 5-limb moduli with 4q < R (the CIOS policy): the two tile launches with any
        // plaintext modulus, taken as the Montgomery constant p_mont
        auto e = mod->bgv_elems.find(log_n);
        T.elems = e != mod->bgv_elems.end() ? e->second : 8;
        job_desc J = bgv_job(T, B, 0, log_n, pk_a, pk_b, messages, c0, c1, static_cast<u64>(n) << log_n, 0, 0);
        std::copy(p_mont, p_mont + mod->limbs, J.p_mont);
        void* w = nullptr;
        rc = ctx->workspace(s, 3 * static_cast<std::size_t>(ws_limbs(mod->limbs)) * (n << log_n) * 4, &w);
        if (rc)
            return rc;
        HMPC_HIP(launch_fast_bgv_passes(mod->limbs, T, J, n, static_cast<u32*>(w), s));
        nonce_add(ctx->nonce, 3 * (static_cast<u64>(n) << log_n));
        return HMPC_OK;
    }
    void* w = nullptr;
    rc = ctx->workspace(s, bgv_workspace_words(mod->limbs, log_n, n) * 4, &w);
    if (rc)
        return rc;
    HMPC_HIP(launch_bgv_encrypt(mod->limbs, T, B, n, pk_a, pk_b, messages, c0, c1, static_cast<u32*>(w), s));
    nonce_add(ctx->nonce, 3 * (static_cast<u64>(n) << log_n));
    return HMPC_OK;
}

int hmpc_bgv_encrypt_plain(hmpc_context ctx, hmpc_modulus mod, int log_n, size_t n, const uint32_t* pk_a,
                           const uint32_t* pk_b, const uint32_t* messages, const uint32_t* p_mont, int bits_u, int bits_v,
                           int bits_w, uint32_t* c0, uint32_t* c1, void* stream)
{
    int rc = bgv_check(ctx, mod, n, pk_a, pk_b, messages, p_mont, bits_u, bits_v, bits_w, c0, c1);
    if (rc)
        return rc;
    transform_args T;
    rc = mod->args(log_n, T);
    if (rc)
        return rc;
    if (n == 0)
        return HMPC_OK;
    device_guard g(ctx->device);
    hipStream_t s = as_stream(stream);
    int L = mod->limbs;
    std::size_t N = std::size_t{1} << log_n;
    std::size_t poly = static_cast<std::size_t>(L) * N;  // words of one polynomial
    void* w = nullptr;
    rc = ctx->workspace(s, 6 * poly * 4, &w);
    if (rc)
        return rc;
    u32* coeff = static_cast<u32*>(w);
    u32* scratch = coeff + poly;
    u32* U = scratch + poly;
    u32* V = U + poly;
    u32* M = V + poly;
    u32* Wt = M + poly;
    bgv_args B = make_bgv_args(ctx, mod, log_n, n, p_mont, bits_u, bits_v, bits_w);
    std::size_t count = n * N;
    // one ciphertext at a time, hmpc's per-op launches: sample, per-stage transforms, then
    // element-wise c0 = (A*u + v*p) + m and c1 = B*u + w*p
    for (std::size_t i = 0; i < n; ++i)
    {
        HMPC_HIP(launch_sample_cbd_capability(L, mod->P, mod->r2, B.key, B.nonce_u, bits_u, log_n, 1, coeff, i, s));
        HMPC_HIP(launch_ntt_plain(L, T, 1, false, coeff, U, scratch, s));
        HMPC_HIP(launch_sample_cbd_capability(L, mod->P, mod->r2, B.key, B.nonce_v, bits_v, log_n, 1, coeff, i, s));
        HMPC_HIP(launch_ntt_plain(L, T, 1, false, coeff, V, scratch, s));
        HMPC_HIP(launch_from_u32(L, mod->P, mod->r2, N, messages + i * N, coeff, s));
        HMPC_HIP(launch_ntt_plain(L, T, 1, false, coeff, M, scratch, s));
        // c0 (written limb-planar into the batch layout through a per-ciphertext temporary)
        HMPC_HIP(launch_pointwise(L, mod->P, HMPC_MUL, N, pk_a, U, N, coeff, s));
        HMPC_HIP(launch_pointwise_scalar(L, mod->P, HMPC_MUL, N, V, B.p_mont, V, s));
        HMPC_HIP(launch_pointwise(L, mod->P, HMPC_ADD, N, coeff, V, N, coeff, s));
        HMPC_HIP(launch_pointwise(L, mod->P, HMPC_ADD, N, coeff, M, N, coeff, s));
        for (int l = 0; l < L; ++l)
            HMPC_HIP(hipMemcpyAsync(c0 + l * count + i * N, coeff + l * N, N * 4, hipMemcpyDeviceToDevice, s));
        HMPC_HIP(launch_sample_cbd_capability(L, mod->P, mod->r2, B.key, B.nonce_w, bits_w, log_n, 1, coeff, i, s));
        HMPC_HIP(launch_ntt_plain(L, T, 1, false, coeff, Wt, scratch, s));
        HMPC_HIP(launch_pointwise(L, mod->P, HMPC_MUL, N, pk_b, U, N, coeff, s));
        HMPC_HIP(launch_pointwise_scalar(L, mod->P, HMPC_MUL, N, Wt, B.p_mont, Wt, s));
        HMPC_HIP(launch_pointwise(L, mod->P, HMPC_ADD, N, coeff, Wt, N, coeff, s));
        for (int l = 0; l < L; ++l)
            HMPC_HIP(hipMemcpyAsync(c1 + l * count + i * N, coeff + l * N, N * 4, hipMemcpyDeviceToDevice, s));
    }
    nonce_add(ctx->nonce, 3 * (static_cast<u64>(n) << log_n));
    return HMPC_OK;
}

int hmpc_bgv_decrypt(hmpc_modulus mod, int log_n, size_t n, const uint32_t* s_ntt, const uint32_t* c0, const uint32_t* c1,
                     uint32_t* messages, void* stream)
{
    if (!mod || (n && (!s_ntt || !c0 || !c1 || !messages)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    transform_args T;
    int rc = mod->args(log_n, T);
    if (rc)
        return rc;
    if (n == 0)
        return HMPC_OK;
    device_guard g(mod->ctx->device);
    hipStream_t s = as_stream(stream);
    std::size_t count = n << log_n;
    void* w = nullptr;
    rc = mod->ctx->workspace(s, count * mod->limbs * 4, &w);
    if (rc)
        return rc;
    u32* x = static_cast<u32*>(w);
    int L = mod->limbs;
    HMPC_HIP(launch_pointwise(L, mod->P, HMPC_MUL, count, c1, s_ntt, std::size_t{1} << log_n, x, s));
    HMPC_HIP(launch_pointwise(L, mod->P, HMPC_SUB, count, c0, x, count, x, s));
    HMPC_HIP(dispatch_ntt(*mod, T, n, true, x, x, s));
    HMPC_HIP(launch_centered_low32(L, mod->P, count, x, messages, s));
    return HMPC_OK;
}

// ------------------------------------------------------------------ general plaintext path
}  // extern "C"

namespace
{
    // plain 2^k mod q (k doublings of 1)
    void pow2_mod(u32* x, int k, u32 const* q, int L)
    {
        std::fill(x, x + L, 0u);
        x[0] = 1u;
        for (int i = 0; i < k; ++i)
            double_mod(x, q, L);
    }

    // plain 2^-k mod q for odd q (k halvings of 1: odd x -> (x + q) / 2)
    void half_pow2_mod(u32* x, int k, u32 const* q, int L)
    {
        u32 t[max_limbs + 1] = {1u};
        for (int i = 0; i < k; ++i)
        {
            if (t[0] & 1u)
            {
                u32 c = 0;
                for (int l = 0; l < L; ++l)
                    t[l] = addc(t[l], q[l], c, c);
                t[L] = c;
            }
            for (int l = 0; l < L; ++l)
                t[l] = (t[l] >> 1) | (t[l + 1] << 31);
            t[L] = 0;
        }
        std::copy(t, t + L, x);
    }

    // constants of the general path's kernels for one modulus
    struct lhe_consts
    {
        u32 c32[max_limbs] = {};   // 2^32 R mod q
        u32 kuni[max_limbs] = {};  // 2^(-32 W) mod q for the uniform sampler's W words
        int uniform_words = 0;
        int uniform_bits = 0;
    };

    lhe_consts make_lhe_consts(hmpc_modulus_s const& m)
    {
        lhe_consts K;
        int L = m.limbs;
        host_limbs(L, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            u32 x[LL], r2v[LL], c[LL];
            pow2_mod(x, 32, m.P.q, LL);
            std::copy(m.r2, m.r2 + LL, r2v);
            mod_mul<LL>(c, x, r2v, m.P);
            std::copy(c, c + LL, K.c32);
            return 0;
        });
        K.uniform_bits = m.bit_size + 80;
        K.uniform_words = (K.uniform_bits + 31) / 32;
        half_pow2_mod(K.kuni, 32 * K.uniform_words, m.P.q, L);
        return K;
    }

    src_modulus make_src(hmpc_modulus_s const& m)
    {
        src_modulus S{};
        S.limbs = m.limbs;
        S.nprime = m.P.nprime;
        for (int l = 0; l < m.limbs; ++l)
            S.q[l] = m.P.q[l];
        for (int l = 0; l < max_limbs; ++l)
            S.half[l] = (S.q[l] >> 1) | (l + 1 < max_limbs ? S.q[l + 1] << 31 : 0u);
        return S;
    }

    // (integer of `words` words mod q) R mod q on the host (the plaintext modulus as a mod q constant)
    void host_mont_from_words(u32* out, u32 const* w, int words, hmpc_modulus_s const& m, lhe_consts const& K)
    {
        host_limbs(m.limbs, [&](auto LC) {
            constexpr int LL = decltype(LC)::value;
            u32 r[LL] = {}, x[LL], r2v[LL], c32[LL];
            std::copy(m.r2, m.r2 + LL, r2v);
            std::copy(K.c32, K.c32 + LL, c32);
            for (int i = words - 1; i >= 0; --i)
            {
                u32 t[LL], d[LL];
                mod_mul<LL>(t, r, c32, m.P);
                for (int l = 0; l < LL; ++l)
                    x[l] = l == 0 ? w[i] : 0u;
                mod_mul<LL>(d, x, r2v, m.P);
                mod_add<LL>(r, t, d, m.P);
            }
            std::copy(r, r + LL, out);
            return 0;
        });
    }

    int check_cap(int dist, int bits, int layout, int log_n, size_t n)
    {
        if (dist < cap_uniform || dist > cap_drown_unsigned)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "unknown distribution");
        if (dist == cap_centered_binomial && !valid_cap_bits(bits))
            return fail(HMPC_ERR_INVALID_ARGUMENT, cap_bits_message);
        if (dist == cap_binomial && (bits < 1 || bits > 256))
            return fail(HMPC_ERR_INVALID_ARGUMENT, "binomial: bits must be in [1, 256]");
        if ((dist == cap_drown_signed && (bits < 2 || bits > 32 * cap_max_words)) ||
            (dist == cap_drown_unsigned && (bits < 1 || bits > 32 * cap_max_words)))
            return fail(HMPC_ERR_INVALID_ARGUMENT, "drowning: bits must be in [1 (2 signed), 384]");
        if (layout < cap_layout_element || layout > cap_layout_intt)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "unknown layout");
        if (log_n < (layout == cap_layout_element ? 0 : 1) || log_n > 30)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "log_n out of range");
        if (n > (std::size_t{1} << 40) >> log_n)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "tensor too large");
        return HMPC_OK;
    }

    // samples of the context's capability (key, current nonce) into out; advances the nonce by n N
    int sample_capability(hmpc_context ctx, hmpc_modulus_s& m, int dist, int bits, int layout, int log_n, size_t n,
                          u32* out, hipStream_t s)
    {
        lhe_consts K = make_lhe_consts(m);
        if (dist == cap_uniform && K.uniform_words > cap_max_words)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "uniform: modulus too wide for the sampler");
        cap_args C{};
        C.key = make_key(ctx->key);
        std::copy(ctx->nonce, ctx->nonce + 3, C.nonce);
        C.dist = dist;
        C.bits = bits;
        C.layout = layout;
        C.log_n = log_n;
        C.total = static_cast<u64>(n) << log_n;
        C.uniform_words = K.uniform_words;
        C.uniform_bits = K.uniform_bits;
        HMPC_HIP(launch_sample_capability(m.limbs, m.P, m.r2, K.c32, K.kuni, C, out, s));
        nonce_add(ctx->nonce, C.total);
        return HMPC_OK;
    }
}

extern "C" {

int hmpc_cast(hmpc_modulus from, hmpc_modulus to, size_t count, const uint32_t* in, uint32_t* out, void* stream)
{
    if (!from || !to || (count && (!in || !out)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (from->ctx != to->ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "cast: both moduli must belong to one context");
    if (count == 0)
        return HMPC_OK;
    device_guard g(to->ctx->device);
    lhe_consts K = make_lhe_consts(*to);
    HMPC_HIP(launch_cast(make_src(*from), to->limbs, to->P, to->r2, K.c32, count, in, out, as_stream(stream)));
    return HMPC_OK;
}

int hmpc_drown_bits(const uint32_t* bound, int bound_words, int is_signed, int security, int* bits)
{
    if (!bound || !bits || bound_words <= 0 || bound_words > 64 || security < 0 || security > 256)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad drowning bound arguments");
    std::vector<u32> bm1(bound, bound + bound_words);
    if (bit_width(bm1.data(), bound_words) == 0)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "drowning bound must be positive");
    u32 br = 1;
    for (auto& w : bm1)
        w = subb(w, 0u, br, br);
    int width = bit_width(bm1.data(), bound_words);
    int b = (width ? width + security : 0) + (is_signed ? 1 : 0);
    if (b < 1 || b > 32 * cap_max_words)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "drowning draw wider than 384 bits");
    *bits = b;
    return HMPC_OK;
}

int hmpc_sample_capability(hmpc_context ctx, hmpc_modulus mod, int distribution, int bits, int layout, int log_n, size_t n,
                           uint32_t* out, void* stream)
{
    if (!ctx || !mod || (n && !out))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (mod->ctx != ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "modulus of another context");
    int rc = check_cap(distribution, bits, layout, log_n, n);
    if (rc)
        return rc;
    if (n == 0)
        return HMPC_OK;
    device_guard g(ctx->device);
    return sample_capability(ctx, *mod, distribution, bits, layout, log_n, n, out, as_stream(stream));
}

int hmpc_bgv_encrypt_plaintext(hmpc_context ctx, hmpc_modulus mod_q, hmpc_modulus mod_p, int log_n, size_t n,
                               const uint32_t* pk_a, const uint32_t* pk_b, const uint32_t* plaintext, int plaintext_ntt,
                               int bits_u, int v_distribution, int bits_v, int bits_w, uint32_t* c0, uint32_t* c1,
                               void* stream)
{
    if (!ctx || !mod_q || !mod_p || (n && (!pk_a || !pk_b || !plaintext || !c0 || !c1)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (mod_q->ctx != ctx || mod_p->ctx != ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "moduli of another context");
    if (!valid_cap_bits(bits_u) || !valid_cap_bits(bits_w))
        return fail(HMPC_ERR_INVALID_ARGUMENT, cap_bits_message);
    if (v_distribution != cap_centered_binomial && v_distribution != cap_drown_signed)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "v: centred binomial (randomness) or signed drowning (drowning_randomness)");
    int rc = check_cap(v_distribution, bits_v, cap_layout_ntt, log_n, n);
    if (rc)
        return rc;
    transform_args Tq, Tp;
    if ((rc = mod_q->args(log_n, Tq)) != 0)
        return rc;
    if (plaintext_ntt && (rc = mod_p->args(log_n, Tp)) != 0)
        return rc;
    if (n == 0)
        return HMPC_OK;
    device_guard g(ctx->device);
    hipStream_t s = as_stream(stream);
    std::size_t count = n << log_n;
    std::size_t Lq = static_cast<std::size_t>(mod_q->limbs), Lp = static_cast<std::size_t>(mod_p->limbs);
    void* w = nullptr;
    if ((rc = ctx->workspace(s, (4 * Lq + Lp) * count * 4, &w)) != 0)
        return rc;
    u32* U = static_cast<u32*>(w);
    u32* V = U + Lq * count;
    u32* W = V + Lq * count;
    u32* M = W + Lq * count;
    u32* X = M + Lq * count;
    // message = NTT_q(cast<mod q>(INTT_p(plaintext))) (hmpc/expr/crypto/lhe/enc.hpp:43-57)
    u32 const* coeffs = plaintext;
    if (plaintext_ntt)
    {
        HMPC_HIP(dispatch_ntt(*mod_p, Tp, n, true, plaintext, X, s));
        coeffs = X;
    }
    lhe_consts Kq = make_lhe_consts(*mod_q);
    HMPC_HIP(launch_cast(make_src(*mod_p), mod_q->limbs, mod_q->P, mod_q->r2, Kq.c32, count, coeffs, M, s));
    HMPC_HIP(dispatch_ntt(*mod_q, Tq, n, false, M, M, s));
    // u, v, w on the capability stream in hmpc's cache order (nonces nonce, + nN, + 2nN), each
    // sampled in the forward transform's stage-0 layout
    HMPC_HIP(launch_sample_cbd_ntt(mod_q->limbs, Tq, make_key(ctx->key), ctx->nonce, bits_u, n, U, s));
    nonce_add(ctx->nonce, count);
    if (v_distribution == cap_centered_binomial)
    {
        HMPC_HIP(launch_sample_cbd_ntt(mod_q->limbs, Tq, make_key(ctx->key), ctx->nonce, bits_v, n, V, s));
        nonce_add(ctx->nonce, count);
    }
    else
    {
        if ((rc = sample_capability(ctx, *mod_q, v_distribution, bits_v, cap_layout_ntt, log_n, n, V, s)) != 0)
            return rc;
        HMPC_HIP(dispatch_ntt(*mod_q, Tq, n, false, V, V, s));
    }
    HMPC_HIP(launch_sample_cbd_ntt(mod_q->limbs, Tq, make_key(ctx->key), ctx->nonce, bits_w, n, W, s));
    nonce_add(ctx->nonce, count);
    // p as a mod q constant (hmpc/expr/crypto/lhe/enc.hpp:59-63)
    u32 p_const[max_limbs] = {};
    host_mont_from_words(p_const, mod_p->P.q, mod_p->limbs, *mod_q, Kq);
    HMPC_HIP(launch_bgv_combine(mod_q->limbs, mod_q->P, count, std::size_t{1} << log_n, pk_a, pk_b, U, V, W, M, p_const, c0, c1,
                                s));
    return HMPC_OK;
}

int hmpc_bgv_decrypt_plaintext(hmpc_modulus mod_q, hmpc_modulus mod_p, int log_n, size_t n, const uint32_t* s_ntt,
                               const uint32_t* c0, const uint32_t* c1, int plaintext_ntt, uint32_t* plaintext, void* stream)
{
    if (!mod_q || !mod_p || (n && (!s_ntt || !c0 || !c1 || !plaintext)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    if (mod_q->ctx != mod_p->ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "moduli of another context");
    transform_args Tq, Tp;
    int rc = mod_q->args(log_n, Tq);
    if (rc)
        return rc;
    if (plaintext_ntt && (rc = mod_p->args(log_n, Tp)) != 0)
        return rc;
    if (n == 0)
        return HMPC_OK;
    hmpc_context ctx = mod_q->ctx;
    device_guard g(ctx->device);
    hipStream_t s = as_stream(stream);
    std::size_t count = n << log_n;
    void* w = nullptr;
    if ((rc = ctx->workspace(s, count * mod_q->limbs * 4, &w)) != 0)
        return rc;
    u32* x = static_cast<u32*>(w);
    int L = mod_q->limbs;
    // hmpc/expr/crypto/lhe/dec.hpp:29-44: INTT(c0 - s c1), cast to mod p, NTT_p for NTT-form plaintexts
    HMPC_HIP(launch_pointwise(L, mod_q->P, HMPC_MUL, count, c1, s_ntt, std::size_t{1} << log_n, x, s));
    HMPC_HIP(launch_pointwise(L, mod_q->P, HMPC_SUB, count, c0, x, count, x, s));
    HMPC_HIP(dispatch_ntt(*mod_q, Tq, n, true, x, x, s));
    lhe_consts Kp = make_lhe_consts(*mod_p);
    HMPC_HIP(launch_cast(make_src(*mod_q), mod_p->limbs, mod_p->P, mod_p->r2, Kp.c32, count, x, plaintext, s));
    if (plaintext_ntt)
        HMPC_HIP(dispatch_ntt(*mod_p, Tp, n, false, plaintext, plaintext, s));
    return HMPC_OK;
}

// ------------------------------------------------------------------ kernel timing
int hmpc_context_set_timing(hmpc_context ctx, int enabled)
{
    if (!ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null context");
    device_guard g(ctx->device);
    ctx->probe.reset();
    ctx->timing = enabled != 0;
    return HMPC_OK;
}

int hmpc_context_timing(hmpc_context ctx, int index, char* name, size_t name_len, size_t* launches, double* avg_ms)
{
    if (!ctx || index < 0)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad arguments");
    device_guard g(ctx->device);
    ctx->probe.collect();
    if (static_cast<std::size_t>(index) >= ctx->probe.stats.size())
        return fail(HMPC_ERR_INVALID_ARGUMENT, "no such kernel record");
    auto const& st = ctx->probe.stats[static_cast<std::size_t>(index)];
    if (name && name_len)
    {
        std::size_t n = std::min(name_len - 1, st.first.size());
        std::memcpy(name, st.first.data(), n);
        name[n] = 0;
    }
    if (launches)
        *launches = st.second.first;
    if (avg_ms)
        *avg_ms = st.second.first ? st.second.second / static_cast<double>(st.second.first) : 0.0;
    return HMPC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ job batches
struct hmpc_batch_s
{
    struct entry
    {
        hmpc_modulus mod;
        job_desc d;
        u64 cost;
    };
    struct job_class
    {
        hmpc_modulus mod = nullptr;
        int elems = 0;
        int log_n = 0;                // any prepared size of the class (kernel parameters)
        std::vector<job_desc> jobs;   // longest first
        std::size_t table_off = 0;    // byte offsets into mem
        std::size_t counter_off = 0;
        std::size_t scratch_off = 0;  // the class's own scratch (byte offset into mem)
        unsigned grid = 0;
        u64 slot_words = 0;
        u64 cost = 0;
    };
    // work item of a launch: a tile chunk (index >= 0) or a persistent class (-1 - index)
    struct lane_item
    {
        int index;
        u64 cost;
    };
    static constexpr int aux_lanes = 3;  // streams beside the caller's (4 hardware queues per process)

    hmpc_context ctx = nullptr;
    int device = 0;  // kept for destruction in any order with the context
    std::vector<entry> entries;
    std::vector<job_class> classes;
    tile_plan tiles;               // jobs of two-pass degrees
    std::size_t tile_jobs_off = 0; // byte offsets into mem
    std::size_t tile_map_off = 0;
    std::size_t tile_ws_off = 0;
    bool built = false;
    void* mem = nullptr;
    std::size_t mem_bytes = 0;
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<std::vector<lane_item>> lanes;  // lane 0: the launch stream, then aux streams
    hipStream_t aux[aux_lanes] = {};
    hipEvent_t fork_event = nullptr;
    hipEvent_t join_events[aux_lanes] = {};

    void release()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (graph)
            (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        if (mem)
            (void)hipFree(mem);
        mem = nullptr;
        mem_bytes = 0;
        classes.clear();
        tiles = tile_plan{};
        lanes.clear();
        built = false;
    }

    ~hmpc_batch_s()
    {
        device_guard g(device);
        (void)hipDeviceSynchronize();
        release();
        if (capture_stream)
            (void)hipStreamDestroy(capture_stream);
        for (int i = 0; i < aux_lanes; ++i)
        {
            if (aux[i])
                (void)hipStreamDestroy(aux[i]);
            if (join_events[i])
                (void)hipEventDestroy(join_events[i]);
        }
        if (fork_event)
            (void)hipEventDestroy(fork_event);
    }

    // Jobs of two-pass degrees go to tile launches over job tables (unless the tuner chose the
    // persistent kernel for their degree); the others are grouped by (modulus, elements per
    // thread), each class ordered longest job first (LPT) for the persistent kernel. Every tile
    // chunk and every class gets its own workspace, so that they are independent of each other:
    // they are spread over the launch stream and up to three more (longest first onto the least
    // loaded), where small launches (a class's persistent kernel holds about one workgroup per CU)
    // and the tails of large ones overlap. Sizes the grids and the workspaces, uploads the tables.
    int build()
    {
        device_guard g(ctx->device);
        std::map<std::pair<hmpc_modulus, int>, std::size_t> index;
        std::vector<std::vector<entry const*>> members;
        std::vector<tile_input> tiled;
        for (auto const& e : entries)
        {
            int log_n = static_cast<int>(e.d.log_n);
            int te = 0;
            if (e.mod->table_job(e.d, &te))
            {
                transform_args T;
                int rc = e.mod->args(log_n, T);
                if (rc)
                    return rc;
                tiled.push_back({e.mod, te, e.d, T});
                continue;
            }
            int elems = e.d.type == job_bgv ? e.mod->bgv_elems_for(log_n) : e.mod->polymul_elems_for(log_n);
            auto key = std::make_pair(e.mod, elems);
            auto f = index.find(key);
            if (f == index.end())
            {
                f = index.emplace(key, classes.size()).first;
                job_class c;
                c.mod = e.mod;
                c.elems = elems;
                c.log_n = log_n;
                classes.push_back(c);
                members.emplace_back();
            }
            members[f->second].push_back(&e);
        }
        std::size_t off = 0;
        for (std::size_t k = 0; k < classes.size(); ++k)
        {
            auto& c = classes[k];
            auto& mem_list = members[k];
            std::stable_sort(mem_list.begin(), mem_list.end(), [](entry const* x, entry const* y) { return x->cost > y->cost; });
            for (auto const* e : mem_list)
            {
                c.jobs.push_back(e->d);
                c.slot_words = std::max(c.slot_words, job_scratch_words(e->d.type, static_cast<int>(e->d.log_n), c.mod->limbs));
                c.cost += static_cast<u64>(e->d.log_n) << e->d.log_n;
            }
            int rc = jobs_grid(*c.mod, c.elems, c.jobs.size(), &c.grid);
            if (rc)
                return rc;
            c.table_off = off;
            off += (c.jobs.size() * sizeof(job_desc) + 255) / 256 * 256;
            c.counter_off = off;
            off += counter_bytes;
        }
        tiles.build(tiled, false, true);
        tile_jobs_off = off;
        off += (tiles.jobs.size() * sizeof(job_desc) + 255) / 256 * 256;
        tile_map_off = off;
        off += (tiles.map.size() * 4 + 255) / 256 * 256;
        tile_ws_off = off;
        off += tiles.ws_words * 4;
        for (auto& c : classes)
        {
            c.scratch_off = off;
            off += (static_cast<u64>(c.grid) * c.slot_words * 4 + 255) / 256 * 256;
        }
        mem_bytes = off;
        // longest work item first onto the least loaded lane
        std::vector<lane_item> items;
        for (std::size_t i = 0; i < tiles.chunks.size(); ++i)
            items.push_back({static_cast<int>(i), tiles.chunks[i].cost});
        for (std::size_t k = 0; k < classes.size(); ++k)
            items.push_back({-1 - static_cast<int>(k), classes[k].cost});
        std::stable_sort(items.begin(), items.end(), [](lane_item const& x, lane_item const& y) { return x.cost > y.cost; });
        lanes.assign(std::min<std::size_t>(items.size(), 1 + aux_lanes), {});
        std::vector<u64> load(lanes.size(), 0);
        for (auto const& it : items)
        {
            std::size_t l = static_cast<std::size_t>(std::min_element(load.begin(), load.end()) - load.begin());
            lanes[l].push_back(it);
            load[l] += it.cost;
        }
        for (std::size_t l = 1; l < lanes.size(); ++l)
        {
            if (!aux[l - 1])
                HMPC_HIP(hipStreamCreateWithFlags(&aux[l - 1], hipStreamNonBlocking));
            if (!join_events[l - 1])
                HMPC_HIP(hipEventCreateWithFlags(&join_events[l - 1], hipEventDisableTiming));
        }
        if (lanes.size() > 1 && !fork_event)
            HMPC_HIP(hipEventCreateWithFlags(&fork_event, hipEventDisableTiming));
        HMPC_HIP(hipMalloc(&mem, mem_bytes));
        for (auto const& c : classes)
            HMPC_HIP(hipMemcpy(static_cast<char*>(mem) + c.table_off, c.jobs.data(), c.jobs.size() * sizeof(job_desc),
                               hipMemcpyHostToDevice));
        if (!tiles.empty())
        {
            HMPC_HIP(hipMemcpy(static_cast<char*>(mem) + tile_jobs_off, tiles.jobs.data(), tiles.jobs.size() * sizeof(job_desc),
                               hipMemcpyHostToDevice));
            HMPC_HIP(hipMemcpy(static_cast<char*>(mem) + tile_map_off, tiles.map.data(), tiles.map.size() * 4,
                               hipMemcpyHostToDevice));
        }
        built = true;
        return HMPC_OK;
    }

    int launch_item(lane_item const& it, hipStream_t s)
    {
        char* base = static_cast<char*>(mem);
        if (it.index >= 0)
            return tiles.launch_chunk(tiles.chunks[static_cast<std::size_t>(it.index)],
                                      reinterpret_cast<job_desc const*>(base + tile_jobs_off),
                                      reinterpret_cast<u32 const*>(base + tile_map_off), reinterpret_cast<u32*>(base + tile_ws_off),
                                      s);
        auto const& c = classes[static_cast<std::size_t>(-1 - it.index)];
        transform_args T;
        int rc = c.mod->args(c.log_n, T);
        if (rc)
            return rc;
        T.elems = c.elems;
        HMPC_HIP(launch_fast_jobs(c.mod->limbs, T, reinterpret_cast<job_desc const*>(base + c.table_off), job_desc{},
                                  static_cast<u32>(c.jobs.size()), c.grid, reinterpret_cast<u32*>(base + c.counter_off),
                                  reinterpret_cast<u32*>(base + c.scratch_off), c.slot_words, s));
        return HMPC_OK;
    }

    // lane 0 on `s`, the others on the auxiliary streams between a fork and a join event on `s`
    // (captured as branches of the graph when `s` is being captured)
    int launch(hipStream_t s)
    {
        if (lanes.size() > 1)
        {
            HMPC_HIP(hipEventRecord(fork_event, s));
            for (std::size_t l = 1; l < lanes.size(); ++l)
                HMPC_HIP(hipStreamWaitEvent(aux[l - 1], fork_event, 0));
        }
        for (std::size_t l = 0; l < lanes.size(); ++l)
        {
            hipStream_t ls = l == 0 ? s : aux[l - 1];
            for (auto const& it : lanes[l])
            {
                int rc = launch_item(it, ls);
                if (rc)
                    return rc;
            }
            if (l > 0)
            {
                HMPC_HIP(hipEventRecord(join_events[l - 1], ls));
                HMPC_HIP(hipStreamWaitEvent(s, join_events[l - 1], 0));
            }
        }
        return HMPC_OK;
    }
};

extern "C" {

int hmpc_batch_create(hmpc_context ctx, hmpc_batch* out)
{
    if (!ctx || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    auto b = std::make_unique<hmpc_batch_s>();
    b->ctx = ctx;
    b->device = ctx->device;
    *out = b.release();
    return HMPC_OK;
}

int hmpc_batch_destroy(hmpc_batch batch)
{
    delete batch;
    return HMPC_OK;
}

namespace
{
    int batch_check(hmpc_batch batch, hmpc_modulus mod, int log_n, transform_args& T)
    {
        if (!batch || !mod || mod->ctx != batch->ctx)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "batch and modulus must belong to one context");
        if (batch->built)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "batch already ran; create a new batch to add jobs");
        int rc = mod->args(log_n, T);
        if (rc)
            return rc;
        if (!mod->fast(log_n))
            return fail(HMPC_ERR_INVALID_ARGUMENT, "job batches need a modulus served by the fast kernels "
                                                   "(q = 2^124 + c with c < 2^24, or 2^64 - 2^32 + 1)");
        if (batch->entries.size() >= 0x7fffffffu)
            return fail(HMPC_ERR_INVALID_ARGUMENT, "batch too large");
        return HMPC_OK;
    }
}

int hmpc_batch_add_polymul(hmpc_batch batch, hmpc_modulus mod, int log_n, const uint32_t* a, const uint32_t* b,
                           uint32_t* out)
{
    transform_args T;
    int rc = batch_check(batch, mod, log_n, T);
    if (rc)
        return rc;
    if (!a || !b || !out)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument");
    u64 N = u64{1} << log_n;
    batch->entries.push_back({mod, polymul_job(T, log_n, a, b, out, N, 0), 3 * N * static_cast<u64>(log_n)});
    return HMPC_OK;
}

int hmpc_batch_add_bgv_encrypt(hmpc_batch batch, hmpc_modulus mod, int log_n, const uint32_t* pk_a, const uint32_t* pk_b,
                               const uint32_t* messages, const uint32_t* p_mont, int bits_u, int bits_v, int bits_w,
                               uint32_t* c0, uint32_t* c1)
{
    transform_args T;
    int rc = batch_check(batch, mod, log_n, T);
    if (rc)
        return rc;
    rc = bgv_check(batch->ctx, mod, 1, pk_a, pk_b, messages, p_mont, bits_u, bits_v, bits_w, c0, c1);
    if (rc)
        return rc;
    u64 p_int = 0;
    if (!small_plaintext(*mod, p_mont, &p_int))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "job batches need a plaintext modulus below 2^40");
    hmpc_context ctx = batch->ctx;
    bgv_args B = make_bgv_args(ctx, mod, log_n, 1, p_mont, bits_u, bits_v, bits_w);
    u64 N = u64{1} << log_n;
    batch->entries.push_back({mod, bgv_job(T, B, p_int, log_n, pk_a, pk_b, messages, c0, c1, N, 0, 0),
                              4 * N * static_cast<u64>(log_n)});
    nonce_add(ctx->nonce, 3 * N);
    return HMPC_OK;
}

int hmpc_batch_run(hmpc_batch batch, int use_graph, void* stream)
{
    if (!batch)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null batch");
    if (batch->entries.empty())
        return HMPC_OK;
    device_guard g(batch->ctx->device);
    hipStream_t s = as_stream(stream);
    if (!batch->built)
    {
        int rc = batch->build();
        if (rc)
        {
            batch->release();
            return rc;
        }
    }
    if (!use_graph)
    {
        probe_guard pg(batch->ctx);
        return batch->launch(s);
    }
    if (!batch->exec)
    {
        // capture on a private stream (the null stream cannot be captured); counters are reset
        // by a kernel node, never a memset node
        if (!batch->capture_stream)
            HMPC_HIP(hipStreamCreateWithFlags(&batch->capture_stream, hipStreamNonBlocking));
        HMPC_HIP(hipStreamBeginCapture(batch->capture_stream, hipStreamCaptureModeRelaxed));
        int rc = batch->launch(batch->capture_stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(batch->capture_stream, &graph);
        if (rc)
        {
            if (graph)
                (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess)
            return hip_fail(e, "hipStreamEndCapture");
        batch->graph = graph;
        HMPC_HIP(hipGraphInstantiate(&batch->exec, graph, nullptr, nullptr, 0));
    }
    HMPC_HIP(hipGraphLaunch(batch->exec, s));
    return HMPC_OK;
}

int hmpc_batch_info(hmpc_batch batch, size_t* jobs, size_t* classes, size_t* device_bytes)
{
    if (!batch)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null batch");
    if (jobs)
        *jobs = batch->entries.size();
    if (classes)
        *classes = batch->classes.size() + batch->tiles.chunks.size();
    if (device_bytes)
        *device_bytes = batch->mem_bytes;
    return HMPC_OK;
}

// ------------------------------------------------------------------ autotuner
int hmpc_autotune(hmpc_modulus mod, int log_n, size_t pairs, void* stream, int* elems_out, int* persistent_out,
                  float* ms_out)
{
    if (!mod || pairs == 0 || pairs > (size_t{1} << 20))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad arguments");
    transform_args T;
    int rc = mod->args(log_n, T);
    if (rc)
        return rc;
    if (!mod->fast(log_n))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "no tunable schedule: this modulus uses the portable kernels");
    device_guard g(mod->ctx->device);
    hipStream_t s = as_stream(stream);
    std::size_t count = pairs << log_n;
    std::size_t bytes = count * mod->limbs * 4;
    u32* buf = nullptr;
    HMPC_HIP(hipMalloc(&buf, 3 * bytes));
    u32* a = buf;
    u32* b = buf + count * mod->limbs;
    u32* out = b + count * mod->limbs;
    chacha_key key{};
    key.k[0] = 0x7475u;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int best_e = 0, best_p = 1;
    float best = 0.0f;
    auto cleanup = [&]() {
        if (e0)
            (void)hipEventDestroy(e0);
        if (e1)
            (void)hipEventDestroy(e1);
        (void)hipFree(buf);
    };
    if (launch_sample_uniform(mod->limbs, mod->P, mod->bit_size, key, 0, count, a, s) != hipSuccess ||
        launch_sample_uniform(mod->limbs, mod->P, mod->bit_size, key, 1, count, b, s) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    {
        cleanup();
        return fail(HMPC_ERR_DEVICE, "autotune set-up failed");
    }
    int const cand_l4[2] = {4, 8}, cand_l2[2] = {8, 16};
    int const* cand = mod->limbs == 4 ? cand_l4 : cand_l2;
    auto saved_e = mod->elems;
    auto saved_p = mod->passes;
    for (int ci = 0; ci < 2; ++ci)
        for (int persistent = 1; persistent >= 0; --persistent)
        {
            mod->elems[log_n] = cand[ci];
            mod->passes[log_n] = persistent ? 0 : 1;
            rc = hmpc_polymul(mod, log_n, pairs, a, b, out, s);  // warm-up (workspace, code objects)
            for (int r = 0; rc == HMPC_OK && r < 3; ++r)
            {
                if (r == 0)
                    (void)hipEventRecord(e0, s);
                rc = hmpc_polymul(mod, log_n, pairs, a, b, out, s);
            }
            float ms = 0.0f;
            if (rc == HMPC_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                                  hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                rc = fail(HMPC_ERR_DEVICE, "autotune timing failed");
            if (rc)
            {
                mod->elems = saved_e;
                mod->passes = saved_p;
                cleanup();
                return rc;
            }
            ms /= 3.0f;
            if (best_e == 0 || ms < best)
            {
                best = ms;
                best_e = cand[ci];
                best_p = persistent;
            }
        }
    mod->elems[log_n] = best_e;
    mod->passes[log_n] = best_p ? 0 : 1;
    cleanup();
    if (elems_out)
        *elems_out = best_e;
    if (persistent_out)
        *persistent_out = best_p;
    if (ms_out)
        *ms_out = best;
    return HMPC_OK;
}

int hmpc_autotune_bgv(hmpc_context ctx, hmpc_modulus mod, int log_n, size_t n, void* stream, int* elems_out, int* tiles_out,
                      float* ms_out)
{
    if (!ctx || !mod || n == 0 || n > (size_t{1} << 16))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad arguments");
    transform_args T;
    int rc = mod->args(log_n, T);
    if (rc)
        return rc;
    if (!mod->fast(log_n))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "no tunable schedule: this modulus uses the portable kernels");
    device_guard g(ctx->device);
    hipStream_t s = as_stream(stream);
    std::size_t N = std::size_t{1} << log_n;
    std::size_t L = static_cast<std::size_t>(mod->limbs);
    u32* buf = nullptr;
    HMPC_HIP(hipMalloc(&buf, (2 * L * N + n * N + 2 * L * n * N) * 4));
    u32* pk_a = buf;
    u32* pk_b = pk_a + L * N;
    u32* msg = pk_b + L * N;
    u32* c0 = msg + n * N;
    u32* c1 = c0 + L * n * N;
    chacha_key key{};
    key.k[0] = 0x7476u;
    u32 p_mont[max_limbs] = {};
    host_limbs(mod->limbs, [&](auto LC) {
        constexpr int LL = decltype(LC)::value;
        u32 v[LL] = {}, r2v[LL], r[LL];
        v[0] = 0;
        if constexpr (LL >= 2)
            v[1] = 1;  // 2^32
        for (int i = 0; i < LL; ++i)
            r2v[i] = mod->r2[i];
        mod_mul<LL>(r, v, r2v, mod->P);
        std::copy(r, r + LL, p_mont);
        return 0;
    });
    hipEvent_t e0 = nullptr, e1 = nullptr;
    u32 saved_nonce[3];
    std::copy(ctx->nonce, ctx->nonce + 3, saved_nonce);
    auto saved = mod->bgv_elems;
    auto cleanup = [&]() {
        std::copy(saved_nonce, saved_nonce + 3, ctx->nonce);
        if (e0)
            (void)hipEventDestroy(e0);
        if (e1)
            (void)hipEventDestroy(e1);
        (void)hipStreamSynchronize(s);
        (void)hipFree(buf);
    };
    if (launch_sample_uniform(mod->limbs, mod->P, mod->bit_size, key, 0, N, pk_a, s) != hipSuccess ||
        launch_sample_uniform(mod->limbs, mod->P, mod->bit_size, key, 1, N, pk_b, s) != hipSuccess ||
        launch_sample_words(key, 2, n * N, msg, s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
    {
        cleanup();
        return fail(HMPC_ERR_DEVICE, "autotune set-up failed");
    }
    int const cand_l4[2] = {4, 8}, cand_l2[2] = {8, 16};
    int const* cand = mod->limbs == 4 ? cand_l4 : cand_l2;
    int best_e = 0, best_tiles = 0;
    float best = 0.0f;
    auto saved_passes = mod->bgv_passes;
    for (int trial = 0; trial < (log_n > 12 ? 4 : 2); ++trial)
    {
        int ci = trial & 1;
        int tiles = trial >> 1;
        mod->bgv_elems[log_n] = cand[ci];
        mod->bgv_passes[log_n] = tiles;
        rc = hmpc_bgv_encrypt(ctx, mod, log_n, n, pk_a, pk_b, msg, p_mont, 2, 40, 40, c0, c1, s);
        for (int r = 0; rc == HMPC_OK && r < 3; ++r)
        {
            if (r == 0)
                (void)hipEventRecord(e0, s);
            rc = hmpc_bgv_encrypt(ctx, mod, log_n, n, pk_a, pk_b, msg, p_mont, 2, 40, 40, c0, c1, s);
        }
        float ms = 0.0f;
        if (rc == HMPC_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(HMPC_ERR_DEVICE, "autotune timing failed");
        if (rc)
        {
            mod->bgv_elems = saved;
            mod->bgv_passes = saved_passes;
            cleanup();
            return rc;
        }
        ms /= 3.0f;
        if (best_e == 0 || ms < best)
        {
            best = ms;
            best_e = cand[ci];
            best_tiles = tiles;
        }
    }
    mod->bgv_elems[log_n] = best_e;
    mod->bgv_passes[log_n] = best_tiles;
    cleanup();
    if (elems_out)
        *elems_out = best_e;
    if (tiles_out)
        *tiles_out = best_tiles;
    if (ms_out)
        *ms_out = best;
    return HMPC_OK;
}

// ------------------------------------------------------------------ timing
int hmpc_event_create(hmpc_context ctx, void** event)
{
    device_guard g(ctx->device);
    hipEvent_t e;
    HMPC_HIP(hipEventCreate(&e));
    *event = e;
    return HMPC_OK;
}

int hmpc_event_destroy(hmpc_context ctx, void* event)
{
    device_guard g(ctx->device);
    HMPC_HIP(hipEventDestroy(static_cast<hipEvent_t>(event)));
    return HMPC_OK;
}

int hmpc_event_record(hmpc_context ctx, void* event, void* stream)
{
    device_guard g(ctx->device);
    HMPC_HIP(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)));
    return HMPC_OK;
}

int hmpc_event_elapsed_ms(hmpc_context ctx, void* start, void* stop, float* ms)
{
    device_guard g(ctx->device);
    HMPC_HIP(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
    HMPC_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)));
    return HMPC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ serving loop
// A worker thread owns a few streams ("slots"). Requests queue in arrival order; whenever a
// slot is free the worker admits queued requests while their outputs fit the HBM budget and
// packs their ciphertexts / pairs into one persistent-kernel launch on that slot. Launches on
// different slots overlap on the device (the next launch's workgroups take CUs as the previous
// launch's workgroups finish: continuous batching at workgroup granularity). A host function
// queued behind each launch timestamps its completion. Memory comes back when the client
// releases a finished request.
namespace
{
    using server_clock = std::chrono::steady_clock;

    int pool_alloc(void*, std::size_t bytes, void** out) { return hipMalloc(out, bytes) == hipSuccess ? 0 : 1; }
    void pool_free(void*, void* p) { (void)hipFree(p); }

    struct server_key
    {
        hmpc_modulus mod;
        int log_n;
        u32 const* pk_a;
        u32 const* pk_b;
        u64 p_int;
        int bits[3];
    };

    enum request_state : int
    {
        req_queued = 0,
        req_running = 1,
        req_done = 2,
        req_failed = 3,
    };

    // A request carries everything the worker needs from its modulus, resolved on the client's thread
    // at submission (arithmetic parameters and tables, launch structure, elements per thread,
    // occupancy): the worker never reads or writes a modulus's state, which the client may change
    // meanwhile (new degrees, schedules).
    struct server_request
    {
        int kind = 0;  // 0 encrypt, 1 product
        hmpc_modulus mod = nullptr;
        transform_args T{};
        bool tile = false;   // tile launches over a job table (two-pass degrees)
        int elems = 0;       // elements per thread of its launch
        int occupancy = 0;   // persistent kernel: workgroups per CU at `elems`
        std::string err;     // why it failed
        int log_n = 0;
        std::size_t n = 0;  // ciphertexts or pairs
        int key = -1;
        u32 const* messages = nullptr;
        u32 nonce[3] = {};
        u32 const* a = nullptr;
        u32 const* b = nullptr;
        std::size_t bytes = 0;
        std::size_t cls = 0;
        void* out = nullptr;
        int state = req_queued;
        server_clock::time_point t_submit{}, t_done{};
    };

    struct server_slot
    {
        hipStream_t stream = nullptr;
        job_desc* d_table = nullptr;
        job_desc* h_table = nullptr;  // pinned staging of the job table
        std::size_t table_cap = 0;
        u32* counters = nullptr;      // one 256-byte counter per class
        u32* scratch = nullptr;
        std::size_t scratch_words = 0;
        tile_plan tiles;              // the current launch's jobs at two-pass degrees
        u32* d_map = nullptr;
        u32* h_map = nullptr;         // pinned staging of the tile map
        std::size_t map_cap = 0;
        std::vector<u64> requests;
        bool busy = false;
        std::atomic<bool> done{false};
        std::atomic<long long> done_ns{0};
        std::condition_variable* wake = nullptr;
    };

    // host function queued behind a slot's launch (runs on the runtime's callback thread: no HIP
    // calls, no server lock)
    void slot_finished(void* p)
    {
        auto* s = static_cast<server_slot*>(p);
        s->done_ns.store(std::chrono::duration_cast<std::chrono::nanoseconds>(server_clock::now().time_since_epoch()).count());
        s->done.store(true);
        s->wake->notify_one();
    }

    constexpr int max_classes_per_launch = 16;

    // workgroups per CU of the persistent kernel at `elems` for this modulus (computed once)
    int jobs_occupancy(hmpc_modulus_s& m, int elems, int* occ_out)
    {
        int& occ = m.occupancy[elems];
        if (occ == 0)
        {
            device_guard g(m.ctx->device);
            hipError_t e = fast_jobs_occupancy(m.limbs, elems, &occ);
            if (e != hipSuccess)
                return hip_fail(e, "fast_jobs_occupancy");
            occ = std::max(occ, 1);
        }
        *occ_out = occ;
        return HMPC_OK;
    }
}

struct hmpc_server_s
{
    hmpc_context ctx = nullptr;
    int device = 0;  // kept for destruction in any order with the context
    std::size_t budget = 0;
    std::size_t in_use = 0;
    std::size_t max_jobs = 4096;
    chacha_key key{};
    std::vector<server_key> keys;
    std::mutex m;
    std::condition_variable work_cv, done_cv;
    std::deque<u64> queue;
    std::unordered_map<u64, server_request> reqs;
    std::vector<std::unique_ptr<server_slot>> slots;
    block_pool pool;
    std::thread worker;
    bool stop = false;
    u64 next_id = 1;
    std::vector<std::pair<u64, double>> completed;  // finished, not yet reported: (request, latency us)
    u64 launches = 0, jobs_run = 0;

    bool idle() const
    {
        for (auto const& s : slots)
            if (s->busy)
                return false;
        return true;
    }

    // worker: completions, then admission into free slots
    void run()
    {
        device_guard g(device);
        std::unique_lock<std::mutex> lk(m);
        for (;;)
        {
            bool finished = false;
            for (auto& sp : slots)
            {
                server_slot& s = *sp;
                if (s.busy && s.done.load())
                {
                    auto t = server_clock::time_point(std::chrono::nanoseconds(s.done_ns.load()));
                    for (u64 id : s.requests)
                    {
                        auto& r = reqs[id];
                        r.state = req_done;
                        r.t_done = t;
                        completed.push_back({id, std::chrono::duration<double, std::micro>(t - r.t_submit).count()});
                    }
                    s.requests.clear();
                    s.busy = false;
                    s.done.store(false);
                    finished = true;
                }
            }
            if (finished)
                done_cv.notify_all();
            if (stop)
            {
                // shutting down: queued requests are dropped, running ones drained
                for (u64 id : queue)
                {
                    reqs[id].state = req_failed;
                    reqs[id].err = "server stopped";
                }
                queue.clear();
                if (idle())
                    break;
            }
            for (auto& sp : slots)
            {
                if (sp->busy || queue.empty())
                    continue;
                if (!admit(*sp))
                    break;
            }
            work_cv.wait_for(lk, std::chrono::microseconds(200));
        }
    }

    // a request that cannot run: its outputs go back to the pool, the client sees it failed with `why`
    void fail_request(u64 id, std::string const& why)
    {
        auto& r = reqs[id];
        if (r.out)
        {
            pool.give(r.cls, r.out);
            in_use -= r.bytes;
            r.out = nullptr;
        }
        r.state = req_failed;
        r.err = why;
        completed.push_back({id, -1.0});
    }

    // moves queued requests that fit into slot s (arrival order: the head waits for memory) and
    // launches them; false if nothing was launched
    bool admit(server_slot& s)
    {
        std::vector<u64> take;
        launch_admission adm;
        adm.max_jobs = max_jobs;
        adm.max_classes = max_classes_per_launch;
        bool failed = false;
        while (!queue.empty())
        {
            u64 id = queue.front();
            auto& r = reqs[id];
            if (!adm.fits(r.n, r.mod, r.elems, !r.tile))
                break;
            if (in_use + r.bytes > budget)
                break;
            void* out = nullptr;
            int took = pool.take(r.cls, &out);
            if (took != block_pool::take_ok)
            {
                // this request alone fails; the server and the requests behind it go on
                queue.pop_front();
                fail_request(id, took == block_pool::take_out_of_memory ? "device out of memory for the request's outputs"
                                                                        : "request outputs exceed the memory budget");
                failed = true;
                continue;
            }
            r.out = out;
            in_use += r.bytes;
            r.state = req_running;
            take.push_back(id);
            adm.add(r.n, r.mod, r.elems, !r.tile);
            queue.pop_front();
        }
        if (failed)
            done_cv.notify_all();
        if (take.empty())
            return false;
        if (launch(s, take, adm.jobs) != HMPC_OK)
        {
            std::string why = last_error;
            for (u64 id : take)
                fail_request(id, why);
            done_cv.notify_all();
            return false;
        }
        return true;
    }

    // Requests at two-pass degrees run as tile launches over a table of all their ciphertexts
    // or pairs; N <= 4096 (and degrees tuned to it) as persistent launches per class. Uses only what
    // the requests captured at submission.
    int launch(server_slot& s, std::vector<u64> const& take, std::size_t jobs)
    {
        struct cls_info
        {
            hmpc_modulus mod;
            int elems;
            int occupancy;
            transform_args T;
            std::vector<std::pair<u64, job_desc>> jobs;  // (cost, job)
        };
        std::vector<cls_info> classes;
        std::vector<tile_input> tiled;
        for (u64 id : take)
        {
            auto& r = reqs[id];
            transform_args const& T = r.T;
            auto it = classes.end();
            if (!r.tile)
            {
                it = std::find_if(classes.begin(), classes.end(), [&](cls_info const& c) { return c.mod == r.mod && c.elems == r.elems; });
                if (it == classes.end())
                {
                    classes.push_back({r.mod, r.elems, r.occupancy, T, {}});
                    it = classes.end() - 1;
                }
            }
            u64 N = u64{1} << r.log_n;
            u64 count = static_cast<u64>(r.n) << r.log_n;
            auto* out = static_cast<u32*>(r.out);
            int limbs = r.T.P.limbs;
            if (r.kind == 0)
            {
                server_key const& k = keys[static_cast<std::size_t>(r.key)];
                bgv_args B{};
                B.key = key;
                std::copy(r.nonce, r.nonce + 3, B.nonce_u);
                std::copy(r.nonce, r.nonce + 3, B.nonce_v);
                nonce_add(B.nonce_v, count);
                std::copy(B.nonce_v, B.nonce_v + 3, B.nonce_w);
                nonce_add(B.nonce_w, count);
                B.bits_u = k.bits[0];
                B.bits_v = k.bits[1];
                B.bits_w = k.bits[2];
                u32* c1 = out + static_cast<std::size_t>(limbs) * count;
                for (std::size_t i = 0; i < r.n; ++i)
                {
                    job_desc J = bgv_job(T, B, k.p_int, r.log_n, k.pk_a, k.pk_b, r.messages, out, c1, count, i * N,
                                         static_cast<u32>(i));
                    if (r.tile)
                        tiled.push_back({r.mod, r.elems, J, T});
                    else
                        it->jobs.push_back({4 * N * r.log_n, J});
                }
            }
            else
            {
                for (std::size_t i = 0; i < r.n; ++i)
                {
                    job_desc J = polymul_job(T, r.log_n, r.a, r.b, out, count, i * N);
                    if (r.tile)
                        tiled.push_back({r.mod, r.elems, J, T});
                    else
                        it->jobs.push_back({3 * N * r.log_n, J});
                }
            }
        }
        // the slot is idle: its table, counters and scratch may be resized
        if (s.table_cap < jobs)
        {
            if (s.d_table)
                HMPC_HIP(hipFree(s.d_table));
            if (s.h_table)
                HMPC_HIP(hipHostFree(s.h_table));
            s.d_table = nullptr;
            s.h_table = nullptr;
            std::size_t cap = std::max(jobs, max_jobs);
            HMPC_HIP(hipMalloc(&s.d_table, cap * sizeof(job_desc)));
            HMPC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_table), cap * sizeof(job_desc)));
            s.table_cap = cap;
        }
        std::size_t need = 0;
        std::vector<unsigned> grids(classes.size());
        std::vector<u64> slot_words(classes.size());
        for (std::size_t c = 0; c < classes.size(); ++c)
        {
            auto& cl = classes[c];
            std::stable_sort(cl.jobs.begin(), cl.jobs.end(), [](auto const& x, auto const& y) { return x.first > y.first; });
            u64 w = 0;
            for (auto const& j : cl.jobs)
                w = std::max(w, job_scratch_words(j.second.type, static_cast<int>(j.second.log_n), cl.T.P.limbs));
            std::size_t g = static_cast<std::size_t>(cl.occupancy) * static_cast<std::size_t>(ctx->props.multiProcessorCount);
            grids[c] = static_cast<unsigned>(std::max<std::size_t>(1, std::min(g, cl.jobs.size())));
            slot_words[c] = w;
            need = std::max<std::size_t>(need, static_cast<std::size_t>(grids[c]) * w);
        }
        s.tiles.build(tiled, true);
        need = std::max<std::size_t>(need, s.tiles.ws_words);
        if (s.map_cap < s.tiles.map.size())
        {
            if (s.d_map)
                HMPC_HIP(hipFree(s.d_map));
            if (s.h_map)
                HMPC_HIP(hipHostFree(s.h_map));
            s.d_map = nullptr;
            s.h_map = nullptr;
            std::size_t cap = std::max<std::size_t>(s.tiles.map.size(), 1 << 16);
            HMPC_HIP(hipMalloc(&s.d_map, cap * 4));
            HMPC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s.h_map), cap * 4));
            s.map_cap = cap;
        }
        if (s.scratch_words < need)
        {
            if (s.scratch)
                HMPC_HIP(hipFree(s.scratch));
            s.scratch = nullptr;
            HMPC_HIP(hipMalloc(&s.scratch, need * 4));
            s.scratch_words = need;
        }
        std::size_t off = 0;
        for (auto const& cl : classes)
            for (auto const& j : cl.jobs)
                s.h_table[off++] = j.second;
        std::size_t tile_off = off;
        std::copy(s.tiles.jobs.begin(), s.tiles.jobs.end(), s.h_table + off);
        off += s.tiles.jobs.size();
        HMPC_HIP(hipMemcpyAsync(s.d_table, s.h_table, off * sizeof(job_desc), hipMemcpyHostToDevice, s.stream));
        if (!s.tiles.empty())
        {
            std::copy(s.tiles.map.begin(), s.tiles.map.end(), s.h_map);
            HMPC_HIP(hipMemcpyAsync(s.d_map, s.h_map, s.tiles.map.size() * 4, hipMemcpyHostToDevice, s.stream));
            int rc = s.tiles.launch(s.d_table + tile_off, s.d_map, s.scratch, s.stream);
            if (rc)
                return rc;
        }
        off = 0;
        for (std::size_t c = 0; c < classes.size(); ++c)
        {
            auto& cl = classes[c];
            transform_args T = cl.T;
            T.elems = cl.elems;
            HMPC_HIP(launch_fast_jobs(T.P.limbs, T, s.d_table + off, job_desc{}, static_cast<u32>(cl.jobs.size()), grids[c],
                                      s.counters + 64 * c, s.scratch, slot_words[c], s.stream));
            off += cl.jobs.size();
        }
        s.done.store(false);
        s.requests = take;
        s.busy = true;
        HMPC_HIP(hipLaunchHostFunc(s.stream, slot_finished, &s));
        launches += 1;
        jobs_run += jobs;
        return HMPC_OK;
    }
};

extern "C" {

int hmpc_server_create(hmpc_context ctx, size_t budget_bytes, int slots, hmpc_server* out)
{
    if (!ctx || !out || budget_bytes == 0 || slots < 1 || slots > 8)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "server: context, budget > 0 and 1..8 slots required");
    device_guard g(ctx->device);
    auto srv = std::make_unique<hmpc_server_s>();
    srv->ctx = ctx;
    srv->device = ctx->device;
    srv->budget = budget_bytes;
    srv->pool.alloc = pool_alloc;
    srv->pool.release = pool_free;
    srv->pool.cap = budget_bytes;
    std::copy(ctx->key, ctx->key + 8, srv->key.k);
    for (int i = 0; i < slots; ++i)
    {
        auto s = std::make_unique<server_slot>();
        s->wake = &srv->work_cv;
        HMPC_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        HMPC_HIP(hipMalloc(&s->counters, 256 * max_classes_per_launch));
        srv->slots.push_back(std::move(s));
    }
    hmpc_server_s* raw = srv.get();
    srv->worker = std::thread([raw]() { raw->run(); });
    *out = srv.release();
    return HMPC_OK;
}

int hmpc_server_destroy(hmpc_server srv)
{
    if (!srv)
        return HMPC_OK;
    {
        std::lock_guard<std::mutex> lk(srv->m);
        srv->stop = true;
    }
    srv->work_cv.notify_all();
    if (srv->worker.joinable())
        srv->worker.join();
    device_guard g(srv->device);
    for (auto& s : srv->slots)
    {
        (void)hipStreamSynchronize(s->stream);
        (void)hipStreamDestroy(s->stream);
        if (s->d_table)
            (void)hipFree(s->d_table);
        if (s->h_table)
            (void)hipHostFree(s->h_table);
        if (s->d_map)
            (void)hipFree(s->d_map);
        if (s->h_map)
            (void)hipHostFree(s->h_map);
        if (s->counters)
            (void)hipFree(s->counters);
        if (s->scratch)
            (void)hipFree(s->scratch);
    }
    for (auto& [id, r] : srv->reqs)
        if (r.out)
            srv->pool.give(r.cls, r.out);
    srv->pool.release_all();
    delete srv;
    return HMPC_OK;
}

int hmpc_server_add_key(hmpc_server srv, hmpc_modulus mod, int log_n, const uint32_t* pk_a, const uint32_t* pk_b,
                        const uint32_t* p_mont, int bits_u, int bits_v, int bits_w, int* key_id)
{
    if (!srv || !mod || !pk_a || !pk_b || !p_mont || !key_id || mod->ctx != srv->ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null argument or foreign modulus");
    if (!valid_cap_bits(bits_u) || !valid_cap_bits(bits_v) || !valid_cap_bits(bits_w))
        return fail(HMPC_ERR_INVALID_ARGUMENT, cap_bits_message);
    int rc = mod->prepare(log_n);
    if (rc)
        return rc;
    if (!mod->fast(log_n))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "serving needs a modulus served by the fast kernels");
    u64 p_int = 0;
    if (!small_plaintext(*mod, p_mont, &p_int))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "serving needs a plaintext modulus below 2^40");
    std::lock_guard<std::mutex> lk(srv->m);
    srv->keys.push_back({mod, log_n, pk_a, pk_b, p_int, {bits_u, bits_v, bits_w}});
    *key_id = static_cast<int>(srv->keys.size() - 1);
    return HMPC_OK;
}

namespace
{
    // the request's modulus-dependent launch parameters, resolved on the client's thread
    int resolve_request(server_request& r)
    {
        int rc = r.mod->args(r.log_n, r.T);
        if (rc)
            return rc;
        if (r.kind == 0)
        {
            r.tile = r.mod->table_tiles(r.log_n);
            r.elems = r.tile ? r.mod->table_elems_for(r.log_n) : r.mod->bgv_elems_for(r.log_n);
        }
        else
        {
            r.tile = r.mod->table_polymul(r.log_n);
            r.elems = r.mod->polymul_elems_for(r.log_n);
        }
        r.T.elems = r.elems;
        return r.tile ? HMPC_OK : jobs_occupancy(*r.mod, r.elems, &r.occupancy);
    }

    int server_enqueue(hmpc_server srv, server_request r, uint64_t* request)
    {
        if (r.bytes > srv->budget)
            return fail(HMPC_ERR_BUDGET, "request outputs exceed the server's memory budget");
        int rc = resolve_request(r);
        if (rc)
            return rc;
        {
            std::lock_guard<std::mutex> lk(srv->m);
            u64 id = srv->next_id++;
            r.t_submit = server_clock::now();
            srv->reqs.emplace(id, r);
            srv->queue.push_back(id);
            *request = id;
        }
        srv->work_cv.notify_one();
        return HMPC_OK;
    }
}

int hmpc_server_submit_encrypt(hmpc_server srv, int key_id, size_t k, const uint32_t* messages, const uint32_t* nonce3,
                               uint64_t* request)
{
    if (!srv || !messages || !nonce3 || !request || k == 0 || k > 65536)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad arguments");
    server_request r;
    {
        std::lock_guard<std::mutex> lk(srv->m);
        if (key_id < 0 || static_cast<std::size_t>(key_id) >= srv->keys.size())
            return fail(HMPC_ERR_INVALID_ARGUMENT, "no such key");
        server_key const& key = srv->keys[static_cast<std::size_t>(key_id)];
        r.mod = key.mod;
        r.log_n = key.log_n;
    }
    r.kind = 0;
    r.key = key_id;
    r.n = k;
    r.messages = messages;
    std::copy(nonce3, nonce3 + 3, r.nonce);
    std::size_t bytes = 2 * (k << r.log_n) * static_cast<std::size_t>(r.mod->limbs) * 4;
    r.cls = block_pool::class_of(bytes);
    r.bytes = r.cls;
    return server_enqueue(srv, r, request);
}

int hmpc_server_submit_polymul(hmpc_server srv, hmpc_modulus mod, int log_n, size_t n, const uint32_t* a, const uint32_t* b,
                               uint64_t* request)
{
    if (!srv || !mod || !a || !b || !request || n == 0 || n > 65536 || mod->ctx != srv->ctx)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad arguments");
    int rc = mod->prepare(log_n);
    if (rc)
        return rc;
    if (!mod->fast(log_n))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "serving needs a modulus served by the fast kernels");
    server_request r;
    r.kind = 1;
    r.mod = mod;
    r.log_n = log_n;
    r.n = n;
    r.a = a;
    r.b = b;
    std::size_t bytes = (n << log_n) * static_cast<std::size_t>(mod->limbs) * 4;
    r.cls = block_pool::class_of(bytes);
    r.bytes = r.cls;
    return server_enqueue(srv, r, request);
}

int hmpc_server_poll(hmpc_server srv, uint64_t request, int* state, double* latency_us)
{
    if (!srv)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(srv->m);
    auto it = srv->reqs.find(request);
    if (it == srv->reqs.end())
        return fail(HMPC_ERR_INVALID_ARGUMENT, "no such request");
    if (state)
        *state = it->second.state;
    if (latency_us)
        *latency_us = it->second.state == req_done
                          ? std::chrono::duration<double, std::micro>(it->second.t_done - it->second.t_submit).count()
                          : 0.0;
    return it->second.state == req_failed ? fail(HMPC_ERR_DEVICE, "server: " + it->second.err) : HMPC_OK;
}

int hmpc_server_wait(hmpc_server srv, uint64_t request)
{
    if (!srv)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(srv->m);
    for (;;)
    {
        auto it = srv->reqs.find(request);
        if (it == srv->reqs.end())
            return fail(HMPC_ERR_INVALID_ARGUMENT, "no such request");
        if (it->second.state == req_done)
            return HMPC_OK;
        if (it->second.state == req_failed)
            return fail(HMPC_ERR_DEVICE, "server: " + it->second.err);
        // nothing runs and the head of the queue waits for memory that only a release frees: waiting
        // would not end
        if (it->second.state == req_queued && srv->idle() && !srv->queue.empty() &&
            srv->in_use + srv->reqs[srv->queue.front()].bytes > srv->budget)
            return fail(HMPC_ERR_BUDGET, "server: queued behind outputs that are not released (hmpc_server_release)");
        srv->done_cv.wait_for(lk, std::chrono::milliseconds(10));
    }
}

int hmpc_server_result(hmpc_server srv, uint64_t request, uint32_t** c0, uint32_t** c1)
{
    if (!srv)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(srv->m);
    auto it = srv->reqs.find(request);
    if (it == srv->reqs.end())
        return fail(HMPC_ERR_INVALID_ARGUMENT, "no such request");
    if (it->second.state != req_done)
        return fail(HMPC_ERR_NOT_READY, "request not finished");
    server_request const& r = it->second;
    auto* out = static_cast<u32*>(r.out);
    if (c0)
        *c0 = out;
    if (c1)
        *c1 = r.kind == 0 ? out + static_cast<std::size_t>(r.mod->limbs) * (r.n << r.log_n) : nullptr;
    return HMPC_OK;
}

int hmpc_server_release(hmpc_server srv, uint64_t request)
{
    if (!srv)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->m);
        auto it = srv->reqs.find(request);
        if (it == srv->reqs.end())
            return fail(HMPC_ERR_INVALID_ARGUMENT, "no such request");
        server_request& r = it->second;
        if (r.state == req_running)
            return fail(HMPC_ERR_NOT_READY, "request still running");
        if (r.state == req_queued)
        {
            srv->queue.erase(std::find(srv->queue.begin(), srv->queue.end(), request));
            srv->reqs.erase(it);
            return HMPC_OK;
        }
        if (r.out)
        {
            srv->pool.give(r.cls, r.out);
            srv->in_use -= r.bytes;
        }
        srv->reqs.erase(it);
    }
    srv->work_cv.notify_one();
    return HMPC_OK;
}

int hmpc_server_completed(hmpc_server srv, uint64_t* requests, double* latency_us, size_t max, size_t* count)
{
    if (!srv || !count || (max && (!requests || !latency_us)))
        return fail(HMPC_ERR_INVALID_ARGUMENT, "bad arguments");
    std::lock_guard<std::mutex> lk(srv->m);
    std::size_t n = std::min(max, srv->completed.size());
    for (std::size_t i = 0; i < n; ++i)
    {
        requests[i] = srv->completed[i].first;
        latency_us[i] = srv->completed[i].second;
    }
    srv->completed.erase(srv->completed.begin(), srv->completed.begin() + static_cast<std::ptrdiff_t>(n));
    *count = n;
    return HMPC_OK;
}

int hmpc_server_stats(hmpc_server srv, size_t* queued, size_t* running, size_t* bytes_in_use, uint64_t* launches,
                      uint64_t* jobs)
{
    if (!srv)
        return fail(HMPC_ERR_INVALID_ARGUMENT, "null server");
    std::lock_guard<std::mutex> lk(srv->m);
    if (queued)
        *queued = srv->queue.size();
    if (running)
    {
        std::size_t r = 0;
        for (auto const& s : srv->slots)
            r += s->requests.size();
        *running = r;
    }
    if (bytes_in_use)
        *bytes_in_use = srv->in_use;
    if (launches)
        *launches = srv->launches;
    if (jobs)
        *jobs = srv->jobs_run;
    return HMPC_OK;
}

}  // extern "C"
