// CPU check of the serving loop's host policy (hmpc_amd/csrc/server_pool.hpp) with an allocator that
// fails on demand: the output pool stays within its cap, an allocation failure fails that request
// only and leaves the pool usable, trimming returns other classes' free blocks before the pool grows
// past the cap, and admission stops at the class and job limits. Prints "ok" or the failed check.
#include "../../hmpc_amd/csrc/server_pool.hpp"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

using hmpc_amd::block_pool;
using hmpc_amd::launch_admission;

namespace
{
    struct fake_device
    {
        std::size_t bytes = 0;      // currently allocated
        std::size_t limit = 0;      // allocations beyond this fail (device memory)
        int fail_next = 0;          // the next n allocations fail regardless
        std::size_t allocs = 0, frees = 0;
        std::map<char*, std::size_t> live;
    };

    int fake_alloc(void* user, std::size_t n, void** out)
    {
        auto* d = static_cast<fake_device*>(user);
        if (d->fail_next > 0)
        {
            --d->fail_next;
            return 1;
        }
        if (d->bytes + n > d->limit)
            return 1;
        char* p = static_cast<char*>(std::malloc(1));
        d->bytes += n;
        d->allocs += 1;
        d->live[p] = n;
        *out = p;
        return 0;
    }


    void fake_free(void* user, void* p)
    {
        auto* d = static_cast<fake_device*>(user);
        d->frees += 1;
        d->bytes -= d->live.at(static_cast<char*>(p));
        d->live.erase(static_cast<char*>(p));
        std::free(p);
    }

    int failures = 0;
    void expect(bool ok, char const* what)
    {
        if (!ok)
        {
            std::printf("FAILED: %s\n", what);
            ++failures;
        }
    }

    constexpr std::size_t MiB = std::size_t{1} << 20;
}

int main()
{
    // --- class sizes
    expect(block_pool::class_of(1) == MiB, "class of 1 byte is 1 MiB");
    expect(block_pool::class_of(MiB + 1) == 2 * MiB, "class rounds up to a power of two");

    // --- the pool never holds more than its cap, over a random take/give sequence
    {
        fake_device dev;
        dev.limit = std::size_t{1} << 40;
        block_pool pool;
        pool.alloc = fake_alloc;
        pool.release = fake_free;
        pool.user = &dev;
        pool.cap = 64 * MiB;
        std::mt19937 rng(5);
        std::vector<std::pair<std::size_t, void*>> held;
        std::size_t in_use = 0;
        for (int step = 0; step < 20000; ++step)
        {
            bool do_take = held.empty() || rng() % 2;
            if (do_take)
            {
                std::size_t cls = MiB << (rng() % 6);  // 1..32 MiB
                if (in_use + cls > pool.cap)
                    continue;  // the server's admission keeps in_use + request within the budget
                void* p = nullptr;
                int rc = pool.take(cls, &p);
                expect(rc == block_pool::take_ok, "take within the budget succeeds");
                if (rc == block_pool::take_ok)
                {
                    held.push_back({cls, p});
                    in_use += cls;
                }
            }
            else
            {
                std::size_t i = rng() % held.size();
                pool.give(held[i].first, held[i].second);
                in_use -= held[i].first;
                held[i] = held.back();
                held.pop_back();
            }
            expect(pool.held <= pool.cap, "pool holds at most its cap");
            expect(pool.held == in_use + pool.free_bytes(), "held = in use + free lists");
            if (failures)
                break;
        }
        for (auto& [cls, p] : held)
            pool.give(cls, p);
        pool.release_all();
        expect(pool.held == 0 && dev.live.empty(), "release_all returns every block");
    }

    // --- an allocation failure fails one take; the pool stays usable and frees its cache first
    {
        fake_device dev;
        dev.limit = 8 * MiB;
        block_pool pool;
        pool.alloc = fake_alloc;
        pool.release = fake_free;
        pool.user = &dev;
        pool.cap = 64 * MiB;
        void *a = nullptr, *b = nullptr, *c = nullptr;
        expect(pool.take(4 * MiB, &a) == block_pool::take_ok, "first block");
        expect(pool.take(2 * MiB, &b) == block_pool::take_ok, "second block");
        pool.give(2 * MiB, b);  // cached, 6 MiB held on an 8 MiB device
        // 4 MiB more does not fit the device: the cached 2 MiB block is returned and the retry fits
        expect(pool.take(4 * MiB, &c) == block_pool::take_ok, "retry after returning the cache");
        expect(dev.frees == 1 && pool.free_bytes() == 0, "the cached block was returned to the device");
        // the device is full now: an 8 MiB block fails, but only that one
        void* d = nullptr;
        expect(pool.take(8 * MiB, &d) == block_pool::take_out_of_memory, "device full: that take fails");
        pool.give(4 * MiB, a);
        void* e = nullptr;
        expect(pool.take(4 * MiB, &e) == block_pool::take_ok && e == a, "a later take reuses a freed block");
        // injected failures of both attempts
        dev.limit = std::size_t{1} << 40;
        dev.fail_next = 2;
        void* f = nullptr;
        expect(pool.take(16 * MiB, &f) == block_pool::take_out_of_memory, "injected failure of both attempts");
        expect(pool.take(16 * MiB, &f) == block_pool::take_ok, "the next take succeeds");
        pool.give(4 * MiB, c);
        pool.give(4 * MiB, e);
        pool.give(16 * MiB, f);
        pool.release_all();
        expect(dev.live.empty(), "everything returned");
    }

    // --- trimming: other classes' free blocks go back before the pool grows past its cap
    {
        fake_device dev;
        dev.limit = std::size_t{1} << 40;
        block_pool pool;
        pool.alloc = fake_alloc;
        pool.release = fake_free;
        pool.user = &dev;
        pool.cap = 16 * MiB;
        void *a = nullptr, *b = nullptr, *c = nullptr;
        pool.take(8 * MiB, &a);
        pool.take(4 * MiB, &b);
        pool.give(8 * MiB, a);
        pool.give(4 * MiB, b);  // 12 MiB cached
        expect(pool.take(16 * MiB, &c) == block_pool::take_ok, "a 16 MiB block after trimming");
        expect(pool.held == 16 * MiB && pool.free_bytes() == 0, "both cached blocks were returned");
        pool.give(16 * MiB, c);
        pool.release_all();
    }

    // --- admission: job limit (a first request may exceed it) and at most max_classes persistent classes
    {
        launch_admission adm;
        adm.max_jobs = 100;
        adm.max_classes = 2;
        int m1 = 0, m2 = 0;
        expect(adm.fits(500, &m1, 4, true), "a first request may exceed the job limit");
        adm.add(500, &m1, 4, true);
        expect(!adm.fits(1, &m1, 4, true), "then the job limit holds");
        launch_admission b;
        b.max_jobs = 1000;
        b.max_classes = 2;
        b.add(1, &m1, 4, true);
        b.add(1, &m1, 8, true);
        expect(b.fits(1, &m1, 4, true) && b.fits(1, &m1, 8, true), "existing classes fit");
        expect(!b.fits(1, &m2, 4, true), "a third class does not");
        expect(b.fits(1, &m2, 4, false), "tile launches open no class");
        expect(b.classes.size() == 2 && b.jobs == 2 && b.requests == 2, "bookkeeping");
    }

    if (failures == 0)
        std::printf("ok\n");
    return failures == 0 ? 0 : 1;
}
