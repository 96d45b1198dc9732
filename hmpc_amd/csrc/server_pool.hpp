// Host-side policy of the serving loop (no device code, no HIP types: the allocator comes in as
// callbacks, so the policy is also checked on the CPU with injected failures).
//
// block_pool: request outputs in power-of-two blocks (>= 1 MiB), allocated one by one and recycled
// through per-class free lists. The bytes held from the device (blocks in use plus free blocks)
// stay within the cap (the server's budget): before a block would take the pool past it, free
// blocks of other classes are returned to the device, largest first. When the device refuses an
// allocation, every free block is returned and the allocation is tried once more; if that fails
// too, the request that needed it fails on its own and the pool stays usable.
//
// launch_admission: which queued requests one launch takes, in arrival order: at most max_jobs
// jobs (a first request may exceed it alone), and at most max_classes persistent-kernel classes
// (modulus, elements per thread); requests of tile launches open no class.
#pragma once

#include <algorithm>
#include <cstddef>
#include <map>
#include <utility>
#include <vector>

namespace hmpc_amd
{
    struct block_pool
    {
        // returns 0 on success
        using alloc_fn = int (*)(void* user, std::size_t bytes, void** out);
        using free_fn = void (*)(void* user, void* p);

        enum take_status
        {
            take_ok = 0,
            take_out_of_memory = 1,  // the device refused the block, even after trimming
            take_over_cap = 2,       // the block does not fit the cap, even with every free block returned
        };

        alloc_fn alloc = nullptr;
        free_fn release = nullptr;
        void* user = nullptr;
        std::size_t cap = 0;
        std::size_t held = 0;  // bytes allocated from the device: in use plus free lists
        std::map<std::size_t, std::vector<void*>> free_lists;

        static std::size_t class_of(std::size_t bytes)
        {
            std::size_t c = std::size_t{1} << 20;
            while (c < bytes)
                c <<= 1;
            return c;
        }

        std::size_t free_bytes() const
        {
            std::size_t b = 0;
            for (auto const& [cls, v] : free_lists)
                b += cls * v.size();
            return b;
        }

        // return free blocks to the device, largest class first (skipping `keep`), until a block of
        // `need` bytes fits the cap (need = 0 with keep = 0: every free block)
        void trim(std::size_t need, std::size_t keep)
        {
            for (auto it = free_lists.rbegin(); it != free_lists.rend(); ++it)
            {
                if (it->first == keep)
                    continue;
                auto& v = it->second;
                while (!v.empty() && (need == 0 || held + need > cap))
                {
                    release(user, v.back());
                    v.pop_back();
                    held -= it->first;
                }
                if (need != 0 && held + need <= cap)
                    return;
            }
        }

        int take(std::size_t cls, void** out)
        {
            auto& fl = free_lists[cls];
            if (!fl.empty())
            {
                *out = fl.back();
                fl.pop_back();
                return take_ok;
            }
            if (held + cls > cap)
                trim(cls, cls);
            if (held + cls > cap)
                return take_over_cap;
            void* p = nullptr;
            if (alloc(user, cls, &p) != 0)
            {
                trim(0, 0);
                if (alloc(user, cls, &p) != 0)
                    return take_out_of_memory;
            }
            held += cls;
            *out = p;
            return take_ok;
        }

        void give(std::size_t cls, void* p) { free_lists[cls].push_back(p); }

        // every free block back to the device (blocks still in use are the caller's to give first)
        void release_all()
        {
            trim(0, 0);
            free_lists.clear();
        }
    };

    struct launch_admission
    {
        std::size_t max_jobs = 0;
        std::size_t max_classes = 0;
        std::size_t jobs = 0;
        std::size_t requests = 0;
        std::vector<std::pair<void const*, int>> classes;

        // a request of n jobs in the persistent class (mod, elems), or of tile launches (persistent
        // false), fits this launch
        bool fits(std::size_t n, void const* mod, int elems, bool persistent) const
        {
            if (requests && jobs + n > max_jobs)
                return false;
            if (persistent &&
                std::find(classes.begin(), classes.end(), std::make_pair(mod, elems)) == classes.end() &&
                classes.size() == max_classes)
                return false;
            return true;
        }

        void add(std::size_t n, void const* mod, int elems, bool persistent)
        {
            jobs += n;
            requests += 1;
            auto key = std::make_pair(mod, elems);
            if (persistent && std::find(classes.begin(), classes.end(), key) == classes.end())
                classes.push_back(key);
        }
    };
}
