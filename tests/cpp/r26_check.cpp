// Host build of the device arithmetic in hmpc_amd/csrc/r26.hpp, driven by tests/test_r26_host.py.
// Reads commands from stdin, prints results; python checks them with big integers.
//   M c a0..a4 b0..b4      -> r26_mont limbs
//   F c x0..x4             -> r26_fold limbs
//   C c x0..x4             -> r26_canonical 32-bit words
//   B c stages x0..x4 y0..y4 w0..w4 (per stage: w limbs)  -> forward butterfly chain
#include "../../hmpc_amd/csrc/r26.hpp"

#include <cstdio>
#include <iostream>
#include <string>

using namespace hmpc_amd;

static p26 make(u32 c, u32 q_unused)
{
    (void)q_unused;
    p26 R{};
    R.c = c;
    u32 inv = 1;
    for (int i = 0; i < 6; ++i)
        inv *= 2u - c * inv;  // q mod 2^32 = c (q = 2^124 + c)
    R.n26 = (0u - inv) & m26;
    u32 const k4[5] = {4 * c + (1u << 26), (1u << 26) - 1, (1u << 26) - 1, (1u << 26) - 1, (1u << 22) - 1};
    u32 const k32[5] = {32 * c + (1u << 29), (1u << 29) - 8, (1u << 29) - 8, (1u << 29) - 8, (1u << 25) - 8};
    u32 const qb[5] = {c + (1u << 26), (1u << 26) - 1, (1u << 26) - 1, (1u << 26) - 1, (1u << 20) - 1};
    for (int i = 0; i < 5; ++i)
    {
        R.k4[i] = k4[i];
        R.k32[i] = k32[i];
        R.qb[i] = qb[i];
    }
    R.q32[0] = c;
    R.q32[1] = 0;
    R.q32[2] = 0;
    R.q32[3] = 0x10000000u;
    R.t20 = 1u << 20;
    return R;
}

int main()
{
    std::string cmd;
    while (std::cin >> cmd)
    {
        u32 c;
        std::cin >> c;
        p26 P = make(c, 0);
        if (cmd == "M")
        {
            u32 a[5], b[5], r[5];
            for (auto& v : a) std::cin >> v;
            for (auto& v : b) std::cin >> v;
            r26_mont(r, a, b, P);
            printf("%u %u %u %u %u\n", r[0], r[1], r[2], r[3], r[4]);
        }
        else if (cmd == "F")
        {
            u32 x[5];
            for (auto& v : x) std::cin >> v;
            r26_fold(x, P);
            printf("%u %u %u %u %u\n", x[0], x[1], x[2], x[3], x[4]);
        }
        else if (cmd == "C")
        {
            u32 x[5], w[4];
            for (auto& v : x) std::cin >> v;
            r26_canonical(w, x, P);
            printf("%u %u %u %u\n", w[0], w[1], w[2], w[3]);
        }
        else if (cmd == "B" || cmd == "I")
        {
            // chain of butterflies on one pair: forward (x, y) -> (x + t, x - t + 4q) keeping the
            // first output and the second alternately, or inverse (fold(x + y), mont(x - y + 32q, w))
            int stages;
            std::cin >> stages;
            u32 x[5], y[5];
            for (auto& v : x) std::cin >> v;
            for (auto& v : y) std::cin >> v;
            for (int s = 0; s < stages; ++s)
            {
                u32 w[5];
                for (auto& v : w) std::cin >> v;
                if (cmd == "B")
                {
                    u32 t[5];
                    r26_mont(t, y, w, P);
                    for (int l = 0; l < 5; ++l)
                    {
                        u32 xl = x[l];
                        x[l] = xl + t[l];
                        y[l] = xl + (P.k4[l] - t[l]);
                    }
                }
                else
                {
                    u32 d[5];
                    for (int l = 0; l < 5; ++l)
                    {
                        d[l] = x[l] + (P.k32[l] - y[l]);
                        x[l] = x[l] + y[l];
                    }
                    r26_fold(x, P);
                    r26_mont(y, d, w, P);
                }
            }
            printf("%u %u %u %u %u %u %u %u %u %u\n", x[0], x[1], x[2], x[3], x[4], y[0], y[1], y[2], y[3], y[4]);
        }
        fflush(stdout);
    }
    return 0;
}
