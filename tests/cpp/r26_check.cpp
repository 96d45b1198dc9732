// Host build of the device arithmetic in hmpc_amd/csrc/r26.hpp, driven by tests/test_r26_host.py.
// Reads commands from stdin, prints results; python checks them with big integers.
//   M c a0..a4 b0..b4      -> r26_mont_signed limbs (a: signed 32-bit limbs; the result's top limb signed)
//   U c a0..a4 b0..b4      -> r26_mont limbs (non-negative a)
//   F c x0..x4             -> r26_fold limbs
//   G c x0..x4             -> r26_fold_signed limbs (x: signed 32-bit limbs)
//   C c x0..x4             -> r26_canonical 32-bit words
//   S c z                  -> r26_from_signed limbs (z a signed 64-bit integer)
//   P c a0..a4 b0..b4      -> r26_point limbs (a, b: signed 32-bit limbs; the result's top limb signed)
//   B c stages x0..x4 y0..y4 w0..w4 (per stage: w limbs)  -> forward butterfly chain (signed limbs)
//   T c a0..a4 W0..W4      -> r26_mont_tw limbs of a (signed limbs) and the expanded twiddle of W
//   X c W0..W4             -> r26_expand_twiddle's 25 words
//   BX / IX                -> B / I chains whose stages name their product: after each stage's
//                             twiddle limbs (W, as stored) a kind, 0 = mont, 1 = mont_tw
#include "../../hmpc_amd/csrc/r26.hpp"

#include <cstdio>
#include <iostream>
#include <string>

using namespace hmpc_amd;

static p26 make(u32 c)
{
    p26 R{};
    r26_constants(R, c);
    return R;
}

int main()
{
    std::string cmd;
    while (std::cin >> cmd)
    {
        u32 c;
        std::cin >> c;
        p26 P = make(c);
        if (cmd == "M")
        {
            u32 a[5], b[5], r[5];
            for (auto& v : a)
            {
                long long z;
                std::cin >> z;
                v = static_cast<u32>(z);
            }
            for (auto& v : b) std::cin >> v;
            r26_mont_signed(r, a, b, P);
            printf("%u %u %u %u %d\n", r[0], r[1], r[2], r[3], static_cast<int>(r[4]));
        }
        else if (cmd == "P")
        {
            u32 a[5], b[5];
            for (auto* p : {&a, &b})
                for (auto& v : *p)
                {
                    long long z;
                    std::cin >> z;
                    v = static_cast<u32>(z);
                }
            r26_point(a, b, P);
            printf("%d %d %d %d %d\n", static_cast<int>(a[0]), static_cast<int>(a[1]), static_cast<int>(a[2]),
                   static_cast<int>(a[3]), static_cast<int>(a[4]));
        }
        else if (cmd == "T" || cmd == "X")
        {
            u32 a[5] = {}, W[5], r[5], V[r26_tw_words];
            if (cmd == "T")
                for (auto& v : a)
                {
                    long long z;
                    std::cin >> z;
                    v = static_cast<u32>(z);
                }
            for (auto& v : W) std::cin >> v;
            r26_expand_twiddle(V, W, P);
            if (cmd == "X")
            {
                for (int i = 0; i < r26_tw_words; ++i)
                    printf("%u%s", V[i], i + 1 < r26_tw_words ? " " : "\n");
            }
            else
            {
                r26_mont_tw(r, a, V, P);
                printf("%u %u %u %u %d\n", r[0], r[1], r[2], r[3], static_cast<int>(r[4]));
            }
        }
        else if (cmd == "U")
        {
            u32 a[5], b[5], r[5];
            for (auto& v : a) std::cin >> v;
            for (auto& v : b) std::cin >> v;
            r26_mont(r, a, b, P);
            printf("%u %u %u %u %u\n", r[0], r[1], r[2], r[3], r[4]);
        }
        else if (cmd == "G")
        {
            u32 x[5];
            for (auto& v : x)
            {
                long long z;
                std::cin >> z;
                v = static_cast<u32>(z);
            }
            r26_fold_signed(x, P);
            printf("%u %u %u %u %u\n", x[0], x[1], x[2], x[3], x[4]);
        }
        else if (cmd == "F")
        {
            u32 x[5];
            for (auto& v : x) std::cin >> v;
            r26_fold(x, P);
            printf("%u %u %u %u %u\n", x[0], x[1], x[2], x[3], x[4]);
        }
        else if (cmd == "S")
        {
            long long z;
            std::cin >> z;
            u32 x[5];
            r26_from_signed(x, z, P);
            printf("%u %u %u %u %u\n", x[0], x[1], x[2], x[3], x[4]);
        }
        else if (cmd == "C")
        {
            u32 x[5], w[4];
            for (auto& v : x) std::cin >> v;
            r26_canonical(w, x, P);
            printf("%u %u %u %u\n", w[0], w[1], w[2], w[3]);
        }
        else if (cmd == "B" || cmd == "I" || cmd == "BX" || cmd == "IX")
        {
            // chain of butterflies on one pair: forward (x, y) -> (x + t, x - t), t = mont(y, w) (signed digits),
            // or inverse (x + y, mont(x - y, w)) with x + y folded on request (signed digits). Per stage: the
            // twiddle limbs, then for "B" a selector of the next stage's pair (0: (x', y'),
            // 1: (y', y'), 2: (x', x')), for "I" the fold flag and a selector (0: (x', y'),
            // 1: (x', x'), 2: (y', y')); prints the pair after every stage
            int stages;
            std::cin >> stages;
            u32 x[5], y[5];
            for (auto* p : {&x, &y})
                for (auto& v : *p)
                {
                    long long z;
                    std::cin >> z;
                    v = static_cast<u32>(z);
                }
            for (int s = 0; s < stages; ++s)
            {
                u32 w[5];
                for (auto& v : w) std::cin >> v;
                int kind = 0;
                if (cmd.size() == 2)
                    std::cin >> kind;
                u32 V[r26_tw_words];
                if (kind)
                    r26_expand_twiddle(V, w, P);
                if (cmd[0] == 'B')
                {
                    int sel;
                    std::cin >> sel;
                    if (kind)
                        r26_forward_tw(x, y, V, P);
                    else
                        r26_forward(x, y, w, P);
                    for (int l = 0; l < 5; ++l)
                    {
                        if (sel == 1)
                            x[l] = y[l];
                        else if (sel == 2)
                            y[l] = x[l];
                    }
                }
                else
                {
                    int fold, sel;
                    std::cin >> fold >> sel;
                    if (kind)
                        r26_inverse_tw(x, y, V, P, fold != 0);
                    else
                        r26_inverse(x, y, w, P, fold != 0);
                    for (int l = 0; l < 5; ++l)
                    {
                        if (sel == 1)
                            y[l] = x[l];
                        else if (sel == 2)
                            x[l] = y[l];
                    }
                }
                // signed digits: every limb prints as a signed 32-bit value
                for (int l = 0; l < 5; ++l)
                    printf("%d ", static_cast<int>(x[l]));
                for (int l = 0; l < 5; ++l)
                    printf("%d%s", static_cast<int>(y[l]), l < 4 ? " " : (s + 1 < stages ? " " : "\n"));
            }
        }
        fflush(stdout);
    }
    return 0;
}
