// Host build of the capability draw helpers in hmpc_amd/csrc/chacha.hpp: reads lines
// "w0 w1 bits" and prints capability_small_words(w0, w1, bits) and capability_two_words(bits).
#include "../../hmpc_amd/csrc/chacha.hpp"

#include <iostream>

int main()
{
    unsigned long long w0, w1;
    int bits;
    while (std::cin >> w0 >> w1 >> bits)
    {
        int v = hmpc_amd::capability_small_words(static_cast<hmpc_amd::u32>(w0), static_cast<hmpc_amd::u32>(w1), bits);
        std::cout << v << ' ' << (hmpc_amd::capability_two_words(bits) ? 1 : 0) << std::endl;
    }
    return 0;
}
