// Host check of the compile-time integer layer of include/hmpc_amd/hmpc.hpp (no device calls):
// prints, per modulus, "limbs bit_size q r1 r2 nprime" and the stored words of converted values
// and of integer_traits<mod>::one, as hexadecimal integers; then literal and shape checks.
#include <hmpc_amd/hmpc.hpp>

#include <cstdio>

namespace
{
    template<typename W>
    void print_words(W const* w, std::size_t n)
    {
        std::printf(" 0x");
        for (std::size_t i = n; i-- > 0;)
            std::printf("%08x", static_cast<unsigned>(w[i]));
    }

    template<typename M>
    void report(char const* name)
    {
        std::printf("%s %zu %zu", name, M::limb_size, M::bit_size);
        print_words(M::q.w, M::limb_size);
        print_words(M::r1.w, M::limb_size);
        print_words(M::r2.w, M::limb_size);
        std::printf(" %u", static_cast<unsigned>(M::nprime));
        // stored words of small values, of a 96-bit value and of one
        for (std::uint64_t v : {0ull, 1ull, 2ull, 12345ull, 0xffffffffull})
        {
            M m{hmpc::ints::ubigint<32>{v}};
            print_words(m.data, M::limb_size);
        }
        hmpc::ints::ubigint<96> wide{};
        wide.data[0] = 0x89abcdefu;
        wide.data[1] = 0x01234567u;
        wide.data[2] = 0xfedcba98u;
        M mw{wide};
        print_words(mw.data, M::limb_size);
        auto one = hmpc::ints::integer_traits<M>::one;
        print_words(one.data, M::limb_size);
        // a Montgomery product of two converted values: (3 R)(5 R) R^-1
        M three{hmpc::ints::ubigint<32>{3}}, five{hmpc::ints::ubigint<32>{5}};
        M prod = three * five;
        print_words(prod.data, M::limb_size);
        std::printf("\n");
    }
}

int main()
{
    using namespace hmpc::ints::literals;
    constexpr auto q279 = 0x59'1f5b'834c'0d96'1f67'343b'cc89'02bd'eda2'771f'5430'6ff1'5116'2ff8'd2b4'0f41'94dc'0001_int;
    constexpr auto p64 = 0x8822'd806'2332'0001_int;
    constexpr auto q150 = 0x2faeadbe7a0195c011ac195ad10269830e8001_int;
    constexpr auto p125 = 0x18f25cd9a75ccbd9c146d4abaec00001_int;
    constexpr auto gold = 18446744069414584321_int;
    constexpr auto small = 12289_int;
    report<hmpc::ints::mod<q279>>("q279");
    report<hmpc::ints::mod<p64>>("p64");
    report<hmpc::ints::mod<q150>>("q150");
    report<hmpc::ints::mod<p125>>("p125");
    report<hmpc::ints::mod<gold>>("gold");
    report<hmpc::ints::mod<small>>("small");
    // literal widths: bit width of the value
    std::printf("widths %zu %zu %zu %zu %zu\n", decltype(q279)::bit_size, decltype(p64)::bit_size, decltype(gold)::bit_size,
                decltype(small)::bit_size, decltype(0b1011_int)::bit_size);
    // shapes: size, axes by constant (negative from the end)
    hmpc::shape s{4, 2, 3};
    std::printf("shape %zu %zu %zu %zu %zu\n", s.size(), s.get(hmpc::constants::zero), s.get(hmpc::constants::minus_one),
                hmpc::shape{}.size(), s.rank);
    return 0;
}
