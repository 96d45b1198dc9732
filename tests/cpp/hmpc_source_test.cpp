// Code on hmpc's NTT / BGV path written in hmpc's own spelling (include/hmpc_amd/hmpc.hpp): compile-
// time moduli from _int literals, poly_mod types, make_tensor<T>(shape), host_accessor, the queue's
// planned expressions with a shared capability sampler, lhe::enc / ciphertext operators /
// dec<plaintext>. Three workloads with the parameters of hmpc's own code:
//   linear_bgv   the linear BGV case of tests/expr/crypto/lhe/enc.cpp:15-72 (q of 279 bits, p of 64
//                bits, N = 2^16, shape {4, 2}, keys and plaintext of raw words i, s = one)
//   bgv_example  the expression of examples/bgv.cpp:25-90 (150-bit q, p as a constant, shared u)
//   ntt_example  the transform of examples/ntt.cpp:22-55 (125-bit p, N = 2^20)
// Each writes its tensors (limb-planar u32 words, a u64 word count first) to the file argv[1];
// tests/test_gpu_hmpc_source.py recomputes them with the oracle.
#include <hmpc_amd/hmpc.hpp>

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

namespace
{
    std::vector<std::vector<std::uint32_t>> dumps;

    template<typename T>
    void dump(hmpc::comp::tensor<T>& t)
    {
        hmpc::comp::host_accessor acc(t, hmpc::access::read);
        using element = typename hmpc::comp::tensor<T>::element_type;
        std::vector<std::uint32_t> words(acc.size() * element::limb_size);
        for (hmpc::size i = 0; i < acc.size(); ++i)
        {
            element v = acc[i];
            for (hmpc::size l = 0; l < element::limb_size; ++l)
                words[l * acc.size() + i] = v.data[l];
        }
        dumps.push_back(std::move(words));
    }

    int failures = 0;
    void expect(bool ok, char const* what)
    {
        if (!ok)
        {
            std::fprintf(stderr, "FAILED: %s\n", what);
            ++failures;
        }
    }

    constexpr std::array<std::uint32_t, 8> queue_key = {42, 0, 0, 0, 0, 0, 0, 1};

    void linear_bgv()
    {
        using namespace hmpc::ints::literals;
        constexpr auto p = 0x8822'd806'2332'0001_int;
        constexpr auto q = 0x59'1f5b'834c'0d96'1f67'343b'cc89'02bd'eda2'771f'5430'6ff1'5116'2ff8'd2b4'0f41'94dc'0001_int;
        constexpr hmpc::size N = hmpc::size{1} << 16;
        using Rq = hmpc::ints::poly_mod<q, N, hmpc::ints::coefficient_representation>;
        using ntt_Rq = hmpc::ints::traits::number_theoretic_transform_type_t<Rq>;
        using mod_q = Rq::element_type;
        using limb = mod_q::limb_type;
        using ntt_Rp = hmpc::ints::traits::number_theoretic_transform_type_t<hmpc::ints::poly_mod<p, N, hmpc::ints::coefficient_representation>>;
        using mod_p = ntt_Rp::element_type;
        static_assert(mod_q::limb_size == 9 && mod_p::limb_size == 2);

        hmpc::comp::queue queue{queue_key};
        hmpc::shape shape{4, 2};

        // public key halves of raw words i (their Montgomery representation is i), the secret one
        auto a = hmpc::comp::make_tensor<ntt_Rq>(hmpc::shape{});
        {
            hmpc::comp::host_accessor acc(a, hmpc::access::discard_write);
            for (hmpc::size i = 0; i < N; ++i)
                acc[i] = hmpc::ints::num::bit_copy<mod_q>(hmpc::core::size_limb_span<limb>(i));
        }
        auto b = a;
        auto s = hmpc::comp::make_tensor<ntt_Rq>(hmpc::shape{});
        {
            hmpc::comp::host_accessor acc(s, hmpc::access::discard_write);
            for (hmpc::size i = 0; i < N; ++i)
                acc[i] = hmpc::ints::integer_traits<mod_q>::one;
        }
        auto x = hmpc::comp::make_tensor<ntt_Rp>(shape);
        {
            hmpc::comp::host_accessor acc(x, hmpc::access::discard_write);
            for (hmpc::size i = 0; i < shape.size() * N; ++i)
                acc[i] = hmpc::ints::num::bit_copy<mod_p>(hmpc::core::size_limb_span<limb>(i));
        }

        auto key = hmpc::comp::crypto::lhe::key{a, b};
        auto r = hmpc::expr::crypto::lhe::randomness<Rq>(shape);
        auto c = queue(hmpc::expr::crypto::lhe::enc(hmpc::expr::crypto::lhe::key(key), hmpc::expr::tensor(x), r));
        auto ct = hmpc::expr::crypto::lhe::ciphertext(c);
        auto [sum, twice] = queue(ct + ct, hmpc::expr::value(mod_q{hmpc::ints::ubigint<2>{2}}) * ct);
        auto y = queue(hmpc::expr::crypto::lhe::dec<ntt_Rp>(hmpc::expr::tensor(s), ct));

        expect(y.shape().rank == 2 && y.shape().get(hmpc::constants::zero) == 4 && y.shape().get(hmpc::constants::one) == 2,
               "dec<plaintext> keeps the plaintext's shape {4, 2}");
        {
            hmpc::comp::host_accessor acc(y, hmpc::access::read);
            bool all = true;
            for (hmpc::size i = 0; i < shape.size() * N && all; ++i)
                all = acc[i] == hmpc::ints::num::bit_copy<mod_p>(hmpc::core::size_limb_span<limb>(i));
            expect(all, "dec<plaintext>(s, enc(x)) == x");
        }
        dump(c.c0);
        dump(c.c1);
        dump(sum.c0);
        dump(sum.c1);
        dump(twice.c0);
        dump(twice.c1);
        dump(y);
    }

    void bgv_example()
    {
        using namespace hmpc::ints::literals;
        using namespace hmpc::expr::operators;
        constexpr auto q = 0x2faeadbe7a0195c011ac195ad10269830e8001_int;
        constexpr hmpc::size N = 1 << 14;
        using R = hmpc::ints::poly_mod<q, N, hmpc::ints::coefficient_representation>;
        using mod_q = R::element_type;
        constexpr auto p = hmpc::expr::constant_of<mod_q{0x89d046e0ae640001_int}>;
        constexpr hmpc::size n = 4;
        constexpr hmpc::size argc = 1;  // the example run without arguments

        hmpc::comp::queue queue{queue_key};
        auto a_tensor = hmpc::comp::make_tensor<R>(hmpc::shape{});
        auto b_tensor = hmpc::comp::make_tensor<R>(hmpc::shape{});
        auto m_tensor = hmpc::comp::make_tensor<R>(hmpc::shape{n});
        {
            hmpc::comp::host_accessor a(a_tensor, hmpc::access::discard_write);
            hmpc::comp::host_accessor b(b_tensor, hmpc::access::discard_write);
            for (hmpc::size i = 0; i < N; ++i)
            {
                a[i] = static_cast<mod_q>(hmpc::ints::ubigint<32>{argc + n + i});
                b[i] = static_cast<mod_q>(hmpc::ints::ubigint<32>{argc + N - i});
            }
            hmpc::comp::host_accessor m(m_tensor, hmpc::access::discard_write);
            for (hmpc::size i = 0; i < N * n; ++i)
                m[i] = static_cast<mod_q>(hmpc::ints::ubigint<32>{argc + n - i});
        }
        auto [ntt_a, ntt_b, ntt_m] = queue(hmpc::expr::number_theoretic_transform(hmpc::expr::tensor(a_tensor)),
                                           hmpc::expr::number_theoretic_transform(hmpc::expr::tensor(b_tensor)),
                                           hmpc::expr::number_theoretic_transform(hmpc::expr::tensor(m_tensor)));
        auto a = hmpc::expr::tensor(ntt_a);
        auto b = hmpc::expr::tensor(ntt_b);
        auto m = hmpc::expr::tensor(ntt_m);
        auto u = hmpc::expr::number_theoretic_transform(hmpc::expr::random::centered_binomial<R>(hmpc::shape{n}, hmpc::constants::half));
        auto v = hmpc::expr::number_theoretic_transform(hmpc::expr::random::centered_binomial<R>(hmpc::shape{n}, hmpc::constants::ten));
        auto w = hmpc::expr::number_theoretic_transform(hmpc::expr::random::centered_binomial<R>(hmpc::shape{n}, hmpc::constants::ten));
        auto [c0, c1] = queue(hmpc::expr::unsqueeze(b, hmpc::constants::minus_one) * u + v * p + m,
                              hmpc::expr::unsqueeze(a, hmpc::constants::minus_one) * u + w * p);
        queue.wait();
        expect(c0.shape() == hmpc::shape{n} && c1.shape() == hmpc::shape{n}, "ciphertext halves of shape {n}");
        dump(ntt_a);
        dump(ntt_b);
        dump(c0);
        dump(c1);
    }

    void ntt_example()
    {
        using namespace hmpc::ints::literals;
        constexpr auto p = 0x18f25cd9a75ccbd9c146d4abaec00001_int;
        constexpr hmpc::size N = 1 << 20;
        using R = hmpc::ints::poly_mod<p, N, hmpc::ints::coefficient_representation>;
        using mod_p = R::element_type;
        constexpr hmpc::size n = 2;
        constexpr hmpc::size argc = 1;

        hmpc::comp::queue queue{queue_key};
        auto tensor = hmpc::comp::make_tensor<R>(hmpc::shape{n});
        {
            hmpc::comp::host_accessor x(tensor, hmpc::access::discard_write);
            for (hmpc::size i = 0; i < N * n; ++i)
                x[i] = static_cast<mod_p>(hmpc::ints::ubigint<32>{argc + N * n - i});
        }
        auto y = queue(hmpc::expr::number_theoretic_transform(hmpc::expr::tensor(tensor)));
        // INTT(NTT(x)) through a second queue call returns x
        auto back = queue(hmpc::expr::inverse_number_theoretic_transform(hmpc::expr::tensor(y)));
        {
            hmpc::comp::host_accessor bx(back, hmpc::access::read);
            hmpc::comp::host_accessor x(tensor, hmpc::access::read);
            bool same = true;
            for (hmpc::size i = 0; i < N * n && same; ++i)
                same = bx[i] == static_cast<mod_p>(x[i]);
            expect(same, "INTT(NTT(x)) == x");
        }
        dump(y);
    }
}

int main(int argc, char** argv)
{
    if (argc < 2)
    {
        std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
        return 2;
    }
    try
    {
        std::printf("linear_bgv\n");
        std::fflush(stdout);
        linear_bgv();
        std::printf("bgv_example\n");
        std::fflush(stdout);
        bgv_example();
        std::printf("ntt_example\n");
        std::fflush(stdout);
        ntt_example();
    }
    catch (std::exception const& e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::ofstream f(argv[1], std::ios::binary);
    for (auto const& v : dumps)
    {
        std::uint64_t count = v.size();
        f.write(reinterpret_cast<char const*>(&count), sizeof count);
        f.write(reinterpret_cast<char const*>(v.data()), static_cast<std::streamsize>(v.size() * 4));
    }
    return failures == 0 ? 0 : 1;
}
