// Code on hmpc's NTT / BGV path written in hmpc's own spelling (include/hmpc_amd/hmpc.hpp): compile-
// time moduli from _int literals, poly_mod types, make_tensor<T>(shape), host_accessor, the queue's
// planned expressions with a shared capability sampler, lhe::enc / ciphertext operators /
// dec<plaintext>. Three workloads with the parameters of hmpc's own code:
//   linear_bgv   the linear BGV case of tests/expr/crypto/lhe/enc.cpp:15-72 (q of 279 bits, p of 64
//                bits, N = 2^16, shape {4, 2}, keys and plaintext of raw words i, s = one)
//   bgv_example  the expression of examples/bgv.cpp:25-90 (150-bit q, p as a constant, shared u)
//   ntt_example  the transform of examples/ntt.cpp:22-55 (125-bit p, N = 2^20)
//   cache        the execution-cache sizes that tests/expr/cache.cpp:13-57 requires (checked here)
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

    // every element i of t set to value(i) through one host accessor
    template<typename T, typename F>
    void fill(hmpc::comp::tensor<T>& t, F value)
    {
        hmpc::comp::host_accessor acc(t, hmpc::access::discard_write);
        for (hmpc::size i = 0; i < acc.size(); ++i)
            acc[i] = value(i);
    }

    // whether every element i of t equals value(i)
    template<typename T, typename F>
    bool all_equal(hmpc::comp::tensor<T>& t, F value)
    {
        hmpc::comp::host_accessor acc(t, hmpc::access::read);
        for (hmpc::size i = 0; i < acc.size(); ++i)
            if (!(acc[i] == value(i)))
                return false;
        return true;
    }

    // hmpc's linear BGV case: 279-bit q, 64-bit p, N = 2^16, a batch of shape {4, 2}
    void linear_bgv()
    {
        using namespace hmpc::ints::literals;
        constexpr auto q = 0x59'1f5b'834c'0d96'1f67'343b'cc89'02bd'eda2'771f'5430'6ff1'5116'2ff8'd2b4'0f41'94dc'0001_int;
        constexpr auto p = 0x8822'd806'2332'0001_int;
        constexpr hmpc::size degree = hmpc::size{1} << 16;
        using coeff_q = hmpc::ints::poly_mod<q, degree, hmpc::ints::coefficient_representation>;
        using eval_q = hmpc::ints::traits::number_theoretic_transform_type_t<coeff_q>;
        using eval_p = hmpc::ints::traits::number_theoretic_transform_type_t<hmpc::ints::poly_mod<p, degree, hmpc::ints::coefficient_representation>>;
        using zq = eval_q::element_type;
        using zp = eval_p::element_type;
        static_assert(zq::limb_size == 9 && zp::limb_size == 2);
        // raw words i as the stored (Montgomery) representation
        auto raw_q = [](hmpc::size i) { return hmpc::ints::num::bit_copy<zq>(hmpc::core::size_limb_span<zq::limb_type>(i)); };
        auto raw_p = [](hmpc::size i) { return hmpc::ints::num::bit_copy<zp>(hmpc::core::size_limb_span<zp::limb_type>(i)); };

        hmpc::comp::queue queue{queue_key};
        hmpc::shape batch{4, 2};
        auto pk_half = hmpc::comp::make_tensor<eval_q>(hmpc::shape{});
        fill(pk_half, raw_q);
        auto secret = hmpc::comp::make_tensor<eval_q>(hmpc::shape{});
        fill(secret, [](hmpc::size) { return hmpc::ints::integer_traits<zq>::one; });
        auto plaintext = hmpc::comp::make_tensor<eval_p>(batch);
        fill(plaintext, raw_p);

        // both key halves are the same tensor
        hmpc::comp::crypto::lhe::key public_key{pk_half, pk_half};
        auto c = queue(hmpc::expr::crypto::lhe::enc(hmpc::expr::crypto::lhe::key(public_key), hmpc::expr::tensor(plaintext),
                                                    hmpc::expr::crypto::lhe::randomness<coeff_q>(batch)));
        auto c_expr = hmpc::expr::crypto::lhe::ciphertext(c);
        auto two = hmpc::expr::value(zq{hmpc::ints::ubigint<2>{2}});
        auto [doubled, scaled] = queue(c_expr + c_expr, two * c_expr);
        auto decrypted = queue(hmpc::expr::crypto::lhe::dec<eval_p>(hmpc::expr::tensor(secret), c_expr));

        auto shape = decrypted.shape();
        expect(shape.rank == 2 && shape.get(hmpc::constants::zero) == 4 && shape.get(hmpc::constants::one) == 2,
               "dec<plaintext> keeps the plaintext's shape {4, 2}");
        expect(all_equal(decrypted, raw_p), "dec<plaintext>(s, enc(x)) == x");
        for (auto* t : {&c.c0, &c.c1, &doubled.c0, &doubled.c1, &scaled.c0, &scaled.c1})
            dump(*t);
        dump(decrypted);
    }

    // the ciphertext expression of hmpc's bgv example: 150-bit q, N = 2^14, p as a constant, one u
    // for both halves; inputs as the example computes them when run with no arguments (argc 1)
    void bgv_example()
    {
        using namespace hmpc::ints::literals;
        using namespace hmpc::expr::operators;
        constexpr auto q = 0x2faeadbe7a0195c011ac195ad10269830e8001_int;
        constexpr hmpc::size degree = 1 << 14;
        constexpr hmpc::size count = 4;
        constexpr hmpc::size argc = 1;
        using poly = hmpc::ints::poly_mod<q, degree, hmpc::ints::coefficient_representation>;
        using zq = poly::element_type;
        auto word = [](hmpc::size v) { return static_cast<zq>(hmpc::ints::ubigint<32>{v}); };

        hmpc::comp::queue queue{queue_key};
        auto in_a = hmpc::comp::make_tensor<poly>(hmpc::shape{});
        auto in_b = hmpc::comp::make_tensor<poly>(hmpc::shape{});
        auto in_m = hmpc::comp::make_tensor<poly>(hmpc::shape{count});
        fill(in_a, [&](hmpc::size i) { return word(argc + count + i); });
        fill(in_b, [&](hmpc::size i) { return word(argc + degree - i); });
        fill(in_m, [&](hmpc::size i) { return word(argc + count - i); });

        using hmpc::expr::number_theoretic_transform;
        auto [a_hat, b_hat, m_hat] = queue(number_theoretic_transform(hmpc::expr::tensor(in_a)),
                                           number_theoretic_transform(hmpc::expr::tensor(in_b)),
                                           number_theoretic_transform(hmpc::expr::tensor(in_m)));
        auto sample = [&](auto variance) {
            return number_theoretic_transform(hmpc::expr::random::centered_binomial<poly>(hmpc::shape{count}, variance));
        };
        auto u = sample(hmpc::constants::half);
        auto v = sample(hmpc::constants::ten);
        auto w = sample(hmpc::constants::ten);
        constexpr auto p = hmpc::expr::constant_of<zq{0x89d046e0ae640001_int}>;
        auto row = [](auto const& t) { return hmpc::expr::unsqueeze(hmpc::expr::tensor(t), hmpc::constants::minus_one); };
        auto [c0, c1] = queue(row(b_hat) * u + v * p + hmpc::expr::tensor(m_hat), row(a_hat) * u + w * p);
        queue.wait();
        expect(c0.shape() == hmpc::shape{count} && c1.shape() == hmpc::shape{count}, "ciphertext halves of shape {n}");
        for (auto* t : {&a_hat, &b_hat, &c0, &c1})
            dump(*t);
    }

    // the transform of hmpc's ntt example: 125-bit p, N = 2^20, two polynomials (argc 1)
    void ntt_example()
    {
        using namespace hmpc::ints::literals;
        constexpr auto p = 0x18f25cd9a75ccbd9c146d4abaec00001_int;
        constexpr hmpc::size degree = 1 << 20;
        constexpr hmpc::size count = 2;
        constexpr hmpc::size argc = 1;
        using poly = hmpc::ints::poly_mod<p, degree, hmpc::ints::coefficient_representation>;
        using zp = poly::element_type;
        auto input = [](hmpc::size i) { return static_cast<zp>(hmpc::ints::ubigint<32>{argc + degree * count - i}); };

        hmpc::comp::queue queue{queue_key};
        auto x = hmpc::comp::make_tensor<poly>(hmpc::shape{count});
        fill(x, input);
        auto x_hat = queue(hmpc::expr::number_theoretic_transform(hmpc::expr::tensor(x)));
        // a second call takes it back
        auto back = queue(hmpc::expr::inverse_number_theoretic_transform(hmpc::expr::tensor(x_hat)));
        expect(all_equal(back, input), "INTT(NTT(x)) == x");
        dump(x_hat);
    }

    // the plan's kernels without running them, for the two cases of hmpc's cache test: a top-level
    // sum with a cache node reached once, given twice (one kernel), and NTTs of three capability
    // samplers in (u + v, w) (u, v, the sum and w: four kernels of two polynomials each)
    void cache()
    {
        using namespace hmpc::expr::operators;
        constexpr auto p = hmpc::ints::ubigint<5>{17};
        using mod_p = hmpc::ints::mod<p>;
        using R = hmpc::ints::poly_mod<p, 8, hmpc::ints::coefficient_representation>;
        hmpc::comp::queue queue{queue_key};

        auto x = hmpc::comp::make_tensor<mod_p>(hmpc::shape{10});
        auto x_expr = hmpc::expr::tensor(x);
        auto sum = x_expr + hmpc::expr::cache(x_expr);
        expect(queue.execution_cache(sum, sum).size == 1, "x + cache(x) twice: one kernel");

        auto shape = hmpc::shape{2};
        auto u = hmpc::expr::number_theoretic_transform(hmpc::expr::random::centered_binomial<R>(shape, hmpc::constants::half));
        auto v = hmpc::expr::number_theoretic_transform(hmpc::expr::random::centered_binomial<R>(shape, hmpc::constants::ten));
        auto w = hmpc::expr::number_theoretic_transform(hmpc::expr::random::centered_binomial<R>(shape, hmpc::constants::ten));
        expect(u.shape().size() == 2, "u has shape {2}");
        auto plan = queue.execution_cache(u + v, w);
        expect(plan.size == 4, "(u + v, w): four kernels");
        expect(plan.get(hmpc::constants::zero).shape().size() == 2 && plan.get(hmpc::constants::one).shape().size() == 2 &&
                   plan.get(hmpc::constants::two).shape().size() == 2,
               "kernels 0, 1, 2 of shape {2}");
        // and the values: (u + v, w) run
        auto [uv, ww] = queue(u + v, w);
        expect(uv.shape() == shape && ww.shape() == shape, "results of shape {2}");
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
        std::printf("cache\n");
        std::fflush(stdout);
        cache();
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
