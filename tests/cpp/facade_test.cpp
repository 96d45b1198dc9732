// Uses the C++ interface (include/hmpc_amd/comp.hpp) the way hmpc code uses hmpc::comp: a queue,
// tensors, NTT expressions with hmpc's rewrites, explicit-key samplers, lhe::enc / dec, ciphertext
// operators, the tensor cipher, sum and matrix_vector_product. Writes
// its results (limb-planar u32 words) to the file named by argv[1] for tests/test_gpu_facade.py,
// which recomputes them with the oracle. Exit code 0 when its own checks pass.
#include "../../include/hmpc_amd/comp.hpp"

#include <cstdio>
#include <fstream>

using namespace hmpc_amd;
using namespace hmpc_amd::expr;

namespace
{
    void put(std::ofstream& f, std::vector<std::uint32_t> const& v)
    {
        std::uint64_t n = v.size();
        f.write(reinterpret_cast<char const*>(&n), sizeof n);
        f.write(reinterpret_cast<char const*>(v.data()), static_cast<std::streamsize>(v.size() * 4));
    }

    int fail(char const* what)
    {
        std::fprintf(stderr, "facade check failed: %s\n", what);
        return 1;
    }
}

int main(int argc, char** argv)
{
    if (argc < 2)
        return fail("usage: facade_test OUTPUT");
    try
    {
        std::array<std::uint32_t, 8> key{42, 0, 0, 0, 0, 0, 0, 0};
        comp::queue q(0, key);
        // q = 2^124 + 2^20 + 2^18 + 1
        comp::modulus m(q, {0x00140001u, 0u, 0u, 0x10000000u});
        int const log_n = 12;
        std::size_t const n = 3;
        auto a = q(random::uniform(m, key, 0, n, log_n));
        auto b = q(random::uniform(m, key, 1, n, log_n));
        // fused product, then the same product node by node
        auto c = q(inverse_number_theoretic_transform(number_theoretic_transform(tensor(a)) * number_theoretic_transform(tensor(b))));
        auto fa = q(number_theoretic_transform(tensor(a)));
        auto fb = q(number_theoretic_transform(tensor(b)));
        auto p = q(tensor(fa) * tensor(fb));
        auto c2 = q(inverse_number_theoretic_transform(tensor(p)));
        if (c.to_host() != c2.to_host())
            return fail("fused product differs from the node-by-node product");
        // hmpc's rewrite: NTT(INTT(x)) is x itself, no kernel
        auto same = q(number_theoretic_transform(inverse_number_theoretic_transform(tensor(a))));
        if (same.data() != a.data())
            return fail("NTT(INTT(x)) was not rewritten to x");
        // BGV with a hand-built key, as in tests/expr/crypto/lhe/enc.cpp:35-47
        auto B = q(random::uniform(m, key, 10, 1, log_n));
        auto s = q(number_theoretic_transform(random::centered_binomial(m, key, 11, 2, 1, log_n)));
        auto e = q(number_theoretic_transform(random::centered_binomial(m, key, 12, 40, 1, log_n)));
        auto t = m.from_unsigned(std::uint64_t{1} << 32);
        auto A = q(tensor(B) * tensor(s) + tensor(e) * t);
        std::size_t const k = 2;
        auto msgs = comp::make_tensor(q, m, k, log_n, comp::representation::words);
        std::vector<std::uint32_t> words(k << log_n);
        for (std::size_t i = 0; i < words.size(); ++i)
            words[i] = static_cast<std::uint32_t>(i * 2654435761u + 12345u);
        msgs.from_host(words.data(), words.size());
        auto ct = q(crypto::lhe::enc({A, B}, msgs, t));
        auto nonce = q.nonce();
        if (nonce[0] != 3u * (k << log_n))
            return fail("capability nonce did not advance by 3 n N");
        auto plain = q(crypto::lhe::dec(s, ct));
        if (plain.to_host() != words)
            return fail("decryption differs from the messages");
        // ciphertext operators: a second encryption, the sum, 7 times the first
        auto msgs2 = comp::make_tensor(q, m, k, log_n, comp::representation::words);
        std::vector<std::uint32_t> words2(k << log_n);
        for (std::size_t i = 0; i < words2.size(); ++i)
            words2[i] = static_cast<std::uint32_t>(i * 40503u + 7u);
        msgs2.from_host(words2.data(), words2.size());
        auto ct2 = q(crypto::lhe::enc({A, B}, msgs2, t));
        auto ct_sum = q(ct + ct2);
        auto ct_7 = q(ct * m.from_unsigned(7));
        auto dec_sum = q(crypto::lhe::dec(s, ct_sum)).to_host();
        auto dec_7 = q(crypto::lhe::dec(s, ct_7)).to_host();
        for (std::size_t i = 0; i < words.size(); ++i)
            if (dec_sum[i] != words[i] + words2[i] || dec_7[i] != 7u * words[i])
                return fail("ciphertext operators do not decrypt to the plaintext operations");
        // tensor cipher round trip
        crypto::cipher ch{key, {1, 0, 0}};
        auto y = q(crypto::enc(ch, a));
        auto a_again = q(crypto::dec(ch, y));
        if (a_again.to_host() != a.to_host())
            return fail("tensor cipher round trip");
        // reductions and a matrix-vector product
        auto total = q(sum(tensor(a)));
        auto M = q(random::uniform(m, key, 20, 4, log_n));
        auto v = q(random::uniform(m, key, 21, 1, log_n));
        auto Mv = q(matrix_vector_product(tensor(M), tensor(v)));
        q.wait();
        std::ofstream f(argv[1], std::ios::binary);
        put(f, a.to_host());
        put(f, b.to_host());
        put(f, c.to_host());
        put(f, t);
        put(f, B.to_host());
        put(f, s.to_host());
        put(f, e.to_host());
        put(f, A.to_host());
        put(f, words);
        put(f, ct.c0.to_host());
        put(f, ct.c1.to_host());
        put(f, words2);
        put(f, ct_sum.c0.to_host());
        put(f, ct_sum.c1.to_host());
        put(f, y.to_host());
        put(f, total.to_host());
        put(f, Mv.to_host());
        std::printf("facade ok\n");
        return 0;
    }
    catch (error const& err)
    {
        std::fprintf(stderr, "%s\n", err.what());
        return 2;
    }
}
