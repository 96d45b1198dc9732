// Uses the C++ interface (include/hmpc_amd/comp.hpp) the way hmpc code uses hmpc::comp: a queue,
// tensors, NTT expressions with hmpc's rewrites, explicit-key samplers, lhe::enc / dec, ciphertext
// operators, the tensor cipher, sum and matrix_vector_product, capability samplers planned as hmpc's
// execution cache plans them (examples/bgv.cpp with a shared u), casts and BGV with an Rp plaintext. Writes
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
        std::vector<std::vector<std::uint32_t>> put_later;
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
        // hmpc's rewrite: NTT(INTT(x)) is x itself, no kernel (the top-level tensor is copied out)
        auto rewritten = number_theoretic_transform(inverse_number_theoretic_transform(tensor(a)));
        if (rewritten->kind != op::leaf)
            return fail("NTT(INTT(x)) was not rewritten to x");
        auto same = q(rewritten);
        if (same.data() == a.data() || same.to_host() != a.to_host())
            return fail("a top-level tensor is not materialised into a new tensor");
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
        // ---- execution plan: examples/bgv.cpp:80-90 with the shared capability u. The cache order is
        // u, v, (c0), w, (c1): u, v, w draw on nonces n0, n0 + kN, n0 + 2kN
        auto nonce_bgv = q.nonce();
        auto mq = comp::make_tensor(q, m, k, log_n);
        check(hmpc_from_u32(m.handle(), mq.size(), msgs.data(), mq.data(), nullptr));
        auto mh = q(number_theoretic_transform(tensor(mq)));
        {
            auto u = number_theoretic_transform(random::centered_binomial(m, 2, k, log_n));
            auto v = number_theoretic_transform(random::centered_binomial(m, 40, k, log_n));
            auto w = number_theoretic_transform(random::centered_binomial(m, 40, k, log_n));
            auto [e0, e1] = q(tensor(A) * u + v * t + tensor(mh), tensor(B) * u + w * t);
            auto after = q.nonce();
            if (after[0] != nonce_bgv[0] + 3u * (k << log_n))
                return fail("shared-u BGV expression: nonce did not advance by 3 n N");
            auto dec_e = q(crypto::lhe::dec(s, {e0, e1})).to_host();
            if (dec_e != words)
                return fail("shared-u BGV expression does not decrypt");
            put_later.push_back(e0.to_host());
            put_later.push_back(e1.to_host());
        }
        // element-wise capability samplers: top-level (its own kernel) and fused into a sum
        auto nonce_elem = q.nonce();
        auto ru = q(random::uniform(m, n, log_n));
        auto rs = q(random::centered_binomial(m, 40, n, log_n) + tensor(a));
        // under an inverse transform (pairs 2t, 2t + 1) and a drowning sampler under a forward one
        auto ri = q(inverse_number_theoretic_transform(random::centered_binomial(m, 40, n, log_n)));
        auto rd = q(number_theoretic_transform(random::drown_signed_uniform(m, {2u}, n, log_n)));
        // a sampler shared by two results: materialised once (element layout), one nonce
        auto shared = random::centered_binomial(m, 2, n, log_n);
        auto [sh0, sh1] = q(shared + tensor(a), shared * tensor(b));
        auto nonce_end = q.nonce();
        if (nonce_end[0] != nonce_elem[0] + 5u * (n << log_n))
            return fail("capability samplers: five kernels should advance the nonce by 5 n N");
        // cast and BGV with an Rp plaintext (NTT form) over the Goldilocks prime
        comp::modulus mp(q, {0x00000001u, 0xffffffffu});
        auto xp = q(random::uniform(mp, key, 30, k, log_n));
        auto xq = q(cast(tensor(xp), m));
        auto Ap = q(tensor(B) * tensor(s) + tensor(e) * m.from_unsigned(0xffffffff00000001ull));
        auto xp_ntt = q(number_theoretic_transform(tensor(xp)));
        auto nonce_plain = q.nonce();
        auto ctp = q(crypto::lhe::enc({Ap, B}, xp_ntt));
        auto back = q(crypto::lhe::dec(s, ctp, mp, comp::representation::number_theoretic_transform));
        if (back.to_host() != xp_ntt.to_host())
            return fail("dec<plaintext>(enc(x)) differs from x");
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
        put(f, std::vector<std::uint32_t>(nonce_bgv.begin(), nonce_bgv.end()));
        put(f, mh.to_host());
        for (auto const& v : put_later)
            put(f, v);
        put(f, std::vector<std::uint32_t>(nonce_elem.begin(), nonce_elem.end()));
        put(f, ru.to_host());
        put(f, rs.to_host());
        put(f, ri.to_host());
        put(f, rd.to_host());
        put(f, sh0.to_host());
        put(f, sh1.to_host());
        put(f, xp.to_host());
        put(f, xq.to_host());
        put(f, Ap.to_host());
        put(f, std::vector<std::uint32_t>(nonce_plain.begin(), nonce_plain.end()));
        put(f, ctp.c0.to_host());
        put(f, ctp.c1.to_host());
        std::printf("facade ok\n");
        return 0;
    }
    catch (error const& err)
    {
        std::fprintf(stderr, "%s\n", err.what());
        return 2;
    }
}
