// C++ host interface over libhmpc_hip.so that mirrors hmpc's hmpc::comp / hmpc::expr surface for
// the NTT and BGV paths, so that hmpc code on these paths ports by renaming:
//
//   hmpc                                              here
//   hmpc::comp::queue q{sycl_queue, key}              hmpc_amd::comp::queue q{device, key}
//     (hmpc/comp/queue.hpp:229-239)
//   mod<q> (compile-time modulus, ints/mod.hpp)       hmpc_amd::comp::modulus m{q, words}  (run time)
//   hmpc::comp::make_tensor<poly>(shape{n, N})        hmpc_amd::comp::make_tensor(q, m, n, log_n)
//     (hmpc/comp/tensor.hpp:86-90)
//   host_accessor(t, access::read / discard_write)    t.to_host() / t.from_host(...)
//     (hmpc/comp/accessor.hpp:282-288)
//   hmpc::expr::tensor(t)                             hmpc_amd::expr::tensor(t)
//   number_theoretic_transform(e), inverse_...(e)     same names (hmpc/expr/number_theoretic_transform.hpp)
//   e1 + e2, e1 - e2, e1 * e2, e * constant           same operators (hmpc/expr/binary_expression.hpp)
//   random::centered_binomial (capability)            expr::random::centered_binomial(m, bits, n, log_n)
//   random::number_generator(key, index) + uniform    expr::random::uniform(m, key, offset, n, log_n)
//   crypto::lhe::enc(key, m, randomness()) / dec      expr::crypto::lhe::enc(...) / dec(...)
//   q(exprs...) -> tensors, q.wait(), q.info()        same (hmpc/comp/queue.hpp:416-597)
//
// Evaluation keeps hmpc's algebraic rewrites (hmpc/expr/number_theoretic_transform.hpp:558-583):
// NTT(INTT(e)) and INTT(NTT(e)) are e. It fuses INTT(NTT(a) o NTT(b)) into the batched product
// kernel and NTT(capability CBD) into the fused sampler; everything else runs as one C-ABI call per
// node. Tensors share ownership of their device buffer (copying aliases, as in hmpc). Errors throw
// hmpc_amd::error (hmpc's FFI wrapper throws on non-ok codes, hmpc/net/ffi.hpp:14).
//
// Header only, C++17; link with -lhmpc_hip.
#pragma once

#include "../hmpc_hip.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

namespace hmpc_amd
{
    struct error : std::runtime_error
    {
        int code;
        error(int c, std::string const& msg) : std::runtime_error("hmpc error " + std::to_string(c) + ": " + msg), code(c) {}
    };

    inline void check(int rc)
    {
        if (rc != HMPC_OK)
            throw error(rc, hmpc_last_error());
    }

    namespace comp
    {
        class queue;

        // a registered modulus q (hmpc::ints::mod<q> constants and resident twiddle tables)
        class modulus
        {
        public:
            modulus() = default;
            modulus(queue& q, std::vector<std::uint32_t> const& q_words);

            hmpc_modulus handle() const { return state_->handle; }
            int limbs() const { return state_->limbs; }
            int bit_size() const { return state_->bit_size; }
            queue& owner() const { return *state_->owner; }
            bool operator==(modulus const& o) const { return state_ == o.state_; }

            // Montgomery form (value R mod q) of a small unsigned integer, as L host words
            std::vector<std::uint32_t> from_unsigned(std::uint64_t value) const;

        private:
            struct state
            {
                hmpc_modulus handle = nullptr;
                int limbs = 0;
                int bit_size = 0;
                queue* owner = nullptr;
                std::vector<std::uint32_t> q, r2;
                std::uint32_t nprime = 0;
                ~state()
                {
                    if (handle)
                        (void)hmpc_modulus_destroy(handle);
                }
            };
            std::shared_ptr<state> state_;
        };

        enum class representation
        {
            coefficient,
            number_theoretic_transform,
            words,         // plain u32 words (messages, decrypted plaintexts)
            cipher_words,  // tensor-cipher output: element-major u32[size][L] (hmpc's shape {..., L})
        };

        // n polynomials of N = 2^log_n coefficients mod q, limb-planar [L][n N] on the device; or
        // n N plain 32-bit words (representation::words)
        class tensor
        {
        public:
            tensor() = default;
            tensor(queue& q, modulus m, std::size_t n, int log_n, representation r);

            std::size_t size() const { return n_ << log_n_; }  // elements
            std::size_t batch() const { return n_; }
            int log_n() const { return log_n_; }
            int limbs() const { return rep_ == representation::words ? 1 : mod_.limbs(); }
            representation rep() const { return rep_; }
            modulus const& mod() const { return mod_; }
            std::uint32_t* data() const { return static_cast<std::uint32_t*>(buf_->ptr); }
            std::size_t bytes() const { return size() * static_cast<std::size_t>(limbs()) * 4; }
            explicit operator bool() const { return static_cast<bool>(buf_); }
            tensor with_rep(representation r) const
            {
                tensor t = *this;
                t.rep_ = r;
                return t;
            }

            // host_accessor(t, hmpc::access::read): limb-planar words, L * size()
            std::vector<std::uint32_t> to_host() const;
            // host_accessor(t, hmpc::access::discard_write)
            void from_host(std::uint32_t const* words, std::size_t count);

        private:
            struct buffer
            {
                hmpc_context ctx = nullptr;
                void* ptr = nullptr;
                ~buffer()
                {
                    if (ptr)
                        (void)hmpc_free(ctx, ptr);
                }
            };
            std::shared_ptr<buffer> buf_;
            modulus mod_;
            std::size_t n_ = 0;
            int log_n_ = 0;
            representation rep_ = representation::coefficient;
            hmpc_context ctx_ = nullptr;
        };
    }

    namespace expr
    {
        struct node;
        using expression = std::shared_ptr<node const>;

        enum class op
        {
            leaf,
            ntt,
            intt,
            add,
            sub,
            mul,
            scalar_mul,
            sum,            // hmpc::expr::sum (reduce with add) to one element
            product,        // hmpc::expr::product (reduce with multiply)
            matvec,         // hmpc::expr::matrix_vector_product
            uniform,        // explicit-key uniform<mod q>
            cbd,            // explicit-key centred binomial
            cbd_capability, // queue-capability centred binomial (only under an NTT)
        };

        struct node
        {
            op kind = op::leaf;
            comp::tensor value;            // leaf
            expression a, b;               // operands
            std::vector<std::uint32_t> c;  // scalar (Montgomery words)
            comp::modulus mod;             // samplers
            std::array<std::uint32_t, 8> key{};
            std::uint64_t offset = 0;
            int bits = 0;
            std::size_t n = 0;
            int log_n = 0;
        };

        inline expression make(node n) { return std::make_shared<node const>(std::move(n)); }

        inline expression tensor(comp::tensor const& t)
        {
            node n;
            n.value = t;
            return make(std::move(n));
        }

        // hmpc's rewrite NTT(INTT(e)) -> e (hmpc/expr/number_theoretic_transform.hpp:558-562)
        inline expression number_theoretic_transform(expression e)
        {
            if (e->kind == op::intt)
                return e->a;
            node n;
            n.kind = op::ntt;
            n.a = std::move(e);
            return make(std::move(n));
        }

        // INTT(NTT(e)) -> e (:579-583)
        inline expression inverse_number_theoretic_transform(expression e)
        {
            if (e->kind == op::ntt)
                return e->a;
            node n;
            n.kind = op::intt;
            n.a = std::move(e);
            return make(std::move(n));
        }

        inline expression binary(op k, expression x, expression y)
        {
            node n;
            n.kind = k;
            n.a = std::move(x);
            n.b = std::move(y);
            return make(std::move(n));
        }

        inline expression operator+(expression x, expression y) { return binary(op::add, std::move(x), std::move(y)); }
        inline expression operator-(expression x, expression y) { return binary(op::sub, std::move(x), std::move(y)); }
        inline expression operator*(expression x, expression y) { return binary(op::mul, std::move(x), std::move(y)); }

        // e * constant (Montgomery words), hmpc's value / constant leaves (hmpc/expr/value.hpp)
        inline expression operator*(expression x, std::vector<std::uint32_t> constant_mont)
        {
            node n;
            n.kind = op::scalar_mul;
            n.a = std::move(x);
            n.c = std::move(constant_mont);
            return make(std::move(n));
        }

        // hmpc/expr/reduce.hpp:264-274: sum / product of every element, a one-element tensor
        inline expression sum(expression x)
        {
            node n;
            n.kind = op::sum;
            n.a = std::move(x);
            return make(std::move(n));
        }
        inline expression product(expression x)
        {
            node n;
            n.kind = op::product;
            n.a = std::move(x);
            return make(std::move(n));
        }

        // hmpc/expr/matrix_vector_product.hpp: M with `batch()` rows of N = cols coefficients, v one
        // row of cols; the result holds one element per row of M (rows a power of two)
        inline expression matrix_vector_product(expression matrix, expression vector)
        {
            return binary(op::matvec, std::move(matrix), std::move(vector));
        }

        namespace random
        {
            // explicit-key uniform<mod q> (hmpc/expr/random/uniform_from_number_generator.hpp)
            inline expression uniform(comp::modulus m, std::array<std::uint32_t, 8> key, std::uint64_t offset, std::size_t n,
                                      int log_n)
            {
                node e;
                e.kind = op::uniform;
                e.mod = std::move(m);
                e.key = key;
                e.offset = offset;
                e.n = n;
                e.log_n = log_n;
                return make(std::move(e));
            }

            // explicit-key centred binomial with `bits` = 4 x variance
            inline expression centered_binomial(comp::modulus m, std::array<std::uint32_t, 8> key, std::uint64_t offset, int bits,
                                                std::size_t n, int log_n)
            {
                node e;
                e.kind = op::cbd;
                e.mod = std::move(m);
                e.key = key;
                e.offset = offset;
                e.bits = bits;
                e.n = n;
                e.log_n = log_n;
                return make(std::move(e));
            }

            // queue-capability centred binomial (hmpc/expr/random/binomial.hpp); consumed by an NTT
            inline expression centered_binomial(comp::modulus m, int bits, std::size_t n, int log_n)
            {
                node e;
                e.kind = op::cbd_capability;
                e.mod = std::move(m);
                e.bits = bits;
                e.n = n;
                e.log_n = log_n;
                return make(std::move(e));
            }
        }

        namespace crypto::lhe
        {
            // hmpc::comp::crypto::lhe::key: A = B s + t e and B, both in NTT form, shape {} (one polynomial)
            struct key
            {
                comp::tensor A, B;
            };
            struct ciphertext
            {
                comp::tensor c0, c1;
            };
            // default randomness of hmpc/expr/crypto/lhe/randomness.hpp:80 (variances 1/2, 10, 10)
            struct randomness
            {
                int bits_u = 2, bits_v = 40, bits_w = 40;
            };
            struct enc_expression
            {
                key k;
                comp::tensor messages;  // representation::words, n N
                std::vector<std::uint32_t> plaintext_modulus;  // t as a Montgomery constant
                randomness r;
            };
            struct dec_expression
            {
                comp::tensor s;  // secret in NTT form
                ciphertext ct;
            };
            inline enc_expression enc(key k, comp::tensor messages, std::vector<std::uint32_t> t_mont, randomness r = {})
            {
                return {std::move(k), std::move(messages), std::move(t_mont), r};
            }
            inline dec_expression dec(comp::tensor s, ciphertext ct) { return {std::move(s), std::move(ct)}; }

            // ciphertext_expression operators (hmpc/expr/crypto/lhe/ciphertext.hpp:48-81): with another
            // ciphertext (+, -), a plaintext polynomial in NTT form broadcast over the batch (*), or a
            // Montgomery constant (*)
            struct ciphertext_expression
            {
                int op = HMPC_ADD;
                ciphertext a, b;
                comp::tensor plain;
                std::vector<std::uint32_t> constant;
            };
            inline ciphertext_expression operator+(ciphertext a, ciphertext b) { return {HMPC_ADD, std::move(a), std::move(b), {}, {}}; }
            inline ciphertext_expression operator-(ciphertext a, ciphertext b) { return {HMPC_SUB, std::move(a), std::move(b), {}, {}}; }
            inline ciphertext_expression operator*(ciphertext a, comp::tensor plain)
            {
                return {HMPC_MUL, std::move(a), {}, std::move(plain), {}};
            }
            inline ciphertext_expression operator*(ciphertext a, std::vector<std::uint32_t> constant_mont)
            {
                return {HMPC_MUL, std::move(a), {}, {}, std::move(constant_mont)};
            }
        }

        // hmpc::expr::crypto::cipher / enc / dec (hmpc/expr/crypto/cipher.hpp): ChaCha20 keystream XOR
        namespace crypto
        {
            struct cipher
            {
                std::array<std::uint32_t, 8> key{};
                std::array<std::uint32_t, 3> nonce{};
            };
            struct cipher_enc_expression
            {
                cipher c;
                comp::tensor value;
            };
            struct cipher_dec_expression
            {
                cipher c;
                comp::tensor words;
                comp::representation rep;
            };
            inline cipher_enc_expression enc(cipher c, comp::tensor value) { return {c, std::move(value)}; }
            // dec<T>: the representation of the decrypted tensor stands for hmpc's value type T
            inline cipher_dec_expression dec(cipher c, comp::tensor words,
                                             comp::representation rep = comp::representation::coefficient)
            {
                return {c, std::move(words), rep};
            }
        }
    }

    namespace comp
    {
        class queue
        {
        public:
            // hmpc/comp/queue.hpp:229-233: key from the OS entropy source
            explicit queue(int device = 0) { check(hmpc_context_create(device, nullptr, &ctx_)); }
            // hmpc/comp/queue.hpp:235-239: explicit 8-word ChaCha20 key, nonce 0
            queue(int device, std::array<std::uint32_t, 8> const& key) { check(hmpc_context_create(device, key.data(), &ctx_)); }
            queue(queue const&) = delete;
            queue& operator=(queue const&) = delete;
            ~queue()
            {
                if (ctx_)
                    (void)hmpc_context_destroy(ctx_);
            }

            hmpc_context handle() const { return ctx_; }
            void wait() { check(hmpc_context_wait(ctx_)); }
            hmpc_device_info info() const
            {
                hmpc_device_info d{};
                check(hmpc_context_info(ctx_, &d));
                return d;
            }
            std::array<std::uint32_t, 3> nonce() const
            {
                std::array<std::uint32_t, 8> k{};
                std::array<std::uint32_t, 3> n{};
                check(hmpc_context_get_rng(ctx_, k.data(), n.data()));
                return n;
            }

            // q(expr): evaluates one expression into a new tensor
            tensor operator()(expr::expression const& e) { return eval(e); }
            // q(enc(...)): one BGV encryption call; the capability nonce advances by 3 n N
            expr::crypto::lhe::ciphertext operator()(expr::crypto::lhe::enc_expression const& e)
            {
                auto const& m = e.k.A.mod();
                tensor c0(*this, m, e.messages.batch(), e.messages.log_n(), representation::number_theoretic_transform);
                tensor c1(*this, m, e.messages.batch(), e.messages.log_n(), representation::number_theoretic_transform);
                check(hmpc_bgv_encrypt(ctx_, m.handle(), e.messages.log_n(), e.messages.batch(), e.k.A.data(), e.k.B.data(),
                                       e.messages.data(), e.plaintext_modulus.data(), e.r.bits_u, e.r.bits_v, e.r.bits_w,
                                       c0.data(), c1.data(), nullptr));
                return {c0, c1};
            }
            // q(ct op ...): a new ciphertext, both parts in one launch
            expr::crypto::lhe::ciphertext operator()(expr::crypto::lhe::ciphertext_expression const& e)
            {
                auto const& m = e.a.c0.mod();
                std::size_t count = e.a.c0.size();
                tensor c0(*this, m, e.a.c0.batch(), e.a.c0.log_n(), representation::number_theoretic_transform);
                tensor c1(*this, m, e.a.c0.batch(), e.a.c0.log_n(), representation::number_theoretic_transform);
                if (!e.constant.empty())
                    check(hmpc_ciphertext_scalar(m.handle(), e.op, count, e.a.c0.data(), e.a.c1.data(), e.constant.data(),
                                                 c0.data(), c1.data(), nullptr));
                else if (e.plain)
                    check(hmpc_ciphertext_op(m.handle(), e.op, count, e.a.c0.data(), e.a.c1.data(), e.plain.data(),
                                             e.plain.data(), e.plain.size(), c0.data(), c1.data(), nullptr));
                else
                    check(hmpc_ciphertext_op(m.handle(), e.op, count, e.a.c0.data(), e.a.c1.data(), e.b.c0.data(),
                                             e.b.c1.data(), count, c0.data(), c1.data(), nullptr));
                return {c0, c1};
            }
            // q(enc(cipher, x)): element-major cipher words
            tensor operator()(expr::crypto::cipher_enc_expression const& e)
            {
                auto const& x = e.value;
                tensor t(*this, x.mod(), x.batch(), x.log_n(), representation::cipher_words);
                check(hmpc_cipher_encrypt(ctx_, x.limbs(), x.size(), e.c.key.data(), e.c.nonce.data(), 0, x.data(), t.data(),
                                          nullptr));
                return t;
            }
            // q(dec<T>(cipher, y)): the limb-planar tensor back
            tensor operator()(expr::crypto::cipher_dec_expression const& e)
            {
                auto const& y = e.words;
                tensor t(*this, y.mod(), y.batch(), y.log_n(), e.rep);
                check(hmpc_cipher_decrypt(ctx_, y.limbs(), y.size(), e.c.key.data(), e.c.nonce.data(), 0, y.data(), t.data(),
                                          nullptr));
                return t;
            }
            // q(dec(...)): plaintext words
            tensor operator()(expr::crypto::lhe::dec_expression const& e)
            {
                auto const& m = e.ct.c0.mod();
                tensor out(*this, m, e.ct.c0.batch(), e.ct.c0.log_n(), representation::words);
                check(hmpc_bgv_decrypt(m.handle(), e.ct.c0.log_n(), e.ct.c0.batch(), e.s.data(), e.ct.c0.data(),
                                       e.ct.c1.data(), out.data(), nullptr));
                return out;
            }
            // several expressions, evaluated in argument order (hmpc's cache order for independent roots)
            template<typename... E>
            auto operator()(E const&... es) -> std::enable_if_t<(sizeof...(E) > 1), std::tuple<decltype((*this)(es))...>>
            {
                return std::tuple<decltype((*this)(es))...>{(*this)(es)...};
            }

        private:
            hmpc_context ctx_ = nullptr;

            tensor eval(expr::expression const& e);
        };

        inline modulus::modulus(queue& q, std::vector<std::uint32_t> const& q_words) : state_(std::make_shared<state>())
        {
            state_->owner = &q;
            check(hmpc_modulus_create(q.handle(), q_words.data(), static_cast<int>(q_words.size()), &state_->handle));
            check(hmpc_modulus_info(state_->handle, &state_->limbs, &state_->bit_size, nullptr, nullptr, nullptr));
            auto L = static_cast<std::size_t>(state_->limbs);
            state_->q.assign(L, 0);
            for (std::size_t i = 0; i < L && i < q_words.size(); ++i)
                state_->q[i] = q_words[i];
            state_->r2.assign(L, 0);
            std::vector<std::uint32_t> r1(L);
            check(hmpc_modulus_info(state_->handle, nullptr, nullptr, &state_->nprime, r1.data(), state_->r2.data()));
        }

        // value R mod q = REDC(value * (R^2 mod q)), host CIOS Montgomery product (hmpc/ints/mod.hpp:165-173)
        inline std::vector<std::uint32_t> modulus::from_unsigned(std::uint64_t value) const
        {
            int const L = limbs();
            std::vector<std::uint32_t> const& q = state_->q;
            std::vector<std::uint32_t> a(static_cast<std::size_t>(L), 0), t(static_cast<std::size_t>(L) + 2, 0);
            a[0] = static_cast<std::uint32_t>(value);
            if (L > 1)
                a[1] = static_cast<std::uint32_t>(value >> 32);
            for (int i = 0; i < L; ++i)
            {
                std::uint64_t c = 0;
                for (int j = 0; j < L; ++j)
                {
                    std::uint64_t s = static_cast<std::uint64_t>(a[static_cast<std::size_t>(j)]) * state_->r2[static_cast<std::size_t>(i)] +
                                      t[static_cast<std::size_t>(j)] + c;
                    t[static_cast<std::size_t>(j)] = static_cast<std::uint32_t>(s);
                    c = s >> 32;
                }
                std::uint64_t s = static_cast<std::uint64_t>(t[static_cast<std::size_t>(L)]) + c;
                t[static_cast<std::size_t>(L)] = static_cast<std::uint32_t>(s);
                t[static_cast<std::size_t>(L) + 1] = static_cast<std::uint32_t>(s >> 32);
                std::uint32_t m = t[0] * state_->nprime;
                c = (static_cast<std::uint64_t>(m) * q[0] + t[0]) >> 32;
                for (int j = 1; j < L; ++j)
                {
                    std::uint64_t v = static_cast<std::uint64_t>(m) * q[static_cast<std::size_t>(j)] + t[static_cast<std::size_t>(j)] + c;
                    t[static_cast<std::size_t>(j) - 1] = static_cast<std::uint32_t>(v);
                    c = v >> 32;
                }
                s = static_cast<std::uint64_t>(t[static_cast<std::size_t>(L)]) + c;
                t[static_cast<std::size_t>(L) - 1] = static_cast<std::uint32_t>(s);
                t[static_cast<std::size_t>(L)] = t[static_cast<std::size_t>(L) + 1] + static_cast<std::uint32_t>(s >> 32);
            }
            // t < 2q: subtract q once if t >= q
            std::vector<std::uint32_t> d(static_cast<std::size_t>(L));
            std::uint64_t borrow = 0;
            for (int j = 0; j < L; ++j)
            {
                std::uint64_t v = static_cast<std::uint64_t>(t[static_cast<std::size_t>(j)]) - q[static_cast<std::size_t>(j)] - borrow;
                d[static_cast<std::size_t>(j)] = static_cast<std::uint32_t>(v);
                borrow = (v >> 63) & 1u;
            }
            bool take = t[static_cast<std::size_t>(L)] != 0 || borrow == 0;
            t.resize(static_cast<std::size_t>(L));
            return take ? d : t;
        }

        inline tensor::tensor(queue& q, modulus m, std::size_t n, int log_n, representation r)
            : buf_(std::make_shared<buffer>()), mod_(std::move(m)), n_(n), log_n_(log_n), rep_(r), ctx_(q.handle())
        {
            buf_->ctx = ctx_;
            check(hmpc_malloc(ctx_, std::max<std::size_t>(bytes(), 4), &buf_->ptr));
        }

        inline std::vector<std::uint32_t> tensor::to_host() const
        {
            std::vector<std::uint32_t> out(size() * static_cast<std::size_t>(limbs()));
            check(hmpc_memcpy_d2h(ctx_, out.data(), buf_->ptr, out.size() * 4));
            return out;
        }

        inline void tensor::from_host(std::uint32_t const* words, std::size_t count)
        {
            if (count != size() * static_cast<std::size_t>(limbs()))
                throw error(HMPC_ERR_INVALID_ARGUMENT, "from_host: expected L * size() words");
            check(hmpc_memcpy_h2d(ctx_, buf_->ptr, words, count * 4));
        }

        inline tensor make_tensor(queue& q, modulus const& m, std::size_t n, int log_n,
                                  representation r = representation::coefficient)
        {
            return tensor(q, m, n, log_n, r);
        }

        inline tensor queue::eval(expr::expression const& e)
        {
            using expr::op;
            switch (e->kind)
            {
            case op::leaf:
                return e->value;
            case op::uniform:
            {
                tensor t(*this, e->mod, e->n, e->log_n, representation::coefficient);
                check(hmpc_sample_uniform(e->mod.handle(), e->key.data(), e->offset, t.size(), t.data(), nullptr));
                return t;
            }
            case op::cbd:
            {
                tensor t(*this, e->mod, e->n, e->log_n, representation::coefficient);
                check(hmpc_sample_cbd(e->mod.handle(), e->key.data(), e->offset, e->bits, t.size(), t.data(), nullptr));
                return t;
            }
            case op::cbd_capability:
                throw error(HMPC_ERR_INVALID_ARGUMENT, "capability samples are produced under a number_theoretic_transform");
            case op::ntt:
            {
                if (e->a->kind == op::cbd_capability)
                {
                    auto const& s = *e->a;
                    tensor t(*this, s.mod, s.n, s.log_n, representation::number_theoretic_transform);
                    check(hmpc_sample_cbd_ntt(ctx_, s.mod.handle(), s.bits, s.log_n, s.n, t.data(), nullptr));
                    return t;
                }
                tensor x = eval(e->a);
                tensor t(*this, x.mod(), x.batch(), x.log_n(), representation::number_theoretic_transform);
                check(hmpc_ntt(x.mod().handle(), x.log_n(), x.batch(), 0, x.data(), t.data(), nullptr));
                return t;
            }
            case op::intt:
            {
                // fused batched product: INTT(NTT(a) o NTT(b))
                auto const& inner = *e->a;
                if (inner.kind == op::mul && inner.a->kind == op::ntt && inner.b->kind == op::ntt)
                {
                    tensor a = eval(inner.a->a);
                    tensor b = eval(inner.b->a);
                    if (a.mod() == b.mod() && a.size() == b.size())
                    {
                        tensor t(*this, a.mod(), a.batch(), a.log_n(), representation::coefficient);
                        check(hmpc_polymul(a.mod().handle(), a.log_n(), a.batch(), a.data(), b.data(), t.data(), nullptr));
                        return t;
                    }
                }
                tensor x = eval(e->a);
                tensor t(*this, x.mod(), x.batch(), x.log_n(), representation::coefficient);
                check(hmpc_ntt(x.mod().handle(), x.log_n(), x.batch(), 1, x.data(), t.data(), nullptr));
                return t;
            }
            case op::add:
            case op::sub:
            case op::mul:
            {
                // a * b + c in one kernel, as hmpc fuses an element-wise tree into one (queue.hpp:387-412)
                if (e->kind == op::add && e->a->kind == op::mul)
                {
                    tensor x = eval(e->a->a);
                    tensor y = eval(e->a->b);
                    tensor z = eval(e->b);
                    if (x.mod() == y.mod() && x.mod() == z.mod() && x.size() == y.size() && x.size() == z.size())
                    {
                        tensor t(*this, x.mod(), x.batch(), x.log_n(), x.rep());
                        check(hmpc_fma(x.mod().handle(), x.size(), x.data(), y.data(), z.data(), t.data(), nullptr));
                        return t;
                    }
                }
                tensor a = eval(e->a);
                tensor b = eval(e->b);
                if (!(a.mod() == b.mod()) || b.size() == 0 || a.size() % b.size() != 0)
                    throw error(HMPC_ERR_INVALID_ARGUMENT, "operands need one modulus and a broadcastable shape");
                int o = e->kind == op::add ? HMPC_ADD : (e->kind == op::sub ? HMPC_SUB : HMPC_MUL);
                tensor t(*this, a.mod(), a.batch(), a.log_n(), a.rep());
                check(hmpc_pointwise(a.mod().handle(), o, a.size(), a.data(), b.data(), b.size(), t.data(), nullptr));
                return t;
            }
            case op::scalar_mul:
            {
                tensor a = eval(e->a);
                tensor t(*this, a.mod(), a.batch(), a.log_n(), a.rep());
                check(hmpc_pointwise_scalar(a.mod().handle(), HMPC_MUL, a.size(), a.data(), e->c.data(), t.data(), nullptr));
                return t;
            }
            case op::sum:
            case op::product:
            {
                tensor a = eval(e->a);
                tensor t(*this, a.mod(), 1, 0, a.rep());
                check(hmpc_reduce(a.mod().handle(), e->kind == op::sum ? HMPC_ADD : HMPC_MUL, a.size(), a.data(), t.data(),
                                  nullptr));
                return t;
            }
            case op::matvec:
            {
                tensor m = eval(e->a);
                tensor v = eval(e->b);
                std::size_t rows = m.batch(), cols = m.size() / m.batch();
                int log_rows = 0;
                while ((std::size_t{1} << log_rows) < rows)
                    ++log_rows;
                if (!(m.mod() == v.mod()) || v.size() != cols || (std::size_t{1} << log_rows) != rows)
                    throw error(HMPC_ERR_INVALID_ARGUMENT, "matrix_vector_product: one modulus, v of M's row length, 2^k rows");
                tensor t(*this, m.mod(), 1, log_rows, m.rep());
                check(hmpc_matvec(m.mod().handle(), 1, rows, cols, m.data(), 0, v.data(), 0, t.data(), nullptr));
                return t;
            }
            }
            throw error(HMPC_ERR_INVALID_ARGUMENT, "unknown expression");
        }
    }
}
