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
//   random::{uniform, centered_binomial, binomial,    expr::random::{uniform, centered_binomial, binomial,
//     drown_signed/unsigned_uniform} (capability)       drown_signed/unsigned_uniform}(m, [bits,] n, log_n)
//   random::number_generator(key, index) + uniform    expr::random::uniform(m, key, offset, n, log_n)
//   cast<T>(e) across moduli                          expr::cast(e, modulus)  (centred lift)
//   crypto::lhe::enc(key, x, randomness()) / dec<T>   expr::crypto::lhe::enc(...) / dec(...)
//   q(exprs...) -> tensors, q.wait(), q.info()        same (hmpc/comp/queue.hpp:416-597)
//
// Evaluation keeps hmpc's algebraic rewrites (hmpc/expr/number_theoretic_transform.hpp:558-583):
// NTT(INTT(e)) and INTT(NTT(e)) are e. q(e1, e2, ...) plans its kernels as hmpc's execution cache
// does (hmpc/expr/cache.hpp:85-253): a cacheable node (transform, reduction, capability sampler) is
// materialised when it is complex, top-level, reached on more than one path or through a broadcast;
// every other node is fused into the kernel of its nearest materialised ancestor; kernels run in
// hmpc's post-order over the arguments. Capability samplers therefore draw on the same nonces and
// in the same stream layouts as in hmpc (element-wise: one stream per element; under a forward or
// inverse transform: the stage-0 pair layouts). Each node is evaluated once per call. INTT(NTT(a) o
// NTT(b)) runs as the batched product kernel and NTT(capability CBD) as the fused sampler, with the
// same results. Tensors share ownership of their device buffer (copying aliases, as in hmpc).
// Errors throw hmpc_amd::error (hmpc's FFI wrapper throws on non-ok codes, hmpc/net/ffi.hpp:14).
//
// Header only, C++17; link with -lhmpc_hip.
#pragma once

#include "../hmpc_hip.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
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
            capability,     // queue-capability sampler (distribution `dist`, width `bits`)
            cast,           // cast to the modulus `mod` (centred lift)
            cache,          // the wrapper of a top-level expression that is not cacheable itself
        };

        struct node
        {
            op kind = op::leaf;
            comp::tensor value;            // leaf
            expression a, b;               // operands
            std::vector<std::uint32_t> c;  // scalar (Montgomery words)
            comp::modulus mod;             // samplers, cast target
            std::array<std::uint32_t, 8> key{};
            std::uint64_t offset = 0;
            int bits = 0;
            int dist = HMPC_CENTERED_BINOMIAL;
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

        // hmpc::expr::cast<mod<Q>> of a mod<P> expression (hmpc/expr/cast.hpp): the centred lift of
        // hmpc/ints/mod.hpp:259-307
        inline expression cast(expression x, comp::modulus to)
        {
            node n;
            n.kind = op::cast;
            n.a = std::move(x);
            n.mod = std::move(to);
            return make(std::move(n));
        }

        // hmpc::expr::cache(e) (hmpc/expr/cache.hpp:71-82): e may be materialised on its own; the plan
        // decides (an entry when top-level, reached on several paths or through a broadcast)
        inline expression cache(expression x)
        {
            node n;
            n.kind = op::cache;
            n.a = std::move(x);
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

            // queue-capability samplers (hmpc/expr/random/uniform.hpp, hmpc/expr/random/binomial.hpp):
            // n polynomials of N = 2^log_n coefficients (log_n = 0: n elements)
            inline expression capability(comp::modulus m, int dist, int bits, std::size_t n, int log_n)
            {
                node e;
                e.kind = op::capability;
                e.mod = std::move(m);
                e.dist = dist;
                e.bits = bits;
                e.n = n;
                e.log_n = log_n;
                return make(std::move(e));
            }
            inline expression uniform(comp::modulus m, std::size_t n, int log_n)
            {
                return capability(std::move(m), HMPC_UNIFORM, 0, n, log_n);
            }
            // bits = 4 x variance
            inline expression centered_binomial(comp::modulus m, int bits, std::size_t n, int log_n)
            {
                return capability(std::move(m), HMPC_CENTERED_BINOMIAL, bits, n, log_n);
            }
            // popcount of `count` random bits
            inline expression binomial(comp::modulus m, int count, std::size_t n, int log_n)
            {
                return capability(std::move(m), HMPC_BINOMIAL, count, n, log_n);
            }
            // drown_uniform of a bound given as words: width bit_width((bound - 1) << security) + signed
            inline int drown_bits(std::vector<std::uint32_t> const& bound, bool is_signed, int security = 80)
            {
                int b = 0;
                check(hmpc_drown_bits(bound.data(), static_cast<int>(bound.size()), is_signed ? 1 : 0, security, &b));
                return b;
            }
            inline expression drown_signed_uniform(comp::modulus m, std::vector<std::uint32_t> const& bound, std::size_t n,
                                                   int log_n, int security = 80)
            {
                return capability(std::move(m), HMPC_DROWN_SIGNED, drown_bits(bound, true, security), n, log_n);
            }
            inline expression drown_unsigned_uniform(comp::modulus m, std::vector<std::uint32_t> const& bound, std::size_t n,
                                                     int log_n, int security = 80)
            {
                return capability(std::move(m), HMPC_DROWN_UNSIGNED, drown_bits(bound, false, security), n, log_n);
            }
        }

        namespace crypto::lhe
        {
            // hmpc::comp::crypto::lhe::key: A = B s + p e and B, both in NTT form, shape {} (one polynomial)
            struct key
            {
                comp::tensor A, B;
            };
            struct ciphertext
            {
                comp::tensor c0, c1;
            };
            // hmpc/expr/crypto/lhe/randomness.hpp: randomness() draws u, v, w centred binomial with
            // variances 1/2, 10, 10 (bits 2, 40, 40, :80-110); drowning_randomness(bound) draws v as a
            // signed drowning uniform (:112-143)
            struct randomness
            {
                int bits_u = 2, bits_v = 40, bits_w = 40;
                int v_distribution = HMPC_CENTERED_BINOMIAL;
            };
            inline randomness drowning_randomness(std::vector<std::uint32_t> const& bound, int security = 80)
            {
                randomness r;
                r.v_distribution = HMPC_DROWN_SIGNED;
                r.bits_v = random::drown_bits(bound, true, security);
                return r;
            }
            // 32-bit messages (representation::words) with p = t as a Montgomery constant, or a
            // plaintext tensor over its own modulus p (coefficient or NTT representation)
            struct enc_expression
            {
                key k;
                comp::tensor messages;
                std::vector<std::uint32_t> plaintext_modulus;  // words path: t as a Montgomery constant
                randomness r;
            };
            struct dec_expression
            {
                comp::tensor s;  // secret in NTT form
                ciphertext ct;
                comp::modulus plaintext_modulus;  // unset: 32-bit words (t = 2^32)
                comp::representation rep = comp::representation::words;
            };
            inline enc_expression enc(key k, comp::tensor messages, std::vector<std::uint32_t> t_mont, randomness r = {})
            {
                return {std::move(k), std::move(messages), std::move(t_mont), r};
            }
            // hmpc/expr/crypto/lhe/enc.hpp:19-69: a plaintext tensor over mod p
            inline enc_expression enc(key k, comp::tensor plaintext, randomness r = {})
            {
                return {std::move(k), std::move(plaintext), {}, r};
            }
            inline dec_expression dec(comp::tensor s, ciphertext ct) { return {std::move(s), std::move(ct), {}, comp::representation::words}; }
            // dec<plaintext> (hmpc/expr/crypto/lhe/dec.hpp:13-45): the plaintext's modulus and representation
            inline dec_expression dec(comp::tensor s, ciphertext ct, comp::modulus p, comp::representation rep)
            {
                return {std::move(s), std::move(ct), std::move(p), rep};
            }

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

            // q(expr): one expression into a new tensor
            tensor operator()(expr::expression const& e) { return std::move(eval({e})[0]); }
            // q(e1, e2, ...) over a list known at run time: one execution plan, one tensor per root
            std::vector<tensor> evaluate(std::vector<expr::expression> const& roots) { return eval(roots); }
            // hmpc::expr::generate_execution_cache(e1, e2, ...) (hmpc/expr/cache.hpp:233-253): the
            // kernels q(e1, e2, ...) would run, in order, without running them
            struct cache_entry
            {
                expr::op kind;
                std::size_t n;  // polynomials (or elements)
                int log_n;
            };
            std::vector<cache_entry> execution_cache(std::vector<expr::expression> const& roots);
            // q(e1, e2, ...): one execution plan over all of them (shared nodes evaluated once, kernels in
            // hmpc's cache order), a tuple of new tensors
            template<typename... E>
            auto operator()(E const&... es)
                -> std::enable_if_t<(sizeof...(E) > 1) && (std::is_convertible_v<E const&, expr::expression> && ...),
                                    std::tuple<decltype((void)std::declval<E>(), tensor())...>>
            {
                std::vector<tensor> out = eval({expr::expression(es)...});
                return unpack<sizeof...(E)>(out, std::make_index_sequence<sizeof...(E)>{});
            }
            // q(enc(...)): one BGV encryption call; the capability nonce advances by 3 n N
            expr::crypto::lhe::ciphertext operator()(expr::crypto::lhe::enc_expression const& e)
            {
                auto const& m = e.k.A.mod();
                tensor c0(*this, m, e.messages.batch(), e.messages.log_n(), representation::number_theoretic_transform);
                tensor c1(*this, m, e.messages.batch(), e.messages.log_n(), representation::number_theoretic_transform);
                if (e.messages.rep() == representation::words)
                    check(hmpc_bgv_encrypt(ctx_, m.handle(), e.messages.log_n(), e.messages.batch(), e.k.A.data(), e.k.B.data(),
                                           e.messages.data(), e.plaintext_modulus.data(), e.r.bits_u, e.r.bits_v, e.r.bits_w,
                                           c0.data(), c1.data(), nullptr));
                else
                {
                    if (e.r.v_distribution == HMPC_CENTERED_BINOMIAL && e.messages.mod() == m)
                        throw error(HMPC_ERR_INVALID_ARGUMENT, "enc: the plaintext needs its own modulus p");
                    bool ntt_form = e.messages.rep() == representation::number_theoretic_transform;
                    check(hmpc_bgv_encrypt_plaintext(ctx_, m.handle(), e.messages.mod().handle(), e.messages.log_n(),
                                                     e.messages.batch(), e.k.A.data(), e.k.B.data(), e.messages.data(),
                                                     ntt_form ? 1 : 0, e.r.bits_u, e.r.v_distribution, e.r.bits_v, e.r.bits_w,
                                                     c0.data(), c1.data(), nullptr));
                }
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
            // q(dec(...)): plaintext words (t = 2^32), or dec<plaintext> over the plaintext's modulus
            tensor operator()(expr::crypto::lhe::dec_expression const& e)
            {
                auto const& m = e.ct.c0.mod();
                if (e.rep == representation::words)
                {
                    tensor out(*this, m, e.ct.c0.batch(), e.ct.c0.log_n(), representation::words);
                    check(hmpc_bgv_decrypt(m.handle(), e.ct.c0.log_n(), e.ct.c0.batch(), e.s.data(), e.ct.c0.data(),
                                           e.ct.c1.data(), out.data(), nullptr));
                    return out;
                }
                tensor out(*this, e.plaintext_modulus, e.ct.c0.batch(), e.ct.c0.log_n(), e.rep);
                check(hmpc_bgv_decrypt_plaintext(m.handle(), e.plaintext_modulus.handle(), e.ct.c0.log_n(), e.ct.c0.batch(),
                                                 e.s.data(), e.ct.c0.data(), e.ct.c1.data(),
                                                 e.rep == representation::number_theoretic_transform ? 1 : 0, out.data(),
                                                 nullptr));
                return out;
            }

        private:
            hmpc_context ctx_ = nullptr;

            std::vector<tensor> eval(std::vector<expr::expression> const& roots);

            template<std::size_t K, std::size_t... I>
            static auto unpack(std::vector<tensor>& v, std::index_sequence<I...>)
            {
                return std::make_tuple(std::move(v[I])...);
            }
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


        // ---------------------------------------------------------------- execution plan
        // A restatement of hmpc's execution cache (hmpc/expr/cache.hpp:85-253) over expression nodes
        // (node identity stands for hmpc's expression type tags):
        //   trace:  for every cacheable node, the distinct paths from its nearest cacheable ancestor
        //           (cache.hpp:86-163); a path through a broadcast (operands of different sizes) is a
        //           "multiple" access (binary_expression.hpp access())
        //   cache:  post-order over the top-level expressions in argument order; a cacheable node is
        //           an entry when it is top-level, complex, traced more than once or through a broadcast,
        //           else it is fused into its ancestor's kernel (cache.hpp:187-231)
        //   run:    entries in order, one kernel each; a capability sampler fused into a kernel draws
        //           with that kernel's nonce (the queue nonce then advances by the kernel's size,
        //           queue.hpp:242-253) in its layout: element-wise, forward-NTT pairs, inverse-NTT pairs
        namespace detail
        {
            using expr::node;
            using expr::op;

            struct path
            {
                std::vector<std::pair<node const*, int>> steps;
                bool multiple = false;
                bool operator<(path const& o) const
                {
                    return std::tie(steps, multiple) < std::tie(o.steps, o.multiple);
                }
            };

            struct shape_info
            {
                modulus mod;
                std::size_t n = 0;
                int log_n = 0;
                representation rep = representation::coefficient;
                std::size_t size() const { return n << log_n; }
            };

            enum class layout
            {
                element = HMPC_LAYOUT_ELEMENT,
                ntt = HMPC_LAYOUT_NTT,
                intt = HMPC_LAYOUT_INTT,
            };

            inline bool cacheable(node const* e)
            {
                switch (e->kind)
                {
                case op::ntt:
                case op::intt:
                case op::sum:
                case op::product:
                case op::capability:
                case op::cache:
                    return true;
                default:
                    return false;
                }
            }

            inline bool complex(node const* e)
            {
                return e->kind == op::ntt || e->kind == op::intt || e->kind == op::sum || e->kind == op::product;
            }

            inline int arity(node const* e)
            {
                if (e->b)
                    return 2;
                return e->a ? 1 : 0;
            }

            inline node const* child(node const* e, int i) { return (i == 0 ? e->a : e->b).get(); }

            class plan
            {
            public:
                explicit plan(queue& q) : q_(q) {}

                std::vector<tensor> run(std::vector<expr::expression> const& roots)
                {
                    std::vector<node const*> tops;
                    for (auto const& r : roots)
                        tops.push_back(top_level(r));
                    for (node const* t : tops)
                        find(t);
                    for (node const* t : tops)
                        cache(t, true);
                    mark_fused_products(tops);
                    for (node const* e : order_)
                        if (!elided_.count(e))
                            results_.emplace(e, kernel(e));
                    std::vector<tensor> out;
                    for (node const* t : tops)
                        out.push_back(results_.at(t));
                    return out;
                }

                std::vector<node const*> const& order() const { return order_; }

                // hmpc's execution cache of the roots (cache.hpp:233-253) without running it: the
                // entries in execution order with their kinds and shapes (before this library's fused
                // products, which hmpc does not have)
                std::vector<std::pair<op, shape_info>> analyze(std::vector<expr::expression> const& roots)
                {
                    std::vector<node const*> tops;
                    for (auto const& r : roots)
                        tops.push_back(top_level(r));
                    for (node const* t : tops)
                        find(t);
                    for (node const* t : tops)
                        cache(t, true);
                    std::vector<std::pair<op, shape_info>> out;
                    for (node const* e : order_)
                        out.emplace_back(e->kind, shape(e));
                    return out;
                }

            private:
                queue& q_;
                std::vector<std::unique_ptr<node>> wrappers_;
                std::map<node const*, node const*> wrapper_of_;
                std::unordered_map<node const*, std::set<path>> traces_;
                std::unordered_set<node const*> found_;
                std::vector<node const*> order_;
                std::unordered_set<node const*> in_order_;
                std::unordered_set<node const*> elided_, fused_products_;
                std::unordered_map<node const*, tensor> results_;
                std::unordered_map<node const*, shape_info> shapes_;

                // the top-level node: cacheable expressions stand for themselves, others are wrapped
                // (hmpc::expr::cache, cache.hpp:71-82; one wrapper per distinct expression)
                node const* top_level(expr::expression const& r)
                {
                    node const* e = r.get();
                    if (cacheable(e))
                        return e;
                    auto it = wrapper_of_.find(e);
                    if (it != wrapper_of_.end())
                        return it->second;
                    auto w = std::make_unique<node>();
                    w->kind = op::cache;
                    w->a = r;
                    node const* wp = w.get();
                    wrappers_.push_back(std::move(w));
                    wrapper_of_.emplace(e, wp);
                    return wp;
                }

                shape_info const& shape(node const* e)
                {
                    auto it = shapes_.find(e);
                    if (it != shapes_.end())
                        return it->second;
                    shape_info s;
                    switch (e->kind)
                    {
                    case op::leaf:
                        s = {e->value.mod(), e->value.batch(), e->value.log_n(), e->value.rep()};
                        break;
                    case op::uniform:
                    case op::cbd:
                    case op::capability:
                        s = {e->mod, e->n, e->log_n, representation::coefficient};
                        break;
                    case op::ntt:
                        s = shape(e->a.get());
                        s.rep = representation::number_theoretic_transform;
                        break;
                    case op::intt:
                        s = shape(e->a.get());
                        s.rep = representation::coefficient;
                        break;
                    case op::cast:
                        s = shape(e->a.get());
                        s.mod = e->mod;
                        break;
                    case op::scalar_mul:
                    case op::cache:
                        s = shape(e->a.get());
                        break;
                    case op::sum:
                    case op::product:
                        s = shape(e->a.get());
                        s.n = 1;
                        s.log_n = 0;
                        break;
                    case op::matvec:
                    {
                        shape_info const& m = shape(e->a.get());
                        s = m;
                        int log_rows = 0;
                        while ((std::size_t{1} << log_rows) < m.n)
                            ++log_rows;
                        s.n = 1;
                        s.log_n = log_rows;
                        break;
                    }
                    default:  // add, sub, mul: the larger operand's shape (broadcast)
                    {
                        shape_info const& a = shape(e->a.get());
                        shape_info const& b = shape(e->b.get());
                        s = a.size() >= b.size() ? a : b;
                        break;
                    }
                    }
                    return shapes_.emplace(e, s).first->second;
                }

                // an operand reached through a broadcast is read for several output elements
                bool multiple_access(node const* e)
                {
                    if (e->kind != op::add && e->kind != op::sub && e->kind != op::mul)
                        return false;
                    return shape(e->a.get()).size() != shape(e->b.get()).size();
                }

                void trace(node const* x, path p)
                {
                    if (cacheable(x))
                    {
                        traces_[x].insert(std::move(p));
                        return;
                    }
                    for (int i = 0; i < arity(x); ++i)
                    {
                        path q = p;
                        q.steps.emplace_back(x, i);
                        q.multiple = q.multiple || multiple_access(x);
                        trace(child(x, i), std::move(q));
                    }
                }

                // cache.hpp:113-147 (find_next_expression_to_trace)
                void find(node const* e)
                {
                    if (!found_.insert(e).second)
                        return;
                    if (cacheable(e))
                        for (int i = 0; i < arity(e); ++i)
                        {
                            path p;
                            p.steps.emplace_back(e, i);
                            trace(child(e, i), std::move(p));
                        }
                    for (int i = 0; i < arity(e); ++i)
                        find(child(e, i));
                }

                void insert(node const* e)
                {
                    if (in_order_.insert(e).second)
                        order_.push_back(e);
                }

                // cache.hpp:187-231 (cache_expression)
                void cache(node const* e, bool top)
                {
                    auto traverse = [&] {
                        for (int i = 0; i < arity(e); ++i)
                            cache(child(e, i), false);
                    };
                    if (!cacheable(e))
                    {
                        traverse();
                        return;
                    }
                    auto const& tr = traces_[e];
                    bool fused = !top && !complex(e) && tr.size() == 1 && !tr.begin()->multiple;
                    traverse();
                    if (!fused)
                        insert(e);
                }

                bool entry(node const* e) const { return in_order_.count(e) != 0; }

                // does the kernel subtree of x (stopping at entries) draw on the capability?
                bool draws(node const* x) const
                {
                    if (entry(x))
                        return false;
                    if (x->kind == op::capability)
                        return true;
                    for (int i = 0; i < arity(x); ++i)
                        if (draws(child(x, i)))
                            return true;
                    return false;
                }

                // INTT(NTT(a) o NTT(b)) whose two forward transforms are used there only, are not
                // top-level and draw no randomness: one batched product kernel (identical results; the
                // transforms of a and b are not materialised)
                void mark_fused_products(std::vector<node const*> const& tops)
                {
                    for (node const* e : order_)
                    {
                        if (e->kind != op::intt || e->a->kind != op::mul)
                            continue;
                        node const* x = e->a->a.get();
                        node const* y = e->a->b.get();
                        if (x->kind != op::ntt || y->kind != op::ntt || x == y)
                            continue;
                        auto single = [&](node const* t) {
                            auto const& tr = traces_[t];
                            return tr.size() == 1 && !tr.begin()->multiple && !draws(t->a.get()) &&
                                   std::find(tops.begin(), tops.end(), t) == tops.end();
                        };
                        shape_info const& sx = shape(x->a.get());
                        shape_info const& sy = shape(y->a.get());
                        if (single(x) && single(y) && sx.mod == sy.mod && sx.size() == sy.size())
                        {
                            elided_.insert(x);
                            elided_.insert(y);
                            fused_products_.insert(e);
                        }
                    }
                }

                // one capability sampler of shape s drawn in layout `lay` at the queue's nonce
                tensor sample(node const* x, layout lay)
                {
                    tensor t = make(shape(x));
                    check(hmpc_sample_capability(q_.handle(), x->mod.handle(), x->dist, x->bits, static_cast<int>(lay), x->log_n,
                                                 x->n, t.data(), nullptr));
                    return t;
                }

                tensor make(shape_info const& s) { return tensor(q_, s.mod, s.n, s.log_n, s.rep); }

                // the value of x inside the current kernel (x fused into it), reading entries' results
                tensor value(node const* x, layout lay, int& capability_draws)
                {
                    if (entry(x))
                        return results_.at(x);
                    shape_info const& s = shape(x);
                    switch (x->kind)
                    {
                    case op::leaf:
                        return x->value;
                    case op::cache:
                        // a cache node fused into its ancestor's kernel: its operand's value
                        return value(x->a.get(), lay, capability_draws);
                    case op::uniform:
                    {
                        tensor t = make(s);
                        check(hmpc_sample_uniform(s.mod.handle(), x->key.data(), x->offset, t.size(), t.data(), nullptr));
                        return t;
                    }
                    case op::cbd:
                    {
                        tensor t = make(s);
                        check(hmpc_sample_cbd(s.mod.handle(), x->key.data(), x->offset, x->bits, t.size(), t.data(), nullptr));
                        return t;
                    }
                    case op::capability:
                    {
                        // one capability per kernel element: a second sampler in the same kernel would
                        // continue the same stream, which this interface does not restate
                        if (++capability_draws > 1)
                            throw error(HMPC_ERR_INVALID_ARGUMENT, "two capability samplers fused into one kernel");
                        return sample(x, lay);
                    }
                    case op::cast:
                    {
                        tensor a = value(x->a.get(), lay, capability_draws);
                        tensor t = make(s);
                        check(hmpc_cast(a.mod().handle(), s.mod.handle(), a.size(), a.data(), t.data(), nullptr));
                        return t;
                    }
                    case op::scalar_mul:
                    {
                        tensor a = value(x->a.get(), lay, capability_draws);
                        tensor t = make(s);
                        check(hmpc_pointwise_scalar(a.mod().handle(), HMPC_MUL, a.size(), a.data(), x->c.data(), t.data(), nullptr));
                        return t;
                    }
                    case op::matvec:
                    {
                        tensor m = value(x->a.get(), lay, capability_draws);
                        tensor v = value(x->b.get(), lay, capability_draws);
                        std::size_t rows = m.batch(), cols = m.size() / m.batch();
                        if (!(m.mod() == v.mod()) || v.size() != cols || (std::size_t{1} << s.log_n) != rows)
                            throw error(HMPC_ERR_INVALID_ARGUMENT, "matrix_vector_product: one modulus, v of M's row length, 2^k rows");
                        tensor t = make(s);
                        check(hmpc_matvec(m.mod().handle(), 1, rows, cols, m.data(), 0, v.data(), 0, t.data(), nullptr));
                        return t;
                    }
                    case op::add:
                    case op::sub:
                    case op::mul:
                    {
                        // a * b + c of one shape: one fused kernel (hmpc fuses the whole element-wise tree)
                        if (x->kind == op::add && x->a->kind == op::mul && !entry(x->a.get()))
                        {
                            node const* m = x->a.get();
                            if (shape(m->a.get()).size() == s.size() && shape(m->b.get()).size() == s.size() &&
                                shape(x->b.get()).size() == s.size())
                            {
                                tensor a = value(m->a.get(), lay, capability_draws);
                                tensor b = value(m->b.get(), lay, capability_draws);
                                tensor c = value(x->b.get(), lay, capability_draws);
                                if (a.mod() == b.mod() && a.mod() == c.mod())
                                {
                                    tensor t = make(s);
                                    check(hmpc_fma(a.mod().handle(), a.size(), a.data(), b.data(), c.data(), t.data(), nullptr));
                                    return t;
                                }
                                throw error(HMPC_ERR_INVALID_ARGUMENT, "operands need one modulus");
                            }
                        }
                        tensor a = value(x->a.get(), lay, capability_draws);
                        tensor b = value(x->b.get(), lay, capability_draws);
                        int o = x->kind == op::add ? HMPC_ADD : (x->kind == op::sub ? HMPC_SUB : HMPC_MUL);
                        if (a.size() < b.size() && o != HMPC_SUB)
                            std::swap(a, b);  // commutative: broadcast the smaller operand
                        if (!(a.mod() == b.mod()) || b.size() == 0 || a.size() % b.size() != 0)
                            throw error(HMPC_ERR_INVALID_ARGUMENT, "operands need one modulus and a broadcastable shape");
                        tensor t = make(s);
                        check(hmpc_pointwise(a.mod().handle(), o, a.size(), a.data(), b.data(), b.size(), t.data(), nullptr));
                        return t;
                    }
                    default:
                        throw error(HMPC_ERR_INVALID_ARGUMENT, "a complex expression was not planned as a kernel");
                    }
                }

                // one entry of the plan: its own kernel over its fused subtree
                tensor kernel(node const* e)
                {
                    int draws_in_kernel = 0;
                    shape_info const& s = shape(e);
                    switch (e->kind)
                    {
                    case op::cache:
                    {
                        // hmpc materialises a top-level expression into a new tensor
                        tensor v = value(e->a.get(), layout::element, draws_in_kernel);
                        if (e->a->kind == op::leaf)
                        {
                            tensor t = make(s);
                            check(hmpc_memcpy_d2d(q_.handle(), t.data(), v.data(), v.bytes(), nullptr));
                            return t;
                        }
                        return v;
                    }
                    case op::capability:
                        // materialised (shared or reached through a broadcast): an element-wise kernel
                        return sample(e, layout::element);
                    case op::ntt:
                    {
                        node const* inner = e->a.get();
                        if (inner->kind == op::capability && !entry(inner) && inner->dist == HMPC_CENTERED_BINOMIAL)
                        {
                            // sampling fused with the transform's first stages (same nonce and layout)
                            tensor t = make(s);
                            check(hmpc_sample_cbd_ntt(q_.handle(), inner->mod.handle(), inner->bits, inner->log_n, inner->n,
                                                      t.data(), nullptr));
                            return t;
                        }
                        tensor x = value(inner, layout::ntt, draws_in_kernel);
                        tensor t = make(s);
                        check(hmpc_ntt(x.mod().handle(), x.log_n(), x.batch(), 0, x.data(), t.data(), nullptr));
                        return t;
                    }
                    case op::intt:
                    {
                        if (fused_products_.count(e))
                        {
                            tensor a = value(e->a->a->a.get(), layout::element, draws_in_kernel);
                            tensor b = value(e->a->b->a.get(), layout::element, draws_in_kernel);
                            tensor t = make(s);
                            check(hmpc_polymul(a.mod().handle(), a.log_n(), a.batch(), a.data(), b.data(), t.data(), nullptr));
                            return t;
                        }
                        tensor x = value(e->a.get(), layout::intt, draws_in_kernel);
                        tensor t = make(s);
                        check(hmpc_ntt(x.mod().handle(), x.log_n(), x.batch(), 1, x.data(), t.data(), nullptr));
                        return t;
                    }
                    case op::sum:
                    case op::product:
                    {
                        if (draws(e->a.get()))
                            throw error(HMPC_ERR_INVALID_ARGUMENT, "capability samplers inside a reduction are not supported");
                        tensor a = value(e->a.get(), layout::element, draws_in_kernel);
                        tensor t = make(s);
                        check(hmpc_reduce(a.mod().handle(), e->kind == op::sum ? HMPC_ADD : HMPC_MUL, a.size(), a.data(), t.data(),
                                          nullptr));
                        return t;
                    }
                    default:
                        throw error(HMPC_ERR_INVALID_ARGUMENT, "unknown plan entry");
                    }
                }
            };
        }

        inline std::vector<tensor> queue::eval(std::vector<expr::expression> const& roots)
        {
            detail::plan p(*this);
            return p.run(roots);
        }

        inline std::vector<queue::cache_entry> queue::execution_cache(std::vector<expr::expression> const& roots)
        {
            detail::plan p(*this);
            std::vector<cache_entry> out;
            for (auto const& [kind, s] : p.analyze(roots))
                out.push_back({kind, s.n, s.log_n});
            return out;
        }
    }
}
