// hmpc's own spelling of the NTT / BGV path over libhmpc_hip.so: the names, types and call forms of
// hmpc::ints, hmpc::comp and hmpc::expr that code on this path uses, so that such code compiles
// against this header with its hmpc includes replaced by this one and its queue built without a
// SYCL queue (there is no SYCL here: hmpc::comp::queue{} or hmpc::comp::queue{key}).
//
//   hmpc::ints::literals::operator""_int        compile-time unsigned integers (hmpc/ints/literals.hpp)
//   hmpc::ints::ubigint<Bits>, mod<q>           q as a template argument; values in Montgomery form,
//   poly_mod<q, N, representation>              a R mod q with R = 2^(32 L) (hmpc/ints/mod.hpp:90-127)
//   traits::number_theoretic_transform_type_t   and coefficient_type_t (hmpc/ints/poly_mod.hpp:78-90)
//   integer_traits<mod<q>>::zero / one, num::bit_copy<T>(core::size_limb_span<limb>(i))
//   hmpc::shape{...}, hmpc::constants::*        extents, axis and variance constants
//   hmpc::comp::make_tensor<T>(shape)           a tensor that belongs to no queue until first used
//   hmpc::comp::host_accessor(t, access::...)   element access on the host (hmpc/comp/accessor.hpp)
//   hmpc::comp::queue, queue(e1, e2, ...)       one execution plan, tuple of tensors (comp/queue.hpp)
//   hmpc::expr::tensor / value / constant_of    leaves (hmpc/expr/tensor.hpp, value.hpp, constant.hpp)
//   number_theoretic_transform, inverse_...,    hmpc/expr/number_theoretic_transform.hpp, cast.hpp,
//   cast<T>, unsqueeze, operators + - *         unsqueeze.hpp, binary_expression.hpp
//   expr::random::uniform<T> / centered_binomial<T>(shape[, variance])   queue capability samplers
//   expr::crypto::lhe::{key, randomness, drowning_randomness, enc, ciphertext, dec<Plaintext>}
//   comp::crypto::lhe::{key, ciphertext}        (hmpc/expr/crypto/lhe/*.hpp, hmpc/comp/crypto/lhe/*.hpp)
//
// Evaluation goes through hmpc_amd::comp (comp.hpp): hmpc's execution-cache plan, its capability
// nonces and stream layouts, its rewrites, on the library's kernels. Expressions are values that
// share their nodes: copies of one sampler expression are one node (as one expression type in hmpc),
// separate calls make separate nodes (hmpc's unique tags). A tensor lives on the host until a queue
// uses it, then on that queue's device; a host_accessor moves it back. Tensors a queue computed must
// be released or read before that queue is destroyed.
//
// Not here: SYCL kernels written by the caller (device_accessor + parallel_for: fill tensors through a
// host_accessor instead), explicit-key samplers (use_number_generator), compile-time tensor
// extents (a shape is a run-time value), and expressions beyond this path.
//
// Header only, C++20 (class-type template arguments, as hmpc); link with -lhmpc_hip.
#pragma once

#if __cplusplus < 202002L
#error "hmpc_amd/hmpc.hpp needs C++20"
#endif

#include "comp.hpp"

#include <algorithm>
#include <array>
#include <bit>
#include <concepts>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

namespace hmpc
{
    using size = std::size_t;
    using signed_size = std::ptrdiff_t;

    template<typename T, T Value>
    struct constant
    {
        using value_type = T;
        static constexpr T value = Value;
        constexpr operator T() const { return Value; }
    };
    template<size V>
    using size_constant = constant<size, V>;
    template<signed_size V>
    using signed_size_constant = constant<signed_size, V>;
    template<size Numerator, size Denominator>
    struct rational_size_constant
    {
        static constexpr size numerator = Numerator;
        static constexpr size denominator = Denominator;
    };

    namespace constants
    {
        inline constexpr size_constant<0> zero{};
        inline constexpr size_constant<1> one{};
        inline constexpr size_constant<2> two{};
        inline constexpr size_constant<3> three{};
        inline constexpr size_constant<4> four{};
        inline constexpr size_constant<5> five{};
        inline constexpr size_constant<6> six{};
        inline constexpr size_constant<7> seven{};
        inline constexpr size_constant<8> eight{};
        inline constexpr size_constant<9> nine{};
        inline constexpr size_constant<10> ten{};
        inline constexpr signed_size_constant<-1> minus_one{};
        inline constexpr signed_size_constant<-2> minus_two{};
        inline constexpr rational_size_constant<1, 2> half{};
    }

    // extents of a tensor of polynomials (or of elements); rank 0 is one polynomial
    struct shape
    {
        static constexpr hmpc::size max_rank = 8;
        hmpc::size rank = 0;
        std::array<hmpc::size, max_rank> extents{};

        constexpr shape() = default;
        template<std::integral... S>
            requires(sizeof...(S) >= 1 && sizeof...(S) <= max_rank)
        constexpr shape(S... s) : rank(sizeof...(S)), extents{static_cast<hmpc::size>(s)...}
        {
        }

        constexpr hmpc::size get(hmpc::size axis) const { return extents[axis]; }
        // an axis given as a constant; negative axes count from the end
        template<typename T, T Axis>
        constexpr hmpc::size get(constant<T, Axis>) const
        {
            auto a = static_cast<signed_size>(Axis);
            return extents[static_cast<hmpc::size>(a < 0 ? a + static_cast<signed_size>(rank) : a)];
        }
        constexpr hmpc::size size() const
        {
            hmpc::size s = 1;
            for (hmpc::size i = 0; i < rank; ++i)
                s *= extents[i];
            return s;
        }
        constexpr bool operator==(shape const& o) const
        {
            if (rank != o.rank)
                return false;
            for (hmpc::size i = 0; i < rank; ++i)
                if (extents[i] != o.extents[i])
                    return false;
            return true;
        }
    };

    namespace access
    {
        enum class mode
        {
            read,
            write,
            read_write,
            discard_write,
        };
        inline constexpr std::integral_constant<mode, mode::read> read{};
        inline constexpr std::integral_constant<mode, mode::write> write{};
        inline constexpr std::integral_constant<mode, mode::read_write> read_write{};
        inline constexpr std::integral_constant<mode, mode::discard_write> discard_write{};
    }

    namespace ints
    {
        using limb = std::uint32_t;

        namespace detail
        {
            constexpr size bit_width_of(limb const* w, size n)
            {
                for (size i = n; i-- > 0;)
                    if (w[i])
                        return 32 * i + static_cast<size>(std::bit_width(w[i]));
                return 0;
            }
        }

        // an unsigned integer of Bits bits in 32-bit limbs, least significant first
        template<size Bits>
        struct ubigint
        {
            static constexpr size bit_size = Bits;
            static constexpr size limb_size = Bits == 0 ? 1 : (Bits + 31) / 32;
            using limb_type = limb;

            limb data[limb_size] = {};

            constexpr ubigint() = default;
            // the low Bits bits of an integer (two's complement for negative values)
            template<std::integral T>
            constexpr ubigint(T v)
            {
                auto x = static_cast<std::uint64_t>(v);
                data[0] = static_cast<limb>(x);
                if constexpr (limb_size > 1)
                    data[1] = static_cast<limb>(x >> 32);
                if constexpr (Bits % 32 != 0)
                    data[limb_size - 1] &= (limb{1} << (Bits % 32)) - 1u;
            }
            constexpr bool operator==(ubigint const&) const = default;
        };

        namespace literals
        {
            namespace detail
            {
                template<size Words>
                struct parsed
                {
                    limb w[Words] = {};
                    size bits = 0;
                };

                constexpr int digit(char c)
                {
                    if (c >= '0' && c <= '9')
                        return c - '0';
                    if (c >= 'a' && c <= 'f')
                        return c - 'a' + 10;
                    if (c >= 'A' && c <= 'F')
                        return c - 'A' + 10;
                    return -1;
                }

                // hexadecimal (0x), binary (0b), octal (leading 0) or decimal digits with ' separators
                template<char... Cs>
                consteval auto parse()
                {
                    constexpr char s[] = {Cs..., '\0'};
                    constexpr size n = sizeof...(Cs);
                    parsed<(4 * n + 31) / 32 + 1> r{};
                    size i = 0;
                    unsigned base = 10;
                    if (n > 1 && s[0] == '0' && (s[1] == 'x' || s[1] == 'X'))
                        base = 16, i = 2;
                    else if (n > 1 && s[0] == '0' && (s[1] == 'b' || s[1] == 'B'))
                        base = 2, i = 2;
                    else if (n > 1 && s[0] == '0')
                        base = 8, i = 1;
                    constexpr size words = (4 * n + 31) / 32 + 1;
                    for (; i < n; ++i)
                    {
                        if (s[i] == '\'')
                            continue;
                        int d = digit(s[i]);
                        if (d < 0 || static_cast<unsigned>(d) >= base)
                            throw "invalid digit in an _int literal";
                        std::uint64_t carry = static_cast<unsigned>(d);
                        for (size k = 0; k < words; ++k)
                        {
                            std::uint64_t v = std::uint64_t{r.w[k]} * base + carry;
                            r.w[k] = static_cast<limb>(v);
                            carry = v >> 32;
                        }
                    }
                    r.bits = ints::detail::bit_width_of(r.w, words);
                    if (r.bits == 0)
                        r.bits = 1;
                    return r;
                }
            }

            // the literal's value as ubigint<bit width of the value>
            template<char... Cs>
            consteval auto operator""_int()
            {
                constexpr auto p = detail::parse<Cs...>();
                ubigint<p.bits> r{};
                for (size i = 0; i < ubigint<p.bits>::limb_size; ++i)
                    r.data[i] = p.w[i];
                return r;
            }
        }

        namespace detail
        {
            // little-endian limb arithmetic on L limbs, usable in constant expressions
            template<size L>
            struct words
            {
                limb w[L] = {};
            };

            template<size L>
            constexpr bool geq(limb const* a, limb const* b)
            {
                for (size i = L; i-- > 0;)
                    if (a[i] != b[i])
                        return a[i] > b[i];
                return true;
            }

            template<size L>
            constexpr void sub_in_place(limb* a, limb const* b)
            {
                std::uint64_t borrow = 0;
                for (size i = 0; i < L; ++i)
                {
                    std::uint64_t v = std::uint64_t{a[i]} - b[i] - borrow;
                    a[i] = static_cast<limb>(v);
                    borrow = (v >> 63) & 1u;
                }
            }

            // x = 2 x mod q for x < q (one more limb holds the doubled value's carry)
            template<size L>
            constexpr void double_mod(limb* x, limb const* q)
            {
                limb top = x[L - 1] >> 31;
                for (size i = L; i-- > 1;)
                    x[i] = (x[i] << 1) | (x[i - 1] >> 31);
                x[0] <<= 1;
                if (top || geq<L>(x, q))
                    sub_in_place<L>(x, q);
            }

            // REDC(a b) for a < R, b < q: CIOS with n' = -q^-1 mod 2^32 and one final subtraction
            template<size L>
            constexpr words<L> montgomery_product(limb const* a, limb const* b, limb const* q, limb nprime)
            {
                limb t[L + 2] = {};
                for (size i = 0; i < L; ++i)
                {
                    std::uint64_t c = 0;
                    for (size j = 0; j < L; ++j)
                    {
                        std::uint64_t s = std::uint64_t{a[j]} * b[i] + t[j] + c;
                        t[j] = static_cast<limb>(s);
                        c = s >> 32;
                    }
                    std::uint64_t s = std::uint64_t{t[L]} + c;
                    t[L] = static_cast<limb>(s);
                    t[L + 1] = static_cast<limb>(s >> 32);
                    limb m = t[0] * nprime;
                    c = (std::uint64_t{m} * q[0] + t[0]) >> 32;
                    for (size j = 1; j < L; ++j)
                    {
                        std::uint64_t v = std::uint64_t{m} * q[j] + t[j] + c;
                        t[j - 1] = static_cast<limb>(v);
                        c = v >> 32;
                    }
                    s = std::uint64_t{t[L]} + c;
                    t[L - 1] = static_cast<limb>(s);
                    t[L] = t[L + 1] + static_cast<limb>(s >> 32);
                }
                words<L> r{};
                for (size j = 0; j < L; ++j)
                    r.w[j] = t[j];
                if (t[L] || geq<L>(r.w, q))
                    sub_in_place<L>(r.w, q);
                return r;
            }
        }

        // an element of Z/qZ for the compile-time modulus Q (a ubigint), stored as hmpc stores it:
        // L = ceil(bit_width(q - 1) / 32) limbs of a R mod q, R = 2^(32 L) (hmpc/ints/mod.hpp:90-127)
        template<auto Q>
        struct mod
        {
            using modulus_type = std::remove_cv_t<decltype(Q)>;
            using limb_type = limb;
            static constexpr modulus_type modulus_constant = Q;

        private:
            static constexpr size compute_bit_size()
            {
                modulus_type m = Q;
                // q - 1 (q is odd, so no borrow beyond limb 0)
                m.data[0] -= 1;
                return ints::detail::bit_width_of(m.data, modulus_type::limb_size);
            }

        public:
            static constexpr size bit_size = compute_bit_size();
            static constexpr size limb_size = (bit_size + 31) / 32;

        private:
            static constexpr detail::words<limb_size> q_words()
            {
                detail::words<limb_size> r{};
                for (size i = 0; i < limb_size && i < modulus_type::limb_size; ++i)
                    r.w[i] = Q.data[i];
                return r;
            }
            static constexpr detail::words<limb_size> power_mod(size doublings)
            {
                // 2^doublings mod q from 1 by repeated doubling
                detail::words<limb_size> x{};
                x.w[0] = 1;
                constexpr auto q = q_words();
                for (size i = 0; i < doublings; ++i)
                    detail::double_mod<limb_size>(x.w, q.w);
                return x;
            }
            static constexpr limb compute_nprime()
            {
                limb q0 = Q.data[0];
                limb inv = 1;
                for (int i = 0; i < 5; ++i)
                    inv *= 2u - q0 * inv;  // Newton: inv = q0^-1 mod 2^32
                return static_cast<limb>(0u - inv);
            }

        public:
            static constexpr detail::words<limb_size> q = q_words();
            static constexpr detail::words<limb_size> r1 = power_mod(32 * limb_size);      // R mod q
            static constexpr detail::words<limb_size> r2 = power_mod(64 * limb_size);      // R^2 mod q
            static constexpr limb nprime = compute_nprime();

            limb data[limb_size] = {};

            constexpr mod() = default;
            // the residue of an unsigned integer (Montgomery form of v mod q)
            template<size Bits>
            constexpr explicit mod(ubigint<Bits> const& v)
            {
                constexpr size VL = ubigint<Bits>::limb_size;
                detail::words<limb_size> a{};
                if constexpr (VL <= limb_size)
                {
                    for (size i = 0; i < VL; ++i)
                        a.w[i] = v.data[i];
                }
                else
                {
                    // Horner from the top limb: a = a 2^32 + limb (mod q), each limb below q after reduction
                    for (size i = VL; i-- > 0;)
                    {
                        for (int k = 0; k < 32; ++k)
                            detail::double_mod<limb_size>(a.w, q.w);
                        detail::words<limb_size> d{};
                        // a limb is below q unless q has one limb (then reduce it first)
                        d.w[0] = limb_size == 1 ? v.data[i] % q.w[0] : v.data[i];
                        // a + d < 2q: one conditional subtraction
                        std::uint64_t carry = 0;
                        for (size j = 0; j < limb_size; ++j)
                        {
                            std::uint64_t s = std::uint64_t{a.w[j]} + d.w[j] + carry;
                            a.w[j] = static_cast<limb>(s);
                            carry = s >> 32;
                        }
                        if (carry || detail::geq<limb_size>(a.w, q.w))
                            detail::sub_in_place<limb_size>(a.w, q.w);
                    }
                    // a = v mod q in plain form: its Montgomery form is mont(a, R^2)
                }
                auto m = detail::montgomery_product<limb_size>(a.w, r2.w, q.w, nprime);
                for (size i = 0; i < limb_size; ++i)
                    data[i] = m.w[i];
            }
            constexpr bool operator==(mod const&) const = default;

            // the stored words (a R mod q)
            static constexpr mod from_words(limb const* w)
            {
                mod r{};
                for (size i = 0; i < limb_size; ++i)
                    r.data[i] = w[i];
                return r;
            }
            // Montgomery product (a R)(b R) R^-1 = ab R: hmpc's mod operator*
            friend constexpr mod operator*(mod const& a, mod const& b)
            {
                auto m = detail::montgomery_product<limb_size>(a.data, b.data, q.w, nprime);
                return from_words(m.w);
            }
            std::vector<limb> words() const { return std::vector<limb>(data, data + limb_size); }
            static std::vector<limb> modulus_words() { return std::vector<limb>(q.w, q.w + limb_size); }
        };

        template<typename T>
        struct integer_traits;
        template<auto Q>
        struct integer_traits<mod<Q>>
        {
            using type = mod<Q>;
            static constexpr size bit_size = type::bit_size;
            static constexpr type zero = {};
            static constexpr type one = type::from_words(type::r1.w);
        };

        enum class polynomial_representation
        {
            coefficient,
            number_theoretic_transform,
        };
        inline constexpr auto coefficient_representation = polynomial_representation::coefficient;
        inline constexpr auto number_theoretic_transform_representation = polynomial_representation::number_theoretic_transform;

        // hmpc/ints/poly_mod.hpp: a polynomial of N coefficients mod q in one representation
        template<auto Modulus, size Degree, polynomial_representation Representation>
        struct poly_mod
        {
            using element_type = mod<Modulus>;
            using limb_type = limb;
            static constexpr size degree = Degree;
            static constexpr auto representation = Representation;
        };

        namespace traits
        {
            template<typename T>
            struct number_theoretic_transform_type;
            template<auto Q, size N, polynomial_representation R>
            struct number_theoretic_transform_type<poly_mod<Q, N, R>>
            {
                using type = poly_mod<Q, N, number_theoretic_transform_representation>;
            };
            template<typename T>
            using number_theoretic_transform_type_t = typename number_theoretic_transform_type<T>::type;

            template<typename T>
            struct coefficient_type;
            template<auto Q, size N, polynomial_representation R>
            struct coefficient_type<poly_mod<Q, N, R>>
            {
                using type = poly_mod<Q, N, coefficient_representation>;
            };
            template<typename T>
            using coefficient_type_t = typename coefficient_type<T>::type;
        }

        namespace num
        {
            // the limbs of a span copied into T's storage (zero-extended or truncated): hmpc's
            // bit_copy puts raw words into a mod's Montgomery representation
            template<typename T, typename Span>
            constexpr T bit_copy(Span const& span)
            {
                T r{};
                for (size i = 0; i < T::limb_size && i < span.size(); ++i)
                    r.data[i] = span[i];
                return r;
            }
        }
    }

    namespace core
    {
        // the limbs of a size value, least significant first (hmpc/core/size_limb_span.hpp)
        template<typename Limb>
        struct size_limb_span
        {
            static constexpr hmpc::size limbs = (sizeof(hmpc::size) + sizeof(Limb) - 1) / sizeof(Limb);
            Limb data[limbs] = {};
            constexpr size_limb_span(hmpc::size value)
            {
                for (hmpc::size i = 0; i < limbs; ++i)
                    data[i] = static_cast<Limb>(value >> (8 * sizeof(Limb) * i));
            }
            constexpr hmpc::size size() const { return limbs; }
            constexpr Limb operator[](hmpc::size i) const { return data[i]; }
        };
    }

    namespace detail
    {
        // what a value type means to the library: its modulus words, degree and representation
        template<typename T>
        struct value_info;
        template<auto Q>
        struct value_info<ints::mod<Q>>
        {
            using element_type = ints::mod<Q>;
            static constexpr int log_n = 0;
            static constexpr auto rep = hmpc_amd::comp::representation::coefficient;
        };
        template<auto Q, size N, ints::polynomial_representation R>
        struct value_info<ints::poly_mod<Q, N, R>>
        {
            static_assert(N >= 1 && (N & (N - 1)) == 0, "the degree is a power of two");
            using element_type = ints::mod<Q>;
            static constexpr int log_n = std::countr_zero(N);
            static constexpr auto rep = R == ints::number_theoretic_transform_representation
                                            ? hmpc_amd::comp::representation::number_theoretic_transform
                                            : hmpc_amd::comp::representation::coefficient;
        };

        // shape of an element-wise result: numpy-style broadcasting from the last axis
        inline hmpc::shape broadcast(hmpc::shape const& a, hmpc::shape const& b)
        {
            hmpc::shape r;
            r.rank = std::max(a.rank, b.rank);
            for (size i = 0; i < r.rank; ++i)
            {
                size ea = i < a.rank ? a.extents[a.rank - 1 - i] : 1;
                size eb = i < b.rank ? b.extents[b.rank - 1 - i] : 1;
                if (ea != eb && ea != 1 && eb != 1)
                    throw hmpc_amd::error(HMPC_ERR_INVALID_ARGUMENT, "shapes do not broadcast");
                r.extents[r.rank - 1 - i] = std::max(ea, eb);
            }
            return r;
        }
    }

    namespace comp
    {
        class queue;

        namespace detail
        {
            // the storage of a tensor: host words and/or the device tensor of the queue that last
            // used it; `host_current` / `device_current` say which copy holds the value
            struct tensor_state
            {
                hmpc::shape shape;
                std::vector<std::uint32_t> host;
                bool host_current = true;
                hmpc_amd::comp::tensor device;
                queue const* owner = nullptr;
                bool device_current = false;
            };

            // the tensor's value on queue q's device (uploaded when q does not hold it)
            template<typename T>
            hmpc_amd::comp::tensor bind(queue& q, std::shared_ptr<tensor_state> const& s);
        }

        template<typename T>
        class tensor
        {
        public:
            using value_type = T;
            using element_type = typename hmpc::detail::value_info<T>::element_type;

            tensor() = default;
            hmpc::shape shape() const { return state_->shape; }
            // number of elements of the element type (coefficients for polynomial tensors)
            hmpc::size element_count() const { return state_->shape.size() << hmpc::detail::value_info<T>::log_n; }

            // internal: construction, binding to a queue and host synchronisation
            explicit tensor(std::shared_ptr<detail::tensor_state> s) : state_(std::move(s)) {}
            detail::tensor_state& state() const { return *state_; }
            std::shared_ptr<detail::tensor_state> const& shared_state() const { return state_; }
            void to_host() const
            {
                auto& s = *state_;
                if (s.host_current)
                    return;
                s.host = s.device.to_host();
                s.host_current = true;
            }
            void ensure_host_storage() const
            {
                auto& s = *state_;
                std::size_t words = element_count() * element_type::limb_size;
                if (s.host.size() != words)
                    s.host.assign(words, 0u);
            }

        private:
            std::shared_ptr<detail::tensor_state> state_;
        };

        // hmpc::comp::make_tensor<T>(shape): zero-filled, on the host until a queue uses it
        template<typename T>
        tensor<T> make_tensor(hmpc::shape shape)
        {
            auto s = std::make_shared<detail::tensor_state>();
            s->shape = shape;
            tensor<T> t(std::move(s));
            t.ensure_host_storage();
            return t;
        }

        // host_accessor(t, hmpc::access::...): element i of the limb-planar [L][count] storage
        template<typename T>
        class host_accessor
        {
        public:
            using element_type = typename tensor<T>::element_type;
            static constexpr hmpc::size L = element_type::limb_size;

            class reference
            {
            public:
                reference(std::uint32_t* base, hmpc::size stride) : base_(base), stride_(stride) {}
                operator element_type() const
                {
                    element_type v{};
                    for (hmpc::size l = 0; l < L; ++l)
                        v.data[l] = base_[l * stride_];
                    return v;
                }
                reference& operator=(element_type const& v)
                {
                    for (hmpc::size l = 0; l < L; ++l)
                        base_[l * stride_] = v.data[l];
                    return *this;
                }
                reference& operator=(reference const& o) { return *this = static_cast<element_type>(o); }
                friend bool operator==(reference const& a, element_type const& b) { return static_cast<element_type>(a) == b; }

            private:
                std::uint32_t* base_;
                hmpc::size stride_;
            };

            template<hmpc::access::mode M>
            host_accessor(tensor<T>& t, std::integral_constant<hmpc::access::mode, M>) : t_(t)
            {
                auto& s = t.state();
                if constexpr (M == hmpc::access::mode::discard_write)
                {
                    t.ensure_host_storage();
                    s.host_current = true;
                }
                else
                {
                    t.to_host();
                    t.ensure_host_storage();
                }
                if constexpr (M != hmpc::access::mode::read)
                    s.device_current = false;
                count_ = t.element_count();
            }

            hmpc::size size() const { return count_; }
            reference operator[](hmpc::size i) { return reference(t_.state().host.data() + i, count_); }
            element_type operator[](hmpc::size i) const
            {
                return static_cast<element_type>(reference(const_cast<std::uint32_t*>(t_.state().host.data()) + i, count_));
            }

        private:
            tensor<T>& t_;
            hmpc::size count_ = 0;
        };
        template<typename T, typename M>
        host_accessor(tensor<T>&, M) -> host_accessor<T>;

        namespace crypto::lhe
        {
            // hmpc/comp/crypto/lhe/key.hpp: a = B s + p e and b, NTT form, one polynomial each
            template<typename Poly>
            struct key
            {
                hmpc::comp::tensor<Poly> a;
                hmpc::comp::tensor<Poly> b;
            };
            template<typename Poly>
            key(hmpc::comp::tensor<Poly>, hmpc::comp::tensor<Poly>) -> key<Poly>;

            // hmpc/comp/crypto/lhe/ciphertext.hpp
            template<typename Poly>
            struct ciphertext
            {
                hmpc::comp::tensor<Poly> c0;
                hmpc::comp::tensor<Poly> c1;
            };
        }
    }

    namespace expr
    {
        namespace detail
        {
            enum class kind
            {
                tensor,
                ntt,
                intt,
                add,
                sub,
                mul,
                scalar_mul,
                capability,
                cast,
                unsqueeze,
                cache,
            };

            struct node
            {
                kind k = kind::tensor;
                std::shared_ptr<node const> a, b;
                hmpc::shape shape;
                // leaves: the tensor and how to bind it to a queue
                std::shared_ptr<comp::detail::tensor_state> state;
                std::function<hmpc_amd::comp::tensor(comp::queue&)> bind;
                // samplers and cast targets: modulus words and the value's degree
                std::vector<std::uint32_t> q;
                int log_n = 0;
                int dist = HMPC_CENTERED_BINOMIAL;
                int bits = 0;
                std::vector<std::uint32_t> constant;  // scalar_mul: Montgomery words
            };
            using node_ptr = std::shared_ptr<node const>;
        }

        // an expression whose value type is T (hmpc's value_type); copies share their node
        template<typename T>
        struct expression
        {
            using value_type = T;
            detail::node_ptr node;
            hmpc::shape shape() const { return node->shape; }
        };

        // a constant of the element type: constant_of<V> (compile time) or value(v)
        template<typename Element>
        struct scalar
        {
            Element value;
        };
        template<auto V>
        inline constexpr scalar<std::remove_cvref_t<decltype(V)>> constant_of{V};
        template<typename Element>
        scalar<Element> value(Element const& v)
        {
            return {v};
        }

        template<typename T>
        expression<T> tensor(comp::tensor<T> const& t)
        {
            auto n = std::make_shared<detail::node>();
            n->k = detail::kind::tensor;
            n->shape = t.shape();
            n->state = t.shared_state();
            n->bind = [state = n->state](comp::queue& q) { return comp::detail::bind<T>(q, state); };
            return {std::move(n)};
        }

        namespace detail
        {
            template<typename T>
            expression<T> unary(kind k, node_ptr a, hmpc::shape s)
            {
                auto n = std::make_shared<node>();
                n->k = k;
                n->a = std::move(a);
                n->shape = s;
                return {std::move(n)};
            }
        }

        // NTT(INTT(e)) = e and INTT(NTT(e)) = e (hmpc/expr/number_theoretic_transform.hpp:558-583)
        template<typename T>
        auto number_theoretic_transform(expression<T> const& e)
        {
            static_assert(T::representation == ints::coefficient_representation, "NTT of a coefficient-form polynomial");
            using R = ints::traits::number_theoretic_transform_type_t<T>;
            if (e.node->k == detail::kind::intt)
                return expression<R>{e.node->a};
            return detail::unary<R>(detail::kind::ntt, e.node, e.shape());
        }
        template<typename T>
        auto inverse_number_theoretic_transform(expression<T> const& e)
        {
            static_assert(T::representation == ints::number_theoretic_transform_representation,
                          "INTT of an NTT-form polynomial");
            using R = ints::traits::coefficient_type_t<T>;
            if (e.node->k == detail::kind::ntt)
                return expression<R>{e.node->a};
            return detail::unary<R>(detail::kind::intt, e.node, e.shape());
        }

        // hmpc/expr/cast.hpp: the centred lift of each element into Target's modulus
        template<typename Target, typename T>
        expression<Target> cast(expression<T> const& e)
        {
            auto n = std::make_shared<detail::node>();
            n->k = detail::kind::cast;
            n->a = e.node;
            n->shape = e.shape();
            n->q = hmpc::detail::value_info<Target>::element_type::modulus_words();
            return {std::move(n)};
        }

        // hmpc/expr/cache.hpp:71-82: e may be materialised on its own (the plan decides)
        template<typename T>
        expression<T> cache(expression<T> const& e)
        {
            return detail::unary<T>(detail::kind::cache, e.node, e.shape());
        }

        // hmpc/expr/unsqueeze.hpp: a new axis of extent 1 at `axis` (negative: from the end)
        template<typename T, typename A, A Axis>
        expression<T> unsqueeze(expression<T> const& e, constant<A, Axis>)
        {
            hmpc::shape s = e.shape();
            signed_size pos = static_cast<signed_size>(Axis);
            if (pos < 0)
                pos += static_cast<signed_size>(s.rank) + 1;
            if (pos < 0 || pos > static_cast<signed_size>(s.rank) || s.rank == hmpc::shape::max_rank)
                throw hmpc_amd::error(HMPC_ERR_INVALID_ARGUMENT, "unsqueeze: axis out of range");
            for (auto i = static_cast<signed_size>(s.rank); i > pos; --i)
                s.extents[static_cast<size>(i)] = s.extents[static_cast<size>(i - 1)];
            s.extents[static_cast<size>(pos)] = 1;
            ++s.rank;
            return detail::unary<T>(detail::kind::unsqueeze, e.node, s);
        }

        namespace operators
        {
            namespace detail_ops
            {
                template<typename T>
                expression<T> binary(expr::detail::kind k, expression<T> const& x, expression<T> const& y)
                {
                    auto n = std::make_shared<expr::detail::node>();
                    n->k = k;
                    n->a = x.node;
                    n->b = y.node;
                    n->shape = hmpc::detail::broadcast(x.shape(), y.shape());
                    return {std::move(n)};
                }
                template<typename T>
                expression<T> times(expression<T> const& x, typename hmpc::detail::value_info<T>::element_type const& c)
                {
                    auto n = std::make_shared<expr::detail::node>();
                    n->k = expr::detail::kind::scalar_mul;
                    n->a = x.node;
                    n->shape = x.shape();
                    n->constant = c.words();
                    return {std::move(n)};
                }
            }
            template<typename T>
            expression<T> operator+(expression<T> const& x, expression<T> const& y)
            {
                return detail_ops::binary(expr::detail::kind::add, x, y);
            }
            template<typename T>
            expression<T> operator-(expression<T> const& x, expression<T> const& y)
            {
                return detail_ops::binary(expr::detail::kind::sub, x, y);
            }
            template<typename T>
            expression<T> operator*(expression<T> const& x, expression<T> const& y)
            {
                return detail_ops::binary(expr::detail::kind::mul, x, y);
            }
            template<typename T>
            expression<T> operator*(expression<T> const& x, scalar<typename hmpc::detail::value_info<T>::element_type> const& c)
            {
                return detail_ops::times(x, c.value);
            }
            template<typename T>
            expression<T> operator*(scalar<typename hmpc::detail::value_info<T>::element_type> const& c, expression<T> const& x)
            {
                return detail_ops::times(x, c.value);
            }
        }

        namespace random
        {
            namespace detail
            {
                template<typename T>
                expression<T> capability(hmpc::shape s, int dist, int bits)
                {
                    auto n = std::make_shared<expr::detail::node>();
                    n->k = expr::detail::kind::capability;
                    n->shape = s;
                    n->q = hmpc::detail::value_info<T>::element_type::modulus_words();
                    n->log_n = hmpc::detail::value_info<T>::log_n;
                    n->dist = dist;
                    n->bits = bits;
                    return {std::move(n)};
                }
                // centred binomial width for a variance: 4 x variance random bits (hmpc/random/binomial.hpp)
                template<size V>
                constexpr int bits_of(size_constant<V>)
                {
                    return static_cast<int>(4 * V);
                }
                template<size Num, size Den>
                constexpr int bits_of(rational_size_constant<Num, Den>)
                {
                    static_assert((4 * Num) % Den == 0, "4 x variance must be an integer");
                    return static_cast<int>(4 * Num / Den);
                }
            }

            // queue-capability uniform<mod q> (hmpc/expr/random/uniform.hpp:41-45)
            template<typename T>
            expression<T> uniform(hmpc::shape s)
            {
                return detail::capability<T>(s, HMPC_UNIFORM, 0);
            }
            // queue-capability centred binomial of a variance (hmpc/expr/random/binomial.hpp:94-110)
            template<typename T, typename Variance = size_constant<1>>
            expression<T> centered_binomial(hmpc::shape s, Variance v = {})
            {
                return detail::capability<T>(s, HMPC_CENTERED_BINOMIAL, detail::bits_of(v));
            }
            // queue-capability binomial with `Count` coin flips (hmpc/expr/random/binomial.hpp:42-46)
            template<typename T, size Count>
            expression<T> binomial(hmpc::shape s, size_constant<Count>)
            {
                return detail::capability<T>(s, HMPC_BINOMIAL, static_cast<int>(Count));
            }
        }

        namespace crypto::lhe
        {
            template<typename Poly>
            struct key_expression
            {
                comp::crypto::lhe::key<Poly> k;
            };
            template<typename Poly>
            key_expression<Poly> key(comp::crypto::lhe::key<Poly> const& k)
            {
                return {k};
            }

            // hmpc/expr/crypto/lhe/randomness.hpp:80-143: u, v, w widths and v's distribution
            template<typename T>
            struct randomness_expression
            {
                hmpc::shape shape;
                int bits_u = 2, bits_v = 40, bits_w = 40;
                int v_distribution = HMPC_CENTERED_BINOMIAL;
            };
            template<typename T, typename VU = rational_size_constant<1, 2>, typename VV = size_constant<10>,
                     typename VW = size_constant<10>>
            randomness_expression<T> randomness(hmpc::shape s, VU u = {}, VV v = {}, VW w = {})
            {
                return {s, expr::random::detail::bits_of(u), expr::random::detail::bits_of(v), expr::random::detail::bits_of(w),
                        HMPC_CENTERED_BINOMIAL};
            }
            // v drawn as a signed drowning uniform of width bit_width((bound - 1) << security) + 1
            template<typename T, auto Bound, typename VU = rational_size_constant<1, 2>, typename VW = size_constant<10>>
            randomness_expression<T> drowning_randomness(hmpc::shape s, constant<decltype(Bound), Bound> = {}, VU u = {},
                                                         VW w = {}, int security = 80)
            {
                std::vector<std::uint32_t> bound(std::begin(Bound.data), std::end(Bound.data));
                int bits = 0;
                hmpc_amd::check(hmpc_drown_bits(bound.data(), static_cast<int>(bound.size()), 1, security, &bits));
                return {s, expr::random::detail::bits_of(u), bits, expr::random::detail::bits_of(w), HMPC_DROWN_SIGNED};
            }

            // a ciphertext as an expression: leaves and + - with ciphertexts, * with constants
            template<typename Poly>
            struct ciphertext_expression
            {
                struct part
                {
                    int op = -1;  // -1: leaf
                    comp::crypto::lhe::ciphertext<Poly> leaf;
                    std::shared_ptr<part const> a, b;
                    std::vector<std::uint32_t> constant;
                };
                std::shared_ptr<part const> p;
            };
            template<typename Poly>
            ciphertext_expression<Poly> ciphertext(comp::crypto::lhe::ciphertext<Poly> const& c)
            {
                auto p = std::make_shared<typename ciphertext_expression<Poly>::part>();
                p->leaf = c;
                return {std::move(p)};
            }
            namespace detail
            {
                template<typename Poly>
                ciphertext_expression<Poly> combine(int op, ciphertext_expression<Poly> const& x, ciphertext_expression<Poly> const* y,
                                                    std::vector<std::uint32_t> constant)
                {
                    auto p = std::make_shared<typename ciphertext_expression<Poly>::part>();
                    p->op = op;
                    p->a = x.p;
                    if (y)
                        p->b = y->p;
                    p->constant = std::move(constant);
                    return {std::move(p)};
                }
            }
            template<typename Poly>
            ciphertext_expression<Poly> operator+(ciphertext_expression<Poly> const& x, ciphertext_expression<Poly> const& y)
            {
                return detail::combine<Poly>(HMPC_ADD, x, &y, {});
            }
            template<typename Poly>
            ciphertext_expression<Poly> operator-(ciphertext_expression<Poly> const& x, ciphertext_expression<Poly> const& y)
            {
                return detail::combine<Poly>(HMPC_SUB, x, &y, {});
            }
            template<typename Poly>
            ciphertext_expression<Poly> operator*(scalar<typename Poly::element_type> const& c, ciphertext_expression<Poly> const& x)
            {
                return detail::combine<Poly>(HMPC_MUL, x, nullptr, c.value.words());
            }
            template<typename Poly>
            ciphertext_expression<Poly> operator*(ciphertext_expression<Poly> const& x, scalar<typename Poly::element_type> const& c)
            {
                return detail::combine<Poly>(HMPC_MUL, x, nullptr, c.value.words());
            }

            // hmpc/expr/crypto/lhe/enc.hpp:19-69 over a plaintext tensor of its own modulus p
            template<typename Poly, typename Plaintext>
            struct enc_expression
            {
                comp::crypto::lhe::key<Poly> k;
                comp::tensor<Plaintext> plaintext;
                randomness_expression<ints::traits::coefficient_type_t<Poly>> r;
            };
            template<typename Poly, typename Plaintext, typename R>
            enc_expression<Poly, Plaintext> enc(key_expression<Poly> const& k, expression<Plaintext> const& x,
                                                randomness_expression<R> const& r)
            {
                static_assert(Poly::representation == ints::number_theoretic_transform_representation, "keys are in NTT form");
                if (x.node->k != expr::detail::kind::tensor)
                    throw hmpc_amd::error(HMPC_ERR_INVALID_ARGUMENT, "enc: the plaintext is a tensor expression");
                return {k.k, comp::tensor<Plaintext>(x.node->state), {r.shape, r.bits_u, r.bits_v, r.bits_w, r.v_distribution}};
            }

            // dec<Plaintext>(s, c) (hmpc/expr/crypto/lhe/dec.hpp:13-45)
            template<typename Plaintext, typename Poly>
            struct dec_expression
            {
                comp::tensor<Poly> s;
                ciphertext_expression<Poly> c;
            };
            template<typename Plaintext, typename Poly>
            dec_expression<Plaintext, Poly> dec(expression<Poly> const& s, ciphertext_expression<Poly> const& c)
            {
                if (s.node->k != expr::detail::kind::tensor)
                    throw hmpc_amd::error(HMPC_ERR_INVALID_ARGUMENT, "dec: the key is a tensor expression");
                return {comp::tensor<Poly>(s.node->state), c};
            }
        }
    }

    namespace comp
    {
        // hmpc::comp::queue (hmpc/comp/queue.hpp:229-239, 416-597) on one device: its random-number
        // capability (key, nonce) and the moduli its expressions use, registered on first use
        class queue
        {
        public:
            // device 0, key from the OS entropy source, nonce 0
            queue() : rq_(0) {}
            explicit queue(int device) : rq_(device) {}
            // an explicit 8-word ChaCha20 key, nonce 0
            explicit queue(std::array<std::uint32_t, 8> const& key) : rq_(0, key) {}
            queue(int device, std::array<std::uint32_t, 8> const& key) : rq_(device, key) {}
            queue(queue const&) = delete;
            queue& operator=(queue const&) = delete;

            void wait() { rq_.wait(); }
            hmpc_device_info info() const { return rq_.info(); }
            hmpc_amd::comp::queue& runtime() { return rq_; }

            hmpc_amd::comp::modulus const& modulus_for(std::vector<std::uint32_t> const& q)
            {
                auto it = mods_.find(q);
                if (it == mods_.end())
                    it = mods_.emplace(q, hmpc_amd::comp::modulus(rq_, q)).first;
                return it->second;
            }

            // hmpc::expr::generate_execution_cache(e1, e2, ...) (hmpc/expr/cache.hpp:233-253), which
            // needs the queue here to register moduli and bind tensors: the kernels queue(e1, e2, ...)
            // would run, in order, without running them; entry i has shape {polynomials}
            struct execution_cache_entry
            {
                hmpc::size polynomials;
                hmpc::shape shape() const { return hmpc::shape{polynomials}; }
            };
            struct execution_cache_result
            {
                hmpc::size size = 0;
                std::vector<execution_cache_entry> entries;
                template<typename T, T I>
                execution_cache_entry const& get(constant<T, I>) const { return entries.at(static_cast<hmpc::size>(I)); }
            };
            template<typename... E>
            execution_cache_result execution_cache(expr::expression<E> const&... es)
            {
                std::vector<hmpc_amd::expr::expression> roots;
                memo_map memo;
                (roots.push_back(translate(es.node.get(), memo)), ...);
                execution_cache_result r;
                for (auto const& e : rq_.execution_cache(roots))
                    r.entries.push_back({e.n});
                r.size = r.entries.size();
                return r;
            }

            // queue(e): the value; queue(e1, e2, ...): a tuple. Expression arguments share one execution
            // plan (hmpc's cache order and nonces); enc, ciphertext and dec arguments run after it, in
            // argument order.
            template<typename... E>
                requires(sizeof...(E) >= 1)
            auto operator()(E const&... es)
            {
                std::vector<hmpc_amd::expr::expression> roots;
                std::unordered_map<expr::detail::node const*, hmpc_amd::expr::expression> memo;
                (collect(es, roots, memo), ...);
                std::vector<hmpc_amd::comp::tensor> planned;
                if (!roots.empty())
                    planned = rq_.evaluate(roots);
                std::size_t next = 0;
                if constexpr (sizeof...(E) == 1)
                    return result(es..., planned, next);
                else
                    return std::tuple{result(es, planned, next)...};
            }

        private:
            hmpc_amd::comp::queue rq_;
            std::map<std::vector<std::uint32_t>, hmpc_amd::comp::modulus> mods_;

            using memo_map = std::unordered_map<expr::detail::node const*, hmpc_amd::expr::expression>;

            template<typename T>
            comp::tensor<T> wrap(hmpc_amd::comp::tensor const& t, hmpc::shape const& s)
            {
                auto st = std::make_shared<detail::tensor_state>();
                st->shape = s;
                st->device = t;
                st->owner = this;
                st->device_current = true;
                st->host_current = false;
                return comp::tensor<T>(std::move(st));
            }

            // the library broadcasts an operand by repeating it: numpy broadcasting where the smaller
            // operand's shape is the larger one's trailing axes behind leading ones of extent 1
            static void check_broadcast(hmpc::shape const& x, hmpc::shape const& y)
            {
                hmpc::shape const& big = x.size() >= y.size() ? x : y;
                hmpc::shape const& small = x.size() >= y.size() ? y : x;
                size r = std::max(big.rank, small.rank);
                bool trailing = true;
                for (size i = 0; i < r; ++i)
                {
                    size eb = i < big.rank ? big.extents[big.rank - 1 - i] : 1;
                    size es = i < small.rank ? small.extents[small.rank - 1 - i] : 1;
                    if (trailing && es != eb)
                        trailing = false;
                    if (!trailing && es != 1)
                        throw hmpc_amd::error(HMPC_ERR_INVALID_ARGUMENT, "broadcast of this shape pair is not supported");
                }
            }

            hmpc_amd::expr::expression translate(expr::detail::node const* n, memo_map& memo)
            {
                auto it = memo.find(n);
                if (it != memo.end())
                    return it->second;
                using expr::detail::kind;
                hmpc_amd::expr::expression r;
                switch (n->k)
                {
                case kind::tensor:
                    r = hmpc_amd::expr::tensor(n->bind(*this));
                    break;
                case kind::ntt:
                    r = hmpc_amd::expr::number_theoretic_transform(translate(n->a.get(), memo));
                    break;
                case kind::intt:
                    r = hmpc_amd::expr::inverse_number_theoretic_transform(translate(n->a.get(), memo));
                    break;
                case kind::add:
                case kind::sub:
                case kind::mul:
                {
                    check_broadcast(n->a->shape, n->b->shape);
                    auto op = n->k == kind::add ? hmpc_amd::expr::op::add
                              : n->k == kind::sub ? hmpc_amd::expr::op::sub
                                                  : hmpc_amd::expr::op::mul;
                    r = hmpc_amd::expr::binary(op, translate(n->a.get(), memo), translate(n->b.get(), memo));
                    break;
                }
                case kind::scalar_mul:
                    r = hmpc_amd::expr::operator*(translate(n->a.get(), memo), n->constant);
                    break;
                case kind::capability:
                    r = hmpc_amd::expr::random::capability(modulus_for(n->q), n->dist, n->bits, n->shape.size(), n->log_n);
                    break;
                case kind::cast:
                    r = hmpc_amd::expr::cast(translate(n->a.get(), memo), modulus_for(n->q));
                    break;
                case kind::unsqueeze:
                    // extents only: the same values
                    r = translate(n->a.get(), memo);
                    break;
                case kind::cache:
                    r = hmpc_amd::expr::cache(translate(n->a.get(), memo));
                    break;
                }
                memo.emplace(n, r);
                return r;
            }

            template<typename T>
            void collect(expr::expression<T> const& e, std::vector<hmpc_amd::expr::expression>& roots, memo_map& memo)
            {
                roots.push_back(translate(e.node.get(), memo));
            }
            template<typename Other>
            void collect(Other const&, std::vector<hmpc_amd::expr::expression>&, memo_map&)
            {
            }

            template<typename T>
            comp::tensor<T> result(expr::expression<T> const& e, std::vector<hmpc_amd::comp::tensor>& planned, std::size_t& next)
            {
                return wrap<T>(planned[next++], e.shape());
            }

            template<typename Poly, typename Plaintext>
            crypto::lhe::ciphertext<Poly> result(expr::crypto::lhe::enc_expression<Poly, Plaintext> const& e,
                                                 std::vector<hmpc_amd::comp::tensor>&, std::size_t&)
            {
                auto a = detail::bind<Poly>(*this, e.k.a.shared_state());
                auto b = detail::bind<Poly>(*this, e.k.b.shared_state());
                auto x = detail::bind<Plaintext>(*this, e.plaintext.shared_state());
                hmpc_amd::expr::crypto::lhe::randomness r;
                r.bits_u = e.r.bits_u;
                r.bits_v = e.r.bits_v;
                r.bits_w = e.r.bits_w;
                r.v_distribution = e.r.v_distribution;
                auto ct = rq_(hmpc_amd::expr::crypto::lhe::enc({a, b}, x, r));
                hmpc::shape s = e.plaintext.shape();
                return {wrap<Poly>(ct.c0, s), wrap<Poly>(ct.c1, s)};
            }

            template<typename Poly>
            hmpc_amd::expr::crypto::lhe::ciphertext evaluate_ciphertext(
                typename expr::crypto::lhe::ciphertext_expression<Poly>::part const& p, hmpc::shape& s)
            {
                if (p.op < 0)
                {
                    s = p.leaf.c0.shape();
                    return {detail::bind<Poly>(*this, p.leaf.c0.shared_state()), detail::bind<Poly>(*this, p.leaf.c1.shared_state())};
                }
                auto a = evaluate_ciphertext<Poly>(*p.a, s);
                if (!p.constant.empty())
                    return rq_(hmpc_amd::expr::crypto::lhe::operator*(a, p.constant));
                hmpc::shape sb;
                auto b = evaluate_ciphertext<Poly>(*p.b, sb);
                if (!(sb == s))
                    throw hmpc_amd::error(HMPC_ERR_INVALID_ARGUMENT, "ciphertexts of different shapes");
                return rq_(p.op == HMPC_ADD ? a + b : a - b);
            }

            template<typename Poly>
            crypto::lhe::ciphertext<Poly> result(expr::crypto::lhe::ciphertext_expression<Poly> const& e,
                                                 std::vector<hmpc_amd::comp::tensor>&, std::size_t&)
            {
                hmpc::shape s;
                auto ct = evaluate_ciphertext<Poly>(*e.p, s);
                return {wrap<Poly>(ct.c0, s), wrap<Poly>(ct.c1, s)};
            }

            template<typename Plaintext, typename Poly>
            comp::tensor<Plaintext> result(expr::crypto::lhe::dec_expression<Plaintext, Poly> const& e,
                                           std::vector<hmpc_amd::comp::tensor>&, std::size_t&)
            {
                using info = hmpc::detail::value_info<Plaintext>;
                hmpc::shape s;
                auto ct = evaluate_ciphertext<Poly>(*e.c.p, s);
                auto key = detail::bind<Poly>(*this, e.s.shared_state());
                auto const& p = modulus_for(info::element_type::modulus_words());
                auto out = rq_(hmpc_amd::expr::crypto::lhe::dec(key, ct, p, info::rep));
                return wrap<Plaintext>(out, s);
            }
        };

        namespace detail
        {
            template<typename T>
            hmpc_amd::comp::tensor bind(queue& q, std::shared_ptr<tensor_state> const& sp)
            {
                auto& s = *sp;
                if (s.device_current && s.owner == &q)
                    return s.device;
                if (!s.host_current)
                {
                    // held by another queue's device: through the host
                    s.host = s.device.to_host();
                    s.host_current = true;
                }
                using info = hmpc::detail::value_info<T>;
                auto const& m = q.modulus_for(info::element_type::modulus_words());
                hmpc_amd::comp::tensor t(q.runtime(), m, s.shape.size(), info::log_n, info::rep);
                std::size_t words = t.size() * static_cast<std::size_t>(t.limbs());
                if (s.host.size() != words)
                    s.host.assign(words, 0u);
                t.from_host(s.host.data(), s.host.size());
                s.device = t;
                s.owner = &q;
                s.device_current = true;
                return t;
            }
        }
    }
}
