/*
 * hmpc_hip.h — C ABI of the MI355X (gfx950) library for hmpc's data-parallel core:
 * the negacyclic NTT over multi-limb Montgomery integers, the BGV (linear HE) encryption
 * built on it, and the ChaCha20 sampling that feeds both.
 *
 * Conventions shared by every entry point
 *  - Plain C types only: opaque handles, device pointers (void* / uint32_t*), sizes.
 *  - Every call returns an hmpc_status; 0 is success. hmpc_last_error() gives a message
 *    for the calling thread. (Same convention as hmpc-rs's FFI, hmpc-rs/src/ffi/errors.rs:191-216.)
 *  - Polynomial tensors are limb-planar: limb l of element e lives at buf[l * count + e],
 *    count = n * N, element index = b * N + k (hmpc/comp/accessor.hpp:191-197,
 *    hmpc/shape.hpp:204-239). Values are canonical Montgomery residues a*R mod q,
 *    R = 2^(32 L), L = ceil(bit_width(q - 1) / 32) (hmpc/ints/mod.hpp:90-93).
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream). Calls are
 *    asynchronous on that stream unless documented otherwise.
 *  - A context is not thread-safe (like hmpc::comp::queue, hmpc/comp/queue.hpp:165-253):
 *    use one host thread per context.
 */
#ifndef HMPC_HIP_H
#define HMPC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hmpc_status {
    HMPC_OK = 0,
    HMPC_ERR_INVALID_ARGUMENT = 1, /* bad size, null pointer, unsupported N or L */
    HMPC_ERR_INVALID_MODULUS = 2,  /* even, too large, or no 2N-th root of unity */
    HMPC_ERR_OUT_OF_MEMORY = 3,
    HMPC_ERR_DEVICE = 4,           /* HIP runtime error */
    HMPC_ERR_BUDGET = 5,           /* admission refused by the memory budget */
    HMPC_ERR_NOT_READY = 6,        /* request still running */
    HMPC_ERR_CANCELLED = 7
} hmpc_status;

typedef struct hmpc_context_s* hmpc_context;
typedef struct hmpc_modulus_s* hmpc_modulus;

/* ------------------------------------------------------------------ context
 * Replaces hmpc::comp::queue's constructors (hmpc/comp/queue.hpp:229-239): `key` is the
 * 8-word ChaCha20 key of the queue's random-number capability; the 3-word nonce starts at
 * zero. key == NULL selects a key from the OS entropy source (hmpc/detail/random.hpp:10-29). */
int hmpc_context_create(int device, const uint32_t* key, hmpc_context* out);
int hmpc_context_destroy(hmpc_context ctx);
/* Queue RNG state (hmpc/comp/queue.hpp:175-180): key[8], nonce[3]. */
int hmpc_context_get_rng(hmpc_context ctx, uint32_t* key8, uint32_t* nonce3);
int hmpc_context_set_rng(hmpc_context ctx, const uint32_t* key8, const uint32_t* nonce3);
/* hmpc::comp::queue::wait (hmpc/comp/queue.hpp:531-534): synchronise the device. */
int hmpc_context_wait(hmpc_context ctx);

typedef struct hmpc_device_info {
    char name[64];
    char arch[32];
    int compute_units;
    int wavefront_size;
    int clock_khz;
    int memory_clock_khz;
    int memory_bus_width;
    size_t global_memory;
    size_t lds_per_workgroup;
} hmpc_device_info;
/* hmpc::comp::queue::info (hmpc/comp/queue.hpp:536-597) */
int hmpc_context_info(hmpc_context ctx, hmpc_device_info* out);

const char* hmpc_last_error(void);

/* ------------------------------------------------------------------ device memory
 * Stand-ins for hmpc::comp::make_tensor / host_accessor (hmpc/comp/tensor.hpp:86-90,
 * hmpc/comp/accessor.hpp:282-288): raw device buffers plus blocking copies. */
int hmpc_malloc(hmpc_context ctx, size_t bytes, void** out);
int hmpc_free(hmpc_context ctx, void* ptr);
int hmpc_memcpy_h2d(hmpc_context ctx, void* dst, const void* src, size_t bytes);
int hmpc_memcpy_d2h(hmpc_context ctx, void* dst, const void* src, size_t bytes);
int hmpc_memcpy_d2d(hmpc_context ctx, void* dst, const void* src, size_t bytes, void* stream);
/* Pinned host staging for the wire format (hmpc-rs sends limb-planar Montgomery bytes,
 * hmpc/net/queue.hpp:79-80, hmpc/net/structure.hpp:292-300): page-locked host memory and
 * asynchronous copies on a stream, so ciphertexts go to and from the network buffers without a
 * relayout or a pageable bounce. The async copies return at once; synchronize the stream (or
 * record an event) before the host touches the buffer. */
int hmpc_host_alloc(hmpc_context ctx, size_t bytes, void** out);
int hmpc_host_free(hmpc_context ctx, void* ptr);
int hmpc_memcpy_h2d_async(hmpc_context ctx, void* dst, const void* src, size_t bytes, void* stream);
int hmpc_memcpy_d2h_async(hmpc_context ctx, void* dst, const void* src, size_t bytes, void* stream);
int hmpc_stream_create(hmpc_context ctx, void** stream);
int hmpc_stream_destroy(hmpc_context ctx, void* stream);
int hmpc_stream_synchronize(hmpc_context ctx, void* stream);

/* ------------------------------------------------------------------ moduli
 * A registered modulus carries what hmpc computes at compile time for mod<q>
 * (hmpc/ints/mod.hpp:98-127): N', R mod q, R^2 mod q; and, per prepared transform size,
 * the resident twiddle tables of basic_number_theoretic_transform::get_roots
 * (hmpc/expr/number_theoretic_transform.hpp:35-88), cached like the queue's extra tensors
 * (hmpc/comp/queue.hpp:364-381). q must be odd with bit_width(q - 1) <= 288. */
int hmpc_modulus_create(hmpc_context ctx, const uint32_t* q, int q_words, hmpc_modulus* out);
/* A modulus (like a batch or a server) keeps its device: it may be destroyed after its context. */
int hmpc_modulus_destroy(hmpc_modulus mod);
/* limbs L, bit size of q - 1, and constants (each array L words, may be NULL) */
int hmpc_modulus_info(hmpc_modulus mod, int* limbs, int* bit_size, uint32_t* nprime, uint32_t* r_mod_q, uint32_t* r2_mod_q);
/* Kernel family for this modulus: 0 (default) = the fast tile kernels (register-resident rounds
 * between LDS exchanges) where supported, else the portable LDS radix-2 kernels; 1 = always the
 * portable kernels. Results are identical. The fast kernels serve every operation for
 * q = 2^124 + c (c < 2^24, radix-2^26 arithmetic) and Goldilocks 2^64 - 2^32 + 1; for any 4- or
 * 5-limb q with 4q < 2^(32 L) (CIOS Montgomery products, lazy butterflies), e.g. the 125-bit prime of
 * hmpc's examples/ntt.cpp and the 150-bit prime of examples/bgv.cpp, they serve transforms and
 * products (hmpc_ntt, hmpc_polymul) and hmpc_bgv_encrypt at N > 4096 (any plaintext modulus). */
int hmpc_modulus_set_kernels(hmpc_modulus mod, int variant);
/* Schedule of the fast kernels for N = 2^log_n: elements per thread (4 or 8 for 4 and 5 limbs, 8
 * or 16 for 2 limbs; 0 restores the default) and launch structure (persistent = 1: one persistent
 * job kernel in which each workgroup runs whole products / encryptions; 0: one launch per
 * transform pass; -1: unchanged; -2: the default, which is one launch per pass for uniform products
 * and the tile launches in job batches and serving; the default elements per thread are 8 at 4 limbs
 * and 16 at 2 per pass, 4 at 4 limbs on the persistent kernel). Chosen by hmpc_autotune; results
 * are identical. Without a set value, an hmpc_ntt call on 2 limbs too small to give every
 * compute unit a tile runs more threads per tile (4 elements per thread for N <= 4096, 8 above)
 * for latency; hmpc_modulus_get_schedule reports the set or default value. */
int hmpc_modulus_set_schedule(hmpc_modulus mod, int log_n, int elems_per_thread, int persistent);
/* Schedule of fast-path BGV encryption for N = 2^log_n: elements per thread (0 = default) and, for
 * N > 4096, tiles = 1 (a column launch and a row launch over all ciphertexts' tiles), 0 (the
 * persistent job kernel, one workgroup per ciphertext) or -1 (default: tiles, at 8 elements per
 * thread). Chosen by hmpc_autotune_bgv; results are identical. */
int hmpc_modulus_set_bgv_schedule(hmpc_modulus mod, int log_n, int elems_per_thread, int tiles);
/* Current schedule and whether the fast kernels serve this (modulus, N)'s transforms and products. */
int hmpc_modulus_get_schedule(hmpc_modulus mod, int log_n, int* elems_per_thread, int* persistent, int* fast);
/* Autotuner: times batched products of `pairs` synthetic polynomial pairs of N = 2^log_n under
 * every schedule of the fast kernels, installs the fastest with hmpc_modulus_set_schedule and
 * reports it (each output may be NULL; ms = time of one batch). Blocking. No counterpart in
 * hmpc, whose kernels have no schedule. */
int hmpc_autotune(hmpc_modulus mod, int log_n, size_t pairs, void* stream, int* elems_per_thread, int* persistent,
                  float* ms);
/* Autotuner of batched BGV encryption on the fast moduli: times n synthetic encryptions of
 * N = 2^log_n per elements-per-thread choice and, for N > 4096, per launch structure (tiles = 1: a
 * column launch and a row launch over every ciphertext's tiles; 0: the persistent job kernel, one
 * workgroup per ciphertext) and installs the fastest for hmpc_bgv_encrypt (batch and serving jobs
 * use the persistent kernel with the chosen elements per thread). Untuned, hmpc_bgv_encrypt uses
 * the tile launches at 8 elements per thread for N > 4096. The context's RNG state is restored
 * afterwards.
 * Blocking. */
int hmpc_autotune_bgv(hmpc_context ctx, hmpc_modulus mod, int log_n, size_t n, void* stream, int* elems_per_thread,
                      int* tiles, float* ms);
/* Build (once) and keep resident the forward/inverse tables for N = 2^log_n (2 <= log_n <= 20),
 * 2N must divide q - 1. Blocking. */
int hmpc_modulus_prepare(hmpc_modulus mod, int log_n);
/* Primitive 2N-th root phi (or phi^-1) in Montgomery form, L words on the host
 * (hmpc/ints/num/theory/root_of_unity.hpp:19-54). */
int hmpc_modulus_root(hmpc_modulus mod, int log_n, int inverse, uint32_t* out);
/* Device pointer of the resident table [L][N] (read only). */
int hmpc_modulus_twiddles(hmpc_modulus mod, int log_n, int inverse, const uint32_t** out);

/* ------------------------------------------------------------------ NTT
 * Forward: number_theoretic_transform_expression (hmpc/expr/number_theoretic_transform.hpp:
 * 239-394): natural-order coefficients -> evaluations in hmpc's bit-reversed order,
 * A[j] = sum_i a_i phi^((2 brv(j) + 1) i). Inverse: the exact inverse (:396-550).
 * n polynomials of N = 2^log_n coefficients; `in` may equal `out`.
 * The fused path keeps stages in LDS (one pass for small N, two passes otherwise). */
int hmpc_ntt(hmpc_modulus mod, int log_n, size_t n, int inverse, const uint32_t* in, uint32_t* out, void* stream);
/* hmpc's own launch structure: stage 0 into a transposed [N][n] scratch, one launch per
 * middle stage, last stage back to [n][N] (K2-K7 of the reference). `scratch` holds
 * L * n * N words (NULL: allocated and freed by the call). */
int hmpc_ntt_plain(hmpc_modulus mod, int log_n, size_t n, int inverse, const uint32_t* in, uint32_t* out,
                   uint32_t* scratch, void* stream);

/* ------------------------------------------------------------------ point-wise ops
 * add/subtract/multiply_expression executed element-wise (hmpc/expr/binary_expression.hpp:
 * 8-163 via hmpc/comp/queue.hpp:387-412). b is broadcast with period b_count (b index =
 * e % b_count; b_count == count for no broadcast, the unsqueeze of hmpc/expr/unsqueeze.hpp). */
typedef enum hmpc_pointwise_op { HMPC_ADD = 0, HMPC_SUB = 1, HMPC_MUL = 2 } hmpc_pointwise_op;
int hmpc_pointwise(hmpc_modulus mod, int op, size_t count, const uint32_t* a, const uint32_t* b, size_t b_count,
                   uint32_t* out, void* stream);
/* out = a * c for a Montgomery constant c (L host words): value/constant leaves
 * (hmpc/expr/value.hpp, hmpc/expr/constant.hpp). */
int hmpc_pointwise_scalar(hmpc_modulus mod, int op, size_t count, const uint32_t* a, const uint32_t* c_host,
                          uint32_t* out, void* stream);
/* Homomorphic operations on BGV ciphertexts (c0, c1) of `count` elements in NTT form, both
 * parts in one launch: ciphertext_expression operator+ / operator- and operator* with an
 * expression (hmpc/expr/crypto/lhe/ciphertext.hpp:48-81). hmpc_ciphertext_op: out_p = a_p op b_p
 * (op = HMPC_ADD / HMPC_SUB with another ciphertext b; HMPC_MUL with b0 = b1 = a plaintext
 * polynomial in NTT form, broadcast with period b_count); hmpc_ciphertext_scalar: out_p = a_p op c
 * for a Montgomery constant c (L host words). Outputs may alias inputs. */
int hmpc_ciphertext_op(hmpc_modulus mod, int op, size_t count, const uint32_t* a0, const uint32_t* a1, const uint32_t* b0,
                       const uint32_t* b1, size_t b_count, uint32_t* out0, uint32_t* out1, void* stream);
int hmpc_ciphertext_scalar(hmpc_modulus mod, int op, size_t count, const uint32_t* a0, const uint32_t* a1,
                           const uint32_t* c_host, uint32_t* out0, uint32_t* out1, void* stream);
/* out = a * b + c element-wise, fused (examples/fma.cpp:56-67) */
int hmpc_fma(hmpc_modulus mod, size_t count, const uint32_t* a, const uint32_t* b, const uint32_t* c, uint32_t* out,
             void* stream);
/* Full reduction of `count` elements to one (out: L device words): op = HMPC_ADD (hmpc::expr::sum)
 * or HMPC_MUL (hmpc::expr::product), hmpc/expr/reduce.hpp:229-274 with sycl::plus / multiplies
 * over mod q. The result is exact whatever the reduction order. */
int hmpc_reduce(hmpc_modulus mod, int op, size_t count, const uint32_t* in, uint32_t* out, void* stream);
/* Matrix-vector product over mod q (hmpc/expr/matrix_vector_product.hpp:92-133): for b < batch,
 * out[b][r] = sum_c M[b][r][c] * v[b][c], with M row-major [batch][rows][cols] and v [batch][cols]
 * (limb-planar); broadcast_matrix / broadcast_vector = 1 uses one matrix / vector for every b
 * (hmpc's common_extent broadcast). out is [batch][rows]. */
int hmpc_matvec(hmpc_modulus mod, size_t batch, size_t rows, size_t cols, const uint32_t* M, int broadcast_matrix,
                const uint32_t* v, int broadcast_vector, uint32_t* out, void* stream);
/* unsigned 32-bit words -> mod q (hmpc/ints/mod.hpp:165-173, the lift used by examples/bgv.cpp:66-69) */
int hmpc_from_u32(hmpc_modulus mod, size_t count, const uint32_t* words, uint32_t* out, void* stream);
/* Montgomery -> canonical integers (hmpc/ints/mod.hpp:356-361), limb-planar in and out */
int hmpc_to_unsigned(hmpc_modulus mod, size_t count, const uint32_t* in, uint32_t* out, void* stream);

/* ------------------------------------------------------------------ tensor cipher
 * hmpc::expr::crypto::enc / dec with cipher(key, nonce) (hmpc/expr/crypto/cipher.hpp:63-318,
 * hmpc/crypto/cipher.hpp:30-45): element i of a tensor of `limbs`-word values takes words
 * [i limbs, (i + 1) limbs) of one ChaCha20 keystream (key8, nonce3, block counter starting at
 * counter0; hmpc's expressions use 0) and is XORed with them limb by limb. Encryption reads the
 * limb-planar tensor u32[limbs][count] and writes the element-major word tensor
 * u32[count][limbs] (hmpc's result shape {..., limbs}); decryption the reverse. The block
 * counter is 32 bits wide and wraps, as in hmpc. */
int hmpc_cipher_encrypt(hmpc_context ctx, int limbs, size_t count, const uint32_t* key8, const uint32_t* nonce3,
                        uint32_t counter0, const uint32_t* in, uint32_t* out, void* stream);
int hmpc_cipher_decrypt(hmpc_context ctx, int limbs, size_t count, const uint32_t* key8, const uint32_t* nonce3,
                        uint32_t counter0, const uint32_t* in, uint32_t* out, void* stream);

/* ------------------------------------------------------------------ sampling
 * Explicit-key streams (hmpc/expr/random/number_generator.hpp:31-74): element e of a tensor
 * with `count` elements uses ChaCha20(key, nonce = e + count * offset (96-bit), counter 0).
 *  uniform: uniform<mod q> with 80-bit drowning (hmpc/random/uniform.hpp:34-66)
 *  cbd:     centred binomial with `bits` = 4 * variance random bits (hmpc/random/binomial.hpp)
 *  words:   the first 32-bit word of each element's stream (BGV messages in [0, 2^32)) */
int hmpc_sample_uniform(hmpc_modulus mod, const uint32_t* key8, uint64_t offset, size_t count, uint32_t* out, void* stream);
int hmpc_sample_cbd(hmpc_modulus mod, const uint32_t* key8, uint64_t offset, int bits, size_t count, uint32_t* out, void* stream);
int hmpc_sample_words(hmpc_context ctx, const uint32_t* key8, uint64_t offset, size_t count, uint32_t* out, void* stream);
/* Width of a capability-stream draw that selects the uniform ternary distribution over {-1, 0, 1}
 * instead of a centred binomial. hmpc has no such sampler: this is the "ternary uniform" secret and
 * ephemeral key of the common BGV parameter sets. Each draw takes two words w0, w1 of the stream
 * (as a centred binomial of more than 32 bits does) and gives floor(3 (w1 2^32 + w0) / 2^64) - 1,
 * within 2^-63 of uniform. Accepted wherever a capability width is (hmpc_sample_cbd_ntt, the BGV
 * entry points' bits_u / bits_v / bits_w, HMPC_CENTERED_BINOMIAL in hmpc_sample_capability);
 * explicit-key samplers (hmpc_sample_cbd) follow hmpc and reject it. */
#define HMPC_BITS_TERNARY 0
/* Capability-path centred binomial consumed by a forward NTT (hmpc/comp/queue.hpp:185-253,
 * hmpc/expr/number_theoretic_transform.hpp:320-333): uses the context's key and current
 * nonce, then advances the nonce by n * N. Output is NTT(sample) (evaluation domain). */
int hmpc_sample_cbd_ntt(hmpc_context ctx, hmpc_modulus mod, int bits, int log_n, size_t n, uint32_t* out, void* stream);

/* ------------------------------------------------------------------ fused batches
 * Batched negacyclic products out_i = INTT(NTT(a_i) o NTT(b_i)), i < n, in one fused pass
 * sequence (LDS-resident stages, no per-stage launches). */
int hmpc_polymul(hmpc_modulus mod, int log_n, size_t n, const uint32_t* a, const uint32_t* b, uint32_t* out, void* stream);
/* hmpc_polymul without any allocation: no counterpart in hmpc. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid allocation). The
 * same results through the per-pass kernels, with their workspace (2 x 5 (4 limbs) or 2 x L words per
 * coefficient for N > 4096; none for N <= 4096) in a fixed 64 MiB device array that the library holds
 * on every GPU it runs on until the process ends. Fails with HMPC_ERR_INVALID_ARGUMENT when the
 * workspace would exceed 64 MiB (e.g. more than 12 pairs at N = 2^16 with 4 limbs) or when the tables
 * of N were not prepared (hmpc_modulus_prepare). Every call on a GPU shares that GPU's array: calls
 * that overlap on one GPU (other streams or threads) race. */
int hmpc_polymul_noalloc(hmpc_modulus mod, int log_n, size_t n, const uint32_t* a, const uint32_t* b, uint32_t* out,
                         void* stream);

/* BGV public-key encryption of n messages (hmpc/expr/crypto/lhe/enc.hpp:19-69 with the
 * default randomness of hmpc/expr/crypto/lhe/randomness.hpp:80-110):
 *   u, v, w  centred binomial with bits_u/v/w random bits (hmpc's defaults: 2, 40, 40 —
 *            variances 1/2, 10, 10; width HMPC_BITS_TERNARY draws uniform ternary values;
 *            eta = 10 in the usual sense is 20 bits), drawn on the context's capability stream with nonces
 *            nonce, nonce + nN, nonce + 2nN; the context nonce advances by 3nN
 *   c0 = A*NTT(u) + NTT(v)*p + NTT(m),  c1 = B*NTT(u) + NTT(w)*p
 * pk_a, pk_b: [L][N] in NTT form; messages: n*N u32 words lifted as unsigned values;
 * p_mont: plaintext modulus as a Montgomery constant (L host words); c0, c1: [L][n*N].
 * Kernels: for q = 2^124 + c or Goldilocks with p < 2^40, fused launches over tiles or the
 * persistent job kernel (hmpc_modulus_set_bgv_schedule); for other 4- and 5-limb q with 4q < R at
 * N > 4096, the tile launches with p as a Montgomery constant; otherwise the portable kernels. */
int hmpc_bgv_encrypt(hmpc_context ctx, hmpc_modulus mod, int log_n, size_t n, const uint32_t* pk_a,
                     const uint32_t* pk_b, const uint32_t* messages, const uint32_t* p_mont, int bits_u,
                     int bits_v, int bits_w, uint32_t* c0, uint32_t* c1, void* stream);
/* Same result, hmpc's launch structure: per ciphertext, per-stage NTT launches and
 * separate element-wise kernels (the plain one-launch-sequence-per-ciphertext path). */
int hmpc_bgv_encrypt_plain(hmpc_context ctx, hmpc_modulus mod, int log_n, size_t n, const uint32_t* pk_a,
                           const uint32_t* pk_b, const uint32_t* messages, const uint32_t* p_mont, int bits_u,
                           int bits_v, int bits_w, uint32_t* c0, uint32_t* c1, void* stream);
/* hmpc/expr/crypto/lhe/dec.hpp:13-45 with plaintext modulus 2^32: x = INTT(c0 - s*c1),
 * centred lift, low 32 bits -> messages (n*N u32). */
int hmpc_bgv_decrypt(hmpc_modulus mod, int log_n, size_t n, const uint32_t* s_ntt, const uint32_t* c0,
                     const uint32_t* c1, uint32_t* messages, void* stream);

/* ------------------------------------------------------------------ general plaintext path
 * Cross-modulus conversion: hmpc::expr::cast<mod<Q>> of a mod<P> tensor (hmpc/expr/cast.hpp:42-45)
 * with the centred lift of hmpc/ints/mod.hpp:259-307: the canonical integer v of each element of
 * `from` is taken as v - P when v > floor(P / 2), then reduced into `to` (Montgomery form).
 * Limb-planar in ([L_from][count]) and out ([L_to][count]); `in` may equal `out`. */
int hmpc_cast(hmpc_modulus from, hmpc_modulus to, size_t count, const uint32_t* in, uint32_t* out, void* stream);

/* Capability-path sampling (hmpc/expr/random/uniform.hpp, hmpc/expr/random/binomial.hpp drawn through
 * the queue's random-number capability, hmpc/comp/queue.hpp:185-253): the context's key and current
 * nonce, then the nonce advances by n N (N = 2^log_n). The capability of linear index i is the
 * ChaCha20 stream (key, nonce, counter = low32(i)). Output: n polynomials of N coefficients mod q
 * in coefficient form, in the order the consuming kernel reads them:
 *   HMPC_LAYOUT_ELEMENT  element-wise kernels: coefficient e from stream e (hmpc/comp/queue.hpp:400-409;
 *                        log_n may be 0 for plain tensors of n elements)
 *   HMPC_LAYOUT_NTT      a following forward NTT: stream b N + t gives t, then t + N/2
 *                        (hmpc/expr/number_theoretic_transform.hpp:320-329)
 *   HMPC_LAYOUT_INTT     a following inverse NTT: stream b N + 2t gives 2t, then 2t + 1 (:480-489)
 * Distributions (bits):
 *   HMPC_UNIFORM            uniform<mod q> with 80-bit drowning (hmpc/random/uniform.hpp:61-66; bits unused)
 *   HMPC_CENTERED_BINOMIAL  bits = 4 x variance, even, 2..64 (hmpc/random/binomial.hpp:88-128), or
 *                        HMPC_BITS_TERNARY for uniform ternary values (not in hmpc)
 *   HMPC_BINOMIAL           popcount of bits = 1..256 random bits (hmpc/random/binomial.hpp:30-58)
 *   HMPC_DROWN_SIGNED / _UNSIGNED  drown_uniform of width bits = 1..384 (hmpc/random/uniform.hpp:34-47;
 *                        hmpc_drown_bits gives the width for a bound) */
typedef enum hmpc_distribution {
    HMPC_UNIFORM = 0,
    HMPC_CENTERED_BINOMIAL = 1,
    HMPC_BINOMIAL = 2,
    HMPC_DROWN_SIGNED = 3,
    HMPC_DROWN_UNSIGNED = 4
} hmpc_distribution;
typedef enum hmpc_layout { HMPC_LAYOUT_ELEMENT = 0, HMPC_LAYOUT_NTT = 1, HMPC_LAYOUT_INTT = 2 } hmpc_layout;
int hmpc_sample_capability(hmpc_context ctx, hmpc_modulus mod, int distribution, int bits, int layout, int log_n, size_t n,
                           uint32_t* out, void* stream);
/* width of drown_uniform's draw for a bound (bound_words u32 words, little-endian):
 * bit_width((bound - 1) << security) + is_signed (hmpc/random/uniform.hpp:39-41; hmpc's default
 * statistical security is 80, hmpc/config.hpp:102) */
int hmpc_drown_bits(const uint32_t* bound, int bound_words, int is_signed, int security, int* bits);

/* BGV public-key encryption of a plaintext over mod p (hmpc/expr/crypto/lhe/enc.hpp:19-69):
 *   message = NTT_q(cast<mod q>(INTT_p(x))) for an NTT-form plaintext (plaintext_ntt = 1), or
 *             NTT_q(cast<mod q>(x)) for a coefficient-form one (centred lift, hmpc_cast)
 *   u = NTT(CBD(bits_u)), v = NTT(v_distribution(bits_v)), w = NTT(CBD(bits_w)) on the context's
 *   capability stream with nonces nonce, nonce + nN, nonce + 2nN (hmpc's cache order); the nonce
 *   advances by 3nN. v_distribution = HMPC_CENTERED_BINOMIAL is lhe::randomness() (hmpc's default
 *   bits 2, 40, 40; randomness.hpp:80-110); HMPC_DROWN_SIGNED with bits_v = hmpc_drown_bits(bound,
 *   signed, 80) is lhe::drowning_randomness(bound) (randomness.hpp:112-143)
 *   c0 = A u + v p + message,  c1 = B u + w p  (p = the plaintext modulus as a mod q constant)
 * pk_a, pk_b: [L_q][N] (NTT form); plaintext: [L_p][n N] mod p; c0, c1: [L_q][n N]. Any odd moduli
 * of 1..9 limbs with 2N | q - 1 (and 2N | p - 1 for an NTT-form plaintext). */
int hmpc_bgv_encrypt_plaintext(hmpc_context ctx, hmpc_modulus mod_q, hmpc_modulus mod_p, int log_n, size_t n,
                               const uint32_t* pk_a, const uint32_t* pk_b, const uint32_t* plaintext, int plaintext_ntt,
                               int bits_u, int v_distribution, int bits_v, int bits_w, uint32_t* c0, uint32_t* c1,
                               void* stream);
/* dec<plaintext> (hmpc/expr/crypto/lhe/dec.hpp:13-45): x = INTT_q(c0 - s c1), cast to mod p with the
 * centred lift, then NTT_p for an NTT-form plaintext (plaintext_ntt = 1). plaintext: [L_p][n N]. */
int hmpc_bgv_decrypt_plaintext(hmpc_modulus mod_q, hmpc_modulus mod_p, int log_n, size_t n, const uint32_t* s_ntt,
                               const uint32_t* c0, const uint32_t* c1, int plaintext_ntt, uint32_t* plaintext, void* stream);

/* ------------------------------------------------------------------ job batches
 * Many independent jobs of uneven degree (2^12 .. 2^16 and beyond) and modulus in persistent
 * launches: one launch per (modulus, schedule) class, in which each workgroup claims jobs
 * longest first and runs each from inputs to outputs. Jobs at N > 4096 instead run as launches
 * over a table of all their tiles (BGV: column pass, row pass; products: forward columns, rows,
 * inverse columns), unless the tuner chose the persistent kernel for that degree
 * (hmpc_autotune_bgv, hmpc_autotune). Each job's polynomials are single
 * polynomials ([L][N] buffers). hmpc has no batch object: a job list stands for the sequence of
 * queue(...) calls it replaces (hmpc/comp/queue.hpp:416-529), with the same results.
 * Moduli must be served by the fast kernels (hmpc_modulus_get_schedule: fast = 1). */
typedef struct hmpc_batch_s* hmpc_batch;
int hmpc_batch_create(hmpc_context ctx, hmpc_batch* out);
int hmpc_batch_destroy(hmpc_batch batch);
/* out = INTT(NTT(a) o NTT(b)) */
int hmpc_batch_add_polymul(hmpc_batch batch, hmpc_modulus mod, int log_n, const uint32_t* a, const uint32_t* b,
                           uint32_t* out);
/* one BGV encryption (hmpc_bgv_encrypt with n = 1): the context nonce is taken now and
 * advanced by 3N, as a queue(enc(...)) call would; p_mont must encode an integer < 2^40 */
int hmpc_batch_add_bgv_encrypt(hmpc_batch batch, hmpc_modulus mod, int log_n, const uint32_t* pk_a, const uint32_t* pk_b,
                               const uint32_t* messages, const uint32_t* p_mont, int bits_u, int bits_v, int bits_w,
                               uint32_t* c0, uint32_t* c1);
/* Run every job. The first run freezes the batch (job tables uploaded, schedules fixed); with
 * use_graph the launches are captured once into a hipGraph and later runs replay it (same
 * inputs, nonces and outputs). Asynchronous on `stream`. */
int hmpc_batch_run(hmpc_batch batch, int use_graph, void* stream);
/* jobs, launch groups after the first run (persistent-kernel classes plus tile-launch chunks of
 * jobs at N > 4096), device bytes held (tables, counters, scratch) */
int hmpc_batch_info(hmpc_batch batch, size_t* jobs, size_t* classes, size_t* device_bytes);

/* ------------------------------------------------------------------ serving
 * A request loop on a native worker thread that admits encryption and product requests as they
 * arrive (continuous batching). Whenever one of `slots` streams is free, queued requests whose
 * outputs fit the HBM budget are taken in arrival order into one set of launches on it: tile
 * launches over a table of every ciphertext or product pair at two-pass degrees (N > 4096, split by
 * degree, so that each runs the kernels with its pass lengths fixed at compile time), and one
 * persistent launch per (modulus, elements per thread) for N <= 4096 and degrees tuned to the
 * persistent kernel (at most 16 such classes per set; a request that would add a 17th waits for the
 * next). Launches on different slots overlap on the device. Keys, twiddles, messages and outputs
 * stay in HBM. Outputs are allocated by the server from a pool whose device memory stays within
 * `budget_bytes` and are held until hmpc_server_release. A request larger than the budget fails with
 * HMPC_ERR_BUDGET at submission; others wait for memory. A request whose launch or output
 * allocation fails fails alone (hmpc_server_poll / _wait report why); later requests run.
 * Randomness: the server's ChaCha key is the context key at creation; an encryption request of k
 * messages with nonce n gives exactly the ciphertexts of hmpc_bgv_encrypt(n = k) on a context with
 * that key and nonce (u, v, w on nonces n, n + kN, n + 2kN), whatever else shares the GPU. Input
 * buffers must stay valid until the request is done, and moduli and keys until the server is
 * destroyed. A request's launch parameters (tables, schedule) are taken from its modulus at
 * submission, on the client's thread: the worker never reads a modulus's state, so the client may
 * use the modulus meanwhile. Thread safety: one client thread; the calls below lock the server. No
 * counterpart in hmpc, whose queue executes synchronously per call (hmpc/comp/queue.hpp:416-529). */
typedef struct hmpc_server_s* hmpc_server;
int hmpc_server_create(hmpc_context ctx, size_t budget_bytes, int slots, hmpc_server* out);
/* drops queued requests, drains running ones, frees all server memory */
int hmpc_server_destroy(hmpc_server server);
/* resident public key (caller-owned device buffers [L][N] in NTT form) and its sampler widths */
int hmpc_server_add_key(hmpc_server server, hmpc_modulus mod, int log_n, const uint32_t* pk_a, const uint32_t* pk_b,
                        const uint32_t* p_mont, int bits_u, int bits_v, int bits_w, int* key_id);
/* k messages (device, k*N u32 words) under key_id; outputs c0, c1 each [L][k*N] */
int hmpc_server_submit_encrypt(hmpc_server server, int key_id, size_t k, const uint32_t* messages, const uint32_t* nonce3,
                               uint64_t* request);
/* n products INTT(NTT(a_i) o NTT(b_i)) of [L][n*N] inputs; output [L][n*N] (returned as c0) */
int hmpc_server_submit_polymul(hmpc_server server, hmpc_modulus mod, int log_n, size_t n, const uint32_t* a,
                               const uint32_t* b, uint64_t* request);
/* state 0 queued, 1 running, 2 done, 3 failed; latency = submission to completion (us, when done) */
int hmpc_server_poll(hmpc_server server, uint64_t request, int* state, double* latency_us);
/* blocks until the request is done or failed; fails with HMPC_ERR_BUDGET instead of waiting forever
 * when nothing runs and the queue's head waits for memory held by unreleased outputs */
int hmpc_server_wait(hmpc_server server, uint64_t request);
/* device pointers of a finished request's outputs (c1 is NULL for products), valid until release */
int hmpc_server_result(hmpc_server server, uint64_t request, uint32_t** c0, uint32_t** c1);
/* returns the outputs' memory to the budget (or cancels a queued request) */
int hmpc_server_release(hmpc_server server, uint64_t request);
/* requests finished since the last call (up to max), with their latencies in us (-1: failed) */
int hmpc_server_completed(hmpc_server server, uint64_t* requests, double* latency_us, size_t max, size_t* count);
int hmpc_server_stats(hmpc_server server, size_t* queued, size_t* running, size_t* bytes_in_use, uint64_t* launches,
                      uint64_t* jobs);

/* ------------------------------------------------------------------ kernel timing
 * Per-kernel durations of the fast kernels, measured with HIP events recorded around every launch
 * on the stream it runs on (not inside graph capture). Enabling (or disabling) clears the
 * records. hmpc_context_timing reports the index-th kernel name seen so far (name_len bytes with
 * the terminating zero), its launch count and average duration in ms; it synchronises on the
 * recorded events and fails with HMPC_ERR_INVALID_ARGUMENT past the last kernel. No counterpart
 * in hmpc (profiling of the SYCL queue is left to the SYCL runtime there). */
int hmpc_context_set_timing(hmpc_context ctx, int enabled);
int hmpc_context_timing(hmpc_context ctx, int index, char* name, size_t name_len, size_t* launches, double* avg_ms);

/* ------------------------------------------------------------------ timing helpers
 * HIP events on a given stream (used by benchmarks to time kernels on the stream they run on). */
int hmpc_event_create(hmpc_context ctx, void** event);
int hmpc_event_destroy(hmpc_context ctx, void* event);
int hmpc_event_record(hmpc_context ctx, void* event, void* stream);
int hmpc_event_elapsed_ms(hmpc_context ctx, void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* HMPC_HIP_H */
