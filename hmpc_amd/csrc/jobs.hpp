// Job descriptors of the persistent fast-path kernel (host and device share this layout).
//
// A job is one independent unit of work that one workgroup runs from inputs in HBM to outputs in
// HBM: a negacyclic polynomial product INTT(NTT(a) o NTT(b)) or one BGV public-key encryption
// (sampling, three forward transforms, combination with the public key). Jobs of one launch may
// differ in degree; they share the arithmetic policy (the modulus form) and the schedule.
#pragma once

#include "chacha.hpp"
#include "modarith.hpp"

namespace hmpc_amd
{
    enum job_type : u32
    {
        job_polymul = 0,
        job_bgv = 1,
    };

    struct job_desc
    {
        u32 type;
        u32 log_n;
        u32 poly;          // bgv: capability polynomial index, ChaCha counter = low32(poly * N + k)
        int bits[3];       // bgv: centred-binomial bits (4 x variance) of u, v, w
        chacha_key key;    // bgv: queue key
        u32 nonce[3][3];   // bgv: nonces of the u, v, w streams
        u64 p_int;         // bgv: plaintext modulus as an integer (< 2^40)
        u32 p_mont[max_limbs];  // bgv on the CIOS policy: the plaintext modulus as a Montgomery constant
        u32 const* twf;    // forward / inverse twiddles of the policy's table format for log_n
        u32 const* twi;
        u32 const* twfx;   // radix-2^26 (non-null there): the first entries of the forward / inverse
        u32 const* twix;   // tables in expanded form (r26_mont_tw)
        // polymul: in0 = a, in1 = b (planes of in_count, polynomial at element in_off)
        // bgv:     in0 = public key A, in1 = B (planes of in_count = N, in_off = 0)
        u32 const* in0;
        u32 const* in1;
        u64 in_count, in_off;
        // polymul: out0 = product; bgv: out0 = c0, out1 = c1
        u32* out0;
        u32* out1;
        u64 out_count, out_off;
        u32 const* msg;    // bgv: messages as u32 words, msg[msg_off + k]
        u64 msg_off;
        u64 ws_off;        // tile launches over a job table: the job's workspace region (words)
    };

    // Words per element of the fast kernels' private workspaces: the raw radix-2^26 limbs (5)
    // for 4-limb moduli, the data format otherwise.
    inline int ws_limbs(int io_limbs) { return io_limbs == 4 ? 5 : io_limbs; }

    // Scratch words one workgroup needs for a job of degree 2^log_n: the transformed first
    // operand (NTT(a), NTT(u)) of whole-tile degrees, and the column-pass results of every
    // polynomial of the two-pass transform.
    inline u64 job_scratch_words(u32 type, int log_n, int io_limbs)
    {
        u64 polys = log_n <= 12 ? 1 : (type == job_bgv ? 3 : 2);
        return polys * static_cast<u64>(ws_limbs(io_limbs)) << log_n;
    }
}
