// ctd_probe.hpp -- VALU throughput probe: the measured integer peak that the
// roofline of the residue kernels is quoted against.
//
// Each lane runs 8 independent dependency chains of one instruction, 16 deep
// per loop trip, with enough waves per SIMD to hide latency; the rate is
// lane-instructions per second.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ctd {

enum ProbeOp : int {
    PROBE_ADD_U32 = 0,
    PROBE_MUL_LO_U32 = 1,
    PROBE_MUL_HI_U32 = 2,
    PROBE_MAD_U64_U32 = 3,
    PROBE_MAD_U32_U24 = 4,
    PROBE_LSHL_ADD_U64 = 5,
    PROBE_CNDMASK_B32 = 6,
    PROBE_ALIGNBIT_B32 = 7,
    PROBE_MIN_U32 = 8,
    PROBE_LSHLREV_B64 = 9,
    PROBE_ADD_F64 = 10,
    PROBE_FMA_F64 = 11,
    PROBE_MUL_F64 = 12,
    PROBE_MAX_F64 = 13,
    PROBE_CMP_EQ_U32 = 14, // compares write a lane mask to scalar registers
    PROBE_CMP_EQ_F64 = 15,
    PROBE_CMP_EQ_U64 = 16,
    PROBE_XOR_B32 = 17,
    PROBE_SUB_U32 = 18,
    PROBE_ADD_F32 = 19,
    PROBE_COUNT = 20
};

constexpr uint32_t PROBE_CHAINS = 8;
constexpr uint32_t PROBE_DEPTH = 16;

#define CTD_P32(INS)                                                                   \
    _Pragma("unroll") for (int d = 0; d < (int)PROBE_DEPTH; ++d) {                     \
        _Pragma("unroll") for (int j = 0; j < (int)PROBE_CHAINS; ++j) asm volatile(INS \
                                                                          : "+v"(a[j]) \
                                                                          : "v"(b));   \
    }

template <int OP>
__global__ void __launch_bounds__(256) valu_probe_kernel(uint32_t *out, uint32_t iters)
{
    const uint32_t b = (blockIdx.x * 2654435761u) | 1u;
    if constexpr (OP >= PROBE_CMP_EQ_U32 && OP <= PROBE_CMP_EQ_U64) {
        // independent compares against lane values; the last mask of each
        // chain is kept live
        uint64_t a[PROBE_CHAINS];
        double f[PROBE_CHAINS];
        uint64_t m[PROBE_CHAINS];
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) {
            a[j] = threadIdx.x * 0x9e3779b97f4a7c15ull + j;
            f[j] = (double)(threadIdx.x + j);
            m[j] = 0;
        }
        const double c = (double)b;
        const uint64_t b64 = b;
        for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
            for (int d = 0; d < (int)PROBE_DEPTH; ++d) {
#pragma unroll
                for (int j = 0; j < (int)PROBE_CHAINS; ++j) {
                    if constexpr (OP == PROBE_CMP_EQ_U32)
                        asm volatile("v_cmp_eq_u32 %0, %1, %2" : "=s"(m[j]) : "v"((uint32_t)a[j]), "v"(b));
                    else if constexpr (OP == PROBE_CMP_EQ_F64)
                        asm volatile("v_cmp_eq_f64 %0, %1, %2" : "=s"(m[j]) : "v"(f[j]), "v"(c));
                    else
                        asm volatile("v_cmp_eq_u64 %0, %1, %2" : "=s"(m[j]) : "v"(a[j]), "v"(b64));
                }
            }
        }
        uint64_t x = 0;
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) x |= m[j];
        if (x == 0x1234567ull) out[0] = (uint32_t)x;
    } else if constexpr (OP == PROBE_ADD_F32) {
        float a[PROBE_CHAINS];
        const float c = 1.0f + 0x1p-20f * (float)(b & 7);
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) a[j] = 1.0f + threadIdx.x * 0x1p-12f + j;
        for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
            for (int d = 0; d < (int)PROBE_DEPTH; ++d) {
#pragma unroll
                for (int j = 0; j < (int)PROBE_CHAINS; ++j) asm volatile("v_add_f32 %0, %0, %1" : "+v"(a[j]) : "v"(c));
            }
        }
        float x = 0;
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) x += a[j];
        if (x == 1234567.0f) out[0] = 1u;
    } else if constexpr (OP >= PROBE_ADD_F64 && OP <= PROBE_MAX_F64) {
        // values stay finite: |a| grows by at most (1 + 2^-20) per step
        const double c = 1.0 + 0x1p-20 * (double)(b & 7);
        double a[PROBE_CHAINS];
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) a[j] = 1.0 + threadIdx.x * 0x1p-12 + j;
        for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
            for (int d = 0; d < (int)PROBE_DEPTH; ++d) {
#pragma unroll
                for (int j = 0; j < (int)PROBE_CHAINS; ++j) {
                    if constexpr (OP == PROBE_ADD_F64)
                        asm volatile("v_add_f64 %0, %0, %1" : "+v"(a[j]) : "v"(c));
                    else if constexpr (OP == PROBE_FMA_F64)
                        asm volatile("v_fma_f64 %0, %0, %1, %1" : "+v"(a[j]) : "v"(c));
                    else if constexpr (OP == PROBE_MUL_F64)
                        asm volatile("v_mul_f64 %0, %0, %1" : "+v"(a[j]) : "v"(c));
                    else
                        asm volatile("v_max_f64 %0, %0, %1" : "+v"(a[j]) : "v"(c));
                }
            }
        }
        double x = 0;
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) x += a[j];
        if (x == 1234567.0) out[0] = 1u;
    } else if constexpr (OP == PROBE_MAD_U64_U32 || OP == PROBE_LSHL_ADD_U64 || OP == PROBE_LSHLREV_B64) {
        uint64_t a[PROBE_CHAINS];
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) a[j] = threadIdx.x * 0x9e3779b97f4a7c15ull + j;
        for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
            for (int d = 0; d < (int)PROBE_DEPTH; ++d) {
#pragma unroll
                for (int j = 0; j < (int)PROBE_CHAINS; ++j) {
                    if constexpr (OP == PROBE_MAD_U64_U32)
                        asm volatile("v_mad_u64_u32 %0, vcc, %1, %1, %0" : "+v"(a[j]) : "v"(b) : "vcc");
                    else if constexpr (OP == PROBE_LSHL_ADD_U64)
                        asm volatile("v_lshl_add_u64 %0, %0, 1, %0" : "+v"(a[j]));
                    else
                        asm volatile("v_lshlrev_b64 %0, 1, %0" : "+v"(a[j]));
                }
            }
        }
        uint64_t x = 0;
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) x ^= a[j];
        if (x == 0x1234567ull) out[0] = (uint32_t)x; // keeps the chains live
    } else {
        uint32_t a[PROBE_CHAINS];
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) a[j] = threadIdx.x * 2246822519u + j;
        for (uint32_t it = 0; it < iters; ++it) {
            if constexpr (OP == PROBE_ADD_U32) CTD_P32("v_add_u32 %0, %0, %1")
            else if constexpr (OP == PROBE_MUL_LO_U32) CTD_P32("v_mul_lo_u32 %0, %0, %1")
            else if constexpr (OP == PROBE_MUL_HI_U32) CTD_P32("v_mul_hi_u32 %0, %0, %1")
            else if constexpr (OP == PROBE_MAD_U32_U24) CTD_P32("v_mad_u32_u24 %0, %0, %1, %0")
            else if constexpr (OP == PROBE_CNDMASK_B32) CTD_P32("v_cndmask_b32 %0, %0, %1, vcc")
            else if constexpr (OP == PROBE_ALIGNBIT_B32) CTD_P32("v_alignbit_b32 %0, %0, %1, 7")
            else if constexpr (OP == PROBE_XOR_B32) CTD_P32("v_xor_b32 %0, %0, %1")
            else if constexpr (OP == PROBE_SUB_U32) CTD_P32("v_sub_u32 %0, %0, %1")
            else CTD_P32("v_min_u32 %0, %0, %1")
        }
        uint32_t x = 0;
        for (uint32_t j = 0; j < PROBE_CHAINS; ++j) x ^= a[j];
        if (x == 0x1234567u) out[0] = x;
    }
}

#undef CTD_P32

} // namespace ctd
