// Device helpers shared by the gfx950 kernels that stream operands straight
// into LDS (buffer loads with the LDS bit) and wait on their own loads.
#pragma once
#include <hip/hip_runtime.h>

namespace f4j {

// Pointer type of the LDS argument of __builtin_amdgcn_raw_ptr_buffer_load_lds.
typedef __attribute__((address_space(3))) void lds_void;

// Operand and accumulator vectors of the INT8 MFMA (v_mfma_i32_32x32x32_i8).
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));

}  // namespace f4j

// s_waitcnt with vmcnt(n) only (gfx9 encoding: vmcnt bits 3:0 and 15:14,
// expcnt and lgkmcnt at their maxima, i.e. not waited for).
#define F4J_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (7 << 4) | (15 << 8) | (((n) >> 4) << 14))
