// pfc_internal.h -- declarations shared by the kernel file and the host
// driver of libpfcsieve.so (not part of the public C ABI).
#ifndef PFC_INTERNAL_H
#define PFC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/pfc.h"

namespace pfc {

extern thread_local char g_err[512];
int set_error(int code, const char* msg);
int hip_fail(hipError_t e, const char* what);

int launch_generate(uint64_t lo, uint64_t hi, uint64_t* d_p, uint32_t cap, pfc_ctrl* d_ctrl,
                    hipStream_t st);
int launch_setup(const uint64_t* d_p, const pfc_ctrl* d_ctrl, uint32_t* d_v, uint64_t* d_one,
                 uint64_t* d_ra, uint64_t* d_rb, const pfc_setup_table* a,
                 const pfc_setup_table* b, uint32_t cap, hipStream_t st);
int launch_walk(uint32_t modes, uint32_t j0, uint32_t j1, const uint32_t* d_list,
                const uint32_t* d_comp, uint32_t nbase, const uint64_t* d_p, const uint32_t* d_v,
                const uint64_t* d_one, uint64_t* d_ra, uint64_t* d_rb, pfc_ctrl* d_ctrl,
                pfc_factor* d_fac, uint32_t fcap, uint32_t cap, uint32_t ppl, uint32_t block,
                hipStream_t st);
int launch_check(uint32_t modes, uint32_t nmax, const uint64_t* d_p, const uint32_t* d_v,
                 const uint64_t* d_one, const uint64_t* d_ra, const uint64_t* d_rb,
                 pfc_ctrl* d_ctrl, uint32_t cap, hipStream_t st);
int probe_mul_rates(int cus, double rates[3]);

}  // namespace pfc

#endif  // PFC_INTERNAL_H
