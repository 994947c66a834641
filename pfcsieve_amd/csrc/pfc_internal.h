// pfc_internal.h -- declarations shared by the kernel file and the host
// driver of libpfcsieve.so (not part of the public C ABI).
#ifndef PFC_INTERNAL_H
#define PFC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "../../include/pfc.h"

// One job's device state as the persistent batch kernel sees it.
struct pfc_batch_job {
    uint32_t modes, nmin, nmax, cap;
    uint64_t* P;
    uint32_t* V;
    uint64_t* ONE;
    uint64_t* RA;
    uint64_t* RB;
    pfc_ctrl* ctrl;
    pfc_factor* fac;
    uint32_t fcap;
    uint32_t forms_off;        // pfc_tuning::walk_forms_off of the job
    pfc_setup_table ta;        // start table of res_a
    pfc_setup_table tb;        // start table of res_b (F|C)
    const uint32_t* list;
    const uint32_t* composite;
    const uint32_t* clist;     // C, F|C (optional): composites in [nmin, nmax), increasing
    const uint32_t* cprefix;   //   and the count of them before each composite bitmap word
};

// One work item: 64*K consecutive candidates of a job from index `first`,
// multiplier steps [j0, j1) whose largest multiplier is bmax (0: not known);
// flags bit0 = set up first, bit1 = fold checksum.
struct pfc_batch_item {
    uint32_t job, first, j0, j1, flags, bmax, blocks, pad;   // blocks: of 64*K candidates, from first
};

// The work items of one job segment in one launch: items [start, end) of the
// launch's queue, where start is the previous range's end (0 for the first);
// item start + k covers candidates from k * 64*K.  Ranges are stored in queue
// order with increasing ends.
struct pfc_batch_range {
    uint32_t job, j0, j1, flags, bmax, end;
    uint32_t blocks;           // an item of this range covers blocks x 64*K candidates
    uint32_t pad;
};

namespace pfc {

// Walk mode of a job's mode set, as the walk kernels are instantiated:
// 0 = F, 1 = P, 2 = F|C, 3 = C; -1 for any other set.
__host__ __device__ constexpr int walk_mode_of(uint32_t modes) {
    return modes == PFC_MODE_FACTORIAL ? 0
         : modes == PFC_MODE_PRIMORIAL ? 1
         : modes == (PFC_MODE_FACTORIAL | PFC_MODE_COMPOSITORIAL) ? 2
         : modes == PFC_MODE_COMPOSITORIAL ? 3 : -1;
}

extern thread_local char g_err[512];
int set_error(int code, const char* msg);
int hip_fail(hipError_t e, const char* what);

// lds_segment: integers sieved per generator workgroup (15360, 30720, 61440; 0 = 15360).
int launch_generate(uint64_t lo, uint64_t hi, uint64_t* d_p, uint32_t cap, pfc_ctrl* d_ctrl,
                    hipStream_t st, uint32_t lds_segment);
int launch_head_dev(const void* d_bounds, uint64_t max_span, uint64_t* d_p, uint32_t cap,
                    pfc_ctrl* d_ctrl, hipStream_t st, uint32_t lds_segment);
int launch_setup(const uint64_t* d_p, const pfc_ctrl* d_ctrl, uint32_t* d_v, uint64_t* d_one,
                 uint64_t* d_ra, uint64_t* d_rb, const pfc_setup_table* a,
                 const pfc_setup_table* b, uint32_t cap, hipStream_t st);
// d_clist / d_cprefix: optional composite list of the n-range and its count
// before each bitmap word (both or neither); C and F|C walk it when given.
int launch_walk(uint32_t modes, uint32_t j0, uint32_t j1, const uint32_t* d_list,
                const uint32_t* d_comp, const uint32_t* d_clist, const uint32_t* d_cprefix, uint32_t nbase,
                const uint64_t* d_p, const uint32_t* d_v, uint64_t* d_ra, uint64_t* d_rb, pfc_ctrl* d_ctrl,
                pfc_factor* d_fac, uint32_t fcap, uint32_t cap, uint32_t ppl, uint32_t block,
                uint32_t bmax, uint32_t forms_off, hipStream_t st);
int launch_check(uint32_t modes, uint32_t nmax, const uint64_t* d_p, const uint32_t* d_v,
                 const uint64_t* d_one, const uint64_t* d_ra, const uint64_t* d_rb,
                 pfc_ctrl* d_ctrl, uint32_t cap, hipStream_t st);
// One batch round's kernels (k_batch phases): set-up (phase 4), the walks of
// walk modes F, P, F|C and C (phases 0..3), fold (phase 5), in that order,
// with nblocks[ph] workgroups for phase ph (0: not launched).  d_words: the
// round's device words, read by the kernels: [0 .. 6) the phases' queue
// heads, [6 .. 12) their range counts, [12] the slot count (for
// launch_gather_ctrl).  d_ranges holds the phases' range lists in phase
// order, each with its own item ends.
constexpr uint32_t kBatchWords = 16;
// A fold item covers kFoldBlocks blocks of 64 * K candidates, a set-up item
// one; a walk item as many as make its steps x blocks reach kWalkItemSteps
// (at most kWalkMaxBlocks), so that short rounds (a serving loop's) do not
// spend their time on queue updates.
constexpr uint32_t kFoldBlocks = 16;
constexpr uint32_t kWalkItemSteps = 32768, kWalkMaxBlocks = 64;
int launch_batch(const pfc_batch_job* d_jobs, const pfc_batch_range* d_ranges, uint32_t* d_words,
                 uint32_t ppl, const uint32_t nblocks[6], hipStream_t st);
// Clears the six queue heads of d_words.
int launch_zero_heads(uint32_t* d_words, hipStream_t st);
// d_out[i] = *d_jobs[i].ctrl for i < *d_nslots (cap: the most slots).
int launch_gather_ctrl(const pfc_batch_job* d_jobs, const uint32_t* d_nslots, pfc_ctrl* d_out, uint32_t cap,
                       hipStream_t st);
// d_work: PFC_VERIFY_WORDS words (block partials in [2, PFC_VERIFY_WORDS));
// the two results go to d_out[0], d_out[1] (d_out may be d_work).
int launch_verify_table(uint32_t kind, uint32_t N, const pfc_setup_table* t, uint64_t* d_work,
                        uint64_t* d_out, hipStream_t st);
int probe_mul_rates(int cus, double rates[3]);
int probe_copy(uint64_t bytes, double* ms);

}  // namespace pfc

#endif  // PFC_INTERNAL_H
