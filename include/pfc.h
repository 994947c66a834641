/*
 * pfc.h -- C ABI of the MI355X (gfx950) PFCSieve sieve library, libpfcsieve.so.
 *
 * The library finds 64-bit prime factors p of n!+-1, n#+-1 and n!/n#+-1 for
 * n in [nmin, nmax) and p in [pmin, pmax), with the reference's candidate set,
 * factor records and BOINC-quorum checksum, bit for bit.
 *
 * It replaces the reference's OpenCL boundary (simpleCL.h:65-119, which
 * cl_sieve.cpp drives) and the device half of its search driver
 * (cl_sieve.cpp:1328-1761).  Plain pointers and sizes only; every call returns
 * an int status (PFC_OK == 0) and never exits the process: the reference's
 * exit(EXIT_FAILURE) points (simpleCL.c:347, cl_sieve.cpp:462-486, 1061-1066)
 * become status codes, with a message in pfc_last_error().
 *
 * Threading: a pfc_device may be shared by threads; each pfc_job owns one HIP
 * stream and must be driven by one thread at a time.
 */
#ifndef PFC_H
#define PFC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define PFC_OK 0
#define PFC_ERR_INVALID (-1)         /* bad argument or mode set (cl_sieve.cpp:600-641) */
#define PFC_ERR_HIP (-2)             /* HIP runtime failure (simpleCL.c:347 analogue) */
#define PFC_ERR_SIEVE (-3)           /* primorial nmax is prime: "CPU sieve failure" (cl_sieve.cpp:1061-1066) */
#define PFC_ERR_FACTOR_OVERFLOW (-4) /* factor records exceeded capacity (cl_sieve.cpp:484-487) */
#define PFC_ERR_VALIDATION (-5)      /* device self-check failed: walk incomplete (check.cl:62-73,
                                        cl_sieve.cpp:476-480), a stored reciprocal, or a start table */
#define PFC_ERR_VERIFY (-6)          /* CPU factor verification failed (cl_sieve.cpp:509-529) */
#define PFC_ERR_NOMEM (-7)           /* host or device allocation failed */
#define PFC_ERR_IO (-8)              /* factors.txt / checkpoint file I/O */
#define PFC_ERR_CAPACITY (-9)        /* a size exceeds a documented limit */

/* Mode bits: -! -# -c (main.cpp:84-100).  Only F|C may combine (cl_sieve.cpp:600-617). */
#define PFC_MODE_FACTORIAL 1u
#define PFC_MODE_PRIMORIAL 2u
#define PFC_MODE_COMPOSITORIAL 4u

/* Factor record types (cl_sieve.h:4-6). */
#define PFC_FACTORIAL 0
#define PFC_PRIMORIAL 1
#define PFC_COMPOSITORIAL 2

/* One factor record, 16 bytes, the reference's `factor` (cl_sieve.h:8-12,
 * iterate.cl:17-21): nc = -n means p | X(n)-1, nc = +n means p | X(n)+1. */
typedef struct pfc_factor {
    uint64_t p;
    int32_t nc;
    int32_t type;
} pfc_factor;

/* A sieve job: the reference's workStatus search fields (cl_sieve.h:14-18). */
typedef struct pfc_job_desc {
    uint32_t modes;            /* PFC_MODE_* bits */
    uint32_t nmin, nmax;       /* 101 <= nmin < nmax < 2^31 (main.cpp:66-72) */
    uint64_t pmin, pmax;       /* 3 <= pmin < pmax <= 2^64-1 (main.cpp:58-64) */
    uint64_t factor_capacity;  /* 0: 1,000,000, or 30,000,000 when pmin < 2^32-1
                                  (main.cpp:215, cl_sieve.cpp:645-647) */
} pfc_job_desc;

/* Launch shapes; the counterpart of the reference's profile-driven
 * range/sstep/nstep (cl_sieve.cpp:655-769, 1642-1694).  Zero fields mean
 * "library default / autotuned". */
typedef struct pfc_tuning {
    uint32_t primes_per_lane;  /* independent residue chains per lane: 1, 2, 4 */
    uint32_t block_size;       /* walk workgroup size: 64, 128, 256 */
    uint32_t walk_chunk;       /* multiplier steps per walk launch (default 2^18) */
    uint32_t seg_candidates;   /* target candidates per p-segment (default 2^24) */
    uint32_t use_graph;        /* 1: capture a job's segment pipeline (clear, generate, setup,
                                  walk chunks, check) once as a HIP graph and replay it per segment;
                                  and a batcher round's device phase (work-item upload, queue clear,
                                  the persistent set-up, walk and fold launches, control-block
                                  readback) once per batcher, replayed every round */
    uint32_t skip_table_check; /* 1: do not self-check the start tables at job creation
                                  (pfc_stage_verify_table); 0 (default): check them */
    uint32_t lds_segment;      /* integers sieved per generator workgroup and staged in LDS:
                                  15360 (default), 30720 or 61440 */
    uint32_t walk_forms_off;   /* walk arithmetic forms not to take (for comparisons): bit 0 the
                                  narrow form (plain residues, 2^34 < p < 2^55, p * n < 2^62), bit 1
                                  the medium form (2^32 < p < 2^63, n < 2^28); 0 (default): each
                                  launch takes the cheapest form its primes and multipliers allow */
    uint32_t batch_walk_blocks; /* blocks of 64 * primes_per_lane candidates per batcher walk item,
                                  1 to 64; 0 (default): as many as bring an item's steps x blocks
                                  to 32768 while the round keeps four items per wavefront slot */
} pfc_tuning;

/* Host-side summary of a finished job (cl_sieve.cpp:1748). */
typedef struct pfc_job_stats {
    uint64_t dev_checksum;     /* sum of the check terms (check.cl:81-91), wrapping */
    uint64_t primecount;       /* candidates tested (g_sum[0], cl_sieve.cpp:456) */
    uint64_t raw_factors;      /* factor records before the isPrime filter */
    uint64_t segments;         /* p-segments processed */
    double device_ms;          /* stream time of the whole job */
    double walk_ms;            /* summed duration of the walk launches (HIP events) */
    uint64_t walk_launches;    /* walk kernel launches */
    uint64_t walk_updates;     /* p*n residue updates done by those launches */
} pfc_job_stats;

typedef struct pfc_device pfc_device;
typedef struct pfc_job pfc_job;

/* Message of the last failing call on this thread ("" if none). */
const char* pfc_last_error(void);

/* ------------------------------------------------------------------ device */
/* Replaces sclGetCLSoftware/clCreateContext/queue setup (main.cpp:252-330,
 * simpleCL.c:631-668): one HIP device, kernels compiled ahead of time. */
int pfc_device_open(int ordinal, pfc_device** out);
int pfc_device_close(pfc_device* dev);
int pfc_device_info(const pfc_device* dev, char* name, size_t name_cap, int* compute_units,
                    uint64_t* hbm_bytes);
int pfc_set_tuning(pfc_device* dev, const pfc_tuning* t);
int pfc_get_tuning(const pfc_device* dev, pfc_tuning* t);
/* Project autotuner: times the walk kernel at every primes_per_lane x
 * block_size shape on a sample of `desc`, then the n-chunk length
 * (walk_chunk 2^14, 2^16, 2^18 steps per launch) at the fastest shape, then
 * the generator's LDS segment (lds_segment), and stores the winners in the
 * device tuning (also returned in *best).
 * Counterpart of the reference's profile-driven sizing (cl_sieve.cpp:669-769,
 * 1642-1694). */
int pfc_autotune(pfc_device* dev, const pfc_job_desc* desc, pfc_tuning* best);

/* ------------------------------------------------------------------ jobs */
/* Validates like setupSearch (cl_sieve.cpp:596-665) and reports the HBM bytes
 * a job will hold (prime state + tables + factor buffer), for admission. */
int pfc_job_footprint(const pfc_device* dev, const pfc_job_desc* desc, uint64_t* bytes);
/* Builds the host tables (cl_sieve.cpp:842-1326) and allocates device state. */
int pfc_job_create(pfc_device* dev, const pfc_job_desc* desc, pfc_job** out);
/* Runs p-segments [p, min(p + span, pmax)) until pmax or until `max_segments`
 * segments (0 = all) have run: generate -> setup -> walk -> check per segment,
 * the reference's main loop (cl_sieve.cpp:1563-1726).  Blocking. */
int pfc_job_run(pfc_job* job, uint64_t max_segments);
/* Next p to be sieved (== pmax when done); the checkpoint position (workStatus.p). */
int pfc_job_position(const pfc_job* job, uint64_t* p);
/* Resume position and accumulated device sums, from a checkpoint
 * (cl_sieve.cpp:1416-1431).  Only valid before the first pfc_job_run. */
int pfc_job_seek(pfc_job* job, uint64_t p, uint64_t dev_checksum, uint64_t primecount);
int pfc_job_stats_get(const pfc_job* job, pfc_job_stats* st);
/* Raw factor records collected so far, sorted by (p, |n|, type)
 * (cl_sieve.cpp:433-447, 496-501).  With out == NULL only *n is set. */
int pfc_job_factors(const pfc_job* job, pfc_factor* out, size_t cap, size_t* n);
/* Drops the collected factor records (after they were harvested to disk). */
int pfc_job_clear_factors(pfc_job* job);
/* Waits for all work queued on the job's stream (the reference's sclFinish
 * before reading results, cl_sieve.cpp:452).  pfc_job_run already returns
 * with its segments complete; this also covers the queued table self-check. */
int pfc_job_sync(pfc_job* job);
/* getResults' device reads in one call (cl_sieve.cpp:452-501): the device
 * checksum part (sum of check terms, before the n+c of kept factors is
 * added), the candidate count, and the raw factor records sorted as
 * pfc_job_factors returns them (out may be NULL to query *n). */
int pfc_job_result(const pfc_job* job, uint64_t* checksum_partial, uint64_t* primecount, pfc_factor* out,
                   size_t cap, size_t* n);
int pfc_job_destroy(pfc_job* job);
/* Test hook of the walk-completion check: from now on the job's walk leaves
 * out the last `steps` multiplier steps of every segment (each segment's
 * last walk launch, plain, graph or batched, ends that many steps early), so
 * the drain must fail with PFC_ERR_VALIDATION.  0 restores the full walk.  No
 * counterpart in the reference; for tests of the guard only. */
int pfc_job_debug_shorten_walk(pfc_job* job, uint32_t steps);

/* ------------------------------------------------------------------ batching */
/* No counterpart in the reference, which runs one p-range per launch
 * sequence.  A batcher runs many jobs together: each round generates new
 * p-segments for jobs that need one, cuts every active segment into work
 * items (blocks of 64*K candidates x up to max_steps multiplier steps,
 * longest first; a walk item covers more blocks when the round's steps are
 * few, pfc_tuning.batch_walk_blocks)
 * and drains them with persistent kernels fed by device work queues: one
 * launch for the set-up of new segments, one per walk mode for the walk, one
 * for the checksum fold of finished segments.  Job state stays resident on
 * the device between rounds; jobs can be added between rounds (continuous
 * batching).  Results land in the jobs exactly as pfc_job_run leaves them. */
typedef struct pfc_batcher pfc_batcher;
int pfc_batcher_create(pfc_device* dev, pfc_batcher** out);
/* The job stays owned by the caller and must outlive its time in the batcher. */
int pfc_batcher_add(pfc_batcher* b, pfc_job* job);
/* Jobs still in the batcher. */
int pfc_batcher_pending(const pfc_batcher* b, size_t* n);
/* One round (blocking).  max_steps == 0: walk each segment to its end in one
 * round.  Jobs that finished in this round are written to done[0..cap) and
 * counted in *ndone, and leave the batcher. */
int pfc_batcher_round(pfc_batcher* b, uint32_t max_steps, pfc_job** done, size_t cap, size_t* ndone);

/* A round sized by work instead of steps: about max_updates p*n updates over
 * all walking segments, an equal share per segment; a segment that needs less
 * than its share walks everything it has left (up to what one wavefront walks
 * in such a round, or the step count every segment would get if all got the
 * same, whichever is more), the others floor(share / its candidates) steps,
 * at least 256.  Rounds, and the wait of a job admitted between them, stay short while
 * large jobs are resident, and a job of few candidates finishes in a round or
 * a few whatever else is resident.  No counterpart in the reference
 * (serving). */
int pfc_batcher_round_work(pfc_batcher* b, uint64_t max_updates, pfc_job** done, size_t cap,
                           size_t* ndone);
int pfc_batcher_destroy(pfc_batcher* b);
/* Runs jobs to completion through one batcher. */
int pfc_batch_run(pfc_job** jobs, size_t njobs);

/* ------------------------------------------------------------------ harvest */
/* getResults' host half (cl_sieve.cpp:496-593): sorts `f` by (p, |n|, type),
 * verifies every record on the CPU with `verify_threads` OpenMP threads
 * (verifyprime.cpp:149-218), keeps records whose p passes the 7-base
 * Miller-Rabin isPrime (verifyprime.cpp:109-145), adds n + c of each kept
 * record to *checksum, compacts kept records to the front of `f` and returns
 * their number in *nkept.  verify_threads == 0 skips the CPU verification. */
int pfc_harvest(pfc_factor* f, size_t n, int verify_threads, uint64_t* checksum, size_t* nkept);
/* factors.txt line for one kept record: "%llu | %u!%+d", "%llu | %u#%+d" or
 * "%llu | %u!/#%+d" plus '\n' (cl_sieve.cpp:565-577).  Returns the length. */
int pfc_format_factor(const pfc_factor* f, char* buf, size_t cap);
/* Deterministic 7-base Miller-Rabin (verifyprime.cpp:109-145). */
int pfc_is_prime(uint64_t p);
/* Slow CPU check that p divides X(n)+c (verifyprime.cpp:149-218). */
int pfc_verify_factor(uint64_t p, uint32_t n, int32_t c, int32_t type);
/* The reference's number grammar: K/M/G/T/P, eN, k/m/g/t/p, bN suffixes
 * (putil.c:34-116).  Returns 0, -1 (cannot parse) or -2 (out of range). */
int pfc_parse_uint64(const char* s, uint64_t lo, uint64_t hi, uint64_t* out);

/* ------------------------------------------------------------------ stages */
/* Stage-level entry points over caller-owned device memory, on a caller HIP
 * stream (hipStream_t passed as void*; NULL = default stream).  Used by tests
 * and by callers that build their own pipelines.  All pointers are device
 * pointers unless stated.
 *
 * Control block, PFC_CTRL_BYTES (48) bytes, zeroed by the caller before a
 * segment:  u32 ncand, nfactor, flags, pad; u64 checksum, primecount, walked,
 * reserved.
 *
 * `walked` is the walk-completion evidence that replaces the reference's
 * per-lane multiplier counter (.s7 advanced by one every step and converted
 * back and compared with nmax-1 at the check, check.cl:62-73; a mismatch is
 * "gpu validation failure", cl_sieve.cpp:476-480): every wavefront counts the
 * multiplier applications it performed, in a scalar register inside the step
 * loops, and adds count x its live candidates once.  After a segment's walk
 * over steps [j0, j1) it must equal ncand x (multipliers in [j0, j1)); the job
 * and batch drains compare the two and fail with PFC_ERR_VALIDATION. */
#define PFC_CTRL_BYTES 48
typedef struct pfc_ctrl {
    uint32_t ncand;      /* candidates appended (g_primecount[0]) */
    uint32_t nfactor;    /* factor records appended (g_primecount[2]); may exceed capacity */
    uint32_t flags;      /* bit0 candidate overflow, bit1 stored reciprocal failed its check */
    uint32_t pad;
    uint64_t checksum;   /* sum of check terms (wrapping) */
    uint64_t primecount; /* sum of ncand over checked segments */
    uint64_t walked;     /* multiplier applications summed over candidates (walk launches) */
    uint64_t reserved;
} pfc_ctrl;

/* Device memory for stage-level callers that have no allocator of their own
 * (plain hipMalloc/hipFree/hipMemcpy on the device's ordinal).  kind: 1 host
 * to device, 2 device to host, 3 device to device.  pfc_device_synchronize
 * waits for all work on the device. */
int pfc_device_malloc(pfc_device* dev, size_t bytes, void** ptr);
int pfc_device_free(pfc_device* dev, void* ptr);
int pfc_memcpy(pfc_device* dev, void* dst, const void* src, size_t bytes, int kind);
int pfc_device_synchronize(pfc_device* dev);

/* (a) Candidates in [lo, hi): primes below 114 plus wheel-30 numbers with no
 * factor 7..113 that pass the base-2 strong PRP (getsegprimes.cl:170-232,
 * addsmallprimes.cl:21-40).  Appends to d_p[0..cap) in no particular order;
 * ctrl->ncand counts all, flag bit0 when more than cap. */
int pfc_stage_generate(uint64_t lo, uint64_t hi, uint64_t* d_p, uint32_t cap, pfc_ctrl* d_ctrl,
                       void* stream);

/* Host-built start-residue tables (replacing cl_sieve.cpp:842-1326).  The
 * exponents e(r) of the primes r <= N in X(N) are written in base B = 2^k
 * (k = digit_bits, chosen per table by the host to minimise the modular
 * products): group (b, t) holds the primes whose exponent digit at position b
 * is t, multiplied into chunks below 2^63.  Groups are stored position-major
 * from the highest position, digits from B-1 down to 1, so group
 * g = (npos-1-b)*(B-1) + (B-1-t) is chunks[groups[g] .. groups[g+1]), and
 *   X = prod_b (prod_t D_{b,t}^t)^(B^b)
 * with D_{b,t} the product of group (b, t).  The set-up evaluates it by
 * Horner's rule over positions, each H_b = prod_t D_{b,t}^t by suffix
 * products (every chunk multiplied once, one more product per digit). */
typedef struct pfc_setup_table {
    const uint64_t* chunks;    /* device */
    const uint32_t* groups;    /* device: npos*(2^digit_bits - 1) + 1 offsets into chunks */
    uint32_t npos;             /* digit positions of the largest exponent (0: X = 1) */
    uint32_t digit_bits;       /* k, 1..8 */
} pfc_setup_table;

/* (b) Per candidate: reciprocal, Montgomery one (2^64 mod p) and start
 * residues X(nmin-1) mod p for table A (F, P or C) and table B (C in F+C
 * mode) (setup.cl:49-288).  Residues are plain (not Montgomery forms) and
 * stored shifted left by clz64(p); the checksum fold forms the Montgomery
 * terms the reference sums. */
int pfc_stage_setup(const uint64_t* d_p, const pfc_ctrl* d_ctrl, uint32_t* d_v, uint64_t* d_one,
                    uint64_t* d_res_a, uint64_t* d_res_b, const pfc_setup_table* a,
                    const pfc_setup_table* b, uint32_t cap, void* stream);

/* (c) Residue walk over multiplier steps [j0, j1) (iterate.cl:59-198):
 *   modes F:    multiplier n = j                          (res_a)
 *   modes C:    n = j when j is composite                 (res_a)
 *   modes F|C:  n = j for res_a; composite j also res_b   (res_b)
 *   modes P:    multiplier n = d_list[j]                  (res_a)
 * Composite flags: bit (n - nbase) of d_composite (u32 words).
 * Records (p, -+n, type) when a residue becomes +-1, up to fcap records, and
 * adds the multiplier applications performed (candidates x steps) to
 * ctrl->walked.  Ranges: F, C and F|C need 2 <= j0 and j1 <= 2^31 - 1 (the
 * multiplier is j); P needs j1 <= 2^31 - 1 and every d_list value below 2^31;
 * anything else is PFC_ERR_INVALID.
 * The residues must be canonical (below p << clz64(p)), as pfc_stage_setup
 * and this call leave them.  (Where every p of a wavefront exceeds twice the
 * largest multiplier j1 - 1, F, C and F|C steps are tested in pairs, and a
 * pair with both multipliers at most 65,536 is one step by their product;
 * such a wavefront walks plain residues below 2p when 2^34 < p < 2^55 and
 * p * (j1 - 1) < 2^62, or half-scale residues when p < 2^63, converting on
 * entry and returning them canonical.  P steps are tested one by one here,
 * in the shifted form, since the call cannot bound d_list's values.  The
 * records are the same either way.) */
int pfc_stage_walk(uint32_t modes, uint32_t j0, uint32_t j1, const uint32_t* d_list,
                   const uint32_t* d_composite, uint32_t nbase, const uint64_t* d_p,
                   const uint32_t* d_v, uint64_t* d_res_a, uint64_t* d_res_b, pfc_ctrl* d_ctrl,
                   pfc_factor* d_fac, uint32_t fcap, uint32_t cap, void* stream);

/* (c) Checksum fold (check.cl:45-204): adds every candidate's term to
 * ctrl->checksum and ncand to ctrl->primecount; validates the reciprocals. */
int pfc_stage_check(uint32_t modes, uint32_t nmax, const uint64_t* d_p, const uint32_t* d_v,
                    const uint64_t* d_one, const uint64_t* d_res_a, const uint64_t* d_res_b,
                    pfc_ctrl* d_ctrl, uint32_t cap, void* stream);

/* Builds the setup table for X(N) on the host: kind 0 = N!, 1 = N#, 2 = N!/N#
 * (layout as pfc_setup_table).  Writes up to chunk_cap chunks and group_cap
 * group offsets and reports the full sizes in *nchunks and *ngroups
 * (= npos*(2^digit_bits - 1) + 1), the positions in *npos and the digit
 * width in *digit_bits; with chunks == NULL and groups == NULL only the sizes
 * are returned. */
int pfc_build_setup_table(uint32_t kind, uint32_t N, uint64_t* chunks, size_t chunk_cap,
                          size_t* nchunks, uint32_t* groups, size_t group_cap, size_t* ngroups,
                          uint32_t* npos, uint32_t* digit_bits);

/* Start-table self-check (verifyslow.cl, verify.cl, verifyresult.cl;
 * cl_sieve.cpp:917-970, 1079-1134, 1246-1304): X(N) of a table of `kind`
 * (0: N!, 1: N#, 2: N!/N#) modulo the test prime 2^64-59, once as the
 * product of every multiplier up to N (primality of each n decided on the
 * device by a deterministic 32-bit Miller-Rabin) and once through the
 * table's start-residue path.  d_work: device buffer of PFC_VERIFY_WORDS u64;
 * afterwards d_work[0] = Mont(X) from the product and d_work[1] = the
 * table's start residue; the table is right when they are equal.
 * pfc_job_create queues this for every table it builds unless
 * pfc_tuning.skip_table_check is set; the job's first finished segment
 * (pfc_job_run or a batcher round) compares the results and fails with
 * PFC_ERR_VALIDATION ("power table verification failed") on a mismatch. */
#define PFC_VERIFY_BLOCKS 256
#define PFC_VERIFY_WORDS (2 + PFC_VERIFY_BLOCKS)
int pfc_stage_verify_table(uint32_t kind, uint32_t N, const pfc_setup_table* t, uint64_t* d_work,
                           void* stream);

/* The same self-check without a workspace argument and without allocating:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The block
 * partials live in a library-owned device array of PFC_NOALLOC_SCRATCH_BYTES
 * bytes per GPU, held until the process ends (its size does not depend on
 * kind, N or the table, so no valid input exceeds it); the results go to
 * d_out[0] (Mont(X) from the product) and d_out[1] (the table's start
 * residue), as d_work[0..1] of pfc_stage_verify_table.  Calls on one GPU
 * share that array: calls that overlap on one GPU (other streams or
 * threads) race, so a caller must order them.  No hipMalloc, hipFree or other
 * allocation is made; the array's device address is looked up once per GPU
 * on the first call there. */
#define PFC_NOALLOC_SCRATCH_BYTES (8 * PFC_VERIFY_WORDS)
int pfc_stage_verify_table_noalloc(uint32_t kind, uint32_t N, const pfc_setup_table* t, uint64_t* d_out,
                                   void* stream);

/* ------------------------------------------------------------------ probe */
/* Measures the lane-op rate of independent v_mad_u64_u32 / v_mul_lo_u32 /
 * v_mul_hi_u32 streams on all CUs (the INT32 multiply roofline).  rates[3]
 * receives lane-ops per second for {mad_u64_u32, mul_lo_u32, mul_hi_u32}. */
int pfc_probe_mul_rates(pfc_device* dev, double rates[3]);
/* One launch of a plain device copy of `bytes` (read + write), the reference
 * kernel for scaling HBM traffic counters; *ms receives its duration. */
int pfc_probe_copy(pfc_device* dev, uint64_t bytes, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* PFC_H */
